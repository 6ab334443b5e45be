// hc_transport.hip -- the photon-packet transport pass on the GPU and the C ABI
// around it (include/hc.h).
//
// Execution model: one persistent launch per pass.  Every lane pulls packet
// indices from a device counter and carries each packet from launch to
// escape, star re-emission or kill (the body of propagate.f90 around the
// reconstructed tauint), so lanes stay busy until the pass drains regardless
// of how unevenly the packets random-walk.  Packet p draws from its own
// segment of the L'Ecuyer-Cote generator, so results do not depend on which
// lane ran it: the deterministic-trace mode compares packet by packet with
// the CPU restatement.
//
// Device layout: per-cell data are cell-major with the active species
// innermost (dens[cell*nact + k]), cell = (ir-1) + nrg*((it-1) + ntg*(ip-1))
// as in the reference's arrays, so one crossing reads one contiguous run of
// nact doubles.  densm holds the same densities with species at or above the
// sublimation temperature zeroed (peeloff_3d.f90:103-110), so a crossing needs
// no temperature read.  Tallies are FP64 device atomics.
#include "hc_device.h"
#include "../../include/hc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace hc {

constexpr int kGuide = 512;  // bins of the opacity-table guide
constexpr int kCdfGuide = 1024;  // probability bins of the emissivity-CDF guide
constexpr int kAtmGuide = 4096;  // probability bins of the stellar-spectrum CDF guide (2^12:
                                 // floor(xi kAtmGuide) and b / kAtmGuide are exact)

struct Params {
  // grid
  int nrg, ntg, npg, nact;
  int act[kMaxSpecies];      // reference species id (1-based) of slot k
  const double *r2arr, *tan2arr, *thetarr, *aarr, *barr, *phiarr;
  const double *dens, *densm, *tdust;
  // dust
  int nlammax, nT, nnu;
  int nlambda[kMaxSpecies];
  const double *nudust, *kappad, *adust, *gdust, *pldust;
  double kappav[kMaxSpecies];
  const double *lTint, *nuint, *kapBint;
  // per species id and row j: number of CDF nodes <= b/kCdfGuide, b = 0..kCdfGuide
  const uint16_t *cdfguide;
  const int32_t *guide;      // [ndg][kGuide] bracket start per log-frequency bin
  double glog0[kMaxSpecies], ginv[kMaxSpecies];
  // sources
  double rstar, rmincgs, tstar, tshock;
  int ips, ispot, nspot, ipeel;
  double fspot, csprad, cspot1, sspot1, scspot1, ssspot1, cspot2, sspot2, scspot2, ssspot2;
  long long np, npacc, npaccstar;
  // binning
  int nfreq, nmu, nph, nap, npeel, nxhst, imcube, iintmax;  // imcube 2: FP32 cube words
  double numin, lnurat, rmaxi, fractxh;
  const double *u, *sini, *aperture2, *coste, *sinte, *cospe, *sinpe;
  // rng
  int32_t g1, g2;
  unsigned long long seg0;
  const unsigned long long *seg_dev;  // segment base on the device (graph replays), or null
  const int32_t *pow1, *pow2;
  // tallies
  double *dtauabs, *nabs, *jmean;
  int32_t *killed, *dbig;
  double *sed, *peel, *img_i, *img_q, *img_u, *img_v, *si_init;
  double *scal;
  unsigned long long *next_packet;
  hc_trace *trace;
  // launch shape
  int regen_lanes;  // waiting lanes (1..64) of one event kind that end a walk phase
  int regen_run_lanes;  // then every kind with this many waiters runs (<= regen_lanes)
  int peel_q;       // peel records per wavefront queue
  int walls_lds;    // bytes of LDS holding the wall arrays
  int opac_base;    // LDS byte offset of the opacity tables
  int guide_off;    // byte offset of the opacset guides within them
  int opac_off[kMaxSpecies];  // table of active species s: offset in doubles
  int emis_base;    // LDS byte offset of lTint[nT] and nuint[nnu]
  int peel_base;    // LDS byte offset of the peel-off queues
  int pool_base;    // LDS byte offset of the workgroup's event pool (pooled passes)
  int pool_min;     // requests of one kind a walking wavefront waits for before it runs them
  int pool_walk;    // below this many walking lanes a wavefront runs any pooled request
  // disk-accretion launch (initpacc.f90, emit_gasacc.f90)
  const double *costarr;  // cos of the theta walls (gridset.f90:580-599)
  double rddustmin, rmaxdmax, rmax, massc, mdot;
  double fmass[kMaxSpecies], z1[kMaxSpecies], bflare[kMaxSpecies];
  int slot_of[kMaxSpecies];  // active slot of reference species id, -1 if none
  uint32_t max_work;  // events + crossings after which a packet is dropped
  // stellar spectrum (atmosinit/atmosfreq): nhnu = 0 samples the Planck
  // function at tstar instead
  int nhnu;
  const double *atm;          // atmosnu, hnuint, lognu, loghnu [nhnu each]
  const int32_t *atm_guide;   // [kAtmGuide + 1]: table nodes with hnuint below b / kAtmGuide
  // Eddington limb darkening (samptabl.f90): xmu[ntab], prob[ntab]
  int limb, ntab;
  const double *limb_tab;
  double delmu;
  // Bjorkman-Wood re-emission (ilucy = 0): each absorption brings its
  // cell's temperature up to date in tdust_w (and the walk's mask densm_w)
  // and re-emits from kapdBint; serial: one lane runs the packets in order
  int bw, serial;
  const double *kapdBint, *kappap, *mass4;
  const uint16_t *cdfguide_bw;
  double *tdust_w, *densm_w;
  double tsub_ev, ltot, nphot;
  // filter-convolved images (peeled passes): filt_mod's curves and the
  // (npeel,nx,nx,nfilt) I, Q, U, V accumulators, one block after the other
  int nfilt, nwfiltmax;
  const int32_t *nwfilt;
  const double *filtwave, *filtphi;
  double *imgb;
  // diffusion approximation (gridset.f90:666-789, emit_common.f90:80-387):
  // the layer mask and direction per cell (int8, cell_of layout; null when
  // off), the radial walls, sin of the theta walls, the Rosseland table
  // kappar[id][nT], and the pass's stellar packet count
  const int8_t *diffus, *diffdir;
  const double *rarr, *sintarr, *kappar;
  double nstar;
};

// Launch parameters are the transport kernel's argument: every wave reads
// them from the kernel-argument segment with scalar loads, no inlined helper
// has to carry them, and passes of several contexts can run at once on their
// own streams.  Only code inlined into the kernel may call kernel_params():
// a function the compiler emits as a call does not receive the segment's
// address (the builtin yields a null pointer there), so every function marked
// noinline takes what it needs as arguments.
//
// A batched launch runs the passes of several models in one grid: the
// kernel's argument then also points at a device table of the models'
// Params and at the model of each workgroup, and every wave reads its
// workgroup's entry, again with scalar loads (the table is read through the
// constant address space, which the kernel never writes).
struct KernArgs {
  Params p;                     // single-model launch
  const Params *table;          // batched launch: the models' parameters, else null
  const uint32_t *block_model;  // batched launch: model of each workgroup
};
static_assert(sizeof(KernArgs) <= 4096, "launch parameters exceed the kernel-argument limit");
typedef const __attribute__((address_space(4))) Params ConstParams;
typedef const __attribute__((address_space(4))) uint32_t ConstU32;
__device__ __forceinline__ const Params &kernel_params() {
#if defined(__HIP_DEVICE_COMPILE__)
  const KernArgs &ka = *reinterpret_cast<const KernArgs *>(__builtin_amdgcn_kernarg_segment_ptr());
  if (ka.table == nullptr) return ka.p;
  ConstU32 *bm = (ConstU32 *)ka.block_model;
  ConstParams *t = (ConstParams *)ka.table + bm[blockIdx.x];
  return *(const Params *)t;
#else
  __builtin_unreachable();  // the host pass never runs device code
#endif
}
#define c_P (::hc::kernel_params())

enum Scalar {
  S_TOT = 0, S_FLUX, S_SFLUX, S_NSCAT, S_SCOUNT, S_ABFLUX, S_AFLUX, S_SUMSUB, S_SUMALL,
  S_HIT, S_HTOT, S_ERR_WALL, S_ERR_DUST, S_WARN, S_NCROSS, S_NSCAT_EV, S_NABS_EV, S_NPEELRAY,
  S_NPEELCROSS, S_NTALLY, S_COUNT
};

// Device counters of a pass (Params::next_packet): the next packet number,
// the longest packet's work, the number of packets dropped at max_work and,
// from kCtrDropped on, the first kDropMax of their packet numbers.
constexpr int kCtrNext = 0, kCtrLongest = 1, kCtrNDropped = 2, kCtrDropped = 4;
constexpr int kDropMax = 16;
constexpr int kCounters = kCtrDropped + kDropMax;

// ---------------------------------------------------------------------------
// packet state (registers)
// ---------------------------------------------------------------------------
template <class R, int NA>
struct Pkt {
  double x, y, z, ux, uy, uz;  // position and direction of the walk: FP64
  Dir<R> d;
  Stokes4<R> s;
  R nub;
  double rsq;
  int ir, it, ip;
  bool on_wall;
  int iflag, i_orig;
  double kapd[NA];  // kappa R* kappa_V per active species (walk opacity)
  R alb[NA];
};

struct Acc {  // per-lane counters, flushed once per lane
  uint32_t nscat, hit, htot, err_wall, err_dust;
  uint32_t ncross, nscat_ev, nabs_ev, npeelray, npeelcross, warn, maxwork;
};

// Counters of the once-per-packet events (launch, exit), one set per
// wavefront in LDS rather than in every lane's registers (registers the walk
// needs: with them the FP64 kernel spilled to scratch, and a scratch reload
// waits on the same in-order counter as the density loads and the tally
// atomics).  Lanes add with LDS atomics; each wavefront flushes its own set.
struct WaveCold {
  unsigned int tot, flux, sflux, scount, abflux, aflux;
  double sumsub, sumall;
};
constexpr int kMaxWaves = 8;  // wavefronts of the largest workgroup (512 lanes)
__shared__ WaveCold hc_cold[kMaxWaves];
__device__ __forceinline__ WaveCold &cold() { return hc_cold[threadIdx.x >> 6]; }
__device__ __forceinline__ void cold_inc(unsigned int &c) { atomicAdd(&c, 1u); }

// Lucy tally of the cell the packet is in: segments accumulate here and go to
// the device arrays with one atomic per species when the packet leaves the
// cell (a packet random-walking in an optically thick cell adds thousands of
// segments to the same cell).
template <class R, int NA>
struct LucyAcc {
  long long cell;
  double tau[NA];
  double len;
  uint32_t nabs[NA];  // absorptions per species (nabs)
  int nabs_any;     // absorptions of any species (dbig)
  uint32_t ntally;  // per-cell tally atomics issued by this lane (all its packets)
};

struct Trace {
  double path, tau_peel;
  long long ncross, npeelcross, nscat, nabs, nstar;
  unsigned long long hash;
};

__device__ __forceinline__ size_t cell_of(int ir, int it, int ip) {
  const Params &P = c_P;
  return (size_t)(ir - 1) + (size_t)P.nrg * ((size_t)(it - 1) + (size_t)P.ntg * (size_t)(ip - 1));
}

__device__ __forceinline__ void hash_cell(Trace &t, int ir, int it, int ip) {
  t.hash ^= (unsigned long long)(unsigned)ir | ((unsigned long long)(unsigned)it << 20) |
            ((unsigned long long)(unsigned)ip << 40);
  t.hash *= 1099511628211ull;
}

__device__ __forceinline__ void atomic_add(double *p, double v) { unsafeAtomicAdd(p, v); }

__device__ __forceinline__ int clampi(int v, int lo, int hi) {
  return v < lo ? lo : (v > hi ? hi : v);
}

// Dynamic LDS of the transport kernel (layout set on the host, Params):
//   [0, walls_lds)            wall arrays r2[nrg] t2[ntg] a[npg] b[npg],
//                             then rr[nrg] tc[ntg] ts[ntg] (wall_clearance)
//   [opac_base, +opac_bytes)  per active species s: nudust, kappad, adust,
//                             gdust, pldust (nlambda doubles each) at
//                             opac_off[s] doubles, then the int32 opacset
//                             guides [nact][kGuide] at guide_off bytes
//   [peel_base, ...)          peel-off queue counters and queues
extern __shared__ __align__(16) unsigned char hc_lds[];

// The wall arrays of the grid as the walk reads them: staged in LDS by the
// transport kernel (r^2 of the radial walls, tan^2 of the theta walls, the
// phi planes' coefficients), so the roots of a crossing wait on no global
// memory.
struct WallView {
  const double *r2, *t2, *a, *b;
  const double *rr, *tc, *ts;  // radii; |cos| and sin of the theta walls
};

// Per-cell densities are stored once, signed: +rho for a species below its
// sublimation temperature, -rho above it.  The walk sees the masked density
// (peeloff_3d.f90:103-110), select_dust the full one (select_dust.f90), from
// the same load.
__device__ __forceinline__ double masked(double d) { return d > 0.0 ? d : 0.0; }

// Masked opacity of a cell: sum over active species in increasing id.
template <class R, int NA>
__device__ __forceinline__ double cell_opac(const Pkt<R, NA> &k, const double (&dm)[NA]) {
  double opac = 0.0;
#pragma unroll
  for (int s = 0; s < NA; ++s)
    if (s < c_P.nact) opac = opac + k.kapd[s] * masked(dm[s]);
  return opac;
}

// On the polar axis (x = y = 0) every phi wall passes through the point, so
// phiface's roots are all 0 and the walk would step from phi cell to phi cell
// without moving; a straight ray leaving the axis keeps the azimuth of its
// direction, so there the phi walls are skipped and the caller has put the
// packet in the phi cell of its direction (axis_phi_cell).
__device__ __forceinline__ bool next_wall(const WallView &W, double ux, double uy, double uz,
                                          double x, double y, double z, int ir, int it, int ip,
                                          bool on_wall, double &dt, int &face) {
  const Params &P = c_P;
  Walls w;
  w.reset();
  radface(w, ux, uy, uz, x, y, z, W.r2[ir - 1], W.r2[ir]);
  if (P.ntg > 2) thetaface(w, ux, uy, uz, x, y, z, W.t2[it - 1], W.t2[it]);
  if (P.npg > 2 && P.ntg > 2 && !(x == 0.0 && y == 0.0))
    phiface(w, W.a[ip - 1], W.b[ip - 1], W.a[ip], W.b[ip], ux, uy, x, y);
  return w.next(on_wall, dt, face);
}

// The phi cell a ray leaving the polar axis enters: the one holding the
// azimuth of its direction (unchanged when the ray runs along the axis).
// Not inlined (a rare path), so it takes the phi walls as arguments: the
// launch parameters are read from the kernel-argument segment, whose address
// only the kernel's own code has.
__device__ __attribute__((noinline)) int axis_phi_cell_call(const double *phiarr, int npg,
                                                            double ux, double uy, int ip) {
  if (npg <= 2 || (ux == 0.0 && uy == 0.0)) return ip;
  double p = hcm_atan2(uy, ux);
  if (p < 0.0) p = p + kR2P;
  return locate(phiarr, npg, p);
}

__device__ __forceinline__ int axis_phi_cell(double ux, double uy, int ip) {
  return axis_phi_cell_call(c_P.phiarr, c_P.npg, ux, uy, ip);
}

// The cell across wall `face`: 1/2 the inner/outer sphere, 3/4 the theta
// walls, 5 the lower phi plane, anything else the upper one; phi is periodic.
// Written as selects: a switch over the three references compiles to a
// pointer chosen at run time, which puts the cell indices in scratch memory,
// and every scratch access then waits on the in-order memory counter that
// the density loads and the tally atomics also count on.
__device__ __forceinline__ void face_step(int face, int &ir, int &it, int &ip) {
  ir += face == 1 ? -1 : (face == 2 ? 1 : 0);
  it += face == 3 ? -1 : (face == 4 ? 1 : 0);
  ip += face == 5 ? -1 : ((face < 1 || face > 5) ? 1 : 0);
  if (ip == 0) ip = c_P.npg - 1;
  if (ip == c_P.npg) ip = 1;
}

enum { ST_INTERACT = 0, ST_EXIT = 1, ST_STAR = 2, ST_ERROR = 3 };

template <class R, int NA>
__device__ __forceinline__ void lucy_flush(LucyAcc<R, NA> &la) {
  const Params &P = c_P;
  if (la.cell < 0) return;
  if (P.dtauabs) {
    double *ta = P.dtauabs + (size_t)la.cell * (size_t)P.nact;
#pragma unroll
    for (int s = 0; s < NA; ++s)
      if (s < P.nact && la.tau[s] != 0.0) {
        atomic_add(&ta[s], la.tau[s]);
        la.ntally++;
      }
  }
  if (P.jmean && la.len != 0.0) {
    atomic_add(&P.jmean[la.cell], la.len * P.rmincgs);
    la.ntally++;
  }
  if (la.nabs_any) {
    double *na = P.nabs + (size_t)la.cell * (size_t)P.nact;
#pragma unroll
    for (int s = 0; s < NA; ++s)
      if (s < P.nact && la.nabs[s] != 0u) {
        atomic_add(&na[s], (double)la.nabs[s]);
        la.ntally++;
      }
    atomicAdd(&P.dbig[la.cell], la.nabs_any);
    la.ntally++;
  }
  la.cell = -1;
}

template <class R, int NA>
__device__ __forceinline__ void lucy_open(LucyAcc<R, NA> &la, size_t cell) {
  if (la.cell != (long long)cell) {
    lucy_flush(la);
    la.cell = (long long)cell;
#pragma unroll
    for (int s = 0; s < NA; ++s) {
      la.tau[s] = 0.0;
      la.nabs[s] = 0u;
    }
    la.len = 0.0;
    la.nabs_any = 0;
  }
}

// Add one chord of cell `cell` to the per-cell accumulator: dtauabs gains
// kapd (1-albedo) rho ds per active species, jmean gains ds R*.
template <class R, int NA>
__device__ __forceinline__ void lucy_add(LucyAcc<R, NA> &la, const Pkt<R, NA> &k,
                                         const double (&dm)[NA], size_t cell, double ds) {
  if (!c_P.dtauabs && !c_P.jmean) return;
  lucy_open(la, cell);
#pragma unroll
  for (int s = 0; s < NA; ++s)
    if (s < c_P.nact && dm[s] > 0.0) la.tau[s] += k.kapd[s] * (1.0 - (double)k.alb[s]) * dm[s] * ds;
  la.len += ds;
}

// The cell walk to optical depth tau (tauint reconstruction), split into
// its entry (the packet on a radial wall steps into the cell its direction
// points to) and single crossings, so the event loop can interleave lanes.
enum { ST_WALK = 4, ST_DONE = 6, ST_ABORT = 7, ST_SCAT = 8, ST_ABS = 9, ST_END = 10,
       ST_RESTART = 11, ST_SCAT_WAIT = 12, ST_ABS_WAIT = 13, ST_SCAT_READY = 14,
       ST_ABS_READY = 15 };

template <class R, int NA>
__device__ __forceinline__ int walk_begin(const WallView &W, Pkt<R, NA> &k) {
  const Params &P = c_P;
  if (k.on_wall) {
    const bool inner = fabs(k.rsq - W.r2[k.ir - 1]) < fabs(k.rsq - W.r2[k.ir]);
    const bool outwards = k.ux * k.x + k.uy * k.y + k.uz * k.z > 0.0;
    if (inner && !outwards) {
      k.ir -= 1;
      if (k.ir == 0) return ST_STAR;
    }
    if (!inner && outwards) k.ir += 1;
    if (k.ir == P.nrg) return ST_EXIT;
  }
  return ST_WALK;
}

// A lower bound of the distance from the packet to every wall surface of its
// cell: |r - R| to the two spheres, the meridional distance
// |rho |cos t| - |z| sin t| to the two double cones (no point of either nappe
// is closer), and |a x + b y| to the two full phi planes.  Every root the
// walk can take lies at least that far along the unit direction, so an
// interaction point closer than the bound is reached before any wall.  The
// margin, 1e-6 r, is far above the rounding of the roots, also for rays
// almost tangent to a wall; a packet off its cell's range gets a negative
// bound.
template <class R, int NA>
__device__ __forceinline__ double wall_clearance(const WallView &W, const Pkt<R, NA> &k) {
  const Params &P = c_P;
  const double r = sqrt(k.rsq);
  double d = fmin(r - W.rr[k.ir - 1], W.rr[k.ir] - r);
  if (P.ntg > 2) {
    const double rho = sqrt(k.x * k.x + k.y * k.y), az = fabs(k.z);
    d = fmin(d, fabs(rho * W.tc[k.it - 1] - az * W.ts[k.it - 1]));
    d = fmin(d, fabs(rho * W.tc[k.it] - az * W.ts[k.it]));
  }
  if (P.npg > 2) {
    d = fmin(d, fabs(W.a[k.ip - 1] * k.x + W.b[k.ip - 1] * k.y));
    d = fmin(d, fabs(W.a[k.ip] * k.x + W.b[k.ip] * k.y));
  }
  return d - 1e-6 * r;
}

// One step of the walk: to the next wall, or to the interaction point when
// the optical depth tau is reached inside the current cell.
// The masked densities of the packet's cell, into registers.
template <class R, int NA>
__device__ __forceinline__ void load_dens(const Pkt<R, NA> &k, double (&dm)[NA]) {
  const Params &P = c_P;
  if (k.ir < 1 || k.ir >= P.nrg) return;  // outside the grid: nothing to read
  const double *d = P.densm + cell_of(k.ir, k.it, k.ip) * (size_t)P.nact;
#pragma unroll
  for (int s = 0; s < NA; ++s) dm[s] = (s < P.nact) ? d[s] : 0.0;
}

// dm holds the masked densities of the packet's cell on entry and those of
// the next cell on return: the load for the next crossing is issued as soon
// as that cell is known, and its latency overlaps the rest of this step.
template <class R, int NA, bool TRACE>
__device__ __forceinline__ int walk_step(const WallView &W, Pkt<R, NA> &k, double tau,
                                         double &tsum, double (&dm)[NA], LucyAcc<R, NA> &la,
                                         Acc &acc, Trace &tr) {
  const Params &P = c_P;
  if (k.x == 0.0 && k.y == 0.0) {
    k.ip = axis_phi_cell(k.ux, k.uy, k.ip);
    load_dens(k, dm);
  }
  const size_t cell = cell_of(k.ir, k.it, k.ip);
  auto interact = [&](double ds) __attribute__((always_inline)) {
    lucy_add(la, k, dm, cell, ds);
    k.x = k.x + ds * k.ux;
    k.y = k.y + ds * k.uy;
    k.z = k.z + ds * k.uz;
    k.on_wall = false;
    k.rsq = k.x * k.x + k.y * k.y + k.z * k.z;
    if constexpr (TRACE) tr.path += ds;
    return ST_INTERACT;
  };
  // After an interaction (the packet is off the walls) the next one is often
  // closer than any wall of the cell: then the walls' roots are not needed,
  // and the point is the one the full step below reaches.
  if (!k.on_wall) {
    const double ds = (tau - tsum) / cell_opac(k, dm);  // opac = 0: inf or NaN, no shortcut
    if (ds < wall_clearance(W, k)) return interact(ds);
  }
  // A packet on a wall has just crossed it: its densities were loaded at the
  // crossing, and the flush atomics issued after that load share its in-order
  // counter.  The roots come first, so that wait overlaps them.
  double dt;
  int face;
  if (!next_wall(W, k.ux, k.uy, k.uz, k.x, k.y, k.z, k.ir, k.it, k.ip, k.on_wall, dt, face)) {
    acc.err_wall++;
    return ST_ERROR;
  }
  const double opac = cell_opac(k, dm);
  const double dtau = opac * dt;
  if (tsum + dtau >= tau) {
    double ds = (tau - tsum) / opac;
    if constexpr (sizeof(R) == sizeof(float)) {
      // an interaction that rounding puts at the wall stays 1e-6 of the step
      // inside the cell, so the walk from it starts, in any new direction,
      // from a point its cell's walls bound (the FP64 path keeps the
      // reference's clamp to the wall itself)
      const double dmax = dt * (1.0 - 1e-6);
      if (ds > dmax) ds = dmax;
    } else {
      if (ds > dt) ds = dt;  // rounding must not carry the packet past the wall
    }
    return interact(ds);
  }
  lucy_add(la, k, dm, cell, dt);
  tsum = tsum + dtau;
  face_step(face, k.ir, k.it, k.ip);
  load_dens(k, dm);
  // the packet leaves the cell: its tallies go out now, after the next
  // cell's density load is issued.  Loads and atomics share one in-order
  // counter, so a load issued after the atomics would also wait for them.
  lucy_flush(la);
  k.x = k.x + dt * k.ux;
  k.y = k.y + dt * k.uy;
  k.z = k.z + dt * k.uz;
  k.on_wall = true;
  k.rsq = k.x * k.x + k.y * k.y + k.z * k.z;
  acc.ncross++;
  if constexpr (TRACE) {
    tr.path += dt;
    tr.ncross++;
    hash_cell(tr, k.ir, k.it, k.ip);
  }
  if (k.ir == P.nrg) return ST_EXIT;
  if (k.ir == 0) return ST_STAR;
  return ST_WALK;
}

// opacset for one active species (dust.f90:13-76), slot s of the active
// list, from its tables in LDS.  The bracket of nu in the decreasing
// frequency table is found from a guide table of log-frequency bins and a
// short forward scan; for a strictly decreasing table that is the same unique
// bracket the reference's bisection finds.
__device__ __forceinline__ Opac1 opacset_species(int s, double nu) {
  const Params &P = c_P;
  const int id = P.act[s] - 1;
  const int nl = P.nlambda[id];
  const double *nud = reinterpret_cast<const double *>(hc_lds + P.opac_base) + P.opac_off[s];
  Opac1 r;
  int i0, i1;
  double f;
  if (nu >= nud[0]) {
    i0 = i1 = 1;
    f = 0.0;
  } else if (nu <= nud[nl - 1]) {
    i0 = i1 = nl;
    f = 0.0;
  } else {
    // the bin only seeds the scan, so a single-precision log suffices
    int b = (int)((P.glog0[id] - (double)__logf((float)nu)) * P.ginv[id]);
    b = clampi(b - 1, 0, kGuide - 1);
    i1 = reinterpret_cast<const int32_t *>(hc_lds + P.opac_base + P.guide_off)[s * kGuide + b];
    while (!(nu > nud[i1 - 1])) ++i1;
    i0 = i1 - 1;
    f = (nu - nud[i1 - 1]) / (nud[i0 - 1] - nud[i1 - 1]);
  }
  const double *kd = nud + nl, *ad = nud + 2 * nl, *gd = nud + 3 * nl, *pd = nud + 4 * nl;
  if (i0 == i1) {
    r.g = gd[i0 - 1];
    r.pl = pd[i0 - 1];
    r.kappa = kd[i0 - 1];
    r.albedo = ad[i0 - 1];
  } else {
    const double fm1 = 1.0 - f;
    r.g = fm1 * gd[i1 - 1] + f * gd[i0 - 1];
    r.pl = fm1 * pd[i1 - 1] + f * pd[i0 - 1];
    r.kappa = fm1 * kd[i1 - 1] + f * kd[i0 - 1];
    r.albedo = fm1 * ad[i1 - 1] + f * ad[i0 - 1];
  }
  return r;
}

// Henyey-Greenstein terms of active species s at frequency nu.
template <class R>
__device__ __forceinline__ HG<R> hg_of(int s, R nu) {
  const Opac1 op = opacset_species(s, (double)nu);
  return HG<R>(R(op.g), R(op.pl));
}

struct PeelOut {
  double tau;
  int ncross;
  int err;
};

// What a peel-off ray needs of the packet at its emission or scattering
// event.  Records wait in a per-wavefront queue in LDS; the wavefront's lanes
// walk their (record, observer) rays (flush_peels).
template <class R>
struct Normal4 {  // the stellar surface normal at a launch point
  R cosb, sinb, coslp, sinlp;
};

// A direct stellar record's Stokes vector is (1,0,0,0): every stellar launch
// and re-emission is unpolarised with I = 1.  Its Stokes slots hold the
// surface normal instead, which keeps the record at 168 B (FP64).
template <class R, int NA>
struct PeelRec {
  double x, y, z, rsq;
  double kapd[NA];
  Dir<R> d;
  union {
    Stokes4<R> s;    // scattered or re-emitted by dust, accretion
    Normal4<R> n;    // direct stellar emission
  } u;
  R nub;
  int ir, it, ip;
  int flags;                   // on_wall | iflag<<1 | i_orig<<2 | slot<<5 | direct<<9
  long long pid;
};

static_assert(sizeof(PeelRec<double, 4>) == 168, "FP64 peel record layout");
static_assert(sizeof(PeelRec<float, 4>) == 128, "FP32 peel record layout");

template <class R, int NA>
__device__ __forceinline__ void make_rec(PeelRec<R, NA> &r, const Pkt<R, NA> &k, int slot,
                                         bool direct, long long pid, const R *nrm) {
  r.x = k.x;
  r.y = k.y;
  r.z = k.z;
  r.rsq = k.rsq;
#pragma unroll
  for (int s = 0; s < NA; ++s) r.kapd[s] = k.kapd[s];
  r.d = k.d;
  if (direct)
    r.u.n = Normal4<R>{nrm[0], nrm[1], nrm[2], nrm[3]};
  else
    r.u.s = k.s;
  r.nub = k.nub;
  r.ir = k.ir;
  r.it = k.it;
  r.ip = k.ip;
  r.flags = (k.on_wall ? 1 : 0) | (k.iflag << 1) | (k.i_orig << 2) | (slot << 5) |
            ((direct ? 1 : 0) << 9);
  r.pid = pid;
}

template <class R, int NA>
__device__ __forceinline__ double rec_opac(const PeelRec<R, NA> &k, const double *dm) {
  double opac = 0.0;
#pragma unroll
  for (int s = 0; s < NA; ++s)
    if (s < c_P.nact) opac = opac + k.kapd[s] * masked(dm[s]);
  return opac;
}

// peeloff_3d.f90:5-288, one (record, observer) ray at a time, split so that
// a wavefront can walk many rays with lane refill: peel_begin sets the ray up
// (peeloff_3d.f90:5-72), peel_step crosses one wall (:74-154), peel_tally
// adds the ray to the peeled SED and the image cube (:155-288).  The
// arithmetic of a ray is that of the serial routine, in the same order.
enum { kPeelIdle = 0, kPeelWalk = 1, kPeelTally = 2 };

struct PeelWalk {
  double x, y, z, ux, uy, uz, tsum, w, tau;
  int ir, it, ip, ncross, err, r, rec, obs;  // r = rec * npeel + obs
  bool on_wall;
};

// Ray r = (record r / npeel, observer r % npeel) of the queue.  Returns
// kPeelIdle for a direct stellar ray the observer cannot see (not a ray) or
// for a ray that ends at once (counted by the caller when S.r >= 0).
template <class R, int NA>
__device__ __forceinline__ int peel_begin(const WallView &W, const PeelRec<R, NA> *q, int r,
                                          PeelWalk &S) {
  const Params &P = c_P;
  const int ri = r / P.npeel, obs = r - ri * P.npeel;
  const PeelRec<R, NA> &k = q[ri];
  S.r = -1;
  if ((k.flags >> 9) & 1) {  // direct stellar emission (newdisktherm.F90:342-377)
    const Normal4<R> &nm = k.u.n;
    if (!P.ips) {
      const R cosnorm =
          nm.cosb * R(P.coste[obs]) +
          (nm.sinb * R(P.sinte[obs]) * (R(P.cospe[obs]) * nm.coslp + R(P.sinpe[obs]) * nm.sinlp));
      if (!(cosnorm > R(0.0))) return kPeelIdle;
      // I = 1; energy per sr ~ mu for isotropic intensity, and another
      // factor mu under Eddington limb darkening (newdisktherm.F90:349-355)
      S.w = P.limb ? (double)(R(12.0 / 5.0) * R(1.0) * (cosnorm + cosnorm * cosnorm))
                   : (double)(R(4.0) * R(1.0) * cosnorm);
    } else {
      S.w = (double)(R(1.0) * R(2.0));
    }
  } else {
    S.w = (double)k.u.s.I;
  }
  S.r = r;
  S.rec = ri;
  S.obs = obs;
  S.tau = 0.0;
  S.ncross = 0;
  S.err = 0;
  S.x = k.x;
  S.y = k.y;
  S.z = k.z;
  S.ir = k.ir;
  S.it = k.it;
  S.ip = k.ip;
  S.on_wall = (k.flags & 1) != 0;
  const double cospe = P.cospe[obs], sinpe = P.sinpe[obs];
  const double coste = P.coste[obs], sinte = P.sinte[obs];
  S.ux = sinte * cospe;
  S.uy = sinte * sinpe;
  S.uz = coste;
  if (S.on_wall) {
    const bool inner = fabs(k.rsq - W.r2[S.ir - 1]) < fabs(k.rsq - W.r2[S.ir]);
    const bool outwards = S.ux * S.x + S.uy * S.y + S.uz * S.z > 0.0;
    if (inner && !outwards) {
      if (S.ir == 1) return kPeelIdle;
      S.ir = S.ir - 1;
    }
    if (!inner && outwards) S.ir = S.ir + 1;
  }
  S.tsum = 0.0;
  return S.ir < P.nrg ? kPeelWalk : kPeelTally;
}

template <class R, int NA>
__device__ __forceinline__ int peel_step(const WallView &W, const PeelRec<R, NA> &k, PeelWalk &S) {
  const Params &P = c_P;
  // a straight ray meets each sphere and each cone at most twice and each
  // phi half-plane at most once: more crossings than that is a stuck ray
  const int max_cross = 2 * (P.nrg + P.ntg + P.npg) + 16;
  double dt;
  int face;
  if (S.x == 0.0 && S.y == 0.0) S.ip = axis_phi_cell(S.ux, S.uy, S.ip);
  // the cell's densities are loaded before its walls are computed, so that
  // the load's latency overlaps the roots (a walking ray is inside the grid);
  // held across the wavefront's steps instead, they cost more spills than
  // the longer overlap gains
  double dm[NA];
  {
    const double *dp = P.densm + cell_of(S.ir, S.it, S.ip) * (size_t)P.nact;
#pragma unroll
    for (int s = 0; s < NA; ++s) dm[s] = (s < P.nact) ? dp[s] : 0.0;
  }
  if (S.ncross > max_cross ||
      !next_wall(W, S.ux, S.uy, S.uz, S.x, S.y, S.z, S.ir, S.it, S.ip, S.on_wall, dt, face)) {
    S.err = 1;
    return kPeelIdle;
  }
  S.tsum = S.tsum + rec_opac(k, dm) * dt;
  if (S.tsum > kTauMax) {
    S.tau = S.tsum;
    return kPeelIdle;
  }
  face_step(face, S.ir, S.it, S.ip);
  S.x = S.x + dt * S.ux;
  S.y = S.y + dt * S.uy;
  S.z = S.z + dt * S.uz;
  S.on_wall = true;
  S.ncross++;
  if (S.ir == 0) {
    S.tau = S.tsum;
    return kPeelIdle;
  }
  if (S.ir < P.nrg) return kPeelWalk;
  S.tau = S.tsum;
  return kPeelTally;
}

template <class R, int NA>
__device__ __forceinline__ void peel_tally(const PeelRec<R, NA> &k, const PeelWalk &S) {
  const Params &P = c_P;
  const int peelid = S.obs;
  const bool direct = ((k.flags >> 9) & 1) != 0;  // Q = U = V = 0
  Stokes4<R> s{R(S.w), direct ? R(0.0) : k.u.s.Q, direct ? R(0.0) : k.u.s.U,
               direct ? R(0.0) : k.u.s.V};
  const int iflag = (k.flags >> 1) & 1, i_orig = (k.flags >> 2) & 7, slot = (k.flags >> 5) & 15;
  const double cospe = P.cospe[peelid], sinpe = P.sinpe[peelid];
  const double coste = P.coste[peelid], sinte = P.sinte[peelid];
  const R factor = m_exp(R(-S.tsum));
  s.I = s.I * factor;
  s.Q = s.Q * factor;
  s.U = s.U * factor;
  s.V = s.V * factor;
  if (iflag == 1)
    peel_weight<R>(k.d, R(coste), R(sinte), R(cospe), R(sinpe), hg_of<R>(slot, k.nub), s);
  const double yimage = S.z * sinte - S.y * coste * sinpe - S.x * coste * cospe;
  const double ximage = S.y * cospe - S.x * sinpe;
  const int ix = (int)((ximage + P.rmaxi) / P.fractxh) + 1;
  const int iy = (int)((yimage + P.rmaxi) / P.fractxh) + 1;
  const double rho2 = ximage * ximage + yimage * yimage;
  const int inub =
      clampi((int)((double)P.nfreq * (double)m_log(R(k.nub / P.numin)) / P.lnurat) + 1, 1, P.nfreq);
  const int i_inter = (iflag == 1) ? 2 : ((i_orig == 1) ? 1 : 3);
  const int indx[3] = {1, 1 + i_orig, 5 + i_inter};
  if (P.peel) {
    const size_t n = (size_t)P.nfreq * P.npeel * P.nap * kNo;
    for (int ia = 0; ia < P.nap; ++ia) {
      if (rho2 < P.aperture2[ia]) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const size_t q = (size_t)(inub - 1) +
                           (size_t)P.nfreq * ((size_t)peelid +
                                              (size_t)P.npeel * ((size_t)ia + (size_t)P.nap *
                                                                                  (size_t)(indx[j] - 1)));
          atomic_add(&P.peel[q], s.I);
          atomic_add(&P.peel[q + n], s.Q);
          atomic_add(&P.peel[q + 2 * n], s.U);
          atomic_add(&P.peel[q + 3 * n], s.V);
          atomic_add(&P.peel[q + 4 * n], s.I * s.I);
          atomic_add(&P.peel[q + 5 * n], s.Q * s.Q);
          atomic_add(&P.peel[q + 6 * n], s.U * s.U);
          atomic_add(&P.peel[q + 7 * n], s.V * s.V);
          atomic_add(&P.peel[q + 8 * n], 1.0);
        }
      }
    }
  }
  // peeloff_3d.f90:179-211: every filter whose node range holds the ray's
  // frequency strictly weights it by the response at the last node below
  // it (the reference scans up to it; the tables increase strictly, checked
  // on the host, so the bisection finds the same node)
  if (P.nfilt > 0 && ix <= P.nxhst && iy <= P.nxhst && ix > 0 && iy > 0) {
    const double wave = 1.2398 / (double)k.nub;
    const double freq = 2.9979e14 / wave;
    const size_t nx = (size_t)P.nxhst;
    const size_t nb = (size_t)P.npeel * nx * nx * (size_t)P.nfilt;
    for (int f = 0; f < P.nfilt; ++f) {
      const double *fw = P.filtwave + (size_t)P.nwfiltmax * f;
      const int nf = P.nwfilt[f] - 1;
      if (!(freq > fw[0] && freq < fw[nf - 1])) continue;
      int lo = 0, hi = nf - 1;  // fw[lo] < freq <= fw[hi]
      while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (freq > fw[mid])
          lo = mid;
        else
          hi = mid;
      }
      const double wt = P.filtphi[(size_t)P.nwfiltmax * f + lo];
      const size_t q = (size_t)peelid +
                       (size_t)P.npeel * ((size_t)(ix - 1) +
                                          nx * ((size_t)(iy - 1) + nx * (size_t)f));
      atomic_add(&P.imgb[q], (double)s.I * wt);
      atomic_add(&P.imgb[q + nb], (double)s.Q * wt);
      atomic_add(&P.imgb[q + 2 * nb], (double)s.U * wt);
      atomic_add(&P.imgb[q + 3 * nb], (double)s.V * wt);
    }
  }
  if (P.imcube && ix <= P.nxhst && iy <= P.nxhst && ix > 0 && iy > 0) {
    const size_t nx = (size_t)P.nxhst;
    const size_t q = (size_t)peelid +
                     (size_t)P.npeel * ((size_t)(ix - 1) + nx * ((size_t)(iy - 1) +
                                                                  nx * (size_t)(inub - 1)));
    const size_t n4 = (size_t)P.npeel * nx * nx * P.nfreq;
    if (P.imcube == 2) {  // the reference's real*4 image arrays
      float *ci = reinterpret_cast<float *>(P.img_i);
#pragma unroll
      for (int j = 0; j < 3; ++j) unsafeAtomicAdd(&ci[q + n4 * (size_t)(indx[j] - 1)], (float)s.I);
      unsafeAtomicAdd(&reinterpret_cast<float *>(P.img_q)[q], (float)s.Q);
      unsafeAtomicAdd(&reinterpret_cast<float *>(P.img_u)[q], (float)s.U);
      unsafeAtomicAdd(&reinterpret_cast<float *>(P.img_v)[q], (float)s.V);
    } else {
#pragma unroll
      for (int j = 0; j < 3; ++j) atomic_add(&P.img_i[q + n4 * (size_t)(indx[j] - 1)], s.I);
      atomic_add(&P.img_q[q], s.Q);
      atomic_add(&P.img_u[q], s.U);
      atomic_add(&P.img_v[q], s.V);
    }
  }
}

__constant__ double c_inczeta[146];

template <class R>
__device__ __forceinline__ R bbfreq(Mrg &rng, R teff) {  // random.f90:49-77
  const R z0 = R(c_inczeta[145]) * R(ran(rng));
  const R z1 = R(ran(rng)), z2 = R(ran(rng)), z3 = R(ran(rng)), z4 = R(ran(rng));
  int i = 1;
  while (z0 > R(c_inczeta[i])) ++i;
  return -m_log(z1 * z2 * z3 * z4) * teff / (R)i;
}

// lucydustfreq.f90:19-157 with its guard chain; j clamped to [1,nT-1].
// Split in two: ldf_begin draws the packet's number and issues the table
// loads, ldf_end consumes them, so that a caller can place independent
// arithmetic between the two while the loads are in flight.
template <class R>
struct LdfState {
  R T0, xi, f;
  const double *row;
  double a0, a1, e0, e1;
  int g0a, g1a, g0b, g1b;
  bool guided;
};

template <class R>
__device__ __forceinline__ LdfState<R> ldf_begin(double T0d, int id, Mrg &rng,
                                                 const double *table, const uint16_t *guide) {
  const Params &P = c_P;
  const int nT = P.nT, nnu = P.nnu;
  // lTint from LDS (staged by the transport kernel)
  const double *lT = reinterpret_cast<const double *>(hc_lds + P.emis_base);
  const double *kb = table + (size_t)nT * nnu * id;
  LdfState<R> L;
  L.T0 = R(T0d);
  const R lT0 = m_log(L.T0);
  L.xi = R(ran(rng));
  int j = (int)((R)(nT - 1) * (lT0 - R(lT[0])) / (R(lT[nT - 1]) - R(lT[0]))) + 1;
  j = clampi(j, 1, nT - 1);
  R f = (lT0 - R(lT[j - 1])) / (R(lT[j]) - R(lT[j - 1]));
  if (f < R(0.0)) f = R(0.0);
  if (f > R(1.0)) f = R(1.0);
  L.f = f;
  L.row = kb + (j - 1);  // KB(j, i) = row[(i-1)*nT], KB(j+1, i) = row[(i-1)*nT + 1]
  // the end nodes and the guide entries do not depend on each other: their
  // loads are issued together before any of them is needed
  const double *pe = L.row + (size_t)nT * (size_t)(nnu - 1);
  L.a0 = L.row[0];
  L.a1 = L.row[1];
  L.e0 = pe[0];
  L.e1 = pe[1];
  const int bq = clampi((int)(L.xi * R(kCdfGuide)), 0, kCdfGuide - 1);
  L.g0a = L.g1a = L.g0b = L.g1b = 0;
  L.guided = guide != nullptr;
  if (guide) {
    const uint16_t *g0 = guide + ((size_t)id * nT + (size_t)(j - 1)) * (kCdfGuide + 1);
    const uint16_t *g1 = g0 + (kCdfGuide + 1);
    L.g0a = g0[bq];
    L.g1a = g1[bq];
    L.g0b = g0[bq + 1];
    L.g1b = g1[bq + 1];
  }
  return L;
}

template <class R>
__device__ __forceinline__ R ldf_end(const LdfState<R> &L) {
  const Params &P = c_P;
  const int nT = P.nT, nnu = P.nnu;
  const double *nuint = reinterpret_cast<const double *>(hc_lds + P.emis_base) + nT;
  const R T0 = L.T0, xi = L.xi;
  R f = L.f;
  const double *row = L.row;
  auto KB = [&](int ii) __attribute__((always_inline)) {
    const double *p = row + (size_t)nT * (size_t)(ii - 1);
    return (R(1.0) - f) * R(p[0]) + f * R(p[1]);
  };
  const double a0 = L.a0, a1 = L.a1, e0 = L.e0, e1 = L.e1;
  const int g0a = L.g0a, g1a = L.g1a, g0b = L.g0b, g1b = L.g1b;
  int i0 = 1, i1 = nnu;
  R P0 = (R(1.0) - f) * R(a0) + f * R(a1);
  if (xi < P0) return (P0 > R(0.0)) ? T0 * xi * R(nuint[0]) / P0 : T0 * R(nuint[0]);
  R P1 = (R(1.0) - f) * R(e0) + f * R(e1);
  if (P1 > R(1.0)) P1 = R(1.0);
  if (xi >= P1) {
    if (P1 < R(1.0) && xi < R(1.0)) return T0 * R(nuint[nnu - 1]) * (R(1.0) - P1) / (R(1.0) - xi);
    return T0 * R(nuint[nnu - 1]);
  }
  // a short bracket's interior nodes pb+1 .. pb+4, read with its ends
  int pb = 0;
  R q1 = R(0.0), q2 = R(0.0), q3 = R(0.0), q4 = R(0.0);
  if (L.guided) {
    // The interpolated CDF lies between rows j and j+1, so the node counts
    // of both rows at the probability bin of xi bound the bracket; the
    // rows are strictly increasing (checked on the host), so the bracket is
    // the unique one the full bisection finds.
    const int lo = max(1, min(g0a, g1a));
    const int hi = min(nnu, max(g0b, g1b) + 1);
    const int a = lo > i0 ? lo : i0, b = hi < i1 ? hi : i1;
    // every node the bisection can visit lies inside (a, b): when there are
    // at most four, all loads are issued before any value is used
    const R Pa = KB(a), Pb = KB(b);
    if (b - a > 1 && b - a <= 5) {
      pb = a;
      q1 = KB(min(a + 1, b));
      q2 = KB(min(a + 2, b));
      q3 = KB(min(a + 3, b));
      q4 = KB(min(a + 4, b));
    }
    if (lo > i0) {
      i0 = lo;
      P0 = Pa;
      if (P0 > R(1.0)) P0 = R(1.0);
    }
    if (hi < i1) {
      i1 = hi;
      P1 = Pb;
      if (P1 > R(1.0)) P1 = R(1.0);
    }
  }
  while ((i1 - i0) > 1) {
    const int i = (i0 + i1) / 2;
    const int d = i - pb;
    R Pm = pb > 0 ? (d == 1 ? q1 : d == 2 ? q2 : d == 3 ? q3 : q4) : KB(i);
    if (Pm > R(1.0)) Pm = R(1.0);
    if (Pm < P0) {
      P0 = Pm;
      i0 = i;
    }
    if (Pm > P1) {
      P1 = Pm;
      i1 = 1;  // lucydustfreq.f90:109 as written
    }
    if (Pm > xi) {
      i1 = i;
      P1 = Pm;
    } else {
      i0 = i;
      P0 = Pm;
    }
  }
  if (xi < P0)
    f = R(0.0);
  else if (xi > P1)
    f = R(1.0);
  else if (P0 <= R(0.0))
    f = R(0.5);
  else if (P1 <= R(0.0) || xi <= R(0.0)) {
  } else if (!(P1 > P0))  // nodes that round to one value (FP32): no log(1) = 0 divisor
    f = R(0.0);
  else
    f = m_log(xi / P0) / m_log(P1 / P0);
  return T0 * R(nuint[i0 - 1]) * m_pow(R(nuint[i1 - 1]) / R(nuint[i0 - 1]), f);
}

template <class R>
__device__ __forceinline__ R lucydustfreq(double T0d, int id, Mrg &rng) {
  const Params &P = c_P;
  return ldf_end<R>(ldf_begin<R>(T0d, id, rng, P.kapBint, P.cdfguide));
}

// Planck-mean opacity at T (dust.f90:114-142) from the kappap table
__device__ __forceinline__ double kapp_at(const double *kappap, const double *lTint, int nT,
                                          double t, int id) {
  const double *kp = kappap + (size_t)nT * id;
  const double lt = hcm_log(t);
  const int k = (int)((double)(nT - 1) * (lt - lTint[0]) / (lTint[nT - 1] - lTint[0])) + 1;
  if (k < 1 || k > nT - 1) return k < 1 ? kp[0] : kp[nT - 1];
  const double f = (lt - lTint[k - 1]) / (lTint[k] - lTint[k - 1]);
  return (1.0 - f) * kp[k - 1] + f * kp[k];
}

// dusttemp.f90:6-58: the temperature at which na absorbed packets balance
// the emission of a cell-species of (rho V)^(1/4) = m4, by fixed-point
// iteration from T0 to 1e-5 (capped at 1000 sweeps where the reference
// would spin), clamped to [0.1, 1600] K
__device__ __forceinline__ double dusttemp_fp(double na, double T0, double ltot, double nphot,
                                              double kappav, const double *kappap,
                                              const double *lTint, int nT, double m4, int id) {
  double Td = (T0 != 0.0) ? T0 : 0.1 / 11605.0;
  double Tprev = 0.5 * Td;
  int iter = 0;
  while (fabs(Td - Tprev) > 1.e-5 * Td && iter < 1000) {
    Tprev = Td;
    Td = hcm_pow((na * ltot) / (4.0 * kSigT * nphot * kappav * kapp_at(kappap, lTint, nT, Td, id)),
                 0.25) /
         11605.0;
    if (m4 > 0.0)
      Td = Td / m4;
    else
      Td = 0.1 / 11605.0;
    if (Td > 1600.0 / 11605.0) Td = 1600.0 / 11605.0;
    Td = fmax(Td, 0.1 / 11605.0);
    ++iter;
  }
  return Td;
}

// Bjorkman-Wood branch of emit_common.f90:57-66 for an absorption of
// species id in tdust word w: the absorption is counted (propagate.f90:123),
// the cell-species temperature is brought up to date with its count so far
// (dusttemp), the walk's sublimation mask follows it, and the packet
// re-emits at the mean of the old and new temperatures (3/4 new) on the
// first absorption, at the new one after.  Packets running side by side
// take their counts in the order of the atomic; the stored temperature is
// the highest computed (it grows with the count), in serial mode the last.
__device__ __forceinline__ double bw_temperature(size_t w, int id) {
  const Params &P = c_P;
  const double na = atomicAdd(P.nabs + w, 1.0) + 1.0;
  const double told = __hip_atomic_load(P.tdust_w + w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  const double tnew = dusttemp_fp(na, told, P.ltot, P.nphot, P.kappav[id], P.kappap, P.lTint,
                                  P.nT, P.mass4[w], id);
  if (P.serial)
    P.tdust_w[w] = tnew;
  else
    atomicMax(reinterpret_cast<unsigned long long *>(P.tdust_w + w),
              (unsigned long long)__double_as_longlong(tnew));
  if (tnew >= P.tsub_ev) P.densm_w[w] = -P.dens[w];
  return na < 2.0 ? 0.25 * told + 0.75 * tnew : tnew;
}

// opacset for all active species: kapd = kappa * R*(cm) * kappa_V and the
// albedo.  launch_order picks newdisktherm's (kappa*rmincgs)*kappav or
// propagate's ((kappa*rstar)*rsol)*kappav grouping.
template <class R, int NA>
__device__ __forceinline__ void set_opacity(Pkt<R, NA> &k, bool launch_order) {
  const Params &P = c_P;
#pragma unroll
  for (int s = 0; s < NA; ++s) {
    if (s < P.nact) {
      const int id = P.act[s] - 1;
      const Opac1 op = opacset_species(s, (double)k.nub);
      k.kapd[s] = launch_order ? op.kappa * P.rmincgs * P.kappav[id]
                               : op.kappa * P.rstar * kRsol * P.kappav[id];
      k.alb[s] = R(op.albedo);
    }
  }
}

// darkening (samptabl.f90:29-66): the draw located in the cumulative
// probability of the emission cosine, mu interpolated linearly from the
// bin's upper node (the table's first node is mu = 0, probability 0).
__device__ __forceinline__ double darkening(double xi) {
  const Params &P = c_P;
  const double *xmu = P.limb_tab, *prob = P.limb_tab + P.ntab;
  const int n = clampi(locate(prob, P.ntab, xi), 1, P.ntab - 1);
  return xmu[n] - (prob[n] - xi) / (prob[n] - prob[n - 1]) * P.delmu;
}

// atmosfreq (starbarb.f90:120-156): the bracket hnuint(i) < xi <= hnuint(i+1)
// of locate2's bisection is the number i of nodes below the draw, found from
// the guide and a forward scan; then log-log interpolation, linear in the
// first bin (hnuint(1) = 0).
__device__ __forceinline__ double atmosfreq(Mrg &rng) {
  const Params &P = c_P;
  const int n = P.nhnu;
  const double *nu = P.atm, *h = P.atm + n, *lnu = P.atm + 2 * n, *lh = P.atm + 3 * n;
  const double xi = ran(rng);
  int i = P.atm_guide[clampi((int)(xi * (double)kAtmGuide), 0, kAtmGuide)];
  while (i < n && h[i] < xi) ++i;
  i = clampi(i, 1, n - 1);
  if (i > 1) {
    const double logatm = lnu[i - 1] + (lnu[i] - lnu[i - 1]) * (hcm_log10(xi) - lh[i - 1]) /
                                           (lh[i] - lh[i - 1]);
    return hcm_exp10(logatm);
  }
  return nu[0] + (nu[1] - nu[0]) * (xi - h[0]) / (h[1] - h[0]);
}

// The photon energy of stellar emission: the atmosphere table when one is
// set (the reference's stellar photons), else the Planck function at tstar.
template <class R>
__device__ __forceinline__ R star_freq(Mrg &rng) {
  return c_P.nhnu > 0 ? R(atmosfreq(rng)) : bbfreq<R>(rng, R(c_P.tstar));
}

template <class R, int NA>
__device__ __forceinline__ void initp_sample(Pkt<R, NA> &k, Mrg &rng) {  // initp_common.f90:1-29
  // isotropic intensity: mu = sqrt(xi); Eddington limb darkening: the table
  k.d.cost = c_P.limb ? R(darkening(ran(rng))) : m_sqrt(R(ran(rng)));
  k.d.sint = m_sqrt(R(1.0) - k.d.cost * k.d.cost);
  k.d.phi = R(kR2P) * R(ran(rng));
  k.d.cosp = m_cos(k.d.phi);
  k.d.sinp = m_sin(k.d.phi);
}

template <class R, int NA>
__device__ __forceinline__ void initp_transform(Pkt<R, NA> &k, R cosb, R sinb, R lp) {
  const R kpi = R(kPi), kr2p = R(kR2P);
  if (fabs(cosb) < R(0.9999999)) {  // initp_common.f90:31-94
    R costi = k.d.cost * cosb - k.d.sint * sinb * k.d.cosp;
    if (fabs(costi) > R(1.0)) {
      if (costi > R(1.0)) costi = R(1.0);
      if (k.d.cost < R(-1.0)) costi = R(-1.0);
    }
    const R sinti = m_sqrt(R(1.0) - costi * costi);
    R phinew;
    if (sinti > R(0.0000001)) {
      R cospn = (k.d.cost - costi * cosb) / (sinti * sinb);
      if (fabs(cospn) > R(1.0)) cospn = (cospn < R(0.0)) ? R(-1.0) : R(1.0);
      if (k.d.phi < kpi) {
        phinew = m_acos(cospn);
        if (phinew > kpi) phinew = kpi;
      } else {
        phinew = kr2p - m_acos(cospn);
        if (phinew < kpi) phinew = kpi;
      }
    } else {
      phinew = R(0.0);
    }
    k.d.phi = phinew;
    k.d.cost = costi;
    k.d.sint = sinti;
  } else if (cosb < R(0.0)) {
    k.d.cost = -k.d.cost;
  }
  k.d.phi = k.d.phi + lp;
  if (k.d.phi > kr2p) k.d.phi = k.d.phi - kr2p;
  if (k.d.phi < R(0.0)) k.d.phi = k.d.phi + kr2p;
  k.d.sinp = m_sin(k.d.phi);
  k.d.cosp = m_cos(k.d.phi);
}

template <class R>
__device__ __forceinline__ R clamp1(R x) {
  return x > R(1.0) ? R(1.0) : (x < R(-1.0) ? R(-1.0) : x);
}

// Walk direction from the packet's angles.
template <class R, int NA>
__device__ __forceinline__ void set_walk_dir(Pkt<R, NA> &k) {
  if constexpr (sizeof(R) == sizeof(double)) {
    k.ux = (double)(k.d.sint * k.d.cosp);
    k.uy = (double)(k.d.sint * k.d.sinp);
    k.uz = (double)k.d.cost;
  } else {
    // FP32 direction cosines are unit only to ~1e-7; the wall roots assume a
    // unit vector, so the walk direction is renormalised in FP64.
    const double ux = (double)k.d.sint * (double)k.d.cosp;
    const double uy = (double)k.d.sint * (double)k.d.sinp;
    const double uz = (double)k.d.cost;
    const double inv = 1.0 / sqrt(ux * ux + uy * uy + uz * uz);
    k.ux = ux * inv;
    k.uy = uy * inv;
    k.uz = uz * inv;
  }
}

template <class R>
__device__ __forceinline__ void bin_init(R nub, int i_orig_init) {
  const Params &P = c_P;
  if (!P.si_init) return;
  const int inub = clampi(
      (int)((double)P.nfreq * (double)m_log(R(nub / P.numin)) / P.lnurat) + 1, 1, P.nfreq);
  atomic_add(&P.si_init[(size_t)(inub - 1) + (size_t)P.nfreq * (size_t)(i_orig_init - 1)], 1.0);
  atomic_add(&P.si_init[(size_t)(inub - 1)], 1.0);
}

// theta and phi cells of the packet's position (FP64, as the oracle's
// star re-emission: cos(theta) = z / r clamped, phi = atan2(y, x) in [0, 2pi)).
template <class R, int NA>
__device__ __forceinline__ void locate_cell_angles(Pkt<R, NA> &k) {
  const Params &P = c_P;
  double cb = k.z / sqrt(k.rsq);
  if (cb > 1.0) cb = 1.0;
  if (cb < -1.0) cb = -1.0;
  double lp = hcm_atan2(k.y, k.x);
  if (lp < 0.0) lp = lp + kR2P;
  k.it = locate(P.thetarr, P.ntg, hcm_acos(cb));
  k.ip = locate(P.phiarr, P.npg, lp);
}

// Species masks of grid_mod.f90:15-20 (0-based ids).
__constant__ const int c_is_disk[8] = {1, 1, 0, 0, 1, 1, 0, 0};
__constant__ const int c_is_thermal_disk[8] = {1, 1, 0, 0, 0, 0, 0, 0};

// initpacc.f90:27-52: the emission point of a disk-accretion packet.  r from
// p(1/r) ~ 1 - sqrt(1/r) by rejection within (rddustmin, rmaxdmax), a
// uniform azimuth (rantheta, random.f90:28-47) and a Gaussian height of the
// mass-weighted scale height (gasdev, random.f90:140-159); redrawn outside
// rmax.  The reference's gasdev keeps the second deviate of a pair for its
// next call, across packets; a packet here owns its generator segment, so
// every call draws a fresh pair and returns the deviate returned first.
__device__ __forceinline__ void acc_position(Mrg &rng, double &x, double &y, double &z) {
  const Params &P = c_P;
  for (;;) {
    double rp = 0.0;
    while (rp >= P.rmaxdmax || rp <= P.rddustmin) {
      double winv = ran(rng);
      while (ran(rng) > 1.0 - sqrt(winv)) winv = ran(rng);
      rp = 1.0 / winv;
    }
    double r2 = 2.0, xx = 0.0, yy = 0.0;
    while (r2 > 1.0) {
      xx = ran(rng);
      xx = xx + xx - 1.0;
      yy = ran(rng);
      yy = yy + yy - 1.0;
      r2 = xx * xx + yy * yy;
    }
    const double rr = sqrt(r2);
    x = rp * (xx / rr);
    y = rp * (yy / rr);
    double zsum = 0.0;
#pragma unroll
    for (int id = 0; id < 8; ++id)
      if (c_is_disk[id]) zsum = zsum + P.fmass[id] * P.z1[id] * hcm_pow(rp, P.bflare[id]);
    double v1, v2, rsq;
    for (;;) {
      v1 = 2.0 * ran(rng) - 1.0;
      v2 = 2.0 * ran(rng) - 1.0;
      rsq = v1 * v1 + v2 * v2;
      if (rsq < 1.0 && rsq > 0.0) break;
    }
    z = zsum * (v2 * sqrt(-2.0 * hcm_log(rsq) / rsq));
    if (z * z + rp * rp < P.rmax * P.rmax) return;
  }
}

// ---------------------------------------------------------------------------
// Diffusion approximation (emit_common.f90:80-387, Lucy mode): a packet
// absorbed in (or launched into) a cell of the optically thick layer that
// gridset marks (gridset.f90:666-789) leaves from the layer's surface.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool dif_at(int ir, int it, int ip) {
  const Params &P = c_P;
  if (ir < 1 || ir > P.nrg || it < 1 || it > P.ntg) return false;
  return P.diffus[cell_of(ir, it, ip)] != 0;
}
__device__ __forceinline__ int dir_at(int ir, int it, int ip) {
  const Params &P = c_P;
  if (ir < 1 || ir > P.nrg || it < 1 || it > P.ntg) return 0;
  return P.diffdir[cell_of(ir, it, ip)];
}
__device__ __forceinline__ double pow4(double x) {  // x**4 as (x*x)*(x*x)
  x = x * x;
  return x * x;
}
// chiR (dust.f90:83-111): the Rosseland table at T, linear in ln T
__device__ __forceinline__ double chir(double T, int id) {
  const Params &P = c_P;
  const int nT = P.nT;
  const double *kr = P.kappar + (size_t)nT * id;
  const double lT = hcm_log(T);
  const int k = (int)((double)(nT - 1) * (lT - P.lTint[0]) / (P.lTint[nT - 1] - P.lTint[0])) + 1;
  if (k < 1 || k > nT - 1) return k < 1 ? kr[0] : kr[nT - 1];
  const double f = (lT - P.lTint[k - 1]) / (P.lTint[k] - P.lTint[k - 1]);
  return (1.0 - f) * kr[k - 1] + f * kr[k];
}
// isotrp.f90:1-36: a direction about s with density ~ |u.s| (rejection in
// the cube, reversed when against s)
__device__ __forceinline__ void isotrp(Mrg &rng, double sux, double suy, double suz, double &ux,
                                       double &uy, double &uz) {
  double x = 0.0, y = 0.0, z = 0.0, r2 = 2.0, mur = 0.0;
  while (r2 * r2 > fabs(mur)) {
    x = ran(rng);
    x = x + x - 1.0;
    y = ran(rng);
    y = y + y - 1.0;
    z = ran(rng);
    z = z + z - 1.0;
    mur = x * sux + y * suy + z * suz;
    r2 = x * x + y * y + z * z;
  }
  double r = sqrt(r2);
  if (mur < 0.0) r = -r;
  ux = x / r;
  uy = y / r;
  uz = z / r;
}
// The sublimation mask follows a temperature the layer raised.
__device__ __forceinline__ void dif_mask(size_t w, double t) {
  const Params &P = c_P;
  if (t >= P.tsub_ev) P.densm_w[w] = -P.dens[w];
}
// The surface cell's T^4 gains c, the temperature capped at sublimation;
// T4 returns the uncapped sum, which starts the column below.  Packets in
// flight apply the same step to the cell one after the other
// (compare-and-swap), so after n hits the temperature is the same n-fold
// step in any order.
__device__ __forceinline__ double dif_surface(size_t w, double c, double tsube, double &T4) {
  const Params &P = c_P;
  double *p = P.tdust_w + w;
  if (P.serial) {
    T4 = pow4(*p) + c;
    const double t = fmin(tsube, hcm_pow(T4, 0.25));
    *p = t;
    dif_mask(w, t);
    return t;
  }
  unsigned long long *pi = reinterpret_cast<unsigned long long *>(p);
  unsigned long long old = __hip_atomic_load(pi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  for (;;) {
    T4 = pow4(__longlong_as_double((long long)old)) + c;
    const double t = fmin(tsube, hcm_pow(T4, 0.25));
    const unsigned long long prev =
        atomicCAS(pi, old, (unsigned long long)__double_as_longlong(t));
    if (prev == old) {
      dif_mask(w, t);
      return t;
    }
    old = prev;
  }
}
// A cell behind the surface takes max(T, tnew).
__device__ __forceinline__ void dif_raise(size_t w, double tnew) {
  const Params &P = c_P;
  if (P.serial) {
    P.tdust_w[w] = fmax(tnew, P.tdust_w[w]);
  } else {
    atomicMax(reinterpret_cast<unsigned long long *>(P.tdust_w + w),
              (unsigned long long)__double_as_longlong(tnew));
  }
  dif_mask(w, tnew);
}

// emit_common.f90:80-387 for a packet in a layer cell (the oracle's
// emit_diffusion states each step): the packet moves to the layer's surface
// (radially inwards onto the inner wall, or in theta to 1.00001 cos of the
// surface cone), leaves about the surface normal with density ~ cos
// (isotrp), the surface cell of the two disk species gains
// 4 pi T*^4 / (nstar dA) in T^4 and the cells behind it are raised to the
// column's T^4 sum; the emitting species is drawn among the surface cell's
// disk species with weights w8 (per species id), and the packet's energy
// from the Planck function at the grey-atmosphere temperature.  Packet
// position, direction (k.ux.. = the isotrp vector) and cell are updated,
// the Stokes vector is kept.  Returns the emitting species id (0-based), or
// -1 (an unknown direction, or a surface cell without disk dust).
template <class R, int NA>
__device__ __forceinline__ int emit_diffusion(Pkt<R, NA> &k, Mrg &rng, const double (&w8)[2]) {
  const Params &P = c_P;
  const double tsube = 1600.0 / 11605.0;
  const double *ra = P.rarr;
  int ir = k.ir, it = k.it;
  const int ip = k.ip;
  const int dd = dir_at(ir, it, ip);
  double sux, suy, suz, ux, uy, uz, dA, nacc = 0.0;
  double T4s = P.tstar * P.tstar;
  T4s = T4s * T4s;
  const double rtot = sqrt(k.rsq);
  const int sl0 = P.slot_of[0], sl1 = P.slot_of[1];
  auto setdir = [&]() __attribute__((always_inline)) {
    k.ux = ux;
    k.uy = uy;
    k.uz = uz;
    const double cost = uz;
    const double sint = sqrt(1.0 - cost * cost);
    const double cosp = ux / sint, sinp = uy / sint;
    k.d.cost = R(cost);
    k.d.sint = R(sint);
    k.d.cosp = R(cosp);
    k.d.sinp = R(sinp);
    k.d.phi = R(hcm_atan2(sinp, cosp));
  };
  // the column sums of one species behind the surface cell (ir, it): radial
  // (dir -1) outwards, or in theta away from the surface (dir -2 / +2)
  auto column = [&](int id, int sl, int dir, double T4) __attribute__((always_inline)) {
    const double cc = 3.0 * kPi * T4s / (P.nstar * dA * 1.0);
    const double kvr = P.kappav[id];
    double T4d = T4;
    nacc = 0.0;
    if (dir == -1) {
      for (int i = ir; dir_at(i, it, ip) == -1 && dif_at(ir, it, ip); ++i)
        nacc = nacc + P.nabs[cell_of(i, it, ip) * (size_t)P.nact + sl];
      for (int i = ir + 1; dir_at(i, it, ip) == -1 && dif_at(ir, it, ip); ++i) {
        const size_t km = cell_of(i - 1, it, ip) * (size_t)P.nact + sl;
        T4d = T4d + cc * nacc * kvr * kRsol * P.rstar * P.dens[km] * chir(P.tdust_w[km], id) *
                        (ra[i - 1] - ra[i - 2]);
        dif_raise(cell_of(i, it, ip) * (size_t)P.nact + sl, fmin(tsube, hcm_pow(T4d, 0.25)));
        nacc = nacc - P.nabs[km];
      }
    } else if (dir == -2) {
      for (int i = it; dir_at(ir, i, ip) == -2 && dif_at(ir, it, ip); ++i)
        nacc = nacc + P.nabs[cell_of(ir, i, ip) * (size_t)P.nact + sl];
      for (int i = it + 1; dir_at(ir, i, ip) == -2 && dif_at(ir, it, ip); ++i) {
        const size_t km = cell_of(ir, i - 1, ip) * (size_t)P.nact + sl;
        T4d = T4d + cc * nacc * kvr * kRsol * P.rstar * P.dens[km] * chir(P.tdust_w[km], id) *
                        0.5 * (ra[ir - 1] + ra[ir]) * (P.thetarr[i - 1] - P.thetarr[i - 2]);
        dif_raise(cell_of(ir, i, ip) * (size_t)P.nact + sl, fmin(tsube, hcm_pow(T4d, 0.25)));
        nacc = nacc - P.nabs[km];
      }
    } else {
      const double cc2 = 3.0 * kPi * T4s / (P.nstar * dA);
      for (int i = it; dir_at(ir, i, ip) == 2 && dif_at(ir, it, ip); --i)
        nacc = nacc + P.nabs[cell_of(ir, i, ip) * (size_t)P.nact + sl];
      for (int i = it - 1; dir_at(ir, i, ip) == 2 && dif_at(ir, it, ip); --i) {
        const size_t kp = cell_of(ir, i + 1, ip) * (size_t)P.nact + sl;
        T4d = T4d + cc2 * nacc * kvr * kRsol * P.rstar * P.dens[kp] * chir(P.tdust_w[kp], id) *
                        0.5 * (ra[ir - 1] + ra[ir]) * (P.thetarr[i + 1] - P.thetarr[i]);
        dif_raise(cell_of(ir, i, ip) * (size_t)P.nact + sl, fmin(tsube, hcm_pow(T4d, 0.25)));
        nacc = nacc - P.nabs[kp];
      }
    }
  };
  // the emitting species and the packet's energy (emit_common.f90:185-205)
  auto finish = [&](double t0, double t1, double dA_) __attribute__((always_inline)) {
    const size_t wc = cell_of(ir, it, ip) * (size_t)P.nact;
    const double xran = ran(rng);
    const double d0 = w8[0] * P.dens[wc + sl0], d1 = w8[1] * P.dens[wc + sl1];
    const double opac = d0 + d1;
    int id2 = -1;
    if (opac != 0.0) {
      if (xran <= d0 / opac + (double)1.e-10f)
        id2 = 0;
      else if (xran <= d0 / opac + d1 / opac + (double)1.e-10f)
        id2 = 1;
    }
    if (id2 < 0) return -1;
    const double T4 = pow4(id2 == 0 ? t0 : t1);
    const double T4eff = 4.0 * kPi * T4s * nacc / (dA_ * P.nstar * 1.0);
    const double mu = ux * sux + uy * suy + uz * suz;
    const double T = fmin(hcm_pow(T4 + 0.75 * mu * T4eff, 0.25), tsube);
    k.nub = bbfreq<R>(rng, R(T));
    return id2;
  };
  if (dd == -1) {
    while (dir_at(ir, it, ip) == -1 && dif_at(ir, it, ip)) --ir;
    ++ir;
    const double rnew = ra[ir - 1];
    k.on_wall = true;
    k.x = k.x * rnew / rtot;
    k.y = k.y * rnew / rtot;
    k.z = k.z * rnew / rtot;
    k.rsq = rnew * rnew;
    sux = -k.x / rnew;
    suy = -k.y / rnew;
    suz = -k.z / rnew;
    isotrp(rng, sux, suy, suz, ux, uy, uz);
    setdir();
    dA = 2.0 * kPi * (ra[ir - 1] * ra[ir - 1]) * (P.costarr[it - 1] - P.costarr[it]);
    const double c = 4.0 * kPi * T4s / (P.nstar * dA * 1.0);
    const size_t wc = cell_of(ir, it, ip) * (size_t)P.nact;
    double T4;
    const double t0 = dif_surface(wc + sl0, c, tsube, T4);
    column(0, sl0, -1, T4);
    const double t1 = dif_surface(wc + sl1, c, tsube, T4);
    column(1, sl1, -1, T4);
    const int id2 = finish(t0, t1, dA);
    k.ir = ir - 1;
    return id2;
  }
  if (dd == -2 || dd == 2) {
    if (dd == -2) {
      while (dif_at(ir, it, ip)) --it;
      ++it;
    } else {
      while (dif_at(ir, it, ip)) ++it;
    }
    const double cosb = k.z / rtot;
    const double sinb = sqrt(1.0 - cosb * cosb);
    const double cosbnew = 1.00001 * P.costarr[it - 1];
    const double sinbnew = sqrt(1.0 - cosbnew * cosbnew);
    k.x = k.x * sinbnew / sinb;
    k.y = k.y * sinbnew / sinb;
    k.z = rtot * cosbnew;
    const double mu = P.costarr[it - 1], sth = P.sintarr[it - 1];
    if (mu > 0.0) {
      sux = -(mu * k.x) / sth / rtot;
      suy = -(mu * k.y) / sth / rtot;
      suz = sth;
    } else {
      sux = mu * k.x / sth / rtot;
      suy = mu * k.y / sth / rtot;
      suz = -sth;
    }
    isotrp(rng, sux, suy, suz, ux, uy, uz);
    setdir();
    if (P.costarr[it - 1] < 0.0) --it;
    dA = kPi * sth * (ra[ir] * ra[ir] - ra[ir - 1] * ra[ir - 1]);
    const double c = 4.0 * kPi * T4s / (P.nstar * dA * 1.0);
    const size_t wc = cell_of(ir, it, ip) * (size_t)P.nact;
    const int dir = dir_at(ir, it, ip) == -2 ? -2 : 2;
    double T4;
    const double t0 = dif_surface(wc + sl0, c, tsube, T4);
    column(0, sl0, dir, T4);
    const double t1 = dif_surface(wc + sl1, c, tsube, T4);
    column(1, sl1, dir, T4);
    const int id2 = finish(t0, t1, dA);
    it = dir_at(ir, it, ip) == -2 ? it - 1 : it + 1;
    k.it = it;
    return id2;
  }
  return -1;
}

// Disk-accretion launch: initpacc.f90:3-122 without the diffusion layer,
// then emit_common.f90:25-78 (Lucy re-emission at the cell's temperature) or
// emit_gasacc.f90:1-44 (Pringle 1981 gas emission where the cell has no
// thermal disk dust), then newdisktherm.F90:219-250.  The reference picks
// the emitting species with the kapd the previous packet left (select_dust
// reads the module global); packets here are independent, so the thermal
// disk species are weighted by kappa_V(id) rho(id), which is that choice
// whenever the species share one dust table.  A packet born in a cell of
// the diffusion layer leaves from the layer's surface (emit_diffusion).
// Returns the emitting slot (0 for gas emission), or -1 for a layer cell
// without disk dust (where the reference stops).
template <class R, int NA>
__device__ __forceinline__ int launch_acc(const WallView &W, Pkt<R, NA> &k, Mrg &rng) {
  const Params &P = c_P;
  double x, y, z;
  acc_position(rng, x, y, z);
  k.x = x;
  k.y = y;
  k.z = z;
  k.s = Stokes4<R>{R(1.0), R(0.0), R(0.0), R(0.0)};
  k.rsq = x * x + y * y + z * z;
  const double rtot = sqrt(k.rsq);
  k.ir = locate(W.r2, P.nrg, k.rsq);
  k.it = locate(P.costarr, P.ntg, z / rtot);
  double lp = hcm_atan2(y, x);
  if (lp < 0.0) lp = lp + kR2P;
  k.ip = locate(P.phiarr, P.npg, lp);
  const size_t cell = cell_of(k.ir, k.it, k.ip);
  const double *dn = P.densm + cell * (size_t)P.nact;  // signed: |.| is the density
  double rho[8];
  double dsum = 0.0;
#pragma unroll
  for (int id = 0; id < 8; ++id) {
    const int sl = P.slot_of[id];
    rho[id] = (c_is_thermal_disk[id] && sl >= 0) ? fabs(dn[sl]) : 0.0;
    if (c_is_thermal_disk[id]) dsum = dsum + rho[id];
  }
  int idust2 = -1;
  if (dsum > 0.0) {  // select_dust(is_thermal_disk), select_dust.f90:17-42
    const double xran = ran(rng);
    double opac = 0.0, frac = 0.0;
#pragma unroll
    for (int id = 0; id < 8; ++id)
      if (c_is_thermal_disk[id]) opac = opac + P.kappav[id] * rho[id];
#pragma unroll
    for (int id = 0; id < 8; ++id) {
      if (c_is_thermal_disk[id] && idust2 < 0) {
        frac = frac + P.kappav[id] * rho[id] / opac;
        if (xran <= frac + (double)1.e-10f) idust2 = id;
      }
    }
  }
  int slot = 0;
  k.on_wall = false;
  if (P.diffus != nullptr && P.diffus[cell] != 0) {
    // initpacc.f90:95-112: the absorption shared by the two disk species in
    // proportion to kappa_V rho (kapd restated as above), then the diffusion
    // branch of emit_common from the layer's surface
    if (idust2 >= 0) atomicAdd(&P.dbig[cell], 1);
    const double op2 = P.kappav[0] * rho[0] + P.kappav[1] * rho[1];
    if (op2 == 0.0) return -1;
    atomic_add(&P.nabs[cell * (size_t)P.nact + P.slot_of[0]], P.kappav[0] * rho[0] / op2);
    atomic_add(&P.nabs[cell * (size_t)P.nact + P.slot_of[1]], P.kappav[1] * rho[1] / op2);
    const double w8[2] = {P.kappav[0], P.kappav[1]};
    const int id2 = emit_diffusion(k, rng, w8);
    if (id2 < 0) return -1;
    bin_init(k.nub, 4);  // newdisktherm.F90:219-250
    set_opacity(k, true);
    k.iflag = 0;
    k.i_orig = 2;
    return P.slot_of[id2];
  }
  if (idust2 >= 0) {  // initpacc.f90:95-115, then the Lucy re-emission
    slot = P.slot_of[idust2];
    const size_t w = cell * (size_t)P.nact + slot;
    atomicAdd(&P.dbig[cell], 1);
    double tem;
    if (P.bw) {
      tem = bw_temperature(w, idust2);
    } else {
      atomic_add(&P.nabs[w], 1.0);
      tem = P.tdust[w];
    }
    k.d.cost = R(1.0) - R(2.0) * R(ran(rng));
    k.d.sint = m_sqrt(R(1.0) - m_pow(k.d.cost, R(2.0)));
    k.d.phi = R(kR2P) * R(ran(rng));
    k.d.cosp = m_cos(k.d.phi);
    k.d.sinp = m_sin(k.d.phi);
    if (P.bw && tem >= P.tsub_ev)  // dustfreq: the Planck function at sublimation
      k.nub = bbfreq<R>(rng, R(tem));
    else
      k.nub = ldf_end<R>(ldf_begin<R>(tem, idust2, rng, P.bw ? P.kapdBint : P.kapBint,
                                      P.bw ? P.cdfguide_bw : P.cdfguide));
  } else {  // emit_gasacc.f90:22-42
    const double t0 = 3.0 * kGn * (kMsol * P.massc) * P.mdot /
                      (8.0 * kPi * kSigT * hcm_pow(rtot * P.rstar * kRsol, 3.0));
    const double tgas = hcm_pow(t0 * (1.0 - sqrt(1.0 / rtot)), 0.25) / 11605.0;
    k.nub = bbfreq<R>(rng, R(tgas));
    k.d.cost = R(1.0) - R(2.0) * R(ran(rng));
    k.d.sint = m_sqrt(R(1.0) - m_pow(k.d.cost, R(2.0)));
    k.d.phi = R(kR2P) * R(ran(rng));
    k.d.cosp = m_cos(k.d.phi);
    k.d.sinp = m_sin(k.d.phi);
  }
  set_walk_dir(k);
  bin_init(k.nub, 4);  // newdisktherm.F90:219-250
  set_opacity(k, true);
  k.iflag = 0;
  k.i_orig = 2;
  return slot;
}

// Star / hot-spot launch (initpspot.f90:3-175 or initp_ps.f90:3-79, then
// newdisktherm.F90:256-377).
// The launch point is FP64 on both paths (cos b, sin b and the longitude lp
// from the FP64 draws, the position from FP64 sin/cos), as the positions of
// the walk are: an FP32 longitude is quantised to 2^-24 of a turn, and the
// native FP32 sine of a longitude that lands on a multiple of pi/2 is exactly
// zero, which put such a packet on a phi wall and on the stellar surface at
// once -- two walls that the walk's single on-wall flag cannot both skip, so
// it stepped between the two phi cells without moving (about one packet in
// 10^6).  The direction frame and the event arithmetic stay in R.
template <class R, int NA>
__device__ __forceinline__ void launch_star(Pkt<R, NA> &k, Mrg &rng, bool hotspot, R *nrm) {
  const Params &P = c_P;
  const R kpi = R(kPi), kr2p = R(kR2P);
  double cosb_d, sinb_d, lp_d;
  R cosb, sinb, lp;
  int ispotph = hotspot ? 1 : 0;
  k.s = Stokes4<R>{R(1.0), R(0.0), R(0.0), R(0.0)};
  if (!P.ips) {
    // a spot point by rejection (initpspot.f90:34-60); after kMaxTrials the
    // last point is kept, a bound that a spot of any area reached by
    // hc_set_sources' checks does not meet in practice
    for (uint32_t ntry = 1;; ++ntry) {
      for (;;) {
        cosb_d = 1.0 - 2.0 * ran(rng);
        sinb_d = sqrt(1.0 - cosb_d * cosb_d);
        if (cosb_d != 0.0 && fabs(cosb_d) < 1.0) break;
      }
      lp_d = kR2P * ran(rng);
      cosb = R(cosb_d);
      sinb = R(sinb_d);
      lp = R(lp_d);
      if (P.ispot == 1 && ispotph == 1) {
        const R cphi = m_cos(lp), sphi = m_sin(lp);
        const R d1 = R(P.scspot1) * sinb * cphi + R(P.ssspot1) * sinb * sphi + R(P.cspot1) * cosb;
        const R d2 = R(P.scspot2) * sinb * cphi + R(P.ssspot2) * sinb * sphi + R(P.cspot2) * cosb;
        const R cs = R(P.csprad);
        if ((P.nspot == 1 ? (d1 < cs) : (d1 < cs && d2 < cs)) && ntry < kMaxTrials) continue;
      }
      break;
    }
    if (P.ispot == 1 && P.fspot < 1.0) {
      R dellon = (lp < kpi) ? lp : kr2p - lp;
      if (R(P.cspot1) * cosb + R(P.sspot1) * sinb * m_cos(dellon) > R(P.csprad)) ispotph = 1;
      if (P.nspot == 2) {
        dellon = (lp < kpi) ? kpi - lp : lp - kpi;
        if (R(P.cspot2) * cosb + R(P.sspot2) * sinb * m_cos(dellon) > R(P.csprad)) ispotph = 1;
      }
    }
    initp_sample(k, rng);
    initp_transform(k, cosb, sinb, lp);
  } else {
    for (;;) {
      cosb_d = clamp1(1.0 - 2.0 * ran(rng));
      sinb_d = clamp1(sqrt(1.0 - cosb_d * cosb_d));
      if (cosb_d != 0.0) break;
    }
    cosb = R(cosb_d);
    sinb = R(sinb_d);
    k.d.cost = cosb;
    k.d.sint = sinb;
    lp_d = kR2P * ran(rng);
    lp = R(lp_d);
    k.d.phi = lp;
    k.d.sinp = clamp1(m_sin(k.d.phi));
    k.d.cosp = clamp1(m_cos(k.d.phi));
  }
  // initpspot.f90:147-159 / initp_ps.f90:59-69: hot-spot photons
  // (iplanckst = 1) take the Planck function at tshock inside a spot and at
  // tstar outside it; stellar photons (iplanckst = 0) take the stellar
  // spectrum outside a spot (everywhere when fspot = 1) and tshock's Planck
  // function inside one
  if (hotspot)
    k.nub = bbfreq<R>(rng, R(((P.fspot == 1.0 && !P.ips) || ispotph) ? P.tshock : P.tstar));
  else if (ispotph && (P.ips || P.fspot != 1.0))
    k.nub = bbfreq<R>(rng, R(P.tshock));
  else
    k.nub = star_freq<R>(rng);
  if (hotspot) {
    if (ran(rng) < 0.5)
      k.nub = R(1.2398 / 0.06) + R(ran(rng)) * R(1.2398 / 0.015 - 1.2398 / 0.06);
    bin_init(k.nub, 3);
  } else {
    bin_init(k.nub, 2);
  }
  k.i_orig = 1;
  set_opacity(k, true);
  const double coslp = hcm_cos(lp_d), sinlp = hcm_sin(lp_d);
  k.z = cosb_d * 1.0;
  const double rp = sinb_d * 1.0;
  k.x = rp * coslp;
  k.y = rp * sinlp;
  k.rsq = k.z * k.z + rp * rp;
  set_walk_dir(k);
  k.iflag = 0;
  k.ir = 1;
  k.on_wall = true;
  k.it = locate(P.thetarr, P.ntg, hcm_acos(cosb_d));
  k.ip = locate(P.phiarr, P.npg, lp_d);
  nrm[0] = cosb;
  nrm[1] = sinb;
  nrm[2] = R(coslp);
  nrm[3] = R(sinlp);
}

// Star re-emission of a packet that reached the stellar surface.
template <class R, int NA>
__device__ __forceinline__ void star_reemit(Pkt<R, NA> &k, Mrg &rng, Acc &acc, R *nrm) {
  cold_inc(cold().aflux);
  const R cosb = clamp1(R(k.z / sqrt(k.rsq)));
  const R sinb = m_sqrt(R(1.0) - cosb * cosb);
  R lp = m_atan2(R(k.y), R(k.x));
  if (lp < R(0.0)) lp = lp + R(kR2P);
  k.s = Stokes4<R>{R(1.0), R(0.0), R(0.0), R(0.0)};
  initp_sample(k, rng);
  initp_transform(k, cosb, sinb, lp);
  k.nub = star_freq<R>(rng);
  k.i_orig = 1;
  k.iflag = 0;
  set_opacity(k, true);
  set_walk_dir(k);
  k.ir = 1;
  k.on_wall = true;
  // the impact point can lie a rounding error outside the theta/phi cell the
  // walk arrived in (where a cone or phi wall meets the stellar surface); the
  // walk cannot find a wall it is already past, so the cell is located anew
  // from the point itself
  locate_cell_angles(k);
  nrm[0] = cosb;
  nrm[1] = sinb;
  nrm[2] = m_cos(lp);
  nrm[3] = m_sin(lp);
}

// select_dust.f90 over the active species (zero-density slots add nothing),
// from the cell's densities already in registers (|signed density|).
template <class R, int NA>
__device__ __forceinline__ int select_dust(const Pkt<R, NA> &k, const double (&dm)[NA], Mrg &rng) {
  const Params &P = c_P;
  const double xran = ran(rng);
  double opac = 0.0, frac = 0.0;
#pragma unroll
  for (int s = 0; s < NA; ++s)
    if (s < P.nact) opac = opac + k.kapd[s] * fabs(dm[s]);
  if (opac == 0.0) return -1;
  const double slack = (double)1.e-10f;
#pragma unroll
  for (int s = 0; s < NA; ++s) {
    if (s < P.nact) {
      frac = frac + k.kapd[s] * fabs(dm[s]) / opac;
      if (xran <= frac + slack) return s;
    }
  }
  return -1;
}

template <class R, int NA>
__device__ __forceinline__ void bin_exit(Pkt<R, NA> &k, double xpold, double ypold, double zpold, int iscat,
                         Acc &acc) {
  const Params &P = c_P;
  int it = (int)((double)P.nmu * (1.0 - (double)k.d.cost) / 2.0) + 1;
  it = clampi(it, 1, P.nmu);
  if (k.d.phi < R(0.0)) k.d.phi = k.d.phi + R(kR2P);
  const int ip = clampi((int)((double)P.nph * (double)k.d.phi / kR2P) + 1, 1, P.nph);
  const double ut = P.u[it - 1];
  const double cp = (double)k.d.cosp, sp = (double)k.d.sinp;
  const double x = ypold * cp - xpold * sp;
  const double y = zpold * P.sini[it - 1] - ypold * ut * sp - xpold * ut * cp;
  const double rho2 = (x * x + y * y);
  if (rho2 < P.aperture2[P.nap - 1] * 1.0001) {
    cold_inc(cold().flux);
    if (k.iflag == 1) {
      cold_inc(cold().sflux);
      acc.nscat += (uint32_t)iscat;
    }
  }
  const int inub =
      clampi((int)((double)P.nfreq * (double)m_log(R(k.nub / P.numin)) / P.lnurat) + 1, 1, P.nfreq);
  const int i_inter = (k.iflag == 1) ? 2 : ((k.i_orig == 1) ? 1 : 3);
  const int indx[3] = {1, 1 + k.i_orig, 5 + i_inter};
  const bool counted = !(k.iflag == 0 && k.i_orig == 4);
  if (P.sed && counted) {
    const size_t n = (size_t)P.nfreq * P.nmu * P.nph * P.nap * kNo;
    for (int ia = 0; ia < P.nap; ++ia) {
      if (rho2 < P.aperture2[ia]) {
#pragma unroll
        for (int j = 0; j < 3; ++j) {
          const size_t q =
              (size_t)(inub - 1) +
              (size_t)P.nfreq *
                  ((size_t)(it - 1) +
                   (size_t)P.nmu * ((size_t)(ip - 1) +
                                    (size_t)P.nph * ((size_t)ia + (size_t)P.nap *
                                                                      (size_t)(indx[j] - 1))));
          const double sI = k.s.I, sQ = k.s.Q, sU = k.s.U, sV = k.s.V;
          atomic_add(&P.sed[q], sI);
          atomic_add(&P.sed[q + n], sQ);
          atomic_add(&P.sed[q + 2 * n], sU);
          atomic_add(&P.sed[q + 3 * n], sV);
          atomic_add(&P.sed[q + 4 * n], sI * sI);
          atomic_add(&P.sed[q + 5 * n], sQ * sQ);
          atomic_add(&P.sed[q + 6 * n], sU * sU);
          atomic_add(&P.sed[q + 7 * n], sV * sV);
          atomic_add(&P.sed[q + 8 * n], 1.0);
        }
      }
    }
  }
  if (counted) atomic_add(&cold().sumsub, (double)k.s.I);
  atomic_add(&cold().sumall, (double)k.s.I);
  if (k.iflag == 1) cold_inc(cold().scount);
}

// Optical depth to the next event, -ln(xi) (propagate.f90:33).  In FP32 a
// draw within 2^-25 of 1 rounds to 1.0f and would give tau = 0, which the
// walk cannot place inside an empty cell (0/0); such draws take the
// smallest FP32 step instead.  The FP64 draw is never 1.
template <class R>
__device__ __forceinline__ double draw_tau(Mrg &rng) {
  const double t = (double)(-m_log(R(ran(rng))));
  if constexpr (sizeof(R) == sizeof(float)) return t > 0.0 ? t : 5.9604644775390625e-08;
  return t;
}

// One lane's packet between events.
template <class R, int NA>
struct Lane {
  Pkt<R, NA> k;
  Mrg rng;
  LucyAcc<R, NA> la;
  double xpold, ypold, zpold;
  long long pid;
  int iscat, iint, pir, pit, pip;
  uint32_t work;  // events + crossings of the packet so far
  uint32_t qpos;  // pool position of the packet's pending scattering or absorption
  double tdust;   // T of the absorbing species in the cell (prefetched for re-emission)
  bool first;
};

__device__ __forceinline__ double wave_sum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Peel records per wavefront queue (Params::peel_q, sized on the host to the
// LDS left after the wall arrays).  An iteration adds at most one record per
// lane, so a queue flushed whenever it holds more than peel_q - 64 records
// never overflows.

// Walk every queued (record, observer) pair of this wavefront: lanes take
// consecutive rays, so the observers of one record run side by side.
// The wavefront walks the rays of its peel queue with lane refill: a lane
// whose ray ends takes the next (record, observer) pair, so the wavefront
// steps while rays remain instead of waiting, batch by batch, for the
// longest ray of 64.  Rays that reach the grid's edge wait for their tally
// until kPeelTallyMin lanes do (or nothing walks), so the tally code also
// runs for many lanes at once.
constexpr int kPeelTallyMin = 16;

template <class R, int NA, bool TRACE>
__device__ __forceinline__ void flush_peels(const WallView &W, const PeelRec<R, NA> *q, int n,
                                            Acc &acc) {
  const Params &P = c_P;
  const int lane = threadIdx.x & 63;
  const unsigned long long below = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  const int total = n * P.npeel;
  int base = 0;  // next pair to hand out (the same on every lane)
  PeelWalk S;
  S.r = -1;
  int st = kPeelIdle;
  auto ray_done = [&]() __attribute__((always_inline)) {
    acc.npeelray++;
    acc.npeelcross += (uint32_t)S.ncross;
    acc.err_wall += (uint32_t)S.err;
    if constexpr (TRACE) {
      const long long pid = q[S.rec].pid;
      atomicAdd(reinterpret_cast<unsigned long long *>(&P.trace[pid].npeelcross),
                (unsigned long long)S.ncross);
      atomic_add(&P.trace[pid].tau_peel, S.err ? 0.0 : S.tau);
    }
    S.r = -1;
  };
  for (;;) {
    const unsigned long long idle = __ballot(st == kPeelIdle);
    if (st == kPeelIdle) {
      const int r = base + __popcll(idle & below);
      if (r < total) {
        st = peel_begin(W, q, r, S);
        if (st == kPeelIdle && S.r >= 0) ray_done();  // ended before its first crossing
      }
    }
    base += __popcll(idle);
    const unsigned long long walking = __ballot(st == kPeelWalk);
    const unsigned long long tally = __ballot(st == kPeelTally);
    if (walking == 0ull && tally == 0ull) {
      if (base >= total) break;
      continue;
    }
    if (tally != 0ull && (__popcll(tally) >= kPeelTallyMin || walking == 0ull)) {
      if (st == kPeelTally) {
        peel_tally(q[S.rec], S);
        ray_done();
        st = kPeelIdle;
      }
      continue;
    }
    if (st == kPeelWalk) {
      st = peel_step(W, q[S.rec], S);
      if (st == kPeelIdle) ray_done();
    }
  }
}

// Event loop of the pass (propagate.f90:1-325 around newdisktherm.F90:194-449)
// with path regeneration and event sorting.  Each lane holds one packet and
// is in one state: walking between events, or waiting for one kind of event
// (scattering, absorption + re-emission, star re-emission, end of the packet
// and launch of the next).  A wavefront alternates two phases:
//  - walk phase: walking lanes cross one cell wall per step.  A lane whose
//    walk reaches an interaction draws, in the reference's order, the dust
//    species and the scatter/absorb test right away, so it waits for the
//    event kind it will perform.  The phase ends as soon as ev_min lanes wait
//    for one kind, or nobody walks;
//  - event phase: every event kind with at least ev_min waiting lanes runs
//    (all kinds when nobody walks), so each event branch executes for many
//    lanes at once; peel-off records queued by the events are walked once the
//    wavefront's queue fills.
// A packet's history and its random numbers are those of the serial loop.
#ifndef HC_BLOCK_LUCY
#define HC_BLOCK_LUCY 256  // lanes per workgroup of the passes without peel-off
#endif
// Peeled passes: four wavefronts (each with its peel queue) per workgroup, one
// workgroup per compute unit, so one wavefront per SIMD.  Compiled for that
// shape the kernel keeps its state in the SIMD's whole register file (VGPRs
// and AGPRs) instead of spilling to scratch at the two-wavefront budget; on
// the 200x199x200-cell model's peeled pass that ran the bulk ~9 % faster than
// the same kernel built for 512-lane workgroups.
constexpr int kBlockPeel = 256;
// Event pool of the pooled Lucy passes (512-lane workgroups, one per compute
// unit): a request record of kPoolFields doubles per lane.
constexpr int kPoolLanes = 512;
constexpr int kPoolFields = 12;
constexpr uint32_t kPoolTagMask = 0x7fffffu;  // queue entry: (position & mask) << 9 | lane
static_assert(kPoolLanes == 512, "queue entries hold 9-bit lane numbers");
constexpr int kBlockLucy = HC_BLOCK_LUCY;

template <class R, int NA, bool TRACE, bool PEEL, bool POOL>
__global__ __launch_bounds__(PEEL ? kBlockPeel : POOL ? kPoolLanes : kBlockLucy,
                             (PEEL || POOL || TRACE || NA > 4) ? 1 : 2) void transport(
    Params, const Params *, const uint32_t * /* read through kernel_params() */) {
  static_assert(!(PEEL && POOL), "the event pool shares the LDS with no peel-off queues");
  const Params &P = c_P;
  // packet p draws from generator segment seg_base + p
  const unsigned long long seg_base = P.seg_dev ? *P.seg_dev : P.seg0;
  const int wave = threadIdx.x >> 6;
  unsigned char *const lds = hc_lds;
  // wall arrays into LDS: r2[nrg] t2[ntg] a[npg] b[npg]
  double *wl = reinterpret_cast<double *>(lds);
  for (int i = threadIdx.x; i < P.nrg; i += blockDim.x) wl[i] = P.r2arr[i];
  for (int i = threadIdx.x; i < P.ntg; i += blockDim.x) wl[P.nrg + i] = P.tan2arr[i];
  for (int i = threadIdx.x; i < P.npg; i += blockDim.x) {
    wl[P.nrg + P.ntg + i] = P.aarr[i];
    wl[P.nrg + P.ntg + P.npg + i] = P.barr[i];
  }
  // radii and the theta walls' |cos t|, sin t (the midplane: 0, 1)
  double *wr = wl + P.nrg + P.ntg + 2 * P.npg;
  for (int i = threadIdx.x; i < P.nrg; i += blockDim.x) wr[i] = sqrt(P.r2arr[i]);
  for (int i = threadIdx.x; i < P.ntg; i += blockDim.x) {
    const double t2 = P.tan2arr[i];
    const double c = t2 < 0.0 ? 0.0 : 1.0 / sqrt(1.0 + t2);
    wr[P.nrg + i] = c;
    wr[P.nrg + P.ntg + i] = t2 < 0.0 ? 1.0 : sqrt(t2) * c;
  }
  const WallView W{wl,         wl + P.nrg,          wl + P.nrg + P.ntg, wl + P.nrg + P.ntg + P.npg,
                   wr,         wr + P.nrg,          wr + P.nrg + P.ntg};
  // opacity tables and guides of the active species into LDS
  {
    double *ot = reinterpret_cast<double *>(lds + P.opac_base);
    int32_t *og = reinterpret_cast<int32_t *>(lds + P.opac_base + P.guide_off);
    for (int s2 = 0; s2 < P.nact; ++s2) {
      const int id = P.act[s2] - 1, nl = P.nlambda[id];
      const size_t o = (size_t)P.nlammax * id;
      double *t = ot + P.opac_off[s2];
      for (int i = threadIdx.x; i < nl; i += blockDim.x) {
        t[i] = P.nudust[o + i];
        t[nl + i] = P.kappad[o + i];
        t[2 * nl + i] = P.adust[o + i];
        t[3 * nl + i] = P.gdust[o + i];
        t[4 * nl + i] = P.pldust[o + i];
      }
      for (int i = threadIdx.x; i < kGuide; i += blockDim.x)
        og[s2 * kGuide + i] = P.guide[(size_t)id * kGuide + i];
    }
    double *em = reinterpret_cast<double *>(lds + P.emis_base);
    for (int i = threadIdx.x; i < P.nT; i += blockDim.x) em[i] = P.lTint[i];
    for (int i = threadIdx.x; i < P.nnu; i += blockDim.x) em[P.nT + i] = P.nuint[i];
  }
  const int kQ = P.peel_q;
  PeelRec<R, NA> *queue = nullptr;
  int *qcount = nullptr;
  if constexpr (PEEL) {
    qcount = reinterpret_cast<int *>(lds + P.peel_base);
    queue = reinterpret_cast<PeelRec<R, NA> *>(lds + P.peel_base + 64) + (size_t)wave * kQ;
    if ((threadIdx.x & 63) == 0) qcount[wave] = 0;
  }
  // the workgroup's event pool (pooled passes): request records (kPoolFields
  // doubles per lane, field-major), two queues of tagged lane numbers
  // (scattering, absorption), the completed position of each lane's request,
  // and the queues' head and tail counters
  double *prec = nullptr;
  uint32_t *pq = nullptr, *pdone = nullptr, *pctr = nullptr;
  if constexpr (POOL) {
    prec = reinterpret_cast<double *>(lds + P.pool_base);
    pq = reinterpret_cast<uint32_t *>(prec + kPoolFields * kPoolLanes);
    pdone = pq + 2 * kPoolLanes;
    pctr = pdone + kPoolLanes;
    for (int i = threadIdx.x; i < 2 * kPoolLanes; i += blockDim.x) pq[i] = 0xffffffffu;
    for (int i = threadIdx.x; i < kPoolLanes; i += blockDim.x) pdone[i] = 0xffffffffu;
    if (threadIdx.x < 4) pctr[threadIdx.x] = 0u;
  }
  __syncthreads();
  auto push = [&](const Pkt<R, NA> &k, int sp, bool direct, long long pid,
                  const R *nrm) __attribute__((always_inline)) {
    if constexpr (PEEL) {
      const int qi = atomicAdd(&qcount[wave], 1);
      PeelRec<R, NA> &r = queue[qi];
      make_rec(r, k, sp, direct, pid, nrm);
    }
  };
  Acc acc{};
  if ((threadIdx.x & 63) == 0) cold() = WaveCold{};  // this wavefront's set, before its adds
  Lane<R, NA> L;
  L.la.ntally = 0;
  Trace tr{0.0, 0.0, 0, 0, 0, 0, 0, 14695981039346656037ull};
  R nrm[4] = {R(0.0), R(0.0), R(0.0), R(0.0)};
  double tau = 0.0, tsum = 0.0;
  double dm[NA];  // masked densities of the packet's cell (walk_step prefetches the next)
#pragma unroll
  for (int s2 = 0; s2 < NA; ++s2) dm[s2] = 0.0;
  // status < 0: no packet yet; in serial mode only the grid's first lane runs
  int st = (P.serial && (blockIdx.x != 0 || threadIdx.x != 0)) ? (int)ST_DONE : (int)ST_END;
  int status = -1, slot = 0;
  bool reload = false;  // the restarting walk enters a new cell (launch, star)
  const int ev_min = P.regen_lanes;
  const int ev_run = min(P.regen_run_lanes, ev_min);

  // After a walk step: an interaction draws its species and the
  // scatter/absorb test (select_dust.f90, propagate.f90:60-104); the packet
  // ends when it escaped, was lost, or interacted iintmax times in one cell.
  auto after_step = [&](int s) __attribute__((always_inline)) {
    Pkt<R, NA> &k = L.k;
    if (s == ST_WALK || s == ST_STAR) return s;
    if (s == ST_EXIT) {
      status = 0;
      return (int)ST_END;
    }
    if (s == ST_ERROR || s == ST_ABORT) {
      if (s == ST_ABORT) {
        acc.warn++;
        const unsigned long long i = atomicAdd(P.next_packet + kCtrNDropped, 1ull);
        if (i < (unsigned long long)kDropMax) P.next_packet[kCtrDropped + i] = (unsigned long long)L.pid;
      }
      status = 2;
      return (int)ST_END;
    }
    const size_t cell = cell_of(k.ir, k.it, k.ip);
    if (!L.first) {
      if (k.ir != L.pir || k.it != L.pit || k.ip != L.pip) L.iint = 0;
      if (L.iint > P.iintmax) {
        if (P.killed) atomicAdd(&P.killed[cell], 1);
        cold_inc(cold().abflux);
        status = 1;
        return (int)ST_END;
      }
    }
    slot = select_dust(k, dm, L.rng);
    if (slot < 0) {
      acc.err_dust++;
      status = 2;
      return (int)ST_END;
    }
    if (L.first) cold_inc(cold().tot);
    L.first = false;
    R alb_slot = R(0.0);  // k.alb[slot] without a dynamically indexed register array
#pragma unroll
    for (int s2 = 0; s2 < NA; ++s2)
      if (s2 == slot) alb_slot = k.alb[s2];
    if (ran(L.rng) <= (double)alb_slot) return (int)ST_SCAT;
    if constexpr (!POOL) L.tdust = P.tdust[cell * (size_t)P.nact + slot];  // used when the absorption runs
    return (int)ST_ABS;
  };
  // the walk after an event: new optical depth, entry into the cell
  // reload: the event moved the packet to another cell (launch, star)
  auto start_walk = [&](bool reload) __attribute__((always_inline)) {
    Pkt<R, NA> &k = L.k;
    tau = draw_tau<R>(L.rng);
    L.xpold = k.x;
    L.ypold = k.y;
    L.zpold = k.z;
    tsum = 0.0;
    const int ir0 = k.ir;
    const int s = after_step(walk_begin(W, k));
    if (s == ST_WALK && (reload || k.ir != ir0)) load_dens(k, dm);
    return s;
  };


  // ---- workgroup event pool (pooled passes) ----
  // A lane whose walk reaches a scattering or an absorption hands the part of
  // the event that only needs the packet's direction, Stokes vector,
  // frequency and generator state to the workgroup: it writes a request
  // record into its LDS slot and queues its lane number.  A wavefront takes
  // up to 64 queued requests of one kind at once and runs them on its own
  // lanes, so the expensive event code (the Stokes rejection loop, the
  // re-emission frequency) runs for many packets together although each
  // wavefront only has a few of them.  The owner lane picks up the result
  // and finishes the event itself.  A packet's draws stay in their order:
  // the owner draws nothing while its request is out.
  const int tid = (int)threadIdx.x;
  const int wlane = (int)__lane_id();
  const unsigned long long below = wlane == 0 ? 0ull : (~0ull >> (64 - wlane));
  auto rec = [&](int f, int lane_no) -> double & { return prec[f * kPoolLanes + lane_no]; };
  auto put_rng = [&](int lane_no, const Mrg &g) __attribute__((always_inline)) {
    rec(11, lane_no) = __builtin_bit_cast(double, ((unsigned long long)(uint32_t)g.s2 << 32) |
                                                      (uint32_t)g.s1);
  };
  auto get_rng = [&](int lane_no) __attribute__((always_inline)) {
    const unsigned long long b = __builtin_bit_cast(unsigned long long, rec(11, lane_no));
    return Mrg{(int32_t)(uint32_t)(b & 0xffffffffull), (int32_t)(uint32_t)(b >> 32)};
  };
  // requests of kind q (0 scattering, 1 absorption) queued and not yet
  // taken: the head is read before the tail, and both only grow, so the
  // difference is a count of requests that exist (at most one per lane)
  auto pool_avail = [&](int q) __attribute__((always_inline)) {
    const uint32_t h = __hip_atomic_load(&pctr[q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t t = __hip_atomic_load(&pctr[2 + q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
    const uint32_t a = t - h;
    return (uint32_t)__builtin_amdgcn_readfirstlane((int)(a > (uint32_t)kPoolLanes ? 0u : a));
  };
  // the wavefront's lanes that reached a scattering or absorption: the
  // owner's part of the event, the request record, the queue entry
  auto pool_enqueue_kind = [&](const int q) __attribute__((always_inline)) {
    Pkt<R, NA> &k = L.k;
    {
      const int want = q == 0 ? (int)ST_SCAT : (int)ST_ABS;
      const unsigned long long m = __ballot(st == want);
      if (m == 0ull) return;
      if (st == want) {
        if (q == 0) {
          k.iflag = 1;
          ++L.iscat;
          ++L.iint;
          L.work++;
          acc.nscat_ev++;
          if constexpr (TRACE) tr.nscat++;
          rec(0, tid) = (double)k.d.cost;
          rec(1, tid) = (double)k.d.sint;
          rec(2, tid) = (double)k.d.cosp;
          rec(3, tid) = (double)k.d.sinp;
          rec(4, tid) = (double)k.d.phi;
          rec(5, tid) = (double)k.s.I;
          rec(6, tid) = (double)k.s.Q;
          rec(7, tid) = (double)k.s.U;
          rec(8, tid) = (double)k.s.V;
          rec(9, tid) = (double)k.nub;
          rec(10, tid) = (double)slot;
        } else {
          const int idust2 = P.act[slot] - 1;
          const size_t cell = cell_of(k.ir, k.it, k.ip);
          L.work++;
          lucy_open(L.la, cell);
          L.la.nabs_any += 1;
          // every slot adds (0 or 1): an indexed add would leave the lane's
          // state in scratch memory
#pragma unroll
          for (int s2 = 0; s2 < NA; ++s2) L.la.nabs[s2] += s2 == slot ? 1u : 0u;
          k.i_orig = (idust2 == 0 || idust2 == 1 || idust2 == 4 || idust2 == 5) ? 2 : 3;
          k.s = Stokes4<R>{R(1.0), R(0.0), R(0.0), R(0.0)};
          ++L.iint;
          acc.nabs_ev++;
          if constexpr (TRACE) tr.nabs++;
          rec(0, tid) = (double)(cell * (size_t)P.nact + (size_t)slot);  // the tdust word
          rec(1, tid) = (double)idust2;
        }
        put_rng(tid, L.rng);
      }
      const int leader = __ffsll((long long)m) - 1;
      uint32_t base = 0;
      if (wlane == leader) base = atomicAdd(&pctr[2 + q], (uint32_t)__popcll(m));
      base = (uint32_t)__shfl((int)base, leader);
      if (st == want) {
        const uint32_t pos = base + (uint32_t)__popcll(m & below);
        // the record is complete before the entry that hands it out
        __hip_atomic_store(&pq[q * kPoolLanes + (pos & (kPoolLanes - 1))],
                           ((pos & kPoolTagMask) << 9) | (uint32_t)tid, __ATOMIC_RELEASE,
                           __HIP_MEMORY_SCOPE_WORKGROUP);
        L.qpos = pos;
        st = q == 0 ? (int)ST_SCAT_WAIT : (int)ST_ABS_WAIT;
      }
    }
  };
  auto pool_enqueue = [&]() __attribute__((always_inline)) {
    pool_enqueue_kind(0);
    pool_enqueue_kind(1);
  };
  // take up to 64 requests of kind q (at least pool_min unless eager) and
  // run them on this wavefront's lanes
  auto pool_consume = [&](int q, bool eager) __attribute__((always_inline)) {
    uint32_t h = 0, n = 0;
    if (wlane == 0) {
      // claim [h, h + n) with n at most the requests reserved so far (the
      // tail read after the head), so that every claimed entry is written
      h = __hip_atomic_load(&pctr[q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
      for (;;) {
        const uint32_t t =
            __hip_atomic_load(&pctr[2 + q], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP);
        const uint32_t a = t - h;
        if (a == 0u || a > (uint32_t)kPoolLanes || (!eager && a < (uint32_t)P.pool_min)) {
          n = 0;
          break;
        }
        n = a < 64u ? a : 64u;
        if (__hip_atomic_compare_exchange_strong(&pctr[q], &h, h + n, __ATOMIC_ACQUIRE,
                                                 __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP))
          break;
      }
    }
    n = (uint32_t)__shfl((int)n, 0);
    if (n == 0u) return false;
    h = (uint32_t)__shfl((int)h, 0);
    if ((uint32_t)wlane < n) {
      const uint32_t pos = h + (uint32_t)wlane;
      // the producer reserved this position before writing its entry
      uint32_t e;
      do {
        e = __hip_atomic_load(&pq[q * kPoolLanes + (pos & (kPoolLanes - 1))], __ATOMIC_ACQUIRE,
                              __HIP_MEMORY_SCOPE_WORKGROUP);
      } while ((e >> 9) != (pos & kPoolTagMask));
      const int o = (int)(e & (uint32_t)(kPoolLanes - 1));
      Mrg g = get_rng(o);
      if (q == 0) {
        Dir<R> d{R(rec(0, o)), R(rec(1, o)), R(rec(2, o)), R(rec(3, o)), R(rec(4, o))};
        Stokes4<R> sv{R(rec(5, o)), R(rec(6, o)), R(rec(7, o)), R(rec(8, o))};
        const HG<R> hg = hg_of<R>((int)rec(10, o), R(rec(9, o)));
        stokes<R>(d, sv, hg, g, acc.htot, acc.hit);
        rec(0, o) = (double)d.cost;
        rec(1, o) = (double)d.sint;
        rec(2, o) = (double)d.cosp;
        rec(3, o) = (double)d.sinp;
        rec(4, o) = (double)d.phi;
        rec(5, o) = (double)sv.I;
        rec(6, o) = (double)sv.Q;
        rec(7, o) = (double)sv.U;
        rec(8, o) = (double)sv.V;
      } else {
        const double tdu = P.tdust[(size_t)rec(0, o)];
        const int idust2 = (int)rec(1, o);
        // the draws in their order (direction, direction, frequency)
        const R u1 = R(ran(g));
        const R u2 = R(ran(g));
        const LdfState<R> ldf = ldf_begin<R>(tdu, idust2, g, P.kapBint, P.cdfguide);
        Dir<R> d;
        d.cost = R(1.0) - R(2.0) * u1;
        d.sint = m_sqrt(R(1.0) - m_pow(d.cost, R(2.0)));
        d.phi = R(kR2P) * u2;
        d.cosp = m_cos(d.phi);
        d.sinp = m_sin(d.phi);
        rec(9, o) = (double)ldf_end<R>(ldf);
        rec(0, o) = (double)d.cost;
        rec(1, o) = (double)d.sint;
        rec(2, o) = (double)d.cosp;
        rec(3, o) = (double)d.sinp;
        rec(4, o) = (double)d.phi;
      }
      put_rng(o, g);
      __hip_atomic_store(&pdone[o], pos, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_WORKGROUP);
    }
    return true;
  };
  // lanes whose request has been run
  auto pool_poll = [&]() __attribute__((always_inline)) {
    if (st == ST_SCAT_WAIT || st == ST_ABS_WAIT) {
      if (__hip_atomic_load(&pdone[tid], __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_WORKGROUP) ==
          L.qpos)
        st = st == ST_SCAT_WAIT ? (int)ST_SCAT_READY : (int)ST_ABS_READY;
    }
  };
  // the owner's part after the pooled event
  auto pool_pickup = [&]() __attribute__((always_inline)) {
    Pkt<R, NA> &k = L.k;
    if (st == ST_SCAT_READY || st == ST_ABS_READY) {
      const bool sc = st == ST_SCAT_READY;
      k.d = Dir<R>{R(rec(0, tid)), R(rec(1, tid)), R(rec(2, tid)), R(rec(3, tid)), R(rec(4, tid))};
      L.rng = get_rng(tid);
      if (sc) {
        k.s = Stokes4<R>{R(rec(5, tid)), R(rec(6, tid)), R(rec(7, tid)), R(rec(8, tid))};
        set_walk_dir(k);
      } else {
        k.nub = R(rec(9, tid));
        set_walk_dir(k);
        k.iflag = 0;
        set_opacity(k, false);
      }
      L.pir = k.ir;
      L.pit = k.it;
      L.pip = k.ip;
      st = ST_RESTART;
      reload = false;
    }
  };
  // star re-emission (the packet hit the stellar surface)
  auto ev_star = [&]() __attribute__((always_inline)) {
    Pkt<R, NA> &k = L.k;
    L.work++;
    star_reemit(k, L.rng, acc, nrm);
    if constexpr (TRACE) tr.nstar++;
    push(k, 0, true, L.pid, nrm);
    st = ST_RESTART;
    reload = true;
  };
  // the end of a packet and the launch of the next
  auto ev_end = [&]() __attribute__((always_inline)) {
    Pkt<R, NA> &k = L.k;
    if (status >= 0) {  // the packet ends (0 escaped, 1 killed, 2 lost)
      acc.maxwork = max(acc.maxwork, L.work);
      lucy_flush(L.la);
      if (status == 0) bin_exit(k, L.xpold, L.ypold, L.zpold, L.iscat, acc);
      if constexpr (TRACE) {
        hc_trace &o = P.trace[L.pid];
        o.x = k.x;
        o.y = k.y;
        o.z = k.z;
        o.cost = k.d.cost;
        o.sint = k.d.sint;
        o.phi = k.d.phi;
        o.I = k.s.I;
        o.Q = k.s.Q;
        o.U = k.s.U;
        o.V = k.s.V;
        o.nub = (double)k.nub;
        o.path = tr.path;
        o.ncross = tr.ncross;
        o.nscat = tr.nscat;
        o.nabs = tr.nabs;
        o.nstar = tr.nstar;
        o.ntrial = 0;
        o.cellhash = tr.hash;
        o.status = status;
        o.ir = k.ir;
        o.it = k.it;
        o.ip = k.ip;
        o.pad = 0;
      }
    }
    // take the next packet and launch it: the lanes of the wavefront that
    // need one share one atomic on the packet counter (a block of packet
    // numbers), each lane taking its rank among them
    const unsigned long long need = __ballot(1);
    const int leader = __ffsll((long long)need) - 1;
    const int lane = (int)__lane_id();
    unsigned long long base = 0;
    if (lane == leader)
      base = atomicAdd(P.next_packet + kCtrNext, (unsigned long long)__popcll(need));
    const unsigned blo = __shfl((unsigned)(base & 0xffffffffull), leader);
    const unsigned bhi = __shfl((unsigned)(base >> 32), leader);
    const unsigned long long p = (((unsigned long long)bhi << 32) | blo) +
                                 (unsigned long long)__popcll(need & ((1ull << lane) - 1ull));
    if (p >= (unsigned long long)P.np) {
      st = ST_DONE;
    } else {
      L.pid = (long long)p;
      L.rng = mrg_segment(P.g1, P.g2, seg_base + p, P.pow1, P.pow2);
      L.la.cell = -1;
      L.iscat = L.iint = L.pir = L.pit = L.pip = 0;
      L.work = 0;
      L.first = true;
      status = -1;
      if constexpr (TRACE) tr = Trace{0.0, 0.0, 0, 0, 0, 0, 0, 14695981039346656037ull};
      const long long iphot = (long long)p + 1;
      st = ST_RESTART;
      reload = true;
      if (iphot < P.npacc) {  // newdisktherm.F90:209
        const int sl = launch_acc(W, k, L.rng);
        if (sl < 0) {
          acc.err_dust++;
          status = 2;
          st = ST_END;
        } else {
          if constexpr (TRACE) hash_cell(tr, k.ir, k.it, k.ip);
          push(k, sl, false, L.pid, nrm);
        }
      } else {
        launch_star(k, L.rng, iphot < P.npacc + P.npaccstar, nrm);
        if constexpr (TRACE) hash_cell(tr, k.ir, k.it, k.ip);
        push(k, 0, true, L.pid, nrm);
      }
    }
  };

  for (;;) {
    // ---------------- event phase ----------------
    if constexpr (POOL) {
      pool_enqueue();
      const unsigned long long walkers0 = __ballot(st == ST_WALK);
      const bool eager = __popcll(walkers0) < P.pool_walk;
      bool ran_pool = pool_consume(0, eager);
      ran_pool = pool_consume(1, eager) || ran_pool;
      pool_poll();
      pool_pickup();
      const unsigned long long walkers = __ballot(st == ST_WALK);
      const int n_r = __popcll(__ballot(st == ST_RESTART));
      const int n_t = __popcll(__ballot(st == ST_STAR));
      const int n_e = __popcll(__ballot(st == ST_END));
      const bool all = walkers == 0ull;
      const bool trig = n_r >= ev_min || n_t >= ev_min || n_e >= ev_min;
      const int lim = all ? 1 : (trig ? ev_run : ev_min);
      if (n_t > 0 && n_t >= lim && st == ST_STAR) ev_star();
      if (n_e > 0 && n_e >= lim && st == ST_END) ev_end();
      const bool any_ev = __ballot(st == ST_RESTART) != 0ull;
      if (st == ST_RESTART) st = start_walk(reload);
      // a wavefront whose lanes all wait for requests that other wavefronts
      // are running has nothing to do until they finish
      if (!ran_pool && !any_ev && all &&
          __ballot(st == ST_WALK || st == ST_SCAT || st == ST_ABS || st == ST_END ||
                   st == ST_STAR) == 0ull)
        __builtin_amdgcn_s_sleep(2);
    } else {
      const unsigned long long walkers = __ballot(st == ST_WALK);
      const int n_s = __popcll(__ballot(st == ST_SCAT));
      const int n_a = __popcll(__ballot(st == ST_ABS));
      const int n_t = __popcll(__ballot(st == ST_STAR));
      const int n_e = __popcll(__ballot(st == ST_END));
      // once one kind has ev_min waiters, every kind with ev_run waiters runs
      const bool all = walkers == 0ull;
      const bool trig = n_s >= ev_min || n_a >= ev_min || n_t >= ev_min || n_e >= ev_min;
      const int lim = all ? 1 : (trig ? ev_run : ev_min);
      Pkt<R, NA> &k = L.k;
      if (n_s > 0 && n_s >= lim && st == ST_SCAT) {
        k.iflag = 1;
        ++L.iscat;
        ++L.iint;
        L.work++;
        const HG<R> h = hg_of<R>(slot, k.nub);
        push(k, slot, false, L.pid, nrm);
        stokes<R>(k.d, k.s, h, L.rng, acc.htot, acc.hit);
        acc.nscat_ev++;
        if constexpr (TRACE) tr.nscat++;
        set_walk_dir(k);
        L.pir = k.ir;
        L.pit = k.it;
        L.pip = k.ip;
        st = ST_RESTART;
        reload = false;
      }
      if (n_a > 0 && n_a >= lim && st == ST_ABS) {
        const int idust2 = P.act[slot] - 1;
        const size_t cell = cell_of(k.ir, k.it, k.ip);
        L.work++;
        lucy_open(L.la, cell);
        L.la.nabs_any += 1;
        k.i_orig = (idust2 == 0 || idust2 == 1 || idust2 == 4 || idust2 == 5) ? 2 : 3;
        const bool layer = P.diffus != nullptr && P.diffus[cell] != 0;
        bool ok = true;
        if (layer) {
          // propagate.f90:123-142: no count in the layer; the packet leaves
          // from its surface (Stokes vector kept)
          double w8[2] = {0.0, 0.0};
#pragma unroll
          for (int s2 = 0; s2 < NA; ++s2) {
            if (s2 == P.slot_of[0]) w8[0] = k.kapd[s2];
            if (s2 == P.slot_of[1]) w8[1] = k.kapd[s2];
          }
          const int id2 = emit_diffusion(k, L.rng, w8);
          ok = id2 >= 0;
          if (ok) slot = P.slot_of[id2];
        } else {
          double tem = L.tdust;
          if (P.bw) {
            tem = bw_temperature(cell * (size_t)P.nact + slot, idust2);
          } else {
#pragma unroll
            for (int s2 = 0; s2 < NA; ++s2)
              if (s2 == slot) L.la.nabs[s2] += 1u;
          }
          k.s = Stokes4<R>{R(1.0), R(0.0), R(0.0), R(0.0)};
          // the draws in their order (direction, direction, frequency); the
          // frequency's table loads fly while the direction is computed
          const R u1 = R(ran(L.rng));
          const R u2 = R(ran(L.rng));
          const bool planck = P.bw && tem >= P.tsub_ev;  // dustfreq at sublimation
          LdfState<R> ldf;
          if (!planck)
            ldf = ldf_begin<R>(tem, idust2, L.rng, P.bw ? P.kapdBint : P.kapBint,
                               P.bw ? P.cdfguide_bw : P.cdfguide);
          k.d.cost = R(1.0) - R(2.0) * u1;
          k.d.sint = m_sqrt(R(1.0) - m_pow(k.d.cost, R(2.0)));
          k.d.phi = R(kR2P) * u2;
          k.d.cosp = m_cos(k.d.phi);
          k.d.sinp = m_sin(k.d.phi);
          k.nub = planck ? bbfreq<R>(L.rng, R(tem)) : ldf_end<R>(ldf);
        }
        if (ok) {
          set_walk_dir(k);
          k.iflag = 0;
          set_opacity(k, false);
          push(k, slot, false, L.pid, nrm);
          ++L.iint;
          acc.nabs_ev++;
          if constexpr (TRACE) tr.nabs++;
          L.pir = k.ir;
          L.pit = k.it;
          L.pip = k.ip;
          st = ST_RESTART;
          // the cell's mask may have changed with its temperature, or the
          // packet left the layer for another cell
          reload = P.bw != 0 || layer;
        } else {
          acc.err_dust++;
          status = 2;
          st = ST_END;
        }
      }
      if (n_t > 0 && n_t >= lim && st == ST_STAR) ev_star();
      if (n_e > 0 && n_e >= lim && st == ST_END) ev_end();
      // every lane whose event ran starts its next walk here, the event
      // kinds together: the draw of the new optical depth and the first
      // walls of the walk are the same code for all of them
      if (st == ST_RESTART) st = start_walk(reload);
    }
    if constexpr (PEEL) {
      const int n = qcount[wave];
      const bool last = __ballot(st != ST_DONE) == 0ull;
      if (n > kQ - 64 || (last && n > 0)) {
        flush_peels<R, NA, TRACE>(W, queue, n, acc);
        if ((threadIdx.x & 63) == 0) qcount[wave] = 0;
      }
    }
    if (__ballot(st != ST_DONE) == 0ull) break;
    // ---------------- walk phase ----------------
    if constexpr (POOL) {
      for (;;) {
        pool_enqueue();
        pool_poll();
        if (__ballot(st == ST_WALK) == 0ull) break;
        if (__popcll(__ballot(st == ST_SCAT_READY || st == ST_ABS_READY)) >= ev_min ||
            __popcll(__ballot(st == ST_END)) >= ev_min || __popcll(__ballot(st == ST_STAR)) >= ev_min ||
            pool_avail(0) >= (uint32_t)P.pool_min || pool_avail(1) >= (uint32_t)P.pool_min)
          break;
        if (st == ST_WALK) {
          int s = walk_step<R, NA, TRACE>(W, L.k, tau, tsum, dm, L.la, acc, tr);
          if (++L.work > P.max_work) s = ST_ABORT;
          st = after_step(s);
        }
      }
      continue;
    }
    for (;;) {
      if (__ballot(st == ST_WALK) == 0ull) break;
      if (__popcll(__ballot(st == ST_SCAT)) >= ev_min || __popcll(__ballot(st == ST_ABS)) >= ev_min ||
          __popcll(__ballot(st == ST_END)) >= ev_min || __popcll(__ballot(st == ST_STAR)) >= ev_min)
        break;
      if (st == ST_WALK) {
        int s = walk_step<R, NA, TRACE>(W, L.k, tau, tsum, dm, L.la, acc, tr);
        if (++L.work > P.max_work) s = ST_ABORT;
        st = after_step(s);
      }
    }
  }
  // (the wavefront's once-per-packet counters are added from its LDS set below)
  const double v[S_COUNT] = {0.0,                0.0,                  0.0,
                             (double)acc.nscat,  0.0,                  0.0,
                             0.0,                0.0,                  0.0,
                             (double)acc.hit,    (double)acc.htot,     (double)acc.err_wall,
                             (double)acc.err_dust, (double)acc.warn,
                             (double)acc.ncross, (double)acc.nscat_ev, (double)acc.nabs_ev,
                             (double)acc.npeelray, (double)acc.npeelcross,
                             (double)L.la.ntally};
  const WaveCold &wc = cold();
  const double vc[S_COUNT] = {(double)wc.tot, (double)wc.flux, (double)wc.sflux, 0.0,
                              (double)wc.scount, (double)wc.abflux, (double)wc.aflux,
                              wc.sumsub, wc.sumall};
#pragma unroll
  for (int i = 0; i < S_COUNT; ++i) {
    const double w = wave_sum(v[i]) + vc[i];
    if ((threadIdx.x & 63) == 0 && w != 0.0) atomic_add(&P.scal[i], w);
  }
  unsigned mw = acc.maxwork;
  for (int off = 32; off > 0; off >>= 1) mw = max(mw, (unsigned)__shfl_xor((int)mw, off, 64));
  if ((threadIdx.x & 63) == 0) atomicMax(P.next_packet + kCtrLongest, (unsigned long long)mw);
}

// ---------------------------------------------------------------------------
// Lucy temperature (tfinal reconstruction with dusttemp.f90:6-58).
// ---------------------------------------------------------------------------
struct TParams {
  int nrg, ntg, npg, nact, nT;
  int act[kMaxSpecies];
  double kappav[kMaxSpecies];
  const double *lTint, *kappap, *dtauabs, *mass4, *tdust;
  double *tdust2;
  double ltot, nphot;
  unsigned long long *nchanged;
};

__device__ __forceinline__ double kapp(const TParams &T, double t, int id) {  // dust.f90:114-142
  const double *kp = T.kappap + (size_t)T.nT * id;
  const double lt = hcm_log(t);
  const int k = (int)((double)(T.nT - 1) * (lt - T.lTint[0]) / (T.lTint[T.nT - 1] - T.lTint[0])) + 1;
  if (k < 1 || k > T.nT - 1) return k < 1 ? kp[0] : kp[T.nT - 1];
  const double f = (lt - T.lTint[k - 1]) / (T.lTint[k] - T.lTint[k - 1]);
  return (1.0 - f) * kp[k - 1] + f * kp[k];
}

__global__ __launch_bounds__(256) void lucy_temperature(TParams T) {
  const size_t ncellw = (size_t)T.nrg * T.ntg * T.npg;
  const size_t n = ncellw * (size_t)T.nact;
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long changed = 0;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) {
    const size_t cell = i / (size_t)T.nact;
    const int s = (int)(i - cell * (size_t)T.nact);
    const int id = T.act[s] - 1;
    const int ir = (int)(cell % (size_t)T.nrg);
    const int it = (int)((cell / (size_t)T.nrg) % (size_t)T.ntg);
    const int ip = (int)(cell / ((size_t)T.nrg * T.ntg));
    const double m4 = T.mass4[i];
    const double told = T.tdust[i];
    double tnew = 0.1 / 11605.0;
    if (ir < T.nrg - 1 && it < T.ntg - 1 && ip < T.npg - 1 && m4 > 0.0) {
      const double na = T.dtauabs[i];
      double Td = (told != 0.0) ? told : 0.1 / 11605.0;
      double Tprev = 0.5 * Td;
      int iter = 0;
      while (fabs(Td - Tprev) > 1.e-5 * Td && iter < 1000) {
        Tprev = Td;
        Td = hcm_pow((na * T.ltot) / (4.0 * kSigT * T.nphot * T.kappav[id] * kapp(T, Td, id)),
                     0.25) /
             11605.0;
        Td = Td / m4;
        if (Td > 1600.0 / 11605.0) Td = 1600.0 / 11605.0;
        Td = fmax(Td, 0.1 / 11605.0);
        ++iter;
      }
      tnew = Td;
      if (fabs(tnew - told) > 0.01 * told) ++changed;
    }
    T.tdust2[i] = tnew;
  }
  for (int off = 32; off > 0; off >>= 1) changed += __shfl_xor(changed, off, 64);
  if ((threadIdx.x & 63) == 0 && changed) atomicAdd(T.nchanged, changed);
}

// The temperatures of the diffusion layer's cells back to 0.1 K (every
// species), as initarr does before each iteration.
__global__ void reset_layer(const int8_t *diffus, double *tdust, size_t ncellw, int nact) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  const size_t n = ncellw * (size_t)nact;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x)
    if (diffus[i / (size_t)nact]) tdust[i] = 0.1 / 11605.0;
}

// Masked density: zero where the species is at or above Tsub.
__global__ void mask_density(const double *dens, const double *tdust, double tsub_ev,
                             double *densm, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) densm[i] = tdust[i] < tsub_ev ? dens[i] : -dens[i];
}

}  // namespace hc

// ===========================================================================
// C ABI
// ===========================================================================
using namespace hc;

struct hc_ctx {
  hc_dims dims;
  int device = 0;
  std::string err;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  bool have_grid = false, have_dust = false, have_src = false;
  int nact = 0;
  int act[kMaxSpecies] = {0};
  int slot[kMaxSpecies];
  size_t ncellw = 0;
  double tsub = 1600.0;
  // device buffers
  double *r2arr = nullptr, *tan2arr = nullptr, *thetarr = nullptr, *aarr = nullptr,
         *barr = nullptr, *phiarr = nullptr, *costarr = nullptr;
  double *dens = nullptr, *densm = nullptr, *tdust = nullptr;
  double *nudust = nullptr, *kappad = nullptr, *adust = nullptr, *gdust = nullptr,
         *pldust = nullptr, *lTint = nullptr, *nuint = nullptr, *kapBint = nullptr,
         *kappap = nullptr;
  double *u = nullptr, *sini = nullptr, *aperture2 = nullptr, *coste = nullptr,
         *sinte = nullptr, *cospe = nullptr, *sinpe = nullptr;
  int32_t *pow1 = nullptr, *pow2 = nullptr, *guide = nullptr;
  uint16_t *cdfguide = nullptr;
  double glog0[kMaxSpecies] = {0}, ginv[kMaxSpecies] = {0};
  double *dtauabs = nullptr, *nabs = nullptr, *jmean = nullptr;
  int32_t *killed = nullptr, *dbig = nullptr;
  double *sed = nullptr, *peel = nullptr, *img_i = nullptr, *img_q = nullptr, *img_u = nullptr,
         *img_v = nullptr, *si_init = nullptr, *scal = nullptr, *mass4 = nullptr,
         *tdust2 = nullptr;
  unsigned long long *counter = nullptr;
  unsigned long long *segdev = nullptr;  // segment base of graph-replayed passes
  // stellar spectrum table (hc_set_atmosphere) and limb-darkening table
  // (hc_set_limb_table)
  int nhnu = 0;
  double *atm = nullptr;
  int32_t *atm_guide = nullptr;
  int ntab = 0;
  double *limb_tab = nullptr;
  hc_trace *trace = nullptr;
  size_t trace_cap = 0;
  // host copies of scalars
  int32_t nlambda[kMaxSpecies] = {0};
  double kappav[kMaxSpecies] = {0};
  hc_sources src;
  // launch shape
  int regen_lanes = 8;  // 12 on the FP32 path (hc_create)
  int regen_run_lanes = 8;
  int blocks_per_cu = 0;
  int peel_queue = 0;  // records per wavefront peel queue, 0 = 128 or what fits
  size_t lds_per_block = 64 * 1024;  // the device's limit (hc_create)
  int pool = 0;        // pooled scattering/absorption in Lucy passes: switched off
  int bw = 0;          // Bjorkman-Wood re-emission (hc_set_bw)
  int serial = 0;      // one lane runs the packets in order (hc_set_option "serial")
  double *kapdBint = nullptr;
  uint16_t *cdfguide_bw = nullptr;
  // filter curves (hc_set_filters) and the filter images of the last pass
  int nfilt = 0, nwfiltmax = 0;
  int32_t *nwfilt = nullptr;
  double *filtwave = nullptr, *filtphi = nullptr, *imgb = nullptr;
  size_t imgb_words = 0;
  bool imgb_valid = false;
  // diffusion layer (hc_set_diffusion)
  int8_t *diffus = nullptr, *diffdir = nullptr;
  double *rarr = nullptr, *sintarr = nullptr, *kappar = nullptr;
  int pool_min = 48;   // requests a walking wavefront takes at once
  int pool_walk = 16;  // walking lanes below which a wavefront takes any request
  uint32_t max_work = 1u << 24;
};

namespace {

int fail(hc_ctx *c, int code, const std::string &msg) {
  if (c) c->err = msg;
  return code;
}

#define HCK(expr)                                                                    \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      return fail(ctx, HC_E_HIP, std::string(#expr) + ": " + hipGetErrorString(e_)); \
  } while (0)

// Copies between host and device go through the context's own (non-blocking)
// stream: the null stream of a plain hipMemcpy would order them after the
// work of every other context on the device, and contexts that run side by
// side would wait for each other's passes.
#define HC_COPY_(dst, src, bytes, kind)                                               \
  do {                                                                               \
    HCK(hipMemcpyAsync((dst), (src), (bytes), (kind), ctx->stream));                 \
    HCK(hipStreamSynchronize(ctx->stream));                                          \
  } while (0)
#define HC_D2H(dst, src, bytes) HC_COPY_(dst, src, bytes, hipMemcpyDeviceToHost)
#define HC_H2D(dst, src, bytes) HC_COPY_(dst, src, bytes, hipMemcpyHostToDevice)

template <class T>
int dalloc(hc_ctx *ctx, T **p, size_t n) {
  if (*p) return HC_OK;
  HCK(hipMalloc((void **)p, std::max<size_t>(n, 1) * sizeof(T)));
  return HC_OK;
}

template <class T>
int upload(hc_ctx *ctx, T **p, const T *h, size_t n) {
  int rc = dalloc(ctx, p, n);
  if (rc) return rc;
  HC_H2D(*p, h, n * sizeof(T));
  return HC_OK;
}

int64_t powmod_host(int64_t a, uint64_t e, int64_t m) {
  int64_t r = 1, b = a % m;
  while (e) {
    if (e & 1u) r = (r * b) % m;
    b = (b * b) % m;
    e >>= 1;
  }
  return r;
}

// The guide of a normalised cumulative emissivity table (kapBint or
// kapdBint, (nT,nnu,ndg)): per species and temperature row, the number of
// nodes <= b/kCdfGuide.  Built only when every row increases strictly (then
// the bracket of an interpolated CDF is unique); *out is null otherwise.
int build_cdf_guide(hc_ctx *ctx, const double *table, uint16_t **out) {
  const hc_dims &d = ctx->dims;
  const int nT = d.nT, nnu = d.nnu;
  std::vector<uint16_t> cg((size_t)d.ndg * nT * (kCdfGuide + 1));
  bool strict = nnu < 65536;
  for (int id = 0; id < d.ndg && strict; ++id)
    for (int j = 0; j < nT && strict; ++j) {
      const double *kb = table + (size_t)nT * nnu * id + j;
      for (int i = 1; i < nnu; ++i)
        if (!(kb[(size_t)nT * i] > kb[(size_t)nT * (i - 1)])) {
          strict = false;
          break;
        }
      uint16_t *g = cg.data() + ((size_t)id * nT + j) * (kCdfGuide + 1);
      int cnt = 0;
      for (int b = 0; b <= kCdfGuide; ++b) {
        const double v = (double)b / (double)kCdfGuide;
        while (cnt < nnu && kb[(size_t)nT * cnt] <= v) ++cnt;
        g[b] = (uint16_t)cnt;
      }
    }
  if (*out) {
    (void)hipFree(*out);
    *out = nullptr;
  }
  if (strict) return upload(ctx, out, cg.data(), cg.size());
  return HC_OK;
}

// Cell-major, active-species-inner copy of a (nrg,ntg,npg,ndg) array.
void to_device_layout(const hc_ctx *ctx, const double *ref, std::vector<double> &out) {
  out.assign(ctx->ncellw * (size_t)ctx->nact, 0.0);
  for (int s = 0; s < ctx->nact; ++s) {
    const double *src = ref + ctx->ncellw * (size_t)(ctx->act[s] - 1);
    for (size_t c = 0; c < ctx->ncellw; ++c) out[c * (size_t)ctx->nact + s] = src[c];
  }
}

int queue_mask(hc_ctx *ctx) {
  const size_t n = ctx->ncellw * (size_t)ctx->nact;
  hipLaunchKernelGGL(mask_density, dim3(1024), dim3(256), 0, ctx->stream, ctx->dens, ctx->tdust,
                     ctx->tsub / 11605.0, ctx->densm, n);
  HCK(hipGetLastError());
  return HC_OK;
}

int refresh_mask(hc_ctx *ctx) {
  const int rc = queue_mask(ctx);
  if (rc) return rc;
  HCK(hipStreamSynchronize(ctx->stream));
  return HC_OK;
}

}  // namespace

extern "C" {

int hc_abi_version(void) { return HC_ABI_VERSION; }

const char *hc_last_error(const hc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int hc_create(const hc_dims *dims, int device, hc_ctx **out) {
  if (!dims || !out) return HC_E_ARG;
  *out = nullptr;
  if (dims->nrg < 2 || dims->ntg < 2 || dims->npg < 2 || dims->ndg < 1 ||
      dims->ndg > kMaxSpecies || dims->nfreq < 1 || dims->nmu < 1 || dims->nph < 1 ||
      dims->nap < 1 || dims->npeel < 0 || dims->nxhst < 1 || dims->nlammax < 2 || dims->nT < 2 ||
      dims->nnu < 2)
    return HC_E_ARG;
  if (dims->precision != HC_FP64 && dims->precision != HC_FP32) return HC_E_UNSUPPORTED;
  if (dims->imcube < 0 || dims->imcube > 2) return HC_E_ARG;
  hc_ctx *ctx = new hc_ctx();
  ctx->dims = *dims;
  if (dims->precision == HC_FP32) ctx->regen_lanes = 12;
  ctx->device = device;
  ctx->ncellw = (size_t)dims->nrg * dims->ntg * dims->npg;
  for (int i = 0; i < kMaxSpecies; ++i) ctx->slot[i] = -1;
  hipError_t e = hipSetDevice(device);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev0);
  if (e == hipSuccess) e = hipEventCreate(&ctx->ev1);
  if (e != hipSuccess) {
    delete ctx;
    return HC_E_HIP;
  }
  int lds = 0;  // the workgroup LDS limit; 64 KiB when the runtime does not say
  if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess &&
      lds >= 64 * 1024)
    ctx->lds_per_block = (size_t)lds;
  (void)hipGetLastError();
  // RNG jump tables: (a^(2^30))^(2^k) mod m for k = 0..63
  std::vector<int32_t> p1(64), p2(64);
  int64_t b1 = 1033780774, b2 = 1494757890;  // a1w, a2w (ecuyer_cote_mod.f90:14-15)
  for (int k = 0; k < 64; ++k) {
    p1[k] = (int32_t)b1;
    p2[k] = (int32_t)b2;
    b1 = (b1 * b1) % 2147483563;
    b2 = (b2 * b2) % 2147483399;
  }
  int rc = upload(ctx, &ctx->pow1, p1.data(), 64);
  if (!rc) rc = upload(ctx, &ctx->pow2, p2.data(), 64);
  if (!rc) rc = dalloc(ctx, &ctx->scal, (size_t)S_COUNT);
  if (!rc) rc = dalloc(ctx, &ctx->counter, kCounters);
  double z[146];
  z[0] = 0.0;
  for (int i = 1; i <= 145; ++i) z[i] = 0.0;
  for (int j = 145; j >= 1; --j) {
    const double t = 1.0 / (double)(j * j * j * j);
    for (int i = 145; i >= j; --i) z[i] = z[i] + t;
  }
  // the table is the same for every context: written once per device, so a
  // context made while another one's pass runs does not write what it reads
  if (!rc) {
    static std::mutex zeta_mu;
    static std::vector<int> zeta_done;
    std::lock_guard<std::mutex> lock(zeta_mu);
    if (std::find(zeta_done.begin(), zeta_done.end(), device) == zeta_done.end()) {
      if (hipMemcpyToSymbol(HIP_SYMBOL(c_inczeta), z, sizeof z) != hipSuccess)
        rc = HC_E_HIP;
      else
        zeta_done.push_back(device);
    }
  }
  if (rc) {
    hc_destroy(ctx);
    return rc;
  }
  *out = ctx;
  return HC_OK;
}

int hc_set_grid(hc_ctx *ctx, const double *r2arr, const double *tan2arr, const double *thetarr,
                const double *aarr, const double *barr, const double *phiarr,
                const double *densarr, const double *tdust, double tsub) {
  if (!ctx) return HC_E_ARG;
  if (!r2arr || !tan2arr || !thetarr || !aarr || !barr || !phiarr || !densarr || !tdust)
    return fail(ctx, HC_E_ARG, "hc_set_grid: null array");
  const hc_dims &d = ctx->dims;
  HCK(hipSetDevice(ctx->device));
  // active species: any non-zero density (fixed at the first call)
  if (!ctx->have_grid) {
    ctx->nact = 0;
    for (int id = 0; id < d.ndg; ++id) {
      const double *p = densarr + ctx->ncellw * (size_t)id;
      bool any = false;
      for (size_t c = 0; c < ctx->ncellw && !any; ++c) any = p[c] != 0.0;
      if (any) {
        ctx->slot[id] = ctx->nact;
        ctx->act[ctx->nact++] = id + 1;
      }
    }
    if (ctx->nact == 0) return fail(ctx, HC_E_ARG, "hc_set_grid: all densities are zero");
  }
  int rc = upload(ctx, &ctx->r2arr, r2arr, (size_t)d.nrg);
  if (!rc) rc = upload(ctx, &ctx->tan2arr, tan2arr, (size_t)d.ntg);
  if (!rc) rc = upload(ctx, &ctx->thetarr, thetarr, (size_t)d.ntg);
  if (!rc) rc = upload(ctx, &ctx->aarr, aarr, (size_t)d.npg);
  if (!rc) rc = upload(ctx, &ctx->barr, barr, (size_t)d.npg);
  if (!rc) rc = upload(ctx, &ctx->phiarr, phiarr, (size_t)d.npg);
  {
    // costarr as gridset.f90:580-599 builds it: exact at the poles (tan2 = 0)
    // and at a midplane wall (tan2 = -1), cos(theta) elsewhere
    std::vector<double> ct(d.ntg);
    for (int i = 0; i < d.ntg; ++i)
      ct[i] = tan2arr[i] == 0.0 ? (i == 0 ? 1.0 : -1.0)
                                : (tan2arr[i] == -1.0 ? 0.0 : std::cos(thetarr[i]));
    if (!rc) rc = upload(ctx, &ctx->costarr, ct.data(), (size_t)d.ntg);
    // sintarr likewise: 0 at the poles, 1 at a midplane wall
    for (int i = 0; i < d.ntg; ++i)
      ct[i] = tan2arr[i] == 0.0 ? 0.0 : (tan2arr[i] == -1.0 ? 1.0 : std::sin(thetarr[i]));
    if (!rc) rc = upload(ctx, &ctx->sintarr, ct.data(), (size_t)d.ntg);
  }
  if (rc) return rc;
  std::vector<double> tmp;
  to_device_layout(ctx, densarr, tmp);
  rc = upload(ctx, &ctx->dens, tmp.data(), tmp.size());
  if (!rc) rc = dalloc(ctx, &ctx->densm, tmp.size());
  to_device_layout(ctx, tdust, tmp);
  if (!rc) rc = upload(ctx, &ctx->tdust, tmp.data(), tmp.size());
  if (rc) return rc;
  ctx->tsub = tsub;
  ctx->have_grid = true;
  return refresh_mask(ctx);
}

int hc_set_temperature(hc_ctx *ctx, const double *tdust) {
  if (!ctx || !tdust) return HC_E_ARG;
  if (!ctx->have_grid) return fail(ctx, HC_E_STATE, "hc_set_temperature before hc_set_grid");
  HCK(hipSetDevice(ctx->device));
  std::vector<double> tmp;
  to_device_layout(ctx, tdust, tmp);
  HC_H2D(ctx->tdust, tmp.data(), tmp.size() * sizeof(double));
  return refresh_mask(ctx);
}

int hc_get_temperature(hc_ctx *ctx, double *tdust) {
  if (!ctx || !tdust) return HC_E_ARG;
  if (!ctx->have_grid) return fail(ctx, HC_E_STATE, "hc_get_temperature before hc_set_grid");
  HCK(hipSetDevice(ctx->device));
  const size_t nc = ctx->ncellw * (size_t)ctx->nact;
  std::vector<double> tmp(nc);
  HC_D2H(tmp.data(), ctx->tdust, nc * sizeof(double));
  for (size_t i = 0; i < ctx->ncellw * (size_t)ctx->dims.ndg; ++i) tdust[i] = 0.1 / 11605.0;
  for (int k = 0; k < ctx->nact; ++k) {
    double *dst = tdust + ctx->ncellw * (size_t)(ctx->act[k] - 1);
    for (size_t c = 0; c < ctx->ncellw; ++c) dst[c] = tmp[c * (size_t)ctx->nact + k];
  }
  return HC_OK;
}

int hc_set_dust(hc_ctx *ctx, const int32_t *nlambda, const double *nudust, const double *kappad,
                const double *adust, const double *gdust, const double *pldust,
                const double *kappav, const double *lTint, const double *nuint,
                const double *kapBint, const double *kappap) {
  if (!ctx) return HC_E_ARG;
  if (!nlambda || !nudust || !kappad || !adust || !gdust || !pldust || !kappav || !lTint ||
      !nuint || !kapBint || !kappap)
    return fail(ctx, HC_E_ARG, "hc_set_dust: null array");
  const hc_dims &d = ctx->dims;
  HCK(hipSetDevice(ctx->device));
  for (int id = 0; id < d.ndg; ++id) {
    if (nlambda[id] < 1 || nlambda[id] > d.nlammax)
      return fail(ctx, HC_E_ARG, "hc_set_dust: nlambda out of range");
    ctx->nlambda[id] = nlambda[id];
    ctx->kappav[id] = kappav[id];
  }
  const size_t nl = (size_t)d.nlammax * d.ndg;
  int rc = upload(ctx, &ctx->nudust, nudust, nl);
  if (!rc) rc = upload(ctx, &ctx->kappad, kappad, nl);
  if (!rc) rc = upload(ctx, &ctx->adust, adust, nl);
  if (!rc) rc = upload(ctx, &ctx->gdust, gdust, nl);
  if (!rc) rc = upload(ctx, &ctx->pldust, pldust, nl);
  if (!rc) rc = upload(ctx, &ctx->lTint, lTint, (size_t)d.nT);
  if (!rc) rc = upload(ctx, &ctx->nuint, nuint, (size_t)d.nnu);
  if (!rc) rc = upload(ctx, &ctx->kapBint, kapBint, (size_t)d.nT * d.nnu * d.ndg);
  if (!rc) rc = upload(ctx, &ctx->kappap, kappap, (size_t)d.nT * d.ndg);
  if (rc) return rc;
  // guide table: for log-frequency bin b of species id, the bracket index i1
  // (1-based, nud[i1-1] < nu) of the bin's largest frequency
  std::vector<int32_t> guide((size_t)kMaxSpecies * kGuide, 2);
  for (int id = 0; id < d.ndg; ++id) {
    const double *nud = nudust + (size_t)d.nlammax * id;
    const int nl = nlambda[id];
    for (int i = 1; i < nl; ++i)
      if (!(nud[i] < nud[i - 1]))
        return fail(ctx, HC_E_ARG, "hc_set_dust: frequencies must decrease strictly");
    if (nl < 2) continue;
    const double L = std::log(nud[0] / nud[nl - 1]);
    ctx->glog0[id] = std::log(nud[0]);
    ctx->ginv[id] = (double)kGuide / L;
    int i1 = 2;
    for (int b = 0; b < kGuide; ++b) {
      const double nub = nud[0] * std::exp(-(double)b * L / (double)kGuide) * (1.0 + 1e-9);
      while (i1 < nl && !(nub > nud[i1 - 1])) ++i1;
      guide[(size_t)id * kGuide + b] = std::max(2, i1 - 1);
    }
  }
  rc = upload(ctx, &ctx->guide, guide.data(), guide.size());
  if (rc) return rc;
  rc = build_cdf_guide(ctx, kapBint, &ctx->cdfguide);
  if (rc) return rc;
  ctx->have_dust = true;
  return HC_OK;
}

int hc_set_sources(hc_ctx *ctx, const hc_sources *src) {
  if (!ctx || !src) return HC_E_ARG;
  const hc_dims &d = ctx->dims;
  if (!src->u || !src->aperture2 || (d.npeel > 0 && (!src->coste || !src->sinte || !src->cospe ||
                                                     !src->sinpe)))
    return fail(ctx, HC_E_ARG, "hc_set_sources: null array");
  if (src->limb != 0 && src->limb != 1)
    return fail(ctx, HC_E_ARG, "hc_set_sources: limb must be 0 (isotropic) or 1 (Eddington)");
  if (src->accfrac < 0.0 || src->accfrac >= 1.0 || src->accfrac2 < 0.0 || src->accfrac2 >= 1.0)
    return fail(ctx, HC_E_ARG, "hc_set_sources: accfrac and accfrac2 must lie in [0, 1)");
  if (src->accfrac > 0.0 &&
      !(src->rddustmin >= 1.0 && src->rmaxdmax > src->rddustmin && src->rmax > 1.0 &&
        src->massc > 0.0 && src->mdot >= 0.0))
    return fail(ctx, HC_E_ARG,
                "hc_set_sources: disk accretion needs 1 <= rddustmin < rmaxdmax, rmax > 1, "
                "massc > 0, mdot >= 0");
  if (src->ispot == 1) {
    // spotset's geometry: cos/sin of the spot's colatitude and their
    // products with cos/sin of its longitude; the launch's rejection loop
    // only ends for a spot that lies on the star
    auto unit = [](double c, double s, double sc, double ss) {
      return std::fabs(c * c + s * s - 1.0) < 1e-6 && std::fabs(sc * sc + ss * ss - s * s) < 1e-6;
    };
    if (!(src->nspot == 1 || src->nspot == 2) || !(src->csprad > -1.0 && src->csprad < 1.0) ||
        !unit(src->cspot1, src->sspot1, src->scspot1, src->ssspot1) ||
        (src->nspot == 2 && !unit(src->cspot2, src->sspot2, src->scspot2, src->ssspot2)))
      return fail(ctx, HC_E_ARG,
                  "hc_set_sources: hot spots need nspot 1 or 2, -1 < csprad < 1 and spot "
                  "directions with cspot^2 + sspot^2 = 1 and scspot^2 + ssspot^2 = sspot^2");
  }
  HCK(hipSetDevice(ctx->device));
  std::vector<double> sini(d.nmu);
  for (int i = 0; i < d.nmu; ++i) sini[i] = std::sqrt(1.0 - src->u[i] * src->u[i]);
  int rc = upload(ctx, &ctx->u, src->u, (size_t)d.nmu);
  if (!rc) rc = upload(ctx, &ctx->sini, sini.data(), (size_t)d.nmu);
  if (!rc) rc = upload(ctx, &ctx->aperture2, src->aperture2, (size_t)d.nap);
  if (d.npeel > 0) {
    if (!rc) rc = upload(ctx, &ctx->coste, src->coste, (size_t)d.npeel);
    if (!rc) rc = upload(ctx, &ctx->sinte, src->sinte, (size_t)d.npeel);
    if (!rc) rc = upload(ctx, &ctx->cospe, src->cospe, (size_t)d.npeel);
    if (!rc) rc = upload(ctx, &ctx->sinpe, src->sinpe, (size_t)d.npeel);
  }
  if (rc) return rc;
  ctx->src = *src;
  ctx->have_src = true;
  return HC_OK;
}

}  // extern "C"

namespace {

// One transport pass, ready to launch: kernel, parameters, launch shape and
// the sizes of its tallies.
typedef void (*TransportKernel)(Params, const Params *, const uint32_t *);
struct PassLaunch {
  TransportKernel kern = nullptr;
  Params P;
  size_t shm = 0;
  unsigned blocks = 0;
  int block = 0;
  bool cube = false;
  size_t nc = 0, nsed = 0, npeelsed = 0, nimg = 0;
};

__global__ void zero_words(uint32_t *p, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

// Zero `bytes` (a multiple of 4) of device memory on the context's stream.
int zero_bytes(hc_ctx *ctx, void *p, size_t bytes) {
  const size_t n = bytes / 4;
  if (!p || n == 0) return HC_OK;
  const unsigned blocks = (unsigned)std::min<size_t>((n + 255) / 256, 2048);
  hipLaunchKernelGGL(zero_words, dim3(blocks), dim3(256), 0, ctx->stream,
                     static_cast<uint32_t *>(p), n);
  HCK(hipGetLastError());
  return HC_OK;
}

// Tallies allocated, launch parameters and LDS layout of one pass; queues no
// device work.
int prepare_pass(hc_ctx *ctx, int64_t np, int32_t ipeel, int32_t seed1, int32_t seed2,
                 int32_t gen, uint64_t seg0, bool trace, PassLaunch &L) {
  if (!ctx->have_grid || !ctx->have_dust || !ctx->have_src)
    return fail(ctx, HC_E_STATE, "transport pass: grid, dust and sources must be set");
  if (ctx->src.limb && ctx->ntab == 0)
    return fail(ctx, HC_E_STATE, "transport pass: limb darkening needs hc_set_limb_table");
  if (np < 0 || gen < 1 || gen > 32 || seed1 <= 0 || seed1 >= 2147483563 || seed2 <= 0 ||
      seed2 >= 2147483399)
    return fail(ctx, HC_E_ARG, "transport pass: bad np or seeds");
  const hc_dims &d = ctx->dims;
  HCK(hipSetDevice(ctx->device));
  const size_t nc = ctx->ncellw * (size_t)ctx->nact;
  const size_t nsed = (size_t)9 * d.nfreq * d.nmu * d.nph * d.nap * kNo;
  const size_t npeelsed = (size_t)9 * d.nfreq * std::max(d.npeel, 1) * d.nap * kNo;
  const size_t nimg = (size_t)std::max(d.npeel, 1) * d.nxhst * d.nxhst * d.nfreq;
  int rc = 0;
  rc = rc ? rc : dalloc(ctx, &ctx->dtauabs, nc);
  rc = rc ? rc : dalloc(ctx, &ctx->nabs, nc);
  rc = rc ? rc : dalloc(ctx, &ctx->jmean, ctx->ncellw);
  rc = rc ? rc : dalloc(ctx, &ctx->killed, ctx->ncellw);
  rc = rc ? rc : dalloc(ctx, &ctx->dbig, ctx->ncellw);
  rc = rc ? rc : dalloc(ctx, &ctx->sed, nsed);
  rc = rc ? rc : dalloc(ctx, &ctx->peel, npeelsed);
  rc = rc ? rc : dalloc(ctx, &ctx->si_init, (size_t)d.nfreq * 5);
  if (d.imcube && ipeel) {
    // the FP32 cube (imcube 2) takes half the words of the FP64 one
    const size_t wi = d.imcube == 2 ? (nimg * kNo + 1) / 2 : nimg * kNo;
    const size_t wq = d.imcube == 2 ? (nimg + 1) / 2 : nimg;
    rc = rc ? rc : dalloc(ctx, &ctx->img_i, wi);
    rc = rc ? rc : dalloc(ctx, &ctx->img_q, wq);
    rc = rc ? rc : dalloc(ctx, &ctx->img_u, wq);
    rc = rc ? rc : dalloc(ctx, &ctx->img_v, wq);
  }
  if (rc) return rc;
  if (trace && (size_t)np > ctx->trace_cap) {
    if (ctx->trace) HCK(hipFree(ctx->trace));
    ctx->trace = nullptr;
    rc = dalloc(ctx, &ctx->trace, (size_t)np);
    if (rc) return rc;
    ctx->trace_cap = (size_t)np;
  }
  const bool cube = d.imcube && ipeel && ctx->img_i;
  Params &P = L.P;
  std::memset(&P, 0, sizeof P);
  P.nrg = d.nrg;
  P.ntg = d.ntg;
  P.npg = d.npg;
  P.nact = ctx->nact;
  for (int i = 0; i < kMaxSpecies; ++i) {
    P.act[i] = ctx->act[i];
    P.nlambda[i] = ctx->nlambda[i];
    P.kappav[i] = ctx->kappav[i];
    P.glog0[i] = ctx->glog0[i];
    P.ginv[i] = ctx->ginv[i];
  }
  P.r2arr = ctx->r2arr;
  P.tan2arr = ctx->tan2arr;
  P.thetarr = ctx->thetarr;
  P.aarr = ctx->aarr;
  P.barr = ctx->barr;
  P.phiarr = ctx->phiarr;
  P.dens = ctx->dens;
  P.densm = ctx->densm;
  P.tdust = ctx->tdust;
  P.nlammax = d.nlammax;
  P.nT = d.nT;
  P.nnu = d.nnu;
  P.nudust = ctx->nudust;
  P.kappad = ctx->kappad;
  P.adust = ctx->adust;
  P.gdust = ctx->gdust;
  P.pldust = ctx->pldust;
  P.lTint = ctx->lTint;
  P.nuint = ctx->nuint;
  P.kapBint = ctx->kapBint;
  P.guide = ctx->guide;
  P.cdfguide = ctx->cdfguide;
  const hc_sources &S = ctx->src;
  P.rstar = S.rstar;
  P.rmincgs = S.rstar * kRsol;
  P.tstar = S.tstar;
  P.tshock = S.tshock;
  P.ips = S.ips;
  P.ispot = S.ispot;
  P.nspot = S.nspot;
  P.ipeel = ipeel && d.npeel > 0;
  P.fspot = S.fspot;
  P.csprad = S.csprad;
  P.cspot1 = S.cspot1;
  P.sspot1 = S.sspot1;
  P.scspot1 = S.scspot1;
  P.ssspot1 = S.ssspot1;
  P.cspot2 = S.cspot2;
  P.sspot2 = S.sspot2;
  P.scspot2 = S.scspot2;
  P.ssspot2 = S.ssspot2;
  // source split (newdisktherm.F90:143-176; no external illumination)
  long long ns = (long long)((double)np * (1.0 - S.accfrac));
  const double rn = (double)(float)np * (1.0 - S.accfrac);
  if (rn - (double)ns > 0.5) ns = ns + 1;
  P.npacc = (long long)((double)np * S.accfrac);
  P.npaccstar = (long long)(S.accfrac2 * (double)np);
  if (ns + P.npacc != np) ns = ns + np - (ns + P.npacc);
  P.nstar = (double)ns;
  P.np = np;
  P.nfreq = d.nfreq;
  P.nmu = d.nmu;
  P.nph = d.nph;
  P.nap = d.nap;
  P.npeel = d.npeel;
  P.nxhst = d.nxhst;
  P.imcube = cube ? d.imcube : 0;
  P.iintmax = S.iintmax > 0 ? S.iintmax : 10000;
  P.numin = S.numin;
  P.lnurat = S.lnurat;
  P.rmaxi = S.rmaxi;
  P.fractxh = 2.0 * S.rmaxi / (double)d.nxhst;
  P.u = ctx->u;
  P.sini = ctx->sini;
  P.aperture2 = ctx->aperture2;
  P.coste = ctx->coste;
  P.sinte = ctx->sinte;
  P.cospe = ctx->cospe;
  P.sinpe = ctx->sinpe;
  // generator gen's initial seeds (set_all_seeds, ecuyer_cote_mod.f90:345-362)
  P.g1 = (int32_t)((powmod_host(2082007225, (uint64_t)(gen - 1), 2147483563) * seed1) % 2147483563);
  P.g2 = (int32_t)((powmod_host(784306273, (uint64_t)(gen - 1), 2147483399) * seed2) % 2147483399);
  P.seg0 = seg0;
  P.pow1 = ctx->pow1;
  P.pow2 = ctx->pow2;
  // the Lucy tallies always accumulate on the device: the host arrays in
  // `out` only choose what is downloaded afterwards
  P.dtauabs = ctx->dtauabs;
  P.nabs = ctx->nabs;
  P.jmean = ctx->jmean;
  P.killed = ctx->killed;
  P.dbig = ctx->dbig;
  P.sed = ctx->sed;
  P.peel = P.ipeel ? ctx->peel : nullptr;
  P.img_i = cube ? ctx->img_i : nullptr;
  P.imcube = cube ? d.imcube : 0;
  P.img_q = ctx->img_q;
  P.img_u = ctx->img_u;
  P.img_v = ctx->img_v;
  P.si_init = ctx->si_init;
  P.scal = ctx->scal;
  P.next_packet = ctx->counter;
  P.trace = trace ? ctx->trace : nullptr;
  P.regen_lanes = ctx->regen_lanes;
  P.regen_run_lanes = ctx->regen_run_lanes;
  P.max_work = ctx->max_work;
  P.costarr = ctx->costarr;
  P.rddustmin = S.rddustmin;
  P.rmaxdmax = S.rmaxdmax;
  P.rmax = S.rmax;
  P.massc = S.massc;
  P.mdot = S.mdot;
  for (int i = 0; i < kMaxSpecies; ++i) {
    P.fmass[i] = i < 8 ? S.fmass[i] : 0.0;
    P.z1[i] = i < 8 ? S.z1[i] : 0.0;
    P.bflare[i] = i < 8 ? S.bflare[i] : 0.0;
    P.slot_of[i] = ctx->slot[i];
  }
  P.nhnu = ctx->nhnu;
  P.atm = ctx->atm;
  P.atm_guide = ctx->atm_guide;
  P.limb = S.limb;
  P.ntab = ctx->ntab;
  P.limb_tab = ctx->limb_tab;
  P.delmu = ctx->ntab > 0 ? 1.0 / (double)ctx->ntab : 0.0;
  if (ctx->bw && !ctx->mass4)
    return fail(ctx, HC_E_STATE, "transport pass: Bjorkman-Wood re-emission needs hc_set_bw");
  P.bw = ctx->bw;
  P.serial = ctx->serial;
  P.kapdBint = ctx->kapdBint;
  P.cdfguide_bw = ctx->cdfguide_bw;
  P.kappap = ctx->kappap;
  P.mass4 = ctx->mass4;
  P.tdust_w = ctx->tdust;
  P.densm_w = ctx->densm;
  P.tsub_ev = ctx->tsub / 11605.0;
  P.ltot = S.ltot;
  P.nphot = (double)np;
  if (P.ipeel && ctx->nfilt > 0) {
    const size_t words = (size_t)4 * std::max(d.npeel, 1) * d.nxhst * d.nxhst * ctx->nfilt;
    if (ctx->imgb && ctx->imgb_words < words) {
      HCK(hipFree(ctx->imgb));
      ctx->imgb = nullptr;
    }
    int rc2 = dalloc(ctx, &ctx->imgb, words);
    if (rc2) return rc2;
    ctx->imgb_words = words;
    P.nfilt = ctx->nfilt;
    P.nwfiltmax = ctx->nwfiltmax;
    P.nwfilt = ctx->nwfilt;
    P.filtwave = ctx->filtwave;
    P.filtphi = ctx->filtphi;
    P.imgb = ctx->imgb;
  }
  if (ctx->diffus) {
    if (ctx->bw)
      return fail(ctx, HC_E_UNSUPPORTED,
                  "transport pass: the diffusion approximation runs with Lucy re-emission only");
    P.diffus = ctx->diffus;
    P.diffdir = ctx->diffdir;
    P.rarr = ctx->rarr;
    P.kappar = ctx->kappar;
  }
  P.sintarr = ctx->sintarr;

  TransportKernel &kern = L.kern;
  size_t rec_bytes = 0;
  const bool fp32 = d.precision == HC_FP32;
  const bool pk = P.ipeel != 0;
  // dynamic LDS: the wall arrays, then (peel-off passes) four wavefront
  // queues of peel records; two workgroups share a compute unit
  const size_t walls =
      ((size_t)(2 * d.nrg + 3 * d.ntg + 2 * d.npg) * sizeof(double) + 63) / 64 * 64;
  P.walls_lds = (int)walls;
  size_t ntab = 0;  // doubles of the active species' opacity tables
  for (int k = 0; k < ctx->nact; ++k) {
    P.opac_off[k] = (int)ntab;
    ntab += 5 * (size_t)ctx->nlambda[ctx->act[k] - 1];
  }
  P.opac_base = (int)walls;
  P.guide_off = (int)((ntab * sizeof(double) + 15) / 16 * 16);
  const size_t opac =
      ((size_t)P.guide_off + (size_t)ctx->nact * kGuide * sizeof(int32_t) + 63) / 64 * 64;
  P.emis_base = (int)(walls + opac);
  const size_t emis = ((size_t)(d.nT + d.nnu) * sizeof(double) + 63) / 64 * 64;
  P.peel_base = (int)(walls + opac + emis);
  size_t shm = walls + opac + emis;
  // Lucy passes pool their scattering and absorption events over a 512-lane
  // workgroup when the pool fits beside the tables
  const size_t pool_bytes = (size_t)kPoolFields * kPoolLanes * sizeof(double) +
                            (size_t)3 * kPoolLanes * sizeof(uint32_t) + 4 * sizeof(uint32_t);
  // (switched off: hc_set_option refuses "pool" 1, and no pooled kernel
  // is instantiated)
  const bool pool = false;
  (void)pool_bytes;
  P.pool_base = (int)shm;
  P.pool_min = ctx->pool_min;
  P.pool_walk = ctx->pool_walk;
#define HC_PICK(R_, NA_)                                                                    \
  kern = trace ? (pk ? transport<R_, NA_, true, true, false>                                 \
                     : transport<R_, NA_, true, false, false>)                               \
               : (pk ? transport<R_, NA_, false, true, false>                                \
                     : transport<R_, NA_, false, false, false>);                             \
  rec_bytes = sizeof(PeelRec<R_, NA_>)
  if (ctx->nact <= 4) {
    if (fp32) {
      HC_PICK(float, 4);
    } else {
      HC_PICK(double, 4);
    }
  } else {
    if (fp32) {
      HC_PICK(float, 8);
    } else {
      HC_PICK(double, 8);
    }
  }
#undef HC_PICK
  P.peel_q = 0;
  int block = pk ? kBlockPeel : (pool ? kPoolLanes : kBlockLucy);
  if (pool) shm += pool_bytes;
  if (pk) {
    // one queue per wavefront, four wavefronts per workgroup
    const size_t budget = ctx->lds_per_block;
    const size_t avail = budget > shm + 64 ? budget - shm - 64 : 0;
    // 128 by default: the depth the autotuner picked on the 200x199x200-cell
    // model's peeled pass (1.08x and 1.11x the 96-record queues)
    const size_t want = ctx->peel_queue > 0 ? (size_t)ctx->peel_queue : 128;
    const size_t nq = kBlockPeel / 64;
    const size_t q = std::min<size_t>(want, avail / (nq * rec_bytes));
    if (q < 65)
      return fail(ctx, HC_E_UNSUPPORTED,
                  "transport pass: the peel-off queues do not fit the LDS left by this grid's "
                  "wall arrays and dust tables");
    P.peel_q = (int)q;
    shm += 64 + nq * q * rec_bytes;
    block = 64 * (int)nq;
  }
  if (shm > ctx->lds_per_block)
    return fail(ctx, HC_E_UNSUPPORTED,
                "transport pass: the grid's wall arrays and the dust tables exceed "
                "the LDS of a workgroup");
  if (shm > 64 * 1024)
    HCK(hipFuncSetAttribute(reinterpret_cast<const void *>(kern),
                            hipFuncAttributeMaxDynamicSharedMemorySize, (int)shm));
  int dev_cus = 0, bpm = 0;
  HCK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  HCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpm, kern, block, shm));
  const int per_cu = ctx->blocks_per_cu > 0 ? std::min(ctx->blocks_per_cu, std::max(bpm, 1))
                                            : std::max(bpm, 1);
  long long blocks = (long long)std::max(dev_cus, 1) * per_cu;
  blocks = std::min<long long>(blocks, std::max<long long>((np + block - 1) / block, 1));
  if (ctx->serial) blocks = 1;
  L.shm = shm;
  L.blocks = (unsigned)blocks;
  L.block = block;
  L.cube = cube;
  L.nc = nc;
  L.nsed = nsed;
  L.npeelsed = npeelsed;
  L.nimg = nimg;
  return HC_OK;
}

// The pass's tallies and counters to zero, with kernels (a graph captured
// from the stream then holds no memset nodes).
int zero_tallies(hc_ctx *ctx, const PassLaunch &L, int64_t np, bool trace) {
  const hc_dims &d = ctx->dims;
  const size_t nc = L.nc;
  int rc = HC_OK;
  auto z = [&](void *p, size_t bytes) {
    if (!rc) rc = zero_bytes(ctx, p, bytes);
  };
  z(ctx->dtauabs, nc * sizeof(double));
  z(ctx->nabs, nc * sizeof(double));
  z(ctx->jmean, ctx->ncellw * sizeof(double));
  z(ctx->killed, ctx->ncellw * sizeof(int32_t));
  z(ctx->dbig, ctx->ncellw * sizeof(int32_t));
  z(ctx->sed, L.nsed * sizeof(double));
  z(ctx->peel, L.npeelsed * sizeof(double));
  z(ctx->si_init, (size_t)d.nfreq * 5 * sizeof(double));
  z(ctx->scal, S_COUNT * sizeof(double));
  z(ctx->counter, kCounters * sizeof(unsigned long long));
  if (trace) z(ctx->trace, (size_t)np * sizeof(hc_trace));
  // only a peeled pass with filters refills the filter images; the
  // images of the last such pass stay readable across Lucy passes
  if (L.P.nfilt > 0) {
    z(ctx->imgb, ctx->imgb_words * sizeof(double));
    ctx->imgb_valid = true;
  }
  if (L.cube) {
    const size_t w = d.imcube == 2 ? sizeof(float) : sizeof(double);
    z(ctx->img_i, L.nimg * kNo * w);
    z(ctx->img_q, L.nimg * w);
    z(ctx->img_u, L.nimg * w);
    z(ctx->img_v, L.nimg * w);
  }
  return rc;
}

// The scalar tallies of the last pass into out.
int download_scalars(hc_ctx *ctx, hc_tallies *out) {
  double scal[S_COUNT];
  HC_D2H(scal, ctx->scal, sizeof scal);
  out->tot = scal[S_TOT];
  out->flux = scal[S_FLUX];
  out->sflux = scal[S_SFLUX];
  out->nscat = scal[S_NSCAT];
  out->scount = scal[S_SCOUNT];
  out->abflux = scal[S_ABFLUX];
  out->aflux = scal[S_AFLUX];
  out->sumsub = scal[S_SUMSUB];
  out->sumall = scal[S_SUMALL];
  out->hit = scal[S_HIT];
  out->htot = scal[S_HTOT];
  out->n_err_nowall = (int64_t)scal[S_ERR_WALL];
  out->n_err_dust = (int64_t)scal[S_ERR_DUST];
  out->n_warn = (int64_t)scal[S_WARN];
  out->ncross = scal[S_NCROSS];
  out->nscat_ev = scal[S_NSCAT_EV];
  out->nabs_ev = scal[S_NABS_EV];
  out->npeelray = scal[S_NPEELRAY];
  out->npeelcross = scal[S_NPEELCROSS];
  out->ntally = scal[S_NTALLY];
  return HC_OK;
}

__global__ void copy_words(const double *src, double *dst, size_t n) {
  size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (; i < n; i += (size_t)gridDim.x * blockDim.x) dst[i] = src[i];
}

__global__ void advance_segment(unsigned long long *seg, unsigned long long np) {
  if (threadIdx.x == 0) *seg += np;
}

// Device temperature update (tfinal): mass4 uploaded when given, tdust2
// allocated, kernel parameters.
int prepare_temperature(hc_ctx *ctx, const double *massarr, double ltot, double nphot,
                        TParams &T) {
  if (!ctx->have_grid || !ctx->have_dust || !ctx->dtauabs)
    return fail(ctx, HC_E_STATE, "temperature update needs a completed pass");
  if (!massarr && !ctx->mass4)
    return fail(ctx, HC_E_ARG, "temperature update: massarr required on the first call");
  const hc_dims &d = ctx->dims;
  HCK(hipSetDevice(ctx->device));
  const size_t nc = ctx->ncellw * (size_t)ctx->nact;
  int rc = 0;
  if (massarr) {
    std::vector<double> tmp;
    to_device_layout(ctx, massarr, tmp);
    rc = upload(ctx, &ctx->mass4, tmp.data(), nc);
    if (rc) return rc;
  }
  rc = dalloc(ctx, &ctx->tdust2, nc);
  if (rc) return rc;
  std::memset(&T, 0, sizeof T);
  T.nrg = d.nrg;
  T.ntg = d.ntg;
  T.npg = d.npg;
  T.nact = ctx->nact;
  T.nT = d.nT;
  for (int i = 0; i < kMaxSpecies; ++i) {
    T.act[i] = ctx->act[i];
    T.kappav[i] = ctx->kappav[i];
  }
  T.lTint = ctx->lTint;
  T.kappap = ctx->kappap;
  T.dtauabs = ctx->dtauabs;
  T.mass4 = ctx->mass4;
  T.tdust = ctx->tdust;
  T.tdust2 = ctx->tdust2;
  T.ltot = ltot;
  T.nphot = nphot;
  T.nchanged = ctx->counter;
  return HC_OK;
}

// Queue the update on the stream (no synchronisation, so it can be captured):
// the changed-cell counter zeroed, the new temperatures into tdust2 and, with
// apply, copied into tdust and the sublimation mask refreshed.
int queue_temperature(hc_ctx *ctx, const TParams &T, bool apply) {
  const size_t nc = ctx->ncellw * (size_t)ctx->nact;
  int rc = zero_bytes(ctx, ctx->counter, sizeof(unsigned long long));
  if (rc) return rc;
  const unsigned blocks = (unsigned)std::min<size_t>((nc + 255) / 256, 4096);
  hipLaunchKernelGGL(lucy_temperature, dim3(std::max(blocks, 1u)), dim3(256), 0, ctx->stream, T);
  HCK(hipGetLastError());
  if (apply) {
    hipLaunchKernelGGL(copy_words, dim3(std::max(std::min(blocks, 2048u), 1u)), dim3(256), 0,
                       ctx->stream, (const double *)ctx->tdust2, ctx->tdust, nc);
    HCK(hipGetLastError());
    if (ctx->diffus) {  // initarr.f90:78-84: the layer restarts at 0.1 K
      hipLaunchKernelGGL(reset_layer, dim3(std::max(std::min(blocks, 2048u), 1u)), dim3(256), 0,
                         ctx->stream, (const int8_t *)ctx->diffus, ctx->tdust, ctx->ncellw,
                         ctx->nact);
      HCK(hipGetLastError());
    }
    rc = queue_mask(ctx);
  }
  return rc;
}

// The scalars, the longest packet and the arrays out points at from the last
// pass of the context (L as prepare_pass laid it out).
int download_pass(hc_ctx *ctx, const PassLaunch &L, int64_t np, hc_tallies *out,
                  hc_trace *trace) {
  const hc_dims &d = ctx->dims;
  const size_t nc = L.nc, nsed = L.nsed, npeelsed = L.npeelsed, nimg = L.nimg;
  const bool cube = L.cube;
  int rc = download_scalars(ctx, out);

  if (rc) return rc;
  unsigned long long counters[2] = {0, 0};
  HC_D2H(counters, ctx->counter, sizeof counters);
  out->max_packet_work = (double)counters[1];
  std::vector<double> tmp;
  auto per_species = [&](double *dev, double *host) -> int {
    if (!host) return HC_OK;
    tmp.resize(nc);
    HC_D2H(tmp.data(), dev, nc * sizeof(double));
    std::memset(host, 0, ctx->ncellw * (size_t)d.ndg * sizeof(double));
    for (int k = 0; k < ctx->nact; ++k) {
      double *dst = host + ctx->ncellw * (size_t)(ctx->act[k] - 1);
      for (size_t c = 0; c < ctx->ncellw; ++c) dst[c] = tmp[c * (size_t)ctx->nact + k];
    }
    return HC_OK;
  };
  if ((rc = per_species(ctx->dtauabs, out->dtauabs))) return rc;
  if ((rc = per_species(ctx->nabs, out->nabs))) return rc;
  if (out->jmean)
    HC_D2H(out->jmean, ctx->jmean, ctx->ncellw * sizeof(double));
  if (out->killed)
    HC_D2H(out->killed, ctx->killed, ctx->ncellw * sizeof(int32_t));
  if (out->dbig)
    HC_D2H(out->dbig, ctx->dbig, ctx->ncellw * sizeof(int32_t));
  if (out->dsg) std::memset(out->dsg, 0, ctx->ncellw * sizeof(int32_t));
  if (out->sed) HC_D2H(out->sed, ctx->sed, nsed * sizeof(double));
  if (out->peel && d.npeel > 0)
    HC_D2H(out->peel, ctx->peel, npeelsed * sizeof(double));
  if (out->si_init)
    HC_D2H(out->si_init, ctx->si_init, (size_t)d.nfreq * 5 * sizeof(double));
  if (cube) {
    // FP32 cube (imcube 2): the caller's img arrays are real*4
    const size_t w = d.imcube == 2 ? sizeof(float) : sizeof(double);
    if (out->img_i)
      HC_D2H(out->img_i, ctx->img_i, nimg * kNo * w);
    if (out->img_q)
      HC_D2H(out->img_q, ctx->img_q, nimg * w);
    if (out->img_u)
      HC_D2H(out->img_u, ctx->img_u, nimg * w);
    if (out->img_v)
      HC_D2H(out->img_v, ctx->img_v, nimg * w);
  }
  if (trace && np > 0)
    HC_D2H(trace, ctx->trace, (size_t)np * sizeof(hc_trace));
  return HC_OK;
}

}  // namespace

extern "C" {

int hc_run_iteration(hc_ctx *ctx, int64_t np, int32_t ipeel, int32_t seed1, int32_t seed2,
                     int32_t gen, uint64_t seg0, hc_tallies *out, hc_trace *trace) {
  if (!ctx || !out) return HC_E_ARG;
  PassLaunch L;
  int rc = prepare_pass(ctx, np, ipeel, seed1, seed2, gen, seg0, trace != nullptr, L);
  if (!rc) rc = zero_tallies(ctx, L, np, trace != nullptr);
  if (rc) return rc;
  hipStream_t s = ctx->stream;
  HCK(hipEventRecord(ctx->ev0, s));
  if (np > 0) hipLaunchKernelGGL(L.kern, dim3(L.blocks), dim3(L.block), L.shm, s, L.P,
                       (const Params *)nullptr, (const uint32_t *)nullptr);
  HCK(hipGetLastError());
  HCK(hipEventRecord(ctx->ev1, s));
  HCK(hipStreamSynchronize(s));
  float ms = 0.f;
  HCK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  out->kernel_ms = ms;
  return download_pass(ctx, L, np, out, trace);
}
int hc_temperature(hc_ctx *ctx, const double *massarr, double ltot, double nphot, double *tdust2,
                   int64_t *nchanged) {
  if (!ctx) return HC_E_ARG;
  TParams T;
  int rc = prepare_temperature(ctx, massarr, ltot, nphot, T);
  if (!rc) rc = queue_temperature(ctx, T, tdust2 == nullptr);
  if (rc) return rc;
  const hc_dims &d = ctx->dims;
  const size_t nc = ctx->ncellw * (size_t)ctx->nact;
  unsigned long long nch = 0;
  HC_D2H(&nch, ctx->counter, sizeof nch);
  if (tdust2) {
    std::vector<double> tmp(nc);
    HC_D2H(tmp.data(), ctx->tdust2, nc * sizeof(double));
    for (size_t i = 0; i < ctx->ncellw * (size_t)d.ndg; ++i) tdust2[i] = 0.1 / 11605.0;
    for (int k = 0; k < ctx->nact; ++k) {
      double *dst = tdust2 + ctx->ncellw * (size_t)(ctx->act[k] - 1);
      for (size_t c = 0; c < ctx->ncellw; ++c) dst[c] = tmp[c * (size_t)ctx->nact + k];
    }
  }
  if (nchanged) *nchanged = (int64_t)nch;
  return HC_OK;
}

int hc_run_lucy(hc_ctx *ctx, int32_t niter, int64_t np, int32_t seed1, int32_t seed2,
                int32_t gen, uint64_t seg0, const double *massarr, double ltot, hc_tallies *out,
                int64_t *nchanged) {
  if (!ctx || !out) return HC_E_ARG;
  if (niter < 1) return fail(ctx, HC_E_ARG, "hc_run_lucy: niter must be >= 1");
  PassLaunch L;
  int rc = prepare_pass(ctx, np, 0, seed1, seed2, gen, seg0, false, L);
  TParams T;
  if (!rc) rc = prepare_temperature(ctx, massarr, ltot, (double)np, T);
  if (!rc) rc = dalloc(ctx, &ctx->segdev, 1);
  if (rc) return rc;
  L.P.seg_dev = ctx->segdev;
  hipStream_t s = ctx->stream;
  const unsigned long long seg_h = seg0;
  HC_H2D(ctx->segdev, &seg_h, sizeof seg_h);
  // one Lucy iteration as a graph: zero the tallies, the pass, the
  // temperature update applied on the device, the next segment base
  HCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int crc = zero_tallies(ctx, L, np, false);
  if (!crc && np > 0) {
    hipLaunchKernelGGL(L.kern, dim3(L.blocks), dim3(L.block), L.shm, s, L.P,
                       (const Params *)nullptr, (const uint32_t *)nullptr);
    if (hipGetLastError() != hipSuccess) crc = fail(ctx, HC_E_HIP, "hc_run_lucy: pass launch");
  }
  if (!crc) crc = queue_temperature(ctx, T, true);
  if (!crc) {
    hipLaunchKernelGGL(advance_segment, dim3(1), dim3(64), 0, s, ctx->segdev,
                       (unsigned long long)np);
    if (hipGetLastError() != hipSuccess) crc = fail(ctx, HC_E_HIP, "hc_run_lucy: launch");
  }
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(s, &graph);
  if (crc || ce != hipSuccess) {
    if (graph) (void)hipGraphDestroy(graph);
    (void)hipGetLastError();
    return crc ? crc : fail(ctx, HC_E_HIP, std::string("hc_run_lucy: capture: ") +
                                               hipGetErrorString(ce));
  }
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev0, s);
  for (int i = 0; i < niter && e == hipSuccess; ++i) e = hipGraphLaunch(exec, s);
  if (e == hipSuccess) e = hipEventRecord(ctx->ev1, s);
  if (e == hipSuccess) e = hipStreamSynchronize(s);
  if (exec) (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess)
    return fail(ctx, HC_E_HIP, std::string("hc_run_lucy: replay: ") + hipGetErrorString(e));
  float ms = 0.f;
  HCK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  out->kernel_ms = ms;
  rc = download_scalars(ctx, out);
  if (rc) return rc;
  unsigned long long counters[2] = {0, 0};
  HC_D2H(counters, ctx->counter, sizeof counters);
  if (nchanged) *nchanged = (int64_t)counters[0];
  out->max_packet_work = (double)counters[1];
  return HC_OK;
}

}  // extern "C"

namespace {

// Contexts whose stream is the batch's while its graph is captured (their
// zeroing and temperature kernels then join the capture); restored on exit.
struct StreamSwap {
  std::vector<hc_ctx *> ctx;
  std::vector<hipStream_t> saved;
  void swap_to(hc_ctx *c, hipStream_t s) {
    ctx.push_back(c);
    saved.push_back(c->stream);
    c->stream = s;
  }
  ~StreamSwap() {
    for (size_t i = 0; i < ctx.size(); ++i) ctx[i]->stream = saved[i];
  }
};

struct DeviceBuffer {
  void *p = nullptr;
  ~DeviceBuffer() {
    if (p) (void)hipFree(p);
  }
};

}  // namespace

extern "C" {

int hc_run_batch(int32_t nmodels, hc_ctx *const *ctxs, int32_t niter, const int64_t *np,
                 int32_t seed1, int32_t seed2, int32_t gen, const uint64_t *seg0,
                 const double *const *massarr, const double *ltot, int32_t oversubscribe,
                 hc_tallies *out, int64_t *nchanged, double *device_ms) {
  if (nmodels < 1 || !ctxs || !np || !seg0 || !ltot || !out) return HC_E_ARG;
  for (int m = 0; m < nmodels; ++m) {
    if (!ctxs[m]) return HC_E_ARG;
    for (int k = 0; k < m; ++k)
      if (ctxs[k] == ctxs[m]) return fail(ctxs[0], HC_E_ARG, "hc_run_batch: a context twice");
  }
  hc_ctx *ctx = ctxs[0];  // errors are reported here
  if (niter < 1) return fail(ctx, HC_E_ARG, "hc_run_batch: niter must be >= 1");
  if (oversubscribe < 0 || oversubscribe > 64)
    return fail(ctx, HC_E_ARG, "hc_run_batch: oversubscribe must be in 0..64");
  // 64 times the resident workgroups: a model whose packets are few for its
  // cost per packet then gets enough workgroups (the 256-model grid ran its
  // four iterations at 1.80e7 packets/s with 4, 2.13e7 with 16, 2.18e7 with 64)
  const int over = oversubscribe > 0 ? oversubscribe : 64;
  const int n = nmodels;
  std::vector<PassLaunch> L(n);
  std::vector<TParams> T(n);
  for (int m = 0; m < n; ++m) {
    hc_ctx *c = ctxs[m];
    if (c->device != ctx->device)
      return fail(ctx, HC_E_ARG, "hc_run_batch: contexts on different devices");
    int rc = prepare_pass(c, np[m], 0, seed1, seed2, gen, seg0[m], false, L[m]);
    if (!rc) rc = prepare_temperature(c, massarr ? massarr[m] : nullptr, ltot[m],
                                      (double)np[m], T[m]);
    if (!rc) rc = dalloc(c, &c->segdev, 1);
    if (rc) return fail(ctx, rc, "hc_run_batch: model " + std::to_string(m) + ": " + c->err);
    L[m].P.seg_dev = c->segdev;
    const unsigned long long sh = seg0[m];
    {
      hc_ctx *ctx = c;  // HC_H2D reports on this context
      HC_H2D(c->segdev, &sh, sizeof sh);
    }
  }
  HCK(hipSetDevice(ctx->device));
  // one launch per kernel variant (precision, species count) over its models
  std::vector<TransportKernel> kern;
  std::vector<std::vector<int>> members;
  for (int m = 0; m < n; ++m) {
    if (np[m] <= 0) continue;
    size_t g = 0;
    while (g < kern.size() && kern[g] != L[m].kern) ++g;
    if (g == kern.size()) {
      kern.push_back(L[m].kern);
      members.emplace_back();
    }
    members[g].push_back(m);
  }
  int dev_cus = 0;
  HCK(hipDeviceGetAttribute(&dev_cus, hipDeviceAttributeMultiprocessorCount, ctx->device));
  struct Group {
    size_t shm = 0;
    int block = 0;
    size_t table_off = 0, map_off = 0;
    std::vector<uint32_t> map;
  };
  std::vector<Group> grp(kern.size());
  size_t bytes = 0;
  for (size_t g = 0; g < kern.size(); ++g) {
    Group &G = grp[g];
    G.block = L[members[g][0]].block;
    long long npk = 0;
    for (int m : members[g]) {
      G.shm = std::max(G.shm, L[m].shm);
      npk += np[m];
    }
    if (G.shm > 64 * 1024)
      HCK(hipFuncSetAttribute(reinterpret_cast<const void *>(kern[g]),
                              hipFuncAttributeMaxDynamicSharedMemorySize, (int)G.shm));
    int bpm = 0;
    HCK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&bpm, kern[g], G.block, G.shm));
    const long long resident = (long long)std::max(dev_cus, 1) * std::max(bpm, 1);
    // workgroups per model in proportion to its packets, `over` times the
    // workgroups the device holds at once: a workgroup runs its model's
    // packets until that model has none left, and the dispatcher then starts
    // the next ones, which join the models still running (or end at once)
    const double target = (double)over * (double)resident;
    std::vector<std::pair<double, uint32_t>> order;
    for (size_t k = 0; k < members[g].size(); ++k) {
      const long long p = np[members[g][k]];
      const long long cap = (p + G.block - 1) / G.block;
      const long long w = std::max(1LL, std::min(cap, (long long)std::llround(
                                                          target * (double)p / (double)npk)));
      // each model's workgroups spread evenly over the dispatch order
      for (long long j = 0; j < w; ++j) order.push_back({((double)j + 0.5) / (double)w, (uint32_t)k});
    }
    std::stable_sort(order.begin(), order.end(),
                     [](const std::pair<double, uint32_t> &a, const std::pair<double, uint32_t> &b) {
                       return a.first < b.first;
                     });
    G.map.resize(order.size());
    for (size_t i = 0; i < order.size(); ++i) G.map[i] = order[i].second;
    bytes = (bytes + 255) / 256 * 256;
    G.table_off = bytes;
    bytes += members[g].size() * sizeof(Params);
    bytes = (bytes + 255) / 256 * 256;
    G.map_off = bytes;
    bytes += G.map.size() * sizeof(uint32_t);
  }
  DeviceBuffer buf;
  hipStream_t s = ctx->stream;
  if (bytes > 0) {
    HCK(hipMalloc(&buf.p, bytes));
    std::vector<unsigned char> host(bytes, 0);
    for (size_t g = 0; g < kern.size(); ++g) {
      Params *t = reinterpret_cast<Params *>(host.data() + grp[g].table_off);
      for (size_t k = 0; k < members[g].size(); ++k) t[k] = L[members[g][k]].P;
      std::memcpy(host.data() + grp[g].map_off, grp[g].map.data(),
                  grp[g].map.size() * sizeof(uint32_t));
    }
    HC_H2D(buf.p, host.data(), bytes);
  }
  for (int m = 0; m < n; ++m) HCK(hipStreamSynchronize(ctxs[m]->stream));
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipError_t e = hipSuccess;
  {
    StreamSwap swap;
    for (int m = 0; m < n; ++m)
      if (ctxs[m] != ctx) swap.swap_to(ctxs[m], s);
    // one Lucy iteration of every model as a graph: the tallies zeroed, the
    // passes, the temperature updates applied on the device, the next
    // segment bases
    HCK(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
    int crc = 0;
    for (int m = 0; m < n && !crc; ++m) crc = zero_tallies(ctxs[m], L[m], np[m], false);
    unsigned char *base = static_cast<unsigned char *>(buf.p);
    for (size_t g = 0; g < kern.size() && !crc; ++g) {
      const Group &G = grp[g];
      hipLaunchKernelGGL(kern[g], dim3((unsigned)G.map.size()), dim3(G.block), G.shm, s,
                         L[members[g][0]].P, reinterpret_cast<const Params *>(base + G.table_off),
                         reinterpret_cast<const uint32_t *>(base + G.map_off));
      if (hipGetLastError() != hipSuccess) crc = fail(ctx, HC_E_HIP, "hc_run_batch: pass launch");
    }
    for (int m = 0; m < n && !crc; ++m) crc = queue_temperature(ctxs[m], T[m], true);
    for (int m = 0; m < n && !crc; ++m) {
      hipLaunchKernelGGL(advance_segment, dim3(1), dim3(64), 0, s, ctxs[m]->segdev,
                         (unsigned long long)std::max<int64_t>(np[m], 0));
      if (hipGetLastError() != hipSuccess) crc = fail(ctx, HC_E_HIP, "hc_run_batch: launch");
    }
    const hipError_t ce = hipStreamEndCapture(s, &graph);
    if (crc || ce != hipSuccess) {
      if (graph) (void)hipGraphDestroy(graph);
      (void)hipGetLastError();
      if (crc && ctxs[0]->err.empty()) ctx->err = "hc_run_batch: capture";
      return crc ? crc : fail(ctx, HC_E_HIP, std::string("hc_run_batch: capture: ") +
                                                 hipGetErrorString(ce));
    }
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev0, s);
    for (int i = 0; i < niter && e == hipSuccess; ++i) e = hipGraphLaunch(exec, s);
    if (e == hipSuccess) e = hipEventRecord(ctx->ev1, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
  }
  if (exec) (void)hipGraphExecDestroy(exec);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess)
    return fail(ctx, HC_E_HIP, std::string("hc_run_batch: replay: ") + hipGetErrorString(e));
  float ms = 0.f;
  HCK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
  if (device_ms) *device_ms = ms;
  for (int m = 0; m < n; ++m) {
    hc_ctx *c = ctxs[m];
    out[m].kernel_ms = ms;
    const int rc = download_pass(c, L[m], np[m], &out[m], nullptr);
    if (rc) return fail(ctx, rc, "hc_run_batch: model " + std::to_string(m) + ": " + c->err);
    if (nchanged) {
      unsigned long long nch = 0;
      hipError_t de = hipMemcpyAsync(&nch, c->counter, sizeof nch, hipMemcpyDeviceToHost,
                                     c->stream);
      if (de == hipSuccess) de = hipStreamSynchronize(c->stream);
      if (de != hipSuccess) return fail(ctx, HC_E_HIP, "hc_run_batch: download");
      nchanged[m] = (int64_t)nch;
    }
  }
  return HC_OK;
}

int hc_set_atmosphere(hc_ctx *ctx, int32_t nhnu, const double *atmosnu, const double *hnuint,
                      const double *lognu, const double *loghnu) {
  if (!ctx) return HC_E_ARG;
  HCK(hipSetDevice(ctx->device));
  if (nhnu == 0) {  // back to the Planck function at tstar
    ctx->nhnu = 0;
    return HC_OK;
  }
  if (nhnu < 2 || nhnu > 60000 || !atmosnu || !hnuint || !lognu || !loghnu)
    return fail(ctx, HC_E_ARG, "hc_set_atmosphere: need 2 <= nhnu <= 60000 and four arrays");
  for (int i = 1; i < nhnu; ++i)
    if (!(atmosnu[i] > atmosnu[i - 1]) || hnuint[i] < hnuint[i - 1])
      return fail(ctx, HC_E_ARG,
                  "hc_set_atmosphere: atmosnu must increase and hnuint must not decrease");
  if (!(hnuint[0] == 0.0 && hnuint[nhnu - 1] == 1.0))
    return fail(ctx, HC_E_ARG, "hc_set_atmosphere: hnuint must run from 0 to 1");
  std::vector<double> tab(4 * (size_t)nhnu);
  std::copy(atmosnu, atmosnu + nhnu, tab.begin());
  std::copy(hnuint, hnuint + nhnu, tab.begin() + nhnu);
  std::copy(lognu, lognu + nhnu, tab.begin() + 2 * (size_t)nhnu);
  std::copy(loghnu, loghnu + nhnu, tab.begin() + 3 * (size_t)nhnu);
  // guide: the nodes with hnuint below b / kAtmGuide
  std::vector<int32_t> guide(kAtmGuide + 1);
  int cnt = 0;
  for (int b = 0; b <= kAtmGuide; ++b) {
    const double v = (double)b / (double)kAtmGuide;
    while (cnt < nhnu && hnuint[cnt] < v) ++cnt;
    guide[b] = cnt;
  }
  for (void *b : {(void *)ctx->atm, (void *)ctx->atm_guide})
    if (b) (void)hipFree(b);
  ctx->atm = nullptr;
  ctx->atm_guide = nullptr;
  ctx->nhnu = 0;
  int rc = upload(ctx, &ctx->atm, tab.data(), tab.size());
  if (!rc) rc = upload(ctx, &ctx->atm_guide, guide.data(), guide.size());
  if (rc) return rc;
  ctx->nhnu = nhnu;
  return HC_OK;
}

int hc_set_limb_table(hc_ctx *ctx, int32_t ntab, const double *xmu, const double *prob) {
  if (!ctx || ntab < 3 || !xmu || !prob) return HC_E_ARG;
  for (int i = 1; i < ntab; ++i)
    if (!(prob[i] > prob[i - 1]) || !(xmu[i] > xmu[i - 1]))
      return fail(ctx, HC_E_ARG, "hc_set_limb_table: xmu and prob must increase");
  if (prob[ntab - 1] != 1.0)
    return fail(ctx, HC_E_ARG, "hc_set_limb_table: prob must end at 1");
  HCK(hipSetDevice(ctx->device));
  std::vector<double> tab(2 * (size_t)ntab);
  std::copy(xmu, xmu + ntab, tab.begin());
  std::copy(prob, prob + ntab, tab.begin() + ntab);
  if (ctx->limb_tab) (void)hipFree(ctx->limb_tab);
  ctx->limb_tab = nullptr;
  ctx->ntab = 0;
  const int rc = upload(ctx, &ctx->limb_tab, tab.data(), tab.size());
  if (rc) return rc;
  ctx->ntab = ntab;
  return HC_OK;
}

int hc_set_bw(hc_ctx *ctx, const double *kapdBint, const double *massarr) {
  if (!ctx) return HC_E_ARG;
  if (!kapdBint) {
    ctx->bw = 0;
    return HC_OK;
  }
  if (!massarr) return fail(ctx, HC_E_ARG, "hc_set_bw: massarr required with kapdBint");
  if (!ctx->have_grid || !ctx->have_dust)
    return fail(ctx, HC_E_STATE, "hc_set_bw: grid and dust must be set first");
  const hc_dims &d = ctx->dims;
  HCK(hipSetDevice(ctx->device));
  const size_t n = (size_t)d.nT * d.nnu * d.ndg;
  for (int id = 0; id < d.ndg; ++id) {
    const double *last = kapdBint + (size_t)d.nT * d.nnu * id + (size_t)d.nT * (d.nnu - 1);
    for (int j = 0; j < d.nT; ++j)
      if (!(last[j] == 1.0)) return fail(ctx, HC_E_ARG, "hc_set_bw: kapdBint rows must end at 1");
  }
  int rc = upload(ctx, &ctx->kapdBint, kapdBint, n);
  if (!rc) rc = build_cdf_guide(ctx, kapdBint, &ctx->cdfguide_bw);
  if (rc) return rc;
  std::vector<double> tmp;
  to_device_layout(ctx, massarr, tmp);
  rc = upload(ctx, &ctx->mass4, tmp.data(), tmp.size());
  if (rc) return rc;
  ctx->bw = 1;
  return HC_OK;
}

int hc_set_filters(hc_ctx *ctx, int32_t nfilt, int32_t nwfiltmax, const int32_t *nwfilt,
                   const double *filtwave, const double *filtphi) {
  if (!ctx) return HC_E_ARG;
  if (nfilt == 0) {
    ctx->nfilt = 0;
    ctx->imgb_valid = false;
    return HC_OK;
  }
  if (nfilt < 0 || nwfiltmax < 2 || !nwfilt || !filtwave || !filtphi)
    return fail(ctx, HC_E_ARG, "hc_set_filters: bad arguments");
  for (int f = 0; f < nfilt; ++f) {
    const int nf = nwfilt[f] - 1;
    if (nf < 2 || nf > nwfiltmax) return fail(ctx, HC_E_ARG, "hc_set_filters: nwfilt out of range");
    const double *fw = filtwave + (size_t)nwfiltmax * f;
    for (int i = 1; i < nf; ++i)
      if (!(fw[i] > fw[i - 1]))
        return fail(ctx, HC_E_ARG, "hc_set_filters: filter frequencies must increase strictly");
  }
  HCK(hipSetDevice(ctx->device));
  const size_t n = (size_t)nwfiltmax * nfilt;
  for (void **p : {(void **)&ctx->nwfilt, (void **)&ctx->filtwave, (void **)&ctx->filtphi})
    if (*p) {
      HCK(hipFree(*p));
      *p = nullptr;
    }
  int rc = upload(ctx, &ctx->nwfilt, nwfilt, (size_t)nfilt);
  if (!rc) rc = upload(ctx, &ctx->filtwave, filtwave, n);
  if (!rc) rc = upload(ctx, &ctx->filtphi, filtphi, n);
  if (rc) return rc;
  ctx->nfilt = nfilt;
  ctx->nwfiltmax = nwfiltmax;
  ctx->imgb_valid = false;
  return HC_OK;
}

int hc_get_filter_images(hc_ctx *ctx, double *image_b_i, double *image_b_q, double *image_b_u,
                         double *image_b_v) {
  if (!ctx) return HC_E_ARG;
  if (!ctx->imgb_valid || !ctx->imgb)
    return fail(ctx, HC_E_STATE, "hc_get_filter_images: no peeled pass with filters yet");
  const size_t nb = ctx->imgb_words / 4;
  double *out[4] = {image_b_i, image_b_q, image_b_u, image_b_v};
  for (int k = 0; k < 4; ++k)
    if (out[k]) HC_D2H(out[k], ctx->imgb + nb * k, nb * sizeof(double));
  return HC_OK;
}

int hc_set_diffusion(hc_ctx *ctx, const int8_t *diffus, const int8_t *diffdir, const double *rarr,
                     const double *kappar) {
  if (!ctx) return HC_E_ARG;
  for (void **p : {(void **)&ctx->diffus, (void **)&ctx->diffdir, (void **)&ctx->rarr,
                   (void **)&ctx->kappar})
    if (*p) {
      HCK(hipFree(*p));
      *p = nullptr;
    }
  if (!diffus) return HC_OK;
  if (!diffdir || !rarr || !kappar) return fail(ctx, HC_E_ARG, "hc_set_diffusion: null array");
  if (!ctx->have_grid || !ctx->have_dust)
    return fail(ctx, HC_E_STATE, "hc_set_diffusion: set the grid and the dust first");
  if (ctx->slot[0] < 0 || ctx->slot[1] < 0)
    return fail(ctx, HC_E_ARG, "hc_set_diffusion: disk species 1 and 2 must both hold dust");
  const hc_dims &d = ctx->dims;
  for (size_t c = 0; c < ctx->ncellw; ++c) {
    if (!diffus[c]) continue;
    if (diffdir[c] != -1 && diffdir[c] != -2 && diffdir[c] != 2)
      return fail(ctx, HC_E_ARG,
                  "hc_set_diffusion: a layer cell's direction must be -1, -2 or 2");
    // a layer cell lies inside the grid with a cell on each side of it
    const int ir = (int)(c % (size_t)d.nrg) + 1;
    const int it = (int)((c / (size_t)d.nrg) % (size_t)d.ntg) + 1;
    const int ip = (int)(c / ((size_t)d.nrg * d.ntg)) + 1;
    if (ir < 2 || ir > d.nrg - 1 || it < 2 || it > d.ntg - 2 || ip > d.npg - 1)
      return fail(ctx, HC_E_ARG,
                  "hc_set_diffusion: layer cells must have 2 <= ir < nrg and 2 <= it <= ntg-2");
  }
  HCK(hipSetDevice(ctx->device));
  int rc = 0;
  for (auto q : {std::make_pair(&ctx->diffus, diffus), std::make_pair(&ctx->diffdir, diffdir)}) {
    if (rc) break;
    HCK(hipMalloc((void **)q.first, ctx->ncellw));
    HCK(hipMemcpy(*q.first, q.second, ctx->ncellw, hipMemcpyHostToDevice));
  }
  // kappar (nT, ndg) column-major is kappar[id][nT] already
  if (!rc) rc = upload(ctx, &ctx->rarr, rarr, (size_t)d.nrg);
  if (!rc) rc = upload(ctx, &ctx->kappar, kappar, (size_t)d.nT * d.ndg);
  return rc;
}

int hc_dropped_packets(hc_ctx *ctx, int64_t *pids, int32_t cap, int64_t *ndropped) {
  if (!ctx || !ndropped || cap < 0 || (cap > 0 && !pids)) return HC_E_ARG;
  if (!ctx->counter) return fail(ctx, HC_E_STATE, "hc_dropped_packets: no pass has run");
  unsigned long long c[kCounters];
  HC_D2H(c, ctx->counter, sizeof c);
  *ndropped = (int64_t)c[kCtrNDropped];
  const int n = (int)std::min<unsigned long long>(c[kCtrNDropped], (unsigned long long)kDropMax);
  for (int i = 0; i < n && i < cap; ++i) pids[i] = (int64_t)c[kCtrDropped + i];
  return HC_OK;
}

int hc_set_option(hc_ctx *ctx, const char *name, int64_t value) {
  if (!ctx || !name) return HC_E_ARG;
  const std::string n(name);
  if (n == "regen_lanes") {
    if (value < 1 || value > 64) return fail(ctx, HC_E_ARG, "regen_lanes must be in 1..64");
    ctx->regen_lanes = (int)value;
  } else if (n == "regen_run_lanes") {
    if (value < 1 || value > 64) return fail(ctx, HC_E_ARG, "regen_run_lanes must be in 1..64");
    ctx->regen_run_lanes = (int)value;
  } else if (n == "max_work") {
    if (value < 1 || value > 0x7fffffff) return fail(ctx, HC_E_ARG, "max_work must be in 1..2^31-1");
    ctx->max_work = (uint32_t)value;
  } else if (n == "peel_queue") {
    if (value != 0 && (value < 65 || value > 1024))
      return fail(ctx, HC_E_ARG, "peel_queue must be 0 or in 65..1024");
    ctx->peel_queue = (int)value;
  } else if (n == "blocks_per_cu") {
    if (value < 0 || value > 64) return fail(ctx, HC_E_ARG, "blocks_per_cu must be in 0..64");
    ctx->blocks_per_cu = (int)value;
  } else if (n == "serial") {
    if (value < 0 || value > 1) return fail(ctx, HC_E_ARG, "serial must be 0 or 1");
    ctx->serial = (int)value;
  } else if (n == "pool") {
    if (value < 0 || value > 1) return fail(ctx, HC_E_ARG, "pool must be 0 or 1");
    if (value == 1)
      return fail(ctx, HC_E_UNSUPPORTED,
                  "pool: the workgroup event pool is switched off (its first GPU run did not "
                  "finish and no cause has been found)");
    ctx->pool = 0;
  } else {
    return fail(ctx, HC_E_ARG, "hc_set_option: unknown option " + n);
  }
  return HC_OK;
}

int hc_destroy(hc_ctx *ctx) {
  if (!ctx) return HC_OK;
  (void)hipSetDevice(ctx->device);
  void *bufs[] = {ctx->r2arr,  ctx->tan2arr, ctx->thetarr, ctx->aarr,    ctx->barr,   ctx->phiarr,
                  ctx->costarr,
                  ctx->dens,   ctx->densm,   ctx->tdust,   ctx->nudust,  ctx->kappad, ctx->adust,
                  ctx->gdust,  ctx->pldust,  ctx->lTint,   ctx->nuint,   ctx->kapBint, ctx->kappap,
                  ctx->u,      ctx->sini,    ctx->aperture2, ctx->coste, ctx->sinte,  ctx->cospe,
                  ctx->sinpe,  ctx->pow1,    ctx->pow2,    ctx->dtauabs, ctx->nabs,   ctx->jmean,
                  ctx->killed, ctx->dbig,    ctx->sed,     ctx->peel,    ctx->img_i,  ctx->img_q,
                  ctx->img_u,  ctx->img_v,   ctx->si_init, ctx->scal,    ctx->mass4,  ctx->counter,
                  ctx->trace,  ctx->guide,   ctx->tdust2,  ctx->cdfguide, ctx->segdev,
                  ctx->atm,    ctx->atm_guide, ctx->limb_tab, ctx->kapdBint, ctx->cdfguide_bw,
                  ctx->nwfilt, ctx->filtwave, ctx->filtphi, ctx->imgb,
                  ctx->diffus, ctx->diffdir, ctx->rarr, ctx->sintarr, ctx->kappar};
  for (void *b : bufs)
    if (b) (void)hipFree(b);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
  return HC_OK;
}

}  // extern "C"
