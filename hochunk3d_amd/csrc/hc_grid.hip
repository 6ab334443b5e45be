// Cell densities, volumes and (rho V)^(1/4) of the benchmark models on the
// GPU: the density part of gridset (gridset.f90:1002-1083) for two flared
// disks (densdiskg.f90), an Ulrich infalling envelope (densenv.f90, the
// streamline cubic solved by bisection) and a polynomial bipolar cavity
// (envset.f90:25-33), with an optional clump multiplier of the envelope
// (fractal_sph.f90 via gridset.f90:796-870, computed by the caller).
//
// One thread per cell evaluates the analytic densities at the cell centre
// (gridset.f90:494-499, 609-620: log-mean radius, root-mean-square theta) and
// the cell volume; the disk species are normalised to their mass with one
// FP64 sum per species (wavefront partial sums, one atomic per wavefront).
// The outputs use the reference's layout, (nrg,ntg,npg[,ndg]) column-major
// with the last wall index of each dimension zero.
//
// Also here: the dust emissivity tables of setemiss, one thread per
// (temperature, frequency bin), then one per temperature row for the
// cumulative sum; and the clump point counts of the envelope multiplier, one
// thread per point.

#include "../../include/hc.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <vector>

namespace hc_grid {

constexpr double kMsol = 1.989e33;

struct Dev {
  int nrg, ntg, npg, ndisk;
  const double *rarr, *thetarr, *phiarr, *mult;
  double rmincgs;
  int disk_id[8];
  double disk_z1[8], disk_b[8], disk_norm[8];
  double a_exp, rmind, rmaxd;
  double rc, rhoe0, rmine, rmax, rhoamb;
  int env_id, cav_id;
  double c1e, cavity_exp, rho_cavity, r1000;
  double *dens, *vol;  // (nrg,ntg,npg,8) and (nrg,ntg,npg)
  double *mass_sum;    // [8] disk species mass before normalisation
};

// mu0 in [0,1] with mu0^3 + (rx - 1) mu0 - rx mu = 0: 60 bisection steps, as
// the host construction does, so both land on the same root.
__device__ double zerod(double xmu, double rx) {
  double lo = 0.0, hi = 1.0;
  for (int i = 0; i < 60; ++i) {
    const double mid = 0.5 * (lo + hi);
    const double f = mid * mid * mid + (rx - 1.0) * mid - rx * xmu;
    if (f > 0.0)
      hi = mid;
    else
      lo = mid;
  }
  return 0.5 * (lo + hi);
}

__global__ __launch_bounds__(256) void densities(Dev D) {
  const size_t nci = (size_t)(D.nrg - 1), ncj = (size_t)(D.ntg - 1), nck = (size_t)(D.npg - 1);
  const size_t nc = nci * ncj * nck;
  const size_t wall = (size_t)D.nrg * D.ntg * D.npg;
  double part[8];
  for (int d = 0; d < 8; ++d) part[d] = 0.0;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
       c += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % nci);
    const int j = (int)((c / nci) % ncj);
    const int k = (int)(c / (nci * ncj));
    const size_t w = (size_t)i + (size_t)D.nrg * ((size_t)j + (size_t)D.ntg * (size_t)k);
    const double r0 = D.rarr[i], r1 = D.rarr[i + 1];
    const double t0 = D.thetarr[j], t1 = D.thetarr[j + 1];
    const double R = pow(10.0, 0.5 * (log10(r0) + log10(r1)));
    const double th = sqrt(0.5 * (t0 * t0 + t1 * t1));
    const double vol = -(r1 * r1 * r1 - r0 * r0 * r0) * (cos(t1) - cos(t0)) *
                       (D.phiarr[k + 1] - D.phiarr[k]) / 3.0 * D.rmincgs * D.rmincgs * D.rmincgs;
    D.vol[w] = vol;
    const double cost = cos(th), sint = sin(th);
    const double rp = R * sint, zp = R * cost;
    for (int d = 0; d < D.ndisk; ++d) {
      const double zmr = D.disk_z1[d] * pow(rp, D.disk_b[d]);
      const double q = zp / zmr;
      const double gexp = q * q;
      double shape = pow(rp, -D.a_exp) * (gexp < 120.0 ? exp(-0.5 * gexp) : 0.0);
      if (!(rp >= D.rmind && rp <= D.rmaxd)) shape = 0.0;
      D.dens[w + wall * (size_t)D.disk_id[d]] = shape;
      part[d] += shape * vol;
    }
    // Ulrich envelope
    const double rx = R / D.rc;
    const double xmu = fabs(cost);
    const double mu0 = fmax(zerod(xmu, rx), 1e-12);
    const double factor = 1.0 / ((xmu / mu0) + 2.0 * (mu0 * mu0) / rx);
    double env = D.rhoe0 * pow(rx, -1.5) * factor / sqrt(1.0 + xmu / mu0);
    env = R > D.rmine ? fmax(env, D.rhoamb) : 0.0;
    if (!(R < D.rmax)) env = 0.0;
    // polynomial cavity
    const double zup = D.c1e * pow(R * sint, D.cavity_exp);
    const bool inhole = R * xmu > zup;
    const double rr = R / D.r1000;
    double cav = D.rho_cavity / (rr * rr);
    if (!(R > D.rmine && R < D.rmax)) cav = 0.0;
    double e = inhole ? 0.0 : env;
    if (D.mult) e *= D.mult[c];
    D.dens[w + wall * (size_t)D.env_id] = e;
    D.dens[w + wall * (size_t)D.cav_id] = inhole ? cav : 0.0;
  }
  // per-species partial sums: wavefront reduction, one atomic per wavefront
  for (int d = 0; d < D.ndisk; ++d) {
    double v = part[d];
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    if ((threadIdx.x & 63) == 0) unsafeAtomicAdd(&D.mass_sum[d], v);
  }
}

__global__ __launch_bounds__(256) void normalise(Dev D, double *massarr) {
  const size_t nci = (size_t)(D.nrg - 1), ncj = (size_t)(D.ntg - 1), nck = (size_t)(D.npg - 1);
  const size_t nc = nci * ncj * nck;
  const size_t wall = (size_t)D.nrg * D.ntg * D.npg;
  for (size_t c = (size_t)blockIdx.x * blockDim.x + threadIdx.x; c < nc;
       c += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(c % nci);
    const int j = (int)((c / nci) % ncj);
    const int k = (int)(c / (nci * ncj));
    const size_t w = (size_t)i + (size_t)D.nrg * ((size_t)j + (size_t)D.ntg * (size_t)k);
    for (int d = 0; d < D.ndisk; ++d) {
      const double m = D.mass_sum[d];
      double &x = D.dens[w + wall * (size_t)D.disk_id[d]];
      x = m > 0.0 ? x * (D.disk_norm[d] / m) : 0.0;
    }
    const double vol = D.vol[w];
    for (int s = 0; s < 8; ++s) massarr[w + wall * (size_t)s] = pow(D.dens[w + wall * (size_t)s] * vol, 0.25);
  }
}

// setemiss (dust.f90:394-474) on a (temperature, frequency bin) grid: the
// emissivity kappa_abs B_nu / B integrated over each bin of nuint (units of
// kT) with the caller's quadrature rule, kappa_abs linear in nu between the
// bin ends (opacset's linear interpolation of the dust table at each end).
struct Emis {
  int nl, nT, nnu, nq;
  const double *nud, *kd, *ad;  // [nl], nud decreasing
  const double *lTint, *nuint, *xq, *wq;
  double *bins;                 // (nT, nnu) bin integrals, then the table
  double *kappap;               // [nT]
};

// np.interp-style linear interpolation in the dust table at nu (eV),
// clamped to the end values; returns (1 - albedo) kappa.
__device__ double kap_abs(const Emis &E, double nu) {
  // the table decreases in nu: index 0 has the highest frequency
  if (nu >= E.nud[0]) return (1.0 - E.ad[0]) * E.kd[0];
  if (nu <= E.nud[E.nl - 1]) return (1.0 - E.ad[E.nl - 1]) * E.kd[E.nl - 1];
  int lo = 0, hi = E.nl - 1;  // nud[lo] > nu >= nud[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (E.nud[mid] > nu)
      lo = mid;
    else
      hi = mid;
  }
  // increasing abscissa x0 = nud[hi] <= nu < x1 = nud[lo]
  const double x0 = E.nud[hi], x1 = E.nud[lo];
  const double sk = (E.kd[lo] - E.kd[hi]) / (x1 - x0);
  const double sa = (E.ad[lo] - E.ad[hi]) / (x1 - x0);
  const double k = sk * (nu - x0) + E.kd[hi];
  const double a = sa * (nu - x0) + E.ad[hi];
  return (1.0 - a) * k;
}

__device__ double planck_norm(double nu, double T) {  // mathsubs.f90:1-24
  const double x = nu / T;
  if (x == 0.0) return 0.0;
  return 0.1539897 * (x * x * x) / expm1(x);
}

__global__ __launch_bounds__(256) void emis_bins(Emis E) {
  const size_t n = (size_t)E.nT * E.nnu;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (size_t)gridDim.x * blockDim.x) {
    const int i = (int)(t % (size_t)E.nT);
    const int j = (int)(t / (size_t)E.nT);
    const double T = exp(E.lTint[i]);
    const double nu1 = E.nuint[j] * T;
    const double nu0 = j > 0 ? E.nuint[j - 1] * T : 0.0;
    const double dnu = nu1 - nu0;
    const double kap1 = kap_abs(E, nu1), kap0 = kap_abs(E, nu0);
    double acc = 0.0;
    for (int q = 0; q < E.nq; ++q) {
      const double f = E.xq[q];
      acc += E.wq[q] * ((1.0 - f) * kap0 + f * kap1) * planck_norm(nu0 + f * dnu, T);
    }
    E.bins[t] = (dnu / T) * acc;
  }
}

// One thread per temperature row: the cumulative sum in frequency, its total
// (the Planck mean, kappap) and the normalised table with a last node of 1.
__global__ __launch_bounds__(256) void emis_cumulate(Emis E) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= E.nT) return;
  double cum = 0.0;
  for (int j = 0; j < E.nnu; ++j) {
    cum += E.bins[i + (size_t)E.nT * j];
    E.bins[i + (size_t)E.nT * j] = cum;
  }
  E.kappap[i] = cum;
  for (int j = 0; j < E.nnu - 1; ++j) E.bins[i + (size_t)E.nT * j] /= cum;
  E.bins[i + (size_t)E.nT * (E.nnu - 1)] = 1.0;
}

// ---------------------------------------------------------------------------
// Clump points per cell for the envelope multiplier: the construction of
// hochunk3d_amd/fractal.py, whose deviates come from NumPy's PCG64 (128-bit
// LCG, XSL-RR output, doubles from the top 53 bits).  Each thread places one
// leaf point: it jumps the stream to the point's draws (the LCG jumped by
// 2^b steps, b = 0..63, is tabulated once), sums the parent coordinate and
// the level offsets in the host's order and counts the point in its cell
// with the host's comparisons.

struct U128 {
  uint64_t lo, hi;
};

__device__ __forceinline__ U128 mul128(U128 a, U128 b) {  // mod 2^128
  U128 r;
  r.lo = a.lo * b.lo;
  r.hi = __umul64hi(a.lo, b.lo) + a.lo * b.hi + a.hi * b.lo;
  return r;
}

__device__ __forceinline__ U128 add128(U128 a, U128 b) {
  U128 r;
  r.lo = a.lo + b.lo;
  r.hi = a.hi + b.hi + (r.lo < a.lo ? 1u : 0u);
  return r;
}

// jump[b] = (multiplier, increment) of 2^b LCG steps
__global__ void pcg_jump_table(U128 mult, U128 inc, U128 *jump) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  U128 a = mult, c = inc;
  const U128 one = {1u, 0u};
  for (int b = 0; b < 64; ++b) {
    jump[2 * b] = a;
    jump[2 * b + 1] = c;
    c = mul128(c, add128(a, one));
    a = mul128(a, a);
  }
}

__device__ __forceinline__ double pcg_double(U128 s) {
  const unsigned rot = (unsigned)(s.hi >> 58);
  const uint64_t x = s.hi ^ s.lo;
  const uint64_t o = (x >> rot) | (x << ((64u - rot) & 63u));
  return (double)(o >> 11) * 0x1.0p-53;
}

struct Frac {
  int nrg, ntg, npg, n, levels;
  uint64_t npts, nlev;                // n^(levels+1), n^levels
  const double *rarr, *mcos_t, *cph, *sph;
  double scale[8], scale_r;
  uint64_t S;                         // draws per parent
  uint64_t offk[9], div[9];           // first draw of level k; n^(levels-k)
  U128 s0;
  const U128 *jump;
  uint32_t *counts;
};

// the draws i, i+1, i+2 of the stream that starts at s0
__device__ __forceinline__ void draw3(const Frac &F, uint64_t i, double d[3]) {
  U128 s = F.s0;
  const uint64_t k = i + 1;
  for (int b = 0; b < 64; ++b)
    if ((k >> b) & 1u) s = add128(mul128(s, F.jump[2 * b]), F.jump[2 * b + 1]);
  for (int c = 0; c < 3; ++c) {
    d[c] = pcg_double(s);
    s = add128(mul128(s, F.jump[0]), F.jump[1]);
  }
}

// number of entries of a[0..n) below v (lower bound)
__device__ __forceinline__ int count_below(const double *a, int n, double v) {
  int lo = 0, hi = n;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (a[mid] < v)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

__global__ __launch_bounds__(256) void fractal_points(Frac F) {
  extern __shared__ double walls[];  // rarr, mcos_t, cph, sph
  double *rw = walls, *tw = walls + F.nrg, *cw = tw + F.ntg, *sw = cw + F.npg;
  for (int i = threadIdx.x; i < F.nrg; i += blockDim.x) rw[i] = F.rarr[i];
  for (int i = threadIdx.x; i < F.ntg; i += blockDim.x) tw[i] = F.mcos_t[i];
  for (int i = threadIdx.x; i < F.npg; i += blockDim.x) {
    cw[i] = F.cph[i];
    sw[i] = F.sph[i];
  }
  __syncthreads();
  const uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= F.npts) return;
  const uint64_t parent = t / F.nlev, rem = t % F.nlev;
  double d[3], x[3];
  draw3(F, 3 * parent, d);
  for (int c = 0; c < 3; ++c) x[c] = d[c] - 0.5;
  for (int k = 1; k <= F.levels; ++k) {
    const uint64_t node = rem / F.div[k];
    draw3(F, 3 * (uint64_t)F.n + parent * F.S + F.offk[k] + 3 * node, d);
    for (int c = 0; c < 3; ++c) x[c] = x[c] + (d[c] - 0.5) * F.scale[k - 1];
  }
  const double r = __builtin_sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
  double cz = x[2] / r;
  cz = cz < -1.0 ? -1.0 : (cz > 1.0 ? 1.0 : cz);
  int ir = count_below(rw, F.nrg, r * F.scale_r);
  if (ir == 0) ir = 1;
  if (ir > F.nrg - 1) return;
  int it = count_below(tw, F.ntg, -cz);
  if (it < 1) it = 1;
  const double px = x[0], py = x[1];
  const bool lower = (py < 0.0) || (py == 0.0 && px < 0.0);
  int lo = 0, hi = F.npg;  // first wall not below the point's azimuth
  while (lo < hi) {
    const int j = (lo + hi) >> 1;
    const bool hw = 2 * j >= F.npg - 1;
    const double cr = cw[j] * py - sw[j] * px;
    const bool below = (hw != lower) ? (lower && !hw) : (cr > 0.0);
    if (below)
      lo = j + 1;
    else
      hi = j;
  }
  const int ip = lo < 1 ? 1 : lo;
  const size_t idx = (size_t)(ir - 1) +
                     (size_t)F.nrg * ((size_t)(it - 1) + (size_t)F.ntg * (size_t)(ip - 1));
  atomicAdd(&F.counts[idx], 1u);
}

}  // namespace hc_grid

extern "C" int hc_fractal_counts(int device, int32_t nrg, int32_t ntg, int32_t npg,
                                 const double *rarr, const double *mcos_theta,
                                 const double *cos_phi, const double *sin_phi, int32_t n,
                                 int32_t levels, const double *scale, double scale_r,
                                 const uint64_t *pcg, uint32_t *counts) {
  using namespace hc_grid;
  if (nrg < 2 || ntg < 2 || npg < 2 || n < 1 || levels < 0 || levels > 8 || !rarr ||
      !mcos_theta || !cos_phi || !sin_phi || (levels > 0 && !scale) || !pcg || !counts)
    return HC_E_ARG;
  if ((pcg[2] & 1u) == 0) return HC_E_ARG;  // the PCG increment is odd
  Frac F{};
  F.nrg = nrg;
  F.ntg = ntg;
  F.npg = npg;
  F.n = n;
  F.levels = levels;
  // n^(levels+1) points, at most 2^31
  uint64_t p = 1;
  for (int k = 0; k <= levels; ++k) {
    p *= (uint64_t)n;
    if (p > (1ull << 31)) return HC_E_ARG;
  }
  F.npts = p;
  F.nlev = p / (uint64_t)n;
  uint64_t nk = 1, acc = 0;
  for (int k = 1; k <= levels; ++k) {
    F.offk[k] = acc;  // 3 (n + ... + n^(k-1))
    nk *= (uint64_t)n;
    acc += 3 * nk;
    F.scale[k - 1] = scale[k - 1];
  }
  F.S = acc;
  for (int k = 1; k <= levels; ++k) {
    uint64_t dv = 1;
    for (int j = 0; j < levels - k; ++j) dv *= (uint64_t)n;
    F.div[k] = dv;
  }
  F.scale_r = scale_r;
  F.s0 = U128{pcg[0], pcg[1]};
  const size_t lds = (size_t)(nrg + ntg + 2 * npg) * sizeof(double);
  if (lds > 64 * 1024) return HC_E_UNSUPPORTED;
  const size_t ncell = (size_t)nrg * ntg * npg;
  if (hipSetDevice(device) != hipSuccess) return HC_E_HIP;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return HC_E_HIP;
  double *d_w = nullptr;
  U128 *d_jump = nullptr;
  uint32_t *d_cnt = nullptr;
  int rc = HC_OK;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == HC_OK) rc = HC_E_HIP;
    return rc == HC_OK;
  };
  ok(hipMalloc((void **)&d_w, lds)) && ok(hipMalloc((void **)&d_jump, 128 * sizeof(U128))) &&
      ok(hipMalloc((void **)&d_cnt, ncell * sizeof(uint32_t))) &&
      ok(hipMemcpyAsync(d_w, rarr, nrg * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_w + nrg, mcos_theta, ntg * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_w + nrg + ntg, cos_phi, npg * sizeof(double), hipMemcpyHostToDevice,
                        s)) &&
      ok(hipMemcpyAsync(d_w + nrg + ntg + npg, sin_phi, npg * sizeof(double),
                        hipMemcpyHostToDevice, s)) &&
      ok(hipMemsetAsync(d_cnt, 0, ncell * sizeof(uint32_t), s));
  if (rc == HC_OK) {
    const U128 mult = {0x4385df649fccf645ull, 0x2360ed051fc65da4ull};  // PCG64's multiplier
    hipLaunchKernelGGL(pcg_jump_table, dim3(1), dim3(64), 0, s, mult, U128{pcg[2], pcg[3]},
                       d_jump);
    ok(hipGetLastError());
  }
  if (rc == HC_OK) {
    F.rarr = d_w;
    F.mcos_t = d_w + nrg;
    F.cph = d_w + nrg + ntg;
    F.sph = d_w + nrg + ntg + npg;
    F.jump = d_jump;
    F.counts = d_cnt;
    const uint64_t blocks = (F.npts + 255) / 256;
    hipLaunchKernelGGL(fractal_points, dim3((unsigned)blocks), dim3(256), lds, s, F);
    ok(hipGetLastError()) &&
        ok(hipMemcpyAsync(counts, d_cnt, ncell * sizeof(uint32_t), hipMemcpyDeviceToHost, s)) &&
        ok(hipStreamSynchronize(s));
  }
  (void)hipStreamSynchronize(s);
  if (d_w) (void)hipFree(d_w);
  if (d_jump) (void)hipFree(d_jump);
  if (d_cnt) (void)hipFree(d_cnt);
  (void)hipStreamDestroy(s);
  return rc;
}

extern "C" int hc_build_emissivity(int device, int32_t nl, const double *nud, const double *kd,
                                   const double *ad, int32_t nT, const double *lTint,
                                   int32_t nnu, const double *nuint, int32_t nq,
                                   const double *xq, const double *wq, double *kapBint,
                                   double *kappap) {
  using namespace hc_grid;
  if (nl < 2 || nT < 1 || nnu < 2 || nq < 1 || nq > 64 || !nud || !kd || !ad || !lTint ||
      !nuint || !xq || !wq || !kapBint || !kappap)
    return HC_E_ARG;
  if (hipSetDevice(device) != hipSuccess) return HC_E_HIP;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return HC_E_HIP;
  const size_t nb = (size_t)nT * nnu;
  double *d_tab = nullptr, *d_ax = nullptr, *d_bins = nullptr, *d_kp = nullptr;
  int rc = HC_OK;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == HC_OK) rc = HC_E_HIP;
    return rc == HC_OK;
  };
  // table: nud, kd, ad; axes: lTint, nuint, xq, wq
  const size_t nax = (size_t)nT + nnu + 2 * (size_t)nq;
  ok(hipMalloc((void **)&d_tab, 3 * (size_t)nl * sizeof(double))) &&
      ok(hipMalloc((void **)&d_ax, nax * sizeof(double))) &&
      ok(hipMalloc((void **)&d_bins, nb * sizeof(double))) &&
      ok(hipMalloc((void **)&d_kp, (size_t)nT * sizeof(double))) &&
      ok(hipMemcpyAsync(d_tab, nud, nl * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_tab + nl, kd, nl * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_tab + 2 * nl, ad, nl * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_ax, lTint, nT * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_ax + nT, nuint, nnu * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_ax + nT + nnu, xq, nq * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_ax + nT + nnu + nq, wq, nq * sizeof(double), hipMemcpyHostToDevice, s));
  if (rc == HC_OK) {
    Emis E;
    E.nl = nl;
    E.nT = nT;
    E.nnu = nnu;
    E.nq = nq;
    E.nud = d_tab;
    E.kd = d_tab + nl;
    E.ad = d_tab + 2 * nl;
    E.lTint = d_ax;
    E.nuint = d_ax + nT;
    E.xq = d_ax + nT + nnu;
    E.wq = d_ax + nT + nnu + nq;
    E.bins = d_bins;
    E.kappap = d_kp;
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((nb + 255) / 256, 8192));
    hipLaunchKernelGGL(emis_bins, dim3(blocks), dim3(256), 0, s, E);
    ok(hipGetLastError());
    if (rc == HC_OK) {
      hipLaunchKernelGGL(emis_cumulate, dim3((nT + 255) / 256), dim3(256), 0, s, E);
      ok(hipGetLastError());
    }
    ok(rc == HC_OK ? hipMemcpyAsync(kapBint, d_bins, nb * sizeof(double), hipMemcpyDeviceToHost, s)
                   : hipSuccess) &&
        ok(hipMemcpyAsync(kappap, d_kp, nT * sizeof(double), hipMemcpyDeviceToHost, s)) &&
        ok(hipStreamSynchronize(s));
  }
  (void)hipStreamSynchronize(s);
  for (double *b : {d_tab, d_ax, d_bins, d_kp})
    if (b) (void)hipFree(b);
  (void)hipStreamDestroy(s);
  return rc;
}

namespace hc_grid {

// The kernel parameters of a density build whose arrays are device pointers
// (model walls and multiplier in, densities, volumes out, mass sums scratch).
static Dev density_params(const hc_density_model *m, const double *rarr, const double *thetarr,
                          const double *phiarr, const double *mult, double *dens, double *vol,
                          double *mass_sum) {
  Dev D;
  std::memset(&D, 0, sizeof D);
  D.nrg = m->nrg;
  D.ntg = m->ntg;
  D.npg = m->npg;
  D.ndisk = m->ndisk;
  D.rarr = rarr;
  D.thetarr = thetarr;
  D.phiarr = phiarr;
  D.mult = mult;
  D.rmincgs = m->rmincgs;
  for (int d = 0; d < m->ndisk; ++d) {
    D.disk_id[d] = m->disk_id[d];
    D.disk_z1[d] = m->disk_z1[d];
    D.disk_b[d] = m->disk_b[d];
    D.disk_norm[d] = m->disk_fmass[d] * m->massd * kMsol;
  }
  D.a_exp = m->a_exp;
  D.rmind = m->rmind;
  D.rmaxd = m->rmaxd;
  D.rc = m->rc;
  D.rhoe0 = m->rhoe0;
  D.rmine = m->rmine;
  D.rmax = m->rmax;
  D.rhoamb = m->rhoamb;
  D.env_id = m->env_id;
  D.cav_id = m->cav_id;
  D.c1e = m->c1e;
  D.cavity_exp = m->cavity_exp;
  D.rho_cavity = m->rho_cavity;
  D.r1000 = m->r1000;
  D.dens = dens;
  D.vol = vol;
  D.mass_sum = mass_sum;
  return D;
}

static int check_density_model(const hc_density_model *m) {
  if (!m) return HC_E_ARG;
  if (m->nrg < 2 || m->ntg < 2 || m->npg < 2 || m->ndisk < 0 || m->ndisk > 8 ||
      m->env_id < 0 || m->env_id > 7 || m->cav_id < 0 || m->cav_id > 7 || !m->rarr ||
      !m->thetarr || !m->phiarr)
    return HC_E_ARG;
  for (int d = 0; d < m->ndisk; ++d)
    if (m->disk_id[d] < 0 || m->disk_id[d] > 7) return HC_E_ARG;
  return HC_OK;
}

__global__ void fill_zero(double *p, size_t n) {
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// The disk-mass sums of hc_build_density_noalloc: one array per GPU that the
// library holds until the process ends, so the call allocates nothing.  Host
// code must not use this name as a device address (in host code it is a
// stand-in); it uses the address hipGetSymbolAddress returns for the current
// device instead (density_scratch()).
constexpr int kDensityScratch = HC_DENSITY_SCRATCH;  // doubles: one mass sum per flared disk
static __device__ __align__(256) double scratch[kDensityScratch];

constexpr int kMaxDevices = 64;

// The device address of `scratch` on `device` (which must be current),
// looked up once per device; null if the lookup fails.
static double *density_scratch(int device) {
  static std::mutex mu;
  static double *ptr[kMaxDevices] = {};
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[device]) {
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) return nullptr;
    ptr[device] = static_cast<double *>(p);
  }
  return ptr[device];
}

}  // namespace hc_grid

extern "C" int hc_build_density_noalloc(const hc_density_model *m, int device, void *stream,
                                        double *densarr, double *massarr, double *vcell) {
  using namespace hc_grid;
  if (!densarr || !massarr || !vcell) return HC_E_ARG;
  if (m && (m->ndisk < 0 || m->ndisk > kDensityScratch)) {
    std::fprintf(stderr,
                 "hc_build_density_noalloc: %d disks exceed the %d mass sums of its device "
                 "scratch array\n",
                 m->ndisk, kDensityScratch);
    return HC_E_ARG;
  }
  const int chk = check_density_model(m);
  if (chk != HC_OK) return chk;
  if (hipSetDevice(device) != hipSuccess) return HC_E_HIP;
  double *sum = density_scratch(device);
  if (!sum) return HC_E_HIP;
  hipStream_t s = static_cast<hipStream_t>(stream);
  const size_t wall = (size_t)m->nrg * m->ntg * m->npg;
  const size_t nc = (size_t)(m->nrg - 1) * (m->ntg - 1) * (m->npg - 1);
  const unsigned fill_blocks =
      (unsigned)std::max<size_t>(1, std::min<size_t>((8 * wall + 255) / 256, 8192));
  // outputs and mass sums to zero with kernels (no memset nodes when the
  // caller captures the stream into a graph)
  hipLaunchKernelGGL(fill_zero, dim3(fill_blocks), dim3(256), 0, s, densarr, 8 * wall);
  hipLaunchKernelGGL(fill_zero, dim3(fill_blocks), dim3(256), 0, s, massarr, 8 * wall);
  hipLaunchKernelGGL(fill_zero, dim3(fill_blocks), dim3(256), 0, s, vcell, wall);
  hipLaunchKernelGGL(fill_zero, dim3(1), dim3(64), 0, s, sum, (size_t)kDensityScratch);
  const Dev D = density_params(m, m->rarr, m->thetarr, m->phiarr, m->fractal, densarr, vcell, sum);
  const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((nc + 255) / 256, 8192));
  hipLaunchKernelGGL(densities, dim3(blocks), dim3(256), 0, s, D);
  hipLaunchKernelGGL(normalise, dim3(blocks), dim3(256), 0, s, D, massarr);
  return hipGetLastError() == hipSuccess ? HC_OK : HC_E_HIP;
}

extern "C" int hc_build_density(const hc_density_model *m, int device, double *densarr,
                                double *massarr, double *vcell) {
  using namespace hc_grid;
  if (!m || !densarr || !massarr || !vcell) return HC_E_ARG;
  if (check_density_model(m) != HC_OK) return HC_E_ARG;
  const size_t wall = (size_t)m->nrg * m->ntg * m->npg;
  const size_t nc = (size_t)(m->nrg - 1) * (m->ntg - 1) * (m->npg - 1);
  if (hipSetDevice(device) != hipSuccess) return HC_E_HIP;
  hipStream_t s = nullptr;
  if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) return HC_E_HIP;
  double *d_r = nullptr, *d_t = nullptr, *d_p = nullptr, *d_mult = nullptr, *d_dens = nullptr,
         *d_vol = nullptr, *d_mass = nullptr, *d_sum = nullptr;
  int rc = HC_OK;
  auto ok = [&](hipError_t e) {
    if (e != hipSuccess && rc == HC_OK) rc = HC_E_HIP;
    return rc == HC_OK;
  };
  ok(hipMalloc((void **)&d_r, m->nrg * sizeof(double))) &&
      ok(hipMalloc((void **)&d_t, m->ntg * sizeof(double))) &&
      ok(hipMalloc((void **)&d_p, m->npg * sizeof(double))) &&
      ok(hipMalloc((void **)&d_dens, wall * 8 * sizeof(double))) &&
      ok(hipMalloc((void **)&d_mass, wall * 8 * sizeof(double))) &&
      ok(hipMalloc((void **)&d_vol, wall * sizeof(double))) &&
      ok(hipMalloc((void **)&d_sum, 8 * sizeof(double))) &&
      (!m->fractal || ok(hipMalloc((void **)&d_mult, nc * sizeof(double)))) &&
      ok(hipMemcpyAsync(d_r, m->rarr, m->nrg * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_t, m->thetarr, m->ntg * sizeof(double), hipMemcpyHostToDevice, s)) &&
      ok(hipMemcpyAsync(d_p, m->phiarr, m->npg * sizeof(double), hipMemcpyHostToDevice, s)) &&
      (!m->fractal ||
       ok(hipMemcpyAsync(d_mult, m->fractal, nc * sizeof(double), hipMemcpyHostToDevice, s))) &&
      ok(hipMemsetAsync(d_dens, 0, wall * 8 * sizeof(double), s)) &&
      ok(hipMemsetAsync(d_mass, 0, wall * 8 * sizeof(double), s)) &&
      ok(hipMemsetAsync(d_vol, 0, wall * sizeof(double), s)) &&
      ok(hipMemsetAsync(d_sum, 0, 8 * sizeof(double), s));
  if (rc == HC_OK) {
    const Dev D = density_params(m, d_r, d_t, d_p, d_mult, d_dens, d_vol, d_sum);
    const unsigned blocks = (unsigned)std::max<size_t>(1, std::min<size_t>((nc + 255) / 256, 8192));
    hipLaunchKernelGGL(densities, dim3(blocks), dim3(256), 0, s, D);
    ok(hipGetLastError());
    if (rc == HC_OK) {
      hipLaunchKernelGGL(normalise, dim3(blocks), dim3(256), 0, s, D, d_mass);
      ok(hipGetLastError());
    }
    ok(rc == HC_OK ? hipMemcpyAsync(densarr, d_dens, wall * 8 * sizeof(double),
                                    hipMemcpyDeviceToHost, s)
                   : hipSuccess) &&
        ok(hipMemcpyAsync(massarr, d_mass, wall * 8 * sizeof(double), hipMemcpyDeviceToHost,
                          s)) &&
        ok(hipMemcpyAsync(vcell, d_vol, wall * sizeof(double), hipMemcpyDeviceToHost, s)) &&
        ok(hipStreamSynchronize(s));
  }
  (void)hipStreamSynchronize(s);
  for (double *b : {d_r, d_t, d_p, d_mult, d_dens, d_vol, d_mass, d_sum})
    if (b) (void)hipFree(b);
  (void)hipStreamDestroy(s);
  return rc;
}
