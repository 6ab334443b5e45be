/*
 * hc.h -- C ABI of the MI355X photon-packet transport library
 * (libhochunk3d_amd.so).  One context per GPU, one host thread per context,
 * synchronous calls.  Every array crosses the boundary in the reference's
 * Fortran column-major layout and dimensions (wall counts nrg/ntg/npg, ndg
 * species), so a Fortran caller passes its module arrays unchanged.
 *
 * The boundary replaces the photon loop and tally reduction of Hochunk3d's
 * disk() (newdisktherm.F90:194-547), called once per Lucy iteration
 * (newtts.F90:312) and once for the final peeled pass (newtts.F90:349), and
 * the temperature update tfinal (newdisktherm.F90:557).
 *
 * Status: every entry point returns 0 on success and a negative HC_E* code on
 * failure; hc_last_error() gives the message.  Conditions on which the
 * reference stops (no wall found, zero-opacity species draw) are counted in
 * hc_tallies instead of aborting.
 */
#ifndef HC_H
#define HC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HC_ABI_VERSION 1

enum {
  HC_OK = 0,
  HC_E_ARG = -1,      /* invalid argument or size */
  HC_E_HIP = -2,      /* HIP runtime failure */
  HC_E_STATE = -3,    /* call order (grid/dust/sources not set) */
  HC_E_UNSUPPORTED = -4
};

enum { HC_FP64 = 0, HC_FP32 = 1 };

typedef struct hc_ctx hc_ctx;

/* Sizes fixed for the life of a context (grid_mod.f90:10-51,
 * dust_mod.f90:7-17, the missing tts_mod's nfreq..nxhst). */
typedef struct {
  int32_t nrg, ntg, npg;   /* wall counts (cells: nrg-1, ntg-1, npg-1) */
  int32_t ndg;             /* species, 8 in the reference */
  int32_t nfreq, nmu, nph, nap, npeel, nxhst;
  int32_t nlammax, nT, nnu;
  int32_t precision;       /* HC_FP64 or HC_FP32 packet state */
  int32_t imcube;          /* the (npeel,nx,nx,nfreq,no) image cube: 0 none, 1 FP64
                            * accumulators, 2 FP32 accumulators returned in the
                            * reference's real*4 arrays (hc_tallies img_* then point
                            * to float) */
} hc_dims;

/* Source and binning scalars: the tts_mod / spot_mod / stokesvar_mod values
 * that the photon loop reads (newdisktherm.F90:60-180, propagate.f90,
 * peeloff_3d.f90).  Pointers are host arrays copied at hc_set_sources. */
typedef struct {
  double rstar;            /* solar radii */
  double tstar, tshock;    /* eV (K / 11605) */
  int32_t limb, ips, ispot, nspot;  /* limb: 0 isotropic, 1 Eddington (hc_set_limb_table) */
  double fspot, csprad, cspot1, sspot1, scspot1, ssspot1, cspot2, sspot2, scspot2, ssspot2;
  double accfrac, accfrac2;
  double numin, lnurat;    /* SED frequency binning (setup.f90:36-40) */
  double rmaxi;            /* image half-size, stellar radii */
  double ltot;             /* erg/s */
  int32_t iintmax;         /* 10000 (propagate.f90:27) */
  int32_t pad;
  const double *u;         /* [nmu] cos(theta) bin centres (initarr.f90:24-28) */
  const double *aperture2; /* [nap] squared aperture radii, stellar radii^2 */
  const double *coste, *sinte, *cospe, *sinpe; /* [npeel] observer directions */
  /* disk-accretion launch (initpacc.f90, emit_gasacc.f90), used when
   * accfrac > 0: emitting radii rddustmin < r < rmaxdmax, the grid's outer
   * radius rmax (stellar radii); per species fmass, z1 (scale height at
   * r = 1) and b (flaring exponent) of the flared disks; the stellar mass
   * massc (Msun) and the disk accretion rate mdot (g/s) of the gas emission */
  double rddustmin, rmaxdmax, rmax;
  double massc, mdot;
  double fmass[8], z1[8], bflare[8];
} hc_sources;

/* Caller-owned host outputs written by hc_run_iteration.  Any pointer may be
 * NULL to skip that tally.  Per-cell arrays are (nrg,ntg,npg[,ndg]);
 * sed is 9 stacked (nfreq,nmu,nph,nap,8) arrays in the order
 * I,Q,U,V,I^2,Q^2,U^2,V^2,N (propagate.f90:276-301); peel is 9 stacked
 * (nfreq,npeel,nap,8) arrays (peeloff_3d.f90:247-262); img_i is
 * (npeel,nx,nx,nfreq,8), img_q/u/v are (npeel,nx,nx,nfreq)
 * (peeloff_3d.f90:270-277).  All accumulators are FP64 (the reference's
 * SED/image arrays are FP32; the caller may narrow them), except the image
 * cube of a context created with imcube = 2, which accumulates and returns
 * FP32 (img_i..img_v are then float arrays). */
typedef struct {
  double *dtauabs, *nabs, *jmean;
  int32_t *killed, *dbig, *dsg;
  double *sed, *peel;
  double *img_i, *img_q, *img_u, *img_v;
  double *si_init;         /* (nfreq,5) */
  double tot, flux, sflux, nscat, scount, abflux, aflux, sumsub, sumall, hit, htot;
  int64_t n_err_nowall, n_err_dust, n_warn;
  double kernel_ms;        /* device time of the transport launch(es) */
  /* work counters of the pass: wall crossings of transport rays, scattering
   * and absorption events, peel-off rays and their wall crossings */
  double ncross, nscat_ev, nabs_ev, npeelray, npeelcross;
  /* the longest packet of the pass: its events + transport-ray crossings
   * (the pass cannot end before that packet's serial walk does) */
  double max_packet_work;
  /* per-cell tally atomics of the pass (dtauabs, jmean, nabs, dbig): each
   * is one 8 B read-modify-write of a tally word */
  double ntally;
} hc_tallies;

/* Per-packet record of the deterministic-trace mode. */
typedef struct {
  double x, y, z, cost, sint, phi, I, Q, U, V, nub, path, tau_peel;
  int64_t ncross, npeelcross, nscat, nabs, nstar, ntrial;
  uint64_t cellhash;
  int32_t status, ir, it, ip, pad;
} hc_trace;

/* Replaces grid_allocate + the module-global setup of the reference
 * (grid_mod.f90:55-92); selects the device. */
int hc_create(const hc_dims *dims, int device, hc_ctx **out);

/* Walls and densities built by gridset (gridset.f90:411-1098):
 * r2arr[nrg], tan2arr/thetarr[ntg], aarr/barr/phiarr[npg],
 * densarr/tdust (nrg,ntg,npg,ndg) with tdust in eV, Tsub in K. */
int hc_set_grid(hc_ctx *ctx, const double *r2arr, const double *tan2arr, const double *thetarr,
                const double *aarr, const double *barr, const double *phiarr,
                const double *densarr, const double *tdust, double tsub);

/* New dust temperatures between Lucy iterations (initarr.f90:78-84). */
int hc_set_temperature(hc_ctx *ctx, const double *tdust);

/* The dust temperatures the device holds (nrg,ntg,npg,ndg) in eV, e.g. after
 * Lucy iterations applied on the device; species without dust read 0.1 K
 * (initarr.f90:53).  Replaces reading tts_mod's tdust after the loop
 * (newtts.F90:312-330). */
int hc_get_temperature(hc_ctx *ctx, double *tdust);

/* Dust tables of dustinit/setemiss (dust.f90:314-474): nlambda[ndg];
 * nudust..pldust (nlammax,ndg); kappav[ndg]; lTint[nT]; nuint[nnu];
 * kapBint (nT,nnu,ndg) normalised cumulative; kappap (nT,ndg). */
int hc_set_dust(hc_ctx *ctx, const int32_t *nlambda, const double *nudust, const double *kappad,
                const double *adust, const double *gdust, const double *pldust,
                const double *kappav, const double *lTint, const double *nuint,
                const double *kapBint, const double *kappap);

int hc_set_sources(hc_ctx *ctx, const hc_sources *src);

/* The stellar spectrum of atmosinit (starbarb.f90:3-118), the atmos_interp
 * module arrays in increasing frequency: atmosnu (eV), the normalised
 * cumulative flux hnuint (0 first, 1 last), lognu = log10(atmosnu), loghnu =
 * log10(hnuint) [nhnu <= 60000].  Stellar (non-spot) photons then take
 * atmosfreq (starbarb.f90:120-156), as the reference's photon loop does for
 * every stellar photon (iplanckst = 0, newdisktherm.F90:283-287); star
 * re-emission too.  nhnu = 0 returns to the Planck function at tstar
 * (bbfreq), the default. */
int hc_set_atmosphere(hc_ctx *ctx, int32_t nhnu, const double *atmosnu, const double *hnuint,
                      const double *lognu, const double *loghnu);

/* The limb-darkening table of `table` (samptabl.f90:1-22): xmu[ntab],
 * prob[ntab] (the cumulative probability (mu^3 + mu^2)/2, increasing to 1;
 * delmu = 1/ntab).  Used by the stellar launch and star re-emission when
 * hc_sources.limb = 1 (initp_common.f90:1-29, darkening samptabl.f90:29-66),
 * with the direct peel weight 12/5 (mu + mu^2) (newdisktherm.F90:349-355).
 * The reference's tabl_mod, which sets ntab, is not in its source tree. */
int hc_set_limb_table(hc_ctx *ctx, int32_t ntab, const double *xmu, const double *prob);

/* Bjorkman-Wood re-emission (CLUCY = NO, ilucy = 0) instead of Lucy's:
 * kapdBint (nT,nnu,ndg) is setemiss's normalised cumulative of
 * kappa_abs dB_nu/dT (dust.f90:426-462, rows ending at 1) and massarr
 * (nrg,ntg,npg,ndg) = (rho V)^(1/4) (gridset).  In the passes that follow,
 * each absorption counts itself in nabs and brings its cell-species
 * temperature up to date with the count so far (dusttemp with
 * nphot = np, emit_common.f90:57-66), in the device temperatures (which
 * hc_get_temperature returns) and in the walk's sublimation mask; the
 * packet re-emits at the mean of the old and new temperatures (3/4 new) on
 * the cell's first absorption, at the new one after, with dustfreq
 * (dust.f90:158-311: the Planck function at or above Tsub).  The reference
 * processes packets one after the other; here packets in flight take their
 * counts in the order of the device's atomics and the stored temperature is
 * the highest computed, so results equal the reference's for the serial
 * order only with hc_set_option(ctx, "serial", 1).  kapdBint = NULL returns
 * to Lucy re-emission. */
int hc_set_bw(hc_ctx *ctx, const double *kapdBint, const double *massarr);

/* Filter curves for the filter-convolved images of peeled passes
 * (IMFILT, peeloff_3d.f90:182-211): filt_mod's arrays after filt()
 * (filt_nu.f90): nwfilt[nfilt] line counts of the filter files (nodes + 1),
 * filtwave (nwfiltmax,nfilt) frequencies in Hz increasing strictly per
 * filter, filtphi (nwfiltmax,nfilt) responses normalised over frequency.
 * nfilt = 0 switches the images off (the default). */
int hc_set_filters(hc_ctx *ctx, int32_t nfilt, int32_t nwfiltmax, const int32_t *nwfilt,
                   const double *filtwave, const double *filtphi);

/* The filter images of the last peeled pass, image_b_{i,q,u,v}
 * (npeel,nxhst,nxhst,nfilt), FP64 (the reference's are real*4): a ray at
 * frequency nu adds its Stokes vector times the response at the last node
 * below nu of every filter whose node range holds nu.  Lucy passes leave
 * them as they are; hc_set_filters discards them (HC_E_STATE until the next
 * peeled pass).  Any pointer may be NULL. */
int hc_get_filter_images(hc_ctx *ctx, double *image_b_i, double *image_b_q, double *image_b_u,
                         double *image_b_v);

/* The diffusion approximation (DIFFUSION=YES; gridset.f90:666-789,
 * emit_common.f90:80-387, propagate.f90:123-142, initpacc.f90:103-112,
 * initarr.f90:78-84), Lucy re-emission only.  diffus/diffdir
 * (nrg,ntg,npg) int8: gridset's layer mask and its direction (-1 radially
 * inwards, -2 up, +2 down); rarr[nrg] the radial walls; kappar (nT,ndg)
 * setemiss's Rosseland-mean table (chiR).  A packet absorbed in a layer
 * cell is not counted in nabs and leaves from the layer's surface at the
 * grey-atmosphere temperature of the surface cell, which every such packet
 * heats (4 pi T*^4 / (nstar dA) in T^4) together with the cells behind it;
 * an accretion packet born in the layer shares its count between the two
 * disk species.  Packets in flight update these temperatures with atomics
 * (the surface by the same step per hit in any order, the cells behind by
 * maximum); hc_set_option(ctx, "serial", 1) runs the reference's order.  The
 * device temperature update (hc_temperature with apply, hc_run_lucy) resets
 * the layer to 0.1 K as initarr does; a caller applying tdust2 itself does
 * the same.  diffus = NULL switches the approximation off.  Needs dust in
 * both disk species (ids 1 and 2). */
int hc_set_diffusion(hc_ctx *ctx, const int8_t *diffus, const int8_t *diffdir, const double *rarr,
                     const double *kappar);

/* One transport pass (the body of disk(), newdisktherm.F90:194-547): np
 * packets, peel-off when ipeel=1.  Packet p (0-based) draws from segment
 * seg0+p of L'Ecuyer-Cote generator gen seeded with (seed1,seed2)
 * (ecuyer_cote_mod.f90).  trace, when non-NULL, receives np records. */
int hc_run_iteration(hc_ctx *ctx, int64_t np, int32_t ipeel, int32_t seed1, int32_t seed2,
                     int32_t gen, uint64_t seg0, hc_tallies *out, hc_trace *trace);

/* Lucy temperature update on the device from the last pass's dtauabs
 * (reconstructed tfinal, newdisktherm.F90:557 with dusttemp.f90:6-58).
 * massarr (nrg,ntg,npg,ndg) = (rho V)^(1/4); NULL reuses the one given to the
 * previous call.  With tdust2 non-NULL the new temperatures are written there
 * (the caller applies them with hc_set_temperature, initarr.f90:78-84); with
 * tdust2 NULL they are applied on the device directly.  nchanged receives
 * the number of dusty (cell, species) whose T moved by more than 1 %. */
int hc_temperature(hc_ctx *ctx, const double *massarr, double ltot, double nphot,
                   double *tdust2, int64_t *nchanged);

/* niter Lucy iterations without peel-off, each one hc_run_iteration pass of
 * np packets followed by hc_temperature applied on the device: the body of
 * the driver's iteration loop (newtts.F90:304-330) without its convergence
 * test, which the caller keeps.  Iteration i (0-based) draws from segments
 * seg0 + i*np + p, so the result is that of the same calls made one by one
 * (up to the order of the floating-point tally sums).  One iteration is
 * captured from the context's stream as a HIP graph and replayed niter
 * times with no host round trip.  out receives the scalar tallies and
 * max_packet_work of the last pass and, in kernel_ms, the device time of
 * all iterations; its array pointers are not used.  nchanged (nullable)
 * receives the last update's changed (cell, species) count.  massarr as in
 * hc_temperature. */
int hc_run_lucy(hc_ctx *ctx, int32_t niter, int64_t np, int32_t seed1, int32_t seed2,
                int32_t gen, uint64_t seg0, const double *massarr, double ltot,
                hc_tallies *out, int64_t *nchanged);

/* niter Lucy iterations of nmodels independent models (a parameter grid)
 * with the passes of all models in one launch per kernel variant (FP64 or
 * FP32, up to 4 or up to 8 active species): hc_run_lucy for many contexts at
 * once.  No counterpart in the reference, which runs one model per process;
 * a grid driver calls it in place of a loop of hc_run_lucy calls.
 *   ctxs[m]        distinct contexts on one device, grid, dust and sources set;
 *   np[m], seg0[m] packets and first draw segment of model m (iteration i
 *                  draws from seg0[m] + i*np[m] + p, as hc_run_lucy);
 *   massarr[m]     as in hc_temperature (massarr itself may be NULL when
 *                  every context already has it);
 *   ltot[m]        the model's luminosity (hc_temperature);
 *   oversubscribe  workgroups launched per workgroup slot of the device,
 *                  0 = 64: each model gets workgroups in proportion to its
 *                  packets; a workgroup runs its model's packets until none
 *                  are left, and later workgroups join the models still
 *                  running, so that the models' longest packets overlap.
 * One iteration of all models (tally zeroing, the passes, every model's
 * temperature update applied on the device) is captured from ctxs[0]'s
 * stream as a HIP graph and replayed niter times.  Each model's results are
 * those of hc_run_lucy on its own context (up to the order of the
 * floating-point tally sums).  out[m] receives model m's last pass like
 * hc_run_iteration (scalars, max_packet_work, and the arrays its non-NULL
 * pointers name), kernel_ms the device time of the whole batch; nchanged[m]
 * (nullable) the last update's changed (cell, species); device_ms
 * (nullable) the device time of all iterations.  Calls that overlap on one
 * device must use disjoint contexts. */
int hc_run_batch(int32_t nmodels, hc_ctx *const *ctxs, int32_t niter, const int64_t *np,
                 int32_t seed1, int32_t seed2, int32_t gen, const uint64_t *seg0,
                 const double *const *massarr, const double *ltot, int32_t oversubscribe,
                 hc_tallies *out, int64_t *nchanged, double *device_ms);

/* Launch-shape knobs of the transport kernel, set by the autotuner
 * (hochunk3d_amd/autotune.py); no counterpart in the reference, whose packet
 * loop is serial.  Names:
 *   "regen_lanes"   1..64: lanes of a wavefront that must wait for one event
 *                   kind before the wavefront stops walking and runs the
 *                   events (default 8 on the FP64 path, 12 on the FP32 path);
 *   "regen_run_lanes" 1..64: once the walk has stopped, every event kind
 *                   with at least this many waiting lanes runs (capped at
 *                   regen_lanes; default 8);
 *   "peel_queue"    0 or 65..1024: peel-off records per wavefront queue in
 *                   LDS (0 = 128, or fewer when the LDS left by the grid's
 *                   walls and dust tables holds fewer);
 *   "blocks_per_cu" 0..64: workgroups per compute unit, 0 = as many as fit
 *                   (default 0);
 *   "serial"        0 or 1: one lane of the device runs the pass's packets
 *                   in their order (the reference's serial loop; for
 *                   Bjorkman-Wood parity, default 0);
 *   "pool"          0 only: the workgroup event pool of the Lucy passes is
 *                   switched off, and 1 is refused (HC_E_UNSUPPORTED);
 *   "max_work"      1..2^31-1: events + cell crossings after which a packet
 *                   is dropped and counted in n_warn (default 2^24, above the
 *                   longest packets of the benchmark models: ~4e5 on the 8M-
 *                   cell fractal model, ~3.4e6 in the optically thick clumps
 *                   of the 31M-cell log-normal model; it bounds the kernel's
 *                   run time against runaway packets).
 * Apart from max_work, which drops the packets that exceed it, results do
 * not depend on them beyond the order of the floating-point tally sums. */
int hc_set_option(hc_ctx *ctx, const char *name, int64_t value);

/* Packets of the last pass dropped at max_work (counted in n_warn): their
 * number in *ndropped and the first min(cap, 16) packet numbers (0-based,
 * draw segment seg0 + pid) in pids, so that a caller can run them again
 * alone, e.g. in the trace mode.  No counterpart in the reference, whose
 * serial loop has no such bound. */
int hc_dropped_packets(hc_ctx *ctx, int64_t *pids, int32_t cap, int64_t *ndropped);

/* Analytic densities of the benchmark models, for grid construction on the
 * GPU: the density part of gridset (gridset.f90:1002-1083) with flared disks
 * (densdiskg.f90), an Ulrich envelope (densenv.f90) and a polynomial bipolar
 * cavity (envset.f90:25-33), evaluated at the cell centres of the walls
 * (gridset.f90:494-499, 609-620).  Lengths in stellar radii. */
typedef struct {
  int32_t nrg, ntg, npg;
  int32_t ndisk;                      /* flared disks, <= 8 */
  const double *rarr, *thetarr, *phiarr; /* walls [nrg], [ntg], [npg] */
  const double *fractal;              /* (nrg-1,ntg-1,npg-1) envelope multiplier or NULL */
  double rmincgs;                     /* stellar radius, cm */
  int32_t disk_id[8];                 /* 0-based species of each disk */
  double disk_fmass[8], disk_z1[8], disk_b[8]; /* mass fraction, h(r=1), flaring */
  double a_exp, rmind, rmaxd, massd;  /* density exponent, radial range, disk mass (Msun) */
  double rc, rhoe0, rmine, rmax, rhoamb; /* centrifugal radius, density scale (g/cm^3) */
  int32_t env_id, cav_id;             /* 0-based species of envelope and cavity */
  double c1e, cavity_exp, rho_cavity, r1000; /* z > c1e rp^ex; rho at r1000 (1000 AU) */
} hc_density_model;

/* Fills densarr and massarr = (rho V)^(1/4), both (nrg,ntg,npg,8), and vcell
 * (nrg,ntg,npg) in cm^3, column-major with the last wall index zero; the
 * disks carry their masses.  No counterpart in the reference's loop: it
 * replaces the host-side evaluation of those formulas in grid setup. */
int hc_build_density(const hc_density_model *model, int device, double *densarr,
                     double *massarr, double *vcell);

/* hc_build_density without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc).  model's rarr,
 * thetarr, phiarr and fractal, and densarr, massarr, vcell (same shapes and
 * values as hc_build_density's) are DEVICE pointers of GPU `device`.  The
 * kernels are queued on `stream` (a hipStream_t; NULL = the null stream) and
 * the call returns without synchronising.  The disk-mass sums live in a
 * device array of HC_DENSITY_SCRATCH doubles that the library holds on each
 * GPU until the process ends (one sum per flared disk: ndisk above it is
 * rejected with HC_E_ARG); every call on one GPU shares that array, so calls
 * that overlap on one GPU race.  Calls no hipMalloc/hipFree. */
#define HC_DENSITY_SCRATCH 8
int hc_build_density_noalloc(const hc_density_model *model, int device, void *stream,
                             double *densarr, double *massarr, double *vcell);

/* setemiss for one species (dust.f90:394-474) on the GPU: kapBint (nT,nnu),
 * the cumulative emissivity kappa_abs B_nu over the frequency bins of nuint
 * (units of kT) normalised to 1 at the last node, and kappap[nT], its total.
 * Each bin is integrated with the caller's rule (nq nodes xq in [0,1], weights
 * wq summing to 1), kappa_abs = (1 - albedo) kappa interpolated linearly in nu
 * between the bin's ends from the dust table nud (decreasing), kd, ad [nl].
 * The reference integrates adaptively to 1e-5 instead. */
int hc_build_emissivity(int device, int32_t nl, const double *nud, const double *kd,
                        const double *ad, int32_t nT, const double *lTint, int32_t nnu,
                        const double *nuint, int32_t nq, const double *xq, const double *wq,
                        double *kapBint, double *kappap);

/* Points per cell of the clumpy-envelope construction (fractal_sph.f90:1-175
 * as restated by hochunk3d_amd/fractal.py), the multiplier's input: n parents
 * uniform in the unit cube, each with n children per level for `levels`
 * levels, offset by (u - 0.5) scale[k-1] at level k; a point's radius is
 * scaled by scale_r and the point counted in its cell.  The deviates are
 * NumPy PCG64 doubles from the stream whose 128-bit state and increment are
 * pcg = {state lo, state hi, inc lo, inc hi} (increment odd), drawn in the
 * host construction's order, so the counts equal the host's.  Cells: radial
 * walls rarr[nrg] below r scale_r, walls with mcos_theta[j] = -cos(theta_j)
 * below -z/r, azimuth walls (uniform, cos_phi/sin_phi[npg]) below the point
 * by half-plane and cross product.  counts: (nrg,ntg,npg) column-major.
 * n^(levels+1) <= 2^31, levels <= 8.  No counterpart in the reference's
 * loop: it replaces the host-side point placement of grid setup. */
int hc_fractal_counts(int device, int32_t nrg, int32_t ntg, int32_t npg, const double *rarr,
                      const double *mcos_theta, const double *cos_phi, const double *sin_phi,
                      int32_t n, int32_t levels, const double *scale, double scale_r,
                      const uint64_t *pcg, uint32_t *counts);

int hc_destroy(hc_ctx *ctx);
const char *hc_last_error(const hc_ctx *ctx);
int hc_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif
