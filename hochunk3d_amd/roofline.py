"""Per-unit algorithmic work of the transport path (frozen constants).

Counted from the FP64 device code (hc_transport.hip, hc_device.h,
hc_math.h) before any GPU timing: one FLOP per IEEE add, subtract, multiply,
divide or square root; comparisons, integer work (the L'Ecuyer-Cote
generator, index arithmetic) and loads are not FLOPs.  The elementary
functions count as the operations of their hc_math.h implementation: log 28,
exp 32, sin or cos 24 (reduction 7 + series 17), acos 55, general pow 65.

Units of work and their FLOPs (NA = 4 active dust species, as in the disk +
envelope + cavity models):

  crossing (transport ray)   radface 34, thetaface 58 (two cones),
                             phiface 14, masked opacity 2 NA, optical depth 3,
                             advance 6, r^2 5, Lucy accumulation 5 NA + 1
                             -> 149
  crossing (peel-off ray)    the same without the Lucy accumulation -> 128
  scattering trial           mu, dustmat, azimuth sin/cos, double angles,
                             acceptance ratio -> 88
  accepted scattering        new direction (acos), Mueller rotation and
                             normalisation, sin/cos of the new azimuth,
                             opacset of the species -> 210
  absorption + re-emission   isotropic direction (sin/cos), lucydustfreq
                             (log, 10-step bisection, log ratio, pow),
                             opacset of NA species (log each) -> 440
  any interaction            select_dust, albedo test, -log(xi), direction
                             -> 60
  packet launch              surface point, limb sampling and frame rotation,
                             blackbody frequency, opacset, cell location -> 450
  peel-off ray               setup, exp(-tau), image projection, phase-matrix
                             weighting for scattered packets -> 250 (mean)

Algorithmic HBM bytes: a transport crossing reads the masked densities of
its cell, NA doubles (32 B at NA = 4, FP64) or floats (16 B, FP32); when the
packet leaves a cell, the cell's Lucy tallies it accumulated (dtauabs of each
species with opacity, jmean, and nabs/dbig after absorptions) are added with
one atomic each, an 8 B read plus an 8 B write of the tally word (B_TALLY
per atomic; the kernel counts them).  The walls and dust tables are small
and cache resident.

MI355X peaks (vector): FP64 78.6 TFLOP/s, FP32 157.3 TFLOP/s; HBM3E 8 TB/s.
"""

F_CROSS = 149
F_PEEL_CROSS = 128
F_TRIAL = 88
F_SCAT = 210
F_ABS = 440
F_EVENT = 60
F_LAUNCH = 450
F_PEEL_RAY = 250
B_CROSS_FP64 = 32
B_CROSS_FP32 = 16
B_TALLY = 16

PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3
PEAK_HBM_GBS = 8000.0


def pass_flops(s: dict, npackets: int) -> float:
    """Algorithmic FLOPs of one pass from its work counters."""
    return (F_CROSS * s["ncross"] + F_TRIAL * s["htot"] + F_SCAT * s["hit"]
            + F_ABS * s["nabs_ev"] + F_EVENT * (s["nscat_ev"] + s["nabs_ev"])
            + F_LAUNCH * npackets + F_PEEL_RAY * s["npeelray"]
            + F_PEEL_CROSS * s["npeelcross"])


def pass_bytes(s: dict, fp32: bool = False) -> float:
    """Algorithmic HBM bytes of one pass from its work counters: the density
    gathers of transport and peel-ray crossings and the tally
    read-modify-writes."""
    b = B_CROSS_FP32 if fp32 else B_CROSS_FP64
    return b * (s["ncross"] + s["npeelcross"]) + B_TALLY * s.get("ntally", 0.0)


# Measured HBM traffic of one plain FP64 Lucy pass of the 200x199x200-cell
# fractal disk + envelope model at 2.388e7 packets (3 x cells), after five
# warm-up iterations: rocprofv3 FETCH_SIZE x 2 plus WRITE_SIZE x 1 (KiB per
# dispatch; the factors are what the two counters read for a 256 MiB device
# copy).  75.1 kB read and 53.3 kB written per packet, 6.5 x the algorithmic
# density bytes: the reads are 64-128 B lines for 32 B of density, the
# writes evictions of the per-cell tally lines the atomics dirty.
CONFIG2_STEP_PACKETS = 23_880_000
CONFIG2_TRAFFIC_PROFILE = "profiles/transport_fp64_config2_pmc.md"
CONFIG2_STEP_HBM_READ = 1.72086e12
CONFIG2_STEP_HBM_WRITE = 1.15943e12
