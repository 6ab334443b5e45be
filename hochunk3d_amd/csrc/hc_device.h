// hc_device.h -- device-side building blocks of the photon-packet transport:
// the L'Ecuyer-Cote generator, wall roots, opacity interpolation, CS-HG
// Stokes scattering and peel-off weighting, Lucy re-emission frequency and the
// radiative-equilibrium temperature.  The FP64 routines evaluate the
// reference's Fortran expressions term by term in the same order (the
// translation unit is compiled with -ffp-contract=off), so that the plain FP64
// path rounds like the reference and its CPU restatement.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#define HCM_FN __host__ __device__ static __forceinline__
#include "hc_math.h"

namespace hc {

constexpr double kPi = 3.1415926535897932384626433;  // constants.f90:5
constexpr double kR2P = 2.0 * kPi;
constexpr double kRsol = 6.9598e10;                   // constants.f90:11
constexpr double kSigT = 5.670400e-5;                 // constants.f90:21
constexpr double kGn = 6.67428e-8;                    // constants.f90:19
constexpr double kMsol = 1.989e33;                    // constants.f90:13
constexpr double kHuge = 1.7976931348623157e308;      // huge(1d0)
constexpr uint32_t kMaxTrials = 1u << 20;             // stokes rejection-loop bound
constexpr int kMaxSpecies = 8;                        // grid_mod.f90:13
constexpr int kNo = 8;                                // SED output slots
constexpr double kTauMax = 25.0;                      // peeloff_3d.f90:40

// ---------------------------------------------------------------------------
// RNG: combined MRG of ecuyer_cote_mod.f90:252-293 (Schrage steps), one
// segment per packet.  The segment start is seeded on the host side of the
// launch from a table of a^(2^w 2^k) mod m.
// ---------------------------------------------------------------------------
struct Mrg {
  int32_t s1, s2;
};

__device__ __forceinline__ double ran(Mrg &r) {
  int32_t k = r.s1 / 53668;
  int32_t s1 = 40014 * (r.s1 - k * 53668) - k * 12211;
  if (s1 < 0) s1 += 2147483563;
  k = r.s2 / 52774;
  int32_t s2 = 40692 * (r.s2 - k * 52774) - k * 3791;
  if (s2 < 0) s2 += 2147483399;
  r.s1 = s1;
  r.s2 = s2;
  int32_t z = s1 - s2;
  if (z < 1) z += 2147483563 - 1;
  return (double)z * 4.656613057e-10;  // random_standard_uniform_mod.f90:43
}

__device__ __forceinline__ int32_t mulmod(int32_t a, int32_t b, int32_t m) {
  return (int32_t)(((uint64_t)(uint32_t)a * (uint64_t)(uint32_t)b) % (uint64_t)(uint32_t)m);
}

// pow1[k] = (a1^(2^w))^(2^k) mod m1, pow2 likewise; seg selects the product.
__device__ __forceinline__ Mrg mrg_segment(int32_t g1, int32_t g2, uint64_t seg,
                                           const int32_t *__restrict__ pow1,
                                           const int32_t *__restrict__ pow2) {
  Mrg r{g1, g2};
  int k = 0;
  while (seg) {
    if (seg & 1u) {
      r.s1 = mulmod(r.s1, pow1[k], 2147483563);
      r.s2 = mulmod(r.s2, pow2[k], 2147483399);
    }
    seg >>= 1;
    ++k;
  }
  return r;
}

// ---------------------------------------------------------------------------
// Wall roots: nearest-wall bookkeeping of closest_wall.f90 and the three
// surface families.
// ---------------------------------------------------------------------------
struct Walls {
  double tn1, tp1, tp2;
  int in1, ip1, ip2;
  __device__ __forceinline__ void reset() {
    tn1 = -kHuge;
    tp1 = kHuge;
    tp2 = kHuge;
    in1 = ip1 = ip2 = 0;
  }
  // closest_wall.f90:24-45 as selects (a NaN root changes nothing)
  __device__ __forceinline__ void insert(double t, int i) {
    const bool neg = t < 0.0;
    const bool c0 = neg && t > tn1;
    const bool c1 = !neg && t < tp1;
    const bool c2 = !neg && !c1 && t < tp2;
    tn1 = c0 ? t : tn1;
    in1 = c0 ? i : in1;
    tp2 = c1 ? tp1 : (c2 ? t : tp2);
    ip2 = c1 ? ip1 : (c2 ? i : ip2);
    tp1 = c1 ? t : tp1;
    ip1 = c1 ? i : ip1;
  }
  // find_next_wall: on a wall, skip the root that is that wall.
  __device__ __forceinline__ bool next(bool on_wall, double &t, int &face) const {
    if (on_wall && !(fabs(tn1) < fabs(tp1))) {
      t = tp2;
      face = ip2;
    } else {
      t = tp1;
      face = ip1;
    }
    return t != kHuge;
  }
};

// radface.f90:44-62.  A ray that misses a sphere (descr <= 0) inserts NaN
// roots, which change nothing.
__device__ __forceinline__ void radface(Walls &w, double ux, double uy, double uz, double x,
                                        double y, double z, double r2a, double r2b) {
  const double nan = __builtin_nan("");
  const double bb = 2.0 * (x * ux + y * uy + uz * z);
  double cc = x * x + y * y + z * z - r2a;
  double descr = bb * bb - 4.0 * cc;
  descr = descr > 0.0 ? sqrt(descr) : nan;
  w.insert((-bb + descr) * 0.5, 1);
  w.insert((-bb - descr) * 0.5, 1);
  cc = x * x + y * y + z * z - r2b;
  descr = bb * bb - 4.0 * cc;
  descr = descr > 0.0 ? sqrt(descr) : nan;
  w.insert((-bb + descr) * 0.5, 2);
  w.insert((-bb - descr) * 0.5, 2);
}

// One theta wall: cone x^2+y^2 = t2 z^2 (t2 > 0), the midplane (t2 = -1) or
// a pole (t2 = 0, no wall).  A cell whose two walls carry the same t2 lies
// across the midplane (even wall counts); its two walls are the two nappes of
// one cone and each root goes to the nappe its z says (split = true).
__device__ __forceinline__ void cone_roots(Walls &w, double ux, double uy, double uz, double x,
                                           double y, double z, double t2, int face,
                                           bool split) {
  if (t2 == 0.0) return;
  if (t2 == -1.0) {
    if (uz != 0.0) w.insert(-z / uz, face);
    return;
  }
  const double a = ux * ux + uy * uy - t2 * uz * uz;
  const double b = 2.0 * (x * ux + y * uy - t2 * z * uz);
  const double c = x * x + y * y - t2 * z * z;
  if (a == 0.0) {
    if (b != 0.0) {
      const double t = -c / b;
      w.insert(t, split ? ((z + t * uz) > 0.0 ? 3 : 4) : face);
    }
    return;
  }
  double descr = b * b - 4.0 * a * c;
  if (descr > 0.0) {
    descr = sqrt(descr);
    const double t1 = (-b + descr) / (2.0 * a);
    const double t2r = (-b - descr) / (2.0 * a);
    if (split) {
      w.insert(t1, (z + t1 * uz) > 0.0 ? 3 : 4);
      w.insert(t2r, (z + t2r * uz) > 0.0 ? 3 : 4);
    } else {
      w.insert(t1, face);
      w.insert(t2r, face);
    }
  }
}

__device__ __forceinline__ void thetaface(Walls &w, double ux, double uy, double uz, double x,
                                          double y, double z, double t2a, double t2b) {
  if (t2a == t2b && t2a > 0.0) {
    cone_roots(w, ux, uy, uz, x, y, z, t2a, 3, true);
  } else {
    cone_roots(w, ux, uy, uz, x, y, z, t2a, 3, false);
    cone_roots(w, ux, uy, uz, x, y, z, t2b, 4, false);
  }
}

// phiface.f90:43-57; a ray parallel to a plane inserts a NaN root.
__device__ __forceinline__ void phiface(Walls &w, double a1, double b1, double a2, double b2,
                                        double ux, double uy, double x, double y) {
  const double nan = __builtin_nan("");
  const double d1 = (a1 * ux + b1 * uy);
  const double t1 = -(a1 * x + b1 * y) / d1;
  w.insert(d1 != 0.0 ? t1 : nan, 5);
  const double d2 = (a2 * ux + b2 * uy);
  const double t2 = -(a2 * x + b2 * y) / d2;
  w.insert(d2 != 0.0 ? t2 : nan, 6);
}

// nr.f:131-153 bisection.
__device__ __forceinline__ int locate(const double *__restrict__ xx, int n, double x) {
  int jl = 0, ju = n + 1;
  const bool incr = xx[n - 1] > xx[0];
  while (ju - jl > 1) {
    const int jm = (ju + jl) >> 1;
    if (incr == (x > xx[jm - 1]))
      jl = jm;
    else
      ju = jm;
  }
  if (jl == 0 && x == xx[0]) jl = 1;
  return jl;
}

// ---------------------------------------------------------------------------
// Dust: one species of opacset (dust.f90:13-76).
// ---------------------------------------------------------------------------
struct Opac1 {
  double kappa, albedo, g, pl;
};

__device__ __forceinline__ Opac1 opacset1(const double *__restrict__ nud,
                                          const double *__restrict__ kd,
                                          const double *__restrict__ ad,
                                          const double *__restrict__ gd,
                                          const double *__restrict__ pd, int nl, double nu) {
  Opac1 o;
  if (nu >= nud[0]) {
    o.g = gd[0];
    o.pl = pd[0];
    o.kappa = kd[0];
    o.albedo = ad[0];
  } else if (nu <= nud[nl - 1]) {
    o.g = gd[nl - 1];
    o.pl = pd[nl - 1];
    o.kappa = kd[nl - 1];
    o.albedo = ad[nl - 1];
  } else {
    int i0 = 1, i1 = nl;
    while (i1 - i0 > 1) {
      const int i = (i1 + i0) >> 1;
      if (nu <= nud[i - 1])
        i0 = i;
      else
        i1 = i;
    }
    const double f = (nu - nud[i1 - 1]) / (nud[i0 - 1] - nud[i1 - 1]);
    const double fm1 = 1.0 - f;
    o.g = fm1 * gd[i1 - 1] + f * gd[i0 - 1];
    o.pl = fm1 * pd[i1 - 1] + f * pd[i0 - 1];
    o.kappa = fm1 * kd[i1 - 1] + f * kd[i0 - 1];
    o.albedo = fm1 * ad[i1 - 1] + f * ad[i0 - 1];
  }
  return o;
}

// ---------------------------------------------------------------------------
// Precision of the event arithmetic.  R = double evaluates every expression
// as the reference does with the deterministic elementary functions of
// hc_math.h (bit-identical to the CPU oracle); R = float runs the same
// expressions in FP32 with the hardware's native elementary functions.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double m_log(double x) { return hcm_log(x); }
__device__ __forceinline__ double m_exp(double x) { return hcm_exp(x); }
__device__ __forceinline__ double m_sin(double x) { return hcm_sin(x); }
__device__ __forceinline__ double m_cos(double x) { return hcm_cos(x); }
__device__ __forceinline__ double m_acos(double x) { return hcm_acos(x); }
__device__ __forceinline__ double m_pow(double x, double y) { return hcm_pow(x, y); }
__device__ __forceinline__ double m_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ double m_atan2(double y, double x) { return hcm_atan2(y, x); }
__device__ __forceinline__ float m_log(float x) { return __logf(x); }
__device__ __forceinline__ float m_exp(float x) { return __expf(x); }
__device__ __forceinline__ float m_sin(float x) { return __sinf(x); }
__device__ __forceinline__ float m_cos(float x) { return __cosf(x); }
__device__ __forceinline__ float m_acos(float x) { return acosf(x); }
__device__ __forceinline__ float m_pow(float x, float y) {
  return y == 2.0f ? x * x : __powf(x, y);
}
__device__ __forceinline__ float m_sqrt(float x) { return __fsqrt_rn(x); }
__device__ __forceinline__ float m_atan2(float y, float x) { return atan2f(y, x); }

// Derived HG terms of a species (dust.f90:68-72).
template <class R>
struct HG {
  R g, g2, g2p1, onemg2, twog, pl;
  __device__ __forceinline__ HG(R g_, R pl_) : g(g_), pl(pl_) {
    g2 = g * g;
    g2p1 = g2 + R(1.0);
    onemg2 = R(1.0) - g2;
    twog = R(2.0) * g;
  }
};

// dustmat.f90:39-66 (p4 = 0).
template <class R>
__device__ __forceinline__ void dustmat(const HG<R> &h, R cost, R cost2, R &p1, R &p2, R &p3) {
  const R temp = h.g2p1 - h.twog * cost;
  p1 = R(1.5) * h.onemg2 / (R(2.0) + h.g2) * (R(1.0) + cost2) / (temp * m_sqrt(temp));
  p2 = -h.pl * p1 * (R(1.0) - cost2) / (R(1.0) + cost2);
  p3 = p1 * R(2.0) * cost / (R(1.0) + cost2);
}

template <class R>
__device__ __forceinline__ R cs_peak(const HG<R> &h) {  // stokes.f90:54-62
  const R temp = (h.g > R(0.0)) ? h.g2p1 - h.twog : h.g2p1 + h.twog;
  return R(1.5) * (h.onemg2) / (R(2.0) + h.g2) * R(2.0) / temp / m_sqrt(temp) * R(1.01);
}

// Packet direction and Stokes vector that scattering rotates.
template <class R>
struct Dir {
  R cost, sint, cosp, sinp, phi;
};
template <class R>
struct Stokes4 {
  R I, Q, U, V;
};

// Mueller rows for the scattering frame (stokes.f90:206-240 / 337-369),
// shared by stokes and stokespeel.  upper selects the ri1 > pi branch.
template <class R>
struct Mueller {
  R a11, a12, a13, a21, a22, a23, a31, a32, a33, a44;
};

template <class R>
__device__ __forceinline__ Mueller<R> mueller(bool upper, R p1, R p2, R p3, R sini2, R cosi2,
                                              R s2a, R c2a) {
  // s2a/c2a are sin(2 i1), cos(2 i1) (lower) or sin(2 i3), cos(2 i3) (upper)
  Mueller<R> m;
  const R sin2i2 = R(2.0) * sini2 * cosi2;
  const R cos2i2 = R(2.0) * (cosi2 * cosi2) - R(1.0);
  const R sin2 = sin2i2 * s2a;
  const R cos2 = cos2i2 * c2a;
  const R sin2cos1 = sin2i2 * c2a;
  const R cos2sin1 = cos2i2 * s2a;
  m.a11 = p1;
  m.a12 = p2 * c2a;
  m.a21 = p2 * cos2i2;
  m.a22 = p1 * cos2 - p3 * sin2;
  m.a33 = -p1 * sin2 + p3 * cos2;
  m.a44 = p3;
  if (upper) {
    m.a13 = p2 * s2a;
    m.a23 = p1 * cos2sin1 + p3 * sin2cos1;
    m.a31 = -p2 * sin2i2;
    m.a32 = -p1 * sin2cos1 - p3 * cos2sin1;
  } else {
    m.a13 = -p2 * s2a;
    m.a23 = -p1 * cos2sin1 - p3 * sin2cos1;
    m.a31 = p2 * sin2i2;
    m.a32 = p1 * sin2cos1 + p3 * cos2sin1;
  }
  return m;
}

// stokes.f90:1-392 (idust = 1): rejection sampling of (mu, i1), new direction
// by spherical trigonometry, Stokes rotation normalised by the accepted
// probability.  ntrial/nhit count trials and acceptances (htot, hit).
template <class R>
__device__ __forceinline__ void stokes(Dir<R> &d, Stokes4<R> &s, const HG<R> &h, Mrg &rng,
                                       uint32_t &ntrial, uint32_t &nhit) {
  const R kpi = R(kPi), kr2p = R(kR2P);
  const R costp = d.cost, sintp = d.sint, phip = d.phi;
  R peak = cs_peak(h);
  R bmu, cosb2, p1, p2, p3, sinbt, ri, c2a, s2a, ci, si, rprob;
  bool upper;
  for (uint32_t tries = 1;; ++tries) {
    ntrial += 1;
    bmu = R(1.0) - R(2.0) * R(ran(rng));
    cosb2 = bmu * bmu;
    dustmat(h, bmu, cosb2, p1, p2, p3);
    sinbt = m_sqrt(R(1.0) - cosb2);
    const R ri1 = kr2p * R(ran(rng));
    upper = ri1 > kpi;
    ri = upper ? kr2p - ri1 : ri1;
    ci = m_cos(ri);
    si = m_sin(ri);
    s2a = R(2.0) * si * ci;
    c2a = R(2.0) * (ci * ci) - R(1.0);
    const R a12 = p2 * c2a;
    const R a13 = upper ? p2 * s2a : -p2 * s2a;
    rprob = (p1 * s.I + a12 * s.Q + a13 * s.U) / s.I;
    if (rprob > peak) peak = rprob;
    // rejected; a Stokes vector that no angle accepts (non-finite after FP32
    // rounding) takes the last trial after kMaxTrials
    if (peak * R(ran(rng)) > rprob && tries < kMaxTrials) continue;
    nhit += 1;
    break;
  }
  if (bmu == R(1.0)) return;  // stokes.f90:102-105
  R cost, sint, sini2, cosi2;
  if (bmu == R(-1.0)) {
    si = R(1.0);
    sini2 = R(1.0);
    ci = R(0.0);
    cosi2 = R(0.0);
    cost = -costp;
    sint = sintp;
  } else {
    cost = costp * bmu + sintp * sinbt * ci;
    if (fabs(cost) < R(1.0)) {
      sint = fabs(m_sqrt(R(1.0) - cost * cost));
      sini2 = si * sintp / sint;
      const R bott = sint * sinbt;
      cosi2 = costp / bott - cost * bmu / bott;
      if (cosi2 >= R(1.0)) {
        cosi2 = R(1.0);
        sini2 = R(0.0);
      }
      if (cosi2 <= R(-1.0)) {
        cosi2 = R(-1.0);
        sini2 = R(0.0);
      }
    } else {
      cosi2 = R(0.0);
      if (cost >= R(1.0)) {
        cost = R(1.0);
        cosi2 = R(-1.0);
      }
      if (cost <= R(-1.0)) {
        cost = R(-1.0);
        cosi2 = R(1.0);
      }
      sint = R(0.0);
      sini2 = R(0.0);
    }
  }
  R cosdph = upper ? -cosi2 * ci + sini2 * si * bmu : -ci * cosi2 + si * sini2 * bmu;
  if (fabs(cosdph) > R(1.0)) cosdph = (cosdph > R(1.0)) ? R(1.0) : R(-1.0);
  R phi = upper ? phip + m_acos(cosdph) : phip - m_acos(cosdph);
  if (phi > kr2p) phi = phi - kr2p;
  if (phi < R(0.0)) phi = phi + kr2p;
  const Mueller<R> m = mueller(upper, p1, p2, p3, sini2, cosi2, s2a, c2a);
  const R a = rprob;
  const R sI = (m.a11 * s.I + m.a12 * s.Q + m.a13 * s.U) / a;
  const R sQ = (m.a21 * s.I + m.a22 * s.Q + m.a23 * s.U + R(0.0) * s.V) / a;
  const R sU = (m.a31 * s.I + m.a32 * s.Q + m.a33 * s.U + R(0.0) * s.V) / a;
  const R sV = (R(0.0) * s.Q + R(0.0) * s.U + m.a44 * s.V) / a;
  s.I = sI;
  s.Q = sQ;
  s.U = sU;
  s.V = sV;
  d.cost = cost;
  d.sint = sint;
  d.phi = phi;
  d.cosp = m_cos(phi);
  d.sinp = m_sin(phi);
}

// findangle3.f90 (observer azimuth phi2 = 0 as peeloff_3d.f90:168 passes
// it) followed by stokespeel.f90 for one observer direction (ce, se).
template <class R>
__device__ __forceinline__ void peel_weight(const Dir<R> &d, R ce, R se, R cpe, R spe,
                                            const HG<R> &h, Stokes4<R> &s) {
  const R kpi = R(kPi), kr2p = R(kR2P);
  R bmu = d.cost * ce + (d.sint * se * (d.cosp * cpe + d.sinp * spe));
  if (bmu > R(1.0)) bmu = R(1.0);
  if (bmu < R(-1.0)) bmu = R(-1.0);
  const R sintnew = m_sqrt(R(1.0) - bmu * bmu);
  R ri1 = R(0.0);
  if (d.sint > R(1.e-8) && sintnew > R(1.e-8)) {
    R cosx = (ce - (d.cost * bmu)) / (d.sint * sintnew);
    if (cosx > R(1.0)) cosx = R(1.0);
    if (cosx < R(-1.0)) cosx = R(-1.0);
    const R x = m_acos(cosx);
    const R phi2 = R(0.0);
    if (phi2 > d.phi)
      ri1 = (phi2 - d.phi > kpi) ? x : kr2p - x;
    else
      ri1 = (d.phi - phi2 > kpi) ? kr2p - x : x;
  }
  const R cosb2 = bmu * bmu;
  R p1, p2, p3;
  dustmat(h, bmu, cosb2, p1, p2, p3);
  const R sinbt = m_sqrt(R(1.0) - cosb2);
  const bool upper = ri1 > kpi;
  const R ri = upper ? kr2p - ri1 : ri1;
  R ci = m_cos(ri), si = m_sin(ri);
  const R s2a = R(2.0) * si * ci;
  const R c2a = R(2.0) * (ci * ci) - R(1.0);
  if (bmu == R(1.0)) return;  // stokespeel.f90:116-119
  R sini2, cosi2;
  if (bmu == R(-1.0)) {
    si = R(1.0);
    sini2 = R(1.0);
    ci = R(0.0);
    cosi2 = R(0.0);
  } else if (fabs(ce) < R(1.0)) {
    sini2 = si * d.sint / se;
    const R bott = se * sinbt;
    cosi2 = d.cost / bott - ce * bmu / bott;
    if (cosi2 >= R(1.0)) {
      cosi2 = R(1.0);
      sini2 = R(0.0);
    }
    if (cosi2 <= R(-1.0)) {
      cosi2 = R(-1.0);
      sini2 = R(0.0);
    }
  } else {
    cosi2 = (ce >= R(1.0)) ? R(-1.0) : R(1.0);
    sini2 = R(0.0);
  }
  const Mueller<R> m = mueller(upper, p1, p2, p3, sini2, cosi2, s2a, c2a);
  const R sI = (m.a11 * s.I + m.a12 * s.Q + m.a13 * s.U);
  const R sQ = (m.a21 * s.I + m.a22 * s.Q + m.a23 * s.U + R(0.0) * s.V);
  const R sU = (m.a31 * s.I + m.a32 * s.Q + m.a33 * s.U + R(0.0) * s.V);
  const R sV = (R(0.0) * s.Q + R(0.0) * s.U + m.a44 * s.V);
  s.I = sI;
  s.Q = sQ;
  s.U = sU;
  s.V = sV;
}

}  // namespace hc
