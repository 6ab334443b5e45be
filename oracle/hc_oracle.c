/*
 * hc_oracle.c -- FP64 CPU restatement of the Hochunk3d photon-packet
 * transport path.  TEST INFRASTRUCTURE: only the tests, the smoke check and
 * the benchmark's CPU-baseline leg load this library, and only as the checker.
 *
 * Arithmetic follows the reference's Fortran expressions term by term, in
 * left-to-right order, so that the restatement rounds like the reference
 * (build with -ffp-contract=off).  Each routine cites the file:line it
 * restates.  Three routines whose sources are absent from the reference tree
 * are reconstructed from their call sites (tauint, thetaface, tfinal); their
 * contracts are stated where they are defined.
 */
#include "hc_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

/* Elementary functions: the C library by default; with HCO_DETMATH the
 * deterministic FP64 set the GPU's plain FP64 path uses (hc_math.h), so a
 * packet's arithmetic is bit-identical on both sides. */
#ifdef HCO_DETMATH
#include "../hochunk3d_amd/csrc/hc_math.h"
#define HCO_LOG hcm_log
#define HCO_EXP hcm_exp
#define HCO_SIN hcm_sin
#define HCO_COS hcm_cos
#define HCO_ACOS hcm_acos
#define HCO_POW hcm_pow
#define HCO_ATAN2 hcm_atan2
#define HCO_LOG10 hcm_log10
#define HCO_EXP10 hcm_exp10
#else
#define HCO_LOG log
#define HCO_EXP exp
#define HCO_SIN sin
#define HCO_COS cos
#define HCO_ACOS acos
#define HCO_POW pow
#define HCO_ATAN2 atan2
#define HCO_LOG10 log10
#define HCO_EXP10(y) pow(10.0, (y))
#endif

#define PI 3.1415926535897932384626433  /* constants.f90:5 */
#define R2P (2.0 * PI)
#define RSOL 6.9598e10                   /* constants.f90:11 */
#define SIGT 5.670400e-5                 /* constants.f90:21 */
#define GN 6.67428e-8                    /* constants.f90:19 */
#define MSOL 1.989e33                    /* constants.f90:13 */
#define HUGE_T DBL_MAX
#define HCO_MAX_TRIALS (1u << 20)  /* the device's rejection-loop bound (kMaxTrials) */

/* ------------------------------------------------------------------------ */
/* RNG: ecuyer_cote_mod.f90 (L'Ecuyer & Cote 1991)                           */
/* ------------------------------------------------------------------------ */
enum {
  A1 = 40014, A2 = 40692, M1 = 2147483563, M2 = 2147483399,
  A1VW = 2082007225, A2VW = 784306273, A1W = 1033780774, A2W = 1494757890
};  /* ecuyer_cote_mod.f90:14-17 */

int32_t hco_mulmod(int32_t a, int32_t s, int32_t m) {
  /* a*s mod m for 0<a,s<m: the value multiply_modulo
   * (ecuyer_cote_mod.f90:131-230) computes; here via a 64-bit product. */
  return (int32_t)(((int64_t)a * (int64_t)s) % (int64_t)m);
}

static int32_t powmod(int32_t a, uint64_t e, int32_t m) {
  int64_t r = 1, b = a;
  while (e) {
    if (e & 1u) r = (r * b) % m;
    b = (b * b) % m;
    e >>= 1;
  }
  return (int32_t)r;
}

/* Packet p of a pass draws from segment `segment` of generator `gen`:
 * generator g starts at seed * a^(2^(v+w) (g-1)) (set_all_seeds,
 * ecuyer_cote_mod.f90:345-362) and each segment is 2^w draws further on
 * (reinitialize_current_generator(1), ecuyer_cote_mod.f90:318-320).  The
 * segment start is therefore seed_g * (a^(2^w))^segment mod m. */
void hco_rng_packet(hco_rng *r, int32_t seed1, int32_t seed2, int gen, uint64_t segment) {
  int32_t g1 = seed1, g2 = seed2;
  if (gen > 1) {
    g1 = hco_mulmod(powmod(A1VW, (uint64_t)(gen - 1), M1), g1, M1);
    g2 = hco_mulmod(powmod(A2VW, (uint64_t)(gen - 1), M2), g2, M2);
  }
  r->s1 = hco_mulmod(powmod(A1W, segment, M1), g1, M1);
  r->s2 = hco_mulmod(powmod(A2W, segment, M2), g2, M2);
}

/* random_large_integer (ecuyer_cote_mod.f90:252-293) scaled by
 * random_standard_uniform_mod.f90:43. */
double hco_ran(hco_rng *r) {
  int32_t s1 = r->s1, s2 = r->s2, k, z;
  k = s1 / 53668;
  s1 = A1 * (s1 - k * 53668) - k * 12211;
  if (s1 < 0) s1 = s1 + M1;
  k = s2 / 52774;
  s2 = A2 * (s2 - k * 52774) - k * 3791;
  if (s2 < 0) s2 = s2 + M2;
  r->s1 = s1;
  r->s2 = s2;
  z = s1 - s2;
  if (z < 1) z = z + M1 - 1;
  return (double)z * 4.656613057e-10;
}

/* ------------------------------------------------------------------------ */
/* wall geometry                                                             */
/* ------------------------------------------------------------------------ */
void hco_reset_t(hco_walls *w) {  /* closest_wall.f90:13-22 */
  w->tn1 = -HUGE_T;
  w->tp1 = HUGE_T;
  w->tp2 = HUGE_T;
  w->in1 = w->ip1 = w->ip2 = 0;
}

void hco_insert_t(hco_walls *w, double t, int i) {  /* closest_wall.f90:24-45 */
  if (t < 0.0) {
    if (t > w->tn1) {
      w->tn1 = t;
      w->in1 = i;
    }
  } else {
    if (t < w->tp1) {
      w->tp2 = w->tp1;
      w->ip2 = w->ip1;
      w->tp1 = t;
      w->ip1 = i;
    } else if (t < w->tp2) {
      w->tp2 = t;
      w->ip2 = i;
    }
  }
}

/* closest_wall.f90:47-66; returns 0 when no wall exists (the reference
 * stops there). */
int hco_find_next_wall(const hco_walls *w, int on_wall, double *t, int *face) {
  if (on_wall) {
    if (fabs(w->tn1) < fabs(w->tp1)) {
      *t = w->tp1;
      *face = w->ip1;
    } else {
      *t = w->tp2;
      *face = w->ip2;
    }
  } else {
    *t = w->tp1;
    *face = w->ip1;
  }
  return *t != HUGE_T;
}

void hco_radface(hco_walls *w, double ux, double uy, double uz, double x1, double y1, double z1,
                 double r2_1, double r2_2) {  /* radface.f90:34-61 */
  double bb, cc, descr;
  bb = 2.0 * (x1 * ux + y1 * uy + uz * z1);
  cc = x1 * x1 + y1 * y1 + z1 * z1 - r2_1;
  descr = bb * bb - 4.0 * cc;
  if (descr > 0.0) {
    descr = sqrt(descr);
    hco_insert_t(w, (-bb + descr) * 0.5, 1);
    hco_insert_t(w, (-bb - descr) * 0.5, 1);
  }
  cc = x1 * x1 + y1 * y1 + z1 * z1 - r2_2;
  descr = bb * bb - 4.0 * cc;
  if (descr > 0.0) {
    descr = sqrt(descr);
    hco_insert_t(w, (-bb + descr) * 0.5, 2);
    hco_insert_t(w, (-bb - descr) * 0.5, 2);
  }
}

/* thetaface: source absent from the reference tree (Makefile.in lists
 * thetaface.o).  Reconstructed from its call (peeloff_3d.f90:98, arguments
 * ux,uy,uz,x,y,z,tan2arr(it),tan2arr(it+1)) and the wall conventions of
 * gridset.f90:581-601:
 *   tan2 > 0 : cone x^2+y^2 = tan2*z^2; both real roots of
 *              (ux^2+uy^2-tan2*uz^2) t^2 + 2(x ux+y uy-tan2 z uz) t
 *              + (x^2+y^2-tan2 z^2) = 0 are inserted (a linear root when the
 *              quadratic coefficient is zero);
 *   tan2 = -1: the midplane z = 0, root -z/uz when uz /= 0;
 *   tan2 = 0 : a pole (theta = 0 or pi), no wall.
 * Face ids 3 (wall it) and 4 (wall it+1) follow peeloff_3d.f90:126-129.
 * Roots on the mirror nappe (pi - theta) are inserted unfiltered: from inside
 * a cell the mirror nappe can only be reached after a nearer crossing of the
 * cell's own walls, so it never becomes the selected wall. */
static void thetaface_one(hco_walls *w, double ux, double uy, double uz, double x1, double y1,
                          double z1, double t2, int face) {
  double a, b, c, descr;
  if (t2 == 0.0) return;
  if (t2 == -1.0) {
    if (uz != 0.0) hco_insert_t(w, -z1 / uz, face);
    return;
  }
  a = ux * ux + uy * uy - t2 * uz * uz;
  b = 2.0 * (x1 * ux + y1 * uy - t2 * z1 * uz);
  c = x1 * x1 + y1 * y1 - t2 * z1 * z1;
  if (a == 0.0) {
    if (b != 0.0) hco_insert_t(w, -c / b, face);
    return;
  }
  descr = b * b - 4.0 * a * c;
  if (descr > 0.0) {
    descr = sqrt(descr);
    hco_insert_t(w, (-b + descr) / (2.0 * a), face);
    hco_insert_t(w, (-b - descr) / (2.0 * a), face);
  }
}

/* Grids with an even wall count (which gridset.f90:529-559 cannot build)
 * have a cell across the midplane whose two walls are the two nappes of one
 * cone (equal tan2).  There each root is assigned to the nappe given by the
 * sign of its z: face 3 above the midplane, face 4 below. */
static void straddle_roots(hco_walls *w, double ux, double uy, double uz, double x1, double y1,
                           double z1, double t2) {
  double a, b, c, descr, t;
  a = ux * ux + uy * uy - t2 * uz * uz;
  b = 2.0 * (x1 * ux + y1 * uy - t2 * z1 * uz);
  c = x1 * x1 + y1 * y1 - t2 * z1 * z1;
  if (a == 0.0) {
    if (b != 0.0) {
      t = -c / b;
      hco_insert_t(w, t, (z1 + t * uz) > 0.0 ? 3 : 4);
    }
    return;
  }
  descr = b * b - 4.0 * a * c;
  if (descr > 0.0) {
    descr = sqrt(descr);
    t = (-b + descr) / (2.0 * a);
    hco_insert_t(w, t, (z1 + t * uz) > 0.0 ? 3 : 4);
    t = (-b - descr) / (2.0 * a);
    hco_insert_t(w, t, (z1 + t * uz) > 0.0 ? 3 : 4);
  }
}

void hco_thetaface(hco_walls *w, double ux, double uy, double uz, double x1, double y1, double z1,
                   double t2a, double t2b) {
  if (t2a == t2b && t2a > 0.0) {
    straddle_roots(w, ux, uy, uz, x1, y1, z1, t2a);
    return;
  }
  thetaface_one(w, ux, uy, uz, x1, y1, z1, t2a, 3);
  thetaface_one(w, ux, uy, uz, x1, y1, z1, t2b, 4);
}

/* phiface.f90:43-57 with C=D=0 (the planes contain the z axis). */
void hco_phiface(hco_walls *w, double a1, double b1, double a2, double b2, double ux, double uy,
                 double x1, double y1) {
  double denom1, denom2;
  denom1 = (a1 * ux + b1 * uy);
  if (denom1 != 0.0) hco_insert_t(w, -(a1 * x1 + b1 * y1) / denom1, 5);
  denom2 = (a2 * ux + b2 * uy);
  if (denom2 != 0.0) hco_insert_t(w, -(a2 * x1 + b2 * y1) / denom2, 6);
}

/* nr.f:131-153 locate: 1-based j with xx(j) < x <= xx(j+1) for an
 * increasing table, 0 or n outside; x == xx(1) gives 1. */
int hco_locate(const double *xx, int n, double x) {
  int jl = 0, ju = n + 1, jm;
  int incr = xx[n - 1] > xx[0];
  while (ju - jl > 1) {
    jm = (ju + jl) / 2;
    if (incr == (x > xx[jm - 1]))
      jl = jm;
    else
      ju = jm;
  }
  if (jl == 0 && x == xx[0]) jl = 1;
  return jl;
}

/* ------------------------------------------------------------------------ */
/* dust                                                                      */
/* ------------------------------------------------------------------------ */
#define IDX2(i, j, n1) ((size_t)(i) + (size_t)(n1) * (size_t)(j))

void hco_opacset(const hco_dust *d, double nu, hco_opac *o) {  /* dust.f90:9-80 */
  int id;
  for (id = 0; id < HCO_NDG; id++) {
    const int nl = d->nlambda[id];
    const double *nud = d->nudust + (size_t)d->nlammax * id;
    const double *kd = d->kappad + (size_t)d->nlammax * id;
    const double *ad = d->adust + (size_t)d->nlammax * id;
    const double *gd = d->gdust + (size_t)d->nlammax * id;
    const double *pd = d->pldust + (size_t)d->nlammax * id;
    if (nu >= nud[0]) {
      o->g[id] = gd[0];
      o->pl[id] = pd[0];
      o->kappa[id] = kd[0];
      o->albedo[id] = ad[0];
    } else if (nu <= nud[nl - 1]) {
      o->g[id] = gd[nl - 1];
      o->pl[id] = pd[nl - 1];
      o->kappa[id] = kd[nl - 1];
      o->albedo[id] = ad[nl - 1];
    } else {
      int i0 = 1, i1 = nl, i;
      double f, fm1;
      while (i1 - i0 > 1) {
        i = (i1 + i0) / 2;
        if (nu <= nud[i - 1])
          i0 = i;
        else
          i1 = i;
      }
      f = (nu - nud[i1 - 1]) / (nud[i0 - 1] - nud[i1 - 1]);
      fm1 = 1.0 - f;
      o->g[id] = fm1 * gd[i1 - 1] + f * gd[i0 - 1];
      o->pl[id] = fm1 * pd[i1 - 1] + f * pd[i0 - 1];
      o->kappa[id] = fm1 * kd[i1 - 1] + f * kd[i0 - 1];
      o->albedo[id] = fm1 * ad[i1 - 1] + f * ad[i0 - 1];
    }
    o->g2[id] = o->g[id] * o->g[id];
    o->g2p1[id] = o->g2[id] + 1.0;
    o->onemg2[id] = 1.0 - o->g2[id];
    o->twog[id] = 2.0 * o->g[id];
  }
}

static double inczeta[146];
static int inczeta_ready = 0;

static void zetainit(void) {  /* random.f90:79-97 */
  int i, j;
  double temp;
  double z[146];
  for (i = 1; i <= 145; i++) z[i] = 0.0;
  for (j = 145; j >= 1; j--) {
    temp = 1.0 / (double)(j * j * j * j);
    for (i = 145; i >= j; i--) z[i] = z[i] + temp;
  }
  memcpy(inczeta, z, sizeof z);
  inczeta_ready = 1;
}

double hco_bbfreq(hco_rng *r, double teff) {  /* random.f90:49-77 */
  double z0, z1, z2, z3, z4;
  int i;
  if (!inczeta_ready) zetainit();
  z0 = inczeta[145] * hco_ran(r);
  z1 = hco_ran(r);
  z2 = hco_ran(r);
  z3 = hco_ran(r);
  z4 = hco_ran(r);
  i = 1;
  while (z0 > inczeta[i]) i = i + 1;
  return -HCO_LOG(z1 * z2 * z3 * z4) * teff / (double)i;
}

/* lucydustfreq.f90:19-157, including its guard chain, on the normalised
 * cumulative table tab (nT,nnu,ndg).  The row index j is clamped to
 * [1,nT-1] where the reference would index out of bounds. */
static double cdf_freq(const hco_dust *d, const double *tab, hco_rng *r, double T0, int id) {
  const int nT = d->nT, nnu = d->nnu;
  const double *lTint = d->lTint, *nuint = d->nuint;
  const double *kb = tab + (size_t)nT * nnu * id;
  double lT0, xi, f, P0, P1, P;
  int j, i0, i1, i;
#define KB(jj, ii) kb[IDX2((jj) - 1, (ii) - 1, nT)]
  lT0 = HCO_LOG(T0);
  xi = hco_ran(r);
  j = (int)((double)(nT - 1) * (lT0 - lTint[0]) / (lTint[nT - 1] - lTint[0])) + 1;
  if (j < 1) j = 1;
  if (j > nT - 1) j = nT - 1;
  f = (lT0 - lTint[j - 1]) / (lTint[j] - lTint[j - 1]);
  if (f < 0.0) f = 0.0;
  if (f > 1.0) f = 1.0;
  i0 = 1;
  i1 = nnu;
  P0 = (1.0 - f) * KB(j, i0) + f * KB(j + 1, i0);
  if (xi < P0) {
    if (P0 > 0.0) return T0 * xi * nuint[0] / P0;
    return T0 * nuint[0];
  }
  P1 = (1.0 - f) * KB(j, i1) + f * KB(j + 1, i1);
  if (P1 > 1.0) P1 = 1.0;
  if (xi >= P1) {
    if (P1 < 1.0) {
      if (xi < 1.0) return T0 * nuint[nnu - 1] * (1.0 - P1) / (1.0 - xi);
      return T0 * nuint[nnu - 1];
    }
    return T0 * nuint[nnu - 1];
  }
  while ((i1 - i0) > 1) {
    i = (i0 + i1) / 2;
    P = (1.0 - f) * KB(j, i) + f * KB(j + 1, i);
    if (P > 1.0) P = 1.0;
    if (P < P0) {
      P0 = P;
      i0 = i;
    }
    if (P > P1) {
      P1 = P;
      i1 = 1; /* lucydustfreq.f90:109 as written */
    }
    if (P > xi) {
      i1 = i;
      P1 = P;
    } else {
      i0 = i;
      P0 = P;
    }
  }
  if (xi < P0)
    f = 0.0;
  else if (xi > P1)
    f = 1.0;
  else if (P0 <= 0.0)
    f = 0.5;
  else if (P1 <= 0.0) {
    /* f keeps its interpolation weight (lucydustfreq.f90:138-139) */
  } else if (xi <= 0.0) {
  } else
    f = HCO_LOG(xi / P0) / HCO_LOG(P1 / P0);
#undef KB
  return T0 * nuint[i0 - 1] * HCO_POW(nuint[i1 - 1] / nuint[i0 - 1], f);
}

double hco_lucydustfreq(const hco_dust *d, hco_rng *r, double T0, int id) {
  return cdf_freq(d, d->kapBint, r, T0, id);
}

/* dustfreq (dust.f90:158-311), the Bjorkman-Wood re-emission frequency:
 * at or above the sublimation temperature tsub_k (K) the Planck function
 * (bbfreq), else lucydustfreq's sampling of the kappa_abs dB/dT table
 * kapdBint.  The reference stops where lucydustfreq only warns (P1 <= 0,
 * xi <= 0, P1 < P0, j out of bounds); here both take lucydustfreq's path. */
double hco_dustfreq(const hco_dust *d, hco_rng *r, double T0, int id, double tsub_k) {
  if (T0 >= tsub_k / 11605.0) return hco_bbfreq(r, T0);
  return cdf_freq(d, d->kapdBint, r, T0, id);
}

/* A per-species (nT) table at temperature T: linear in ln T between the
 * nodes of lTint, the end values outside (KapP, dust.f90:114-142, and chiR,
 * dust.f90:83-111, read kappap and kappar this way). */
static double lT_interp(const hco_dust *d, const double *tab, double T, int id) {
  const int nT = d->nT;
  const double *kp = tab + (size_t)nT * id;
  double lT = HCO_LOG(T), f;
  int k = (int)((double)(nT - 1) * (lT - d->lTint[0]) / (d->lTint[nT - 1] - d->lTint[0])) + 1;
  if (k < 1 || k > nT - 1) return k < 1 ? kp[0] : kp[nT - 1];
  f = (lT - d->lTint[k - 1]) / (d->lTint[k] - d->lTint[k - 1]);
  return (1.0 - f) * kp[k - 1] + f * kp[k];
}

double hco_kapp(const hco_dust *d, double T, int id) { return lT_interp(d, d->kappap, T, id); }

double hco_chir(const hco_dust *d, double T, int id) { return lT_interp(d, d->kappar, T, id); }

/* dusttemp.f90:6-58 (mass4 = massarr = (rho V)^(1/4)); the fixed-point
 * loop is capped at 1000 sweeps where the reference would spin. */
double hco_dusttemp(const hco_dust *d, double na, double T0, double ltot, double nphot,
                    double mass4, int id) {
  const double eps = 1.e-5;
  double Td, Told;
  int n = 0;
  Td = (T0 != 0.0) ? T0 : 0.1 / 11605.0;
  Told = 0.5 * Td;
  while (fabs(Td - Told) > eps * Td && n < 1000) {
    Told = Td;
    Td = HCO_POW((na * ltot) / (4.0 * SIGT * nphot * d->kappav[id] * hco_kapp(d, Td, id)), 0.25) /
         11605.0;
    if (mass4 > 0.0)
      Td = Td / mass4;
    else
      Td = 0.1 / 11605.0;
    if (Td > 1600.0 / 11605.0) Td = 1600.0 / 11605.0;
    Td = fmax(Td, 0.1 / 11605.0);
    n++;
  }
  return Td;
}

/* tfinal: source absent (Makefile.in lists tfinal.o; called at
 * newdisktherm.F90:557).  Reconstructed as the Lucy (1999) update the
 * reference's own pieces imply: dtauabs holds the absorbed optical depth
 * sum_k kapd*(1-albedo)*rho*ds per cell and species, so the cell absorbs
 * (L/N)*dtauabs, and dusttemp's balance with na -> dtauabs gives T.  Cells
 * outside the (nrg-1,ntg-1,npg-1) block and species with zero mass keep
 * 0.1 K, the initarr.f90:53 initial value.  Returns the number of dusty
 * (cell, species) pairs whose T moved by more than 1 %. */
int64_t hco_tfinal(const hco_grid *g, const hco_dust *d, const double *dtauabs, double ltot,
                   double nphot, double *tdust2) {
  const size_t ncell = (size_t)g->nrg * g->ntg * g->npg;
  int64_t nchanged = 0;
  int ir, it, ip, id;
  for (id = 0; id < HCO_NDG; id++)
    for (ip = 0; ip < g->npg; ip++)
      for (it = 0; it < g->ntg; it++)
        for (ir = 0; ir < g->nrg; ir++) {
          size_t c = (size_t)ir + (size_t)g->nrg * ((size_t)it + (size_t)g->ntg * ip);
          size_t k = c + ncell * id;
          double tnew = 0.1 / 11605.0;
          if (ir < g->nrg - 1 && it < g->ntg - 1 && ip < g->npg - 1 && g->massarr[k] > 0.0) {
            tnew = hco_dusttemp(d, dtauabs[k], g->tdust[k], ltot, nphot, g->massarr[k], id);
            if (fabs(tnew - g->tdust[k]) > 0.01 * g->tdust[k]) nchanged++;
          }
          tdust2[k] = tnew;
        }
  return nchanged;
}

/* ------------------------------------------------------------------------ */
/* scattering                                                                */
/* ------------------------------------------------------------------------ */
static void dustmat(double *p1, double *p2, double *p3, double *p4, double cost, double cost2,
                    const hco_opac *o, int id) {  /* dustmat.f90:39-66 */
  double temp = o->g2p1[id] - o->twog[id] * cost;
  *p1 = 1.5 * o->onemg2[id] / (2.0 + o->g2[id]) * (1.0 + cost2) / (temp * sqrt(temp));
  *p2 = -o->pl[id] * *p1 * (1.0 - cost2) / (1.0 + cost2);
  *p3 = *p1 * 2.0 * cost / (1.0 + cost2);
  *p4 = 0.0;
}

static double cs_peak(const hco_opac *o, int id) {  /* stokes.f90:54-62 */
  double temp;
  if (o->g[id] > 0.0)
    temp = o->g2p1[id] - o->twog[id];
  else
    temp = o->g2p1[id] + o->twog[id];
  return 1.5 * (o->onemg2[id]) / (2.0 + o->g2[id]) * 2.0 / temp / sqrt(temp) * 1.01;
}

/* stokes.f90:1-392: rejection-sample the scattering angle and azimuth, then
 * rotate the Stokes vector and the direction. */
void hco_stokes(hco_scat *s, const hco_opac *o, int id, hco_rng *r) {
  const double costp = s->cost, sintp = s->sint, phip = s->phi;
  const double sip = s->I, sqp = s->Q, sup = s->U, svp = s->V;
  double peak = cs_peak(o, id);
  double bmu, cosb2, b, p1, p2, p3, p4, sinbt, ri1, ri3, xran, rprob, a;
  double cosi1 = 0, sini1 = 0, cosi2 = 0, sini2 = 0, cosi3 = 0, sini3 = 0;
  double sin2i1, cos2i1, sin2i2, cos2i2, sin2i3, cos2i3, bott, cosdph;
  double sin2, cos2, sin2cos1, cos2sin1;
  double a11, a12, a13, a21, a22, a23, a24, a31, a32, a33, a34, a42, a43, a44;
  double cost = s->cost, sint = s->sint, phi = s->phi;
  for (;;) {
    s->htot = s->htot + 1.0;
    xran = hco_ran(r);
    bmu = 1.0 - 2.0 * xran;
    cosb2 = bmu * bmu;
    b = cosb2 - 1.0;
    (void)b;
    dustmat(&p1, &p2, &p3, &p4, bmu, cosb2, o, id);
    if (fabs(bmu) > 1.0) {
      if (bmu > 1.0) {
        bmu = 1.0;
        cosb2 = 1.0;
      } else {
        bmu = -1.0;
        cosb2 = 1.0;
      }
    }
    sinbt = sqrt(1.0 - cosb2);
    xran = hco_ran(r);
    ri1 = R2P * xran;
    if (ri1 > PI) {
      ri3 = R2P - ri1;
      cosi3 = HCO_COS(ri3);
      sini3 = HCO_SIN(ri3);
      sin2i3 = 2.0 * sini3 * cosi3;
      cos2i3 = 2.0 * (cosi3 * cosi3) - 1.0;
      a11 = p1;
      a12 = p2 * cos2i3;
      a13 = p2 * sin2i3;
      rprob = (a11 * sip + a12 * sqp + a13 * sup) / sip;
      if (rprob > peak) peak = rprob;
      xran = hco_ran(r);
      if (peak * xran > rprob) continue;
      s->hit = s->hit + 1.0;
      a = rprob;
      if (bmu == 1.0) return; /* stokes.f90:102-105: leave the packet as is */
      if (bmu == -1.0) {
        sini3 = 1.0;
        sini2 = 1.0;
        cosi3 = 0.0;
        cosi2 = 0.0;
        cost = -costp;
        sint = sintp;
      } else {
        cost = costp * bmu + sintp * sinbt * cosi3;
        if (fabs(cost) < 1.0) {
          sint = fabs(sqrt(1.0 - cost * cost));
          sini2 = sini3 * sintp / sint;
          bott = sint * sinbt;
          cosi2 = costp / bott - cost * bmu / bott;
          if (cosi2 >= 1.0) {
            cosi2 = 1.0;
            sini2 = 0.0;
          }
          if (cosi2 <= -1.0) {
            cosi2 = -1.0;
            sini2 = 0.0;
          }
        } else {
          if (cost >= 1.0) {
            cost = 1.0;
            cosi2 = -1.0;
          }
          if (cost <= -1.0) {
            cost = -1.0;
            cosi2 = 1.0;
          }
          sint = 0.0;
          sini2 = 0.0;
        }
      }
      cosdph = -cosi2 * cosi3 + sini2 * sini3 * bmu;
      if (fabs(cosdph) > 1.0) cosdph = (cosdph > 1.0) ? 1.0 : -1.0;
      phi = phip + HCO_ACOS(cosdph);
      if (phi > R2P) phi = phi - R2P;
      if (phi < 0.0) phi = phi + R2P;
      sin2i2 = 2.0 * sini2 * cosi2;
      cos2i2 = 2.0 * (cosi2 * cosi2) - 1.0;
      sin2 = sin2i2 * sin2i3;
      cos2 = cos2i2 * cos2i3;
      sin2cos1 = sin2i2 * cos2i3;
      cos2sin1 = cos2i2 * sin2i3;
      a21 = p2 * cos2i2;
      a22 = p1 * cos2 - p3 * sin2;
      a23 = p1 * cos2sin1 + p3 * sin2cos1;
      a24 = -p4 * sin2i2;
      a31 = -p2 * sin2i2;
      a32 = -p1 * sin2cos1 - p3 * cos2sin1;
      a33 = -p1 * sin2 + p3 * cos2;
      a34 = -p4 * cos2i2;
      a42 = -p4 * sin2i3;
      a43 = p4 * cos2i3;
      a44 = p3;
    } else {
      cosi1 = HCO_COS(ri1);
      sini1 = HCO_SIN(ri1);
      sin2i1 = 2.0 * sini1 * cosi1;
      cos2i1 = 2.0 * (cosi1 * cosi1) - 1.0;
      a11 = p1;
      a12 = p2 * cos2i1;
      a13 = -p2 * sin2i1;
      rprob = (a11 * sip + a12 * sqp + a13 * sup) / sip;
      if (rprob > peak) peak = rprob;
      xran = hco_ran(r);
      if (peak * xran > rprob) continue;
      s->hit = s->hit + 1.0;
      a = rprob;
      if (bmu == 1.0) return;
      if (bmu == -1.0) {
        sini1 = 1.0;
        sini2 = 1.0;
        cosi1 = 0.0;
        cosi2 = 0.0;
        cost = -costp;
        sint = sintp;
      } else {
        cost = costp * bmu + sintp * sinbt * cosi1;
        if (fabs(cost) < 1.0) {
          sint = fabs(sqrt(1.0 - cost * cost));
          sini2 = sini1 * sintp / sint;
          bott = sint * sinbt;
          cosi2 = costp / bott - cost * bmu / bott;
          if (cosi2 >= 1.0) {
            cosi2 = 1.0;
            sini2 = 0.0;
          }
          if (cosi2 <= -1.0) {
            cosi2 = -1.0;
            sini2 = 0.0;
          }
        } else {
          if (cost >= 1.0) {
            cost = 1.0;
            cosi2 = -1.0;
          }
          if (cost <= -1.0) {
            cost = -1.0;
            cosi2 = 1.0;
          }
          sint = 0.0;
          sini2 = 0.0;
        }
      }
      cosdph = -cosi1 * cosi2 + sini1 * sini2 * bmu;
      if (fabs(cosdph) > 1.0) cosdph = (cosdph > 1.0) ? 1.0 : -1.0;
      phi = phip - HCO_ACOS(cosdph);
      if (phi > R2P) phi = phi - R2P;
      if (phi < 0.0) phi = phi + R2P;
      sin2i2 = 2.0 * sini2 * cosi2;
      cos2i2 = 2.0 * (cosi2 * cosi2) - 1.0;
      sin2 = sin2i2 * sin2i1;
      cos2 = cos2i2 * cos2i1;
      sin2cos1 = sin2i2 * cos2i1;
      cos2sin1 = cos2i2 * sin2i1;
      a21 = p2 * cos2i2;
      a22 = p1 * cos2 - p3 * sin2;
      a23 = -p1 * cos2sin1 - p3 * sin2cos1;
      a24 = p4 * sin2i2;
      a31 = p2 * sin2i2;
      a32 = p1 * sin2cos1 + p3 * cos2sin1;
      a33 = -p1 * sin2 + p3 * cos2;
      a34 = -p4 * cos2i2;
      a42 = p4 * sin2i1;
      a43 = p4 * cos2i1;
      a44 = p3;
    }
    break;
  }
  s->I = (a11 * sip + a12 * sqp + a13 * sup) / a;
  s->Q = (a21 * sip + a22 * sqp + a23 * sup + a24 * svp) / a;
  s->U = (a31 * sip + a32 * sqp + a33 * sup + a34 * svp) / a;
  s->V = (a42 * sqp + a43 * sup + a44 * svp) / a;
  s->cost = cost;
  s->sint = sint;
  s->phi = phi;
  s->cosp = HCO_COS(phi);
  s->sinp = HCO_SIN(phi);
}

/* stokespeel.f90:1-382: the un-normalised phase-matrix weighting of a peeled
 * packet for the given scattering cosine bmu and rotation angle ri1.  cost
 * and sint are the peel direction (clamped in place like the reference). */
void hco_stokespeel(double *I, double *Q, double *U, double *V, double *cost, double *sint,
                    double ri1, double bmu, double costp, double sintp, const hco_opac *o, int id) {
  const double sip = *I, sqp = *Q, sup = *U, svp = *V;
  double cosb2, p1, p2, p3, p4, sinbt;
  double cosi1 = 0, sini1 = 0, cosi2 = 0, sini2 = 0, cosi3 = 0, sini3 = 0, ri3;
  double sin2i1, cos2i1, sin2i2, cos2i2, sin2i3, cos2i3, bott;
  double sin2, cos2, sin2cos1, cos2sin1;
  double a11, a12, a13, a21, a22, a23, a24, a31, a32, a33, a34, a42, a43, a44;
  cosb2 = bmu * bmu;
  dustmat(&p1, &p2, &p3, &p4, bmu, cosb2, o, id);
  if (fabs(bmu) > 1.0) {
    bmu = (bmu > 1.0) ? 1.0 : -1.0;
    cosb2 = 1.0;
  }
  sinbt = sqrt(1.0 - cosb2);
  if (ri1 > PI) {
    ri3 = R2P - ri1;
    cosi3 = HCO_COS(ri3);
    sini3 = HCO_SIN(ri3);
    sin2i3 = 2.0 * sini3 * cosi3;
    cos2i3 = 2.0 * (cosi3 * cosi3) - 1.0;
    a11 = p1;
    a12 = p2 * cos2i3;
    a13 = p2 * sin2i3;
    if (bmu == 1.0) return; /* stokespeel.f90:116-119 */
    if (bmu == -1.0) {
      sini3 = 1.0;
      sini2 = 1.0;
      cosi3 = 0.0;
      cosi2 = 0.0;
    } else if (fabs(*cost) < 1.0) {
      sini2 = sini3 * sintp / *sint;
      bott = *sint * sinbt;
      cosi2 = costp / bott - *cost * bmu / bott;
      if (cosi2 >= 1.0) {
        cosi2 = 1.0;
        sini2 = 0.0;
      }
      if (cosi2 <= -1.0) {
        cosi2 = -1.0;
        sini2 = 0.0;
      }
    } else {
      if (*cost >= 1.0) {
        *cost = 1.0;
        cosi2 = -1.0;
      }
      if (*cost <= -1.0) {
        *cost = -1.0;
        cosi2 = 1.0;
      }
      *sint = 0.0;
      sini2 = 0.0;
    }
    sin2i2 = 2.0 * sini2 * cosi2;
    cos2i2 = 2.0 * (cosi2 * cosi2) - 1.0;
    sin2 = sin2i2 * sin2i3;
    cos2 = cos2i2 * cos2i3;
    sin2cos1 = sin2i2 * cos2i3;
    cos2sin1 = cos2i2 * sin2i3;
    a21 = p2 * cos2i2;
    a22 = p1 * cos2 - p3 * sin2;
    a23 = p1 * cos2sin1 + p3 * sin2cos1;
    a24 = -p4 * sin2i2;
    a31 = -p2 * sin2i2;
    a32 = -p1 * sin2cos1 - p3 * cos2sin1;
    a33 = -p1 * sin2 + p3 * cos2;
    a34 = -p4 * cos2i2;
    a42 = -p4 * sin2i3;
    a43 = p4 * cos2i3;
    a44 = p3;
  } else {
    cosi1 = HCO_COS(ri1);
    sini1 = HCO_SIN(ri1);
    sin2i1 = 2.0 * sini1 * cosi1;
    cos2i1 = 2.0 * (cosi1 * cosi1) - 1.0;
    a11 = p1;
    a12 = p2 * cos2i1;
    a13 = -p2 * sin2i1;
    if (bmu == 1.0) return;
    if (bmu == -1.0) {
      sini1 = 1.0;
      sini2 = 1.0;
      cosi1 = 0.0;
      cosi2 = 0.0;
    } else if (fabs(*cost) < 1.0) {
      sini2 = sini1 * sintp / *sint;
      bott = *sint * sinbt;
      cosi2 = costp / bott - *cost * bmu / bott;
      if (cosi2 >= 1.0) {
        cosi2 = 1.0;
        sini2 = 0.0;
      }
      if (cosi2 <= -1.0) {
        cosi2 = -1.0;
        sini2 = 0.0;
      }
    } else {
      if (*cost >= 1.0) {
        *cost = 1.0;
        cosi2 = -1.0;
      }
      if (*cost <= -1.0) {
        *cost = -1.0;
        cosi2 = 1.0;
      }
      *sint = 0.0;
      sini2 = 0.0;
    }
    sin2i2 = 2.0 * sini2 * cosi2;
    cos2i2 = 2.0 * (cosi2 * cosi2) - 1.0;
    sin2 = sin2i2 * sin2i1;
    cos2 = cos2i2 * cos2i1;
    sin2cos1 = sin2i2 * cos2i1;
    cos2sin1 = cos2i2 * sin2i1;
    a21 = p2 * cos2i2;
    a22 = p1 * cos2 - p3 * sin2;
    a23 = -p1 * cos2sin1 - p3 * sin2cos1;
    a24 = p4 * sin2i2;
    a31 = p2 * sin2i2;
    a32 = p1 * sin2cos1 + p3 * cos2sin1;
    a33 = -p1 * sin2 + p3 * cos2;
    a34 = -p4 * cos2i2;
    a42 = p4 * sin2i1;
    a43 = p4 * cos2i1;
    a44 = p3;
  }
  *I = (a11 * sip + a12 * sqp + a13 * sup);
  *Q = (a21 * sip + a22 * sqp + a23 * sup + a24 * svp);
  *U = (a31 * sip + a32 * sqp + a33 * sup + a34 * svp);
  *V = (a42 * sqp + a43 * sup + a44 * svp);
}

static double clamp1(double x) {  /* check.f90:1-30 */
  if (x > 1.0) return 1.0;
  if (x < -1.0) return -1.0;
  return x;
}

/* findangle3.f90:1-70.  ri1 is left at 0 in the near-pole branches where the
 * reference leaves it unset. */
void hco_findangle3(double cost1, double sint1, double phi1, double cosp1, double sinp1,
                    double cost2, double sint2, double phi2, double cosp2, double sinp2,
                    double *costnew, double *sintnew, double *phinew, double *ri1) {
  double cosx, x;
  *ri1 = 0.0;
  *costnew = cost1 * cost2 + (sint1 * sint2 * (cosp1 * cosp2 + sinp1 * sinp2));
  *costnew = clamp1(*costnew);
  *sintnew = sqrt(1.0 - *costnew * *costnew);
  if (sint1 > 1.e-8) {
    if (*sintnew > 1.e-8) {
      cosx = (cost2 - (cost1 * *costnew)) / (sint1 * *sintnew);
      cosx = clamp1(cosx);
      x = HCO_ACOS(cosx);
      if (phi2 > phi1) {
        if (phi2 - phi1 > PI) {
          *ri1 = x;
          *phinew = PI + x;
        } else {
          *phinew = PI - x;
          *ri1 = R2P - x;
        }
      } else {
        if (phi1 - phi2 > PI) {
          *phinew = PI - x;
          *ri1 = R2P - x;
        } else {
          *phinew = PI + x;
          *ri1 = x;
        }
      }
    } else {
      *phinew = 0.0;
    }
  } else {
    *phinew = phi2;
  }
}

/* ------------------------------------------------------------------------ */
/* packet transport                                                          */
/* ------------------------------------------------------------------------ */
typedef struct {
  double xp, yp, zp, ux, uy, uz;
  double cost, sint, cosp, sinp, phi;
  double sip, sqp, sup, svp, nub;
  double rsq, rtot;
  int ir, it, ip, on_wall;
  int iflag, i_orig;
  double kapd[HCO_NDG];
  hco_opac op;
} pkt_t;

typedef struct {
  const hco_grid *g;
  const hco_dust *d;
  const hco_setup *s;
  hco_tallies *t;
  hco_rng rng;
  hco_trace *tr;
  int ipeel;
  double rmincgs;
  double tsub_ev;
  size_t ncell;
  int nact;               /* species with any density */
  int act[HCO_NDG];
  double nphot;           /* packets of the pass (dusttemp's nphot) */
  double nstar;           /* stellar packets of the pass (emit_common's nstar) */
} ctx_t;

static inline size_t cell_index(const hco_grid *g, int ir, int it, int ip) {
  return (size_t)(ir - 1) + (size_t)g->nrg * ((size_t)(it - 1) + (size_t)g->ntg * (size_t)(ip - 1));
}

static inline void hash_cell(hco_trace *tr, int ir, int it, int ip) {
  if (!tr) return;
  tr->cellhash ^= (uint64_t)(uint32_t)ir | ((uint64_t)(uint32_t)it << 20) |
                  ((uint64_t)(uint32_t)ip << 40);
  tr->cellhash *= 1099511628211ull;
}

/* Masked opacity of a cell (peeloff_3d.f90:103-110). */
static inline double cell_opac(const ctx_t *c, const pkt_t *P, size_t cell) {
  double opac = 0.0;
  int id;
  for (id = 0; id < HCO_NDG; id++) {
    size_t k = cell + c->ncell * id;
    if (c->g->tdust[k] < c->tsub_ev) opac = opac + P->kapd[id] * c->g->densarr[k];
  }
  return opac;
}

/* Walls of the current cell and the next one along (ux,uy,uz)
 * (peeloff_3d.f90:92-103).  On the polar axis (x = y = 0) every phi wall
 * passes through the point, so phiface's roots are all 0 and the walk would
 * step from phi cell to phi cell without moving; a straight ray leaving the
 * axis keeps the azimuth of its direction, so there the phi walls are
 * skipped and the caller has put the packet in the phi cell of its
 * direction (axis_phi_cell). */
static inline int next_wall(const ctx_t *c, double ux, double uy, double uz, double x, double y,
                            double z, int ir, int it, int ip, int on_wall, double *dt, int *face) {
  const hco_grid *g = c->g;
  hco_walls w;
  hco_reset_t(&w);
  hco_radface(&w, ux, uy, uz, x, y, z, g->r2arr[ir - 1], g->r2arr[ir]);
  if (g->ntg > 2) hco_thetaface(&w, ux, uy, uz, x, y, z, g->tan2arr[it - 1], g->tan2arr[it]);
  if (g->npg > 2 && g->ntg > 2 && !(x == 0.0 && y == 0.0))
    hco_phiface(&w, g->aarr[ip - 1], g->barr[ip - 1], g->aarr[ip], g->barr[ip], ux, uy, x, y);
  return hco_find_next_wall(&w, on_wall, dt, face);
}

/* The phi cell a ray leaving the polar axis enters: the one holding the
 * azimuth of its direction (unchanged when the ray runs along the axis). */
static inline int axis_phi_cell(const hco_grid *g, double ux, double uy, int ip) {
  double p;
  if (g->npg <= 2 || (ux == 0.0 && uy == 0.0)) return ip;
  p = HCO_ATAN2(uy, ux);
  if (p < 0.0) p = p + R2P;
  return hco_locate(g->phiarr, g->npg, p);
}

static inline void face_step(const hco_grid *g, int face, int *ir, int *it, int *ip) {
  switch (face) {  /* peeloff_3d.f90:120-139 */
    case 1: *ir -= 1; break;
    case 2: *ir += 1; break;
    case 3: *it -= 1; break;
    case 4: *it += 1; break;
    case 5: *ip -= 1; break;
    default: *ip += 1; break;
  }
  if (*ip == 0) *ip = g->npg - 1;
  if (*ip == g->npg) *ip = 1;
}

/* A straight ray through the empty grid from (x,y,z), starting in cell
 * (ir,it,ip) (on a radial wall when on_wall), to the outer sphere or the
 * star: the wall-crossing geometry of tauint/peeloff without opacity, for
 * tests.  Returns the number of crossings (-1 when no wall is found or the
 * ray makes more crossings than a straight line can), and the last cell. */
int hco_walk_ray(const hco_grid *g, double x, double y, double z, double ux, double uy, double uz,
                 int *ir, int *it, int *ip, int on_wall) {
  ctx_t c;
  int n = 0, face;
  double dt;
  memset(&c, 0, sizeof c);
  c.g = g;
  while (*ir >= 1 && *ir < g->nrg) {
    if (n > 2 * (g->nrg + g->ntg + g->npg) + 16) return -1;
    if (x == 0.0 && y == 0.0) *ip = axis_phi_cell(g, ux, uy, *ip);
    if (!next_wall(&c, ux, uy, uz, x, y, z, *ir, *it, *ip, on_wall, &dt, &face)) return -1;
    face_step(g, face, ir, it, ip);
    x = x + dt * ux;
    y = y + dt * uy;
    z = z + dt * uz;
    on_wall = 1;
    n++;
  }
  return n;
}

enum { ST_INTERACT = 0, ST_EXIT = 1, ST_STAR = 2, ST_ERROR = 3 };

/* tauint: source absent (Makefile.in lists tauint_3d.o; called at
 * propagate.f90:50,185).  Reconstructed on the peel-off walk of
 * peeloff_3d.f90:74-154: walk cells from (xp,yp,zp) along (ux,uy,uz) summing
 * the masked opacity times the chord until the sum reaches tau.
 *  - Interaction inside a cell: advance by (tau-tsum)/opac, on_wall=false.
 *  - ir reaches nrg: escape (exitflag).  ir reaches 0: the packet hit the
 *    star (the caller re-emits it from the stellar surface).
 *  - Lucy tallies per traversed chord ds (stellar radii):
 *    dtauabs(cell,id) += kapd(id)*(1-albedo(id))*rho(cell,id)*ds for the
 *    unmasked species, jmean(cell) += ds*R*(cm).
 *  - The radial on_wall re-orientation at entry is the one of
 *    peeloff_3d.f90:58-72. */
static int tauint(ctx_t *c, pkt_t *P, double tau) {
  const hco_grid *g = c->g;
  hco_tallies *t = c->t;
  double tsum = 0.0, dt, opac, dtau;
  int face;
  if (P->on_wall) {
    int inner = fabs(P->rsq - g->r2arr[P->ir - 1]) < fabs(P->rsq - g->r2arr[P->ir]);
    int outwards = P->ux * P->xp + P->uy * P->yp + P->uz * P->zp > 0.0;
    if (inner && !outwards) {
      P->ir -= 1;
      if (P->ir == 0) return ST_STAR;
    }
    if (!inner && outwards) P->ir += 1;
    if (P->ir == g->nrg) return ST_EXIT;
  }
  for (;;) {
    size_t cell;
    int id;
    if (P->xp == 0.0 && P->yp == 0.0) P->ip = axis_phi_cell(g, P->ux, P->uy, P->ip);
    cell = cell_index(g, P->ir, P->it, P->ip);
    if (!next_wall(c, P->ux, P->uy, P->uz, P->xp, P->yp, P->zp, P->ir, P->it, P->ip, P->on_wall,
                   &dt, &face)) {
      t->n_err_nowall++;
      return ST_ERROR;
    }
    opac = cell_opac(c, P, cell);
    dtau = opac * dt;
    if (tsum + dtau >= tau) {
      double ds = (tau - tsum) / opac;
      if (ds > dt) ds = dt; /* rounding must not carry the packet past the wall */
      if (t->dtauabs)
        for (id = 0; id < HCO_NDG; id++) {
          size_t k = cell + c->ncell * id;
          if (g->tdust[k] < c->tsub_ev)
            t->dtauabs[k] += P->kapd[id] * (1.0 - P->op.albedo[id]) * g->densarr[k] * ds;
        }
      if (t->jmean) t->jmean[cell] += ds * c->rmincgs;
      P->xp = P->xp + ds * P->ux;
      P->yp = P->yp + ds * P->uy;
      P->zp = P->zp + ds * P->uz;
      P->on_wall = 0;
      P->rsq = P->xp * P->xp + P->yp * P->yp + P->zp * P->zp;
      P->rtot = sqrt(P->rsq);
      if (c->tr) c->tr->path += ds;
      return ST_INTERACT;
    }
    if (t->dtauabs)
      for (id = 0; id < HCO_NDG; id++) {
        size_t k = cell + c->ncell * id;
        if (g->tdust[k] < c->tsub_ev)
          t->dtauabs[k] += P->kapd[id] * (1.0 - P->op.albedo[id]) * g->densarr[k] * dt;
      }
    if (t->jmean) t->jmean[cell] += dt * c->rmincgs;
    tsum = tsum + dtau;
    face_step(g, face, &P->ir, &P->it, &P->ip);
    P->xp = P->xp + dt * P->ux;
    P->yp = P->yp + dt * P->uy;
    P->zp = P->zp + dt * P->uz;
    P->on_wall = 1;
    P->rsq = P->xp * P->xp + P->yp * P->yp + P->zp * P->zp;
    P->rtot = sqrt(P->rsq);
    if (c->tr) {
      c->tr->path += dt;
      c->tr->ncross++;
      hash_cell(c->tr, P->ir, P->it, P->ip);
    }
    if (P->ir == g->nrg) return ST_EXIT;
    if (P->ir == 0) return ST_STAR;
  }
}

static inline int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

/* peeloff_3d.f90:5-288 for one observer direction. */
static void peeloff(ctx_t *c, const pkt_t *P0, double sip, int idust2, int peelid) {
  const hco_grid *g = c->g;
  const hco_setup *s = c->s;
  hco_tallies *t = c->t;
  double sqp = P0->sqp, sup = P0->sup, svp = P0->svp;
  double xp1 = P0->xp, yp1 = P0->yp, zp1 = P0->zp, rsq1 = P0->rsq;
  int ir = P0->ir, it = P0->it, ip = P0->ip, on_wall = P0->on_wall;
  double cospe = s->cospe[peelid], sinpe = s->sinpe[peelid];
  double coste = s->coste[peelid], sinte = s->sinte[peelid];
  double ux1 = sinte * cospe, uy1 = sinte * sinpe, uz1 = coste;
  double tsum = 0.0, dt, opac, factor, ximage, yimage, rho2;
  int face, ix, iy, inub, ia, i_inter, k, indx[3], nstep = 0;
  if (on_wall) {
    int inner = fabs(rsq1 - g->r2arr[ir - 1]) < fabs(rsq1 - g->r2arr[ir]);
    int outwards = ux1 * xp1 + uy1 * yp1 + uz1 * zp1 > 0.0;
    if (inner && !outwards) {
      if (ir == 1) return;
      ir = ir - 1;
    }
    if (!inner && outwards) ir = ir + 1;
  }
  while (ir < g->nrg) {
    /* a straight ray meets each sphere and each cone at most twice and each
     * phi half-plane at most once: more crossings than that is a stuck ray */
    if (++nstep > 2 * (g->nrg + g->ntg + g->npg) + 17) {
      t->n_err_nowall++;
      return;
    }
    if (xp1 == 0.0 && yp1 == 0.0) ip = axis_phi_cell(g, ux1, uy1, ip);
    if (!next_wall(c, ux1, uy1, uz1, xp1, yp1, zp1, ir, it, ip, on_wall, &dt, &face)) {
      t->n_err_nowall++;
      return;
    }
    opac = cell_opac(c, P0, cell_index(g, ir, it, ip));
    tsum = tsum + opac * dt;
    if (tsum > HCO_TAUMAX) {
      if (c->tr) c->tr->tau_peel += tsum;
      return;
    }
    face_step(g, face, &ir, &it, &ip);
    xp1 = xp1 + dt * ux1;
    yp1 = yp1 + dt * uy1;
    zp1 = zp1 + dt * uz1;
    on_wall = 1;
    if (c->tr) c->tr->npeelcross++;
    if (ir == g->nrg) break;
    if (ir == 0) {
      if (c->tr) c->tr->tau_peel += tsum;
      return;
    }
  }
  if (c->tr) c->tr->tau_peel += tsum;
  factor = HCO_EXP(-tsum);
  sip = sip * factor;
  sqp = sqp * factor;
  sup = sup * factor;
  svp = svp * factor;
  if (P0->iflag == 1) {
    double bmu, sinbm, phinew, ri1;
    /* phie = 0 (reapar.f90:494) is what peeloff_3d.f90:168 passes as the
     * observer azimuth whatever PHIE holds. */
    hco_findangle3(P0->cost, P0->sint, P0->phi, P0->cosp, P0->sinp, coste, sinte, 0.0, cospe,
                   sinpe, &bmu, &sinbm, &phinew, &ri1);
    hco_stokespeel(&sip, &sqp, &sup, &svp, &coste, &sinte, ri1, bmu, P0->cost, P0->sint, &P0->op,
                   idust2);
  }
  yimage = zp1 * sinte - yp1 * coste * sinpe - xp1 * coste * cospe;
  ximage = yp1 * cospe - xp1 * sinpe;
  ix = (int)((ximage + s->rmaxi) / (2.0 * s->rmaxi / (double)s->nxhst)) + 1;
  iy = (int)((yimage + s->rmaxi) / (2.0 * s->rmaxi / (double)s->nxhst)) + 1;
  /* peeloff_3d.f90:179-211: the filter-convolved images.  For every filter
   * whose node range holds the ray's frequency strictly, the weight is the
   * response at the last node below it (the reference's upward scan) */
  if (s->nfilt > 0 && t->image_b_i && ix <= s->nxhst && iy <= s->nxhst && ix > 0 && iy > 0) {
    const double wave = 1.2398 / P0->nub;
    const double freq = 2.9979e14 / wave;
    int ifilt;
    for (ifilt = 0; ifilt < s->nfilt; ifilt++) {
      const double *fw = s->filtwave + (size_t)s->nwfiltmax * ifilt;
      const int nf = s->nwfilt[ifilt] - 1;
      if (freq > fw[0] && freq < fw[nf - 1]) {
        int iw = 1;
        while (freq > fw[iw - 1]) iw++;
        iw = iw - 1;
        const double wt = s->filtphi[(size_t)s->nwfiltmax * ifilt + (size_t)(iw - 1)];
        const size_t q = (size_t)peelid +
                         (size_t)s->npeel * ((size_t)(ix - 1) +
                                             (size_t)s->nxhst * ((size_t)(iy - 1) +
                                                                 (size_t)s->nxhst * ifilt));
        t->image_b_i[q] += sip * wt;
        t->image_b_q[q] += sqp * wt;
        t->image_b_u[q] += sup * wt;
        t->image_b_v[q] += svp * wt;
      }
    }
  }
  rho2 = ximage * ximage + yimage * yimage;
  inub = clampi((int)((double)s->nfreq * HCO_LOG(P0->nub / s->numin) / s->lnurat) + 1, 1, s->nfreq);
  if (P0->iflag == 1)
    i_inter = 2;
  else
    i_inter = (P0->i_orig == 1) ? 1 : 3;
  indx[0] = 1;
  indx[1] = 1 + P0->i_orig;
  indx[2] = 5 + i_inter;
  if (t->peel) {
    const size_t n = (size_t)s->nfreq * s->npeel * s->nap * HCO_NO;
    for (ia = 0; ia < s->nap; ia++) {
      if (rho2 < s->aperture2[ia]) {
        for (k = 0; k < 3; k++) {
          size_t q = (size_t)(inub - 1) +
                     (size_t)s->nfreq * ((size_t)peelid + (size_t)s->npeel *
                                         ((size_t)ia + (size_t)s->nap * (size_t)(indx[k] - 1)));
          t->peel[q] += sip;
          t->peel[q + n] += sqp;
          t->peel[q + 2 * n] += sup;
          t->peel[q + 3 * n] += svp;
          t->peel[q + 4 * n] += sip * sip;
          t->peel[q + 5 * n] += sqp * sqp;
          t->peel[q + 6 * n] += sup * sup;
          t->peel[q + 7 * n] += svp * svp;
          t->peel[q + 8 * n] += 1.0;
        }
      }
    }
  }
  if (s->imcube && t->img_i && ix <= s->nxhst && iy <= s->nxhst && ix > 0 && iy > 0) {
    const size_t nx = (size_t)s->nxhst;
    size_t q = (size_t)peelid + (size_t)s->npeel * ((size_t)(ix - 1) + nx * ((size_t)(iy - 1) +
                                                     nx * (size_t)(inub - 1)));
    const size_t n4 = (size_t)s->npeel * nx * nx * s->nfreq;
    for (k = 0; k < 3; k++) t->img_i[q + n4 * (size_t)(indx[k] - 1)] += sip;
    t->img_q[q] += sqp;
    t->img_u[q] += sup;
    t->img_v[q] += svp;
  }
}

static void peel_all(ctx_t *c, const pkt_t *P, double sip, int idust2) {
  int k;
  for (k = 0; k < c->s->npeel; k++) peeloff(c, P, sip, idust2, k);
}

static void set_kapd_launch(ctx_t *c, pkt_t *P) {  /* newdisktherm.F90:310-312 */
  int id;
  for (id = 0; id < HCO_NDG; id++) P->kapd[id] = P->op.kappa[id] * c->rmincgs * c->d->kappav[id];
}

static void set_kapd_reemit(ctx_t *c, pkt_t *P) {  /* propagate.f90:149-151 */
  int id;
  for (id = 0; id < HCO_NDG; id++)
    P->kapd[id] = P->op.kappa[id] * c->s->rstar * RSOL * c->d->kappav[id];
}

/* initp_sample with limb=0 (initp_common.f90:1-29). */
/* darkening (samptabl.f90:29-66): locate the draw in the table of the
 * cumulative probability prob(mu) = (mu^3 + mu^2) / 2, then interpolate
 * linearly in mu from the bin's upper node (xmu(1) = 0, xmu(i) = i delmu
 * otherwise, as `table` fills it). */
double hco_darkening(const hco_setup *s, double xi) {
  int n = hco_locate(s->prob, s->ntab, xi);
  if (n < 1) n = 1;
  if (n > s->ntab - 1) n = s->ntab - 1;  /* a draw of 1 would read past the table */
  return s->xmu[n] - (s->prob[n] - xi) / (s->prob[n] - s->prob[n - 1]) * s->delmu;
}

/* initp_sample (initp_common.f90:1-29): mu = sqrt(xi) for isotropic
 * intensity, or from the limb-darkening table. */
static void initp_sample(ctx_t *c, pkt_t *P) {
  double xran = hco_ran(&c->rng);
  P->cost = (c->s && c->s->limb) ? hco_darkening(c->s, xran) : sqrt(xran);
  P->sint = sqrt(1.0 - P->cost * P->cost);
  xran = hco_ran(&c->rng);
  P->phi = R2P * xran;
  P->cosp = HCO_COS(P->phi);
  P->sinp = HCO_SIN(P->phi);
}

/* initp_transform (initp_common.f90:31-94), including its clamp of costi
 * guarded by cost < -1 as written. */
static void initp_transform(pkt_t *P, double cosb, double sinb, double lp) {
  if (fabs(cosb) < 0.9999999) {
    double costi, sinti, cospn, phinew;
    costi = P->cost * cosb - P->sint * sinb * P->cosp;
    if (fabs(costi) > 1.0) {
      if (costi > 1.0) costi = 1.0;
      if (P->cost < -1.0) costi = -1.0;
    }
    sinti = sqrt(1.0 - costi * costi);
    if (sinti > 0.0000001) {
      cospn = (P->cost - costi * cosb) / (sinti * sinb);
      if (fabs(cospn) > 1.0) cospn = (cospn < 0.0) ? -1.0 : 1.0;
      if (P->phi < PI) {
        phinew = HCO_ACOS(cospn);
        if (phinew > PI) phinew = PI;
      } else {
        phinew = R2P - HCO_ACOS(cospn);
        if (phinew < PI) phinew = PI;
      }
    } else {
      phinew = 0.0;
    }
    P->phi = phinew;
    P->cost = costi;
    P->sint = sinti;
  } else {
    if (cosb < 0.0) P->cost = -P->cost;
  }
  P->phi = P->phi + lp;
  if (P->phi > R2P) P->phi = P->phi - R2P;
  if (P->phi < 0.0) P->phi = P->phi + R2P;
  P->sinp = HCO_SIN(P->phi);
  P->cosp = HCO_COS(P->phi);
}

/* Leaf entry for the tests: the emission direction of a packet leaving the
 * stellar surface at (cosb, sinb, lp), initp_sample then initp_transform. */
void hco_initp_direction(hco_rng *r, double cosb, double sinb, double lp, double *cost,
                         double *sint, double *phi) {
  ctx_t c;
  pkt_t P;
  memset(&c, 0, sizeof c);
  memset(&P, 0, sizeof P);
  c.rng = *r;
  initp_sample(&c, &P);
  initp_transform(&P, cosb, sinb, lp);
  *r = c.rng;
  *cost = P.cost;
  *sint = P.sint;
  *phi = P.phi;
}

static void bin_init(ctx_t *c, double nub, int i_orig_init) {
  const hco_setup *s = c->s;
  int inub;
  if (!c->t->si_init) return;
  inub = clampi((int)((double)s->nfreq * HCO_LOG(nub / s->numin) / s->lnurat) + 1, 1, s->nfreq);
  c->t->si_init[(size_t)(inub - 1) + (size_t)s->nfreq * (size_t)(i_orig_init - 1)] += 1.0;
  c->t->si_init[(size_t)(inub - 1)] += 1.0;
}

/* Direct stellar peel (newdisktherm.F90:342-377). */
static void peel_direct(ctx_t *c, pkt_t *P, double cosb, double sinb, double coslp,
                        double sinlp) {
  const hco_setup *s = c->s;
  int k;
  for (k = 0; k < s->npeel; k++) {
    if (!s->ips) {
      double cosnorm = cosb * s->coste[k] +
                       (sinb * s->sinte[k] * (s->cospe[k] * coslp + s->sinpe[k] * sinlp));
      /* isotropic intensity: energy per sr ~ mu; Eddington limb darkening:
       * intensity ~ (1 + 3/2 mu), another factor mu (newdisktherm.F90:349-355) */
      double sipnew = s->limb ? 12.0 / 5.0 * P->sip * (cosnorm + cosnorm * cosnorm)
                              : 4.0 * P->sip * cosnorm;
      if (cosnorm > 0.0) peeloff(c, P, sipnew, 0, k);
    } else {
      peeloff(c, P, P->sip * 2.0, 0, k);
    }
  }
}

/* Species masks of grid_mod.f90:15-20 (0-based ids). */
static const int IS_DISK[HCO_NDG] = {1, 1, 0, 0, 1, 1, 0, 0};
static const int IS_THERMAL_DISK[HCO_NDG] = {1, 1, 0, 0, 0, 0, 0, 0};

/* rantheta, random.f90:28-47: a uniform azimuth from a point in the unit
 * disk (sin and cos of it). */
static void rantheta(hco_rng *r, double *st, double *ct) {
  double x, y, x2, y2, r2 = 2.0, rr;
  while (r2 > 1.0) {
    x = hco_ran(r);
    x = x + x - 1.0;
    y = hco_ran(r);
    y = y + y - 1.0;
    x2 = x * x;
    y2 = y * y;
    r2 = x2 + y2;
  }
  rr = sqrt(r2);
  *st = y / rr;
  *ct = x / rr;
}

/* gasdev, random.f90:140-159 (polar Box-Muller).  The reference keeps the
 * second deviate of a pair for its next call, across packets; each packet
 * here draws from its own generator segment, so every call draws a fresh
 * pair and returns the deviate the reference returns first. */
static double gasdev(hco_rng *r) {
  double v1, v2, rsq, fac;
  for (;;) {
    v1 = 2.0 * hco_ran(r) - 1.0;
    v2 = 2.0 * hco_ran(r) - 1.0;
    rsq = v1 * v1 + v2 * v2;
    if (rsq < 1.0 && rsq > 0.0) break;
  }
  fac = sqrt(-2.0 * HCO_LOG(rsq) / rsq);
  return v2 * fac;
}

void hco_acc_position(const hco_setup *s, hco_rng *r, double *xyz) {
  double rp, winv, sp, cp, zsum, zp;
  int id;
  for (;;) {
    rp = 0.0;
    while (rp >= s->rmaxdmax || rp <= s->rddustmin) {
      winv = hco_ran(r);
      while (hco_ran(r) > 1.0 - sqrt(winv)) winv = hco_ran(r);
      rp = 1.0 / winv;
    }
    rantheta(r, &sp, &cp);
    xyz[0] = rp * cp;
    xyz[1] = rp * sp;
    zsum = 0.0;
    for (id = 0; id < HCO_NDG; id++)
      if (IS_DISK[id]) zsum = zsum + s->fmass[id] * s->z1[id] * HCO_POW(rp, s->bflare[id]);
    zp = zsum * gasdev(r);
    xyz[2] = zp;
    if (zp * zp + rp * rp < s->rmax * s->rmax) break;
  }
}

/* costarr of gridset.f90:580-599: exact at the poles and at a midplane
 * wall, the C library's cos elsewhere (as the library computes it). */
static double cost_wall(const hco_grid *g, int it) {
  if (g->tan2arr[it] == 0.0) return it == 0 ? 1.0 : -1.0;
  if (g->tan2arr[it] == -1.0) return 0.0;
  return cos(g->thetarr[it]);
}

static int locate_cost(const hco_grid *g, double cosb) {  /* nr.f:131-153 on costarr */
  int jl = 0, ju = g->ntg + 1, jm;
  while (ju - jl > 1) {
    jm = (ju + jl) / 2;
    if ((cost_wall(g, g->ntg - 1) > cost_wall(g, 0)) == (cosb > cost_wall(g, jm - 1)))
      jl = jm;
    else
      ju = jm;
  }
  if (jl == 0 && cosb == cost_wall(g, 0)) jl = 1;
  return jl;
}

/* emit_common.f90:57-70: the temperature a thermal packet re-emits at, for
 * an absorption already counted in nabs[k].  Lucy branch: the cell's
 * temperature.  Bjorkman-Wood branch (ilucy = 0): the cell's temperature is
 * first brought up to date with its absorptions so far, in place (the pass
 * reads it from then on, the walk's sublimation mask included), and the
 * packet re-emits at the mean of the old and new temperatures (3/4 new) on
 * the cell's first absorption, at the new one after. */
static double reemit_temperature(ctx_t *c, size_t k, int id) {
  const hco_grid *g = c->g;
  if (!c->s->bw) return g->tdust[k];
  double *td = (double *)g->tdust;
  const double na = c->t->nabs[k], told = td[k];
  const double tnew =
      hco_dusttemp(c->d, na, told, c->s->ltot, c->nphot, g->massarr[k], id);
  td[k] = tnew;
  return na < 2.0 ? 0.25 * told + 0.75 * tnew : tnew;
}

/* reemit.f90:30-34: lucydustfreq, or dustfreq in the Bjorkman-Wood branch */
static double reemit_freq(ctx_t *c, double tem, int id) {
  return c->s->bw ? hco_dustfreq(c->d, &c->rng, tem, id, c->g->tsub)
                  : hco_lucydustfreq(c->d, &c->rng, tem, id);
}

static void set_kapd_launch(ctx_t *c, pkt_t *P);
static void bin_init(ctx_t *c, double nub, int i_orig_init);
static void peel_all(ctx_t *c, const pkt_t *P, double sip, int idust2);

/* isotrp.f90:1-36: a direction about the unit vector s with density
 * proportional to |u.s|: points drawn in the cube [-1,1]^3 until
 * |p|^4 <= |p.s| (inside the unit ball, |p|^3 <= cos of its angle to s),
 * reversed when they point against s, normalised. */
void hco_isotrp(hco_rng *r, double sux, double suy, double suz, double *u) {
  double x = 0.0, y = 0.0, z = 0.0, r2 = 2.0, mur = 0.0, rr;
  while (r2 * r2 > fabs(mur)) {
    x = hco_ran(r);
    x = x + x - 1.0;
    y = hco_ran(r);
    y = y + y - 1.0;
    z = hco_ran(r);
    z = z + z - 1.0;
    mur = x * sux + y * suy + z * suz;
    r2 = x * x + y * y + z * z;
  }
  rr = sqrt(r2);
  if (mur < 0.0) rr = -rr;
  u[0] = x / rr;
  u[1] = y / rr;
  u[2] = z / rr;
}

/* The diffusion mask and direction of a cell; cells off the grid's theta
 * range count as outside the layer. */
static inline int dif_at(const hco_grid *g, int ir, int it, int ip) {
  if (ir < 1 || ir > g->nrg || it < 1 || it > g->ntg) return 0;
  return g->diffus[cell_index(g, ir, it, ip)] != 0;
}
static inline int dir_at(const hco_grid *g, int ir, int it, int ip) {
  if (ir < 1 || ir > g->nrg || it < 1 || it > g->ntg) return 0;
  return g->diffdir[cell_index(g, ir, it, ip)];
}
static double sint_wall(const hco_grid *g, int it) {  /* gridset.f90:581-599 */
  if (g->tan2arr[it] == 0.0) return 0.0;
  if (g->tan2arr[it] == -1.0) return 1.0;
  return sin(g->thetarr[it]);
}
static inline double pow4(double x) {  /* x**4 as (x*x)*(x*x) */
  x = x * x;
  return x * x;
}

/* select_dust(is_thermal_disk) (select_dust.f90) with per-species weights w
 * (the packet's kapd); 0 when the cell holds no thermal disk dust. */
static int select_thermal_disk(ctx_t *c, size_t cell, const double *w) {
  const double *dens = c->g->densarr;
  double xran = hco_ran(&c->rng), opac = 0.0, frac = 0.0;
  int id;
  for (id = 0; id < HCO_NDG; id++)
    if (IS_THERMAL_DISK[id]) opac = opac + w[id] * dens[cell + c->ncell * id];
  if (opac == 0.0) return 0;
  for (id = 0; id < HCO_NDG; id++)
    if (IS_THERMAL_DISK[id]) {
      frac = frac + w[id] * dens[cell + c->ncell * id] / opac;
      if (xran <= frac + (double)1.e-10f) return id + 1;
    }
  return 0;
}

static void set_direction(pkt_t *P, const double *u) {  /* emit_common.f90:116-121 */
  P->ux = u[0];
  P->uy = u[1];
  P->uz = u[2];
  P->cost = P->uz;
  P->sint = sqrt(1.0 - P->cost * P->cost);
  P->cosp = P->ux / P->sint;
  P->sinp = P->uy / P->sint;
  P->phi = HCO_ATAN2(P->sinp, P->cosp);
}

/* emit_common.f90:80-387, the diffusion branch (Lucy mode, ilucy = 1) for a
 * packet absorbed in (or launched into) a cell of the diffusion layer:
 *  - radial direction (diffdir -1): inwards to the layer's inner radial
 *    wall; the packet is moved radially onto it (on_wall) and leaves in a
 *    direction about the inward normal with density ~ cos (isotrp);
 *  - theta direction (diffdir -2 above the midplane, +2 below): along the
 *    theta cells to the layer's surface cone, the packet moved in theta to
 *    cos(theta) = 1.00001 cos of that wall (just outside the layer), and
 *    leaving about the cone's outward normal;
 *  - the surface cell of the two disk species gains
 *    4 pi T*^4 / (nstar dA) in T^4 (dA the surface's ring area), capped at
 *    the sublimation temperature, and the cells behind it along the
 *    direction take max(T, T_surface^4 + the column sum of
 *    3 pi T*^4 / (nstar dA) nacc kappa_V R* rho chiR(T) ds) (nacc the absorbed
 *    counts still behind, which only accretion photons add);
 *  - the emitting species is chosen among the disk species of the surface
 *    cell with the weights w (the packet's kapd), and the packet's energy
 *    from the Planck function (bbfreq) at the grey-atmosphere temperature
 *    (T_surface^4 + 3/4 mu T_eff^4)^(1/4), mu = the cosine to the normal.
 * The reference's loops over the layer test diffus of the surface cell,
 * which every branch reaches, so they run on diffdir alone; that is kept.
 * Returns the emitting species (1-based) or 0 on an unknown direction or
 * a surface cell without disk dust (where the reference stops). */
static int emit_diffusion(ctx_t *c, pkt_t *P, const double *w) {
  const hco_grid *g = c->g;
  const hco_dust *d = c->d;
  const hco_setup *s = c->s;
  const double tsube = 1600.0 / 11605.0;
  double *td = (double *)g->tdust;
  const double *nabs = c->t->nabs;
  const double *ra = g->rarr;
  int ir = P->ir, it = P->it, ip = P->ip, id, i, idust2;
  double sux, suy, suz, u[3], dA, T4s, T4d, T4, T4eff, nacc = 0.0, mu, T;
  const int dd = dir_at(g, ir, it, ip);
  T4s = s->tstar * s->tstar;
  T4s = T4s * T4s;
  if (dd == -1) {
    double rnew;
    while (dir_at(g, ir, it, ip) == -1 && dif_at(g, ir, it, ip)) ir = ir - 1;
    ir = ir + 1;
    rnew = ra[ir - 1];
    P->on_wall = 1;
    P->xp = P->xp * rnew / P->rtot;
    P->yp = P->yp * rnew / P->rtot;
    P->zp = P->zp * rnew / P->rtot;
    P->rtot = rnew;
    P->rsq = rnew * rnew;
    sux = -P->xp / P->rtot;
    suy = -P->yp / P->rtot;
    suz = -P->zp / P->rtot;
    hco_isotrp(&c->rng, sux, suy, suz, u);
    set_direction(P, u);
    dA = 2.0 * PI * (ra[ir - 1] * ra[ir - 1]) * (cost_wall(g, it - 1) - cost_wall(g, it));
    for (id = 0; id < 2; id++) {
      const size_t k0 = cell_index(g, ir, it, ip) + c->ncell * id;
      T4 = pow4(td[k0]) + 4.0 * PI * T4s / (c->nstar * dA * 1.0);
      td[k0] = fmin(tsube, HCO_POW(T4, 0.25));
      nacc = 0.0;
      i = ir;
      while (dir_at(g, i, it, ip) == -1 && dif_at(g, ir, it, ip)) {
        nacc = nacc + nabs[cell_index(g, i, it, ip) + c->ncell * id];
        i = i + 1;
      }
      T4d = T4;
      i = ir + 1;
      while (dir_at(g, i, it, ip) == -1 && dif_at(g, ir, it, ip)) {
        const size_t km = cell_index(g, i - 1, it, ip) + c->ncell * id;
        const size_t ki = cell_index(g, i, it, ip) + c->ncell * id;
        double told, tnew;
        T4d = T4d + (3.0 * PI * T4s / (c->nstar * dA * 1.0)) * nacc * d->kappav[id] * RSOL *
                        s->rstar * g->densarr[km] * hco_chir(d, td[km], id) * (ra[i - 1] - ra[i - 2]);
        told = td[ki];
        tnew = fmin(tsube, HCO_POW(T4d, 0.25));
        td[ki] = fmax(tnew, told);
        nacc = nacc - nabs[km];
        i = i + 1;
      }
    }
    idust2 = select_thermal_disk(c, cell_index(g, ir, it, ip), w);
    if (idust2 == 0) return 0;
    T4 = pow4(td[cell_index(g, ir, it, ip) + c->ncell * (idust2 - 1)]);
    T4eff = 4.0 * PI * T4s * nacc / (1.0 * dA * c->nstar);
    mu = P->ux * sux + P->uy * suy + P->uz * suz;
    T = fmin(HCO_POW(T4 + 0.75 * mu * T4eff, 0.25), tsube);
    P->nub = hco_bbfreq(&c->rng, T);
    P->ir = ir - 1;
    return idust2;
  }
  if (dd == -2 || dd == 2) {
    double cosb, sinb, cosbnew, sinbnew, sth;
    if (dd == -2) {
      while (dif_at(g, ir, it, ip)) it = it - 1;
      it = it + 1;
    } else {
      while (dif_at(g, ir, it, ip)) it = it + 1;
    }
    cosb = P->zp / P->rtot;
    sinb = sqrt(1.0 - cosb * cosb);
    cosbnew = 1.00001 * cost_wall(g, it - 1);
    sinbnew = sqrt(1.0 - cosbnew * cosbnew);
    P->xp = P->xp * sinbnew / sinb;
    P->yp = P->yp * sinbnew / sinb;
    P->zp = P->rtot * cosbnew;
    mu = cost_wall(g, it - 1);
    sth = sint_wall(g, it - 1);
    if (mu > 0.0) {
      sux = -(mu * P->xp) / sth / P->rtot;
      suy = -(mu * P->yp) / sth / P->rtot;
      suz = sth;
    } else {
      sux = mu * P->xp / sth / P->rtot;
      suy = mu * P->yp / sth / P->rtot;
      suz = -sth;
    }
    hco_isotrp(&c->rng, sux, suy, suz, u);
    set_direction(P, u);
    if (cost_wall(g, it - 1) < 0.0) it = it - 1;
    dA = PI * sth * (ra[ir] * ra[ir] - ra[ir - 1] * ra[ir - 1]);
    for (id = 0; id < 2; id++) {
      const size_t k0 = cell_index(g, ir, it, ip) + c->ncell * id;
      T4 = pow4(td[k0]) + 4.0 * PI * T4s / (c->nstar * dA * 1.0);
      td[k0] = fmin(tsube, HCO_POW(T4, 0.25));
      T4d = T4;
      if (dir_at(g, ir, it, ip) == -2) {
        nacc = 0.0;
        i = it;
        while (dir_at(g, ir, i, ip) == -2 && dif_at(g, ir, it, ip)) {
          nacc = nacc + nabs[cell_index(g, ir, i, ip) + c->ncell * id];
          i = i + 1;
        }
        i = it + 1;
        while (dir_at(g, ir, i, ip) == -2 && dif_at(g, ir, it, ip)) {
          const size_t km = cell_index(g, ir, i - 1, ip) + c->ncell * id;
          const size_t ki = cell_index(g, ir, i, ip) + c->ncell * id;
          double told, tnew;
          T4d = T4d + (3.0 * PI * T4s / (c->nstar * dA * 1.0)) * nacc * d->kappav[id] * RSOL *
                          s->rstar * g->densarr[km] * hco_chir(d, td[km], id) * 0.5 *
                          (ra[ir - 1] + ra[ir]) * (g->thetarr[i - 1] - g->thetarr[i - 2]);
          told = td[ki];
          tnew = fmin(tsube, HCO_POW(T4d, 0.25));
          td[ki] = fmax(told, tnew);
          nacc = nacc - nabs[km];
          i = i + 1;
        }
      } else {
        nacc = 0.0;
        i = it;
        while (dir_at(g, ir, i, ip) == 2 && dif_at(g, ir, it, ip)) {
          nacc = nacc + nabs[cell_index(g, ir, i, ip) + c->ncell * id];
          i = i - 1;
        }
        i = it - 1;
        while (dir_at(g, ir, i, ip) == 2 && dif_at(g, ir, it, ip)) {
          const size_t kp = cell_index(g, ir, i + 1, ip) + c->ncell * id;
          const size_t ki = cell_index(g, ir, i, ip) + c->ncell * id;
          double told, tnew;
          T4d = T4d + (3.0 * PI * T4s / (c->nstar * dA)) * nacc * d->kappav[id] * RSOL *
                          s->rstar * g->densarr[kp] * hco_chir(d, td[kp], id) * 0.5 *
                          (ra[ir - 1] + ra[ir]) * (g->thetarr[i + 1] - g->thetarr[i]);
          told = td[ki];
          tnew = fmin(tsube, HCO_POW(T4d, 0.25));
          td[ki] = fmax(told, tnew);
          nacc = nacc - nabs[kp];
          i = i - 1;
        }
      }
    }
    idust2 = select_thermal_disk(c, cell_index(g, ir, it, ip), w);
    if (idust2 == 0) return 0;
    T4 = pow4(td[cell_index(g, ir, it, ip) + c->ncell * (idust2 - 1)]);
    T4eff = 4.0 * PI * T4s * nacc / (dA * c->nstar * 1.0);
    mu = P->ux * sux + P->uy * suy + P->uz * suz;
    T = fmin(HCO_POW(T4 + 0.75 * mu * T4eff, 0.25), tsube);
    P->nub = hco_bbfreq(&c->rng, T);
    if (dir_at(g, ir, it, ip) == -2)
      it = it - 1;
    else
      it = it + 1;
    P->it = it;
    return idust2;
  }
  return 0;
}

/* Disk-accretion launch: initpacc.f90:3-122 without the diffusion layer,
 * emit_common.f90:25-78 (Lucy branch) or emit_gasacc.f90:1-44, then
 * newdisktherm.F90:219-250.  The reference picks the emitting species with
 * the kapd the previous packet left (select_dust reads the module global);
 * a packet-independent restatement weights the thermal disk species by
 * kappa_V(id) * rho(id), which is that choice whenever the species share one
 * dust table (kapd(id) = kappa(nu) R* kappa_V(id) for any nu). */
static int launch_acc(ctx_t *c, pkt_t *P) {
  const hco_setup *s = c->s;
  const hco_grid *g = c->g;
  const hco_dust *d = c->d;
  hco_tallies *t = c->t;
  double xyz[3], cosb, lp, dsum = 0.0, opac = 0.0, frac = 0.0, xran, tgas, t0;
  int ir, it, ip, id, idust2 = -1;
  size_t cell;
  hco_acc_position(s, &c->rng, xyz);
  P->xp = xyz[0];
  P->yp = xyz[1];
  P->zp = xyz[2];
  P->sip = 1.0;
  P->sqp = P->sup = P->svp = 0.0;
  P->rsq = P->xp * P->xp + P->yp * P->yp + P->zp * P->zp;
  P->rtot = sqrt(P->rsq);
  ir = hco_locate(g->r2arr, g->nrg, P->rsq);
  cosb = P->zp / P->rtot;
  it = locate_cost(g, cosb);
  lp = HCO_ATAN2(P->yp, P->xp);
  if (lp < 0.0) lp = lp + R2P;
  ip = hco_locate(g->phiarr, g->npg, lp);
  P->ir = ir;
  P->it = it;
  P->ip = ip;
  cell = cell_index(g, ir, it, ip);
  for (id = 0; id < HCO_NDG; id++)
    if (IS_THERMAL_DISK[id]) dsum = dsum + g->densarr[cell + c->ncell * id];
  if (dsum > 0.0) {  /* select_dust(is_thermal_disk), select_dust.f90:17-42 */
    xran = hco_ran(&c->rng);
    for (id = 0; id < HCO_NDG; id++)
      if (IS_THERMAL_DISK[id]) opac = opac + d->kappav[id] * g->densarr[cell + c->ncell * id];
    for (id = 0; id < HCO_NDG; id++) {
      if (IS_THERMAL_DISK[id]) {
        frac = frac + d->kappav[id] * g->densarr[cell + c->ncell * id] / opac;
        if (xran <= frac + (double)1.e-10f) {
          idust2 = id;
          break;
        }
      }
    }
  }
  P->on_wall = 0;
  if (idust2 >= 0 && t->dbig) t->dbig[cell] += 1;  /* initpacc.f90:95-101 */
  if (g->diffus && g->diffus[cell]) {
    /* initpacc.f90:103-112: the absorption shared by the two disk species in
     * proportion to kappa rho (kapd restated as kappa_V, as above), then
     * emit_common's diffusion branch from the layer's surface */
    double w[HCO_NDG] = {0};
    const double op2 = d->kappav[0] * g->densarr[cell] + d->kappav[1] * g->densarr[cell + c->ncell];
    if (op2 == 0.0) return -1;
    for (id = 0; id < 2; id++) {
      w[id] = d->kappav[id];
      if (t->nabs) t->nabs[cell + c->ncell * id] += d->kappav[id] * g->densarr[cell + c->ncell * id] / op2;
    }
    if (emit_diffusion(c, P, w) == 0) return -1;
  } else if (idust2 >= 0) {  /* initpacc.f90:113-115 (the thermal disk species are not VSG/PAH) */
    if (t->nabs) t->nabs[cell + c->ncell * idust2] += 1.0;
    /* emit_common.f90:57-70 -> reemit.f90:15-33 */
    const double tem = reemit_temperature(c, cell + c->ncell * idust2, idust2);
    P->cost = 1.0 - 2.0 * hco_ran(&c->rng);
    P->sint = sqrt(1.0 - HCO_POW(P->cost, 2.0));
    P->phi = R2P * hco_ran(&c->rng);
    P->cosp = HCO_COS(P->phi);
    P->sinp = HCO_SIN(P->phi);
    P->nub = reemit_freq(c, tem, idust2);
  } else {
    /* emit_gasacc.f90:22-42: Pringle (1981) surface temperature */
    t0 = 3.0 * GN * (MSOL * s->massc) * s->mdot /
         (8.0 * PI * SIGT * HCO_POW(P->rtot * s->rstar * RSOL, 3.0));
    tgas = HCO_POW(t0 * (1.0 - sqrt(1.0 / P->rtot)), 0.25) / 11605.0;
    P->nub = hco_bbfreq(&c->rng, tgas);
    P->cost = 1.0 - 2.0 * hco_ran(&c->rng);
    P->sint = sqrt(1.0 - HCO_POW(P->cost, 2.0));
    P->phi = R2P * hco_ran(&c->rng);
    P->cosp = HCO_COS(P->phi);
    P->sinp = HCO_SIN(P->phi);
  }
  if (!(g->diffus && g->diffus[cell])) {  /* emit_common.f90:74-76 */
    P->uy = P->sint * P->sinp;
    P->ux = P->sint * P->cosp;
    P->uz = P->cost;
  }
  /* newdisktherm.F90:219-250 */
  hco_opacset(d, P->nub, &P->op);
  bin_init(c, P->nub, 4);
  set_kapd_launch(c, P);
  P->iflag = 0;
  P->i_orig = 2;
  if (c->ipeel) peel_all(c, P, P->sip, idust2 < 0 ? 0 : idust2);
  return 0;
}

/* Star / hot-spot launch: initpspot.f90:3-175 (or initp_ps.f90:3-79 for the
 * point-source mode), then newdisktherm.F90:256-337.  hotspot selects the
 * hot-spot branch of the photon loop (newdisktherm.F90:256-273): the packet
 * starts inside the spot when ispot=1, is emitted at tshock and half of such
 * packets are X-rays.  Stellar packets take the stellar spectrum: the
 * atmosphere table when one is set (atmosfreq), else the tstar blackbody. */
/* The surface point of a stellar packet (initpspot.f90:3-120): uniform on
 * the star, or, for a hot-spot packet with ispot=1, uniform within the spot
 * caps by rejection; a stellar packet that lands in a cap becomes a spot
 * packet.  Returns ispotph. */
static int star_point(const hco_setup *s, hco_rng *r, int hotspot, double *cosb_out,
                      double *sinb_out, double *lp_out) {
  double cosb, sinb, lp, xran;
  int ispotph = hotspot;
  /* the spot rejection ends after HCO_MAX_TRIALS tries (the device's
   * kMaxTrials), keeping the last point */
  for (unsigned ntry = 1;; ++ntry) {
    for (;;) {
      xran = hco_ran(r);
      cosb = 1.0 - 2.0 * xran;
      sinb = sqrt(1.0 - cosb * cosb);
      if (cosb != 0.0 && fabs(cosb) < 1.0) break;
    }
    xran = hco_ran(r);
    lp = R2P * xran;
    if (s->ispot == 1 && ispotph == 1 && ntry < HCO_MAX_TRIALS) {
      double cphi = HCO_COS(lp), sphi = HCO_SIN(lp);
      double dotsp1 = s->scspot1 * sinb * cphi + s->ssspot1 * sinb * sphi + s->cspot1 * cosb;
      double dotsp2 = s->scspot2 * sinb * cphi + s->ssspot2 * sinb * sphi + s->cspot2 * cosb;
      if (s->nspot == 1) {
        if (dotsp1 < s->csprad) continue;
      } else {
        if (dotsp1 < s->csprad && dotsp2 < s->csprad) continue;
      }
    }
    break;
  }
  if (s->ispot == 1 && s->fspot < 1.0) {
    double dellon = (lp < PI) ? lp : R2P - lp;
    double costhp = s->cspot1 * cosb + s->sspot1 * sinb * HCO_COS(dellon);
    if (costhp > s->csprad) ispotph = 1;
    if (s->nspot == 2) {
      dellon = (lp < PI) ? PI - lp : lp - PI;
      costhp = s->cspot2 * cosb + s->sspot2 * sinb * HCO_COS(dellon);
      if (costhp > s->csprad) ispotph = 1;
    }
  }
  *cosb_out = cosb;
  *sinb_out = sinb;
  *lp_out = lp;
  return ispotph;
}

int hco_star_point(const hco_setup *s, hco_rng *r, int hotspot, double *cosb, double *sinb,
                   double *lp) {
  return star_point(s, r, hotspot, cosb, sinb, lp);
}

/* atmosfreq (starbarb.f90:120-156): the bracket hnuint(i) < xi <= hnuint(i+1)
 * by locate2's bisection, then log-log interpolation (linear in the first
 * bin, where hnuint(1) = 0). */
double hco_atmosfreq_xi(const hco_setup *s, double xi) {
  int i = hco_locate(s->hnuint, s->nhnu, xi);
  if (i < 1) i = 1;
  if (i > s->nhnu - 1) i = s->nhnu - 1;  /* xi < 1 = hnuint(nhnu): never taken */
  if (i > 1) {
    double logatm = s->lognu[i - 1] + (s->lognu[i] - s->lognu[i - 1]) *
                                          (HCO_LOG10(xi) - s->loghnu[i - 1]) /
                                          (s->loghnu[i] - s->loghnu[i - 1]);
    return HCO_EXP10(logatm);
  }
  return s->atmosnu[i - 1] +
         (s->atmosnu[i] - s->atmosnu[i - 1]) * (xi - s->hnuint[i - 1]) /
             (s->hnuint[i] - s->hnuint[i - 1]);
}

/* The photon energy of stellar emission: the atmosphere table when one is
 * set (the reference's stellar photons, iplanckst = 0), else the Planck
 * function at tstar. */
static double star_freq(ctx_t *c) {
  if (c->s->nhnu > 0) return hco_atmosfreq_xi(c->s, hco_ran(&c->rng));
  return hco_bbfreq(&c->rng, c->s->tstar);
}

static void launch_star(ctx_t *c, pkt_t *P, int hotspot) {
  const hco_setup *s = c->s;
  const hco_grid *g = c->g;
  double cosb, sinb, lp, coslp, sinlp, rp, xran;
  int ispotph = hotspot;
  P->sip = 1.0;
  P->sqp = P->sup = P->svp = 0.0;
  if (!s->ips) {
    ispotph = star_point(s, &c->rng, hotspot, &cosb, &sinb, &lp);
    initp_sample(c, P);
    initp_transform(P, cosb, sinb, lp);
  } else {
    for (;;) {
      xran = hco_ran(&c->rng);
      cosb = clamp1(1.0 - 2.0 * xran);
      sinb = clamp1(sqrt(1.0 - cosb * cosb));
      if (cosb != 0.0) break;
    }
    P->cost = cosb;
    P->sint = sinb;
    xran = hco_ran(&c->rng);
    lp = R2P * xran;
    P->phi = lp;
    P->sinp = clamp1(HCO_SIN(P->phi));
    P->cosp = clamp1(HCO_COS(P->phi));
  }
  /* initpspot.f90:147-159 / initp_ps.f90:59-69: hot-spot photons
   * (iplanckst = 1) take the Planck function at tshock inside a spot and at
   * tstar outside it; stellar photons (iplanckst = 0) take the stellar
   * spectrum outside a spot (and everywhere when fspot = 1) and tshock's
   * Planck function inside one */
  if (hotspot)
    P->nub = hco_bbfreq(&c->rng, ((s->fspot == 1.0 && !s->ips) || ispotph) ? s->tshock
                                                                           : s->tstar);
  else if (ispotph && (s->ips || s->fspot != 1.0))
    P->nub = hco_bbfreq(&c->rng, s->tshock);
  else
    P->nub = star_freq(c);
  if (hotspot) {
    if (hco_ran(&c->rng) < 0.5)
      P->nub = 1.2398 / 0.06 + hco_ran(&c->rng) * (1.2398 / 0.015 - 1.2398 / 0.06);
    bin_init(c, P->nub, 3);
  } else {
    bin_init(c, P->nub, 2);
  }
  P->i_orig = 1;
  hco_opacset(c->d, P->nub, &P->op);
  set_kapd_launch(c, P);
  coslp = HCO_COS(lp);
  sinlp = HCO_SIN(lp);
  P->zp = cosb * 1.0;
  rp = sinb * 1.0;
  P->xp = rp * coslp;
  P->yp = rp * sinlp;
  P->rsq = P->zp * P->zp + rp * rp;
  P->rtot = sqrt(P->rsq);
  P->ux = P->sint * P->cosp;
  P->uy = P->sint * P->sinp;
  P->uz = P->cost;
  P->iflag = 0;
  P->ir = 1;
  P->on_wall = 1;
  P->it = hco_locate(g->thetarr, g->ntg, HCO_ACOS(cosb));
  P->ip = hco_locate(g->phiarr, g->npg, lp);
  if (c->ipeel) peel_direct(c, P, cosb, sinb, coslp, sinlp);
}

/* A packet that reaches the stellar surface is absorbed and re-emitted by
 * the star (output.F90:378 "flux which gets absorbed (and reemitted) by
 * star").  Reconstruction: re-launch from the impact point with the stellar
 * launch's direction and spectrum sampling, counted in aflux. */
static inline const hco_grid *g_of(const ctx_t *c) { return c->g; }

static void star_reemit(ctx_t *c, pkt_t *P) {
  double cosb, sinb, lp;
  c->t->aflux += 1.0;
  cosb = clamp1(P->zp / P->rtot);
  sinb = sqrt(1.0 - cosb * cosb);
  lp = HCO_ATAN2(P->yp, P->xp);
  if (lp < 0.0) lp = lp + R2P;
  P->sip = 1.0;
  P->sqp = P->sup = P->svp = 0.0;
  initp_sample(c, P);
  initp_transform(P, cosb, sinb, lp);
  P->nub = star_freq(c);
  P->i_orig = 1;
  P->iflag = 0;
  hco_opacset(c->d, P->nub, &P->op);
  set_kapd_launch(c, P);
  P->ux = P->sint * P->cosp;
  P->uy = P->sint * P->sinp;
  P->uz = P->cost;
  P->ir = 1;
  P->on_wall = 1;
  /* the impact point can lie a rounding error outside the theta/phi cell the
   * walk arrived in (where a cone or phi wall meets the stellar surface); the
   * walk cannot find a wall it is already past, so the cell is located anew
   * from the point itself */
  P->it = hco_locate(g_of(c)->thetarr, g_of(c)->ntg, HCO_ACOS(cosb));
  P->ip = hco_locate(g_of(c)->phiarr, g_of(c)->npg, lp);
  if (c->ipeel) peel_direct(c, P, cosb, sinb, HCO_COS(lp), HCO_SIN(lp));
  if (c->tr) c->tr->nstar++;
}

/* select_dust.f90:1-67 with mask is_any; 0 when the cell has no opacity. */
int hco_select_dust(const double *kapd, const double *dens, hco_rng *r) {
  double xran = hco_ran(r), opac = 0.0, frac = 0.0;
  int id;
  for (id = 0; id < HCO_NDG; id++) opac = opac + kapd[id] * dens[id];
  if (opac == 0.0) return 0;
  for (id = 0; id < HCO_NDG; id++) {
    frac = frac + kapd[id] * dens[id] / opac;
    if (xran <= frac + (double)1.e-10f) return id + 1;
  }
  return 0;
}

void hco_dustmat(const hco_opac *o, int id, double cost, double cost2, double *p1, double *p2,
                 double *p3) {
  double p4;
  dustmat(p1, p2, p3, &p4, cost, cost2, o, id);
}

static int select_dust(ctx_t *c, const pkt_t *P, size_t cell) {
  double xran = hco_ran(&c->rng), opac = 0.0, frac = 0.0;
  int id;
  for (id = 0; id < HCO_NDG; id++) opac = opac + P->kapd[id] * c->g->densarr[cell + c->ncell * id];
  if (opac == 0.0) return 0;
  for (id = 0; id < HCO_NDG; id++) {
    frac = frac + P->kapd[id] * c->g->densarr[cell + c->ncell * id] / opac;
    if (xran <= frac + (double)1.e-10f) return id + 1;
  }
  return 0;
}

static void bin_exit(ctx_t *c, pkt_t *P, double xpold, double ypold, double zpold, int iscat) {
  const hco_setup *s = c->s;
  hco_tallies *t = c->t;
  double x, y, rho2, sini;
  int it, ip, inub, i_inter, ia, k, indx[3];
  it = (int)((double)s->nmu * (1.0 - P->cost) / 2.0) + 1;
  it = clampi(it, 1, s->nmu);
  if (P->phi < 0.0) P->phi = P->phi + R2P;
  ip = clampi((int)((double)s->nph * P->phi / R2P) + 1, 1, s->nph);
  sini = sqrt(1.0 - s->u[it - 1] * s->u[it - 1]);
  x = ypold * P->cosp - xpold * P->sinp;
  y = zpold * sini - ypold * s->u[it - 1] * P->sinp - xpold * s->u[it - 1] * P->cosp;
  rho2 = (x * x + y * y);
  if (rho2 < s->aperture2[s->nap - 1] * 1.0001) {
    t->flux = t->flux + 1.0;
    if (P->iflag == 1) {
      t->sflux = t->sflux + 1.0;
      t->nscat = t->nscat + (double)iscat;
    }
  }
  inub = clampi((int)((double)s->nfreq * HCO_LOG(P->nub / s->numin) / s->lnurat) + 1, 1, s->nfreq);
  if (P->iflag == 1)
    i_inter = 2;
  else
    i_inter = (P->i_orig == 1) ? 1 : 3;
  indx[0] = 1;
  indx[1] = 1 + P->i_orig;
  indx[2] = 5 + i_inter;
  if (t->sed) {
    const size_t n = (size_t)s->nfreq * s->nmu * s->nph * s->nap * HCO_NO;
    for (ia = 0; ia < s->nap; ia++) {
      if (rho2 < s->aperture2[ia] && !(P->iflag == 0 && P->i_orig == 4)) {
        for (k = 0; k < 3; k++) {
          size_t q = (size_t)(inub - 1) +
                     (size_t)s->nfreq *
                         ((size_t)(it - 1) +
                          (size_t)s->nmu * ((size_t)(ip - 1) +
                                            (size_t)s->nph * ((size_t)ia + (size_t)s->nap *
                                                                             (size_t)(indx[k] - 1))));
          t->sed[q] += P->sip;
          t->sed[q + n] += P->sqp;
          t->sed[q + 2 * n] += P->sup;
          t->sed[q + 3 * n] += P->svp;
          t->sed[q + 4 * n] += P->sip * P->sip;
          t->sed[q + 5 * n] += P->sqp * P->sqp;
          t->sed[q + 6 * n] += P->sup * P->sup;
          t->sed[q + 7 * n] += P->svp * P->svp;
          t->sed[q + 8 * n] += 1.0;
        }
      }
    }
  }
  if (!(P->iflag == 0 && P->i_orig == 4)) t->sumsub = t->sumsub + P->sip;
  t->sumall = t->sumall + P->sip;
  if (P->iflag == 1) t->scount = t->scount + 1.0;
}

static void finish_trace(ctx_t *c, const pkt_t *P, int status) {
  hco_trace *tr = c->tr;
  if (!tr) return;
  tr->x = P->xp;
  tr->y = P->yp;
  tr->z = P->zp;
  tr->cost = P->cost;
  tr->sint = P->sint;
  tr->phi = P->phi;
  tr->I = P->sip;
  tr->Q = P->sqp;
  tr->U = P->sup;
  tr->V = P->svp;
  tr->nub = P->nub;
  tr->status = status;
  tr->ir = P->ir;
  tr->it = P->it;
  tr->ip = P->ip;
}

/* propagate.f90:1-325: the event loop of one packet. */
static int propagate(ctx_t *c, pkt_t *P) {
  const hco_setup *s = c->s;
  const hco_grid *g = c->g;
  hco_tallies *t = c->t;
  double xpold = P->xp, ypold = P->yp, zpold = P->zp, tau;
  int iscat = 0, iint = 0, idust2 = 0, st, first = 1;
  int pir = 0, pit = 0, pip = 0;
  tau = -HCO_LOG(hco_ran(&c->rng));
  st = tauint(c, P, tau);
  for (;;) {
    while (st == ST_STAR) {
      star_reemit(c, P);
      tau = -HCO_LOG(hco_ran(&c->rng));
      xpold = P->xp;
      ypold = P->yp;
      zpold = P->zp;
      st = tauint(c, P, tau);
    }
    if (st == ST_ERROR) return HCO_ERROR;
    if (st == ST_EXIT) break;
    if (!first) {
      if (P->ir != pir || P->it != pit || P->ip != pip) iint = 0;
      if (iint > s->iintmax) {
        size_t cell = cell_index(g, P->ir, P->it, P->ip);
        if (t->killed) t->killed[cell] += 1;
        t->abflux = t->abflux + 1.0;
        return HCO_KILLED;
      }
    }
    {
      size_t cell = cell_index(g, P->ir, P->it, P->ip);
      idust2 = select_dust(c, P, cell);
      if (idust2 == 0) {
        t->n_err_dust++;
        return HCO_ERROR;
      }
      if (first) t->tot = t->tot + 1.0;
      first = 0;
      if (hco_ran(&c->rng) <= P->op.albedo[idust2 - 1]) {
        hco_scat sc;
        P->iflag = 1;
        iscat++;
        iint++;
        if (c->ipeel) peel_all(c, P, P->sip, idust2 - 1);
        sc.cost = P->cost;
        sc.sint = P->sint;
        sc.cosp = P->cosp;
        sc.sinp = P->sinp;
        sc.phi = P->phi;
        sc.I = P->sip;
        sc.Q = P->sqp;
        sc.U = P->sup;
        sc.V = P->svp;
        sc.hit = t->hit;
        sc.htot = t->htot;
        hco_stokes(&sc, &P->op, idust2 - 1, &c->rng);
        t->hit = sc.hit;
        t->htot = sc.htot;
        P->cost = sc.cost;
        P->sint = sc.sint;
        P->cosp = sc.cosp;
        P->sinp = sc.sinp;
        P->phi = sc.phi;
        P->sip = sc.I;
        P->sqp = sc.Q;
        P->sup = sc.U;
        P->svp = sc.V;
        if (c->tr) c->tr->nscat++;
      } else {
        size_t k = cell + c->ncell * (size_t)(idust2 - 1);
        if (t->dbig) t->dbig[cell] += 1;
        P->i_orig = (idust2 == 1 || idust2 == 2 || idust2 == 5 || idust2 == 6) ? 2 : 3;
        if (g->diffus && g->diffus[cell]) {
          /* propagate.f90:123-142: no count in the layer; emit_common's
           * diffusion branch moves the packet to the layer's surface */
          idust2 = emit_diffusion(c, P, P->kapd);
          if (idust2 == 0) {
            t->n_err_dust++;
            return HCO_ERROR;
          }
        } else {
          if (t->nabs) t->nabs[k] += 1.0;
          /* emit_common.f90:57-70 -> reemit.f90:15-33 */
          const double tem = reemit_temperature(c, k, idust2 - 1);
          P->sip = 1.0;
          P->sqp = P->sup = P->svp = 0.0;
          P->cost = 1.0 - 2.0 * hco_ran(&c->rng);
          P->sint = sqrt(1.0 - HCO_POW(P->cost, 2.0));
          P->phi = R2P * hco_ran(&c->rng);
          P->cosp = HCO_COS(P->phi);
          P->sinp = HCO_SIN(P->phi);
          P->nub = reemit_freq(c, tem, idust2 - 1);
          P->uy = P->sint * P->sinp;
          P->ux = P->sint * P->cosp;
          P->uz = P->cost;
        }
        P->iflag = 0;
        hco_opacset(c->d, P->nub, &P->op);
        set_kapd_reemit(c, P);
        if (c->ipeel) peel_all(c, P, P->sip, idust2 - 1);
        iint++;
        if (c->tr) c->tr->nabs++;
      }
    }
    tau = -HCO_LOG(hco_ran(&c->rng));
    xpold = P->xp;
    ypold = P->yp;
    zpold = P->zp;
    P->ux = P->sint * P->cosp;
    P->uy = P->sint * P->sinp;
    P->uz = P->cost;
    pir = P->ir;
    pit = P->it;
    pip = P->ip;
    st = tauint(c, P, tau);
  }
  bin_exit(c, P, xpold, ypold, zpold, iscat);
  return HCO_ESCAPED;
}

int hco_run_pass(const hco_grid *g, const hco_dust *d, const hco_setup *s, hco_tallies *t,
                 int64_t np, int64_t p_begin, int64_t p_end, int ipeel, int32_t seed1,
                 int32_t seed2, int gen, uint64_t seg0, hco_trace *trace) {
  ctx_t c;
  int64_t p, ns, npacc, npaccstar;
  double rn;
  memset(&c, 0, sizeof c);
  c.g = g;
  c.d = d;
  c.s = s;
  c.t = t;
  c.ipeel = ipeel;
  c.rmincgs = s->rstar * RSOL;
  c.tsub_ev = g->tsub / 11605.0;
  c.ncell = (size_t)g->nrg * g->ntg * g->npg;
  c.nphot = (double)np;
  if (s->bw && (!d->kapdBint || !t->nabs || !g->massarr)) return -1;
  if (g->diffus && (s->bw || !g->diffdir || !g->rarr || !d->kappar || !t->nabs)) return -1;
  /* source split, newdisktherm.F90:143-176 (no external illumination) */
  ns = (int64_t)((double)np * (1.0 - s->accfrac));
  rn = (double)(float)np * (1.0 - s->accfrac);
  if (rn - (double)ns > 0.5) ns = ns + 1;
  npacc = (int64_t)((double)np * s->accfrac);
  npaccstar = (int64_t)(s->accfrac2 * (double)np);
  if (ns + npacc != np) ns = ns + np - (ns + npacc);
  c.nstar = (double)ns;
  for (p = p_begin; p < p_end; p++) {
    pkt_t P;
    int64_t iphot = p + 1;
    int status;
    memset(&P, 0, sizeof P);
    hco_rng_packet(&c.rng, seed1, seed2, gen, seg0 + (uint64_t)p);
    c.tr = trace ? &trace[p - p_begin] : NULL;
    if (c.tr) {
      memset(c.tr, 0, sizeof *c.tr);
      c.tr->cellhash = 14695981039346656037ull;
    }
    if (iphot < npacc) { /* newdisktherm.F90:209 */
      if (launch_acc(&c, &P)) {  /* a layer cell without disk dust */
        t->n_err_dust++;
        finish_trace(&c, &P, HCO_ERROR);
        continue;
      }
    } else {
      launch_star(&c, &P, iphot < npacc + npaccstar ? 1 : 0);
    }
    hash_cell(c.tr, P.ir, P.it, P.ip);
    status = propagate(&c, &P);
    finish_trace(&c, &P, status);
  }
  return 0;
}

/* Leaf entries of tauint and peeloff for the closed-form tests: a packet
 * state built by the caller instead of a launch. */
static void leaf_ctx(ctx_t *c, const hco_grid *g, const hco_setup *s, hco_tallies *t) {
  memset(c, 0, sizeof *c);
  c->g = g;
  c->s = s;
  c->t = t;
  c->rmincgs = s->rstar * RSOL;
  c->tsub_ev = g->tsub / 11605.0;
  c->ncell = (size_t)g->nrg * g->ntg * g->npg;
}

static void leaf_packet(pkt_t *P, const double *xyz, const double *u, const int *cell,
                        int on_wall, const double *kapd, const double *albedo) {
  int id;
  memset(P, 0, sizeof *P);
  P->xp = xyz[0];
  P->yp = xyz[1];
  P->zp = xyz[2];
  P->ux = u[0];
  P->uy = u[1];
  P->uz = u[2];
  P->rsq = P->xp * P->xp + P->yp * P->yp + P->zp * P->zp;
  P->rtot = sqrt(P->rsq);
  P->ir = cell[0];
  P->it = cell[1];
  P->ip = cell[2];
  P->on_wall = on_wall;
  for (id = 0; id < HCO_NDG; id++) {
    P->kapd[id] = kapd[id];
    P->op.albedo[id] = albedo[id];
  }
}

int hco_tauint_ray(const hco_grid *g, const hco_setup *s, hco_tallies *t, const double *xyz,
                   const double *u, int *cell, int on_wall, const double *kapd,
                   const double *albedo, double tau, double *xyz_out) {
  ctx_t c;
  pkt_t P;
  int st;
  leaf_ctx(&c, g, s, t);
  leaf_packet(&P, xyz, u, cell, on_wall, kapd, albedo);
  st = tauint(&c, &P, tau);
  xyz_out[0] = P.xp;
  xyz_out[1] = P.yp;
  xyz_out[2] = P.zp;
  cell[0] = P.ir;
  cell[1] = P.it;
  cell[2] = P.ip;
  return st;
}

void hco_peel_ray(const hco_grid *g, const hco_setup *s, hco_tallies *t, const double *xyz,
                  const int *cell, int on_wall, const double *kapd, const double *stokes,
                  const double *dir, double nub, int iflag, int i_orig, int peelid) {
  ctx_t c;
  pkt_t P;
  const double u0[3] = {0.0, 0.0, 1.0};
  static const double alb0[HCO_NDG] = {0};
  int cl[3] = {cell[0], cell[1], cell[2]};
  leaf_ctx(&c, g, s, t);
  leaf_packet(&P, xyz, u0, cl, on_wall, kapd, alb0);
  P.sqp = stokes[1];
  P.sup = stokes[2];
  P.svp = stokes[3];
  P.cost = dir[0];
  P.sint = dir[1];
  P.phi = dir[2];
  P.cosp = HCO_COS(dir[2]);
  P.sinp = HCO_SIN(dir[2]);
  P.nub = nub;
  P.iflag = iflag;
  P.i_orig = i_orig;
  peeloff(&c, &P, stokes[0], 0, peelid);
}

/* Leaf entry of the diffusion branch (emit_diffusion) for the closed-form
 * tests: a packet at xyz in cell (1-based) with walk opacities kapd leaves
 * from the layer's surface; the grid's tdust is updated in place, the
 * tallies' nabs read.  out: x, y, z, ux, uy, uz, nub; cell and on_wall
 * updated; returns the emitting species (1-based) or 0. */
int hco_emit_diffusion(const hco_grid *g, const hco_dust *d, const hco_setup *s, hco_tallies *t,
                       hco_rng *r, double nstar, const double *xyz, int *cell, int *on_wall,
                       const double *kapd, double *out) {
  ctx_t c;
  pkt_t P;
  const double u0[3] = {0.0, 0.0, 1.0};
  static const double alb0[HCO_NDG] = {0};
  int id2;
  leaf_ctx(&c, g, s, t);
  c.d = d;
  c.nstar = nstar;
  c.rng = *r;
  leaf_packet(&P, xyz, u0, cell, *on_wall, kapd, alb0);
  id2 = emit_diffusion(&c, &P, P.kapd);
  *r = c.rng;
  out[0] = P.xp;
  out[1] = P.yp;
  out[2] = P.zp;
  out[3] = P.ux;
  out[4] = P.uy;
  out[5] = P.uz;
  out[6] = P.nub;
  cell[0] = P.ir;
  cell[1] = P.it;
  cell[2] = P.ip;
  *on_wall = P.on_wall;
  return id2;
}

/* Elementary-function probe for the accuracy tests: fn 0 log, 1 exp, 2 sin,
 * 3 cos, 4 acos, 5 pow(x,y), 6 atan2(x,y), 7 log10, 8 10^x. */
void hco_math(int fn, const double *x, const double *y, double *out, int64_t n) {
  int64_t i;
  for (i = 0; i < n; i++) {
    switch (fn) {
      case 0: out[i] = HCO_LOG(x[i]); break;
      case 1: out[i] = HCO_EXP(x[i]); break;
      case 2: out[i] = HCO_SIN(x[i]); break;
      case 3: out[i] = HCO_COS(x[i]); break;
      case 4: out[i] = HCO_ACOS(x[i]); break;
      case 5: out[i] = HCO_POW(x[i], y[i]); break;
      case 7: out[i] = HCO_LOG10(x[i]); break;
      case 8: out[i] = HCO_EXP10(x[i]); break;
      default: out[i] = HCO_ATAN2(x[i], y[i]); break;
    }
  }
}
