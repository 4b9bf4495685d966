/*
 * cpc_oracle.c -- TEST INFRASTRUCTURE ONLY: the parity oracle and CPU baseline.
 *
 * A plain-C restatement of corr_pc's cross-pair loop, written from the reference's behaviour (not
 * its text).  It follows, step for step:
 *   corels2 outer loop, image loops and inner sweep    corr.cpp:219-459 (do_pair_sorted corr.cpp:67-86)
 *   corels1 (auto pairs, params->auto_pairs)           corr.cpp:103-216: scans from j = i+1, or j = 0
 *                                                      for PBz > 0, each ending at the first included
 *                                                      galaxy outside the z window
 *   per-outer window and image flags (box)             PB_calcs.cpp:12-61 (lim_calc_xyz, lim_calc_PB)
 *   per-outer window (sky, coordinates 0-5)            sky_calcs.cpp:33-81, 105-190 (lim_calc_sky,
 *                                                      lim_calc_ra_dec); no z test for 4-5, corr.cpp:69
 *   pi and rp of a pair (r and mu, coordinates 1, 7)   PB_calcs.cpp:87-98; calcs.cpp:108-140
 *                                                      (pi_calc, pi_calc_projected, pi_calc_phi,
 *                                                      r_mu_calc); sky_calcs.cpp:24-31, 84-102
 *                                                      (dA_calc, rp_calc_sky, _projected, _angular);
 *                                                      choose_rp_pi_func, initialization.cpp:545-571
 *   bin indices, pair weight                           corels.cpp:21-46 (corel_bin)
 *   the pair kinds                                     corels.cpp:48-77, 79-120, 122-204, 206-244, 246-262
 *   ED: cos^2 of axis and separation                   PB_calcs.cpp:101-134
 *   orientation angle, rotation, ephi                  calcs.cpp:54-64, 97-105; sky_calcs.cpp:13-22
 *   jackknife touch slots and leave-one-out            jk.cpp:53-81, 138-171, 201-223; outp.cpp:242-288
 * in FP64 with the host libm (log10, sqrt, atan, sin, cos, acos, floor) and no FMA contraction, like
 * the reference's -O2 x86-64 build.  OpenMP threads own private histograms, merged in thread order.
 *
 * One deliberate difference: the reference finds the first inner galaxy of each outer z window with
 * a per-thread pointer that only moves forward (corr.cpp:361-408).  When some window holds no inner
 * galaxy but later galaxies exist, that pointer runs past the end and the reference's result depends
 * on the OpenMP schedule.  Here every window starts at its binary-searched lower bound, which equals
 * the reference whenever no such window occurs; such windows are counted in n_sparse_windows.  The
 * PBz>0 image keeps the reference's start at j=0 and break at the first galaxy out of the window.
 *
 * Nothing in the product may load this file: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker.
 */
#include <math.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <sys/time.h>

#include "../include/cpc.h"

#define NF CPC_NFIELD

static const double R2D = 180.0 / M_PI; /* r2d of calcs.cpp:19, sky_calcs.cpp:10 */

typedef struct {
  int n_rp, n_pi, R;
  int64_t *num;
  double *wt;
  double *val;
  double *err;
  int64_t *num_rp;
} ohist;

static int hist_alloc(ohist *h, int n_rp, int n_pi, int R) {
  size_t cells = (size_t)R * n_rp * n_pi;
  h->n_rp = n_rp;
  h->n_pi = n_pi;
  h->R = R;
  h->num = calloc(cells, sizeof(int64_t));
  h->wt = calloc(cells, sizeof(double));
  h->val = calloc(cells * NF, sizeof(double));
  h->err = calloc(cells * NF, sizeof(double));
  h->num_rp = calloc((size_t)R * n_rp, sizeof(int64_t));
  return h->num && h->wt && h->val && h->err && h->num_rp;
}

static void hist_free(ohist *h) {
  free(h->num);
  free(h->wt);
  free(h->val);
  free(h->err);
  free(h->num_rp);
}

/* absolute() of calcs.cpp:50-53: returns -n for n<0, n otherwise (so -0.0 stays -0.0) */
static inline double absolute_(double n) { return n < 0 ? -1 * n : n; }

/* ephi_calc, calcs.cpp:97-105 */
static inline double ephi_(double phi, double theta) {
  double t = absolute_(phi - theta);
  if (t > M_PI / 2.0) t = M_PI - t;
  return t;
}

/* elip_rotate, calcs.cpp:60-64 */
static inline void rotate_(double e1, double e2, double theta, double *ep, double *ex) {
  ep[0] = cos(2.0 * theta) * e1 - sin(2.0 * theta) * e2;
  ex[0] = sin(2.0 * theta) * e1 + cos(2.0 * theta) * e2;
}

static inline double cw(const cpc_catalog_desc *c, int64_t i) { return c->w ? c->w[i] : 1.0; }
static inline double cphi(const cpc_catalog_desc *c, int64_t i) { return c->phi ? c->phi[i] : -10.0; }
static inline double cjp(const cpc_catalog_desc *c, int64_t i) { return c->jk_prob ? c->jk_prob[i] : 0.0; }
static inline double cinc(const cpc_catalog_desc *c, int64_t i) {
  return c->include_prob ? c->include_prob[i] : 0.0;
}

/* first index j of the z-sorted catalog with z[j] + shift >= zmin (all galaxies, as the reference's
 * starts pointer scans every galaxy, corr.cpp:361-369, 395-404) */
static int64_t lower_bound_(const double *z, int64_t n, double shift, double zmin) {
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    int64_t mid = lo + (hi - lo) / 2;
    if (z[mid] + shift < zmin)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

typedef struct {
  int geom;
  double L, pmin, pmax, psize, rmin, rmax, rsize;
  int n_rp, n_pi, lin, kind, gws, est, which_corr, jk_on;
  int autop; /* corels1 auto pass */
  int unsorted; /* data_sorted 0: no scan stops (do_pair_unsorted); auto pairs j > i under every image */
  int rmu;   /* r-mu bins (coordinates 7; on the sky coordinates 1) */
  double inc_max;
  /* sky */
  const double *H, *dA;
  int64_t ntab;
  double dz, z_max, z_sep_max, z_min, z_sep_min;
  int coord; /* data_info.coordinates on the sky, 0-5 */
  int rpm;   /* rp: 0 rp_calc_sky, 1 rp_calc_sky_projected, 2 rp_calc_sky_angular */
  int pim;   /* pi: 0 pi_calc, 1 pi_calc_projected, 2 pi_calc_phi */
} oparm;

/* Shape quantities of a binned pair (corels.cpp:122-244), theta = the orientation angle.
 * v[f], e[f]: the values added to data[f].val / data[f].err; unset ones stay 0, as calc_temp_ini
 * leaves the per-thread ep/ex/epp/exx/epx/etheta that a kind never assigns. */
static inline void kind_values(const oparm *P, const cpc_catalog_desc *g1, int64_t i, const cpc_catalog_desc *g2,
                               int64_t j, double wt, double theta, double *v, double *e) {
  const double sig_crit = 1.0; /* calc_temp_ini, initialization.cpp:596 */
  if (P->kind == CPC_KIND_DENSITY_CONVERGENCE) {
    /* density_convergence_corel, corels.cpp:79-91: kappa sits in e[0]; no err */
    double ep = 0;
    if (P->gws == 1)
      ep = wt * g1->e1[i] * sig_crit;
    else if (P->gws == 2)
      ep = wt * g2->e1[j] * sig_crit;
    v[1] = ep;
  } else if (P->kind == CPC_KIND_SHAPE_CONVERGENCE) {
    /* shape_convergence_corel, corels.cpp:93-120: rotate the g_with_shape side, kappa from the other */
    double ep = 0, r0 = 0, r1 = 0;
    if (P->gws == 2) {
      rotate_(g2->e1[j], g2->e2[j], theta, &r0, &r1);
      ep = wt * g1->e1[i] * sig_crit;
    } else if (P->gws == 1) {
      rotate_(g1->e1[i], g1->e2[i], theta, &r0, &r1);
      ep = wt * g2->e1[j] * sig_crit;
    }
    v[1] = ep;
    v[3] = ep * r0;
    v[5] = ep * r1;
  } else if (P->kind == CPC_KIND_SHAPE_DENSITY) {
    /* shape_density_corel, corels.cpp:206-244 */
    double ep, ex;
    if (P->gws == 2)
      rotate_(g2->e1[j], g2->e2[j], theta, &ep, &ex);
    else
      rotate_(g1->e1[i], g1->e2[i], theta, &ep, &ex);
    v[1] = 1.0 * wt * ep;
    e[1] = (wt * ep) * (wt * ep);
    v[2] = 1.0 * wt * ex;
    e[2] = (wt * ex) * (wt * ex);
    v[6] = ephi_(cphi(g1, i), theta); /* always g1's phi, corels.cpp:229 */
    e[6] = v[6] * v[6];
  } else if (P->kind == CPC_KIND_SHAPE_SHAPE) {
    /* shape_shape_corel, corels.cpp:122-204 */
    double epi, exi, epj, exj;
    rotate_(g1->e1[i], g1->e2[i], theta, &epi, &exi);
    rotate_(g2->e1[j], g2->e2[j], theta, &epj, &exj);
    if (P->est == 0 || P->coord == 2 || P->coord == 4) { /* corels.cpp:139 */
      if (P->gws == 2) {
        v[1] = 1.0 * wt * (epj);
        e[1] = (wt * epj) * (wt * epj);
        v[6] = ephi_(cphi(g2, j), theta);
        e[6] = v[6] * v[6];
      } else if (P->gws == 1) {
        v[1] = 1.0 * wt * (epi);
        e[1] = (wt * epi) * (wt * epi);
        v[6] = ephi_(cphi(g1, i), theta);
        e[6] = v[6] * v[6];
      }
      /* ex is never assigned on this branch: data[2] receives 0 */
    } else {
      v[1] = 1.0 * wt * (epi + epj) / 2.0;
      e[1] = (wt * epi / 2.0) * (wt * epi / 2.0) + (wt * epj / 2.0) * (wt * epj / 2.0);
      v[2] = 1.0 * wt * (exi + exj) / 2.0;
      e[2] = (wt * exi / 2.0) * (wt * exi / 2.0) + (wt * exj / 2.0) * (wt * exj / 2.0);
      double t1_ = ephi_(cphi(g1, i), theta);
      double t2_ = ephi_(cphi(g2, j), theta);
      e[6] = t1_ * t1_ + t2_ * t2_;
      v[6] = t1_ + t2_;
    }
    v[3] = 1.0 * wt * (epi * epj);
    e[3] = (wt * (epi * epj)) * (wt * (epi * epj));
    v[4] = 1.0 * wt * (exi * exj);
    e[4] = (wt * (exi * exj)) * (wt * (exi * exj));
    v[5] = 1.0 * wt * (exi * epj + epi * exj) / 2.0;
    {
      double a = wt * (exi * epj) / 2.0, b = wt * epi * exj / 2.0;
      e[5] = a * a + b * b;
    }
  }
}

/* corel_bin's index tests (corels.cpp:26-38) given pi (d_P) and rp (da_R); returns 0 if rejected. */
static inline int bin_pair(const oparm *P, double d_P, double da_R, int *n1, int *n2) {
  double q2 = floor((d_P - P->pmin) / P->psize);
  if (!(q2 >= 0 && q2 < P->n_pi)) return 0;
  double q1;
  if (P->lin)
    q1 = floor((da_R - P->rmin) / P->rsize);
  else
    q1 = floor(log10(da_R / P->rmin) / P->rsize);
  if (!(q1 >= 0 && q1 < P->n_rp)) return 0;
  *n1 = (int)q1;
  *n2 = (int)q2;
  return 1;
}

/* density_density_corel's counters (corels.cpp:67-73), the kind's data fields, then corel_jk with
 * jk_check_PB or jk_check_sky and corel_jk1 (jk.cpp:53-71, 138-171, 201-223). */
static inline void accumulate(const oparm *P, ohist *h, int n1, int n2, double wt, const double *v, const double *e,
                              int jk1, int jk2, double jp1, double jp2) {
  const size_t nc = (size_t)P->n_rp * P->n_pi;
  size_t c = (size_t)n1 * P->n_pi + n2;
  h->num_rp[n1]++;
  h->num[c]++;
  h->wt[c] += wt;
  if (P->kind != CPC_KIND_DD) {
    for (int f = 1; f < NF; f++) {
      h->val[c * NF + f] += v[f];
      h->err[c * NF + f] += e[f];
    }
  }
  if (!P->jk_on) return;
  int gjk[2] = {jk1, jk2};
  double jp[2] = {jp1, jp2};
  if (gjk[0] == gjk[1]) {
    if (P->geom == CPC_GEOM_BOX) jp[0] = 1; /* jk_check_PB */
    jp[1] = 0;
  }
  for (int s = 0; s < 2; s++) {
    if (jp[s] < 0.5) continue;
    int r = 1 + gjk[s];
    size_t cr = (size_t)r * nc + c;
    h->num_rp[(size_t)r * P->n_rp + n1]++;
    h->num[cr]++;
    h->wt[cr] += wt;
    if (P->which_corr == 0) continue;
    for (int f = 1; f < NF; f++) h->val[cr * NF + f] += v[f];
  }
}

/* Box pair that passed the window predicates (pi_calc_PB, rp_calc_PB, r_mu_calc for coordinates 7,
 * corel, corel_jk). */
static inline int corel_box(const oparm *P, const cpc_catalog_desc *g1, int64_t i, const cpc_catalog_desc *g2,
                            int64_t j, int PBx, int PBy, int PBz, ohist *h) {
  const double L = P->L;
  double d_P = (g2->z[j] + L * PBz) - g1->z[i];
  double t0 = g1->x[i] - (g2->x[j] + L * PBx);
  double t1 = g1->y[i] - (g2->y[j] + L * PBy);
  double da_R = sqrt(t0 * t0 + t1 * t1);
  if (P->rmu) { /* r_mu_calc, calcs.cpp:135-140 */
    da_R = sqrt(da_R * da_R + d_P * d_P);
    d_P /= da_R;
  }
  int n1, n2;
  if (!bin_pair(P, d_P, da_R, &n1, &n2)) return 0;
  double wt = cw(g1, i) * cw(g2, j);
  double v[NF] = {0}, e[NF] = {0};
  if (P->kind == CPC_KIND_ED) {
    /* ED_corel / ED_calc_PB (corels.cpp:246-262, PB_calcs.cpp:101-134) */
    double xyz[3];
    xyz[0] = g1->x[i] - (g2->x[j] + L * PBx);
    xyz[1] = g1->y[i] - (g2->y[j] + L * PBy);
    xyz[2] = g1->z[i] - (g2->z[j] + L * PBz);
    double xyz_norm = sqrt(xyz[0] * xyz[0] + xyz[1] * xyz[1] + xyz[2] * xyz[2]);
    double cos_ang;
    if (P->gws == 2)
      cos_ang = g2->e1[j] * xyz[0] + g2->e2[j] * xyz[1] + g2->e3[j] * xyz[2];
    else
      cos_ang = g1->e1[i] * xyz[0] + g1->e2[i] * xyz[1] + g1->e3[i] * xyz[2];
    cos_ang /= xyz_norm;
    v[6] = cos_ang * cos_ang;
    e[6] = pow(v[6], 2.0);
  } else if (P->kind != CPC_KIND_DD) {
    double theta = atan((g2->y[j] - g1->y[i]) / (g2->x[j] - g1->x[i])); /* calcs.cpp:54-57 */
    kind_values(P, g1, i, g2, j, wt, theta, v, e);
  }
  accumulate(P, h, n1, n2, wt, v, e, P->jk_on ? g1->jk[i] : 0, P->jk_on ? g2->jk[j] : 0, cjp(g1, i), cjp(g2, j));
  return 1;
}

/* Sky pair that passed the window predicates (pi_calc, rp_calc_sky/dA_calc, corel, corel_jk). */
static inline int corel_sky(const oparm *P, const cpc_catalog_desc *g1, int64_t i, const cpc_catalog_desc *g2,
                            int64_t j, ohist *h) {
  const double dec1 = g1->y[i] / R2D, dec2 = g2->y[j] / R2D, ra1 = g1->x[i] / R2D, ra2 = g2->x[j] / R2D;
  /* angle_orient_axis_sky, sky_calcs.cpp:13-22 (pi_calc_phi and the shape kinds) */
  double theta = 0;
  if (P->kind != CPC_KIND_DD || P->pim == 2)
    theta = atan((g2->y[j] - g1->y[i]) / ((g2->x[j] - g1->x[i]) * cos((dec1 + dec2) / 2.0)));
  /* ct.pi_calc (calcs.cpp:108-132) */
  double d_P;
  if (P->pim == 0) {
    d_P = g2->dc[j] - g1->dc[i];
  } else if (P->pim == 1) {
    d_P = 0;
  } else {
    /* ephi_calc of the g_with_shape galaxy (g1 unless it is 2), calcs.cpp:97-105 */
    double t = (P->gws == 2 ? cphi(g2, j) : cphi(g1, i)) - theta;
    t = t < 0 ? -1 * t : t;
    if (t > M_PI / 2.0) t = M_PI - t;
    d_P = t;
  }
  /* ct.rp_calc: dA_calc, then the distance (sky_calcs.cpp:24-31, 84-102) */
  double sum = sin(dec1) * sin(dec2) + cos(dec1) * cos(dec2) * cos((ra1 - ra2));
  double da = (sum > 0.9999999999) ? acos(1.0) : acos(sum);
  double da_R;
  if (P->rpm == 0)
    da_R = da * (g1->da[i] + g2->da[j]) / 2.;
  else if (P->rpm == 1)
    da_R = da * g1->da[i];
  else
    da_R = da * R2D;
  if (P->rmu) { /* r_mu_calc, calcs.cpp:135-140 */
    da_R = sqrt(da_R * da_R + d_P * d_P);
    d_P /= da_R;
  }
  int n1, n2;
  if (!bin_pair(P, d_P, da_R, &n1, &n2)) return 0;
  double wt = cw(g1, i) * cw(g2, j);
  double v[NF] = {0}, e[NF] = {0};
  if (P->kind != CPC_KIND_DD) kind_values(P, g1, i, g2, j, wt, theta, v, e);
  accumulate(P, h, n1, n2, wt, v, e, P->jk_on ? g1->jk[i] : 0, P->jk_on ? g2->jk[j] : 0, cjp(g1, i), cjp(g2, j));
  return 1;
}

/* Linear interpolation of a distance-table column, as dAR_to_dA/p_to_dz (sky_calcs.cpp:71-81). */
static inline double tab_interp(const double *tab, double dz, double z) {
  int i_t = (int)floor(z / dz);
  return tab[i_t] + ((tab[i_t + 1] - tab[i_t]) * (z - dz * i_t) / dz);
}

static inline int tab_ok(const oparm *P, double z) {
  double f = floor(z / P->dz);
  return f >= 0 && f + 1 < (double)P->ntab;
}

/* lim_calc_sky + lim_calc_ra_dec (sky_calcs.cpp:105-190).  Returns 0 on the conditions where the
 * reference exits or reads outside its table.  Coordinates 4 and 5 make no z test (corr.cpp:69):
 * their window is unbounded in z. */
static int sky_window(const oparm *P, double ra, double dec, double z, double *w) {
  double z_max = z + P->z_sep_max, z_min, da_max;
  if (P->coord >= 2) {
    /* cases 2-5: the RA/Dec reach is taken at z_min = z (clamped), then the z window */
    z_min = z;
    if (z_min < 0.001) z_min = 0.001;
    if (z_min > P->z_max) return 0;
    if (P->coord >= 4) {
      da_max = P->rmax / R2D;
    } else {
      if (!tab_ok(P, z_min)) return 0;
      da_max = P->rmax / tab_interp(P->dA, P->dz, z_min);
    }
  } else {
    if (!tab_ok(P, z_max) || !tab_ok(P, z)) return 0;
    if (P->coord == 1) { /* case 1 */
      z_max = z + P->rmax * tab_interp(P->H, P->dz, z_max) / 300000.0;
      z_min = z * 2 - z_max;
    } else {
      z_max = z + P->pmax * tab_interp(P->H, P->dz, z_max) / 300000.0;
      z_min = z + P->pmin * tab_interp(P->H, P->dz, z) / 300000.0;
    }
    if (z_min < 0.001) z_min = 0.001;
    if (z_min > P->z_max) return 0; /* the reference calls exit(1) */
    if (!tab_ok(P, z_min)) return 0;
    da_max = P->rmax / tab_interp(P->dA, P->dz, z_min);
  }
  double dec_rad = dec / R2D;
  double sd = sin(dec_rad), cd = cos(dec_rad);
  double ra_lim = acos((cos(da_max) - sd * sd) / (cd * cd)) * 1.02 * R2D;
  double ra_max = ra + ra_lim, ra_min = ra - ra_lim;
  if (ra_max > 360) ra_min = 0;
  if (ra_min < 0) ra_max = 360;
  if (P->coord >= 2) {
    z_min = P->z_sep_min == 0 ? P->z_min : z + P->z_sep_min;
    z_max = P->z_sep_max == 0 ? P->z_max : z + P->z_sep_max;
  }
  if (P->coord >= 4) {
    z_min = -INFINITY;
    z_max = INFINITY;
  }
  w[0] = z_min;
  w[1] = z_max;
  w[2] = ra_min;
  w[3] = ra_max;
  w[4] = dec - da_max * R2D * 1.05;
  w[5] = dec + da_max * R2D * 1.05;
  return 1;
}

static double now_ms(void) {
  struct timeval tv;
  gettimeofday(&tv, NULL);
  return tv.tv_sec * 1e3 + tv.tv_usec * 1e-3;
}

/* Validate and derive the data_info quantities (initialization.cpp:305-311). */
static int make_params(const cpc_params *p, oparm *P) {
  if (p->geometry != CPC_GEOM_BOX && p->geometry != CPC_GEOM_SKY) return CPC_ERR_UNSUPPORTED;
  if (p->n_rp <= 0 || p->n_pi <= 0 || !(p->rp_max > 0) || p->rp_min > p->rp_max) return CPC_ERR_INVALID;
  if (p->do_jk != 0 && p->do_jk != 1) return CPC_ERR_UNSUPPORTED;
  P->geom = p->geometry;
  P->L = p->box_size;
  P->pmin = p->pi_min;
  P->pmax = p->pi_max;
  P->psize = (p->pi_max - p->pi_min) / (p->n_pi);
  P->rmin = p->rp_min;
  P->rmax = p->rp_max;
  if (p->lin_bin)
    P->rsize = (p->rp_max - p->rp_min) / p->n_rp;
  else
    P->rsize = (log10(p->rp_max) - log10(p->rp_min)) / p->n_rp;
  P->n_rp = p->n_rp;
  P->n_pi = p->n_pi;
  P->lin = p->lin_bin;
  P->kind = p->kind;
  P->gws = p->g_with_shape;
  P->est = p->estimator;
  P->which_corr = p->which_corr;
  P->jk_on = (p->do_jk == 1 && p->n_jk > 0);
  P->inc_max = p->include_prob_max;
  P->H = p->sky_H;
  P->dA = p->sky_dA;
  P->ntab = p->sky_n;
  P->dz = p->sky_dz;
  P->z_max = p->z_max;
  P->z_sep_max = p->z_sep_max;
  P->z_min = p->z_min;
  P->z_sep_min = p->z_sep_min;
  P->autop = p->auto_pairs != 0;
  P->unsorted = p->unsorted != 0;
  P->rmu = p->r_mu != 0;
  P->coord = p->coordinates;
  if (P->geom == CPC_GEOM_SKY && (!P->H || !P->dA || P->ntab < 2 || !(P->dz > 0))) return CPC_ERR_INVALID;
  if (P->geom == CPC_GEOM_SKY && P->rmu) return CPC_ERR_INVALID; /* the sky's r-mu is coordinates 1 */
  if (P->geom == CPC_GEOM_BOX && P->coord != 0) return CPC_ERR_INVALID;
  if (P->coord < 0 || P->coord > 5) return CPC_ERR_UNSUPPORTED;
  /* choose_rp_pi_func (initialization.cpp:545-571), data_ini's p_size (initialization.cpp:305-308) */
  const int proj = P->pmax > 600 || P->coord == 3 || P->coord == 5;
  P->rpm = (P->coord == 4 || P->coord == 5) ? 2 : (proj ? 1 : 0);
  P->pim = (P->coord == 2 || P->coord == 4) ? 2 : (proj ? 1 : 0);
  if (P->coord == 1) P->rmu = 1;
  if (P->coord == 3 || P->coord == 5) P->psize = 1;
  return CPC_OK;
}

/*
 * orc_count: the cross pass of corels2 for outer = g1, inner = g2 (after the swap).
 * Writes `out` in cpc_hist layout; returns a cpc_status.
 */
int orc_count(const cpc_catalog_desc *g1, const cpc_catalog_desc *g2, const cpc_params *params, cpc_hist *out,
              cpc_stats *st, int n_threads) {
  oparm P;
  int rc = make_params(params, &P);
  if (rc != CPC_OK) return rc;
  const int R = P.jk_on ? 1 + params->n_jk : 1;
  const size_t nc = (size_t)P.n_rp * P.n_pi;
  const int64_t n1 = g1->n, n2 = g2->n;
  if (P.jk_on) {
    if ((n1 && !g1->jk) || (n2 && !g2->jk)) return CPC_ERR_INVALID;
    for (int64_t i = 0; i < n1; i++)
      if (g1->jk[i] < 0 || g1->jk[i] >= params->n_jk) return CPC_ERR_INVALID;
    for (int64_t j = 0; j < n2; j++)
      if (g2->jk[j] < 0 || g2->jk[j] >= params->n_jk) return CPC_ERR_INVALID;
  }
  const int sky = P.geom == CPC_GEOM_SKY;
  if (P.autop && (g1->n != g2->n || g1->z != g2->z)) return CPC_ERR_INVALID; /* one catalog */
  if (sky && ((n1 && (!g1->dc || !g1->da)) || (n2 && (!g2->dc || !g2->da)))) return CPC_ERR_INVALID;
  /* the shape kinds read e[] of the shape side (both sides for shape_shape); the reference would
   * dereference an unallocated pointer here (corels.cpp:216-219, read_dat.cpp:126) */
  if (P.kind == CPC_KIND_SHAPE_DENSITY) {
    const cpc_catalog_desc *s = P.gws == 2 ? g2 : g1;
    if (s->n && (!s->e1 || !s->e2)) return CPC_ERR_INVALID;
  }
  if (P.kind == CPC_KIND_SHAPE_SHAPE && ((n1 && !g1->e1) || (n2 && !g2->e1))) return CPC_ERR_INVALID;
  if (P.kind == CPC_KIND_ED) {
    const cpc_catalog_desc *ws = P.gws == 2 ? g2 : g1;
    if (sky) return CPC_ERR_UNSUPPORTED; /* ED_calc_PB: the periodic box */
    if (ws->n && (!ws->e1 || !ws->e2 || !ws->e3)) return CPC_ERR_INVALID;
  }
  if ((P.kind == CPC_KIND_DENSITY_CONVERGENCE || P.kind == CPC_KIND_SHAPE_CONVERGENCE) && P.gws != 0) {
    const cpc_catalog_desc *ws = P.gws == 2 ? g2 : g1, *ot = P.gws == 2 ? g1 : g2;
    const cpc_catalog_desc *kap = P.kind == CPC_KIND_DENSITY_CONVERGENCE ? ws : ot;
    if (kap->n && !kap->e1) return CPC_ERR_INVALID;
    if (P.kind == CPC_KIND_SHAPE_CONVERGENCE && ws->n && (!ws->e1 || !ws->e2)) return CPC_ERR_INVALID;
  }
  if (n_threads <= 0) n_threads = omp_get_num_procs();
  double t0 = now_ms();

  ohist *th = calloc(n_threads, sizeof(ohist));
  int64_t *tbinned = calloc(n_threads, sizeof(int64_t));
  int64_t *tcand = calloc(n_threads, sizeof(int64_t));
  int64_t *tsparse = calloc(n_threads, sizeof(int64_t));
  int *tbad = calloc(n_threads, sizeof(int));
  int ok = th && tbinned && tcand && tsparse && tbad;
  for (int t = 0; ok && t < n_threads; t++) ok = hist_alloc(&th[t], P.n_rp, P.n_pi, R);
  if (!ok) return CPC_ERR_NOMEM;

  const double L = P.L;
  long chunk = (long)floor((double)n1 / (n_threads * 40)); /* schedule of corr.cpp:331 */
  if (chunk < 1) chunk = 1;

  if (n1 > 0 && n2 > 0) {
#pragma omp parallel num_threads(n_threads)
    {
      int t = omp_get_thread_num();
      ohist *h = &th[t];
      int64_t binned = 0, cand = 0, sparse = 0;
#pragma omp for schedule(dynamic, chunk)
      for (int64_t i = 0; i < n1; i++) {
        if (cinc(g1, i) > P.inc_max) continue;
        if (sky) {
          double w[6];
          if (!sky_window(&P, g1->x[i], g1->y[i], g1->z[i], w)) {
            tbad[t] = 1;
            continue;
          }
          int64_t j;
          if (P.autop) {
            j = i + 1; /* corels1: scan from j = i+1 (corr.cpp:168) */
          } else if (P.unsorted) {
            j = 0; /* do_pair_unsorted scans every inner galaxy (corr.cpp:36-52) */
          } else {
            j = lower_bound_(g2->z, n2, 0.0, w[0]);
            if (j < n2 && g2->z[j] > w[1]) sparse++;
          }
          for (; j < n2; j++) {
            if (cinc(g2, j) > P.inc_max) continue;
            double zz = g2->z[j];
            if (zz > w[1] || zz < w[0]) {
              if (P.unsorted) continue;
              break;
            }
            if (g2->y[j] > w[5] || g2->y[j] < w[4]) continue;
            if (g2->x[j] < w[2] || g2->x[j] > w[3]) continue;
            cand++;
            binned += corel_sky(&P, g1, i, g2, j, h);
          }
          continue;
        }
        /* lim_calc_xyz + lim_calc_PB, PB_calcs.cpp:12-61: z +- rp_max for r-mu (coordinates 7) */
        double z_max = g1->z[i] + (P.rmu ? P.rmax : P.pmax), z_min = g1->z[i] + (P.rmu ? -P.rmax : P.pmin);
        double ra_max = g1->x[i] + P.rmax, ra_min = g1->x[i] - P.rmax;
        double dec_max = g1->y[i] + P.rmax, dec_min = g1->y[i] - P.rmax;
        int PBx = 0, PBy = 0, PBz = 0;
        if (z_min < 0) PBz = -1;
        if (z_max > L) PBz = abs(PBz) + 1;
        if (ra_min < 0) PBx = -1;
        if (ra_max > L) PBx = abs(PBx) + 1;
        if (dec_min < 0) PBy = -1;
        if (dec_max > L) PBy = abs(PBy) + 1;
        int xt = PBx, yt = PBy, zt = PBz;
        /* image loops, corr.cpp:380-423 */
        int cy = yt > 0;
        for (int yi = 0; yi <= abs(yt); yi++, cy--) {
          int cx = xt > 0;
          for (int xi = 0; xi <= abs(xt); xi++, cx--) {
            int cz = zt > 0;
            for (int zi = 0; zi <= abs(zt); zi++, cz--) {
              const double sz = L * cz, sy = L * cy, sx = L * cx;
              int64_t j;
              if (P.unsorted) {
                j = P.autop ? i + 1 : 0; /* corels1 / corels2 with data_sorted 0 (corr.cpp:167-182, 380-423) */
              } else if (cz > 0) {
                j = 0;
              } else if (P.autop) {
                j = i + 1; /* corels1: PBz <= 0 scans from j = i+1 (corr.cpp:168-176) */
              } else {
                j = lower_bound_(g2->z, n2, sz, z_min);
                if (j < n2 && g2->z[j] + sz > z_max) sparse++;
              }
              for (; j < n2; j++) {
                if (cinc(g2, j) > P.inc_max) continue;
                double zz = g2->z[j] + sz;
                if (zz > z_max || zz < z_min) {
                  if (P.unsorted) continue; /* do_pair_unsorted, corr.cpp:38-43 */
                  break;                    /* do_pair_sorted, corr.cpp:71-73 */
                }
                double yy = g2->y[j] + sy;
                if (yy > dec_max || yy < dec_min) continue;
                double xx = g2->x[j] + sx;
                if (xx < ra_min || xx > ra_max) continue;
                cand++;
                binned += corel_box(&P, g1, i, g2, j, cx, cy, cz, h);
              }
            }
          }
        }
      }
      tbinned[t] = binned;
      tcand[t] = cand;
      tsparse[t] = sparse;
    }
  }

  int bad = 0;
  for (int t = 0; t < n_threads; t++) bad |= tbad[t];
  /* merge in thread order (comb_thread_bins, corr.cpp:88-100) */
  ohist acc;
  if (!hist_alloc(&acc, P.n_rp, P.n_pi, R)) return CPC_ERR_NOMEM;
  int64_t binned = 0, cand = 0, sparse = 0;
  for (int t = 0; t < n_threads; t++) {
    for (size_t k = 0; k < (size_t)R * nc; k++) {
      acc.num[k] += th[t].num[k];
      acc.wt[k] += th[t].wt[k];
      for (int f = 0; f < NF; f++) {
        acc.val[k * NF + f] += th[t].val[k * NF + f];
        acc.err[k * NF + f] += th[t].err[k * NF + f];
      }
    }
    for (size_t k = 0; k < (size_t)R * P.n_rp; k++) acc.num_rp[k] += th[t].num_rp[k];
    binned += tbinned[t];
    cand += tcand[t];
    sparse += tsparse[t];
    hist_free(&th[t]);
  }
  free(th);
  free(tbinned);
  free(tcand);
  free(tsparse);
  free(tbad);
  if (bad) {
    hist_free(&acc);
    return CPC_ERR_INVALID;
  }

  /* final_jk_bins (jk.cpp:74-81): jk_r = full - touch_r over every field, add_bins(sub=true) */
  if (P.jk_on && params->jk_mode == CPC_JK_LEAVE_ONE_OUT) {
    for (int r = 1; r < R; r++) {
      for (size_t k = 0; k < nc; k++) {
        size_t a = (size_t)r * nc + k;
        acc.num[a] = acc.num[k] + acc.num[a] * -1;
        acc.wt[a] = acc.wt[k] + acc.wt[a] * -1;
        for (int f = 0; f < NF; f++) {
          acc.val[a * NF + f] = acc.val[k * NF + f] + acc.val[a * NF + f] * -1;
          acc.err[a * NF + f] = acc.err[k * NF + f] + acc.err[a * NF + f] * -1;
        }
      }
      for (int k = 0; k < P.n_rp; k++)
        acc.num_rp[(size_t)r * P.n_rp + k] = acc.num_rp[k] + acc.num_rp[(size_t)r * P.n_rp + k] * -1;
    }
  }
  memcpy(out->num, acc.num, (size_t)R * nc * sizeof(int64_t));
  memcpy(out->wt_num, acc.wt, (size_t)R * nc * sizeof(double));
  memcpy(out->val, acc.val, (size_t)R * nc * NF * sizeof(double));
  memcpy(out->err, acc.err, (size_t)R * nc * NF * sizeof(double));
  memcpy(out->num_rp, acc.num_rp, (size_t)R * P.n_rp * sizeof(int64_t));
  hist_free(&acc);
  if (st) {
    memset(st, 0, sizeof(*st));
    st->n_binned = binned;
    st->n_candidates = cand;
    st->n_sparse_windows = sparse;
    st->ms_total = now_ms() - t0;
  }
  return CPC_OK;
}

/* The reference's rp bin index for a squared separation s (corels.cpp:31-34 after rp_calc_PB's sqrt),
 * as a double so that out-of-range values stay comparable.  Used by tests to pin the threshold table. */
double orc_rp_bin(const cpc_params *p, double s) {
  oparm P;
  if (make_params(p, &P) != CPC_OK) return NAN;
  double rp = sqrt(s);
  if (P.lin) return floor((rp - P.rmin) / P.rsize);
  return floor(log10(rp / P.rmin) / P.rsize);
}

/* The reference's pi bin index (corels.cpp:26). */
double orc_pi_bin(const cpc_params *p, double d_P) {
  oparm P;
  if (make_params(p, &P) != CPC_OK) return NAN;
  return floor((d_P - P.pmin) / P.psize);
}

/* The sky window of one outer galaxy (lim_calc_sky): w = z_min, z_max, ra_min, ra_max, dec_min,
 * dec_max.  Returns 0 where the reference would exit. */
int orc_sky_window(const cpc_params *p, double ra, double dec, double z, double *w) {
  oparm P;
  if (make_params(p, &P) != CPC_OK) return 0;
  return sky_window(&P, ra, dec, z, w);
}
