// cpc_device.h -- device helpers shared by the box (cpc_kernels.hip) and sky (cpc_sky.hip) pair
// kernels: pair kinds, binning against host-calibrated thresholds, image flags, z-range search.
#pragma once

#include <hip/hip_runtime.h>

#include "cpc_internal.h"

namespace cpc {

enum { kDD = 0, kSD = 1, kSS0 = 2, kSS1 = 3, kDK = 4, kSK = 5, kED = 6 };

// Per pair kind: NV data[] values accumulated, and NE of them with their err sums (the first NE).
template <int K>
struct KindNV;
template <>
struct KindNV<kDD> { static constexpr int NV = 0, NE = 0; };
template <>
struct KindNV<kSD> { static constexpr int NV = 3, NE = 3; };  // data[1], data[2], data[6]
template <>
struct KindNV<kSS0> { static constexpr int NV = 5, NE = 5; };  // data[1], data[3..6]
template <>
struct KindNV<kSS1> { static constexpr int NV = 6, NE = 6; };  // data[1..6]
template <>
struct KindNV<kDK> { static constexpr int NV = 1, NE = 0; };  // data[1]; no err (corels.cpp:79-91)
template <>
struct KindNV<kSK> { static constexpr int NV = 3, NE = 0; };  // data[1], data[3], data[5]; no err (:93-120)
template <>
struct KindNV<kED> { static constexpr int NV = 1, NE = 1; };  // data[6] (corels.cpp:246-262)

// ED_calc_PB (PB_calcs.cpp:101-134): cos^2 of the angle between the imaged separation
// (sx, sy, sz) = g1 - (g2 + L PB) and the 3-D axis (e1, e2, e3) of g2 if g_with_shape is 2, else of g1;
// the axis travels with e3 in the record's phi slot.  q = [wt, cos^2, cos^4].
__device__ __forceinline__ void ed_values(double w1, double w2, double ax, double ay, double az, double sx,
                                          double sy, double sz, double *q) {
  const double wt = w1 * w2;
  const double norm = sqrt(sx * sx + sy * sy + sz * sz);
  double c = ax * sx + ay * sy + az * sz;
  c /= norm;
  const double t = c * c;
  q[0] = wt;
  q[1] = t;
  q[2] = t * t;
}

// 1/x from the hardware reciprocal estimate and one Newton step: the relative error squares, to
// below 1e-13 even from a 2^-22 estimate (x finite, != 0).  Used only where the result feeds a
// weighted sum, whose tolerance is 1e-10.
__device__ __forceinline__ double rcp_refined(double x) {
  const double r = __builtin_amdgcn_rcp(x);
  return fma(r, fma(-x, r, 1.0), r);
}

struct Gal {
  double x, y, w, e1, e2, phi;
};

// atan(y/x) to within 1e-14 absolute, in registers only and without branches.  With
// a = min(|x|,|y|), b = max, the ratio r = a/b is reduced about the nearest quarter c = k/4 (k from an
// FP32 estimate):
//   atan(r) = atan(c) + atan(d),  d = (a - c b)/(b + c a),  |d| <= 0.126,
// atan(d) is its odd Taylor series to d^13 (truncation below |d|^15/15 < 3e-15) and atan(k/4) is the
// quartic in k through the five values atan(0), atan(1/4), .., atan(1) (coefficients solved in
// 50-digit arithmetic; within 6e-17 of each value at k = 0..4).  |y| > |x| uses
// pi/2 - atan(x/y); the sign is that of the IEEE quotient y/x; x = y = 0 gives NaN like atan(0/0).
__device__ __forceinline__ double atan_ratio(double y, double x) {
#pragma clang fp contract(fast)
  const double ax = fabs(x), ay = fabs(y);
  const bool sw = ay > ax;
  const double a = sw ? ax : ay, b = sw ? ay : ax;
  const float rf = (float)a * __builtin_amdgcn_rcpf((float)b);
  const int k = (int)rintf(fminf(fmaxf(rf * 4.0f, 0.0f), 4.0f));
  const double c = (double)k * 0.25;
  const double d = fma(-c, b, a) * rcp_refined(fma(c, a, b));
  const double d2 = d * d;
  double p = fma(d2, 1.0 / 13.0, -1.0 / 11.0);
  p = fma(d2, p, 1.0 / 9.0);
  p = fma(d2, p, -1.0 / 7.0);
  p = fma(d2, p, 1.0 / 5.0);
  p = fma(d2, p, -1.0 / 3.0);
  const double kd = (double)k;
  const double base =
      kd * fma(kd, fma(kd, fma(kd, 0x1.23f4fc5bb1b6dp-11, -0x1.63904aa83c890p-8), -0x1.971b236888072p-11),
               0x1.00a384078b643p-2);
  const double t0 = base + fma(d * d2, p, d);
  const double t = sw ? M_PI_2 - t0 : t0;
  return (__builtin_signbit(x) != __builtin_signbit(y)) ? -t : t;
}

// ephi_calc, calcs.cpp:97-105 (with absolute() of calcs.cpp:50-53)
__device__ __forceinline__ double ephi(double phi, double theta) {
  double t = phi - theta;
  t = t < 0 ? -1 * t : t;
  if (t > M_PI / 2.0) t = M_PI - t;
  return t;
}

// Shape quantities of one binned pair, g1 = outer, g2 = inner (corels.cpp:79-244).
// q = [wt, vals..., errs...] in the slot order of KindNV.  The convergence kinds read the convergence
// kappa where the reference keeps it, in e[0] (e1 here) of the galaxy that carries it.
// (dxr, dyr) is the orientation vector whose atan(dyr/dxr) is the reference's angle: raw coordinate
// differences in the box (angle_orient_axis_PB, calcs.cpp:54-57), the cos(mean Dec)-scaled RA/Dec
// differences on the sky (angle_orient_axis_sky, sky_calcs.cpp:13-22).
template <int K, bool SPLIT>
__device__ __forceinline__ void pair_values_v(const Gal &g1, const Gal &g2, int gws, double dxr, double dyr,
                                              double *q) {
  // weighted sums only (tolerance 1e-10 relative): products and sums may fuse into FMAs
#pragma clang fp contract(fast)
  const double wt = g1.w * g2.w;
  q[0] = wt;
  if constexpr (K == kDD) {
    return;
  } else if constexpr (K == kDK) {
    // density_convergence_corel: w kappa of the g_with_shape galaxy (sig_crit = 1); g_with_shape 0
    // assigns nothing and adds the initial 0
    q[1] = gws == 1 ? wt * g1.e1 : (gws == 2 ? wt * g2.e1 : 0.0);
  } else {
    double theta, c2, s2;
    if constexpr (SPLIT) {
      // theta = atan(dy/dx) by argument reduction; cos 2theta and sin 2theta algebraically (about
      // one ulp each, far inside the weighted-sum tolerance)
      theta = atan_ratio(dyr, dxr);
      const double xx = dxr * dxr, yy = dyr * dyr;
      const double inv = rcp_refined(xx + yy);
      c2 = (xx - yy) * inv;
      s2 = (2.0 * dxr * dyr) * inv;
    } else {
      theta = atan(dyr / dxr);
      c2 = cos(2.0 * theta);
      s2 = sin(2.0 * theta);
    }
    if constexpr (K == kSK) {
      // shape_convergence_corel: rotate the g_with_shape galaxy's shape, kappa from the other one;
      // ep = w kappa, e++ = ep e+, e+x = ep ex; g_with_shape 0 adds the initial zeros
      if (gws == 1 || gws == 2) {
        const Gal &s = (gws == 2) ? g2 : g1;
        const double kap = (gws == 2) ? g1.e1 : g2.e1;
        const double ep = wt * kap;
        q[1] = ep;
        q[2] = ep * (c2 * s.e1 - s2 * s.e2);
        q[3] = ep * (s2 * s.e1 + c2 * s.e2);
      } else {
        q[1] = q[2] = q[3] = 0.0;
      }
      (void)theta;
    } else if constexpr (K == kSD) {
      // shape_density_corel: rotate the shape galaxy, theta column from g1's phi
      const Gal &s = (gws == 2) ? g2 : g1;
      const double ep = c2 * s.e1 - s2 * s.e2;
      const double ex = s2 * s.e1 + c2 * s.e2;
      const double a = wt * ep, b = wt * ex;
      const double t = ephi(g1.phi, theta);
      q[1] = a;
      q[2] = b;
      q[3] = t;
      q[4] = a * a;
      q[5] = b * b;
      q[6] = t * t;
    } else {
      const double epi = c2 * g1.e1 - s2 * g1.e2, exi = s2 * g1.e1 + c2 * g1.e2;
      const double epj = c2 * g2.e1 - s2 * g2.e2, exj = s2 * g2.e1 + c2 * g2.e2;
      // data[5] = w (ex_i e+_j + e+_i ex_j)/2 with err (w ex_i e+_j/2)^2 + (w e+_i ex_j/2)^2
      const double pp = wt * (epi * epj), xx2 = wt * (exi * exj);
      const double hw = 0.5 * wt;
      const double ea = hw * (exi * epj), eb = hw * (epi * exj);
      const double px = ea + eb;
      if constexpr (K == kSS0) {
        // estimator 0: e+ of the shape-side galaxy only, ex never assigned (data[2] += 0)
        const double ep = (gws == 2) ? epj : epi;
        const double a = wt * ep;
        const double t = ephi((gws == 2) ? g2.phi : g1.phi, theta);
        q[1] = a;
        q[2] = pp;
        q[3] = xx2;
        q[4] = px;
        q[5] = t;
        q[6] = a * a;
        q[7] = pp * pp;
        q[8] = xx2 * xx2;
        q[9] = ea * ea + eb * eb;
        q[10] = t * t;
      } else {
        const double ep = wt * (epi + epj) / 2.0;
        const double ex = wt * (exi + exj) / 2.0;
        const double ai = wt * epi / 2.0, aj = wt * epj / 2.0;
        const double bi = wt * exi / 2.0, bj = wt * exj / 2.0;
        const double t1 = ephi(g1.phi, theta), t2 = ephi(g2.phi, theta);
        q[1] = ep;
        q[2] = ex;
        q[3] = pp;
        q[4] = xx2;
        q[5] = px;
        q[6] = t1 + t2;
        q[7] = ai * ai + aj * aj;
        q[8] = bi * bi + bj * bj;
        q[9] = pp * pp;
        q[10] = xx2 * xx2;
        q[11] = ea * ea + eb * eb;
        q[12] = t1 * t1 + t2 * t2;
      }
    }
  }
}

template <int K, bool SPLIT>
__device__ __forceinline__ void pair_values(const Gal &g1, const Gal &g2, int gws, double *q) {
  pair_values_v<K, SPLIT>(g1, g2, gws, g2.x - g1.x, g2.y - g1.y, q);
}

// The KArgs fields the candidate loop of the box kernel reads, copied into registers once per work
// item (the kernel reads KArgs from a table in constant memory).
struct HotArgs {
  double L, pmin, pmax, rmax, psize;
  float gz, pioff32, s_out_lo, s_out_hi, s_cert_lo, linv, lrmin, la, lb, gA, gB, gC, gD, inv_psize32;
  int n_pi, n_rp, lin, gws, auto_mode;
  int lbase, lcell;  // first LDS cell (row0 * n_pi) and LDS cells per copy
  double wlo, whi;   // z window offsets of an outer galaxy
  int rmu;           // r-mu bins (coordinates 7)
  float cull_r2, cull_d;  // tile cull against the wave's bounding box
};

// rp bin of a squared separation against the FP64 thresholds thr[0..n_rp]: bin k iff
// thr[k] <= s < thr[k+1].  An FP32 logarithm gives the starting guess.
template <typename A>
__device__ __forceinline__ int rp_bin64(const A &P, const double *thr, double s) {
  const float sf = (float)s;
  float kf = P.lin ? (sqrtf(sf) - P.lrmin) * P.linv : __log2f(sf) * P.la + P.lb;
  kf = fminf(fmaxf(kf, -1.0f), (float)P.n_rp);
  int k = (int)floorf(kf);
  while (k >= 0 && s < thr[k]) --k;
  while (k < P.n_rp && s >= thr[k + 1]) ++k;
  return (k >= 0 && k < P.n_rp) ? k : -1;
}

// The reference's candidate test and binning for outer (x1,y1,z1), inner (x2,y2,z2) seen through
// image (PBx, PBy, PBz): do_pair_sorted (corr.cpp:67-86) then corel_bin (corels.cpp:21-46), with
// r_mu_calc (calcs.cpp:135-140) for r-mu bins.  Returns the histogram cell n1*n_pi + n2, or -1.
// `A` is HotArgs (arguments held in registers: the FP64 kernel, where this is the hot path) or the
// job's KArgs in constant memory (the split kernel, where only guard-band candidates come here).
template <bool RMU, typename A>
__device__ __forceinline__ int exact_cell(const A &P, const double *thr, double x1, double y1, double z1,
                                          double x2, double y2, double z2, int PBx, int PBy, int PBz) {
  const double L = P.L;
  const double zz = z2 + L * PBz;
  const double zmax = z1 + P.whi, zmin = z1 + P.wlo;
  if (zz > zmax || zz < zmin) return -1;
  const double yy = y2 + L * PBy;
  if (yy > y1 + P.rmax || yy < y1 - P.rmax) return -1;
  const double xx = x2 + L * PBx;
  if (xx < x1 - P.rmax || xx > x1 + P.rmax) return -1;
  const double t0 = x1 - xx, t1 = y1 - yy;
  double d_P = zz - z1, s = t0 * t0 + t1 * t1;
  if constexpr (RMU) {
    // rp rounded first, then r = sqrt(rp*rp + pi*pi); the thresholds bin sqrt of their argument
    const double rp = sqrt(s);
    s = rp * rp + d_P * d_P;
    d_P /= sqrt(s);  // NaN for r = 0: no bin, as the reference's int conversion gives INT_MIN
  }
  const double q2 = floor((d_P - P.pmin) / P.psize);
  if (!(q2 >= 0.0 && q2 < (double)P.n_pi)) return -1;
  const int k = rp_bin64(P, thr, s);
  if (k < 0) return -1;
  return k * P.n_pi + (int)q2;
}

// Whether the outer galaxy with image flags `meta` visits image (PBx, PBy, PBz) (corr.cpp:380-423).
__device__ __forceinline__ bool image_allowed(int meta, int PBx, int PBy, int PBz) {
  bool ok = true;
  ok &= (PBx == 0) || (PBx < 0 ? (meta & kMetaXm) : (meta & kMetaXp));
  ok &= (PBy == 0) || (PBy < 0 ? (meta & kMetaYm) : (meta & kMetaYp));
  ok &= ((PBz == 0 ? kMetaZ0 : PBz < 0 ? kMetaZm : kMetaZp) & meta) != 0;
  return ok;
}

__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// first index in [lo, hi) of the z-sorted records with z >= v (upper=false) or z > v (upper=true)
template <bool UPPER>
__device__ __forceinline__ int zsearch(const double *__restrict__ zs, int lo, int hi, double v) {
  while (lo < hi) {
    const int mid = uni(lo + ((hi - lo) >> 1));
    const double z = zs[mid];
    const bool go_right = UPPER ? (z <= v) : (z < v);
    if (go_right)
      lo = mid + 1;
    else
      hi = mid;
  }
  return lo;
}

using lds_u32 = __attribute__((address_space(3))) unsigned;
using lds_f64 = __attribute__((address_space(3))) double;

// A closed run of one lane into the cell-major LDS histogram copy (lacc[cell][NS]: one address and
// immediate slot offsets), or, for cells below the LDS rows (CUT only), straight into the target's
// global histogram ([NS][ncell]).
template <int NS, bool CUT>
__device__ __forceinline__ void flush_cm(lds_u32 *mynum, lds_f64 *myacc, int lbase, unsigned long long *gnum,
                                         double *gacc, int ncell, int cell, unsigned n, const double *v) {
  const int lc = cell - lbase;
  if (!CUT || lc >= 0) {
    __hip_atomic_fetch_add(&mynum[lc], n, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
    lds_f64 *const a = myacc + lc * NS;
#pragma unroll
    for (int s = 0; s < NS; ++s) __hip_atomic_fetch_add(&a[s], v[s], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  } else {
    atomicAdd(&gnum[cell], (unsigned long long)n);
#pragma unroll
    for (int s = 0; s < NS; ++s) unsafeAtomicAdd(&gacc[(size_t)s * ncell + cell], v[s]);
  }
}

}  // namespace cpc
