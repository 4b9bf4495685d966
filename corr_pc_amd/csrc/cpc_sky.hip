// cpc_sky.hip -- pair kernel of the sky geometry (RA/Dec/redshift, rp-pi, corr_pc coordinates 0).
//
// Same decomposition as the box kernel: one workgroup per run of block galaxies in one
// (RA/Dec cell, region, coin) group, a private (rp, pi) histogram in LDS, lanes = 64 block galaxies
// contiguous in redshift, tile galaxies streamed with wave-uniform loads.  Differences:
//   * each outer galaxy carries the window lim_calc_sky computed on the host (sky_calcs.cpp:105-190);
//     the candidate test compares against it exactly as do_pair_sorted does (corr.cpp:67-86);
//   * pi = DC_j - DC_i (calcs.cpp:126-132) is exact; rp = acos(sum) (DA_i + DA_j)/2
//     (sky_calcs.cpp:24-31, 91-96) uses sin/cos of Dec from the host and the device cos/acos, so a
//     pair whose rp lies within a proven bound of a bin edge (or whose sum lies near the 0.9999999999
//     clamp) is handed to the host, which bins it with the reference's libm;
//   * jackknife: the outer galaxy's region is touched iff its jk_prob >= 0.5 whatever the inner region
//     (jk_check_sky, jk.cpp:53-57), the inner galaxy's iff its jk_prob >= 0.5 and the regions differ
//     (corel_jk, jk.cpp:201-223).  As in the box kernel, each item's tile filter (tmode, treg) takes
//     only tile groups whose pairs share one set of touched regions, and the item counts into that
//     set's target; a reduction adds the targets into the full sample and the touch histograms.

#include <hip/hip_runtime.h>

#include "cpc_device.h"
#include "cpc_internal.h"

namespace cpc {

__device__ __forceinline__ double wave_min(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ double wave_max(double v) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
  return v;
}

// cos(x) for |x| <= pi/2 by its Taylor series to x^18 (truncation below 4e-15 there); feeds
// only the orientation angle of a binned pair, whose weighted sums have a 1e-10 tolerance.
__device__ __forceinline__ double cos_poly(double x) {
  const double t = x * x;
  double p = fma(t, -1.0 / 6402373705728000.0, 1.0 / 20922789888000.0);
  p = fma(t, p, -1.0 / 87178291200.0);
  p = fma(t, p, 1.0 / 479001600.0);
  p = fma(t, p, -1.0 / 3628800.0);
  p = fma(t, p, 1.0 / 40320.0);
  p = fma(t, p, -1.0 / 720.0);
  p = fma(t, p, 1.0 / 24.0);
  p = fma(t, p, -0.5);
  return fma(t, p, 1.0);
}

// r2d of sky_calcs.cpp:10
constexpr double kSkyR2D = 180.0 / M_PI;

__device__ __forceinline__ int rp_bin_sky(const SkyArgs &P, const double *thr, double rp, double band,
                                          bool *unc) {
  // bin k iff thr[k] <= rp < thr[k+1]; uncertain if rp +- band straddles an edge
  int k = -1;
  while (k < P.n_rp && rp >= thr[k + 1]) ++k;
  const bool lo_ok = (k < 0) ? true : (rp - band >= thr[k]);
  const bool hi_ok = (k >= P.n_rp) ? true : (rp + band < thr[k + 1]);
  *unc = !(lo_ok && hi_ok);
  return (k >= 0 && k < P.n_rp) ? k : -1;
}

// GEN compiles in the other coordinates' pi and rp (SkyArgs pim, rpm, rmu); without it the kernel is
// that of coordinates 0, whose registers allow two waves per SIMD (three for density pairs).  The
// split density kernel is held to 128 registers for four waves per SIMD: its few spills sit in the
// FP64 path of the uncertain candidates (device cos and acos), and it runs 10% faster than at 168
// registers and three waves; the shape kinds spill inside the tile loop at three waves and run 27%
// slower, so they keep the compiler's choice.
template <int K, bool SPLIT, bool GEN>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu((SPLIT && !GEN && K == kDD) ? 4 : 1)))
sky_kernel(SkyArgs P, SkySide B, SkySide T) {
  const PRec *__restrict__ brec = B.rec;
  const SRec *__restrict__ bsr = B.sr;
  const Win *__restrict__ bwin = B.win;
  const PRec *__restrict__ trec = T.rec;
  const SRec *__restrict__ tsr = T.sr;
  constexpr int NV = KindNV<K>::NV;
  constexpr int NE = KindNV<K>::NE;
  constexpr int NS = 1 + NV + NE;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int next_chunk;  // next chunk of the item that no wave has taken
  // LDS as in the box kernel: ncopy lane-selected copies of the rp rows from row0 up, lnum[ncopy][lcell]
  // and cell-major lacc[ncopy][lcell][NS], then the thresholds; runs of the rows below go to the global
  // histogram.  Compiled in where the flush dominates (7 or more slots per cell).
  constexpr bool CUT = NS >= 7;
  const int ncell = P.ncell, ncopy = CUT ? P.ncopy : 1;
  const int lbase = CUT ? P.row0 * P.n_pi : 0, lcell = CUT ? ncell - lbase : ncell;
  unsigned *lnum = reinterpret_cast<unsigned *>(smem);
  double *lacc = reinterpret_cast<double *>(smem + ((ncopy * lcell * 4 + 15) & ~15));
  double *lthr = lacc + (size_t)ncopy * lcell * NS;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = uni(tid >> 6);
  const int nwave = blockDim.x >> 6;
  const int cp = lane % ncopy;
  lds_u32 *const mynum = (lds_u32 *)(lnum + cp * lcell);
  lds_f64 *const myacc = (lds_f64 *)(lacc + (size_t)cp * lcell * NS);
  for (int k = tid; k <= P.n_rp; k += blockDim.x) lthr[k] = P.thr[k];

  const Item it = P.items[blockIdx.x];
  const Group bg = P.bgrp[it.group0];  // region and coin of every chunk of the item
  const int tgt = it.target, tmode = it.tmode, treg = it.treg;
  unsigned long long *const gnum = P.accnum + (size_t)tgt * ncell;
  double *const gacc = P.acc + (size_t)tgt * NS * ncell;
  unsigned long long ncand = 0;
  for (int c = tid; c < ncopy * lcell; c += blockDim.x) {
    lnum[c] = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) lacc[(size_t)c * NS + s] = 0.0;
  }
  if (tid == 0) next_chunk = nwave;
  __syncthreads();

  // chunks come sorted by decreasing cost; each wave starts on its own and then takes the next
  // unclaimed one, so the waves of the group finish close together
  int next_nx = 0;
  for (int ch = wave; ch < it.n; ch = uni(next_nx)) {
    const Chunk ck = P.chunks[it.begin + ch];
    const int base = uni(ck.first);
    const int cnt = uni(ck.count);
    const bool valid = lane < cnt;
    const int p = base + (valid ? lane : 0);
    const PRec me = brec[p];
    const double mez = B.z[p];
    const SRec ms = bsr[p];
    const Win mw = bwin[p];
    // search rectangle of the wave: the union of its galaxies' windows
    const double zlo = wave_min(valid ? mw.zmin : INFINITY);
    const double zhi = wave_max(valid ? mw.zmax : -INFINITY);
    const double declo = wave_min(valid ? mw.decmin : INFINITY);
    const double dechi = wave_max(valid ? mw.decmax : -INFINITY);
    const bool ra_all = __ballot(valid && !(mw.ramin <= mw.ramax)) != 0;
    const double ralo = wave_min(valid ? mw.ramin : INFINITY);
    const double rahi = wave_max(valid ? mw.ramax : -INFINITY);
    const double eps = 1e-9;
    int cx0 = 0, cx1 = P.nra - 1;
    if (!ra_all) {
      cx0 = uni((int)fmax(0.0, fmin((double)(P.nra - 1), floor((ralo - eps - P.ra0) / P.hra))));
      cx1 = uni((int)fmax(0.0, fmin((double)(P.nra - 1), floor((rahi + eps - P.ra0) / P.hra))));
    }
    const int cy0 = uni((int)fmax(0.0, fmin((double)(P.ndec - 1), floor((declo - eps - P.dec0) / P.hdec))));
    const int cy1 = uni((int)fmax(0.0, fmin((double)(P.ndec - 1), floor((dechi + eps - P.dec0) / P.hdec))));
    const double zsl = zlo - 1e-12, zsh = zhi + 1e-12;

    // (the sums share a struct with the count: an array of doubles alone would be promoted to one
    // vector register tuple, which every join of the loop's control flow copies whole)
    int run_cell = -1;
    struct {
      double v[NS];
      unsigned n;
    } R;
    R.n = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) R.v[s] = 0.0;
    double *const run = R.v;
    unsigned &run_n = R.n;

    for (int cy = cy0; cy <= cy1; ++cy) {
      for (int cx = cx0; cx <= cx1; ++cx) {
        const int col = cy * P.nra + cx;
        const int g0 = P.tcol[col], g1 = P.tcol[col + 1];
        for (int g = g0; g < g1; ++g) {
          const Group tg = P.tgrp[g];
          // the item's tile filter: every pair it meets belongs to its target
          const bool same = tg.region == bg.region;
          const bool take = tmode == kTileAll || (tmode == kTileSame && same) ||
                            (tmode == kTileOtherC0 && !same && tg.coin == 0) ||
                            (tmode == kTileSameOrC0 && (same || tg.coin == 0)) ||
                            (tmode == kTileRegionC1 && tg.region == treg && tg.coin == 1);
          if (!take) continue;
          if (zsh < tg.zlo || zsl > tg.zhi) continue;
          const int j0 = zsearch<false>(T.z, tg.start, tg.end, zsl);
          const int j1 = zsearch<true>(T.z, j0, tg.end, zsh);
          for (int j = j0; j < j1; ++j) {
            const PRec t = trec[j];
            const double tzj = T.z[j];
            const SRec ts = tsr[j];
            // outer (o: block lane) and inner (i: tile) roles of the pair
            const PRec &po = me;
            const PRec &pi = t;
            const SRec &so = ms;
            const SRec &si = ts;
            const Win &w = mw;
            // do_pair_sorted's tests against the outer galaxy's window (PBz = PBy = PBx = 0)
            const double piz = tzj;
            bool ok = valid && !(piz > w.zmax || piz < w.zmin) && !(pi.y > w.decmax || pi.y < w.decmin) &&
                      !(pi.x < w.ramin || pi.x > w.ramax);
            if (P.auto_mode) {
              // corels1 pairs: the inner galaxy after the outer one in catalog order, which is (z,
              // index) order for z-sorted input (1) and the index alone for unsorted input (2)
              const int io = B.idx[p], ii = T.idx[j];
              ok = ok && (P.auto_mode == 2 ? ii > io : (piz > mez || (piz == mez && ii > io)));
            }
            ncand += ok;
            int cell = -1;
            if (ok) {
              // pi_calc or pi_calc_projected, corel_bin; ephi (pi_calc_phi) and r-mu bin pi late
              double dP = (!GEN || P.pim == 0) ? si.dc - so.dc : 0.0;
              double q2;
              if constexpr (SPLIT && !GEN) {
                // the pi bin from a product with 1/psize: it differs from the IEEE quotient by about
                // one ulp, so where its fraction lies at least pguard from both edges its floor is
                // the reference's; nearer an edge the quotient itself decides
                const double u = (dP - P.pmin) * P.inv_psize;
                q2 = floor(u);
                const double fr = u - q2;
                if (!(fr >= P.pguard && fr <= 1.0 - P.pguard && fabs(u) < 1048576.0))
                  q2 = floor((dP - P.pmin) / P.psize);
              } else {
                q2 = floor((dP - P.pmin) / P.psize);
              }
              const bool late = GEN && (P.pim == 2 || P.rmu);
              bool exact = true;
              if (SPLIT && P.split_on && q2 >= 0.0 && q2 < (double)P.n_pi) {
                // rp^2 from the chord between the unit vectors: theta^2 = 4 asin^2(c/2), with
                // asin^2(x) = x^2 (1 + x^2/3 + 8x^4/45 + 4x^6/35 + 128x^8/1575 + 128x^10/2079 + ...),
                // truncated below 5e-14 relative for c < 0.2 (x^2 < 0.01)
                const double ex = so.ux - si.ux, ey = so.uy - si.uy, ez = so.sd - si.sd;
                const double c2 = fma(ex, ex, fma(ey, ey, ez * ez));
                const double x2 = 0.25 * c2;
                double poly = fma(x2, 128.0 / 2079.0, 128.0 / 1575.0);
                poly = fma(x2, poly, 4.0 / 35.0);
                poly = fma(x2, poly, 8.0 / 45.0);
                poly = fma(x2, poly, 1.0 / 3.0);
                poly = fma(x2, poly, 1.0);
                const double th2 = c2 * poly;
                const double dbar = (so.da + si.da) * 0.5;
                const double s64 = th2 * dbar * dbar;
                // below theta^2 = 2.2e-10 the reference's clamp (sum > 0.9999999999) may apply
                if (th2 > 2.2e-10 && x2 < 0.01) {
                  const float s32 = (float)s64;
                  if (s32 < P.s_out_lo || s32 >= P.s_out_hi) {
                    exact = false;
                  } else {
                    const float kf = __log2f(s32) * P.la + P.lb;
                    const float fk = floorf(kf);
                    const float fr = kf - fk;
                    if (fr >= P.g && fr <= 1.0f - P.g) {
                      exact = false;
                      if (fk >= 0.0f && fk < (float)P.n_rp) cell = (int)fk * P.n_pi + (int)q2;
                    }
                  }
                }
              }
              if (exact && (late || (q2 >= 0.0 && q2 < (double)P.n_pi))) {
                // dA_calc: sin d1 sin d2 + cos d1 cos d2 cos(a1 - a2), clamp, acos; then rp_calc_sky,
                // rp_calc_sky_projected (outer DA) or rp_calc_sky_angular (degrees)
                const double sum = so.sd * si.sd + so.cd * si.cd * cos((so.ra_rad - si.ra_rad));
                const bool clamped = sum > 0.9999999999;
                const double dA = clamped ? acos(1.0) : acos(sum);
                const int rpm = GEN ? P.rpm : 0;
                const double dsc = rpm == 0 ? 0.5 * (so.da + si.da) : (rpm == 1 ? so.da : kSkyR2D);
                const double rp = rpm == 0 ? dA * (so.da + si.da) / 2. : dA * dsc;
                // the host's dA moves by dsum / sqrt(1 - sum^2) from the sum's difference, and by the
                // acos's own rounding; a clamped sum (more than dsum above the clamp, checked below)
                // is clamped on the host too and gives dA = acos(1) there as well, exactly
                const double one_m = fmax(1.0 - sum * sum, 1e-300);
                const double dda = clamped ? 0.0 : P.dsum / sqrt(one_m) + 1e-15 * dA;
                double band = 1.5 * dsc * dda + 4e-16 * rp;
                double r = rp;
                bool unc = false;
                if (late) {
                  double pb = 0.0;  // bound on the device/host difference of the binned pi
                  if (P.pim == 2) {
                    // ephi_calc of the g_with_shape galaxy against angle_orient_axis_sky
                    const double th = atan((pi.y - po.y) / ((pi.x - po.x) * cos((so.dec_rad + si.dec_rad) / 2.0)));
                    double t = (P.gws == 2 ? pi.phi : po.phi) - th;
                    t = t < 0 ? -1 * t : t;
                    if (t > M_PI / 2.0) t = M_PI - t;
                    dP = t;
                    pb = P.dphi;
                  }
                  if (P.rmu) {
                    // r_mu_calc: r = sqrt(rp^2 + pi^2) moves by at most rp's band, mu = pi / r by
                    // |pi| band / r^2 plus its rounding
                    r = sqrt(rp * rp + dP * dP);
                    band += 4e-16 * r;
                    const double mu = dP / r;
                    pb = fabs(dP) * band / (r * r) + 4e-16;
                    dP = mu;
                  }
                  q2 = floor((dP - P.pmin) / P.psize);
                  unc = floor((dP - pb - P.pmin) / P.psize) != floor((dP + pb - P.pmin) / P.psize);
                }
                bool unc_r;
                const int k = rp_bin_sky(P, lthr, r, band, &unc_r);
                unc = unc || unc_r || fabs(sum - 0.9999999999) <= P.dsum;
                if (unc) {
                  const unsigned long long slot = atomicAdd(&P.counters[3], 1ull);
                  if (slot < (unsigned long long)P.hq_cap) P.hq[slot] = make_uint2(B.idx[p], T.idx[j]);
                } else if (k >= 0 && q2 >= 0.0 && q2 < (double)P.n_pi) {
                  cell = k * P.n_pi + (int)q2;
                }
              }
            }
            if (cell >= 0) {
              const Gal go{po.x, po.y, po.w, po.e1, po.e2, po.phi};
              const Gal gi{pi.x, pi.y, pi.w, pi.e1, pi.e2, pi.phi};
              // angle_orient_axis_sky: atan((dec_i - dec_o) / ((ra_i - ra_o) cos((dec_o + dec_i)/2)))
              const double X = (pi.x - po.x) * (SPLIT ? cos_poly((so.dec_rad + si.dec_rad) / 2.0)
                                                      : cos((so.dec_rad + si.dec_rad) / 2.0));
              const double Y = pi.y - po.y;
              double qv[NS];
              pair_values_v<K, SPLIT>(go, gi, P.gws, X, Y, qv);
              if constexpr (NS == 1) {
                // one slot (density pairs): the run restarts through selects, and only the flush
                // branches (no copies of the run at the join of two branches)
                const bool nw = cell != run_cell;
                if (nw && run_cell >= 0) flush_cm<NS, CUT>(mynum, myacc, lbase, gnum, gacc, ncell, run_cell, run_n, run);
                run[0] = (nw ? 0.0 : run[0]) + qv[0];
                run_n = (nw ? 0u : run_n) + 1u;
                run_cell = cell;
              } else if (cell != run_cell) {
                if (run_cell >= 0) flush_cm<NS, CUT>(mynum, myacc, lbase, gnum, gacc, ncell, run_cell, run_n, run);
                // the new run starts from this pair's values
#pragma unroll
                for (int s = 0; s < NS; ++s) run[s] = qv[s];
                run_n = 1;
                run_cell = cell;
              } else {
#pragma unroll
                for (int s = 0; s < NS; ++s) run[s] += qv[s];
                ++run_n;
              }
            }
          }
        }
      }
    }
    if (run_cell >= 0) flush_cm<NS, CUT>(mynum, myacc, lbase, gnum, gacc, ncell, run_cell, run_n, run);
    int nx = 0;
    if (lane == 0) nx = atomicAdd(&next_chunk, 1);
    next_nx = __builtin_amdgcn_readfirstlane(nx);
  }
  __syncthreads();
  for (int c = tid; c < lcell; c += blockDim.x) {
    unsigned n = 0;
    for (int k = 0; k < ncopy; ++k) n += lnum[k * lcell + c];
    if (n) {
      atomicAdd(&gnum[lbase + c], (unsigned long long)n);
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        double v = 0.0;
        for (int k = 0; k < ncopy; ++k) v += lacc[((size_t)k * lcell + c) * NS + s];
        unsafeAtomicAdd(&gacc[(size_t)s * ncell + lbase + c], v);
      }
    }
  }
  if (ncand) atomicAdd(&P.counters[0], ncand);
}

__global__ void k_gather_sky(const int32_t *__restrict__ sidx, int64_t n, const SRec *__restrict__ src_sr,
                             const Win *__restrict__ src_win, SRec *__restrict__ sr, Win *__restrict__ win) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int32_t i = sidx[p];
  sr[p] = src_sr[i];
  if (src_win) win[p] = src_win[i];
}

// The device cos (fn 0) or acos (fn 1) of the sky kernels, for checks against the host C library
__global__ void k_libm(int fn, const double *__restrict__ x, int64_t n, double *__restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  y[i] = fn == 0 ? cos(x[i]) : acos(x[i]);
}

hipError_t launch_libm(int fn, const double *x, int64_t n, double *y, hipStream_t st) {
  if (n == 0) return hipSuccess;
  k_libm<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(fn, x, n, y);
  return hipGetLastError();
}

hipError_t launch_gather_sky(const int32_t *sidx, int64_t n, const SRec *src_sr, const Win *src_win, SRec *sr,
                             Win *win, hipStream_t st) {
  if (n == 0) return hipSuccess;
  k_gather_sky<<<(unsigned)((n + 255) / 256), 256, 0, st>>>(sidx, n, src_sr, src_win, sr, win);
  return hipGetLastError();
}

template <int K, bool SPLIT, bool GEN = false>
static hipError_t launch_sky_t(const SkyArgs &P, int threads, size_t lds, const SkySide &B, const SkySide &T,
                               hipStream_t st) {
  auto fn = sky_kernel<K, SPLIT, GEN>;
  hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  fn<<<P.n_items, threads, lds, st>>>(P, B, T);
  return hipGetLastError();
}

hipError_t launch_sky(int kind, bool split, const SkyArgs &P, int threads, const SkySide &B, const SkySide &T,
                      hipStream_t st) {
  if (P.n_items == 0) return hipSuccess;
  const size_t lds = pair_lds_bytes(P.ncell - P.row0 * P.n_pi, P.nslots, P.n_rp, P.ncopy);
  // the other coordinates classify every candidate in FP64 (split_on is 0 there): one kernel for both
  const bool gen = P.pim != 0 || P.rpm != 0 || P.rmu != 0;
#define CPC_SKY(KK)                                                            \
  if (kind == KK) {                                                            \
    if (gen) return launch_sky_t<KK, false, true>(P, threads, lds, B, T, st);  \
    if (split) return launch_sky_t<KK, true>(P, threads, lds, B, T, st);       \
    return launch_sky_t<KK, false>(P, threads, lds, B, T, st);                 \
  }
  CPC_SKY(kDD)
  CPC_SKY(kSD)
  CPC_SKY(kSS0)
  CPC_SKY(kSS1)
  CPC_SKY(kDK)
  CPC_SKY(kSK)
#undef CPC_SKY
  return hipErrorInvalidValue;
}

}  // namespace cpc
