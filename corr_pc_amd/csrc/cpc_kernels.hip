// cpc_kernels.hip -- CDNA4 (gfx950) kernels of the rp-pi pair engine.
//
// Work decomposition.  Both catalogs are gathered into 64-byte records ordered by
// (xy column, jackknife region, coin, z).  A work item is a run of up to a few hundred block-side
// galaxies of one (column, region, coin) group; one workgroup owns an item and keeps a private
// (rp bin x pi bin) histogram in LDS.  Each wave holds 64 consecutive block galaxies (one per lane,
// contiguous in z) and walks the neighbour columns; inside a neighbour group it binary-searches the
// z range its lanes can reach and then streams the tile galaxies with wave-uniform (scalar) loads,
// so every tile galaxy is fetched once per wave and evaluated against 64 block galaxies.
//
// Exactness.  The FP64 path evaluates corr_pc's predicates in the reference's operation order
// (do_pair_sorted corr.cpp:67-86, pi_calc_PB/rp_calc_PB PB_calcs.cpp:87-98, corel_bin
// corels.cpp:21-46) with contraction disabled, and bins rp by comparing s = t0*t0 + t1*t1 against
// host-calibrated thresholds (the host libm's log10 decides them), so bin indices are bit-identical.
// The split path classifies in FP32 on column-local coordinates and sends every candidate that lies
// within a proven error bound of a bin edge or window edge to the FP64 path.
//
// Jackknife.  Every pair is accumulated once, into a histogram selected per (item, phase):
//   primary pass (block = outer catalog):
//     target 2a+1: pairs that touch region a through the outer galaxy (same region, or coin >= 0.5)
//     target 2a  : pairs that do not (different region and outer coin < 0.5)
//   touch pass (block = inner galaxies of region b with coin >= 0.5, tiles = outer galaxies of other
//     regions): target b.
// The host rebuilds full = sum of primary targets and touch_r = target 2r+1 + touch-pass target r
// (jk.cpp:201-223), then leave-one-out jk_r = full - touch_r (jk.cpp:74-81).

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "cpc_device.h"
#include "cpc_internal.h"

namespace cpc {

template <int K, bool SPLIT, bool SWAP, int ACC>
__global__ void __launch_bounds__(256) pair_kernel(KArgs P, const PRec *__restrict__ brec,
                                                   const FRec *__restrict__ bfr, const PRec *__restrict__ trec,
                                                   const FRec *__restrict__ tfr) {
  constexpr int NV = KindNV<K>::NV;
  constexpr int NS = SWAP ? 1 + NV : 1 + 2 * NV;  // f64 slots accumulated
  constexpr int NQ = 1 + 2 * NV;                  // quantities computed
  constexpr int QCAP = 128;                       // compaction ring per wave
  extern __shared__ __align__(16) unsigned char smem[];
  const int ncell = P.ncell;
  unsigned *lnum = reinterpret_cast<unsigned *>(smem);
  double *lacc = reinterpret_cast<double *>(smem + ((ncell * 4 + 15) & ~15));
  double *lthr = lacc + (size_t)NS * ncell;
  float *llo = reinterpret_cast<float *>(lthr + P.n_rp + 1);
  float *lhi = llo + P.n_rp + 1;
  unsigned long long *lq =
      reinterpret_cast<unsigned long long *>(reinterpret_cast<uintptr_t>(lhi + P.n_rp + 1 + 1) & ~uintptr_t(7));

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = uni(tid >> 6);
  const int nwave = blockDim.x >> 6;
  unsigned long long *wq = lq + wave * QCAP;

  for (int k = tid; k <= P.n_rp; k += blockDim.x) {
    lthr[k] = P.thr[k];
    if (SPLIT) {
      llo[k] = P.thr_lo[k];
      lhi[k] = P.thr_hi[k];
    }
  }

  // outermost split-precision edges, wave-uniform: s below thr[0] or above thr[n_rp] for sure
  const float hi_bottom = SPLIT ? P.thr_hi[0] : 0.0f;
  const float lo_top = SPLIT ? P.thr_lo[P.n_rp] : 0.0f;
  const Item it = P.items[blockIdx.x];
  const Group bg = P.bgrp[it.group];
  const int bcx = bg.col % P.nc, bcy = bg.col / P.nc;
  // phase_mode: 0 primary without jackknife, 1 primary with jackknife, 2 touch pass
  const bool two_phase = (P.phase_mode == 1) && (bg.coin == 0);
  const int nphase = two_phase ? 2 : 1;

  unsigned long long ncand = 0, nre = 0;

  for (int ph = 0; ph < nphase; ++ph) {
    // filter on tile region: 0 all, 1 same region as block, 2 other regions
    const int filt = (P.phase_mode == 2) ? 2 : (two_phase ? (ph == 0 ? 1 : 2) : 0);
    const int tgt = (P.phase_mode == 2) ? it.target : (two_phase ? it.target + (ph == 0 ? 1 : 0) : it.target);
    for (int c = tid; c < ncell; c += blockDim.x) {
      lnum[c] = 0;
#pragma unroll
      for (int s = 0; s < NS; ++s) lacc[s * ncell + c] = 0.0;
    }
    __syncthreads();

    const int nchunk = (it.count + 63) >> 6;
    for (int ch = wave; ch < nchunk; ch += nwave) {
      const int base = uni(it.first + ch * 64);
      const int cnt = uni(min(64, it.first + it.count - base));
      const bool valid = lane < cnt;
      const int p = base + (valid ? lane : 0);
      const PRec me = brec[p];
      FRec mf;
      if (SPLIT) mf = bfr[p];
      const double wzlo = brec[base].z;
      const double wzhi = brec[base + cnt - 1].z;

      // ACC 0: a lane keeps the sums of its current histogram cell in registers and adds them to LDS
      // only when its cell changes.  Lanes of a wave hold neighbouring galaxies, so their pairs crowd
      // into few cells; flushing per run instead of per pair divides the same-address LDS atomics.
      int run_cell = -1;
      unsigned run_n = 0;
      double run[ACC == 0 ? NS : 1];
#pragma unroll
      for (int s = 0; s < (ACC == 0 ? NS : 1); ++s) run[s] = 0.0;
      // ACC 1: binned pairs (tile index, lane, cell) go to a per-wave ring in LDS; every 64 entries
      // the wave evaluates them with all lanes busy, so the FP64 shape arithmetic runs on full waves.
      int qn = 0, qr = 0;

      auto run_batch = [&](int nvalid) {
        const bool act = lane < nvalid;
        const unsigned long long e = wq[(qr + lane) & (QCAP - 1)];
        if (act) {
          const int j = (int)(e >> 32);
          const int src = (int)((e >> 16) & 63);
          const int cell = (int)(e & 0xffff);
          const PRec b = brec[base + src];
          const PRec t = trec[j];
          const Gal gb{b.x, b.y, b.w, b.e1, b.e2, b.phi};
          const Gal gt{t.x, t.y, t.w, t.e1, t.e2, t.phi};
          double qv[NQ];
          if (!SWAP)
            pair_values<K, SPLIT>(gb, gt, P.gws, qv);
          else
            pair_values<K, SPLIT>(gt, gb, P.gws, qv);
          lds_add<NS>(lnum, lacc, ncell, cell, 1u, qv);
        }
        qr = (qr + nvalid) & (QCAP - 1);
        qn -= nvalid;
      };

      for (int q = 0; q < P.n_nbr; ++q) {
        const int2 o = P.nbr[q];
        const int ux = bcx + o.x, uy = bcy + o.y;
        const int sx = ux >= 0 ? ux / P.nc : -((-ux + P.nc - 1) / P.nc);
        const int sy = uy >= 0 ? uy / P.nc : -((-uy + P.nc - 1) / P.nc);
        const int tcx = ux - sx * P.nc, tcy = uy - sy * P.nc;
        const int tcoln = tcy * P.nc + tcx;
        const int g0 = P.tcol[tcoln], g1 = P.tcol[tcoln + 1];
        const float Ox = (float)((double)(tcx - bcx) * P.h + P.L * sx);
        const float Oy = (float)((double)(tcy - bcy) * P.h + P.L * sy);
        for (int g = g0; g < g1; ++g) {
          const Group tg = P.tgrp[g];
          if (filt == 1 && tg.region != bg.region) continue;
          if (filt == 2 && tg.region == bg.region) continue;
          for (int sz = -1; sz <= 1; ++sz) {
            double lo, hi;
            if (!SWAP) {
              lo = wzlo + P.pmin - P.L * sz - P.zmargin;
              hi = wzhi + P.pmax - P.L * sz + P.zmargin;
            } else {
              lo = wzlo - P.pmax - P.L * sz - P.zmargin;
              hi = wzhi - P.pmin - P.L * sz + P.zmargin;
            }
            if (hi < tg.zlo || lo > tg.zhi) continue;
            const int j0 = zsearch<false>(trec, tg.start, tg.end, lo);
            const int j1 = zsearch<true>(trec, j0, tg.end, hi);
            if (j0 >= j1) continue;
            // image of the inner galaxy in the reference's terms
            const int PBx = SWAP ? -sx : sx, PBy = SWAP ? -sy : sy, PBz = SWAP ? -sz : sz;
            const bool lane_ok = valid && (SWAP || image_allowed(me.meta, PBx, PBy, PBz));
            const float Oz = (float)(P.L * sz);
            for (int jb = j0; jb < j1; jb += 4) {
              F4 f4;
              if (SPLIT) f4 = *reinterpret_cast<const F4 *>(tfr + jb);  // padded past the end
#pragma unroll
              for (int u = 0; u < 4; ++u) {
                const int j = jb + u;
                if (j >= j1) break;
                bool ok = lane_ok;
                int cell = -1;
                bool unc = ok;
                if (SPLIT) {
                  const FRec tf = f4.r[u];
                  if (SWAP) ok = ok && image_allowed(tf.meta, PBx, PBy, PBz);
                  unc = ok;
                  if (ok) {
                    const float dx = (tf.xl - mf.xl) + Ox;
                    const float dy = (tf.yl - mf.yl) + Oy;
                    const float dz = (tf.zf - mf.zf) + Oz;
                    const float pif = SWAP ? -dz : dz;
                    const float s32 = fmaf(dx, dx, dy * dy);
                    const float uu = (pif - P.pmin32) * P.inv_psize32;
                    const float fu = floorf(uu);
                    const float fr = uu - fu;
                    const bool near_pi = (fr < P.gz) || (fr > 1.0f - P.gz);
                    const bool pi_in = (fu >= 0.0f) && (fu < (float)P.n_pi);
                    float kf = P.lin ? (sqrtf(s32) - P.lrmin) * P.linv : __log2f(s32) * P.la + P.lb;
                    kf = fminf(fmaxf(kf, -1.0f), (float)P.n_rp);
                    const int k = (int)floorf(kf);
                    const int kk = min(max(k, 0), P.n_rp - 1);
                    const bool rp_out = (s32 < hi_bottom) || (s32 >= lo_top);
                    const bool rp_in = (k == kk) && (s32 >= llo[kk]) && (s32 < lhi[kk + 1]);
                    if ((!near_pi && !pi_in) || rp_out) {
                      unc = false;
                    } else if (!near_pi && rp_in) {
                      unc = false;
                      cell = k * P.n_pi + (int)fu;
                    }
                  }
                } else if (SWAP) {
                  ok = ok && image_allowed(trec[j].meta, PBx, PBy, PBz);
                  unc = ok;
                }
                ncand += ok;
                if (unc) {
                  nre += SPLIT;
                  const PRec t = trec[j];
                  if (!SWAP)
                    cell = exact_cell(P, lthr, me.x, me.y, me.z, t.x, t.y, t.z, PBx, PBy, PBz);
                  else
                    cell = exact_cell(P, lthr, t.x, t.y, t.z, me.x, me.y, me.z, PBx, PBy, PBz);
                }
                if constexpr (ACC == 0) {
                  if (cell >= 0) {
                    if (cell != run_cell) {
                      if (run_cell >= 0) lds_add<NS>(lnum, lacc, ncell, run_cell, run_n, run);
                      run_cell = cell;
                      run_n = 0;
#pragma unroll
                      for (int s = 0; s < NS; ++s) run[s] = 0.0;
                    }
                    ++run_n;
                    const PRec t = trec[j];
                    const Gal gme{me.x, me.y, me.w, me.e1, me.e2, me.phi};
                    const Gal gt{t.x, t.y, t.w, t.e1, t.e2, t.phi};
                    double qv[NQ];
                    if (!SWAP)
                      pair_values<K, SPLIT>(gme, gt, P.gws, qv);
                    else
                      pair_values<K, SPLIT>(gt, gme, P.gws, qv);
#pragma unroll
                    for (int s = 0; s < NS; ++s) run[s] += qv[s];
                  }
                } else {
                  const unsigned long long m = __ballot(cell >= 0);
                  if (m) {
                    const int pos = __builtin_amdgcn_mbcnt_hi((unsigned)(m >> 32),
                                                              __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
                    if (cell >= 0)
                      wq[(qr + qn + pos) & (QCAP - 1)] =
                          ((unsigned long long)(unsigned)j << 32) | ((unsigned long long)lane << 16) | (unsigned)cell;
                    qn += __popcll(m);
                    if (qn >= 64) {
                      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                      __builtin_amdgcn_wave_barrier();
                      run_batch(64);
                      __builtin_amdgcn_wave_barrier();
                    }
                  }
                }
              }
            }
          }
        }
      }
      if constexpr (ACC == 0) {
        if (run_cell >= 0) lds_add<NS>(lnum, lacc, ncell, run_cell, run_n, run);
      } else {
        if (qn > 0) {
          __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
          __builtin_amdgcn_wave_barrier();
          run_batch(qn);
          __builtin_amdgcn_wave_barrier();
        }
      }
    }
    __syncthreads();
    for (int c = tid; c < ncell; c += blockDim.x) {
      const unsigned n = lnum[c];
      if (n) {
        atomicAdd(&P.accnum[(size_t)tgt * ncell + c], (unsigned long long)n);
#pragma unroll
        for (int s = 0; s < NS; ++s) unsafeAtomicAdd(&P.acc[((size_t)tgt * NS + s) * ncell + c], lacc[s * ncell + c]);
      }
    }
    __syncthreads();
  }
  if (ncand) atomicAdd(&P.counters[0], ncand);
  if (nre) atomicAdd(&P.counters[1], nre);
}

// ---------------------------------------------------------------------------------------------
// Cell-list construction


// key = ((column * nreg + region) * 2 + coin), or 0xffffffff for galaxies the pass skips
// (include_prob > include_prob_max, corr.cpp:352,414)
__global__ void k_keys(KeyArgs A, const int32_t *__restrict__ order, uint32_t *__restrict__ keys,
                       int32_t *__restrict__ vals) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  const int32_t i = order ? order[p] : (int32_t)p;
  uint32_t key = 0xffffffffu;
  const double inc = A.inc ? A.inc[i] : 0.0;
  if (!(inc > A.inc_max)) {
    int cx = (int)floor((A.x[i] - A.x0) / A.hx), cy = (int)floor((A.y[i] - A.y0) / A.hy);
    cx = min(max(cx, 0), A.ncx - 1);
    cy = min(max(cy, 0), A.ncy - 1);
    const int reg = (A.nreg > 1) ? A.jk[i] : 0;
    const int coin = A.jkp ? (A.jkp[i] >= 0.5) : 0;
    key = ((uint32_t)(cy * A.ncx + cx) * (uint32_t)A.nreg + (uint32_t)reg) * 2u + (uint32_t)coin;
  }
  keys[p] = key;
  vals[p] = i;
}


__global__ void k_gather(GatherArgs A, const int32_t *__restrict__ sidx, PRec *__restrict__ rec,
                         FRec *__restrict__ fr) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  const int32_t i = sidx[p];
  PRec r;
  r.x = A.x[i];
  r.y = A.y[i];
  r.z = A.z[i];
  r.w = A.w ? A.w[i] : 1.0;
  r.e1 = A.e1 ? A.e1[i] : 0.0;
  r.e2 = A.e2 ? A.e2[i] : 0.0;
  r.phi = A.phi ? A.phi[i] : -10.0;
  r.idx = i;
  const double z_max = r.z + A.pmax, z_min = r.z + A.pmin;
  const double ra_max = r.x + A.rmax, ra_min = r.x - A.rmax;
  const double dec_max = r.y + A.rmax, dec_min = r.y - A.rmax;
  int m = 0;
  if (ra_min < 0) m |= kMetaXm;
  if (ra_max > A.L) m |= kMetaXp;
  if (dec_min < 0) m |= kMetaYm;
  if (dec_max > A.L) m |= kMetaYp;
  if (z_min < 0) m |= kMetaZm;
  // PBz>0 sweeps from j = 0 and stops at the first included galaxy outside the window (corr.cpp:407-416)
  if (z_max > A.L && A.has_zfirst && !(A.zfirst_plus_L < z_min)) m |= kMetaZp;
  r.meta = m;
  rec[p] = r;
  int cx = (int)floor((r.x - A.x0) / A.hx), cy = (int)floor((r.y - A.y0) / A.hy);
  cx = min(max(cx, 0), A.ncx - 1);
  cy = min(max(cy, 0), A.ncy - 1);
  FRec f;
  f.xl = (float)(r.x - (A.x0 + cx * A.hx));
  f.yl = (float)(r.y - (A.y0 + cy * A.hy));
  f.zf = (float)r.z;
  f.meta = m;
  fr[p] = f;
}

// Outer galaxies whose PBz = 0 or PBz = -1 window holds no inner galaxy while later ones exist: the
// reference's forward-only start pointer runs off in that case (corr.cpp:361-369, 395-404).

__global__ void k_sparse(SparseArgs A, unsigned long long *count) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned c = 0;
  if (i < A.no && !((A.inco ? A.inco[i] : 0.0) > A.inc_max)) {
    const double z_max = A.zo[i] + A.pmax, z_min = A.zo[i] + A.pmin;
    for (int PBz = -1; PBz <= 0; ++PBz) {
      if (PBz < 0 && !(z_min < 0)) continue;
      const double sh = A.L * PBz;
      int64_t lo = 0, hi = A.ni;
      while (lo < hi) {
        const int64_t mid = lo + (hi - lo) / 2;
        if (A.zi[mid] + sh < z_min)
          lo = mid + 1;
        else
          hi = mid;
      }
      if (lo < A.ni && A.zi[lo] + sh > z_max) ++c;
    }
  }
  if (c) atomicAdd(count, (unsigned long long)c);
}

// ---------------------------------------------------------------------------------------------
// Host-side launch helpers (called from cpc_api.cpp)

hipError_t launch_keys(const KeyArgs &A, const int32_t *order, uint32_t *keys, int32_t *vals, hipStream_t st) {
  if (A.n == 0) return hipSuccess;
  const int bs = 256;
  k_keys<<<(unsigned)((A.n + bs - 1) / bs), bs, 0, st>>>(A, order, keys, vals);
  return hipGetLastError();
}

hipError_t launch_gather(const GatherArgs &A, const int32_t *sidx, PRec *rec, FRec *fr, hipStream_t st) {
  if (A.n == 0) return hipSuccess;
  const int bs = 256;
  k_gather<<<(unsigned)((A.n + bs - 1) / bs), bs, 0, st>>>(A, sidx, rec, fr);
  return hipGetLastError();
}

hipError_t launch_sparse(const SparseArgs &A, unsigned long long *count, hipStream_t st) {
  if (A.no == 0 || A.ni == 0) return hipSuccess;
  const int bs = 256;
  k_sparse<<<(unsigned)((A.no + bs - 1) / bs), bs, 0, st>>>(A, count);
  return hipGetLastError();
}

// Stable sort of (key, value) pairs; temporary storage grows as needed.
template <typename KeyT>
static hipError_t sort_pairs_impl(void *&tmp, size_t &tmp_bytes, const KeyT *kin, KeyT *kout, const int32_t *vin,
                                  int32_t *vout, int64_t n, int end_bit, hipStream_t st) {
  size_t need = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, (int)n, 0, end_bit, st);
  if (e != hipSuccess) return e;
  if (need > tmp_bytes) {
    if (tmp) (void)hipFree(tmp);
    tmp = nullptr;
    tmp_bytes = 0;
    e = hipMalloc(&tmp, need);
    if (e != hipSuccess) return e;
    tmp_bytes = need;
  }
  return hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, kin, kout, vin, vout, (int)n, 0, end_bit, st);
}

hipError_t sort_u32(void *&tmp, size_t &tmp_bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                    int32_t *vout, int64_t n, hipStream_t st) {
  return sort_pairs_impl<uint32_t>(tmp, tmp_bytes, kin, kout, vin, vout, n, 32, st);
}

hipError_t sort_f64(void *&tmp, size_t &tmp_bytes, const double *kin, double *kout, const int32_t *vin,
                    int32_t *vout, int64_t n, hipStream_t st) {
  return sort_pairs_impl<double>(tmp, tmp_bytes, kin, kout, vin, vout, n, 64, st);
}

template <int K, bool SPLIT, bool SWAP, int ACC>
static hipError_t launch_pair_t(const KArgs &P, int threads, size_t lds, const PRec *brec, const FRec *bfr,
                                const PRec *trec, const FRec *tfr, hipStream_t st) {
  auto fn = pair_kernel<K, SPLIT, SWAP, ACC>;
  hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  fn<<<P.n_items, threads, lds, st>>>(P, brec, bfr, trec, tfr);
  return hipGetLastError();
}

// LDS bytes a pair kernel needs for histograms of `ncell` cells and `nslots` f64 slots.
size_t pair_lds_bytes(int ncell, int nslots, int n_rp, int nwave, int acc_mode) {
  size_t b = (size_t)((ncell * 4 + 15) & ~15) + (size_t)nslots * ncell * 8 + (size_t)(n_rp + 1) * 8 +
             (size_t)2 * (n_rp + 1) * 4;
  b = (b + 4 + 7) & ~(size_t)7;
  if (acc_mode == 1) b += (size_t)nwave * 128 * 8;
  return b;
}

hipError_t launch_pair(int kind, bool split, bool swap, const KArgs &P, int threads, const PRec *brec,
                       const FRec *bfr, const PRec *trec, const FRec *tfr, hipStream_t st) {
  if (P.n_items == 0) return hipSuccess;
  const size_t lds = pair_lds_bytes(P.ncell, P.nslots, P.n_rp, threads / 64, P.acc_mode);
#define CPC_DISPATCH2(KK, AA)                                                                          \
  if (split && swap) return launch_pair_t<KK, true, true, AA>(P, threads, lds, brec, bfr, trec, tfr, st); \
  if (split) return launch_pair_t<KK, true, false, AA>(P, threads, lds, brec, bfr, trec, tfr, st);        \
  if (swap) return launch_pair_t<KK, false, true, AA>(P, threads, lds, brec, bfr, trec, tfr, st);         \
  return launch_pair_t<KK, false, false, AA>(P, threads, lds, brec, bfr, trec, tfr, st);
#define CPC_DISPATCH(KK)                    \
  if (kind == KK) {                         \
    if (P.acc_mode == 1) {                  \
      CPC_DISPATCH2(KK, 1)                  \
    } else {                                \
      CPC_DISPATCH2(KK, 0)                  \
    }                                       \
  }
  CPC_DISPATCH(kDD)
  CPC_DISPATCH(kSD)
  CPC_DISPATCH(kSS0)
  CPC_DISPATCH(kSS1)
#undef CPC_DISPATCH
#undef CPC_DISPATCH2
  return hipErrorInvalidValue;
}

}  // namespace cpc
