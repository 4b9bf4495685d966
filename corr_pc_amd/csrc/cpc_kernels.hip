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
// Jackknife.  Every pair is evaluated once (block = outer catalog) and accumulated into the target
// of its item.  An item's block galaxies share a region a and coin, and its tile filter (Item tmode,
// treg) takes only tile groups whose pairs touch the same regions, so the target fixes them:
//   target 2a+1   : pairs that touch region a only (through the outer galaxy)
//   target 2a     : pairs that touch no region
//   target Y(a,x,b): pairs with an inner galaxy of region b != a and coin >= 0.5, which touch b, and
//                    a as well when x = 1
// (jk.cpp:201-223).  k_reduce then adds every target into the full sample and every target that
// touches r into touch_r; the host forms leave-one-out jk_r = full - touch_r (jk.cpp:74-81).

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "cpc_device.h"
#include "cpc_internal.h"

namespace cpc {

// Tiles reach the wave in batches of kBatch z-consecutive records, loaded with coalesced vector loads
// one batch ahead of use: lanes 4t..4t+3 hold the four 16-byte quarters of tile t's record, and the
// wave reads a tile's fields with v_readlane, so no tile load stands between two candidate tests.
constexpr int kBatch = 16;
static_assert(kRecPad >= kBatch, "record padding must cover a batch read past the end of a range");

__device__ __forceinline__ float rl_f(unsigned v, int l) { return __uint_as_float(__builtin_amdgcn_readlane(v, l)); }
__device__ __forceinline__ int rl_i(unsigned v, int l) { return (int)__builtin_amdgcn_readlane(v, l); }
__device__ __forceinline__ double rl_d(unsigned lo, unsigned hi, int l) {
  return __hiloint2double((int)__builtin_amdgcn_readlane(hi, l), (int)__builtin_amdgcn_readlane(lo, l));
}
__device__ __forceinline__ unsigned long long popc_ballot(bool b) { return (unsigned long long)__popcll(__ballot(b)); }
__device__ __forceinline__ float wave_min(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o));
  return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o));
  return v;
}

using CKArgs = __attribute__((address_space(4))) const KArgs;

// The box pair kernel.  SPLIT classifies candidates in FP32 (log rp bins) and re-evaluates the
// guard band in FP64; without it every candidate takes the FP64 path.  AUTO compiles in corels1's
// order predicate (auto_mode 1 and 2).  RMU compiles in the r-mu binning of coordinates 7 (with
// SPLIT false only: r-mu candidates all take the FP64 path).  Kinds, AUTO and RMU are separate
// instantiations so that each keeps only the registers and branches it needs.  Occupancy: the split
// kernels keep 4 waves per SIMD (128 registers), the split density kernel 5 (96 registers: with
// 256-thread groups a CU then holds 5 groups), the FP64 kernels 2.
template <int K, bool SPLIT, bool AUTO, bool RMU>
__global__ void __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(SPLIT ? (KindNV<K>::NV == 0 ? 5 : 4) : 2)))
pair_kernel(const KArgs *__restrict__ jobs, const Item *__restrict__ items) {
  constexpr int NV = KindNV<K>::NV;
  constexpr int NE = KindNV<K>::NE;
  constexpr int NS = 1 + NV + NE;  // f64 slots accumulated per cell: wt, vals, errs
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int next_chunk;  // next chunk of the item that no wave has taken
  const Item it = items[blockIdx.x];
  // the pass arguments of the item's job, read as constant memory: wave-uniform scalar loads the
  // compiler may repeat instead of holding them in registers, as it does with kernel arguments
  const CKArgs &P = ((const CKArgs *)jobs)[it.job];
  // the arguments of the inner loops, read once: the compiler keeps them in registers instead of
  // reloading them inside the loop, where each reload would also wait for the LDS atomics in flight
  const HotArgs H{P.L,         P.pmin, P.pmax,  P.rmax, P.psize, P.gz, P.pioff32, P.s_out_lo, P.s_out_hi,
                  P.s_cert_lo, P.linv, P.lrmin, P.la, P.lb, P.gA, P.gB, P.gC, P.gD, P.inv_psize32, P.n_pi, P.n_rp,
                  P.lin, P.gws, P.auto_mode, P.row0 * P.n_pi, (P.n_rp - P.row0) * P.n_pi, P.wlo, P.whi, P.rmu,
                  P.cull_r2, P.cull_d};
  // three addends of the FP32 classification, held in vector registers where the kind has
  // registers to spare (fewer than 7 slots): an FMA reads at most one scalar register, so with two
  // scalar operands the loop copies one to a VGPR on every tile
  float pioff_v = H.pioff32, gB_v = H.gB, lb_v = H.lb;
  if constexpr (NS < 7) asm volatile("" : "+v"(pioff_v), "+v"(gB_v), "+v"(lb_v));
  const PRec *__restrict__ brec = P.brec;
  const double *__restrict__ bz = P.bz;
  const double *__restrict__ tz = P.tz;
  const uint4 *__restrict__ tq = reinterpret_cast<const uint4 *>(P.trec);  // records as 16-byte quarters
  // LDS: ncopy histogram copies of the cells from lbase on, lnum[ncopy][lcell] and
  // lacc[ncopy][lcell][NS], then the thresholds.  Neighbouring lanes often close runs of the same cell
  // in the same instruction, and LDS atomics to one address serialise (about 3 cycles per lane,
  // against about 6 cycles for a whole instruction without collisions); copies spread those lanes.
  // Row cut and copies are compiled in only where the flush dominates: 7 or more slots per cell.
  constexpr bool CUT = NS >= 7;
  const int ncell = P.ncell, ncopy = CUT ? P.ncopy : 1, lcell = CUT ? H.lcell : P.ncell, lbase = CUT ? H.lbase : 0;
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

  for (int k = tid; k <= H.n_rp; k += blockDim.x) lthr[k] = P.thr[k];
  for (int c = tid; c < ncopy * lcell; c += blockDim.x) {
    lnum[c] = 0;
#pragma unroll
    for (int s = 0; s < NS; ++s) lacc[(size_t)c * NS + s] = 0.0;
  }
  if (tid == 0) next_chunk = nwave;
  __syncthreads();

  const Group bg = P.bgrp[it.group0];  // region and coin of every chunk of the item
  const int tmode = it.tmode, treg = it.treg;
  unsigned long long *const gnum = P.accnum + (size_t)it.target * ncell;
  double *const gacc = P.acc + (size_t)it.target * NS * ncell;
  unsigned long long ncand = 0, nre = 0;  // wave totals
  unsigned nre_l = 0;  // small kinds: this lane's FP64 re-evaluations (summed over the wave at the end)
  int next_nx = 0;

  // chunks come sorted by decreasing cost; each wave starts on its own and then takes the next
  // unclaimed one, so the waves of the group finish close together
  for (int ch = wave; ch < it.n; ch = uni(next_nx)) {
    const Chunk ck = P.chunks[it.begin + ch];
    const int col = uni(P.bgrp[ck.group].col);
    const int bcx = col % P.nc, bcy = col / P.nc;
    const int base = uni(ck.first);
    const int cnt = uni(ck.count);
    const bool valid = lane < cnt;
    const int p = base + (valid ? lane : 0);
    const PRec me = brec[p];
    const double mz = bz[p];
    const int midx = AUTO ? P.bidx[p] : 0;
    const double wzlo = bz[base];
    const double wzhi = bz[base + cnt - 1];
    // bounding box of the wave's block galaxies in column-local FP32 coordinates (tile cull)
    const float bx0 = wave_min(valid ? me.xl : INFINITY), bx1 = wave_max(valid ? me.xl : -INFINITY);
    const float by0 = wave_min(valid ? me.yl : INFINITY), by1 = wave_max(valid ? me.yl : -INFINITY);

    // A lane keeps the sums of its current histogram cell in registers and adds them to LDS only
    // when its cell changes: lanes of a wave hold neighbouring galaxies, so their pairs crowd into
    // few cells, and flushing per run instead of per pair divides the same-address LDS atomics.
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

    for (int q = 0; q < P.n_nbr; ++q) {
      const int2 o = P.nbr[q];
      const int ux = bcx + o.x, uy = bcy + o.y;
      const int sx = ux >= 0 ? ux / P.nc : -((-ux + P.nc - 1) / P.nc);
      const int sy = uy >= 0 ? uy / P.nc : -((-uy + P.nc - 1) / P.nc);
      const int tcx = ux - sx * P.nc, tcy = uy - sy * P.nc;
      const int tcoln = tcy * P.nc + tcx;
      const int g0 = P.tcol[tcoln], g1 = P.tcol[tcoln + 1];
      const float Ox = (float)((double)(tcx - bcx) * P.h + H.L * sx);
      const float Oy = (float)((double)(tcy - bcy) * P.h + H.L * sy);
      for (int g = g0; g < g1; ++g) {
        const Group tg = P.tgrp[g];
        // the item's tile filter: every pair it meets belongs to its target
        const bool same = tg.region == bg.region;
        const bool take = tmode == kTileAll || (tmode == kTileSame && same) ||
                          (tmode == kTileOtherC0 && !same && tg.coin == 0) ||
                          (tmode == kTileSameOrC0 && (same || tg.coin == 0)) ||
                          (tmode == kTileRegionC1 && tg.region == treg && tg.coin == 1);
        if (!take) continue;
        for (int sz = -1; sz <= 1; ++sz) {
          const double lo = wzlo + H.wlo - H.L * sz - P.zmargin;
          const double hi = wzhi + H.whi - H.L * sz + P.zmargin;
          if (hi < tg.zlo || lo > tg.zhi) continue;
          const int j0 = zsearch<false>(tz, tg.start, tg.end, lo);
          const int j1 = zsearch<true>(tz, j0, tg.end, hi);
          if (j0 >= j1) continue;
          // (sx, sy, sz) is the inner galaxy's image in the reference's terms
          const bool lane_ok = valid && image_allowed(me.meta, sx, sy, sz);
          // corels1: the inner galaxy must come after the outer one (sorted: images with PBz <= 0,
          // in (z, index) order; unsorted: every image, in index order)
          const bool ordered = AUTO && (H.auto_mode == 2 || (H.auto_mode == 1 && sz <= 0));
          const bool unsorted_off = H.auto_mode != 2;
          const unsigned long long ok_lanes = AUTO ? 0ull : popc_ballot(lane_ok);
          // FP32 offsets of this lane to the tile column: d = tile local + offset
          const float ax = Ox - me.xl, ay = Oy - me.yl, az = (float)(H.L * sz) - me.zf;

          uint4 cq = tq[(size_t)j0 * 4 + lane];
          double cz = tz[j0 + (lane & (kBatch - 1))];  // z of tile t in lane t
          for (int jb = j0; jb < j1; jb += kBatch) {
            uint4 nq = cq;
            double nz = cz;
            if (jb + kBatch < j1) {
              nq = tq[(size_t)(jb + kBatch) * 4 + lane];
              nz = tz[jb + kBatch + (lane & (kBatch - 1))];
            }
            const int nt = min(kBatch, j1 - jb);
            // tiles within reach of the wave's bounding box: lane 4t+3 holds tile t's column-local
            // {xl, yl}; the others cannot bin with any lane and are skipped by the whole wave
            unsigned long long tmask;
            {
              const float X = __uint_as_float(cq.x) + Ox, Y = __uint_as_float(cq.y) + Oy;
              const float ex = fmaxf(fmaxf(bx0 - X, X - bx1) - H.cull_d, 0.0f);
              const float ey = fmaxf(fmaxf(by0 - Y, Y - by1) - H.cull_d, 0.0f);
              tmask = __ballot((lane & 3) == 3 && (lane >> 2) < nt && fmaf(ex, ex, ey * ey) <= H.cull_r2);
            }
            // without the order predicate every tile of the batch meets the same candidate lanes
            if (!AUTO) ncand += ok_lanes * (unsigned long long)__popcll(tmask);
            for (; tmask; tmask &= tmask - 1) {
              const int t = uni(__builtin_ctzll(tmask) >> 2);
              const int l3 = 4 * t + 3;  // lane of the quarter {xl, yl, zf, meta}
              bool ok = lane_ok;
              if (AUTO && ordered) {
                const double zi = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(cz), t),
                                                   __builtin_amdgcn_readlane(__double2loint(cz), t));
                bool after = zi > mz && unsorted_off;
                const bool tie = zi == mz || !unsorted_off;
                if (__ballot(ok && tie) != 0ull) {
                  const int tix = P.tidx[jb + t];
                  after = after || (tie && tix > midx);
                }
                ok = ok && after;
              }
              int cell = -1;
              bool unc = ok;
              if (SPLIT) {
                // FP32 classification, branch-free: a candidate is certainly outside every bin
                // (`out`), certainly in the bin (fk, fu) (`cert`: pi and rp bin fractions outside
                // their guard bands), or uncertain and re-evaluated in FP64
                const float dx = rl_f(cq.x, l3) + ax;
                const float dy = rl_f(cq.y, l3) + ay;
                const float dz = rl_f(cq.z, l3) + az;
                const float s32 = fmaf(dx, dx, dy * dy);
                const float uu = fmaf(dz, H.inv_psize32, pioff_v);
                const float fu = floorf(uu);
                const float fr = uu - fu;
                const bool near_pi = (fr < H.gz) || (fr > 1.0f - H.gz);
                const bool pi_in = (fu >= 0.0f) && (fu < (float)H.n_pi);
                const bool out = (!near_pi && !pi_in) || (s32 < H.s_out_lo) || (s32 >= H.s_out_hi);
                // s32 >= s_cert_lo > 0 wherever cert can hold, so kf and gk are finite there
                const float rinv = __builtin_amdgcn_rsqf(s32);
                // one FMA where lb sits in a VGPR (one rounding fewer than the guard band allows for)
                const float kf = NS < 7 ? fmaf(__log2f(s32), H.la, lb_v) : __log2f(s32) * H.la + H.lb;
                const float fk = floorf(kf);
                const float frk = kf - fk;
                const float gk = fmaf(fmaf(H.gC, rinv, gB_v), rinv, H.gA);
                const bool cert = !near_pi && s32 >= H.s_cert_lo && frk >= gk && frk <= 1.0f - gk;
                unc = ok && !out && !cert;
                if (ok && !out && cert && fk >= 0.0f && fk < (float)H.n_rp)
                  cell = (int)fmaf(fk, (float)H.n_pi, fu);
              }
              if (AUTO) ncand += popc_ballot(ok);
              // the small kinds take a plain branch, which the wave skips when no lane is uncertain,
              // and count per lane (a ballot of a lane mask is rebuilt from a 0/1 vector on every
              // tile; a wave total updated inside the branch would become a per-lane value); the
              // shape kinds keep the wave-uniform test, with which their register allocation spills
              // nothing inside the tile loops
              constexpr bool kPlain = NS < 7;
              if (kPlain ? unc : (__ballot(unc) != 0ull)) {
                if (SPLIT) {
                  if (kPlain)
                    ++nre_l;
                  else
                    nre += popc_ballot(unc);
                }
                const double tx = rl_d(cq.x, cq.y, 4 * t), ty = rl_d(cq.z, cq.w, 4 * t);
                const double tzz = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(cz), t),
                                                    __builtin_amdgcn_readlane(__double2loint(cz), t));
                if (unc) {
                  if constexpr (SPLIT)
                    cell = exact_cell<RMU>(P, lthr, me.x, me.y, mz, tx, ty, tzz, sx, sy, sz);
                  else
                    cell = exact_cell<RMU>(H, lthr, me.x, me.y, mz, tx, ty, tzz, sx, sy, sz);
                }
              }
              if (__ballot(cell >= 0) != 0ull) {
                const Gal gt{rl_d(cq.x, cq.y, 4 * t),     rl_d(cq.z, cq.w, 4 * t),
                             rl_d(cq.x, cq.y, 4 * t + 1), rl_d(cq.z, cq.w, 4 * t + 1),
                             rl_d(cq.x, cq.y, 4 * t + 2), rl_d(cq.z, cq.w, 4 * t + 2)};
                if (cell >= 0) {
                  const Gal gme{me.x, me.y, me.w, me.e1, me.e2, me.phi};
                  double qv[NS];
                  if constexpr (K == kED) {
                    // g1 = outer, g2 = inner in the reference's roles; (sx, sy, sz) is g2's image
                    const double tzv = __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(cz), t),
                                                        __builtin_amdgcn_readlane(__double2loint(cz), t));
                    const Gal &ax3 = H.gws == 2 ? gt : gme;
                    ed_values(gme.w, gt.w, ax3.e1, ax3.e2, ax3.phi, gme.x - (gt.x + H.L * sx),
                              gme.y - (gt.y + H.L * sy), mz - (tzv + H.L * sz), qv);
                  } else {
                    pair_values<K, SPLIT>(gme, gt, H.gws, qv);
                  }
                  if constexpr (NS == 1) {
                    // one slot (density pairs): the run restarts through selects, and only the
                    // flush branches (no copies of the run at the join of two branches)
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
            cq = nq;
            cz = nz;
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
  if constexpr (NS < 7) {
    unsigned long long t = nre_l;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o);
    nre += t;
  }
  if (lane == 0) {
    if (ncand) atomicAdd(&P.counters[0], ncand);
    if (nre) atomicAdd(&P.counters[1], nre);
  }
}

// ---------------------------------------------------------------------------------------------
// Cell-list construction


// key = ((column * nreg + region) * 2 + coin), or 0xffffffff for galaxies the pass skips
// (include_prob > include_prob_max, corr.cpp:352,414)
__device__ __forceinline__ uint32_t cell_key(const KeyArgs &A, int32_t i) {
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
  return key;
}

__global__ void k_keys(KeyArgs A, const int32_t *__restrict__ order, uint32_t *__restrict__ keys,
                       int32_t *__restrict__ vals) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  const int32_t i = order ? order[p] : (int32_t)p;
  keys[p] = cell_key(A, i);
  vals[p] = i;
}


// Record p of a side from catalog galaxy i (position, weight, shape, image flags, column-local
// FP32 coordinates)
__device__ __forceinline__ void gather_rec(const GatherArgs &A, int32_t i, int64_t p, PRec *__restrict__ rec,
                                           double *__restrict__ zs, int32_t *__restrict__ idx) {
  PRec r;
  r.x = A.x[i];
  r.y = A.y[i];
  const double z = A.z[i];
  r.w = A.w ? A.w[i] : 1.0;
  r.e1 = A.e1 ? A.e1[i] : 0.0;
  r.e2 = A.e2 ? A.e2[i] : 0.0;
  r.phi = A.phi ? A.phi[i] : -10.0;
  const double z_max = z + A.pmax, z_min = z + A.pmin;
  const double ra_max = r.x + A.rmax, ra_min = r.x - A.rmax;
  const double dec_max = r.y + A.rmax, dec_min = r.y - A.rmax;
  const int af = A.aflags ? A.aflags[i] : 3;
  int m = 0;
  if (ra_min < 0) m |= kMetaXm;
  if (ra_max > A.L) m |= kMetaXp;
  if (dec_min < 0) m |= kMetaYm;
  if (dec_max > A.L) m |= kMetaYp;
  if (z_min < 0 && (af & 2)) m |= kMetaZm;
  if (af & 1) m |= kMetaZ0;
  // PBz>0 sweeps from j = 0 and stops at the first included galaxy outside the window (corr.cpp:407-416)
  if (z_max > A.L && A.has_zfirst && !(A.zfirst_plus_L < z_min)) m |= kMetaZp;
  r.meta = m;
  int cx = (int)floor((r.x - A.x0) / A.hx), cy = (int)floor((r.y - A.y0) / A.hy);
  cx = min(max(cx, 0), A.ncx - 1);
  cy = min(max(cy, 0), A.ncy - 1);
  r.xl = (float)(r.x - (A.x0 + cx * A.hx));
  r.yl = (float)(r.y - (A.y0 + cy * A.hy));
  r.zf = (float)z;
  rec[p] = r;
  zs[p] = z;
  idx[p] = i;
}

__global__ void k_gather(GatherArgs A, const int32_t *__restrict__ sidx, PRec *__restrict__ rec,
                         double *__restrict__ zs, int32_t *__restrict__ idx) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= A.n) return;
  gather_rec(A, sidx[p], p, rec, zs, idx);
}

// Group boundaries of a side from its sorted keys: for every possible key k < nkeys the records of
// group k are [out[k].x, out[k].y), with the catalog indices of its first and last galaxy in .z and
// .w (-1 when the group is empty); out[nkeys].x is the number of included records (excluded
// galaxies carry key 0xffffffff >= nkeys).
// Boundaries of group k over n sorted keys whose low 32 bits are the cell keys (all equal above)
template <typename KeyT>
__device__ __forceinline__ int4 group_bounds_at(const KeyT *__restrict__ keys, const int32_t *__restrict__ vals,
                                                int64_t n, uint32_t nkeys, int64_t k) {
  auto first_at_least = [&](uint64_t v) {
    int64_t lo = 0, hi = n;
    while (lo < hi) {
      const int64_t mid = lo + ((hi - lo) >> 1);
      if ((uint64_t)(uint32_t)keys[mid] < v)
        lo = mid + 1;
      else
        hi = mid;
    }
    return lo;
  };
  const int64_t b = first_at_least((uint64_t)k);
  const int64_t e = k < (int64_t)nkeys ? first_at_least((uint64_t)k + 1) : b;
  return make_int4((int)b, (int)e, b < e ? vals[b] : -1, b < e ? vals[e - 1] : -1);
}

__global__ void k_group_bounds(const uint32_t *__restrict__ keys, const int32_t *__restrict__ vals, int64_t n,
                               uint32_t nkeys, int4 *__restrict__ out) {
  const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (k > (int64_t)nkeys) return;
  out[k] = group_bounds_at(keys, vals, n, nkeys, k);
}

// Outer galaxies whose PBz = 0 or PBz = -1 window holds no inner galaxy while later ones exist: the
// reference's forward-only start pointer runs off in that case (corr.cpp:361-369, 395-404).

__device__ __forceinline__ void sparse_at(const SparseArgs &A, int64_t i) {
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
  if (c) atomicAdd(A.count, (unsigned long long)c);
}

// Entry e of a table of launches, served by blocks [off[e], off[e+1]) (the block's own entry is
// found by binary search; off is ascending with off[0] = 0)
__device__ __forceinline__ int table_entry(const int64_t *__restrict__ off, int n, int64_t b) {
  int lo = 0, hi = n;  // off[lo] <= b < off[hi]
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (off[mid] <= b)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

__global__ void k_sparse_many(const SparseArgs *__restrict__ tab, const int64_t *__restrict__ off, int n) {
  const int e = table_entry(off, n, blockIdx.x);
  sparse_at(tab[e], (int64_t)(blockIdx.x - off[e]) * blockDim.x + threadIdx.x);
}

// Batched cell-list build (SideBuild): every side's sort keys, entry number above the cell key
__global__ void k_keys_many(const SideBuild *__restrict__ tab, const int64_t *__restrict__ blk, int n,
                            uint64_t *__restrict__ keys, int32_t *__restrict__ vals) {
  const int e = table_entry(blk, n, blockIdx.x);
  const SideBuild &S = tab[e];
  const int64_t p = (int64_t)(blockIdx.x - blk[e]) * blockDim.x + threadIdx.x;
  if (p >= S.ka.n) return;
  keys[S.off + p] = ((uint64_t)e << 32) | cell_key(S.ka, (int32_t)p);
  vals[S.off + p] = (int32_t)p;
}

// ... each side's group boundaries within its run of the sorted arrays
__global__ void k_bounds_many(const SideBuild *__restrict__ tab, const int64_t *__restrict__ blk, int n,
                              const uint64_t *__restrict__ keys, const int32_t *__restrict__ vals,
                              int4 *__restrict__ bounds) {
  const int e = table_entry(blk, n, blockIdx.x);
  const SideBuild &S = tab[e];
  const int64_t k = (int64_t)(blockIdx.x - blk[e]) * blockDim.x + threadIdx.x;
  if (k > (int64_t)S.nkeys) return;
  bounds[S.boff + k] = group_bounds_at(keys + S.off, vals + S.off, S.ka.n, S.nkeys, k);
}

// ... and each side's records in sorted order (the sky's SRec and windows too)
__global__ void k_gather_many(const SideBuild *__restrict__ tab, const int64_t *__restrict__ blk, int n,
                              const int32_t *__restrict__ vals) {
  const int e = table_entry(blk, n, blockIdx.x);
  const SideBuild &S = tab[e];
  const int64_t p = (int64_t)(blockIdx.x - blk[e]) * blockDim.x + threadIdx.x;
  if (p >= S.ga.n) return;
  const int32_t i = vals[S.off + p];
  gather_rec(S.ga, i, p, S.rec, S.zs, S.idx);
  if (S.src_sr) {
    S.sr[p] = S.src_sr[i];
    if (S.src_win) S.win[p] = S.src_win[i];
  }
}

// The split kernels' FP32 approximations against FP64 over the FP32 values b0 .. b0+n-1 (as bit
// patterns): out[0] the scaled log2 error, out[1] the rsq error, each a maximum of non-negative
// doubles kept as their bit patterns (which order like the values).
__global__ void k_fp32_error(uint32_t b0, uint32_t n, unsigned long long *out) {
  double e_log = 0.0, e_rsq = 0.0;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float s = __uint_as_float(b0 + i);
    const double lg = log2((double)s);
    e_log = fmax(e_log, fabs((double)__log2f(s) - lg) / fmax(1.0, fabs(lg)));
    e_rsq = fmax(e_rsq, fabs((double)__builtin_amdgcn_rsqf(s) * sqrt((double)s) - 1.0));
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    e_log = fmax(e_log, __shfl_xor(e_log, o));
    e_rsq = fmax(e_rsq, __shfl_xor(e_rsq, o));
  }
  if ((threadIdx.x & 63) == 0) {
    atomicMax(&out[0], (unsigned long long)__double_as_longlong(e_log));
    atomicMax(&out[1], (unsigned long long)__double_as_longlong(e_rsq));
  }
}

// One thread per (output, slot, cell); slot -1 is the count.  Summation follows the list order, so
// the result does not depend on the launch.
__device__ __forceinline__ void reduce_at(const ReduceArgs &A, int64_t i) {
  const int64_t per_out = (int64_t)(A.ns + 1) * A.ncell;
  if (i >= (int64_t)A.n_out * per_out) return;
  const int o = (int)(i / per_out);
  const int s = (int)((i % per_out) / A.ncell) - 1;
  const int c = (int)(i % A.ncell);
  const int b = A.beg[o], e = A.beg[o + 1];
  if (s < 0) {
    unsigned long long n = 0;
    for (int k = b; k < e; ++k) n += A.accn[(size_t)A.lst[k] * A.ncell + c];
    A.outn[(size_t)o * A.ncell + c] = n;
  } else {
    double v = 0.0;
    for (int k = b; k < e; ++k) v += A.acc[((size_t)A.lst[k] * A.ns + s) * A.ncell + c];
    A.out[((size_t)o * A.ns + s) * A.ncell + c] = v;
  }
}

__global__ void k_reduce(ReduceArgs A) { reduce_at(A, (int64_t)blockIdx.x * blockDim.x + threadIdx.x); }

__global__ void k_reduce_many(const ReduceArgs *__restrict__ tab, const int64_t *__restrict__ off, int n) {
  const int e = table_entry(off, n, blockIdx.x);
  reduce_at(tab[e], (int64_t)(blockIdx.x - off[e]) * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------------------------
// Host-side launch helpers (called from cpc_api.cpp)

hipError_t launch_reduce(const ReduceArgs &A, hipStream_t st) {
  const int64_t n = (int64_t)A.n_out * (A.ns + 1) * A.ncell;
  if (n == 0) return hipSuccess;
  const int bs = 256;
  k_reduce<<<(unsigned)((n + bs - 1) / bs), bs, 0, st>>>(A);
  return hipGetLastError();
}

hipError_t launch_fp32_error(uint32_t b0, uint32_t n, unsigned long long *out, hipStream_t st) {
  if (n == 0) return hipSuccess;
  k_fp32_error<<<2048, 256, 0, st>>>(b0, n, out);
  return hipGetLastError();
}

hipError_t launch_keys(const KeyArgs &A, const int32_t *order, uint32_t *keys, int32_t *vals, hipStream_t st) {
  if (A.n == 0) return hipSuccess;
  const int bs = 256;
  k_keys<<<(unsigned)((A.n + bs - 1) / bs), bs, 0, st>>>(A, order, keys, vals);
  return hipGetLastError();
}

hipError_t launch_gather(const GatherArgs &A, const int32_t *sidx, PRec *rec, double *zs, int32_t *idx,
                         hipStream_t st) {
  if (A.n == 0) return hipSuccess;
  const int bs = 256;
  k_gather<<<(unsigned)((A.n + bs - 1) / bs), bs, 0, st>>>(A, sidx, rec, zs, idx);
  return hipGetLastError();
}

hipError_t launch_group_bounds(const uint32_t *keys, const int32_t *vals, int64_t n, uint32_t nkeys, int4 *out,
                               hipStream_t st) {
  const int bs = 256;
  k_group_bounds<<<(unsigned)(((int64_t)nkeys + 1 + bs - 1) / bs), bs, 0, st>>>(keys, vals, n, nkeys, out);
  return hipGetLastError();
}

int64_t reduce_blocks(const ReduceArgs &A) { return ((int64_t)A.n_out * (A.ns + 1) * A.ncell + 255) / 256; }

int64_t sparse_blocks(const SparseArgs &A) { return (A.no == 0 || A.ni == 0) ? 0 : (A.no + 255) / 256; }

hipError_t launch_reduce_many(const ReduceArgs *tab, const int64_t *off, int n, int64_t blocks, hipStream_t st) {
  if (n == 0 || blocks == 0) return hipSuccess;
  k_reduce_many<<<(unsigned)blocks, 256, 0, st>>>(tab, off, n);
  return hipGetLastError();
}

hipError_t launch_sparse_many(const SparseArgs *tab, const int64_t *off, int n, int64_t blocks, hipStream_t st) {
  if (n == 0 || blocks == 0) return hipSuccess;
  k_sparse_many<<<(unsigned)blocks, 256, 0, st>>>(tab, off, n);
  return hipGetLastError();
}

// Stable sort of (key, value) pairs; temporary storage grows as needed, unless `may_alloc` is false
// (then a too small `tmp` is an error).
template <typename KeyT>
static hipError_t sort_pairs_impl(void *&tmp, size_t &tmp_bytes, const KeyT *kin, KeyT *kout, const int32_t *vin,
                                  int32_t *vout, int64_t n, int end_bit, hipStream_t st, bool may_alloc) {
  size_t need = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, need, kin, kout, vin, vout, (int)n, 0, end_bit, st);
  if (e != hipSuccess) return e;
  if (need > tmp_bytes && !may_alloc) return hipErrorOutOfMemory;
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
                    int32_t *vout, int64_t n, hipStream_t st, bool may_alloc) {
  return sort_pairs_impl<uint32_t>(tmp, tmp_bytes, kin, kout, vin, vout, n, 32, st, may_alloc);
}

hipError_t sort_f64(void *&tmp, size_t &tmp_bytes, const double *kin, double *kout, const int32_t *vin,
                    int32_t *vout, int64_t n, hipStream_t st, bool may_alloc) {
  return sort_pairs_impl<double>(tmp, tmp_bytes, kin, kout, vin, vout, n, 64, st, may_alloc);
}

hipError_t build_sides_many(const SideBuild *tab, int n, const int64_t *blk_el, int64_t nblk_el,
                            const int64_t *blk_key, int64_t nblk_key, uint64_t *keys_a, uint64_t *keys_b,
                            int32_t *vals_a, int32_t *vals_b, int64_t total, int4 *bounds, void *&tmp,
                            size_t &tmp_bytes, hipStream_t st) {
  if (n == 0 || total == 0) return hipSuccess;
  k_keys_many<<<(unsigned)nblk_el, 256, 0, st>>>(tab, blk_el, n, keys_a, vals_a);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  // entry numbers take the bits above the 32-bit cell key
  int end_bit = 32;
  while (end_bit < 64 && ((uint64_t)(n - 1) >> (end_bit - 32)) != 0) ++end_bit;
  e = sort_pairs_impl<uint64_t>(tmp, tmp_bytes, keys_a, keys_b, vals_a, vals_b, total, end_bit, st, true);
  if (e != hipSuccess) return e;
  k_bounds_many<<<(unsigned)nblk_key, 256, 0, st>>>(tab, blk_key, n, keys_b, vals_b, bounds);
  e = hipGetLastError();
  if (e != hipSuccess) return e;
  k_gather_many<<<(unsigned)nblk_el, 256, 0, st>>>(tab, blk_el, n, vals_b);
  return hipGetLastError();
}

// Temporary storage both sorts need for n pairs (a query: hipcub allocates nothing without storage).
hipError_t sort_temp_bytes(int64_t n, size_t *bytes) {
  size_t a = 0, b = 0;
  hipError_t e = hipcub::DeviceRadixSort::SortPairs(nullptr, a, (const uint32_t *)nullptr, (uint32_t *)nullptr,
                                                    (const int32_t *)nullptr, (int32_t *)nullptr, (int)n, 0, 32);
  if (e != hipSuccess) return e;
  e = hipcub::DeviceRadixSort::SortPairs(nullptr, b, (const double *)nullptr, (double *)nullptr,
                                         (const int32_t *)nullptr, (int32_t *)nullptr, (int)n, 0, 64);
  *bytes = a > b ? a : b;
  return e;
}

template <int K, bool SPLIT, bool AUTO, bool RMU>
static hipError_t launch_pair_t(const KArgs *jobs, const Item *items, int n_items, int threads, size_t lds,
                                hipStream_t st) {
  auto fn = pair_kernel<K, SPLIT, AUTO, RMU>;
  hipError_t e = hipFuncSetAttribute((const void *)fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  fn<<<n_items, threads, lds, st>>>(jobs, items);
  return hipGetLastError();
}

// LDS bytes a pair kernel needs for `ncopy` histograms of `ncell` cells and `nslots` f64 slots.
size_t pair_lds_bytes(int ncell, int nslots, int n_rp, int ncopy) {
  return (size_t)((ncopy * ncell * 4 + 15) & ~15) + (size_t)ncopy * nslots * ncell * 8 + (size_t)(n_rp + 1) * 8;
}

hipError_t launch_pair(int kind, bool split, bool autop, const KArgs *jobs, const Item *items, int n_items,
                       size_t lds, int threads, hipStream_t st) {
  if (n_items == 0) return hipSuccess;
  const bool rmu = kind >= kRmuKind;
  if (rmu) kind -= kRmuKind;
#define CPC_DISPATCH(KK)                                                                                  \
  if (kind == KK) {                                                                                       \
    if (rmu && autop) return launch_pair_t<KK, false, true, true>(jobs, items, n_items, threads, lds, st);  \
    if (rmu) return launch_pair_t<KK, false, false, true>(jobs, items, n_items, threads, lds, st);          \
    if (split && autop) return launch_pair_t<KK, true, true, false>(jobs, items, n_items, threads, lds, st); \
    if (split) return launch_pair_t<KK, true, false, false>(jobs, items, n_items, threads, lds, st);        \
    if (autop) return launch_pair_t<KK, false, true, false>(jobs, items, n_items, threads, lds, st);        \
    return launch_pair_t<KK, false, false, false>(jobs, items, n_items, threads, lds, st);                  \
  }
  CPC_DISPATCH(kDD)
  CPC_DISPATCH(kSD)
  CPC_DISPATCH(kSS0)
  CPC_DISPATCH(kSS1)
  CPC_DISPATCH(kDK)
  CPC_DISPATCH(kSK)
  CPC_DISPATCH(kED)
#undef CPC_DISPATCH
  return hipErrorInvalidValue;
}

}  // namespace cpc
