// cpc_internal.h -- layouts shared by the host orchestration (cpc_api.cpp) and the HIP kernels
// (cpc_kernels.hip) of the rp-pi pair engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cpc {

// A galaxy as the pair kernels read it: one 64-byte record whose 16-byte quarters are {x, y},
// {w, e1}, {e2, phi} and {xl, yl, zf, meta}, so that four lanes fetch a galaxy with one coalesced
// load each.  Gathered into (cell, region, coin, z) order per pass; z and the catalog index live in
// side arrays of the same order.
struct alignas(64) PRec {
  double x, y;         // gal.ra.val_deg, gal.dec.val_deg (box: x, y)
  double w;            // gal.wt
  double e1;           // gal.e[0] (0 when the catalog has no shapes)
  double e2;           // gal.e[1]
  double phi;          // gal.phi
  float xl, yl;        // box: x, y relative to the galaxy's column corner (FP32)
  float zf;            // box: z (FP32)
  int32_t meta;        // box: image flags of this galaxy as an outer galaxy (kMeta*)
};
static_assert(sizeof(PRec) == 64, "PRec is four 16-byte quarters");

// Sorted record arrays carry this many padding records so that kernels may load a whole batch of
// records past the end of a range.
constexpr int kRecPad = 16;

// Image flags of an outer galaxy i, computed exactly as lim_calc_PB (PB_calcs.cpp:30-61) decides
// which periodic images of the inner catalog it visits (corr.cpp:380-423).
enum : int32_t {
  kMetaXm = 1,   // ra_min < 0  -> PBx = -1 visited
  kMetaXp = 2,   // ra_max > L  -> PBx = +1 visited
  kMetaYm = 4,
  kMetaYp = 8,
  kMetaZm = 16,
  kMetaZp = 32,  // z_max > L and the sweep from j = 0 does not stop at once (corr.cpp:389-408)
  kMetaZ0 = 64,  // PBz = 0 visited: always in a cross pass; in an auto pass when the scan from
                 // j = i+1 does not stop at once (corr.cpp:168-186)
};

// One run of galaxies sharing (column, jackknife region, coin); z-sorted inside.
struct alignas(8) Group {
  int32_t start, end;  // record range
  int32_t region, coin;
  int32_t col, pad;    // column index cy*nc + cx
  double zlo, zhi;     // z of the first and last record
};

// One wave's block galaxies: `count` <= 64 consecutive records of one block-side group from `first`.
struct alignas(16) Chunk {
  int32_t group, first, count, pad;
};

// A unit of work (one workgroup): chunks [begin, begin + n) of its job's chunk table, all from
// groups of the same jackknife region and coin, so that they share one histogram target; group0 is
// one of those groups.  `job` selects the pass arguments (KArgs) in the launch's job table.
// Items also select the tile groups they meet (`tmode`, `treg`, see TileFilter), so that every
// pair an item counts belongs to its one target.
struct alignas(16) Item {
  int32_t begin, n, target, group0;
  int32_t job, tmode, treg, pad;
};

// Tile-group filter of a box item whose block galaxies lie in region a:
enum TileFilter : int32_t {
  kTileAll = 0,         // every tile group (no jackknife)
  kTileSame = 1,        // region a
  kTileOtherC0 = 2,     // another region, coin < 0.5
  kTileSameOrC0 = 3,    // region a, or coin < 0.5
  kTileRegionC1 = 4,    // region treg (!= a), coin >= 0.5
};

// Histogram slots per pair kind: every target accumulates [wt, vals..., errs...].  Quantity codes:
// 0 = wt, 1..6 = data[f].val, 7..12 = data[f-6].err.
constexpr int kMaxSlots = 13;

// Arguments of one pass of one job.  A launch carries a device table of them and every work item
// names its entry, so one launch serves the items of many jobs (catalogs, boxes, binnings).
struct KArgs {
  const PRec *brec;    // block-side records and their z
  const double *bz;
  const PRec *trec;    // tile-side records and their z
  const double *tz;
  double L, pmin, pmax, psize, rmax;
  // z window of an outer galaxy: [z + wlo, z + whi] = pi range (rp-pi) or +- rmax (r-mu, coordinates 7)
  double wlo, whi;
  int rmu;  // r-mu bins: r = sqrt(rp*rp + pi*pi) against the thresholds, mu = pi / r in the pi bins
  double h;            // column width
  double zmargin;      // slack of the tile z-range search
  int nc;              // columns per side
  int n_rp, n_pi, ncell;
  // LDS histogram: cells from row0 * n_pi on (the rp bins that hold nearly all pairs), in ncopy
  // copies (lane l accumulates into copy l % ncopy); cells below go to the global histogram directly
  int row0, ncopy;
  int gws, n_nbr;
  int nslots;          // f64 slots accumulated by this pass
  int auto_mode;       // corels1 pass: 1 (sorted) pairs through PBz <= 0 images only with the inner
                       // galaxy after the outer one (z, then catalog index); 2 (unsorted) through every
                       // image only with the inner galaxy after the outer one in catalog order
  const int32_t *bidx, *tidx;  // catalog index of the block-side and tile-side records
  int lin;
  // split-precision classification
  float pioff32, inv_psize32, gz;  // pi bin: u = dz*inv_psize + pioff (pioff = -pmin/psize), guard gz in bin units
  float la, lb;                    // rp bin estimate from log2(s) (log bins) ...
  float lrmin, linv;               // ... or from sqrt(s) (linear bins)
  // tables (device)
  const double *thr;   // n_rp+1 FP64 thresholds on s
  float s_out_lo, s_out_hi;  // FP32 s below s_out_lo or from s_out_hi on is outside every rp bin
  float s_cert_lo;           // FP32 s below this is never classified in FP32
  float gA, gB, gC, gD;      // guard band of the rp bin index: g = gA + (gB + gC/r)/r + gD r, r = sqrt(s)
  // tile cull (both precisions): a tile whose FP32 column-relative position lies farther than
  // sqrt(cull_r2) from the bounding box of a wave's block galaxies, after taking cull_d off each axis
  // distance, cannot pair with any of them; cull_d bounds the FP32 error of those distances
  float cull_r2, cull_d;
  const int2 *nbr;     // neighbour column offsets
  const Chunk *chunks;
  const Group *bgrp;   // groups of the block side
  const Group *tgrp;   // groups of the tile side
  const int *tcol;     // [nc*nc+1] first group of each tile-side column
  double *acc;         // [n_target][nslots][ncell]
  unsigned long long *accnum;  // [n_target][ncell]
  unsigned long long *counters;  // [0] candidates, [1] FP64 re-evaluations
};


// Sky geometry: per-galaxy quantities the pair test needs besides PRec, all computed on the host
// with the reference's libm so that the device reproduces its sums term by term.
struct alignas(64) SRec {
  double dc, da;          // gal.DC, gal.DA
  double sd, cd;          // sin and cos of gal.dec.val_rad
  double ra_rad, dec_rad; // gal.ra.val_rad, gal.dec.val_rad (= degrees / r2d)
  double ux, uy;          // cd cos(ra), cd sin(ra): with sd, the unit vector of the direction
};

// The window lim_calc_sky gives an outer galaxy (sky_calcs.cpp:105-190).
struct alignas(64) Win {
  double zmin, zmax, ramin, ramax, decmin, decmax;
  double pad0, pad1;
};

struct SkyArgs {
  double pmin, pmax, psize;
  double inv_psize, pguard;  // split kernels: pi bin from (dP - pmin) * inv_psize, certain pguard from an edge
  int n_rp, n_pi, ncell, gws, n_items, nslots, lin;
  // cell grid over RA/Dec (degrees)
  double ra0, dec0, hra, hdec;
  int nra, ndec;
  double dsum;             // bound on the device/host difference of the great-circle cosine sum
  // the reference's pi and rp of a sky pair (choose_rp_pi_func): pim 0 DC_j - DC_i, 1 zero, 2 ephi of
  // the g_with_shape galaxy; rpm 0 dA (DA_i + DA_j)/2, 1 dA DA_i, 2 dA in degrees; rmu bins r and mu
  int pim, rpm, rmu;
  double dphi;             // bound on the device/host difference of ephi
  // split classification (log rp bins): rp^2 from the chord of the unit vectors, its FP32 log bin
  // estimate kf = log2(s32) la + lb, certain when frac(kf) is at least g from 0 and 1; s32 below
  // s_out_lo or from s_out_hi on is outside every bin; split_on = 0 sends every candidate to the
  // reference's formula
  int split_on;
  float la, lb, g, s_out_lo, s_out_hi;
  int auto_mode;           // corels1 pass: pairs only with the inner galaxy after the outer one, in
                           // (z, catalog index) order (1) or catalog order (2, unsorted)
  int row0, ncopy;         // LDS holds rp rows row0.. in ncopy lane-selected copies (7+ slots per cell)
  const double *thr;       // n_rp+1 thresholds on rp (host libm log10)
  const Item *items;
  const Chunk *chunks;
  const Group *bgrp;
  const Group *tgrp;
  const int *tcol;
  double *acc;
  unsigned long long *accnum;
  unsigned long long *counters;  // [0] candidates, [1] FP64 re-evaluations, [3] pairs for the host
  uint2 *hq;               // pairs (outer index, inner index) the host bins with its libm
  int hq_cap;
};

// Arguments of the cell-list construction kernels.
struct KeyArgs {
  const double *x, *y, *inc;
  const int32_t *jk;
  const double *jkp;
  int64_t n;
  int ncx, ncy;      // cell grid: cx = floor((x - x0)/hx) clamped to [0, ncx)
  double x0, y0, hx, hy;
  int nreg;          // regions in the key (1 when the jackknife is off)
  double inc_max;
};

struct GatherArgs {
  const double *x, *y, *z, *w, *e1, *e2, *phi;
  int64_t n;  // included galaxies
  int ncx, ncy;
  double x0, y0, hx, hy;
  // image flags (lim_calc_PB, PB_calcs.cpp:30-61)
  double L, pmin, pmax, rmax;
  double zfirst_plus_L;  // fl(z_first + L) of the inner catalog's first included galaxy
  int has_zfirst;
  // auto pass: per catalog galaxy, bit 0 = the PBz = 0 scan from j = i+1 does not stop at once,
  // bit 1 = nor does the PBz = -1 scan (nullptr in a cross pass)
  const uint8_t *aflags;
};

struct SparseArgs {
  const double *zo, *inco;  // outer z, include_prob
  int64_t no;
  const double *zi;         // inner z, ascending, every galaxy
  int64_t ni;
  double L, pmin, pmax, inc_max;
  unsigned long long *count;  // outer galaxies found
};

// Sums of raw pass targets into output histograms: output o is the sum of targets
// lst[beg[o] .. beg[o+1]) in list order, for the counts and every f64 slot.
struct ReduceArgs {
  const double *acc;                // raw targets [n_target][ns][ncell]
  const unsigned long long *accn;   // raw counts [n_target][ncell]
  double *out;                      // [n_out][ns][ncell]
  unsigned long long *outn;         // [n_out][ncell]
  const int32_t *beg, *lst;
  int n_out, ns, ncell;
};

// One side of a batched cell-list build (every eligible side of a call is sorted in one launch):
// its galaxies occupy [off, off + ka.n) of the concatenated key and index arrays, sort key
// (entry << 32 | cell key) with entry its place in the table, and its group boundaries occupy
// [boff, boff + nkeys] of the bounds array (as k_group_bounds writes them, positions relative to
// the side).  The records of its ga.n = ka.n galaxies go to rec, zs, idx (and, for a sky side,
// src_sr / src_win gathered into sr / win; src_sr null otherwise).
struct SideBuild {
  KeyArgs ka;
  GatherArgs ga;
  PRec *rec;
  double *zs;
  int32_t *idx;
  const SRec *src_sr;
  const Win *src_win;
  SRec *sr;
  Win *win;
  int64_t off, boff;
  uint32_t nkeys, pad;
};

// Launch helpers implemented in cpc_kernels.hip.
hipError_t launch_keys(const KeyArgs &A, const int32_t *order, uint32_t *keys, int32_t *vals, hipStream_t st);
hipError_t launch_gather(const GatherArgs &A, const int32_t *sidx, PRec *rec, double *zs, int32_t *idx,
                         hipStream_t st);
// Many reductions (sparse-window counts) in one launch: entry e of the device table `tab` is served
// by blocks [off[e], off[e+1]) of 256 threads, off[e+1] - off[e] = reduce_blocks (sparse_blocks) of
// it; `blocks` = off[n].
int64_t reduce_blocks(const ReduceArgs &A);
int64_t sparse_blocks(const SparseArgs &A);
hipError_t launch_reduce_many(const ReduceArgs *tab, const int64_t *off, int n, int64_t blocks, hipStream_t st);
hipError_t launch_sparse_many(const SparseArgs *tab, const int64_t *off, int n, int64_t blocks, hipStream_t st);
hipError_t launch_group_bounds(const uint32_t *keys, const int32_t *vals, int64_t n, uint32_t nkeys, int4 *out,
                               hipStream_t st);
hipError_t launch_fp32_error(uint32_t b0, uint32_t n, unsigned long long *out, hipStream_t st);
hipError_t launch_reduce(const ReduceArgs &A, hipStream_t st);
// The batched build of the n sides of the device table `tab`: keys, one stable radix sort of all
// `total` galaxies (entries above cell keys, so sides stay apart and each keeps its z order),
// group boundaries and records.  blk_el / blk_key: block offsets of each entry for the per-galaxy
// and per-key kernels (256 threads a block).  Grows `tmp` as the sort needs.
hipError_t build_sides_many(const SideBuild *tab, int n, const int64_t *blk_el, int64_t nblk_el,
                            const int64_t *blk_key, int64_t nblk_key, uint64_t *keys_a, uint64_t *keys_b,
                            int32_t *vals_a, int32_t *vals_b, int64_t total, int4 *bounds, void *&tmp,
                            size_t &tmp_bytes, hipStream_t st);
hipError_t sort_u32(void *&tmp, size_t &tmp_bytes, const uint32_t *kin, uint32_t *kout, const int32_t *vin,
                    int32_t *vout, int64_t n, hipStream_t st, bool may_alloc);
hipError_t sort_f64(void *&tmp, size_t &tmp_bytes, const double *kin, double *kout, const int32_t *vin,
                    int32_t *vout, int64_t n, hipStream_t st, bool may_alloc);
hipError_t sort_temp_bytes(int64_t n, size_t *bytes);
size_t pair_lds_bytes(int ncell, int nslots, int n_rp, int ncopy = 1);
// One launch over `n_items` items whose KArgs live in the device table `jobs` (all with the same
// kind, precision and pass direction); `lds` covers the largest histogram among them.
// kind + kRmuKind selects the r-mu instantiation (FP64 only) of the kind.
constexpr int kRmuKind = 16;
hipError_t launch_pair(int kind, bool split, bool autop, const KArgs *jobs, const Item *items, int n_items,
                       size_t lds, int threads, hipStream_t st);
// Record arrays of one side of a sky pass.
struct SkySide {
  const PRec *rec;
  const double *z;
  const int32_t *idx;
  const SRec *sr;
  const Win *win;  // outer side only
};
hipError_t launch_sky(int kind, bool split, const SkyArgs &P, int threads, const SkySide &B,
                      const SkySide &T, hipStream_t st);
// y = cos(x) (fn 0) or acos(x) (fn 1) with the device functions of the sky kernels
hipError_t launch_libm(int fn, const double *x, int64_t n, double *y, hipStream_t st);
// ulps by which the device cos of the sky kernels may differ from the host's (cpc_libm_error checks
// it on the device); the great-circle sum bound dsum follows from it
constexpr double kSkyCosUlp = 2.0;
hipError_t launch_gather_sky(const int32_t *sidx, int64_t n, const SRec *src_sr, const Win *src_win, SRec *sr,
                             Win *win, hipStream_t st);

}  // namespace cpc
