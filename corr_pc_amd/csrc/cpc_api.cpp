// cpc_api.cpp -- host side of the C ABI declared in include/cpc.h.
//
// Responsibilities: argument validation with the reference's preconditions, the data_info-derived
// binning quantities (initialization.cpp:305-311), the rp thresholds calibrated against the host
// libm, the per-galaxy sky windows (lim_calc_sky), cell-list orchestration on the device, work
// lists, the pair-kernel launches, the host binning of the few sky pairs the device cannot decide
// bit-exactly, and the final assembly of the reference's histogram layout.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <string>
#include <vector>
#include <functional>
#include <limits>
#include <type_traits>
#include <thread>
#include <atomic>

#include "../../include/cpc.h"
#include "cpc_internal.h"

using namespace cpc;

namespace {

const char *kVersion = "corr_pc_amd 0.2 (gfx950)";
const double kR2D = 180.0 / M_PI;  // r2d of calcs.cpp:19

// Device memory of one cpc_count_noalloc call: consecutive 256-byte aligned pieces of the library's
// static scratch array, handed out in order and all returned when the call ends.
struct Arena {
  char *base = nullptr;
  size_t cap = 0, off = 0;
  bool overflow = false;
  void *take(size_t bytes) {
    const size_t at = (off + 255) & ~(size_t)255;
    if (at + bytes > cap) {
      overflow = true;
      return nullptr;
    }
    off = at + bytes;
    return base + at;
  }
};
// Set while a cpc_count_noalloc call runs on this thread: device buffers then come from it.
thread_local Arena *t_arena = nullptr;

template <typename T>
struct DevBuf {
  T *p = nullptr;
  size_t cap = 0;         // elements
  bool borrowed = false;  // p lies in the no-alloc scratch (never freed)
  hipError_t ensure(size_t n) {
    if (n <= cap && p) return hipSuccess;
    const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (t_arena) {
      // a no-alloc call only ever grows buffers that it borrowed itself
      if (p && !borrowed) return hipErrorInvalidValue;
      void *q = t_arena->take(bytes);
      if (!q) return hipErrorOutOfMemory;
      p = (T *)q;
      cap = std::max<size_t>(n, 1);
      borrowed = true;
      return hipSuccess;
    }
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    borrowed = false;
    hipError_t e = hipMalloc((void **)&p, bytes);
    if (e == hipSuccess) cap = std::max<size_t>(n, 1);
    return e;
  }
  void release() {
    if (p && !borrowed) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    borrowed = false;
  }
};

// Page-locked host memory kept by the context between calls: copies to and from it are DMA
// transfers that the host does not wait for.  Not used by no-alloc calls.
struct HostBuf {
  char *p = nullptr;
  size_t cap = 0;  // bytes
  hipError_t ensure(size_t bytes) {
    if (bytes <= cap && p) return hipSuccess;
    release();
    hipError_t e = hipHostMalloc((void **)&p, std::max<size_t>(bytes, 1), hipHostMallocDefault);
    if (e == hipSuccess)
      cap = std::max<size_t>(bytes, 1);
    else
      p = nullptr;
    return e;
  }
  void release() {
    if (p) (void)hipHostFree(p);
    p = nullptr;
    cap = 0;
  }
};

// fn(i) for i in [0, n) on up to `nthreads` host threads (the calling one among them)
template <typename Fn>
void parallel_for(int64_t n, int nthreads, Fn fn) {
  nthreads = (int)std::min<int64_t>(std::max(1, nthreads), n);
  if (nthreads <= 1) {
    for (int64_t i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int64_t> next{0};
  auto work = [&] {
    for (int64_t i = next++; i < n; i = next++) fn(i);
  };
  std::vector<std::thread> pool;
  for (int t = 1; t < nthreads; ++t) pool.emplace_back(work);
  work();
  for (auto &th : pool) th.join();
}

int host_threads() { return (int)std::min(16u, std::max(1u, std::thread::hardware_concurrency())); }

struct Status {
  cpc_status code = CPC_OK;
  std::string msg;
};

// A 2-D cell grid: box columns (x, y) or sky cells (RA, Dec).
struct Grid {
  int ncx = 1, ncy = 1;
  double x0 = 0, y0 = 0, hx = 1, hy = 1;
};

// One catalog ordered for a pass: records in (cell, region, coin, z) order plus group tables.
struct Side {
  int64_t n_incl = 0;
  DevBuf<PRec> rec;
  DevBuf<double> zs;    // z in record order
  DevBuf<int32_t> idx;  // catalog index in record order
  DevBuf<SRec> sr;
  DevBuf<Win> win;
  DevBuf<Group> d_groups;
  DevBuf<int> d_colbeg;
  std::vector<Group> groups;
  std::vector<int> colbeg;        // [ncells+1]
  std::vector<int64_t> colcount;  // galaxies per cell
  // group boundaries found on the device, read back after every side of a call has been launched
  DevBuf<int4> d_bounds;
  std::vector<int4> h_bounds;  // host copy when no page-locked buffer is used
  const int4 *hb = nullptr;    // the host copy read by finish_side
  bool pending = false;
  uint64_t nkeys = 0;
  int nreg = 1;
  const cpc_catalog *cat = nullptr;
  void release() {
    d_bounds.release();
    rec.release();
    zs.release();
    idx.release();
    sr.release();
    win.release();
    d_groups.release();
    d_colbeg.release();
  }
};

SkySide sky_side(const Side &S) { return SkySide{S.rec.p, S.zs.p, S.idx.p, S.sr.p, S.win.p}; }

double ms_since(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// ---------------------------------------------------------------------------------------------
// rp bin thresholds.  The reference's rp bin index (corels.cpp:31-34) kept as a double so that
// out-of-range values stay ordered; as a function of the squared separation s (box: rp_calc_PB's
// sqrt, PB_calcs.cpp:97) or of rp itself (sky: rp_calc_sky, sky_calcs.cpp:91-96).

double bin_value(double v, bool squared, double rmin, double rsize, int lin) {
  const double rp = squared ? std::sqrt(v) : v;
  if (lin) return std::floor((rp - rmin) / rsize);
  return std::floor(std::log10(rp / rmin) / rsize);
}

uint64_t dbits(double d) {
  uint64_t u;
  std::memcpy(&u, &d, 8);
  return u;
}
double from_bits(uint64_t u) {
  double d;
  std::memcpy(&d, &u, 8);
  return d;
}

// Smallest non-negative double v with bin_value(v) >= k, by bisection over the ordered bit patterns
// of non-negative doubles; then a scan of `window` neighbours on each side confirms that the index
// is monotone around the edge (a composition of correctly rounded operations and the host log10).
bool threshold(bool squared, double rmin, double rsize, int lin, int k, int window, double *out) {
  uint64_t lo = 0, hi = dbits(INFINITY);
  while (lo < hi) {
    uint64_t mid = lo + (hi - lo) / 2;
    if (bin_value(from_bits(mid), squared, rmin, rsize, lin) >= k)
      hi = mid;
    else
      lo = mid + 1;
  }
  for (int d = 1; d <= window; ++d) {
    if (lo + d <= dbits(INFINITY) && !(bin_value(from_bits(lo + d), squared, rmin, rsize, lin) >= k)) return false;
    if (lo >= (uint64_t)d && (bin_value(from_bits(lo - d), squared, rmin, rsize, lin) >= k)) return false;
  }
  *out = from_bits(lo);
  return true;
}

struct Derived {
  double psize, rsize;
};

// z window of an outer galaxy, [z + lo, z + hi]: the pi range in rp-pi, +- rp_max in r-mu
// (lim_calc_xyz, PB_calcs.cpp:14-23).
double window_lo(const cpc_params &p) { return p.r_mu ? -p.rp_max : p.pi_min; }
double window_hi(const cpc_params &p) { return p.r_mu ? p.rp_max : p.pi_max; }

// The sky's pi and rp formulas (choose_rp_pi_func, initialization.cpp:545-571): pi 0 pi_calc, 1
// pi_calc_projected (0), 2 pi_calc_phi (ephi); rp 0 rp_calc_sky, 1 rp_calc_sky_projected, 2
// rp_calc_sky_angular; coordinates 1 bins r and mu (r_mu_calc).
struct SkyMode {
  int pim, rpm;
  bool rmu;
};
SkyMode sky_mode(const cpc_params &p) {
  const int c = p.coordinates;
  const bool proj = p.pi_max > 600 || c == 3 || c == 5;
  return {(c == 2 || c == 4) ? 2 : (proj ? 1 : 0), (c == 4 || c == 5) ? 2 : (proj ? 1 : 0), c == 1};
}

// shape_shape_corel's estimator-0 branch (corels.cpp:139): estimator 0, or the sky's coordinates 2 and 4
bool ss_shape_side_only(const cpc_params &p) {
  return p.estimator == 0 || (p.geometry == CPC_GEOM_SKY && (p.coordinates == 2 || p.coordinates == 4));
}

Derived derive(const cpc_params &p) {
  Derived d;
  d.psize = (p.pi_max - p.pi_min) / (p.n_pi);  // initialization.cpp:305
  if (p.geometry == CPC_GEOM_SKY && (p.coordinates == 3 || p.coordinates == 5)) d.psize = 1;  // :307-308
  if (p.lin_bin)
    d.rsize = (p.rp_max - p.rp_min) / p.n_rp;  // initialization.cpp:310
  else
    d.rsize = (std::log10(p.rp_max) - std::log10(p.rp_min)) / p.n_rp;  // initialization.cpp:311
  return d;
}

Status thresholds(const cpc_params &p, bool squared, int window, std::vector<double> &thr) {
  Derived d = derive(p);
  thr.assign(p.n_rp + 1, 0.0);
  for (int k = 0; k <= p.n_rp; ++k) {
    if (!threshold(squared, p.rp_min, d.rsize, p.lin_bin, k, window, &thr[k]))
      return {CPC_ERR_UNSUPPORTED,
              "rp bin index is not monotone near edge " + std::to_string(k) + " with this host libm"};
  }
  return {};
}

float round_up_f(double v) {
  float f = (float)v;
  if ((double)f < v) f = std::nextafter(f, INFINITY);
  return f;
}
float round_down_f(double v) {
  float f = (float)v;
  if ((double)f > v) f = std::nextafter(f, -INFINITY);
  return f;
}

// Guard band of the split-precision rp classification.  The kernel computes, in FP32, s32 (the
// squared separation from column-local coordinates, |s32 - s| <= E(s) with
// E(s) = 2 (2 sqrt2 eps r + 2 eps^2 + s 2^-21), r = sqrt(s), eps the bound on the FP32 error of dx
// and dy) and the bin index estimate kf (log2 of s32 through v_log_f32, or sqrt of s32 through
// v_rsq_f32).  A pair is binned in FP32 only when the fractional part of kf is at least g from 0
// and 1, g = gA + (gB + gC/r)/r + gD r bounding |kf - f(s)| with f the real-valued bin index whose
// integer crossings are the FP64 thresholds; other pairs (and s32 below s_cert_lo, where the
// bound's linearisation does not hold) go to the FP64 path.
struct Guard {
  float s_out_lo, s_out_hi, s_cert_lo, gA, gB, gC, gD;
};

Guard rp_guard(const cpc_params &p, const Derived &dv, const std::vector<double> &thr, double eps) {
  Guard g{};
  auto E = [&](double s) { return 2.0 * (2 * std::sqrt(2.0) * std::sqrt(s) * eps + 2 * eps * eps + std::ldexp(s, -21)); };
  g.s_out_lo = round_down_f(thr[0] - E(thr[0]));
  g.s_out_hi = round_up_f(thr[p.n_rp] + E(thr[p.n_rp]));
  // below r = 300 eps the relative error E(s)/s exceeds 3%: never certain there
  g.s_cert_lo = round_up_f(std::pow(300.0 * eps, 2));
  const double t24 = std::ldexp(1.0, -24), t21 = std::ldexp(1.0, -21), t20 = std::ldexp(1.0, -20);
  const double margin = 2.0;  // s32 in place of s in the bound, and slack
  const double shi = std::max((double)g.s_out_hi, 1e-30);
  double a, b, c, d;
  if (p.lin_bin) {
    // f(s) = (sqrt(s) - rmin)/rsize:  |f(s32) - f(s)| <= E(s)/(r rsize)
    const double k = margin / dv.rsize;
    a = k * 2 * 2 * std::sqrt(2.0) * eps;
    b = k * 2 * 2 * eps * eps;
    c = 0.0;
    d = k * 2 * t21;
    // evaluation of kf: sqrt through rsq (2^-21 relative), subtraction, product, constants
    const double rhi = std::sqrt(shi);
    a += margin * ((2 * rhi + std::fabs(p.rp_min)) * t20 / dv.rsize + (p.n_rp + 2) * t24 * 4);
  } else {
    // f(s) = 0.5 log10(s)/rsize - log10(rmin)/rsize:  |f(s32) - f(s)| <= 0.2172/rsize * 1.03 E(s)/s
    const double k = margin * 0.5 * std::log10(std::exp(1.0)) * 1.03 / dv.rsize;
    a = k * 2 * t21;
    b = k * 2 * 2 * std::sqrt(2.0) * eps;
    c = k * 2 * 2 * eps * eps;
    d = 0.0;
    // evaluation of kf = log2(s32) la + lb: v_log_f32 error (taken as 2^-21 max(1, |log2 s|)),
    // rounding of la, lb, the product and the sum
    const double la = 0.5 * std::log10(2.0) / dv.rsize, lb = -std::log10(p.rp_min) / dv.rsize;
    const double l2 = std::max({1.0, std::fabs(std::log2(std::max((double)g.s_out_lo, 1e-30))), std::fabs(std::log2(shi))});
    a += margin * (la * l2 * (t21 + 2 * t24) + std::fabs(lb) * t24 + (p.n_rp + 2) * t24 * 2);
  }
  a += 1e-12 * (p.n_rp + 1);  // the reference's own rounding of its bin index
  g.gA = round_up_f(a);
  g.gB = round_up_f(b);
  g.gC = round_up_f(c);
  g.gD = round_up_f(d);
  return g;
}

// Records the first failing HIP call of a step in `st`; true when `e` is an error (so that steps
// chain as `if (check(..) || check(..)) return st;` or run unconditionally and are tested once).
bool check(hipError_t e, Status &st, const char *what) {
  if (e != hipSuccess && st.code == CPC_OK) {
    st.code = (e == hipErrorOutOfMemory) ? CPC_ERR_NOMEM : CPC_ERR_DEVICE;
    st.msg = std::string(what) + ": " + hipGetErrorString(e);
  }
  return e != hipSuccess;
}

// ---------------------------------------------------------------------------------------------
// Sky helpers: the distance-table interpolation and per-galaxy window of lim_calc_sky
// (sky_calcs.cpp:33-81, 105-190), restated with the host libm.

struct SkyTable {
  const double *H, *dA;
  int64_t n;
  double dz, z_max, z_sep_max;
};

bool tab_ok(const SkyTable &T, double z) {
  const double f = std::floor(z / T.dz);
  return f >= 0 && f + 1 < (double)T.n;
}

double tab_interp(const double *tab, double dz, double z) {
  const int i_t = (int)std::floor(z / dz);
  return tab[i_t] + ((tab[i_t + 1] - tab[i_t]) * (z - dz * i_t) / dz);
}

// Window of an outer galaxy; false where the reference exits (z_min > z_max of the data,
// sky_calcs.cpp:109-113) or would read outside its table.  Coordinates 4 and 5 make no z test
// (corr.cpp:69): their window is unbounded in z.
bool sky_window(const SkyTable &T, const cpc_params &p, double ra, double dec, double z, Win &w) {
  const double rmax = p.rp_max;
  double z_max = z + T.z_sep_max, z_min, da_max;
  if (p.coordinates >= 2) {
    // cases 2-5 (sky_calcs.cpp:158-175): the RA/Dec reach at z_min = z, clamped; then the z window
    z_min = z;
    if (z_min < 0.001) z_min = 0.001;
    if (z_min > T.z_max) return false;
    if (p.coordinates >= 4) {
      da_max = rmax / kR2D;
    } else {
      if (!tab_ok(T, z_min)) return false;
      da_max = rmax / tab_interp(T.dA, T.dz, z_min);
    }
  } else {
    if (!tab_ok(T, z_max) || !tab_ok(T, z)) return false;
    if (p.coordinates == 1) {
      z_max = z + rmax * tab_interp(T.H, T.dz, z_max) / 300000.0;  // case 1: p_to_dz(bin_r_max, z + z_sep_max)
      z_min = z * 2 - z_max;
    } else {
      z_max = z + p.pi_max * tab_interp(T.H, T.dz, z_max) / 300000.0;  // p_to_dz(p_max, z + z_sep_max)
      z_min = z + p.pi_min * tab_interp(T.H, T.dz, z) / 300000.0;      // p_to_dz(p_min, z)
    }
    if (z_min < 0.001) z_min = 0.001;
    if (z_min > T.z_max || !tab_ok(T, z_min)) return false;
    da_max = rmax / tab_interp(T.dA, T.dz, z_min);  // dAR_to_dA(bin_r_max, z_min)
  }
  const double dec_rad = dec / kR2D;
  const double sd = std::sin(dec_rad), cd = std::cos(dec_rad);
  const double ra_lim = std::acos((std::cos(da_max) - sd * sd) / (cd * cd)) * 1.02 * kR2D;
  double ra_max = ra + ra_lim, ra_min = ra - ra_lim;
  if (ra_max > 360) ra_min = 0;
  if (ra_min < 0) ra_max = 360;
  if (p.coordinates >= 2) {
    z_min = p.z_sep_min == 0 ? p.z_min : z + p.z_sep_min;
    z_max = T.z_sep_max == 0 ? T.z_max : z + T.z_sep_max;
  }
  if (p.coordinates >= 4) {
    z_min = -INFINITY;
    z_max = INFINITY;
  }
  w.zmin = z_min;
  w.zmax = z_max;
  w.ramin = ra_min;
  w.ramax = ra_max;
  w.decmin = dec - da_max * kR2D * 1.05;
  w.decmax = dec + da_max * kR2D * 1.05;
  w.pad0 = w.pad1 = 0;
  return true;
}

}  // namespace

// ---------------------------------------------------------------------------------------------

struct cpc_catalog {
  int64_t n = 0;
  double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr, *d_w = nullptr, *d_e1 = nullptr, *d_e2 = nullptr,
         *d_phi = nullptr, *d_jkp = nullptr, *d_inc = nullptr, *d_e3 = nullptr;
  int32_t *d_jk = nullptr;
  double *d_zsorted = nullptr;  // every galaxy's z ascending (only when the upload order is not sorted)
  SRec *d_sr = nullptr;         // sky quantities per galaxy (upload order), when DC/DA were given
  bool z_sorted = true, has_e = false, has_jk = false, has_sky = false, has_e3 = false;
  double lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
  int32_t jk_lo = 0, jk_hi = -1;
  // host copies: z and include_prob (first-included-galaxy rule); for sky catalogs also what the
  // host needs to bin guard-band pairs
  std::vector<double> h_z, h_inc, h_x, h_y, h_w, h_e1, h_e2, h_phi, h_jkp, h_dc, h_da;
  std::vector<int32_t> h_jk;
};

// Device state of one prepared pass, kept in the context's pool (entry i serves job i of a call) and
// reused by later calls.
struct JobBufs {
  Side so, si;  // outer and inner side (si unused when both are one catalog)
  DevBuf<double> thr;
  DevBuf<int2> nbr;
  DevBuf<Chunk> chunks;
  DevBuf<uint8_t> aflags;
  DevBuf<Win> win_raw;
  DevBuf<uint2> hq;
  DevBuf<int32_t> red_beg, red_lst;  // target lists of the device-side jackknife reduction
  HostBuf win_host;                   // sky windows, page-locked for the copy (not in no-alloc calls)
  void release() {
    win_host.release();
    so.release();
    si.release();
    thr.release();
    nbr.release();
    chunks.release();
    aflags.release();
    win_raw.release();
    hq.release();
    red_beg.release();
    red_lst.release();
  }
};

struct cpc_ctx {
  int device = 0;
  char *scratch = nullptr;  // this device's copy of the no-alloc scratch array (hipGetSymbolAddress)
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, ev2 = nullptr;
  hipEvent_t evc0 = nullptr, evc1 = nullptr;  // around the batched cell-list build (ms_cells)
  bool cells_timed = false;
  std::string err;
  size_t max_lds = 65536;
  int n_cu = 256;
  void *sort_tmp = nullptr;
  size_t sort_tmp_bytes = 0;
  double col_frac = 0.0;
  int item_size = 0, block_threads = 0;
  // scratch of the cell-list construction (sides are built one after another)
  DevBuf<uint32_t> keys_a, keys_b;
  DevBuf<int32_t> vals_a, vals_b, order;
  DevBuf<double> zk_b;
  DevBuf<char> tables;  // every plan's small tables (stage_tables)
  DevBuf<char> aux;     // per call: the tables of the batched reduction and sparse-window launches
  HostBuf hb_bounds, hb_stage;  // page-locked group-boundary readback and table staging
  // batched cell-list build: the sides queued by setup and their concatenated sort arrays, group
  // boundaries and table (never used by a no-alloc call)
  std::vector<SideBuild> side_q;
  std::vector<Side *> side_q_side;
  int64_t side_q_total = 0;
  DevBuf<uint64_t> k64_a, k64_b;
  DevBuf<int32_t> sv_a, sv_b;
  DevBuf<int4> bounds_all;
  DevBuf<char> side_tab;
  // per call: work items of every launch, box pass arguments (two per job), and the accumulator
  // arenas (f64 histogram slots; counts and counters)
  DevBuf<Item> items;
  DevBuf<KArgs> jobs;
  DevBuf<double> acc;
  DevBuf<unsigned long long> accn;
  // raw pass targets that are reduced on the device before the download (box passes with jackknife)
  DevBuf<double> acc_raw;
  DevBuf<unsigned long long> accn_raw;
  std::vector<std::unique_ptr<JobBufs>> pool;
  std::map<std::tuple<double, double, int, int, int>, std::vector<double>> thr_cache;
  // every device buffer a pass uses, exchanged wholesale around a no-alloc call
  struct Swap {
    void *sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    DevBuf<uint32_t> keys_a, keys_b;
    DevBuf<int32_t> vals_a, vals_b, order;
    DevBuf<double> zk_b;
    DevBuf<char> tables, aux;
    DevBuf<Item> items;
    DevBuf<KArgs> jobs;
    DevBuf<double> acc, acc_raw;
    DevBuf<unsigned long long> accn, accn_raw;
    std::vector<std::unique_ptr<JobBufs>> pool;
    // borrowed buffers hold no memory of their own: forget them
    void drop() {
      sort_tmp = nullptr;
      sort_tmp_bytes = 0;
      keys_a = {};
      keys_b = {};
      vals_a = {};
      vals_b = {};
      order = {};
      zk_b = {};
      tables = {};
      aux = {};
      items = {};
      jobs = {};
      acc = {};
      acc_raw = {};
      accn = {};
      accn_raw = {};
      for (auto &b : pool) *b = JobBufs();
      pool.clear();
    }
  };
  void swap_buffers(Swap &o) {
    std::swap(sort_tmp, o.sort_tmp);
    std::swap(sort_tmp_bytes, o.sort_tmp_bytes);
    std::swap(keys_a, o.keys_a);
    std::swap(keys_b, o.keys_b);
    std::swap(vals_a, o.vals_a);
    std::swap(vals_b, o.vals_b);
    std::swap(order, o.order);
    std::swap(zk_b, o.zk_b);
    std::swap(tables, o.tables);
    std::swap(aux, o.aux);
    std::swap(items, o.items);
    std::swap(jobs, o.jobs);
    std::swap(acc, o.acc);
    std::swap(acc_raw, o.acc_raw);
    std::swap(accn, o.accn);
    std::swap(accn_raw, o.accn_raw);
    std::swap(pool, o.pool);
  }
};

namespace {

cpc_status fail(cpc_ctx *ctx, const Status &s) {
  if (ctx) ctx->err = s.msg;
  return s.code;
}
cpc_status fail(cpc_ctx *ctx, cpc_status code, const std::string &msg) { return fail(ctx, Status{code, msg}); }

const std::vector<double> *cached_thresholds(cpc_ctx *ctx, const cpc_params &p, bool squared, Status &st) {
  auto key = std::make_tuple(p.rp_min, p.rp_max, p.n_rp, p.lin_bin, squared ? 1 : 0);
  auto it = ctx->thr_cache.find(key);
  if (it == ctx->thr_cache.end()) {
    std::vector<double> thr;
    st = thresholds(p, squared, 256, thr);
    if (st.code) return nullptr;
    it = ctx->thr_cache.emplace(key, thr).first;
  }
  return &it->second;
}

// Build one side: keys, stable sort, gather, group tables.  `zfirst_plus_L` feeds the PBz>0 flag
// (box only).  For the sky, `win_raw` (per original galaxy, outer side) is gathered as well.
Status side_tables(cpc_ctx *ctx, Side &S, size_t ncells);

Status build_side(cpc_ctx *ctx, const cpc_catalog *c, const Grid &G, int nreg, bool jk_on, const cpc_params &p,
                  bool has_zfirst, double zfirst_plus_L, bool sky, const Win *win_raw, const uint8_t *aflags,
                  Side &S) {
  Status st;
  hipStream_t s = ctx->stream;
  const int64_t n = c->n;
  const size_t ncells = (size_t)G.ncx * G.ncy;
  S.groups.clear();
  S.colbeg.assign(ncells + 1, 0);
  S.colcount.assign(ncells, 0);
  S.n_incl = 0;
  S.pending = false;
  if (n == 0) {
    check(S.rec.ensure(kRecPad), st, "alloc");
    check(S.zs.ensure(kRecPad), st, "alloc");
    check(S.idx.ensure(kRecPad), st, "alloc");
    check(S.sr.ensure(1), st, "alloc");
    check(S.win.ensure(1), st, "alloc");
    return st;
  }
  if (check(ctx->keys_a.ensure(n), st, "alloc") || check(ctx->keys_b.ensure(n), st, "alloc") ||
      check(ctx->vals_a.ensure(n), st, "alloc") || check(ctx->vals_b.ensure(n), st, "alloc"))
    return st;
  const int32_t *order = nullptr;
  if (!c->z_sorted) {
    // galaxies into ascending z first, so that the stable key sort keeps z order inside a group
    if (check(ctx->order.ensure(n), st, "alloc") || check(ctx->zk_b.ensure(n), st, "alloc")) return st;
    std::vector<int32_t> iota(n);
    std::iota(iota.begin(), iota.end(), 0);
    check(hipMemcpyAsync(ctx->vals_a.p, iota.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, s), st, "copy");
    check(sort_f64(ctx->sort_tmp, ctx->sort_tmp_bytes, c->d_z, ctx->zk_b.p, ctx->vals_a.p, ctx->order.p, n, s,
                   t_arena == nullptr),
          st,
          "sort z");
    check(hipStreamSynchronize(s), st, "sync");
    if (st.code) return st;
    order = ctx->order.p;
  }
  KeyArgs ka{c->d_x, c->d_y, c->d_inc, c->d_jk, jk_on ? c->d_jkp : nullptr, n, G.ncx, G.ncy,
             G.x0, G.y0, G.hx, G.hy, nreg, p.include_prob_max};
  check(launch_keys(ka, order, ctx->keys_a.p, ctx->vals_a.p, s), st, "keys");
  check(sort_u32(ctx->sort_tmp, ctx->sort_tmp_bytes, ctx->keys_a.p, ctx->keys_b.p, ctx->vals_a.p, ctx->vals_b.p, n, s,
                 t_arena == nullptr),
        st, "sort keys");
  if (st.code) return st;
  // groups: runs of equal keys.  Found on the device by one binary search per possible key; the
  // boundaries are read back once every side of the call has been launched (finish_side), and the
  // records of every galaxy are gathered now (the excluded ones, last in key order, lie past the
  // groups and are never read).  With very many possible keys the sorted keys are scanned here.
  const uint64_t nkeys = (uint64_t)ncells * (uint64_t)nreg * 2u;
  int64_t ng = 0;  // records gathered
  S.pending = false;
  if (nkeys <= (1u << 23)) {
    if (check(S.d_bounds.ensure(nkeys + 1), st, "alloc")) return st;
    check(launch_group_bounds(ctx->keys_b.p, ctx->vals_b.p, n, (uint32_t)nkeys, S.d_bounds.p, s), st, "groups");
    S.pending = true;
    S.nkeys = nkeys;
    S.nreg = nreg;
    S.cat = c;
    S.hb = nullptr;
    ng = n;
  } else {
    std::vector<uint32_t> hkeys(n);
    std::vector<int32_t> hidx(n);
    check(hipMemcpyAsync(hkeys.data(), ctx->keys_b.p, n * sizeof(uint32_t), hipMemcpyDeviceToHost, s), st, "copy");
    check(hipMemcpyAsync(hidx.data(), ctx->vals_b.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, s), st, "copy");
    check(hipStreamSynchronize(s), st, "sync");
    if (st.code) return st;
    const int64_t ni = std::lower_bound(hkeys.begin(), hkeys.end(), 0xffffffffu) - hkeys.begin();
    for (int64_t a = 0; a < ni;) {
      int64_t b = a + 1;
      while (b < ni && hkeys[b] == hkeys[a]) ++b;
      const uint32_t key = hkeys[a];
      Group g{};
      g.start = (int32_t)a;
      g.end = (int32_t)b;
      g.coin = key & 1u;
      g.region = (int32_t)((key >> 1) % (uint32_t)nreg);
      g.col = (int32_t)((key >> 1) / (uint32_t)nreg);
      g.zlo = c->h_z[hidx[a]];
      g.zhi = c->h_z[hidx[b - 1]];
      S.groups.push_back(g);
      S.colcount[g.col] += b - a;
      a = b;
    }
    S.n_incl = ni;
    ng = ni;
  }
  if (check(S.rec.ensure(ng + kRecPad), st, "alloc") || check(S.zs.ensure(ng + kRecPad), st, "alloc") ||
      check(S.idx.ensure(ng + kRecPad), st, "alloc"))
    return st;
  // ED passes carry the third axis component in the record's phi slot (ED never reads phi)
  const double *phi_src = p.kind == CPC_KIND_ED ? c->d_e3 : c->d_phi;
  GatherArgs ga{c->d_x, c->d_y, c->d_z, c->d_w, c->d_e1, c->d_e2, phi_src, ng, G.ncx, G.ncy, G.x0, G.y0, G.hx, G.hy,
                p.box_size, window_lo(p), window_hi(p), p.rp_max, zfirst_plus_L, has_zfirst ? 1 : 0, aflags};
  check(launch_gather(ga, ctx->vals_b.p, S.rec.p, S.zs.p, S.idx.p, s), st, "gather");
  if (sky) {
    if (check(S.sr.ensure(std::max<int64_t>(ng, 1)), st, "alloc") || check(S.win.ensure(std::max<int64_t>(ng, 1)), st, "alloc"))
      return st;
    check(launch_gather_sky(ctx->vals_b.p, ng, c->d_sr, win_raw, S.sr.p, S.win.p, s), st, "gather sky");
  }
  if (!S.pending) st = side_tables(ctx, S, ncells);
  return st;
}

// The group table of a side whose boundaries came back from the device (build_side).
Status finish_side(cpc_ctx *ctx, Side &S, size_t ncells) {
  if (!S.pending) return {};
  S.pending = false;
  const int4 *hb = S.hb;
  const cpc_catalog *c = S.cat;
  for (uint64_t key = 0; key < S.nkeys; ++key) {
    const int4 q = hb[key];
    if (q.x >= q.y) continue;
    Group g{};
    g.start = q.x;
    g.end = q.y;
    g.coin = (int32_t)(key & 1u);
    g.region = (int32_t)((key >> 1) % (uint64_t)S.nreg);
    g.col = (int32_t)((key >> 1) / (uint64_t)S.nreg);
    g.zlo = c->h_z[q.z];
    g.zhi = c->h_z[q.w];
    S.groups.push_back(g);
    S.colcount[g.col] += q.y - q.x;
  }
  S.n_incl = hb[S.nkeys].x;
  return side_tables(ctx, S, ncells);
}

// A side of the call: queued for the batched build (build_queued_sides) when its catalog is in z
// order and the call may allocate, else built now (build_side).
Status place_side(cpc_ctx *ctx, const cpc_catalog *c, const Grid &G, int nreg, bool jk_on, const cpc_params &p,
                  bool has_zfirst, double zfirst_plus_L, bool sky, const Win *win_raw, const uint8_t *aflags,
                  Side &S) {
  const int64_t n = c->n;
  const size_t ncells = (size_t)G.ncx * G.ncy;
  const uint64_t nkeys = (uint64_t)ncells * (uint64_t)nreg * 2u;
  const bool batched = !t_arena && n > 0 && c->z_sorted && nkeys <= (1u << 23) &&
                       ctx->side_q_total + n < (int64_t)INT32_MAX && ctx->side_q.size() < (1u << 20);
  if (!batched) return build_side(ctx, c, G, nreg, jk_on, p, has_zfirst, zfirst_plus_L, sky, win_raw, aflags, S);
  Status st;
  S.groups.clear();
  S.colbeg.assign(ncells + 1, 0);
  S.colcount.assign(ncells, 0);
  S.n_incl = 0;
  if (check(S.rec.ensure(n + kRecPad), st, "alloc") || check(S.zs.ensure(n + kRecPad), st, "alloc") ||
      check(S.idx.ensure(n + kRecPad), st, "alloc"))
    return st;
  if (sky && (check(S.sr.ensure(n), st, "alloc") || check(S.win.ensure(n), st, "alloc"))) return st;
  SideBuild b{};
  b.ka = KeyArgs{c->d_x, c->d_y, c->d_inc, c->d_jk, jk_on ? c->d_jkp : nullptr, n, G.ncx, G.ncy,
                 G.x0, G.y0, G.hx, G.hy, nreg, p.include_prob_max};
  // ED passes carry the third axis component in the record's phi slot (ED never reads phi)
  const double *phi_src = p.kind == CPC_KIND_ED ? c->d_e3 : c->d_phi;
  b.ga = GatherArgs{c->d_x, c->d_y, c->d_z, c->d_w, c->d_e1, c->d_e2, phi_src, n, G.ncx, G.ncy, G.x0, G.y0, G.hx, G.hy,
                    p.box_size, window_lo(p), window_hi(p), p.rp_max, zfirst_plus_L, has_zfirst ? 1 : 0, aflags};
  b.rec = S.rec.p;
  b.zs = S.zs.p;
  b.idx = S.idx.p;
  if (sky) {
    b.src_sr = c->d_sr;
    b.src_win = win_raw;
    b.sr = S.sr.p;
    b.win = S.win.p;
  }
  b.nkeys = (uint32_t)nkeys;
  ctx->side_q.push_back(b);
  ctx->side_q_side.push_back(&S);
  ctx->side_q_total += n;
  S.pending = true;
  S.nkeys = nkeys;
  S.nreg = nreg;
  S.cat = c;
  S.hb = nullptr;
  return st;
}

// Every queued side in one build: keys, one sort, group boundaries and records (build_sides_many),
// then the boundaries copied back behind it; each side's `hb` points at its part of the copy.
Status build_queued_sides(cpc_ctx *ctx) {
  Status st;
  std::vector<SideBuild> &q = ctx->side_q;
  if (q.empty()) return st;
  const int n = (int)q.size();
  std::vector<int64_t> blk_el{0}, blk_key{0};
  int64_t total = 0, nb = 0;
  for (SideBuild &b : q) {
    b.off = total;
    total += b.ka.n;
    b.boff = nb;
    nb += (int64_t)b.nkeys + 1;
    blk_el.push_back(blk_el.back() + (b.ka.n + 255) / 256);
    blk_key.push_back(blk_key.back() + ((int64_t)b.nkeys + 1 + 255) / 256);
  }
  const size_t o_el = q.size() * sizeof(SideBuild), o_key = o_el + blk_el.size() * sizeof(int64_t),
               bytes = o_key + blk_key.size() * sizeof(int64_t);
  static_assert(sizeof(SideBuild) % 8 == 0, "table of 8-byte aligned records");
  std::vector<char> h(bytes);
  std::memcpy(h.data(), q.data(), o_el);
  std::memcpy(h.data() + o_el, blk_el.data(), blk_el.size() * sizeof(int64_t));
  std::memcpy(h.data() + o_key, blk_key.data(), blk_key.size() * sizeof(int64_t));
  hipStream_t s = ctx->stream;
  if (check(ctx->k64_a.ensure(total), st, "alloc") || check(ctx->k64_b.ensure(total), st, "alloc") ||
      check(ctx->sv_a.ensure(total), st, "alloc") || check(ctx->sv_b.ensure(total), st, "alloc") ||
      check(ctx->bounds_all.ensure(nb), st, "alloc") || check(ctx->side_tab.ensure(bytes), st, "alloc"))
    return st;
  check(hipEventRecord(ctx->evc0, s), st, "event");
  check(hipMemcpyAsync(ctx->side_tab.p, h.data(), bytes, hipMemcpyHostToDevice, s), st, "copy");
  const char *tab = ctx->side_tab.p;
  check(build_sides_many(reinterpret_cast<const SideBuild *>(tab), n, reinterpret_cast<const int64_t *>(tab + o_el),
                         blk_el.back(), reinterpret_cast<const int64_t *>(tab + o_key), blk_key.back(), ctx->k64_a.p,
                         ctx->k64_b.p, ctx->sv_a.p, ctx->sv_b.p, total, ctx->bounds_all.p, ctx->sort_tmp,
                         ctx->sort_tmp_bytes, s),
        st, "side build");
  if (st.code) return st;
  if (check(ctx->hb_bounds.ensure(nb * sizeof(int4)), st, "host alloc")) return st;
  const int4 *hb = reinterpret_cast<const int4 *>(ctx->hb_bounds.p);
  check(hipMemcpyAsync(ctx->hb_bounds.p, ctx->bounds_all.p, nb * sizeof(int4), hipMemcpyDeviceToHost, s), st, "copy");
  check(hipEventRecord(ctx->evc1, s), st, "event");
  ctx->cells_timed = !st.code;
  for (int k = 0; k < n; ++k) ctx->side_q_side[k]->hb = hb + q[k].boff;
  q.clear();
  ctx->side_q_side.clear();
  ctx->side_q_total = 0;
  return st;
}

// Column starts of a side's groups (host; the tables reach the device with the call's other small
// tables, stage_tables).
Status side_tables(cpc_ctx *, Side &S, size_t ncells) {
  size_t gi = 0;
  for (size_t col = 0; col < ncells; ++col) {
    S.colbeg[col] = (int)gi;
    while (gi < S.groups.size() && (size_t)S.groups[gi].col == col) ++gi;
  }
  S.colbeg[ncells] = (int)S.groups.size();
  return {};
}

// Descending order of cost in steps of one octave (the integer part of its binary logarithm), ties
// in input order: a stable counting sort over the steps the vector uses (the order run_box gives the
// items of several plans).  With `exact` the order of the costs themselves (a stable comparison
// sort).  Box passes run faster in steps, where ties keep target and column order (the 1M-galaxy w++
// pass: 574 ms in exact order, 564 in 1/16-octave steps, 557 in octaves, 595 in half octaves; the
// wgg box, the large box and the batched mocks gain from octaves as well); sky passes run faster in
// exact order (759 against 776 ms in 1/16-octave steps on the 2e6-galaxy survey).
int cost_step(double c) {
  const double l = c > 0 ? std::log2(c) : 0.0;
  return (int)std::min(4095.0, std::max(0.0, l));
}

template <typename T>
void sort_by_cost_desc(std::vector<std::pair<double, T>> &v, bool exact = false) {
  if (v.size() < 2) return;
  if (exact) {
    std::stable_sort(v.begin(), v.end(), [](const auto &x, const auto &y) { return x.first > y.first; });
    return;
  }
  std::vector<int> st(v.size());
  int lo = 4095, hi = 0;
  for (size_t i = 0; i < v.size(); ++i) {
    st[i] = cost_step(v[i].first);
    lo = std::min(lo, st[i]);
    hi = std::max(hi, st[i]);
  }
  std::vector<size_t> pos(hi - lo + 2, 0);  // pos[hi - s + 1]: items of steps above s
  for (int x : st) ++pos[hi - x + 1];
  for (size_t b = 1; b < pos.size(); ++b) pos[b] += pos[b - 1];
  std::vector<std::pair<double, T>> out(v.size());
  for (size_t i = 0; i < v.size(); ++i) out[pos[hi - st[i]]++] = v[i];
  v.swap(out);
}

// Work items of a pass, with the jackknife decomposed by the tile side so that every pair is
// evaluated once.  A pair of an outer galaxy (region a, coin ci) and an inner galaxy (region b, coin
// cj) touches region a when ci >= 0.5, or in the box also when b == a (`same_touches`: jk_check_PB
// against jk_check_sky, jk.cpp:53-71), and region b when b != a and cj >= 0.5 (corel_jk,
// jk.cpp:201-223).  Each item takes block chunks of one (a, ci) and the tile groups of one filter,
// which fixes both flags, and counts into one target (sky with ci = 0: region a and coin-0 tiles ->
// target 2a):
//   ci = 1: tiles of region a or coin 0      -> target 2a+1      (touches a only)
//           tiles of region b != a, coin 1   -> target Y(a, 1, b) (touches a and b)
//   ci = 0: tiles of region a                -> target 2a+1      (touches a only)
//           tiles of other regions, coin 0   -> target 2a        (touches nothing)
//           tiles of region b != a, coin 1   -> target Y(a, 0, b) (touches b only)
// Y targets follow the 2 nreg others in first-use order; `ytar` receives (a, x, b) of each, x the
// flag of touching a.  Without the jackknife every item takes every tile group into target 1.
// Chunks of one (target, filter) are ordered by cost (sort_by_cost_desc; `exact_order` for sky
// passes) and packed `per_item` at a time; items run most expensive first.  The cost of a chunk is its galaxies times the tile galaxies its filter
// takes in the neighbour columns.
template <typename ForNbr>
void make_items_jk(const Side &B, const Side &T, bool jk_on, bool same_touches, bool exact_order, int nreg,
                   int per_item, size_t ncols, ForNbr for_nbr, std::vector<Item> &items, std::vector<double> &costs,
                   std::vector<Chunk> &chunks, std::vector<int3> &ytar, int threads) {
  struct Key {
    int target, mode, reg;
    bool operator<(const Key &o) const { return std::tie(target, mode, reg) < std::tie(o.target, o.mode, o.reg); }
  };
  std::map<Key, std::vector<std::pair<double, Chunk>>> by_key;
  std::vector<int> yidx((size_t)nreg * 2 * nreg, -1);  // (a, x, b) -> Y target number
  ytar.clear();
  // tile galaxies in reach of each column, by tile (region, coin): [col][2 * region + coin], filled
  // on first use (block groups of one column share it)
  const int nr2 = 2 * nreg;
  std::vector<double> colsum;
  std::vector<char> have(ncols, 0);
  colsum.assign(ncols * nr2, 0.0);
  std::vector<std::pair<Key, double>> reach;
  for (size_t g = 0; g < B.groups.size(); ++g) {
    const Group &G = B.groups[g];
    const int a = G.region, ci = G.coin;
    double *tc = &colsum[(size_t)G.col * nr2];
    if (!have[G.col]) {
      have[G.col] = 1;
      for_nbr(G.col, [&](int col) {
        for (int tg = T.colbeg[col]; tg < T.colbeg[col + 1]; ++tg) {
          const Group &H = T.groups[tg];
          tc[2 * H.region + H.coin] += (double)(H.end - H.start);
        }
      });
    }
    reach.clear();
    if (!jk_on) {
      double c = 0;
      for (int k = 0; k < nr2; ++k) c += tc[k];
      if (c > 0) reach.push_back({Key{1, kTileAll, 0}, c});
    } else {
      // tiles that touch region a only: region a, and (ci = 1) other regions' coin-0 tiles
      double touch_a = tc[2 * a] + tc[2 * a + 1], other0 = 0;
      for (int b = 0; b < nreg; ++b)
        if (b != a) other0 += tc[2 * b];
      if (ci) {
        touch_a += other0;
        if (touch_a > 0) reach.push_back({Key{2 * a + 1, kTileSameOrC0, 0}, touch_a});
      } else if (!same_touches) {
        // jk_check_sky: with coin 0 the outer galaxy touches nothing, its own region included
        touch_a += other0;
        if (touch_a > 0) reach.push_back({Key{2 * a, kTileSameOrC0, 0}, touch_a});
      } else {
        if (touch_a > 0) reach.push_back({Key{2 * a + 1, kTileSame, 0}, touch_a});
        if (other0 > 0) reach.push_back({Key{2 * a, kTileOtherC0, 0}, other0});
      }
      for (int b = 0; b < nreg; ++b) {
        if (b == a || !(tc[2 * b + 1] > 0)) continue;
        int &y = yidx[((size_t)a * 2 + ci) * nreg + b];
        if (y < 0) {
          y = (int)ytar.size();
          ytar.push_back(make_int3(a, ci, b));
        }
        reach.push_back({Key{2 * nreg + y, kTileRegionC1, b}, tc[2 * b + 1]});
      }
    }
    for (const auto &kv : reach) {
      auto &v = by_key[kv.first];
      for (int f = G.start; f < G.end; f += 64) {
        const Chunk c{(int32_t)g, f, std::min(64, G.end - f), 0};
        v.push_back({(double)c.count * kv.second, c});
      }
    }
  }
  // each key's chunks ordered by cost and packed into items on host threads; then every item, most
  // expensive first
  std::vector<std::pair<const Key, std::vector<std::pair<double, Chunk>>> *> keys;
  for (auto &kv : by_key) keys.push_back(&kv);
  std::vector<std::vector<Chunk>> kchunks(keys.size());
  std::vector<std::vector<double>> kcost(keys.size());  // per item of the key
  parallel_for((int64_t)keys.size(), threads, [&](int64_t k) {
    auto &v = keys[k]->second;
    sort_by_cost_desc(v, exact_order);
    std::vector<Chunk> &cs = kchunks[k];
    cs.reserve(v.size());
    for (size_t a = 0; a < v.size(); a += per_item) {
      double c = 0.0;
      for (size_t b = a; b < std::min(v.size(), a + per_item); ++b) {
        c += v[b].first;
        cs.push_back(v[b].second);
      }
      kcost[k].push_back(c);
    }
  });
  std::vector<std::pair<double, int2>> order;  // (cost, (key, item of the key))
  for (size_t k = 0; k < keys.size(); ++k)
    for (size_t j = 0; j < kcost[k].size(); ++j) order.push_back({kcost[k][j], make_int2((int)k, (int)j)});
  sort_by_cost_desc(order, exact_order);
  items.clear();
  costs.clear();
  items.reserve(order.size());
  costs.reserve(order.size());
  for (const auto &o : order) {
    const int k = o.second.x, j = o.second.y;
    const std::vector<Chunk> &cs = kchunks[k];
    const size_t a = (size_t)j * per_item, e = std::min(cs.size(), a + per_item);
    const Key &key = keys[k]->first;
    items.push_back(Item{(int32_t)chunks.size(), (int32_t)(e - a), key.target, cs[a].group, 0, key.mode, key.reg, 0});
    costs.push_back(o.first);
    chunks.insert(chunks.end(), cs.begin() + a, cs.begin() + e);
  }
}

// Contributions binned on the host, in the reference layout: region 0 the full sample (all
// fields), regions 1..n_jk the touch sums (num, wt_num, data[].val).
struct HostHist {
  int R = 0, ncell = 0;
  std::vector<int64_t> num;
  std::vector<double> wt, val, err;
  void init(int R_, int ncell_) {
    R = R_;
    ncell = ncell_;
    num.assign((size_t)R * ncell, 0);
    wt.assign((size_t)R * ncell, 0.0);
    val.assign((size_t)R * ncell * CPC_NFIELD, 0.0);
    err.assign((size_t)R * ncell * CPC_NFIELD, 0.0);
  }
};

double h_or(const std::vector<double> &v, int64_t i, double d) { return v.empty() ? d : v[i]; }

// One sky pair binned on the host: its cell, the regions it touches (-1: none) and its sums.
struct HostPair {
  int c;
  int reg[2];
  double wt;
  double v[CPC_NFIELD], e[CPC_NFIELD];
};

// The reference's sky pair with the host libm (pi_calc, rp_calc_sky/dA_calc, corel_bin, the kind,
// corel_jk with jk_check_sky): corels.cpp:21-244, calcs.cpp:60-64,97-105,126-132,
// sky_calcs.cpp:13-31,91-96, jk.cpp:53-57,138-171,201-223.  False when the pair lands in no bin.
bool host_sky_pair(const cpc_params &p, const Derived &dv, const cpc_catalog *g1, int64_t i, const cpc_catalog *g2,
                   int64_t j, bool jk_on, HostPair &out) {
  const SkyMode m = sky_mode(p);
  const double dec1 = g1->h_y[i] / kR2D, dec2 = g2->h_y[j] / kR2D;
  const double ra1 = g1->h_x[i] / kR2D, ra2 = g2->h_x[j] / kR2D;
  // angle_orient_axis_sky (sky_calcs.cpp:13-22): pi_calc_phi and the shape kinds
  double theta = 0;
  if (p.kind != CPC_KIND_DD || m.pim == 2)
    theta = std::atan((g2->h_y[j] - g1->h_y[i]) / ((g2->h_x[j] - g1->h_x[i]) * std::cos((dec1 + dec2) / 2.0)));
  auto ephi = [&](double phi) {  // ephi_calc, calcs.cpp:97-105
    double t = phi - theta;
    t = t < 0 ? -1 * t : t;
    if (t > M_PI / 2.0) t = M_PI - t;
    return t;
  };
  double d_P = 0;  // pi_calc_projected
  if (m.pim == 0)
    d_P = g2->h_dc[j] - g1->h_dc[i];
  else if (m.pim == 2)
    d_P = ephi(p.g_with_shape == 2 ? h_or(g2->h_phi, j, -10.0) : h_or(g1->h_phi, i, -10.0));
  const double sum = std::sin(dec1) * std::sin(dec2) + std::cos(dec1) * std::cos(dec2) * std::cos((ra1 - ra2));
  const double da = (sum > 0.9999999999) ? std::acos(1.0) : std::acos(sum);
  double da_R = m.rpm == 0 ? da * (g1->h_da[i] + g2->h_da[j]) / 2. : (m.rpm == 1 ? da * g1->h_da[i] : da * kR2D);
  if (m.rmu) {  // r_mu_calc, calcs.cpp:135-140
    da_R = std::sqrt(da_R * da_R + d_P * d_P);
    d_P /= da_R;
  }
  const double q2 = std::floor((d_P - p.pi_min) / dv.psize);
  if (!(q2 >= 0 && q2 < p.n_pi)) return false;
  const double q1 = p.lin_bin ? std::floor((da_R - p.rp_min) / dv.rsize)
                              : std::floor(std::log10(da_R / p.rp_min) / dv.rsize);
  if (!(q1 >= 0 && q1 < p.n_rp)) return false;
  const int c = (int)q1 * p.n_pi + (int)q2;
  const double wt = h_or(g1->h_w, i, 1.0) * h_or(g2->h_w, j, 1.0);
  double *v = out.v, *e = out.e;
  std::fill(v, v + CPC_NFIELD, 0.0);
  std::fill(e, e + CPC_NFIELD, 0.0);
  if (p.kind != CPC_KIND_DD) {
    auto rot = [&](double e1, double e2, double &ep, double &ex) {
      ep = std::cos(2.0 * theta) * e1 - std::sin(2.0 * theta) * e2;
      ex = std::sin(2.0 * theta) * e1 + std::cos(2.0 * theta) * e2;
    };
    if (p.kind == CPC_KIND_DENSITY_CONVERGENCE) {
      // density_convergence_corel (corels.cpp:79-91): kappa in e[0] of the g_with_shape galaxy
      if (p.g_with_shape == 1) v[1] = wt * g1->h_e1[i] * 1.0;
      else if (p.g_with_shape == 2) v[1] = wt * g2->h_e1[j] * 1.0;
    } else if (p.kind == CPC_KIND_SHAPE_CONVERGENCE) {
      // shape_convergence_corel (corels.cpp:93-120)
      double ep = 0, epr = 0, exr = 0;
      if (p.g_with_shape == 2) {
        rot(g2->h_e1[j], g2->h_e2[j], epr, exr);
        ep = wt * g1->h_e1[i] * 1.0;
      } else if (p.g_with_shape == 1) {
        rot(g1->h_e1[i], g1->h_e2[i], epr, exr);
        ep = wt * g2->h_e1[j] * 1.0;
      }
      v[1] = ep;
      v[3] = ep * epr;
      v[5] = ep * exr;
    } else if (p.kind == CPC_KIND_SHAPE_DENSITY) {
      double ep, ex;
      if (p.g_with_shape == 2)
        rot(g2->h_e1[j], g2->h_e2[j], ep, ex);
      else
        rot(g1->h_e1[i], g1->h_e2[i], ep, ex);
      v[1] = 1.0 * wt * ep;
      e[1] = (wt * ep) * (wt * ep);
      v[2] = 1.0 * wt * ex;
      e[2] = (wt * ex) * (wt * ex);
      v[6] = ephi(h_or(g1->h_phi, i, -10.0));
      e[6] = v[6] * v[6];
    } else {
      double epi, exi, epj, exj;
      rot(g1->h_e1[i], g1->h_e2[i], epi, exi);
      rot(g2->h_e1[j], g2->h_e2[j], epj, exj);
      if (ss_shape_side_only(p)) {
        const double ep = p.g_with_shape == 2 ? epj : epi;
        v[1] = 1.0 * wt * ep;
        e[1] = (wt * ep) * (wt * ep);
        v[6] = ephi(p.g_with_shape == 2 ? h_or(g2->h_phi, j, -10.0) : h_or(g1->h_phi, i, -10.0));
        e[6] = v[6] * v[6];
      } else {
        v[1] = 1.0 * wt * (epi + epj) / 2.0;
        e[1] = (wt * epi / 2.0) * (wt * epi / 2.0) + (wt * epj / 2.0) * (wt * epj / 2.0);
        v[2] = 1.0 * wt * (exi + exj) / 2.0;
        e[2] = (wt * exi / 2.0) * (wt * exi / 2.0) + (wt * exj / 2.0) * (wt * exj / 2.0);
        const double t1 = ephi(h_or(g1->h_phi, i, -10.0)), t2 = ephi(h_or(g2->h_phi, j, -10.0));
        e[6] = t1 * t1 + t2 * t2;
        v[6] = t1 + t2;
      }
      v[3] = 1.0 * wt * (epi * epj);
      e[3] = (wt * (epi * epj)) * (wt * (epi * epj));
      v[4] = 1.0 * wt * (exi * exj);
      e[4] = (wt * (exi * exj)) * (wt * (exi * exj));
      v[5] = 1.0 * wt * (exi * epj + epi * exj) / 2.0;
      const double a = wt * (exi * epj) / 2.0, b = wt * epi * exj / 2.0;
      e[5] = a * a + b * b;
    }
  }
  out.c = c;
  out.wt = wt;
  out.reg[0] = out.reg[1] = -1;
  if (!jk_on) return true;
  int gjk[2] = {g1->h_jk[i], g2->h_jk[j]};
  double jp[2] = {h_or(g1->h_jkp, i, 0.0), h_or(g2->h_jkp, j, 0.0)};
  if (gjk[0] == gjk[1]) jp[1] = 0;
  for (int s = 0; s < 2; ++s)
    if (jp[s] >= 0.5) out.reg[s] = gjk[s];
  return true;
}

// Adds a host-binned pair to the full sample and to the touch sums of its regions.
void host_add(const cpc_params &p, const HostPair &q, HostHist &H) {
  const int c = q.c;
  H.num[c]++;
  H.wt[c] += q.wt;
  for (int f = 1; f < CPC_NFIELD; ++f) {
    H.val[(size_t)c * CPC_NFIELD + f] += q.v[f];
    H.err[(size_t)c * CPC_NFIELD + f] += q.e[f];
  }
  for (int s = 0; s < 2; ++s) {
    if (q.reg[s] < 0) continue;
    const size_t o = (size_t)(1 + q.reg[s]) * H.ncell + c;
    H.num[o]++;
    H.wt[o] += q.wt;
    if (p.which_corr != 0)
      for (int f = 1; f < CPC_NFIELD; ++f) H.val[o * CPC_NFIELD + f] += q.v[f];
  }
}

// ---------------------------------------------------------------------------------------------
// Passes.  One pass (a cpc_count call, or one job of cpc_count_batch) is checked and prepared on
// the host into a Plan: the derived binning, rp thresholds, cell grid, both sides' records on the
// device, the work items and the kernel arguments.  Box plans then run together, one launch per
// (kind, precision, pass direction) over the items of all of them; sky plans run one at a time.
// Each plan's histograms are assembled into the reference layout last.

struct Plan {
  // arguments
  const cpc_catalog *co = nullptr, *ci = nullptr;
  cpc_params p{};
  cpc_hist *out = nullptr;
  cpc_stats *stats = nullptr;
  // derived
  bool sky = false, jk_on = false, autop = false, split = false, use_split = false;
  int R = 1, nreg = 1, ncell = 0, kind_code = 0, NV = 0, NE = 0, NSP = 1, threads = 256;
  int row0_p = 0, ncopy_p = 1;  // LDS rows and copies
  Derived dv{};
  const std::vector<double> *thr = nullptr;
  // device state and the host arrays copied into it (alive until the call's last synchronisation)
  JobBufs *B = nullptr;
  Side *SO = nullptr, *SI = nullptr;
  std::vector<int2> nbr;
  std::vector<Chunk> chunks;
  std::vector<uint8_t> aflags;
  Win *wins = nullptr;  // sky windows of the outer galaxies: in JobBufs::win_host, or wins_vec
  std::vector<Win> wins_vec;
  std::vector<double2> reach;  // sky: RA/Dec half-widths of the outer windows per redshift bin
  // work
  std::vector<Item> items_p;
  std::vector<int3> ytar;  // (outer region, touches-outer flag, inner region) of targets 2 nreg + k
  Grid G{};                // cell grid
  int Rn = 0;              // box: neighbour reach in columns
  int per_item = 8;        // chunks per work item
  // with the jackknife: the nP raw targets stay on the device and are reduced there into
  // 1 + n_jk downloaded histograms, the full sample and touch_r (lists rbeg/rlst)
  bool reduce = false;
  std::vector<int32_t> rbeg, rlst;
  size_t off_raw = 0, off_rawn = 0;
  std::vector<double> cost_p;
  KArgs kp{};    // box: pass arguments (accumulator pointers set at launch)
  SkyArgs sa{};  // sky: pass arguments
  int hq_cap = 1;
  // accumulators: element offsets into the context's arenas
  size_t nP = 0;
  size_t off_p = 0, off_pn = 0, off_c = 0;
  // results
  double ms_build = 0;
  float ms_cells = 0;                   // device time of the cell-list build (first plan of a call)
  float kms = 0, kms_p = 0, kms_r = 0;  // pass, pair kernel, reduction
  std::vector<uint2> hq_pairs;
};

// Pair-kind code of the kernels and the data[] fields each kind fills (corels.cpp:48-244):
// 0 DD, 1 shape-density, 2 shape-shape estimator 0, 3 shape-shape other estimators,
// 4 density-convergence, 5 shape-convergence, 6 ED.  kNE: how many of the fields (the first ones) also
// carry their err sums.
const int kNV[7] = {0, 3, 5, 6, 1, 3, 1};
const int kNE[7] = {0, 3, 5, 6, 0, 0, 1};
const int kValField[7][6] = {{0}, {1, 2, 6}, {1, 3, 4, 5, 6}, {1, 2, 3, 4, 5, 6}, {1}, {1, 3, 5}, {6}};

// The reference's own preconditions and the engine's, checked before any device work.
Status validate(const cpc_job &job, Plan &P) {
  P.co = job.outer;
  P.ci = job.inner;
  P.out = job.out;
  P.stats = job.stats;
  if (!P.co || !P.ci || !job.params || !P.out) return {CPC_ERR_INVALID, "null argument"};
  P.p = *job.params;
  const cpc_params &p = P.p;
  const cpc_catalog *co = P.co, *ci = P.ci;
  const bool sky = p.geometry == CPC_GEOM_SKY;
  P.sky = sky;
  if (p.geometry != CPC_GEOM_BOX && !sky) return {CPC_ERR_UNSUPPORTED, "unknown geometry"};
  if (!sky && !(p.box_size > 0)) return {CPC_ERR_INVALID, "box_size must be > 0"};
  if (p.n_rp <= 0 || p.n_pi <= 0) return {CPC_ERR_INVALID, "n_rp and n_pi must be > 0"};
  if (!(p.rp_max > 0) || p.rp_min > p.rp_max) return {CPC_ERR_INVALID, "Bad Bin limits"};  // initialization.cpp:299-303
  const bool unit_pi = sky && (p.coordinates == 3 || p.coordinates == 5);  // pi bins of width 1 (pi = 0)
  if (!unit_pi && !(p.pi_max > p.pi_min)) return {CPC_ERR_INVALID, "pi_max must exceed pi_min"};
  if (!p.lin_bin && !(p.rp_min > 0)) return {CPC_ERR_INVALID, "log rp bins need rp_min > 0"};
  if (p.r_mu != 0 && p.r_mu != 1) return {CPC_ERR_INVALID, "r_mu must be 0 or 1"};
  if (sky && p.r_mu) return {CPC_ERR_INVALID, "on the sky r-mu bins are coordinates 1; r_mu must be 0"};
  if (!sky && p.coordinates != 0) return {CPC_ERR_INVALID, "coordinates selects the sky formulas: 0 in the box"};
  if (sky && (p.coordinates < 0 || p.coordinates > 5))
    return {CPC_ERR_UNSUPPORTED, "sky coordinates must be 0-5 (6 and 7 are the periodic box)"};
  // rp windows narrower than half the box (each x-y neighbour column is met through one image); the
  // z window may be as wide as the reference allows: it tests the images -1, 0, +1 of lim_calc_PB
  // (PB_calcs.cpp:30-61) one by one, and so does the engine (the generator notebook's box of 75
  // with |pi| < 75 counts every pair under two z images)
  if (!sky && 2 * p.rp_max >= p.box_size)
    return {CPC_ERR_UNSUPPORTED, "rp_max must be below half the box size"};
  if (p.kind < CPC_KIND_DD || p.kind > CPC_KIND_ED) return {CPC_ERR_INVALID, "unknown kind"};
  if (p.kind == CPC_KIND_ED) {
    if (sky) return {CPC_ERR_UNSUPPORTED, "the ED kind is defined for the periodic box only (ED_calc_PB)"};
    const cpc_catalog *ws = p.g_with_shape == 2 ? ci : co;  // g2 if g_with_shape is 2, else g1
    if (ws->n && !ws->has_e3)
      return {CPC_ERR_INVALID, "the ED kind needs e1, e2 and e3 on the catalog whose axis it reads (the "
                               "reference would dereference an unallocated e pointer, PB_calcs.cpp:116-120)"};
  }
  const bool conv = p.kind == CPC_KIND_DENSITY_CONVERGENCE || p.kind == CPC_KIND_SHAPE_CONVERGENCE;
  if (p.kind != CPC_KIND_DD && p.kind != CPC_KIND_ED && !conv && p.g_with_shape != 1 && p.g_with_shape != 2)
    return {CPC_ERR_INVALID, "g_with_shape must be 1 or 2 for shape kinds"};
  if (conv && (p.g_with_shape < 0 || p.g_with_shape > 2))
    return {CPC_ERR_INVALID, "g_with_shape must be 0, 1 or 2 for the convergence kinds"};
  if (p.do_jk != 0 && p.do_jk != 1) return {CPC_ERR_UNSUPPORTED, "do_jk=2 (region pairs) is not supported"};
  if (p.auto_pairs != 0 && p.auto_pairs != 1) return {CPC_ERR_INVALID, "auto_pairs must be 0 or 1"};
  P.autop = p.auto_pairs == 1;
  if (P.autop && co != ci) return {CPC_ERR_INVALID, "the auto pass (corels1) takes one catalog on both sides"};
  if (p.unsorted < 0 || p.unsorted > 2) return {CPC_ERR_INVALID, "unsorted must be 0, 1 or 2"};
  if (p.unsorted == 2 && !sky)
    return {CPC_ERR_INVALID, "unsorted = 2 (data_sorted 2, Dec order) is for the sky (initialization.cpp:158-160)"};
  if (P.autop && !p.unsorted && co->n > 0 && !co->z_sorted)
    return {CPC_ERR_INVALID, "the auto pass needs the catalog sorted by z (data_sorted=1, corr.cpp:129-132)"};
  if (p.precision != CPC_PREC_SPLIT && p.precision != CPC_PREC_FP64) return {CPC_ERR_INVALID, "unknown precision"};
  P.jk_on = p.do_jk == 1 && p.n_jk > 0;
  if (P.jk_on) {
    for (const cpc_catalog *c : {co, ci}) {
      if (c->n > 0 && !c->has_jk) return {CPC_ERR_INVALID, "jackknife on but a catalog has no jk array"};
      if (c->n > 0 && (c->jk_lo < 0 || c->jk_hi >= p.n_jk)) return {CPC_ERR_INVALID, "jk region outside [0, n_jk)"};
    }
  }
  if (p.kind == CPC_KIND_SHAPE_SHAPE && ((co->n && !co->has_e) || (ci->n && !ci->has_e)))
    return {CPC_ERR_INVALID, "shape_shape needs e1/e2 on both catalogs"};
  if (p.kind == CPC_KIND_SHAPE_DENSITY) {
    const cpc_catalog *sc = p.g_with_shape == 2 ? ci : co;
    if (sc->n && !sc->has_e)
      return {CPC_ERR_INVALID, "shape_density needs e1/e2 on the shape side (the reference would "
                               "dereference an unallocated e pointer, corels.cpp:216-219)"};
  }
  if (conv && p.g_with_shape != 0) {
    // the galaxy read for kappa (e[0]) and, for shape-convergence, the one rotated (corels.cpp:79-120)
    const cpc_catalog *ws = p.g_with_shape == 2 ? ci : co, *other = p.g_with_shape == 2 ? co : ci;
    const cpc_catalog *kap = p.kind == CPC_KIND_DENSITY_CONVERGENCE ? ws : other;
    if (kap->n && !kap->has_e)
      return {CPC_ERR_INVALID, "the convergence kinds need kappa (e1) on the catalog that carries it (the "
                               "reference would dereference an unallocated e pointer, corels.cpp:79-120)"};
    if (p.kind == CPC_KIND_SHAPE_CONVERGENCE && ws->n && !ws->has_e)
      return {CPC_ERR_INVALID, "shape_convergence needs e1/e2 on the shape side"};
  }
  if (sky) {
    if ((co->n && !co->has_sky) || (ci->n && !ci->has_sky))
      return {CPC_ERR_INVALID, "the sky geometry needs dc and da on both catalogs"};
    if (!p.sky_H || !p.sky_dA || p.sky_n < 2 || !(p.sky_dz > 0))
      return {CPC_ERR_INVALID, "the sky geometry needs the distance table"};
  } else {
    for (const cpc_catalog *c : {co, ci})
      for (int a = 0; a < 3; ++a)
        if (c->n > 0 && (c->lo[a] < 0 || c->hi[a] > p.box_size))
          return {CPC_ERR_INVALID, "coordinates must lie in [0, box_size]"};
  }
  P.nreg = P.jk_on ? p.n_jk : 1;
  P.R = P.jk_on ? 1 + p.n_jk : 1;
  P.ncell = p.n_rp * p.n_pi;
  P.dv = derive(p);
  if (p.kind == CPC_KIND_DD)
    P.kind_code = 0;
  else if (p.kind == CPC_KIND_SHAPE_DENSITY)
    P.kind_code = 1;
  else if (p.kind == CPC_KIND_SHAPE_SHAPE)
    P.kind_code = ss_shape_side_only(p) ? 2 : 3;
  else if (p.kind == CPC_KIND_ED)
    P.kind_code = 6;
  else
    P.kind_code = p.kind == CPC_KIND_DENSITY_CONVERGENCE ? 4 : 5;
  P.NV = kNV[P.kind_code];
  P.NE = kNE[P.kind_code];
  P.NSP = 1 + P.NV + P.NE;
  P.nP = (size_t)2 * P.nreg;
  P.split = p.precision == CPC_PREC_SPLIT;
  return {};
}

// Thresholds, grid, windows, both sides, work items, device tables and kernel arguments.
Status setup(cpc_ctx *ctx, Plan &P) {
  const cpc_params &p = P.p;
  const cpc_catalog *co = P.co, *ci = P.ci;
  const bool sky = P.sky, jk_on = P.jk_on, autop = P.autop;
  const int nreg = P.nreg, ncell = P.ncell;
  JobBufs &B = *P.B;
  Status st;
  hipStream_t s = ctx->stream;

  P.thr = cached_thresholds(ctx, p, !sky, st);
  if (!P.thr) return st;
  const std::vector<double> &thr = *P.thr;
  // box kernel: up to 512 threads (the split kernel keeps 4 waves per SIMD with 8 waves sharing a
  // histogram; the FP64 kernel needs more registers and runs 4 waves per group; the split density
  // kernel runs 5 waves per SIMD, which 256-thread groups fill: 145 against 170 ms on the 1M-galaxy
  // wgg pass); sky kernel: up to 256
  const int box_threads = P.split ? (P.NSP == 1 ? 256 : 512) : 256;
  P.threads = sky ? std::min(ctx->block_threads ? ctx->block_threads : 256, 256)
                  : (ctx->block_threads ? ctx->block_threads : box_threads);
  const size_t lds_p = pair_lds_bytes(ncell, P.NSP, p.n_rp);
  if (lds_p > ctx->max_lds)
    return {CPC_ERR_UNSUPPORTED, "histogram of " + std::to_string(ncell) + " cells does not fit in LDS (" +
                                     std::to_string(lds_p) + " bytes)"};
  {
    // LDS histogram: passes with 7 or more slots per cell (shape kinds), whose run flushes
    // dominate, get as many copies (at most 4) of the top rp rows as fit in the group's share of the
    // CU's LDS, keeping in LDS every row from rp >= rp_max / 20 up (rows below hold about 0.25% of
    // the pairs of a homogeneous sample and go to the global histogram directly: device-wide atomics
    // cost far more than LDS ones, 3% of the runs there cost as much as the whole LDS flush);
    // other passes keep the whole histogram in one copy.  Waves per CU: the box split kernel keeps
    // 4 per SIMD, the box FP64 and the sky kernels 2.
    const int waves_cu = (!sky && P.split) ? 16 : 8;
    const int groups_cu = std::max(1, waves_cu / std::max(1, P.threads / 64));
    const size_t share = ctx->max_lds / groups_cu;
    auto choose = [&](int ns, int &row0, int &ncopy) {
      row0 = 0;
      ncopy = 1;
      if (ns < 7) return;  // the kernel keeps the whole histogram in one copy (compiled that way)
      for (int c = 4; c >= 1; --c) {
        int rows = p.n_rp;
        while (rows > 0 && pair_lds_bytes(rows * p.n_pi, ns, p.n_rp, c) > share) --rows;
        if (rows == 0) continue;
        const int r0 = p.n_rp - rows;
        const double rp_edge = sky ? thr[r0] : std::sqrt(thr[r0]);  // sky thresholds are on rp, box on rp^2
        if (r0 == 0 || rp_edge <= 0.05 * p.rp_max) {
          row0 = r0;
          ncopy = c;
          return;
        }
      }
    };
    choose(P.NSP, P.row0_p, P.ncopy_p);
  }
  // block galaxies per work item: four chunks per wave of the group (waves claim chunks as they
  // finish, so a few chunks each keep the group's waves busy until close to its end)
  const int item_size = ctx->item_size ? ctx->item_size : (sky ? 512 : 4 * P.threads);
  const int64_t nbig = std::max(co->n, ci->n);

  // ---- geometry-specific set-up: grid, neighbour cells, windows
  Grid G;
  int Rn = 0;
  bool has_zfirst = false;
  double zfirst_plus_L = 0;
  SkyArgs SA{};
  auto t_build = std::chrono::steady_clock::now();
  if (!sky) {
    const double L = p.box_size;
    double frac = ctx->col_frac > 0 ? ctx->col_frac : 0.25;
    int nc = std::max(1, (int)std::floor(L / (p.rp_max * frac)));
    // keep enough galaxies per column for full waves
    const int nc_dense = std::max(1, (int)std::floor(std::sqrt((double)std::max<int64_t>(nbig, 1) / 256.0)));
    if (ctx->col_frac <= 0) nc = std::min(nc, nc_dense);
    nc = std::min(nc, 1024);
    G.ncx = G.ncy = nc;
    G.hx = G.hy = L / nc;
    const double h = G.hx;
    // neighbour column offsets (circle cut at the largest binned separation)
    const double rcut = std::sqrt(thr[p.n_rp]) * (1 + 1e-9) + 1e-9 * L;
    Rn = (int)std::floor(p.rp_max / h) + 1;
    for (int dy = -Rn; dy <= Rn; ++dy)
      for (int dx = -Rn; dx <= Rn; ++dx) {
        const double ex = std::max(0, std::abs(dx) - 1) * h, ey = std::max(0, std::abs(dy) - 1) * h;
        if (ex * ex + ey * ey <= rcut * rcut) P.nbr.push_back(make_int2(dx, dy));
      }
    // first included galaxy of the inner catalog in upload order (PBz>0 sweep, corr.cpp:407-416)
    for (int64_t j = 0; j < ci->n; ++j) {
      const double inc = ci->h_inc.empty() ? 0.0 : ci->h_inc[j];
      if (inc > p.include_prob_max) continue;
      has_zfirst = true;
      // an unsorted scan (do_pair_unsorted) does not stop at that first galaxy
      zfirst_plus_L = p.unsorted ? INFINITY : ci->h_z[j] + L * 1;
      break;
    }
    if (autop && !p.unsorted && co->n > 0) {
      // corels1 scans for PBz = 0 and PBz = -1 start at j = i+1 and stop at once when that next
      // included galaxy lies below the window (do_pair_sorted, corr.cpp:71-73): bit 0 / bit 1
      std::vector<uint8_t> &af = P.aflags;
      af.assign(co->n, 0);
      int64_t nxt = -1;
      for (int64_t i = co->n - 1; i >= 0; --i) {
        if (nxt >= 0) {
          const double z_min = co->h_z[i] + window_lo(p);
          if (!(co->h_z[nxt] + L * 0 < z_min)) af[i] |= 1;
          if (!(co->h_z[nxt] + L * -1 < z_min)) af[i] |= 2;
        }
        if (!(!co->h_inc.empty() && co->h_inc[i] > p.include_prob_max)) nxt = i;
      }
      if (check(B.aflags.ensure(co->n), st, "alloc")) return st;
      check(hipMemcpyAsync(B.aflags.p, af.data(), co->n, hipMemcpyHostToDevice, s), st, "copy");
      if (st.code) return st;
    }
  } else {
    // per-galaxy windows of the outer catalog (lim_calc_sky), with the host libm
    const SkyTable T{p.sky_H, p.sky_dA, p.sky_n, p.sky_dz, p.z_max, p.z_sep_max};
    if (t_arena || B.win_host.ensure(std::max<int64_t>(co->n, 1) * sizeof(Win)) != hipSuccess) {
      P.wins_vec.resize(std::max<int64_t>(co->n, 1));
      P.wins = P.wins_vec.data();
    } else {
      P.wins = reinterpret_cast<Win *>(B.win_host.p);
    }
    Win *const wins = P.wins;
    std::vector<double> dec_half;
    // RA/Dec reach per bin of the outer redshift
    double ozlo = INFINITY, ozhi = -INFINITY;
    for (int64_t i = 0; i < co->n; ++i) {
      if (!co->h_inc.empty() && co->h_inc[i] > p.include_prob_max) continue;
      ozlo = std::min(ozlo, co->h_z[i]);
      ozhi = std::max(ozhi, co->h_z[i]);
    }
    const int nb = 512;
    if (!(ozlo <= ozhi)) ozlo = ozhi = 0.0;
    const double rz0 = ozlo, rinv = nb / std::max(ozhi - ozlo, 1e-9);
    std::vector<double2> &reach = P.reach;
    reach.assign(nb, make_double2(0.0, 0.0));
    dec_half.reserve(co->n);
    // on host threads, a fixed range of galaxies a chunk, each with its own reach table and
    // half-widths, merged in chunk order
    const int64_t nchunk = std::min<int64_t>(64, std::max<int64_t>(1, co->n / 4096));
    std::vector<std::vector<double2>> creach(nchunk, std::vector<double2>(nb, make_double2(0.0, 0.0)));
    std::vector<std::vector<double>> chalf(nchunk);
    std::vector<int64_t> cbad(nchunk, -1);  // first galaxy whose window fails
    parallel_for(nchunk, host_threads(), [&](int64_t k) {
      const int64_t i0 = co->n * k / nchunk, i1 = co->n * (k + 1) / nchunk;
      std::vector<double2> &rr = creach[k];
      for (int64_t i = i0; i < i1; ++i) {
        const double inc = co->h_inc.empty() ? 0.0 : co->h_inc[i];
        Win &w = wins[i];
        if (inc > p.include_prob_max) {
          w = Win{1, 0, 0, 0, 0, 0, 0, 0};  // never used: the galaxy is not in the pass
          continue;
        }
        if (!sky_window(T, p, co->h_x[i], co->h_y[i], co->h_z[i], w)) {
          cbad[k] = i;
          return;
        }
        const int b = (int)std::fmax(0.0, std::fmin((double)(nb - 1), std::floor((co->h_z[i] - rz0) * rinv)));
        double2 &rb = rr[b];
        if (!(w.ramin <= w.ramax) || w.ramin <= 0 || w.ramax >= 360)
          rb.x = INFINITY;
        else
          rb.x = std::max(rb.x, std::max(co->h_x[i] - w.ramin, w.ramax - co->h_x[i]) * (1 + 1e-12) + 1e-9);
        const double dh = std::max(co->h_y[i] - w.decmin, w.decmax - co->h_y[i]);
        rb.y = std::max(rb.y, dh * (1 + 1e-12) + 1e-9);
        chalf[k].push_back(dh);
      }
    });
    for (int64_t k = 0; k < nchunk; ++k) {
      if (cbad[k] >= 0)
        return {CPC_ERR_INVALID, "galaxy " + std::to_string(cbad[k]) +
                                     ": its redshift window leaves the distance table or exceeds z_max "
                                     "(the reference exits here, sky_calcs.cpp:109-113)"};
      for (int b = 0; b < nb; ++b) {
        reach[b].x = std::max(reach[b].x, creach[k][b].x);
        reach[b].y = std::max(reach[b].y, creach[k][b].y);
      }
      dec_half.insert(dec_half.end(), chalf[k].begin(), chalf[k].end());
    }
    // RA/Dec grid over both catalogs, cells about half the median window
    double ralo = 360, rahi = 0, declo = 90, dechi = -90;
    for (const cpc_catalog *c : {co, ci})
      if (c->n) {
        ralo = std::min(ralo, c->lo[0]);
        rahi = std::max(rahi, c->hi[0]);
        declo = std::min(declo, c->lo[1]);
        dechi = std::max(dechi, c->hi[1]);
      }
    double hd = 1.0;
    if (!dec_half.empty()) {
      // the median Dec half-width sets the cell size only (never which pairs are counted): from at
      // most 65536 evenly strided galaxies
      const size_t stride = dec_half.size() / 65536 + 1;
      std::vector<double> smp;
      smp.reserve(dec_half.size() / stride + 1);
      for (size_t i = 0; i < dec_half.size(); i += stride) smp.push_back(dec_half[i]);
      std::nth_element(smp.begin(), smp.begin() + smp.size() / 2, smp.end());
      hd = smp[smp.size() / 2] * (ctx->col_frac > 0 ? ctx->col_frac : 0.5);
    }
    const double span_ra = std::max(rahi - ralo, 1e-6), span_dec = std::max(dechi - declo, 1e-6);
    hd = std::max(hd, span_dec / 2048);
    const double cdec = std::max(0.2, std::cos(0.5 * (declo + dechi) / kR2D));
    const double hr = std::max(hd / cdec, span_ra / 2048);
    G.ncx = std::max(1, (int)std::ceil(span_ra / hr));
    G.ncy = std::max(1, (int)std::ceil(span_dec / hd));
    G.hx = span_ra / G.ncx * (1 + 1e-12);
    G.hy = span_dec / G.ncy * (1 + 1e-12);
    G.x0 = ralo;
    G.y0 = declo;
    if (autop && !p.unsorted) {
      // corels1's scan from j = i+1 stops at once when that next included galaxy lies below the
      // window (corr.cpp:168-186): such outer galaxies take no pairs
      int64_t nxt = -1;
      for (int64_t i = co->n - 1; i >= 0; --i) {
        const bool inc_i = !(!co->h_inc.empty() && co->h_inc[i] > p.include_prob_max);
        if (inc_i && (nxt < 0 || co->h_z[nxt] < wins[i].zmin)) {
          wins[i].zmin = 1.0;
          wins[i].zmax = 0.0;
        }
        if (inc_i) nxt = i;
      }
    }
    if (check(B.win_raw.ensure(std::max<int64_t>(co->n, 1)), st, "alloc")) return st;
    if (co->n)
      check(hipMemcpyAsync(B.win_raw.p, wins, co->n * sizeof(Win), hipMemcpyHostToDevice, s), st, "copy");
    if (st.code) return st;
  }

  // ---- sides (sort keys are (cell, region, coin) packed in 32 bits, all ones meaning excluded)
  if ((double)G.ncx * G.ncy * nreg * 2.0 >= 4294967295.0)
    return {CPC_ERR_UNSUPPORTED, "too many cells x jackknife regions for 32-bit sort keys"};
  st = place_side(ctx, co, G, nreg, jk_on, p, has_zfirst, zfirst_plus_L, sky, sky ? B.win_raw.p : nullptr,
                  (autop && !sky && !p.unsorted) ? B.aflags.p : nullptr, B.so);
  if (st.code) return st;
  const bool same = (co == ci);
  if (!same) {
    st = place_side(ctx, ci, G, nreg, jk_on, p, has_zfirst, zfirst_plus_L, sky, nullptr, nullptr, B.si);
    if (st.code) return st;
  }
  P.SO = &B.so;
  P.SI = same ? &B.so : &B.si;
  P.ms_build = ms_since(t_build);

  P.G = G;
  P.Rn = Rn;
  P.sa = SA;
  P.per_item = std::max(1, item_size / 64);
  return st;
}

// The plan's work items, from its sides' group tables: host work only, so the plans of a batch
// build theirs concurrently.
// (threads: host threads for this plan's own work; 1 when plans are set up on threads already)
void setup_items(Plan &P, int threads) {
  const bool sky = P.sky, jk_on = P.jk_on;
  const int nreg = P.nreg, per_item = P.per_item;
  const Side &SO = *P.SO, &SI = *P.SI;
  {
    const int ncx = P.G.ncx, ncy = P.G.ncy;
    const size_t ncols = (size_t)ncx * ncy;
    if (!sky) {
      // periodic neighbour columns of the box
      const std::vector<int2> &nbr = P.nbr;
      auto for_nbr = [&](int col, auto &&fn) {
        const int bcx = col % ncx, bcy = col / ncx;
        for (const int2 &o : nbr) {
          const int ux = ((bcx + o.x) % ncx + ncx) % ncx, uy = ((bcy + o.y) % ncy + ncy) % ncy;
          fn(uy * ncx + ux);
        }
      };
      make_items_jk(SO, SI, jk_on, true, false, nreg, per_item, ncols, for_nbr, P.items_p, P.cost_p, P.chunks, P.ytar,
                    threads);
    } else {
      // the sky: cells within the widest window reach of any outer galaxy (every RA cell when some
      // window covers the full circle)
      double rra = 0, rdec = 0;
      for (const double2 &r : P.reach) {
        rra = std::max(rra, r.x);
        rdec = std::max(rdec, r.y);
      }
      const int dx = std::isfinite(rra) ? (int)std::min<double>(ncx, std::ceil(rra / P.G.hx) + 1) : ncx;
      const int dy = (int)std::min<double>(ncy, std::ceil(rdec / P.G.hy) + 1);
      auto for_nbr = [&](int col, auto &&fn) {
        const int bcx = col % ncx, bcy = col / ncx;
        for (int y = std::max(0, bcy - dy); y <= std::min(ncy - 1, bcy + dy); ++y)
          for (int x = std::max(0, bcx - dx); x <= std::min(ncx - 1, bcx + dx); ++x) fn(y * ncx + x);
      };
      make_items_jk(SO, SI, jk_on, false, true, nreg, per_item, ncols, for_nbr, P.items_p, P.cost_p, P.chunks, P.ytar,
                    threads);
    }
    P.nP = (size_t)2 * nreg + P.ytar.size();
    if (jk_on) {
      // output 0: every target; output r+1: the targets whose pairs touch region r (2r+1 and the
      // Y targets through the outer galaxy, x = 1, or the inner one, b = r)
      std::vector<std::vector<int32_t>> lists(1 + nreg);
      for (int t = 0; t < (int)P.nP; ++t) lists[0].push_back(t);
      for (int r = 0; r < nreg; ++r) lists[1 + r].push_back(2 * r + 1);
      for (size_t k = 0; k < P.ytar.size(); ++k) {
        const int3 y = P.ytar[k];
        if (y.y) lists[1 + y.x].push_back(2 * nreg + (int)k);
        lists[1 + y.z].push_back(2 * nreg + (int)k);
      }
      P.rbeg.assign(1, 0);
      P.rlst.clear();
      for (const auto &l : lists) {
        P.rlst.insert(P.rlst.end(), l.begin(), l.end());
        P.rbeg.push_back((int32_t)P.rlst.size());
      }
      P.reduce = true;
    }
  }
}

// Device tables of the plan (thresholds, neighbour offsets, chunks, reduction lists), its guard
// bands and its kernel arguments.
Status setup_tables(cpc_ctx *ctx, Plan &P) {
  const cpc_params &p = P.p;
  const cpc_catalog *co = P.co, *ci = P.ci;
  const bool sky = P.sky, jk_on = P.jk_on, autop = P.autop;
  const int ncell = P.ncell;
  const Derived &dv = P.dv;
  JobBufs &B = *P.B;
  Status st;
  hipStream_t s = ctx->stream;
  const Grid &G = P.G;
  const int Rn = P.Rn;
  SkyArgs &SA = P.sa;
  const std::vector<double> &thr = *P.thr;
  Side &SO = *P.SO, &SI = *P.SI;
  const int64_t nbig = std::max(co->n, ci->n);
  (void)jk_on;
  (void)s;
  // (the device tables themselves were placed by stage_tables)
  const double two24 = std::ldexp(1.0, -24);
  double gz = 1.0;
  Guard gr{};
  float cull_r2 = INFINITY, cull_d = 0.0f;
  if (!sky) {
    // guard bands of the FP32 classification from bounds on the FP32 error of dx, dy (column-local
    // coordinates plus a rounded column offset) and of dz (absolute coordinates)
    const double L = p.box_size, h = G.hx;
    const double amax = 1.01 * h + 1e-9 * L;
    const double omax = (Rn + 1) * h + 1e-9 * L;
    const double eps_xy = 2.0 * (6 * amax + 2 * omax) * two24;
    gr = rp_guard(p, dv, thr, eps_xy);
    // tile cull radius: the neighbour-column cut (no pair beyond it reaches a bin), widened by the
    // FP32 rounding of the squares and their sum; cull_d covers the rounding of the tile's and the
    // lanes' column-relative coordinates and of the axis distances (four roundings of magnitude
    // at most omax + amax), doubled
    const double rcut_c = std::sqrt(thr[p.n_rp]) * (1 + 1e-9) + 1e-9 * L;
    cull_r2 = round_up_f(rcut_c * rcut_c * (1 + 1e-6));
    cull_d = round_up_f(2.0 * 4 * (omax + amax) * two24);
    const double eps_z = 2.0 * 5 * L * two24;
    const double pabs = std::max(std::fabs(p.pi_min), std::fabs(p.pi_max));
    gz = eps_z / dv.psize + 8.0 * (p.n_pi + 2 * pabs / dv.psize + 2) * two24;
  }
  // r-mu passes classify every candidate in FP64 (the FP64 kernel with the r-mu code compiled in),
  // and so do box passes with linear rp bins: the box split classification is for log bins
  P.use_split = P.split && (sky || gz < 0.25) && (sky || (!p.r_mu && !p.lin_bin));

  if (!sky) {
    KArgs K{};
    K.L = p.box_size;
    K.pmin = p.pi_min;
    K.pmax = p.pi_max;
    K.psize = dv.psize;
    K.rmax = p.rp_max;
    K.wlo = window_lo(p);
    K.whi = window_hi(p);
    K.rmu = p.r_mu ? 1 : 0;
    K.h = G.hx;
    K.zmargin = 1e-7 * (p.box_size + std::max(std::fabs(K.wlo), std::fabs(K.whi)));
    K.nc = G.ncx;
    K.n_rp = p.n_rp;
    K.n_pi = p.n_pi;
    K.ncell = ncell;
    K.gws = p.g_with_shape;
    K.n_nbr = (int)P.nbr.size();
    K.lin = p.lin_bin;
    K.pioff32 = (float)(-p.pi_min / dv.psize);
    K.inv_psize32 = (float)(1.0 / dv.psize);
    K.gz = (float)gz;
    K.la = (float)(0.5 * std::log10(2.0) / dv.rsize);
    K.lb = (float)(-std::log10(p.rp_min > 0 ? p.rp_min : 1.0) / dv.rsize);
    K.lrmin = (float)p.rp_min;
    K.linv = (float)(1.0 / dv.rsize);
    K.thr = B.thr.p;
    K.s_out_lo = gr.s_out_lo;
    K.s_out_hi = gr.s_out_hi;
    K.s_cert_lo = gr.s_cert_lo;
    K.gA = gr.gA;
    K.gB = gr.gB;
    K.gC = gr.gC;
    K.gD = gr.gD;
    K.cull_r2 = cull_r2;
    K.cull_d = cull_d;
    K.nbr = B.nbr.p;
    K.auto_mode = autop ? (p.unsorted ? 2 : 1) : 0;
    K.chunks = B.chunks.p;
    // the pass: blocks = outer galaxies, tiles = inner galaxies
    P.kp = K;
    P.kp.brec = SO.rec.p;
    P.kp.bz = SO.zs.p;
    P.kp.trec = SI.rec.p;
    P.kp.tz = SI.zs.p;
    P.kp.nslots = P.NSP;
    P.kp.row0 = P.row0_p;
    P.kp.ncopy = P.ncopy_p;
    P.kp.bgrp = SO.d_groups.p;
    P.kp.tgrp = SI.d_groups.p;
    P.kp.tcol = SI.d_colbeg.p;
    P.kp.bidx = SO.idx.p;
    P.kp.tidx = SI.idx.p;
  } else {
    SkyArgs &A = SA;
    A.pmin = p.pi_min;
    A.pmax = p.pi_max;
    A.psize = dv.psize;
    A.n_rp = p.n_rp;
    A.n_pi = p.n_pi;
    A.ncell = ncell;
    A.gws = p.g_with_shape;
    A.lin = p.lin_bin;
    A.ra0 = G.x0;
    A.dec0 = G.y0;
    A.hra = G.hx;
    A.hdec = G.hy;
    A.nra = G.ncx;
    A.ndec = G.ncy;
    // sum = fl(fl(sd sd) + fl(fl(cd cd) c)) with sd, cd from the host: only c = cos(a1 - a2) differs
    // between device and host, by at most kSkyCosUlp ulps (at most 2^-53 each, c <= 1); the product
    // and the sum can then round apart by one more ulp each: |dsum| <= (k + 2) 2^-53, doubled
    A.dsum = 2.0 * (kSkyCosUlp + 2.0) * std::ldexp(1.0, -53);
    const SkyMode sm = sky_mode(p);
    A.pim = sm.pim;
    A.rpm = sm.rpm;
    A.rmu = sm.rmu ? 1 : 0;
    A.dphi = 1e-13;
    A.auto_mode = autop ? (p.unsorted ? 2 : 1) : 0;
    // Split classification.  The estimate rp^2 = theta^2 dbar^2 (theta^2 = 4 asin^2(c/2) from the
    // chord c of FP64 unit vectors, a series with relative error below 1e-12) differs from the
    // reference's acos(sum) dbar, squared, by at most 2 (5e-16 / theta^2) relative (the rounding of
    // sum near 1, about 4 ulp, propagated through acos) plus 1e-12; theta >= rp_min / DA_max for pairs that can reach
    // a bin.  Rounding s to FP32 adds 2^-24.  In bin units: la / ln 2 per unit relative error, plus
    // v_log_f32 (2^-21 max(1, |log2 s|)) times la, the rounding of la and lb, and a margin of 2.
    // pi bin by product: (dP - pmin) * fl(1/psize) is within 2^-51 |u| of the quotient the reference
    // rounds (one rounding of 1/psize, one of the product, half an ulp of the IEEE quotient); with
    // |u| below 2^20 (the kernel checks it) that is under 2^-31, so a fraction 1e-9 from an edge is certain
    A.inv_psize = 1.0 / dv.psize;
    A.pguard = 1e-9;
    A.split_on = 0;
    if (P.split && !p.lin_bin && p.rp_min > 0 && sm.pim == 0 && sm.rpm == 0 && !sm.rmu) {
      double da_max = 0;
      for (const cpc_catalog *c : {co, ci})
        for (double v : c->h_da) da_max = std::max(da_max, v);
      const double th_min = p.rp_min / std::max(da_max, 1e-300);
      const double rel = 2 * 5e-16 / (th_min * th_min) + 1e-12 + std::ldexp(1.0, -24);
      const double la = 0.5 * std::log10(2.0) / dv.rsize, lb = -std::log10(p.rp_min) / dv.rsize;
      const double s_lo = thr[0] * thr[0] * (1 - 4 * rel), s_hi = thr[p.n_rp] * thr[p.n_rp] * (1 + 4 * rel);
      const double l2 = std::max({1.0, std::fabs(std::log2(std::max(s_lo, 1e-300))), std::fabs(std::log2(s_hi))});
      const double t24 = std::ldexp(1.0, -24), t21 = std::ldexp(1.0, -21);
      const double g = 2.0 * (la * rel / std::log(2.0) + la * l2 * (t21 + 2 * t24) + std::fabs(lb) * t24 +
                              (p.n_rp + 2) * t24 * 2) + 1e-12 * (p.n_rp + 1);
      if (g < 0.1) {
        A.split_on = 1;
        A.la = (float)la;
        A.lb = (float)lb;
        A.g = round_up_f(g);
        A.s_out_lo = round_down_f(s_lo);
        A.s_out_hi = round_up_f(s_hi);
      }
    }
    A.thr = B.thr.p;
    A.chunks = B.chunks.p;
    P.sa = A;
    // host-queue capacity for sky guard-band pairs; grown and the pass rerun if ever exceeded
    P.hq_cap = (int)std::min<int64_t>(std::max<int64_t>(1 << 16, nbig), 1 << 24);
  }
  return st;
}

// The small device tables of every plan of a call (group tables and column starts of both sides,
// reduction lists, thresholds, neighbour offsets, chunks) in one device region, filled by one copy
// from a host staging buffer: a batch of many passes would otherwise pay thousands of small copies.
// The tables' DevBufs then point into the region (borrowed: the region owns the memory).
Status stage_tables(cpc_ctx *ctx, std::vector<Plan> &plans) {
  Status st;
  size_t total = 0;
  struct Copy {
    size_t at;
    const void *src;
    size_t bytes;
  };
  std::vector<Copy> copies;
  std::vector<std::function<void(char *)>> binders;
  auto add = [&](auto &buf, const void *src, size_t n) {
    using T = typename std::remove_pointer<decltype(buf.p)>::type;
    const size_t cap = std::max<size_t>(n, 1);
    const size_t at = (total + 255) & ~(size_t)255;
    total = at + cap * sizeof(T);
    if (n) copies.push_back({at, src, n * sizeof(T)});
    binders.push_back([&buf, at, cap](char *base) {
      if (buf.p && !buf.borrowed && !t_arena) (void)hipFree(buf.p);
      buf.p = reinterpret_cast<T *>(base + at);
      buf.cap = cap;
      buf.borrowed = true;
    });
  };
  for (Plan &P : plans) {
    JobBufs &B = *P.B;
    // the outer side, and the inner side when it is another catalog (an auto pass has one side)
    add(P.SO->d_groups, P.SO->groups.data(), P.SO->groups.size());
    add(P.SO->d_colbeg, P.SO->colbeg.data(), P.SO->colbeg.size());
    if (P.SI != P.SO) {
      add(P.SI->d_groups, P.SI->groups.data(), P.SI->groups.size());
      add(P.SI->d_colbeg, P.SI->colbeg.data(), P.SI->colbeg.size());
    }
    if (P.reduce) {
      add(B.red_beg, P.rbeg.data(), P.rbeg.size());
      add(B.red_lst, P.rlst.data(), P.rlst.size());
    }
    add(B.thr, P.thr->data(), P.thr->size());
    add(B.nbr, P.nbr.data(), P.nbr.size());
    add(B.chunks, P.chunks.data(), P.chunks.size());
  }
  if (check(ctx->tables.ensure(std::max<size_t>(total, 1)), st, "alloc")) return st;
  // staged in page-locked memory unless the call may not allocate; the gaps between tables are
  // never read
  std::vector<char> hv;
  char *h;
  if (!t_arena && ctx->hb_stage.ensure(total) == hipSuccess) {
    h = ctx->hb_stage.p;
  } else {
    hv.resize(total);
    h = hv.data();
  }
  parallel_for((int64_t)copies.size(), copies.size() > 256 ? host_threads() : 1,
               [&](int64_t k) { std::memcpy(h + copies[k].at, copies[k].src, copies[k].bytes); });
  if (total) check(hipMemcpyAsync(ctx->tables.p, h, total, hipMemcpyHostToDevice, ctx->stream), st, "copy");
  // the staging memory is reused by the next call: wait for the copy
  check(hipStreamSynchronize(ctx->stream), st, "sync");
  for (auto &b : binders) b(ctx->tables.p);
  return st;
}

// All box plans: one launch per (kind, precision, auto), each over the items of every plan it
// serves ordered by decreasing cost, so that the device starts the longest items first whichever
// job they belong to; then the jackknife reductions.  Kernel times are those of the whole group of
// launches.
Status run_box(cpc_ctx *ctx, const std::vector<Plan *> &plans) {
  Status st;
  if (plans.empty()) return st;
  hipStream_t s = ctx->stream;
  // job table: entry i is plan i's pass
  std::vector<KArgs> table(plans.size());
  for (size_t i = 0; i < plans.size(); ++i) {
    Plan &P = *plans[i];
    KArgs &a = table[i];
    a = P.kp;
    a.acc = P.reduce ? ctx->acc_raw.p + P.off_raw : ctx->acc.p + P.off_p;
    a.accnum = P.reduce ? ctx->accn_raw.p + P.off_rawn : ctx->accn.p + P.off_pn;
    a.counters = ctx->accn.p + P.off_c;
  }
  struct Launch {
    int kind;
    bool split;
    int threads;
    size_t lds, first, n;
  };
  std::map<std::pair<int, int>, std::vector<size_t>> groups;  // (kind, split) -> plans
  // (kind, split, auto) -> plans; r-mu box passes run their own instantiation (kind + kRmuKind),
  // auto passes (corels1) theirs (+ kAutoKey)
  constexpr int kAutoKey = 64;
  for (size_t i = 0; i < plans.size(); ++i)
    groups[{plans[i]->kind_code + (!plans[i]->sky && plans[i]->p.r_mu ? kRmuKind : 0) + (plans[i]->autop ? kAutoKey : 0),
            plans[i]->use_split ? 1 : 0}]
        .push_back(i);
  std::vector<Item> all;
  std::vector<Launch> launches;
  for (const auto &g : groups) {
    std::vector<std::pair<double, Item>> v;
    size_t lds = 0, nv = 0;
    int threads = 0;
    for (size_t i : g.second) nv += plans[i]->items_p.size();
    v.reserve(nv);
    for (size_t i : g.second) {
      const Plan &P = *plans[i];
      for (size_t k = 0; k < P.items_p.size(); ++k) {
        Item it = P.items_p[k];
        it.job = (int32_t)i;
        v.push_back({P.cost_p[k], it});
      }
      lds = std::max(lds, pair_lds_bytes((P.p.n_rp - P.row0_p) * P.p.n_pi, P.NSP, P.p.n_rp, P.ncopy_p));
      threads = P.threads;
    }
    if (v.empty()) continue;
    launches.push_back(Launch{g.first.first, g.first.second != 0, threads, lds, all.size(), v.size()});
    if (g.second.size() == 1) {
      for (const auto &x : v) all.push_back(x.second);  // one plan: its items are in cost order already
      continue;
    }
    // items of several plans, most expensive first (the grid runs in that order): a stable counting
    // sort on the cost's octave, ties in plan order
    sort_by_cost_desc(v);
    for (const auto &x : v) all.push_back(x.second);
  }
  // every plan's jackknife reduction and sparse-window count, one launch each: the tables and their
  // block offsets in one device region
  std::vector<ReduceArgs> rtab;
  std::vector<SparseArgs> stab;
  std::vector<int64_t> roff{0}, soff{0};
  for (Plan *Pp : plans) {
    const Plan &P = *Pp;
    if (P.reduce) {  // raw targets into the full sample and touch histograms
      rtab.push_back(ReduceArgs{ctx->acc_raw.p + P.off_raw, ctx->accn_raw.p + P.off_rawn, ctx->acc.p + P.off_p,
                                ctx->accn.p + P.off_pn, P.B->red_beg.p, P.B->red_lst.p, 1 + P.nreg, P.NSP, P.ncell});
      roff.push_back(roff.back() + reduce_blocks(rtab.back()));
    }
    if (!P.autop && P.co->n > 0 && P.ci->n > 0) {  // sparse-window diagnostic (corels2's starts pointer)
      stab.push_back(SparseArgs{P.co->d_z, P.co->d_inc, P.co->n, P.ci->z_sorted ? P.ci->d_z : P.ci->d_zsorted,
                                P.ci->n, P.p.box_size, window_lo(P.p), window_hi(P.p), P.p.include_prob_max,
                                ctx->accn.p + P.off_c + 2});
      soff.push_back(soff.back() + sparse_blocks(stab.back()));
    }
  }
  const size_t o_roff = rtab.size() * sizeof(ReduceArgs), o_stab = o_roff + roff.size() * sizeof(int64_t),
               o_soff = o_stab + stab.size() * sizeof(SparseArgs), aux_bytes = o_soff + soff.size() * sizeof(int64_t);
  std::vector<char> aux(aux_bytes);
  std::memcpy(aux.data(), rtab.data(), rtab.size() * sizeof(ReduceArgs));
  std::memcpy(aux.data() + o_roff, roff.data(), roff.size() * sizeof(int64_t));
  std::memcpy(aux.data() + o_stab, stab.data(), stab.size() * sizeof(SparseArgs));
  std::memcpy(aux.data() + o_soff, soff.data(), soff.size() * sizeof(int64_t));
  static_assert(sizeof(ReduceArgs) % 8 == 0 && sizeof(SparseArgs) % 8 == 0, "tables of 8-byte aligned records");
  if (check(ctx->jobs.ensure(table.size()), st, "alloc") || check(ctx->items.ensure(all.size()), st, "alloc") ||
      check(ctx->aux.ensure(aux_bytes), st, "alloc"))
    return st;
  check(hipMemcpyAsync(ctx->jobs.p, table.data(), table.size() * sizeof(KArgs), hipMemcpyHostToDevice, s), st,
        "copy");
  if (!all.empty())
    check(hipMemcpyAsync(ctx->items.p, all.data(), all.size() * sizeof(Item), hipMemcpyHostToDevice, s), st, "copy");
  check(hipMemcpyAsync(ctx->aux.p, aux.data(), aux_bytes, hipMemcpyHostToDevice, s), st, "copy");
  check(hipEventRecord(ctx->ev0, s), st, "event");
  for (const Launch &l : launches)
    check(launch_pair(l.kind % kAutoKey, l.split, l.kind >= kAutoKey, ctx->jobs.p, ctx->items.p + l.first, (int)l.n,
                      l.lds, l.threads, s),
          st, "pair kernel");
  check(hipEventRecord(ctx->ev2, s), st, "event");
  check(launch_reduce_many(reinterpret_cast<const ReduceArgs *>(ctx->aux.p),
                           reinterpret_cast<const int64_t *>(ctx->aux.p + o_roff), (int)rtab.size(), roff.back(), s),
        st, "reduce");
  check(hipEventRecord(ctx->ev1, s), st, "event");
  check(launch_sparse_many(reinterpret_cast<const SparseArgs *>(ctx->aux.p + o_stab),
                           reinterpret_cast<const int64_t *>(ctx->aux.p + o_soff), (int)stab.size(), soff.back(), s),
        st, "sparse");
  check(hipStreamSynchronize(s), st, "sync");
  if (st.code) return st;
  float kms = 0, kms_p = 0, kms_r = 0;
  (void)hipEventElapsedTime(&kms, ctx->ev0, ctx->ev1);
  (void)hipEventElapsedTime(&kms_p, ctx->ev0, ctx->ev2);
  (void)hipEventElapsedTime(&kms_r, ctx->ev2, ctx->ev1);
  for (Plan *P : plans) {
    P->kms = kms;
    P->kms_p = kms_p;
    P->kms_r = kms_r;
  }
  return st;
}

// One sky plan: the pair kernel, rerun with a larger host queue if guard-band pairs overflowed it,
// and the jackknife reduction; the queued pairs are copied back for host binning.
Status run_sky(cpc_ctx *ctx, Plan &P) {
  Status st;
  hipStream_t s = ctx->stream;
  JobBufs &B = *P.B;
  const size_t np = P.nP * P.NSP * P.ncell, npn = P.nP * P.ncell;
  if (check(ctx->items.ensure(P.items_p.size()), st, "alloc")) return st;
  if (!P.items_p.empty())
    check(hipMemcpyAsync(ctx->items.p, P.items_p.data(), P.items_p.size() * sizeof(Item), hipMemcpyHostToDevice, s),
          st, "copy");
  unsigned long long counters[4] = {0, 0, 0, 0};
  for (int attempt = 0; attempt < 2; ++attempt) {
    if (check(B.hq.ensure(P.hq_cap), st, "alloc")) return st;
    double *const accp = P.reduce ? ctx->acc_raw.p + P.off_raw : ctx->acc.p + P.off_p;
    unsigned long long *const accpn = P.reduce ? ctx->accn_raw.p + P.off_rawn : ctx->accn.p + P.off_pn;
    check(hipMemsetAsync(accp, 0, np * 8, s), st, "memset");
    check(hipMemsetAsync(accpn, 0, npn * 8, s), st, "memset");
    check(hipMemsetAsync(ctx->accn.p + P.off_c, 0, 4 * 8, s), st, "memset");
    SkyArgs A = P.sa;
    A.counters = ctx->accn.p + P.off_c;
    A.hq = B.hq.p;
    A.hq_cap = P.hq_cap;
    check(hipEventRecord(ctx->ev0, s), st, "event");
    if (!P.items_p.empty()) {
      SkyArgs C = A;
      C.n_items = (int)P.items_p.size();
      C.nslots = P.NSP;
      C.items = ctx->items.p;
      C.bgrp = P.SO->d_groups.p;
      C.tgrp = P.SI->d_groups.p;
      C.tcol = P.SI->d_colbeg.p;
      C.acc = accp;
      C.accnum = accpn;
      C.row0 = P.row0_p;
      C.ncopy = P.ncopy_p;
      check(launch_sky(P.kind_code, P.use_split, C, P.threads, sky_side(*P.SO), sky_side(*P.SI), s), st,
            "sky pair kernel");
    }
    check(hipEventRecord(ctx->ev2, s), st, "event");
    if (st.code) return st;
    check(hipMemcpyAsync(counters, ctx->accn.p + P.off_c, 4 * 8, hipMemcpyDeviceToHost, s), st, "copy");
    check(hipStreamSynchronize(s), st, "sync");
    if (st.code) return st;
    if (counters[3] > (unsigned long long)P.hq_cap) {
      if (counters[3] > (unsigned long long)INT32_MAX) return {CPC_ERR_NOMEM, "guard-band queue overflow"};
      P.hq_cap = (int)counters[3];
      continue;  // rerun with a queue large enough for every guard-band pair
    }
    break;
  }
  if (P.reduce) {  // raw targets into the full sample and touch histograms
    ReduceArgs ra{ctx->acc_raw.p + P.off_raw, ctx->accn_raw.p + P.off_rawn, ctx->acc.p + P.off_p,
                  ctx->accn.p + P.off_pn, P.B->red_beg.p, P.B->red_lst.p, 1 + P.nreg, P.NSP, P.ncell};
    check(launch_reduce(ra, s), st, "reduce");
  }
  check(hipEventRecord(ctx->ev1, s), st, "event");
  check(hipStreamSynchronize(s), st, "sync");
  if (st.code) return st;
  // the pass is deterministic, so the rerun queue holds every pair; a still larger count would mean
  // pairs lost, never a partial copy
  if (counters[3] > (unsigned long long)P.hq_cap)
    return {CPC_ERR_DEVICE, "guard-band queue still overflowing after a rerun with room for every pair"};
  (void)hipEventElapsedTime(&P.kms, ctx->ev0, ctx->ev1);
  (void)hipEventElapsedTime(&P.kms_p, ctx->ev0, ctx->ev2);
  (void)hipEventElapsedTime(&P.kms_r, ctx->ev2, ctx->ev1);
  if (counters[3] > 0) {
    P.hq_pairs.resize(std::min<unsigned long long>(counters[3], (unsigned long long)P.hq_cap));
    check(hipMemcpyAsync(P.hq_pairs.data(), B.hq.p, P.hq_pairs.size() * sizeof(uint2), hipMemcpyDeviceToHost, s), st,
          "copy");
    check(hipStreamSynchronize(s), st, "sync");
  }
  return st;
}

// The reference layout from a plan's device histograms (host copies of the arenas): the full
// sample, the touch or leave-one-out jackknife regions, num_rp, and the plan's stats.
// Returns the plan's sparse-window count.
int64_t assemble(const Plan &P, const std::vector<double> &acc, const std::vector<unsigned long long> &accn,
                 double ms_total, int threads) {
  const cpc_params &p = P.p;
  const int ncell = P.ncell, R = P.R, NV = P.NV, NE = P.NE, NSP = P.NSP, kc = P.kind_code;
  const double *accp = acc.data() + P.off_p;
  const unsigned long long *accpn = accn.data() + P.off_pn;
  const unsigned long long *counters = accn.data() + P.off_c;
  cpc_hist *out = P.out;
  const cpc_catalog *co = P.co, *ci = P.ci;

  // ---- pairs the device could not bin bit-exactly: bin them with the host libm
  HostHist HH;
  const bool host_pairs = !P.hq_pairs.empty();
  if (host_pairs) {
    // pairs evaluated on host threads a block at a time, then added in queue order
    HH.init(R, ncell);
    const int64_t nq = (int64_t)P.hq_pairs.size(), blk = 1 << 16;
    std::vector<HostPair> hp((size_t)std::min(nq, blk));
    std::vector<char> ok(hp.size());
    for (int64_t b0 = 0; b0 < nq; b0 += blk) {
      const int64_t nb = std::min(blk, nq - b0);
      parallel_for(nb, nb >= 4096 ? threads : 1, [&](int64_t k) {
        const uint2 q = P.hq_pairs[b0 + k];
        ok[k] = host_sky_pair(p, P.dv, co, (int64_t)q.x, ci, (int64_t)q.y, P.jk_on, hp[k]);
      });
      for (int64_t k = 0; k < nb; ++k)
        if (ok[k]) host_add(p, hp[k], HH);
    }
  }
  int64_t sparse = (int64_t)counters[2];
  if (P.sky && p.unsorted == 2 && !P.autop && ci->n > 0) {
    // data_sorted 2: corels2 starts each outer galaxy's scan at a per-thread pointer into the
    // Dec-ordered inner catalog that only moves forward, past galaxies outside the Dec window
    // (corr.cpp:356-362), and a scan ends at its first galaxy above the window (ret_dec,
    // corr.cpp:74-76).  The engine counts every pair of the window instead.  Replayed here in
    // catalog order, as one thread runs it: an outer galaxy whose window begins below the pointer
    // (windows narrow and widen with redshift) or after the pointer ran off an empty window loses
    // the galaxies in between; those outer galaxies are counted.
    const std::vector<double> &dec = ci->h_y;
    int64_t starts = 0;
    for (int64_t i = 0; i < co->n; ++i) {
      if (!co->h_inc.empty() && co->h_inc[i] > p.include_prob_max) continue;
      const Win &w = P.wins[i];
      while (starts < ci->n && (dec[starts] > w.decmax || dec[starts] < w.decmin)) ++starts;
      const int64_t lo = std::lower_bound(dec.begin(), dec.end(), w.decmin) - dec.begin();
      const int64_t hi = std::upper_bound(dec.begin(), dec.end(), w.decmax) - dec.begin();
      if (lo < hi && starts > lo) ++sparse;
    }
  }
  if (P.sky && ci->z_sorted && ci->n > 0) {  // sparse-window diagnostic on host threads
    const int64_t blk = 1 << 16, nblk = (co->n + blk - 1) / blk;
    std::vector<int64_t> cnt(nblk, 0);
    parallel_for(nblk, threads, [&](int64_t b) {
      for (int64_t i = b * blk; i < std::min(co->n, (b + 1) * blk); ++i) {
        const double inc = co->h_inc.empty() ? 0.0 : co->h_inc[i];
        if (inc > p.include_prob_max) continue;
        auto it = std::lower_bound(ci->h_z.begin(), ci->h_z.end(), P.wins[i].zmin);
        if (it != ci->h_z.end() && *it > P.wins[i].zmax) ++cnt[b];
      }
    });
    for (int64_t c : cnt) sparse += c;
  }

  const size_t cells_all = (size_t)R * ncell;
  std::fill(out->num, out->num + cells_all, 0);
  std::fill(out->wt_num, out->wt_num + cells_all, 0.0);
  std::fill(out->val, out->val + cells_all * CPC_NFIELD, 0.0);
  std::fill(out->err, out->err + cells_all * CPC_NFIELD, 0.0);
  std::fill(out->num_rp, out->num_rp + (size_t)R * p.n_rp, 0);
  // full sample = every primary target (+ host-binned pairs); reduced on the device into target 0
  const int n_full = P.reduce ? 1 : (int)P.nP;
  for (int t = 0; t < n_full; ++t)
    for (int c = 0; c < ncell; ++c) {
      const unsigned long long n = accpn[(size_t)t * ncell + c];
      if (!n) continue;
      out->num[c] += (int64_t)n;
      out->wt_num[c] += accp[((size_t)t * NSP + 0) * ncell + c];
      for (int v = 0; v < NV; ++v) {
        const int f = kValField[kc][v];
        out->val[(size_t)c * CPC_NFIELD + f] += accp[((size_t)t * NSP + 1 + v) * ncell + c];
        if (v < NE) out->err[(size_t)c * CPC_NFIELD + f] += accp[((size_t)t * NSP + 1 + NV + v) * ncell + c];
      }
    }
  if (host_pairs)
    for (int c = 0; c < ncell; ++c) {
      out->num[c] += HH.num[c];
      out->wt_num[c] += HH.wt[c];
      for (int f = 0; f < CPC_NFIELD; ++f) {
        out->val[(size_t)c * CPC_NFIELD + f] += HH.val[(size_t)c * CPC_NFIELD + f];
        out->err[(size_t)c * CPC_NFIELD + f] += HH.err[(size_t)c * CPC_NFIELD + f];
      }
    }
  if (P.jk_on) {
    std::vector<double> tv((size_t)ncell * CPC_NFIELD), tw(ncell);
    std::vector<int64_t> tn(ncell);
    for (int r = 0; r < p.n_jk; ++r) {
      // touch_r (jk.cpp:138-171): num, wt_num and data[].val; shape vals only when which_corr != 0.
      // The device reduced every target whose pairs touch r into downloaded histogram 1 + r.
      std::fill(tv.begin(), tv.end(), 0.0);
      std::fill(tw.begin(), tw.end(), 0.0);
      std::fill(tn.begin(), tn.end(), 0);
      const int t = 1 + r;
      for (int c = 0; c < ncell; ++c) {
        const unsigned long long n = accpn[(size_t)t * ncell + c];
        if (!n) continue;
        tn[c] += (int64_t)n;
        tw[c] += accp[((size_t)t * NSP + 0) * ncell + c];
        for (int v = 0; v < NV; ++v)
          tv[(size_t)c * CPC_NFIELD + kValField[kc][v]] += accp[((size_t)t * NSP + 1 + v) * ncell + c];
      }
      if (host_pairs)
        for (int c = 0; c < ncell; ++c) {
          const size_t o = (size_t)(r + 1) * ncell + c;
          tn[c] += HH.num[o];
          tw[c] += HH.wt[o];
          for (int f = 0; f < CPC_NFIELD; ++f) tv[(size_t)c * CPC_NFIELD + f] += HH.val[o * CPC_NFIELD + f];
        }
      if (p.which_corr == 0) std::fill(tv.begin(), tv.end(), 0.0);
      const size_t reg = (size_t)(r + 1);
      for (int c = 0; c < ncell; ++c) {
        const size_t o = reg * ncell + c;
        if (p.jk_mode == CPC_JK_TOUCH) {
          out->num[o] = tn[c];
          out->wt_num[o] = tw[c];
          for (int f = 0; f < CPC_NFIELD; ++f) out->val[o * CPC_NFIELD + f] = tv[(size_t)c * CPC_NFIELD + f];
        } else {
          // final_jk_bins: add_bins(full, touch, jk, subtract) over every field (outp.cpp:242-288)
          out->num[o] = out->num[c] + tn[c] * -1;
          out->wt_num[o] = out->wt_num[c] + tw[c] * -1;
          for (int f = 0; f < CPC_NFIELD; ++f) {
            out->val[o * CPC_NFIELD + f] = out->val[(size_t)c * CPC_NFIELD + f] + tv[(size_t)c * CPC_NFIELD + f] * -1;
            out->err[o * CPC_NFIELD + f] = out->err[(size_t)c * CPC_NFIELD + f];
          }
        }
      }
    }
  }
  int64_t nbinned = 0;
  for (int r = 0; r < R; ++r)
    for (int k = 0; k < p.n_rp; ++k) {
      int64_t sum = 0;
      for (int q = 0; q < p.n_pi; ++q) sum += out->num[((size_t)r * p.n_rp + k) * p.n_pi + q];
      out->num_rp[(size_t)r * p.n_rp + k] = sum;
      if (r == 0) nbinned += sum;
    }
  if (P.stats) {
    cpc_stats *st = P.stats;
    st->n_binned = nbinned;
    st->n_candidates = (int64_t)counters[0];
    st->n_fp64_reeval = (int64_t)counters[1];
    st->n_sparse_windows = sparse;
    st->ms_kernel = P.kms;
    st->ms_primary = P.kms_p;
    st->ms_reduce = P.kms_r;
    st->ms_build = P.ms_build;
    st->ms_cells = P.ms_cells;
    st->n_items_primary = (int64_t)P.items_p.size();
    st->n_host_pairs = (int64_t)P.hq_pairs.size();
    st->ms_total = ms_total;
  }
  return sparse;
}

// cpc_count and cpc_count_batch: check every job, prepare each, run box plans together and sky
// plans in turn, then copy the accumulators back once and assemble every output.
cpc_status run_jobs(cpc_ctx *ctx, int n_jobs, const cpc_job *jobs, cpc_stats *total, bool batch) {
  auto t_start = std::chrono::steady_clock::now();
  if (!ctx) return CPC_ERR_INVALID;
  if (n_jobs < 0 || (n_jobs > 0 && !jobs)) return fail(ctx, CPC_ERR_INVALID, "bad job list");
  auto tag = [&](int i, Status st) {
    if (batch) st.msg = "job " + std::to_string(i) + ": " + st.msg;
    return st;
  };
  std::vector<Plan> plans(n_jobs);
  for (int i = 0; i < n_jobs; ++i) {
    Status st = validate(jobs[i], plans[i]);
    if (st.code) return fail(ctx, tag(i, st));
  }
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, CPC_ERR_DEVICE, "hipSetDevice failed");
  while (ctx->pool.size() < (size_t)n_jobs) ctx->pool.push_back(std::unique_ptr<JobBufs>(new JobBufs()));
  ctx->side_q.clear();
  ctx->side_q_side.clear();
  ctx->side_q_total = 0;
  for (int i = 0; i < n_jobs; ++i) {
    plans[i].B = ctx->pool[i].get();
    Status st = setup(ctx, plans[i]);
    if (st.code) return fail(ctx, tag(i, st));
  }
  ctx->cells_timed = false;
  if (n_jobs > 0) {
    auto t_b = std::chrono::steady_clock::now();
    Status st = build_queued_sides(ctx);
    if (st.code) return fail(ctx, st);
    plans[0].ms_build += ms_since(t_b);
  }
  {
    // every side's group boundaries, once all sides are launched: the batched build copied its
    // sides' back already; the sides built one at a time copy theirs, then one wait and the group
    // tables of all sides on host threads
    Status st;
    std::vector<std::pair<Side *, size_t>> pend;  // side, its grid's cells
    for (Plan &P : plans)
      for (Side *S : {P.SO, P.SI})
        if (S->pending && (pend.empty() || pend.back().first != S)) pend.push_back({S, (size_t)P.G.ncx * P.G.ncy});
    for (const auto &q : pend) {
      Side *S = q.first;
      if (S->hb) continue;
      S->h_bounds.resize(S->nkeys + 1);
      S->hb = S->h_bounds.data();
      check(hipMemcpyAsync(S->h_bounds.data(), S->d_bounds.p, (S->nkeys + 1) * sizeof(int4), hipMemcpyDeviceToHost,
                           ctx->stream),
            st, "copy");
    }
    check(hipStreamSynchronize(ctx->stream), st, "sync");
    if (st.code) return fail(ctx, st);
    if (ctx->cells_timed && n_jobs > 0) (void)hipEventElapsedTime(&plans[0].ms_cells, ctx->evc0, ctx->evc1);
    std::vector<Status> sts(pend.size());
    parallel_for((int64_t)pend.size(), host_threads(),
                 [&](int64_t k) { sts[k] = finish_side(ctx, *pend[k].first, pend[k].second); });
    for (const Status &x : sts)
      if (x.code) return fail(ctx, x);
  }
  // work items of every plan: host work, spread over threads when the batch is large
  parallel_for(n_jobs, host_threads(), [&](int64_t i) { setup_items(plans[i], n_jobs > 1 ? 1 : host_threads()); });
  {
    Status st = stage_tables(ctx, plans);
    if (st.code) return fail(ctx, st);
  }
  for (int i = 0; i < n_jobs; ++i) {
    Status st = setup_tables(ctx, plans[i]);
    if (st.code) return fail(ctx, tag(i, st));
  }
  // accumulator arenas: per plan the downloaded f64 histograms and [counts | 4 counters]; with the
  // jackknife the raw targets in arenas of their own
  size_t nd = 0, nu = 0, nr = 0, nrn = 0;
  for (Plan &P : plans) {
    const size_t nP_dl = P.reduce ? (size_t)(1 + P.nreg) : P.nP;  // targets downloaded
    if (P.reduce) {
      P.off_raw = nr;
      nr += P.nP * P.NSP * P.ncell;
      P.off_rawn = nrn;
      nrn += P.nP * P.ncell;
    }
    P.off_p = nd;
    nd += nP_dl * P.NSP * P.ncell;
    P.off_pn = nu;
    nu += nP_dl * P.ncell;
    P.off_c = nu;
    nu += 4;
  }
  Status st;
  hipStream_t s = ctx->stream;
  if (check(ctx->acc.ensure(std::max<size_t>(nd, 1)), st, "alloc") ||
      check(ctx->accn.ensure(std::max<size_t>(nu, 1)), st, "alloc"))
    return fail(ctx, st);
  if (nd) check(hipMemsetAsync(ctx->acc.p, 0, nd * 8, s), st, "memset");
  if (nu) check(hipMemsetAsync(ctx->accn.p, 0, nu * 8, s), st, "memset");
  if (nr) {
    if (check(ctx->acc_raw.ensure(nr), st, "alloc") || check(ctx->accn_raw.ensure(nrn), st, "alloc"))
      return fail(ctx, st);
    check(hipMemsetAsync(ctx->acc_raw.p, 0, nr * 8, s), st, "memset");
    check(hipMemsetAsync(ctx->accn_raw.p, 0, nrn * 8, s), st, "memset");
  }
  if (st.code) return fail(ctx, st);
  std::vector<Plan *> box;
  for (Plan &P : plans)
    if (!P.sky) box.push_back(&P);
  st = run_box(ctx, box);
  if (st.code) return fail(ctx, st);
  for (int i = 0; i < n_jobs; ++i)
    if (plans[i].sky) {
      st = run_sky(ctx, plans[i]);
      if (st.code) return fail(ctx, tag(i, st));
    }
  std::vector<double> acc(nd);
  std::vector<unsigned long long> accn(nu);
  if (nd) check(hipMemcpyAsync(acc.data(), ctx->acc.p, nd * 8, hipMemcpyDeviceToHost, s), st, "copy");
  if (nu) check(hipMemcpyAsync(accn.data(), ctx->accn.p, nu * 8, hipMemcpyDeviceToHost, s), st, "copy");
  check(hipStreamSynchronize(s), st, "sync");
  if (st.code) return fail(ctx, st);
  const double ms_total = ms_since(t_start);
  std::vector<int64_t> sparse_of(n_jobs, 0);
  // plans on host threads in a large batch, else each plan's host-binned pairs on them
  const bool outer_par = n_jobs > 16;
  parallel_for(n_jobs, outer_par ? host_threads() : 1, [&](int64_t i) {
    sparse_of[i] = assemble(plans[i], acc, accn, ms_total, outer_par ? 1 : host_threads());
  });
  int64_t sparse = 0;
  for (int64_t v : sparse_of) sparse += v;
  const double ms_all = ms_since(t_start);  // the call's wall time, host binning and assembly included
  for (Plan &P : plans)
    if (P.stats) P.stats->ms_total = ms_all;
  if (total) {
    cpc_stats t{};
    for (const Plan &P : plans) {
      const unsigned long long *c = accn.data() + P.off_c;
      for (int k = 0; k < P.p.n_rp; ++k) t.n_binned += P.out->num_rp[k];
      t.n_candidates += (int64_t)c[0];
      t.n_fp64_reeval += (int64_t)c[1];
      t.n_items_primary += (int64_t)P.items_p.size();
      t.n_host_pairs += (int64_t)P.hq_pairs.size();
      t.ms_build += P.ms_build;
      t.ms_cells += P.ms_cells;
      if (P.sky) {
        t.ms_kernel += P.kms;
        t.ms_primary += P.kms_p;
        t.ms_reduce += P.kms_r;
      }
    }
    if (!box.empty()) {  // box plans share their launches
      t.ms_kernel += box[0]->kms;
      t.ms_primary += box[0]->kms_p;
      t.ms_reduce += box[0]->kms_r;
    }
    t.n_sparse_windows = sparse;
    t.ms_total = ms_since(t_start);
    *total = t;
  }
  ctx->err.clear();
  return CPC_OK;
}

}  // namespace

// The no-alloc scratch (cpc_count_noalloc): every device the library runs on holds this array
// until the process ends.  Host code reaches it only through cpc_ctx::scratch.
constexpr size_t kScratchBytes = (size_t)CPC_NOALLOC_SCRATCH_MIB << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

extern "C" {

const char *cpc_version(void) { return kVersion; }

const char *cpc_last_error(const cpc_ctx *ctx) { return ctx ? ctx->err.c_str() : "null context"; }

cpc_status cpc_init(cpc_ctx **out, int device) {
  if (!out) return CPC_ERR_INVALID;
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return CPC_ERR_DEVICE;
  if (device < 0 || device >= ndev) return CPC_ERR_INVALID;
  if (hipSetDevice(device) != hipSuccess) return CPC_ERR_DEVICE;
  cpc_ctx *ctx = new cpc_ctx();
  ctx->device = device;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) == hipSuccess) {
    ctx->n_cu = prop.multiProcessorCount;
    ctx->max_lds = prop.sharedMemPerBlock;
    if (prop.maxSharedMemoryPerMultiProcessor > ctx->max_lds) ctx->max_lds = prop.maxSharedMemoryPerMultiProcessor;
  }
  if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess ||
      hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess ||
      hipEventCreate(&ctx->ev2) != hipSuccess || hipEventCreate(&ctx->evc0) != hipSuccess ||
      hipEventCreate(&ctx->evc1) != hipSuccess) {
    delete ctx;
    return CPC_ERR_DEVICE;
  }
  // the no-alloc scratch: in host code `scratch` names a host-side stand-in, so its device address
  // on this device (each device holds its own copy) is looked up once, with the device current
  // (a failed lookup leaves it null: cpc_count_noalloc then reports the error, the rest works)
  void *sp = nullptr;
  if (hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch)) == hipSuccess && sp)
    ctx->scratch = (char *)sp;
  else
    (void)hipGetLastError();
  *out = ctx;
  return CPC_OK;
}

void cpc_destroy(cpc_ctx *ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  (void)hipStreamSynchronize(ctx->stream);
  ctx->keys_a.release();
  ctx->keys_b.release();
  ctx->vals_a.release();
  ctx->vals_b.release();
  ctx->order.release();
  ctx->zk_b.release();
  ctx->tables.release();
  ctx->aux.release();
  ctx->k64_a.release();
  ctx->k64_b.release();
  ctx->sv_a.release();
  ctx->sv_b.release();
  ctx->bounds_all.release();
  ctx->side_tab.release();
  ctx->hb_bounds.release();
  ctx->hb_stage.release();
  ctx->items.release();
  ctx->jobs.release();
  ctx->acc.release();
  ctx->accn.release();
  ctx->acc_raw.release();
  ctx->accn_raw.release();
  for (auto &b : ctx->pool) b->release();
  ctx->pool.clear();
  if (ctx->sort_tmp) (void)hipFree(ctx->sort_tmp);
  (void)hipEventDestroy(ctx->ev0);
  (void)hipEventDestroy(ctx->ev1);
  (void)hipEventDestroy(ctx->ev2);
  (void)hipEventDestroy(ctx->evc0);
  (void)hipEventDestroy(ctx->evc1);
  (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

cpc_status cpc_set_tuning(cpc_ctx *ctx, double col_frac, int item_size, int block_threads) {
  if (!ctx) return CPC_ERR_INVALID;
  if (col_frac < 0 || item_size < 0 || block_threads < 0) return fail(ctx, CPC_ERR_INVALID, "negative tuning value");
  if (block_threads && (block_threads % 64 || block_threads > 512))
    return fail(ctx, CPC_ERR_INVALID, "block_threads must be a multiple of 64 and at most 512");
  ctx->col_frac = col_frac;
  ctx->item_size = item_size;
  ctx->block_threads = block_threads;
  return CPC_OK;
}

cpc_status cpc_catalog_upload(cpc_ctx *ctx, const cpc_catalog_desc *d, cpc_catalog **out) {
  if (!ctx || !d || !out) return ctx ? fail(ctx, CPC_ERR_INVALID, "null argument") : CPC_ERR_INVALID;
  *out = nullptr;
  if (d->n < 0 || d->n > (int64_t)0x7fffffff - 1)
    return fail(ctx, CPC_ERR_INVALID, "catalog size must be in [0, 2^31-1)");
  if (d->n > 0 && (!d->x || !d->y || !d->z)) return fail(ctx, CPC_ERR_INVALID, "x, y and z are required");
  if ((d->e1 == nullptr) != (d->e2 == nullptr)) return fail(ctx, CPC_ERR_INVALID, "e1 and e2 must both be given");
  if ((d->dc == nullptr) != (d->da == nullptr)) return fail(ctx, CPC_ERR_INVALID, "dc and da must both be given");
  if (d->e3 && !d->e1) return fail(ctx, CPC_ERR_INVALID, "e3 needs e1 and e2");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, CPC_ERR_DEVICE, "hipSetDevice failed");
  const int64_t n = d->n;
  cpc_catalog *c = new cpc_catalog();
  c->n = n;
  c->has_e = d->e1 != nullptr;
  c->has_jk = d->jk != nullptr;
  c->has_sky = d->dc != nullptr;
  c->has_e3 = d->e3 != nullptr;
  Status st;
  auto up = [&](const double *src, double **dst) {
    if (!src || n == 0) return;
    if (check(hipMalloc((void **)dst, n * sizeof(double)), st, "alloc")) return;
    check(hipMemcpyAsync(*dst, src, n * sizeof(double), hipMemcpyHostToDevice, ctx->stream), st, "copy");
  };
  up(d->x, &c->d_x);
  up(d->y, &c->d_y);
  up(d->z, &c->d_z);
  up(d->w, &c->d_w);
  up(d->e1, &c->d_e1);
  up(d->e2, &c->d_e2);
  up(d->phi, &c->d_phi);
  up(d->e3, &c->d_e3);
  up(d->jk_prob, &c->d_jkp);
  up(d->include_prob, &c->d_inc);
  if (d->jk && n > 0) {
    if (!check(hipMalloc((void **)&c->d_jk, n * sizeof(int32_t)), st, "alloc"))
      check(hipMemcpyAsync(c->d_jk, d->jk, n * sizeof(int32_t), hipMemcpyHostToDevice, ctx->stream), st, "copy");
  }
  std::vector<SRec> sr;
  if (n > 0) {
    const double *v[3] = {d->x, d->y, d->z};
    for (int a = 0; a < 3; ++a) {
      double lo = v[a][0], hi = v[a][0];
      for (int64_t i = 1; i < n; ++i) {
        lo = std::min(lo, v[a][i]);
        hi = std::max(hi, v[a][i]);
      }
      if (!(lo == lo) || !(hi == hi)) st = {CPC_ERR_INVALID, "NaN coordinate"};
      c->lo[a] = lo;
      c->hi[a] = hi;
    }
    for (int64_t i = 1; i < n; ++i)
      if (d->z[i] < d->z[i - 1]) {
        c->z_sorted = false;
        break;
      }
    c->h_z.assign(d->z, d->z + n);
    if (d->include_prob) c->h_inc.assign(d->include_prob, d->include_prob + n);
    if (d->jk) {
      c->jk_lo = *std::min_element(d->jk, d->jk + n);
      c->jk_hi = *std::max_element(d->jk, d->jk + n);
    }
    if (!c->z_sorted && st.code == CPC_OK) {
      std::vector<double> zs(d->z, d->z + n);
      std::sort(zs.begin(), zs.end());
      if (!check(hipMalloc((void **)&c->d_zsorted, n * sizeof(double)), st, "alloc"))
        check(hipMemcpy(c->d_zsorted, zs.data(), n * sizeof(double), hipMemcpyHostToDevice), st, "copy");
    }
    if (c->has_sky && st.code == CPC_OK) {
      // per-galaxy sky quantities with the host libm (read_dat.cpp:221-222, sky_calcs.cpp:26)
      sr.resize(n);
      for (int64_t i = 0; i < n; ++i) {
        SRec &r = sr[i];
        r.dc = d->dc[i];
        r.da = d->da[i];
        r.ra_rad = d->x[i] / kR2D;
        r.dec_rad = d->y[i] / kR2D;
        r.sd = std::sin(r.dec_rad);
        r.cd = std::cos(r.dec_rad);
        r.ux = r.cd * std::cos(r.ra_rad);
        r.uy = r.cd * std::sin(r.ra_rad);
      }
      if (!check(hipMalloc((void **)&c->d_sr, n * sizeof(SRec)), st, "alloc"))
        check(hipMemcpy(c->d_sr, sr.data(), n * sizeof(SRec), hipMemcpyHostToDevice), st, "copy");
      auto keep = [&](const double *src, std::vector<double> &dst) {
        if (src) dst.assign(src, src + n);
      };
      keep(d->x, c->h_x);
      keep(d->y, c->h_y);
      keep(d->w, c->h_w);
      keep(d->e1, c->h_e1);
      keep(d->e2, c->h_e2);
      keep(d->phi, c->h_phi);
      keep(d->jk_prob, c->h_jkp);
      keep(d->dc, c->h_dc);
      keep(d->da, c->h_da);
      if (d->jk) c->h_jk.assign(d->jk, d->jk + n);
    }
  }
  check(hipStreamSynchronize(ctx->stream), st, "sync");
  if (st.code != CPC_OK) {
    cpc_catalog_free(ctx, c);
    return fail(ctx, st);
  }
  *out = c;
  return CPC_OK;
}

void cpc_catalog_free(cpc_ctx *ctx, cpc_catalog *c) {
  if (!c) return;
  if (ctx) (void)hipSetDevice(ctx->device);
  double *bufs[] = {c->d_x,   c->d_y,   c->d_z,   c->d_w,       c->d_e1, c->d_e2,
                    c->d_phi, c->d_jkp, c->d_inc, c->d_zsorted, c->d_e3};
  for (double *b : bufs)
    if (b) (void)hipFree(b);
  if (c->d_jk) (void)hipFree(c->d_jk);
  if (c->d_sr) (void)hipFree(c->d_sr);
  delete c;
}

cpc_status cpc_rand_stream(uint32_t seed, int64_t skip, int64_t n, double *out) {
  if (skip < 0 || n < 0 || (n > 0 && !out)) return CPC_ERR_INVALID;
  srand(seed);
  for (int64_t k = 0; k < skip; ++k) (void)rand();
  for (int64_t k = 0; k < n; ++k) out[k] = rand() / (RAND_MAX + 1.);
  return CPC_OK;
}

cpc_status cpc_bin_edges(const cpc_params *p, double *rp_edges, double *pi_edges) {
  if (!p || p->n_rp <= 0 || p->n_pi <= 0) return CPC_ERR_INVALID;
  if (!(p->rp_max > 0) || p->rp_min > p->rp_max) return CPC_ERR_INVALID;  // initialization.cpp:299-303
  if (rp_edges) {
    if (p->lin_bin) {
      // bin_lin: b_size = float((rmax - rmin) / (float)n), running FP32 sum from rmin
      const float b_size = (float)((p->rp_max - p->rp_min) / (float)p->n_rp);
      float temp = (float)p->rp_min;
      rp_edges[0] = temp;
      for (int i = 0; i < p->n_rp; ++i) {
        temp += b_size;
        rp_edges[i + 1] = temp;
      }
    } else {
      // bin_log: FP32 log10 limits and step, edges pow(10, running FP32 log)
      if (!(p->rp_min > 0)) return CPC_ERR_INVALID;
      float d = (float)std::log10(p->rp_min);
      const float d_max = (float)std::log10(p->rp_max);
      const float b_size = (d_max - d) / (float)p->n_rp;
      rp_edges[0] = std::pow(10.0, (double)d);
      for (int i = 0; i < p->n_rp; ++i) {
        d += b_size;
        rp_edges[i + 1] = std::pow(10.0, (double)d);
      }
    }
  }
  if (pi_edges) {
    // P_bin_ini: p_size = float(data_info.p_size), running FP32 sum from p_min
    const float p_size = (float)derive(*p).psize;
    float temp = (float)p->pi_min;
    pi_edges[0] = temp;
    for (int i = 0; i < p->n_pi; ++i) {
      temp += p_size;
      pi_edges[i + 1] = temp;
    }
  }
  return CPC_OK;
}

cpc_status cpc_box_rr(const cpc_params *p, double *wt_num) {
  if (!p || !wt_num || !(p->box_size > 0)) return CPC_ERR_INVALID;
  std::vector<double> re(p->n_rp + 1), pe(p->n_pi + 1);
  const cpc_status st = cpc_bin_edges(p, re.data(), pe.data());
  if (st != CPC_OK) return st;
  const double total_volume = std::pow(p->box_size, 3);
  for (int i = 0; i < p->n_rp; ++i) {
    // coordinates 6: a ring times the pi width; coordinates 7: a spherical shell, half of it per unit mu
    const double area = p->r_mu ? 4. / 3. * M_PI * (re[i + 1] * re[i + 1] * re[i + 1] - re[i] * re[i] * re[i])
                                : M_PI * (re[i + 1] * re[i + 1] - re[i] * re[i]);
    for (int k = 0; k < p->n_pi; ++k) {
      const double volume = p->r_mu ? area / 2 * (pe[k + 1] - pe[k]) : area * (pe[k + 1] - pe[k]);
      wt_num[(size_t)i * p->n_pi + k] = volume / total_volume;
    }
  }
  return CPC_OK;
}

cpc_status cpc_rp_thresholds(const cpc_params *p, double *thr) {
  if (!p || !thr || p->n_rp <= 0) return CPC_ERR_INVALID;
  std::vector<double> t;
  Status st = thresholds(*p, p->geometry != CPC_GEOM_SKY, 64, t);
  if (st.code) return st.code;
  std::copy(t.begin(), t.end(), thr);
  return CPC_OK;
}

cpc_status cpc_count(cpc_ctx *ctx, const cpc_catalog *co, const cpc_catalog *ci, const cpc_params *pp, cpc_hist *out,
                     cpc_stats *stats) {
  const cpc_job job{co, ci, pp, out, stats};
  return run_jobs(ctx, 1, &job, nullptr, false);
}

cpc_status cpc_count_noalloc(cpc_ctx *ctx, const cpc_catalog *co, const cpc_catalog *ci, const cpc_params *pp,
                             cpc_hist *out, cpc_stats *stats) {
  if (!ctx) return CPC_ERR_INVALID;
  if (!ctx->scratch) return fail(ctx, CPC_ERR_DEVICE, "no-alloc scratch unavailable on this device");
  if (t_arena) return fail(ctx, CPC_ERR_INVALID, "cpc_count_noalloc called from inside another no-alloc call");
  Arena A;
  A.base = ctx->scratch;
  A.cap = kScratchBytes;
  // the context's own buffers step aside for the call: every device buffer of the pass is then
  // borrowed from the scratch, and none of them outlives the call
  cpc_ctx::Swap saved;
  ctx->swap_buffers(saved);
  const int64_t nmax = std::max(co ? co->n : 0, ci ? ci->n : 0);
  size_t sort_need = 0;
  cpc_status rc = CPC_OK;
  if (sort_temp_bytes(std::max<int64_t>(nmax, 1), &sort_need) != hipSuccess) {
    rc = fail(ctx, CPC_ERR_DEVICE, "sort size query failed");
  } else {
    ctx->sort_tmp = A.take(sort_need);
    ctx->sort_tmp_bytes = ctx->sort_tmp ? sort_need : 0;
    t_arena = &A;
    const cpc_job job{co, ci, pp, out, stats};
    rc = run_jobs(ctx, 1, &job, nullptr, false);
    t_arena = nullptr;
  }
  ctx->swap_buffers(saved);  // the borrowed buffers go back with `saved` and are dropped
  saved.drop();
  if (A.overflow) {
    (void)hipStreamSynchronize(ctx->stream);  // work launched before the shortfall ends first
    return fail(ctx, CPC_ERR_NOMEM,
                "the pass needs more device memory than the no-alloc scratch holds (" +
                    std::to_string(kScratchBytes >> 20) + " MiB); use cpc_count");
  }
  return rc;
}

cpc_status cpc_count_batch(cpc_ctx *ctx, int n_jobs, const cpc_job *jobs, cpc_stats *total) {
  return run_jobs(ctx, n_jobs, jobs, total, true);
}

cpc_status cpc_autotune(cpc_ctx *ctx, int n_jobs, const cpc_job *jobs, cpc_tuning *best, int *tried) {
  if (!ctx) return CPC_ERR_INVALID;
  if (!best) return fail(ctx, CPC_ERR_INVALID, "null argument");
  bool any_box = false, any_sky = false;
  for (int i = 0; i < n_jobs && jobs; ++i)
    if (jobs[i].params) (jobs[i].params->geometry == CPC_GEOM_SKY ? any_sky : any_box) = true;
  // candidate values per knob; the sky kernel runs at most 256 threads per group
  const std::vector<double> fracs = any_sky && !any_box ? std::vector<double>{0.0, 0.35, 0.75}
                                                        : std::vector<double>{0.0, 0.35, 0.5};
  const std::vector<int> items = any_sky && !any_box ? std::vector<int>{0, 256, 1024} : std::vector<int>{0, 512, 4096};
  const std::vector<int> threads = any_sky ? std::vector<int>{0, 128} : std::vector<int>{0, 256};
  cpc_tuning cur{0.0, 0, 0};
  double cur_ms = 0;
  int runs = 0;
  auto time_of = [&](const cpc_tuning &t, double &ms) -> cpc_status {
    ctx->col_frac = t.col_frac;
    ctx->item_size = t.item_size;
    ctx->block_threads = t.block_threads;
    cpc_stats tot{};
    const cpc_status rc = run_jobs(ctx, n_jobs, jobs, &tot, true);
    ++runs;
    ms = tot.ms_kernel;
    return rc;
  };
  cpc_status rc = time_of(cur, cur_ms);
  for (int knob = 0; knob < 3 && rc == CPC_OK; ++knob) {
    const size_t nv = knob == 0 ? fracs.size() : knob == 1 ? items.size() : threads.size();
    cpc_tuning base = cur;
    for (size_t v = 1; v < nv && rc == CPC_OK; ++v) {
      cpc_tuning t = base;
      if (knob == 0) t.col_frac = fracs[v];
      if (knob == 1) t.item_size = items[v];
      if (knob == 2) t.block_threads = threads[v];
      double ms = 0;
      rc = time_of(t, ms);
      if (rc == CPC_OK && ms < 0.97 * cur_ms) {
        cur = t;
        cur_ms = ms;
      }
    }
  }
  if (rc != CPC_OK) {
    ctx->col_frac = 0.0;
    ctx->item_size = 0;
    ctx->block_threads = 0;
    return rc;
  }
  // leave the outputs as the chosen tuning computes them
  ctx->col_frac = cur.col_frac;
  ctx->item_size = cur.item_size;
  ctx->block_threads = cur.block_threads;
  double ms = 0;
  rc = time_of(cur, ms);
  *best = cur;
  if (tried) *tried = runs;
  return rc;
}

cpc_status cpc_libm_error(cpc_ctx *ctx, int fn, double lo, double hi, int64_t n, double *max_ulp) {
  if (!ctx || !max_ulp) return fail(ctx, CPC_ERR_INVALID, "null argument");
  if ((fn != 0 && fn != 1) || !(lo <= hi) || !std::isfinite(lo) || !std::isfinite(hi) || n < 1 || n > (1 << 26))
    return fail(ctx, CPC_ERR_INVALID, "cpc_libm_error needs fn 0 or 1, finite lo <= hi and 1 <= n <= 2^26");
  if (fn == 1 && (lo < -1 || hi > 1)) return fail(ctx, CPC_ERR_INVALID, "acos arguments lie in [-1, 1]");
  if (hipSetDevice(ctx->device) != hipSuccess) return fail(ctx, CPC_ERR_DEVICE, "hipSetDevice failed");
  std::vector<double> x(n), y(n);
  for (int64_t k = 0; k < n; ++k) x[k] = n == 1 ? lo : lo + (hi - lo) * ((double)k / (double)(n - 1));
  Status st;
  DevBuf<double> dx, dy;
  if (check(dx.ensure(n), st, "alloc") || check(dy.ensure(n), st, "alloc")) {
    dx.release();
    dy.release();
    return fail(ctx, st);
  }
  check(hipMemcpyAsync(dx.p, x.data(), n * sizeof(double), hipMemcpyHostToDevice, ctx->stream), st, "copy");
  check(launch_libm(fn, dx.p, n, dy.p, ctx->stream), st, "libm kernel");
  check(hipMemcpyAsync(y.data(), dy.p, n * sizeof(double), hipMemcpyDeviceToHost, ctx->stream), st, "copy");
  check(hipStreamSynchronize(ctx->stream), st, "sync");
  dx.release();
  dy.release();
  if (st.code) return fail(ctx, st);
  double worst = 0;
  for (int64_t k = 0; k < n; ++k) {
    const double h = fn == 0 ? std::cos(x[k]) : std::acos(x[k]);
    const double a = std::fabs(h);
    const double ulp = std::nextafter(a, INFINITY) - a;  // spacing above |h|
    worst = std::max(worst, std::fabs(y[k] - h) / (ulp > 0 ? ulp : std::numeric_limits<double>::denorm_min()));
  }
  *max_ulp = worst;
  return CPC_OK;
}

cpc_status cpc_fp32_error(cpc_ctx *ctx, float s_lo, float s_hi, double out[2]) {
  if (!ctx || !out) return fail(ctx, CPC_ERR_INVALID, "null argument");
  if (!(s_lo > 0) || !(s_hi >= s_lo) || !std::isfinite(s_hi))
    return fail(ctx, CPC_ERR_INVALID, "cpc_fp32_error needs 0 < s_lo <= s_hi < inf");
  Status st;
  uint32_t b0, b1;
  std::memcpy(&b0, &s_lo, 4);
  std::memcpy(&b1, &s_hi, 4);
  DevBuf<unsigned long long> d;
  if (check(d.ensure(2), st, "alloc")) return fail(ctx, st);
  unsigned long long h[2] = {0, 0};
  check(hipMemcpyAsync(d.p, h, sizeof(h), hipMemcpyHostToDevice, ctx->stream), st, "copy");
  check(launch_fp32_error(b0, b1 - b0 + 1, d.p, ctx->stream), st, "fp32 error kernel");
  check(hipMemcpyAsync(h, d.p, sizeof(h), hipMemcpyDeviceToHost, ctx->stream), st, "copy");
  check(hipStreamSynchronize(ctx->stream), st, "sync");
  d.release();
  if (st.code) return fail(ctx, st);
  for (int k = 0; k < 2; ++k) std::memcpy(&out[k], &h[k], 8);
  return CPC_OK;
}

}  // extern "C"
