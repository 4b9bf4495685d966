// corr_gpu.cpp -- the reference-side binding of the MI355X pair engine.
//
// A corr_pc maintainer adds this file to corr_pc's sources (compiled with -DCPC_GPU, include path
// <repo>/include, linked with -lcpc) and calls corels2_gpu / corels1_gpu where do_corr.cpp calls
// corels2 / corels1 (do_corr.cpp:63,70,76,78,110,115,125,131).  Both keep the reference's
// signatures (corr.h:27-30) and semantics: the operand roles and the larger-catalog-outer swap of
// corels2 (corr.cpp:226-296), the pair kind of choose_corel_func (initialization.cpp:490-543),
// include_prob_max of calc_temp_ini (initialization.cpp:597-599), accumulation into `bjk` as
// comb_thread_bins does (corr.cpp:88-100), and final_jk_bins (jk.cpp:74-81) at the end.  The return
// value is the reference's n_in, the number of binned pairs.
//
// It needs only data_def.h and jk.h of the reference (both free of MPI and GSL) and include/cpc.h.
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <utility>
#include <vector>

#include "cpc.h"
#include "data_def.h"
#include "jk.h"

namespace {

cpc_ctx *g_ctx = nullptr;

void ensure_context() {
  if (!g_ctx && cpc_init(&g_ctx, 0) != CPC_OK) throw std::runtime_error("cpc_init: no usable HIP device");
}

void check(cpc_status rc) {
  if (rc != CPC_OK) throw std::runtime_error(cpc_last_error(g_ctx));
}

// The AoS `gal` records (data_def.h:23-44) as the engine's SoA descriptor, uploaded once.  `ne` is
// the number of e[] components the catalog carries: 0 (no e pointer), 1 (a convergence), 2 (a
// shape) or 3 (an ED axis).
cpc_catalog *upload(gal *g, int n, int ne, bool sky) {
  std::vector<double> x(n), y(n), z(n), w(n), phi(n), jkp(n), inc(n), e1, e2, e3, dc, da;
  std::vector<int32_t> jk(n);
  if (ne >= 1) {
    e1.assign(n, 0.0);
    e2.assign(n, 0.0);
  }
  if (ne >= 3) e3.resize(n);
  if (sky) {
    dc.resize(n);
    da.resize(n);
  }
  for (int i = 0; i < n; ++i) {
    x[i] = g[i].ra.val_deg;
    y[i] = g[i].dec.val_deg;
    z[i] = g[i].redshift;
    w[i] = g[i].wt;
    phi[i] = g[i].phi;
    jk[i] = g[i].jk;
    jkp[i] = g[i].jk_prob;
    inc[i] = g[i].include_prob;
    if (ne >= 1) e1[i] = g[i].e[0];
    if (ne >= 2) e2[i] = g[i].e[1];  // a convergence keeps e2 = 0
    if (ne >= 3) e3[i] = g[i].e[2];
    if (sky) {
      dc[i] = g[i].DC;
      da[i] = g[i].DA;
    }
  }
  cpc_catalog_desc d{};
  d.n = n;
  d.x = x.data();
  d.y = y.data();
  d.z = z.data();
  d.w = w.data();
  d.e1 = ne >= 1 ? e1.data() : nullptr;
  d.e2 = ne >= 1 ? e2.data() : nullptr;
  d.phi = phi.data();
  d.jk = jk.data();
  d.jk_prob = jkp.data();
  d.include_prob = inc.data();
  d.dc = sky ? dc.data() : nullptr;
  d.da = sky ? da.data() : nullptr;
  d.e3 = ne >= 3 ? e3.data() : nullptr;
  cpc_catalog *c = nullptr;
  check(cpc_catalog_upload(g_ctx, &d, &c));
  return c;
}

// e[] components of a catalog as corr_pc's reader allocates them (read_dat.cpp:274-304): a shape
// carries two, an ED axis (which_corr 5) three, a convergence one; catalogs read without an e file
// carry none (read_dat.cpp:126).  For which_corr 4 only the g_with_shape side is a shape (`shaped`).
int components(const gal *g, int n, const data_info &di, bool shaped) {
  if (n == 0 || g[0].e == nullptr) return 0;
  if (di.which_corr == 5) return 3;
  if (di.which_corr == 3) return 1;
  if (di.which_corr == 4) return shaped ? 2 : 1;
  return 2;
}

// choose_corel_func (initialization.cpp:490-543) for which_corr 0-5
int pair_kind(const data_info &di) {
  if (di.do_RsRs || di.do_RsRd) return CPC_KIND_DD;
  switch (di.which_corr) {
    case 0:
      return CPC_KIND_DD;
    case 1:
      return di.do_SS ? CPC_KIND_SHAPE_SHAPE : CPC_KIND_SHAPE_DENSITY;
    case 2: {
      int k = (!di.do_SD && !di.do_SS) ? CPC_KIND_SHAPE_DENSITY : CPC_KIND_SHAPE_SHAPE;
      if (di.coordinates == 2 || di.coordinates == 4) k = (di.do_DRs || di.do_SD) ? CPC_KIND_SHAPE_SHAPE : CPC_KIND_DD;
      return k;
    }
    case 3:
      return CPC_KIND_DENSITY_CONVERGENCE;
    case 4:
      return CPC_KIND_SHAPE_CONVERGENCE;
    case 5:
      return CPC_KIND_ED;
    default:
      throw std::runtime_error("which_corr outside the engine (the WL modes 7-10 stay on the CPU)");
  }
}

// The engine's parameters for the pass data_info describes; g_with_shape as the caller leaves it.
cpc_params make_params(const data_info &di, bool autop) {
  const bool sky = !di.periodic_box;
  cpc_params p{};
  p.geometry = sky ? CPC_GEOM_SKY : CPC_GEOM_BOX;
  p.box_size = di.periodic_box_size;
  p.kind = pair_kind(di);
  p.g_with_shape = di.g_with_shape;
  // shape_density_corel rotates g1's shape unless g_with_shape is 2 (corels.cpp:216-219)
  if (p.kind == CPC_KIND_SHAPE_DENSITY && p.g_with_shape == 0) p.g_with_shape = 1;
  p.estimator = di.estimator;
  p.which_corr = di.which_corr;
  p.n_rp = di.n_bins;
  p.rp_min = di.bin_r_min;
  p.rp_max = di.bin_r_max;
  p.lin_bin = di.lin_bin ? 1 : 0;
  p.n_pi = di.n_p_bin;
  p.pi_min = di.p_min;
  p.pi_max = di.p_max;
  p.do_jk = di.do_jk;
  p.n_jk = di.n_jk;
  // calc_temp_ini (initialization.cpp:597-599)
  p.include_prob_max = (di.do_RsRs || di.do_RsRd) ? di.include_prob_RR : 1.0;
  p.precision = CPC_PREC_SPLIT;
  p.jk_mode = CPC_JK_TOUCH;  // raw touch sums; final_jk_bins runs below, as in corels2
  p.auto_pairs = autop ? 1 : 0;
  p.unsorted = di.data_sorted == 0 ? 1 : 0;
  if (sky) {
    // the distance table da_Z filled (initialization.cpp:23-60, 255-268)
    p.sky_H = di.H_z;
    p.sky_dA = di.dA_z;
    p.sky_n = (int64_t)std::ceil((di.z_max + 1) / di.dz);
    p.sky_dz = di.dz;
    p.z_max = di.z_max;
    p.z_sep_max = di.z_sep_max;
    p.coordinates = di.coordinates;
    p.z_min = di.z_min;
    p.z_sep_min = di.z_sep_min;
  } else {
    p.r_mu = di.coordinates == 7 ? 1 : 0;
  }
  return p;
}

// One pass on the device, added into bjk as comb_thread_bins adds a thread's bins (corr.cpp:88-100)
// (region 0 is bjk.b, region r the touch sums of jackknife region r-1), then final_jk_bins.
long count_into(gal *g1, int n1, gal *g2, int n2, bool same, bin_jk &bjk, data_info &di, bool autop) {
  ensure_context();
  const bool sky = !di.periodic_box;
  const cpc_params p = make_params(di, autop);
  cpc_catalog *co = upload(g1, n1, components(g1, n1, di, di.g_with_shape == 1), sky);
  cpc_catalog *ci = same ? co : upload(g2, n2, components(g2, n2, di, di.g_with_shape == 2), sky);
  const int R = di.do_jk == 1 ? 1 + di.n_jk : 1;
  const size_t cells = (size_t)R * di.n_bins * di.n_p_bin;
  std::vector<int64_t> num(cells), num_rp((size_t)R * di.n_bins);
  std::vector<double> wt(cells), val(cells * CPC_NFIELD), err(cells * CPC_NFIELD);
  cpc_hist h{num.data(), wt.data(), val.data(), err.data(), num_rp.data()};
  cpc_stats st{};
  const cpc_status rc = cpc_count(g_ctx, co, ci, &p, &h, &st);
  if (ci != co) cpc_catalog_free(g_ctx, ci);
  cpc_catalog_free(g_ctx, co);
  check(rc);
  for (int r = 0; r < R; ++r) {
    bin *b = r == 0 ? bjk.b : bjk.jk[r - 1].b;
    for (int k = 0; k < di.n_bins; ++k) {
      b[k].num += num_rp[(size_t)r * di.n_bins + k];
      for (int q = 0; q < di.n_p_bin; ++q) {
        const size_t c = ((size_t)r * di.n_bins + k) * di.n_p_bin + q;
        P_bin &pb = b[k].p_bin[q];
        pb.num += num[c];
        pb.wt_num += wt[c];
        for (int f = 1; f < CPC_NFIELD; ++f) {
          pb.data[f].val += val[c * CPC_NFIELD + f];
          pb.data[f].err += err[c * CPC_NFIELD + f];
        }
      }
    }
  }
  if (di.do_jk == 1) final_jk_bins(bjk, di);
  return (long)st.n_binned;
}

}  // namespace

// corels2 (corr.cpp:219-459) on the GPU.
long corels2_gpu(gal g1[], gal g2[], bin_jk &bjk, data_info &di) {
  // operand roles (corr.cpp:226-271): catalog sizes and which side carries the shape
  int n1 = 0, n2 = 0;
  if (di.do_DRs) {
    n2 = di.n_density;
    n1 = di.n_shape_rand;
    di.g_with_shape = 0;
    if (di.coordinates == 2 || di.coordinates == 4) di.g_with_shape = 1;
  } else if (di.do_RsRd) {
    n2 = di.n_shape_rand;
    n1 = di.n_density_rand;
    di.g_with_shape = 0;
    if (di.coordinates == 2 || di.coordinates == 4) di.g_with_shape = 2;
  } else if (di.do_RsRs) {
    n1 = di.n_shape_rand;
    n2 = n1;
    di.g_with_shape = 0;
    if (di.coordinates == 2 || di.coordinates == 4) di.g_with_shape = 1;
  } else if (di.do_SRd) {
    n2 = di.n_shape;
    n1 = di.n_density_rand;
    di.g_with_shape = 2;
  } else if (di.do_SD) {
    n2 = di.n_shape;
    n1 = di.n_density;
    di.g_with_shape = 2;
  } else if (di.do_SRs) {
    n2 = di.n_shape;
    n1 = di.n_shape_rand;
    di.g_with_shape = 2;
  }
  if (n1 == 0 || n2 == 0) return 0;  // corr.cpp:273-277
  // the larger catalog becomes outer, which flips the shape side (corr.cpp:281-296)
  if (n2 > n1 && !(di.coordinates == 3 && di.sig_crit != 0)) {
    std::swap(g1, g2);
    std::swap(n1, n2);
    if (di.g_with_shape == 2)
      di.g_with_shape = 1;
    else if (di.g_with_shape == 1)
      di.g_with_shape = 2;
  }
  return count_into(g1, n1, g2, n2, g1 == g2, bjk, di, false);
}

// corels1 (corr.cpp:103-216) on the GPU: each unordered pair of one catalog once, with its roles
// (corr.cpp:112-118): g_with_shape 1, the shape catalog or the shape randoms.
long corels1_gpu(gal g1[], bin_jk &bjk, data_info &di) {
  di.g_with_shape = 1;
  int n = 0;
  if (di.do_SS)
    n = di.n_shape;
  else if (di.do_RsRs)
    n = di.n_shape_rand;
  if (n == 0) return 0;
  return count_into(g1, n, g1, n, true, bjk, di, true);
}
