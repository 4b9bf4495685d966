// cpc_out.cpp -- the output side of corr_pc (do_bining), on the histograms cpc_count returns.
//
// Restates, for which_corr 0-5 and estimators 0-4:
//   final_calc / final_calc_bins / int_calc    bins_calcs.cpp:30-52, 94-256, 331-383
//   final_calc_auto / re_bin                   bins_calcs.cpp:54-91, 465-527
//   final_calc_mean_projected (estimator 4)    bins_calcs.cpp:258-329
//   RR_PB_calc (analytic periodic-box RR)      PB_calcs.cpp:63-85
//   jk_final / jk_final2D                      bins_calcs.cpp:393-462
//   outp_bins / outp_2Dbins                    outp.cpp:45-74, 291-321
//   do_bining, both branches                   do_corr.cpp:203-433 (file names and order)
// with the reference's operation order and its C++ stream formatting (default precision: six
// significant digits), so that the same histograms give the same files.  Host code only.

#include <cmath>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cpc.h"

namespace {

constexpr int kF = 7;  // data[] fields: 0 = W, 1 = e+, 2 = ex, 3 = e++, 4 = exx, 5 = e+x, 6 = theta

// data_info.p_size: the pi range over its bins, 1 for the sky's projected coordinates 3 and 5
// (initialization.cpp:305-308)
double pi_size(const cpc_params &p) {
  if (p.geometry == CPC_GEOM_SKY && (p.coordinates == 3 || p.coordinates == 5)) return 1;
  return (p.pi_max - p.pi_min) / p.n_pi;
}

struct PBin {
  double p_min = 0, p_max = 0;
  long num = 0;
  double wt_num = 0;
  double val[kF] = {0}, err[kF] = {0};
  double ls_err = 0;
};

// The reference's `bin` (data_def.h:109-121) with its P_bin row.
struct Bin {
  double b_min = 0, b_max = 0;
  long num = 0;
  double wt_num = 0, sig_crit = 0;
  std::vector<PBin> p;
  double val[kF] = {0}, err[kF] = {0};
  double ls_err_int = 0;
};

using Bins = std::vector<Bin>;

struct Wt {
  double val = 0, err = 0;
};

// Empty bins with the reference's FP32-stepped edges (bin_log / bin_lin / P_bin_ini).
Bins empty_bins(const cpc_params &p) {
  std::vector<double> re(p.n_rp + 1), pe(p.n_pi + 1);
  cpc_bin_edges(&p, re.data(), pe.data());
  Bins b(p.n_rp);
  for (int i = 0; i < p.n_rp; ++i) {
    b[i].b_min = re[i];
    b[i].b_max = re[i + 1];
    b[i].p.resize(p.n_pi);
    for (int k = 0; k < p.n_pi; ++k) {
      b[i].p[k].p_min = pe[k];
      b[i].p[k].p_max = pe[k + 1];
    }
  }
  return b;
}

// Region r of a cpc_hist as the reference holds it after corels2: P_bin num, wt_num, data[] and the
// per-rp pair count (its other per-rp fields are still zero at that point).
Bins load_region(const cpc_params &p, const cpc_hist *h, int r) {
  Bins b = empty_bins(p);
  for (int i = 0; i < p.n_rp; ++i) {
    b[i].num = (long)h->num_rp[(size_t)r * p.n_rp + i];
    for (int k = 0; k < p.n_pi; ++k) {
      const size_t c = ((size_t)r * p.n_rp + i) * p.n_pi + k;
      PBin &q = b[i].p[k];
      q.num = (long)h->num[c];
      q.wt_num = h->wt_num[c];
      for (int j = 0; j < kF; ++j) {
        q.val[j] = h->val[c * kF + j];
        q.err[j] = h->err[c * kF + j];
      }
    }
  }
  return b;
}

// RR_PB_calc (PB_calcs.cpp:63-85): the expected fraction of pairs per bin, a ring times the pi width
// (coordinates 6) or half a spherical shell per unit mu (coordinates 7, the reference's formula).
void rr_periodic(Bins &rr, const cpc_params &p) {
  const double total_volume = std::pow(p.box_size, 3);
  for (Bin &b : rr) {
    const double area = p.r_mu ? 4. / 3. * M_PI * (b.b_max * b.b_max * b.b_max - b.b_min * b.b_min * b.b_min)
                               : M_PI * (b.b_max * b.b_max - b.b_min * b.b_min);
    b.num = 1;
    for (PBin &q : b.p) {
      const double volume = p.r_mu ? area / 2 * (q.p_max - q.p_min) : area * (q.p_max - q.p_min);
      q.wt_num = volume / total_volume;
      q.num = 1;
    }
  }
}

struct Norm {
  Wt SD, SR, DR, RR;
};

// final_calc_bins (bins_calcs.cpp:94-256) for the kinds of which_corr 0-5.
void final_calc_bins(Bins &sd, Bins &sr, Bins &dr, Bins &rr, Bins &fin, const Norm &N, const cpc_params &p,
                     bool periodic) {
  const double SD_wt = N.SD.val, SR_wt = N.SR.val, DR_wt = N.DR.val, RR_wt = N.RR.val;
  for (int i = 0; i < p.n_rp; ++i) {
    for (int k = 0; k < p.n_pi; ++k) {
      fin[i].num += sd[i].p[k].num;
      const double sd_num = (double)sd[i].p[k].num;
      const double sr_num = (double)sr[i].p[k].num;
      const double rr_num = (double)rr[i].p[k].num;
      const double sd_wt_num = sd[i].p[k].wt_num;
      const double sr_wt_num = sr[i].p[k].wt_num;
      const double dr_wt_num = dr[i].p[k].wt_num;
      const double rr_wt_num = rr[i].p[k].wt_num;
      if (rr_num == 0) continue;
      const double sd_num_rel = (sd_wt_num / rr_wt_num) * (RR_wt / SD_wt);
      const double sr_num_rel = (sr_wt_num / rr_wt_num) * (RR_wt / SR_wt);
      const double dr_num_rel = (dr_wt_num / rr_wt_num) * (RR_wt / DR_wt);
      double w_dp = sd_num_rel - sr_num_rel - dr_num_rel + 1.0;
      if (periodic) w_dp = sd_num_rel - 1.0;
      fin[i].p[k].val[0] = w_dp;
      sd[i].p[k].val[0] = sd_num_rel;
      sr[i].p[k].val[0] = sr_num_rel;
      dr[i].p[k].val[0] = dr_num_rel;
      rr[i].p[k].val[0] = 1.0;
      if (p.which_corr == 0) continue;
      for (int j = 1; j < kF; ++j) {
        const double sd_val = sd[i].p[k].val[j];
        const double sr_val = sr[i].p[k].val[j];
        double sd_val_norm = 0, sr_val_norm = 0;
        if (p.which_corr == 5 && j == 6) {
          // ED: the mean cos^2 over the SD pairs minus its isotropic value 1/3 (bins_calcs.cpp:190-196)
          if (sd_num > 0) sd_val_norm = (sd_val / sd_wt_num);
          sd_val_norm -= 1. / 3;
        } else {
          sd_val_norm = (sd_val / rr_wt_num) * (RR_wt / SD_wt);
          sr_val_norm = (sr_val / rr_wt_num) * (RR_wt / SR_wt);
        }
        double val_f = sd_val_norm - sr_val_norm;
        if (sd_num == 0 || sr_num == 0 || rr_num == 0) {
          val_f = 0;
          if (periodic) val_f = sd_val_norm;
        }
        fin[i].p[k].val[j] = val_f;
        sd[i].p[k].val[j] = sd_val;  // the raw sum, not the mean (bins_calcs.cpp:189,235)
        sr[i].p[k].val[j] = sr_val_norm;
        if (j == 6) {
          sd[i].p[k].val[j] = sd_num == 0 ? 0 : sd_val;
          sr[i].p[k].val[j] = sr_num == 0 ? 0 : sr_val;
        }
      }
    }
  }
}

// int_calc (bins_calcs.cpp:331-358): pi integration of one bin array.
void int_calc(Bins &dat, const cpc_params &p) {
  const double dC = pi_size(p);  // data_inf.p_size
  for (Bin &b : dat) {
    for (const PBin &q : b.p) {
      b.wt_num += q.wt_num;
      for (int j = 0; j < kF; ++j) {
        b.val[j] += q.val[j] * dC;
        b.err[j] += std::pow(q.err[j] * dC, 2);
      }
      b.ls_err_int += std::pow(q.ls_err * dC, 2);
    }
    for (int j = 0; j < kF; ++j) b.err[j] = std::sqrt(b.err[j]);
    b.ls_err_int = std::sqrt(b.ls_err_int);
  }
}

// final_calc with four bin arrays (bins_calcs.cpp:30-52): pair-count normalisations from the
// catalogs' weight sums, the analytic RR in a periodic box, the estimator, the pi integration.
// N receives the normalisations, which the reference keeps in data_info.{SD,DRs,SRd,RsRd}_wt.
void final_calc(Bins &sd, Bins &sr, Bins &dr, Bins &rr, Bins &fin, const cpc_norm &n, const cpc_params &p, Norm &N) {
  const bool periodic = p.geometry == CPC_GEOM_BOX;
  N.SD.val = n.S.val * n.D.val;
  N.SD.err = N.SD.val * std::sqrt(n.S.err / std::pow(n.S.val, 2) + n.D.err / std::pow(n.D.val, 2));
  N.DR.val = n.Rs.val * n.D.val;
  N.DR.err = N.DR.val * std::sqrt(n.Rs.err / std::pow(n.Rs.val, 2) + n.D.err / std::pow(n.D.val, 2));
  N.SR.val = n.S.val * n.Rd.val;
  N.SR.err = N.SR.val * std::sqrt(n.S.err / std::pow(n.S.val, 2) + n.Rd.err / std::pow(n.Rd.val, 2));
  N.RR.val = n.Rs.val * n.Rd.val * n.include_prob_RR * n.include_prob_RR;
  N.RR.err = N.RR.val * std::sqrt(n.Rs.err / std::pow(n.Rs.val, 2) + n.Rd.err / std::pow(n.Rd.val, 2));
  if (periodic) {
    N.RR.val = 1;
    N.RR.err = 0;
    rr_periodic(rr, p);
  }
  final_calc_bins(sd, sr, dr, rr, fin, N, p, periodic);
  // int_calc(sd, sr, dr, rr, final) (bins_calcs.cpp:360-371)
  int_calc(sd, p);
  int_calc(sr, p);
  int_calc(dr, p);
  int_calc(rr, p);
  int_calc(fin, p);
}

// jk_final and jk_final2D (bins_calcs.cpp:393-462): value of the full sample, error
// sqrt((n-1)/n * sum over regions of (x_r - mean)^2).
void jk_final(Bins &out, const Bins &fin, const std::vector<Bins> &jk, const cpc_params &p) {
  const int n = (int)jk.size();
  for (int i = 0; i < p.n_rp; ++i) {
    for (int j = 0; j < kF; ++j) {  // data_inf.n_bin_var = 7 (initialization.cpp:141)
      double var = 0, mean = 0;
      for (int k = 0; k < n; ++k) mean += jk[k][i].val[j];
      mean /= n;
      for (int k = 0; k < n; ++k) var += std::pow(jk[k][i].val[j] - mean, 2.0);
      var *= ((n - 1.0) / (n));
      out[i].val[j] = fin[i].val[j];
      out[i].err[j] = std::sqrt(var);
    }
    for (int l = 0; l < p.n_pi; ++l)
      for (int j = 0; j < kF; ++j) {
        double var = 0, mean = 0;
        for (int k = 0; k < n; ++k) mean += jk[k][i].p[l].val[j];
        mean /= n;
        for (int k = 0; k < n; ++k) var += std::pow(jk[k][i].p[l].val[j] - mean, 2.0);
        var *= ((n - 1.0) / (n));
        out[i].p[l].val[j] = fin[i].p[l].val[j];
        out[i].p[l].err[j] = std::sqrt(var);
      }
  }
}

// re_bin (bins_calcs.cpp:465-527): folds every negative-pi bin of each rp row onto the positive bin
// that holds its mirrored centre, the bins' stored (FP32-stepped) edges deciding which one.  A pair
// of bins whose edges do not mirror to within p_size/100 is reported and left alone, as there.
void re_bin(Bins &ss, const cpc_params &p) {
  if (p.n_pi <= 1) return;
  const double p_size = pi_size(p);
  for (int i = 0; i < p.n_rp; ++i) {
    std::vector<PBin> &q = ss[i].p;
    for (int k = 0; k < p.n_pi; ++k) {
      if (q[k].p_min >= 0) break;  // only the negative side moves
      const double p_val = 0.0 - (q[k].p_min + q[k].p_max) / 2.0;
      int t = (int)std::floor((p_val - p.pi_min) / p_size);
      // walk to the bin whose edges hold p_val; the edges are contiguous, so the walk is monotone
      // and ends within n_pi steps (the bound only guards against a malformed edge table)
      bool found = false;
      for (int step = 0; step <= p.n_pi + 1; ++step) {
        if (t < 0 || t >= p.n_pi) break;
        if (p_val < q[t].p_min) {
          --t;
        } else if (p_val > q[t].p_max) {
          ++t;
        } else {
          found = true;
          break;
        }
      }
      if (!found) break;  // the reference leaves the row here (bins_calcs.cpp:483-486)
      if (std::fabs(q[k].p_min + q[t].p_max) > (p_size / 100.) || std::fabs(q[k].p_max + q[t].p_min) > (p_size / 100.) ||
          q[t].p_max < 0) {
        std::cout << "rebinning problem:" << i << "   " << t << "   " << k << "   " << p_val << "   " << q[t].p_min
                  << std::endl;
        continue;
      }
      q[t].num += q[k].num;
      q[k].num = 0;
      q[t].wt_num += q[k].wt_num;
      q[k].wt_num = 0;
      q[t].ls_err += q[k].ls_err;
      q[k].ls_err = 0;
      for (int j = 0; j < kF; ++j) {
        q[t].val[j] = q[t].val[j] + q[k].val[j];
        q[k].val[j] = 0;
        q[t].err[j] = q[t].err[j] + q[k].err[j];
        q[k].err[j] = 0;
      }
    }
  }
}

// final_calc_auto (bins_calcs.cpp:54-91): the auto correlation's normalisations (SS pairs counted
// once, RR halved unless it came from the ordered RsRd-same pass), the analytic RR of a periodic box
// (folded like the data when re-binning), then the estimator with SR standing in for DR and the pi
// integration of ss, sr, rr and the result.
void final_calc_auto(Bins &ss, Bins &sr, Bins &rr, Bins &fin, const cpc_norm &n, const cpc_params &p, bool rsrd_same,
                     bool rebin, Norm &out) {
  const bool periodic = p.geometry == CPC_GEOM_BOX;
  const double SS_wt = n.S.val * n.S.val / 2.0;
  const double SS_err = SS_wt * std::sqrt(n.S.err / std::pow(n.S.val, 2) + n.S.err / std::pow(n.S.val, 2));
  const double SR_wt = n.S.val * n.Rs.val;
  const double SR_err = SR_wt * std::sqrt(n.S.err / std::pow(n.S.val, 2) + n.Rs.err / std::pow(n.Rs.val, 2));
  double RR_wt = n.Rs.val * n.Rs.val * n.include_prob_RR * n.include_prob_RR;
  double RR_err = RR_wt * std::sqrt(n.Rs.err / std::pow(n.Rs.val, 2) + n.Rs.err / std::pow(n.Rs.val, 2));
  if (!rsrd_same) {
    RR_wt /= 2.0;
    RR_err /= 2.0;
  }
  if (periodic) {
    RR_wt = 1;
    RR_err = 0;
    rr_periodic(rr, p);
    if (rebin) re_bin(rr, p);
  }
  out.SD = {SS_wt, SS_err};
  out.SR = {SR_wt, SR_err};
  out.RR = {RR_wt, RR_err};
  out.DR = {SR_wt, SR_err};
  final_calc_bins(ss, sr, sr, rr, fin, out, p, periodic);
  // int_calc(ss, rr, sr, final) (bins_calcs.cpp:373-383): each array integrated once
  int_calc(ss, p);
  int_calc(rr, p);
  int_calc(sr, p);
  int_calc(fin, p);
}

// final_calc_mean_projected (bins_calcs.cpp:258-329), estimator 4: per rp bin the pi-integrated
// mean of each field over the SD pairs minus the same over the SR pairs (the boost factor applies
// only to the lensing kinds, which_corr >= 7, which this engine does not count).  Only the rp rows
// of the result are set; its (rp, pi) cells stay zero.  The error column is val * sqrt(0), as there.
void final_calc_mean_projected(Bins &sd, Bins &sr, Bins &fin, const cpc_params &p) {
  int_calc(sd, p);
  int_calc(sr, p);
  const double final_var = 0;
  for (int i = 0; i < p.n_rp; ++i) {
    if (sd[i].num > 0 || sr[i].num > 0) {
      const double sd_num = (double)sd[i].num;
      const double sr_num = (double)sr[i].num;
      if (sd_num > 0) fin[i].sig_crit = sd[i].sig_crit / sd[i].wt_num;
      for (int j = 0; j < kF; ++j) {
        fin[i].val[j] = 0;
        if (sd_num > 0) fin[i].val[j] = sd[i].val[j] / sd[i].wt_num;
        fin[i].err[j] = fin[i].val[j] * std::sqrt(final_var);
        if (sr_num > 0) fin[i].val[j] -= sr[i].val[j] / sr[i].wt_num;
      }
    } else {
      for (int j = 0; j < kF; ++j) {
        fin[i].val[j] = 0;
        fin[i].err[j] = 0;
      }
    }
    fin[i].num = sd[i].num;
    fin[i].wt_num = sd[i].wt_num;
    if (fin[i].val[0] != fin[i].val[0]) std::cout << "NaN Problem in bin " << i << "  " << std::endl;
  }
}

// outp_bins (outp.cpp:45-74): one line per rp bin.
bool write_bins(const Bins &b, const std::string &name) {
  std::ofstream w(name.c_str());
  if (!w.is_open()) return false;
  w << "# rp\tnpairs\twgg\twgg_err\twgp\twgp_err\twgx wgx_err\twpp \twpp_err\t\twxx\twxx_err\t\twpx  wpx_err \ttheta   "
       "theta_err   LS_err     wt_npair    rp_low   rp_high sig_crit"
    << std::endl;
  for (const Bin &x : b) {
    w << (x.b_max + x.b_min) / 2 << "\t" << x.num;
    for (int j = 0; j < kF; ++j) w << "\t" << x.val[j] << "\t" << x.err[j];
    w << "\t" << x.ls_err_int << "   " << x.wt_num << "      " << x.b_min << "    " << x.b_max << "    " << x.sig_crit
      << std::endl;
  }
  return true;
}

// outp_2Dbins (outp.cpp:291-321): one line per (rp, pi) bin.
bool write_bins2d(const Bins &b, const std::string &name) {
  std::ofstream w(name.c_str());
  if (!w.is_open()) return false;
  w << "# rp     pi   npairs     wgg      wgg_err  wgp     wgp_err  wgx  wgx_err     wpp  wpp_err    wxx   wxx_err  "
       "wpx   wpx_err    theta    theta_err  wt_npairs "
    << std::endl;
  for (const Bin &x : b)
    for (const PBin &q : x.p) {
      w << (x.b_max + x.b_min) / 2 << "        " << (q.p_min + q.p_max) / 2.0 << "         " << q.num;
      for (int j = 0; j < kF; ++j) w << "       " << q.val[j] << "      " << q.err[j];
      w << "   " << q.wt_num << std::endl;
    }
  return true;
}

// Files of one do_bining call under the prefix `of`, counted; the first failure to open a file
// stops the writing (the reference prints "<name> not open" and goes on, outp.cpp:50-54).
struct Writer {
  std::string of;
  int files = 0;
  bool ok = true;
  void put(const Bins &b, const std::string &name, bool two_d) {
    if (!ok) return;
    ok = two_d ? write_bins2d(b, of + name) : write_bins(b, of + name);
    if (ok) ++files;
  }
};

// A pair set of do_bining: region r of the caller's histogram, or zero bins for a set that was not
// counted (bins_ini allocates every set, initialization.cpp:472-488).
Bins region_or_empty(const cpc_params &p, const cpc_hist *h, int r) {
  return h ? load_region(p, h, r) : empty_bins(p);
}

std::string jk_tag(int r) {
  std::ostringstream s;
  s << r;  // do_jk == 1 (do_corr.cpp:278-284)
  return s.str();
}

cpc_wsum to_wsum(const Wt &w) { return {w.val, w.err}; }

// The auto branch of do_bining (do_corr.cpp:220-324).
void auto_branch(Writer &W, const cpc_params &p, const cpc_norm &n, const cpc_bining &b, Norm &last) {
  const bool jk_on = p.do_jk == 1 && p.n_jk > 0;
  Bins SS = region_or_empty(p, b.SS, 0), SR = region_or_empty(p, b.SRs, 0), RR = region_or_empty(p, b.RsRs, 0),
       F = empty_bins(p);
  if (b.do_auto_rebin) {
    re_bin(SS, p);
    re_bin(SR, p);
    re_bin(RR, p);
    W.put(SR, "bins_auto_SR2.temp", true);
  }
  final_calc_auto(SS, SR, RR, F, n, p, b.rsrd_same != 0, b.do_auto_rebin != 0, last);
  W.put(F, "bins_auto_final.dat", false);
  W.put(SS, "bins_auto_SS.dat", false);
  W.put(SR, "bins_auto_SR.dat", false);
  W.put(RR, "bins_auto_RR.dat", false);
  W.put(F, "bins2D_auto_final.dat", true);
  W.put(SS, "bins2D_auto_SS.dat", true);
  W.put(SR, "bins2D_auto_SR.dat", true);
  W.put(RR, "bins2D_auto_RR.dat", true);
  if (!jk_on) return;
  std::vector<Bins> jk;
  jk.reserve(p.n_jk);
  for (int r = 0; r < p.n_jk && W.ok; ++r) {
    Bins jSS = region_or_empty(p, b.SS, r + 1), jSR = region_or_empty(p, b.SRs, r + 1),
         jRR = region_or_empty(p, b.RsRs, r + 1), jF = empty_bins(p);
    if (b.do_auto_rebin) {
      re_bin(jSS, p);
      re_bin(jRR, p);
      re_bin(jSR, p);
    }
    final_calc_auto(jSS, jSR, jRR, jF, n, p, b.rsrd_same != 0, b.do_auto_rebin != 0, last);
    const std::string t = jk_tag(r);
    W.put(jSS, "bins_auto_jk" + t + "_SS.dat", false);
    W.put(jRR, "bins_auto_jk" + t + "_RR.dat", false);
    W.put(jSR, "bins_auto_jk" + t + "_SR.dat", false);
    W.put(jF, "bins_auto_jk" + t + "_final.dat", false);
    W.put(jF, "bins2D_auto_jk" + t + "_final.dat", true);
    W.put(jSS, "bins2D_auto_jk" + t + "_SS.dat", true);
    W.put(jRR, "bins2D_auto_jk" + t + "_RR.dat", true);
    W.put(jSR, "bins2D_auto_jk" + t + "_SR.dat", true);
    jk.push_back(std::move(jF));
  }
  if (!W.ok) return;
  Bins J = empty_bins(p);
  jk_final(J, F, jk, p);
  W.put(J, "bins_auto_jk_final.dat", false);
  W.put(J, "bins2D_auto_jk_final.dat", true);
}

// One region of the cross estimator: final_calc_mean_projected for estimator 4, else final_calc.
void cross_calc(Bins &sd, Bins &sr, Bins &dr, Bins &rr, Bins &fin, const cpc_norm &n, const cpc_params &p,
                Norm &last) {
  if (p.estimator == 4)
    final_calc_mean_projected(sd, sr, fin, p);
  else
    final_calc(sd, sr, dr, rr, fin, n, p, last);
}

// The cross branch of do_bining (do_corr.cpp:326-431).
void cross_branch(Writer &W, const cpc_params &p, const cpc_norm &n, const cpc_bining &b, Norm &last) {
  const bool jk_on = p.do_jk == 1 && p.n_jk > 0;
  Bins SD = region_or_empty(p, b.SD, 0), SR = region_or_empty(p, b.SRd, 0), DR = region_or_empty(p, b.DRs, 0),
       RR = region_or_empty(p, b.RsRd, 0), F = empty_bins(p);
  cross_calc(SD, SR, DR, RR, F, n, p, last);
  W.put(F, "bins_cross_final.dat", false);
  W.put(F, "bins2D_cross_final.dat", true);
  W.put(SD, "bins_cross_SD.dat", false);
  W.put(SD, "bins2D_cross_SD.dat", true);
  W.put(SR, "bins_cross_SR.dat", false);
  W.put(SR, "bins2D_cross_SR.dat", true);
  W.put(DR, "bins_cross_DR.dat", false);
  W.put(DR, "bins2D_cross_DR.dat", true);
  W.put(RR, "bins_cross_RR.dat", false);
  W.put(RR, "bins2D_cross_RR.dat", true);
  if (!jk_on) return;
  std::vector<Bins> jk;
  jk.reserve(p.n_jk);
  for (int r = 0; r < p.n_jk && W.ok; ++r) {
    Bins jSD = region_or_empty(p, b.SD, r + 1), jSR = region_or_empty(p, b.SRd, r + 1),
         jDR = region_or_empty(p, b.DRs, r + 1), jRR = region_or_empty(p, b.RsRd, r + 1), jF = empty_bins(p);
    cross_calc(jSD, jSR, jDR, jRR, jF, n, p, last);
    const std::string t = jk_tag(r);
    W.put(jSD, "bins_cross_jk" + t + "_SD.dat", false);
    W.put(jRR, "bins_cross_jk" + t + "_RR.dat", false);
    W.put(jDR, "bins_cross_jk" + t + "_DR.dat", false);
    W.put(jSR, "bins_cross_jk" + t + "_SR.dat", false);
    W.put(jF, "bins_cross_jk" + t + "_final.dat", false);
    W.put(jF, "bins2D_cross_jk" + t + "_final.dat", true);
    W.put(jSD, "bins2D_cross_jk" + t + "_SD.dat", true);
    W.put(jRR, "bins2D_cross_jk" + t + "_RR.dat", true);
    W.put(jDR, "bins2D_cross_jk" + t + "_DR.dat", true);
    W.put(jSR, "bins2D_cross_jk" + t + "_SR.dat", true);
    jk.push_back(std::move(jF));
  }
  if (!W.ok) return;
  Bins J = empty_bins(p);
  jk_final(J, F, jk, p);
  W.put(J, "bins_cross_jk_final.dat", false);
  W.put(J, "bins2D_cross_jk_final.dat", true);
}

const char *check_output_params(const cpc_params &p) {
  if (p.estimator < 0 || p.estimator > 4) return "estimator must be 0-4";
  if (p.which_corr < 0 || p.which_corr > 5) return "which_corr must be 0-5";
  const bool unit_pi = p.geometry == CPC_GEOM_SKY && (p.coordinates == 3 || p.coordinates == 5);
  if (p.n_rp <= 0 || p.n_pi <= 0 || (!unit_pi && !(p.pi_max > p.pi_min))) return "bad binning";
  if (p.do_jk != 0 && p.do_jk != 1) return "do_jk must be 0 or 1";
  if (p.do_jk == 1 && p.n_jk > 0 && p.jk_mode != CPC_JK_LEAVE_ONE_OUT) return "histograms must be leave-one-out";
  return nullptr;
}

}  // namespace

extern "C" {

cpc_status cpc_weight_sum(const double *w, int64_t n, cpc_wsum *out) {
  if (!out || n < 0) return CPC_ERR_INVALID;
  double s = 0, s2 = 0;
  for (int64_t i = 0; i < n; ++i) {
    const double x = w ? w[i] : 1.0;
    s += x;
    s2 += x * x;
  }
  out->val = s;
  out->err = s2;
  return CPC_OK;
}

cpc_status cpc_do_bining(const cpc_params *pp, const cpc_norm *norm, const cpc_bining *sets, const char *out_file,
                         cpc_pair_norm *pair_norm, int *n_files) {
  if (n_files) *n_files = 0;
  if (!pp || !norm || !sets || !out_file) return CPC_ERR_INVALID;
  const cpc_params &p = *pp;
  if (p.estimator < 0 || p.estimator > 4 || p.which_corr < 0 || p.which_corr > 5) return CPC_ERR_UNSUPPORTED;
  if (check_output_params(p)) return CPC_ERR_INVALID;
  const cpc_bining &b = *sets;
  Writer W;
  W.of = out_file;
  Norm last;  // data_info.{SD,SRd,DRs,RsRd}_wt start at zero (initialization.cpp:322-338)
  W.put(region_or_empty(p, b.SRs, 0), "bins_auto_SR.temp", true);
  W.put(region_or_empty(p, b.SRd, 0), "bins_cross_SR.temp", true);
  if (b.do_auto_corr) auto_branch(W, p, *norm, b, last);
  if (b.do_cross_corr) cross_branch(W, p, *norm, b, last);
  if (pair_norm) *pair_norm = {to_wsum(last.SD), to_wsum(last.SR), to_wsum(last.DR), to_wsum(last.RR)};
  if (n_files) *n_files = W.files;
  return W.ok ? CPC_OK : CPC_ERR_INVALID;
}

cpc_status cpc_write_cross(const cpc_params *pp, const cpc_norm *norm, const cpc_hist *sd, const cpc_hist *sr,
                           const cpc_hist *dr, const cpc_hist *rr, const char *out_file, int *n_files) {
  if (n_files) *n_files = 0;
  if (!pp || !norm || !sd || !sr || !dr || !rr || !out_file) return CPC_ERR_INVALID;
  const cpc_params &p = *pp;
  if (p.estimator != 0 && p.estimator != 4) return CPC_ERR_UNSUPPORTED;
  if (p.which_corr < 0 || p.which_corr > 5) return CPC_ERR_UNSUPPORTED;
  if (check_output_params(p)) return CPC_ERR_INVALID;
  cpc_bining b{};
  b.do_cross_corr = 1;
  b.SD = sd;
  b.SRd = sr;
  b.DRs = dr;
  b.RsRd = rr;
  Writer W;
  W.of = out_file;
  Norm last;
  cross_branch(W, p, *norm, b, last);
  if (n_files) *n_files = W.files;
  return W.ok ? CPC_OK : CPC_ERR_INVALID;
}

}  // extern "C"
