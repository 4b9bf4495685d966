// cpc_main.cpp -- the corr_pc program on the MI355X engine.
//
// `corr_pc <file.inp> [--precision split|fp64] [--device N] [--plan]` takes the reference's parameter
// file unchanged, reads the catalogs it names, runs the reference's pair passes through the engine's
// C ABI (include/cpc.h) and writes the reference's output files: inp.info, every bins*_*.dat /
// bins2D_*.dat / *.temp file of do_bining, and inp2.info.  Host side only; the pair loops run on the
// GPU inside cpc_count_batch.
//
// Restates, for which_corr 0-5 on a periodic box (periodic_box 1, coordinates 6 or 7) or a survey
// (periodic_box 0, coordinates 0), data_sorted 1 (or 0 without auto pairs), do_jk 0 or 1:
//   main                                   main.cpp:27-82
//   data_ini: the positional parser and    initialization.cpp:62-350
//             the derived flags
//   da_Z: the distance table               initialization.cpp:23-60
//   jk_initilize, seed                     jk.cpp:22-43
//   read_data: catalogs, in read order     read_dat.cpp:18-347
//             with the C library's rand()
//   elip_angle, arctan                     calcs.cpp:35-46, 73-83
//   do_corr: pass list, operand order,     do_corr.cpp:24-158
//            the RsRd_same merge
//   corels2 / corels1 roles and swap,      corr.cpp:226-296, 111; initialization.cpp:490-543, 597-599
//     choose_corel_func, include_prob_max
//   outp_data_info                         outp.cpp:77-162
// and hands the histograms to cpc_do_bining (do_bining, do_corr.cpp:203-433).
//
// `--plan` parses, reads and prints what would run (catalog sums, coins, passes) without touching
// the GPU.

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/cpc.h"

namespace {

struct VarSt {
  double val = 0, err = 0;
};

// data_info (data_def.h:46-96): the fields data_ini fills and the program reads.
struct DataInfo {
  bool do_auto_corr = false, do_cross_corr = false, do_auto_rebin = false, use_comoving = false, RsRd_same = false,
       periodic_box = false, rand_subsample = false;
  int do_jk = 0, data_sorted = 0, which_corr = 0, coordinates = 0, sig_crit = 0, estimator = 0;
  std::string shape_pos_file, shape_z_file, shape_e_file, shape_wt_file, shape_jk_file, shape_patch_file;
  std::string density_pos_file, density_z_file, density_wt_file, density_jk_file, density_e_file;
  std::string shape_rand_pos_file, shape_rand_z_file, shape_rand_wt_file, shape_rand_jk_file, shape_rand_patch_file;
  std::string density_rand_pos_file, density_rand_z_file, density_rand_wt_file, density_rand_jk_file;
  std::string distance_file, patches_file, out_file;
  int n_bins = 0, n_threads = 0, n_jk = 0, n_jk_regions = 0, n_bin_var = 7, n_patch = 0;
  double bin_r_min = 0, bin_r_max = 0, r_size = 0;
  bool lin_bin = false;
  double dz = 0, z_min = 0, z_max = 0, z_sep_max = 0, z_sep_min = 0;
  std::vector<double> dA_z, H_z, dC_z;
  int64_t table_rows = 0;  // rows of the distance table actually read
  int n_p_bin = 0;
  double p_max = 0, p_min = 0, p_size = 0, periodic_box_size = 0;
  int n_shape = 0, n_density = 0, n_density_rand = 0, n_shape_rand = 0;
  VarSt DD_wt, RsRs_wt, DRs_wt, SS_wt, SD_wt, SRs_wt, SRd_wt, RsRd_wt, D_wt, S_wt, Rs_wt, Rd_wt;
  double include_prob_RR = 1;
};

constexpr uint32_t kSeed = 124693517;  // seed(), jk.cpp:40-43

bool fail(const std::string &msg) {
  std::cout << msg << std::endl;
  return false;
}

// da_Z (initialization.cpp:23-60): rows z, H, then dA and dC (use_comoving 0: z H dA dC x;
// use_comoving 1: z H x dC dA) at z = k*dz, at most ceil((z_max+1)/dz)+1 rows.
bool da_Z(DataInfo &d) {
  std::ifstream in(d.distance_file.c_str());
  if (!in.is_open()) return fail("da_Z:: distances file (dA vs Z) not opened: " + d.distance_file);
  const double limit = std::ceil((d.z_max + 1) / d.dz) + 1;
  int i = 0;
  while (!in.eof() && i < limit) {
    double z = 0, H = 0, a = 0, c = 0, b = 0;
    in >> z >> H >> a >> c >> b;
    if (in) {
      d.H_z.push_back(H);
      d.dC_z.push_back(c);
      d.dA_z.push_back(d.use_comoving ? b : a);
    }
    ++i;  // the reference counts the read that met the end of the file too
  }
  if (i < std::ceil(d.z_max / d.dz) + 1) return fail("da_Z:: distances file (dA vs Z) ended unexpectedly");
  d.table_rows = (int64_t)d.H_z.size();
  return true;
}

template <class T>
void field(std::istream &in, T &v) {
  std::string junk;
  in >> junk >> v;
}

// data_ini (initialization.cpp:62-350).
bool data_ini(DataInfo &d, const std::string &name) {
  std::ifstream in(name.c_str());
  if (!in.is_open()) return fail("data_ini:: initialisation File not open" + name);
  field(in, d.which_corr);
  field(in, d.coordinates);
  field(in, d.estimator);
  field(in, d.data_sorted);
  field(in, d.use_comoving);
  field(in, d.do_jk);
  field(in, d.sig_crit);
  field(in, d.shape_pos_file);
  field(in, d.shape_z_file);
  field(in, d.shape_e_file);
  field(in, d.shape_wt_file);
  field(in, d.shape_jk_file);
  field(in, d.shape_patch_file);
  field(in, d.density_pos_file);
  field(in, d.density_z_file);
  field(in, d.density_wt_file);
  field(in, d.density_jk_file);
  field(in, d.density_e_file);
  field(in, d.shape_rand_pos_file);
  field(in, d.shape_rand_z_file);
  field(in, d.shape_rand_wt_file);
  field(in, d.shape_rand_jk_file);
  field(in, d.shape_rand_patch_file);
  field(in, d.density_rand_pos_file);
  field(in, d.density_rand_z_file);
  field(in, d.density_rand_wt_file);
  field(in, d.density_rand_jk_file);
  field(in, d.distance_file);
  field(in, d.patches_file);
  field(in, d.out_file);
  field(in, d.n_threads);
  field(in, d.n_shape);
  field(in, d.n_density);
  field(in, d.n_shape_rand);
  field(in, d.n_density_rand);
  field(in, d.rand_subsample);
  field(in, d.n_jk_regions);
  field(in, d.n_patch);
  field(in, d.bin_r_min);
  field(in, d.bin_r_max);
  field(in, d.n_bins);
  field(in, d.lin_bin);
  field(in, d.n_p_bin);
  field(in, d.p_min);
  field(in, d.p_max);
  field(in, d.z_min);
  field(in, d.z_max);
  field(in, d.dz);
  field(in, d.z_sep_min);
  field(in, d.z_sep_max);
  field(in, d.periodic_box);
  field(in, d.periodic_box_size);
  // The reference reads on past a malformed or short file; here that is an error.
  if (in.fail()) return fail("data_ini:: could not read every parameter of " + name);

  // modes this engine runs
  if (d.which_corr < 0 || d.which_corr > 5)
    return fail("which_corr " + std::to_string(d.which_corr) +
                " is not supported (0-2: density and IA pairs, 3-4: convergence pairs, 5: ED)");
  if (d.periodic_box && d.coordinates != 6 && d.coordinates != 7) {
    std::cout << " wrong coordinates choice for periodic box:" << d.periodic_box_size << " " << d.coordinates
              << std::endl;
    return false;
  }
  if (!d.periodic_box && (d.coordinates < 0 || d.coordinates > 5))
    return fail("coordinates " + std::to_string(d.coordinates) +
                " is not supported without a periodic box (supported: 0-5 on the sky, 6 and 7 in a box)");
  if (d.which_corr == 5 && !d.periodic_box) return fail("which_corr 5 (ED) is defined for the periodic box only");
  if (d.periodic_box && d.data_sorted == 2) return fail("data_sorted==2 doesnot work with periodic box.. use 1");
  if (d.data_sorted < 0 || d.data_sorted > 2) return fail("data_sorted must be 0, 1 or 2");
  if (d.do_jk != 0 && d.do_jk != 1) return fail("do_jk must be 0 or 1");
  if (d.estimator < 0 || d.estimator > 4) return fail("estimator must be 0-4");
  // sig_crit > 0 weights every pair by Sigma_crit^-2 and scales its shape sums by Sigma_crit
  // (corels.cpp:52-61, calcs.cpp:142-176), the weak-lensing weighting; counting without it would
  // silently give other numbers
  if (d.sig_crit != 0)
    return fail("sig_crit " + std::to_string(d.sig_crit) + " is not supported (the weak-lensing Sigma_crit weights)");

  d.n_bin_var = 7;
  if (d.periodic_box) {
    d.dz = 0;
    std::cout << "doing xyz box: size=" << d.periodic_box_size << "   periodic=" << d.periodic_box
              << " njk: " << d.n_jk_regions << std::endl;
    if (d.bin_r_max > d.periodic_box_size || d.p_max > d.periodic_box_size)
      std::cout << "bin limits > the periodic box size" << std::endl;
  }
  switch (d.estimator) {  // initialization.cpp:162-185
    case 0:
      d.do_cross_corr = true;
      d.do_auto_corr = false;
      break;
    case 2:
      d.do_auto_rebin = true;
      d.do_auto_corr = true;
      d.do_cross_corr = false;
      break;
    case 1:
      d.do_auto_corr = true;
      d.do_cross_corr = false;
      break;
    case 3:
      d.do_auto_corr = true;
      d.do_cross_corr = true;
      d.do_auto_rebin = true;
      break;
    default:
      d.do_cross_corr = true;
      d.do_auto_corr = false;
      d.do_auto_rebin = false;
      break;
  }
  // data_sorted 0 reaches the engine as cpc_params.unsorted = 1: do_pair_unsorted (corr.cpp:36-52)
  // ends no scan, and the auto pass starts every image's scan at j = i+1 (corr.cpp:167-182).
  // data_sorted 2 (the sky only, Dec order) reaches it as unsorted = 2: the z test ends no scan
  // (ret_z = 1) and the Dec test ends a scan at its first galaxy above the window (ret_dec,
  // initialization.cpp:607-612), so a scan from the window's first galaxy finds every pair of the
  // window, which the engine counts; but corels2 starts scans at a forward-only pointer
  // (corr.cpp:356-362), which skips galaxies when Dec windows are not ordered like the galaxies (their
  // half-width changes with redshift) or runs off an empty window.  The engine reports the outer
  // galaxies where one thread's pointer would do so (n_sparse_windows), and the program warns.
  d.RsRd_same = false;
  if (d.n_density_rand == 0 && d.which_corr <= 2) {
    d.RsRd_same = true;
    d.n_density_rand = d.n_shape_rand;
  }
  d.include_prob_RR = 1;
  if (d.rand_subsample)
    d.include_prob_RR = std::sqrt(std::sqrt(double(d.n_density) / d.n_density_rand * double(d.n_shape) /
                                            d.n_shape_rand)) * 1.0;
  std::cout << "RsRd same  " << d.RsRd_same << "  RR-prob:" << d.include_prob_RR << std::endl;
  d.n_jk = d.n_jk_regions;  // jk_initilize (jk.cpp:22-38)

  if (d.z_sep_max == 0) d.z_sep_max = 9;
  if (d.dz == 0 && !d.periodic_box) return fail("dz=0");
  if (!d.periodic_box && !da_Z(d)) return false;
  if (d.coordinates == 0 || d.coordinates == 1) {
    const int z_idx = (int)std::ceil(d.z_max / d.dz);
    if (z_idx >= d.table_rows) return fail("da_Z:: distances file (dA vs Z) ended unexpectedly");
    if (d.coordinates == 0) {
      d.z_sep_max = d.p_max * d.H_z[z_idx] / 300000.0 + 0.001;
      d.z_sep_min = d.p_min * d.H_z[z_idx] / 300000.0 - 0.001;
    } else {
      d.z_sep_max = d.bin_r_max * d.H_z[z_idx] / 300000.0 + 0.001;
      d.z_sep_min = -1 * d.z_sep_max;
    }
  } else if (d.coordinates == 6) {
    d.z_sep_max = d.p_max + 1;
    d.z_sep_min = d.p_min - 1;
  } else if (d.coordinates == 7) {
    d.z_sep_max = d.bin_r_max + 0.1;
    d.z_sep_min = d.z_sep_max * -1;
  }  // coordinates 2-5 keep the file's z_sep_min and z_sep_max
  if (d.bin_r_max == 0 || d.bin_r_min > d.bin_r_max) return fail("Bad Bin limits");
  d.p_size = (d.p_max - d.p_min) / (d.n_p_bin);
  if (d.coordinates == 3 || d.coordinates == 5) d.p_size = 1;  // initialization.cpp:307-308
  if (d.lin_bin)
    d.r_size = (d.bin_r_max - d.bin_r_min) / d.n_bins;
  else
    d.r_size = (std::log10(d.bin_r_max) - std::log10(d.bin_r_min)) / d.n_bins;
  return true;
}

// outp_data_info (outp.cpp:77-162).
bool outp_data_info(const DataInfo &d, const std::string &name) {
  std::ofstream o(name.c_str());
  if (!o.is_open()) return fail(name + " not open");
  o << "do_auto_corr         " << d.do_auto_corr << std::endl;
  o << "do_cross_corr        " << d.do_cross_corr << std::endl;
  o << "data_sorted          " << d.data_sorted << std::endl;
  o << "use_comoving         " << d.use_comoving << std::endl;
  o << "do_auto_rebin        " << d.do_auto_rebin << std::endl;
  o << "do_jk                " << d.do_jk << std::endl;
  o << "shape_pos            " << d.shape_pos_file << std::endl;
  o << "shape_z              " << d.shape_z_file << std::endl;
  o << "shape_e              " << d.shape_e_file << std::endl;
  o << "shape_wt             " << d.shape_wt_file << std::endl;
  o << "shape_jk             " << d.shape_jk_file << std::endl;
  o << "density_pos          " << d.density_pos_file << std::endl;
  o << "density_z            " << d.density_z_file << std::endl;
  o << "density_wt           " << d.density_wt_file << std::endl;
  o << "density_jk           " << d.density_jk_file << std::endl;
  o << "density_e            " << d.density_e_file << std::endl;
  o << "shape_rand_pos       " << d.shape_rand_pos_file << std::endl;
  o << "shape_rand_z         " << d.shape_rand_z_file << std::endl;
  o << "shape_rand_wt        " << d.shape_rand_wt_file << std::endl;
  o << "shape_rand_jk        " << d.shape_rand_jk_file << std::endl;
  o << "density_rand_pos     " << d.density_rand_pos_file << std::endl;
  o << "density_rand_z       " << d.density_rand_z_file << std::endl;
  o << "density_rand_wt      " << d.density_rand_wt_file << std::endl;
  o << "density_rand_jk      " << d.density_rand_jk_file << std::endl;
  o << "distances            " << d.distance_file << std::endl;
  o << "output               " << d.out_file << std::endl;
  o << "n_threads            " << d.n_threads << std::endl;
  o << "n_shape              " << d.n_shape << std::endl;
  o << "n_density            " << d.n_density << std::endl;
  o << "n_shape_rand         " << d.n_shape_rand << std::endl;
  o << "n_density_rand       " << d.n_density_rand << std::endl;
  o << "n_jk                 " << d.n_jk << std::endl;
  o << "bin_r_min            " << d.bin_r_min << std::endl;
  o << "bin_r_max            " << d.bin_r_max << std::endl;
  o << "n_bins               " << d.n_bins << std::endl;
  o << "lin_bin              " << d.lin_bin << std::endl;
  o << "n_p_bin              " << d.n_p_bin << std::endl;
  o << "p_min                " << d.p_min << std::endl;
  o << "p_max                " << d.p_max << std::endl;
  o << "z_min                " << d.z_min << std::endl;
  o << "z_max                " << d.z_max << std::endl;
  o << "dz                   " << d.dz << std::endl;
  o << "p_bin_siz            " << d.p_size << std::endl;
  o << "z_sep_max            " << d.z_sep_max << std::endl;
  o << "z_sep_min            " << d.z_sep_min << std::endl;
  o << "DD_Wt                " << d.DD_wt.val << std::endl;
  o << "DRs_wt               " << d.DRs_wt.val << std::endl;
  o << "SS_wt                " << d.SS_wt.val << std::endl;
  o << "SRs_wt               " << d.SRs_wt.val << std::endl;
  o << "SD_wt                " << d.SD_wt.val << std::endl;
  o << "RsRs_Wt              " << d.RsRs_wt.val << std::endl;
  o << "RsRd_wt              " << d.RsRd_wt.val << std::endl;
  o << "S_wt                 " << d.S_wt.val << std::endl;
  o << "D_wt                 " << d.D_wt.val << std::endl;
  o << "Rs_wt                " << d.Rs_wt.val << std::endl;
  o << "Rd_wt                " << d.Rd_wt.val << std::endl;
  o << "RsRd_same            " << d.RsRd_same << std::endl;
  return true;
}

// One catalog as read_data leaves it, in SoA form.
struct Cat {
  std::string name;
  bool read = false;  // false: never read (an empty catalog for every pass)
  std::vector<double> x, y, z, w, e1, e2, e3, phi, jkp, inc, dc, da;
  std::vector<int32_t> jk;
  bool has_e = false;  // e1 (and e2) carry values: shapes, or a convergence in e1
  bool kappa = false;  // e1 is a convergence kappa read from a one-column file, e2 is zero
  bool has_e3 = false; // (e1, e2, e3) is a unit 3-D axis (which_corr 5)
  int64_t n() const { return (int64_t)x.size(); }
};

enum Role { kShape, kDensity, kShapeRand, kDensityRand };

// arctan and elip_angle (calcs.cpp:35-46, 73-83): phi = atan2-like angle of (e1, -e2), halved; NaN -> 0.
double elip_angle(double e0, double e1_in) {
  double x = e0, y = -1 * e1_in;
  double t = std::atan(y / x);
  if (x < 0) {
    if (y < 0)
      t -= M_PI;
    else
      t += M_PI;
  }
  double phi = 0.5 * t;
  if (phi != phi) phi = 0;
  return phi;
}

// Sky coordinates whose pi is the ephi angle of the g_with_shape galaxy (pi_calc_phi).
bool phi_bins(const DataInfo &d) { return !d.periodic_box && (d.coordinates == 2 || d.coordinates == 4); }

// read_data (read_dat.cpp:18-347) for coordinates 0-7, which_corr 0-5.
bool read_data(DataInfo &d, Role role, Cat &c) {
  std::string p_f, z_f, e_f, w_f, jk_f;
  int n_gal = 0;
  VarSt *wt = nullptr;
  // what the e file holds, per which_corr (read_dat.cpp:28-68): shapes (e1 e2), or a convergence
  bool read_e = true, read_kappa = false, read_phi = false, read_e3 = false, cross_shape_shape = false,
       read_jk = true;
  switch (d.which_corr) {
    case 0:
      read_e = false;
      break;
    case 2:
      cross_shape_shape = true;
      break;
    case 3:
      read_e = false;
      read_kappa = true;
      break;
    case 4:
      read_e = role == kShape;  // the shape sample has e, the density sample a convergence
      read_kappa = role != kShape;
      cross_shape_shape = true;
      break;
    case 5:  // the shape sample reads 3-D axes; the case falls into case 8 (read_dat.cpp:47-58)
      read_e = false;
      read_e3 = role == kShape;
      cross_shape_shape = true;
      break;
    default:
      break;
  }
  const bool is_rand = role == kShapeRand || role == kDensityRand;
  // coordinates 4 and 5 read no redshift (read_dat.cpp:27-28); with 2 and 4 the shape sample reads
  // its e file for the angle phi, and random catalogs draw phi (read_dat.cpp:91-92, 127-131, 146-149)
  const bool read_z = !(d.coordinates == 4 || d.coordinates == 5);
  const bool rand_phi = phi_bins(d) && (role == kShapeRand || (role == kDensityRand && d.which_corr == 2));
  if (role == kShape && phi_bins(d)) read_e = true;
  switch (role) {
    case kShape:
      p_f = d.shape_pos_file, z_f = d.shape_z_file, e_f = d.shape_e_file, w_f = d.shape_wt_file,
      jk_f = d.shape_jk_file, n_gal = d.n_shape, wt = &d.S_wt;
      break;
    case kDensity:
      p_f = d.density_pos_file, z_f = d.density_z_file, w_f = d.density_wt_file, jk_f = d.density_jk_file,
      n_gal = d.n_density, wt = &d.D_wt;
      if (cross_shape_shape)
        e_f = d.density_e_file;
      else
        read_e = read_kappa = false;
      read_e3 = false;
      break;
    case kShapeRand:
      p_f = d.shape_rand_pos_file, z_f = d.shape_rand_z_file, w_f = d.shape_rand_wt_file,
      jk_f = d.shape_rand_jk_file, n_gal = d.n_shape_rand, wt = &d.Rs_wt, read_e = read_kappa = false;
      break;
    case kDensityRand:
      p_f = d.density_rand_pos_file, z_f = d.density_rand_z_file, w_f = d.density_rand_wt_file,
      jk_f = d.density_rand_jk_file, n_gal = d.n_density_rand, wt = &d.Rd_wt, read_e = read_kappa = false;
      break;
  }
  c.read = true;
  if (n_gal == 0) {
    std::cout << "n gal zero. Nothing to read: " << p_f << std::endl;
    return true;
  }
  std::ifstream data_p(p_f.c_str()), data_z, data_e, data_wt, data_jk;
  if (read_z) data_z.open(z_f.c_str());
  if (jk_f != "0") {
    data_jk.open(jk_f.c_str());
  } else {
    read_jk = false;
    std::cout << "no jk file given" << std::endl;
  }
  if (!d.do_jk) read_jk = false;
  if (w_f != "0") {
    std::cout << "doing weighted calcs" << std::endl;
    data_wt.open(w_f.c_str());
    if (!data_wt.is_open()) return fail("wt file not open: " + w_f);
  }
  if (!data_p.is_open()) return fail("initialisation::pos File not open: " + p_f);
  if (read_z && !data_z.is_open()) return fail("initialisation::z File not open:" + z_f);
  if (d.do_jk && !data_jk.is_open() && read_jk) return fail("initialisation::jk File not open: " + jk_f);
  if (is_rand) read_e3 = false;
  if (read_e || read_kappa || read_e3) {
    data_e.open(e_f.c_str());
    if (!data_e.is_open()) return fail("initialisation:: e File not open:" + e_f);
    if (read_e3 && (e_f.find("phi") != std::string::npos || e_f.find("kappa") != std::string::npos))
      return fail("e file " + e_f + ": a one-column phi / kappa name with the 3-D axes of which_corr 5");
    // the file's name can switch it to one column of phi or of kappa (read_dat.cpp:189-204)
    if (e_f.find("phi") != std::string::npos) read_phi = true, read_e = read_kappa = false;
    if (e_f.find("kappa") != std::string::npos || read_kappa) read_kappa = true, read_e = read_phi = false;
    // One column of phi leaves the shapes unset, and so does a convergence where the pair kind
    // rotates shapes (which_corr 1-2); the reference's pair functions would read them anyway.
    if (read_phi) return fail("e file " + e_f + ": one-column phi inputs are not supported");
    if (read_kappa && d.which_corr < 3)
      return fail("e file " + e_f + ": read as one column of kappa, which leaves the shapes of which_corr " +
                  std::to_string(d.which_corr) + " unset");
  }
  c.has_e = read_e || read_kappa || read_e3;
  c.kappa = read_kappa;
  c.has_e3 = read_e3;
  c.x.reserve(n_gal), c.y.reserve(n_gal), c.z.reserve(n_gal), c.w.reserve(n_gal), c.jkp.reserve(n_gal);
  int i = 0;
  while (data_p.good() && (data_z.good() || !read_z) && (data_jk.good() || !read_jk) &&
         (data_wt.good() || w_f == "0") && i < n_gal) {
    double ra = 0, dec = 0, redshift = 0;
    data_p >> ra >> dec;
    if (read_z) {
      data_z >> redshift;
      if (redshift < 0) {
        redshift = 0;
        std::cout << "gal redshift <0. changed to z=0" << std::endl;
      }
      if (redshift > d.z_max) {
        std::cout << "initialisation:: wrong z at data " << i << "     z==" << redshift << "  vs  " << d.z_max
                  << std::endl;
        return false;
      }
    }
    double w = 1, jk_prob = 1, DC = 0, DA = 0;
    if (d.periodic_box) {
      jk_prob = rand() / (RAND_MAX + 1.);
      DC = redshift;
      DA = 0;
    } else {
      const int it = (int)std::floor(redshift / d.dz);
      if (it >= d.table_rows) return fail("read_data:: redshift beyond the distance table");
      DC = d.dC_z[it];
      DA = d.dA_z[it];
    }
    if (d.do_jk == 0 || d.n_jk == 0) jk_prob = 0;
    if (w_f != "0") data_wt >> w;
    int jk = -1;
    if (read_jk) data_jk >> jk;
    double phi = -10;
    if (read_e) {
      double a = 0, b = 0;
      data_e >> a >> b;
      c.e1.push_back(a);
      c.e2.push_back(b);
      phi = elip_angle(a, b);
    } else if (read_kappa) {
      double k = 0;
      data_e >> k;
      c.e1.push_back(k);
      c.e2.push_back(0.0);
    } else if (read_e3) {
      // three components, scaled to unit length when not zero (read_dat.cpp:283-293)
      double a = 0, b = 0, e = 0;
      data_e >> a >> b >> e;
      const double temp_en = std::sqrt(a * a + b * b + e * e);
      if (temp_en > 0) {
        a /= temp_en;
        b /= temp_en;
        e /= temp_en;
      }
      c.e1.push_back(a);
      c.e2.push_back(b);
      c.e3.push_back(e);
    }
    if (rand_phi) phi = rand() / (RAND_MAX + 1.) * M_PI / 2 - M_PI / 2;
    double include_prob = 0;
    if (is_rand) include_prob = rand() / (RAND_MAX + 1.);
    c.x.push_back(ra);
    c.y.push_back(dec);
    c.z.push_back(redshift);
    c.w.push_back(w);
    c.jk.push_back(jk);
    c.jkp.push_back(jk_prob);
    c.phi.push_back(phi);
    c.inc.push_back(include_prob);
    c.dc.push_back(DC);
    c.da.push_back(DA);
    wt->val += w;
    wt->err += w * w;
    ++i;
  }
  if (i < n_gal) {
    std::cout << "gal data reading ended unexpectedly: " << i << "   " << n_gal << std::endl;
    return false;
  }
  for (size_t k = 1; d.data_sorted == 1 && k < c.z.size(); ++k)
    if (c.z[k] < c.z[k - 1])
      return fail("data_sorted=1 but " + z_f + " is not sorted by redshift (row " + std::to_string(k) + ")");
  for (size_t k = 1; d.data_sorted == 2 && k < c.y.size(); ++k)
    if (c.y[k] < c.y[k - 1])
      return fail("data_sorted=2 but " + p_f + " is not sorted by Dec (row " + std::to_string(k) + ")");
  return true;
}

// A pass of do_corr (do_corr.cpp:59-131): which sets of corels1 / corels2 run, with which catalogs.
struct Pass {
  std::string name;   // histogram it fills (SS, SRs, RsRs, SRd, RsRd, SD, DRs)
  std::string flag;   // the data_info do_* flag set around the call
  const Cat *g1, *g2; // arguments in the reference's order (g2 unused for corels1)
  bool auto_pass;
};

// Roles and kind of one corels call (corr.cpp:226-296 and 111; choose_corel_func,
// initialization.cpp:490-543; include_prob_max, :597-599).
struct Setup {
  const Cat *outer = nullptr, *inner = nullptr;
  int kind = CPC_KIND_DD, gws = 0;
  bool swapped = false, empty = false, shapes_undefined = false;
  double include_prob_max = 1;
};

Setup setup_pass(const DataInfo &d, const Pass &ps) {
  Setup s;
  const std::string &f = ps.flag;
  const bool rr = f == "RsRs" || f == "RsRd";
  if (rr || d.which_corr == 0) {
    s.kind = CPC_KIND_DD;
  } else if (d.which_corr == 1) {
    s.kind = f == "SS" ? CPC_KIND_SHAPE_SHAPE : CPC_KIND_SHAPE_DENSITY;
  } else if (d.which_corr == 2) {
    s.kind = (f == "SD" || f == "SS") ? CPC_KIND_SHAPE_SHAPE : CPC_KIND_SHAPE_DENSITY;
    if (phi_bins(d)) s.kind = (f == "DRs" || f == "SD") ? CPC_KIND_SHAPE_SHAPE : CPC_KIND_DD;
  } else if (d.which_corr == 5) {
    s.kind = CPC_KIND_ED;
  } else {
    s.kind = d.which_corr == 3 ? CPC_KIND_DENSITY_CONVERGENCE : CPC_KIND_SHAPE_CONVERGENCE;
  }
  s.include_prob_max = rr ? d.include_prob_RR : 1.0;
  if (ps.auto_pass) {
    s.outer = s.inner = ps.g1;
    s.gws = 1;  // corels1, corr.cpp:111
    s.empty = ps.g1->n() == 0;
    return s;
  }
  int n1 = 0, n2 = 0;  // sizes from the role flags, as corels2 takes them
  // coordinates 2 and 4 give the random passes a g_with_shape for pi_calc_phi (corr.cpp:226-249)
  if (f == "DRs") n2 = d.n_density, n1 = d.n_shape_rand, s.gws = phi_bins(d) ? 1 : 0;
  else if (f == "RsRd") n2 = d.n_shape_rand, n1 = d.n_density_rand, s.gws = phi_bins(d) ? 2 : 0;
  else if (f == "RsRs") n1 = n2 = d.n_shape_rand, s.gws = phi_bins(d) ? 1 : 0;
  else if (f == "SRd") n2 = d.n_shape, n1 = d.n_density_rand, s.gws = 2;
  else if (f == "SD") n2 = d.n_shape, n1 = d.n_density, s.gws = 2;
  else if (f == "SRs") n2 = d.n_shape, n1 = d.n_shape_rand, s.gws = 2;
  s.outer = ps.g1;
  s.inner = ps.g2;
  s.empty = n1 == 0 || n2 == 0 || ps.g1->n() == 0 || ps.g2->n() == 0;
  if (n2 > n1) {
    std::swap(s.outer, s.inner);
    s.swapped = true;
    s.gws = s.gws == 2 ? 1 : (s.gws == 1 ? 2 : s.gws);
  }
  if (s.kind == CPC_KIND_SHAPE_DENSITY && s.gws == 0) {
    // shape_density_corel rotates g1's shape when g_with_shape is neither 2 nor 1 (corels.cpp:216-219)
    s.gws = 1;
    if (!s.outer->has_e) {
      // g1 carries no shape: the reference reads an unset pointer there; these shape columns feed
      // no estimator, so the pass runs as density pairs
      s.kind = CPC_KIND_DD;
      s.shapes_undefined = true;
    }
  }
  if ((s.kind == CPC_KIND_SHAPE_DENSITY && !(s.gws == 2 ? s.inner : s.outer)->has_e) ||
      (s.kind == CPC_KIND_SHAPE_SHAPE && (!s.outer->has_e || !s.inner->has_e))) {
    // a shape kind on a catalog without shapes (coordinates 2 and 4 give such roles to random
    // catalogs): the reference reads an unset pointer; density pairs instead
    s.kind = CPC_KIND_DD;
    s.shapes_undefined = true;
  }
  if (s.kind == CPC_KIND_ED && !(s.gws == 2 ? s.inner : s.outer)->has_e3) {
    // ED_calc_PB reads the axis of g2 (g_with_shape 2) or else of g1; a catalog without axes there
    // is a reference read of an unset pointer: density pairs instead
    s.kind = CPC_KIND_DD;
    s.shapes_undefined = true;
  }
  if ((s.kind == CPC_KIND_DENSITY_CONVERGENCE || s.kind == CPC_KIND_SHAPE_CONVERGENCE) && s.gws != 0) {
    // kappa from the g_with_shape galaxy (density-convergence) or from the other one (shape-
    // convergence, which rotates the g_with_shape galaxy's shape) (corels.cpp:79-120); a random
    // catalog there carries none and the reference reads an unset pointer: density pairs instead
    const Cat *ws = s.gws == 2 ? s.inner : s.outer, *ot = s.gws == 2 ? s.outer : s.inner;
    const Cat *kap = s.kind == CPC_KIND_DENSITY_CONVERGENCE ? ws : ot;
    if (!kap->has_e || (s.kind == CPC_KIND_SHAPE_CONVERGENCE && (!ws->has_e || ws->kappa))) {
      s.kind = CPC_KIND_DD;
      s.shapes_undefined = true;
    }
  }
  if (s.kind == CPC_KIND_DD && s.gws == 0) s.gws = 2;
  return s;
}

// Histogram storage behind a cpc_hist.
struct Hist {
  std::vector<int64_t> num, num_rp;
  std::vector<double> wt_num, val, err;
  cpc_hist h{};
  void alloc(int R, int n_rp, int n_pi) {
    const size_t cells = (size_t)R * n_rp * n_pi;
    num.assign(cells, 0);
    wt_num.assign(cells, 0);
    val.assign(cells * CPC_NFIELD, 0);
    err.assign(cells * CPC_NFIELD, 0);
    num_rp.assign((size_t)R * n_rp, 0);
    h = {num.data(), wt_num.data(), val.data(), err.data(), num_rp.data()};
  }
  void add(const Hist &o) {  // add_bins (outp.cpp:242-288) over every array
    for (size_t i = 0; i < num.size(); ++i) num[i] += o.num[i];
    for (size_t i = 0; i < wt_num.size(); ++i) wt_num[i] += o.wt_num[i];
    for (size_t i = 0; i < val.size(); ++i) val[i] += o.val[i];
    for (size_t i = 0; i < err.size(); ++i) err[i] += o.err[i];
    for (size_t i = 0; i < num_rp.size(); ++i) num_rp[i] += o.num_rp[i];
  }
};

cpc_params base_params(const DataInfo &d, int precision) {
  cpc_params p{};
  p.geometry = d.periodic_box ? CPC_GEOM_BOX : CPC_GEOM_SKY;
  p.box_size = d.periodic_box_size;
  p.estimator = d.estimator;
  p.which_corr = d.which_corr;
  p.n_rp = d.n_bins;
  p.rp_min = d.bin_r_min;
  p.rp_max = d.bin_r_max;
  p.lin_bin = d.lin_bin;
  p.n_pi = d.n_p_bin;
  p.pi_min = d.p_min;
  p.pi_max = d.p_max;
  p.do_jk = d.do_jk;
  p.n_jk = d.n_jk;
  p.include_prob_max = 1;
  p.precision = precision;
  p.jk_mode = CPC_JK_LEAVE_ONE_OUT;
  p.r_mu = d.coordinates == 7 ? 1 : 0;
  if (!d.periodic_box) {
    p.sky_H = d.H_z.data();
    p.sky_dA = d.dA_z.data();
    p.sky_n = d.table_rows;
    p.sky_dz = d.dz;
    p.z_max = d.z_max;
    p.z_sep_max = d.z_sep_max;
    p.coordinates = d.coordinates;
    p.z_min = d.z_min;
    p.z_sep_min = d.z_sep_min;
  }
  p.unsorted = d.data_sorted == 1 ? 0 : (d.data_sorted == 2 ? 2 : 1);
  return p;
}

cpc_catalog_desc describe(const Cat &c) {
  cpc_catalog_desc s{};
  s.n = c.n();
  s.x = c.x.data();
  s.y = c.y.data();
  s.z = c.z.data();
  s.w = c.w.data();
  s.e1 = c.has_e ? c.e1.data() : nullptr;
  s.e2 = c.has_e ? c.e2.data() : nullptr;
  s.phi = c.phi.data();
  s.jk = c.jk.data();
  s.jk_prob = c.jkp.data();
  s.include_prob = c.inc.data();
  s.dc = c.dc.data();
  s.da = c.da.data();
  s.e3 = c.has_e3 ? c.e3.data() : nullptr;
  return s;
}

const char *kind_name(int k) {
  switch (k) {
    case CPC_KIND_DD: return "dd";
    case CPC_KIND_SHAPE_DENSITY: return "shape_density";
    case CPC_KIND_SHAPE_SHAPE: return "shape_shape";
    case CPC_KIND_DENSITY_CONVERGENCE: return "density_convergence";
    case CPC_KIND_SHAPE_CONVERGENCE: return "shape_convergence";
    default: return "ed";
  }
}

struct Options {
  std::string inp;
  int precision = CPC_PREC_SPLIT;
  int device = 0;
  bool plan = false;
};

bool parse_args(int argc, char **argv, Options &o) {
  for (int i = 1; i < argc; ++i) {
    const std::string a = argv[i];
    if (a == "--plan") {
      o.plan = true;
    } else if (a == "--precision" && i + 1 < argc) {
      const std::string v = argv[++i];
      if (v == "split") o.precision = CPC_PREC_SPLIT;
      else if (v == "fp64") o.precision = CPC_PREC_FP64;
      else return fail("--precision takes split or fp64");
    } else if (a == "--device" && i + 1 < argc) {
      o.device = std::atoi(argv[++i]);
    } else if (o.inp.empty() && a.rfind("--", 0) != 0) {
      o.inp = a;
    } else {
      return fail("unknown argument " + a);
    }
  }
  if (o.inp.empty()) return fail("no input file given!!");
  return true;
}

int run(int argc, char **argv) {
  std::cout << "Beginning corr_pc (MI355X engine " << cpc_version() << ")" << std::endl;
  Options opt;
  DataInfo d;
  if (!parse_args(argc, argv, opt) || !data_ini(d, opt.inp)) {
    std::cout << "Bad initialization!! Bye" << std::endl;
    return 1;
  }
  if (!outp_data_info(d, d.out_file + "inp.info")) {
    std::cout << "OOPS!! Something went wrong... Bye" << std::endl;
    return 1;
  }

  // Catalogs in do_corr's read order (do_corr.cpp:40-96), one rand() stream: seeded in box mode
  // (seed(), called from data_ini), the C library's initial state otherwise.
  srand(d.periodic_box ? kSeed : 1u);
  Cat S, Rs, D, Rd;
  S.name = "S", Rs.name = "Rs", D.name = "D", Rd.name = "Rd";
  bool ok = read_data(d, kShape, S);
  if (!d.periodic_box || d.n_jk > 0) ok = read_data(d, kShapeRand, Rs) && ok;
  if (!ok) {
    std::cout << "auto_corr:: OOPS!! something went wrong. Bye" << std::endl;
    return 1;
  }
  if (d.do_cross_corr) {
    ok = read_data(d, kDensity, D);
    if (!d.RsRd_same) ok = read_data(d, kDensityRand, Rd) && ok;
    if (!ok) {
      std::cout << "cross_corr:: OOPS!! something went wrong. Bye" << std::endl;
      return 1;
    }
  }
  const Cat &Rdens = d.RsRd_same ? Rs : Rd;  // grand_density = grand_shape (do_corr.cpp:133-134)

  // The passes of do_corr (do_corr.cpp:59-131), in its order.
  std::vector<Pass> passes;
  if (d.do_auto_corr) passes.push_back({"SS", "SS", &S, &S, true});
  if ((d.do_auto_corr || d.RsRd_same) && !d.periodic_box) {
    passes.push_back({"SRs", "SRs", &Rs, &S, false});
    if (!d.RsRd_same)
      passes.push_back({"RsRs", "RsRs", &Rs, &Rs, true});
    else
      passes.push_back({"RsRd", "RsRs", &Rs, &Rs, false});
  }
  if (d.do_cross_corr) {
    if (!d.RsRd_same) {
      passes.push_back({"SRd", "SRd", &Rdens, &S, false});
      passes.push_back({"RsRd", "RsRd", &Rdens, &Rs, false});
    }
    passes.push_back({"SD", "SD", &D, &S, false});
    if (!d.periodic_box) passes.push_back({"DRs", "DRs", &Rs, &D, false});
  }

  const int R = (d.do_jk == 1 && d.n_jk > 0) ? 1 + d.n_jk : 1;
  std::vector<Setup> setups;
  for (const Pass &ps : passes) setups.push_back(setup_pass(d, ps));

  if (opt.plan) {
    for (const Cat *c : {&S, &Rs, &D, &Rd}) {
      std::printf("catalog %s read=%d n=%lld has_e=%d", c->name.c_str(), (int)c->read, (long long)c->n(),
                  (int)c->has_e);
      if (c->n() > 0)
        std::printf(" jk_prob0=%.17g include_prob0=%.17g phi0=%.17g jk0=%d dc0=%.17g da0=%.17g", c->jkp[0],
                    c->inc[0], c->phi[0], c->jk[0], c->dc[0], c->da[0]);
      std::printf("\n");
    }
    std::printf("weights S=%.17g,%.17g D=%.17g,%.17g Rs=%.17g,%.17g Rd=%.17g,%.17g\n", d.S_wt.val, d.S_wt.err,
                d.D_wt.val, d.D_wt.err, d.Rs_wt.val, d.Rs_wt.err, d.Rd_wt.val, d.Rd_wt.err);
    std::printf("flags auto=%d cross=%d rebin=%d rsrd_same=%d n_jk=%d regions=%d include_prob_RR=%.17g "
                "z_sep_max=%.17g table_rows=%lld\n",
                (int)d.do_auto_corr, (int)d.do_cross_corr, (int)d.do_auto_rebin, (int)d.RsRd_same, d.n_jk, R,
                d.include_prob_RR, d.z_sep_max, (long long)d.table_rows);
    for (size_t k = 0; k < passes.size(); ++k) {
      const Setup &s = setups[k];
      std::printf("pass %s flag=%s auto=%d outer=%s inner=%s kind=%s gws=%d swapped=%d empty=%d "
                  "include_prob_max=%.17g\n",
                  passes[k].name.c_str(), passes[k].flag.c_str(), (int)passes[k].auto_pass, s.outer->name.c_str(),
                  s.inner->name.c_str(), kind_name(s.kind), s.gws, (int)s.swapped, (int)s.empty,
                  s.include_prob_max);
    }
    return 0;
  }

  // Upload, then all passes in one batch.
  cpc_ctx *ctx = nullptr;
  if (cpc_init(&ctx, opt.device) != CPC_OK) {
    std::cout << "corr_pc: no usable GPU device " << opt.device << std::endl;
    return 1;
  }
  struct Closer {
    cpc_ctx *c;
    std::vector<cpc_catalog *> cats;
    ~Closer() {
      for (cpc_catalog *x : cats) cpc_catalog_free(c, x);
      cpc_destroy(c);
    }
  } closer{ctx, {}};
  std::vector<const Cat *> hosts = {&S, &Rs, &D, &Rd};
  std::vector<cpc_catalog *> dev(hosts.size(), nullptr);
  for (size_t k = 0; k < hosts.size(); ++k) {
    if (!hosts[k]->read || hosts[k]->n() == 0) continue;
    const cpc_catalog_desc desc = describe(*hosts[k]);
    if (cpc_catalog_upload(ctx, &desc, &dev[k]) != CPC_OK) {
      std::cout << "corr_pc: upload of catalog " << hosts[k]->name << " failed: " << cpc_last_error(ctx) << std::endl;
      return 1;
    }
    closer.cats.push_back(dev[k]);
  }
  auto handle = [&](const Cat *c) {
    for (size_t k = 0; k < hosts.size(); ++k)
      if (hosts[k] == c) return dev[k];
    return (cpc_catalog *)nullptr;
  };

  std::vector<Hist> hist(passes.size());
  std::vector<cpc_params> params(passes.size());
  std::vector<cpc_stats> stats(passes.size());
  std::vector<cpc_job> jobs;
  std::vector<size_t> job_pass;
  for (size_t k = 0; k < passes.size(); ++k) {
    hist[k].alloc(R, d.n_bins, d.n_p_bin);
    const Setup &s = setups[k];
    std::cout << (passes[k].auto_pass ? "Doing auto " : "Doing ") << passes[k].name << ": outer " << s.outer->name
              << " (" << s.outer->n() << "), inner " << s.inner->name << " (" << s.inner->n() << "), "
              << kind_name(s.kind) << (s.swapped ? ", order changed" : "") << std::endl;
    if (s.shapes_undefined)
      std::cout << "  " << passes[k].name << ": the shape side carries no shapes; counted as density pairs"
                << std::endl;
    if (s.empty) continue;  // corels returns 0 pairs (corr.cpp:273-277)
    if (!d.periodic_box && d.coordinates >= 4 && d.data_sorted == 1 && !passes[k].auto_pass) {
      // no z test with coordinates 4 and 5, but corels2's sorted start pointer still skips inner
      // galaxies outside the window; every z is 0, so it passes all of them unless 0 lies inside
      // (corr.cpp:361-369, sky_calcs.cpp:158-175)
      const double zlo = d.z_sep_min == 0 ? d.z_min : d.z_sep_min, zhi = d.z_sep_max == 0 ? d.z_max : d.z_sep_max;
      if (!(0 <= zhi && 0 >= zlo)) {
        std::cout << "  " << passes[k].name << ": z = 0 lies outside the window [" << zlo << ", " << zhi
                  << "]: the sorted scan finds no pair" << std::endl;
        continue;
      }
    }
    cpc_params &p = params[k];
    p = base_params(d, opt.precision);
    p.kind = s.kind;
    p.g_with_shape = s.gws;
    p.include_prob_max = s.include_prob_max;
    p.auto_pairs = passes[k].auto_pass ? 1 : 0;
    jobs.push_back({handle(s.outer), handle(s.inner), &p, &hist[k].h, &stats[k]});
    job_pass.push_back(k);
  }
  if (!jobs.empty()) {
    cpc_stats total{};
    if (cpc_count_batch(ctx, (int)jobs.size(), jobs.data(), &total) != CPC_OK) {
      std::cout << "corr_pc: pair counting failed: " << cpc_last_error(ctx) << std::endl;
      return 1;
    }
    for (size_t j = 0; j < jobs.size(); ++j) {
      const size_t k = job_pass[j];
      std::cout << passes[k].name << ": pairs in= " << stats[k].n_binned << std::endl;
      if (stats[k].n_sparse_windows > 0)
        std::cout << "  warning: " << passes[k].name << ": " << stats[k].n_sparse_windows
                  << " outer galaxies have windows that corr_pc's forward-only start pointer would cut short "
                     "(its count then depends on its thread schedule); every pair of those windows is counted"
                  << std::endl;
    }
    std::cout << "pair counting: " << total.ms_total << " ms (" << total.ms_kernel << " ms on the GPU)" << std::endl;
  }

  auto find = [&](const std::string &name) -> Hist * {
    for (size_t k = 0; k < passes.size(); ++k)
      if (passes[k].name == name) return &hist[k];
    return nullptr;
  };
  // The RsRd_same merge (do_corr.cpp:135-154): RsRs += RsRd, SRd += SRs.
  Hist merged_RsRs, merged_SRd;
  Hist *hRsRs = find("RsRs"), *hSRd = find("SRd");
  if (d.RsRd_same && !d.periodic_box) {
    merged_RsRs.alloc(R, d.n_bins, d.n_p_bin);
    merged_SRd.alloc(R, d.n_bins, d.n_p_bin);
    if (hRsRs) merged_RsRs.add(*hRsRs);
    if (Hist *h = find("RsRd")) merged_RsRs.add(*h);
    if (hSRd) merged_SRd.add(*hSRd);
    if (Hist *h = find("SRs")) merged_SRd.add(*h);
    hRsRs = &merged_RsRs;
    hSRd = &merged_SRd;
    d.RsRs_wt.val *= 2.0;
    d.RsRs_wt.err *= 2.0;
    d.Rd_wt = d.Rs_wt;
  }

  std::cout << "Bin calculations" << std::endl;
  cpc_params out_p = base_params(d, opt.precision);
  cpc_norm norm{{d.S_wt.val, d.S_wt.err}, {d.D_wt.val, d.D_wt.err}, {d.Rs_wt.val, d.Rs_wt.err},
                {d.Rd_wt.val, d.Rd_wt.err}, d.include_prob_RR};
  auto view = [](Hist *h) { return h ? &h->h : (const cpc_hist *)nullptr; };
  cpc_bining b{};
  b.do_auto_corr = d.do_auto_corr;
  b.do_cross_corr = d.do_cross_corr;
  b.do_auto_rebin = d.do_auto_rebin;
  b.rsrd_same = d.RsRd_same;
  b.SS = view(find("SS"));
  b.SRs = view(find("SRs"));
  b.RsRs = view(hRsRs);
  b.SD = view(find("SD"));
  b.SRd = view(hSRd);
  b.DRs = view(find("DRs"));
  b.RsRd = view(find("RsRd"));
  cpc_pair_norm pn{};
  int n_files = 0;
  const cpc_status st = cpc_do_bining(&out_p, &norm, &b, d.out_file.c_str(), &pn, &n_files);
  if (st != CPC_OK) {
    std::cout << "corr_pc: writing the output files failed (status " << st << ", " << n_files << " written)"
              << std::endl;
    return 1;
  }
  d.SD_wt = {pn.SD.val, pn.SD.err};
  d.SRd_wt = {pn.SRd.val, pn.SRd.err};
  d.DRs_wt = {pn.DRs.val, pn.DRs.err};
  d.RsRd_wt = {pn.RsRd.val, pn.RsRd.err};
  std::cout << "Binning done (" << n_files << " files)" << std::endl;
  outp_data_info(d, d.out_file + "inp2.info");
  return 0;
}

}  // namespace

int main(int argc, char **argv) { return run(argc, argv); }
