/*
 * cpc.h -- C ABI of the MI355X rp-pi pair-counting engine.
 *
 * This header is the drop-in boundary for corr_pc's cross-pair loop.  The reference seam it replaces is
 *     long corels2(gal g1[], gal g2[], bin_jk &bjk, data_info &data_inf);      (corr.h:30, corr.cpp:219-459)
 * which callers reach from do_corr.cpp:70,78,110,115,125,131.  corels2 takes AoS `gal` records
 * (data_def.h:23-44), a `bin_jk` histogram tree (data_def.h:98-132) and the `data_info` mode flags
 * (data_def.h:46-96).  Here the same information crosses as plain pointers and sizes:
 *
 *   gal g1[], gal g2[]      -> cpc_catalog_desc (host SoA, one per catalog) uploaded once with
 *                              cpc_catalog_upload; the library keeps the device copy.
 *   data_info (hot fields)  -> cpc_params (geometry, kind, binning, jackknife, precision)
 *   bin_jk &bjk             -> cpc_hist (caller-owned host arrays in the reference's logical layout
 *                              [region][rp][pi], region 0 = full sample, 1..n_jk = jackknife regions)
 *   return value n_in       -> cpc_stats.n_binned
 *
 * The host applies corels2's operand-role table and larger-catalog-outer swap (corr.cpp:226-296)
 * before calling cpc_count, so `outer` is the reference's g1 after the swap and `inner` its g2.
 *
 * No function here calls exit(); every failure is a negative cpc_status with a message retrievable
 * through cpc_last_error().  A context is bound to one device and is not thread safe.
 */
#ifndef CPC_H
#define CPC_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum cpc_status {
  CPC_OK = 0,
  CPC_ERR_INVALID = -1,     /* an argument violates the documented contract */
  CPC_ERR_DEVICE = -2,      /* the HIP runtime reported an error */
  CPC_ERR_NOMEM = -3,       /* device or host allocation failed */
  CPC_ERR_UNSUPPORTED = -4  /* a mode the engine does not implement */
} cpc_status;

/* geometry */
#define CPC_GEOM_BOX 0 /* periodic box with x, y, z (data_info.periodic_box=1, coordinates=6) */
#define CPC_GEOM_SKY 1 /* RA/Dec/redshift survey, rp-pi (periodic_box=0, coordinates=0) */

/* pair kind: which per-pair function of corels.cpp runs (choose_corel_func, initialization.cpp:490-543) */
#define CPC_KIND_DD 0            /* density_density_corel, corels.cpp:48-77 */
#define CPC_KIND_SHAPE_DENSITY 1 /* shape_density_corel,   corels.cpp:206-244 */
#define CPC_KIND_SHAPE_SHAPE 2   /* shape_shape_corel,     corels.cpp:122-204 */
/* convergence kinds (which_corr 3 and 4): the convergence kappa travels in e1 (the reference's e[0]);
 * a catalog that carries only kappa passes e2 as zeros.  g_with_shape may be 0 for these kinds: the
 * reference then assigns nothing and adds the zero values its per-pass state starts with. */
#define CPC_KIND_DENSITY_CONVERGENCE 3 /* density_convergence_corel, corels.cpp:79-91 */
#define CPC_KIND_SHAPE_CONVERGENCE 4   /* shape_convergence_corel,   corels.cpp:93-120 */
/* ED, which_corr 5 (periodic box only): cos^2 of the angle between the pair's imaged 3-D separation
 * and the 3-D axis (e1, e2, e3) of the g_with_shape galaxy (of g1 unless g_with_shape is 2), into
 * data[6] with its square as err (ED_corel / ED_calc_PB, corels.cpp:246-262, PB_calcs.cpp:101-134) */
#define CPC_KIND_ED 5

/* arithmetic of the pair kernel */
#define CPC_PREC_SPLIT 0 /* FP32 classification, FP64 re-evaluation inside a guard band */
#define CPC_PREC_FP64 1  /* every candidate in FP64, shape angles through atan/cos/sin */

/* what the jackknife part of cpc_hist holds */
#define CPC_JK_LEAVE_ONE_OUT 0 /* full - touch_r, as after final_jk_bins (jk.cpp:74-81) */
#define CPC_JK_TOUCH 1         /* the raw per-region touch sums before final_jk_bins */

#define CPC_NFIELD 7 /* data[7]: 0=W, 1=e+, 2=ex, 3=e++, 4=exx, 5=e+x, 6=theta (data_def.h:105) */

/* One catalog as host SoA arrays; caller-owned, read only during cpc_catalog_upload.
 * Mirrors the fields of `gal` that the pair loop reads (data_def.h:23-44, filled in read_dat.cpp:216-325). */
typedef struct cpc_catalog_desc {
  int64_t n;                  /* number of galaxies (>= 0) */
  const double *x;            /* gal.ra.val_deg: box x coordinate */
  const double *y;            /* gal.dec.val_deg: box y coordinate */
  const double *z;            /* gal.redshift: box z coordinate (line of sight) */
  const double *w;            /* gal.wt; NULL means every weight is 1 */
  const double *e1;           /* gal.e[0]; NULL for catalogs without shapes */
  const double *e2;           /* gal.e[1]; NULL iff e1 is NULL */
  const double *phi;          /* gal.phi (elip_angle, calcs.cpp:73-83); NULL means -10 (read_dat.cpp:269) */
  const int32_t *jk;          /* gal.jk, 0 <= jk < n_jk; NULL allowed only when jackknife is off */
  const double *jk_prob;      /* gal.jk_prob (read_dat.cpp:239-255); NULL means 0 */
  const double *include_prob; /* gal.include_prob (read_dat.cpp:318-320); NULL means 0 */
  /* sky geometry only: x = RA and y = Dec in degrees, z = redshift, and the comoving distances the
   * reader looked up, gal.DC = dC_z[floor(z/dz)] and gal.DA = dA_z[floor(z/dz)] (read_dat.cpp:249-251) */
  const double *dc;
  const double *da;
  /* which_corr 5 (ED) only: the third component gal.e[2] of a 3-D axis whose first two are e1 and e2,
   * normalised to unit length as read_data does (read_dat.cpp:283-293); NULL otherwise */
  const double *e3;
} cpc_catalog_desc;

/* The data_info fields the pair loop reads (data_def.h:46-96) plus engine choices. */
typedef struct cpc_params {
  int geometry;            /* CPC_GEOM_BOX or CPC_GEOM_SKY */
  double box_size;         /* data_info.periodic_box_size */
  int kind;                /* CPC_KIND_* */
  int g_with_shape;        /* data_info.g_with_shape after the swap: 1 = outer, 2 = inner carries the shape */
  int estimator;           /* data_info.estimator (selects the e+ branch of shape_shape_corel) */
  int which_corr;          /* data_info.which_corr (0 disables the shape fields of corel_jk1, jk.cpp:150) */
  int n_rp;                /* data_info.n_bins */
  double rp_min, rp_max;   /* data_info.bin_r_min, bin_r_max */
  int lin_bin;             /* data_info.lin_bin */
  int n_pi;                /* data_info.n_p_bin */
  double pi_min, pi_max;   /* data_info.p_min, p_max */
  int do_jk;               /* data_info.do_jk: 0 or 1 */
  int n_jk;                /* data_info.n_jk */
  double include_prob_max; /* calc_temp.include_prob_max (initialization.cpp:597-599) */
  int precision;           /* CPC_PREC_* */
  int jk_mode;             /* CPC_JK_* */
  /* sky geometry only: the distance table of da_Z (initialization.cpp:23-60) that lim_calc_sky
   * interpolates (sky_calcs.cpp:33-81, 105-190): H(z) and the angular-diameter column dA_z, sampled
   * at z = k*dz for k < sky_n; data_info.z_max and data_info.z_sep_max (initialization.cpp:272-276) */
  const double *sky_H;
  const double *sky_dA;
  int64_t sky_n;
  double sky_dz;
  double z_max;
  double z_sep_max;
  /* 0: the cross pass of corels2 (corr.cpp:219-459).  1: the auto pass of corels1 (corr.cpp:103-216)
   * over one z-sorted catalog (outer and inner the same handle): each unordered pair once, found from
   * the lower galaxy i through images with PBz <= 0 (scan from j = i+1) or from any galaxy through
   * PBz = +1 (scan from j = 0), each scan ending at its first included galaxy outside the z window. */
  int auto_pairs;
  /* 0: rp-pi bins (coordinates 6).  1: r-mu bins of a periodic box (coordinates 7): the z window is
   * z_i +- rp_max (lim_calc_xyz, PB_calcs.cpp:14-23), r = sqrt(rp*rp + pi*pi) with rp = sqrt(t0*t0 +
   * t1*t1) and mu = pi / r (r_mu_calc, calcs.cpp:135-140); rp_* bin r and pi_* bin mu.  Must be 0
   * on the sky, where coordinates 1 selects r-mu. */
  int r_mu;
  /* Sky geometry only (0 in the box): data_info.coordinates, which picks the window, pi and rp of a
   * pair (lim_calc_sky, sky_calcs.cpp:105-190; choose_rp_pi_func, initialization.cpp:545-571):
   *   0  rp-pi:     pi = DC_j - DC_i, rp = dA (DA_i + DA_j)/2, z window from p_to_dz(p_min/p_max);
   *   1  r-mu:      as 0, then r = sqrt(rp^2 + pi^2) and mu = pi / r are binned (r_mu_calc), z window
   *                 z_i +- p_to_dz(rp_max) as lim_calc_sky's case 1 writes it;
   *   2  rp-ephi:   pi = ephi of the g_with_shape galaxy (pi_calc_phi, calcs.cpp:115-123);
   *   3  projected: pi = 0 (pi_calc_projected), rp = dA DA_i of the outer galaxy, pi bins of width 1;
   *   4  angular-ephi: rp = dA in degrees (rp_calc_sky_angular), pi = ephi, no z test (corr.cpp:69);
   *   5  angular projected: rp in degrees, pi = 0, pi bins of width 1, no z test.
   * With 2-5 the z window is [z_min or z_i + z_sep_min, z_i + z_sep_max] (sky_calcs.cpp:158-175);
   * pi_max > 600 with 0 or 2 takes the projected rp, and with 0 also pi = 0 (initialization.cpp:550-554).
   * With 2 and 4 the shape_shape kind takes its estimator-0 branch whatever `estimator` says
   * (corels.cpp:139). */
  int coordinates;
  double z_min;     /* data_info.z_min: the window's lower edge for coordinates 2-5 when z_sep_min is 0 */
  double z_sep_min; /* data_info.z_sep_min (coordinates 2-5) */
  /* data_sorted 0: do_pair_unsorted (corr.cpp:36-52) ends no scan, so every image's scan meets every
   * inner galaxy: the PBz = +1 image takes no first-galaxy rule, and the auto pass (corels1) pairs
   * the galaxies after the outer one in catalog order under every image (corr.cpp:167-182).  The
   * catalogs need not be sorted.  0: data_sorted 1.  2: data_sorted 2 (the sky only, catalogs in
   * Dec order): the pairs of 1, that is every pair of each outer galaxy's window; corels2's Dec start
   * pointer only moves forward (corr.cpp:356-362), so where windows are not ordered like the
   * galaxies (their Dec half-width changes with redshift) or a window is empty, the reference skips
   * some of those pairs, depending on its OpenMP schedule: cpc_stats.n_sparse_windows counts the
   * outer galaxies whose window one thread's pointer would cut short. */
  int unsorted;
} cpc_params;

/* Output histograms, caller-allocated host memory, overwritten by cpc_count.
 * R = 1 + n_jk when jackknife is on, else 1.  Index (r, k, p) = (r*n_rp + k)*n_pi + p.
 * Mirrors bin_jk: b[k].p_bin[p] is region 0, jk[r-1].b[k].p_bin[p] is region r (data_def.h:98-132). */
typedef struct cpc_hist {
  int64_t *num;    /* [R*n_rp*n_pi]   P_bin.num */
  double *wt_num;  /* [R*n_rp*n_pi]   P_bin.wt_num */
  double *val;     /* [R*n_rp*n_pi*7] P_bin.data[f].val */
  double *err;     /* [R*n_rp*n_pi*7] P_bin.data[f].err */
  int64_t *num_rp; /* [R*n_rp]        bin.num */
} cpc_hist;

typedef struct cpc_stats {
  int64_t n_binned;         /* pairs that landed in a valid (rp, pi) bin: corels2's n_in (corr.cpp:455) */
  int64_t n_candidates;     /* candidate pairs the kernel evaluated */
  int64_t n_fp64_reeval;    /* split precision: candidates re-evaluated in FP64 */
  int64_t n_sparse_windows; /* outer galaxies whose z window held no inner galaxy while later ones exist */
  double ms_total;          /* wall time of cpc_count */
  double ms_kernel;         /* device time of the pair kernels (HIP events on the engine's stream) */
  double ms_primary;        /* device time of the pair kernel (block side = outer catalog) */
  double ms_reduce;         /* device time of the jackknife reduction (raw targets into the full sample
                               and the touch histograms) */
  double ms_build;          /* wall time of the cell-list construction */
  int64_t n_items_primary;  /* work items (workgroups) of the primary kernel */
  int64_t n_host_pairs;     /* sky: guard-band pairs binned on the host with the host libm */
  double ms_cells;          /* device time of the cell-list construction (HIP events around its key, sort,
                               bounds and gather kernels and the copies of its tables); in a batch the
                               first job's stats and the total carry it */
} cpc_stats;

typedef struct cpc_ctx cpc_ctx;
typedef struct cpc_catalog cpc_catalog;

/* Context on HIP device `device`. */
cpc_status cpc_init(cpc_ctx **ctx, int device);
void cpc_destroy(cpc_ctx *ctx);
/* Message of the last failure on this context (empty string if none). */
const char *cpc_last_error(const cpc_ctx *ctx);
/* Library version string. */
const char *cpc_version(void);

/* Copy one catalog to the device.  Replaces the reference's `new gal[n]` + read_data
 * (do_corr.cpp:41,47,85,94); the arrays are not referenced after the call returns. */
cpc_status cpc_catalog_upload(cpc_ctx *ctx, const cpc_catalog_desc *desc, cpc_catalog **cat);
void cpc_catalog_free(cpc_ctx *ctx, cpc_catalog *cat);

/* One cross pass: replaces corels2(g1=outer, g2=inner, bjk, data_inf) (corr.cpp:219-459) after the
 * host has resolved roles and the swap.  Writes `out` (overwrite, not accumulate) and, if non-NULL,
 * `stats`.  outer and inner may be the same catalog handle. */
cpc_status cpc_count(cpc_ctx *ctx, const cpc_catalog *outer, const cpc_catalog *inner,
                     const cpc_params *params, cpc_hist *out, cpc_stats *stats);

/* One pass of a batch: the arguments of one cpc_count call. */
typedef struct cpc_job {
  const cpc_catalog *outer;
  const cpc_catalog *inner;
  const cpc_params *params;
  cpc_hist *out;
  cpc_stats *stats; /* may be NULL */
} cpc_job;

/* Size of the library's static device scratch array used by cpc_count_noalloc, in MiB. */
#define CPC_NOALLOC_SCRATCH_MIB 256

/* cpc_count without device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory during the call (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Arguments, results and errors are those of cpc_count; the
 * catalogs are uploaded beforehand (cpc_catalog_upload allocates).  Every device buffer of the pass
 * (cell lists, sort space, work items, accumulators) is carved from one static device array of
 * CPC_NOALLOC_SCRATCH_MIB MiB that each device the library runs on holds until the process ends; a
 * pass that needs more fails with CPC_ERR_NOMEM and a message (use cpc_count then).  The call
 * allocates and frees no device memory.  All calls on one device share that array: calls that overlap
 * on one device (two contexts, or two threads) race and must be serialised by the caller. */
cpc_status cpc_count_noalloc(cpc_ctx *ctx, const cpc_catalog *outer, const cpc_catalog *inner,
                             const cpc_params *params, cpc_hist *out, cpc_stats *stats);

/* Many passes at once: every job's `out` and `stats` as cpc_count(job) would write them (integer
 * counts identical; FP64 sums equal up to summation order).  The reference runs its passes one after
 * another (do_corr.cpp:63-131, once per catalog pair); here the cell lists of all jobs are built
 * by one sort, and the work items of all box jobs are merged into one launch (one more per further
 * pair kind, precision or auto pass in the batch), ordered by cost across jobs, so small catalogs
 * fill the device together and only one tail remains; one more launch reduces every job's
 * jackknife targets.  Sky jobs run one after another
 * within the call.  Jobs write their outputs concurrently, so no two jobs may share an `out` or
 * `stats`.  Jobs are validated before any work starts: on failure no output is written and
 * the message names the job.  The records of all jobs are resident on the device together.
 * `total`, if non-NULL, receives summed counters; its times are those of the whole call.  In each
 * job's stats the kernel times are those of the launches that served it (shared by the batch). */
cpc_status cpc_count_batch(cpc_ctx *ctx, int n_jobs, const cpc_job *jobs, cpc_stats *total);

/* The thresholds the engine bins rp with.  Box: thr[k] (k = 0..n_rp) is the smallest double s >= 0
 * for which the reference's bin index of rp = sqrt(s) (corels.cpp:31-34, with the host libm's log10)
 * is >= k.  Sky: the same on rp itself (rp_calc_sky, sky_calcs.cpp:91-96).  Exposed so that callers
 * and tests can check them against the host libm. */
cpc_status cpc_rp_thresholds(const cpc_params *params, double *thr);

/* The output bin edges corr_pc reports and uses for its analytic RR: bin_log / bin_lin and P_bin_ini
 * (initialization.cpp:357-444) step the edges as FP32 running sums (log edges as pow(10, running FP32
 * log10)); they are returned here as the doubles the reference stores.  rp_edges[n_rp + 1] and
 * pi_edges[n_pi + 1]; bin k spans [edges[k], edges[k+1]].  These edges label the output only: the
 * pair kernels bin with the FP64 arithmetic of corel_bin. */
cpc_status cpc_bin_edges(const cpc_params *params, double *rp_edges, double *pi_edges);

/* Analytic RR of a periodic box in rp-pi (RR_PB_calc, PB_calcs.cpp:63-85, coordinates 6), which the
 * reference puts in place of the counted RR in box mode (bins_calcs.cpp:44-49, 71-77):
 * wt_num[k*n_pi + q] = pi (b_max^2 - b_min^2) (p_max - p_min) / L^3 over the edges of cpc_bin_edges;
 * the reference sets num = 1 in every bin. */
cpc_status cpc_box_rr(const cpc_params *params, double *wt_num);

/* ---- Output side of the cross correlation (estimator 0) ---- */

/* Sum of a catalog's weights and of their squares, in read order, as read_data accumulates them
 * into data_info.{S,D,Rs,Rd}_wt (read_dat.cpp:322-323); w == NULL means unit weights. */
typedef struct cpc_wsum {
  double val;
  double err;
} cpc_wsum;
cpc_status cpc_weight_sum(const double *w, int64_t n, cpc_wsum *out);

/* The data_info inputs of final_calc besides the binning in cpc_params (bins_calcs.cpp:30-52). */
typedef struct cpc_norm {
  cpc_wsum S, D, Rs, Rd;  /* shape, density, shape-random and density-random catalogs */
  double include_prob_RR; /* data_info.include_prob_RR (1 unless the randoms are subsampled) */
} cpc_norm;

/* The cross branch of do_bining (do_corr.cpp:326-431) on the four pair-set histograms that
 * cpc_count returns in CPC_JK_LEAVE_ONE_OUT layout (SD, SRd, DRs, RsRd): final_calc with
 * final_calc_bins and int_calc (bins_calcs.cpp:30-52, 94-256, 331-371; in the box geometry the
 * analytic RR of RR_PB_calc, PB_calcs.cpp:63-85, replaces rr as the reference does), then every
 * file of that branch under the prefix `out_file` (data_info.out_file): bins_cross_{final,SD,SR,
 * DR,RR}.dat and their bins2D_ versions (outp_bins / outp_2Dbins, outp.cpp:45-74, 291-321), per
 * jackknife region bins_cross_jk<r>_*.dat and bins2D_cross_jk<r>_*.dat, and the jackknife summary
 * bins_cross_jk_final.dat / bins2D_cross_jk_final.dat (jk_final / jk_final2D,
 * bins_calcs.cpp:393-462).  Numbers are written as the reference's streams write them (six
 * significant digits).  Estimator 0, or 4 (final_calc_mean_projected, bins_calcs.cpp:258-329), and
 * which_corr 0-2 only (CPC_ERR_UNSUPPORTED otherwise).  `n_files` (may be NULL) receives the number
 * of files written. */
cpc_status cpc_write_cross(const cpc_params *params, const cpc_norm *norm, const cpc_hist *sd, const cpc_hist *sr,
                           const cpc_hist *dr, const cpc_hist *rr, const char *out_file, int *n_files);

/* The pair-set histograms and mode flags that do_bining reads from `bins B` and data_info
 * (do_corr.cpp:203-433).  Each histogram is as cpc_count returns it (CPC_JK_LEAVE_ONE_OUT), after the
 * caller's RsRd_same merge (do_corr.cpp:135-154: RsRs += RsRd and SRd += SRs, elementwise over every
 * array); NULL stands for a set that was not counted, which the reference holds as zero bins. */
typedef struct cpc_bining {
  int do_auto_corr;  /* data_info.do_auto_corr: estimators 1, 2, 3 (initialization.cpp:163-185) */
  int do_cross_corr; /* data_info.do_cross_corr: estimators 0, 3, 4 */
  int do_auto_rebin; /* data_info.do_auto_rebin: estimators 2, 3 */
  int rsrd_same;     /* data_info.RsRd_same: no density randoms (initialization.cpp:187-192) */
  const cpc_hist *SS, *SRs, *RsRs; /* auto pair sets */
  const cpc_hist *SD, *SRd, *DRs, *RsRd; /* cross pair sets */
} cpc_bining;

/* The pair normalisations data_info holds after do_bining: {SD,SRd,DRs,RsRd}_wt as the last
 * final_calc / final_calc_auto call left them (zero when none ran), which outp_data_info reports. */
typedef struct cpc_pair_norm {
  cpc_wsum SD, SRd, DRs, RsRd;
} cpc_pair_norm;

/* do_bining (do_corr.cpp:203-433): the raw bins_auto_SR.temp / bins_cross_SR.temp (SRs, SRd), then
 * the auto branch (re_bin, bins_calcs.cpp:465-527, when do_auto_rebin; final_calc_auto,
 * bins_calcs.cpp:54-91; bins_auto_*.dat, bins2D_auto_*.dat, per-region and jackknife-summary files)
 * and the cross branch (as cpc_write_cross), under the prefix `out_file`, with the reference's file
 * names and formatting.  which_corr 0-2, estimators 0-4, do_jk 0 or 1.  `pair_norm` and `n_files`
 * may be NULL. */
cpc_status cpc_do_bining(const cpc_params *params, const cpc_norm *norm, const cpc_bining *sets,
                         const char *out_file, cpc_pair_norm *pair_norm, int *n_files);

/* The reference's per-galaxy random stream: out[k] = rand() / (RAND_MAX + 1.) for the draws numbered
 * skip .. skip+n-1 after srand(seed), with the host C library's rand().  corr_pc seeds 124693517 in
 * seed() (jk.cpp:40-43) and draws jk_prob (read_dat.cpp:243), a random phi (read_dat.cpp:307) and
 * include_prob (read_dat.cpp:318-320) in catalog read order (do_corr.cpp:40-96). */
cpc_status cpc_rand_stream(uint32_t seed, int64_t skip, int64_t n, double *out);

/* Engine tuning knobs: column width in units of rp_max, block galaxies per work item and threads per
 * work group (a multiple of 64, at most 512 for the box and 256 for the sky kernel; 0 keeps the
 * automatic choice).  Used by the autotuner; results do not depend on them beyond floating-point
 * summation order. */
cpc_status cpc_set_tuning(cpc_ctx *ctx, double col_frac, int item_size, int block_threads);

/* A choice of the cpc_set_tuning knobs (0 = automatic). */
typedef struct cpc_tuning {
  double col_frac;
  int item_size;
  int block_threads;
} cpc_tuning;

/* Autotuner of the decomposition (no counterpart in the reference, whose OpenMP loop has no such
 * knobs): coordinate descent from the automatic choice over column width, block galaxies per work
 * item and threads per work group, each candidate timed by the device time of cpc_count_batch on
 * `jobs` (pass a representative subset to tune quickly).  A candidate replaces the current best
 * only when at least 3% faster.  On return the context uses `best` (as after cpc_set_tuning), and
 * every job's output and stats hold a valid result of the jobs.  `tried` (may be NULL) receives
 * the number of batches run. */
cpc_status cpc_autotune(cpc_ctx *ctx, int n_jobs, const cpc_job *jobs, cpc_tuning *best, int *tried);

/* Device check of the FP32 hardware approximations the split kernels classify with (no counterpart
 * in the reference, which bins in FP64 with the host libm, corels.cpp:31-34).  For every FP32 value s
 * in [s_lo, s_hi] (both positive and finite) the device evaluates the same instructions as the pair
 * kernels: out[0] = max |log2~(s) - log2(s)| / max(1, |log2(s)|) for the hardware logarithm
 * (v_log_f32), out[1] = max |rsq~(s) sqrt(s) - 1| for the hardware reciprocal square root
 * (v_rsq_f32), against FP64 references.  The split guard bands take both as at most 2^-21 (about
 * 4.8e-7); a larger value means the guard bands do not cover this device. */
cpc_status cpc_fp32_error(cpc_ctx *ctx, float s_lo, float s_hi, double out[2]);

/* No counterpart in the reference: a check of the device functions the sky kernels bin with.  The
 * device cos (fn 0) or acos (fn 1) of the sky kernels against the host C library's on n points
 * spread evenly over [lo, hi] (n <= 2^26): *max_ulp = max |device - host| in units of the last
 * place of the host value.  The sky guard band takes the cosine's as at most 2 ulps (its bound on
 * the device/host difference of the great-circle sum follows from that) and the arc cosine's as
 * within 1e-15 relative. */
cpc_status cpc_libm_error(cpc_ctx *ctx, int fn, double lo, double hi, int64_t n, double *max_ulp);

#ifdef __cplusplus
}
#endif

#endif /* CPC_H */
