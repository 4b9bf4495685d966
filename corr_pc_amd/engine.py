"""ctypes bindings of the engine's C ABI (include/cpc.h) and numpy-facing helpers.

The product path is libcpc.so (HIP kernels for gfx950).  Loading fails loudly when the library is
missing; there is no CPU fallback.
"""
import ctypes
import os
from ctypes import POINTER, c_char_p, c_double, c_int, c_int32, c_int64, c_void_p
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcpc.so")

CPC_OK = 0
CPC_ERR_INVALID = -1
CPC_ERR_DEVICE = -2
CPC_ERR_NOMEM = -3
CPC_ERR_UNSUPPORTED = -4

GEOM_BOX = 0
GEOM_SKY = 1
KIND_DD = 0
KIND_SHAPE_DENSITY = 1
KIND_SHAPE_SHAPE = 2
KIND_DENSITY_CONVERGENCE = 3
KIND_SHAPE_CONVERGENCE = 4
KIND_ED = 5
PREC_SPLIT = 0
PREC_FP64 = 1
JK_LEAVE_ONE_OUT = 0
JK_TOUCH = 1
NFIELD = 7

_dp = POINTER(c_double)
_ip = POINTER(c_int32)
_lp = POINTER(c_int64)


class CatalogDesc(ctypes.Structure):
    _fields_ = [("n", c_int64), ("x", _dp), ("y", _dp), ("z", _dp), ("w", _dp), ("e1", _dp), ("e2", _dp),
                ("phi", _dp), ("jk", _ip), ("jk_prob", _dp), ("include_prob", _dp), ("dc", _dp), ("da", _dp),
                ("e3", _dp)]


class Params(ctypes.Structure):
    _fields_ = [("geometry", c_int), ("box_size", c_double), ("kind", c_int), ("g_with_shape", c_int),
                ("estimator", c_int), ("which_corr", c_int), ("n_rp", c_int), ("rp_min", c_double),
                ("rp_max", c_double), ("lin_bin", c_int), ("n_pi", c_int), ("pi_min", c_double),
                ("pi_max", c_double), ("do_jk", c_int), ("n_jk", c_int), ("include_prob_max", c_double),
                ("precision", c_int), ("jk_mode", c_int), ("sky_H", _dp), ("sky_dA", _dp), ("sky_n", c_int64),
                ("sky_dz", c_double), ("z_max", c_double), ("z_sep_max", c_double), ("auto_pairs", c_int),
                ("r_mu", c_int), ("coordinates", c_int), ("z_min", c_double), ("z_sep_min", c_double),
                ("unsorted", c_int)]

    def attach_table(self, H, dA, dz, z_max, z_sep_max):
        """Sky geometry: keep references to the distance table arrays and point the struct at them."""
        self._table = (np.ascontiguousarray(H, np.float64), np.ascontiguousarray(dA, np.float64))
        self.sky_H = self._table[0].ctypes.data_as(_dp)
        self.sky_dA = self._table[1].ctypes.data_as(_dp)
        self.sky_n = len(self._table[0])
        self.sky_dz, self.z_max, self.z_sep_max = dz, z_max, z_sep_max


class HistC(ctypes.Structure):
    _fields_ = [("num", _lp), ("wt_num", _dp), ("val", _dp), ("err", _dp), ("num_rp", _lp)]


class Stats(ctypes.Structure):
    _fields_ = [("n_binned", c_int64), ("n_candidates", c_int64), ("n_fp64_reeval", c_int64),
                ("n_sparse_windows", c_int64), ("ms_total", c_double), ("ms_kernel", c_double),
                ("ms_primary", c_double), ("ms_reduce", c_double), ("ms_build", c_double),
                ("n_items_primary", c_int64), ("n_host_pairs", c_int64), ("ms_cells", c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class WSum(ctypes.Structure):
    """cpc_wsum: sum of weights and of their squares."""
    _fields_ = [("val", c_double), ("err", c_double)]


class Norm(ctypes.Structure):
    """cpc_norm: the weight sums of the S, D, Rs, Rd catalogs and include_prob_RR."""
    _fields_ = [("S", WSum), ("D", WSum), ("Rs", WSum), ("Rd", WSum), ("include_prob_RR", c_double)]


class Bining(ctypes.Structure):
    """cpc_bining: do_bining's mode flags and its seven pair-set histograms (NULL = not counted)."""
    _fields_ = [("do_auto_corr", c_int), ("do_cross_corr", c_int), ("do_auto_rebin", c_int), ("rsrd_same", c_int),
                ("SS", POINTER(HistC)), ("SRs", POINTER(HistC)), ("RsRs", POINTER(HistC)), ("SD", POINTER(HistC)),
                ("SRd", POINTER(HistC)), ("DRs", POINTER(HistC)), ("RsRd", POINTER(HistC))]


class PairNorm(ctypes.Structure):
    """cpc_pair_norm: data_info.{SD,SRd,DRs,RsRd}_wt after do_bining."""
    _fields_ = [("SD", WSum), ("SRd", WSum), ("DRs", WSum), ("RsRd", WSum)]


class TuningC(ctypes.Structure):
    _fields_ = [("col_frac", c_double), ("item_size", c_int), ("block_threads", c_int)]


class JobC(ctypes.Structure):
    """cpc_job: one pass of a cpc_count_batch call."""
    _fields_ = [("outer", c_void_p), ("inner", c_void_p), ("params", POINTER(Params)), ("out", POINTER(HistC)),
                ("stats", POINTER(Stats))]


class EngineError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"cpc status {status}: {msg}")
        self.status = status


@dataclass
class Catalog:
    """One catalog in SoA form; the fields of corr_pc's `gal` that the pair loop reads."""
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    w: Optional[np.ndarray] = None
    e1: Optional[np.ndarray] = None
    e2: Optional[np.ndarray] = None
    phi: Optional[np.ndarray] = None
    jk: Optional[np.ndarray] = None
    jk_prob: Optional[np.ndarray] = None
    include_prob: Optional[np.ndarray] = None
    dc: Optional[np.ndarray] = None
    da: Optional[np.ndarray] = None
    e3: Optional[np.ndarray] = None  # third component of a 3-D axis (which_corr 5, ED)
    _keep: list = field(default_factory=list, repr=False)

    def __post_init__(self):
        for name in ("x", "y", "z", "w", "e1", "e2", "phi", "jk_prob", "include_prob", "dc", "da", "e3"):
            v = getattr(self, name)
            if v is not None:
                setattr(self, name, np.ascontiguousarray(v, dtype=np.float64))
        if self.jk is not None:
            self.jk = np.ascontiguousarray(self.jk, dtype=np.int32)
        n = len(self.x)
        for name in ("y", "z", "w", "e1", "e2", "phi", "jk", "jk_prob", "include_prob", "dc", "da", "e3"):
            v = getattr(self, name)
            if v is not None and len(v) != n:
                raise ValueError(f"catalog field {name} has length {len(v)} != {n}")

    @property
    def n(self):
        return len(self.x)

    def subset(self, idx):
        """Catalog restricted to rows idx (keeps order)."""
        kw = {}
        for name in ("x", "y", "z", "w", "e1", "e2", "phi", "jk", "jk_prob", "include_prob", "dc", "da", "e3"):
            v = getattr(self, name)
            kw[name] = None if v is None else v[idx]
        return Catalog(**kw)

    def desc(self):
        # fields may have been replaced after construction: pass contiguous buffers of the C types,
        # kept alive with the catalog
        self._keep = []

        def p(a, t=_dp):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.int32 if t is _ip else np.float64)
            self._keep.append(a)
            return a.ctypes.data_as(t)
        return CatalogDesc(self.n, p(self.x), p(self.y), p(self.z), p(self.w), p(self.e1), p(self.e2),
                           p(self.phi), p(self.jk, _ip), p(self.jk_prob), p(self.include_prob), p(self.dc),
                           p(self.da), p(self.e3))


def make_params(box_size, n_rp, rp_min, rp_max, n_pi, pi_min, pi_max, kind=KIND_DD, g_with_shape=2,
                estimator=0, which_corr=None, lin_bin=0, do_jk=0, n_jk=0, include_prob_max=1.0,
                precision=PREC_SPLIT, jk_mode=JK_LEAVE_ONE_OUT, r_mu=0):
    if which_corr is None:
        which_corr = {KIND_DD: 0, KIND_SHAPE_DENSITY: 1, KIND_SHAPE_SHAPE: 2, KIND_DENSITY_CONVERGENCE: 3,
                      KIND_SHAPE_CONVERGENCE: 4, KIND_ED: 5}[kind]
    p = Params(GEOM_BOX, box_size, kind, g_with_shape, estimator, which_corr, n_rp, rp_min, rp_max, lin_bin,
               n_pi, pi_min, pi_max, do_jk, n_jk, include_prob_max, precision, jk_mode)
    p.r_mu = r_mu
    return p


def make_sky_params(table, n_rp, rp_min, rp_max, n_pi, pi_min, pi_max, coordinates=0, z_min=0.0, z_sep_min=0.0,
                    **kw):
    """Sky geometry (RA, Dec in degrees and redshift): `table` is (H, dA, dz, z_max, z_sep_max) as the
    reference's data_ini derives them (initialization.cpp:223, 272-276); `coordinates` is the
    reference's sky choice 0-5 (cpc_params.coordinates), `z_min` and `z_sep_min` the data_info values
    that coordinates 2-5 read; the remaining keywords are those of make_params."""
    p = make_params(0.0, n_rp, rp_min, rp_max, n_pi, pi_min, pi_max, **kw)
    p.geometry = GEOM_SKY
    p.attach_table(*table)
    p.coordinates, p.z_min, p.z_sep_min = coordinates, z_min, z_sep_min
    return p


class Histogram:
    """The reference's bin_jk in array form: region 0 is the full sample, 1..n_jk the jackknife regions."""

    def __init__(self, n_regions, n_rp, n_pi):
        self.R, self.n_rp, self.n_pi = n_regions, n_rp, n_pi
        self.num = np.zeros((n_regions, n_rp, n_pi), np.int64)
        self.wt_num = np.zeros((n_regions, n_rp, n_pi), np.float64)
        self.val = np.zeros((n_regions, n_rp, n_pi, NFIELD), np.float64)
        self.err = np.zeros((n_regions, n_rp, n_pi, NFIELD), np.float64)
        self.num_rp = np.zeros((n_regions, n_rp), np.int64)

    @classmethod
    def for_params(cls, p):
        R = 1 + p.n_jk if (p.do_jk == 1 and p.n_jk > 0) else 1
        return cls(R, p.n_rp, p.n_pi)

    @classmethod
    def many(cls, params_list):
        """Histograms of many passes as views of two zeroed buffers (one int64, one float64), with
        their addresses kept for c_struct: one allocation for a whole batch."""
        shapes = [(1 + p.n_jk if (p.do_jk == 1 and p.n_jk > 0) else 1, p.n_rp, p.n_pi) for p in params_list]
        ibuf = np.zeros(sum(R * r * q + R * r for R, r, q in shapes), np.int64)
        fbuf = np.zeros(sum(R * r * q * (1 + 2 * NFIELD) for R, r, q in shapes), np.float64)
        ib, fb = ibuf.ctypes.data, fbuf.ctypes.data
        out, io, fo = [], 0, 0
        for R, r, q in shapes:
            h = cls.__new__(cls)
            h.R, h.n_rp, h.n_pi = R, r, q
            n = R * r * q
            h.num = ibuf[io:io + n].reshape(R, r, q)
            h.num_rp = ibuf[io + n:io + n + R * r].reshape(R, r)
            h.wt_num = fbuf[fo:fo + n].reshape(R, r, q)
            h.val = fbuf[fo + n:fo + n * (1 + NFIELD)].reshape(R, r, q, NFIELD)
            h.err = fbuf[fo + n * (1 + NFIELD):fo + n * (1 + 2 * NFIELD)].reshape(R, r, q, NFIELD)
            h._addr = (ib + 8 * io, fb + 8 * fo, fb + 8 * (fo + n), fb + 8 * (fo + n * (1 + NFIELD)), ib + 8 * (io + n))
            io += n + R * r
            fo += n * (1 + 2 * NFIELD)
            out.append(h)
        return out

    def c_struct(self):
        a = getattr(self, "_addr", None)
        if a is not None:
            return HistC(ctypes.cast(a[0], _lp), ctypes.cast(a[1], _dp), ctypes.cast(a[2], _dp), ctypes.cast(a[3], _dp),
                         ctypes.cast(a[4], _lp))
        return HistC(self.num.ctypes.data_as(_lp), self.wt_num.ctypes.data_as(_dp), self.val.ctypes.data_as(_dp),
                     self.err.ctypes.data_as(_dp), self.num_rp.ctypes.data_as(_lp))


_lib = None


def load_library(path=LIB_PATH):
    """Load libcpc.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(f"{path} is missing: build it with `python -m corr_pc_amd.build`")
    lib = ctypes.CDLL(path)
    lib.cpc_init.argtypes = [POINTER(c_void_p), c_int]
    lib.cpc_init.restype = c_int
    lib.cpc_destroy.argtypes = [c_void_p]
    lib.cpc_destroy.restype = None
    lib.cpc_last_error.argtypes = [c_void_p]
    lib.cpc_last_error.restype = c_char_p
    lib.cpc_version.argtypes = []
    lib.cpc_version.restype = c_char_p
    lib.cpc_catalog_upload.argtypes = [c_void_p, POINTER(CatalogDesc), POINTER(c_void_p)]
    lib.cpc_catalog_upload.restype = c_int
    lib.cpc_catalog_free.argtypes = [c_void_p, c_void_p]
    lib.cpc_catalog_free.restype = None
    lib.cpc_count.argtypes = [c_void_p, c_void_p, c_void_p, POINTER(Params), POINTER(HistC), POINTER(Stats)]
    lib.cpc_count.restype = c_int
    lib.cpc_count_batch.argtypes = [c_void_p, c_int, POINTER(JobC), POINTER(Stats)]
    lib.cpc_count_batch.restype = c_int
    lib.cpc_autotune.argtypes = [c_void_p, c_int, POINTER(JobC), POINTER(TuningC), POINTER(c_int)]
    lib.cpc_autotune.restype = c_int
    lib.cpc_rp_thresholds.argtypes = [POINTER(Params), _dp]
    lib.cpc_rp_thresholds.restype = c_int
    lib.cpc_rand_stream.argtypes = [ctypes.c_uint32, c_int64, c_int64, _dp]
    lib.cpc_rand_stream.restype = c_int
    lib.cpc_set_tuning.argtypes = [c_void_p, c_double, c_int, c_int]
    lib.cpc_set_tuning.restype = c_int
    lib.cpc_bin_edges.argtypes = [POINTER(Params), _dp, _dp]
    lib.cpc_bin_edges.restype = c_int
    lib.cpc_box_rr.argtypes = [POINTER(Params), _dp]
    lib.cpc_box_rr.restype = c_int
    lib.cpc_weight_sum.argtypes = [_dp, c_int64, POINTER(WSum)]
    lib.cpc_weight_sum.restype = c_int
    lib.cpc_write_cross.argtypes = [POINTER(Params), POINTER(Norm), POINTER(HistC), POINTER(HistC), POINTER(HistC),
                                    POINTER(HistC), c_char_p, POINTER(c_int)]
    lib.cpc_write_cross.restype = c_int
    lib.cpc_do_bining.argtypes = [POINTER(Params), POINTER(Norm), POINTER(Bining), c_char_p, POINTER(PairNorm),
                                  POINTER(c_int)]
    lib.cpc_do_bining.restype = c_int
    lib.cpc_count_noalloc.argtypes = [c_void_p, c_void_p, c_void_p, POINTER(Params), POINTER(HistC), POINTER(Stats)]
    lib.cpc_count_noalloc.restype = c_int
    lib.cpc_fp32_error.argtypes = [c_void_p, ctypes.c_float, ctypes.c_float, _dp]
    lib.cpc_fp32_error.restype = c_int
    lib.cpc_libm_error.argtypes = [c_void_p, c_int, c_double, c_double, ctypes.c_int64, _dp]
    lib.cpc_libm_error.restype = c_int
    _lib = lib
    return lib


EXPORTED_SYMBOLS = ["cpc_init", "cpc_destroy", "cpc_last_error", "cpc_version", "cpc_catalog_upload",
                    "cpc_catalog_free", "cpc_count", "cpc_count_batch", "cpc_rp_thresholds", "cpc_rand_stream",
                    "cpc_set_tuning", "cpc_autotune", "cpc_bin_edges", "cpc_box_rr", "cpc_weight_sum",
                    "cpc_write_cross", "cpc_do_bining", "cpc_fp32_error", "cpc_count_noalloc", "cpc_libm_error"]


REFERENCE_SEED = 124693517  # seed(), jk.cpp:40-43


def rand_stream(n, skip=0, seed=REFERENCE_SEED):
    """Draws skip..skip+n-1 of rand()/(RAND_MAX+1.) after srand(seed), host C library."""
    lib = load_library()
    out = np.zeros(int(n), np.float64)
    rc = lib.cpc_rand_stream(seed, int(skip), int(n), out.ctypes.data_as(_dp))
    if rc != CPC_OK:
        raise EngineError(rc, "cpc_rand_stream failed")
    return out


def bin_edges(params):
    """(rp_edges[n_rp+1], pi_edges[n_pi+1]): the reference's FP32-stepped output edges, see include/cpc.h."""
    lib = load_library()
    re = np.zeros(params.n_rp + 1, np.float64)
    pe = np.zeros(params.n_pi + 1, np.float64)
    rc = lib.cpc_bin_edges(ctypes.byref(params), re.ctypes.data_as(_dp), pe.ctypes.data_as(_dp))
    if rc != CPC_OK:
        raise EngineError(rc, "cpc_bin_edges failed")
    return re, pe


def box_rr(params):
    """Analytic periodic-box RR wt_num[n_rp, n_pi] (RR_PB_calc); the reference sets num = 1."""
    lib = load_library()
    out = np.zeros((params.n_rp, params.n_pi), np.float64)
    rc = lib.cpc_box_rr(ctypes.byref(params), out.ctypes.data_as(_dp))
    if rc != CPC_OK:
        raise EngineError(rc, "cpc_box_rr failed")
    return out


def weight_sum(cat):
    """cpc_weight_sum over a Catalog's weights (unit weights when it has none)."""
    lib = load_library()
    out = WSum()
    w = None if cat.w is None else cat.w.ctypes.data_as(_dp)
    rc = lib.cpc_weight_sum(w, cat.n, ctypes.byref(out))
    if rc != CPC_OK:
        raise EngineError(rc, "cpc_weight_sum failed")
    return out


def write_cross(params, norm, sd, sr, dr, rr, out_prefix):
    """cpc_write_cross: the cross branch's estimator and every .dat file under `out_prefix`, from the
    four pair-set Histograms (leave-one-out layout).  Returns the number of files written."""
    lib = load_library()
    hs = [h.c_struct() for h in (sd, sr, dr, rr)]
    n = c_int(0)
    rc = lib.cpc_write_cross(ctypes.byref(params), ctypes.byref(norm), *[ctypes.byref(h) for h in hs],
                             out_prefix.encode(), ctypes.byref(n))
    if rc != CPC_OK:
        raise EngineError(rc, "cpc_write_cross failed")
    return n.value


AUTO_SETS = ("SS", "SRs", "RsRs")
CROSS_SETS = ("SD", "SRd", "DRs", "RsRd")


def do_bining(params, norm, hists, out_prefix, do_auto_corr=0, do_cross_corr=1, do_auto_rebin=0, rsrd_same=0):
    """cpc_do_bining: do_bining's estimators and every output file under `out_prefix`.

    `hists` maps pair-set names (AUTO_SETS, CROSS_SETS) to Histograms in leave-one-out layout; a
    missing set is one that was not counted.  Returns (number of files written, PairNorm)."""
    lib = load_library()
    unknown = set(hists) - set(AUTO_SETS + CROSS_SETS)
    if unknown:
        raise ValueError(f"unknown pair sets {sorted(unknown)}")
    b = Bining(do_auto_corr, do_cross_corr, do_auto_rebin, rsrd_same)
    keep = {}
    for name, h in hists.items():
        if h is not None:
            keep[name] = h.c_struct()
            setattr(b, name, ctypes.pointer(keep[name]))
    pn = PairNorm()
    n = c_int(0)
    rc = lib.cpc_do_bining(ctypes.byref(params), ctypes.byref(norm), ctypes.byref(b), out_prefix.encode(),
                           ctypes.byref(pn), ctypes.byref(n))
    if rc != CPC_OK:
        raise EngineError(rc, "cpc_do_bining failed")
    return n.value, pn


def rp_thresholds(params):
    """Squared-separation edges thr[0..n_rp] of the rp bins (host libm), see include/cpc.h."""
    lib = load_library()
    out = np.zeros(params.n_rp + 1, np.float64)
    rc = lib.cpc_rp_thresholds(ctypes.byref(params), out.ctypes.data_as(_dp))
    if rc != CPC_OK:
        raise EngineError(rc, "cpc_rp_thresholds failed")
    return out


class Engine:
    """One device context.  Catalog handles are cached per Python Catalog object."""

    def __init__(self, device=0):
        self.lib = load_library()
        h = c_void_p()
        rc = self.lib.cpc_init(ctypes.byref(h), device)
        if rc != CPC_OK:
            raise EngineError(rc, f"cpc_init(device={device}) failed")
        self.ctx = h
        self._cats = {}

    def _check(self, rc, what):
        if rc != CPC_OK:
            msg = self.lib.cpc_last_error(self.ctx).decode()
            raise EngineError(rc, f"{what}: {msg}")

    def upload(self, cat: Catalog):
        key = id(cat)
        if key in self._cats:
            return self._cats[key][0]
        h = c_void_p()
        d = cat.desc()
        self._check(self.lib.cpc_catalog_upload(self.ctx, ctypes.byref(d), ctypes.byref(h)), "upload")
        self._cats[key] = (h, cat)
        return h

    def free(self, cat: Catalog):
        ent = self._cats.pop(id(cat), None)
        if ent is not None:
            self.lib.cpc_catalog_free(self.ctx, ent[0])

    def set_tuning(self, col_frac=0.0, item_size=0, block_threads=0):
        self._check(self.lib.cpc_set_tuning(self.ctx, col_frac, item_size, block_threads), "set_tuning")

    def count(self, outer: Catalog, inner: Catalog, params: Params, hist: Histogram = None):
        """One cross pass (the reference's corels2 after the role swap).  Returns (Histogram, stats dict)."""
        ho = self.upload(outer)
        hi = ho if inner is outer else self.upload(inner)
        if hist is None:
            hist = Histogram.for_params(params)
        hc = hist.c_struct()
        st = Stats()
        self._check(self.lib.cpc_count(self.ctx, ho, hi, ctypes.byref(params), ctypes.byref(hc), ctypes.byref(st)),
                    "count")
        return hist, st.as_dict()

    def count_noalloc(self, outer: Catalog, inner: Catalog, params: Params, hist: Histogram = None):
        """The same pass through cpc_count_noalloc: no device allocation during the call (catalogs are
        uploaded first), every device buffer from the library's static scratch array."""
        ho = self.upload(outer)
        hi = ho if inner is outer else self.upload(inner)
        if hist is None:
            hist = Histogram.for_params(params)
        hc = hist.c_struct()
        st = Stats()
        self._check(self.lib.cpc_count_noalloc(self.ctx, ho, hi, ctypes.byref(params), ctypes.byref(hc),
                                               ctypes.byref(st)), "count_noalloc")
        return hist, st.as_dict()

    def _jobs(self, jobs):
        n = len(jobs)
        arr = (JobC * max(n, 1))()
        keep = []
        hists = Histogram.many([params for _, _, params in jobs])
        for k, (outer, inner, params) in enumerate(jobs):
            ho = self.upload(outer)
            hi = ho if inner is outer else self.upload(inner)
            hist = hists[k]
            hc = hist.c_struct()
            st = Stats()
            keep.append((params, hist, hc, st))
            arr[k] = JobC(ho, hi, ctypes.pointer(params), ctypes.pointer(hc), ctypes.pointer(st))
        return n, arr, keep

    def count_batch(self, jobs):
        """Many passes in one call (cpc_count_batch): `jobs` is a list of (outer, inner, params).
        Returns ([(Histogram, stats dict)] in job order, total stats dict)."""
        n, arr, keep = self._jobs(jobs)
        tot = Stats()
        self._check(self.lib.cpc_count_batch(self.ctx, n, arr, ctypes.byref(tot)), "count_batch")
        return [(h, st.as_dict()) for (_, h, _, st) in keep], tot.as_dict()

    def autotune(self, jobs):
        """cpc_autotune over (outer, inner, params) jobs; the context keeps the chosen tuning.
        Returns (tuning dict, number of batches run, [(Histogram, stats dict)])."""
        n, arr, keep = self._jobs(jobs)
        best = TuningC()
        tried = c_int(0)
        self._check(self.lib.cpc_autotune(self.ctx, n, arr, ctypes.byref(best), ctypes.byref(tried)), "autotune")
        tun = {"col_frac": best.col_frac, "item_size": best.item_size, "block_threads": best.block_threads}
        return tun, tried.value, [(h, st.as_dict()) for (_, h, _, st) in keep]

    def fp32_error(self, s_lo, s_hi):
        """(log2 error, rsq error) of the device's FP32 approximations over every FP32 value in
        [s_lo, s_hi] (cpc_fp32_error)."""
        out = np.zeros(2, np.float64)
        self._check(self.lib.cpc_fp32_error(self.ctx, float(s_lo), float(s_hi), out.ctypes.data_as(_dp)),
                    "cpc_fp32_error")
        return float(out[0]), float(out[1])

    def libm_error(self, fn, lo, hi, n):
        """Max ulps between the sky kernels' device cos (fn "cos") or acos (fn "acos") and the host C
        library's on n points spread evenly over [lo, hi] (cpc_libm_error)."""
        out = np.zeros(1, np.float64)
        code = {"cos": 0, "acos": 1}[fn]
        self._check(self.lib.cpc_libm_error(self.ctx, code, float(lo), float(hi), int(n), out.ctypes.data_as(_dp)),
                    "cpc_libm_error")
        return float(out[0])

    def close(self):
        if self.ctx is not None:
            for h, _ in self._cats.values():
                self.lib.cpc_catalog_free(self.ctx, h)
            self._cats.clear()
            self.lib.cpc_destroy(self.ctx)
            self.ctx = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
