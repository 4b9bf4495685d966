"""TEST INFRASTRUCTURE ONLY: ctypes wrapper of the CPU oracle (oracle/cpc_oracle.c).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as the checker; the product
never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

from corr_pc_amd.engine import CPC_OK, CatalogDesc, EngineError, Histogram, HistC, Params, Stats

HERE = os.path.dirname(os.path.abspath(__file__))
# CPC_ORACLE_LIB selects another build of the same source (the sanitizer build of oracle/Makefile)
LIB_PATH = os.environ.get("CPC_ORACLE_LIB") or os.path.join(HERE, "libcpc_oracle.so")

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        lib = ctypes.CDLL(LIB_PATH)
        lib.orc_count.argtypes = [ctypes.POINTER(CatalogDesc), ctypes.POINTER(CatalogDesc), ctypes.POINTER(Params),
                                  ctypes.POINTER(HistC), ctypes.POINTER(Stats), ctypes.c_int]
        lib.orc_count.restype = ctypes.c_int
        lib.orc_rp_bin.argtypes = [ctypes.POINTER(Params), ctypes.c_double]
        lib.orc_rp_bin.restype = ctypes.c_double
        lib.orc_pi_bin.argtypes = [ctypes.POINTER(Params), ctypes.c_double]
        lib.orc_pi_bin.restype = ctypes.c_double
        _lib = lib
    return _lib


def count(outer, inner, params, n_threads=0):
    """The oracle's corels2 pass: returns (Histogram, stats dict)."""
    lib = load()
    hist = Histogram.for_params(params)
    hc = hist.c_struct()
    st = Stats()
    do, di = outer.desc(), inner.desc()
    rc = lib.orc_count(ctypes.byref(do), ctypes.byref(di), ctypes.byref(params), ctypes.byref(hc), ctypes.byref(st),
                       n_threads)
    if rc != CPC_OK:
        raise EngineError(rc, "orc_count failed")
    return hist, st.as_dict()


def rp_bin(params, s):
    return load().orc_rp_bin(ctypes.byref(params), float(s))


def pi_bin(params, d):
    return load().orc_pi_bin(ctypes.byref(params), float(d))
