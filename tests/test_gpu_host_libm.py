"""The arithmetic of the machine that runs the kernels, on which bin assignment rests.

* The host C library of the machine that drives the GPU: its log10 decides the rp thresholds
  (corels.cpp:34) and its rand() the jackknife coins (jk.cpp:40-43, read_dat.cpp:243); both are
  compared with the committed fixtures (the same checks as test_host.py, run where the kernels run).
* The device's FP32 approximations that the split kernels classify with: v_log_f32 (log rp bins) and
  v_rsq_f32 (linear bins and the guard band), swept over every FP32 squared separation the split
  classification can see for each workload binning, against FP64.  The guard bands take both errors
  as at most 2^-21 (relative to max(1, |log2 s|) for the logarithm); the sweep proves it for this
  device.
"""
import math

import numpy as np
import pytest

import test_host as host_checks
from corr_pc_amd.engine import make_params, rp_thresholds

pytestmark = pytest.mark.gpu

BOUND = 2.0 ** -21


def test_rand_stream_on_gpu_host():
    host_checks.test_rand_stream_matches_glibc_generator()


@pytest.mark.parametrize("name", ["config0", "config1", "config4", "linear"])
def test_rp_thresholds_on_gpu_host(oracle_lib, name):
    host_checks.test_rp_thresholds(oracle_lib, name)


# (n_rp, rp_min, rp_max, lin_bin): the box workloads' binnings (configs 0 and 3; 1 and 2; 4) and a
# linear binning from 0
BINNINGS = {"config0": (10, 0.1, 30.0, 0), "config1": (20, 0.1, 100.0, 0), "config4": (25, 0.1, 200.0, 0),
            "linear": (7, 0.0, 9.0, 1)}


@pytest.mark.parametrize("name", sorted(BINNINGS))
def test_fp32_approximation_errors(engine, name):
    n_rp, rmin, rmax, lin = BINNINGS[name]
    p = make_params(1000.0, n_rp, rmin, rmax, 10, -10.0, 10.0, lin_bin=lin)
    thr = rp_thresholds(p)
    # every FP32 value around the binned range of s = rp^2 (a factor 2 beyond either end)
    s_lo = max(float(thr[0]) / 2, 1e-6)
    s_hi = float(thr[-1]) * 2
    e_log, e_rsq = engine.fp32_error(s_lo, s_hi)
    print(f"[fp32] {name}: s in [{s_lo:.3g}, {s_hi:.3g}]: log2 error {e_log:.3e} "
          f"(x2^21 = {e_log / BOUND:.3f}), rsq error {e_rsq:.3e} (x2^21 = {e_rsq / BOUND:.3f})", flush=True)
    assert e_log <= BOUND, f"v_log_f32 error {e_log} exceeds the guard band's 2^-21"
    assert e_rsq <= BOUND, f"v_rsq_f32 error {e_rsq} exceeds the guard band's 2^-21"
    # the sweep does see the approximations' rounding (a zero would mean it compared nothing)
    assert e_log > 0 and e_rsq > 0
    assert math.isfinite(e_log) and math.isfinite(e_rsq)


def test_fp32_error_rejects_bad_range(engine):
    from corr_pc_amd.engine import EngineError
    with pytest.raises(EngineError):
        engine.fp32_error(0.0, 1.0)
    with pytest.raises(EngineError):
        engine.fp32_error(2.0, 1.0)
    assert np.isfinite(engine.fp32_error(1.0, 1.0)[0])


# The sky kernels' device cos and acos against the host C library.  The sky guard band bounds the
# device/host difference of the great-circle sum sin d1 sin d2 + cos d1 cos d2 cos(a1 - a2) from a
# cosine within 2 ulps of the host's (sky_calcs.cpp:24-31), and takes the arc cosine within 1e-15
# relative (about 4.5 ulps).  Ranges: RA differences of one field up to the full circle, and the
# sums of separations up to a few degrees, where acos is steepest.
@pytest.mark.parametrize("lo,hi", [(0.0, 0.05), (0.0, 1.6), (-7.0, 7.0)])
def test_sky_cos_within_bound(engine, lo, hi):
    ulp = engine.libm_error("cos", lo, hi, 1 << 22)
    print(f"[libm] cos on [{lo}, {hi}]: {ulp:.3f} ulp", flush=True)
    assert ulp <= 2.0


@pytest.mark.parametrize("lo,hi", [(0.99, 1.0), (-1.0, 1.0)])
def test_sky_acos_within_bound(engine, lo, hi):
    ulp = engine.libm_error("acos", lo, hi, 1 << 22)
    print(f"[libm] acos on [{lo}, {hi}]: {ulp:.3f} ulp", flush=True)
    assert ulp <= 4.0


def test_libm_error_rejects_bad_arguments(engine):
    from corr_pc_amd.engine import EngineError
    with pytest.raises(EngineError):
        engine.libm_error("acos", 0.5, 1.5, 16)
    with pytest.raises(EngineError):
        engine.libm_error("cos", 1.0, 0.0, 16)
