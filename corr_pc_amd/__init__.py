"""MI355X-native rp-pi pair-counting engine: a drop-in for corr_pc's cross-pair loop (corels2).

The compute path is libcpc.so (hand-written HIP kernels for gfx950 behind the C ABI in
include/cpc.h); this package holds its ctypes bindings and the host-side mirror of corr_pc's pass
set-up.
"""
from .engine import (Catalog, Engine, EngineError, Histogram, Params, load_library, make_params, rand_stream,
                     rp_thresholds)
from .engine import (GEOM_BOX, JK_LEAVE_ONE_OUT, JK_TOUCH, KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE,
                     PREC_FP64, PREC_SPLIT)

__all__ = ["Catalog", "Engine", "EngineError", "Histogram", "Params", "load_library", "make_params",
           "rand_stream", "rp_thresholds", "GEOM_BOX", "JK_LEAVE_ONE_OUT", "JK_TOUCH", "KIND_DD",
           "KIND_SHAPE_DENSITY", "KIND_SHAPE_SHAPE", "PREC_FP64", "PREC_SPLIT"]
