"""Host mirror of the pass set-up of corels2 (corr.cpp:219-296) and corels1 (corr.cpp:103-111) and of
choose_corel_func (initialization.cpp:490-543): which catalog is outer, which carries the shape, which
pair kind runs.

corels2(g1, g2, bjk, data_inf) takes its catalog sizes from the do_* role flags, makes the larger
catalog the outer one (ties keep g1), and flips g_with_shape when it swaps.  `setup_pass` returns
the engine-level (outer, inner, params) of such a call, ready for Engine.count.
"""
from dataclasses import dataclass

from .engine import (JK_LEAVE_ONE_OUT, KIND_DD, KIND_DENSITY_CONVERGENCE, KIND_ED, KIND_SHAPE_CONVERGENCE,
                     KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, PREC_SPLIT, make_params)

PASSES = ("SD", "SRd", "RsRd", "SRs", "DRs", "RsRs", "SS")


@dataclass
class DataInfo:
    """The data_info fields a pass needs (data_def.h:46-96); box geometry."""
    which_corr: int
    estimator: int
    box_size: float
    n_rp: int
    rp_min: float
    rp_max: float
    n_pi: int
    pi_min: float
    pi_max: float
    lin_bin: int = 0
    do_jk: int = 1
    n_jk: int = 0
    include_prob_RR: float = 1.0
    r_mu: int = 0  # coordinates 7: r-mu bins
    coordinates: int = 0  # the sky's coordinates 0-5 (0 in the box)

    @property
    def phi_bins(self):
        """Coordinates 2 and 4: pi is the ephi angle of the g_with_shape galaxy (pi_calc_phi)."""
        return self.coordinates in (2, 4)


def pass_kind(which_corr, pass_name, coordinates=0):
    """choose_corel_func, initialization.cpp:490-543; pass_name is the do_* flag."""
    if pass_name in ("RsRs", "RsRd"):
        return KIND_DD
    if which_corr == 0:
        return KIND_DD
    if which_corr == 1:
        return KIND_SHAPE_SHAPE if pass_name == "SS" else KIND_SHAPE_DENSITY
    if which_corr == 2:
        if coordinates in (2, 4):
            return KIND_SHAPE_SHAPE if pass_name in ("DRs", "SD") else KIND_DD
        return KIND_SHAPE_SHAPE if pass_name in ("SD", "SS") else KIND_SHAPE_DENSITY
    if which_corr == 3:
        return KIND_DENSITY_CONVERGENCE
    if which_corr == 4:
        return KIND_SHAPE_CONVERGENCE
    if which_corr == 5:
        return KIND_ED
    raise ValueError(f"which_corr={which_corr} is outside the pair kinds this engine runs")


def setup_pass(pass_name, g1, g2, di: DataInfo, precision=PREC_SPLIT, jk_mode=JK_LEAVE_ONE_OUT, auto=False):
    """(outer, inner, params, order_changed) of corels2(g1, g2, ...) for pass `pass_name` (the do_*
    flag set around the call), or of corels1(g1, ...) when `auto`; None when a catalog is empty.

    g1, g2 are the arguments in the order do_corr passes them (e.g. SD: corels2(gdensity, gshape),
    do_corr.cpp:125; SRd: corels2(grand_density, gshape), :110; RsRd: corels2(grand_density,
    grand_shape), :115)."""
    if pass_name not in PASSES:
        raise ValueError(pass_name)
    kind = pass_kind(di.which_corr, pass_name, di.coordinates)
    inc_max = di.include_prob_RR if pass_name in ("RsRs", "RsRd") else 1.0  # initialization.cpp:597-599

    def params(kind, gws):
        return make_params(di.box_size, di.n_rp, di.rp_min, di.rp_max, di.n_pi, di.pi_min, di.pi_max, kind=kind,
                           g_with_shape=gws, estimator=di.estimator, which_corr=di.which_corr,
                           lin_bin=di.lin_bin, do_jk=di.do_jk, n_jk=di.n_jk, include_prob_max=inc_max,
                           precision=precision, jk_mode=jk_mode, r_mu=di.r_mu)

    if auto:  # corels1: one catalog, g_with_shape = 1 (corr.cpp:111)
        if g1.n == 0:
            return None
        p = params(kind, 1)
        p.auto_pairs = 1
        return g1, g1, p, False
    # operand roles, corr.cpp:226-271: g_with_shape before the swap; coordinates 2 and 4 give the
    # random passes one for pi_calc_phi
    gws = 2 if pass_name in ("SRd", "SD", "SRs") else 0
    if di.phi_bins:
        gws = {"DRs": 1, "RsRd": 2, "RsRs": 1}.get(pass_name, gws)
    n1, n2 = g1.n, g2.n
    if n1 == 0 or n2 == 0:
        return None  # corels2 returns 0 pairs (corr.cpp:273-277)
    outer, inner, swapped = g1, g2, False
    if n2 > n1:  # corr.cpp:281-296
        outer, inner, swapped = g2, g1, True
        gws = {2: 1, 1: 2}.get(gws, gws)
    if kind == KIND_SHAPE_DENSITY and gws == 0:
        # shape_density_corel rotates g1's shape unless g_with_shape is 2 (corels.cpp:216-219); with no
        # shapes on g1 the reference reads an unset pointer, and such columns feed no estimator
        gws = 1
        if outer.e1 is None:
            kind = KIND_DD
    if (kind == KIND_SHAPE_DENSITY and (inner if gws == 2 else outer).e1 is None) or \
            (kind == KIND_SHAPE_SHAPE and (outer.e1 is None or inner.e1 is None)):
        kind = KIND_DD  # a shape kind on a catalog without shapes: a reference read of an unset pointer
    if kind == KIND_ED and (inner if gws == 2 else outer).e3 is None:
        kind = KIND_DD  # ED_calc_PB reads the axis of g2 (g_with_shape 2) or g1: none there
    if kind in (KIND_DENSITY_CONVERGENCE, KIND_SHAPE_CONVERGENCE):
        if gws == 0:
            return outer, inner, params(kind, 0), swapped  # assigns nothing: zero values (corels.cpp:79-120)
        # kappa (e1) from the g_with_shape side (density-convergence) or the other side (shape-
        # convergence); a catalog without it is a reference read of an unset pointer: density pairs
        ws, ot = (inner, outer) if gws == 2 else (outer, inner)
        kap = ws if kind == KIND_DENSITY_CONVERGENCE else ot
        if kap.e1 is None or (kind == KIND_SHAPE_CONVERGENCE and (ws.e1 is None or getattr(ws, "kappa", False))):
            kind = KIND_DD
    return outer, inner, params(kind, gws if gws else 2), swapped


def config_datainfo(cfg, do_jk=1):
    """DataInfo of a synthetic box configuration (synth.CONFIGS)."""
    return DataInfo(which_corr=cfg.which_corr, estimator=0, box_size=cfg.box, n_rp=cfg.n_rp, rp_min=cfg.rp_min,
                    rp_max=cfg.rp_max, n_pi=cfg.n_pi, pi_min=-cfg.pi_max, pi_max=cfg.pi_max, do_jk=do_jk,
                    n_jk=cfg.jk_side ** 3 if do_jk else 0)
