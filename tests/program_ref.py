"""Test-side restatement of the corr_pc program around the pair passes, written independently of
corr_pc_amd/csrc/cpc_main.cpp so the two can be compared: the derived flags of data_ini
(initialization.cpp:162-198), the reader's per-galaxy fields and rand() draws in read order
(read_dat.cpp:216-325, do_corr.cpp:40-96), do_corr's pass list and RsRd_same merge (do_corr.cpp:59-154),
then do_bining through cpc_do_bining.  The pair counts come from the `count` callable (the oracle in
the tests)."""
import math

import numpy as np

from corr_pc_amd.engine import (GEOM_SKY, Catalog, Histogram, Norm, WSum, do_bining, make_params, rand_stream,
                                weight_sum)
from corr_pc_amd.passes import DataInfo, setup_pass
from corr_pc_amd.synth import elip_angle

BOX_SEED = 124693517  # seed(), jk.cpp:40-43
DEFAULT_SEED = 1      # rand() before any srand()


def derive(inp):
    """Flags data_ini derives (initialization.cpp:162-198)."""
    est = int(inp["estimator"])
    auto = est in (1, 2, 3)
    cross = est in (0, 3) or est not in (1, 2, 3)
    rebin = est in (2, 3)
    n_dr = int(inp["n_density_rand"])
    rsrd_same = n_dr == 0 and int(inp["which_corr"]) <= 2
    if rsrd_same:
        n_dr = int(inp["n_shape_rand"])
    ip_rr = 1.0
    if int(inp.get("rand_subsample", 0)):
        ip_rr = math.sqrt(math.sqrt(int(inp["n_density"]) / n_dr * int(inp["n_shape"]) / int(inp["n_shape_rand"])))
    return {"auto": auto, "cross": cross, "rebin": rebin, "rsrd_same": rsrd_same, "n_density_rand": n_dr,
            "include_prob_RR": ip_rr}


def read_order(inp, flags):
    """Catalog roles in the order do_corr reads them."""
    box = bool(int(inp["periodic_box"]))
    order = ["shape"]
    if not box or int(inp["n_jk_regions"]) > 0:
        order.append("shape_rand")
    if flags["cross"]:
        order.append("density")
        if not flags["rsrd_same"]:
            order.append("density_rand")
    return order


def build_catalogs(inp, cols, table=None):
    """Catalog objects with the reader's derived fields and the rand() stream drawn in read order.
    `cols[role]` holds the arrays written to that role's files (x, y, z and optional e1, e2, w, jk)."""
    flags = derive(inp)
    box = bool(int(inp["periodic_box"]))
    do_jk, n_jk = int(inp["do_jk"]), int(inp["n_jk_regions"])
    which = int(inp["which_corr"])
    counts = {"shape": int(inp["n_shape"]), "density": int(inp["n_density"]),
              "shape_rand": int(inp["n_shape_rand"]), "density_rand": flags["n_density_rand"]}
    order = read_order(inp, flags)
    coords = int(inp.get("coordinates", 6 if box else 0))
    phi_bins = not box and coords in (2, 4)
    # random catalogs draw phi with coordinates 2 and 4 (shape randoms always, density randoms for
    # which_corr 2) before include_prob (read_dat.cpp:127-131, 146-149, 305-318)
    rand_phi = {r: phi_bins and (r == "shape_rand" or (r == "density_rand" and which == 2)) for r in order}
    per_gal = {r: (1 if box else 0) + (1 if rand_phi[r] else 0) + (1 if r.endswith("rand") else 0) for r in order}
    total = sum(counts[r] * per_gal[r] for r in order)
    stream = rand_stream(total, seed=BOX_SEED if box else DEFAULT_SEED) if total else np.zeros(0)
    cats, k = {}, 0
    for role in order:
        n = counts[role]
        c = cols[role]
        draws = stream[k:k + n * per_gal[role]].reshape(n, per_gal[role]) if per_gal[role] else np.zeros((n, 0))
        k += n * per_gal[role]
        jk_prob = draws[:, 0].copy() if box else np.ones(n)
        if do_jk == 0 or n_jk == 0:
            jk_prob[:] = 0
        inc = draws[:, -1].copy() if role.endswith("rand") else None
        kap = (which == 3 and role == "shape") or (which == 4 and role == "density")
        axis = which == 5 and role == "shape"  # 3-D axes, scaled to unit length (read_dat.cpp:283-293)
        with_e = not kap and not axis and ((role == "shape" and (which != 0 or phi_bins)) or
                                           (role == "density" and which == 2))
        e1 = c["e1"][:n] if with_e else (c["kappa"][:n] if kap else None)
        e2 = c["e2"][:n] if with_e else (np.zeros(n) if kap else None)
        e3 = None
        if axis:
            a, b, e = c["e1"][:n], c["e2"][:n], c["e3"][:n]
            norm = np.sqrt(a * a + b * b + e * e)
            pos = norm > 0
            e1, e2, e3 = (np.where(pos, v / np.where(pos, norm, 1.0), v) for v in (a, b, e))
        kw = {}
        z = c["z"][:n] if coords not in (4, 5) else np.zeros(n)  # 4 and 5 read no redshift
        if not box:
            it = np.floor(z / table["dz"]).astype(np.int64)
            kw = {"dc": table["dC"][it], "da": table["dA"][it]}
        phi = elip_angle(e1, e2) if with_e else None
        if rand_phi[role]:
            phi = draws[:, (1 if box else 0)] * np.pi / 2 - np.pi / 2
        jk = c["jk"][:n].astype(np.int32) if (do_jk and c.get("jk") is not None) else np.full(n, -1, np.int32)
        cat = Catalog(x=c["x"][:n], y=c["y"][:n], z=z, w=None if c.get("w") is None else c["w"][:n],
                      e1=e1, e2=e2, phi=phi, jk=jk, jk_prob=jk_prob,
                      include_prob=inc, e3=e3, **kw)
        cats[role] = cat
    return cats, flags


def run(inp, cols, count, out_prefix, table=None, precision=0):
    """Every pass of do_corr counted by `count(outer, inner, params) -> (Histogram, stats)`, the
    RsRd_same merge and do_bining's files under `out_prefix`.  Returns (pass histograms, n_files)."""
    cats, flags = build_catalogs(inp, cols, table)
    box = bool(int(inp["periodic_box"]))
    n_jk = int(inp["n_jk_regions"])
    di = DataInfo(which_corr=int(inp["which_corr"]), estimator=int(inp["estimator"]),
                  box_size=float(inp["periodic_box_size"]), n_rp=int(inp["n_bins"]), rp_min=float(inp["bin_r_min"]),
                  rp_max=float(inp["bin_r_max"]), n_pi=int(inp["n_p_bin"]), pi_min=float(inp["p_min"]),
                  pi_max=float(inp["p_max"]), lin_bin=int(inp["lin_bin"]), do_jk=int(inp["do_jk"]), n_jk=n_jk,
                  include_prob_RR=flags["include_prob_RR"], r_mu=int(int(inp.get("coordinates", 6)) == 7),
                  coordinates=0 if box else int(inp.get("coordinates", 0)))
    empty = Catalog(x=np.zeros(0), y=np.zeros(0), z=np.zeros(0))
    S = cats.get("shape", empty)
    Rs = cats.get("shape_rand", empty)
    D = cats.get("density", empty)
    Rd = Rs if flags["rsrd_same"] else cats.get("density_rand", empty)
    passes = []  # (histogram name, do_* flag, g1, g2, auto)
    if flags["auto"]:
        passes.append(("SS", "SS", S, S, True))
    if (flags["auto"] or flags["rsrd_same"]) and not box:
        passes.append(("SRs", "SRs", Rs, S, False))
        passes.append(("RsRs", "RsRs", Rs, Rs, True) if not flags["rsrd_same"] else ("RsRd", "RsRs", Rs, Rs, False))
    if flags["cross"]:
        if not flags["rsrd_same"]:
            passes.append(("SRd", "SRd", Rd, S, False))
            passes.append(("RsRd", "RsRd", Rd, Rs, False))
        passes.append(("SD", "SD", D, S, False))
        if not box:
            passes.append(("DRs", "DRs", Rs, D, False))
    z_sep_max = None
    if not box:
        dz, z_max = table["dz"], float(inp["z_max"])
        rows = int(min(len(table["z"]), np.ceil((z_max + 1) / dz) + 1))
        # data_ini (initialization.cpp:241-281): from p_max (0) or rp_max (1); the file's value, 9
        # when 0, otherwise
        h_top = table["H"][int(np.ceil(z_max / dz))]
        z_sep_max = float(inp.get("z_sep_max", 0)) or 9.0
        z_sep_min = float(inp.get("z_sep_min", 0))
        if di.coordinates == 0:
            z_sep_max = float(inp["p_max"]) * h_top / 300000.0 + 0.001
        elif di.coordinates == 1:
            z_sep_max = float(inp["bin_r_max"]) * h_top / 300000.0 + 0.001
    hists = {}
    shape = None
    for name, flag, g1, g2, auto in passes:
        st = setup_pass(flag, g1, g2, di, precision=precision, auto=auto)
        if st is None:
            hists[name] = None
            continue
        outer, inner, p, _ = st
        # data_sorted 0, and 2 on the sky (Dec order: the z test ends no scan, the Dec test only
        # past the top of its two-sided window), scan like do_pair_unsorted in catalog order
        p.unsorted = int(int(inp.get("data_sorted", 1)) != 1)
        if not box:
            p.geometry = GEOM_SKY
            p.attach_table(table["H"][:rows], table["dA"][:rows], dz, z_max, z_sep_max)
            p.coordinates, p.z_min, p.z_sep_min = di.coordinates, float(inp.get("z_min", 0)), z_sep_min
            if di.coordinates >= 4 and int(inp.get("data_sorted", 1)) == 1 and not auto:
                # corels2's sorted start pointer skips inner galaxies (all at z = 0) outside the window
                zlo = z_sep_min if z_sep_min != 0 else float(inp.get("z_min", 0))
                if not zlo <= 0 <= z_sep_max:
                    hists[name] = None
                    continue
        h, _ = count(outer, inner, p)
        hists[name] = h
        shape = h.num.shape
    for k, v in list(hists.items()):
        if v is None and shape is not None:
            hists[k] = Histogram(shape[0], shape[1], shape[2])
    ws = {r: (weight_sum(c) if c.n else WSum(0.0, 0.0)) for r, c in cats.items()}
    zero = WSum(0.0, 0.0)
    S_w, Rs_w, D_w = ws.get("shape", zero), ws.get("shape_rand", zero), ws.get("density", zero)
    Rd_w = ws.get("density_rand", zero)
    bining = dict(hists)
    if flags["rsrd_same"] and not box:  # do_corr.cpp:135-154
        def add(a, b):
            out = Histogram(*shape)
            for h in (a, b):
                if h is None:
                    continue
                out.num += h.num
                out.wt_num += h.wt_num
                out.val += h.val
                out.err += h.err
                out.num_rp += h.num_rp
            return out
        bining["RsRs"] = add(hists.get("RsRs"), hists.get("RsRd"))
        bining["SRd"] = add(hists.get("SRd"), hists.get("SRs"))
        Rd_w = Rs_w
    norm = Norm(S_w, D_w, Rs_w, Rd_w, flags["include_prob_RR"])
    po = make_params(di.box_size, di.n_rp, di.rp_min, di.rp_max, di.n_pi, di.pi_min, di.pi_max,
                     estimator=di.estimator, which_corr=di.which_corr, lin_bin=di.lin_bin, do_jk=di.do_jk,
                     n_jk=di.n_jk, r_mu=di.r_mu)
    if not box:
        po.geometry = GEOM_SKY
        po.coordinates = di.coordinates
    n_files, pn = do_bining(po, norm, {k: v for k, v in bining.items() if v is not None}, out_prefix,
                            do_auto_corr=int(flags["auto"]), do_cross_corr=int(flags["cross"]),
                            do_auto_rebin=int(flags["rebin"]), rsrd_same=int(flags["rsrd_same"]))
    return hists, n_files, pn, {"S": S_w, "D": D_w, "Rs": Rs_w, "Rd": Rd_w}
