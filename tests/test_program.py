"""CPU tests of the corr_pc program (corr_pc_amd/corr_pc, corr_pc_amd/csrc/cpc_main.cpp) that need no
GPU: `--plan` parses the .inp file, reads the catalogs and prints what it read and which passes it
would run, then stops before the engine starts.

Pinned by:
  * the rand() stream (glibc, seeded 124693517 in a box, default-seeded on the sky): each catalog's
    first jk_prob / include_prob is the draw its position in do_corr's read order gives
    (do_corr.cpp:40-96, read_dat.cpp:243, 318-320), computed here from the catalog sizes alone;
  * the weight sums against a sequential sum of the written weights;
  * the distance-table step lookup DC = dC_z[floor(z/dz)] and z_sep_max = p_max H(z_max)/3e5 + 0.001
    against the table written;
  * the pass list, operand roles, swap and pair kinds against the reference's rules as restated in
    corr_pc_amd/passes.py and tests/program_ref.py (do_corr.cpp:59-131, corr.cpp:226-296);
  * the error paths: unsupported modes, a truncated .inp, unsorted catalogs, a missing file.
"""
import os
import subprocess

import numpy as np
import pytest

from corr_pc_amd import inputs, synth
from corr_pc_amd.engine import KIND_DD, KIND_SHAPE_DENSITY, KIND_SHAPE_SHAPE, rand_stream, weight_sum
from corr_pc_amd.passes import DataInfo, setup_pass
import program_ref

# CPC_PROG selects another build of the program (the sanitizer build of corr_pc_amd/build.py)
PROG = os.environ.get("CPC_PROG") or os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                                                  "corr_pc_amd", "corr_pc")
KINDS = {"dd": KIND_DD, "shape_density": KIND_SHAPE_DENSITY, "shape_shape": KIND_SHAPE_SHAPE}


def run_prog(inp_path, *args):
    return subprocess.run([PROG, inp_path] + list(args), capture_output=True, text=True, timeout=120)


def parse_plan(out):
    cats, passes, info = {}, [], {}
    for line in out.splitlines():
        tok = line.split()
        if not tok:
            continue
        kv = dict(t.split("=", 1) for t in tok[1:] if "=" in t)
        if tok[0] == "catalog":
            cats[tok[1]] = kv
        elif tok[0] == "pass":
            kv["name"] = tok[1]
            passes.append(kv)
        elif tok[0] in ("weights", "flags"):
            info[tok[0]] = kv
    return cats, passes, info


def kappa_role(which, role):
    """The catalogs whose e file is one column of convergence (read_dat.cpp:36-46, 97-99)."""
    return (which == 3 and role == "shape") or (which == 4 and role == "density")


def box_case(tmp_path, which=2, estimator=0, n_s=1500, n_d=1200, n_rs=2500, n_rd=2000, do_jk=1, jk_side=2, seed=3,
             sorted_z=True, L=60.0, **extra):
    """A small periodic box in the reference's file format: shape and density samples with shapes,
    weights and jk files, shape and density randoms."""
    rng = np.random.default_rng(seed)
    cols, files = {}, {}
    for role, n, with_e, with_w in (("shape", n_s, True, True), ("density", n_d, True, True),
                                    ("shape_rand", n_rs, False, False), ("density_rand", n_rd, False, False)):
        pos = rng.uniform(0, L, (n, 3))
        if sorted_z:
            pos = pos[np.argsort(pos[:, 2], kind="stable")]
        jk = synth.jk_grid(pos[:, 0], pos[:, 1], pos[:, 2], L, jk_side)
        c = {"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2], "jk": jk,
             "w": rng.lognormal(0, 0.3, n) if with_w else None}
        if with_e and kappa_role(which, role):
            c["kappa"] = rng.normal(0, 0.05, n)
        elif with_e and which == 5 and role == "shape":
            c["e1"], c["e2"], c["e3"] = rng.normal(0, 1, n), rng.normal(0, 1, n), rng.normal(0, 1, n)
        elif with_e:
            c["e1"], c["e2"] = rng.normal(0, 0.25, n), rng.normal(0, 0.25, n)
        cols[role] = c
        files[role] = inputs.write_catalog(str(tmp_path / role), c["x"], c["y"], c["z"], c.get("e1"), c.get("e2"),
                                           c["w"], jk, kappa=c.get("kappa"), e3=c.get("e3"))
    inp = dict(which_corr=which, coordinates=6, estimator=estimator, do_jk=do_jk, out_file=str(tmp_path / "out_"),
               n_jk_regions=jk_side ** 3, bin_r_min=0.5, bin_r_max=12.0, n_bins=6, lin_bin=0, n_p_bin=8, p_min=-16.0,
               p_max=16.0, z_min=0, z_max=L, periodic_box=1, periodic_box_size=L)
    for role, n in (("shape", n_s), ("density", n_d), ("shape_rand", n_rs), ("density_rand", n_rd)):
        inp.update(inputs.catalog_params(role, files[role], n))
    inp.update(extra)
    path = inputs.write_inp(str(tmp_path / "run.inp"), **inp)
    return path, inp, cols


def sky_case(tmp_path, which=1, estimator=0, n_s=1200, n_d=1000, n_rs=2400, n_rd=0, seed=5, sorted_z=True,
             sort_dec=False, **extra):
    """A small survey field: RA 10-14 deg, Dec -2..2 deg, z in (0.2, 0.5), with a flat LCDM table."""
    rng = np.random.default_rng(seed)
    table = synth.distance_table(0.3, 1e-4, 0.5)
    cols, files = {}, {}
    for role, n, with_e, with_w in (("shape", n_s, True, True), ("density", n_d, which in (2, 4), True),
                                    ("shape_rand", n_rs, False, False), ("density_rand", n_rd, False, False)):
        ra = rng.uniform(10.0, 14.0, n)
        dec = rng.uniform(-2.0, 2.0, n)
        z = rng.uniform(0.2, 0.5, n)
        if sorted_z:
            z = np.sort(z)
        if sort_dec:
            o = np.argsort(dec, kind="stable")
            ra, dec, z = ra[o], dec[o], z[o]
        jk = (np.minimum(((ra - 10.0) / 2.0).astype(int), 1) * 2 + np.minimum(((dec + 2.0) / 2.0).astype(int), 1))
        c = {"x": ra, "y": dec, "z": z, "jk": jk, "w": rng.lognormal(0, 0.3, n) if with_w else None}
        if with_e and kappa_role(which, role):
            c["kappa"] = rng.normal(0, 0.05, n)
        elif with_e:
            c["e1"], c["e2"] = rng.normal(0, 0.25, n), rng.normal(0, 0.25, n)
        cols[role] = c
        if n:
            files[role] = inputs.write_catalog(str(tmp_path / role), ra, dec, z, c.get("e1"), c.get("e2"), c["w"], jk,
                                               kappa=c.get("kappa"))
        else:
            files[role] = ["0"] * 5
    dist = inputs.write_distance_table(str(tmp_path / "dist.dat"), table)
    inp = dict(which_corr=which, coordinates=0, estimator=estimator, do_jk=1, out_file=str(tmp_path / "out_"),
               n_jk_regions=4, bin_r_min=0.5, bin_r_max=15.0, n_bins=5, lin_bin=0, n_p_bin=6, p_min=-30.0,
               p_max=30.0, z_min=0, z_max=0.5, dz=0.0001, periodic_box=0, periodic_box_size=0, distance_file=dist)
    for role, n in (("shape", n_s), ("density", n_d), ("shape_rand", n_rs), ("density_rand", n_rd)):
        inp.update(inputs.catalog_params(role, files[role], n))
    inp.update(extra)
    path = inputs.write_inp(str(tmp_path / "run.inp"), **inp)
    return path, inp, cols, table


def test_inp_roundtrip_fields(tmp_path):
    """write_inp puts one "<label> <value>" per parameter in data_ini's order: 52 lines."""
    path = inputs.write_inp(str(tmp_path / "a.inp"), which_corr=1, periodic_box_size=75)
    lines = open(path).read().splitlines()
    assert len(lines) == len(inputs.INP_FIELDS) == 52
    assert lines[0].split() == ["which_corr", "1"] and lines[-1].split() == ["box_size", "75"]


@pytest.mark.parametrize("which", [0, 1, 2])
def test_box_plan(tmp_path, which):
    path, inp, cols = box_case(tmp_path, which=which)
    r = run_prog(path, "--plan")
    assert r.returncode == 0, r.stdout + r.stderr
    cats, passes, info = parse_plan(r.stdout)
    nS, nD, nRs, nRd = 1500, 1200, 2500, 2000
    s = rand_stream(nS + 2 * nRs + nD + 2 * nRd, seed=program_ref.BOX_SEED)
    # read order S, Rs, D, Rd; one jk_prob per galaxy, then include_prob for randoms
    assert float(cats["S"]["jk_prob0"]) == s[0]
    assert float(cats["Rs"]["jk_prob0"]) == s[nS] and float(cats["Rs"]["include_prob0"]) == s[nS + 1]
    assert float(cats["D"]["jk_prob0"]) == s[nS + 2 * nRs]
    assert float(cats["Rd"]["jk_prob0"]) == s[nS + 2 * nRs + nD]
    assert float(cats["Rd"]["include_prob0"]) == s[nS + 2 * nRs + nD + 1]
    assert cats["S"]["has_e"] == str(int(which != 0)) and cats["D"]["has_e"] == str(int(which == 2))
    assert cats["Rs"]["has_e"] == "0" and cats["Rd"]["has_e"] == "0"
    for k, role in (("S", "shape"), ("D", "density")):
        acc = acc2 = 0.0
        for x in cols[role]["w"]:
            acc += x
            acc2 += x * x
        assert tuple(float(v) for v in info["weights"][k].split(",")) == (acc, acc2)
    assert tuple(float(v) for v in info["weights"]["Rs"].split(",")) == (nRs, nRs)
    # pass list of do_corr for estimator 0 in a box with density randoms: SRd, RsRd, SD
    assert [p["name"] for p in passes] == ["SRd", "RsRd", "SD"]
    pcats, _ = program_ref.build_catalogs(inp, cols)
    di = DataInfo(which_corr=which, estimator=0, box_size=60.0, n_rp=6, rp_min=0.5, rp_max=12.0, n_pi=8,
                  pi_min=-16.0, pi_max=16.0, do_jk=1, n_jk=8)
    args = {"SRd": (pcats["density_rand"], pcats["shape"]), "RsRd": (pcats["density_rand"], pcats["shape_rand"]),
            "SD": (pcats["density"], pcats["shape"])}
    names = {id(pcats["shape"]): "S", id(pcats["density"]): "D", id(pcats["shape_rand"]): "Rs",
             id(pcats["density_rand"]): "Rd"}
    for p in passes:
        outer, inner, pp, swapped = setup_pass(p["flag"], *args[p["name"]], di)
        assert (p["outer"], p["inner"]) == (names[id(outer)], names[id(inner)]), p
        assert KINDS[p["kind"]] == pp.kind and int(p["gws"]) == pp.g_with_shape, p
        assert int(p["swapped"]) == int(swapped)
    if which:
        phi0 = synth.elip_angle(cols["shape"]["e1"][:1], cols["shape"]["e2"][:1])[0]
        np.testing.assert_allclose(float(cats["S"]["phi0"]), phi0, rtol=1e-15)
    else:
        assert float(cats["S"]["phi0"]) == -10.0  # no shapes read (read_dat.cpp:269)
    info_lines = open(inp["out_file"] + "inp.info").read().splitlines()
    assert "n_jk                 8" in info_lines and "RsRd_same            0" in info_lines
    assert "do_cross_corr        1" in info_lines and "S_wt                 0" in info_lines


def test_box_plan_without_jk_skips_shape_randoms(tmp_path):
    """do_jk = 0 and n_jk = 0 in a box: the shape randoms are not read (do_corr.cpp:45), so the density
    sample's jk_prob is the draw right after the shape sample's."""
    path, inp, cols = box_case(tmp_path, which=0, do_jk=0, jk_side=1)
    inputs.write_inp(path, **dict(inp, n_jk_regions=0))
    r = run_prog(path, "--plan")
    assert r.returncode == 0, r.stdout
    cats, passes, _ = parse_plan(r.stdout)
    assert cats["Rs"]["read"] == "0"
    assert float(cats["D"]["jk_prob0"]) == 0.0  # drawn, then zeroed without jackknife
    # RsRd runs on unread shape randoms: empty
    assert [(p["name"], p["empty"]) for p in passes] == [("SRd", "0"), ("RsRd", "1"), ("SD", "0")]


@pytest.mark.parametrize("estimator", [0, 2])
def test_sky_plan(tmp_path, estimator):
    path, inp, cols, table = sky_case(tmp_path, which=1, estimator=estimator)
    r = run_prog(path, "--plan")
    assert r.returncode == 0, r.stdout + r.stderr
    cats, passes, info = parse_plan(r.stdout)
    nS, nRs = 1200, 2400
    # sky: no jk_prob draws (jk_prob = 1); randoms draw include_prob from the default-seeded stream
    s = rand_stream(nRs, seed=program_ref.DEFAULT_SEED)
    assert float(cats["S"]["jk_prob0"]) == 1.0 and float(cats["Rs"]["include_prob0"]) == s[0]
    it = int(np.floor(cols["shape"]["z"][0] / 1e-4))
    assert float(cats["S"]["dc0"]) == table["dC"][it] and float(cats["S"]["da0"]) == table["dA"][it]
    flags = info["flags"]
    z_sep = 30.0 * table["H"][int(np.ceil(0.5 / 1e-4))] / 300000.0 + 0.001
    assert float(flags["z_sep_max"]) == z_sep
    assert int(flags["table_rows"]) == int(np.ceil(1.5 / 1e-4)) + 1
    assert flags["rsrd_same"] == "1"  # no density randoms
    if estimator == 0:
        # RsRd_same on the sky: SRs and the ordered Rs x Rs pass into RsRd, then SD and DRs
        assert [(p["name"], p["flag"], p["auto"]) for p in passes] == [
            ("SRs", "SRs", "0"), ("RsRd", "RsRs", "0"), ("SD", "SD", "0"), ("DRs", "DRs", "0")]
        drs = passes[-1]
        assert drs["kind"] == "dd"  # which_corr 1: the shape-less outer side (see setup_pass)
    else:
        assert flags["auto"] == "1" and flags["rebin"] == "1" and flags["cross"] == "0"
        assert [(p["name"], p["auto"]) for p in passes] == [("SS", "1"), ("SRs", "0"), ("RsRd", "0")]
        assert passes[0]["kind"] == "shape_shape"
        assert cats["D"]["read"] == "0"


@pytest.mark.parametrize("coordinates", [2, 4])
def test_sky_plan_angle_bins(tmp_path, coordinates):
    """Coordinates 2 and 4 (read_dat.cpp:27-28, 91-92, 127-131, 146-149; corr.cpp:226-249;
    initialization.cpp:510-520): random catalogs draw phi before include_prob, coordinates 4 reads
    no redshift (table row 0), the random passes get a g_with_shape, and which_corr 2 runs
    shape_shape for SD and DRs only (DRs falls back to density pairs: the shape randoms carry no
    shapes) and density pairs otherwise."""
    path, inp, cols, table = sky_case(tmp_path, which=2, estimator=0, n_rd=900, coordinates=coordinates,
                                      p_min=0.0, p_max=1.6, n_p_bin=4)
    r = run_prog(path, "--plan")
    assert r.returncode == 0, r.stdout + r.stderr
    cats, passes, info = parse_plan(r.stdout)
    nRs, nRd = 2400, 900
    s = rand_stream(2 * nRs + 2 * nRd, seed=program_ref.DEFAULT_SEED)
    # read order S, Rs, D, Rd: (phi, include_prob) per random galaxy
    assert float(cats["Rs"]["phi0"]) == s[0] * np.pi / 2 - np.pi / 2
    assert float(cats["Rs"]["include_prob0"]) == s[1]
    assert float(cats["Rd"]["phi0"]) == s[2 * nRs] * np.pi / 2 - np.pi / 2
    assert float(cats["Rd"]["include_prob0"]) == s[2 * nRs + 1]
    assert cats["S"]["has_e"] == "1" and cats["D"]["has_e"] == "1" and cats["Rs"]["has_e"] == "0"
    if coordinates == 4:
        assert float(cats["S"]["dc0"]) == table["dC"][0] and float(cats["S"]["da0"]) == table["dA"][0]
    else:
        it = int(np.floor(cols["shape"]["z"][0] / 1e-4))
        assert float(cats["S"]["dc0"]) == table["dC"][it]
    by = {p["name"]: p for p in passes}
    assert by["SD"]["kind"] == "shape_shape"
    assert by["DRs"]["kind"] == "dd" and by["SRd"]["kind"] == "dd" and by["RsRd"]["kind"] == "dd"
    # DRs: corels2(grand_shape, gdensity) with g_with_shape 1 (the outer Rs, 2400 > 1000 galaxies)
    assert by["DRs"]["outer"] == "Rs" and by["DRs"]["gws"] == "1"
    # RsRd: corels2(grand_density, grand_shape), g_with_shape 2, swapped (Rs is larger): 1
    assert by["RsRd"]["swapped"] == "1" and by["RsRd"]["gws"] == "1"


def test_errors(tmp_path):
    path, inp, cols = box_case(tmp_path, which=0)
    bad = inputs.write_inp(str(tmp_path / "w7.inp"), **dict(inp, which_corr=7))
    r = run_prog(bad, "--plan")
    assert r.returncode == 1 and "which_corr 7 is not supported" in r.stdout
    lines = open(path).read().splitlines()
    short = tmp_path / "short.inp"
    short.write_text("\n".join(lines[:-3]) + "\n")
    r = run_prog(str(short), "--plan")
    assert r.returncode == 1 and "could not read every parameter" in r.stdout
    missing = inputs.write_inp(str(tmp_path / "m.inp"), **dict(inp, shape_pos_file=str(tmp_path / "nope.dat")))
    r = run_prog(missing, "--plan")
    assert r.returncode == 1 and "pos File not open" in r.stdout
    r = subprocess.run([PROG], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "no input file given" in r.stdout


def test_unsorted_catalog_rejected(tmp_path):
    path, inp, cols = box_case(tmp_path, which=0, sorted_z=False)
    r = run_prog(path, "--plan")
    assert r.returncode == 1 and "not sorted by redshift" in r.stdout


def test_unsorted_input(tmp_path):
    """data_sorted 0: every pass takes catalogs in any order, the auto estimators included (their
    corels1 scans start at j = i+1 under every image, corr.cpp:167-182)."""
    path, inp, cols = box_case(tmp_path, which=0, sorted_z=False, data_sorted=0)
    r = run_prog(path, "--plan")
    assert r.returncode == 0, r.stdout
    inputs.write_inp(path, **dict(inp, estimator=1))
    r = run_prog(path, "--plan")
    assert r.returncode == 0, r.stdout
    cats, passes, info = parse_plan(r.stdout)
    assert [(p["name"], p["auto"]) for p in passes][0] == ("SS", "1")


def test_rand_subsample_include_prob(tmp_path):
    """rand_subsample 1: include_prob_RR = (nD/nRd * nS/nRs)^(1/4) (initialization.cpp:195-196) caps the
    randoms of the RsRd pass and no other."""
    path, inp, cols, table = sky_case(tmp_path, which=0, n_rd=1800, rand_subsample=1)
    r = run_prog(path, "--plan")
    assert r.returncode == 0, r.stdout
    cats, passes, info = parse_plan(r.stdout)
    expect = ((1000 / 1800 * 1200 / 2400) ** 0.5) ** 0.5
    assert float(info["flags"]["include_prob_RR"]) == pytest.approx(expect, rel=1e-15)
    for p in passes:
        assert float(p["include_prob_max"]) == (float(info["flags"]["include_prob_RR"]) if p["flag"] == "RsRd" else 1.0)


def test_box_without_randoms_plan(tmp_path):
    """The generator notebook's box run (no randoms): RsRd_same, the shape randoms read as empty, and
    only the SD pass counted; the analytic RR stands in for the randoms."""
    path, inp, cols = box_case(tmp_path, which=2, n_shape_rand=0, n_density_rand=0)
    r = run_prog(path, "--plan")
    assert r.returncode == 0, r.stdout
    cats, passes, info = parse_plan(r.stdout)
    assert info["flags"]["rsrd_same"] == "1" and cats["Rs"]["n"] == "0" and cats["Rd"]["read"] == "0"
    assert [p["name"] for p in passes] == ["SD"]
    s = rand_stream(1500 + 1200, seed=program_ref.BOX_SEED)
    assert float(cats["D"]["jk_prob0"]) == s[1500]  # no draws for the empty shape randoms


@pytest.mark.parametrize("which", [3, 4])
def test_convergence_plan(tmp_path, which):
    """which_corr 3: the shape sample's e file is one column of kappa and every non-random pass is
    density-convergence; which_corr 4: the density sample carries kappa, the shape sample e, and the
    passes whose kappa side is a random catalog run as density pairs."""
    path, inp, cols = box_case(tmp_path, which=which)
    r = run_prog(path, "--plan")
    assert r.returncode == 0, r.stdout
    cats, passes, _ = parse_plan(r.stdout)
    assert cats["S"]["has_e"] == "1" and cats["D"]["has_e"] == str(int(which == 4))
    assert float(cats["S"]["phi0"]) == (-10.0 if which == 3 else float(cats["S"]["phi0"]))
    kinds = {p["name"]: p["kind"] for p in passes}
    if which == 3:
        assert kinds == {"SRd": "density_convergence", "RsRd": "dd", "SD": "density_convergence"}
    else:
        assert kinds == {"SRd": "dd", "RsRd": "dd", "SD": "shape_convergence"}


def test_more_error_paths(tmp_path):
    """The reference's own input checks, each with its message and exit status 1:
    a redshift above z_max (read_dat.cpp:234), bad bin limits (initialization.cpp:299-303), dz = 0 on
    the sky (:249-253), a distance table shorter than z_max / dz (:48-52), data_sorted 2 in a box
    (:158-160)."""
    spath, sinp, scols, table = sky_case(tmp_path, which=0)
    # a galaxy above z_max: z_max below the catalogs' largest redshift
    p = inputs.write_inp(str(tmp_path / "zmax.inp"), **dict(sinp, z_max=0.3))
    r = run_prog(p, "--plan")
    assert r.returncode == 1 and "wrong z at data" in r.stdout, r.stdout
    p = inputs.write_inp(str(tmp_path / "dz.inp"), **dict(sinp, dz=0))
    r = run_prog(p, "--plan")
    assert r.returncode == 1 and "dz=0" in r.stdout, r.stdout
    short = tmp_path / "short_dist.dat"
    rows = open(sinp["distance_file"]).read().splitlines()
    need = int(np.ceil(sinp["z_max"] / sinp["dz"])) + 1  # rows da_Z must read (initialization.cpp:48)
    assert len(rows) >= need
    short.write_text("\n".join(rows[: need - 1000]) + "\n")
    p = inputs.write_inp(str(tmp_path / "table.inp"), **dict(sinp, distance_file=str(short)))
    r = run_prog(p, "--plan")
    assert r.returncode == 1 and "ended unexpectedly" in r.stdout, r.stdout
    bpath, binp, bcols = box_case(tmp_path, which=0)
    for bad in (dict(bin_r_min=5.0, bin_r_max=2.0), dict(bin_r_max=0)):
        p = inputs.write_inp(str(tmp_path / "bins.inp"), **dict(binp, **bad))
        r = run_prog(p, "--plan")
        assert r.returncode == 1 and "Bad Bin limits" in r.stdout, (bad, r.stdout)
    p = inputs.write_inp(str(tmp_path / "ds2.inp"), **dict(binp, data_sorted=2))
    r = run_prog(p, "--plan")
    assert r.returncode == 1 and "data_sorted==2 doesnot work with periodic box" in r.stdout, r.stdout
    # sig_crit weighting (corels.cpp:52-61) is refused rather than silently left out
    p = inputs.write_inp(str(tmp_path / "sc.inp"), **dict(binp, sig_crit=1))
    r = run_prog(p, "--plan")
    assert r.returncode == 1 and "sig_crit 1 is not supported" in r.stdout, r.stdout
