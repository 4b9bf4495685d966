"""Benchmark: binned pair evaluations per second of the split-precision pair kernel against the plain
FP64 kernel, on the 1M-galaxy wg+/w++ periodic box (L = 500 Mpc/h; 70% uniform + 30% in 5000
clumps; shapes and lognormal weights; 20 log rp bins in [0.1, 100); 40 pi bins in [-100, 100);
27 jackknife regions).  One step is one full DD pass of corr_pc's which_corr=2 run (the SD pass:
shape_shape_corel, 1M outer x 1M inner), from catalogs resident on the device to the finished
histogram in host memory: cell lists, the pair kernel, the device-side target reduction and the
assembly of the jackknife layout.

The line also carries the measured error of the weighted bins (split against FP64 on the full pass,
and against the CPU oracle on the cpu_baseline sample), and, under "secondary", the other workloads
at their real shapes: the 2e6-galaxy sky survey (split and FP64), the 256 batched mocks in one call
against sequential calls, and the 2e7-galaxy box on every 8th outer galaxy.  `--workload` runs one
of these alone, or a full-size run no default line includes: `large --large-full` (the large box's
whole SD or SRd pass) and `skyfull` (the sky survey's four passes with its 2e7 randoms).

Prints one JSON line on stdout; progress goes to stderr.
"""
import argparse
import hashlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

# algorithmic FP32-VALU flops per binned pair of the shape-shape kind (w++): distances and bins 12,
# rotation/projection of one shape 21 (wg+), second shape, products and error squares 20
FLOPS_PER_PAIR = {"DD": 12, "SD": 33, "SS": 53}
PEAK_TFLOPS = 157.3  # MI355X FP32 vector peak

TRAFFIC_SUMMARY = os.path.join(ROOT, "profiles", "pair_kernel_ss_split_box1m_traffic.md")
# the box kernel's sources and the host code that decomposes the pass into work items (their order
# changes which tile records the caches hold, hence the traffic)
KERNEL_SOURCES = [os.path.join(ROOT, "corr_pc_amd", "csrc", f) for f in ("cpc_kernels.hip", "cpc_device.h",
                                                                          "cpc_internal.h", "cpc_api.cpp")]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def kernel_source_sha():
    """First 16 hex digits of the SHA-256 of the box kernel's sources and of the host code that
    decomposes the pass, as the traffic summary records the sources it was measured on."""
    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(f, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic():
    """(HBM bytes per launch of the split kernel, note): from the committed rocprofv3 FETCH_SIZE and
    WRITE_SIZE summary of this workload, only when it was measured on the kernel sources built now."""
    try:
        vals = {}
        with open(TRAFFIC_SUMMARY) as f:
            for line in f:
                if ":" in line and not line.startswith(("|", "#", " ", "`")):
                    k, v = line.split(":", 1)
                    vals[k.strip()] = v.strip()
    except OSError:
        return None, "no traffic summary under profiles/"
    if vals.get("kernel_source_sha") != kernel_source_sha():
        return None, (f"the traffic summary was measured on kernel sources {vals.get('kernel_source_sha')}, "
                      f"not the built {kernel_source_sha()}")
    try:
        return float(vals["traffic_bytes_per_launch"]), "measured on these kernel sources"
    except (KeyError, ValueError):
        return None, "traffic summary without traffic_bytes_per_launch"


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def mocks_workload(args, steps=1, warmup=0):
    """The batched-mocks workload: independent wgg boxes (L = 250, data sizes log-uniform in
    [1e4, 5e5], clustered, 2x randoms, 2x2x2 jackknife; 10 log rp bins in [0.1, 30), 20 pi bins in
    [-60, 60)), each with its DD, DR and RR passes (SD, SRd, RsRd), all in one cpc_count_batch call
    with the autotuned decomposition, against the same passes as sequential cpc_count calls."""
    import numpy as np

    from corr_pc_amd import synth
    from corr_pc_amd.engine import PREC_SPLIT, Engine
    from corr_pc_amd.passes import config_datainfo, setup_pass

    t = time.perf_counter()
    sizes = synth.mock_sizes(args.mocks)
    di = config_datainfo(synth.MOCK_BINNING)
    jobs, tune_jobs = [], []
    for k, n in enumerate(sizes):
        c = synth.mock_catalogs(k, int(n))
        mine = []
        for name, g1, g2 in (("SD", c["D"], c["S"]), ("SRd", c["Rd"], c["S"]), ("RsRd", c["Rd"], c["Rs"])):
            outer, inner, p, _ = setup_pass(name, g1, g2, di, precision=PREC_SPLIT)
            mine.append((outer, inner, p))
        jobs += mine
        if k % args.tune_stride == 0:
            tune_jobs += mine
    log(f"mocks: {len(sizes)} boxes, {int(sizes.sum())} data galaxies, {len(jobs)} passes ready in "
        f"{time.perf_counter() - t:.1f}s")
    eng = Engine(0)
    for o, i, _ in jobs:
        eng.upload(o)
        eng.upload(i)
    t0 = time.perf_counter()
    tun, tried, _ = eng.autotune(tune_jobs)
    log(f"mocks autotune: {tun} after {tried} batches of {len(tune_jobs)} passes in {time.perf_counter() - t0:.1f}s")
    for k in range(warmup):
        _, tot = eng.count_batch(jobs)
    t0 = time.perf_counter()
    for k in range(steps):
        res, tot = eng.count_batch(jobs)
        log(f"mocks batch step {k}: {tot['ms_total']:.0f} ms (pair kernels {tot['ms_kernel']:.0f} ms, cell lists "
            f"{tot['ms_cells']:.0f} ms on the device), binned {tot['n_binned']}")
    dt = (time.perf_counter() - t0) / steps
    n_binned = tot["n_binned"]
    # the sequential baseline: each pass alone with the automatic decomposition (the batch's tuning
    # is chosen for the batch; applied to one small pass it can leave most of the GPU idle)
    eng.set_tuning(0.0, 0, 0)
    t0 = time.perf_counter()
    seq_dev = 0.0
    identical = True
    for (o, i, p), (h, _) in zip(jobs, res):
        hs, st = eng.count(o, i, p)
        seq_dev += st["ms_kernel"] + st["ms_cells"]
        identical = identical and bool(np.array_equal(hs.num, h.num))
    dt_seq = time.perf_counter() - t0
    log(f"mocks sequential: {dt_seq * 1e3:.0f} ms (device {seq_dev:.0f} ms), counts identical: {identical}")
    eng.close()
    return {
        "workload": f"{len(sizes)} wgg boxes L=250, data log-uniform in [1e4, 5e5] (clustered), 2x randoms, 8 "
                    f"jackknife regions, DD+DR+RR passes, 10 log rp bins [0.1,30) x 20 pi bins [-60,60)",
        "value": n_binned / dt, "unit": "pair evals/s", "ms_per_step": dt * 1e3, "n_binned": n_binned,
        "n_passes": len(jobs), "tuning": tun, "batch_device_ms": tot["ms_kernel"] + tot["ms_cells"],
        "batch_kernel_ms": tot["ms_kernel"], "batch_cells_ms": tot["ms_cells"],
        "batch_host_ms": dt * 1e3 - tot["ms_kernel"] - tot["ms_cells"],
        "sequential": {"value": n_binned / dt_seq, "ms": dt_seq * 1e3, "device_ms": seq_dev,
                       "counts_identical": identical, "tuning": "automatic per pass"},
        "batch_vs_sequential": dt_seq / dt,
    }


def wgg_workload(args, steps=1, warmup=1):
    """The density kernel at the headline box's shape: the SD pass of the 1M-galaxy box (L = 500, 20
    log rp bins in [0.1, 100), 40 pi bins in [-100, 100), 27 jackknife regions) as which_corr = 0
    counts it (density pairs, wgg), split precision and plain FP64."""
    import dataclasses

    import numpy as np

    from corr_pc_amd import synth
    from corr_pc_amd.engine import PREC_FP64, PREC_SPLIT, Engine
    from corr_pc_amd.passes import config_datainfo, setup_pass

    cats = synth.catalogs(1)
    di = dataclasses.replace(config_datainfo(cats["config"]), which_corr=0)
    eng = Engine(0)
    res = {}
    for prec, label in ((PREC_SPLIT, "split"), (PREC_FP64, "fp64")):
        outer, inner, p, _ = setup_pass("SD", cats["D"], cats["S"], di, precision=prec)
        eng.upload(outer)
        eng.upload(inner)
        for k in range(warmup):
            eng.count(outer, inner, p)
        t0 = time.perf_counter()
        for k in range(steps):
            h, st = eng.count(outer, inner, p)
            log(f"wgg {label} step {k}: {st['ms_total']:.0f} ms (kernel {st['ms_primary']:.0f} ms), binned "
                f"{st['n_binned']}")
        res[label] = ((time.perf_counter() - t0) / steps, st, h)
    eng.close()
    dt, st, h = res["split"]
    dt64, st64, h64 = res["fp64"]
    achieved = FLOPS_PER_PAIR["DD"] * st["n_binned"] / (st["ms_primary"] * 1e-3) / 1e12
    return {
        "workload": "wgg periodic box L=500, 1M galaxies (30% in 5000 clumps) x 1M, SD pass as density pairs, "
                    "20 log rp bins [0.1,100) x 40 pi bins [-100,100), 27 jackknife regions",
        "value": st["n_binned"] / dt, "unit": "pair evals/s", "ms_per_step": dt * 1e3, "n_binned": st["n_binned"],
        "n_candidates": st["n_candidates"], "kernel_ms": st["ms_primary"], "frac_fp32_valu": achieved / PEAK_TFLOPS,
        "fp64": {"value": st64["n_binned"] / dt64, "ms_per_step": dt64 * 1e3, "kernel_ms": st64["ms_primary"],
                 "counts_identical": bool(np.array_equal(h.num, h64.num))},
        "split_vs_fp64": dt64 / dt,
    }


def large_workload(args, steps=1, warmup=0):
    """The large box at its real shape: L = 1000, 2e7 galaxies (30% in 40,000 clumps) with shapes and
    weights, 4e7 randoms, 25 log rp bins in [0.1, 200), 60 pi bins in [-150, 150), 64 jackknife
    regions; the SD pass (or with `--large-pass SRd` the randoms against the shapes) of every
    `--large-stride`-th outer galaxy against all inner galaxies.  With `--large-full` every residue
    class of the stride runs in turn (one call each), which together are the whole pass."""
    import numpy as np

    from corr_pc_amd import synth
    from corr_pc_amd.engine import PREC_SPLIT, Engine
    from corr_pc_amd.passes import config_datainfo, setup_pass

    t = time.perf_counter()
    srd = args.large_pass == "SRd"
    cats = synth.catalogs(4, n_rand=None if srd else 10)  # the SD pass reads no randoms
    di = config_datainfo(cats["config"])
    if srd:  # the density randoms against the shape sample (do_corr.cpp:125)
        outer0, inner, p, _ = setup_pass("SRd", cats["Rd"], cats["S"], di, precision=PREC_SPLIT)
    else:
        outer0, inner, p, _ = setup_pass("SD", cats["D"], cats["S"], di, precision=PREC_SPLIT)
    offsets = range(args.large_stride) if args.large_full else range(1)
    parts = [outer0.subset(np.arange(o, outer0.n, args.large_stride)) for o in offsets]
    log(f"large: catalogs ready in {time.perf_counter() - t:.1f}s: outer {sum(q.n for q in parts)} of {outer0.n} "
        f"in {len(parts)} call(s), inner {inner.n}")
    eng = Engine(0)
    eng.upload(inner)
    for q in parts:
        eng.upload(q)
    for k in range(warmup):
        eng.count(parts[0], inner, p)
    n_binned = n_cand = n_re = 0
    ms_kernel = 0.0
    t0 = time.perf_counter()
    for k in range(steps):
        n_binned = n_cand = n_re = 0
        ms_kernel = 0.0
        for i, q in enumerate(parts):
            _, st = eng.count(q, inner, p)
            n_binned += st["n_binned"]
            n_cand += st["n_candidates"]
            n_re += st["n_fp64_reeval"]
            ms_kernel += st["ms_primary"]
            log(f"large step {k} part {i}: {st['ms_total']:.0f} ms (kernel {st['ms_primary']:.0f} ms, build "
                f"{st['ms_build']:.0f} ms), binned {st['n_binned']}")
    dt = (time.perf_counter() - t0) / steps
    eng.close()
    achieved = FLOPS_PER_PAIR["SD" if srd else "SS"] * n_binned / (ms_kernel * 1e-3) / 1e12
    outer_n = sum(q.n for q in parts)
    return {
        "workload": f"wg+/w++ periodic box L=1000, 2e7 galaxies (40,000 clumps), 4e7 randoms, {args.large_pass} pass, "
                    f"{outer_n} of {outer0.n} outer galaxies x all {inner.n} inner, 25 log rp bins [0.1,200) x 60 pi "
                    f"bins [-150,150), 64 jackknife regions",
        "value": n_binned / dt, "unit": "pair evals/s", "ms_per_step": dt * 1e3, "n_binned": n_binned,
        "n_candidates": n_cand, "n_fp64_reeval": n_re, "n_calls": len(parts),
        "kernel_ms": ms_kernel, "frac_fp32_valu": achieved / PEAK_TFLOPS,
    }


def sky_workload(args, steps=1, warmup=1):
    """The sky survey at its real shape: 2e6 galaxies in six RA/Dec fields (about 150 deg^2),
    0.2 < z < 1 with n(z) ~ z^2 exp(-(z/0.4)^1.5), flat LCDM distances, 100 jackknife regions, 20 log
    rp bins in [0.1, 100), 40 pi bins in [-100, 100); the SD pass of which_corr = 1 (shape_density,
    wg+), 2e6 x 2e6, split precision and plain FP64."""
    import numpy as np

    from corr_pc_amd import synth
    from corr_pc_amd.engine import KIND_SHAPE_DENSITY, PREC_FP64, PREC_SPLIT, Catalog, Engine, make_sky_params

    t = time.perf_counter()
    table = synth.distance_table(0.3, 1e-4, 1.0)
    S, n_jk, _ = synth.sky_catalog(2, "data", table=table)
    D = Catalog(x=S.x, y=S.y, z=S.z, w=S.w, jk=S.jk, jk_prob=S.jk_prob, dc=S.dc, da=S.da)
    pt = synth.sky_params_table(2, table)
    params = {prec: make_sky_params(pt, 20, 0.1, 100.0, 40, -100.0, 100.0, kind=KIND_SHAPE_DENSITY, g_with_shape=2,
                                    which_corr=1, do_jk=1, n_jk=n_jk, precision=prec)
              for prec in (PREC_SPLIT, PREC_FP64)}
    log(f"sky: catalogs ready in {time.perf_counter() - t:.1f}s: {S.n} galaxies, {n_jk} jackknife regions")
    eng = Engine(0)
    eng.upload(D)
    eng.upload(S)
    res = {}
    for prec, label in ((PREC_SPLIT, "split"), (PREC_FP64, "fp64")):
        for k in range(warmup):
            eng.count(D, S, params[prec])
        t0 = time.perf_counter()
        for k in range(steps):
            h, st = eng.count(D, S, params[prec])
            log(f"sky {label} step {k}: {st['ms_total']:.0f} ms (pair kernel {st['ms_primary']:.0f} ms, reduction "
                f"{st['ms_reduce']:.1f} ms, build {st['ms_build']:.0f} ms), binned {st['n_binned']}, host-binned "
                f"{st['n_host_pairs']}")
        res[label] = ((time.perf_counter() - t0) / steps, st, h)
    eng.close()
    dt, st, h = res["split"]
    dt64, st64, h64 = res["fp64"]
    return {
        "workload": f"sky survey, 2e6 galaxies in 6 RA/Dec fields (~150 deg^2), 0.2<z<1, SD pass (shape_density) "
                    f"2e6 x 2e6, 20 log rp bins [0.1,100) x 40 pi bins [-100,100), {n_jk} jackknife regions",
        "value": st["n_binned"] / dt, "unit": "pair evals/s", "ms_per_step": dt * 1e3, "n_binned": st["n_binned"],
        "n_candidates": st["n_candidates"], "n_host_pairs": st["n_host_pairs"], "ms_primary": st["ms_primary"],
        "ms_reduce": st["ms_reduce"],
        "fp64": {"value": st64["n_binned"] / dt64, "ms_per_step": dt64 * 1e3,
                 "counts_identical": bool(np.array_equal(h.num, h64.num))},
        "split_vs_fp64": dt64 / dt,
    }


def sky_passes_workload(args, steps=1, warmup=0):
    """The sky survey's whole pass set at its real shape: 2e6 galaxies and 2e7 randoms in the six
    fields, which_corr = 1 (wg+) with one random catalog for both roles, as do_corr runs it
    (do_corr.cpp:66-78, 110-131 with RsRd_same): SD = corels2(D, S) (shape_density, 2e6 x 2e6),
    SRs = corels2(Rs, S) (randoms against the shapes, 2e7 x 2e6), DRs = corels2(Rs, D) (density
    pairs, 2e7 x 2e6) and RsRd = corels2(Rs, Rs) (density pairs, 2e7 x 2e7).  Each pass runs in
    `--sky-stride` calls, one per residue class of its outer galaxies, in split precision and, with
    `--sky-fp64`, in plain FP64 as well (counts compared call by call)."""
    import numpy as np

    from corr_pc_amd import synth
    from corr_pc_amd.engine import PREC_FP64, PREC_SPLIT, Catalog, Engine, make_sky_params
    from corr_pc_amd.passes import DataInfo, setup_pass

    t = time.perf_counter()
    table = synth.distance_table(0.3, 1e-4, 1.0)
    S, n_jk, _ = synth.sky_catalog(2, "data", table=table)
    R, n_jk_r, _ = synth.sky_catalog(2, "rand", table=table)
    if n_jk_r != n_jk:
        raise RuntimeError(f"the randoms span {n_jk_r} jackknife regions, the data {n_jk}")
    D = Catalog(x=S.x, y=S.y, z=S.z, w=S.w, jk=S.jk, jk_prob=S.jk_prob, dc=S.dc, da=S.da)
    pt = synth.sky_params_table(2, table)
    # the role logic of corels2 (larger catalog outer, g_with_shape flip, pair kind per do_* flag);
    # box_size is not read by it
    di = DataInfo(which_corr=1, estimator=0, box_size=1.0, n_rp=20, rp_min=0.1, rp_max=100.0, n_pi=40,
                  pi_min=-100.0, pi_max=100.0, do_jk=1, n_jk=n_jk)
    passes = [("SD", D, S), ("SRs", R, S), ("DRs", R, D), ("RsRd", R, R)]
    log(f"skyfull: catalogs ready in {time.perf_counter() - t:.1f}s: {S.n} galaxies, {R.n} randoms, {n_jk} "
        f"jackknife regions")
    eng = Engine(0)
    for c in (D, S, R):
        eng.upload(c)
    precs = ((PREC_SPLIT, "split"), (PREC_FP64, "fp64")) if args.sky_fp64 else ((PREC_SPLIT, "split"),)
    out = {}
    for name, g1, g2 in passes:
        outer, inner, pb, _ = setup_pass(name, g1, g2, di)
        prm = {prec: make_sky_params(pt, di.n_rp, di.rp_min, di.rp_max, di.n_pi, di.pi_min, di.pi_max, kind=pb.kind,
                                     g_with_shape=pb.g_with_shape, which_corr=1, do_jk=1, n_jk=n_jk,
                                     include_prob_max=pb.include_prob_max, precision=prec)
               for prec, _ in precs}
        stride = max(1, args.sky_stride)
        parts = [outer if stride == 1 else outer.subset(np.arange(o, outer.n, stride)) for o in range(stride)]
        for q in parts:
            if q is not outer:
                eng.upload(q)
        rec = {"outer": outer.n, "inner": inner.n, "kind": int(pb.kind), "g_with_shape": int(pb.g_with_shape),
               "n_calls": len(parts)}
        for k in range(warmup):
            eng.count(parts[0], inner, prm[PREC_SPLIT])
        for prec, label in precs:
            n_binned = n_cand = n_host = 0
            ms_kernel = 0.0
            hists = []
            t0 = time.perf_counter()
            for i, q in enumerate(parts):
                h, st = eng.count(q, inner, prm[prec])
                hists.append(h.num.copy())
                n_binned += st["n_binned"]
                n_cand += st["n_candidates"]
                n_host += st["n_host_pairs"]
                ms_kernel += st["ms_primary"]
                log(f"skyfull {name} {label} part {i}: {st['ms_total']:.0f} ms (kernel {st['ms_primary']:.0f} ms, "
                    f"build {st['ms_build']:.0f} ms), binned {st['n_binned']}, host-binned {st['n_host_pairs']}")
            dt = time.perf_counter() - t0
            rec[label] = {"ms": dt * 1e3, "kernel_ms": ms_kernel, "n_binned": n_binned, "n_candidates": n_cand,
                          "n_host_pairs": n_host, "value": n_binned / dt}
            if label == "split":
                split_hists = hists
            else:
                rec["counts_identical"] = all(np.array_equal(a, b) for a, b in zip(split_hists, hists))
                rec["split_vs_fp64"] = rec["fp64"]["ms"] / rec["split"]["ms"]
        out[name] = rec
        for q in parts:
            if q is not outer:
                eng.free(q)
    eng.close()
    n_all = sum(r["split"]["n_binned"] for r in out.values())
    ms_all = sum(r["split"]["ms"] for r in out.values())
    return {
        "workload": f"sky survey, 2e6 galaxies and 2e7 randoms in 6 RA/Dec fields (~150 deg^2), 0.2<z<1, which_corr 1: "
                    f"passes SD, SRs, DRs, RsRd (DD, RD, DR, RR), 20 log rp bins [0.1,100) x 40 pi bins [-100,100), "
                    f"{n_jk} jackknife regions",
        "value": n_all / (ms_all * 1e-3), "unit": "pair evals/s", "ms_per_step": ms_all, "n_binned": n_all,
        "passes": out,
    }


def standalone(name, fn, args):
    """--workload mocks|large|sky|wgg|skyfull: one workload as its own JSON line."""
    r = fn(args, steps=args.steps, warmup=args.warmup)
    out = {"metric": f"pair evals/sec: {name}", "value": r["value"], "unit": "pair evals/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": r["ms_per_step"], "higher_is_better": True,
           "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
           "config": {"workload": r["workload"], "parallelism": "single GPU"}, "detail": r}
    print(json.dumps(out), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cpu-stride", type=int, default=40,
                    help="CPU baseline: every n-th outer galaxy of the same pass")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-secondary", action="store_true", help="skip the sky, mocks and large-box workloads")
    ap.add_argument("--workload", choices=("wgplus", "mocks", "large", "sky", "wgg", "skyfull"), default="wgplus",
                    help="wgplus: the 1M-galaxy wg+/w++ box (default, with the others as secondary); mocks, "
                         "large, sky, wgg: one of the secondary workloads alone; skyfull: the sky survey's four "
                         "passes with its 2e7 randoms")
    ap.add_argument("--sky-stride", type=int, default=1,
                    help="skyfull workload: calls per pass, one per residue class of the outer galaxies")
    ap.add_argument("--sky-fp64", action="store_true", help="skyfull workload: also run every call in plain FP64")
    ap.add_argument("--large-stride", type=int, default=8, help="large workload: every n-th outer galaxy")
    ap.add_argument("--large-full", action="store_true",
                    help="large workload: every residue class of the stride in turn, i.e. the whole pass")
    ap.add_argument("--large-pass", choices=("SD", "SRd"), default="SD",
                    help="large workload: the SD pass (default) or the SRd pass (4e7 randoms x 2e7 shapes)")
    ap.add_argument("--mocks", type=int, default=256, help="number of mocks of the mocks workload")
    ap.add_argument("--tune-stride", type=int, default=16, help="mocks workload: autotune on every n-th mock")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB mebibytes (to scale rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n * 4, "launches": 3}), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}): more GPUs would each take a share "
              f"of the work items and sum their histograms in one reduction, which is not launched here.",
              file=sys.stderr)
        sys.exit(2)

    if args.workload == "mocks":
        standalone("batched wgg mocks in one call vs sequential calls", mocks_workload, args)
        return
    if args.workload == "large":
        standalone("split-FP32 kernel, large wg+/w++ box at real shape", large_workload, args)
        return
    if args.workload == "sky":
        standalone("split-FP32 vs FP64 sky kernel, wg+ 2M-galaxy survey", sky_workload, args)
        return
    if args.workload == "skyfull":
        standalone("split-FP32 sky kernel, the 2M-galaxy survey's SD, SRs, DRs and RsRd passes", sky_passes_workload,
                   args)
        return
    if args.workload == "wgg":
        standalone("split-FP32 vs FP64 density kernel, wgg 1M-gal box", wgg_workload, args)
        return

    import numpy as np

    from corr_pc_amd import synth
    from corr_pc_amd.compare import field_errors, kind_fields
    from corr_pc_amd.engine import PREC_FP64, PREC_SPLIT, Engine
    from corr_pc_amd.passes import config_datainfo, setup_pass

    t = time.perf_counter()
    cats = synth.catalogs(1)
    cfg = cats["config"]
    di = config_datainfo(cfg)
    outer, inner, p_split, _ = setup_pass("SD", cats["D"], cats["S"], di, precision=PREC_SPLIT)
    _, _, p_fp64, _ = setup_pass("SD", cats["D"], cats["S"], di, precision=PREC_FP64)
    fields = kind_fields(p_split.kind, p_split.estimator)
    log(f"catalogs ready in {time.perf_counter() - t:.1f}s: outer {outer.n}, inner {inner.n}")

    eng = Engine(0)
    eng.upload(outer)
    eng.upload(inner)

    def run(p, steps, label):
        stats = []
        h = None
        t0 = time.perf_counter()
        for k in range(steps):
            h, st = eng.count(outer, inner, p)
            stats.append(st)
            log(f"{label} step {k}: {st['ms_total']:.1f} ms (kernel {st['ms_primary']:.1f} ms, build "
                f"{st['ms_build']:.1f} ms), binned {st['n_binned']}")
        return time.perf_counter() - t0, stats, h

    run(p_split, args.warmup, "warmup split")
    dt, st_split, h_split = run(p_split, args.steps, "split")
    ms_step = dt / args.steps * 1e3
    n_binned = st_split[-1]["n_binned"]
    value = n_binned / (dt / args.steps)
    ms_kernel = float(np.mean([s["ms_primary"] for s in st_split]))

    fp64 = None
    errors = {}
    if not args.no_fp64:
        run(p_fp64, 1, "warmup fp64")
        dt64, st64, h64 = run(p_fp64, args.steps, "fp64")
        v64 = st64[-1]["n_binned"] / (dt64 / args.steps)
        fp64 = {"value": v64, "ms_per_step": dt64 / args.steps * 1e3,
                "ms_kernel": float(np.mean([s["ms_primary"] for s in st64])),
                "counts_identical": bool(np.array_equal(h64.num, h_split.num))}
        # the stated error of every weighted bin (wgg wt_num; wg+ and w++ val and err sums) of the full
        # pass: split against the FP64 kernel, per field, max over regions and cells
        errors["split_vs_fp64_full_pass"] = field_errors(h_split, h64, fields)

    # the same SD pass as which_corr = 1 runs it (shape_density_corel: wg+ alone), one step per precision
    wg_only = None
    if not args.no_fp64:
        import dataclasses
        di1 = dataclasses.replace(di, which_corr=1)
        _, _, q_split, _ = setup_pass("SD", cats["D"], cats["S"], di1, precision=PREC_SPLIT)
        _, _, q_fp64, _ = setup_pass("SD", cats["D"], cats["S"], di1, precision=PREC_FP64)
        run(q_split, 1, "warmup wg+ split")
        ds, sts, hs = run(q_split, 1, "wg+ split")
        run(q_fp64, 1, "warmup wg+ fp64")
        d64, st64b, h64b = run(q_fp64, 1, "wg+ fp64")
        wg_only = {"kind": "shape_density (which_corr 1)", "split_ms_per_step": ds * 1e3,
                   "fp64_ms_per_step": d64 * 1e3, "split_ms_kernel": sts[-1]["ms_primary"],
                   "fp64_ms_kernel": st64b[-1]["ms_primary"], "split_vs_fp64": d64 / ds,
                   "n_binned": sts[-1]["n_binned"], "counts_identical": bool(np.array_equal(hs.num, h64b.num))}

    traffic, traffic_note = measured_traffic()
    achieved = FLOPS_PER_PAIR["SS"] * n_binned / (ms_kernel * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": achieved, "peak": PEAK_TFLOPS, "unit": "TFLOP/s",
                "frac": achieved / PEAK_TFLOPS, "traffic": traffic, "traffic_note": traffic_note,
                "kernel": "pair_kernel<shape_shape, split>", "flops_per_pair": FLOPS_PER_PAIR["SS"],
                "ms_per_launch": ms_kernel, "kernel_source_sha": kernel_source_sha()}

    cpu = None
    if not args.no_cpu:
        from oracle import oracle
        threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
        sub = outer.subset(np.arange(0, outer.n, args.cpu_stride))
        t0 = time.perf_counter()
        href, cst = oracle.count(sub, inner, p_split, n_threads=threads)
        cdt = time.perf_counter() - t0
        cpu = {"value": cst["n_binned"] / cdt, "unit": "pair evals/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(),
               "sample": f"same SD pass, every {args.cpu_stride}th outer galaxy ({sub.n} of {outer.n}) against all "
                         f"{inner.n} inner galaxies: {cst['n_binned']} binned pairs in {cdt:.1f} s"}
        log(f"cpu baseline: {cpu['value']:.3e} pairs/s on {threads} threads")
        # the split kernel on the same sample against the oracle (corr_pc's CPU arithmetic)
        hsub, _ = eng.count(sub, inner, p_split)
        errors["split_vs_oracle_cpu_sample"] = field_errors(hsub, href, fields)
    eng.close()

    secondary = None
    if not args.no_secondary:
        secondary = {}
        # one untimed call first where the workload's own timing is short (allocations of the first call)
        for name, fn, kw in (("sky", sky_workload, {}), ("mocks", mocks_workload, {"warmup": 1}),
                             ("large", large_workload, {})):
            t0 = time.perf_counter()
            secondary[name] = fn(args, **kw)
            log(f"secondary {name}: {time.perf_counter() - t0:.1f} s")

    out = {
        "metric": "pair evals/sec: split-FP32 vs FP64 GPU kernel, wg+ 1M-gal box, bit-exact bins",
        "value": value,
        "unit": "pair evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "wg+/w++ periodic box L=500, 1M galaxies (30% in 5000 clumps) x 1M, SD pass "
                               "(shape_shape, estimator 0), 20 log rp bins [0.1,100) x 40 pi bins [-100,100), "
                               "27 jackknife regions",
                   "global_batch": outer.n, "parallelism": "single GPU"},
        "n_binned": n_binned,
        "n_candidates": st_split[-1]["n_candidates"],
        "n_fp64_reeval": st_split[-1]["n_fp64_reeval"],
        "fp64": fp64,
        "split_vs_fp64": (value / fp64["value"]) if fp64 else None,
        "wg_plus_only": wg_only,
        "weighted_error": errors or None,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
