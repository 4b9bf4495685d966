"""Per-field error statistics between two histograms of one pass (reference layout
[region][rp][pi], cpc_hist): the stated error of the weighted bins.

For every accumulated field (wt_num; data[f].val and data[f].err for the kind's fields, i.e. the
sums of corels.cpp:48-244 and jk.cpp:150-170) over every region and cell:
  max_rel      max |got - ref| / |ref| over the bins with ref != 0
  max_rel_mag  max |got - ref| / M, M the bin's accumulated magnitude: |ref| for sums of non-negative
               terms (wt_num, err) and |ref| + sqrt(num * err) of the full sample for the val sums
               (Cauchy-Schwarz: sqrt(num * sum q^2) >= sum |q|), the denominator of the 1e-10 tolerance
Integer counts are compared exactly (counts_identical).
"""
import numpy as np

# data[] fields each pair kind fills (corels.cpp:48-262), by the cpc_params kind and estimator
_FIELDS = {0: (), 1: (1, 2, 6), 3: (1,), 4: (1, 3, 5), 5: (6,)}


def kind_fields(kind, estimator=0):
    if kind == 2:
        return (1, 3, 4, 5, 6) if estimator == 0 else (1, 2, 3, 4, 5, 6)
    return _FIELDS[kind]


def _stats(d, ref, mag):
    nz = ref != 0
    rel = float(np.max(d[nz] / np.abs(ref[nz]))) if np.any(nz) else 0.0
    relm = float(np.max(d / np.maximum(mag, 1e-300))) if d.size else 0.0
    return {"max_rel": rel, "max_rel_mag": relm}


def field_errors(got, ref, fields):
    """Error statistics of `got` against `ref` (Histogram objects or dicts with num, wt_num, val, err)."""
    g = got if isinstance(got, dict) else got.__dict__
    r = ref if isinstance(ref, dict) else ref.__dict__
    out = {"counts_identical": bool(np.array_equal(g["num"], r["num"]))}
    rw = r["wt_num"]
    out["wt_num"] = _stats(np.abs(g["wt_num"] - rw), rw, np.abs(rw))
    num0 = r["num"][0].astype(np.float64)
    for f in fields:
        rv, re_ = r["val"][..., f], r["err"][..., f]
        mag_v = np.abs(rv) + np.sqrt(num0 * np.abs(re_[0]))[None]
        out[f"val{f}"] = _stats(np.abs(g["val"][..., f] - rv), rv, mag_v)
        if np.any(re_ != 0):
            out[f"err{f}"] = _stats(np.abs(g["err"][..., f] - re_), re_, np.abs(re_))
    worst = max((v["max_rel_mag"] for k, v in out.items() if isinstance(v, dict)), default=0.0)
    out["max_rel_mag_all"] = worst
    return out
