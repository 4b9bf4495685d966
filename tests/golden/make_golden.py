"""Writes the committed fixtures of tests/golden/ (run by hand; the outputs are checked in).

glibc_rand_124693517.json
    The first draws of corr_pc's coin stream, rand()/(RAND_MAX+1.) after srand(124693517)
    (jk.cpp:40-43, read_dat.cpp:243), computed here from glibc's published TYPE_3 additive feedback
    generator (r[i] = r[i-3] + r[i-31] mod 2^32 seeded by the 16807 LCG, 310 draws discarded,
    output r >> 1) -- not from any C library call -- so that a host whose rand() differs is detected.
rp_thresholds.json
    The squared-separation edges of the rp bins of the benchmark binnings as computed by this
    container's libm (glibc 2.35), in hex, so that a machine whose log10 moves an edge is detected.
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def glibc_rand(seed, n):
    """glibc random()/rand() TYPE_3 state machine (stdlib/random_r.c) restated."""
    r = [0] * (344 + n)
    r[0] = seed if seed != 0 else 1
    word = r[0]
    for i in range(1, 31):
        hi, lo = word // 127773, word % 127773
        word = 16807 * lo - 2836 * hi
        if word < 0:
            word += 2147483647
        r[i] = word
    for i in range(31, 34):
        r[i] = r[i - 31]
    for i in range(34, 344 + n):
        r[i] = (r[i - 31] + r[i - 3]) & 0xFFFFFFFF
    return [r[i] >> 1 for i in range(344, 344 + n)]


BINNINGS = {
    "config0": dict(n_rp=10, rp_min=0.1, rp_max=30.0, lin_bin=0),
    "config1": dict(n_rp=20, rp_min=0.1, rp_max=100.0, lin_bin=0),
    "config4": dict(n_rp=25, rp_min=0.1, rp_max=200.0, lin_bin=0),
    "linear": dict(n_rp=8, rp_min=0.0, rp_max=8.0, lin_bin=1),
}


def main():
    sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
    draws = glibc_rand(124693517, 64)
    with open(os.path.join(HERE, "glibc_rand_124693517.json"), "w") as f:
        json.dump({"seed": 124693517, "rand": draws, "RAND_MAX": 2147483647}, f, indent=0)
    from corr_pc_amd.engine import make_params, rp_thresholds
    out = {}
    for name, b in BINNINGS.items():
        p = make_params(1000.0, b["n_rp"], b["rp_min"], b["rp_max"], 10, -10.0, 10.0, lin_bin=b["lin_bin"])
        out[name] = {"binning": b, "thresholds_hex": [float(t).hex() for t in rp_thresholds(p)]}
    with open(os.path.join(HERE, "rp_thresholds.json"), "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
