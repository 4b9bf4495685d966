"""Pure-Python O(N^2) evaluation of corr_pc's box cross pass, for tiny catalogs.

Written independently of oracle/cpc_oracle.c from the reference's semantics: for every ordered
(outer i, inner j) and every periodic image the reference visits for i (lim_calc_PB,
PB_calcs.cpp:30-61), apply the window predicates (corr.cpp:67-86), the bin formulas
(corels.cpp:21-46, PB_calcs.cpp:87-98) and the pair kinds (corels.cpp:48-244) with Python floats
(IEEE double, libm transcendentals).  No z-sorted sweep: every (i, j, image) is enumerated, so this
pins the sweep's start/break logic as well.
"""
import math

import numpy as np

NF = 7


def _absolute(n):
    return -1 * n if n < 0 else n


def _ephi(phi, theta):
    t = _absolute(phi - theta)
    if t > math.pi / 2.0:
        t = math.pi - t
    return t


def _rot(e1, e2, th):
    return (math.cos(2.0 * th) * e1 - math.sin(2.0 * th) * e2, math.sin(2.0 * th) * e1 + math.cos(2.0 * th) * e2)


def _kind_fields(p, outer, i, inner, j, w, th):
    """The per-pair fields of the shape kinds (corels.cpp:171-244) for position angle th."""
    def g(c, name, k, default):
        a = getattr(c, name)
        return default if a is None else float(a[k])
    v = [0.0] * NF
    e = [0.0] * NF
    if p.kind == 1:
        src = (inner, j) if p.g_with_shape == 2 else (outer, i)
        ep, ex = _rot(float(src[0].e1[src[1]]), float(src[0].e2[src[1]]), th)
        v[1], e[1] = w * ep, (w * ep) ** 2
        v[2], e[2] = w * ex, (w * ex) ** 2
        v[6] = _ephi(g(outer, "phi", i, -10.0), th)
        e[6] = v[6] ** 2
        return v, e
    epi, exi = _rot(float(outer.e1[i]), float(outer.e2[i]), th)
    epj, exj = _rot(float(inner.e1[j]), float(inner.e2[j]), th)
    if p.estimator == 0:
        if p.g_with_shape == 2:
            v[1], e[1] = w * epj, (w * epj) ** 2
            v[6] = _ephi(g(inner, "phi", j, -10.0), th)
        else:
            v[1], e[1] = w * epi, (w * epi) ** 2
            v[6] = _ephi(g(outer, "phi", i, -10.0), th)
        e[6] = v[6] ** 2
    else:
        v[1] = w * (epi + epj) / 2.0
        e[1] = (w * epi / 2.0) ** 2 + (w * epj / 2.0) ** 2
        v[2] = w * (exi + exj) / 2.0
        e[2] = (w * exi / 2.0) ** 2 + (w * exj / 2.0) ** 2
        a = _ephi(g(outer, "phi", i, -10.0), th)
        b = _ephi(g(inner, "phi", j, -10.0), th)
        v[6], e[6] = a + b, a * a + b * b
    v[3], e[3] = w * (epi * epj), (w * (epi * epj)) ** 2
    v[4], e[4] = w * (exi * exj), (w * (exi * exj)) ** 2
    v[5] = w * (exi * epj + epi * exj) / 2.0
    e[5] = (w * (exi * epj) / 2.0) ** 2 + (w * epi * exj / 2.0) ** 2
    return v, e


def images(zi, xi, yi, L, pmin, pmax, rmax):
    """Image shifts (PBx, PBy, PBz) visited for an outer galaxy: 0 always, -1 if the window starts
    below 0, +1 if it ends above L."""
    def axis(lo, hi):
        s = [0]
        if lo < 0:
            s.append(-1)
        if hi > L:
            s.append(1)
        return s
    return [(a, b, c) for b in axis(yi - rmax, yi + rmax) for a in axis(xi - rmax, xi + rmax)
            for c in axis(zi + pmin, zi + pmax)]


def count(outer, inner, p):
    """corels2(outer, inner) after the swap; with p.auto_pairs, corels1 over `outer` (inner ignored):
    each pair (i, j) through an image with PBz <= 0 only for j after i, and an image only when its
    scan would not stop at once, i.e. when the first included galaxy the scan meets (the next one
    after i for PBz <= 0, the first of the catalog for PBz > 0) lies at or above the window's lower
    edge (corr.cpp:159-190 with do_pair_sorted's break, corr.cpp:71-73)."""
    auto = bool(getattr(p, "auto_pairs", 0))
    # data_sorted 0 (do_pair_unsorted, corr.cpp:36-52): no scan stops, and corels1 takes j > i under
    # every image (corr.cpp:167-182)
    unsorted = bool(getattr(p, "unsorted", 0))
    if auto:
        inner = outer
    L = p.box_size
    pmin, pmax, rmin, rmax = p.pi_min, p.pi_max, p.rp_min, p.rp_max
    psize = (pmax - pmin) / p.n_pi
    rsize = (rmax - rmin) / p.n_rp if p.lin_bin else (math.log10(rmax) - math.log10(rmin)) / p.n_rp
    jk_on = p.do_jk == 1 and p.n_jk > 0
    R = 1 + p.n_jk if jk_on else 1
    num = np.zeros((R, p.n_rp, p.n_pi), np.int64)
    wt = np.zeros((R, p.n_rp, p.n_pi))
    val = np.zeros((R, p.n_rp, p.n_pi, NF))
    err = np.zeros((R, p.n_rp, p.n_pi, NF))

    def g(c, name, i, default):
        a = getattr(c, name)
        return default if a is None else float(a[i])

    inc_o = [g(outer, "include_prob", i, 0.0) for i in range(outer.n)]
    inc_i = [g(inner, "include_prob", j, 0.0) for j in range(inner.n)]
    # the first included inner galaxy: the PBz=+1 sweep starts at j=0 and stops at once if it is
    # below the window (corr.cpp:407-416)
    first = next((j for j in range(inner.n) if not inc_i[j] > p.include_prob_max), None)
    for i in range(outer.n):
        if inc_o[i] > p.include_prob_max:
            continue
        xi, yi, zi = float(outer.x[i]), float(outer.y[i]), float(outer.z[i])
        zmin, zmax = zi + pmin, zi + pmax
        nxt = next((j for j in range(i + 1, inner.n) if not inc_i[j] > p.include_prob_max), None) if auto else None
        for (PBx, PBy, PBz) in images(zi, xi, yi, L, pmin, pmax, rmax):
            if not unsorted and PBz == 1 and (first is None or float(inner.z[first]) + L * 1 < zmin):
                continue
            if not unsorted and auto and PBz <= 0 and (nxt is None or float(inner.z[nxt]) + L * PBz < zmin):
                continue
            for j in range(inner.n):
                if inc_i[j] > p.include_prob_max:
                    continue
                if auto and (PBz <= 0 or unsorted) and j <= i:
                    continue
                xj, yj, zj = float(inner.x[j]), float(inner.y[j]), float(inner.z[j])
                zz = zj + L * PBz
                if zz > zmax or zz < zmin:
                    continue
                yy = yj + L * PBy
                if yy > yi + rmax or yy < yi - rmax:
                    continue
                xx = xj + L * PBx
                if xx < xi - rmax or xx > xi + rmax:
                    continue
                dP = zz - zi
                q2 = math.floor((dP - pmin) / psize)
                if not (0 <= q2 < p.n_pi):
                    continue
                t0, t1 = xi - xx, yi - yy
                rp = math.sqrt(t0 * t0 + t1 * t1)
                if p.lin_bin:
                    q1 = math.floor((rp - rmin) / rsize)
                else:
                    if rp == 0.0:
                        continue
                    q1 = math.floor(math.log10(rp / rmin) / rsize)
                if not (0 <= q1 < p.n_rp):
                    continue
                w = g(outer, "w", i, 1.0) * g(inner, "w", j, 1.0)
                v, e = [0.0] * NF, [0.0] * NF
                if p.kind != 0:
                    dx, dy = xj - xi, yj - yi
                    th = math.atan(dy / dx) if dx != 0 else math.copysign(math.pi / 2, dy)
                    v, e = _kind_fields(p, outer, i, inner, j, w, th)
                num[0, q1, q2] += 1
                wt[0, q1, q2] += w
                val[0, q1, q2] += v
                err[0, q1, q2] += e
                if jk_on:
                    a, b = int(outer.jk[i]), int(inner.jk[j])
                    pa, pb = g(outer, "jk_prob", i, 0.0), g(inner, "jk_prob", j, 0.0)
                    if a == b:
                        pa, pb = 1.0, 0.0
                    for r, pr in ((a, pa), (b, pb)):
                        if pr < 0.5:
                            continue
                        num[1 + r, q1, q2] += 1
                        wt[1 + r, q1, q2] += w
                        if p.which_corr != 0:
                            val[1 + r, q1, q2] += v
    if jk_on and p.jk_mode == 0:
        for r in range(1, R):
            num[r] = num[0] - num[r]
            wt[r] = wt[0] - wt[r]
            val[r] = val[0] - val[r]
            err[r] = err[0] - err[r]
    return {"num": num, "wt_num": wt, "val": val, "err": err, "num_rp": num.sum(axis=2)}


R2D = 180.0 / math.pi


def _interp(tab, dz, z):
    i = int(math.floor(z / dz))
    return tab[i] + ((tab[i + 1] - tab[i]) * (z - dz * i) / dz)


def sky_window(p, ra, dec, z):
    """lim_calc_sky + lim_calc_ra_dec for coordinates 0 (sky_calcs.cpp:105-190)."""
    H, dA, dz = p._table[0], p._table[1], p.sky_dz
    zmax = z + p.z_sep_max
    zmax = z + p.pi_max * _interp(H, dz, zmax) / 300000.0
    zmin = z + p.pi_min * _interp(H, dz, z) / 300000.0
    if zmin < 0.001:
        zmin = 0.001
    da_max = p.rp_max / _interp(dA, dz, zmin)
    dr = dec / R2D
    arg = (math.cos(da_max) - math.sin(dr) ** 2) / math.cos(dr) ** 2
    ra_lim = (math.acos(arg) if -1.0 <= arg <= 1.0 else float("nan")) * 1.02 * R2D
    ra_max, ra_min = ra + ra_lim, ra - ra_lim
    if ra_max > 360:
        ra_min = 0
    if ra_min < 0:
        ra_max = 360
    return zmin, zmax, ra_min, ra_max, dec - da_max * R2D * 1.05, dec + da_max * R2D * 1.05


def count_sky(outer, inner, p):
    """Every ordered (outer, inner) pair of the sky pass, no z-sorted sweep; with p.auto_pairs, the
    corels1 pairs (i, j > i) of `outer`, for i whose next included galaxy is not below its window."""
    auto = bool(getattr(p, "auto_pairs", 0))
    unsorted = bool(getattr(p, "unsorted", 0))
    if auto:
        inner = outer
    pmin, pmax, rmin, rmax = p.pi_min, p.pi_max, p.rp_min, p.rp_max
    psize = (pmax - pmin) / p.n_pi
    rsize = (rmax - rmin) / p.n_rp if p.lin_bin else (math.log10(rmax) - math.log10(rmin)) / p.n_rp
    jk_on = p.do_jk == 1 and p.n_jk > 0
    R = 1 + p.n_jk if jk_on else 1
    num = np.zeros((R, p.n_rp, p.n_pi), np.int64)
    wt = np.zeros((R, p.n_rp, p.n_pi))
    val = np.zeros((R, p.n_rp, p.n_pi, NF))
    err = np.zeros((R, p.n_rp, p.n_pi, NF))

    def g(c, name, i, default):
        a = getattr(c, name)
        return default if a is None else float(a[i])

    for i in range(outer.n):
        if g(outer, "include_prob", i, 0.0) > p.include_prob_max:
            continue
        rai, deci, zi = float(outer.x[i]), float(outer.y[i]), float(outer.z[i])
        zmin, zmax, ramin, ramax, decmin, decmax = sky_window(p, rai, deci, zi)
        if auto and not unsorted:
            nxt = next((j for j in range(i + 1, inner.n) if not g(inner, "include_prob", j, 0.0) > p.include_prob_max),
                       None)
            if nxt is None or float(inner.z[nxt]) < zmin:
                continue
        for j in range(inner.n):
            if g(inner, "include_prob", j, 0.0) > p.include_prob_max:
                continue
            if auto and j <= i:
                continue
            raj, decj, zj = float(inner.x[j]), float(inner.y[j]), float(inner.z[j])
            if zj > zmax or zj < zmin or decj > decmax or decj < decmin or raj < ramin or raj > ramax:
                continue
            dP = float(inner.dc[j]) - float(outer.dc[i])
            d1, d2 = deci / R2D, decj / R2D
            s = math.sin(d1) * math.sin(d2) + math.cos(d1) * math.cos(d2) * math.cos((rai / R2D - raj / R2D))
            da = math.acos(1.0) if s > 0.9999999999 else math.acos(s)
            rp = da * (float(outer.da[i]) + float(inner.da[j])) / 2.0
            q2 = math.floor((dP - pmin) / psize)
            if not (0 <= q2 < p.n_pi):
                continue
            if p.lin_bin:
                q1 = math.floor((rp - rmin) / rsize)
            else:
                if rp == 0.0:
                    continue
                q1 = math.floor(math.log10(rp / rmin) / rsize)
            if not (0 <= q1 < p.n_rp):
                continue
            w = g(outer, "w", i, 1.0) * g(inner, "w", j, 1.0)
            v, e = [0.0] * NF, [0.0] * NF
            if p.kind != 0:
                den = (raj - rai) * math.cos((d1 + d2) / 2.0)
                th = math.atan((decj - deci) / den) if den != 0 else math.copysign(math.pi / 2, decj - deci)
                v, e = _kind_fields(p, outer, i, inner, j, w, th)
            num[0, q1, q2] += 1
            wt[0, q1, q2] += w
            val[0, q1, q2] += v
            err[0, q1, q2] += e
            if jk_on:
                a, b = int(outer.jk[i]), int(inner.jk[j])
                pa, pb = g(outer, "jk_prob", i, 0.0), g(inner, "jk_prob", j, 0.0)
                if a == b:
                    pb = 0.0  # jk_check_sky
                for r, pr in ((a, pa), (b, pb)):
                    if pr < 0.5:
                        continue
                    num[1 + r, q1, q2] += 1
                    wt[1 + r, q1, q2] += w
                    if p.which_corr != 0:
                        val[1 + r, q1, q2] += v
    if jk_on and p.jk_mode == 0:
        for r in range(1, R):
            num[r] = num[0] - num[r]
            wt[r] = wt[0] - wt[r]
            val[r] = val[0] - val[r]
            err[r] = err[0] - err[r]
    return {"num": num, "wt_num": wt, "val": val, "err": err, "num_rp": num.sum(axis=2)}
