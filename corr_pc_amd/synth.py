"""Synthetic periodic-box workloads with the shapes of the benchmark configurations.

Every catalog is generated with numpy's PCG64 seeded 20261203 + 100*config + catalog_id
(data = 0, randoms = 1), stably sorted by z (corr_pc's data_sorted=1 input), and carries the
per-galaxy jk_prob / include_prob values that corr_pc's reader would draw from glibc rand() after
srand(124693517) in its read order S, Rs, D, Rd (do_corr.cpp:40-96, read_dat.cpp:243,318-320).
"""
from dataclasses import dataclass

import numpy as np

from .engine import Catalog, rand_stream

BASE_SEED = 20261203


@dataclass
class BoxConfig:
    name: str
    box: float
    n_data: int
    n_rand: int
    n_rp: int
    rp_min: float
    rp_max: float
    n_pi: int
    pi_max: float
    jk_side: int
    clustered: bool
    n_clumps: int
    shapes: bool
    weights: bool
    which_corr: int
    shape_randoms: int  # n_shape_rand (Rs); 0 when the configuration has none


CONFIGS = {
    # plain wgg box: 2e4 uniform data, 4e4 uniform randoms, 10 log rp bins in [0.1, 30), 20 pi bins
    # in [-60, 60), 2x2x2 jackknife
    0: BoxConfig("wgg box L=250, 2e4 D / 4e4 R", 250.0, 20_000, 40_000, 10, 0.1, 30.0, 20, 60.0, 2, False, 0,
                 False, False, 0, 40_000),
    # wg+/w++ box: 1e6 data (70% uniform, 30% in 5000 clumps), 3e6 randoms, 20 log rp bins in
    # [0.1, 100), 40 pi bins in [-100, 100), 3x3x3 jackknife, shapes and lognormal weights
    1: BoxConfig("wg+/w++ box L=500, 1e6 D / 3e6 R", 500.0, 1_000_000, 3_000_000, 20, 0.1, 100.0, 40, 100.0, 3,
                 True, 5000, True, True, 2, 0),
    # large box: 2e7 data in 40000 clumps, 4e7 randoms, 25 log rp bins in [0.1, 200), 60 pi bins in
    # [-150, 150), 4x4x4 jackknife
    4: BoxConfig("wg+/w++ box L=1000, 2e7 D / 4e7 R", 1000.0, 20_000_000, 40_000_000, 25, 0.1, 200.0, 60, 150.0, 4,
                 True, 40000, True, True, 2, 0),
}


def _rng(config, catalog_id):
    return np.random.Generator(np.random.PCG64(BASE_SEED + 100 * config + catalog_id))


def jk_grid(x, y, z, L, k):
    """3-D grid region ix*k^2 + iy*k + iz with ix = min(floor(x*k/L), k-1)."""
    ix = np.minimum(np.floor(x * k / L).astype(np.int64), k - 1)
    iy = np.minimum(np.floor(y * k / L).astype(np.int64), k - 1)
    iz = np.minimum(np.floor(z * k / L).astype(np.int64), k - 1)
    return (ix * k * k + iy * k + iz).astype(np.int32)


def positions(rng, n, L, clustered=False, n_clumps=0, frac_clumped=0.3, sigma=1.0):
    """Uniform positions, or 70% uniform + 30% Gaussian clumps (members wrapped mod L).
    Returns (pos[n,3], host clump centre or NaN)."""
    pos = rng.uniform(0.0, L, size=(n, 3))
    centre = np.full((n, 3), np.nan)
    if clustered and n_clumps > 0:
        n_c = int(round(frac_clumped * n))
        centres = rng.uniform(0.0, L, size=(n_clumps, 3))
        host = rng.integers(0, n_clumps, size=n_c)
        off = rng.normal(0.0, sigma, size=(n_c, 3))
        idx = rng.permutation(n)[:n_c]
        pos[idx] = np.mod(centres[host] + off, L)
        centre[idx] = centres[host]
    # fl(x mod L) can round to L; keep coordinates in [0, L)
    pos[pos >= L] = 0.0
    return pos, centre


def shapes(rng, pos, centre, L, sigma_e=0.25, align=0.02):
    """(e1, e2): Gaussian per component, |e| < 1, plus a radial alignment of `align` toward the host
    clump centre in the x-y plane, signed so that corr_pc's e+ = cos2t e1 - sin2t e2 (calcs.cpp:60-64)
    sees +align for galaxies pointing at their centre."""
    n = len(pos)
    d = centre[:, :2] - pos[:, :2]
    d = d - L * np.round(d / L)
    has = np.isfinite(d[:, 0])
    alpha = np.where(has, np.arctan2(np.nan_to_num(d[:, 1]), np.nan_to_num(d[:, 0])), 0.0)
    a1 = np.where(has, align * np.cos(2 * alpha), 0.0)
    a2 = np.where(has, -align * np.sin(2 * alpha), 0.0)
    e1 = np.empty(n)
    e2 = np.empty(n)
    todo = np.arange(n)
    while len(todo):
        r1 = rng.normal(0.0, sigma_e, len(todo)) + a1[todo]
        r2 = rng.normal(0.0, sigma_e, len(todo)) + a2[todo]
        ok = r1 * r1 + r2 * r2 < 1.0
        e1[todo[ok]] = r1[ok]
        e2[todo[ok]] = r2[ok]
        todo = todo[~ok]
    return e1, e2


def elip_angle(e1, e2):
    """Per-galaxy phi of elip_angle (calcs.cpp:73-83): 0.5*arctan(e1, -e2) in [-pi/2, pi/2], NaN -> 0."""
    x = e1
    y = -e2
    with np.errstate(divide="ignore", invalid="ignore"):
        t = np.arctan(y / x)
    t = np.where(x < 0, np.where(y < 0, t - np.pi, t + np.pi), t)
    phi = 0.5 * t
    return np.where(np.isnan(phi), 0.0, phi)


def lognormal_weights(rng, n, sigma=0.3):
    w = np.exp(rng.normal(0.0, sigma, n))
    return w / w.mean()


def _sorted(cols):
    order = np.argsort(cols["z"], kind="stable")
    return {k: (None if v is None else v[order]) for k, v in cols.items()}


def box_data(config, n=None):
    """The data catalog of a box configuration (columns, z-sorted, no coins yet)."""
    c = CONFIGS[config]
    n = c.n_data if n is None else n
    rng = _rng(config, 0)
    pos, centre = positions(rng, n, c.box, c.clustered, c.n_clumps)
    cols = {"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2], "w": None, "e1": None, "e2": None, "phi": None}
    if c.shapes:
        e1, e2 = shapes(rng, pos, centre, c.box)
        cols.update(e1=e1, e2=e2, phi=elip_angle(e1, e2))
    if c.weights:
        cols["w"] = lognormal_weights(rng, n)
    cols["jk"] = jk_grid(pos[:, 0], pos[:, 1], pos[:, 2], c.box, c.jk_side)
    return _sorted(cols)


def box_randoms(config, n=None):
    c = CONFIGS[config]
    n = c.n_rand if n is None else n
    rng = _rng(config, 1)
    pos, _ = positions(rng, n, c.box)
    cols = {"x": pos[:, 0], "y": pos[:, 1], "z": pos[:, 2], "w": None, "e1": None, "e2": None, "phi": None}
    cols["jk"] = jk_grid(pos[:, 0], pos[:, 1], pos[:, 2], c.box, c.jk_side)
    return _sorted(cols)


def catalogs(config, n_data=None, n_rand=None):
    """Catalogs S, Rs, D, Rd of a box configuration with the reference's coin stream.

    S and D are the same galaxies (shape and density samples of one file), Rs and Rd the same
    randoms; each copy gets its own draws, as corr_pc reads each file once per role."""
    c = CONFIGS[config]
    data = box_data(config, n_data)
    rand = box_randoms(config, n_rand)
    nd = len(data["x"])
    nr = len(rand["x"])
    n_rs = nr if c.shape_randoms else 0
    # draws in read order: S (jk_prob), Rs (jk_prob, include_prob), D (jk_prob), Rd (jk_prob, include_prob)
    total = nd + 2 * n_rs + nd + 2 * nr
    stream = rand_stream(total)
    k = 0
    s_jkp = stream[k:k + nd]
    k += nd
    rs_draw = stream[k:k + 2 * n_rs]
    k += 2 * n_rs
    d_jkp = stream[k:k + nd]
    k += nd
    rd_draw = stream[k:k + 2 * nr]
    keep_e = c.which_corr == 2  # density catalog keeps shapes for shape-shape (read_dat.cpp:105-112)

    def cat(cols, jkp, inc, with_e):
        return Catalog(x=cols["x"], y=cols["y"], z=cols["z"], w=cols["w"],
                       e1=cols["e1"] if with_e else None, e2=cols["e2"] if with_e else None,
                       phi=cols["phi"] if with_e else None, jk=cols["jk"], jk_prob=jkp, include_prob=inc)

    S = cat(data, s_jkp, None, c.shapes)
    D = cat(data, d_jkp, None, c.shapes and keep_e)
    Rs = cat(rand, rs_draw[0::2], rs_draw[1::2], False) if n_rs else None
    Rd = cat(rand, rd_draw[0::2], rd_draw[1::2], False)
    return {"S": S, "Rs": Rs, "D": D, "Rd": Rd, "config": c}


# ---------------------------------------------------------------------------------------------
# Batched mocks: many independent wgg boxes of uneven size

MOCK_CONFIG = 3
MOCK_BOX = 250.0
MOCK_N_MIN, MOCK_N_MAX = 10_000, 500_000
MOCK_PER_CLUMP = 60
# binning of the plain wgg box (10 log rp bins in [0.1, 30), 20 pi bins in [-60, 60)), 2x2x2 jackknife
MOCK_BINNING = CONFIGS[0]


def mock_sizes(n_mocks=256, n_min=MOCK_N_MIN, n_max=MOCK_N_MAX):
    """Data sizes of the mocks, log-uniform in [n_min, n_max] (catalog id 0 of the mock configuration)."""
    rng = _rng(MOCK_CONFIG, 0)
    return np.exp(rng.uniform(np.log(n_min), np.log(n_max), n_mocks)).astype(np.int64)


def mock_catalogs(k, n):
    """Catalogs S, Rs, D, Rd of mock k with n data galaxies: clustered as the wg+ box (30% of the
    galaxies in Gaussian clumps of sigma 1, MOCK_PER_CLUMP members per clump), 2n uniform randoms,
    unit weights, 2x2x2 jackknife; positions from catalog id 2 + k.  Every mock is its own corr_pc
    run, so each draws its coins from the start of the reference's stream (read order S, Rs, D, Rd)."""
    L = MOCK_BOX
    rng = _rng(MOCK_CONFIG, 2 + k)
    n_clumps = max(1, int(round(0.3 * n / MOCK_PER_CLUMP)))
    pos, _ = positions(rng, n, L, True, n_clumps)
    rpos, _ = positions(rng, 2 * n, L)

    def cols(p):
        c = {"x": p[:, 0], "y": p[:, 1], "z": p[:, 2], "w": None, "e1": None, "e2": None, "phi": None,
             "jk": jk_grid(p[:, 0], p[:, 1], p[:, 2], L, 2)}
        return _sorted(c)

    data, rand = cols(pos), cols(rpos)
    nr = 2 * n
    stream = rand_stream(n + 2 * nr + n + 2 * nr)
    s_jkp, rs_draw = stream[:n], stream[n:n + 2 * nr]
    d_jkp, rd_draw = stream[n + 2 * nr:2 * n + 2 * nr], stream[2 * n + 2 * nr:]

    def cat(c, jkp, inc):
        return Catalog(x=c["x"], y=c["y"], z=c["z"], jk=c["jk"], jk_prob=jkp, include_prob=inc)

    return {"S": cat(data, s_jkp, None), "Rs": cat(rand, rs_draw[0::2], rs_draw[1::2]),
            "D": cat(data, d_jkp, None), "Rd": cat(rand, rd_draw[0::2], rd_draw[1::2]), "config": MOCK_BINNING}


# ---------------------------------------------------------------------------------------------
# Sky survey (RA/Dec/redshift)

C_LIGHT = 299792.458  # km/s


@dataclass
class SkyConfig:
    name: str
    n_data: int
    n_rand: int
    fields: tuple  # (ra_lo, ra_hi, dec_lo, dec_hi, n_jk_ra, n_jk_dec) in degrees
    z_lo: float
    z_hi: float
    omega_m: float
    dz: float
    n_rp: int
    rp_min: float
    rp_max: float
    n_pi: int
    pi_max: float
    which_corr: int


# six 5x5 degree fields (about 150 deg^2 in total), |Dec| <= 30; 100 jackknife regions on an RA/Dec
# grid (4x4 in each of the first four fields, 6x3 in the last two)
SKY_FIELDS = ((20.0, 25.0, -20.0, -15.0, 4, 4), (80.0, 85.0, 0.0, 5.0, 4, 4), (140.0, 145.0, 10.0, 15.0, 4, 4),
              (200.0, 205.0, -10.0, -5.0, 4, 4), (260.0, 265.0, 5.0, 10.0, 6, 3), (320.0, 325.0, -25.0, -20.0, 6, 3))

SKY_CONFIGS = {
    2: SkyConfig("sky survey, 2e6 D / 2e7 R, 6 fields ~150 deg^2", 2_000_000, 20_000_000, SKY_FIELDS, 0.2, 1.0, 0.3,
                 1e-4, 20, 0.1, 100.0, 40, 100.0, 1),
}


def distance_table(omega_m=0.3, dz=1e-4, z_max=1.0):
    """Flat LCDM distance table in corr_pc's 5-column layout with use_comoving=1
    (initialization.cpp:37-44): z, H(z) [km/s/(Mpc/h)], unused, comoving DC, transverse comoving DA
    (equal to DC when flat), rows k*dz for k = 0 .. ceil((z_max+1)/dz)+20 (the reference reads
    ceil((z_max+1)/dz)+1 rows and allocates 20 more)."""
    n = int(np.ceil((z_max + 1) / dz)) + 21
    z = np.arange(n) * dz
    H = 100.0 * np.sqrt(omega_m * (1 + z) ** 3 + (1 - omega_m))
    # comoving distance by the trapezoid rule on a 10x finer grid
    zf = np.linspace(0.0, z[-1], 10 * (n - 1) + 1)
    inv = C_LIGHT / (100.0 * np.sqrt(omega_m * (1 + zf) ** 3 + (1 - omega_m)))
    chi_f = np.concatenate([[0.0], np.cumsum(0.5 * (inv[1:] + inv[:-1]) * np.diff(zf))])
    dc = chi_f[::10]
    return {"z": z, "H": H, "dC": dc, "dA": dc.copy(), "dz": dz}


def _sky_positions(rng, n, fields):
    """Uniform on the sphere inside the union of RA/Dec rectangles (area-weighted)."""
    areas = np.array([(f[1] - f[0]) * (np.sin(np.radians(f[3])) - np.sin(np.radians(f[2]))) for f in fields])
    which = rng.choice(len(fields), size=n, p=areas / areas.sum())
    ra = np.empty(n)
    dec = np.empty(n)
    jk = np.empty(n, np.int32)
    base = 0
    for k, f in enumerate(fields):
        m = which == k
        c = int(m.sum())
        ra[m] = rng.uniform(f[0], f[1], c)
        s = rng.uniform(np.sin(np.radians(f[2])), np.sin(np.radians(f[3])), c)
        dec[m] = np.degrees(np.arcsin(s))
        ia = np.minimum(((ra[m] - f[0]) / (f[1] - f[0]) * f[4]).astype(int), f[4] - 1)
        ib = np.minimum(((dec[m] - f[2]) / (f[3] - f[2]) * f[5]).astype(int), f[5] - 1)
        jk[m] = base + ia * f[5] + ib
        base += f[4] * f[5]
    return ra, dec, jk, base


def _nz_draw(rng, n, z_lo, z_hi):
    """n(z) proportional to z^2 exp(-(z/0.4)^1.5) on (z_lo, z_hi), by inverse CDF."""
    zg = np.linspace(z_lo, z_hi, 200001)
    pdf = zg ** 2 * np.exp(-(zg / 0.4) ** 1.5)
    cdf = np.concatenate([[0.0], np.cumsum(0.5 * (pdf[1:] + pdf[:-1]) * np.diff(zg))])
    cdf /= cdf[-1]
    z = np.interp(rng.uniform(0.0, 1.0, n), cdf, zg)
    return np.clip(z, z_lo, np.nextafter(z_hi, 0))


def sky_catalog(config, which, n=None, table=None):
    """Data (which='data') or randoms of a sky configuration, z-sorted, with the reader's derived
    fields: DC = dC_z[floor(z/dz)], DA = dA_z[floor(z/dz)] (read_dat.cpp:249-251), jk_prob = 1
    (read_dat.cpp:239) and, for data, shapes and lognormal weights."""
    c = SKY_CONFIGS[config]
    table = table or distance_table(c.omega_m, c.dz, c.z_hi)
    n = (c.n_data if which == "data" else c.n_rand) if n is None else n
    rng = _rng(config, 0 if which == "data" else 1)
    ra, dec, jk, n_jk = _sky_positions(rng, n, c.fields)
    z = _nz_draw(rng, n, c.z_lo, c.z_hi)
    cols = {"x": ra, "y": dec, "z": z, "jk": jk, "w": None, "e1": None, "e2": None, "phi": None}
    if which == "data":
        e1, e2 = shapes(rng, np.zeros((n, 3)), np.full((n, 3), np.nan), 1.0)
        cols.update(e1=e1, e2=e2, phi=elip_angle(e1, e2), w=lognormal_weights(rng, n))
    cols = _sorted(cols)
    it = np.floor(cols["z"] / table["dz"]).astype(np.int64)
    cols["dc"] = table["dC"][it]
    cols["da"] = table["dA"][it]
    return Catalog(x=cols["x"], y=cols["y"], z=cols["z"], w=cols["w"], e1=cols["e1"], e2=cols["e2"],
                   phi=cols["phi"], jk=cols["jk"], jk_prob=np.ones(n), dc=cols["dc"], da=cols["da"]), n_jk, table


def sky_params_table(config, table):
    """(H, dA, dz, z_max, z_sep_max) as data_ini derives them for coordinates 0
    (initialization.cpp:223, 272-276): z_sep_max = p_max*H[ceil(z_max/dz)]/300000 + 0.001."""
    c = SKY_CONFIGS[config]
    z_idx = int(np.ceil(c.z_hi / table["dz"]))
    z_sep_max = c.pi_max * table["H"][z_idx] / 300000.0 + 0.001
    return table["H"], table["dA"], table["dz"], c.z_hi, z_sep_max
