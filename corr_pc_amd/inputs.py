"""Input files of the corr_pc program: the positional .inp parameter file and the catalog text files.

The layout follows the reference's reader and its generator notebook: data_ini reads one
"<label> <value>" pair per parameter in a fixed order (initialization.cpp:62-130; labels are not
checked), and read_data reads <prefix>_pos.dat (x y, or RA Dec), _z.dat, _e.dat (e1 e2), _wt.dat and
_jk.dat as whitespace-separated text (read_dat.cpp:164-300), which the notebook writes with numpy's
savetxt (Gen_inp_Periodic_box.ipynb, write_data).  "0" in place of a file name means "none".
"""
import os

import numpy as np

# The parameters in data_ini's order, with the labels the generator notebook writes
# (Gen_inp_Periodic_box.ipynb, inp_gen); the labels are informational only.
INP_FIELDS = (
    ("which_corr", "which_corr"), ("coordinates", "coordinates"), ("estimator", "estimator"),
    ("data_sorted", "data_sorted"), ("use_comoving", "use_comoving"), ("do_jk", "do_jk"),
    ("sig_crit", "sig_crit"),
    ("shape_pos_file", "shape_pos"), ("shape_z_file", "shape_z"), ("shape_e_file", "shape_e"),
    ("shape_wt_file", "shape_wt"), ("shape_jk_file", "shape_jk"), ("shape_patch_file", "density_patch"),
    ("density_pos_file", "density_pos"), ("density_z_file", "density_z"), ("density_wt_file", "density_wt"),
    ("density_jk_file", "density_jk"), ("density_e_file", "density_e"),
    ("shape_rand_pos_file", "Srandom_pos"), ("shape_rand_z_file", "Srandom_z"),
    ("shape_rand_wt_file", "Srandoms_wt"), ("shape_rand_jk_file", "Srand_jk"),
    ("shape_rand_patch_file", "Drandom_patch"),
    ("density_rand_pos_file", "Drandom_pos"), ("density_rand_z_file", "Drandom_z"),
    ("density_rand_wt_file", "Drandoms_wt"), ("density_rand_jk_file", "drand_jk"),
    ("distance_file", "distances"), ("patches_file", "patch_file"), ("out_file", "out_file"),
    ("n_threads", "n_threads"), ("n_shape", "n_shape"), ("n_density", "n_density"),
    ("n_shape_rand", "n_Srand"), ("n_density_rand", "n_Drand"), ("rand_subsample", "rand_subsample"),
    ("n_jk_regions", "n_jk"), ("n_patch", "n_patch"),
    ("bin_r_min", "binR_min"), ("bin_r_max", "binR_max"), ("n_bins", "n_bins"), ("lin_bin", "lin_bin"),
    ("n_p_bin", "n_p_bin"), ("p_min", "pmin"), ("p_max", "pmax"), ("z_min", "z_min"), ("z_max", "z_max"),
    ("dz", "dz"), ("z_sep_min", "z_sep_min"), ("z_sep_max", "z_sep_max"), ("periodic_box", "periodic_box"),
    ("periodic_box_size", "box_size"),
)

DEFAULTS = {name: "0" for name, _ in INP_FIELDS}
DEFAULTS.update({"data_sorted": 1, "use_comoving": 1, "dz": 0.0001})

CATALOG_ROLES = {  # role -> (.inp fields of its files, count field)
    "shape": (("shape_pos_file", "shape_z_file", "shape_e_file", "shape_wt_file", "shape_jk_file"), "n_shape"),
    "density": (("density_pos_file", "density_z_file", "density_e_file", "density_wt_file", "density_jk_file"),
                "n_density"),
    "shape_rand": (("shape_rand_pos_file", "shape_rand_z_file", None, "shape_rand_wt_file", "shape_rand_jk_file"),
                   "n_shape_rand"),
    "density_rand": (("density_rand_pos_file", "density_rand_z_file", None, "density_rand_wt_file",
                      "density_rand_jk_file"), "n_density_rand"),
}


def write_inp(path, **params):
    """Write a .inp file; parameters not given take DEFAULTS (0 / "0" = none)."""
    unknown = set(params) - {name for name, _ in INP_FIELDS}
    if unknown:
        raise ValueError(f"unknown parameters {sorted(unknown)}")
    vals = dict(DEFAULTS)
    vals.update(params)
    with open(path, "w") as f:
        for name, label in INP_FIELDS:
            f.write(f"{label:<16}{vals[name]}\n")
    return path


def write_catalog(prefix, x, y, z, e1=None, e2=None, w=None, jk=None, kappa=None, e3=None):
    """Write <prefix>_pos.dat, _z.dat and, when given, _e.dat (e1 e2, or e1 e2 e3 for the 3-D axes
    of which_corr 5) or _kappa.dat (one column of convergence, which_corr 3-4), _wt.dat, _jk.dat
    (write_data's layout).  Returns the .inp file entries (pos, z, e, wt, jk), "0" for the files not
    written."""
    os.makedirs(os.path.dirname(prefix) or ".", exist_ok=True)
    np.savetxt(prefix + "_pos.dat", np.column_stack((x, y)))
    np.savetxt(prefix + "_z.dat", np.asarray(z))
    files = [prefix + "_pos.dat", prefix + "_z.dat", "0", "0", "0"]
    if e1 is not None:
        np.savetxt(prefix + "_e.dat", np.column_stack((e1, e2) if e3 is None else (e1, e2, e3)))
        files[2] = prefix + "_e.dat"
    elif kappa is not None:
        np.savetxt(prefix + "_kappa.dat", np.asarray(kappa))
        files[2] = prefix + "_kappa.dat"
    if w is not None:
        np.savetxt(prefix + "_wt.dat", np.asarray(w))
        files[3] = prefix + "_wt.dat"
    if jk is not None:
        np.savetxt(prefix + "_jk.dat", np.asarray(jk), fmt="%i")
        files[4] = prefix + "_jk.dat"
    return files


def catalog_params(role, files, n):
    """The .inp entries of catalog `role` from write_catalog's file list."""
    fields, count = CATALOG_ROLES[role]
    out = {count: n}
    for field, f in zip(fields, files):
        if field is not None:
            out[field] = f
    return out


def write_distance_table(path, table):
    """The 5-column distance table da_Z reads with use_comoving 1 (initialization.cpp:37-44):
    z, H(z), an unused column, comoving DC, transverse comoving DA."""
    np.savetxt(path, np.column_stack((table["z"], table["H"], np.zeros_like(table["z"]), table["dC"], table["dA"])))
    return path
