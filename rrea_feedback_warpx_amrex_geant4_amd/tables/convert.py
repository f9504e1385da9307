"""Convert the reference's schema-6 production bundle into one committed
``.npz`` of raw columns (``data/schema6_air.npz``).

The bundle (``rrea_transport_tables/schema6/production``) is RREATBL data, not
code.  This converter keeps every value bit-for-bit (float64 in, float64 out)
and regroups it the way ``RreaInteractionTables::Load``
(native_amrex_rrea/src/RreaInteractionTables.cpp:1659-2230) regroups it:

* rate/stopping tables keep their x grid and value column
  (``load_table`` :484-557, plus ``Finalize`` :1632-1657 for the mode);
* conditional inverse CDFs are split into primary-energy groups
  (``load_conditional_cdf`` :559-619);
* GS / WentzelVI angle tables become dense (energy, path, u) cubes, ordered
  by sorted energy then sorted path with the u nodes in file order
  (``load_gs_angle_table`` :687-772);
* elastic parameter tables keep their four columns
  (``load_elastic_mfp_table`` :650-685).

Derived quantities (node logarithms, the hard-Moller mean secondary energy,
conditional-CDF group means) are NOT stored: the native host loader computes
them at table creation, exactly as ``Load`` does.

Run once, in a checkout that has the reference bundle:

    python -m rrea_feedback_warpx_amrex_geant4_amd.tables.convert \
        /path/to/rrea_transport_tables/schema6/production
"""

from __future__ import annotations

import json
import os
import sys

import numpy as np

from . import rreatbl

# Interpolation mode ids shared with the native loader (csrc/rrea_tables.h).
MODES = {
    "loglog_linear": 0,
    "loglog_positive_with_zero_threshold": 1,
    "logx_linear": 2,
    "linear": 3,
}

RATE_TABLES = [
    # (image key, bundle table name)
    ("e_col", "electron_collision_stopping_air"),
    ("e_soft", "electron_soft_radiative_stopping_air"),
    ("e_moller", "electron_hard_moller_inverse_length_air"),
    ("e_brems", "electron_tracked_brems_inverse_length_air"),
    ("e_eltr", "electron_elastic_transport_inverse_length_air"),
    ("p_col", "positron_collision_stopping_air"),
    ("p_soft", "positron_soft_radiative_stopping_air"),
    ("p_bhabha", "positron_hard_bhabha_inverse_length_air"),
    ("p_brems", "positron_tracked_brems_inverse_length_air"),
    ("p_eltr", "positron_elastic_transport_inverse_length_air"),
    ("p_annih", "positron_annihilation_inverse_length_air"),
    ("g_compton", "photon_compton_inverse_length_air"),
    ("g_pe", "photon_photoelectric_inverse_length_air"),
    ("g_pairn", "photon_pair_nuclear_inverse_length_air"),
    ("g_pairt", "photon_pair_triplet_inverse_length_air"),
]

CDF_TABLES = [
    ("moller", "electron_moller_secondary_cdf_air"),
    ("bhabha", "positron_hard_bhabha_secondary_cdf_air"),
    ("ebrems", "brems_photon_energy_cdf_air"),
    ("pbrems", "positron_brems_photon_energy_cdf_air"),
]

ANGLE_TABLES = [
    ("gs_e", "electron_gs_angle_cdf_air"),
    ("gs_p", "positron_gs_angle_cdf_air"),
    ("wz_e", "electron_wentzel_vi_angle_cdf_air"),
    ("wz_p", "positron_wentzel_vi_angle_cdf_air"),
]

ELASTIC_TABLES = [
    ("el_e", "electron_elastic_mfp_air"),
    ("el_p", "positron_elastic_mfp_air"),
]


def _angle_cube(container):
    energies = container.column("primary_energy_eV")
    paths = container.column("path_over_elastic_mfp")
    us = container.column("cdf")
    xs = container.column("one_minus_cos_theta")
    e_grid = np.unique(energies)
    first = energies == e_grid[0]
    p_grid = np.unique(paths[first])
    first_group = first & (paths == p_grid[0])
    u_grid = us[first_group]
    ne, npth, nu = len(e_grid), len(p_grid), len(u_grid)
    cube = np.empty((ne, npth, nu), dtype=np.float64)
    e_index = np.searchsorted(e_grid, energies)
    p_index = np.searchsorted(p_grid, paths)
    counts = np.zeros((ne, npth), dtype=np.int64)
    for row in range(len(energies)):
        ie, ip = e_index[row], p_index[row]
        iu = counts[ie, ip]
        if us[row] != u_grid[iu]:
            raise ValueError("angle CDF u-nodes differ across (E,n) groups")
        cube[ie, ip, iu] = xs[row]
        counts[ie, ip] += 1
    if not np.all(counts == nu):
        raise ValueError("angle CDF node count varies across groups")
    return e_grid, p_grid, u_grid, cube


def convert(bundle_dir: str, out_path: str) -> dict:
    with open(os.path.join(bundle_dir, "transport_physics.json")) as handle:
        config = json.load(handle)
    if int(config["schema_version"]) != 6:
        raise ValueError("only schema 6 is supported")
    tables = {t["name"]: t for t in config["tables"]}
    out: dict[str, np.ndarray] = {}

    for key, name in RATE_TABLES:
        spec = tables[name]
        c = rreatbl.read(os.path.join(bundle_dir, spec["path"]))
        out[key + "_x"] = c.column(spec.get("x_column", "energy_eV"))
        out[key + "_y"] = c.column(spec["value_column"])
        out[key + "_mode"] = np.array(MODES[spec.get("interpolation", "loglog_linear")], dtype=np.int64)

    for key, name in CDF_TABLES:
        spec = tables[name]
        c = rreatbl.read(os.path.join(bundle_dir, spec["path"]))
        primary = c.column("primary_energy_eV")
        cdf = c.column("cdf_u")
        values = c.column(spec["value_column"])
        starts = np.flatnonzero(np.r_[True, primary[1:] != primary[:-1]])
        out[key + "_primary"] = primary[starts]
        out[key + "_start"] = np.r_[starts, len(primary)].astype(np.int64)
        out[key + "_u"] = cdf
        out[key + "_v"] = values

    for key, name in ANGLE_TABLES:
        spec = tables[name]
        c = rreatbl.read(os.path.join(bundle_dir, spec["path"]))
        e_grid, p_grid, u_grid, cube = _angle_cube(c)
        out[key + "_e"] = e_grid
        out[key + "_n"] = p_grid
        out[key + "_u"] = u_grid
        out[key + "_omc"] = cube

    for key, name in ELASTIC_TABLES:
        spec = tables[name]
        c = rreatbl.read(os.path.join(bundle_dir, spec["path"]))
        out[key + "_e"] = c.column("primary_energy_eV")
        out[key + "_mfp"] = c.column("elastic_mfp_m_stp")
        out[key + "_eta"] = c.column("screening_eta")
        out[key + "_trmfp"] = c.column("transport_mfp_m_stp")

    # Photoelectric subshell CDF (load_photoelectric_subshell_cdf :893-990): a
    # new group starts at a new primary energy or where the cdf resets at a
    # duplicated absorption-edge energy.
    spec = tables["photoelectric_subshell_cdf_air"]
    c = rreatbl.read(os.path.join(bundle_dir, spec["path"]))
    primary = c.column("primary_energy_eV")
    cdf = c.column("cdf_u")
    starts = [0]
    for i in range(1, len(primary)):
        if primary[i] != primary[i - 1] or cdf[i] <= cdf[i - 1]:
            starts.append(i)
    starts = np.array(starts, dtype=np.int64)
    out["pe_primary"] = primary[starts]
    out["pe_start"] = np.r_[starts, len(primary)].astype(np.int64)
    out["pe_u"] = cdf
    out["pe_binding"] = c.column("binding_energy_eV")
    out["pe_relaxation"] = c.column("relaxation_energy_eV")

    scattering = config["elastic_scattering_model"]
    out["scalars"] = np.array([
        config["hard_moller_secondary_threshold_eV"],
        config["brems_absolute_cutoff_eV"],
        config["reference_density_kg_m3"],
        config["charged_energy_range_eV"][0],
        config["charged_energy_range_eV"][1],
        config["photon_energy_range_eV"][0],
        config["photon_energy_range_eV"][1],
        scattering["few_collision_threshold_n"],
        scattering["isotropic_above_path_n"],
        scattering["transition_energy_eV"],
    ], dtype=np.float64)
    np.savez_compressed(out_path, **out)
    return out


def convert_pair_bank(bundle_dir: str, out_path: str) -> dict:
    """The training blocks of the pair final-state record bank
    (load_pair_empirical_blocks NA/src/RreaInteractionTables.cpp:1187-1369),
    each record's leptons in canonical order: the first pdg 11 slot (pair
    electron), the pdg -11 slot (positron), the second pdg 11 slot (triplet
    recoil; zero for nuclear records) -- the slot choice of
    BuildRreaPairEmpiricalFinalState :1384-1420, made once here."""
    bank = rreatbl.read(os.path.join(bundle_dir, "final_state/pair_final_state_samples_air.bin.rtb"))
    with open(os.path.join(bundle_dir, "final_state/pair_final_state_samples_air.index.json")) as handle:
        index = json.load(handle)
    if index.get("failures"):
        raise ValueError("pair bank index reports producer failures")
    pdg = np.stack([bank.column(f"p{s}_pdg") for s in range(3)], 1)
    frac = np.stack([bank.column(f"p{s}_kinetic_fraction") for s in range(3)], 1)
    dirs = np.stack([np.stack([bank.column(f"p{s}_dir_{a}") for a in "xyz"], 1) for s in range(3)], 1)
    channel_code = bank.column("channel_code")
    blocks = [b for b in index["blocks"] if b["split"] == "training" and int(b["count"]) > 0]
    code = {"nuclear": 1, "triplet": 2}
    blocks.sort(key=lambda b: (code[b["channel"]], float(b["primary_energy_eV"])))
    rec_frac, rec_dir, starts, energies, channels = [], [], [0], [], []
    for b in blocks:
        lo, n = int(b["record_offset"]), int(b["count"])
        rows = np.arange(lo, lo + n)
        if np.any(channel_code[rows] != code[b["channel"]]):
            raise ValueError("pair record channel differs from its block")
        for row in rows:
            electrons = [s for s in range(3) if pdg[row, s] == 11]
            positrons = [s for s in range(3) if pdg[row, s] == -11]
            order = [electrons[0], positrons[0], electrons[1] if len(electrons) > 1 else -1]
            rec_frac.append([frac[row, s] if s >= 0 else 0.0 for s in order])
            rec_dir.append(np.concatenate([dirs[row, s] if s >= 0 else np.zeros(3) for s in order]))
        starts.append(starts[-1] + n)
        energies.append(float(b["primary_energy_eV"]))
        channels.append(code[b["channel"]])
    out = {
        "pair_block_channel": np.array(channels, dtype=np.int64),
        "pair_block_energy": np.array(energies, dtype=np.float64),
        "pair_block_start": np.array(starts, dtype=np.int64),
        "pair_frac": np.array(rec_frac, dtype=np.float64),
        "pair_dir": np.array(rec_dir, dtype=np.float64),
    }
    np.savez_compressed(out_path, **out)
    return out


def convert_validation(bundle_dir: str, out_path: str) -> None:
    """Quarterpoint holdouts: the bundle's own (energy, channel, runtime_value)
    known-answer rows for the rate interpolation."""
    c = rreatbl.read(os.path.join(bundle_dir, "validation/charged_transport_quarterpoints_air.csv.rtb"))
    p = rreatbl.read(os.path.join(bundle_dir, "validation/photon_process_quarterpoints_air.csv.rtb"))
    np.savez_compressed(
        out_path,
        charged_energy=c.column("energy_eV"),
        charged_species=np.array(c.text_column("species")),
        charged_channel=np.array(c.text_column("channel")),
        charged_runtime=c.column("runtime_value"),
        charged_kind=np.array(c.text_column("validation_point_kind")),
        photon_energy=p.column("energy_eV"),
        photon_channel=np.array(p.text_column("channel")),
        photon_runtime=p.column("runtime_value"),
        photon_kind=np.array(p.text_column("validation_point_kind")),
    )


if __name__ == "__main__":
    bundle = sys.argv[1]
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    convert(bundle, os.path.join(here, "data", "schema6_air.npz"))
    convert_pair_bank(bundle, os.path.join(here, "data", "pair_bank_air.npz"))
