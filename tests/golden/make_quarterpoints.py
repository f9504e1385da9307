"""Writes tests/golden/quarterpoints.npz from the reference bundle's own
validation holdouts (data files, not code):

  rrea_transport_tables/schema6/production/validation/charged_transport_quarterpoints_air.csv.rtb
  rrea_transport_tables/schema6/production/validation/photon_process_quarterpoints_air.csv.rtb

Each row holds an energy, a channel and the `runtime_value` the reference's
interpolation produced there; they pin the schema-6 rate interpolation.

    python tests/golden/make_quarterpoints.py /root/reference/rrea_transport_tables/schema6/production
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", ".."))
from rrea_feedback_warpx_amrex_geant4_amd.tables import rreatbl  # noqa: E402

CHARGED = {
    ("electron", "collision_stopping_eV_per_m_stp"): "e_col",
    ("electron", "soft_radiative_stopping_eV_per_m_stp"): "e_soft",
    ("electron", "hard_moller_inverse_m_stp"): "e_moller",
    ("electron", "tracked_brems_inverse_m_stp"): "e_brems",
    ("electron", "elastic_transport_inverse_m_stp"): "e_eltr",
    ("positron", "collision_stopping_eV_per_m_stp"): "p_col",
    ("positron", "soft_radiative_stopping_eV_per_m_stp"): "p_soft",
    ("positron", "hard_bhabha_inverse_m_stp"): "p_bhabha",
    ("positron", "tracked_brems_inverse_m_stp"): "p_brems",
    ("positron", "elastic_transport_inverse_m_stp"): "p_eltr",
    ("positron", "annihilation_inverse_m_stp"): "p_annih",
}
PHOTON = {"compton": "g_compton", "photoelectric": "g_pe", "pair_nuclear": "g_pairn", "pair_triplet": "g_pairt"}
KEYS = list(CHARGED.values()) + list(PHOTON.values())


def main(bundle):
    c = rreatbl.read(os.path.join(bundle, "validation/charged_transport_quarterpoints_air.csv.rtb"))
    p = rreatbl.read(os.path.join(bundle, "validation/photon_process_quarterpoints_air.csv.rtb"))
    table = [KEYS.index(CHARGED[(s, ch)]) for s, ch in zip(c.text_column("species"), c.text_column("channel"))]
    table += [KEYS.index(PHOTON[ch]) for ch in p.text_column("channel")]
    energy = np.concatenate([c.column("energy_eV"), p.column("energy_eV")])
    runtime = np.concatenate([c.column("runtime_value"), p.column("runtime_value")])
    out = os.path.join(os.path.dirname(__file__), "quarterpoints.npz")
    np.savez_compressed(out, energy=energy, table=np.array(table, dtype=np.int8), runtime=runtime,
                        keys=np.array(KEYS))
    print(out, len(energy))


if __name__ == "__main__":
    main(sys.argv[1])
