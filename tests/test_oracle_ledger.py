"""The oracle's event ledger (RreaEventLedger, native_amrex_rrea/include/rrea/
RreaEventLedger.H:18-96, recorded where the reference's transport records it)
pinned by what the step's other outputs fix independently:

- event counts equal the step's own event tallies and its secondaries by
  process (hard Moller, Bhabha, brems, Compton, photoelectric, pair), and the
  photon interaction count is their sum;
- the running maxima equal the maxima over the step's outputs (the largest
  hard-Moller secondary, the largest substep count, the largest transported
  energy);
- sums over the secondaries (Moller and Bhabha secondary energy, kinetic
  pair positrons, kinetic Compton electrons) and the escaped photon energy
  from the final states;
- the ionization records split exactly into low and kinetic secondaries,
  and differ from the ion-pair tally only by the Bhabha events (a positive
  ion, no ionization source) and the triplet recoils (an ionization source
  with no transport-ionization record);
- the expected hard-Moller event weight (the per-segment competing-risk
  probability) agrees with the sampled count within 4 sigma;
- the fold of two chunks is a sum except for the ten maxima.
A planted mistake in any one record breaks one of these."""
import numpy as np
import pytest

from oracle import binding as ob
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig

MOLLER, BHABHA, BREMS, COMPTON, PHOTOELECTRIC, PAIR_N, PAIR_T, ANNIH = 2, 3, 4, 5, 6, 7, 8, 9
T_MOLLER, T_BREMS, T_SUBSTEPS, T_ION, T_BHABHA, T_COMPTON, T_PE, T_PAIR = 3, 4, 6, 8, 11, 13, 14, 15
ME = 510998.95


def kinetic(ux, uy, uz):
    c = 299792458.0
    u2 = ux * ux + uy * uy + uz * uz
    gamma = np.sqrt(1.0 + u2 / (c * c))
    return u2 * ME / (c * c * (gamma + 1.0))


@pytest.fixture(scope="module")
def c2():
    return wl.c2_fields()


@pytest.fixture(scope="module")
def electron_step(otables, c2):
    cfg = StepConfig(rng_seed=3, dt_s=1e-8)
    e = wl.c2_electrons(8000, c2, seed=2)
    return e, ob.electron_step(otables, c2, cfg, e.as_dict(), nthreads=8, sec_cap=1 << 20)


def test_electron_ledger_against_step_outputs(electron_step):
    e, o = electron_step
    L = ob.ledger(o["tallies"])
    t = o["tallies"]
    sec = o["secondaries"]
    mol = sec[sec["process"] == MOLLER]
    assert L["hard_moller_event_count"] == t[T_MOLLER] == len(mol) > 100
    assert L["hard_moller_event_weight"] == L["hard_moller_secondary_energy_weight_sum"] == mol["weight"].sum()
    assert L["hard_moller_secondary_energy_max_ev"] == mol["energy_eV"].max()
    assert np.isclose(L["hard_moller_secondary_energy_weighted_sum_ev"], (mol["weight"] * mol["energy_eV"]).sum(),
                      rtol=1e-12)
    assert L["direct_hard_moller_electron_weight"] == mol["weight"].sum()  # threshold above the cutoff
    assert L["electron_brems_photon_count"] == L["total_brems_photon_count"] == t[T_BREMS]
    tracked = sec[sec["process"] == BREMS]
    assert L["electron_brems_photon_energy_ev"] >= (tracked["weight"] * tracked["energy_eV"]).sum()
    assert L["max_transport_subcycles"] == o["substeps"].max() > 1
    k0 = kinetic(e.ux, e.uy, e.uz)
    moved = o["substeps"] > 0
    assert L["max_transport_electron_energy_ev"] == k0[moved].max()
    # ionization records: one flag each; Moller is the only kinetic one here
    assert L["transport_ionization_weight"] == pytest.approx(
        L["transport_low_secondary_weight"] + L["transport_kinetic_secondary_weight"], rel=1e-12)
    assert L["transport_kinetic_secondary_weight"] == len(mol)
    assert np.isclose(L["transport_ionization_weight"], t[T_ION], rtol=1e-12)
    # restricted collision loss is the ion-pair source loss
    assert L["ion_pair_source_loss_ev"] == L["restricted_collision_loss_ev"] > 0.0
    assert L["parent_drag_loss_ev"] == pytest.approx(L["restricted_collision_loss_ev"]
                                                     + L["photon_continuous_soft_brems_energy_ev"], rel=1e-9)
    # the expected count of the competing-risk draws against the sampled count
    exp = L["hard_moller_expected_event_weight"]
    assert abs(L["hard_moller_event_count"] - exp) < 4.0 * np.sqrt(exp), (L["hard_moller_event_count"], exp)
    # electrons only
    for k in ("photon_interaction_count", "positron_bhabha_event_count", "max_transport_positron_tau",
              "max_transport_photon_tau", "positron_transport_energy_loss_ev"):
        assert L[k] == 0.0
    # the maximum of a sum lies between the larger maximum and the sum of the maxima
    hm, br, tot = (L[k] for k in ("max_transport_hard_moller_tau", "max_transport_brems_tau",
                                  "max_transport_electron_total_competing_tau"))
    assert 0.0 < max(hm, br) <= tot <= (hm + br) * (1.0 + 1e-12)
    assert tot <= 0.05 * (1.0 + 1e-9)  # the guard keeps every substep under the optical-depth target


def test_positron_ledger_against_step_outputs(otables, c2):
    cfg = StepConfig(rng_seed=4, dt_s=1e-8)
    p = wl.c2_electrons(6000, c2, seed=3, first_id=1 << 40, charge_sign=1.0)
    o = ob.positron_step(otables, c2, cfg, p.as_dict(), nthreads=8, sec_cap=1 << 20)
    L = ob.ledger(o["tallies"])
    t = o["tallies"]
    sec = o["secondaries"]
    bh = sec[sec["process"] == BHABHA]
    assert L["positron_bhabha_event_count"] == t[T_BHABHA] == len(bh) > 10
    assert L["positron_bhabha_kinetic_secondary_weight"] == L["positron_bhabha_event_weight"] == bh["weight"].sum()
    assert np.isclose(L["positron_bhabha_secondary_energy_ev"], (bh["weight"] * bh["energy_eV"]).sum(), rtol=1e-12)
    assert L["positron_brems_event_count"] == L["total_brems_photon_count"] == t[T_BREMS]
    assert L["electron_brems_photon_count"] == 0.0 and L["hard_moller_event_count"] == 0.0
    # positrons record no loss partition; their loss is the transport loss
    assert L["parent_drag_loss_ev"] == 0.0 and L["explicit_hard_moller_energy_loss_ev"] == 0.0
    assert L["positron_transport_energy_loss_ev"] > 0.0
    assert L["max_transport_subcycles"] == o["substeps"].max()
    moved = o["substeps"] > 0
    assert L["max_transport_positron_energy_ev"] == kinetic(p.ux, p.uy, p.uz)[moved].max()
    assert L["max_transport_positron_tau"] > 0.0 and L["max_transport_hard_moller_tau"] == 0.0
    # Bhabha records a kinetic ionization but books a positive ion, not an ion pair
    assert L["transport_kinetic_secondary_weight"] == len(bh)
    assert np.isclose(L["transport_ionization_weight"] - len(bh), t[T_ION], rtol=1e-12)
    # positrons stopped in the chord annihilate at rest, thermalized
    annihilated_at_rest = int(((o["status"] == 6) & (o["substeps"] > 0)).sum())
    assert L["positron_thermalized_count"] <= int((o["status"] == 6).sum())
    assert L["positron_thermalized_count"] >= annihilated_at_rest - int((sec["process"] == ANNIH).sum() // 2)


def test_photon_ledger_against_step_outputs(otables, c2):
    dt = 1e-7
    cfg = StepConfig(rng_seed=5, dt_s=dt)
    g = wl.c2_photons(20000, c2, dt_s=dt, seed=4)
    p = g.as_dict()
    p.update(g.extra)
    o = ob.photon_step(otables, c2, cfg, p, nthreads=8, sec_cap=1 << 20)
    L = ob.ledger(o["tallies"])
    t = o["tallies"]
    sec = o["secondaries"]
    assert L["compton_interaction_count"] == t[T_COMPTON] > 100
    assert L["photoelectric_interaction_count"] == t[T_PE]
    assert L["pair_production_count"] == t[T_PAIR] == L["pair_nuclear_count"] + L["pair_triplet_count"] > 0
    assert L["photon_interaction_count"] == (L["compton_interaction_count"] + L["photoelectric_interaction_count"]
                                             + L["pair_production_count"])
    pair_pos = sec[np.isin(sec["process"], [PAIR_N, PAIR_T]) & (sec["species"] == 1)]
    assert L["pair_kinetic_positron_weight"] == L["photon_seeded_kinetic_positron_weight"] == len(pair_pos)
    assert L["photon_seeded_positron_weight"] == L["pair_production_weight"]
    ce = sec[(sec["process"] == COMPTON) & (sec["species"] == 0)]
    assert L["compton_kinetic_electron_weight"] == len(ce)
    assert L["compton_kinetic_electron_weight"] + L["compton_low_electron_weight"] == L["compton_interaction_weight"]
    assert L["photon_seeded_electron_weight"] == pytest.approx(
        L["photon_seeded_kinetic_electron_weight"] + L["photon_seeded_low_electron_weight"], rel=1e-15)
    # escaped photons carry their final energy out
    esc = o["status"] == 1
    e_out = np.sqrt(o["ux"] ** 2 + o["uy"] ** 2 + o["uz"] ** 2)
    assert np.isclose(L["photon_escaped_energy_ev"], (p["w"][esc] * e_out[esc]).sum(), rtol=1e-12)
    assert L["max_transport_subcycles"] == 1.0
    e_in = np.sqrt(p["ux"] ** 2 + p["uy"] ** 2 + p["uz"] ** 2)
    assert L["max_transport_photon_energy_seen_ev"] <= e_in.max()
    assert L["max_transport_photon_tau"] > 0.0
    # ion pairs: every photon interaction but the nuclear pair ionizes the
    # target once; the triplet recoil's ionization has no transport record
    assert np.isclose(L["transport_ionization_weight"] + L["pair_triplet_weight"], t[T_ION], rtol=1e-12)


def test_fold_sums_and_maxima():
    a = np.arange(ob.N_TALLIES, dtype=float)
    b = np.full(ob.N_TALLIES, 50.0)
    out = a.copy()
    for i, n in enumerate(ob.LEDGER_NAMES):
        if n in ob.LEDGER_MAX:
            out[ob.LEDGER_OFFSET + i] = max(a[ob.LEDGER_OFFSET + i], 50.0)
        else:
            out[ob.LEDGER_OFFSET + i] = a[ob.LEDGER_OFFSET + i] + 50.0
    out[:ob.LEDGER_OFFSET] = a[:ob.LEDGER_OFFSET] + 50.0
    assert len(ob.LEDGER_MAX) == 10
    # the drain folds its per-record chunks with the same rule
    got = ob.fold_tallies(a, b)
    assert np.array_equal(got, out)
