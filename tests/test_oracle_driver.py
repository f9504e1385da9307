"""The oracle's multi-step loop (oracle/driver.py) on the CPU: the commit
rules of CommitLocalParticleBatch (RreaCouplingSources.cpp:280-307) on hand
cases, and the loop's bookkeeping (ids unique and consecutive in commit
order, committed counts equal to what was appended, invalidated particles
removed, photon push at c)."""
import numpy as np
import pytest

from oracle.driver import ELECTRON, PHOTON, POSITRON, OracleSimulation, commit_arrays
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig
from rrea_feedback_warpx_amrex_geant4_amd.transport import SECONDARY_DTYPE

C_LIGHT = 299792458.0


def record(species, d, e, w=1.0, pos=(3.0, 4.0, 7.0)):
    r = np.zeros(1, dtype=SECONDARY_DTYPE)
    r["species"] = species
    r["x"], r["y"], r["z"] = pos
    r["dir_x"], r["dir_y"], r["dir_z"] = d
    r["weight"] = w
    r["energy_eV"] = e
    return r


def test_commit_rules():
    sv = np.concatenate([
        record(ELECTRON, (0.0, 0.0, 0.0), 1e6),        # zero direction: u along +z
        record(ELECTRON, (0.0, 3.0, 4.0), 1e6),        # |dir| = 5: normalised
        record(ELECTRON, (0.1, 0.2, 0.3), 0.0),        # no energy: the raw direction is the momentum
        record(ELECTRON, (0.0, 0.0, 1.0), 1e6, w=0.0),  # zero weight: skipped
        record(PHOTON, (0.0, 0.0, 0.0), 2e6),          # zero direction: p = E along +z
        record(PHOTON, (0.0, 0.0, 2.0), 2e6),          # p = E dir / |dir|
        record(PHOTON, (1.0, 0.0, 0.0), 0.0),          # no energy: skipped
        record(POSITRON, (1.0, 0.0, 0.0), 5e5, w=-1.0),  # negative weight: skipped
    ])
    e = commit_arrays(sv, ELECTRON)
    g = commit_arrays(sv, PHOTON)
    p = commit_arrays(sv, POSITRON)
    assert len(e["r"]) == 3 and len(g["r"]) == 2 and len(p["r"]) == 0
    gamma = 1.0 + 1e6 / 510998.95
    u = C_LIGHT * np.sqrt(gamma * gamma - 1.0)
    assert np.allclose([e["ux"][0], e["uy"][0], e["uz"][0]], [0.0, 0.0, u], rtol=1e-15)
    assert np.allclose([e["ux"][1], e["uy"][1], e["uz"][1]], [0.0, 0.6 * u, 0.8 * u], rtol=1e-15)
    assert [e["ux"][2], e["uy"][2], e["uz"][2]] == [0.1, 0.2, 0.3]
    assert [g["ux"][0], g["uy"][0], g["uz"][0]] == [0.0, 0.0, 2e6]
    assert [g["ux"][1], g["uy"][1], g["uz"][1]] == [0.0, 0.0, 2e6]
    assert np.allclose(e["r"], 5.0) and np.allclose(e["theta"], np.arctan2(4.0, 3.0)) and np.all(e["z"] == 7.0)


def test_loop_bookkeeping(otables):
    fields, e = wl.c1_plain(300)
    sim = OracleSimulation(otables, fields, StepConfig(rng_seed=3, dt_s=1e-10), e, next_id=10 ** 9, nthreads=4)
    for _ in range(30):
        sim.step()
        for s in sim.sp:
            assert np.all(sim.sp[s]["id"] != np.iinfo(np.int64).min)
    ids = np.concatenate([sim.sp[s]["id"] for s in (ELECTRON, PHOTON, POSITRON)])
    assert len(np.unique(ids)) == len(ids)
    assert sim.next_id == 10 ** 9 + int(sim.created.sum())
    assert sim.created[ELECTRON] > 0
    new = np.sort(ids[ids >= 10 ** 9])
    # the survivors of every step are still alive or were removed; none is duplicated
    assert np.all(np.diff(new) > 0)


def test_photon_push_is_straight_at_c(otables):
    fields, e = wl.c1_plain(1)
    g = wl.c2_photons(50, fields, seed=9, first_id=100)
    sim = OracleSimulation(otables, fields, StepConfig(rng_seed=3, dt_s=1e-12), e, photons=g)
    d = sim.sp[PHOTON]
    x0 = d["r"] * np.cos(d["theta"]), d["r"] * np.sin(d["theta"]), d["z"].copy()
    sim._push_photons(1e-12)
    x1 = d["r"] * np.cos(d["theta"]), d["r"] * np.sin(d["theta"]), d["z"]
    step = np.sqrt(sum((a - b) ** 2 for a, b in zip(x1, x0)))
    assert np.allclose(step, C_LIGHT * 1e-12, rtol=1e-9)
    assert np.array_equal(d["prev_z"], x0[2])
