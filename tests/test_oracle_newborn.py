"""Pins the oracle's newborn residual advance and transient FIFO
(OV/RreaNewbornAdvance.cpp, OV/RreaSecondaryCommit.cpp:338-468): the
reference's own packing and demotion-gate cases
(native_amrex_rrea/src/RreaSecondaryGroupSmoke.cpp:766-885), straight-line
known answers in near-vacuum, at-birth dispositions, and the FIFO's lineage
and generation invariants over a multi-generation step.  CPU only."""
import ctypes as C

import numpy as np
import pytest

from oracle import binding as ob
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
from rrea_feedback_warpx_amrex_geant4_amd.config import RzFields, StepConfig

ME = 510998.95
W_AIR = 33.97
C_LIGHT = 299792458.0
INVALID = 2 ** 64


def test_transient_index_packing_pins():
    """RreaSecondaryGroupSmoke.cpp:766-828: the all-maximum value fits 52
    bits, every field separates, the live continuation cannot alias a first
    hazard, and every overflow is refused."""
    ok, vmax = ob.pack_transient_index(True, 1, 16383, 16383, 8191, 255)
    assert ok and vmax < 2 ** 52
    base = ob.pack_transient_index(True, 0, 2, 3, 4, 5)
    live_root = ob.pack_transient_index(False, 0, 1, 3, 4, 5)
    others = [ob.pack_transient_index(True, 2, 2, 3, 4, 5), ob.pack_transient_index(True, 0, 3, 3, 4, 5),
              ob.pack_transient_index(True, 0, 2, 4, 4, 5), ob.pack_transient_index(True, 0, 2, 3, 5, 5),
              ob.pack_transient_index(True, 0, 2, 3, 4, 6)]
    assert base[0] and live_root[0] and all(o[0] for o in others)
    assert live_root[1] != base[1] and all(o[1] != base[1] for o in others)
    assert live_root[1] >= 2 ** 14
    assert ob.live_hazard_index(0, 8191, 0) == (True, 8191)
    ok2, second = ob.live_hazard_index(0, 0, 1)
    assert ok2 and second >= 2 ** 14
    for args in ((True, 0, 16384, 0, 0, 0), (True, 0, 0, 16384, 0, 0), (True, 0, 0, 0, 8192, 0),
                 (True, 0, 0, 0, 0, 256), (True, 3, 0, 0, 0, 0)):
        assert not ob.pack_transient_index(*args)[0]
    assert not ob.live_hazard_index(3, 0, 0)[0]
    # the closed form of the layout
    assert base[1] == (1 << 51) | (0 << 49) | (2 << 35) | (3 << 21) | (4 << 8) | 5


@pytest.mark.parametrize("energy,field,demote", [(9.0e2, 1.0e6, True), (5.0e4, 6.0e4, True), (5.0e4, 8.0e4, False),
                                                 (2.0e5, 6.0e4, False)])
def test_state_policy_cases(energy, field, demote):
    """RreaSecondaryGroupSmoke.cpp:861-883 with cutoff 1 keV, threshold
    100 keV and drag 70 kV/m."""
    assert ob.electron_state_demotes(energy, 1.0e3, 1.0e5, field, 7.0e4) == demote


def test_elapsed_mapping():
    """RreaNewbornElapsedStepFraction: birth + (1 - birth) (completed + f) / n, capped at 1."""
    assert ob.newborn_elapsed(0.25, 0, 4, 0.0) == 0.25
    assert ob.newborn_elapsed(0.25, 2, 4, 0.5) == pytest.approx(0.25 + 0.75 * 2.5 / 4, abs=1e-15)
    assert ob.newborn_elapsed(0.25, 4, 4, 1.0) == 1.0
    assert ob.newborn_elapsed(0.25, 1, 4, -1.0) == pytest.approx(0.25 + 0.75 * 0.25, abs=1e-15)


def record(species, x, y, z, d, energy, birth, root_key=77, lineage=0, root_species=0, w=1.0, parent=0, process=2):
    r = np.zeros(1, dtype=ob.SECONDARY_DTYPE)
    d = np.asarray(d, dtype=float) / np.linalg.norm(d)
    r["parent_index"], r["species"], r["process"] = parent, species, process
    r["x"], r["y"], r["z"] = x, y, z
    r["dir_x"], r["dir_y"], r["dir_z"] = d
    r["weight"], r["energy_eV"], r["birth_fraction"] = w, energy, birth
    r["root_key"], r["lineage"], r["root_species"] = root_key, lineage, root_species
    return r


@pytest.fixture(scope="module")
def vacuum():
    # a near-vacuum column with no field: no interaction, no drag worth the name
    return RzFields.uniform(32, 64, 10.0, 10.0, density_ratio=1e-12)


def test_vacuum_straight_lines(otables, vacuum):
    """In near-vacuum a newborn flies the rest of the step in a straight line:
    an electron at v = u / gamma over (1 - f) dt, a photon at c; a record
    already at the end of the step is not advanced; the electron's net chord
    from its birth point to its end point is the only current deposited."""
    dt = 1e-9
    k = 1.0e6
    ini = np.concatenate([
        record(0, 30.0, 40.0, 100.0, [0.3, 0.4, 0.866], k, 0.25, root_key=5, lineage=0),
        record(2, 20.0, -10.0, 200.0, [0.0, 0.6, 0.8], 2.0e6, 0.5, root_key=5, lineage=1, process=4),
        record(0, 50.0, 0.0, 300.0, [1.0, 0.0, 0.0], k, 1.0, root_key=9, lineage=0),
    ])
    o = ob.newborn_drain(otables, vacuum, StepConfig(rng_seed=3, dt_s=dt), ini)
    assert o["code"] == 0
    assert list(o["status"]) == [0, 0, 0] and len(o["work"]) == 3
    s = o["survivors"]
    assert len(s) == 3 and np.all(s["birth_fraction"] == 1.0)
    u = wl.proper_velocity(np.array([k]))[0]
    v = u / np.sqrt(1.0 + (u / C_LIGHT) ** 2)
    for i, speed, f in ((0, v, 0.25), (1, C_LIGHT, 0.5)):
        d = np.array([ini[i]["dir_x"], ini[i]["dir_y"], ini[i]["dir_z"]])
        want = np.array([ini[i]["x"], ini[i]["y"], ini[i]["z"]]) + speed * (1.0 - f) * dt * d
        got = np.array([s[i]["x"], s[i]["y"], s[i]["z"]])
        assert np.abs(got - want).max() <= 1e-12 * 400.0, (i, got - want)
        assert np.allclose([s[i]["dir_x"], s[i]["dir_y"], s[i]["dir_z"]], d, atol=1e-9)
    assert s[0]["energy_eV"] == pytest.approx(k, rel=1e-9)
    assert s[1]["energy_eV"] == 2.0e6
    assert s[2] == ini[2] or (s[2]["x"] == ini[2]["x"] and s[2]["energy_eV"] == ini[2]["energy_eV"])
    # the electron's net chord, deposited with the padded grid's own routine
    nrp, nzp = vacuum.padded_shape()
    rho, jr, jz = np.zeros(nrp * nzp), np.zeros((nrp + 1) * nzp), np.zeros(nrp * (nzp + 1))
    ob.lib().orc_deposit_segment(nrp, nzp, vacuum.dr, vacuum.dz, 0.0, vacuum.z_lo - 2 * vacuum.dz,
                                 *(x.ctypes.data_as(ob.P_D) for x in (rho, jr, jz)), -1.0, dt,
                                 float(np.hypot(30.0, 40.0)), 100.0, float(np.hypot(s[0]["x"], s[0]["y"])),
                                 float(s[0]["z"]))
    assert np.array_equal(o["jr"], jr) and np.array_equal(o["jz"], jz) and np.array_equal(o["rho"], rho)
    t = o["tallies"]
    assert t[17] == 2 and t[18] == 2 and t[19] == 0 and t[20] == 0


def test_birth_dispositions(otables, vacuum):
    """A sub-cutoff electron newborn is demoted where it is born (K / W_air ion
    pairs, a low-energy source of its weight, no current); a sub-cutoff
    positron annihilates at rest there into two back-to-back 511 keV photons
    that inherit its root and birth time and then fly the rest of the step."""
    dt = 1e-9
    cfg = StepConfig(rng_seed=4, dt_s=dt)
    ini = np.concatenate([
        record(0, 30.0, 0.0, 100.0, [0, 0, 1], 5.0e3, 0.4, root_key=11, lineage=0, w=2.0),
        record(1, 60.0, 0.0, 150.0, [1, 0, 0], 5.0e3, 0.6, root_key=12, lineage=0, root_species=2, process=7),
    ])
    o = ob.newborn_drain(otables, vacuum, cfg, ini)
    assert o["code"] == 0
    w = o["work"]
    assert list(o["status"][:2]) == [2, 6]
    assert o["ion_pairs"][0] == pytest.approx(2.0 * 5.0e3 / W_AIR, rel=1e-14)
    assert o["tallies"][16] == 2.0
    assert not np.any(o["jr"]) and not np.any(o["jz"])
    kids = w[2:]
    assert len(kids) == 2 and np.all(kids["species"] == 2) and np.all(kids["process"] == 9)
    assert np.all(kids["generation"] == 1) and np.all(kids["parent_index"] == 1)
    assert np.all(kids["root_key"] == 12) and np.all(kids["root_species"] == 2)
    assert sorted(kids["lineage"]) == [1, 2]
    assert np.all(kids["birth_fraction"] == 0.6)
    assert np.allclose(kids["energy_eV"], ME, rtol=1e-12)
    d0 = np.array([kids[0]["dir_x"], kids[0]["dir_y"], kids[0]["dir_z"]])
    d1 = np.array([kids[1]["dir_x"], kids[1]["dir_y"], kids[1]["dir_z"]])
    assert np.allclose(d0, -d1, atol=1e-15)
    # both photons survive, a c * 0.4 dt flight from the annihilation point
    assert list(o["status"][2:]) == [0, 0]
    for s, kid in zip(o["survivors"], kids):
        want = np.array([kid["x"], kid["y"], kid["z"]]) + C_LIGHT * 0.4 * dt * np.array(
            [kid["dir_x"], kid["dir_y"], kid["dir_z"]])
        assert np.abs(np.array([s["x"], s["y"], s["z"]]) - want).max() < 1e-9


def test_escape(otables, vacuum):
    """A newborn heading out through r_hi within its residual time escapes and
    books its net chord from the birth point to the crossing."""
    dt = 1e-6
    ini = record(0, vacuum.r_hi - 1.0, 0.0, 300.0, [1, 0, 0], 1.0e6, 0.5, root_key=3)
    o = ob.newborn_drain(otables, vacuum, StepConfig(rng_seed=5, dt_s=dt), ini)
    assert o["code"] == 0 and list(o["status"]) == [1] and len(o["survivors"]) == 0
    assert o["tallies"][19] == 1
    assert np.any(o["jr"])


@pytest.fixture(scope="module")
def long_step(otables):
    """The step's staged secondaries of a C2 sample at dt = 100 ns (Apply's
    order), and the drained FIFO."""
    fields = wl.c2_fields()
    cfg = StepConfig(rng_seed=5, step=0, dt_s=1e-7)
    e = wl.c2_electrons(1500, fields, seed=2).as_dict()
    g = wl.c2_photons(700, fields, dt_s=1e-7, seed=4, first_id=1 << 41)
    gp = g.as_dict()
    gp.update(g.extra)
    p = wl.c2_electrons(300, fields, seed=3, first_id=1 << 40, charge_sign=1.0).as_dict()
    staged = []
    for name, t in (("electron", e), ("photon", gp), ("positron", p)):
        if name == "photon":
            o = ob.photon_step(otables, fields, cfg, t, nthreads=8, sec_cap=1 << 18)
        else:
            o = ob.charged_step(name, otables, fields, cfg, t, nthreads=8, sec_cap=1 << 18)
        assert o["code"] == 0
        staged.append(o["secondaries"])
    ini = np.concatenate(staged)
    return fields, cfg, ini, ob.newborn_drain(otables, fields, cfg, ini)


def test_fifo_invariants(long_step):
    """Initial entries are kept as staged; every later entry is a child of an
    earlier one, one generation on, with its root; each root's lineages are
    exactly 0..m-1 (the spawn counter, then the FIFO's next free lineage);
    the survivors are the status-0 entries in FIFO order at birth fraction 1."""
    fields, cfg, ini, o = long_step
    assert o["code"] == 0
    w = o["work"]
    n0 = len(ini)
    assert len(w) > n0 and w["generation"].max() >= 3
    assert (w[:n0] == ini).all()
    kids = w[n0:]
    par = kids["parent_index"]
    assert np.all(par < np.arange(n0, len(w)))
    assert np.all(kids["generation"] == w["generation"][par] + 1)
    assert np.all(kids["root_key"] == w["root_key"][par]) and np.all(kids["root_species"] == w["root_species"][par])
    assert np.all(kids["birth_fraction"] >= w["birth_fraction"][par]) and np.all(kids["birth_fraction"] <= 1.0)
    roots = {}
    for r in w:
        roots.setdefault((int(r["root_key"]), int(r["root_species"])), []).append(int(r["lineage"]))
    for lins in roots.values():
        assert sorted(lins) == list(range(len(lins)))
    # children of one root are allocated in FIFO order
    for key, lins in roots.items():
        sel = (w["root_key"] == key[0]) & (w["root_species"] == key[1])
        assert np.all(np.diff(w["lineage"][sel][w["generation"][sel] > 0]) > 0)
    st = o["status"]
    assert set(np.unique(st)) <= {0, 1, 2, 6, 7}
    s = o["survivors"]
    assert len(s) == int((st == 0).sum()) and np.all(s["birth_fraction"] == 1.0)
    alive = np.flatnonzero(st == 0)
    for f in ("root_key", "lineage", "species", "generation"):
        assert np.array_equal(s[f], w[f][alive])
    rr = np.hypot(s["x"], s["y"])
    assert np.all(rr < fields.r_hi) and np.all((s["z"] >= fields.z_lo) & (s["z"] < fields.z_hi))
    charged = s["species"] != 2
    assert np.all(s["energy_eV"][charged] >= cfg.low_energy_cutoff_eV)
    t = o["tallies"]
    advanced = int((w["birth_fraction"] < 1.0).sum())
    assert t[17] == advanced
    assert t[18] + t[19] + t[20] == advanced


def test_drain_is_deterministic(otables, long_step):
    fields, cfg, ini, o = long_step
    o2 = ob.newborn_drain(otables, fields, cfg, ini)
    assert (o2["work"] == o["work"]).all() and np.array_equal(o2["jz"], o["jz"])
