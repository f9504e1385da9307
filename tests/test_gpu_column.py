"""GPU runs of the feedback column (column.FeedbackColumn, BatchedColumn):
injection, the full transport step, the Maxwell feedback and population
control together.

- One column instance against the oracle's loop over 20 steps: the same
  injections, the oracle's step, drain and commit, and the oracle's current
  conversion and Maxwell step driving the next step's field.
- The loop's own bookkeeping: every injected macroparticle is accounted for,
  population control engages above 1.25 x target and brings the species
  back to the band, the deposit drives a scattered field, and no device
  error is raised.
- The configuration's scale: 16 instances holding 5e7 live
  macroparticles."""
import dataclasses

import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import maxwell as OM  # noqa: E402
from oracle.driver import ELECTRON, PHOTON, POSITRON, OracleSimulation  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def test_feedback_column_runs_and_controls_its_population(tables):
    from rrea_feedback_warpx_amrex_geant4_amd.column import FeedbackColumn, column_fields
    fields = column_fields(nr=64, nz=256)
    col = FeedbackColumn(tables, fields, dt_s=1e-9, rate_per_s=1e14, inject_macros=128, targets=(1500, 200),
                         control_every=10, capacity=1 << 16)
    print("case feedback column: 64 x 256 grid, 128 macros/step, 40 steps", flush=True)
    history = []
    for _ in range(4):
        col.run(10)
        history.append(col.observe())
        print(history[-1], flush=True)
    obs = history[-1]
    assert obs["injected"] == 40 * 128
    assert obs["electrons"] > 0
    # control engaged once the electrons passed 1.25 x target, and after each
    # control pass the live count sits at or below the band
    assert obs["removed_electrons"] > 0
    for h in history:
        if h["removed_electrons"] > 0:
            assert h["electrons"] <= 1500 + 1500 // 4 + 1
    # the column's current built a scattered field
    assert obs["max_scattered_ez_v_per_m"] > 0.0
    # weights stay positive and finite
    e = col.sim.electrons
    w = e.w[: e.n][e.id[: e.n] != -(2 ** 63)].cpu().numpy()
    assert np.all(np.isfinite(w)) and np.all(w > 0)


def test_batched_column_steps_every_instance(tables):
    """Four columns of differing E_z peaks in one batch: every instance gets
    its injections, keeps its particles (instance indices survive commit,
    compaction and control), is thinned on its own mesh above its band, and
    builds its own scattered field."""
    from rrea_feedback_warpx_amrex_geant4_amd.column import BatchedColumn
    peaks = [3.0e5, 4.0e5, 5.0e5, 6.0e5]
    col = BatchedColumn(tables, peaks, nr=64, nz=256, inject_macros=96, targets=(800, 100), control_every=10,
                        capacity=1 << 17)
    print("case batched column: 4 instances, 64 x 256 grids, 96 macros/instance/step, 30 steps", flush=True)
    for _ in range(3):
        col.run(10)
        obs = col.observe()
        print(obs, flush=True)
    assert obs["injected"] == 30 * 96 * 4
    assert all(n > 0 for n in obs["electrons"])
    assert all(n <= 800 + 800 // 4 + 1 for n in obs["electrons"])
    assert obs["removed_electrons"] > 0
    assert all(f > 0.0 for f in obs["max_scattered_ez_v_per_m"])
    e = col.stores[0]
    inst = e.inst[: e.n].cpu().numpy()
    assert inst.min() >= 0 and inst.max() < 4
    w = e.w[: e.n].cpu().numpy()
    assert np.all(np.isfinite(w)) and np.all(w > 0)
    col.close()


def test_final_window_sensitivity(tables):
    """The adjoint of the column's final-window current w.r.t. its Maxwell
    field: finite, non-zero where the electrons drift, and equal to a
    Richardson-extrapolated central difference of the same drift along a
    random field direction.  This checks the wiring (a wrong field, sign or
    weight gives an O(1) mismatch); the difference of a kinked loss (cell
    crossings, the cutoff) over every face at once carries ~1e-4 noise, so
    the bar is 2e-3.  The drift adjoint itself is pinned face by face against
    an FP64 reference gradient to 3e-15 in test_gpu_adjoint.py."""
    from rrea_feedback_warpx_amrex_geant4_amd.adjoint import DriftConfig, DriftProblem, drift_current
    from rrea_feedback_warpx_amrex_geant4_amd.column import FeedbackColumn, column_fields
    fields = column_fields(nr=64, nz=256)
    col = FeedbackColumn(tables, fields, dt_s=1e-9, inject_macros=256, targets=(1 << 20, 1 << 18), capacity=1 << 16)
    col.run(5)
    print("case final-window sensitivity: 5 column steps, then 4 drift steps", flush=True)
    nrp, nzp = fields.padded_shape()
    g = torch.Generator(device="cpu").manual_seed(3)
    wts = torch.randn(nrp * (nzp + 1), generator=g, dtype=torch.float64).cuda()
    L, ger, gez = col.final_window_sensitivity(wts, steps=4, checkpoint_every=2, dt_s=1e-10)
    assert torch.isfinite(ger).all() and torch.isfinite(gez).all()
    assert float(gez.abs().max()) > 0.0
    # directional derivative check along a random direction of E_z
    import dataclasses
    f = dataclasses.replace(fields, er=col.maxwell.er.cpu().numpy(), ez=col.maxwell.ez.cpu().numpy(),
                            btheta=col.maxwell.bt.cpu().numpy())
    e = col.sim.electrons
    h = e.host()
    live = h["id"] != np.iinfo(np.int64).min
    from rrea_feedback_warpx_amrex_geant4_amd.config import ParticleTile
    parts = ParticleTile(*(h[k][live] for k in ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu")))
    prob = DriftProblem(tables, f, parts, DriftConfig(steps=4, dt_s=1e-10, checkpoint_every=2))
    d = torch.randn(prob.ez0.numel(), generator=g, dtype=torch.float64).cuda()
    h_step = 1e-3 * float(prob.ez0.abs().max() + 1.0) + 1.0

    def loss(ez):
        with torch.no_grad():
            return float((drift_current(prob.er0, ez, prob)[1] * wts).sum())
    def central(hh):
        return (loss(prob.ez0 + hh * d) - loss(prob.ez0 - hh * d)) / (2 * hh)
    fd = (4.0 * central(h_step) - central(2.0 * h_step)) / 3.0
    ad = float((gez * d).sum())
    assert abs(fd - ad) <= 2e-3 * abs(ad), (fd, ad)
    print(f"final-window sensitivity: L = {float(L):.4e}, directional adjoint {ad:.6e} vs difference {fd:.6e}",
          flush=True)


def test_column_matches_oracle_loop(tables, otables):
    """One column of the feedback configuration (256 x 1024 mesh over 4-14 km
    of exponential atmosphere, E_z peak 450 kV/m, 64 macroparticles injected
    per 1 ns step) over 20 steps, against the oracle: the same injections
    (the column's draws replayed from a generator of the same seed), the
    oracle's loop (photon push, the three species steps, the drain, the
    commit; oracle/driver.py) and the oracle's current conversion and TM
    Maxwell step (oracle/maxwell.py) feeding each step's deposit back into
    the next step's field.  Population control is off (its plan is held to
    the reference's own in test_gpu_population.py).  At the end: the same
    particles, ids and order per species, states to 1e-9 of |u| and 1e-9 m,
    and the Maxwell fields to 1e-9 of their maximum."""
    from rrea_feedback_warpx_amrex_geant4_amd.column import FeedbackColumn, column_fields, injection_tile
    from rrea_feedback_warpx_amrex_geant4_amd.config import ParticleTile
    fields = column_fields()
    n_inj, seed, steps = 64, 5, 20
    col = FeedbackColumn(tables, fields, dt_s=1e-9, inject_macros=n_inj, control_every=0, capacity=1 << 16,
                         seed=seed)
    print(f"case column vs oracle loop: 256 x 1024, {n_inj} macros/step, {steps} steps of 1 ns", flush=True)
    empty = ParticleTile(*(np.zeros(0) for _ in range(7)), id=np.zeros(0, dtype=np.int64),
                         cpu=np.zeros(0, dtype=np.int32))
    osim = OracleSimulation(otables, fields, col.cfg, empty, next_id=1)
    ofield = OM.Field(fields.nr, fields.nz, fields.dr, fields.dz, fields.er, fields.ez, fields.btheta)
    rng = np.random.default_rng(seed)
    dt = col.cfg.dt_s
    for k in range(steps):
        col.step()
        # the oracle: the same injection, then its loop on the field of the
        # previous step's feedback
        t = injection_tile(rng, n_inj, col.rate * dt / n_inj, col.inject_radius, col.inject_z, osim.next_id,
                           col.index, col.e_min, col.e_max)
        osim.next_id += n_inj
        e = osim.sp[ELECTRON]
        for q in ("r", "z", "theta", "ux", "uy", "uz", "w", "id", "cpu"):
            e[q] = np.concatenate([e[q], getattr(t, q)])
        osim.fields = dataclasses.replace(fields, er=ofield.er.copy(), ez=ofield.ez.copy(), btheta=ofield.bt.copy())
        osim.step()
        jr, jz = OM.current_to_physical(fields.nr, fields.nz, fields.dr, fields.dz, osim.jr, osim.jz)
        ofield.apply(0, dt, jr=jr, jz=jz)
    torch.cuda.synchronize()
    assert np.array_equal(col.sim.created, osim.created), (col.sim.created, osim.created)
    assert col.sim.created[ELECTRON] + col.sim.created[PHOTON] > 0
    for s, store in ((ELECTRON, col.sim.electrons), (PHOTON, col.sim.photons), (POSITRON, col.sim.positrons)):
        gd = store.host()
        od = osim.sp[s]
        assert np.array_equal(gd["id"], od["id"]), s
        umag = np.sqrt(od["ux"] ** 2 + od["uy"] ** 2 + od["uz"] ** 2)
        du = np.sqrt((gd["ux"] - od["ux"]) ** 2 + (gd["uy"] - od["uy"]) ** 2 + (gd["uz"] - od["uz"]) ** 2)
        assert np.all(du <= 1e-9 * np.maximum(umag, 1.0)), (s, float(np.max(du / np.maximum(umag, 1.0))))
        xg = gd["r"] * np.cos(gd["theta"]), gd["r"] * np.sin(gd["theta"])
        xo = od["r"] * np.cos(od["theta"]), od["r"] * np.sin(od["theta"])
        dx = np.sqrt((xg[0] - xo[0]) ** 2 + (xg[1] - xo[1]) ** 2 + (gd["z"] - od["z"]) ** 2)
        assert np.all(dx <= 1e-9), (s, float(dx.max()))
    for q in ("er", "ez", "bt"):
        want = getattr(ofield, q)
        scale = max(float(np.abs(want).max()), 1e-300)
        got = getattr(col.maxwell, q).cpu().numpy()
        assert np.allclose(got, want, rtol=0, atol=1e-9 * scale), (q, float(np.abs(got - want).max() / scale))
    assert float(np.abs(ofield.ez).max()) > 0.0
    print(f"column vs oracle: {steps} steps, live e- {col.sim.electrons.live()}, gamma {col.sim.photons.live()}, "
          f"e+ {col.sim.positrons.live()}, created {col.sim.created.tolist()}, scattered |Ez| up to "
          f"{float(np.abs(ofield.ez).max()):.3e} V/m", flush=True)


def test_batched_column_at_configuration_scale(tables):
    """The feedback configuration at its size: 16 columns (E_z peaks 300-600
    kV/m, 256 x 1024 meshes over 4-14 km of exponential atmosphere) seeded
    with 3,125,000 electrons each -- 5e7 live macroparticles -- then three
    steps of injection (4,096 per instance), the batched species steps, the
    drain, the commit, compaction and each instance's Maxwell feedback, then
    one control pass at 2e6 electrons per instance.  No device error; every
    instance keeps its particles and builds its own field; at least 5e7 live
    before control; after it, every instance at or below its band and every
    weight positive and finite."""
    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    from rrea_feedback_warpx_amrex_geant4_amd.column import BatchedColumn
    peaks = np.linspace(3.0e5, 6.0e5, 16)
    per = 3_125_000
    col = BatchedColumn(tables, peaks, inject_macros=4096, targets=(2_000_000, 1 << 18), control_every=0,
                        capacity=(16 * per) + (1 << 22))
    print(f"case batched column at scale: 16 instances x {per} seeded electrons, 3 steps, one control pass",
          flush=True)
    col.seed_population(per)
    assert col.live() == 16 * per
    for k in range(3):
        col.step()
        print(f"batched column at scale: step {k}, live {col.live()}", flush=True)
    live = col.live()
    assert live >= 50_000_000
    obs = col.observe()
    assert all(n >= per for n in obs["electrons"])
    assert all(f > 0.0 for f in obs["max_scattered_ez_v_per_m"])
    col.control()
    obs = col.observe()
    print(f"batched column at scale: live {live} before control, electrons per instance after {obs['electrons']}, "
          f"removed {obs['removed_electrons']}", flush=True)
    assert obs["removed_electrons"] > 0
    assert all(n <= 2_000_000 + 2_000_000 // 4 + 1 for n in obs["electrons"])
    e = col.stores[0]
    valid = e.id[: e.n] != _lib.INVALID_ID
    w = e.w[: e.n][valid]
    assert bool(torch.isfinite(w).all()) and bool((w > 0).all())
    inst = e.inst[: e.n]
    assert int(inst.min()) >= 0 and int(inst.max()) < 16
    col.out[0].check_error()
    col.close()
