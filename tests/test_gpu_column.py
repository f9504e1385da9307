"""GPU run of the feedback column (column.FeedbackColumn): injection, the
full transport step, the Maxwell feedback and population control together.
The parts are each held to the oracle or the reference elsewhere; here the
loop's own bookkeeping is checked: every injected macroparticle is
accounted for, population control engages above 1.25 x target and brings the
species back to the band without touching the CIC charge it reports as
conserved, the deposit drives a scattered field, and no device error is
raised."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

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
