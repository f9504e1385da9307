"""GPU checks of the drift segment and its adjoint (rrea_drift_forward /
rrea_drift_adjoint, adjoint.DriftCurrent):

- the drift forward is the transport step's one-substep branch: on a
  one-step sample, every particle the step transported without an event in
  one substep has the same position bit for bit and the same energy, and
  most (those with no elastic scatter) the same momentum bit for bit;
- the adjoint gradient equals the FP64 reference gradient (the drift step
  restated in torch float64, differentiated by autograd) to 1e-9;
- on the differentiation configuration (2^20 electrons, 64 steps, 128 x 512
  grid, checkpoints every 8) it agrees with central finite differences at 32
  seeded grid points;
- the checkpoint interval does not change the gradient, and the autograd
  binding returns what the adjoint does."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from drift_reference import DriftReference  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def kinetic(ux, uy, uz):
    c = 299792458.0
    u2 = ux * ux + uy * uy + uz * uz
    return u2 * 510998.95 / ((c * c) * (np.sqrt(1.0 + u2 / (c * c)) + 1.0))


def loss_weights(prob, seed):
    g = torch.Generator(device="cpu").manual_seed(seed)
    gjr = torch.randn(prob.jr_size, generator=g, dtype=torch.float64)
    gjz = torch.randn(prob.jz_size, generator=g, dtype=torch.float64)
    return gjr, gjz


def test_drift_is_the_one_substep_branch(tables):
    from rrea_feedback_warpx_amrex_geant4_amd.adjoint import DriftConfig, DriftProblem
    from rrea_feedback_warpx_amrex_geant4_amd.transport import DeviceTile, RreaParticleInteraction, StepOutputs, Coupling
    fields = wl.c2_fields()
    n = 1 << 15
    host = wl.c2_electrons(n, fields, seed=21)
    print(f"case drift vs step: {n} C2 electrons, one 1 ps step", flush=True)
    cfg = StepConfig(rng_seed=77, dt_s=1e-12)
    dev = DeviceTile.from_host(host)
    out = StepOutputs(n, fields, secondary_capacity=1 << 14)
    RreaParticleInteraction.apply_electrons(tables, Coupling(fields, tables), dev, cfg, out)
    prob = DriftProblem(tables, fields, host, DriftConfig(steps=1, dt_s=1e-12, checkpoint_every=1))
    res = prob.forward(prob.er0, prob.ez0, checkpoints=False)
    torch.cuda.synchronize()
    st = out.status.cpu().numpy()
    sub = out.substeps.cpu().numpy()
    parents = np.zeros(n, dtype=bool)
    parents[out.secondaries_host()["parent_index"]] = True
    plain = (st == 0) & (sub == 1) & ~parents
    alive = res["alive"].cpu().numpy().astype(bool)
    assert np.all(alive[plain]), "the drift ended a particle the step transported in one substep"
    a = {k: getattr(res["tile"], k).cpu().numpy()[plain] for k in ("r", "z", "theta", "ux", "uy", "uz")}
    b = {k: getattr(dev, k).cpu().numpy()[plain] for k in ("r", "z", "theta", "ux", "uy", "uz")}
    for k in ("r", "z", "theta"):
        assert np.array_equal(a[k], b[k]), k
    ka, kb = kinetic(a["ux"], a["uy"], a["uz"]), kinetic(b["ux"], b["uy"], b["uz"])
    # |u| = u(K) |d|: a scattered direction is renormalised, a few ulps from 1
    assert np.all(np.abs(ka - kb) <= 2e-15 * kb), float(np.max(np.abs(ka - kb) / kb))
    same = (a["ux"] == b["ux"]) & (a["uy"] == b["uy"]) & (a["uz"] == b["uz"])
    assert same.mean() > 0.5, same.mean()
    print(f"drift vs step: {plain.sum()} one-substep event-free particles, positions and energies equal, "
          f"{same.mean():.3f} of them (no elastic scatter) bit-identical in u", flush=True)


def small_problem(tables, steps=6, every=4, n=16, seed=31):
    from rrea_feedback_warpx_amrex_geant4_amd.adjoint import DriftConfig, DriftProblem
    fields, parts = wl.c4_drift(n, nr=16, nz=32, cell_m=0.25, seed=seed)
    cfg = DriftConfig(steps=steps, dt_s=1e-10, checkpoint_every=every)
    return fields, parts, DriftProblem(tables, fields, parts, cfg)


def test_adjoint_matches_fp64_reference_gradient(tables, bundle):
    fields, parts, prob = small_problem(tables)
    print("case adjoint vs torch reference: 16 electrons, 6 steps of 100 ps, 16 x 32 grid", flush=True)
    gjr, gjz = loss_weights(prob, 5)
    out = prob.forward(prob.er0, prob.ez0, gjr.cuda(), gjz.cuda())
    g_er, g_ez = prob.adjoint(prob.er0, prob.ez0, out["ck"], out["ck_alive"], gjr.cuda(), gjz.cuda())
    torch.cuda.synchronize()
    assert bool(out["alive"].all())
    ref = DriftReference(fields, bundle, dt=1e-10)
    er = torch.tensor(fields.er, dtype=torch.float64, requires_grad=True)
    ez = torch.tensor(fields.ez, dtype=torch.float64, requires_grad=True)
    L = ref.loss(er, ez, parts, 6, gjr, gjz)
    L.backward()
    l_gpu = float(out["loss"].sum().item())
    l_ref = float(L.detach())
    assert abs(l_gpu - l_ref) <= 1e-12 * abs(l_ref), (l_gpu, l_ref)
    for name, a, b in (("Er", g_er.cpu(), er.grad), ("Ez", g_ez.cpu(), ez.grad)):
        scale = float(b.abs().max())
        assert scale > 0.0
        err = float((a - b).abs().max())
        assert err <= 1e-9 * scale, (name, err, scale)
        print(f"adjoint vs reference {name}: max |diff| {err:.3e} of max |grad| {scale:.3e}, "
              f"{int((b != 0).sum())} nonzero faces", flush=True)


def test_checkpoint_interval_does_not_change_the_gradient(tables):
    grads = []
    for every in (1, 5, 16):
        fields, parts, prob = small_problem(tables, steps=16, every=every, n=64, seed=33)
        gjr, gjz = loss_weights(prob, 6)
        out = prob.forward(prob.er0, prob.ez0)
        grads.append(prob.adjoint(prob.er0, prob.ez0, out["ck"], out["ck_alive"], gjr.cuda(), gjz.cuda()))
    for g in grads[1:]:
        for a, b in zip(g, grads[0]):
            assert torch.allclose(a, b, rtol=1e-12, atol=1e-12 * float(b.abs().max()))


def test_autograd_binding(tables):
    from rrea_feedback_warpx_amrex_geant4_amd.adjoint import drift_current
    fields, parts, prob = small_problem(tables, steps=10, every=3, n=128, seed=34)
    gjr, gjz = loss_weights(prob, 7)
    er = prob.er0.clone().requires_grad_(True)
    ez = prob.ez0.clone().requires_grad_(True)
    jr, jz = drift_current(er, ez, prob)
    L = (jr * gjr.cuda()).sum() + (jz * gjz.cuda()).sum()
    L.backward()
    out = prob.forward(prob.er0, prob.ez0, gjr.cuda(), gjz.cuda())
    g_er, g_ez = prob.adjoint(prob.er0, prob.ez0, out["ck"], out["ck_alive"], gjr.cuda(), gjz.cuda())
    assert torch.allclose(er.grad, g_er, rtol=1e-12, atol=1e-12 * float(g_er.abs().max()))
    assert torch.allclose(ez.grad, g_ez, rtol=1e-12, atol=1e-12 * float(g_ez.abs().max()))
    assert abs(float(L) - float(out["loss"].sum())) <= 1e-9 * abs(float(L))


def test_autograd_binding_gridded_fields(tables):
    """Fields passed as (nz, nr + 1) / (nz + 1, nr) grids: the gradients come
    back in the same shapes, equal to the flat call's up to the summation
    order of the atomic scatter."""
    from rrea_feedback_warpx_amrex_geant4_amd.adjoint import drift_current
    fields, parts, prob = small_problem(tables, steps=6, every=2, n=128, seed=35)
    gjr, gjz = loss_weights(prob, 8)
    nr, nz = fields.nr, fields.nz
    er = prob.er0.clone().view(nz, nr + 1).requires_grad_(True)
    ez = prob.ez0.clone().view(nz + 1, nr).requires_grad_(True)
    jr, jz = drift_current(er, ez, prob)
    ((jr * gjr.cuda()).sum() + (jz * gjz.cuda()).sum()).backward()
    assert er.grad.shape == er.shape and ez.grad.shape == ez.shape
    er1 = prob.er0.clone().requires_grad_(True)
    ez1 = prob.ez0.clone().requires_grad_(True)
    jr1, jz1 = drift_current(er1, ez1, prob)
    ((jr1 * gjr.cuda()).sum() + (jz1 * gjz.cuda()).sum()).backward()
    # the field adjoints are scattered with FP64 atomics: equal to summation order
    for a, b in ((er.grad.reshape(-1), er1.grad), (ez.grad.reshape(-1), ez1.grad)):
        assert torch.allclose(a, b, rtol=0.0, atol=1e-12 * float(b.abs().max()))


def test_adjoint_matches_finite_differences(tables):
    """The differentiation configuration: 2^20 electrons, 64 drift steps of
    100 ps, 128 x 512 grid, checkpoints every 8; L = sum(J_z W) with
    W ~ N(0, 1); finite differences at 32 seeded faces (16 of Er, 16 of Ez)
    where the gradient is not zero, to 1e-6.  A field change moves the
    particles by ~ dt^2 x change, and a difference quotient's noise is the
    positions' rounding (~1e-15 m at r ~ 10 m) over that displacement: a step
    of 1e-6 of the field resolves the gradient only to ~1e-5 here (and not at
    all at dt = 1 ps).  So the steps are h = 1e-3 and 2e-3 of the ambient
    field, and the two central differences are Richardson-extrapolated,
    (4 D(h) - D(2h)) / 3, which cancels their h^2 truncation term."""
    from rrea_feedback_warpx_amrex_geant4_amd.adjoint import DriftConfig, DriftProblem
    fields, parts = wl.c4_drift()
    prob = DriftProblem(tables, fields, parts, DriftConfig(steps=64, dt_s=1e-10, checkpoint_every=8))
    print("case adjoint vs finite differences: 2^20 electrons, 64 steps of 100 ps, 128 x 512", flush=True)
    g = torch.Generator(device="cpu").manual_seed(11)
    gjr = torch.zeros(prob.jr_size, dtype=torch.float64).cuda()
    gjz = torch.randn(prob.jz_size, generator=g, dtype=torch.float64).cuda()
    out = prob.forward(prob.er0, prob.ez0, gjr, gjz)
    assert float(out["alive"].float().mean()) > 0.99
    g_er, g_ez = prob.adjoint(prob.er0, prob.ez0, out["ck"], out["ck_alive"], gjr, gjz)
    torch.cuda.synchronize()
    rng = np.random.default_rng(12)
    worst = 0.0

    def central(name, base, f, h):
        ls = []
        for sgn in (1.0, -1.0):
            e = base.clone()
            e[f] += sgn * h
            er, ez = (e, prob.ez0) if name == "Er" else (prob.er0, e)
            ls.append(prob.forward(er, ez, gjr, gjz, checkpoints=False, deposit=False)["loss"])
        # per-particle differences: particles the face does not reach cancel exactly
        return float((ls[0] - ls[1]).sum()) / (2.0 * h)

    h = 1e-3 * 4.0e5
    for name, grad, base in (("Er", g_er, prob.er0), ("Ez", g_ez, prob.ez0)):
        cand = torch.nonzero(grad != 0).flatten().cpu().numpy()
        assert len(cand) > 1000
        for f in rng.choice(cand, 16, replace=False):
            fd = (4.0 * central(name, base, f, h) - central(name, base, f, 2.0 * h)) / 3.0
            ad = float(grad[f])
            rel = abs(fd - ad) / max(abs(ad), 1e-300)
            worst = max(worst, rel)
            assert rel <= 1e-6, (name, int(f), fd, ad, rel)
    print(f"adjoint vs finite differences: 32 faces, worst relative difference {worst:.3e}", flush=True)
