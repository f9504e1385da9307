"""The concurrent Apply (ConcurrentApply: the three species passes on three
streams from three host threads) against the sequential Apply on the same
inputs: particle states and statuses bit for bit, the staged secondaries and
the drain's FIFO and survivors record for record, the J grid and the
material sources to FP64 summation order, for several steps in a row."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


# (4, "pre", positron early) is the bench's default: the photon push of the
# step at the head of the photon stream, the positron pass starting with the
# electron pass, the photon pass after the electron fast kernel, the drain
# with its counts on the device (four generations); (4, "after", positron
# early) pushes the photons on the main stream after the drain; (0, "after",
# none) is the host-driven drain on the electron pass's handle (whose
# ordering workspace the electron pass also uses) with both passes after the
# electron fast kernel; (16, "stream", both early) drains on the device and
# pushes the photons for the next step on the photon stream after its pass.
@pytest.mark.parametrize("device_drain,push,early", [(4, "pre", ("positron",)), (4, "after", ("positron",)),
                                                     (0, "after", ()), (16, "stream", ("positron", "photon"))])
def test_concurrent_apply_matches_sequential(tables, device_drain, push, early):
    from rrea_feedback_warpx_amrex_geant4_amd.transport import (ConcurrentApply, Coupling, DeviceTile, DrainOutputs,
                                                                PhotonTile, RreaParticleInteraction, StepOutputs)
    fields = wl.c2_fields()
    n = 1 << 15
    e = wl.c2_electrons(n, fields, seed=2)
    g = wl.c2_photons(n // 2, fields, dt_s=1e-9, seed=4, first_id=1 << 41)
    p = wl.c2_electrons(n // 4, fields, seed=3, first_id=1 << 40, charge_sign=1.0)
    print(f"case concurrent apply: {n} e-, {n // 2} gamma, {n // 4} e+, 3 steps of 1 ns, "
          f"device_drain={device_drain}, push={push}, early={early}", flush=True)

    def make():
        tiles = (DeviceTile.from_host(e), PhotonTile.from_host(g), DeviceTile.from_host(p))
        oe = StepOutputs(n, fields, secondary_capacity=1 << 15, sources=True)
        og = StepOutputs(n // 2, fields, secondary_capacity=1 << 15, share_grid=oe)
        op = StepOutputs(n // 4, fields, secondary_capacity=1 << 15, share_grid=oe)
        return tiles, (oe, og, op), DrainOutputs(1 << 17, oe)

    seq_t, seq_o, seq_d = make()
    con_t, con_o, con_d = make()
    coupling = Coupling(fields, tables)
    conc = ConcurrentApply(tables, fields, early=early)
    apply = RreaParticleInteraction
    for step in range(3):
        cfg = StepConfig(rng_seed=81, step=step, dt_s=1e-9)
        for outs in (seq_o, con_o):
            for o in outs:
                o.secondary_count.zero_()
            for t in (outs[0].jr, outs[0].jz, outs[0].rho, outs[0].sources):
                t.zero_()
        if push == "pre":
            seq_t[1].push(cfg.dt_s)
        apply.apply_electrons(tables, coupling, seq_t[0], cfg, seq_o[0])
        apply.apply_photons(tables, coupling, seq_t[1], cfg, seq_o[1])
        apply.apply_positrons(tables, coupling, seq_t[2], cfg, seq_o[2])
        apply.drain_secondaries(tables, coupling, cfg, list(seq_o), seq_d)
        conc.run(con_t[0], con_t[1], con_t[2], cfg, con_o[0], con_o[1], con_o[2], con_d, check=True,
                 push_photons=push == "stream", pre_push=push == "pre", device_drain=device_drain)
        if push == "after":
            con_t[1].push(cfg.dt_s)
        if push != "pre":
            seq_t[1].push(cfg.dt_s)
        torch.cuda.synchronize()
        for a, b in ((seq_t[0], con_t[0]), (seq_t[1].cur, con_t[1].cur), (seq_t[2], con_t[2])):
            for k in ("r", "z", "theta", "ux", "uy", "uz", "w", "id"):
                assert torch.equal(getattr(a, k), getattr(b, k)), (step, k)
        for a, b in zip(seq_o, con_o):
            assert torch.equal(a.status, b.status)
            sa, sb = a.secondaries_host(), b.secondaries_host()
            assert len(sa) == len(sb) and (sa == sb).all()
        wa, wb = seq_d.work_host(), con_d.work_host()
        assert len(wa) == len(wb) and (wa == wb).all()
        assert (seq_d.survivors_host() == con_d.survivors_host()).all()
        for k in ("jr", "jz", "sources"):
            a, b = getattr(seq_o[0], k), getattr(con_o[0], k)
            assert torch.allclose(a, b, rtol=0.0, atol=1e-12 * float(a.abs().max())), k
        assert torch.allclose(seq_o[0].tallies, con_o[0].tallies, rtol=1e-12, atol=0.0)
        for k in ("prev_r", "prev_z", "prev_theta"):
            assert torch.equal(getattr(seq_t[1], k), getattr(con_t[1], k)), (step, k)
    print(f"concurrent apply: 3 steps identical ({len(wa)} FIFO entries in the last)", flush=True)
    conc.close()
