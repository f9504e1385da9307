"""GPU parity of the field side of the feedback loop against the oracle
(oracle/rrea_maxwell_oracle.cpp, itself pinned bit for bit to the reference's
header): the deposit-to-current conversion and every Maxwell update are
bit-identical; the energy ledger agrees to round-off of its summation order;
and a transport <-> Maxwell feedback loop -- each step's electron transport
deposits J, J drives a Maxwell step, the total field feeds the next step's
gather -- tracks the oracle's loop step by step."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

from oracle import maxwell as OM  # noqa: E402
from parity import compare_step, gpu_step  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd.config import RzFields, StepConfig  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tables(bundle):
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import InteractionTables
    return InteractionTables(bundle)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def random_fields(rng, nr, nz, dr, dz):
    n_er, n_ez, n_bt = OM.shapes(nr, nz)
    return RzFields(nr, nz, dr, dz, 0.0, 1e3 * rng.standard_normal(n_er), 1e3 * rng.standard_normal(n_ez),
                    1e5 * rng.standard_normal(n_er), 1e5 * rng.standard_normal(n_ez),
                    3e-6 * rng.standard_normal(n_bt))


@pytest.mark.parametrize("nr,nz", [(5, 3), (37, 53), (256, 1024)])
def test_kernels_are_bit_identical(nr, nz):
    from rrea_feedback_warpx_amrex_geant4_amd.maxwell import MaxwellField
    print(f"case maxwell kernels {nr} x {nz}", flush=True)
    rng = np.random.default_rng(nr + nz)
    dr, dz, dt = 10.0, 9.765625, 1e-9
    f = random_fields(rng, nr, nz, dr, dz)
    n_er, n_ez, _ = OM.shapes(nr, nz)
    # the conversion of a padded index-space deposit, zeros included
    jr_i = rng.standard_normal((nr + 3) * (nz + 4))
    jz_i = rng.standard_normal((nr + 2) * (nz + 5))
    jr_i[rng.random(jr_i.size) < 0.5] = 0.0
    jz_i[rng.random(jz_i.size) < 0.5] = 0.0
    m = MaxwellField(f)
    m.current_from_deposit(dev(jr_i), dev(jz_i))
    want_r, want_z = OM.current_to_physical(nr, nz, dr, dz, jr_i, jz_i)
    assert np.array_equal(m.jr.cpu().numpy(), want_r)
    assert np.array_equal(m.jz.cpu().numpy(), want_z)
    # every update, with conduction, current and ambient
    sr, sz = 1e-3 * rng.random(n_er), 1e-3 * rng.random(n_ez)
    o = OM.Field(nr, nz, dr, dz, f.er, f.ez, f.btheta)
    for _ in range(3):
        m.step(dt, current=True, sigma_r=dev(sr), sigma_z=dev(sz))
        o.apply(0, dt, sigma_r=sr, sigma_z=sz, jr=want_r, jz=want_z, amb_er=f.amb_er, amb_ez=f.amb_ez)
    for k, want in (("er", o.er), ("ez", o.ez), ("bt", o.bt)):
        assert np.array_equal(getattr(m, k).cpu().numpy(), want), k
    # the separate updates
    import ctypes as C
    from rrea_feedback_warpx_amrex_geant4_amd import _lib
    L = _lib.lib()
    g = C.byref(m.geom)
    _lib.check(L.rrea_maxwell_advance_b(g, m.er.data_ptr(), m.ez.data_ptr(), m.bt.data_ptr(), dt, None), "b")
    o.apply(1, dt)
    assert np.array_equal(m.bt.cpu().numpy(), o.bt)
    _lib.check(L.rrea_maxwell_silver_mueller(g, m.er.data_ptr(), m.ez.data_ptr(), m.bt.data_ptr(), None), "sm")
    o.apply(2)
    assert np.array_equal(m.bt.cpu().numpy(), o.bt)
    _lib.check(L.rrea_maxwell_advance_e(g, m.er.data_ptr(), m.ez.data_ptr(), m.bt.data_ptr(), dt, None, None,
                                        m.jr.data_ptr(), m.jz.data_ptr(), None, None, None), "e")
    o.apply(3, dt, jr=want_r, jz=want_z)
    assert np.array_equal(m.er.cpu().numpy(), o.er)
    assert np.array_equal(m.ez.cpu().numpy(), o.ez)
    # the ledger, to summation order
    got = m.ledger(True).cpu().numpy()
    joule, moment = o.joule(want_r, want_z, f.amb_er, f.amb_ez)
    want = [o.energy(f.amb_er, f.amb_ez), joule, o.poynting(f.amb_er, f.amb_ez), moment]
    # random fields: each sum is ~sqrt(N) terms' size, its rounding ~eps N
    for q in range(4):
        assert abs(got[q] - want[q]) <= 1e-11 * abs(want[q]), (q, got[q], want[q])


def test_feedback_loop_tracks_the_oracle(tables, otables):
    """Electrons of weight 1e9 in a uniform -400 kV/m field: each step's
    deposit drives the field the next step gathers, on the device and in the
    oracle, compared every step (particles by the transport parity bar, the
    fields to the J grid's summation-order round-off)."""
    from oracle import binding as ob
    from rrea_feedback_warpx_amrex_geant4_amd.maxwell import MaxwellField
    from rrea_feedback_warpx_amrex_geant4_amd.transport import Coupling, DeviceTile, StepOutputs
    nr, nz, dr, dz = 24, 96, 1.0, 1.0
    fields = RzFields.uniform(nr, nz, dr, dz, ez_v_per_m=-4.0e5, density_ratio=1.225 / 1.20479)
    rng = np.random.default_rng(17)
    _, tile = wl.c1_plain(2000)
    tile.r[:] = 5.0 * np.sqrt(rng.random(len(tile.r)))
    tile.z[:] = 10.0 + 20.0 * rng.random(len(tile.r))
    tile.w[:] = 1e9
    dt = 1e-10
    cfg = StepConfig(rng_seed=99, dt_s=dt, low_energy_cutoff_eV=1e4)
    cp = Coupling(fields, tables)
    m = MaxwellField(fields)
    o = OM.Field(nr, nz, dr, dz, fields.er, fields.ez, fields.btheta)
    host = tile
    stats = {}
    field_scale = 0.0
    for s in range(8):
        print(f"case feedback step {s}: {len(host)} electrons", flush=True)
        cfg.step = s
        f_o = RzFields(nr, nz, dr, dz, 0.0, o.er.copy(), o.ez.copy(), fields.amb_er, fields.amb_ez, o.bt.copy(),
                       density_ratio=fields.density_ratio)
        res = ob.electron_step(otables, f_o, cfg, host.as_dict(), nthreads=8)
        assert res["code"] == 0
        tile_dev = DeviceTile.from_host(host)
        out = StepOutputs(len(host), fields, secondary_capacity=max(4096, len(host)))
        gpu_step(tables, cp, tile_dev, cfg, out)
        compare_step(f_o, host, res, tile_dev, out, label=f"feedback step {s}", dt=dt, stats=stats)
        # the field side: device from the device's deposit, oracle from its own
        m.feedback_step(out.jr, out.jz, dt, coupling=cp)
        jr_o, jz_o = OM.current_to_physical(nr, nz, dr, dz, res["jr"], res["jz"])
        o.apply(0, dt, jr=jr_o, jz=jz_o)
        torch.cuda.synchronize()
        j_scale = max(np.abs(jr_o).max(), np.abs(jz_o).max())
        assert j_scale > 0.0
        assert np.allclose(m.jr.cpu().numpy(), jr_o, rtol=0, atol=1e-10 * j_scale)
        assert np.allclose(m.jz.cpu().numpy(), jz_o, rtol=0, atol=1e-10 * j_scale)
        for k in ("er", "ez", "bt"):
            want = getattr(o, k)
            field_scale = max(np.abs(want).max(), 1e-300)
            assert np.allclose(getattr(m, k).cpu().numpy(), want, rtol=0, atol=1e-10 * field_scale), (s, k)
        keep = res["status"] == 0
        host = host.take(np.flatnonzero(keep))
        for k in ("r", "z", "theta", "ux", "uy", "uz"):
            getattr(host, k)[:] = res[k][keep]
    # the current has built a scattered field the gather now sees
    assert field_scale > 0.0
    print(f"feedback loop: 8 steps, scattered |Ez| up to {np.abs(o.ez).max():.3e} V/m; "
          f"{stats.get('within_1e-12', 0)} of {stats.get('n', 0)} particle-steps within 1e-12", flush=True)
