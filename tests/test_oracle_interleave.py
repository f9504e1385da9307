"""The oracle's charged-step primitives pinned by the reference's own
behavioural cases, restated here as inputs and expected values:

- warpx_overlay/Source/Rrea/RreaChargedInterleaveSmoke.cpp: unit cycle and
  normalisation (:31-48), the Boris rotation (:50-61), cone directions
  (:63-94), the two-body Moller cosines and lab directions (:96-184), the
  photon recoil and axis rotation (:186-217), the ion-pair partition
  (:219-272), the stop/escape predicate (:305-314), the channel gate
  (:350-383), the Strang half kick (:385-430), the capped Poisson draw
  (:432-446), the event interleave on its analytic recording medium
  (:448-685), the substep skeleton (:687-803) and the exact boundary
  crossings (:805-880);
- native_amrex_rrea/src/RreaTransportPathSmoke.cpp:238-424: the chord clip
  (the non-periodic cases; periodic z is out of scope);

plus closed forms the reference code fixes: the competing-hazard selection
(RreaParticleInteraction.H:1402-1454) and the cutoff limiter on a linear loss
(RreaParticleInteraction.H:203-265).  Every probe runs the function the
oracle's electron, positron and newborn loops call (oracle/rrea_oracle.cpp,
the orc_* probes)."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle import binding as ob

ME = 510998.95  # eV, the reference's constant


def d3(*v):
    return (C.c_double * 3)(*v)


def dot(a, b):
    return sum(a[i] * b[i] for i in range(3))


def close(a, b, rel):
    assert abs(a - b) <= rel * max(abs(b), 1e-300), (a, b)


L = ob.lib()


# ---------------------------------------------------------------------------
# small kinematic helpers
# ---------------------------------------------------------------------------
def test_unit_cycle_and_normalize():
    assert L.orc_unit_cycle(-0.25) == 0.75
    assert L.orc_unit_cycle(1.25) == 0.25
    assert L.orc_unit_cycle(1.0) == 0.0
    assert L.orc_unit_cycle(0.0) == 0.0
    v = d3(0.0, 0.0, 0.0)
    L.orc_normalize(v)
    assert tuple(v) == (0.0, 0.0, 1.0)
    v = d3(3.0, 0.0, 4.0)
    L.orc_normalize(v)
    close(v[0], 0.6, 1e-15)
    close(v[2], 0.8, 1e-15)


def test_boris_rotation_keeps_the_energy():
    u = d3(1.0e8, 2.0e7, -3.0e7)
    before = dot(u, u)
    L.orc_boris_rotate(u, -1.75882001076e11 * 2.5e-9, d3(0.0, 18.0e-6, 0.0))
    close(dot(u, u), before, 2e-15)
    assert u[0] != 1.0e8  # it did rotate


@pytest.mark.parametrize("axis", [(0.1, 0.2, 0.97), (0.9, 0.1, 0.3)])
def test_cone_direction(axis):
    n = math.sqrt(dot(axis, axis))
    for ct in (-0.8, -0.2, 0.3, 0.7, 0.999):
        for az in (0.0, 0.31, 0.77):
            o = d3()
            L.orc_cone_direction(d3(*axis), ct, az, o)
            close(dot(o, o), 1.0, 1e-15)
            close(dot(o, [a / n for a in axis]), ct, 1e-14)
    o = d3()
    L.orc_cone_direction(d3(0, 0, 1), 1.5, 0.4, o)
    close(o[2], 1.0, 1e-15)
    L.orc_cone_direction(d3(0, 0, 1), -2.0, 0.4, o)
    close(o[2], -1.0, 1e-15)


def test_moller_cosines_closed_form_and_limits():
    for t in (2.0e5, 1.0e6, 7.2e6, 5.0e7, 1.0e9):
        for f in (0.05, 0.25, 0.5):
            eps = f * t
            expected = math.sqrt(eps * (t + 2 * ME) / (t * (eps + 2 * ME)))
            close(L.orc_moller_cos(t, eps, 1), expected, 1e-14)
    assert L.orc_moller_cos(2.0e5, 1.0e5, 1) < 0.80
    assert L.orc_moller_cos(1.0e6, 1.0e-6, 1) < 1.0e-3
    assert L.orc_moller_cos(1.0e6, 1.0e6, 1) == 1.0
    close(L.orc_moller_cos(1.0e6, 5.0e5, 1), L.orc_moller_cos(1.0e6, 5.0e5, 0), 1e-15)
    assert L.orc_moller_cos(1.0e6, 1.0e6, 0) == 1.0


def test_moller_lab_directions():
    i = (0.267261241912424, 0.534522483824849, 0.801783725737273)
    for t in (3.0e5, 2.0e6, 3.0e7):
        for f in (0.1, 0.5, 0.8, 0.97):
            for az in (0.05, 0.6):
                eps = f * t
                p, s = d3(), d3()
                L.orc_moller_lab_directions(d3(*i), t, eps, az, p, s)
                close(dot(p, p), 1.0, 1e-15)
                close(dot(s, s), 1.0, 1e-15)
                cs, cp = dot(s, i), dot(p, i)
                close(cs, L.orc_moller_cos(t, eps, 1), 1e-14)
                close(cp, L.orc_moller_cos(t, eps, 0), 1e-14)
                ts = [s[k] - cs * i[k] for k in range(3)]
                tp = [p[k] - cp * i[k] for k in range(3)]
                ns, nq = math.sqrt(dot(ts, ts)), math.sqrt(dot(tp, tp))
                if ns > 1e-12 and nq > 1e-12:
                    close(dot([x / ns for x in ts], [x / nq for x in tp]), -1.0, 1e-13)


def test_photon_recoil_and_axis_rotation():
    c = d3(0.0, 0.0, 1.0)
    L.orc_photon_recoil(1.0e6, 2.0e5, d3(0, 0, 1), c)
    close(c[2], 1.0, 1e-15)
    c = d3(0.0, 0.0, 1.0)
    L.orc_photon_recoil(1.0e6, 2.0e5, d3(1, 0, 0), c)
    assert c[0] < 0.0 and c[2] > 0.0
    p0 = math.sqrt(1.0e6 * (1.0e6 + 2 * ME))
    close(c[0], -2.0e5 / math.sqrt(2.0e5 ** 2 + p0 * p0), 1e-14)
    o = d3()
    L.orc_rotate_to_axis(d3(0.6, 0.0, 0.8), d3(0.0, 0.0, 1.0), o)
    close(o[0], 0.6, 1e-15)
    close(o[2], 0.8, 1e-15)
    L.orc_rotate_to_axis(d3(0.6, 0.0, 0.8), d3(1.0, 0.0, 0.0), o)
    assert abs(dot(o, (0.6, 0.0, 0.8))) <= 1e-15


# ---------------------------------------------------------------------------
# the continuous segment: ion pairs, cutoff limiter
# ---------------------------------------------------------------------------
def partition(ib, ie, ir, full, b, e, seg, raw, w, wair, stp):
    n = len(ib)
    arr = lambda v: (C.c_double * max(n, 1))(*v)  # noqa: E731
    of, ow = (C.c_double * 8)(), (C.c_double * 8)()
    m = L.orc_partition_ion_pairs(arr(ib), arr(ie), arr(ir), n, full, b, e, seg, raw, w, wair, stp, of, ow, 8)
    return list(of[:m]), list(ow[:m])


def test_ion_pair_partition():
    ib, ie, ir = [0.0, 0.4, 0.9], [0.4, 1.0, 0.95], [1.0, 0.5, 2.0]
    f, w = partition(ib, ie, ir, 2.0, 0.25, 0.85, 6.0e4, 7.5e4, 100.0, 34.0, 1.0e5)
    assert len(f) == 2
    close(f[0], 0.325, 1e-15)
    close(f[1], 0.625, 1e-15)
    close(w[0], 100.0 * 2.4e4 / 34.0, 1e-12)
    close(w[1], 100.0 * 3.6e4 / 34.0, 1e-12)
    close((w[0] + w[1]) * 34.0 / 100.0, 6.0e4, 1e-12)
    # a closed segment emits nothing; a sub-unity W_air divides by the 1.0 floor
    assert partition(ib, ie, ir, 2.0, 0.25, 0.85, 0.0, 7.5e4, 100.0, 34.0, 1.0e5) == ([], [])
    f, w = partition(ib, ie, ir, 2.0, 0.25, 0.4, 3.0e4, 3.0e4, 100.0, 0.5, 1.0e5)
    close(w[-1], 100.0 * 3.0e4 / 1.0, 1e-12)


@pytest.mark.parametrize("b,e,k,cut,rate", [(0.0, 1.0, 1.0e6, 1.0e4, 2.0e6), (0.2, 0.9, 5.0e5, 1.0e4, 1.0e6),
                                            (0.0, 1.0, 3.0e4, 1.0e4, 1.0e3)])
def test_cutoff_limiter_on_a_linear_loss(b, e, k, cut, rate):
    """loss_to(f) = rate (f - b): the stop lies at b + (k - cut) / rate when
    that is inside [b, e] (to the 56-step bisection's resolution), else e."""
    end = C.c_double()
    reached = L.orc_limit_at_cutoff_linear(b, e, k, cut, rate, C.byref(end))
    f_star = b + (k - cut) / rate
    if f_star <= e:
        assert reached == 1
        assert abs(end.value - f_star) <= (e - b) * 2.0 ** -55
    else:
        assert reached == 0 and end.value == e


def test_cutoff_limiter_edges():
    end = C.c_double()
    assert L.orc_limit_at_cutoff_linear(0.3, 0.3, 1.0e4, 1.0e4, 1.0, C.byref(end)) == 1 and end.value == 0.3
    assert L.orc_limit_at_cutoff_linear(0.3, 0.3, 2.0e4, 1.0e4, 1.0, C.byref(end)) == 0 and end.value == 0.3
    # already at the cutoff: the stop is at the segment start
    assert L.orc_limit_at_cutoff_linear(0.1, 0.8, 1.0e4, 1.0e4, 1.0, C.byref(end)) == 1 and end.value == 0.1
    # fail-closed inputs: below the cutoff, reversed segment, negative loss
    assert L.orc_limit_at_cutoff_linear(0.0, 1.0, 5.0e3, 1.0e4, 1.0, C.byref(end)) == -4
    assert L.orc_limit_at_cutoff_linear(0.5, 0.4, 5.0e4, 1.0e4, 1.0, C.byref(end)) == -4
    assert L.orc_limit_at_cutoff_linear(0.0, 1.0, 5.0e4, 1.0e4, -1.0, C.byref(end)) == -4


# ---------------------------------------------------------------------------
# decision helpers, the channel gate, the half kick, the Poisson draw
# ---------------------------------------------------------------------------
def test_stop_is_escape():
    assert L.orc_stop_is_escape(1, 0.8, 0.8) == 1
    assert L.orc_stop_is_escape(1, 0.80000001, 0.8) == 1
    assert L.orc_stop_is_escape(1, 0.79, 0.8) == 0
    assert L.orc_stop_is_escape(0, 0.9, 0.8) == 0


def test_channel_gate():
    calls = C.c_int32(0)
    close(L.orc_gated_inverse_length(1, 8.0, C.byref(calls)), 1.0 / 8.0, 1e-14)
    close(L.orc_gated_inverse_length(1, 8.0 / 0.25, C.byref(calls)), 0.25 / 8.0, 1e-14)
    before = calls.value
    assert L.orc_gated_inverse_length(0, 8.0, C.byref(calls)) == 0.0
    assert calls.value == before  # a closed channel does not evaluate its mean free path
    assert L.orc_gated_inverse_length(1, 0.0, C.byref(calls)) == 0.0
    assert L.orc_gated_inverse_length(1, -1.0, C.byref(calls)) == 0.0


def half_kick(field, half_dt, q, k, d):
    f = (C.c_double * 7)(*field)
    kk = C.c_double(k)
    dd = d3(*d)
    hw = L.orc_half_kick(f, half_dt, q, C.byref(kk), dd)
    return hw, kk.value, tuple(dd)


def test_half_kick():
    # a transverse field with zero work turns the direction and keeps the energy
    hw, k, d = half_kick([0.0, 0.0, 5.0e6, 0.0, 0.0, 0.0, 0.0], 1.0e-9, -1.0, 3.0e6, (0, 0, 1))
    assert hw == 0.0
    close(k, 3.0e6, 1e-14)
    close(dot(d, d), 1.0, 1e-14)
    assert d[1] != 0.0
    # half the work, with its sign
    hw, k, _ = half_kick([4.0e5, 0, 0, 0, 0, 0, 0], 1.0e-12, -1.0, 1.0e6, (0, 0, 1))
    close(hw, 2.0e5, 1e-14)
    close(k, 1.2e6, 1e-14)
    # floored at zero
    _, k, _ = half_kick([-1.0e7, 0, 0, 0, 0, 0, 0], 1.0e-12, -1.0, 1.0e5, (0, 0, 1))
    assert k == 0.0
    # the charge sign reverses the impulse
    _, _, de = half_kick([0.0, 0.0, 5.0e6, 0, 0, 0, 0], 1.0e-9, -1.0, 3.0e6, (0, 0, 1))
    _, _, dp = half_kick([0.0, 0.0, 5.0e6, 0, 0, 0, 0], 1.0e-9, +1.0, 3.0e6, (0, 0, 1))
    close(de[1], -dp[1], 1e-14)


def test_poisson_capped():
    assert L.orc_poisson_capped(0.0, 0.7, 12) == 0
    assert L.orc_poisson_capped(3.0, 1.0 - 1.0e-16, 9) == 9
    n = 10000
    mean = sum(L.orc_poisson_capped(3.0, (i + 0.5) / n, 30) for i in range(n)) / n
    close(mean, 3.0, 1e-2)


def select(taus, nch, u1, u2):
    out = (C.c_double * 5)()
    L.orc_select_hazard((C.c_double * 3)(*(list(taus) + [0.0] * (3 - len(taus)))), nch, u1, u2, out)
    return bool(out[0]), bool(out[1]), int(out[2]), out[3], out[4]


def test_competing_hazard_selection_closed_forms():
    taus = (0.2, 0.3, 0.5)
    # the event iff -ln(1 - u1) < sum(taus); the channel by u2 * sum
    v, ev, ch, s, tot = select(taus, 3, -math.expm1(-0.9), 0.25)
    assert v and ev and ch == 1 and tot == 1.0 and abs(s - 0.9) <= 1e-15
    v, ev, ch, s, _ = select(taus, 3, -math.expm1(-1.1), 0.25)
    assert v and not ev
    assert select(taus, 3, 0.1, 0.1)[2] == 0
    assert select(taus, 3, 0.1, 0.9)[2] == 2
    # a draw of exactly one picks the last open channel
    assert select((0.3, 0.0, 0.0), 3, 0.1, 1.0)[2] == 0
    assert select((0.3, 0.4, 0.0), 3, 0.1, 1.0)[2] == 1
    # invalid: a negative depth, a channel count outside 1..3, a non-finite draw
    assert not select((0.3, -0.1, 0.0), 2, 0.1, 0.1)[0]
    assert not select(taus, 0, 0.1, 0.1)[0]
    assert not select(taus, 3, float("nan"), 0.1)[0]
    # u1 -> 1 is floored at 1e-30 inside the log
    assert select((100.0,), 1, 1.0, 0.5)[3] == -math.log(1.0e-30)


# ---------------------------------------------------------------------------
# the event interleave on the reference's recording medium
# ---------------------------------------------------------------------------
def u_for_tau(tau):
    return -math.expm1(-tau)


def interleave(kinetic, cutoff, open_above, draws, losses, consume=False):
    rec = ob.InterleaveRecord()
    flat = [v for d in draws for v in d]
    dr = (C.c_double * max(len(flat), 1))(*flat)
    lo = (C.c_double * max(len(losses), 1))(*losses)
    code = L.orc_interleave_probe(kinetic, cutoff, open_above, dr, len(draws), lo, len(losses), int(consume),
                                  C.byref(rec))
    assert code == 0
    return rec


def test_interleave_refreezes_at_the_survivor_energy():
    r = interleave(1.0e6, 1.0e3, 0.0, [(u_for_tau(0.25), 0.1), (u_for_tau(10.0), 0.1)], [1.0e5])
    assert r.n_frozen == 2
    close(r.frozen[0], 1.0e6, 1e-15)
    close(r.frozen[1], 7.75e5, 1e-12)
    close(r.taus[1], 0.58125, 1e-12)
    assert r.n_segments == 2
    close(r.seg_begin[0], 0.0, 1e-15)
    close(r.seg_end[0], 0.25, 1e-12)
    close(r.seg_frozen[0], 1.0e6, 1e-15)
    close(r.open_energies[0], 8.75e5, 1e-12)
    assert r.n_final == 1 and r.n_stops == 0
    close(r.seg_begin[1], 0.25, 1e-12)
    close(r.seg_frozen[1], 7.75e5, 1e-12)


def test_interleave_post_event_loss_at_the_survivor_energy():
    r = interleave(1.0e6, 1.0e3, 0.0, [(u_for_tau(0.25), 0.1), (u_for_tau(0.09375), 0.1), (u_for_tau(10.0), 0.1)],
                   [5.0e5, 1.0e4])
    assert r.n_final == 2
    close(r.seg_begin[1], 0.25, 1e-12)
    close(r.seg_end[1], 0.5, 1e-11)
    close(r.seg_frozen[1], 3.75e5, 1e-12)
    close(0.5 * (r.seg_end[1] - r.seg_begin[1]) * r.seg_frozen[1], 46875.0, 1e-11)


def test_interleave_drag_closed_channel_is_a_null_event():
    r = interleave(9.0e5, 1.0e3, 8.0e5, [(u_for_tau(0.54), 0.1), (u_for_tau(10.0), 0.1)], [])
    assert r.n_final == 0
    assert r.n_draws == 2 and r.draw_ordinals[0] == 0 and r.draw_ordinals[1] == 1
    close(r.taus[1], 0.252, 1e-11)
    assert r.n_open == 2
    close(r.open_energies[1], 9.0e5, 1e-15)


def test_interleave_consumed_parent_ends_at_once():
    r = interleave(1.0e6, 1.0e3, 0.0, [(u_for_tau(0.25), 0.1)], [1.0e5], consume=True)
    assert r.n_final == 1 and r.n_segments == 1 and r.n_stops == 0 and r.n_draws == 1


def test_interleave_cutoff_stop_at_the_limiter_fraction():
    r = interleave(1.5e3, 1.0e3, 0.0, [(u_for_tau(10.0), 0.1)], [])
    assert r.n_stops == 1
    close(r.stops[0], 0.5, 1e-15)
    assert r.n_final == 0


# ---------------------------------------------------------------------------
# the substep skeleton
# ---------------------------------------------------------------------------
NEVER = 1000


def substeps(n, build_ok=True, refuse_at=NEVER, exit_at=NEVER, stop_after=False, kinetic=1.0e6):
    tags = (C.c_int32 * 64)()
    done = C.c_uint64()
    ex = C.c_int32()
    m = L.orc_substep_probe(n, kinetic, int(build_ok), refuse_at, exit_at, int(stop_after), tags, 64, C.byref(done),
                            C.byref(ex))
    names = {1: "build", 2: "events", 3: "finish", 4: "advance"}
    return [f"{names[t // 1000]}{t % 1000}" for t in tags[:m]], done.value, bool(ex.value)


def test_substep_skeleton():
    tags, done, ex = substeps(3)
    assert tags == [f"{k}{i}" for i in range(3) for k in ("build", "events", "finish", "advance")]
    assert done == 3 and not ex
    tags, done, ex = substeps(3, refuse_at=1)
    assert ex and done == 1 and tags[-1] == "advance1" and len(tags) == 8
    tags, done, ex = substeps(3, exit_at=0)
    assert ex and done == 0
    tags, done, ex = substeps(3, stop_after=True)
    assert tags == ["build0", "events0"] and done == 0 and not ex
    tags, _, _ = substeps(3, build_ok=False)
    assert tags == ["build0"]
    tags, done, _ = substeps(3, kinetic=0.5e3)
    assert tags == [] and done == 0


# ---------------------------------------------------------------------------
# exact boundary crossings and the running-point commit (domain r < 8,
# 0 <= z < 10)
# ---------------------------------------------------------------------------
def crossing(r0, z0, r1, z1):
    out = (C.c_double * 8)()
    assert L.orc_boundary_crossing(r0, 0.0, z0, r1, 0.0, z1, 0.0, 8.0, 0.0, 10.0, out) == 0
    return out


@pytest.mark.parametrize("r0,z0,r1,z1,cr,cz,faces", [
    (7.0, 4.0, 9.0, 5.0, 8.0, 4.5, 1),          # r_hi
    (3.0, 1.0, 3.0, -1.0, 3.0, 0.0, 2),         # z_lo
    (3.0, 9.0, 3.0, 11.0, 3.0, 10.0, 4),        # z_hi
    (7.0, 9.0, 9.0, 11.0, 8.0, 10.0, 1 | 4),    # r_hi-z_hi corner
    (7.0, 1.0, 9.0, -1.0, 8.0, 0.0, 1 | 2),     # r_hi-z_lo corner
    (7.0, 0.0, 9.0, 0.0, 8.0, 0.0, 1),          # radial grazing on z_lo
    (7.0, 4.0, 8.0, 4.0, 8.0, 4.0, 1),          # endpoint exactly on the excluded r_hi face
    (3.0, 9.0, 3.0, 10.0, 3.0, 10.0, 4),        # endpoint exactly on the excluded z_hi face
    (3.0, 0.0, 3.0, -1.0, 3.0, 0.0, 2),         # newborn on the included z_lo face, leaving
])
def test_boundary_crossings(r0, z0, r1, z1, cr, cz, faces):
    out = crossing(r0, z0, r1, z1)
    assert out[0] == 1.0 and int(out[1]) == faces
    close(out[2], r0, 2e-14)
    close(out[3], z0, 2e-14) if z0 != 0.0 else None
    close(out[4], cr, 2e-14)
    assert abs(out[5] - cz) <= 2e-14 * max(abs(cz), 1.0)


def test_exact_upper_face_is_refused_and_newborn_exit_at_zero():
    out = crossing(7.0, 4.0, 8.0, 4.0)
    assert out[6] == 0.0  # an endpoint exactly on r_hi is not a geometric overshoot
    run = d3(7.0, 0.0, 4.0)
    assert L.orc_advance_running_position(7.0, 0.0, 4.0, 8.0, 0.0, 4.0, 0.0, 8.0, 0.0, 10.0, out[7], run) == 0
    assert tuple(run) == (7.0, 0.0, 4.0)
    out = crossing(3.0, 0.0, 3.0, -1.0)
    assert out[6] == 1.0 and out[7] == 0.0


# ---------------------------------------------------------------------------
# the chord clip (RreaTransportPathSmoke.cpp, non-periodic cases)
# ---------------------------------------------------------------------------
def clip(sr, st, sz, er, et, ez, r_lo, r_hi, z_lo, z_hi):
    out = (C.c_double * 12)()
    assert L.orc_clip_path(sr, st, sz, er, et, ez, r_lo, r_hi, z_lo, z_hi, out) == 0
    keys = ("sx", "sy", "sz", "ex", "ey", "ez", "full", "begin", "end", "in_len", "exits", "pieces")
    return dict(zip(keys, out))


def test_clip_interior_axial_radial_and_axis_crossing():
    p = clip(0.5, 0.0, 0.0, 0.5, 0.0, 0.25, 0.0, 1.0, -1.0, 1.0)
    assert abs(p["full"] - 0.25) <= 1e-14 and abs(p["in_len"] - 0.25) <= 1e-14 and p["exits"] == 0
    assert p["pieces"] == 1 and p["begin"] == 0.0
    p = clip(0.5, 0.0, 0.5, 0.5, 0.0, 1.5, 0.0, 1.0, -1.0, 1.0)
    assert abs(p["end"] - 0.5) <= 1e-14 and abs(p["in_len"] - 0.5) <= 1e-14 and p["exits"] == 1
    p = clip(0.25, 0.0, 0.0, 1.25, 0.0, 0.0, 0.0, 1.0, -1.0, 1.0)
    assert abs(p["end"] - 0.75) <= 1e-14 and abs(p["in_len"] - 0.75) <= 1e-14 and p["exits"] == 1
    p = clip(0.5, 0.0, 0.0, 0.5, math.pi, 0.0, 0.0, 1.0, -1.0, 1.0)
    assert abs(p["full"] - 1.0) <= 1e-14 and abs(p["in_len"] - 1.0) <= 1e-14 and p["pieces"] == 1


def test_clip_compton_continuation_and_annulus():
    inc = clip(0.5, 0.0, 0.5, 0.5, 0.0, 1.5, 0.0, 1.0, 0.0, 1.0)
    assert abs(inc["full"] - 1.0) <= 1e-14 and abs(inc["in_len"] - 0.5) <= 1e-14
    unused = inc["full"] - 0.25
    cont = clip(0.5, 0.0, 0.75, 0.5, 0.0, 0.75 - unused, 0.0, 1.0, 0.0, 1.0)
    assert abs(cont["in_len"] - 0.75) <= 1e-14 and cont["exits"] == 0
    ann = clip(0.75, 0.0, 0.5, 0.75, math.pi, 0.5, 0.5, 1.0, 0.0, 1.0)
    assert ann["exits"] == 1 and abs(ann["in_len"] - 0.25) <= 1e-14


def test_clip_exact_upper_faces_keep_half_open_ownership():
    p = clip(0.5, 0.0, 0.0, 0.5, 0.0, 1.0, 0.0, 1.0, -1.0, 1.0)
    assert p["exits"] == 0 and L.orc_outside_certified(0.0, 1.0, -1.0, 1.0, 0.5, 1.0) == 1
    p = clip(0.25, 0.0, 0.0, 1.0, 0.0, 0.0, 0.0, 1.0, -1.0, 1.0)
    assert p["exits"] == 0 and L.orc_outside_certified(0.0, 1.0, -1.0, 1.0, 1.0, 0.0) == 1
    assert L.orc_outside_certified(0.0, 1.0, -1.0, 1.0, 0.5, -1.0) == 0


def test_clip_rejects_a_start_outside_the_domain():
    out = (C.c_double * 12)()
    assert L.orc_clip_path(2.0, 0.0, 0.0, 2.0, 0.0, 0.5, 0.0, 1.0, -1.0, 1.0, out) == 2
    assert L.orc_clip_path(-0.5, 0.0, 0.0, 0.5, 0.0, 0.5, 0.0, 1.0, -1.0, 1.0, out) == 4


def test_chord_points_are_rz_round_trips():
    """Cartesian end points from (r, theta): a random clip's points equal
    r cos theta, r sin theta to the last place, and the clipped length is the
    chord's length times the in-domain fraction."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        sr, er = rng.uniform(0, 2, 2)
        st, et = rng.uniform(-math.pi, math.pi, 2)
        sz, ez = rng.uniform(-2, 2, 2)
        if not (sr < 1.5 and -1.0 <= sz <= 1.5):
            continue
        p = clip(sr, st, sz, er, et, ez, 0.0, 1.5, -1.0, 1.5)
        assert p["sx"] == sr * math.cos(st) and p["sy"] == sr * math.sin(st)
        full = math.sqrt((p["ex"] - p["sx"]) ** 2 + (p["ey"] - p["sy"]) ** 2 + (p["ez"] - p["sz"]) ** 2)
        assert abs(p["full"] - full) <= 1e-15 * max(full, 1.0)
        assert abs(p["in_len"] - p["full"] * (p["end"] - p["begin"])) <= 1e-15 * max(full, 1.0)
        if p["exits"]:
            x = p["sx"] + p["end"] * (p["ex"] - p["sx"])
            y = p["sy"] + p["end"] * (p["ey"] - p["sy"])
            z = p["sz"] + p["end"] * (p["ez"] - p["sz"])
            on_face = min(abs(math.hypot(x, y) - 1.5), abs(z - 1.5), abs(z + 1.0))
            assert on_face <= 1e-12
