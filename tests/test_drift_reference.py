"""The FP64 reference gradient of the drift segment (tests/drift_reference.py)
on the CPU: its autograd gradient against its own Richardson-extrapolated
central differences, and the drift invariants it must keep (a field-free
drift moves each particle along its direction at its speed and loses exactly
the CSDA drag; the deposit is charge-conserving, so with uniform loss weights
on the jz faces of one column it telescopes to the net z-crossings)."""
import math

import numpy as np
import torch

from drift_reference import KC, KMEC2, DriftReference
from rrea_feedback_warpx_amrex_geant4_amd import workloads as wl
from rrea_feedback_warpx_amrex_geant4_amd.config import RzFields


def test_reference_gradient_matches_its_finite_differences(bundle):
    fields, parts = wl.c4_drift(6, nr=16, nz=32, cell_m=0.25, seed=41)
    ref = DriftReference(fields, bundle, dt=1e-10)
    nrp, nzp = fields.padded_shape()
    g = torch.Generator().manual_seed(3)
    gjr = torch.randn((nrp + 1) * nzp, generator=g, dtype=torch.float64)
    gjz = torch.randn(nrp * (nzp + 1), generator=g, dtype=torch.float64)
    er = torch.tensor(fields.er, dtype=torch.float64, requires_grad=True)
    ez = torch.tensor(fields.ez, dtype=torch.float64)
    L = ref.loss(er, ez, parts, 4, gjr, gjz)
    L.backward()
    faces = torch.argsort(er.grad.abs(), descending=True)[:3].tolist()
    h = 400.0

    def central(f, step):
        with torch.no_grad():
            p, m = er.detach().clone(), er.detach().clone()
            p[f] += step
            m[f] -= step
            return float(ref.loss(p, ez, parts, 4, gjr, gjz) - ref.loss(m, ez, parts, 4, gjr, gjz)) / (2 * step)

    for f in faces:
        fd = (4.0 * central(f, h) - central(f, 2 * h)) / 3.0
        ad = float(er.grad[f])
        assert abs(fd - ad) <= 1e-6 * abs(ad), (f, fd, ad)


def test_field_free_drift_and_deposit_continuity(bundle):
    f = RzFields.uniform(16, 64, 0.25, 0.25, z_lo=0.0, ez_v_per_m=0.0, density_ratio=1.0)
    ref = DriftReference(f, bundle, dt=1e-10)
    k0 = 5.0e6
    gamma = 1.0 + k0 / KMEC2
    u = KC * math.sqrt(gamma * gamma - 1.0)
    d = np.array([0.3, 0.0, 0.9])
    d /= np.linalg.norm(d)
    s = tuple(torch.tensor(x, dtype=torch.float64) for x in (2.0, 0.0, 5.0, u * d[0], u * d[1], u * d[2]))
    nrp, nzp = f.padded_shape()
    # weights 1 on every jz face of row fj: the deposit then sums the net
    # z-flux through that row of faces
    gjr = torch.zeros((nrp + 1) * nzp, dtype=torch.float64)
    cer, cez = ref.cells(torch.zeros((f.nr + 1) * f.nz, dtype=torch.float64),
                         torch.zeros(f.nr * (f.nz + 1), dtype=torch.float64))
    out, _ = ref.step(s, 1.0, cer, cez, gjr, torch.zeros(nrp * (nzp + 1), dtype=torch.float64))
    r1, th1, z1 = (float(x) for x in out[:3])
    x1, y1 = r1 * math.cos(th1), r1 * math.sin(th1)
    beta = math.sqrt(1.0 - 1.0 / (gamma * gamma))
    L = KC * beta * 1e-10
    assert math.isclose(x1, 2.0 + L * d[0], rel_tol=1e-13) and math.isclose(z1, 5.0 + L * d[2], rel_tol=1e-13)
    k1 = float(ref.kinetic(*out[3:]))
    drag = (float(ref.value("col", torch.tensor(k0, dtype=torch.float64)))
            + float(ref.value("soft", torch.tensor(k0, dtype=torch.float64)))) * L * f.density_ratio
    assert math.isclose(k0 - k1, drag, rel_tol=1e-9)
    # continuity: the z-flux a chord deposits on the jz faces of row fj is
    # q w / dt times the part of it (in index units) inside [fj - 1, fj], and
    # over all rows the whole chord
    zlo_p = f.z_lo - 2 * f.dz
    xz0, xz1 = (5.0 - zlo_p) / f.dz - 0.5, (z1 - zlo_p) / f.dz - 0.5
    qdt = -1.0 / 1e-10
    ones = torch.ones(nrp * (nzp + 1), dtype=torch.float64)
    _, total = ref.step(s, 1.0, cer, cez, gjr, ones)
    assert math.isclose(float(total), qdt * (xz1 - xz0), rel_tol=1e-12)
    for fj in sorted({math.floor(xz0) + 1, math.floor(xz1) + 1}):
        gjz = torch.zeros(nrp * (nzp + 1), dtype=torch.float64)
        gjz[nrp * fj:nrp * (fj + 1)] = 1.0
        _, lz = ref.step(s, 1.0, cer, cez, gjr, gjz)
        lo, hi = fj - 1.0, float(fj)
        expected = qdt * (min(max(xz1, lo), hi) - min(max(xz0, lo), hi))
        assert math.isclose(float(lz), expected, rel_tol=1e-12), (fj, float(lz), expected)
