"""FP64 reference gradient of the drift segment: the drift step restated in
scalar torch float64 operations, differentiated by torch.autograd -- test
infrastructure for rrea_drift_adjoint.

Each function restates the device drift step's arithmetic (the charged
step's one-substep branch with events and elastic deflection frozen out):
the chord and its RZ round trip (RreaBuildRunningChargedChord,
OV/RreaParticleInteraction.H:110-161; ClipRzTransportPath
NA/src/RreaTransportPath.cpp:74-293), the GL2 field work and average
(OV/RreaPathFieldWork.H:57-150) over the bilinear gather of cell-centred
fields (NA/include/rrea/RreaFieldGatherStencil.H:49-130), the Strang half
kicks (OV/RreaChargedInterleave.H:100-138), the CSDA drag over the table
interpolation (NA/src/RreaInteractionTables.cpp:1580-1630) and the net-chord
deposit (NA/include/rrea/RreaSegmentDepositionRZ.H:96-192), here contracted
with loss weights.  Slow (one particle-step at a time) and meant for a few
hundred particle-steps.
"""

from __future__ import annotations

import bisect
import math

import numpy as np
import torch

KMEC2 = 510998.95
KC = 299792458.0
KEPS = 2.220446049250313e-16
GL2 = 0.57735026918962576451


def tmax(a, b):
    return b if a < b else a


def tmin(a, b):
    return b if b < a else a


def v(x):
    return x.detach().item() if torch.is_tensor(x) else x


class DriftReference:
    def __init__(self, fields, bundle, dt, cutoff=1.0e4, species="electron"):
        self.f = fields
        self.dt = dt
        self.cutoff = cutoff
        self.q = -1.0 if species == "electron" else 1.0
        pre = "e" if species == "electron" else "p"
        self.cx = [float(x) for x in bundle[pre + "_col_x"]]
        self.clogx = [math.log(x) for x in self.cx]
        self.tables = {}
        for name in ("col", "soft"):
            y = [float(a) for a in bundle[f"{pre}_{name}_y"]]
            self.tables[name] = (y, [math.log(a) if a > 0.0 else 0.0 for a in y], int(bundle[f"{pre}_{name}_mode"]))
        f = fields
        self.D = (0.0, f.nr * f.dr, f.z_lo, f.z_lo + f.nz * f.dz)
        nrp, nzp = fields.padded_shape()
        self.G = (nrp, nzp, f.dr, f.dz, 0.0, f.z_lo - 2 * f.dz)
        self.ratio = f.density_ratio
        self.bt = np.asarray(f.btheta, dtype=np.float64)

    # -- fields ---------------------------------------------------------------
    def cells(self, er, ez):
        """cell-centred totals from the Maxwell face arrays (+ ambient)."""
        f = self.f
        nr, nz = f.nr, f.nz
        aer = torch.as_tensor(f.amb_er, dtype=torch.float64)
        aez = torch.as_tensor(f.amb_ez, dtype=torch.float64)
        er2, aer2 = er.reshape(nz, nr + 1), aer.reshape(nz, nr + 1)
        ez2, aez2 = ez.reshape(nz + 1, nr), aez.reshape(nz + 1, nr)
        cer = 0.5 * ((er2[:, :-1] + aer2[:, :-1]) + (er2[:, 1:] + aer2[:, 1:]))
        cez = 0.5 * ((ez2[:-1, :] + aez2[:-1, :]) + (ez2[1:, :] + aez2[1:, :]))
        return cer.reshape(-1), cez.reshape(-1)

    def gather(self, cer, cez, r, z):
        f = self.f
        nr, nz = f.nr, f.nz
        fx = (r - 0.0) / f.dr + 0.0 - 0.5
        fz = (z - f.z_lo) / f.dz + 0.0 - 0.5
        ir0, iz0 = math.floor(v(fx)), math.floor(v(fz))
        tx, tz = fx - float(ir0), fz - float(iz0)
        i0, i1 = max(0, min(nr - 1, ir0)), max(0, min(nr - 1, ir0 + 1))
        j0, j1 = max(0, min(nz - 1, iz0)), max(0, min(nz - 1, iz0 + 1))
        first_center = 0.0 + 0.5 * f.dr
        axis = 1.0 if v(r) >= first_center else tmax(0.0, (r - 0.0) / (0.5 * f.dr))

        def bil(c):
            f00, f01, f10, f11 = c[i0 + nr * j0], c[i0 + nr * j1], c[i1 + nr * j0], c[i1 + nr * j1]
            return (1.0 - tx) * ((1.0 - tz) * f00 + tz * f01) + tx * ((1.0 - tz) * f10 + tz * f11)

        er = bil(cer) * axis
        ez = bil(cez)
        if v(r) <= 0.0:
            bt = 0.0
        else:
            frr = r / f.dr
            frr = 0.0 if v(frr) < 0.0 else (float(nr) if v(frr) > float(nr) else frr)
            fzz = (z - f.z_lo) / f.dz
            fzz = 0.0 if v(fzz) < 0.0 else (float(nz) if v(fzz) > float(nz) else fzz)
            a0, c0 = min(math.floor(v(frr)), nr - 1), min(math.floor(v(fzz)), nz - 1)
            a1, c1 = a0 + 1, c0 + 1
            trr, tzz = frr - float(a0), fzz - float(c0)
            w = nr + 1
            B = self.bt
            bt = (1.0 - trr) * ((1.0 - tzz) * B[a0 + w * c0] + tzz * B[a0 + w * c1]) + trr * (
                (1.0 - tzz) * B[a1 + w * c0] + tzz * B[a1 + w * c1])
        return er, ez, bt

    # -- kinematics ------------------------------------------------------------
    @staticmethod
    def kinetic(ux, uy, uz):
        u2 = ux * ux + uy * uy + uz * uz
        gamma = torch.sqrt(1.0 + u2 / (KC * KC))
        return u2 * KMEC2 / ((KC * KC) * (gamma + 1.0))

    @staticmethod
    def speed(k):
        gamma = 1.0 + k / KMEC2
        beta2 = tmax(0.0, 1.0 - 1.0 / (gamma * gamma))
        return KC * torch.sqrt(beta2)

    @staticmethod
    def proper_velocity(k):
        if v(k) <= 0.0:
            return 0.0
        gamma = 1.0 + k / KMEC2
        return KC * torch.sqrt(tmax(gamma * gamma - 1.0, 0.0))

    def value(self, name, q):
        y, ly, mode = self.tables[name]
        cx, clogx = self.cx, self.clogx
        qv = v(q)
        assert cx[0] <= qv <= cx[-1]
        up = bisect.bisect_left(cx, qv)
        if up == 0:
            return y[0]
        if up == len(cx):
            return y[-1]
        if cx[up] == qv:
            return y[up]
        lo, hi = up - 1, up
        if mode == 3:
            t = (q - cx[lo]) / (cx[hi] - cx[lo])
            return y[lo] + t * (y[hi] - y[lo])
        t = (torch.log(q) - clogx[lo]) / (clogx[hi] - clogx[lo])
        positive = y[lo] > 0.0 and y[hi] > 0.0
        if mode == 2 or (mode == 1 and not positive):
            return y[lo] + t * (y[hi] - y[lo])
        return torch.exp(ly[lo] + t * (ly[hi] - ly[lo]))

    def half_kick(self, f, half_dt, k, d):
        dx, dy, dz = d
        umag = self.proper_velocity(k)
        impulse = self.q * (KC * KC / KMEC2) * half_dt
        ux = umag * dx + impulse * f["ex"]
        uy = umag * dy + impulse * f["ey"]
        uz = umag * dz + impulse * f["ez"]
        gamma = torch.sqrt(1.0 + (ux * ux + uy * uy + uz * uz) / (KC * KC))
        tx, ty, tz = (impulse * f[c] / (2.0 * gamma) for c in ("bx", "by", "bz"))
        t2 = tx * tx + ty * ty + tz * tz
        sx, sy, sz = 2.0 * tx / (1.0 + t2), 2.0 * ty / (1.0 + t2), 2.0 * tz / (1.0 + t2)
        px = ux + uy * tz - uz * ty
        py = uy + uz * tx - ux * tz
        pz = uz + ux * ty - uy * tx
        ox = ux + py * sz - pz * sy
        oy = uy + pz * sx - px * sz
        oz = uz + px * sy - py * sx
        n = torch.sqrt(ox * ox + oy * oy + oz * oz)
        if v(n) > 0.0:
            d = (ox / n, oy / n, oz / n)
        return tmax(0.0, k + 0.5 * f["work"]), d

    # -- deposit -----------------------------------------------------------------
    def deposit_loss(self, gjr, gjz, qw, r0, z0, r1, z1):
        nrp, nzp, dr, dz, rlo, zlo = self.G
        qdt = qw / self.dt
        xr0, xz0 = (r0 - rlo) / dr - 0.5, (z0 - zlo) / dz - 0.5
        xr1, xz1 = (r1 - rlo) / dr - 0.5, (z1 - zlo) / dz - 0.5

        def cuts(a, b):
            av, bv = v(a), v(b)
            if av == bv:
                return []
            lo, hi = min(av, bv), max(av, bv)
            ks = range(math.ceil(lo), math.floor(hi) + 1)
            out = []
            for k in (ks if bv > av else reversed(ks)):
                s = (k - a) / (b - a)
                if 0.0 < v(s) < 1.0:
                    out.append(s)
            return out

        a_list, b_list = cuts(xr0, xr1), cuts(xz0, xz1)
        ai = bi = 0
        s0 = 0.0
        acc = 0.0
        while True:
            a = a_list[ai] if ai < len(a_list) else 2.0
            b = b_list[bi] if bi < len(b_list) else 2.0
            if v(a) <= v(b) and v(a) < 2.0:
                s1 = a
                ai += 1
            elif v(b) < 2.0:
                s1 = b
                bi += 1
            else:
                s1 = 1.0
            if v(s1) > v(s0):
                acc = acc + self.piece(gjr, gjz, qdt, xr0 + s0 * (xr1 - xr0), xz0 + s0 * (xz1 - xz0),
                                       xr0 + s1 * (xr1 - xr0), xz0 + s1 * (xz1 - xz0))
            s0 = s1
            if v(a) >= 2.0 and v(b) >= 2.0:
                break
        return acc

    def piece(self, gjr, gjz, qdt, xr0, xz0, xr1, xz1):
        nrp, nzp = self.G[0], self.G[1]
        acc = 0.0
        fr_flux = qdt * (xr1 - xr0)
        fz_flux = qdt * (xz1 - xz0)
        mr, mz = 0.5 * (xr0 + xr1), 0.5 * (xz0 + xz1)
        jf = math.floor(v(mz))
        fz = mz - float(jf)
        fi = math.floor(v(mr)) + 1
        if fi > 0:
            i = min(fi, nrp)
            for b in range(2):
                j = min(max(jf + b, 0), nzp - 1)
                acc = acc + fr_flux * ((1.0 - fz) if b == 0 else fz) * float(gjr[i + (nrp + 1) * j])
        ifc = math.floor(v(mr))
        fr = mr - float(ifc)
        fj = min(max(math.floor(v(mz)) + 1, 0), nzp)
        for a in range(2):
            i = ifc + a
            if i < 0:
                i = -1 - i
            if i >= nrp:
                i = nrp - 1
            i = max(i, 0)
            acc = acc + fz_flux * ((1.0 - fr) if a == 0 else fr) * float(gjz[i + nrp * fj])
        return acc

    # -- one step ------------------------------------------------------------------
    def step(self, s, w, cer, cez, gjr, gjz):
        """(new state, loss contribution), or (None, 0) when the drift ends."""
        r, th, z, ux, uy, uz = s
        rlo, rhi, zlo, zhi = self.D
        if not (w > 0.0) or v(r) > rhi + 1.0e-9 * rhi:
            return None, 0.0
        k = self.kinetic(ux, uy, uz)
        n = torch.sqrt(ux * ux + uy * uy + uz * uz)
        dx, dy, dz = ux / n, uy / n, uz / n
        if not (v(k) >= self.cutoff):
            return None, 0.0
        x0, y0 = r * torch.cos(th), r * torch.sin(th)
        speed = self.speed(k)
        norm = torch.sqrt(dx * dx + dy * dy + dz * dz)
        scale = speed * self.dt / norm
        ex, ey, ez = x0 + scale * dx, y0 + scale * dy, z + scale * dz
        sr, st, er_, et = torch.hypot(x0, y0), torch.atan2(y0, x0), torch.hypot(ex, ey), torch.atan2(ey, ex)
        sx, sy, sz = sr * torch.cos(st), sr * torch.sin(st), z
        pex, pey, pez = er_ * torch.cos(et), er_ * torch.sin(et), ez
        ddx, ddy, ddz = pex - sx, pey - sy, pez - sz
        full = torch.sqrt(ddx * ddx + ddy * ddy + ddz * ddz)
        # exits end the drift (checked on values, as the device does)
        if v(ddz) > 0.0 and (zhi - v(sz)) / v(ddz) < 1.0:
            return None, 0.0
        if v(ddz) < 0.0 and (zlo - v(sz)) / v(ddz) < 1.0:
            return None, 0.0
        a = v(ddx) ** 2 + v(ddy) ** 2
        if a > 0.0:
            b = 2.0 * (v(sx) * v(ddx) + v(sy) * v(ddy))
            c = v(sx) ** 2 + v(sy) ** 2 - rhi * rhi
            disc = b * b - 4.0 * a * c
            if not (disc < 0.0):
                root = math.sqrt(max(disc, 0.0))
                for cand in ((-b - root) * (0.5 / a), (-b + root) * (0.5 / a)):
                    if cand <= 0.0 or cand > 1.0:
                        continue
                    if 2.0 * a * cand + b > 0.0 and cand < 1.0:
                        return None, 0.0
        e_r = math.hypot(v(pex), v(pey))
        if e_r >= rhi or v(pez) < zlo or v(pez) >= zhi:
            return None, 0.0
        # density path over the single piece
        zm = tmax(zlo, tmin(zhi, sz + 0.5 * (pez - sz) + 0.0))
        rho = self.ratio
        dl = full * (1.0 - 0.0) * rho
        ratio = dl / (full * (1.0 - 0.0))
        # GL2 field work and averages
        pdx, pdy, pdz = ddx / full, ddy / full, ddz / full
        off = 0.5 * GL2
        acc = {c: [] for c in range(7)}
        for f_node in (0.5 - off, 0.5 + off):
            fv = max(0.0, min(1.0, f_node))
            x = sx + fv * (pex - sx)
            y = sy + fv * (pey - sy)
            zz = tmax(zlo, tmin(zhi, sz + fv * (pez - sz) + 0.0))
            rr = torch.hypot(x, y)
            rad, rc, rs = (pdx * x + pdy * y) / rr, x / rr, y / rr
            g_er, g_ez, bt = self.gather(cer, cez, rr, zz)
            wk = self.q * (g_er * rad + g_ez * pdz) * 1.0
            vals = [wk, g_er * rc, g_er * rs, g_ez, -bt * rs, bt * rc, 0.0]
            for c in range(7):
                acc[c].append(vals[c])
        integ = [0.5 * (acc[c][0] + acc[c][1]) for c in range(7)]
        length = 0.0 + full * (1.0 - 0.0)
        fld = {"work": 0.0 + full * integ[0]}
        for c, name in zip(range(1, 7), ("ex", "ey", "ez", "bx", "by", "bz")):
            fld[name] = (0.0 + full * integ[c]) / length
        k1, d1 = self.half_kick(fld, 0.5 * self.dt, k, (dx, dy, dz))
        if not (v(k1) >= self.cutoff):
            return None, 0.0
        L = full * (1.0 - 0.0) * ratio
        col = self.value("col", k1) * 1.0
        soft = self.value("soft", k1) * 1.0
        req = (col + soft) * L
        if not (v(req) < v(k1) - self.cutoff):
            return None, 0.0
        drag = tmin(col * L + soft * L, tmax(0.0, k1 - self.cutoff))
        k2 = tmax(0.0, k1 - drag)
        if not (v(k2) >= self.cutoff):
            return None, 0.0
        k3, d2 = self.half_kick(fld, 0.5 * self.dt, k2, d1)
        if not (v(k3) >= self.cutoff):
            return None, 0.0
        x = sx + 1.0 * (pex - sx)
        y = sy + 1.0 * (pey - sy)
        zz = sz + 1.0 * (pez - sz)
        end_r = torch.hypot(x, y)
        um = self.proper_velocity(k3)
        out = (end_r, torch.atan2(y, x), zz, um * d2[0], um * d2[1], um * d2[2])
        loss = self.deposit_loss(gjr, gjz, self.q * w, r, z, end_r, zz)
        return out, loss

    def loss(self, er, ez, particles, steps, gjr, gjz):
        """L = sum over particles and steps of the deposit contracted with the
        loss weights, as a differentiable function of the face fields."""
        cer, cez = self.cells(er, ez)
        total = 0.0
        for p in range(len(particles.r)):
            s = tuple(torch.tensor(float(getattr(particles, k)[p]), dtype=torch.float64)
                      for k in ("r", "theta", "z", "ux", "uy", "uz"))
            w = float(particles.w[p])
            for _ in range(steps):
                s_new, lp = self.step(s, w, cer, cez, gjr, gjz)
                if s_new is None:
                    break
                total = total + lp
                s = s_new
        return total
