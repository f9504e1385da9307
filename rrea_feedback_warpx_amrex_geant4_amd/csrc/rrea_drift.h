// The deterministic drift segment of the charged step and its adjoint.
//
// The drift step is the charged step's fast branch (RreaElectronTransport.cpp
// one-substep path: RreaBuildRunningChargedChord OV/RreaParticleInteraction.H:
// 110-161, path_field_work_and_average OV/RreaPathFieldWork.H:57-150, the
// Strang half kicks RreaApplyChargedFieldHalfKick OV/RreaChargedInterleave.H:
// 100-138, the CSDA drag RreaPlanContinuousDragSegment :239-292, the net-chord
// deposit DepositSegmentRZ NA/include/rrea/RreaSegmentDepositionRZ.H:159-192)
// with the discrete events and the elastic deflection frozen out: one Strang
// step per dt, push + continuous loss + deposition.  Written once over a
// scalar type T: T = double is the forward model, with the same operations in
// the same order as the transport kernels' fast branch (so a particle that
// has no event and no elastic scatter there gets bit-identical results), and
// T = Dual<N> gives the step's local Jacobian for the reverse sweep.
//
// A particle's drift ends (it is removed, with no deposit for that step)
// where the transport step would leave the one-substep branch for a boundary
// or the cutoff: its chord reaches the domain boundary, its end point is
// outside the half-open domain, or its energy falls below the cutoff.
#pragma once

#include "rrea_device.h"
#include "rrea_dual.h"

#include <type_traits>

namespace rrea_dev {

template <class T>
struct DState { T r, th, z, ux, uy, uz; };

// where a gathered value came from: the four cell-centred values and their
// weights (the bilinear weights times the odd-radial axis scale for Er)
struct GatherStencil {
    int32_t c[4];
    double w_er[4];
    double w_ez[4];
};

// Tangent slots of the gathered fields: Er and Ez at each GL node.  A double
// evaluation has no slots; a dual one adds a unit tangent (value unchanged).
template <class T> struct FieldSlots {
    __device__ static void add(T&, int) {}
};
template <int N> struct FieldSlots<Dual<N>> {
    __device__ static void add(Dual<N>& x, int slot) { x.d[slot] += 1.0; }
};
constexpr int kStateSlots = 6;
constexpr int kDriftSlots = 10;  // 6 state + (Er, Ez) at 2 GL nodes
using D10 = Dual<kDriftSlots>;

template <class T>
__device__ __forceinline__ T speed_t(T k)
{
    T const gamma = 1.0 + k / kMeC2;
    T const beta2 = tmax(0.0, 1.0 - 1.0 / (gamma * gamma));
    return kC * sqrt(beta2);
}

template <class T>
__device__ __forceinline__ T proper_velocity_t(T k)
{
    if (val(k) <= 0.0) return T(0.0);
    T const gamma = 1.0 + k / kMeC2;
    return kC * sqrt(tmax(gamma * gamma - 1.0, 0.0));
}

template <class T>
__device__ __forceinline__ T kinetic_t(T ux, T uy, T uz)
{
    T const u2 = ux * ux + uy * uy + uz * uz;
    T const gamma = sqrt(1.0 + u2 / (kC * kC));
    return u2 * kMeC2 / ((kC * kC) * (gamma + 1.0));
}

template <class T>
__device__ __forceinline__ void normalize_t(T& x, T& y, T& z)
{
    T const n = sqrt(x * x + y * y + z * z);
    if (val(n) <= 0.0) { x = T(0.0); y = T(0.0); z = T(1.0); return; }
    x = x / n; y = y / n; z = z / n;
}

template <class T>
struct DPath { T sx, sy, sz, ex, ey, ez, full; };

// The running chord of a particle that stays inside the domain over dt
// (build_chord + clip_rz_path with no exit), or false.
template <class T>
__device__ bool chord_t(T x, T y, T z, T dx, T dy, T dz, T k, double dt, Domain const& d, DPath<T>& p)
{
    T const speed = speed_t(k);
    T const norm = sqrt(dx * dx + dy * dy + dz * dz);
    if (!finite(val(speed)) || !finite(val(norm)) || !(val(norm) > 0.0)) return false;
    T const scale = speed * dt / norm;
    T const ex = x + scale * dx;
    T const ey = y + scale * dy;
    T const ez = z + scale * dz;
    T const sr = hypot(x, y), st = atan2(y, x), er = hypot(ex, ey), et = atan2(ey, ex);
    p.sx = sr * cos(st);
    p.sy = sr * sin(st);
    p.sz = z;
    p.ex = er * cos(et);
    p.ey = er * sin(et);
    p.ez = ez;
    T const ddx = p.ex - p.sx, ddy = p.ey - p.sy, ddz = p.ez - p.sz;
    p.full = sqrt(ddx * ddx + ddy * ddy + ddz * ddz);
    // the clip's exit candidates, on values: any exit ends the drift
    double const scl = dmax(1.0, dmax(d.r_hi, dmax(fabs(d.z_lo), fabs(d.z_hi))));
    double const tol = 64.0 * kEps * scl;
    double const vsr = val(sr), vsz = val(z);
    if (vsr < d.r_lo - tol || vsr > d.r_hi + tol || vsz < d.z_lo - tol || vsz > d.z_hi + tol) return false;
    double const vdx = val(ddx), vdy = val(ddy), vdz = val(ddz);
    if (vdz > 0.0 && (d.z_hi - vsz) / vdz < 1.0) return false;
    if (vdz < 0.0 && (d.z_lo - vsz) / vdz < 1.0) return false;
    double const a = vdx * vdx + vdy * vdy;
    if (a > 0.0) {
        double const psx = val(p.sx), psy = val(p.sy);
        double const b = 2.0 * (psx * vdx + psy * vdy);
        double const c = psx * psx + psy * psy - d.r_hi * d.r_hi;
        double const disc = b * b - 4.0 * a * c;
        if (!(disc < 0.0)) {
            double const root = sqrt(dmax(disc, 0.0));
            double const inv_2a = 0.5 / a;
            double const roots[2] = {(-b - root) * inv_2a, (-b + root) * inv_2a};
            for (int q = 0; q < 2; ++q) {
                double const cand = roots[q];
                if (cand <= 0.0 || cand > 1.0) continue;
                if (2.0 * a * cand + b > 0.0 && cand < 1.0) return false;
            }
        }
    }
    if (!finite(val(p.full)) || !(val(p.full) > 0.0)) return false;
    // clipped_escapes: the chord end must lie in the half-open domain
    double const er2 = ::hypot(val(p.ex), val(p.ey)), ez2 = val(p.ez);
    if (!finite(er2) || er2 < d.r_lo || er2 >= d.r_hi || ez2 < d.z_lo || ez2 >= d.z_hi) return false;
    return true;
}

template <class T>
__device__ __forceinline__ void gather_t(DevFields const& F, T r, T z, T& er, T& ez, T& bt, GatherStencil* st)
{
    int32_t const nr = F.nr, nz = F.nz;
    T const fx = (r - F.r_lo) / F.dr + 0.0 - 0.5;
    T const fz = (z - F.z_lo) / F.dz + 0.0 - 0.5;
    int32_t const ir0 = static_cast<int32_t>(floor(val(fx)));
    int32_t const iz0 = static_cast<int32_t>(floor(val(fz)));
    T const tx = fx - static_cast<double>(ir0);
    T const tz = fz - static_cast<double>(iz0);
    int32_t const i0 = max(0, min(nr - 1, ir0));
    int32_t const i1 = max(0, min(nr - 1, ir0 + 1));
    int32_t const j0 = max(0, min(nz - 1, iz0));
    int32_t const j1 = max(0, min(nz - 1, iz0 + 1));
    double const first_center = F.r_lo + 0.5 * F.dr;
    T const axis = val(r) >= first_center ? T(1.0) : tmax(0.0, (r - F.r_lo) / (0.5 * F.dr));
    int32_t const c00 = i0 + nr * j0, c01 = i0 + nr * j1, c10 = i1 + nr * j0, c11 = i1 + nr * j1;
    auto bilinear = [&](const double* c) {
        double const f00 = c[c00], f01 = c[c01], f10 = c[c10], f11 = c[c11];
        return (1.0 - tx) * ((1.0 - tz) * f00 + tz * f01) + tx * ((1.0 - tz) * f10 + tz * f11);
    };
    er = bilinear(F.cell_er) * axis;
    ez = bilinear(F.cell_ez);
    if (st) {
        double const a = val(tx), b = val(tz), s = val(axis);
        st->c[0] = c00; st->c[1] = c01; st->c[2] = c10; st->c[3] = c11;
        double const w[4] = {(1.0 - a) * (1.0 - b), (1.0 - a) * b, a * (1.0 - b), a * b};
        for (int q = 0; q < 4; ++q) { st->w_ez[q] = w[q]; st->w_er[q] = w[q] * s; }
    }
    if (val(r) <= 0.0) {
        bt = T(0.0);
    } else {
        T frr = r / F.dr;
        frr = val(frr) < 0.0 ? T(0.0) : (val(frr) > static_cast<double>(nr) ? T(static_cast<double>(nr)) : frr);
        T fzz = (z - F.z_lo) / F.dz;
        fzz = val(fzz) < 0.0 ? T(0.0) : (val(fzz) > static_cast<double>(nz) ? T(static_cast<double>(nz)) : fzz);
        int32_t const a0 = min(static_cast<int32_t>(floor(val(frr))), nr - 1);
        int32_t const b0 = min(static_cast<int32_t>(floor(val(fzz))), nz - 1);
        int32_t const a1 = a0 + 1, b1 = b0 + 1;
        T const trr = frr - static_cast<double>(a0), tzz = fzz - static_cast<double>(b0);
        int32_t const w = nr + 1;
        bt = (1.0 - trr) * ((1.0 - tzz) * F.btheta[a0 + w * b0] + tzz * F.btheta[a0 + w * b1])
            + trr * ((1.0 - tzz) * F.btheta[a1 + w * b0] + tzz * F.btheta[a1 + w * b1]);
    }
}

template <class T>
struct DField { T work, ex, ey, ez, bx, by, bz, length; };

// path_field_work over the whole in-domain chord [0, 1] (GL2)
template <class T>
__device__ DField<T> field_work_t(DevFields const& F, DPath<T> const& p, Domain const& d, double q, bool& ok,
                                  GatherStencil* st)
{
    DField<T> res{T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0), T(0.0)};
    T const pdx = (p.ex - p.sx) / p.full;
    T const pdy = (p.ey - p.sy) / p.full;
    T const pdz = (p.ez - p.sz) / p.full;
    double const pb = 0.0, pe = 1.0;
    double const mid = 0.5 * (pb + pe);
    double const half = 0.5 * (pe - pb);
    double const off = half * kGl2Node;
    T s[2][7];
    for (int node = 0; node < 2; ++node) {
        double const f = dmax(0.0, dmin(1.0, node == 0 ? mid - off : mid + off));
        T const x = p.sx + f * (p.ex - p.sx);
        T const y = p.sy + f * (p.ey - p.sy);
        T z = p.sz + f * (p.ez - p.sz) + 0.0;
        z = tmax(d.z_lo, tmin(d.z_hi, z));
        T const r = hypot(x, y);
        T rad(0.0), rc(1.0), rs(0.0);
        if (val(r) > 0.0) { rad = (pdx * x + pdy * y) / r; rc = x / r; rs = y / r; }
        T er, ez, bt;
        gather_t(F, r, z, er, ez, bt, st ? st + node : nullptr);
        FieldSlots<T>::add(er, kStateSlots + 2 * node);
        FieldSlots<T>::add(ez, kStateSlots + 2 * node + 1);
        T const wk = q * (er * rad + ez * pdz) * 1.0;
        if (!finite(val(wk)) || !finite(val(er)) || !finite(val(ez)) || !finite(val(bt))) ok = false;
        s[node][0] = wk; s[node][1] = er * rc; s[node][2] = er * rs; s[node][3] = ez;
        s[node][4] = -bt * rs; s[node][5] = bt * rc; s[node][6] = T(0.0);
    }
    T integ[7];
    for (int c = 0; c < 7; ++c) integ[c] = half * (s[0][c] + s[1][c]);
    res.work = res.work + p.full * integ[0];
    res.ex = res.ex + p.full * integ[1];
    res.ey = res.ey + p.full * integ[2];
    res.ez = res.ez + p.full * integ[3];
    res.bx = res.bx + p.full * integ[4];
    res.by = res.by + p.full * integ[5];
    res.bz = res.bz + p.full * integ[6];
    res.length = res.length + p.full * (pe - pb);
    if (val(res.length) > 0.0) {
        res.ex = res.ex / res.length; res.ey = res.ey / res.length; res.ez = res.ez / res.length;
        res.bx = res.bx / res.length; res.by = res.by / res.length; res.bz = res.bz / res.length;
    }
    return res;
}

template <class T>
__device__ __forceinline__ void half_kick_t(DField<T> const& f, double half_dt, double q, T& k, T& dx, T& dy, T& dz)
{
    T const umag = proper_velocity_t(k);
    double const impulse = q * (kC * kC / kMeC2) * half_dt;
    T ux = umag * dx + impulse * f.ex;
    T uy = umag * dy + impulse * f.ey;
    T uz = umag * dz + impulse * f.ez;
    {
        T const gamma = sqrt(1.0 + (ux * ux + uy * uy + uz * uz) / (kC * kC));
        T const tx = impulse * f.bx / (2.0 * gamma);
        T const ty = impulse * f.by / (2.0 * gamma);
        T const tz = impulse * f.bz / (2.0 * gamma);
        T const t2 = tx * tx + ty * ty + tz * tz;
        T const sx = 2.0 * tx / (1.0 + t2);
        T const sy = 2.0 * ty / (1.0 + t2);
        T const sz = 2.0 * tz / (1.0 + t2);
        T const px = ux + uy * tz - uz * ty;
        T const py = uy + uz * tx - ux * tz;
        T const pz = uz + ux * ty - uy * tx;
        T const ox = ux + py * sz - pz * sy;
        T const oy = uy + pz * sx - px * sz;
        T const oz = uz + px * sy - py * sx;
        ux = ox; uy = oy; uz = oz;
    }
    T const n = sqrt(ux * ux + uy * uy + uz * uz);
    if (val(n) > 0.0 && finite(val(n))) { dx = ux / n; dy = uy / n; dz = uz / n; }
    T const hw = 0.5 * f.work;
    k = tmax(0.0, k + hw);
}

// RreaInteractionTable::Interpolate of a charged column at q
template <class T>
__device__ __forceinline__ T charged_value_t(DevTables const& Tb, int table, T q, bool& ok)
{
    Bracket b;
    if (!charged_bracket(Tb, val(q), b)) { ok = false; return T(0.0); }
    const double* y = Tb.cy[table];
    if (b.kind == 1) return T(y[0]);
    if (b.kind == 2) return T(y[Tb.nc - 1]);
    if (b.kind == 3) return T(y[b.hi]);
    int32_t const mode = Tb.cmode[table];
    double const ylo = y[b.lo], yhi = y[b.hi];
    if (mode == 3) {
        T const lin_t = (q - Tb.cx[b.lo]) / (Tb.cx[b.hi] - Tb.cx[b.lo]);
        return ylo + lin_t * (yhi - ylo);
    }
    T const t = (log(q) - Tb.clogx[b.lo]) / (Tb.clogx[b.hi] - Tb.clogx[b.lo]);
    bool const positive = ylo > 0.0 && yhi > 0.0;
    if (mode == 2 || (mode == 1 && !positive)) return ylo + t * (yhi - ylo);
    if (!positive) { ok = false; return T(0.0); }
    const double* ly = Tb.clogy[table];
    return exp(ly[b.lo] + t * (ly[b.hi] - ly[b.lo]));
}

// The net-chord deposit over T end points into a sink of T face values.
template <class T, class Sink>
__device__ void deposit_piece_t(DevGrid const& g, Sink& sink, double qdt, T xr0, T xz0, T xr1, T xz1)
{
    T fr_flux = qdt * (xr1 - xr0);
    T const fz_flux = qdt * (xz1 - xz0);
    T const mr = 0.5 * (xr0 + xr1);
    T const mz = 0.5 * (xz0 + xz1);
    {
        int32_t const jf = static_cast<int32_t>(floor(val(mz)));
        T const fz = mz - static_cast<double>(jf);
        int32_t const fi = static_cast<int32_t>(floor(val(mr))) + 1;
        if (fi <= 0) fr_flux = T(0.0);
        T const wz0 = 1.0 - fz;
        if (fi > 0) {
            int32_t const i = fi > g.nr ? g.nr : fi;
            for (int b = 0; b < 2; ++b) {
                int32_t j = jf + b;
                if (j < 0) j = 0;
                if (j >= g.nz) j = g.nz - 1;
                sink.jr(i, j, fr_flux * (b == 0 ? wz0 : fz));
            }
        }
    }
    {
        int32_t const ifc = static_cast<int32_t>(floor(val(mr)));
        T const fr = mr - static_cast<double>(ifc);
        int32_t fj = static_cast<int32_t>(floor(val(mz))) + 1;
        T const wr0 = 1.0 - fr;
        if (fj < 0) fj = 0;
        if (fj > g.nz) fj = g.nz;
        for (int a = 0; a < 2; ++a) {
            int32_t i = ifc + a;
            if (i < 0) i = -1 - i;
            if (i >= g.nr) i = g.nr - 1;
            if (i < 0) i = 0;
            sink.jz(i, fj, fz_flux * (a == 0 ? wr0 : fr));
        }
    }
}

// One axis' cell-boundary crossings in increasing s, s = (k - a) / (b - a).
template <class T>
struct CutStreamT {
    T a, b;
    double k, k_end, dk;
    bool active;
    __device__ void init(T a_, T b_)
    {
        a = a_;
        b = b_;
        active = val(a) != val(b);
        if (!active) return;
        double const lo = dmin(val(a), val(b)), hi = dmax(val(a), val(b));
        if (val(b) > val(a)) { k = ceil(lo); k_end = floor(hi); dk = 1.0; }
        else { k = floor(hi); k_end = ceil(lo); dk = -1.0; }
        advance_to_valid();
    }
    __device__ bool in_range() const { return active && (dk > 0.0 ? k <= k_end : k >= k_end); }
    __device__ T value() const { return (k - a) / (b - a); }
    __device__ void advance_to_valid()
    {
        while (in_range()) {
            double const s = val(value());
            if (s > 0.0 && s < 1.0) return;
            k += dk;
        }
    }
    __device__ T peek() const { return in_range() ? value() : T(2.0); }
    __device__ void pop() { k += dk; advance_to_valid(); }
};

template <class T, class Sink>
__device__ bool deposit_segment_t(DevGrid const& g, Sink& sink, double q, double dt, T r0, T z0, T r1, T z1)
{
    double const qdt = q / dt;
    T const xr0 = (r0 - g.r_lo) / g.dr - 0.5;
    T const xz0 = (z0 - g.z_lo) / g.dz - 0.5;
    T const xr1 = (r1 - g.r_lo) / g.dr - 0.5;
    T const xz1 = (z1 - g.z_lo) / g.dz - 0.5;
    double const lim_r = static_cast<double>(g.nr) + 2.0, lim_z = static_cast<double>(g.nz) + 2.0;
    if (!(fabs(val(xr0)) <= lim_r && fabs(val(xr1)) <= lim_r && fabs(val(xz0)) <= lim_z
          && fabs(val(xz1)) <= lim_z)) {
        return false;
    }
    CutStreamT<T> cr, cz;
    cr.init(xr0, xr1);
    cz.init(xz0, xz1);
    T s0(0.0);
    for (;;) {
        T const a = cr.peek(), b = cz.peek();
        T s1;
        if (val(a) <= val(b) && val(a) < 2.0) { s1 = a; cr.pop(); }
        else if (val(b) < 2.0) { s1 = b; cz.pop(); }
        else { s1 = T(1.0); }
        if (val(s1) > val(s0)) {
            deposit_piece_t(g, sink, qdt, xr0 + s0 * (xr1 - xr0), xz0 + s0 * (xz1 - xz0), xr0 + s1 * (xr1 - xr0),
                            xz0 + s1 * (xz1 - xz0));
        }
        s0 = s1;
        if (val(a) >= 2.0 && val(b) >= 2.0) break;
    }
    return true;
}

// Sinks of the drift deposit: device atomics into the J grid, the dot product
// with loss weights (dL/dJ), or both.
struct DriftSinks {
    double* jr;        // may be null
    double* jz;
    const double* gjr; // may be null
    const double* gjz;
};
template <class T>
struct DriftSink {
    DevGrid const& g;
    DriftSinks const& s;
    T acc;
    __device__ void jr(int32_t i, int32_t j, T v)
    {
        int64_t const f = i + static_cast<int64_t>(g.nr + 1) * j;
        if constexpr (std::is_same<T, double>::value) { if (s.jr) atomic_add(&s.jr[f], v); }
        if (s.gjr) acc = acc + v * s.gjr[f];
    }
    __device__ void jz(int32_t i, int32_t j, T v)
    {
        int64_t const f = i + static_cast<int64_t>(g.nr) * j;
        if constexpr (std::is_same<T, double>::value) { if (s.jz) atomic_add(&s.jz[f], v); }
        if (s.gjz) acc = acc + v * s.gjz[f];
    }
};

struct DriftEnv {
    DevFields F;
    Domain D;
    DevGrid G;          // the padded net-chord grid's index space
    double fallback_ratio;
    double q;           // charge number: -1 e-, +1 e+
    int32_t col, soft;  // stopping-power columns
    double dt, cutoff;
};

// One drift step of one particle of weight w; false when its drift ends.
template <class T>
__device__ bool drift_step(DriftEnv const& E, DevTables const& Tb, double w, DState<T>& s, DriftSink<T>& sink,
                           GatherStencil* st)
{
    Domain const& D = E.D;
    if (!(w > 0.0) || val(s.r) > D.r_hi + 1.0e-9 * D.r_hi) return false;
    T const pr = s.r, pz = s.z;
    T dx = s.ux, dy = s.uy, dz = s.uz;
    T k = kinetic_t(dx, dy, dz);
    normalize_t(dx, dy, dz);
    if (!(val(k) >= E.cutoff)) return false;
    T const x0 = s.r * cos(s.th);
    T const y0 = s.r * sin(s.th);
    DPath<T> p;
    if (!chord_t(x0, y0, s.z, dx, dy, dz, k, E.dt, D, p)) return false;
    // build_interval over the single piece: the density ratio at its midpoint
    int32_t err = 0;
    T ratio;
    {
        T const zm = tmax(D.z_lo, tmin(D.z_hi, p.sz + 0.5 * (p.ez - p.sz) + 0.0));
        T rho;
        if (E.F.density_mode == 0) {
            rho = T(E.F.density_ratio);
        } else {
            double const alt_v = E.F.altitude_at_z0 + val(zm);
            int32_t const n = E.F.prof_n;
            const double* a = E.F.prof_alt;
            if (!(alt_v >= a[0]) || !(alt_v <= a[n - 1])) return false;
            int32_t const first = guided_upper_bound(a, n, alt_v, E.F.prof_guide, alt_v);
            T v;
            if (first == 0) v = T(E.F.prof_rho[0]);
            else if (first == n) v = T(E.F.prof_rho[n - 1]);
            else {
                int32_t const lo = first - 1;
                double const span = a[first] - a[lo];
                T const alt = E.F.altitude_at_z0 + zm;
                T const t = span > 0.0 ? (alt - a[lo]) / span : T(0.0);
                v = (1.0 - t) * E.F.prof_rho[lo] + t * E.F.prof_rho[first];
            }
            rho = v / E.F.reference_density;
        }
        T const dl = p.full * (1.0 - 0.0) * (val(rho) > 0.0 ? rho : T(E.fallback_ratio));
        T const piece_len = p.full * (1.0 - 0.0);
        T const rr = dl / piece_len;
        ratio = val(rr) > 0.0 ? rr : T(E.fallback_ratio);
    }
    (void)err;
    bool ok = true;
    DField<T> const fr = field_work_t(E.F, p, D, E.q, ok, st);
    if (!ok) return false;
    double const substep_dt = E.dt * (1.0 - 0.0);
    T k1 = k, d1x = dx, d1y = dy, d1z = dz;
    half_kick_t(fr, 0.5 * substep_dt, E.q, k1, d1x, d1y, d1z);
    if (!(val(k1) >= E.cutoff)) return false;
    T const L = p.full * (1.0 - 0.0) * ratio;
    T const col = charged_value_t(Tb, E.col, k1, ok) * 1.0;
    T const soft = charged_value_t(Tb, E.soft, k1, ok) * 1.0;
    if (!ok) return false;
    T const total = col + soft;
    T const to_cut = k1 - E.cutoff;
    T const req = total * L;
    if (!finite(val(req)) || val(req) < 0.0 || !(val(req) < val(to_cut)) || !(val(k1) > 0.0)) return false;
    T const raw_col = col * L;
    T const raw_soft = soft * L;
    T const avail = tmax(0.0, k1 - E.cutoff);
    T const drag = tmin(raw_col + raw_soft, avail);
    T k2 = tmax(0.0, k1 - drag);
    if (!(val(k2) >= E.cutoff)) return false;
    half_kick_t(fr, 0.5 * substep_dt, E.q, k2, d1x, d1y, d1z);
    if (!(val(k2) >= E.cutoff)) return false;
    // the end point at the full chord fraction
    T const x = p.sx + 1.0 * (p.ex - p.sx);
    T const y = p.sy + 1.0 * (p.ey - p.sy);
    T const z = p.sz + 1.0 * (p.ez - p.sz);
    if (!in_half_open(val(x), val(y), val(z), D)) return false;
    T const end_r = hypot(x, y);
    if (!deposit_segment_check(E.G, val(pr), val(pz), val(end_r), val(z))) return false;
    T const um = proper_velocity_t(k2);
    s.ux = um * d1x; s.uy = um * d1y; s.uz = um * d1z;
    s.r = end_r;
    s.th = atan2(y, x);
    s.z = z;
    deposit_segment_t(E.G, sink, E.q * w, E.dt, pr, pz, end_r, z);
    return true;
}

}  // namespace rrea_dev
