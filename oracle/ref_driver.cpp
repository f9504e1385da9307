// Thin C wrappers around the reference's own AMReX-free sources, compiled
// from the reference checkout by `make -C oracle ref` into oracle/_ref/.
// Test infrastructure only: used to pin the oracle's restatement of the
// counter RNG (OV/RreaRng.cpp), the uniform stream
// (NA/include/rrea/RreaUniformStream.H), the two-point Gauss-Legendre rule
// (NA/include/rrea/RreaGaussLegendre8.H), the RZ segment deposition and its
// conversion to physical face currents (NA/include/rrea/
// RreaSegmentDepositionRZ.H) and the TM Maxwell leapfrog
// (NA/include/rrea/RreaMaxwellTMRZ.H).
#include "RreaRng.H"
#include "rrea/RreaCicConservingResample.H"
#include "rrea/RreaGaussLegendre8.H"
#include "rrea/RreaMaxwellTMRZ.H"
#include "rrea/RreaSegmentDepositionRZ.H"
#include "rrea/RreaUniformStream.H"

#include <array>
#include <cstdint>
#include <cstring>
#include <vector>

extern "C" {

uint64_t ref_rng_counter(uint64_t seed, uint64_t pid, uint64_t step, uint64_t idx, uint64_t chan)
{
    return rrea::warpx::RreaRng::Counter(rrea::warpx::RreaRngKey{seed, pid, step, idx, chan});
}

double ref_rng_uniform(uint64_t seed, uint64_t pid, uint64_t step, uint64_t idx, uint64_t chan)
{
    return rrea::warpx::RreaRng::Uniform01(rrea::warpx::RreaRngKey{seed, pid, step, idx, chan});
}

void ref_uniform_stream(double u1, double u2, int64_t n, double* out)
{
    rrea::RreaUniformStream s(u1, u2);
    for (int64_t i = 0; i < n; ++i) out[i] = s.Next();
}

// GL2 integral of the cubic c0 + c1 f + c2 f^2 + c3 f^3 over [a, b].
double ref_gl2_cubic(double a, double b, const double* c)
{
    auto const r = rrea::RreaGaussLegendre2IntegrateArray<double, 1>(a, b, [&](double f) {
        return std::array<double, 1>{c[0] + f * (c[1] + f * (c[2] + f * c[3]))};
    });
    return r[0];
}

static rrea::SegmentGridRZ make_grid(int nr, int nz, double dr, double dz, double r_lo, double z_lo,
                                     const double* rho, const double* jr, const double* jz)
{
    rrea::SegmentGridRZ g;
    g.nr = nr; g.nz = nz; g.dr = dr; g.dz = dz; g.r_lo = r_lo; g.z_lo = z_lo;
    g.resize();
    std::memcpy(g.rho.data(), rho, g.rho.size() * sizeof(double));
    std::memcpy(g.jr.data(), jr, g.jr.size() * sizeof(double));
    std::memcpy(g.jz.data(), jz, g.jz.size() * sizeof(double));
    return g;
}

static void copy_back(rrea::SegmentGridRZ const& g, double* rho, double* jr, double* jz)
{
    std::memcpy(rho, g.rho.data(), g.rho.size() * sizeof(double));
    std::memcpy(jr, g.jr.data(), g.jr.size() * sizeof(double));
    std::memcpy(jz, g.jz.data(), g.jz.size() * sizeof(double));
}

void ref_deposit_segment(int nr, int nz, double dr, double dz, double r_lo, double z_lo, double* rho, double* jr,
                         double* jz, double q, double dt, double r0, double z0, double r1, double z1)
{
    auto g = make_grid(nr, nz, dr, dz, r_lo, z_lo, rho, jr, jz);
    rrea::DepositSegmentRZ(g, q, dt, r0, z0, r1, z1);
    copy_back(g, rho, jr, jz);
}

int ref_deposit_escaping_segment(int nr, int nz, double dr, double dz, double r_lo, double z_lo, double* rho,
                                 double* jr, double* jz, double q, double dt, const double* a7, int pnr, int pzb,
                                 int pnz, uint32_t faces)
{
    auto g = make_grid(nr, nz, dr, dz, r_lo, z_lo, rho, jr, jz);
    auto const res = rrea::DepositEscapingSegmentRZ(g, q, dt, a7[0], a7[1], a7[2], a7[3], a7[4], a7[5], pnr, pzb,
                                                    pnz, faces);
    copy_back(g, rho, jr, jz);
    return res.valid ? 1 : 0;
}

void ref_current_to_physical(int nr, int nz, double dr, double dz, int pad, double q, const double* jr_index,
                             const double* jz_index, double* jr, double* jz)
{
    std::vector<double> const a(jr_index, jr_index + static_cast<size_t>(nr + pad + 1) * (nz + 2 * pad));
    std::vector<double> const b(jz_index, jz_index + static_cast<size_t>(nr + pad) * (nz + 2 * pad + 1));
    std::vector<double> out_r, out_z;
    rrea::ConvertPaddedIndexCurrentToPhysicalRZ(nr, nz, dr, dz, pad, q, a, b, out_r, out_z);
    std::memcpy(jr, out_r.data(), out_r.size() * sizeof(double));
    std::memcpy(jz, out_z.data(), out_z.size() * sizeof(double));
}

static rrea::MaxwellTMRZ make_maxwell(int nr, int nz, double dr, double dz, const double* er, const double* ez,
                                      const double* bt)
{
    rrea::MaxwellTMRZ f;
    f.nr = nr; f.nz = nz; f.dr = dr; f.dz = dz;
    f.resize();
    std::memcpy(f.er.data(), er, f.er.size() * sizeof(double));
    std::memcpy(f.ez.data(), ez, f.ez.size() * sizeof(double));
    std::memcpy(f.btheta.data(), bt, f.btheta.size() * sizeof(double));
    return f;
}

static const std::vector<double>* opt(std::vector<double>& v, const double* p, size_t n)
{
    if (!p) return nullptr;
    v.assign(p, p + n);
    return &v;
}

// op: 0 MaxwellStep, 1 MaxwellAdvanceB, 2 MaxwellSilverMueller, 3 MaxwellAdvanceE
void ref_maxwell(int op, int nr, int nz, double dr, double dz, double* er, double* ez, double* bt, double dt,
                 const double* sr, const double* sz, const double* jr, const double* jz, const double* aer,
                 const double* aez)
{
    auto f = make_maxwell(nr, nz, dr, dz, er, ez, bt);
    size_t const ner = f.er.size(), nez = f.ez.size();
    std::vector<double> vsr, vsz, vjr, vjz, vaer, vaez;
    auto const* psr = opt(vsr, sr, ner);
    auto const* psz = opt(vsz, sz, nez);
    auto const* pjr = opt(vjr, jr, ner);
    auto const* pjz = opt(vjz, jz, nez);
    auto const* paer = opt(vaer, aer, ner);
    auto const* paez = opt(vaez, aez, nez);
    if (op == 0) rrea::MaxwellStep(f, dt, psr, psz, pjr, pjz, paer, paez);
    else if (op == 1) rrea::MaxwellAdvanceB(f, dt);
    else if (op == 2) rrea::MaxwellSilverMueller(f);
    else rrea::MaxwellAdvanceE(f, dt, psr, psz, pjr, pjz, paer, paez);
    std::memcpy(er, f.er.data(), ner * sizeof(double));
    std::memcpy(ez, f.ez.data(), nez * sizeof(double));
    std::memcpy(bt, f.btheta.data(), f.btheta.size() * sizeof(double));
}

// out: field energy (ambient-relative when aer/aez given), Joule power (jr/jz
// given), boundary Poynting power, and the Jz moment
void ref_maxwell_diagnostics(int nr, int nz, double dr, double dz, const double* er, const double* ez,
                             const double* bt, const double* aer, const double* aez, const double* jr,
                             const double* jz, double* out)
{
    auto f = make_maxwell(nr, nz, dr, dz, er, ez, bt);
    rrea::MaxwellTMRZ amb = f;
    bool const has_amb = aer && aez;
    if (has_amb) {
        std::memcpy(amb.er.data(), aer, amb.er.size() * sizeof(double));
        std::memcpy(amb.ez.data(), aez, amb.ez.size() * sizeof(double));
        std::fill(amb.btheta.begin(), amb.btheta.end(), 0.0);
    }
    out[0] = rrea::MaxwellFieldEnergy(f, has_amb ? &amb : nullptr);
    out[1] = 0.0;
    out[3] = 0.0;
    if (jr && jz) {
        std::vector<double> const vjr(jr, jr + f.er.size()), vjz(jz, jz + f.ez.size());
        out[1] = rrea::MaxwellJoulePower(f, has_amb ? &amb : nullptr, vjr, vjz, &out[3]);
    }
    out[2] = rrea::MaxwellBoundaryPoyntingPower(f, has_amb ? &amb : nullptr);
}

// The single-rank flow of the adaptive_resample_v1 thinning plan
// (RreaCouplingPopulation.cpp:321-520 with ranks = 1) over records whose
// support, phase and fractions the caller formed: groups by support in
// (support_r, support_z, phase, candidate, id) order, exact removal capacity
// per support, the greedy request per support in support order, and
// RreaPlanGlobalCicThinning per requested support with the controller's
// counter draw.  out_w receives every record's new weight; counts: global
// capacity, geometric floor, nominal request, removed.  Returns 0, or the
// failing step (1 invalid plan, 2 insufficient progress).
int ref_population_thin(int64_t n, const int32_t* sr, const int32_t* sz, const int32_t* phase, const double* fr,
                        const double* fz, const double* w, const uint8_t* candidate, const uint64_t* high,
                        const uint64_t* low, int64_t target, uint64_t seed, uint64_t salt, uint64_t step,
                        int species, double* out_w, int64_t* counts)
{
    std::vector<rrea::RreaGlobalCicParticle> owned(static_cast<size_t>(n));
    for (int64_t i = 0; i < n; ++i) {
        auto& p = owned[static_cast<size_t>(i)];
        p.support_r = sr[i]; p.support_z = sz[i]; p.phase = phase[i];
        p.source_rank = 0;
        p.stable_id = low[i];
        p.stable_id_high = high ? high[i] : 0;
        p.candidate = candidate ? candidate[i] != 0 : false;
        p.fraction_r = fr[i]; p.fraction_z = fz[i]; p.weight = w[i];
        out_w[i] = w[i];
    }
    std::vector<size_t> order(owned.size());
    for (size_t i = 0; i < order.size(); ++i) order[i] = i;
    auto less = [&](size_t a, size_t b) {
        auto const& l = owned[a];
        auto const& r = owned[b];
        if (l.support_r != r.support_r) return l.support_r < r.support_r;
        if (l.support_z != r.support_z) return l.support_z < r.support_z;
        if (l.phase != r.phase) return l.phase < r.phase;
        if (l.candidate != r.candidate) return l.candidate < r.candidate;
        if (l.stable_id_high != r.stable_id_high) return l.stable_id_high < r.stable_id_high;
        return l.stable_id < r.stable_id;
    };
    std::stable_sort(order.begin(), order.end(), less);
    uint64_t const stream_key = species == 0 ? 0x656c656374726f6eULL : 0x706f736974726f6eULL;
    struct G { size_t begin, end, capacity, floor; };
    std::vector<G> groups;
    int64_t capacity = 0, floor = 0;
    for (size_t b = 0; b < order.size();) {
        size_t e = b + 1;
        while (e < order.size() && owned[order[e]].support_r == owned[order[b]].support_r
               && owned[order[e]].support_z == owned[order[b]].support_z) ++e;
        std::vector<rrea::RreaCicResampleParticle> support;
        for (size_t i = b; i < e; ++i) {
            auto const& p = owned[order[i]];
            support.push_back({p.fraction_r, p.fraction_z, p.weight});
        }
        size_t const cap = rrea::RreaCicResampleRemovalCapacity(support);
        groups.push_back(G{b, e, cap, support.size() - cap});
        capacity += static_cast<int64_t>(cap);
        floor += static_cast<int64_t>(support.size() - cap);
        b = e;
    }
    bool const due = n > target + target / 4;
    int64_t const nominal = due ? std::min<int64_t>(n - target, capacity) : 0;
    int64_t remaining = nominal, removed = 0;
    rrea::warpx::RreaAdaptiveResampleSpecies const sp =
        species == 0 ? rrea::warpx::RreaAdaptiveResampleSpecies::Electron
                     : rrea::warpx::RreaAdaptiveResampleSpecies::Positron;
    int status = 0;
    for (auto const& g : groups) {
        int64_t const request = std::min<int64_t>(remaining, static_cast<int64_t>(g.capacity));
        remaining -= request;
        if (request == 0) continue;
        std::vector<rrea::RreaGlobalCicParticle> support;
        for (size_t i = g.begin; i < g.end; ++i) support.push_back(owned[order[i]]);
        auto plan = rrea::RreaPlanGlobalCicThinning<double>(
            support, support.size() - static_cast<size_t>(request), 1, stream_key,
            [&](std::uint64_t group_id, std::size_t reduction) {
                return rrea::warpx::RreaAdaptiveResampleCicUniform(seed, group_id ^ salt, step, reduction, sp);
            });
        if (!plan.valid() || plan.removed == 0) { status = 1; break; }
        removed += static_cast<int64_t>(plan.removed);
        for (auto const& u : plan.updates) {
            for (size_t i = g.begin; i < g.end; ++i) {
                auto const& p = owned[order[i]];
                if (p.stable_id == u.stable_id && p.stable_id_high == u.stable_id_high && p.candidate == u.candidate) {
                    out_w[order[i]] = u.stored_weight;
                    break;
                }
            }
        }
    }
    if (status == 0 && nominal > 0 && removed < nominal) status = 2;
    counts[0] = capacity; counts[1] = floor; counts[2] = nominal; counts[3] = removed;
    return status;
}

double ref_maxwell_mixed_magnetic_energy(int nr, int nz, double dr, double dz, const double* bt,
                                         const double* bt_previous)
{
    std::vector<double> zr(static_cast<size_t>(nr + 1) * nz, 0.0), zz(static_cast<size_t>(nr) * (nz + 1), 0.0);
    auto f = make_maxwell(nr, nz, dr, dz, zr.data(), zz.data(), bt);
    std::vector<double> const prev(bt_previous, bt_previous + f.btheta.size());
    return rrea::MaxwellMixedMagneticEnergy(f, prev);
}

}  // extern "C"
