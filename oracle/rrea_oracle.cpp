// CPU oracle -- TEST INFRASTRUCTURE ONLY.
//
// A plain C++17 restatement of the reference's per-step electron transport:
// the algorithm of RreaParticleInteraction::Apply's electron phase
// (native_amrex_rrea/warpx_overlay/Source/Rrea/RreaElectronTransport.cpp:9-902)
// with every primitive it calls, each function citing the reference lines it
// follows.  It is the checker of the HIP path, never part of it: the shipped
// library does not link this file, and only tests/, the smoke check and the
// CPU-baseline leg of bench.py load it.
//
// Abbreviations in citations:
//   OV/ = native_amrex_rrea/warpx_overlay/Source/Rrea/
//   NA/ = native_amrex_rrea/
//
// Scope notes (documented, deliberate):
//   * non-periodic z only -- the reference's net-chord deposit refuses periodic
//     z (OV/RreaWarpXCoupling.cpp:420-424), so the production step never runs
//     with it;
//   * r_lo = 0 (the RZ axis), the reference's startup contract;
//   * C&D plane diagnostics are off (the reference default).
//
// Compile with -ffp-contract=off so that no multiply-add is fused.

#include "rrea_oracle.h"

#include <algorithm>
#include <array>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

constexpr double kMeC2 = 510998.95000;           // NA/include/rrea/RreaConstants.H:9
constexpr double kC = 299792458.0;               // RreaConstants.H:10
constexpr double kWAir = 33.97;                  // RreaConstants.H:30
constexpr double kTwoPi = 6.283185307179586476925286766559;

struct OracleError : std::runtime_error {
    int code;
    OracleError(int c, const char* what) : std::runtime_error(what), code(c) {}
};

[[noreturn]] inline void fail(int code, const char* what) { throw OracleError(code, what); }

// ---------------------------------------------------------------------------
// Counter RNG -- OV/RreaRng.cpp:9-44
// ---------------------------------------------------------------------------
inline uint64_t splitmix64(uint64_t v)
{
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30U)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27U)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31U);
}

inline uint64_t rng_counter(uint64_t seed, uint64_t pid, uint64_t step, uint64_t idx, uint64_t chan)
{
    uint64_t s = splitmix64(seed);
    s ^= splitmix64(pid + 0x100000001b3ULL);
    s ^= splitmix64(step + 0x9e3779b97f4a7c15ULL);
    s ^= splitmix64(idx + 0xbf58476d1ce4e5b9ULL);
    s ^= splitmix64(chan + 0x94d049bb133111ebULL);
    return splitmix64(s);
}

inline double rng_uniform(uint64_t seed, uint64_t pid, uint64_t step, uint64_t idx, uint64_t chan)
{
    return (static_cast<double>(rng_counter(seed, pid, step, idx, chan) >> 12U) + 0.5) * 0x1.0p-52;
}

// NA/include/rrea/RreaUniformStream.H:21-53
struct UniformStream {
    uint64_t state;
    static uint64_t rotl(uint64_t x, int r) { return (x << r) | (x >> (64 - r)); }
    UniformStream(double u1, double u2)
    {
        uint64_t b1, b2;
        std::memcpy(&b1, &u1, 8);
        std::memcpy(&b2, &u2, 8);
        state = splitmix64(b1) ^ rotl(splitmix64(b2), 32);
    }
    double next()
    {
        state += 0x9E3779B97F4A7C15ULL;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z = z ^ (z >> 31);
        return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

// RNG channel ids -- OV/RreaRng.H:16-98 (the subset the charged steps use)
enum : uint64_t {
    kChHardMollerEnergy = 22,
    kChHardMollerAzimuth = 23,
    kChBremsPhotonEnergy = 42,
    kChBremsPhotonAngle = 43,
    kChBremsPhotonAzimuth = 44,
    kChGsAngle = 89,
    kChGsAzimuth = 90,
    kChGsFewCount = 91,
    kChGsFewScatter = 92,
    kChGsFewAzimuth = 93,
    kChCompetingOpticalDepth = 107,
    kChCompetingChannel = 108,
    kChContinuationOpticalDepth = 124,
    kChContinuationChannel = 125,
    kChPositronBhabhaEnergy = 74,
    kChPositronBremsPhotonEnergy = 76,
    kChPositronBremsPhotonAngle = 77,
    kChPositronBremsPhotonAzimuth = 78,
    kChPositronAnnihilationPolar = 80,
    kChPositronAnnihilationAzimuth = 81,
    kChPositronGsAngle = 94,
    kChPositronGsAzimuth = 95,
    kChPositronGsFewCount = 96,
    kChPositronGsFewScatter = 97,
    kChPositronGsFewAzimuth = 98,
    kChPositronCompetingOpticalDepth = 109,
    kChPositronCompetingChannel = 110,
    kChPositronContinuationOpticalDepth = 126,
    kChPositronContinuationChannel = 127,
    kChPositronBhabhaAzimuth = 130,
    kChPhotonInteraction = 51,
    kChPhotonChannel = 52,
    kChComptonRecoil = 53,
    kChComptonScatter = 54,
    kChPhotoelectric = 55,
    kChPairAngle = 57,
    kChComptonScatterAzimuth = 58,
    kChPairAzimuth = 59,
    kChPhotoelectricAngle = 60,
    kChPhotoelectricAzimuth = 65,
    kChPairInvariant = 66,
    kChPairLeptonPolar = 67,
    kChPairLeptonAzimuth = 68,
    kChPhotonContinuationOpticalDepth = 128,
    kChPhotonContinuationChannel = 129,
};

// RreaSecondaryProcess ids, OV/RreaParticleInteraction.H:34-48
enum : int32_t {
    kProcHardMoller = 2, kProcHardBhabha = 3, kProcBrems = 4, kProcCompton = 5, kProcPhotoelectric = 6,
    kProcPairNuclear = 7, kProcPairTriplet = 8, kProcAnnihilation = 9,
};

// OV/RreaParticleInteraction.H:642-710
constexpr uint64_t kLineageCap = uint64_t{1} << 14;
constexpr uint64_t kSegmentCap = uint64_t{1} << 14;
constexpr uint64_t kEventCap = uint64_t{1} << 13;
constexpr uint64_t kInnerCap = uint64_t{1} << 8;

inline bool pack_transient_index(bool newborn, uint64_t species, uint64_t lineage, uint64_t segment,
                                 uint64_t event, uint64_t inner, uint64_t& out)
{
    if (species >= 3U || lineage >= kLineageCap || segment >= kSegmentCap || event >= kEventCap
        || inner >= kInnerCap) {
        return false;
    }
    out = (static_cast<uint64_t>(newborn) << 51) | (species << 49) | (lineage << 35) | (segment << 21)
        | (event << 8) | inner;
    return true;
}

inline bool live_hazard_index(uint64_t species, uint64_t segment, uint64_t event, uint64_t& out)
{
    if (species >= 3U) return false;
    if (event == 0U) {
        if (segment >= kSegmentCap) return false;
        out = segment;
        return true;
    }
    return pack_transient_index(false, species, 1U, segment, event, 0U, out);
}

// ---------------------------------------------------------------------------
// Schema-6 tables -- NA/src/RreaInteractionTables.cpp
// ---------------------------------------------------------------------------
enum Mode { LogLog = 0, LogLogZero = 1, LogXLinear = 2, Linear = 3 };

struct RateTable {
    std::vector<double> x, y, logx, logy;
    int mode = LogLog;

    // Finalize :1632-1657
    void finalize()
    {
        if (mode == Linear) return;
        logx.resize(x.size());
        logy.resize(y.size());
        for (size_t i = 0; i < x.size(); ++i) {
            logx[i] = std::log(x[i]);
            logy[i] = y[i] > 0.0 ? std::log(y[i]) : 0.0;
        }
    }

    // Interpolate :1580-1630
    double interpolate(double q) const
    {
        if (!std::isfinite(q) || q < x.front() || q > x.back()) {
            fail(ORC_ERR_TABLE_RANGE, "interaction table query outside certified range");
        }
        auto const upper = std::lower_bound(x.begin(), x.end(), q);
        if (upper == x.begin()) return y.front();
        if (upper == x.end()) return y.back();
        size_t const hi = static_cast<size_t>(upper - x.begin());
        size_t const lo = hi - 1U;
        if (x[hi] == q) return y[hi];
        if (mode == Linear) {
            double const t = (q - x[lo]) / (x[hi] - x[lo]);
            return y[lo] + t * (y[hi] - y[lo]);
        }
        double const t = (std::log(q) - logx[lo]) / (logx[hi] - logx[lo]);
        bool const positive = y[lo] > 0.0 && y[hi] > 0.0;
        if (mode == LogXLinear || (mode == LogLogZero && !positive)) {
            return y[lo] + t * (y[hi] - y[lo]);
        }
        if (!positive) fail(ORC_ERR_INVALID, "loglog_linear requires positive bracket");
        return std::exp(logy[lo] + t * (logy[hi] - logy[lo]));
    }
};

struct CdfGroups {
    std::vector<double> primary;
    std::vector<std::vector<double>> u, v;
};

struct AngleTable {
    std::vector<double> e, n, u, omc;  // omc flattened (ie*np + ip)*nu + iu
};

struct ElasticTable {
    std::vector<double> e, mfp, eta, trmfp;
};

enum RateId {
    kECol = 0, kESoft, kEMoller, kEBrems, kEEltr,
    kPCol, kPSoft, kPBhabha, kPBrems, kPEltr, kPAnnih,
    kGCompton, kGPe, kGPairN, kGPairT, kRateCount
};

// interp_semilogx_strict :624-648
inline double interp_semilogx_strict(std::vector<double> const& x, std::vector<double> const& y, double q)
{
    if (x.size() < 2) return y.empty() ? 0.0 : y.front();
    if (!std::isfinite(q) || q < x.front() || q > x.back()) {
        fail(ORC_ERR_TABLE_RANGE, "elastic-table query outside range");
    }
    if (q == x.front()) return y.front();
    if (q == x.back()) return y.back();
    auto const upper = std::lower_bound(x.begin(), x.end(), q);
    size_t const hi = static_cast<size_t>(upper - x.begin());
    size_t const lo = hi - 1U;
    if (x[hi] == q) return y[hi];
    double const t = (std::log(q) - std::log(x[lo])) / (std::log(x[hi]) - std::log(x[lo]));
    return y[lo] + t * (y[hi] - y[lo]);
}

// log_bracket :807-829
inline void log_bracket(std::vector<double> const& g, double q, int& i, double& frac)
{
    int const last = static_cast<int>(g.size()) - 1;
    if (!std::isfinite(q) || q < g[0] || q > g[last]) {
        fail(ORC_ERR_TABLE_RANGE, "elastic-angle query outside range");
    }
    if (q == g[0]) { i = 0; frac = 0.0; return; }
    if (q == g[last]) { i = last - 1; frac = 1.0; return; }
    int lo = 0, hi = last;
    while (hi - lo > 1) { int const m = (lo + hi) / 2; if (g[m] <= q) lo = m; else hi = m; }
    i = lo;
    double const log_lo = std::log(g[lo]);
    double const log_hi = std::log(g[hi]);
    if (log_hi == log_lo) { frac = (q == g[lo]) ? 0.0 : 1.0; return; }
    frac = (std::log(q) - log_lo) / (log_hi - log_lo);
}

// sample_gs_angle_cos :831-891
inline double sample_gs_angle_cos(AngleTable const& t, double energy, double path_n, double u)
{
    int const np = static_cast<int>(t.n.size());
    int const nu = static_cast<int>(t.u.size());
    int ie = 0, ip = 0;
    double te = 0.0, tn = 0.0;
    log_bracket(t.e, energy, ie, te);
    log_bracket(t.n, path_n, ip, tn);
    auto cos_at_node = [&](int iu) {
        auto val = [&](int e_i, int p_i) { return t.omc[(static_cast<size_t>(e_i) * np + p_i) * nu + iu]; };
        double const la = std::log(val(ie, ip));
        double const lb = std::log(val(ie + 1, ip));
        double const lc = std::log(val(ie, ip + 1));
        double const ld = std::log(val(ie + 1, ip + 1));
        double const lx = (1 - te) * (1 - tn) * la + te * (1 - tn) * lb + (1 - te) * tn * lc + te * tn * ld;
        return 1.0 - std::exp(lx);
    };
    auto omc_at = [&](int iu) { return std::max(1.0 - cos_at_node(iu), std::numeric_limits<double>::min()); };
    int const first = 1;
    int const last = nu - 2;
    double one_minus_cos;
    if (last <= first) {
        one_minus_cos = omc_at(first);
    } else if (u <= t.u[first]) {
        double const d1 = omc_at(first);
        double const d2 = omc_at(first + 1);
        double const u1 = t.u[first];
        double const u2 = t.u[first + 1];
        double const slope = std::log(d2 / d1) / std::log(u2 / u1);
        double const u_eff = std::max(u, u1 * 1.0e-6);
        one_minus_cos = d1 * std::exp(slope * std::log(u_eff / u1));
    } else if (u >= t.u[last]) {
        one_minus_cos = omc_at(last);
    } else {
        int lo = first, hi = last;
        while (hi - lo > 1) { int const m = (lo + hi) / 2; if (t.u[m] <= u) lo = m; else hi = m; }
        double const d0 = omc_at(lo);
        double const d1 = omc_at(hi);
        double const fu = std::log(u / t.u[lo]) / std::log(t.u[hi] / t.u[lo]);
        one_minus_cos = d0 * std::exp(fu * std::log(d1 / d0));
    }
    double const cos_val = 1.0 - std::min(2.0, one_minus_cos);
    return std::min(1.0, std::max(-1.0, cos_val));
}

// sample_conditional_cdf :995-1069
inline double sample_cdf_group(std::vector<double> const& gu, std::vector<double> const& gv, double cdf_u)
{
    auto const upper = std::lower_bound(gu.begin(), gu.end(), cdf_u);
    if (upper == gu.begin()) return gv.front();
    if (upper == gu.end()) return gv.back();
    size_t const hi = static_cast<size_t>(upper - gu.begin());
    size_t const lo = hi - 1U;
    double const denom = gu[hi] - gu[lo];
    double const t = denom > 0.0 ? (cdf_u - gu[lo]) / denom : 0.0;
    return gv[lo] + t * (gv[hi] - gv[lo]);
}

inline double sample_conditional_cdf(CdfGroups const& g, double primary, double cdf_u, bool log_interp)
{
    cdf_u = std::max(0.0, std::min(1.0, cdf_u));
    if (!std::isfinite(primary) || primary < g.primary.front() || primary > g.primary.back()) {
        fail(ORC_ERR_TABLE_RANGE, "conditional-CDF primary energy outside range");
    }
    size_t const n = g.primary.size();
    if (primary == g.primary.front()) return sample_cdf_group(g.u[0], g.v[0], cdf_u);
    if (primary == g.primary.back()) return sample_cdf_group(g.u[n - 1], g.v[n - 1], cdf_u);
    auto const upper = std::lower_bound(g.primary.begin(), g.primary.end(), primary);
    if (upper == g.primary.begin()) return sample_cdf_group(g.u[0], g.v[0], cdf_u);
    if (upper == g.primary.end()) return sample_cdf_group(g.u[n - 1], g.v[n - 1], cdf_u);
    size_t const hi = static_cast<size_t>(upper - g.primary.begin());
    size_t const lo = hi - 1U;
    if (g.primary[hi] == primary) return sample_cdf_group(g.u[hi], g.v[hi], cdf_u);
    double const y0 = sample_cdf_group(g.u[lo], g.v[lo], cdf_u);
    double const y1 = sample_cdf_group(g.u[hi], g.v[hi], cdf_u);
    double const t = (std::log(primary) - std::log(g.primary[lo])) / (std::log(g.primary[hi]) - std::log(g.primary[lo]));
    if (log_interp && y0 > 0.0 && y1 > 0.0) return std::exp(std::log(y0) + t * (std::log(y1) - std::log(y0)));
    return y0 + t * (y1 - y0);
}

// mean_conditional_cdf :1071-1110
inline double cdf_group_mean(std::vector<double> const& gu, std::vector<double> const& gv)
{
    double mean = gu.front() * gv.front();
    for (size_t i = 1; i < gu.size(); ++i) mean += 0.5 * (gv[i - 1] + gv[i]) * (gu[i] - gu[i - 1]);
    mean += (1.0 - gu.back()) * gv.back();
    return mean;
}

inline double mean_conditional_cdf(CdfGroups const& g, double primary)
{
    if (g.primary.empty() || primary < g.primary.front() || primary > g.primary.back()) {
        fail(ORC_ERR_TABLE_RANGE, "conditional-CDF mean query outside range");
    }
    size_t const n = g.primary.size();
    auto const upper = std::lower_bound(g.primary.begin(), g.primary.end(), primary);
    if (upper == g.primary.begin()) return cdf_group_mean(g.u[0], g.v[0]);
    if (upper == g.primary.end()) return cdf_group_mean(g.u[n - 1], g.v[n - 1]);
    size_t const hi = static_cast<size_t>(upper - g.primary.begin());
    if (g.primary[hi] == primary) return cdf_group_mean(g.u[hi], g.v[hi]);
    size_t const lo = hi - 1U;
    double const y0 = cdf_group_mean(g.u[lo], g.v[lo]);
    double const y1 = cdf_group_mean(g.u[hi], g.v[hi]);
    double const t = (std::log(primary) - std::log(g.primary[lo])) / (std::log(g.primary[hi]) - std::log(g.primary[lo]));
    return y0 > 0.0 && y1 > 0.0 ? std::exp(std::log(y0) + t * (std::log(y1) - std::log(y0))) : y0 + t * (y1 - y0);
}

struct Tables {
    RateTable rate[kRateCount];
    CdfGroups moller, bhabha, ebrems, pbrems;
    AngleTable gs_e, gs_p, wz_e, wz_p;
    ElasticTable el_e, el_p;
    std::vector<double> hm_mean_secondary;  // on the hard-Moller rate grid
    double hard_moller_threshold = 0.0;
    double brems_cutoff = 0.0;
    double reference_density = 0.0;
    double charged_min = 0.0, charged_max = 0.0, photon_min = 0.0, photon_max = 0.0;
    double gs_few = 10.0, gs_iso = 1.0e5, gs_transition = 0.0;
    // photoelectric subshell CDF groups (load_photoelectric_subshell_cdf :893-990)
    std::vector<double> pe_primary;
    std::vector<int64_t> pe_start;
    std::vector<double> pe_u, pe_binding;
    // pair final-state bank, training blocks by (channel, energy); records hold
    // the pair electron, positron and triplet recoil (fraction, direction)
    std::vector<int32_t> pair_channel;   // 1 nuclear, 2 triplet
    std::vector<double> pair_energy;
    std::vector<int64_t> pair_start;
    std::vector<double> pair_frac;       // 3 per record
    std::vector<double> pair_dir;        // 9 per record

    void require_charged(double e) const
    {
        if (!std::isfinite(e) || e < charged_min || e > charged_max) {
            fail(ORC_ERR_TABLE_RANGE, "schema-6 charged sampler energy outside certified range");
        }
    }

    // SampleMollerSecondaryEnergy :2584-2653
    double sample_moller(double primary, double cdf_u) const
    {
        require_charged(primary);
        if (!std::isfinite(primary) || primary < moller.primary.front() || primary > moller.primary.back()) {
            fail(ORC_ERR_TABLE_RANGE, "Moller CDF primary energy outside range");
        }
        cdf_u = std::max(0.0, std::min(1.0, cdf_u));
        size_t const n = moller.primary.size();
        auto const upper = std::lower_bound(moller.primary.begin(), moller.primary.end(), primary);
        if (upper == moller.primary.begin()) return sample_cdf_group(moller.u[0], moller.v[0], cdf_u);
        if (upper == moller.primary.end()) return sample_cdf_group(moller.u[n - 1], moller.v[n - 1], cdf_u);
        size_t const hi = static_cast<size_t>(upper - moller.primary.begin());
        size_t const lo = hi - 1U;
        if (moller.primary[hi] == primary) return sample_cdf_group(moller.u[hi], moller.v[hi], cdf_u);
        double const y0 = sample_cdf_group(moller.u[lo], moller.v[lo], cdf_u);
        double const y1 = sample_cdf_group(moller.u[hi], moller.v[hi], cdf_u);
        double const x0 = std::log(moller.primary[lo]);
        double const x1 = std::log(moller.primary[hi]);
        double const t = (std::log(primary) - x0) / (x1 - x0);
        return std::exp(std::log(y0) + t * (std::log(y1) - std::log(y0)));
    }

    // Load :2146-2179 (hard-Moller mean secondary by 4096-point midpoint quadrature)
    void build_hm_mean()
    {
        RateTable const& r = rate[kEMoller];
        hm_mean_secondary.assign(r.x.size(), 0.0);
        int const samples = 4096;
        for (size_t i = 0; i < r.x.size(); ++i) {
            if (!(r.y[i] > 0.0)) continue;
            double mean = 0.0;
            for (int s = 0; s < samples; ++s) {
                mean += sample_moller(r.x[i], (static_cast<double>(s) + 0.5) / samples);
            }
            hm_mean_secondary[i] = mean / static_cast<double>(samples);
        }
    }

    // ElectronHardMollerEnergyLoss :2414-2440
    double hard_moller_energy_loss(double e, double density_ratio) const
    {
        RateTable const& r = rate[kEMoller];
        double const inv = r.interpolate(e);
        if (!(inv > 0.0)) return 0.0;
        double const mean = interp_semilogx_strict(r.x, hm_mean_secondary, e);
        if (!(mean > 0.0)) return 0.0;
        return mean * inv * density_ratio;
    }

    // ElectronTrackedBremsEnergyLoss :2678-2696
    double tracked_brems_energy_loss(double e, double density_ratio) const
    {
        double const inv = rate[kEBrems].interpolate(e) * density_ratio;
        if (!(inv > 0.0)) return 0.0;
        return inv * mean_conditional_cdf(ebrems, e);
    }

    // ElasticMeanFreePath :2477-2493
    double elastic_mfp(bool positron, double e, double density_ratio) const
    {
        ElasticTable const& t = positron ? el_p : el_e;
        return interp_semilogx_strict(t.e, t.mfp, e) / density_ratio;
    }

    // ScreeningEta :2495-2502
    double screening_eta(bool positron, double e) const
    {
        ElasticTable const& t = positron ? el_p : el_e;
        return interp_semilogx_strict(t.e, t.eta, e);
    }

    // SampleGsCosTheta :2504-2544
    double sample_gs_cos(bool positron, double e, double path_n, double cdf_u) const
    {
        require_charged(e);
        if (!std::isfinite(cdf_u) || cdf_u < 0.0 || cdf_u > 1.0) fail(ORC_ERR_INVALID, "GS cdf out of [0,1]");
        bool const use_gs = e <= gs_transition;
        AngleTable const& t = positron ? (use_gs ? gs_p : wz_p) : (use_gs ? gs_e : wz_e);
        if (!std::isfinite(e) || e < t.e.front() || e > t.e.back()) fail(ORC_ERR_TABLE_RANGE, "GS energy outside branch");
        if (!std::isfinite(path_n) || path_n < gs_few) fail(ORC_ERR_INVALID, "GS path below multiple-scattering branch");
        if (path_n >= gs_iso) return 2.0 * cdf_u - 1.0;
        return sample_gs_angle_cos(t, e, path_n, cdf_u);
    }

    // SampleSingleElasticCosTheta :2546-2560
    double sample_single_elastic_cos(bool positron, double e, double cdf_u) const
    {
        require_charged(e);
        double const eta = screening_eta(positron, e);
        double const u = std::min(std::max(cdf_u, 0.0), 1.0 - 1.0e-15);
        return 1.0 - 2.0 * eta * u / (1.0 + eta - u);
    }

    // SampleBremsstrahlungPhotonEnergy :2698-2712
    double sample_brems_energy(double primary, double cdf_u) const
    {
        require_charged(primary);
        return sample_conditional_cdf(ebrems, primary, cdf_u, true);
    }

    // SamplePositronBremsstrahlungPhotonEnergy :2562-2575
    double sample_positron_brems_energy(double primary, double cdf_u) const
    {
        require_charged(primary);
        return sample_conditional_cdf(pbrems, primary, cdf_u, true);
    }

    // SamplePositronBhabhaSecondaryEnergy :3180-3194
    double sample_bhabha(double primary, double cdf_u) const
    {
        require_charged(primary);
        return sample_conditional_cdf(bhabha, primary, cdf_u, true);
    }

    // SampleBremsstrahlungCosTheta :2714-2760 (bounded ModifiedTsai by bisection)
    double sample_brems_cos(double primary, double cdf_u) const
    {
        require_charged(primary);
        cdf_u = std::max(0.0, std::min(1.0, cdf_u));
        if (cdf_u == 0.0) return 1.0;
        if (cdf_u == 1.0) return -1.0;
        constexpr double a = 0.625;
        auto g2 = [](double x) {
            if (x < 1.0e-3) {
                return x * x * (0.5 + x * (-1.0 / 3.0 + x * (1.0 / 8.0 + x * (-1.0 / 30.0 + x * (1.0 / 144.0)))));
            }
            return -std::expm1(-x) - x * std::exp(-x);
        };
        auto mix = [&](double u) { return 0.25 * g2(a * u) + 0.75 * g2(3.0 * a * u); };
        double const u_max = 2.0 * (1.0 + primary / kMeC2);
        double const target = cdf_u * mix(u_max);
        double lo = 0.0, hi = u_max;
        for (int it = 0; it < 60; ++it) {
            double const mid = 0.5 * (lo + hi);
            if (mix(mid) < target) lo = mid; else hi = mid;
        }
        double const f = 0.5 * (lo + hi) / u_max;
        return 1.0 - 2.0 * f * f;
    }
};

// ---------------------------------------------------------------------------
// Geometry: the RZ chord -- NA/src/RreaTransportPath.cpp, NA/include/rrea/RreaTransportPath.H
// ---------------------------------------------------------------------------
struct Piece { double begin = 0, end = 0, length = 0, shift = 0; };

struct Path {
    double sx = 0, sy = 0, sz = 0, ex = 0, ey = 0, ez = 0;
    double full_length = 0, begin_f = 0, end_f = 0, in_length = 0;
    bool exits = false;
    double r_lo = 0, r_hi = 0, z_lo = 0, z_hi = 0;
    std::vector<Piece> pieces;

    // PointAtFullPathFraction :42-53 (WrapZ is the identity when not periodic)
    void point_at_full(double f, double& x, double& y, double& z) const
    {
        f = std::max(0.0, std::min(1.0, f));
        x = sx + f * (ex - sx);
        y = sy + f * (ey - sy);
        z = sz + f * (ez - sz);
    }
    // PointAtPieceFraction :55-72
    void point_at_piece(Piece const& p, double f, double& x, double& y, double& z) const
    {
        f = std::max(p.begin, std::min(p.end, f));
        x = sx + f * (ex - sx);
        y = sy + f * (ey - sy);
        z = sz + f * (ez - sz) + p.shift;
        z = std::max(z_lo, std::min(z_hi, z));
    }
    // OutsideCertifiedDomain :28-40
    bool outside_certified(double r, double z) const
    {
        if (!std::isfinite(r) || !std::isfinite(z)) return true;
        if (r < r_lo || r >= r_hi) return true;
        return z < z_lo || z >= z_hi;
    }
};

// ClipRzTransportPath :74-293, non-periodic branch
inline Path clip_rz_path(double sr, double st, double sz, double er_, double et, double ez_,
                         double r_lo, double r_hi, double z_lo, double z_hi)
{
    auto finite = [](double v) { return std::isfinite(v); };
    if (!finite(sr) || !finite(st) || !finite(sz) || !finite(er_) || !finite(et) || !finite(ez_)
        || sr < 0.0 || er_ < 0.0 || r_lo < 0.0 || !(r_lo < r_hi) || !(z_lo < z_hi)) {
        fail(ORC_ERR_INVALID, "ClipRzTransportPath received invalid geometry");
    }
    Path p;
    p.r_lo = r_lo; p.r_hi = r_hi; p.z_lo = z_lo; p.z_hi = z_hi;
    p.sx = sr * std::cos(st);
    p.sy = sr * std::sin(st);
    p.sz = sz;
    p.ex = er_ * std::cos(et);
    p.ey = er_ * std::sin(et);
    p.ez = ez_;
    double const dx = p.ex - p.sx;
    double const dy = p.ey - p.sy;
    double const dz = p.ez - p.sz;
    p.full_length = std::sqrt(dx * dx + dy * dy + dz * dz);
    p.begin_f = 0.0;
    double const scale = std::max(1.0, std::max(r_hi, std::max(std::abs(z_lo), std::abs(z_hi))));
    double const tol = 64.0 * std::numeric_limits<double>::epsilon() * scale;
    if (sr < r_lo - tol || sr > r_hi + tol || sz < z_lo - tol || sz > z_hi + tol) {
        fail(ORC_ERR_OUTSIDE_DOMAIN, "previous particle position outside the RZ domain");
    }
    double exit_f = 1.0;
    auto retain = [&](double c) { if (c >= 0.0 && c < exit_f) exit_f = c; };
    if (dz > 0.0) retain((z_hi - sz) / dz);
    else if (dz < 0.0) retain((z_lo - sz) / dz);
    double const a = dx * dx + dy * dy;
    if (a > 0.0) {
        double const b = 2.0 * (p.sx * dx + p.sy * dy);
        auto cyl = [&](double radius, bool outer) {
            double const c = p.sx * p.sx + p.sy * p.sy - radius * radius;
            double const disc = b * b - 4.0 * a * c;
            if (disc < 0.0) return;
            double const root = std::sqrt(std::max(disc, 0.0));
            double const inv_2a = 0.5 / a;
            double const roots[2] = {(-b - root) * inv_2a, (-b + root) * inv_2a};
            for (double cand : roots) {
                if (cand <= 0.0 || cand > 1.0) continue;
                double const deriv = 2.0 * a * cand + b;
                bool const leaves = outer ? deriv > 0.0 : deriv < 0.0;
                if (leaves) retain(cand);
            }
        };
        cyl(r_hi, true);
        if (r_lo > 0.0) cyl(r_lo, false);
    }
    p.end_f = std::max(0.0, std::min(1.0, exit_f));
    p.in_length = p.full_length * (p.end_f - p.begin_f);
    p.exits = p.end_f < 1.0 - 64.0 * std::numeric_limits<double>::epsilon();
    // Non-periodic: the breakpoints are {begin, end}; one piece when end > begin
    // (the unique() merge at :258-270 cannot fire on two distinct entries that
    // the piece loop would keep).
    double const b0 = p.begin_f, b1 = p.end_f;
    bool merged = std::abs(b0 - b1) <= 128.0 * std::numeric_limits<double>::epsilon()
                                       * std::max(1.0, std::max(std::abs(b0), std::abs(b1)));
    if (!merged && b1 > b0) {
        p.pieces.push_back(Piece{b0, b1, p.full_length * (b1 - b0), 0.0});
    }
    return p;
}

struct Chord { Path path; bool valid = false; };

// RreaChargedSpeedFromKineticEnergy OV/RreaParticleInteraction.H:92-105
inline double charged_speed(double k)
{
    if (!std::isfinite(k) || k < 0.0) return std::numeric_limits<double>::quiet_NaN();
    double const gamma = 1.0 + k / kMeC2;
    double const beta2 = std::max(0.0, 1.0 - 1.0 / (gamma * gamma));
    return kC * std::sqrt(beta2);
}

// RreaBuildRunningChargedChord OV/RreaParticleInteraction.H:110-161
inline Chord build_chord(double x, double y, double z, double dx, double dy, double dz, double k,
                         double req_dt, double outer_dt, double r_lo, double r_hi, double z_lo, double z_hi)
{
    Chord c;
    if (!std::isfinite(req_dt) || !std::isfinite(outer_dt) || !(req_dt >= 0.0) || !(outer_dt > 0.0)) return c;
    double const speed = charged_speed(k);
    double const norm = std::sqrt(dx * dx + dy * dy + dz * dz);
    if (!std::isfinite(speed) || !std::isfinite(norm) || !(norm > 0.0)) return c;
    double const scale = speed * req_dt / norm;
    double const ex = x + scale * dx;
    double const ey = y + scale * dy;
    double const ez = z + scale * dz;
    c.path = clip_rz_path(std::hypot(x, y), std::atan2(y, x), z, std::hypot(ex, ey), std::atan2(ey, ex), ez,
                          r_lo, r_hi, z_lo, z_hi);
    c.valid = std::isfinite(c.path.full_length) && c.path.begin_f == 0.0;
    return c;
}

struct EscapePolicy { bool escapes = false; bool invalidate = false; };

// RreaEndpointEscapePolicyForPath / RreaClippedTransportEscapePolicyForPath
// OV/RreaParticleInteraction.H:409-452
inline EscapePolicy clipped_escape_policy(Path const& p, double er_, double ez_)
{
    if (!std::isfinite(er_) || !std::isfinite(ez_)) return {true, true};
    bool const outside = p.outside_certified(er_, ez_);
    bool const escapes = p.exits || outside;
    EscapePolicy pol{escapes, escapes};
    return pol;
}

// RreaCartesianPositionInHalfOpenRzDomain :318-354
inline bool in_half_open(double x, double y, double z, double r_lo, double r_hi, double z_lo, double z_hi)
{
    if (!std::isfinite(x) || !std::isfinite(y)) return false;
    double const r = std::hypot(x, y);
    return std::isfinite(r) && std::isfinite(z) && r_hi > r_lo && z_hi > z_lo && r >= r_lo && r < r_hi
        && z >= z_lo && z < z_hi;
}

// RreaAdvanceRunningChargedPosition :368-392
inline bool advance_running_position(Path const& p, double f, double& rx, double& ry, double& rz)
{
    double x, y, z;
    p.point_at_full(f, x, y, z);
    if (!in_half_open(x, y, z, p.r_lo, p.r_hi, p.z_lo, p.z_hi)) return false;
    rx = x; ry = y; rz = z;
    return true;
}

struct BoundaryCrossing {
    bool valid = false;
    double start_r = 0, start_z = 0, cross_r = 0, cross_z = 0;
    uint32_t faces = 0;
};

// RreaChargedBoundaryCrossingForPath :468-540 (non-periodic)
inline BoundaryCrossing boundary_crossing(Path const& p)
{
    BoundaryCrossing res;
    double cx, cy;
    p.point_at_full(p.end_f, cx, cy, res.cross_z);
    res.cross_r = std::hypot(cx, cy);
    res.start_r = std::hypot(p.sx, p.sy);
    res.start_z = p.sz;
    double const scale = std::max(1.0, std::max(p.r_hi, std::max(std::abs(p.z_lo), std::abs(p.z_hi))));
    double const tol = 512.0 * std::numeric_limits<double>::epsilon() * scale;
    bool const on_r_hi = std::abs(res.cross_r - p.r_hi) <= tol;
    bool const on_z_lo = std::abs(res.cross_z - p.z_lo) <= tol;
    bool const on_z_hi = std::abs(res.cross_z - p.z_hi) <= tol;
    double const dx = p.ex - p.sx, dy = p.ey - p.sy, dz = p.ez - p.sz;
    double const outward = cx * dx + cy * dy;
    if (on_r_hi && outward > -tol * scale) { res.faces |= 1U; res.cross_r = p.r_hi; }
    if (on_z_lo && dz < 0.0) { res.faces |= 2U; res.cross_z = p.z_lo; }
    if (on_z_hi && dz > 0.0) { res.faces |= 4U; res.cross_z = p.z_hi; }
    if (res.faces == 0U) {
        if (on_r_hi) { res.faces |= 1U; res.cross_r = p.r_hi; }
        if (on_z_hi) { res.faces |= 4U; res.cross_z = p.z_hi; }
    }
    res.valid = res.faces != 0U && std::isfinite(res.start_r) && std::isfinite(res.start_z)
        && std::isfinite(res.cross_r) && std::isfinite(res.cross_z);
    return res;
}

// ---------------------------------------------------------------------------
// Coupling services: density ratio and field gather
// ---------------------------------------------------------------------------
struct Coupling {
    orc_geometry g;
    double reference_density;
    double r_hi, z_hi;

    // TransportDensityRatioAtZ OV/RreaWarpXCoupling.cpp:79-95 +
    // RreaAltitudeProfile::Interpolate NA/src/RreaAltitudeProfile.cpp:125-151
    double density_ratio_at_z(double z) const
    {
        if (g.density_mode == 0) return g.density_ratio;
        double const alt = g.altitude_at_z0_m + z;
        int const n = g.profile_n;
        double const* a = g.profile_alt_m;
        if (alt < a[0] || alt > a[n - 1]) fail(ORC_ERR_TABLE_RANGE, "altitude outside profile");
        double const* up = std::upper_bound(a, a + n, alt);
        double v;
        if (up == a) v = g.profile_rho[0];
        else if (up == a + n) v = g.profile_rho[n - 1];
        else {
            size_t const hi = static_cast<size_t>(up - a);
            size_t const lo = hi - 1U;
            double const span = a[hi] - a[lo];
            double const t = span > 0.0 ? (alt - a[lo]) / span : 0.0;
            v = (1.0 - t) * g.profile_rho[lo] + t * g.profile_rho[hi];
        }
        return v / reference_density;
    }

    // GatherTotalFieldRZ NA/src/RreaAmrexAdvance.cpp:532-579 with the stencil of
    // NA/include/rrea/RreaFieldGatherStencil.H:49-130, and MaxwellBthetaAt :511-530
    void gather(double r, double z, double& er, double& ez, double& bt) const
    {
        int const nr = g.nr, nz = g.nz;
        double const fx = (r - g.r_lo) / g.dr + 0.0 - 0.5;
        double const fz = (z - g.z_lo) / g.dz + 0.0 - 0.5;
        int const ir0 = static_cast<int>(std::floor(fx));
        int const iz0 = static_cast<int>(std::floor(fz));
        double const tx = fx - static_cast<double>(ir0);
        double const tz = fz - static_cast<double>(iz0);
        int const i0 = std::max(0, std::min(nr - 1, ir0));
        int const i1 = std::max(0, std::min(nr - 1, ir0 + 1));
        int const j0 = std::max(0, std::min(nz - 1, iz0));
        int const j1 = std::max(0, std::min(nz - 1, iz0 + 1));
        auto er_idx = [&](int i, int j) { return static_cast<size_t>(i) + static_cast<size_t>(nr + 1) * j; };
        auto ez_idx = [&](int i, int j) { return static_cast<size_t>(i) + static_cast<size_t>(nr) * j; };
        auto cell_er = [&](int i, int j) {
            double const lo = g.er[er_idx(i, j)] + g.amb_er[er_idx(i, j)];
            double const hi = g.er[er_idx(i + 1, j)] + g.amb_er[er_idx(i + 1, j)];
            return 0.5 * (lo + hi);
        };
        auto cell_ez = [&](int i, int j) {
            double const lo = g.ez[ez_idx(i, j)] + g.amb_ez[ez_idx(i, j)];
            double const hi = g.ez[ez_idx(i, j + 1)] + g.amb_ez[ez_idx(i, j + 1)];
            return 0.5 * (lo + hi);
        };
        auto bilinear = [](double f00, double f01, double f10, double f11, double tx_, double tz_) {
            return (1.0 - tx_) * ((1.0 - tz_) * f00 + tz_ * f01) + tx_ * ((1.0 - tz_) * f10 + tz_ * f11);
        };
        double const first_center = g.r_lo + 0.5 * g.dr;
        double const axis = r >= first_center ? 1.0 : std::max(0.0, (r - g.r_lo) / (0.5 * g.dr));
        er = bilinear(cell_er(i0, j0), cell_er(i0, j1), cell_er(i1, j0), cell_er(i1, j1), tx, tz) * axis;
        ez = bilinear(cell_ez(i0, j0), cell_ez(i0, j1), cell_ez(i1, j0), cell_ez(i1, j1), tx, tz);
        // MaxwellBthetaAt
        if (r <= 0.0) {
            bt = 0.0;
        } else {
            double const frr = std::clamp(r / g.dr, 0.0, static_cast<double>(nr));
            double const fzz = std::clamp((z - g.z_lo) / g.dz, 0.0, static_cast<double>(nz));
            int const a0 = std::min(static_cast<int>(std::floor(frr)), nr - 1);
            int const c0 = std::min(static_cast<int>(std::floor(fzz)), nz - 1);
            int const a1 = a0 + 1, c1 = c0 + 1;
            double const trr = frr - a0, tzz = fzz - c0;
            auto B = [&](int i, int j) { return g.btheta[er_idx(i, j)]; };
            bt = (1.0 - trr) * ((1.0 - tzz) * B(a0, c0) + tzz * B(a0, c1))
                + trr * ((1.0 - tzz) * B(a1, c0) + tzz * B(a1, c1));
        }
    }
};

// ---------------------------------------------------------------------------
// Material path -- OV/RreaMaterialPath.H
// ---------------------------------------------------------------------------
struct Interval { double begin, end, length, ratio; };

inline double positive_or(double v, double fallback) { return v > 0.0 ? v : fallback; }

// integrate_density_piece_m :60-84 + build_material_path_intervals :86-114
inline std::vector<Interval> build_intervals(Coupling const& cp, Path const& p, double fallback)
{
    std::vector<Interval> out;
    for (auto const& pc : p.pieces) {
        if (!(pc.length > 0.0)) continue;
        double dl = 0.0;
        double const b = std::max(pc.begin, pc.begin), e = std::min(pc.end, pc.end);
        if (e > b && p.full_length > 0.0) {
            double x, y, z;
            p.point_at_piece(pc, 0.5 * (b + e), x, y, z);
            dl = p.full_length * (e - b) * positive_or(cp.density_ratio_at_z(z), fallback);
        }
        out.push_back(Interval{pc.begin, pc.end, pc.length, positive_or(dl / pc.length, fallback)});
    }
    return out;
}

// density_path_m :117-137
inline double density_path(std::vector<Interval> const& iv, Path const& p, double b, double e)
{
    b = std::max(p.begin_f, b);
    e = std::min(p.end_f, e);
    if (!(e > b)) return 0.0;
    double s = 0.0;
    for (auto const& it : iv) {
        double const lb = std::max(b, it.begin), le = std::min(e, it.end);
        if (le > lb) s += p.full_length * (le - lb) * it.ratio;
    }
    return s;
}

// fraction_at_density_path_m :143-172
inline double fraction_at_density_path(std::vector<Interval> const& iv, Path const& p, double b, double e, double target)
{
    b = std::max(p.begin_f, b);
    e = std::min(p.end_f, e);
    if (!std::isfinite(target) || target < 0.0) fail(ORC_ERR_INVALID, "sampled density path invalid");
    double rem = target;
    for (auto const& it : iv) {
        double const lb = std::max(b, it.begin), le = std::min(e, it.end);
        if (!(le > lb) || !(it.ratio > 0.0)) continue;
        double const contrib = p.full_length * (le - lb) * it.ratio;
        if (rem <= contrib) return lb + rem / (p.full_length * it.ratio);
        rem -= contrib;
    }
    return e;
}

// max_uniform_subcycle_density_path_m :177-201
inline double max_uniform_slice(std::vector<Interval> const& iv, Path const& p, uint64_t n)
{
    if (n == 0U) fail(ORC_ERR_INVALID, "zero subcycles");
    double const span = p.end_f - p.begin_f;
    double const den = static_cast<double>(n);
    double m = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        m = std::max(m, density_path(iv, p, p.begin_f + static_cast<double>(i) / den * span,
                                     p.begin_f + static_cast<double>(i + 1U) / den * span));
    }
    return m;
}

// RreaFrozenStpRate :216-223
inline double frozen(double v)
{
    if (!std::isfinite(v) || v < 0.0) fail(ORC_ERR_INVALID, "frozen rate negative or non-finite");
    return v;
}

// ---------------------------------------------------------------------------
// Field work -- OV/RreaPathFieldWork.H:57-150 with the GL2 rule of
// NA/include/rrea/RreaGaussLegendre8.H (two-point node 1/sqrt(3), weights 1)
// ---------------------------------------------------------------------------
constexpr double kGl2Node = static_cast<double>(0.577350269189625764509148780501957L);

struct FieldRange { double work = 0, ex = 0, ey = 0, ez = 0, bx = 0, by = 0, bz = 0, length = 0; };

inline FieldRange path_field_work(Coupling const& cp, Path const& p, double q, double b, double e)
{
    FieldRange res;
    if (!(p.full_length > 0.0)) return res;
    b = std::max(p.begin_f, b);
    e = std::min(p.end_f, e);
    if (!(e > b)) return res;
    double const pdx = (p.ex - p.sx) / p.full_length;
    double const pdy = (p.ey - p.sy) / p.full_length;
    double const pdz = (p.ez - p.sz) / p.full_length;
    for (auto const& pc : p.pieces) {
        double const pb = std::max(b, pc.begin), pe = std::min(e, pc.end);
        if (!(pe > pb)) continue;
        auto sample = [&](double f) {
            double x, y, z;
            p.point_at_piece(pc, f, x, y, z);
            double const r = std::hypot(x, y);
            double rad_dir = 0.0, rc = 1.0, rs = 0.0;
            if (r > 0.0) { rad_dir = (pdx * x + pdy * y) / r; rc = x / r; rs = y / r; }
            double er, ez, bt;
            cp.gather(r, z, er, ez, bt);
            double const w = q * (er * rad_dir + ez * pdz) * 1.0;
            if (!std::isfinite(w) || !std::isfinite(er) || !std::isfinite(ez) || !std::isfinite(bt)) {
                fail(ORC_ERR_INVALID, "substep field-impulse gather non-finite");
            }
            return std::array<double, 7>{w, er * rc, er * rs, ez, -bt * rs, bt * rc, 0.0};
        };
        std::array<double, 7> integ{};
        double const mid = 0.5 * (pb + pe);
        double const half = 0.5 * (pe - pb);
        double const off = half * kGl2Node;
        auto const lo = sample(mid - off);
        auto const hi = sample(mid + off);
        for (int c = 0; c < 7; ++c) integ[c] = half * (lo[c] + hi[c]);
        res.work += p.full_length * integ[0];
        res.ex += p.full_length * integ[1];
        res.ey += p.full_length * integ[2];
        res.ez += p.full_length * integ[3];
        res.bx += p.full_length * integ[4];
        res.by += p.full_length * integ[5];
        res.bz += p.full_length * integ[6];
        res.length += p.full_length * (pe - pb);
    }
    if (res.length > 0.0) {
        res.ex /= res.length; res.ey /= res.length; res.ez /= res.length;
        res.bx /= res.length; res.by /= res.length; res.bz /= res.length;
    }
    return res;
}

// ---------------------------------------------------------------------------
// Kinematics -- OV/RreaChargedInterleave.H, NA/include/rrea/RreaLorentzPush.H,
// NA/include/rrea/RreaSeedSchedule.H:30-43, OV/RreaParticleInteraction.H:1494-1506
// ---------------------------------------------------------------------------
inline double proper_velocity_from_kinetic(double k)
{
    if (k <= 0.0) return 0.0;
    double const gamma = 1.0 + k / kMeC2;
    return kC * std::sqrt(std::max(gamma * gamma - 1.0, 0.0));
}

inline double kinetic_from_proper_velocity(double ux, double uy, double uz)
{
    double const u2 = ux * ux + uy * uy + uz * uz;
    double const gamma = std::sqrt(1.0 + u2 / (kC * kC));
    return u2 * kMeC2 / ((kC * kC) * (gamma + 1.0));
}

inline void normalize(double& x, double& y, double& z)
{
    double const n = std::sqrt(x * x + y * y + z * z);
    if (n <= 0.0) { x = 0.0; y = 0.0; z = 1.0; return; }
    x /= n; y /= n; z /= n;
}

// RreaBorisMagneticRotate NA/include/rrea/RreaLorentzPush.H:21-46
inline void boris_rotate(double& ux, double& uy, double& uz, double qmdt, double bx, double by, double bz)
{
    double const gamma = std::sqrt(1.0 + (ux * ux + uy * uy + uz * uz) / (kC * kC));
    double const tx = qmdt * bx / (2.0 * gamma);
    double const ty = qmdt * by / (2.0 * gamma);
    double const tz = qmdt * bz / (2.0 * gamma);
    double const t2 = tx * tx + ty * ty + tz * tz;
    double const sx = 2.0 * tx / (1.0 + t2);
    double const sy = 2.0 * ty / (1.0 + t2);
    double const sz = 2.0 * tz / (1.0 + t2);
    double const px = ux + uy * tz - uz * ty;
    double const py = uy + uz * tx - ux * tz;
    double const pz = uz + ux * ty - uy * tx;
    double const ox = ux + py * sz - pz * sy;
    double const oy = uy + pz * sx - px * sz;
    double const oz = uz + px * sy - py * sx;
    ux = ox; uy = oy; uz = oz;
}

// RreaApplyChargedFieldHalfKick OV/RreaChargedInterleave.H:100-138
inline double half_kick(FieldRange const& f, double half_dt, double q, double& k, double& dx, double& dy, double& dz)
{
    double const umag = proper_velocity_from_kinetic(k);
    double const impulse = q * (kC * kC / kMeC2) * half_dt;
    double ux = umag * dx + impulse * f.ex;
    double uy = umag * dy + impulse * f.ey;
    double uz = umag * dz + impulse * f.ez;
    boris_rotate(ux, uy, uz, impulse, f.bx, f.by, f.bz);
    double const n = std::sqrt(ux * ux + uy * uy + uz * uz);
    if (n > 0.0 && std::isfinite(n)) { dx = ux / n; dy = uy / n; dz = uz / n; }
    double const hw = 0.5 * f.work;
    k = std::max(0.0, k + hw);
    return hw;
}

// RreaTransverseBasis :309-338
inline void transverse_basis(double dx, double dy, double dz, double& b1x, double& b1y, double& b1z,
                             double& b2x, double& b2y, double& b2z)
{
    normalize(dx, dy, dz);
    double rx = 0.0, ry = 0.0, rz = 1.0;
    if (std::abs(dz) > 0.9) { rx = 1.0; rz = 0.0; }
    b1x = ry * dz - rz * dy;
    b1y = rz * dx - rx * dz;
    b1z = rx * dy - ry * dx;
    normalize(b1x, b1y, b1z);
    b2x = dy * b1z - dz * b1y;
    b2y = dz * b1x - dx * b1z;
    b2z = dx * b1y - dy * b1x;
    normalize(b2x, b2y, b2z);
}

// RreaSampleConeDirection :340-364
inline void cone_direction(double ax, double ay, double az, double ct, double uphi, double& ox, double& oy, double& oz)
{
    normalize(ax, ay, az);
    double b1x, b1y, b1z, b2x, b2y, b2z;
    transverse_basis(ax, ay, az, b1x, b1y, b1z, b2x, b2y, b2z);
    ct = std::min(1.0, std::max(-1.0, ct));
    double const st = std::sqrt(std::max(0.0, 1.0 - ct * ct));
    double const phi = kTwoPi * uphi;
    double const cp = std::cos(phi), sp = std::sin(phi);
    ox = ax * ct + st * (cp * b1x + sp * b2x);
    oy = ay * ct + st * (cp * b1y + sp * b2y);
    oz = az * ct + st * (cp * b1z + sp * b2z);
    normalize(ox, oy, oz);
}

// RreaPoissonSampleCapped :371-383
inline int poisson_capped(double mean, double u, int kmax)
{
    if (mean <= 0.0) return 0;
    double p = std::exp(-mean);
    double cum = p;
    int k = 0;
    while (u > cum && k < kmax) { ++k; p *= mean / static_cast<double>(k); cum += p; }
    return k;
}

inline double unit_cycle(double u)
{
    u -= std::floor(u);
    return u < 0.0 ? u + 1.0 : u;
}

// Moller two-body lab cosines :391-423
inline double moller_secondary_cos(double kp, double ks)
{
    kp = std::max(kp, 0.0);
    ks = std::max(ks, 0.0);
    if (kp <= 0.0 || ks <= 0.0) return 1.0;
    double const v = ks * (kp + 2.0 * kMeC2) / (kp * (ks + 2.0 * kMeC2));
    return std::sqrt(std::min(1.0, std::max(0.0, v)));
}

inline double moller_primary_cos(double kp, double ks)
{
    double const sc = std::max(0.0, kp - ks);
    if (kp <= 0.0 || sc <= 0.0) return 1.0;
    double const v = sc * (kp + 2.0 * kMeC2) / (kp * (sc + 2.0 * kMeC2));
    return std::sqrt(std::min(1.0, std::max(0.0, v)));
}

// RreaSampleMollerLabDirections :429-463
inline void moller_lab_directions(double pdx, double pdy, double pdz, double kp, double ks, double az,
                                  double& spx, double& spy, double& spz, double& sdx, double& sdy, double& sdz)
{
    normalize(pdx, pdy, pdz);
    az = unit_cycle(az);
    cone_direction(pdx, pdy, pdz, moller_secondary_cos(kp, ks), az, sdx, sdy, sdz);
    cone_direction(pdx, pdy, pdz, moller_primary_cos(kp, ks), unit_cycle(az + 0.5), spx, spy, spz);
}

// RreaApplyPhotonRecoilToChargedDirection :465-496
inline void photon_recoil(double k0, double eg, double gx, double gy, double gz, double& cx, double& cy, double& cz)
{
    eg = std::max(eg, 0.0);
    if (k0 <= 0.0 || eg <= 0.0) return;
    normalize(cx, cy, cz);
    normalize(gx, gy, gz);
    double const p0 = std::sqrt(std::max(0.0, k0 * (k0 + 2.0 * kMeC2)));
    double const px = p0 * cx - eg * gx;
    double const py = p0 * cy - eg * gy;
    double const pz = p0 * cz - eg * gz;
    double const n = std::sqrt(px * px + py * py + pz * pz);
    if (n > 0.0) { cx = px / n; cy = py / n; cz = pz / n; normalize(cx, cy, cz); }
}

// ---------------------------------------------------------------------------
// Secondary-group conservation -- RreaValidateSecondaryGroupConservation
// OV/RreaParticleInteraction.H:1099-1251: charge (e x weight) and total
// energy (kinetic + rest, weighted eV) of one interaction, input against
// output, each within 64 eps of its throughput.  A neutral target electron
// consumed by an ionization event or a positive ion brings its rest energy
// in and leaves +e; a sub-cutoff ionization, or a low-energy electron
// source, releases -e and that rest energy into the fluid.
// ---------------------------------------------------------------------------
inline double species_q(int s) { return s == 0 ? -1.0 : s == 1 ? 1.0 : 0.0; }
inline double species_m(int s) { return s == 2 ? 0.0 : kMeC2; }

struct Closure {
    // 'through' holds the particles' and sources' charge magnitudes; the
    // parent enters as |q_in| + |q_out + q_escaped|
    double in_q = 0, out_q = 0, in_e = 0, out_e = 0, through = 0;
    double parent_in_q = 0, parent_out_q = 0, escaped_q = 0;
    bool finite = true;
    void parent(int s, double w, double before, double after, bool survives)
    {
        finite = finite && std::isfinite(w) && std::isfinite(before) && std::isfinite(after) && w > 0.0
            && before >= 0.0 && after >= 0.0;
        double const q = species_q(s) * w;
        in_q += q;
        out_q += survives ? q : 0.0;
        in_e += w * (before + species_m(s));
        out_e += survives ? w * (after + species_m(s)) : 0.0;
        parent_in_q = q;
        parent_out_q = survives ? q : 0.0;
    }
    void local(double e) { finite = finite && std::isfinite(e) && e >= 0.0; out_e += e; }
    void escaped(double e, double q)
    {
        finite = finite && std::isfinite(e) && std::isfinite(q) && e >= 0.0;
        out_e += e;
        out_q += q;
        escaped_q += q;
    }
    // a group record's local and escaped terms, summed as one output
    // (RreaParticleInteraction.H:1172-1179)
    void residual_terms(double local, double esc_e, double esc_q)
    {
        finite = finite && std::isfinite(local) && std::isfinite(esc_e) && std::isfinite(esc_q) && local >= 0.0
            && esc_e >= 0.0;
        out_q += esc_q;
        escaped_q += esc_q;
        out_e += local + esc_e;
    }
    void particle(int s, double w, double e)
    {
        finite = finite && std::isfinite(w) && std::isfinite(e) && w > 0.0 && e >= 0.0;
        out_q += species_q(s) * w;
        through += std::abs(species_q(s) * w);
        out_e += w * (e + species_m(s));
    }
    void ionization(double w, double secondary, double cutoff)
    {
        finite = finite && std::isfinite(w) && std::isfinite(secondary) && w > 0.0;
        in_e += w * kMeC2;
        out_q += w;
        through += w;
        if (secondary < cutoff) { out_q -= w; through += w; out_e += w * kMeC2; }
    }
    void positive_ion(double w) { finite = finite && std::isfinite(w) && w > 0.0; in_e += w * kMeC2; out_q += w; through += w; }
    void low_electron(double w) { finite = finite && std::isfinite(w) && w > 0.0; out_q -= w; through += w; out_e += w * kMeC2; }
    // 0 valid, 1 non-finite or out of range, 2 charge, 3 energy
    int verdict(double cutoff, double* res = nullptr) const
    {
        if (!(finite && std::isfinite(cutoff) && cutoff > 0.0 && std::isfinite(in_q) && std::isfinite(out_q)
              && std::isfinite(in_e) && std::isfinite(out_e))) return 1;
        double const eps = std::numeric_limits<double>::epsilon();
        if (res) { res[0] = out_q - in_q; res[1] = out_e - in_e; }
        double const q_through = std::abs(parent_in_q) + std::abs(parent_out_q + escaped_q) + through;
        if (std::abs(out_q - in_q) > 64.0 * eps * std::max(1.0, q_through)) return 2;
        if (std::abs(out_e - in_e) > 64.0 * eps * std::max({1.0, std::abs(in_e), std::abs(out_e)})) return 3;
        return 0;
    }
};

// the transport's own check of a group it just made (fails the step)
inline void require_conserved(Closure const& c, double cutoff)
{
    if (c.verdict(cutoff) != 0) fail(ORC_ERR_INVALID, "interaction group does not conserve charge and energy");
}

// ---------------------------------------------------------------------------
// Two-photon positron annihilation -- NA/src/RreaInteractionTables.cpp
// ---------------------------------------------------------------------------
// clamp_unit / wrap_unit_interval :43-61
inline double clamp_unit(double v)
{
    if (!std::isfinite(v)) return 0.5;
    return std::max(0.0, std::min(1.0, v));
}

inline double wrap_unit_interval(double v)
{
    if (!std::isfinite(v)) return 0.0;
    v = v - std::floor(v);
    if (v < 0.0) v += 1.0;
    return v;
}

// sample_annihilation_photon_energy_fraction :63-108: the Heitler spectrum
// inverted by 64 bisections of its antiderivative a ln e - e + 1/((g+1)^2 e)
// over the centre-of-mass support [(1 - b_cm)/2, (1 + b_cm)/2].
inline double annihilation_energy_fraction(double gamma, double draw)
{
    if (gamma <= 1.0 + 1.0e-12) return 0.5;
    double const beta_cm = std::sqrt(std::max(0.0, (gamma - 1.0) / (gamma + 1.0)));
    double const eps_min = 0.5 * (1.0 - beta_cm);
    double const eps_max = 0.5 * (1.0 + beta_cm);
    double const u = clamp_unit(draw);
    double const gp1 = gamma + 1.0;
    double const a = 1.0 + 2.0 * gamma / (gp1 * gp1);
    auto anti = [a, gp1](double e) { return a * std::log(e) - e + 1.0 / (gp1 * gp1 * e); };
    double const f_lo = anti(eps_min);
    double const f_hi = anti(eps_max);
    double const span = f_hi - f_lo;
    if (!(span > 0.0)) return 0.5;
    double const target = f_lo + u * span;
    double lo = eps_min, hi = eps_max;
    for (int it = 0; it < 64; ++it) {
        double const mid = 0.5 * (lo + hi);
        if (anti(mid) < target) lo = mid; else hi = mid;
    }
    return std::max(eps_min, std::min(eps_max, 0.5 * (lo + hi)));
}

struct AnnihilationState {
    double e1 = 0.0, e2 = 0.0, cos1 = 1.0, az1 = 0.0;
    bool valid = false, at_rest = false;
};

// SamplePositronAnnihilationFinalState :3208-3253 (the energy range check,
// which admits an exact zero, is the caller's)
inline AnnihilationState sample_annihilation(double k, double draw_energy, double draw_azimuth)
{
    AnnihilationState s;
    k = std::max(0.0, k);
    double const gamma = 1.0 + k / kMeC2;
    double const total = 2.0 * kMeC2 + k;
    if (gamma <= 1.0 + 1.0e-10) {
        s.e1 = kMeC2;
        s.e2 = kMeC2;
        s.cos1 = 2.0 * clamp_unit(draw_energy) - 1.0;
        s.az1 = wrap_unit_interval(draw_azimuth);
        s.valid = true;
        s.at_rest = true;
        return s;
    }
    double const beta = std::sqrt(std::max(0.0, 1.0 - 1.0 / (gamma * gamma)));
    double const eps = annihilation_energy_fraction(gamma, draw_energy);
    s.e1 = eps * total;
    s.e2 = std::max(0.0, total - s.e1);
    if (s.e1 <= 0.0 || s.e2 <= 0.0) return s;
    double const denom = eps * gamma * beta;
    if (denom <= 0.0) return s;
    s.cos1 = std::max(-1.0, std::min(1.0, (eps * (gamma + 1.0) - 1.0) / denom));
    s.az1 = wrap_unit_interval(draw_azimuth);
    s.valid = true;
    s.at_rest = false;
    return s;
}

// ---------------------------------------------------------------------------
// Competing hazard -- OV/RreaParticleInteraction.H:1357-1492
// ---------------------------------------------------------------------------
struct Hazard { bool valid = false, event = false; int channel = -1; double sampled = 0, total = 0; };

inline Hazard select_hazard(double const taus[3], int nch, double u1, double u2)
{
    Hazard h;
    if (nch < 1 || nch > 3 || !std::isfinite(u1) || !std::isfinite(u2)) return h;
    for (int i = 0; i < nch; ++i) {
        if (!std::isfinite(taus[i]) || taus[i] < 0.0) return h;
        h.total += taus[i];
    }
    h.valid = std::isfinite(h.total);
    if (!h.valid) return h;
    u1 = std::max(0.0, std::min(1.0, u1));
    h.sampled = -std::log(std::max(1.0e-30, 1.0 - u1));
    if (!(h.sampled < h.total)) return h;
    h.event = true;
    u2 = std::max(0.0, std::min(1.0, u2));
    double const target = u2 * h.total;
    double cum = 0.0;
    int last_open = -1;
    for (int i = 0; i < nch; ++i) {
        if (taus[i] > 0.0) last_open = i;
        cum += taus[i];
        if (target < cum) { h.channel = i; return h; }
    }
    h.channel = last_open;
    h.valid = h.channel >= 0;
    h.event = h.valid;
    return h;
}

// ---------------------------------------------------------------------------
// Net-chord deposit -- NA/include/rrea/RreaSegmentDepositionRZ.H
// ---------------------------------------------------------------------------
struct Grid {
    int nr = 0, nz = 0;
    double dr = 0, dz = 0, r_lo = 0, z_lo = 0;
    double* rho = nullptr;
    double* jr = nullptr;
    double* jz = nullptr;
};

// DepositChargeCIC :68-90
inline void deposit_cic(Grid& g, double q, double r, double z)
{
    double const xr = (r - g.r_lo) / g.dr - 0.5;
    double const xz = (z - g.z_lo) / g.dz - 0.5;
    int const i0 = static_cast<int>(std::floor(xr));
    int const j0 = static_cast<int>(std::floor(xz));
    double const fr = xr - i0, fz = xz - j0;
    double const wr[2] = {1.0 - fr, fr};
    double const wz[2] = {1.0 - fz, fz};
    for (int a = 0; a < 2; ++a) {
        for (int b = 0; b < 2; ++b) {
            int i = i0 + a, j = j0 + b;
            if (i < 0) i = -1 - i;
            if (i >= g.nr) i = g.nr - 1;
            if (j < 0) j = 0;
            if (j >= g.nz) j = g.nz - 1;
            g.rho[static_cast<size_t>(i) + static_cast<size_t>(g.nr) * j] += q * wr[a] * wz[b];
        }
    }
}

// detail::deposit_piece :96-152
inline void deposit_piece(Grid& g, double qdt, double xr0, double xz0, double xr1, double xz1)
{
    double fr_flux = qdt * (xr1 - xr0);
    double const fz_flux = qdt * (xz1 - xz0);
    double const mr = 0.5 * (xr0 + xr1);
    double const mz = 0.5 * (xz0 + xz1);
    {
        int const jf = static_cast<int>(std::floor(mz));
        double const fz = mz - jf;
        int const fi = static_cast<int>(std::floor(mr)) + 1;
        if (fi <= 0) fr_flux = 0.0;
        double const wz0 = 1.0 - fz;
        for (int b = 0; b < 2; ++b) {
            int j = jf + b;
            if (j < 0) j = 0;
            if (j >= g.nz) j = g.nz - 1;
            int i = fi;
            if (i <= 0) continue;
            if (i > g.nr) i = g.nr;
            g.jr[static_cast<size_t>(i) + static_cast<size_t>(g.nr + 1) * j] += fr_flux * (b == 0 ? wz0 : fz);
        }
    }
    {
        int const ifc = static_cast<int>(std::floor(mr));
        double const fr = mr - ifc;
        int const fj = static_cast<int>(std::floor(mz)) + 1;
        double const wr0 = 1.0 - fr;
        for (int a = 0; a < 2; ++a) {
            int i = ifc + a;
            if (i < 0) i = -1 - i;
            if (i >= g.nr) i = g.nr - 1;
            int j = fj;
            if (j < 0) j = 0;
            if (j > g.nz) j = g.nz;
            g.jz[static_cast<size_t>(i) + static_cast<size_t>(g.nr) * j] += fz_flux * (a == 0 ? wr0 : fr);
        }
    }
}

// DepositSegmentRZ :159-192
inline void deposit_segment(Grid& g, double q, double dt, double r0, double z0, double r1, double z1)
{
    double const qdt = q / dt;
    double const xr0 = (r0 - g.r_lo) / g.dr - 0.5;
    double const xz0 = (z0 - g.z_lo) / g.dz - 0.5;
    double const xr1 = (r1 - g.r_lo) / g.dr - 0.5;
    double const xz1 = (z1 - g.z_lo) / g.dz - 0.5;
    std::vector<double> cuts = {0.0, 1.0};
    auto add = [&](double a, double b) {
        if (a == b) return;
        double const lo = std::min(a, b), hi = std::max(a, b);
        for (double k = std::ceil(lo); k <= std::floor(hi); ++k) {
            double const s = (k - a) / (b - a);
            if (s > 0.0 && s < 1.0) cuts.push_back(s);
        }
    };
    add(xr0, xr1);
    add(xz0, xz1);
    std::sort(cuts.begin(), cuts.end());
    for (size_t p = 0; p + 1 < cuts.size(); ++p) {
        double const s0 = cuts[p], s1 = cuts[p + 1];
        if (s1 <= s0) continue;
        deposit_piece(g, qdt, xr0 + s0 * (xr1 - xr0), xz0 + s0 * (xz1 - xz0),
                      xr0 + s1 * (xr1 - xr0), xz0 + s1 * (xz1 - xz0));
    }
}

// DepositEscapingSegmentRZ :229-389
inline bool deposit_escaping_segment(Grid& g, double q, double dt, double sr, double sz, double fr_, double fz_,
                                     double cr, double cz, int pnr, int pzb, int pnz, uint32_t faces)
{
    uint32_t constexpr known = 1U | 2U | 4U;
    if (!(dt > 0.0) || !std::isfinite(q) || !std::isfinite(dt) || !std::isfinite(sr) || !std::isfinite(sz)
        || !std::isfinite(fr_) || !std::isfinite(fz_) || !std::isfinite(cr) || !std::isfinite(cz)
        || pnr <= 0 || pnr >= g.nr || pnz <= 0 || pzb <= 0 || pzb + pnz >= g.nz || faces == 0U
        || (faces & ~known) != 0U || ((faces & 2U) != 0U && (faces & 4U) != 0U)) {
        return false;
    }
    if (faces & 1U) cr = g.r_lo + pnr * g.dr;
    if (faces & 2U) cz = g.z_lo + pzb * g.dz;
    else if (faces & 4U) cz = g.z_lo + (pzb + pnz) * g.dz;
    struct Move { size_t from, to, face; double charge; bool radial; int orient; };
    std::vector<Move> moves;
    double const xr = (cr - g.r_lo) / g.dr - 0.5;
    double const xz = (cz - g.z_lo) / g.dz - 0.5;
    int const i0 = static_cast<int>(std::floor(xr));
    int const j0 = static_cast<int>(std::floor(xz));
    double const frc = xr - i0, fzc = xz - j0;
    double const wr[2] = {1.0 - frc, frc};
    double const wz[2] = {1.0 - fzc, fzc};
    for (int a = 0; a < 2; ++a) {
        for (int b = 0; b < 2; ++b) {
            int i = i0 + a, j = j0 + b;
            if (i < 0) i = -1 - i;
            if (i < 0 || i >= g.nr || j < 0 || j >= g.nz) return false;
            bool const in_phys = i < pnr && j >= pzb && j < pzb + pnz;
            if (!in_phys) continue;
            uint32_t cf[3] = {0, 0, 0};
            int cc = 0;
            if ((faces & 1U) && i == pnr - 1) cf[cc++] = 1U;
            if ((faces & 2U) && j == pzb) cf[cc++] = 2U;
            if ((faces & 4U) && j == pzb + pnz - 1) cf[cc++] = 4U;
            if (cc == 0) return false;
            double const cell_charge = q * wr[a] * wz[b];
            double const share = cell_charge / cc;
            size_t const from = static_cast<size_t>(i) + static_cast<size_t>(g.nr) * j;
            for (int c = 0; c < cc; ++c) {
                if (cf[c] == 1U) {
                    moves.push_back(Move{from, static_cast<size_t>(i + 1) + static_cast<size_t>(g.nr) * j,
                                         static_cast<size_t>(pnr) + static_cast<size_t>(g.nr + 1) * j, share, true, 1});
                } else {
                    int const face_j = cf[c] == 2U ? pzb : pzb + pnz;
                    int const to_j = cf[c] == 2U ? j - 1 : j + 1;
                    moves.push_back(Move{from, static_cast<size_t>(i) + static_cast<size_t>(g.nr) * to_j,
                                         static_cast<size_t>(i) + static_cast<size_t>(g.nr) * face_j, share, false,
                                         cf[c] == 2U ? -1 : 1});
                }
            }
        }
    }
    if (moves.empty()) return false;
    deposit_cic(g, -q, sr, sz);
    deposit_cic(g, q, cr, cz);
    deposit_segment(g, q, dt, sr, sz, fr_, fz_);
    deposit_segment(g, q, dt, fr_, fz_, cr, cz);
    for (auto const& m : moves) {
        double const flux = m.orient * m.charge / dt;
        if (m.radial) g.jr[m.face] += flux; else g.jz[m.face] += flux;
        g.rho[m.from] -= m.charge;
        g.rho[m.to] += m.charge;
    }
    return true;
}

// ---------------------------------------------------------------------------
// Per-particle side effects (the oracle's per-chunk log)
// ---------------------------------------------------------------------------
struct DepositOp {
    int kind;  // 0 segment, 1 escape
    double a[7];
    uint32_t faces;
};

// One material-source record of the ledger (TransportSideEffects::
// Accumulate*, OV/RreaTransportLog.H:449-490): kind 0 ionization event,
// 1 positive ion only, 2 low-energy electron.
struct SourceOp {
    int kind;
    double r, z, w, secondary, loss;
};

struct ChunkLog {
    std::vector<orc_secondary> secondaries;
    std::vector<DepositOp> deposits;
    std::vector<SourceOp> sources;
    double tallies[ORC_T_COUNT] = {0};
    int error = 0;
};

// RreaFoldEventLedger OV/RreaTransportLog.H:45-127: sums add, the running
// maxima take the larger value
inline void fold_tallies(double* into, double const* from)
{
    for (int q = 0; q < ORC_T_COUNT; ++q) {
        int const l = q - ORC_LEDGER_OFFSET;
        if (l >= 0 && ORC_LEDGER_IS_MAX(l)) into[q] = std::max(into[q], from[q]);
        else into[q] += from[q];
    }
}

// The record operations of RreaEventLedger (NA/include/rrea/RreaEventLedger.H:
// 98-413) on a chunk's ledger slots; 'cutoff' is the ledger's
// low_energy_cutoff_eV.
struct Ledger {
    double* t;
    double cutoff;
    double& at(int q) { return t[ORC_LEDGER_OFFSET + q]; }
    void add(int q, double v) { at(q) += v; }
    void up(int q, double v) { at(q) = (at(q) < v) ? v : at(q); }
    static double pos(double v) { return std::max(v, 0.0); }

    void transport_ionization(double w, double loss, bool low, bool kinetic)  // :98-116
    {
        if (w <= 0.0) return;
        add(ORC_L_TRANSPORT_IONIZATION_WEIGHT, w);
        add(ORC_L_TRANSPORT_ENERGY_LOSS_EV, w * pos(loss));
        if (low) add(ORC_L_TRANSPORT_LOW_SECONDARY_WEIGHT, w);
        if (kinetic) add(ORC_L_TRANSPORT_KINETIC_SECONDARY_WEIGHT, w);
    }
    void loss_partition(double w, double drag, double col, double hard)  // :118-134
    {
        if (w <= 0.0) return;
        add(ORC_L_PARENT_DRAG_LOSS_EV, w * pos(drag));
        add(ORC_L_ION_PAIR_SOURCE_LOSS_EV, w * pos(col));
        add(ORC_L_RESTRICTED_COLLISION_LOSS_EV, w * pos(col));
        add(ORC_L_EXPLICIT_HARD_MOLLER_ENERGY_LOSS_EV, w * pos(hard));
    }
    void soft_brems(double w, double e)  // :136-143
    {
        if (w > 0.0 && e > 0.0) add(ORC_L_PHOTON_CONTINUOUS_SOFT_BREMS_ENERGY_EV, w * e);
    }
    void hm_expectation(double x)  // :145-150
    {
        if (x > 0.0) add(ORC_L_HARD_MOLLER_EXPECTED_EVENT_WEIGHT, x);
    }
    void hm_event(double w, double sec)  // :152-168
    {
        if (w <= 0.0) return;
        add(ORC_L_HARD_MOLLER_EVENT_COUNT, 1.0);
        add(ORC_L_HARD_MOLLER_EVENT_WEIGHT, w);
        if (sec >= cutoff) add(ORC_L_DIRECT_HARD_MOLLER_ELECTRON_WEIGHT, w);
        add(ORC_L_HARD_MOLLER_SECONDARY_ENERGY_WEIGHTED_SUM_EV, w * pos(sec));
        add(ORC_L_HARD_MOLLER_SECONDARY_ENERGY_WEIGHT_SUM, w);
        up(ORC_L_HARD_MOLLER_SECONDARY_ENERGY_MAX_EV, sec);
    }
    void photon_interaction(double w)  // :170-180
    {
        if (w <= 0.0) return;
        add(ORC_L_PHOTON_INTERACTION_COUNT, 1.0);
        add(ORC_L_PHOTON_INTERACTION_WEIGHT, w);
    }
    void photon_escape(double w, double e)  // :182-189
    {
        if (w > 0.0 && e > 0.0) add(ORC_L_PHOTON_ESCAPED_ENERGY_EV, w * e);
    }
    void brems_photon(double w, double e)  // :191-204
    {
        if (w <= 0.0 || e <= 0.0) return;
        add(ORC_L_ELECTRON_BREMS_PHOTON_COUNT, 1.0);
        add(ORC_L_TOTAL_BREMS_PHOTON_COUNT, 1.0);
        add(ORC_L_ELECTRON_BREMS_PHOTON_WEIGHT, w);
        add(ORC_L_ELECTRON_BREMS_PHOTON_ENERGY_EV, w * e);
        add(ORC_L_TOTAL_BREMS_PHOTON_WEIGHT, w);
        add(ORC_L_TOTAL_BREMS_PHOTON_ENERGY_EV, w * e);
    }
    // a photon-seeded electron of kinetic energy e, kinetic or low-energy
    void seeded_electron(double w, double e, int kinetic_slot, int low_slot)
    {
        if (e >= 0.0) {
            add(ORC_L_PHOTON_SEEDED_ELECTRON_WEIGHT, w);
            if (e >= cutoff) {
                add(ORC_L_PHOTON_SEEDED_KINETIC_ELECTRON_WEIGHT, w);
                add(kinetic_slot, w);
            } else {
                add(ORC_L_PHOTON_SEEDED_LOW_ELECTRON_WEIGHT, w);
                add(low_slot, w);
            }
        }
    }
    void compton(double w, double recoil, double charged, double local)  // :206-235
    {
        photon_interaction(w);
        add(ORC_L_COMPTON_INTERACTION_COUNT, 1.0);
        add(ORC_L_COMPTON_INTERACTION_WEIGHT, w);
        seeded_electron(w, recoil, ORC_L_COMPTON_KINETIC_ELECTRON_WEIGHT, ORC_L_COMPTON_LOW_ELECTRON_WEIGHT);
        add(ORC_L_PHOTON_TO_CHARGED_ENERGY_EV, w * pos(charged));
        add(ORC_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV, w * pos(local));
        add(ORC_L_TRANSPORT_ENERGY_LOSS_EV, w * (pos(charged) + pos(local)));
    }
    void photoelectric(double w, double photon, double electron, double charged, double local)  // :237-269
    {
        photon_interaction(w);
        add(ORC_L_PHOTOELECTRIC_INTERACTION_COUNT, 1.0);
        add(ORC_L_PHOTOELECTRIC_INTERACTION_WEIGHT, w);
        seeded_electron(w, electron, ORC_L_PHOTOELECTRIC_KINETIC_ELECTRON_WEIGHT,
                        ORC_L_PHOTOELECTRIC_LOW_ELECTRON_WEIGHT);
        add(ORC_L_PHOTON_ABSORBED_ENERGY_EV, w * pos(photon));
        add(ORC_L_PHOTON_TO_CHARGED_ENERGY_EV, w * pos(charged));
        add(ORC_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV, w * pos(local));
        add(ORC_L_TRANSPORT_ENERGY_LOSS_EV, w * pos(photon));
    }
    void pair(double w, double photon, double electron, double positron, double recoil, double charged,
              double local, bool triplet)  // :271-340
    {
        photon_interaction(w);
        add(ORC_L_PAIR_PRODUCTION_COUNT, 1.0);
        add(ORC_L_PAIR_PRODUCTION_WEIGHT, w);
        add(triplet ? ORC_L_PAIR_TRIPLET_COUNT : ORC_L_PAIR_NUCLEAR_COUNT, 1.0);
        add(triplet ? ORC_L_PAIR_TRIPLET_WEIGHT : ORC_L_PAIR_NUCLEAR_WEIGHT, w);
        seeded_electron(w, electron, ORC_L_PAIR_KINETIC_ELECTRON_WEIGHT, ORC_L_PAIR_LOW_ELECTRON_WEIGHT);
        add(ORC_L_PHOTON_SEEDED_POSITRON_WEIGHT, w);
        if (positron >= cutoff) {
            add(ORC_L_PHOTON_SEEDED_KINETIC_POSITRON_WEIGHT, w);
            add(ORC_L_PAIR_KINETIC_POSITRON_WEIGHT, w);
        } else {
            add(ORC_L_PHOTON_SEEDED_BELOW_CUTOFF_POSITRON_WEIGHT, w);
            add(ORC_L_PAIR_BELOW_CUTOFF_POSITRON_WEIGHT, w);
        }
        if (triplet) {
            add(ORC_L_PHOTON_SEEDED_ELECTRON_WEIGHT, w);
            if (recoil >= cutoff) {
                add(ORC_L_PHOTON_SEEDED_KINETIC_ELECTRON_WEIGHT, w);
                add(ORC_L_PAIR_TRIPLET_RECOIL_KINETIC_ELECTRON_WEIGHT, w);
            } else {
                add(ORC_L_PHOTON_SEEDED_LOW_ELECTRON_WEIGHT, w);
                add(ORC_L_PAIR_TRIPLET_RECOIL_LOW_ELECTRON_WEIGHT, w);
            }
        }
        add(ORC_L_PAIR_RECOIL_ELECTRON_ENERGY_EV, w * pos(recoil));
        add(ORC_L_PAIR_LOCAL_DEPOSIT_ENERGY_EV, w * pos(local));
        add(ORC_L_PHOTON_ABSORBED_ENERGY_EV, w * pos(photon));
        add(ORC_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV, w * pos(local));
        add(ORC_L_PHOTON_TO_CHARGED_ENERGY_EV, w * pos(charged));
        add(ORC_L_TRANSPORT_ENERGY_LOSS_EV, w * pos(charged + local - 2.0 * kMeC2));
    }
    void positron_transport(double w, double loss, bool thermalized)  // :342-354
    {
        if (w <= 0.0) return;
        add(ORC_L_POSITRON_TRANSPORT_ENERGY_LOSS_EV, w * pos(loss));
        if (thermalized) add(ORC_L_POSITRON_THERMALIZED_COUNT, 1.0);
    }
    void bhabha(double w, double sec, bool kinetic)  // :356-370
    {
        if (w <= 0.0) return;
        add(ORC_L_POSITRON_BHABHA_EVENT_COUNT, 1.0);
        add(ORC_L_POSITRON_BHABHA_EVENT_WEIGHT, w);
        add(ORC_L_POSITRON_BHABHA_SECONDARY_ENERGY_EV, w * pos(sec));
        if (kinetic) add(ORC_L_POSITRON_BHABHA_KINETIC_SECONDARY_WEIGHT, w);
    }
    void positron_brems(double w, double e)  // :372-385
    {
        if (w <= 0.0 || e <= 0.0) return;
        add(ORC_L_POSITRON_BREMS_EVENT_COUNT, 1.0);
        add(ORC_L_TOTAL_BREMS_PHOTON_COUNT, 1.0);
        add(ORC_L_POSITRON_BREMS_EVENT_WEIGHT, w);
        add(ORC_L_POSITRON_BREMS_PHOTON_ENERGY_EV, w * e);
        add(ORC_L_TOTAL_BREMS_PHOTON_WEIGHT, w);
        add(ORC_L_TOTAL_BREMS_PHOTON_ENERGY_EV, w * e);
    }
    void guard(double hm_tau, double brems_tau, double total_tau, double photon_tau, double positron_tau,
               uint64_t subcycles, double electron_e, double photon_e, double positron_e, bool violation)  // :387-413
    {
        up(ORC_L_MAX_TRANSPORT_HARD_MOLLER_TAU, hm_tau);
        up(ORC_L_MAX_TRANSPORT_BREMS_TAU, brems_tau);
        up(ORC_L_MAX_TRANSPORT_ELECTRON_TOTAL_COMPETING_TAU, total_tau);
        up(ORC_L_MAX_TRANSPORT_PHOTON_TAU, photon_tau);
        up(ORC_L_MAX_TRANSPORT_POSITRON_TAU, positron_tau);
        up(ORC_L_MAX_TRANSPORT_SUBCYCLES, static_cast<double>(subcycles));
        up(ORC_L_MAX_TRANSPORT_ELECTRON_ENERGY_EV, electron_e);
        up(ORC_L_MAX_TRANSPORT_PHOTON_ENERGY_SEEN_EV, photon_e);
        up(ORC_L_MAX_TRANSPORT_POSITRON_ENERGY_EV, positron_e);
        if (violation) add(ORC_L_TRANSPORT_OPTICAL_DEPTH_VIOLATION_COUNT, 1.0);
    }
    // RreaWarpXCoupling::RecordPhotonLocalDeposit OV/RreaCouplingSources.cpp:518-525,
    // the ledger part of RecordPositronAnnihilation :527-553
    void photon_local_deposit(double w, double e)
    {
        if (w > 0.0 && e > 0.0) add(ORC_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV, w * e);
    }
};

// Probability of a competing-risk optical depth, OV/RreaMaterialPath.H:203-210
inline double probability_from_optical_depth(double tau)
{
    if (!(tau > 0.0)) return 0.0;
    return std::min(1.0, -std::expm1(-tau));
}

// The photon energy flows the interaction records carry (charged kinetic plus
// rest mass, local), RreaComptonPhotonEnergyFlow / RreaPhotoelectricPhoton
// EnergyFlow / RreaPairPhotonEnergyFlow, OV/RreaParticleInteraction.H:1014-1092
struct EnergyFlow { double charged = 0.0, local = 0.0; };

inline EnergyFlow compton_flow(double recoil, double scattered, double cutoff, double gcut)
{
    EnergyFlow f;
    double const r = std::max(recoil, 0.0), s = std::max(scattered, 0.0);
    if (r >= cutoff) f.charged = r;
    else f.local = r;
    if (s < gcut) f.local += s;
    return f;
}

inline EnergyFlow photoelectric_flow(double electron, double binding, double cutoff)
{
    EnergyFlow f;
    double const e = std::max(electron, 0.0);
    f.local = std::max(binding, 0.0);
    if (e >= cutoff) f.charged = e;
    else f.local += e;
    return f;
}

inline EnergyFlow pair_flow(double electron, double positron, double recoil, double local, double cutoff,
                            bool has_recoil)
{
    EnergyFlow f;
    f.charged = 2.0 * kMeC2;
    f.local = std::max(local, 0.0);
    auto part = [&](double k) {
        double const e = std::max(k, 0.0);
        if (e >= cutoff) f.charged += e;
        else f.local += e;
    };
    part(electron);
    part(positron);
    if (has_recoil) part(recoil);
    return f;
}

struct StepContext {
    Tables const* t;
    Coupling cp;
    orc_config cfg;
    double r_lo, r_hi, z_lo, z_hi;
    double density_ratio;  // reference-z fallback (RreaParticleInteraction.cpp:33-35)
};

// ---------------------------------------------------------------------------
// Material sources: RreaEventAccumulator (NA/src/RreaEventAccumulator.cpp)
// fed by RreaWarpXCoupling::Accumulate* (OV/RreaCouplingSources.cpp:435-516)
// ---------------------------------------------------------------------------
// AddToField :386-512 on a non-periodic RZ grid from the axis
inline void source_add(orc_geometry const* g, int field, double r, double z, double value)
{
    if (!std::isfinite(value) || value == 0.0) return;
    int const ic = static_cast<int>(std::floor((r - g->r_lo) / g->dr));
    int const jc = static_cast<int>(std::floor((z - g->z_lo) / g->dz));
    if (ic < 0 || ic >= g->nr || jc < 0 || jc >= g->nz) return;
    double const r_cell = (r - g->r_lo) / g->dr + static_cast<double>(0) - 0.5;
    int const i0 = static_cast<int>(std::floor(r_cell));
    double const wr1 = r_cell - static_cast<double>(i0);
    double const wr0 = 1.0 - wr1;
    double const z_cell = (z - g->z_lo) / g->dz + static_cast<double>(0) - 0.5;
    int const j0 = static_cast<int>(std::floor(z_cell));
    double const wz1 = z_cell - static_cast<double>(j0);
    double const wz0 = 1.0 - wz1;
    double* const dst = g->sources + static_cast<int64_t>(field) * g->nr * g->nz;
    for (int rz = 0; rz < 2; ++rz) {
        double const wz = rz == 0 ? wz0 : wz1;
        if (wz == 0.0) continue;
        int const j = j0 + rz;
        if (j < 0 || j > g->nz - 1) continue;  // an absorbing z boundary keeps it in its ghost cell
        for (int rr = 0; rr < 2; ++rr) {
            double const wr = rr == 0 ? wr0 : wr1;
            double const v = value * wr * wz;
            if (v == 0.0) continue;
            int i = i0 + rr;
            if (i < 0 || i > g->nr - 1) {
                if (i == -1 && g->r_lo == 0.0) i = 0;  // the RZ mode-0 fold across the axis
                else continue;
            }
            dst[i + static_cast<int64_t>(g->nr) * j] += v;
        }
    }
}

// AddIonizationDiagnostic :156-167
inline void source_diagnostic(orc_geometry const* g, double r, double z, double weight, double loss)
{
    source_add(g, 3, r, z, weight);
    source_add(g, 4, r, z, std::max(weight, 0.0) * std::max(loss, 0.0));
}

// the coupling's three accumulators; a record outside the half-open physical
// domain is the reference's abort (EventSourcePositionInDomain)
inline int accumulate_source(orc_geometry const* g, double cutoff, SourceOp const& op)
{
    if (op.w <= 0.0) return 0;
    double const r_hi = g->r_lo + g->nr * g->dr, z_hi = g->z_lo + g->nz * g->dz;
    if (!(op.r >= g->r_lo && op.r < r_hi && op.z >= g->z_lo && op.z < z_hi)) return ORC_ERR_OUTSIDE_DOMAIN;
    if (op.kind == 0) {
        source_add(g, 1, op.r, op.z, op.w);
        if (op.secondary < cutoff) source_add(g, 0, op.r, op.z, op.w);
        source_diagnostic(g, op.r, op.z, op.w, op.loss);
    } else if (op.kind == 1) {
        source_add(g, 1, op.r, op.z, op.w);
        source_diagnostic(g, op.r, op.z, op.w, op.loss);
    } else {
        source_add(g, 0, op.r, op.z, op.w);
        source_diagnostic(g, op.r, op.z, op.w, std::max(op.loss, 0.0));
    }
    return 0;
}

// the replay of one chunk's material-source ledger, in order
inline int replay_sources(orc_geometry const* g, double cutoff, std::vector<SourceOp> const& ops)
{
    if (!g->sources) return 0;
    for (auto const& op : ops) {
        int const e = accumulate_source(g, cutoff, op);
        if (e) return e;
    }
    return 0;
}

// substep_driver_count OV/RreaTransportStep.H:106-110
inline uint64_t driver_count(double amount, double per)
{
    return static_cast<uint64_t>(std::max(1.0, std::ceil(amount / per)));
}

// selected_subcycles :125-141 + selected_material_subcycles :148-184
inline uint64_t select_material_subcycles(double max_tau, std::vector<Interval> const& iv, Path const& p,
                                          orc_config const& cfg, bool& violation, double total_stp)
{
    if (!(std::isfinite(cfg.max_optical_depth_per_substep) && cfg.max_optical_depth_per_substep > 0.0
          && cfg.max_subcycles > 0U)) {
        fail(ORC_ERR_INVALID, "invalid optical-depth subcycle controls");
    }
    double const target = cfg.max_optical_depth_per_substep;
    uint64_t const cap = cfg.max_subcycles;
    uint64_t const requested = static_cast<uint64_t>(std::max(1.0, std::ceil(max_tau / target)));
    violation = requested > cap;
    uint64_t n = std::min(requested, cap);
    for (;;) {
        double const maximum = total_stp * max_uniform_slice(iv, p, n);
        double const tol = 64.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, target);
        if (maximum <= target + tol) return n;
        if (n >= cap) { violation = true; return n; }
        long double const est = std::ceil(static_cast<long double>(n) * static_cast<long double>(maximum / target));
        uint64_t next = n + 1U;
        if (est >= static_cast<long double>(cap)) next = cap;
        else if (est > static_cast<long double>(next)) next = static_cast<uint64_t>(est);
        n = std::min(next, cap);
    }
}

// The per-particle side-effect sink of one transport transaction: ion-pair
// events (AccumulateIonizationEvent weights, OV/RreaTransportLog.H:449-462),
// local deposits as E / W_air pairs (deposit_local_energy_as_ion_pairs
// OV/RreaTransportStep.H:537-561) and secondaries with the parent's spawn
// lineage counter (AddSecondary OV/RreaTransportLog.H:437-447).
struct Emitter {
    ChunkLog& fx;
    int64_t ip;
    uint64_t pkey;
    double w;
    double& ion;
    int32_t root_species;     // RreaSecondaryParticle::root_species of every record it makes
    int32_t generation = 0;   // FIFO generation of the records it makes
    uint64_t lineage = 0;

    void ion_event(double ev_w)
    {
        ion += ev_w;
        fx.tallies[ORC_T_ION_PAIRS] += ev_w;
    }
    // material-source records at (r, z)
    void ionization(double r, double z, double ew, double secondary, double loss)
    {
        fx.sources.push_back(SourceOp{0, r, z, ew, secondary, loss});
    }
    void positive_ion(double r, double z, double ew, double loss)
    {
        fx.sources.push_back(SourceOp{1, r, z, ew, 0.0, loss});
    }
    void low_electron(double r, double z, double ew, double loss)
    {
        fx.sources.push_back(SourceOp{2, r, z, ew, 0.0, loss});
    }
    // deposit_local_energy_as_ion_pairs: E / W_air pairs, one ionization
    // event of secondary energy 0 and diagnostic loss W_air at (r, z)
    void deposit_local(double energy, double r, double z)
    {
        if (w <= 0.0 || energy <= 0.0) return;
        double const pw = w * energy / std::max(kWAir, 1.0);
        if (pw <= 0.0) return;
        ion_event(pw);
        ionization(r, z, pw, 0.0, kWAir);
        Ledger{fx.tallies, 0.0}.transport_ionization(pw, kWAir, true, false);
    }
    void secondary(int sp, int process, double x, double y, double z, double sdx, double sdy, double sdz,
                   double energy, double birth)
    {
        orc_secondary s{ip, sp, process, x, y, z, sdx, sdy, sdz, w, energy, birth, pkey, lineage++,
                        root_species, generation, -1, 0};
        fx.secondaries.push_back(s);
    }
};

// spawn_positron_annihilation OV/RreaTransportStep.H:563-700: two photons
// (tracked at or above the photon cutoff, else deposited locally), the
// residual kinetic energy of an at-rest annihilation deposited as ion pairs,
// and the positive ion the annihilated target electron leaves behind.  The
// photons join the enclosing group, so they carry its process id.
inline void spawn_annihilation(StepContext const& ctx, Emitter& em, int32_t process, double x, double y, double z,
                               double r, double ax, double ay, double az, double k, bool at_rest, uint64_t index,
                               double birth)
{
    Tables const& t = *ctx.t;
    double const gcut = ctx.cfg.photon_cutoff_eV;
    uint64_t const step = static_cast<uint64_t>(std::max<int64_t>(ctx.cfg.step, 0));
    if (em.w <= 0.0) return;
    normalize(ax, ay, az);
    double const ka = at_rest ? 0.0 : std::max(k, 0.0);
    if (!(ka == 0.0) && (!std::isfinite(ka) || ka < t.charged_min || ka > t.charged_max)) {
        fail(ORC_ERR_TABLE_RANGE, "annihilation energy outside certified range");
    }
    AnnihilationState const st = sample_annihilation(
        ka, rng_uniform(ctx.cfg.rng_seed, em.pkey, step, index, kChPositronAnnihilationPolar),
        rng_uniform(ctx.cfg.rng_seed, em.pkey, step, index, kChPositronAnnihilationAzimuth));
    double g1x, g1y, g1z;
    cone_direction(ax, ay, az, st.cos1, st.az1, g1x, g1y, g1z);
    double const e1 = std::max(st.e1, 0.0);
    double const e2 = std::max(st.e2, 0.0);
    double g2x = -g1x, g2y = -g1y, g2z = -g1z;
    if (!st.at_rest && e2 > 0.0) {
        double const p = std::sqrt(std::max(0.0, ka * (ka + 2.0 * kMeC2)));
        g2x = (p * ax - e1 * g1x) / e2;
        g2y = (p * ay - e1 * g1y) / e2;
        g2z = (p * az - e1 * g1z) / e2;
        normalize(g2x, g2y, g2z);
    }
    double local = 0.0;
    if (e1 >= gcut) em.secondary(2, process, x, y, z, g1x, g1y, g1z, e1, birth);
    else local += e1;
    if (e2 >= gcut) em.secondary(2, process, x, y, z, g2x, g2y, g2z, e2, birth);
    else local += e2;
    if (at_rest && k > 0.0) em.deposit_local(k, r, z);
    if (local > 0.0) em.deposit_local(local, r, z);
    // the annihilated target electron leaves one positive ion
    em.positive_ion(r, z, em.w, 0.0);
    em.fx.tallies[ORC_T_POSITIVE_IONS] += em.w;
    em.fx.tallies[ORC_T_ANNIHILATIONS] += 1.0;
    Ledger{em.fx.tallies, 0.0}.photon_local_deposit(em.w, local);
    if (em.w > 0.0) {
        Closure c;
        c.parent(1, em.w, std::max(k, 0.0), 0.0, false);
        if (e1 >= ctx.cfg.photon_cutoff_eV) c.particle(2, em.w, e1);
        if (e2 >= ctx.cfg.photon_cutoff_eV) c.particle(2, em.w, e2);
        c.local(em.w * ((at_rest && k > 0.0 ? k : 0.0) + local));
        c.positive_ion(em.w);
        require_conserved(c, ctx.cfg.low_energy_cutoff_eV);
    }
}

// mean_free_path_from_inverse_length NA/src/RreaInteractionTables.cpp:110-119
inline double mfp_from_inverse(double inverse_per_m)
{
    if (!std::isfinite(inverse_per_m) || inverse_per_m < 0.0) fail(ORC_ERR_INVALID, "inverse length invalid");
    return inverse_per_m > 0.0 ? 1.0 / inverse_per_m : 1.0e99;
}

// RreaChannelInverseLength OV/RreaChargedInterleave.H:69-78: a closed channel
// is exactly zero and never evaluates its mean free path (the gate keeps the
// lookup inside the channel's energy range); a non-positive mean free path
// is a table hole, zero.
template <class Mfp>
inline double gated_inverse_length(bool open, Mfp&& mfp_m)
{
    if (!open) return 0.0;
    double const mfp = mfp_m();
    return mfp > 0.0 ? 1.0 / mfp : 0.0;
}

// RreaChargedCompetingStpRates OV/RreaChargedFinalStates.H:107-152 (electron):
// hard Moller (ElectronHardMollerMeanFreePath :2394-2412, open above 2 T_cut)
// and tracked brems (ElectronBremsstrahlungMeanFreePath :2666-2676, open
// above the photon cutoff).
inline void electron_stp_rates(Tables const& t, double thr, double gcut, double e, bool hm_on, double out[3])
{
    bool const hm_open = hm_on && thr > 0.0 && e > 2.0 * thr;
    bool const br_open = std::isfinite(e) && std::isfinite(gcut) && e > gcut;
    double const hm = gated_inverse_length(hm_open, [&] { return mfp_from_inverse(t.rate[kEMoller].interpolate(e) * 1.0); });
    double const br = gated_inverse_length(br_open, [&] { return mfp_from_inverse(t.rate[kEBrems].interpolate(e) * 1.0); });
    out[0] = frozen(hm);
    out[1] = frozen(br);
    out[2] = 0.0;
}

// RreaChargedCompetingStpRates OV/RreaChargedFinalStates.H:114-133 (positron):
// hard Bhabha (PositronHardBhabhaMeanFreePath :3163-3178, open unless
// K <= T_cut), tracked brems (PositronBremsstrahlungMeanFreePath :3151-3161,
// K > photon cutoff), in-flight annihilation
// (PositronAnnihilationMeanFreePath :3196-3206, always open).
inline void positron_stp_rates(Tables const& t, double thr, double gcut, double e, double out[3])
{
    double const bh = gated_inverse_length(!(thr > 0.0 && e <= thr),
                                           [&] { return mfp_from_inverse(t.rate[kPBhabha].interpolate(e) * 1.0); });
    double const br = gated_inverse_length(e > gcut, [&] { return mfp_from_inverse(t.rate[kPBrems].interpolate(e) * 1.0); });
    double const an = gated_inverse_length(true, [&] { return mfp_from_inverse(t.rate[kPAnnih].interpolate(e) * 1.0); });
    out[0] = frozen(bh);
    out[1] = frozen(br);
    out[2] = frozen(an);
}

// RreaChargedChannelOpenAt OV/RreaChargedFinalStates.H:56-96; the
// Electron/PositronTransportCoefficients call range-checks and evaluates
// every rate of the species.
inline bool charged_channel_open(Tables const& t, bool positron, double thr, double gcut, int ch, double e, bool hm_on)
{
    if (!std::isfinite(e) || e < t.charged_min || e > t.charged_max) {
        fail(ORC_ERR_TABLE_RANGE, "schema-6 charged energy outside certified range");
    }
    if (positron) {
        (void)t.rate[kPCol].interpolate(e);
        (void)t.rate[kPSoft].interpolate(e);
        double const bh_inv = t.rate[kPBhabha].interpolate(e) * 1.0;
        double const br_inv = t.rate[kPBrems].interpolate(e) * 1.0;
        (void)t.rate[kPEltr].interpolate(e);
        double const an_inv = t.rate[kPAnnih].interpolate(e) * 1.0;
        if (ch == 0) return !(thr > 0.0 && e <= thr) && bh_inv > 0.0;
        if (ch == 1) return e > gcut && br_inv > 0.0;
        if (ch == 2) return an_inv > 0.0;
        return false;
    }
    double const hm_inv = t.rate[kEMoller].interpolate(e) * 1.0;
    double const br_inv = t.rate[kEBrems].interpolate(e) * 1.0;
    (void)t.rate[kECol].interpolate(e);
    (void)t.rate[kESoft].interpolate(e);
    (void)t.rate[kEEltr].interpolate(e);
    if (ch == 0) return hm_on && thr > 0.0 && e > 2.0 * thr && hm_inv > 0.0;
    if (ch == 1) return std::isfinite(e) && std::isfinite(gcut) && e > gcut && br_inv > 0.0;
    return false;
}

// elastic_inverse_length_stp OV/RreaTransportStep.H:199-208
inline double elastic_inverse_length(Tables const& t, bool positron, double e)
{
    double const mfp = t.elastic_mfp(positron, e, 1.0);
    return frozen(mfp > 0.0 ? 1.0 / mfp : 0.0);
}

// electron_realized_continuous_loss_ev_per_m :210-224
inline double realized_loss(Tables const& t, double e, double ratio)
{
    if (!(e > 0.0) || !(ratio > 0.0)) return 0.0;
    return t.rate[kECol].interpolate(e) * ratio + t.rate[kESoft].interpolate(e) * ratio;
}

// electron_ensemble_mean_drag_ev_per_m :226-247
inline double ensemble_drag(Tables const& t, double e, double ratio, bool hm_on)
{
    if (!(e > 0.0) || !(ratio > 0.0)) return 0.0;
    double const dd = hm_on ? t.hard_moller_energy_loss(e, ratio) : 0.0;
    return realized_loss(t, e, ratio) + dd + t.tracked_brems_energy_loss(e, ratio);
}

// ---------------------------------------------------------------------------
// The charged interleave core shared by the live and the newborn loops
// (OV/RreaChargedInterleave.H, OV/RreaParticleInteraction.H).  The policies
// are callables, so the exported probes (orc_interleave_probe,
// orc_substep_probe, orc_partition_ion_pairs, orc_limit_at_cutoff) run the
// very code the steps run on the reference's own behavioural cases.
// ---------------------------------------------------------------------------
struct SegLimit { double end; bool cutoff; };

// RreaChargedStopIsEscape OV/RreaChargedInterleave.H:140-146: a stop at or past
// the clip root of an exiting chord is an escape.
inline bool stop_is_escape(bool path_exits, double f, double end_f) { return path_exits && f >= end_f; }

// RreaLimitContinuousSegmentAtCutoff OV/RreaParticleInteraction.H:203-265:
// where the loss accumulated from b, loss_to(f), first reaches k - cutoff
// (56 bisection steps), or e when it does not.
template <class LossTo>
inline SegLimit limit_at_cutoff(double b, double e, double k, double cutoff, LossTo&& loss_to)
{
    if (!std::isfinite(b) || !std::isfinite(e) || !std::isfinite(k) || !std::isfinite(cutoff) || !(e >= b)
        || !(k >= cutoff) || !(cutoff >= 0.0)) {
        fail(ORC_ERR_INVALID, "continuous stopping could not locate a cutoff point");
    }
    if (e == b) return {b, k == cutoff};
    double const to_cut = k - cutoff;
    double const req = loss_to(e);
    if (!std::isfinite(req) || req < 0.0) fail(ORC_ERR_INVALID, "continuous stopping invalid");
    if (req < to_cut) return {e, false};
    if (to_cut == 0.0) return {b, true};
    double lo = b, hi = e;
    for (int it = 0; it < 56; ++it) {
        double const mid = 0.5 * (lo + hi);
        double const l = loss_to(mid);
        if (!std::isfinite(l) || l < 0.0) fail(ORC_ERR_INVALID, "continuous stopping invalid");
        if (l < to_cut) lo = mid; else hi = mid;
    }
    return {0.5 * (lo + hi), true};
}

// RreaPartitionContinuousIonPairs OV/RreaChargedInterleave.H:157-196: the
// segment's collision loss prorated over the material intervals it overlaps,
// each share emitted as ion pairs at the overlap's midpoint fraction.
template <class Emit>
inline void partition_ion_pairs(std::vector<Interval> const& iv, double full_length, double b, double e,
                                double segment_loss, double raw_loss, double w, double w_air, double stp, Emit&& emit)
{
    if (!(segment_loss > 0.0) || !(raw_loss > 0.0)) return;
    for (auto const& it : iv) {
        double const lb = std::max(b, it.begin), le = std::min(e, it.end);
        if (!(le > lb)) continue;
        double const raw_i = full_length * (le - lb) * it.ratio * stp;
        double const loss_i = segment_loss * raw_i / raw_loss;
        double const pw = w * loss_i / std::max(w_air, 1.0);
        if (!(pw > 0.0)) continue;
        emit(0.5 * (lb + le), pw);
    }
}

// RreaPlanContinuousDragSegment OV/RreaChargedInterleave.H:239-292 and its
// application in apply_electron_continuous_segment
// (OV/RreaElectronTransport.cpp:294-400) / the positron and newborn
// equivalents: collision + soft radiative stopping frozen at stop_e over
// [b, e], limited at the cutoff, its collision share emitted as ion pairs;
// k is reduced by the drag (set to the cutoff when it is reached).
// The plan's loss partition for the ledger: drag, its collision and soft
// radiative shares, and the mean hard-Moller transfer over the same density
// path (zero where the policy passes no hard-transfer rate).
struct DragPlan { double drag = 0.0, collision = 0.0, soft = 0.0, hard = 0.0; };

template <class HardStp, class Emit>
inline SegLimit continuous_drag(Tables const& t, bool pos, Path const& path, std::vector<Interval> const& iv,
                                double b, double e, double stop_e, double cutoff, double w, double& k, DragPlan& plan,
                                HardStp&& hard_stp, Emit&& emit)
{
    plan = DragPlan{};
    if (!(e > b) || !(k > 0.0) || stop_e < cutoff) return {b, false};
    double const col = frozen(t.rate[pos ? kPCol : kECol].interpolate(stop_e) * 1.0);
    double const soft = frozen(t.rate[pos ? kPSoft : kESoft].interpolate(stop_e) * 1.0);
    double const total = col + soft;
    SegLimit const lim = limit_at_cutoff(b, e, k, cutoff, [&](double f) { return total * density_path(iv, path, b, f); });
    double const pm = density_path(iv, path, b, lim.end);
    double const raw_col = col * pm;
    double const raw_soft = soft * pm;
    double const avail = std::max(0.0, k - cutoff);
    double const drag = lim.cutoff ? avail : std::min(raw_col + raw_soft, avail);
    double const col_loss = std::min(raw_col, drag);
    plan.drag = drag;
    plan.collision = col_loss;
    plan.soft = std::max(0.0, drag - col_loss);
    plan.hard = std::min(frozen(hard_stp()) * pm, avail);
    partition_ion_pairs(iv, path.full_length, b, lim.end, col_loss, raw_col, w, kWAir, col, emit);
    k = std::max(0.0, k - drag);
    if (lim.cutoff) k = cutoff;
    return lim;
}

// RreaRunChargedEventInterleave OV/RreaChargedInterleave.H:546-664 with
// RreaResolveRemainingSegmentHazard OV/RreaParticleInteraction.H:1467-1492.
// Per iteration: rates refrozen at the survivor energy, one competing-hazard
// draw; no event: the continuous tail to se; an event: continuous loss to its
// exact depth, the channel revalidated at the post-drag energy (a channel
// closed by the drag is a null event with a new ordinal), the final state
// (false: the parent was consumed).
template <class Taus, class Draws, class DepthToFraction, class Continuous, class MarkStopped, class EventPoint,
          class ChannelOpen, class FinalState>
inline void run_event_interleave(double sb, double se, double const& k, double cutoff, int nch, uint64_t max_ordinal,
                                 Taus&& taus_at, Draws&& draws, DepthToFraction&& depth_to_fraction,
                                 Continuous&& continuous, MarkStopped&& mark_stopped, EventPoint&& event_point,
                                 ChannelOpen&& channel_open, FinalState&& final_state)
{
    double rem = sb;
    uint64_t ord = 0;
    while (rem < se && k >= cutoff) {
        if (ord >= max_ordinal) fail(ORC_ERR_INVALID, "event ordinal overflow");
        double const frozen_e = k;
        std::array<double, 3> const taus = taus_at(frozen_e, rem, se);
        std::pair<double, double> const u = draws(ord);
        Hazard const h = select_hazard(taus.data(), nch, u.first, u.second);
        if (!h.valid) fail(ORC_ERR_INVALID, "competing-hazard selection invalid");
        if (!h.event) {
            SegLimit const sl = continuous(rem, se, frozen_e);
            if (sl.cutoff) mark_stopped(sl.end);
            break;
        }
        double const ef = depth_to_fraction(h.sampled, rem, se);
        if (!(std::isfinite(rem) && std::isfinite(ef) && std::isfinite(se) && se > rem && ef >= rem && ef <= se)) {
            fail(ORC_ERR_INVALID, "sampled event depth outside its subcycle");
        }
        event_point(ef);
        SegLimit const sl = continuous(rem, ef, frozen_e);
        if (sl.cutoff) { mark_stopped(sl.end); break; }
        bool const open_ev = k >= cutoff && channel_open(h.channel, k);
        if (!open_ev) {
            bool const open_fr = frozen_e >= cutoff && channel_open(h.channel, frozen_e);
            if (!open_fr) fail(ORC_ERR_INVALID, "selected channel closed before its sampled depth");
            rem = ef;
            ++ord;
            continue;
        }
        if (!final_state(h.channel, ef)) return;
        if (k <= cutoff) { mark_stopped(ef); break; }
        rem = ef;
        ++ord;
    }
}

// RreaRunChargedSubstepInterleave OV/RreaChargedInterleave.H:686-732: per
// substep build -> events -> finish -> advance; a stop skips finish and the
// commit; a refused endpoint or an exiting chord sets `exited` and ends the
// loop; `completed` counts only committed substeps.
template <class Build, class Events, class Stopped, class Finish, class Advance, class Exits>
inline void run_substep_interleave(uint64_t subcycles, double const& k, double cutoff, uint64_t& completed,
                                   bool& exited, Build&& build, Events&& events, Stopped&& stopped, Finish&& finish,
                                   Advance&& advance, Exits&& exits)
{
    for (uint64_t sc = 0; sc < subcycles; ++sc) {
        if (k < cutoff) break;
        double sb = 0.0, se = 0.0;
        if (!build(sc, sb, se)) break;
        events(sc, sb, se);
        if (stopped()) break;
        if (finish(sc, sb, se)) break;
        if (!advance(se)) { exited = true; break; }
        if (exits()) { exited = true; break; }
        completed = sc + 1U;
    }
}

// One charged particle: the body of the particle loop of
// OV/RreaElectronTransport.cpp:59-889 (positron = false) or
// OV/RreaPositronTransport.cpp:58-862 (positron = true), with the shared
// prologue BeginRunningChargedParticle OV/RreaTransportStep.H:720-820, the
// substep skeleton RreaRunChargedSubstepInterleave and the event skeleton
// RreaRunChargedEventInterleave (OV/RreaChargedInterleave.H:546-732).
struct ChargedOut {
    double r, z, th, ux, uy, uz, ion_pairs;
    int status;
    uint64_t substeps;
};

inline void charged_particle(StepContext const& ctx, ChunkLog& fx, bool const pos, int64_t ip, double pr, double pz,
                             double pth, double pux, double puy, double puz, double w, int64_t id, int32_t cpu,
                             ChargedOut& out)
{
    Tables const& t = *ctx.t;
    orc_config const& cfg = ctx.cfg;
    double const cutoff = cfg.low_energy_cutoff_eV;
    double const gcut = cfg.photon_cutoff_eV;
    double const thr = t.hard_moller_threshold;
    bool const hm_on = cfg.disable_hard_moller == 0;
    double const dt = cfg.dt_s;
    double const charge = pos ? 1.0 : -1.0;
    uint64_t const species = pos ? 1U : 0U;
    uint64_t const step = static_cast<uint64_t>(std::max<int64_t>(cfg.step, 0));
    uint64_t const seed = cfg.rng_seed;

    out.ion_pairs = 0.0;
    out.substeps = 0;
    uint64_t const pkey = static_cast<uint64_t>(std::llabs(id)) ^ (static_cast<uint64_t>(static_cast<uint32_t>(cpu)) << 40);
    if (w <= 0.0) { out.status = ORC_ZERO_WEIGHT; return; }
    if (pr > ctx.r_hi + 1.0e-9 * ctx.r_hi) fail(ORC_ERR_OUTSIDE_DOMAIN, "charged particle survived outside the domain");

    double dx = pux, dy = puy, dz = puz;
    double k = kinetic_from_proper_velocity(dx, dy, dz);
    normalize(dx, dy, dz);
    double run_x = pr * std::cos(pth);
    double run_y = pr * std::sin(pth);
    double run_z = pz;
    double const x_particle = run_x, y_particle = run_y;
    double const r_particle = pr, z_particle = pz;

    Chord const driver = build_chord(run_x, run_y, run_z, dx, dy, dz, k, dt, dt, ctx.r_lo, ctx.r_hi, ctx.z_lo, ctx.z_hi);
    Path const& dpath = driver.path;
    EscapePolicy escape = clipped_escape_policy(dpath, std::hypot(dpath.ex, dpath.ey), dpath.ez);
    double const ds = driver.valid ? dpath.in_length : 0.0;
    std::vector<Interval> const div = build_intervals(ctx.cp, dpath, ctx.density_ratio);
    double const span = std::max(dpath.end_f, 1.0e-6);
    double const field_work = path_field_work(ctx.cp, dpath, charge, 0.0, 1.0).work / span;

    double const chord_dp = density_path(div, dpath, dpath.begin_f, dpath.end_f);
    // the electron demotion gate's density, OV/RreaElectronTransport.cpp:102-105
    double particle_ratio = 0.0;
    if (!pos) particle_ratio = ds > 0.0 ? chord_dp / ds : positive_or(ctx.cp.density_ratio_at_z(z_particle), ctx.density_ratio);
    double tr_dp = 0.0, tr_len = 0.0;
    Path path = dpath;
    std::vector<Interval> iv = div;

    Emitter em{fx, ip, pkey, w, out.ion_pairs, pos ? 1 : 0};
    Ledger L{fx.tallies, cutoff};
    auto ion_event = [&](double ev_w) { em.ion_event(ev_w); };
    auto deposit_local = [&](double energy, double r, double z) { em.deposit_local(energy, r, z); };
    // the net chord carries the species' signed charge (-w for e-, +w for e+)
    auto segment = [&](double r0, double z0, double r1, double z1) {
        DepositOp op{0, {r0, z0, r1, z1, charge * w, 0.0, 0.0}, 0U};
        fx.deposits.push_back(op);
    };
    auto commit_escape = [&]() {
        // CommitChargedEscape OV/RreaTransportStep.H:829-863
        double const um = proper_velocity_from_kinetic(k);
        out.ux = um * dx; out.uy = um * dy; out.uz = um * dz;
        out.r = pr; out.z = pz; out.th = pth;
        BoundaryCrossing const bc = boundary_crossing(path);
        if (!bc.valid) fail(ORC_ERR_INVALID, "charged escape has no certified absorbing-face crossing");
        DepositOp op{1, {r_particle, z_particle, bc.start_r, bc.start_z, bc.cross_r, bc.cross_z, charge * w}, bc.faces};
        fx.deposits.push_back(op);
        out.status = ORC_ESCAPE;
        fx.tallies[ORC_T_ESCAPES] += 1.0;
    };
    auto demote = [&](double rr, double zz) {
        // a low-energy electron source of the parent's weight
        // (AccumulateLowEnergyElectron, OV/RreaElectronTransport.cpp:152, 857)
        fx.tallies[ORC_T_LOW_ENERGY_ELECTRONS] += w;
        em.low_electron(rr, zz, w, 0.0);
        deposit_local(k, rr, zz);
        segment(r_particle, z_particle, rr, zz);
        out.r = pr; out.z = pz; out.th = pth; out.ux = pux; out.uy = puy; out.uz = puz;
        out.status = ORC_DEMOTED;
        fx.tallies[ORC_T_DEMOTIONS] += 1.0;
    };
    auto push_secondary = [&](int sp, int process, double x, double y, double z, double sdx, double sdy, double sdz,
                              double energy, double birth) {
        em.secondary(sp, process, x, y, z, sdx, sdy, sdz, energy, birth);
    };
    auto annihilate = [&](double x, double y, double z, double r, bool at_rest, uint64_t index, double birth) {
        spawn_annihilation(ctx, em, kProcAnnihilation, x, y, z, r, dx, dy, dz, k, at_rest, index, birth);
    };
    auto annihilated_at = [&](double r, double z) {
        segment(r_particle, z_particle, r, z);
        out.r = pr; out.z = pz; out.th = pth; out.ux = pux; out.uy = puy; out.uz = puz;
        out.status = ORC_ANNIHILATED;
    };

    // Initial terminal disposition (RreaResolveChargedTerminalDisposition
    // OV/RreaParticleInteraction.H:267-285)
    bool const below0 = k < cutoff;
    if (escape.escapes && below0) { commit_escape(); return; }
    if (!escape.escapes && below0) {
        if (pos) {
            // at rest where it stands, before any transport: a zero-length net chord
            annihilate(x_particle, y_particle, z_particle, r_particle, true, 0U, 1.0);
            L.positron_transport(w, k, true);  // OV/RreaPositronTransport.cpp:152
            annihilated_at(r_particle, z_particle);
        } else {
            demote(r_particle, z_particle);
        }
        return;
    }
    if (ds <= 0.0) {
        if (escape.escapes) { commit_escape(); return; }
        out.r = pr; out.z = pz; out.th = pth; out.ux = pux; out.uy = puy; out.uz = puz;
        out.status = ORC_NO_WRITEBACK;
        return;
    }
    if (div.empty()) fail(ORC_ERR_INVALID, "positive-length path without material segments");

    // Substep sizing, OV/RreaElectronTransport.cpp:187-279 / OV/RreaPositronTransport.cpp:177-245
    double stp0[3];
    if (pos) positron_stp_rates(t, thr, gcut, k, stp0);
    else electron_stp_rates(t, thr, gcut, k, hm_on, stp0);
    double const ddp = chord_dp / span;
    bool violation = false;
    uint64_t subcycles;
    if (pos) {
        double const total = stp0[0] + stp0[1] + stp0[2];
        subcycles = select_material_subcycles(total * ddp, div, dpath, cfg, violation, total);
    } else {
        subcycles = select_material_subcycles(stp0[0] * ddp + stp0[1] * ddp, div, dpath, cfg, violation,
                                              stp0[0] + stp0[1]);
    }
    {
        double const gs_n_total = elastic_inverse_length(t, pos, k) * ddp;
        double const drag_est = pos ? t.rate[kPCol].interpolate(std::max(k, cutoff)) * 1.0 * ddp
                                    : realized_loss(t, std::max(k, cutoff), 1.0) * ddp;
        double const impulse_scale = std::max(cfg.max_field_impulse_fraction * k, std::max(cfg.field_impulse_floor_eV, 1.0));
        uint64_t const gsd = driver_count(std::min(gs_n_total, cfg.gs_total_collision_cap),
                                          std::max(cfg.max_gs_collisions_per_substep, 1.0));
        uint64_t const imd = driver_count(std::abs(field_work) + drag_est, impulse_scale);
        subcycles = std::min<uint64_t>(std::max({subcycles, gsd, imd}), cfg.max_subcycles);
    }
    // the reported per-substep taus: uniform slices of the driver chord over the
    // clipped span (OV/RreaElectronTransport.cpp:250-278, OV/RreaPositronTransport.cpp:241-259)
    if (pos) {
        double const max_positron_tau = (stp0[0] + stp0[1] + stp0[2]) * max_uniform_slice(div, dpath, subcycles) / span;
        L.guard(0.0, 0.0, 0.0, 0.0, max_positron_tau, subcycles, 0.0, 0.0, k, violation);
    } else {
        double const slice = max_uniform_slice(div, dpath, subcycles) / span;
        L.guard(stp0[0] * slice, stp0[1] * slice, (stp0[0] + stp0[1]) * slice, 0.0, 0.0, subcycles, k, 0.0, 0.0,
                violation);
    }
    if (cfg.strict_transport_guards && violation) fail(ORC_ERR_GUARD, "strict transport guard: optical-depth cap exceeded");
    out.substeps = subcycles;
    fx.tallies[ORC_T_SUBSTEPS] += static_cast<double>(subcycles);
    double positron_total_loss = 0.0;

    bool exited = false;
    bool terminated = false;
    bool annihilated = false;
    double mid_x = x_particle, mid_y = y_particle, mid_r = r_particle, mid_z = z_particle;
    double stop_elapsed = 1.0;
    uint64_t completed = 0;
    auto elapsed_at = [&](double f) {
        double const within = std::min(1.0, std::max(0.0, f));
        return std::min(1.0, (static_cast<double>(completed) + within) / static_cast<double>(subcycles));
    };
    auto mark_stopped = [&](double f) {
        if (stop_is_escape(path.exits, f, path.end_f)) { exited = true; return; }
        double x, y, z;
        path.point_at_full(f, x, y, z);
        double const r = std::hypot(x, y);
        if (!in_half_open(x, y, z, ctx.r_lo, ctx.r_hi, ctx.z_lo, ctx.z_hi)) { exited = true; return; }
        mid_x = x; mid_y = y; mid_z = z; mid_r = r; terminated = true;
        stop_elapsed = elapsed_at(f);
    };

    // apply_electron_continuous_segment OV/RreaElectronTransport.cpp:294-400 /
    // apply_positron_continuous_segment OV/RreaPositronTransport.cpp:278-360:
    // collision + soft radiative stopping frozen at the segment's energy.
    auto continuous = [&](double b, double e, double stop_e) -> SegLimit {
        DragPlan plan;
        SegLimit const sl = continuous_drag(
            t, pos, path, iv, b, e, stop_e, cutoff, w, k, plan,
            [&] { return (!pos && hm_on && thr > 0.0 && stop_e > 2.0 * thr) ? t.hard_moller_energy_loss(stop_e, 1.0) : 0.0; },
            [&](double mid_f, double pw) {
                ion_event(pw);
                double mx, my, mz;
                path.point_at_full(mid_f, mx, my, mz);
                em.ionization(std::hypot(mx, my), mz, pw, 0.5 * cutoff, kWAir);
                L.transport_ionization(pw, kWAir, true, false);
            });
        fx.tallies[ORC_T_DRAG_LOSS_WEV] += w * plan.drag;
        L.soft_brems(w, plan.soft);
        if (pos) positron_total_loss += plan.drag;
        else L.loss_partition(w, plan.drag, plan.collision, plan.hard);
        return sl;
    };

    double const sub_dt = dt / static_cast<double>(subcycles);
    FieldRange sfield;
    double substep_dt = 0.0;
    int const nch = pos ? 3 : 2;
    uint64_t const ch_depth0 = pos ? kChPositronCompetingOpticalDepth : kChCompetingOpticalDepth;
    uint64_t const ch_pick0 = pos ? kChPositronCompetingChannel : kChCompetingChannel;
    uint64_t const ch_depth1 = pos ? kChPositronContinuationOpticalDepth : kChContinuationOpticalDepth;
    uint64_t const ch_pick1 = pos ? kChPositronContinuationChannel : kChContinuationChannel;

    // build_substep: the microchord from the running point, its material
    // intervals and field sub-range, the leading half kick
    auto build_substep = [&](uint64_t, double& sb, double& se) -> bool {
        Chord const ch = build_chord(run_x, run_y, run_z, dx, dy, dz, k, sub_dt, dt, ctx.r_lo, ctx.r_hi, ctx.z_lo, ctx.z_hi);
        if (!ch.valid) return false;
        path = ch.path;
        if (!(path.full_length > 0.0) || !(path.in_length > 0.0)) { exited = path.exits; return false; }
        iv = build_intervals(ctx.cp, path, ctx.density_ratio);
        if (iv.empty()) fail(ORC_ERR_INVALID, "microchord without material intervals");
        for (auto const& it : iv) { tr_dp += it.length * it.ratio; tr_len += it.length; }
        sb = path.begin_f;
        se = path.end_f;
        sfield = path_field_work(ctx.cp, path, charge, sb, se);
        substep_dt = sub_dt * (se - sb);
        fx.tallies[ORC_T_FIELD_WORK_WEV] += w * half_kick(sfield, 0.5 * substep_dt, charge, k, dx, dy, dz);
        if (k < cutoff) { mark_stopped(sb); return false; }
        return true;
    };
    // run_events: the competing hazards of the substep
    auto run_events = [&](uint64_t sc, double sb, double se) {
        uint64_t idx = 0;
        double stp[3] = {0, 0, 0};
        double ex_ = 0.0, ey_ = 0.0, ez_ = 0.0, er_ = 0.0;
        run_event_interleave(
            sb, se, k, cutoff, nch, kEventCap,
            [&](double frozen_e, double rem, double end) {
                if (pos) positron_stp_rates(t, thr, gcut, frozen_e, stp);
                else electron_stp_rates(t, thr, gcut, frozen_e, hm_on, stp);
                double const dp = density_path(iv, path, rem, end);
                if (!pos) {
                    double const hard_tau = stp[0] * dp, brems_tau = stp[1] * dp;
                    double const total_tau = hard_tau + brems_tau;
                    double const p_hard = total_tau > 0.0
                        ? probability_from_optical_depth(total_tau) * hard_tau / total_tau : 0.0;
                    L.hm_expectation(w * p_hard);
                }
                return std::array<double, 3>{stp[0] * dp, stp[1] * dp, stp[2] * dp};
            },
            [&](uint64_t ord) {
                if (!live_hazard_index(species, sc, ord, idx)) fail(ORC_ERR_INVALID, "hazard RNG index overflow");
                return std::pair<double, double>{rng_uniform(seed, pkey, step, idx, ord == 0U ? ch_depth0 : ch_depth1),
                                                 rng_uniform(seed, pkey, step, idx, ord == 0U ? ch_pick0 : ch_pick1)};
            },
            [&](double tau, double rem, double end) {
                double const total_stp = pos ? stp[0] + stp[1] + stp[2] : stp[0] + stp[1];
                return fraction_at_density_path(iv, path, rem, end, tau / total_stp);
            },
            continuous, mark_stopped,
            [&](double ef) {
                path.point_at_full(ef, ex_, ey_, ez_);
                er_ = std::hypot(ex_, ey_);
            },
            [&](int channel, double e) { return charged_channel_open(t, pos, thr, gcut, channel, e, hm_on); },
            [&](int channel, double ef) -> bool {
                double const efrac = elapsed_at(ef);
                if (!pos && channel == 0) {
                    // RreaApplyHardMollerFinalState OV/RreaChargedFinalStates.H:164-225
                    double const before = k;
                    double T = t.sample_moller(before, rng_uniform(seed, pkey, step, idx, kChHardMollerEnergy));
                    T = std::max(thr, std::min(T, 0.5 * before));
                    double pdx, pdy, pdz, sdx, sdy, sdz;
                    moller_lab_directions(dx, dy, dz, before, T, rng_uniform(seed, pkey, step, idx, kChHardMollerAzimuth),
                                          pdx, pdy, pdz, sdx, sdy, sdz);
                    k = std::max(0.0, before - T);
                    dx = pdx; dy = pdy; dz = pdz;
                    push_secondary(0, kProcHardMoller, ex_, ey_, ez_, sdx, sdy, sdz, T, efrac);
                    ion_event(w);
                    em.ionization(std::hypot(ex_, ey_), ez_, w, T, T);
                    fx.tallies[ORC_T_MOLLER_EVENTS] += 1.0;
                    L.hm_event(w, T);
                    L.transport_ionization(w, T, false, true);
                    Closure c;
                    c.parent(0, w, before, k, true);
                    c.particle(0, w, T);
                    c.ionization(w, T, cutoff);
                    require_conserved(c, cutoff);
                } else if (pos && channel == 0) {
                    // RreaApplyBhabhaFinalState OV/RreaChargedFinalStates.H:238-297: a
                    // secondary below the production threshold discards this event only
                    double const before = k;
                    double const T = std::min(t.sample_bhabha(before, rng_uniform(seed, pkey, step, idx, kChPositronBhabhaEnergy)),
                                              std::max(0.0, before));
                    if (T > 0.0 && !(thr > 0.0 && T < thr)) {
                        double pdx, pdy, pdz, sdx, sdy, sdz;
                        moller_lab_directions(dx, dy, dz, before, T,
                                              rng_uniform(seed, pkey, step, idx, kChPositronBhabhaAzimuth),
                                              pdx, pdy, pdz, sdx, sdy, sdz);
                        k = std::max(0.0, before - T);
                        dx = pdx; dy = pdy; dz = pdz;
                        push_secondary(0, kProcHardBhabha, ex_, ey_, ez_, sdx, sdy, sdz, T, efrac);
                        em.positive_ion(std::hypot(ex_, ey_), ez_, w, T);
                        fx.tallies[ORC_T_POSITIVE_IONS] += w;
                        fx.tallies[ORC_T_BHABHA_EVENTS] += 1.0;
                        L.transport_ionization(w, T, false, true);
                        L.bhabha(w, T, true);
                        Closure c;
                        c.parent(1, w, before, k, true);
                        c.particle(0, w, T);
                        c.positive_ion(w);
                        require_conserved(c, cutoff);
                    }
                } else if (channel == 1) {
                    // RreaApplyBremsFinalState OV/RreaChargedFinalStates.H:329-394
                    double const before = k;
                    double const eu = rng_uniform(seed, pkey, step, idx, pos ? kChPositronBremsPhotonEnergy : kChBremsPhotonEnergy);
                    double const ks = pos ? t.sample_positron_brems_energy(before, eu) : t.sample_brems_energy(before, eu);
                    // RreaClassifyBremsEnergyDisposition OV/RreaParticleInteraction.H:1552-1574
                    double const pe = std::min(std::max(ks, 0.0), std::max(before, 0.0));
                    double const after = std::max(0.0, std::max(before, 0.0) - pe);
                    bool const tracked = std::isfinite(gcut) && gcut >= 0.0 && pe >= gcut;
                    double const local = tracked ? 0.0 : pe;
                    double gx, gy, gz;
                    cone_direction(dx, dy, dz,
                                   t.sample_brems_cos(before, rng_uniform(seed, pkey, step, idx,
                                                                          pos ? kChPositronBremsPhotonAngle : kChBremsPhotonAngle)),
                                   rng_uniform(seed, pkey, step, idx, pos ? kChPositronBremsPhotonAzimuth : kChBremsPhotonAzimuth),
                                   gx, gy, gz);
                    photon_recoil(before, pe, gx, gy, gz, dx, dy, dz);
                    k = after;
                    if (pos) L.positron_brems(w, pe);
                    else L.brems_photon(w, pe);
                    if (tracked) {
                        push_secondary(2, kProcBrems, ex_, ey_, ez_, gx, gy, gz, pe, efrac);
                        fx.tallies[ORC_T_BREMS_PHOTONS] += 1.0;
                    } else if (local > 0.0) {
                        deposit_local(local, std::hypot(ex_, ey_), ez_);
                    }
                    fx.tallies[ORC_T_BREMS_EVENTS] += 1.0;
                    {
                        Closure c;
                        c.parent(pos ? 1 : 0, w, before, after, true);
                        if (tracked) c.particle(2, w, pe);
                        else c.local(w * local);
                        require_conserved(c, cutoff);
                    }
                } else if (pos && channel == 2) {
                    // in-flight annihilation, OV/RreaPositronTransport.cpp:710-745: the
                    // parent is consumed at the event point
                    annihilate(ex_, ey_, ez_, std::hypot(ex_, ey_), false, idx, efrac);
                    segment(r_particle, z_particle, er_, ez_);
                    annihilated = true;
                    return false;
                }
                return true;
            });
    };
    // finish_interleaved_substep: elastic scattering, then the trailing half kick
    auto finish_substep = [&](uint64_t sc, double sb, double se) -> bool {
        if (k >= cutoff) {
            // apply_gs_elastic_scatter OV/RreaTransportStep.H:372-423
            double const gs_n = elastic_inverse_length(t, pos, k) * density_path(iv, path, sb, se);
            if (gs_n > 0.0) {
                auto U = [&](uint64_t chan, uint64_t inner) {
                    return rng_uniform(seed, pkey, step, (sc << 8) | inner, chan);
                };
                auto deflect = [&](double ct, double uphi) {
                    double nx, ny, nz;
                    cone_direction(dx, dy, dz, ct, uphi, nx, ny, nz);
                    dx = nx; dy = ny; dz = nz;
                };
                if (gs_n < t.gs_few) {
                    int const cap = static_cast<int>(std::ceil(3.0 * t.gs_few));
                    if (cap >= 255) fail(ORC_ERR_INVALID, "GS few-collision cap does not pack");
                    int const cnt = poisson_capped(gs_n, U(pos ? kChPositronGsFewCount : kChGsFewCount, 0), cap);
                    for (int q = 0; q < cnt; ++q) {
                        uint64_t const inner = static_cast<uint64_t>(q + 1);
                        double const ct = t.sample_single_elastic_cos(
                            pos, k, U(pos ? kChPositronGsFewScatter : kChGsFewScatter, inner));
                        deflect(ct, U(pos ? kChPositronGsFewAzimuth : kChGsFewAzimuth, inner));
                    }
                } else {
                    double const ct = t.sample_gs_cos(pos, k, gs_n, U(pos ? kChPositronGsAngle : kChGsAngle, 0));
                    deflect(ct, U(pos ? kChPositronGsAzimuth : kChGsAzimuth, 0));
                }
            }
        }
        fx.tallies[ORC_T_FIELD_WORK_WEV] += w * half_kick(sfield, 0.5 * substep_dt, charge, k, dx, dy, dz);
        if (k < cutoff) { mark_stopped(se); return true; }
        return false;
    };

    // RreaRunChargedSubstepInterleave OV/RreaChargedInterleave.H:686-732
    run_substep_interleave(
        subcycles, k, cutoff, completed, exited, build_substep, run_events,
        [&] { return terminated || annihilated; }, finish_substep,
        [&](double se) { return advance_running_position(path, se, run_x, run_y, run_z); },
        [&] { return path.exits; });

    if (exited) escape = clipped_escape_policy(path, std::hypot(path.ex, path.ey), path.ez);
    if (pos) L.positron_transport(w, positron_total_loss, false);  // OV/RreaPositronTransport.cpp:776
    if (pos) {
        // OV/RreaPositronTransport.cpp:764-846: a consumed parent outranks every
        // other outcome; then escape; then a stop or a below-cutoff endpoint
        // annihilates at rest where the positron ended up.
        if (annihilated) {
            out.r = pr; out.z = pz; out.th = pth; out.ux = pux; out.uy = puy; out.uz = puz;
            out.status = ORC_ANNIHILATED;
            return;
        }
        if (exited) { commit_escape(); return; }
        if (terminated || k < cutoff) {
            double const ax = terminated ? mid_x : run_x;
            double const ay = terminated ? mid_y : run_y;
            double const az = terminated ? mid_z : run_z;
            double const ar = terminated ? mid_r : std::hypot(run_x, run_y);
            annihilate(ax, ay, az, std::hypot(ax, ay), true, subcycles, terminated ? stop_elapsed : 1.0);
            L.positron_transport(w, k, true);  // OV/RreaPositronTransport.cpp:839
            annihilated_at(ar, az);
            return;
        }
    } else {
        if (tr_len > 0.0) particle_ratio = tr_dp / tr_len;
        double const end_r = std::hypot(run_x, run_y);
        double fer, fez, fbt;
        ctx.cp.gather(end_r, run_z, fer, fez, fbt);
        // RreaEvaluateElectronStatePolicy OV/RreaParticleInteraction.H:1279-1317.  The
        // ensemble drag is read only when K <= the hard-Moller threshold; it has no
        // side effect, so evaluating it only there is exact.
        bool const below = k < cutoff;
        bool nonrunaway = false;
        if (thr > 0.0 && k <= thr) {
            double const thr_v = ensemble_drag(t, std::max(k, cutoff), particle_ratio, hm_on);
            nonrunaway = std::hypot(fer, fez) <= thr_v;
        }
        // RreaResolveChargedTransportTerminal :292-312
        if (exited) { commit_escape(); return; }
        if (terminated || below || nonrunaway) {
            demote(terminated ? mid_r : end_r, terminated ? mid_z : run_z);
            return;
        }
    }
    // CommitSurvivorChargedWriteback OV/RreaTransportStep.H:874-905
    double const end_r = std::hypot(run_x, run_y);
    double const um = proper_velocity_from_kinetic(k);
    out.ux = um * dx; out.uy = um * dy; out.uz = um * dz;
    out.r = end_r;
    out.th = std::atan2(run_y, run_x);
    out.z = run_z;
    segment(r_particle, z_particle, end_r, run_z);
    out.status = ORC_SURVIVE;
    fx.tallies[ORC_T_SURVIVORS] += 1.0;
}

// ---------------------------------------------------------------------------
// Photon step -- OV/RreaPhotonTransport.cpp:9-461, OV/RreaPhotonFinalStates.H,
// samplers NA/src/RreaInteractionTables.cpp:2272-2296, 2762-3127
// ---------------------------------------------------------------------------
enum PhotonChannel { kCompton = 0, kPhotoelectric = 1, kPairNuclear = 2, kPairTriplet = 3 };

struct PhotonRates { double compton, pe, pair_n, pair_t; };

// PhotonProcessInverseLengths :2272-2296 (density ratio 1)
inline PhotonRates photon_rates(Tables const& t, double e)
{
    if (!std::isfinite(e) || e < t.photon_min || e > t.photon_max) {
        fail(ORC_ERR_TABLE_RANGE, "schema-6 photon energy outside certified range");
    }
    return {t.rate[kGCompton].interpolate(e) * 1.0, t.rate[kGPe].interpolate(e) * 1.0,
            t.rate[kGPairN].interpolate(e) * 1.0, t.rate[kGPairT].interpolate(e) * 1.0};
}

inline double photon_total(PhotonRates const& r) { return r.compton + r.pe + r.pair_n + r.pair_t; }

// PhotonFeedbackMeanFreePath :2762-2771
inline double photon_mfp(Tables const& t, double e)
{
    double const v = photon_total(photon_rates(t, e));
    if (!std::isfinite(v) || v < 0.0) fail(ORC_ERR_INVALID, "photon inverse length invalid");
    return v > 0.0 ? 1.0 / v : 1.0e99;
}

// SamplePhotonChannel :2778-2804
inline int sample_photon_channel(Tables const& t, double e, double u)
{
    if (!std::isfinite(e) || e < t.photon_min || e > t.photon_max) {
        fail(ORC_ERR_TABLE_RANGE, "photon channel energy outside certified range");
    }
    u = std::max(0.0, std::min(1.0, u));
    PhotonRates const r = photon_rates(t, e);
    double const total = photon_total(r);
    if (!(total > 0.0) || !std::isfinite(total)) fail(ORC_ERR_INVALID, "photon inverse lengths sum to zero");
    double const target = u * total;
    if (target < r.compton) return kCompton;
    if (target < r.compton + r.pe) return kPhotoelectric;
    if (target < r.compton + r.pe + r.pair_n) return kPairNuclear;
    return kPairTriplet;
}

struct ComptonState { double scattered, recoil, cos_theta; };

// SampleComptonKleinNishinaFinalState :2806-2871: Butcher-Messel rejection,
// round 0 on the two drawn uniforms, later draws from the uniform stream
// seeded by them
inline ComptonState sample_compton(Tables const& t, double e, double draw_mu, double draw_accept)
{
    if (!std::isfinite(e) || e < t.photon_min || e > t.photon_max) {
        fail(ORC_ERR_TABLE_RANGE, "Compton energy outside certified range");
    }
    if (e <= 0.0) fail(ORC_ERR_INVALID, "Compton sampler requires positive photon energy");
    draw_mu = std::max(0.0, std::min(1.0, draw_mu));
    draw_accept = std::max(0.0, std::min(1.0, draw_accept));
    UniformStream stream(draw_mu, draw_accept);
    double const alpha = e / kMeC2;
    double const eps0 = 1.0 / (1.0 + 2.0 * alpha);
    double const eps0sq = eps0 * eps0;
    double const a1 = -std::log(eps0);
    double const a2 = 0.5 * (1.0 - eps0sq);
    double epsilon = eps0;
    double onecost = 0.0;
    for (int iter = 0; iter < 128; ++iter) {
        double const u1 = iter == 0 ? draw_mu : stream.next();
        double const u2 = iter == 0 ? draw_accept : stream.next();
        double epsilonsq;
        if (a1 > (a1 + a2) * u1) {
            epsilon = std::exp(-a1 * u2);
            epsilonsq = epsilon * epsilon;
        } else {
            epsilonsq = eps0sq + (1.0 - eps0sq) * u2;
            epsilon = std::sqrt(epsilonsq);
        }
        onecost = (1.0 - epsilon) / (epsilon * alpha);
        double const sint2 = onecost * (2.0 - onecost);
        double const greject = 1.0 - epsilon * sint2 / (1.0 + epsilonsq);
        if (greject >= stream.next()) break;
    }
    double mu = 1.0 - onecost;
    mu = std::max(-1.0, std::min(1.0, mu));
    ComptonState st;
    st.scattered = std::max(0.0, epsilon * e);
    st.recoil = std::max(0.0, e - st.scattered);
    st.cos_theta = mu;
    return st;
}

struct PhotoelectricState { double binding, electron; };

// SamplePhotoelectricFinalState :2878-2950
inline PhotoelectricState sample_photoelectric(Tables const& t, double e, double u)
{
    if (!std::isfinite(e) || e < t.photon_min || e > t.photon_max) {
        fail(ORC_ERR_TABLE_RANGE, "photoelectric energy outside certified range");
    }
    if (e <= 0.0) fail(ORC_ERR_INVALID, "photoelectric sampler requires positive photon energy");
    if (t.pe_primary.empty() || e < t.pe_primary.front() || e > t.pe_primary.back()) {
        fail(ORC_ERR_TABLE_RANGE, "photoelectric subshell primary energy out of range");
    }
    u = std::max(0.0, std::min(1.0, u));
    size_t const n = t.pe_primary.size();
    size_t const up = static_cast<size_t>(std::lower_bound(t.pe_primary.begin(), t.pe_primary.end(), e)
                                          - t.pe_primary.begin());
    size_t g;
    if (up == 0) g = 0;
    else if (up == n) g = n - 1;
    else if (t.pe_primary[up] == e) g = up;  // the low side at a duplicated edge
    else g = up - 1;
    auto const b = t.pe_u.begin() + t.pe_start[g], en = t.pe_u.begin() + t.pe_start[g + 1];
    auto const row_it = std::lower_bound(b, en, u);
    size_t const row = static_cast<size_t>((row_it == en ? en - 1 : row_it) - t.pe_u.begin());
    double const binding = t.pe_binding[row];
    if (binding > e) fail(ORC_ERR_INVALID, "photoelectric binding exceeds photon energy");
    return {binding, std::max(0.0, e - binding)};
}

// SamplePhotoelectronCosTheta :2957-3044: Sauter-Gavrila K-shell polar angle
// by bisection of its analytic primitive in t = 1 - beta mu, in long double
inline double sample_photoelectron_cos(Tables const& t, double ke, double draw)
{
    if (!(ke == 0.0) && (!std::isfinite(ke) || ke < t.charged_min || ke > t.charged_max)) {
        fail(ORC_ERR_TABLE_RANGE, "photoelectron energy outside certified range");
    }
    long double const u = static_cast<long double>(clamp_unit(draw));
    if (u <= 0.0L) return -1.0;
    if (u >= 1.0L) return 1.0;
    long double const gamma = 1.0L + static_cast<long double>(ke) / static_cast<long double>(kMeC2);
    long double const inv_gamma_sq = 1.0L / (gamma * gamma);
    long double const beta = std::sqrt(std::max(0.0L, 1.0L - inv_gamma_sq));
    if (beta < 1.0e-4L) {
        long double lo = -1.0L, hi = 1.0L;
        for (int it = 0; it < 96; ++it) {
            long double const mid = 0.5L * (lo + hi);
            if (mid == lo || mid == hi) break;
            long double const cdf = (2.0L + mid * (3.0L - mid * mid)) / 4.0L;
            if (cdf < u) lo = mid; else hi = mid;
        }
        return static_cast<double>(0.5L * (lo + hi));
    }
    long double const b = 0.5L * gamma * (gamma - 1.0L) * (gamma - 2.0L);
    long double const c = -inv_gamma_sq;
    long double const t_lo = inv_gamma_sq / (1.0L + beta);
    long double const t_hi = 1.0L + beta;
    auto primitive = [c, b](long double x) {
        return -c / (3.0L * x * x * x) - (2.0L + b * c) / (2.0L * x * x) - (2.0L * b - 1.0L) / x - b * std::log(x);
    };
    long double const f_lo = primitive(t_lo);
    long double const f_hi = primitive(t_hi);
    long double const total = f_hi - f_lo;
    if (!(total > 0.0L) || !std::isfinite(total)) fail(ORC_ERR_INVALID, "Sauter-Gavrila normalization invalid");
    long double const target = (1.0L - u) * f_hi + u * f_lo;
    long double lo = t_lo, hi = t_hi;
    for (int it = 0; it < 128; ++it) {
        long double const mid = 0.5L * (lo + hi);
        if (mid == lo || mid == hi) break;
        if (primitive(mid) < target) lo = mid; else hi = mid;
    }
    long double const tt = 0.5L * (lo + hi);
    long double const mu = std::max(-1.0L, std::min(1.0L, (1.0L - tt) / beta));
    return static_cast<double>(mu);
}

struct PairState {
    bool valid = false;
    double electron = 0, positron = 0, recoil = 0, local = 0;
    double edx = 0, edy = 0, edz = 0, pdx = 0, pdy = 0, pdz = 0, rdx = 0, rdy = 0, rdz = 0;
    bool has_recoil = false;
};

// normalize_local_direction NA/src/RreaInteractionTables.cpp:121-133
inline void normalize_local(double& x, double& y, double& z)
{
    double const n = std::sqrt(x * x + y * y + z * z);
    if (n <= 0.0 || !std::isfinite(n)) { x = 0.0; y = 0.0; z = 1.0; return; }
    x /= n; y /= n; z /= n;
}

// SamplePairFinalState :3053-3127 + BuildRreaPairEmpiricalFinalState
// :1384-1525: replay of one training record of the bracketing energy block
inline PairState sample_pair(Tables const& t, double e, int channel, double draw_invariant, double draw_block,
                             double draw_azimuth)
{
    if (!std::isfinite(e) || e < t.photon_min || e > t.photon_max) {
        fail(ORC_ERR_TABLE_RANGE, "pair energy outside certified range");
    }
    PairState st;
    bool const triplet = channel == kPairTriplet;
    double const threshold = triplet ? 4.0 * kMeC2 : 2.0 * kMeC2;  // RreaPairChannelHasPhaseSpace :1373-1382
    if (!(std::isfinite(e) && e > threshold)) return st;
    int32_t const code = triplet ? 2 : 1;
    std::vector<size_t> blocks;
    for (size_t i = 0; i < t.pair_channel.size(); ++i) if (t.pair_channel[i] == code) blocks.push_back(i);
    if (blocks.empty()) fail(ORC_ERR_INVALID, "pair channel has no training blocks");
    size_t sel = blocks.front();
    if (e >= t.pair_energy[blocks.back()]) {
        sel = blocks.back();
    } else if (e > t.pair_energy[blocks.front()]) {
        size_t up = 0;
        while (t.pair_energy[blocks[up]] < e) ++up;  // lower_bound over the channel's blocks
        if (t.pair_energy[blocks[up]] == e) {
            sel = blocks[up];
        } else {
            double const lo_e = t.pair_energy[blocks[up - 1]], hi_e = t.pair_energy[blocks[up]];
            double const tt = (std::log(e) - std::log(lo_e)) / (std::log(hi_e) - std::log(lo_e));
            sel = clamp_unit(draw_block) < tt ? blocks[up] : blocks[up - 1];
        }
    }
    int64_t const size = t.pair_start[sel + 1] - t.pair_start[sel];
    if (size <= 0) fail(ORC_ERR_INVALID, "selected pair block is empty");
    size_t const rec_in_block = std::min(static_cast<size_t>(size - 1),
                                         static_cast<size_t>(clamp_unit(draw_invariant) * static_cast<double>(size)));
    size_t const rec = static_cast<size_t>(t.pair_start[sel]) + rec_in_block;
    double const* f = &t.pair_frac[3 * rec];
    double const* d = &t.pair_dir[9 * rec];
    double const available = e - 2.0 * kMeC2;
    double const rotation = kTwoPi * clamp_unit(draw_azimuth);
    double const cr = std::cos(rotation), sr = std::sin(rotation);
    auto rotated = [&](double const* v, double& x, double& y, double& z) {
        x = cr * v[0] - sr * v[1];
        y = sr * v[0] + cr * v[1];
        z = v[2];
        normalize_local(x, y, z);
    };
    st.valid = true;
    st.electron = available * f[0];
    st.positron = available * f[1];
    rotated(d + 0, st.edx, st.edy, st.edz);
    rotated(d + 3, st.pdx, st.pdy, st.pdz);
    if (triplet) {
        st.has_recoil = true;
        st.recoil = available * f[2];
        rotated(d + 6, st.rdx, st.rdy, st.rdz);
    }
    double const tol = 64.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, available);
    auto sum = [&]() { return st.electron + st.positron + st.recoil; };
    double expl = sum();
    if (!std::isfinite(expl) || st.electron < 0.0 || st.positron < 0.0 || st.recoil < 0.0) {
        fail(ORC_ERR_INVALID, "pair record produced a negative/nonfinite energy");
    }
    if (expl > available) {
        if (expl - available > tol) fail(ORC_ERR_INVALID, "pair record failed energy closure");
        double const scale = available / expl;
        st.electron *= scale;
        st.positron *= scale;
        st.recoil *= scale;
        expl = sum();
    }
    st.local = available - expl;
    auto reduce_largest = [&](double correction) {
        double* e3[3] = {&st.electron, &st.positron, &st.recoil};
        double** largest = std::max_element(e3, e3 + 3, [](double const* a, double const* b) { return *a < *b; });
        if (**largest < correction) fail(ORC_ERR_INVALID, "pair record cannot close nonnegative remainder");
        **largest -= correction;
    };
    if (st.local < 0.0) {
        double const correction = -st.local;
        if (correction > tol) fail(ORC_ERR_INVALID, "pair record remainder is negative");
        reduce_largest(correction);
        st.local = 0.0;
    }
    double accounted = sum() + st.local;
    double residual = available - accounted;
    if (std::abs(residual) > tol) fail(ORC_ERR_INVALID, "pair record failed energy closure");
    if (residual >= 0.0) st.local += residual;
    else if (st.local >= -residual) st.local += residual;
    else reduce_largest(-residual);
    accounted = sum() + st.local;
    if (st.electron < 0.0 || st.positron < 0.0 || st.recoil < 0.0 || st.local < 0.0
        || std::abs(available - accounted) > tol) {
        fail(ORC_ERR_INVALID, "pair record closure correction failed");
    }
    return st;
}

// RreaRotateLocalDirectionToAxis OV/RreaChargedInterleave.H:498-517
inline void rotate_to_axis(double ax, double ay, double az, double lx, double ly, double lz, double& ox, double& oy,
                           double& oz)
{
    normalize(ax, ay, az);
    normalize(lx, ly, lz);
    double b1x, b1y, b1z, b2x, b2y, b2z;
    transverse_basis(ax, ay, az, b1x, b1y, b1z, b2x, b2y, b2z);
    ox = lx * b1x + ly * b2x + lz * ax;
    oy = lx * b1y + ly * b2y + lz * ay;
    oz = lx * b1z + ly * b2z + lz * az;
    normalize(ox, oy, oz);
}

// The photon final states at an interaction site, for the live loop and the
// newborn residual advance alike: Compton (RreaApplyComptonFinalState
// OV/RreaPhotonFinalStates.H:54-140), photoelectric absorption
// (RreaApplyPhotoelectricFinalState :153-224) and pair production
// (RreaApplyPairFinalState :239-363).  Draws are keyed by the emitter's RNG
// key and the pass's interaction index.  Returns true when the photon
// survives (Compton at or above the photon cutoff), with its new energy and
// direction; false when it is absorbed.
inline bool photon_final_state(StepContext const& ctx, Emitter& em, int channel, uint64_t idx, double x, double y,
                               double z, double birth, double& energy, double& dx, double& dy, double& dz)
{
    Tables const& t = *ctx.t;
    ChunkLog& fx = em.fx;
    double const cutoff = ctx.cfg.low_energy_cutoff_eV;
    double const gcut = ctx.cfg.photon_cutoff_eV;
    double const w = em.w;
    uint64_t const step = static_cast<uint64_t>(std::max<int64_t>(ctx.cfg.step, 0));
    auto U = [&](uint64_t chan) { return rng_uniform(ctx.cfg.rng_seed, em.pkey, step, idx, chan); };
    double const sr = std::hypot(x, y);  // RreaInteractionSite::r_m of the event
    Ledger L{fx.tallies, cutoff};
    if (channel == kCompton) {
        double const incident = energy;
        ComptonState const st = sample_compton(t, incident, U(kChComptonRecoil), U(kChComptonScatter));
        double sx, sy, sz;
        cone_direction(dx, dy, dz, st.cos_theta, U(kChComptonScatterAzimuth), sx, sy, sz);
        double rx = incident * dx - st.scattered * sx;
        double ry = incident * dy - st.scattered * sy;
        double rz = incident * dz - st.scattered * sz;
        normalize(rx, ry, rz);
        if (st.recoil >= cutoff) {
            em.secondary(0, kProcCompton, x, y, z, rx, ry, rz, st.recoil, birth);
            em.ion_event(w);
            em.ionization(sr, z, w, st.recoil, st.recoil);
        } else {
            em.ion_event(w);
            em.ionization(sr, z, w, st.recoil, 0.0);
            em.deposit_local(st.recoil, sr, z);
        }
        fx.tallies[ORC_T_COMPTON_EVENTS] += 1.0;
        L.transport_ionization(w, 0.0, !(st.recoil >= cutoff), st.recoil >= cutoff);
        {
            EnergyFlow const f = compton_flow(st.recoil, st.scattered, cutoff, gcut);
            L.compton(w, st.recoil, f.charged, f.local);
        }
        {
            Closure c;
            bool const survives = !(st.scattered < gcut);
            c.parent(2, w, incident, st.scattered, survives);
            if (!survives) c.local(w * st.scattered);
            if (st.recoil >= cutoff) c.particle(0, w, st.recoil);
            else c.local(w * st.recoil);
            c.ionization(w, st.recoil, cutoff);
            require_conserved(c, cutoff);
        }
        energy = st.scattered;
        if (energy < gcut) {
            em.deposit_local(energy, sr, z);
            return false;
        }
        dx = sx; dy = sy; dz = sz;
        return true;
    }
    if (channel == kPhotoelectric) {
        PhotoelectricState const st = sample_photoelectric(t, energy, U(kChPhotoelectric));
        double const ke = std::min(st.electron, energy);
        if (ke >= cutoff) {
            double ex, ey, ez;
            cone_direction(dx, dy, dz, sample_photoelectron_cos(t, ke, U(kChPhotoelectricAngle)),
                           U(kChPhotoelectricAzimuth), ex, ey, ez);
            em.secondary(0, kProcPhotoelectric, x, y, z, ex, ey, ez, ke, birth);
            em.ion_event(w);
            em.ionization(sr, z, w, ke, ke);
        } else {
            em.ion_event(w);
            em.ionization(sr, z, w, ke, 0.0);
            em.deposit_local(ke, sr, z);
        }
        L.transport_ionization(w, 0.0, !(ke >= cutoff), ke >= cutoff);
        if (st.binding > 0.0) em.deposit_local(st.binding, sr, z);
        fx.tallies[ORC_T_PHOTOELECTRIC_EVENTS] += 1.0;
        {
            EnergyFlow const f = photoelectric_flow(ke, st.binding, cutoff);
            L.photoelectric(w, energy, ke, f.charged, f.local);
        }
        Closure c;
        c.parent(2, w, energy, 0.0, false);
        if (ke >= cutoff) c.particle(0, w, ke);
        else c.local(w * ke);
        c.ionization(w, ke, cutoff);
        c.local(w * std::max(st.binding, 0.0));
        require_conserved(c, cutoff);
        return false;
    }
    PairState const pair = sample_pair(t, energy, channel, U(kChPairInvariant), U(kChPairAngle), U(kChPairAzimuth));
    if (!pair.valid) fail(ORC_ERR_INVALID, "pair channel without valid final-state phase space");
    bool const triplet = channel == kPairTriplet;
    if (triplet && !pair.has_recoil) fail(ORC_ERR_INVALID, "triplet final state without its recoil");
    int32_t const process = triplet ? kProcPairTriplet : kProcPairNuclear;
    double ex, ey, ez, px, py, pz2;
    rotate_to_axis(dx, dy, dz, pair.edx, pair.edy, pair.edz, ex, ey, ez);
    rotate_to_axis(dx, dy, dz, pair.pdx, pair.pdy, pair.pdz, px, py, pz2);
    if (pair.electron >= cutoff) {
        em.secondary(0, process, x, y, z, ex, ey, ez, pair.electron, birth);
    } else {
        fx.tallies[ORC_T_LOW_ENERGY_ELECTRONS] += w;
        em.low_electron(sr, z, w, 0.0);
        em.deposit_local(pair.electron, sr, z);
    }
    if (pair.positron >= cutoff) {
        em.secondary(1, process, x, y, z, px, py, pz2, pair.positron, birth);
    } else {
        spawn_annihilation(ctx, em, process, x, y, z, sr, px, py, pz2, pair.positron, true, idx, birth);
        L.positron_transport(w, pair.positron, true);  // OV/RreaPhotonFinalStates.H:320
    }
    if (pair.has_recoil) {
        double qx, qy, qz;
        rotate_to_axis(dx, dy, dz, pair.rdx, pair.rdy, pair.rdz, qx, qy, qz);
        if (pair.recoil >= cutoff) em.secondary(0, process, x, y, z, qx, qy, qz, pair.recoil, birth);
        em.ion_event(w);
        em.ionization(sr, z, w, pair.recoil, 0.0);
        if (pair.recoil < cutoff) em.deposit_local(pair.recoil, sr, z);
    }
    if (pair.local > 0.0) em.deposit_local(pair.local, sr, z);
    fx.tallies[ORC_T_PAIR_EVENTS] += 1.0;
    {
        EnergyFlow const f = pair_flow(pair.electron, pair.positron, pair.recoil, pair.local, cutoff, pair.has_recoil);
        L.pair(w, energy, pair.electron, pair.positron, pair.recoil, f.charged, f.local, triplet);
    }
    {
        // a sub-cutoff positron leaves as the particle its at-rest
        // annihilation (a group checked of its own) consumes
        Closure c;
        c.parent(2, w, energy, 0.0, false);
        if (pair.electron >= cutoff) c.particle(0, w, pair.electron);
        else { c.low_electron(w); c.local(w * pair.electron); }
        c.particle(1, w, pair.positron);
        if (pair.has_recoil) {
            if (pair.recoil >= cutoff) c.particle(0, w, pair.recoil);
            else c.local(w * pair.recoil);
            c.ionization(w, pair.recoil, cutoff);
        }
        c.local(w * std::max(pair.local, 0.0));
        require_conserved(c, cutoff);
    }
    return false;
}

// One photon: the body of the particle loop of OV/RreaPhotonTransport.cpp:56-438.
// The chord is the captured pre-push point to WarpX's pushed point; the loop
// samples one exponential depth over the remaining chord per pass, and a
// Compton survivor continues from its interaction point along the scattered
// direction for the unused path length.
inline void photon_particle(StepContext const& ctx, ChunkLog& fx, int64_t ip, double pr, double pth, double pz,
                            double cr, double cth, double cz, double ux, double uy, double uz, double w, int64_t id,
                            int32_t cpu, ChargedOut& out)
{
    Tables const& t = *ctx.t;
    orc_config const& cfg = ctx.cfg;
    double const gcut = cfg.photon_cutoff_eV;
    double const dt = cfg.dt_s;
    uint64_t const step = static_cast<uint64_t>(std::max<int64_t>(cfg.step, 0));
    uint64_t const seed = cfg.rng_seed;
    uint64_t const pkey = static_cast<uint64_t>(std::llabs(id)) ^ (static_cast<uint64_t>(static_cast<uint32_t>(cpu)) << 40);
    out.ion_pairs = 0.0;
    out.substeps = 0;
    out.r = cr; out.z = cz; out.th = cth; out.ux = ux; out.uy = uy; out.uz = uz;
    Emitter em{fx, ip, pkey, w, out.ion_pairs, 2};

    double energy = std::sqrt(ux * ux + uy * uy + uz * uz);
    double dx = ux, dy = uy, dz = uz;
    normalize(dx, dy, dz);
    if (pr > ctx.r_hi + 1.0e-9 * ctx.r_hi || pz < ctx.z_lo - 1.0e-9 || pz > ctx.z_hi + 1.0e-9 * ctx.z_hi) {
        fail(ORC_ERR_OUTSIDE_DOMAIN, "photon survived outside the domain");
    }
    Path path = clip_rz_path(pr, pth, pz, cr, cth, cz, ctx.r_lo, ctx.r_hi, ctx.z_lo, ctx.z_hi);
    uint64_t ordinal = 0;
    double elapsed_start = 0.0;
    for (;;) {
        if (ordinal >= kEventCap) fail(ORC_ERR_INVALID, "photon event ordinal overflow");
        EscapePolicy const escape = clipped_escape_policy(path, std::hypot(path.ex, path.ey), path.ez);
        double const ds = path.in_length;
        std::vector<Interval> const iv = build_intervals(ctx.cp, path, ctx.density_ratio);
        double const chord_dp = density_path(iv, path, path.begin_f, path.end_f);
        if (w <= 0.0 || energy < gcut) {
            // a photon below the production cutoff deposits at its path start
            double x, y, z;
            path.point_at_full(path.begin_f, x, y, z);
            em.deposit_local(energy, std::hypot(x, y), z);
            out.status = w <= 0.0 ? ORC_ZERO_WEIGHT : ORC_DEMOTED;
            if (w > 0.0) fx.tallies[ORC_T_DEMOTIONS] += 1.0;
            return;
        }
        auto commit_no_event = [&]() {
            double x, y, z;
            path.point_at_full(path.end_f, x, y, z);
            out.r = std::hypot(x, y);
            out.th = std::atan2(y, x);
            out.z = z;
            if (escape.escapes) {
                out.status = ORC_ESCAPE;
                fx.tallies[ORC_T_ESCAPES] += 1.0;
                Ledger{fx.tallies, 0.0}.photon_escape(w, energy);  // OV/RreaPhotonTransport.cpp:192
            } else {
                out.status = ORC_SURVIVE;
                fx.tallies[ORC_T_SURVIVORS] += 1.0;
            }
        };
        if (ds <= 0.0) { commit_no_event(); return; }
        double const mfp = photon_mfp(t, energy);
        double const tau = chord_dp / mfp;
        // one exponential depth per pass (OV/RreaPhotonTransport.cpp:222-233)
        Ledger{fx.tallies, 0.0}.guard(0.0, 0.0, 0.0, tau, 0.0, 1U, 0.0, energy, 0.0, false);
        uint64_t idx = 0;
        if (!live_hazard_index(2U, 0U, ordinal, idx)) fail(ORC_ERR_INVALID, "photon hazard RNG index overflow");
        double const draw = rng_uniform(seed, pkey, step, idx,
                                        ordinal == 0U ? kChPhotonInteraction : kChPhotonContinuationOpticalDepth);
        double const taus[3] = {tau, 0.0, 0.0};
        Hazard const h = select_hazard(taus, 1, draw, 0.0);
        if (!h.event) { commit_no_event(); return; }
        double const f = fraction_at_density_path(iv, path, path.begin_f, path.end_f, h.sampled * mfp);
        double const dist = path.full_length * (f - path.begin_f);
        double x, y, z;
        path.point_at_full(f, x, y, z);
        double const remaining = std::max(0.0, path.full_length - dist);
        double const birth = std::min(1.0, elapsed_start + dist / std::max(kC * dt, std::numeric_limits<double>::min()));
        double const r_site = std::hypot(x, y);
        int const channel = sample_photon_channel(
            t, energy, rng_uniform(seed, pkey, step, idx, ordinal == 0U ? kChPhotonChannel : kChPhotonContinuationChannel));
        out.substeps = ordinal + 1U;
        if (!photon_final_state(ctx, em, channel, idx, x, y, z, birth, energy, dx, dy, dz)) {
            out.status = ORC_ABSORBED;
            return;
        }
        // a Compton survivor continues from the interaction point along its
        // scattered direction for the unused path length
        out.ux = energy * dx; out.uy = energy * dy; out.uz = energy * dz;
        double const xt = x + dx * remaining, yt = y + dy * remaining, zt = z + dz * remaining;
        Path const cont = clip_rz_path(r_site, std::atan2(y, x), z, std::hypot(xt, yt), std::atan2(yt, xt), zt,
                                       ctx.r_lo, ctx.r_hi, ctx.z_lo, ctx.z_hi);
        double xn, yn, zn;
        cont.point_at_full(cont.end_f, xn, yn, zn);
        out.r = std::hypot(xn, yn);
        out.th = std::atan2(yn, xn);
        out.z = zn;
        if (!(remaining > 0.0)) {
            out.status = ORC_SURVIVE;
            fx.tallies[ORC_T_SURVIVORS] += 1.0;
            return;
        }
        path = cont;
        elapsed_start = birth;
        ++ordinal;
    }
}


// ---------------------------------------------------------------------------
// Newborn residual advance -- OV/RreaNewbornAdvance.cpp -- and the transient
// FIFO of DrainAndCommitSecondaries (OV/RreaSecondaryCommit.cpp:338-468).
// A secondary born at elapsed fraction f < 1 is flown through the remaining
// (1 - f) dt with the live loops' physics; its draws come from the ROOT
// particle's stream, indexed by the packed transient lineage.
// ---------------------------------------------------------------------------
enum : uint64_t {
    kChNewbornElectronGsAngle = 112,
    kChNewbornElectronGsAzimuth = 113,
    kChNewbornElectronGsFewCount = 114,
    kChNewbornElectronGsFewScatter = 115,
    kChNewbornElectronGsFewAzimuth = 116,
    kChNewbornPositronGsAngle = 117,
    kChNewbornPositronGsAzimuth = 118,
    kChNewbornPositronGsFewCount = 119,
    kChNewbornPositronGsFewScatter = 120,
    kChNewbornPositronGsFewAzimuth = 121,
    kChNewbornCompetingOpticalDepth = 122,
    kChNewbornCompetingChannel = 123,
};

// RreaNewbornResidualFraction OV/RreaParticleInteraction.H:171-179
inline double newborn_residual_fraction(double birth)
{
    return std::min(1.0, std::max(0.0, 1.0 - birth));
}

// RreaNewbornElapsedStepFraction OV/RreaParticleInteraction.H:181-197
inline double newborn_elapsed(double birth, double residual, uint64_t completed, uint64_t substeps, double f)
{
    double const within = std::min(1.0, std::max(0.0, f));
    double const consumed = (static_cast<double>(completed) + within) / static_cast<double>(substeps);
    return std::min(1.0, birth + residual * consumed);
}

// TransportStep::NewbornFinalPoint OV/RreaNewbornAdvance.cpp:20-48: the point
// at `fraction`, pulled back along the chord (1e-9, 1e-7, 1e-5, 1e-3 of the
// in-domain span) until it lies in the half-open domain.
inline bool newborn_final_point(Path const& path, double fraction, double& ox, double& oy, double& oz)
{
    double const span = fraction - path.begin_f;
    double pullback = 1.0e-9;
    for (int attempt = 0; attempt < 5; ++attempt) {
        double const trial = attempt == 0 ? fraction
                                          : std::max(path.begin_f, fraction - pullback * std::max(span, 1.0e-12));
        path.point_at_full(trial, ox, oy, oz);
        if (in_half_open(ox, oy, oz, path.r_lo, path.r_hi, path.z_lo, path.z_hi)) return true;
        if (attempt > 0) pullback *= 100.0;
    }
    return false;
}

// The newborn's electron_state_policy_at (OV/RreaNewbornAdvance.cpp:79-96):
// RreaEvaluateElectronStatePolicy (OV/RreaParticleInteraction.H:1296-1317)
// with the field gathered at the point and the density ratio at its z.
inline bool newborn_electron_demotes(StepContext const& ctx, double x, double y, double z, double k, bool hm_on)
{
    double const cutoff = ctx.cfg.low_energy_cutoff_eV;
    double const thr = ctx.t->hard_moller_threshold;
    double er, ez, bt;
    ctx.cp.gather(std::hypot(x, y), z, er, ez, bt);
    double const ratio = positive_or(ctx.cp.density_ratio_at_z(z), 1.0);
    double const drag = ensemble_drag(*ctx.t, std::max(k, cutoff), ratio, hm_on);
    bool const below = k < cutoff;
    bool const nonrunaway = thr > 0.0 && k <= thr && std::hypot(er, ez) <= drag;
    return below || nonrunaway;
}

struct NewbornOut {
    int status = ORC_SURVIVE;
    double ion_pairs = 0.0;
    uint64_t substeps = 0;  // charged: the residual substep count; photon: interactions
};

// TransportStep::AdvanceChargedNewborn OV/RreaNewbornAdvance.cpp:50-777.  On
// survival `p` holds the state at the end of the step (birth fraction 1).
inline void charged_newborn(StepContext const& ctx, ChunkLog& fx, orc_secondary& p, int64_t ip, NewbornOut& out)
{
    Tables const& t = *ctx.t;
    orc_config const& cfg = ctx.cfg;
    bool const pos = p.species == 1;
    double const cutoff = cfg.low_energy_cutoff_eV;
    double const gcut = cfg.photon_cutoff_eV;
    double const thr = t.hard_moller_threshold;
    bool const hm_on = cfg.disable_hard_moller == 0;
    double const dt = cfg.dt_s;
    double const charge = pos ? 1.0 : -1.0;
    uint64_t const step = static_cast<uint64_t>(std::max<int64_t>(cfg.step, 0));
    uint64_t const seed = cfg.rng_seed;
    uint64_t const root = p.root_key;
    uint64_t const rsp = static_cast<uint64_t>(p.root_species);
    uint64_t const lin = p.lineage;
    double const w = p.weight;
    Emitter em{fx, ip, root, w, out.ion_pairs, p.root_species, p.generation + 1};
    Ledger L{fx.tallies, cutoff};
    auto pack = [&](uint64_t segment, uint64_t ordinal, uint64_t inner) {
        uint64_t v = 0;
        if (!pack_transient_index(true, rsp, lin, segment, ordinal, inner, v)) {
            fail(ORC_ERR_INVALID, "newborn RNG index overflow");
        }
        return v;
    };
    double const birth_r = std::hypot(p.x, p.y);
    double const birth_z = p.z;
    double const residual = newborn_residual_fraction(p.birth_fraction);
    double k = p.energy_eV;
    if (!in_half_open(p.x, p.y, p.z, ctx.r_lo, ctx.r_hi, ctx.z_lo, ctx.z_hi)) {
        fail(ORC_ERR_OUTSIDE_DOMAIN, "charged transient born outside the half-open domain");
    }
    double dx = p.dir_x, dy = p.dir_y, dz = p.dir_z;
    normalize(dx, dy, dz);
    // the net chord from the birth point
    auto segment = [&](double r1, double z1) {
        fx.deposits.push_back(DepositOp{0, {birth_r, birth_z, r1, z1, charge * w, 0.0, 0.0}, 0U});
    };
    auto escape_along = [&](Path const& path) {
        BoundaryCrossing const bc = boundary_crossing(path);
        if (!bc.valid) fail(ORC_ERR_INVALID, "newborn escape has no certified absorbing-face crossing");
        fx.deposits.push_back(
            DepositOp{1, {birth_r, birth_z, bc.start_r, bc.start_z, bc.cross_r, bc.cross_z, charge * w}, bc.faces});
        out.status = ORC_ESCAPE;
        fx.tallies[ORC_T_NEWBORN_ESCAPES] += 1.0;
    };
    auto in_range = [&]() {
        if (!std::isfinite(k) || k < t.charged_min || k > t.charged_max) {
            fail(ORC_ERR_TABLE_RANGE, "newborn charged energy left the certified range");
        }
    };

    bool const birth_demote = pos ? k < cutoff : newborn_electron_demotes(ctx, p.x, p.y, p.z, k, hm_on);
    if (birth_demote) {
        if (pos) {
            spawn_annihilation(ctx, em, kProcAnnihilation, p.x, p.y, p.z, std::hypot(p.x, p.y), dx, dy, dz, k, true,
                               pack(0, 0, 0), p.birth_fraction);
            L.positron_transport(w, k, true);  // OV/RreaNewbornAdvance.cpp:123
            out.status = ORC_ANNIHILATED;
        } else {
            fx.tallies[ORC_T_LOW_ENERGY_ELECTRONS] += w;
            em.low_electron(std::hypot(p.x, p.y), p.z, w, 0.0);
            em.deposit_local(k, std::hypot(p.x, p.y), p.z);
            out.status = ORC_DEMOTED;
        }
        fx.tallies[ORC_T_NEWBORN_TERMINATED] += 1.0;
        return;
    }
    if (!(residual > 0.0)) return;
    double const birth_u = proper_velocity_from_kinetic(k);
    double const birth_gamma = std::sqrt(1.0 + birth_u * birth_u / (kC * kC));
    double const residual_dt = dt * residual;
    double const chord = (birth_u / birth_gamma) * residual_dt;
    if (!(chord > 0.0)) {
        if (k < cutoff) out.status = ORC_DEMOTED;
        return;
    }
    double const end_x = p.x + chord * dx;
    double const end_y = p.y + chord * dy;
    double const end_z = p.z + chord * dz;
    // provisional whole-residual chord, used only to size the substeps
    Path const driver = clip_rz_path(std::hypot(p.x, p.y), std::atan2(p.y, p.x), p.z, std::hypot(end_x, end_y),
                                     std::atan2(end_y, end_x), end_z, ctx.r_lo, ctx.r_hi, ctx.z_lo, ctx.z_hi);
    if (!(driver.full_length > 0.0) || !(driver.in_length > 0.0)) {
        if (driver.exits) escape_along(driver);
        return;
    }
    std::vector<Interval> const div = build_intervals(ctx.cp, driver, ctx.density_ratio);
    if (div.empty()) fail(ORC_ERR_INVALID, "newborn path without material intervals");
    Path path = driver;
    std::vector<Interval> iv = div;
    double const driver_dp = density_path(div, driver, driver.begin_f, driver.end_f);
    double const gs_n_total = elastic_inverse_length(t, pos, k) * driver_dp;
    double const drag_est = (pos ? t.rate[kPCol].interpolate(k) * 1.0 : realized_loss(t, k, 1.0)) * driver_dp;
    double const whole_work = path_field_work(ctx.cp, driver, charge, 0.0, 1.0).work;
    double const impulse = std::max(cfg.max_field_impulse_fraction * k, std::max(cfg.field_impulse_floor_eV, 1.0));
    double const span = std::max(driver.end_f, 1.0e-6);
    uint64_t const substeps = std::min<uint64_t>(
        std::max(driver_count(std::min(gs_n_total / span, cfg.gs_total_collision_cap),
                              std::max(cfg.max_gs_collisions_per_substep, 1.0)),
                 driver_count((std::abs(whole_work) + drag_est) / span, impulse)),
        cfg.max_subcycles);
    if (substeps >= kSegmentCap) fail(ORC_ERR_INVALID, "newborn substep count outside its transient RNG budget");
    out.substeps = substeps;

    double run_x = p.x, run_y = p.y, run_z = p.z;
    double const sub_dt = residual_dt / static_cast<double>(substeps);
    double final_f = 0.0;
    bool terminal_interaction = false;
    bool terminal_at_point = false;
    bool exited = false;
    uint64_t completed = 0;
    auto elapsed_at = [&](double f) { return newborn_elapsed(p.birth_fraction, residual, completed, substeps, f); };
    auto mark_stopped = [&](double f) {
        final_f = f;
        terminal_at_point = true;
    };
    // RreaPlanContinuousDragSegment OV/RreaChargedInterleave.H:239-292 with the
    // cutoff limiter :203-265 and the ion-pair partition :157-196
    double positron_total_loss = 0.0;
    auto continuous = [&](double b, double e, double stop_e) -> SegLimit {
        DragPlan plan;
        SegLimit const sl = continuous_drag(
            t, pos, path, iv, b, e, stop_e, cutoff, w, k, plan,
            [&] { return (!pos && hm_on && thr > 0.0 && stop_e > 2.0 * thr) ? t.hard_moller_energy_loss(stop_e, 1.0) : 0.0; },
            [&](double mid_f, double pw) {
                em.ion_event(pw);
                double mx, my, mz;
                path.point_at_full(mid_f, mx, my, mz);
                em.ionization(std::hypot(mx, my), mz, pw, 0.5 * cutoff, kWAir);
                L.transport_ionization(pw, kWAir, true, false);
            });
        fx.tallies[ORC_T_DRAG_LOSS_WEV] += w * plan.drag;
        // OV/RreaNewbornAdvance.cpp:309-320
        L.soft_brems(w, plan.soft);
        if (pos) positron_total_loss += plan.drag;
        else L.loss_partition(w, plan.drag, plan.collision, plan.hard);
        return sl;
    };

    FieldRange sfield;
    double substep_dt = 0.0;
    int const nch = pos ? 3 : 2;
    auto build_substep = [&](uint64_t, double& sb, double& se) -> bool {
        Chord const ch = build_chord(run_x, run_y, run_z, dx, dy, dz, k, sub_dt, residual_dt, ctx.r_lo, ctx.r_hi,
                                     ctx.z_lo, ctx.z_hi);
        if (!ch.valid) return false;
        path = ch.path;
        if (!(path.full_length > 0.0) || !(path.in_length > 0.0)) {
            exited = path.exits;
            final_f = 0.0;
            return false;
        }
        iv = build_intervals(ctx.cp, path, ctx.density_ratio);
        if (iv.empty()) fail(ORC_ERR_INVALID, "newborn microchord without material intervals");
        sb = path.begin_f;
        se = path.end_f;
        final_f = se;
        sfield = path_field_work(ctx.cp, path, charge, sb, se);
        substep_dt = sub_dt * (se - sb);
        (void)half_kick(sfield, 0.5 * substep_dt, charge, k, dx, dy, dz);
        if (k < cutoff) {
            final_f = sb;
            terminal_at_point = true;
            return false;
        }
        in_range();
        return true;
    };
    auto run_events = [&](uint64_t sc, double sb, double se) {
        uint64_t idx = 0;
        double stp[3] = {0, 0, 0};
        double ex_ = 0.0, ey_ = 0.0, ez_ = 0.0, er_ = 0.0;
        run_event_interleave(
            sb, se, k, cutoff, nch, kEventCap,
            [&](double frozen_e, double rem, double end) {
                if (pos) positron_stp_rates(t, thr, gcut, frozen_e, stp);
                else electron_stp_rates(t, thr, gcut, frozen_e, hm_on, stp);
                double const dp = density_path(iv, path, rem, end);
                std::array<double, 3> const taus{stp[0] * dp, stp[1] * dp, stp[2] * dp};
                if (!pos) {
                    // OV/RreaNewbornAdvance.cpp:536-548
                    double const total_tau = taus[0] + taus[1];
                    L.hm_expectation(w * (total_tau > 0.0 ? probability_from_optical_depth(total_tau) * taus[0] / total_tau
                                                          : 0.0));
                }
                return taus;
            },
            [&](uint64_t ord) {
                idx = pack(sc, ord, 0U);
                return std::pair<double, double>{rng_uniform(seed, root, step, idx, kChNewbornCompetingOpticalDepth),
                                                 rng_uniform(seed, root, step, idx, kChNewbornCompetingChannel)};
            },
            [&](double tau, double rem, double end) {
                double const total_stp = stp[0] + stp[1] + stp[2];
                return fraction_at_density_path(iv, path, rem, end, tau / total_stp);
            },
            continuous, mark_stopped,
            [&](double ef) {
                path.point_at_full(ef, ex_, ey_, ez_);
                er_ = std::hypot(ex_, ey_);
            },
            [&](int channel, double e) { return charged_channel_open(t, pos, thr, gcut, channel, e, hm_on); },
            [&](int channel, double ef) -> bool {
                double const efrac = elapsed_at(ef);
                if (!pos && channel == 0) {
                    double const before = k;
                    double T = t.sample_moller(before, rng_uniform(seed, root, step, idx, kChHardMollerEnergy));
                    T = std::max(thr, std::min(T, 0.5 * before));
                    double pdx, pdy, pdz, sdx, sdy, sdz;
                    moller_lab_directions(dx, dy, dz, before, T, rng_uniform(seed, root, step, idx, kChHardMollerAzimuth),
                                          pdx, pdy, pdz, sdx, sdy, sdz);
                    k = std::max(0.0, before - T);
                    dx = pdx; dy = pdy; dz = pdz;
                    em.secondary(0, kProcHardMoller, ex_, ey_, ez_, sdx, sdy, sdz, T, efrac);
                    em.ion_event(w);
                    em.ionization(std::hypot(ex_, ey_), ez_, w, T, T);
                    fx.tallies[ORC_T_MOLLER_EVENTS] += 1.0;
                    L.hm_event(w, T);
                    L.transport_ionization(w, T, false, true);
                    Closure c;
                    c.parent(0, w, before, k, true);
                    c.particle(0, w, T);
                    c.ionization(w, T, cutoff);
                    require_conserved(c, cutoff);
                } else if (pos && channel == 0) {
                    double const before = k;
                    double const T = std::min(t.sample_bhabha(before, rng_uniform(seed, root, step, idx, kChPositronBhabhaEnergy)),
                                              std::max(0.0, before));
                    if (T > 0.0 && !(thr > 0.0 && T < thr)) {
                        double pdx, pdy, pdz, sdx, sdy, sdz;
                        moller_lab_directions(dx, dy, dz, before, T,
                                              rng_uniform(seed, root, step, idx, kChPositronBhabhaAzimuth),
                                              pdx, pdy, pdz, sdx, sdy, sdz);
                        k = std::max(0.0, before - T);
                        dx = pdx; dy = pdy; dz = pdz;
                        em.secondary(0, kProcHardBhabha, ex_, ey_, ez_, sdx, sdy, sdz, T, efrac);
                        em.positive_ion(std::hypot(ex_, ey_), ez_, w, T);
                        fx.tallies[ORC_T_POSITIVE_IONS] += w;
                        fx.tallies[ORC_T_BHABHA_EVENTS] += 1.0;
                        L.transport_ionization(w, T, false, true);
                        L.bhabha(w, T, true);
                        Closure c;
                        c.parent(1, w, before, k, true);
                        c.particle(0, w, T);
                        c.positive_ion(w);
                        require_conserved(c, cutoff);
                    }
                } else if (channel == 1) {
                    double const before = k;
                    double const eu = rng_uniform(seed, root, step, idx, pos ? kChPositronBremsPhotonEnergy : kChBremsPhotonEnergy);
                    double const ks = pos ? t.sample_positron_brems_energy(before, eu) : t.sample_brems_energy(before, eu);
                    double const pe = std::min(std::max(ks, 0.0), std::max(before, 0.0));
                    double const after = std::max(0.0, std::max(before, 0.0) - pe);
                    bool const tracked = std::isfinite(gcut) && gcut >= 0.0 && pe >= gcut;
                    double const local = tracked ? 0.0 : pe;
                    double gx, gy, gz;
                    cone_direction(dx, dy, dz,
                                   t.sample_brems_cos(before, rng_uniform(seed, root, step, idx,
                                                                          pos ? kChPositronBremsPhotonAngle : kChBremsPhotonAngle)),
                                   rng_uniform(seed, root, step, idx, pos ? kChPositronBremsPhotonAzimuth : kChBremsPhotonAzimuth),
                                   gx, gy, gz);
                    photon_recoil(before, pe, gx, gy, gz, dx, dy, dz);
                    k = after;
                    if (pos) L.positron_brems(w, pe);
                    else L.brems_photon(w, pe);
                    if (tracked) {
                        em.secondary(2, kProcBrems, ex_, ey_, ez_, gx, gy, gz, pe, efrac);
                        fx.tallies[ORC_T_BREMS_PHOTONS] += 1.0;
                    } else if (local > 0.0) {
                        em.deposit_local(local, std::hypot(ex_, ey_), ez_);
                    }
                    fx.tallies[ORC_T_BREMS_EVENTS] += 1.0;
                    {
                        Closure c;
                        c.parent(pos ? 1 : 0, w, before, after, true);
                        if (tracked) c.particle(2, w, pe);
                        else c.local(w * local);
                        require_conserved(c, cutoff);
                    }
                } else if (pos && channel == 2) {
                    // in flight: the positron is consumed at the event point
                    spawn_annihilation(ctx, em, kProcAnnihilation, ex_, ey_, ez_, std::hypot(ex_, ey_), dx, dy, dz, k,
                                       false, idx, efrac);
                    segment(er_, ez_);
                    terminal_interaction = true;
                    final_f = ef;
                    return false;
                }
                return true;
            });
    };
    auto finish_substep = [&](uint64_t sc, double sb, double se) -> bool {
        if (k >= cutoff) {
            // apply_gs_elastic_scatter OV/RreaTransportStep.H:372-423 on the newborn channels
            double const gs_n = elastic_inverse_length(t, pos, k) * density_path(iv, path, sb, se);
            if (gs_n > 0.0) {
                auto U = [&](uint64_t chan, uint64_t inner) {
                    return rng_uniform(seed, root, step, pack(sc, 0U, inner), chan);
                };
                auto deflect = [&](double ct, double uphi) {
                    double nx, ny, nz;
                    cone_direction(dx, dy, dz, ct, uphi, nx, ny, nz);
                    dx = nx; dy = ny; dz = nz;
                };
                if (gs_n < t.gs_few) {
                    int const cap = static_cast<int>(std::ceil(3.0 * t.gs_few));
                    if (cap >= 255) fail(ORC_ERR_INVALID, "GS few-collision cap does not pack");
                    int const cnt = poisson_capped(
                        gs_n, U(pos ? kChNewbornPositronGsFewCount : kChNewbornElectronGsFewCount, 0), cap);
                    for (int q = 0; q < cnt; ++q) {
                        uint64_t const inner = static_cast<uint64_t>(q + 1);
                        double const ct = t.sample_single_elastic_cos(
                            pos, k, U(pos ? kChNewbornPositronGsFewScatter : kChNewbornElectronGsFewScatter, inner));
                        deflect(ct, U(pos ? kChNewbornPositronGsFewAzimuth : kChNewbornElectronGsFewAzimuth, inner));
                    }
                } else {
                    double const ct = t.sample_gs_cos(pos, k, gs_n,
                                                      U(pos ? kChNewbornPositronGsAngle : kChNewbornElectronGsAngle, 0));
                    deflect(ct, U(pos ? kChNewbornPositronGsAzimuth : kChNewbornElectronGsAzimuth, 0));
                }
            }
        }
        (void)half_kick(sfield, 0.5 * substep_dt, charge, k, dx, dy, dz);
        if (k < cutoff) {
            final_f = se;
            terminal_at_point = true;
            return true;
        }
        in_range();
        return false;
    };
    // RreaRunChargedSubstepInterleave OV/RreaChargedInterleave.H:693-732
    run_substep_interleave(
        substeps, k, cutoff, completed, exited, build_substep, run_events,
        [&] { return terminal_interaction || terminal_at_point; }, finish_substep,
        [&](double se) { return advance_running_position(path, se, run_x, run_y, run_z); },
        [&] { return path.exits; });
    if (pos && positron_total_loss > 0.0) L.positron_transport(w, positron_total_loss, false);  // :686-689
    if (terminal_interaction) {
        out.status = ORC_ANNIHILATED;
        fx.tallies[ORC_T_NEWBORN_TERMINATED] += 1.0;
        return;
    }
    double fx_ = 0.0, fy_ = 0.0, fz_ = 0.0;
    if (!newborn_final_point(path, final_f, fx_, fy_, fz_)) {
        fail(ORC_ERR_INVALID, "newborn could not place its final point in the half-open domain");
    }
    double const final_r = std::hypot(fx_, fy_);
    double const final_el = elapsed_at(final_f);
    bool const demote = pos ? k < cutoff : newborn_electron_demotes(ctx, fx_, fy_, fz_, k, hm_on);
    if (terminal_at_point || demote) {
        if (pos) {
            spawn_annihilation(ctx, em, kProcAnnihilation, fx_, fy_, fz_, final_r, dx, dy, dz, k, true,
                               pack(std::min<uint64_t>(substeps, kSegmentCap - 1U), 0U, 0U), final_el);
            L.positron_transport(w, k, true);  // OV/RreaNewbornAdvance.cpp:729
            out.status = ORC_ANNIHILATED;
        } else {
            fx.tallies[ORC_T_LOW_ENERGY_ELECTRONS] += w;
            em.low_electron(final_r, fz_, w, 0.0);
            em.deposit_local(k, final_r, fz_);
            out.status = ORC_DEMOTED;
        }
        segment(final_r, fz_);
        fx.tallies[ORC_T_NEWBORN_TERMINATED] += 1.0;
        return;
    }
    if (exited) {
        escape_along(path);
        return;
    }
    p.x = fx_; p.y = fy_; p.z = fz_;
    p.dir_x = dx; p.dir_y = dy; p.dir_z = dz;
    p.energy_eV = k;
    p.birth_fraction = 1.0;
    segment(final_r, fz_);
}

// TransportStep::AdvancePhotonNewborn OV/RreaNewbornAdvance.cpp:779-994.
inline void photon_newborn(StepContext const& ctx, ChunkLog& fx, orc_secondary& p, int64_t ip, NewbornOut& out)
{
    Tables const& t = *ctx.t;
    orc_config const& cfg = ctx.cfg;
    double const gcut = cfg.photon_cutoff_eV;
    double const dt = cfg.dt_s;
    uint64_t const step = static_cast<uint64_t>(std::max<int64_t>(cfg.step, 0));
    uint64_t const seed = cfg.rng_seed;
    uint64_t const root = p.root_key;
    uint64_t const rsp = static_cast<uint64_t>(p.root_species);
    double const w = p.weight;
    Emitter em{fx, ip, root, w, out.ion_pairs, p.root_species, p.generation + 1};
    double const residual = newborn_residual_fraction(p.birth_fraction);
    if (!(residual > 0.0)) {
        if (!(p.energy_eV >= gcut)) out.status = ORC_DEMOTED;
        return;
    }
    if (!in_half_open(p.x, p.y, p.z, ctx.r_lo, ctx.r_hi, ctx.z_lo, ctx.z_hi)) {
        fail(ORC_ERR_OUTSIDE_DOMAIN, "photon transient born outside the half-open domain");
    }
    double dx = p.dir_x, dy = p.dir_y, dz = p.dir_z;
    normalize(dx, dy, dz);
    double energy = p.energy_eV;
    if (energy < gcut) {
        em.deposit_local(energy, std::hypot(p.x, p.y), p.z);
        out.status = ORC_DEMOTED;
        fx.tallies[ORC_T_NEWBORN_TERMINATED] += 1.0;
        return;
    }
    auto escape = [&]() {
        out.status = ORC_ESCAPE;
        fx.tallies[ORC_T_NEWBORN_ESCAPES] += 1.0;
        Ledger{fx.tallies, 0.0}.photon_escape(w, energy);
    };
    auto survive_at = [&](double x, double y, double z) {
        p.x = x; p.y = y; p.z = z;
        p.dir_x = dx; p.dir_y = dy; p.dir_z = dz;
        p.energy_eV = energy;
        p.birth_fraction = 1.0;
    };
    double const chord = kC * dt * residual;
    double const end_x = p.x + chord * dx;
    double const end_y = p.y + chord * dy;
    double const end_z = p.z + chord * dz;
    Path path = clip_rz_path(std::hypot(p.x, p.y), std::atan2(p.y, p.x), p.z, std::hypot(end_x, end_y),
                             std::atan2(end_y, end_x), end_z, ctx.r_lo, ctx.r_hi, ctx.z_lo, ctx.z_hi);
    if (!(path.full_length > 0.0) || !(path.in_length > 0.0)) {
        if (path.exits) escape();
        return;
    }
    double elapsed_start = p.birth_fraction;
    uint64_t ord = 0;
    for (;;) {
        uint64_t idx = 0;
        if (!pack_transient_index(true, rsp, p.lineage, 0U, ord, 0U, idx)) {
            fail(ORC_ERR_INVALID, "newborn photon RNG index overflow");
        }
        std::vector<Interval> const iv = build_intervals(ctx.cp, path, ctx.density_ratio);
        double const mfp = photon_mfp(t, energy);
        double const tau = density_path(iv, path, path.begin_f, path.end_f) / mfp;
        double const taus[3] = {tau, 0.0, 0.0};
        Hazard const h = select_hazard(taus, 1, rng_uniform(seed, root, step, idx, kChNewbornCompetingOpticalDepth), 0.0);
        if (!h.event) {
            double x, y, z;
            if (!newborn_final_point(path, path.end_f, x, y, z)) {
                fail(ORC_ERR_INVALID, "newborn photon could not place its final point");
            }
            if (path.exits) { escape(); return; }
            survive_at(x, y, z);
            return;
        }
        double const f = fraction_at_density_path(iv, path, path.begin_f, path.end_f, h.sampled * mfp);
        double const dist = path.full_length * (f - path.begin_f);
        double const remaining = std::max(0.0, path.full_length - dist);
        double x, y, z;
        path.point_at_full(f, x, y, z);
        double const r_site = std::hypot(x, y);
        double const el = std::min(1.0, elapsed_start + dist / std::max(kC * dt, std::numeric_limits<double>::min()));
        int const channel = sample_photon_channel(t, energy, rng_uniform(seed, root, step, idx, kChNewbornCompetingChannel));
        out.substeps = ord + 1U;
        if (!photon_final_state(ctx, em, channel, idx, x, y, z, el, energy, dx, dy, dz)) {
            out.status = ORC_ABSORBED;
            fx.tallies[ORC_T_NEWBORN_TERMINATED] += 1.0;
            return;
        }
        if (!(remaining > 0.0)) { survive_at(x, y, z); return; }
        double const tx = x + remaining * dx, ty = y + remaining * dy, tz = z + remaining * dz;
        path = clip_rz_path(r_site, std::atan2(y, x), z, std::hypot(tx, ty), std::atan2(ty, tx), tz, ctx.r_lo, ctx.r_hi,
                            ctx.z_lo, ctx.z_hi);
        elapsed_start = el;
        ++ord;
    }
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C entry points (ctypes)
// ---------------------------------------------------------------------------
using namespace orc;

extern "C" {

uint64_t orc_rng_counter(uint64_t seed, uint64_t pid, uint64_t step, uint64_t idx, uint64_t chan)
{
    return rng_counter(seed, pid, step, idx, chan);
}

double orc_rng_uniform(uint64_t seed, uint64_t pid, uint64_t step, uint64_t idx, uint64_t chan)
{
    return rng_uniform(seed, pid, step, idx, chan);
}

void orc_uniform_stream(double u1, double u2, int64_t n, double* out)
{
    UniformStream s(u1, u2);
    for (int64_t i = 0; i < n; ++i) out[i] = s.next();
}

void* orc_tables_create() { return new Tables(); }
void orc_tables_destroy(void* h) { delete static_cast<Tables*>(h); }

// Rate table slot `id` (RateId order), x/y arrays of length n, interpolation mode.
void orc_tables_set_rate(void* h, int id, const double* x, const double* y, int64_t n, int mode)
{
    auto* t = static_cast<Tables*>(h);
    RateTable& r = t->rate[id];
    r.x.assign(x, x + n);
    r.y.assign(y, y + n);
    r.mode = mode;
    r.finalize();
}

// Conditional CDF `which`: 0 moller, 1 bhabha, 2 e- brems, 3 e+ brems.
void orc_tables_set_cdf(void* h, int which, const double* primary, int64_t ngroups, const int64_t* start,
                        const double* u, const double* v)
{
    auto* t = static_cast<Tables*>(h);
    CdfGroups* g = which == 0 ? &t->moller : which == 1 ? &t->bhabha : which == 2 ? &t->ebrems : &t->pbrems;
    g->primary.assign(primary, primary + ngroups);
    g->u.clear();
    g->v.clear();
    for (int64_t i = 0; i < ngroups; ++i) {
        g->u.emplace_back(u + start[i], u + start[i + 1]);
        g->v.emplace_back(v + start[i], v + start[i + 1]);
    }
}

// Angle table `which`: 0 GS e-, 1 GS e+, 2 Wentzel e-, 3 Wentzel e+.
void orc_tables_set_angle(void* h, int which, const double* e, int64_t ne, const double* n, int64_t nn,
                          const double* u, int64_t nu, const double* omc)
{
    auto* t = static_cast<Tables*>(h);
    AngleTable* a = which == 0 ? &t->gs_e : which == 1 ? &t->gs_p : which == 2 ? &t->wz_e : &t->wz_p;
    a->e.assign(e, e + ne);
    a->n.assign(n, n + nn);
    a->u.assign(u, u + nu);
    a->omc.assign(omc, omc + ne * nn * nu);
}

void orc_tables_set_elastic(void* h, int positron, const double* e, const double* mfp, const double* eta,
                            const double* trmfp, int64_t n)
{
    auto* t = static_cast<Tables*>(h);
    ElasticTable& el = positron ? t->el_p : t->el_e;
    el.e.assign(e, e + n);
    el.mfp.assign(mfp, mfp + n);
    el.eta.assign(eta, eta + n);
    el.trmfp.assign(trmfp, trmfp + n);
}

// scalars: hard-Moller threshold, brems cutoff, reference density, charged
// min/max, photon min/max, GS few threshold, isotropic n, transition energy.
int orc_tables_finalize(void* h, const double* s)
{
    auto* t = static_cast<Tables*>(h);
    t->hard_moller_threshold = s[0];
    t->brems_cutoff = s[1];
    t->reference_density = s[2];
    t->charged_min = s[3];
    t->charged_max = s[4];
    t->photon_min = s[5];
    t->photon_max = s[6];
    t->gs_few = s[7];
    t->gs_iso = s[8];
    t->gs_transition = s[9];
    try {
        t->build_hm_mean();
    } catch (OracleError const& e) {
        return e.code;
    }
    return 0;
}

const double* orc_tables_hm_mean(void* h) { return static_cast<Tables*>(h)->hm_mean_secondary.data(); }

#define ORC_GUARD(expr) \
    try { expr; } catch (OracleError const& e) { if (err) *err = e.code; return 0.0; }

double orc_interpolate(void* h, int id, double e, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->rate[id].interpolate(e));
}

double orc_sample_moller(void* h, double e, double u, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->sample_moller(e, u));
}

double orc_sample_brems_energy(void* h, double e, double u, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->sample_brems_energy(e, u));
}

double orc_sample_brems_cos(void* h, double e, double u, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->sample_brems_cos(e, u));
}

double orc_sample_gs_cos(void* h, int positron, double e, double n, double u, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->sample_gs_cos(positron != 0, e, n, u));
}

double orc_sample_single_elastic_cos(void* h, int positron, double e, double u, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->sample_single_elastic_cos(positron != 0, e, u));
}

double orc_elastic_mfp(void* h, int positron, double e, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->elastic_mfp(positron != 0, e, 1.0));
}

double orc_ensemble_drag(void* h, double e, double ratio, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return ensemble_drag(*static_cast<Tables*>(h), e, ratio, true));
}

double orc_hard_moller_energy_loss(void* h, double e, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->hard_moller_energy_loss(e, 1.0));
}

// Chord clip: out = [sx,sy,sz,ex,ey,ez,full,begin,end,in_len,exits,npieces]
int orc_clip_path(double sr, double st, double sz, double er_, double et, double ez_, double r_lo, double r_hi,
                  double z_lo, double z_hi, double* out)
{
    try {
        Path const p = clip_rz_path(sr, st, sz, er_, et, ez_, r_lo, r_hi, z_lo, z_hi);
        double const v[12] = {p.sx, p.sy, p.sz, p.ex, p.ey, p.ez, p.full_length, p.begin_f, p.end_f,
                              p.in_length, p.exits ? 1.0 : 0.0, static_cast<double>(p.pieces.size())};
        std::memcpy(out, v, sizeof(v));
        return 0;
    } catch (OracleError const& e) {
        return e.code;
    }
}

void orc_gather(const orc_geometry* g, double r, double z, double* out3)
{
    Coupling cp{*g, 1.0, 0.0, 0.0};
    cp.gather(r, z, out3[0], out3[1], out3[2]);
}

// Index-space deposit of one segment on a grid of nr x nz cells.
void orc_deposit_segment(int nr, int nz, double dr, double dz, double r_lo, double z_lo, double* rho, double* jr,
                         double* jz, double q, double dt, double r0, double z0, double r1, double z1)
{
    Grid g{nr, nz, dr, dz, r_lo, z_lo, rho, jr, jz};
    deposit_segment(g, q, dt, r0, z0, r1, z1);
}

int orc_deposit_escaping_segment(int nr, int nz, double dr, double dz, double r_lo, double z_lo, double* rho,
                                 double* jr, double* jz, double q, double dt, const double* a7, int pnr, int pzb,
                                 int pnz, uint32_t faces)
{
    Grid g{nr, nz, dr, dz, r_lo, z_lo, rho, jr, jz};
    return deposit_escaping_segment(g, q, dt, a7[0], a7[1], a7[2], a7[3], a7[4], a7[5], pnr, pzb, pnz, faces) ? 1 : 0;
}

// One charged transport step over n particles of one species.  prev_* is the
// pre-push snapshot the reference reads; out_* receives the written-back
// state.  The padded net-chord grid is (nr+2) x (nz+4) with two low-z pad
// cells (OV/RreaWarpXCoupling.cpp:417-444); jr/jz/rho are accumulated (+=).
static int charged_step(bool positron, void* h, const orc_geometry* g, const orc_config* cfg, int64_t n,
                        const double* pr, const double* pz, const double* pth, const double* pux, const double* puy,
                        const double* puz, const double* w, const int64_t* id, const int32_t* cpu, double* out_r,
                        double* out_z, double* out_th, double* out_ux, double* out_uy, double* out_uz,
                        int32_t* status, double* ion_pairs, int64_t* substeps, int64_t sec_cap, orc_secondary* sec,
                        int64_t* n_sec, double* jr, double* jz, double* rho, double* tallies, int nthreads)
{
    auto const* t = static_cast<Tables const*>(h);
    if (g->r_lo != 0.0) return ORC_ERR_INVALID;
    StepContext ctx;
    ctx.t = t;
    ctx.cp = Coupling{*g, t->reference_density, 0.0, 0.0};
    ctx.cfg = *cfg;
    ctx.r_lo = g->r_lo;
    ctx.r_hi = g->r_lo + g->nr * g->dr;
    ctx.z_lo = g->z_lo;
    ctx.z_hi = g->z_lo + g->nz * g->dz;
    ctx.cp.r_hi = ctx.r_hi;
    ctx.cp.z_hi = ctx.z_hi;
    // RreaParticleInteraction::Apply :29-35: ratio at the domain's mid z, or 1.
    try {
        ctx.density_ratio = positive_or(ctx.cp.density_ratio_at_z(0.5 * (ctx.z_lo + ctx.z_hi)), 1.0);
    } catch (OracleError const& e) {
        *n_sec = 0;
        return e.code;
    }
    if (!(cfg->dt_s > 0.0)) { *n_sec = 0; return 0; }

    constexpr int64_t chunk = 64;
    int64_t const nchunks = (n + chunk - 1) / chunk;
    std::vector<ChunkLog> logs(static_cast<size_t>(std::max<int64_t>(nchunks, 0)));
    std::atomic<int> error{0};
#ifdef _OPENMP
    int const threads = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (threads > 1 && nchunks > 1)
#endif
    for (int64_t c = 0; c < nchunks; ++c) {
        ChunkLog& fx = logs[static_cast<size_t>(c)];
        int64_t const b = c * chunk, e = std::min(n, b + chunk);
        for (int64_t i = b; i < e; ++i) {
            if (error.load(std::memory_order_relaxed) != 0) break;
            ChargedOut o{};
            if (id[i] == std::numeric_limits<int64_t>::min()) {
                // an invalidated particle: the reference loop `continue`s
                // (RreaElectronTransport.cpp:60-63, RreaPositronTransport.cpp:59-62)
                out_r[i] = pr[i]; out_z[i] = pz[i]; out_th[i] = pth[i];
                out_ux[i] = pux[i]; out_uy[i] = puy[i]; out_uz[i] = puz[i];
                status[i] = 5;
                ion_pairs[i] = 0.0;
                substeps[i] = 0;
                continue;
            }
            try {
                charged_particle(ctx, fx, positron, i, pr[i], pz[i], pth[i], pux[i], puy[i], puz[i], w[i], id[i], cpu[i], o);
            } catch (OracleError const& err) {
                int expected = 0;
                error.compare_exchange_strong(expected, err.code);
                break;
            }
            if (o.status == ORC_ZERO_WEIGHT) {
                o.r = pr[i]; o.z = pz[i]; o.th = pth[i]; o.ux = pux[i]; o.uy = puy[i]; o.uz = puz[i];
            }
            out_r[i] = o.r; out_z[i] = o.z; out_th[i] = o.th;
            out_ux[i] = o.ux; out_uy[i] = o.uy; out_uz[i] = o.uz;
            status[i] = o.status;
            ion_pairs[i] = o.ion_pairs;
            substeps[i] = static_cast<int64_t>(o.substeps);
        }
    }
    (void)nthreads;
    if (error.load() != 0) { *n_sec = 0; return error.load(); }

    // Deterministic merge in chunk order, then replay of the deposits
    // (OV/RreaElectronTransport.cpp:892-897, OV/RreaTransportLog.H:603-653).
    int const pnr = g->nr, pnz = g->nz, pad = 2;
    Grid grid{pnr + pad, pnz + 2 * pad, g->dr, g->dz, g->r_lo, g->z_lo - pad * g->dz, rho, jr, jz};
    int64_t ns = 0;
    for (auto const& fx : logs) {
        for (auto const& s : fx.secondaries) {
            if (ns >= sec_cap) { *n_sec = ns; return ORC_ERR_SECONDARY_CAPACITY; }
            sec[ns++] = s;
        }
        fold_tallies(tallies, fx.tallies);
        for (auto const& op : fx.deposits) {
            if (op.kind == 0) {
                if (op.a[4] != 0.0) deposit_segment(grid, op.a[4], cfg->dt_s, op.a[0], op.a[1], op.a[2], op.a[3]);
            } else {
                if (op.a[6] != 0.0
                    && !deposit_escaping_segment(grid, op.a[6], cfg->dt_s, op.a[0], op.a[1], op.a[2], op.a[3],
                                                 op.a[4], op.a[5], pnr, pad, pnz, op.faces)) {
                    *n_sec = ns;
                    return ORC_ERR_INVALID;
                }
            }
        }
        int const se = replay_sources(g, cfg->low_energy_cutoff_eV, fx.sources);
        if (se) { *n_sec = ns; return se; }
    }
    *n_sec = ns;
    return 0;
}

int orc_electron_step(void* h, const orc_geometry* g, const orc_config* cfg, int64_t n, const double* pr,
                      const double* pz, const double* pth, const double* pux, const double* puy, const double* puz,
                      const double* w, const int64_t* id, const int32_t* cpu, double* out_r, double* out_z,
                      double* out_th, double* out_ux, double* out_uy, double* out_uz, int32_t* status,
                      double* ion_pairs, int64_t* substeps, int64_t sec_cap, orc_secondary* sec, int64_t* n_sec,
                      double* jr, double* jz, double* rho, double* tallies, int nthreads)
{
    return charged_step(false, h, g, cfg, n, pr, pz, pth, pux, puy, puz, w, id, cpu, out_r, out_z, out_th, out_ux,
                        out_uy, out_uz, status, ion_pairs, substeps, sec_cap, sec, n_sec, jr, jz, rho, tallies,
                        nthreads);
}

int orc_positron_step(void* h, const orc_geometry* g, const orc_config* cfg, int64_t n, const double* pr,
                      const double* pz, const double* pth, const double* pux, const double* puy, const double* puz,
                      const double* w, const int64_t* id, const int32_t* cpu, double* out_r, double* out_z,
                      double* out_th, double* out_ux, double* out_uy, double* out_uz, int32_t* status,
                      double* ion_pairs, int64_t* substeps, int64_t sec_cap, orc_secondary* sec, int64_t* n_sec,
                      double* jr, double* jz, double* rho, double* tallies, int nthreads)
{
    return charged_step(true, h, g, cfg, n, pr, pz, pth, pux, puy, puz, w, id, cpu, out_r, out_z, out_th, out_ux,
                        out_uy, out_uz, status, ion_pairs, substeps, sec_cap, sec, n_sec, jr, jz, rho, tallies,
                        nthreads);
}

// One photon transport step over n photons.  prev_* is the captured
// pre-push point, cur_* WarpX's pushed point (the chord end), u = E * dir.
// out_* receives the written-back state (cur_* and u where nothing is written).
int orc_photon_step(void* h, const orc_geometry* g, const orc_config* cfg, int64_t n, const double* pr,
                    const double* pz, const double* pth, const double* cr, const double* cz, const double* cth,
                    const double* ux, const double* uy, const double* uz, const double* w, const int64_t* id,
                    const int32_t* cpu, double* out_r, double* out_z, double* out_th, double* out_ux,
                    double* out_uy, double* out_uz, int32_t* status, double* ion_pairs, int64_t* events,
                    int64_t sec_cap, orc_secondary* sec, int64_t* n_sec, double* tallies, int nthreads)
{
    auto const* t = static_cast<Tables const*>(h);
    if (g->r_lo != 0.0) return ORC_ERR_INVALID;
    StepContext ctx;
    ctx.t = t;
    ctx.cp = Coupling{*g, t->reference_density, 0.0, 0.0};
    ctx.cfg = *cfg;
    ctx.r_lo = g->r_lo;
    ctx.r_hi = g->r_lo + g->nr * g->dr;
    ctx.z_lo = g->z_lo;
    ctx.z_hi = g->z_lo + g->nz * g->dz;
    ctx.cp.r_hi = ctx.r_hi;
    ctx.cp.z_hi = ctx.z_hi;
    try {
        ctx.density_ratio = positive_or(ctx.cp.density_ratio_at_z(0.5 * (ctx.z_lo + ctx.z_hi)), 1.0);
    } catch (OracleError const& e) {
        *n_sec = 0;
        return e.code;
    }
    if (!(cfg->dt_s > 0.0)) { *n_sec = 0; return 0; }
    constexpr int64_t chunk = 64;
    int64_t const nchunks = (n + chunk - 1) / chunk;
    std::vector<ChunkLog> logs(static_cast<size_t>(std::max<int64_t>(nchunks, 0)));
    std::atomic<int> error{0};
#ifdef _OPENMP
    int const threads = nthreads > 0 ? nthreads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads) if (threads > 1 && nchunks > 1)
#endif
    for (int64_t c = 0; c < nchunks; ++c) {
        ChunkLog& fx = logs[static_cast<size_t>(c)];
        int64_t const b = c * chunk, e = std::min(n, b + chunk);
        for (int64_t i = b; i < e; ++i) {
            if (error.load(std::memory_order_relaxed) != 0) break;
            ChargedOut o{};
            if (id[i] == std::numeric_limits<int64_t>::min()) {
                // an invalidated photon: the reference loop `continue`s (RreaPhotonTransport.cpp:52-55)
                out_r[i] = cr[i]; out_z[i] = cz[i]; out_th[i] = cth[i];
                out_ux[i] = ux[i]; out_uy[i] = uy[i]; out_uz[i] = uz[i];
                status[i] = ORC_SKIPPED_INVALID;
                ion_pairs[i] = 0.0;
                events[i] = 0;
                continue;
            }
            try {
                photon_particle(ctx, fx, i, pr[i], pth[i], pz[i], cr[i], cth[i], cz[i], ux[i], uy[i], uz[i], w[i], id[i],
                                cpu[i], o);
            } catch (OracleError const& err) {
                int expected = 0;
                error.compare_exchange_strong(expected, err.code);
                break;
            }
            out_r[i] = o.r; out_z[i] = o.z; out_th[i] = o.th;
            out_ux[i] = o.ux; out_uy[i] = o.uy; out_uz[i] = o.uz;
            status[i] = o.status;
            ion_pairs[i] = o.ion_pairs;
            events[i] = static_cast<int64_t>(o.substeps);
        }
    }
    (void)nthreads;
    if (error.load() != 0) { *n_sec = 0; return error.load(); }
    int64_t ns = 0;
    for (auto const& fx : logs) {
        for (auto const& s : fx.secondaries) {
            if (ns >= sec_cap) { *n_sec = ns; return ORC_ERR_SECONDARY_CAPACITY; }
            sec[ns++] = s;
        }
        fold_tallies(tallies, fx.tallies);
        int const se = replay_sources(g, cfg->low_energy_cutoff_eV, fx.sources);
        if (se) { *n_sec = ns; return se; }
    }
    *n_sec = ns;
    return 0;
}

void orc_tables_set_photoelectric(void* h, const double* primary, int64_t ngroups, const int64_t* start,
                                  const double* u, const double* binding)
{
    auto* t = static_cast<Tables*>(h);
    t->pe_primary.assign(primary, primary + ngroups);
    t->pe_start.assign(start, start + ngroups + 1);
    t->pe_u.assign(u, u + start[ngroups]);
    t->pe_binding.assign(binding, binding + start[ngroups]);
}

void orc_tables_set_pair(void* h, const int64_t* channel, const double* energy, int64_t nblocks, const int64_t* start,
                         const double* frac, const double* dir)
{
    auto* t = static_cast<Tables*>(h);
    t->pair_channel.assign(channel, channel + nblocks);
    t->pair_energy.assign(energy, energy + nblocks);
    t->pair_start.assign(start, start + nblocks + 1);
    t->pair_frac.assign(frac, frac + 3 * start[nblocks]);
    t->pair_dir.assign(dir, dir + 9 * start[nblocks]);
}

// Standalone photon samplers: out = [scattered, recoil, cos] / [binding, electron] / cos.
int orc_sample_compton(void* h, double e, double u1, double u2, double* out3)
{
    try {
        ComptonState const s = sample_compton(*static_cast<Tables*>(h), e, u1, u2);
        out3[0] = s.scattered; out3[1] = s.recoil; out3[2] = s.cos_theta;
        return 0;
    } catch (OracleError const& err) {
        return err.code;
    }
}

int orc_sample_photoelectric(void* h, double e, double u, double* out2)
{
    try {
        PhotoelectricState const s = sample_photoelectric(*static_cast<Tables*>(h), e, u);
        out2[0] = s.binding; out2[1] = s.electron;
        return 0;
    } catch (OracleError const& err) {
        return err.code;
    }
}

double orc_sample_photoelectron_cos(void* h, double ke, double u, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return sample_photoelectron_cos(*static_cast<Tables*>(h), ke, u));
}

int orc_sample_photon_channel(void* h, double e, double u)
{
    try {
        return sample_photon_channel(*static_cast<Tables*>(h), e, u);
    } catch (OracleError const& err) {
        return -err.code;
    }
}

// out = [valid, electron, positron, recoil, local, e dir(3), p dir(3), r dir(3)]
int orc_sample_pair(void* h, double e, int channel, double u_inv, double u_block, double u_az, double* out14)
{
    try {
        PairState const s = sample_pair(*static_cast<Tables*>(h), e, channel, u_inv, u_block, u_az);
        double const v[14] = {s.valid ? 1.0 : 0.0, s.electron, s.positron, s.recoil, s.local, s.edx, s.edy, s.edz,
                              s.pdx, s.pdy, s.pdz, s.rdx, s.rdy, s.rdz};
        std::memcpy(out14, v, sizeof(v));
        return 0;
    } catch (OracleError const& err) {
        return err.code;
    }
}

// SamplePositronAnnihilationFinalState as a standalone sampler:
// out = [E1, E2, cos1, azimuth_u, valid, at_rest].
void orc_sample_annihilation(double k, double draw_energy, double draw_azimuth, double* out6)
{
    AnnihilationState const s = sample_annihilation(k, draw_energy, draw_azimuth);
    out6[0] = s.e1; out6[1] = s.e2; out6[2] = s.cos1; out6[3] = s.az1;
    out6[4] = s.valid ? 1.0 : 0.0; out6[5] = s.at_rest ? 1.0 : 0.0;
}

double orc_sample_bhabha(void* h, double e, double u, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->sample_bhabha(e, u));
}

double orc_sample_positron_brems_energy(void* h, double e, double u, int* err)
{
    if (err) *err = 0;
    ORC_GUARD(return static_cast<Tables*>(h)->sample_positron_brems_energy(e, u));
}

}  // extern "C"

// The transient FIFO of DrainAndCommitSecondaries
// (OV/RreaSecondaryCommit.cpp:338-468): the step's staged secondaries in
// Apply's order (electron, photon, then positron loop, each merged in chunk
// order) seed a FIFO; every entry born before the end of the step is flown
// through its residual time, and the children it makes join the back of the
// FIFO with the root's next free transient lineage.  `work` receives the
// whole FIFO (initial entries first), `status` and `ion_pairs` one value per
// FIFO entry, and `survivors` the entries that reach the end of the step, in
// FIFO order, at their transported state with birth fraction 1.  Net chords
// deposit into jr/jz/rho; tallies accumulate.
extern "C" int orc_newborn_drain(void* h, const orc_geometry* g, const orc_config* cfg, const orc_secondary* initial,
                                 int64_t n_initial, int64_t cap, orc_secondary* work, int64_t* n_work,
                                 int32_t* status, double* ion_pairs, int64_t* substeps, orc_secondary* survivors,
                                 int64_t* n_surv,
                                 double* jr, double* jz, double* rho, double* tallies)
{
    *n_work = 0;
    *n_surv = 0;
    auto const* t = static_cast<Tables const*>(h);
    if (g->r_lo != 0.0) return ORC_ERR_INVALID;
    StepContext ctx;
    ctx.t = t;
    ctx.cp = Coupling{*g, t->reference_density, 0.0, 0.0};
    ctx.cfg = *cfg;
    ctx.r_lo = g->r_lo;
    ctx.r_hi = g->r_lo + g->nr * g->dr;
    ctx.z_lo = g->z_lo;
    ctx.z_hi = g->z_lo + g->nz * g->dz;
    ctx.cp.r_hi = ctx.r_hi;
    ctx.cp.z_hi = ctx.z_hi;
    int const pnr = g->nr, pnz = g->nz, pad = 2;
    Grid grid{pnr + pad, pnz + 2 * pad, g->dr, g->dz, g->r_lo, g->z_lo - pad * g->dz, rho, jr, jz};
    try {
        ctx.density_ratio = positive_or(ctx.cp.density_ratio_at_z(0.5 * (ctx.z_lo + ctx.z_hi)), 1.0);
        if (!(cfg->dt_s > 0.0)) return 0;
        std::vector<orc_secondary> fifo(initial, initial + std::max<int64_t>(n_initial, 0));
        // the FIFO's birth-payload validation, OV/RreaSecondaryCommit.cpp:338-376
        for (auto const& p : fifo) {
            bool const finite = std::isfinite(p.x) && std::isfinite(p.y) && std::isfinite(p.z) && std::isfinite(p.dir_x)
                && std::isfinite(p.dir_y) && std::isfinite(p.dir_z) && std::isfinite(p.weight)
                && std::isfinite(p.energy_eV) && std::isfinite(p.birth_fraction);
            bool const energy_ok = p.species == 2 ? p.energy_eV > 0.0 : p.energy_eV >= 0.0;
            if (p.species < 0 || p.species > 2 || p.root_species < 0 || p.root_species > 2 || !finite
                || !(p.weight > 0.0) || !energy_ok || p.birth_fraction < 0.0 || p.birth_fraction > 1.0
                || p.lineage >= kLineageCap) {
                fail(ORC_ERR_INVALID, "transient FIFO received an invalid birth payload");
            }
        }
        // next free lineage per (root key, root species), :390-406
        std::map<std::pair<uint64_t, int32_t>, uint64_t> next_lineage;
        for (auto const& p : fifo) {
            auto& nx = next_lineage[{p.root_key, p.root_species}];
            nx = std::max(nx, p.lineage + 1U);
        }
        std::vector<int32_t> st;
        std::vector<double> ion;
        std::vector<int64_t> sub;
        std::vector<orc_secondary> surv;
        for (size_t i = 0; i < fifo.size(); ++i) {
            orc_secondary p = fifo[i];
            // transient_energy_is_certified :148-180
            bool const photon = p.species == 2;
            double const lo = photon ? t->photon_min : t->charged_min;
            double const hi = photon ? t->photon_max : t->charged_max;
            if (!std::isfinite(p.energy_eV) || p.energy_eV < lo || p.energy_eV > hi) {
                fail(ORC_ERR_TABLE_RANGE, "transient FIFO particle outside its certified energy range");
            }
            ChunkLog fx;
            NewbornOut o;
            if (p.birth_fraction < 1.0) {
                fx.tallies[ORC_T_NEWBORN_ADVANCED] += 1.0;
                if (photon) photon_newborn(ctx, fx, p, static_cast<int64_t>(i), o);
                else charged_newborn(ctx, fx, p, static_cast<int64_t>(i), o);
            }
            for (auto c : fx.secondaries) {
                auto& nx = next_lineage[{c.root_key, c.root_species}];
                if (nx >= kLineageCap) fail(ORC_ERR_INVALID, "transient FIFO lineage overflow");
                c.lineage = nx++;
                fifo.push_back(c);
            }
            fold_tallies(tallies, fx.tallies);
            for (auto const& op : fx.deposits) {
                if (op.kind == 0) {
                    if (op.a[4] != 0.0) deposit_segment(grid, op.a[4], cfg->dt_s, op.a[0], op.a[1], op.a[2], op.a[3]);
                } else if (op.a[6] != 0.0
                           && !deposit_escaping_segment(grid, op.a[6], cfg->dt_s, op.a[0], op.a[1], op.a[2], op.a[3],
                                                        op.a[4], op.a[5], pnr, pad, pnz, op.faces)) {
                    fail(ORC_ERR_INVALID, "newborn escape deposit rejected");
                }
            }
            if (int const se = replay_sources(g, cfg->low_energy_cutoff_eV, fx.sources)) {
                fail(se, "newborn material source outside the half-open domain");
            }
            st.push_back(o.status);
            ion.push_back(o.ion_pairs);
            sub.push_back(static_cast<int64_t>(o.substeps));
            if (o.status == ORC_SURVIVE) {
                p.birth_fraction = 1.0;
                surv.push_back(p);
                if (fifo[i].birth_fraction < 1.0) tallies[ORC_T_NEWBORN_SURVIVORS] += 1.0;
            }
        }
        if (static_cast<int64_t>(fifo.size()) > cap) return ORC_ERR_SECONDARY_CAPACITY;
        for (size_t i = 0; i < fifo.size(); ++i) {
            work[i] = fifo[i];
            status[i] = st[i];
            ion_pairs[i] = ion[i];
            substeps[i] = sub[i];
        }
        for (size_t i = 0; i < surv.size(); ++i) survivors[i] = surv[i];
        *n_work = static_cast<int64_t>(fifo.size());
        *n_surv = static_cast<int64_t>(surv.size());
    } catch (OracleError const& e) {
        return e.code;
    }
    return 0;
}

// RreaTryPackTransientRngIndex / RreaTryLiveHazardRngIndex
// (OV/RreaParticleInteraction.H:660-710): 1 and the packed value, or 0.
extern "C" int orc_pack_transient_index(int newborn, uint64_t species, uint64_t lineage, uint64_t segment,
                                        uint64_t event, uint64_t inner, uint64_t* out)
{
    return pack_transient_index(newborn != 0, species, lineage, segment, event, inner, *out) ? 1 : 0;
}

extern "C" int orc_live_hazard_index(uint64_t species, uint64_t segment, uint64_t event, uint64_t* out)
{
    return live_hazard_index(species, segment, event, *out) ? 1 : 0;
}

// RreaEvaluateElectronStatePolicy OV/RreaParticleInteraction.H:1296-1317 as
// the newborn evaluates it: 1 when the state demotes.
extern "C" int orc_electron_state_demotes(double energy, double cutoff, double threshold, double field, double drag)
{
    bool const below = energy < cutoff;
    bool const nonrunaway = threshold > 0.0 && energy <= threshold && field <= drag;
    return below || nonrunaway ? 1 : 0;
}

// RreaNewbornElapsedStepFraction OV/RreaParticleInteraction.H:181-197
extern "C" double orc_newborn_elapsed(double birth, uint64_t completed, uint64_t substeps, double within)
{
    return newborn_elapsed(birth, newborn_residual_fraction(birth), completed, substeps, within);
}

extern "C" int orc_accumulate_source(const orc_geometry* g, double cutoff, int kind, double r, double z, double weight,
                                     double secondary, double loss)
{
    if (!g || !g->sources || kind < 0 || kind > 2) return ORC_ERR_INVALID;
    return accumulate_source(g, cutoff, SourceOp{kind, r, z, weight, secondary, loss});
}

// NormalizeAndSync :280-345: every field over its RZ shell cell volume
// (rz_shell_volume, NA/include/rrea/RreaConstants.H:37-43), the three
// source rates also over dt
extern "C" int orc_sources_normalize(const orc_geometry* g, double dt, int diagnostics)
{
    if (!g || !g->sources || !(dt > 0.0)) return ORC_ERR_INVALID;
    int64_t const cells = static_cast<int64_t>(g->nr) * g->nz;
    double const pi = 3.141592653589793238462643383279502884;
    for (int j = 0; j < g->nz; ++j) {
        for (int i = 0; i < g->nr; ++i) {
            double const r_inner = g->r_lo + static_cast<double>(i - 0) * g->dr;
            double const inner = r_inner > 0.0 ? r_inner : 0.0;
            double const outer = r_inner + g->dr > 0.0 ? r_inner + g->dr : 0.0;
            double const volume = pi * (outer * outer - inner * inner) * g->dz;
            for (int f = 0; f < 5; ++f) {
                bool const rate = f <= 2;
                if (!rate && !diagnostics) continue;
                double const denom = rate ? volume * dt : volume;
                double& v = g->sources[f * cells + i + static_cast<int64_t>(g->nr) * j];
                v = (denom > 0.0) ? v / denom : 0.0;
            }
        }
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Probes of the step's primitives: the functions above, called on explicit
// inputs, so the reference's behavioural cases (its charged-interleave,
// transport-path and sampler smokes) and closed forms can be checked against
// the code the oracle's steps run.
// ---------------------------------------------------------------------------
extern "C" {

double orc_unit_cycle(double u) { return unit_cycle(u); }

void orc_normalize(double* v) { normalize(v[0], v[1], v[2]); }

void orc_boris_rotate(double* u, double qmdt, const double* b) { boris_rotate(u[0], u[1], u[2], qmdt, b[0], b[1], b[2]); }

void orc_cone_direction(const double* axis, double ct, double uphi, double* out)
{
    cone_direction(axis[0], axis[1], axis[2], ct, uphi, out[0], out[1], out[2]);
}

double orc_moller_cos(double kp, double ks, int secondary)
{
    return secondary ? moller_secondary_cos(kp, ks) : moller_primary_cos(kp, ks);
}

void orc_moller_lab_directions(const double* d, double kp, double ks, double az, double* parent, double* secondary)
{
    moller_lab_directions(d[0], d[1], d[2], kp, ks, az, parent[0], parent[1], parent[2], secondary[0], secondary[1],
                          secondary[2]);
}

void orc_photon_recoil(double k0, double eg, const double* g, double* c)
{
    photon_recoil(k0, eg, g[0], g[1], g[2], c[0], c[1], c[2]);
}

void orc_rotate_to_axis(const double* axis, const double* local, double* out)
{
    rotate_to_axis(axis[0], axis[1], axis[2], local[0], local[1], local[2], out[0], out[1], out[2]);
}

int orc_poisson_capped(double mean, double u, int kmax) { return poisson_capped(mean, u, kmax); }

// out = [valid, event, channel, sampled, total]
void orc_select_hazard(const double* taus, int nch, double u1, double u2, double* out)
{
    Hazard const h = select_hazard(taus, nch, u1, u2);
    out[0] = h.valid ? 1.0 : 0.0;
    out[1] = h.event ? 1.0 : 0.0;
    out[2] = h.channel;
    out[3] = h.sampled;
    out[4] = h.total;
}

// field = [work, ex, ey, ez, bx, by, bz]; returns the half work
double orc_half_kick(const double* field, double half_dt, double q, double* k, double* d)
{
    FieldRange f;
    f.work = field[0]; f.ex = field[1]; f.ey = field[2]; f.ez = field[3];
    f.bx = field[4]; f.by = field[5]; f.bz = field[6];
    return half_kick(f, half_dt, q, *k, d[0], d[1], d[2]);
}

// The gate of a channel's inverse length; *calls counts the mean free path's
// evaluations.
double orc_gated_inverse_length(int open, double mfp, int* calls)
{
    return gated_inverse_length(open != 0, [&] { ++*calls; return mfp; });
}

int orc_stop_is_escape(int exits, double f, double end_f) { return stop_is_escape(exits != 0, f, end_f) ? 1 : 0; }

// The cutoff limiter on a linear loss loss_to(f) = rate (f - b): 1 when the
// cutoff is reached, 0 when not, -code on a fail-closed input.
int orc_limit_at_cutoff_linear(double b, double e, double k, double cutoff, double rate, double* end)
{
    try {
        SegLimit const l = limit_at_cutoff(b, e, k, cutoff, [&](double f) { return rate * (f - b); });
        *end = l.end;
        return l.cutoff ? 1 : 0;
    } catch (OracleError const& err) {
        return -err.code;
    }
}

// The ion-pair partition over n intervals [begin, end) of density ratio
// `ratio`; writes the emitted (midpoint fraction, pair weight) and returns
// their count.
int orc_partition_ion_pairs(const double* ib, const double* ie, const double* ir, int n, double full_length, double b,
                            double e, double segment_loss, double raw_loss, double w, double w_air, double stp,
                            double* out_f, double* out_w, int cap)
{
    std::vector<Interval> iv;
    for (int i = 0; i < n; ++i) iv.push_back(Interval{ib[i], ie[i], full_length * (ie[i] - ib[i]), ir[i]});
    int m = 0;
    partition_ion_pairs(iv, full_length, b, e, segment_loss, raw_loss, w, w_air, stp, [&](double f, double pw) {
        if (m < cap) { out_f[m] = f; out_w[m] = pw; }
        ++m;
    });
    return m;
}

// The event interleave on the analytic medium of the reference's recorder:
// rate sigma(E) = E / 1e6 per unit chord fraction, stopping 0.5 E, a cutoff
// stop at the segment midpoint, a channel open above open_above, scripted
// draw pairs and per-event losses.  Records what the core handed each
// policy.  0, or the error code of a fail-closed exit.
typedef struct orc_interleave_record {
    int32_t n_frozen, n_segments, n_draws, n_open, n_final, n_stops;
    double frozen[16], taus[16];
    double seg_begin[16], seg_end[16], seg_frozen[16];
    uint64_t draw_ordinals[16];
    double open_energies[16];
    double final_fractions[16];
    double stops[16];
    double kinetic_end;
} orc_interleave_record;

int orc_interleave_probe(double kinetic, double cutoff, double open_above, const double* draws, int n_draws,
                         const double* losses, int n_losses, int consume, orc_interleave_record* rec)
{
    std::memset(rec, 0, sizeof(*rec));
    double k = kinetic;
    int dc = 0, lc = 0;
    auto put = [](double* a, int32_t& n, double v) { if (n < 16) a[n] = v; ++n; };
    try {
        run_event_interleave(
            0.0, 1.0, k, cutoff, 2, 64,
            [&](double frozen_e, double b, double e) {
                double const tau = (e - b) * frozen_e / 1.0e6;
                int32_t nf = rec->n_frozen;
                put(rec->frozen, nf, frozen_e);
                put(rec->taus, rec->n_frozen, tau);
                return std::array<double, 3>{tau, 0.0, 0.0};
            },
            [&](uint64_t ord) {
                if (rec->n_draws < 16) rec->draw_ordinals[rec->n_draws] = ord;
                ++rec->n_draws;
                if (dc >= n_draws) fail(ORC_ERR_INVALID, "the script holds no more draws");
                std::pair<double, double> const u{draws[2 * dc], draws[2 * dc + 1]};
                ++dc;
                return u;
            },
            [&](double tau, double b, double) { return b + tau * 1.0e6 / k; },
            [&](double b, double e, double frozen_e) -> SegLimit {
                int32_t ns = rec->n_segments, ns2 = rec->n_segments;
                put(rec->seg_begin, ns, b);
                put(rec->seg_end, ns2, e);
                put(rec->seg_frozen, rec->n_segments, frozen_e);
                k -= 0.5 * (e - b) * frozen_e;
                if (k < cutoff) return {0.5 * (b + e), true};
                return {e, false};
            },
            [&](double f) { put(rec->stops, rec->n_stops, f); },
            [](double) {},
            [&](int, double e) {
                put(rec->open_energies, rec->n_open, e);
                return e > open_above;
            },
            [&](int channel, double f) -> bool {
                if (channel != 0) fail(ORC_ERR_INVALID, "the single-rate medium selected another channel");
                put(rec->final_fractions, rec->n_final, f);
                if (lc >= n_losses) fail(ORC_ERR_INVALID, "the script holds no more event losses");
                k -= losses[lc++];
                return consume == 0;
            });
    } catch (OracleError const& err) {
        rec->kinetic_end = k;
        return err.code;
    }
    rec->kinetic_end = k;
    return 0;
}

// The substep skeleton with recording policies: tags kind * 1000 + substep
// (1 build, 2 events, 3 finish, 4 advance); a build that fails when
// build_ok == 0, an endpoint refused at substep refuse_at, an exiting chord
// at substep exit_at, a stop after the events when stop_after != 0.
int orc_substep_probe(uint64_t subcycles, double kinetic, int build_ok, uint64_t refuse_at, uint64_t exit_at,
                      int stop_after, int32_t* tags, int cap, uint64_t* completed, int* exited)
{
    int n = 0;
    auto tag = [&](int kind, uint64_t sc) {
        if (n < cap) tags[n] = kind * 1000 + static_cast<int32_t>(sc);
        ++n;
    };
    uint64_t done = 0;
    bool ex = false;
    double const k = kinetic;
    run_substep_interleave(
        subcycles, k, 1.0e3, done, ex,
        [&](uint64_t sc, double& b, double& e) {
            tag(1, sc);
            b = 0.0;
            e = 1.0;
            return build_ok != 0;
        },
        [&](uint64_t sc, double, double) { tag(2, sc); },
        [&] { return stop_after != 0; },
        [&](uint64_t sc, double, double) {
            tag(3, sc);
            return false;
        },
        [&](double) {
            uint64_t const sc = done;
            tag(4, sc);
            return sc != refuse_at;
        },
        [&] { return done == exit_at; });
    *completed = done;
    *exited = ex ? 1 : 0;
    return n;
}

// The clipped chord's boundary crossing: out = [valid, faces, start_r,
// start_z, cross_r, cross_z, exits, end_f]; 0 or an error code.
int orc_boundary_crossing(double sr, double st, double sz, double er_, double et, double ez_, double r_lo, double r_hi,
                          double z_lo, double z_hi, double* out)
{
    try {
        Path const p = clip_rz_path(sr, st, sz, er_, et, ez_, r_lo, r_hi, z_lo, z_hi);
        BoundaryCrossing const bc = boundary_crossing(p);
        double const v[8] = {bc.valid ? 1.0 : 0.0, static_cast<double>(bc.faces), bc.start_r, bc.start_z,
                             bc.cross_r, bc.cross_z, p.exits ? 1.0 : 0.0, p.end_f};
        std::memcpy(out, v, sizeof(v));
        return 0;
    } catch (OracleError const& err) {
        return err.code;
    }
}

// Commit of the running point at fraction f of the clipped chord: 1 when the
// point is in the half-open domain (run <- point), 0 when refused (run kept).
int orc_advance_running_position(double sr, double st, double sz, double er_, double et, double ez_, double r_lo,
                                 double r_hi, double z_lo, double z_hi, double f, double* run)
{
    try {
        Path const p = clip_rz_path(sr, st, sz, er_, et, ez_, r_lo, r_hi, z_lo, z_hi);
        return advance_running_position(p, f, run[0], run[1], run[2]) ? 1 : 0;
    } catch (OracleError const& err) {
        return -err.code;
    }
}

// OutsideCertifiedDomain of a chord clipped to the given domain.
int orc_outside_certified(double r_lo, double r_hi, double z_lo, double z_hi, double r, double z)
{
    Path p;
    p.r_lo = r_lo; p.r_hi = r_hi; p.z_lo = z_lo; p.z_hi = z_hi;
    return p.outside_certified(r, z) ? 1 : 0;
}

}  // extern "C"

// RreaValidateSecondaryGroupConservation, OV/RreaParticleInteraction.H:1099-1252
extern "C" int orc_validate_group(const orc_group* g, double cutoff, double* res)
{
    if (res) res[0] = res[1] = 0.0;
    bool const required = !(g->process == 0 || g->process == 1);
    if (!required) return 1 | 2 | 4 | 8;
    int out = 1;
    if (g->closure_mode == 0) return out;
    out |= 2;
    if (g->closure_mode == 2) {
        bool const zero = g->n_particles == 0 && g->n_sources == 0 && g->remove_parent && !g->parent_still_valid
            && !g->parent_survives && g->parent_weight == 0.0 && g->parent_energy_before_eV == 0.0
            && g->parent_energy_after_eV == 0.0 && g->local_energy_eV == 0.0 && g->escaped_energy_eV == 0.0
            && g->escaped_charge_e == 0.0;
        return zero ? (out | 4 | 8) : out;
    }
    if (g->closure_mode != 1 || g->n_particles < 0 || g->n_particles > 4 || g->n_sources < 0 || g->n_sources > 4) {
        return out;
    }
    Closure c;
    c.finite = (g->parent_survives != 0) == (g->remove_parent == 0) && g->parent_species >= 0 && g->parent_species <= 2;
    c.parent(g->parent_species, g->parent_weight, g->parent_energy_before_eV, g->parent_energy_after_eV,
             g->parent_survives != 0);
    c.residual_terms(g->local_energy_eV, g->escaped_energy_eV, g->escaped_charge_e);
    for (int i = 0; i < g->n_particles; ++i) {
        orc_group_entry const& p = g->particles[i];
        c.finite = c.finite && p.kind >= 0 && p.kind <= 2;
        c.particle(p.kind, p.weight, p.energy_eV);
    }
    for (int i = 0; i < g->n_sources; ++i) {
        orc_group_entry const& e = g->sources[i];
        c.finite = c.finite && std::isfinite(e.energy_eV);
        if (e.kind == 0) c.ionization(e.weight, e.energy_eV, cutoff);
        else if (e.kind == 1) c.positive_ion(e.weight);
        else if (e.kind == 2) c.low_electron(e.weight);
        else c.finite = false;
    }
    int const v = c.verdict(cutoff, res);
    if (v == 1) return out;
    out |= 4;
    return v == 0 ? (out | 8) : out;
}

extern "C" void orc_fold_tallies(double* into, const double* from) { fold_tallies(into, from); }
