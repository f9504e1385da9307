// Device-side physics of the RREA transport step, written for gfx950.
//
// One lane owns one particle for the whole step.  The arithmetic follows the
// reference expression by expression (the citations give file:line) because
// the step must agree with the reference CPU path to 1e-12 per particle-step;
// the structure does not: the chord is non-periodic and has at most one
// in-domain piece, so the reference's heap vectors become scalars; the
// charged tables share one energy grid, so one bracket search serves every
// table read at one energy; every node logarithm the reference takes at
// query time is precomputed on the host with the same libm.
//
// Citation abbreviations: OV/ = native_amrex_rrea/warpx_overlay/Source/Rrea/,
// NA/ = native_amrex_rrea/.
#pragma once

#include <hip/hip_runtime.h>

#include <stdint.h>

namespace rrea_dev {

constexpr double kMeC2 = 510998.95000;           // NA/include/rrea/RreaConstants.H:9
constexpr double kC = 299792458.0;               // RreaConstants.H:10
constexpr double kWAir = 33.97;                  // RreaConstants.H:30
constexpr double kTwoPi = 6.283185307179586476925286766559;
constexpr double kPi = 3.141592653589793238462643383279502884;  // NA/include/rrea/RreaConstants.H:8
constexpr double kEps = 2.220446049250313e-16;   // DBL_EPSILON
constexpr double kDblMin = 2.2250738585072014e-308;
// Two-point Gauss-Legendre node 1/sqrt(3) (NA/include/rrea/RreaGaussLegendre8.H:29-30),
// the long double literal rounded to double.
constexpr double kGl2Node = 0.57735026918962576451;

enum Err : int32_t { kOk = 0, kErrRange = 1, kErrDomain = 2, kErrGuard = 3, kErrInvalid = 4, kErrCapacity = 5 };

// std::max / std::min semantics (first argument on ties and NaN), which
// fmax/fmin do not guarantee.
__device__ __forceinline__ double dmax(double a, double b) { return (a < b) ? b : a; }
__device__ __forceinline__ double dmin(double a, double b) { return (b < a) ? b : a; }
__device__ __forceinline__ bool finite(double v) { return isfinite(v); }

// sin and cos of one angle through one argument reduction.  The device
// library's sincos returns the bits its sin and cos return (measured on
// gfx950 over 1.3e8 arguments, |x| from below 3.2 up to 1e300, with signed
// zeros, +-pi, the smallest subnormal, infinity and NaN), so the pair costs
// one reduction instead of two and every result is unchanged.
__device__ __forceinline__ void sin_cos(double x, double& s, double& c) { sincos(x, &s, &c); }

// ---------------------------------------------------------------------------
// Counter RNG, counter_splitmix64_v1 -- OV/RreaRng.cpp:9-44
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint64_t splitmix64(uint64_t v)
{
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30U)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27U)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31U);
}

// The seed/particle/step words of a particle's key are fixed for the whole
// step, so their mixed XOR is formed once per particle.
__device__ __forceinline__ uint64_t rng_prefix(uint64_t seed, uint64_t pid, uint64_t step)
{
    return splitmix64(seed) ^ splitmix64(pid + 0x100000001b3ULL) ^ splitmix64(step + 0x9e3779b97f4a7c15ULL);
}

__device__ __forceinline__ double rng_uniform(uint64_t prefix, uint64_t idx, uint64_t chan)
{
    uint64_t s = prefix;
    s ^= splitmix64(idx + 0xbf58476d1ce4e5b9ULL);
    s ^= splitmix64(chan + 0x94d049bb133111ebULL);
    return (static_cast<double>(splitmix64(s) >> 12U) + 0.5) * 0x1.0p-52;
}

// RNG channel ids, OV/RreaRng.H:16-98
enum : uint64_t {
    kChHardMollerEnergy = 22, kChHardMollerAzimuth = 23,
    kChBremsPhotonEnergy = 42, kChBremsPhotonAngle = 43, kChBremsPhotonAzimuth = 44,
    kChGsAngle = 89, kChGsAzimuth = 90, kChGsFewCount = 91, kChGsFewScatter = 92, kChGsFewAzimuth = 93,
    kChCompetingOpticalDepth = 107, kChCompetingChannel = 108,
    kChContinuationOpticalDepth = 124, kChContinuationChannel = 125,
    kChPositronBhabhaEnergy = 74, kChPositronBhabhaAzimuth = 130,
    kChPositronBremsPhotonEnergy = 76, kChPositronBremsPhotonAngle = 77, kChPositronBremsPhotonAzimuth = 78,
    kChPositronAnnihilationPolar = 80, kChPositronAnnihilationAzimuth = 81,
    kChPositronGsAngle = 94, kChPositronGsAzimuth = 95, kChPositronGsFewCount = 96, kChPositronGsFewScatter = 97,
    kChPositronGsFewAzimuth = 98,
    kChPositronCompetingOpticalDepth = 109, kChPositronCompetingChannel = 110,
    kChPositronContinuationOpticalDepth = 126, kChPositronContinuationChannel = 127,
    kChPhotonInteraction = 51, kChPhotonChannel = 52, kChComptonRecoil = 53, kChComptonScatter = 54,
    kChPhotoelectric = 55, kChPairAngle = 57, kChComptonScatterAzimuth = 58, kChPairAzimuth = 59,
    kChPhotoelectricAngle = 60, kChPhotoelectricAzimuth = 65, kChPairInvariant = 66,
    kChPhotonContinuationOpticalDepth = 128, kChPhotonContinuationChannel = 129,
};

// RreaSecondaryProcess ids, OV/RreaParticleInteraction.H:34-48
enum : int32_t {
    kProcHardMoller = 2, kProcHardBhabha = 3, kProcBrems = 4, kProcCompton = 5, kProcPhotoelectric = 6,
    kProcPairNuclear = 7, kProcPairTriplet = 8, kProcAnnihilation = 9,
};

// ---------------------------------------------------------------------------
// Secondary-group conservation -- RreaValidateSecondaryGroupConservation
// OV/RreaParticleInteraction.H:1099-1251.  One interaction's charge (in e x
// macro weight) and total energy (kinetic + rest, macro-weighted eV) before
// and after: the parent before and, if it survives, after; its particles;
// its material sources (an ionization event or a positive ion consumes one
// neutral target electron: its rest energy in, +e out; a sub-cutoff
// ionization also releases that electron into the fluid: -e and its rest
// energy out; a low-energy electron source: -e and its rest energy out);
// the weighted local deposit and escaped energy and charge.  Valid when
// every input is finite and in range and both residuals are within
// 64 eps x max(1, throughput).  Species: 0 e-, 1 e+, 2 photon.
// ---------------------------------------------------------------------------
__host__ __device__ inline double species_charge(int species) { return species == 0 ? -1.0 : species == 1 ? 1.0 : 0.0; }
__host__ __device__ inline double species_rest(int species) { return species == 2 ? 0.0 : kMeC2; }

struct GroupClosure {
    // q_through holds the particles' and sources' charge magnitudes; the
    // parent's input and its output plus any escaped charge enter as the two
    // magnitudes |q_in| + |q_out + q_escaped|.
    double in_q = 0.0, out_q = 0.0, in_e = 0.0, out_e = 0.0, q_through = 0.0;
    double parent_in_q = 0.0, parent_out_q = 0.0, escaped_q = 0.0;
    bool ok = true;

    __host__ __device__ static bool fin(double v) { return v == v && v - v == 0.0; }

    __host__ __device__ void parent(int species, double w, double before, double after, bool survives)
    {
        ok = ok && fin(w) && fin(before) && fin(after) && w > 0.0 && before >= 0.0 && after >= 0.0;
        double const q = species_charge(species) * w;
        double const m = species_rest(species);
        in_q += q;
        double const oq = survives ? q : 0.0;
        out_q += oq;
        in_e += w * (before + m);
        out_e += survives ? w * (after + m) : 0.0;
        parent_in_q = q;
        parent_out_q = oq;
    }
    __host__ __device__ void local(double weighted_eV)
    {
        ok = ok && fin(weighted_eV) && weighted_eV >= 0.0;
        out_e += weighted_eV;
    }
    __host__ __device__ void escaped(double weighted_eV, double charge_e)
    {
        ok = ok && fin(weighted_eV) && fin(charge_e) && weighted_eV >= 0.0;
        out_e += weighted_eV;
        out_q += charge_e;
        escaped_q += charge_e;
    }
    // a group record's local and escaped terms, summed as one output
    __host__ __device__ void residual_terms(double local_eV, double escaped_eV, double charge_e)
    {
        ok = ok && fin(local_eV) && fin(escaped_eV) && fin(charge_e) && local_eV >= 0.0 && escaped_eV >= 0.0;
        out_q += charge_e;
        escaped_q += charge_e;
        out_e += local_eV + escaped_eV;
    }
    __host__ __device__ void particle(int species, double w, double e)
    {
        ok = ok && fin(w) && fin(e) && w > 0.0 && e >= 0.0;
        double const q = species_charge(species) * w;
        out_q += q;
        q_through += q < 0.0 ? -q : q;
        out_e += w * (e + species_rest(species));
    }
    __host__ __device__ void ionization(double w, double secondary_eV, double cutoff)
    {
        ok = ok && fin(w) && fin(secondary_eV) && w > 0.0;
        in_e += w * kMeC2;
        out_q += w;
        q_through += w;
        if (secondary_eV < cutoff) {
            out_q -= w;
            q_through += w;
            out_e += w * kMeC2;
        }
    }
    __host__ __device__ void positive_ion(double w)
    {
        ok = ok && fin(w) && w > 0.0;
        in_e += w * kMeC2;
        out_q += w;
        q_through += w;
    }
    __host__ __device__ void low_electron(double w)
    {
        ok = ok && fin(w) && w > 0.0;
        out_q -= w;
        q_through += w;
        out_e += w * kMeC2;
    }
    // 0: valid; otherwise why not (1 non-finite or out of range, 2 charge, 3 energy)
    __host__ __device__ int verdict(double cutoff, double* q_res = nullptr, double* e_res = nullptr) const
    {
        if (!(ok && fin(cutoff) && cutoff > 0.0 && fin(in_q) && fin(out_q) && fin(in_e) && fin(out_e))) return 1;
        double const dq = out_q - in_q;
        double const de = out_e - in_e;
        if (q_res) *q_res = dq;
        if (e_res) *e_res = de;
        double const po = parent_out_q + escaped_q;
        double const through = (parent_in_q < 0.0 ? -parent_in_q : parent_in_q) + (po < 0.0 ? -po : po) + q_through;
        double const tq = 64.0 * kEps * (through > 1.0 ? through : 1.0);
        double const ai = in_e < 0.0 ? -in_e : in_e, ao = out_e < 0.0 ? -out_e : out_e;
        double const te = 64.0 * kEps * (ai > ao ? (ai > 1.0 ? ai : 1.0) : (ao > 1.0 ? ao : 1.0));
        if ((dq < 0.0 ? -dq : dq) > tq) return 2;
        if ((de < 0.0 ? -de : de) > te) return 3;
        return 0;
    }
};

constexpr uint64_t kSegmentCap = uint64_t{1} << 14;  // OV/RreaParticleInteraction.H:642-650
constexpr uint64_t kEventCap = uint64_t{1} << 13;

// RreaTryLiveHazardRngIndex OV/RreaParticleInteraction.H:664-710 for a live
// root (lineage 1) of species 0 (e-) or 1 (e+)
__device__ __forceinline__ bool live_hazard_index(uint64_t segment, uint64_t event, uint64_t& out,
                                                  uint64_t species = 0U)
{
    if (event == 0U) {
        if (segment >= kSegmentCap) return false;
        out = segment;
        return true;
    }
    if (segment >= kSegmentCap || event >= kEventCap) return false;
    out = (species << 49) | (uint64_t{1} << 35) | (segment << 21) | (event << 8);
    return true;
}

// ---------------------------------------------------------------------------
// Tables (device view)
// ---------------------------------------------------------------------------
enum ChargedTable : int {
    kECol = 0, kESoft, kEMoller, kEBrems, kEEltr, kPCol, kPSoft, kPBhabha, kPBrems, kPEltr, kPAnnih, kChargedCount
};

struct DevCdf {
    const double* primary;
    const double* log_primary;
    const int32_t* start;
    const double* u;
    const double* v;
    const double* mean;       // group means, mean_conditional_cdf's group_mean
    const double* log_mean;
    int32_t n_groups;
};

struct DevAngle {
    const double* e; const double* log_e;
    const double* n; const double* log_n;
    const double* u;
    const double* log_u_ratio;  // log(u[i+1] / u[i])
    const double* log_omc;      // log(one_minus_cos)
    int32_t n_e, n_n, n_u;
};

// A bin table of starting indices for a search over a nondecreasing node
// array: bin b covers keys [lo + b/scale, lo + (b+1)/scale) and holds the
// lower bound of its left edge.  It only shortens the search; the scans in
// guided_lower_bound / guided_upper_bound return exactly std::lower_bound /
// std::upper_bound from any start.
struct Guide {
    const int32_t* idx;
    int32_t nbins;
    double lo, scale;
};

struct DevElastic {
    const double* e; const double* log_e;
    const double* mfp; const double* eta;
    int32_t n;
    Guide guide;                // over log e
};

struct DevTables {
    const double* cx;               // shared charged energy grid
    const double* clogx;
    int32_t nc;
    const double* cy[kChargedCount];
    const double* clogy[kChargedCount];
    int32_t cmode[kChargedCount];
    const double* hm_mean;          // hard-Moller mean secondary on the grid
    Guide cguide;                   // over log of the charged grid
    DevCdf moller, ebrems, bhabha, pbrems;
    DevAngle gs_e, wz_e, gs_p, wz_p;
    DevElastic el_e, el_p;
    double hard_moller_threshold;
    double charged_min, charged_max;
    double gs_few, gs_iso, gs_transition;
    // photon process inverse lengths: one grid (with duplicated edge nodes)
    // shared by the Compton, photoelectric, pair-nuclear and pair-triplet columns
    const double* gx;
    const double* glogx;
    int32_t ng;
    Guide gguide;
    const double* gy[4];
    const double* glogy[4];
    int32_t gmode[4];
    double photon_min, photon_max;
    // photoelectric subshell CDF groups
    const double* pe_primary;
    const int32_t* pe_start;
    const double* pe_u;
    const double* pe_binding;
    int32_t pe_n;
    // pair final-state bank: training blocks by (channel, energy); records of
    // (pair electron, positron, triplet recoil) fractions and directions
    const int32_t* pair_channel;
    const double* pair_energy;
    const int32_t* pair_start;
    const double* pair_frac;
    const double* pair_dir;
    int32_t pair_nblocks;
};

// One bracket of the shared charged grid for one query energy: the
// std::lower_bound + node-exact test of RreaInteractionTable::Interpolate
// (NA/src/RreaInteractionTables.cpp:1580-1630), computed once.
struct Bracket {
    int32_t lo, hi;
    int32_t kind;   // 0 interior, 1 first node, 2 last node, 3 exact node hi
    double t;       // (log q - log x_lo) / (log x_hi - log x_lo)
    double lin_t;   // (q - x_lo) / (x_hi - x_lo), for linear tables
};

__device__ __forceinline__ int32_t guide_start(Guide const& g, double key)
{
    double const f = (key - g.lo) * g.scale;
    int32_t b = 0;
    if (f > 0.0) b = f < static_cast<double>(g.nbins - 1) ? static_cast<int32_t>(f) : g.nbins - 1;
    return g.idx[b];
}

// std::lower_bound (first x[i] >= q) from the guide's start for `key`.
__device__ __forceinline__ int32_t guided_lower_bound(const double* x, int32_t n, double q, Guide const& g, double key)
{
    int32_t i = guide_start(g, key);
    while (i > 0 && !(x[i - 1] < q)) --i;
    while (i < n && x[i] < q) ++i;
    return i;
}

// std::upper_bound (first x[i] > q) from the guide's start for `key`.
__device__ __forceinline__ int32_t guided_upper_bound(const double* x, int32_t n, double q, Guide const& g, double key)
{
    int32_t i = guide_start(g, key);
    while (i > 0 && q < x[i - 1]) --i;
    while (i < n && !(q < x[i])) ++i;
    return i;
}

__device__ __forceinline__ int32_t lower_bound(const double* x, int32_t n, double q)
{
    int32_t first = 0, count = n;
    while (count > 0) {
        int32_t const step = count >> 1;
        int32_t const it = first + step;
        if (x[it] < q) { first = it + 1; count -= step + 1; } else { count = step; }
    }
    return first;
}

__device__ __forceinline__ bool charged_bracket(DevTables const& T, double q, Bracket& b)
{
    if (!finite(q) || q < T.cx[0] || q > T.cx[T.nc - 1]) {
        // a safe bracket (first node) so that a caller that carries on to
        // its error exit reads nothing out of bounds
        b.lo = 0; b.hi = 0; b.kind = 1; b.t = 0.0; b.lin_t = 0.0;
        return false;
    }
    double const lq = log(q);
    int32_t const up = guided_lower_bound(T.cx, T.nc, q, T.cguide, lq);
    b.lo = up - 1;
    b.hi = up;
    b.t = 0.0;
    b.lin_t = 0.0;
    if (up == 0) { b.kind = 1; return true; }
    if (up == T.nc) { b.kind = 2; return true; }
    if (T.cx[up] == q) { b.kind = 3; return true; }
    b.kind = 0;
    b.t = (lq - T.clogx[b.lo]) / (T.clogx[b.hi] - T.clogx[b.lo]);
    b.lin_t = (q - T.cx[b.lo]) / (T.cx[b.hi] - T.cx[b.lo]);
    return true;
}

// charged_bracket for an energy q near one already bracketed: when q lies in
// (x[lo], x[hi]] of `near` (an interior or exact-node bracket), lower_bound
// is near's hi, so the bracket follows from those two nodes with the same
// arithmetic and no guided search (one round of loads instead of the
// guide's and the search's); otherwise the full search.
__device__ __forceinline__ bool charged_bracket_near(DevTables const& T, double q, Bracket const& near, Bracket& b)
{
    if (near.kind == 0 || near.kind == 3) {
        int32_t const lo = near.lo, hi = near.hi;
        double const xlo = T.cx[lo], xhi = T.cx[hi];
        double const llo = T.clogx[lo], lhi = T.clogx[hi];
        if (xlo < q && q <= xhi) {
            b.lo = lo;
            b.hi = hi;
            b.t = 0.0;
            b.lin_t = 0.0;
            if (xhi == q) { b.kind = 3; return true; }
            double const lq = log(q);
            b.kind = 0;
            b.t = (lq - llo) / (lhi - llo);
            b.lin_t = (q - xlo) / (xhi - xlo);
            return true;
        }
    }
    return charged_bracket(T, q, b);
}

// The same bracket on the photon rate grid, whose duplicated absorption-edge
// nodes resolve to the low side at an exact edge and to the high side just
// above it (lower_bound + node-exact test).
__device__ __forceinline__ bool photon_bracket(DevTables const& T, double q, Bracket& b)
{
    if (!finite(q) || q < T.gx[0] || q > T.gx[T.ng - 1]) {
        b.lo = 0; b.hi = 0; b.kind = 1; b.t = 0.0; b.lin_t = 0.0;
        return false;
    }
    double const lq = log(q);
    int32_t const up = guided_lower_bound(T.gx, T.ng, q, T.gguide, lq);
    b.lo = up - 1;
    b.hi = up;
    b.t = 0.0;
    b.lin_t = 0.0;
    if (up == 0) { b.kind = 1; return true; }
    if (up == T.ng) { b.kind = 2; return true; }
    if (T.gx[up] == q) { b.kind = 3; return true; }
    b.kind = 0;
    b.t = (lq - T.glogx[b.lo]) / (T.glogx[b.hi] - T.glogx[b.lo]);
    b.lin_t = (q - T.gx[b.lo]) / (T.gx[b.hi] - T.gx[b.lo]);
    return true;
}

__device__ __forceinline__ double photon_value(DevTables const& T, int col, Bracket const& b, int32_t& err)
{
    const double* y = T.gy[col];
    if (b.kind == 1) return y[0];
    if (b.kind == 2) return y[T.ng - 1];
    if (b.kind == 3) return y[b.hi];
    int32_t const mode = T.gmode[col];
    double const ylo = y[b.lo], yhi = y[b.hi];
    if (mode == 3) return ylo + b.lin_t * (yhi - ylo);
    bool const positive = ylo > 0.0 && yhi > 0.0;
    if (mode == 2 || (mode == 1 && !positive)) return ylo + b.t * (yhi - ylo);
    if (!positive) { if (!err) err = kErrInvalid; return 0.0; }
    const double* ly = T.glogy[col];
    return exp(ly[b.lo] + b.t * (ly[b.hi] - ly[b.lo]));
}

// RreaInteractionTable::Interpolate on a precomputed bracket.
__device__ __forceinline__ double charged_value(DevTables const& T, int table, Bracket const& b, int32_t& err)
{
    const double* y = T.cy[table];
    if (b.kind == 1) return y[0];
    if (b.kind == 2) return y[T.nc - 1];
    if (b.kind == 3) return y[b.hi];
    int32_t const mode = T.cmode[table];
    double const ylo = y[b.lo], yhi = y[b.hi];
    if (mode == 3) return ylo + b.lin_t * (yhi - ylo);
    bool const positive = ylo > 0.0 && yhi > 0.0;
    if (mode == 2 || (mode == 1 && !positive)) return ylo + b.t * (yhi - ylo);
    if (!positive) { if (!err) err = kErrInvalid; return 0.0; }
    const double* ly = T.clogy[table];
    return exp(ly[b.lo] + b.t * (ly[b.hi] - ly[b.lo]));
}

// interp_semilogx_strict NA/src/RreaInteractionTables.cpp:624-648 on the
// charged grid (ElectronHardMollerEnergyLoss :2414-2440).
__device__ __forceinline__ double charged_semilog(DevTables const& T, const double* y, Bracket const& b, double q)
{
    if (b.kind == 1) return y[0];            // q == x.front()
    if (b.kind == 2 || q == T.cx[T.nc - 1]) return y[T.nc - 1];
    if (b.kind == 3) return y[b.hi];
    double const t = (log(q) - T.clogx[b.lo]) / (T.clogx[b.hi] - T.clogx[b.lo]);
    return y[b.lo] + t * (y[b.hi] - y[b.lo]);
}

// interp_semilogx_strict on the elastic parameter grid.
__device__ __forceinline__ double elastic_semilog(DevElastic const& E, const double* y, double q, int32_t& err)
{
    if (!finite(q) || q < E.e[0] || q > E.e[E.n - 1]) { if (!err) err = kErrRange; return 0.0; }
    if (q == E.e[0]) return y[0];
    if (q == E.e[E.n - 1]) return y[E.n - 1];
    double const lq = log(q);
    int32_t const hi = guided_lower_bound(E.e, E.n, q, E.guide, lq);
    int32_t const lo = hi - 1;
    if (E.e[hi] == q) return y[hi];
    double const t = (lq - E.log_e[lo]) / (E.log_e[hi] - E.log_e[lo]);
    return y[lo] + t * (y[hi] - y[lo]);
}

// interp_semilogx_strict on the elastic grid as a reusable bracket: the
// collision mfp and the screening eta of one energy share it.
struct EBracket { int32_t lo, hi, kind; double t; };  // kind 0 interior, 1 first, 2 last, 3 exact node

__device__ __forceinline__ bool elastic_bracket(DevElastic const& E, double q, EBracket& b)
{
    b.lo = 0; b.hi = 0; b.kind = 1; b.t = 0.0;
    if (!finite(q) || q < E.e[0] || q > E.e[E.n - 1]) return false;
    if (q == E.e[0]) return true;
    if (q == E.e[E.n - 1]) { b.kind = 2; return true; }
    double const lq = log(q);
    int32_t const hi = guided_lower_bound(E.e, E.n, q, E.guide, lq);
    b.hi = hi;
    b.lo = hi - 1;
    if (E.e[hi] == q) { b.kind = 3; return true; }
    b.kind = 0;
    b.t = (lq - E.log_e[b.lo]) / (E.log_e[b.hi] - E.log_e[b.lo]);
    return true;
}

// elastic_bracket for an energy q near one already bracketed, as
// charged_bracket_near: q in (e[lo], e[hi]) of an interior or exact-node
// `near` gives its nodes; q == e[hi] is that node (the last node's own kind
// when hi is the last); otherwise the full search.
__device__ __forceinline__ bool elastic_bracket_near(DevElastic const& E, double q, EBracket const& near, EBracket& b)
{
    if (near.kind == 0 || near.kind == 3) {
        int32_t const lo = near.lo, hi = near.hi;
        double const xlo = E.e[lo], xhi = E.e[hi];
        double const llo = E.log_e[lo], lhi = E.log_e[hi];
        if (xlo < q && q <= xhi) {
            b.lo = 0; b.hi = 0; b.kind = 1; b.t = 0.0;
            if (q == xhi && hi == E.n - 1) { b.kind = 2; return true; }
            b.hi = hi;
            b.lo = lo;
            if (q == xhi) { b.kind = 3; return true; }
            double const lq = log(q);
            b.kind = 0;
            b.t = (lq - llo) / (lhi - llo);
            return true;
        }
    }
    return elastic_bracket(E, q, b);
}

__device__ __forceinline__ double elastic_value(DevElastic const& E, const double* y, EBracket const& b)
{
    if (b.kind == 1) return y[0];
    if (b.kind == 2) return y[E.n - 1];
    if (b.kind == 3) return y[b.hi];
    return y[b.lo] + b.t * (y[b.hi] - y[b.lo]);
}

// sample_conditional_cdf :995-1069 -- one group
__device__ __forceinline__ double cdf_group_sample(DevCdf const& G, int32_t g, double cu)
{
    int32_t const s = G.start[g];
    int32_t const n = G.start[g + 1] - s;
    const double* u = G.u + s;
    const double* v = G.v + s;
    int32_t const up = lower_bound(u, n, cu);
    if (up == 0) return v[0];
    if (up == n) return v[n - 1];
    int32_t const lo = up - 1;
    double const denom = u[up] - u[lo];
    double const t = denom > 0.0 ? (cu - u[lo]) / denom : 0.0;
    return v[lo] + t * (v[up] - v[lo]);
}

// sample_conditional_cdf (log_interp) :995-1069, and the Moller variant
// SampleMollerSecondaryEnergy :2584-2653 (force_log: no positivity test).
__device__ __forceinline__ double cdf_sample(DevCdf const& G, double primary, double cu, bool force_log, int32_t& err)
{
    cu = dmax(0.0, dmin(1.0, cu));
    int32_t const ng = G.n_groups;
    if (!finite(primary) || primary < G.primary[0] || primary > G.primary[ng - 1]) {
        if (!err) err = kErrRange;
        return 0.0;
    }
    if (!force_log) {
        if (primary == G.primary[0]) return cdf_group_sample(G, 0, cu);
        if (primary == G.primary[ng - 1]) return cdf_group_sample(G, ng - 1, cu);
    }
    int32_t const up = lower_bound(G.primary, ng, primary);
    if (up == 0) return cdf_group_sample(G, 0, cu);
    if (up == ng) return cdf_group_sample(G, ng - 1, cu);
    int32_t const lo = up - 1;
    if (G.primary[up] == primary) return cdf_group_sample(G, up, cu);
    double const y0 = cdf_group_sample(G, lo, cu);
    double const y1 = cdf_group_sample(G, up, cu);
    double const t = (log(primary) - G.log_primary[lo]) / (G.log_primary[up] - G.log_primary[lo]);
    if (force_log || (y0 > 0.0 && y1 > 0.0)) return exp(log(y0) + t * (log(y1) - log(y0)));
    return y0 + t * (y1 - y0);
}

// mean_conditional_cdf :1071-1110 with the group means precomputed
__device__ __forceinline__ double cdf_mean(DevCdf const& G, double primary, int32_t& err)
{
    int32_t const ng = G.n_groups;
    if (primary < G.primary[0] || primary > G.primary[ng - 1]) { if (!err) err = kErrRange; return 0.0; }
    int32_t const up = lower_bound(G.primary, ng, primary);
    if (up == 0) return G.mean[0];
    if (up == ng) return G.mean[ng - 1];
    if (G.primary[up] == primary) return G.mean[up];
    int32_t const lo = up - 1;
    double const y0 = G.mean[lo], y1 = G.mean[up];
    double const t = (log(primary) - G.log_primary[lo]) / (G.log_primary[up] - G.log_primary[lo]);
    return y0 > 0.0 && y1 > 0.0 ? exp(G.log_mean[lo] + t * (G.log_mean[up] - G.log_mean[lo])) : y0 + t * (y1 - y0);
}

// log_bracket :807-829
__device__ __forceinline__ bool log_bracket(const double* g, const double* lg, int32_t n, double q, int32_t& i, double& frac)
{
    int32_t const last = n - 1;
    if (!finite(q) || q < g[0] || q > g[last]) return false;
    if (q == g[0]) { i = 0; frac = 0.0; return true; }
    if (q == g[last]) { i = last - 1; frac = 1.0; return true; }
    int32_t lo = 0, hi = last;
    while (hi - lo > 1) { int32_t const m = (lo + hi) / 2; if (g[m] <= q) lo = m; else hi = m; }
    i = lo;
    if (lg[hi] == lg[lo]) { frac = (q == g[lo]) ? 0.0 : 1.0; return true; }
    frac = (log(q) - lg[lo]) / (lg[hi] - lg[lo]);
    return true;
}

// sample_gs_angle_cos :831-891
__device__ double sample_gs_angle_cos(DevAngle const& A, double energy, double path_n, double u, int32_t& err)
{
    int32_t ie = 0, ip = 0;
    double te = 0.0, tn = 0.0;
    if (!log_bracket(A.e, A.log_e, A.n_e, energy, ie, te) || !log_bracket(A.n, A.log_n, A.n_n, path_n, ip, tn)) {
        if (!err) err = kErrRange;
        return 1.0;
    }
    int32_t const np = A.n_n, nu = A.n_u;
    auto omc_at = [&](int32_t iu) {
        double const la = A.log_omc[(ie * np + ip) * nu + iu];
        double const lb = A.log_omc[((ie + 1) * np + ip) * nu + iu];
        double const lc = A.log_omc[(ie * np + ip + 1) * nu + iu];
        double const ld = A.log_omc[((ie + 1) * np + ip + 1) * nu + iu];
        double const lx = (1 - te) * (1 - tn) * la + te * (1 - tn) * lb + (1 - te) * tn * lc + te * tn * ld;
        double const cos_node = 1.0 - exp(lx);
        return dmax(1.0 - cos_node, kDblMin);
    };
    int32_t const first = 1, last = nu - 2;
    double omc;
    if (last <= first) {
        omc = omc_at(first);
    } else if (u <= A.u[first]) {
        double const d1 = omc_at(first);
        double const d2 = omc_at(first + 1);
        double const u1 = A.u[first];
        double const slope = log(d2 / d1) / A.log_u_ratio[first];
        double const u_eff = dmax(u, u1 * 1.0e-6);
        omc = d1 * exp(slope * log(u_eff / u1));
    } else if (u >= A.u[last]) {
        omc = omc_at(last);
    } else {
        int32_t lo = first, hi = last;
        while (hi - lo > 1) { int32_t const m = (lo + hi) / 2; if (A.u[m] <= u) lo = m; else hi = m; }
        double const d0 = omc_at(lo);
        double const d1 = omc_at(hi);
        // hi == lo + 1, so log(u[hi]/u[lo]) is the precomputed node ratio.
        double const fu = log(u / A.u[lo]) / A.log_u_ratio[lo];
        omc = d0 * exp(fu * log(d1 / d0));
    }
    double const c = 1.0 - dmin(2.0, omc);
    return dmin(1.0, dmax(-1.0, c));
}

// ---------------------------------------------------------------------------
// Geometry -- NA/src/RreaTransportPath.cpp, non-periodic: one piece at most
// ---------------------------------------------------------------------------
struct Domain { double r_lo, r_hi, z_lo, z_hi; };

struct Path {
    double sx, sy, sz, ex, ey, ez;
    double full;      // full_length_m
    double end_f;     // in_domain_end_fraction (begin is always 0)
    double in_len;
    double piece_len; // length of the single piece (valid when has_piece)
    bool exits;
    bool has_piece;
};

// PointAtFullPathFraction :42-53
__device__ __forceinline__ void point_at_full(Path const& p, double f, double& x, double& y, double& z)
{
    f = dmax(0.0, dmin(1.0, f));
    x = p.sx + f * (p.ex - p.sx);
    y = p.sy + f * (p.ey - p.sy);
    z = p.sz + f * (p.ez - p.sz);
}

// PointAtPieceFraction :55-72 (the piece is [0, end_f], shift 0)
__device__ __forceinline__ void point_at_piece(Path const& p, Domain const& d, double f, double& x, double& y, double& z)
{
    f = dmax(0.0, dmin(p.end_f, f));
    x = p.sx + f * (p.ex - p.sx);
    y = p.sy + f * (p.ey - p.sy);
    z = p.sz + f * (p.ez - p.sz) + 0.0;
    z = dmax(d.z_lo, dmin(d.z_hi, z));
}

// ClipRzTransportPath :74-293, non-periodic branch.  Returns false for the
// reference's Abort cases.
__device__ bool clip_rz_path(double sr, double st, double sz, double er_, double et, double ez_, Domain const& d,
                             Path& p, int32_t& err)
{
    if (!finite(sr) || !finite(st) || !finite(sz) || !finite(er_) || !finite(et) || !finite(ez_) || sr < 0.0
        || er_ < 0.0) {
        if (!err) err = kErrInvalid;
        return false;
    }
    double s_st, c_st, s_et, c_et;
    sin_cos(st, s_st, c_st);
    sin_cos(et, s_et, c_et);
    p.sx = sr * c_st;
    p.sy = sr * s_st;
    p.sz = sz;
    p.ex = er_ * c_et;
    p.ey = er_ * s_et;
    p.ez = ez_;
    double const dx = p.ex - p.sx;
    double const dy = p.ey - p.sy;
    double const dz = p.ez - p.sz;
    p.full = sqrt(dx * dx + dy * dy + dz * dz);
    double const scale = dmax(1.0, dmax(d.r_hi, dmax(fabs(d.z_lo), fabs(d.z_hi))));
    double const tol = 64.0 * kEps * scale;
    if (sr < d.r_lo - tol || sr > d.r_hi + tol || sz < d.z_lo - tol || sz > d.z_hi + tol) {
        if (!err) err = kErrDomain;
        return false;
    }
    double exit_f = 1.0;
    auto retain = [&](double c) { if (c >= 0.0 && c < exit_f) exit_f = c; };
    if (dz > 0.0) retain((d.z_hi - sz) / dz);
    else if (dz < 0.0) retain((d.z_lo - sz) / dz);
    double const a = dx * dx + dy * dy;
    if (a > 0.0) {
        double const b = 2.0 * (p.sx * dx + p.sy * dy);
        double const c = p.sx * p.sx + p.sy * p.sy - d.r_hi * d.r_hi;
        double const disc = b * b - 4.0 * a * c;
        if (!(disc < 0.0)) {
            double const root = sqrt(dmax(disc, 0.0));
            double const inv_2a = 0.5 / a;
            double const roots[2] = {(-b - root) * inv_2a, (-b + root) * inv_2a};
            for (int k = 0; k < 2; ++k) {
                double const cand = roots[k];
                if (cand <= 0.0 || cand > 1.0) continue;
                if (2.0 * a * cand + b > 0.0) retain(cand);
            }
        }
    }
    p.end_f = dmax(0.0, dmin(1.0, exit_f));
    p.in_len = p.full * (p.end_f - 0.0);
    p.exits = p.end_f < 1.0 - 64.0 * kEps;
    double const b0 = 0.0, b1 = p.end_f;
    bool const merged = fabs(b0 - b1) <= 128.0 * kEps * dmax(1.0, dmax(fabs(b0), fabs(b1)));
    p.has_piece = !merged && b1 > b0;
    p.piece_len = p.full * (b1 - b0);
    return true;
}

// RreaChargedSpeedFromKineticEnergy OV/RreaParticleInteraction.H:92-105
__device__ __forceinline__ double charged_speed(double k)
{
    if (!finite(k) || k < 0.0) return __builtin_nan("");
    double const gamma = 1.0 + k / kMeC2;
    double const beta2 = dmax(0.0, 1.0 - 1.0 / (gamma * gamma));
    return kC * sqrt(beta2);
}

// RreaBuildRunningChargedChord OV/RreaParticleInteraction.H:110-161.  Returns
// the chord's validity; an invalid chord leaves p zero-initialised.
__device__ bool build_chord(double x, double y, double z, double dx, double dy, double dz, double k, double req_dt,
                            double outer_dt, Domain const& d, Path& p, int32_t& err)
{
    p = Path{};
    if (!finite(req_dt) || !finite(outer_dt) || !(req_dt >= 0.0) || !(outer_dt > 0.0)) return false;
    double const speed = charged_speed(k);
    double const norm = sqrt(dx * dx + dy * dy + dz * dz);
    if (!finite(speed) || !finite(norm) || !(norm > 0.0)) return false;
    double const scale = speed * req_dt / norm;
    double const ex = x + scale * dx;
    double const ey = y + scale * dy;
    double const ez = z + scale * dz;
    if (!clip_rz_path(hypot(x, y), atan2(y, x), z, hypot(ex, ey), atan2(ey, ex), ez, d, p, err)) return false;
    return finite(p.full);
}

// OutsideCertifiedDomain :28-40 + RreaClippedTransportEscapePolicyForPath
// OV/RreaParticleInteraction.H:409-452: escapes iff the chord exits or its
// end is outside the half-open domain; RREA then owns the invalidation.
__device__ __forceinline__ bool clipped_escapes(Path const& p, Domain const& d)
{
    double const r = hypot(p.ex, p.ey);
    double const z = p.ez;
    if (!finite(r) || !finite(z)) return true;
    bool const outside = r < d.r_lo || r >= d.r_hi || z < d.z_lo || z >= d.z_hi;
    return p.exits || outside;
}

// RreaCartesianPositionInHalfOpenRzDomain :318-354
__device__ __forceinline__ bool in_half_open(double x, double y, double z, Domain const& d)
{
    if (!finite(x) || !finite(y)) return false;
    double const r = hypot(x, y);
    return finite(r) && finite(z) && d.r_hi > d.r_lo && d.z_hi > d.z_lo && r >= d.r_lo && r < d.r_hi
        && z >= d.z_lo && z < d.z_hi;
}

struct Crossing { double start_r, start_z, cross_r, cross_z; uint32_t faces; bool valid; };

// RreaChargedBoundaryCrossingForPath OV/RreaParticleInteraction.H:468-540
__device__ Crossing boundary_crossing(Path const& p, Domain const& d)
{
    Crossing res;
    double cx, cy;
    point_at_full(p, p.end_f, cx, cy, res.cross_z);
    res.cross_r = hypot(cx, cy);
    res.start_r = hypot(p.sx, p.sy);
    res.start_z = p.sz;
    res.faces = 0U;
    double const scale = dmax(1.0, dmax(d.r_hi, dmax(fabs(d.z_lo), fabs(d.z_hi))));
    double const tol = 512.0 * kEps * scale;
    bool const on_r_hi = fabs(res.cross_r - d.r_hi) <= tol;
    bool const on_z_lo = fabs(res.cross_z - d.z_lo) <= tol;
    bool const on_z_hi = fabs(res.cross_z - d.z_hi) <= tol;
    double const dx = p.ex - p.sx, dy = p.ey - p.sy, dz = p.ez - p.sz;
    double const outward = cx * dx + cy * dy;
    if (on_r_hi && outward > -tol * scale) { res.faces |= 1U; res.cross_r = d.r_hi; }
    if (on_z_lo && dz < 0.0) { res.faces |= 2U; res.cross_z = d.z_lo; }
    if (on_z_hi && dz > 0.0) { res.faces |= 4U; res.cross_z = d.z_hi; }
    if (res.faces == 0U) {
        if (on_r_hi) { res.faces |= 1U; res.cross_r = d.r_hi; }
        if (on_z_hi) { res.faces |= 4U; res.cross_z = d.z_hi; }
    }
    res.valid = res.faces != 0U && finite(res.start_r) && finite(res.start_z) && finite(res.cross_r)
        && finite(res.cross_z);
    return res;
}

// ---------------------------------------------------------------------------
// Coupling services on the device: density law and field gather
// ---------------------------------------------------------------------------
struct DevFields {
    int32_t nr, nz;
    double dr, dz, r_lo, z_lo;
    const double* cell_er;   // nr * nz cell-centred total Er, c = i + nr*j
    const double* cell_ez;   // nr * nz cell-centred total Ez
    const double* btheta;    // (nr+1) * (nz+1)
    int32_t density_mode;    // 0 constant, 1 profile
    double density_ratio;
    const double* prof_alt;
    const double* prof_rho;
    int32_t prof_n;
    double altitude_at_z0;
    double reference_density;
    Guide prof_guide;        // over altitude
};

// TransportDensityRatioAtZ OV/RreaWarpXCoupling.cpp:79-95 +
// RreaAltitudeProfile::Interpolate NA/src/RreaAltitudeProfile.cpp:125-151
__device__ __forceinline__ double density_ratio_at_z(DevFields const& F, double z, int32_t& err)
{
    if (F.density_mode == 0) return F.density_ratio;
    double const alt = F.altitude_at_z0 + z;
    int32_t const n = F.prof_n;
    const double* a = F.prof_alt;
    if (!(alt >= a[0]) || !(alt <= a[n - 1])) { if (!err) err = kErrRange; return 1.0; }
    int32_t const first = guided_upper_bound(a, n, alt, F.prof_guide, alt);
    double v;
    if (first == 0) v = F.prof_rho[0];
    else if (first == n) v = F.prof_rho[n - 1];
    else {
        int32_t const lo = first - 1;
        double const span = a[first] - a[lo];
        double const t = span > 0.0 ? (alt - a[lo]) / span : 0.0;
        v = (1.0 - t) * F.prof_rho[lo] + t * F.prof_rho[first];
    }
    return v / F.reference_density;
}

// GatherTotalFieldRZ NA/src/RreaAmrexAdvance.cpp:532-579 (stencil
// NA/include/rrea/RreaFieldGatherStencil.H:49-130) and MaxwellBthetaAt :511-530.
// The cell values are the face averages 0.5*((m_lo + a_lo) + (m_hi + a_hi))
// formed once per field update.
__device__ __forceinline__ void gather(DevFields const& F, double r, double z, double& er, double& ez, double& bt)
{
    int32_t const nr = F.nr, nz = F.nz;
    double const fx = (r - F.r_lo) / F.dr + 0.0 - 0.5;
    double const fz = (z - F.z_lo) / F.dz + 0.0 - 0.5;
    int32_t const ir0 = static_cast<int32_t>(floor(fx));
    int32_t const iz0 = static_cast<int32_t>(floor(fz));
    double const tx = fx - static_cast<double>(ir0);
    double const tz = fz - static_cast<double>(iz0);
    int32_t const i0 = max(0, min(nr - 1, ir0));
    int32_t const i1 = max(0, min(nr - 1, ir0 + 1));
    int32_t const j0 = max(0, min(nz - 1, iz0));
    int32_t const j1 = max(0, min(nz - 1, iz0 + 1));
    double const first_center = F.r_lo + 0.5 * F.dr;
    double const axis = r >= first_center ? 1.0 : dmax(0.0, (r - F.r_lo) / (0.5 * F.dr));
    auto bilinear = [&](const double* c) {
        double const f00 = c[i0 + nr * j0], f01 = c[i0 + nr * j1], f10 = c[i1 + nr * j0], f11 = c[i1 + nr * j1];
        return (1.0 - tx) * ((1.0 - tz) * f00 + tz * f01) + tx * ((1.0 - tz) * f10 + tz * f11);
    };
    er = bilinear(F.cell_er) * axis;
    ez = bilinear(F.cell_ez);
    if (r <= 0.0) {
        bt = 0.0;
    } else {
        double frr = r / F.dr;
        frr = frr < 0.0 ? 0.0 : (frr > static_cast<double>(nr) ? static_cast<double>(nr) : frr);
        double fzz = (z - F.z_lo) / F.dz;
        fzz = fzz < 0.0 ? 0.0 : (fzz > static_cast<double>(nz) ? static_cast<double>(nz) : fzz);
        int32_t const a0 = min(static_cast<int32_t>(floor(frr)), nr - 1);
        int32_t const c0 = min(static_cast<int32_t>(floor(fzz)), nz - 1);
        int32_t const a1 = a0 + 1, c1 = c0 + 1;
        double const trr = frr - a0, tzz = fzz - c0;
        int32_t const w = nr + 1;
        bt = (1.0 - trr) * ((1.0 - tzz) * F.btheta[a0 + w * c0] + tzz * F.btheta[a0 + w * c1])
            + trr * ((1.0 - tzz) * F.btheta[a1 + w * c0] + tzz * F.btheta[a1 + w * c1]);
    }
}

// ---------------------------------------------------------------------------
// Material path (OV/RreaMaterialPath.H) on the single-piece chord
// ---------------------------------------------------------------------------
struct Interval { double end, length, ratio; bool valid; };  // begin is 0

__device__ __forceinline__ double positive_or(double v, double fb) { return v > 0.0 ? v : fb; }

// integrate_density_piece_m :60-84 + build_material_path_intervals :86-114
__device__ __forceinline__ Interval build_interval(DevFields const& F, Path const& p, Domain const& d, double fallback,
                                                   int32_t& err)
{
    Interval iv{0.0, 0.0, 1.0, false};
    if (!p.has_piece || !(p.piece_len > 0.0)) return iv;
    double dl = 0.0;
    double const b = 0.0, e = p.end_f;
    if (e > b && p.full > 0.0) {
        double x, y, z;
        point_at_piece(p, d, 0.5 * (b + e), x, y, z);
        dl = p.full * (e - b) * positive_or(density_ratio_at_z(F, z, err), fallback);
    }
    iv.end = p.end_f;
    iv.length = p.piece_len;
    iv.ratio = positive_or(dl / p.piece_len, fallback);
    iv.valid = true;
    return iv;
}

// density_path_m :117-137
__device__ __forceinline__ double density_path(Interval const& iv, Path const& p, double b, double e)
{
    b = dmax(0.0, b);
    e = dmin(p.end_f, e);
    if (!(e > b)) return 0.0;
    double s = 0.0;
    if (iv.valid) {
        double const lb = dmax(b, 0.0), le = dmin(e, iv.end);
        if (le > lb) s += p.full * (le - lb) * iv.ratio;
    }
    return s;
}

// fraction_at_density_path_m :143-172
__device__ __forceinline__ double fraction_at_density_path(Interval const& iv, Path const& p, double b, double e,
                                                           double target, int32_t& err)
{
    b = dmax(0.0, b);
    e = dmin(p.end_f, e);
    if (!finite(target) || target < 0.0) { if (!err) err = kErrInvalid; return e; }
    double rem = target;
    if (iv.valid) {
        double const lb = dmax(b, 0.0), le = dmin(e, iv.end);
        if (le > lb && iv.ratio > 0.0) {
            double const contrib = p.full * (le - lb) * iv.ratio;
            if (rem <= contrib) return lb + rem / (p.full * iv.ratio);
            rem -= contrib;
        }
    }
    return e;
}

// max_uniform_subcycle_density_path_m :177-201
__device__ __forceinline__ double max_uniform_slice(Interval const& iv, Path const& p, uint64_t n)
{
    double const span = p.end_f - 0.0;
    double const den = static_cast<double>(n);
    double m = 0.0;
    for (uint64_t i = 0; i < n; ++i) {
        m = dmax(m, density_path(iv, p, 0.0 + static_cast<double>(i) / den * span,
                                 0.0 + static_cast<double>(i + 1U) / den * span));
    }
    return m;
}

// ---------------------------------------------------------------------------
// Field work and kinematics
// ---------------------------------------------------------------------------
struct FieldRange { double work, ex, ey, ez, bx, by, bz, length; };

// path_field_work_and_average OV/RreaPathFieldWork.H:57-150 (GL2 per piece).
// With a `hold` pointer (kHold) the first node's six products wait in LDS
// (hold[c * kHoldStride], zero again on return) while the second node is
// gathered, so a caller short of registers does not spill them; the values
// and the arithmetic are the same.
template <bool kHold = false, int kHoldStride = 256>
__device__ FieldRange path_field_work(DevFields const& F, Path const& p, Domain const& d, double q, double b, double e,
                                      int32_t& err, double* hold = nullptr)
{
    FieldRange res{0, 0, 0, 0, 0, 0, 0, 0};
    if (!(p.full > 0.0)) return res;
    b = dmax(0.0, b);
    e = dmin(p.end_f, e);
    if (!(e > b) || !p.has_piece) return res;
    double const pdx = (p.ex - p.sx) / p.full;
    double const pdy = (p.ey - p.sy) / p.full;
    double const pdz = (p.ez - p.sz) / p.full;
    double const pb = dmax(b, 0.0), pe = dmin(e, p.end_f);
    if (!(pe > pb)) return res;
    double s[2][7];
    double const mid = 0.5 * (pb + pe);
    double const half = 0.5 * (pe - pb);
    double const off = half * kGl2Node;
    for (int node = 0; node < 2; ++node) {
        double x, y, z;
        point_at_piece(p, d, node == 0 ? mid - off : mid + off, x, y, z);
        double const r = hypot(x, y);
        double rad = 0.0, rc = 1.0, rs = 0.0;
        if (r > 0.0) { rad = (pdx * x + pdy * y) / r; rc = x / r; rs = y / r; }
        double er, ez, bt;
        gather(F, r, z, er, ez, bt);
        double const wk = q * (er * rad + ez * pdz) * 1.0;
        if (!finite(wk) || !finite(er) || !finite(ez) || !finite(bt)) { if (!err) err = kErrInvalid; }
        s[node][0] = wk; s[node][1] = er * rc; s[node][2] = er * rs; s[node][3] = ez;
        s[node][4] = -bt * rs; s[node][5] = bt * rc; s[node][6] = 0.0;
        if constexpr (kHold) {
            if (node == 0) {
                for (int c = 0; c < 6; ++c) hold[c * kHoldStride] = s[0][c];
            }
        }
    }
    if constexpr (kHold) {
        // relaxed LDS loads, which the compiler does not fold into the stores
        for (int c = 0; c < 6; ++c) {
            s[0][c] = __hip_atomic_load(hold + c * kHoldStride, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            hold[c * kHoldStride] = 0.0;
        }
    }
    double integ[7];
    for (int c = 0; c < 7; ++c) integ[c] = half * (s[0][c] + s[1][c]);
    res.work += p.full * integ[0];
    res.ex += p.full * integ[1];
    res.ey += p.full * integ[2];
    res.ez += p.full * integ[3];
    res.bx += p.full * integ[4];
    res.by += p.full * integ[5];
    res.bz += p.full * integ[6];
    res.length += p.full * (pe - pb);
    if (res.length > 0.0) {
        res.ex /= res.length; res.ey /= res.length; res.ez /= res.length;
        res.bx /= res.length; res.by /= res.length; res.bz /= res.length;
    }
    return res;
}

// RreaElectronProperVelocityFromKineticEv NA/include/rrea/RreaSeedSchedule.H:30-43
__device__ __forceinline__ double proper_velocity(double k)
{
    if (k <= 0.0) return 0.0;
    double const gamma = 1.0 + k / kMeC2;
    return kC * sqrt(dmax(gamma * gamma - 1.0, 0.0));
}

// RreaKineticEnergyEvFromProperVelocity OV/RreaParticleInteraction.H:1494-1506
__device__ __forceinline__ double kinetic_from_u(double ux, double uy, double uz)
{
    double const u2 = ux * ux + uy * uy + uz * uz;
    double const gamma = sqrt(1.0 + u2 / (kC * kC));
    return u2 * kMeC2 / ((kC * kC) * (gamma + 1.0));
}

// RreaNormalizeDirection OV/RreaChargedInterleave.H:294-307
__device__ __forceinline__ void normalize(double& x, double& y, double& z)
{
    double const n = sqrt(x * x + y * y + z * z);
    if (n <= 0.0) { x = 0.0; y = 0.0; z = 1.0; return; }
    x /= n; y /= n; z /= n;
}

// RreaApplyChargedFieldHalfKick OV/RreaChargedInterleave.H:100-138 with
// RreaBorisMagneticRotate NA/include/rrea/RreaLorentzPush.H:21-46
__device__ __forceinline__ double half_kick(FieldRange const& f, double half_dt, double q, double& k, double& dx,
                                            double& dy, double& dz)
{
    double const umag = proper_velocity(k);
    double const impulse = q * (kC * kC / kMeC2) * half_dt;
    double ux = umag * dx + impulse * f.ex;
    double uy = umag * dy + impulse * f.ey;
    double uz = umag * dz + impulse * f.ez;
    {
        double const gamma = sqrt(1.0 + (ux * ux + uy * uy + uz * uz) / (kC * kC));
        double const tx = impulse * f.bx / (2.0 * gamma);
        double const ty = impulse * f.by / (2.0 * gamma);
        double const tz = impulse * f.bz / (2.0 * gamma);
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
    double const n = sqrt(ux * ux + uy * uy + uz * uz);
    if (n > 0.0 && finite(n)) { dx = ux / n; dy = uy / n; dz = uz / n; }
    double const hw = 0.5 * f.work;
    k = dmax(0.0, k + hw);
    return hw;
}

// RreaTransverseBasis + RreaSampleConeDirection OV/RreaChargedInterleave.H:309-364
__device__ __forceinline__ void cone_direction(double ax, double ay, double az, double ct, double uphi, double& ox,
                                               double& oy, double& oz)
{
    normalize(ax, ay, az);
    double dx = ax, dy = ay, dz = az;
    normalize(dx, dy, dz);
    double rx = 0.0, ry = 0.0, rz = 1.0;
    if (fabs(dz) > 0.9) { rx = 1.0; rz = 0.0; }
    double b1x = ry * dz - rz * dy;
    double b1y = rz * dx - rx * dz;
    double b1z = rx * dy - ry * dx;
    normalize(b1x, b1y, b1z);
    double b2x = dy * b1z - dz * b1y;
    double b2y = dz * b1x - dx * b1z;
    double b2z = dx * b1y - dy * b1x;
    normalize(b2x, b2y, b2z);
    ct = dmin(1.0, dmax(-1.0, ct));
    double const st = sqrt(dmax(0.0, 1.0 - ct * ct));
    double const phi = kTwoPi * uphi;
    double sp, cp;
    sin_cos(phi, sp, cp);
    ox = ax * ct + st * (cp * b1x + sp * b2x);
    oy = ay * ct + st * (cp * b1y + sp * b2y);
    oz = az * ct + st * (cp * b1z + sp * b2z);
    normalize(ox, oy, oz);
}

// RreaPoissonSampleCapped :371-383
__device__ __forceinline__ int poisson_capped(double mean, double u, int kmax)
{
    if (mean <= 0.0) return 0;
    double p = exp(-mean);
    double cum = p;
    int k = 0;
    while (u > cum && k < kmax) { ++k; p *= mean / static_cast<double>(k); cum += p; }
    return k;
}

__device__ __forceinline__ double unit_cycle(double u)
{
    u -= floor(u);
    return u < 0.0 ? u + 1.0 : u;
}

// Moller two-body lab cosines :391-423
__device__ __forceinline__ double moller_cos(double kp, double ks_or_scattered)
{
    double const v = ks_or_scattered * (kp + 2.0 * kMeC2) / (kp * (ks_or_scattered + 2.0 * kMeC2));
    return sqrt(dmin(1.0, dmax(0.0, v)));
}

// RreaSelectCompetingHazard OV/RreaParticleInteraction.H:1402-1454 (2 channels)
struct Hazard { bool valid, event; int channel; double sampled; };

__device__ __forceinline__ Hazard select_hazard(double t0, double t1, double u1, double u2)
{
    Hazard h{false, false, -1, 0.0};
    if (!finite(u1) || !finite(u2)) return h;
    if (!finite(t0) || t0 < 0.0 || !finite(t1) || t1 < 0.0) return h;
    double total = 0.0;
    total += t0;
    total += t1;
    h.valid = finite(total);
    if (!h.valid) return h;
    u1 = dmax(0.0, dmin(1.0, u1));
    h.sampled = -log(dmax(1.0e-30, 1.0 - u1));
    if (!(h.sampled < total)) return h;
    h.event = true;
    u2 = dmax(0.0, dmin(1.0, u2));
    double const target = u2 * total;
    double cum = 0.0;
    int last_open = -1;
    if (t0 > 0.0) last_open = 0;
    cum += t0;
    if (target < cum) { h.channel = 0; return h; }
    if (t1 > 0.0) last_open = 1;
    cum += t1;
    if (target < cum) { h.channel = 1; return h; }
    h.channel = last_open;
    h.valid = h.channel >= 0;
    h.event = h.valid;
    return h;
}

// The same selection over three channels (the positron's Bhabha, brems,
// annihilation); the running sums are formed in the same order.
__device__ __forceinline__ Hazard select_hazard3(double t0, double t1, double t2, double u1, double u2)
{
    Hazard h{false, false, -1, 0.0};
    if (!finite(u1) || !finite(u2)) return h;
    if (!finite(t0) || t0 < 0.0 || !finite(t1) || t1 < 0.0 || !finite(t2) || t2 < 0.0) return h;
    double total = 0.0;
    total += t0;
    total += t1;
    total += t2;
    h.valid = finite(total);
    if (!h.valid) return h;
    u1 = dmax(0.0, dmin(1.0, u1));
    h.sampled = -log(dmax(1.0e-30, 1.0 - u1));
    if (!(h.sampled < total)) return h;
    h.event = true;
    u2 = dmax(0.0, dmin(1.0, u2));
    double const target = u2 * total;
    double cum = 0.0;
    int last_open = -1;
    if (t0 > 0.0) last_open = 0;
    cum += t0;
    if (target < cum) { h.channel = 0; return h; }
    if (t1 > 0.0) last_open = 1;
    cum += t1;
    if (target < cum) { h.channel = 1; return h; }
    if (t2 > 0.0) last_open = 2;
    cum += t2;
    if (target < cum) { h.channel = 2; return h; }
    h.channel = last_open;
    h.valid = h.channel >= 0;
    h.event = h.valid;
    return h;
}

// ---------------------------------------------------------------------------
// Two-photon annihilation -- NA/src/RreaInteractionTables.cpp:43-108, 3208-3253
// ---------------------------------------------------------------------------
__device__ __forceinline__ double clamp_unit(double v)
{
    if (!finite(v)) return 0.5;
    return dmax(0.0, dmin(1.0, v));
}

__device__ __forceinline__ double wrap_unit_interval(double v)
{
    if (!finite(v)) return 0.0;
    v = v - floor(v);
    if (v < 0.0) v += 1.0;
    return v;
}

// Heitler photon-energy fraction: 64 bisections of the antiderivative
// a ln e - e + 1/((g+1)^2 e) over the centre-of-mass support.
__device__ double annihilation_energy_fraction(double gamma, double draw)
{
    if (gamma <= 1.0 + 1.0e-12) return 0.5;
    double const beta_cm = sqrt(dmax(0.0, (gamma - 1.0) / (gamma + 1.0)));
    double const eps_min = 0.5 * (1.0 - beta_cm);
    double const eps_max = 0.5 * (1.0 + beta_cm);
    double const u = clamp_unit(draw);
    double const gp1 = gamma + 1.0;
    double const a = 1.0 + 2.0 * gamma / (gp1 * gp1);
    double const f_lo = a * log(eps_min) - eps_min + 1.0 / (gp1 * gp1 * eps_min);
    double const f_hi = a * log(eps_max) - eps_max + 1.0 / (gp1 * gp1 * eps_max);
    double const span = f_hi - f_lo;
    if (!(span > 0.0)) return 0.5;
    double const target = f_lo + u * span;
    double lo = eps_min, hi = eps_max;
    for (int it = 0; it < 64; ++it) {
        double const mid = 0.5 * (lo + hi);
        if (a * log(mid) - mid + 1.0 / (gp1 * gp1 * mid) < target) lo = mid; else hi = mid;
    }
    return dmax(eps_min, dmin(eps_max, 0.5 * (lo + hi)));
}

struct Annihilation { double e1, e2, cos1, az1; bool valid, at_rest; };

__device__ __forceinline__ Annihilation sample_annihilation(double k, double draw_energy, double draw_azimuth)
{
    Annihilation s{0.0, 0.0, 1.0, 0.0, false, false};
    k = dmax(0.0, k);
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
    double const beta = sqrt(dmax(0.0, 1.0 - 1.0 / (gamma * gamma)));
    double const eps = annihilation_energy_fraction(gamma, draw_energy);
    s.e1 = eps * total;
    s.e2 = dmax(0.0, total - s.e1);
    if (s.e1 <= 0.0 || s.e2 <= 0.0) return s;
    double const denom = eps * gamma * beta;
    if (denom <= 0.0) return s;
    s.cos1 = dmax(-1.0, dmin(1.0, (eps * (gamma + 1.0) - 1.0) / denom));
    s.az1 = wrap_unit_interval(draw_azimuth);
    s.valid = true;
    s.at_rest = false;
    return s;
}

// ---------------------------------------------------------------------------
// Net-chord deposition -- NA/include/rrea/RreaSegmentDepositionRZ.H
// ---------------------------------------------------------------------------
struct DevGrid { int32_t nr, nz; double dr, dz, r_lo, z_lo; double* rho; double* jr; double* jz; };

__device__ __forceinline__ void atomic_add(double* p, double v) { unsafeAtomicAdd(p, v); }

// DepositChargeCIC :68-90
__device__ void deposit_cic(DevGrid const& g, double q, double r, double z)
{
    double const xr = (r - g.r_lo) / g.dr - 0.5;
    double const xz = (z - g.z_lo) / g.dz - 0.5;
    int32_t const i0 = static_cast<int32_t>(floor(xr));
    int32_t const j0 = static_cast<int32_t>(floor(xz));
    double const fr = xr - i0, fz = xz - j0;
    double const wr[2] = {1.0 - fr, fr};
    double const wz[2] = {1.0 - fz, fz};
    for (int a = 0; a < 2; ++a) {
        for (int b = 0; b < 2; ++b) {
            int32_t i = i0 + a, j = j0 + b;
            if (i < 0) i = -1 - i;
            if (i >= g.nr) i = g.nr - 1;
            if (i < 0) i = 0;  // only reachable for a non-finite position
            if (j < 0) j = 0;
            if (j >= g.nz) j = g.nz - 1;
            atomic_add(&g.rho[i + g.nr * j], q * wr[a] * wz[b]);
        }
    }
}

// Where the face currents of a deposit go: straight to the grid with device
// atomics, or into a workgroup's LDS window of it (LdsWindow below).
struct GlobalSink {
    DevGrid const& g;
    __device__ void jr(int32_t i, int32_t j, double v) const { atomic_add(&g.jr[i + (g.nr + 1) * j], v); }
    __device__ void jz(int32_t i, int32_t j, double v) const { atomic_add(&g.jz[i + g.nr * j], v); }
};

// detail::deposit_piece :96-152
template <class Sink>
__device__ void deposit_piece(DevGrid const& g, Sink const& sink, double qdt, double xr0, double xz0, double xr1,
                              double xz1)
{
    double fr_flux = qdt * (xr1 - xr0);
    double const fz_flux = qdt * (xz1 - xz0);
    double const mr = 0.5 * (xr0 + xr1);
    double const mz = 0.5 * (xz0 + xz1);
    {
        int32_t const jf = static_cast<int32_t>(floor(mz));
        double const fz = mz - jf;
        int32_t const fi = static_cast<int32_t>(floor(mr)) + 1;
        if (fi <= 0) fr_flux = 0.0;
        double const wz0 = 1.0 - fz;
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
        int32_t const ifc = static_cast<int32_t>(floor(mr));
        double const fr = mr - ifc;
        int32_t fj = static_cast<int32_t>(floor(mz)) + 1;
        double const wr0 = 1.0 - fr;
        if (fj < 0) fj = 0;
        if (fj > g.nz) fj = g.nz;
        for (int a = 0; a < 2; ++a) {
            int32_t i = ifc + a;
            if (i < 0) i = -1 - i;
            if (i >= g.nr) i = g.nr - 1;
            if (i < 0) i = 0;  // only reachable for a non-finite position
            sink.jz(i, fj, fz_flux * (a == 0 ? wr0 : fr));
        }
    }
}

// The cell-boundary crossings of one axis, generated in increasing s: the
// same values (k - a) / (b - a) the reference collects and sorts.
struct CutStream {
    double a, b, k, k_end, dk;
    bool active;
    __device__ void init(double a_, double b_)
    {
        a = a_;
        b = b_;
        active = a != b;
        if (!active) return;
        double const lo = dmin(a, b), hi = dmax(a, b);
        if (b > a) { k = ceil(lo); k_end = floor(hi); dk = 1.0; }
        else { k = floor(hi); k_end = ceil(lo); dk = -1.0; }
        advance_to_valid();
    }
    __device__ bool in_range() const { return active && (dk > 0.0 ? k <= k_end : k >= k_end); }
    __device__ double value() const { return (k - a) / (b - a); }
    __device__ void advance_to_valid()
    {
        while (in_range()) {
            double const s = value();
            if (s > 0.0 && s < 1.0) return;
            k += dk;
        }
    }
    __device__ double peek() const { return in_range() ? value() : 2.0; }
    __device__ void pop() { k += dk; advance_to_valid(); }
};

// DepositSegmentRZ :159-192, merging the two sorted cut streams in place of
// the reference's collect-and-sort.
template <class Sink>
__device__ bool deposit_segment_to(DevGrid const& g, Sink const& sink, double q, double dt, double r0, double z0,
                                   double r1, double z1)
{
    double const qdt = q / dt;
    double const xr0 = (r0 - g.r_lo) / g.dr - 0.5;
    double const xz0 = (z0 - g.z_lo) / g.dz - 0.5;
    double const xr1 = (r1 - g.r_lo) / g.dr - 0.5;
    double const xz1 = (z1 - g.z_lo) / g.dz - 0.5;
    // Both ends lie inside the padded grid for every segment the step makes;
    // anything else (a non-finite or runaway coordinate) is refused so that
    // the crossing walk below is always bounded by the grid size.
    double const lim_r = static_cast<double>(g.nr) + 2.0, lim_z = static_cast<double>(g.nz) + 2.0;
    if (!(fabs(xr0) <= lim_r && fabs(xr1) <= lim_r && fabs(xz0) <= lim_z && fabs(xz1) <= lim_z)) return false;
    CutStream cr, cz;
    cr.init(xr0, xr1);
    cz.init(xz0, xz1);
    double s0 = 0.0;
    for (;;) {
        double const a = cr.peek(), b = cz.peek();
        double s1;
        if (a <= b && a < 2.0) { s1 = a; cr.pop(); }
        else if (b < 2.0) { s1 = b; cz.pop(); }
        else { s1 = 1.0; }
        if (s1 > s0) {
            deposit_piece(g, sink, qdt, xr0 + s0 * (xr1 - xr0), xz0 + s0 * (xz1 - xz0), xr0 + s1 * (xr1 - xr0),
                          xz0 + s1 * (xz1 - xz0));
        }
        s0 = s1;
        if (a >= 2.0 && b >= 2.0) break;
    }
    return true;
}

__device__ __forceinline__ bool deposit_segment(DevGrid const& g, double q, double dt, double r0, double z0, double r1,
                                                double z1)
{
    return deposit_segment_to(g, GlobalSink{g}, q, dt, r0, z0, r1, z1);
}

// A workgroup's window of the face grids in LDS: faces (i, j) with
// i0 <= i < i0 + wi and j0 <= j < j0 + wj, for jr and for jz.  Every face a
// segment between two points can touch lies within [floor(x) - 1, floor(x) + 2]
// of its end points' index coordinates (clamped to the grid), which is how
// the window is sized.
struct LdsWindow {
    double* jr_w;
    double* jz_w;
    int32_t i0, j0, wi;
    __device__ void jr(int32_t i, int32_t j, double v) const { atomicAdd(&jr_w[(i - i0) + wi * (j - j0)], v); }
    __device__ void jz(int32_t i, int32_t j, double v) const { atomicAdd(&jz_w[(i - i0) + wi * (j - j0)], v); }
};

// The index-space face range [lo, hi] one segment end point can reach.
__device__ __forceinline__ void segment_face_range(DevGrid const& g, double r, double z, int32_t& ilo, int32_t& ihi,
                                                   int32_t& jlo, int32_t& jhi)
{
    double const xr = (r - g.r_lo) / g.dr - 0.5;
    double const xz = (z - g.z_lo) / g.dz - 0.5;
    int32_t const fi = static_cast<int32_t>(floor(xr));
    int32_t const fj = static_cast<int32_t>(floor(xz));
    ilo = fi - 1 < 0 ? 0 : fi - 1;
    ihi = fi + 2 > g.nr ? g.nr : fi + 2;
    jlo = fj - 1 < 0 ? 0 : fj - 1;
    jhi = fj + 2 > g.nz ? g.nz : fj + 2;
}

// The range test of deposit_segment alone, so that a caller can decide
// before committing anything.
__device__ __forceinline__ bool deposit_segment_check(DevGrid const& g, double r0, double z0, double r1, double z1)
{
    double const xr0 = (r0 - g.r_lo) / g.dr - 0.5;
    double const xz0 = (z0 - g.z_lo) / g.dz - 0.5;
    double const xr1 = (r1 - g.r_lo) / g.dr - 0.5;
    double const xz1 = (z1 - g.z_lo) / g.dz - 0.5;
    double const lim_r = static_cast<double>(g.nr) + 2.0, lim_z = static_cast<double>(g.nz) + 2.0;
    return fabs(xr0) <= lim_r && fabs(xr1) <= lim_r && fabs(xz0) <= lim_z && fabs(xz1) <= lim_z;
}

// DepositEscapingSegmentRZ :229-389
__device__ bool deposit_escaping_segment(DevGrid const& g, double q, double dt, double sr, double sz, double fr_,
                                         double fz_, double cr, double cz, int32_t pnr, int32_t pzb, int32_t pnz,
                                         uint32_t faces)
{
    if (!(dt > 0.0) || !finite(q) || !finite(dt) || !finite(sr) || !finite(sz) || !finite(fr_) || !finite(fz_)
        || !finite(cr) || !finite(cz) || pnr <= 0 || pnr >= g.nr || pnz <= 0 || pzb <= 0 || pzb + pnz >= g.nz
        || faces == 0U || (faces & ~7U) != 0U || ((faces & 2U) != 0U && (faces & 4U) != 0U)) {
        return false;
    }
    if (faces & 1U) cr = g.r_lo + pnr * g.dr;
    if (faces & 2U) cz = g.z_lo + pzb * g.dz;
    else if (faces & 4U) cz = g.z_lo + (pzb + pnz) * g.dz;
    int32_t mfrom[8], mto[8], mface[8], morient[8];
    double mcharge[8];
    bool mradial[8];
    int nm = 0;
    double const xr = (cr - g.r_lo) / g.dr - 0.5;
    double const xz = (cz - g.z_lo) / g.dz - 0.5;
    int32_t const i0 = static_cast<int32_t>(floor(xr));
    int32_t const j0 = static_cast<int32_t>(floor(xz));
    double const frc = xr - i0, fzc = xz - j0;
    double const wr[2] = {1.0 - frc, frc};
    double const wz[2] = {1.0 - fzc, fzc};
    for (int a = 0; a < 2; ++a) {
        for (int b = 0; b < 2; ++b) {
            int32_t i = i0 + a, j = j0 + b;
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
            int32_t const from = i + g.nr * j;
            for (int c = 0; c < cc; ++c) {
                if (nm >= 8) return false;
                if (cf[c] == 1U) {
                    mfrom[nm] = from; mto[nm] = (i + 1) + g.nr * j; mface[nm] = pnr + (g.nr + 1) * j;
                    mradial[nm] = true; morient[nm] = 1;
                } else {
                    int32_t const face_j = cf[c] == 2U ? pzb : pzb + pnz;
                    int32_t const to_j = cf[c] == 2U ? j - 1 : j + 1;
                    mfrom[nm] = from; mto[nm] = i + g.nr * to_j; mface[nm] = i + g.nr * face_j;
                    mradial[nm] = false; morient[nm] = cf[c] == 2U ? -1 : 1;
                }
                mcharge[nm] = share;
                ++nm;
            }
        }
    }
    if (nm == 0) return false;
    deposit_cic(g, -q, sr, sz);
    deposit_cic(g, q, cr, cz);
    if (!deposit_segment(g, q, dt, sr, sz, fr_, fz_) || !deposit_segment(g, q, dt, fr_, fz_, cr, cz)) return false;
    for (int m = 0; m < nm; ++m) {
        double const flux = morient[m] * mcharge[m] / dt;
        if (mradial[m]) atomic_add(&g.jr[mface[m]], flux); else atomic_add(&g.jz[mface[m]], flux);
        atomic_add(&g.rho[mfrom[m]], -mcharge[m]);
        atomic_add(&g.rho[mto[m]], mcharge[m]);
    }
    return true;
}

// ---------------------------------------------------------------------------
// Photon samplers -- NA/src/RreaInteractionTables.cpp:2272-2296, 2762-3127,
// OV/RreaChargedInterleave.H:309-338, 498-517
// ---------------------------------------------------------------------------
enum PhotonChannel : int { kCompton = 0, kPhotoelectric = 1, kPairNuclear = 2, kPairTriplet = 3 };

// RreaUniformStream NA/include/rrea/RreaUniformStream.H:21-53: a splitmix64
// sequence seeded from the bit patterns of two uniforms
struct UniformStream {
    uint64_t state;
    __device__ UniformStream(double u1, double u2)
    {
        uint64_t const b1 = static_cast<uint64_t>(__double_as_longlong(u1));
        uint64_t const b2 = static_cast<uint64_t>(__double_as_longlong(u2));
        uint64_t const m2 = splitmix64(b2);
        state = splitmix64(b1) ^ ((m2 << 32) | (m2 >> 32));
    }
    __device__ double next()
    {
        state += 0x9E3779B97F4A7C15ULL;
        uint64_t z = state;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        z = z ^ (z >> 31);
        return static_cast<double>(z >> 11) * (1.0 / 9007199254740992.0);
    }
};

// RreaTransverseBasis OV/RreaChargedInterleave.H:309-338
__device__ __forceinline__ void transverse_basis(double dx, double dy, double dz, double& b1x, double& b1y,
                                                 double& b1z, double& b2x, double& b2y, double& b2z)
{
    normalize(dx, dy, dz);
    double rx = 0.0, ry = 0.0, rz = 1.0;
    if (fabs(dz) > 0.9) { rx = 1.0; rz = 0.0; }
    b1x = ry * dz - rz * dy;
    b1y = rz * dx - rx * dz;
    b1z = rx * dy - ry * dx;
    normalize(b1x, b1y, b1z);
    b2x = dy * b1z - dz * b1y;
    b2y = dz * b1x - dx * b1z;
    b2z = dx * b1y - dy * b1x;
    normalize(b2x, b2y, b2z);
}

// RreaRotateLocalDirectionToAxis OV/RreaChargedInterleave.H:498-517
__device__ __forceinline__ void rotate_to_axis(double ax, double ay, double az, double lx, double ly, double lz,
                                               double& ox, double& oy, double& oz)
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

struct PhotonRates { double compton, pe, pair_n, pair_t; };

// PhotonProcessInverseLengths :2272-2296 at density ratio 1
__device__ __forceinline__ bool photon_rates(DevTables const& T, double e, PhotonRates& r, int32_t& err)
{
    Bracket b;
    if (!(finite(e) && e >= T.photon_min && e <= T.photon_max) || !photon_bracket(T, e, b)) {
        if (!err) err = kErrRange;
        r = PhotonRates{0.0, 0.0, 0.0, 0.0};
        return false;
    }
    r.compton = photon_value(T, 0, b, err) * 1.0;
    r.pe = photon_value(T, 1, b, err) * 1.0;
    r.pair_n = photon_value(T, 2, b, err) * 1.0;
    r.pair_t = photon_value(T, 3, b, err) * 1.0;
    return true;
}

// SampleComptonKleinNishinaFinalState :2806-2871 (Butcher-Messel rejection;
// round 0 on the two drawn uniforms, later draws from their uniform stream)
__device__ void sample_compton(double e, double draw_mu, double draw_accept, double& scattered, double& recoil,
                               double& cos_theta)
{
    draw_mu = dmax(0.0, dmin(1.0, draw_mu));
    draw_accept = dmax(0.0, dmin(1.0, draw_accept));
    UniformStream stream(draw_mu, draw_accept);
    double const alpha = e / kMeC2;
    double const eps0 = 1.0 / (1.0 + 2.0 * alpha);
    double const eps0sq = eps0 * eps0;
    double const a1 = -log(eps0);
    double const a2 = 0.5 * (1.0 - eps0sq);
    double epsilon = eps0;
    double onecost = 0.0;
    for (int iter = 0; iter < 128; ++iter) {
        double const u1 = iter == 0 ? draw_mu : stream.next();
        double const u2 = iter == 0 ? draw_accept : stream.next();
        double epsilonsq;
        if (a1 > (a1 + a2) * u1) {
            epsilon = exp(-a1 * u2);
            epsilonsq = epsilon * epsilon;
        } else {
            epsilonsq = eps0sq + (1.0 - eps0sq) * u2;
            epsilon = sqrt(epsilonsq);
        }
        onecost = (1.0 - epsilon) / (epsilon * alpha);
        double const sint2 = onecost * (2.0 - onecost);
        double const greject = 1.0 - epsilon * sint2 / (1.0 + epsilonsq);
        if (greject >= stream.next()) break;
    }
    double const mu = dmax(-1.0, dmin(1.0, 1.0 - onecost));
    scattered = dmax(0.0, epsilon * e);
    recoil = dmax(0.0, e - scattered);
    cos_theta = mu;
}

// SamplePhotoelectricFinalState :2878-2950: the subshell row and its binding
__device__ bool sample_photoelectric(DevTables const& T, double e, double u, double& binding, int32_t& err)
{
    if (T.pe_n <= 0 || !(e >= T.pe_primary[0] && e <= T.pe_primary[T.pe_n - 1])) {
        if (!err) err = kErrRange;
        return false;
    }
    u = dmax(0.0, dmin(1.0, u));
    int32_t const up = lower_bound(T.pe_primary, T.pe_n, e);
    int32_t g;
    if (up == 0) g = 0;
    else if (up == T.pe_n) g = T.pe_n - 1;
    else if (T.pe_primary[up] == e) g = up;  // the low side at a duplicated edge
    else g = up - 1;
    int32_t const s = T.pe_start[g], n = T.pe_start[g + 1] - s;
    int32_t row = lower_bound(T.pe_u + s, n, u);
    if (row == n) row = n - 1;
    binding = T.pe_binding[s + row];
    if (binding > e) { if (!err) err = kErrInvalid; return false; }
    return true;
}

// SamplePhotoelectronCosTheta :2957-3044: the Sauter-Gavrila K-shell polar
// angle by bisection of its analytic primitive in t = 1 - beta mu.  The
// reference bisects in long double; the device has no wider type, so it
// bisects in double to the same stopping rule (agreement to a few ulp of mu).
__device__ double sample_photoelectron_cos(double ke, double draw)
{
    double const u = clamp_unit(draw);
    if (u <= 0.0) return -1.0;
    if (u >= 1.0) return 1.0;
    double const gamma = 1.0 + ke / kMeC2;
    double const inv_gamma_sq = 1.0 / (gamma * gamma);
    double const beta = sqrt(dmax(0.0, 1.0 - inv_gamma_sq));
    if (beta < 1.0e-4) {
        double lo = -1.0, hi = 1.0;
        for (int it = 0; it < 96; ++it) {
            double const mid = 0.5 * (lo + hi);
            if (mid == lo || mid == hi) break;
            double const cdf = (2.0 + mid * (3.0 - mid * mid)) / 4.0;
            if (cdf < u) lo = mid; else hi = mid;
        }
        return 0.5 * (lo + hi);
    }
    double const b = 0.5 * gamma * (gamma - 1.0) * (gamma - 2.0);
    double const c = -inv_gamma_sq;
    double const t_lo = inv_gamma_sq / (1.0 + beta);
    double const t_hi = 1.0 + beta;
    auto primitive = [c, b](double x) {
        return -c / (3.0 * x * x * x) - (2.0 + b * c) / (2.0 * x * x) - (2.0 * b - 1.0) / x - b * log(x);
    };
    double const f_lo = primitive(t_lo);
    double const f_hi = primitive(t_hi);
    double const target = (1.0 - u) * f_hi + u * f_lo;
    double lo = t_lo, hi = t_hi;
    for (int it = 0; it < 128; ++it) {
        double const mid = 0.5 * (lo + hi);
        if (mid == lo || mid == hi) break;
        if (primitive(mid) < target) lo = mid; else hi = mid;
    }
    double const tt = 0.5 * (lo + hi);
    return dmax(-1.0, dmin(1.0, (1.0 - tt) / beta));
}

struct PairState {
    double electron, positron, recoil, local;
    double edx, edy, edz, pdx, pdy, pdz, rdx, rdy, rdz;
    bool valid, has_recoil;
};

__device__ __forceinline__ void normalize_local(double& x, double& y, double& z)
{
    double const n = sqrt(x * x + y * y + z * z);
    if (n <= 0.0 || !finite(n)) { x = 0.0; y = 0.0; z = 1.0; return; }
    x /= n; y /= n; z /= n;
}

// SamplePairFinalState :3053-3127 + BuildRreaPairEmpiricalFinalState
// :1384-1525: replay of one training record of the bracketing energy block,
// rotated about the pair axis, with the record's exact energy closure
__device__ PairState sample_pair(DevTables const& T, double e, int channel, double draw_invariant, double draw_block,
                                 double draw_azimuth, int32_t& err)
{
    PairState st{};
    st.valid = false;
    bool const triplet = channel == kPairTriplet;
    double const threshold = triplet ? 4.0 * kMeC2 : 2.0 * kMeC2;
    if (!(finite(e) && e > threshold)) return st;
    int32_t const code = triplet ? 2 : 1;
    int32_t first = -1, last = -1;
    for (int32_t i = 0; i < T.pair_nblocks; ++i) {
        if (T.pair_channel[i] == code) { if (first < 0) first = i; last = i; }
    }
    if (first < 0) { if (!err) err = kErrInvalid; return st; }
    int32_t sel = first;
    if (e >= T.pair_energy[last]) {
        sel = last;
    } else if (e > T.pair_energy[first]) {
        int32_t up = first;
        while (T.pair_energy[up] < e) ++up;  // blocks of one channel are contiguous and increasing
        if (T.pair_energy[up] == e) {
            sel = up;
        } else {
            double const lo_e = T.pair_energy[up - 1], hi_e = T.pair_energy[up];
            double const tt = (log(e) - log(lo_e)) / (log(hi_e) - log(lo_e));
            sel = clamp_unit(draw_block) < tt ? up : up - 1;
        }
    }
    int32_t const size = T.pair_start[sel + 1] - T.pair_start[sel];
    if (size <= 0) { if (!err) err = kErrInvalid; return st; }
    double const scaled = clamp_unit(draw_invariant) * static_cast<double>(size);
    int32_t idx = scaled < static_cast<double>(size - 1) ? static_cast<int32_t>(scaled) : size - 1;
    int64_t const rec = static_cast<int64_t>(T.pair_start[sel]) + idx;
    const double* f = T.pair_frac + 3 * rec;
    const double* d = T.pair_dir + 9 * rec;
    double const available = e - 2.0 * kMeC2;
    double const rotation = kTwoPi * clamp_unit(draw_azimuth);
    double sr, cr;
    sin_cos(rotation, sr, cr);
    auto rotated = [&](const double* v, double& x, double& y, double& z) {
        x = cr * v[0] - sr * v[1];
        y = sr * v[0] + cr * v[1];
        z = v[2];
        normalize_local(x, y, z);
    };
    st.valid = true;
    st.has_recoil = false;
    st.electron = available * f[0];
    st.positron = available * f[1];
    st.recoil = 0.0;
    rotated(d + 0, st.edx, st.edy, st.edz);
    rotated(d + 3, st.pdx, st.pdy, st.pdz);
    st.rdx = 0.0; st.rdy = 0.0; st.rdz = 0.0;
    if (triplet) {
        st.has_recoil = true;
        st.recoil = available * f[2];
        rotated(d + 6, st.rdx, st.rdy, st.rdz);
    }
    double const tol = 64.0 * kEps * dmax(1.0, available);
    double expl = st.electron + st.positron + st.recoil;
    if (!finite(expl) || st.electron < 0.0 || st.positron < 0.0 || st.recoil < 0.0) {
        if (!err) err = kErrInvalid;
        return st;
    }
    if (expl > available) {
        if (expl - available > tol) { if (!err) err = kErrInvalid; return st; }
        double const scale = available / expl;
        st.electron *= scale;
        st.positron *= scale;
        st.recoil *= scale;
        expl = st.electron + st.positron + st.recoil;
    }
    st.local = available - expl;
    // std::max_element with `<`: the first of the largest, in (e-, e+, recoil) order
    auto reduce_largest = [&](double correction) {
        double* largest = &st.electron;
        if (*largest < st.positron) largest = &st.positron;
        if (*largest < st.recoil) largest = &st.recoil;
        if (*largest < correction) { if (!err) err = kErrInvalid; return; }
        *largest -= correction;
    };
    if (st.local < 0.0) {
        double const correction = -st.local;
        if (correction > tol) { if (!err) err = kErrInvalid; return st; }
        reduce_largest(correction);
        st.local = 0.0;
    }
    double accounted = st.electron + st.positron + st.recoil + st.local;
    double const residual = available - accounted;
    if (fabs(residual) > tol) { if (!err) err = kErrInvalid; return st; }
    if (residual >= 0.0) st.local += residual;
    else if (st.local >= -residual) st.local += residual;
    else reduce_largest(-residual);
    accounted = st.electron + st.positron + st.recoil + st.local;
    if (st.electron < 0.0 || st.positron < 0.0 || st.recoil < 0.0 || st.local < 0.0
        || fabs(available - accounted) > tol) {
        if (!err) err = kErrInvalid;
    }
    return st;
}

}  // namespace rrea_dev
