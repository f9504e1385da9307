// MI355X-native population control: the charge-exact thinning plan of the
// overlay's adaptive_resample_v1 controller for one charged species on one
// rank.  Follows
//   the support record and phase bin   RreaCouplingPopulation.cpp:233-277
//   grouping, capacity, greedy request RreaCouplingPopulation.cpp:321-456
//   the per-support plan               RreaPlanGlobalCicThinning
//                                       (native_amrex_rrea/include/rrea/
//                                       RreaCicConservingResample.H:482-617)
//   the thinning primitive             RreaThinCicWeights (:236-352),
//                                       RreaCicResampleRemovalCapacity
//                                       (:222-234), null_direction (:62-155),
//                                       constraint_rank (:157-203),
//                                       RreaRoundAndVerifyCicWeights (:440-476)
//   the controller's draw              RreaAdaptiveResampleCicUniform
//                                       (warpx_overlay/Source/Rrea/RreaRng.H:112-129)
// Records are sorted on the device (hipcub radix passes, least significant
// key first); one lane owns one support and runs its thinning sequentially,
// as the reference does per support.  The reference eliminates in x87 long
// double (64-bit mantissa); the device has none and uses double-double
// (106 bits) with the reference's tolerances, so its decisions are the
// reference's and its weights agree to the last place of a double (a value
// within ~1e-19 of a rounding boundary can land one ulp apart).
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off.

#include "../../include/rrea_hip.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <vector>

namespace {

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t const e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "rrea_hip: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return RREA_ERR_HIP;                                                   \
        }                                                                          \
    } while (0)

constexpr int kThreads = 256;
constexpr int64_t kInvalidId = INT64_MIN;
// 512 * LDBL_EPSILON of x87 long double (2^-63): the reference's rank and
// null-space tolerance
constexpr double kRankTol = 0x1.0p-54;
constexpr double kDblEps = 0x1.0p-52;

// ---------------------------------------------------------------------------
// double-double arithmetic (Dekker/Knuth error-free transformations)
// ---------------------------------------------------------------------------
struct DD {
    double hi, lo;
};

__device__ __forceinline__ DD dd(double a) { return DD{a, 0.0}; }

__device__ __forceinline__ DD two_sum(double a, double b)
{
    double const s = a + b;
    double const bb = s - a;
    double const e = (a - (s - bb)) + (b - bb);
    return DD{s, e};
}

__device__ __forceinline__ DD quick_two_sum(double a, double b)
{
    double const s = a + b;
    return DD{s, b - (s - a)};
}

__device__ __forceinline__ DD operator+(DD a, DD b)
{
    DD s = two_sum(a.hi, b.hi);
    DD t = two_sum(a.lo, b.lo);
    s.lo += t.hi;
    s = quick_two_sum(s.hi, s.lo);
    s.lo += t.lo;
    return quick_two_sum(s.hi, s.lo);
}

__device__ __forceinline__ DD operator-(DD a) { return DD{-a.hi, -a.lo}; }
__device__ __forceinline__ DD operator-(DD a, DD b) { return a + (-b); }

__device__ __forceinline__ DD operator*(DD a, DD b)
{
    double const p = a.hi * b.hi;
    double const e = fma(a.hi, b.hi, -p);
    double const cross = a.hi * b.lo + a.lo * b.hi;
    return quick_two_sum(p, e + cross);
}

__device__ __forceinline__ DD operator/(DD a, DD b)
{
    double const q1 = a.hi / b.hi;
    DD r = a - b * dd(q1);
    double const q2 = r.hi / b.hi;
    r = r - b * dd(q2);
    double const q3 = r.hi / b.hi;
    return quick_two_sum(q1, q2) + dd(q3);
}

__device__ __forceinline__ bool operator<(DD a, DD b) { return a.hi < b.hi || (a.hi == b.hi && a.lo < b.lo); }
__device__ __forceinline__ bool operator>(DD a, DD b) { return b < a; }
__device__ __forceinline__ bool is_zero(DD a) { return a.hi == 0.0 && a.lo == 0.0; }
__device__ __forceinline__ DD dd_abs(DD a) { return a.hi < 0.0 || (a.hi == 0.0 && a.lo < 0.0) ? -a : a; }
__device__ __forceinline__ double to_double(DD a) { return a.hi + a.lo; }

// ---------------------------------------------------------------------------
// The thinning primitive on one support (cic_resample_detail)
// ---------------------------------------------------------------------------
struct Col {
    DD v[4];
};

__device__ __forceinline__ Col column(double fr, double fz)
{
    DD const a = dd(1.0) - dd(fr);
    DD const b = dd(1.0) - dd(fz);
    return Col{{a * b, dd(fr) * b, a * dd(fz), dd(fr) * dd(fz)}};
}

// null_direction: reduced-row-echelon elimination on the 4 x n matrix of the
// columns (n <= 5); a mixed-sign null vector normalised to max |d| = 1.
__device__ bool null_direction(Col const* cols, int n, DD* dir)
{
    DD m[4][5];
    for (int c = 0; c < n; ++c) {
        for (int r = 0; r < 4; ++r) m[r][c] = cols[c].v[r];
    }
    int pivot_column[4] = {-1, -1, -1, -1};
    bool is_pivot[5] = {false, false, false, false, false};
    int rank = 0;
    DD const tol = dd(kRankTol);
    for (int c = 0; c < n && rank < 4; ++c) {
        int pivot = rank;
        DD pivot_abs = dd(0.0);
        for (int r = rank; r < 4; ++r) {
            DD const cand = dd_abs(m[r][c]);
            if (cand > pivot_abs) {
                pivot = r;
                pivot_abs = cand;
            }
        }
        if (!(pivot_abs > tol)) continue;
        if (pivot != rank) {
            for (int j = 0; j < 5; ++j) {
                DD const t = m[pivot][j];
                m[pivot][j] = m[rank][j];
                m[rank][j] = t;
            }
        }
        DD const scale = m[rank][c];
        for (int j = 0; j < n; ++j) m[rank][j] = m[rank][j] / scale;
        for (int r = 0; r < 4; ++r) {
            if (r == rank) continue;
            DD const factor = m[r][c];
            if (is_zero(factor)) continue;
            for (int j = 0; j < n; ++j) m[r][j] = m[r][j] - factor * m[rank][j];
        }
        pivot_column[rank] = c;
        is_pivot[c] = true;
        ++rank;
    }
    if (rank == n) return false;
    int free_column = n;
    for (int c = n; c-- > 0;) {
        if (!is_pivot[c]) {
            free_column = c;
            break;
        }
    }
    if (free_column == n) return false;
    for (int c = 0; c < 5; ++c) dir[c] = dd(0.0);
    dir[free_column] = dd(1.0);
    for (int rr = rank; rr-- > 0;) {
        int const pivot = pivot_column[rr];
        DD value = dd(0.0);
        for (int c = 0; c < n; ++c) {
            if (c != pivot) value = value + m[rr][c] * dir[c];
        }
        dir[pivot] = -value;
    }
    DD max_abs = dd(0.0);
    for (int c = 0; c < n; ++c) {
        DD const a = dd_abs(dir[c]);
        if (a > max_abs) max_abs = a;
    }
    if (!(max_abs.hi > 0.0) || !isfinite(max_abs.hi)) return false;
    bool positive = false, negative = false;
    for (int c = 0; c < n; ++c) {
        dir[c] = dir[c] / max_abs;
        positive = positive || dir[c] > tol;
        negative = negative || dir[c] < -tol;
    }
    return positive && negative;
}

// constraint_rank over the positive-weight particles of [b, e)
__device__ int constraint_rank(const double* fr, const double* fz, const double* w, int64_t b, int64_t e)
{
    DD basis[4][4];
    int rank = 0;
    DD const tol = dd(kRankTol);
    for (int64_t p = b; p < e && rank < 4; ++p) {
        if (!(w[p] > 0.0)) continue;
        Col v = column(fr[p], fz[p]);
        for (int r = 0; r < rank; ++r) {
            int pivot = 0;
            while (pivot < 4 && !(dd_abs(basis[r][pivot]) > tol)) ++pivot;
            if (pivot == 4) continue;
            DD const factor = v.v[pivot] / basis[r][pivot];
            for (int j = pivot; j < 4; ++j) v.v[j] = v.v[j] - factor * basis[r][j];
        }
        int pivot = 0;
        while (pivot < 4 && !(dd_abs(v.v[pivot]) > tol)) ++pivot;
        if (pivot == 4) continue;
        DD const pv = v.v[pivot];
        for (int j = pivot; j < 4; ++j) v.v[j] = v.v[j] / pv;
        for (int j = 0; j < 4; ++j) basis[rank][j] = v.v[j];
        ++rank;
    }
    return rank;
}

// splitmix64 and the counter uniform of OV/RreaRng.cpp:9-44 (the same
// arithmetic as the transport's counter RNG)
__device__ __forceinline__ uint64_t splitmix64(uint64_t v)
{
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

__device__ __forceinline__ double counter_uniform(uint64_t seed, uint64_t pid, uint64_t step, uint64_t idx,
                                                  uint64_t chan)
{
    uint64_t s = splitmix64(seed) ^ splitmix64(pid + 0x100000001b3ULL) ^ splitmix64(step + 0x9e3779b97f4a7c15ULL);
    s ^= splitmix64(idx + 0xbf58476d1ce4e5b9ULL);
    s ^= splitmix64(chan + 0x94d049bb133111ebULL);
    return (static_cast<double>(splitmix64(s) >> 12U) + 0.5) * 0x1.0p-52;
}

// RreaCicMix64 (RreaCicConservingResample.H:413-419)
__device__ __forceinline__ uint64_t cic_mix(uint64_t v)
{
    v += 0x9e3779b97f4a7c15ULL;
    v = (v ^ (v >> 30)) * 0xbf58476d1ce4e5b9ULL;
    v = (v ^ (v >> 27)) * 0x94d049bb133111ebULL;
    return v ^ (v >> 31);
}

struct PlanArgs {
    int64_t n_groups;
    const int64_t* start;      // n_groups + 1 group offsets into the sorted records
    const int32_t* sr;         // sorted records
    const int32_t* sz;
    const double* fr;
    const double* fz;
    double* w;                 // in/out, sorted order
    const uint8_t* cand;       // may be NULL
    const uint64_t* high;      // may be NULL
    const uint64_t* low;
    const int64_t* request;    // per group
    uint64_t seed, salt, step, chan, stream_key;
    unsigned long long* removed;
    int32_t* error;
};

__device__ __forceinline__ void fail(int32_t* error, int32_t code, int64_t index)
{
    if (atomicCAS(&error[0], 0, code) == 0) error[1] = static_cast<int32_t>(index);
}

// Removal capacity and live count per support (RreaCicResampleRemovalCapacity)
__global__ void capacity_kernel(int64_t n_groups, const int64_t* __restrict__ start, const double* __restrict__ fr,
                                const double* __restrict__ fz, const double* __restrict__ w,
                                int64_t* __restrict__ capacity)
{
    int64_t const g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    int64_t const b = start[g], e = start[g + 1];
    int64_t live = 0;
    for (int64_t p = b; p < e; ++p) live += w[p] > 0.0 ? 1 : 0;
    int64_t const rank = constraint_rank(fr, fz, w, b, e);
    capacity[g] = live > rank ? live - rank : 0;
}

// The greedy request in support order: min(capacity, what the earlier
// supports left of the nominal request)
__global__ void request_kernel(int64_t n_groups, const int64_t* __restrict__ capacity,
                               const int64_t* __restrict__ before, int64_t nominal, int64_t* __restrict__ request)
{
    int64_t const g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= n_groups) return;
    int64_t const left = nominal - before[g];
    int64_t const r = left < 0 ? 0 : left;
    request[g] = capacity[g] < r ? capacity[g] : r;
}

// One lane per requested support: RreaPlanGlobalCicThinning over that support
// alone, i.e. RreaThinCicWeights with the controller's draws, then the
// stored-moment verification.
__global__ void thin_kernel(const PlanArgs* __restrict__ args)
{
    PlanArgs const& A = *args;
    int64_t const g = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (g >= A.n_groups) return;
    int64_t const request = A.request[g];
    if (request <= 0) return;
    int64_t const b = A.start[g], e = A.start[g + 1];
    // the group's draw id from its first record (RreaCicConservingResample.H:553-560)
    uint64_t const cells = (static_cast<uint64_t>(static_cast<uint32_t>(A.sr[b])) << 32)
        | static_cast<uint32_t>(A.sz[b]);
    uint64_t const support_key = cic_mix(cells ^ cic_mix(A.stream_key));
    uint64_t const first_high = A.high ? A.high[b] : 0ULL;
    uint64_t const first_cand = A.cand ? static_cast<uint64_t>(A.cand[b] != 0) : 0ULL;
    uint64_t const group_id = cic_mix(support_key ^ cic_mix(first_high) ^ cic_mix(A.low[b]) ^ first_cand);
    // the moments before (RreaRoundAndVerifyCicWeights)
    DD before[4] = {dd(0.0), dd(0.0), dd(0.0), dd(0.0)};
    for (int64_t p = b; p < e; ++p) {
        double const f_r = A.fr[p], f_z = A.fz[p], wp = A.w[p];
        if (!isfinite(f_r) || !isfinite(f_z) || f_r < 0.0 || f_r > 1.0 || f_z < 0.0 || f_z > 1.0 || !isfinite(wp)
            || !(wp > 0.0)) {
            fail(A.error, RREA_ERR_INVALID, p);
            return;
        }
        before[0] = before[0] + dd(wp * (1.0 - f_r) * (1.0 - f_z));
        before[1] = before[1] + dd(wp * f_r * (1.0 - f_z));
        before[2] = before[2] + dd(wp * (1.0 - f_r) * f_z);
        before[3] = before[3] + dd(wp * f_r * f_z);
    }
    // active: the first five positive weights in order (a window) and the
    // untouched records behind it, all positive
    int64_t win[5];
    int nw = 0;
    int64_t next = b;
    while (nw < 5 && next < e) win[nw++] = next++;
    int64_t removed = 0;
    uint64_t draws = 0;
    while (removed < request && nw >= 2) {
        int const n = nw;
        Col cols[5];
        for (int c = 0; c < n; ++c) cols[c] = column(A.fr[win[c]], A.fz[win[c]]);
        DD dir[5];
        if (!null_direction(cols, n, dir)) {
            if (next < e) {  // five columns in four dimensions are always dependent
                fail(A.error, RREA_ERR_INVALID, win[0]);
                return;
            }
            break;
        }
        DD alpha_plus = dd(INFINITY), alpha_minus = dd(INFINITY);
        int hit_plus = n, hit_minus = n;
        for (int c = 0; c < n; ++c) {
            DD const wc = dd(A.w[win[c]]);
            if (dir[c] < dd(0.0)) {
                DD const bound = wc / -dir[c];
                if (bound < alpha_plus) {
                    alpha_plus = bound;
                    hit_plus = c;
                }
            } else if (dir[c] > dd(0.0)) {
                DD const bound = wc / dir[c];
                if (bound < alpha_minus) {
                    alpha_minus = bound;
                    hit_minus = c;
                }
            }
        }
        if (!isfinite(alpha_plus.hi) || !isfinite(alpha_minus.hi) || alpha_plus.hi < 0.0 || alpha_minus.hi < 0.0
            || hit_plus == n || hit_minus == n) {
            fail(A.error, RREA_ERR_INVALID, win[0]);
            return;
        }
        double const u = counter_uniform(A.seed, group_id ^ A.salt, A.step, draws, A.chan);
        ++draws;
        DD const plus_probability = alpha_minus / (alpha_plus + alpha_minus);
        bool const take_plus = dd(u) < plus_probability;
        DD const alpha = take_plus ? alpha_plus : -alpha_minus;
        int const zero = take_plus ? hit_plus : hit_minus;
        double updated_w[5];
        for (int c = 0; c < n; ++c) {
            double const w0 = A.w[win[c]];
            DD updated = dd(w0) + alpha * dir[c];
            if (c == zero) updated = dd(0.0);
            // a round-off negative is the analytic zero (512 ulp x support size)
            double const slack = 512.0 * kDblEps * static_cast<double>(n) * fmax(1.0, fabs(w0));
            if (updated < dd(0.0) && -updated.hi <= slack) updated = dd(0.0);
            if (updated < dd(0.0) || !isfinite(updated.hi)) {
                fail(A.error, RREA_ERR_INVALID, win[c]);
                return;
            }
            double const d = to_double(updated);
            updated_w[c] = d > 0.0 ? d : 0.0;
        }
        int keep = 0;
        int64_t now = 0;
        for (int c = 0; c < n; ++c) {
            A.w[win[c]] = updated_w[c];
            if (updated_w[c] == 0.0) ++now;
            else win[keep++] = win[c];
        }
        if (now == 0) {
            fail(A.error, RREA_ERR_INVALID, win[0]);
            return;
        }
        nw = keep;
        while (nw < 5 && next < e) win[nw++] = next++;
        removed += now;
    }
    if (removed == 0) {
        fail(A.error, RREA_ERR_INVALID, b);
        return;
    }
    // the four deposited moments, rechecked after the weights are stored
    DD after[4] = {dd(0.0), dd(0.0), dd(0.0), dd(0.0)};
    for (int64_t p = b; p < e; ++p) {
        double const f_r = A.fr[p], f_z = A.fz[p], wp = A.w[p];
        after[0] = after[0] + dd(wp * (1.0 - f_r) * (1.0 - f_z));
        after[1] = after[1] + dd(wp * f_r * (1.0 - f_z));
        after[2] = after[2] + dd(wp * (1.0 - f_r) * f_z);
        after[3] = after[3] + dd(wp * f_r * f_z);
    }
    for (int k = 0; k < 4; ++k) {
        double const scale = fmax(1.0, fabs(to_double(before[k])));
        double const tol = 512.0 * kDblEps * static_cast<double>(e - b) * scale;
        if (fabs(to_double(after[k] - before[k])) > tol) {
            fail(A.error, RREA_ERR_INVALID, b);
            return;
        }
    }
    atomicAdd(A.removed, static_cast<unsigned long long>(removed));
}

// The support record of RreaCouplingPopulation.cpp:252-277 (non-periodic,
// domain small end 0): cell-centred CIC coordinates, the support's lower-left
// cell, the fractions, and phase_bin (:233-247).
__global__ void record_kernel(int64_t n, const double* __restrict__ r, const double* __restrict__ z,
                              const double* __restrict__ theta, const double* __restrict__ ux,
                              const double* __restrict__ uy, const double* __restrict__ uz, double z_lo, double dr,
                              double dz, int32_t* __restrict__ sr, int32_t* __restrict__ sz,
                              int32_t* __restrict__ phase, double* __restrict__ fr, double* __restrict__ fz,
                              int32_t* __restrict__ error)
{
    int64_t const p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double const rc = (r[p] - 0.0) / dr + 0.0 - 0.5;
    double const zc = (z[p] - z_lo) / dz + 0.0 - 0.5;
    double const i0 = floor(rc), j0 = floor(zc);
    if (!isfinite(rc) || !isfinite(zc) || i0 < -2147483648.0 || i0 > 2147483647.0 || j0 < -2147483648.0
        || j0 > 2147483647.0) {
        fail(error, RREA_ERR_INVALID, p);
        return;
    }
    // the sort key holds support_r + 1 in 26 bits and support_z + 1 in 24
    if (i0 < -1.0 || j0 < -1.0 || i0 + 1.0 >= 67108864.0 || j0 + 1.0 >= 16777216.0) {
        fail(error, RREA_ERR_INVALID, p);
        return;
    }
    sr[p] = static_cast<int32_t>(i0);
    sz[p] = static_cast<int32_t>(j0);
    fr[p] = rc - i0;
    fz[p] = zc - j0;
    double const c = cos(theta[p]), s = sin(theta[p]);
    double const ur = ux[p] * c + uy[p] * s;
    double const ut = -ux[p] * s + uy[p] * c;
    double const mag = sqrt(ux[p] * ux[p] + uy[p] * uy[p] + uz[p] * uz[p]);
    int ex = mag > 0.0 && isfinite(mag) ? ilogb(mag) : -1024;
    ex = ex < -1024 ? -1024 : (ex > 1023 ? 1023 : ex);
    phase[p] = (ex + 1024) * 8 + (ur >= 0.0 ? 1 : 0) + (ut >= 0.0 ? 2 : 0) + (uz[p] >= 0.0 ? 4 : 0);
}

// Sort keys, least significant first: the low id, the high id, the
// candidate flag, then (support_r, support_z, phase).
__global__ void key_kernel(int64_t n, int pass, const int32_t* __restrict__ perm, const uint64_t* __restrict__ low,
                           const uint64_t* __restrict__ high, const uint8_t* __restrict__ cand,
                           const int32_t* __restrict__ sr, const int32_t* __restrict__ sz,
                           const int32_t* __restrict__ phase, uint64_t* __restrict__ key)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t const p = perm[i];
    uint64_t k;
    if (pass == 0) k = low[p];
    else if (pass == 1) k = high[p];
    else if (pass == 2) k = cand[p] != 0 ? 1ULL : 0ULL;
    else {
        k = (static_cast<uint64_t>(static_cast<uint32_t>(sr[p] + 1)) << 38)
            | (static_cast<uint64_t>(static_cast<uint32_t>(sz[p] + 1)) << 14) | static_cast<uint64_t>(phase[p]);
    }
    key[i] = k;
}

__global__ void iota_kernel(int64_t n, int32_t* __restrict__ p)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) p[i] = static_cast<int32_t>(i);
}

// The records in sorted order, and a flag at each support's first record.
__global__ void gather_sorted_kernel(int64_t n, const int32_t* __restrict__ perm, const int32_t* __restrict__ sr,
                                     const int32_t* __restrict__ sz, const double* __restrict__ fr,
                                     const double* __restrict__ fz, const double* __restrict__ w,
                                     const uint8_t* __restrict__ cand, const uint64_t* __restrict__ high,
                                     const uint64_t* __restrict__ low, int32_t* __restrict__ o_sr,
                                     int32_t* __restrict__ o_sz, double* __restrict__ o_fr, double* __restrict__ o_fz,
                                     double* __restrict__ o_w, uint8_t* __restrict__ o_cand,
                                     uint64_t* __restrict__ o_high, uint64_t* __restrict__ o_low,
                                     int32_t* __restrict__ head)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t const p = perm[i];
    o_sr[i] = sr[p]; o_sz[i] = sz[p]; o_fr[i] = fr[p]; o_fz[i] = fz[p]; o_w[i] = w[p];
    o_cand[i] = cand ? cand[p] : 0;
    o_high[i] = high ? high[p] : 0ULL;
    o_low[i] = low[p];
    int32_t const q = i > 0 ? perm[i - 1] : -1;
    head[i] = (i == 0 || sr[q] != sr[p] || sz[q] != sz[p]) ? 1 : 0;
}

// gidx is the exclusive scan of the head flags: a head's own group index,
// one more than the group index elsewhere; so the end of the last group is
// start[gidx[n-1] + head[n-1]] = start[n_groups].
__global__ void group_start_kernel(int64_t n, const int32_t* __restrict__ head, const int32_t* __restrict__ gidx,
                                   int64_t* __restrict__ start)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (head[i]) start[gidx[i]] = i;
    if (i == n - 1) start[gidx[i] + head[i]] = n;
}

// New weights back to the caller's order; a particle whose weight reached
// zero is invalidated (id given).
__global__ void scatter_kernel(int64_t n, const int32_t* __restrict__ perm, const double* __restrict__ sw,
                               double* __restrict__ w, int64_t* __restrict__ id)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int32_t const p = perm[i];
    w[p] = sw[i];
    if (id && sw[i] == 0.0) id[p] = kInvalidId;
}

unsigned blocks(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

// Device buffers of one call, released on scope exit.
struct Buffers {
    std::vector<void*> ptrs;
    ~Buffers() { for (void* p : ptrs) (void)hipFree(p); }
    template <class T>
    hipError_t alloc(T** p, size_t count)
    {
        void* q = nullptr;
        hipError_t const e = hipMalloc(&q, std::max<size_t>(count, 1) * sizeof(T));
        if (e == hipSuccess) ptrs.push_back(q);
        *p = static_cast<T*>(q);
        return e;
    }
};

}  // namespace

extern "C" {

rrea_status rrea_population_thin(const rrea_population_records* rec, const rrea_population_config* cfg,
                                 int64_t* counts, int32_t* error, rrea_stream_t stream)
{
    if (!rec || !cfg || !counts || !error) return RREA_ERR_ARGUMENT;
    int64_t const n = rec->n;
    if (n < 0 || n > INT32_MAX || cfg->target < 0 || !(cfg->dr > 0.0) || !(cfg->dz > 0.0)
        || (cfg->species != 0 && cfg->species != 1)) {
        return RREA_ERR_ARGUMENT;
    }
    counts[0] = counts[1] = counts[2] = counts[3] = 0;
    if (n == 0) return RREA_OK;
    if (!rec->r || !rec->z || !rec->theta || !rec->ux || !rec->uy || !rec->uz || !rec->w || !rec->stable_id) {
        return RREA_ERR_ARGUMENT;
    }
    hipStream_t const s = static_cast<hipStream_t>(stream);
    Buffers B;
    int32_t *sr, *sz, *phase, *perm, *perm2, *head, *gidx, *o_sr, *o_sz;
    double *fr, *fz, *o_fr, *o_fz, *o_w;
    uint64_t *key, *key2, *o_high, *o_low;
    uint8_t* o_cand;
    HIP_TRY(B.alloc(&sr, n)); HIP_TRY(B.alloc(&sz, n)); HIP_TRY(B.alloc(&phase, n));
    HIP_TRY(B.alloc(&perm, n)); HIP_TRY(B.alloc(&perm2, n)); HIP_TRY(B.alloc(&head, n)); HIP_TRY(B.alloc(&gidx, n));
    HIP_TRY(B.alloc(&o_sr, n)); HIP_TRY(B.alloc(&o_sz, n));
    HIP_TRY(B.alloc(&fr, n)); HIP_TRY(B.alloc(&fz, n)); HIP_TRY(B.alloc(&o_fr, n)); HIP_TRY(B.alloc(&o_fz, n));
    HIP_TRY(B.alloc(&o_w, n));
    HIP_TRY(B.alloc(&key, n)); HIP_TRY(B.alloc(&key2, n)); HIP_TRY(B.alloc(&o_high, n)); HIP_TRY(B.alloc(&o_low, n));
    HIP_TRY(B.alloc(&o_cand, n));
    unsigned const nb = blocks(n);
    record_kernel<<<dim3(nb), dim3(kThreads), 0, s>>>(n, rec->r, rec->z, rec->theta, rec->ux, rec->uy, rec->uz,
                                                       cfg->z_lo, cfg->dr, cfg->dz, sr, sz, phase, fr, fz, error);
    iota_kernel<<<dim3(nb), dim3(kThreads), 0, s>>>(n, perm);
    HIP_TRY(hipGetLastError());
    // radix passes, least significant key first (each pass is stable)
    size_t tmp_bytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, key, key2, perm, perm2, static_cast<int>(n), 0,
                                               64, s));
    void* tmp = nullptr;
    HIP_TRY(B.alloc(reinterpret_cast<char**>(&tmp), tmp_bytes));
    for (int pass = 0; pass < 4; ++pass) {
        if (pass == 1 && !rec->stable_id_high) continue;
        if (pass == 2 && !rec->candidate) continue;
        key_kernel<<<dim3(nb), dim3(kThreads), 0, s>>>(n, pass, perm, rec->stable_id, rec->stable_id_high,
                                                        rec->candidate, sr, sz, phase, key);
        HIP_TRY(hipGetLastError());
        int const end_bit = pass == 2 ? 1 : 64;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, key, key2, perm, perm2, static_cast<int>(n), 0,
                                                   end_bit, s));
        std::swap(perm, perm2);
    }
    gather_sorted_kernel<<<dim3(nb), dim3(kThreads), 0, s>>>(n, perm, sr, sz, fr, fz, rec->w, rec->candidate,
                                                              rec->stable_id_high, rec->stable_id, o_sr, o_sz, o_fr,
                                                              o_fz, o_w, o_cand, o_high, o_low, head);
    HIP_TRY(hipGetLastError());
    // group index per record, group starts, the group count
    size_t scan_bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, scan_bytes, head, gidx, static_cast<int>(n), s));
    void* scan_tmp = nullptr;
    HIP_TRY(B.alloc(reinterpret_cast<char**>(&scan_tmp), scan_bytes));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(scan_tmp, scan_bytes, head, gidx, static_cast<int>(n), s));
    int32_t last_g = 0, last_head = 0;
    int32_t host_err[2] = {0, 0};
    HIP_TRY(hipMemcpyAsync(&last_g, gidx + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&last_head, head + n - 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(host_err, error, sizeof(host_err), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    if (host_err[0] != 0) return RREA_OK;  // an invalid record: nothing planned
    int64_t const n_groups = static_cast<int64_t>(last_g) + (last_head ? 1 : 0);
    int64_t *start, *capacity, *before, *request;
    HIP_TRY(B.alloc(&start, n_groups + 1)); HIP_TRY(B.alloc(&capacity, n_groups));
    HIP_TRY(B.alloc(&before, n_groups)); HIP_TRY(B.alloc(&request, n_groups));
    group_start_kernel<<<dim3(nb), dim3(kThreads), 0, s>>>(n, head, gidx, start);
    capacity_kernel<<<dim3(blocks(n_groups)), dim3(kThreads), 0, s>>>(n_groups, start, o_fr, o_fz, o_w, capacity);
    HIP_TRY(hipGetLastError());
    size_t cs_bytes = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, cs_bytes, capacity, before, static_cast<int>(n_groups), s));
    void* cs_tmp = nullptr;
    HIP_TRY(B.alloc(reinterpret_cast<char**>(&cs_tmp), cs_bytes));
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(cs_tmp, cs_bytes, capacity, before, static_cast<int>(n_groups), s));
    // the global capacity and geometric floor
    int64_t last_before = 0, last_cap = 0;
    HIP_TRY(hipMemcpyAsync(&last_before, before + n_groups - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(&last_cap, capacity + n_groups - 1, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    int64_t const total_capacity = last_before + last_cap;
    bool const due = n > cfg->target + cfg->target / 4;
    int64_t const nominal = due ? std::min<int64_t>(n - cfg->target, total_capacity) : 0;
    counts[0] = total_capacity;
    counts[1] = n - total_capacity;
    counts[2] = nominal;
    if (nominal == 0) return RREA_OK;
    request_kernel<<<dim3(blocks(n_groups)), dim3(kThreads), 0, s>>>(n_groups, capacity, before, nominal, request);
    HIP_TRY(hipGetLastError());
    unsigned long long* removed;
    PlanArgs* d_args;
    HIP_TRY(B.alloc(&removed, 1));
    HIP_TRY(B.alloc(&d_args, 1));
    HIP_TRY(hipMemsetAsync(removed, 0, sizeof(unsigned long long), s));
    PlanArgs A{n_groups, start, o_sr, o_sz, o_fr, o_fz, o_w, rec->candidate ? o_cand : nullptr,
               rec->stable_id_high ? o_high : nullptr, o_low, request, cfg->rng_seed, cfg->rng_salt,
               static_cast<uint64_t>(cfg->step), cfg->species == 0 ? 131ULL : 132ULL,
               cfg->species == 0 ? 0x656c656374726f6eULL : 0x706f736974726f6eULL, removed, error};
    // pageable source: the copy has consumed A when hipMemcpyAsync returns
    HIP_TRY(hipMemcpyAsync(d_args, &A, sizeof(PlanArgs), hipMemcpyHostToDevice, s));
    // one lane per support, 64-lane workgroups
    thin_kernel<<<dim3(static_cast<unsigned>((n_groups + 63) / 64)), dim3(64), 0, s>>>(d_args);
    HIP_TRY(hipGetLastError());
    unsigned long long removed_host = 0;
    HIP_TRY(hipMemcpyAsync(&removed_host, removed, sizeof(removed_host), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipMemcpyAsync(host_err, error, sizeof(host_err), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    counts[3] = static_cast<int64_t>(removed_host);
    if (host_err[0] != 0) return RREA_OK;  // a support rejected its plan: no weight is written
    if (static_cast<int64_t>(removed_host) < nominal) {
        // insufficient progress: the reference aborts before applying
        int32_t const e[2] = {RREA_ERR_INVALID, -1};
        HIP_TRY(hipMemcpyAsync(error, e, sizeof(e), hipMemcpyHostToDevice, s));
        HIP_TRY(hipStreamSynchronize(s));
        return RREA_OK;
    }
    scatter_kernel<<<dim3(nb), dim3(kThreads), 0, s>>>(n, perm, o_w, rec->w, rec->id);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipStreamSynchronize(s));
    return RREA_OK;
}

}  // extern "C"
