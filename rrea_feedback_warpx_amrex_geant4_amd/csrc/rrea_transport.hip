// MI355X-native RREA transport step: kernels and the C-ABI of include/rrea_hip.h.
//
// The electron kernel is the per-particle body of
// TransportStep::ApplyElectronTransport
// (native_amrex_rrea/warpx_overlay/Source/Rrea/RreaElectronTransport.cpp:59-889)
// with the shared prologue BeginRunningChargedParticle
// (warpx_overlay/Source/Rrea/RreaTransportStep.H:720-820), run one lane per
// particle.  Side effects the reference logs per 64-particle chunk and
// replays later are applied directly: net-chord and escape deposits go to the
// padded J grid with FP64 atomics, secondaries are appended to a staging
// buffer and then ordered by (parent index, spawn lineage) -- the reference's
// chunk-order merge -- and tallies are reduced per workgroup.
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off (no fused
// multiply-add anywhere: the step must agree with the reference to 1e-12).

#include "rrea_device.h"
#include "rrea_drift.h"
#include "../../include/rrea_hip.h"

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <mutex>
#include <new>
#include <vector>

using namespace rrea_dev;

namespace {

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t const e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "rrea_hip: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return RREA_ERR_HIP;                                                   \
        }                                                                          \
    } while (0)

// ---------------------------------------------------------------------------
// Host-side table derivation (the parts of RreaInteractionTables::Load that
// compute: NA/src/RreaInteractionTables.cpp:1632-1657, 2146-2179, 1071-1110)
// ---------------------------------------------------------------------------
struct HostCdf {
    std::vector<double> primary, log_primary, u, v, mean, log_mean;
    std::vector<int32_t> start;
};

HostCdf host_cdf(rrea_cdf_groups const& g)
{
    HostCdf h;
    h.primary.assign(g.primary, g.primary + g.n_groups);
    for (double p : h.primary) h.log_primary.push_back(std::log(p));
    int64_t const rows = g.start[g.n_groups];
    h.u.assign(g.u, g.u + rows);
    h.v.assign(g.v, g.v + rows);
    for (int64_t i = 0; i <= g.n_groups; ++i) h.start.push_back(static_cast<int32_t>(g.start[i]));
    for (int64_t k = 0; k < g.n_groups; ++k) {
        // group_mean of mean_conditional_cdf, same summation order
        int64_t const s = g.start[k], e = g.start[k + 1];
        double mean = h.u[s] * h.v[s];
        for (int64_t i = s + 1; i < e; ++i) mean += 0.5 * (h.v[i - 1] + h.v[i]) * (h.u[i] - h.u[i - 1]);
        mean += (1.0 - h.u[e - 1]) * h.v[e - 1];
        h.mean.push_back(mean);
        h.log_mean.push_back(mean > 0.0 ? std::log(mean) : 0.0);
    }
    return h;
}

// SampleMollerSecondaryEnergy :2584-2653 on the host, for the Load-time mean.
double host_sample_cdf_group(HostCdf const& h, int64_t g, double cu)
{
    int64_t const s = h.start[g], e = h.start[g + 1];
    auto const b = h.u.begin() + s, en = h.u.begin() + e;
    auto const up = std::lower_bound(b, en, cu);
    if (up == b) return h.v[s];
    if (up == en) return h.v[e - 1];
    int64_t const hi = up - h.u.begin();
    int64_t const lo = hi - 1;
    double const denom = h.u[hi] - h.u[lo];
    double const t = denom > 0.0 ? (cu - h.u[lo]) / denom : 0.0;
    return h.v[lo] + t * (h.v[hi] - h.v[lo]);
}

double host_sample_moller(HostCdf const& h, double primary, double cu)
{
    cu = std::max(0.0, std::min(1.0, cu));
    int64_t const n = static_cast<int64_t>(h.primary.size());
    auto const up = std::lower_bound(h.primary.begin(), h.primary.end(), primary);
    if (up == h.primary.begin()) return host_sample_cdf_group(h, 0, cu);
    if (up == h.primary.end()) return host_sample_cdf_group(h, n - 1, cu);
    int64_t const hi = up - h.primary.begin();
    int64_t const lo = hi - 1;
    if (h.primary[hi] == primary) return host_sample_cdf_group(h, hi, cu);
    double const y0 = host_sample_cdf_group(h, lo, cu);
    double const y1 = host_sample_cdf_group(h, hi, cu);
    double const x0 = std::log(h.primary[lo]);
    double const x1 = std::log(h.primary[hi]);
    double const t = (std::log(primary) - x0) / (x1 - x0);
    return std::exp(std::log(y0) + t * (std::log(y1) - std::log(y0)));
}

// A bin table of lower-bound starts over nondecreasing keys (the Guide of
// rrea_device.h): any start is corrected by the device scans, so the table
// only has to be close.
struct HostGuide {
    std::vector<int32_t> idx;
    double lo = 0.0, scale = 0.0;
};

HostGuide host_guide(std::vector<double> const& key, int32_t nbins)
{
    HostGuide g;
    g.lo = key.front();
    double const span = key.back() - key.front();
    g.scale = span > 0.0 ? static_cast<double>(nbins) / span : 0.0;
    g.idx.resize(nbins);
    for (int32_t b = 0; b < nbins; ++b) {
        double const edge = g.scale > 0.0 ? g.lo + static_cast<double>(b) / g.scale : g.lo;
        g.idx[b] = static_cast<int32_t>(std::lower_bound(key.begin(), key.end(), edge) - key.begin());
    }
    return g;
}

constexpr int32_t kChargedGuideBins = 8192;
constexpr int32_t kElasticGuideBins = 1024;
constexpr int32_t kProfileGuideBins = 4096;

struct DeviceBlob {
    std::vector<double> f64;
    std::vector<int32_t> i32;
    size_t add(std::vector<double> const& v)
    {
        size_t const off = f64.size();
        f64.insert(f64.end(), v.begin(), v.end());
        while (f64.size() % 32) f64.push_back(0.0);  // 256-byte alignment of every array
        return off;
    }
    size_t add_i(std::vector<int32_t> const& v)
    {
        size_t const off = i32.size();
        i32.insert(i32.end(), v.begin(), v.end());
        while (i32.size() % 64) i32.push_back(0);
        return off;
    }
};

}  // namespace

struct rrea_tables_s {
    double* d_f64 = nullptr;
    int32_t* d_i32 = nullptr;
    DevTables view{};
    std::vector<double> hm_mean;
    double reference_density = 0.0;
    double brems_cutoff = 0.0;
    int device = 0;
};

struct rrea_fields_s {
    rrea_rz_geometry geom{};
    double* d_cell_er = nullptr;
    double* d_cell_ez = nullptr;
    double* d_btheta = nullptr;
    double* d_stage = nullptr;   // staging for host-pointer updates
    size_t stage_len = 0;
    double* d_prof = nullptr;    // altitude | rho
    std::vector<double> h_prof_alt, h_prof_rho;  // host copy, for the reference-z ratio
    int64_t prof_n = 0;
    int32_t* d_prof_guide = nullptr;
    double prof_guide_lo = 0.0, prof_guide_scale = 0.0;
    int density_mode = 0;
    double density_ratio = 1.0;
    double altitude_at_z0 = 0.0;
    double reference_density = 1.0;
    bool fields_set = false;
    // secondary ordering workspace
    rrea_secondary* d_stage_sec = nullptr;
    int64_t stage_sec_cap = 0;
    uint64_t* d_keys = nullptr;
    uint64_t* d_keys_out = nullptr;
    int64_t* d_idx = nullptr;
    int64_t* d_idx_out = nullptr;
    void* d_sort_tmp = nullptr;
    size_t sort_tmp_bytes = 0;
    int64_t* d_count = nullptr;
    void* d_args = nullptr;      // device copy of the kernel's argument block
    int64_t* d_base = nullptr;       // the caller's secondary count at the start of a call
    int32_t* d_slow_list = nullptr;  // fast-path hand-over list
    int32_t* d_slow_count = nullptr;
    int64_t slow_cap = 0;
    double* d_partials = nullptr;    // the fast paths' per-workgroup partial rows (slow_cap's workgroups)
    // tile ordering workspace
    uint32_t* d_tkeys = nullptr;
    uint32_t* d_tkeys_out = nullptr;
    int32_t* d_tidx = nullptr;
    int32_t* d_tidx_out = nullptr;
    void* d_tsort_tmp = nullptr;
    size_t tsort_tmp_bytes = 0;
    void* d_tbuf = nullptr;
    int64_t tsort_cap = 0;
    int64_t* d_live = nullptr;       // rrea_tile_compact's live count
    // drift adjoint workspace: the argument block, the cell-centred gradients
    void* d_dargs = nullptr;
    double* d_gcell = nullptr;
    // newborn FIFO workspace
    void* d_nargs = nullptr;
    rrea_secondary* d_fin = nullptr;
    uint8_t* d_fst = nullptr;
    double* d_fion = nullptr;
    int32_t* d_fsub = nullptr;
    int64_t* d_anchor = nullptr;
    unsigned long long* d_next = nullptr;
    int32_t* d_flag = nullptr;
    int64_t* d_nsel = nullptr;
    void* d_sel_tmp = nullptr;
    size_t sel_tmp_bytes = 0;
    int64_t fifo_cap = 0;
    bool timing = false;
    // signalled after each fast-path kernel when enabled (rrea_fields_signal_fast_path)
    hipEvent_t fast_event = nullptr;
    bool fast_signal = false;
    int32_t block = 256;             // the fast kernels' workgroup size (rrea_fields_set_block)
    void* d_dstate = nullptr;        // the device-driven drain's generation state
    struct Timed { hipEvent_t start, stop; int32_t kernel; };
    std::vector<Timed> events;   // pool; the first events_used are recorded launches
    size_t events_used = 0;
};

// A batch of independent simulation instances (one fields handle, J grid and
// RNG seed each) stepped together.
struct rrea_batch_s {
    std::vector<rrea_fields_t> fields;
    std::vector<double*> jr, jz, rho;
    std::vector<uint64_t> seeds;
    std::vector<double*> src;       // each instance's material-source grids (empty: none)
    rrea_fields_t work = nullptr;   // workspace of the batched calls (kernel arguments, lists, timing)
    void* d_envs = nullptr;         // one environment per instance
    std::vector<unsigned char> h_envs;
};

namespace {

// ---------------------------------------------------------------------------
// Field preparation: the cell-centred totals GatherTotalFieldRZ forms per
// query (RreaAmrexAdvance.cpp:556-567), formed once per update, same order.
// ---------------------------------------------------------------------------
__global__ void cell_fields_kernel(int32_t nr, int32_t nz, const double* __restrict__ er,
                                   const double* __restrict__ ez, const double* __restrict__ aer,
                                   const double* __restrict__ aez, double* __restrict__ cer, double* __restrict__ cez)
{
    int64_t const c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (c >= static_cast<int64_t>(nr) * nz) return;
    int32_t const i = static_cast<int32_t>(c % nr);
    int32_t const j = static_cast<int32_t>(c / nr);
    int64_t const fr0 = i + static_cast<int64_t>(nr + 1) * j;
    double const lo_r = er[fr0] + aer[fr0];
    double const hi_r = er[fr0 + 1] + aer[fr0 + 1];
    cer[c] = 0.5 * (lo_r + hi_r);
    int64_t const fz0 = i + static_cast<int64_t>(nr) * j;
    double const lo_z = ez[fz0] + aez[fz0];
    double const hi_z = ez[fz0 + nr] + aez[fz0 + nr];
    cez[c] = 0.5 * (lo_z + hi_z);
}

// ---------------------------------------------------------------------------
// Workgroup reduction of the per-particle tallies
// ---------------------------------------------------------------------------
constexpr int kBlock = 256;
constexpr int kTallyUsed = 17;

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off > 0; off >>= 1) {
        double const o = __shfl_xor(v, off, 64);
        v = v < o ? o : v;
    }
    return v;
}

// A running maximum of non-negative doubles, whose bit patterns order as
// unsigned integers (a zeroed slot is the empty maximum).
__device__ __forceinline__ void atomic_max_nonneg(double* p, double v)
{
    if (v > 0.0) {
        atomicMax(reinterpret_cast<unsigned long long*>(p), static_cast<unsigned long long>(__double_as_longlong(v)));
    }
}

__host__ __device__ constexpr bool tally_is_max(int q)
{
    return q >= RREA_LEDGER_OFFSET && RREA_LEDGER_IS_MAX(q - RREA_LEDGER_OFFSET);
}

// The workgroup's event ledger in LDS: zeroed before any record ...
__device__ __forceinline__ void ledger_begin(double* ledger)
{
    for (int q = threadIdx.x; q < RREA_LEDGER_COUNT; q += blockDim.x) ledger[q] = 0.0;
    __syncthreads();
}

// ... and folded into the device tallies once every record is in (call after a
// workgroup barrier that follows the last record)
__device__ __forceinline__ void ledger_flush(double const* ledger, double* tallies)
{
    for (int q = threadIdx.x; q < RREA_LEDGER_COUNT; q += blockDim.x) {
        double const v = ledger[q];
        double* const out = tallies + RREA_LEDGER_OFFSET + q;
        if (RREA_LEDGER_IS_MAX(q)) atomic_max_nonneg(out, v);
        else if (v != 0.0) atomic_add(out, v);
    }
}

// The record operations of RreaEventLedger (NA/include/rrea/RreaEventLedger.H:
// 98-413) on the workgroup's ledger in LDS (RREA_LEDGER_COUNT doubles, zeroed
// at the start of the kernel and folded into the device tallies at its end);
// cutoff is the ledger's low_energy_cutoff_eV.  Sums add atomically; the
// running maxima keep the larger value (atomic max on the bit pattern of a
// non-negative double; a zero, negative or NaN value leaves the maximum as
// it is, as the reference's max does for a ledger that starts at zero).
struct Ledger {
    double* t;
    double cutoff;
    __device__ void add(int q, double v)
    {
        if (v != 0.0) atomicAdd(t + q, v);
    }
    __device__ void up(int q, double v)
    {
        if (v > 0.0) {
            atomicMax(reinterpret_cast<unsigned long long*>(t + q), static_cast<unsigned long long>(__double_as_longlong(v)));
        }
    }
    __device__ static double pos(double v) { return dmax(v, 0.0); }

    __device__ void transport_ionization(double w, double loss, bool low, bool kinetic)  // :98-116
    {
        if (w <= 0.0) return;
        add(RREA_L_TRANSPORT_IONIZATION_WEIGHT, w);
        add(RREA_L_TRANSPORT_ENERGY_LOSS_EV, w * pos(loss));
        if (low) add(RREA_L_TRANSPORT_LOW_SECONDARY_WEIGHT, w);
        if (kinetic) add(RREA_L_TRANSPORT_KINETIC_SECONDARY_WEIGHT, w);
    }
    __device__ void loss_partition(double w, double drag, double col, double hard)  // :118-134
    {
        if (w <= 0.0) return;
        add(RREA_L_PARENT_DRAG_LOSS_EV, w * pos(drag));
        add(RREA_L_ION_PAIR_SOURCE_LOSS_EV, w * pos(col));
        add(RREA_L_RESTRICTED_COLLISION_LOSS_EV, w * pos(col));
        add(RREA_L_EXPLICIT_HARD_MOLLER_ENERGY_LOSS_EV, w * pos(hard));
    }
    __device__ void soft_brems(double w, double e)  // :136-143
    {
        if (w > 0.0 && e > 0.0) add(RREA_L_PHOTON_CONTINUOUS_SOFT_BREMS_ENERGY_EV, w * e);
    }
    __device__ void hm_expectation(double x)  // :145-150
    {
        if (x > 0.0) add(RREA_L_HARD_MOLLER_EXPECTED_EVENT_WEIGHT, x);
    }
    __device__ void hm_event(double w, double sec)  // :152-168
    {
        if (w <= 0.0) return;
        add(RREA_L_HARD_MOLLER_EVENT_COUNT, 1.0);
        add(RREA_L_HARD_MOLLER_EVENT_WEIGHT, w);
        if (sec >= cutoff) add(RREA_L_DIRECT_HARD_MOLLER_ELECTRON_WEIGHT, w);
        add(RREA_L_HARD_MOLLER_SECONDARY_ENERGY_WEIGHTED_SUM_EV, w * pos(sec));
        add(RREA_L_HARD_MOLLER_SECONDARY_ENERGY_WEIGHT_SUM, w);
        up(RREA_L_HARD_MOLLER_SECONDARY_ENERGY_MAX_EV, sec);
    }
    __device__ void photon_interaction(double w)  // :170-180
    {
        if (w <= 0.0) return;
        add(RREA_L_PHOTON_INTERACTION_COUNT, 1.0);
        add(RREA_L_PHOTON_INTERACTION_WEIGHT, w);
    }
    __device__ void photon_escape(double w, double e)  // :182-189
    {
        if (w > 0.0 && e > 0.0) add(RREA_L_PHOTON_ESCAPED_ENERGY_EV, w * e);
    }
    __device__ void brems_photon(double w, double e)  // :191-204
    {
        if (w <= 0.0 || e <= 0.0) return;
        add(RREA_L_ELECTRON_BREMS_PHOTON_COUNT, 1.0);
        add(RREA_L_TOTAL_BREMS_PHOTON_COUNT, 1.0);
        add(RREA_L_ELECTRON_BREMS_PHOTON_WEIGHT, w);
        add(RREA_L_ELECTRON_BREMS_PHOTON_ENERGY_EV, w * e);
        add(RREA_L_TOTAL_BREMS_PHOTON_WEIGHT, w);
        add(RREA_L_TOTAL_BREMS_PHOTON_ENERGY_EV, w * e);
    }
    // a photon-seeded electron of kinetic energy e, kinetic or low-energy
    __device__ void seeded_electron(double w, double e, int kinetic_slot, int low_slot)
    {
        if (e >= 0.0) {
            add(RREA_L_PHOTON_SEEDED_ELECTRON_WEIGHT, w);
            if (e >= cutoff) {
                add(RREA_L_PHOTON_SEEDED_KINETIC_ELECTRON_WEIGHT, w);
                add(kinetic_slot, w);
            } else {
                add(RREA_L_PHOTON_SEEDED_LOW_ELECTRON_WEIGHT, w);
                add(low_slot, w);
            }
        }
    }
    __device__ void compton(double w, double recoil, double charged, double local)  // :206-235
    {
        photon_interaction(w);
        add(RREA_L_COMPTON_INTERACTION_COUNT, 1.0);
        add(RREA_L_COMPTON_INTERACTION_WEIGHT, w);
        seeded_electron(w, recoil, RREA_L_COMPTON_KINETIC_ELECTRON_WEIGHT, RREA_L_COMPTON_LOW_ELECTRON_WEIGHT);
        add(RREA_L_PHOTON_TO_CHARGED_ENERGY_EV, w * pos(charged));
        add(RREA_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV, w * pos(local));
        add(RREA_L_TRANSPORT_ENERGY_LOSS_EV, w * (pos(charged) + pos(local)));
    }
    __device__ void photoelectric(double w, double photon, double electron, double charged, double local)  // :237-269
    {
        photon_interaction(w);
        add(RREA_L_PHOTOELECTRIC_INTERACTION_COUNT, 1.0);
        add(RREA_L_PHOTOELECTRIC_INTERACTION_WEIGHT, w);
        seeded_electron(w, electron, RREA_L_PHOTOELECTRIC_KINETIC_ELECTRON_WEIGHT,
                        RREA_L_PHOTOELECTRIC_LOW_ELECTRON_WEIGHT);
        add(RREA_L_PHOTON_ABSORBED_ENERGY_EV, w * pos(photon));
        add(RREA_L_PHOTON_TO_CHARGED_ENERGY_EV, w * pos(charged));
        add(RREA_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV, w * pos(local));
        add(RREA_L_TRANSPORT_ENERGY_LOSS_EV, w * pos(photon));
    }
    __device__ void pair(double w, double photon, double electron, double positron, double recoil, double charged,
                         double local, bool triplet)  // :271-340
    {
        photon_interaction(w);
        add(RREA_L_PAIR_PRODUCTION_COUNT, 1.0);
        add(RREA_L_PAIR_PRODUCTION_WEIGHT, w);
        add(triplet ? RREA_L_PAIR_TRIPLET_COUNT : RREA_L_PAIR_NUCLEAR_COUNT, 1.0);
        add(triplet ? RREA_L_PAIR_TRIPLET_WEIGHT : RREA_L_PAIR_NUCLEAR_WEIGHT, w);
        seeded_electron(w, electron, RREA_L_PAIR_KINETIC_ELECTRON_WEIGHT, RREA_L_PAIR_LOW_ELECTRON_WEIGHT);
        add(RREA_L_PHOTON_SEEDED_POSITRON_WEIGHT, w);
        if (positron >= cutoff) {
            add(RREA_L_PHOTON_SEEDED_KINETIC_POSITRON_WEIGHT, w);
            add(RREA_L_PAIR_KINETIC_POSITRON_WEIGHT, w);
        } else {
            add(RREA_L_PHOTON_SEEDED_BELOW_CUTOFF_POSITRON_WEIGHT, w);
            add(RREA_L_PAIR_BELOW_CUTOFF_POSITRON_WEIGHT, w);
        }
        if (triplet) {
            add(RREA_L_PHOTON_SEEDED_ELECTRON_WEIGHT, w);
            if (recoil >= cutoff) {
                add(RREA_L_PHOTON_SEEDED_KINETIC_ELECTRON_WEIGHT, w);
                add(RREA_L_PAIR_TRIPLET_RECOIL_KINETIC_ELECTRON_WEIGHT, w);
            } else {
                add(RREA_L_PHOTON_SEEDED_LOW_ELECTRON_WEIGHT, w);
                add(RREA_L_PAIR_TRIPLET_RECOIL_LOW_ELECTRON_WEIGHT, w);
            }
        }
        add(RREA_L_PAIR_RECOIL_ELECTRON_ENERGY_EV, w * pos(recoil));
        add(RREA_L_PAIR_LOCAL_DEPOSIT_ENERGY_EV, w * pos(local));
        add(RREA_L_PHOTON_ABSORBED_ENERGY_EV, w * pos(photon));
        add(RREA_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV, w * pos(local));
        add(RREA_L_PHOTON_TO_CHARGED_ENERGY_EV, w * pos(charged));
        add(RREA_L_TRANSPORT_ENERGY_LOSS_EV, w * pos(charged + local - 2.0 * kMeC2));
    }
    __device__ void positron_transport(double w, double loss, bool thermalized)  // :342-354
    {
        if (w <= 0.0) return;
        add(RREA_L_POSITRON_TRANSPORT_ENERGY_LOSS_EV, w * pos(loss));
        if (thermalized) add(RREA_L_POSITRON_THERMALIZED_COUNT, 1.0);
    }
    __device__ void bhabha(double w, double sec, bool kinetic)  // :356-370
    {
        if (w <= 0.0) return;
        add(RREA_L_POSITRON_BHABHA_EVENT_COUNT, 1.0);
        add(RREA_L_POSITRON_BHABHA_EVENT_WEIGHT, w);
        add(RREA_L_POSITRON_BHABHA_SECONDARY_ENERGY_EV, w * pos(sec));
        if (kinetic) add(RREA_L_POSITRON_BHABHA_KINETIC_SECONDARY_WEIGHT, w);
    }
    __device__ void positron_brems(double w, double e)  // :372-385
    {
        if (w <= 0.0 || e <= 0.0) return;
        add(RREA_L_POSITRON_BREMS_EVENT_COUNT, 1.0);
        add(RREA_L_TOTAL_BREMS_PHOTON_COUNT, 1.0);
        add(RREA_L_POSITRON_BREMS_EVENT_WEIGHT, w);
        add(RREA_L_POSITRON_BREMS_PHOTON_ENERGY_EV, w * e);
        add(RREA_L_TOTAL_BREMS_PHOTON_WEIGHT, w);
        add(RREA_L_TOTAL_BREMS_PHOTON_ENERGY_EV, w * e);
    }
    __device__ void guard(double hm_tau, double brems_tau, double total_tau, double photon_tau, double positron_tau,
                          uint64_t subcycles, double electron_e, double photon_e, double positron_e,
                          bool violation)  // :387-413
    {
        up(RREA_L_MAX_TRANSPORT_HARD_MOLLER_TAU, hm_tau);
        up(RREA_L_MAX_TRANSPORT_BREMS_TAU, brems_tau);
        up(RREA_L_MAX_TRANSPORT_ELECTRON_TOTAL_COMPETING_TAU, total_tau);
        up(RREA_L_MAX_TRANSPORT_PHOTON_TAU, photon_tau);
        up(RREA_L_MAX_TRANSPORT_POSITRON_TAU, positron_tau);
        up(RREA_L_MAX_TRANSPORT_SUBCYCLES, static_cast<double>(subcycles));
        up(RREA_L_MAX_TRANSPORT_ELECTRON_ENERGY_EV, electron_e);
        up(RREA_L_MAX_TRANSPORT_PHOTON_ENERGY_SEEN_EV, photon_e);
        up(RREA_L_MAX_TRANSPORT_POSITRON_ENERGY_EV, positron_e);
        if (violation) add(RREA_L_TRANSPORT_OPTICAL_DEPTH_VIOLATION_COUNT, 1.0);
    }
    // the ledger part of RreaWarpXCoupling::RecordPositronAnnihilation
    // (RecordPhotonLocalDeposit), OV/RreaCouplingSources.cpp:518-553
    __device__ void photon_local_deposit(double w, double e)
    {
        if (w > 0.0 && e > 0.0) add(RREA_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV, w * e);
    }
};

// Probability of a competing-risk optical depth, OV/RreaMaterialPath.H:203-210
__device__ __forceinline__ double probability_from_optical_depth(double tau)
{
    if (!(tau > 0.0)) return 0.0;
    return dmin(1.0, -expm1(-tau));
}

// The photon energy flows the interaction records carry: charged (kinetic
// plus created rest mass) and local, RreaComptonPhotonEnergyFlow /
// RreaPhotoelectricPhotonEnergyFlow / RreaPairPhotonEnergyFlow,
// OV/RreaParticleInteraction.H:1014-1092
struct EnergyFlow { double charged, local; };

__device__ __forceinline__ EnergyFlow compton_flow(double recoil, double scattered, double cutoff, double gcut)
{
    EnergyFlow f{0.0, 0.0};
    double const r = dmax(recoil, 0.0), sc = dmax(scattered, 0.0);
    if (r >= cutoff) f.charged = r;
    else f.local = r;
    if (sc < gcut) f.local += sc;
    return f;
}

__device__ __forceinline__ EnergyFlow photoelectric_flow(double electron, double binding, double cutoff)
{
    EnergyFlow f{0.0, dmax(binding, 0.0)};
    double const e = dmax(electron, 0.0);
    if (e >= cutoff) f.charged = e;
    else f.local += e;
    return f;
}

__device__ __forceinline__ EnergyFlow pair_flow(double electron, double positron, double recoil, double local,
                                                double cutoff, bool has_recoil)
{
    EnergyFlow f{2.0 * kMeC2, dmax(local, 0.0)};
    double const e[3] = {dmax(electron, 0.0), dmax(positron, 0.0), has_recoil ? dmax(recoil, 0.0) : -1.0};
    for (int q = 0; q < 3; ++q) {
        if (e[q] < 0.0) continue;
        if (e[q] >= cutoff) f.charged += e[q];
        else f.local += e[q];
    }
    return f;
}

// ---------------------------------------------------------------------------
// The electron step kernel
// ---------------------------------------------------------------------------
// One simulation's environment: its fields, domain, padded J grid, the
// reference-z density fallback and its RNG seed.  A plain step has one; a
// batched step one per instance, selected by each particle's instance id.
struct Env {
    DevFields F;
    Domain D;
    DevGrid G;
    double density_ratio;     // reference-z fallback ratio
    int32_t pnr, pzb, pnz;
    int32_t instance;         // -1 for a plain step
    uint64_t rng_seed;
    double* src;              // material-source grids (5 x nr x nz), or null: none accumulated
};

// ---------------------------------------------------------------------------
// Material sources: the coupling's AccumulateIonizationEvent /
// AccumulatePositiveIonSourceOnly / AccumulateLowEnergyElectron
// (OV/RreaCouplingSources.cpp:435-516) into the event accumulator's fields
// (RreaEventAccumulator, NA/src/RreaEventAccumulator.cpp): 0 low-energy
// electron source, 1 positive-ion source, 2 direct negative-ion source,
// 3 ionization event count, 4 energy loss.  Raw weights here; the division
// by cell volume (and dt) is rrea_sources_normalize.
// ---------------------------------------------------------------------------
enum : int32_t { kSrcLowElectron = 0, kSrcPositiveIon, kSrcNegativeIon, kSrcEventCount, kSrcEnergyLoss, kSrcFields };

// AddToField :386-512 on a non-periodic RZ grid from the axis: the order-1
// shape at cell centres, a radial ghost below the axis folded onto the first
// cell, support beyond the other boundaries dropped.
__device__ void source_add(Env const& E, int32_t field, double r, double z, double value)
{
    if (!finite(value) || value == 0.0) return;
    DevFields const& F = E.F;
    int32_t const ic = static_cast<int32_t>(floor((r - F.r_lo) / F.dr));
    int32_t const jc = static_cast<int32_t>(floor((z - F.z_lo) / F.dz));
    if (ic < 0 || ic >= F.nr || jc < 0 || jc >= F.nz) return;
    double const r_cell = (r - F.r_lo) / F.dr + 0.0 - 0.5;
    int32_t const i0 = static_cast<int32_t>(floor(r_cell));
    double const wr1 = r_cell - static_cast<double>(i0);
    double const wr0 = 1.0 - wr1;
    double const z_cell = (z - F.z_lo) / F.dz + 0.0 - 0.5;
    int32_t const j0 = static_cast<int32_t>(floor(z_cell));
    double const wz1 = z_cell - static_cast<double>(j0);
    double const wz0 = 1.0 - wz1;
    double* const dst = E.src + static_cast<int64_t>(field) * F.nr * F.nz;
    for (int rz = 0; rz < 2; ++rz) {
        double const wz = rz == 0 ? wz0 : wz1;
        if (wz == 0.0) continue;
        int32_t const j = j0 + rz;
        if (j < 0 || j > F.nz - 1) continue;
        for (int rr = 0; rr < 2; ++rr) {
            double const wr = rr == 0 ? wr0 : wr1;
            double const v = value * wr * wz;
            if (v == 0.0) continue;
            int32_t i = i0 + rr;
            if (i < 0 || i > F.nr - 1) {
                if (i == -1 && F.r_lo == 0.0) i = 0;
                else continue;
            }
            atomic_add(&dst[i + static_cast<int64_t>(F.nr) * j], v);
        }
    }
}

// RreaRzPositionInHalfOpenDomain: a source outside the physical domain is
// the reference's abort, reported as an outside-domain error
__device__ __forceinline__ bool source_in_domain(Env const& E, double r, double z)
{
    return r >= E.D.r_lo && r < E.D.r_hi && z >= E.D.z_lo && z < E.D.z_hi;
}

__device__ void source_diagnostic(Env const& E, double r, double z, double weight, double loss)
{
    source_add(E, kSrcEventCount, r, z, weight);
    source_add(E, kSrcEnergyLoss, r, z, dmax(weight, 0.0) * dmax(loss, 0.0));
}

// AccumulateIonizationEvent :435-464
__device__ void source_ionization(Env const& E, double cutoff, double r, double z, double weight, double secondary,
                                  double loss, int32_t& err)
{
    if (!E.src || weight <= 0.0) return;
    if (!source_in_domain(E, r, z)) { if (!err) err = kErrDomain; return; }
    source_add(E, kSrcPositiveIon, r, z, weight);
    if (secondary < cutoff) source_add(E, kSrcLowElectron, r, z, weight);
    source_diagnostic(E, r, z, weight, loss);
}

// AccumulatePositiveIonSourceOnly :466-487
__device__ void source_positive_ion(Env const& E, double r, double z, double weight, double loss, int32_t& err)
{
    if (!E.src || weight <= 0.0) return;
    if (!source_in_domain(E, r, z)) { if (!err) err = kErrDomain; return; }
    source_add(E, kSrcPositiveIon, r, z, weight);
    source_diagnostic(E, r, z, weight, loss);
}

// AccumulateLowEnergyElectron :489-516
__device__ void source_low_electron(Env const& E, double r, double z, double weight, double loss, int32_t& err)
{
    if (!E.src || weight <= 0.0) return;
    if (!source_in_domain(E, r, z)) { if (!err) err = kErrDomain; return; }
    source_add(E, kSrcLowElectron, r, z, weight);
    source_diagnostic(E, r, z, weight, dmax(loss, 0.0));
}

struct StepArgs {
    DevTables T;
    Env env;                  // the plain step's environment
    const Env* batch;         // batched step: one environment per instance
    const int32_t* inst;      // batched step: each particle's instance
    rrea_step_config cfg;
    int64_t n;
    const double* prev_r; const double* prev_z; const double* prev_th;
    const double* prev_ux; const double* prev_uy; const double* prev_uz;
    const double* w; int64_t* id; const int32_t* cpu;
    double* r; double* z; double* th; double* ux; double* uy; double* uz;
    uint8_t* status;
    double* ion_pairs;
    int32_t* substeps;
    rrea_secondary* sec; int64_t sec_cap; int64_t* sec_count;
    double* tallies;
    int32_t* error;
};

__device__ __forceinline__ Env const& env_of(StepArgs const& A, int64_t ip)
{
    return A.inst ? A.batch[A.inst[ip]] : A.env;
}

// A FIFO entry's environment: its instance's in a batched drain
__device__ __forceinline__ Env const& env_of_record(StepArgs const& A, int32_t instance)
{
    return (A.batch && instance >= 0) ? A.batch[instance] : A.env;
}

__device__ __forceinline__ void report(int32_t* error, int32_t code, int64_t ip)
{
    if (atomicCAS(&error[0], 0, code) == 0) error[1] = static_cast<int32_t>(ip);
}

__device__ __forceinline__ void push_secondary(StepArgs const& A, rrea_secondary const& s, int32_t& err)
{
    int64_t const slot = static_cast<int64_t>(atomicAdd(reinterpret_cast<unsigned long long*>(A.sec_count), 1ULL));
    if (slot >= A.sec_cap) { if (!err) err = kErrCapacity; return; }
    A.sec[slot] = s;
}

// The transport's own run of RreaValidateSecondaryGroupConservation
// (OV/RreaParticleInteraction.H:1099-1251) on an interaction it just made: a
// group whose charge or energy does not close fails the step with
// RREA_ERR_INVALID, and nothing of the step is committed.
__device__ __forceinline__ void check_group(GroupClosure const& c, double cutoff, int32_t& err)
{
    if (c.verdict(cutoff) != 0 && !err) err = kErrInvalid;
}

// The per-particle side-effect sink of one transaction: local deposits as
// E / W_air ion pairs (deposit_local_energy_as_ion_pairs
// OV/RreaTransportStep.H:537-561), ion-pair events, and secondaries with the
// parent's spawn lineage counter (AddSecondary OV/RreaTransportLog.H:437-447).
struct Emitter {
    StepArgs const& A;
    int64_t ip;
    uint64_t pkey;
    double w;
    double& ion;
    uint64_t& lineage;
    int32_t& err;
    int32_t root_species;   // species of the tracked root of every record made here
    int32_t generation;     // FIFO generation of the records made here
    int32_t instance;       // the simulation instance (-1: a plain step)
    Env const* env;         // where material sources go
    double* ledger;         // the workgroup's event ledger (LDS)

    // deposit_local_energy_as_ion_pairs: E / W_air pairs booked as one
    // ionization event of diagnostic loss W_air at (r, z)
    __device__ void deposit_local(double energy, double r, double z)
    {
        if (w <= 0.0 || energy <= 0.0) return;
        double const pw = w * energy / dmax(kWAir, 1.0);
        if (pw <= 0.0) return;
        ion += pw;
        source_ionization(*env, A.cfg.low_energy_cutoff_eV, r, z, pw, 0.0, kWAir, err);
        Ledger{ledger, 0.0}.transport_ionization(pw, kWAir, true, false);
    }
    __device__ void ionization(double r, double z, double weight, double secondary, double loss)
    {
        source_ionization(*env, A.cfg.low_energy_cutoff_eV, r, z, weight, secondary, loss, err);
    }
    __device__ void positive_ion(double r, double z, double weight, double loss)
    {
        source_positive_ion(*env, r, z, weight, loss, err);
    }
    __device__ void low_electron(double r, double z, double weight, double loss)
    {
        source_low_electron(*env, r, z, weight, loss, err);
    }
    // Out of line: inlined into the photon step, the species word of the
    // record was taken from a register shared across the call sites and came
    // out wrong on some paths (non-deterministic pair-electron records); one
    // out-of-line store sequence per record keeps every field its own value.
    __device__ __noinline__ void secondary(int32_t species, int32_t process, double x, double y, double z, double sdx,
                                           double sdy, double sdz, double energy, double birth)
    {
        rrea_secondary s;
        s.parent_index = ip; s.species = species; s.process = process;
        s.x = x; s.y = y; s.z = z;
        s.dir_x = sdx; s.dir_y = sdy; s.dir_z = sdz;
        s.weight = w; s.energy_eV = energy; s.birth_fraction = birth;
        s.root_key = pkey; s.lineage = lineage++;
        s.root_species = root_species; s.generation = generation;
        s.instance = instance; s.reserved = 0;
        push_secondary(A, s, err);
    }
};

// spawn_positron_annihilation OV/RreaTransportStep.H:563-700: two photons
// (tracked at or above the photon cutoff, else deposited locally), the
// residual energy of an at-rest annihilation as ion pairs, and the positive
// ion the annihilated target electron leaves.  The photons join the enclosing
// group, so they carry its process id.
__device__ void spawn_annihilation(DevTables const& T, double gcut, uint64_t rng, Emitter& em, int32_t process,
                                   double x, double y, double z, double ax, double ay, double az, double k,
                                   bool at_rest, uint64_t index, double birth, double* tally)
{
    if (em.w <= 0.0) return;
    normalize(ax, ay, az);
    double const ka = at_rest ? 0.0 : dmax(k, 0.0);
    if (!(ka == 0.0) && !(finite(ka) && ka >= T.charged_min && ka <= T.charged_max)) {
        if (!em.err) em.err = kErrRange;
        return;
    }
    Annihilation const st = sample_annihilation(ka, rng_uniform(rng, index, kChPositronAnnihilationPolar),
                                                rng_uniform(rng, index, kChPositronAnnihilationAzimuth));
    double g1x, g1y, g1z;
    cone_direction(ax, ay, az, st.cos1, st.az1, g1x, g1y, g1z);
    double const e1 = dmax(st.e1, 0.0);
    double const e2 = dmax(st.e2, 0.0);
    double g2x = -g1x, g2y = -g1y, g2z = -g1z;
    if (!st.at_rest && e2 > 0.0) {
        double const p = sqrt(dmax(0.0, ka * (ka + 2.0 * kMeC2)));
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
    double const sr = hypot(x, y);
    if (at_rest && k > 0.0) em.deposit_local(k, sr, z);
    if (local > 0.0) em.deposit_local(local, sr, z);
    // the annihilated target electron leaves one positive ion
    em.positive_ion(sr, z, em.w, 0.0);
    tally[12] += em.w;
    tally[10] += 1.0;
    Ledger{em.ledger, 0.0}.photon_local_deposit(em.w, local);
    {
        GroupClosure c;
        c.parent(1, em.w, dmax(k, 0.0), 0.0, false);
        if (e1 >= gcut) c.particle(2, em.w, e1);
        if (e2 >= gcut) c.particle(2, em.w, e2);
        c.local(em.w * ((at_rest && k > 0.0 ? k : 0.0) + local));
        c.positive_ion(em.w);
        check_group(c, em.A.cfg.low_energy_cutoff_eV, em.err);
    }
}

// RreaChargedCompetingStpRates OV/RreaChargedFinalStates.H:107-152 (electron)
__device__ __forceinline__ void electron_stp_rates(DevTables const& T, Bracket const& b, double thr, double gcut,
                                                   double e, bool hm_on, double& hm, double& br, int32_t& err,
                                                   double* hm_inv = nullptr)
{
    hm = 0.0;
    br = 0.0;
    if (hm_on && thr > 0.0 && e > 2.0 * thr) {
        // ElectronHardMollerMeanFreePath :2394-2412, then 1/mfp
        double const v = charged_value(T, kEMoller, b, err) * 1.0;
        if (hm_inv) *hm_inv = v;
        double const mfp = v > 0.0 ? 1.0 / v : 1.0e99;
        hm = mfp > 0.0 ? 1.0 / mfp : 0.0;
    }
    if (finite(e) && finite(gcut) && e > gcut) {
        double const v = charged_value(T, kEBrems, b, err) * 1.0;
        double const mfp = v > 0.0 ? 1.0 / v : 1.0e99;
        br = mfp > 0.0 ? 1.0 / mfp : 0.0;
    }
}

// RreaChargedChannelOpenAt OV/RreaChargedFinalStates.H:56-96 (electron)
__device__ __forceinline__ bool electron_channel_open(DevTables const& T, double thr, double gcut, int ch, double e,
                                                      bool hm_on, int32_t& err)
{
    Bracket b;
    if (!(finite(e) && e >= T.charged_min && e <= T.charged_max) || !charged_bracket(T, e, b)) {
        if (!err) err = kErrRange;
        return false;
    }
    if (ch == 0) return hm_on && thr > 0.0 && e > 2.0 * thr && charged_value(T, kEMoller, b, err) * 1.0 > 0.0;
    if (ch == 1) return finite(e) && finite(gcut) && e > gcut && charged_value(T, kEBrems, b, err) * 1.0 > 0.0;
    return false;
}

// elastic_inverse_length_stp OV/RreaTransportStep.H:199-208
__device__ __forceinline__ double elastic_inverse_length(DevElastic const& E, double e, int32_t& err)
{
    double const mfp = elastic_semilog(E, E.mfp, e, err) / 1.0;
    return mfp > 0.0 ? 1.0 / mfp : 0.0;
}

// mean_free_path_from_inverse_length NA/src/RreaInteractionTables.cpp:110-119,
// then RreaChannelInverseLength OV/RreaChargedInterleave.H:69-78
__device__ __forceinline__ double channel_inverse_length(double v)
{
    double const mfp = v > 0.0 ? 1.0 / v : 1.0e99;
    return mfp > 0.0 ? 1.0 / mfp : 0.0;
}

// RreaChargedCompetingStpRates OV/RreaChargedFinalStates.H:114-133 (positron):
// hard Bhabha (open unless K <= T_cut), tracked brems (K > photon cutoff),
// in-flight annihilation (always open)
__device__ __forceinline__ void positron_stp_rates(DevTables const& T, Bracket const& b, double thr, double gcut,
                                                   double e, double& bh, double& br, double& an, int32_t& err)
{
    bh = 0.0;
    br = 0.0;
    if (!(thr > 0.0 && e <= thr)) bh = channel_inverse_length(charged_value(T, kPBhabha, b, err) * 1.0);
    if (e > gcut) br = channel_inverse_length(charged_value(T, kPBrems, b, err) * 1.0);
    an = channel_inverse_length(charged_value(T, kPAnnih, b, err) * 1.0);
}

// RreaChargedChannelOpenAt OV/RreaChargedFinalStates.H:56-96 (positron)
__device__ __forceinline__ bool positron_channel_open(DevTables const& T, double thr, double gcut, int ch, double e,
                                                      int32_t& err)
{
    Bracket b;
    if (!(finite(e) && e >= T.charged_min && e <= T.charged_max) || !charged_bracket(T, e, b)) {
        if (!err) err = kErrRange;
        return false;
    }
    if (ch == 0) return !(thr > 0.0 && e <= thr) && charged_value(T, kPBhabha, b, err) * 1.0 > 0.0;
    if (ch == 1) return e > gcut && charged_value(T, kPBrems, b, err) * 1.0 > 0.0;
    if (ch == 2) return charged_value(T, kPAnnih, b, err) * 1.0 > 0.0;
    return false;
}

// ElectronHardMollerEnergyLoss NA/src/RreaInteractionTables.cpp:2414-2440 at
// unit density ratio: the mean drag of the discrete delta-ray channel
// given the channel's inverse length there (inv); the mean secondary is the
// semilog interpolation on the charged grid, whose t the bracket holds
__device__ __forceinline__ double hard_moller_loss_at(DevTables const& T, Bracket const& b, double inv)
{
    if (!(inv > 0.0)) return 0.0;
    double mean;
    if (b.kind == 1) mean = T.hm_mean[0];
    else if (b.kind == 2) mean = T.hm_mean[T.nc - 1];
    else if (b.kind == 3) mean = T.hm_mean[b.hi];
    else mean = T.hm_mean[b.lo] + b.t * (T.hm_mean[b.hi] - T.hm_mean[b.lo]);
    if (!(mean > 0.0)) return 0.0;
    return mean * inv;
}

__device__ double hard_moller_energy_loss(DevTables const& T, Bracket const& b, double e, int32_t& err)
{
    (void)e;
    return hard_moller_loss_at(T, b, charged_value(T, kEMoller, b, err));
}

// electron_ensemble_mean_drag_ev_per_m OV/RreaTransportStep.H:226-247
__device__ double ensemble_drag(DevTables const& T, double e, double ratio, bool hm_on, int32_t& err)
{
    if (!(e > 0.0) || !(ratio > 0.0)) return 0.0;
    Bracket b;
    if (!charged_bracket(T, e, b)) { if (!err) err = kErrRange; return 0.0; }
    double dd = 0.0;
    if (hm_on) dd = hard_moller_energy_loss(T, b, e, err) * ratio;
    double const realized = charged_value(T, kECol, b, err) * ratio + charged_value(T, kESoft, b, err) * ratio;
    // ElectronTrackedBremsEnergyLoss :2678-2696
    double tb = 0.0;
    double const inv_b = charged_value(T, kEBrems, b, err) * ratio;
    if (inv_b > 0.0) tb = inv_b * cdf_mean(T.ebrems, e, err);
    return realized + dd + tb;
}

// RreaSampleMollerLabDirections OV/RreaChargedInterleave.H:429-463: the
// secondary on a cone of the two-body cosine, the primary on the opposite azimuth
__device__ __forceinline__ void moller_lab_directions(double dx, double dy, double dz, double before, double t2,
                                                      double u_az, double& sdx, double& sdy, double& sdz,
                                                      double& ndx, double& ndy, double& ndz)
{
    double pdx = dx, pdy = dy, pdz = dz;
    normalize(pdx, pdy, pdz);
    double const az = unit_cycle(u_az);
    double const kp = dmax(before, 0.0), ks = dmax(t2, 0.0);
    double const cos_s = (kp <= 0.0 || ks <= 0.0) ? 1.0 : moller_cos(kp, ks);
    double const sc_e = dmax(0.0, before - t2);
    double const cos_p = (before <= 0.0 || sc_e <= 0.0) ? 1.0 : moller_cos(before, sc_e);
    cone_direction(pdx, pdy, pdz, cos_s, az, sdx, sdy, sdz);
    cone_direction(pdx, pdy, pdz, cos_p, unit_cycle(az + 0.5), ndx, ndy, ndz);
}

// SampleBremsstrahlungCosTheta NA/src/RreaInteractionTables.cpp:2714-2760:
// the bounded ModifiedTsai polar angle by 60 bisections
__device__ double brems_cos(double before, double u)
{
    double const cu = dmax(0.0, dmin(1.0, u));
    if (cu == 0.0) return 1.0;
    if (cu == 1.0) return -1.0;
    constexpr double a = 0.625;
    auto g2 = [](double x) {
        if (x < 1.0e-3) {
            return x * x * (0.5 + x * (-1.0 / 3.0 + x * (1.0 / 8.0 + x * (-1.0 / 30.0 + x * (1.0 / 144.0)))));
        }
        return -expm1(-x) - x * exp(-x);
    };
    auto mix = [&](double v) { return 0.25 * g2(a * v) + 0.75 * g2(3.0 * a * v); };
    double const u_max = 2.0 * (1.0 + before / kMeC2);
    double const target = cu * mix(u_max);
    double lo = 0.0, hi = u_max;
    for (int it = 0; it < 60; ++it) {
        double const mid = 0.5 * (lo + hi);
        if (mix(mid) < target) lo = mid; else hi = mid;
    }
    double const f = 0.5 * (lo + hi) / u_max;
    return 1.0 - 2.0 * f * f;
}

// Per-species constants of the charged step.
template <bool kPos> struct Species;
template <> struct Species<false> {
    static constexpr double kCharge = -1.0;
    static constexpr uint64_t kId = 0;
    static constexpr int kCol = kECol, kSoft = kESoft;
    static constexpr uint64_t kDepth0 = kChCompetingOpticalDepth, kPick0 = kChCompetingChannel;
    static constexpr uint64_t kDepth1 = kChContinuationOpticalDepth, kPick1 = kChContinuationChannel;
    static constexpr uint64_t kBremsE = kChBremsPhotonEnergy, kBremsA = kChBremsPhotonAngle;
    static constexpr uint64_t kBremsPhi = kChBremsPhotonAzimuth;
    static constexpr uint64_t kGsAngle = kChGsAngle, kGsPhi = kChGsAzimuth, kFewCount = kChGsFewCount;
    static constexpr uint64_t kFewScatter = kChGsFewScatter, kFewPhi = kChGsFewAzimuth;
};
template <> struct Species<true> {
    static constexpr double kCharge = 1.0;
    static constexpr uint64_t kId = 1;
    static constexpr int kCol = kPCol, kSoft = kPSoft;
    static constexpr uint64_t kDepth0 = kChPositronCompetingOpticalDepth, kPick0 = kChPositronCompetingChannel;
    static constexpr uint64_t kDepth1 = kChPositronContinuationOpticalDepth;
    static constexpr uint64_t kPick1 = kChPositronContinuationChannel;
    static constexpr uint64_t kBremsE = kChPositronBremsPhotonEnergy, kBremsA = kChPositronBremsPhotonAngle;
    static constexpr uint64_t kBremsPhi = kChPositronBremsPhotonAzimuth;
    static constexpr uint64_t kGsAngle = kChPositronGsAngle, kGsPhi = kChPositronGsAzimuth;
    static constexpr uint64_t kFewCount = kChPositronGsFewCount, kFewScatter = kChPositronGsFewScatter;
    static constexpr uint64_t kFewPhi = kChPositronGsFewAzimuth;
};

// The complete per-particle step (every branch of the reference loop body):
// OV/RreaElectronTransport.cpp:59-889 for electrons, OV/RreaPositronTransport.cpp:58-862
// for positrons, with the shared prologue BeginRunningChargedParticle
// (OV/RreaTransportStep.H:720-820) and the substep and event skeletons of
// OV/RreaChargedInterleave.H:546-732.
template <bool kPos>
__device__ __noinline__ void charged_full_particle(StepArgs const& A, int64_t ip, double* tally, double* ledger)
{
    using S = Species<kPos>;
    int32_t err = kOk;
    {
        DevTables const& T = A.T;
        Env const& E = env_of(A, ip);
        DevFields const& F = E.F;
        Domain const& D = E.D;
        rrea_step_config const& cfg = A.cfg;
        double const cutoff = cfg.low_energy_cutoff_eV;
        double const gcut = cfg.photon_cutoff_eV;
        double const thr = T.hard_moller_threshold;
        bool const hm_on = cfg.disable_hard_moller == 0;
        double const dt = cfg.dt_s;
        uint64_t const step = static_cast<uint64_t>(cfg.step > 0 ? cfg.step : 0);
        DevElastic const& EL = kPos ? T.el_p : T.el_e;

        int64_t const id = A.id[ip];
        if (id == RREA_INVALID_ID) {
            // not a valid particle: the reference loop `continue`s
            if (A.status) A.status[ip] = static_cast<uint8_t>(RREA_SKIPPED_INVALID);
            if (A.ion_pairs) A.ion_pairs[ip] = 0.0;
            if (A.substeps) A.substeps[ip] = 0;
            return;
        }
        {
        double const pr = A.prev_r[ip], pz = A.prev_z[ip], pth = A.prev_th[ip];
        double const pux = A.prev_ux[ip], puy = A.prev_uy[ip], puz = A.prev_uz[ip];
        double const w = A.w[ip];
        int32_t const cpu = A.cpu ? A.cpu[ip] : 0;
        uint64_t const pkey = static_cast<uint64_t>(id < 0 ? -id : id)
            ^ (static_cast<uint64_t>(static_cast<uint32_t>(cpu)) << 40);

        double o_r = pr, o_z = pz, o_th = pth, o_ux = pux, o_uy = puy, o_uz = puz;
        int status = RREA_NO_WRITEBACK;
        double ion = 0.0;
        uint64_t subcycles = 0;
        uint64_t lineage = 0;

        // BeginRunningChargedParticle :720-820
        if (w <= 0.0) {
            status = RREA_ZERO_WEIGHT;
        } else if (pr > D.r_hi + 1.0e-9 * D.r_hi) {
            err = kErrDomain;
        } else {
            uint64_t const rng = rng_prefix(E.rng_seed, pkey, step);
            double dx = pux, dy = puy, dz = puz;
            double k = kinetic_from_u(dx, dy, dz);
            normalize(dx, dy, dz);
            double s_th, c_th;
            sin_cos(pth, s_th, c_th);
            double run_x = pr * c_th;
            double run_y = pr * s_th;
            double run_z = pz;
            double const x_particle = run_x, y_particle = run_y;
            double const r_particle = pr, z_particle = pz;

            Path path;
            bool const driver_valid = build_chord(run_x, run_y, run_z, dx, dy, dz, k, dt, dt, D, path, err);
            // An invalid driver chord is the reference's default-constructed path,
            // whose zero domain puts every endpoint outside: it escapes.
            bool escapes = driver_valid ? clipped_escapes(path, D) : true;
            double const ds = driver_valid ? path.in_len : 0.0;
            Interval iv = build_interval(F, path, D, E.density_ratio, err);
            double const span = dmax(path.end_f, 1.0e-6);
            double const field_work = path_field_work(F, path, D, S::kCharge, 0.0, 1.0, err).work / span;
            double const chord_dp = density_path(iv, path, 0.0, path.end_f);
            double particle_ratio = 0.0;
            if constexpr (!kPos) {
                particle_ratio = ds > 0.0 ? chord_dp / ds : positive_or(density_ratio_at_z(F, z_particle, err), E.density_ratio);
            }
            double tr_dp = 0.0, tr_len = 0.0;

            Emitter em{A, ip, pkey, w, ion, lineage, err, kPos ? 1 : 0, 0, E.instance, &E, ledger};
            auto deposit_local = [&](double energy, double r, double z) { em.deposit_local(energy, r, z); };
            auto push = [&](int32_t species, int32_t process, double x, double y, double z, double sdx, double sdy,
                            double sdz, double energy, double birth) {
                em.secondary(species, process, x, y, z, sdx, sdy, sdz, energy, birth);
            };
            auto annihilate = [&](double x, double y, double z, bool at_rest, uint64_t index, double birth) {
                spawn_annihilation(T, gcut, rng, em, kProcAnnihilation, x, y, z, dx, dy, dz, k, at_rest, index, birth,
                                   tally);
            };
            enum { kNone, kEscape, kDemote, kAnnihilated, kSurvive } terminal = kNone;
            double demote_r = r_particle, demote_z = z_particle;

            bool const below0 = k < cutoff;
            if (err) {
                // fall through to the error report
            } else if (escapes && below0) {
                terminal = kEscape;
            } else if (!escapes && below0) {
                if constexpr (kPos) {
                    // at rest where it stands, before any transport: a zero-length net chord
                    annihilate(x_particle, y_particle, z_particle, true, 0U, 1.0);
                    Ledger{ledger, cutoff}.positron_transport(w, k, true);  // OV/RreaPositronTransport.cpp:152
                    terminal = kAnnihilated;
                } else {
                    terminal = kDemote;
                }
            } else if (ds <= 0.0) {
                terminal = escapes ? kEscape : kNone;
            } else if (!iv.valid) {
                err = kErrInvalid;
            } else {
                // Substep sizing, OV/RreaElectronTransport.cpp:187-279 /
                // OV/RreaPositronTransport.cpp:177-245
                Bracket b0;
                bool const b0_ok = charged_bracket(T, k, b0);
                if (!b0_ok) err = kErrRange;
                double s00, s01, s02 = 0.0;
                if constexpr (kPos) positron_stp_rates(T, b0, thr, gcut, k, s00, s01, s02, err);
                else electron_stp_rates(T, b0, thr, gcut, k, hm_on, s00, s01, err);
                double const ddp = chord_dp / span;
                bool violation = false;
                {
                    // selected_subcycles :125-141 + selected_material_subcycles :148-184
                    double const target = cfg.max_optical_depth_per_substep;
                    uint64_t const cap = cfg.max_subcycles;
                    double total_stp;
                    double tau;
                    if constexpr (kPos) {
                        total_stp = s00 + s01 + s02;
                        tau = total_stp * ddp;
                    } else {
                        total_stp = s00 + s01;
                        tau = s00 * ddp + s01 * ddp;
                    }
                    uint64_t const requested = static_cast<uint64_t>(dmax(1.0, ceil(tau / target)));
                    violation = requested > cap;
                    uint64_t n = requested < cap ? requested : cap;
                    for (;;) {
                        double const maximum = total_stp * max_uniform_slice(iv, path, n);
                        double const tol = 64.0 * kEps * dmax(1.0, target);
                        if (maximum <= target + tol) break;
                        // a non-finite rate can only come from a corrupt table: stop at once
                        // rather than walk every count up to the cap
                        if (!finite(maximum)) { violation = true; n = cap; break; }
                        if (n >= cap) { violation = true; break; }
                        // the reference estimates in long double; the estimate only picks the next
                        // trial count and the loop re-tests it, so double is used here
                        double const est = ceil(static_cast<double>(n) * (maximum / target));
                        uint64_t next = n + 1U;
                        if (est >= static_cast<double>(cap)) next = cap;
                        else if (est > static_cast<double>(next)) next = static_cast<uint64_t>(est);
                        n = next < cap ? next : cap;
                    }
                    subcycles = n;
                }
                {
                    double const gs_n_total = elastic_inverse_length(EL, k, err) * ddp;
                    double const kc = dmax(k, cutoff);
                    // kc is k unless k is below the cutoff: then b0 is its
                    // bracket (the same search on the same energy)
                    Bracket bc = b0;
                    if (!(kc == k ? b0_ok : charged_bracket(T, kc, bc))) err = kErrRange;
                    double drag_est;
                    if constexpr (kPos) {
                        // PositronCollisionLoss only
                        drag_est = charged_value(T, kPCol, bc, err) * 1.0 * ddp;
                    } else {
                        drag_est = (kc > 0.0 ? charged_value(T, kECol, bc, err) * 1.0
                                                   + charged_value(T, kESoft, bc, err) * 1.0 : 0.0) * ddp;
                    }
                    double const impulse_scale = dmax(cfg.max_field_impulse_fraction * k,
                                                      dmax(cfg.field_impulse_floor_eV, 1.0));
                    uint64_t const gsd = static_cast<uint64_t>(dmax(1.0, ceil(dmin(gs_n_total, cfg.gs_total_collision_cap)
                                                                            / dmax(cfg.max_gs_collisions_per_substep, 1.0))));
                    uint64_t const imd = static_cast<uint64_t>(dmax(1.0, ceil((fabs(field_work) + drag_est) / impulse_scale)));
                    uint64_t m = subcycles;
                    if (gsd > m) m = gsd;
                    if (imd > m) m = imd;
                    subcycles = m < cfg.max_subcycles ? m : cfg.max_subcycles;
                }
                // the reported per-substep taus: uniform slices of the driver chord over its
                // clipped span (OV/RreaElectronTransport.cpp:250-278, OV/RreaPositronTransport.cpp:241-259)
                if constexpr (kPos) {
                    double const max_positron_tau = (s00 + s01 + s02) * max_uniform_slice(iv, path, subcycles) / span;
                    Ledger{ledger, cutoff}.guard(0.0, 0.0, 0.0, 0.0, max_positron_tau, subcycles, 0.0, 0.0, k, violation);
                } else {
                    double const slice = max_uniform_slice(iv, path, subcycles) / span;
                    Ledger{ledger, cutoff}.guard(s00 * slice, s01 * slice, (s00 + s01) * slice, 0.0, 0.0, subcycles, k, 0.0,
                                                0.0, violation);
                }
                if (cfg.strict_transport_guards && violation && !err) err = kErrGuard;
                tally[6] += static_cast<double>(subcycles);
                double positron_total_loss = 0.0;

                bool exited = false, terminated = false, annihilated = false;
                double mid_x = x_particle, mid_y = y_particle, mid_r = r_particle, mid_z = z_particle;
                double stop_elapsed = 1.0;
                uint64_t completed = 0;
                double const sub_dt = dt / static_cast<double>(subcycles);
                double ex_ = 0.0, ey_ = 0.0, ez_ = 0.0;

                auto elapsed_at = [&](double f) {
                    double const within = dmin(1.0, dmax(0.0, f));
                    return dmin(1.0, (static_cast<double>(completed) + within) / static_cast<double>(subcycles));
                };
                auto mark_stopped = [&](double f) {
                    if (path.exits && f >= path.end_f) { exited = true; return; }
                    double x, y, z;
                    point_at_full(path, f, x, y, z);
                    double const r = hypot(x, y);
                    if (!in_half_open(x, y, z, D)) { exited = true; return; }
                    mid_x = x; mid_y = y; mid_z = z; mid_r = r; terminated = true;
                    if constexpr (kPos) stop_elapsed = elapsed_at(f);
                };
                // apply_electron_continuous_segment OV/RreaElectronTransport.cpp:294-400 /
                // apply_positron_continuous_segment OV/RreaPositronTransport.cpp:278-360 via
                // RreaPlanContinuousDragSegment OV/RreaChargedInterleave.H:239-292
                // the last event-interleave bracket: the next one (the energy
                // after a kick, drag or event) is usually in its interval; the drag segment brackets the same energy
                Bracket hint{0, 0, 1, 0.0, 0.0};
                auto continuous = [&](double b, double e, double stop_e, double& eff) -> bool {
                    eff = b;
                    if (!(e > b) || !(k > 0.0) || stop_e < cutoff) return false;
                    Bracket bs;
                    if (!charged_bracket_near(T, stop_e, hint, bs)) { if (!err) err = kErrRange; return false; }
                    double const col = charged_value(T, S::kCol, bs, err) * 1.0;
                    double const soft = charged_value(T, S::kSoft, bs, err) * 1.0;
                    double const total = col + soft;
                    bool reaches = false;
                    double const to_cut = k - cutoff;
                    double const req = total * density_path(iv, path, b, e);
                    if (!finite(req) || req < 0.0) { if (!err) err = kErrInvalid; return false; }
                    if (req < to_cut) {
                        eff = e;
                    } else {
                        reaches = true;
                        if (to_cut != 0.0) {
                            double lo = b, hi = e;
                            for (int it = 0; it < 56; ++it) {
                                double const mid = 0.5 * (lo + hi);
                                if (total * density_path(iv, path, b, mid) < to_cut) lo = mid; else hi = mid;
                            }
                            eff = 0.5 * (lo + hi);
                        }
                    }
                    double const pm = density_path(iv, path, b, eff);
                    double const raw_col = col * pm;
                    double const raw_soft = soft * pm;
                    double const avail = dmax(0.0, k - cutoff);
                    double const drag = reaches ? avail : dmin(raw_col + raw_soft, avail);
                    double const col_loss = dmin(raw_col, drag);
                    // RreaPartitionContinuousIonPairs OV/RreaChargedInterleave.H:157-196
                    if (col_loss > 0.0 && raw_col > 0.0 && iv.valid) {
                        double const lb = dmax(b, 0.0), le = dmin(eff, iv.end);
                        if (le > lb) {
                            double const raw_i = path.full * (le - lb) * iv.ratio * col;
                            double const loss_i = col_loss * raw_i / raw_col;
                            double const pw = w * loss_i / dmax(kWAir, 1.0);
                            if (pw > 0.0) {
                                ion += pw;
                                double mx, my, mz;
                                point_at_full(path, 0.5 * (lb + le), mx, my, mz);
                                em.ionization(hypot(mx, my), mz, pw, 0.5 * cutoff, kWAir);
                                Ledger{ledger, cutoff}.transport_ionization(pw, kWAir, true, false);
                            }
                        }
                    }
                    tally[7] += w * drag;
                    {
                        // the plan's partition: soft radiative share, and the mean hard-Moller
                        // transfer over the same density path (electrons; OV/RreaElectronTransport.cpp:306-338)
                        Ledger led{ledger, cutoff};
                        led.soft_brems(w, dmax(0.0, drag - col_loss));
                        if constexpr (kPos) {
                            positron_total_loss += drag;
                        } else {
                            double const hstp = (hm_on && thr > 0.0 && stop_e > 2.0 * thr)
                                ? hard_moller_energy_loss(T, bs, stop_e, err) : 0.0;
                            led.loss_partition(w, drag, col_loss, dmin(hstp * pm, avail));
                        }
                    }
                    k = dmax(0.0, k - drag);
                    if (reaches) k = cutoff;
                    return reaches;
                };

                FieldRange sfield{0, 0, 0, 0, 0, 0, 0, 0};
                double substep_dt = 0.0;
                // RreaRunChargedSubstepInterleave OV/RreaChargedInterleave.H:686-732
                for (uint64_t sc = 0; sc < subcycles && !err; ++sc) {
                    if (k < cutoff) break;
                    double sb = 0.0, se = 0.0;
                    {
                        // build_substep
                        {
                            // an invalid chord leaves the previous microchord in place
                            Path next;
                            if (!build_chord(run_x, run_y, run_z, dx, dy, dz, k, sub_dt, dt, D, next, err)) break;
                            path = next;
                        }
                        if (!(path.full > 0.0) || !(path.in_len > 0.0)) { exited = path.exits; break; }
                        iv = build_interval(F, path, D, E.density_ratio, err);
                        if (!iv.valid) { if (!err) err = kErrInvalid; break; }
                        tr_dp += iv.length * iv.ratio;
                        tr_len += iv.length;
                        sb = 0.0;
                        se = path.end_f;
                        sfield = path_field_work(F, path, D, S::kCharge, sb, se, err);
                        substep_dt = sub_dt * (se - sb);
                        tally[9] += w * half_kick(sfield, 0.5 * substep_dt, S::kCharge, k, dx, dy, dz);
                        if (k < cutoff) { mark_stopped(sb); break; }
                    }
                    {
                        // RreaRunChargedEventInterleave OV/RreaChargedInterleave.H:546-664
                        double rem = sb;
                        uint64_t ord = 0;
                        while (rem < se && k >= cutoff && !err) {
                            if (ord >= kEventCap) { err = kErrInvalid; break; }
                            double const frozen_e = k;
                            Bracket bf;
                            if (!charged_bracket_near(T, frozen_e, hint, bf)) { err = kErrRange; break; }
                            hint = bf;
                            double s0, s1, s2 = 0.0;
                            if constexpr (kPos) positron_stp_rates(T, bf, thr, gcut, frozen_e, s0, s1, s2, err);
                            else electron_stp_rates(T, bf, thr, gcut, frozen_e, hm_on, s0, s1, err);
                            double const L = density_path(iv, path, rem, se);
                            if constexpr (!kPos) {
                                double const hard_tau = s0 * L, brems_tau = s1 * L;
                                double const total_tau = hard_tau + brems_tau;
                                Ledger{ledger, cutoff}.hm_expectation(
                                    w * (total_tau > 0.0 ? probability_from_optical_depth(total_tau) * hard_tau / total_tau
                                                         : 0.0));
                            }
                            uint64_t idx = 0;
                            if (!live_hazard_index(sc, ord, idx, S::kId)) { err = kErrInvalid; break; }
                            double const u1 = rng_uniform(rng, idx, ord == 0U ? S::kDepth0 : S::kDepth1);
                            double const u2 = rng_uniform(rng, idx, ord == 0U ? S::kPick0 : S::kPick1);
                            Hazard h;
                            double total_stp;
                            if constexpr (kPos) {
                                h = select_hazard3(s0 * L, s1 * L, s2 * L, u1, u2);
                                total_stp = s0 + s1 + s2;
                            } else {
                                h = select_hazard(s0 * L, s1 * L, u1, u2);
                                total_stp = s0 + s1;
                            }
                            if (!h.valid) { err = kErrInvalid; break; }
                            if (!h.event) {
                                double eff;
                                if (continuous(rem, se, frozen_e, eff)) mark_stopped(eff);
                                break;
                            }
                            double const ef = fraction_at_density_path(iv, path, rem, se, h.sampled / total_stp, err);
                            if (!(finite(rem) && finite(ef) && finite(se) && se > rem && ef >= rem && ef <= se)) {
                                if (!err) err = kErrInvalid;
                                break;
                            }
                            point_at_full(path, ef, ex_, ey_, ez_);
                            {
                                double eff;
                                if (continuous(rem, ef, frozen_e, eff)) { mark_stopped(eff); break; }
                            }
                            bool open_ev, open_fr = false;
                            if constexpr (kPos) {
                                open_ev = k >= cutoff && positron_channel_open(T, thr, gcut, h.channel, k, err);
                                if (!open_ev) {
                                    open_fr = frozen_e >= cutoff
                                        && positron_channel_open(T, thr, gcut, h.channel, frozen_e, err);
                                }
                            } else {
                                open_ev = k >= cutoff && electron_channel_open(T, thr, gcut, h.channel, k, hm_on, err);
                                if (!open_ev) {
                                    open_fr = frozen_e >= cutoff
                                        && electron_channel_open(T, thr, gcut, h.channel, frozen_e, hm_on, err);
                                }
                            }
                            if (!open_ev) {
                                if (!open_fr) { if (!err) err = kErrInvalid; break; }
                                rem = ef;
                                ++ord;
                                continue;
                            }
                            double const efrac = elapsed_at(ef);
                            if (h.channel == 0) {
                                double const before = k;
                                if (!(finite(before) && before >= T.charged_min && before <= T.charged_max)) { err = kErrRange; break; }
                                if constexpr (kPos) {
                                    // RreaApplyBhabhaFinalState OV/RreaChargedFinalStates.H:238-297: a
                                    // secondary below the production threshold discards this event only
                                    double const T2 = dmin(cdf_sample(T.bhabha, before, rng_uniform(rng, idx, kChPositronBhabhaEnergy),
                                                                      false, err),
                                                           dmax(0.0, before));
                                    if (T2 > 0.0 && !(thr > 0.0 && T2 < thr)) {
                                        double sdx, sdy, sdz, ndx, ndy, ndz;
                                        moller_lab_directions(dx, dy, dz, before, T2,
                                                              rng_uniform(rng, idx, kChPositronBhabhaAzimuth),
                                                              sdx, sdy, sdz, ndx, ndy, ndz);
                                        k = dmax(0.0, before - T2);
                                        dx = ndx; dy = ndy; dz = ndz;
                                        push(0, kProcHardBhabha, ex_, ey_, ez_, sdx, sdy, sdz, T2, efrac);
                                        em.positive_ion(hypot(ex_, ey_), ez_, w, T2);
                                        tally[12] += w;
                                        tally[11] += 1.0;
                                        {
                                            Ledger led{ledger, cutoff};
                                            led.transport_ionization(w, T2, false, true);
                                            led.bhabha(w, T2, true);
                                        }
                                        GroupClosure c;
                                        c.parent(1, w, before, k, true);
                                        c.particle(0, w, T2);
                                        c.positive_ion(w);
                                        check_group(c, cutoff, err);
                                    }
                                } else {
                                    // RreaApplyHardMollerFinalState OV/RreaChargedFinalStates.H:164-225
                                    double T2 = cdf_sample(T.moller, before, rng_uniform(rng, idx, kChHardMollerEnergy), true, err);
                                    T2 = dmax(thr, dmin(T2, 0.5 * before));
                                    double sdx, sdy, sdz, ndx, ndy, ndz;
                                    moller_lab_directions(dx, dy, dz, before, T2, rng_uniform(rng, idx, kChHardMollerAzimuth),
                                                          sdx, sdy, sdz, ndx, ndy, ndz);
                                    k = dmax(0.0, before - T2);
                                    dx = ndx; dy = ndy; dz = ndz;
                                    push(0, kProcHardMoller, ex_, ey_, ez_, sdx, sdy, sdz, T2, efrac);
                                    ion += w;
                                    em.ionization(hypot(ex_, ey_), ez_, w, T2, T2);
                                    tally[3] += 1.0;
                                    {
                                        Ledger led{ledger, cutoff};
                                        led.hm_event(w, T2);
                                        led.transport_ionization(w, T2, false, true);
                                    }
                                    GroupClosure c;
                                    c.parent(0, w, before, k, true);
                                    c.particle(0, w, T2);
                                    c.ionization(w, T2, cutoff);
                                    check_group(c, cutoff, err);
                                }
                            } else if (h.channel == 1) {
                                // RreaApplyBremsFinalState OV/RreaChargedFinalStates.H:329-394
                                double const before = k;
                                if (!(finite(before) && before >= T.charged_min && before <= T.charged_max)) { err = kErrRange; break; }
                                double const ks = cdf_sample(kPos ? T.pbrems : T.ebrems, before,
                                                             rng_uniform(rng, idx, S::kBremsE), false, err);
                                double const pe = dmin(dmax(ks, 0.0), dmax(before, 0.0));
                                double const after = dmax(0.0, dmax(before, 0.0) - pe);
                                bool const tracked = finite(gcut) && gcut >= 0.0 && pe >= gcut;
                                double const ct = brems_cos(before, rng_uniform(rng, idx, S::kBremsA));
                                double gx, gy, gz;
                                cone_direction(dx, dy, dz, ct, rng_uniform(rng, idx, S::kBremsPhi), gx, gy, gz);
                                // RreaApplyPhotonRecoilToChargedDirection :465-496
                                {
                                    double const eg = dmax(pe, 0.0);
                                    if (before > 0.0 && eg > 0.0) {
                                        normalize(dx, dy, dz);
                                        double hx = gx, hy = gy, hz = gz;
                                        normalize(hx, hy, hz);
                                        double const p0 = sqrt(dmax(0.0, before * (before + 2.0 * kMeC2)));
                                        double const px = p0 * dx - eg * hx;
                                        double const py = p0 * dy - eg * hy;
                                        double const pz2 = p0 * dz - eg * hz;
                                        double const nn = sqrt(px * px + py * py + pz2 * pz2);
                                        if (nn > 0.0) { dx = px / nn; dy = py / nn; dz = pz2 / nn; normalize(dx, dy, dz); }
                                    }
                                }
                                k = after;
                                if constexpr (kPos) Ledger{ledger, cutoff}.positron_brems(w, pe);
                                else Ledger{ledger, cutoff}.brems_photon(w, pe);
                                if (tracked) {
                                    push(2, kProcBrems, ex_, ey_, ez_, gx, gy, gz, pe, efrac);
                                    tally[5] += 1.0;
                                } else if (pe > 0.0) {
                                    deposit_local(pe, hypot(ex_, ey_), ez_);
                                }
                                tally[4] += 1.0;
                                {
                                    GroupClosure c;
                                    c.parent(kPos ? 1 : 0, w, before, after, true);
                                    if (tracked) c.particle(2, w, pe);
                                    else c.local(w * pe);
                                    check_group(c, cutoff, err);
                                }
                            } else if (kPos && h.channel == 2) {
                                // in-flight annihilation, OV/RreaPositronTransport.cpp:710-745:
                                // the parent is consumed at the event point
                                annihilate(ex_, ey_, ez_, false, idx, efrac);
                                demote_r = hypot(ex_, ey_);
                                demote_z = ez_;
                                annihilated = true;
                                break;
                            }
                            if (k <= cutoff) { mark_stopped(ef); break; }
                            rem = ef;
                            ++ord;
                        }
                    }
                    if (terminated || annihilated || err) break;
                    {
                        // finish_interleaved_substep
                        if (k >= cutoff) {
                            // apply_gs_elastic_scatter OV/RreaTransportStep.H:372-423
                            double const gs_n = elastic_inverse_length(EL, k, err) * density_path(iv, path, sb, se);
                            if (gs_n > 0.0) {
                                if (gs_n < T.gs_few) {
                                    int const cap = static_cast<int>(ceil(3.0 * T.gs_few));
                                    if (cap >= 255) { err = kErrInvalid; break; }
                                    int const cnt = poisson_capped(gs_n, rng_uniform(rng, (sc << 8) | 0U, S::kFewCount), cap);
                                    for (int q = 0; q < cnt; ++q) {
                                        uint64_t const inner = static_cast<uint64_t>(q + 1);
                                        // SampleSingleElasticCosTheta :2546-2560
                                        if (!(finite(k) && k >= T.charged_min && k <= T.charged_max)) { err = kErrRange; break; }
                                        double const eta = elastic_semilog(EL, EL.eta, k, err);
                                        double const u = dmin(dmax(rng_uniform(rng, (sc << 8) | inner, S::kFewScatter), 0.0),
                                                              1.0 - 1.0e-15);
                                        double const ct = 1.0 - 2.0 * eta * u / (1.0 + eta - u);
                                        double nx, ny, nz;
                                        cone_direction(dx, dy, dz, ct, rng_uniform(rng, (sc << 8) | inner, S::kFewPhi), nx, ny, nz);
                                        dx = nx; dy = ny; dz = nz;
                                    }
                                } else {
                                    // SampleGsCosTheta :2504-2544
                                    double const u = rng_uniform(rng, (sc << 8) | 0U, S::kGsAngle);
                                    double ct;
                                    if (!(finite(k) && k >= T.charged_min && k <= T.charged_max)) { err = kErrRange; break; }
                                    DevAngle const& A2 = k <= T.gs_transition ? (kPos ? T.gs_p : T.gs_e)
                                                                               : (kPos ? T.wz_p : T.wz_e);
                                    if (!(k >= A2.e[0] && k <= A2.e[A2.n_e - 1])) { err = kErrRange; break; }
                                    if (gs_n >= T.gs_iso) ct = 2.0 * u - 1.0;
                                    else ct = sample_gs_angle_cos(A2, k, gs_n, u, err);
                                    double nx, ny, nz;
                                    cone_direction(dx, dy, dz, ct, rng_uniform(rng, (sc << 8) | 0U, S::kGsPhi), nx, ny, nz);
                                    dx = nx; dy = ny; dz = nz;
                                }
                            }
                        }
                        if (err) break;
                        tally[9] += w * half_kick(sfield, 0.5 * substep_dt, S::kCharge, k, dx, dy, dz);
                        if (k < cutoff) { mark_stopped(se); break; }
                    }
                    {
                        // RreaAdvanceRunningChargedPosition OV/RreaParticleInteraction.H:368-392
                        double x, y, z;
                        point_at_full(path, se, x, y, z);
                        if (!in_half_open(x, y, z, D)) { exited = true; break; }
                        run_x = x; run_y = y; run_z = z;
                    }
                    if (path.exits) { exited = true; break; }
                    completed = sc + 1U;
                }

                // OV/RreaPositronTransport.cpp:776
                if constexpr (kPos) Ledger{ledger, cutoff}.positron_transport(w, positron_total_loss, false);
                if (!err) {
                    if (exited) escapes = clipped_escapes(path, D);
                    double const end_r = hypot(run_x, run_y);
                    bool const below = k < cutoff;
                    bool resolved = false;
                    if constexpr (kPos) {
                        // OV/RreaPositronTransport.cpp:764-846: a consumed parent outranks
                        // every other outcome; then escape; then a stop or a below-cutoff
                        // endpoint annihilates at rest where the positron ended up
                        resolved = true;
                        if (annihilated) {
                            terminal = kAnnihilated;
                        } else if (exited) {
                            terminal = kEscape;
                        } else if (terminated || below) {
                            double const ax = terminated ? mid_x : run_x;
                            double const ay = terminated ? mid_y : run_y;
                            double const az = terminated ? mid_z : run_z;
                            annihilate(ax, ay, az, true, subcycles, terminated ? stop_elapsed : 1.0);
                            Ledger{ledger, cutoff}.positron_transport(w, k, true);  // OV/RreaPositronTransport.cpp:839
                            demote_r = terminated ? mid_r : end_r;
                            demote_z = az;
                            terminal = kAnnihilated;
                        } else {
                            resolved = false;
                        }
                    } else {
                        if (tr_len > 0.0) particle_ratio = tr_dp / tr_len;
                        bool nonrunaway = false;
                        if (!exited && !terminated && !below && thr > 0.0 && k <= thr) {
                            // RreaEvaluateElectronStatePolicy OV/RreaParticleInteraction.H:1279-1317; the
                            // ensemble drag has no side effect, so it is read only where the gate uses it.
                            double fer, fez, fbt;
                            gather(F, end_r, run_z, fer, fez, fbt);
                            double const thr_v = ensemble_drag(T, dmax(k, cutoff), particle_ratio, hm_on, err);
                            nonrunaway = hypot(fer, fez) <= thr_v;
                        }
                        // RreaResolveChargedTransportTerminal :292-312
                        resolved = true;
                        if (exited) {
                            terminal = kEscape;
                        } else if (terminated || below || nonrunaway) {
                            terminal = kDemote;
                            demote_r = terminated ? mid_r : end_r;
                            demote_z = terminated ? mid_z : run_z;
                        } else {
                            resolved = false;
                        }
                    }
                    if (!resolved) {
                        // CommitSurvivorChargedWriteback OV/RreaTransportStep.H:874-905
                        double const um = proper_velocity(k);
                        o_ux = um * dx; o_uy = um * dy; o_uz = um * dz;
                        o_r = end_r;
                        o_th = atan2(run_y, run_x);
                        o_z = run_z;
                        terminal = kSurvive;
                        if (w != 0.0 && !deposit_segment(E.G, S::kCharge * w, dt, r_particle, z_particle, end_r, run_z)) {
                            err = kErrInvalid;
                        }
                    }
                }
            }
            if (!err) {
                if (terminal == kEscape) {
                    // CommitChargedEscape OV/RreaTransportStep.H:829-863
                    double const um = proper_velocity(k);
                    o_ux = um * dx; o_uy = um * dy; o_uz = um * dz;
                    Crossing const bc = boundary_crossing(path, D);
                    double const q = S::kCharge * w;
                    if (!bc.valid) {
                        err = kErrInvalid;
                    } else if (q != 0.0 && !deposit_escaping_segment(E.G, q, dt, r_particle, z_particle, bc.start_r,
                                                                     bc.start_z, bc.cross_r, bc.cross_z, E.pnr, E.pzb,
                                                                     E.pnz, bc.faces)) {
                        err = kErrInvalid;
                    }
                    status = RREA_ESCAPE;
                    tally[1] += 1.0;
                } else if (terminal == kDemote) {
                    // a low-energy electron source of the parent's weight
                    // (AccumulateLowEnergyElectron, OV/RreaElectronTransport.cpp:152, 857)
                    tally[16] += w;
                    em.low_electron(demote_r, demote_z, w, 0.0);
                    deposit_local(k, demote_r, demote_z);
                    if (w != 0.0 && !deposit_segment(E.G, -w, dt, r_particle, z_particle, demote_r, demote_z)) err = kErrInvalid;
                    status = RREA_DEMOTED;
                    tally[2] += 1.0;
                } else if (terminal == kAnnihilated) {
                    // net chord from the captured start to the annihilation point
                    if (w != 0.0 && !deposit_segment(E.G, w, dt, r_particle, z_particle, demote_r, demote_z)) err = kErrInvalid;
                    status = RREA_ANNIHILATED;
                } else if (terminal == kSurvive) {
                    status = RREA_SURVIVE;
                    tally[0] += 1.0;
                }
            }
        }
        if (err) {
            report(A.error, err, ip);
        } else {
            if (status == RREA_SURVIVE || status == RREA_ESCAPE) {
                A.r[ip] = o_r; A.z[ip] = o_z; A.th[ip] = o_th;
                A.ux[ip] = o_ux; A.uy[ip] = o_uy; A.uz[ip] = o_uz;
            }
            if (status == RREA_ESCAPE || status == RREA_DEMOTED || status == RREA_ZERO_WEIGHT
                || status == RREA_ANNIHILATED) {
                A.id[ip] = RREA_INVALID_ID;  // particle_id.make_invalid()
            }
            if (A.status) A.status[ip] = static_cast<uint8_t>(status);
            if (A.ion_pairs) A.ion_pairs[ip] = ion;
            if (A.substeps) A.substeps[ip] = static_cast<int32_t>(subcycles);
            tally[8] += ion;
        }
        }
    }
}

// The photon final states at an interaction site, shared by the live loop
// and the newborn residual advance: Compton (RreaApplyComptonFinalState
// OV/RreaPhotonFinalStates.H:54-140), photoelectric absorption
// (RreaApplyPhotoelectricFinalState :153-224) and pair production
// (RreaApplyPairFinalState :239-363).  `rng` is the emitter's key prefix and
// `idx` the pass's interaction index.  Returns true when the photon survives
// (Compton at or above the photon cutoff) with its new energy and direction;
// false when it is absorbed or on error (em.err).
__device__ __noinline__ bool photon_final_state(DevTables const& T, double cutoff, double gcut, uint64_t rng,
                                                Emitter& em, int channel, uint64_t idx, double x, double y, double z,
                                                double birth, double& energy, double& dx, double& dy, double& dz,
                                                double* tally)
{
    int32_t& err = em.err;
    double const w = em.w;
    double& ion = em.ion;
    double const sr = hypot(x, y);  // the event site's radius (material sources)
    if (channel == kCompton) {
        double const incident = energy;
        double scattered, recoil, ct;
        sample_compton(incident, rng_uniform(rng, idx, kChComptonRecoil), rng_uniform(rng, idx, kChComptonScatter),
                       scattered, recoil, ct);
        double sx, sy, sz;
        cone_direction(dx, dy, dz, ct, rng_uniform(rng, idx, kChComptonScatterAzimuth), sx, sy, sz);
        double rx = incident * dx - scattered * sx;
        double ry = incident * dy - scattered * sy;
        double rz = incident * dz - scattered * sz;
        normalize(rx, ry, rz);
        if (recoil >= cutoff) {
            em.secondary(0, kProcCompton, x, y, z, rx, ry, rz, recoil, birth);
            ion += w;
            em.ionization(sr, z, w, recoil, recoil);
        } else {
            ion += w;
            em.ionization(sr, z, w, recoil, 0.0);
            em.deposit_local(recoil, sr, z);
        }
        tally[13] += 1.0;
        {
            Ledger led{em.ledger, cutoff};
            led.transport_ionization(w, 0.0, !(recoil >= cutoff), recoil >= cutoff);
            EnergyFlow const f = compton_flow(recoil, scattered, cutoff, gcut);
            led.compton(w, recoil, f.charged, f.local);
        }
        {
            GroupClosure c;
            bool const survives = !(scattered < gcut);
            c.parent(2, w, incident, scattered, survives);
            if (!survives) c.local(w * scattered);
            if (recoil >= cutoff) c.particle(0, w, recoil);
            else c.local(w * recoil);
            c.ionization(w, recoil, cutoff);
            check_group(c, cutoff, err);
        }
        energy = scattered;
        if (energy < gcut) {
            em.deposit_local(energy, sr, z);
            return false;
        }
        dx = sx; dy = sy; dz = sz;
        return true;
    }
    if (channel == kPhotoelectric) {
        double binding;
        if (!sample_photoelectric(T, energy, rng_uniform(rng, idx, kChPhotoelectric), binding, err)) return false;
        double const ke = dmin(dmax(0.0, energy - binding), energy);
        if (ke >= cutoff) {
            if (!(ke == 0.0) && !(finite(ke) && ke >= T.charged_min && ke <= T.charged_max)) {
                err = kErrRange;
                return false;
            }
            double ex, ey, ez;
            cone_direction(dx, dy, dz, sample_photoelectron_cos(ke, rng_uniform(rng, idx, kChPhotoelectricAngle)),
                           rng_uniform(rng, idx, kChPhotoelectricAzimuth), ex, ey, ez);
            em.secondary(0, kProcPhotoelectric, x, y, z, ex, ey, ez, ke, birth);
            ion += w;
            em.ionization(sr, z, w, ke, ke);
        } else {
            ion += w;
            em.ionization(sr, z, w, ke, 0.0);
            em.deposit_local(ke, sr, z);
        }
        Ledger{em.ledger, cutoff}.transport_ionization(w, 0.0, !(ke >= cutoff), ke >= cutoff);
        if (binding > 0.0) em.deposit_local(binding, sr, z);
        tally[14] += 1.0;
        {
            EnergyFlow const f = photoelectric_flow(ke, binding, cutoff);
            Ledger{em.ledger, cutoff}.photoelectric(w, energy, ke, f.charged, f.local);
        }
        GroupClosure c;
        c.parent(2, w, energy, 0.0, false);
        if (ke >= cutoff) c.particle(0, w, ke);
        else c.local(w * ke);
        c.ionization(w, ke, cutoff);
        c.local(w * dmax(binding, 0.0));
        check_group(c, cutoff, err);
        return false;
    }
    PairState const pair = sample_pair(T, energy, channel, rng_uniform(rng, idx, kChPairInvariant),
                                       rng_uniform(rng, idx, kChPairAngle), rng_uniform(rng, idx, kChPairAzimuth), err);
    if (err) return false;
    if (!pair.valid) { err = kErrInvalid; return false; }
    bool const triplet = channel == kPairTriplet;
    if (triplet && !pair.has_recoil) { err = kErrInvalid; return false; }
    int32_t const process = triplet ? kProcPairTriplet : kProcPairNuclear;
    double ex, ey, ez, px, py, pz2;
    rotate_to_axis(dx, dy, dz, pair.edx, pair.edy, pair.edz, ex, ey, ez);
    rotate_to_axis(dx, dy, dz, pair.pdx, pair.pdy, pair.pdz, px, py, pz2);
    if (pair.electron >= cutoff) {
        em.secondary(0, process, x, y, z, ex, ey, ez, pair.electron, birth);
    } else {
        tally[16] += w;
        em.low_electron(sr, z, w, 0.0);
        em.deposit_local(pair.electron, sr, z);
    }
    if (pair.positron >= cutoff) {
        em.secondary(1, process, x, y, z, px, py, pz2, pair.positron, birth);
    } else {
        spawn_annihilation(T, gcut, rng, em, process, x, y, z, px, py, pz2, pair.positron, true, idx, birth, tally);
        Ledger{em.ledger, cutoff}.positron_transport(w, pair.positron, true);  // OV/RreaPhotonFinalStates.H:320
    }
    if (pair.has_recoil) {
        double qx, qy, qz;
        rotate_to_axis(dx, dy, dz, pair.rdx, pair.rdy, pair.rdz, qx, qy, qz);
        if (pair.recoil >= cutoff) em.secondary(0, process, x, y, z, qx, qy, qz, pair.recoil, birth);
        ion += w;
        em.ionization(sr, z, w, pair.recoil, 0.0);
        if (pair.recoil < cutoff) em.deposit_local(pair.recoil, sr, z);
    }
    if (pair.local > 0.0) em.deposit_local(pair.local, sr, z);
    tally[15] += 1.0;
    {
        EnergyFlow const f = pair_flow(pair.electron, pair.positron, pair.recoil, pair.local, cutoff, pair.has_recoil);
        Ledger{em.ledger, cutoff}.pair(w, energy, pair.electron, pair.positron, pair.recoil, f.charged, f.local, triplet);
    }
    {
        // the pair group; a sub-cutoff positron goes on as the particle its
        // at-rest annihilation (checked as its own group) consumes
        GroupClosure c;
        c.parent(2, w, energy, 0.0, false);
        if (pair.electron >= cutoff) c.particle(0, w, pair.electron);
        else { c.low_electron(w); c.local(w * pair.electron); }
        c.particle(1, w, pair.positron);
        if (pair.has_recoil) {
            if (pair.recoil >= cutoff) c.particle(0, w, pair.recoil);
            else c.local(w * pair.recoil);
            c.ionization(w, pair.recoil, cutoff);
        }
        c.local(w * dmax(pair.local, 0.0));
        check_group(c, cutoff, err);
    }
    return false;
}

// One photon: the body of the particle loop of OV/RreaPhotonTransport.cpp:56-438.
// The chord runs from the captured pre-push point to WarpX's pushed point; one
// exponential depth is sampled over the remaining chord per pass, and a
// Compton survivor continues from its interaction point along the scattered
// direction for the unused path length.
__device__ __noinline__ void photon_particle(StepArgs const& A, int64_t ip, double* tally, double* ledger)
{
    int32_t err = kOk;
    DevTables const& T = A.T;
    Env const& E = env_of(A, ip);
    DevFields const& F = E.F;
    Domain const& D = E.D;
    rrea_step_config const& cfg = A.cfg;
    double const cutoff = cfg.low_energy_cutoff_eV;
    double const gcut = cfg.photon_cutoff_eV;
    double const dt = cfg.dt_s;
    uint64_t const step = static_cast<uint64_t>(cfg.step > 0 ? cfg.step : 0);

    int64_t const id = A.id[ip];
    if (id == RREA_INVALID_ID) {
        // not a valid photon: the reference loop `continue`s
        if (A.status) A.status[ip] = static_cast<uint8_t>(RREA_SKIPPED_INVALID);
        if (A.ion_pairs) A.ion_pairs[ip] = 0.0;
        if (A.substeps) A.substeps[ip] = 0;
        return;
    }
    double const pr = A.prev_r[ip], pth = A.prev_th[ip], pz = A.prev_z[ip];
    double const cr = A.r[ip], cth = A.th[ip], cz = A.z[ip];
    double const ux = A.ux[ip], uy = A.uy[ip], uz = A.uz[ip];
    double const w = A.w[ip];
    int32_t const cpu = A.cpu ? A.cpu[ip] : 0;
    uint64_t const pkey = static_cast<uint64_t>(id < 0 ? -id : id) ^ (static_cast<uint64_t>(static_cast<uint32_t>(cpu)) << 40);
    uint64_t const rng = rng_prefix(E.rng_seed, pkey, step);

    double o_r = cr, o_th = cth, o_z = cz, o_ux = ux, o_uy = uy, o_uz = uz;
    bool write_pos = false, write_u = false;
    int status = RREA_SURVIVE;
    double ion = 0.0;
    uint64_t lineage = 0;
    uint64_t events = 0;
    Emitter em{A, ip, pkey, w, ion, lineage, err, 2, 0, E.instance, &E, ledger};

    double energy = sqrt(ux * ux + uy * uy + uz * uz);
    double dx = ux, dy = uy, dz = uz;
    normalize(dx, dy, dz);
    if (pr > D.r_hi + 1.0e-9 * D.r_hi || pz < D.z_lo - 1.0e-9 || pz > D.z_hi + 1.0e-9 * D.z_hi) {
        err = kErrDomain;
    }
    Path path;
    if (!err) clip_rz_path(pr, pth, pz, cr, cth, cz, D, path, err);
    uint64_t ordinal = 0;
    double elapsed_start = 0.0;
    while (!err) {
        if (ordinal >= kEventCap) { err = kErrInvalid; break; }
        bool const escapes = clipped_escapes(path, D);
        double const ds = path.in_len;
        Interval const iv = build_interval(F, path, D, E.density_ratio, err);
        double const chord_dp = density_path(iv, path, 0.0, path.end_f);
        if (err) break;
        if (w <= 0.0 || energy < gcut) {
            // below the production cutoff: deposited where the path starts
            em.deposit_local(energy, hypot(path.sx, path.sy), path.sz);
            status = w <= 0.0 ? RREA_ZERO_WEIGHT : RREA_DEMOTED;
            if (w > 0.0) tally[2] += 1.0;
            break;
        }
        auto commit_no_event = [&]() {
            double x, y, z;
            point_at_full(path, path.end_f, x, y, z);
            o_r = hypot(x, y);
            o_th = atan2(y, x);
            o_z = z;
            write_pos = true;
            if (escapes) {
                status = RREA_ESCAPE;
                tally[1] += 1.0;
                Ledger{ledger, 0.0}.photon_escape(w, energy);  // OV/RreaPhotonTransport.cpp:192
            } else {
                status = RREA_SURVIVE;
                tally[0] += 1.0;
            }
        };
        if (ds <= 0.0) { commit_no_event(); break; }
        PhotonRates rates;
        if (!photon_rates(T, energy, rates, err)) break;
        double const total = rates.compton + rates.pe + rates.pair_n + rates.pair_t;
        // PhotonFeedbackMeanFreePath :2762-2771 via mean_free_path_from_inverse_length :110-119
        if (!finite(total) || total < 0.0) { err = kErrInvalid; break; }
        double const mfp = total > 0.0 ? 1.0 / total : 1.0e99;
        double const tau = chord_dp / mfp;
        // one exponential depth per pass (OV/RreaPhotonTransport.cpp:222-233)
        Ledger{ledger, 0.0}.guard(0.0, 0.0, 0.0, tau, 0.0, 1U, 0.0, energy, 0.0, false);
        uint64_t idx = 0;
        if (!live_hazard_index(0U, ordinal, idx, 2U)) { err = kErrInvalid; break; }
        double const draw = rng_uniform(rng, idx, ordinal == 0U ? kChPhotonInteraction : kChPhotonContinuationOpticalDepth);
        Hazard const h = select_hazard(tau, 0.0, draw, 0.0);
        if (!h.event) { commit_no_event(); break; }
        double const f = fraction_at_density_path(iv, path, 0.0, path.end_f, h.sampled * mfp, err);
        if (err) break;
        double const dist = path.full * (f - 0.0);
        double x, y, z;
        point_at_full(path, f, x, y, z);
        double const remaining = dmax(0.0, path.full - dist);
        double const birth = dmin(1.0, elapsed_start + dist / dmax(kC * dt, kDblMin));
        double const r_site = hypot(x, y);
        // SamplePhotonChannel :2778-2804
        int channel;
        {
            double const cu = dmax(0.0, dmin(1.0, rng_uniform(rng, idx, ordinal == 0U ? kChPhotonChannel
                                                                                     : kChPhotonContinuationChannel)));
            if (!(total > 0.0) || !finite(total)) { err = kErrInvalid; break; }
            double const target = cu * total;
            if (target < rates.compton) channel = kCompton;
            else if (target < rates.compton + rates.pe) channel = kPhotoelectric;
            else if (target < rates.compton + rates.pe + rates.pair_n) channel = kPairNuclear;
            else channel = kPairTriplet;
        }
        events = ordinal + 1U;
        bool const survives = photon_final_state(T, cutoff, gcut, rng, em, channel, idx, x, y, z, birth, energy, dx, dy,
                                                 dz, tally);
        if (err) break;
        if (!survives) {
            status = RREA_ABSORBED;
            break;
        }
        // a Compton survivor continues from the interaction point along its
        // scattered direction for the unused path length
        o_ux = energy * dx; o_uy = energy * dy; o_uz = energy * dz;
        write_u = true;
        double const xt = x + dx * remaining, yt = y + dy * remaining, zt = z + dz * remaining;
        Path cont;
        if (!clip_rz_path(r_site, atan2(y, x), z, hypot(xt, yt), atan2(yt, xt), zt, D, cont, err)) break;
        double xn, yn, zn;
        point_at_full(cont, cont.end_f, xn, yn, zn);
        o_r = hypot(xn, yn);
        o_th = atan2(yn, xn);
        o_z = zn;
        write_pos = true;
        if (!(remaining > 0.0)) {
            status = RREA_SURVIVE;
            tally[0] += 1.0;
            break;
        }
        path = cont;
        elapsed_start = birth;
        ++ordinal;
    }
    if (err) {
        report(A.error, err, ip);
        return;
    }
    if (write_pos) { A.r[ip] = o_r; A.th[ip] = o_th; A.z[ip] = o_z; }
    if (write_u) { A.ux[ip] = o_ux; A.uy[ip] = o_uy; A.uz[ip] = o_uz; }
    if (status == RREA_ESCAPE || status == RREA_DEMOTED || status == RREA_ZERO_WEIGHT || status == RREA_ABSORBED) {
        A.id[ip] = RREA_INVALID_ID;
    }
    if (A.status) A.status[ip] = static_cast<uint8_t>(status);
    if (A.ion_pairs) A.ion_pairs[ip] = ion;
    if (A.substeps) A.substeps[ip] = static_cast<int32_t>(events);
    tally[8] += ion;
}

// workgroup reduction of the tallies, one atomic per tally per workgroup
// Returns thread q's workgroup total of slot q (zero for threads past the slots).
template <int kSlots = kTallyUsed, int kThreads = kBlock>
__device__ __forceinline__ double reduce_tallies(double* tally, double* out)
{
    __shared__ double red[kSlots][kThreads / 64];
    int const lane = threadIdx.x & 63;
    int const wave = threadIdx.x >> 6;
    for (int q = 0; q < kSlots; ++q) {
        double const v = tally[q];
        double s = 0.0;
        if (__ballot(v != 0.0) != 0) s = tally_is_max(q) ? wave_max(v) : wave_sum(v);
        if (lane == 0) red[q][wave] = s;
    }
    __syncthreads();
    double mine = 0.0;
    for (int q = threadIdx.x; q < kSlots; q += kThreads) {
        bool const mx = tally_is_max(q);
        double s = 0.0;
        for (int wv = 0; wv < kThreads / 64; ++wv) {
            double const r = red[q][wv];
            s = mx ? (s < r ? r : s) : s + r;
        }
        if (mx) atomic_max_nonneg(&out[q], s);
        else if (s != 0.0) atomic_add(&out[q], s);
        mine = s;
    }
    return mine;
}

// The workgroup's net-chord deposits.  When the chords of the workgroup's
// survivors lie close together (a cell-ordered tile), the face currents are
// accumulated in an LDS window of the grid and flushed with one device atomic
// per touched face; otherwise every lane deposits with device atomics.  Every
// thread of the workgroup calls this.
// The LDS deposit window of a fast-path workgroup of kB lanes: 8 faces per
// lane (2048 at 256 lanes), so the LDS per wave is the same at every size.
template <int kB> constexpr int window_faces() { return 8 * kB; }
constexpr int kWindowFaces = window_faces<256>();
// the fast-path workgroup sizes the library is built for
constexpr bool fast_block_ok(int b) { return b == 128 || b == 256 || b == 512; }

__device__ __forceinline__ int32_t wave_min_i32(int32_t v)
{
    for (int off = 32; off > 0; off >>= 1) { int32_t const o = __shfl_xor(v, off, 64); v = o < v ? o : v; }
    return v;
}

__device__ __forceinline__ int32_t wave_max_i32(int32_t v)
{
    for (int off = 32; off > 0; off >>= 1) { int32_t const o = __shfl_xor(v, off, 64); v = o > v ? o : v; }
    return v;
}

// A fast-path lane's ion-pair record located on its grid (AddToField's
// order-1 shape: the left cell index and upper weight in r and z); w = 0: none.
struct SourceCell {
    double w, wr1, wz1;
    int32_t i0, j0;
};

__device__ __forceinline__ SourceCell source_cell(Env const& E, double r, double z, double w)
{
    SourceCell c{0.0, 0.0, 0.0, 0, 0};
    DevFields const& F = E.F;
    int32_t const ic = static_cast<int32_t>(floor((r - F.r_lo) / F.dr));
    int32_t const jc = static_cast<int32_t>(floor((z - F.z_lo) / F.dz));
    if (ic < 0 || ic >= F.nr || jc < 0 || jc >= F.nz) return c;  // AddToField keeps nothing outside
    double const r_cell = (r - F.r_lo) / F.dr + 0.0 - 0.5;
    c.i0 = static_cast<int32_t>(floor(r_cell));
    c.wr1 = r_cell - static_cast<double>(c.i0);
    double const z_cell = (z - F.z_lo) / F.dz + 0.0 - 0.5;
    c.j0 = static_cast<int32_t>(floor(z_cell));
    c.wz1 = z_cell - static_cast<double>(c.j0);
    c.w = w;
    return c;
}

// The record's shape into the source fields with device atomics: the low-
// energy electron, positive-ion and event-count fields get w x shape, the
// energy-loss field W_air w x shape (AddToField :386-512 per field).
__device__ void source_cell_add(Env const& E, SourceCell const& c)
{
    int32_t const nr = E.F.nr, nz = E.F.nz;
    int64_t const cells = static_cast<int64_t>(nr) * nz;
    double const v4 = dmax(c.w, 0.0) * dmax(kWAir, 0.0);
    for (int rz = 0; rz < 2; ++rz) {
        double const wz = rz == 0 ? 1.0 - c.wz1 : c.wz1;
        if (wz == 0.0) continue;
        int32_t const j = c.j0 + rz;
        if (j < 0 || j > nz - 1) continue;
        for (int rr = 0; rr < 2; ++rr) {
            double const wr = rr == 0 ? 1.0 - c.wr1 : c.wr1;
            double const v = c.w * wr * wz;
            if (v == 0.0) continue;
            int32_t i = c.i0 + rr;
            if (i < 0 || i > nr - 1) {
                if (i == -1) i = 0;  // the axis fold (the grid starts at r = 0)
                else continue;
            }
            int64_t const cell = i + static_cast<int64_t>(nr) * j;
            atomic_add(&E.src[kSrcLowElectron * cells + cell], v);
            atomic_add(&E.src[kSrcPositiveIon * cells + cell], v);
            atomic_add(&E.src[kSrcEventCount * cells + cell], v);
            atomic_add(&E.src[kSrcEnergyLoss * cells + cell], v4 * wr * wz);
        }
    }
}

// The workgroup's deferred side effects of the fast path, through one LDS
// window at a time: first every survivor's net chord into the J faces, then
// every survivor's continuous ion-pair record (weight sw at (sr, sz); sw = 0:
// none) into the material-source cells, each window flushed with one device
// atomic per touched face or cell.  A window that is too large, or whose
// lanes span several batched instances, falls back to per-lane atomics.
// jr_w, jz_w: the kernel's LDS window (kWindowFaces doubles each), which it
// may use for other per-lane data as long as that is read before this call.
template <int kB>
__device__ __forceinline__ void deposit_block(StepArgs const& A, double* jr_w, double* jz_w, bool dep, int32_t instance, double q,
                              double dt, double r0, double z0, double r1, double z1, SourceCell const& sc)
{
    __shared__ int32_t win[6];
    auto reduce_window = [&](int32_t ilo, int32_t ihi, int32_t jlo, int32_t jhi, int32_t klo, int32_t khi) {
        __syncthreads();  // the previous window's readers are done
        if (threadIdx.x == 0) {
            win[0] = INT32_MAX; win[1] = INT32_MIN; win[2] = INT32_MAX; win[3] = INT32_MIN;
            win[4] = INT32_MAX; win[5] = INT32_MIN;
        }
        ilo = wave_min_i32(ilo);
        ihi = wave_max_i32(ihi);
        jlo = wave_min_i32(jlo);
        jhi = wave_max_i32(jhi);
        klo = wave_min_i32(klo);
        khi = wave_max_i32(khi);
        __syncthreads();
        if ((threadIdx.x & 63) == 0) {
            atomicMin(&win[0], ilo);
            atomicMax(&win[1], ihi);
            atomicMin(&win[2], jlo);
            atomicMax(&win[3], jhi);
            atomicMin(&win[4], klo);
            atomicMax(&win[5], khi);
        }
        __syncthreads();
    };
    // ---- the net chords ----
    {
        // a batched workgroup can hold several instances: the window needs one grid
        DevGrid const& g = instance >= 0 ? A.batch[instance].G : A.env.G;
        int32_t ilo = INT32_MAX, ihi = INT32_MIN, jlo = INT32_MAX, jhi = INT32_MIN;
        if (dep) {
            int32_t a0, a1, b0, b1, c0, c1, d0, d1;
            segment_face_range(g, r0, z0, a0, a1, b0, b1);
            segment_face_range(g, r1, z1, c0, c1, d0, d1);
            ilo = a0 < c0 ? a0 : c0;
            ihi = a1 > c1 ? a1 : c1;
            jlo = b0 < d0 ? b0 : d0;
            jhi = b1 > d1 ? b1 : d1;
        }
        reduce_window(ilo, ihi, jlo, jhi, dep ? instance : INT32_MAX, dep ? instance : INT32_MIN);
        int32_t const i0 = win[0], i1 = win[1], j0 = win[2], j1 = win[3];
        int32_t const wi = i1 - i0 + 1, nw = i0 > i1 ? 0 : wi * (j1 - j0 + 1);
        if (nw > window_faces<kB>() || (nw > 0 && win[4] != win[5])) {
            if (dep) deposit_segment(g, q, dt, r0, z0, r1, z1);
        } else if (nw > 0) {
            // the window's one grid: every lane flushes into it, depositing or not
            DevGrid const& gw = win[4] >= 0 ? A.batch[win[4]].G : A.env.G;
            for (int c = threadIdx.x; c < nw; c += kB) { jr_w[c] = 0.0; jz_w[c] = 0.0; }
            __syncthreads();
            if (dep) deposit_segment_to(gw, LdsWindow{jr_w, jz_w, i0, j0, wi}, q, dt, r0, z0, r1, z1);
            __syncthreads();
            for (int c = threadIdx.x; c < nw; c += kB) {
                int32_t const i = i0 + c % wi, j = j0 + c / wi;
                double const vr = jr_w[c], vz = jz_w[c];
                if (vr != 0.0) atomic_add(&gw.jr[i + (gw.nr + 1) * j], vr);
                if (vz != 0.0) atomic_add(&gw.jz[i + gw.nr * j], vz);
            }
        }
    }
    // ---- the ion-pair records: an ionization event of secondary energy
    // cutoff / 2 and diagnostic loss W_air, so the low-energy electron,
    // positive-ion and event-count fields receive the same shape-weighted
    // value and the energy-loss field W_air times it ----
    {
        bool const has = dep && sc.w > 0.0;
        int32_t const ci0 = sc.i0, cj0 = sc.j0;
        Env const& El = instance >= 0 ? A.batch[instance] : A.env;
        int32_t ilo = INT32_MAX, ihi = INT32_MIN, jlo = INT32_MAX, jhi = INT32_MIN;
        if (has) {
            ilo = ci0 < 0 ? 0 : ci0;
            ihi = ci0 + 1 > El.F.nr - 1 ? El.F.nr - 1 : ci0 + 1;
            jlo = cj0 < 0 ? 0 : cj0;
            jhi = cj0 + 1 > El.F.nz - 1 ? El.F.nz - 1 : cj0 + 1;
        }
        reduce_window(ilo, ihi, jlo, jhi, has ? instance : INT32_MAX, has ? instance : INT32_MIN);
        int32_t const i0 = win[0], i1 = win[1], j0 = win[2], j1 = win[3];
        int32_t const wi = i1 - i0 + 1, nw = i0 > i1 ? 0 : wi * (j1 - j0 + 1);
        if (nw > window_faces<kB>() || (nw > 0 && win[4] != win[5])) {
            if (has) source_cell_add(El, sc);
        } else if (nw > 0) {
            Env const& Ew = win[4] >= 0 ? A.batch[win[4]] : A.env;
            for (int c = threadIdx.x; c < nw; c += kB) { jr_w[c] = 0.0; jz_w[c] = 0.0; }
            __syncthreads();
            if (has) {
                double const v4 = dmax(sc.w, 0.0) * dmax(kWAir, 0.0);
                for (int rz = 0; rz < 2; ++rz) {
                    double const wz = rz == 0 ? 1.0 - sc.wz1 : sc.wz1;
                    if (wz == 0.0) continue;
                    int32_t const j = cj0 + rz;
                    if (j < 0 || j > Ew.F.nz - 1) continue;
                    for (int rr = 0; rr < 2; ++rr) {
                        double const wr = rr == 0 ? 1.0 - sc.wr1 : sc.wr1;
                        double const v = sc.w * wr * wz;
                        if (v == 0.0) continue;
                        int32_t i = ci0 + rr;
                        if (i < 0 || i > Ew.F.nr - 1) {
                            if (i == -1) i = 0;  // the axis fold (the grid starts at r = 0)
                            else continue;
                        }
                        int32_t const c = (i - i0) + wi * (j - j0);
                        atomicAdd(&jr_w[c], v);
                        atomicAdd(&jz_w[c], v4 * wr * wz);
                    }
                }
            }
            __syncthreads();
            int64_t const cells = static_cast<int64_t>(Ew.F.nr) * Ew.F.nz;
            for (int c = threadIdx.x; c < nw; c += kB) {
                int64_t const cell = (i0 + c % wi) + static_cast<int64_t>(Ew.F.nr) * (j0 + c / wi);
                double const v = jr_w[c], v4 = jz_w[c];
                if (v != 0.0) {
                    atomic_add(&Ew.src[kSrcLowElectron * cells + cell], v);
                    atomic_add(&Ew.src[kSrcPositiveIon * cells + cell], v);
                    atomic_add(&Ew.src[kSrcEventCount * cells + cell], v);
                }
                if (v4 != 0.0) atomic_add(&Ew.src[kSrcEnergyLoss * cells + cell], v4);
            }
        }
    }
}

// The particles the fast path handed over, strided over a fixed grid.
// One wavefront per workgroup: a hand-over wave (~290 registers) is placed
// on any SIMD with room for it, not only where a whole CU's four SIMDs are
// free, while the other species' passes fill the chip.
constexpr int kSlowThreads = 64;

template <bool kPos>
__global__ void __launch_bounds__(kSlowThreads) charged_slow_kernel(const StepArgs* __restrict__ args,
                                                            const int32_t* __restrict__ list,
                                                            int32_t* __restrict__ count)
{
    StepArgs const& A = *args;
    __shared__ double ledger[RREA_LEDGER_COUNT];
    ledger_begin(ledger);
    double tally[kTallyUsed];
    for (int q = 0; q < kTallyUsed; ++q) tally[q] = 0.0;
    // a work queue over the two classes of append_slow_split.  Each class is
    // cut into chunks of about half a wavefront's share of it (so a heavy
    // chunk holds a few divergent paths, not 64), heavy chunks numbered first
    // so that the long paths start earliest; count[2] hands out the chunks
    // after each wavefront's first, and every wavefront ends when they run
    // out.
    static_assert(kSlowThreads == 64, "one wavefront per workgroup");
    int64_t const cl = count[0], ch = count[1];
    int64_t const waves = gridDim.x, lane = threadIdx.x;
    int64_t lh = (2 * ch + waves - 1) / waves, ll = (2 * cl + waves - 1) / waves;
    lh = lh < 1 ? 1 : (lh > 64 ? 64 : lh);
    ll = ll < 1 ? 1 : (ll > 64 ? 64 : ll);
    int64_t const hc = (ch + lh - 1) / lh, lc = (cl + ll - 1) / ll;
    // the first chunk is the wavefront's own index, later ones from the queue
    for (int64_t c = blockIdx.x; c < hc + lc;) {
        if (c < hc) {
            int64_t const e = c * lh + lane;
            if (lane < lh && e < ch) charged_full_particle<kPos>(A, list[A.n - 1 - e], tally, ledger);
        } else {
            int64_t const e = (c - hc) * ll + lane;
            if (lane < ll && e < cl) charged_full_particle<kPos>(A, list[e], tally, ledger);
        }
        int32_t q = 0;
        if (lane == 0) q = atomicAdd(count + 2, 1);
        c = waves + __shfl(q, 0, 64);
    }
    reduce_tallies<kTallyUsed, kSlowThreads>(tally, A.tallies);
    ledger_flush(ledger, A.tallies);
}

// The workgroup fold of a fast path's per-lane values, held in LDS columns
// (value q of lane t at col[q * kBlock + t], kQ <= 16 columns): the first
// kSums columns are sums, the rest running maxima, and emit(q, total) books
// column q's workgroup total in one thread.  Sixteen lanes reduce a column,
// sixteen entries each (consecutive lanes read consecutive entries), then
// combine within their group of sixteen.  The whole workgroup calls this once
// its columns are written.
template <int kQ, int kSums, int kB, class Emit>
__device__ __forceinline__ void fold_columns(double const* col, Emit&& emit)
{
    // kB / 16 lanes per column (a group within one wavefront), each summing
    // sixteen entries
    constexpr int kL = kB / 16;
    static_assert(kQ <= 16 && kL <= 64 && (kL & (kL - 1)) == 0, "sixteen column groups of a power of two");
    __syncthreads();
    int const q = threadIdx.x / kL;
    int const part = threadIdx.x % kL;
    bool const mx = q >= kSums;
    double s = 0.0;
    if (q < kQ) {
        double const* c = col + q * kB + part;
        for (int i = 0; i < 16; ++i) {
            double const v = c[i * kL];
            s = mx ? (s < v ? v : s) : s + v;
        }
    }
    for (int off = kL / 2; off > 0; off >>= 1) {
        double const o = __shfl_xor(s, off, 64);
        s = mx ? (s < o ? o : s) : s + o;
    }
    if (q < kQ && part == 0) emit(q, s);
}

// The columns of the fast paths' per-lane values.  Charged: the first
// kColSums are sums, then the guard taus and the transported energy (maxima);
// kColHazardTotal is a lane's scratch column, not folded.
template <bool kPos>
struct ChargedColumns {
    enum {
        kColSurvivor, kColSubstep, kColDrag, kColIon, kColFieldWork, kColSoft,
        kColCol = kColSoft + 1, kColHard = kColSoft + 2, kColHmx = kColSoft + 3,  // electrons only
        kColSums = kPos ? kColSoft + 1 : kColSoft + 4,
        kColTauA = kColSums, kColTauB = kColSums + 1, kColTauTot = kColSums + 2,  // positrons: A only
        kColEnergy = kPos ? kColSums + 1 : kColSums + 3,
        kCols = kColEnergy + 1,
        // the hazard's total tau, with its hard-Moller tau in kColHmx until
        // the epilogue forms the expected weight (electrons)
        kColHazardTotal = kCols,
    };
};
struct PhotonColumns {
    enum { kColSurvivor, kColEscape, kColEscapedEnergy, kColTau, kColEnergy, kCols, kColSums = kColTau };
};

// Books column q's total v of a charged fast-path fold: the tallies and the
// committed steps' ledger records (OV/RreaElectronTransport.cpp:250-400,
// OV/RreaPositronTransport.cpp:241-360, 776).
template <bool kPos>
__device__ void book_charged_column(double* tallies, int q, double v)
{
    using Col = ChargedColumns<kPos>;
    double* const led = tallies + RREA_LEDGER_OFFSET;
    if (q < Col::kColSums) {
        if (v == 0.0) return;
        switch (q) {
        case Col::kColSurvivor: atomic_add(&tallies[0], v); break;
        case Col::kColSubstep: atomic_add(&tallies[6], v); break;
        case Col::kColDrag:
            atomic_add(&tallies[7], v);
            atomic_add(led + (kPos ? RREA_L_POSITRON_TRANSPORT_ENERGY_LOSS_EV : RREA_L_PARENT_DRAG_LOSS_EV), v);
            break;
        case Col::kColIon:
            // every ion-pair piece is a low-energy ionization record of diagnostic loss W_air
            atomic_add(&tallies[8], v);
            atomic_add(led + RREA_L_TRANSPORT_IONIZATION_WEIGHT, v);
            atomic_add(led + RREA_L_TRANSPORT_LOW_SECONDARY_WEIGHT, v);
            atomic_add(led + RREA_L_TRANSPORT_ENERGY_LOSS_EV, v * kWAir);
            break;
        case Col::kColFieldWork: atomic_add(&tallies[9], v); break;
        case Col::kColSoft: atomic_add(led + RREA_L_PHOTON_CONTINUOUS_SOFT_BREMS_ENERGY_EV, v); break;
        default:
            if constexpr (!kPos) {
                if (q == Col::kColCol) {
                    atomic_add(led + RREA_L_ION_PAIR_SOURCE_LOSS_EV, v);
                    atomic_add(led + RREA_L_RESTRICTED_COLLISION_LOSS_EV, v);
                } else if (q == Col::kColHard) {
                    atomic_add(led + RREA_L_EXPLICIT_HARD_MOLLER_ENERGY_LOSS_EV, v);
                } else if (q == Col::kColHmx) {
                    atomic_add(led + RREA_L_HARD_MOLLER_EXPECTED_EVENT_WEIGHT, v);
                }
            }
            break;
        }
        return;
    }
    if (q == Col::kColEnergy) {
        atomic_max_nonneg(led + (kPos ? RREA_L_MAX_TRANSPORT_POSITRON_ENERGY_EV
                                      : RREA_L_MAX_TRANSPORT_ELECTRON_ENERGY_EV), v);
        if (v > 0.0) atomic_max_nonneg(led + RREA_L_MAX_TRANSPORT_SUBCYCLES, 1.0);
    } else if constexpr (kPos) {
        atomic_max_nonneg(led + RREA_L_MAX_TRANSPORT_POSITRON_TAU, v);
    } else {
        atomic_max_nonneg(led + (q == Col::kColTauA ? RREA_L_MAX_TRANSPORT_HARD_MOLLER_TAU
                                 : q == Col::kColTauB ? RREA_L_MAX_TRANSPORT_BREMS_TAU
                                                 : RREA_L_MAX_TRANSPORT_ELECTRON_TOTAL_COMPETING_TAU), v);
    }
}

// Books column q's total v of a photon fast-path fold (OV/RreaPhotonTransport.cpp:192, 222-233).
__device__ void book_photon_column(double* tallies, int q, double v)
{
    using Col = PhotonColumns;
    double* const led = tallies + RREA_LEDGER_OFFSET;
    if (q == Col::kColSurvivor && v != 0.0) atomic_add(&tallies[0], v);
    if (q == Col::kColEscape && v != 0.0) atomic_add(&tallies[1], v);
    if (q == Col::kColEscapedEnergy && v != 0.0) atomic_add(led + RREA_L_PHOTON_ESCAPED_ENERGY_EV, v);
    if (q == Col::kColTau) atomic_max_nonneg(led + RREA_L_MAX_TRANSPORT_PHOTON_TAU, v);
    if (q == Col::kColEnergy) {
        atomic_max_nonneg(led + RREA_L_MAX_TRANSPORT_PHOTON_ENERGY_SEEN_EV, v);
        if (v > 0.0) atomic_max_nonneg(led + RREA_L_MAX_TRANSPORT_SUBCYCLES, 1.0);
    }
}

// A fast-path workgroup's folded columns go to row blockIdx.x of the
// partials (one 128-byte line per workgroup, every workgroup of the launch
// writing its columns), and a fold kernel books the rows into the tallies and
// ledger: same-address atomics from every workgroup of a launch would
// serialise on the tally lines, where the rows take plain stores.
constexpr int kPartialStride = 16;
static_assert(ChargedColumns<false>::kCols <= kPartialStride && PhotonColumns::kCols <= kPartialStride,
              "a workgroup's columns exceed its partial row");

// The fold of `rows` partial rows: sixteen rows per workgroup pass over a grid
// stride (sixteen lanes a row, consecutive lanes consecutive columns), the
// wave's row groups and then the workgroup's waves combined, and book(q, v)
// called once per nonzero column total of the workgroup.
template <class Col, class Book>
__device__ __forceinline__ void fold_partial_rows(const double* __restrict__ partials, int64_t rows, Book&& book)
{
    constexpr int kRowsPerPass = kBlock / kPartialStride;
    int const q = threadIdx.x & (kPartialStride - 1);
    bool const mx = q >= Col::kColSums;
    double s = 0.0;
    if (q < Col::kCols) {
        int64_t const stride = static_cast<int64_t>(gridDim.x) * kRowsPerPass;
        for (int64_t r = static_cast<int64_t>(blockIdx.x) * kRowsPerPass + threadIdx.x / kPartialStride; r < rows;
             r += stride) {
            double const v = partials[r * kPartialStride + q];
            s = mx ? (s < v ? v : s) : s + v;
        }
    }
    for (int off = kPartialStride; off < 64; off <<= 1) {
        double const o = __shfl_xor(s, off, 64);
        s = mx ? (s < o ? o : s) : s + o;
    }
    __shared__ double wave_total[kBlock / 64][kPartialStride];
    if ((threadIdx.x & 63) < kPartialStride) wave_total[threadIdx.x >> 6][threadIdx.x & 63] = s;
    __syncthreads();
    if (threadIdx.x < Col::kCols) {
        double t = wave_total[0][threadIdx.x];
        for (int w = 1; w < kBlock / 64; ++w) {
            double const o = wave_total[w][threadIdx.x];
            t = mx ? (t < o ? o : t) : t + o;
        }
        if (t != 0.0) book(static_cast<int>(threadIdx.x), t);
    }
}

template <bool kPos>
__global__ void __launch_bounds__(kBlock) charged_fold_kernel(const StepArgs* __restrict__ args,
                                                              const double* __restrict__ partials, int64_t rows)
{
    double* const tallies = args->tallies;
    fold_partial_rows<ChargedColumns<kPos>>(partials, rows,
                                            [&](int q, double v) { book_charged_column<kPos>(tallies, q, v); });
}

__global__ void __launch_bounds__(kBlock) photon_fold_kernel(const StepArgs* __restrict__ args,
                                                             const double* __restrict__ partials, int64_t rows)
{
    double* const tallies = args->tallies;
    fold_partial_rows<PhotonColumns>(partials, rows, [&](int q, double v) { book_photon_column(tallies, q, v); });
}

// ---------------------------------------------------------------------------
// The fast path.  The dominant particle-step of the benchmark configurations
// -- one substep, no discrete event, at most two single elastic scatters, no
// boundary, survivor above the demotion gate -- is the same arithmetic as the
// full path restricted to that branch; with one substep the substep chord is
// bit-identical to the driver chord (same state, dt / 1 == dt), so its
// geometry, density and field work are formed once.  Any other outcome is
// detected before a single side effect and the particle index is appended to
// the slow list, which the complete path then finishes from the unchanged
// inputs.
// ---------------------------------------------------------------------------
// The charged fast paths hand over in two classes: light lanes (a particle
// below the cutoff from the start, an electron ending the step at or below
// the demotion gate: a short complete path) from the front of the list with
// count[0], the rest from the back (list[cap - 1 - j]) with count[1], so the
// hand-over kernel can give each class its own wavefronts.
__device__ __forceinline__ void append_slow_split(bool slow, bool light, int64_t ip, int32_t* list, int32_t* count,
                                                  int64_t cap)
{
    int const lane = threadIdx.x & 63;
    uint64_t const below = (uint64_t{1} << lane) - 1U;
    for (int cls = 0; cls < 2; ++cls) {
        bool const mine = slow && (light == (cls == 0));
        uint64_t const mask = __ballot(mine);
        if (mask == 0) continue;
        int const leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
        int32_t base = 0;
        if (lane == leader) base = atomicAdd(count + cls, static_cast<int32_t>(__popcll(mask)));
        base = __shfl(base, leader, 64);
        if (mine) {
            int64_t const j = base + __popcll(mask & below);
            list[cls == 0 ? j : cap - 1 - j] = static_cast<int32_t>(ip);
        }
    }
}

__device__ __forceinline__ void append_slow(bool slow, int64_t ip, int32_t* list, int32_t* count)
{
    uint64_t const mask = __ballot(slow);
    if (mask == 0) return;
    int const lane = threadIdx.x & 63;
    int const leader = __ffsll(static_cast<unsigned long long>(mask)) - 1;
    int32_t base = 0;
    if (lane == leader) base = atomicAdd(count, static_cast<int32_t>(__popcll(mask)));
    base = __shfl(base, leader, 64);
    if (slow) {
        uint64_t const below = mask & ((uint64_t{1} << lane) - 1U);
        list[base + __popcll(below)] = static_cast<int32_t>(ip);
    }
}

// four waves per SIMD (128 VGPRs): the plain kernels fit without spilling;
// the batched ones spill 60 bytes per lane and still run 7 % faster than at
// three waves (143 VGPRs, no spill): 7.8 against 8.4 ms per launch over the
// 256-instance batch
template <bool kPos, bool kBatched, int kB = 256>
__global__ void __launch_bounds__(kB) __attribute__((amdgpu_waves_per_eu(4)))
charged_fast_kernel(const StepArgs* __restrict__ args, int32_t* __restrict__ slow_list,
                    int32_t* __restrict__ slow_count, double* __restrict__ partials)
{
    using S = Species<kPos>;
    StepArgs const& A = *args;
    int64_t const ip = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    bool slow = false;
    bool light = false;  // a short complete path: below the cutoff from the start, or at the demotion gate
    // the survivor's net chord, deposited by the whole workgroup at the end;
    // the chord and its ion-pair source record are set with dep and read only
    // where it is set, so they carry no value on the paths that leave the
    // fast branch early (no zeroing of them at each of its exits)
    bool dep = false;
    int32_t dep_inst = -1;
    double dep_q, dep_r0, dep_z0, dep_r1, dep_z1;
    SourceCell src;
    // The LDS window of the deposits at the end; until then it holds each
    // lane's tallies and ledger records in columns of its own (column q of
    // the lane at window[q * kB + lane]), folded per workgroup at the end:
    //   sums -- the tallies of a committed step (survivor, substep, drag, ion
    //     pairs, field work), and its soft radiative loss, collision loss,
    //     mean hard-Moller transfer and expected hard-Moller weight, written as
    //     they are formed and cleared for a lane handed over;
    //   maxima -- the guard's per-substep taus and the transported energy of a
    //     lane whose drivers all asked for one subcycle (a running maximum
    //     takes the same value again if the complete path later steps the
    //     particle, which it sizes the same way, so recording them before the
    //     commit is exact).
    // Past the window, three more columns hold the lane's start point and
    // weight while the chord's field work and the kicks need the registers;
    // they are read back (relaxed LDS loads, which the compiler does not fold
    // into the stores) where the drag, the commit and the deposit use them.
    __shared__ double window[2 * window_faces<kB>() + 3 * kB];
    using Col = ChargedColumns<kPos>;
    static_assert((Col::kCols + 1) * kB <= 2 * window_faces<kB>(), "the columns exceed the deposit window");
    double* const lane_col = window + threadIdx.x;
    double* const stash = window + 2 * window_faces<kB>() + threadIdx.x;
    auto unstash = [&](int q) { return __hip_atomic_load(stash + q * kB, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP); };
    for (int q = 0; q < Col::kCols + 1; ++q) lane_col[q * kB] = 0.0;
    // a batched wave takes the fast path only when all its particles share one
    // instance, whose environment is then wave-uniform; a wave that straddles
    // an instance boundary goes to the complete path
    int32_t wave_inst = 0;
    bool mixed = false;
    if constexpr (kBatched) {
        int32_t const mine = ip < A.n ? A.inst[ip] : -1;
        int32_t const first = __builtin_amdgcn_readfirstlane(mine);
        mixed = __ballot(ip < A.n && mine != first) != 0;
        wave_inst = first < 0 ? 0 : first;
    }
    if (ip < A.n) {
        slow = true;  // cleared only when the whole fast branch completes
        DevTables const& T = A.T;
        Env const* __restrict__ const envs = A.batch;
        Env const& E = kBatched ? envs[wave_inst] : A.env;
        DevFields const& F = E.F;
        Domain const& D = E.D;
        rrea_step_config const& cfg = A.cfg;
        double const cutoff = cfg.low_energy_cutoff_eV;
        double const gcut = cfg.photon_cutoff_eV;
        double const thr = T.hard_moller_threshold;
        bool const hm_on = cfg.disable_hard_moller == 0;
        double const dt = cfg.dt_s;
        DevElastic const& EL = kPos ? T.el_p : T.el_e;
        int32_t err = kOk;
        int64_t const id = A.id[ip];
        if (kBatched && mixed) {
            // handed over as a whole wave
        } else if (id == RREA_INVALID_ID) {
            if (A.status) A.status[ip] = static_cast<uint8_t>(RREA_SKIPPED_INVALID);
            if (A.ion_pairs) A.ion_pairs[ip] = 0.0;
            if (A.substeps) A.substeps[ip] = 0;
            slow = false;
        } else do {
            double const w = A.w[ip];
            double const pr = A.prev_r[ip];
            if (!(w > 0.0) || pr > D.r_hi + 1.0e-9 * D.r_hi) break;
            double const pz = A.prev_z[ip], pth = A.prev_th[ip];
            stash[0] = pr;
            stash[kB] = pz;
            stash[2 * kB] = w;
            double dx = A.prev_ux[ip], dy = A.prev_uy[ip], dz = A.prev_uz[ip];
            double k = kinetic_from_u(dx, dy, dz);
            normalize(dx, dy, dz);
            if (!(k >= cutoff)) { light = true; break; }
            double s_th, c_th;
            sin_cos(pth, s_th, c_th);
            double run_x = pr * c_th;
            double run_y = pr * s_th;
            double run_z = pz;
            Path path;
            if (!build_chord(run_x, run_y, run_z, dx, dy, dz, k, dt, dt, D, path, err) || err) break;
            if (clipped_escapes(path, D) || !(path.in_len > 0.0)) break;
            Interval const iv = build_interval(F, path, D, E.density_ratio, err);
            if (!iv.valid || err) break;
            double const span = dmax(path.end_f, 1.0e-6);
            // batched: the first GL node's products wait in the (still empty)
            // ledger columns, which keeps that kernel from spilling them
            static_assert(Col::kCols >= 6, "the hold columns are the ledger's");
            FieldRange const fr = path_field_work<kBatched, kB>(F, path, D, S::kCharge, 0.0, 1.0, err, lane_col);
            if (err) break;
            double const field_work = fr.work / span;
            double const chord_dp = density_path(iv, path, 0.0, path.end_f);
            // substep sizing: all three drivers must ask for one substep
            Bracket b0;
            if (!charged_bracket(T, k, b0)) break;
            double const ddp = chord_dp / span;
            double const target = cfg.max_optical_depth_per_substep;
            double tau0, total0;
            double g_a = 0.0, g_b = 0.0;
            if constexpr (kPos) {
                double bh0, br0, an0;
                positron_stp_rates(T, b0, thr, gcut, k, bh0, br0, an0, err);
                total0 = bh0 + br0 + an0;
                tau0 = total0 * ddp;
            } else {
                double hm0, br0;
                electron_stp_rates(T, b0, thr, gcut, k, hm_on, hm0, br0, err);
                total0 = hm0 + br0;
                tau0 = hm0 * ddp + br0 * ddp;
                g_a = hm0;
                g_b = br0;
            }
            if (dmax(1.0, ceil(tau0 / target)) != 1.0) break;
            double const mus = max_uniform_slice(iv, path, 1);
            if (!(total0 * mus <= target + 64.0 * kEps * dmax(1.0, target))) break;
            // the guard's taus for one subcycle; they count only if every driver
            // below asks for one subcycle too, which the energy column marks
            if constexpr (kPos) {
                lane_col[Col::kColTauA * kB] = total0 * mus / span;
            } else {
                double const slice = mus / span;
                lane_col[Col::kColTauA * kB] = g_a * slice;
                lane_col[Col::kColTauB * kB] = g_b * slice;
                lane_col[Col::kColTauTot * kB] = (g_a + g_b) * slice;
            }
            EBracket eb0;
            if (!elastic_bracket(EL, k, eb0)) break;
            double const mfp0 = elastic_value(EL, EL.mfp, eb0) / 1.0;
            double const gs_n_total = (mfp0 > 0.0 ? 1.0 / mfp0 : 0.0) * ddp;
            // the driver's drag estimate: collision + soft (e-), collision only (e+)
            double const drag_est = kPos ? charged_value(T, kPCol, b0, err) * 1.0 * ddp
                                         : (charged_value(T, kECol, b0, err) * 1.0
                                            + charged_value(T, kESoft, b0, err) * 1.0) * ddp;
            double const impulse_scale = dmax(cfg.max_field_impulse_fraction * k, dmax(cfg.field_impulse_floor_eV, 1.0));
            if (dmax(1.0, ceil(dmin(gs_n_total, cfg.gs_total_collision_cap) / dmax(cfg.max_gs_collisions_per_substep, 1.0))) != 1.0) break;
            if (dmax(1.0, ceil((fabs(field_work) + drag_est) / impulse_scale)) != 1.0) break;
            if (err) break;
            // every driver asked for one subcycle: the guard's record of this step
            lane_col[Col::kColEnergy * kB] = k;
            // the single substep: the driver chord again (sub_dt = dt / 1)
            double const se = path.end_f;
            double const substep_dt = dt * (se - 0.0);
            double k1 = k, d1x = dx, d1y = dy, d1z = dz;
            double const hw1 = half_kick(fr, 0.5 * substep_dt, S::kCharge, k1, d1x, d1y, d1z);
            if (!(k1 >= cutoff) || !(0.0 < se)) break;
            // the first hazard attempt: no event
            Bracket b1;
            if (!charged_bracket_near(T, k1, b0, b1)) break;
            double const L = density_path(iv, path, 0.0, se);
            uint64_t const rng = rng_prefix(E.rng_seed, static_cast<uint64_t>(id < 0 ? -id : id)
                                                ^ (static_cast<uint64_t>(static_cast<uint32_t>(A.cpu ? A.cpu[ip] : 0)) << 40),
                                            static_cast<uint64_t>(cfg.step > 0 ? cfg.step : 0));
            Hazard h;
            double hm_inv1 = 0.0;
            if constexpr (kPos) {
                double s0, s1, s2;
                positron_stp_rates(T, b1, thr, gcut, k1, s0, s1, s2, err);
                h = select_hazard3(s0 * L, s1 * L, s2 * L, rng_uniform(rng, 0, S::kDepth0), rng_uniform(rng, 0, S::kPick0));
            } else {
                double s0, s1;
                electron_stp_rates(T, b1, thr, gcut, k1, hm_on, s0, s1, err, &hm_inv1);
                double const hard_tau = s0 * L, brems_tau = s1 * L;
                h = select_hazard(hard_tau, brems_tau, rng_uniform(rng, 0, S::kDepth0), rng_uniform(rng, 0, S::kPick0));
                lane_col[Col::kColHmx * kB] = hard_tau;
                lane_col[Col::kColHazardTotal * kB] = hard_tau + brems_tau;
            }
            if (!h.valid || h.event || err) break;
            // continuous drag over [0, se] at the frozen energy k1, not reaching the cutoff
            double const col = charged_value(T, S::kCol, b1, err) * 1.0;
            double const soft = charged_value(T, S::kSoft, b1, err) * 1.0;
            // the plan's mean hard-Moller transfer rate at the frozen energy (its
            // channel gate is the rate's own, so the inverse length is at hand)
            double const hstp = kPos ? 0.0 : hard_moller_loss_at(T, b1, hm_inv1);
            double const total = col + soft;
            double const to_cut = k1 - cutoff;
            double const req = total * L;
            if (!finite(req) || req < 0.0 || !(req < to_cut) || !(k1 > 0.0) || err) break;
            double const wq = unstash(2);
            double const raw_col = col * L;
            double const raw_soft = soft * L;
            double const avail = dmax(0.0, k1 - cutoff);
            double const drag = dmin(raw_col + raw_soft, avail);
            double const col_loss = dmin(raw_col, drag);
            lane_col[Col::kColSoft * kB] = wq * dmax(0.0, drag - col_loss);
            if constexpr (!kPos) {
                lane_col[Col::kColCol * kB] = wq * col_loss;
                lane_col[Col::kColHard * kB] = wq * dmin(hstp * L, avail);
            }
            double ion = 0.0;
            if (col_loss > 0.0 && raw_col > 0.0) {
                double const le = dmin(se, iv.end);
                if (le > 0.0) {
                    double const raw_i = path.full * (le - 0.0) * iv.ratio * col;
                    double const loss_i = col_loss * raw_i / raw_col;
                    double const pw = wq * loss_i / dmax(kWAir, 1.0);
                    if (pw > 0.0) ion += pw;
                }
            }
            double k2 = dmax(0.0, k1 - drag);
            if (!(k2 >= cutoff)) break;
            // elastic scattering: the few-collision branch with at most two scatters
            EBracket eb2;
            if (!elastic_bracket_near(EL, k2, eb0, eb2)) break;
            double const mfp2 = elastic_value(EL, EL.mfp, eb2) / 1.0;
            double const gs_n = (mfp2 > 0.0 ? 1.0 / mfp2 : 0.0) * L;
            if (gs_n > 0.0) {
                if (!(gs_n < T.gs_few) || ceil(3.0 * T.gs_few) >= 255.0) break;
                int const cnt = poisson_capped(gs_n, rng_uniform(rng, 0U, S::kFewCount), static_cast<int>(ceil(3.0 * T.gs_few)));
                if (cnt > 2) break;
                if (cnt > 0 && !(k2 >= T.charged_min && k2 <= T.charged_max)) break;
                for (int q = 0; q < cnt; ++q) {
                    uint64_t const inner = static_cast<uint64_t>(q + 1);
                    double const eta = elastic_value(EL, EL.eta, eb2);
                    double const u = dmin(dmax(rng_uniform(rng, inner, S::kFewScatter), 0.0), 1.0 - 1.0e-15);
                    double const ct = 1.0 - 2.0 * eta * u / (1.0 + eta - u);
                    double nx, ny, nz;
                    cone_direction(d1x, d1y, d1z, ct, rng_uniform(rng, inner, S::kFewPhi), nx, ny, nz);
                    d1x = nx; d1y = ny; d1z = nz;
                }
            }
            double const hw2 = half_kick(fr, 0.5 * substep_dt, S::kCharge, k2, d1x, d1y, d1z);
            if (!(k2 >= cutoff)) break;
            double x, y, z;
            point_at_full(path, se, x, y, z);
            if (!in_half_open(x, y, z, D) || path.exits) break;
            // terminal: a survivor (for e-, above the demotion gate's energy)
            if (!kPos && thr > 0.0 && k2 <= thr) { light = true; break; }
            double const end_r = hypot(x, y);
            double const pr0 = unstash(0), pz0 = unstash(1);
            if (!deposit_segment_check(E.G, pr0, pz0, end_r, z)) break;
            // ---- commit: nothing above has written anything ----
            slow = false;
            dep = true;
            src.w = 0.0;  // no record unless one is formed below
            if (ion > 0.0 && E.src) {
                // the ion pairs' record at the interval midpoint (the one
                // interval ends where the substep does: iv.end == se here),
                // deposited with the workgroup's other records at the end
                double mx, my, mz;
                point_at_full(path, 0.5 * (0.0 + se), mx, my, mz);
                double const mr = hypot(mx, my);
                if (source_in_domain(E, mr, mz)) src = source_cell(E, mr, mz, ion);
                else report(A.error, kErrDomain, ip);
            }
            dep_inst = E.instance;
            dep_q = S::kCharge * wq;
            dep_r0 = pr0; dep_z0 = pz0; dep_r1 = end_r; dep_z1 = z;
            double const um = proper_velocity(k2);
            A.ux[ip] = um * d1x; A.uy[ip] = um * d1y; A.uz[ip] = um * d1z;
            A.r[ip] = end_r;
            A.th[ip] = atan2(y, x);
            A.z[ip] = z;
            if (A.status) A.status[ip] = static_cast<uint8_t>(RREA_SURVIVE);
            if (A.ion_pairs) A.ion_pairs[ip] = ion;
            if (A.substeps) A.substeps[ip] = 1;
            lane_col[Col::kColSurvivor * kB] = 1.0;
            lane_col[Col::kColSubstep * kB] = 1.0;
            lane_col[Col::kColDrag * kB] = wq * drag;
            lane_col[Col::kColIon * kB] = ion;
            lane_col[Col::kColFieldWork * kB] = wq * hw1 + wq * hw2;
        } while (false);
    }
    append_slow_split(slow, light, ip, slow_list, slow_count, A.n);
    if (!(lane_col[Col::kColEnergy * kB] > 0.0)) {
        // a lane stopped by a substep driver records no guard taus
        for (int q = Col::kColSums; q < Col::kColEnergy; ++q) lane_col[q * kB] = 0.0;
    }
    if (slow) {
        // a lane handed over counts none of the sums it formed
        for (int q = Col::kColSoft; q < Col::kColSums; ++q) lane_col[q * kB] = 0.0;
    } else if constexpr (!kPos) {
        if (ip < A.n) {
            // the expected hard-Moller weight of the committed step's one hazard
            // draw, w P(total) hard / total (OV/RreaElectronTransport.cpp:630-646)
            double const hard_tau = lane_col[Col::kColHmx * kB];
            double const total_tau = lane_col[Col::kColHazardTotal * kB];
            lane_col[Col::kColHmx * kB] = A.w[ip] * (total_tau > 0.0
                ? probability_from_optical_depth(total_tau) * hard_tau / total_tau : 0.0);
        }
    }
    // the workgroup's tallies and ledger records, booked by the fold kernel
    fold_columns<Col::kCols, Col::kColSums, kB>(window, [&](int q, double v) {
        partials[static_cast<int64_t>(blockIdx.x) * kPartialStride + q] = v;
    });
    deposit_block<kB>(A, window, window + window_faces<kB>(), dep, dep_inst, dep_q, A.cfg.dt_s, dep_r0, dep_z0, dep_r1, dep_z1,
                  src);
}

// Every particle through the complete path (used when the fast path is off).
__global__ void __launch_bounds__(kBlock) electron_full_kernel(const StepArgs* __restrict__ args)
{
    StepArgs const& A = *args;
    __shared__ double ledger[RREA_LEDGER_COUNT];
    ledger_begin(ledger);
    double tally[kTallyUsed];
    for (int q = 0; q < kTallyUsed; ++q) tally[q] = 0.0;
    int64_t const ip = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (ip < A.n) charged_full_particle<false>(A, ip, tally, ledger);
    reduce_tallies(tally, A.tallies);
    ledger_flush(ledger, A.tallies);
}

// The straight-line photon push at c along u / |u| that WarpX applies before
// the hook: the pushed point becomes the chord end, the old point the
// captured pre-push snapshot.
// n_dev (device, may be NULL): the live length, read instead of n (a tile
// growing inside a captured graph).
__global__ void photon_push_kernel(int64_t n, const int64_t* __restrict__ id, double* __restrict__ prev_r,
                                   double* __restrict__ prev_z, double* __restrict__ prev_th, double* __restrict__ r,
                                   double* __restrict__ z, double* __restrict__ th, const double* __restrict__ ux,
                                   const double* __restrict__ uy, const double* __restrict__ uz, double c_dt,
                                   const int64_t* __restrict__ n_dev = nullptr)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const len = n_dev ? (*n_dev < n ? *n_dev : n) : n;
    if (i >= len || id[i] == RREA_INVALID_ID) return;
    double const r0 = r[i], t0 = th[i], z0 = z[i];
    prev_r[i] = r0; prev_th[i] = t0; prev_z[i] = z0;
    double const un = sqrt(ux[i] * ux[i] + uy[i] * uy[i] + uz[i] * uz[i]);
    double const s = c_dt / (un > 0.0 ? un : 1.0);
    double s_t0, c_t0;
    sin_cos(t0, s_t0, c_t0);
    double const x = r0 * c_t0 + ux[i] * s;
    double const y = r0 * s_t0 + uy[i] * s;
    r[i] = hypot(x, y);
    th[i] = atan2(y, x);
    z[i] = z0 + uz[i] * s;
}

// One lane per photon.
__global__ void __launch_bounds__(kBlock) photon_kernel(const StepArgs* __restrict__ args)
{
    StepArgs const& A = *args;
    __shared__ double ledger[RREA_LEDGER_COUNT];
    ledger_begin(ledger);
    double tally[kTallyUsed];
    for (int q = 0; q < kTallyUsed; ++q) tally[q] = 0.0;
    int64_t const ip = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (ip < A.n) photon_particle(A, ip, tally, ledger);
    reduce_tallies(tally, A.tallies);
    ledger_flush(ledger, A.tallies);
}

// The photon fast path.  At small dt nearly every photon is valid, weighted,
// above the photon cutoff and draws a depth beyond its chord: it survives at
// (or escapes through) the end of the clipped chord from the captured point to
// the pushed point.  That branch of the complete step is taken here with its
// own arithmetic, in the same order; any other outcome -- an interaction, a
// sub-cutoff or weightless photon, an error -- is detected before the first
// write and the photon is handed over to the complete step.
template <bool kBatched, int kB = 256>
__global__ void __launch_bounds__(kB) photon_fast_kernel(const StepArgs* __restrict__ args,
                                                           int32_t* __restrict__ slow_list,
                                                           int32_t* __restrict__ slow_count,
                                                           double* __restrict__ partials)
{
    StepArgs const& A = *args;
    // each lane's tallies and ledger records in LDS columns of its own, folded
    // per workgroup at the end: survivors, escapes, escaped energy (sums); the
    // pass's tau and energy (maxima) -- OV/RreaPhotonTransport.cpp:192, 222-233
    using Col = PhotonColumns;
    __shared__ double cols[Col::kCols * kB];
    double* const lane_col = cols + threadIdx.x;
    for (int q = 0; q < Col::kCols; ++q) lane_col[q * kB] = 0.0;
    int64_t const ip = static_cast<int64_t>(blockIdx.x) * kB + threadIdx.x;
    bool slow = false;
    // a batched wave takes the fast path only when all its particles share one
    // instance, whose environment is then wave-uniform; a wave that straddles
    // an instance boundary goes to the complete path
    int32_t wave_inst = 0;
    bool mixed = false;
    if constexpr (kBatched) {
        int32_t const mine = ip < A.n ? A.inst[ip] : -1;
        int32_t const first = __builtin_amdgcn_readfirstlane(mine);
        mixed = __ballot(ip < A.n && mine != first) != 0;
        wave_inst = first < 0 ? 0 : first;
    }
    if (ip < A.n) {
        slow = true;
        DevTables const& T = A.T;
        Env const* __restrict__ const envs = A.batch;
        Env const& E = kBatched ? envs[wave_inst] : A.env;
        Domain const& D = E.D;
        rrea_step_config const& cfg = A.cfg;
        int64_t const id = A.id[ip];
        if (kBatched && mixed) {
            // handed over as a whole wave
        } else if (id == RREA_INVALID_ID) {
            if (A.status) A.status[ip] = static_cast<uint8_t>(RREA_SKIPPED_INVALID);
            if (A.ion_pairs) A.ion_pairs[ip] = 0.0;
            if (A.substeps) A.substeps[ip] = 0;
            slow = false;
        } else do {
            double const w = A.w[ip];
            if (!(w > 0.0)) break;
            double const ux = A.ux[ip], uy = A.uy[ip], uz = A.uz[ip];
            double const energy = sqrt(ux * ux + uy * uy + uz * uz);
            if (!(energy >= cfg.photon_cutoff_eV)) break;
            double const pr = A.prev_r[ip], pth = A.prev_th[ip], pz = A.prev_z[ip];
            if (pr > D.r_hi + 1.0e-9 * D.r_hi || pz < D.z_lo - 1.0e-9 || pz > D.z_hi + 1.0e-9 * D.z_hi) break;
            int32_t err = kOk;
            Path path;
            if (!clip_rz_path(pr, pth, pz, A.r[ip], A.th[ip], A.z[ip], D, path, err) || err) break;
            bool const escapes = clipped_escapes(path, D);
            Interval const iv = build_interval(E.F, path, D, E.density_ratio, err);
            double const chord_dp = density_path(iv, path, 0.0, path.end_f);
            if (err) break;
            double pass_tau = 0.0;
            if (path.in_len > 0.0) {
                PhotonRates rates;
                if (!photon_rates(T, energy, rates, err) || err) break;
                double const total = rates.compton + rates.pe + rates.pair_n + rates.pair_t;
                if (!finite(total) || total < 0.0) break;
                double const mfp = total > 0.0 ? 1.0 / total : 1.0e99;
                int32_t const cpu = A.cpu ? A.cpu[ip] : 0;
                uint64_t const pkey = static_cast<uint64_t>(id < 0 ? -id : id)
                    ^ (static_cast<uint64_t>(static_cast<uint32_t>(cpu)) << 40);
                uint64_t const step = static_cast<uint64_t>(cfg.step > 0 ? cfg.step : 0);
                // the first pass's live index is segment 0, event 0: index 0
                double const draw = rng_uniform(rng_prefix(E.rng_seed, pkey, step), 0U, kChPhotonInteraction);
                pass_tau = chord_dp / mfp;
                if (select_hazard(pass_tau, 0.0, draw, 0.0).event) break;
            }
            // ---- commit: nothing above has written anything ----
            slow = false;
            double x, y, z;
            point_at_full(path, path.end_f, x, y, z);
            A.r[ip] = hypot(x, y);
            A.th[ip] = atan2(y, x);
            A.z[ip] = z;
            if (escapes) {
                A.id[ip] = RREA_INVALID_ID;
                lane_col[Col::kColEscape * kB] = 1.0;
                lane_col[Col::kColEscapedEnergy * kB] = w * energy;
            } else {
                lane_col[Col::kColSurvivor * kB] = 1.0;
            }
            if (path.in_len > 0.0) {
                lane_col[Col::kColTau * kB] = pass_tau;
                lane_col[Col::kColEnergy * kB] = energy;
            }
            if (A.status) A.status[ip] = static_cast<uint8_t>(escapes ? RREA_ESCAPE : RREA_SURVIVE);
            if (A.ion_pairs) A.ion_pairs[ip] = 0.0;
            if (A.substeps) A.substeps[ip] = 0;
        } while (false);
    }
    append_slow(slow, ip, slow_list, slow_count);
    fold_columns<Col::kCols, Col::kColSums, kB>(cols, [&](int q, double v) {
        partials[static_cast<int64_t>(blockIdx.x) * kPartialStride + q] = v;
    });
}

// The photons the fast path handed over, strided over a fixed grid.
__global__ void __launch_bounds__(kSlowThreads) photon_slow_kernel(const StepArgs* __restrict__ args,
                                                           const int32_t* __restrict__ list,
                                                           const int32_t* __restrict__ count)
{
    StepArgs const& A = *args;
    __shared__ double ledger[RREA_LEDGER_COUNT];
    ledger_begin(ledger);
    double tally[kTallyUsed];
    for (int q = 0; q < kTallyUsed; ++q) tally[q] = 0.0;
    int64_t const n = *count;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * kSlowThreads + threadIdx.x; i < n;
         i += static_cast<int64_t>(gridDim.x) * kSlowThreads) {
        photon_particle(A, list[i], tally, ledger);
    }
    reduce_tallies<kTallyUsed, kSlowThreads>(tally, A.tallies);
    ledger_flush(ledger, A.tallies);
}

// Every positron through the complete path (used when the fast path is off).
__global__ void __launch_bounds__(kBlock) positron_full_kernel(const StepArgs* __restrict__ args)
{
    StepArgs const& A = *args;
    __shared__ double ledger[RREA_LEDGER_COUNT];
    ledger_begin(ledger);
    double tally[kTallyUsed];
    for (int q = 0; q < kTallyUsed; ++q) tally[q] = 0.0;
    int64_t const ip = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x;
    if (ip < A.n) charged_full_particle<true>(A, ip, tally, ledger);
    reduce_tallies(tally, A.tallies);
    ledger_flush(ledger, A.tallies);
}

// ---------------------------------------------------------------------------
// Deterministic secondary order: key = (parent index, lineage), the order of
// the reference's chunk-ordered merge (RreaElectronTransport.cpp:892-897).
// ---------------------------------------------------------------------------
// ===========================================================================
// Newborn residual advance (OV/RreaNewbornAdvance.cpp) and the transient FIFO
// of DrainAndCommitSecondaries (OV/RreaSecondaryCommit.cpp:338-468).  A
// secondary born at elapsed fraction f < 1 is flown through the remaining
// (1 - f) dt with the live loops' physics; its draws come from the ROOT
// particle's stream at the packed transient lineage.  The FIFO runs one
// generation per launch: within a generation every root's entries are
// contiguous (the step's output is ordered by parent, and children are
// ordered by parent), which is what lets lineages be allocated per run.
// ===========================================================================
constexpr int kTallyNewborn = 21;
constexpr uint64_t kLineageCap = uint64_t{1} << 14;
constexpr uint64_t kInnerCap = uint64_t{1} << 8;

// RNG channels of the residual advance, OV/RreaRng.H:65-81
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

// One generation of the FIFO: entries [gb, ge) of `work`.  The children the
// generation makes are staged through S.sec / S.sec_count.
struct NewbornArgs {
    StepArgs S;
    const rrea_secondary* work;
    int64_t gb, ge;
    rrea_secondary* fin;   // per entry: the state at the end of the step
    uint8_t* status;       // per entry
    double* ion;           // per entry: ion-pair weight
    int32_t* substeps;     // per entry: residual substeps (charged) or interactions (photon)
};

// RreaTryPackTransientRngIndex OV/RreaParticleInteraction.H:660-686 with the
// newborn marker (bit 51)
__device__ __forceinline__ uint64_t newborn_index(uint64_t rsp, uint64_t lineage, uint64_t segment, uint64_t event,
                                                  uint64_t inner, int32_t& err)
{
    if (rsp >= 3U || lineage >= kLineageCap || segment >= kSegmentCap || event >= kEventCap || inner >= kInnerCap) {
        if (!err) err = kErrInvalid;
        return 0;
    }
    return (uint64_t{1} << 51) | (rsp << 49) | (lineage << 35) | (segment << 21) | (event << 8) | inner;
}

// RreaNewbornResidualFraction / RreaNewbornElapsedStepFraction
// OV/RreaParticleInteraction.H:171-197
__device__ __forceinline__ double newborn_residual(double birth) { return dmin(1.0, dmax(0.0, 1.0 - birth)); }

__device__ __forceinline__ double newborn_elapsed(double birth, double residual, uint64_t completed,
                                                  uint64_t substeps, double f)
{
    double const within = dmin(1.0, dmax(0.0, f));
    double const consumed = (static_cast<double>(completed) + within) / static_cast<double>(substeps);
    return dmin(1.0, birth + residual * consumed);
}

// TransportStep::NewbornFinalPoint OV/RreaNewbornAdvance.cpp:20-48: the point
// at `fraction`, pulled back along the chord until it is in the half-open domain
__device__ bool newborn_final_point(Path const& p, Domain const& D, double fraction, double& x, double& y, double& z)
{
    double const span = fraction - 0.0;
    double pullback = 1.0e-9;
    for (int attempt = 0; attempt < 5; ++attempt) {
        double const trial = attempt == 0 ? fraction : dmax(0.0, fraction - pullback * dmax(span, 1.0e-12));
        point_at_full(p, trial, x, y, z);
        if (in_half_open(x, y, z, D)) return true;
        if (attempt > 0) pullback *= 100.0;
    }
    return false;
}

// The newborn's electron_state_policy_at (OV/RreaNewbornAdvance.cpp:79-96):
// the demotion gate with the field gathered at the point and the density
// ratio at its z
__device__ bool newborn_electron_demotes(StepArgs const& A, Env const& E, double x, double y, double z, double k,
                                         int32_t& err)
{
    double const cutoff = A.cfg.low_energy_cutoff_eV;
    double const thr = A.T.hard_moller_threshold;
    double er, ez, bt;
    gather(E.F, hypot(x, y), z, er, ez, bt);
    double const ratio = positive_or(density_ratio_at_z(E.F, z, err), 1.0);
    double const drag = ensemble_drag(A.T, dmax(k, cutoff), ratio, A.cfg.disable_hard_moller == 0, err);
    return k < cutoff || (thr > 0.0 && k <= thr && hypot(er, ez) <= drag);
}

// TransportStep::AdvanceChargedNewborn OV/RreaNewbornAdvance.cpp:50-777
template <bool kPos>
__device__ __noinline__ void charged_newborn(NewbornArgs const& N, int64_t i, double* tally, double* ledger)
{
    using S = Species<kPos>;
    StepArgs const& A = N.S;
    DevTables const& T = A.T;
    Env const& E = env_of_record(A, N.work[i].instance);
    DevFields const& F = E.F;
    Domain const& D = E.D;
    rrea_step_config const& cfg = A.cfg;
    double const cutoff = cfg.low_energy_cutoff_eV;
    double const gcut = cfg.photon_cutoff_eV;
    double const thr = T.hard_moller_threshold;
    bool const hm_on = cfg.disable_hard_moller == 0;
    double const dt = cfg.dt_s;
    uint64_t const step = static_cast<uint64_t>(cfg.step > 0 ? cfg.step : 0);
    DevElastic const& EL = kPos ? T.el_p : T.el_e;
    constexpr uint64_t kFewCount = kPos ? kChNewbornPositronGsFewCount : kChNewbornElectronGsFewCount;
    constexpr uint64_t kFewScatter = kPos ? kChNewbornPositronGsFewScatter : kChNewbornElectronGsFewScatter;
    constexpr uint64_t kFewPhi = kPos ? kChNewbornPositronGsFewAzimuth : kChNewbornElectronGsFewAzimuth;
    constexpr uint64_t kGsAngle = kPos ? kChNewbornPositronGsAngle : kChNewbornElectronGsAngle;
    constexpr uint64_t kGsPhi = kPos ? kChNewbornPositronGsAzimuth : kChNewbornElectronGsAzimuth;

    rrea_secondary p = N.work[i];
    int32_t err = kOk;
    double ion = 0.0;
    uint64_t lineage = 0;
    uint64_t nsub = 0;
    int status = RREA_SURVIVE;
    uint64_t const rsp = static_cast<uint64_t>(p.root_species);
    uint64_t const lin = p.lineage;
    double const w = p.weight;
    uint64_t const rng = rng_prefix(E.rng_seed, p.root_key, step);
    Emitter em{A, i, p.root_key, w, ion, lineage, err, p.root_species, p.generation + 1, E.instance, &E, ledger};
    double const birth_r = hypot(p.x, p.y), birth_z = p.z;
    double const residual = newborn_residual(p.birth_fraction);
    double k = p.energy_eV;
    double dx = p.dir_x, dy = p.dir_y, dz = p.dir_z;
    normalize(dx, dy, dz);
    auto segment_to = [&](double r1, double z1) {
        if (w != 0.0 && !deposit_segment(E.G, S::kCharge * w, dt, birth_r, birth_z, r1, z1) && !err) err = kErrInvalid;
    };
    auto escape_along = [&](Path const& path) {
        Crossing const bc = boundary_crossing(path, D);
        double const q = S::kCharge * w;
        if (!bc.valid) {
            if (!err) err = kErrInvalid;
        } else if (q != 0.0 && !deposit_escaping_segment(E.G, q, dt, birth_r, birth_z, bc.start_r, bc.start_z,
                                                         bc.cross_r, bc.cross_z, E.pnr, E.pzb, E.pnz, bc.faces)) {
            if (!err) err = kErrInvalid;
        }
        status = RREA_ESCAPE;
        tally[19] += 1.0;
    };
    auto terminate_at = [&](double x, double y, double z, uint64_t index, double elapsed) {
        if constexpr (kPos) {
            spawn_annihilation(T, gcut, rng, em, kProcAnnihilation, x, y, z, dx, dy, dz, k, true, index, elapsed,
                               tally);
            // OV/RreaNewbornAdvance.cpp:123, 729
            Ledger{ledger, cutoff}.positron_transport(w, k, true);
            status = RREA_ANNIHILATED;
        } else {
            tally[16] += w;
            em.low_electron(hypot(x, y), z, w, 0.0);
            em.deposit_local(k, hypot(x, y), z);
            status = RREA_DEMOTED;
        }
        tally[20] += 1.0;
    };

    if (!in_half_open(p.x, p.y, p.z, D)) {
        err = kErrDomain;
    } else {
        bool const birth_demote = kPos ? k < cutoff : newborn_electron_demotes(A, E, p.x, p.y, p.z, k, err);
        if (err) {
        } else if (birth_demote) {
            terminate_at(p.x, p.y, p.z, newborn_index(rsp, lin, 0U, 0U, 0U, err), p.birth_fraction);
        } else if (residual > 0.0) {
            double const birth_u = proper_velocity(k);
            double const birth_gamma = sqrt(1.0 + birth_u * birth_u / (kC * kC));
            double const residual_dt = dt * residual;
            double const chord = (birth_u / birth_gamma) * residual_dt;
            Path driver;
            bool driver_ok = false;
            if (!(chord > 0.0)) {
                if (k < cutoff) status = RREA_DEMOTED;
            } else {
                double const end_x = p.x + chord * dx, end_y = p.y + chord * dy, end_z = p.z + chord * dz;
                if (clip_rz_path(hypot(p.x, p.y), atan2(p.y, p.x), p.z, hypot(end_x, end_y), atan2(end_y, end_x), end_z,
                                 D, driver, err)) {
                    if (!(driver.full > 0.0) || !(driver.in_len > 0.0)) {
                        if (driver.exits) escape_along(driver);
                    } else {
                        driver_ok = true;
                    }
                }
            }
            if (driver_ok && !err) {
                Interval iv = build_interval(F, driver, D, E.density_ratio, err);
                if (!iv.valid && !err) err = kErrInvalid;
                Path path = driver;
                uint64_t substeps = 1;
                if (!err) {
                    double const driver_dp = density_path(iv, driver, 0.0, driver.end_f);
                    double const gs_n_total = elastic_inverse_length(EL, k, err) * driver_dp;
                    Bracket bk;
                    if (!charged_bracket(T, k, bk)) err = kErrRange;
                    double drag_est;
                    if constexpr (kPos) {
                        drag_est = charged_value(T, kPCol, bk, err) * 1.0 * driver_dp;
                    } else {
                        drag_est = (k > 0.0 ? charged_value(T, kECol, bk, err) * 1.0 + charged_value(T, kESoft, bk, err) * 1.0
                                            : 0.0) * driver_dp;
                    }
                    double const whole_work = path_field_work(F, driver, D, S::kCharge, 0.0, 1.0, err).work;
                    double const impulse = dmax(cfg.max_field_impulse_fraction * k, dmax(cfg.field_impulse_floor_eV, 1.0));
                    double const span = dmax(driver.end_f, 1.0e-6);
                    uint64_t const gsd = static_cast<uint64_t>(dmax(1.0, ceil(dmin(gs_n_total / span, cfg.gs_total_collision_cap)
                                                                            / dmax(cfg.max_gs_collisions_per_substep, 1.0))));
                    uint64_t const imd = static_cast<uint64_t>(dmax(1.0, ceil(((fabs(whole_work) + drag_est) / span) / impulse)));
                    substeps = gsd > imd ? gsd : imd;
                    if (substeps > cfg.max_subcycles) substeps = cfg.max_subcycles;
                    if (substeps >= kSegmentCap && !err) err = kErrInvalid;
                    nsub = substeps;
                }
                double run_x = p.x, run_y = p.y, run_z = p.z;
                double const sub_dt = residual_dt / static_cast<double>(substeps);
                double final_f = 0.0;
                bool terminal_interaction = false, terminal_at_point = false, exited = false;
                uint64_t completed = 0;
                auto elapsed_at = [&](double f) {
                    return newborn_elapsed(p.birth_fraction, residual, completed, substeps, f);
                };
                auto mark_stopped = [&](double f) { final_f = f; terminal_at_point = true; };
                auto in_range = [&]() {
                    if (!(finite(k) && k >= T.charged_min && k <= T.charged_max) && !err) err = kErrRange;
                };
                double positron_total_loss = 0.0;
                // RreaPlanContinuousDragSegment OV/RreaChargedInterleave.H:239-292
                // the last event-interleave bracket: the next one (the energy
                // after a kick, drag or event) is usually in its interval; the drag segment brackets the same energy
                Bracket hint{0, 0, 1, 0.0, 0.0};
                auto continuous = [&](double b, double e, double stop_e, double& eff) -> bool {
                    eff = b;
                    if (!(e > b) || !(k > 0.0) || stop_e < cutoff) return false;
                    Bracket bs;
                    if (!charged_bracket_near(T, stop_e, hint, bs)) { if (!err) err = kErrRange; return false; }
                    double const col = charged_value(T, S::kCol, bs, err) * 1.0;
                    double const soft = charged_value(T, S::kSoft, bs, err) * 1.0;
                    double const total = col + soft;
                    bool reaches = false;
                    double const to_cut = k - cutoff;
                    double const req = total * density_path(iv, path, b, e);
                    if (!finite(req) || req < 0.0) { if (!err) err = kErrInvalid; return false; }
                    if (req < to_cut) {
                        eff = e;
                    } else {
                        reaches = true;
                        if (to_cut != 0.0) {
                            double lo = b, hi = e;
                            for (int it = 0; it < 56; ++it) {
                                double const mid = 0.5 * (lo + hi);
                                if (total * density_path(iv, path, b, mid) < to_cut) lo = mid; else hi = mid;
                            }
                            eff = 0.5 * (lo + hi);
                        }
                    }
                    double const pm = density_path(iv, path, b, eff);
                    double const raw_col = col * pm;
                    double const raw_soft = soft * pm;
                    double const avail = dmax(0.0, k - cutoff);
                    double const drag = reaches ? avail : dmin(raw_col + raw_soft, avail);
                    double const col_loss = dmin(raw_col, drag);
                    // RreaPartitionContinuousIonPairs OV/RreaChargedInterleave.H:157-196
                    if (col_loss > 0.0 && raw_col > 0.0 && iv.valid) {
                        double const lb = dmax(b, 0.0), le = dmin(eff, iv.end);
                        if (le > lb) {
                            double const raw_i = path.full * (le - lb) * iv.ratio * col;
                            double const loss_i = col_loss * raw_i / raw_col;
                            double const pw = w * loss_i / dmax(kWAir, 1.0);
                            if (pw > 0.0) {
                                ion += pw;
                                double mx, my, mz;
                                point_at_full(path, 0.5 * (lb + le), mx, my, mz);
                                em.ionization(hypot(mx, my), mz, pw, 0.5 * cutoff, kWAir);
                                Ledger{ledger, cutoff}.transport_ionization(pw, kWAir, true, false);
                            }
                        }
                    }
                    tally[7] += w * drag;
                    {
                        // the plan's partition: soft radiative share, and the mean hard-Moller
                        // transfer over the same density path (electrons; OV/RreaElectronTransport.cpp:306-338)
                        Ledger led{ledger, cutoff};
                        led.soft_brems(w, dmax(0.0, drag - col_loss));
                        if constexpr (kPos) {
                            positron_total_loss += drag;
                        } else {
                            double const hstp = (hm_on && thr > 0.0 && stop_e > 2.0 * thr)
                                ? hard_moller_energy_loss(T, bs, stop_e, err) : 0.0;
                            led.loss_partition(w, drag, col_loss, dmin(hstp * pm, avail));
                        }
                    }
                    k = dmax(0.0, k - drag);
                    if (reaches) k = cutoff;
                    return reaches;
                };
                FieldRange sfield{0, 0, 0, 0, 0, 0, 0, 0};
                double substep_dt = 0.0;
                // RreaRunChargedSubstepInterleave OV/RreaChargedInterleave.H:693-732
                for (uint64_t sc = 0; sc < substeps && !err; ++sc) {
                    if (k < cutoff) break;
                    double sb = 0.0, se = 0.0;
                    {
                        Path next;
                        if (!build_chord(run_x, run_y, run_z, dx, dy, dz, k, sub_dt, residual_dt, D, next, err)) break;
                        path = next;
                        if (!(path.full > 0.0) || !(path.in_len > 0.0)) {
                            exited = path.exits;
                            final_f = 0.0;
                            break;
                        }
                        iv = build_interval(F, path, D, E.density_ratio, err);
                        if (!iv.valid) { if (!err) err = kErrInvalid; break; }
                        sb = 0.0;
                        se = path.end_f;
                        final_f = se;
                        sfield = path_field_work(F, path, D, S::kCharge, sb, se, err);
                        substep_dt = sub_dt * (se - sb);
                        (void)half_kick(sfield, 0.5 * substep_dt, S::kCharge, k, dx, dy, dz);
                        if (k < cutoff) {
                            final_f = sb;
                            terminal_at_point = true;
                            break;
                        }
                        in_range();
                    }
                    {
                        // RreaRunChargedEventInterleave OV/RreaChargedInterleave.H:546-664
                        double rem = sb;
                        uint64_t ord = 0;
                        double ex_ = 0.0, ey_ = 0.0, ez_ = 0.0;
                        while (rem < se && k >= cutoff && !err) {
                            if (ord >= kEventCap) { err = kErrInvalid; break; }
                            double const frozen_e = k;
                            Bracket bf;
                            if (!charged_bracket_near(T, frozen_e, hint, bf)) { err = kErrRange; break; }
                            hint = bf;
                            double s0, s1, s2 = 0.0;
                            if constexpr (kPos) positron_stp_rates(T, bf, thr, gcut, frozen_e, s0, s1, s2, err);
                            else electron_stp_rates(T, bf, thr, gcut, frozen_e, hm_on, s0, s1, err);
                            double const total_stp = s0 + s1 + s2;
                            double const L = density_path(iv, path, rem, se);
                            if constexpr (!kPos) {
                                // OV/RreaNewbornAdvance.cpp:536-548
                                double const hard_tau = s0 * L, brems_tau = s1 * L;
                                double const total_tau = hard_tau + brems_tau;
                                Ledger{ledger, cutoff}.hm_expectation(
                                    w * (total_tau > 0.0 ? probability_from_optical_depth(total_tau) * hard_tau / total_tau
                                                         : 0.0));
                            }
                            uint64_t const idx = newborn_index(rsp, lin, sc, ord, 0U, err);
                            if (err) break;
                            double const u1 = rng_uniform(rng, idx, kChNewbornCompetingOpticalDepth);
                            double const u2 = rng_uniform(rng, idx, kChNewbornCompetingChannel);
                            Hazard h;
                            if constexpr (kPos) h = select_hazard3(s0 * L, s1 * L, s2 * L, u1, u2);
                            else h = select_hazard(s0 * L, s1 * L, u1, u2);
                            if (!h.valid) { err = kErrInvalid; break; }
                            if (!h.event) {
                                double eff;
                                if (continuous(rem, se, frozen_e, eff)) mark_stopped(eff);
                                break;
                            }
                            double const ef = fraction_at_density_path(iv, path, rem, se, h.sampled / total_stp, err);
                            if (!(finite(rem) && finite(ef) && finite(se) && se > rem && ef >= rem && ef <= se)) {
                                if (!err) err = kErrInvalid;
                                break;
                            }
                            point_at_full(path, ef, ex_, ey_, ez_);
                            double const efrac = elapsed_at(ef);
                            {
                                double eff;
                                if (continuous(rem, ef, frozen_e, eff)) { mark_stopped(eff); break; }
                            }
                            bool open_ev, open_fr = false;
                            if constexpr (kPos) {
                                open_ev = k >= cutoff && positron_channel_open(T, thr, gcut, h.channel, k, err);
                                if (!open_ev) {
                                    open_fr = frozen_e >= cutoff
                                        && positron_channel_open(T, thr, gcut, h.channel, frozen_e, err);
                                }
                            } else {
                                open_ev = k >= cutoff && electron_channel_open(T, thr, gcut, h.channel, k, hm_on, err);
                                if (!open_ev) {
                                    open_fr = frozen_e >= cutoff
                                        && electron_channel_open(T, thr, gcut, h.channel, frozen_e, hm_on, err);
                                }
                            }
                            if (!open_ev) {
                                if (!open_fr) { if (!err) err = kErrInvalid; break; }
                                rem = ef;
                                ++ord;
                                continue;
                            }
                            if (h.channel == 0) {
                                double const before = k;
                                if (!(finite(before) && before >= T.charged_min && before <= T.charged_max)) { err = kErrRange; break; }
                                if constexpr (kPos) {
                                    // RreaApplyBhabhaFinalState OV/RreaChargedFinalStates.H:238-297
                                    double const T2 = dmin(cdf_sample(T.bhabha, before, rng_uniform(rng, idx, kChPositronBhabhaEnergy),
                                                                      false, err),
                                                           dmax(0.0, before));
                                    if (T2 > 0.0 && !(thr > 0.0 && T2 < thr)) {
                                        double sdx, sdy, sdz, ndx, ndy, ndz;
                                        moller_lab_directions(dx, dy, dz, before, T2,
                                                              rng_uniform(rng, idx, kChPositronBhabhaAzimuth),
                                                              sdx, sdy, sdz, ndx, ndy, ndz);
                                        k = dmax(0.0, before - T2);
                                        dx = ndx; dy = ndy; dz = ndz;
                                        em.secondary(0, kProcHardBhabha, ex_, ey_, ez_, sdx, sdy, sdz, T2, efrac);
                                        em.positive_ion(hypot(ex_, ey_), ez_, w, T2);
                                        tally[12] += w;
                                        tally[11] += 1.0;
                                        {
                                            Ledger led{ledger, cutoff};
                                            led.transport_ionization(w, T2, false, true);
                                            led.bhabha(w, T2, true);
                                        }
                                        GroupClosure c;
                                        c.parent(1, w, before, k, true);
                                        c.particle(0, w, T2);
                                        c.positive_ion(w);
                                        check_group(c, cutoff, err);
                                    }
                                } else {
                                    // RreaApplyHardMollerFinalState OV/RreaChargedFinalStates.H:164-225
                                    double T2 = cdf_sample(T.moller, before, rng_uniform(rng, idx, kChHardMollerEnergy), true, err);
                                    T2 = dmax(thr, dmin(T2, 0.5 * before));
                                    double sdx, sdy, sdz, ndx, ndy, ndz;
                                    moller_lab_directions(dx, dy, dz, before, T2, rng_uniform(rng, idx, kChHardMollerAzimuth),
                                                          sdx, sdy, sdz, ndx, ndy, ndz);
                                    k = dmax(0.0, before - T2);
                                    dx = ndx; dy = ndy; dz = ndz;
                                    em.secondary(0, kProcHardMoller, ex_, ey_, ez_, sdx, sdy, sdz, T2, efrac);
                                    ion += w;
                                    em.ionization(hypot(ex_, ey_), ez_, w, T2, T2);
                                    tally[3] += 1.0;
                                    {
                                        Ledger led{ledger, cutoff};
                                        led.hm_event(w, T2);
                                        led.transport_ionization(w, T2, false, true);
                                    }
                                    GroupClosure c;
                                    c.parent(0, w, before, k, true);
                                    c.particle(0, w, T2);
                                    c.ionization(w, T2, cutoff);
                                    check_group(c, cutoff, err);
                                }
                            } else if (h.channel == 1) {
                                // RreaApplyBremsFinalState OV/RreaChargedFinalStates.H:329-394
                                double const before = k;
                                if (!(finite(before) && before >= T.charged_min && before <= T.charged_max)) { err = kErrRange; break; }
                                double const ks = cdf_sample(kPos ? T.pbrems : T.ebrems, before,
                                                             rng_uniform(rng, idx, S::kBremsE), false, err);
                                double const pe = dmin(dmax(ks, 0.0), dmax(before, 0.0));
                                double const after = dmax(0.0, dmax(before, 0.0) - pe);
                                bool const tracked = finite(gcut) && gcut >= 0.0 && pe >= gcut;
                                double const ct = brems_cos(before, rng_uniform(rng, idx, S::kBremsA));
                                double gx, gy, gz;
                                cone_direction(dx, dy, dz, ct, rng_uniform(rng, idx, S::kBremsPhi), gx, gy, gz);
                                // RreaApplyPhotonRecoilToChargedDirection :465-496
                                {
                                    double const eg = dmax(pe, 0.0);
                                    if (before > 0.0 && eg > 0.0) {
                                        normalize(dx, dy, dz);
                                        double hx = gx, hy = gy, hz = gz;
                                        normalize(hx, hy, hz);
                                        double const p0 = sqrt(dmax(0.0, before * (before + 2.0 * kMeC2)));
                                        double const px = p0 * dx - eg * hx;
                                        double const py = p0 * dy - eg * hy;
                                        double const pz2 = p0 * dz - eg * hz;
                                        double const nn = sqrt(px * px + py * py + pz2 * pz2);
                                        if (nn > 0.0) { dx = px / nn; dy = py / nn; dz = pz2 / nn; normalize(dx, dy, dz); }
                                    }
                                }
                                k = after;
                                if constexpr (kPos) Ledger{ledger, cutoff}.positron_brems(w, pe);
                                else Ledger{ledger, cutoff}.brems_photon(w, pe);
                                if (tracked) {
                                    em.secondary(2, kProcBrems, ex_, ey_, ez_, gx, gy, gz, pe, efrac);
                                    tally[5] += 1.0;
                                } else if (pe > 0.0) {
                                    em.deposit_local(pe, hypot(ex_, ey_), ez_);
                                }
                                tally[4] += 1.0;
                                {
                                    GroupClosure c;
                                    c.parent(kPos ? 1 : 0, w, before, after, true);
                                    if (tracked) c.particle(2, w, pe);
                                    else c.local(w * pe);
                                    check_group(c, cutoff, err);
                                }
                            } else if (kPos && h.channel == 2) {
                                // in flight: the positron is consumed at the event point
                                spawn_annihilation(T, gcut, rng, em, kProcAnnihilation, ex_, ey_, ez_, dx, dy, dz, k, false,
                                                   idx, efrac, tally);
                                segment_to(hypot(ex_, ey_), ez_);
                                terminal_interaction = true;
                                final_f = ef;
                                break;
                            }
                            if (k <= cutoff) { mark_stopped(ef); break; }
                            rem = ef;
                            ++ord;
                        }
                    }
                    if (terminal_interaction || terminal_at_point || err) break;
                    {
                        if (k >= cutoff) {
                            // apply_gs_elastic_scatter OV/RreaTransportStep.H:372-423 on the newborn channels
                            double const gs_n = elastic_inverse_length(EL, k, err) * density_path(iv, path, sb, se);
                            if (gs_n > 0.0) {
                                if (gs_n < T.gs_few) {
                                    int const cap = static_cast<int>(ceil(3.0 * T.gs_few));
                                    if (cap >= 255) { err = kErrInvalid; break; }
                                    int const cnt = poisson_capped(gs_n, rng_uniform(rng, newborn_index(rsp, lin, sc, 0U, 0U, err),
                                                                                     kFewCount), cap);
                                    for (int q = 0; q < cnt; ++q) {
                                        uint64_t const ix = newborn_index(rsp, lin, sc, 0U, static_cast<uint64_t>(q + 1), err);
                                        if (!(finite(k) && k >= T.charged_min && k <= T.charged_max)) { err = kErrRange; break; }
                                        double const eta = elastic_semilog(EL, EL.eta, k, err);
                                        double const u = dmin(dmax(rng_uniform(rng, ix, kFewScatter), 0.0), 1.0 - 1.0e-15);
                                        double const ct = 1.0 - 2.0 * eta * u / (1.0 + eta - u);
                                        double nx, ny, nz;
                                        cone_direction(dx, dy, dz, ct, rng_uniform(rng, ix, kFewPhi), nx, ny, nz);
                                        dx = nx; dy = ny; dz = nz;
                                    }
                                } else {
                                    uint64_t const ix = newborn_index(rsp, lin, sc, 0U, 0U, err);
                                    double const u = rng_uniform(rng, ix, kGsAngle);
                                    double ct;
                                    if (!(finite(k) && k >= T.charged_min && k <= T.charged_max)) { err = kErrRange; break; }
                                    DevAngle const& A2 = k <= T.gs_transition ? (kPos ? T.gs_p : T.gs_e)
                                                                               : (kPos ? T.wz_p : T.wz_e);
                                    if (!(k >= A2.e[0] && k <= A2.e[A2.n_e - 1])) { err = kErrRange; break; }
                                    if (gs_n >= T.gs_iso) ct = 2.0 * u - 1.0;
                                    else ct = sample_gs_angle_cos(A2, k, gs_n, u, err);
                                    double nx, ny, nz;
                                    cone_direction(dx, dy, dz, ct, rng_uniform(rng, ix, kGsPhi), nx, ny, nz);
                                    dx = nx; dy = ny; dz = nz;
                                }
                            }
                        }
                        if (err) break;
                        (void)half_kick(sfield, 0.5 * substep_dt, S::kCharge, k, dx, dy, dz);
                        if (k < cutoff) {
                            final_f = se;
                            terminal_at_point = true;
                            break;
                        }
                        in_range();
                    }
                    {
                        double x, y, z;
                        point_at_full(path, se, x, y, z);
                        if (!in_half_open(x, y, z, D)) { exited = true; break; }
                        run_x = x; run_y = y; run_z = z;
                    }
                    if (path.exits) { exited = true; break; }
                    completed = sc + 1U;
                }
                // OV/RreaNewbornAdvance.cpp:686-689
                if constexpr (kPos) {
                    if (positron_total_loss > 0.0) Ledger{ledger, cutoff}.positron_transport(w, positron_total_loss, false);
                }
                if (!err) {
                    if (terminal_interaction) {
                        status = RREA_ANNIHILATED;
                        tally[20] += 1.0;
                    } else {
                        double fx_ = 0.0, fy_ = 0.0, fz_ = 0.0;
                        if (!newborn_final_point(path, D, final_f, fx_, fy_, fz_)) {
                            err = kErrInvalid;
                        } else {
                            double const final_r = hypot(fx_, fy_);
                            double const final_el = elapsed_at(final_f);
                            bool const demote = kPos ? k < cutoff : newborn_electron_demotes(A, E, fx_, fy_, fz_, k, err);
                            if (err) {
                            } else if (terminal_at_point || demote) {
                                terminate_at(fx_, fy_, fz_,
                                             newborn_index(rsp, lin, substeps < kSegmentCap - 1U ? substeps : kSegmentCap - 1U,
                                                           0U, 0U, err),
                                             final_el);
                                segment_to(final_r, fz_);
                            } else if (exited) {
                                escape_along(path);
                            } else {
                                p.x = fx_; p.y = fy_; p.z = fz_;
                                p.dir_x = dx; p.dir_y = dy; p.dir_z = dz;
                                p.energy_eV = k;
                                segment_to(final_r, fz_);
                            }
                        }
                    }
                }
            }
        }
    }
    if (err) {
        report(A.error, err, i);
        return;
    }
    p.birth_fraction = 1.0;
    N.fin[i] = p;
    N.status[i] = static_cast<uint8_t>(status);
    N.ion[i] = ion;
    N.substeps[i] = static_cast<int32_t>(nsub);
    tally[8] += ion;
    if (status == RREA_SURVIVE) tally[18] += 1.0;
}

// TransportStep::AdvancePhotonNewborn OV/RreaNewbornAdvance.cpp:779-994
__device__ __noinline__ void photon_newborn(NewbornArgs const& N, int64_t i, double* tally, double* ledger)
{
    StepArgs const& A = N.S;
    DevTables const& T = A.T;
    Env const& E = env_of_record(A, N.work[i].instance);
    DevFields const& F = E.F;
    Domain const& D = E.D;
    rrea_step_config const& cfg = A.cfg;
    double const cutoff = cfg.low_energy_cutoff_eV;
    double const gcut = cfg.photon_cutoff_eV;
    double const dt = cfg.dt_s;
    uint64_t const step = static_cast<uint64_t>(cfg.step > 0 ? cfg.step : 0);

    rrea_secondary p = N.work[i];
    int32_t err = kOk;
    double ion = 0.0;
    uint64_t lineage = 0;
    uint64_t nsub = 0;
    int status = RREA_SURVIVE;
    uint64_t const rsp = static_cast<uint64_t>(p.root_species);
    double const w = p.weight;
    uint64_t const rng = rng_prefix(E.rng_seed, p.root_key, step);
    Emitter em{A, i, p.root_key, w, ion, lineage, err, p.root_species, p.generation + 1, E.instance, &E, ledger};
    double const residual = newborn_residual(p.birth_fraction);
    double dx = p.dir_x, dy = p.dir_y, dz = p.dir_z;
    normalize(dx, dy, dz);
    double energy = p.energy_eV;
    auto escape = [&]() {
        status = RREA_ESCAPE;
        tally[19] += 1.0;
        Ledger{ledger, 0.0}.photon_escape(w, energy);
    };
    auto survive_at = [&](double x, double y, double z) {
        p.x = x; p.y = y; p.z = z;
        p.dir_x = dx; p.dir_y = dy; p.dir_z = dz;
        p.energy_eV = energy;
    };
    if (!(residual > 0.0)) {
        if (!(energy >= gcut)) status = RREA_DEMOTED;
    } else if (!in_half_open(p.x, p.y, p.z, D)) {
        err = kErrDomain;
    } else if (energy < gcut) {
        em.deposit_local(energy, hypot(p.x, p.y), p.z);
        status = RREA_DEMOTED;
        tally[20] += 1.0;
    } else {
        double const chord = kC * dt * residual;
        double const end_x = p.x + chord * dx, end_y = p.y + chord * dy, end_z = p.z + chord * dz;
        Path path;
        if (clip_rz_path(hypot(p.x, p.y), atan2(p.y, p.x), p.z, hypot(end_x, end_y), atan2(end_y, end_x), end_z, D,
                         path, err)) {
            if (!(path.full > 0.0) || !(path.in_len > 0.0)) {
                if (path.exits) escape();
            } else {
                double elapsed_start = p.birth_fraction;
                uint64_t ord = 0;
                while (!err) {
                    uint64_t const idx = newborn_index(rsp, p.lineage, 0U, ord, 0U, err);
                    if (err) break;
                    Interval const iv = build_interval(F, path, D, E.density_ratio, err);
                    PhotonRates rates;
                    if (!photon_rates(T, energy, rates, err)) break;
                    double const total = rates.compton + rates.pe + rates.pair_n + rates.pair_t;
                    if (!finite(total) || total < 0.0) { err = kErrInvalid; break; }
                    double const mfp = total > 0.0 ? 1.0 / total : 1.0e99;
                    double const tau = density_path(iv, path, 0.0, path.end_f) / mfp;
                    if (err) break;
                    Hazard const h = select_hazard(tau, 0.0, rng_uniform(rng, idx, kChNewbornCompetingOpticalDepth), 0.0);
                    if (!h.event) {
                        double x, y, z;
                        if (!newborn_final_point(path, D, path.end_f, x, y, z)) { err = kErrInvalid; break; }
                        if (path.exits) escape();
                        else survive_at(x, y, z);
                        break;
                    }
                    double const f = fraction_at_density_path(iv, path, 0.0, path.end_f, h.sampled * mfp, err);
                    if (err) break;
                    double const dist = path.full * (f - 0.0);
                    double const remaining = dmax(0.0, path.full - dist);
                    double x, y, z;
                    point_at_full(path, f, x, y, z);
                    double const r_site = hypot(x, y);
                    double const el = dmin(1.0, elapsed_start + dist / dmax(kC * dt, kDblMin));
                    // SamplePhotonChannel NA/src/RreaInteractionTables.cpp:2778-2804
                    int channel;
                    {
                        double const cu = dmax(0.0, dmin(1.0, rng_uniform(rng, idx, kChNewbornCompetingChannel)));
                        if (!(total > 0.0)) { err = kErrInvalid; break; }
                        double const target = cu * total;
                        if (target < rates.compton) channel = kCompton;
                        else if (target < rates.compton + rates.pe) channel = kPhotoelectric;
                        else if (target < rates.compton + rates.pe + rates.pair_n) channel = kPairNuclear;
                        else channel = kPairTriplet;
                    }
                    nsub = ord + 1U;
                    bool const survives = photon_final_state(T, cutoff, gcut, rng, em, channel, idx, x, y, z, el, energy,
                                                             dx, dy, dz, tally);
                    if (err) break;
                    if (!survives) {
                        status = RREA_ABSORBED;
                        tally[20] += 1.0;
                        break;
                    }
                    if (!(remaining > 0.0)) { survive_at(x, y, z); break; }
                    double const tx = x + remaining * dx, ty = y + remaining * dy, tz = z + remaining * dz;
                    Path cont;
                    if (!clip_rz_path(r_site, atan2(y, x), z, hypot(tx, ty), atan2(ty, tx), tz, D, cont, err)) break;
                    path = cont;
                    elapsed_start = el;
                    ++ord;
                }
            }
        }
    }
    if (err) {
        report(A.error, err, i);
        return;
    }
    p.birth_fraction = 1.0;
    N.fin[i] = p;
    N.status[i] = static_cast<uint8_t>(status);
    N.ion[i] = ion;
    N.substeps[i] = static_cast<int32_t>(nsub);
    tally[8] += ion;
    if (status == RREA_SURVIVE) tally[18] += 1.0;
}

// One lane per FIFO entry of the generation.  Entries already at the end of
// the step (birth fraction 1) survive untouched.
__global__ void __launch_bounds__(kBlock) newborn_kernel(const NewbornArgs* __restrict__ args)
{
    NewbornArgs const& N = *args;
    // an error recorded before the launch (the step's, or a rejected birth
    // payload) stops the drain: nothing is advanced
    bool const stopped = __hip_atomic_load(N.S.error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    // grid-stride: a launch sized for the FIFO's capacity covers a generation
    // whose bounds only the device knows.  A generation smaller than the
    // launch's lanes is spread over every wavefront, `lpw` lanes of each
    // (one entry per wavefront when there are no more entries than
    // wavefronts), so a wavefront runs its own few entries' paths rather
    // than the union of 64 divergent ones; the other lanes idle.
    int64_t const ge = stopped ? N.gb : N.ge;
    constexpr int kWaves = kBlock / 64;
    int64_t const waves = static_cast<int64_t>(gridDim.x) * kWaves;
    int64_t lpw = (ge - N.gb + waves - 1) / waves;
    lpw = lpw < 1 ? 1 : (lpw > 64 ? 64 : lpw);
    int64_t const wave = static_cast<int64_t>(blockIdx.x) * kWaves + (threadIdx.x >> 6);
    int64_t const lane = threadIdx.x & 63;
    // a workgroup with no entry (an empty generation of the device-driven
    // drain) leaves at once; the test is the same for all its lanes
    if (N.gb + static_cast<int64_t>(blockIdx.x) * kWaves * lpw >= ge) return;
    __shared__ double ledger[RREA_LEDGER_COUNT];
    ledger_begin(ledger);
    double tally[kTallyNewborn];
    for (int q = 0; q < kTallyNewborn; ++q) tally[q] = 0.0;
    for (int64_t base = N.gb + wave * lpw; base < ge; base += waves * lpw) {
        int64_t const i = base + lane;
        if (lane >= lpw || i >= ge) continue;
        rrea_secondary const& p = N.work[i];
        bool const photon = p.species == 2;
        // transient_energy_is_certified OV/RreaSecondaryCommit.cpp:148-180
        double const lo = photon ? N.S.T.photon_min : N.S.T.charged_min;
        double const hi = photon ? N.S.T.photon_max : N.S.T.charged_max;
        if (!(finite(p.energy_eV) && p.energy_eV >= lo && p.energy_eV <= hi)) {
            report(N.S.error, kErrRange, i);
        } else if (p.birth_fraction < 1.0) {
            tally[17] += 1.0;
            if (photon) photon_newborn(N, i, tally, ledger);
            else if (p.species == 1) charged_newborn<true>(N, i, tally, ledger);
            else charged_newborn<false>(N, i, tally, ledger);
        } else {
            N.fin[i] = p;
            N.status[i] = static_cast<uint8_t>(RREA_SURVIVE);
            N.ion[i] = 0.0;
            N.substeps[i] = 0;
        }
    }
    reduce_tallies<kTallyNewborn>(tally, N.S.tallies);
    ledger_flush(ledger, N.S.tallies);
}

// The FIFO's birth-payload validation (OV/RreaSecondaryCommit.cpp:338-376),
// each entry's root anchor (the first entry of its root's contiguous run) and
// the root's next free lineage (:390-394).
// `dn` (device, may be NULL) replaces n.  Grid-stride, so a launch sized for
// the capacity covers a count known only on the device.
__global__ void fifo_init_kernel(const rrea_secondary* __restrict__ work, int64_t n, const int64_t* __restrict__ dn,
                                 int64_t* __restrict__ anchor, unsigned long long* __restrict__ next,
                                 int32_t* __restrict__ error)
{
    if (dn) n = *dn;
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    rrea_secondary const& p = work[i];
    bool const finite_payload = finite(p.x) && finite(p.y) && finite(p.z) && finite(p.dir_x) && finite(p.dir_y)
        && finite(p.dir_z) && finite(p.weight) && finite(p.energy_eV) && finite(p.birth_fraction);
    bool const energy_ok = p.species == 2 ? p.energy_eV > 0.0 : p.energy_eV >= 0.0;
    if (p.species < 0 || p.species > 2 || p.root_species < 0 || p.root_species > 2 || !finite_payload
        || !(p.weight > 0.0) || !energy_ok || p.birth_fraction < 0.0 || p.birth_fraction > 1.0
        || p.lineage >= kLineageCap) {
        report(error, kErrInvalid, i);
    }
    int64_t a = i;
    while (a > 0 && work[a - 1].root_key == p.root_key && work[a - 1].root_species == p.root_species) --a;
    anchor[i] = a;
    atomicMax(&next[a], static_cast<unsigned long long>(p.lineage + 1U));
    }
}

// Children of one generation, already in FIFO order: inherit the parent's
// anchor.  The generation's FIFO range is [b, e), or be[0..1] on the device
// when be is given; grid-stride as fifo_init_kernel.
__global__ void fifo_anchor_kernel(const rrea_secondary* __restrict__ work, int64_t b, int64_t e,
                                   const int64_t* __restrict__ be, int64_t* __restrict__ anchor)
{
    if (be) { b = be[0]; e = be[1]; }
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t j = b + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < e; j += stride) {
        anchor[j] = anchor[work[j].parent_index];
    }
}

// allocate_lineage (OV/RreaSecondaryCommit.cpp:395-406) for one generation:
// the root's next free lineage plus the child's rank within its root's run
__global__ void fifo_lineage_kernel(rrea_secondary* __restrict__ work, int64_t b, int64_t e,
                                    const int64_t* __restrict__ be, const int64_t* __restrict__ anchor,
                                    const unsigned long long* __restrict__ next, int32_t* __restrict__ error)
{
    if (be) { b = be[0]; e = be[1]; }
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t j = b + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < e; j += stride) {
        int64_t const a = anchor[j];
        int64_t s = j;
        while (s > b && anchor[s - 1] == a) --s;
        uint64_t const lin = static_cast<uint64_t>(next[a]) + static_cast<uint64_t>(j - s);
        if (lin >= kLineageCap) report(error, kErrInvalid, j);
        work[j].lineage = lin;
    }
}

// The last child of each root's run advances the root's next free lineage.
__global__ void fifo_advance_kernel(int64_t b, int64_t e, const int64_t* __restrict__ be,
                                    const int64_t* __restrict__ anchor, unsigned long long* __restrict__ next)
{
    if (be) { b = be[0]; e = be[1]; }
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t j = b + static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; j < e; j += stride) {
        int64_t const a = anchor[j];
        if (j + 1 < e && anchor[j + 1] == a) continue;
        int64_t s = j;
        while (s > b && anchor[s - 1] == a) --s;
        next[a] += static_cast<unsigned long long>(j - s + 1);
    }
}

// `dn` (device, may be NULL) replaces n; grid-stride.
__global__ void fill_u64_kernel(unsigned long long* __restrict__ p, int64_t n, unsigned long long v,
                                const int64_t* __restrict__ dn = nullptr)
{
    if (dn) n = *dn;
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = v;
}

// Survivors in FIFO order: the end-of-step states of the entries that survive.
__global__ void survivor_flag_kernel(const uint8_t* __restrict__ status, int64_t n, int32_t* __restrict__ flag)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) flag[i] = status[i] == RREA_SURVIVE ? 1 : 0;
}

// Tile ordering keys: cell of the captured position (r fastest), then a
// kinetic-energy band within the cell; invalid particles last.
__global__ void tile_key_kernel(int64_t n, const double* __restrict__ r, const double* __restrict__ z,
                                const double* __restrict__ ux, const double* __restrict__ uy,
                                const double* __restrict__ uz, const int64_t* __restrict__ id, int32_t photon,
                                int32_t nr, int32_t nz, double r_lo, double z_lo, double dr, double dz, int32_t ebits,
                                uint32_t* __restrict__ keys, int32_t* __restrict__ idx)
{
    int64_t const p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    idx[p] = static_cast<int32_t>(p);
    if (id[p] == RREA_INVALID_ID) { keys[p] = 0xffffffffU; return; }
    double const fr = (r[p] - r_lo) / dr, fz = (z[p] - z_lo) / dz;
    int32_t i = finite(fr) ? static_cast<int32_t>(dmin(dmax(fr, 0.0), static_cast<double>(nr - 1))) : nr - 1;
    int32_t j = finite(fz) ? static_cast<int32_t>(dmin(dmax(fz, 0.0), static_cast<double>(nz - 1))) : nz - 1;
    double const u2 = ux[p] * ux[p] + uy[p] * uy[p] + uz[p] * uz[p];
    double const e = photon ? sqrt(u2) : kinetic_from_u(ux[p], uy[p], uz[p]);
    // four bands per octave above 1 keV at 6 bits; a narrower field merges
    // 2^(6 - ebits) neighbouring bands, so the bands still span 16 octaves
    double const band = finite(e) && e > 1.0e3 ? 4.0 * log2(e * 1.0e-3) : 0.0;
    uint32_t const emax = (1U << ebits) - 1U;
    uint32_t const q = static_cast<uint32_t>(dmin(band, 63.0)) >> (6 - ebits);
    uint32_t const eb = ebits > 0 ? (q < emax ? q : emax) : 0U;
    uint32_t const cell = static_cast<uint32_t>(i) + static_cast<uint32_t>(nr) * static_cast<uint32_t>(j);
    keys[p] = (cell << ebits) | eb;
}

// rrea_batch_tile_sort's key: the particle's instance, then its cell on
// that instance's own mesh, then the energy band (as tile_key_kernel);
// invalidated particles last.  The instances' meshes come from their
// environments.
__global__ void batch_tile_key_kernel(int64_t n, const double* __restrict__ r, const double* __restrict__ z,
                                      const double* __restrict__ ux, const double* __restrict__ uy,
                                      const double* __restrict__ uz, const int64_t* __restrict__ id,
                                      const int32_t* __restrict__ inst, int32_t photon,
                                      const Env* __restrict__ envs, int32_t n_inst, int32_t cbits, int32_t ebits,
                                      uint32_t* __restrict__ keys, int32_t* __restrict__ idx)
{
    int64_t const p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p >= n) return;
    idx[p] = static_cast<int32_t>(p);
    int32_t const k = inst[p];
    if (id[p] == RREA_INVALID_ID || k < 0 || k >= n_inst) { keys[p] = 0xffffffffU; return; }
    DevFields const& F = envs[k].F;
    double const fr = (r[p] - F.r_lo) / F.dr, fz = (z[p] - F.z_lo) / F.dz;
    int32_t const i = finite(fr) ? static_cast<int32_t>(dmin(dmax(fr, 0.0), static_cast<double>(F.nr - 1))) : F.nr - 1;
    int32_t const j = finite(fz) ? static_cast<int32_t>(dmin(dmax(fz, 0.0), static_cast<double>(F.nz - 1))) : F.nz - 1;
    double const u2 = ux[p] * ux[p] + uy[p] * uy[p] + uz[p] * uz[p];
    double const e = photon ? sqrt(u2) : kinetic_from_u(ux[p], uy[p], uz[p]);
    double const band = finite(e) && e > 1.0e3 ? 4.0 * log2(e * 1.0e-3) : 0.0;
    uint32_t const emax = (1U << ebits) - 1U;
    uint32_t const q = static_cast<uint32_t>(dmin(band, 63.0)) >> (6 - ebits);
    uint32_t const eb = ebits > 0 ? (q < emax ? q : emax) : 0U;
    uint32_t const cell = static_cast<uint32_t>(i) + static_cast<uint32_t>(F.nr) * static_cast<uint32_t>(j);
    keys[p] = (((static_cast<uint32_t>(k) << cbits) | cell) << ebits) | eb;
}

// rrea_tile_compact's key: live particles (0) before invalidated ones (1),
// and the live count (wave-aggregated).
__global__ void live_key_kernel(int64_t n, const int64_t* __restrict__ id, uint32_t* __restrict__ keys,
                                int32_t* __restrict__ idx, unsigned long long* __restrict__ live)
{
    int64_t const p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool const ok = p < n && id[p] != RREA_INVALID_ID;
    if (p < n) {
        idx[p] = static_cast<int32_t>(p);
        keys[p] = ok ? 0U : 1U;
    }
    uint64_t const m = __ballot(ok);
    if ((threadIdx.x & 63) == 0 && m) atomicAdd(live, static_cast<unsigned long long>(__popcll(m)));
}

// One species' particle store for rrea_commit_survivors (device arrays with
// room for cap particles; prev_u* null for photons; inst null for a plain run).
struct CommitStore {
    double *prev_r, *prev_z, *prev_th, *prev_ux, *prev_uy, *prev_uz;
    double *r, *z, *th, *ux, *uy, *uz, *w;
    int64_t* id;
    int32_t* cpu;
    int32_t* inst;
    int64_t cap;
};
struct CommitStores { CommitStore s[3]; };

// The species slot a survivor commits to, or -1 when CommitLocalParticleBatch
// skips it (RreaCouplingSources.cpp:280-283: non-positive weight; a photon
// with non-positive energy).
__device__ __forceinline__ int32_t commit_slot(rrea_secondary const& q)
{
    if (q.species < 0 || q.species > 2) return -1;
    if (!(q.weight > 0.0)) return -1;
    if (q.species == 2 && !(q.energy_eV > 0.0)) return -1;
    return q.species;
}

// AddNParticles of the drain's survivors, one workgroup: each species'
// accepted survivors are appended behind its count in FIFO order, and ids
// are taken from *next_id in the order the reference commits the batches --
// electrons, then photons, then positrons (RreaSecondaryCommit.cpp:594-629).
// Momentum and position as CommitLocalParticleBatch and
// append_rrea_previous_position_attributes form them
// (RreaCouplingSources.cpp:284-307, 68-74).  Nothing is written when a store
// would overflow.
__global__ void __launch_bounds__(1024) commit_kernel(const rrea_secondary* __restrict__ sv,
                                                      const int64_t* __restrict__ sv_count, CommitStores S,
                                                      int64_t* __restrict__ counts, int64_t* __restrict__ next_id,
                                                      int32_t cpu, int32_t* __restrict__ error)
{
    __shared__ unsigned long long tot[3];
    __shared__ int32_t wave_n[16][3];
    __shared__ int64_t running[3];
    __shared__ int32_t abort_flag;
    int32_t const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // fail closed: a step or drain that reported an error commits nothing
    // (RreaSecondaryCommit.cpp:240-265 refuses the transaction)
    if (__hip_atomic_load(&error[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return;
    int64_t const n = *sv_count;
    if (tid < 3) { tot[tid] = 0ULL; running[tid] = 0; }
    if (tid == 0) abort_flag = 0;
    __syncthreads();
    {
        unsigned long long c[3] = {0ULL, 0ULL, 0ULL};
        for (int64_t i = tid; i < n; i += 1024) {
            int32_t const q = commit_slot(sv[i]);
            if (q >= 0) c[q] += 1ULL;
        }
        for (int q = 0; q < 3; ++q) if (c[q]) atomicAdd(&tot[q], c[q]);
    }
    __syncthreads();
    if (tid == 0) {
        for (int q = 0; q < 3; ++q) {
            if (counts[q] + static_cast<int64_t>(tot[q]) > S.s[q].cap) abort_flag = 1;
        }
        if (abort_flag && atomicCAS(&error[0], 0, kErrCapacity) == 0) error[1] = -1;
    }
    __syncthreads();
    if (abort_flag) return;
    int64_t const id0 = *next_id;
    // commit order electrons (0), photons (2), positrons (1)
    int64_t const id_base[3] = {id0, id0 + static_cast<int64_t>(tot[0] + tot[2]), id0 + static_cast<int64_t>(tot[0])};
    for (int64_t start = 0; start < n; start += 1024) {
        int64_t const i = start + tid;
        int32_t const q = i < n ? commit_slot(sv[i]) : -1;
        uint64_t const below = (lane == 0) ? 0ULL : (~0ULL >> (64 - lane));
        int32_t rank_in_wave = 0;
        for (int k = 0; k < 3; ++k) {
            uint64_t const m = __ballot(q == k);
            if (q == k) rank_in_wave = __popcll(m & below);
            if (lane == 0) wave_n[wave][k] = __popcll(m);
        }
        __syncthreads();
        if (q >= 0) {
            int64_t rank = rank_in_wave;
            for (int w = 0; w < wave; ++w) rank += wave_n[w][q];
            CommitStore const& D = S.s[q];
            int64_t const slot = counts[q] + running[q] + rank;
            rrea_secondary const& t = sv[i];
            double const px = t.dir_x, py = t.dir_y, pz = t.dir_z;
            double const norm = sqrt(px * px + py * py + pz * pz);
            double ux = px, uy = py, uz = pz;
            if (q == 2) {
                double const scale = norm > 0.0 ? t.energy_eV / norm : t.energy_eV;
                ux = px * scale;
                uy = py * scale;
                uz = norm > 0.0 ? pz * scale : t.energy_eV;
            } else if (t.energy_eV > 0.0) {
                double const um = proper_velocity(t.energy_eV);
                if (norm > 0.0) {
                    ux = um * px / norm;
                    uy = um * py / norm;
                    uz = um * pz / norm;
                } else {
                    ux = 0.0;
                    uy = 0.0;
                    uz = um;
                }
            }
            double const r = hypot(t.x, t.y), th = atan2(t.y, t.x);
            D.prev_r[slot] = r; D.prev_z[slot] = t.z; D.prev_th[slot] = th;
            if (D.prev_ux) { D.prev_ux[slot] = ux; D.prev_uy[slot] = uy; D.prev_uz[slot] = uz; }
            D.r[slot] = r; D.z[slot] = t.z; D.th[slot] = th;
            D.ux[slot] = ux; D.uy[slot] = uy; D.uz[slot] = uz;
            D.w[slot] = t.weight;
            D.id[slot] = id_base[q] + running[q] + rank;
            if (D.cpu) D.cpu[slot] = cpu;
            if (D.inst) D.inst[slot] = t.instance;
        }
        __syncthreads();
        if (tid < 3) {
            int64_t add = 0;
            for (int w = 0; w < 16; ++w) add += wave_n[w][tid];
            running[tid] += add;
        }
        __syncthreads();
    }
    if (tid < 3) counts[tid] += static_cast<int64_t>(tot[tid]);
    if (tid == 0) *next_id = id0 + static_cast<int64_t>(tot[0] + tot[1] + tot[2]);
}

template <class V>
__global__ void gather_kernel(int64_t n, const int32_t* __restrict__ idx, const V* __restrict__ src, V* __restrict__ dst)
{
    int64_t const p = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (p < n) dst[p] = src[idx[p]];
}

__global__ void secondary_keys_kernel(const rrea_secondary* __restrict__ s, const int64_t* __restrict__ count,
                                      int64_t cap, uint64_t* __restrict__ keys, int64_t* __restrict__ idx)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const n = *count < cap ? *count : cap;
    if (i >= cap) return;
    if (i < n) {
        keys[i] = (static_cast<uint64_t>(s[i].parent_index) << 16) | (s[i].lineage & 0xffffULL);
    } else {
        keys[i] = ~0ULL;
    }
    idx[i] = i;
}

__global__ void secondary_gather_kernel(const rrea_secondary* __restrict__ in, const int64_t* __restrict__ idx,
                                        const int64_t* __restrict__ count, int64_t cap, rrea_secondary* __restrict__ out)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const n = *count < cap ? *count : cap;
    if (i < n) out[i] = in[idx[i]];
}

// The ordered records appended behind the caller's count (read on the
// device); records beyond the caller's capacity are dropped and reported.
__global__ void secondary_append_kernel(const rrea_secondary* __restrict__ in, const int64_t* __restrict__ idx,
                                        int64_t n, const int64_t* __restrict__ base, int64_t cap,
                                        rrea_secondary* __restrict__ out, int32_t* __restrict__ error)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t const at = *base + i;
    if (at < cap) out[at] = in[idx[i]];
    else if (error) report(error, kErrCapacity, -1);
}

__global__ void append_count_kernel(int64_t* __restrict__ count, const int64_t* __restrict__ base, int64_t n)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = *base + n;
}

// A host argument block put into device memory by a one-workgroup kernel
// that takes it by value: the launch copies it into the kernel's argument
// segment, so the host block is consumed when the launch returns and no
// staging copy of pageable memory runs on the stream.
template <typename T>
__global__ void put_args_kernel(T* __restrict__ dst, T const value)
{
    static_assert(sizeof(T) % sizeof(unsigned long long) == 0 && sizeof(T) <= 4000, "argument block shape");
    auto const* src = reinterpret_cast<const unsigned long long*>(&value);
    auto* out = reinterpret_cast<unsigned long long*>(dst);
    for (unsigned i = threadIdx.x; i < sizeof(T) / sizeof(unsigned long long); i += blockDim.x) out[i] = src[i];
}

__global__ void set_i64_kernel(int64_t* __restrict__ p, int64_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

__global__ void report_kernel(int32_t* __restrict__ error, int32_t code, int64_t index)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) report(error, code, index);
}

__global__ void copy_i64_kernel(int64_t* __restrict__ dst, const int64_t* __restrict__ src)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *dst = *src;
}

__global__ void set_i32_kernel(int32_t* __restrict__ p, int32_t v)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *p = v;
}

// The caller's count after an ordered append: its value at the start of the
// call plus the records staged (the staged count is not clipped, as in the
// host-ordered path).
__global__ void append_count_dev_kernel(int64_t* __restrict__ count, const int64_t* __restrict__ base,
                                        const int64_t* __restrict__ staged)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) *count = *base + *staged;
}

// Clears `n` doubles (graph-safe: a kernel, not a memset node).
__global__ void fill_f64_kernel(double* __restrict__ p, int64_t n)
{
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) p[i] = 0.0;
}

// ---------------------------------------------------------------------------
// Record ordering with the count on the device (graph-captured steps).  One
// workgroup sorts the staged records by the key the host-ordered path sorts
// on, (parent_index << 16) | (lineage & 0xffff): a stable LSD radix sort
// over 8-bit digits, as many digits as the largest key has, a pass whose
// digit is the same for every record skipped.  Within each 1024-record tile
// a record's rank among equal digits comes from wavefront ballots over the
// digit's bits and per-wavefront digit counts in LDS, so the scatter keeps
// the input order of equal keys (the order hipcub's radix sort also keeps).
// Up to kSortLds records are sorted in LDS; more go through global ping-pong
// buffers of the workspace.  Writes the permutation to `order`.
// ---------------------------------------------------------------------------
constexpr int kSortThreads = 1024;
constexpr int kSortWaves = kSortThreads / 64;
constexpr int kSortLds = 2048;

__device__ __forceinline__ uint64_t record_key(rrea_secondary const& r)
{
    return (static_cast<uint64_t>(r.parent_index) << 16) | (r.lineage & 0xffffULL);
}

// The number of records to order: the device count clipped to the staging
// capacity (host value, or the device value at cap_dev when given).
__device__ __forceinline__ int64_t staged_records(const int64_t* count, const int64_t* cap_dev, int64_t cap)
{
    int64_t const c = cap_dev ? *cap_dev : cap;
    int64_t const n = *count < c ? *count : c;
    return n > 0 ? n : 0;
}

// The sort of n records by one workgroup of kSortThreads lanes; every lane
// returns after its last write to `order` (no barrier after it).
__device__ __forceinline__ void sort_records_block(const rrea_secondary* __restrict__ in, int64_t n,
                                                   uint64_t* __restrict__ gk0, uint64_t* __restrict__ gk1,
                                                   int32_t* __restrict__ gi0, int32_t* __restrict__ gi1,
                                                   int32_t* __restrict__ order)
{
    __shared__ uint64_t lk[2][kSortLds];
    __shared__ int32_t li[2][kSortLds];
    __shared__ uint16_t wc[kSortWaves][256];
    __shared__ uint32_t hist[256], dbase[256], tot[256];
    __shared__ uint64_t wmax[kSortWaves];
    __shared__ int32_t skip[8];  // per pass, so no reset races a slower wavefront's read
    int const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    if (n == 0) return;
    bool const in_lds = n <= kSortLds;
    uint64_t* ks = in_lds ? lk[0] : gk0;
    uint64_t* kd = in_lds ? lk[1] : gk1;
    int32_t* is = in_lds ? li[0] : gi0;
    int32_t* id = in_lds ? li[1] : gi1;
    uint64_t mx = 0;
    for (int64_t i = tid; i < n; i += kSortThreads) {
        uint64_t const k = record_key(in[i]);
        ks[i] = k;
        is[i] = static_cast<int32_t>(i);
        mx = k > mx ? k : mx;
    }
    for (int off = 32; off > 0; off >>= 1) {
        uint64_t const o = __shfl_xor(mx, off, 64);
        mx = o > mx ? o : mx;
    }
    if (lane == 0) wmax[wave] = mx;
    if (tid < 8) skip[tid] = 0;
    for (int j = tid; j < kSortWaves * 256; j += kSortThreads) (&wc[0][0])[j] = 0;
    __syncthreads();
    mx = 0;
    for (int w = 0; w < kSortWaves; ++w) mx = wmax[w] > mx ? wmax[w] : mx;
    int const bits = mx ? 64 - __clzll(static_cast<long long>(mx)) : 0;
    int const passes = (bits + 7) / 8;
    uint64_t const lt = (1ULL << lane) - 1ULL;
    for (int pass = 0; pass < passes; ++pass) {
        int const shift = 8 * pass;
        if (tid < 256) hist[tid] = 0;
        __syncthreads();
        for (int64_t i = tid; i < n; i += kSortThreads) atomicAdd(&hist[(ks[i] >> shift) & 255U], 1U);
        __syncthreads();
        if (tid < 256 && static_cast<int64_t>(hist[tid]) == n) skip[pass] = 1;
        if (tid < 64) {
            // exclusive scan of the 256 digit counts, four per lane
            uint32_t const h0 = hist[4 * tid], h1 = hist[4 * tid + 1], h2 = hist[4 * tid + 2], h3 = hist[4 * tid + 3];
            uint32_t const own = h0 + h1 + h2 + h3;
            uint32_t inc = own;
            for (int off = 1; off < 64; off <<= 1) {
                uint32_t const o = __shfl_up(inc, off, 64);
                if (lane >= off) inc += o;
            }
            uint32_t const ex = inc - own;
            dbase[4 * tid] = ex;
            dbase[4 * tid + 1] = ex + h0;
            dbase[4 * tid + 2] = ex + h0 + h1;
            dbase[4 * tid + 3] = ex + h0 + h1 + h2;
        }
        __syncthreads();
        if (skip[pass]) continue;  // every record has this digit: the order stands
        for (int64_t b0 = 0; b0 < n; b0 += kSortThreads) {
            int64_t const i = b0 + tid;
            bool const valid = i < n;
            uint64_t const key = valid ? ks[i] : 0ULL;
            int32_t const idx = valid ? is[i] : 0;
            uint32_t const d = static_cast<uint32_t>(key >> shift) & 255U;
            uint64_t m = __ballot(valid);
            for (int b = 0; b < 8; ++b) {
                bool const bit = (d >> b) & 1U;
                uint64_t const bb = __ballot(bit);
                m &= bit ? bb : ~bb;
            }
            uint32_t const rank = static_cast<uint32_t>(__popcll(m & lt));
            if (valid && (m & lt) == 0ULL) wc[wave][d] = static_cast<uint16_t>(__popcll(m));
            __syncthreads();
            if (tid < 256) {
                uint32_t run = 0;
                for (int w = 0; w < kSortWaves; ++w) {
                    uint32_t const c = wc[w][tid];
                    wc[w][tid] = static_cast<uint16_t>(run);
                    run += c;
                }
                tot[tid] = run;
            }
            __syncthreads();
            if (valid) {
                uint32_t const pos = dbase[d] + wc[wave][d] + rank;
                kd[pos] = key;
                id[pos] = idx;
            }
            __syncthreads();
            if (tid < 256) {
                dbase[tid] += tot[tid];
                for (int w = 0; w < kSortWaves; ++w) wc[w][tid] = 0;
            }
            __syncthreads();
        }
        uint64_t* const tk = ks; ks = kd; kd = tk;
        int32_t* const ti = is; is = id; id = ti;
    }
    for (int64_t i = tid; i < n; i += kSortThreads) order[i] = is[i];
}

__global__ void __launch_bounds__(kSortThreads) sort_records_kernel(
    const rrea_secondary* __restrict__ in, const int64_t* __restrict__ count, const int64_t* __restrict__ cap_dev,
    int64_t cap, uint64_t* __restrict__ gk0, uint64_t* __restrict__ gk1, int32_t* __restrict__ gi0,
    int32_t* __restrict__ gi1, int32_t* __restrict__ order)
{
    sort_records_block(in, staged_records(count, cap_dev, cap), gk0, gk1, gi0, gi1, order);
}

// The ordered records appended behind the count at `base` (device; NULL: 0),
// within the destination's capacity; records beyond it are dropped and
// reported.  With `be` (device) the destination range is [be[0], be[1])
// instead: a drain generation's children, empty when the drain stopped.
__global__ void gather_records_kernel(const rrea_secondary* __restrict__ in, const int32_t* __restrict__ order,
                                      const int64_t* __restrict__ count, const int64_t* __restrict__ cap_dev,
                                      int64_t cap, const int64_t* __restrict__ base, int64_t dst_cap,
                                      rrea_secondary* __restrict__ dst, int32_t* __restrict__ error,
                                      const int64_t* __restrict__ be = nullptr)
{
    int64_t const n = be ? be[1] - be[0] : staged_records(count, cap_dev, cap);
    int64_t const b = be ? be[0] : (base ? *base : 0);
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
        int64_t const at = b + i;
        if (at < dst_cap) dst[at] = in[order[i]];
        else report(error, kErrCapacity, -1);
    }
}

// ---------------------------------------------------------------------------
// The device-driven newborn drain (graph-captured steps).  The host-driven
// drain reads each generation's child count to size the next launches; here
// the generation bounds live on the device (DrainState) and every launch is
// sized for the FIFO's capacity, finishing early when a generation is empty.
// ---------------------------------------------------------------------------
struct DrainState {
    int64_t gb, ge;         // the generation being advanced
    int64_t stage_cap;      // room for its children: capacity - ge
    int64_t be[2];          // its children's FIFO range once ordered
    int64_t cap;            // FIFO capacity
};

// The step's ordered records are generation 0.
// One workgroup: the drain's first generation bounds, the lineage counters
// of the initial entries cleared (fifo_init_kernel takes their maxima) and
// the staging count cleared for the first generation's children.
__global__ void drain_begin_kernel(NewbornArgs* __restrict__ N, DrainState* __restrict__ g,
                                   const int64_t* __restrict__ n_initial, int64_t cap,
                                   unsigned long long* __restrict__ next, int64_t* __restrict__ count)
{
    if (blockIdx.x != 0) return;
    int64_t n0 = *n_initial;
    n0 = n0 < 0 ? 0 : (n0 > cap ? cap : n0);
    for (int64_t i = threadIdx.x; i < n0; i += blockDim.x) next[i] = 0ULL;
    if (threadIdx.x != 0) return;
    g->gb = 0; g->ge = n0; g->stage_cap = cap - n0; g->cap = cap;
    g->be[0] = n0; g->be[1] = n0;
    N->gb = 0; N->ge = n0; N->S.sec_cap = cap - n0;
    *count = 0;
}

// The FIFO's length; a generation left unadvanced after the last captured
// one is a capacity error (index -2).
__global__ void drain_end_kernel(const DrainState* __restrict__ g, int64_t* __restrict__ work_count,
                                 int32_t* __restrict__ error)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    *work_count = g->ge;
    if (g->gb < g->ge) report(error, kErrCapacity, -2);
}

// One generation's children linked into the FIFO by one workgroup, after the
// generation's newborn kernel: the work of sort_records, drain_bounds,
// gather_records, fifo_anchor, fifo_lineage, fifo_advance and drain_advance
// (each phase as those kernels do it, a barrier between phases), then the
// child count cleared for the next generation.  Two launches per captured
// generation (this and the newborn kernel) instead of nine.
__global__ void __launch_bounds__(kSortThreads) link_generation_kernel(
    NewbornArgs* __restrict__ N, DrainState* __restrict__ g, const rrea_secondary* __restrict__ stage,
    int64_t* __restrict__ count, uint64_t* __restrict__ gk0, uint64_t* __restrict__ gk1, int32_t* __restrict__ gi0,
    int32_t* __restrict__ gi1, int32_t* __restrict__ order, rrea_secondary* __restrict__ work,
    int64_t* __restrict__ anchor, unsigned long long* __restrict__ next, int32_t* __restrict__ error)
{
    __shared__ int64_t range[2];
    int const tid = threadIdx.x;
    int64_t const stage_cap = g->stage_cap, cap = g->cap, ge = g->ge;
    int64_t const c = *count;
    int64_t const n = c < stage_cap ? (c > 0 ? c : 0) : stage_cap;
    sort_records_block(stage, n, gk0, gk1, gi0, gi1, order);
    if (tid == 0) {
        // drain_bounds: empty when the generation made no children, an error
        // is recorded, or they overflowed the FIFO (the newborn kernel
        // reported that)
        bool const stop = __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0 || c <= 0
            || c > stage_cap;
        range[0] = ge;
        range[1] = stop ? ge : ge + c;
    }
    __syncthreads();
    int64_t const b = range[0], e = range[1];
    for (int64_t j = b + tid; j < e; j += kSortThreads) {
        if (j < cap) work[j] = stage[order[j - b]];
        else report(error, kErrCapacity, -1);
    }
    __syncthreads();
    for (int64_t j = b + tid; j < e; j += kSortThreads) anchor[j] = anchor[work[j].parent_index];
    __syncthreads();
    for (int64_t j = b + tid; j < e; j += kSortThreads) {
        int64_t const a = anchor[j];
        int64_t s = j;
        while (s > b && anchor[s - 1] == a) --s;
        uint64_t const lin = static_cast<uint64_t>(next[a]) + static_cast<uint64_t>(j - s);
        if (lin >= kLineageCap) report(error, kErrInvalid, j);
        work[j].lineage = lin;
    }
    __syncthreads();
    for (int64_t j = b + tid; j < e; j += kSortThreads) {
        int64_t const a = anchor[j];
        if (j + 1 < e && anchor[j + 1] == a) continue;
        int64_t s = j;
        while (s > b && anchor[s - 1] == a) --s;
        next[a] += static_cast<unsigned long long>(j - s + 1);
    }
    if (tid == 0) {
        g->be[0] = b; g->be[1] = e;
        g->gb = b; g->ge = e;
        g->stage_cap = cap - e;
        N->gb = b; N->ge = e; N->S.sec_cap = cap - e;
        *count = 0;
    }
}

// Survivors in FIFO order (one workgroup): the end-of-step states of the
// entries with RREA_SURVIVE, up to the capacity (beyond it: reported, the
// count clipped); none after an error, as in the host-driven drain.
__global__ void __launch_bounds__(1024) survivors_kernel(const rrea_secondary* __restrict__ fin,
                                                         const uint8_t* __restrict__ status,
                                                         const DrainState* __restrict__ g,
                                                         rrea_secondary* __restrict__ out, int64_t out_cap,
                                                         int64_t* __restrict__ out_count, int32_t* __restrict__ error)
{
    __shared__ uint32_t wsum[16];
    __shared__ int64_t base;
    int const tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    bool const failed = __hip_atomic_load(error, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
    int64_t const total = failed ? 0 : g->ge;
    if (tid == 0) base = 0;
    __syncthreads();
    uint64_t const lt = (1ULL << lane) - 1ULL;
    for (int64_t b0 = 0; b0 < total; b0 += 1024) {
        int64_t const i = b0 + tid;
        bool const keep = i < total && status[i] == RREA_SURVIVE;
        uint64_t const m = __ballot(keep);
        if (lane == 0) wsum[wave] = static_cast<uint32_t>(__popcll(m));
        __syncthreads();
        int64_t at = base;
        for (int w = 0; w < wave; ++w) at += wsum[w];
        at += __popcll(m & lt);
        if (keep) {
            if (at < out_cap) out[at] = fin[i];
            else report(error, kErrCapacity, -1);
        }
        __syncthreads();
        if (tid == 0) {
            int64_t t = 0;
            for (int w = 0; w < 16; ++w) t += wsum[w];
            base += t;
        }
        __syncthreads();
    }
    if (tid == 0) *out_count = base < out_cap ? base : out_cap;
}

// The step number of every argument block a captured step reads.
__global__ void set_step_kernel(StepArgs* a0, StepArgs* a1, StepArgs* a2, NewbornArgs* nb, int64_t step)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (a0) a0->cfg.step = step;
    if (a1) a1->cfg.step = step;
    if (a2) a2->cfg.step = step;
    if (nb) nb->S.cfg.step = step;
}

// The live length of each growing tile at the start of a captured step:
// counts[k] clipped to the tile's capacity, into species k's argument block.
__global__ void set_lengths_kernel(StepArgs* a0, StepArgs* a1, StepArgs* a2, const int64_t* counts,
                                   int64_t cap0, int64_t cap1, int64_t cap2)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    StepArgs* const a[3] = {a0, a1, a2};
    int64_t const cap[3] = {cap0, cap1, cap2};
    for (int k = 0; k < 3; ++k) {
        if (!a[k]) continue;
        int64_t const c = counts[k];
        a[k]->n = c < 0 ? 0 : (c < cap[k] ? c : cap[k]);
    }
}

__global__ void next_step_kernel(StepArgs* a0, StepArgs* a1, StepArgs* a2, NewbornArgs* nb)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    if (a0) a0->cfg.step += 1;
    if (a1) a1->cfg.step += 1;
    if (a2) a2->cfg.step += 1;
    if (nb) nb->S.cfg.step += 1;
}

// RreaValidateSecondaryGroupConservation over caller-built groups.
__device__ int32_t validate_group(rrea_secondary_group const& g, double cutoff, double* res)
{
    bool const required = !(g.process == 0 || g.process == 1);  // continuous transport, seed
    if (!required) return 1 | 2 | 4 | 8;
    int32_t out = 1;
    if (g.closure_mode == 0) return out;  // missing
    out |= 2;
    if (g.closure_mode == 2) {
        // an explicit zero: a payload-free removal of a parent already invalid
        bool const zero = g.n_particles == 0 && g.n_sources == 0 && g.remove_parent && !g.parent_still_valid
            && !g.parent_survives && g.parent_weight == 0.0 && g.parent_energy_before_eV == 0.0
            && g.parent_energy_after_eV == 0.0 && g.local_energy_eV == 0.0 && g.escaped_energy_eV == 0.0
            && g.escaped_charge_e == 0.0;
        return zero ? (out | 4 | 8) : out;
    }
    if (g.closure_mode != 1 || g.n_particles < 0 || g.n_particles > 4 || g.n_sources < 0 || g.n_sources > 4) {
        return out;
    }
    GroupClosure c;
    c.ok = (g.parent_survives != 0) == (g.remove_parent == 0) && g.parent_species >= 0 && g.parent_species <= 2;
    c.parent(g.parent_species, g.parent_weight, g.parent_energy_before_eV, g.parent_energy_after_eV,
             g.parent_survives != 0);
    c.residual_terms(g.local_energy_eV, g.escaped_energy_eV, g.escaped_charge_e);
    for (int i = 0; i < g.n_particles; ++i) {
        rrea_group_entry const& p = g.particles[i];
        c.ok = c.ok && p.kind >= 0 && p.kind <= 2;
        c.particle(p.kind, p.weight, p.energy_eV);
    }
    for (int i = 0; i < g.n_sources; ++i) {
        rrea_group_entry const& s = g.sources[i];
        c.ok = c.ok && GroupClosure::fin(s.energy_eV);
        if (s.kind == 0) c.ionization(s.weight, s.energy_eV, cutoff);
        else if (s.kind == 1) c.positive_ion(s.weight);
        else if (s.kind == 2) c.low_electron(s.weight);
        else c.ok = false;
    }
    double qr = 0.0, er = 0.0;
    int const v = c.verdict(cutoff, &qr, &er);
    if (res) { res[0] = qr; res[1] = er; }
    if (v == 1) return out;
    out |= 4;
    return v == 0 ? (out | 8) : out;
}

__global__ void validate_groups_kernel(const rrea_secondary_group* __restrict__ g, int64_t n, double cutoff,
                                       int32_t* __restrict__ result, double* __restrict__ residuals)
{
    int64_t const i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n) return;
    result[i] = validate_group(g[i], cutoff, residuals ? residuals + 2 * i : nullptr);
}

int64_t blocks_for(int64_t n, int b) { return (n + b - 1) / b; }

}  // namespace

// ===========================================================================
// C-ABI
// ===========================================================================
extern "C" {

const char* rrea_version(void) { return "rrea-mi355x 0.1 (gfx950, fp64)"; }

rrea_status rrea_tables_create(const rrea_table_image* im, rrea_tables_t* out)
{
    if (!im || !out) return RREA_ERR_ARGUMENT;
    *out = nullptr;
    // Every charged column shares one energy grid in the production bundle;
    // the shared bracket relies on it.
    for (int t = 1; t < 11; ++t) {
        if (im->rate[t].n != im->rate[0].n
            || std::memcmp(im->rate[t].x, im->rate[0].x, sizeof(double) * im->rate[0].n) != 0) {
            return RREA_ERR_ARGUMENT;
        }
    }
    std::unique_ptr<rrea_tables_s> T(new (std::nothrow) rrea_tables_s());
    if (!T) return RREA_ERR_ARGUMENT;
    HIP_TRY(hipGetDevice(&T->device));
    DeviceBlob blob;
    int64_t const nc = im->rate[0].n;
    std::vector<double> cx(im->rate[0].x, im->rate[0].x + nc), clogx(nc);
    for (int64_t i = 0; i < nc; ++i) clogx[i] = std::log(cx[i]);  // Finalize :1650-1656
    size_t const o_cx = blob.add(cx), o_clogx = blob.add(clogx);
    HostGuide const cguide = host_guide(clogx, kChargedGuideBins);
    size_t const o_cguide = blob.add_i(cguide.idx);
    size_t o_cy[11], o_clogy[11];
    for (int t = 0; t < 11; ++t) {
        std::vector<double> y(im->rate[t].y, im->rate[t].y + nc), ly(nc);
        for (int64_t i = 0; i < nc; ++i) ly[i] = y[i] > 0.0 ? std::log(y[i]) : 0.0;
        o_cy[t] = blob.add(y);
        o_clogy[t] = blob.add(ly);
    }
    // Load :2146-2179 -- mean hard-Moller secondary by 4096-point midpoint rule
    HostCdf const moller = host_cdf(im->moller);
    std::vector<double> hm_mean(nc, 0.0);
    for (int64_t i = 0; i < nc; ++i) {
        if (!(im->rate[2].y[i] > 0.0)) continue;
        if (cx[i] < moller.primary.front() || cx[i] > moller.primary.back()) return RREA_ERR_ARGUMENT;
        double mean = 0.0;
        for (int s = 0; s < 4096; ++s) mean += host_sample_moller(moller, cx[i], (static_cast<double>(s) + 0.5) / 4096);
        hm_mean[i] = mean / 4096.0;
    }
    T->hm_mean = hm_mean;
    size_t const o_hm = blob.add(hm_mean);
    // photon rates: the four columns share one grid (duplicated edge nodes kept)
    int64_t const ng = im->rate[11].n;
    for (int t = 12; t < 15; ++t) {
        if (im->rate[t].n != ng || std::memcmp(im->rate[t].x, im->rate[11].x, sizeof(double) * ng) != 0) {
            return RREA_ERR_ARGUMENT;
        }
    }
    std::vector<double> gx(im->rate[11].x, im->rate[11].x + ng), glogx(ng);
    for (int64_t i = 0; i < ng; ++i) glogx[i] = std::log(gx[i]);
    size_t const o_gx = blob.add(gx), o_glogx = blob.add(glogx);
    HostGuide const gguide = host_guide(glogx, kElasticGuideBins);
    size_t const o_gguide = blob.add_i(gguide.idx);
    size_t o_gy[4], o_glogy[4];
    for (int t = 0; t < 4; ++t) {
        rrea_rate_column const& c = im->rate[11 + t];
        std::vector<double> y(c.y, c.y + ng), ly(ng);
        for (int64_t i = 0; i < ng; ++i) ly[i] = y[i] > 0.0 ? std::log(y[i]) : 0.0;
        o_gy[t] = blob.add(y);
        o_glogy[t] = blob.add(ly);
    }
    // photoelectric subshell groups
    rrea_photoelectric_table const& pe = im->photoelectric;
    if (pe.n_groups < 1 || !pe.primary || !pe.start || !pe.cdf_u || !pe.binding_eV) return RREA_ERR_ARGUMENT;
    int64_t const pe_rows = pe.start[pe.n_groups];
    size_t const o_pep = blob.add(std::vector<double>(pe.primary, pe.primary + pe.n_groups));
    size_t const o_peu = blob.add(std::vector<double>(pe.cdf_u, pe.cdf_u + pe_rows));
    size_t const o_peb = blob.add(std::vector<double>(pe.binding_eV, pe.binding_eV + pe_rows));
    std::vector<int32_t> pe_start(pe.start, pe.start + pe.n_groups + 1);
    size_t const o_pes = blob.add_i(pe_start);
    // pair bank
    rrea_pair_bank const& pb = im->pair;
    if (pb.n_blocks < 1 || !pb.channel || !pb.energy_eV || !pb.start || !pb.fraction || !pb.direction) {
        return RREA_ERR_ARGUMENT;
    }
    int64_t const pair_rec = pb.start[pb.n_blocks];
    if (pair_rec > INT32_MAX) return RREA_ERR_ARGUMENT;
    size_t const o_pbe = blob.add(std::vector<double>(pb.energy_eV, pb.energy_eV + pb.n_blocks));
    size_t const o_pbf = blob.add(std::vector<double>(pb.fraction, pb.fraction + 3 * pair_rec));
    size_t const o_pbd = blob.add(std::vector<double>(pb.direction, pb.direction + 9 * pair_rec));
    size_t const o_pbc = blob.add_i(std::vector<int32_t>(pb.channel, pb.channel + pb.n_blocks));
    size_t const o_pbs = blob.add_i(std::vector<int32_t>(pb.start, pb.start + pb.n_blocks + 1));
    HostCdf const ebrems = host_cdf(im->electron_brems);
    struct CdfOff { size_t p, lp, u, v, m, lm, s; int32_t n; };
    auto add_cdf = [&](HostCdf const& h) {
        CdfOff o;
        o.p = blob.add(h.primary); o.lp = blob.add(h.log_primary); o.u = blob.add(h.u); o.v = blob.add(h.v);
        o.m = blob.add(h.mean); o.lm = blob.add(h.log_mean); o.s = blob.add_i(h.start);
        o.n = static_cast<int32_t>(h.primary.size());
        return o;
    };
    HostCdf const bhabha = host_cdf(im->bhabha), pbrems = host_cdf(im->positron_brems);
    CdfOff const om = add_cdf(moller), ob = add_cdf(ebrems), obh = add_cdf(bhabha), opb = add_cdf(pbrems);
    struct AngOff { size_t e, le, n, ln, u, lur, lomc; int32_t ne, nn, nu; };
    auto add_angle = [&](rrea_angle_table const& a) {
        AngOff o;
        std::vector<double> e(a.e, a.e + a.n_e), n(a.n, a.n + a.n_n), u(a.u, a.u + a.n_u);
        std::vector<double> le(a.n_e), ln(a.n_n), lur(a.n_u, 0.0), lomc(a.n_e * a.n_n * a.n_u);
        for (int64_t i = 0; i < a.n_e; ++i) le[i] = std::log(e[i]);
        for (int64_t i = 0; i < a.n_n; ++i) ln[i] = std::log(n[i]);
        for (int64_t i = 0; i + 1 < a.n_u; ++i) lur[i] = std::log(u[i + 1] / u[i]);
        for (int64_t i = 0; i < a.n_e * a.n_n * a.n_u; ++i) lomc[i] = std::log(a.omc[i]);
        o.e = blob.add(e); o.le = blob.add(le); o.n = blob.add(n); o.ln = blob.add(ln);
        o.u = blob.add(u); o.lur = blob.add(lur); o.lomc = blob.add(lomc);
        o.ne = static_cast<int32_t>(a.n_e); o.nn = static_cast<int32_t>(a.n_n); o.nu = static_cast<int32_t>(a.n_u);
        return o;
    };
    AngOff const og = add_angle(im->gs_electron), ow = add_angle(im->wentzel_electron);
    AngOff const ogp = add_angle(im->gs_positron), owp = add_angle(im->wentzel_positron);
    struct ElOff { size_t e, le, mfp, eta, guide; int32_t n; double glo, gscale; };
    auto add_elastic = [&](rrea_elastic_table const& el) {
        ElOff o;
        std::vector<double> e(el.e, el.e + el.n), le(el.n);
        for (int64_t i = 0; i < el.n; ++i) le[i] = std::log(e[i]);
        o.e = blob.add(e); o.le = blob.add(le);
        HostGuide const g = host_guide(le, kElasticGuideBins);
        o.guide = blob.add_i(g.idx); o.glo = g.lo; o.gscale = g.scale;
        o.mfp = blob.add(std::vector<double>(el.mfp_m_stp, el.mfp_m_stp + el.n));
        o.eta = blob.add(std::vector<double>(el.screening_eta, el.screening_eta + el.n));
        o.n = static_cast<int32_t>(el.n);
        return o;
    };
    ElOff const oel = add_elastic(im->elastic_electron), oelp = add_elastic(im->elastic_positron);

    HIP_TRY(hipMalloc(&T->d_f64, blob.f64.size() * sizeof(double)));
    HIP_TRY(hipMemcpy(T->d_f64, blob.f64.data(), blob.f64.size() * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&T->d_i32, std::max<size_t>(blob.i32.size(), 1) * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(T->d_i32, blob.i32.data(), blob.i32.size() * sizeof(int32_t), hipMemcpyHostToDevice));
    double* const f = T->d_f64;
    int32_t* const ii = T->d_i32;
    DevTables& V = T->view;
    V.cx = f + o_cx; V.clogx = f + o_clogx; V.nc = static_cast<int32_t>(nc);
    V.cguide = Guide{ii + o_cguide, kChargedGuideBins, cguide.lo, cguide.scale};
    for (int t = 0; t < 11; ++t) { V.cy[t] = f + o_cy[t]; V.clogy[t] = f + o_clogy[t]; V.cmode[t] = im->rate[t].mode; }
    V.hm_mean = f + o_hm;
    V.gx = f + o_gx; V.glogx = f + o_glogx; V.ng = static_cast<int32_t>(ng);
    V.gguide = Guide{ii + o_gguide, kElasticGuideBins, gguide.lo, gguide.scale};
    for (int t = 0; t < 4; ++t) { V.gy[t] = f + o_gy[t]; V.glogy[t] = f + o_glogy[t]; V.gmode[t] = im->rate[11 + t].mode; }
    V.photon_min = im->photon_energy_range_eV[0];
    V.photon_max = im->photon_energy_range_eV[1];
    V.pe_primary = f + o_pep; V.pe_u = f + o_peu; V.pe_binding = f + o_peb; V.pe_start = ii + o_pes;
    V.pe_n = static_cast<int32_t>(pe.n_groups);
    V.pair_energy = f + o_pbe; V.pair_frac = f + o_pbf; V.pair_dir = f + o_pbd;
    V.pair_channel = ii + o_pbc; V.pair_start = ii + o_pbs; V.pair_nblocks = static_cast<int32_t>(pb.n_blocks);
    auto set_cdf = [&](DevCdf& d, CdfOff const& o) {
        d.primary = f + o.p; d.log_primary = f + o.lp; d.u = f + o.u; d.v = f + o.v;
        d.mean = f + o.m; d.log_mean = f + o.lm; d.start = ii + o.s; d.n_groups = o.n;
    };
    set_cdf(V.moller, om);
    set_cdf(V.ebrems, ob);
    set_cdf(V.bhabha, obh);
    set_cdf(V.pbrems, opb);
    auto set_angle = [&](DevAngle& d, AngOff const& o) {
        d.e = f + o.e; d.log_e = f + o.le; d.n = f + o.n; d.log_n = f + o.ln; d.u = f + o.u;
        d.log_u_ratio = f + o.lur; d.log_omc = f + o.lomc; d.n_e = o.ne; d.n_n = o.nn; d.n_u = o.nu;
    };
    set_angle(V.gs_e, og);
    set_angle(V.wz_e, ow);
    set_angle(V.gs_p, ogp);
    set_angle(V.wz_p, owp);
    auto set_elastic = [&](DevElastic& d, ElOff const& o) {
        d.e = f + o.e; d.log_e = f + o.le; d.mfp = f + o.mfp; d.eta = f + o.eta; d.n = o.n;
        d.guide = Guide{ii + o.guide, kElasticGuideBins, o.glo, o.gscale};
    };
    set_elastic(V.el_e, oel);
    set_elastic(V.el_p, oelp);
    V.hard_moller_threshold = im->hard_moller_secondary_threshold_eV;
    V.charged_min = im->charged_energy_range_eV[0];
    V.charged_max = im->charged_energy_range_eV[1];
    V.gs_few = im->gs_few_collision_threshold_n;
    V.gs_iso = im->gs_isotropic_above_path_n;
    V.gs_transition = im->gs_wentzel_transition_energy_eV;
    T->reference_density = im->reference_density_kg_m3;
    T->brems_cutoff = im->brems_absolute_cutoff_eV;
    *out = T.release();
    return RREA_OK;
}

rrea_status rrea_tables_destroy(rrea_tables_t t)
{
    if (!t) return RREA_OK;
    if (t->d_f64) (void)hipFree(t->d_f64);
    if (t->d_i32) (void)hipFree(t->d_i32);
    delete t;
    return RREA_OK;
}

rrea_status rrea_tables_hm_mean(rrea_tables_t t, double* out, int64_t n)
{
    if (!t || !out || n != static_cast<int64_t>(t->hm_mean.size())) return RREA_ERR_ARGUMENT;
    std::memcpy(out, t->hm_mean.data(), sizeof(double) * n);
    return RREA_OK;
}

rrea_status rrea_fields_create(const rrea_rz_geometry* g, rrea_fields_t* out)
{
    if (!g || !out || g->nr <= 0 || g->nz <= 0 || !(g->dr > 0.0) || !(g->dz > 0.0) || g->r_lo != 0.0) {
        return RREA_ERR_ARGUMENT;
    }
    auto* F = new (std::nothrow) rrea_fields_s();
    if (!F) return RREA_ERR_ARGUMENT;
    F->geom = *g;
    size_t const cells = static_cast<size_t>(g->nr) * g->nz;
    size_t const nodes = static_cast<size_t>(g->nr + 1) * (g->nz + 1);
    if (hipMalloc(&F->d_cell_er, cells * sizeof(double)) != hipSuccess
        || hipMalloc(&F->d_cell_ez, cells * sizeof(double)) != hipSuccess
        || hipMalloc(&F->d_btheta, nodes * sizeof(double)) != hipSuccess
        || hipMalloc(&F->d_count, sizeof(int64_t)) != hipSuccess
        || hipMalloc(&F->d_args, sizeof(StepArgs)) != hipSuccess) {
        rrea_fields_destroy(F);
        return RREA_ERR_HIP;
    }
    *out = F;
    return RREA_OK;
}

rrea_status rrea_fields_destroy(rrea_fields_t F)
{
    if (!F) return RREA_OK;
    void* ptrs[] = {F->d_cell_er, F->d_cell_ez, F->d_btheta, F->d_stage, F->d_prof, F->d_stage_sec, F->d_keys,
                    F->d_keys_out, F->d_idx, F->d_idx_out, F->d_sort_tmp, F->d_count, F->d_args,
                    F->d_slow_list, F->d_slow_count, F->d_partials, F->d_prof_guide, F->d_nargs, F->d_fin, F->d_fst,
                    F->d_fion, F->d_fsub, F->d_anchor, F->d_next, F->d_flag, F->d_nsel, F->d_sel_tmp,
                    F->d_tkeys, F->d_tkeys_out, F->d_tidx, F->d_tidx_out, F->d_tsort_tmp, F->d_tbuf, F->d_base,
                    F->d_live, F->d_dargs, F->d_gcell};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    for (auto& ev : F->events) { (void)hipEventDestroy(ev.start); (void)hipEventDestroy(ev.stop); }
    if (F->fast_event) (void)hipEventDestroy(F->fast_event);
    if (F->d_dstate) (void)hipFree(F->d_dstate);
    delete F;
    return RREA_OK;
}

rrea_status rrea_fields_signal_fast_path(rrea_fields_t F, int32_t enable)
{
    if (!F) return RREA_ERR_ARGUMENT;
    if (enable && !F->fast_event) HIP_TRY(hipEventCreateWithFlags(&F->fast_event, hipEventDisableTiming));
    F->fast_signal = enable != 0;
    return RREA_OK;
}

rrea_status rrea_fields_set_block(rrea_fields_t F, int32_t threads)
{
    if (!F || !fast_block_ok(threads)) return RREA_ERR_ARGUMENT;
    F->block = threads;
    return RREA_OK;
}

rrea_status rrea_stream_wait_fast_path(rrea_fields_t F, rrea_stream_t stream)
{
    if (!F || !F->fast_event) return RREA_ERR_ARGUMENT;
    HIP_TRY(hipStreamWaitEvent(static_cast<hipStream_t>(stream), F->fast_event, 0));
    return RREA_OK;
}

rrea_status rrea_kernel_timing(rrea_fields_t F, int32_t enable)
{
    if (!F) return RREA_ERR_ARGUMENT;
    F->timing = enable != 0;
    F->events_used = 0;
    return RREA_OK;
}

rrea_status rrea_kernel_timing_read(rrea_fields_t F, int32_t kernel, double* total_ms, int64_t* launches)
{
    if (!F || !total_ms || !launches || kernel < -1 || kernel > 6) return RREA_ERR_ARGUMENT;
    double sum = 0.0;
    int64_t count = 0;
    for (size_t i = 0; i < F->events_used; ++i) {
        if (kernel >= 0 && F->events[i].kernel != kernel) continue;
        HIP_TRY(hipEventSynchronize(F->events[i].stop));
        float ms = 0.0f;
        HIP_TRY(hipEventElapsedTime(&ms, F->events[i].start, F->events[i].stop));
        sum += ms;
        ++count;
    }
    *total_ms = sum;
    *launches = count;
    return RREA_OK;
}

rrea_status rrea_fields_update(rrea_fields_t F, const double* er, const double* ez, const double* amb_er,
                               const double* amb_ez, const double* btheta, int32_t device_pointers,
                               rrea_stream_t stream)
{
    if (!F || !er || !ez || !amb_er || !amb_ez || !btheta) return RREA_ERR_ARGUMENT;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    int32_t const nr = F->geom.nr, nz = F->geom.nz;
    size_t const n_er = static_cast<size_t>(nr + 1) * nz, n_ez = static_cast<size_t>(nr) * (nz + 1);
    size_t const n_bt = static_cast<size_t>(nr + 1) * (nz + 1);
    const double *d_er = er, *d_ez = ez, *d_aer = amb_er, *d_aez = amb_ez;
    if (!device_pointers) {
        size_t const need = 2 * n_er + 2 * n_ez;
        if (F->stage_len < need) {
            if (F->d_stage) (void)hipFree(F->d_stage);
            F->d_stage = nullptr;
            HIP_TRY(hipMalloc(&F->d_stage, need * sizeof(double)));
            F->stage_len = need;
        }
        double* st = F->d_stage;
        HIP_TRY(hipMemcpyAsync(st, er, n_er * sizeof(double), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(st + n_er, amb_er, n_er * sizeof(double), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(st + 2 * n_er, ez, n_ez * sizeof(double), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(st + 2 * n_er + n_ez, amb_ez, n_ez * sizeof(double), hipMemcpyHostToDevice, s));
        HIP_TRY(hipMemcpyAsync(F->d_btheta, btheta, n_bt * sizeof(double), hipMemcpyHostToDevice, s));
        d_er = st; d_aer = st + n_er; d_ez = st + 2 * n_er; d_aez = st + 2 * n_er + n_ez;
    } else {
        HIP_TRY(hipMemcpyAsync(F->d_btheta, btheta, n_bt * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    int64_t const cells = static_cast<int64_t>(nr) * nz;
    cell_fields_kernel<<<dim3(static_cast<unsigned>(blocks_for(cells, 256))), dim3(256), 0, s>>>(
        nr, nz, d_er, d_ez, d_aer, d_aez, F->d_cell_er, F->d_cell_ez);
    HIP_TRY(hipGetLastError());
    if (!device_pointers) HIP_TRY(hipStreamSynchronize(s));
    F->fields_set = true;
    return RREA_OK;
}

rrea_status rrea_density_set_constant(rrea_fields_t F, double ratio)
{
    if (!F || !std::isfinite(ratio)) return RREA_ERR_ARGUMENT;
    F->density_mode = 0;
    F->density_ratio = ratio;
    return RREA_OK;
}

rrea_status rrea_density_set_profile(rrea_fields_t F, const double* alt, const double* rho, int64_t n, double z0_alt,
                                     double reference_density)
{
    if (!F || !alt || !rho || n < 2 || !(reference_density > 0.0)) return RREA_ERR_ARGUMENT;
    for (int64_t i = 1; i < n; ++i) if (!(alt[i] >= alt[i - 1])) return RREA_ERR_ARGUMENT;
    if (F->d_prof) (void)hipFree(F->d_prof);
    F->d_prof = nullptr;
    HIP_TRY(hipMalloc(&F->d_prof, 2 * n * sizeof(double)));
    HIP_TRY(hipMemcpy(F->d_prof, alt, n * sizeof(double), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(F->d_prof + n, rho, n * sizeof(double), hipMemcpyHostToDevice));
    HostGuide const g = host_guide(std::vector<double>(alt, alt + n), kProfileGuideBins);
    if (!F->d_prof_guide) HIP_TRY(hipMalloc(&F->d_prof_guide, kProfileGuideBins * sizeof(int32_t)));
    HIP_TRY(hipMemcpy(F->d_prof_guide, g.idx.data(), kProfileGuideBins * sizeof(int32_t), hipMemcpyHostToDevice));
    F->prof_guide_lo = g.lo;
    F->prof_guide_scale = g.scale;
    F->h_prof_alt.assign(alt, alt + n);
    F->h_prof_rho.assign(rho, rho + n);
    F->prof_n = n;
    F->density_mode = 1;
    F->altitude_at_z0 = z0_alt;
    F->reference_density = reference_density;
    return RREA_OK;
}

}  // extern "C"

namespace {

// One charged step (either species): argument checks, the argument block,
// the launch, and the deterministic secondary ordering.
enum SpeciesId : int { kElectron = 0, kPositron = 1, kPhoton = 2 };

// The secondary-ordering workspace (staging, keys, indices, sort scratch),
// grown to `cap` records.
rrea_status order_workspace(rrea_fields_t F, int64_t cap)
{
    if (F->stage_sec_cap >= cap) return RREA_OK;
    if (cap > INT32_MAX) return RREA_ERR_ARGUMENT;
    void* ptrs[] = {F->d_stage_sec, F->d_keys, F->d_keys_out, F->d_idx, F->d_idx_out, F->d_sort_tmp};
    for (void* p : ptrs) if (p) (void)hipFree(p);
    F->d_stage_sec = nullptr; F->d_keys = nullptr; F->d_keys_out = nullptr;
    F->d_idx = nullptr; F->d_idx_out = nullptr; F->d_sort_tmp = nullptr;
    F->stage_sec_cap = 0;
    HIP_TRY(hipMalloc(&F->d_stage_sec, cap * sizeof(rrea_secondary)));
    HIP_TRY(hipMalloc(&F->d_keys, cap * sizeof(uint64_t)));
    HIP_TRY(hipMalloc(&F->d_keys_out, cap * sizeof(uint64_t)));
    HIP_TRY(hipMalloc(&F->d_idx, cap * sizeof(int64_t)));
    HIP_TRY(hipMalloc(&F->d_idx_out, cap * sizeof(int64_t)));
    size_t bytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, F->d_keys, F->d_keys_out, F->d_idx, F->d_idx_out,
                                               static_cast<int>(cap)));
    HIP_TRY(hipMalloc(&F->d_sort_tmp, bytes));
    F->sort_tmp_bytes = bytes;
    F->stage_sec_cap = cap;
    return RREA_OK;
}

// Sorts the n records staged in F->d_stage_sec (count at F->d_count) by
// (parent index, lineage) into `dst`.
rrea_status order_staged(rrea_fields_t F, int64_t n, rrea_secondary* dst, hipStream_t s,
                         const int64_t* base = nullptr, int64_t cap = 0, int32_t* error = nullptr)
{
    if (n <= 0) return RREA_OK;
    secondary_keys_kernel<<<dim3(static_cast<unsigned>(blocks_for(n, 256))), dim3(256), 0, s>>>(
        F->d_stage_sec, F->d_count, n, F->d_keys, F->d_idx);
    HIP_TRY(hipGetLastError());
    size_t bytes = F->sort_tmp_bytes;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(F->d_sort_tmp, bytes, F->d_keys, F->d_keys_out, F->d_idx, F->d_idx_out,
                                               static_cast<int>(n), 0, 64, s));
    if (base) {
        secondary_append_kernel<<<dim3(static_cast<unsigned>(blocks_for(n, 256))), dim3(256), 0, s>>>(
            F->d_stage_sec, F->d_idx_out, n, base, cap, dst, error);
    } else {
        secondary_gather_kernel<<<dim3(static_cast<unsigned>(blocks_for(n, 256))), dim3(256), 0, s>>>(
            F->d_stage_sec, F->d_idx_out, F->d_count, n, dst);
    }
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

// The fast paths' hand-over list and its counts (int32: the light and the
// heavy class of append_slow_split, or one count for the photon path, then the
// hand-over kernel's chunk cursor), sized for n particles (the counts are
// cleared by the launch sequence).
rrea_status slow_list_workspace(rrea_fields_t F, int64_t n)
{
    if (F->slow_cap < n) {
        if (F->d_slow_list) (void)hipFree(F->d_slow_list);
        if (F->d_partials) (void)hipFree(F->d_partials);
        F->d_slow_list = nullptr;
        F->d_partials = nullptr;
        F->slow_cap = 0;
        if (n > INT32_MAX) return RREA_ERR_ARGUMENT;
        HIP_TRY(hipMalloc(&F->d_slow_list, n * sizeof(int32_t)));
        // one partial row per fast-path workgroup, at the smallest workgroup size
        HIP_TRY(hipMalloc(&F->d_partials, blocks_for(n, 128) * kPartialStride * sizeof(double)));
        F->slow_cap = n;
    }
    if (!F->d_slow_count) HIP_TRY(hipMalloc(&F->d_slow_count, 4 * sizeof(int32_t)));
    return RREA_OK;
}

// One simulation's environment from its fields handle: fields, domain, the
// reference-z density ratio, the padded J grid and the RNG seed.
rrea_status make_env(rrea_fields_t F, double* jr, double* jz, double* rho, int32_t instance, uint64_t seed, Env& E,
                     double* src = nullptr)
{
    rrea_rz_geometry const& g = F->geom;
    E.F.nr = g.nr; E.F.nz = g.nz; E.F.dr = g.dr; E.F.dz = g.dz; E.F.r_lo = g.r_lo; E.F.z_lo = g.z_lo;
    E.F.cell_er = F->d_cell_er; E.F.cell_ez = F->d_cell_ez; E.F.btheta = F->d_btheta;
    E.F.density_mode = F->density_mode; E.F.density_ratio = F->density_ratio;
    E.F.prof_alt = F->d_prof; E.F.prof_rho = F->d_prof ? F->d_prof + F->prof_n : nullptr;
    E.F.prof_n = static_cast<int32_t>(F->prof_n); E.F.altitude_at_z0 = F->altitude_at_z0;
    E.F.prof_guide = Guide{F->d_prof_guide, kProfileGuideBins, F->prof_guide_lo, F->prof_guide_scale};
    E.F.reference_density = F->reference_density;
    E.D = Domain{g.r_lo, g.r_lo + g.nr * g.dr, g.z_lo, g.z_lo + g.nz * g.dz};
    // RreaParticleInteraction::Apply :29-35: the ratio at the domain's mid z.
    {
        double ratio = F->density_ratio;
        if (F->density_mode == 1) {
            // the host copy of the profile kept by rrea_density_set_profile
            std::vector<double> const& alt = F->h_prof_alt;
            std::vector<double> const& rhov = F->h_prof_rho;
            double const a = F->altitude_at_z0 + 0.5 * (E.D.z_lo + E.D.z_hi);
            if (a < alt.front() || a > alt.back()) return RREA_ERR_TABLE_RANGE;
            auto const up = std::upper_bound(alt.begin(), alt.end(), a);
            double v;
            if (up == alt.begin()) v = rhov.front();
            else if (up == alt.end()) v = rhov.back();
            else {
                size_t const hi = up - alt.begin(), lo = hi - 1;
                double const span = alt[hi] - alt[lo];
                double const t = span > 0.0 ? (a - alt[lo]) / span : 0.0;
                v = (1.0 - t) * rhov[lo] + t * rhov[hi];
            }
            ratio = v / F->reference_density;
        }
        E.density_ratio = ratio > 0.0 ? ratio : 1.0;
    }
    int32_t const pad = 2;
    E.G = DevGrid{g.nr + pad, g.nz + 2 * pad, g.dr, g.dz, g.r_lo, g.z_lo - pad * g.dz, rho, jr, jz};
    E.pnr = g.nr; E.pzb = pad; E.pnz = g.nz;
    E.instance = instance;
    E.rng_seed = seed;
    E.src = src;
    return RREA_OK;
}

// The argument block every transport kernel shares: tables, the plain step's
// environment and the step configuration.
rrea_status common_args(rrea_tables_t T, rrea_fields_t F, const rrea_step_config* cfg, double* jr, double* jz,
                        double* rho, StepArgs& A, double* src = nullptr)
{
    A.T = T->view;
    A.batch = nullptr;
    A.inst = nullptr;
    A.cfg = *cfg;
    return make_env(F, jr, jz, rho, -1, cfg->rng_seed, A.env, src);
}

// Every instance's environment, formed on the host from its fields handle
// and copied to the device (fields may have changed since the last call).
rrea_status upload_envs(rrea_batch_t B, const rrea_step_config* cfg, hipStream_t s)
{
    (void)cfg;
    size_t const n = B->fields.size();
    Env* h = reinterpret_cast<Env*>(B->h_envs.data());
    for (size_t i = 0; i < n; ++i) {
        if (!B->fields[i]->fields_set) return RREA_ERR_ARGUMENT;
        rrea_status const st = make_env(B->fields[i], B->jr[i], B->jz[i], B->rho[i], static_cast<int32_t>(i),
                                        B->seeds[i], h[i], B->src.empty() ? nullptr : B->src[i]);
        if (st != RREA_OK) return st;
    }
    // pageable source: the copy has consumed the host array when hipMemcpyAsync returns
    HIP_TRY(hipMemcpyAsync(B->d_envs, h, n * sizeof(Env), hipMemcpyHostToDevice, s));
    return RREA_OK;
}

// A batched step's instances: their environments on the device, each
// particle's instance, and the workspace handle the step runs in.
struct BatchCtx {
    rrea_fields_t work;
    const Env* d_envs;
    const int32_t* inst;
};

// What a species step's launches need beyond its device argument block.
struct SpeciesLaunch {
    int species = 0;
    bool batched = false;
    bool run = false;       // false: nothing to launch (dt <= 0 or an empty tile)
    bool order = false;     // the staged secondaries are ordered into the caller's buffer
    bool complete = false;  // every particle through the complete step
    int64_t n = 0;
    int64_t cap = 0;        // the caller's secondary capacity
};

// A species step's host half: argument checks, the argument block A and the
// workspace it needs (grown here, before any launch, so that the launches
// can be captured into a graph).
// The device buffers a species step works in beyond its inputs and outputs:
// the staging buffer of the secondaries with its count, the caller's count at
// the start of the call, the ordering's keys and indices (and hipcub's
// temporary storage for the host-counted ordering), and the fast path's
// hand-over list with its count.
struct StepWork {
    rrea_secondary* stage = nullptr;
    int64_t* count = nullptr;
    int64_t* base = nullptr;
    uint64_t* keys = nullptr;
    uint64_t* keys_out = nullptr;
    int64_t* idx = nullptr;
    int64_t* idx_out = nullptr;
    int32_t* slow_list = nullptr;
    int32_t* slow_count = nullptr;
    double* partials = nullptr;
};

StepWork work_of(rrea_fields_t W)
{
    StepWork K;
    K.stage = W->d_stage_sec; K.count = W->d_count; K.base = W->d_base;
    K.keys = W->d_keys; K.keys_out = W->d_keys_out; K.idx = W->d_idx; K.idx_out = W->d_idx_out;
    K.slow_list = W->d_slow_list; K.slow_count = W->d_slow_count; K.partials = W->d_partials;
    return K;
}

// `fixed`: the step works in these buffers, which hold room for the tile and
// the caller's secondary capacity (checked by the caller); nothing is
// allocated.  Otherwise the handle's workspace is grown as needed.
rrea_status species_prepare(int species, rrea_tables_t T, rrea_fields_t F, const rrea_tile* tile,
                            const rrea_step_config* cfg, const rrea_step_out* out, const BatchCtx* batch,
                            StepArgs& A, SpeciesLaunch& L, const StepWork* fixed = nullptr)
{
    bool const batched = batch != nullptr;
    bool const positron = species == kPositron;
    bool const photon = species == kPhoton;
    if (!T || !F || !tile || !cfg || !out || !F->fields_set) return RREA_ERR_ARGUMENT;
    if (!out->secondaries || !out->secondary_count || !out->tallies || !out->error || out->secondary_capacity < 0) {
        return RREA_ERR_ARGUMENT;
    }
    // the photon step deposits no current; a batched step deposits into its instances' grids
    if (!photon && !batched && (!out->jr || !out->jz || !out->rho)) return RREA_ERR_ARGUMENT;
    if (batched && (!batch->work || !batch->d_envs || (tile->n > 0 && !batch->inst))) return RREA_ERR_ARGUMENT;
    if (tile->n < 0 || (tile->n > 0 && (!tile->prev_r || !tile->prev_z || !tile->prev_theta || !tile->w || !tile->id
                                         || !tile->r || !tile->z || !tile->theta || !tile->ux || !tile->uy || !tile->uz
                                         || (!photon && (!tile->prev_ux || !tile->prev_uy || !tile->prev_uz))))) {
        return RREA_ERR_ARGUMENT;
    }
    // a photon's chord end is the tile's pushed point, read before it is written
    if (photon && tile->n > 0 && (tile->prev_r == tile->r || tile->prev_z == tile->z || tile->prev_theta == tile->theta)) {
        return RREA_ERR_ARGUMENT;
    }
    L = SpeciesLaunch{};
    L.species = species;
    L.batched = batched;
    // RreaParticleInteraction::Apply :38-40 -- a non-positive dt does nothing.
    if (!(cfg->dt_s > 0.0) || tile->n == 0) return RREA_OK;
    if (!(cfg->max_optical_depth_per_substep > 0.0) || cfg->max_subcycles == 0 || cfg->max_subcycles > 16383) {
        return RREA_ERR_ARGUMENT;
    }
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != T->device) return RREA_ERR_ARGUMENT;

    rrea_fields_t const W = batched ? batch->work : F;  // the call's workspace
    {
        rrea_status const cs = common_args(T, F, cfg, out->jr, out->jz, out->rho, A, batched ? nullptr : out->sources);
        if (cs != RREA_OK) return cs;
    }
    if (batched) {
        A.batch = batch->d_envs;
        A.inst = batch->inst;
    }
    A.n = tile->n;
    A.prev_r = tile->prev_r; A.prev_z = tile->prev_z; A.prev_th = tile->prev_theta;
    A.prev_ux = tile->prev_ux; A.prev_uy = tile->prev_uy; A.prev_uz = tile->prev_uz;
    A.w = tile->w; A.id = tile->id; A.cpu = tile->cpu;
    A.r = tile->r; A.z = tile->z; A.th = tile->theta; A.ux = tile->ux; A.uy = tile->uy; A.uz = tile->uz;
    A.status = out->status; A.ion_pairs = out->ion_pairs; A.substeps = out->substeps;
    A.tallies = out->tallies; A.error = out->error;

    bool const order = out->order_secondaries != 0 && out->secondary_capacity > 0;
    int64_t const cap = out->secondary_capacity;
    if (!fixed) {
        if (order) {
            rrea_status const ws = order_workspace(W, cap);
            if (ws != RREA_OK) return ws;
            if (!W->d_base) HIP_TRY(hipMalloc(&W->d_base, sizeof(int64_t)));
        }
        if (!cfg->complete_path_only) {
            rrea_status const ls = slow_list_workspace(W, tile->n);
            if (ls != RREA_OK) return ls;
        }
    }
    StepWork const K = fixed ? *fixed : work_of(W);
    if (order) {
        // stage into the workspace (count from 0), then sort into the caller's
        // buffer behind its current count, which stays on the device
        A.sec = K.stage;
        A.sec_count = K.count;
    } else {
        A.sec = out->secondaries;
        A.sec_count = out->secondary_count;
    }
    A.sec_cap = cap;
    L.run = true;
    L.order = order;
    L.complete = cfg->complete_path_only != 0;
    L.n = tile->n;
    L.cap = cap;
    return RREA_OK;
}

// The hand-over kernels run at one wave per SIMD (they hold ~290 registers),
// so 1024 one-wave workgroups are all the chip runs of them at once; they
// stride over the hand-over list, and a larger grid would only queue empty
// workgroups behind the busy ones.
constexpr int64_t kSlowBlocks = 1024;

// The fold kernel's grid over `rows` partial rows: sixteen passes of sixteen
// rows per workgroup, at most 64 workgroups.  Each fold workgroup books its
// column totals with one atomic per column, all workgroups on the same
// addresses, so few workgroups (each striding over many rows) keep that
// contention low.
int64_t fold_blocks(int64_t rows)
{
    int64_t const b = blocks_for(rows, 16 * (kBlock / kPartialStride));
    return b < 64 ? b : 64;
}

// The fast kernels at the workgroup size kb (128, 256 or 512 lanes) over nb
// workgroups.
template <bool kPos, bool kBatched>
void launch_charged_fast(int32_t kb, int64_t nb, hipStream_t s, const StepArgs* d_args, StepWork const& K)
{
    dim3 const g(static_cast<unsigned>(nb));
    if (kb == 128) {
        charged_fast_kernel<kPos, kBatched, 128><<<g, dim3(128), 0, s>>>(d_args, K.slow_list, K.slow_count, K.partials);
    } else if (kb == 512) {
        charged_fast_kernel<kPos, kBatched, 512><<<g, dim3(512), 0, s>>>(d_args, K.slow_list, K.slow_count, K.partials);
    } else {
        charged_fast_kernel<kPos, kBatched, 256><<<g, dim3(256), 0, s>>>(d_args, K.slow_list, K.slow_count, K.partials);
    }
}

template <bool kBatched>
void launch_photon_fast(int32_t kb, int64_t nb, hipStream_t s, const StepArgs* d_args, StepWork const& K)
{
    dim3 const g(static_cast<unsigned>(nb));
    if (kb == 128) {
        photon_fast_kernel<kBatched, 128><<<g, dim3(128), 0, s>>>(d_args, K.slow_list, K.slow_count, K.partials);
    } else if (kb == 512) {
        photon_fast_kernel<kBatched, 512><<<g, dim3(512), 0, s>>>(d_args, K.slow_list, K.slow_count, K.partials);
    } else {
        photon_fast_kernel<kBatched, 256><<<g, dim3(256), 0, s>>>(d_args, K.slow_list, K.slow_count, K.partials);
    }
}

// A species step's launches on `s`, reading the argument block at d_args.
// With device_order the staged secondaries are ordered with their count on
// the device (sort_records_kernel), so the sequence makes no host round trip
// and can be captured into a graph; otherwise the count is read back once and
// hipcub sorts exactly that many records.
// W: the handle whose workspace the launches use, and whose kernel timing and
// fast-path signal apply; nullptr with `fixed` (no timing, no signal).
rrea_status species_launch(rrea_fields_t W, const StepArgs* d_args, SpeciesLaunch const& L, const rrea_step_out* out,
                           hipStream_t s, bool device_order, const StepWork* fixed = nullptr)
{
    StepWork const K = fixed ? *fixed : work_of(W);
    bool const timing = W && W->timing;
    bool const signal = W && W->fast_signal;
    if (!L.run) return RREA_OK;
    bool const positron = L.species == kPositron;
    bool const photon = L.species == kPhoton;
    bool const batched = L.batched;
    bool const order = L.order;
    int64_t const cap = L.cap;
    if (order) {
        copy_i64_kernel<<<dim3(1), dim3(64), 0, s>>>(K.base, out->secondary_count);
        fill_u64_kernel<<<dim3(1), dim3(64), 0, s>>>(reinterpret_cast<unsigned long long*>(K.count), 1, 0ULL);
        HIP_TRY(hipGetLastError());
    }
    if (!L.complete) {
        // both classes' counts and the chunk cursor (int32 in two aligned words)
        fill_u64_kernel<<<dim3(1), dim3(64), 0, s>>>(reinterpret_cast<unsigned long long*>(K.slow_count), 2, 0ULL);
        HIP_TRY(hipGetLastError());
    }
    // the fast kernels' workgroup size: the handle's setting (the fixed
    // workspace of the no-allocation step is laid out for 256)
    int32_t const kb = (W && !fixed) ? W->block : 256;
    int64_t const nb = blocks_for(L.n, L.complete ? kBlock : kb);
    // HIP events around each transport kernel while timing is enabled
    auto timed = [&](int32_t kernel, auto&& launch) -> rrea_status {
        hipEvent_t stop = nullptr;
        if (timing) {
            if (W->events_used == W->events.size()) {
                hipEvent_t a, b;
                HIP_TRY(hipEventCreate(&a));
                HIP_TRY(hipEventCreate(&b));
                W->events.push_back({a, b, 0});
            }
            rrea_fields_s::Timed& t = W->events[W->events_used++];
            t.kernel = kernel;
            HIP_TRY(hipEventRecord(t.start, s));
            stop = t.stop;
        }
        launch();
        HIP_TRY(hipGetLastError());
        if (stop) HIP_TRY(hipEventRecord(stop, s));
        return RREA_OK;
    };
    rrea_status st = RREA_OK;
    int32_t const kid = positron ? 2 : 0;   // timing ids: main kernel, then its hand-over kernel
    if (photon && L.complete) {
        st = timed(4, [&] { photon_kernel<<<dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s>>>(d_args); });
    } else if (photon) {
        st = timed(4, [&] {
            if (batched) launch_photon_fast<true>(kb, nb, s, d_args, K);
            else launch_photon_fast<false>(kb, nb, s, d_args, K);
        });
        if (st != RREA_OK) return st;
        if (signal) HIP_TRY(hipEventRecord(W->fast_event, s));
        int64_t const slow_blocks = nb < kSlowBlocks ? nb : kSlowBlocks;
        st = timed(5, [&] {
            photon_fold_kernel<<<dim3(static_cast<unsigned>(fold_blocks(nb))), dim3(kBlock), 0, s>>>(
                d_args, K.partials, nb);
            photon_slow_kernel<<<dim3(static_cast<unsigned>(slow_blocks)), dim3(kSlowThreads), 0, s>>>(
                d_args, K.slow_list, K.slow_count);
        });
    } else if (L.complete) {
        st = timed(kid, [&] {
            if (positron) positron_full_kernel<<<dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s>>>(d_args);
            else electron_full_kernel<<<dim3(static_cast<unsigned>(nb)), dim3(kBlock), 0, s>>>(d_args);
        });
    } else {
        st = timed(kid, [&] {
            if (positron) {
                if (batched) launch_charged_fast<true, true>(kb, nb, s, d_args, K);
                else launch_charged_fast<true, false>(kb, nb, s, d_args, K);
            } else {
                if (batched) launch_charged_fast<false, true>(kb, nb, s, d_args, K);
                else launch_charged_fast<false, false>(kb, nb, s, d_args, K);
            }
        });
        if (st != RREA_OK) return st;
        if (signal) HIP_TRY(hipEventRecord(W->fast_event, s));
        int64_t const slow_blocks = nb < kSlowBlocks ? nb : kSlowBlocks;
        st = timed(kid + 1, [&] {
            if (positron) {
                charged_fold_kernel<true><<<dim3(static_cast<unsigned>(fold_blocks(nb))), dim3(kBlock), 0, s>>>(
                    d_args, K.partials, nb);
                charged_slow_kernel<true><<<dim3(static_cast<unsigned>(slow_blocks)), dim3(kSlowThreads), 0, s>>>(
                    d_args, K.slow_list, K.slow_count);
            } else {
                charged_fold_kernel<false><<<dim3(static_cast<unsigned>(fold_blocks(nb))), dim3(kBlock), 0, s>>>(
                    d_args, K.partials, nb);
                charged_slow_kernel<false><<<dim3(static_cast<unsigned>(slow_blocks)), dim3(kSlowThreads), 0, s>>>(
                    d_args, K.slow_list, K.slow_count);
            }
        });
    }
    if (st != RREA_OK) return st;

    if (order && device_order) {
        sort_records_kernel<<<dim3(1), dim3(kSortThreads), 0, s>>>(
            K.stage, K.count, nullptr, cap, K.keys, K.keys_out, reinterpret_cast<int32_t*>(K.idx),
            reinterpret_cast<int32_t*>(K.idx) + cap, reinterpret_cast<int32_t*>(K.idx_out));
        HIP_TRY(hipGetLastError());
        gather_records_kernel<<<dim3(static_cast<unsigned>(std::min<int64_t>(blocks_for(cap, 256), 1024))), dim3(256),
                                0, s>>>(K.stage, reinterpret_cast<const int32_t*>(K.idx_out), K.count,
                                        nullptr, cap, K.base, cap, out->secondaries, out->error);
        HIP_TRY(hipGetLastError());
        append_count_dev_kernel<<<dim3(1), dim3(64), 0, s>>>(out->secondary_count, K.base, K.count);
        HIP_TRY(hipGetLastError());
    } else if (order) {
        if (!W || fixed) return RREA_ERR_ARGUMENT;  // the host-counted ordering sorts in the handle's workspace
        // the one host round trip of the call: the sort's item count
        int64_t nstaged = 0;
        HIP_TRY(hipMemcpyAsync(&nstaged, W->d_count, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        int64_t const nsort = nstaged < cap ? nstaged : cap;
        {
            rrea_status const os = order_staged(W, nsort, out->secondaries, s, W->d_base, cap, out->error);
            if (os != RREA_OK) return os;
        }
        append_count_kernel<<<dim3(1), dim3(64), 0, s>>>(out->secondary_count, W->d_base, nstaged);
        HIP_TRY(hipGetLastError());
    }
    return RREA_OK;
}

rrea_status species_step(int species, rrea_tables_t T, rrea_fields_t F, const rrea_tile* tile,
                         const rrea_step_config* cfg, const rrea_step_out* out, rrea_stream_t stream,
                         const BatchCtx* batch = nullptr)
{
    StepArgs A;
    SpeciesLaunch L;
    rrea_status const ps = species_prepare(species, T, F, tile, cfg, out, batch, A, L);
    if (ps != RREA_OK || !L.run) return ps;
    rrea_fields_t const W = batch ? batch->work : F;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    put_args_kernel<StepArgs><<<dim3(1), dim3(256), 0, s>>>(static_cast<StepArgs*>(W->d_args), A);
    HIP_TRY(hipGetLastError());
    // order_secondaries 2: the count stays on the device (no host round trip)
    return species_launch(W, static_cast<const StepArgs*>(W->d_args), L, out, s, out->order_secondaries == 2);
}

// ---------------------------------------------------------------------------
// The species step without allocation (rrea_transport_step_noalloc).  Its
// workspace -- the argument block, the counts, the hand-over list, the fast
// path's partial rows and the ordering buffers -- lives in one file-scope device array of fixed size,
// `scratch`, so the call makes no hipMalloc.  Each device holds its own copy
// of the array; its address is looked up once per device with
// hipGetSymbolAddress and only that address is handed to kernels and copies.
// Every call on a device shares that device's array: calls that overlap on
// one device race.  The counts that must start at zero are cleared by kernels
// at the start of each call's launch sequence.
// ---------------------------------------------------------------------------
constexpr int64_t kNoAllocMaxParticles = int64_t(1) << 20;   // particles per tile
constexpr int64_t kNoAllocMaxSecondaries = int64_t(1) << 15; // secondary_capacity when ordering
constexpr size_t kNoAllocMaxDevices = 64;

constexpr size_t round256(size_t b) { return (b + 255) & ~static_cast<size_t>(255); }
constexpr size_t kNaArgs = 0;
constexpr size_t kNaCount = kNaArgs + round256(sizeof(StepArgs));
constexpr size_t kNaBase = kNaCount + 256;
constexpr size_t kNaSlowCount = kNaBase + 256;
constexpr size_t kNaSlowList = kNaSlowCount + 256;
constexpr size_t kNaPartials = kNaSlowList + round256(kNoAllocMaxParticles * sizeof(int32_t));
constexpr size_t kNaStage = kNaPartials + round256((kNoAllocMaxParticles / kBlock) * kPartialStride * sizeof(double));
constexpr size_t kNaKeys = kNaStage + round256(kNoAllocMaxSecondaries * sizeof(rrea_secondary));
constexpr size_t kNaKeysOut = kNaKeys + round256(kNoAllocMaxSecondaries * sizeof(uint64_t));
constexpr size_t kNaIdx = kNaKeysOut + round256(kNoAllocMaxSecondaries * sizeof(uint64_t));
constexpr size_t kNaIdxOut = kNaIdx + round256(kNoAllocMaxSecondaries * sizeof(int64_t));
constexpr size_t kNoAllocBytes = kNaIdxOut + round256(kNoAllocMaxSecondaries * sizeof(int64_t));  // ~9 MiB

static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

// The device address of this device's `scratch`, looked up on the first call
// on the device (with that device current) and kept; nullptr on failure.
unsigned char* noalloc_scratch(int dev)
{
    static std::mutex mu;
    static unsigned char* ptr[kNoAllocMaxDevices] = {};
    if (dev < 0 || static_cast<size_t>(dev) >= kNoAllocMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return nullptr;
        ptr[dev] = static_cast<unsigned char*>(p);
    }
    return ptr[dev];
}

rrea_status noalloc_step(int species, rrea_tables_t T, rrea_fields_t F, const rrea_tile* tile,
                         const rrea_step_config* cfg, const rrea_step_out* out, rrea_stream_t stream)
{
    if (species < kElectron || species > kPhoton || !tile || !out || !cfg) return RREA_ERR_ARGUMENT;
    if (tile->n > kNoAllocMaxParticles) return RREA_ERR_ARGUMENT;
    // the ordering always keeps its count on the device (order_secondaries 2)
    bool const order = out->order_secondaries != 0 && out->secondary_capacity > 0;
    if (order && out->secondary_capacity > kNoAllocMaxSecondaries) return RREA_ERR_ARGUMENT;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    unsigned char* const base = noalloc_scratch(dev);
    if (!base) return RREA_ERR_HIP;
    StepWork K;
    K.count = reinterpret_cast<int64_t*>(base + kNaCount);
    K.base = reinterpret_cast<int64_t*>(base + kNaBase);
    K.slow_count = reinterpret_cast<int32_t*>(base + kNaSlowCount);
    K.slow_list = reinterpret_cast<int32_t*>(base + kNaSlowList);
    K.partials = reinterpret_cast<double*>(base + kNaPartials);
    K.stage = reinterpret_cast<rrea_secondary*>(base + kNaStage);
    K.keys = reinterpret_cast<uint64_t*>(base + kNaKeys);
    K.keys_out = reinterpret_cast<uint64_t*>(base + kNaKeysOut);
    K.idx = reinterpret_cast<int64_t*>(base + kNaIdx);
    K.idx_out = reinterpret_cast<int64_t*>(base + kNaIdxOut);
    StepArgs A;
    SpeciesLaunch L;
    rrea_status const ps = species_prepare(species, T, F, tile, cfg, out, nullptr, A, L, &K);
    if (ps != RREA_OK || !L.run) return ps;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    StepArgs* const d_args = reinterpret_cast<StepArgs*>(base + kNaArgs);
    put_args_kernel<StepArgs><<<dim3(1), dim3(256), 0, s>>>(d_args, A);
    HIP_TRY(hipGetLastError());
    return species_launch(nullptr, d_args, L, out, s, true, &K);
}


// The drain's workspace for a FIFO of `cap` records: the ordering workspace,
// the per-entry end states, statuses, ion pairs and substeps, the lineage
// anchors and counters, the survivor selection and the argument block.
rrea_status fifo_workspace(rrea_fields_t W, int64_t cap)
{
    {
        rrea_status const ws = order_workspace(W, cap);
        if (ws != RREA_OK) return ws;
    }
    if (W->fifo_cap < cap) {
        void* ptrs[] = {W->d_fin, W->d_fst, W->d_fion, W->d_fsub, W->d_anchor, W->d_next, W->d_flag, W->d_sel_tmp};
        for (void* p : ptrs) if (p) (void)hipFree(p);
        W->d_fin = nullptr; W->d_fst = nullptr; W->d_fion = nullptr; W->d_fsub = nullptr;
        W->d_anchor = nullptr; W->d_next = nullptr;
        W->d_flag = nullptr; W->d_sel_tmp = nullptr;
        W->fifo_cap = 0;
        HIP_TRY(hipMalloc(&W->d_fin, cap * sizeof(rrea_secondary)));
        HIP_TRY(hipMalloc(&W->d_fst, cap * sizeof(uint8_t)));
        HIP_TRY(hipMalloc(&W->d_fion, cap * sizeof(double)));
        HIP_TRY(hipMalloc(&W->d_fsub, cap * sizeof(int32_t)));
        HIP_TRY(hipMalloc(&W->d_anchor, cap * sizeof(int64_t)));
        HIP_TRY(hipMalloc(&W->d_next, cap * sizeof(unsigned long long)));
        HIP_TRY(hipMalloc(&W->d_flag, cap * sizeof(int32_t)));
        size_t bytes = 0;
        HIP_TRY(hipcub::DeviceSelect::Flagged(nullptr, bytes, W->d_fin, W->d_flag, W->d_stage_sec, W->d_nsel,
                                              static_cast<int>(cap)));
        HIP_TRY(hipMalloc(&W->d_sel_tmp, bytes));
        W->sel_tmp_bytes = bytes;
        W->fifo_cap = cap;
    }
    if (!W->d_nsel) HIP_TRY(hipMalloc(&W->d_nsel, sizeof(int64_t)));
    if (!W->d_nargs) HIP_TRY(hipMalloc(&W->d_nargs, sizeof(NewbornArgs)));
    return RREA_OK;
}

// Up to three species' ordered records, concatenated on the device in the
// order given (Apply's: electrons, photons, positrons) -- what the host-side
// staging does with counts read back (DrainOutputs.stage).
struct StagedLists {
    const rrea_secondary* rec[3];
    const int64_t* count[3];
    int64_t cap[3];
    int32_t n;
};

__global__ void concat_staged_kernel(StagedLists S, rrea_secondary* __restrict__ work, int64_t work_cap,
                                     int64_t* __restrict__ total, int32_t* __restrict__ error)
{
    int64_t off[4] = {0, 0, 0, 0};
    for (int k = 0; k < S.n; ++k) {
        int64_t c = *S.count[k];
        c = c < 0 ? 0 : (c > S.cap[k] ? S.cap[k] : c);
        off[k + 1] = off[k] + c;
    }
    int64_t const all = off[S.n];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *total = all < work_cap ? all : work_cap;
        if (all > work_cap) report(error, kErrCapacity, -1);
    }
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t i = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < all && i < work_cap;
         i += stride) {
        int k = 0;
        while (k + 1 < S.n && i >= off[k + 1]) ++k;
        work[i] = S.rec[k][i - off[k]];
    }
}

// The device-driven drain's launches: no host round trip, so the sequence
// can be captured into a graph.  The argument block at d_nargs holds all but
// the generation bounds, which drain_begin/advance_kernel write on the
// device; max_gen generations are launched, each sized for the FIFO's
// capacity and finishing at once when its generation is empty.  The FIFO's
// records, order, lineages and survivors are those of drain_impl; the
// per-entry status, ion pairs and substeps stay in the workspace.
rrea_status drain_device_launch(rrea_fields_t W, NewbornArgs* d_nargs, DrainState* d_state, rrea_secondary* work,
                                int64_t cap, const int64_t* n_initial, int64_t* work_count,
                                rrea_secondary* survivors, int64_t survivor_cap, int64_t* survivor_count,
                                int32_t* error, int32_t max_gen, hipStream_t s)
{
    unsigned const g256 = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(blocks_for(cap, 256), 1024)));
    // 256 workgroups stride over a generation of any size (the drain's
    // generations are thousands of entries at most in the configurations);
    // an empty generation's launch then costs a few microseconds
    unsigned const gnb = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(blocks_for(cap, kBlock), 256)));
    int32_t* const order = reinterpret_cast<int32_t*>(W->d_idx_out);
    drain_begin_kernel<<<dim3(1), dim3(1024), 0, s>>>(d_nargs, d_state, n_initial, cap, W->d_next, W->d_count);
    fifo_init_kernel<<<dim3(g256), dim3(256), 0, s>>>(work, 0, &d_state->ge, W->d_anchor, W->d_next, error);
    HIP_TRY(hipGetLastError());
    for (int32_t gen = 0; gen < max_gen; ++gen) {
        newborn_kernel<<<dim3(gnb), dim3(kBlock), 0, s>>>(d_nargs);
        link_generation_kernel<<<dim3(1), dim3(kSortThreads), 0, s>>>(
            d_nargs, d_state, W->d_stage_sec, W->d_count, W->d_keys, W->d_keys_out,
            reinterpret_cast<int32_t*>(W->d_idx), reinterpret_cast<int32_t*>(W->d_idx) + cap, order, work,
            W->d_anchor, W->d_next, error);
        HIP_TRY(hipGetLastError());
    }
    drain_end_kernel<<<dim3(1), dim3(64), 0, s>>>(d_state, work_count, error);
    survivors_kernel<<<dim3(1), dim3(1024), 0, s>>>(W->d_fin, W->d_fst, d_state, survivors, survivor_cap,
                                                    survivor_count, error);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

// The drain of one plain step (batch == nullptr) or of a batched step, whose
// records carry their instance.
rrea_status drain_impl(rrea_tables_t T, rrea_fields_t F, const rrea_step_config* cfg, const rrea_drain_io* io,
                       rrea_stream_t stream, const BatchCtx* batch)
{
    if (!T || !F || !cfg || !io || !F->fields_set) return RREA_ERR_ARGUMENT;
    if (!io->work || !io->work_count || !io->survivors || !io->survivor_count || !io->tallies || !io->error
        || (!batch && (!io->jr || !io->jz || !io->rho))) {
        return RREA_ERR_ARGUMENT;
    }
    int64_t const cap = io->work_capacity;
    if (cap < 0 || cap > INT32_MAX || io->n_initial < 0 || io->n_initial > cap || io->survivor_capacity < 0) {
        return RREA_ERR_ARGUMENT;
    }
    hipStream_t const s = static_cast<hipStream_t>(stream);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != T->device) return RREA_ERR_ARGUMENT;
    // RreaParticleInteraction::Apply :38-40 -- a non-positive dt does nothing.
    if (!(cfg->dt_s > 0.0) || io->n_initial == 0) {
        set_i64_kernel<<<dim3(1), dim3(64), 0, s>>>(io->work_count, 0);
        set_i64_kernel<<<dim3(1), dim3(64), 0, s>>>(io->survivor_count, 0);
        HIP_TRY(hipGetLastError());
        return RREA_OK;
    }
    if (!(cfg->max_optical_depth_per_substep > 0.0) || cfg->max_subcycles == 0 || cfg->max_subcycles > 16383) {
        return RREA_ERR_ARGUMENT;
    }
    rrea_fields_t const W = batch ? batch->work : F;  // the call's workspace
    StepArgs A{};
    {
        rrea_status const cs = common_args(T, F, cfg, io->jr, io->jz, io->rho, A, batch ? nullptr : io->sources);
        if (cs != RREA_OK) return cs;
    }
    A.tallies = io->tallies;
    A.error = io->error;
    if (batch) A.batch = batch->d_envs;
    {
        rrea_status const ws = fifo_workspace(W, cap);
        if (ws != RREA_OK) return ws;
    }
    int64_t const n0 = io->n_initial;
    fill_u64_kernel<<<dim3(static_cast<unsigned>(blocks_for(n0, 256))), dim3(256), 0, s>>>(W->d_next, n0, 0ULL);
    HIP_TRY(hipGetLastError());
    fifo_init_kernel<<<dim3(static_cast<unsigned>(blocks_for(n0, 256))), dim3(256), 0, s>>>(
        io->work, n0, nullptr, W->d_anchor, W->d_next, io->error);
    HIP_TRY(hipGetLastError());
    // One launch per generation; the host reads the child count in between.
    // The generation kernel advances nothing once an error is recorded (a
    // rejected birth payload included).
    int32_t err_host[2] = {0, 0};
    int64_t gb = 0, ge = n0;
    while (gb < ge) {
        fill_u64_kernel<<<dim3(1), dim3(64), 0, s>>>(reinterpret_cast<unsigned long long*>(W->d_count), 1, 0ULL);
        HIP_TRY(hipGetLastError());
        A.sec = W->d_stage_sec;
        A.sec_cap = cap - ge;
        A.sec_count = W->d_count;
        NewbornArgs N{A, io->work, gb, ge, W->d_fin, W->d_fst, W->d_fion, W->d_fsub};
        put_args_kernel<NewbornArgs><<<dim3(1), dim3(256), 0, s>>>(static_cast<NewbornArgs*>(W->d_nargs), N);
        HIP_TRY(hipGetLastError());
        {
            hipEvent_t stop = nullptr;
            if (W->timing) {
                if (W->events_used == W->events.size()) {
                    hipEvent_t a, b;
                    HIP_TRY(hipEventCreate(&a));
                    HIP_TRY(hipEventCreate(&b));
                    W->events.push_back({a, b, 0});
                }
                rrea_fields_s::Timed& t = W->events[W->events_used++];
                t.kernel = 6;
                HIP_TRY(hipEventRecord(t.start, s));
                stop = t.stop;
            }
            newborn_kernel<<<dim3(static_cast<unsigned>(blocks_for(ge - gb, kBlock))), dim3(kBlock), 0, s>>>(
                static_cast<const NewbornArgs*>(W->d_nargs));
            HIP_TRY(hipGetLastError());
            if (stop) HIP_TRY(hipEventRecord(stop, s));
        }
        int64_t c = 0;
        HIP_TRY(hipMemcpyAsync(&c, W->d_count, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipMemcpyAsync(err_host, io->error, sizeof(err_host), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (err_host[0] != 0 || c <= 0) break;
        if (c > cap - ge) break;  // the kernel reported the capacity error
        {
            rrea_status const os = order_staged(W, c, io->work + ge, s);
            if (os != RREA_OK) return os;
        }
        unsigned const cb = static_cast<unsigned>(blocks_for(c, 256));
        fifo_anchor_kernel<<<dim3(cb), dim3(256), 0, s>>>(io->work, ge, ge + c, nullptr, W->d_anchor);
        HIP_TRY(hipGetLastError());
        fifo_lineage_kernel<<<dim3(cb), dim3(256), 0, s>>>(io->work, ge, ge + c, nullptr, W->d_anchor, W->d_next,
                                                           io->error);
        HIP_TRY(hipGetLastError());
        fifo_advance_kernel<<<dim3(cb), dim3(256), 0, s>>>(ge, ge + c, nullptr, W->d_anchor, W->d_next);
        HIP_TRY(hipGetLastError());
        gb = ge;
        ge += c;
    }
    int64_t const total = ge;
    set_i64_kernel<<<dim3(1), dim3(64), 0, s>>>(io->work_count, total);
    HIP_TRY(hipGetLastError());
    if (err_host[0] != 0) {
        set_i64_kernel<<<dim3(1), dim3(64), 0, s>>>(io->survivor_count, 0);
        HIP_TRY(hipGetLastError());
        return RREA_OK;
    }
    survivor_flag_kernel<<<dim3(static_cast<unsigned>(blocks_for(total, 256))), dim3(256), 0, s>>>(W->d_fst, total,
                                                                                                  W->d_flag);
    HIP_TRY(hipGetLastError());
    size_t bytes = W->sel_tmp_bytes;
    if (io->survivor_capacity >= total) {
        // every entry fits: select straight into the caller's buffer, no round trip
        HIP_TRY(hipcub::DeviceSelect::Flagged(W->d_sel_tmp, bytes, W->d_fin, W->d_flag, io->survivors,
                                              io->survivor_count, static_cast<int>(total), s));
    } else {
        int64_t nsurv = 0;
        HIP_TRY(hipcub::DeviceSelect::Flagged(W->d_sel_tmp, bytes, W->d_fin, W->d_flag, W->d_stage_sec, W->d_nsel,
                                              static_cast<int>(total), s));
        HIP_TRY(hipMemcpyAsync(&nsurv, W->d_nsel, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (nsurv > io->survivor_capacity) {
            report_kernel<<<dim3(1), dim3(64), 0, s>>>(io->error, kErrCapacity, -1);
            HIP_TRY(hipGetLastError());
            nsurv = io->survivor_capacity;
        }
        if (nsurv > 0) {
            HIP_TRY(hipMemcpyAsync(io->survivors, W->d_stage_sec, nsurv * sizeof(rrea_secondary),
                                   hipMemcpyDeviceToDevice, s));
        }
        set_i64_kernel<<<dim3(1), dim3(64), 0, s>>>(io->survivor_count, nsurv);
        HIP_TRY(hipGetLastError());
    }
    if (io->status) HIP_TRY(hipMemcpyAsync(io->status, W->d_fst, total, hipMemcpyDeviceToDevice, s));
    if (io->ion_pairs) {
        HIP_TRY(hipMemcpyAsync(io->ion_pairs, W->d_fion, total * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    if (io->substeps) {
        HIP_TRY(hipMemcpyAsync(io->substeps, W->d_fsub, total * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    }
    return RREA_OK;
}

// The tile ordering workspace for n particles (keys, permutation, sort
// temporary, one staging array).
rrea_status tile_sort_workspace(rrea_fields_t F, int64_t n)
{
    if (F->tsort_cap < n) {
        void* ptrs[] = {F->d_tkeys, F->d_tkeys_out, F->d_tidx, F->d_tidx_out, F->d_tsort_tmp, F->d_tbuf};
        for (void* p : ptrs) if (p) (void)hipFree(p);
        F->d_tkeys = nullptr; F->d_tkeys_out = nullptr; F->d_tidx = nullptr; F->d_tidx_out = nullptr;
        F->d_tsort_tmp = nullptr; F->d_tbuf = nullptr;
        F->tsort_cap = 0;
        HIP_TRY(hipMalloc(&F->d_tkeys, n * sizeof(uint32_t)));
        HIP_TRY(hipMalloc(&F->d_tkeys_out, n * sizeof(uint32_t)));
        HIP_TRY(hipMalloc(&F->d_tidx, n * sizeof(int32_t)));
        HIP_TRY(hipMalloc(&F->d_tidx_out, n * sizeof(int32_t)));
        HIP_TRY(hipMalloc(&F->d_tbuf, n * sizeof(double)));
        size_t bytes = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, F->d_tkeys, F->d_tkeys_out, F->d_tidx,
                                                   F->d_tidx_out, static_cast<int>(n)));
        HIP_TRY(hipMalloc(&F->d_tsort_tmp, bytes));
        F->tsort_tmp_bytes = bytes;
        F->tsort_cap = n;
    }
    return RREA_OK;
}

// Gather every distinct array of the tile (and the instance array, when
// given) through F->d_tidx_out, the permutation a sort left there.
// The commit kernel's view of three particle stores (electrons, positrons,
// photons), checked.
rrea_status commit_stores(const rrea_particle_store* stores, CommitStores& S)
{
    for (int q = 0; q < 3; ++q) {
        rrea_particle_store const& P = stores[q];
        rrea_tile const& t = P.tile;
        bool const photon = q == 2;
        if (P.capacity < 0) return RREA_ERR_ARGUMENT;
        if (P.capacity > 0 && (!t.prev_r || !t.prev_z || !t.prev_theta || !t.w || !t.id || !t.r || !t.z || !t.theta
                               || !t.ux || !t.uy || !t.uz || (!photon && (!t.prev_ux || !t.prev_uy || !t.prev_uz)))) {
            return RREA_ERR_ARGUMENT;
        }
        CommitStore& D = S.s[q];
        D.prev_r = const_cast<double*>(t.prev_r); D.prev_z = const_cast<double*>(t.prev_z);
        D.prev_th = const_cast<double*>(t.prev_theta);
        D.prev_ux = photon ? nullptr : const_cast<double*>(t.prev_ux);
        D.prev_uy = photon ? nullptr : const_cast<double*>(t.prev_uy);
        D.prev_uz = photon ? nullptr : const_cast<double*>(t.prev_uz);
        D.r = t.r; D.z = t.z; D.th = t.theta; D.ux = t.ux; D.uy = t.uy; D.uz = t.uz;
        D.w = const_cast<double*>(t.w); D.id = t.id; D.cpu = const_cast<int32_t*>(t.cpu);
        D.inst = P.instance;
        D.cap = P.capacity;
    }
    return RREA_OK;
}

rrea_status permute_tile(rrea_fields_t F, const rrea_tile* tile, int32_t photon, int32_t* instance, hipStream_t s)
{
    int64_t const n = tile->n;
    unsigned const nb = static_cast<unsigned>(blocks_for(n, 256));
    std::vector<double*> f64 = {const_cast<double*>(tile->prev_r), const_cast<double*>(tile->prev_z),
                                const_cast<double*>(tile->prev_theta), const_cast<double*>(tile->w)};
    if (!photon) {
        f64.push_back(const_cast<double*>(tile->prev_ux));
        f64.push_back(const_cast<double*>(tile->prev_uy));
        f64.push_back(const_cast<double*>(tile->prev_uz));
    }
    for (double* p : {tile->r, tile->z, tile->theta, tile->ux, tile->uy, tile->uz}) {
        if (std::find(f64.begin(), f64.end(), p) == f64.end()) f64.push_back(p);
    }
    double* const buf = static_cast<double*>(F->d_tbuf);
    for (double* p : f64) {
        gather_kernel<double><<<dim3(nb), dim3(256), 0, s>>>(n, F->d_tidx_out, p, buf);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(p, buf, n * sizeof(double), hipMemcpyDeviceToDevice, s));
    }
    {
        int64_t* const ib = static_cast<int64_t*>(F->d_tbuf);
        gather_kernel<int64_t><<<dim3(nb), dim3(256), 0, s>>>(n, F->d_tidx_out, tile->id, ib);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(tile->id, ib, n * sizeof(int64_t), hipMemcpyDeviceToDevice, s));
    }
    if (tile->cpu) {
        int32_t* const cb = static_cast<int32_t*>(F->d_tbuf);
        gather_kernel<int32_t><<<dim3(nb), dim3(256), 0, s>>>(n, F->d_tidx_out, tile->cpu, cb);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(const_cast<int32_t*>(tile->cpu), cb, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    }
    if (instance) {
        int32_t* const ib = static_cast<int32_t*>(F->d_tbuf);
        gather_kernel<int32_t><<<dim3(nb), dim3(256), 0, s>>>(n, F->d_tidx_out, instance, ib);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipMemcpyAsync(instance, ib, n * sizeof(int32_t), hipMemcpyDeviceToDevice, s));
    }
    return RREA_OK;
}

// ---------------------------------------------------------------------------
// The drift segment (forward model, checkpoints, reverse sweep)
// ---------------------------------------------------------------------------
constexpr int32_t kDriftMaxEvery = 16;

struct DriftArgs {
    DevTables T;
    DriftEnv E;
    int64_t n;
    int32_t steps, every, n_ck;
    double *r, *z, *th, *ux, *uy, *uz;
    const double* w;
    const int64_t* id;
    double* ck;              // [checkpoint][component][particle], components r, theta, z, ux, uy, uz
    uint8_t* ck_alive;       // [checkpoint][particle]
    uint8_t* alive;          // [particle], after the last step
    DriftSinks sinks;        // forward: J grid and/or loss weights; adjoint: loss weights
    double* loss;            // [particle]: the particle's sum over its deposits of J x dL/dJ
    double* g_er;            // adjoint: dL/d(cell-centred Er), nr*nz
    double* g_ez;
};

__global__ void __launch_bounds__(256) drift_forward_kernel(const DriftArgs* __restrict__ args)
{
    DriftArgs const& A = *args;
    int64_t const ip = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= A.n) return;
    DState<double> s{A.r[ip], A.th[ip], A.z[ip], A.ux[ip], A.uy[ip], A.uz[ip]};
    bool alive = A.id[ip] != RREA_INVALID_ID;
    double const w = A.w[ip];
    DriftSink<double> sink{A.E.G, A.sinks, 0.0};
    for (int32_t t = 0; t < A.steps; ++t) {
        if (A.ck && t % A.every == 0) {
            int64_t const c = t / A.every;
            double* const o = A.ck + c * 6 * A.n + ip;
            o[0] = s.r; o[A.n] = s.th; o[2 * A.n] = s.z; o[3 * A.n] = s.ux; o[4 * A.n] = s.uy; o[5 * A.n] = s.uz;
            A.ck_alive[c * A.n + ip] = alive ? 1 : 0;
        }
        if (alive) alive = drift_step(A.E, A.T, w, s, sink, nullptr);
    }
    A.r[ip] = s.r; A.th[ip] = s.th; A.z[ip] = s.z; A.ux[ip] = s.ux; A.uy[ip] = s.uy; A.uz[ip] = s.uz;
    if (A.alive) A.alive[ip] = alive ? 1 : 0;
    if (A.loss) A.loss[ip] = sink.acc;
}

// The reverse sweep, one particle per lane: per checkpoint segment (last
// first), the segment's states are recomputed from its checkpoint, then each
// step is evaluated once more in dual numbers over (state, gathered fields)
// and the adjoint of its input state is formed from the local Jacobian, the
// field adjoints scattered to the cells of the gather stencils.
__global__ void __launch_bounds__(64) drift_adjoint_kernel(const DriftArgs* __restrict__ args)
{
    DriftArgs const& A = *args;
    int64_t const ip = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (ip >= A.n) return;
    double const w = A.w[ip];
    DriftSinks const none{nullptr, nullptr, nullptr, nullptr};
    DriftSinks const weights{nullptr, nullptr, A.sinks.gjr, A.sinks.gjz};
    double adj[kStateSlots] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int32_t c = A.n_ck - 1; c >= 0; --c) {
        int32_t const t0 = c * A.every;
        int32_t const m = min(A.steps - t0, A.every);
        DState<double> st[kDriftMaxEvery];
        bool al[kDriftMaxEvery];
        {
            const double* const o = A.ck + static_cast<int64_t>(c) * 6 * A.n + ip;
            st[0] = DState<double>{o[0], o[A.n], o[2 * A.n], o[3 * A.n], o[4 * A.n], o[5 * A.n]};
            al[0] = A.ck_alive[static_cast<int64_t>(c) * A.n + ip] != 0;
        }
        for (int32_t q = 0; q + 1 < m; ++q) {
            st[q + 1] = st[q];
            DriftSink<double> sk{A.E.G, none, 0.0};
            al[q + 1] = al[q] && drift_step(A.E, A.T, w, st[q + 1], sk, nullptr);
        }
        for (int32_t q = m - 1; q >= 0; --q) {
            if (!al[q]) continue;
            DState<D10> x{D10::seed(st[q].r, 0), D10::seed(st[q].th, 1), D10::seed(st[q].z, 2),
                          D10::seed(st[q].ux, 3), D10::seed(st[q].uy, 4), D10::seed(st[q].uz, 5)};
            DriftSink<D10> sk{A.E.G, weights, D10(0.0)};
            GatherStencil gs[2];
            if (!drift_step(A.E, A.T, w, x, sk, gs)) continue;
            D10 const* out[kStateSlots] = {&x.r, &x.th, &x.z, &x.ux, &x.uy, &x.uz};
            double g[kDriftSlots];
            for (int j = 0; j < kDriftSlots; ++j) {
                double v = sk.acc.d[j];
                for (int i = 0; i < kStateSlots; ++i) v += adj[i] * out[i]->d[j];
                g[j] = v;
            }
            for (int i = 0; i < kStateSlots; ++i) adj[i] = g[i];
            for (int node = 0; node < 2; ++node) {
                double const ger = g[kStateSlots + 2 * node], gez = g[kStateSlots + 2 * node + 1];
                for (int k = 0; k < 4; ++k) {
                    int32_t const cell = gs[node].c[k];
                    if (ger != 0.0) atomic_add(&A.g_er[cell], ger * gs[node].w_er[k]);
                    if (gez != 0.0) atomic_add(&A.g_ez[cell], gez * gs[node].w_ez[k]);
                }
            }
        }
    }
}

// dL/d(face) from dL/d(cell): each cell-centred value is the average of its
// two faces (cell_fields_kernel), so a face gets half of each adjacent
// cell's gradient.  Accumulates into the caller's face arrays.
__global__ void face_gradient_kernel(int32_t nr, int32_t nz, const double* __restrict__ g_er,
                                     const double* __restrict__ g_ez, double* __restrict__ f_er,
                                     double* __restrict__ f_ez)
{
    int64_t const f = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const n_er = static_cast<int64_t>(nr + 1) * nz, n_ez = static_cast<int64_t>(nr) * (nz + 1);
    if (f < n_er) {
        int32_t const i = static_cast<int32_t>(f % (nr + 1)), j = static_cast<int32_t>(f / (nr + 1));
        double v = 0.0;
        if (i >= 1) v += 0.5 * g_er[(i - 1) + static_cast<int64_t>(nr) * j];
        if (i < nr) v += 0.5 * g_er[i + static_cast<int64_t>(nr) * j];
        f_er[f] += v;
    }
    if (f < n_ez) {
        int32_t const i = static_cast<int32_t>(f % nr), j = static_cast<int32_t>(f / nr);
        double v = 0.0;
        if (j >= 1) v += 0.5 * g_ez[i + static_cast<int64_t>(nr) * (j - 1)];
        if (j < nz) v += 0.5 * g_ez[i + static_cast<int64_t>(nr) * j];
        f_ez[f] += v;
    }
}

// RreaEventAccumulator::NormalizeAndSync :280-345: each source field over
// its RZ shell cell volume pi (r_o^2 - r_i^2) dz (rz_shell_volume,
// NA/include/rrea/RreaConstants.H:37-43), the three source rates also over dt.
__global__ void sources_normalize_kernel(int32_t nr, int32_t nz, double r_lo, double dr, double dz, double dt,
                                         int32_t diagnostics, double* __restrict__ src)
{
    int64_t const c = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int64_t const cells = static_cast<int64_t>(nr) * nz;
    if (c >= cells) return;
    int32_t const i = static_cast<int32_t>(c % nr);
    double const r_inner = r_lo + static_cast<double>(i - 0) * dr;
    double const inner = r_inner > 0.0 ? r_inner : 0.0;
    double const outer = r_inner + dr > 0.0 ? r_inner + dr : 0.0;
    double const volume = kPi * (outer * outer - inner * inner) * dz;
    for (int32_t f = 0; f < kSrcFields; ++f) {
        bool const rate = f <= kSrcNegativeIon;
        if (!rate && !diagnostics) continue;
        double const denom = rate ? volume * dt : volume;
        double& v = src[f * cells + c];
        v = (denom > 0.0) ? v / denom : 0.0;
    }
}

rrea_status drift_args(rrea_tables_t T, rrea_fields_t F, const rrea_drift_config* cfg, DriftArgs& A)
{
    if (!T || !F || !cfg || !F->fields_set) return RREA_ERR_ARGUMENT;
    if (cfg->species != 0 && cfg->species != 1) return RREA_ERR_ARGUMENT;
    if (cfg->steps < 0 || cfg->checkpoint_every < 1 || cfg->checkpoint_every > kDriftMaxEvery) {
        return RREA_ERR_ARGUMENT;
    }
    if (!(cfg->dt_s > 0.0) || !std::isfinite(cfg->dt_s) || !(cfg->low_energy_cutoff_eV > 0.0)) {
        return RREA_ERR_ARGUMENT;
    }
    Env env;
    rrea_status const st = make_env(F, nullptr, nullptr, nullptr, -1, 0, env);
    if (st != RREA_OK) return st;
    A = DriftArgs{};
    A.T = T->view;
    A.E.F = env.F;
    A.E.D = env.D;
    A.E.G = env.G;
    A.E.fallback_ratio = env.density_ratio;
    bool const pos = cfg->species == 1;
    A.E.q = pos ? 1.0 : -1.0;
    A.E.col = pos ? kPCol : kECol;
    A.E.soft = pos ? kPSoft : kESoft;
    A.E.dt = cfg->dt_s;
    A.E.cutoff = cfg->low_energy_cutoff_eV;
    A.steps = cfg->steps;
    A.every = cfg->checkpoint_every;
    A.n_ck = (cfg->steps + cfg->checkpoint_every - 1) / cfg->checkpoint_every;
    return RREA_OK;
}

}  // namespace

extern "C" {

rrea_status rrea_batch_create(rrea_fields_t const* fields, int32_t n, double* const* jr, double* const* jz,
                              double* const* rho, const uint64_t* rng_seeds, rrea_batch_t* out)
{
    if (!fields || n <= 0 || !jr || !jz || !rho || !rng_seeds || !out) return RREA_ERR_ARGUMENT;
    for (int32_t i = 0; i < n; ++i) {
        if (!fields[i] || !jr[i] || !jz[i] || !rho[i]) return RREA_ERR_ARGUMENT;
    }
    auto* B = new (std::nothrow) rrea_batch_s();
    if (!B) return RREA_ERR_HIP;
    B->fields.assign(fields, fields + n);
    B->jr.assign(jr, jr + n);
    B->jz.assign(jz, jz + n);
    B->rho.assign(rho, rho + n);
    B->seeds.assign(rng_seeds, rng_seeds + n);
    rrea_status st = rrea_fields_create(&fields[0]->geom, &B->work);
    if (st == RREA_OK && hipMalloc(&B->d_envs, static_cast<size_t>(n) * sizeof(Env)) != hipSuccess) st = RREA_ERR_HIP;
    if (st != RREA_OK) {
        rrea_batch_destroy(B);
        return st;
    }
    B->h_envs.resize(static_cast<size_t>(n) * sizeof(Env));
    *out = B;
    return RREA_OK;
}

rrea_status rrea_batch_destroy(rrea_batch_t B)
{
    if (!B) return RREA_OK;
    if (B->d_envs) (void)hipFree(B->d_envs);
    if (B->work) rrea_fields_destroy(B->work);
    delete B;
    return RREA_OK;
}

rrea_status rrea_batch_workspace(rrea_batch_t B, rrea_fields_t* out)
{
    if (!B || !out) return RREA_ERR_ARGUMENT;
    *out = B->work;
    return RREA_OK;
}

rrea_status rrea_transport_step_batched(rrea_tables_t T, rrea_batch_t B, int32_t species, const rrea_tile* tile,
                                        const int32_t* instance, const rrea_step_config* cfg, const rrea_step_out* out,
                                        rrea_stream_t stream)
{
    if (!T || !B || !cfg || species < kElectron || species > kPhoton) return RREA_ERR_ARGUMENT;
    rrea_status const es = upload_envs(B, cfg, static_cast<hipStream_t>(stream));
    if (es != RREA_OK) return es;
    BatchCtx const ctx{B->work, static_cast<const Env*>(B->d_envs), instance};
    return species_step(species, T, B->fields[0], tile, cfg, out, stream, &ctx);
}

rrea_status rrea_newborn_drain_batched(rrea_tables_t T, rrea_batch_t B, const rrea_step_config* cfg,
                                       const rrea_drain_io* io, rrea_stream_t stream)
{
    if (!T || !B || !cfg) return RREA_ERR_ARGUMENT;
    rrea_status const es = upload_envs(B, cfg, static_cast<hipStream_t>(stream));
    if (es != RREA_OK) return es;
    BatchCtx const ctx{B->work, static_cast<const Env*>(B->d_envs), nullptr};
    return drain_impl(T, B->fields[0], cfg, io, stream, &ctx);
}

rrea_status rrea_transport_step_electrons(rrea_tables_t T, rrea_fields_t F, const rrea_tile* tile,
                                          const rrea_step_config* cfg, const rrea_step_out* out, rrea_stream_t stream)
{
    return species_step(kElectron, T, F, tile, cfg, out, stream);
}

rrea_status rrea_transport_step_positrons(rrea_tables_t T, rrea_fields_t F, const rrea_tile* tile,
                                          const rrea_step_config* cfg, const rrea_step_out* out, rrea_stream_t stream)
{
    return species_step(kPositron, T, F, tile, cfg, out, stream);
}

rrea_status rrea_transport_step_photons(rrea_tables_t T, rrea_fields_t F, const rrea_tile* tile,
                                        const rrea_step_config* cfg, const rrea_step_out* out, rrea_stream_t stream)
{
    return species_step(kPhoton, T, F, tile, cfg, out, stream);
}

rrea_status rrea_validate_secondary_groups(const rrea_secondary_group* groups, int64_t n, double cutoff,
                                           int32_t* result, double* residuals, rrea_stream_t stream)
{
    if (n < 0 || (n > 0 && (!groups || !result))) return RREA_ERR_ARGUMENT;
    if (n == 0) return RREA_OK;
    validate_groups_kernel<<<dim3(static_cast<unsigned>(blocks_for(n, 256))), dim3(256), 0,
                             static_cast<hipStream_t>(stream)>>>(groups, n, cutoff, result, residuals);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

rrea_status rrea_transport_step_noalloc(int32_t species, rrea_tables_t T, rrea_fields_t F, const rrea_tile* tile,
                                        const rrea_step_config* cfg, const rrea_step_out* out, rrea_stream_t stream)
{
    return noalloc_step(species, T, F, tile, cfg, out, stream);
}

rrea_status rrea_newborn_drain(rrea_tables_t T, rrea_fields_t F, const rrea_step_config* cfg, const rrea_drain_io* io,
                               rrea_stream_t stream)
{
    return drain_impl(T, F, cfg, io, stream, nullptr);
}

rrea_status rrea_newborn_drain_staged(rrea_tables_t T, rrea_fields_t F, const rrea_step_config* cfg,
                                      const rrea_secondary* const* staged, const int64_t* const* staged_counts,
                                      const int64_t* staged_caps, int32_t n_staged, const rrea_drain_io* io,
                                      int32_t max_generations, rrea_stream_t stream)
{
    if (!T || !F || !cfg || !io || !F->fields_set || !staged || !staged_counts || !staged_caps || n_staged < 1
        || n_staged > 3 || max_generations < 1 || max_generations > 64) {
        return RREA_ERR_ARGUMENT;
    }
    for (int32_t k = 0; k < n_staged; ++k) {
        if (!staged[k] || !staged_counts[k] || staged_caps[k] < 0) return RREA_ERR_ARGUMENT;
    }
    if (!io->work || !io->work_count || !io->survivors || !io->survivor_count || !io->tallies || !io->error
        || !io->jr || !io->jz || !io->rho) {
        return RREA_ERR_ARGUMENT;
    }
    int64_t const cap = io->work_capacity;
    if (cap <= 0 || cap > INT32_MAX || io->survivor_capacity < 0) return RREA_ERR_ARGUMENT;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != T->device) return RREA_ERR_ARGUMENT;
    // RreaParticleInteraction::Apply :38-40 -- a non-positive dt does nothing.
    if (!(cfg->dt_s > 0.0)) {
        set_i64_kernel<<<dim3(1), dim3(64), 0, s>>>(io->work_count, 0);
        set_i64_kernel<<<dim3(1), dim3(64), 0, s>>>(io->survivor_count, 0);
        HIP_TRY(hipGetLastError());
        return RREA_OK;
    }
    if (!(cfg->max_optical_depth_per_substep > 0.0) || cfg->max_subcycles == 0 || cfg->max_subcycles > 16383) {
        return RREA_ERR_ARGUMENT;
    }
    rrea_fields_t const W = F;
    {
        rrea_status const ws = fifo_workspace(W, cap);
        if (ws != RREA_OK) return ws;
    }
    if (!W->d_dstate) HIP_TRY(hipMalloc(&W->d_dstate, sizeof(DrainState)));
    StepArgs A{};
    {
        rrea_status const cs = common_args(T, F, cfg, io->jr, io->jz, io->rho, A, io->sources);
        if (cs != RREA_OK) return cs;
    }
    A.tallies = io->tallies;
    A.error = io->error;
    A.sec = W->d_stage_sec;
    A.sec_cap = 0;  // written on the device per generation
    A.sec_count = W->d_count;
    NewbornArgs N{A, io->work, 0, 0, W->d_fin, W->d_fst, W->d_fion, W->d_fsub};
    put_args_kernel<NewbornArgs><<<dim3(1), dim3(256), 0, s>>>(static_cast<NewbornArgs*>(W->d_nargs), N);
    HIP_TRY(hipGetLastError());
    StagedLists S{};
    S.n = n_staged;
    int64_t most = 0;
    for (int32_t k = 0; k < n_staged; ++k) {
        S.rec[k] = staged[k];
        S.count[k] = staged_counts[k];
        S.cap[k] = staged_caps[k];
        most += staged_caps[k];
    }
    unsigned const cb = static_cast<unsigned>(std::max<int64_t>(1, std::min<int64_t>(blocks_for(std::min(most, cap), 256),
                                                                                      1024)));
    // the FIFO's initial length goes through work_count, which the drain's end overwrites
    concat_staged_kernel<<<dim3(cb), dim3(256), 0, s>>>(S, io->work, cap, io->work_count, io->error);
    HIP_TRY(hipGetLastError());
    return drain_device_launch(W, static_cast<NewbornArgs*>(W->d_nargs), static_cast<DrainState*>(W->d_dstate),
                               io->work, cap, io->work_count, io->work_count, io->survivors, io->survivor_capacity,
                               io->survivor_count, io->error, max_generations, s);
}


rrea_status rrea_tile_sort(rrea_fields_t F, const rrea_tile* tile, int32_t photon, rrea_stream_t stream)
{
    return rrea_tile_sort_banded(F, tile, photon, -1, stream);
}

rrea_status rrea_tile_sort_banded(rrea_fields_t F, const rrea_tile* tile, int32_t photon, int32_t energy_bits,
                                  rrea_stream_t stream)
{
    if (!F || !tile || tile->n < 0 || energy_bits < -1 || energy_bits > 6) return RREA_ERR_ARGUMENT;
    int64_t const n = tile->n;
    if (n == 0) return RREA_OK;
    if (n > INT32_MAX || !tile->prev_r || !tile->prev_z || !tile->prev_theta || !tile->w || !tile->id
        || !tile->r || !tile->z || !tile->theta || !tile->ux || !tile->uy || !tile->uz
        || (!photon && (!tile->prev_ux || !tile->prev_uy || !tile->prev_uz))) {
        return RREA_ERR_ARGUMENT;
    }
    rrea_rz_geometry const& g = F->geom;
    uint64_t const cells = static_cast<uint64_t>(g.nr) * static_cast<uint64_t>(g.nz);
    int32_t cbits = 0;
    while ((uint64_t{1} << cbits) < cells) ++cbits;
    if (cbits > 31) return RREA_ERR_ARGUMENT;
    int32_t const emax = 32 - cbits < 6 ? 32 - cbits : 6;
    int32_t const ebits = energy_bits < 0 || energy_bits > emax ? emax : energy_bits;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    {
        rrea_status const ws = tile_sort_workspace(F, n);
        if (ws != RREA_OK) return ws;
    }
    // charged tiles key on the captured state; photons on their current point
    const double* ku_x = photon ? tile->ux : tile->prev_ux;
    const double* ku_y = photon ? tile->uy : tile->prev_uy;
    const double* ku_z = photon ? tile->uz : tile->prev_uz;
    unsigned const nb = static_cast<unsigned>(blocks_for(n, 256));
    tile_key_kernel<<<dim3(nb), dim3(256), 0, s>>>(n, tile->prev_r, tile->prev_z, ku_x, ku_y, ku_z, tile->id,
                                                   photon ? 1 : 0, g.nr, g.nz, g.r_lo, g.z_lo, g.dr, g.dz, ebits,
                                                   F->d_tkeys, F->d_tidx);
    HIP_TRY(hipGetLastError());
    size_t bytes = F->tsort_tmp_bytes;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(F->d_tsort_tmp, bytes, F->d_tkeys, F->d_tkeys_out, F->d_tidx,
                                               F->d_tidx_out, static_cast<int>(n), 0, cbits + ebits, s));
    return permute_tile(F, tile, photon, nullptr, s);
}

rrea_status rrea_batch_tile_sort(rrea_batch_t B, const rrea_tile* tile, int32_t photon, int32_t* instance,
                                 int32_t energy_bits, rrea_stream_t stream)
{
    if (!B || !tile || tile->n < 0 || energy_bits < -1 || energy_bits > 6) return RREA_ERR_ARGUMENT;
    int64_t const n = tile->n;
    if (n == 0) return RREA_OK;
    if (n > INT32_MAX || !instance || !tile->prev_r || !tile->prev_z || !tile->prev_theta || !tile->w || !tile->id
        || !tile->r || !tile->z || !tile->theta || !tile->ux || !tile->uy || !tile->uz
        || (!photon && (!tile->prev_ux || !tile->prev_uy || !tile->prev_uz))) {
        return RREA_ERR_ARGUMENT;
    }
    int32_t const n_inst = static_cast<int32_t>(B->fields.size());
    int32_t ibits = 0;
    while ((int64_t{1} << ibits) < n_inst) ++ibits;
    int32_t cbits = 0;
    for (rrea_fields_t f : B->fields) {
        uint64_t const cells = static_cast<uint64_t>(f->geom.nr) * static_cast<uint64_t>(f->geom.nz);
        while ((uint64_t{1} << cbits) < cells) ++cbits;
    }
    // the invalid key (all ones) must stay above every live key
    if (ibits + cbits > 31) return RREA_ERR_ARGUMENT;
    int32_t const emax = 31 - ibits - cbits < 6 ? 31 - ibits - cbits : 6;
    int32_t const ebits = energy_bits < 0 || energy_bits > emax ? emax : energy_bits;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    rrea_status const es = upload_envs(B, nullptr, s);
    if (es != RREA_OK) return es;
    rrea_fields_t const W = B->work;
    {
        rrea_status const ws = tile_sort_workspace(W, n);
        if (ws != RREA_OK) return ws;
    }
    // charged tiles key on the captured state; photons on their current point
    const double* ku_x = photon ? tile->ux : tile->prev_ux;
    const double* ku_y = photon ? tile->uy : tile->prev_uy;
    const double* ku_z = photon ? tile->uz : tile->prev_uz;
    unsigned const nb = static_cast<unsigned>(blocks_for(n, 256));
    batch_tile_key_kernel<<<dim3(nb), dim3(256), 0, s>>>(n, tile->prev_r, tile->prev_z, ku_x, ku_y, ku_z, tile->id,
                                                         instance, photon ? 1 : 0, static_cast<const Env*>(B->d_envs),
                                                         n_inst, cbits, ebits, W->d_tkeys, W->d_tidx);
    HIP_TRY(hipGetLastError());
    size_t bytes = W->tsort_tmp_bytes;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(W->d_tsort_tmp, bytes, W->d_tkeys, W->d_tkeys_out, W->d_tidx,
                                               W->d_tidx_out, static_cast<int>(n), 0, 32, s));
    return permute_tile(W, tile, photon, instance, s);
}

rrea_status rrea_tile_compact(rrea_fields_t F, const rrea_tile* tile, int32_t photon, int32_t* instance,
                              int64_t* n_live, rrea_stream_t stream)
{
    if (!F || !tile || tile->n < 0 || !n_live) return RREA_ERR_ARGUMENT;
    int64_t const n = tile->n;
    *n_live = 0;
    if (n == 0) return RREA_OK;
    if (n > INT32_MAX || !tile->prev_r || !tile->prev_z || !tile->prev_theta || !tile->w || !tile->id
        || !tile->r || !tile->z || !tile->theta || !tile->ux || !tile->uy || !tile->uz
        || (!photon && (!tile->prev_ux || !tile->prev_uy || !tile->prev_uz))) {
        return RREA_ERR_ARGUMENT;
    }
    hipStream_t const s = static_cast<hipStream_t>(stream);
    {
        rrea_status const ws = tile_sort_workspace(F, n);
        if (ws != RREA_OK) return ws;
    }
    if (!F->d_live) HIP_TRY(hipMalloc(&F->d_live, sizeof(int64_t)));
    fill_u64_kernel<<<dim3(1), dim3(64), 0, s>>>(reinterpret_cast<unsigned long long*>(F->d_live), 1, 0ULL);
    HIP_TRY(hipGetLastError());
    unsigned const nb = static_cast<unsigned>(blocks_for(n, 256));
    live_key_kernel<<<dim3(nb), dim3(256), 0, s>>>(n, tile->id, F->d_tkeys, F->d_tidx,
                                                   reinterpret_cast<unsigned long long*>(F->d_live));
    HIP_TRY(hipGetLastError());
    size_t bytes = 0;
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, F->d_tkeys, F->d_tkeys_out, F->d_tidx,
                                               F->d_tidx_out, static_cast<int>(n), 0, 1));
    if (bytes > F->tsort_tmp_bytes) {
        if (F->d_tsort_tmp) (void)hipFree(F->d_tsort_tmp);
        F->d_tsort_tmp = nullptr;
        F->tsort_tmp_bytes = 0;
        HIP_TRY(hipMalloc(&F->d_tsort_tmp, bytes));
        F->tsort_tmp_bytes = bytes;
    }
    bytes = F->tsort_tmp_bytes;
    // one key bit: a stable partition, live particles keep their order
    HIP_TRY(hipcub::DeviceRadixSort::SortPairs(F->d_tsort_tmp, bytes, F->d_tkeys, F->d_tkeys_out, F->d_tidx,
                                               F->d_tidx_out, static_cast<int>(n), 0, 1, s));
    rrea_status const ps = permute_tile(F, tile, photon, instance, s);
    if (ps != RREA_OK) return ps;
    HIP_TRY(hipMemcpyAsync(n_live, F->d_live, sizeof(int64_t), hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    return RREA_OK;
}

rrea_status rrea_commit_survivors(const rrea_secondary* survivors, const int64_t* survivor_count,
                                  const rrea_particle_store* stores, int64_t* counts, int64_t* next_id, int32_t cpu,
                                  int32_t* error, rrea_stream_t stream)
{
    if (!survivors || !survivor_count || !stores || !counts || !next_id || !error) return RREA_ERR_ARGUMENT;
    CommitStores S{};
    rrea_status const cs = commit_stores(stores, S);
    if (cs != RREA_OK) return cs;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    commit_kernel<<<dim3(1), dim3(1024), 0, s>>>(survivors, survivor_count, S, counts, next_id, cpu, error);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

rrea_status rrea_photon_push(const rrea_tile* tile, double dt_s, rrea_stream_t stream)
{
    if (!tile || tile->n < 0 || !(dt_s >= 0.0)) return RREA_ERR_ARGUMENT;
    if (tile->n == 0) return RREA_OK;
    if (!tile->prev_r || !tile->prev_z || !tile->prev_theta || !tile->r || !tile->z || !tile->theta || !tile->ux
        || !tile->uy || !tile->uz || !tile->id) {
        return RREA_ERR_ARGUMENT;
    }
    hipStream_t const s = static_cast<hipStream_t>(stream);
    photon_push_kernel<<<dim3(static_cast<unsigned>(blocks_for(tile->n, 256))), dim3(256), 0, s>>>(
        tile->n, tile->id, const_cast<double*>(tile->prev_r), const_cast<double*>(tile->prev_z),
        const_cast<double*>(tile->prev_theta), tile->r, tile->z, tile->theta, tile->ux, tile->uy, tile->uz,
        kC * dt_s);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}


rrea_status rrea_drift_forward(rrea_tables_t T, rrea_fields_t F, const rrea_tile* tile, const rrea_drift_config* cfg,
                               double* checkpoints, uint8_t* checkpoint_alive, uint8_t* alive, double* jr,
                               double* jz, const double* gjr, const double* gjz, double* loss, rrea_stream_t stream)
{
    DriftArgs A;
    {
        rrea_status const st = drift_args(T, F, cfg, A);
        if (st != RREA_OK) return st;
    }
    if (!tile || tile->n < 0 || (!checkpoints) != (!checkpoint_alive) || (!gjr) != (!gjz) || (loss && !gjr)
        || (!jr) != (!jz)) {
        return RREA_ERR_ARGUMENT;
    }
    if (tile->n == 0) return RREA_OK;
    if (!tile->r || !tile->z || !tile->theta || !tile->ux || !tile->uy || !tile->uz || !tile->w || !tile->id) {
        return RREA_ERR_ARGUMENT;
    }
    hipStream_t const s = static_cast<hipStream_t>(stream);
    A.n = tile->n;
    A.r = tile->r; A.z = tile->z; A.th = tile->theta; A.ux = tile->ux; A.uy = tile->uy; A.uz = tile->uz;
    A.w = tile->w; A.id = tile->id;
    A.ck = checkpoints; A.ck_alive = checkpoint_alive; A.alive = alive;
    A.sinks = DriftSinks{jr, jz, gjr, gjz};
    A.loss = loss;
    if (!F->d_dargs) HIP_TRY(hipMalloc(&F->d_dargs, sizeof(DriftArgs)));
    HIP_TRY(hipMemcpyAsync(F->d_dargs, &A, sizeof(DriftArgs), hipMemcpyHostToDevice, s));
    drift_forward_kernel<<<dim3(static_cast<unsigned>(blocks_for(A.n, 256))), dim3(256), 0, s>>>(
        static_cast<const DriftArgs*>(F->d_dargs));
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

rrea_status rrea_drift_adjoint(rrea_tables_t T, rrea_fields_t F, const rrea_tile* tile, const rrea_drift_config* cfg,
                               const double* checkpoints, const uint8_t* checkpoint_alive, const double* gjr,
                               const double* gjz, double* grad_er, double* grad_ez, rrea_stream_t stream)
{
    DriftArgs A;
    {
        rrea_status const st = drift_args(T, F, cfg, A);
        if (st != RREA_OK) return st;
    }
    if (!tile || tile->n < 0 || !gjr || !gjz || !grad_er || !grad_ez) return RREA_ERR_ARGUMENT;
    if (tile->n == 0 || cfg->steps == 0) return RREA_OK;
    if (!checkpoints || !checkpoint_alive || !tile->w) return RREA_ERR_ARGUMENT;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    int32_t const nr = F->geom.nr, nz = F->geom.nz;
    int64_t const cells = static_cast<int64_t>(nr) * nz;
    if (!F->d_gcell) HIP_TRY(hipMalloc(&F->d_gcell, 2 * cells * sizeof(double)));
    fill_u64_kernel<<<dim3(static_cast<unsigned>(blocks_for(2 * cells, 256))), dim3(256), 0, s>>>(
        reinterpret_cast<unsigned long long*>(F->d_gcell), 2 * cells, 0ULL);
    HIP_TRY(hipGetLastError());
    A.n = tile->n;
    A.w = tile->w;
    A.ck = const_cast<double*>(checkpoints);
    A.ck_alive = const_cast<uint8_t*>(checkpoint_alive);
    A.sinks = DriftSinks{nullptr, nullptr, gjr, gjz};
    A.g_er = F->d_gcell;
    A.g_ez = F->d_gcell + cells;
    if (!F->d_dargs) HIP_TRY(hipMalloc(&F->d_dargs, sizeof(DriftArgs)));
    HIP_TRY(hipMemcpyAsync(F->d_dargs, &A, sizeof(DriftArgs), hipMemcpyHostToDevice, s));
    drift_adjoint_kernel<<<dim3(static_cast<unsigned>(blocks_for(A.n, 64))), dim3(64), 0, s>>>(
        static_cast<const DriftArgs*>(F->d_dargs));
    HIP_TRY(hipGetLastError());
    int64_t const faces = static_cast<int64_t>(nr + 1) * (nz + 1);
    face_gradient_kernel<<<dim3(static_cast<unsigned>(blocks_for(faces, 256))), dim3(256), 0, s>>>(
        nr, nz, A.g_er, A.g_ez, grad_er, grad_ez);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}


rrea_status rrea_batch_bind_sources(rrea_batch_t B, double* const* sources)
{
    if (!B) return RREA_ERR_ARGUMENT;
    if (!sources) { B->src.clear(); return RREA_OK; }
    for (size_t i = 0; i < B->fields.size(); ++i) if (!sources[i]) return RREA_ERR_ARGUMENT;
    B->src.assign(sources, sources + B->fields.size());
    return RREA_OK;
}

rrea_status rrea_sources_normalize(rrea_fields_t F, double* sources, double dt_s, int32_t diagnostics,
                                   rrea_stream_t stream)
{
    if (!F || !sources || !(dt_s > 0.0) || !std::isfinite(dt_s)) return RREA_ERR_ARGUMENT;
    rrea_rz_geometry const& g = F->geom;
    int64_t const cells = static_cast<int64_t>(g.nr) * g.nz;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    sources_normalize_kernel<<<dim3(static_cast<unsigned>(blocks_for(cells, 256))), dim3(256), 0, s>>>(
        g.nr, g.nz, g.r_lo, g.dr, g.dz, dt_s, diagnostics, sources);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}


// ---------------------------------------------------------------------------
// Graph-captured batched steps
// ---------------------------------------------------------------------------
struct rrea_batch_graph_s {
    rrea_batch_t batch = nullptr;
    rrea_tile tiles[3] = {};
    bool present[3] = {false, false, false};
    void* d_sargs[3] = {nullptr, nullptr, nullptr};  // StepArgs per species
    void* d_nargs = nullptr;                         // NewbornArgs of the drain
    void* d_state = nullptr;                         // DrainState
    rrea_fields_t work = nullptr;                    // the graph's own workspace: its buffers are baked in
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int device = 0;
    // the avalanche: each step ends with the commit of its survivors into the tiles
    bool grow = false;
    CommitStores commit{};
    int64_t* counts = nullptr;
    int64_t* next_id = nullptr;
    int32_t cpu = 0;
    int64_t cap[3] = {0, 0, 0};
};

rrea_status rrea_batch_graph_destroy(rrea_batch_graph_t G)
{
    if (!G) return RREA_OK;
    if (G->exec) (void)hipGraphExecDestroy(G->exec);
    if (G->graph) (void)hipGraphDestroy(G->graph);
    for (void* p : G->d_sargs) if (p) (void)hipFree(p);
    if (G->d_nargs) (void)hipFree(G->d_nargs);
    if (G->d_state) (void)hipFree(G->d_state);
    if (G->work) rrea_fields_destroy(G->work);
    delete G;
    return RREA_OK;
}

rrea_status rrea_batch_graph_create(rrea_tables_t T, rrea_batch_t B, const rrea_batch_graph_desc* d,
                                    rrea_batch_graph_t* out)
{
    if (!T || !B || !d || !out) return RREA_ERR_ARGUMENT;
    *out = nullptr;
    if (!d->records || !d->record_count || d->record_capacity <= 0 || d->record_capacity > INT32_MAX
        || !d->survivors || !d->survivor_count || d->survivor_capacity < 0 || !d->tallies || !d->error
        || d->max_generations < 1 || d->max_generations > 64 || d->steps < 1 || d->steps > 1024) {
        return RREA_ERR_ARGUMENT;
    }
    for (int k = 0; k < 4; ++k) {
        if (d->clear_n[k] < 0 || (d->clear_n[k] > 0 && !d->clear[k])) return RREA_ERR_ARGUMENT;
    }
    bool any = false;
    for (int k = 0; k < 3; ++k) {
        if (d->tiles[k] && d->tiles[k]->n > 0) {
            if (!d->instance[k]) return RREA_ERR_ARGUMENT;
            any = true;
        }
    }
    if (!any || !(d->config.dt_s > 0.0)) return RREA_ERR_ARGUMENT;
    if (d->counts && !d->next_id) return RREA_ERR_ARGUMENT;
    if (d->push_photons && !(d->tiles[kPhoton] && d->tiles[kPhoton]->n > 0)) return RREA_ERR_ARGUMENT;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != T->device) return RREA_ERR_ARGUMENT;

    auto* G = new (std::nothrow) rrea_batch_graph_s();
    if (!G) return RREA_ERR_HIP;
    G->batch = B;
    G->device = dev;
    if (d->counts) {
        // the tiles are stores: n is their capacity, counts[k] their length
        G->grow = true;
        G->counts = d->counts;
        G->next_id = d->next_id;
        G->cpu = d->cpu;
        rrea_particle_store stores[3] = {};
        for (int k = 0; k < 3; ++k) {
            if (!(d->tiles[k] && d->tiles[k]->n > 0)) continue;
            stores[k].tile = *d->tiles[k];
            stores[k].instance = const_cast<int32_t*>(d->instance[k]);
            stores[k].capacity = d->tiles[k]->n;
            G->cap[k] = d->tiles[k]->n;
        }
        rrea_status const cs = commit_stores(stores, G->commit);
        if (cs != RREA_OK) {
            delete G;
            return cs;
        }
    }
    rrea_status st = rrea_fields_create(&B->fields[0]->geom, &G->work);
    if (st != RREA_OK) {
        rrea_batch_graph_destroy(G);
        return st;
    }
    // a workspace no other call grows or frees: the captured launches hold its pointers
    rrea_fields_t const W = G->work;
    // the fast kernels' workgroup size the batch's workspace is set to
    W->block = B->work->block;
    hipStream_t cs = nullptr;
    auto fail = [&](rrea_status e) {
        if (cs) (void)hipStreamDestroy(cs);
        rrea_batch_graph_destroy(G);
        return e;
    };
    if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return fail(RREA_ERR_HIP);
    st = upload_envs(B, &d->config, cs);
    if (st != RREA_OK) return fail(st);

    // every species' argument block and launch plan, formed on the host once
    rrea_step_out so{};
    so.secondaries = d->records;
    so.secondary_capacity = d->record_capacity;
    so.secondary_count = d->record_count;
    so.tallies = d->tallies;
    so.error = d->error;
    so.order_secondaries = 1;
    SpeciesLaunch L[3];
    for (int k = 0; k < 3; ++k) {
        if (!(d->tiles[k] && d->tiles[k]->n > 0)) continue;
        G->tiles[k] = *d->tiles[k];
        G->present[k] = true;
        BatchCtx const ctx{W, static_cast<const Env*>(B->d_envs), d->instance[k]};
        StepArgs A;
        st = species_prepare(k, T, B->fields[0], &G->tiles[k], &d->config, &so, &ctx, A, L[k]);
        if (st != RREA_OK) return fail(st);
        if (hipMalloc(&G->d_sargs[k], sizeof(StepArgs)) != hipSuccess) return fail(RREA_ERR_HIP);
        if (hipMemcpy(G->d_sargs[k], &A, sizeof(StepArgs), hipMemcpyHostToDevice) != hipSuccess) {
            return fail(RREA_ERR_HIP);
        }
    }
    // the drain's argument block: the step's records are the FIFO
    {
        st = fifo_workspace(W, d->record_capacity);
        if (st != RREA_OK) return fail(st);
        StepArgs A{};
        st = common_args(T, B->fields[0], &d->config, nullptr, nullptr, nullptr, A, nullptr);
        if (st != RREA_OK) return fail(st);
        A.tallies = d->tallies;
        A.error = d->error;
        A.batch = static_cast<const Env*>(B->d_envs);
        A.sec = W->d_stage_sec;
        A.sec_cap = 0;  // written on the device per generation
        A.sec_count = W->d_count;
        NewbornArgs N{A, d->records, 0, 0, W->d_fin, W->d_fst, W->d_fion, W->d_fsub};
        if (hipMalloc(&G->d_nargs, sizeof(NewbornArgs)) != hipSuccess
            || hipMalloc(&G->d_state, sizeof(DrainState)) != hipSuccess
            || hipMemcpy(G->d_nargs, &N, sizeof(NewbornArgs), hipMemcpyHostToDevice) != hipSuccess) {
            return fail(RREA_ERR_HIP);
        }
    }
    if (hipStreamSynchronize(cs) != hipSuccess) return fail(RREA_ERR_HIP);

    // capture `steps` steps: clear, (photon push), Apply's species order, the
    // drain, the next step number
    StepArgs* const sa[3] = {static_cast<StepArgs*>(G->d_sargs[0]), static_cast<StepArgs*>(G->d_sargs[1]),
                             static_cast<StepArgs*>(G->d_sargs[2])};
    NewbornArgs* const na = static_cast<NewbornArgs*>(G->d_nargs);
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) return fail(RREA_ERR_HIP);
    rrea_status cst = RREA_OK;
    for (int32_t step = 0; step < d->steps && cst == RREA_OK; ++step) {
        set_i64_kernel<<<dim3(1), dim3(64), 0, cs>>>(d->record_count, 0);
        if (G->grow) {
            set_lengths_kernel<<<dim3(1), dim3(64), 0, cs>>>(sa[0], sa[1], sa[2], G->counts, G->cap[0], G->cap[1],
                                                             G->cap[2]);
        }
        for (int k = 0; k < 4; ++k) {
            if (d->clear_n[k] == 0) continue;
            unsigned const b = static_cast<unsigned>(std::min<int64_t>(blocks_for(d->clear_n[k], 256), 2048));
            fill_f64_kernel<<<dim3(b), dim3(256), 0, cs>>>(d->clear[k], d->clear_n[k]);
        }
        if (hipGetLastError() != hipSuccess) { cst = RREA_ERR_HIP; break; }
        if (d->push_photons && G->grow) {
            rrea_tile const& t = G->tiles[kPhoton];
            photon_push_kernel<<<dim3(static_cast<unsigned>(blocks_for(t.n, 256))), dim3(256), 0, cs>>>(
                t.n, t.id, const_cast<double*>(t.prev_r), const_cast<double*>(t.prev_z),
                const_cast<double*>(t.prev_theta), t.r, t.z, t.theta, t.ux, t.uy, t.uz, kC * d->config.dt_s,
                G->counts + kPhoton);
        } else if (d->push_photons) {
            cst = rrea_photon_push(&G->tiles[kPhoton], d->config.dt_s, cs);
            if (cst != RREA_OK) break;
        }
        for (int k : {static_cast<int>(kElectron), static_cast<int>(kPhoton), static_cast<int>(kPositron)}) {
            if (!G->present[k]) continue;
            cst = species_launch(W, sa[k], L[k], &so, cs, true);
            if (cst != RREA_OK) break;
        }
        if (cst != RREA_OK) break;
        cst = drain_device_launch(W, na, static_cast<DrainState*>(G->d_state), d->records, d->record_capacity,
                                  d->record_count, d->record_count, d->survivors, d->survivor_capacity,
                                  d->survivor_count, d->error, d->max_generations, cs);
        if (cst != RREA_OK) break;
        if (G->grow) {
            // AddNParticles of the drain's survivors (RreaSecondaryCommit.cpp:594-629)
            commit_kernel<<<dim3(1), dim3(1024), 0, cs>>>(d->survivors, d->survivor_count, G->commit, G->counts,
                                                          G->next_id, G->cpu, d->error);
        }
        next_step_kernel<<<dim3(1), dim3(64), 0, cs>>>(sa[0], sa[1], sa[2], na);
        if (hipGetLastError() != hipSuccess) cst = RREA_ERR_HIP;
    }
    hipGraph_t graph = nullptr;
    hipError_t const ce = hipStreamEndCapture(cs, &graph);
    G->graph = graph;
    if (cst != RREA_OK) return fail(cst);
    if (ce != hipSuccess || !graph) return fail(RREA_ERR_HIP);
    if (hipGraphInstantiate(&G->exec, G->graph, nullptr, nullptr, 0) != hipSuccess) return fail(RREA_ERR_HIP);
    (void)hipStreamDestroy(cs);
    *out = G;
    return RREA_OK;
}

rrea_status rrea_batch_graph_launch(rrea_batch_graph_t G, int64_t first_step, rrea_stream_t stream)
{
    if (!G || !G->exec) return RREA_ERR_ARGUMENT;
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (dev != G->device) return RREA_ERR_ARGUMENT;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    // the instances' environments as their fields handles stand now
    rrea_status const es = upload_envs(G->batch, nullptr, s);
    if (es != RREA_OK) return es;
    set_step_kernel<<<dim3(1), dim3(64), 0, s>>>(static_cast<StepArgs*>(G->d_sargs[0]),
                                                 static_cast<StepArgs*>(G->d_sargs[1]),
                                                 static_cast<StepArgs*>(G->d_sargs[2]),
                                                 static_cast<NewbornArgs*>(G->d_nargs), first_step);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipGraphLaunch(G->exec, s));
    return RREA_OK;
}

}  // extern "C"
