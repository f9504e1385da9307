// MI355X-native field side of the RREA feedback loop: the conversion of the
// transport step's net-chord deposit (padded index space) to physical face
// currents, and the axisymmetric TM Maxwell leapfrog that those currents
// drive.  Follows
//   ConvertPaddedIndexCurrentToPhysicalRZ
//     native_amrex_rrea/include/rrea/RreaSegmentDepositionRZ.H:399-453
//   MaxwellAdvanceB / MaxwellSilverMueller / MaxwellAdvanceE / MaxwellStep and
//   the energy ledger (MaxwellFieldEnergy, MaxwellJoulePower,
//   MaxwellBoundaryPoyntingPower)
//     native_amrex_rrea/include/rrea/RreaMaxwellTMRZ.H:64-357
// Every face and corner update is independent of the others, so one lane
// computes one output with the reference's own operations in its order: the
// fields are bit-identical to the serial loops.  The interior Faraday update
// and the Silver-Mueller boundary corners write disjoint corners from the
// same E, so one kernel does both.  The ledger sums are reduced per
// workgroup and added with FP64 atomics (summation order differs from the
// serial loops).
//
// Build: hipcc --offload-arch=gfx950 -O3 -ffp-contract=off.

#include "../../include/rrea_hip.h"

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>

namespace {

#define HIP_TRY(expr)                                                              \
    do {                                                                           \
        hipError_t const e_ = (expr);                                              \
        if (e_ != hipSuccess) {                                                    \
            std::fprintf(stderr, "rrea_hip: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return RREA_ERR_HIP;                                                   \
        }                                                                          \
    } while (0)

constexpr double kPi = 3.141592653589793238462643383279502884;  // RreaConstants.H:8
constexpr double kEps0 = 8.8541878128e-12;                        // RreaMaxwellTMRZ.H:58
constexpr double kMu0 = 1.25663706212e-6;                         // :59
constexpr double kC = 299792458.0;                                // :60
constexpr int kThreads = 256;

struct Mesh {
    int nr, nz;
    double dr, dz;
    __device__ __forceinline__ int64_t er(int i, int j) const { return i + static_cast<int64_t>(nr + 1) * j; }
    __device__ __forceinline__ int64_t ez(int i, int j) const { return i + static_cast<int64_t>(nr) * j; }
    // rz_shell_volume (RreaConstants.H:37-43) and the duals of :98-118
    __device__ __forceinline__ static double shell(double a, double b, double dz)
    {
        double const in = a > 0.0 ? a : 0.0;
        double const out = b > 0.0 ? b : 0.0;
        return kPi * (out * out - in * in) * dz;
    }
    __device__ __forceinline__ double cell_volume(int i) const { return shell(i * dr, (i + 1) * dr, dz); }
    __device__ __forceinline__ double radial_dual(int i) const
    {
        double const hi = (i + 0.5) * dr;
        double const wall = nr * dr;
        return shell((i - 0.5) * dr, hi < wall ? hi : wall, dz);
    }
    __device__ __forceinline__ double axial_dual(int i, int j) const
    {
        return cell_volume(i) * ((j == 0 || j == nz) ? 0.5 : 1.0);
    }
    __device__ __forceinline__ double corner_dual(int i, int j) const
    {
        return radial_dual(i) * ((j == 0 || j == nz) ? 0.5 : 1.0);
    }
};

// One lane per physical face of either family (jr faces first).
__global__ void current_to_physical_kernel(int nr, int nz, double dr, double dz, int pad, double q,
                                           const double* __restrict__ jr_index, const double* __restrict__ jz_index,
                                           double* __restrict__ jr, double* __restrict__ jz)
{
    int64_t const n_er = static_cast<int64_t>(nr + 1) * nz;
    int64_t const n_ez = static_cast<int64_t>(nr) * (nz + 1);
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    int const nr_pad = nr + pad;
    if (k < n_er) {
        int const i = static_cast<int>(k % (nr + 1));
        int const jp = static_cast<int>(k / (nr + 1));
        double const flux = jr_index[i + static_cast<int64_t>(nr_pad + 1) * (jp + pad)];
        double const r_face = i * dr;
        // the axis face has zero area and carries no flux
        jr[k] = (flux == 0.0 || r_face <= 0.0) ? 0.0 : q * flux / (2.0 * kPi * r_face * dz);
    } else if (k < n_er + n_ez) {
        int64_t const m = k - n_er;
        int const i = static_cast<int>(m % nr);
        int const jp = static_cast<int>(m / nr);
        double const flux = jz_index[i + static_cast<int64_t>(nr_pad) * (jp + pad)];
        double const lo = i * dr;
        double const hi = (i + 1) * dr;
        double const ring = kPi * (hi * hi - lo * lo);
        jz[m] = flux == 0.0 ? 0.0 : q * flux / ring;
    }
}

// MaxwellAdvanceB (interior corners, -curl E) and/or MaxwellSilverMueller
// (boundary corners from the tangential E, then the axis) -- one lane per
// corner; `which` bit 0: Faraday, bit 1: the absorber.
__global__ void corner_kernel(Mesh m, const double* __restrict__ er, const double* __restrict__ ez,
                              double* __restrict__ bt, double dt, int which)
{
    int64_t const n = static_cast<int64_t>(m.nr + 1) * (m.nz + 1);
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int const i = static_cast<int>(k % (m.nr + 1));
    int const j = static_cast<int>(k / (m.nr + 1));
    bool const interior = i >= 1 && i < m.nr && j >= 1 && j < m.nz;
    if (interior) {
        if (which & 1) {
            bt[k] += dt * ((ez[m.ez(i, j)] - ez[m.ez(i - 1, j)]) / m.dr - (er[m.er(i, j)] - er[m.er(i, j - 1)]) / m.dz);
        }
        return;
    }
    if (!(which & 2)) return;
    // the reference's write order decides the shared corners: the outer
    // column, then the z rows (i >= 1), then the axis
    if (i == 0) bt[k] = 0.0;
    else if (j == 0) bt[k] = -er[m.er(i, 0)] / kC;
    else if (j == m.nz) bt[k] = er[m.er(i, m.nz - 1)] / kC;
    else bt[k] = -ez[m.ez(m.nr - 1, j)] / kC;  // i == nr
}

// MaxwellAdvanceE: one lane per face (Er faces first); sigma, J and the
// ambient may be NULL.
__global__ void advance_e_kernel(Mesh m, double* __restrict__ er, double* __restrict__ ez,
                                 const double* __restrict__ bt, double dt, const double* __restrict__ sigma_r,
                                 const double* __restrict__ sigma_z, const double* __restrict__ jr,
                                 const double* __restrict__ jz, const double* __restrict__ amb_er,
                                 const double* __restrict__ amb_ez)
{
    int64_t const n_er = static_cast<int64_t>(m.nr + 1) * m.nz;
    int64_t const n_ez = static_cast<int64_t>(m.nr) * (m.nz + 1);
    int64_t const k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    double const c2dt = dt / (kEps0 * kMu0);
    if (k < n_er) {
        int const i = static_cast<int>(k % (m.nr + 1));
        int const j = static_cast<int>(k / (m.nr + 1));
        if (i == 0) {  // the axis, pinned to 0
            er[k] = 0.0;
            return;
        }
        double const sigma = sigma_r ? sigma_r[k] : 0.0;
        double const a = sigma * dt / (2.0 * kEps0);
        double drive = -c2dt * (bt[m.er(i, j + 1)] - bt[m.er(i, j)]) / m.dz;
        if (jr) drive -= dt / kEps0 * jr[k];
        if (amb_er) drive -= dt / kEps0 * sigma * amb_er[k];
        er[k] = ((1.0 - a) * er[k] + drive) / (1.0 + a);
    } else if (k < n_er + n_ez) {
        int64_t const f = k - n_er;
        int const i = static_cast<int>(f % m.nr);
        int const j = static_cast<int>(f / m.nr);
        double const sigma = sigma_z ? sigma_z[f] : 0.0;
        double const a = sigma * dt / (2.0 * kEps0);
        double drive = c2dt * ((i + 1) * bt[m.er(i + 1, j)] - i * bt[m.er(i, j)]) / ((i + 0.5) * m.dr);
        if (jz) drive -= dt / kEps0 * jz[f];
        if (amb_ez) drive -= dt / kEps0 * sigma * amb_ez[f];
        ez[f] = ((1.0 - a) * ez[f] + drive) / (1.0 + a);
    }
}

__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}

// Adds the workgroup's sums of four terms to out[0..3].
__device__ __forceinline__ void block_add(double (&t)[4], double* out)
{
    __shared__ double red[4][kThreads / 64];
    int const lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int q = 0; q < 4; ++q) {
        double const s = wave_sum(t[q]);
        if (lane == 0) red[q][wave] = s;
    }
    __syncthreads();
    if (threadIdx.x < 4) {
        double s = 0.0;
        for (int w = 0; w < kThreads / 64; ++w) s += red[threadIdx.x][w];
        if (s != 0.0) atomicAdd(&out[threadIdx.x], s);
    }
}

// The energy ledger, grid-stride over every face and corner:
// out[0] MaxwellFieldEnergy (ambient-relative when the ambient is given),
// out[1] MaxwellJoulePower (J given), out[2] the boundary-face part of
// MaxwellBoundaryPoyntingPower's bracket, out[3] the Jz moment.  The host
// divides out[2] by -mu0.
__global__ void __launch_bounds__(kThreads) ledger_kernel(Mesh m, const double* __restrict__ er,
                                                          const double* __restrict__ ez,
                                                          const double* __restrict__ bt,
                                                          const double* __restrict__ amb_er,
                                                          const double* __restrict__ amb_ez,
                                                          const double* __restrict__ jr,
                                                          const double* __restrict__ jz, double* __restrict__ out)
{
    int64_t const n_er = static_cast<int64_t>(m.nr + 1) * m.nz;
    int64_t const n_ez = static_cast<int64_t>(m.nr) * (m.nz + 1);
    int64_t const n_bt = static_cast<int64_t>(m.nr + 1) * (m.nz + 1);
    int64_t const total = n_er + n_ez + n_bt;
    double t[4] = {0.0, 0.0, 0.0, 0.0};
    auto const ert = [&](int i, int j) { return er[m.er(i, j)] + (amb_er ? amb_er[m.er(i, j)] : 0.0); };
    auto const ezt = [&](int i, int j) { return ez[m.ez(i, j)] + (amb_ez ? amb_ez[m.ez(i, j)] : 0.0); };
    int64_t const stride = static_cast<int64_t>(gridDim.x) * blockDim.x;
    for (int64_t k = static_cast<int64_t>(blockIdx.x) * blockDim.x + threadIdx.x; k < total; k += stride) {
        if (k < n_er) {
            int const i = static_cast<int>(k % (m.nr + 1));
            int const j = static_cast<int>(k / (m.nr + 1));
            double const esc = er[k];
            double const amb = amb_er ? amb_er[k] : 0.0;
            double const v = m.radial_dual(i);
            t[0] += kEps0 * (0.5 * esc * esc + esc * amb) * v;
            if (jr) t[1] += jr[k] * (esc + amb) * v;
        } else if (k < n_er + n_ez) {
            int64_t const f = k - n_er;
            int const i = static_cast<int>(f % m.nr);
            int const j = static_cast<int>(f / m.nr);
            double const esc = ez[f];
            double const amb = amb_ez ? amb_ez[f] : 0.0;
            double const v = m.axial_dual(i, j);
            t[0] += kEps0 * (0.5 * esc * esc + esc * amb) * v;
            if (jz) {
                t[1] += jz[f] * (esc + amb) * v;
                t[3] += jz[f] * v;
            }
        } else {
            int64_t const c = k - n_er - n_ez;
            int const i = static_cast<int>(c % (m.nr + 1));
            int const j = static_cast<int>(c / (m.nr + 1));
            double const b = bt[c];
            t[0] += 0.5 / kMu0 * b * b * m.corner_dual(i, j);
            // the Poynting bracket's terms that involve corner (i, j)
            double br = 0.0;
            if (i >= 1) {
                double const area = m.radial_dual(i) / m.dz;
                if (j == m.nz) br -= area * (ert(i, m.nz - 1) * b);
                if (j == 0) br += area * (ert(i, 0) * b);
                if (i == m.nr && j >= 1 && j < m.nz) br += area * b * (ert(i, j) - ert(i, j - 1));
            }
            double const ring = 2.0 * kPi * m.dz * ((j == 0 || j == m.nz) ? 0.5 : 1.0);
            if (i == m.nr) br += ring * m.nr * m.dr * b * ezt(m.nr - 1, j);
            if ((j == 0 || j == m.nz) && i >= 1 && i < m.nr) br -= ring * i * m.dr * b * (ezt(i, j) - ezt(i - 1, j));
            t[2] += br;
        }
    }
    block_add(t, out);
}

__global__ void ledger_finish_kernel(double* __restrict__ out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) out[2] = -out[2] / kMu0;
}

__global__ void zero4_kernel(double* __restrict__ out)
{
    if (threadIdx.x < 4 && blockIdx.x == 0) out[threadIdx.x] = 0.0;
}

bool mesh_of(const rrea_rz_geometry* g, Mesh& m)
{
    if (!g || g->nr < 1 || g->nz < 1 || !(g->dr > 0.0) || !(g->dz > 0.0) || g->r_lo != 0.0) return false;
    m = Mesh{g->nr, g->nz, g->dr, g->dz};
    return true;
}

unsigned blocks(int64_t n) { return static_cast<unsigned>((n + kThreads - 1) / kThreads); }

}  // namespace

extern "C" {

rrea_status rrea_current_to_physical(const rrea_rz_geometry* g, int32_t pad_cells, double charge_per_index_flux_c,
                                     const double* jr_index, const double* jz_index, double* jr, double* jz,
                                     rrea_stream_t stream)
{
    Mesh m;
    if (!mesh_of(g, m) || pad_cells < 0 || !jr_index || !jz_index || !jr || !jz) return RREA_ERR_ARGUMENT;
    int64_t const n = static_cast<int64_t>(m.nr + 1) * m.nz + static_cast<int64_t>(m.nr) * (m.nz + 1);
    hipStream_t const s = static_cast<hipStream_t>(stream);
    current_to_physical_kernel<<<dim3(blocks(n)), dim3(kThreads), 0, s>>>(m.nr, m.nz, m.dr, m.dz, pad_cells,
                                                                         charge_per_index_flux_c, jr_index, jz_index,
                                                                         jr, jz);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

rrea_status rrea_maxwell_advance_b(const rrea_rz_geometry* g, const double* er, const double* ez, double* btheta,
                                   double dt, rrea_stream_t stream)
{
    Mesh m;
    if (!mesh_of(g, m) || !er || !ez || !btheta) return RREA_ERR_ARGUMENT;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    corner_kernel<<<dim3(blocks(static_cast<int64_t>(m.nr + 1) * (m.nz + 1))), dim3(kThreads), 0, s>>>(
        m, er, ez, btheta, dt, 1);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

rrea_status rrea_maxwell_silver_mueller(const rrea_rz_geometry* g, const double* er, const double* ez,
                                        double* btheta, rrea_stream_t stream)
{
    Mesh m;
    if (!mesh_of(g, m) || !er || !ez || !btheta) return RREA_ERR_ARGUMENT;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    corner_kernel<<<dim3(blocks(static_cast<int64_t>(m.nr + 1) * (m.nz + 1))), dim3(kThreads), 0, s>>>(
        m, er, ez, btheta, 0.0, 2);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

rrea_status rrea_maxwell_advance_e(const rrea_rz_geometry* g, double* er, double* ez, const double* btheta, double dt,
                                   const double* sigma_r, const double* sigma_z, const double* jr, const double* jz,
                                   const double* amb_er, const double* amb_ez, rrea_stream_t stream)
{
    Mesh m;
    if (!mesh_of(g, m) || !er || !ez || !btheta) return RREA_ERR_ARGUMENT;
    int64_t const n = static_cast<int64_t>(m.nr + 1) * m.nz + static_cast<int64_t>(m.nr) * (m.nz + 1);
    hipStream_t const s = static_cast<hipStream_t>(stream);
    advance_e_kernel<<<dim3(blocks(n)), dim3(kThreads), 0, s>>>(m, er, ez, btheta, dt, sigma_r, sigma_z, jr, jz,
                                                               amb_er, amb_ez);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

rrea_status rrea_maxwell_step(const rrea_rz_geometry* g, double* er, double* ez, double* btheta, double dt,
                              const double* sigma_r, const double* sigma_z, const double* jr, const double* jz,
                              const double* amb_er, const double* amb_ez, rrea_stream_t stream)
{
    Mesh m;
    if (!mesh_of(g, m) || !er || !ez || !btheta) return RREA_ERR_ARGUMENT;
    hipStream_t const s = static_cast<hipStream_t>(stream);
    // B -> absorber -> E (MaxwellStep :221-230); the first two write
    // disjoint corners from the same E
    corner_kernel<<<dim3(blocks(static_cast<int64_t>(m.nr + 1) * (m.nz + 1))), dim3(kThreads), 0, s>>>(
        m, er, ez, btheta, dt, 3);
    int64_t const n = static_cast<int64_t>(m.nr + 1) * m.nz + static_cast<int64_t>(m.nr) * (m.nz + 1);
    advance_e_kernel<<<dim3(blocks(n)), dim3(kThreads), 0, s>>>(m, er, ez, btheta, dt, sigma_r, sigma_z, jr, jz,
                                                               amb_er, amb_ez);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

rrea_status rrea_maxwell_ledger(const rrea_rz_geometry* g, const double* er, const double* ez, const double* btheta,
                                const double* amb_er, const double* amb_ez, const double* jr, const double* jz,
                                double* out, rrea_stream_t stream)
{
    Mesh m;
    if (!mesh_of(g, m) || !er || !ez || !btheta || !out || (!amb_er) != (!amb_ez) || (!jr) != (!jz)) {
        return RREA_ERR_ARGUMENT;
    }
    hipStream_t const s = static_cast<hipStream_t>(stream);
    int64_t const n = static_cast<int64_t>(m.nr + 1) * m.nz + static_cast<int64_t>(m.nr) * (m.nz + 1)
        + static_cast<int64_t>(m.nr + 1) * (m.nz + 1);
    zero4_kernel<<<dim3(1), dim3(64), 0, s>>>(out);
    ledger_kernel<<<dim3(std::min<unsigned>(blocks(n), 1024U)), dim3(kThreads), 0, s>>>(m, er, ez, btheta, amb_er,
                                                                                       amb_ez, jr, jz, out);
    ledger_finish_kernel<<<dim3(1), dim3(64), 0, s>>>(out);
    HIP_TRY(hipGetLastError());
    return RREA_OK;
}

}  // extern "C"
