// Thin C wrappers around the reference's own AMReX-free sources, compiled
// from the reference checkout by `make -C oracle ref` into oracle/_ref/.
// Test infrastructure only: used to pin the oracle's restatement of the
// counter RNG (OV/RreaRng.cpp), the uniform stream
// (NA/include/rrea/RreaUniformStream.H), the two-point Gauss-Legendre rule
// (NA/include/rrea/RreaGaussLegendre8.H) and the RZ segment deposition
// (NA/include/rrea/RreaSegmentDepositionRZ.H).
#include "RreaRng.H"
#include "rrea/RreaGaussLegendre8.H"
#include "rrea/RreaSegmentDepositionRZ.H"
#include "rrea/RreaUniformStream.H"

#include <array>
#include <cstdint>
#include <cstring>

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

}  // extern "C"
