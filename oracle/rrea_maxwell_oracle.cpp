// CPU oracle for the field side of the feedback loop -- TEST INFRASTRUCTURE.
//
// A plain restatement of the reference's conversion of the net-chord deposit
// to physical face currents and of its axisymmetric TM Maxwell leapfrog:
//   ConvertPaddedIndexCurrentToPhysicalRZ
//     native_amrex_rrea/include/rrea/RreaSegmentDepositionRZ.H:399-453
//   MaxwellTMRZ metric, MaxwellAdvanceB, MaxwellSilverMueller,
//   MaxwellAdvanceE, MaxwellStep, MaxwellFieldEnergy, MaxwellJoulePower,
//   MaxwellBoundaryPoyntingPower
//     native_amrex_rrea/include/rrea/RreaMaxwellTMRZ.H:64-357
// Only tests/ load this (through liboracle.so); the HIP path never does.
// Layouts are the reference's: Er (nr+1)*nz, f = i + (nr+1)*j; Ez nr*(nz+1),
// f = i + nr*j; Btheta (nr+1)*(nz+1), c = i + (nr+1)*j.
#include "rrea_oracle.h"

#include <cstddef>
#include <cstdint>

namespace {

constexpr double kPi = 3.141592653589793238462643383279502884;   // RreaConstants.H:8
constexpr double kEps0 = 8.8541878128e-12;                         // RreaMaxwellTMRZ.H:58
constexpr double kMu0 = 1.25663706212e-6;                          // :59
constexpr double kC = 299792458.0;                                 // :60

// rz_shell_volume RreaConstants.H:37-43: both radii clamped to the axis
double shell(double r_in, double r_out, double dz)
{
    double const a = r_in > 0.0 ? r_in : 0.0;
    double const b = r_out > 0.0 ? r_out : 0.0;
    return kPi * (b * b - a * a) * dz;
}

struct Mesh {
    int nr, nz;
    double dr, dz;
    size_t er(int i, int j) const { return static_cast<size_t>(i) + static_cast<size_t>(nr + 1) * j; }
    size_t ez(int i, int j) const { return static_cast<size_t>(i) + static_cast<size_t>(nr) * j; }
    // RreaMaxwellTMRZ.H:98-118
    double cell_volume(int i) const { return shell(i * dr, (i + 1) * dr, dz); }
    double radial_dual(int i) const
    {
        double const hi = (i + 0.5) * dr;
        double const wall = nr * dr;
        return shell((i - 0.5) * dr, wall < hi ? wall : hi, dz);
    }
    double axial_dual(int i, int j) const { return cell_volume(i) * ((j == 0 || j == nz) ? 0.5 : 1.0); }
    double corner_dual(int i, int j) const { return radial_dual(i) * ((j == 0 || j == nz) ? 0.5 : 1.0); }
};

}  // namespace

extern "C" {

// RreaSegmentDepositionRZ.H:399-453.  The padded index grid is
// (nr + pad) x (nz + 2 pad) cells; its jr is (nr+pad+1) wide, its jz (nr+pad).
// Only the physical faces are kept; the axis face carries no flux.
void orc_current_to_physical(int nr, int nz, double dr, double dz, int pad, double charge_per_index_flux,
                             const double* jr_index, const double* jz_index, double* jr, double* jz)
{
    int const nr_pad = nr + pad;
    for (size_t k = 0; k < static_cast<size_t>(nr + 1) * nz; ++k) jr[k] = 0.0;
    for (size_t k = 0; k < static_cast<size_t>(nr) * (nz + 1); ++k) jz[k] = 0.0;
    for (int j = pad; j < pad + nz; ++j) {
        int const jp = j - pad;
        for (int i = 0; i <= nr; ++i) {
            double const flux = jr_index[static_cast<size_t>(i) + static_cast<size_t>(nr_pad + 1) * j];
            double const r_face = i * dr;
            if (flux == 0.0 || r_face <= 0.0) continue;
            jr[static_cast<size_t>(i) + static_cast<size_t>(nr + 1) * jp] +=
                charge_per_index_flux * flux / (2.0 * kPi * r_face * dz);
        }
    }
    for (int j = pad; j <= pad + nz; ++j) {
        int const jp = j - pad;
        for (int i = 0; i < nr; ++i) {
            double const flux = jz_index[static_cast<size_t>(i) + static_cast<size_t>(nr_pad) * j];
            if (flux == 0.0) continue;
            double const lo = i * dr;
            double const hi = (i + 1) * dr;
            double const ring = kPi * (hi * hi - lo * lo);
            jz[static_cast<size_t>(i) + static_cast<size_t>(nr) * jp] += charge_per_index_flux * flux / ring;
        }
    }
}

// MaxwellAdvanceB :147-156: interior corners, -curl E
void orc_maxwell_advance_b(int nr, int nz, double dr, double dz, const double* er, const double* ez, double* bt,
                           double dt)
{
    Mesh const m{nr, nz, dr, dz};
    for (int j = 1; j < nz; ++j) {
        for (int i = 1; i < nr; ++i) {
            bt[m.er(i, j)] += dt * ((ez[m.ez(i, j)] - ez[m.ez(i - 1, j)]) / dr
                                    - (er[m.er(i, j)] - er[m.er(i, j - 1)]) / dz);
        }
    }
}

// MaxwellSilverMueller :165-177: outgoing-wave corners, then the axis
void orc_maxwell_silver_mueller(int nr, int nz, const double* er, const double* ez, double* bt)
{
    Mesh const m{nr, nz, 0.0, 0.0};
    for (int j = 0; j <= nz; ++j) bt[m.er(nr, j)] = -ez[m.ez(nr - 1, j)] / kC;
    for (int i = 1; i <= nr; ++i) {
        bt[m.er(i, 0)] = -er[m.er(i, 0)] / kC;
        bt[m.er(i, nz)] = er[m.er(i, nz - 1)] / kC;
    }
    for (int j = 0; j <= nz; ++j) bt[m.er(0, j)] = 0.0;
}

// MaxwellAdvanceE :180-216: semi-implicit conduction (sigma may be NULL),
// the current and the ambient's conduction drive; the axis Er pinned to 0
void orc_maxwell_advance_e(int nr, int nz, double dr, double dz, double* er, double* ez, const double* bt, double dt,
                           const double* sigma_r, const double* sigma_z, const double* jr, const double* jz,
                           const double* amb_er, const double* amb_ez)
{
    Mesh const m{nr, nz, dr, dz};
    double const c2dt = dt / (kEps0 * kMu0);
    for (int j = 0; j < nz; ++j) {
        for (int i = 1; i <= nr; ++i) {
            size_t const f = m.er(i, j);
            double const sigma = sigma_r ? sigma_r[f] : 0.0;
            double const a = sigma * dt / (2.0 * kEps0);
            double drive = -c2dt * (bt[m.er(i, j + 1)] - bt[m.er(i, j)]) / dz;
            if (jr) drive -= dt / kEps0 * jr[f];
            if (amb_er) drive -= dt / kEps0 * sigma * amb_er[f];
            er[f] = ((1.0 - a) * er[f] + drive) / (1.0 + a);
        }
        er[m.er(0, j)] = 0.0;
    }
    for (int j = 0; j <= nz; ++j) {
        for (int i = 0; i < nr; ++i) {
            size_t const f = m.ez(i, j);
            double const sigma = sigma_z ? sigma_z[f] : 0.0;
            double const a = sigma * dt / (2.0 * kEps0);
            double drive = c2dt * ((i + 1) * bt[m.er(i + 1, j)] - i * bt[m.er(i, j)]) / ((i + 0.5) * dr);
            if (jz) drive -= dt / kEps0 * jz[f];
            if (amb_ez) drive -= dt / kEps0 * sigma * amb_ez[f];
            ez[f] = ((1.0 - a) * ez[f] + drive) / (1.0 + a);
        }
    }
}

// MaxwellStep :221-230
void orc_maxwell_step(int nr, int nz, double dr, double dz, double* er, double* ez, double* bt, double dt,
                      const double* sigma_r, const double* sigma_z, const double* jr, const double* jz,
                      const double* amb_er, const double* amb_ez)
{
    orc_maxwell_advance_b(nr, nz, dr, dz, er, ez, bt, dt);
    orc_maxwell_silver_mueller(nr, nz, er, ez, bt);
    orc_maxwell_advance_e(nr, nz, dr, dz, er, ez, bt, dt, sigma_r, sigma_z, jr, jz, amb_er, amb_ez);
}

// MaxwellFieldEnergy :241-268 (amb_er/amb_ez NULL: the scattered field's own)
double orc_maxwell_field_energy(int nr, int nz, double dr, double dz, const double* er, const double* ez,
                                const double* bt, const double* amb_er, const double* amb_ez)
{
    Mesh const m{nr, nz, dr, dz};
    double energy = 0.0;
    for (int j = 0; j < nz; ++j) {
        for (int i = 0; i <= nr; ++i) {
            double const esc = er[m.er(i, j)];
            double const amb = amb_er ? amb_er[m.er(i, j)] : 0.0;
            energy += kEps0 * (0.5 * esc * esc + esc * amb) * m.radial_dual(i);
        }
    }
    for (int j = 0; j <= nz; ++j) {
        for (int i = 0; i < nr; ++i) {
            double const esc = ez[m.ez(i, j)];
            double const amb = amb_ez ? amb_ez[m.ez(i, j)] : 0.0;
            energy += kEps0 * (0.5 * esc * esc + esc * amb) * m.axial_dual(i, j);
        }
        for (int i = 0; i <= nr; ++i) {
            double const b = bt[m.er(i, j)];
            energy += 0.5 / kMu0 * b * b * m.corner_dual(i, j);
        }
    }
    return energy;
}

// MaxwellMixedMagneticEnergy :273-285
double orc_maxwell_mixed_magnetic_energy(int nr, int nz, double dr, double dz, const double* bt,
                                         const double* bt_previous)
{
    Mesh const m{nr, nz, dr, dz};
    double energy = 0.0;
    for (int j = 1; j < nz; ++j) {
        for (int i = 1; i < nr; ++i) {
            energy += 0.5 / kMu0 * bt[m.er(i, j)] * bt_previous[m.er(i, j)] * m.corner_dual(i, j);
        }
    }
    return energy;
}

// MaxwellJoulePower :291-316: sum J . E_total dV (jz_moment may be NULL)
double orc_maxwell_joule_power(int nr, int nz, double dr, double dz, const double* er, const double* ez,
                               const double* amb_er, const double* amb_ez, const double* jr, const double* jz,
                               double* jz_moment)
{
    Mesh const m{nr, nz, dr, dz};
    double power = 0.0;
    for (int j = 0; j < nz; ++j) {
        for (int i = 0; i <= nr; ++i) {
            size_t const k = m.er(i, j);
            power += jr[k] * (er[k] + (amb_er ? amb_er[k] : 0.0)) * m.radial_dual(i);
        }
    }
    for (int j = 0; j <= nz; ++j) {
        for (int i = 0; i < nr; ++i) {
            size_t const k = m.ez(i, j);
            double const v = m.axial_dual(i, j);
            power += jz[k] * (ez[k] + (amb_ez ? amb_ez[k] : 0.0)) * v;
            if (jz_moment) *jz_moment += jz[k] * v;
        }
    }
    return power;
}

// MaxwellBoundaryPoyntingPower :326-357: the exact discrete boundary power
// of the total field
double orc_maxwell_poynting_power(int nr, int nz, double dr, double dz, const double* er, const double* ez,
                                  const double* bt, const double* amb_er, const double* amb_ez)
{
    Mesh const m{nr, nz, dr, dz};
    auto const ert = [&](int i, int j) { return er[m.er(i, j)] + (amb_er ? amb_er[m.er(i, j)] : 0.0); };
    auto const ezt = [&](int i, int j) { return ez[m.ez(i, j)] + (amb_ez ? amb_ez[m.ez(i, j)] : 0.0); };
    auto const b = [&](int i, int j) { return bt[m.er(i, j)]; };
    double bracket = 0.0;
    for (int i = 1; i <= nr; ++i) {
        double const area = m.radial_dual(i) / dz;
        bracket -= area * (ert(i, nz - 1) * b(i, nz) - ert(i, 0) * b(i, 0));
        if (i == nr) {
            for (int j = 1; j < nz; ++j) bracket += area * b(i, j) * (ert(i, j) - ert(i, j - 1));
        }
    }
    for (int j = 0; j <= nz; ++j) {
        double const ring = 2.0 * kPi * dz * ((j == 0 || j == nz) ? 0.5 : 1.0);
        bracket += ring * nr * dr * b(nr, j) * ezt(nr - 1, j);
        if (j == 0 || j == nz) {
            for (int i = 1; i < nr; ++i) bracket -= ring * i * dr * b(i, j) * (ezt(i, j) - ezt(i - 1, j));
        }
    }
    return -bracket / kMu0;
}

}  // extern "C"
