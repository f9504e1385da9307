/* CPU oracle for the RREA per-step charged transport -- TEST INFRASTRUCTURE.
 *
 * A plain C++17 restatement of the reference's electron, positron and
 * photon transport steps (native_amrex_rrea/warpx_overlay/Source/Rrea/
 * RreaElectronTransport.cpp, RreaPositronTransport.cpp,
 * RreaPhotonTransport.cpp and the headers they include).  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library; the shipped HIP path
 * never links or calls it.
 *
 * Every struct here is the oracle's own; the product's C-ABI declares its
 * own equivalents in include/rrea_hip.h.  Tests feed both the same inputs.
 */
#ifndef RREA_ORACLE_H
#define RREA_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_geometry {
    int32_t nr;          /* physical cells in r */
    int32_t nz;          /* physical cells in z */
    double dr;
    double dz;
    double r_lo;         /* must be 0 (the RZ axis) */
    double z_lo;
    const double* er;    /* Maxwell Er on r-faces, (nr+1)*nz, f = i + (nr+1)*j */
    const double* ez;    /* Maxwell Ez on z-faces, nr*(nz+1), f = i + nr*j */
    const double* amb_er;/* ambient Er, same layout as er */
    const double* amb_ez;/* ambient Ez, same layout as ez */
    const double* btheta;/* Btheta on nodes, (nr+1)*(nz+1), c = i + (nr+1)*j */
    /* density: mode 0 = constant ratio; 1 = altitude profile (absolute kg/m^3) */
    int32_t density_mode;
    int32_t profile_n;
    double density_ratio;
    const double* profile_alt_m;
    const double* profile_rho;
    double altitude_at_z0_m;
    /* material-source grids (5 x nr x nz: low-energy electron, positive ion,
       direct negative ion, event count, energy loss; c = i + nr*j), raw
       weights accumulated by the steps and the drain; NULL: none */
    double* sources;
} orc_geometry;

typedef struct orc_config {
    uint64_t rng_seed;
    int64_t step;
    double dt_s;
    double low_energy_cutoff_eV;
    double photon_cutoff_eV;
    double max_optical_depth_per_substep;
    uint64_t max_subcycles;
    int32_t strict_transport_guards;
    int32_t disable_hard_moller;
    double max_gs_collisions_per_substep;
    double gs_total_collision_cap;
    double max_field_impulse_fraction;
    double field_impulse_floor_eV;
} orc_config;

typedef struct orc_secondary {
    int64_t parent_index;
    int32_t species;      /* 0 e-, 1 e+, 2 photon */
    int32_t process;      /* RreaSecondaryProcess id */
    double x, y, z;
    double dir_x, dir_y, dir_z;
    double weight;
    double energy_eV;
    double birth_fraction;
    uint64_t root_key;
    uint64_t lineage;
    int32_t root_species; /* species of the tracked root particle: 0 e-, 1 e+, 2 photon */
    int32_t generation;   /* 0: staged by a live step; g: created by a generation-g newborn */
    int32_t instance;     /* -1: one simulation (the oracle runs instances one at a time) */
    int32_t reserved;
} orc_secondary;

/* Per-particle status written by the step. */
enum {
    ORC_SURVIVE = 0,
    ORC_ESCAPE = 1,
    ORC_DEMOTED = 2,
    ORC_ZERO_WEIGHT = 3,
    ORC_NO_WRITEBACK = 4,
    ORC_SKIPPED_INVALID = 5,
    ORC_ANNIHILATED = 6,   /* positron consumed: in flight, stopped, or at rest */
    ORC_ABSORBED = 7,      /* photon consumed: photoelectric, pair, or Compton below the cutoff */
};

/* Tally slots (sums over the step). */
enum {
    ORC_T_SURVIVORS = 0,
    ORC_T_ESCAPES,
    ORC_T_DEMOTIONS,
    ORC_T_MOLLER_EVENTS,
    ORC_T_BREMS_EVENTS,
    ORC_T_BREMS_PHOTONS,
    ORC_T_SUBSTEPS,
    ORC_T_DRAG_LOSS_WEV,
    ORC_T_ION_PAIRS,
    ORC_T_FIELD_WORK_WEV,
    ORC_T_ANNIHILATIONS,
    ORC_T_BHABHA_EVENTS,
    ORC_T_POSITIVE_IONS,   /* weight of positive-ion-only sources (Bhabha, annihilation) */
    ORC_T_COMPTON_EVENTS,
    ORC_T_PHOTOELECTRIC_EVENTS,
    ORC_T_PAIR_EVENTS,
    ORC_T_LOW_ENERGY_ELECTRONS, /* weight booked as low-energy electron sources: demoted electrons and
                                   sub-cutoff pair electrons */
    ORC_T_NEWBORN_ADVANCED,     /* newborns flown through their residual (1 - f) dt */
    ORC_T_NEWBORN_SURVIVORS,    /* newborns that reach the end of the step */
    ORC_T_NEWBORN_ESCAPES,      /* newborns that leave the domain */
    ORC_T_NEWBORN_TERMINATED,   /* newborns demoted, annihilated or absorbed */
    ORC_T_COUNT = 24 + 75,      /* 21-23 reserved, then the event ledger */
};

/* The event ledger (RreaEventLedger, NA/include/rrea/RreaEventLedger.H:18-96)
   from tallies[ORC_LEDGER_OFFSET] on, in the reference's field order; the
   ORC_LEDGER_IS_MAX slots are running maxima, the rest sums. */
#define ORC_LEDGER_OFFSET 24
enum {
    ORC_L_PHOTON_INTERACTION_COUNT = 0,
    ORC_L_ELECTRON_BREMS_PHOTON_COUNT,
    ORC_L_TOTAL_BREMS_PHOTON_COUNT,
    ORC_L_PHOTON_INTERACTION_WEIGHT,
    ORC_L_ELECTRON_BREMS_PHOTON_WEIGHT,
    ORC_L_ELECTRON_BREMS_PHOTON_ENERGY_EV,
    ORC_L_TOTAL_BREMS_PHOTON_WEIGHT,
    ORC_L_TOTAL_BREMS_PHOTON_ENERGY_EV,
    ORC_L_PHOTON_CONTINUOUS_SOFT_BREMS_ENERGY_EV,
    ORC_L_COMPTON_INTERACTION_COUNT,
    ORC_L_PHOTOELECTRIC_INTERACTION_COUNT,
    ORC_L_PAIR_PRODUCTION_COUNT,
    ORC_L_PAIR_NUCLEAR_COUNT,
    ORC_L_PAIR_TRIPLET_COUNT,
    ORC_L_COMPTON_INTERACTION_WEIGHT,
    ORC_L_PHOTOELECTRIC_INTERACTION_WEIGHT,
    ORC_L_PAIR_PRODUCTION_WEIGHT,
    ORC_L_PAIR_NUCLEAR_WEIGHT,
    ORC_L_PAIR_TRIPLET_WEIGHT,
    ORC_L_PAIR_RECOIL_ELECTRON_ENERGY_EV,
    ORC_L_PAIR_LOCAL_DEPOSIT_ENERGY_EV,
    ORC_L_PHOTON_SEEDED_ELECTRON_WEIGHT,
    ORC_L_PHOTON_SEEDED_POSITRON_WEIGHT,
    ORC_L_PHOTON_SEEDED_KINETIC_ELECTRON_WEIGHT,
    ORC_L_PHOTON_SEEDED_LOW_ELECTRON_WEIGHT,
    ORC_L_PHOTON_SEEDED_KINETIC_POSITRON_WEIGHT,
    ORC_L_PHOTON_SEEDED_BELOW_CUTOFF_POSITRON_WEIGHT,
    ORC_L_COMPTON_KINETIC_ELECTRON_WEIGHT,
    ORC_L_COMPTON_LOW_ELECTRON_WEIGHT,
    ORC_L_PHOTOELECTRIC_KINETIC_ELECTRON_WEIGHT,
    ORC_L_PHOTOELECTRIC_LOW_ELECTRON_WEIGHT,
    ORC_L_PAIR_KINETIC_ELECTRON_WEIGHT,
    ORC_L_PAIR_LOW_ELECTRON_WEIGHT,
    ORC_L_PAIR_KINETIC_POSITRON_WEIGHT,
    ORC_L_PAIR_BELOW_CUTOFF_POSITRON_WEIGHT,
    ORC_L_PAIR_TRIPLET_RECOIL_KINETIC_ELECTRON_WEIGHT,
    ORC_L_PAIR_TRIPLET_RECOIL_LOW_ELECTRON_WEIGHT,
    ORC_L_PHOTON_ABSORBED_ENERGY_EV,
    ORC_L_PHOTON_ESCAPED_ENERGY_EV,
    ORC_L_PHOTON_TO_CHARGED_ENERGY_EV,
    ORC_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV,
    ORC_L_POSITRON_TRANSPORT_ENERGY_LOSS_EV,
    ORC_L_POSITRON_THERMALIZED_COUNT,
    ORC_L_POSITRON_BHABHA_EVENT_COUNT,
    ORC_L_POSITRON_BHABHA_EVENT_WEIGHT,
    ORC_L_POSITRON_BHABHA_SECONDARY_ENERGY_EV,
    ORC_L_POSITRON_BHABHA_KINETIC_SECONDARY_WEIGHT,
    ORC_L_POSITRON_BREMS_EVENT_COUNT,
    ORC_L_POSITRON_BREMS_EVENT_WEIGHT,
    ORC_L_POSITRON_BREMS_PHOTON_ENERGY_EV,
    ORC_L_TRANSPORT_IONIZATION_WEIGHT,
    ORC_L_TRANSPORT_LOW_SECONDARY_WEIGHT,
    ORC_L_TRANSPORT_KINETIC_SECONDARY_WEIGHT,
    ORC_L_TRANSPORT_ENERGY_LOSS_EV,
    ORC_L_HARD_MOLLER_EVENT_COUNT,
    ORC_L_HARD_MOLLER_EVENT_WEIGHT,
    ORC_L_DIRECT_HARD_MOLLER_ELECTRON_WEIGHT,
    ORC_L_HARD_MOLLER_EXPECTED_EVENT_WEIGHT,
    ORC_L_HARD_MOLLER_SECONDARY_ENERGY_WEIGHTED_SUM_EV,
    ORC_L_HARD_MOLLER_SECONDARY_ENERGY_WEIGHT_SUM,
    ORC_L_HARD_MOLLER_SECONDARY_ENERGY_MAX_EV,
    ORC_L_PARENT_DRAG_LOSS_EV,
    ORC_L_ION_PAIR_SOURCE_LOSS_EV,
    ORC_L_RESTRICTED_COLLISION_LOSS_EV,
    ORC_L_EXPLICIT_HARD_MOLLER_ENERGY_LOSS_EV,
    ORC_L_MAX_TRANSPORT_HARD_MOLLER_TAU,
    ORC_L_MAX_TRANSPORT_BREMS_TAU,
    ORC_L_MAX_TRANSPORT_ELECTRON_TOTAL_COMPETING_TAU,
    ORC_L_MAX_TRANSPORT_PHOTON_TAU,
    ORC_L_MAX_TRANSPORT_POSITRON_TAU,
    ORC_L_MAX_TRANSPORT_SUBCYCLES,
    ORC_L_TRANSPORT_OPTICAL_DEPTH_VIOLATION_COUNT,
    ORC_L_MAX_TRANSPORT_ELECTRON_ENERGY_EV,
    ORC_L_MAX_TRANSPORT_PHOTON_ENERGY_SEEN_EV,
    ORC_L_MAX_TRANSPORT_POSITRON_ENERGY_EV,
    ORC_LEDGER_COUNT
};
#define ORC_LEDGER_IS_MAX(q) ((q) == ORC_L_HARD_MOLLER_SECONDARY_ENERGY_MAX_EV \
    || ((q) >= ORC_L_MAX_TRANSPORT_HARD_MOLLER_TAU && (q) <= ORC_L_MAX_TRANSPORT_SUBCYCLES) \
    || (q) >= ORC_L_MAX_TRANSPORT_ELECTRON_ENERGY_EV)

/* Error codes; 0 means success. */
enum {
    ORC_OK = 0,
    ORC_ERR_TABLE_RANGE = 1,
    ORC_ERR_OUTSIDE_DOMAIN = 2,
    ORC_ERR_GUARD = 3,
    ORC_ERR_INVALID = 4,
    ORC_ERR_SECONDARY_CAPACITY = 5,
};

/* One interaction group and its conservation closure (RreaSecondaryGroup,
   OV/RreaParticleInteraction.H:709-717, 777-806): entries are particles
   (kind = species 0 e-, 1 e+, 2 photon) or material sources (kind 0
   ionization, 1 positive ion, 2 low-energy electron). */
typedef struct orc_group_entry {
    int32_t kind;
    int32_t reserved;
    double weight;
    double energy_eV;
} orc_group_entry;

typedef struct orc_group {
    int32_t process, parent_species, remove_parent, parent_still_valid, closure_mode, parent_survives;
    double parent_weight, parent_energy_before_eV, parent_energy_after_eV;
    double local_energy_eV, escaped_energy_eV, escaped_charge_e;
    int32_t n_particles, n_sources;
    orc_group_entry particles[4];
    orc_group_entry sources[4];
} orc_group;

/* RreaValidateSecondaryGroupConservation: bit 0 required, 1 present, 2 finite,
   3 valid; res[2] = charge, energy residuals (may be NULL). */
int orc_validate_group(const orc_group* g, double low_energy_cutoff_eV, double* res);

/* RreaFoldEventLedger: into += from, the ledger's running maxima by max
   (ORC_T_COUNT doubles each). */
void orc_fold_tallies(double* into, const double* from);

/* One material-source record into g->sources: kind 0 ionization event,
   1 positive ion only, 2 low-energy electron (RreaWarpXCoupling::
   Accumulate*, OV/RreaCouplingSources.cpp:435-516).  0, or the error code of
   a record outside the physical domain. */
int orc_accumulate_source(const orc_geometry* g, double cutoff, int kind, double r, double z, double weight,
                          double secondary, double loss);
/* RreaEventAccumulator::NormalizeAndSync on g->sources. */
int orc_sources_normalize(const orc_geometry* g, double dt, int diagnostics);

/* The field side of the feedback loop (rrea_maxwell_oracle.cpp):
 * ConvertPaddedIndexCurrentToPhysicalRZ and the TM Maxwell leapfrog of
 * NA/include/rrea/RreaMaxwellTMRZ.H, in the reference's layouts.  Optional
 * arrays may be NULL. */
void orc_current_to_physical(int nr, int nz, double dr, double dz, int pad, double charge_per_index_flux,
                             const double* jr_index, const double* jz_index, double* jr, double* jz);
void orc_maxwell_advance_b(int nr, int nz, double dr, double dz, const double* er, const double* ez, double* bt,
                           double dt);
void orc_maxwell_silver_mueller(int nr, int nz, const double* er, const double* ez, double* bt);
void orc_maxwell_advance_e(int nr, int nz, double dr, double dz, double* er, double* ez, const double* bt, double dt,
                           const double* sigma_r, const double* sigma_z, const double* jr, const double* jz,
                           const double* amb_er, const double* amb_ez);
void orc_maxwell_step(int nr, int nz, double dr, double dz, double* er, double* ez, double* bt, double dt,
                      const double* sigma_r, const double* sigma_z, const double* jr, const double* jz,
                      const double* amb_er, const double* amb_ez);
double orc_maxwell_field_energy(int nr, int nz, double dr, double dz, const double* er, const double* ez,
                                const double* bt, const double* amb_er, const double* amb_ez);
double orc_maxwell_mixed_magnetic_energy(int nr, int nz, double dr, double dz, const double* bt,
                                         const double* bt_previous);
double orc_maxwell_joule_power(int nr, int nz, double dr, double dz, const double* er, const double* ez,
                               const double* amb_er, const double* amb_ez, const double* jr, const double* jz,
                               double* jz_moment);
double orc_maxwell_poynting_power(int nr, int nz, double dr, double dz, const double* er, const double* ez,
                                  const double* bt, const double* amb_er, const double* amb_ez);

#ifdef __cplusplus
}
#endif

#endif
