/* rrea_hip.h -- C-ABI of the MI355X-native RREA transport step.
 *
 * The drop-in boundary for the reference's per-step hook
 *   RreaParticleInteraction::Apply(WarpX&, RreaWarpXCoupling&,
 *       RreaInteractionTables const&, int step, Real time_s, Real dt_s)
 *   (native_amrex_rrea/warpx_overlay/Source/Rrea/RreaParticleInteraction.H:1609-1623,
 *    RreaParticleInteraction.cpp:12-62)
 * and for the services it reads from the coupling and the table library.
 * Plain C: pointers and sizes, no torch or AMReX types.  Every entry point
 * returns an rrea_status (0 = success); device-side failures are reported
 * through the step's error word and nothing is committed by the host glue on
 * error, as the reference fails closed before any commit.
 *
 * Threading: stateless apart from the handles; one HIP stream per call; no
 * internal threads.  Calls on one handle must not overlap.
 */
#ifndef RREA_HIP_H
#define RREA_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t rrea_status;
typedef struct rrea_tables_s* rrea_tables_t;
typedef struct rrea_fields_s* rrea_fields_t;
typedef void* rrea_stream_t; /* a hipStream_t; NULL = the default stream */

enum {
    RREA_OK = 0,
    RREA_ERR_TABLE_RANGE = 1,    /* a certified-range table query: the reference Aborts */
    RREA_ERR_OUTSIDE_DOMAIN = 2, /* a survivor stored outside the domain */
    RREA_ERR_GUARD = 3,          /* strict optical-depth subcycle guard */
    RREA_ERR_INVALID = 4,        /* any other fail-closed invariant */
    RREA_ERR_SECONDARY_CAPACITY = 5,
    RREA_ERR_ARGUMENT = 16,      /* host-side argument check */
    RREA_ERR_HIP = 17,           /* a HIP runtime call failed */
};

/* ------------------------------------------------------------------------
 * Tables.  Replaces RreaInteractionTables::Load + its accessors
 * (native_amrex_rrea/src/RreaInteractionTables.cpp:1659-2230, :2272-3265).
 * The host loader fills this image with the bundle's columns exactly as Load
 * groups them; rrea_tables_create derives the node logarithms, the
 * hard-Moller mean secondary energy (Load :2146-2179) and the brems CDF
 * group means on the host, then uploads everything to the current device.
 * ------------------------------------------------------------------------ */
typedef struct rrea_rate_column {
    const double* x;   /* energy nodes [eV], nondecreasing */
    const double* y;   /* values */
    int64_t n;
    int32_t mode;      /* 0 loglog, 1 loglog with zero threshold, 2 logx-linear, 3 linear */
} rrea_rate_column;

typedef struct rrea_cdf_groups {
    const double* primary; /* group primary energies, strictly increasing */
    int64_t n_groups;
    const int64_t* start;  /* n_groups + 1 row offsets into u/v */
    const double* u;       /* cdf nodes per group, strictly increasing */
    const double* v;       /* sampled values */
} rrea_cdf_groups;

typedef struct rrea_angle_table {
    const double* e; int64_t n_e;  /* energy grid */
    const double* n; int64_t n_n;  /* path n = s / lambda_el grid */
    const double* u; int64_t n_u;  /* shared cdf nodes */
    const double* omc;             /* one_minus_cos, (ie*n_n + ip)*n_u + iu */
} rrea_angle_table;

typedef struct rrea_elastic_table {
    const double* e;
    const double* mfp_m_stp;
    const double* screening_eta;
    int64_t n;
} rrea_elastic_table;

/* Photoelectric subshell CDF (load_photoelectric_subshell_cdf,
 * NA/src/RreaInteractionTables.cpp:893-990): groups by primary energy (a
 * duplicated absorption-edge energy opens a second group), each a strictly
 * increasing cdf over the subshell rows with their binding energies. */
typedef struct rrea_photoelectric_table {
    const double* primary;   /* n_groups, nondecreasing */
    int64_t n_groups;
    const int64_t* start;    /* n_groups + 1 row offsets */
    const double* cdf_u;
    const double* binding_eV;
} rrea_photoelectric_table;

/* Pair final-state bank, training blocks only (load_pair_empirical_blocks
 * :1187-1369), sorted by (channel, energy); each record's leptons in the
 * order pair electron, positron, triplet recoil (zero for nuclear). */
typedef struct rrea_pair_bank {
    const int64_t* channel;  /* per block: 1 nuclear, 2 triplet */
    const double* energy_eV; /* per block */
    int64_t n_blocks;
    const int64_t* start;    /* n_blocks + 1 record offsets */
    const double* fraction;  /* 3 per record: kinetic fraction of (E - 2 m c^2) */
    const double* direction; /* 9 per record: unit directions in the photon frame */
} rrea_pair_bank;

typedef struct rrea_table_image {
    /* e- collision, soft radiative, hard-Moller 1/m, tracked brems 1/m, elastic transport 1/m,
       e+ collision, soft radiative, hard-Bhabha, tracked brems, elastic transport, annihilation,
       gamma Compton, photoelectric, pair nuclear, pair triplet */
    rrea_rate_column rate[15];
    rrea_cdf_groups moller, bhabha, electron_brems, positron_brems;
    rrea_angle_table gs_electron, gs_positron, wentzel_electron, wentzel_positron;
    rrea_elastic_table elastic_electron, elastic_positron;
    double hard_moller_secondary_threshold_eV;
    double brems_absolute_cutoff_eV;
    double reference_density_kg_m3;
    double charged_energy_range_eV[2];
    double photon_energy_range_eV[2];
    double gs_few_collision_threshold_n;
    double gs_isotropic_above_path_n;
    double gs_wentzel_transition_energy_eV;
    rrea_photoelectric_table photoelectric;
    rrea_pair_bank pair;
} rrea_table_image;

rrea_status rrea_tables_create(const rrea_table_image* image, rrea_tables_t* out);
rrea_status rrea_tables_destroy(rrea_tables_t tables);
/* Host copy of the derived hard-Moller mean secondary energy on the rate grid
 * (for checks); n must equal the hard-Moller rate column length. */
rrea_status rrea_tables_hm_mean(rrea_tables_t tables, double* out, int64_t n);

/* ------------------------------------------------------------------------
 * Fields and density.  Replaces RreaWarpXCoupling::GatherFieldRZ
 * (warpx_overlay/Source/Rrea/RreaWarpXCoupling.cpp:556-576) over the
 * replicated Maxwell + ambient state of RreaAmrexAdvance::GatherTotalFieldRZ
 * (native_amrex_rrea/src/RreaAmrexAdvance.cpp:532-579), and
 * TransportDensityRatioAtZ (RreaWarpXCoupling.cpp:79-95).
 * ------------------------------------------------------------------------ */
typedef struct rrea_rz_geometry {
    int32_t nr, nz;     /* physical cells */
    double dr, dz;
    double r_lo;        /* must be 0: the axis */
    double z_lo;
} rrea_rz_geometry;

rrea_status rrea_fields_create(const rrea_rz_geometry* geometry, rrea_fields_t* out);
rrea_status rrea_fields_destroy(rrea_fields_t fields);
/* Er/Ez faces of the Maxwell state and of the ambient field, Btheta nodes, in
 * the RreaMaxwellTMRZ layouts: er (nr+1)*nz, ez nr*(nz+1), btheta
 * (nr+1)*(nz+1).  `device_pointers` says whether they are device memory.  The
 * cell-centred totals the gather reads are formed on the device, in the
 * reference's order of operations. */
rrea_status rrea_fields_update(rrea_fields_t fields, const double* er, const double* ez,
                               const double* amb_er, const double* amb_ez, const double* btheta,
                               int32_t device_pointers, rrea_stream_t stream);
rrea_status rrea_density_set_constant(rrea_fields_t fields, double density_ratio);
/* Absolute density [kg/m^3] against altitude [m]; the ratio divides by the
 * table set's reference density.  altitude(z) = altitude_at_z0_m + z. */
rrea_status rrea_density_set_profile(rrea_fields_t fields, const double* altitude_m, const double* rho_kg_m3,
                                     int64_t n, double altitude_at_z0_m, double reference_density_kg_m3);

/* ------------------------------------------------------------------------
 * One transport step of one species' tile.
 * Replaces the electron phase of RreaParticleInteraction::Apply,
 * TransportStep::ApplyElectronTransport
 * (warpx_overlay/Source/Rrea/RreaElectronTransport.cpp:9-902), including the
 * net-chord replay into the padded SegmentGridRZ
 * (RreaWarpXCoupling.cpp:417-544).
 * ------------------------------------------------------------------------ */
typedef struct rrea_step_config {
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
    int32_t complete_path_only; /* nonzero: every particle through the complete
                                   step, no fast-path specialisation (testing) */
    int32_t reserved;
} rrea_step_config;

/* Device SoA of one tile (the ChargedSoAView of RreaTransportStep.H:454-468).
 * prev_* is the pre-push snapshot the step reads; r..uz receive the
 * written-back state and may alias prev_* (fused in-place mode). */
typedef struct rrea_tile {
    int64_t n;
    const double* prev_r; const double* prev_z; const double* prev_theta;
    const double* prev_ux; const double* prev_uy; const double* prev_uz;
    const double* w;
    int64_t* id;         /* in/out: an invalidated particle gets RREA_INVALID_ID and is
                            skipped by later steps (particle_id.make_invalid() /
                            is_valid(), RreaElectronTransport.cpp:60-63) */
    const int32_t* cpu;  /* may be NULL: cpu 0 */
    double* r; double* z; double* theta; double* ux; double* uy; double* uz;
} rrea_tile;

/* One staged secondary, before the newborn residual advance
 * (RreaSecondaryParticle, RreaParticleInteraction.H:612-635). */
typedef struct rrea_secondary {
    int64_t parent_index;
    int32_t species;        /* 0 e-, 1 e+, 2 photon */
    int32_t process;        /* RreaSecondaryProcess id: 2 hard Moller, 3 hard Bhabha,
                               4 brems, 9 positron annihilation */
    double x, y, z;
    double dir_x, dir_y, dir_z;
    double weight;
    double energy_eV;
    double birth_fraction;  /* elapsed fraction of the outer step at birth */
    uint64_t root_key;      /* |id| ^ (cpu << 40) of the tracked root particle */
    uint64_t lineage;       /* transient lineage: the parent's spawn counter for a live parent,
                               the root's next free lineage for a newborn's child */
    int32_t root_species;   /* species of the tracked root: 0 e-, 1 e+, 2 photon */
    int32_t generation;     /* 0: staged by a live step; g: made by a generation-g newborn */
    int32_t instance;       /* the simulation instance of a batched step; -1 for a plain step */
    int32_t reserved;
} rrea_secondary;

enum {
    RREA_SURVIVE = 0,
    RREA_ESCAPE = 1,
    RREA_DEMOTED = 2,
    RREA_ZERO_WEIGHT = 3,
    RREA_NO_WRITEBACK = 4,
    RREA_SKIPPED_INVALID = 5,
    RREA_ANNIHILATED = 6,   /* positron consumed: in flight, stopped mid-chord, or at rest */
    RREA_ABSORBED = 7,      /* photon consumed: photoelectric, pair, or Compton below the cutoff */
};

#define RREA_INVALID_ID INT64_MIN

#define RREA_TALLY_COUNT (RREA_LEDGER_OFFSET + RREA_LEDGER_COUNT)
/* Tally slots: 0 survivors, 1 escapes, 2 demotions, 3 hard-Moller events,
 * 4 brems events, 5 tracked brems photons, 6 substeps, 7 continuous drag
 * (weight x eV), 8 ion pairs, 9 field work (weight x eV), 10 annihilations,
 * 11 hard-Bhabha events, 12 positive-ion-only sources (weight), 13 Compton
 * events, 14 photoelectric events, 15 pair events, 16 low-energy electron
 * sources (weight: demoted electrons and sub-cutoff pair electrons);
 * rrea_newborn_drain adds 17 newborns advanced, 18 newborn survivors,
 * 19 newborn escapes, 20 newborns demoted, annihilated or absorbed;
 * 21-23 reserved; then the event ledger. */

/* The event ledger (RreaEventLedger, native_amrex_rrea/include/rrea/
 * RreaEventLedger.H:18-96), in the reference's field order, from
 * tallies[RREA_LEDGER_OFFSET] on.  Counts are kept as exact double sums.
 * The RREA_LEDGER_IS_MAX slots are running maxima of non-negative values
 * (folded with max, as RreaFoldEventLedger, OV/RreaTransportLog.H:45-127);
 * every other slot is a sum.  A zeroed tally array is the empty ledger. */
#define RREA_LEDGER_OFFSET 24
enum {
    RREA_L_PHOTON_INTERACTION_COUNT = 0,
    RREA_L_ELECTRON_BREMS_PHOTON_COUNT,
    RREA_L_TOTAL_BREMS_PHOTON_COUNT,
    RREA_L_PHOTON_INTERACTION_WEIGHT,
    RREA_L_ELECTRON_BREMS_PHOTON_WEIGHT,
    RREA_L_ELECTRON_BREMS_PHOTON_ENERGY_EV,
    RREA_L_TOTAL_BREMS_PHOTON_WEIGHT,
    RREA_L_TOTAL_BREMS_PHOTON_ENERGY_EV,
    RREA_L_PHOTON_CONTINUOUS_SOFT_BREMS_ENERGY_EV,
    RREA_L_COMPTON_INTERACTION_COUNT,
    RREA_L_PHOTOELECTRIC_INTERACTION_COUNT,          /* 10 */
    RREA_L_PAIR_PRODUCTION_COUNT,
    RREA_L_PAIR_NUCLEAR_COUNT,
    RREA_L_PAIR_TRIPLET_COUNT,
    RREA_L_COMPTON_INTERACTION_WEIGHT,
    RREA_L_PHOTOELECTRIC_INTERACTION_WEIGHT,
    RREA_L_PAIR_PRODUCTION_WEIGHT,
    RREA_L_PAIR_NUCLEAR_WEIGHT,
    RREA_L_PAIR_TRIPLET_WEIGHT,
    RREA_L_PAIR_RECOIL_ELECTRON_ENERGY_EV,
    RREA_L_PAIR_LOCAL_DEPOSIT_ENERGY_EV,             /* 20 */
    RREA_L_PHOTON_SEEDED_ELECTRON_WEIGHT,
    RREA_L_PHOTON_SEEDED_POSITRON_WEIGHT,
    RREA_L_PHOTON_SEEDED_KINETIC_ELECTRON_WEIGHT,
    RREA_L_PHOTON_SEEDED_LOW_ELECTRON_WEIGHT,
    RREA_L_PHOTON_SEEDED_KINETIC_POSITRON_WEIGHT,
    RREA_L_PHOTON_SEEDED_BELOW_CUTOFF_POSITRON_WEIGHT,
    RREA_L_COMPTON_KINETIC_ELECTRON_WEIGHT,
    RREA_L_COMPTON_LOW_ELECTRON_WEIGHT,
    RREA_L_PHOTOELECTRIC_KINETIC_ELECTRON_WEIGHT,
    RREA_L_PHOTOELECTRIC_LOW_ELECTRON_WEIGHT,        /* 30 */
    RREA_L_PAIR_KINETIC_ELECTRON_WEIGHT,
    RREA_L_PAIR_LOW_ELECTRON_WEIGHT,
    RREA_L_PAIR_KINETIC_POSITRON_WEIGHT,
    RREA_L_PAIR_BELOW_CUTOFF_POSITRON_WEIGHT,
    RREA_L_PAIR_TRIPLET_RECOIL_KINETIC_ELECTRON_WEIGHT,
    RREA_L_PAIR_TRIPLET_RECOIL_LOW_ELECTRON_WEIGHT,
    RREA_L_PHOTON_ABSORBED_ENERGY_EV,
    RREA_L_PHOTON_ESCAPED_ENERGY_EV,
    RREA_L_PHOTON_TO_CHARGED_ENERGY_EV,
    RREA_L_PHOTON_LOCAL_DEPOSIT_ENERGY_EV,           /* 40 */
    RREA_L_POSITRON_TRANSPORT_ENERGY_LOSS_EV,
    RREA_L_POSITRON_THERMALIZED_COUNT,
    RREA_L_POSITRON_BHABHA_EVENT_COUNT,
    RREA_L_POSITRON_BHABHA_EVENT_WEIGHT,
    RREA_L_POSITRON_BHABHA_SECONDARY_ENERGY_EV,
    RREA_L_POSITRON_BHABHA_KINETIC_SECONDARY_WEIGHT,
    RREA_L_POSITRON_BREMS_EVENT_COUNT,
    RREA_L_POSITRON_BREMS_EVENT_WEIGHT,
    RREA_L_POSITRON_BREMS_PHOTON_ENERGY_EV,
    RREA_L_TRANSPORT_IONIZATION_WEIGHT,              /* 50 */
    RREA_L_TRANSPORT_LOW_SECONDARY_WEIGHT,
    RREA_L_TRANSPORT_KINETIC_SECONDARY_WEIGHT,
    RREA_L_TRANSPORT_ENERGY_LOSS_EV,
    RREA_L_HARD_MOLLER_EVENT_COUNT,
    RREA_L_HARD_MOLLER_EVENT_WEIGHT,
    RREA_L_DIRECT_HARD_MOLLER_ELECTRON_WEIGHT,
    RREA_L_HARD_MOLLER_EXPECTED_EVENT_WEIGHT,
    RREA_L_HARD_MOLLER_SECONDARY_ENERGY_WEIGHTED_SUM_EV,
    RREA_L_HARD_MOLLER_SECONDARY_ENERGY_WEIGHT_SUM,
    RREA_L_HARD_MOLLER_SECONDARY_ENERGY_MAX_EV,      /* 60, max */
    RREA_L_PARENT_DRAG_LOSS_EV,
    RREA_L_ION_PAIR_SOURCE_LOSS_EV,
    RREA_L_RESTRICTED_COLLISION_LOSS_EV,
    RREA_L_EXPLICIT_HARD_MOLLER_ENERGY_LOSS_EV,
    RREA_L_MAX_TRANSPORT_HARD_MOLLER_TAU,            /* max */
    RREA_L_MAX_TRANSPORT_BREMS_TAU,                  /* max */
    RREA_L_MAX_TRANSPORT_ELECTRON_TOTAL_COMPETING_TAU, /* max */
    RREA_L_MAX_TRANSPORT_PHOTON_TAU,                 /* max */
    RREA_L_MAX_TRANSPORT_POSITRON_TAU,               /* max */
    RREA_L_MAX_TRANSPORT_SUBCYCLES,                  /* 70, max */
    RREA_L_TRANSPORT_OPTICAL_DEPTH_VIOLATION_COUNT,
    RREA_L_MAX_TRANSPORT_ELECTRON_ENERGY_EV,         /* max */
    RREA_L_MAX_TRANSPORT_PHOTON_ENERGY_SEEN_EV,      /* max */
    RREA_L_MAX_TRANSPORT_POSITRON_ENERGY_EV,         /* max */
    RREA_LEDGER_COUNT                                /* 75 */
};
#define RREA_LEDGER_IS_MAX(slot) ((slot) == RREA_L_HARD_MOLLER_SECONDARY_ENERGY_MAX_EV \
    || ((slot) >= RREA_L_MAX_TRANSPORT_HARD_MOLLER_TAU && (slot) <= RREA_L_MAX_TRANSPORT_SUBCYCLES) \
    || (slot) >= RREA_L_MAX_TRANSPORT_ELECTRON_ENERGY_EV)

typedef struct rrea_step_out {
    uint8_t* status;          /* per particle, RREA_SURVIVE.. (may be NULL) */
    double* ion_pairs;        /* per particle ion-pair weight (may be NULL) */
    int32_t* substeps;        /* per particle substep count (may be NULL) */
    rrea_secondary* secondaries; int64_t secondary_capacity;
    int64_t* secondary_count; /* device int64; the step adds to it */
    double* jr; double* jz; double* rho; /* padded net-chord grid, accumulated */
    double* tallies;          /* device double[RREA_TALLY_COUNT], accumulated */
    int32_t* error;           /* device int32[2]: first error code, particle index */
    int32_t order_secondaries;/* nonzero: sort the staged secondaries by (parent, lineage); 1: the call
                                 reads the staged count back once to size the sort; 2: the count stays
                                 on the device (one-workgroup sort, no host round trip) */
    double* sources;          /* material-source grids (RREA_SOURCE_FIELDS x nr x nz, cell c = i + nr*j),
                                 accumulated raw (see rrea_sources_normalize); NULL: not accumulated */
} rrea_step_out;

rrea_status rrea_transport_step_electrons(rrea_tables_t tables, rrea_fields_t fields, const rrea_tile* tile,
                                          const rrea_step_config* config, const rrea_step_out* out,
                                          rrea_stream_t stream);

/* One positron transport step over a tile: the particle loop of
 * TransportStep::ApplyPositronTransport
 * (warpx_overlay/Source/Rrea/RreaPositronTransport.cpp:10-871) -- hard
 * Bhabha, tracked brems and in-flight annihilation as competing hazards,
 * at-rest annihilation of a stopped or sub-cutoff positron.  Same arguments
 * and error behaviour as rrea_transport_step_electrons. */
rrea_status rrea_transport_step_positrons(rrea_tables_t tables, rrea_fields_t fields, const rrea_tile* tile,
                                          const rrea_step_config* config, const rrea_step_out* out,
                                          rrea_stream_t stream);

/* One photon transport step over a tile: the particle loop of
 * TransportStep::ApplyPhotonTransport
 * (warpx_overlay/Source/Rrea/RreaPhotonTransport.cpp:9-461).  The chord runs
 * from the captured pre-push point prev_(r, theta, z) to WarpX's pushed point
 * (r, theta, z), both read from the tile; u = E * dir in ux, uy, uz.  One
 * exponential depth is sampled over the remaining chord per pass: Compton
 * (a survivor continues from the interaction point along the scattered
 * direction for its unused path), photoelectric absorption, or pair
 * production (nuclear or triplet, a replay of the pair bank).  Written back:
 * r, theta, z and ux, uy, uz; the tile's prev_u* are not read.  Photons carry
 * no charge: jr, jz and rho of `out` may be NULL and are not touched;
 * `substeps` receives the number of interactions. */
rrea_status rrea_transport_step_photons(rrea_tables_t tables, rrea_fields_t fields, const rrea_tile* tile,
                                        const rrea_step_config* config, const rrea_step_out* out,
                                        rrea_stream_t stream);

/* Secondary-group conservation.  Replaces RreaValidateSecondaryGroupConservation
 * (warpx_overlay/Source/Rrea/RreaParticleInteraction.H:1099-1251) over the
 * RreaSecondaryGroup of :777-806 and its closure (:709-717).  The transport
 * kernels run the same check on every interaction they make (hard Moller,
 * hard Bhabha, brems, annihilation, Compton, photoelectric, pair) and report
 * a violation as RREA_ERR_INVALID, which stops the commit of the step's
 * survivors (rrea_commit_survivors commits nothing once the error word is
 * set), as the reference refuses to commit (RreaSecondaryCommit.cpp:240-265).
 * This entry point checks caller-built groups, one result per group. */
typedef struct rrea_group_entry {
    int32_t kind;        /* a particle: species 0 e-, 1 e+, 2 photon; a material source:
                            0 ionization event, 1 positive ion, 2 low-energy electron */
    int32_t reserved;
    double weight;       /* macro weight (a source's event weight) */
    double energy_eV;    /* kinetic (photon) energy; a source's secondary energy */
} rrea_group_entry;

typedef struct rrea_secondary_group {
    int32_t process;            /* RreaSecondaryProcess id: 0 continuous transport, 1 seed (no closure
                                   required), 2 hard Moller .. 12 low-energy demotion */
    int32_t parent_species;     /* 0 e-, 1 e+, 2 photon */
    int32_t remove_parent;
    int32_t parent_still_valid; /* the parent was still valid at preflight */
    int32_t closure_mode;       /* 0 missing, 1 detailed, 2 explicit zero */
    int32_t parent_survives;
    double parent_weight;
    double parent_energy_before_eV;
    double parent_energy_after_eV;
    double local_energy_eV;     /* weighted local, recoil, binding and sub-cutoff energy */
    double escaped_energy_eV;   /* weighted */
    double escaped_charge_e;
    int32_t n_particles;        /* 0..4 */
    int32_t n_sources;          /* 0..4 */
    rrea_group_entry particles[4];
    rrea_group_entry sources[4];
} rrea_secondary_group;

/* result (device int32 per group): bit 0 closure required, bit 1 present,
 * bit 2 finite, bit 3 valid; residuals (device double[2] per group, may be
 * NULL): charge (e) and energy (eV) residuals of a detailed closure. */
rrea_status rrea_validate_secondary_groups(const rrea_secondary_group* groups, int64_t n,
                                           double low_energy_cutoff_eV, int32_t* result, double* residuals,
                                           rrea_stream_t stream);

/* The species step without allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory during a step (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The same step as rrea_transport_step_electrons (species 0),
 * rrea_transport_step_positrons (1) or rrea_transport_step_photons (2), with
 * the same arguments, results and error behaviour, except:
 *   - the call allocates nothing: its workspace (argument block, hand-over
 *     list, partial rows, staging and ordering buffers) is one fixed device
 *     array per device, about 9 MiB, held until the process ends;
 *   - tile->n is at most 1,048,576 and, when ordering, secondary_capacity is
 *     at most 32,768 (RREA_ERR_ARGUMENT beyond either);
 *   - a nonzero order_secondaries always orders with the count on the device
 *     (the order_secondaries = 2 mode); kernel timing and the fast-path signal
 *     of the fields handle do not apply;
 *   - every call on a device shares that device's workspace, so calls that
 *     overlap on one device (other streams, other threads) race. */
rrea_status rrea_transport_step_noalloc(int32_t species, rrea_tables_t tables, rrea_fields_t fields,
                                        const rrea_tile* tile, const rrea_step_config* config,
                                        const rrea_step_out* out, rrea_stream_t stream);

/* The straight-line photon push at c along u / |u| that WarpX applies before
 * the hook (no counterpart in the transport path: a stand-in for the caller's
 * own push so that a photon tile arrives as rrea_transport_step_photons
 * expects).  For every valid photon: prev_(r, theta, z) <- (r, theta, z), then
 * (r, theta, z) advance by c dt in Cartesian.  prev_* are written through the
 * tile's const pointers. */
rrea_status rrea_photon_push(const rrea_tile* tile, double dt_s, rrea_stream_t stream);

/* Reorders one tile's particles by the cell of their captured position (r
 * fastest) and, within a cell, by kinetic-energy band (photon energy for
 * photon tiles, photon != 0); invalid ids go last.  Every distinct SoA array
 * of the tile (prev_*, w, id, cpu and r..uz) is permuted together, written
 * through the tile's pointers.  No counterpart in the overlay: this is the
 * particle sort WarpX applies to its tiles, offered so that a caller can keep
 * a tile cell-ordered, which lets the transport kernels accumulate a
 * workgroup's current in LDS and gather fields from cache.  Results of a
 * step do not depend on the order except through the order of the staged
 * secondaries (by parent index) and floating-point summation order. */
rrea_status rrea_tile_sort(rrea_fields_t fields, const rrea_tile* tile, int32_t photon, rrea_stream_t stream);
/* The same with the energy key's width chosen: energy_bits bits of energy
 * bands above 1 keV, 2^(6 - energy_bits) quarter-octaves wide, so the bands
 * span 16 octaves at every width (0: cell order only; -1: the default, 6, or
 * what the cell key leaves of 32 bits).  Tuning knob (no reference
 * counterpart). */
rrea_status rrea_tile_sort_banded(rrea_fields_t fields, const rrea_tile* tile, int32_t photon, int32_t energy_bits,
                                  rrea_stream_t stream);

/* The newborn residual advance of one step.  Replaces the transient FIFO of
 * TransportStep::DrainAndCommitSecondaries
 * (warpx_overlay/Source/Rrea/RreaSecondaryCommit.cpp:338-468) with its
 * per-entry advance TransportStep::AdvanceChargedNewborn /
 * AdvancePhotonNewborn (RreaNewbornAdvance.cpp:50-994): every secondary born
 * at elapsed fraction f < 1 is flown through the remaining (1 - f) dt with
 * the live loops' physics, its draws taken from the root particle's stream
 * at its transient lineage; the children it makes join the back of the FIFO
 * with the root's next free lineage.  Validation failures that the reference
 * aborts on are reported through `error` before anything is returned.
 *
 * `work[0, n_initial)` holds the step's staged secondaries in Apply's order
 * (the electron, photon and positron steps' ordered outputs, concatenated);
 * each root's records must be contiguous, as an ordered step output is.  The
 * call appends every later generation behind them, so work_capacity bounds
 * the whole FIFO.  Synchronises `stream` once per generation. */
typedef struct rrea_drain_io {
    rrea_secondary* work;        /* device FIFO: in [0, n_initial), out the whole FIFO */
    int64_t work_capacity;
    int64_t n_initial;
    int64_t* work_count;         /* device int64: FIFO length after the drain */
    uint8_t* status;             /* device, per FIFO entry (may be NULL; work_capacity entries):
                                    RREA_SURVIVE, RREA_ESCAPE, RREA_DEMOTED, RREA_ANNIHILATED or RREA_ABSORBED */
    double* ion_pairs;           /* device, per FIFO entry ion-pair weight (may be NULL) */
    int32_t* substeps;           /* device, per FIFO entry (may be NULL): residual substeps of a charged
                                    newborn, interactions of a photon newborn */
    rrea_secondary* survivors;   /* device: the entries alive at the end of the step, in FIFO order,
                                    at their end-of-step state with birth_fraction 1 */
    int64_t survivor_capacity;
    int64_t* survivor_count;     /* device int64 */
    double* jr; double* jz; double* rho;  /* the step's padded net-chord grid, accumulated */
    double* tallies;             /* device double[RREA_TALLY_COUNT], accumulated */
    int32_t* error;              /* device int32[2]: first error code, FIFO index */
    double* sources;             /* material-source grids as rrea_step_out.sources (NULL: none) */
} rrea_drain_io;

rrea_status rrea_newborn_drain(rrea_tables_t tables, rrea_fields_t fields, const rrea_step_config* config,
                               const rrea_drain_io* io, rrea_stream_t stream);

/* Batched transport: independent simulation instances stepped together, one
 * launch per species over the particles of every instance (no counterpart in
 * the overlay, which runs one simulation per process).  Each instance keeps
 * its own fields handle (geometry, fields, density), its own padded J grid
 * (jr, jz, rho device pointers, as in rrea_step_out) and its own RNG seed; a
 * batched tile holds the particles of all instances with an instance index
 * per particle (device int32).  An instance's particle-steps are those of a
 * plain step of that instance alone with rng_seed = its seed, except for the
 * summation order of its J grid, the tallies (summed over the batch) and the
 * parent indices of the staged secondaries (indices into the batched tile;
 * each record carries its instance).  The batch's workspace handle
 * (rrea_batch_workspace) is where kernel timing is enabled and read. */
typedef struct rrea_batch_s* rrea_batch_t;
rrea_status rrea_batch_create(rrea_fields_t const* fields, int32_t n, double* const* jr, double* const* jz,
                              double* const* rho, const uint64_t* rng_seeds, rrea_batch_t* out);
rrea_status rrea_batch_destroy(rrea_batch_t batch);
rrea_status rrea_batch_workspace(rrea_batch_t batch, rrea_fields_t* out);
/* species: 0 electrons, 1 positrons, 2 photons; out->jr/jz/rho are not used. */
rrea_status rrea_transport_step_batched(rrea_tables_t tables, rrea_batch_t batch, int32_t species,
                                        const rrea_tile* tile, const int32_t* instance,
                                        const rrea_step_config* config, const rrea_step_out* out,
                                        rrea_stream_t stream);
/* The newborn drain of a batched step: each FIFO record advances in its own
 * instance (rrea_secondary.instance); io->jr/jz/rho are not used. */
rrea_status rrea_newborn_drain_batched(rrea_tables_t tables, rrea_batch_t batch, const rrea_step_config* config,
                                       const rrea_drain_io* io, rrea_stream_t stream);
/* rrea_tile_sort_banded for a batched tile: particles ordered by instance,
 * then by the cell of their captured position on that instance's own mesh
 * (r fastest), then by energy band; invalid ids and instances out of range
 * go last.  Every distinct SoA array of the tile and the instance array
 * (device int32, written) are permuted together.  Keeps a batched
 * wavefront's lanes in one instance (the fast path's uniform environment)
 * and a workgroup's chords in a few cells (its LDS deposit window) as an
 * avalanche appends newborns of every instance behind the seeds.  No
 * counterpart in the overlay (WarpX's particle sort, per instance);
 * energy_bits as rrea_tile_sort_banded, within what the instance and cell
 * keys leave of 31 bits. */
rrea_status rrea_batch_tile_sort(rrea_batch_t batch, const rrea_tile* tile, int32_t photon, int32_t* instance,
                                 int32_t energy_bits, rrea_stream_t stream);

/* Graph-captured batched steps (no counterpart in the overlay, which steps one
 * simulation per process and returns to the host between phases).  One graph
 * holds `steps` whole steps of a batch: per step, the listed device ranges
 * cleared (the instances' J and source grids), the record count zeroed, the
 * optional photon push, the species steps in Apply's order (electrons,
 * photons, positrons; RreaParticleInteraction.cpp:12-62) with their
 * secondaries ordered into `records`, then the newborn drain of those records
 * as the FIFO (rrea_newborn_drain_batched), then the next step number.
 * Nothing in a step reads a count on the host: the records are ordered with
 * their count on the device, and the drain's generation bounds live on the
 * device, up to max_generations generations per step (a FIFO still holding
 * an unadvanced generation after the last is RREA_ERR_SECONDARY_CAPACITY,
 * index -2).  The particle states, records, FIFO and survivors equal those of
 * rrea_transport_step_batched + rrea_newborn_drain_batched on the same inputs
 * (the J grids to summation order).  Tiles keep their size: invalidated
 * particles stay in place and are skipped, as in a plain step.  Per-particle
 * and per-entry outputs (status, ion pairs, substeps) are not written.  The
 * tiles, instance arrays and buffers named here are bound at creation and
 * must outlive the graph; every launch uploads the instances' environments
 * from their fields handles, so field and density updates between launches
 * are seen.  Errors accumulate in `error` and are read by the caller.  With
 * `counts` set (below) the tiles grow instead: each step commits its
 * survivors, the avalanche of the reference's Apply + AddNParticles. */
typedef struct rrea_batch_graph_desc {
    const rrea_tile* tiles[3];       /* electrons, positrons, photons (index = species); NULL or n == 0: absent */
    const int32_t* instance[3];      /* each tile's per-particle instance index (device int32) */
    rrea_step_config config;         /* step is replaced by the launch's step numbers */
    rrea_secondary* records;         /* device: the step's ordered secondaries, then the whole FIFO */
    int64_t record_capacity;
    int64_t* record_count;           /* device int64: FIFO length at the end of each step */
    rrea_secondary* survivors;       /* device: the drain's survivors, FIFO order */
    int64_t survivor_capacity;
    int64_t* survivor_count;         /* device int64 */
    double* tallies;                 /* device double[RREA_TALLY_COUNT], accumulated */
    int32_t* error;                  /* device int32[2] */
    double* clear[4];                /* device double ranges zeroed at the start of every step */
    int64_t clear_n[4];
    int32_t max_generations;         /* newborn FIFO generations per step, 1..64 */
    int32_t steps;                   /* steps per launch, 1..1024 */
    int32_t push_photons;            /* nonzero: the straight photon push (rrea_photon_push) opens each step */
    int32_t reserved;
    /* The avalanche.  With counts != NULL every step ends with the commit of
     * the drain's survivors into the tiles (rrea_commit_survivors, the tiles
     * as its stores, instance arrays included), so each instance's newborns
     * are transported from the next step on: tiles[k]->n is then tile k's
     * capacity and counts[k] (device int64[3]) its length, the particles a
     * step transports; a tile of capacity 0 is absent.  Ids continue from
     * *next_id (device int64); committed particles get the cpu word `cpu`.
     * Survivors beyond a tile's capacity: RREA_ERR_SECONDARY_CAPACITY and
     * nothing committed that step. */
    int64_t* counts;
    int64_t* next_id;
    int32_t cpu;
    int32_t reserved2;
} rrea_batch_graph_desc;
typedef struct rrea_batch_graph_s* rrea_batch_graph_t;
rrea_status rrea_batch_graph_create(rrea_tables_t tables, rrea_batch_t batch, const rrea_batch_graph_desc* desc,
                                    rrea_batch_graph_t* out);
/* Enqueues the graph's steps first_step, first_step + 1, ... on `stream`. */
rrea_status rrea_batch_graph_launch(rrea_batch_graph_t graph, int64_t first_step, rrea_stream_t stream);
rrea_status rrea_batch_graph_destroy(rrea_batch_graph_t graph);

/* The newborn drain of one plain step without a host round trip: the
 * species' ordered records (staged[k], staged_counts[k] device int64,
 * staged_caps[k]; Apply's order) are concatenated into io->work on the
 * device, then the FIFO is advanced with its generation bounds on the
 * device, up to max_generations generations (a FIFO still holding an
 * unadvanced generation is RREA_ERR_SECONDARY_CAPACITY, index -2); the FIFO,
 * its lineages and the survivors are rrea_newborn_drain's; io->n_initial is
 * not read, and the per-entry status, ion pairs and substeps are not
 * written.  Replaces the same FIFO as rrea_newborn_drain
 * (RreaSecondaryCommit.cpp:338-468). */
rrea_status rrea_newborn_drain_staged(rrea_tables_t tables, rrea_fields_t fields, const rrea_step_config* config,
                                      const rrea_secondary* const* staged, const int64_t* const* staged_counts,
                                      const int64_t* staged_caps, int32_t n_staged, const rrea_drain_io* io,
                                      int32_t max_generations, rrea_stream_t stream);

/* Scheduling support (no counterpart in the overlay, which runs the species
 * loops one after another on the host).  With `enable`, every species step
 * on this handle records the handle's event on its stream right after the
 * fast-path kernel, before the hand-over kernel and the ordering;
 * rrea_stream_wait_fast_path makes `stream` wait for the last record, so
 * other work (another species' pass) can start once the bulk of a step is
 * done and run under the latency-bound hand-over.  Enqueue the wait after
 * the step call that records it. */
rrea_status rrea_fields_signal_fast_path(rrea_fields_t fields, int32_t enable);
rrea_status rrea_stream_wait_fast_path(rrea_fields_t fields, rrea_stream_t stream);
/* Tuning knob (no counterpart in the overlay): the workgroup size of the
 * fast-path kernels of the species steps on this handle -- 128, 256 (the
 * default) or 512 lanes.  Each workgroup keeps an LDS deposit window of 8
 * faces per lane and folds its tallies into one partial row, so the size
 * trades window flushes and fold rows against scheduling granularity.
 * Results do not depend on it (the J grid and tallies to summation order).
 * For a batch, set it on the batch's workspace handle
 * (rrea_batch_workspace). */
rrea_status rrea_fields_set_block(rrea_fields_t fields, int32_t threads);

/* ------------------------------------------------------------------------
 * The field side of the feedback loop: the step's net-chord deposit as
 * physical face currents, and the axisymmetric TM Maxwell leapfrog those
 * currents drive.  Layouts are the reference's (rrea_fields_update's): Er on
 * radial faces (nr+1)*nz, Ez on axial faces nr*(nz+1), Btheta on corners
 * (nr+1)*(nz+1); all pointers are device memory; geometry r_lo must be 0.
 * Every field value is bit-identical to the reference's serial loops.
 * ------------------------------------------------------------------------ */
/* Replaces ConvertPaddedIndexCurrentToPhysicalRZ
 * (native_amrex_rrea/include/rrea/RreaSegmentDepositionRZ.H:399-453), called
 * from RreaAmrexAdvance::AssembleTotalFaceCurrent
 * (native_amrex_rrea/src/RreaAmrexAdvance.cpp:860-863) with pad_cells 2 and
 * charge_per_index_flux_c = qe (RreaWarpXCoupling.cpp:715-717): the padded
 * index-space jr/jz of rrea_step_out to A/m^2 on the physical faces (the pad
 * support past absorbing faces is dropped; the axis face carries none). */
rrea_status rrea_current_to_physical(const rrea_rz_geometry* geometry, int32_t pad_cells,
                                     double charge_per_index_flux_c, const double* jr_index, const double* jz_index,
                                     double* jr, double* jz, rrea_stream_t stream);
/* Replace MaxwellAdvanceB, MaxwellSilverMueller, MaxwellAdvanceE and
 * MaxwellStep (native_amrex_rrea/include/rrea/RreaMaxwellTMRZ.H:147-230):
 * the Faraday update of the interior corners, the first-order Silver-Mueller
 * boundary corners (and the axis), the semi-implicit Ampere update (sigma_r/
 * sigma_z conductivity, jr/jz current, amb_er/amb_ez the static ambient the
 * scattered field rides on; each may be NULL), and the three in order. */
rrea_status rrea_maxwell_advance_b(const rrea_rz_geometry* geometry, const double* er, const double* ez,
                                   double* btheta, double dt, rrea_stream_t stream);
rrea_status rrea_maxwell_silver_mueller(const rrea_rz_geometry* geometry, const double* er, const double* ez,
                                        double* btheta, rrea_stream_t stream);
rrea_status rrea_maxwell_advance_e(const rrea_rz_geometry* geometry, double* er, double* ez, const double* btheta,
                                   double dt, const double* sigma_r, const double* sigma_z, const double* jr,
                                   const double* jz, const double* amb_er, const double* amb_ez,
                                   rrea_stream_t stream);
rrea_status rrea_maxwell_step(const rrea_rz_geometry* geometry, double* er, double* ez, double* btheta, double dt,
                              const double* sigma_r, const double* sigma_z, const double* jr, const double* jz,
                              const double* amb_er, const double* amb_ez, rrea_stream_t stream);
/* The energy ledger of RreaAmrexAdvance::AdvanceOneStep (:694-846) on the
 * device: out[0] MaxwellFieldEnergy (:241-268; ambient-relative when amb_er
 * and amb_ez are given), out[1] MaxwellJoulePower (:291-316; jr and jz
 * given, else 0), out[2] MaxwellBoundaryPoyntingPower (:326-357), out[3] the
 * Jz moment sum Jz dV.  `out` is device double[4], overwritten; the sums are
 * workgroup-reduced, so they agree with the serial loops to round-off. */
rrea_status rrea_maxwell_ledger(const rrea_rz_geometry* geometry, const double* er, const double* ez,
                                const double* btheta, const double* amb_er, const double* amb_ez, const double* jr,
                                const double* jz, double* out, rrea_stream_t stream);

/* ------------------------------------------------------------------------
 * Population control: the charge-exact thinning of the overlay's
 * adaptive_resample_v1 controller for one charged species on one rank.
 * Replaces the thinning plan of
 * RreaWarpXCoupling::PreflightAndApplySecondaryPopulationControl
 * (warpx_overlay/Source/Rrea/RreaCouplingPopulation.cpp:175-520, ranks = 1)
 * over RreaPlanGlobalCicThinning / RreaThinCicWeights
 * (native_amrex_rrea/include/rrea/RreaCicConservingResample.H:236-617).
 * Each record's order-1 cell-centred CIC support and phase bin are formed
 * as the coupling's record lambda does (:233-277); records are grouped by
 * support in (support, phase, candidate, id) order; thinning engages above
 * 1.25 x target and removes min(count - target, exact capacity) records in
 * support order, each support's weights moved along null directions of its
 * four CIC columns with the controller's counter draws
 * (RreaAdaptiveResampleCicUniform, RreaRng.H:112-129), so every support's
 * four deposited charges are preserved.  Weights are rewritten in place;
 * with `id` given, a record whose weight reached zero gets RREA_INVALID_ID.
 * All pointers are device memory; every record's weight must be positive.
 * Synchronises `stream` (a host-planned pass, not part of the step).
 * ------------------------------------------------------------------------ */
typedef struct rrea_population_records {
    int64_t n;
    const double* r; const double* z; const double* theta;   /* position (z_lo-relative domain) */
    const double* ux; const double* uy; const double* uz;    /* u = gamma v [m/s] */
    double* w;                       /* in/out: macro weights */
    const uint64_t* stable_id;       /* the record's stable id (low word) */
    const uint64_t* stable_id_high;  /* may be NULL: 0 (live particles) */
    const uint8_t* candidate;        /* may be NULL: 0 (live particles; 1 = a staged secondary) */
    int64_t* id;                     /* may be NULL: invalidated where the weight reaches zero */
} rrea_population_records;

typedef struct rrea_population_config {
    double z_lo, dr, dz;             /* the mesh (r_lo = 0, non-periodic) */
    int64_t target;                  /* the species' target macro count */
    uint64_t rng_seed;               /* rrea.rng_seed */
    uint64_t rng_salt;               /* population_controller_rng_salt */
    int64_t step;
    int32_t species;                 /* 0 electrons, 1 positrons (the draw channel and support stream) */
    int32_t reserved;
} rrea_population_config;

/* counts (host int64[4]): exact removal capacity, geometric floor, the
 * nominal request, records removed.  Device errors (an invalid record or
 * support plan, insufficient progress) go to `error` (device int32[2]) and
 * leave every weight unchanged, as the reference aborts before applying. */
rrea_status rrea_population_thin(const rrea_population_records* records, const rrea_population_config* config,
                                 int64_t* counts, int32_t* error, rrea_stream_t stream);

/* Device glue on either side of the hook (the WarpX calls the reference
 * makes around Apply, done here without leaving HBM). */

/* One species' particle store: tile arrays with room for `capacity`
 * particles (tile.n is not used; charged stores may alias prev_* and r..uz,
 * photon stores need no prev_u*), and, for a batched store, each particle's
 * instance (NULL otherwise). */
typedef struct rrea_particle_store {
    rrea_tile tile;
    int32_t* instance;
    int64_t capacity;
} rrea_particle_store;

/* Commit of the drain's survivors: CreateLocal{KineticElectrons,Photons,
 * KineticPositrons} -> CommitLocalParticleBatch -> AddNParticles
 * (RreaSecondaryCommit.cpp:594-629, RreaCouplingSources.cpp:239-370).
 * stores[0..2] are the electron, positron and photon stores (the species
 * codes of rrea_secondary).  Each species' survivors, in FIFO order, are
 * appended behind counts[species] (device int64[3], updated); a survivor of
 * non-positive weight, or a photon of non-positive energy, is skipped.
 * Charged momentum u = u(K) dir/|dir| (+z for a zero direction; K <= 0 keeps
 * the raw direction), photon u = E dir/|dir|; r = hypot(x, y), theta =
 * atan2(y, x); previous state = the new state.  Ids are consecutive from
 * *next_id (device int64, advanced) in the reference's commit order:
 * electrons, then photons, then positrons; cpu is the caller's rank.  If a
 * store would overflow, nothing is written and error[0] = 5 (capacity).
 * Fail-closed: when error[0] is already set (the step or the drain reported
 * an error, a group that did not conserve among them), nothing is committed
 * (RreaSecondaryCommit.cpp:240-265). */
rrea_status rrea_commit_survivors(const rrea_secondary* survivors, const int64_t* survivor_count,
                                  const rrea_particle_store* stores, int64_t* counts, int64_t* next_id,
                                  int32_t cpu, int32_t* error, rrea_stream_t stream);

/* deleteInvalidParticles (RreaSecondaryCommit.cpp:578-592): a stable
 * partition of the tile, live particles (id != RREA_INVALID_ID) first in
 * their order, invalidated ones behind; `instance` (may be NULL) is permuted
 * with the tile.  *n_live (host) receives the live count; the call
 * synchronises its stream once to return it.  Transport results do not
 * depend on it: steps skip invalidated particles wherever they are. */
rrea_status rrea_tile_compact(rrea_fields_t fields, const rrea_tile* tile, int32_t photon, int32_t* instance,
                              int64_t* n_live, rrea_stream_t stream);

/* Differentiation of the drift segment (no counterpart in the reference,
 * which has no adjoint).  The drift step is the charged step's one-substep
 * branch -- chord, GL2 field work, Strang half kicks, CSDA drag, net-chord
 * deposit -- with the discrete events and the elastic deflection frozen
 * out; a particle whose chord reaches the domain boundary or whose energy
 * falls below the cutoff ends its drift there (removed, no deposit for that
 * step).  rrea_drift_forward runs `steps` drift steps over the tile in place
 * (r..uz), depositing into jr/jz (the padded net-chord grid of rrea_step_out,
 * accumulated over all steps; may be NULL) and, with loss weights gjr/gjz
 * (dL/dJ on the same faces), each particle's sum of J x dL/dJ into loss[n]
 * (an order-independent per-particle loss, for finite differences).  It
 * writes the state every checkpoint_every steps into checkpoints
 * ([n_ck][6][n] doubles: r, theta, z, ux, uy, uz; n_ck = ceil(steps /
 * checkpoint_every)) and the live flags into checkpoint_alive ([n_ck][n]).
 * rrea_drift_adjoint then forms dL/dE for L = sum(J x dL/dJ) over the
 * fields' Maxwell Er (r-faces, (nr+1)*nz) and Ez (z-faces, nr*(nz+1)),
 * accumulated into grad_er / grad_ez: a reverse sweep over the checkpoint
 * segments with recompute, each step's local Jacobian from a tangent-linear
 * (dual-number) evaluation of the same step. */
typedef struct rrea_drift_config {
    int32_t species;            /* 0 electrons, 1 positrons */
    int32_t steps;
    int32_t checkpoint_every;   /* 1..16 */
    int32_t reserved;
    double dt_s;
    double low_energy_cutoff_eV;
} rrea_drift_config;

rrea_status rrea_drift_forward(rrea_tables_t tables, rrea_fields_t fields, const rrea_tile* tile,
                               const rrea_drift_config* config, double* checkpoints, uint8_t* checkpoint_alive,
                               uint8_t* alive, double* jr, double* jz, const double* gjr, const double* gjz,
                               double* loss, rrea_stream_t stream);
rrea_status rrea_drift_adjoint(rrea_tables_t tables, rrea_fields_t fields, const rrea_tile* tile,
                               const rrea_drift_config* config, const double* checkpoints,
                               const uint8_t* checkpoint_alive, const double* gjr, const double* gjz,
                               double* grad_er, double* grad_ez, rrea_stream_t stream);

/* Material sources (TransportSideEffects::Accumulate* replayed into
 * RreaWarpXCoupling::Accumulate{IonizationEvent,PositiveIonSourceOnly,
 * LowEnergyElectron}, OV/RreaCouplingSources.cpp:435-516, into
 * RreaEventAccumulator, NA/src/RreaEventAccumulator.cpp).  With
 * rrea_step_out.sources / rrea_drain_io.sources set, every ionization,
 * positive-ion and low-energy-electron record of the step -- continuous ion
 * pairs at their interval midpoints, hard Moller and Bhabha events, local
 * deposits as E / W_air pairs, demotions, annihilations, Compton,
 * photoelectric and pair events -- is deposited where it occurs with the
 * accumulator's order-1 cell-centred shape (AddToField :386-512: a radial
 * ghost below the axis folds onto the first cell; support past the other
 * boundaries is dropped) into the fields below, as raw weights.  A record
 * outside the physical domain is reported as error 2 (the reference aborts). */
#define RREA_SOURCE_FIELDS 5
enum {
    RREA_SOURCE_LOW_ELECTRON = 0,   /* low-energy electron source */
    RREA_SOURCE_POSITIVE_ION = 1,   /* positive-ion source */
    RREA_SOURCE_NEGATIVE_ION = 2,   /* direct negative-ion source (no transport record feeds it) */
    RREA_SOURCE_EVENT_COUNT = 3,    /* ionization event count (diagnostic) */
    RREA_SOURCE_ENERGY_LOSS = 4,    /* energy loss, weight x eV (diagnostic) */
};
/* NormalizeAndSync (:280-345) on one grid: every field over its RZ shell
 * cell volume pi (r_o^2 - r_i^2) dz, the three source fields also over dt;
 * with diagnostics == 0 the two diagnostic fields are left raw. */
rrea_status rrea_sources_normalize(rrea_fields_t fields, double* sources, double dt_s, int32_t diagnostics,
                                   rrea_stream_t stream);
/* Each instance's source grids for batched steps and drains (NULL: none). */
rrea_status rrea_batch_bind_sources(rrea_batch_t batch, double* const* sources);

/* Kernel timing with HIP events recorded around every transport-kernel
 * launch made through `fields` while enabled (no counterpart in the reference:
 * measurement support).  Enabling or disabling clears the record.  _read
 * waits for the recorded launches of one kernel and returns their summed
 * milliseconds and count: per species, the main kernel (the fast path, or the
 * complete path when complete_path_only is set) and the complete-path kernel
 * over the fast path's hand-over list; -1 = all. */
enum {
    RREA_KERNEL_ALL = -1,
    RREA_KERNEL_ELECTRON = 0,
    RREA_KERNEL_ELECTRON_HANDOVER = 1,
    RREA_KERNEL_POSITRON = 2,
    RREA_KERNEL_POSITRON_HANDOVER = 3,
    RREA_KERNEL_PHOTON = 4,
    RREA_KERNEL_PHOTON_HANDOVER = 5,
    RREA_KERNEL_NEWBORN = 6,
};
rrea_status rrea_kernel_timing(rrea_fields_t fields, int32_t enable);
rrea_status rrea_kernel_timing_read(rrea_fields_t fields, int32_t kernel, double* total_ms, int64_t* launches);

/* The library version string. */
const char* rrea_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RREA_HIP_H */
