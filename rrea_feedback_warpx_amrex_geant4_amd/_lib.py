"""ctypes binding of ``librrea_hip.so`` -- the C-ABI declared in
``include/rrea_hip.h``.

The structs below mirror the header field by field; a CPU test compiles the
header with the host compiler and checks every size and offset against them.
There is no fallback: if the library is missing the import of the transport
module fails with the build command to run.
"""

from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "librrea_hip.so")

P_D = C.POINTER(C.c_double)
P_I64 = C.POINTER(C.c_int64)


class RateColumn(C.Structure):
    _fields_ = [("x", P_D), ("y", P_D), ("n", C.c_int64), ("mode", C.c_int32)]


class CdfGroups(C.Structure):
    _fields_ = [("primary", P_D), ("n_groups", C.c_int64), ("start", P_I64), ("u", P_D), ("v", P_D)]


class AngleTable(C.Structure):
    _fields_ = [("e", P_D), ("n_e", C.c_int64), ("n", P_D), ("n_n", C.c_int64), ("u", P_D), ("n_u", C.c_int64),
                ("omc", P_D)]


class ElasticTable(C.Structure):
    _fields_ = [("e", P_D), ("mfp_m_stp", P_D), ("screening_eta", P_D), ("n", C.c_int64)]


class PhotoelectricTable(C.Structure):
    _fields_ = [("primary", P_D), ("n_groups", C.c_int64), ("start", P_I64), ("cdf_u", P_D), ("binding_eV", P_D)]


class PairBank(C.Structure):
    _fields_ = [("channel", P_I64), ("energy_eV", P_D), ("n_blocks", C.c_int64), ("start", P_I64),
                ("fraction", P_D), ("direction", P_D)]


class TableImage(C.Structure):
    _fields_ = [
        ("rate", RateColumn * 15),
        ("moller", CdfGroups), ("bhabha", CdfGroups), ("electron_brems", CdfGroups), ("positron_brems", CdfGroups),
        ("gs_electron", AngleTable), ("gs_positron", AngleTable),
        ("wentzel_electron", AngleTable), ("wentzel_positron", AngleTable),
        ("elastic_electron", ElasticTable), ("elastic_positron", ElasticTable),
        ("hard_moller_secondary_threshold_eV", C.c_double),
        ("brems_absolute_cutoff_eV", C.c_double),
        ("reference_density_kg_m3", C.c_double),
        ("charged_energy_range_eV", C.c_double * 2),
        ("photon_energy_range_eV", C.c_double * 2),
        ("gs_few_collision_threshold_n", C.c_double),
        ("gs_isotropic_above_path_n", C.c_double),
        ("gs_wentzel_transition_energy_eV", C.c_double),
        ("photoelectric", PhotoelectricTable),
        ("pair", PairBank),
    ]


class RzGeometry(C.Structure):
    _fields_ = [("nr", C.c_int32), ("nz", C.c_int32), ("dr", C.c_double), ("dz", C.c_double),
                ("r_lo", C.c_double), ("z_lo", C.c_double)]


class StepConfigC(C.Structure):
    _fields_ = [
        ("rng_seed", C.c_uint64), ("step", C.c_int64), ("dt_s", C.c_double),
        ("low_energy_cutoff_eV", C.c_double), ("photon_cutoff_eV", C.c_double),
        ("max_optical_depth_per_substep", C.c_double), ("max_subcycles", C.c_uint64),
        ("strict_transport_guards", C.c_int32), ("disable_hard_moller", C.c_int32),
        ("max_gs_collisions_per_substep", C.c_double), ("gs_total_collision_cap", C.c_double),
        ("max_field_impulse_fraction", C.c_double), ("field_impulse_floor_eV", C.c_double),
        ("complete_path_only", C.c_int32), ("reserved", C.c_int32),
    ]


class Tile(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in (
        "prev_r", "prev_z", "prev_theta", "prev_ux", "prev_uy", "prev_uz", "w", "id", "cpu",
        "r", "z", "theta", "ux", "uy", "uz")]


class Secondary(C.Structure):
    _fields_ = [
        ("parent_index", C.c_int64), ("species", C.c_int32), ("process", C.c_int32),
        ("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
        ("dir_x", C.c_double), ("dir_y", C.c_double), ("dir_z", C.c_double),
        ("weight", C.c_double), ("energy_eV", C.c_double), ("birth_fraction", C.c_double),
        ("root_key", C.c_uint64), ("lineage", C.c_uint64),
        ("root_species", C.c_int32), ("generation", C.c_int32), ("instance", C.c_int32), ("reserved", C.c_int32),
    ]


class ParticleStore(C.Structure):
    _fields_ = [("tile", Tile), ("instance", C.c_void_p), ("capacity", C.c_int64)]


class DriftConfig(C.Structure):
    _fields_ = [("species", C.c_int32), ("steps", C.c_int32), ("checkpoint_every", C.c_int32),
                ("reserved", C.c_int32), ("dt_s", C.c_double), ("low_energy_cutoff_eV", C.c_double)]


class StepOut(C.Structure):
    _fields_ = [
        ("status", C.c_void_p), ("ion_pairs", C.c_void_p), ("substeps", C.c_void_p),
        ("secondaries", C.c_void_p), ("secondary_capacity", C.c_int64), ("secondary_count", C.c_void_p),
        ("jr", C.c_void_p), ("jz", C.c_void_p), ("rho", C.c_void_p), ("tallies", C.c_void_p),
        ("error", C.c_void_p), ("order_secondaries", C.c_int32), ("sources", C.c_void_p),
    ]


class DrainIO(C.Structure):
    _fields_ = [
        ("work", C.c_void_p), ("work_capacity", C.c_int64), ("n_initial", C.c_int64), ("work_count", C.c_void_p),
        ("status", C.c_void_p), ("ion_pairs", C.c_void_p), ("substeps", C.c_void_p), ("survivors", C.c_void_p),
        ("survivor_capacity", C.c_int64), ("survivor_count", C.c_void_p),
        ("jr", C.c_void_p), ("jz", C.c_void_p), ("rho", C.c_void_p), ("tallies", C.c_void_p),
        ("error", C.c_void_p), ("sources", C.c_void_p),
    ]


class BatchGraphDesc(C.Structure):
    _fields_ = [
        ("tiles", C.c_void_p * 3), ("instance", C.c_void_p * 3), ("config", StepConfigC),
        ("records", C.c_void_p), ("record_capacity", C.c_int64), ("record_count", C.c_void_p),
        ("survivors", C.c_void_p), ("survivor_capacity", C.c_int64), ("survivor_count", C.c_void_p),
        ("tallies", C.c_void_p), ("error", C.c_void_p), ("clear", C.c_void_p * 4), ("clear_n", C.c_int64 * 4),
        ("max_generations", C.c_int32), ("steps", C.c_int32), ("push_photons", C.c_int32), ("reserved", C.c_int32),
        ("counts", C.c_void_p), ("next_id", C.c_void_p), ("cpu", C.c_int32), ("reserved2", C.c_int32),
    ]


class GroupEntry(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("weight", C.c_double), ("energy_eV", C.c_double)]


class SecondaryGroup(C.Structure):
    _fields_ = [("process", C.c_int32), ("parent_species", C.c_int32), ("remove_parent", C.c_int32),
                ("parent_still_valid", C.c_int32), ("closure_mode", C.c_int32), ("parent_survives", C.c_int32),
                ("parent_weight", C.c_double), ("parent_energy_before_eV", C.c_double),
                ("parent_energy_after_eV", C.c_double), ("local_energy_eV", C.c_double),
                ("escaped_energy_eV", C.c_double), ("escaped_charge_e", C.c_double),
                ("n_particles", C.c_int32), ("n_sources", C.c_int32),
                ("particles", GroupEntry * 4), ("sources", GroupEntry * 4)]


class PopulationRecords(C.Structure):
    _fields_ = [("n", C.c_int64)] + [(k, C.c_void_p) for k in (
        "r", "z", "theta", "ux", "uy", "uz", "w", "stable_id", "stable_id_high", "candidate", "id")]


class PopulationConfig(C.Structure):
    _fields_ = [("z_lo", C.c_double), ("dr", C.c_double), ("dz", C.c_double), ("target", C.c_int64),
                ("rng_seed", C.c_uint64), ("rng_salt", C.c_uint64), ("step", C.c_int64), ("species", C.c_int32),
                ("reserved", C.c_int32)]


TALLY_COUNT = 24 + 75  # RREA_TALLY_COUNT: the counts, then the event ledger
LEDGER_OFFSET = 24
SOURCE_FIELDS = 5
STATUS = {0: "survive", 1: "escape", 2: "demoted", 3: "zero_weight", 4: "no_writeback", 5: "skipped_invalid",
          6: "annihilated", 7: "absorbed"}
INVALID_ID = -(2**63)
ERRORS = {0: "ok", 1: "table range", 2: "outside domain", 3: "strict guard", 4: "invalid", 5: "secondary capacity",
          16: "argument", 17: "hip"}

# (symbol, restype, argtypes) for every entry point of include/rrea_hip.h
SIGNATURES = [
    ("rrea_tables_create", C.c_int32, [C.POINTER(TableImage), C.POINTER(C.c_void_p)]),
    ("rrea_tables_destroy", C.c_int32, [C.c_void_p]),
    ("rrea_tables_hm_mean", C.c_int32, [C.c_void_p, P_D, C.c_int64]),
    ("rrea_fields_create", C.c_int32, [C.POINTER(RzGeometry), C.POINTER(C.c_void_p)]),
    ("rrea_fields_destroy", C.c_int32, [C.c_void_p]),
    ("rrea_fields_update", C.c_int32, [C.c_void_p, P_D, P_D, P_D, P_D, P_D, C.c_int32, C.c_void_p]),
    ("rrea_density_set_constant", C.c_int32, [C.c_void_p, C.c_double]),
    ("rrea_density_set_profile", C.c_int32, [C.c_void_p, P_D, P_D, C.c_int64, C.c_double, C.c_double]),
    ("rrea_transport_step_electrons", C.c_int32,
     [C.c_void_p, C.c_void_p, C.POINTER(Tile), C.POINTER(StepConfigC), C.POINTER(StepOut), C.c_void_p]),
    ("rrea_transport_step_positrons", C.c_int32,
     [C.c_void_p, C.c_void_p, C.POINTER(Tile), C.POINTER(StepConfigC), C.POINTER(StepOut), C.c_void_p]),
    ("rrea_transport_step_photons", C.c_int32,
     [C.c_void_p, C.c_void_p, C.POINTER(Tile), C.POINTER(StepConfigC), C.POINTER(StepOut), C.c_void_p]),
    ("rrea_transport_step_noalloc", C.c_int32,
     [C.c_int32, C.c_void_p, C.c_void_p, C.POINTER(Tile), C.POINTER(StepConfigC), C.POINTER(StepOut), C.c_void_p]),
    ("rrea_validate_secondary_groups", C.c_int32,
     [C.c_void_p, C.c_int64, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rrea_photon_push", C.c_int32, [C.POINTER(Tile), C.c_double, C.c_void_p]),
    ("rrea_tile_sort", C.c_int32, [C.c_void_p, C.POINTER(Tile), C.c_int32, C.c_void_p]),
    ("rrea_tile_sort_banded", C.c_int32, [C.c_void_p, C.POINTER(Tile), C.c_int32, C.c_int32, C.c_void_p]),
    ("rrea_batch_create", C.c_int32, [C.POINTER(C.c_void_p), C.c_int32, C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                      C.POINTER(C.c_void_p), C.POINTER(C.c_uint64), C.POINTER(C.c_void_p)]),
    ("rrea_batch_destroy", C.c_int32, [C.c_void_p]),
    ("rrea_batch_tile_sort", C.c_int32, [C.c_void_p, C.POINTER(Tile), C.c_int32, C.c_void_p, C.c_int32, C.c_void_p]),
    ("rrea_batch_workspace", C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("rrea_transport_step_batched", C.c_int32,
     [C.c_void_p, C.c_void_p, C.c_int32, C.POINTER(Tile), C.c_void_p, C.POINTER(StepConfigC), C.POINTER(StepOut),
      C.c_void_p]),
    ("rrea_newborn_drain_batched", C.c_int32,
     [C.c_void_p, C.c_void_p, C.POINTER(StepConfigC), C.POINTER(DrainIO), C.c_void_p]),
    ("rrea_newborn_drain", C.c_int32,
     [C.c_void_p, C.c_void_p, C.POINTER(StepConfigC), C.POINTER(DrainIO), C.c_void_p]),
    ("rrea_commit_survivors", C.c_int32, [C.c_void_p, C.c_void_p, C.POINTER(ParticleStore), C.c_void_p, C.c_void_p,
                                          C.c_int32, C.c_void_p, C.c_void_p]),
    ("rrea_tile_compact", C.c_int32, [C.c_void_p, C.POINTER(Tile), C.c_int32, C.c_void_p, P_I64, C.c_void_p]),
    ("rrea_drift_forward", C.c_int32,
     [C.c_void_p, C.c_void_p, C.POINTER(Tile), C.POINTER(DriftConfig)] + [C.c_void_p] * 9),
    ("rrea_drift_adjoint", C.c_int32,
     [C.c_void_p, C.c_void_p, C.POINTER(Tile), C.POINTER(DriftConfig)] + [C.c_void_p] * 7),
    ("rrea_sources_normalize", C.c_int32, [C.c_void_p, C.c_void_p, C.c_double, C.c_int32, C.c_void_p]),
    ("rrea_batch_bind_sources", C.c_int32, [C.c_void_p, C.POINTER(C.c_void_p)]),
    ("rrea_batch_graph_create", C.c_int32,
     [C.c_void_p, C.c_void_p, C.POINTER(BatchGraphDesc), C.POINTER(C.c_void_p)]),
    ("rrea_batch_graph_launch", C.c_int32, [C.c_void_p, C.c_int64, C.c_void_p]),
    ("rrea_batch_graph_destroy", C.c_int32, [C.c_void_p]),
    ("rrea_current_to_physical", C.c_int32,
     [C.POINTER(RzGeometry), C.c_int32, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]),
    ("rrea_maxwell_advance_b", C.c_int32,
     [C.POINTER(RzGeometry), C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]),
    ("rrea_maxwell_silver_mueller", C.c_int32, [C.POINTER(RzGeometry)] + [C.c_void_p] * 4),
    ("rrea_maxwell_advance_e", C.c_int32,
     [C.POINTER(RzGeometry), C.c_void_p, C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 7),
    ("rrea_maxwell_step", C.c_int32,
     [C.POINTER(RzGeometry), C.c_void_p, C.c_void_p, C.c_void_p, C.c_double] + [C.c_void_p] * 7),
    ("rrea_maxwell_ledger", C.c_int32, [C.POINTER(RzGeometry)] + [C.c_void_p] * 9),
    ("rrea_population_thin", C.c_int32,
     [C.POINTER(PopulationRecords), C.POINTER(PopulationConfig), P_I64, C.c_void_p, C.c_void_p]),
    ("rrea_newborn_drain_staged", C.c_int32,
     [C.c_void_p, C.c_void_p, C.POINTER(StepConfigC), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), P_I64,
      C.c_int32, C.POINTER(DrainIO), C.c_int32, C.c_void_p]),
    ("rrea_fields_signal_fast_path", C.c_int32, [C.c_void_p, C.c_int32]),
    ("rrea_stream_wait_fast_path", C.c_int32, [C.c_void_p, C.c_void_p]),
    ("rrea_fields_set_block", C.c_int32, [C.c_void_p, C.c_int32]),
    ("rrea_kernel_timing", C.c_int32, [C.c_void_p, C.c_int32]),
    ("rrea_kernel_timing_read", C.c_int32, [C.c_void_p, C.c_int32, P_D, P_I64]),
    ("rrea_version", C.c_char_p, []),
]


class RreaHipError(RuntimeError):
    pass


_lib = None


def lib():
    """Load the in-tree library; raise if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RreaHipError(
                f"{LIB_PATH} is missing; build it with "
                "`python -m rrea_feedback_warpx_amrex_geant4_amd.build` (no CPU fallback exists)")
        L = C.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(code: int, what: str) -> None:
    if code != 0:
        raise RreaHipError(f"{what} failed: {ERRORS.get(code, code)} ({code})")
