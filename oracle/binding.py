"""ctypes binding of the CPU oracle (``oracle/build/liboracle.so``) and of
the reference-compiled pins (``oracle/_ref/libref.so``) -- test infrastructure.
"""

from __future__ import annotations

import ctypes as C
import os
import re
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# RREA_ORACLE_LIB names another build of the oracle (the planted-mistake
# checks load deliberately broken copies through it)
LIB_PATH = os.environ.get("RREA_ORACLE_LIB") or os.path.join(HERE, "build", "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libref.so")

P_D = C.POINTER(C.c_double)
P_I64 = C.POINTER(C.c_int64)
P_I32 = C.POINTER(C.c_int32)


class Geometry(C.Structure):
    _fields_ = [
        ("nr", C.c_int32), ("nz", C.c_int32), ("dr", C.c_double), ("dz", C.c_double),
        ("r_lo", C.c_double), ("z_lo", C.c_double),
        ("er", P_D), ("ez", P_D), ("amb_er", P_D), ("amb_ez", P_D), ("btheta", P_D),
        ("density_mode", C.c_int32), ("profile_n", C.c_int32), ("density_ratio", C.c_double),
        ("profile_alt_m", P_D), ("profile_rho", P_D), ("altitude_at_z0_m", C.c_double),
        ("sources", P_D),
    ]


class Config(C.Structure):
    _fields_ = [
        ("rng_seed", C.c_uint64), ("step", C.c_int64), ("dt_s", C.c_double),
        ("low_energy_cutoff_eV", C.c_double), ("photon_cutoff_eV", C.c_double),
        ("max_optical_depth_per_substep", C.c_double), ("max_subcycles", C.c_uint64),
        ("strict_transport_guards", C.c_int32), ("disable_hard_moller", C.c_int32),
        ("max_gs_collisions_per_substep", C.c_double), ("gs_total_collision_cap", C.c_double),
        ("max_field_impulse_fraction", C.c_double), ("field_impulse_floor_eV", C.c_double),
    ]


class Secondary(C.Structure):
    _fields_ = [
        ("parent_index", C.c_int64), ("species", C.c_int32), ("process", C.c_int32),
        ("x", C.c_double), ("y", C.c_double), ("z", C.c_double),
        ("dir_x", C.c_double), ("dir_y", C.c_double), ("dir_z", C.c_double),
        ("weight", C.c_double), ("energy_eV", C.c_double), ("birth_fraction", C.c_double),
        ("root_key", C.c_uint64), ("lineage", C.c_uint64),
        ("root_species", C.c_int32), ("generation", C.c_int32), ("instance", C.c_int32), ("reserved", C.c_int32),
    ]


SECONDARY_DTYPE = np.dtype([
    ("parent_index", "<i8"), ("species", "<i4"), ("process", "<i4"),
    ("x", "<f8"), ("y", "<f8"), ("z", "<f8"),
    ("dir_x", "<f8"), ("dir_y", "<f8"), ("dir_z", "<f8"),
    ("weight", "<f8"), ("energy_eV", "<f8"), ("birth_fraction", "<f8"),
    ("root_key", "<u8"), ("lineage", "<u8"), ("root_species", "<i4"), ("generation", "<i4"),
    ("instance", "<i4"), ("reserved", "<i4"),
])
assert SECONDARY_DTYPE.itemsize == C.sizeof(Secondary)

RATE_KEYS = ["e_col", "e_soft", "e_moller", "e_brems", "e_eltr",
             "p_col", "p_soft", "p_bhabha", "p_brems", "p_eltr", "p_annih",
             "g_compton", "g_pe", "g_pairn", "g_pairt"]
CDF_KEYS = ["moller", "bhabha", "ebrems", "pbrems"]
ANGLE_KEYS = ["gs_e", "gs_p", "wz_e", "wz_p"]
N_TALLIES = 24 + 75
LEDGER_OFFSET = 24


def _ledger_names():
    """The event-ledger slot names, in order, from the oracle header."""
    txt = open(os.path.join(HERE, "rrea_oracle.h")).read()
    blk = txt[txt.index("ORC_L_PHOTON_INTERACTION_COUNT = 0"):txt.index("    ORC_LEDGER_COUNT")]
    return [n.lower() for n in re.findall(r"ORC_L_([A-Z0-9_]+)", blk)]


LEDGER_NAMES = _ledger_names()
LEDGER_MAX = {n for n in LEDGER_NAMES if n.startswith("max_")} | {"hard_moller_secondary_energy_max_ev"}
assert len(LEDGER_NAMES) == N_TALLIES - LEDGER_OFFSET


def ledger(tallies) -> dict:
    """The event ledger of a tally array, by field name."""
    return {n: float(tallies[LEDGER_OFFSET + i]) for i, n in enumerate(LEDGER_NAMES)}


class InterleaveRecord(C.Structure):
    """orc_interleave_record: what the event interleave handed its policies."""
    _fields_ = [("n_frozen", C.c_int32), ("n_segments", C.c_int32), ("n_draws", C.c_int32), ("n_open", C.c_int32),
                ("n_final", C.c_int32), ("n_stops", C.c_int32),
                ("frozen", C.c_double * 16), ("taus", C.c_double * 16),
                ("seg_begin", C.c_double * 16), ("seg_end", C.c_double * 16), ("seg_frozen", C.c_double * 16),
                ("draw_ordinals", C.c_uint64 * 16), ("open_energies", C.c_double * 16),
                ("final_fractions", C.c_double * 16), ("stops", C.c_double * 16), ("kinetic_end", C.c_double)]


class GroupEntry(C.Structure):
    """orc_group_entry: a particle (kind = species) or a material source."""
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("weight", C.c_double), ("energy_eV", C.c_double)]


class Group(C.Structure):
    """orc_group: one interaction group and its conservation closure."""
    _fields_ = [("process", C.c_int32), ("parent_species", C.c_int32), ("remove_parent", C.c_int32),
                ("parent_still_valid", C.c_int32), ("closure_mode", C.c_int32), ("parent_survives", C.c_int32),
                ("parent_weight", C.c_double), ("parent_energy_before_eV", C.c_double),
                ("parent_energy_after_eV", C.c_double), ("local_energy_eV", C.c_double),
                ("escaped_energy_eV", C.c_double), ("escaped_charge_e", C.c_double),
                ("n_particles", C.c_int32), ("n_sources", C.c_int32),
                ("particles", GroupEntry * 4), ("sources", GroupEntry * 4)]


def validate_group(g: Group, low_energy_cutoff_eV: float):
    """orc_validate_group: (result bits 1 required, 2 present, 4 finite,
    8 valid; charge residual; energy residual)."""
    res = np.zeros(2)
    bits = lib().orc_validate_group(C.byref(g), low_energy_cutoff_eV, _ptr(res))
    return int(bits), float(res[0]), float(res[1])


def fold_tallies(into: np.ndarray, frm: np.ndarray) -> np.ndarray:
    """orc_fold_tallies on copies: the fold of two tally arrays."""
    a = np.ascontiguousarray(into, dtype=np.float64).copy()
    b = np.ascontiguousarray(frm, dtype=np.float64)
    assert a.shape == b.shape == (N_TALLIES,)
    lib().orc_fold_tallies(_ptr(a), _ptr(b))
    return a


def build(quiet: bool = True) -> None:
    """Compile the oracle (and the reference pins when the checkout exists)."""
    out = subprocess.DEVNULL if quiet else None
    subprocess.check_call(["make", "-C", HERE, "all"], stdout=out)
    if os.path.isdir("/root/reference/native_amrex_rrea"):
        subprocess.check_call(["make", "-C", HERE, "ref"], stdout=out)


def _ptr(a: np.ndarray, ctype=P_D):
    return a.ctypes.data_as(ctype)


_lib = None
_ref = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        L.orc_rng_counter.restype = C.c_uint64
        L.orc_rng_counter.argtypes = [C.c_uint64] * 5
        L.orc_rng_uniform.restype = C.c_double
        L.orc_rng_uniform.argtypes = [C.c_uint64] * 5
        L.orc_uniform_stream.argtypes = [C.c_double, C.c_double, C.c_int64, P_D]
        L.orc_tables_create.restype = C.c_void_p
        L.orc_tables_destroy.argtypes = [C.c_void_p]
        L.orc_tables_set_rate.argtypes = [C.c_void_p, C.c_int, P_D, P_D, C.c_int64, C.c_int]
        L.orc_tables_set_cdf.argtypes = [C.c_void_p, C.c_int, P_D, C.c_int64, P_I64, P_D, P_D]
        L.orc_tables_set_angle.argtypes = [C.c_void_p, C.c_int, P_D, C.c_int64, P_D, C.c_int64, P_D, C.c_int64, P_D]
        L.orc_tables_set_elastic.argtypes = [C.c_void_p, C.c_int, P_D, P_D, P_D, P_D, C.c_int64]
        L.orc_tables_finalize.argtypes = [C.c_void_p, P_D]
        L.orc_tables_finalize.restype = C.c_int
        L.orc_tables_hm_mean.argtypes = [C.c_void_p]
        L.orc_tables_hm_mean.restype = P_D
        for name, extra in [("orc_interpolate", [C.c_int, C.c_double]),
                            ("orc_sample_moller", [C.c_double, C.c_double]),
                            ("orc_sample_brems_energy", [C.c_double, C.c_double]),
                            ("orc_sample_brems_cos", [C.c_double, C.c_double]),
                            ("orc_sample_gs_cos", [C.c_int, C.c_double, C.c_double, C.c_double]),
                            ("orc_sample_single_elastic_cos", [C.c_int, C.c_double, C.c_double]),
                            ("orc_elastic_mfp", [C.c_int, C.c_double]),
                            ("orc_ensemble_drag", [C.c_double, C.c_double]),
                            ("orc_hard_moller_energy_loss", [C.c_double])]:
            f = getattr(L, name)
            f.restype = C.c_double
            f.argtypes = [C.c_void_p] + extra + [P_I32]
        L.orc_clip_path.argtypes = [C.c_double] * 10 + [P_D]
        L.orc_clip_path.restype = C.c_int
        L.orc_gather.argtypes = [C.POINTER(Geometry), C.c_double, C.c_double, P_D]
        L.orc_deposit_segment.argtypes = [C.c_int, C.c_int] + [C.c_double] * 4 + [P_D] * 3 + [C.c_double] * 6
        L.orc_deposit_escaping_segment.argtypes = ([C.c_int, C.c_int] + [C.c_double] * 4 + [P_D] * 3
                                                   + [C.c_double] * 2 + [P_D, C.c_int, C.c_int, C.c_int, C.c_uint32])
        L.orc_deposit_escaping_segment.restype = C.c_int
        for name in ("orc_electron_step", "orc_positron_step"):
            f = getattr(L, name)
            f.restype = C.c_int
            f.argtypes = ([C.c_void_p, C.POINTER(Geometry), C.POINTER(Config), C.c_int64]
                          + [P_D] * 7 + [P_I64, P_I32] + [P_D] * 6 + [P_I32, P_D, P_I64]
                          + [C.c_int64, C.POINTER(Secondary), P_I64] + [P_D] * 4 + [C.c_int])
        L.orc_sample_annihilation.argtypes = [C.c_double, C.c_double, C.c_double, P_D]
        L.orc_photon_step.restype = C.c_int
        L.orc_photon_step.argtypes = ([C.c_void_p, C.POINTER(Geometry), C.POINTER(Config), C.c_int64]
                                      + [P_D] * 10 + [P_I64, P_I32] + [P_D] * 6 + [P_I32, P_D, P_I64]
                                      + [C.c_int64, C.POINTER(Secondary), P_I64, P_D, C.c_int])
        L.orc_newborn_drain.restype = C.c_int
        L.orc_accumulate_source.restype = C.c_int
        L.orc_accumulate_source.argtypes = [C.POINTER(Geometry), C.c_double, C.c_int] + [C.c_double] * 5
        L.orc_sources_normalize.restype = C.c_int
        L.orc_sources_normalize.argtypes = [C.POINTER(Geometry), C.c_double, C.c_int]
        L.orc_newborn_drain.argtypes = [C.c_void_p, C.POINTER(Geometry), C.POINTER(Config), C.POINTER(Secondary),
                                        C.c_int64, C.c_int64, C.POINTER(Secondary), P_I64, P_I32, P_D, P_I64,
                                        C.POINTER(Secondary), P_I64, P_D, P_D, P_D, P_D]
        L.orc_pack_transient_index.argtypes = [C.c_int] + [C.c_uint64] * 5 + [C.POINTER(C.c_uint64)]
        L.orc_pack_transient_index.restype = C.c_int
        L.orc_live_hazard_index.argtypes = [C.c_uint64] * 3 + [C.POINTER(C.c_uint64)]
        L.orc_live_hazard_index.restype = C.c_int
        L.orc_electron_state_demotes.argtypes = [C.c_double] * 5
        L.orc_electron_state_demotes.restype = C.c_int
        L.orc_newborn_elapsed.argtypes = [C.c_double, C.c_uint64, C.c_uint64, C.c_double]
        L.orc_newborn_elapsed.restype = C.c_double
        L.orc_tables_set_photoelectric.argtypes = [C.c_void_p, P_D, C.c_int64, P_I64, P_D, P_D]
        L.orc_tables_set_pair.argtypes = [C.c_void_p, P_I64, P_D, C.c_int64, P_I64, P_D, P_D]
        L.orc_sample_compton.argtypes = [C.c_void_p, C.c_double, C.c_double, C.c_double, P_D]
        L.orc_sample_compton.restype = C.c_int
        L.orc_sample_photoelectric.argtypes = [C.c_void_p, C.c_double, C.c_double, P_D]
        L.orc_sample_photoelectric.restype = C.c_int
        L.orc_fold_tallies.argtypes = [P_D, P_D]
        L.orc_fold_tallies.restype = None
        L.orc_validate_group.argtypes = [C.POINTER(Group), C.c_double, P_D]
        L.orc_validate_group.restype = C.c_int
        L.orc_sample_photoelectron_cos.argtypes = [C.c_void_p, C.c_double, C.c_double, P_I32]
        L.orc_sample_photoelectron_cos.restype = C.c_double
        L.orc_sample_photon_channel.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_sample_photon_channel.restype = C.c_int
        L.orc_sample_pair.argtypes = [C.c_void_p, C.c_double, C.c_int, C.c_double, C.c_double, C.c_double, P_D]
        L.orc_sample_pair.restype = C.c_int
        for name in ("orc_sample_bhabha", "orc_sample_positron_brems_energy"):
            f = getattr(L, name)
            f.restype = C.c_double
            f.argtypes = [C.c_void_p, C.c_double, C.c_double, P_I32]
        # probes of the step's primitives
        P_U64 = C.POINTER(C.c_uint64)
        L.orc_unit_cycle.restype = C.c_double
        L.orc_unit_cycle.argtypes = [C.c_double]
        L.orc_normalize.argtypes = [P_D]
        L.orc_boris_rotate.argtypes = [P_D, C.c_double, P_D]
        L.orc_cone_direction.argtypes = [P_D, C.c_double, C.c_double, P_D]
        L.orc_moller_cos.restype = C.c_double
        L.orc_moller_cos.argtypes = [C.c_double, C.c_double, C.c_int]
        L.orc_moller_lab_directions.argtypes = [P_D, C.c_double, C.c_double, C.c_double, P_D, P_D]
        L.orc_photon_recoil.argtypes = [C.c_double, C.c_double, P_D, P_D]
        L.orc_rotate_to_axis.argtypes = [P_D, P_D, P_D]
        L.orc_poisson_capped.restype = C.c_int
        L.orc_poisson_capped.argtypes = [C.c_double, C.c_double, C.c_int]
        L.orc_select_hazard.argtypes = [P_D, C.c_int, C.c_double, C.c_double, P_D]
        L.orc_half_kick.restype = C.c_double
        L.orc_half_kick.argtypes = [P_D, C.c_double, C.c_double, P_D, P_D]
        L.orc_gated_inverse_length.restype = C.c_double
        L.orc_gated_inverse_length.argtypes = [C.c_int, C.c_double, P_I32]
        L.orc_stop_is_escape.restype = C.c_int
        L.orc_stop_is_escape.argtypes = [C.c_int, C.c_double, C.c_double]
        L.orc_limit_at_cutoff_linear.restype = C.c_int
        L.orc_limit_at_cutoff_linear.argtypes = [C.c_double] * 5 + [P_D]
        L.orc_partition_ion_pairs.restype = C.c_int
        L.orc_partition_ion_pairs.argtypes = [P_D, P_D, P_D, C.c_int] + [C.c_double] * 8 + [P_D, P_D, C.c_int]
        L.orc_interleave_probe.restype = C.c_int
        L.orc_interleave_probe.argtypes = [C.c_double] * 3 + [P_D, C.c_int, P_D, C.c_int, C.c_int,
                                                              C.POINTER(InterleaveRecord)]
        L.orc_substep_probe.restype = C.c_int
        L.orc_substep_probe.argtypes = [C.c_uint64, C.c_double, C.c_int, C.c_uint64, C.c_uint64, C.c_int, P_I32,
                                        C.c_int, P_U64, P_I32]
        L.orc_boundary_crossing.restype = C.c_int
        L.orc_boundary_crossing.argtypes = [C.c_double] * 10 + [P_D]
        L.orc_advance_running_position.restype = C.c_int
        L.orc_advance_running_position.argtypes = [C.c_double] * 11 + [P_D]
        L.orc_outside_certified.restype = C.c_int
        L.orc_outside_certified.argtypes = [C.c_double] * 6
        _lib = L
    return _lib


def ref_available() -> bool:
    """Whether the reference-compiled pins exist or can be built here,
    without loading them."""
    return os.path.exists(REF_PATH) or os.path.isdir("/root/reference/native_amrex_rrea")


def ref():
    """The reference-compiled pins, or None where the reference is absent."""
    global _ref
    if _ref is None:
        if not os.path.exists(REF_PATH):
            if not os.path.isdir("/root/reference/native_amrex_rrea"):
                return None
            subprocess.check_call(["make", "-C", HERE, "ref"], stdout=subprocess.DEVNULL)
        R = C.CDLL(REF_PATH)
        R.ref_rng_counter.restype = C.c_uint64
        R.ref_rng_counter.argtypes = [C.c_uint64] * 5
        R.ref_rng_uniform.restype = C.c_double
        R.ref_rng_uniform.argtypes = [C.c_uint64] * 5
        R.ref_uniform_stream.argtypes = [C.c_double, C.c_double, C.c_int64, P_D]
        R.ref_gl2_cubic.restype = C.c_double
        R.ref_gl2_cubic.argtypes = [C.c_double, C.c_double, P_D]
        R.ref_deposit_segment.argtypes = [C.c_int, C.c_int] + [C.c_double] * 4 + [P_D] * 3 + [C.c_double] * 6
        R.ref_deposit_escaping_segment.argtypes = ([C.c_int, C.c_int] + [C.c_double] * 4 + [P_D] * 3
                                                   + [C.c_double] * 2 + [P_D, C.c_int, C.c_int, C.c_int, C.c_uint32])
        R.ref_deposit_escaping_segment.restype = C.c_int
        _ref = R
    return _ref


class OracleError(RuntimeError):
    pass


class Tables:
    """The oracle's own schema-6 table set, filled from the bundle arrays."""

    def __init__(self, bundle: dict):
        L = lib()
        self._keep = []
        self.h = L.orc_tables_create()
        for i, key in enumerate(RATE_KEYS):
            x = np.ascontiguousarray(bundle[key + "_x"], dtype=np.float64)
            y = np.ascontiguousarray(bundle[key + "_y"], dtype=np.float64)
            L.orc_tables_set_rate(self.h, i, _ptr(x), _ptr(y), len(x), int(bundle[key + "_mode"]))
        for i, key in enumerate(CDF_KEYS):
            prim = np.ascontiguousarray(bundle[key + "_primary"])
            start = np.ascontiguousarray(bundle[key + "_start"], dtype=np.int64)
            u = np.ascontiguousarray(bundle[key + "_u"])
            v = np.ascontiguousarray(bundle[key + "_v"])
            L.orc_tables_set_cdf(self.h, i, _ptr(prim), len(prim), _ptr(start, P_I64), _ptr(u), _ptr(v))
        for i, key in enumerate(ANGLE_KEYS):
            e = np.ascontiguousarray(bundle[key + "_e"])
            n = np.ascontiguousarray(bundle[key + "_n"])
            u = np.ascontiguousarray(bundle[key + "_u"])
            omc = np.ascontiguousarray(bundle[key + "_omc"])
            L.orc_tables_set_angle(self.h, i, _ptr(e), len(e), _ptr(n), len(n), _ptr(u), len(u), _ptr(omc))
        for i, key in enumerate(["el_e", "el_p"]):
            arrs = [np.ascontiguousarray(bundle[key + s]) for s in ("_e", "_mfp", "_eta", "_trmfp")]
            L.orc_tables_set_elastic(self.h, i, *[_ptr(a) for a in arrs], len(arrs[0]))
        pe = [np.ascontiguousarray(bundle[k]) for k in ("pe_primary", "pe_start", "pe_u", "pe_binding")]
        self._keep += pe
        L.orc_tables_set_photoelectric(self.h, _ptr(pe[0]), len(pe[0]), _ptr(pe[1], P_I64), _ptr(pe[2]), _ptr(pe[3]))
        pb = [np.ascontiguousarray(bundle[k]) for k in ("pair_block_channel", "pair_block_energy", "pair_block_start",
                                                        "pair_frac", "pair_dir")]
        self._keep += pb
        L.orc_tables_set_pair(self.h, _ptr(pb[0], P_I64), _ptr(pb[1]), len(pb[1]), _ptr(pb[2], P_I64), _ptr(pb[3]),
                              _ptr(pb[4]))
        scalars = np.ascontiguousarray(bundle["scalars"], dtype=np.float64)
        code = L.orc_tables_finalize(self.h, _ptr(scalars))
        if code:
            raise OracleError(f"oracle table finalize failed: {code}")
        self.scalars = scalars
        self.n_moller = len(bundle["e_moller_x"])

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_tables_destroy(self.h)
            self.h = None

    def hm_mean(self) -> np.ndarray:
        p = lib().orc_tables_hm_mean(self.h)
        return np.ctypeslib.as_array(p, shape=(self.n_moller,)).copy()

    def _call(self, name, *args):
        err = C.c_int32(0)
        v = getattr(lib(), name)(self.h, *args, C.byref(err))
        if err.value:
            raise OracleError(f"{name} failed with code {err.value}")
        return v

    def interpolate(self, table: str, e: float) -> float:
        return self._call("orc_interpolate", RATE_KEYS.index(table), e)

    def sample_moller(self, e, u):
        return self._call("orc_sample_moller", e, u)

    def sample_brems_energy(self, e, u):
        return self._call("orc_sample_brems_energy", e, u)

    def sample_bhabha(self, e, u):
        return self._call("orc_sample_bhabha", e, u)

    def sample_positron_brems_energy(self, e, u):
        return self._call("orc_sample_positron_brems_energy", e, u)

    def sample_compton(self, e, u1, u2):
        out = np.zeros(3)
        code = lib().orc_sample_compton(self.h, e, u1, u2, _ptr(out))
        if code:
            raise OracleError(f"orc_sample_compton failed with code {code}")
        return {"scattered": out[0], "recoil": out[1], "cos": out[2]}

    def sample_photoelectric(self, e, u):
        out = np.zeros(2)
        code = lib().orc_sample_photoelectric(self.h, e, u, _ptr(out))
        if code:
            raise OracleError(f"orc_sample_photoelectric failed with code {code}")
        return {"binding": out[0], "electron": out[1]}

    def sample_photoelectron_cos(self, ke, u):
        return self._call("orc_sample_photoelectron_cos", ke, u)

    def sample_photon_channel(self, e, u):
        v = lib().orc_sample_photon_channel(self.h, e, u)
        if v < 0:
            raise OracleError(f"orc_sample_photon_channel failed with code {-v}")
        return v

    def sample_pair(self, e, channel, u_inv, u_block, u_az):
        out = np.zeros(14)
        code = lib().orc_sample_pair(self.h, e, channel, u_inv, u_block, u_az, _ptr(out))
        if code:
            raise OracleError(f"orc_sample_pair failed with code {code}")
        return {"valid": bool(out[0]), "electron": out[1], "positron": out[2], "recoil": out[3], "local": out[4],
                "e_dir": out[5:8].copy(), "p_dir": out[8:11].copy(), "r_dir": out[11:14].copy()}

    def sample_brems_cos(self, e, u):
        return self._call("orc_sample_brems_cos", e, u)

    def sample_gs_cos(self, positron, e, n, u):
        return self._call("orc_sample_gs_cos", int(positron), e, n, u)

    def sample_single_elastic_cos(self, positron, e, u):
        return self._call("orc_sample_single_elastic_cos", int(positron), e, u)

    def elastic_mfp(self, positron, e):
        return self._call("orc_elastic_mfp", int(positron), e)

    def ensemble_drag(self, e, ratio):
        return self._call("orc_ensemble_drag", e, ratio)

    def hard_moller_energy_loss(self, e):
        return self._call("orc_hard_moller_energy_loss", e)


def make_geometry(fields, sources: np.ndarray | None = None) -> tuple[Geometry, list]:
    """Pack a transport field/density description into the oracle struct;
    ``sources`` (float64, 5 * nr * nz) receives the material sources."""
    keep = []

    def arr(a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        keep.append(a)
        return _ptr(a)

    g = Geometry()
    g.nr, g.nz = fields.nr, fields.nz
    g.dr, g.dz = fields.dr, fields.dz
    g.r_lo, g.z_lo = 0.0, fields.z_lo
    g.er, g.ez = arr(fields.er), arr(fields.ez)
    g.amb_er, g.amb_ez = arr(fields.amb_er), arr(fields.amb_ez)
    g.btheta = arr(fields.btheta)
    g.density_mode = 1 if fields.profile_alt_m is not None else 0
    g.density_ratio = fields.density_ratio
    if fields.profile_alt_m is not None:
        g.profile_n = len(fields.profile_alt_m)
        g.profile_alt_m = arr(fields.profile_alt_m)
        g.profile_rho = arr(fields.profile_rho)
        g.altitude_at_z0_m = fields.altitude_at_z0_m
    if sources is not None:
        assert sources.dtype == np.float64 and sources.flags.c_contiguous and sources.size == 5 * fields.nr * fields.nz
        g.sources = _ptr(sources)
    return g, keep


def new_sources(fields) -> np.ndarray:
    """Zeroed material-source grids (5 * nr * nz)."""
    return np.zeros(5 * fields.nr * fields.nz)


def accumulate_source(fields, sources: np.ndarray, cutoff: float, kind: int, r: float, z: float, weight: float,
                      secondary: float = 0.0, loss: float = 0.0) -> int:
    """One record (0 ionization, 1 positive ion, 2 low-energy electron) into ``sources``."""
    g, keep = make_geometry(fields, sources)
    code = lib().orc_accumulate_source(C.byref(g), cutoff, kind, r, z, weight, secondary, loss)
    del keep
    return code


def sources_normalize(fields, sources: np.ndarray, dt: float, diagnostics: bool = True) -> int:
    g, keep = make_geometry(fields, sources)
    code = lib().orc_sources_normalize(C.byref(g), dt, int(diagnostics))
    del keep
    return code


def make_config(cfg) -> Config:
    c = Config()
    c.rng_seed = cfg.rng_seed
    c.step = cfg.step
    c.dt_s = cfg.dt_s
    c.low_energy_cutoff_eV = cfg.low_energy_cutoff_eV
    c.photon_cutoff_eV = cfg.photon_cutoff_eV
    c.max_optical_depth_per_substep = cfg.max_optical_depth_per_substep
    c.max_subcycles = cfg.max_subcycles
    c.strict_transport_guards = int(cfg.strict_transport_guards)
    c.disable_hard_moller = int(cfg.debug_disable_hard_moller_secondaries)
    c.max_gs_collisions_per_substep = cfg.transport_max_gs_collisions_per_substep
    c.gs_total_collision_cap = cfg.transport_gs_total_collision_cap
    c.max_field_impulse_fraction = cfg.transport_max_field_impulse_fraction
    c.field_impulse_floor_eV = cfg.transport_field_impulse_floor_eV
    return c


def photon_step(tables: Tables, fields, cfg, particles: dict, nthreads: int = 1, sec_cap: int | None = None,
                sources: np.ndarray | None = None):
    """Run the oracle photon step.  ``particles`` holds prev r, z, theta, the
    pushed (current) r_cur, z_cur, theta_cur, u = E dir (ux, uy, uz), w, id,
    cpu.  Returns a dict of outputs (r, z, theta, ux, uy, uz written back)."""
    L = lib()
    n = len(particles["r"])
    g, keep = make_geometry(fields, sources)
    c = make_config(cfg)
    ins = [np.ascontiguousarray(particles[k], dtype=np.float64)
           for k in ("r", "z", "theta", "r_cur", "z_cur", "theta_cur", "ux", "uy", "uz", "w")]
    pid = np.ascontiguousarray(particles["id"], dtype=np.int64)
    cpu = np.ascontiguousarray(particles["cpu"], dtype=np.int32)
    outs = {k: np.zeros(n) for k in ("r", "z", "theta", "ux", "uy", "uz")}
    status = np.zeros(n, dtype=np.int32)
    ion = np.zeros(n)
    events = np.zeros(n, dtype=np.int64)
    cap = sec_cap if sec_cap is not None else max(1024, 4 * n)
    sec = (Secondary * cap)()
    nsec = C.c_int64(0)
    tallies = np.zeros(N_TALLIES)
    code = L.orc_photon_step(
        tables.h, C.byref(g), C.byref(c), n, *[_ptr(a) for a in ins], _ptr(pid, P_I64), _ptr(cpu, P_I32),
        *[_ptr(outs[k]) for k in ("r", "z", "theta", "ux", "uy", "uz")], _ptr(status, P_I32), _ptr(ion),
        _ptr(events, P_I64), cap, sec, C.byref(nsec), _ptr(tallies), nthreads)
    del keep
    secs = np.frombuffer(bytes(sec), dtype=SECONDARY_DTYPE, count=cap)[: nsec.value].copy()
    return {"code": code, **outs, "status": status, "ion_pairs": ion, "substeps": events, "secondaries": secs,
            "tallies": tallies}


def newborn_drain(tables: Tables, fields, cfg, initial: np.ndarray, cap: int | None = None,
                  sources: np.ndarray | None = None):
    """Run the oracle's transient FIFO over the step's staged secondaries
    (SECONDARY_DTYPE records, Apply's order).  Returns the whole FIFO, one
    status and ion-pair weight per entry, the survivors, and the J grid and
    tallies the newborns add."""
    L = lib()
    initial = np.ascontiguousarray(initial, dtype=SECONDARY_DTYPE)
    n0 = len(initial)
    cap = cap if cap is not None else max(1024, 8 * n0)
    g, keep = make_geometry(fields, sources)
    c = make_config(cfg)
    src = (Secondary * max(n0, 1)).from_buffer_copy(initial.tobytes() + bytes(SECONDARY_DTYPE.itemsize * (n0 == 0)))
    work = (Secondary * cap)()
    surv = (Secondary * cap)()
    status = np.zeros(cap, dtype=np.int32)
    ion = np.zeros(cap)
    sub = np.zeros(cap, dtype=np.int64)
    nw = C.c_int64(0)
    ns = C.c_int64(0)
    nrp, nzp = fields.padded_shape()
    jr = np.zeros((nrp + 1) * nzp)
    jz = np.zeros(nrp * (nzp + 1))
    rho = np.zeros(nrp * nzp)
    tallies = np.zeros(N_TALLIES)
    code = L.orc_newborn_drain(tables.h, C.byref(g), C.byref(c), src, n0, cap, work, C.byref(nw), _ptr(status, P_I32),
                               _ptr(ion), _ptr(sub, P_I64), surv, C.byref(ns), _ptr(jr), _ptr(jz), _ptr(rho),
                               _ptr(tallies))
    del keep
    w = np.frombuffer(bytes(work), dtype=SECONDARY_DTYPE, count=cap)[: nw.value].copy()
    sv = np.frombuffer(bytes(surv), dtype=SECONDARY_DTYPE, count=cap)[: ns.value].copy()
    return {"code": code, "work": w, "status": status[: nw.value].copy(), "ion_pairs": ion[: nw.value].copy(),
            "substeps": sub[: nw.value].copy(), "survivors": sv, "jr": jr, "jz": jz, "rho": rho, "tallies": tallies}


def pack_transient_index(newborn: bool, species: int, lineage: int, segment: int, event: int, inner: int):
    """(valid, value) of the packed transient RNG index."""
    out = C.c_uint64(0)
    ok = lib().orc_pack_transient_index(int(newborn), species, lineage, segment, event, inner, C.byref(out))
    return bool(ok), int(out.value)


def live_hazard_index(species: int, segment: int, event: int):
    out = C.c_uint64(0)
    ok = lib().orc_live_hazard_index(species, segment, event, C.byref(out))
    return bool(ok), int(out.value)


def electron_state_demotes(energy, cutoff, threshold, field, drag) -> bool:
    return bool(lib().orc_electron_state_demotes(energy, cutoff, threshold, field, drag))


def newborn_elapsed(birth: float, completed: int, substeps: int, within: float) -> float:
    return lib().orc_newborn_elapsed(birth, completed, substeps, within)


def sample_annihilation(k: float, draw_energy: float, draw_azimuth: float) -> dict:
    """SamplePositronAnnihilationFinalState of the oracle."""
    out = np.zeros(6)
    lib().orc_sample_annihilation(k, draw_energy, draw_azimuth, _ptr(out))
    return {"e1": out[0], "e2": out[1], "cos1": out[2], "azimuth_u": out[3], "valid": bool(out[4]),
            "at_rest": bool(out[5])}


def electron_step(tables: Tables, fields, cfg, particles: dict, nthreads: int = 1, sec_cap: int | None = None):
    """Run the oracle electron step.  ``particles`` holds prev r, z, theta, ux,
    uy, uz, w (float64), id (int64) and cpu (int32).  Returns a dict of outputs."""
    return charged_step("electron", tables, fields, cfg, particles, nthreads, sec_cap)


def positron_step(tables: Tables, fields, cfg, particles: dict, nthreads: int = 1, sec_cap: int | None = None):
    """Run the oracle positron step (same inputs and outputs as electron_step)."""
    return charged_step("positron", tables, fields, cfg, particles, nthreads, sec_cap)


def charged_step(species: str, tables: Tables, fields, cfg, particles: dict, nthreads: int = 1,
                 sec_cap: int | None = None, sources: np.ndarray | None = None):
    L = lib()
    n = len(particles["r"])
    g, keep = make_geometry(fields, sources)
    c = make_config(cfg)
    ins = [np.ascontiguousarray(particles[k], dtype=np.float64) for k in ("r", "z", "theta", "ux", "uy", "uz", "w")]
    pid = np.ascontiguousarray(particles["id"], dtype=np.int64)
    cpu = np.ascontiguousarray(particles["cpu"], dtype=np.int32)
    outs = {k: np.zeros(n) for k in ("r", "z", "theta", "ux", "uy", "uz")}
    status = np.zeros(n, dtype=np.int32)
    ion = np.zeros(n)
    sub = np.zeros(n, dtype=np.int64)
    cap = sec_cap if sec_cap is not None else max(1024, n)
    sec = (Secondary * cap)()
    nsec = C.c_int64(0)
    nrp, nzp = fields.nr + 2, fields.nz + 4
    jr = np.zeros((nrp + 1) * nzp)
    jz = np.zeros(nrp * (nzp + 1))
    rho = np.zeros(nrp * nzp)
    tallies = np.zeros(N_TALLIES)
    code = getattr(L, f"orc_{species}_step")(
        tables.h, C.byref(g), C.byref(c), n, *[_ptr(a) for a in ins], _ptr(pid, P_I64), _ptr(cpu, P_I32),
        *[_ptr(outs[k]) for k in ("r", "z", "theta", "ux", "uy", "uz")], _ptr(status, P_I32), _ptr(ion),
        _ptr(sub, P_I64), cap, sec, C.byref(nsec), _ptr(jr), _ptr(jz), _ptr(rho), _ptr(tallies), nthreads)
    del keep
    secs = np.frombuffer(bytes(sec), dtype=SECONDARY_DTYPE, count=cap)[: nsec.value].copy()
    return {"code": code, **outs, "status": status, "ion_pairs": ion, "substeps": sub,
            "secondaries": secs, "jr": jr, "jz": jz, "rho": rho, "tallies": tallies}
