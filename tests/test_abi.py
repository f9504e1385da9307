"""The C-ABI library loads, exports every symbol include/rrea_hip.h declares,
and the ctypes mirrors match the header's layout (CPU only, no device call)."""
import ctypes as C
import json
import os
import re
import subprocess
import tempfile

import pytest

from rrea_feedback_warpx_amrex_geant4_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "rrea_hip.h")

STRUCTS = {
    "rrea_rate_column": _lib.RateColumn,
    "rrea_cdf_groups": _lib.CdfGroups,
    "rrea_angle_table": _lib.AngleTable,
    "rrea_elastic_table": _lib.ElasticTable,
    "rrea_photoelectric_table": _lib.PhotoelectricTable,
    "rrea_pair_bank": _lib.PairBank,
    "rrea_table_image": _lib.TableImage,
    "rrea_rz_geometry": _lib.RzGeometry,
    "rrea_step_config": _lib.StepConfigC,
    "rrea_tile": _lib.Tile,
    "rrea_secondary": _lib.Secondary,
    "rrea_step_out": _lib.StepOut,
    "rrea_drain_io": _lib.DrainIO,
    "rrea_particle_store": _lib.ParticleStore,
    "rrea_drift_config": _lib.DriftConfig,
    "rrea_batch_graph_desc": _lib.BatchGraphDesc,
    "rrea_population_records": _lib.PopulationRecords,
    "rrea_population_config": _lib.PopulationConfig,
    "rrea_group_entry": _lib.GroupEntry,
    "rrea_secondary_group": _lib.SecondaryGroup,
}


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^(?:rrea_status|const char\*)\s+(rrea_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 9
    lib = _lib.lib()
    for name in names:
        assert hasattr(lib, name), name
    assert {n for n, _, _ in _lib.SIGNATURES} == set(names)
    assert lib.rrea_version().startswith(b"rrea-mi355x")


def test_ctypes_layout_matches_header():
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {", 'printf("{");']
    first = True
    for cname, cls in STRUCTS.items():
        for fname, _ in cls._fields_:
            sep = "" if first else ","
            first = False
            lines.append(f'printf("{sep}\\"{cname}.{fname}\\": %zu", offsetof({cname}, {fname}));')
        lines.append(f'printf(",\\"{cname}\\": %zu", sizeof({cname}));')
    lines += ['printf("}");', "return 0; }"]
    with tempfile.TemporaryDirectory() as d:
        src = os.path.join(d, "layout.c")
        exe = os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.check_call(["gcc", "-std=c11", "-o", exe, src])
        layout = json.loads(subprocess.check_output([exe]))
    for cname, cls in STRUCTS.items():
        assert layout[cname] == C.sizeof(cls), cname
        for fname, _ in cls._fields_:
            assert layout[f"{cname}.{fname}"] == getattr(cls, fname).offset, (cname, fname)


def test_argument_checks_need_no_device():
    lib = _lib.lib()
    assert lib.rrea_tables_create(None, None) == 16
    assert lib.rrea_fields_create(None, None) == 16
    bad = _lib.RzGeometry(0, 10, 1.0, 1.0, 0.0, 0.0)
    h = C.c_void_p()
    assert lib.rrea_fields_create(C.byref(bad), C.byref(h)) == 16
    assert lib.rrea_transport_step_electrons(None, None, None, None, None, None) == 16
    assert lib.rrea_transport_step_noalloc(0, None, None, None, None, None, None) == 16


def test_event_ledger_layout_matches_the_oracle():
    """The header's ledger slots are the reference's RreaEventLedger fields in
    order (the oracle's list), the tally array holds the counts and then the
    ledger, and exactly the ten running maxima fold by max."""
    from oracle import binding as ob
    text = open(HEADER).read()
    blk = text[text.index("RREA_L_PHOTON_INTERACTION_COUNT = 0"):text.index("    RREA_LEDGER_COUNT ")]
    names = [n.lower() for n in re.findall(r"RREA_L_([A-Z0-9_]+)", blk)]
    assert names == ob.LEDGER_NAMES
    assert _lib.TALLY_COUNT == ob.N_TALLIES == _lib.LEDGER_OFFSET + len(names)
    src = ('#include <stdio.h>\n#include "rrea_hip.h"\n'
           'int main(void) { printf("%d %d\\n", (int)RREA_TALLY_COUNT, (int)RREA_LEDGER_OFFSET);'
           ' for (int q = 0; q < RREA_LEDGER_COUNT; ++q) printf("%d", RREA_LEDGER_IS_MAX(q) ? 1 : 0);'
           ' printf("\\n"); return 0; }\n')
    with tempfile.TemporaryDirectory() as d:
        c, exe = os.path.join(d, "l.c"), os.path.join(d, "l")
        open(c, "w").write(src)
        subprocess.check_call(["gcc", "-std=c99", "-I", os.path.dirname(HEADER), "-o", exe, c])
        out = subprocess.check_output([exe], text=True).split()
    assert int(out[0]) == _lib.TALLY_COUNT and int(out[1]) == _lib.LEDGER_OFFSET
    is_max = {n for n, b in zip(names, out[2]) if b == "1"}
    assert is_max == ob.LEDGER_MAX and len(is_max) == 10
