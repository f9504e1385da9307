"""rrea_validate_secondary_groups -- the conservation closure of interaction
groups (RreaValidateSecondaryGroupConservation, OV/RreaParticleInteraction.H:
1099-1252) evaluated on the device -- against the oracle's restatement: the
reference smoke cases (native_amrex_rrea/src/RreaSecondaryGroupSmoke.cpp:
342-523) and 4,000 random groups on either side of the tolerance, result
bits equal and the charge and energy residuals bit for bit."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")

import group_cases as gc  # noqa: E402
from oracle import binding as ob  # noqa: E402
from rrea_feedback_warpx_amrex_geant4_amd import _lib  # noqa: E402

pytestmark = pytest.mark.gpu


def test_device_closure_equals_oracle():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from rrea_feedback_warpx_amrex_geant4_amd.transport import validate_secondary_groups
    rng = np.random.default_rng(11)
    cases = ([g for g in gc.VALID.values()] + [g for g, _ in gc.INVALID.values()]
             + [gc.group(p, mode=gc.MISSING) for p in (gc.CONTINUOUS, gc.SEED)]
             + [g for g, _ in gc.random_groups(rng, 4000)])
    print(f"case group closure: {len(cases)} groups", flush=True)
    arr = (_lib.SecondaryGroup * len(cases))(*[gc.build(_lib.SecondaryGroup, _lib.GroupEntry, g) for g in cases])
    bits, res = validate_secondary_groups(arr, gc.CUT)
    for i, g in enumerate(cases):
        ob_bits, dq, de = ob.validate_group(gc.build(ob.Group, ob.GroupEntry, g), gc.CUT)
        assert bits[i] == ob_bits, (i, g, bits[i], ob_bits)
        assert res[i, 0] == dq and res[i, 1] == de, (i, g, res[i], dq, de)
    n_valid = int((bits == 15).sum())
    assert n_valid > 2000 and (bits != 15).sum() > 500
    assert (bits[:len(gc.VALID)] == 15).all()
    empty, _ = validate_secondary_groups((_lib.SecondaryGroup * 0)(), gc.CUT)
    assert len(empty) == 0
