"""The Skyrme parameter -> EDF coefficient map and constants of
mocca_amd.functional against the values the reference's own Force module
computed for SLy4 (stored with the golden vectors)."""
import numpy as np
import pytest

from mocca_amd import functional as fn
from oracle_lib import golden


@pytest.mark.parametrize("case", ["scf_trc_fd", "scf_todd_fd", "ev8_trc_fd", "ev8_todd_lag"])
def test_sly4_coefficients_and_constants_match_reference(case):
    g = golden(case)
    m = g["meta"]
    a = fn.scf_arguments("SLy4", m["neutrons"], m["protons"], m["trc"])
    assert np.array_equal(a["B"], g["coef"])
    c = g["consts"]
    assert np.array_equal(np.array(a["hbm"]), c[0:2])
    assert np.allclose(a["mass"], c[2:4], rtol=1e-15, atol=0)
    assert a["e2"] == c[4] and fn.HBAR == c[5]
    assert a["com1body"] == m["com1body"]
