"""GPU parity of IterativeEstimation (ImaginaryTime.f90:28-211) through the
C ABI: the power iteration of the extra state, relE, minh and the dt /
momentum it returns, against the oracle on the same start vector.

Tolerances: FP64 dt, momentum, maxE, relE within 1e-11 relative and the same
iteration count, for the cold start (iteration 1) and the warm start that
continues from it; mixed path (FP32 state) dt within 1e-4 relative.
"""
import numpy as np
import pytest

import mocca_amd
import oracle_lib as ol
from oracle_lib import GOLDEN_CASES, Mesh, fields_of, golden, kin_coefs, relerr

pytestmark = pytest.mark.gpu
HBAR = 6.58211928


def _solve(g, precision):
    m = g["meta"]
    s = mocca_amd.Solve(m["nx"], m["ny"], m["nz"], m["dx"], g["consts"][6], m["nwt"],
                        time_reversal=m["trc"], lagrange=m["lagrange"], precision=precision)
    s.set_states_meta(g["qn"], g["occ"])
    s.upload_states(g["psi"])
    s.set_potentials(g["U"], g["B"], g["NablaB"], g["W"], g["DivW"], kin_coefs(g),
                     S=g["S"], A=g["A"], DivA=g["DivA"])
    return s


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("prec", [mocca_amd.FP64, mocca_amd.MIXED])
def test_estimate_parameters_vs_oracle(case, prec):
    g = golden(case)
    m = g["meta"]
    N = m["nx"] * m["ny"] * m["nz"]
    dv = g["consts"][6]
    print(f"estimation: case {case} precision {prec}", flush=True)
    s = _solve(g, prec)
    # single-particle energies with particle-hole gaps in both isospins
    energy = np.linspace(-40.0, -5.0, m["nwt"])
    s.set_energies(energy)
    seed = np.random.default_rng(7).uniform(0.0, 1.0, 4 * N)
    s.estimation_seed(seed)
    mesh = Mesh(m["nx"], m["ny"], m["nz"], m["dx"], m["lagrange"])
    F = fields_of(g)
    kin2 = kin_coefs(g)
    qn = g["qn"]
    qe = qn[-1]
    kin = kin2[(qe[0] + 3) // 2 - 1]
    psi = seed.copy()
    ref = ol.estimate(mesh, psi, 1, qe, m["trc"], kin, F, 0.012, 0.3, dv, HBAR, qn, g["occ"],
                      energy)
    got = s.estimate_parameters(1, 0.012, 0.3)
    if prec == mocca_amd.FP64:
        assert relerr(got[:4], ref[:4]) < 1e-11
        assert got[5] == ref[5]
    else:
        assert relerr(got[:2], ref[:2]) < 1e-4
    assert abs(got[4]) < 1e-2 and got[5] > 1
    assert np.isfinite(got[:2]).all() and 0 < got[1] < 1
    # a warm start continues from the state the first call left
    ref2 = ol.estimate(mesh, psi, 0, qe, m["trc"], kin, F, ref[0], ref[1], dv, HBAR, qn,
                       g["occ"], energy)
    got2 = s.estimate_parameters(2, got[0], got[1])
    if prec == mocca_amd.FP64:
        assert relerr(got2[:4], ref2[:4]) < 1e-11
        assert got2[5] == ref2[5]
        assert got2[5] < got[5]
    else:
        assert relerr(got2[:2], ref2[:2]) < 1e-4
    s.close()


def test_estimation_random_start_converges():
    """Without a seed the state starts from the library's own uniform
    numbers; maxE still converges to the same eigenvalue (within the 1e-2
    MeV stopping rule) as from the oracle's start."""
    g = golden("ev8_trc_fd")
    energy = np.linspace(-40.0, -5.0, g["meta"]["nwt"])
    s = _solve(g, mocca_amd.FP64)
    s.set_energies(energy)
    a = s.estimate_parameters(1, 0.012, 0.3)
    b = s.estimate_parameters(2, a[0], a[1])
    s.close()
    t = _solve(g, mocca_amd.FP64)
    t.set_energies(energy)
    N = g["meta"]["nx"] * g["meta"]["ny"] * g["meta"]["nz"]
    t.estimation_seed(np.random.default_rng(11).uniform(0.0, 1.0, 4 * N))
    c = t.estimate_parameters(1, 0.012, 0.3)
    d = t.estimate_parameters(2, c[0], c[1])
    t.close()
    assert abs(b[2] - d[2]) < 1e-3 * abs(d[2])
    assert a[3] == c[3]


def test_estimation_rejects_bad_arguments():
    g = golden("ev8_trc_fd")
    s = _solve(g, mocca_amd.FP64)
    with pytest.raises(mocca_amd.MoccaError):
        s.estimate_parameters(0, 0.012, 0.3)
    with pytest.raises(ValueError):
        s.estimation_seed(np.zeros(5))
    s.close()
