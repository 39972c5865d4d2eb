"""ConvergenceCheck and FinalIteration through the C ABI, against the
reference's own records (oracle/ref_driver.f90, tasks conv and final).

- mocca_convergence_check after every iteration of the reference's runs
  (HF on finite-difference and Lagrange meshes, alternating and
  augmented-Lagrangian Q20/Q22 constraints with a switch-off, HFB pairing):
  its energy, moment and Fermi tests at the precisions the reference driver
  records (energy 1e-9..1e-2, moments 1e-4..0.5, Fermi 1e-4..100) equal the
  reference's, flag for flag.  The README sample (100 iterations, estimation,
  HFB, Q20 switched off after 10) is checked after its last iteration.
- mocca_final_iteration: the energy row with Lagrange derivatives equals the
  reference's FinalIteration row (total energy 1e-9 relative, every column
  1e-10 of the row's scale), and the convergence history takes that
  CompEnergy as the reference does.
"""
import numpy as np
import pytest

import mocca_amd
import oracle_lib as ol
from oracle_lib import golden
from test_gpu_scf import _pairing, _solve, _traj_cols

pytestmark = pytest.mark.gpu


def _flags(s, pairing):
    out = [int(s.convergence_check(energy_prec=e)["energy"]) for e in ol.CONV_EPREC]
    out += [int(s.convergence_check(moment_prec=p)["moments"]) for p in ol.CONV_MPREC]
    out += [int(s.convergence_check(pairing_prec=p)["fermi"]) for p in ol.CONV_PPREC]
    return out


def _check_final(s, g):
    row = s.final_iteration()
    ref = g["final_row"]
    got = _traj_cols(row[None, :])[0]
    rel = abs(got[0] - ref[0]) / abs(ref[0])
    scale = np.max(np.abs(ref))
    print(f"  FinalIteration: total energy {got[0]:.9f} (reference {ref[0]:.9f}), "
          f"relative difference {rel:.2e}, row {np.max(np.abs(got - ref)) / scale:.2e}",
          flush=True)
    assert rel < 1e-9
    assert np.max(np.abs(got - ref)) / scale < 1e-10


@pytest.mark.parametrize("case", ["scf_trc_fd", "scf_trc_lag", "scfcon_trc_fd",
                                  "scfcon_todd_lag", "scfhfb_trc_fd"])
def test_convergence_check_and_final_iteration_vs_reference(case):
    g = golden(case)
    m = g["meta"]
    n = m["niter"]
    hfb = m["pairing_type"] == 2
    print(f"convergence tests after each iteration: case {case} niter {n}", flush=True)
    s = _solve(g, np.zeros(m["nwt"]))
    if hfb:
        _pairing(s, g)
    s.scf_start()
    if hfb:
        # the HFB start-up row is complete at once (HF rows after the next
        # iteration's hPsi pass)
        assert _flags(s, hfb) == g["conv"][0].tolist()
    for k in range(1, n + 1):
        s.iterate(1, m["dt"] / g["consts"][5], m["momentum"], fetch=False)
        got = _flags(s, hfb)
        assert got == g["conv"][k].tolist(), f"iteration {k}: {got} vs {g['conv'][k].tolist()}"
    c = s.convergence_check()
    assert c["positive_energy"] == (g["traj"][-1, 0] >= 0.0)
    _check_final(s, g)
    s.close()


def test_sample_convergence_and_final_iteration_vs_reference():
    g = golden("sample_mg24")
    m = g["meta"]
    n = m["niter"]
    print("sample input: convergence tests after the last iteration, FinalIteration",
          flush=True)
    s = _solve(g, np.zeros(m["nwt"]))
    s.set_pairing(type=2, strength=m["pairing_strength"], alpha=m["pairing_alpha"],
                  cut=m["pairing_cut"], mu=m["pairing_mu"], rhosat=m["rhosat"],
                  hfb_mix=m["hfb_mix"], iterations=m["hfb_iter"], guess_kappa=True)
    s.estimation_seed(g["est_seed"])
    s.scf_start()
    s.iterate_estimated(n, m["dt"], m["momentum"])
    assert _flags(s, True) == g["conv"][n].tolist()
    c = s.convergence_check(energy_prec=1e-4, moment_prec=1e-4, pairing_prec=100.0)
    assert c["converged"] == bool(g["conv"][n][2] and g["conv"][n][5] and g["conv"][n][12])
    assert not c["positive_energy"]
    _check_final(s, g)
    s.close()
