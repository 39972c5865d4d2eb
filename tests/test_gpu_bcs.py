"""BCS pairing (PairingType 1) on the device through the C ABI, against a
run of the reference (tests/golden scfbcs_trc_fd: Type="BCS", PairingIter
50, FermiPrec 1e-12, the sample's contact pairing, 6 + 4 nucleons in 6 + 6
states with time reversal).

- The start-up SolvePairing: Fermi energies, quasiparticle energies and
  occupations of the HF states to 1e-10.
- Ten iterations: every row's total energy 1e-9 MeV, the row 1e-11 of its
  scale, Fermi and pairing energies 1e-9 MeV, the final states 1e-10 and
  occupations 1e-10; in mixed precision the total energy within 1 keV.
"""
import numpy as np
import pytest

import mocca_amd
from oracle_lib import golden, relerr
from test_gpu_scf import _solve, _traj_cols

pytestmark = pytest.mark.gpu


def _bcs(s, g):
    m = g["meta"]
    s.set_energies(g["start_energies"])
    s.set_pairing(type=1, strength=m["pairing_strength"], alpha=m["pairing_alpha"],
                  cut=m["pairing_cut"], mu=m["pairing_mu"], rhosat=m["rhosat"],
                  iterations=m["pairing_iter"], fermi_prec=m["fermi_prec"],
                  fermi=g["start_fermi"], guess_kappa=False)


@pytest.mark.parametrize("prec", [mocca_amd.FP64, mocca_amd.MIXED])
def test_iterate_bcs_vs_reference(prec):
    g = golden("scfbcs_trc_fd")
    m = g["meta"]
    n = m["niter"]
    print(f"scf loop with BCS pairing: precision {prec}", flush=True)
    s = _solve(g, np.zeros(m["nwt"]), precision=prec)
    _bcs(s, g)
    s.scf_start()
    fermi, epair, qpe, occ = s.pairing()
    st = g["bcs_start"]
    tol = 1e-10 if prec == mocca_amd.FP64 else 1e-4
    assert np.max(np.abs(fermi - g["bcs_start_fermi"])) < tol * 100
    assert np.max(np.abs(qpe - st[2])) < tol * 100
    assert np.max(np.abs(occ - st[3])) < tol
    rows = s.iterate(n, m["dt"] / g["consts"][5], m["momentum"])
    traj = _traj_cols(rows)
    de = np.max(np.abs(traj[:, 0] - g["traj"][:, 0]))
    dp = np.max(np.abs(rows[:, [45, 46, 43, 44]] - g["pairing"]))
    print(f"  max total-energy difference {de:.2e} MeV, Fermi/pairing {dp:.2e} MeV", flush=True)
    if prec == mocca_amd.FP64:
        assert de < 1e-9
        assert relerr(traj, g["traj"]) < 1e-11
        assert dp < 1e-9
        assert relerr(s.download_states().reshape(-1), g["final_psi"].reshape(-1)) < 1e-10
        assert np.max(np.abs(s.occupations() - g["final_eo"][:, 1])) < 1e-10
    else:
        assert de < 1e-3
    s.close()
