"""The large-block HFB step (Broyden or Newton Fermi solver over Jacobi
diagonalisations in LDS with rotation logs, pair-Gram gaps and dense
matrix-core applies for the pairing field and the canonical states) through
the C ABI.

- Forced onto the reference's own HFB fixtures, whose blocks the in-LDS
  solver also takes: the large-block Broyden iteration with Jacobi
  diagonalisations (solver 2) or with the matrix sign function (solver 4)
  and Newton's method (solver 3) give every iteration's total energy within 1e-9
  relative of the reference run, pairing and Fermi energies within 1e-8 MeV
  (both stop at the reference's FermiPrec, 1e-12 in the particle number);
  the README sample (24Mg, 100 iterations with IterativeEstimation,
  MixingScheme 2 and a Q20 constraint, whose start-up Fermi solve the
  reference leaves unconverged) with the large-block Broyden iteration at
  the same 1e-9.
- Blocks of 100 states (200 x 200 HFB matrices, the size of the 40^3 /
  400-state workload's blocks) against a run of the reference itself
  (scfhfb_large), four iterations with pairing active throughout: total
  energy 1e-9 relative, pairing and Fermi energies 1e-8 MeV; mixed precision
  within 1 keV.
"""
import numpy as np
import pytest

import mocca_amd
from oracle_lib import golden

pytestmark = pytest.mark.gpu


def _traj_cols(rows):
    return np.concatenate([rows[:, 0:1], rows[:, 1:2] + rows[:, 2:3], rows[:, 3:5],
                           rows[:, 5:6] + rows[:, 6:7], rows[:, 7:43]], axis=1)


@pytest.mark.parametrize("solver", [2, 3, 4])
def test_large_block_solvers_on_the_reference_hfb_fixture(solver):
    from test_gpu_scf import _solve
    g = golden("scfhfb_trc_fd")
    m = g["meta"]
    n = m["niter"]
    print(f"scf loop with HFB pairing, large-block solver {solver}", flush=True)
    s = _solve(g, np.zeros(m["nwt"]), coulomb=1)
    s.set_energies(g["start_energies"])
    s.set_pairing(type=2, strength=m["pairing_strength"], alpha=m["pairing_alpha"],
                  cut=m["pairing_cut"], mu=m["pairing_mu"], rhosat=m["rhosat"],
                  hfb_mix=m["hfb_mix"], iterations=m["hfb_iter"], fermi=g["start_fermi"],
                  guess_kappa=True, solver=solver)
    s.scf_start()
    rows = s.iterate(n, m["dt"] / g["consts"][5], m["momentum"])
    traj = _traj_cols(rows)
    rel = np.abs(traj[:, 0] - g["traj"][:, 0]) / np.abs(g["traj"][:, 0])
    df = np.max(np.abs(rows[:, 45:47] - g["pairing"][:, 0:2]))
    de = np.max(np.abs(rows[:, 43:45] - g["pairing"][:, 2:4]))
    print(f"  max relative energy difference {rel.max():.2e}, Fermi {df:.2e} MeV, "
          f"pairing energy {de:.2e} MeV", flush=True)
    assert rel.max() < 1e-9
    assert df < 1e-8
    assert de < 1e-8
    fermi, epair, qpe, occ = s.pairing()
    assert np.all(occ >= 0) and np.all(occ <= 2)
    assert abs(occ.sum() - (m["neutrons"] + m["protons"])) < 1e-8
    s.close()


@pytest.mark.parametrize("solver", [2, 4])
def test_large_block_broyden_on_the_sample_input(solver):
    """The sample's start-up step has every single-particle energy at 0, where
    the reference's Broyden iteration stops unconverged after HFBIter
    iterations; the large-block path replays that iteration exactly (with
    the sign function, solver 4, the quasiparticle energies that
    IterativeEstimation reads come from one diagonalisation per step)."""
    from test_gpu_scf import _solve
    g = golden("sample_mg24")
    m = g["meta"]
    n = m["niter"]
    print(f"sample input, large-block Broyden solver {solver}", flush=True)
    s = _solve(g, np.zeros(m["nwt"]))
    s.set_pairing(type=2, strength=m["pairing_strength"], alpha=m["pairing_alpha"],
                  cut=m["pairing_cut"], mu=m["pairing_mu"], rhosat=m["rhosat"],
                  hfb_mix=m["hfb_mix"], iterations=m["hfb_iter"], guess_kappa=True,
                  solver=solver)
    s.estimation_seed(g["est_seed"])
    s.scf_start()
    rows, prm = s.iterate_estimated(n, m["dt"], m["momentum"])
    traj = _traj_cols(rows)
    diff = np.abs(traj[:, 0] - g["traj"][:, 0])
    scale = np.maximum(np.abs(g["traj"][:, 0]), 100.0)
    dp = np.max(np.abs(rows[:, 43:47] - g["pairing"][:, [2, 3, 0, 1]]))
    print(f"  max scaled energy difference {(diff / scale).max():.2e}, final "
          f"{diff[-1] / abs(g['traj'][-1, 0]):.2e}, pairing/Fermi {dp:.2e} MeV", flush=True)
    # The start-up Fermi solve (every single-particle energy 0) is the
    # reference's unconverged 50-step Broyden iteration on a particle number
    # that rises from 0 to 2h within ~0.002 MeV: rounding-level differences
    # of the eigen-solver move its end point by ~1e-8 MeV of total energy in
    # the first rows, hence 2e-9 per row; the converged end of the run
    # agrees to 1e-10.
    assert (diff / scale).max() < 2e-9
    assert diff[-1] / abs(g["traj"][-1, 0]) < 1e-10
    assert dp < 1e-7
    s.close()


@pytest.mark.parametrize("prec,solver", [(mocca_amd.FP64, 0), (mocca_amd.FP64, 2),
                                         (mocca_amd.MIXED, 0)])
def test_blocks_of_100_states_vs_reference(prec, solver):
    """The reference's run with HFB blocks of 100 states (200 x 200 HFB
    matrices): 200 neutron and 40 proton states, 50 + 14 nucleons, 4
    iterations.  The automatic solver takes the large-block path with the
    sign function here; solver 2 diagonalises every evaluation."""
    from test_gpu_scf import _solve
    g = golden("scfhfb_large")
    m = g["meta"]
    n = m["niter"]
    print(f"HFB blocks of 100 states vs the reference: precision {prec}, solver {solver}",
          flush=True)
    s = _solve(g, np.zeros(m["nwt"]), precision=prec, coulomb=1)
    s.set_energies(g["start_energies"])
    s.set_pairing(type=2, strength=m["pairing_strength"], alpha=m["pairing_alpha"],
                  cut=m["pairing_cut"], mu=m["pairing_mu"], rhosat=m["rhosat"],
                  hfb_mix=m["hfb_mix"], iterations=m["hfb_iter"], fermi=g["start_fermi"],
                  guess_kappa=True, solver=solver)
    s.scf_start()
    rows = s.iterate(n, m["dt"] / g["consts"][5], m["momentum"])
    traj = _traj_cols(rows)
    rel = np.abs(traj[:, 0] - g["traj"][:, 0]) / np.abs(g["traj"][:, 0])
    df = np.max(np.abs(rows[:, 45:47] - g["pairing"][:, 0:2]))
    de = np.max(np.abs(rows[:, 43:45] - g["pairing"][:, 2:4]))
    print(f"  energies {rows[:, 0].tolist()}", flush=True)
    print(f"  max relative energy difference {rel.max():.2e}, Fermi {df:.2e} MeV, "
          f"pairing energy {de:.2e} MeV", flush=True)
    fermi, epair, qpe, occ = s.pairing()
    s.close()
    assert np.all(np.isfinite(rows))
    assert np.all(occ >= 0) and np.all(occ <= 2)
    assert abs(occ.sum() - (m["neutrons"] + m["protons"])) < 1e-6
    if prec == mocca_amd.FP64:
        assert rel.max() < 1e-9
        assert df < 1e-8
        assert de < 1e-8
    else:
        assert np.max(np.abs(traj[:, 0] - g["traj"][:, 0])) < 1e-3


def test_large_block_path_on_a_mesh_of_odd_size():
    """A mesh whose point count is not a multiple of 4 (10 x 9 x 9 = 810,
    the reference's HFB fixture cut to it): the large-block path (its pair
    Gram then reads the states point by point) against the in-LDS path,
    whose gaps and pairing field are separate kernels pinned to the
    reference, three iterations each: total energies 1e-10 relative,
    Fermi and pairing energies 1e-9 MeV."""
    from test_gpu_scf import _solve
    g0 = golden("scfhfb_trc_fd")
    m = dict(g0["meta"])
    m["ny"], m["nz"] = 9, 9
    g = dict(g0)
    g["meta"] = m
    g["psi"] = np.ascontiguousarray(g0["psi"][:, :, :9, :9, :])
    print("HFB on a 10 x 9 x 9 mesh: in-LDS path vs large-block solvers 2 and 4", flush=True)
    out = {}
    for solver in (1, 2, 4):
        s = _solve(g, np.zeros(m["nwt"]), coulomb=1)
        s.set_energies(g0["start_energies"])
        s.set_pairing(type=2, strength=m["pairing_strength"], alpha=m["pairing_alpha"],
                      cut=m["pairing_cut"], mu=m["pairing_mu"], rhosat=m["rhosat"],
                      hfb_mix=m["hfb_mix"], iterations=m["hfb_iter"],
                      fermi=g0["start_fermi"], guess_kappa=True, solver=solver)
        s.scf_start()
        out[solver] = s.iterate(3, m["dt"] / g0["consts"][5], m["momentum"])
        s.close()
    ref = out[1]
    assert np.all(np.isfinite(ref))
    for solver in (2, 4):
        rows = out[solver]
        rel = np.max(np.abs(rows[:, 0] - ref[:, 0]) / np.abs(ref[:, 0]))
        dp = np.max(np.abs(rows[:, 43:47] - ref[:, 43:47]))
        print(f"  solver {solver}: relative energy difference {rel:.2e}, "
              f"pairing/Fermi {dp:.2e} MeV", flush=True)
        assert rel < 1e-10
        assert dp < 1e-9
