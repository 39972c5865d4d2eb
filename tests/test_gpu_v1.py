"""MOCCa v1 wavefunction files from and into a solve on the GPU
(mocca_v1_save / mocca_v1_load), against files the reference wrote and
restarted from (oracle/ref_driver.f90, tasks wv1 and rv1).

- mocca_v1_save after the same run as the reference's (HF with time
  reversal, HF with time-odd fields + Lagrange mesh + constrained moments,
  HFB): every record against the reference's file.  Integers, symmetries,
  quantum numbers, sizes, the functional and the HF KappaHFB guess exactly;
  energies 1e-9 MeV; states, densities 1e-10 of their scale; single-particle
  energies 1e-10; Gram-Schmidt norms, angular momenta, cutoff factors, Fermi
  energies, RhoHFB / KappaHFB and the constrained moments' values 1e-9.  U,
  V and CanTransfo are written as zeros (a restart rebuilds them).
- mocca_v1_load of the reference's file, and of the library's own file,
  then the start-up with the file's densities and 3 iterations: energies of
  every iteration 1e-9 MeV from the reference's restart (Fermi and pairing
  energies too with HFB).
"""
import numpy as np
import pytest

import mocca_amd
import oracle_lib as ol
from oracle_lib import golden, relerr
from test_gpu_bcs import _bcs
from test_gpu_scf import _constraints, _functional, _pairing, _solve, _traj_cols

pytestmark = pytest.mark.gpu
V1_CASES = ["scf_trc_fd", "scfcon_todd_lag", "scfhfb_trc_fd", "scfbcs_trc_fd"]


def _run(g):
    m = g["meta"]
    s = _solve(g, np.zeros(m["nwt"]))
    if m["pairing_type"] == 2:
        _pairing(s, g)
    elif m["pairing_type"] == 1:
        _bcs(s, g)
    s.scf_start()
    s.iterate(m["niter"], m["dt"] / g["consts"][5], m["momentum"], fetch=False)
    return s


def _save(s, ref, path):
    s.save_v1(str(path), fd_order=list(ref["fd_order"]), force=ref["force"],
              skyrme=list(ref["skyrme"]), hbar=float(ref["hbar"][0]), com2body=int(ref["com"][1]))


@pytest.mark.parametrize("case", V1_CASES)
def test_save_matches_the_reference_checkpoint(case, tmp_path):
    g = golden(case)
    m = g["meta"]
    hfb = m["pairing_type"] != 0      # HFB or BCS
    print(f"v1 save: case {case}", flush=True)
    ref = ol.v1_parse(g["v1_file"].tobytes())
    s = _run(g)
    _save(s, ref, tmp_path / "lib.v1")
    s.close()
    f = ol.v1_parse((tmp_path / "lib.v1").read_bytes())
    for k in ("version", "n", "dx", "sym", "nwt", "particles", "fd_order", "force", "skyrme",
              "force2", "hbar", "com", "B", "pairing_type", "strength", "alpha", "cut",
              "hfb_size", "blocksizes", "omega", "crank_values", "state_qn"):
        assert np.array_equal(np.asarray(f[k]), np.asarray(ref[k])), k
    assert np.allclose(f["hbm"], ref["hbm"], rtol=1e-15, atol=0)
    assert abs(f["energy"] - ref["energy"]) < 1e-9
    assert np.max(np.abs(f["denergy"] - ref["denergy"])) < 2e-9
    assert relerr(f["grid"].reshape(-1), ref["grid"].reshape(-1)) < 1e-10
    assert relerr(f["state_energy"], ref["state_energy"]) < 1e-10
    assert relerr(f["state_dispersion"], ref["state_dispersion"]) < 1e-7
    assert np.max(np.abs(f["state_norm"] - ref["state_norm"])) < 1e-9
    assert np.max(np.abs(f["state_J"] - ref["state_J"])) < 1e-9
    if hfb:
        assert np.max(np.abs(f["state_occupation"] - ref["state_occupation"])) < 1e-9
    else:
        assert np.array_equal(f["state_occupation"], ref["state_occupation"])
    for k, (a, b) in enumerate(zip(f["dens_rec"], ref["dens_rec"])):
        assert a.size == b.size, k
        if a.size:
            assert relerr(a, b) < 1e-10, k
    assert f["cutoffs"].size == ref["cutoffs"].size
    if hfb:
        assert np.max(np.abs(f["cutoffs"] - ref["cutoffs"])) < 1e-9
        assert np.max(np.abs(f["fermi"] - ref["fermi"])) < 1e-9
        assert np.array_equal(f["hfb_columns"], ref["hfb_columns"])
        for k in ("rho_hfb", "kappa_hfb"):
            assert np.max(np.abs(f[k] - ref[k])) < 1e-9, k
    else:
        for k in ("fermi", "hfb_columns", "rho_hfb", "kappa_hfb"):
            assert np.array_equal(f[k], ref[k]), k
    assert np.array_equal(f["lnlambda"], ref["lnlambda"])
    for k in ("U", "V", "can_transfo"):
        assert f[k].shape == ref[k].shape and not f[k].any(), k
    assert len(f["moments"]) == len(ref["moments"])
    for a, b in zip(f["moments"], ref["moments"]):
        for k in ("constraint_type", "isoswitch", "l", "m", "impart", "true_constraint"):
            assert np.array_equal(np.asarray(a[k]), np.asarray(b[k])), k
        for k in ("intensity", "constraint", "value", "old_value"):
            assert np.max(np.abs(a[k] - b[k])) < 1e-9 * max(1.0, np.abs(b[k]).max()), k


@pytest.mark.parametrize("source", ["reference", "library"])
@pytest.mark.parametrize("case", V1_CASES)
def test_restart_from_a_checkpoint_vs_reference(case, source, tmp_path):
    g = golden(case)
    m = g["meta"]
    print(f"v1 restart: case {case} from the {source}'s file", flush=True)
    path = tmp_path / "restart.v1"
    if source == "reference":
        path.write_bytes(g["v1_file"].tobytes())
    else:
        s0 = _run(g)
        _save(s0, ol.v1_parse(g["v1_file"].tobytes()), path)
        s0.close()
    s = mocca_amd.Solve(m["nx"], m["ny"], m["nz"], m["dx"], g["consts"][6], m["nwt"],
                        time_reversal=m["trc"], lagrange=m["lagrange"])
    s.set_states_meta(g["qn"], np.zeros(m["nwt"]))
    _functional(s, g)
    if m.get("constraints"):
        s.set_constraints(_constraints(m))
    if m["pairing_type"] == 2:
        s.set_pairing(type=2, strength=m["pairing_strength"], alpha=m["pairing_alpha"],
                      cut=m["pairing_cut"], mu=m["pairing_mu"], rhosat=m["rhosat"],
                      hfb_mix=m["hfb_mix"], iterations=m["hfb_iter"], guess_kappa=False)
    elif m["pairing_type"] == 1:
        s.set_pairing(type=1, strength=m["pairing_strength"], alpha=m["pairing_alpha"],
                      cut=m["pairing_cut"], mu=m["pairing_mu"], rhosat=m["rhosat"],
                      iterations=m["pairing_iter"], fermi_prec=m["fermi_prec"],
                      guess_kappa=False)
    info = s.load_v1(str(path))
    assert info["nwt"] == m["nwt"] and info["pairing_type"] == m["pairing_type"]
    # what the handle took from the file, written back before any step
    src = ol.v1_parse(path.read_bytes())
    _save(s, src, tmp_path / "back.v1")
    back = ol.v1_parse((tmp_path / "back.v1").read_bytes())
    for k in ("grid", "state_qn", "state_energy", "state_dispersion", "state_occupation",
              "state_norm", "fermi"):
        assert np.array_equal(back[k], src[k]), k
    if m["pairing_type"] == 2:
        assert np.array_equal(back["rho_hfb"], src["rho_hfb"])
        assert np.array_equal(back["kappa_hfb"], src["kappa_hfb"])
    for a, b in zip(back["dens_rec"][:3], src["dens_rec"][:3]):
        assert np.array_equal(a, b)
    s.scf_start()
    rows = s.iterate(3, m["dt"] / g["consts"][5], m["momentum"])
    s.close()
    traj = _traj_cols(rows)
    ref = g["v1_restart_traj"]
    print(f"  energies {traj[:, 0]} (reference {ref[:, 0]})", flush=True)
    assert np.max(np.abs(traj[:, 0] - ref[:, 0])) < 1e-9
    assert relerr(traj, ref) < 1e-10
    if m["pairing_type"] != 0:
        pr = g["v1_restart_pairing"]                   # Fermi(2), E_pair(2)
        assert np.max(np.abs(rows[:, 45:47] - pr[:, :2])) < 1e-9
        assert np.max(np.abs(rows[:, 43:45] - pr[:, 2:])) < 1e-9


def test_load_refuses_a_different_mesh_or_broken_symmetry(tmp_path):
    g = golden("scfcon_todd_lag")          # time reversal broken on file
    m = g["meta"]
    path = tmp_path / "todd.v1"
    path.write_bytes(g["v1_file"].tobytes())
    s = mocca_amd.Solve(m["nx"], m["ny"], m["nz"], m["dx"], g["consts"][6], m["nwt"],
                        time_reversal=True, lagrange=True)
    s.set_states_meta(g["qn"], np.zeros(m["nwt"]))
    with pytest.raises(mocca_amd.MoccaError, match="time reversal broken on file"):
        s.load_v1(str(path))
    s.close()
    s = mocca_amd.Solve(m["nx"] + 2, m["ny"], m["nz"], m["dx"], g["consts"][6], m["nwt"],
                        time_reversal=False, lagrange=True)
    s.set_states_meta(g["qn"], np.zeros(m["nwt"]))
    with pytest.raises(mocca_amd.MoccaError, match="mesh"):
        s.load_v1(str(path))
    s.close()
