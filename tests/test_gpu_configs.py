"""GPU parity at the sizes of the benchmark configurations, through the C ABI.

- Lagrange meshes with 48- and 64-point lines (the 48^3 / 600-state
  Lagrange workload's matrix-core derivative kernels, k_lag_axis with padded
  line lengths 48 and 64): hPsi and one GradDesc + Gram-Schmidt against the
  oracle's Lagrange derivatives.  FP64 hPsi 1e-11, energies 1e-11, states
  after the step 1e-10; mixed precision 2e-5.
- The Coulomb box of that workload (48^3 stored points at dx = 0.7 fm, a
  50^3 Coulomb box) with a Woods-Saxon proton charge of Z = 12: the direct
  solve meets the discrete equation to 1e-10 (relative max-norm residual)
  and matches e^2 Z / r beyond the charge to the discretisation error of
  the 5-point operator (1e-3 relative).
- The 40^3 mesh (6-row y tiles of the pipelined FP64 step) with a block of
  110 states: GradDesc + Gram-Schmidt and the densities against the oracle
  (FP64 1e-11 energies and densities, 1e-10 states; mixed 2e-5).
- The 40^3 / 400-state self-consistent workload: the mixed-precision total
  energy stays within 1 keV of the FP64 one along 200 iterations.
- The batched constraint scan (64 solves, 16^3..32^3, 20..160 states,
  alternating Q20): every solve's energy rows equal, bit for bit, the same
  solve iterated alone, and the smallest and the largest solve's first two
  iterations match the oracle's Evolve loop (1e-9 relative: the oracle
  iterates the reference's Coulomb CG, the batch solves it directly).
"""
import numpy as np
import pytest

import mocca_amd
import oracle_lib as ol
from mocca_amd import functional, workloads
from oracle_lib import Mesh, density, graddesc, hpsi as or_hpsi, relerr

pytestmark = pytest.mark.gpu


def _fields_case(nx, ny, nz, dx, M, seed, time_odd, lagrange, qn=None, ho=False):
    rng = np.random.default_rng(seed)
    if qn is None:
        qn = workloads.balanced_qn(M, time_reversal=not time_odd)
    if ho:
        psi = workloads.ho_states(rng, nx, ny, nz, dx, [tuple(q) for q in qn], b=2.4)
    else:
        psi = workloads.plane_wave_states(rng, nx, ny, nz, dx, M)
    dv = dx**3 * 8
    psi /= np.sqrt((psi.reshape(M, -1) ** 2).sum(1) * dv)[:, None, None]
    F = workloads.woods_saxon_fields(nx, ny, nz, dx, time_odd=time_odd)
    if not time_odd:
        N = nx * ny * nz
        F.update(S=np.zeros((2, 3, N)), A=np.zeros((2, 3, N)), DivA=np.zeros((2, N)))
    kin = np.array([workloads.HBM_HALF, workloads.HBM_HALF])
    return dict(nx=nx, ny=ny, nz=nz, dx=dx, dv=dv, M=M, qn=np.asarray(qn, np.int32), psi=psi,
                F=F, kin=kin, occ=rng.uniform(0.5, 1.0, M), dt=0.010 / functional.HBAR,
                mu=0.5, trc=not time_odd, lagrange=lagrange)


def _solve_of(c, prec):
    s = mocca_amd.Solve(c["nx"], c["ny"], c["nz"], c["dx"], c["dv"], c["M"],
                        time_reversal=c["trc"], lagrange=c["lagrange"], precision=prec)
    s.set_states_meta(c["qn"], c["occ"])
    s.upload_states(c["psi"])
    F = c["F"]
    s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], c["kin"],
                     S=F["S"], A=F["A"], DivA=F["DivA"])
    return s


def _oracle_derivatives(c):
    mesh = Mesh(c["nx"], c["ny"], c["nz"], c["dx"], c["lagrange"])
    M, N = c["M"], c["nx"] * c["ny"] * c["nz"]
    der = np.zeros((M, 3, 4, N))
    lap = np.zeros((M, 4, N))
    for i, q in enumerate(c["qn"]):
        der[i], lap[i] = mesh.derive_spinor(c["psi"][i], q[1], q[2], q[3])
    return der, lap


@pytest.mark.parametrize("shape", [(48, 64, 16), (40, 56, 20)])
@pytest.mark.parametrize("prec,tol_e,tol_psi", [(mocca_amd.FP64, 1e-11, 1e-10),
                                                (mocca_amd.MIXED, 2e-5, 2e-5)])
def test_lagrange_48_and_64_point_lines_vs_oracle(shape, prec, tol_e, tol_psi):
    nx, ny, nz = shape
    print(f"lagrange lines: mesh {nx}x{ny}x{nz} precision {prec}", flush=True)
    c = _fields_case(nx, ny, nz, 0.7, 10, seed=21, time_odd=True, lagrange=True)
    M, N = c["M"], nx * ny * nz
    s = _solve_of(c, prec)
    der, lap = _oracle_derivatives(c)
    Ff = {k: np.ascontiguousarray(v).reshape(-1) for k, v in c["F"].items()}
    if prec == mocca_amd.FP64:
        for i in (0, 3, M - 1):
            it = (c["qn"][i][0] + 3) // 2 - 1
            ref = or_hpsi(c["psi"][i], der[i], lap[i], it, False, c["kin"][it], Ff, N)
            assert relerr(s.hpsi(i), ref) < 1e-11
    s.grad_desc(c["dt"], c["mu"])
    psi = s.download_states()
    e, _ = s.energies()
    ref_psi, _, ref_e, _ = graddesc(c["psi"], der, lap, np.zeros_like(c["psi"]), c["qn"],
                                    False, c["kin"], Ff, c["dt"], c["mu"], c["dv"])
    err_e, err_p = relerr(e, ref_e), relerr(psi, ref_psi.reshape(M, 4, N))
    print(f"  energies {err_e:.2e} states {err_p:.2e}", flush=True)
    assert err_e < tol_e
    assert err_p < tol_psi
    s.close()


@pytest.mark.parametrize("shape", [(48, 64, 16), (40, 56, 20), (48, 48, 48)])
def test_lagrange_fp32_split_matrix_cores_vs_oracle_and_fp64(shape):
    """Mixed precision with the Lagrange derivatives on the FP32 matrix cores
    (matrices split into high and low FP32 parts, FP32 sums), against the
    oracle's FP64 derivatives: hPsi within 2e-5 of the largest value, the step
    + Gram-Schmidt within 2e-5 (the tolerance of the mixed path).  Its error
    is printed beside that of the FP64 matrix cores on the same FP32 states."""
    nx, ny, nz = shape
    print(f"lagrange lines, FP32 split matrix cores: mesh {nx}x{ny}x{nz}", flush=True)
    c = _fields_case(nx, ny, nz, 0.7, 10, seed=23, time_odd=True, lagrange=True)
    M, N = c["M"], nx * ny * nz
    der, lap = _oracle_derivatives(c)
    Ff = {k: np.ascontiguousarray(v).reshape(-1) for k, v in c["F"].items()}
    ref_psi, _, ref_e, _ = graddesc(c["psi"], der, lap, np.zeros_like(c["psi"]), c["qn"],
                                    False, c["kin"], Ff, c["dt"], c["mu"], c["dv"])
    errs = {}
    for mode in (0, 1):
        s = _solve_of(c, mocca_amd.MIXED)
        s.set_lagrange_precision(mode)
        hp = []
        for i in (0, M - 1):
            it = (c["qn"][i][0] + 3) // 2 - 1
            ref = or_hpsi(c["psi"][i], der[i], lap[i], it, False, c["kin"][it], Ff, N)
            hp.append(relerr(s.hpsi(i), ref))
        s.grad_desc(c["dt"], c["mu"])
        psi = s.download_states()
        e, _ = s.energies()
        errs[mode] = (max(hp), relerr(e, ref_e), relerr(psi, ref_psi.reshape(M, 4, N)))
        s.close()
    print("  FP64 matrix cores: hpsi %.2e energies %.2e states %.2e" % errs[0], flush=True)
    print("  FP32 split:        hpsi %.2e energies %.2e states %.2e" % errs[1], flush=True)
    assert max(errs[1]) < 2e-5
    assert max(errs[0]) < 2e-5


def _woods_saxon_proton_case(n=48, dx=0.7, Z=12.0, A=24.0, a=0.6):
    """One proton state carrying a Woods-Saxon density of Z protons (and one
    neutron state so that both isospins exist): rho_p = occ psi^2 with psi in
    the all-even component (Re up, P = S = TS = +1)."""
    X, Y, Zc = workloads.mesh(n, n, n, dx)
    r = np.sqrt(X**2 + Y**2 + Zc**2)
    R = 1.2 * A ** (1.0 / 3.0)
    f = 1.0 / (1.0 + np.exp((r - R) / a))
    dv = dx**3 * 8
    rho = f / (f.sum() * dv) * Z               # normalised to Z on the full box
    psi = np.zeros((2, 4, n**3))
    psi[1, 0] = np.sqrt(rho / 2.0).reshape(-1)
    psi[0, 0] = np.sqrt(rho / 2.0).reshape(-1)
    qn = np.array([(-1, 1, 1, 1), (1, 1, 1, 1)], np.int32)
    return dict(n=n, dx=dx, dv=dv, psi=psi, qn=qn, rho=rho.reshape(-1), r=r.reshape(-1),
                R=R, a=a, Z=Z)


def test_coulomb_box_of_the_lagrange_workload():
    from test_gpu_scf import _box_residual
    c = _woods_saxon_proton_case()
    n = c["n"]
    print(f"coulomb: {n}^3 stored, {n + 2}^3 box, Woods-Saxon Z = {c['Z']}", flush=True)
    s = mocca_amd.Solve(n, n, n, c["dx"], c["dv"], 2, time_reversal=True, lagrange=True)
    s.set_states_meta(c["qn"], np.array([2.0, 2.0]))
    s.upload_states(c["psi"])
    s.scf_setup(**functional.scf_arguments("SLy4", 12, 12, True, coulomb=2))
    s.update_densities()
    s.construct_potentials()
    vc, _, cp, _ = s.coulomb()
    d = s.compute_density()
    assert relerr(d["rho"][1], c["rho"]) < 1e-13
    res = _box_residual(cp, d["rho"][1], n, n, n, c["dx"], functional.E2)
    print(f"  residual {res:.2e}", flush=True)
    assert res < 1e-10
    # beyond the charge the potential is e^2 Z / r (spherical charge)
    out = c["r"] > c["R"] + 12 * c["a"]
    exact = functional.E2 * c["Z"] / c["r"][out]
    err = np.max(np.abs(vc[out] - exact) / exact)
    print(f"  outside the charge: max relative difference from e2 Z / r {err:.2e}", flush=True)
    assert err < 1e-3
    s.close()


@pytest.mark.parametrize("prec,tol_e,tol_psi,tol_d", [(mocca_amd.FP64, 1e-11, 1e-10, 1e-11),
                                                      (mocca_amd.MIXED, 2e-5, 2e-5, 2e-5)])
def test_40_cubed_mesh_with_a_110_state_block_vs_oracle(prec, tol_e, tol_psi, tol_d):
    n, M = 40, 128
    qn = np.array([(-1, 1, 1, 1)] * 110 + [(1, 1 if j % 2 == 0 else -1, 1, 1)
                                           for j in range(M - 110)], np.int32)
    print(f"40^3 mesh, {M} states, blocks of 110/9/9, precision {prec}", flush=True)
    c = _fields_case(n, n, n, 0.8, M, seed=31, time_odd=False, lagrange=False, qn=qn)
    N = n**3
    s = _solve_of(c, prec)
    der, lap = _oracle_derivatives(c)
    d = s.compute_density()
    rho, tau, nj, _, _ = density(c["psi"], der, c["qn"], c["occ"], True, N)
    for k, ref in (("rho", rho), ("tau", tau), ("nablaJ", nj)):
        err = relerr(d[k], ref)
        print(f"  density {k} {err:.2e}", flush=True)
        assert err < tol_d, k
    Ff = {k: np.ascontiguousarray(v).reshape(-1) for k, v in c["F"].items()}
    s.grad_desc(c["dt"], c["mu"])
    psi = s.download_states()
    e, _ = s.energies()
    ref_psi, _, ref_e, _ = graddesc(c["psi"], der, lap, np.zeros_like(c["psi"]), c["qn"],
                                    True, c["kin"], Ff, c["dt"], c["mu"], c["dv"])
    err_e, err_p = relerr(e, ref_e), relerr(psi, ref_psi.reshape(M, 4, N))
    print(f"  energies {err_e:.2e} states {err_p:.2e}", flush=True)
    assert err_e < tol_e
    assert err_p < tol_psi
    s.close()


def test_full_iteration_mixed_within_1keV_of_fp64():
    w = workloads.full_iteration()
    rows = {}
    for prec in (mocca_amd.FP64, mocca_amd.MIXED):
        s = mocca_amd.Solve(w["nx"], w["ny"], w["nz"], w["dx"], w["dv"], w["M"],
                            time_reversal=w["trc"], lagrange=w["lagrange"], precision=prec)
        s.set_states_meta(w["qn"], w["occ"])
        s.upload_states(w["psi"])
        s.scf_setup(**w["functional"])
        s.gram_schmidt()
        s.scf_start()
        rows[prec] = s.iterate(200, w["dt_over_hbar"], w["momentum"])
        s.close()
    d = np.abs(rows[mocca_amd.MIXED][:, 0] - rows[mocca_amd.FP64][:, 0])
    print(f"40^3 / 400 states, 200 iterations: E(FP64) {rows[mocca_amd.FP64][-1, 0]:.6f} MeV, "
          f"max |E(mixed) - E(FP64)| {d.max() * 1e3:.3f} keV (last {d[-1] * 1e3:.3f} keV)",
          flush=True)
    assert np.all(np.isfinite(rows[mocca_amd.MIXED]))
    assert d.max() < 1e-3


def _start(w, precision=mocca_amd.FP64):
    s = mocca_amd.Solve(w["nx"], w["ny"], w["nz"], w["dx"], w["dv"], w["M"],
                        time_reversal=w["trc"], lagrange=w["lagrange"], precision=precision)
    s.set_states_meta(w["qn"], w["occ"])
    s.upload_states(w["psi"])
    s.scf_setup(**w["functional"])
    s.set_constraints(w["constraints"])
    s.gram_schmidt()
    s.scf_start()
    return s


def _oracle_case(w, psi):
    """The batch workload as the oracle's Evolve loop reads a reference run."""
    f = w["functional"]
    consts = np.array([f["hbm"][0], f["hbm"][1], f["mass"][0], f["mass"][1], f["e2"],
                       functional.HBAR, w["dv"], f["neutrons"], f["protons"]])
    q = w["constraints"][0]
    meta = dict(nx=w["nx"], ny=w["ny"], nz=w["nz"], dx=w["dx"], lagrange=w["lagrange"],
                nwt=w["M"], trc=w["trc"], com1body=f["com1body"], neutrons=f["neutrons"],
                protons=f["protons"], dt=w["dt_over_hbar"] * functional.HBAR,
                momentum=w["momentum"],
                constraints=[f"l={q['l']}, m={q['m']}, Constraint={q['value']!r}, "
                             f"ConstraintType={q['type']}"])
    return dict(meta=meta, coef=np.asarray(f["B"]), consts=consts, qn=w["qn"], psi=psi)


def test_batch_scan_matches_solo_runs_and_oracle():
    ws = workloads.batch_scan()
    niter = 6
    sizes = [w["nx"] ** 3 * w["M"] for w in ws]
    print(f"batch scan: {len(ws)} solves, meshes {min(w['nx'] for w in ws)}^3.."
          f"{max(w['nx'] for w in ws)}^3, states {min(w['M'] for w in ws)}.."
          f"{max(w['M'] for w in ws)}", flush=True)
    alone, start_psi = [], []
    for w in ws:
        s = _start(w)
        start_psi.append(s.download_states())
        alone.append(s.iterate(niter, w["dt_over_hbar"], w["momentum"]))
        s.close()
    solves = [_start(w) for w in ws]
    rows = mocca_amd.batch_run(solves, niter, [w["dt_over_hbar"] for w in ws],
                               [w["momentum"] for w in ws])
    for s in solves:
        s.close()
    for k, (r, a) in enumerate(zip(rows, alone)):
        assert np.all(np.isfinite(r)), k
        assert np.array_equal(r, a), k
    print("  every solve's rows equal its solo run", flush=True)
    for k in (int(np.argmin(sizes)), int(np.argmax(sizes))):
        w = ws[k]
        ref = ol.scf_run(_oracle_case(w, start_psi[k]), 2)["traj"]
        rel = np.abs(rows[k][:3, 0] - ref[:, 0]) / np.abs(ref[:, 0])
        print(f"  solve {k} ({w['nx']}^3, {w['M']} states, Q20 {w['constraints'][0]['value']:.1f}):"
              f" max relative energy difference vs oracle {rel.max():.2e}", flush=True)
        assert rel.max() < 1e-9
