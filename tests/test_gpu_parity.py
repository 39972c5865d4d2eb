"""GPU parity of the HIP library (through the C ABI) against the reference's
own golden vectors and the C restatement (oracle).

Tolerances (max-norm relative to the largest reference value):
  FP64 path   derivatives, hPsi, densities: 1e-11; states after the step +
              Gram-Schmidt: 1e-10; energies 1e-11.
  mixed path  FP32 stencils/hPsi/densities: 2e-5 (a few FP32 ulps of the
              accumulated sums); Gram-Schmidt stays FP64.
"""
import numpy as np
import pytest

import mocca_amd
from mocca_amd import workloads
from oracle_lib import (GOLDEN_CASES, Mesh, density, density_derived, fields_of,
                        golden, graddesc, hpsi as or_hpsi, kin_coefs, relerr)

pytestmark = pytest.mark.gpu


def _solve(g, precision=mocca_amd.FP64):
    m = g["meta"]
    s = mocca_amd.Solve(m["nx"], m["ny"], m["nz"], m["dx"], g["consts"][6], m["nwt"],
                        time_reversal=m["trc"], lagrange=m["lagrange"],
                        precision=precision)
    s.set_states_meta(g["qn"], g["occ"])
    s.upload_states(g["psi"])
    F = {k: g[k] for k in ("U", "B", "NablaB", "W", "DivW", "S", "A", "DivA")}
    s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], kin_coefs(g),
                     S=F["S"], A=F["A"], DivA=F["DivA"])
    return s


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("prec,tol", [(mocca_amd.FP64, 1e-11), (mocca_amd.MIXED, 2e-5)])
def test_derivatives_hpsi_density_vs_reference(case, prec, tol):
    g = golden(case)
    m = g["meta"]
    N = m["nx"] * m["ny"] * m["nz"]
    s = _solve(g, prec)
    s.derive_all()
    for i in range(m["nwt"]):
        der, lap = s.derivatives(i)
        assert relerr(der, g["der"][i].reshape(3, 4, N)) < tol
        assert relerr(lap, g["lap"][i].reshape(4, N)) < tol
        assert relerr(s.hpsi(i), g["hpsi"][i].reshape(4, N)) < tol
    d = s.compute_density()
    for k in ("rho", "tau", "nablaJ", "derrho", "laprho"):
        assert relerr(d[k], g[k].reshape(d[k].shape)) < tol, k
    if not m["trc"]:
        for k, gk in (("vecj", "vecj"), ("vecs", "vecs"), ("rotvecj", "rotvecj"),
                      ("rots", "rots"), ("ders", "ders")):
            assert relerr(d[k], g[gk].reshape(d[k].shape)) < tol, k
    s.close()


@pytest.mark.parametrize("case", GOLDEN_CASES)
@pytest.mark.parametrize("prec,tol", [(mocca_amd.FP64, 1e-11), (mocca_amd.MIXED, 2e-5)])
def test_angular_momentum_vs_reference(case, prec, tol):
    """UpdateAM (DiagSpin, DiagAng of every state; totals over the occupied
    states with the density integrals when time reversal is broken) against
    the reference's values for the same states."""
    g = golden(case)
    m = g["meta"]
    print(f"angular momentum: case {case} precision {prec}", flush=True)
    s = _solve(g, prec)
    s.compute_density()
    am = s.angular_momentum()
    s.close()
    assert relerr(am["states"], g["am_states"]) < tol
    if m["trc"]:
        assert not np.any(am["totals"]) and not np.any(am["isoblock"])
    else:
        assert relerr(am["totals"], g["am_totals"]) < tol
        assert relerr(am["isoblock"], g["am_isoblock"]) < tol


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_grad_desc_gram_schmidt_vs_reference(case):
    g = golden(case)
    m = g["meta"]
    N = m["nx"] * m["ny"] * m["nz"]
    s = _solve(g)
    dth = m["dt"] / g["consts"][5]
    for step in (1, 2):
        s.grad_desc(dth, m["momentum"])
        psi = s.download_states()
        e, d = s.energies()
        assert relerr(psi, g[f"step{step}_psi"].reshape(m["nwt"], 4, N)) < 1e-10
        assert relerr(e, g[f"step{step}_ed"][:, 0]) < 1e-11
        assert relerr(d, g[f"step{step}_ed"][:, 1]) < 1e-8
    s.close()


def test_gram_schmidt_standalone_orthonormal():
    g = golden("ev8_todd_fd")
    m = g["meta"]
    s = _solve(g)
    e = np.linspace(-30, 5, m["nwt"])
    s.set_energies(e)
    s.gram_schmidt()
    q = s.download_states().reshape(m["nwt"], -1)
    dv = g["consts"][6]
    qn = g["qn"]
    for key in set(map(tuple, qn[:, :3])):
        idx = [i for i in range(m["nwt"]) if tuple(qn[i, :3]) == key]
        Q = q[idx]
        assert np.allclose(Q @ Q.T * dv, np.eye(len(idx)), atol=1e-12)
    s.close()


def test_microbench_workload_vs_oracle():
    """The full 32^3 / 200-state workload: one GradDesc (hPsi of every state,
    step, Gram-Schmidt) and the densities, against the oracle at full size."""
    w = workloads.hpsi_microbench()
    n, M, N = w["nx"], w["M"], w["nx"] ** 3
    s = mocca_amd.Solve(n, n, n, w["dx"], w["dv"], M, time_reversal=False)
    s.set_states_meta(w["qn"], w["occ"])
    s.upload_states(w["psi"])
    F = w["fields"]
    s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], w["kin"],
                     S=F["S"], A=F["A"], DivA=F["DivA"])
    # densities of the input states
    d = s.compute_density()
    mesh = Mesh(n, n, n, w["dx"], False)
    der = np.zeros((M, 3, 4, N))
    lap = np.zeros((M, 4, N))
    for i, q in enumerate(w["qn"]):
        der[i], lap[i] = mesh.derive_spinor(w["psi"][i], q[1], q[2], q[3])
    rho, tau, nj, vj, vs = density(w["psi"], der, w["qn"], w["occ"], False, N)
    for k, ref in (("rho", rho), ("tau", tau), ("nablaJ", nj), ("vecj", vj), ("vecs", vs)):
        assert relerr(d[k], ref) < 1e-11, k
    # hPsi of a few states
    Ff = {k: np.ascontiguousarray(v).reshape(-1) for k, v in F.items()}
    for i in (0, 57, 133, M - 1):
        it = (w["qn"][i][0] + 3) // 2 - 1
        ref = or_hpsi(w["psi"][i], der[i], lap[i], it, False, w["kin"][it], Ff, N)
        assert relerr(s.hpsi(i), ref) < 1e-11
    # one GradDesc
    s.grad_desc(w["dt_over_hbar"], w["momentum"])
    psi = s.download_states()
    e, _ = s.energies()
    ref_psi, _, ref_e, _ = graddesc(w["psi"], der, lap, np.zeros_like(w["psi"]), w["qn"],
                                    False, w["kin"], Ff, w["dt_over_hbar"],
                                    w["momentum"], w["dv"])
    assert relerr(e, ref_e) < 1e-11
    assert relerr(psi, ref_psi.reshape(M, 4, N)) < 1e-10
    s.close()


@pytest.mark.parametrize("m,prec", [(5, mocca_amd.FP64), (25, mocca_amd.FP64),
                                    (60, mocca_amd.FP64),
                                    (100, mocca_amd.FP64), (140, mocca_amd.FP64),
                                    (150, mocca_amd.FP64), (150, mocca_amd.MIXED),
                                    (190, mocca_amd.FP64), (250, mocca_amd.FP64)])
def test_gram_schmidt_block_sizes_vs_householder_qr(m, prec):
    """Gram-Schmidt of one block of m states (energy order) through every
    kernel family: MFMA Gram up to 128 states in one wave per K slice (one
    16-row tile from 1 to 16), 129-160 states as three jobs (140: nine row
    tiles, the second job half empty; 150: ten), MFMA apply up to 192, the
    register Cholesky up to 63 states (one or two elements per thread: 25,
    60), the LDS one up to ~137, the packed-LDS one up to 192, vector
    kernels and the global-memory Cholesky beyond.  Modified Gram-Schmidt in a given order equals the QR
    factorisation with a positive-diagonal R, so the result is checked
    against numpy's Householder QR of the same energy-sorted states (an
    independent algorithm): 1e-10 in FP64; in mixed precision (FP32 states,
    FP64 Gram, Cholesky and accumulation) against the QR of the FP32-rounded
    states, 1e-5."""
    n, dx = 8, 0.8
    N, dv = n ** 3, dx ** 3 * 8
    rng = np.random.default_rng(100 + m)
    psi = rng.standard_normal((m, 4, N)) * 0.1
    if prec == mocca_amd.MIXED:
        psi = psi.astype(np.float32).astype(np.float64)
    tol, otol = (1e-10, 1e-11) if prec == mocca_amd.FP64 else (1e-5, 1e-5)
    qn = np.tile(np.array([-1, 1, 1, 1], np.int32), (m, 1))
    e = rng.permutation(np.linspace(-40.0, 10.0, m))
    s = mocca_amd.Solve(n, n, n, dx, dv, m, time_reversal=False, precision=prec)
    s.set_states_meta(qn, np.ones(m))
    s.upload_states(psi)
    s.set_energies(e)
    s.gram_schmidt()
    q = s.download_states().reshape(m, -1).astype(np.float64)
    s.close()
    order = np.argsort(e, kind="stable")
    A = psi.reshape(m, -1)[order] * np.sqrt(dv)
    Qh, R = np.linalg.qr(A.T)
    Qh = Qh * np.sign(np.diag(R))
    ref = Qh.T / np.sqrt(dv)
    assert relerr(q[order], ref) < tol
    assert np.allclose(q @ q.T * dv, np.eye(m), atol=otol)


@pytest.mark.parametrize("nx,ny,nz,trc", [(16, 40, 9, True), (32, 20, 13, True),
                                          (15, 21, 10, False), (24, 18, 7, False)])
@pytest.mark.parametrize("prec,tol", [(mocca_amd.FP64, 1e-11), (mocca_amd.MIXED, 2e-5)])
def test_density_and_step_multitile_vs_oracle(nx, ny, nz, trc, prec, tol):
    """Densities and one GradDesc on FD meshes with several y tiles (the last
    one partial) and z extents that leave short slabs: the pipelined density
    kernel (time reversal, even nx) and the two-barrier one (odd nx, time-odd
    fields), the pipelined FP64 step (even nx) and the register-queue step,
    against the oracle.  FP64 1e-11 (states after the step and Gram-Schmidt
    1e-10), mixed 2e-5."""
    rng = np.random.default_rng(nx * 1000 + ny * 10 + nz)
    dx, M = 0.8, 12
    N = nx * ny * nz
    qn = workloads.balanced_qn(M, time_reversal=trc)
    psi = workloads.plane_wave_states(rng, nx, ny, nz, dx, M)
    dv = dx ** 3 * 8
    psi /= np.sqrt((psi.reshape(M, -1) ** 2).sum(1) * dv)[:, None, None]
    occ = rng.uniform(0.5, 1.0, M)
    occ[3] = 0.0   # an empty state is skipped
    F = workloads.woods_saxon_fields(nx, ny, nz, dx, time_odd=not trc)
    kin = np.array([workloads.HBM_HALF, workloads.HBM_HALF])
    s = mocca_amd.Solve(nx, ny, nz, dx, dv, M, time_reversal=trc, precision=prec)
    s.set_states_meta(qn, occ)
    s.upload_states(psi)
    if trc:
        s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], kin)
    else:
        s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], kin,
                         S=F["S"], A=F["A"], DivA=F["DivA"])
    mesh = Mesh(nx, ny, nz, dx, False)
    der = np.zeros((M, 3, 4, N))
    lap = np.zeros((M, 4, N))
    for i, q in enumerate(qn):
        der[i], lap[i] = mesh.derive_spinor(psi[i], q[1], q[2], q[3])
    d = s.compute_density()
    rho, tau, nj, vj, vs = density(psi, der, qn, occ, trc, N)
    refs = [("rho", rho), ("tau", tau), ("nablaJ", nj)]
    if not trc:
        refs += [("vecj", vj), ("vecs", vs)]
    for k, ref in refs:
        assert relerr(d[k], ref) < tol, k
    if trc:   # no time-odd fields: zeros for the oracle
        F = dict(F, S=np.zeros_like(F["NablaB"]), A=np.zeros_like(F["NablaB"]),
                 DivA=np.zeros_like(F["U"]))
    Ff = {k: np.ascontiguousarray(v).reshape(-1) for k, v in F.items()}
    dt, mu = 0.010 / 6.58211928, 0.5
    s.grad_desc(dt, mu)
    out = s.download_states()
    e, _ = s.energies()
    ref_psi, _, ref_e, _ = graddesc(psi, der, lap, np.zeros_like(psi), qn, trc, kin, Ff, dt,
                                    mu, dv)
    assert relerr(e, ref_e) < tol
    assert relerr(out, ref_psi.reshape(M, 4, N)) < (1e-10 if prec == mocca_amd.FP64 else tol)
    s.close()


@pytest.mark.parametrize("trc", [True, False])
@pytest.mark.parametrize("prec,tol_e,tol_psi", [(mocca_amd.FP64, 1e-11, 1e-10),
                                                (mocca_amd.MIXED, 2e-5, 2e-5)])
def test_step_general_w_vs_oracle(trc, prec, tol_e, tol_psi):
    """The step kernels (pipelined FP64, register-queue FP32) read only
    W[0][1], W[0][2], W[1][2] when the uploaded spin-orbit field is
    antisymmetric with a zero diagonal (as without tensor terms), all 9
    components otherwise.  Both against the oracle on a 32 x 20 x 13 mesh:
    the antisymmetric Woods-Saxon W, and the same W plus a random symmetric
    part.  FP64: states after the step and Gram-Schmidt 1e-10, energies
    1e-11; mixed 2e-5."""
    nx, ny, nz, dx, M = 32, 20, 13, 0.8, 10
    N = nx * ny * nz
    dv = dx ** 3 * 8
    rng = np.random.default_rng(77)
    qn = workloads.balanced_qn(M, time_reversal=trc)
    psi = workloads.plane_wave_states(rng, nx, ny, nz, dx, M)
    psi /= np.sqrt((psi.reshape(M, -1) ** 2).sum(1) * dv)[:, None, None]
    F0 = workloads.woods_saxon_fields(nx, ny, nz, dx, time_odd=not trc)
    kin = np.array([workloads.HBM_HALF, workloads.HBM_HALF])
    mesh = Mesh(nx, ny, nz, dx, False)
    der = np.zeros((M, 3, 4, N))
    lap = np.zeros((M, 4, N))
    for i, q in enumerate(qn):
        der[i], lap[i] = mesh.derive_spinor(psi[i], q[1], q[2], q[3])
    dt, mu = 0.010 / 6.58211928, 0.5
    for general in (False, True):
        F = dict(F0)
        if general:
            sym = rng.standard_normal((2, 3, 3, N)) * 0.3
            F["W"] = F0["W"] + sym + sym.transpose(0, 2, 1, 3)
        s = mocca_amd.Solve(nx, ny, nz, dx, dv, M, time_reversal=trc, precision=prec)
        s.set_states_meta(qn, np.ones(M))
        s.upload_states(psi)
        if trc:
            s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], kin)
            F.update(S=np.zeros_like(F["NablaB"]), A=np.zeros_like(F["NablaB"]),
                     DivA=np.zeros_like(F["U"]))
        else:
            s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], kin,
                             S=F["S"], A=F["A"], DivA=F["DivA"])
        Ff = {k: np.ascontiguousarray(v).reshape(-1) for k, v in F.items()}
        s.grad_desc(dt, mu)
        out = s.download_states()
        e, _ = s.energies()
        ref_psi, _, ref_e, _ = graddesc(psi, der, lap, np.zeros_like(psi), qn, trc, kin, Ff,
                                        dt, mu, dv)
        assert relerr(e, ref_e) < tol_e, general
        assert relerr(out, ref_psi.reshape(M, 4, N)) < tol_psi, general
        s.close()
