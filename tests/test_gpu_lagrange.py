"""Lagrange meshes with 16..64 points per axis take their derivatives from
the matrix cores (k_lag_axis: each line's first and second derivative as a
small GEMM with the folded Lagrange matrix of the component's reflection
sign), then hPsi + the heavy-ball step + Gram-Schmidt.  Checked against the
oracle's Lagrange derivatives (the reference's Opt_LAG*_EV8 / Lapla_LAG_EV8
order of operations) on a non-cubic mesh (24 x 20 x 16: axes padded to 32,
32 and 16 points), with time-odd fields and mixed parities and signatures.

Tolerances (max-norm relative to the largest reference value): FP64 energies
1e-11, states after the step and Gram-Schmidt 1e-10; mixed precision (FP32
states, FP64 matrix-core derivatives) 2e-5.  The single-state hPsi entry point
(on-the-fly derivatives) is checked at the same size, 1e-11.
"""
import numpy as np
import pytest

import mocca_amd
from mocca_amd import workloads
from oracle_lib import Mesh, graddesc, hpsi as or_hpsi, relerr

pytestmark = pytest.mark.gpu


def _case(seed=11, nx=24, ny=20, nz=16, dx=0.8, M=12):
    rng = np.random.default_rng(seed)
    qn = workloads.balanced_qn(M, time_reversal=False)
    psi = workloads.plane_wave_states(rng, nx, ny, nz, dx, M)
    dv = dx**3 * 8
    psi /= np.sqrt((psi.reshape(M, -1) ** 2).sum(1) * dv)[:, None, None]
    F = workloads.woods_saxon_fields(nx, ny, nz, dx, time_odd=True)
    kin = np.array([workloads.HBM_HALF, workloads.HBM_HALF])
    return dict(nx=nx, ny=ny, nz=nz, dx=dx, dv=dv, M=M, qn=qn, psi=psi, F=F, kin=kin,
                occ=rng.uniform(0.5, 1.0, M), dt=0.010 / 6.58211928, mu=0.5)


@pytest.mark.parametrize("prec,tol_e,tol_psi", [(mocca_amd.FP64, 1e-11, 1e-10),
                                                (mocca_amd.MIXED, 2e-5, 2e-5)])
def test_lagrange_matrix_core_graddesc_vs_oracle(prec, tol_e, tol_psi):
    c = _case()
    nx, ny, nz, M = c["nx"], c["ny"], c["nz"], c["M"]
    N = nx * ny * nz
    s = mocca_amd.Solve(nx, ny, nz, c["dx"], c["dv"], M, time_reversal=False, lagrange=True,
                        precision=prec)
    s.set_states_meta(c["qn"], c["occ"])
    s.upload_states(c["psi"])
    F = c["F"]
    s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], c["kin"],
                     S=F["S"], A=F["A"], DivA=F["DivA"])
    mesh = Mesh(nx, ny, nz, c["dx"], True)
    der = np.zeros((M, 3, 4, N))
    lap = np.zeros((M, 4, N))
    for i, q in enumerate(c["qn"]):
        der[i], lap[i] = mesh.derive_spinor(c["psi"][i], q[1], q[2], q[3])
    Ff = {k: np.ascontiguousarray(v).reshape(-1) for k, v in F.items()}
    if prec == mocca_amd.FP64:
        for i in (0, 5, M - 1):
            it = (c["qn"][i][0] + 3) // 2 - 1
            ref = or_hpsi(c["psi"][i], der[i], lap[i], it, False, c["kin"][it], Ff, N)
            assert relerr(s.hpsi(i), ref) < 1e-11
    s.grad_desc(c["dt"], c["mu"])
    psi = s.download_states()
    e, _ = s.energies()
    ref_psi, _, ref_e, _ = graddesc(c["psi"], der, lap, np.zeros_like(c["psi"]), c["qn"],
                                    False, c["kin"], Ff, c["dt"], c["mu"], c["dv"])
    assert relerr(e, ref_e) < tol_e
    assert relerr(psi, ref_psi.reshape(M, 4, N)) < tol_psi
    # a second step uses the heavy-ball memory of the first
    s.grad_desc(c["dt"], c["mu"])
    assert np.all(np.isfinite(s.download_states()))
    s.close()
