"""Pin the C restatement (oracle) against the reference's own outputs
(golden vectors written by the reference modules) and against closed forms.

Tolerances: the oracle keeps the reference's operation order, so agreement
is at rounding level; 1e-12 relative (max-norm) is asserted throughout.
"""
import numpy as np
import pytest

import oracle_lib as ol
from oracle_lib import (GOLDEN_CASES, Mesh, density, density_derived,
                        fields_of, golden, gram_schmidt, graddesc, hpsi,
                        kin_coefs, relerr)

TOL = 1e-12


def _mesh(g):
    m = g["meta"]
    return Mesh(m["nx"], m["ny"], m["nz"], m["dx"], m["lagrange"])


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_derivatives_vs_reference(case):
    g = golden(case)
    mesh = _mesh(g)
    for i, (iso, P_, S_, TS_) in enumerate(g["qn"]):
        der, lap = mesh.derive_spinor(g["psi"][i], P_, S_, TS_)
        assert relerr(der, g["der"][i].reshape(3, 4, -1)) < TOL
        assert relerr(lap, g["lap"][i].reshape(4, -1)) < TOL


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_hpsi_vs_reference(case):
    g = golden(case)
    N = _mesh(g).N
    F = fields_of(g)
    kin = kin_coefs(g)
    trc = g["meta"]["trc"]
    for i, q in enumerate(g["qn"]):
        it = (q[0] + 3) // 2 - 1
        h = hpsi(g["psi"][i], g["der"][i], g["lap"][i], it, trc, kin[it], F, N)
        assert relerr(h, g["hpsi"][i].reshape(4, -1)) < TOL


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_densities_vs_reference(case):
    g = golden(case)
    mesh = _mesh(g)
    N = mesh.N
    trc = g["meta"]["trc"]
    rho, tau, nj, vj, vs = density(g["psi"], g["der"], g["qn"], g["occ"], trc, N)
    assert relerr(rho, g["rho"].reshape(2, N)) < TOL
    assert relerr(tau, g["tau"].reshape(2, N)) < TOL
    assert relerr(nj, g["nablaJ"].reshape(2, N)) < TOL
    derrho, laprho, rotj, rots, ders = density_derived(mesh, rho, vj, vs, trc)
    assert relerr(derrho, g["derrho"].reshape(2, 3, N)) < TOL
    assert relerr(laprho, g["laprho"].reshape(2, N)) < TOL
    if not trc:
        assert relerr(vj, g["vecj"].reshape(2, 3, N)) < TOL
        assert relerr(vs, g["vecs"].reshape(2, 3, N)) < TOL
        assert relerr(rotj, g["rotvecj"].reshape(2, 3, N)) < TOL
        assert relerr(rots, g["rots"].reshape(2, 3, N)) < TOL
        assert relerr(ders, g["ders"].reshape(2, 3, 3, N)) < TOL


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_graddesc_gram_schmidt_vs_reference(case):
    """Two GradDesc calls (step + energy-ordered Gram-Schmidt) with the
    heavy-ball memory carried over, as the reference's driver did."""
    g = golden(case)
    mesh = _mesh(g)
    m = g["meta"]
    N = mesh.N
    F = fields_of(g)
    kin2 = kin_coefs(g)
    hbar = g["consts"][5]
    dv = g["consts"][6]
    dth = m["dt"] / hbar
    psi = g["psi"].reshape(m["nwt"], 4, N)
    upd = np.zeros_like(psi)
    qn = g["qn"]
    for step in (1, 2):
        der = np.zeros((m["nwt"], 3, 4, N))
        lap = np.zeros((m["nwt"], 4, N))
        for i, q in enumerate(qn):
            der[i], lap[i] = mesh.derive_spinor(psi[i], q[1], q[2], q[3])
        psi, upd, e, d = graddesc(psi, der, lap, upd, qn, m["trc"], kin2, F, dth,
                                  m["momentum"], dv)
        ref = g[f"step{step}_psi"].reshape(m["nwt"], 4, N)
        assert relerr(psi, ref) < 1e-11
        assert relerr(e, g[f"step{step}_ed"][:, 0]) < 1e-12
        assert relerr(d, g[f"step{step}_ed"][:, 1]) < 1e-9


# ---------------------------------------------------------------- closed forms
@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_angular_momentum_vs_reference(case):
    """UpdateAM with DiagSpin/DiagAng of every state (and the totals over the
    occupied states when time reversal is broken) against the reference's
    own values for the same states and densities."""
    g = golden(case)
    m = g["meta"]
    mesh = Mesh(m["nx"], m["ny"], m["nz"], m["dx"], m["lagrange"])
    dens = None if m["trc"] else dict(vecj=g["vecj"], vecs=g["vecs"])
    st, tot, iso = ol.angular_momentum(mesh, g["psi"].reshape(m["nwt"], 4, -1), g["qn"],
                                       g["occ"], g["consts"][6], m["trc"], dens)
    assert relerr(st, g["am_states"]) < 1e-13
    if m["trc"]:
        assert not np.any(g["am_totals"]) and not np.any(tot)
    else:
        assert relerr(tot, g["am_totals"]) < 1e-13
        assert relerr(iso, g["am_isoblock"]) < 1e-13
        # the two routes to J_z agree: sum of the states and the integral of
        # the spin and current densities
        assert abs(tot[0, 2] - tot[4, 2]) < 1e-10 * abs(tot[0, 2])


def test_lagrange_exact_on_band_limited_functions():
    """Lagrange-mesh derivatives are exact for the trigonometric functions of
    their basis (Baye & Heenen 1986): with the csc-type coefficients the basis
    on the full line of length L = 2 n dx is anti-periodic, i.e. spanned by
    cos/sin((2k+1) pi x / L); even (cos) and odd (sin) functions exercise the
    two reflection signs of the octant storage."""
    n, dx = 10, 0.8
    mesh = Mesh(n, n, n, dx, True)
    L = 2 * n * dx
    x = (np.arange(n) + 0.5) * dx
    for k in (0, 1, 3):
        kk = np.pi * (2 * k + 1) / L
        for axis in range(3):
            shp = [1, 1, 1]
            shp[2 - axis] = n
            even = np.broadcast_to(np.cos(kk * x).reshape(shp), (n, n, n))
            odd = np.broadcast_to(np.sin(kk * x).reshape(shp), (n, n, n))
            d_even = mesh.derive_field(even, axis, +1)
            d_odd = mesh.derive_field(odd, axis, -1)
            assert np.allclose(d_even, (-kk * np.sin(kk * x).reshape(shp)
                                        * np.ones((n, n, n))).reshape(-1), atol=1e-10)
            assert np.allclose(d_odd, (kk * np.cos(kk * x).reshape(shp)
                                       * np.ones((n, n, n))).reshape(-1), atol=1e-10)
    # Laplacian: exact on products of basis functions, one sign per axis.
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    k = [np.pi * (2 * a + 1) / L for a in (0, 2, 3)]
    f = np.cos(k[0] * X) * np.cos(k[1] * Y) * np.sin(k[2] * Z)
    lap = mesh.laplace_field(f, 1, 1, -1)
    assert np.allclose(lap, (-(k[0]**2 + k[1]**2 + k[2]**2) * f).reshape(-1),
                       atol=1e-11)


def test_fd_stencil_orders_on_polynomials():
    """Away from the box edge, the 7-point first derivative is exact for
    polynomials of degree <= 6 and the 9-point second derivative for degree
    <= 8; the reflection ghost makes x^2 even / x odd behave correctly at the
    symmetry plane."""
    n, dx = 16, 0.5
    mesh = Mesh(n, n, n, dx, False)
    x = (np.arange(n) + 0.5) * dx
    f_even = np.broadcast_to((x**6 - 3 * x**4 + x**2).reshape(1, 1, n), (n, n, n))
    d = mesh.derive_field(f_even, 0, +1).reshape(n, n, n)[0, 0]
    exact = 6 * x**5 - 12 * x**3 + 2 * x
    assert np.allclose(d[:n - 3], exact[:n - 3], rtol=1e-10, atol=1e-9)
    f_odd = np.broadcast_to((x**5 - x).reshape(1, n, 1), (n, n, n))
    d = mesh.derive_field(f_odd, 1, -1).reshape(n, n, n)[:, :, 0][0]
    assert np.allclose(d[:n - 3], (5 * x**4 - 1)[:n - 3], rtol=1e-10, atol=1e-9)
    f8 = np.broadcast_to((x**8).reshape(n, 1, 1), (n, n, n))
    lap = mesh.laplace_field(f8, 1, 1, 1).reshape(n, n, n)[:, 0, 0]
    assert np.allclose(lap[:n - 4], (56 * x**6)[:n - 4], rtol=1e-9, atol=1e-7)


def test_gram_schmidt_is_orthonormal_and_triangular():
    """QR invariants: Q^T Q = I within each block, and the span of the first
    k states (energy order) is preserved."""
    rng = np.random.default_rng(7)
    M, N, dv = 10, 300, 0.4
    qn = np.array([[-1, 1, 1, 1]] * 6 + [[1, 1, 1, 1]] * 4, np.int32)
    psi = rng.standard_normal((M, 4, N))
    e = rng.standard_normal(M)
    q = gram_schmidt(psi, qn, e, dv).reshape(M, -1)
    for iso in (-1, 1):
        idx = [i for i in range(M) if qn[i, 0] == iso]
        idx.sort(key=lambda i: e[i])
        Q = q[idx]
        assert np.allclose(Q @ Q.T * dv, np.eye(len(idx)), atol=1e-13)
        A = psi.reshape(M, -1)[idx]
        R = Q @ A.T * dv            # Q^T A in the dv inner product
        assert np.allclose(np.tril(R, -1), 0, atol=1e-10)
        assert np.all(np.diag(R) > 0)


def test_planted_sign_error_is_caught():
    """A wrong ghost sign (planted in the oracle input) must break agreement
    with the reference's derivative -- the fixture is not blind to it."""
    g = golden("ev8_trc_fd")
    mesh = _mesh(g)
    iso, P_, S_, TS_ = g["qn"][0]
    der, _ = mesh.derive_spinor(g["psi"][0], -P_, S_, TS_)
    assert relerr(der, g["der"][0].reshape(3, 4, -1)) > 1e-6
