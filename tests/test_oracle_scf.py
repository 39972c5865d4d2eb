"""Pin the self-consistent part of the C restatement (moments, Coulomb,
mean-field potentials, energy, density mixing, HF filling, and the whole
iteration loop) against the reference's own outputs, and against what the
mathematics fixes independently of the reference (spherical harmonics from
scipy, the potential of a Gaussian charge, a brute-force filling).

Tolerances: rounding level (1e-12) where the oracle keeps the reference's
operation order; 1e-11 for the Coulomb solution, whose conjugate-gradient
iterate is only defined to the solver's stopping residual; energies along an
iteration trajectory to 1e-9 MeV absolute.
"""
import numpy as np
import pytest
from scipy import special

import oracle_lib as ol
from oracle_lib import GOLDEN_CASES, Mesh, golden, relerr

TOL = 1e-12
SCF_CASES = ["scf_trc_fd", "scf_trc_lag", "scf_todd_fd"]


def _mesh(g):
    m = g["meta"]
    return Mesh(m["nx"], m["ny"], m["nz"], m["dx"], m["lagrange"])


def _dens(g, prefix="pot_"):
    keys = ol.DENSITY_KEYS + (() if g["meta"]["trc"] else ol.DENSITY_KEYS_ODD)
    return {k: g[prefix + k] for k in keys}


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_moments_vs_reference(case):
    g = golden(case)
    lm, val = ol.moments(g["pot_rho"], _mesh(g), g["consts"][6])
    assert np.array_equal(lm, g["pot_moment_lm"][:, :2])
    assert not g["pot_moment_lm"][:, 2].any()      # real parts only in EV8
    assert relerr(val, g["pot_moment_val"]) < TOL


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_coulomb_vs_reference(case):
    """First SolveCoulomb call (initial guess + multipole boundary + CG)."""
    g = golden(case)
    mesh, c = _mesh(g), g["consts"]
    cs = ol.Coulomb(mesh, c[4], c[7] + c[8], c[8])
    vc, vx = cs.solve(g["pot_rho"][1], g["pot_moment_lm"][:, :2], g["pot_moment_val"][:, 1])
    assert 0 < cs.iters < 500
    assert relerr(vc, g["pot_vcoul"].reshape(-1)) < 1e-11
    assert relerr(cs.cpot, g["pot_cpot"].reshape(-1)) < 1e-11
    assert relerr(vx, g["pot_vcoulx"].reshape(-1)) < TOL


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_potentials_vs_reference(case):
    g = golden(case)
    trc = g["meta"]["trc"]
    F = ol.potentials(g["coef"], _mesh(g), trc, _dens(g), g["pot_vcoul"], g["pot_vcoulx"])
    pairs = [("U", "Upot"), ("B", "Bpot"), ("NablaB", "NBpot"), ("W", "Wpot"),
             ("DivW", "DWpot")]
    if not trc:
        pairs += [("S", "Spot"), ("A", "Apot"), ("DivA", "DApot")]
    for k, gk in pairs:
        assert relerr(F[k].reshape(-1), g["pot_" + gk].reshape(-1)) < TOL, k


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_energy_vs_reference(case):
    g = golden(case)
    m, c = g["meta"], g["consts"]
    ks = ol.kinetic_sums(g["psi"], g["lap"], g["qn"], g["occ"], c[6])
    E = ol.energy(g["coef"], c, m["com1body"], m["trc"], _mesh(g), _dens(g),
                  g["pot_vcoul"], ks)
    ge = g["pot_energy"]   # total, kin(2), E_C, E_Cx, CoM(2,2) (Fortran order), terms
    assert not ge[6] and not ge[8]                 # no two-body c.o.m. term
    ref = np.concatenate([ge[0:6], [ge[7]], ge[9:]])
    assert np.max(np.abs(E - ref) / np.maximum(np.abs(ref), 1e-300)) < TOL


@pytest.mark.parametrize("case", SCF_CASES)
def test_scf_trajectory_vs_reference(case):
    """The whole loop: GradDesc, HF filling, densities, mixing, moments,
    Coulomb, potentials, energy; every iteration's energy decomposition."""
    g = golden(case)
    r = ol.scf_run(g, g["meta"]["niter"])
    assert r["traj"].shape == g["traj"].shape
    assert np.max(np.abs(r["traj"][:, 0] - g["traj"][:, 0])) < 1e-9
    assert relerr(r["traj"], g["traj"]) < 1e-12
    assert relerr(r["psi"].reshape(-1), g["final_psi"].reshape(-1)) < 1e-11
    assert relerr(r["energies"], g["final_eo"][:, 0]) < 1e-11
    assert np.array_equal(r["occ"], g["final_eo"][:, 1])
    assert relerr(r["fields"]["U"].reshape(-1), g["final_Upot"].reshape(-1)) < 1e-11
    assert relerr(r["cpot"], g["final_cpot"].reshape(-1)) < 1e-11
    # the energy goes down along the imaginary-time trajectory
    assert g["traj"][-1, 0] < g["traj"][0, 0]


@pytest.mark.parametrize("case", ["scfest_trc_fd", "scfest_todd_lag"])
def test_scf_estimation_trajectory_vs_reference(case):
    """ParameterEstimation = 1: IterativeEstimation before every GradDesc
    (its first call started from the seed the reference run used).  The dt
    and momentum of every iteration and the trajectory match the reference."""
    g = golden(case)
    r = ol.scf_run(g, g["meta"]["niter"])
    assert relerr(r["params"], g["params"]) < 1e-11
    assert np.max(np.abs(r["traj"][:, 0] - g["traj"][:, 0])) < 1e-9
    assert relerr(r["traj"], g["traj"]) < 1e-11
    assert relerr(r["psi"].reshape(-1), g["final_psi"].reshape(-1)) < 1e-10
    # dt grows from the first estimate (relE = 10^4 MeV while every energy is
    # still 0) and the momentum stays in (0, 1)
    assert g["params"][0, 0] < g["params"][1, 0]
    assert np.all((g["params"][:, 1] > 0) & (g["params"][:, 1] < 1))


@pytest.mark.parametrize("case", ["scfpc_trc_fd", "scfpc_todd_lag"])
def test_scf_potential_preconditioning_vs_reference(case):
    """MixingScheme 2: no density mixing; U (and S without time reversal)
    <- old + (1 - Lap)^-1 (new - old) by the reference's CG."""
    g = golden(case)
    r = ol.scf_run(g, g["meta"]["niter"])
    assert np.max(np.abs(r["traj"][:, 0] - g["traj"][:, 0])) < 1e-9
    assert relerr(r["traj"], g["traj"]) < 1e-11
    assert relerr(r["psi"].reshape(-1), g["final_psi"].reshape(-1)) < 1e-10
    assert relerr(r["fields"]["U"].reshape(-1), g["final_Upot"].reshape(-1)) < 1e-11
    if not g["meta"]["trc"]:
        assert relerr(r["fields"]["S"].reshape(-1), g["final_Spot"].reshape(-1)) < 1e-10


@pytest.mark.parametrize("case", ["scfcon_trc_fd", "scfcon_todd_lag"])
def test_scf_constraints_vs_reference(case):
    """Multipole constraints: the AlternateStep before GradDesc (whose hPsi
    then uses the derivatives of the states before that step, as in the
    reference), the constraint field in U, readjusted targets, estimated
    intensities and the switch-off after the given iteration."""
    g = golden(case)
    r = ol.scf_run(g, g["meta"]["niter"])
    assert np.max(np.abs(r["traj"][:, 0] - g["traj"][:, 0])) < 1e-9
    assert relerr(r["traj"], g["traj"]) < 1e-12
    assert relerr(r["psi"].reshape(-1), g["final_psi"].reshape(-1)) < 1e-11
    assert relerr(r["fields"]["U"].reshape(-1), g["final_Upot"].reshape(-1)) < 1e-11


@pytest.mark.parametrize("case", GOLDEN_CASES)
def test_rutz_cutoff_vs_reference(case):
    g = golden(case)
    m = g["meta"]
    mesh = ol.Mesh(m["nx"], m["ny"], m["nz"], m["dx"], m["lagrange"])
    cut = ol.rutz_cutoff(mesh, g["rho"], g["consts"][6])
    assert relerr(cut.reshape(-1), g["pot_cutoff"]) < 1e-14


def test_scf_hfb_vs_reference():
    """HFB pairing (the sample input's pairing): the pairing field from kappa,
    gaps, the Broyden Fermi solver over the signature-reduced HFB matrices,
    the canonical basis as density basis, rho/kappa mixing, pairing energy."""
    g = golden("scfhfb_trc_fd")
    r = ol.scf_run(g, g["meta"]["niter"])
    assert np.max(np.abs(r["traj"][:, 0] - g["traj"][:, 0])) < 1e-9
    assert relerr(r["traj"], g["traj"]) < 1e-11
    assert np.max(np.abs(r["pairing"] - g["pairing"])) < 1e-9       # Fermi, E_pair
    assert relerr(r["psi"].reshape(-1), g["final_psi"].reshape(-1)) < 1e-11
    assert np.all(g["pairing"][:, 2:] < -0.5)                        # pairing stays on


def test_hfb_start_matrices_vs_reference():
    """The start-up SolvePairing: gaps Delta, mixed kappa and rho of every
    (parity, isospin) block against the reference's arrays."""
    g = golden("scfhfb_trc_fd")
    m = g["meta"]
    M, N = m["nwt"], m["nx"] ** 3
    h = ol.HFB(g["qn"], True, m)
    h.guess()
    h.fermi = np.array(g["start_fermi"], float)
    h.solve(g["psi"].reshape(M, 4, N), g["start_energies"], np.zeros((2, N)), g["consts"][6])
    for (P, it) in h.keys:
        n = len(h.blocks[(P, it)])
        for name, mine in (("delta", h.delta), ("kappa", h.kappa), ("rho", h.rho)):
            ref = g["start_" + name][it - 1, P - 1][:n, :n].T      # Fortran (i, j) order
            assert np.max(np.abs(mine[(P, it)] - ref)) < 1e-11 * max(1.0, np.abs(ref).max())


def test_sample_input_vs_reference():
    """The reference's README sample (24Mg, Lagrange mesh, IterativeEstimation,
    MixingScheme 2, HFB with GuessKappa, a Q20 constraint for 10 iterations):
    all 100 iterations, energy to 1e-9 of the run's energy scale, the final
    one to 1e-10 relative."""
    g = golden("sample_mg24")
    r = ol.scf_run(g, g["meta"]["niter"])
    diff = np.abs(r["traj"][:, 0] - g["traj"][:, 0])
    assert (diff / np.maximum(np.abs(g["traj"][:, 0]), 100.0)).max() < 1e-9
    assert diff[-1] / abs(g["traj"][-1, 0]) < 1e-10
    assert np.max(np.abs(r["pairing"] - g["pairing"])) < 1e-8
    assert relerr(r["params"], g["params"]) < 1e-9
    assert relerr(r["psi"].reshape(-1), g["final_psi"].reshape(-1)) < 1e-9
    assert g["traj"][-1, 0] < -190.0             # bound 24Mg after 100 iterations


@pytest.mark.parametrize("case", ["scf_trc_fd", "scf_trc_lag", "scfcon_trc_fd",
                                  "scfcon_todd_lag", "scfhfb_trc_fd", "sample_mg24"])
def test_convergence_tests_and_final_iteration_vs_reference(case):
    """ConvergenceCheck's tests after the start-up and every iteration, at
    the precisions the reference driver records, flag for flag; the energy
    row of FinalIteration (Lagrange derivatives, unmixed densities) to 1e-12
    of the row's scale and the finite-difference densities after it."""
    g = golden(case)
    r = ol.scf_run(g, g["meta"]["niter"])
    assert np.array_equal(r["conv"], g["conv"])
    row, d = ol.final_iteration(g, r)
    assert np.max(np.abs(row - g["final_row"])) / np.max(np.abs(g["final_row"])) < 1e-12
    assert relerr(d["rho"].reshape(2, -1), g["final_iteration_rho"]) < 1e-12


def test_convergence_tests_pick_out_planted_mistakes():
    """The oracle's tests fail on each of the three mistakes a restatement
    could make: ConverEnergy's OldEnergy shift off by one row, and
    ConverFermi reading the last Fermi energies instead of the reference's
    zero history."""
    g = golden("sample_mg24")
    E = g["traj"][:, 0]
    n = len(E) - 1
    differs = False
    for k in range(9, n + 1):
        good = np.concatenate([[E[k]], E[k - 7:k][::-1]])       # E_k, then E_{k-1} .. E_{k-7}
        late = np.concatenate([[E[k]], E[k - 8:k - 1][::-1]])   # the shift one row late
        f1 = [ol.convergence_tests(good, [0, 0], False, None, 1, e, 1)[1]
              for e in ol.CONV_EPREC]
        f2 = [ol.convergence_tests(late, [0, 0], False, None, 1, e, 1)[1]
              for e in ol.CONV_EPREC]
        assert f1 == [bool(x) for x in g["conv"][k][:5]]
        differs = differs or f1 != f2
    assert differs
    good = np.concatenate([[E[n]], E[n - 7:n][::-1]])
    fermi = g["pairing"][-1, 0:2]
    ok = [ol.convergence_tests(good, fermi, True, None, 1, 1, p)[2] for p in ol.CONV_PPREC]
    assert ok == [bool(x) for x in g["conv"][-1][9:]]
    # a history that remembered the last Fermi energies would always pass
    assert not all(ok)


# ---- checks fixed by the mathematics ---------------------------------------

def test_hfb_generalized_density_is_a_projector():
    """For every lambda the quasiparticle vacuum's rho and kappa satisfy
    rho - rho^2 = kappa kappa^T (R^2 = R for the generalised density), the
    selected quasiparticle energies are positive and N(lambda) grows with
    lambda."""
    g = golden("scfhfb_trc_fd")
    m = g["meta"]
    M, N = m["nwt"], m["nx"] ** 3
    h = ol.HFB(g["qn"], True, m)
    h.guess()
    psi = g["psi"].reshape(M, 4, N)
    e = g["start_energies"]
    pc = h.cutoffs(e, np.array(g["start_fermi"]))
    h.gaps(psi, h.pairing_field(psi, pc, h.density_factor(np.zeros((2, N)))), pc, g["consts"][6])
    last = None
    for lam in np.linspace(30.0, 60.0, 7):
        n = h.number(np.array([lam, lam]), e)
        for k in h.keys:
            rho, kap = h.rho[k], h.kappa[k]
            assert np.max(np.abs(rho - rho @ rho - kap @ kap.T)) < 1e-12
            assert np.all(h.qpe[k] > 0)
        if last is not None:
            assert np.all(n >= last - 1e-12)
        last = n

def test_rutz_cutoff_of_a_spherical_density():
    """For rho = exp(-r^2/b^2) the equidensity surface at max/10 is the sphere
    r0 = b sqrt(ln 10 + ...) through the mesh maximum; DeltaR is the distance
    to it, |r - r0|, up to the linear interpolation and the surface sampling
    (a fraction of dx)."""
    n, dx, b = 14, 0.5, 2.5
    mesh = ol.Mesh(n, n, n, dx, 0)
    x = (np.arange(n) + 0.5) * dx
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z).reshape(-1)
    rho = np.exp(-(r / b) ** 2)
    r0 = b * np.sqrt(np.log(10.0) - (r.min() / b) ** 2)    # max of the mesh is at r.min()
    cut = ol.rutz_cutoff(mesh, np.stack([rho, 0.5 * rho]), 8 * dx ** 3, radd=4.0, acut=0.4)
    dr = 4.0 + 0.4 * np.log(1.0 / cut[0] - 1.0)            # invert the Fermi function
    inner = r < 0.8 * n * dx                                # away from the box faces
    assert np.max(np.abs(dr[inner] - (r[inner] - r0))) < 0.6 * dx
    assert np.array_equal(cut[0], cut[1])                   # scale-free threshold
    # no particles: cutoff 0
    assert not ol.rutz_cutoff(mesh, np.zeros((2, mesh.N)), 8 * dx ** 3).any()

@pytest.mark.parametrize("lag", [0, 1])
def test_precondition_solves_the_screened_equation(lag):
    """The CG of PreconditionPotential solves (1 - Lap) x = f to its stopping
    rule dv |r|^2 < 1e-8: checked by applying the operator to x, and against
    a dense solve of the same operator (assembled column by column)."""
    n, dx = 5, 0.8
    mesh = ol.Mesh(n, n, n, dx, lag)
    N, dv = n ** 3, 8 * dx ** 3
    rng = np.random.default_rng(5 + lag)
    f = rng.uniform(-3, 3, (2, N))
    for signs in [ol.PRECON_SIGNS_U] + list(ol.PRECON_SIGNS_S):
        x, iters = ol.precondition(mesh, f, signs, dv)
        assert (iters > 2).all() and (iters < 300).all()
        A = np.zeros((N, N))
        for j in range(N):
            e = np.zeros(N)
            e[j] = 1.0
            A[:, j] = e - mesh.laplace_field(e, *signs)
        for it in range(2):
            r = f[it] - A @ x[it]
            assert np.sum(r * r) * dv < 1e-8
            exact = np.linalg.solve(A, f[it])
            assert np.max(np.abs(x[it] - exact)) < 1e-3
        # the screened operator is symmetric positive definite
        assert np.allclose(A, A.T, atol=1e-10)
        assert np.linalg.eigvalsh(A).min() > 0.99

def test_estimation_finds_largest_eigenvalue():
    """The power iteration of IterativeEstimation converges to the largest
    eigenvalue of h, from below (a Rayleigh quotient): the dense matrix of h
    on a 5^3 octant is built column by column and diagonalised.  dt and the
    momentum are the heavy-ball optimum for the spectrum [relE, maxE - minh]:
    dt/hbar (sqrt(maxE') + sqrt(relE))^2 = 4 * 0.9, momentum =
    ((1 - sqrt(k)) / (1 + sqrt(k)))^2 with k = relE / maxE'."""
    n, dx = 5, 0.8
    mesh = ol.Mesh(n, n, n, dx, 0)
    N = n ** 3
    x = (np.arange(n) + 0.5) * dx
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    f = 1 / (1 + np.exp((r - 3) / 0.6))
    df = -f * (1 - f) / 0.6
    grad = [df * X / r, df * Y / r, df * Z / r]
    F = dict(U=np.stack([-50 * f, -45 * f]).reshape(-1), B=np.stack([10 * f, 11 * f]).reshape(-1),
             NablaB=np.stack([[10 * g for g in grad], [11 * g for g in grad]]).reshape(-1))
    for k, cnt in (("W", 18), ("DivW", 6), ("S", 6), ("A", 6), ("DivA", 2)):
        F[k] = np.zeros(cnt * N)
    kin, hbar, dv = 20.7355191, 6.58211928, 8 * dx ** 3
    qe = np.array([-1, 1, 1, 1], np.int32)
    H = np.zeros((4 * N, 4 * N))
    for j in range(4 * N):
        e = np.zeros(4 * N)
        e[j] = 1
        der, lap = mesh.derive_spinor(e, 1, 1, 1)
        H[:, j] = ol.hpsi(e, der, lap, 0, 1, kin, F, N).reshape(-1)
    lmax = np.linalg.eigvals(H).real.max()
    psi = np.random.default_rng(3).uniform(0, 1, 4 * N)
    qn = np.array([[-1, 1, 1, 1], [-1, 1, 1, 1], [1, 1, 1, 1]], np.int32)
    energy = np.array([-30.0, -10.0, -40.0])      # relE 20 (neutrons), minh -40
    dt, mom, maxE, relE, con, iters = ol.estimate(mesh, psi, 1, qe, 1, kin, F, 0.01, 0.3, dv,
                                                  hbar, qn, np.array([1.0, 0.0, 1.0]), energy)
    assert maxE <= lmax * (1 + 1e-12) and lmax - maxE < 0.05
    assert abs(con) < 1e-2 and 1 < iters < 500
    assert relE == 20.0
    me = maxE + 40.0
    assert abs(dt / hbar * (np.sqrt(me) + np.sqrt(relE)) ** 2 - 4 * np.float32(0.9)) < 1e-12
    k = relE / me
    assert abs(mom - ((1 - np.sqrt(k)) / (1 + np.sqrt(k))) ** 2) < 1e-14
    assert abs(np.sum(psi * psi) * dv - 1) < 1e-12

@pytest.mark.parametrize("l,m", [(0, 0), (2, 0), (2, 2), (4, 0), (4, 2), (4, 4),
                                 (6, 4), (8, 6), (10, 10)])
def test_rlylm_vs_scipy(l, m):
    """r^l Re Y_lm with the Condon-Shortley phase, as scipy defines it."""
    rng = np.random.default_rng(l * 11 + m)
    for X, Y, Z in rng.uniform(0.1, 5.0, size=(20, 3)):
        r = np.sqrt(X * X + Y * Y + Z * Z)
        y = special.sph_harm_y(l, m, np.arccos(Z / r), np.arctan2(Y, X))
        assert abs(ol.rlylm(l, m, X, Y, Z) - r**l * y.real) < 1e-12 * max(1.0, r**l)


def _gaussian_charge_error(n, dx, a=2.0, Z=20.0, e2=1.43996446):
    mesh = Mesh(n, n, n, dx, False)
    x = (np.arange(n) + 0.5) * dx
    zz, yy, xx = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(xx**2 + yy**2 + zz**2)
    rho_p = Z / (np.pi**1.5 * a**3) * np.exp(-(r / a) ** 2)
    rho = np.stack([np.zeros_like(rho_p), rho_p]).reshape(2, -1)
    lm, val = ol.moments(rho, mesh, 8 * dx**3)
    assert abs(val[0, 1] * np.sqrt(4 * np.pi) - Z) < 1e-6 * Z     # Q00 = Z/sqrt(4 pi)
    cs = ol.Coulomb(mesh, e2, 2 * Z, Z)
    vc, _ = cs.solve(rho[1], lm, val[:, 1])
    return relerr(vc, (e2 * Z * special.erf(r / a) / r).reshape(-1))


def test_coulomb_gaussian_charge():
    """Potential of a Gaussian proton distribution, V = e2 Z erf(r/a)/r:
    fixes the source sign and scale, the Laplacian and the multipole
    boundary; the error falls at close to fourth order in dx, as the
    5-point stencil gives."""
    coarse = _gaussian_charge_error(20, 0.8)
    fine = _gaussian_charge_error(40, 0.4)
    assert coarse < 1.5e-3
    assert fine < coarse / 10


def test_hf_fill_brute_force():
    rng = np.random.default_rng(5)
    M = 14
    qn = np.array([[(-1) ** (i % 2), 1, 1, 1] for i in range(M)], np.int32)
    e = rng.normal(size=M)
    e[3] = e[5]                                     # a tie keeps index order
    for trc, (nn, zz) in [(True, (6, 4)), (False, (3, 5)), (True, (7, 5))]:
        occ = ol.hf_fill(e, qn, nn, zz, trc)
        for iso, count in ((-1, nn), (1, zz)):
            k = count // 2 if trc else count
            idx = [i for i in sorted(range(M), key=lambda i: (e[i], i)) if qn[i, 0] == iso][:k]
            want = np.zeros(M)
            want[idx] = 2.0 if trc else 1.0
            sel = qn[:, 0] == iso
            assert np.array_equal(occ[sel], want[sel])


def test_mix_linear():
    rng = np.random.default_rng(9)
    d, h = rng.normal(size=100), rng.normal(size=100)
    want = 0.25 * d + 0.75 * h
    ol.lib().or_mix(ol.P(d), ol.P(h), ol.ctypes.c_size_t(100), ol.ctypes.c_double(0.75))
    assert np.max(np.abs(d - want)) < 1e-15


def test_scf_hfb_blocks_of_100_states_vs_reference():
    """HFB blocks of 100 states (200 x 200 signature-reduced matrices): the
    reference's run of 200 neutron and 40 proton harmonic-oscillator states,
    4 iterations with pairing on throughout."""
    g = golden("scfhfb_large")
    r = ol.scf_run(g, g["meta"]["niter"])
    rel = np.abs(r["traj"][:, 0] - g["traj"][:, 0]) / np.abs(g["traj"][:, 0])
    assert rel.max() < 1e-9
    assert relerr(r["traj"], g["traj"]) < 1e-10
    assert np.max(np.abs(r["pairing"] - g["pairing"])) < 1e-8       # Fermi, E_pair
    assert np.all(np.abs(g["pairing"][:, 2:]) > 0.1)                 # pairing stays on


def test_scf_bcs_vs_reference():
    """BCS pairing (PairingType 1, PairingIter 50, FermiPrec 1e-12): the
    start-up step's gaps, cutoff factors, quasiparticle energies,
    occupations and Fermi energies, then 10 iterations against the
    reference run."""
    g = golden("scfbcs_trc_fd")
    m = g["meta"]
    M, N = m["nwt"], m["nx"] ** 3
    b = ol.BCS(g["qn"], m)
    b.fermi = np.array(g["start_fermi"], float)
    occ = b.solve(g["psi"].reshape(M, 4, N), g["start_energies"], np.zeros((2, N)),
                  g["consts"][6])
    ref = g["bcs_start"]
    for mine, theirs in ((b.delta, ref[0]), (b.pc, ref[1]), (b.eqp, ref[2]), (occ, ref[3])):
        assert np.max(np.abs(mine - theirs)) < 1e-12 * max(1.0, np.abs(theirs).max())
    assert np.max(np.abs(b.fermi - g["bcs_start_fermi"])) < 1e-12
    r = ol.scf_run(g, m["niter"])
    assert np.max(np.abs(r["traj"][:, 0] - g["traj"][:, 0])) < 1e-10
    assert relerr(r["traj"], g["traj"]) < 1e-12
    assert np.max(np.abs(r["pairing"] - g["pairing"])) < 1e-11       # Fermi, E_pair
    assert relerr(r["psi"].reshape(-1), g["final_psi"].reshape(-1)) < 1e-12
    assert np.max(np.abs(r["occ"] - g["final_eo"][:, 1])) < 1e-12
    assert np.all(g["pairing"][:, 2:] < 0.0)                         # pairing on
