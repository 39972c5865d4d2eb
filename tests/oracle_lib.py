"""ctypes view of the C restatement (oracle/liboracle.so) and the golden
fixtures.  Test infrastructure: the oracle is the checker, never the product.
"""
import ctypes
import json
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so missing: run `make -C oracle`")
        _LIB = ctypes.CDLL(path)
        _LIB.or_rlylm.restype = ctypes.c_double
    return _LIB


def P(a):
    """Pointer to a C-contiguous float64 / int32 numpy array (or NULL)."""
    if a is None:
        return None
    assert a.flags.c_contiguous
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def golden(name):
    z = np.load(os.path.join(GOLDEN, name + ".npz"), allow_pickle=False)
    d = {k: z[k] for k in z.files}
    d["meta"] = json.loads(str(d["meta"]))
    return d


GOLDEN_CASES = ["ev8_trc_fd", "ev8_trc_lag", "ev8_todd_fd", "ev8_todd_lag"]


class Mesh:
    """Mesh + derivative scheme, holding the Lagrange matrices if needed."""

    def __init__(self, nx, ny, nz, dx, lag, full=(2, 2, 2)):
        self.nx, self.ny, self.nz, self.dx, self.lag = nx, ny, nz, dx, int(lag)
        self.N = nx * ny * nz
        self.lagx = np.zeros(4 * nx * nx)
        self.lagy = np.zeros(4 * ny * ny)
        self.lagz = np.zeros(4 * nz * nz)
        lib().or_lag_build(nx, ny, nz, ctypes.c_double(dx), nx * full[0],
                           ny * full[1], nz * full[2], P(self.lagx),
                           P(self.lagy), P(self.lagz))

    def lagmats(self):
        return (P(self.lagx), P(self.lagy), P(self.lagz)) if self.lag else (None,) * 3

    def derive_spinor(self, psi, P_, S_, TS_):
        der = np.zeros((3, 4, self.N))
        lap = np.zeros((4, self.N))
        psi = np.ascontiguousarray(psi.reshape(4, self.N))
        lib().or_derive_spinor(P(psi), P(der), P(lap), self.nx, self.ny, self.nz,
                               ctypes.c_double(self.dx), int(P_), int(S_),
                               int(TS_), self.lag, *self.lagmats())
        return der, lap

    def derive_field(self, f, axis, sign):
        f = np.ascontiguousarray(f.reshape(self.N), dtype=np.float64)
        out = np.zeros(self.N)
        m = [self.lagx, self.lagy, self.lagz][axis]
        lib().or_derive_field(P(f), P(out), self.nx, self.ny, self.nz,
                              ctypes.c_double(self.dx), int(axis), int(sign),
                              self.lag,
                              P(m) if self.lag else None)
        return out

    def laplace_field(self, f, sx, sy, sz):
        f = np.ascontiguousarray(f.reshape(self.N), dtype=np.float64)
        out = np.zeros(self.N)
        lib().or_laplace_field(P(f), P(out), self.nx, self.ny, self.nz,
                               ctypes.c_double(self.dx), int(sx), int(sy),
                               int(sz), self.lag,
                               *self.lagmats())
        return out


def fields_of(g):
    N = g["meta"]["nx"] * g["meta"]["ny"] * g["meta"]["nz"]
    keys = ("U", "B", "NablaB", "W", "DivW", "S", "A", "DivA")
    return {k: np.ascontiguousarray(g[k].reshape(-1, N) if k in ("U", "B", "DivA")
                                    else g[k]).reshape(-1) for k in keys}


def kin_coefs(g):
    """hbm(it)/2 * reducedmass(it) from the reference's constants
    (MeanFields.f90:827-846)."""
    consts = g["consts"]
    hbm, mass = consts[0:2], consts[2:4]
    nn, zz = consts[7], consts[8]
    if g["meta"]["com1body"] == 2:
        red = 1.0 - mass / (nn * mass[0] + zz * mass[1])
    else:
        red = np.ones(2)
    return hbm / 2.0 * red


def hpsi(psi, der, lap, it, trc, kin, F, N):
    out = np.zeros((4, N))
    lib().or_hpsi(P(np.ascontiguousarray(psi.reshape(4, N))),
                  P(np.ascontiguousarray(der.reshape(3, 4, N))),
                  P(np.ascontiguousarray(lap.reshape(4, N))), int(N), int(it),
                  int(trc),
                  ctypes.c_double(kin), P(F["U"]), P(F["B"]), P(F["NablaB"]),
                  P(F["W"]), P(F["DivW"]), P(F["S"]), P(F["A"]), P(F["DivA"]),
                  P(out))
    return out


def density(psi, der, qn, occ, trc, N):
    M = psi.shape[0]
    rho, tau, nj = np.zeros((2, N)), np.zeros((2, N)), np.zeros((2, N))
    vj, vs = np.zeros((2, 3, N)), np.zeros((2, 3, N))
    lib().or_density(P(np.ascontiguousarray(psi.reshape(M, 4, N))),
                     P(np.ascontiguousarray(der.reshape(M, 3, 4, N))), M, N,
                     P(np.ascontiguousarray(qn, dtype=np.int32)),
                     P(np.ascontiguousarray(occ, dtype=np.float64)), int(trc),
                     P(rho), P(tau), P(nj), P(vj), P(vs))
    return rho, tau, nj, vj, vs


def density_derived(mesh, rho, vj, vs, trc):
    N = mesh.N
    derrho, laprho = np.zeros((2, 3, N)), np.zeros((2, N))
    rotj, rots, ders = np.zeros((2, 3, N)), np.zeros((2, 3, N)), np.zeros((2, 3, 3, N))
    lib().or_density_derived(P(rho), P(vj), P(vs), mesh.nx, mesh.ny, mesh.nz,
                             ctypes.c_double(mesh.dx), mesh.lag, *mesh.lagmats(),
                             int(trc), P(derrho), P(laprho), P(rotj), P(rots),
                             P(ders))
    return derrho, laprho, rotj, rots, ders


def gram_schmidt(psi, qn, energy, dv):
    M = psi.shape[0]
    N = psi[0].size // 4
    psi = np.ascontiguousarray(psi.reshape(M, 4, N)).copy()
    lib().or_gram_schmidt(P(psi), M, N, P(np.ascontiguousarray(qn, dtype=np.int32)),
                          P(np.ascontiguousarray(energy, dtype=np.float64)),
                          ctypes.c_double(dv))
    return psi


def graddesc(psi, der, lap, upd, qn, trc, kin2, F, dt_over_hbar, mom, dv):
    M = psi.shape[0]
    N = psi[0].size // 4
    psi = np.ascontiguousarray(psi.reshape(M, 4, N)).copy()
    upd = np.ascontiguousarray(upd.reshape(M, 4, N)).copy()
    e, d = np.zeros(M), np.zeros(M)
    lib().or_graddesc(P(psi), P(np.ascontiguousarray(der.reshape(M, 3, 4, N))),
                      P(np.ascontiguousarray(lap.reshape(M, 4, N))), P(upd), M, N,
                      P(np.ascontiguousarray(qn, dtype=np.int32)), int(trc),
                      P(np.ascontiguousarray(kin2, dtype=np.float64)),
                      P(F["U"]), P(F["B"]), P(F["NablaB"]), P(F["W"]),
                      P(F["DivW"]), P(F["S"]), P(F["A"]), P(F["DivA"]),
                      ctypes.c_double(dt_over_hbar), ctypes.c_double(mom),
                      ctypes.c_double(dv), P(e), P(d))
    return psi, upd, e, d


def relerr(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))


# ---- self-consistent part -------------------------------------------------

def rlylm(l, m, X, Y, Z):
    return lib().or_rlylm(int(l), int(m), ctypes.c_double(X), ctypes.c_double(Y),
                          ctypes.c_double(Z))


def moments(rho, mesh, dv, lmax=10):
    lm = np.zeros(2 * 64, np.int32)
    val = np.zeros(2 * 64)
    rho = np.ascontiguousarray(rho.reshape(2, mesh.N))
    cnt = lib().or_moments(P(rho), mesh.nx, mesh.ny, mesh.nz, ctypes.c_double(mesh.dx),
                           ctypes.c_double(dv), int(lmax), P(lm), P(val))
    return lm[:2 * cnt].reshape(cnt, 2), val[:2 * cnt].reshape(cnt, 2)


class Coulomb:
    """Persistent Coulomb state (CPotComplete) of SolveCoulomb."""

    def __init__(self, mesh, e2, nucleons, protons, maxit=500):
        self.mesh, self.e2, self.A, self.Z, self.maxit = mesh, e2, nucleons, protons, maxit
        self.cpot = np.zeros((mesh.nz + 2) * (mesh.ny + 2) * (mesh.nx + 2))
        self.first = True
        self.iters = 0

    def solve(self, rho_p, lm, qp):
        m = self.mesh
        vc, vx = np.zeros(m.N), np.zeros(m.N)
        lm = np.ascontiguousarray(lm, dtype=np.int32)
        qp = np.ascontiguousarray(qp, dtype=np.float64)
        self.iters = lib().or_coulomb(
            P(np.ascontiguousarray(rho_p.reshape(m.N))), m.nx, m.ny, m.nz,
            ctypes.c_double(m.dx), ctypes.c_double(self.e2), ctypes.c_double(self.A),
            ctypes.c_double(self.Z), len(qp), P(lm), P(qp), P(self.cpot),
            int(self.first), P(vc), P(vx), int(self.maxit))
        self.first = False
        return vc, vx


def potentials(coef, mesh, trc, d, vcoul, vcoulx, cexchange=2):
    """d: dict with rho, tau, nablaJ, derrho, laprho (+ vecj, vecs, rotvecj,
    rots when trc is False), shapes as in the reference dumps."""
    N = mesh.N
    c = lambda k, shp: (np.ascontiguousarray(d[k].reshape(shp)) if k in d and d[k] is not None
                        else None)
    out = {"U": np.zeros((2, N)), "B": np.zeros((2, N)), "NablaB": np.zeros((2, 3, N)),
           "W": np.zeros((2, 3, 3, N)), "DivW": np.zeros((2, 3, N)),
           "S": np.zeros((2, 3, N)), "A": np.zeros((2, 3, N)), "DivA": np.zeros((2, N))}
    lib().or_potentials(
        P(np.ascontiguousarray(coef)), mesh.nx, mesh.ny, mesh.nz, ctypes.c_double(mesh.dx),
        mesh.lag, *mesh.lagmats(), int(trc), P(c("rho", (2, N))), P(c("tau", (2, N))),
        P(c("nablaJ", (2, N))), P(c("derrho", (2, 3, N))), P(c("laprho", (2, N))),
        P(c("vecj", (2, 3, N))), P(c("vecs", (2, 3, N))), P(c("rotvecj", (2, 3, N))),
        P(c("rots", (2, 3, N))), P(np.ascontiguousarray(vcoul.reshape(N))),
        P(np.ascontiguousarray(vcoulx.reshape(N))), int(cexchange),
        P(out["U"]), P(out["B"]), P(out["NablaB"]), P(out["W"]), P(out["DivW"]),
        P(out["S"]), P(out["A"]), P(out["DivA"]))
    return out


def kinetic_sums(psi, lap, qn, occ, dv):
    M = psi.shape[0]
    N = psi[0].size // 4
    k = np.zeros(2)
    lib().or_kinetic_sums(P(np.ascontiguousarray(psi.reshape(M, 4, N))),
                          P(np.ascontiguousarray(lap.reshape(M, 4, N))), M, N,
                          P(np.ascontiguousarray(qn, dtype=np.int32)),
                          P(np.ascontiguousarray(occ, dtype=np.float64)),
                          ctypes.c_double(dv), P(k))
    return k


def energy(coef, consts, com1body, trc, mesh, d, vcoul, kin_sp, cexchange=2):
    """Returns [total, kin_n, kin_p, E_C, E_Cx, com_n, com_p, terms(36)]."""
    N = mesh.N
    c = lambda k, shp: (np.ascontiguousarray(d[k].reshape(shp)) if k in d and d[k] is not None
                        else None)
    out = np.zeros(43)
    lib().or_energy(P(np.ascontiguousarray(coef)), P(np.ascontiguousarray(consts)),
                    int(com1body), int(trc), mesh.nx, mesh.ny, mesh.nz,
                    P(c("rho", (2, N))), P(c("tau", (2, N))), P(c("nablaJ", (2, N))),
                    P(c("laprho", (2, N))), P(c("vecj", (2, 3, N))), P(c("vecs", (2, 3, N))),
                    P(c("rots", (2, 3, N))), P(np.ascontiguousarray(vcoul.reshape(N))),
                    P(np.ascontiguousarray(kin_sp, dtype=np.float64)), int(cexchange), P(out))
    return out


def hf_fill(energies, qn, neutrons, protons, trc):
    M = len(energies)
    occ = np.zeros(M)
    lib().or_hf_fill(P(np.ascontiguousarray(energies, dtype=np.float64)),
                     P(np.ascontiguousarray(qn, dtype=np.int32)), M,
                     ctypes.c_double(neutrons), ctypes.c_double(protons), int(trc), P(occ))
    return occ


DENSITY_KEYS = ("rho", "tau", "nablaJ", "derrho", "laprho")
DENSITY_KEYS_ODD = ("vecj", "rotvecj", "vecs", "rots")


def estimate(mesh, psi, start, qn_est, trc, kin, F, dt, mom, dv, hbar, qn, occ, energy):
    """IterativeEstimation on maxspwf psi (updated in place); returns
    (dt, momentum, maxE, relE, con, iterations)."""
    out = np.zeros(6)
    M = len(qn)
    lib().or_estimate(P(psi), int(start), mesh.nx, mesh.ny, mesh.nz, ctypes.c_double(mesh.dx),
                      mesh.lag, *mesh.lagmats(),
                      P(np.ascontiguousarray(qn_est, dtype=np.int32)), int(trc),
                      ctypes.c_double(kin), P(F["U"]), P(F["B"]), P(F["NablaB"]),
                      P(F["W"]), P(F["DivW"]), P(F["S"]), P(F["A"]), P(F["DivA"]),
                      ctypes.c_double(dt), ctypes.c_double(mom), ctypes.c_double(dv),
                      ctypes.c_double(hbar), M, P(np.ascontiguousarray(qn, dtype=np.int32)),
                      P(np.ascontiguousarray(occ, dtype=np.float64)),
                      P(np.ascontiguousarray(energy, dtype=np.float64)), P(out))
    return out


def precondition(mesh, pot, signs, dv, alpha=-1.0, beta=1.0):
    """PreconditionPotential: (beta + alpha Lap)^-1 pot for both isospins by
    the reference's CG; returns (x [2][N], iterations [2])."""
    x = np.zeros((2, mesh.N))
    iters = np.zeros(2, np.int32)
    lib().or_precondition(P(np.ascontiguousarray(pot, dtype=np.float64).reshape(2, mesh.N)),
                          P(x), mesh.nx, mesh.ny, mesh.nz, ctypes.c_double(mesh.dx), mesh.lag,
                          *mesh.lagmats(), int(signs[0]), int(signs[1]), int(signs[2]),
                          ctypes.c_double(alpha), ctypes.c_double(beta), ctypes.c_double(dv),
                          P(iters))
    return x, iters


# Ghost signs (x, y, z) of the Laplacian for the potential components that
# MixingScheme 2 preconditions (PotentialMixing.f90:143-198 with the
# component table of Lapla_EV8, OptimizedDerivatives.f90:533-542), for
# P = S = TS = 1: U; S_x, S_y, S_z; W_{mu nu} as [mu][nu].
PRECON_SIGNS_U = (1, 1, 1)
PRECON_SIGNS_S = ((-1, 1, -1), (1, -1, -1), (1, 1, 1))
PRECON_SIGNS_W = (((-1, -1, -1), (1, 1, -1), (1, -1, 1)),
                  ((1, 1, -1), (-1, -1, -1), (-1, 1, 1)),
                  ((1, -1, 1), (-1, 1, 1), (-1, -1, -1)))


def precondition_potentials(mesh, F, Fold, trc, dv, preconw):
    """The MixingScheme 2 step of ConstructPotentials (PotentialMixing.f90:
    143-198): U, S (and W when PreConW > 0) <- old + precond(new - old)."""
    N = mesh.N
    U = F["U"].reshape(2, N)
    x, _ = precondition(mesh, U - Fold["U"].reshape(2, N), PRECON_SIGNS_U, dv)
    F["U"] = (Fold["U"].reshape(2, N) + x).reshape(F["U"].shape)
    if not trc:
        S, So = F["S"].reshape(2, 3, N).copy(), Fold["S"].reshape(2, 3, N)
        for c in range(3):
            x, _ = precondition(mesh, S[:, c] - So[:, c], PRECON_SIGNS_S[c], dv)
            S[:, c] = So[:, c] + x
        F["S"] = S.reshape(F["S"].shape)
    if preconw > 0:
        W, Wo = F["W"].reshape(2, 3, 3, N).copy(), Fold["W"].reshape(2, 3, 3, N)
        for mu in range(3):
            for nu in range(3):   # W[it][nu][mu] holds Wpot(:,:,:,mu,nu,it)
                x, _ = precondition(mesh, W[:, nu, mu] - Wo[:, nu, mu],
                                    PRECON_SIGNS_W[mu][nu], dv)
                W[:, nu, mu] = Wo[:, nu, mu] + x
        F["W"] = W.reshape(F["W"].shape)


def rutz_cutoff(mesh, rho, dv, radd=4.0, acut=0.4):
    out = np.zeros((2, mesh.N))
    lib().or_rutz_cutoff(P(np.ascontiguousarray(rho, dtype=np.float64).reshape(2, mesh.N)),
                         mesh.nx, mesh.ny, mesh.nz, ctypes.c_double(mesh.dx),
                         ctypes.c_double(dv), ctypes.c_double(radd), ctypes.c_double(acut),
                         P(out))
    return out


def ylm_field(mesh, l, m):
    """SpherHarm of the moment (l, m) on the octant mesh (r^l Re Y_lm;
    1/sqrt(4 pi) for l = 0), points at (i + 1/2) dx."""
    out = np.zeros(mesh.N)
    q = 0
    for k in range(mesh.nz):
        for j in range(mesh.ny):
            for i in range(mesh.nx):
                X, Y, Z = (i + 0.5) * mesh.dx, (j + 0.5) * mesh.dx, (k + 0.5) * mesh.dx
                out[q] = 1.0 / np.sqrt(4 * np.pi) if l == 0 else rlylm(l, m, X, Y, Z)
                q += 1
    return out


def parse_namelist_items(text):
    """'l=2, m=0, Constraint=30.0' -> {'l': 2, 'm': 0, 'constraint': 30.0}."""
    out = {}
    for item in text.split(","):
        k, v = item.split("=")
        k, v = k.strip().lower(), v.strip()
        out[k] = int(v) if v.lstrip("-").isdigit() else float(v)
    return out


D0 = float(np.float32(0.1))     # d0 of Moments.f90:258, a default-real literal


class Constraints:
    """Multipole constraints on the density moments (Moments.f90):
    Calculate_Multipole values and <Q^2> (:1907-2004), the first-call targets
    of CalculateAllMoments (:2250-2334), CalcConstraintEnergy (:2368-2531),
    Readjust (:2929-3058), TurnOffConstraints (:3516-3549) and the multipole
    field of AlternateStep (ImaginaryTime.f90:306-437).  Types 1 and 2,
    Isoswitch 1 and 2; Rutz cutoff, damping 0, ReadjustSlowDown 0.01."""

    def __init__(self, specs, mesh, lm, dv):
        self.mesh, self.dv = mesh, dv
        self.damping, self.slowdown = 0.0, 0.01
        self.items = []
        for text in specs:
            p = parse_namelist_items(text)
            l, m = p["l"], p.get("m", 0)
            idx = [k for k, (a, b) in enumerate(lm) if a == l and b == m][0]
            iso = p.get("isoswitch", 1)
            if iso == 1:
                c = np.array([p.get("constraint", 0.0)] * 2)
            else:
                c = np.array([p.get("constraintneutrons", 0.0), p.get("constraintprotons", 0.0)])
            self.items.append(dict(idx=idx, type=p.get("constrainttype", 2), iso=iso,
                                   iteration=p.get("iteration", -1), con=c.copy(),
                                   true=c.copy(), inten=np.array([p.get("intensity", 0.0)] * 2),
                                   Y=ylm_field(mesh, l, m)))
        self.items.sort(key=lambda c: c["idx"])     # the moment list's order
        self.ce = np.zeros((2, mesh.N))
        self.first = True

    def calculate(self, rho, val):
        rho = rho.reshape(2, self.mesh.N)
        self.cut = rutz_cutoff(self.mesh, rho, self.dv)
        for c in self.items:
            c["value"] = np.array(val[c["idx"]])
            c["sq"] = np.array([np.sum(c["Y"] ** 2 * rho[it]) * self.dv for it in range(2)])
            if self.first and c["type"] != 0:
                c["con"] = (np.array([c["value"].sum()] * 2) if c["iso"] == 1
                            else c["value"].copy())
        self.first = False
        self.ce = self.damping * self.ce
        for c in self.items:
            if c["type"] == 0:
                continue
            if c["iso"] == 1:
                fac = np.array([2.0 * c["inten"][0]] * 2)
                value = np.array([c["value"].sum()] * 2)
                des = np.array([c["con"][0]] * 2)
            else:
                fac, value, des = 2.0 * c["inten"], c["value"], c["con"]
            if np.any(self.ce != 0.0):
                fac = fac * (1.0 - self.damping)
            for it in range(2):
                self.ce[it] = self.ce[it] + fac[it] * (value[it] - des[it]) * c["Y"] * \
                    self.cut[it]

    def readjust(self, momentum):
        for c in self.items:
            if c["type"] == 1:
                if c["iso"] == 1:
                    if c["inten"][0] == 0.0:
                        o2 = np.array([c["sq"].sum()] * 2)
                        c["inten"][:] = 2 / o2.sum()
                        self.slowdown = 2 * c["inten"][0]
                    c["con"][:] = c["con"][0] - self.slowdown * (c["value"].sum() - c["true"][0])
                else:
                    c["con"] = c["con"] - self.slowdown * (c["value"] - c["true"])
            elif c["type"] == 2:
                if c["inten"][0] == 0.0:
                    if c["iso"] == 1:
                        c["inten"][:] = 2 / c["sq"].sum()
                        if momentum != 0.0:
                            c["inten"][:] = 0.5 * c["inten"][0]
                    else:
                        c["inten"] = 2 / c["sq"]
                        if momentum != 0.0:
                            c["inten"] = 0.5 * c["inten"]
                if c["iso"] == 1:
                    c["con"][:] = c["con"][0] - (c["value"].sum() - c["true"][0])
                else:
                    c["con"] = c["con"] - (c["value"] - c["true"])

    def turn_off(self, iteration):
        for c in self.items:
            if c["type"] != 0 and c["iteration"] > 0 and iteration > c["iteration"]:
                c["type"] = 0

    def alternating(self):
        return any(c["type"] == 2 for c in self.items)

    def alternate_field(self):
        mult = np.zeros((2, self.mesh.N))
        for c in self.items:
            if c["type"] != 2:
                continue
            o2 = c["sq"].sum()
            for it in range(2):
                v = c["value"].sum() if c["iso"] == 1 else c["value"][it]
                d = c["true"][0] if c["iso"] == 1 else c["true"][it]
                mult[it] = mult[it] + 0.5 * (v - d) / (o2 + D0) * self.cut[it] * c["Y"]
        return mult


def all_densities(mesh, psi, qn, occ, trc):
    """ComputeDensity + the derived fields, and each state's derivatives."""
    M, N = psi.shape[0], mesh.N
    der = np.zeros((M, 3, 4, N))
    lap = np.zeros((M, 4, N))
    for j, q in enumerate(qn):
        der[j], lap[j] = mesh.derive_spinor(psi[j], q[1], q[2], q[3])
    rho, tau, nj, vj, vs = density(psi, der, qn, occ, trc, N)
    derrho, laprho, rotj, rots, _ = density_derived(mesh, rho, vj, vs, trc)
    d = dict(rho=rho, tau=tau, nablaJ=nj, derrho=derrho, laprho=laprho)
    if not trc:
        d.update(vecj=vj, vecs=vs, rotvecj=rotj, rots=rots)
    return d, der, lap


def scf_run(g, niter):
    """The Evolve loop of the reference (Main.f90:172-225, 255-328) for HF with
    fixed dt/momentum and linear density mixing, restated on the oracle.
    Returns the per-iteration [E, sum kin, E_C, E_Cx, sum com, terms(36)]
    rows (start-up row first), the final states and potentials."""
    m = g["meta"]
    mesh = Mesh(m["nx"], m["ny"], m["nz"], m["dx"], m["lagrange"])
    coef, consts = g["coef"], g["consts"]
    dv, trc, qn = consts[6], m["trc"], g["qn"]
    M, N = m["nwt"], mesh.N
    kin2 = kin_coefs(g)
    psi = np.ascontiguousarray(g["psi"].reshape(M, 4, N)).copy()
    coul = Coulomb(mesh, consts[4], consts[7] + consts[8], consts[8])
    mixing = 0.75

    def rows(E):
        return np.concatenate([[E[0], E[1] + E[2], E[3], E[4], E[5] + E[6]], E[7:]])

    cons = None

    def construct(d, calc=1):
        lm, val = moments(d["rho"], mesh, dv)
        if cons is not None:
            for _ in range(calc):
                cons.calculate(d["rho"], val)
        vc, vx = coul.solve(d["rho"][1], lm, val[:, 1])
        F = potentials(coef, mesh, trc, d, vc, vx)
        if cons is not None:
            F["U"] = (F["U"].reshape(2, N) + cons.ce).reshape(F["U"].shape)
        return F, vc, vx

    if m.get("constraints"):
        lm0, _ = moments(np.zeros((2, N)), mesh, dv)
        cons = Constraints(m["constraints"], mesh, [tuple(x) for x in lm0], dv)

    scheme = m.get("mixing_scheme", 0)
    preconw = m.get("preconw", 0.0)
    occ = hf_fill(np.zeros(M), qn, m["neutrons"], m["protons"], trc)
    d, der, lap = all_densities(mesh, psi, qn, occ, trc)
    F, vc, vx = construct(d, calc=3)      # CalculateAllMoments three times (Main.f90:213-215)
    E = energy(coef, consts, m["com1body"], trc, mesh, d, vc,
               kinetic_sums(psi, lap, qn, occ, dv))
    traj = [rows(E)]
    altcheck = cons is not None and cons.alternating()
    upd = np.zeros_like(psi)
    hbar = consts[5]
    dt, mom = m["dt"], m["momentum"]
    dt_over_hbar = dt / hbar
    est = m.get("parameter_estimation", 0) == 1
    if est:
        est_psi = np.ascontiguousarray(g["est_seed"], dtype=np.float64).reshape(-1).copy()
        qn_est = qn[M - 1]
    e_sp = np.zeros(M)
    params = []
    for k in range(niter):
        if altcheck:
            mult = cons.alternate_field()
            for i in range(M):
                it = (qn[i][0] + 3) // 2 - 1
                psi[i] = psi[i] - mult[it] * psi[i]
            # AlternateStep (ImaginaryTime.f90:306-437) changes the values
            # only: neither SetGrid nor Gramschmidt recomputes the
            # derivatives, so the next GradDesc's hPsi combines the new
            # values with the derivatives of the states before the step
            psi = gram_schmidt(psi, qn, e_sp, dv)
        Ff = {key: np.ascontiguousarray(v).reshape(-1) for key, v in F.items()}
        if est:
            it_est = (qn_est[0] + 3) // 2 - 1
            r = estimate(mesh, est_psi, k == 0, qn_est, trc, kin2[it_est], Ff, dt, mom, dv,
                         hbar, qn, occ, e_sp)
            dt, mom = r[0], r[1]
            dt_over_hbar = dt / hbar
            params.append((dt, mom))
        psi, upd, e_sp, _ = graddesc(psi, der, lap, upd, qn, trc, kin2, Ff, dt_over_hbar,
                                      mom, dv)
        occ = hf_fill(e_sp, qn, m["neutrons"], m["protons"], trc)
        if cons is not None:
            altcheck = cons.alternating()
        old = d
        d, der, lap = all_densities(mesh, psi, qn, occ, trc)
        if scheme == 0:
            for key in d:
                lib().or_mix(P(d[key]), P(np.ascontiguousarray(old[key])),
                             ctypes.c_size_t(d[key].size), ctypes.c_double(mixing))
        Fold = F
        if cons is not None:
            lm_, val_ = moments(d["rho"], mesh, dv)
            cons.calculate(d["rho"], val_)
            cons.readjust(mom)
            cons.turn_off(k + 1)
            F, vc, vx = construct(d, calc=0)
        else:
            F, vc, vx = construct(d)
        if scheme == 2:
            precondition_potentials(mesh, F, Fold, trc, dv, preconw)
        E = energy(coef, consts, m["com1body"], trc, mesh, d, vc,
                   kinetic_sums(psi, lap, qn, occ, dv))
        traj.append(rows(E))
    return dict(traj=np.array(traj), psi=psi, energies=e_sp, occ=occ, fields=F, dens=d,
                vcoul=vc, vcoulx=vx, cpot=coul.cpot, params=np.array(params),
                constraints=[] if cons is None else
                [dict(type=c["type"], con=c["con"].copy(), inten=c["inten"].copy(),
                      sq=c["sq"].copy()) for c in cons.items])
