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

    def derive_states(self, psi, qn):
        """CompDer of every state: der[M][3][4][N], lap[M][4][N]."""
        M = psi.shape[0]
        psi = np.ascontiguousarray(psi.reshape(M, 4, self.N), dtype=np.float64)
        der = np.zeros((M, 3, 4, self.N))
        lap = np.zeros((M, 4, self.N))
        lib().or_derive_states(P(psi), P(der), P(lap), int(M),
                               P(np.ascontiguousarray(qn, dtype=np.int32)), self.nx, self.ny,
                               self.nz, ctypes.c_double(self.dx), self.lag, *self.lagmats())
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


def estimate(mesh, psi, start, qn_est, trc, kin, F, dt, mom, dv, hbar, qn, occ, energy,
             rel_hfb=0.0):
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
                      P(np.ascontiguousarray(energy, dtype=np.float64)),
                      ctypes.c_double(rel_hfb), P(out))
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
                                   Y=ylm_field(mesh, l, m), hist=np.zeros(8)))
        self.items.sort(key=lambda c: c["idx"])     # the moment list's order
        self.ce = np.zeros((2, mesh.N))
        self.first = True

    def calculate(self, rho, val):
        rho = rho.reshape(2, self.mesh.N)
        self.cut = rutz_cutoff(self.mesh, rho, self.dv)
        for c in self.items:
            c["value"] = np.array(val[c["idx"]])
            # Value and OldValue(:,1:7) summed over isospins (Moments.f90:1920-1932)
            c["hist"] = np.concatenate([[c["value"].sum()], c["hist"][:7]])
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


def angular_momentum(mesh, psi, qn, occ, dv, trc, dens=None):
    """UpdateAM (Cranking.f90:135-232) with DiagSpin and DiagAng
    (Spwf.f90:977-1489) of every state under the EV8 symmetries (signature
    and time simplex conserved: J_x, J_y, Im (J T)_x, Re (J T)_y, the z parts
    of J T and the matching spin parts stay zero).  Returns per state
    [M][7][3] = J, JTR, JTI, J2, Sp, SpTR, SpTI, and with time reversal
    broken the totals [5][3] = TotalAngMom, JTR, JTI, J2Total,
    TotalAngMom_dens and AMIsoblock [3][2][2][2] (direction, isospin, parity,
    signature) over the states with their occupations."""
    M = psi.shape[0]
    der, _ = mesh.derive_states(psi, qn)
    nx, ny, nz, dx = mesh.nx, mesh.ny, mesh.nz, mesh.dx
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    X, Y, Z = ((a.reshape(-1) + 0.5) * dx for a in (x, y, z))
    out = np.zeros((M, 7, 3))
    for w in range(M):
        p1, p2, p3, p4 = psi[w].reshape(4, -1)
        (x1, x2, x3, x4), (y1, y2, y3, y4), (z1, z2, z3, z4) = der[w]
        sz = 0.5 * (p1 * p1 + p2 * p2 - p3 * p3 - p4 * p4)
        stx = 0.5 * (-p1 * p1 + p2 * p2 + p3 * p3 - p4 * p4)
        sty = 0.5 * (-p1 * p1 + p2 * p2 - p3 * p3 + p4 * p4)
        J3 = sz + p2 * Y * x1 - p2 * X * y1 - p1 * Y * x2 + p1 * X * y2 \
            + p4 * Y * x3 - p4 * X * y3 - p3 * Y * x4 + p3 * X * y4
        JTR1 = stx + p2 * Z * y3 - p2 * Y * z3 + p1 * Z * y4 - p1 * Y * z4 \
            - p4 * Z * y1 + p4 * Y * z1 - p3 * Z * y2 + p3 * Y * z2
        JTI2 = sty + p1 * Z * x3 - p1 * X * z3 - p2 * Z * x4 + p2 * X * z4 \
            - p3 * Z * x1 + p3 * X * z1 + p4 * Z * x2 - p4 * X * z2
        r = (-Z * y2 + Y * z2 + 0.5 * p3, Z * y1 - Y * z1 + 0.5 * p4,
             -Z * y4 + Y * z4 + 0.5 * p1, Z * y3 - Y * z3 + 0.5 * p2)
        J2x = sum(a * a for a in r)
        r = (-X * z2 + Z * x2 + 0.5 * p4, X * z1 - Z * x1 - 0.5 * p3,
             -X * z4 + Z * x4 - 0.5 * p2, X * z3 - Z * x3 + 0.5 * p1)
        J2y = sum(a * a for a in r)
        r = (-Y * x2 + X * y2 + 0.5 * p1, Y * x1 - X * y1 + 0.5 * p2,
             -Y * x4 + X * y4 - 0.5 * p3, Y * x3 - X * y3 - 0.5 * p4)
        J2z = sum(a * a for a in r)
        out[w, 0, 2] = J3.sum() * dv
        out[w, 1, 0] = JTR1.sum() * dv
        out[w, 2, 1] = JTI2.sum() * dv
        out[w, 3] = [J2x.sum() * dv, J2y.sum() * dv, J2z.sum() * dv]
        out[w, 4, 2] = sz.sum() * dv
        out[w, 5, 0] = stx.sum() * dv
        out[w, 6, 1] = sty.sum() * dv
    tot = np.zeros((5, 3))
    iso = np.zeros((3, 2, 2, 2))
    if not trc:
        for w in range(M):
            tot[0:4] += occ[w] * out[w, 0:4]
            it, P_, S_ = ((qn[w][k] + 3) // 2 - 1 for k in (0, 1, 2))
            iso[:, it, P_, S_] += occ[w] * out[w, 0]
        if dens is not None:
            jx, jy, szd = dens["vecj"][:, 0], dens["vecj"][:, 1], dens["vecs"][:, 2]
            jx, jy, szd = (a.reshape(2, -1) for a in (jx, jy, szd))
            tot[4, 2] = sum(0.5 * szd[it].sum() + (-Y * jx[it] + X * jy[it]).sum()
                            for it in range(2)) * dv
    return out, tot, iso


def all_densities(mesh, psi, qn, occ, trc):
    """ComputeDensity + the derived fields, and each state's derivatives."""
    M, N = psi.shape[0], mesh.N
    der, lap = mesh.derive_states(psi, qn)
    rho, tau, nj, vj, vs = density(psi, der, qn, occ, trc, N)
    derrho, laprho, rotj, rots, _ = density_derived(mesh, rho, vj, vs, trc)
    d = dict(rho=rho, tau=tau, nablaJ=nj, derrho=derrho, laprho=laprho)
    if not trc:
        d.update(vecj=vj, vecs=vs, rotvecj=rotj, rots=rots)
    return d, der, lap


# The precisions at which the reference driver records ConvergenceCheck's
# tests (oracle/ref_driver.f90 dump_conv): energy, moments, Fermi.
CONV_EPREC = (1e-9, 1e-6, 1e-4, 1e-3, 1e-2)
CONV_MPREC = (1e-4, 1e-2, 1e-1, 5e-1)
CONV_PPREC = (1e-4, 1.0, 10.0, 100.0)


def convergence_tests(ehist, fermi, pairing, cons, mprec, eprec, pprec):
    """The tests of ConvergenceCheck (Main.f90:368-397): ConverMultipoleAll /
    ConverMultipole (Moments.f90:2628-2701: the last 8 values of each
    constrained moment, summed over isospins, against sum(Constraint),
    relative unless sum(TrueConstraint) is 0, failing at >= prec),
    ConverEnergy (Energy.f90:1458-1481: the total energy against the 7 before
    it, relative, passing at <= prec) and ConverFermi (Pairing.f90:1048-1070:
    HF passes; with pairing each Fermi energy against FermiHistory, which the
    reference never updates from its initial zeros).  ehist = [E, OldEnergy(1:7)]."""
    mom = True
    for c in (cons.items if cons is not None else []):
        if c["type"] == 0:
            continue
        con = c["con"][0] if c["iso"] == 1 else c["con"].sum()
        tru = c["true"][0] if c["iso"] == 1 else c["true"].sum()
        d = np.abs((con - c["hist"]) / con) if tru != 0.0 else np.abs(con - c["hist"])
        if np.any(d >= mprec):
            mom = False
    en = all(abs((ehist[i] - ehist[0]) / ehist[0]) <= eprec for i in range(1, 8))
    fe = (not pairing) or all(abs(f) <= pprec for f in fermi)
    return mom, en, fe


def conv_flags(ehist, fermi, pairing, cons):
    """convergence_tests at the driver's precisions: 13 flags."""
    out = [convergence_tests(ehist, fermi, pairing, cons, 1.0, e, 1.0)[1] for e in CONV_EPREC]
    out += [convergence_tests(ehist, fermi, pairing, cons, mp, 1.0, 1.0)[0] for mp in CONV_MPREC]
    out += [convergence_tests(ehist, fermi, pairing, cons, 1.0, 1.0, pp)[2] for pp in CONV_PPREC]
    return [int(x) for x in out]


def scf_run(g, niter, on_iteration=None):
    """The Evolve loop of the reference (Main.f90:172-225, 255-328) for HF with
    fixed dt/momentum and linear density mixing, restated on the oracle.
    Returns the per-iteration [E, sum kin, E_C, E_Cx, sum com, terms(36)]
    rows (start-up row first), the final states and potentials.
    on_iteration(k) is called after the start-up (k = 0) and after each
    iteration k."""
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
    hfb = HFB(qn, trc, m) if m.get("pairing_type", 0) == 2 else None
    bcs = BCS(qn, m) if m.get("pairing_type", 0) == 1 else None
    pair = hfb if hfb is not None else bcs
    e_start = np.asarray(g["start_energies"]) if "start_energies" in g else np.zeros(M)
    pairs = []

    def pairing_step(psi, e_sp, rho):
        """SolvePairing: HF filling, BCS in the HF basis, or HFB with the
        canonical basis as the density basis (Pairing.f90:812)."""
        if bcs is not None:
            return psi, bcs.solve(psi.reshape(M, 4, N), e_sp, rho, dv)
        if hfb is None:
            return psi, hf_fill(e_sp, qn, m["neutrons"], m["protons"], trc)
        can, occ_c = hfb.solve(psi.reshape(M, 4, N), e_sp, rho, dv)
        return can, occ_c

    if hfb is not None:
        hfb.guess()
    if pair is not None and "start_fermi" in g:
        pair.fermi = np.array(g["start_fermi"], dtype=float)
    dbasis, occ = pairing_step(psi, e_start, np.zeros((2, N)))
    d, der, lap = all_densities(mesh, dbasis, qn, occ, trc)
    F, vc, vx = construct(d, calc=3)      # CalculateAllMoments three times (Main.f90:213-215)
    E = energy(coef, consts, m["com1body"], trc, mesh, d, vc,
               kinetic_sums(dbasis, lap, qn, occ, dv))
    if pair is not None:
        ep = pair.energy()
        E[0] += ep.sum()
        pairs.append(np.concatenate([pair.fermi, ep]))
    if hfb is not None:
        der, lap = mesh.derive_states(psi, qn)
    traj = [rows(E)]
    ehist = np.zeros(8)
    conv = []

    def check(E):
        nonlocal ehist
        ehist = np.concatenate([[E[0]], ehist[:7]])    # CompEnergy's OldEnergy shift
        fermi = pair.fermi if pair is not None else np.zeros(2)
        conv.append(conv_flags(ehist, fermi, pair is not None, cons))

    check(E)
    if on_iteration is not None:
        on_iteration(0)
    altcheck = cons is not None and cons.alternating()
    upd = np.zeros_like(psi)
    hbar = consts[5]
    dt, mom = m["dt"], m["momentum"]
    dt_over_hbar = dt / hbar
    est = m.get("parameter_estimation", 0) == 1
    if est:
        est_psi = np.ascontiguousarray(g["est_seed"], dtype=np.float64).reshape(-1).copy()
        qn_est = qn[M - 1]
    e_sp = e_start.copy()
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
            rel = 0.0
            if hfb is not None:
                pos = np.concatenate([q[q > 0.0] for q in hfb.qpe.values()])
                rel = min(10000.0, pos.min()) if pos.size else 10000.0
            elif bcs is not None:
                rel = min(10000.0, bcs.eqp.min())     # ImaginaryTime.f90 case(1)
            r = estimate(mesh, est_psi, k == 0, qn_est, trc, kin2[it_est], Ff, dt, mom, dv,
                         hbar, qn, occ, e_sp, rel)
            dt, mom = r[0], r[1]
            dt_over_hbar = dt / hbar
            params.append((dt, mom))
        psi, upd, e_sp, _ = graddesc(psi, der, lap, upd, qn, trc, kin2, Ff, dt_over_hbar,
                                      mom, dv)
        dbasis, occ = pairing_step(psi, e_sp, d["rho"])
        if cons is not None:
            altcheck = cons.alternating()
        old = d
        d, der, lap = all_densities(mesh, dbasis, qn, occ, trc)
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
                   kinetic_sums(dbasis, lap, qn, occ, dv))
        if pair is not None:
            ep = pair.energy()
            E[0] += ep.sum()
            pairs.append(np.concatenate([pair.fermi, ep]))
        if hfb is not None:
            der, lap = mesh.derive_states(psi, qn)
        traj.append(rows(E))
        check(E)
        if on_iteration is not None:
            on_iteration(k + 1)
    return dict(traj=np.array(traj), psi=psi, energies=e_sp, occ=occ, fields=F, dens=d,
                conv=np.array(conv, np.int32), dbasis=dbasis, mesh=mesh,
                vcoul=vc, vcoulx=vx, cpot=coul.cpot, params=np.array(params),
                pairing=np.array(pairs), hfb=hfb, bcs=bcs,
                constraints=[] if cons is None else
                [dict(type=c["type"], con=c["con"].copy(), inten=c["inten"].copy(),
                      sq=c["sq"].copy()) for c in cons.items])


def final_iteration(g, r):
    """FinalIteration (Main.f90:482-529) after scf_run's loop: the density
    basis differentiated with the Lagrange scheme, UpdateDensities(0) and
    CompEnergy with the last Coulomb potential; the energy row in the layout
    of scf_run's rows, and the unmixed densities of the run's own scheme."""
    m = g["meta"]
    coef, consts = g["coef"], g["consts"]
    dv, trc, qn = consts[6], m["trc"], g["qn"]
    mesh = r["mesh"]
    lag = Mesh(mesh.nx, mesh.ny, mesh.nz, m["dx"], True)
    d, der, lap = all_densities(lag, r["dbasis"], qn, r["occ"], trc)
    E = energy(coef, consts, m["com1body"], trc, lag, d, r["vcoul"],
               kinetic_sums(r["dbasis"], lap, qn, r["occ"], dv))
    if r["hfb"] is not None:
        E[0] += r["hfb"].energy().sum()
    row = np.concatenate([[E[0], E[1] + E[2], E[3], E[4], E[5] + E[6]], E[7:]])
    d_own, _, _ = all_densities(mesh, r["dbasis"], qn, r["occ"], trc)
    return row, d_own


# ---- HFB pairing (HFB.f90, Pairing.f90, PairingInteraction.f90) -----------

def pair_density(w1, w2, tr):
    """GetPairDensity (PairingInteraction.f90:101-180) of two spinors [4][N]:
    complex pair amplitude; tr selects the time-reversed form."""
    if not tr:
        re = w1[0] * w2[2] - w1[1] * w2[3] - w1[2] * w2[0] + w1[3] * w2[1]
        im = w1[0] * w2[3] + w1[1] * w2[2] - w1[2] * w2[1] - w1[3] * w2[0]
    else:
        re = -w1[0] * w2[0] - w1[1] * w2[1] - w1[2] * w2[2] - w1[3] * w2[3]
        im = w1[0] * w2[1] - w1[1] * w2[0] + w1[2] * w2[3] - w1[3] * w2[2]
    return re + 1j * im


def time_reverse(w):
    """TimeReverse (Spinor.f90:371-389): -i sigma_y K psi."""
    return np.stack([w[2], -w[3], -w[0], w[1]])


HFB_NUM_CUT = 1e-12


class HFB:
    """The HFB pairing step of SolvePairing (Pairing.f90:743-890) under time
    reversal with conserved signature: blocks (parity, isospin) of the states
    and their time-reversed partners (PrepareHFBModule, HFB.f90:248-295),
    GuessHFBMatrices (:3645-3836), the pairing field and gaps (:350-632),
    the Broyden Fermi solver (:1849-2161) over HFBNumberofParticles
    (:810-1073, ConstructHFBHamiltonian :2554-2631,
    DiagonaliseHFBHamiltonian_Signature :2695-2801, constructRhoHFB,
    ConstructKappaHFB :3128-3360), HFBOccupations with the canonical basis
    (:3361-3644) and HFBEnergy (:296-348).  Dense diagonalisations use
    numpy's eigh (eigenvalues ascending, as diagoncr8 sorts them)."""

    def __init__(self, qn, trc, m):
        assert trc, "HFB restatement covers time-reversal-conserving runs"
        self.qn = np.asarray(qn)
        self.nwt = len(qn)
        self.size = 2 * self.nwt
        self.blocks = {}
        for i in range(self.size):
            ii = i % self.nwt
            P = (self.qn[ii][1] + 3) // 2
            it = (self.qn[ii][0] + 3) // 2
            self.blocks.setdefault((P, it), []).append(i)
        self.keys = sorted(self.blocks, key=lambda k: (k[1], k[0]))   # it outer, P inner
        self.G = np.array(m["pairing_strength"])
        self.alpha = np.array(m["pairing_alpha"])
        self.cut = np.array(m["pairing_cut"])
        self.mu = np.array(m["pairing_mu"])
        self.rhosat = m["rhosat"]
        self.mix = m["hfb_mix"]
        self.iters = m["hfb_iter"]
        self.particles = np.array([m["neutrons"], m["protons"]], dtype=float)
        z = {k: np.zeros((len(v), len(v))) for k, v in self.blocks.items()}
        self.rho = {k: v.copy() for k, v in z.items()}
        self.kappa = {k: v.copy() for k, v in z.items()}
        self.old_rho = {k: v.copy() for k, v in z.items()}
        self.old_kappa = {k: v.copy() for k, v in z.items()}
        self.delta = {k: v.copy() for k, v in z.items()}
        self.fermi = np.array([-10.0, -10.0])

    def sig(self, i):
        s = self.qn[i % self.nwt][2]
        return -s if i >= self.nwt else s

    def guess(self):
        for k, idx in self.blocks.items():
            N = len(idx)
            ovl = np.zeros((N, N))
            for i in range(N // 2):
                ovl[i, N // 2 + i] = ovl[N // 2 + i, i] = 1.0
            kap = np.zeros((N, N))
            found = np.zeros(N, bool)
            for i in range(N):
                best, maxo = 0, -9999.999
                for j in range(N):
                    if found[j]:
                        continue
                    if abs(ovl[i, j]) > maxo:
                        best, maxo = j, abs(ovl[i, j])
                found[best] = True
                kap[i, best] = 0.25
                kap[best, i] = -0.25
            self.kappa[k] = kap

    def cutoffs(self, energy, fermi):
        pc = np.zeros(self.nwt)
        for i in range(self.nwt):
            it = (self.qn[i][0] + 3) // 2 - 1
            up = (energy[i] - fermi[it] - self.cut[it]) / self.mu[it]
            down = -(energy[i] - fermi[it] + self.cut[it]) / self.mu[it]
            c = np.sqrt(np.sqrt(1.0 / (1.0 + np.exp(up))))
            pc[i] = c * np.sqrt(np.sqrt(1.0 / (1.0 + np.exp(down))))
        return pc

    def density_factor(self, rho):
        rho0 = rho[0] + rho[1]
        return np.stack([-self.G[it] * (1.0 - self.alpha[it] / self.rhosat * rho0)
                         for it in range(2)])

    def _block(self, k):
        """Positions of block k: original state of each (idx mod nwt), the
        reversed flag, and the signature of each position."""
        idx = np.array(self.blocks[k])
        orig = idx % self.nwt
        rev = idx >= self.nwt
        sig = np.where(rev, -1, 1) * self.qn[orig, 2]
        return idx, orig, rev, sig

    @staticmethod
    def _partner(y, F, tr):
        """Y with sum_g x_g Y_g = Re P(x, y) Re F + Im P(x, y) Im F for the pair
        amplitude P of pair_density (y: [..., 4, N], F complex [N])."""
        fr, fi = F.real, F.imag
        if tr:
            return np.stack([-fr * y[..., 0, :] + fi * y[..., 1, :],
                             -fr * y[..., 1, :] - fi * y[..., 0, :],
                             -fr * y[..., 2, :] + fi * y[..., 3, :],
                             -fr * y[..., 3, :] - fi * y[..., 2, :]], axis=-2)
        return np.stack([fr * y[..., 2, :] + fi * y[..., 3, :],
                         -fr * y[..., 3, :] + fi * y[..., 2, :],
                         -fr * y[..., 0, :] - fi * y[..., 1, :],
                         fr * y[..., 1, :] - fi * y[..., 0, :]], axis=-2)

    def pairing_field(self, psi, pc, df):
        """HFBPairingField: field_it = -sum over pairs (i > j, opposite
        signatures, pc_i pc_j |kappa_ij| above the cut) of pc_i pc_j kappa_ij
        P(psi_j, psi_i); P is bilinear, so the sum runs as
        sum_o P(psi_o, phi_o) with phi = W psi over the block's states."""
        N = psi.shape[-1]
        field = np.zeros((2, N), complex)
        for k in self.keys:
            it = k[1] - 1
            idx, orig, rev, sig = self._block(k)
            n = len(idx)
            pcs = pc[orig]
            kap = self.kappa[k]
            lower = np.tril(np.ones((n, n), bool), -1)             # i > j
            mask = lower & (pcs[None, :] >= HFB_NUM_CUT) & (sig[:, None] == -sig[None, :])
            mask &= pcs[None, :] * pcs[:, None] * np.abs(kap) >= HFB_NUM_CUT
            uo, inv = np.unique(orig, return_inverse=True)
            base = psi[uo]
            for tr in (False, True):
                sel = mask & ((rev[:, None] ^ rev[None, :]) == tr)
                if not sel.any():
                    continue
                W = np.where(sel, pcs[:, None] * pcs[None, :] * kap, 0.0)   # W[i, j]
                # Wm[o_j, o_i] = sum of W[i, j] over positions of those states
                Wm = np.zeros((len(uo), len(uo)))
                np.add.at(Wm, (inv[None, :].repeat(n, 0), inv[:, None].repeat(n, 1)), W)
                phi = np.tensordot(Wm, base, axes=(1, 0))                # [o][4][N]
                x, y = base, phi
                if tr:
                    re = -(x[:, 0] * y[:, 0] + x[:, 1] * y[:, 1] + x[:, 2] * y[:, 2] +
                           x[:, 3] * y[:, 3])
                    im = x[:, 0] * y[:, 1] - x[:, 1] * y[:, 0] + x[:, 2] * y[:, 3] - \
                        x[:, 3] * y[:, 2]
                else:
                    re = x[:, 0] * y[:, 2] - x[:, 1] * y[:, 3] - x[:, 2] * y[:, 0] + \
                        x[:, 3] * y[:, 1]
                    im = x[:, 0] * y[:, 3] + x[:, 1] * y[:, 2] - x[:, 2] * y[:, 1] - \
                        x[:, 3] * y[:, 0]
                field[it] = field[it] - (re.sum(0) + 1j * im.sum(0))
        return field * df

    def gaps(self, psi, field, pc, dv):
        """HFBGaps: Delta_ij = dv pc_i pc_j sum [Re P(psi_i, psi_j) Re F + Im P
        Im F] for i < j of opposite signatures above the cut, Delta_ji =
        -Delta_ij; the sums over points as products psi . Y(psi)^T."""
        for k in self.keys:
            it = k[1] - 1
            idx, orig, rev, sig = self._block(k)
            n = len(idx)
            pcs = pc[orig]
            uo, inv = np.unique(orig, return_inverse=True)
            base = psi[uo].reshape(len(uo), -1)
            upper = np.triu(np.ones((n, n), bool), 1)
            mask = upper & (sig[:, None] == -sig[None, :])
            mask &= pcs[:, None] * pcs[None, :] >= HFB_NUM_CUT
            d = np.zeros((n, n))
            for tr in (False, True):
                sel = mask & ((rev[:, None] ^ rev[None, :]) == tr)
                if not sel.any():
                    continue
                Y = self._partner(psi[uo], field[it], tr).reshape(len(uo), -1)
                Q = base @ Y.T                                           # Q[o, o']
                d = np.where(sel, dv * pcs[:, None] * pcs[None, :] * Q[inv][:, inv], d)
            self.delta[k] = d - d.T

    def number(self, lam, energy):
        """HFBNumberofParticles: rho, kappa, quasiparticle energies at lam."""
        n = np.zeros(2)
        self.qpe = {}
        for k in self.keys:
            it = k[1] - 1
            idx = np.array(self.blocks[k])
            N = len(idx)
            h = N // 2
            e = energy[idx % self.nwt] - lam[it]
            # the signature-reduced matrix of ConstructHFBHamiltonian:
            # [[e_top - lam, Delta_top,bottom], [Delta^T, -(e_bottom - lam)]]
            T = np.zeros((N, N))
            ar = np.arange(h)
            T[ar, ar] = e[:h]
            T[h + ar, h + ar] = -e[h:]
            T[:h, h:] = self.delta[k][:h, h:]
            T[h:, :h] = self.delta[k][:h, h:].T
            w, vec = np.linalg.eigh(T)
            U = np.zeros((N, 2 * N))
            V = np.zeros((N, 2 * N))
            U[:h, :N] = vec[:h, :N]
            V[h:, :N] = vec[h:, :N]
            U[h:, N:] = -vec[:h, :N]
            V[:h, N:] = vec[h:, :N]
            qpe = np.concatenate([w, w])
            cols = [2 * N - i + 1 for i in range(1, h + 1)] + list(range(h + 1, N + 1))
            cols = np.array(cols) - 1
            rho = V[:, cols] @ V[:, cols].T
            K = np.tril(V[:, cols] @ U[:, cols].T, -1)
            self.rho[k], self.kappa[k] = rho, K - K.T
            self.qpe[k] = qpe[cols]
            n[it] += np.trace(rho)
        return n

    def broyden(self, energy, prec=1e-12):
        fermi = self.fermi.copy()
        hist = fermi + 0.01
        N = self.number(fermi, energy) - self.particles
        conv = np.abs(N) < prec
        NX = self.number(hist, energy) - self.particles
        jac = (N - NX) / (fermi - hist)
        upd = np.zeros(2)
        norm = np.zeros(2)
        for it_ in range(1, self.iters + 1):
            det = jac.copy()
            with np.errstate(divide="ignore"):
                inv = np.where(np.abs(det) < 0.001, np.copysign(0.99, det), 1.0 / det)
            for it in range(2):
                if conv[it]:
                    continue
                upd[it] = -inv[it] * N[it]
                norm[it] = upd[it] ** 2
            upd = 0.5 * upd
            for it in range(2):
                if conv[it]:
                    continue
                hist[it] = fermi[it]
                fermi[it] = fermi[it] + upd[it]
            N = self.number(fermi, energy) - self.particles
            conv = conv | (np.abs(N) < prec)
            if conv.all():
                break
            if it_ > 1:
                for it in range(2):
                    if conv[it]:
                        continue
                    jac[it] = jac[it] + N[it] * upd[it] / norm[it]
        self.fermi = fermi
        self.iterations = it_

    def occupations(self, psi):
        """HFBOccupations: canonical basis and occupations (TRC: twice the
        eigenvalues of rho), placed in the slots of the original states of
        each block; then the mixing of rho and kappa."""
        can = np.zeros_like(psi)
        occ = np.zeros(self.nwt)
        for k in self.keys:
            idx = self.blocks[k]
            N = len(idx)
            h = N // 2
            T = self.rho[k].copy()
            for i in range(N):
                if idx[i] >= self.nwt or self.qn[idx[i] % self.nwt][2] == -1:
                    T[i, i] -= 2.0
            w, vec = np.linalg.eigh(T)
            vec = np.where(np.abs(vec) < 1e-11, 0.0, vec)
            w = np.where(w < -0.1, w + 2, w)
            w = np.where(w < 0.0, 0.0, w)
            w = np.minimum(w, 1.0)
            w = 2.0 * w
            cols = [c for c in range(N) if not np.any(np.abs(vec[h:, c]) > HFB_NUM_CUT)]
            slots = [i for i in idx if i < self.nwt]
            assert len(cols) == len(slots)
            # canonical state c = sum_j vec[j, c] (psi_j, or its time reverse
            # for a reversed position); time reversal is linear
            ia = np.array(idx)
            cv = np.where(np.abs(vec[:, cols]) < HFB_NUM_CUT, 0.0, vec[:, cols])   # [j][c]
            fw, bw = ia < self.nwt, ia >= self.nwt
            v = np.tensordot(cv[fw].T, psi[ia[fw] % self.nwt], axes=(1, 0))
            if bw.any():
                v = v + time_reverse(np.moveaxis(
                    np.tensordot(cv[bw].T, psi[ia[bw] % self.nwt], axes=(1, 0)), 1, 0)
                ).swapaxes(0, 1)
            can[slots] = v
            occ[slots] = w[cols]
        if any(np.any(self.old_rho[k] != 0.0) for k in self.keys):
            for k in self.keys:
                self.rho[k] = self.mix * self.rho[k] + (1.0 - self.mix) * self.old_rho[k]
        if any(np.any(self.kappa[k] != 0.0) for k in self.keys):
            for k in self.keys:
                self.kappa[k] = self.mix * self.kappa[k] + (1.0 - self.mix) * self.old_kappa[k]
        for k in self.keys:
            self.old_rho[k] = self.rho[k].copy()
            self.old_kappa[k] = self.kappa[k].copy()
        return can, occ

    def energy(self):
        e = np.zeros(2)
        for k in self.keys:
            e[k[1] - 1] -= np.sum(self.delta[k] * self.kappa[k])
        return -0.5 * e

    def solve(self, psi, energy, rho, dv):
        """SolvePairing for PairingType 2 with PairingIter = 1."""
        if all(not np.any(self.old_rho[k]) for k in self.keys):
            for k in self.keys:
                self.old_rho[k] = self.rho[k].copy()
                self.old_kappa[k] = self.kappa[k].copy()
        df = self.density_factor(rho)
        pc = self.cutoffs(energy, self.fermi)
        field = self.pairing_field(psi, pc, df)
        self.gaps(psi, field, pc, dv)
        self.broyden(energy)
        return self.occupations(psi)


# ---------------------------------------------------------------------------
# MOCCa v1 wavefunction files (WriteMOCCa_v1 / ReadMOCCa_v1,
# src/InOut.f90:905-1416): Fortran sequential unformatted records, each
# framed by its byte count (int32 little endian) before and after.  Parsed
# here record by record into named numpy arrays, and rebuilt from them.

def v1_records(buf):
    out, p = [], 0
    while p < len(buf):
        n = int(np.frombuffer(buf[p:p + 4], "<i4")[0])
        assert int(np.frombuffer(buf[p + 4 + n:p + 8 + n], "<i4")[0]) == n
        out.append(bytes(buf[p + 4:p + 4 + n]))
        p += 8 + n
    assert p == len(buf)
    return out


def _fields(rec, spec):
    """Split one record by spec = [(name, dtype, count)]."""
    d, p = {}, 0
    for name, dt, cnt in spec:
        size = np.dtype(dt).itemsize * cnt
        d[name] = np.frombuffer(rec[p:p + size], dt).copy()
        p += size
    assert p == len(rec), (p, len(rec))
    return d


V1_STATE = [("occupation", "<f8", 1), ("energy", "<f8", 1), ("dispersion", "<f8", 1),
            ("norm", "<f8", 1), ("J", "<f8", 3), ("qn", "<i4", 5)]   # Spwf.f90:2257-2260
V1_MOMENT = [("intensity", "<f8", 2), ("constraint", "<f8", 2), ("true_constraint", "<f8", 2),
             ("value", "<f8", 2), ("old_value", "<f8", 14)]           # Moments.f90:3341-3359


def v1_parse(buf):
    """The contents of a v1 file in the order WriteMOCCa_v1 writes them
    (InOut.f90:957-1036).  Density arrays keep Fortran order flattened."""
    r = v1_records(buf)
    f = {}
    f["version"] = int(np.frombuffer(r[0], "<i4")[0])
    e = np.frombuffer(r[1], "<f8")
    f["energy"], f["denergy"] = float(e[0]), e[1:].copy()               # :960
    f.update(_fields(r[2], [("n", "<i4", 3), ("dx", "<f8", 1)]))       # :962
    f["sym"] = np.frombuffer(r[3], "<i4").copy()                        # TRC PC SC TSC IC
    f.update(_fields(r[4], [("nwt", "<i4", 1), ("particles", "<f8", 2)]))
    nwt = int(f["nwt"][0])
    nx, ny, nz = (int(v) for v in f["n"])
    N = nx * ny * nz
    k = 5
    grid, st = [], []
    for _ in range(nwt):                                                # Spwf.f90:2255-2260
        grid.append(np.frombuffer(r[k], "<f8").reshape(4, N))
        st.append(_fields(r[k + 1], V1_STATE))
        k += 2
    f["grid"] = np.array(grid)
    for name in ("occupation", "energy", "dispersion", "norm", "J", "qn"):
        f["state_" + name] = np.array([s[name] for s in st]).reshape(nwt, -1).squeeze()
    # WriteDensity (Densities.f90:1189-1240)
    f["dens_rec"] = [np.frombuffer(x, "<f8").copy() for x in r[k:k + 9]]
    k += 9
    f["fd_order"] = np.frombuffer(r[k], "<i4").copy()
    f["force"] = r[k + 1]
    f["skyrme"] = np.frombuffer(r[k + 2], "<f8").copy()
    f.update(_fields(r[k + 3], [("force2", "S200", 1), ("hbar", "<f8", 1), ("hbm", "<f8", 2),
                                ("com", "<i4", 2)]))
    f["B"] = np.frombuffer(r[k + 4], "<f8").copy()
    f.update(_fields(r[k + 5], [("pairing_type", "<i4", 1), ("strength", "<f8", 2),
                                ("alpha", "<f8", 2), ("cut", "<f8", 2)]))
    f["cutoffs"] = np.frombuffer(r[k + 6], "<f8").copy()
    fl = np.frombuffer(r[k + 7], "<f8")
    f["fermi"], f["lnlambda"] = fl[:2].copy(), fl[2:].copy()
    k += 8
    S = 2 * nwt if f["sym"][0] else nwt
    pi = 2 if f["sym"][1] else 1
    ii = 2 if f["sym"][4] else 1
    f["hfb_size"] = S
    f["blocksizes"] = np.frombuffer(r[k], "<i4").reshape(ii, pi).copy()   # (P, it) Fortran
    uv = np.frombuffer(r[k + 1], "<c16")
    f["U"] = uv[:uv.size // 2].reshape(-1, 2 * S, S).copy()              # U(S, 2S, :, :)
    f["V"] = uv[uv.size // 2:].reshape(-1, 2 * S, S).copy()
    f["hfb_columns"] = np.frombuffer(r[k + 2], "<i4").reshape(-1, S).copy()
    rk = np.frombuffer(r[k + 3], "<c16")
    f["rho_hfb"] = rk[:rk.size // 2].reshape(-1, S, S).copy()            # (S, S, 2, 2)
    f["kappa_hfb"] = rk[rk.size // 2:].reshape(-1, S, S).copy()
    f["can_transfo"] = np.frombuffer(r[k + 4], "<c16").reshape(-1, S, S).copy()
    oc = np.frombuffer(r[k + 5], "<f8")
    f["omega"], f["crank_values"] = oc[:3].copy(), oc[3:].copy()
    k += 6
    f["moments"] = []
    while k < len(r):                                                     # WriteMoment
        h = np.frombuffer(r[k], "<i4")
        m = dict(constraint_type=int(h[0]), isoswitch=int(h[1]), l=int(h[2]), m=int(h[3]),
                 impart=int(h[4]))
        m.update(_fields(r[k + 1], V1_MOMENT))
        f["moments"].append(m)
        k += 2
    return f


def v1_build(f):
    """Bytes of a v1 file from v1_parse's dictionary."""
    recs = []
    put = recs.append
    put(np.array([f["version"]], "<i4").tobytes())
    put(np.concatenate([[f["energy"]], f["denergy"]]).astype("<f8").tobytes())
    put(np.asarray(f["n"], "<i4").tobytes() + np.asarray(f["dx"], "<f8").tobytes())
    put(np.asarray(f["sym"], "<i4").tobytes())
    put(np.asarray(f["nwt"], "<i4").tobytes() + np.asarray(f["particles"], "<f8").tobytes())
    nwt = int(np.asarray(f["nwt"]).reshape(-1)[0])
    for i in range(nwt):
        put(np.asarray(f["grid"][i], "<f8").tobytes())
        put(np.array([f["state_occupation"][i], f["state_energy"][i],
                      f["state_dispersion"][i], f["state_norm"][i]], "<f8").tobytes() +
            np.asarray(f["state_J"][i], "<f8").tobytes() +
            np.asarray(f["state_qn"][i], "<i4").tobytes())
    for d in f["dens_rec"]:
        put(np.asarray(d, "<f8").tobytes())
    put(np.asarray(f["fd_order"], "<i4").tobytes())
    put(f["force"])
    put(np.asarray(f["skyrme"], "<f8").tobytes())
    put(np.asarray(f["force2"], "S200").tobytes() + np.asarray(f["hbar"], "<f8").tobytes() +
        np.asarray(f["hbm"], "<f8").tobytes() + np.asarray(f["com"], "<i4").tobytes())
    put(np.asarray(f["B"], "<f8").tobytes())
    put(np.asarray(f["pairing_type"], "<i4").tobytes() +
        np.concatenate([f["strength"], f["alpha"], f["cut"]]).astype("<f8").tobytes())
    put(np.asarray(f["cutoffs"], "<f8").tobytes())
    put(np.concatenate([f["fermi"], f["lnlambda"]]).astype("<f8").tobytes())
    put(np.asarray(f["blocksizes"], "<i4").tobytes())
    put(np.asarray(f["U"], "<c16").tobytes() + np.asarray(f["V"], "<c16").tobytes())
    put(np.asarray(f["hfb_columns"], "<i4").tobytes())
    put(np.asarray(f["rho_hfb"], "<c16").tobytes() + np.asarray(f["kappa_hfb"], "<c16").tobytes())
    put(np.asarray(f["can_transfo"], "<c16").tobytes())
    put(np.concatenate([f["omega"], f["crank_values"]]).astype("<f8").tobytes())
    for m in f["moments"]:
        put(np.array([m["constraint_type"], m["isoswitch"], m["l"], m["m"], m["impart"]],
                     "<i4").tobytes())
        put(b"".join(np.asarray(m[k], "<f8").tobytes() for k, _, _ in V1_MOMENT))
    out = bytearray()
    for x in recs:
        n = np.array([len(x)], "<i4").tobytes()
        out += n + x + n
    return bytes(out)


def v1_blocks(qn, trc):
    """Rows of the reference's HFB blocks (PrepareHFBModule,
    HFB.f90:248-294): rows i = 0..S-1 hold state i mod nwt, the
    time-reversed copy for i >= nwt; block (P, it) lists its rows in order.
    qn: [nwt][4] (isospin, parity, signature, time simplex)."""
    nwt = len(qn)
    S = 2 * nwt if trc else nwt
    rows = {}
    for i in range(S):
        iso, P = qn[i % nwt][0], qn[i % nwt][1]
        rows.setdefault(((P + 3) // 2, (iso + 3) // 2), []).append(i)
    return S, rows


def v1_hf_kappa_guess(qn, trc):
    """KappaHFB of GuessHFBMatrices for PairingType 0 (HFB.f90:3686-3715):
    in each block, +0.1 at (i, j) and -0.1 at (j, i) for rows i < j of
    opposite signature (a time-reversed copy has the opposite one).
    Returned as complex [it][P][j][i] (the file's Fortran order)."""
    nwt = len(qn)
    S, rows = v1_blocks(qn, trc)
    K = np.zeros((2, 2, S, S), np.complex128)

    def sig(i):
        s = qn[i % nwt][2]
        return -s if i >= nwt else s

    for (P, it), r in rows.items():
        # the reference's loop order (j outer, i inner): the later write wins
        for jj in range(len(r)):
            for ii in range(len(r)):
                if sig(r[jj]) * sig(r[ii]) > 0:
                    continue
                K[it - 1, P - 1, jj, ii] = 0.1      # KappaHFB(i, j) = Occ
                K[it - 1, P - 1, ii, jj] = -0.1     # KappaHFB(j, i) = -Occ
    return K


class BCS:
    """BCS pairing (PairingType 1): the loop of SolvePairing (Pairing.f90:
    826-858) over BCSPairingField, BCSGaps, BCSFindFermiEnergy,
    BCSQPEnergies and BCSOccupations (BCS.f90:16-247), per state i with
    cutoff factor c_i, gap Delta_i, quasiparticle energy E_i.  The pair
    density of a state with its own time reverse is -rho_i (GetPairDensity,
    PairingInteraction.f90:150-178), so everything is real."""

    def __init__(self, qn, m):
        self.qn = np.asarray(qn)
        self.nwt = len(qn)
        self.it = (self.qn[:, 0] + 3) // 2 - 1
        self.G = np.asarray(m["pairing_strength"], float)
        self.alpha = np.asarray(m["pairing_alpha"], float)
        self.cut = np.asarray(m["pairing_cut"], float)
        self.mu = np.asarray(m["pairing_mu"], float)
        self.rhosat = float(m["rhosat"])
        self.prec = float(m.get("fermi_prec", 1e-12))
        self.iters = int(m.get("pairing_iter", 50))
        self.particles = np.array([m["neutrons"], m["protons"]], float)
        self.delta = np.zeros(self.nwt)
        self.eqp = np.ones(self.nwt)           # Spwf's Eqp default
        self.fermi = np.array([-10.0, -10.0])
        self.pc = np.zeros(self.nwt)
        self.occ = np.zeros(self.nwt)

    cutoffs = HFB.cutoffs
    density_factor = HFB.density_factor

    def solve(self, psi, energy, rho, dv):
        """SolvePairing for PairingType 1; the occupations of the HF basis."""
        energy = np.asarray(energy, float)
        if np.all(np.abs(self.delta) < 1e-5):   # SolvePairing :812
            self.delta[:] = 1.0
        df = self.density_factor(rho)
        r = np.einsum("icp,icp->ip", psi, psi)   # GetDensity of each state
        N = r.shape[1]
        for _ in range(self.iters):
            pc = self.cutoffs(energy, self.fermi)
            field = np.zeros((2, N))
            for i in range(self.nwt):
                field[self.it[i]] += 0.5 * pc[i] ** 2 * self.delta[i] / self.eqp[i] * (-r[i])
            field *= df
            for i in range(self.nwt):
                self.delta[i] = pc[i] ** 2 * dv * np.sum(r[i] * field[self.it[i]])
            old = self.fermi.copy()
            lam = old[self.it]
            self.eqp = np.sqrt((energy - lam) ** 2 + self.delta ** 2)
            for t in range(2):
                sel = self.it == t
                s1 = np.sum(1.0 - energy[sel] / self.eqp[sel])
                s2 = np.sum(1.0 / self.eqp[sel])
                self.fermi[t] = (self.particles[t] - s1) / s2
            self.pc = self.cutoffs(energy, self.fermi)
            lam = self.fermi[self.it]
            self.eqp = np.sqrt((energy - lam) ** 2 + self.delta ** 2)
            self.occ = np.maximum(0.0, 1.0 - (energy - lam) / self.eqp)
            if np.all(np.abs(self.fermi - old) < self.prec):
                break
        return self.occ.copy()

    def energy(self):
        e = np.zeros(2)
        for i in range(self.nwt):
            e[self.it[i]] += self.delta[i] ** 2 / self.eqp[i]
        return -0.5 * e
