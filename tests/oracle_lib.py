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
