"""Python view of the mocca_amd C ABI (include/mocca_amd.h).

The HIP library ``libmocca_amd.so`` is built in-tree by ``mocca_amd.build``.
Nothing here computes: every call goes to the HIP library, and importing
the bindings without the built library raises at once (no CPU fallback).

The class :class:`Solve` mirrors the reference's module procedures for the
imaginary-time path (SpwfStorage, ImaginaryTime, Densities), one method per
C entry point.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.path.join(HERE, "libmocca_amd.so")

FP64 = 0
MIXED = 1

# Every symbol declared in include/mocca_amd.h.
EXPORTS = (
    "mocca_solve_create", "mocca_solve_destroy", "mocca_last_error",
    "mocca_synchronize", "mocca_set_states_meta", "mocca_set_occupations",
    "mocca_set_energies", "mocca_get_energies", "mocca_upload_states",
    "mocca_download_states", "mocca_reset_updates", "mocca_set_potentials",
    "mocca_derive_all", "mocca_download_derivatives", "mocca_hpsi",
    "mocca_grad_desc", "mocca_gram_schmidt", "mocca_compute_density",
    "mocca_profile_enable", "mocca_profile_get",
    "mocca_scf_setup", "mocca_hf_fill", "mocca_get_occupations", "mocca_update_densities",
    "mocca_mix_densities", "mocca_construct_potentials", "mocca_comp_energy",
    "mocca_convergence_check", "mocca_final_iteration", "mocca_angular_momentum",
    "mocca_set_lagrange_precision",
    "mocca_scf_start", "mocca_iterate", "mocca_get_moments", "mocca_get_coulomb",
    "mocca_get_potentials", "mocca_estimate_parameters", "mocca_estimation_seed",
    "mocca_iterate_estimated", "mocca_set_constraints", "mocca_readjust_constraints",
    "mocca_alternate_step", "mocca_get_constraints", "mocca_set_pairing",
    "mocca_solve_pairing", "mocca_get_pairing", "mocca_batch_run",
    "mocca_gram_schmidt_noalloc",
    "mocca_v1_read", "mocca_v1_write", "mocca_v1_save", "mocca_v1_load",
)

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


class MeshC(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dv", ctypes.c_double),
                ("time_reversal", ctypes.c_int), ("lagrange", ctypes.c_int)]


class FieldsC(ctypes.Structure):
    _fields_ = [(k, _dp) for k in ("U", "B", "NablaB", "W", "DivW", "S", "A", "DivA")] + \
        [("kin", ctypes.c_double * 2)]


class DensityC(ctypes.Structure):
    _fields_ = [(k, _dp) for k in ("rho", "tau", "nablaJ", "derrho", "laprho",
                                   "vecj", "vecs", "rotvecj", "rots", "ders")]


class FunctionalC(ctypes.Structure):
    _fields_ = [("B", ctypes.c_double * 28), ("hbm", ctypes.c_double * 2),
                ("mass", ctypes.c_double * 2), ("e2", ctypes.c_double),
                ("neutrons", ctypes.c_double), ("protons", ctypes.c_double),
                ("com1body", ctypes.c_int), ("cexchange", ctypes.c_int),
                ("coulomb", ctypes.c_int), ("coulomb_maxit", ctypes.c_int),
                ("max_moment", ctypes.c_int), ("damping", ctypes.c_double),
                ("mixing_scheme", ctypes.c_int), ("pad_", ctypes.c_int),
                ("preconu", ctypes.c_double)]


class ConstraintC(ctypes.Structure):
    _fields_ = [("l", ctypes.c_int), ("m", ctypes.c_int), ("type", ctypes.c_int),
                ("isoswitch", ctypes.c_int), ("iteration", ctypes.c_int), ("pad_", ctypes.c_int),
                ("value", ctypes.c_double * 2), ("intensity", ctypes.c_double)]


# Energy row layout of mocca_comp_energy / mocca_iterate.
ENERGY_FIELDS = ("total", "kinetic_n", "kinetic_p", "coulomb", "coulomb_exchange",
                 "com_n", "com_p") + tuple(f"term{k}" for k in range(1, 37)) + \
    ("pairing_n", "pairing_p", "fermi_n", "fermi_p")
ROW_LEN = len(ENERGY_FIELDS)


class PairingC(ctypes.Structure):
    _fields_ = [("type", ctypes.c_int), ("guess_kappa", ctypes.c_int),
                ("iterations", ctypes.c_int), ("solver", ctypes.c_int),
                ("strength", ctypes.c_double * 2), ("alpha", ctypes.c_double * 2),
                ("cut", ctypes.c_double * 2), ("mu", ctypes.c_double * 2),
                ("rhosat", ctypes.c_double), ("hfb_mix", ctypes.c_double),
                ("fermi_prec", ctypes.c_double), ("fermi", ctypes.c_double * 2)]


V1_MAX_MOMENTS = 32


class V1MomentC(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int) for k in ("constraint_type", "isoswitch", "l", "m", "impart")] + \
        [(k, ctypes.c_double * 2) for k in ("intensity", "constraint", "true_constraint",
                                            "value")] + [("old_value", ctypes.c_double * 14)]


class V1InfoC(ctypes.Structure):
    _fields_ = [("energy", ctypes.c_double), ("denergy", ctypes.c_double * 7),
                ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double)] + \
        [(k, ctypes.c_int) for k in ("trc", "pc", "sc", "tsc", "ic", "nwt")] + \
        [("neutrons", ctypes.c_double), ("protons", ctypes.c_double),
         ("fd_order", ctypes.c_int * 3), ("force", ctypes.c_char * 200),
         ("skyrme", ctypes.c_double * 14), ("hbar", ctypes.c_double),
         ("hbm", ctypes.c_double * 2), ("com1body", ctypes.c_int), ("com2body", ctypes.c_int),
         ("B", ctypes.c_double * 28), ("pairing_type", ctypes.c_int),
         ("pairing_strength", ctypes.c_double * 2), ("pairing_alpha", ctypes.c_double * 2),
         ("pairing_cut", ctypes.c_double * 2), ("has_cutoffs", ctypes.c_int),
         ("fermi", ctypes.c_double * 2), ("lnlambda", ctypes.c_double * 2),
         ("hfb_size", ctypes.c_int), ("npar", ctypes.c_int), ("niso", ctypes.c_int),
         ("blocksizes", ctypes.c_int * 4), ("omega", ctypes.c_double * 3),
         ("crank_values", ctypes.c_double * 3), ("nmoments", ctypes.c_int),
         ("moments", V1MomentC * V1_MAX_MOMENTS)]


class MoccaError(RuntimeError):
    pass


_LIB = None


def load():
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIBPATH):
        raise MoccaError(f"{LIBPATH} is missing: build it with `python -m mocca_amd.build`")
    lib = ctypes.CDLL(LIBPATH)
    lib.mocca_last_error.restype = ctypes.c_char_p
    lib.mocca_last_error.argtypes = [ctypes.c_void_p]
    lib.mocca_solve_create.argtypes = [ctypes.POINTER(MeshC), ctypes.c_int, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_void_p)]
    lib.mocca_gram_schmidt_noalloc.argtypes = [
        ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_long, _ip, ctypes.c_void_p,
        ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p]
    for name in EXPORTS:
        f = getattr(lib, name)
        if name not in ("mocca_last_error", "mocca_solve_create"):
            f.restype = ctypes.c_int
    _LIB = lib
    return lib


def gram_schmidt_noalloc(psi, out, nwt, n, qn, energy, dv, status=None, stream=None):
    """mocca_gram_schmidt_noalloc: psi, out, energy, status are device
    addresses (ints), qn a host array [nwt][4], stream a hipStream_t handle
    (None: the default stream)."""
    lib = load()
    qn = np.ascontiguousarray(qn, dtype=np.int32).reshape(nwt, 4)
    rc = lib.mocca_gram_schmidt_noalloc(psi, out, nwt, n, _p(qn), energy, dv, status, stream)
    if rc != 0:
        msg = lib.mocca_last_error(None)
        raise MoccaError(f"mocca error {rc}: {msg.decode() if msg else ''}")


class V1ArraysC(ctypes.Structure):
    pass


def _v1_arrays_type():
    if not hasattr(V1ArraysC, "_fields_"):
        V1ArraysC._fields_ = [("grid", _dp), ("state", _dp), ("qn", _ip), ("cutoffs", _dp),
                              ("dens", DensityC), ("u", _dp), ("v", _dp),
                              ("hfb_columns", _ip), ("rho_hfb", _dp), ("kappa_hfb", _dp),
                              ("can_transfo", _dp)]
    return V1ArraysC


def _v1_chk(lib, rc):
    if rc != 0:
        msg = lib.mocca_last_error(None)
        raise MoccaError(f"mocca error {rc}: {msg.decode() if msg else ''}")


def _info_dict(c):
    """V1InfoC -> dict of plain values (the constrained moments as a list)."""
    d = {}
    for name, _ in V1InfoC._fields_:
        if name == "moments":
            continue
        v = getattr(c, name)
        d[name] = bytes(v) if name == "force" else (list(v) if hasattr(v, "__len__") else v)
    d["force"] = bytes(c.force).ljust(200, b" ")
    d["moments"] = [{k: (list(getattr(m, k)) if hasattr(getattr(m, k), "__len__") else
                         getattr(m, k)) for k, _ in V1MomentC._fields_}
                    for m in c.moments[:c.nmoments]]
    return d


def _info_c(d):
    c = V1InfoC()
    for name, typ in V1InfoC._fields_:
        if name == "moments" or name not in d:
            continue
        v = d[name]
        if name == "force":
            c.force = bytes(v)[:200].ljust(200, b" ")
        elif hasattr(typ, "_length_"):
            getattr(c, name)[:] = list(v)
        else:
            setattr(c, name, v)
    ms = d.get("moments", [])
    c.nmoments = len(ms)
    for k, m in enumerate(ms):
        for name, typ in V1MomentC._fields_:
            if hasattr(typ, "_length_"):
                getattr(c.moments[k], name)[:] = list(m[name])
            else:
                setattr(c.moments[k], name, int(m[name]))
    return c


def _v1_buffers(info, N):
    """Host arrays of a v1 file's per-state and array records."""
    M, S = info["nwt"], info["hfb_size"]
    nb = info["npar"] * info["niso"]
    a = dict(grid=np.zeros((M, 4, N)), state=np.zeros((M, 7)), qn=np.zeros((M, 5), np.int32),
             cutoffs=np.zeros(M), rho=np.zeros((2, N)), derrho=np.zeros((2, 3, N)),
             laprho=np.zeros((2, N)), tau=np.zeros((2, N)), nablaJ=np.zeros((2, N)),
             u=np.zeros((4, 2 * S, S), np.complex128), v=np.zeros((4, 2 * S, S), np.complex128),
             hfb_columns=np.zeros((nb, S), np.int32),
             rho_hfb=np.zeros((4, S, S), np.complex128),
             kappa_hfb=np.zeros((4, S, S), np.complex128),
             can_transfo=np.zeros((nb, S, S), np.complex128))
    if not info["trc"]:
        a.update(vecj=np.zeros((2, 3, N)), rotvecj=np.zeros((2, 3, N)), vecs=np.zeros((2, 3, N)),
                 rots=np.zeros((2, 3, N)), ders=np.zeros((2, 3, 3, N)))
    return a


def _v1_arrays_c(a):
    T = _v1_arrays_type()
    c = T()

    def dp(x):
        if x is None:
            return None
        return (x.view(np.float64) if x.dtype == np.complex128 else x).ctypes.data_as(_dp)

    c.grid, c.state, c.cutoffs = dp(a.get("grid")), dp(a.get("state")), dp(a.get("cutoffs"))
    c.qn = a["qn"].ctypes.data_as(_ip) if a.get("qn") is not None else None
    for k in ("rho", "tau", "nablaJ", "derrho", "laprho", "vecj", "vecs", "rotvecj", "rots",
              "ders"):
        setattr(c.dens, k, dp(a.get(k)))
    c.u, c.v = dp(a.get("u")), dp(a.get("v"))
    c.hfb_columns = (a["hfb_columns"].ctypes.data_as(_ip)
                     if a.get("hfb_columns") is not None else None)
    c.rho_hfb, c.kappa_hfb = dp(a.get("rho_hfb")), dp(a.get("kappa_hfb"))
    c.can_transfo = dp(a.get("can_transfo"))
    return c


def v1_read(path, arrays=True):
    """mocca_v1_read: (info dict, dict of host arrays or None).  Array
    shapes: grid [nwt][4][N], state [nwt][7] (Occupation, Energy,
    Dispersion, Norm, J), qn [nwt][5], densities as mocca_density, the HFB
    matrices complex [block][column][row] (Fortran order)."""
    lib = load()
    c = V1InfoC()
    _v1_chk(lib, lib.mocca_v1_read(os.fsencode(path), ctypes.byref(c), None))
    info = _info_dict(c)
    if not arrays:
        return info, None
    a = _v1_buffers(info, info["nx"] * info["ny"] * info["nz"])
    ac = _v1_arrays_c(a)
    _v1_chk(lib, lib.mocca_v1_read(os.fsencode(path), ctypes.byref(c), ctypes.byref(ac)))
    if not info["has_cutoffs"]:
        a["cutoffs"] = None
    return info, a


def v1_write(path, info, arrays):
    """mocca_v1_write from v1_read's pair (info dict, arrays dict)."""
    lib = load()
    c = _info_c(info)
    a = {k: (np.ascontiguousarray(v) if v is not None else None) for k, v in arrays.items()}
    _v1_chk(lib, lib.mocca_v1_write(os.fsencode(path), ctypes.byref(c),
                                    ctypes.byref(_v1_arrays_c(a))))


def _p(a):
    if a is None:
        return None
    assert a.flags.c_contiguous
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def _f64(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64)


class Solve:
    """A device-resident set of nwt states on one mesh (one HIP stream)."""

    def __init__(self, nx, ny, nz, dx, dv, nwt, time_reversal=True, lagrange=False,
                 precision=FP64):
        self.lib = load()
        self.nx, self.ny, self.nz, self.N, self.M = nx, ny, nz, nx * ny * nz, nwt
        self.trc = bool(time_reversal)
        mesh = MeshC(nx, ny, nz, dx, dv, int(bool(time_reversal)), int(bool(lagrange)))
        h = ctypes.c_void_p()
        self._chk(self.lib.mocca_solve_create(ctypes.byref(mesh), nwt, precision,
                                              ctypes.byref(h)), None)
        self.h = h
        self._keep = []

    def _chk(self, rc, h=-1):
        if rc != 0:
            msg = self.lib.mocca_last_error(self.h if h == -1 else h)
            raise MoccaError(f"mocca error {rc}: {msg.decode() if msg else ''}")

    def close(self):
        if getattr(self, "h", None):
            self.lib.mocca_solve_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # SpwfStorage --------------------------------------------------------
    def set_states_meta(self, qn, occ):
        qn = np.ascontiguousarray(qn, dtype=np.int32).reshape(self.M, 4)
        occ = _f64(occ)
        self._chk(self.lib.mocca_set_states_meta(self.h, _p(qn), _p(occ)))

    def set_occupations(self, occ):
        self._chk(self.lib.mocca_set_occupations(self.h, _p(_f64(occ))))

    def set_energies(self, e):
        self._chk(self.lib.mocca_set_energies(self.h, _p(_f64(e))))

    def energies(self):
        e, d = np.zeros(self.M), np.zeros(self.M)
        self._chk(self.lib.mocca_get_energies(self.h, _p(e), _p(d)))
        return e, d

    def upload_states(self, psi, first=0):
        psi = _f64(psi).reshape(-1, 4 * self.N)
        self._chk(self.lib.mocca_upload_states(self.h, first, psi.shape[0], _p(psi)))

    def download_states(self, first=0, count=None):
        count = self.M - first if count is None else count
        out = np.zeros((count, 4, self.N))
        self._chk(self.lib.mocca_download_states(self.h, first, count, _p(out)))
        return out

    def reset_updates(self):
        self._chk(self.lib.mocca_reset_updates(self.h))

    def derive_all(self):
        self._chk(self.lib.mocca_derive_all(self.h))

    def derivatives(self, state):
        der, lap = np.zeros((3, 4, self.N)), np.zeros((4, self.N))
        self._chk(self.lib.mocca_download_derivatives(self.h, state, _p(der), _p(lap)))
        return der, lap

    def gram_schmidt(self):
        self._chk(self.lib.mocca_gram_schmidt(self.h))

    # MeanFields / ImaginaryTime -------------------------------------------
    def set_potentials(self, U, B, NablaB, W, DivW, kin, S=None, A=None, DivA=None):
        arrs = [_f64(x) for x in (U, B, NablaB, W, DivW, S, A, DivA)]
        self._keep = arrs
        f = FieldsC(*[_p(a) if a is not None else None for a in arrs],
                    (ctypes.c_double * 2)(*kin))
        self._chk(self.lib.mocca_set_potentials(self.h, ctypes.byref(f)))

    def hpsi(self, state):
        out = np.zeros((4, self.N))
        self._chk(self.lib.mocca_hpsi(self.h, state, _p(out)))
        return out

    def grad_desc(self, dt_over_hbar, momentum):
        self._chk(self.lib.mocca_grad_desc(self.h, ctypes.c_double(dt_over_hbar),
                                           ctypes.c_double(momentum)))

    # Densities -----------------------------------------------------------
    def compute_density(self, fetch=True):
        N = self.N
        if not fetch:
            self._chk(self.lib.mocca_compute_density(self.h, None))
            return None
        out = dict(rho=np.zeros((2, N)), tau=np.zeros((2, N)), nablaJ=np.zeros((2, N)),
                   derrho=np.zeros((2, 3, N)), laprho=np.zeros((2, N)))
        if not self.trc:
            out.update(vecj=np.zeros((2, 3, N)), vecs=np.zeros((2, 3, N)),
                       rotvecj=np.zeros((2, 3, N)), rots=np.zeros((2, 3, N)),
                       ders=np.zeros((2, 3, 3, N)))
        keys = ("rho", "tau", "nablaJ", "derrho", "laprho", "vecj", "vecs", "rotvecj",
                "rots", "ders")
        d = DensityC(*[_p(out[k]) if k in out else None for k in keys])
        self._chk(self.lib.mocca_compute_density(self.h, ctypes.byref(d)))
        return out

    def synchronize(self):
        self._chk(self.lib.mocca_synchronize(self.h))

    # Self-consistent part -------------------------------------------------
    def scf_setup(self, B, hbm, mass, e2, neutrons, protons, com1body=2, cexchange=2,
                  coulomb=1, coulomb_maxit=500, max_moment=10, damping=0.75, mixing_scheme=0,
                  preconu=1.0):
        f = FunctionalC((ctypes.c_double * 28)(*[float(x) for x in B]),
                        (ctypes.c_double * 2)(*hbm), (ctypes.c_double * 2)(*mass),
                        float(e2), float(neutrons), float(protons), int(com1body),
                        int(cexchange), int(coulomb), int(coulomb_maxit), int(max_moment),
                        float(damping), int(mixing_scheme), 0, float(preconu))
        self._chk(self.lib.mocca_scf_setup(self.h, ctypes.byref(f)))

    def hf_fill(self):
        self._chk(self.lib.mocca_hf_fill(self.h))

    def occupations(self):
        occ = np.zeros(self.M)
        self._chk(self.lib.mocca_get_occupations(self.h, _p(occ)))
        return occ

    def update_densities(self):
        self._chk(self.lib.mocca_update_densities(self.h))

    def mix_densities(self):
        self._chk(self.lib.mocca_mix_densities(self.h))

    def construct_potentials(self):
        self._chk(self.lib.mocca_construct_potentials(self.h))

    def comp_energy(self):
        e = np.zeros(ROW_LEN)
        self._chk(self.lib.mocca_comp_energy(self.h, _p(e)))
        return e

    def scf_start(self):
        self._chk(self.lib.mocca_scf_start(self.h))

    def convergence_check(self, moment_prec=1e-4, energy_prec=1e-9, pairing_prec=1e-4,
                          crank_prec=1e-4):
        """ConvergenceCheck after the last CompEnergy: dict with converged,
        positive_energy (the reference's stop condition) and the four tests."""
        prec = np.array([moment_prec, energy_prec, pairing_prec, crank_prec], np.float64)
        out = (ctypes.c_int * 6)()
        self._chk(self.lib.mocca_convergence_check(self.h, _p(prec), out))
        keys = ("converged", "positive_energy", "moments", "energy", "fermi", "cranking")
        return {k: bool(v) for k, v in zip(keys, out)}

    def set_lagrange_precision(self, mode):
        """Lagrange derivatives on the matrix cores: 0 FP64, 1 FP32 with split
        matrices (mixed precision only)."""
        self._chk(self.lib.mocca_set_lagrange_precision(self.h, int(mode)))

    def angular_momentum(self, canonical=False):
        """UpdateAM: per-state [M][7][3] (J, JTR, JTI, J2, Sp, SpTR, SpTI) of
        the HF basis (and of the canonical basis when asked), and the totals
        over the density basis: [5][3] (TotalAngMom, JTR, JTI, J2Total,
        TotalAngMom_dens) and AMIsoblock [3][2][2][2] (direction, isospin,
        parity, signature)."""
        st = np.zeros((self.M, 21))
        cn = np.zeros((self.M, 21)) if canonical else None
        tot = np.zeros(39)
        self._chk(self.lib.mocca_angular_momentum(self.h, _p(st), _p(cn) if canonical else None,
                                                  _p(tot)))
        out = dict(states=st.reshape(self.M, 7, 3), totals=tot[:15].reshape(5, 3),
                   isoblock=tot[15:].reshape(3, 2, 2, 2))
        if canonical:
            out["canonical"] = cn.reshape(self.M, 7, 3)
        return out

    def final_iteration(self):
        """FinalIteration: the energy row with Lagrange derivatives; the
        handle's densities are then the unmixed ones of its own scheme."""
        e = np.zeros(ROW_LEN)
        self._chk(self.lib.mocca_final_iteration(self.h, _p(e)))
        return e

    def save_v1(self, path, **extra):
        """mocca_v1_save: a checkpoint in the MOCCa v1 format.  extra: any
        mocca_v1_info member the handle does not hold (fd_order, force,
        skyrme, hbar, com2body, omega, crank_values, fermi for HF)."""
        if extra:
            base = dict(fd_order=[0, 0, 2], force=b"", hbar=6.58211928, fermi=[-10.0, -10.0])
            base.update(extra)
            c = _info_c(base)
            self._chk(self.lib.mocca_v1_save(self.h, os.fsencode(path), ctypes.byref(c)))
        else:
            self._chk(self.lib.mocca_v1_save(self.h, os.fsencode(path), None))

    def load_v1(self, path, recalc=False):
        """mocca_v1_load: restart from a v1 file; returns its header dict."""
        c = V1InfoC()
        self._chk(self.lib.mocca_v1_load(self.h, os.fsencode(path), int(bool(recalc)),
                                         ctypes.byref(c)))
        return _info_dict(c)

    def iterate(self, niter, dt_over_hbar, momentum, fetch=True):
        e = np.zeros((niter + 1, ROW_LEN)) if fetch else None
        self._chk(self.lib.mocca_iterate(self.h, int(niter), ctypes.c_double(dt_over_hbar),
                                         ctypes.c_double(momentum), _p(e)))
        return e

    def set_constraints(self, constraints, radd=4.0, acut=0.4, damping=0.0,
                        readjust_slowdown=0.01):
        """constraints: dicts with l, m, and optionally type (2), isoswitch (1),
        iteration (-1: never switched off), value (isoswitch 1) or
        value_n, value_p (isoswitch 2), intensity (0: estimated)."""
        arr = (ConstraintC * max(1, len(constraints)))()
        self._nconstraints = len(constraints)
        for k, c in enumerate(constraints):
            iso = int(c.get("isoswitch", 1))
            v = ((c.get("value", 0.0), 0.0) if iso == 1
                 else (c.get("value_n", 0.0), c.get("value_p", 0.0)))
            arr[k] = ConstraintC(int(c["l"]), int(c.get("m", 0)), int(c.get("type", 2)), iso,
                                 int(c.get("iteration", -1)), 0,
                                 (ctypes.c_double * 2)(*[float(x) for x in v]),
                                 float(c.get("intensity", 0.0)))
        self._chk(self.lib.mocca_set_constraints(self.h, arr, len(constraints),
                                                 ctypes.c_double(radd), ctypes.c_double(acut),
                                                 ctypes.c_double(damping),
                                                 ctypes.c_double(readjust_slowdown)))

    def readjust_constraints(self, iteration, momentum):
        self._chk(self.lib.mocca_readjust_constraints(self.h, int(iteration),
                                                      ctypes.c_double(momentum)))

    def alternate_step(self):
        self._chk(self.lib.mocca_alternate_step(self.h))

    def constraints(self):
        """[count][10]: l, m, type, target(2), intensity(2), <Q^2>(2), isoswitch."""
        n = getattr(self, "_nconstraints", 0)
        out = np.zeros((max(n, 1), 10))
        self._chk(self.lib.mocca_get_constraints(self.h, _p(out)))
        return out[:n]

    def set_pairing(self, type=2, strength=(1250.0, 1250.0), alpha=(1.0, 1.0), cut=(5.0, 5.0),
                    mu=(0.5, 0.5), rhosat=0.16, hfb_mix=0.1, iterations=50, fermi_prec=1e-12,
                    fermi=(-10.0, -10.0), guess_kappa=True, solver=0):
        """Pairing of SolvePairing: type 0 HF, 1 BCS (iterations = PairingIter),
        2 HFB (namelist Pairing defaults);
        solver 0 automatic, 1 in-LDS Broyden (blocks <= 32 states), 2 large-block
        Broyden with Jacobi diagonalisations, 3 large-block Newton, 4 large-block
        Broyden with the sign function (blocks <= 128; the automatic choice above
        32 states)."""
        d2 = lambda v: (ctypes.c_double * 2)(*[float(x) for x in v])
        p = PairingC(int(type), int(bool(guess_kappa)), int(iterations), int(solver),
                     d2(strength),
                     d2(alpha), d2(cut), d2(mu), float(rhosat), float(hfb_mix),
                     float(fermi_prec), d2(fermi))
        self._chk(self.lib.mocca_set_pairing(self.h, ctypes.byref(p)))

    def solve_pairing(self):
        self._chk(self.lib.mocca_solve_pairing(self.h))

    def pairing(self):
        """(fermi[2], pairing energy[2], qp energies[nwt], canonical occupations[nwt])."""
        out = np.zeros(4 + self.M)
        occ = np.zeros(self.M)
        self._chk(self.lib.mocca_get_pairing(self.h, _p(out), _p(occ)))
        return out[0:2], out[2:4], out[4:], occ

    def estimate_parameters(self, iteration, dt, momentum):
        """IterativeEstimation: (dt, momentum, maxE, relE, con, iterations)."""
        out = np.zeros(6)
        self._chk(self.lib.mocca_estimate_parameters(self.h, int(iteration), ctypes.c_double(dt),
                                                     ctypes.c_double(momentum), _p(out)))
        return out

    def estimation_seed(self, grid):
        """Start vector [4][N] of the estimation state for its next start."""
        g = None if grid is None else np.ascontiguousarray(grid, dtype=np.float64)
        if g is not None and g.size != 4 * self.N:
            raise ValueError("seed grid must hold 4 N values")
        self._chk(self.lib.mocca_estimation_seed(self.h, _p(g)))

    def iterate_estimated(self, niter, dt, momentum, first_iteration=1):
        """mocca_iterate with IterativeEstimation before every step: returns
        (energy rows [niter+1][47], params [niter][2] = dt, momentum used)."""
        e = np.zeros((niter + 1, ROW_LEN))
        prm = np.zeros((max(niter, 1), 2))
        self._chk(self.lib.mocca_iterate_estimated(self.h, int(first_iteration), int(niter),
                                                   ctypes.c_double(dt), ctypes.c_double(momentum),
                                                   _p(e), _p(prm)))
        return e, prm[:niter]

    def moments(self):
        lm = np.zeros(2 * 28, np.int32)
        q = np.zeros(2 * 28)
        cnt = ctypes.c_int()
        self._chk(self.lib.mocca_get_moments(self.h, _p(lm), _p(q), ctypes.byref(cnt)))
        n = cnt.value
        return lm[:2 * n].reshape(n, 2), q[:2 * n].reshape(n, 2)

    def coulomb(self):
        N = self.N
        vc, vx = np.zeros(N), np.zeros(N)
        cp = np.zeros((self.nz + 2) * (self.ny + 2) * (self.nx + 2))
        it = ctypes.c_int()
        self._chk(self.lib.mocca_get_coulomb(self.h, _p(vc), _p(vx), _p(cp), ctypes.byref(it)))
        return vc, vx, cp, it.value

    def potentials(self):
        N = self.N
        out = dict(U=np.zeros((2, N)), B=np.zeros((2, N)), NablaB=np.zeros((2, 3, N)),
                   W=np.zeros((2, 3, 3, N)), DivW=np.zeros((2, 3, N)), S=np.zeros((2, 3, N)),
                   A=np.zeros((2, 3, N)), DivA=np.zeros((2, N)))
        keys = ("U", "B", "NablaB", "W", "DivW", "S", "A", "DivA")
        self._chk(self.lib.mocca_get_potentials(self.h, *[_p(out[k]) for k in keys]))
        return out

    def profile_enable(self, on=True):
        self._chk(self.lib.mocca_profile_enable(self.h, int(on)))

    def profile_get(self, name):
        ms, cnt = ctypes.c_double(), ctypes.c_long()
        self._chk(self.lib.mocca_profile_get(self.h, name.encode(), ctypes.byref(ms),
                                             ctypes.byref(cnt)))
        return ms.value, cnt.value


def batch_run(solves, niter, dt_over_hbar, momentum, fetch=True):
    """mocca_batch_run: niter Evolve iterations of every solve (each after
    scf_start), issued in turn on the solves' own streams.  dt_over_hbar and
    momentum: one value or one per solve.  Returns the energy rows
    [niter+1][ROW_LEN] per solve (or None)."""
    lib = load()
    n = len(solves)
    dts = np.broadcast_to(np.asarray(dt_over_hbar, dtype=np.float64), (n,)).copy()
    mus = np.broadcast_to(np.asarray(momentum, dtype=np.float64), (n,)).copy()
    hs = (ctypes.c_void_p * n)(*[s.h for s in solves])
    rows = [np.zeros((niter + 1, ROW_LEN)) for _ in range(n)] if fetch else None
    ptrs = (_dp * n)(*[_p(r) for r in rows]) if fetch else None
    rc = lib.mocca_batch_run(hs, n, int(niter), _p(dts), _p(mus), ptrs)
    if rc != 0:
        msg = lib.mocca_last_error(None)
        raise MoccaError(f"mocca error {rc}: {msg.decode() if msg else ''}")
    return rows
