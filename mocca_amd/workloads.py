"""Synthetic inputs of the benchmark workloads (numpy, seeded PCG64).

``hpsi_microbench`` is the 32^3 / 200-state FP64 workload: states are a
Gaussian envelope (width 3 fm) times a seeded sum of 8 plane waves with
|k| <= 1.5 fm^-1; the mean fields are built from a Woods-Saxon profile
(R = 4 fm, a = 0.6 fm): U = -50 f, B = 10 f, W_{mu nu} = -eps_{k mu nu} w d_k f,
S = 2 f z-hat (time-odd, so time reversal is broken while parity, signature
and time simplex are kept: the EV8 octant).  ``ho_states`` gives
harmonic-oscillator seeds with the per-axis parity each spinor component
needs; ``full_iteration`` is the 40^3 / 400-state self-consistent workload.
"""
import numpy as np

HBM_HALF = 20.73551910       # hbar^2/2m (MeV fm^2), Force.f90 default hbm/2
EPS = np.zeros((3, 3, 3))
EPS[0, 1, 2] = EPS[1, 2, 0] = EPS[2, 0, 1] = 1
EPS[0, 2, 1] = EPS[2, 1, 0] = EPS[1, 0, 2] = -1


def mesh(nx, ny, nz, dx):
    """Cell-centred octant coordinates x_i = (i - 1/2) dx (Mesh.f90:49-73)."""
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    z = (np.arange(nz) + 0.5) * dx
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def comp_signs(P, S, TS=1):
    """Reflection sign of each spinor component along x, y, z in the EV8
    octant (Derivatives.f90:348-439)."""
    return ([S * TS, -S * TS, -S * TS, S * TS], [TS, -TS, TS, -TS],
            [S * P, S * P, -S * P, -S * P])


def woods_saxon_fields(nx, ny, nz, dx, R=4.0, a=0.6, time_odd=True):
    X, Y, Z = mesh(nx, ny, nz, dx)
    r = np.sqrt(X**2 + Y**2 + Z**2)
    f = 1.0 / (1.0 + np.exp((r - R) / a))
    df = -f * (1 - f) / a
    grad = np.stack([df * X / r, df * Y / r, df * Z / r])
    sh = (nz, ny, nx)
    N = nx * ny * nz
    U = np.stack([-50 * f, -50 * f])
    B = np.stack([10 * f, 10 * f])
    NablaB = np.stack([10 * grad, 10 * grad])
    w = 60 * 0.16
    W = np.zeros((2, 3, 3) + sh)     # [it][nu][mu]
    for mu in range(3):
        for nu in range(3):
            W[:, nu, mu] = -sum(EPS[k, mu, nu] * w * grad[k] for k in range(3))
    DivW = np.zeros((2, 3) + sh)     # sum_mu d_mu W_mu,nu = 0 for this W
    out = dict(U=U, B=B, NablaB=NablaB, W=W, DivW=DivW)
    if time_odd:
        S = np.zeros((2, 3) + sh)
        S[:, 2] = 2 * f
        out.update(S=S, A=np.zeros((2, 3) + sh), DivA=np.zeros((2,) + sh))
    return {k: v.reshape(v.shape[:-3] + (N,)) for k, v in out.items()}


def plane_wave_states(rng, nx, ny, nz, dx, M, width=3.0, nwaves=8, kmax=1.5):
    X, Y, Z = mesh(nx, ny, nz, dx)
    env = np.exp(-(X**2 + Y**2 + Z**2) / (2 * width**2))
    out = np.empty((M, 4, nz * ny * nx))
    for i in range(M):
        for c in range(4):
            k = rng.uniform(-1, 1, (nwaves, 3))
            k *= (kmax * rng.uniform(0, 1, (nwaves, 1)) /
                  np.maximum(np.linalg.norm(k, axis=1, keepdims=True), 1e-12))
            a = rng.uniform(-1, 1, nwaves)
            ph = rng.uniform(0, 2 * np.pi, nwaves)
            f = np.zeros_like(X)
            for q in range(nwaves):
                f += a[q] * np.cos(k[q, 0] * X + k[q, 1] * Y + k[q, 2] * Z + ph[q])
            out[i, c] = (env * f).reshape(-1)
    return out


def hermite(n, t):
    h0, h1 = np.ones_like(t), 2 * t
    if n == 0:
        return h0
    for k in range(1, n):
        h0, h1 = h1, 2 * t * h1 - 2 * k * h0
    return h1


def ho_states(rng, nx, ny, nz, dx, qns, b=1.7, noise=1e-3, nmax=10, terms=3):
    """Harmonic-oscillator seeds: per spinor component, a random combination
    of `terms` Hermite products H_a(x/b) H_b(y/b) H_c(z/b) exp(-r^2/2b^2)
    with degrees below nmax whose per-axis parity matches the component's
    reflection sign, plus uniform noise."""
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    z = (np.arange(nz) + 0.5) * dx
    hx = [hermite(k, x / b) * np.exp(-x * x / (2 * b * b)) for k in range(nmax)]
    hy = [hermite(k, y / b) * np.exp(-y * y / (2 * b * b)) for k in range(nmax)]
    hz = [hermite(k, z / b) * np.exp(-z * z / (2 * b * b)) for k in range(nmax)]
    # scale each degree to comparable size (H_n grows like 2^n n!)
    norm = [1.0 / np.sqrt(2.0**k * np.prod(np.arange(1, k + 1))) for k in range(nmax)]
    pick = {s: [k for k in range(nmax) if (1 if k % 2 == 0 else -1) == s] for s in (1, -1)}
    out = np.empty((len(qns), 4, nz * ny * nx))
    for i, (iso, P, S, TS) in enumerate(qns):
        sx, sy, sz = comp_signs(P, S, TS)
        for c in range(4):
            f = np.zeros((nz, ny, nx))
            for _ in range(terms):
                a, bb, cc = (rng.choice(pick[sx[c]]), rng.choice(pick[sy[c]]),
                             rng.choice(pick[sz[c]]))
                f += (rng.uniform(-1, 1) * norm[a] * norm[bb] * norm[cc] *
                      hz[cc][:, None, None] * hy[bb][None, :, None] * hx[a][None, None, :])
            out[i, c] = (f + noise * rng.uniform(-1, 1, f.shape)).reshape(-1)
    return out


def ho_eigen_states(rng, nx, ny, nz, dx, qns, b=2.4, noise=1e-4):
    """Harmonic-oscillator eigenfunctions H_a(x/b) H_b(y/b) H_c(z/b)
    exp(-r^2/2b^2), shell by shell (a + b + c ascending, then
    lexicographically), one per state: the k-th state of an (isospin, parity)
    takes the k-th function of that parity, in the one spinor component whose
    reflection signs in x, y match the function's parities (component 0
    even-even, 1 odd-odd, 2 odd-even, 3 even-odd), plus uniform noise: the
    filled oscillator shells a mesh solver starts a nucleus from."""
    X, Y, Z = mesh(nx, ny, nz, dx)
    env = np.exp(-(X**2 + Y**2 + Z**2) / (2 * b * b))
    funcs = {1: [], -1: []}
    for N in range(24):
        for a in range(N + 1):
            for bb in range(N + 1 - a):
                funcs[1 if N % 2 == 0 else -1].append((a, bb, N - a - bb))
    comp = {(0, 0): 0, (1, 1): 1, (1, 0): 2, (0, 1): 3}
    count = {}
    out = np.zeros((len(qns), 4, nz * ny * nx))
    for i, (iso, P, S, TS) in enumerate(qns):
        k = count.get((iso, P), 0)
        count[(iso, P)] = k + 1
        a, bb, c = funcs[P][k]
        f = hermite(a, X / b) * hermite(bb, Y / b) * hermite(c, Z / b) * env
        f /= np.sqrt(2.0 ** (a + bb + c) * np.prod(np.arange(1, a + 1)) *
                     np.prod(np.arange(1, bb + 1)) * np.prod(np.arange(1, c + 1)))
        out[i, comp[(a % 2, bb % 2)]] = f.reshape(-1)
        out[i] += noise * rng.uniform(-1, 1, size=out[i].shape)
    return out


def balanced_qn(M, time_reversal):
    """Quantum numbers spread evenly over isospin, parity (and signature
    when time reversal is broken); time simplex +1 (EV8)."""
    qn = []
    for i in range(M):
        iso = -1 if i < M // 2 else 1
        j = i % (M // 2) if M >= 2 else 0
        P = 1 if j % 2 == 0 else -1
        S = 1 if (time_reversal or (j // 2) % 2 == 0) else -1
        qn.append((iso, P, S, 1))
    return np.array(qn, np.int32)


def hpsi_microbench(seed=2, n=32, dx=0.8, M=200):
    """The 32^3 / 200-state FP64 hPsi workload (time-odd S field)."""
    rng = np.random.default_rng(seed)
    qn = balanced_qn(M, time_reversal=False)
    occ = rng.uniform(0.5, 1.0, M)
    psi = plane_wave_states(rng, n, n, n, dx, M)
    fields = woods_saxon_fields(n, n, n, dx, time_odd=True)
    dv = dx**3 * 8
    # orthonormalise-friendly scale: unit norm per state
    psi /= np.sqrt((psi.reshape(M, -1)**2).sum(1) * dv)[:, None, None]
    kin = np.array([HBM_HALF, HBM_HALF])
    return dict(nx=n, ny=n, nz=n, dx=dx, dv=dv, M=M, trc=False, lagrange=False,
                qn=qn, occ=occ, psi=psi, fields=fields, kin=kin,
                dt_over_hbar=0.010 / 6.58211928, momentum=0.5)


def full_iteration(seed=3, n=40, dx=0.8, M=400, neutrons=126, protons=82, b=2.4,
                   seeds="random"):
    """The 40^3 / 400-state self-consistent workload: 200 neutron and 200
    proton states (time reversal, parity, signature and time simplex
    conserved; parity alternating), harmonic-oscillator seeds, SLy4 with
    Coulomb, HF filling of 126 neutrons and 82 protons, fixed dt and
    momentum, linear density mixing."""
    from . import functional
    rng = np.random.default_rng(seed)
    qn = np.array([(iso, 1 if j % 2 == 0 else -1, 1, 1)
                   for iso in (-1, 1) for j in range(M // 2)], np.int32)
    if seeds == "eigen":
        psi = ho_eigen_states(rng, n, n, n, dx, [tuple(q) for q in qn], b=b)
    else:
        psi = ho_states(rng, n, n, n, dx, [tuple(q) for q in qn], b=b)
    return dict(nx=n, ny=n, nz=n, dx=dx, dv=dx**3 * 8, M=M, trc=True, lagrange=False,
                qn=qn, occ=np.zeros(M), psi=psi,
                functional=functional.scf_arguments("SLy4", neutrons, protons, True),
                dt_over_hbar=0.010 / functional.HBAR, momentum=0.3)


def lagrange_48(seed=5, n=48, dx=0.7, M=600, neutrons=12, protons=12, b=1.7):
    """The 48^3 / 600-state Lagrange-mesh workload (Lagrange derivatives of
    48 points per line): 300 neutron and 300 proton states (time reversal,
    parity, signature and time simplex conserved), harmonic-oscillator seeds,
    SLy4 with Coulomb, HF filling of 12 + 12 nucleons."""
    w = full_iteration(seed=seed, n=n, dx=dx, M=M, neutrons=neutrons, protons=protons, b=b)
    w["lagrange"] = True
    return w


def batch_scan(seed=4, nsolves=64, dx=0.8, nmin=16, nmax=32, mmin=20, mmax=160,
               neutrons=12, protons=12, q20=(-300.0, 600.0)):
    """The batched constraint scan: nsolves independent solves of 12 + 12
    nucleons, cubic meshes n ~ U{nmin..nmax}, nwt ~ U{mmin..mmax} (even, half
    neutrons, half protons), harmonic-oscillator seeds (b = 1.7 fm), SLy4 with
    Coulomb, HF, and an alternating (type 2) Q20 constraint spread uniformly
    over q20 (solve k: q20[0] + (q20[1] - q20[0]) k / (nsolves - 1))."""
    from . import functional
    rng = np.random.default_rng(seed)
    out = []
    for k in range(nsolves):
        n = int(rng.integers(nmin, nmax + 1))
        M = 2 * int(rng.integers(mmin // 2, mmax // 2 + 1))
        qn = np.array([(iso, 1 if j % 2 == 0 else -1, 1, 1)
                       for iso in (-1, 1) for j in range(M // 2)], np.int32)
        psi = ho_states(rng, n, n, n, dx, [tuple(q) for q in qn], b=1.7)
        q = q20[0] + (q20[1] - q20[0]) * k / max(nsolves - 1, 1)
        out.append(dict(nx=n, ny=n, nz=n, dx=dx, dv=dx**3 * 8, M=M, trc=True, lagrange=False,
                        qn=qn, occ=np.zeros(M), psi=psi,
                        functional=functional.scf_arguments("SLy4", neutrons, protons, True),
                        constraints=[dict(l=2, m=0, type=2, value=q)],
                        dt_over_hbar=0.010 / functional.HBAR, momentum=0.3))
    return out
