"""Write the golden fixtures under tests/golden/ from the reference's own code.

Runs only in a container that holds the reference sources: the driver
oracle/ref_driver.f90 is linked against the reference's Fortran modules by
oracle/ref.mk (``make -f oracle/ref.mk``).  The fixtures it writes are plain
numbers (inputs and the reference's outputs) and are what the tests read; the
reference itself never travels with them.

    python tests/golden/make_golden.py            # all cases
    python tests/golden/make_golden.py ev8_trc_fd # one case

Inputs are seeded (numpy PCG64): harmonic-oscillator products (b = 1.7 fm)
whose per-axis parity follows each spinor component's reflection sign, plus
1e-3 uniform noise, and smooth Woods-Saxon mean fields (R = 4 fm, a = 0.6 fm).
Meshes are deliberately non-cubic so that axis mix-ups show.
"""
import json
import os
import shutil
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
DRIVER = os.path.join(ROOT, "oracle", "_ref", "mocca_ref_driver")
RUNDIR = os.path.join(ROOT, "oracle", "_ref", "run")
FORCES = "/root/reference/forces.param"
OUTDIR = os.path.dirname(os.path.abspath(__file__))

# Per-component reflection signs (EV8 rules, Derivatives.f90:348-439 and
# OptimizedDerivatives.f90:533-542) for a state with parity P, signature S,
# time simplex TS = +1.
def comp_signs(P, S, TS=1):
    sx = [S * TS, -S * TS, -S * TS, S * TS]
    sy = [TS, -TS, TS, -TS]
    sz = [S * P, S * P, -S * P, -S * P]
    return sx, sy, sz


def hermite(n, t):
    h0, h1 = np.ones_like(t), 2 * t
    if n == 0:
        return h0
    for k in range(1, n):
        h0, h1 = h1, 2 * t * h1 - 2 * k * h0
    return h1


def mesh(nx, ny, nz, dx):
    x = (np.arange(nx) + 0.5) * dx
    y = (np.arange(ny) + 0.5) * dx
    z = (np.arange(nz) + 0.5) * dx
    Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
    return X, Y, Z


def seeded_states(rng, nx, ny, nz, dx, qns, b=1.7, noise=1e-3):
    X, Y, Z = mesh(nx, ny, nz, dx)
    env = np.exp(-(X**2 + Y**2 + Z**2) / (2 * b * b))
    out = np.zeros((len(qns), 4, nz, ny, nx))
    for i, (iso, P, S, TS) in enumerate(qns):
        sx, sy, sz = comp_signs(P, S, TS)
        for c in range(4):
            f = np.zeros_like(X)
            for _ in range(2):
                ex = [n for n in range(4) if (1 if n % 2 == 0 else -1) == sx[c]]
                ey = [n for n in range(4) if (1 if n % 2 == 0 else -1) == sy[c]]
                ez = [n for n in range(4) if (1 if n % 2 == 0 else -1) == sz[c]]
                a = rng.uniform(-1, 1)
                f += a * hermite(rng.choice(ex), X / b) * hermite(rng.choice(ey), Y / b) \
                    * hermite(rng.choice(ez), Z / b)
            out[i, c] = f * env + noise * rng.uniform(-1, 1, size=f.shape)
    return out


def ho_basis_states(rng, n, dx, qns, b=1.5, noise=1e-4):
    """Harmonic-oscillator eigenfunctions H_a(x/b) H_b(y/b) H_c(z/b)
    exp(-r^2/2b^2), shell by shell (a + b + c ascending, then
    lexicographically), one per state: the k-th state of an (isospin, parity)
    takes the k-th function of that parity, in the one spinor component whose
    reflection signs match the function's per-axis parities (for P = S = TS =
    +1 or P = -1: component 0 even-even, 1 odd-odd, 2 odd-even, 3 even-odd in
    x, y), plus uniform noise."""
    X, Y, Z = mesh(n, n, n, dx)
    env = np.exp(-(X**2 + Y**2 + Z**2) / (2 * b * b))
    funcs = {1: [], -1: []}
    for N in range(16):
        for a in range(N + 1):
            for bb in range(N + 1 - a):
                c = N - a - bb
                funcs[1 if N % 2 == 0 else -1].append((a, bb, c))
    comp = {(0, 0): 0, (1, 1): 1, (1, 0): 2, (0, 1): 3}
    count = {}
    out = np.zeros((len(qns), 4, n, n, n))
    for i, (iso, P, S, TS) in enumerate(qns):
        k = count.get((iso, P), 0)
        count[(iso, P)] = k + 1
        a, bb, c = funcs[P][k]
        f = hermite(a, X / b) * hermite(bb, Y / b) * hermite(c, Z / b) * env
        f /= np.sqrt(2.0 ** (a + bb + c) * np.prod(np.arange(1, a + 1)) *
                     np.prod(np.arange(1, bb + 1)) * np.prod(np.arange(1, c + 1)))
        out[i, comp[(a % 2, bb % 2)]] = f
        out[i] += noise * rng.uniform(-1, 1, size=out[i].shape)
    return out


def synthetic_fields(rng, nx, ny, nz, dx, R=4.0, a=0.6):
    X, Y, Z = mesh(nx, ny, nz, dx)
    r = np.sqrt(X**2 + Y**2 + Z**2)
    f = 1.0 / (1.0 + np.exp((r - R) / a))
    df = -f * (1 - f) / a                      # d f / d r
    grad = [df * X / r, df * Y / r, df * Z / r]
    eps = np.zeros((3, 3, 3))
    eps[0, 1, 2] = eps[1, 2, 0] = eps[2, 0, 1] = 1
    eps[0, 2, 1] = eps[2, 1, 0] = eps[1, 0, 2] = -1
    sh = (nz, ny, nx)
    U = np.stack([-50 * f, -45 * f]) + 0.1 * rng.uniform(-1, 1, (2,) + sh)
    B = np.stack([10 * f, 11 * f])
    NablaB = np.stack([[10 * g for g in grad], [11 * g for g in grad]])
    w = 60 * 0.16
    W = np.zeros((2, 3, 3) + sh)               # [it][nu][mu]
    for it in range(2):
        for mu in range(3):
            for nu in range(3):
                W[it, nu, mu] = -sum(eps[k, mu, nu] * w * (1 + 0.1 * it) * grad[k]
                                     for k in range(3))
    DivW = np.stack([np.stack(grad) * 3.0, np.stack(grad) * 2.5]) + \
        0.05 * rng.uniform(-1, 1, (2, 3) + sh)
    S = np.zeros((2, 3) + sh)
    S[:, 2] = 2 * f
    S[:, 0] = 0.3 * f * X
    S[:, 1] = -0.2 * f * Y
    A = np.stack([np.stack(grad) * 0.7, np.stack(grad) * -0.4])
    DivA = np.stack([0.5 * f, -0.3 * f])
    return dict(U=U, B=B, NablaB=NablaB, W=W, DivW=DivW, S=S, A=A, DivA=DivA)


CASES = {
    # name: (nx, ny, nz, dx, trc, lagrange)
    "ev8_trc_fd": (8, 7, 6, 0.8, True, False),
    "ev8_trc_lag": (8, 7, 6, 0.8, True, True),
    "ev8_todd_fd": (7, 8, 6, 0.75, False, False),
    "ev8_todd_lag": (7, 6, 8, 0.75, False, True),
}


def namelists(nx, ny, nz, dx, trc, lag, nwt, neutrons, protons, dt, mom, mixing=0, pe=0,
              preconw=0.0, constraints=(), pairing='Type="HF"'):
    fd = "-1" if lag else "3"
    fdl = "-1" if lag else "4"
    return f"""&GenInfo
TimeReversal={1 if trc else 0}, neutrons={neutrons}, protons={protons}
nx={nx}, ny={ny}, nz={nz}, dx={dx}
MaxIter=2, ParameterEstimation={pe}, dt={dt}, momentum={mom}
/
&SpwfStorage
nwt={nwt}
/
&Densit
MixingScheme={mixing}, PreConW={preconw}
/
&Pairing
{pairing}
/
&Derivatives
MaxFDOrder={fd}, MaxFDLapOrder={fdl}, BStack=.false.
/
&MomentParam
{"MoreConstraints=.true." if constraints else ""}
/
{"".join("&MomentConstraint" + chr(10) + c + (", MoreConstraints=.true." if k + 1 < len(constraints) else "") + chr(10) + "/" + chr(10) for k, c in enumerate(constraints))}&InAndOutput
/
&Force
afor="SLy4"
/
&Coulomb
/
&Cranking
/
"""


def run_case(name):
    nx, ny, nz, dx, trc, lag = CASES[name]
    seed = sum(map(ord, name))
    rng = np.random.default_rng(seed)
    if trc:
        qns = [(-1, 1, 1, 1), (-1, -1, 1, 1), (-1, 1, 1, 1), (-1, -1, 1, 1),
               (1, 1, 1, 1), (1, -1, 1, 1), (1, 1, 1, 1), (1, -1, 1, 1)]
        occ = [2.0, 2.0, 1.3, 0.4, 2.0, 1.7, 0.6, 0.0]
        neutrons = protons = 4
    else:
        qns = [(-1, 1, 1, 1), (-1, 1, -1, 1), (-1, -1, 1, 1), (-1, 1, 1, 1),
               (1, 1, 1, 1), (1, 1, -1, 1), (1, -1, -1, 1), (1, 1, 1, 1)]
        occ = [1.0, 1.0, 0.8, 0.3, 1.0, 0.9, 0.5, 0.2]
        neutrons = protons = 3
    nwt = len(qns)
    psi = seeded_states(rng, nx, ny, nz, dx, qns)
    fields = synthetic_fields(rng, nx, ny, nz, dx)
    dt, mom = 0.015, 0.2

    wd = os.path.join(RUNDIR, name)
    os.makedirs(wd, exist_ok=True)
    if not os.path.exists(os.path.join(wd, "forces.param")):
        os.symlink(FORCES, os.path.join(wd, "forces.param"))
    with open(os.path.join(wd, "input.bin"), "wb") as fh:
        np.array([nwt], "<i4").tofile(fh)
        np.array(qns, "<i4").tofile(fh)
        np.array(occ, "<f8").tofile(fh)
        psi.astype("<f8").tofile(fh)
        for k in ("U", "B", "NablaB", "W", "DivW", "S", "A", "DivA"):
            fields[k].astype("<f8").tofile(fh)
    nml = namelists(nx, ny, nz, dx, trc, lag, nwt, neutrons, protons, dt, mom)
    with open(os.path.join(wd, "input.nml"), "w") as fh:
        fh.write(nml)
    with open(os.path.join(wd, "input.nml")) as fin:
        subprocess.run([DRIVER, "input.bin", "out", "force,derive,hpsi,dens,am,step"],
                       cwd=wd, stdin=fin, check=True, stdout=subprocess.DEVNULL)
    with open(os.path.join(wd, "input.nml")) as fin:
        subprocess.run([DRIVER, "input.bin", "out", "pot"],
                       cwd=wd, stdin=fin, check=True, stdout=subprocess.DEVNULL)

    def rd(fn, count=-1, offset=0):
        return np.fromfile(os.path.join(wd, fn), "<f8", count=count, offset=offset)

    sp = (4, nz, ny, nx)
    n = nx * ny * nz
    d = rd("out_derive.bin").reshape(nwt, 4, *sp)
    h = rd("out_hpsi.bin").reshape(nwt, *sp)
    dens = rd("out_dens.bin")
    k = 0

    def take(shape):
        nonlocal k
        cnt = int(np.prod(shape))
        v = dens[k:k + cnt].reshape(shape)
        k += cnt
        return v
    res = dict(rho=take((2, nz, ny, nx)), tau=take((2, nz, ny, nx)),
               nablaJ=take((2, nz, ny, nx)), derrho=take((2, 3, nz, ny, nx)),
               laprho=take((2, nz, ny, nx)))
    if not trc:
        res.update(vecj=take((2, 3, nz, ny, nx)), vecs=take((2, 3, nz, ny, nx)),
                   rotvecj=take((2, 3, nz, ny, nx)), rots=take((2, 3, nz, ny, nx)),
                   ders=take((2, 3, 3, nz, ny, nx)))
    assert k == dens.size
    # UpdateAM: per state J, JTR, JTI, J2, Sp, SpTR, SpTI (3 each), then
    # TotalAngMom, JTR, JTI, J2Total, TotalAngMom_dens (3 each), AMIsoblock
    am = rd("out_am.bin")
    res["am_states"] = am[:nwt * 21].reshape(nwt, 7, 3)
    res["am_totals"] = am[nwt * 21:nwt * 21 + 15].reshape(5, 3)
    res["am_isoblock"] = am[nwt * 21 + 15:].reshape(3, 2, 2, 2)   # [i][it][P][S]
    st = rd("out_step.bin")
    per = nwt * 4 * n + 2 * nwt
    steps = []
    for it in range(2):
        blk = st[it * per:(it + 1) * per]
        steps.append((blk[:nwt * 4 * n].reshape(nwt, *sp),
                      blk[nwt * 4 * n:].reshape(nwt, 2)))
    fb = open(os.path.join(wd, "out_force.bin"), "rb").read()
    coef = np.frombuffer(fb[:28 * 8], "<f8")
    rest = np.frombuffer(fb[28 * 8:37 * 8], "<f8")
    com1 = int(np.frombuffer(fb[37 * 8:37 * 8 + 4], "<i4")[0])
    meta = dict(nx=nx, ny=ny, nz=nz, dx=dx, trc=trc, lagrange=lag, nwt=nwt,
                dt=dt, momentum=mom, neutrons=neutrons, protons=protons,
                com1body=com1, seed=seed,
                source="reference MOCCa modules via oracle/ref_driver.f90 "
                       "(amdflang -O2)")
    potres = read_pot(wd, "out", nx, ny, nz, trc)
    np.savez_compressed(
        os.path.join(OUTDIR, name + ".npz"), meta=json.dumps(meta), **potres,
        qn=np.array(qns, np.int32), occ=np.array(occ), psi=psi, **fields,
        der=d[:, :3], lap=d[:, 3], hpsi=h, **res,
        step1_psi=steps[0][0], step1_ed=steps[0][1],
        step2_psi=steps[1][0], step2_ed=steps[1][1],
        coef=coef, consts=rest)
    print(name, "written")


def read_fields(fn, nx, ny, nz, trc):
    """Parse dump_fields of oracle/ref_driver.f90."""
    raw = open(fn, "rb").read()
    n = nx * ny * nz
    k = 0
    out = {}

    def take(name, shape):
        nonlocal k
        cnt = int(np.prod(shape))
        out[name] = np.frombuffer(raw[k:k + 8 * cnt], "<f8").reshape(shape).copy()
        k += 8 * cnt
    sh = (nz, ny, nx)
    take("rho", (2,) + sh); take("tau", (2,) + sh); take("nablaJ", (2,) + sh)
    take("derrho", (2, 3) + sh); take("laprho", (2,) + sh)
    if not trc:
        take("vecj", (2, 3) + sh); take("vecs", (2, 3) + sh)
        take("rotvecj", (2, 3) + sh); take("rots", (2, 3) + sh)
    take("vcoul", sh); take("vcoulx", sh)
    mx, my, mz = np.frombuffer(raw[k:k + 12], "<i4")
    k += 12
    take("cpot", (mz, my, mx))
    take("Upot", (2,) + sh); take("Bpot", (2,) + sh); take("NBpot", (2, 3) + sh)
    take("Wpot", (2, 3, 3) + sh); take("DWpot", (2, 3) + sh)
    if not trc:
        take("Spot", (2, 3) + sh); take("Apot", (2, 3) + sh); take("DApot", (2,) + sh)
    assert k == len(raw), (k, len(raw))
    return out


def read_pot(wd, prefix, nx, ny, nz, trc):
    f = read_fields(os.path.join(wd, prefix + "_pot.bin"), nx, ny, nz, trc)
    res = {"pot_" + key: v for key, v in f.items()}
    raw = open(os.path.join(wd, prefix + "_moments.bin"), "rb").read()
    cnt = int(np.frombuffer(raw[:4], "<i4")[0])
    rec = np.dtype([("l", "<i4"), ("m", "<i4"), ("im", "<i4"), ("v", "<f8", 2)])
    mom = np.frombuffer(raw[4:4 + cnt * rec.itemsize], rec)
    res["pot_moment_lm"] = np.stack([mom["l"], mom["m"], mom["im"]], 1).astype(np.int32)
    res["pot_moment_val"] = mom["v"].copy()
    res["pot_cutoff"] = np.fromfile(os.path.join(wd, prefix + "_cutoff.bin"), "<f8")
    e = np.fromfile(os.path.join(wd, prefix + "_energy.bin"), "<f8")
    # TotalEnergy, Kinetic(2), CoulombEnergy, CoulombExchange, CoMCorrection(2,2), terms(36)
    res["pot_energy"] = e
    return res


SCF_CASES = {
    # name: (n, dx, trc, lag, neutrons, protons, nwt_per_isospin, niter,
    #        ParameterEstimation)
    "scf_trc_fd": (10, 0.8, True, False, 8, 8, 6, 12, 0, dict(v1=True)),
    "scf_trc_lag": (10, 0.8, True, True, 8, 8, 6, 12, 0),
    "scf_todd_fd": (8, 0.9, False, False, 4, 4, 6, 8, 0),
    # IterativeEstimation before every step; maxspwf starts from seed.bin
    "scfest_trc_fd": (10, 0.8, True, False, 8, 8, 6, 10, 1),
    "scfest_todd_lag": (8, 0.9, False, True, 4, 4, 6, 8, 1),
    # MixingScheme 2: potentials preconditioned by (1 - Lap)^-1, no density
    # mixing
    "scfpc_trc_fd": (10, 0.8, True, False, 8, 8, 6, 10, 0, dict(mixing=2)),
    "scfpc_todd_lag": (8, 0.9, False, True, 4, 4, 6, 8, 0, dict(mixing=2)),
    # multipole constraints: alternating (ConstraintType 2) on Q20 of the
    # sum of both isospins, switched off after iteration 6; then neutron and
    # proton Q20 separately (alternating) with an augmented-Lagrangian
    # (ConstraintType 1) Q22
    "scfcon_trc_fd": (10, 0.8, True, False, 8, 8, 6, 10, 0,
                      dict(constraints=["l=2, m=0, Constraint=30.0, iteration=6"])),
    "scfcon_todd_lag": (8, 0.9, False, True, 4, 4, 6, 8, 0,
                        dict(constraints=["l=2, m=0, Isoswitch=2, ConstraintNeutrons=9.0, "
                                          "ConstraintProtons=7.0",
                                          "l=2, m=2, ConstraintType=1, Constraint=2.0"],
                             v1=True)),
    # HFB pairing (the sample input's pairing namelist): 4 + 4 particles in
    # 6 + 6 states and their time-reversed partners
    "scfhfb_trc_fd": (10, 0.8, True, False, 6, 4, 6, 10, 0,
                      dict(pairing='Type="HFB", PairingNeutron=1250, AlphaNeutron=1.0, '
                                   'GuessKappa=.true.', start="hf", v1=True)),
    # BCS pairing (Type="BCS", PairingIter 50 by default, the sample's
    # strength): the same nucleus and start data as scfhfb_trc_fd
    "scfbcs_trc_fd": (10, 0.8, True, False, 6, 4, 6, 10, 0,
                      dict(pairing='Type="BCS", PairingNeutron=1250, AlphaNeutron=1.0',
                           start="hf", v1=True)),
    # The sample input of the reference's README (24Mg): 12 + 12 nucleons,
    # 20 + 20 states on the 10^3 octant of a 20^3 mesh at dx = 0.8 fm,
    # Lagrange derivatives, ParameterEstimation 1, MixingScheme 2, HFB with
    # 1250 MeV fm^3 and GuessKappa, Q20 = 50 fm^2 constrained for the first 10
    # iterations, SLy4; 100 iterations from fresh states (energies 0).
    "sample_mg24": (10, 0.8, True, True, 12, 12, 20, 100, 1,
                    dict(mixing=2, constraints=["l=2, m=0, Constraint=50, iteration=10"],
                         pairing='Type="HFB", PairingNeutron=1250, AlphaNeutron=1.0, '
                                 'GuessKappa=.true.')),
    # HFB blocks of 100 states (200 x 200 HFB matrices, the block size of the
    # 40^3 / 400-state workload): 200 neutron and 40 proton states
    # (harmonic-oscillator eigenfunctions, shells 0-9) for 50 neutrons and 14
    # protons (open shells) on an 8^3 octant; the sample's pairing
    "scfhfb_large": (8, 0.8, True, False, 50, 14, (200, 40), 4, 0,
                     dict(pairing='Type="HFB", PairingNeutron=1250, AlphaNeutron=1.0, '
                                  'GuessKappa=.true.', start="hf", seeds="ho_basis",
                          keep_final=False)),
}


def run_scf(name):
    n, dx, trc, lag, neutrons, protons, per_iso, niter, pe = SCF_CASES[name][:9]
    opts = SCF_CASES[name][9] if len(SCF_CASES[name]) > 9 else {}
    mixing, preconw = opts.get("mixing", 0), opts.get("preconw", 0.0)
    constraints = opts.get("constraints", ())
    pairing = opts.get("pairing", 'Type="HF"')
    e0 = f0 = None
    seed = sum(map(ord, name))
    rng = np.random.default_rng(seed)
    counts = per_iso if isinstance(per_iso, tuple) else (per_iso, per_iso)
    pattern = [(1, 1), (-1, 1)] * (max(counts) // 2) if trc else \
        [(1, 1), (1, -1), (-1, 1), (-1, -1), (1, 1), (1, -1)]
    qns = [(iso, P, S, 1) for iso, cnt in ((-1, counts[0]), (1, counts[1]))
           for (P, S) in pattern[:cnt]]
    nwt = len(qns)
    occ = [0.0] * nwt
    if opts.get("seeds") == "ho_basis":
        psi = ho_basis_states(rng, n, dx, qns)
    else:
        psi = seeded_states(rng, n, n, n, dx, qns)
    # Orthonormalise the seeds per (isospin, parity, signature) block in the
    # octant inner product (dv = 8 dx^3), as a start-up would.
    dv8 = 8 * dx**3
    flat = psi.reshape(nwt, -1)
    for key in sorted(set(q[:3] for q in qns)):
        idx = [i for i, q in enumerate(qns) if q[:3] == key]
        for a, i in enumerate(idx):
            for j in idx[:a]:
                flat[i] -= (flat[i] @ flat[j]) * dv8 * flat[j]
            flat[i] /= np.sqrt((flat[i] @ flat[i]) * dv8)
    psi = flat.reshape(psi.shape)
    dt, mom = 0.010, 0.3
    wd = os.path.join(RUNDIR, name)
    os.makedirs(wd, exist_ok=True)
    if not os.path.exists(os.path.join(wd, "forces.param")):
        os.symlink(FORCES, os.path.join(wd, "forces.param"))
    for fn in ("energies.bin", "fermi.bin"):
        if os.path.exists(os.path.join(wd, fn)):
            os.remove(os.path.join(wd, fn))
    extra = {}
    with open(os.path.join(wd, "input.bin"), "wb") as fh:
        np.array([nwt], "<i4").tofile(fh)
        np.array(qns, "<i4").tofile(fh)
        np.array(occ, "<f8").tofile(fh)
        psi.astype("<f8").tofile(fh)
    if pe:
        # stands in for random_number in IterativeEstimation's first call
        seed_grid = rng.uniform(0.0, 1.0, size=(4, n, n, n))
        seed_grid.astype("<f8").tofile(os.path.join(wd, "seed.bin"))
        extra["est_seed"] = seed_grid
    if opts.get("start") == "hf":
        # Start data as a wavefunction file holds them: the single-particle
        # energies <h> of the start states (the energies of a one-step HF run)
        # and Fermi energies halfway between the last filled and the first
        # empty level of the HF filling.
        nml = namelists(n, n, n, dx, trc, lag, nwt, neutrons, protons, dt, mom)
        with open(os.path.join(wd, "input.nml"), "w") as fh:
            fh.write(nml)
        with open(os.path.join(wd, "input.nml")) as fin:
            subprocess.run([DRIVER, "input.bin", "hf1", "scf", "1"], cwd=wd, stdin=fin,
                           check=True, stdout=subprocess.DEVNULL)
        st1 = np.fromfile(os.path.join(wd, "hf1_scfstates.bin"), "<f8")
        e0 = st1[nwt * 4 * n ** 3:].reshape(nwt, 2)[:, 0].copy()
        f0 = []
        for iso, npart in ((-1, neutrons), (1, protons)):
            es = np.sort([e for e, q in zip(e0, qns) if q[0] == iso])
            k = int(npart // 2 if trc else npart)
            f0.append(0.5 * (es[k - 1] + es[k]))
        f0 = np.array(f0)
        e0.astype("<f8").tofile(os.path.join(wd, "energies.bin"))
        f0.astype("<f8").tofile(os.path.join(wd, "fermi.bin"))
        extra["start_energies"] = e0
        extra["start_fermi"] = f0
    nml = namelists(n, n, n, dx, trc, lag, nwt, neutrons, protons, dt, mom, pe=pe,
                    mixing=mixing, preconw=preconw, constraints=constraints, pairing=pairing)
    with open(os.path.join(wd, "input.nml"), "w") as fh:
        fh.write(nml)
    tasks = "force,scf,conv,final" + (",wv1" if opts.get("v1") else "")
    with open(os.path.join(wd, "input.nml")) as fin:
        subprocess.run([DRIVER, "input.bin", "out", tasks, str(niter)],
                       cwd=wd, stdin=fin, check=True, stdout=subprocess.DEVNULL)
    if opts.get("v1"):
        # The wavefunction file the reference writes at the end of the run
        # (WriteMOCCa_v1), and a restart from it: ReadMOCCa_v1, the start-up
        # with the file's densities (Recalc false) and 3 iterations, with
        # the same namelists except GuessKappa (a restart keeps the file's
        # kappa)
        raw = open(os.path.join(wd, "out.v1"), "rb").read()
        extra["v1_file"] = np.frombuffer(raw, np.uint8)
        wd2 = wd + "_rv1"
        os.makedirs(wd2, exist_ok=True)
        for fn in ("energies.bin", "fermi.bin"):
            if os.path.exists(os.path.join(wd2, fn)):
                os.remove(os.path.join(wd2, fn))
        if not os.path.exists(os.path.join(wd2, "forces.param")):
            os.symlink(FORCES, os.path.join(wd2, "forces.param"))
        shutil.copy(os.path.join(wd, "input.bin"), os.path.join(wd2, "input.bin"))
        open(os.path.join(wd2, "restart.v1"), "wb").write(raw)
        with open(os.path.join(wd2, "input.nml"), "w") as fh:
            fh.write(nml.replace("GuessKappa=.true.", "GuessKappa=.false."))
        with open(os.path.join(wd2, "input.nml")) as fin:
            subprocess.run([DRIVER, "input.bin", "rs", "force,scf,rv1", "3"], cwd=wd2,
                           stdin=fin, check=True, stdout=subprocess.DEVNULL)
        extra["v1_restart_traj"] = np.fromfile(os.path.join(wd2, "rs_scf.bin"),
                                               "<f8").reshape(4, 41)
        extra["v1_restart_pairing"] = np.fromfile(os.path.join(wd2, "rs_pairing.bin"),
                                                  "<f8").reshape(4, 4)
    if pe:
        extra["params"] = np.fromfile(os.path.join(wd, "out_params.bin"),
                                      "<f8").reshape(niter, 2)
    traj = np.fromfile(os.path.join(wd, "out_scf.bin"), "<f8").reshape(niter + 1, 41)
    # ConvergenceCheck's tests after every CompEnergy (start-up first) and
    # FinalIteration's Lagrange energy row, then the finite-difference rho
    extra["conv"] = np.fromfile(os.path.join(wd, "out_conv.bin"), "<i4").reshape(niter + 1, 13)
    fraw = np.fromfile(os.path.join(wd, "out_final.bin"), "<f8")
    extra["final_row"] = fraw[:41]
    extra["final_iteration_rho"] = fraw[41:].reshape(2, n ** 3)
    st = np.fromfile(os.path.join(wd, "out_scfstates.bin"), "<f8")
    N = n ** 3
    fin_psi = st[:nwt * 4 * N].reshape(nwt, 4, n, n, n)
    fin_eo = st[nwt * 4 * N:].reshape(nwt, 2)
    f = read_fields(os.path.join(wd, "out_scfpot.bin"), n, n, n, trc)
    fb = open(os.path.join(wd, "out_force.bin"), "rb").read()
    coef = np.frombuffer(fb[:28 * 8], "<f8")
    rest = np.frombuffer(fb[28 * 8:37 * 8], "<f8")
    com1 = int(np.frombuffer(fb[37 * 8:37 * 8 + 4], "<i4")[0])
    ptype, cuttype, hfbiter = (int(x) for x in np.frombuffer(fb[37 * 8 + 4:37 * 8 + 16], "<i4"))
    pp = np.frombuffer(fb[37 * 8 + 16:], "<f8")
    if ptype == 1:
        raw = np.fromfile(os.path.join(wd, "out_bcsstart.bin"), "<f8")
        extra["bcs_start"] = raw[:4 * nwt].reshape(4, nwt)   # Delta, PCutoffs, Eqp, Occ
        extra["bcs_start_fermi"] = raw[4 * nwt:]
    if ptype:
        extra["pairing"] = np.fromfile(os.path.join(wd, "out_pairing.bin"),
                                       "<f8").reshape(niter + 1, 4)
    if ptype == 2:
        raw = np.fromfile(os.path.join(wd, "out_hfbstart.bin"), "<f8")
        S = 2 * nwt if trc else nwt
        k = 0
        for key, cnt in (("delta", S * S * 4), ("kappa", S * S * 4), ("rho", S * S * 4),
                         ("dfac", 2 * n ** 3), ("pcut", nwt)):
            v = raw[k:k + cnt]
            k += cnt
            extra["start_" + key] = (v.reshape(2, 2, S, S) if cnt == S * S * 4 else
                                     v.reshape(2, n ** 3) if key == "dfac" else v)
        assert k == raw.size
    meta = dict(nx=n, ny=n, nz=n, dx=dx, trc=trc, lagrange=lag, nwt=nwt, dt=dt,
                momentum=mom, neutrons=neutrons, protons=protons, niter=niter,
                com1body=com1, seed=seed,
                mixing=("scheme 0, DampingParam 0.75" if mixing == 0 else
                        f"scheme {mixing}, PreConW {preconw}"),
                mixing_scheme=mixing, preconw=preconw, parameter_estimation=pe,
                constraints=list(constraints), pairing_type=ptype, cut_type=cuttype,
                hfb_iter=hfbiter, pairing_iter=int(pp[12]), fermi_prec=float(pp[13]),
                pairing_strength=list(pp[0:2]), pairing_alpha=list(pp[2:4]),
                pairing_ddexp=list(pp[4:6]), pairing_cut=list(pp[6:8]),
                pairing_mu=list(pp[8:10]), rhosat=float(pp[10]), hfb_mix=float(pp[11]),
                source="reference MOCCa modules via oracle/ref_driver.f90 (amdflang -O2)")
    if not opts.get("keep_final", True):
        # (the large case keeps its inputs and energies only)
        fin_psi = np.zeros(0)
        extra = {k: v for k, v in extra.items() if not k.startswith("start_") or
                 k in ("start_energies", "start_fermi")}
        f = {k: v for k, v in f.items() if k in ("rho", "vcoul")}
    np.savez_compressed(os.path.join(OUTDIR, name + ".npz"), meta=json.dumps(meta),
                        qn=np.array(qns, np.int32), psi=psi, traj=traj,
                        final_psi=fin_psi, final_eo=fin_eo, coef=coef, consts=rest,
                        **extra, **{"final_" + k: v for k, v in f.items()})
    print(name, "written", traj[:, 0])


if __name__ == "__main__":
    for c in (sys.argv[1:] or list(CASES) + list(SCF_CASES)):
        if c in CASES:
            run_case(c)
        else:
            run_scf(c)
