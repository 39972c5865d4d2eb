"""Time the reference build (oracle/_ref/mocca_ref_driver, the reference's own
Fortran modules compiled by oracle/ref.mk) on the 40^3 / 400-state
self-consistent workload with HFB pairing, in this container.

The driver runs the Evolve loop (Main.f90:255-328) from the same start twice,
with 1 and with 1 + K iterations; the per-iteration time is the difference of
the two wall-clock times over K, so start-up and file I/O cancel.  The
reference is serial (MKL only inside off-path LAPACK calls): one core, with
MKL and OpenMP pinned to one thread.

    python oracle/time_reference.py [K [neutrons protons]] > profiles/<summary>.md

Needs the reference sources (to have built the driver) and forces.param;
the GPU side never runs this.
"""
import os
import platform
import resource
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests", "golden"))

import make_golden as mg                      # noqa: E402
from mocca_amd import functional, workloads   # noqa: E402

PAIRING = 'Type="HFB", PairingNeutron=1250, AlphaNeutron=1.0, GuessKappa=.true.'


def main():
    K = int(sys.argv[1]) if len(sys.argv) > 1 else 2
    nz = (int(sys.argv[2]), int(sys.argv[3])) if len(sys.argv) > 3 else (146, 94)
    w = workloads.full_iteration(neutrons=nz[0], protons=nz[1], seeds="eigen")
    n, dx, M = w["nx"], w["dx"], w["M"]
    f = w["functional"]
    qns = [tuple(int(x) for x in q) for q in w["qn"]]
    # the seeds orthonormalised per (isospin, parity, signature) block in the
    # octant inner product, as the start-up's Gram-Schmidt leaves them
    flat = np.ascontiguousarray(w["psi"]).reshape(M, -1).copy()
    for key in sorted(set(q[:3] for q in qns)):
        idx = [i for i, q in enumerate(qns) if q[:3] == key]
        for a, i in enumerate(idx):
            for j in idx[:a]:
                flat[i] -= (flat[i] @ flat[j]) * w["dv"] * flat[j]
            flat[i] /= np.sqrt((flat[i] @ flat[i]) * w["dv"])
    wd = os.path.join(mg.RUNDIR, "time_full_iteration")
    os.makedirs(wd, exist_ok=True)
    if not os.path.exists(os.path.join(wd, "forces.param")):
        os.symlink(mg.FORCES, os.path.join(wd, "forces.param"))
    with open(os.path.join(wd, "input.bin"), "wb") as fh:
        np.array([M], "<i4").tofile(fh)
        np.array(qns, "<i4").tofile(fh)
        np.zeros(M).astype("<f8").tofile(fh)
        flat.astype("<f8").tofile(fh)
    dt = w["dt_over_hbar"] * functional.HBAR
    env = dict(os.environ, OMP_NUM_THREADS="1", MKL_NUM_THREADS="1")

    def big_stack():
        # the reference's automatic arrays at this size exceed a default stack
        resource.setrlimit(resource.RLIMIT_STACK,
                           (resource.RLIM_INFINITY, resource.RLIM_INFINITY))

    # start data as a wavefunction file gives them (the benchmark's start):
    # the energies <h> of one HF iteration of the reference from the same
    # states, Fermi energies halfway between the last filled and the first
    # empty level
    for fn in ("energies.bin", "fermi.bin"):
        if os.path.exists(os.path.join(wd, fn)):
            os.remove(os.path.join(wd, fn))
    nml = mg.namelists(n, n, n, dx, True, False, M, int(f["neutrons"]), int(f["protons"]), dt,
                       w["momentum"])
    with open(os.path.join(wd, "input.nml"), "w") as fh:
        fh.write(nml)
    with open(os.path.join(wd, "input.nml")) as fin:
        subprocess.run([mg.DRIVER, "input.bin", "hf1", "scf", "1"], cwd=wd, stdin=fin,
                       check=True, stdout=subprocess.DEVNULL, env=env, preexec_fn=big_stack)
    st1 = np.fromfile(os.path.join(wd, "hf1_scfstates.bin"), "<f8")
    e0 = st1[M * 4 * n ** 3:].reshape(M, 2)[:, 0].copy()
    f0 = []
    for iso, npart in ((-1, f["neutrons"]), (1, f["protons"])):
        es = np.sort([e for e, q in zip(e0, qns) if q[0] == iso])
        k = int(npart // 2)
        f0.append(0.5 * (es[k - 1] + es[k]))
    e0.astype("<f8").tofile(os.path.join(wd, "energies.bin"))
    np.array(f0).astype("<f8").tofile(os.path.join(wd, "fermi.bin"))
    nml = mg.namelists(n, n, n, dx, True, False, M, int(f["neutrons"]), int(f["protons"]), dt,
                       w["momentum"], pairing=PAIRING)
    with open(os.path.join(wd, "input.nml"), "w") as fh:
        fh.write(nml)
    times = {}
    for niter in (1, 1 + K):
        with open(os.path.join(wd, "input.nml")) as fin:
            t0 = time.perf_counter()
            subprocess.run([mg.DRIVER, "input.bin", "out", "force,scf", str(niter)], cwd=wd,
                           stdin=fin, check=True, stdout=subprocess.DEVNULL, env=env,
                           preexec_fn=big_stack)
            times[niter] = time.perf_counter() - t0
        print(f"<!-- {niter} iteration(s): {times[niter]:.2f} s -->", file=sys.stderr, flush=True)
    traj = np.fromfile(os.path.join(wd, "out_scf.bin"), "<f8").reshape(2 + K, 41)
    per_iter = (times[1 + K] - times[1]) / K
    cpu = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                cpu = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    print("# Reference build, 40^3 mesh / 400 states with HFB pairing, one core\n")
    print("Command: `python oracle/time_reference.py %d %d %d` (build container; the "
          "reference\nsources are not on the GPU host)\n" % (K, nz[0], nz[1]))
    print("| item | value |\n|---|---|")
    print(f"| binary | oracle/_ref/mocca_ref_driver (amdflang, oracle/ref.mk flags) |")
    print(f"| host | {cpu}, {os.cpu_count()} logical CPUs, {platform.machine()} |")
    print("| threads | 1 (OMP_NUM_THREADS=MKL_NUM_THREADS=1; the Evolve loop is serial) |")
    print("| workload | 40x40x40 octant mesh (EV8), dx=0.8 fm, 400 states (200 n + 200 p), "
          f"harmonic-oscillator eigenfunction seeds with the energies of one HF iteration as start "
          f"data (workloads.full_iteration, seeds='eigen'), SLy4, N={nz[0]} Z={nz[1]}, HFB 1250 "
          "MeV fm^3 (GuessKappa), reference Coulomb default, fixed dt/momentum, MixingScheme 0 |")
    print(f"| wall time, 1 iteration (start-up included) | {times[1]:.2f} s |")
    print(f"| wall time, {1 + K} iterations | {times[1 + K]:.2f} s |")
    print(f"| per iteration | {per_iter:.3f} s ({1.0 / per_iter:.4f} iters/s) |")
    print(f"| total energy after {1 + K} iterations | {traj[-1, 0]:.6f} MeV |")


if __name__ == "__main__":
    main()
