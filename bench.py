"""Benchmark: imaginary-time iterations per second on one MI355X.

One step = one imaginary-time iteration of the hot path over all states of
the 32^3 / 200-state hPsi workload (EV8 octant, finite differences,
time-odd S field):
    GradDesc  -- hPsi of every state, heavy-ball step, energy-ordered
                 Gram-Schmidt (Cholesky QR), heavy-ball memory update
    densities -- rho, tau, nabla.J, j, s and their derivative fields
with the states resident in HBM.  The FP64 rate is `value`; the mixed
(FP32 stencils/hPsi/densities, FP64 Gram-Schmidt) rate and its energy error
against FP64 are reported beside it.

The self-consistent 40^3 / 400-state workload (SLy4 + Coulomb, HF filling,
density mixing; one step = the whole Evolve iteration on the device) is
timed after it in FP64 and mixed precision from the same start, with the
total-energy difference after the same number of iterations
(``full_iteration`` in the JSON line).  The 48^3 / 600-state Lagrange-mesh
workload (derivatives on the matrix cores) follows (``lagrange_48``), then
64 independent constrained solves through mocca_batch_run (``batch_scan``).

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
                    [--full-steps K3] [--skip-full]
                    [--lagrange-steps K5] [--skip-lagrange]
                    [--batch-steps K4] [--skip-batch]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = ("imag-time iters/sec (hPsi+dens+GS) FP64 vs mixed at ≤1keV err; "
          "GB/s vs HBM peak")
HBM_PEAK_GBS = 8000.0


def make_solve(w, precision):
    import mocca_amd
    s = mocca_amd.Solve(w["nx"], w["ny"], w["nz"], w["dx"], w["dv"], w["M"],
                        time_reversal=w["trc"], lagrange=w["lagrange"],
                        precision=precision)
    s.set_states_meta(w["qn"], w["occ"])
    s.upload_states(w["psi"])
    F = w["fields"]
    s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], w["kin"],
                     S=F.get("S"), A=F.get("A"), DivA=F.get("DivA"))
    return s


def step(s, w):
    s.grad_desc(w["dt_over_hbar"], w["momentum"])
    s.compute_density(fetch=False)


def timed_run(w, precision, steps, warmup, profile):
    s = make_solve(w, precision)
    for _ in range(warmup):
        step(s, w)
    s.synchronize()
    if profile:
        s.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        step(s, w)
    s.synchronize()
    t1 = time.perf_counter()
    prof = {}
    if profile:
        for k in ("hpsi_step", "reduce_ed", "gram", "gram_reduce", "cholesky",
                  "gs_apply", "density", "density_reduce", "field_jobs"):
            ms, cnt = s.profile_get(k)
            prof[k] = (ms, cnt)
        s.profile_enable(False)
    e, _ = s.energies()
    s.close()
    return (t1 - t0) / steps, prof, e


def full_iteration_run(w, precision, steps, warmup, profile):
    """The self-consistent loop on the device: start-up, warmup iterations,
    then `steps` timed iterations (mocca_iterate)."""
    import mocca_amd
    s = mocca_amd.Solve(w["nx"], w["ny"], w["nz"], w["dx"], w["dv"], w["M"],
                        time_reversal=w["trc"], lagrange=w["lagrange"], precision=precision)
    s.set_states_meta(w["qn"], w["occ"])
    s.upload_states(w["psi"])
    s.scf_setup(**w["functional"])
    s.gram_schmidt()
    s.scf_start()
    s.iterate(warmup, w["dt_over_hbar"], w["momentum"], fetch=False)
    s.synchronize()
    if profile:
        s.profile_enable(True)
    t0 = time.perf_counter()
    rows = s.iterate(steps, w["dt_over_hbar"], w["momentum"])
    t1 = time.perf_counter()
    prof = {}
    if profile:
        for k in ("lag_derivs", "hpsi_step", "reduce_ed", "gram", "gram_reduce", "cholesky",
                  "gs_apply", "hf_fill", "density", "density_reduce", "field_jobs", "mix",
                  "moments",
                  "coulomb_bc", "coulomb_cg", "coulomb_direct", "potentials", "pot_jobs",
                  "pot_convert",
                  "energy", "energy_kinetic", "kinetic"):
            ms, cnt = s.profile_get(k)
            if cnt:
                prof[k] = ms / steps
        s.profile_enable(False)
    _, _, _, cg_iters = s.coulomb()
    s.close()
    return (t1 - t0) / steps, prof, rows, cg_iters


def batch_scan_run(steps, warmup):
    """64 independent constrained solves (uneven meshes 16^3..32^3 and
    20..160 states, alternating Q20 constraints) through mocca_batch_run, in
    FP64 and mixed precision from the same start; the same solves iterated
    one after another give the sequential rate."""
    import mocca_amd
    from mocca_amd import FP64, MIXED, workloads
    ws = workloads.batch_scan()

    def start(w, precision):
        s = mocca_amd.Solve(w["nx"], w["ny"], w["nz"], w["dx"], w["dv"], w["M"],
                            time_reversal=w["trc"], lagrange=w["lagrange"],
                            precision=precision)
        s.set_states_meta(w["qn"], w["occ"])
        s.upload_states(w["psi"])
        s.scf_setup(**w["functional"])
        s.set_constraints(w["constraints"])
        s.gram_schmidt()
        s.scf_start()
        return s

    dts = [w["dt_over_hbar"] for w in ws]
    mus = [w["momentum"] for w in ws]
    res, last = {}, {}
    for name, prec in (("fp64", FP64), ("mixed", MIXED)):
        solves = [start(w, prec) for w in ws]
        # the warmup call also captures each solve's iterations as graphs
        mocca_amd.batch_run(solves, max(warmup, 4), dts, mus, fetch=False)
        for s in solves:
            s.synchronize()
        t0 = time.perf_counter()
        rows = mocca_amd.batch_run(solves, steps, dts, mus)
        t1 = time.perf_counter()
        last[name] = np.array([r[-1, 0] for r in rows])
        t2 = time.perf_counter()
        for s, w in zip(solves, ws):
            s.iterate(steps, w["dt_over_hbar"], w["momentum"])
        t3 = time.perf_counter()
        for s in solves:
            s.close()
        res[name] = {"value": len(ws) * steps / (t1 - t0), "unit": "solve-iters/s",
                     "ms_per_batch_step": (t1 - t0) / steps * 1e3,
                     "sequential_value": len(ws) * steps / (t3 - t2),
                     "speedup_vs_sequential": (t3 - t2) / (t1 - t0)}
    return {
        "workload": "64 solves of 12 n + 12 p, cubic meshes n ~ U{16..32} (dx=0.8 fm), "
                    "nwt ~ U{20..160}, HO seeds, SLy4 + Coulomb, HF, alternating Q20 "
                    "constraint over -300..600 fm^2; step = one Evolve iteration of "
                    "every solve, replayed as a captured HIP graph on the solve's own "
                    "stream (the first iteration of each call runs eagerly)",
        "solves": len(ws), "steps": steps, "warmup": warmup,
        "fp64": res["fp64"], "mixed": res["mixed"],
        "mixed_speedup_vs_fp64": res["mixed"]["value"] / res["fp64"]["value"],
        "max_abs_total_energy_err_keV": float(np.max(np.abs(last["mixed"] - last["fp64"])) * 1e3),
    }


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the counter table of the
    committed rocprofv3 summary (2 x FETCH_SIZE + WRITE_SIZE, in KiB per
    dispatch; the factor 2 is the FETCH_SIZE calibration against a plain
    copy), or None."""
    path = os.path.join(ROOT, "profiles", "hpsi_step_gram_density_fd32_200states.md")
    if not os.path.exists(path):
        return None
    in_counters = False
    for line in open(path):
        if line.startswith("## "):
            in_counters = "traffic counters" in line
        if not in_counters:
            continue
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 5 and cells[0].strip("`") == kernel:
            try:
                return (2 * float(cells[1]) + float(cells[3])) * 1024
            except ValueError:
                return None
    return None


def cpu_baseline(w, min_seconds=10.0, max_seconds=30.0):
    """The C restatement (oracle, one core) on one Gram-Schmidt block of the
    same workload: derivatives, hPsi + step + Gram-Schmidt, densities of its
    25 states.  Blocks are independent and every cost is per block, so the
    full-iteration time is (blocks) x (block time)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    qn = w["qn"]
    key = tuple(qn[0, :3])
    idx = [i for i in range(w["M"]) if tuple(qn[i, :3]) == key]
    nblocks = len(set(map(tuple, qn[:, :3])))
    n, N = w["nx"], w["nx"] ** 3
    mesh = ol.Mesh(n, n, n, w["dx"], w["lagrange"])
    psi = np.ascontiguousarray(w["psi"][idx])
    sub_qn = np.ascontiguousarray(qn[idx])
    occ = np.ascontiguousarray(w["occ"][idx])
    F = {k: np.ascontiguousarray(v).reshape(-1) for k, v in w["fields"].items()}
    upd = np.zeros_like(psi)
    reps, t_total = 0, 0.0
    while t_total < min_seconds and (reps == 0 or t_total * (reps + 1) / reps < max_seconds):
        t0 = time.perf_counter()
        der = np.zeros((len(idx), 3, 4, N))
        lap = np.zeros((len(idx), 4, N))
        for j, q in enumerate(sub_qn):
            der[j], lap[j] = mesh.derive_spinor(psi[j], q[1], q[2], q[3])
        psi2, upd, _, _ = ol.graddesc(psi, der, lap, upd, sub_qn, w["trc"], w["kin"], F,
                                      w["dt_over_hbar"], w["momentum"], w["dv"])
        for j, q in enumerate(sub_qn):
            der[j], lap[j] = mesh.derive_spinor(psi2[j], q[1], q[2], q[3])
        ol.density(psi2, der, sub_qn, occ, w["trc"], N)
        t_total += time.perf_counter() - t0
        reps += 1
    per_iter = t_total / reps * nblocks
    return dict(value=1.0 / per_iter, unit="iters/s", cores=1, kind="port",
                sample=f"{reps} x one Gram-Schmidt block ({len(idx)} of {w['M']} states, "
                       f"{nblocks} blocks) of the same workload: derivatives + hPsi/step/"
                       f"Gram-Schmidt + densities, scaled x{nblocks}; oracle/liboracle.so, "
                       f"{t_total:.1f} s of CPU time")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--full-steps", type=int, default=200)
    ap.add_argument("--skip-full", action="store_true")
    ap.add_argument("--lagrange-steps", type=int, default=20)
    ap.add_argument("--skip-lagrange", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 256 MiB device copy three times (the "
                         "reference kernel for the FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        a = torch.ones(256 * 1024 * 1024 // 8, dtype=torch.float64, device="cuda")
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": a.numel() * 8, "copies": 3}), flush=True)
        return
    if args.gpus != 1:
        print("bench.py measures one GPU only: the hot path runs on a single MI355X "
              "(multi-GPU sharding is a design note, not built)", file=sys.stderr)
        sys.exit(2)
    from mocca_amd import FP64, MIXED, workloads
    w = workloads.hpsi_microbench()
    M, N = w["M"], w["nx"] ** 3

    t64, prof, e64 = timed_run(w, FP64, args.steps, args.warmup, True)
    t32, prof32, e32 = timed_run(w, MIXED, args.steps, args.warmup, True)

    # roofline of the dominant kernel (hpsi_step): algorithmic bytes per launch
    # = read psi + read the states before the last step (the heavy-ball
    # memory) + write psi' for all states, plus the potentials of both
    # isospins once (time-odd: U B NB(3) W DW(3) S(3) A(3) DA; the workload's
    # W is antisymmetric with a zero diagonal, so 3 of its 9 components are
    # independent: 18 fields per isospin, 11 with time reversal).
    s_bytes = 4 * N * 8
    nfields = 11 if w["trc"] else 18
    alg_bytes = 3 * M * s_bytes + 2 * nfields * N * 8
    ms, cnt = prof["hpsi_step"]
    t_kernel = ms / cnt / 1e3
    achieved = alg_bytes / t_kernel / 1e9
    tot_ms = sum(v[0] for v in prof.values()) / args.steps

    # seven state passes per iteration: hPsi/step 3, Gram 1, apply 2, densities 1
    iter_bytes = 7 * M * s_bytes
    out = {
        "metric": METRIC,
        "value": 1.0 / t64,
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t64 * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "hpsi_microbench: 32x32x32 octant mesh (EV8), dx=0.8 fm, "
                        "200 states, 7/9-point FD, Woods-Saxon U/B/W + time-odd S; "
                        "step = GradDesc (hPsi + heavy-ball step + Gram-Schmidt) + densities",
            "mesh": [w["nx"], w["ny"], w["nz"]], "states": M, "gs_blocks": 8,
        },
        "roofline": {
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(
                "mocca::k_hpsi_fd_dma<double, %s, true>" % ("true" if w["trc"] else "false")),
            "kernel": "hpsi_step", "kernel_ms": t_kernel * 1e3,
            "alg_bytes_per_launch": alg_bytes,
        },
        "iteration_gbs": iter_bytes / t64 / 1e9,
        "kernel_ms_per_step": {k: v[0] / max(v[1], 1) * (v[1] / args.steps)
                               for k, v in prof.items()},
        "kernel_sum_ms_per_step": tot_ms,
        "mixed": {
            "value": 1.0 / t32, "ms_per_step": t32 * 1e3, "speedup_vs_fp64": t64 / t32,
            "max_abs_sp_energy_err_keV": float(np.max(np.abs(e32 - e64)) * 1e3),
            "note": "single-particle energies <psi|h|psi> after warmup+steps iterations "
                    "from the same start, vs the FP64 run (frozen mean fields)",
        },
    }
    if not args.skip_full:
        wf = workloads.full_iteration()
        k3 = args.full_steps
        tf64, pf64, r64, cg64 = full_iteration_run(wf, FP64, k3, args.warmup, True)
        tf32, pf32, r32, cg32 = full_iteration_run(wf, MIXED, k3, args.warmup, True)
        out["full_iteration"] = {
            "workload": "40x40x40 octant mesh, dx=0.8 fm, 400 states (200 n + 200 p), "
                        "HO seeds, SLy4 + Coulomb (direct solve), HF filling N=126 Z=82 (no pairing), "
                        "MixingScheme 0; step = the whole Evolve iteration on the device "
                        "(GradDesc + fill + densities + mixing + moments + Coulomb + "
                        "potentials + energy)",
            "steps": k3, "warmup": args.warmup,
            "fp64": {"value": 1.0 / tf64, "unit": "iters/s", "ms_per_step": tf64 * 1e3,
                     "kernel_ms_per_step": pf64, "coulomb_cg_iterations_last": cg64},
            "mixed": {"value": 1.0 / tf32, "unit": "iters/s", "ms_per_step": tf32 * 1e3,
                      "speedup_vs_fp64": tf64 / tf32, "kernel_ms_per_step": pf32,
                      "coulomb_cg_iterations_last": cg32},
            "energy_fp64_MeV": float(r64[-1, 0]), "energy_mixed_MeV": float(r32[-1, 0]),
            "total_energy_err_keV": float(abs(r32[-1, 0] - r64[-1, 0]) * 1e3),
            "max_total_energy_err_keV_along_trajectory":
                float(np.max(np.abs(r32[:, 0] - r64[:, 0])) * 1e3),
        }
    if not args.skip_lagrange:
        wl = workloads.lagrange_48()
        kl = args.lagrange_steps
        tl64, pl64, rl64, _ = full_iteration_run(wl, FP64, kl, args.warmup, True)
        tl32, pl32, rl32, _ = full_iteration_run(wl, MIXED, kl, args.warmup, True)
        out["lagrange_48"] = {
            "workload": "48x48x48 Lagrange octant mesh (48-point derivative matrices), "
                        "dx=0.7 fm, 600 states (300 n + 300 p), HO seeds, SLy4 + Coulomb, "
                        "HF filling N=Z=12; step = the whole Evolve iteration on the device",
            "steps": kl, "warmup": args.warmup,
            "fp64": {"value": 1.0 / tl64, "unit": "iters/s", "ms_per_step": tl64 * 1e3,
                     "kernel_ms_per_step": pl64},
            "mixed": {"value": 1.0 / tl32, "unit": "iters/s", "ms_per_step": tl32 * 1e3,
                      "speedup_vs_fp64": tl64 / tl32, "kernel_ms_per_step": pl32},
            "total_energy_err_keV": float(abs(rl32[-1, 0] - rl64[-1, 0]) * 1e3),
        }
    if not args.skip_batch:
        out["batch_scan"] = batch_scan_run(args.batch_steps, args.warmup)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(w)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
