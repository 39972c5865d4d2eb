"""Benchmark: imaginary-time iterations per second on one MI355X.

Headline (`value`): the 40^3 / 400-state self-consistent iteration with the
HFB pairing update (40x40x40 octant mesh, dx = 0.8 fm, 200 neutron + 200
proton harmonic-oscillator seeds, SLy4 + Coulomb, N = 146, Z = 94, HFB with
a 1250 MeV fm^3 contact pairing and GuessKappa), in FP64.  One step is one
whole Evolve iteration on the device:
    GradDesc  -- hPsi of every state, heavy-ball step, energy-ordered
                 Gram-Schmidt (Cholesky QR), heavy-ball memory update
    pairing   -- pairing field, gaps, the Broyden Fermi solve (each
                 evaluation's rho, kappa and particle number from the sign
                 function of the 200 x 200 HFB matrices), canonical basis
                 and states, rho/kappa mixing
    densities, mixing, moments, Coulomb, potentials, energy.
The mixed run (FP32 stencils/hPsi/densities, FP64 Gram-Schmidt, energy,
Coulomb and pairing) starts from the same states; its total-energy
difference from FP64 after the same iterations is reported beside it.

Further legs: the same workload with HF filling (`hf_no_pairing`), the
32^3 / 200-state hPsi microbench (`hpsi_microbench`), the README sample
input (`readme_sample`, 10^3 octant / 40 states, HFB, estimation, Q20), the
48^3 / 600-state Lagrange workload (`lagrange_48`), 64 constrained solves
through mocca_batch_run (`batch_scan`) and a MOCCa v1 checkpoint and restart
of the headline solve (`checkpoint_v1`).  `cpu_baseline` is the oracle's
restatement of the same iteration on the host (1 thread and all threads);
`reference_container` is the reference build timed in the build container.

    python bench.py [--steps K] [--warmup W] [--gpus 1] [--no-cpu]
                    [--skip-hf] [--skip-micro] [--skip-sample]
                    [--skip-lagrange] [--skip-batch] [--skip-checkpoint]
"""
import argparse
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
PROFILES = os.path.join(ROOT, "profiles")
HEADLINE_PROFILE = "full_iteration_hfb_40cubed_400states.md"
MICRO_PROFILE = "hpsi_step_gram_density_fd32_200states.md"
REFERENCE_TIMING = "cpu_reference_full_iteration_hfb_40cubed_400states.md"

# stages timed by the library's own HIP events (mocca_profile_get)
STAGES = ("lag_derivs", "derive_all", "hpsi_step", "hpsi", "reduce_ed", "gram", "gram_reduce",
          "cholesky", "gs_apply", "hf_fill", "hfb_mix", "hfb_field", "hfb_gaps", "hfb_fermi",
          "hfb_jacobi", "hfb_jacobi_apply", "hfb_sign", "hfb_canonical", "hfb_energy", "density",
          "density_reduce", "field_jobs", "mix", "moments", "constraints", "alternate",
          "coulomb_bc", "coulomb_cg", "coulomb_direct", "potentials", "precondition",
          "pot_jobs", "pot_convert", "energy", "energy_kinetic", "kinetic", "cutoff",
          "est_start", "est_reset", "est_hpsi", "est_step", "est_finalize")

# the main kernel of a stage and what bounds it (for the dominant stage)
STAGE_KERNELS = {
    "hfb_sign": ("mocca::k_hs_gemm<false>",
                 "launch and memory latency: 20 dependent 200 x 200 x 200 FP64 matrix "
                 "products per Broyden evaluation (about 6 us each, a few % of the FP64 "
                 "matrix-core peak)"),
    "hfb_jacobi": ("mocca::k_lj_jacobi_slots",
                   "LDS bandwidth and barrier latency of one workgroup per 100 x 100 "
                   "density block (4 blocks: 4 of 256 CUs)"),
    "hfb_gaps": ("mocca::k_hl_pair_gram<double, 7, 4, true>", "FP64 matrix cores"),
    "hfb_field": ("mocca::k_apply_direct<double, 7, 1, 16, true, true>", "FP64 matrix cores"),
    "hpsi_step": ("mocca::k_hpsi_fd_dma<double, true, true, false>",
                  "on-chip issue (VALU, LDS, one barrier per plane)"),
}

# The headline nucleus: open shells in both isospins (146 neutrons, 94
# protons), so that the pairing stays on through the run (a doubly magic
# filling such as 126 / 82 loses its pairing and the HFB step turns trivial),
# started from filled harmonic-oscillator shells (b = 2.4 fm), bound from
# the first iteration on
NUCLEUS = dict(neutrons=146, protons=94, seeds="eigen")

# HFB pairing of the headline workload (the Pairing namelist of the README
# sample: contact pairing 1250 MeV fm^3, alpha 1, GuessKappa)
PAIRING = dict(type=2, strength=(1250.0, 1250.0), alpha=(1.0, 1.0), cut=(5.0, 5.0),
               mu=(0.5, 0.5), rhosat=0.16, hfb_mix=0.1, iterations=50, fermi_prec=1e-12,
               fermi=(-10.0, -10.0), guess_kappa=True)


def profile_table(s, steps):
    out = {}
    for k in STAGES:
        ms, cnt = s.profile_get(k)
        if cnt:
            out[k] = dict(ms_per_step=ms / steps, launches_per_step=cnt / steps,
                          ms_per_launch=ms / cnt)
    return out


def scf_solve(w, precision):
    import mocca_amd
    s = mocca_amd.Solve(w["nx"], w["ny"], w["nz"], w["dx"], w["dv"], w["M"],
                        time_reversal=w["trc"], lagrange=w["lagrange"], precision=precision)
    if w.get("lagrange_split"):
        s.set_lagrange_precision(1)
    s.set_states_meta(w["qn"], w["occ"])
    s.upload_states(w["psi"])
    s.scf_setup(**w["functional"])
    if w.get("constraints"):
        s.set_constraints(w["constraints"])
    s.gram_schmidt()
    return s


def start_data(w):
    """Single-particle energies and Fermi energies of the start states, as a
    wavefunction file gives them to the reference: the energies <h> of one HF
    iteration from the same seeds (FP64), and each Fermi energy halfway
    between the last filled and the first empty level of the HF filling."""
    from mocca_amd import FP64
    s = scf_solve(w, FP64)
    s.scf_start()
    s.iterate(1, w["dt_over_hbar"], w["momentum"], fetch=False)
    e, _ = s.energies()
    s.close()
    f = w["functional"]
    fermi = []
    for iso, npart in ((-1, f["neutrons"]), (1, f["protons"])):
        es = np.sort(e[w["qn"][:, 0] == iso])
        k = int(npart // 2)          # time reversal: two nucleons per state
        fermi.append(0.5 * (es[k - 1] + es[k]))
    return e, np.array(fermi)


def full_iteration_run(w, precision, steps, warmup, pairing, keep_start=False, start=None):
    """Start-up, `warmup` iterations (the first one fetched, for the oracle's
    check), then `steps` timed iterations with the stage events on.  With
    pairing, start = (energies, Fermi energies) of start_data."""
    s = scf_solve(w, precision)
    state0 = s.download_states() if keep_start else None
    if pairing:
        e0, f0 = start
        s.set_energies(e0)
        s.set_pairing(**dict(PAIRING, fermi=tuple(f0)))
    s.scf_start()
    first = s.iterate(1, w["dt_over_hbar"], w["momentum"])
    if warmup > 1:
        s.iterate(warmup - 1, w["dt_over_hbar"], w["momentum"], fetch=False)
    s.synchronize()
    s.profile_enable(True)
    t0 = time.perf_counter()
    rows = s.iterate(steps, w["dt_over_hbar"], w["momentum"])
    t1 = time.perf_counter()
    prof = profile_table(s, steps)
    s.profile_enable(False)
    s.close()
    return dict(t=(t1 - t0) / steps, prof=prof, rows=rows, first=first, start=state0)


def hpsi_roofline(M, N, trc, stage):
    """Roofline of the hPsi/step kernel (k_hpsi_fd_dma): algorithmic bytes
    per launch = read psi + read the states before the last step (the
    heavy-ball memory) + write psi' of all states, plus the potentials of
    both isospins once (time-even: U B NB(3) W(3 independent of the
    antisymmetric 9) DW(3) = 11 fields; time-odd adds S(3) A(3) DA: 18)."""
    s_bytes = 4 * N * 8
    nfields = 11 if trc else 18
    alg = 3 * M * s_bytes + 2 * nfields * N * 8
    t = stage["ms_per_launch"] / 1e3
    achieved = alg / t / 1e9
    return alg, achieved, t


def measured_traffic(profile, kernel):
    """HBM bytes per launch from the counter table of a committed rocprofv3
    summary (FETCH_SIZE and WRITE_SIZE in KiB per dispatch, each multiplied
    by the calibration factor the summary states for a plain copy), or None."""
    path = os.path.join(PROFILES, profile)
    if not os.path.exists(path):
        return None
    in_counters, fetch_cal, write_cal = False, None, None
    for line in open(path):
        if line.startswith("Calibration:"):
            for tok in line.replace(",", " ").split():
                if tok.startswith("FETCH="):
                    fetch_cal = float(tok.split("=")[1])
                if tok.startswith("WRITE="):
                    write_cal = float(tok.split("=")[1])
        if line.startswith("## "):
            in_counters = "traffic counters" in line
        if not in_counters:
            continue
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) >= 4 and cells[0].strip("`") == kernel:
            try:
                return (float(cells[1]) * (fetch_cal or 1.0) +
                        float(cells[2]) * (write_cal or 1.0)) * 1024
            except ValueError:
                return None
    return None


def headline(args):
    from mocca_amd import FP64, MIXED, workloads
    w = workloads.full_iteration(**NUCLEUS)
    M, N = w["M"], w["nx"] ** 3
    start = start_data(w)
    r64 = full_iteration_run(w, FP64, args.steps, args.warmup, True, keep_start=True,
                             start=start)
    r32 = full_iteration_run(w, MIXED, args.steps, args.warmup, True, start=start)
    prof = r64["prof"]
    tot = sum(v["ms_per_step"] for v in prof.values())
    shares = sorted(((k, v["ms_per_step"] / tot) for k, v in prof.items()),
                    key=lambda kv: -kv[1])
    alg, achieved, t_k = hpsi_roofline(M, N, w["trc"], prof["hpsi_step"])
    kname = "mocca::k_hpsi_fd_dma<double, true, true, false>"
    jac = prof.get("hfb_jacobi", {})
    e64, e32 = r64["rows"][:, 0], r32["rows"][:, 0]
    out = dict(
        w=w, r64=r64, r32=r32, start=start,
        roofline={
            "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(HEADLINE_PROFILE, kname),
            "kernel": kname, "stage": "hpsi_step", "kernel_ms": t_k * 1e3,
            "alg_bytes_per_launch": alg,
            "traffic_source": "profiles/" + HEADLINE_PROFILE + " (rocprofv3 --pmc FETCH_SIZE, "
                              "WRITE_SIZE of the same workload, copy-calibrated)",
            "note": "the largest HBM-bound kernel of the hPsi + densities + Gram-Schmidt "
                    "path; the iteration's largest time share is the HFB pairing update "
                    "(dominant_stage), whose kernels are latency- or matrix-core-bound",
        },
        dominant_stage={
            "stage": shares[0][0], "share": shares[0][1],
            "kernel": STAGE_KERNELS.get(shares[0][0], (None, None))[0],
            "ms_per_step": prof[shares[0][0]]["ms_per_step"],
            "launches_per_step": prof[shares[0][0]]["launches_per_step"],
            "bound": STAGE_KERNELS.get(shares[0][0], (None, None))[1],
        },
        stage_shares={k: round(v, 4) for k, v in shares},
        mixed={
            "value": 1.0 / r32["t"], "ms_per_step": r32["t"] * 1e3,
            "speedup_vs_fp64": r64["t"] / r32["t"],
            "total_energy_err_keV": float(abs(e32[-1] - e64[-1]) * 1e3),
            "max_total_energy_err_keV_along_trajectory": float(np.max(np.abs(e32 - e64)) * 1e3),
            "kernel_ms_per_step": {k: v["ms_per_step"] for k, v in r32["prof"].items()},
        },
        energy_fp64_MeV=float(e64[-1]), energy_mixed_MeV=float(e32[-1]),
        hfb_jacobi_ms_per_launch=jac.get("ms_per_launch"),
    )
    return out


def checkpoint(w, start):
    """The headline solve (FP64, HFB) after the start-up and 2 iterations
    written to a MOCCa v1 file (mocca_v1_save: states, densities, HFB
    matrices of the 800-row blocks, energies) and a second solve restarted
    from it (mocca_v1_load); wall-clock seconds of each call, the file size,
    and whether the restarted states equal the saved ones bit for bit.  The
    file goes to the host's temporary directory, so the rates include the
    device-host copies and the page cache, not a disk."""
    import shutil
    import tempfile
    from mocca_amd import FP64
    e0, f0 = start
    s = scf_solve(w, FP64)
    s.set_energies(e0)
    s.set_pairing(**dict(PAIRING, fermi=tuple(f0)))
    s.scf_start()
    s.iterate(2, w["dt_over_hbar"], w["momentum"], fetch=False)
    s.synchronize()
    d = tempfile.mkdtemp()
    path = os.path.join(d, "checkpoint.v1")
    try:
        t0 = time.perf_counter()
        s.save_v1(path)
        t1 = time.perf_counter()
        size = os.path.getsize(path)
        psi = s.download_states()
        s.close()
        s2 = scf_solve(w, FP64)
        s2.set_pairing(**dict(PAIRING, fermi=tuple(f0), guess_kappa=False))
        s2.synchronize()
        t2 = time.perf_counter()
        s2.load_v1(path)
        s2.synchronize()
        t3 = time.perf_counter()
        same = bool(np.array_equal(s2.download_states(), psi))
        s2.close()
    finally:
        shutil.rmtree(d, ignore_errors=True)
    return {"file_bytes": size, "save_s": t1 - t0, "load_s": t3 - t2,
            "save_GB_per_s": size / (t1 - t0) / 1e9, "load_GB_per_s": size / (t3 - t2) / 1e9,
            "restart_states_bit_exact": same,
            "note": "mocca_v1_save / mocca_v1_load of the 40^3 / 400-state HFB solve, host "
                    "temporary directory (page cache), wall clock"}


def hf_no_pairing(steps, warmup):
    from mocca_amd import FP64, MIXED, workloads
    w = workloads.full_iteration(**NUCLEUS)
    r64 = full_iteration_run(w, FP64, steps, warmup, False)
    r32 = full_iteration_run(w, MIXED, steps, warmup, False)
    return {
        "workload": "HF, no pairing: the headline workload with HF filling of N=146, Z=94 in "
                    "place of the HFB step; SLy4 + Coulomb (direct solve); step = the whole "
                    "Evolve iteration on the device",
        "steps": steps, "warmup": warmup,
        "fp64": {"value": 1.0 / r64["t"], "unit": "iters/s", "ms_per_step": r64["t"] * 1e3,
                 "kernel_ms_per_step": {k: v["ms_per_step"] for k, v in r64["prof"].items()}},
        "mixed": {"value": 1.0 / r32["t"], "unit": "iters/s", "ms_per_step": r32["t"] * 1e3,
                  "speedup_vs_fp64": r64["t"] / r32["t"],
                  "kernel_ms_per_step": {k: v["ms_per_step"] for k, v in r32["prof"].items()}},
        "total_energy_err_keV": float(abs(r32["rows"][-1, 0] - r64["rows"][-1, 0]) * 1e3),
    }


def micro_run(w, precision, steps, warmup):
    import mocca_amd
    s = mocca_amd.Solve(w["nx"], w["ny"], w["nz"], w["dx"], w["dv"], w["M"],
                        time_reversal=w["trc"], lagrange=w["lagrange"], precision=precision)
    if w.get("lagrange_split"):
        s.set_lagrange_precision(1)
    s.set_states_meta(w["qn"], w["occ"])
    s.upload_states(w["psi"])
    F = w["fields"]
    s.set_potentials(F["U"], F["B"], F["NablaB"], F["W"], F["DivW"], w["kin"],
                     S=F.get("S"), A=F.get("A"), DivA=F.get("DivA"))

    def step():
        s.grad_desc(w["dt_over_hbar"], w["momentum"])
        s.compute_density(fetch=False)

    for _ in range(warmup):
        step()
    s.synchronize()
    s.profile_enable(True)
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
    s.synchronize()
    t1 = time.perf_counter()
    prof = profile_table(s, steps)
    s.profile_enable(False)
    e, _ = s.energies()
    s.close()
    return (t1 - t0) / steps, prof, e


def hpsi_microbench(steps, warmup):
    from mocca_amd import FP64, MIXED, workloads
    w = workloads.hpsi_microbench()
    M, N = w["M"], w["nx"] ** 3
    t64, p64, e64 = micro_run(w, FP64, steps, warmup)
    t32, p32, e32 = micro_run(w, MIXED, steps, warmup)
    alg, achieved, t_k = hpsi_roofline(M, N, w["trc"], p64["hpsi_step"])
    kname = "mocca::k_hpsi_fd_dma<double, false, true>"
    return {
        "workload": "32x32x32 octant mesh (EV8), dx=0.8 fm, 200 states, 7/9-point FD, "
                    "Woods-Saxon U/B/W + time-odd S; step = GradDesc (hPsi + heavy-ball step "
                    "+ Gram-Schmidt) + densities, mean fields frozen",
        "steps": steps, "warmup": warmup,
        "fp64": {"value": 1.0 / t64, "unit": "iters/s", "ms_per_step": t64 * 1e3,
                 "kernel_ms_per_step": {k: v["ms_per_step"] for k, v in p64.items()}},
        "mixed": {"value": 1.0 / t32, "unit": "iters/s", "ms_per_step": t32 * 1e3,
                  "speedup_vs_fp64": t64 / t32,
                  "max_abs_sp_energy_err_keV": float(np.max(np.abs(e32 - e64)) * 1e3)},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(MICRO_PROFILE, kname), "kernel": kname,
                     "kernel_ms": t_k * 1e3, "alg_bytes_per_launch": alg},
        # seven state passes per iteration: hPsi/step 3, Gram 1, apply 2, densities 1
        "iteration_gbs": 7 * M * 4 * N * 8 / t64 / 1e9,
    }


def readme_sample(warmup):
    """The reference's README sample input (24Mg): its start states from the
    reference's own run (tests/golden/sample_mg24.npz), 100 iterations with
    IterativeEstimation, MixingScheme 2, HFB and a Q20 constraint for the
    first 10; timed as one run (the estimation and the constraint switch-off
    make every iteration different)."""
    import mocca_amd
    z = np.load(os.path.join(ROOT, "tests", "golden", "sample_mg24.npz"), allow_pickle=False)
    g = {k: z[k] for k in z.files}
    m = json.loads(str(g["meta"]))
    c = g["consts"]
    res = {}
    for name, prec in (("fp64", mocca_amd.FP64), ("mixed", mocca_amd.MIXED)):
        best = None
        for rep in range(1 + warmup // 5):
            s = mocca_amd.Solve(m["nx"], m["ny"], m["nz"], m["dx"], c[6], m["nwt"],
                                time_reversal=m["trc"], lagrange=m["lagrange"], precision=prec)
            s.set_states_meta(g["qn"], np.zeros(m["nwt"]))
            s.upload_states(g["psi"])
            s.scf_setup(g["coef"], c[0:2], c[2:4], c[4], c[7], c[8], com1body=m["com1body"],
                        cexchange=2, coulomb=1, coulomb_maxit=500, max_moment=10,
                        damping=0.75, mixing_scheme=m.get("mixing_scheme", 0))
            s.set_constraints([dict(l=2, m=0, type=2, isoswitch=1, iteration=10, value=50.0)])
            s.set_pairing(type=2, strength=m["pairing_strength"], alpha=m["pairing_alpha"],
                          cut=m["pairing_cut"], mu=m["pairing_mu"], rhosat=m["rhosat"],
                          hfb_mix=m["hfb_mix"], iterations=m["hfb_iter"], guess_kappa=True)
            s.estimation_seed(g["est_seed"])
            s.synchronize()
            t0 = time.perf_counter()
            s.scf_start()
            rows, _ = s.iterate_estimated(m["niter"], m["dt"], m["momentum"])
            t1 = time.perf_counter()
            s.close()
            if best is None or t1 - t0 < best[0]:
                best = (t1 - t0, rows)
        t, rows = best
        res[name] = {"value": m["niter"] / t, "unit": "iters/s",
                     "ms_per_iteration": t / m["niter"] * 1e3,
                     "final_energy_MeV": float(rows[-1, 0]),
                     "final_energy_diff_vs_reference_keV":
                         float(abs(rows[-1, 0] - g["traj"][-1, 0]) * 1e3)}
    return {
        "workload": "README sample (24Mg): 10x10x10 octant of a 20^3 Lagrange mesh, dx=0.8 "
                    "fm, 40 states (20 n + 20 p), SLy4, HFB 1250 MeV fm^3, IterativeEstimation, "
                    "MixingScheme 2, Q20 = 50 fm^2 for the first 10 iterations; start-up + 100 "
                    "iterations timed as one run (the fastest of the repetitions)",
        "fp64": res["fp64"], "mixed": res["mixed"],
    }


def lagrange_48(steps, warmup):
    from mocca_amd import FP64, MIXED, workloads
    w = workloads.lagrange_48()
    r64 = full_iteration_run(w, FP64, steps, warmup, False)
    r32 = full_iteration_run(w, MIXED, steps, warmup, False)
    rsp = full_iteration_run(dict(w, lagrange_split=True), MIXED, steps, warmup, False)
    return {
        "workload": "48x48x48 Lagrange octant mesh (48-point derivative matrices on the FP64 "
                    "matrix cores; mixed_fp32_split: on the FP32 matrix cores with the "
                    "matrices split into high and low FP32 parts), dx=0.7 fm, 600 states "
                    "(300 n + 300 p), HO seeds, SLy4 + Coulomb (direct solve on the 50^3 box), "
                    "HF filling N=Z=12; step = the whole Evolve iteration on the device",
        "steps": steps, "warmup": warmup,
        "fp64": {"value": 1.0 / r64["t"], "unit": "iters/s", "ms_per_step": r64["t"] * 1e3,
                 "kernel_ms_per_step": {k: v["ms_per_step"] for k, v in r64["prof"].items()}},
        "mixed": {"value": 1.0 / r32["t"], "unit": "iters/s", "ms_per_step": r32["t"] * 1e3,
                  "speedup_vs_fp64": r64["t"] / r32["t"],
                  "kernel_ms_per_step": {k: v["ms_per_step"] for k, v in r32["prof"].items()}},
        "total_energy_err_keV": float(abs(r32["rows"][-1, 0] - r64["rows"][-1, 0]) * 1e3),
        "mixed_fp32_split": {
            "value": 1.0 / rsp["t"], "unit": "iters/s", "ms_per_step": rsp["t"] * 1e3,
            "speedup_vs_fp64": r64["t"] / rsp["t"],
            "lag_derivs_ms_per_step": rsp["prof"].get("lag_derivs", {}).get("ms_per_step"),
            "total_energy_err_keV": float(abs(rsp["rows"][-1, 0] - r64["rows"][-1, 0]) * 1e3),
            "max_total_energy_err_keV_along_trajectory":
                float(np.max(np.abs(rsp["rows"][:, 0] - r64["rows"][:, 0])) * 1e3)},
    }


def batch_scan(steps, warmup):
    """64 independent constrained solves (uneven meshes 16^3..32^3 and
    20..160 states, alternating Q20 constraints) through mocca_batch_run, in
    FP64 and mixed precision from the same start; the same solves iterated
    one after another give the sequential rate."""
    import mocca_amd
    from mocca_amd import FP64, MIXED, workloads
    ws = workloads.batch_scan()

    def start(w, precision):
        s = scf_solve(w, precision)
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
        rep0 = sum(s.profile_get("graph_replay")[1] for s in solves)
        t0 = time.perf_counter()
        rows = mocca_amd.batch_run(solves, steps, dts, mus)
        t1 = time.perf_counter()
        replayed = sum(s.profile_get("graph_replay")[1] for s in solves) - rep0
        failed = sum(s.profile_get("graph_capture_failed")[1] for s in solves)
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
                     "speedup_vs_sequential": (t3 - t2) / (t1 - t0),
                     "replayed_fraction": replayed / (len(ws) * steps),
                     "captures_failed": failed}
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


def cpu_baseline(w, start, gpu_first, threads_all, start_ef=None):
    """The oracle's restatement (C restatement for derivatives, hPsi, step,
    Gram-Schmidt, densities, potentials, Coulomb, energy; numpy/LAPACK for the
    HFB matrices) of the same Evolve iteration with HFB pairing, from the same
    orthonormalised start states: one start-up and one timed iteration, at 1
    thread and at `threads_all` threads.  Also returns the oracle's energies
    of the start-up and the first iteration against the FP64 device run."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as ol
    from threadpoolctl import threadpool_limits
    from mocca_amd import functional
    f = w["functional"]
    consts = np.array([f["hbm"][0], f["hbm"][1], f["mass"][0], f["mass"][1], f["e2"],
                       functional.HBAR, w["dv"], f["neutrons"], f["protons"]])
    meta = dict(nx=w["nx"], ny=w["ny"], nz=w["nz"], dx=w["dx"], lagrange=w["lagrange"],
                nwt=w["M"], trc=w["trc"], com1body=f["com1body"], neutrons=f["neutrons"],
                protons=f["protons"], dt=w["dt_over_hbar"] * functional.HBAR,
                momentum=w["momentum"], pairing_type=2,
                pairing_strength=list(PAIRING["strength"]), pairing_alpha=list(PAIRING["alpha"]),
                pairing_cut=list(PAIRING["cut"]), pairing_mu=list(PAIRING["mu"]),
                rhosat=PAIRING["rhosat"], hfb_mix=PAIRING["hfb_mix"],
                hfb_iter=PAIRING["iterations"])
    g = dict(meta=meta, coef=np.asarray(f["B"]), consts=consts, qn=w["qn"], psi=start)
    if start_ef is not None:
        g["start_energies"], g["start_fermi"] = start_ef
    out = {}
    traj = None
    for nthr in (1, threads_all):
        ol.lib().or_set_threads(nthr)
        marks = []
        with threadpool_limits(nthr):
            r = ol.scf_run(g, 1, on_iteration=lambda k: marks.append(time.perf_counter()))
        out[nthr] = marks[1] - marks[0]
        traj = r["traj"][:, 0]
    rel = np.abs(traj - gpu_first[:2, 0]) / np.abs(traj)
    return dict(
        value=1.0 / out[threads_all], unit="iters/s", cores=threads_all, kind="port",
        sample=f"one Evolve iteration (after the start-up) of the headline workload on the "
               f"oracle (oracle/liboracle.so with OpenMP over states and points, numpy/LAPACK "
               f"for the HFB matrices) at {threads_all} threads: {out[threads_all]:.1f} s",
        one_thread={"value": 1.0 / out[1], "unit": "iters/s", "cores": 1,
                    "seconds_per_iteration": out[1]},
        energy_rel_diff_vs_device_fp64={"start_up": float(rel[0]), "iteration_1": float(rel[1])},
    )


def reference_container():
    path = os.path.join(PROFILES, REFERENCE_TIMING)
    if not os.path.exists(path):
        return None
    val = None
    for line in open(path):
        if line.startswith("| per iteration |"):
            val = float(line.split("|")[2].split()[0])
    if val is None:
        return None
    return {"value": 1.0 / val, "unit": "iters/s", "cores": 1, "kind": "reference",
            "seconds_per_iteration": val,
            "note": "the reference's own modules (oracle/_ref, amdflang) timed in the build "
                    "container on the same workload; the reference does not travel to the "
                    "GPU host, so this is read from profiles/" + REFERENCE_TIMING}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--hf-steps", type=int, default=100)
    ap.add_argument("--skip-hf", action="store_true")
    ap.add_argument("--micro-steps", type=int, default=50)
    ap.add_argument("--skip-micro", action="store_true")
    ap.add_argument("--skip-sample", action="store_true")
    ap.add_argument("--lagrange-steps", type=int, default=20)
    ap.add_argument("--skip-lagrange", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=20)
    ap.add_argument("--skip-batch", action="store_true")
    ap.add_argument("--skip-checkpoint", action="store_true")
    ap.add_argument("--legs", default="all",
                    help="comma list of the legs to run (headline, hf, micro, sample, "
                         "lagrange, batch, checkpoint) or 'all'; without 'headline' only "
                         "the chosen legs run and print their JSON (for profiling them "
                         "alone)")
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
              "(sharding over several GPUs is a design note, not built)", file=sys.stderr)
        sys.exit(2)
    legs = {"headline", "hf", "micro", "sample", "lagrange", "batch", "checkpoint"}
    if args.legs != "all":
        legs = set(args.legs.split(","))
    if "headline" not in legs:
        part = {}
        if "hf" in legs:
            part["hf_no_pairing"] = hf_no_pairing(args.hf_steps, args.warmup)
        if "micro" in legs:
            part["hpsi_microbench"] = hpsi_microbench(args.micro_steps, args.warmup)
        if "sample" in legs:
            part["readme_sample"] = readme_sample(args.warmup)
        if "lagrange" in legs:
            part["lagrange_48"] = lagrange_48(args.lagrange_steps, args.warmup)
        if "batch" in legs:
            part["batch_scan"] = batch_scan(args.batch_steps, args.warmup)
        if "checkpoint" in legs:
            from mocca_amd import workloads
            w = workloads.full_iteration(**NUCLEUS)
            part["checkpoint_v1"] = checkpoint(w, start_data(w))
        print(json.dumps(part), flush=True)
        return
    h = headline(args)
    w, r64 = h["w"], h["r64"]
    out = {
        "metric": METRIC,
        "value": 1.0 / r64["t"],
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r64["t"] * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "full_iteration_hfb: 40x40x40 octant mesh (EV8), dx=0.8 fm, 400 states "
                        "(200 n + 200 p), harmonic-oscillator eigenfunction seeds (b=2.4 fm) "
                        "with the single-particle and Fermi energies of one HF iteration as "
                        "start data, SLy4 + Coulomb (direct solve), "
                        "N=146 Z=94, "
                        "HFB pairing 1250 MeV fm^3 (100-state blocks, 200 x 200 HFB matrices, "
                        "the reference's Broyden Fermi solve), MixingScheme 0; step = the whole "
                        "Evolve iteration on the device",
            "mesh": [w["nx"], w["ny"], w["nz"]], "states": w["M"], "pairing": "HFB",
        },
        "roofline": h["roofline"],
        "dominant_stage": h["dominant_stage"],
        "stage_shares": h["stage_shares"],
        "kernel_ms_per_step": {k: v["ms_per_step"] for k, v in r64["prof"].items()},
        "launches_per_step": {k: v["launches_per_step"] for k, v in r64["prof"].items()},
        "mixed": h["mixed"],
        "energy_fp64_MeV": h["energy_fp64_MeV"],
        "energy_mixed_MeV": h["energy_mixed_MeV"],
    }
    print(f"# headline done: {out['value']:.2f} iters/s FP64", file=sys.stderr, flush=True)
    if not args.skip_hf and "hf" in legs:
        out["hf_no_pairing"] = hf_no_pairing(args.hf_steps, args.warmup)
        print("# HF leg done", file=sys.stderr, flush=True)
    if not args.skip_micro and "micro" in legs:
        out["hpsi_microbench"] = hpsi_microbench(args.micro_steps, args.warmup)
        print("# microbench done", file=sys.stderr, flush=True)
    if not args.skip_sample and "sample" in legs:
        out["readme_sample"] = readme_sample(args.warmup)
        print("# sample done", file=sys.stderr, flush=True)
    if not args.skip_lagrange and "lagrange" in legs:
        out["lagrange_48"] = lagrange_48(args.lagrange_steps, args.warmup)
        print("# lagrange leg done", file=sys.stderr, flush=True)
    if not args.skip_batch and "batch" in legs:
        out["batch_scan"] = batch_scan(args.batch_steps, args.warmup)
        print("# batch leg done", file=sys.stderr, flush=True)
    if not args.skip_checkpoint and "checkpoint" in legs:
        out["checkpoint_v1"] = checkpoint(w, h["start"])
        print("# checkpoint leg done", file=sys.stderr, flush=True)
    if not args.no_cpu:
        threads = max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16"))))
        out["cpu_baseline"] = cpu_baseline(w, r64["start"], r64["first"], threads,
                                           start_ef=h["start"])
        print("# cpu baseline done", file=sys.stderr, flush=True)
    out["reference_container"] = reference_container()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
