"""Langevin steps/s of the TIS-RNA + explicit-ion loop on one MI355X.

Workload (the benchmark's batched configuration): 16 replicas of the 16SCD
system, replica k in a box of 350 + 100 k / 15 A with c_Mg = 0.5 * 40^(k/15) mM
MgCl2 and 50 mM KCl (3,308 .. 9,469 particles, 86,655 in all), T = 0.596
kcal/mol (300 K), seeds 1..16, FP32 pair and multi-body forces with FP64
positions, velocities and sums, counter-based Philox noise, launch shapes
(hipGraph length, pair-gather lanes and block, list-build grid) from
tis_autotune run on a scratch copy before the timed region.  The replicas
differ in ion count, so no exchange moves are
attempted between them (an exchange needs two copies of one system).

Printed: one JSON line.  value = aggregate replica-steps per second of the
batch, its timed steps replayed from a hipGraph captured before the timer
(launch_mode says how the timed steps ran).  Beside it: the plain FP64
one-replica-per-launch path on the 3,659 particle 5 mM system (no graph, FP64
forces), the roofline of the dominant kernel from HIP-event timings, and the
reference's own step loop (compiled by oracle/Makefile) on this host's cores:
the 16 configurations as concurrent single-rank processes, one after another
on one core, and one configuration under its own MPI decomposition.

--workload highsalt measures one 16SCD replica at 20 mM MgCl2 + 100 mM KCl in a
450 A box (14,957 particles) with the per-term error of its FP32 forces.

--workload chain1500 measures the large synthetic chain instead: one replica
of a seeded 1,500-nt random-sequence chain (self-avoiding walk in a sphere of
75 A, secondary stacks on ~60% of adjacent pairs, 400 HB candidates) in a
600 A box with 10 mM MgCl2 + 100 mM KCl (32,914 particles), same precision,
noise and tuning; its line reports steps/s, ns/day, the whole-step
HBM rate and, from the committed counter summary, the VALU issue fraction.
"""
import argparse
import json
import os
import re
import shutil
import subprocess
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "16s-central-folding_amd"))

import tis_ctypes as T  # noqa: E402

METRIC = "Langevin steps/s, 16SCD+ions: batched FP32 replicas vs plain FP64 single"
FS_PER_STEP = 2.444            # h = 0.05 tau, tau = 48.89 fs (A, kcal/mol, amu)
HBM_PEAK_GBS = 8000.0


def batched_systems(lib, n_rep=16, T_kcal=0.596):
    out = []
    for k in range(n_rep):
        c_mg = 0.5e-3 * 40.0 ** (k / 15.0)
        box = 350.0 + 100.0 * k / 15.0
        d, _ = T.build_16scd(lib, T=T_kcal, c_mg=c_mg, c_k=0.050, seed=k + 1, box=box)
        out.append(d)
    return out


def algorithmic_bytes(eng, handles, kernel, lr_every=10):
    """Bytes one launch of `kernel` must move over the whole batch, from the
    per-unit figures (each input read once, each output written once):
      k_pairs      80 B per particle (position 32, type + list counts 16, force 32)
                   + 4 B per SR list entry + 4 B per LR entry / lr_every
      k_multibody  52 B per stack, 16 B per bond, 20 B per angle (+ 32 B force
                   per particle touched, counted as 32 B per particle)
      k_hbres      147 B per live HB candidate: its gate and energy, 120 B (6
                   positions of 32 B would be 192 B; the two shared with the gate
                   pair are counted once: 4 bead positions 128 B minus the 8 B the
                   gate reuses), and the resolution's 27 B (energy 8, gate 1, id 4,
                   outcome 2, incidence offsets 4, two incidence entries 8); 12 B
                   per site; 144 B per kept candidate (its 6 bead forces)
      k_integrate  312 B per particle (x, x_old, v read and written 192, pair forces
                   and random force read 96, cell-ordered copy written 24) + 32 B per
                   bead (multi-body force); the close of each replica's step is a few
                   hundred bytes per replica
    """
    tot = 0.0
    for r, h in enumerate(handles):
        e = eng.energies(r)
        d = h.d
        n = d["n_particles"]
        nact = e["n_hb_active"]
        if kernel == "k_pairs":
            tot += 80 * n + 4 * e["n_sr_pairs"] * 2 + 4.0 * e["n_lr_pairs"] * 2 / lr_every
        elif kernel == "k_multibody":
            tot += 52 * (d["n_st3"] + d["n_st2"]) + 16 * d["n_bonds"] + 20 * d["n_angles"] + 32 * d["n_beads"]
        elif kernel == "k_hbres":
            tot += 147 * nact + 12 * d["n_hb_sites"] + 144 * int(np.sum(eng.hb_outcome(r) == 1))
        elif kernel == "k_integrate":
            tot += 312 * n + 32 * d["n_beads"]
        else:
            return None
    return tot


def batched_params(n_rep=16):
    return [(0.5e-3 * 40.0 ** (k / 15.0), 350.0 + 100.0 * k / 15.0, k + 1) for k in range(n_rep)]


REF_EXE = os.path.join(REPO, "oracle", "_ref", "ref_harness")
REF_DATA = os.path.join(REPO, "16s-central-folding_amd", "data", "16SCD")


def _ref_cmd(c_mg, box, seed, steps, c_k=0.050):
    return [REF_EXE, "-D0.850", "-E3.975", "-d6.5", "-T0.596", f"-c{c_mg:.10f}", f"-f{c_k:.3f}", f"-s{seed}",
            f"-L{box:.6f}", f"-N{steps}", "-P"]


def _ref_rate(out):
    """steps/s of the reference's timed step loop from the harness's output line."""
    m = re.search(r"steps (\d+) seconds ([0-9.eE+-]+)", out)
    if not m:
        raise RuntimeError("unexpected reference harness output: " + out[-200:])
    return int(m.group(1)) / float(m.group(2))


def host_cores():
    """Host cores this process may use, at most 16 (the CPU share of one GPU)."""
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(16, n))


def cpu_baseline_reference(steps_one, steps_all, cases=None):
    """The reference's own step sequence (its translation unit compiled by
    oracle/Makefile into oracle/_ref/ref_harness, one MPI rank per process) on
    the replica configurations `cases` ((c_Mg, box, seed, c_K); default the 16
    of the batch).  Returns None when that binary was not built, else a dict:
      one_core   replica-steps/s with the configurations run one after another
                 on one core (`steps_one` steps each);
      all_cores  replica-steps/s with every configuration its own single-rank
                 process, all at once on up to host_cores() cores (`steps_all`
                 steps each; the reference runs its replicas as separate jobs).
                 With C >= n configurations it is the sum of the concurrent
                 rates, with fewer cores C times their harmonic mean (the
                 configurations packed onto C cores);
      mpi        steps/s of the first configuration under the reference's own
                 intra-replica decomposition (deterministic_forces_MPI,
                 fast.cpp:2524-2586, MPI_Allreduce :2572) for 1, 2, 4, ... ranks
                 (mpiexec of the installed MPICH), or None without mpiexec."""
    if not os.access(REF_EXE, os.X_OK):
        return None
    cases = cases or [(c, b, s, 0.050) for c, b, s in batched_params()]
    run = lambda cmd: subprocess.run(cmd, cwd=REF_DATA, capture_output=True, text=True, timeout=900,
                                     check=True).stdout
    rates_one = []
    for c_mg, box, seed, c_k in cases:
        rates_one.append(_ref_rate(run(_ref_cmd(c_mg, box, seed, steps_one, c_k))))
        log(f"reference CPU, one core: c_Mg {c_mg * 1e3:.2f} mM box {box:.1f} A: {rates_one[-1]:.0f} steps/s")
    one_core = len(cases) / sum(1.0 / r for r in rates_one)
    cores = host_cores()
    rates_all = []
    pending = list(cases)
    while pending:
        batch, pending = pending[:cores], pending[cores:]
        procs = [subprocess.Popen(_ref_cmd(c_mg, box, seed, steps_all, c_k), cwd=REF_DATA, stdout=subprocess.PIPE,
                                  stderr=subprocess.DEVNULL, text=True) for c_mg, box, seed, c_k in batch]
        for p in procs:
            out, _ = p.communicate(timeout=900)
            if p.returncode != 0:
                raise RuntimeError("reference harness failed")
            rates_all.append(_ref_rate(out))
    if cores >= len(cases):
        all_cores = sum(rates_all)
    else:
        all_cores = cores * len(rates_all) / sum(1.0 / r for r in rates_all)
    log(f"reference CPU, {min(cores, len(cases))} concurrent processes: {all_cores:.0f} replica-steps/s")
    mpi = None
    mpiexec = shutil.which("mpiexec") or ("/opt/conda/bin/mpiexec" if os.access("/opt/conda/bin/mpiexec", os.X_OK)
                                           else None)
    if mpiexec:
        mpi = {}
        c_mg, box, seed, c_k = cases[0]
        n = 1
        while n <= cores:
            try:
                out = run([mpiexec, "-n", str(n)] + _ref_cmd(c_mg, box, seed, steps_one, c_k))
                mpi[str(n)] = min(_ref_rate(line) for line in out.splitlines() if line.startswith("steps"))
            except (subprocess.SubprocessError, RuntimeError, ValueError) as e:
                log(f"mpiexec -n {n}: {e}")
                break
            log(f"reference CPU, mpiexec -n {n}: {mpi[str(n)]:.0f} steps/s")
            n *= 2
    return {"one_core": one_core, "all_cores": all_cores, "cores": min(cores, len(cases)), "mpi": mpi,
            "steps_one": steps_one, "steps_all": steps_all}


def reference_cpu_line(ref, what):
    """cpu_baseline entry from cpu_baseline_reference's result."""
    return {"value": ref["all_cores"], "unit": "steps/s", "cores": ref["cores"], "kind": "reference",
            "sample": f"the reference's step loop (fast.cpp:6380-6386, g++ -O3, one MPI rank per process): "
                      f"{what} as {ref['cores']} concurrent processes, {ref['steps_all']} steps each "
                      f"(each replica its own job, as the reference runs its replicas)",
            "one_core": {"value": ref["one_core"], "cores": 1,
                         "sample": f"the same configurations one after another on one core, "
                                   f"{ref['steps_one']} steps each"},
            "mpi_ranks_one_replica": ref["mpi"]}


def step_bytes(eng, handles):
    """Algorithmic bytes of one batched step, summed over replicas:
    B_step = 148 N + 8 P_SR + 0.8 P_LR + 56 H + 52 S + 16 n_bond + 20 n_angle
             + (40 N + 8 (P_SR + P_LR)) / R_rebuild
    with P the pair counts, H the live HB candidates, S the stacks and R_rebuild
    the mean steps between list rebuilds of the replica."""
    tot = 0.0
    for r, h in enumerate(handles):
        e = eng.energies(r)
        d = h.d
        n = d["n_particles"]
        psr, plr = e["n_sr_pairs"], e["n_lr_pairs"]
        rb = max(1.0, (e["klok"] - 1) / max(1, e["n_rebuilds"]))
        tot += (148 * n + 8 * psr + 0.8 * plr + 56 * e["n_hb_active"] + 52 * (d["n_st3"] + d["n_st2"])
                + 16 * d["n_bonds"] + 20 * d["n_angles"] + (40 * n + 8 * (psr + plr)) / rb)
    return tot


def copy_calibration(mib):
    """A plain device copy of `mib` MiB, five times: run under rocprofv3 --pmc to
    relate FETCH_SIZE / WRITE_SIZE to bytes moved."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": n * 4, "copies": 5}))


def measured_traffic(kernel, summary="pmc_traffic_batched16_16SCD.json"):
    """HBM bytes per launch of `kernel` from a committed counter summary
    (FETCH_SIZE and WRITE_SIZE runs scaled by a plain copy of known size), or None."""
    path = os.path.join(REPO, "profiles", summary)
    try:
        with open(path) as f:
            k = json.load(f)["kernels"].get(kernel)
    except (OSError, ValueError, KeyError):
        return None
    return None if k is None else k.get("hbm_bytes")


def cpu_baseline(systems, steps_each):
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_ctypes as O
    total_steps, total_t = 0, 0.0
    for d in systems:
        h = T.SystemHandle(d)
        o = O.Oracle(h)
        o.set_state(d["pos"], d["pos_old"], d["vel"])
        o.set_stream_glibc(1, (steps_each + 1) * 5000)
        t0 = time.perf_counter()
        rc = o.step(steps_each, 1.0)
        total_t += time.perf_counter() - t0
        if rc != 0:
            raise RuntimeError("oracle step failed")
        total_steps += steps_each
    return total_steps / total_t


def measured_valu(summary="pmc_chain1500.json"):
    """Per-kernel VALU issue fraction from a committed counter summary
    (SQ_INSTS_VALU per dispatch over the dispatch's duration times 256 CUs x
    2.4 GHz, one wave64 VALU instruction per CU per cycle), or None."""
    path = os.path.join(REPO, "profiles", summary)
    try:
        with open(path) as f:
            return {k: v for k, v in json.load(f)["kernels"].items() if "valu_frac" in v}
    except (OSError, ValueError, KeyError):
        return None


def run_chain(args, lib):
    d = T.build_synthetic(lib, n_nt=1500, stack_frac=0.6, n_hb=400, radius=75.0, T=0.596,
                          c_mg=0.010, c_k=0.100, seed=1, box=600.0)
    log(f"built the synthetic chain: {d['n_particles']} particles")
    h = T.SystemHandle(d)
    launch, tuning = tune(lib, [h], args)
    eng, dt, mode = timed_run(lib, [h], launch, args)
    value = args.steps / dt
    kt, dominant, avg_ms, bytes_launch, achieved = {}, None, None, None, None
    if args.profile_steps > 0:
        eng.set_profiling(True)
        eng.step(args.profile_steps)
        kt = eng.kernel_times()
        eng.set_profiling(False)
        dominant = max(kt, key=lambda k: kt[k][0] * kt[k][1])
        _, avg_ms = kt[dominant]
        bytes_launch = algorithmic_bytes(eng, [h], dominant)
        achieved = None if bytes_launch is None else bytes_launch / (avg_ms * 1e-3) / 1e9
    hb_trace = hb_trace_summary(eng, args.hb_trace) if args.hb_trace > 0 else None
    e = eng.energies()
    b_step = step_bytes(eng, [h])
    eng.close()
    cpu = None
    if not args.no_cpu:
        log("CPU baseline leg (oracle)")
        v = cpu_baseline([d], args.cpu_steps)
        cpu = {"value": v, "unit": "steps/s", "cores": 1, "kind": "port",
               "sample": f"FP64 CPU oracle (reference cell lists and force routines, glibc-stream noise), "
                         f"{args.cpu_steps} steps of the same chain system, single thread; the reference's "
                         f"own set-up reads only the 16SCD inputs, so its step loop cannot run this system"}
    valu = measured_valu()
    out = {
        "metric": "Langevin steps/s, large synthetic chain + ions, single replica",
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces / f64 state",
        "data": "synthetic: seeded random-sequence self-avoiding chain, ions placed by the reference's seeded rule",
        "config": {"workload": "1,500-nt synthetic chain, 600 A box, 10 mM MgCl2 + 100 mM KCl, T=0.596, "
                               "stacks on ~60% of adjacent pairs, 400 HB candidates",
                   "replicas": 1, "particles": d["n_particles"], "precision": "mixed", "rng": "philox",
                   "launch": launch, "autotune": tuning},
        "launch_mode": mode,
        "ns_per_day": value * FS_PER_STEP * 86400 * 1e-6,
        "roofline": None if dominant is None else {
            "bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dominant, "pmc_chain1500.json"), "bytes_per_launch": bytes_launch,
            "avg_ms": avg_ms},
        "roofline_step": {"bound": "hbm", "bytes_per_step": b_step, "achieved": b_step * value / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b_step * value / 1e9 / HBM_PEAK_GBS},
        "valu_issue_frac": None if valu is None else {k: v["valu_frac"] for k, v in valu.items()},
        "kernel_avg_ms": {k: v[1] for k, v in kt.items()},
        "hb_active": e["n_hb_active"],
        "rebuilds": e["n_rebuilds"],
        "hb_trace": hb_trace,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


TERM_NAMES = (("stack", T.TERM_STACK), ("backbone", T.TERM_BACKBONE), ("sr", T.TERM_SR), ("lr", T.TERM_LR),
              ("hb", T.TERM_HB))


def per_term_errors(lib, h, seed=2024):
    """FP32 (mixed-precision) forces of each term against the FP64 engine at the
    same state, same Philox draws (so the HB valence resolution takes the same
    branch): |dF| / |F| and max |dF| / RMS(F).  The FP64 engine equals the FP64
    CPU oracle to 1e-10 per term (tests/test_gpu_parity.py)."""
    e64 = T.Engine(lib, [h], precision=T.PREC_FP64, rng=T.RNG_PHILOX, seed=seed, use_graph=0)
    e32 = T.Engine(lib, [h], precision=T.PREC_MIXED, rng=T.RNG_PHILOX, seed=seed, use_graph=0)
    out = {}
    for name, mask in TERM_NAMES:
        f64, _ = e64.eval(mask)
        f32, _ = e32.eval(mask)
        rms = float(np.sqrt(np.mean(np.sum(f64 ** 2, axis=1))))
        out[name] = {"rel_norm": float(np.linalg.norm(f32 - f64) / np.linalg.norm(f64)),
                     "max_over_rms": float(np.abs(f32 - f64).max() / rms)}
    e64.close()
    e32.close()
    return out


def run_highsalt(args, lib):
    """One replica of the 16SCD chain at high salt: 450 A box, 20 mM MgCl2 + 100
    mM KCl (14,957 particles), FP32 forces / FP64 state, Philox noise; steps/s,
    ns/day and the per-term force error of the FP32 arithmetic."""
    d, _ = T.build_16scd(lib, T=0.596, c_mg=0.020, c_k=0.100, seed=1, box=450.0)
    h = T.SystemHandle(d)
    log(f"built the high-salt system: {d['n_particles']} particles")
    errs = per_term_errors(lib, h)
    log(f"per-term FP32 force error: {errs}")
    launch, tuning = tune(lib, [h], args)
    eng, dt, mode = timed_run(lib, [h], launch, args)
    value = args.steps / dt
    kt = {}
    if args.profile_steps > 0:
        eng.set_profiling(True)
        eng.step(args.profile_steps)
        kt = eng.kernel_times()
        eng.set_profiling(False)
    b_step = step_bytes(eng, [h])
    eng.close()
    cpu = None
    if not args.no_cpu:
        ref = cpu_baseline_reference(10 * args.cpu_steps, 10 * args.cpu_steps, [(0.020, 450.0, 1, 0.100)])
        if ref is not None:
            best = max([("1", ref["one_core"])] + list((ref["mpi"] or {}).items()), key=lambda kv: kv[1])
            cpu = {"value": best[1], "unit": "steps/s", "cores": int(best[0]), "kind": "reference",
                   "sample": f"the reference's step loop (fast.cpp:6380-6386, g++ -O3) on this system, "
                             f"{ref['steps_one']} steps, the fastest of one rank and mpiexec -n 2, 4, ... of its "
                             f"own MPI decomposition",
                   "one_core": {"value": ref["one_core"], "cores": 1},
                   "mpi_ranks_one_replica": ref["mpi"]}
    print(json.dumps({
        "metric": "Langevin steps/s, 16SCD at 20 mM MgCl2 + 100 mM KCl (450 A), single replica",
        "value": value, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 forces / f64 state",
        "data": "synthetic: 16SCD chain from the reference PDB, ions placed by the reference's seeded rule",
        "config": {"workload": "16SCD, 450 A box, 20 mM MgCl2 + 100 mM KCl, T=0.596", "replicas": 1,
                   "particles": d["n_particles"], "precision": "mixed", "rng": "philox", "launch": launch,
                   "autotune": tuning},
        "launch_mode": mode,
        "ns_per_day": value * FS_PER_STEP * 86400 * 1e-6,
        "per_term_force_error_fp32_vs_fp64": errs,
        "roofline": None,
        "roofline_step": {"bound": "hbm", "bytes_per_step": b_step, "achieved": b_step * value / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": b_step * value / 1e9 / HBM_PEAK_GBS},
        "kernel_avg_ms": {k: v[1] for k, v in kt.items()},
        "cpu_baseline": cpu,
    }))


def run_force_match(args, lib, n_snap=256):
    """Force-matching iterations over the 256-snapshot set of the 5 mM system:
    beads displaced by seeded N(0, 0.5^2) noise, ions as placed, targets from
    parameters scaled by 1 +- 10 %.  One iteration = new parameters in
    (tis_set_force_field) + loss and every gradient out (tis_force_match: forces,
    dL/dx for all 936,704 coordinates, dL/d hb_E, st3_E, st2_E, ion radii)."""
    import force_match as FM
    d, _ = T.build_16scd(lib, T=0.596, c_mg=0.005, c_k=0.050, seed=1)
    h = T.SystemHandle(d)
    snaps = FM.snapshots(d, n_snap, sigma=0.5, seed=0)
    tp = FM.Problem(lib, h, snaps, T.PREC_MIXED)
    n = d["n_particles"] * n_snap
    target = tp.run(snaps, FM.perturbed_params(d), np.zeros((n, 3)), grads=False)["forces"]
    base = FM.base_params(d)
    tp.set(snaps, base)
    log(f"{n_snap} snapshots set ({n} particles)")
    eng = tp.eng
    iters = max(1, args.steps // 100)
    for _ in range(max(1, args.warmup // 100)):
        eng.set_force_field(**base)
        eng.force_match(target)
    t0 = time.perf_counter()
    for _ in range(iters):
        eng.set_force_field(**base)
        out = eng.force_match(target)
    dt = time.perf_counter() - t0
    tp.close()
    value = iters / dt
    log(f"timed {iters} iterations in {dt:.2f} s")
    cpu = None
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(REPO, "oracle"))
        import oracle_ctypes as O
        o = O.Oracle(T.SystemHandle(d))
        k = 8
        t0 = time.perf_counter()
        for x in snaps[:k]:
            o.set_state(x, x, np.zeros_like(x))
            o.eval(T.TERM_ALL)
        per_snap = (time.perf_counter() - t0) / k
        cpu = {"value": 1.0 / (per_snap * n_snap), "unit": "loss evaluations/s", "cores": 1, "kind": "port",
               "sample": f"FP64 CPU oracle: lists and all force terms of {k} of the snapshots (forces and loss "
                         f"only; the reference has no gradient), scaled to the {n_snap}-snapshot set, one thread"}
    print(json.dumps({
        "metric": "force-matching iterations/s (loss + gradients, 256 snapshots of 16SCD at 5 mM)",
        "value": value, "unit": "iterations/s", "n_gpus": 1, "steps": iters, "warmup": max(1, args.warmup // 100),
        "ms_per_step": dt / iters * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 forces / f64 tangents and sums",
        "data": "synthetic: 16SCD beads from the reference PDB + N(0, 0.5^2) noise, ions by the reference's rule",
        "config": {"workload": "256 snapshots x 3,659 particles, targets from parameters scaled by 1 +- 10 %",
                   "snapshots": n_snap, "particles": n, "precision": "mixed", "rng": "philox"},
        "snapshot_gradients_per_s": value * n_snap,
        "loss": out["loss"],
        "roofline": None,
        "cpu_baseline": cpu,
    }))


def run_remd(args, lib, n_rep=16, every=1000):
    """Temperature replica exchange on the device: 16 copies of the 5 mM 16SCD
    system (3,659 particles each) on a geometric ladder T_k = 0.596 * 1.25^(k/15)
    kcal/mol (300-375 K), swap attempts between neighbours every 1,000 steps
    inside tis_step (energies, swap, Metropolis decision, restore: no host round
    trip)."""
    temps = [0.596 * 1.25 ** (k / (n_rep - 1)) for k in range(n_rep)]
    handles = []
    for k, t in enumerate(temps):
        d, _ = T.build_16scd(lib, T=t, c_mg=0.005, c_k=0.050, seed=1, box=350.0)
        handles.append(T.SystemHandle(d))
    launch, tuning = tune(lib, handles, args)
    eng = T.Engine(lib, handles, precision=T.PREC_MIXED, rng=T.RNG_PHILOX, seed=2024, exchange_every=every,
                   **launch)
    eng.step(max(every, args.warmup))
    a0, c0, _ = eng.exchange_stats()
    t0 = time.perf_counter()
    eng.step(args.steps)
    dt = time.perf_counter() - t0
    att, acc, _ = eng.exchange_stats()
    att, acc = att - a0, acc - c0
    value = n_rep * args.steps / dt
    eng.close()
    print(json.dumps({
        "metric": "Langevin steps/s, temperature replica exchange on the device (16 x 16SCD at 5 mM)",
        "value": value, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": max(every, args.warmup),
        "ms_per_step": dt / args.steps * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 forces / f64 state",
        "data": "synthetic: 16SCD chain from the reference PDB, ions placed by the reference's seeded rule",
        "config": {"workload": "16 x 16SCD, 5 mM MgCl2 + 50 mM KCl, 350 A, T ladder 0.596 * 1.25^(k/15), "
                               "exchange attempts every 1,000 steps", "replicas": n_rep,
                   "precision": "mixed", "rng": "philox", "launch": launch, "autotune": tuning},
        "exchange_attempts": int(att.sum()), "exchange_accepts": int(acc.sum()),
        "acceptance_per_pair": [None if a == 0 else round(float(c) / float(a), 3) for a, c in zip(att, acc)],
        "roofline": None, "cpu_baseline": None,
    }))


def timed_run(lib, handles, launch, args, **engine_kw):
    """Warm-up, then exactly args.steps timed steps of an Engine over `handles`.
    The hipGraph the timed steps replay (its length the tuned one, or the step
    count when that is shorter) is captured before the timer starts; the launch
    counts of the timed call say how the steps ran.  Returns (engine, seconds,
    launch mode)."""
    launch = dict(launch)
    use_graph = 0 if args.no_graph else 1
    if use_graph:
        launch["graph_steps"] = max(1, min(launch["graph_steps"], args.steps))
    kw = dict(precision=T.PREC_MIXED, rng=T.RNG_PHILOX, seed=2024)
    kw.update(engine_kw)
    eng = T.Engine(lib, handles, use_graph=use_graph, **kw, **launch)
    eng.step(args.warmup)
    log(f"warmup {args.warmup} steps done")
    if use_graph:
        eng.prepare_graph(launch["graph_steps"])
    c0 = eng.launch_counts()
    t0 = time.perf_counter()
    eng.step(args.steps)          # returns after a stream synchronize
    dt = time.perf_counter() - t0
    c1 = eng.launch_counts()
    mode = {k: c1[k] - c0[k] for k in c0}
    mode["graph_length"] = launch["graph_steps"] if use_graph else 0
    mode["mode"] = "graph" if mode["eager_steps"] == 0 else ("eager" if mode["graph_steps"] == 0 else "mixed")
    log(f"timed {args.steps} steps in {dt:.3f} s ({mode})")
    return eng, dt, mode


HB_PHASES = ("tables", "excess", "components", "roots", "resolution", "outcome_forces", "end")


def hb_trace_summary(eng, rounds):
    """Phase clock of the HB valence resolution (tis_trace_hb) over `rounds` x 64
    graph-replayed steps: per phase of a (replica, group) workgroup the mean, p50,
    p90 and max duration (us), the same for the group's whole span, the spread of
    the groups' start times within a step, the dispatch span (first start to last
    end of the step's resolution), the phases of the group that ends last, and the
    sizes (over-subscribed sites, components, live candidates, serial depth)."""
    eng.trace_hb(True)
    eng.prepare_graph(64)
    tr = []
    for _ in range(rounds):
        eng.step(64)
        tr.append(eng.hb_trace().astype(np.int64))
    eng.trace_hb(False)
    tr = np.concatenate(tr)[:, :, :eng.opt.hb_groups, :]          # (steps, replicas, groups, words)
    st = tr[..., :8]
    us = 0.01                                                        # 100 MHz clock
    ok = (st[..., 0] > 0) & (st[..., 7] >= st[..., 0])

    def stats(v):
        v = np.asarray(v, dtype=np.float64)
        return {"mean": float(v.mean()), "p50": float(np.percentile(v, 50)),
                "p90": float(np.percentile(v, 90)), "max": float(v.max())}

    d = np.diff(st, axis=-1) * us
    out = {"steps": int(tr.shape[0]), "workgroups_per_step": int(tr.shape[1] * tr.shape[2])}
    out["phase_us"] = {name: stats(d[..., k][ok]) for k, name in enumerate(HB_PHASES)}
    out["group_span_us"] = stats(((st[..., 7] - st[..., 0]) * us)[ok])
    t0 = np.where(ok, st[..., 0], np.iinfo(np.int64).max).reshape(tr.shape[0], -1).min(axis=1)
    t1 = np.where(ok, st[..., 7], 0).reshape(tr.shape[0], -1).max(axis=1)
    out["dispatch_span_us"] = stats((t1 - t0) * us)
    out["start_spread_us"] = stats(((st[..., 0] - t0[:, None, None]) * us)[ok])
    flat = st.reshape(tr.shape[0], -1, 8)
    last = np.argmax(np.where(ok.reshape(tr.shape[0], -1), flat[..., 7], 0), axis=1)
    crit = flat[np.arange(tr.shape[0]), last]
    out["last_group_phase_us"] = {name: float((np.diff(crit, axis=-1)[:, k] * us).mean())
                                  for k, name in enumerate(HB_PHASES)}
    out["last_group_start_us"] = float(((crit[:, 0] - t0) * us).mean())
    sz = tr[..., 9]
    out["serial_depth"] = stats(tr[..., 8][ok])
    out["over_subscribed_sites"] = stats((sz & 0xffff)[ok])
    out["components"] = stats(((sz >> 16) & 0xffff)[ok])
    out["live_candidates"] = stats((sz >> 32)[ok])
    return out


def tune(lib, handles, args):
    """Launch shapes for the timed run: tis_autotune's choice on a scratch copy of
    the batch (its trials are outside the timed region), or the defaults."""
    if args.launch is not None or args.no_autotune or args.no_graph:
        o = T.default_options(lib)
        launch = {k: getattr(o, k) for k in T.LAUNCH_FIELDS}
        launch.update(args.launch or {})
        return launch, None
    launch, rep = T.autotune(lib, handles, trial_steps=args.tune_steps, precision=T.PREC_MIXED,
                             rng=T.RNG_PHILOX, seed=2024)
    log(f"autotune: {launch} ({rep['trials']} settings, {rep['ms_per_step_start']:.4f} -> "
        f"{rep['ms_per_step_best']:.4f} ms/step)")
    return launch, rep


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", choices=["batched16", "chain1500", "forcematch256", "remd16", "highsalt"],
                    default="batched16")
    ap.add_argument("--steps", type=int, default=5000)
    ap.add_argument("--warmup", type=int, default=250)
    ap.add_argument("--profile-steps", type=int, default=300)
    ap.add_argument("--plain-steps", type=int, default=1000)
    ap.add_argument("--cpu-steps", type=int, default=200,
                    help="CPU steps per replica configuration, one core (cpu_baseline.one_core)")
    ap.add_argument("--cpu-steps-all", type=int, default=1000,
                    help="CPU steps per replica configuration, concurrent processes (cpu_baseline)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-graph", action="store_true", help="eager launches (for counter collection)")
    ap.add_argument("--no-autotune", action="store_true", help="default launch shapes instead of tis_autotune's")
    ap.add_argument("--tune-steps", type=int, default=400, help="steps per autotuner trial")
    ap.add_argument("--launch", type=json.loads, default=None, metavar="JSON",
                    help="launch shapes to use instead of tis_autotune's (e.g. for profiler runs)")
    ap.add_argument("--hb-trace", type=int, default=0, metavar="ROUNDS",
                    help="after the timed run, ROUNDS x 64 steps with the HB resolution's phase clock on "
                         "(hb_trace in the line)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one MI355X only: replicas are batched inside one launch on one GPU; "
              "multi-GPU runs (replica sharding over RCCL) are not part of this build.", file=sys.stderr)
        sys.exit(2)
    lib = T.load_library()
    if args.workload == "chain1500":
        run_chain(args, lib)
        return
    if args.workload == "forcematch256":
        run_force_match(args, lib)
        return
    if args.workload == "remd16":
        run_remd(args, lib)
        return
    if args.workload == "highsalt":
        run_highsalt(args, lib)
        return
    systems = batched_systems(lib)
    log(f"built {len(systems)} replicas")
    handles = [T.SystemHandle(d) for d in systems]
    launch, tuning = tune(lib, handles, args)
    eng, dt, mode = timed_run(lib, handles, launch, args)
    n_rep = len(handles)
    value = n_rep * args.steps / dt
    # per-kernel device times over a separate event-timed run of the same batch
    kt, dominant, avg_ms, bytes_launch, achieved = {}, None, None, None, None
    if args.profile_steps > 0:
        eng.set_profiling(True)
        eng.step(args.profile_steps)
        kt = eng.kernel_times()
        eng.set_profiling(False)
        dominant = max(kt, key=lambda k: kt[k][0] * kt[k][1])
        n_launch, avg_ms = kt[dominant]
        bytes_launch = algorithmic_bytes(eng, handles, dominant)
        achieved = None if bytes_launch is None else bytes_launch / (avg_ms * 1e-3) / 1e9
    hb_trace = hb_trace_summary(eng, args.hb_trace) if args.hb_trace > 0 else None
    energies = [eng.energies(r) for r in range(n_rep)]
    b_step = step_bytes(eng, handles)
    eng.close()
    # plain FP64 one-replica-per-launch path (5 mM, 350 A, 3,659 particles)
    d1, _ = T.build_16scd(lib, T=0.596, c_mg=0.005, c_k=0.050, seed=1, box=350.0)
    plain = None
    if args.plain_steps > 0:
        h1 = T.SystemHandle(d1)
        e1 = T.Engine(lib, [h1], precision=T.PREC_FP64, rng=T.RNG_PHILOX, use_graph=0)
        log("plain FP64 single-replica leg")
        e1.step(max(50, args.warmup // 5))
        t0 = time.perf_counter()
        e1.step(args.plain_steps)
        plain = args.plain_steps / (time.perf_counter() - t0)
        e1.close()
    cpu = None
    if not args.no_cpu:
        log("CPU baseline leg")
        ref = cpu_baseline_reference(args.cpu_steps, args.cpu_steps_all)
        if ref is not None:
            cpu = reference_cpu_line(ref, "the 16 replica configurations")
        else:
            v = cpu_baseline(systems, args.cpu_steps)
            cpu = {"value": v, "unit": "steps/s", "cores": 1, "kind": "port",
                   "sample": f"FP64 CPU oracle (reference cell lists, glibc-stream noise), "
                             f"{args.cpu_steps} steps of each of the 16 replicas, single thread"}
    valu = measured_valu("pmc_traffic_batched16_16SCD.json")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": dt / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 forces / f64 state",
        "data": "synthetic: 16SCD chain from the reference PDB, ions placed by the reference's seeded rule",
        "config": {"workload": "16 x 16SCD replicas, c_Mg = 0.5*40^(k/15) mM MgCl2 + 50 mM KCl, "
                               "box 350+100k/15 A, T=0.596, independent (no exchange)",
                   "replicas": n_rep, "particles": int(sum(d["n_particles"] for d in systems)),
                   "precision": "mixed", "rng": "philox", "launch": launch, "autotune": tuning},
        "launch_mode": mode,
        "ns_per_day_aggregate": value * FS_PER_STEP * 86400 * 1e-6,
        "plain_fp64_single": None if plain is None else {
            "value": plain, "unit": "steps/s", "particles": d1["n_particles"],
            "ns_per_day": plain * FS_PER_STEP * 86400 * 1e-6},
        "batched_vs_plain": None if plain is None else value / plain,
        "roofline": None if dominant is None else {
            "bound": "hbm", "kernel": dominant, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": None if achieved is None else achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dominant),
            "bytes_per_launch": bytes_launch, "avg_ms": avg_ms},
        "roofline_step": {"bound": "hbm", "bytes_per_step": b_step, "achieved": b_step * (value / n_rep) / 1e9,
                          "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": b_step * (value / n_rep) / 1e9 / HBM_PEAK_GBS},
        "valu_issue_frac": None if valu is None else {k: v["valu_frac"] for k, v in valu.items()},
        "kernel_avg_ms": {k: v[1] for k, v in kt.items()},
        "hb_active": [e["n_hb_active"] for e in energies],
        "hb_trace": hb_trace,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
