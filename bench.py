"""Throughput of the MI355X local MC sweep on the headline workload.

Workload (one "step" = one ParallelSweep::apply + Microstate::increment_step,
the body of HardSphereSim::advance, benchmarks/src/mc/hard_sphere.rs:90-111):
N = 1,000,188 unit-diameter hard spheres on an FCC lattice (4 x 63^3 sites)
scaled to packing fraction 0.50 in a periodic cubic box, translation moves with
maximum distance 0.08, seed 42.  Inputs are resident in HBM before the timed
region.  The headline value is the mixed FP32/FP64 mode; the plain FP64 mode is
timed the same way and must make identical decisions (checked here: identical
counts and bit-identical final positions).

Prints one JSON line (rank 0).
"""
import argparse
import json
import math
import os
import pathlib
import sys
import time

import numpy as np

ROOT = pathlib.Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "tests"))

METRIC = "trial moves/s, hard spheres N=1M φ=0.5; mixed vs plain FP64, identical accepts"
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=800, help="timed applies (800 = 200 sweeps x 4)")
    ap.add_argument("--warmup", type=int, default=40)
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="CPU baseline sample length")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-fp64", action="store_true", help="skip the plain FP64 comparison run")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--no-other", action="store_true", help="skip the short runs of the other configurations")
    ap.add_argument("--pass-layout", choices=["percolour", "persistent"], default="persistent",
                    help="mixed precision: one launch per colour pass, or one persistent launch per apply")
    ap.add_argument("--profile-period", type=int, default=8,
                    help="time the kernels of every k-th timed apply with HIP events")
    ap.add_argument("--placement-only", action="store_true",
                    help="only run the QuickCompress placement of 1M spheres and print its figures")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain device copy of known size (HBM counter calibration)")
    return ap.parse_args()


def copy_calibration():
    """A plain device-to-device copy of 1 GiB, for scaling the FETCH_SIZE and
    WRITE_SIZE counters of a profiled run to bytes."""
    import torch
    n = 1 << 30
    x = torch.empty(n, dtype=torch.uint8, device="cuda")
    y = torch.empty_like(x)
    x.fill_(1)
    torch.cuda.synchronize()
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": n, "copy_bytes_written": n, "launches": 3}), flush=True)


def pass_bytes(precision, n, n_spaces, nonempty, accepted_per_pass, dim=3):
    """Algorithmic HBM bytes of one pass launch (one colour of one cycle).

    Reads: the cell start offsets of every space and the neighbour copy of
    every site (the 3^D neighbourhoods of one colour's spaces cover the whole
    system once; 16 B float4 in mixed mode, 32 B double4 in FP64), and for each
    trial the chosen body's index (4 B) and FP64 position (32 B).  Writes: per
    accepted move the body position and site (2 x 32 B) and the neighbour copy.
    """
    colours = 2**dim
    active = nonempty / colours
    site = 16 if precision == "mixed" else 32
    write = 64 + site
    return 4 * (n_spaces + 1) + site * n + 36 * active + write * accepted_per_pass


def run_gpu(hmc, L, body, site, precision, steps, warmup, seed, kT, d, period, persistent=False, sdf=False):
    sim = hmc.GpuParallelSweep(L, body, sites=site, precision=precision, persistent=persistent)
    # warm-up applies: steps 0 .. warmup-1
    sim.run(0, seed, kT, d, warmup)
    sim.profile(True, period)
    t0 = time.perf_counter()
    acc, rej = sim.run(warmup, seed, kT, d, steps)  # returns after the stream is synchronised
    t1 = time.perf_counter()
    pass_ms, pass_n, build_ms, build_n = sim.read_profile()
    stats = sim.stats()
    final = sim.positions()
    sdf_out = None
    if sdf:
        # SDF of the final state (x_max 0.02, dx 1e-4): one call, then timed calls
        hist, ov = sim.sdf(0.02, 1e-4)
        t2 = time.perf_counter()
        for _ in range(5):
            hist, ov = sim.sdf(0.02, 1e-4)
        ms = (time.perf_counter() - t2) / 5 * 1e3
        sdf_out = {"ms_per_call": ms, "bodies_per_s": sim.n / (ms * 1e-3), "counted": int(hist.sum()),
                   "overlaps": ov, "beta_p_over_rho_one_frame": hmc.sdf_beta_p_over_rho(hist, sim.n, 1e-4, 3)}
    sim.close()
    return dict(elapsed=t1 - t0, accepted=acc, rejected=rej, pass_ms=pass_ms, pass_n=pass_n,
                build_ms=build_ms, build_n=build_n, stats=stats, final=final, sdf=sdf_out)


def measured_traffic(kernel):
    """HBM bytes per launch of the dominant kernel from the committed rocprofv3
    FETCH_SIZE / WRITE_SIZE summary (calibrated against a plain copy), or None."""
    name = {"k_pass_tile": "pass_tile_c2_hbm_traffic.md", "k_apply_tile": "apply_tile_c2_hbm_traffic.md"}
    f = ROOT / "profiles" / name.get(kernel.split("<")[0], "")
    if not f.is_file():
        return None
    for line in f.read_text().splitlines():
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def apply_bytes(n, n_spaces, trials, accepted):
    """Algorithmic HBM bytes of one persistent apply launch (every colour pass
    of every cycle).

    Reads: the cell start offsets of every space, every site's FP32 copy and
    body index once (the region boxes are staged once; what the halo reloads
    after each pass re-read is not counted), and for each trial the chosen
    body's FP64 position (32 B).  Writes: per accepted move the body position
    and site (2 x 32 B) and the FP32 copy (16 B)."""
    return 4 * (n_spaces + 1) + 20 * n + 32 * trials + 80 * accepted


def other_workloads(hmc, wl):
    """Short timed runs of the other configurations (trial moves/s; inputs
    resident on the device; a few warm-up steps first)."""
    res = {}

    def timed(fn, warm, steps):
        fn(0, warm)
        t0 = time.perf_counter()
        trials = fn(warm, steps)
        return trials / (time.perf_counter() - t0), trials

    # C1: hard disks, 128 x 128 square lattice at packing fraction 0.70, d = 0.15
    L, pos = wl.square_lattice_disks()
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="mixed")
    v, t = timed(lambda s0, k: sum(sim.run(s0, 42, 1.0, 0.15, k)), 20, 400)
    res["C1_hard_disks_16384"] = {"value": v, "unit": "trial moves/s", "steps": 400, "precision": "mixed",
                                  "pass_layout": {2: "persistent", 1: "percolour", 0: "global"}[sim.stats()["tiled"]]}
    sim.close()
    # C4: shifted LJ, FCC 64 x 64 x 32 at number density 0.8, kT 1.2, r_cut 2.5, d = 0.1
    L, pos = wl.fcc_number_density((64, 64, 32), 0.8)
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="lj_shifted", params=(1.0, 1.0, 2.5, 2.5),
                               precision="mixed")
    sim.run(0, 42, 1.2, 0.1, 1)
    r0 = sim.stats()["band_retests"]
    t0 = time.perf_counter()
    t = sum(sim.run(1, 42, 1.2, 0.1, 4))
    v = t / (time.perf_counter() - t0)
    res["C4_shifted_lj_524288"] = {"value": v, "unit": "trial moves/s", "steps": 4,
                                   "precision": "mixed: FP32 pair energies, FP64 re-test inside the error bound",
                                   "fp64_retest_fraction": (sim.stats()["band_retests"] - r0) / max(t, 1)}
    sim.close()
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="lj_shifted", params=(1.0, 1.0, 2.5, 2.5))
    v64, _ = timed(lambda s0, k: sum(sim.run(s0, 42, 1.2, 0.1, k)), 1, 4)
    res["C4_shifted_lj_524288"]["fp64_value"] = v64
    sim.close()
    # FP32 error of dH against FP64 (every trial of two applies, diagnostic run)
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="lj_shifted", params=(1.0, 1.0, 2.5, 2.5),
                               precision="mixed", fp32_error=True)
    sim.run(0, 42, 1.2, 0.1, 2)
    st = sim.stats()
    res["C4_shifted_lj_524288"]["fp32_dH_error"] = {"max_abs": st["fp32_error_max"],
                                                    "max_over_bound": st["fp32_error_bound"]}
    sim.close()
    # C3: unit cubes, simple cubic 64^3 at packing fraction 0.55; a step is a
    # translate apply (d = 0.05) and a rotate apply (0.1), as the reference's
    # polyhedron benchmark; counts are read back after every apply
    L, pos = wl.simple_cubic()
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="convex", vertices=wl.UNIT_CUBE,
                               orientations=np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1)))
    state = {"sub": 0}

    def c3(s0, k):
        tot = 0
        for step in range(s0, s0 + k):
            a, r, used = sim.apply(step, 0, 42, 1.0, 0.05)
            a2, r2, _ = sim.apply_rotate(step, used, 42, 1.0, 0.1)
            tot += a + r + a2 + r2
        return tot
    v, t = timed(c3, 1, 4)
    res["C3_cubes_262144"] = {"value": v, "unit": "trial moves/s", "steps": 4, "precision": "fp64",
                              "moves": "translate 0.05 + rotate 0.1 per step"}
    sim.close()
    # C5: 256 replicas of hard spheres (FCC, N 2,048 .. 62,500, packing
    # fractions 0.30 .. 0.54), seeds 42 + r, d = 0.08, one graph per step
    systems = wl.batched_replicas()
    batch = hmc.GpuBatch([(L, wl.wrap(L, p), None) for L, p in systems], precision="mixed")
    seeds = [42 + r for r in range(len(systems))]
    tune_ms, policy = batch.autotune(0, seeds, 1.0, 0.08, 2)  # tile layout (state left unchanged)
    v, t = timed(lambda s0, k: sum(a + r for a, r in batch.run(s0, seeds, 1.0, 0.08, k)), 2, 40)
    res["C5_replicas_256"] = {"value": v, "unit": "trial moves/s", "steps": 40, "precision": "mixed",
                              "bodies": int(sum(len(p) for _, p in systems)),
                              "pass_layout": "tiles" if batch.stats(0)["tiled"] else "lanes",
                              "autotuned_policy_cus": policy}
    # SDF of every replica (x_max 0.02, dx 1e-4): one call after the run,
    # then timed calls; beta P / rho from the single frame
    nb = sum(len(p) for _, p in systems)
    hist, ov = batch.sdf(0.02, 1e-4)
    t0 = time.perf_counter()
    for _ in range(5):
        hist, ov = batch.sdf(0.02, 1e-4)
    ms = (time.perf_counter() - t0) / 5 * 1e3
    z = [hmc.sdf_beta_p_over_rho(hist[r], len(p), 1e-4, 3) for r, (_L, p) in enumerate(systems)]
    res["C5_replicas_256"]["sdf"] = {"ms_per_call": ms, "bodies_per_s": nb / (ms * 1e-3), "overlaps": int(ov.sum()),
                                     "beta_p_over_rho_one_frame": {"phi_0.30": z[0], "phi_0.54": z[-1]}}
    batch.close()
    # the same batch in plain FP64 (identical decisions)
    batch = hmc.GpuBatch([(L, wl.wrap(L, p), None) for L, p in systems], precision="fp64")
    v64, _ = timed(lambda s0, k: sum(a + r for a, r in batch.run(s0, seeds, 1.0, 0.08, k)), 2, 40)
    res["C5_replicas_256"]["fp64_value"] = v64
    res["C5_replicas_256"]["mixed_over_fp64"] = res["C5_replicas_256"]["value"] / v64
    batch.close()
    res["placement_quick_compress_1M"] = placement(hmc, wl)
    return res


def placement(hmc, wl):
    """Placement (benchmarks/src/place.rs:75-99 after QuickInsert): FCC 4*63^3
    spheres at packing fraction 0.45 compressed to 0.50 by QuickCompress under
    the overlap penalty (OverlapPenalty::default(), delta 1; FP64), one
    QuickCompress call every 20 steps and a translate sweep (d 0.1) each step."""
    L, pos = wl.fcc_spheres(packing_fraction=0.45)
    n = len(pos)
    vt = n * (np.pi / 6.0) / 0.50
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="overlap_penalty")
    qc = hmc.QuickCompress(vt)
    qc_s = sweep_s = 0.0
    calls = trials = 0
    step = 0
    t0 = time.perf_counter()
    while not qc.is_complete() and step < 1000:
        sub = 0
        if step % 20 == 0:
            t1 = time.perf_counter()
            qc.apply(sim, step, 0, 42)
            qc_s += time.perf_counter() - t1
            calls += 1
            sub = 1
        t1 = time.perf_counter()
        a, r, _ = sim.apply(step, sub, 42, 1.0, 0.1)
        sweep_s += time.perf_counter() - t1
        trials += a + r
        step += 1
    total = time.perf_counter() - t0
    res = {
        "bodies": n, "from_packing_fraction": 0.45, "to_packing_fraction": 0.50,
        "complete": qc.is_complete(), "steps": step, "seconds": total,
        "final_packing_fraction": n * (np.pi / 6.0) / float(np.prod(sim.box_edges)),
        "quick_compress_calls": calls, "ms_per_quick_compress_call": qc_s / calls * 1e3,
        "overlap_penalty_sweep_trial_moves_per_s": trials / sweep_s, "ambiguous": qc.ambiguous}
    sim.close()
    return res


def cpu_baseline(L, pos, seconds, seed, kT, d, threads=None, model="hard_sphere", params=(1.0,)):
    import _oracle as O
    threads = threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()
    O.set_threads(threads)
    sys_ = O.OracleSystem(L, pos, model=model, params=params, seed=seed)
    sys_.time_applies(True, kT, d, 0.5)  # warm caches
    moves, elapsed, applies = sys_.time_applies(True, kT, d, seconds)
    return dict(value=moves / elapsed, unit="trial moves/s", cores=threads, kind="port",
                sample=f"{applies} ParallelSweep applies of the same N=1,000,188 system ({elapsed:.1f} s), "
                       f"C++ restatement of hoomd-rs 1.4.0 with OpenMP over same-colour spaces")


def main():
    args = parse()
    if args.copy_calibration:
        copy_calibration()
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import _pkg
    hmc = _pkg.pkg()
    wl = _pkg.workloads()
    if args.placement_only:
        print(json.dumps({"placement_quick_compress_1M": placement(hmc, wl)}), flush=True)
        return
    kT, d = 1.0, 0.08
    L, pos = wl.fcc_spheres((63, 63, 63), packing_fraction=0.50)
    n = len(pos)
    # the microstate's stored bodies (Microstate::add_body wraps them); the
    # library derives the sites as wrap(body + 0)
    body, site = wl.wrap(L, pos), None

    mixed = run_gpu(hmc, L, body, site, "mixed", args.steps, args.warmup, args.seed, kT, d, args.profile_period,
                    persistent=args.pass_layout == "persistent", sdf=True)
    trials = mixed["accepted"] + mixed["rejected"]
    value = trials / mixed["elapsed"]
    out = {
        "metric": METRIC, "value": value, "unit": "trial moves/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": 1e3 * mixed["elapsed"] / args.steps, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64+f32", "data": "synthetic",
        "config": {"workload": "hard spheres 3D, FCC 4*63^3, packing fraction 0.50, translate d=0.08, "
                               "ParallelSweep apply per step", "n_bodies": n, "box_edge": float(L[0]),
                   "precision": "mixed FP32 pair test with FP64 band re-test", "seed": args.seed,
                   "pass_layout": args.pass_layout},
        "sweeps_per_s": value / n,
        "acceptance": mixed["accepted"] / max(trials, 1),
    }
    # roofline of the dominant kernel: the persistent apply (one launch per
    # apply) or the per-colour trial pass
    n_spaces = 100**3
    cycles = mixed["stats"]["cycles"] / max(mixed["stats"]["applies"], 1)
    passes = args.steps * 8 * cycles
    nonempty = trials / (args.steps * cycles)
    acc_per_pass = mixed["accepted"] / passes
    if mixed["pass_n"]:
        avg_ms = mixed["pass_ms"] / mixed["pass_n"]
        if mixed["stats"]["tiled"] == 2:
            b = apply_bytes(n, n_spaces, trials / args.steps, mixed["accepted"] / args.steps)
            kernel = "k_apply_tile<3>"
        else:
            b = pass_bytes("mixed", n, n_spaces, nonempty, acc_per_pass)
            kernel = "k_pass_tile<3,MIXED>" if mixed["stats"]["tiled"] else "k_pass<3,MIXED>"
        achieved = b / (avg_ms * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                           "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic(kernel),
                           "kernel": kernel, "avg_launch_ms": avg_ms,
                           "algorithmic_bytes_per_launch": b, "launches": mixed["pass_n"],
                           "build_ms_per_apply": mixed["build_ms"] / max(mixed["build_n"], 1)}
    out["sdf"] = mixed["sdf"]  # the scale distribution function of the final state (hmc_sdf), timed apart
    # FP32 error of the pair test against FP64: every pair within 1.01 sigma
    # of four applies measured both ways (a diagnostic run on the per-colour
    # tile passes, not timed)
    diag = hmc.GpuParallelSweep(L, body, sites=site, precision="mixed", fp32_error=True)
    diag.run(0, args.seed, kT, d, 4)
    dst = diag.stats()
    out["fp32_d2_error"] = {"max_abs": dst["fp32_error_max"], "max_over_band_bound": dst["fp32_error_bound"]}
    diag.close()
    out["band_retests"] = mixed["stats"]["band_retests"]  # FP64 re-tests over warm-up + timed applies
    if not args.no_fp64:
        f64 = run_gpu(hmc, L, body, site, "fp64", args.steps, args.warmup, args.seed, kT, d, args.profile_period)
        same = (f64["accepted"], f64["rejected"]) == (mixed["accepted"], mixed["rejected"])
        import numpy as np
        same = same and np.array_equal(f64["final"][0], mixed["final"][0]) and np.array_equal(
            f64["final"][1], mixed["final"][1])
        v64 = (f64["accepted"] + f64["rejected"]) / f64["elapsed"]
        out["fp64"] = {"value": v64, "ms_per_step": 1e3 * f64["elapsed"] / args.steps,
                       "avg_pass_ms": f64["pass_ms"] / max(f64["pass_n"], 1),
                       "identical_decisions_and_positions": bool(same)}
        out["mixed_over_fp64"] = value / v64
    if not args.no_other:
        out["other_workloads"] = other_workloads(hmc, wl)
    if not args.no_cpu_baseline:
        cb = cpu_baseline(L, pos, args.cpu_seconds, args.seed, kT, d)
        out["cpu_baseline"] = cb
        out["gpu_over_cpu"] = value / cb["value"]
        # the same restatement on one core (a shorter sample)
        out["cpu_baseline_1core"] = cpu_baseline(L, pos, args.cpu_seconds / 3.0, args.seed, kT, d, threads=1)
        if not args.no_other:
            # the overlap-penalty sweeps of the placement (FCC 4*63^3 at packing
            # fraction 0.45, d 0.1) on the same host cores, a shorter sample
            Lp, pp = wl.fcc_spheres(packing_fraction=0.45)
            pc = cpu_baseline(Lp, wl.wrap(Lp, pp), args.cpu_seconds / 3.0, args.seed, 1.0, 0.1,
                              model="overlap_penalty", params=(1e4, 0.2, 100.0, 1.0))
            pc["sample"] = pc["sample"].replace("ParallelSweep applies", "overlap-penalty ParallelSweep applies")
            pl = out["other_workloads"]["placement_quick_compress_1M"]
            pl["cpu_baseline_sweeps"] = pc
            pl["sweeps_gpu_over_cpu"] = pl["overlap_penalty_sweep_trial_moves_per_s"] / pc["value"]
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
