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
    ap.add_argument("--only", choices=["c1", "c3", "c4", "c5", "placement", "gsd", "tune"],
                    help="only run this other configuration (with its CPU baseline unless --no-cpu-baseline)")
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
    """HBM bytes per launch of the dominant kernel (per step for the batch's
    whole-step roofline) from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE
    summary (calibrated against a plain copy), or None."""
    name = {"k_pass_tile": "pass_tile_c2_hbm_traffic.md", "k_apply_tile": "apply_tile_c2_hbm_traffic.md",
            "k_bapply_tile": "bapply_tile_c5_hbm_traffic.md"}
    f = ROOT / "profiles" / name.get(kernel.split("<")[0], "")
    if not f.is_file():
        return None
    for line in f.read_text().splitlines():
        if line.startswith(("traffic_bytes_per_launch:", "traffic_bytes_per_step:")):
            return float(line.split(":", 1)[1])
    return None


def model_apply_bytes(n, n_spaces, cycles, accepted, retests, dim=3, mixed=True):
    """Compulsory-traffic model of one apply, the per-trial figure the
    roofline uses: every colour pass (cycles x 2^D of them) reads every body
    position once (P = 4 B per FP32 coordinate of the cell-relative copy in
    mixed mode, 8 B per FP64 coordinate otherwise) with its cell index (4 B),
    and the cell offsets (4 (S + 1) B); each accepted move writes its FP64
    position (8 B per coordinate) and, in mixed mode, its FP32 copy; each
    FP64 band re-test reads one FP64 position."""
    p = (4 if mixed else 8) * dim
    w = 8 * dim + (4 * dim if mixed else 0)
    return cycles * 2**dim * (n * (p + 4) + 4 * (n_spaces + 1)) + w * accepted + 8 * dim * retests


def apply_bytes(n, n_spaces, trials, accepted):
    """Algorithmic HBM bytes of one persistent apply launch (every colour pass
    of every cycle).

    Reads: the cell start offsets of every space, every site's FP32 copy and
    body index once (the region boxes are staged once; what the halo reloads
    after each pass re-read is not counted), and for each trial the chosen
    body's FP64 position (32 B).  Writes: per accepted move the body position
    and site (2 x 32 B) and the FP32 copy (16 B)."""
    return 4 * (n_spaces + 1) + 20 * n + 32 * trials + 80 * accepted


PEAK_FP64_TFLOPS = 78.6  # MI355X vector FP64 (no MFMA on this path)
CPU_WARM_S, CPU_TIMED_S = 2.0, 4.0  # the reference harness (benchmarks/src/benchmark.rs:29-35, 46-89)


def lj_trial_flops(n, n_spaces, density, r_cut, dim=3):
    """Algorithmic flops of one LJ trial: the final and initial energies each
    test every candidate of the 3^D neighbour spaces (3 sub, 3 mul-add and a
    compare: 8 flops) and evaluate the pairs within r_cut (a division, 4
    multiplies for s^12 and s^6, a subtraction, a multiply and the sum: 10)."""
    cand = 3**dim * n / n_spaces
    within = (4.0 / 3.0) * math.pi * r_cut**3 * density if dim == 3 else math.pi * r_cut**2 * density
    return 2.0 * (8.0 * cand + 10.0 * within)


def convex_trial_flops(n, n_spaces, density):
    """Algorithmic flops of one trial on unit cubes (an estimate from the
    operation counts of the code path): the circumsphere test of every
    candidate of the 27 neighbour spaces (8 flops), and XenoCollide for every
    candidate within the circumsphere cutoff sqrt(3) (density x 21.8 of
    them): pair_system_to_local and the rotation matrix (about 90 flops), two
    discovery supports (76 each: R^T n, 8 vertex dot products, R s) and 1.6
    refinement steps on average (116 each), about 450 flops per pair."""
    cand = 27.0 * n / n_spaces
    pairs = density * (4.0 / 3.0) * math.pi * math.sqrt(3.0) ** 3
    return 8.0 * cand + 450.0 * pairs


def cpu_sample(make, kT, d, threads, parallel, what, warm=CPU_WARM_S, seconds=CPU_TIMED_S):
    """Trial moves/s of the oracle (the C++ restatement of hoomd-rs 1.4.0) on
    the host cores: `warm` seconds untimed, then `seconds` timed, as the
    reference harness times its benchmarks."""
    import _oracle as O
    O.set_threads(threads)
    sys_ = make()
    sys_.time_applies(parallel, kT, d, warm)
    moves, elapsed, applies = sys_.time_applies(parallel, kT, d, seconds)
    del sys_
    kind = "ParallelSweep applies (OpenMP over same-colour spaces)" if parallel else "serial Sweep applies"
    return dict(value=moves / elapsed, unit="trial moves/s", cores=threads, kind="port",
                sample=f"{applies} {kind} of {what} ({elapsed:.1f} s after {warm:.0f} s warm-up), "
                       f"C++ restatement of hoomd-rs 1.4.0")


def all_cores():
    return int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or os.cpu_count()


def other_workloads(hmc, wl, cpu=True, only=None, fp64=True):
    """Timed runs of the other configurations (trial moves/s; inputs resident
    on the device; warm-up steps first), each with the roofline of its
    dominant kernel and the oracle timed on the host cores beside it."""
    import _oracle as O
    res = {}

    def timed(fn, warm, steps):
        fn(0, warm)
        t0 = time.perf_counter()
        trials = fn(warm, steps)
        return trials / (time.perf_counter() - t0), trials

    def want(k):
        return only is None or only == k

    if want("c1"):
        # C1: hard disks, 128 x 128 square lattice at packing fraction 0.70, d = 0.15
        L, pos = wl.square_lattice_disks()
        sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="mixed")
        sim.run(0, 42, 1.0, 0.15, 20)
        sim.profile(True, 4)
        t0 = time.perf_counter()
        acc, rej = sim.run(20, 42, 1.0, 0.15, 400)
        v = (acc + rej) / (time.perf_counter() - t0)
        pm, pn, _bm, _bn = sim.read_profile()
        st = sim.stats()
        layout = {2: "persistent", 1: "percolour", 0: "global"}[st["tiled"]]
        r = {"value": v, "unit": "trial moves/s", "steps": 400, "precision": "mixed", "pass_layout": layout}
        if pn and st["tiled"] == 2:
            ms = pm / pn
            S = 134 * 134
            cyc = st["cycles"] / max(st["applies"], 1)
            b = model_apply_bytes(len(pos), S, cyc, acc / 400, st["band_retests"] / 420, dim=2)
            bp = apply_bytes(len(pos), S, (acc + rej) / 400, acc / 400)
            r["roofline"] = {"bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": b / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS, "traffic": None,
                             "kernel": "k_apply_tile<2>", "avg_launch_ms": ms, "algorithmic_bytes_per_launch": b,
                             "persistent_model_bytes_per_launch": bp,
                             "persistent_model_frac": bp / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
        sim.close()
        if cpu:
            mk = lambda: O.OracleSystem(L, wl.wrap(L, pos), seed=42)
            r["cpu_baseline"] = cpu_sample(mk, 1.0, 0.15, 1, False, "the same 16,384 disks")
            r["cpu_baseline_all_cores"] = cpu_sample(mk, 1.0, 0.15, all_cores(), True, "the same 16,384 disks")
            r["gpu_over_cpu"] = v / r["cpu_baseline"]["value"]
        res["C1_hard_disks_16384"] = r
    if want("c4"):
        # C4: shifted LJ, FCC 64 x 64 x 32 at number density 0.8, kT 1.2, r_cut 2.5, d = 0.1
        prm = (1.0, 1.0, 2.5, 2.5)
        L, pos = wl.fcc_number_density((64, 64, 32), 0.8)
        sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="lj_shifted", params=prm, precision="mixed")
        sim.run(0, 42, 1.2, 0.1, 2)
        r0 = sim.stats()["band_retests"]
        sim.profile(True, 1)
        t0 = time.perf_counter()
        acc, rej = sim.run(2, 42, 1.2, 0.1, 20)
        t = acc + rej
        v = t / (time.perf_counter() - t0)
        pm, pn, bm, bn = sim.read_profile()
        st = sim.stats()
        e, _bound = sim.total_energy()
        r = {"value": v, "unit": "trial moves/s", "steps": 20,
             "precision": "mixed: FP32 pair energies, FP64 re-test inside the error bound",
             "pass_layout": {2: "persistent", 1: "percolour", 0: "percolour"}[st["tiled"]],
             "fp64_retest_fraction": (st["band_retests"] - r0) / max(t, 1),
             "u_per_body_final": e / len(pos)}
        shape = [int(np.floor(x / 2.5)) for x in L]
        shape = [k - (k % 2) for k in shape]
        n_spaces = int(np.prod(shape))
        if pn:
            ms = pm / pn
            trials_per_launch = t / pn
            fl = lj_trial_flops(len(pos), n_spaces, 0.8, 2.5) * trials_per_launch
            ach = fl / (ms * 1e-3) / 1e12
            r["roofline"] = {"bound": "valu_fp64", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                             "frac": ach / PEAK_FP64_TFLOPS, "traffic": None,
                             "kernel": "k_apply_lj<3>" if st["tiled"] == 2 else "k_pass_lj<3,MIXED>",
                             "avg_launch_ms": ms,
                             "algorithmic_flops_per_launch": fl, "launches": pn,
                             "flops_per_trial": lj_trial_flops(len(pos), n_spaces, 0.8, 2.5)}
        sim.close()
        if fp64:
            sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="lj_shifted", params=prm)
            v64, _ = timed(lambda s0, k: sum(sim.run(s0, 42, 1.2, 0.1, k)), 1, 4)
            r["fp64_value"] = v64
            r["mixed_over_fp64"] = v / v64
            sim.close()
            # FP32 error of dH against FP64 (every trial of two applies, diagnostic run)
            sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="lj_shifted", params=prm, precision="mixed",
                                       fp32_error=True)
            sim.run(0, 42, 1.2, 0.1, 2)
            st = sim.stats()
            r["fp32_dH_error"] = {"max_abs": st["fp32_error_max"], "max_over_bound": st["fp32_error_bound"]}
            sim.close()
        if cpu:
            mk = lambda: O.OracleSystem(L, wl.wrap(L, pos), model="lj_shifted", params=prm, seed=42)
            r["cpu_baseline"] = cpu_sample(mk, 1.2, 0.1, all_cores(), True, "the same 524,288 LJ bodies")
            r["gpu_over_cpu"] = v / r["cpu_baseline"]["value"]
        res["C4_shifted_lj_524288"] = r
    if want("c3"):
        # C3: unit cubes, simple cubic 64^3 at packing fraction 0.55; a step is a
        # translate apply (d = 0.05) and a rotate apply (0.1), as the reference's
        # polyhedron benchmark (hmc_run with both moves: no host reads between)
        L, pos = wl.simple_cubic()
        q0 = np.tile([1.0, 0.0, 0.0, 0.0], (len(pos), 1))
        sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="convex", vertices=wl.UNIT_CUBE, orientations=q0)
        moves = [("translate", 0.05), ("rotate", 0.1)]
        # orientations decorrelated from the lattice before timing: 30 steps
        # of translate 0.05 + rotate 0.5, then one step of the timed moves
        sim.run_moves(0, 42, 1.0, [("translate", 0.05), ("rotate", 0.5)], 30)
        sim.run_moves(30, 42, 1.0, moves, 1)
        t0 = time.perf_counter()
        cnt = sim.run_moves(31, 42, 1.0, moves, 4)
        t = sum(a + b for a, b in cnt)
        v = t / (time.perf_counter() - t0)
        # pass timing: one more step with every launch timed by HIP events
        sim.profile(True, 1)
        t_prof = sum(a + b for a, b in sim.run_moves(35, 42, 1.0, moves, 1))
        pm, pn, _bm, _bn = sim.read_profile()
        bq_now, sq = sim.orientations()
        bp_now = sim.positions()[0]
        sim_state = sim.positions() + sim.orientations()
        r = {"value": v, "unit": "trial moves/s", "steps": 4, "precision": "fp64",
             "moves": "translate 0.05 + rotate 0.1 per step, after 30 steps of translate 0.05 + rotate 0.5",
             "orientation_mean_abs_s": float(np.mean(np.abs(sq[:, 0]))),
             "acceptance": {"translate": cnt[0][0] / max(sum(cnt[0]), 1), "rotate": cnt[1][0] / max(sum(cnt[1]), 1)}}
        if pn:
            ms = pm / pn
            fpt = convex_trial_flops(len(pos), 44**3, len(pos) / float(np.prod(L)))
            fl = fpt * t_prof / pn
            ach = fl / (ms * 1e-3) / 1e12
            r["roofline"] = {"bound": "valu_fp64", "achieved": ach, "peak": PEAK_FP64_TFLOPS, "unit": "TFLOP/s",
                             "frac": ach / PEAK_FP64_TFLOPS, "traffic": None, "kernel": "k_pass_convex",
                             "avg_launch_ms": ms, "algorithmic_flops_per_launch": fl, "launches": pn,
                             "flops_per_trial": fpt}
        sim.close()
        if fp64:
            # the mixed mode (FP32 separating axes, FP64 XenoCollide inside
            # their margin) over the same steps: the same decisions, so the
            # same final state bit for bit
            sm = hmc.GpuParallelSweep(L, wl.wrap(L, pos), model="convex", vertices=wl.UNIT_CUBE, orientations=q0,
                                      precision="mixed")
            sm.run_moves(0, 42, 1.0, [("translate", 0.05), ("rotate", 0.5)], 30)
            sm.run_moves(30, 42, 1.0, moves, 1)
            b0 = sm.stats()["band_retests"]
            t0 = time.perf_counter()
            cm = sm.run_moves(31, 42, 1.0, moves, 4)
            vm = sum(a + b for a, b in cm) / (time.perf_counter() - t0)
            band = sm.stats()["band_retests"] - b0
            sm.run_moves(35, 42, 1.0, moves, 1)
            same = cm == cnt and all(np.array_equal(x.view(np.uint64), y.view(np.uint64))
                                     for x, y in zip(sm.positions() + sm.orientations(),
                                                     sim_state))
            r["mixed"] = {"value": vm, "mixed_over_fp64": vm / v, "identical_to_fp64": bool(same),
                          "fp64_pairs_per_trial": band / max(sum(a + b for a, b in cm), 1)}
            sm.close()
        if cpu:
            def c3_cpu():
                O.set_threads(all_cores())
                # from the GPU's decorrelated state
                ref = O.ConvexOracle(L, bp_now, bq_now, wl.UNIT_CUBE, seed=42)
                ref.time_steps(0.05, 0.1, CPU_WARM_S)
                moves_, el, steps = ref.time_steps(0.05, 0.1, CPU_TIMED_S)
                return dict(value=moves_ / el, unit="trial moves/s", cores=all_cores(), kind="port",
                            sample=f"{steps} steps of ParallelSweep translate + rotate applies of the same 262,144 "
                                   f"cubes from the GPU's decorrelated state ({el:.1f} s after {CPU_WARM_S:.0f} s "
                                   f"warm-up), C++ restatement of hoomd-rs 1.4.0 with OpenMP over same-colour spaces")
            r["cpu_baseline"] = c3_cpu()
            r["gpu_over_cpu"] = v / r["cpu_baseline"]["value"]
        res["C3_cubes_262144"] = r
    if want("c5"):
        # C5: 256 replicas of hard spheres (FCC, N 2,048 .. 62,500, packing
        # fractions 0.30 .. 0.54), seeds 42 + r, d = 0.08; mixed precision
        # runs persistent applies (one cooperative launch per group of
        # replicas per step), FP64 the per-colour tile passes (one graph per
        # step)
        systems = wl.batched_replicas()
        batch = hmc.GpuBatch([(L, wl.wrap(L, p), None) for L, p in systems], precision="mixed")
        seeds = [42 + r for r in range(len(systems))]
        tune_ms, policy = batch.autotune(0, seeds, 1.0, 0.08, 2)  # tile layout (state left unchanged)
        batch.run(0, seeds, 1.0, 0.08, 2)
        cyc0 = [batch.stats(i)["cycles"] for i in range(len(systems))]
        t0 = time.perf_counter()
        cnt = batch.run(2, seeds, 1.0, 0.08, 40)
        el = time.perf_counter() - t0
        t = sum(a + b for a, b in cnt)
        acc = sum(a for a, _ in cnt)
        v = t / el
        nb = int(sum(len(p) for _, p in systems))
        st0 = batch.stats(0)
        layout = {2: "persistent", 1: "tiles", 0: "lanes"}[st0["tiled"]]
        r = {"value": v, "unit": "trial moves/s", "steps": 40, "precision": "mixed", "bodies": nb,
             "pass_layout": layout, "autotuned_policy_cus": policy}
        if layout == "persistent":
            r["regions"] = {"all": st0["regions"], "threads": st0["region_threads"], "per_cu": st0["regions_per_cu"],
                            "rounds": st0["rounds"],
                            "overflows": int(sum(batch.stats(i)["tile_overflows"] for i in range(len(systems)))),
                            "shapes": sorted({batch.stats(i)["tile_shape"][0] for i in range(len(systems))})}
        final_mixed = [batch.positions(i)[1] for i in range(0, len(systems), 51)]
        # whole-step roofline: the build (bodies read twice, 4 index arrays
        # written and read) and the passes (every colour pass reads every
        # site copy and cell offset once) against the step's wall time
        spaces = 0
        pass_b = 0.0
        for i, (L_, p) in enumerate(systems):
            k = [int(np.floor(x / 1.0)) for x in L_]
            s_i = int(np.prod([x - (x % 2) for x in k]))
            spaces += s_i
            cycles_i = (batch.stats(i)["cycles"] - cyc0[i]) / 40.0  # colour cycles per step
            pass_b += 8 * cycles_i * (4 * s_i + 16 * len(p))
        step_b = (2 * 32 * nb + 4 * 4 * nb + 3 * 4 * spaces) + pass_b + (36 * t + 80 * acc) / 40
        ach = step_b / (el / 40) / 1e9
        r["roofline"] = {"bound": "hbm", "achieved": ach, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": ach / PEAK_HBM_GBS, "traffic": measured_traffic("k_bapply_tile"),
                         "kernel": "whole step (build + passes)",
                         "ms_per_step": el / 40 * 1e3, "algorithmic_bytes_per_step": step_b}
        # SDF of every replica (x_max 0.02, dx 1e-4): one call after the run,
        # then timed calls; beta P / rho from the single frame
        hist, ov = batch.sdf(0.02, 1e-4)
        t0 = time.perf_counter()
        for _ in range(5):
            hist, ov = batch.sdf(0.02, 1e-4)
        ms = (time.perf_counter() - t0) / 5 * 1e3
        z = [hmc.sdf_beta_p_over_rho(hist[i], len(p), 1e-4, 3) for i, (_L, p) in enumerate(systems)]
        r["sdf"] = {"ms_per_call": ms, "bodies_per_s": nb / (ms * 1e-3), "overlaps": int(ov.sum()),
                    "beta_p_over_rho_one_frame": {"phi_0.30": z[0], "phi_0.54": z[-1]}}
        batch.close()
        if fp64:
            # the same batch in plain FP64 (identical decisions)
            batch = hmc.GpuBatch([(L, wl.wrap(L, p), None) for L, p in systems], precision="fp64")
            v64, _ = timed(lambda s0, k: sum(a + b for a, b in batch.run(s0, seeds, 1.0, 0.08, k)), 2, 40)
            r["fp64_value"] = v64
            r["mixed_over_fp64"] = v / v64
            r["identical_to_fp64"] = all(np.array_equal(batch.positions(i)[1].view(np.uint64), x.view(np.uint64))
                                         for i, x in zip(range(0, len(systems), 51), final_mixed))
            batch.close()
            if layout == "persistent":
                # the layout's share: mixed precision on the per-colour tile passes
                batch = hmc.GpuBatch([(L, wl.wrap(L, p), None) for L, p in systems], precision="mixed",
                                     persistent=False)
                vpc, _ = timed(lambda s0, k: sum(a + b for a, b in batch.run(s0, seeds, 1.0, 0.08, k)), 2, 40)
                r["mixed_percolour_value"] = vpc
                r["mixed_over_fp64_same_layout"] = vpc / v64
                batch.close()
        if cpu:
            # every 16th replica, each for a share of the sample, one after
            # another with all cores (the replicas are independent)
            O.set_threads(all_cores())
            moves = elapsed = 0.0
            picks = list(range(0, len(systems), 16))
            for i in picks:
                L_, p = systems[i]
                ref = O.OracleSystem(L_, wl.wrap(L_, p), seed=42 + i)
                ref.time_applies(True, 1.0, 0.08, CPU_WARM_S / len(picks))
                m_, e_, _n = ref.time_applies(True, 1.0, 0.08, CPU_TIMED_S / len(picks))
                moves += m_
                elapsed += e_
                del ref
            r["cpu_baseline"] = dict(value=moves / elapsed, unit="trial moves/s", cores=all_cores(), kind="port",
                                     sample=f"{len(picks)} of the 256 replicas (every 16th), ParallelSweep applies "
                                            f"of each in turn ({elapsed:.1f} s), C++ restatement of hoomd-rs 1.4.0 "
                                            f"with OpenMP over same-colour spaces")
            r["gpu_over_cpu"] = v / r["cpu_baseline"]["value"]
        res["C5_replicas_256"] = r
    if want("gsd"):
        res["gsd_frames_1M"] = gsd_frames(hmc, wl, cpu)
    if want("tune"):
        res["tune_local_c2"] = tune_bench(hmc, wl, cpu)
    if want("placement"):
        res["placement_quick_compress_1M"] = placement(hmc, wl)
        if cpu:
            # the overlap-penalty sweeps of the placement (FCC 4*63^3 at packing
            # fraction 0.45, d 0.1) on the same host cores
            Lp, pp = wl.fcc_spheres(packing_fraction=0.45)
            mk = lambda: O.OracleSystem(Lp, wl.wrap(Lp, pp), model="overlap_penalty", params=(1e4, 0.2, 100.0, 1.0),
                                        seed=42)
            pc = cpu_sample(mk, 1.0, 0.1, all_cores(), True, "the same 1,000,188 spheres under the overlap penalty")
            pl = res["placement_quick_compress_1M"]
            pl["cpu_baseline_sweeps"] = pc
            pl["sweeps_gpu_over_cpu"] = pl["overlap_penalty_sweep_trial_moves_per_s"] / pc["value"]
    return res


def tune_bench(hmc, wl, cpu=True):
    """ParallelSweep::tune with Translate and TuneOptions::default() (target
    0.2, 8,000 samples, 32 rounds) on the C2 system (N=1,000,188 hard
    spheres at packing fraction 0.50), from d = 0.08: the GPU's time and
    the oracle's serial time for the same rounds, and whether the tuned
    sizes agree bit for bit."""
    L, pos = wl.fcc_spheres((63, 63, 63), packing_fraction=0.50)
    body = wl.wrap(L, pos)
    sim = hmc.GpuParallelSweep(L, body)
    sim.tune_translate(0, 0, 42, 1.0, 0.2, 8000, 2, 0.08)  # warm-up
    t0 = time.perf_counter()
    m, cnt, amb = sim.tune_translate(0, 0, 42, 1.0, 0.2, 8000, 32, 0.08)
    el = time.perf_counter() - t0
    sim.close()
    r = {"value": 32 * 8000 / el, "unit": "tuning samples/s", "seconds": el, "maximum_distance": [0.08, m],
         "last_round": cnt[-1], "ambiguous": amb}
    if cpu:
        import _oracle as O
        O.set_threads(1)
        ref = O.OracleSystem(L, body, seed=42)
        t0 = time.perf_counter()
        m2, c2 = ref.tune_translate(1.0, 0.2, 8000, 32, 0.08)
        el2 = time.perf_counter() - t0
        r["cpu_baseline"] = dict(value=32 * 8000 / el2, unit="tuning samples/s", cores=1, kind="port",
                                 sample="the same 32 rounds of 8,000 samples, serial as tune_local_trial is "
                                        f"({el2:.2f} s), C++ restatement of hoomd-rs 1.4.0")
        r["identical_to_cpu"] = bool(m2 == m and c2 == cnt)
    return r


def gsd_frames(hmc, wl, cpu=True):
    """HOOMD GSD frames of the C2 system (N=1,000,188) written from the
    device state (hmc_gsd_append_frame: sites packed to FP32 on the device,
    copied back, appended with the index): frames/s and MB/s of file bytes,
    beside the reference's path timed on the host (the FP64 sites converted
    to FP32 with numpy and written with the same file layer)."""
    import tempfile
    L, pos = wl.fcc_spheres()
    sim = hmc.GpuParallelSweep(L, wl.wrap(L, pos), precision="mixed")
    sim.run(0, 42, 1.0, 0.08, 4)
    out = {}
    with tempfile.TemporaryDirectory() as d:
        f = hmc.GsdFile.create(f"{d}/gpu.gsd")
        f.append_frame(sim, 0)
        n = 10
        t0 = time.perf_counter()
        for k in range(n):
            f.append_frame(sim, k + 1)
        el = time.perf_counter() - t0
        f.close()
        frame_bytes = 12 * len(pos) + 8 + 24 + 1 + 4
        out = {"bodies": len(pos), "frames": n, "ms_per_frame": el / n * 1e3, "frames_per_s": n / el,
               "file_MB_per_s": frame_bytes * n / el / 1e6}
        if cpu:
            sites = sim.positions()[1]
            g = hmc.GsdFile.create(f"{d}/cpu.gsd")
            t0 = time.perf_counter()
            for k in range(n):
                g.write_chunk("configuration/step", np.array([k], np.uint64))
                g.write_chunk("configuration/box", np.array([L[0], L[1], L[2], 0, 0, 0], np.float32))
                g.write_chunk("configuration/dimensions", np.array([3], np.uint8))
                g.write_chunk("particles/N", np.array([len(pos)], np.uint32))
                g.write_chunk("particles/position", sites.astype(np.float32))
                g.end_frame()
            el_c = time.perf_counter() - t0
            g.close()
            out["cpu_baseline"] = dict(value=n / el_c, unit="frames/s", cores=1, kind="port",
                                       sample=f"{n} frames of the host's FP64 sites converted with numpy and "
                                              f"written by the same file layer ({el_c:.2f} s)")
    sim.close()
    return out


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
    if args.only:
        print(json.dumps(other_workloads(hmc, wl, cpu=not args.no_cpu_baseline, only=args.only,
                                         fp64=not args.no_fp64)), flush=True)
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
        extra = {}
        if mixed["stats"]["tiled"] == 2:
            retests = mixed["stats"]["band_retests"] / max(mixed["stats"]["applies"], 1)
            b = model_apply_bytes(n, n_spaces, cycles, mixed["accepted"] / args.steps, retests)
            bp = apply_bytes(n, n_spaces, trials / args.steps, mixed["accepted"] / args.steps)
            extra = {"persistent_model_bytes_per_launch": bp,
                     "persistent_model_frac": bp / (avg_ms * 1e-3) / 1e9 / PEAK_HBM_GBS}
            kernel = "k_apply_tile<3>"
        else:
            b = pass_bytes("mixed", n, n_spaces, nonempty, acc_per_pass)
            kernel = "k_pass_tile<3,MIXED>" if mixed["stats"]["tiled"] else "k_pass<3,MIXED>"
        achieved = b / (avg_ms * 1e-3) / 1e9
        out["roofline"] = {"bound": "hbm", "achieved": achieved, "peak": PEAK_HBM_GBS, "unit": "GB/s",
                           "frac": achieved / PEAK_HBM_GBS, "traffic": measured_traffic(kernel),
                           "kernel": kernel, "avg_launch_ms": avg_ms,
                           "algorithmic_bytes_per_launch": b, "launches": mixed["pass_n"],
                           "build_ms_per_apply": mixed["build_ms"] / max(mixed["build_n"], 1), **extra}
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
        # the layout's share and the precision's share of that ratio: mixed
        # precision in the same per-colour layout as the FP64 run
        if args.pass_layout == "persistent":
            pc = run_gpu(hmc, L, body, site, "mixed", args.steps // 4, args.warmup, args.seed, kT, d,
                         args.profile_period, persistent=False)
            vpc = (pc["accepted"] + pc["rejected"]) / pc["elapsed"]
            out["mixed_percolour"] = {"value": vpc, "ms_per_step": 1e3 * pc["elapsed"] / (args.steps // 4),
                                      "avg_pass_ms": pc["pass_ms"] / max(pc["pass_n"], 1),
                                      "steps": args.steps // 4}
            out["mixed_over_fp64_same_layout"] = vpc / v64
            out["persistent_over_percolour_mixed"] = value / vpc
    if not args.no_cpu_baseline:
        import _oracle as O
        mk = lambda: O.OracleSystem(L, pos, seed=args.seed)
        what = "the same N=1,000,188 system"
        out["cpu_baseline"] = cpu_sample(mk, kT, d, all_cores(), True, what)
        out["gpu_over_cpu"] = value / out["cpu_baseline"]["value"]
        out["cpu_baseline_1core"] = cpu_sample(mk, kT, d, 1, False, what)
    if not args.no_other:
        out["other_workloads"] = other_workloads(hmc, wl, cpu=not args.no_cpu_baseline)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
