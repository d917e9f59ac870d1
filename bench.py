"""Benchmark: bundle-adjustment time-to-converge on the 1,000-camera scene,
mixed precision vs FP64 plain, on one MI355X; and the 256-window batch.

A step is one full Levenberg-Marquardt solve to convergence (the reference's
stopping rule, crates/sfm-ba/src/lib.rs:1188-1221) of the 1,000-PINHOLE-camera
/ 500k-point / ~3M-observation scene (1% outliers, Huber 2 px), starting
from the same perturbed state, with the problem already resident in HBM
(created once, untimed).

Prints one JSON line.  ``value`` is the mixed-precision solve's time to
converge by that rule, in seconds (lower is better), the mean over the
timed steps.  Beside it: the FP64 plain solve timed the same way, the
speed-up, per-try damping traces of both, Jacobian and Schur-assembly
throughput (observations/s), residual statistics (RMS from the residuals,
mean |r|, inlier RMS) and pose error of both forms, the roofline of the
dominant kernel (the factorization's trailing update, timed per launch
with HIP events on its stream) with the whole factorization, the Schur
assembly and the Jacobian stage beside it, a CPU baseline from the oracle
(all host cores), the 256-window batched solve (windows/s) and the
5,000-camera scene solved by PCG and by Cholesky.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--scene c2] [--no-cpu-baseline] [--no-windows]
                  [--no-large] [--copy-calibration]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM = 8.0e12          # B/s, MI355X HBM3E
PEAK_FP64_MFMA = 78.6e12   # FLOP/s, MI355X dense FP64 matrix


def log(*a):
    print(*a, file=sys.stderr, flush=True)


TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_syrk_persist_hbm_traffic.txt")
ASSEMBLY_TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "assembly_hbm_traffic.txt")


def measured_traffic(path=TRAFFIC_PROFILE, key="traffic_bytes_per_launch"):
    """HBM bytes per launch of this workload from a committed counter-run
    summary (FETCH_SIZE + WRITE_SIZE per dispatch, each scaled by the same
    counter's reading for a 1 GiB copy): k_syrk_persist by default, the
    assembly kernels by key; None when the summary is absent."""
    try:
        with open(path) as f:
            for line in f:
                if line.startswith(key + ":"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def traffic_source(path=TRAFFIC_PROFILE):
    """Where a roofline's `traffic` comes from: the committed counter-run
    summary (not this run) and the commit it was measured at."""
    commit = None
    try:
        with open(path) as f:
            for line in f:
                if line.startswith("measured_at_commit:"):
                    commit = line.split(":", 1)[1].strip()
    except OSError:
        return None
    rel = os.path.relpath(path, ROOT)
    return (f"{rel}: FETCH_SIZE and WRITE_SIZE counter runs of this bench command (separate rocprofv3 runs, "
            f"not this run)" + (f", measured at commit {commit}" if commit else ""))


def copy_calibration():
    """Three device-to-device copies of 1 GiB (hipMemcpy: the runtime's copy
    kernel reads 1 GiB and writes 1 GiB each): the plain copy of known size
    that a FETCH_SIZE / WRITE_SIZE counter run divides by."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so")
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    hip.hipFree.argtypes = [C.c_void_p]
    n = 1 << 30
    src, dst = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(src), n) or hip.hipMalloc(C.byref(dst), n):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(src, 1, n):
            raise RuntimeError("hipMemset failed")
        for _ in range(3):
            if hip.hipMemcpy(dst, src, n, 3):  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        if hip.hipDeviceSynchronize():
            raise RuntimeError("hipDeviceSynchronize failed")
    finally:
        hip.hipFree(src)
        hip.hipFree(dst)
    log("copy calibration: 3 x 1 GiB device-to-device copies done")


def quat_to_R(q):
    """Rotation matrices of poses given as rows (qw, qx, qy, qz, tx, ty, tz)."""
    w, x, y, z = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    n = np.sqrt(w * w + x * x + y * y + z * z)
    w, x, y, z = w / n, x / n, y / n, z / n
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)], -1),
        np.stack([2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)], -1),
        np.stack([2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)], -1)], -2)


def pose_error(poses, poses_true):
    """Camera centres aligned to ground truth by a similarity (Umeyama; scale
    is a free gauge): RMS centre error and mean rotation error (deg)."""
    def centres(p):
        p = np.asarray(p, dtype=np.float64).reshape(-1, 7)
        R = quat_to_R(p)
        t = p[:, 4:7]
        return R, -np.einsum("nji,nj->ni", R, t)
    R_e, C_e = centres(poses)
    R_t, C_t = centres(poses_true)
    me, mt = C_e.mean(0), C_t.mean(0)
    Xe, Xt = C_e - me, C_t - mt
    U, S, Vt = np.linalg.svd(Xt.T @ Xe / len(Xe))
    Dg = np.diag([1.0, 1.0, np.sign(np.linalg.det(U @ Vt))])
    R0 = U @ Dg @ Vt
    s = np.trace(np.diag(S) @ Dg) / (Xe ** 2).sum(1).mean()
    C_al = s * (R0 @ Xe.T).T + mt
    rel = np.einsum("nji,njk->nik", R_t, R_e @ R0.T)
    ang = np.degrees(np.arccos(np.clip((np.trace(rel, axis1=1, axis2=2) - 1) / 2, -1, 1)))
    return {"centre_rms": float(np.sqrt(((C_al - C_t) ** 2).sum(1).mean())),
            "rotation_mean_deg": float(ang.mean()), "scale": float(s)}


def residual_stats(ba, inp, out, inlier_px=10.0):
    """Plain residual norms at the solved state: RMS, mean (the reference's
    mean_reprojection_error) and the RMS over |r| < inlier_px."""
    st = ba.BaInput(inp.camera_of_image, out.cameras, out.poses, out.points, inp.observations)
    j = ba.observation_jacobians(st)
    n = np.linalg.norm(j["r"], axis=1)[j["valid"]]
    inl = n[n < inlier_px]
    return {"rms_px": float(np.sqrt((n ** 2).mean())), "mean_px": float(n.mean()),
            "inlier_rms_px": float(np.sqrt((inl ** 2).mean())), "inlier_fraction": float(len(inl) / len(n))}


def windows_section(ba, scenes, threads, cpu=True):
    """Config 3: 256 independent local-BA windows (10-40 SIMPLE_RADIAL
    cameras, 2k-20k points each) solved by the persistent batched launch:
    one-off batched calls (packing and upload included), a batch plan whose
    launch shape the autotuner picks (3-iteration timing runs) and whose
    solve is a captured HIP graph replayed from the resident initial state,
    next to a sequential GPU loop (16-window sample) and the oracle on the
    host (4-window sample)."""
    wins = scenes.window_scenes(256)
    inps = [w.inp for w in wins]
    p = ba.BaParams(max_iterations=50)
    res = {"count": len(inps), "observations": int(sum(len(i.observations) for i in inps))}
    for prec in ("fp64", "mixed"):
        ba.bundle_adjust_batch(inps[:8], p, precision=prec)
        t = time.perf_counter()
        outs = ba.bundle_adjust_batch(inps, p, precision=prec)
        wall = time.perf_counter() - t
        n = 16
        ba.bundle_adjust(inps[0], p, precision=prec)
        t = time.perf_counter()
        seq = [ba.bundle_adjust(i, p, precision=prec) for i in inps[:n]]
        seq_s = time.perf_counter() - t
        t = time.perf_counter()
        plan = ba.BatchPlan(inps, p, precision=prec, autotune_iterations=3)
        plan_create_s = time.perf_counter() - t
        info = plan.info()
        plan.solve()
        reps = 3
        t = time.perf_counter()
        for _ in range(reps):
            pouts = plan.solve()
        graph_s = (time.perf_counter() - t) / reps
        plan.close()
        res[prec] = {"graph_windows_per_s": len(inps) / graph_s, "graph_solve_s": graph_s,
                     "plan_create_s": plan_create_s,
                     "launch_shape": {"threads_per_window": info["threads_per_window"],
                                      "pairs_in_flight": info["pairs_in_flight"]},
                     "autotune_seconds_per_shape": {f"{x['threads']}x{x['pairs']}": x["seconds"]
                                                    for x in info["shapes"]},
                     "graph_vs_batch_max_rel_cost_diff": float(max(abs(a.final_cost - b.final_cost) / b.final_cost
                                                                   for a, b in zip(pouts, outs))),
                     "batch_windows_per_s": len(inps) / wall,
                     "batch_kernel_windows_per_s": len(inps) / outs[0].stats["time_total_s"],
                     "sequential_gpu_windows_per_s": n / seq_s,
                     "batch_wall_s": wall, "batch_kernel_s": outs[0].stats["time_total_s"],
                     "mean_iterations": float(np.mean([o.iterations_run for o in outs])),
                     "max_rel_cost_diff_vs_sequential": float(max(abs(a.final_cost - b.final_cost) / b.final_cost
                                                                  for a, b in zip(outs[:n], seq)))}
        log(f"windows {prec}: graph {res[prec]['graph_windows_per_s']:.1f}/s "
            f"(shape {info['threads_per_window']}x{info['pairs_in_flight']}), "
            f"batch {res[prec]['batch_windows_per_s']:.1f}/s, "
            f"sequential {res[prec]['sequential_gpu_windows_per_s']:.1f}/s")
    if cpu:
        import oracle
        k = 4
        t = time.perf_counter()
        for i in inps[:k]:
            oracle.bundle_adjust(i, p, nthreads=threads)
        res["cpu_oracle_windows_per_s"] = k / (time.perf_counter() - t)
        res["cpu_oracle_sample"] = f"first {k} windows, {threads} threads"
        log(f"windows cpu oracle {res['cpu_oracle_windows_per_s']:.2f}/s")
    return res


# large_scene keys: solver name, then the assembly precision
LARGE_KEYS = {"pcg": "pcg", "pcg_mixed": "pcg_fp32_matrix", "cholesky": "cholesky"}


def large_section(ba, scenes):
    """Config 4: 5,000 RADIAL cameras, 2M points (D ~ 45k): the reduced
    system solved by block-Jacobi PCG (FP64 matrix, and FP32-stored with
    FP64 residual replacement) vs the FP64 MFMA Cholesky, with FP64 and
    mixed assembly, one solve each to the reference's stopping rule."""
    sc = scenes.generate("c4")
    t = time.perf_counter()
    prob = ba.Problem(sc.inp)
    res = {"observations": len(sc.inp.observations), "problem_create_s": time.perf_counter() - t}
    p = ba.BaParams(max_iterations=50)
    for solver in ("pcg", "pcg_mixed", "cholesky"):
        for prec in ("fp64", "mixed"):
            t = time.perf_counter()
            out = prob.solve(p, precision=prec, solver=solver, profile=True)
            wall = time.perf_counter() - t
            st = out.stats
            r = {"time_to_converge_s": wall, "iterations": out.iterations_run, "tries": st["num_factorizations"],
                 "final_cost": out.final_cost, "pcg_iterations": st["num_pcg_iterations"],
                 "schur_assembly_ms": per_unit(st, "time_schur_s", "count_schur") * 1e3}
            if solver == "cholesky":
                r["factor_ms"] = per_unit(st, "time_factor_s", "count_factor") * 1e3
                r["trailing_update_tflops"] = st["flop_trailing_update"] / st["time_trailing_update_s"] / 1e12
            else:
                r["solve_ms_per_try"] = st["time_solve_s"] / max(1, st["num_factorizations"]) * 1e3
                r["solve_ms_per_pcg_iteration"] = st["time_solve_s"] / max(1, st["num_pcg_iterations"]) * 1e3
            res[f"{LARGE_KEYS[solver]}_{prec}"] = r
            res["dim"] = st["dim"]
            log(f"large scene {solver} {prec}: {wall:.2f}s, {out.iterations_run} iterations, cost {out.final_cost:.6f}")
    D = res["dim"]
    f = res["cholesky_fp64"]["factor_ms"] * 1e-3
    res["roofline_factorization"] = {"bound": "mfma", "achieved": (D ** 3 / 3.0 + 2.0 * D ** 2) / f / 1e12,
                                     "peak": PEAK_FP64_MFMA / 1e12, "unit": "TFLOP/s",
                                     "frac": (D ** 3 / 3.0 + 2.0 * D ** 2) / f / PEAK_FP64_MFMA}
    res["pcg_speedup_vs_cholesky_fp64"] = res["cholesky_fp64"]["time_to_converge_s"] / res["pcg_fp64"]["time_to_converge_s"]
    res["pcg_fp32_matrix_speedup_vs_pcg_fp64"] = (res["pcg_fp64"]["time_to_converge_s"]
                                                  / res["pcg_fp32_matrix_fp64"]["time_to_converge_s"])
    prob.close()
    return res


def trace(st):
    """Per-try damping and outcome of a solve: A accepted, R rejected, N no
    step; suffix 3 for a reduced system assembled in FP32, L for the LU
    fallback."""
    return {"lambdas": [float(x) for x in st["try_lambdas"]],
            "outcomes": " ".join("ARN?"[o & 3] + ("L" if o & 4 else "") + ("3" if o & 8 else "")
                                 for o in (int(x) for x in st["try_outcomes"])),
            "fp32_tries": int(st["num_fp32_tries"])}


def per_unit(stats, key, count_key):
    n = stats[count_key]
    return stats[key] / n if n else float("nan")


def host_info():
    """Host cores this process may use, CPU model and memory."""
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    mem = None
    try:
        with open("/proc/meminfo") as f:
            for line in f:
                if line.startswith("MemTotal:"):
                    mem = int(line.split()[1]) * 1024
                    break
    except OSError:
        pass
    return {"nproc": cores, "os_cpu_count": os.cpu_count(), "cpu_model": model, "mem_total_bytes": mem}


def cpu_baseline(inp, params, fp64_stats, threads):
    """Oracle (the C++ FP64 restatement of the reference, threaded as the
    reference is: parallel Jacobians, serial accumulation, per-thread D x D
    Schur accumulators) timed on this host at full size: one outer
    iteration's stages on the full scene with `threads` threads, and one
    single-threaded Cholesky factorization + solve of the reference-layout
    reduced system's size D (nalgebra's Cholesky::new is single-threaded,
    lib.rs:1081).  Combined with the GPU FP64 solve's linearization and try
    counts (the same trajectory) into the CPU time-to-converge."""
    import oracle

    t = oracle.stage_times(inp, params, nthreads=threads)
    D = int(t["dim"])
    t0 = time.perf_counter()
    t_chol = oracle.cholesky_seconds(D)
    chol_wall = time.perf_counter() - t0
    n_lin = fp64_stats["num_linearizations"]
    n_try = fp64_stats["num_factorizations"]
    est = t["cost"] + n_lin * (t["jacobians"] + t["accumulate"]) + n_try * (t["schur"] + t_chol + t["cost"])
    sample = (f"oracle on the full scene, timed on this host: one outer iteration's stages with {threads} "
              f"threads (Jacobians {t['jacobians']:.2f}s, serial accumulation {t['accumulate']:.2f}s, "
              f"Schur assembly {t['schur']:.2f}s, cost {t['cost']:.2f}s) and one single-threaded D={D} "
              f"Cholesky + solve ({t_chol:.1f}s); time-to-converge = cost + {n_lin} x (Jacobians + accumulation) "
              f"+ {n_try} x (Schur + Cholesky + cost), the GPU FP64 solve's linearization and try counts")
    return {"value": est, "unit": "s", "cores": threads, "kind": "port", "sample": sample,
            "stage_seconds": {k: float(v) for k, v in t.items() if k != "dim"},
            "cholesky_seconds": t_chol, "cholesky_dim": D, "timed_seconds": sum(float(v) for k, v in t.items()
                                                                                if k != "dim") + chol_wall,
            "host": host_info()}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--scene", default="c2")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-windows", action="store_true", help="skip the 256-window batched section")
    ap.add_argument("--no-large", action="store_true", help="skip the 5,000-camera PCG vs Cholesky section")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="end with three 1 GiB device-to-device copies (the known-size kernel that scales "
                         "FETCH_SIZE / WRITE_SIZE in a counter run)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (the engine has no multi-GPU path); got --gpus {args.gpus}",
              file=sys.stderr)
        return 2

    from sfmtory_amd import ba, scenes

    if ba.device_count() < 1:
        print("no HIP device", file=sys.stderr)
        return 1
    t0 = time.time()
    sc = scenes.generate(args.scene)
    inp = sc.inp
    n_obs, n_pts = len(inp.observations), len(inp.points)
    log(f"scene {args.scene}: {len(inp.poses)} images, {n_pts} points, {n_obs} observations "
        f"({time.time() - t0:.1f}s)")
    params = ba.BaParams(max_iterations=50)
    t0 = time.time()
    prob = ba.Problem(inp)
    create_s = time.time() - t0
    log(f"problem resident on device in {create_s:.2f}s")

    def timed(precision, k, w):
        for i in range(w):
            prob.solve(params, precision=precision)
            log(f"warmup {precision} {i + 1}/{w} done")
        walls, outs = [], []
        for i in range(k):
            t = time.perf_counter()
            out = prob.solve(params, precision=precision)
            walls.append(time.perf_counter() - t)
            outs.append(out)
            log(f"{precision} step {i + 1}/{k}: {walls[-1]:.3f}s, {out.iterations_run} iterations, "
                f"cost {out.final_cost:.6g}")
        return walls, outs

    mixed_walls, mixed_outs = timed("mixed", args.steps, args.warmup)
    fp64_walls, fp64_outs = timed("fp64", args.steps, min(args.warmup, 1))
    # Profiled solves (HIP events on the solve stream) for per-phase times.
    prof_mixed = prob.solve(params, precision="mixed", profile=True).stats
    prof_fp64 = prob.solve(params, precision="fp64", profile=True).stats
    log("profiled solves done")
    # The same scene with the reduced system solved by PCG instead of the
    # factorization (FP64 assembly; FP64 matrix, then the FP32 copy with FP64
    # residual replacement): one untimed solve each, then one timed.
    solvers = {}
    for solver in ("pcg", "pcg_mixed"):
        prob.solve(params, precision="fp64", solver=solver)
        t = time.perf_counter()
        out = prob.solve(params, precision="fp64", solver=solver)
        wall = time.perf_counter() - t
        st = out.stats
        solvers[LARGE_KEYS[solver] + "_fp64"] = {
            "time_to_converge_s": wall, "iterations": out.iterations_run, "tries": st["num_factorizations"],
            "pcg_iterations": st["num_pcg_iterations"], "final_cost": out.final_cost}
        log(f"{solver} fp64: {wall:.3f}s, {out.iterations_run} iterations, {st['num_pcg_iterations']} PCG "
            f"iterations, cost {out.final_cost:.6f}")

    mixed_t = float(np.mean(mixed_walls))
    fp64_t = float(np.mean(fp64_walls))
    q_fp64 = residual_stats(ba, inp, fp64_outs[-1])
    rms_fp64 = q_fp64["rms_px"]
    D = prof_fp64["dim"]
    q_mixed = residual_stats(ba, inp, mixed_outs[-1])
    rms_mixed = q_mixed["rms_px"]
    pe = {"mixed": pose_error(mixed_outs[-1].poses, sc.poses_true),
          "fp64": pose_error(fp64_outs[-1].poses, sc.poses_true),
          "initial": pose_error(inp.poses, sc.poses_true)}

    def rates(st):
        t_lin = per_unit(st, "time_linearize_s", "count_linearize")
        t_schur = per_unit(st, "time_schur_s", "count_schur")
        return {"jac_obs_per_s": n_obs / t_lin, "schur_obs_per_s": n_obs / (t_lin + t_schur),
                "linearize_ms": t_lin * 1e3, "schur_assembly_ms": t_schur * 1e3,
                "factor_ms": per_unit(st, "time_factor_s", "count_factor") * 1e3,
                "solve_ms": st["time_solve_s"] / max(1, st["count_factor"]) * 1e3,
                "update_cost_ms": st["time_update_s"] / max(1, st["count_factor"]) * 1e3,
                "iterations": st["num_linearizations"], "tries": st["num_factorizations"],
                "cholesky_failures": st["num_cholesky_failures"]}

    r_mixed, r_fp64 = rates(prof_mixed), rates(prof_fp64)
    # the mixed solve's FP32 stages alone (its later tries assemble in FP64)
    r_mixed["fp32_linearize_ms"] = per_unit(prof_mixed, "time_linearize_fp32_s", "count_linearize_fp32") * 1e3
    r_mixed["fp32_schur_assembly_ms"] = per_unit(prof_mixed, "time_schur_fp32_s", "count_schur_fp32") * 1e3
    r_mixed["fp32_jac_obs_per_s"] = n_obs / (r_mixed["fp32_linearize_ms"] * 1e-3)
    r_mixed["fp32_tries"] = prof_mixed["num_fp32_tries"]
    # Dominant kernel: the factorization's trailing update (k_syrk_persist),
    # timed per launch with HIP events on the solve stream; its algorithmic
    # FLOP per launch (lower-triangle update + rhs row + next diagonal block)
    # come from the library beside the times.
    n_tu = prof_fp64["count_trailing_update"]
    t_tu = prof_fp64["time_trailing_update_s"] / max(1, n_tu)
    f_tu = prof_fp64["flop_trailing_update"] / max(1, n_tu)
    roof = {"bound": "mfma", "kernel": "k_syrk_persist (FP64 trailing update + lookahead diagonal factorization)",
            "achieved": f_tu / t_tu / 1e12, "peak": PEAK_FP64_MFMA / 1e12, "unit": "TFLOP/s",
            "frac": f_tu / t_tu / PEAK_FP64_MFMA, "traffic": measured_traffic(),
            "traffic_source": traffic_source(),
            "flop_per_launch": f_tu, "avg_launch_us": t_tu * 1e6, "launches": n_tu}
    # The whole factorization (D^3/3 + 2 D^2 FP64 FLOP) as a kernel group.
    flop = D ** 3 / 3.0 + 2.0 * D ** 2
    t_fact = r_fp64["factor_ms"] * 1e-3
    roof_fact = {"bound": "mfma", "kernel": "dense FP64 Cholesky, all launches (D=%d)" % D,
                 "achieved": flop / t_fact / 1e12, "peak": PEAK_FP64_MFMA / 1e12, "unit": "TFLOP/s",
                 "frac": flop / t_fact / PEAK_FP64_MFMA}
    # Schur assembly, mixed: 12 N_obs + 88 N_pts + 4 D(D+1) + 8 D algorithmic bytes per try.
    bytes_schur = 12 * n_obs + 88 * n_pts + 4 * D * (D + 1) + 8 * D
    t_as = r_mixed["fp32_schur_assembly_ms"] * 1e-3
    roof_schur = {"bound": "hbm", "kernel": "Schur assembly (FP32 tries of the mixed solve)",
                  "achieved": bytes_schur / t_as / 1e9,
                  "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": bytes_schur / t_as / PEAK_HBM,
                  "traffic": measured_traffic(ASSEMBLY_TRAFFIC_PROFILE, "schur_traffic_bytes"),
                  "traffic_source": traffic_source(ASSEMBLY_TRAFFIC_PROFILE),
                  "algorithmic_bytes": bytes_schur}
    if roof_schur["traffic"]:
        roof_schur["measured_GB_per_s"] = roof_schur["traffic"] / t_as / 1e9
    # Jacobian stage, mixed: 112 + 8 k_f algorithmic bytes per observation
    # (k_f = 2 free intrinsics per PINHOLE camera).
    bytes_lin = (112 + 8 * 2) * n_obs
    t_lin = r_mixed["fp32_linearize_ms"] * 1e-3
    roof_lin = {"bound": "hbm", "kernel": "Jacobian stage: k_linearize_staged + k_point_accumulate (FP32 linearizations of the "
                "mixed solve)", "achieved": bytes_lin / t_lin / 1e9,
                "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": bytes_lin / t_lin / PEAK_HBM,
                "traffic": measured_traffic(ASSEMBLY_TRAFFIC_PROFILE, "linearize_traffic_bytes"),
                "traffic_source": traffic_source(ASSEMBLY_TRAFFIC_PROFILE),
                "algorithmic_bytes": bytes_lin}
    if roof_lin["traffic"]:
        roof_lin["measured_GB_per_s"] = roof_lin["traffic"] / t_lin / 1e9

    # The same stages in FP64 (the FP64 plain solve, and the mixed solve's
    # tries below the FP32 damping floor): 20 N_obs + 100 N_pts + 4 D(D+1)
    # + 8 D bytes per assembly, 216 + 16 k_f bytes per observation.
    bytes_schur64 = 20 * n_obs + 100 * n_pts + 4 * D * (D + 1) + 8 * D
    t_as64 = r_fp64["schur_assembly_ms"] * 1e-3
    roof_schur64 = {"bound": "hbm", "kernel": "Schur assembly (FP64)", "achieved": bytes_schur64 / t_as64 / 1e9,
                    "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": bytes_schur64 / t_as64 / PEAK_HBM,
                    "algorithmic_bytes": bytes_schur64}
    bytes_lin64 = (216 + 16 * 2) * n_obs
    t_lin64 = r_fp64["linearize_ms"] * 1e-3
    roof_lin64 = {"bound": "hbm", "kernel": "Jacobian stage (FP64)", "achieved": bytes_lin64 / t_lin64 / 1e9,
                  "peak": PEAK_HBM / 1e9, "unit": "GB/s", "frac": bytes_lin64 / t_lin64 / PEAK_HBM,
                  "algorithmic_bytes": bytes_lin64}

    windows = None
    if not args.no_windows:
        windows = windows_section(ba, scenes, host_info()["nproc"], cpu=not args.no_cpu_baseline)

    large = None
    if not args.no_large:
        prob.close()
        large = large_section(ba, scenes)

    cpu = None
    if not args.no_cpu_baseline:
        threads = host_info()["nproc"]
        log(f"cpu baseline: oracle stages with {threads} threads and one full-size Cholesky ...")
        cpu = cpu_baseline(inp, params, prof_fp64, threads)
        log(f"cpu baseline {cpu['value']:.1f}s (Cholesky D={cpu['cholesky_dim']}: {cpu['cholesky_seconds']:.1f}s)")

    # value: the mixed form's time to converge by the reference's stopping
    # rule (mean of the timed steps); the FP64 plain solve beside it.
    value = mixed_t
    line = {
        "metric": "BA time-to-converge (s) + Jac/Schur obs/s, mixed vs FP64 plain, 1000-cam scene",
        "value": value,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": value * 1e3,
        "higher_is_better": False,
        "scaling": "strong",
        "vs_baseline": None,
        "dtype": "f32 assembly while lambda >= 1e-5, then f64; f64 point inverses, cost and Cholesky (mixed); "
                 "f64 (plain)",
        "data": "synthetic (seeded scene generator)",
        "config": {"workload": "1000 PINHOLE cameras, 500k points, %d observations (mean track 6), 1%% outliers, "
                               "Huber 2px, pose 0 fixed, principal points pinned, D=%d" % (n_obs, D),
                   "scene": args.scene, "max_iterations": 50},
        "value_definition": "mixed-precision solve, device-resident problem, run to the reference's stopping "
                            "rule (relative cost decrease < 1e-6, no accepted step in 8 tries, or 50 "
                            "iterations); wall-clock mean over the timed steps",
        "fp64_time_to_converge_s": fp64_t,
        "mixed_speedup_vs_fp64": fp64_t / value,
        "rms_criterion_met": abs(rms_mixed - rms_fp64) <= 1e-4,
        "step_walls_s": {"mixed": mixed_walls, "fp64": fp64_walls},
        "traces": {"mixed": trace(prof_mixed), "fp64": trace(prof_fp64)},
        "iterations": {"mixed": mixed_outs[-1].iterations_run, "fp64": fp64_outs[-1].iterations_run},
        "device_resident_solve_s": {"mixed": float(np.mean([o.stats["time_total_s"] for o in mixed_outs])),
                                    "fp64": float(np.mean([o.stats["time_total_s"] for o in fp64_outs]))},
        "problem_create_s": create_s,
        "rms_px": {"mixed": rms_mixed, "fp64": rms_fp64, "diff": abs(rms_mixed - rms_fp64)},
        "residuals": {"mixed": q_mixed, "fp64": q_fp64},
        "pose_error": pe,
        "final_cost": {"mixed": mixed_outs[-1].final_cost, "fp64": fp64_outs[-1].final_cost},
        "mixed": r_mixed,
        "fp64": r_fp64,
        "solvers": solvers,
        "roofline": roof,
        "roofline_factorization": roof_fact,
        "roofline_schur": roof_schur,
        "roofline_linearize": roof_lin,
        "roofline_schur_fp64": roof_schur64,
        "roofline_linearize_fp64": roof_lin64,
        "cpu_baseline": cpu,
        "windows": windows,
        "large_scene": large,
    }
    print(json.dumps(line), flush=True)
    if args.copy_calibration:
        copy_calibration()
    return 0


if __name__ == "__main__":
    sys.exit(main())
