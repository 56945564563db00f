#!/usr/bin/env python3
"""Benchmark: GN semi-Lagrangian Hessian matvecs per second at 256³ (FP32).

Workload (one "step" = one CLAIRE::HessianMatVec, src/CLAIRE.cpp:1311-1379,
as timed by the reference harness apps/benchmark.cpp:234-283):
  grid 256³ periodic on [0,2π)³, nt = 4, GN, semi-Lagrangian, nc = 1;
  velocity band-limited (|w|∞ <= 8, N(0,1) coefficients decayed by
  1/(1+|w|²)) scaled to CFL 2; template = 24 seeded periodic Gaussian blobs;
  H1SN regularization β = 1e-2 (the harness default, src/RegOpt.cpp:1071-1073);
  incremental velocity ṽ = v as in the reference harness
  (SetIncControlVariable(v), apps/benchmark.cpp:252).
Inputs are resident in HBM before the timed region; the Newton-invariant
data (trajectories, state, ∇m^j) are built once before it, like the
reference's warm-up (forward solve + gradient + one matvec,
apps/benchmark.cpp:257-260).

Prints one JSON line.  Only N = 1 GPU is supported.
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

METRIC = "Hessian matvec/s + interp GB/s @256^3, mixed vs plain FP32, relerr≤1e-5"
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=50)
    p.add_argument("--warmup", type=int, default=5)
    p.add_argument("--n", type=int, default=256, help="grid size per axis")
    p.add_argument("--nt", type=int, default=4)
    p.add_argument("--no-cpu-baseline", action="store_true")
    p.add_argument("--cpu-reps", type=int, default=5,
                   help="timed CPU matvecs after one warm-up matvec (mean and median reported)")
    p.add_argument("--no-parity", action="store_true")
    p.add_argument("--no-registration", action="store_true",
                   help="skip the 64^3 Gauss-Newton-Krylov registration leg")
    p.add_argument("--no-batched", action="store_true",
                   help="skip the 48-registration batched leg")
    p.add_argument("--batch-workers", type=int, default=8)
    p.add_argument("--batch-graphs", type=int, choices=(0, 1), default=1,
                   help="batched leg: replay the Krylov iterations from captured graphs (1) or launch them one by one (0)")
    p.add_argument("--no-variants", action="store_true",
                   help="skip the option variants (matvec types, models, RK2/RK4, continuity) at the same size")
    p.add_argument("--no-large", action="store_true",
                   help="skip the 512^3 leg (40-ellipsoid phantom, nt=8): footprint, matvec/s, 5 GN iterations, "
                        "error vs FP64 on the 128^3 subsample")
    p.add_argument("--no-fp64", action="store_true",
                   help="skip scoring the GPU (plain and mixed) and CPU FP32 matvecs against the FP64 oracle")
    p.add_argument("--calib-copy", action="store_true",
                   help="end with three 256 MiB device copies (reference kernel for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    return p.parse_args()


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), or None."""
    try:
        with open("/proc/cpuinfo", encoding="utf-8") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_threads():
    """Threads of the CPU reference path: OMP_NUM_THREADS where set (the
    host's share of cores for one GPU), else the logical cores, at most 16."""
    return max(1, min(int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)), 16))


def measured_traffic(kernel, n, nt):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 summary of
    this workload (FETCH_SIZE and WRITE_SIZE passes, each scaled by what the
    counter reads for a device copy of known size), or None."""
    path = os.path.join(ROOT, "profiles", f"hbm_traffic_hessian_matvec_{n}.md")
    if not os.path.exists(path):
        return None, None
    for line in open(path, encoding="utf-8"):
        if line.startswith("traffic_bytes_per_launch:"):
            table = json.loads(line.split(":", 1)[1])
            if table.get("nt") == nt and kernel in table.get("kernels", {}):
                return float(table["kernels"][kernel]), os.path.relpath(path, ROOT)
    return None, None


def registration_leg(args):
    """The reference's CPU-runnable configuration: synthetic problem 4 at 64³
    (template (sin²x1+sin²x2+sin²x3)/3, reference = template transported by
    the benchmark velocity, rescaled to [0,1]), H1SN β = 1e-2, nt = 4,
    Gauss–Newton–Krylov with PCG, INVREG, Eisenstat–Walker forcing and
    Armijo, at most 10 Newton iterations.  GPU wall time of the whole solve
    (Hessian matvecs replayed as HIP graphs, and eagerly), beside the FP32
    CPU restatement of the same driver on the host cores."""
    from claire_amd import Claire, synth

    shape, nt, beta = (64, 64, 64), 4, (1e-2, 0.0, 0.0)
    mT, vstar = synth.smooth_template_problem(shape)
    c = Claire(shape, nt=nt, regnorm="h1sn", beta=beta)
    res = {"workload": "GN-Krylov registration, 64^3 FP32, synthetic problem 4, H1SN beta=1e-2, nt=4, maxit=10"}
    try:
        c.set_velocity(vstar)
        c.set_template_image(mT)
        mR = synth.rescale(c.solve_state()[nt])
        c.set_reference_image(mR)
        c.solve_gn(maxit=1)  # first-use allocations and code-object loads
        for graphs in (False, True, False, True):  # best of two solves per launch mode
            c.set_matvec_graphs(graphs)
            c.sync()
            t0 = time.perf_counter()
            v, rep, hist = c.solve_gn(maxit=10)
            c.sync()
            dt = time.perf_counter() - t0
            key = "gpu_graphs" if graphs else "gpu_eager"
            if key in res and res[key]["seconds"] <= dt:
                continue
            res[key] = {"seconds": round(dt, 4), "iterations": rep["iterations"], "matvecs": rep["matvecs"],
                        "objective_evals": rep["objective_evals"], "reason": rep["reason"],
                        "grad_reduction": rep["gnorm"] / rep["gnorm0"], "J": rep["J"]}
            log(f"registration 64^3 ({key}): {dt:.4f}s, {rep['iterations']} its, {rep['matvecs']} matvecs")
        res["seconds"] = min(res["gpu_graphs"]["seconds"], res["gpu_eager"]["seconds"])
        c.set_velocity(v)
        m1 = c.solve_state()[nt]
        res["mismatch_gpu"] = float(np.linalg.norm(mR - m1) / np.linalg.norm(mR - mT))
    finally:
        c.close()
    if not args.no_cpu_baseline:
        from oracle import oracle_py

        threads = cpu_threads()
        oracle_py.set_num_threads(threads)
        t0 = time.perf_counter()
        v32, rep32, _ = oracle_py.gn_solve(np.zeros((3,) + shape, np.float32), mT, mR, nt, "h1sn", beta, maxit=10,
                                           prec="f32")
        dt = time.perf_counter() - t0
        m1 = oracle_py.solve_state(v32, mT, nt, prec="f32")[nt]
        res["cpu_fp32"] = {"seconds": round(dt, 3), "cores": threads, "kind": "port",
                           "iterations": rep32["iterations"], "matvecs": rep32["matvecs"],
                           "reason": rep32["reason"], "grad_reduction": rep32["gnorm"] / rep32["gnorm0"],
                           "J": rep32["J"],
                           "mismatch": float(np.linalg.norm(mR - m1) / np.linalg.norm(mR - mT))}
        res["speedup_vs_cpu"] = round(dt / res["seconds"], 1)
        log(f"registration 64^3 CPU FP32 port ({threads} threads): {dt:.2f}s")
    return res


def guarded(leg, args):
    """Runs a secondary leg; a failure is reported in its place instead of
    losing the headline line."""
    try:
        return leg(args)
    except Exception as e:  # noqa: BLE001 -- reported, not hidden
        log(f"{leg.__name__} failed: {e!r}")
        return {"error": repr(e)}


def batched_leg(args):
    """48 independent registrations: 32×64³, 8×96³, 8×128³.  Each pair is a
    30-ellipsoid phantom (radii 4–20 voxels, intensity U[0,1], σ = 1 voxel
    smoothing, normalized to [0,1]) and its warp by a band-limited velocity
    (|k| <= 8, CFL 2, nt = 4); H1 regularization β = (1e-2, 1e-4), nt = 4,
    Gauss–Newton–Krylov with the default stopping rules.  All 48 solves run
    through clr_solve_gn_batch (worker threads with one HIP stream each,
    matvecs replayed from captured graphs); the same solves one after another
    on one stream are timed beside it and must give identical velocities."""
    from claire_amd import Claire, solve_gn_batch, synth

    sizes = [64] * 32 + [96] * 8 + [128] * 8
    nt, beta = 4, (1e-2, 1e-4, 0.0)
    t0 = time.time()
    problems, refs = [], []
    for i, n in enumerate(sizes):
        shape = (n, n, n)
        mT = synth.ellipsoid_phantom(shape, count=30, seed=1000 + i)
        v = synth.bandlimited_velocity(shape, kmax=8, cfl=2.0, nt=nt, seed=2000 + i)
        c = Claire(shape, nt=nt, regnorm="h1", beta=beta, graphs=bool(args.batch_graphs))
        c.set_velocity(v)
        c.set_template_image(mT)
        mR = c.solve_state()[nt]
        c.set_reference_image(mR)
        problems.append(c)
        refs.append((mT, mR))
    log(f"batched: {len(sizes)} pairs set up in {time.time() - t0:.1f}s")
    res = {"workload": "48 GN-Krylov registrations (32x64^3, 8x96^3, 8x128^3), ellipsoid phantoms warped by "
                       "band-limited v |k|<=8 CFL 2, H1 beta=1e-2, nt=4",
           "points": int(sum(n ** 3 for n in sizes))}
    try:
        solve_gn_batch(problems, workers=args.batch_workers, maxit=1)  # first-use allocations
        for c in problems:
            c.sync()
        t0 = time.perf_counter()
        got = solve_gn_batch(problems, workers=args.batch_workers)
        dt_batch = time.perf_counter() - t0
        t0 = time.perf_counter()
        seq = solve_gn_batch(problems, workers=1)
        dt_seq = time.perf_counter() - t0
        same = all(np.array_equal(a[0], b[0]) and a[1] == b[1] for a, b in zip(got, seq))
        mism = []
        for c, (v, rep), (mT, mR) in zip(problems, got, refs):
            c.set_velocity(v)
            m1 = c.solve_state()[nt]
            mism.append(float(np.linalg.norm(mR - m1) / np.linalg.norm(mR - mT)))
        its = [rep["iterations"] for _, rep in got]
        reasons = sorted({rep["reason"] for _, rep in got})
        res.update({"workers": args.batch_workers, "graphs": bool(args.batch_graphs), "seconds": round(dt_batch, 3),
                    "registrations_per_s": round(len(sizes) / dt_batch, 2),
                    "sequential_seconds": round(dt_seq, 3),
                    "speedup_vs_sequential": round(dt_seq / dt_batch, 2),
                    "batched_equals_sequential": bool(same),
                    "newton_iterations": {"min": min(its), "max": max(its), "total": int(sum(its))},
                    "reasons": reasons,
                    "mismatch": {"min": min(mism), "median": float(np.median(mism)), "max": max(mism)}})
        log(f"batched: {len(sizes)} registrations in {dt_batch:.2f}s ({len(sizes) / dt_batch:.1f}/s), "
            f"sequential {dt_seq:.2f}s, identical={same}")
        if not args.no_cpu_baseline:
            from oracle import oracle_py

            threads = cpu_threads()
            oracle_py.set_num_threads(threads)
            mT, mR = refs[0]
            t0 = time.perf_counter()
            v32, rep32, _ = oracle_py.gn_solve(np.zeros((3,) + mT.shape, np.float32), mT, mR, nt, "h1", beta,
                                               prec="f32")
            dt = time.perf_counter() - t0
            vg, repg = got[0][0].astype(np.float64), got[0][1]
            m1 = oracle_py.solve_state(v32, mT, nt, prec="f32")[nt]
            res["cpu_fp32_sample"] = {
                "pair": 0, "grid": list(mT.shape), "seconds": round(dt, 3), "cores": threads, "kind": "port",
                "iterations": rep32["iterations"], "gpu_iterations": repg["iterations"],
                "pcg_iterations": rep32["pcg_iterations"], "gpu_pcg_iterations": repg["pcg_iterations"],
                "mismatch": float(np.linalg.norm(mR - m1) / np.linalg.norm(mR - mT)), "gpu_mismatch": mism[0],
                "grad_reduction": rep32["gnorm"] / rep32["gnorm0"],
                "gpu_grad_reduction": repg["gnorm"] / repg["gnorm0"],
                "velocity_relerr_gpu_vs_cpu": float(np.linalg.norm(vg - v32) / np.linalg.norm(v32))}
            log(f"batched: CPU FP32 port on pair 0 (64^3): {dt:.2f}s")
            res["pairs_vs_fp64"] = golden_pairs_check(nt, beta)
    finally:
        for c in problems:
            c.close()
    return res


def golden_pairs_check(nt, beta):
    """Per-pair error of the batch's registrations against the stored FP64
    end states (tests/golden/batch_pairs.json: pairs 0, 32 and 40, i.e. one
    of each size) and the spread of the reference FP32 path on the same
    pair.  The golden's reference image is the FP64 state solve rounded to
    FP32 (recomputed here by the oracle, as the checker): these end states
    move by ~1% with the rounding of the reference image, so the GPU solve
    is rerun on exactly that image."""
    from claire_amd import Claire, synth
    from oracle import oracle_py

    path = os.path.join(ROOT, "tests", "golden", "batch_pairs.json")
    if not os.path.exists(path):
        return None
    with open(path, encoding="utf-8") as f:
        pairs = json.load(f)["pairs"]
    out = []
    for key in sorted(pairs, key=int):
        g = pairs[key]
        i, n = int(key), g["n"]
        shape = (n, n, n)
        mT = synth.ellipsoid_phantom(shape, count=30, seed=1000 + i)
        v = synth.bandlimited_velocity(shape, kmax=8, cfl=2.0, nt=nt, seed=2000 + i)
        mR = oracle_py.solve_state(v.astype(np.float64), mT.astype(np.float64), nt)[nt].astype(np.float32)
        c = Claire(shape, nt=nt, regnorm="h1", beta=beta, graphs=True)
        try:
            c.set_template_image(mT)
            c.set_reference_image(mR)
            vs, rep, _ = c.solve_gn()
            c.set_velocity(vs)
            m1 = c.solve_state()[nt]
        finally:
            c.close()
        mis = float(np.linalg.norm(mR - m1) / np.linalg.norm(mR - mT))
        red = rep["gnorm"] / rep["gnorm0"]
        ref = g["oracle_fp64"]
        out.append({"pair": i, "grid": n, "gpu_iterations": rep["iterations"], "fp64_iterations": ref["iterations"],
                    "gpu_mismatch": mis, "fp64_mismatch": ref["mismatch"],
                    "mismatch_relerr": abs(mis - ref["mismatch"]) / ref["mismatch"],
                    "gpu_grad_reduction": red, "fp64_grad_reduction": ref["grad_reduction"],
                    "grad_reduction_relerr": abs(red - ref["grad_reduction"]) / ref["grad_reduction"],
                    "fp32_path_mismatch_relerr_max": max(abs(s["mismatch"] - ref["mismatch"]) / ref["mismatch"]
                                                         for s in g["fp32_spread"]),
                    "fp32_path_grad_reduction_relerr_max": max(
                        abs(s["grad_reduction"] - ref["grad_reduction"]) / ref["grad_reduction"]
                        for s in g["fp32_spread"])})
    return out


def variants_leg(c, v_slow, vt, hv, args):
    """The reference's option variants on the same 256³ FP32 context, each
    timed over a few steps after its own setup: the matvec operator types,
    the two other registration models, the RK2 integrator (on a CFL 0.5
    velocity: it is conditionally stable), the RK4 characteristic (its cost
    is in the trajectories, so the timed step is the trajectory pair plus the
    state solve) and the continuity-equation state solve.  Each entry gives
    ms per step and the dominant kernel class with its algorithmic GB/s
    (HIP events around every launch)."""
    import claire_amd
    from claire_amd import DeviceArray

    steps = max(2, min(args.steps, 8))
    lib = claire_amd.lib()
    m = DeviceArray((args.nt + 1,) + c.nx)

    def matvec():
        c.hessian_matvec_device(hv, vt)

    def state():
        claire_amd._check(lib.clr_solve_state(c.ctx, claire_amd._vp(m.ptr)))

    def setup_and_state():
        # a new velocity drops the trajectories; vt holds the benchmark's v
        claire_amd._check(lib.clr_set_velocity(c.ctx, claire_amd._vp(vt.ptr)))
        state()

    def timed(name, step, prepare=None):
        if prepare:
            prepare()
        step()
        c.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            step()
        c.sync()
        ms = 1e3 * (time.perf_counter() - t0) / steps
        c.profile_reset()
        c.profile_enable(True)
        step()
        c.profile_enable(False)
        prof = c.profile()
        dname, (dms, dcnt, dby) = max(prof.items(), key=lambda kv: kv[1][0])
        res = {"ms_per_step": round(ms, 3), "per_s": round(1e3 / ms, 2), "dominant": dname,
               "dominant_GBps": round(dby / (dms * 1e-3) / 1e9, 1), "dominant_share": round(
                   dms / sum(p[0] for p in prof.values()), 3),
               "kernels_us": {k: round(1e3 * p[0] / p[1], 1) for k, p in
                              sorted(prof.items(), key=lambda kv: -kv[1][0])[:6]}}
        log(f"variant {name}: {ms:.3f} ms, dominant {dname} {res['dominant_GBps']} GB/s")
        return res

    out = {}
    try:
        for kind in ("precond_sym", "precond"):
            c.set_matvec_type(kind)
            out[f"matvec_{kind}"] = timed(f"matvec {kind}", matvec, c.prepare_hessian)
        c.set_matvec_type("default")
        for model in ("stokes", "relaxed_stokes"):
            c.set_reg_model(model)
            out[f"matvec_{model}"] = timed(f"matvec {model}", matvec, c.prepare_hessian)
        c.set_reg_model("compressible")
        c.set_pde_type("continuity")
        out["state_continuity"] = timed("continuity state", state, c.prepare_hessian)
        c.set_pde_type("transport")
        out["state_transport"] = timed("transport state", state, c.prepare_hessian)
        c.set_rk_order(4)
        out["trajectories_and_state_rk4"] = timed("rk4 trajectories + state", setup_and_state)
        c.set_rk_order(2)
        out["trajectories_and_state_rk2"] = timed("rk2 trajectories + state", setup_and_state)
        c.set_pde_solver("rk2")
        c.set_velocity(v_slow)
        out["matvec_pdesolver_rk2"] = timed("matvec -pdesolver rk2", matvec, c.prepare_hessian)
        out["matvec_pdesolver_rk2"]["velocity"] = "band-limited |k|<=8, CFL 0.5"
    finally:
        c.set_pde_solver("sl")
        c.set_rk_order(2)
        c.set_pde_type("transport")
        c.set_reg_model("compressible")
        c.set_matvec_type("default")
        claire_amd._check(lib.clr_set_velocity(c.ctx, claire_amd._vp(vt.ptr)))
        m.free()
    out["steps"] = steps
    return out


def large_leg(args):
    """512³ single registration: a 40-ellipsoid phantom (radii 8–40 voxels,
    σ = 1 voxel smoothing) warped by a band-limited velocity (|k| <= 16,
    CFL 3, nt = 8).  Reports the device footprint, Hessian matvecs per second
    with RegularizationH2 β = 1e-3 (plain and mixed), and five Gauss–Newton
    iterations with H2SN β = 1e-3 (RegularizationH2 as written stalls its
    line search after one step: its gradient omits the hd of its
    functional)."""
    from claire_amd import Claire, DeviceArray, synth

    n, nt = 512, 8
    shape = (n, n, n)
    t0 = time.time()
    mT = synth.ellipsoid_phantom(shape, count=40, seed=7, radii=(8.0, 40.0))
    log(f"large: phantom in {time.time() - t0:.1f}s")
    v = synth.bandlimited_velocity(shape, kmax=16, cfl=3.0, nt=nt, seed=8)
    log(f"large: inputs {shape} in {time.time() - t0:.1f}s")
    res = {"workload": "512^3, 40-ellipsoid phantom warped by band-limited v |k|<=16 CFL 3, nt=8"}
    c = Claire(shape, nt=nt, regnorm="h2", beta=(1e-3, 0.0, 0.0))
    try:
        c.set_velocity(v)
        c.set_template_image(mT)
        mR = c.solve_state()[nt]
        c.set_reference_image(mR)
        c.prepare_hessian()
        vt = DeviceArray.from_host(v)
        hv = DeviceArray((3,) + shape)
        for prec in ("fp32", "mixed"):
            c.set_precision(prec)
            c.prepare_hessian()
            c.hessian_matvec_device(hv, vt)
            c.sync()
            t0 = time.perf_counter()
            for _ in range(5):
                c.hessian_matvec_device(hv, vt)
            c.sync()
            dt = (time.perf_counter() - t0) / 5
            res[f"matvec_{prec}"] = {"regnorm": "h2", "beta": 1e-3, "matvec_per_s": round(1.0 / dt, 2),
                                     "ms_per_matvec": round(1e3 * dt, 2)}
            log(f"large: {prec} matvec {1e3 * dt:.1f} ms")
        res["device_bytes_problem"] = c.device_bytes()
        vt.free()
        hv.free()
    finally:
        c.close()
    if not args.no_cpu_baseline:
        # error against FP64 on the 128^3 subsample of the same problem (every
        # fourth point of the template and the velocity, a trigonometric
        # polynomial of degree 16 < 64), next to the reference FP32 path's
        from oracle import oracle_py

        oracle_py.set_num_threads(cpu_threads())
        ms, vs = np.ascontiguousarray(mT[::4, ::4, ::4]), np.ascontiguousarray(v[:, ::4, ::4, ::4])
        cs = Claire(ms.shape, nt=nt, regnorm="h2", beta=(1e-3, 0.0, 0.0))
        try:
            cs.set_velocity(vs)
            cs.set_template_image(ms)
            Hs = cs.hessian_matvec(vs).astype(np.float64)
            cs.set_precision("mixed")
            Hm = cs.hessian_matvec(vs).astype(np.float64)
        finally:
            cs.close()
        b = (1e-3, 0.0, 0.0)
        H64 = oracle_py.hessian_matvec(vs, oracle_py.solve_state(vs, ms, nt), vs, nt, "h2", b)
        H32 = oracle_py.hessian_matvec(vs, oracle_py.solve_state(vs, ms, nt, prec="f32"), vs, nt, "h2", b, prec="f32")

        def rel(a, r):
            return float(np.linalg.norm(a - r) / np.linalg.norm(r))

        res["relerr_vs_fp64_subsample_128"] = {"gpu_fp32": rel(Hs, H64), "gpu_mixed": rel(Hm, H64),
                                               "cpu_fp32": rel(H32.astype(np.float64), H64)}
        log(f"large: 128^3 subsample relerr vs FP64 gpu {rel(Hs, H64):.3e} mixed {rel(Hm, H64):.3e} "
            f"cpu {rel(H32.astype(np.float64), H64):.3e}")
    c = Claire(shape, nt=nt, regnorm="h2", beta=(1e-3, 0.0, 0.0))
    try:
        # RegularizationH2 as written: its line search stalls after one step
        # (its gradient omits the hd of its functional), as in the FP64 oracle
        c.set_template_image(mT)
        c.set_reference_image(mR)
        t0 = time.perf_counter()
        _, rep, _ = c.solve_gn(maxit=5)
        c.sync()
        res["gn_h2"] = {"regnorm": "h2", "beta": 1e-3, "seconds": round(time.perf_counter() - t0, 2),
                        "iterations": rep["iterations"], "matvecs": rep["matvecs"], "reason": rep["reason"],
                        "grad_reduction": rep["gnorm"] / rep["gnorm0"]}
        log(f"large: GN with H2 {rep['iterations']} its ({rep['reason']})")
    finally:
        c.close()
    c = Claire(shape, nt=nt, regnorm="h2sn", beta=(1e-3, 0.0, 0.0))
    try:
        c.set_template_image(mT)
        c.set_reference_image(mR)
        t0 = time.perf_counter()
        vsol, rep, hist = c.solve_gn(maxit=5)  # 5 Newton iterations (the limit test is iter >= maxit)
        c.sync()
        dt = time.perf_counter() - t0
        c.set_velocity(vsol)
        m1 = c.solve_state()[nt]
        res["gn"] = {"regnorm": "h2sn", "beta": 1e-3, "seconds": round(dt, 2), "iterations": rep["iterations"],
                     "matvecs": rep["matvecs"], "reason": rep["reason"],
                     "grad_reduction": rep["gnorm"] / rep["gnorm0"],
                     "mismatch": float(np.linalg.norm(mR - m1) / np.linalg.norm(mR - mT)),
                     "device_bytes_problem_and_solver": c.device_bytes() + 8 * 3 * 4 * n ** 3}
        log(f"large: GN {rep['iterations']} its in {dt:.1f}s ({rep['reason']})")
    finally:
        c.close()
    return res


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only",
                          "n_gpus": args.gpus}))
        sys.exit(2)

    import claire_amd
    from claire_amd import Claire, DeviceArray, synth

    n, nt = args.n, args.nt
    shape = (n, n, n)
    beta = (1e-2, 1e-4, 0.0)
    regnorm = "h1sn"
    t0 = time.time()
    v = synth.bandlimited_velocity(shape, kmax=8, cfl=2.0, nt=nt, seed=0)
    mT = synth.gaussian_blobs(shape, count=24, seed=1)
    log(f"synthetic inputs {shape} in {time.time() - t0:.1f}s")

    c = Claire(shape, nt=nt, regnorm=regnorm, beta=beta)
    vd = DeviceArray.from_host(v)
    c.set_template_image(mT)
    # per-velocity setup (trajectories, plans, adjoint coefficient, state,
    # the ∇m^j and (∇m^j)(X_s) caches): what a new Newton iterate costs
    # before its first matvec; best of three, inputs resident
    setup_ms = []
    for _ in range(3):
        c.sync()
        t0 = time.perf_counter()
        c.set_velocity_device(vd)
        c.prepare_hessian()
        c.sync()
        setup_ms.append(1e3 * (time.perf_counter() - t0))
    vt = DeviceArray.from_host(v)
    hv = DeviceArray((3,) + shape)
    c.sync()

    for _ in range(args.warmup):
        c.hessian_matvec_device(hv, vt)
    c.sync()

    # timed region: exactly K matvecs, synchronized on both sides
    t_start = time.perf_counter()
    for _ in range(args.steps):
        c.hessian_matvec_device(hv, vt)
    c.sync()
    elapsed = time.perf_counter() - t_start
    ms_per_step = 1e3 * elapsed / args.steps
    value = args.steps / elapsed
    log(f"{args.steps} matvecs in {elapsed * 1e3:.1f} ms -> {value:.1f} matvec/s")

    # per-kernel times: the same K matvecs with HIP events around every launch
    c.profile_reset()
    c.profile_enable(True)
    for _ in range(args.steps):
        c.hessian_matvec_device(hv, vt)
    c.profile_enable(False)
    prof = c.profile()
    total_ms = sum(p[0] for p in prof.values())
    kernels = {}
    for name, (ms, cnt, by) in sorted(prof.items(), key=lambda kv: -kv[1][0]):
        avg_us = 1e3 * ms / cnt
        kernels[name] = {"avg_us": round(avg_us, 2), "launches_per_step": cnt / args.steps,
                         "share": round(ms / total_ms, 4), "GBps": round(by / (ms * 1e-3) / 1e9, 1)}
    dom = max(prof.items(), key=lambda kv: kv[1][0])
    dname, (dms, dcnt, dbytes) = dom
    avg_s = dms * 1e-3 / dcnt
    achieved = (dbytes / dcnt) / avg_s / 1e9
    traffic, traffic_src = measured_traffic(dname, n, nt)
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic, "kernel": dname,
                "algorithmic_bytes_per_launch": dbytes / dcnt, "avg_us": round(avg_s * 1e6, 2),
                "traffic_source": traffic_src}

    # interpolation throughput: scalar gather along the state trajectory (20 B/point)
    f = DeviceArray.from_host(mT)
    fo = DeviceArray(shape)
    c.profile_reset()
    c.profile_enable(True)
    for _ in range(args.steps):
        claire_amd.lib().clr_interp_scalar(c.ctx, 0, claire_amd._vp(f.ptr), claire_amd._vp(fo.ptr))
    c.profile_enable(False)
    g = c.profile()["gather_scalar"]
    interp = {"kernel": "gather_scalar", "GBps": round(g[2] / (g[0] * 1e-3) / 1e9, 1),
              "frac": round(g[2] / (g[0] * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "avg_us": round(1e3 * g[0] / g[1], 2),
              "bytes_per_point": 20}

    # The whole matvec against HBM: the compulsory traffic of the reference's
    # structure (every array read and written once per operation, no caches:
    # incremental state 36 + 152·nt B/pt, incremental adjoint 100 + 92·nt,
    # the hd scaling 24, the regularization 24, the final AXPY 36, the
    # flat/component copies 48), and the algorithmic bytes of this path's
    # launches (the per-kernel sums above, with the Newton-invariant caches).
    ref_bytes = (268.0 + 244.0 * nt) * n ** 3
    path_bytes = sum(p[2] for p in prof.values()) / args.steps
    matvec_roofline = {
        "reference_structure_bytes": ref_bytes,
        "reference_structure_bound_per_s": round(HBM_PEAK_GBS * 1e9 / ref_bytes, 1),
        "reference_structure_GBps": round(ref_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "path_bytes": path_bytes,
        "path_GBps": round(path_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "path_frac": round(path_bytes / (ms_per_step * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
    log(f"matvec: {ref_bytes / 1e9:.1f} GB reference-structure traffic (bound "
        f"{matvec_roofline['reference_structure_bound_per_s']} matvec/s at 8 TB/s), "
        f"{path_bytes / 1e9:.1f} GB moved by this path at {matvec_roofline['path_GBps']} GB/s")

    setup = min(setup_ms)
    per_velocity = {"setup_ms": round(setup, 3),
                    "what": "set_velocity + prepare_hessian: RK2 state and adjoint trajectories, plans, adjoint "
                            "coefficient, state solve, grad m^j and (grad m^j)(X_s) caches",
                    "amortised": {str(k): {"matvecs_per_newton_step": k,
                                           "matvec_per_s": round(k / (1e-3 * setup + k * 1e-3 * ms_per_step), 2)}
                                  for k in (1, 2, 12, 29)},
                    "note": "matvecs per Newton step: 2 on the 64^3 problem-4 registration, 12 on the 128^3 "
                            "H2SN phantom, 29 on the 64^3 H1 phantom pairs"}
    log(f"per-velocity setup {setup:.2f} ms")

    out = {"metric": METRIC, "value": round(value, 3), "unit": "matvec/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "f32", "data": "synthetic",
           "config": {"workload": f"GN-SL Hessian matvec, {n}^3 FP32, nt={nt}, H1SN beta=1e-2, band-limited v "
                                  f"|k|<=8 CFL 2, 24 Gaussian blobs",
                      "grid": list(shape), "nt": nt, "regnorm": regnorm, "beta": list(beta[:2]), "precision": "fp32"},
           "roofline": roofline, "matvec_roofline": matvec_roofline, "interp": interp, "kernels": kernels,
           "per_velocity": per_velocity, "device_bytes": c.device_bytes()}

    Hgpu = hv.download()

    # mixed form: the same matvec with 16-bit block-floating-point storage of the Newton-invariant
    # gradient fields (FP32 arithmetic)
    c.set_precision("mixed")
    c.prepare_hessian()
    for _ in range(args.warmup):
        c.hessian_matvec_device(hv, vt)
    c.sync()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        c.hessian_matvec_device(hv, vt)
    c.sync()
    elapsed_m = time.perf_counter() - t_start
    value_m = args.steps / elapsed_m
    log(f"mixed: {args.steps} matvecs in {elapsed_m * 1e3:.1f} ms -> {value_m:.1f} matvec/s")
    Hmix = hv.download()
    c.set_precision("fp32")
    out["mixed"] = {"value": round(value_m, 3), "unit": "matvec/s", "ms_per_step": round(1e3 * elapsed_m / args.steps, 4),
                    "speedup_vs_fp32": round(value_m / value, 3),
                    "form": "16-bit block floating point (int16 mantissas, one power-of-two scale per 256 "
                            "values) storage of grad m^j and (grad m^j)(X_s), FP32 arithmetic, FP64 reductions"}

    if not args.no_cpu_baseline:
        from oracle import oracle_py

        threads = cpu_threads()
        oracle_py.set_num_threads(threads)
        log(f"CPU reference path: FP32 port, {threads} threads")
        m32 = oracle_py.solve_state(v, mT, nt, prec="f32")
        _, warm = oracle_py.hessian_bench(v, m32, v, nt, regnorm, beta, reps=1, prec="f32")
        reps = max(1, args.cpu_reps)
        Hcpu, secs = oracle_py.hessian_bench(v, m32, v, nt, regnorm, beta, reps=reps, prec="f32")
        mean_s, median_s = float(np.mean(secs)), float(np.median(secs))
        out["cpu_baseline"] = {"value": round(1.0 / mean_s, 5), "unit": "matvec/s", "cores": threads,
                               "kind": "port", "cpu_model": cpu_model(),
                               "seconds_mean": round(mean_s, 3), "seconds_median": round(median_s, 3),
                               "seconds_each": [round(s, 3) for s in secs],
                               "sample": f"{reps} Hessian matvecs of the same {n}^3 workload after one warm-up "
                                         f"matvec ({warm[0]:.1f}s), FP32 OpenMP restatement of the reference path "
                                         f"(its own mixed-radix FFT, not FFTW), trajectories precomputed; threads = "
                                         f"OMP_NUM_THREADS, the host's core share for one GPU"}
        if not args.no_parity:
            def rel(a, b):
                a = a.astype(np.float64)
                b = b.astype(np.float64)
                return float(np.linalg.norm(a - b) / np.linalg.norm(b))

            out["relerr_vs_cpu_fp32"] = rel(Hgpu, Hcpu)
            out["mixed"]["relerr_vs_cpu_fp32"] = rel(Hmix, Hcpu)
            if not args.no_fp64:
                log("FP64 oracle for the same matvec")
                m64 = oracle_py.solve_state(v, mT, nt, prec="f64")
                H64 = oracle_py.hessian_matvec(v, m64, v, nt, regnorm, beta, prec="f64")
                e = {"gpu_fp32": rel(Hgpu, H64), "cpu_fp32": rel(Hcpu, H64), "gpu_mixed": rel(Hmix, H64)}
                out["relerr_vs_fp64"] = e
                out["mixed"]["relerr_vs_fp64"] = e["gpu_mixed"]
                out["parity"] = {
                    "bar": "GPU FP32 Hessian no further from FP64 than the reference FP32 path (1e-5 against that "
                           "path is out of reach of any other FP32 implementation at 256^3: foot points quantized "
                           "to 2^-24 of the domain put every FP32 path ~5e-5 from FP64, and two of them differ by "
                           "the root sum of squares); mixed form <= 1e-4 vs FP64",
                    "gpu_fp32_vs_fp64": e["gpu_fp32"], "cpu_fp32_vs_fp64": e["cpu_fp32"],
                    "gpu_fp32_vs_cpu_fp32": out["relerr_vs_cpu_fp32"], "gpu_mixed_vs_fp64": e["gpu_mixed"],
                    "met": bool(e["gpu_fp32"] <= e["cpu_fp32"] * 1.05 and e["gpu_mixed"] <= 1e-4)}
    if not args.no_variants:
        v_slow = synth.bandlimited_velocity(shape, kmax=8, cfl=0.5, nt=nt, seed=5)

        def variants(a):
            return variants_leg(c, v_slow, vt, hv, a)

        out["variants"] = guarded(variants, args)
    if not args.no_registration:
        out["registration"] = guarded(registration_leg, args)
    if not args.no_batched:
        out["batched"] = guarded(batched_leg, args)
    if not args.no_large:
        out["large"] = guarded(large_leg, args)
    if args.calib_copy:
        big = DeviceArray((64 * 1024 * 1024,))
        big2 = DeviceArray((64 * 1024 * 1024,))
        for _ in range(3):
            claire_amd.lib().clr_memcpy_dtod(claire_amd._vp(big2.ptr), claire_amd._vp(big.ptr),
                                             ctypes.c_size_t(big.nbytes))
        claire_amd.lib().clr_device_synchronize()
        log("calibration copies: 3 x 256 MiB device to device")
    c.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
