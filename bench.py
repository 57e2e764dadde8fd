#!/usr/bin/env python3
"""Part-reduction throughput of the MI355X reducer on one GPU.

Workload: the 48x24x24 HEX8 steel block, 0.48x0.24x0.24 m, nodal jitter
+-5 % from seed 0, 12 external nodes on the x=0 / x=max faces (72 boundary
DOFs), 20 fixed-interface component modes, FP64 throughout.

One "step" is one part through the whole per-part path of the reducer:
the host analysis (SAM numbering, nested-dissection ordering with the
boundary last, the fronts), the upload of the part and of its structure,
the device build of the CSR pattern, assembly contribution lists and front
map, then element matrices, assembly, the LDL^T factorisation with the Schur
complement, B, the block-Lanczos component modes, the mass / stiffness /
gravity reduction and the reduced matrices copied back to the host.  The
part's coordinates, topology and properties are in host memory when a step
starts (the reducer's input).  Parts stream through --pipelines worker
threads (default 6), each with its own reducer handle and HIP stream, so
that a part's host analysis and its latency-bound solve phases overlap
other parts' device work.  Reported beside it: the same path with nothing
overlapped (per-part latency) and the numeric re-reduction of an already
analysed part.

Prints one JSON line (see the keys below).  --gpus N > 1 exits: the
reducer runs on one GPU.
"""
import argparse
import json
import os
import subprocess
import sys
import threading
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fedem-solvers_amd"))

PEAK_FP64_MFMA = 78.6  # TFLOP/s, MI355X dense FP64 matrix


def cpu_baseline(seed=0):
    """One full reduction of the bench workload on the host cores, in a
    subprocess: oracle/multifrontal.py (the same multifrontal algorithm with
    the boundary last, over this library's host ordering; dense front work
    in numpy/OpenBLAS, which uses every core the process may run on),
    from the part's coordinates to SK, SM, G and the 20 modes."""
    code = r"""
import os, sys, time, json
sys.path.insert(0, %r); sys.path.insert(0, %r)
from fedempy_amd import gpu, parts
from oracle import multifrontal as mf
part = parts.config1(seed=%d)
t = time.perf_counter()
R = gpu.GpuReducer(); R.set_part(part); R.analyze(ngen=20)
mf.reduce(part, R.supernodes(), ngen=20)
print(json.dumps({"s": time.perf_counter() - t, "cores": len(os.sched_getaffinity(0))}))
""" % (ROOT, os.path.join(ROOT, "fedem-solvers_amd"), seed)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=900)
    if out.returncode != 0:
        raise RuntimeError(out.stderr[-2000:])
    return json.loads(out.stdout.strip().splitlines()[-1])


def cpu_baselines(seed=0, budget_s=15.0):
    """The other configurations on the host cores, in a subprocess, by the
    same CPU port (oracle/multifrontal.py over this library's host ordering):
    the 20x10 cantilever plate (median of 5 runs), the 200x200 plate (FP64,
    one run), and bounded samples of the batched and served mixes (the same
    seeded parts / requests the GPU legs draw, in their order, until about
    `budget_s` seconds of work, at least 3): reductions, and for requests
    their load cases and internal-displacement recovery."""
    code = r"""
import os, sys, time, json
import numpy as np
sys.path.insert(0, %r); sys.path.insert(0, %r)
from fedempy_amd import gpu, parts
from oracle import multifrontal as mf
budget = %r
seed = %d
def one(part, ngen, loads=None, ub=None):
    t = time.perf_counter()
    R = gpu.GpuReducer(); R.set_part(part); neq = R.analyze(ngen=ngen)["neq"]
    sn = R.supernodes(); R.close()
    if loads is None:
        mf.reduce(part, sn, ngen=ngen)
    else:
        mf.serve(part, sn, ngen, loads, ub)
    return time.perf_counter() - t, neq
out = {"cores": len(os.sched_getaffinity(0))}
out["config0"] = float(np.median([one(parts.config0(), 0)[0] for _ in range(5)]))
out["config2"] = one(parts.config2(seed=seed), 0)[0]
rng = np.random.default_rng(seed)
mix = [parts.random_part(rng) for _ in range(256)]
t, n, dof = 0.0, 0, 0
for p, g in mix:
    dt, neq = one(p, g); t += dt; n += 1; dof += neq
    if n >= 3 and t >= budget:
        break
out["config3"] = {"parts": n, "s": t, "mean_neq": dof / n}
rng = np.random.default_rng(seed)
reqs = [parts.served_request(rng) for _ in range(160)]
t, n, dof = 0.0, 0, 0
for p, g, loads, ub in reqs:
    dt, neq = one(p, g, loads, ub); t += dt; n += 1; dof += neq
    if n >= 3 and t >= budget:
        break
out["config4"] = {"requests": n, "s": t, "mean_neq": dof / n}
print(json.dumps(out))
""" % (ROOT, os.path.join(ROOT, "fedem-solvers_amd"), budget_s, seed)
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=1200)
    if out.returncode != 0:
        raise RuntimeError(out.stderr[-2000:])
    return json.loads(out.stdout.strip().splitlines()[-1])


def measured_traffic():
    """HBM bytes per k_update launch (FETCH_SIZE + WRITE_SIZE, copy-kernel
    calibrated) from the committed PMC summary, or None without one."""
    path = os.path.join(ROOT, "profiles", "k_update_hex8_48x24x24_traffic.md")
    try:
        for ln in open(path):
            if ln.startswith("traffic_bytes_per_launch:"):
                return float(ln.split(":", 1)[1])
    except OSError:
        pass
    return None


def batched(n_parts, slots, seed):
    """Config-3 mix (plates, HEX8 blocks, beam frames; uneven sizes, 0-10
    modes) through the continuous-batching server with `slots` concurrent
    streams, and through one slot (the plain sequential GPU path)."""
    from fedempy_amd import parts
    from fedempy_amd.server import Server
    rng = np.random.default_rng(seed)
    reqs = [parts.random_part(rng) for _ in range(n_parts)]
    out, nmixed = {}, 0
    for k, mixed in ((1, False), (slots, False), (slots, True)):
        srv = Server(slots=k)
        if mixed:
            srv.set_option("mixedPrecision", 1)
        for p, g in reqs[:4]:  # warm the slots (first-touch allocations)
            srv.result(srv.submit(p, g))
        t0 = time.perf_counter()
        ids = [srv.submit(p, g) for p, g in reqs]
        for i in ids:
            nmixed += int(srv.result(i)["mixed"])
            srv.release(i)
        out[(k, mixed)] = n_parts / (time.perf_counter() - t0)
        srv.close()
    return {"workload": "%d parts: shell plates 10-80 squared, HEX8 blocks 6-24 cubed, beam frames 50-500 "
                        "elements, 2-16 external nodes, 0-10 modes, FP64" % n_parts,
            "slots": slots, "parts_per_s": out[(slots, False)], "plain_one_stream_parts_per_s": out[(1, False)],
            "speedup": out[(slots, False)] / out[(1, False)],
            "mixed_precision": {"parts_per_s": out[(slots, True)],
                                "vs_plain_fp64_one_stream": out[(slots, True)] / out[(1, False)],
                                "parts_in_mixed_precision": nmixed,
                                "parts_redone_in_fp64": n_parts - nmixed}}


def mixed_plate(seed, reps=3):
    """The 200 x 200 ANDES4 plate with lognormal thickness and 64 edge nodes
    (242,406 DOF, 384 boundary DOF): the FP64 path and the mixed-precision
    path (FP32 fronts, FP32 MFMA updates, FP64 diagonal blocks, refined B),
    each reduction timed end to end; the Schur/factor rate is the symbolic
    factor flop count over the measured factor phase; error of every mixed
    output against the FP64 path."""
    from fedempy_amd import gpu, parts
    part = parts.config2(seed=seed)
    out, res = {}, {}
    for mode in ("fp64", "mixed"):
        R = gpu.GpuReducer()
        R.set_part(part)
        if mode == "mixed":
            R.set_option("mixedPrecision", 1)
        R.analyze(ngen=0)
        R.reduce()  # warm-up (allocations, graph capture)
        t0 = time.perf_counter()
        for _ in range(reps):
            R.reduce()
            SK, SM, G = R.reduced()
        el = (time.perf_counter() - t0) / reps
        tt, info = R.times(), R.info()
        res[mode] = (SK, SM, G, R.B())
        out[mode] = {"ms_per_reduction": 1e3 * el, "reductions_per_s": 1.0 / el,
                     "factor_ms": tt["factor"], "bsolve_ms": tt["bsolve"], "mass_ms": tt["mass"],
                     "schur_tflops": info["factor_flops"] / (tt["factor"] * 1e-3) / 1e12,
                     "refine_steps": int(tt["refine_steps"])}
    rel = lambda a, b: float(np.linalg.norm(a - b) / np.linalg.norm(b))
    out["mixed"]["rel_error_vs_fp64"] = {k: rel(a, b) for k, a, b in zip(("SK", "SM", "G", "B"), res["mixed"],
                                                                           res["fp64"])}
    out["workload"] = "200x200 ANDES4 plate, lognormal thickness, 64 edge nodes (242,406 DOF, 384 boundary DOF)"
    return out


def _drain(srv, reqs, gaps=None, keep=0):
    """Submit reqs (back to back, or at the given inter-arrival gaps) and
    wait for all; returns (wall seconds, latencies in ms, the results of the
    first `keep` requests, how many ran in mixed precision)."""
    t0 = time.perf_counter()
    ids, due = [], t0
    for k, r in enumerate(reqs):
        if gaps is not None:
            due += gaps[k]
            while time.perf_counter() < due:
                time.sleep(min(0.001, max(0.0, due - time.perf_counter())))
        ids.append(srv.submit(*r))
    lat, kept, nmixed = [], [], 0
    for k, i in enumerate(ids):
        res = srv.result(i)
        lat.append(res["latency_ms"])
        nmixed += int(res["mixed"])
        if k < keep:
            kept.append(res)
        srv.release(i)
    return time.perf_counter() - t0, np.array(lat), kept, nmixed


def served(n_req, slots, seed):
    """Config-4 serving mix: DOF count lognormal (median 20k, cap 300k),
    every request a reduction plus 1-8 load cases (reduced loads and
    recovered internal displacements).  Capacity = requests/s with the queue
    kept full, on `slots` concurrent streams and on one (the plain sequential
    FP64 path); latency p50/p99 under Poisson arrivals at 70 % of capacity."""
    from fedempy_amd import parts
    from fedempy_amd.server import Server
    rng = np.random.default_rng(seed)
    reqs = [parts.served_request(rng) for _ in range(n_req)]
    nkeep = 24
    cap, kept = {}, {}

    def capacity(k, mixed=False):
        srv = Server(slots=k)
        if mixed:
            srv.set_option("mixedPrecision", 1)
        for r in reqs[:4]:  # warm the slots
            srv.result(srv.submit(*r))
        el, _, res, nmixed = _drain(srv, reqs, keep=nkeep)
        srv.close()
        return n_req / el, res, nmixed

    def poisson(rate, mixed=False):
        srv = Server(slots=slots)
        if mixed:
            srv.set_option("mixedPrecision", 1)
        for r in reqs[:4]:
            srv.result(srv.submit(*r))
        el, lat, _, _ = _drain(srv, reqs, rng.exponential(1.0 / rate, n_req))
        srv.close()
        return n_req / el, lat

    for k in (1, slots):
        cap[k], kept[k], _ = capacity(k)
    rate = 0.7 * cap[slots]
    served_rate, lat = poisson(rate)
    # the same mix with mixedPrecision set on the server: capacity, latency
    # at 70 % of that capacity, and the error of the first requests'
    # mode-independent outputs against their FP64 results
    cap_m, kept_m, nmixed = capacity(slots, mixed=True)
    rate_m = 0.7 * cap_m
    served_m, lat_m = poisson(rate_m, mixed=True)
    rel = lambda a, b: float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
    err = {"SK": 0.0, "SM": 0.0, "G": 0.0, "Fred": 0.0, "Ui": 0.0}
    for a, m in zip(kept[slots], kept_m):
        if not m["mixed"]:
            continue  # redone in FP64
        n2 = a["ndof2"]
        err["SK"] = max(err["SK"], rel(m["SK"][:n2, :n2], a["SK"][:n2, :n2]))
        err["SM"] = max(err["SM"], rel(m["SM"][:n2, :n2], a["SM"][:n2, :n2]))
        err["G"] = max(err["G"], rel(m["G"][:n2], a["G"][:n2]))
        if a["Fred"].size:
            err["Fred"] = max(err["Fred"], rel(m["Fred"][:n2], a["Fred"][:n2]))
            err["Ui"] = max(err["Ui"], rel(m["Ui"], a["Ui"]))
    return {"workload": "%d requests, DOF lognormal median 20k cap 300k (mean %.0f), 1-8 load cases + "
                        "recovery each, FP64" % (n_req, np.mean([6 * len(r[0]["xyz"]) for r in reqs])),
            "slots": slots, "requests_per_s": cap[slots], "plain_one_stream_requests_per_s": cap[1],
            "speedup": cap[slots] / cap[1], "poisson_rate_per_s": rate, "poisson_requests_per_s": served_rate,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "mixed_precision": {
                "requests_per_s": cap_m, "vs_plain_fp64_one_stream": cap_m / cap[1],
                "vs_fp64_same_slots": cap_m / cap[slots], "poisson_rate_per_s": rate_m,
                "poisson_requests_per_s": served_m, "p50_ms": float(np.percentile(lat_m, 50)),
                "p99_ms": float(np.percentile(lat_m, 99)),
                "requests_in_mixed_precision": nmixed,
                "requests_redone_in_fp64": n_req - nmixed,
                "max_rel_error_vs_fp64": err,
                "error_sample": "those of the first %d requests that ran in mixed precision; SK, SM, G: boundary "
                                "(Guyan) blocks, Fred: boundary rows, Ui: all (the mode blocks depend on the "
                                "modes' basis)" % nkeep}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=48)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--slots", type=int, default=16, help="server slots of the batched and served measurements")
    ap.add_argument("--pipelines", type=int, default=6,
                    help="worker threads, each with its own reducer handle, taking the parts of the stream")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-serving", action="store_true", help="skip the batched and served measurements")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision plate measurement")
    ap.add_argument("--autotune", action="store_true",
                    help="tune the factor panel width and solve split on the part before timing (default: the "
                         "library defaults, which the committed profiles were measured with)")
    ap.add_argument("--only", choices=("stream", "mixed", "serving", "cpu"), default=None,
                    help="run one measurement alone and print it as a JSON line (for profiling); 'stream' is "
                         "the headline's pipelined part stream without the single-handle phases after it")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run one plain device copy (torch.neg over N MiB of FP64) whose HBM "
                         "counters calibrate FETCH_SIZE/WRITE_SIZE of the profiled kernels")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this reducer runs on one GPU; --gpus %d is not supported" % a.gpus, file=sys.stderr)
        sys.exit(2)
    from fedempy_amd import gpu, parts

    if a.only == "mixed":
        print(json.dumps({"mixed_precision": mixed_plate(a.seed)}))
        return
    if a.only == "serving":
        print(json.dumps({"batched": batched(256, a.slots, a.seed), "served": served(160, a.slots, a.seed)}))
        return
    if a.only == "cpu":
        print(json.dumps({"cpu_baselines": cpu_baselines(a.seed)}))
        return
    if a.calib_copy_mib > 0:
        import torch
        src = torch.ones(a.calib_copy_mib * 2**20 // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        torch.neg(src, out=dst)
        torch.cuda.synchronize()
        del src, dst

    part = parts.config1(seed=a.seed)
    # ---- headline: a stream of parts through the whole per-part path.  A
    # step is one part: host analysis (SAM numbering, nested-dissection
    # ordering, fronts: csAllocPointerArrays/csRenumber/csPreassemble), the
    # upload of the part and its structure, the device build of the CSR
    # pattern, contribution lists and front map, then the numeric reduction
    # and the reduced matrices back on the host.  `pipelines` worker threads,
    # each with its own reducer handle (and HIP stream), take the parts one
    # after the other, so that one part's host analysis and its latency-bound
    # solve phases overlap other parts' device work.
    npipe = max(1, a.pipelines)
    Rs = [gpu.GpuReducer() for _ in range(npipe)]
    for R in Rs:
        R.set_part(part)
    tuned = None
    if a.autotune:  # one-time tile tuning on this part, outside the timed region
        Rs[0].analyze(ngen=20)
        tuned = Rs[0].autotune()
        for R in Rs:
            R.set_option("factorPanel", tuned["factorPanel"])
            R.set_option("solveSplitK", tuned["solveSplitK"])
    pool = ThreadPoolExecutor(npipe)

    def stream(n):
        """n parts through the pipelines (each part: analyze, reduce, results)"""
        left = [n]
        lock = threading.Lock()
        sizes = {}

        def worker(i):
            R = Rs[i]
            while True:
                with lock:
                    if left[0] == 0:
                        return
                    left[0] -= 1
                sizes[i] = R.analyze(ngen=20)
                R.reduce()
                R.reduced()
        for f in [pool.submit(worker, i) for i in range(npipe)]:
            f.result()
        return next(iter(sizes.values()))

    stream(max(2 * npipe, a.warmup))
    t0 = time.perf_counter()
    sizes = stream(a.steps)
    el = time.perf_counter() - t0
    if a.only == "stream":
        print(json.dumps({"stream": {"value": a.steps / el, "unit": "reductions/s", "steps": a.steps,
                                     "ms_per_step": 1e3 * el / a.steps, "pipelines": npipe}}))
        return
    R = Rs[0]
    phases = R.times()
    info = R.info()
    # the same path with nothing overlapped: per-part latency
    t0 = time.perf_counter()
    nser = max(2, a.steps // 2)
    for _ in range(nser):
        R.analyze(ngen=20)
        R.reduce()
        R.reduced()
    serial_ms = 1e3 * (time.perf_counter() - t0) / nser
    # numeric re-reduction of an analysed part (structure resident, factor
    # schedule replayed as a captured graph)
    for _ in range(2):
        R.reduce()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        R.reduce()
        R.reduced()
    numeric_ms = 1e3 * (time.perf_counter() - t0) / a.steps
    numeric_phases = R.times()
    # roofline of the dominant MFMA kernel (trailing LDL^T update), timed per
    # launch with HIP events on the reducer's stream in one extra reduction
    R.set_option("timeUpdates", 1)
    R.analyze(ngen=20)
    R.reduce()
    tt = R.times()
    upd_ms, upd_n = tt["update_kernel_ms"], int(tt["update_launches"])
    upd_flops = info["update_flops"]
    achieved = (upd_flops / upd_n) / ((upd_ms / upd_n) * 1e-3) / 1e12
    nlz, resid = R.lanczos()
    line = {
        "metric": "part reductions/s",
        "value": a.steps / el,
        "unit": "reductions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * el / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded part generator)",
        "config": {"workload": "48x24x24 HEX8 block, 91,911 DOF, 72 boundary DOF, 20 component modes, FP64",
                   "neq": sizes["neq"], "ndof2": sizes["ndof2"], "modes": 20, "seed": a.seed,
                   "parallelism": "1 GPU, %d part pipelines" % npipe, "tiles": tuned or "defaults"},
        "roofline": {"bound": "mfma", "kernel": "k_update (trailing LDL^T update, FP64 MFMA)",
                     "achieved": achieved, "peak": PEAK_FP64_MFMA, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_FP64_MFMA, "traffic": measured_traffic(),
                     "launches": upd_n, "avg_launch_us": 1e3 * upd_ms / upd_n,
                     "flops_per_launch": upd_flops / upd_n},
        "phases_ms": {k: round(v, 3) for k, v in phases.items() if k not in ("update_kernel_ms", "update_launches")},
        "per_part_serial_ms": serial_ms,
        "numeric_only": {"what": "numeric re-reduction of an analysed part (structure resident)",
                         "ms_per_reduction": numeric_ms, "reductions_per_s": 1e3 / numeric_ms,
                         "phases_ms": {k: round(v, 3) for k, v in numeric_phases.items()
                                       if k not in ("update_kernel_ms", "update_launches", "analyze_host",
                                                    "upload")}},
        "symbolic": {"analyze_host_ms": round(phases["analyze_host"], 3), "upload_ms": round(phases["upload"], 3),
                     "factor_gflop": info["factor_flops"] / 1e9,
                     "nnzL": info["nnzL"], "fronts": info["nsup"], "lanczos_vectors": nlz,
                     "lanczos_resid": resid},
    }
    if not a.no_mixed:
        line["mixed_precision"] = mixed_plate(a.seed)
    if not a.no_serving:
        line["batched"] = batched(256, a.slots, a.seed)
        line["served"] = served(160, a.slots, a.seed)
    if not a.no_cpu_baseline:
        cb = cpu_baseline(a.seed)
        line["cpu_baseline"] = {
            "value": 1.0 / cb["s"], "unit": "reductions/s", "cores": cb["cores"], "kind": "port",
            "sample": "one full reduction of the same 48x24x24 workload (SAM numbering and ordering, "
                      "element matrices and assembly, multifrontal Cholesky with the Schur front, B, "
                      "20 ARPACK shift-invert modes, mass and gravity reduction) by oracle/multifrontal.py "
                      "on numpy/OpenBLAS over %d host cores: %.1f s" % (cb["cores"], cb["s"])}
        cbs = cpu_baselines(a.seed)
        c = cbs["cores"]
        other = {
            "config0": {"value": 1.0 / cbs["config0"], "unit": "reductions/s", "cores": c, "kind": "port",
                        "sample": "20x10 ANDES4 cantilever plate, 12 boundary DOF: median of 5 reductions, "
                                  "%.3f s" % cbs["config0"]},
            "config2": {"value": 1.0 / cbs["config2"], "unit": "reductions/s", "cores": c, "kind": "port",
                        "sample": "200x200 ANDES4 plate, 384 boundary DOF, FP64: one reduction, %.1f s"
                                  % cbs["config2"]},
            "config3": {"value": cbs["config3"]["parts"] / cbs["config3"]["s"], "unit": "parts/s", "cores": c,
                        "kind": "port",
                        "sample": "the first %d of the 256 batched parts (mean %.0f equations), %.1f s"
                                  % (cbs["config3"]["parts"], cbs["config3"]["mean_neq"],
                                     cbs["config3"]["s"])},
            "config4": {"value": cbs["config4"]["requests"] / cbs["config4"]["s"], "unit": "requests/s",
                        "cores": c, "kind": "port",
                        "sample": "the first %d of the 160 served requests (mean %.0f equations) with their load "
                                  "cases and recovery, one after the other, %.1f s"
                                  % (cbs["config4"]["requests"], cbs["config4"]["mean_neq"],
                                     cbs["config4"]["s"])},
        }
        line["cpu_baselines_other_configs"] = other
    print(json.dumps(line))


if __name__ == "__main__":
    main()
