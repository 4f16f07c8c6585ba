#!/usr/bin/env python3
"""Benchmark of the hot path on one MI355X: the unsteady pitch + heave workload.

Workload (the unsteady configuration the headline metric is quoted on): flat rectangular wing,
AR 4 (chord 1, half span 2), 10 chordwise x 40 spanwise vortex-ring panels (N = 400), heave
0.1 c sin(t), pitch 2 deg sin(t + pi/2) about the quarter chord, reduced frequency k = 0.5,
dt = c / (10 u_inf) = last panel chord / u_inf, 500 time steps; the wake grows to 500 x 40 rings.

A bench "step" is one batch: 64 simulations of this workload (independent requests of the same
configuration, each computed in full) held by one plan and advanced together by every launch (motion
evaluation and upload before timing; then per batch: pose, influence matrix, LU, inverse, geometry of
every step, and 500 UVLM time steps with wake induction, solve, forces, CL and CM for all 64), all
inside the HIP library.  value = UVLM time steps per second = 64 * 500 * K / (time of K batches),
inputs resident on the device.  The served path evaluates the wake induction with FP64 arithmetic (1/sqrt
and 1/x from the FP64 hardware estimates plus one Newton step) over the shared-segment lattice of the
wake, with the far wake aggregated (ring rows older than 80 / 160 / 320 steps merged into 2x4 / 4x8 /
8x8 rings of mean circulation; --exact-wake turns this off), solves with the FP64 inverse and replays
the loop as one hipGraph; it is accepted only if |dCL| <= 1e-6 at every step of every simulation
against the plain FP64 path (exact FP64 Biot-Savart ring by ring over the whole wake, LU substitution
each step, one simulation at a time), which is timed too, and so is the served batch with the exact
wake ("served_exact_wake") and one simulation per plan ("single_simulation": the latency a lone
request sees).

Also reported under "served_uvlm": 64 such simulations requested at once through the serving loop
(4 lanes, each running up to 8 simulations in one plan).  --batched-only prints only a batch's rate
and wake roofline (for profiler runs).

Also reported: the roofline of the dominant kernel (wake induction, 398 algorithmic flops per
receiver x wake-ring pair -- the ring-by-ring count of the reference algorithm -- timed with HIP
events on the plan's stream, against the FP64 VALU peak since it computes in FP64) and the CPU
baseline (the oracle's FP64 restatement of the reference CPU backend, OpenMP on the host cores).

Also reported under "steady": the steady configuration -- elliptic wing mesh/elliptic_64x64.x
(N = 4096), u_inf = 1, rho = 1 -- as served batches of 64 requests (alpha = 5 deg and 63 more drawn
from U[-5, 10] deg, seed 0: seeded FP64 influence matrices, one batched FP32 LU, FP64 refinement)
against the plain FP64 path (one request at a time, exact FP64 assembly, FP64 LU), with |dCL| between
them on the same angles, per-batch latency, and one CPU solve of the oracle.

Also reported under "assembly": the influence-matrix assembly of the same elliptic mesh in FP32
(rsqrt form) and FP64 (exact form), TFLOP/s at 398 flops per pair against the VALU rates.

Also reported under "hb": the harmonic-balance driver (the libhbvlm3 interface) on a rectangular
wing/flap section of chord 2b = 0.254 and span 8 chords (wing 20 + flap 10 chordwise x 20 spanwise =
600 panels, H = 5: 11 time instances), heave 0.05 c and pitch 1 deg, omega = 0.8, as a continuation
over airspeeds U = linspace(0.5, 1.5) * 6 (each solve warm-started from the previous one), with one
cold-start CPU solve of the oracle and the load difference against it.  Under "two_dof": the
coupled pitch/heave driver (tests/uvlm_2dof.cpp) on the reference's wing/flap section at U_a = 7 in
FP32, time steps per second (each step runs its Newton load iterations), against the oracle's; under
"two_dof_600" the same on a 600-panel wing/flap (20 + 10 chordwise x 20 spanwise, half span 4).

--probe runs one served simulation without graph replay and one 1 GiB device copy (the
calibration kernel for HBM byte counters), for profiler counter passes.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FLOPS_PER_PAIR = 398  # algorithmic flops per (receiver, ring) pair, both mirror images
PEAK_FP64_VALU_TFLOPS = 78.6  # MI355X FP64 vector peak (also non-packed FP32)
COPY_PROBE_BYTES = 1 << 30
# HBM bytes per k_wake_lattice launch of this workload (mean over its 499 launches), from the
# FETCH_SIZE / WRITE_SIZE counter passes over `bench.py --probe`, each scaled by the same counter's
# reading for a 1 GiB device copy (FETCH x 2.000, WRITE x 1.000); and the executed FP64 flops per
# pair from the SQ_INSTS_VALU_{ADD,MUL,FMA,TRANS}_F64 pass.  Counters cannot be read in-process.
# "exact": the served loop without far-wake aggregation, measured when each launch still chose its
# rows per tile alone (38.6 us per launch then, 38.0 us with the joint launch shape).
# "*_sims64": one plan of 64 simulations (the bench step), each launch covering all 64, over
# `bench.py --probe` and `bench.py --probe --exact-wake` (one batch of 64).
# The executed flops per pair were measured again after a segment and its image came to share their
# y-free terms and the 1/(4 pi) factor moved out of the segment loop (the traffic passes precede
# those changes, which moved no memory access).
WAKE_COUNTERS = {
    "exact": {"traffic_bytes": 3008144, "algorithmic_bytes": 349384, "executed_flops_per_pair": 237.9},
    "far": {"traffic_bytes": 2465922, "algorithmic_bytes": 188327, "executed_flops_per_pair": 284.9},
    "far_sims64": {"traffic_bytes": 121996697, "algorithmic_bytes": 12052939, "executed_flops_per_pair": 226.0},
    "exact_sims64": {"traffic_bytes": 174869909, "algorithmic_bytes": 22360576, "executed_flops_per_pair": 222.7},
}
NS, NC, HALF_SPAN, CHORD = 40, 10, 2.0, 1.0
T_STEPS = 501  # UVLM::run loops t_steps-1 times -> 500 time steps
# Far-wake aggregation of the served loop (tvlm_uvlm_set_far_wake): ring rows older than 80 steps in
# blocks of 2 rows x 4 spans, older than 160 in 4 x 8, older than 320 in 8 x 8.  On this workload the
# oracle's restatement moves CL by at most 2.65e-7 over the 500 steps (bar 1e-6); every run checks
# the served result against the plain FP64 exact wake again.
FAR_WAKE = [(80, 2, 4), (160, 4, 8), (320, 8, 8)]
A_AMP = float(np.float32(2) * np.float32(3.141593) / np.float32(180))
MOTION = dict(u=1.0, h_amp=0.1, h_phase=0.0, a_amp=A_AMP, a_phase=math.pi / 2, omega=1.0, x_pivot=0.25)


def rectangular_wing_f32(ns, nc, chord, half_span):
    """Flat rectangular half wing as a plot3d file would hold it (float32), quarter-chord shifted
    in float32 like every reference mesh (vlm/src/vlm_mesh.cpp:15-32), widened to float64."""
    f = np.float32
    v = np.zeros((ns + 1, nc + 1, 4), dtype=f)
    ys = np.linspace(0.0, half_span, ns + 1)
    xs = np.linspace(0.0, chord, nc + 1)
    for i in range(ns + 1):
        v[i, :, 0] = xs
        v[i, :, 1] = ys[i]
    v[:, :, 3] = 1
    for j in range(nc):
        v[:, j, :3] = f(0.75) * v[:, j, :3] + f(0.25) * v[:, j + 1, :3]
    v[:, nc, :3] = f(4.0 / 3.0) * v[:, nc, :3] - f(1.0 / 3.0) * v[:, nc - 1, :3]
    return np.asfortranarray(v.astype(np.float64))


def run_plan(plan, k, use_graph, motion=True):
    """k whole simulations (or batches of them) one after another; motion=True evaluates and uploads the
    motion before each (the single-simulation form), False keeps the motion already on the device."""
    import torch

    lat = []
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    with torch.cuda.stream(plan.stream):
        start.record(plan.stream)
        for _ in range(k):
            t1 = time.perf_counter()
            if motion:
                plan.set_pitch_heave(**MOTION)
            plan.run(use_graph=use_graph)
            plan.stream.synchronize()
            lat.append((time.perf_counter() - t1) * 1e3)
        stop.record(plan.stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    return wall, start.elapsed_time(stop) / 1e3, lat


def cpu_baseline(verts, budget_steps):
    from oracle import cpuref as C

    mot = [MOTION[k] for k in ("u", "h_amp", "h_phase", "a_amp", "a_phase", "omega", "x_pivot")]
    C.uvlm_run([verts], mot, t_steps=3)  # warm-up (thread pool)
    t0 = time.perf_counter()
    cl, _, _, n = C.uvlm_run([verts], mot, t_steps=budget_steps + 1)
    el = time.perf_counter() - t0
    return budget_steps / el, el, C.lib().orc_num_threads(), cl


STEADY_MESH = os.path.join(ROOT, "mesh", "elliptic_64x64.x")


def steady_alphas(batch):
    """alpha = 5 deg first (the steady configuration), then U[-5, 10] deg, PCG64 seed 0; float32
    degrees -> radians as the reference's to_radians (vlm/include/vlm_utils.hpp:68-70)."""
    deg = np.random.default_rng(0).uniform(-5.0, 10.0, batch).astype(np.float32)
    deg[0] = 5.0
    return [float(np.float32(d) * np.float32(3.141593) / np.float32(180)) for d in deg]


def steady_bench(batch, reps, plain_n, with_cpu, lanes):
    """Steady elliptic 64 x 64.  Served: lanes x batch x reps requests (alpha from steady_alphas) through
    the serving loop (`lanes` lanes, batches of up to `batch`), all waiting at the start; value =
    requests / (last completion - start), latency = completion - start per request.  one_plan: the same
    batches through one served plan, one after another.  Plain: plain_n requests through the plain FP64
    plan.  Accuracy: the served results of the first plain_n requests against plain FP64."""
    from tinyvlm_amd import serve as S
    from tinyvlm_amd.mesh_io import read_plot3d
    from tinyvlm_amd.steady_plan import MODE_PLAIN_FP64, MODE_SERVED, SteadyPlan

    verts = read_plot3d(STEADY_MESH)
    alphas = steady_alphas(batch)
    served = SteadyPlan([verts], batch, mode=MODE_SERVED)
    served.solve(alphas)
    served.results(gamma=False)
    lat = []
    t0 = time.perf_counter()
    for _ in range(reps):
        t1 = time.perf_counter()
        served.solve(alphas)
        res_s = served.results(gamma=True)
        lat.append((time.perf_counter() - t1) * 1e3)
    wall_s = time.perf_counter() - t0
    served.close()
    n_req = lanes * batch * reps
    req_alphas = steady_alphas(n_req)  # its first `batch` values are `alphas`
    with S.Server({"elliptic_64": [verts]}, lanes=lanes, max_batch=batch, keep_gamma=True) as srv:
        srv.run([S.Request(-1 - k, "elliptic_64", req_alphas[k]) for k in range(lanes * batch)],
                realtime=False)  # warm-up: one plan per lane
        srv.stats = S.ServeStats()
        reqs = [S.Request(i, "elliptic_64", a) for i, a in enumerate(req_alphas)]
        st = srv.run(reqs, realtime=False)
    span = max(r.finished for r in reqs)
    p50, p99 = S.latency_percentiles(reqs)
    served_value = n_req / span
    plain = SteadyPlan([verts], plain_n, mode=MODE_PLAIN_FP64)
    plain.solve(alphas[:plain_n])
    plain.results(gamma=False)
    t0 = time.perf_counter()
    plain.solve(alphas[:plain_n])
    res_p = plain.results(gamma=True)
    wall_p = time.perf_counter() - t0
    plain.close()
    gmax = np.max(np.abs(res_p["gamma"]))
    dcl = max(float(np.max(np.abs(res_s["cl"][:plain_n] - res_p["cl"]))),
              max(abs(reqs[j].cl - float(res_p["cl"][j])) for j in range(plain_n)))
    dg = max(float(np.max(np.abs(res_s["gamma"][:plain_n] - res_p["gamma"]))),
             max(float(np.max(np.abs(reqs[j].gamma - res_p["gamma"][j]))) for j in range(plain_n))) / gmax
    out = {
        "workload": "steady_elliptic_64x64_N4096",
        "mesh": "mesh/elliptic_64x64.x", "panels": int(verts.shape[0] - 1) * int(verts.shape[1] - 1),
        "served": {"value": round(served_value, 2), "unit": "steady solves/s", "lanes": lanes, "batch": batch,
                   "requests": n_req, "mean_batch": round(float(np.mean(st.batch_sizes)), 2),
                   "max_lanes_busy": st.max_running,
                   "latency_ms_per_request": {"p50": round(p50 * 1e3, 2), "p99": round(p99 * 1e3, 2)},
                   "note": "all requests waiting at the start (capacity)"},
        "one_plan": {"value": round(batch * reps / wall_s, 2), "unit": "steady solves/s", "batch": batch,
                     "batches": reps, "latency_ms_per_batch": {"p50": round(float(np.percentile(lat, 50)), 3),
                                                               "p99": round(float(np.percentile(lat, 99)), 3)}},
        "plain_fp64": {"value": round(plain_n / wall_p, 2), "unit": "steady solves/s", "solves": plain_n,
                       "latency_ms_per_solve": round(wall_p / plain_n * 1e3, 3)},
        "speedup_of_served_path": round(served_value / (plain_n / wall_p), 3),
        "accuracy": {"max_abs_dCL_vs_plain_fp64": dcl, "max_rel_dgamma_vs_plain_fp64": dg, "bar": 1e-6,
                     "accepted": bool(dcl <= 1e-6)},
        "cl_alpha5": float(res_s["cl"][0]),
    }
    if with_cpu:
        from oracle import cpuref as C

        v = C.read_mesh(STEADY_MESH)
        C.vlm_run([C.read_mesh(os.path.join(ROOT, "mesh", "rectangular_4x4.x"))], alphas[0])  # warm-up
        t0 = time.perf_counter()
        _, cl_cpu, _, _ = C.vlm_run([v], alphas[0])
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(1.0 / el, 4), "unit": "steady solves/s", "cores": C.lib().orc_num_threads(),
                               "kind": "port", "sample": f"one oracle FP64 VLM::run at alpha 5 deg ({el:.2f} s)",
                               "abs_dCL_plain_gpu_vs_cpu": abs(float(res_p["cl"][0]) - cl_cpu)}
    return out


def serve_mix_bench(span_hi_s, span_lo_s, sat_requests, lanes):
    """Continuous batching over the committed steady-solve mix (18 meshes, N = 256 .. 16384).

    plain FP64 rate of the mix: one plain solve per (family, n) cell (the cell's first request of
    the trace), each timed after an untimed solve on the same plan; the mix is uniform over the
    cells, so the rate is 18 / (sum of the cell times).  Served: one Server (`lanes` lanes, 16 GiB plans,
    plans kept warm between runs) serves the first requests of the trace with arrivals at
    lambda = 1.5x and 0.5x that plain rate (a prefix spanning about span_hi_s / span_lo_s seconds
    of arrivals), then the first sat_requests requests all at once (capacity).  Latency = completion
    seen by the host - arrival."""
    import tempfile

    from tinyvlm_amd import serve as S
    from tinyvlm_amd.steady_plan import MODE_PLAIN_FP64, SteadyPlan

    trace = S.read_trace()
    d = tempfile.mkdtemp(prefix="tvlm_mix_")
    meshes, first = {}, {}
    reqs = []
    for i in range(len(trace["gap"])):
        fam, n = S.MIX_FAMILIES[trace["family"][i]], int(trace["n"][i])
        key = f"{fam}_{n}"
        if key not in meshes:
            meshes[key] = [S.mix_mesh(fam, n, d)]
        a = float(np.deg2rad(np.float64(trace["alpha_deg"][i])))
        first.setdefault(key, (i, a))
        reqs.append((i, key, a, float(trace["gap"][i])))
    plain_ms, plain_res = {}, {}
    for key, (i, a) in first.items():
        p = SteadyPlan(meshes[key], 1, mode=MODE_PLAIN_FP64)
        p.solve([a])
        p.results(gamma=False)
        t0 = time.perf_counter()
        p.solve([a])
        plain_res[i] = p.results(gamma=True)
        plain_ms[key] = (time.perf_counter() - t0) * 1e3
        p.close()
    plain_rate = len(first) / (sum(plain_ms.values()) / 1e3)

    def requests(count, lam):
        out, t = [], 0.0
        for rid, key, a, g in reqs[:count]:
            t += g / lam if lam else 0.0
            out.append(S.Request(rid, key, a, arrival=t))
        return out

    def summary(rs, st, lam):
        p50, p99 = S.latency_percentiles(rs)
        span = max(r.finished for r in rs) - min(r.arrival for r in rs)
        return {"requests": len(rs), "lambda_per_s": None if lam is None else round(lam, 3),
                "value": round(len(rs) / span, 2), "unit": "steady solves/s",
                "latency_ms": {"p50": round(p50 * 1e3, 2), "p99": round(p99 * 1e3, 2)},
                "batches": st.batches, "mean_batch": round(float(np.mean(st.batch_sizes)), 2),
                "max_lanes_busy": st.max_running, "plans_created": st.plans_created,
                "plans_evicted": st.plans_evicted, "peak_plan_GiB": round(st.peak_bytes / 2**30, 2)}

    out = {"workload": "steady_mix_4096_trace (18 meshes: elliptic fixture / rectangular / 25deg swept, "
                       "n x n panels, n in 16..128)",
           "trace": "tinyvlm_amd/data/steady_mix_4096.csv (PCG64 seed 0)", "lanes": lanes,
           "plain_fp64": {"value": round(plain_rate, 3), "unit": "steady solves/s",
                          "sample": "one timed plain solve per (family, n) cell",
                          "ms_per_cell": {k: round(v, 2) for k, v in sorted(plain_ms.items())}}}
    with S.Server(meshes, lanes=lanes, keep_gamma=True) as srv:
        # warm-up: one request per mesh creates the plans and sizes the lanes' scratch buffers
        srv.run([S.Request(i, k, a) for k, (i, a) in first.items()], realtime=False)
        runs = {}
        for name, factor, span in (("lambda_1.5x_plain", 1.5, span_hi_s), ("lambda_0.5x_plain", 0.5, span_lo_s)):
            lam = factor * plain_rate
            count = int(min(len(reqs), max(32, span * lam)))
            rs = requests(count, lam)
            srv.stats = S.ServeStats()
            st = srv.run(rs)
            runs[name] = summary(rs, st, lam)
            if name == "lambda_1.5x_plain":
                hi = rs
        rs = requests(min(len(reqs), sat_requests), None)
        srv.stats = S.ServeStats()
        st = srv.run(rs, realtime=False)
        runs["saturation"] = summary(rs, st, None)
        runs["saturation"]["speedup_vs_plain_fp64"] = round(runs["saturation"]["value"] / plain_rate, 3)
    out.update(runs)
    dcl, dg, info = 0.0, 0.0, 0
    for r in hi:
        if r.rid in plain_res:
            pr = plain_res[r.rid]
            dcl = max(dcl, abs(r.cl - float(pr["cl"][0])))
            dg = max(dg, float(np.max(np.abs(r.gamma - pr["gamma"][0])) / np.max(np.abs(pr["gamma"][0]))))
        info = max(info, abs(r.info))
    out["accuracy"] = {"compared_requests": sum(1 for r in hi if r.rid in plain_res),
                       "max_abs_dCL_vs_plain_fp64": dcl, "max_rel_dgamma_vs_plain_fp64": dg, "bar": 1e-6,
                       "accepted": bool(dcl <= 1e-6 and info == 0)}
    return out


def lu_probe(n, reps=2):
    """Single-matrix getrf (LU::factorize) of a random n x n matrix, FP32 and FP64, on a non-default
    stream: one untimed factorisation, then `reps` timed ones between HIP events."""
    import torch

    from tinyvlm_amd import capi as H

    out = {"probe": "getrf", "n": n}
    st = torch.cuda.Stream()
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        g = torch.Generator(device="cuda")
        g.manual_seed(0)
        a0 = torch.randn(n, n, dtype=dt, device="cuda", generator=g).t().contiguous().t()
        a = a0.clone()
        ipiv = torch.zeros(n, dtype=torch.int32, device="cuda")
        info = torch.zeros(1, dtype=torch.int32, device="cuda")
        torch.cuda.synchronize()
        with torch.cuda.stream(st):
            a.copy_(a0)
            H.getrf(a, ipiv, info, stream=st)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            ms = 0.0
            for _ in range(reps):
                a.copy_(a0)
                e0.record(st)
                H.getrf(a, ipiv, info, stream=st)
                e1.record(st)
                e1.synchronize()
                ms += e0.elapsed_time(e1)
        ms /= reps
        out[name] = {"ms": round(ms, 3), "tflops": round(2.0 / 3.0 * n ** 3 / (ms * 1e-3) / 1e12, 2),
                     "info": int(info.item())}
        del a, a0
    return out


def serve_uvlm_bench(verts, dt, cl_alone, plain_value, n_requests, lanes, far=None, batch=1):
    """Unsteady requests through the serving loop: n_requests simulations of the headline workload
    (same wing, motion and 500 steps, the headline's wake: far-wake levels `far` or exact) all waiting
    at the start, `lanes` lanes, each lane running up to `batch` simulations in one plan (hipGraph
    replay).  value = n_requests x 500 / (last completion - start); each CL(t) history is compared
    with the single served simulation of the headline measurement."""
    from tinyvlm_amd import serve as S

    reqs = [S.Request(i, "ar4_10x40", kind="unsteady", motion=dict(MOTION)) for i in range(n_requests)]
    cfg = (T_STEPS, dt, far) if far else (T_STEPS, dt)
    with S.Server({"ar4_10x40": [verts]}, lanes=lanes, unsteady={"ar4_10x40": cfg}, unsteady_batch=batch) as srv:
        # warm-up: every lane's plan of `batch` simulations created and its graph captured
        srv.run([S.Request(-1 - k, "ar4_10x40", kind="unsteady", motion=dict(MOTION)) for k in range(lanes * batch)],
                realtime=False)
        srv.stats = S.ServeStats()
        st = srv.run(reqs, realtime=False)
    span = max(r.finished for r in reqs)
    p50, p99 = S.latency_percentiles(reqs)
    value = n_requests * (T_STEPS - 1) / span
    return {"workload": "uvlm_pitch_heave_ar4_10c_x_40s_500steps, concurrent requests", "requests": n_requests,
            "lanes": lanes, "simulations_per_plan": batch, "batch_sizes": sorted(set(st.batch_sizes)),
            "value": round(value, 1), "unit": "UVLM steps/s (all requests)",
            "latency_ms_per_sim": {"p50": round(p50 * 1e3, 2), "p99": round(p99 * 1e3, 2)},
            "max_lanes_busy": st.max_running, "speedup_vs_plain_fp64": round(value / plain_value, 3),
            "wake": far if far else "exact",
            "max_abs_dCL_vs_alone": max(float(np.max(np.abs(r.cl_t - cl_alone))) for r in reqs),
            "cl_history_bitwise_equal_to_alone": bool(all(np.array_equal(r.cl_t, cl_alone) for r in reqs))}


def wake_roofline(plan, sim_s):
    """Roofline of the dominant kernel: live HIP events around each wake launch on the plan's stream."""
    wake_ms, launches, pairs = plan.time_wake()
    avg_ms = wake_ms / launches
    achieved = FLOPS_PER_PAIR * (pairs / launches) / (avg_ms * 1e-3) / 1e12
    return {
        "bound": "valu",
        "kernel": "k_wake_lattice (FP64, seeded rsqrt/rcp, shared-segment lattice)",
        "achieved": round(achieved, 3),
        "peak": PEAK_FP64_VALU_TFLOPS,
        "unit": "TFLOP/s",
        "frac": round(achieved / PEAK_FP64_VALU_TFLOPS, 4),
        "avg_launch_us": round(avg_ms * 1e3, 3),
        "launches": launches,
        "pairs_per_launch": pairs / launches,
        "flops_per_pair": FLOPS_PER_PAIR,
        "wake_share_of_sim": round(wake_ms / 1e3 / sim_s, 4),
    }


def batched_uvlm_bench(verts, dt, cl_alone, sims, reps, far):
    """One plan holding `sims` simulations of the workload (tvlm_uvlm_create_batched; sims = 1: the
    single-simulation plan), its lattice target picked by the autotuner as in the full bench, then the
    whole batch replayed `reps` times; value = sims x 500 x reps / wall.  The wake roofline is taken
    over the batched launches (every launch covers all simulations).  For profiler runs."""
    import torch

    from tinyvlm_amd import autotune as AT
    from tinyvlm_amd.uvlm_plan import MODE_FAST, UVLMPlan

    p = UVLMPlan([verts], T_STEPS, dt, mode=MODE_FAST, sims=sims)
    if far:
        p.set_far_wake(far)
    p.set_pitch_heave(**MOTION)
    tuned = AT.tune_uvlm(p, cache=AT.Cache(None))
    p.run(use_graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        p.run(use_graph=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    value = sims * (T_STEPS - 1) * reps / wall
    dcl = max(float(np.max(np.abs(p.results(sim=z)["cl"] - cl_alone))) for z in range(sims))
    rl = wake_roofline(p, wall / reps)
    p.close()
    return {"workload": "uvlm_pitch_heave_ar4_10c_x_40s_500steps, one plan of several simulations",
            "simulations": sims, "batches_timed": reps, "value": round(value, 1), "unit": "UVLM steps/s",
            "ms_per_batch": round(wall / reps * 1e3, 3), "lattice_target": tuned["best"],
            "max_abs_dCL_vs_alone": dcl, "wake": far if far else "exact", "roofline": rl}


PEAK_FP32_VALU_PACKED_TFLOPS = 157.3  # MI355X FP32 vector peak with packed FP32 (not used: see DESIGN)


def assembly_bench(reps=5):
    """Influence-matrix assembly (tvlm_lhs_assemble_*, the reference's lhs_assemble) of the elliptic
    64 x 64 mesh (N = 4096, 16.8 M receiver x ring pairs, 398 flops each) in FP32 (rsqrt form) and
    FP64 (exact form), timed with HIP events on torch's current stream."""
    import torch

    from tinyvlm_amd import capi as H
    from tinyvlm_amd.mesh_io import read_plot3d

    v = read_plot3d(STEADY_MESH)
    ns, nc = v.shape[0] - 1, v.shape[1] - 1
    N = ns * nc
    out = {"workload": "lhs_assemble_elliptic_64x64_N4096", "pairs": N * N, "flops_per_pair": FLOPS_PER_PAIR}
    exact64 = None
    for name, dt, peak in (("f32", torch.float32, PEAK_FP64_VALU_TFLOPS), ("f64", torch.float64, PEAK_FP64_VALU_TFLOPS),
                           ("f64_seeded", torch.float64, PEAK_FP64_VALU_TFLOPS)):
        verts = [H.to_device(v, dt)]
        colloc, normals, areas = [H.colmajor((ns, nc, 3), dt)], [H.colmajor((ns, nc, 3), dt)], [H.colmajor((ns, nc), dt)]
        H.mesh_metrics(0.0, verts, colloc, normals, areas)
        wake = [H.colmajor((ns + 1, 2, 4), dt)]
        lhs = H.colmajor((N, N), dt)

        def assemble():
            lhs.zero_()
            if name == "f64_seeded":
                H.lhs_assemble_seeded(lhs, colloc, normals, verts, wake, 0)
            else:
                H.lhs_assemble(lhs, colloc, normals, verts, wake, 0)

        assemble()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        z0, z1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        z0.record()
        for _ in range(reps):
            lhs.zero_()
        z1.record()
        e0.record()
        for _ in range(reps):
            assemble()
        e1.record()
        torch.cuda.synchronize()
        ms = (e0.elapsed_time(e1) - z0.elapsed_time(z1)) / reps  # the zeroing fills timed alone, subtracted
        tf = N * N * FLOPS_PER_PAIR / (ms * 1e-3) / 1e12
        out[name] = {"ms": round(ms, 4), "achieved_tflops": round(tf, 2), "peak_tflops": peak,
                     "frac": round(tf / peak, 4)}
        if name == "f64":
            exact64 = lhs.clone()
        if name == "f64_seeded":
            out[name]["max_rel_diff_vs_exact_f64"] = float((lhs - exact64).abs().max() / exact64.abs().max())
            out[name]["note"] = "the kernel the served steady path assembles with (tvlm_lhs_assemble_seeded_f64)"
        if name == "f32":
            out[name]["frac_of_packed_fp32_peak"] = round(tf / PEAK_FP32_VALU_PACKED_TFLOPS, 4)
        del lhs
    out["note"] = ("FP32 rsqrt form against the non-packed FP32 VALU rate (78.6 TFLOP/s: packed FP32 "
                   "instructions are switched off in every kernel); FP64 exact form and the FP64 seeded form (the "
                   "served steady path's) against the FP64 VALU peak")
    out["grouped"] = grouped_assembly_bench(reps)
    return out


def grouped_assembly_bench(reps=5):
    """Grouped seeded-FP64 assembly (tvlm_lhs_group_*): one influence matrix of every cell of the served
    mix (3 families x n in {16, 32, 45, 64, 90, 128}, N = 256 ... 16384, one wake row as the served
    path assembles them) in ONE launch, against the same 18 matrices assembled one call after another;
    and the same for the 9 small cells (n <= 45), where one call underfills the device."""
    import torch

    from tinyvlm_amd import capi as H
    from tinyvlm_amd import serve as S

    def setup(cells):
        items, pairs = [], 0
        for fam, n in cells:
            v = S.mix_mesh(fam, n)
            ns, nc = v.shape[0] - 1, v.shape[1] - 1
            w = np.zeros((ns + 1, 2, 4))
            w[:, 0, :] = v[:, -1, :]
            w[:, 1, :] = v[:, -1, :]
            w[:, 1, 0] += 100.0
            verts, wake = [H.to_device(v, torch.float64)], [H.to_device(w, torch.float64)]
            colloc = [H.colmajor((ns, nc, 3), torch.float64)]
            normals = [H.colmajor((ns, nc, 3), torch.float64)]
            areas = [H.colmajor((ns, nc), torch.float64)]
            H.mesh_metrics(0.05, verts, colloc, normals, areas)
            N = ns * nc
            items.append((H.colmajor((N, N), torch.float64), colloc, normals, verts, wake, 1))
            pairs += N * N
        return items, pairs

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps

    out = {}
    for tag, sizes in (("mix_18_cells", S.MIX_SIZES), ("small_9_cells", (16, 32, 45))):
        items, pairs = setup([(f, n) for f in S.MIX_FAMILIES for n in sizes])
        g = H.LhsGroup(items)
        ms_g = timed(lambda: g.assemble(seeded=True))
        ms_1 = timed(lambda: [H.lhs_assemble_seeded(*it) for it in items])
        tf = pairs * FLOPS_PER_PAIR / (ms_g * 1e-3) / 1e12
        out[tag] = {"matrices": len(items), "N": sorted({int(it[0].shape[0]) for it in items}), "pairs": pairs,
                    "grouped_ms": round(ms_g, 4), "per_matrix_calls_ms": round(ms_1, 4),
                    "grouped_tflops": round(tf, 2), "frac": round(tf / PEAK_FP64_VALU_TFLOPS, 4),
                    "speedup_of_grouping": round(ms_1 / ms_g, 3)}
        g.close()
        del items
    out["note"] = ("accumulating launches timed back to back (the values grow; the timing does not depend on "
                   "them); per_matrix_calls = tvlm_lhs_assemble_seeded_f64 per matrix, on the same stream")
    return out


def write_plot3d(path, xs, ys):
    """Single-block 2-D plot3d file of the flat grid xs (chordwise) x ys (spanwise), z = 0."""
    ni, nj = len(xs), len(ys)
    with open(path, "w") as f:
        f.write(f"1\n{ni} {nj} 1\n")
        for plane in (np.tile(xs, nj), np.repeat(ys, ni), np.zeros(ni * nj)):
            f.write("\n".join(f"{v:.17g}" for v in plane) + "\n")


HB_B = 0.127  # semi-chord of the HB binding (lib/libhbvlm3.cpp:560)
HB_H = 5
HB_OMEGA = 0.8
HB_U_REF = 6.0


def hb_served(wing, flap, X, n_requests, workers, processes=True, wake="lattice"):
    """The harmonic-balance configuration served: n_requests airspeeds U = linspace(0.5, 1.5) U_ref as
    concurrent cold-start requests through Server on `workers` HB worker processes (threads with
    processes=False; one HBVLM and stream each), against the same cold solves one after another on
    one solver; both with the wake induction `wake` (HBVLM.wake_form)."""
    from tinyvlm_amd import serve as S
    from tinyvlm_amd.hb import HBVLM

    K = 2 * HB_H + 1
    speeds = np.linspace(0.5, 1.5, n_requests) * HB_U_REF
    cfg = {"wf": {"meshes": [wing, flap], "harmonics": HB_H, "scaling": 1.0 / HB_B, "wake": wake}}

    def req(i, U):
        return S.Request(i, "wf", kind="hb", hb={"omega": HB_OMEGA, "U": float(U), "dyn": X})

    # one after another, cold (a fresh start for every airspeed): a sample of the same requests
    seq = HBVLM("hip", [wing, flap])
    seq.wake_form = wake
    seq.init(HB_H, 1.0 / HB_B)
    F = np.zeros((3, K))
    seq.run(HB_OMEGA, float(speeds[0]), X, F)  # warm-up
    sample = speeds[:: max(1, n_requests // 8)]
    seq_F = {}
    t0 = time.perf_counter()
    for U in sample:
        seq.reset()
        F = np.zeros((3, K))
        seq.run(HB_OMEGA, float(U), X, F)
        seq_F[float(U)] = F
    seq_rate = len(sample) / (time.perf_counter() - t0)
    reqs = [req(i, U) for i, U in enumerate(speeds)]
    with S.Server({}, lanes=1, hb=cfg, hb_workers=workers, hb_processes=processes) as srv:
        srv.run([req(-1 - k, speeds[0]) for k in range(2 * workers)], realtime=False)  # solvers, scratch
        t0 = time.perf_counter()
        srv.run(reqs, realtime=False)
        wall = time.perf_counter() - t0
    lat = [r.latency * 1e3 for r in reqs]
    same = all(np.array_equal(r.forces, seq_F[r.hb["U"]]) for r in reqs if r.hb["U"] in seq_F)
    return {"requests": n_requests, "workers": workers, "worker_kind": "processes" if processes else "threads",
            "wake_form": wake, "value": round(n_requests / wall, 3), "unit": "HB solves/s",
            "sequential_cold": {"value": round(seq_rate, 3), "unit": "HB solves/s", "solves": len(sample)},
            "speedup_vs_sequential_cold": round(n_requests / wall / seq_rate, 3),
            "latency_ms": {"p50": round(float(np.percentile(lat, 50)), 1), "p99": round(float(np.percentile(lat, 99)), 1)},
            "anderson_iterations_mean": round(float(np.mean([r.iterations for r in reqs])), 1),
            "served_equals_alone_bitwise": bool(same)}


def hb_bench(n_speeds, with_cpu, n_served=64, workers=5, processes=True, wake="lattice"):
    import tempfile

    import torch

    from tinyvlm_amd.hb import HBVLM

    d = tempfile.mkdtemp(prefix="tvlm_hb_")
    chord, half_span = 2 * HB_B, 4 * 2 * HB_B
    ys = np.linspace(0.0, half_span, 21)
    wing, flap = os.path.join(d, "wing_20x20.x"), os.path.join(d, "flap_10x20.x")
    write_plot3d(wing, np.linspace(0.0, chord * 2 / 3, 21), ys)
    write_plot3d(flap, np.linspace(chord * 2 / 3, chord, 11), ys)
    K = 2 * HB_H + 1
    X = np.zeros((3, K))
    X[0, 1] = 0.05 * 2.0  # heave 0.05 c = 0.1 b
    X[1, 2] = math.radians(1.0)  # pitch 1 deg
    speeds = np.linspace(0.5, 1.5, n_speeds) * HB_U_REF
    hb = HBVLM("hip", [wing, flap])

    def continuation(form):
        """Cold start at the first airspeed (also the warm-up), then the timed continuation."""
        hb.wake_form = form
        hb.init(HB_H, 1.0 / HB_B)
        F_cold = np.zeros((3, K))
        hb.run(HB_OMEGA, float(speeds[0]), X, F_cold)
        hb.init(HB_H, 1.0 / HB_B)
        torch.cuda.synchronize()
        lat, iters, loads = [], [], []
        t0 = time.perf_counter()
        for U in speeds:
            t1 = time.perf_counter()
            F = np.zeros((3, K))
            hb.run(HB_OMEGA, float(U), X, F)
            lat.append((time.perf_counter() - t1) * 1e3)
            iters.append(hb.iterations)
            loads.append(F)
        return n_speeds / (time.perf_counter() - t0), lat, iters, F_cold, loads

    rate_x, lat_x, iters_x, F_cold_x, loads_x = continuation("exact")
    rate, lat, iters, F_cold, loads = continuation(wake) if wake != "exact" else \
        (rate_x, lat_x, iters_x, F_cold_x, loads_x)
    d_forms = max(float(np.max(np.abs(a - b))) for a, b in zip(loads, loads_x))
    out = {
        "workload": "hb_rect_wing_flap_20+10c_x_20s_H5", "panels": hb.N, "harmonics": HB_H, "time_instances": K,
        "wake_form": wake, "value": round(rate, 3), "unit": "HB solves/s", "airspeeds": n_speeds,
        "latency_ms_per_solve": {"p50": round(float(np.percentile(lat, 50)), 2),
                                 "p99": round(float(np.percentile(lat, 99)), 2)},
        "anderson_iterations": iters, "dtype": "f64",
        "exact_wake": {"value": round(rate_x, 3), "unit": "HB solves/s", "anderson_iterations": iters_x,
                       "max_abs_dF_vs_this_form": d_forms},
        "note": "continuation: each airspeed warm-started from the previous solution; the Anderson map "
                "(replayed from a hipGraph) is bound by its 11 FP64 wake inductions",
    }
    if n_served:
        out["served"] = hb_served(wing, flap, X, n_served, workers, processes, wake)
    if with_cpu:
        from oracle import hb as ohb
        from oracle import cpuref as C

        o = ohb.HBVLMOracle([wing, flap])
        o.init(HB_H, 1.0 / HB_B)
        Fo = np.zeros((3, K))
        t0 = time.perf_counter()
        o.run(HB_OMEGA, float(speeds[0]), X, Fo)
        el = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": round(1.0 / el, 5), "unit": "HB solves/s", "cores": C.lib().orc_num_threads(),
                               "kind": "port", "sample": f"one cold-start oracle HBVLM::run at the first airspeed "
                                                         f"({el:.1f} s, {o.iterations} Anderson iterations)"}
        out["accuracy"] = {"max_abs_dF_cold_gpu_vs_cpu": float(np.max(np.abs(F_cold - Fo))),
                           "max_abs_dF_cold_exact_wake_gpu_vs_cpu": float(np.max(np.abs(F_cold_x - Fo))),
                           "max_abs_F": float(np.max(np.abs(Fo)))}
    return out


TWO_DOF_MESHES = [("mesh/infinite_rectangular_2x2_c0.x", False), ("mesh/infinite_rectangular_2x2_c1.x", True)]


def two_dof_bench(steps, with_cpu, large=False):
    """The reference's 2 x 2 + 2 x 2 section, or (large) a chord-1 wing/flap of 20 + 10 chordwise x
    20 spanwise panels over a half span of 4 chords (600 panels)."""
    import tempfile

    from tinyvlm_amd.two_dof import UVLM2DOF, Vars

    rel = TWO_DOF_MESHES
    if large:
        d = tempfile.mkdtemp(prefix="tvlm_2dof_")
        ys = np.linspace(0.0, 4.0, 21)
        write_plot3d(os.path.join(d, "wing_20x20.x"), np.linspace(0.0, 2.0 / 3.0, 21), ys)
        write_plot3d(os.path.join(d, "flap_10x20.x"), np.linspace(2.0 / 3.0, 1.0, 11), ys)
        rel = [(os.path.join(d, "wing_20x20.x"), False), (os.path.join(d, "flap_10x20.x"), True)]
    meshes = [(os.path.join(ROOT, p), l) for p, l in rel]
    sim = UVLM2DOF("hip", meshes, dtype=np.float32)
    sim.run(Vars(U_a=7.0), 1000.0, max_steps=5)  # warm-up
    t0 = time.perf_counter()
    r = sim.run(Vars(U_a=7.0), 1000.0, max_steps=steps)
    wall = time.perf_counter() - t0
    name = "uvlm_2dof_wing_flap_20+10c_x_20s_Ua7_fp32" if large else "uvlm_2dof_wing_flap_2x2+2x2_Ua7_fp32"
    out = {"workload": name, "panels": sim.N, "steps": steps,
           "value": round(steps / wall, 2), "unit": "coupled UVLM steps/s",
           "newton_iterations_mean": round(float(np.mean(r["iters"])), 3), "dtype": "f32",
           "note": "each step: Newmark increment + Newton load iterations, launch-bound driver"}
    if with_cpu:
        from oracle import two_dof as o2

        cwd = os.getcwd()
        os.chdir(ROOT)
        try:
            t0 = time.perf_counter()
            o = o2.run(rel, o2.Vars(U_a=7.0), 1000.0, np.float32, max_steps=steps)
            el = time.perf_counter() - t0
        finally:
            os.chdir(cwd)
        n = min(60, steps)
        from oracle import cpuref as C

        out["cpu_baseline"] = {"value": round(steps / el, 2), "unit": "coupled UVLM steps/s",
                               "cores": C.lib().orc_num_threads(),
                               "kind": "port", "sample": f"oracle FP32 restatement, {steps} steps ({el:.2f} s)"}
        out["accuracy"] = {f"max_rel_du_first_{n}_steps_vs_cpu_fp32": float(
            np.max(np.abs(r["u"][:, :n].astype(np.float64) - o["u"][:, :n])) / np.max(np.abs(o["u"][:, :n])))}
    return out


FLUTTER_PARAMS = dict(mu=20.0, omega=0.5, a_h=-0.3, x_a=0.25, r_a=0.5)  # EA at 0.35 c


def flutter_bench(n_airspeeds, steps, window, plain_steps, with_cpu):
    """The flutter sweep: rectangular wing 80 chordwise x 100 spanwise (chord 1, half span 5: 8000
    panels), wake ring of 200 rows (20000 rings once full), mass ratio 20, heave/pitch frequency ratio
    0.5, elastic axis 0.35 c, x_a 0.25, r_a 0.5, cubic pitch spring of the reference, 3-degree initial
    pitch; n_airspeeds reduced velocities over [0.6, 1.2] (onset between 0.7 and 1.0) as one resident
    batch, `steps` coupled time steps each (every step: the fixed-point iterations of the reference
    driver).  Each step is timed on its own (the host reads the convergence flags after each
    iteration).  value = airspeed-steps / s over the whole sweep; `full_ring` = the same over the steps
    at the full 200-row ring; plain FP64 = the drop-in driver (tinyvlm_amd.two_dof over the Backend
    entry points, exact Biot-Savart, LU substitution) on the first airspeed, per-step timed, compared
    over the same full-ring steps; accuracy = |dCL| of that airspeed against it."""
    from tinyvlm_amd import flutter_plan as FP
    from tinyvlm_amd.two_dof import UVLM2DOF, Vars

    R = 200
    v = FP.rectangular_wing(100, 80, 1.0, 5.0)
    speeds = FP.sweep_airspeeds(n_airspeeds)
    t0 = time.perf_counter()
    plan = FP.FlutterPlan(v, speeds, FP.params(**FLUTTER_PARAMS), wake_rows=R, max_steps=steps)
    setup_s = time.perf_counter() - t0
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        t1 = time.perf_counter()
        plan.run(1)
        lat.append((time.perf_counter() - t1) * 1e3)
    wall = time.perf_counter() - t0
    res = plan.results()
    ms_wake, pairs = plan.time_wake(5)
    plan.close()
    lat = np.asarray(lat)
    full = lat[R:R + window] if steps >= R + window else lat[-window:]
    value = n_airspeeds * steps / wall
    full_value = n_airspeeds * len(full) / (full.sum() / 1e3)
    achieved = FLOPS_PER_PAIR * pairs / (ms_wake * 1e-3) / 1e12
    nplain = min(plain_steps, steps)
    dt = res["dt"]
    drop = UVLM2DOF("hip", [(v, True)], dtype=np.float64)
    plat = []
    t0 = time.perf_counter()
    pr = drop.run(Vars(U_a=float(speeds[0]), **FLUTTER_PARAMS), t_final_nd=(nplain + 2) * dt + 0.5 * dt, wake_rows=R,
                  step_times=plat)
    plain_wall = time.perf_counter() - t0
    plat = np.asarray(plat) * 1e3
    pfull = plat[R:R + window] if nplain >= R + window else plat[-window:]
    plain_full = len(pfull) / (pfull.sum() / 1e3)
    cl = -res["F"][: nplain + 1, 0, 0] * (FP.PI_F * FLUTTER_PARAMS["mu"])
    cl_p = -pr["F"][0] * (FP.PI_F * FLUTTER_PARAMS["mu"])
    out = {
        "workload": "uvlm_2dof_flutter_sweep_rect_80c_x_100s_R200",
        "panels": 8000, "wake_rows": R, "airspeeds": n_airspeeds, "U_a": [round(float(speeds[0]), 4),
                                                                         round(float(speeds[-1]), 4)],
        "steps": steps, "dt": dt, "setup_s": round(setup_s, 3),
        "value": round(value, 2), "unit": "airspeed coupled UVLM steps/s (whole sweep)",
        "full_ring": {"value": round(full_value, 2), "unit": "airspeed steps/s", "steps": int(len(full)),
                      "speedup_vs_plain_fp64": round(full_value / plain_full, 3)},
        "latency_ms_per_step": {"p50": round(float(np.percentile(lat, 50)), 3),
                                "p99": round(float(np.percentile(lat, 99)), 3),
                                "full_ring_p50": round(float(np.percentile(full, 50)), 3)},
        "iterations_mean": round(float(res["iters"].mean()), 3),
        "plain_fp64": {"value": round(plain_full, 3), "unit": "airspeed steps/s at the full ring",
                       "steps": nplain, "wall_s": round(plain_wall, 2),
                       "note": "drop-in driver (Backend entry points), first airspeed"},
        "accuracy": {"max_abs_dCL_vs_plain_fp64": float(np.max(np.abs(cl - cl_p))), "bar": 1e-6,
                     "iterations_equal": bool(np.array_equal(res["iters"][:nplain, 0], pr["iters"])),
                     "accepted": bool(np.max(np.abs(cl - cl_p)) <= 1e-6)},
        "roofline": {"bound": "valu", "kernel": "k_fl_wake<4> (FP64 seeded lattice, ring buffer, all airspeeds)",
                     "achieved": round(achieved, 3), "peak": PEAK_FP64_VALU_TFLOPS, "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_FP64_VALU_TFLOPS, 4), "avg_launch_us": round(ms_wake * 1e3, 1),
                     "pairs_per_launch": pairs, "flops_per_pair": FLOPS_PER_PAIR,
                     "wake_share_of_step": round(2 * ms_wake / float(np.percentile(full, 50)), 3)},
    }
    if with_cpu:
        out["cpu_baseline"] = flutter_cpu_sample(v, R, dt, float(res["iters"].mean()))
    return out


def flutter_cpu_sample(v, R, dt, iters_mean):
    """CPU cost of one airspeed-step of the flutter sweep at the full ring: the oracle's wake induction
    over the 200-row ring (8000 receivers x 20000 rings) and one getrs of order 8000 (timed on unit
    triangular factors: the substitution's cost does not depend on the values), times the sweep's mean
    iterations per step.  The one-time LU of the influence matrix is not part of the sample."""
    from oracle import cpuref as C

    w = np.asfortranarray(v.copy())
    w[:, :, :3] /= 0.5
    s = C.Surface(verts=w)
    C.mesh_metrics(0.0, [s])
    ns = w.shape[0] - 1
    wake = np.zeros((ns + 1, R + 1, 4), order="F")
    for r in range(R + 1):
        wake[:, r] = w[:, -1]
        wake[:, r, 0] += dt * r
    s.verts_wake = wake
    s.gamma_wake = np.asfortranarray(np.random.default_rng(0).standard_normal((ns, R)))
    n = ns * (w.shape[1] - 1)
    lu = np.asfortranarray(np.eye(n))
    piv = np.arange(1, n + 1, dtype=np.int32)  # 1-based, as LAPACK's getrs reads them
    rhs = np.zeros(n)
    C.rhs_assemble_wake_influence(np.zeros(n), [s], 2)  # warm-up (thread pool)
    t0 = time.perf_counter()
    C.rhs_assemble_wake_influence(rhs, [s], R)
    C.getrs(lu, piv, rhs)
    it_s = time.perf_counter() - t0
    step_s = it_s * iters_mean
    return {"value": round(1.0 / step_s, 4), "unit": "airspeed steps/s at the full ring",
            "cores": C.lib().orc_num_threads(), "kind": "port",
            "sample": f"one fixed-point iteration at the full {R}-row ring (oracle wake induction {n} x {ns * R} "
                      f"pairs + getrs of order {n}, {it_s:.2f} s) x {iters_mean:.2f} iterations per step; "
                      f"one-time LU excluded"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20, help="timed batches (of --uvlm-sims simulations of 500 steps)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--plain-steps", type=int, default=3, help="timed plain FP64 simulations")
    ap.add_argument("--cpu-steps", type=int, default=500, help="UVLM steps of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--exact-wake", action="store_true", help="served UVLM loop without far-wake aggregation")
    ap.add_argument("--probe", action="store_true", help="one served simulation without graph + a 1 GiB copy")
    ap.add_argument("--steady-batch", type=int, default=64, help="requests per served steady batch")
    ap.add_argument("--steady-reps", type=int, default=3, help="served steady batches per lane (and one-plan batches)")
    ap.add_argument("--steady-plain", type=int, default=4, help="timed plain FP64 steady solves")
    ap.add_argument("--no-steady", action="store_true")
    ap.add_argument("--steady-lanes", type=int, default=3, help="lanes of the served steady measurement")
    ap.add_argument("--mix-span-hi", type=float, default=20.0, help="seconds of arrivals at 1.5x the plain rate")
    ap.add_argument("--mix-span-lo", type=float, default=15.0, help="seconds of arrivals at 0.5x the plain rate")
    ap.add_argument("--mix-saturation", type=int, default=384, help="requests of the saturation run")
    ap.add_argument("--mix-lanes", type=int, default=3)
    ap.add_argument("--uvlm-lanes", type=int, default=4, help="lanes of the served_uvlm run")
    ap.add_argument("--uvlm-batch", type=int, default=8, help="simulations per plan of the served_uvlm run")
    ap.add_argument("--uvlm-sims", type=int, default=64, help="simulations per plan of the bench step (the batch)")
    ap.add_argument("--batched-only", action="store_true",
                    help="only one plan of --uvlm-sims simulations: its rate and wake roofline (for profiler runs)")
    ap.add_argument("--no-mix", action="store_true", help="skip the served mix and the getrf probe")
    ap.add_argument("--no-served-uvlm", action="store_true")
    ap.add_argument("--uvlm-requests", type=int, default=64, help="concurrent UVLM requests of the served_uvlm run")
    ap.add_argument("--lu-probe", type=int, default=0, help="only time getrf of one N x N matrix (FP32, FP64)")
    ap.add_argument("--mix-only", action="store_true", help="only the served_mix section")
    ap.add_argument("--hb-speeds", type=int, default=8, help="airspeeds of the timed HB continuation")
    ap.add_argument("--no-hb", action="store_true")
    ap.add_argument("--hb-served", type=int, default=64, help="airspeeds served as concurrent HB requests (0: none)")
    ap.add_argument("--hb-workers", type=int, default=5, help="HB worker processes of the served run")
    ap.add_argument("--hb-threads", action="store_true", help="HB workers as threads of this process")
    ap.add_argument("--hb-wake", choices=("exact", "lattice"), default="lattice",
                    help="wake induction of the HB map (the other form's continuation is timed beside it)")
    ap.add_argument("--hb-only", action="store_true", help="only the hb section")
    ap.add_argument("--two-dof-steps", type=int, default=400, help="timed coupled pitch/heave steps")
    ap.add_argument("--two-dof-large-steps", type=int, default=100, help="timed steps of the 600-panel section")
    ap.add_argument("--no-two-dof", action="store_true")
    ap.add_argument("--flutter-airspeeds", type=int, default=32, help="airspeeds of the flutter sweep")
    ap.add_argument("--flutter-steps", type=int, default=1000, help="coupled steps of the flutter sweep")
    ap.add_argument("--flutter-plain-steps", type=int, default=220, help="steps of the plain FP64 comparison")
    ap.add_argument("--no-flutter", action="store_true")
    ap.add_argument("--flutter-only", action="store_true", help="only the flutter section")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the requests are independent and would shard across GPUs by request, one process per GPU)",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from tinyvlm_amd import capi
    from tinyvlm_amd.uvlm_plan import MODE_FAST, MODE_PLAIN_FP64, UVLMPlan, time_step

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    capi.init(0)
    if args.lu_probe:
        print(json.dumps(lu_probe(args.lu_probe)))
        return
    if args.mix_only:
        print(json.dumps(serve_mix_bench(args.mix_span_hi, args.mix_span_lo, args.mix_saturation, args.mix_lanes)))
        return
    if args.hb_only:
        print(json.dumps(hb_bench(args.hb_speeds, not args.no_cpu, args.hb_served, args.hb_workers,
                                  not args.hb_threads, args.hb_wake)))
        return
    if args.flutter_only:
        print(json.dumps(flutter_bench(args.flutter_airspeeds, args.flutter_steps, 20, args.flutter_plain_steps,
                                       not args.no_cpu)))
        return
    verts = rectangular_wing_f32(NS, NC, CHORD, HALF_SPAN)
    dt = time_step(verts, MOTION["u"])
    steps_per_sim = T_STEPS - 1
    if args.batched_only:
        one = UVLMPlan([verts], T_STEPS, dt, mode=MODE_FAST)
        if not args.exact_wake:
            one.set_far_wake(FAR_WAKE)
        one.set_pitch_heave(**MOTION)
        one.run(use_graph=True)
        cl_one = one.results()["cl"]
        one.close()
        sec = batched_uvlm_bench(verts, dt, cl_one, args.uvlm_sims, 2 if args.uvlm_sims > 1 else 10,
                                 None if args.exact_wake else FAR_WAKE)
        print(json.dumps(sec))
        return

    B = args.uvlm_sims
    far = None if args.exact_wake else FAR_WAKE
    if args.probe:
        # one served batch without graph replay + a device copy (the calibration kernel of the counter passes)
        pr = UVLMPlan([verts], T_STEPS, dt, mode=MODE_FAST, sims=B)
        if far:
            pr.set_far_wake(far)
        pr.set_pitch_heave(**MOTION)
        run_plan(pr, 1, use_graph=False, motion=False)
        src = torch.ones(COPY_PROBE_BYTES // 8, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"probe": f"uvlm served batch of {B} simulations (no graph) + device copy",
                          "copy_bytes_each_way": COPY_PROBE_BYTES, "cl_last": float(pr.results()["cl"][-1])}))
        pr.close()
        return
    from tinyvlm_amd import autotune as AT

    # served path (the bench step): one plan holding B simulations of the workload -- B independent
    # requests of the same configuration -- advanced together by every launch: seeded FP64 wake induction
    # with far-wake aggregation, FP64 inverse solve, hipGraph replay.  The motion is uploaded once
    # before timing (inputs resident); each timed batch redoes the setup (pose, LHS, LU, inverse,
    # geometry of every step) and the 500 steps of all B simulations.
    fast = UVLMPlan([verts], T_STEPS, dt, mode=MODE_FAST, sims=B)
    if far:
        fast.set_far_wake(far)
    fast.set_pitch_heave(**MOTION)
    tuned = AT.tune_uvlm(fast, cache=AT.Cache(None))
    run_plan(fast, max(args.warmup, 1), use_graph=True, motion=False)
    wall, gpu_s, lat = run_plan(fast, args.steps, use_graph=True, motion=False)
    res_all = [fast.results(sim=z) for z in range(B)]
    value = B * steps_per_sim * args.steps / wall

    # plain FP64 path (exact FP64 kernels, LU substitution each step, no graph), one simulation at a time
    plain = UVLMPlan([verts], T_STEPS, dt, mode=MODE_PLAIN_FP64)
    run_plan(plain, 1, use_graph=False)
    wall_p, _, lat_p = run_plan(plain, args.plain_steps, use_graph=False)
    res_plain = plain.results()
    plain_value = steps_per_sim * args.plain_steps / wall_p
    dcl = max(float(np.max(np.abs(r["cl"] - res_plain["cl"]))) for r in res_all)
    dcm = max(float(np.max(np.abs(r["cm"] - res_plain["cm"]))) for r in res_all)

    def counters(rl, key):
        c = WAKE_COUNTERS.get(key)
        if c is None:
            rl.update({"traffic": None, "traffic_note": "no counter pass recorded for this wake form"})
            return rl
        rl.update({
            "traffic": c["traffic_bytes"],
            "traffic_unit": "bytes per launch (rocprofv3 counters, copy-calibrated)",
            "algorithmic_bytes_per_launch": c["algorithmic_bytes"],
            "executed_fp64_flops_per_pair": c["executed_flops_per_pair"],
            "executed_frac": round(rl["achieved"] * c["executed_flops_per_pair"] / FLOPS_PER_PAIR
                                   / PEAK_FP64_VALU_TFLOPS, 4),
        })
        return rl

    wkey = ("exact" if args.exact_wake else "far") + (f"_sims{B}" if B > 1 else "")
    roofline = counters(wake_roofline(fast, wall / args.steps), wkey)
    roofline["pairs_note"] = (f"pairs = receivers x wake rings evaluated per launch, over all {B} simulations "
                              "(with far-wake aggregation: the exact rings plus the coarse rings); 398 flops per "
                              "pair, the ring-by-ring count")
    wake_ms_p, launches_p, pairs_p = plain.time_wake()
    roofline["plain_fp64_wake_tflops"] = round(FLOPS_PER_PAIR * pairs_p / (wake_ms_p * 1e-3) / 1e12, 3)
    fast.close()

    # the latency path: one simulation per plan (what a lone request sees)
    one = UVLMPlan([verts], T_STEPS, dt, mode=MODE_FAST)
    if far:
        one.set_far_wake(far)
    one.set_pitch_heave(**MOTION)
    tuned1 = AT.tune_uvlm(one, cache=AT.Cache(None), apply_globally=True)
    run_plan(one, max(args.warmup, 1), use_graph=True)
    wall1, _, lat1 = run_plan(one, args.steps, use_graph=True)
    res_one = one.results()
    value1 = steps_per_sim * args.steps / wall1
    single = {"value": round(value1, 2), "unit": "UVLM steps/s", "sims_timed": args.steps,
              "speedup_vs_plain_fp64": round(value1 / plain_value, 3),
              "latency_ms_per_sim": {"p50": round(float(np.percentile(lat1, 50)), 4),
                                     "p99": round(float(np.percentile(lat1, 99)), 4)},
              "max_abs_dCL_vs_plain_fp64": float(np.max(np.abs(res_one["cl"] - res_plain["cl"]))),
              "note": "motion evaluated and uploaded before each simulation, inside the timed loop",
              "roofline": counters(wake_roofline(one, wall1 / args.steps), "exact" if args.exact_wake else "far"),
              "autotune": {"lattice_target": tuned1["best"], "sim_ms": tuned1["sim_ms"]}}
    one.close()

    # the same served batch with the exact wake (every ring evaluated), for comparison
    exact_sec = None
    if not args.exact_wake:
        ex = UVLMPlan([verts], T_STEPS, dt, mode=MODE_FAST, sims=B)
        ex.set_pitch_heave(**MOTION)
        tuned_x = AT.tune_uvlm(ex, cache=AT.Cache(None), reps=1)
        kx = max(2, args.steps // 4)
        run_plan(ex, 1, use_graph=True, motion=False)
        wall_x, _, lat_x = run_plan(ex, kx, use_graph=True, motion=False)
        res_x = [ex.results(sim=z) for z in range(B)]
        value_x = B * steps_per_sim * kx / wall_x
        exact_sec = {"value": round(value_x, 2), "unit": "UVLM steps/s", "simulations_per_batch": B,
                     "batches_timed": kx, "speedup_vs_plain_fp64": round(value_x / plain_value, 3),
                     "max_abs_dCL_vs_plain_fp64": max(float(np.max(np.abs(r["cl"] - res_plain["cl"])))
                                                      for r in res_x),
                     "ms_per_batch": round(wall_x / kx * 1e3, 3),
                     "roofline": counters(wake_roofline(ex, wall_x / kx), f"exact_sims{B}" if B > 1 else "exact"),
                     "autotune": {"lattice_target": tuned_x["best"], "batch_ms": tuned_x["sim_ms"]}}
        ex.close()

    out = {
        "metric": "steady solves/s & UVLM steps/s vs plain FP64 at |dCL|<=1e-6; p50/p99 latency",
        "value": round(value, 2),
        "unit": "UVLM steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "uvlm_pitch_heave_ar4_10c_x_40s_500steps",
            "panels": NS * NC, "chordwise": NC, "spanwise": NS, "time_steps_per_sim": steps_per_sim,
            "wake_rings_final": NS * steps_per_sim, "heave": "0.1c sin(t)", "pitch": "2deg sin(t+pi/2)",
            "reduced_frequency": 0.5, "dt": dt, "simulations_per_batch": B,
            "bench_step": f"one batch: {B} simulations of this workload (independent requests of the same "
                          "configuration, each computed in full) advanced together by one plan, 500 steps each",
            "wake": "exact" if args.exact_wake else {"far_wake_aggregation": FAR_WAKE,
                                                     "levels": "(age steps, row group, span group)"},
        },
        "plain_fp64": {"value": round(plain_value, 2), "unit": "UVLM steps/s", "sims": args.plain_steps,
                       "note": "one simulation at a time (the drop-in path)",
                       "speedup_of_served_path": round(value / plain_value, 3)},
        "accuracy": {"max_abs_dCL_vs_plain_fp64": dcl, "max_abs_dCM_vs_plain_fp64": dcm, "bar": 1e-6,
                     "accepted": bool(dcl <= 1e-6), "over": f"every one of the {B} simulations"},
        "latency_ms_per_sim": {"p50": round(float(np.percentile(lat, 50)), 4),
                               "p99": round(float(np.percentile(lat, 99)), 4),
                               "note": f"a simulation of the batch completes with its batch of {B}",
                               "single_simulation_p50": single["latency_ms_per_sim"]["p50"],
                               "plain_fp64_p50": round(float(np.percentile(lat_p, 50)), 4)},
        "gpu_event_s": round(gpu_s, 6),
        "roofline": roofline,
    }
    out["autotune"] = {"lattice_target": tuned["best"], "batch_ms_per_candidate": tuned["sim_ms"],
                       "note": "tinyvlm_amd.autotune: whole batches per work-group target; the fastest is used"}
    out["single_simulation"] = single
    if exact_sec is not None:
        out["served_exact_wake"] = exact_sec
    if not args.no_cpu:
        cpu_sps, cpu_el, threads, cl_cpu = cpu_baseline(verts, args.cpu_steps)
        n = min(len(cl_cpu), len(res_plain["cl"]))
        out["cpu_baseline"] = {
            "value": round(cpu_sps, 3), "unit": "UVLM steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle FP64 restatement of the reference CPU backend, first {args.cpu_steps} steps of the "
                      f"same workload ({cpu_el:.1f} s, OpenMP {threads} threads)",
            "max_abs_dCL_plain_gpu_vs_cpu": float(np.max(np.abs(cl_cpu[:n] - res_plain["cl"][:n]))),
        }
    else:
        out["cpu_baseline"] = None
    plain.close()
    if not args.no_served_uvlm:
        out["served_uvlm"] = serve_uvlm_bench(verts, dt, res_one["cl"], plain_value, args.uvlm_requests,
                                              args.uvlm_lanes, None if args.exact_wake else FAR_WAKE,
                                              args.uvlm_batch)
    if not args.no_steady:
        out["steady"] = steady_bench(args.steady_batch, args.steady_reps, args.steady_plain, not args.no_cpu,
                                     args.steady_lanes)
    if not args.no_steady:
        out["assembly"] = assembly_bench()
    if not args.no_mix:
        lu = lu_probe(16384)
        for name, peak in (("f32", 157.3), ("f64", 78.6)):
            lu[name]["peak_mfma_tflops"] = peak
            lu[name]["frac"] = round(lu[name]["tflops"] / peak, 4)
        lu["note"] = ("single-matrix getrf (LU::factorize) at the mix's largest size; the served N = 16384 "
                      "requests factor in FP32, the plain path in FP64; panels are latency-bound")
        out["lu"] = lu
        out["served_mix"] = serve_mix_bench(args.mix_span_hi, args.mix_span_lo, args.mix_saturation, args.mix_lanes)
    if not args.no_hb:
        out["hb"] = hb_bench(args.hb_speeds, not args.no_cpu, args.hb_served, args.hb_workers, not args.hb_threads,
                             args.hb_wake)
    if not args.no_two_dof:
        out["two_dof"] = two_dof_bench(args.two_dof_steps, not args.no_cpu)
        out["two_dof_600"] = two_dof_bench(args.two_dof_large_steps, not args.no_cpu, large=True)
    if not args.no_flutter:
        out["flutter"] = flutter_bench(args.flutter_airspeeds, args.flutter_steps, 20, args.flutter_plain_steps,
                                       not args.no_cpu)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
