"""Continuous batching of steady and unsteady solves over a mix of meshes: the serving form of the
hot path's callers VLM::run (vlm/src/vlm.cpp:97-151; a steady request = one (mesh, angle of attack)
solve) and UVLM::run (vlm/src/vlm.cpp:358-509; an unsteady request = one whole prescribed-motion
simulation).

Structure
- lanes: `lanes` streams (default 3), each with a hardware queue of its own (tvlm_stream_create).
  A lane runs one batch at a time; batches on different lanes run at the same time on the device.
- plans: per mesh, a small pool of served-mode plans (tvlm_vlm_create, mode 1: seeded FP64
  matrices, one batched FP32 LU, FP64 refinement), created when first needed and moved to whichever
  lane is free (tvlm_vlm_set_stream).  A plan's batch cap follows from its mesh size (batch_cap):
  min(max_batch, plan_bytes // (12 N^2)) requests (FP64 + FP32 copy of each request's matrix;
  plan_bytes default 16 GiB).
- memory: the device bytes of all plans (tvlm_vlm_footprint) stay under `budget_bytes`.  A new plan
  that would exceed it first evicts the least recently used idle plans; if that is not enough, its
  mesh waits for memory while other meshes are served.
- unsteady requests: device-resident UVLM plans (tvlm_uvlm_*, served mode, hipGraph of the time
  loop) per (mesh, lane), created on that lane when first needed and kept there (its graph and the
  scratch buffers it captured belong to the lane's stream).  With unsteady_batch = B > 1 a lane
  takes up to B waiting simulations of one mesh into one plan that holds several simulations
  (tvlm_uvlm_create_batched): every launch of the time loop advances all of them.  Plans hold a
  power of two of simulations (at most B); a batch of fewer requests fills the rest of its plan
  with copies of its first motion, whose results are dropped.
- admission: whenever a lane is free, the queue (steady: per mesh; unsteady: per mesh) whose oldest
  waiting request arrived first is served: for steady requests an idle plan of that mesh (or a new
  one) takes up to its cap of that mesh's waiting requests, in arrival order.  Requests that arrive
  while others are on the device are admitted as soon as a lane and a plan (memory) are free,
  whatever else is running.

A request's result does not depend on the batch it shared: each matrix of a batched LU is factored
independently, and the FP64 refinement drives every solution to the FP64 solution of its own FP64
system; only FP32 rounding inside the LU (panel width, GEMM kernel choice, which follow the batch
size) differs, and the refinement removes it to ~1e-14 relative.  An unsteady simulation runs the
same kernels on the same operands whatever runs beside it: its CL(t) is bitwise the alone result
when it runs in a one-simulation plan; in a plan of several simulations the setup's batched LU may
block its updates differently, and CL(t) agrees with the alone result to ~1e-13 relative.

The request mix used to measure the server is `steady_mix_trace`: 4096 requests, PCG64 seed 0,
exponential inter-arrival gaps (unit rate, scaled by 1/lambda when run), mesh family in
{elliptic fixture, rectangular, 25-degree leading-edge sweep; chord 1, half-span 5} x n x n panels
with n in {16, 32, 45, 64, 90, 128}, alpha ~ U[-5, 10] degrees drawn in float32.
"""
from __future__ import annotations

import ctypes
import math
import os
import tempfile
import threading
import time
from collections import deque
from concurrent.futures import ProcessPoolExecutor, ThreadPoolExecutor
from dataclasses import dataclass, field

import numpy as np

from . import ROOT_DIR
from . import capi as H
from .mesh_io import read_plot3d
from .steady_plan import MODE_SERVED, SteadyPlan
from .uvlm_plan import MODE_FAST, UVLMPlan

MIX_FAMILIES = ("elliptic", "rectangular", "swept25")
MIX_SIZES = (16, 32, 45, 64, 90, 128)
MIX_CHORD, MIX_HALF_SPAN, MIX_SWEEP_DEG = 1.0, 5.0, 25.0
TRACE_PATH = os.path.join(ROOT_DIR, "tinyvlm_amd", "data", "steady_mix_4096.csv")


# ---- request mix --------------------------------------------------------------------------------
def steady_mix_trace(n_requests: int = 4096, seed: int = 0) -> dict:
    """The request mix: arrays gap (unit-rate exponential inter-arrival gaps), family (index into
    MIX_FAMILIES), n (panels per side), alpha_deg (float32 degrees).  One PCG64 generator, drawn in
    the order gaps, families, sizes, angles."""
    rng = np.random.Generator(np.random.PCG64(seed))
    gap = rng.exponential(1.0, n_requests)
    fam = rng.integers(0, len(MIX_FAMILIES), n_requests)
    size = np.asarray(MIX_SIZES)[rng.integers(0, len(MIX_SIZES), n_requests)]
    alpha = rng.uniform(-5.0, 10.0, n_requests).astype(np.float32)
    return {"gap": gap, "family": fam.astype(np.int64), "n": size.astype(np.int64), "alpha_deg": alpha}


def write_trace(path: str, trace: dict) -> None:
    with open(path, "w") as f:
        f.write("rid,family,n,alpha_deg,gap\n")
        for i in range(len(trace["gap"])):
            f.write(f"{i},{MIX_FAMILIES[trace['family'][i]]},{trace['n'][i]},"
                    f"{float(trace['alpha_deg'][i])!r},{float(trace['gap'][i])!r}\n")


def read_trace(path: str = TRACE_PATH) -> dict:
    fam, n, alpha, gap = [], [], [], []
    with open(path) as f:
        header = f.readline().strip().split(",")
        if header != ["rid", "family", "n", "alpha_deg", "gap"]:
            raise ValueError(f"{path}: unexpected header {header}")
        for line in f:
            _, fa, nn, al, ga = line.strip().split(",")
            fam.append(MIX_FAMILIES.index(fa))
            n.append(int(nn))
            alpha.append(float(al))
            gap.append(float(ga))
    return {"gap": np.asarray(gap), "family": np.asarray(fam, np.int64), "n": np.asarray(n, np.int64),
            "alpha_deg": np.asarray(alpha, np.float32)}


def write_grid_plot3d(path: str, x: np.ndarray, y: np.ndarray) -> None:
    """Single-block plot3d text file of the flat grid x[i, j], y[i, j] (i chordwise, fastest;
    j spanwise), z = 0."""
    ni, nj = x.shape
    with open(path, "w") as f:
        f.write(f"1\n{ni} {nj} 1\n")
        for plane in (x.T.ravel(), y.T.ravel(), np.zeros(ni * nj)):
            f.write("\n".join(f"{v:.17g}" for v in plane) + "\n")


def mix_mesh(family: str, n: int, cache_dir: str | None = None) -> np.ndarray:
    """verts of one mesh of the mix, through the plot3d reader (float32 parse, quarter-chord shift,
    as VLM::run reads its mesh).  elliptic: the committed fixture mesh/elliptic_{n}x{n}.x;
    rectangular / swept25: n x n uniform panels, chord 1, half-span 5, leading edge at
    x = y tan(sweep)."""
    if family == "elliptic":
        return read_plot3d(os.path.join(ROOT_DIR, "mesh", f"elliptic_{n}x{n}.x"))
    if family not in ("rectangular", "swept25"):
        raise ValueError(f"unknown mesh family {family!r}")
    d = cache_dir or tempfile.gettempdir()
    path = os.path.join(d, f"tvlm_mix_{family}_{n}x{n}.x")
    xs = np.linspace(0.0, MIX_CHORD, n + 1)
    ys = np.linspace(0.0, MIX_HALF_SPAN, n + 1)
    x, y = np.meshgrid(xs, ys, indexing="ij")
    if family == "swept25":
        x = x + y * math.tan(math.radians(MIX_SWEEP_DEG))
    write_grid_plot3d(path, x, y)
    return read_plot3d(path)


# ---- harmonic-balance worker processes ----------------------------------------------------------
_HB_SOLVERS: dict = {}  # in a worker process: configuration name -> its HBVLM


def _hb_proc_init(cfgs: dict) -> None:
    """Worker-process start: the HIP context and one HBVLM per configuration (graph replay of the map
    allowed: no other thread of the process captures)."""
    from . import capi as H
    from .hb import HBVLM

    H.init(0)
    for name, cfg in cfgs.items():
        h = HBVLM("hip", list(cfg["meshes"]))
        h.wake_form = cfg.get("wake", "exact")
        h.init(int(cfg["harmonics"]), float(cfg["scaling"]))
        _HB_SOLVERS[name] = h


def _hb_proc_job(name: str, omega: float, U: float, dyn: np.ndarray):
    """One cold-start HB solve in a worker process: (forces [3, K], iterations, wall-clock start)."""
    import torch

    start = time.time()
    h = _HB_SOLVERS[name]
    h.reset()
    forces = np.zeros((3, h.K))
    with torch.cuda.stream(h.stream):
        h.run(omega, U, dyn, forces)
    return forces, int(h.iterations), start


# ---- requests and lanes -------------------------------------------------------------------------
@dataclass
class Request:
    """A steady request (alpha) or, with kind = "unsteady", one UVLM simulation of `mesh` under the
    pitch/heave motion `motion` (UVLMPlan.set_pitch_heave's arguments); its results are the CL / CM
    histories cl_t [t_steps - 1] and cm_t [t_steps - 1, 3].  With kind = "hb", one harmonic-balance
    solve of the HB configuration `mesh` at hb = {"omega", "U", "dyn"} (HBVLM.run's arguments, dyn
    the 3 x (2H+1) Fourier coefficients of the motion), started cold; its results are forces [3, 2H+1]
    (CL, CM_y about the elastic axis, the flap's CM_y at the collocation times) and iterations."""
    rid: int
    mesh: str
    alpha: float = 0.0  # radians (steady)
    arrival: float = 0.0  # seconds after the start of the run
    kind: str = "steady"
    motion: dict | None = None
    admitted: float = math.nan
    finished: float = math.nan
    batch: int = 0  # requests in the batch it ran in
    lane: int = -1
    cl: float = math.nan
    cm: np.ndarray | None = None
    gamma: np.ndarray | None = None
    info: int = 0
    cl_t: np.ndarray | None = None
    cm_t: np.ndarray | None = None
    refine: int = 0  # served refinement status bits (SteadyPlan.refinement)
    hb: dict | None = None
    forces: np.ndarray | None = None
    iterations: int = 0

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def batch_cap(N: int, max_batch: int = 32, plan_bytes: int = 16 << 30) -> int:
    """Requests per plan for a mesh of N panels: as many as plan_bytes holds (12 N^2 bytes each),
    at most max_batch.  Below 8 requests, tall matrices (N > 1024) keep the batch small enough for
    the LU's cooperative panels (a batch of b matrices with ceil(N / 512) work-groups each, at most
    64 work-groups in all); 8 or more take the batched look-ahead LU."""
    cap = int(max(1, min(max_batch, plan_bytes // (12 * N * N))))
    if N > 1024 and cap < 8:
        cap = max(1, min(cap, 64 // ((N + 511) // 512)))
    return cap


def stream_create() -> int:
    s = ctypes.c_void_p()
    H.check(H.lib().tvlm_stream_create(ctypes.byref(s)), "tvlm_stream_create")
    return int(s.value)


def stream_destroy(s: int) -> None:
    H.check(H.lib().tvlm_stream_destroy(ctypes.c_void_p(s)), "tvlm_stream_destroy")


@dataclass
class _Slot:
    plan: object  # SteadyPlan, or UVLMPlan for unsteady requests
    bytes: int
    busy: bool = False
    last_used: float = 0.0
    unsteady: bool = False


@dataclass
class ServeStats:
    makespan: float = 0.0
    batches: int = 0
    plans_created: int = 0
    plans_evicted: int = 0
    peak_bytes: int = 0
    max_running: int = 0  # most lanes busy at once
    batch_sizes: list = field(default_factory=list)
    more_corrections: int = 0  # steady requests whose refinement needed more than the fixed corrections
    fallbacks: int = 0  # steady requests solved by the FP64 LU after the refinement failed
    refactored: int = 0  # steady requests re-factored after a stalled cooperative LU panel
    unresolved: int = 0  # steady requests whose FP64 matrix is singular


class Server:
    def __init__(self, meshes: dict, lanes: int = 3, max_batch: int = 32, plan_bytes: int = 16 << 30,
                 budget_bytes: int | None = None, plans_per_mesh: int | None = None, u_inf: float = 1.0,
                 rho: float = 1.0, keep_gamma: bool = False, unsteady: dict | None = None,
                 hb: dict | None = None, hb_workers: int = 1, unsteady_batch: int = 1,
                 hb_processes: bool = False):
        """meshes: name -> list of verts arrays (one per surface, as read_plot3d returns them).
        unsteady: name -> (t_steps, dt) or (t_steps, dt, far_wake) for the meshes that take unsteady
        requests; far_wake = [(age, row_group, span_group), ...] aggregates the wake beyond `age` steps
        (UVLMPlan.set_far_wake; accept it only where it keeps |dCL| within the bar).
        unsteady_batch: most simulations of one mesh a lane runs together in one plan.
        hb: name -> {"meshes": [wing_path, flap_path], "harmonics": H, "scaling": s[, "wake": form]}
        for the HB configurations (form: HBVLM.wake_form, "exact" unless given) that take kind = "hb"
        requests; they run on hb_workers worker threads, each with its own HBVLM (and stream) per
        configuration, beside the lanes.  An HB solve's Anderson
        loop is driven from Python launch by launch, so more workers in one process contend for the
        interpreter (measured: 1 worker 8.8, 2 workers 7.6, 4 workers 6.9 solves/s).  hb_processes:
        the workers are processes instead (spawned, each with its own HIP context, solvers and graph
        of the HB map; at most 15, one GPU process each beside this one), whose Anderson loops run in
        parallel; a request's loads are the same as from a thread worker's solver."""
        if lanes < 1 or max_batch < 1 or unsteady_batch < 1:
            raise ValueError("lanes, max_batch and unsteady_batch must be positive")
        self.unsteady_batch = int(unsteady_batch)
        self.meshes = {k: [np.asarray(v, dtype=np.float64) for v in vl] for k, vl in meshes.items()}
        self.N = {k: sum((v.shape[0] - 1) * (v.shape[1] - 1) for v in vl) for k, vl in self.meshes.items()}
        self.cap = {k: batch_cap(n, max_batch, plan_bytes) for k, n in self.N.items()}
        if budget_bytes is None:
            import torch

            budget_bytes = int(0.6 * torch.cuda.mem_get_info()[1])
        self.budget = int(budget_bytes)
        self.plans_per_mesh = plans_per_mesh or lanes
        self.u_inf, self.rho, self.keep_gamma = float(u_inf), float(rho), bool(keep_gamma)
        self.lanes = [stream_create() for _ in range(lanes)]
        self.slots: dict[str, list[_Slot]] = {k: [] for k in self.meshes}
        self.unsteady = dict(unsteady or {})
        for k in self.unsteady:
            if k not in self.meshes:
                raise KeyError(f"unsteady mesh {k!r} is not among the meshes")
        self.uslots: dict[tuple[str, int, int], _Slot] = {}  # (mesh, lane, simulations) -> plan
        self.used_bytes = 0
        self.stats = ServeStats()
        self.hb = dict(hb or {})
        self._hb_pool = None
        self._hb_local = threading.local()
        self._hb_all: list = []  # every worker's HBVLM, kept until close
        self._hb_procs = bool(hb_processes)
        if self.hb:
            if hb_workers < 1 or (hb_processes and hb_workers > 15):
                raise ValueError("hb_workers must be positive (and at most 15 worker processes)")
            if hb_processes:
                import multiprocessing

                cfgs = {k: {"meshes": list(c["meshes"]), "harmonics": int(c["harmonics"]),
                            "scaling": float(c["scaling"]), "wake": str(c.get("wake", "exact"))}
                        for k, c in self.hb.items()}
                self._hb_pool = ProcessPoolExecutor(max_workers=hb_workers,
                                                    mp_context=multiprocessing.get_context("spawn"),
                                                    initializer=_hb_proc_init, initargs=(cfgs,))
            else:
                self._hb_pool = ThreadPoolExecutor(max_workers=hb_workers, thread_name_prefix="tvlm-hb")

    # -- memory ------------------------------------------------------------------------------------
    def _estimate(self, key: str) -> int:
        n, b = self.N[key], self.cap[key]
        return 12 * n * n * b + 400 * n * b

    def _uestimate(self, key: str, sims: int = 1) -> int:
        """Device bytes of an unsteady plan of `sims` simulations: per simulation the LHS and its
        inverse, every step's geometry (vertices, collocation points, normals, velocities, areas,
        rhs), the wake and its circulations, and the wake-induction partials (one row of N per
        64-span tile and step, at most)."""
        t_steps = self.unsteady[key][0]
        n = self.N[key]
        vl = self.meshes[key]
        nv = sum(v.shape[0] * v.shape[1] for v in vl)
        ns1 = sum(v.shape[0] for v in vl)
        tiles = sum(-(-(v.shape[0] - 1) // 64) for v in vl) * t_steps + 64
        per = 16 * n * n + 8 * t_steps * (3 * nv + 12 * n + 5 * ns1) + 8 * tiles * n + 512 * n
        return sims * per + (1 << 20)

    def _evict_for(self, need: int, now: float) -> bool:
        """Frees the least recently used idle plans (steady and unsteady) until `need` bytes fit the
        budget; False (nothing evicted) when even all idle plans would not make room."""
        idle = [(s.last_used, 0, k, i) for k, ss in self.slots.items() for i, s in enumerate(ss) if not s.busy]
        idle += [(s.last_used, 1, k, 0) for k, s in self.uslots.items() if not s.busy]
        idle.sort(key=lambda v: (v[0], v[1], str(v[2]), v[3]))
        victims = []
        free = self.budget - self.used_bytes
        for _, unst, k, i in idle:
            if free >= need:
                break
            victims.append((unst, k, i))
            free += (self.uslots[k] if unst else self.slots[k][i]).bytes
        if free < need:
            return False
        for unst, k, i in sorted(victims, key=lambda v: -v[2]):
            s = self.uslots.pop(k) if unst else self.slots[k].pop(i)
            s.plan.close()
            self.used_bytes -= s.bytes
            self.stats.plans_evicted += 1
        return True

    def _slot_for(self, key: str, lane: int, now: float) -> _Slot | None:
        for s in self.slots[key]:
            if not s.busy:
                return s
        if len(self.slots[key]) >= self.plans_per_mesh:
            return None
        need = self._estimate(key)
        if need > self.budget:
            raise RuntimeError(f"mesh {key}: one plan ({need} bytes) exceeds the memory budget {self.budget}")
        if self.used_bytes + need > self.budget and not self._evict_for(need, now):
            return None
        plan = SteadyPlan(self.meshes[key], self.cap[key], mode=MODE_SERVED, stream=self.lanes[lane])
        s = _Slot(plan, plan.footprint())
        self.slots[key].append(s)
        self.used_bytes += s.bytes
        self.stats.plans_created += 1
        self.stats.peak_bytes = max(self.stats.peak_bytes, self.used_bytes)
        return s

    def usims(self, waiting: int) -> int:
        """Simulations of the plan that takes `waiting` requests of one mesh: the smallest power of two
        that holds min(waiting, unsteady_batch) of them, at most unsteady_batch."""
        b = 1
        while b < min(waiting, self.unsteady_batch):
            b *= 2
        return min(b, self.unsteady_batch)

    def _uslot_for(self, key: str, lane: int, now: float, sims: int = 1) -> _Slot | None:
        s = self.uslots.get((key, lane, sims))
        if s is not None:
            return None if s.busy else s
        t_steps, dt = self.unsteady[key][:2]
        far = self.unsteady[key][2] if len(self.unsteady[key]) > 2 else None
        need = self._uestimate(key, sims)
        if need > self.budget:
            raise RuntimeError(f"mesh {key}: one unsteady plan ({need} bytes) exceeds the memory budget {self.budget}")
        if self.used_bytes + need > self.budget and not self._evict_for(need, now):
            return None
        plan = UVLMPlan(self.meshes[key], t_steps, dt, mode=MODE_FAST, stream=self.lanes[lane], sims=sims)
        if far:
            plan.set_far_wake(far)
        s = _Slot(plan, need, unsteady=True)
        self.uslots[(key, lane, sims)] = s
        self.used_bytes += need
        self.stats.plans_created += 1
        self.stats.peak_bytes = max(self.stats.peak_bytes, self.used_bytes)
        return s

    # -- harmonic-balance requests (worker threads) --------------------------------------------------
    def _hb_solver(self, name: str):
        from .hb import HBVLM

        cache = getattr(self._hb_local, "solvers", None)
        if cache is None:
            cache = self._hb_local.solvers = {}
        if name not in cache:
            cfg = self.hb[name]
            h = HBVLM("hip", list(cfg["meshes"]))
            h.wake_form = cfg.get("wake", "exact")
            h.init(int(cfg["harmonics"]), float(cfg["scaling"]))
            # no graph capture here: capturing synchronizes the device first, which another thread's
            # capture (a UVLM lane, another worker) forbids; the map's launches run one by one
            h.use_graph = False
            cache[name] = h
            self._hb_all.append(h)
        return cache[name]

    def _hb_job(self, r: Request, t0: float):
        import torch

        h = self._hb_solver(r.mesh)
        r.admitted = time.perf_counter() - t0
        h.reset()
        forces = np.zeros((3, h.K))
        # the worker's current stream is the solver's own, so that nothing here waits on the null stream
        with torch.cuda.stream(h.stream):
            h.run(float(r.hb["omega"]), float(r.hb["U"]), np.asarray(r.hb["dyn"], dtype=np.float64), forces)
        return forces, int(h.iterations)

    # -- serving loop ------------------------------------------------------------------------------
    def run(self, requests: list[Request], realtime: bool = True, poll_s: float = 20e-6) -> ServeStats:
        """Serves every request; fills in admitted / finished (seconds after the start), cl, cm,
        info (and gamma with keep_gamma; cl_t, cm_t for unsteady requests).  realtime: request r
        becomes visible at r.arrival on the wall clock; otherwise all requests wait from the start
        (a saturation run)."""
        for r in requests:
            if r.kind == "steady" and r.mesh not in self.meshes:
                raise KeyError(f"request {r.rid}: unknown mesh {r.mesh!r}")
            if r.kind == "unsteady" and (r.mesh not in self.unsteady or r.motion is None):
                raise KeyError(f"request {r.rid}: mesh {r.mesh!r} takes no unsteady requests, or no motion")
            if r.kind == "hb" and (r.mesh not in self.hb or not r.hb):
                raise KeyError(f"request {r.rid}: {r.mesh!r} is not an HB configuration, or no hb parameters")
            if r.kind not in ("steady", "unsteady", "hb"):
                raise ValueError(f"request {r.rid}: unknown kind {r.kind!r}")
        order = sorted(requests, key=lambda r: (r.arrival, r.rid)) if realtime else list(requests)
        waiting: dict[tuple[str, str], deque] = {}
        free_lanes = deque(range(len(self.lanes)))
        running: dict[int, tuple[_Slot, list[Request]]] = {}
        nxt, finished = 0, 0
        hb_running: list = []  # (future, request)
        st = self.stats
        t0 = time.perf_counter()
        t0_wall = time.time()
        while finished < len(order):
            now = time.perf_counter() - t0
            while nxt < len(order) and (not realtime or order[nxt].arrival <= now):
                r = order[nxt]
                if r.kind == "hb":  # straight to the HB workers (their pool queues the excess)
                    if self._hb_procs:
                        fut = self._hb_pool.submit(_hb_proc_job, r.mesh, float(r.hb["omega"]), float(r.hb["U"]),
                                                   np.asarray(r.hb["dyn"], dtype=np.float64))
                    else:
                        fut = self._hb_pool.submit(self._hb_job, r, t0)
                    hb_running.append((fut, r))
                else:
                    waiting.setdefault((r.kind, r.mesh), deque()).append(r)
                nxt += 1
            progressed = False
            for fut, r in [x for x in hb_running if x[0].done()]:
                res = fut.result()  # re-raises a worker's error
                r.forces, r.iterations = res[0], res[1]
                if self._hb_procs:  # the worker's start on the wall clock
                    r.admitted = res[2] - t0_wall
                r.finished = time.perf_counter() - t0
                r.batch = 1
                r.cl = float(r.forces[0, 0])
                hb_running.remove((fut, r))
                finished += 1
                progressed = True
            for lane in list(running):
                slot, batch = running[lane]
                if not slot.plan.done():
                    continue
                if slot.unsteady:
                    for j, r in enumerate(batch):
                        res = slot.plan.results(sim=j)
                        r.finished = time.perf_counter() - t0
                        r.cl_t, r.cm_t = res["cl"], res["cm"]
                        r.cl = float(res["cl"][-1])
                        r.cm = res["cm"][-1].copy()
                        if self.keep_gamma:
                            r.gamma = res["gamma"]
                    slot.busy = False
                    slot.last_used = batch[-1].finished
                    del running[lane]
                    free_lanes.append(lane)
                    finished += len(batch)
                    progressed = True
                    continue
                res = slot.plan.results(gamma=self.keep_gamma)
                ref = slot.plan.refinement()["status"]
                t = time.perf_counter() - t0
                for j, r in enumerate(batch):
                    r.finished = t
                    r.cl = float(res["cl"][j])
                    r.cm = res["cm"][j].copy()
                    r.info = int(res["info"][j])
                    r.refine = int(ref[j])
                    st.more_corrections += bool(r.refine & 1)
                    st.fallbacks += bool(r.refine & 2)
                    st.refactored += bool(r.refine & 4)
                    st.unresolved += bool(r.refine & 8)
                    if self.keep_gamma:
                        r.gamma = res["gamma"][j].copy()
                slot.busy = False
                slot.last_used = t
                del running[lane]
                free_lanes.append(lane)
                finished += len(batch)
                progressed = True
            while free_lanes:
                heads = sorted((q[0].arrival, q[0].rid, k) for k, q in waiting.items() if q)
                chosen = None
                for _, _, k in heads:
                    kind, mesh = k
                    if kind == "unsteady":
                        slot = self._uslot_for(mesh, free_lanes[0], now, self.usims(len(waiting[k])))
                    else:
                        slot = self._slot_for(mesh, free_lanes[0], now)
                    if slot is not None:
                        chosen = (k, slot)
                        break
                if chosen is None:
                    break
                (kind, mesh), slot = chosen
                lane = free_lanes.popleft()
                q = waiting[(kind, mesh)]
                if kind == "unsteady":
                    sims = slot.plan.sims
                    batch = [q.popleft() for _ in range(min(sims, len(q)))]
                    slot.busy = True
                    for j in range(sims):  # unused simulations repeat the first motion
                        slot.plan.set_pitch_heave(**batch[j if j < len(batch) else 0].motion, sim=j)
                    slot.plan.run(use_graph=True)
                    t = time.perf_counter() - t0
                    for r in batch:
                        r.admitted = t
                        r.batch, r.lane = len(batch), lane
                    running[lane] = (slot, batch)
                    st.batches += 1
                    st.batch_sizes.append(len(batch))
                    st.max_running = max(st.max_running, len(running))
                    progressed = True
                    continue
                batch = [q.popleft() for _ in range(min(self.cap[mesh], len(q)))]
                slot.plan.set_stream(self.lanes[lane])
                slot.busy = True
                slot.plan.solve(np.asarray([r.alpha for r in batch]), u_inf=self.u_inf, rho=self.rho)
                t = time.perf_counter() - t0
                for r in batch:
                    r.admitted = t
                    r.batch = len(batch)
                    r.lane = lane
                running[lane] = (slot, batch)
                st.batches += 1
                st.batch_sizes.append(len(batch))
                st.max_running = max(st.max_running, len(running))
                progressed = True
            if not progressed:
                if not running and any(waiting.values()):
                    # nothing runs, so nothing will free a lane or memory: a waiting head that cannot
                    # be placed now never will be
                    heads = sorted(k for k, q in waiting.items() if q)
                    raise RuntimeError(f"serving stalled: no plan can be placed for the waiting queues {heads} "
                                       f"(budget {self.budget} bytes, {self.used_bytes} held by idle plans)")
                if not running and not hb_running and nxt < len(order) and not any(waiting.values()):
                    time.sleep(max(0.0, order[nxt].arrival - (time.perf_counter() - t0)))
                else:
                    time.sleep(poll_s)
        st.makespan = time.perf_counter() - t0
        return st

    def close(self):
        if self._hb_pool is not None:
            self._hb_pool.shutdown(wait=True)
            self._hb_pool = None
        self._hb_all.clear()
        for ss in self.slots.values():
            for s in ss:
                s.plan.close()
            ss.clear()
        for s in self.uslots.values():
            s.plan.close()
        self.uslots.clear()
        self.used_bytes = 0
        for s in self.lanes:
            stream_destroy(s)
        self.lanes = []

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def latency_percentiles(requests: list[Request]) -> tuple[float, float]:
    lat = np.asarray([r.latency for r in requests])
    return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))
