"""Served δ⁺-SPH simulations: continuous batching on one GPU.

``BatchEngine`` wraps libtsph's batch engine (tsph_batch_*): simulations of
uneven size are admitted by their row footprint into fixed-capacity HBM
arrays and advanced together, one captured step graph per step; finished
ones are read back and released, and waiting ones take their room.

Particles of a built case are ordered along a Z-order curve of their lattice
indices, so consecutive rows form compact patches (cache locality of the
neighbour gathers).  ``serve`` runs a request stream against it: requests arrive at seeded Poisson
times, are admitted as slots and rows free up, and every batch step's
latency is recorded.  ``drop_case`` / ``impact_case`` build the two δ⁺ cases
of the served workload: a free elliptical drop (benchmarks/delta_plus/
elliptical_drop.py:34-48, generalised to radius R and strain rate A) and a
drop falling onto a pool in a walled tank (hydrostatic pool as in
spheric19.py:124-131, 4 wall layers with floor / side normals).
"""
from __future__ import annotations

import ctypes as C
import math
import time
from collections import deque
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib

STATE_ACTIVE, STATE_DONE, STATE_FAILED = 1, 2, 3


@dataclass
class DpCase:
    """One δ⁺ simulation request: initial state (numpy, FP64) and parameters."""
    pos: np.ndarray
    vel: np.ndarray
    rho: np.ndarray
    mass: np.ndarray
    wall: np.ndarray
    wall_vel: np.ndarray
    wall_normal: np.ndarray
    dx: float
    sound_speed: float
    reference_speed: float
    steps: int
    rho0: float = 1000.0
    gravity: float = 9.81
    viscosity: float = 1e-6
    diffusion: float = 0.1
    alpha: float = 0.02
    courant: float = 0.2
    kind: str = ""

    @property
    def n(self) -> int:
        return int(self.pos.shape[0])

    @property
    def n_wall(self) -> int:
        return int(self.wall.shape[0])


def _morton_order(pts: np.ndarray, dx: float) -> np.ndarray:
    """Permutation along a Z-order curve of the lattice indices: a run of
    consecutive rows is a compact 2-D patch, so the neighbours one wave of
    targets gathers stay in few cache lines."""
    if pts.shape[0] == 0:
        return np.zeros(0, dtype=np.int64)
    ij = np.floor((pts - pts.min(0)) / dx + 0.5).astype(np.uint64)
    key = np.zeros(pts.shape[0], dtype=np.uint64)
    for bit in range(24):
        key |= ((ij[:, 0] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(2 * bit + 1)
        key |= ((ij[:, 1] >> np.uint64(bit)) & np.uint64(1)) << np.uint64(2 * bit)
    return np.argsort(key, kind="stable")


def _local(case: "DpCase") -> "DpCase":
    f = _morton_order(case.pos, case.dx)
    w = _morton_order(case.wall, case.dx)
    case.pos, case.vel, case.rho, case.mass = case.pos[f], case.vel[f], case.rho[f], case.mass[f]
    if w.size:
        case.wall, case.wall_vel, case.wall_normal = case.wall[w], case.wall_vel[w], case.wall_normal[w]
    return case


def _lattice(nx: int, ny: int, dx: float, x0: float = 0.0, y0: float = 0.0) -> np.ndarray:
    ax = (np.arange(nx) + 0.5) * dx + x0
    ay = (np.arange(ny) + 0.5) * dx + y0
    return np.stack(np.meshgrid(ax, ay, indexing="ij"), -1).reshape(-1, 2)


def drop_case(n_target: int, radius: float, strain: float, viscosity: float, c_mult: float,
              steps: int) -> DpCase:
    """Free elliptical drop: u = A (x, -y), p = rho0 A^2 (R^2 - r^2) / 2."""
    dx = radius * math.sqrt(math.pi / n_target)
    cells = int(math.ceil(radius / dx))
    pts = _lattice(2 * cells, 2 * cells, dx, -cells * dx, -cells * dx)
    r2 = (pts ** 2).sum(1)
    pts = pts[r2 < radius * radius]
    r2 = (pts ** 2).sum(1)
    speed = strain * radius
    c0 = c_mult * speed
    rho0 = 1000.0
    rho = rho0 * (1.0 + 0.5 * strain * strain * (radius * radius - r2) / c0 ** 2)
    vel = pts * np.array([strain, -strain])
    empty = np.zeros((0, 2))
    return _local(DpCase(pts, vel, rho, rho * dx * dx, empty, empty.copy(), empty.copy(), dx=dx,
                         sound_speed=c0, reference_speed=speed, steps=steps, rho0=rho0, gravity=0.0,
                         viscosity=viscosity, kind="drop"))


def impact_case(n_target: int, radius: float, speed: float, viscosity: float, c_mult: float,
                steps: int, width: float = 1.0, depth: float = 0.2, layers: int = 4) -> DpCase:
    """A drop of radius R falling at speed V onto a pool of depth D in a tank."""
    g = 9.81
    area = width * depth + math.pi * radius * radius
    dx = math.sqrt(area / n_target)
    nx, ny = int(round(width / dx)), int(round(depth / dx))
    dx = width / nx
    pool = _lattice(nx, ny, dx)
    cy = ny * dx + radius + 3.0 * dx
    cx = 0.5 * width
    cells = int(math.ceil(radius / dx))
    blob = _lattice(2 * cells, 2 * cells, dx, cx - cells * dx, cy - cells * dx)
    blob = blob[((blob - [cx, cy]) ** 2).sum(1) < radius * radius]
    top = cy + radius + 2.0 * dx
    c0 = c_mult * max(speed, math.sqrt(2.0 * g * depth))
    rho0 = 1000.0
    h_pool = ny * dx
    rho_pool = rho0 * np.exp(g * (h_pool - pool[:, 1]) / c0 ** 2)
    rho_blob = np.full(blob.shape[0], rho0)
    pos = np.concatenate([pool, blob])
    rho = np.concatenate([rho_pool, rho_blob])
    vel = np.zeros_like(pos)
    vel[pool.shape[0]:, 1] = -speed
    # wall stations: `layers` rows under the floor and beside both sides, on the fluid lattice
    ny_t = int(math.ceil(top / dx))
    floor = _lattice(nx + 2 * layers, layers, dx, -layers * dx, -layers * dx)
    left = _lattice(layers, ny_t, dx, -layers * dx, 0.0)
    right = _lattice(layers, ny_t, dx, width, 0.0)
    wall = np.concatenate([floor, left, right])
    normal = np.where((wall[:, 1] < 0.0)[:, None], np.array([0.0, 1.0]), np.array([1.0, 0.0]))
    return _local(DpCase(pos, vel, rho, rho * dx * dx, wall, np.zeros_like(wall), normal, dx=dx,
                         sound_speed=c0, reference_speed=max(speed, math.sqrt(2.0 * g * depth)),
                         steps=steps, rho0=rho0, gravity=g, viscosity=viscosity, kind="impact"))


def request_stream(count: int = 512, seed: int = 512, steps: int = 500, n_min: int = 2000,
                   n_max: int = 250000):
    """The served workload: `count` drop / impact requests with log-uniform
    particle counts in [n_min, n_max], nu log-uniform in [1e-6, 1e-3] and a
    sound-speed multiplier in [10, 20], all drawn from PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        nu = math.exp(rng.uniform(math.log(1e-6), math.log(1e-3)))
        mult = rng.uniform(10.0, 20.0)
        if rng.uniform() < 0.5:
            out.append(("drop", dict(n_target=n, radius=rng.uniform(0.05, 0.2),
                                     strain=rng.uniform(2.0, 10.0), viscosity=nu, c_mult=mult,
                                     steps=steps)))
        else:
            out.append(("impact", dict(n_target=n, radius=rng.uniform(0.04, 0.1),
                                       speed=rng.uniform(0.5, 2.0), viscosity=nu, c_mult=mult,
                                       steps=steps)))
    return out


def build_case(kind: str, kw: dict) -> DpCase:
    return drop_case(**kw) if kind == "drop" else impact_case(**kw)


# device bytes one fluid row / one wall station occupy in the engine (state in
# both ping-pong sides, sweep intermediates, cell and list bookkeeping with the
# default 64 / 40 list slots per row)
FLUID_ROW_BYTES = 2 * 16 * 2 + 2 * 8 + 8 + 4 + 4 + 4 + 4 + 64 * 4 + 16 + 8 + 32 + 16 + 16 + 8 + 16 + 1 + 8 + 8
WALL_ROW_BYTES = 3 * 16 + 4 + 4 + 4 + 4 + 40 * 4 + 8 + 8 + 16 + 8


# launch shapes the autotuner tries: (sweep rows per workgroup, list-build
# rows per workgroup, XCD-chunked sweep blocks)
AUTOTUNE_CANDIDATES = ((256, 256, True), (128, 256, True), (64, 256, True), (192, 256, True),
                       (256, 128, True), (128, 128, True), (256, 256, False), (128, 256, False))


class BatchEngine:
    """Resident batch of δ⁺ simulations on one GPU (tsph_batch_*)."""

    def __init__(self, fluid: int, wall: int, sims: int = 256, cells: int | None = None,
                 fluid_entries: int | None = None, wall_entries: int | None = None, device=None,
                 sorted_lists: bool = False):
        self.device = torch.device(device) if device is not None else torch.device("cuda", 0)
        # graph capture needs a stream other than the legacy default stream
        self.stream = torch.cuda.Stream(device=self.device)
        caps = _lib.BatchCaps()
        caps.fluid, caps.wall, caps.sims = int(fluid), int(wall), int(sims)
        caps.cells = int(cells if cells is not None else fluid + wall)
        caps.fluid_entries = int(fluid_entries if fluid_entries is not None else 64 * fluid)
        caps.wall_entries = int(wall_entries if wall_entries is not None else 40 * max(wall, 1))
        caps.sorted_lists = 1 if sorted_lists else 0
        self.caps = caps
        h = C.c_void_p()
        with torch.cuda.device(self.device):
            _lib.check(_lib.lib().tsph_batch_create(C.byref(caps), self._s(), C.byref(h)))
        self._h = h
        self.cases: dict[int, DpCase] = {}
        self.launch = (256, 256, True)

    def _s(self):
        return C.c_void_p(self.stream.cuda_stream)

    def admit(self, case: DpCase):
        """Admit a simulation; returns its slot, or None when it does not fit now."""
        dev = self.device
        f = lambda a: torch.as_tensor(np.ascontiguousarray(a, np.float64)).to(dev)
        tensors = [f(case.pos), f(case.vel), f(case.rho), f(case.mass), f(case.wall),
                   f(case.wall_vel), f(case.wall_normal)]
        torch.cuda.synchronize(dev)
        c = _lib.DpCase()
        c.n, c.n_wall, c.steps = case.n, case.n_wall, int(case.steps)
        c.dx, c.sound_speed, c.reference_speed = case.dx, case.sound_speed, case.reference_speed
        c.rho0, c.gravity, c.viscosity = case.rho0, case.gravity, case.viscosity
        c.diffusion, c.alpha, c.courant = case.diffusion, case.alpha, case.courant
        slot = C.c_int32(-1)
        st = _lib.lib().tsph_batch_admit(self._h, C.byref(c), *[_lib.ptr(t) for t in tensors],
                                         C.byref(slot), self._s())
        if st == _lib.TSPH_ERR_OOM:
            return None
        _lib.check(st)
        self.cases[slot.value] = case
        return slot.value

    def step(self, n: int = 1) -> None:
        _lib.check(_lib.lib().tsph_batch_step(self._h, int(n), self._s()))

    def set_launch(self, sweep_block: int = 256, list_block: int = 256, xcd_chunks: bool = True) -> None:
        """Launch shape of the step: rows per workgroup of the four δ⁺ pair
        sweeps and of the list build (64, 128, 192 or 256), and XCD-chunked
        or round-robin row blocks for the sweeps.  Results are the same for
        every shape; the step graph is captured again at the next step."""
        _lib.check(_lib.lib().tsph_batch_launch(self._h, int(sweep_block), int(list_block),
                                                1 if xcd_chunks else 0))
        self.launch = (int(sweep_block), int(list_block), bool(xcd_chunks))

    def autotune(self, candidates=None, steps: int = 3) -> dict:
        """Times ``steps`` batch steps of the resident mix for each launch
        shape (after one untimed step that captures its graph), keeps the
        fastest and returns {"best": shape, "ms": {shape: ms per step},
        "steps": batch steps taken}.  The trial steps advance the resident
        simulations like any other step."""
        cands = list(candidates) if candidates is not None else list(AUTOTUNE_CANDIDATES)
        ms, taken = {}, 0
        for shape in cands:
            self.set_launch(*shape)
            self.step(1)                          # captures this shape's graph
            t0 = time.perf_counter()
            self.step(steps)
            ms[shape] = 1e3 * (time.perf_counter() - t0) / steps
            taken += 1 + steps
        best = min(ms, key=ms.get)
        self.set_launch(*best)
        return {"best": best, "ms": {f"{a}/{b}/{'xcd' if c else 'rr'}": v for (a, b, c), v in ms.items()},
                "steps": taken}

    def poll(self):
        buf_s = (C.c_int32 * 1024)()
        buf_t = (C.c_int32 * 1024)()
        k = C.c_int32(0)
        _lib.check(_lib.lib().tsph_batch_poll(self._h, buf_s, buf_t, 1024, C.byref(k)))
        return [(buf_s[i], buf_t[i]) for i in range(min(k.value, 1024))]

    def read(self, slot: int) -> dict:
        case = self.cases[slot]
        pos = torch.empty((case.n, 2), dtype=torch.float64, device=self.device)
        vel = torch.empty_like(pos)
        rho = torch.empty(case.n, dtype=torch.float64, device=self.device)
        steps = C.c_int64(0)
        _lib.check(_lib.lib().tsph_batch_read(self._h, slot, _lib.ptr(pos), _lib.ptr(vel), _lib.ptr(rho),
                                              C.byref(steps), self._s()))
        self.stream.synchronize()
        return {"position": pos, "velocity": vel, "density": rho, "steps": steps.value}

    def release(self, slot: int) -> None:
        _lib.check(_lib.lib().tsph_batch_release(self._h, slot, self._s()))
        self.cases.pop(slot, None)

    def footprint_bytes(self) -> int:
        return self.caps.fluid * FLUID_ROW_BYTES + self.caps.wall * WALL_ROW_BYTES

    def stats(self) -> dict:
        c = (C.c_int64 * 4)()
        _lib.check(_lib.lib().tsph_batch_stats(self._h, c))
        return {"steps": c[0], "resident": c[1], "fluid_rows": c[2], "wall_rows": c[3]}

    def close(self) -> None:
        if getattr(self, "_h", None):
            _lib.lib().tsph_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ServeReport:
    sims: int
    failed: int
    wall_s: float
    step_ms: list = field(default_factory=list)
    particle_updates: int = 0
    completion_s: list = field(default_factory=list)
    tuning: dict | None = None       # the autotuner's result, when serve ran it

    def summary(self) -> dict:
        ms = np.asarray(self.step_ms) if self.step_ms else np.zeros(1)
        return {"sims": self.sims, "failed": self.failed, "wall_s": self.wall_s,
                "sims_per_s": self.sims / self.wall_s if self.wall_s > 0 else 0.0,
                "batch_steps": len(self.step_ms),
                "step_ms_p50": float(np.percentile(ms, 50)), "step_ms_p99": float(np.percentile(ms, 99)),
                "particle_updates_per_s": self.particle_updates / self.wall_s if self.wall_s > 0 else 0.0,
                "completion_s_p50": float(np.percentile(self.completion_s, 50)) if self.completion_s else 0.0,
                "completion_s_p99": float(np.percentile(self.completion_s, 99)) if self.completion_s else 0.0,
                "autotune": self.tuning}


def serve(engine: BatchEngine, requests, rate_per_s: float, seed: int = 1, on_result=None,
          tune: bool = False, tune_steps: int = 3) -> ServeReport:
    """Continuous batching: requests (built DpCase, or (kind, kwargs) built on
    arrival) arrive at Poisson times of
    the given rate; each is admitted when a slot and rows are free, every
    active simulation advances one step per batch step, finished ones are read
    back (``on_result(index, state, result)``) and released.  With ``tune``
    the engine's launch shape is autotuned once, on the resident mix of the
    first time the engine is full (its trial steps are batch steps of the
    stream and count as such)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    arrivals = np.cumsum(rng.exponential(1.0 / rate_per_s, size=len(requests)))
    waiting = deque()
    running: dict[int, tuple[int, float]] = {}
    rep = ServeReport(sims=len(requests), failed=0, wall_s=0.0)
    nxt = 0
    t0 = time.perf_counter()
    done = 0
    while done < len(requests):
        now = time.perf_counter() - t0
        while nxt < len(requests) and arrivals[nxt] <= now:
            item = requests[nxt]
            waiting.append((nxt, arrivals[nxt], item if isinstance(item, DpCase) else build_case(*item)))
            nxt += 1
        full = False
        while waiting:
            idx, t_arr, case = waiting[0]
            slot = engine.admit(case)
            if slot is None:
                full = True
                break
            waiting.popleft()
            running[slot] = (idx, t_arr)
        if tune and full and rep.tuning is None and running:
            active_rows = sum(engine.cases[s].n for s in running)
            ts = time.perf_counter()
            rep.tuning = engine.autotune(steps=tune_steps)
            per = 1e3 * (time.perf_counter() - ts) / rep.tuning["steps"]
            rep.step_ms.extend([per] * rep.tuning["steps"])
            rep.particle_updates += active_rows * rep.tuning["steps"]
        if not running:
            if nxt < len(requests):
                time.sleep(max(0.0, min(arrivals[nxt] - (time.perf_counter() - t0), 0.01)))
            continue
        active_rows = sum(engine.cases[s].n for s in running)
        ts = time.perf_counter()
        engine.step(1)
        rep.step_ms.append(1e3 * (time.perf_counter() - ts))
        rep.particle_updates += active_rows
        for slot, state in engine.poll():
            idx, t_arr = running.pop(slot)
            if state == STATE_FAILED:
                rep.failed += 1
            if on_result is not None:
                on_result(idx, state, engine.read(slot))
            engine.release(slot)
            rep.completion_s.append(time.perf_counter() - t0 - t_arr)
            done += 1
    rep.wall_s = time.perf_counter() - t0
    return rep


def run_sequential(requests, device=None, launch=None) -> tuple[float, list]:
    """Single-case runs one after another, each on an engine sized for that
    case alone (with the launch shape ``launch``, if given) and its step
    graphs captured before the clock starts: returns (seconds, per-case
    seconds)."""
    times = []
    for kind, kw in requests:
        case = build_case(kind, kw)
        eng = BatchEngine(case.n + 8, case.n_wall + 8, sims=1, device=device)
        try:
            if launch is not None:
                eng.set_launch(*launch)
            eng.step(0)                       # capture the step graphs
            t0 = time.perf_counter()
            slot = eng.admit(case)
            if slot is None:
                raise MemoryError("case does not fit its own engine")
            while not eng.poll():
                eng.step(1)
            eng.release(slot)
            times.append(time.perf_counter() - t0)
        finally:
            eng.close()
    return sum(times), times


def run_single_case_api(case: DpCase, device=None) -> dict:
    """One case through the single-case API in the reference's own driver loop
    (benchmarks/delta_plus/elliptical_drop.py:44-57, spheric19.py:139-170):
    per step the adaptive dt from max |v|, a PairListCache(2h, 0.4 dx) lookup
    over cat(pos, wall), then delta_plus_step.  The state is on the device
    before the clock starts.  Returns the final state and the seconds taken."""
    from .delta_plus import delta_plus_step
    from .neighbor_search import PairListCache
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    f = lambda a: torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64)).to(dev)
    pos, vel, rho, mass = f(case.pos), f(case.vel), f(case.rho), f(case.mass)
    walls = {}
    if case.n_wall:
        walls = dict(wall=f(case.wall), wall_velocity=f(case.wall_vel), wall_normal=f(case.wall_normal))
    wall = walls.get("wall", pos.new_empty((0, 2)))
    cache = PairListCache(4.0 * case.dx, 0.4 * case.dx, fluid_count=case.n)
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    for _ in range(case.steps):
        vmax = float(torch.sqrt((vel * vel).sum(1)).max())
        dt = case.courant * 2.0 * case.dx / (case.sound_speed + vmax)
        pairs, _ = cache.get(torch.cat((pos, wall)))
        st = delta_plus_step(pos, vel, rho, mass, dx=case.dx, sound_speed=case.sound_speed,
                             reference_speed=case.reference_speed, dt=dt, rho0=case.rho0,
                             gravity=case.gravity, viscosity=case.viscosity, diffusion=case.diffusion,
                             alpha=case.alpha, pairs=pairs, **walls)
        pos, vel, rho = st.position, st.velocity, st.density
    torch.cuda.synchronize(dev)
    return {"position": pos, "velocity": vel, "density": rho, "seconds": time.perf_counter() - t0,
            "rebuilds": cache.rebuilds}


def run_sequential_api(requests, device=None) -> tuple[float, list]:
    """``run_single_case_api`` over the requests one after another: returns
    (seconds, per-case seconds)."""
    times = []
    for kind, kw in requests:
        times.append(run_single_case_api(build_case(kind, kw), device)["seconds"])
    return sum(times), times
