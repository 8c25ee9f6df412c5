"""Serving loop: independent exoplanet-ocean runs admitted continuously onto one GPU.

Each request is a run (truncation N, KM layers, rotation rate Omega, number of steps).
Runs of the same truncation share a slot pool: device buffers of `capacity` layer columns
that hold every run's spectral state resident in HBM.  A run is admitted when a
contiguous column range is free, the total footprint of admitted runs stays within the
memory budget and (optionally) the pool's predicted step time stays within a latency
budget; requests that arrive while others are on the GPU wait in a queue and are admitted
as slots free up (continuous batching).  Each pool has its own HIP stream and steps
asynchronously: whenever a pool's previous step has completed, one set of transform
launches advances all of that pool's admitted runs, while the other pools' steps run
concurrently on their own streams.  Each column is computed independently, so a run's
result does not depend on what else is in the pool.

The step per layer is a reduced form of the hydrostatic-Boussinesq transform workload of
dogcm's HSPM solver (doc/description/ogcm/space_discretization.tex:78-120,
model/dogcm/src/dynamics/hspm_vfvm/HBEBaroc_hspm_vfvm_mod.f90:321-443):
  U, V  = calc_VorDiv2UV(vor, div) * cos(phi)             (4 inverse units)
  vor, T, S on the grid (xy_w)                            (3 inverse units)
  F_A = (f + vor) U, F_B = (f + vor) V, f = 2 Omega mu    (grid space)
  d vor/dt = -H(F_A, F_B), d div/dt = H(F_B, -F_A) - Lapla(|u|^2/2)
  d T/dt = -H(T U, T V), d S/dt = -H(S U, S V)           (9 forward units)
then forward Euler with backward-Euler hyperdiffusion 1/(1 - dt c_n)
(DOGCM_Phys_hspm_vfvm_mod.f90:636-644, LPhys_DIFF_spm_mod.f90:140-167).
16 transform units (8 pairs) per layer per step.
"""
import time

import numpy as np

from . import (COEF_BEDIFF, DshtError, Plan, VAL)

UNITS_PER_LAYER_STEP = 16
EARTH_OMEGA = 7.292115e-5
RADIUS = 6.371e6
GRIDS = {21: (64, 32), 42: (128, 64), 85: (256, 128), 170: (512, 256)}
GRAPH_COLS = 32  # step graphs are captured for active-column counts rounded up to this


class Request:
    def __init__(self, rid, nmax, km, omega_scale, nsteps, arrival, seed):
        self.rid, self.nmax, self.km = rid, nmax, km
        self.omega = EARTH_OMEGA * omega_scale
        self.nsteps, self.arrival, self.seed = nsteps, arrival, seed
        self.steps_done = 0
        self.admitted = None
        self.finished = None
        self.col0 = None
        self.step_ms = []


def poisson_stream(n, rate, seed=0x0D0C2026, nsteps=200, truncations=(21, 42, 85, 170), kms=(20, 30, 60)):
    """n requests with exponential inter-arrival times (mean 1/rate s)."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        out.append(Request(i, int(rng.choice(truncations)), int(rng.choice(kms)),
                           float(rng.uniform(0.25, 4.0)), nsteps, t, int(rng.integers(1 << 30))))
    return out


def initial_state(nmax, km, seed, out):
    """Red-spectrum vorticity/divergence/T'/S' with depth-dependent T/S means (SPMODEL order),
    generated on the device into `out` (4, L, km) on the current stream."""
    import torch

    dev = out.device
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    L = (nmax + 1) ** 2
    n = torch.floor(torch.sqrt(torch.arange(L, device=dev, dtype=torch.float64)))
    sd = torch.sqrt(0.5 * (n + 1.0) ** -3)
    st = torch.randn((4, L, km), generator=g, device=dev, dtype=torch.float64) * sd[None, :, None]
    st[0] *= 1e-5 / torch.sqrt((st[0] ** 2).sum(0)).mean()
    st[1] *= 1e-6 / torch.sqrt((st[1] ** 2).sum(0)).mean()
    st[3] *= 0.05
    st[0, 0] = 0.0
    st[1, 0] = 0.0
    z = -(torch.arange(km, device=dev, dtype=torch.float64) + 0.5) * (5000.0 / km)
    st[2, 0] = 2.0 + 23.0 * torch.exp(z / 1000.0)
    st[3, 0] = 34.5 + 1.0 * torch.exp(z / 1000.0)
    out.copy_(st)


class Pool:
    """Slot pool of one truncation: `capacity` layer columns of resident state."""

    def __init__(self, nmax, capacity, dt, dev, precision=0):
        import torch

        self.stream = torch.cuda.Stream(device=dev)
        self.graphs = {}
        self.mempool = None
        self.max_cols = capacity  # latency cap on the active columns (set by calibrate)
        im, jm = GRIDS[nmax]
        self.plan = Plan(nmax, im, jm, kmax=capacity, radius=RADIUS)
        self.plan.set_precision(precision)
        self.nmax, self.cap, self.dt = nmax, capacity, dt
        L, self.L = self.plan.lm, self.plan.lm
        f64 = dict(dtype=torch.float64, device=dev)
        self.state = torch.zeros((4, L, capacity), **f64)  # vor, div, T, S (LC layout per field)
        self.tend = torch.zeros((5, L, capacity), **f64)   # d vor, d div (H parts), T, S, KE
        self.g = torch.zeros((11, capacity, jm, im), **f64)  # u, v, vor, T, S, FA, FB, TU, TV, SU, SV
        self.omega = torch.zeros(capacity, **f64)
        mu = torch.tensor(np.sin(self.plan.y_Lat), **f64)
        self.mu = mu.view(1, jm, 1)
        self.cos = torch.sqrt((1 - mu) * (1 + mu)).view(1, jm, 1)
        # implicit hyperdiffusion (ord 8, tau 1 day) for momentum and tracers
        dvisc = self.plan.coef(COEF_BEDIFF, [dt, 0.0, -1.0, 8, 1, 86400.0])
        ddiff = self.plan.coef(COEF_BEDIFF, [dt, 0.0, -1.0, 8, 0, 86400.0])
        lapla = self.plan.coef(0)
        self.dvisc = torch.tensor(dvisc, **f64).view(L, 1)
        self.ddiff = torch.tensor(ddiff, **f64).view(L, 1)
        self.lapla_r2 = torch.tensor(lapla / RADIUS ** 2, **f64).view(L, 1)
        self.free = [(0, capacity)]  # free column ranges
        self.runs = []
        self.bytes_per_col = 8 * (9 * L + 11 * im * jm)
        torch.cuda.synchronize(dev)  # buffers are initialised before the pool stream uses them
        # one full-width step on the zero state sizes every workspace of the pool's stream
        with torch.cuda.stream(self.stream):
            self.step(capacity, self.stream.cuda_stream)
        self.stream.synchronize()

    def alloc(self, km):
        for i, (a, b) in enumerate(self.free):
            if b - a >= km:
                self.free[i] = (a + km, b)
                if self.free[i][0] == b:
                    self.free.pop(i)
                return a
        return None

    def release(self, col0, km):
        self.free.append((col0, col0 + km))
        self.free.sort()
        merged = []
        for a, b in self.free:
            if merged and merged[-1][1] == a:
                merged[-1] = (merged[-1][0], b)
            else:
                merged.append((a, b))
        self.free = merged

    def active_hi(self):
        return max((r.col0 + r.km for r in self.runs), default=0)

    def active_cols(self):
        return sum(r.km for r in self.runs)

    def load(self, req):
        """Initial state of an admitted run, written on the pool's stream so that it lands
        after any step still in flight over the same columns."""
        import torch
        with torch.cuda.stream(self.stream):
            initial_state(self.nmax, req.km, req.seed, self.state[:, :, req.col0:req.col0 + req.km])
            self.omega[req.col0:req.col0 + req.km] = req.omega

    def calibrate(self, ref_cols, factor, reps=3):
        """Step time t(c) ~ a + b c from two column counts; caps the active columns so that
        t stays within `factor` x t(ref_cols) (the step time of one run of ref_cols layers)."""
        import torch
        c1, c2 = min(ref_cols, self.cap), self.cap
        ts = []
        for c in (c1, c2):
            with torch.cuda.stream(self.stream):
                self.step(c, self.stream.cuda_stream)
                e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
                e0.record(self.stream)
                for _ in range(reps):
                    self.step(c, self.stream.cuda_stream)
                e1.record(self.stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1) / reps)
        with torch.cuda.stream(self.stream):
            self.state.zero_()
        b = max((ts[1] - ts[0]) / max(c2 - c1, 1), 1e-9)
        a = ts[0] - b * c1
        self.t_ref = ts[0]
        self.max_cols = max(ref_cols, min(self.cap, int((factor * ts[0] - a) / b)))
        return self.t_ref, self.max_cols

    def launch(self, hi):
        """Enqueue one step of columns [0, hi) on the pool's stream as a hipGraph replay.
        hi is rounded up to a multiple of GRAPH_COLS (idle columns hold zeros or a retired
        run's last state and are never read back), so a pool keeps a handful of graphs.
        The library's workspaces were sized by the full-width step at construction, so a
        capture allocates nothing."""
        import torch
        hi = min(self.cap, -(-hi // GRAPH_COLS) * GRAPH_COLS)
        g = self.graphs.get(hi)
        with torch.cuda.stream(self.stream):
            if g is None:
                g = torch.cuda.CUDAGraph()
                # one memory pool per slot pool: a step's temporaries are dead when it ends
                # and the pool's graphs never run concurrently (one stream)
                if self.mempool is None:
                    self.mempool = torch.cuda.graph_pool_handle()
                with torch.cuda.graph(g, pool=self.mempool, stream=self.stream):
                    self.step(hi, self.stream.cuda_stream)
                self.graphs[hi] = g
            g.replay()

    def warm(self):
        """Capture the step graph of every active width up front, so that no capture lands
        inside a served step.  Replays nothing: the state is untouched."""
        import torch
        for hi in range(GRAPH_COLS, self.cap + GRAPH_COLS, GRAPH_COLS):
            hi = min(self.cap, hi)
            if hi in self.graphs:
                continue
            g = torch.cuda.CUDAGraph()
            if self.mempool is None:
                self.mempool = torch.cuda.graph_pool_handle()
            with torch.cuda.stream(self.stream):
                with torch.cuda.graph(g, pool=self.mempool, stream=self.stream):
                    self.step(hi, self.stream.cuda_stream)
            self.graphs[hi] = g
        self.stream.synchronize()

    def step(self, hi, stream):
        """One step for columns [0, hi) of the pool."""
        p, dt = self.plan, self.dt
        s = self.state[:, :, :hi]
        g = self.g[:, :hi]
        u, v, zg, tg, sg, fa, fb, tu, tv, su, sv = g
        p.vordiv2uv_d(s[0], s[1], u, v, stream=stream)
        p.inv_d(VAL, s[0], zg, stream=stream)
        p.inv_d(VAL, s[2], tg, stream=stream)
        p.inv_d(VAL, s[3], sg, stream=stream)
        f = 2.0 * self.omega[:hi].view(hi, 1, 1) * self.mu
        U = u * self.cos
        V = v * self.cos
        fa.copy_((f + zg) * U)
        fb.copy_((f + zg) * V)
        tu.copy_(tg * U)
        tv.copy_(tg * V)
        su.copy_(sg * U)
        sv.copy_(sg * V)
        ke = zg  # reuse: kinetic energy per unit mass
        ke.copy_(0.5 * (u * u + v * v))
        t = self.tend[:, :, :hi]
        p.uv2vordiv_d(fa, fb, t[0], t[1], stream=stream)   # t0 = H(FB, -FA), t1 = H(FA, FB)
        p.alpha_fwd_d(tu, tv, t[2], stream=stream)
        p.alpha_fwd_d(su, sv, t[3], stream=stream)
        p.fwd_d(VAL, ke, t[4], stream=stream)
        s[0].copy_(self.dvisc * (s[0] - dt * t[1]))
        s[1].copy_(self.dvisc * (s[1] + dt * (t[0] - self.lapla_r2 * t[4])))
        s[2].copy_(self.ddiff * (s[2] - dt * t[2]))
        s[3].copy_(self.ddiff * (s[3] - dt * t[3]))


class Server:
    """Continuous-batching request loop over per-truncation slot pools, one stream each."""

    def __init__(self, capacity=None, budget_bytes=64 << 30, dt=600.0, precision=0, dev=None,
                 latency_factor=None, ref_km=60, serialize=True):
        """serialize: at most one pool step on the GPU at a time (pools take turns round robin),
        so a step never shares the GPU with another truncation's kernels; measured against the
        concurrent form (serialize=False: every pool steps on its own stream whenever its
        previous step has finished), it holds the p99 step latency near the latency cap at
        the same throughput."""
        self.dev = dev
        self.capacity = capacity or {21: 240, 42: 240, 85: 240, 170: 240}
        self.budget = budget_bytes
        self.dt, self.precision = dt, precision
        self.latency_factor, self.ref_km = latency_factor, ref_km
        self.serialize = serialize
        self._turn = 0
        self.pools = {}
        self.used = 0

    def pool(self, nmax):
        import torch

        if nmax not in self.pools:
            dev = self.dev or torch.device("cuda:0")
            p = Pool(nmax, self.capacity[nmax], self.dt, dev, self.precision)
            if self.latency_factor:
                p.calibrate(self.ref_km, self.latency_factor)
            p.warm()
            self.pools[nmax] = p
        return self.pools[nmax]

    def try_admit(self, req, now):
        pool = self.pool(req.nmax)
        need = req.km * pool.bytes_per_col
        if self.used + need > self.budget:
            return False
        if pool.runs and pool.active_cols() + req.km > pool.max_cols:
            return False
        c0 = pool.alloc(req.km)
        if c0 is None:
            return False
        req.col0, req.admitted = c0, now
        self.used += need
        pool.runs.append(req)
        pool.load(req)
        return True

    def retire(self, req, now):
        pool = self.pools[req.nmax]
        pool.runs.remove(req)
        pool.release(req.col0, req.km)
        self.used -= req.km * pool.bytes_per_col
        req.finished = now

    def run(self, requests, batched=True, time_scale=1.0, on_finish=None):
        """Serve `requests` (arrival times in seconds from start); batched=False runs them
        one at a time in arrival order.  A run's step latency is the wall time from the
        launch of its pool's step to the observed completion.  Returns summary statistics."""
        import torch

        pending = sorted(requests, key=lambda r: r.arrival)
        queue = []
        t0 = time.perf_counter()
        now = lambda: (time.perf_counter() - t0) / time_scale
        units = 0
        active = []
        max_active = 0
        inflight = {}  # nmax -> (event, launch time, runs in the step)
        while pending or queue or active:
            tnow = now()
            while pending and pending[0].arrival <= tnow:
                queue.append(pending.pop(0))
            progressed = False
            # completions
            for nmax, (ev, tl, runs) in list(inflight.items()):
                if not ev.query():
                    continue
                del inflight[nmax]
                progressed = True
                ms = 1e3 * (time.perf_counter() - tl)
                tnow = now()
                for r in runs:
                    r.step_ms.append(ms)
                    r.steps_done += 1
                    units += UNITS_PER_LAYER_STEP * r.km
                    if r.steps_done >= r.nsteps:
                        if on_finish:
                            torch.cuda.current_stream().wait_stream(self.pools[r.nmax].stream)
                            on_finish(r, self.pools[r.nmax])
                        self.retire(r, tnow)
                        active.remove(r)
            # admission (first fit over the queue; one run at a time when unbatched)
            if batched:
                for r in list(queue):
                    if self.try_admit(r, tnow):
                        queue.remove(r)
                        active.append(r)
                        progressed = True
            elif not active and queue:
                r = queue.pop(0)
                if not self.try_admit(r, tnow):
                    raise DshtError(2, "run does not fit in an empty pool")
                active.append(r)
                progressed = True
            max_active = max(max_active, len(active))
            # launches: every pool with runs and no step in flight (serialized: the next pool
            # in turn, once nothing is in flight)
            order = list(self.pools.items())
            if self.serialize and order:
                k = self._turn % len(order)
                order = order[k:] + order[:k]
            for nmax, pool in order:
                if self.serialize and inflight:
                    break
                if nmax in inflight or not pool.runs:
                    continue
                if self.serialize:
                    self._turn = (list(self.pools).index(nmax) + 1) % len(self.pools)
                runs = list(pool.runs)
                hi = pool.active_hi()
                tl = time.perf_counter()
                pool.launch(hi)
                ev = torch.cuda.Event()
                ev.record(pool.stream)
                inflight[nmax] = (ev, tl, runs)
                progressed = True
            if not progressed:
                if not inflight and not active and not queue and pending:
                    wait = (pending[0].arrival - now()) * time_scale
                    if wait > 0:
                        time.sleep(wait)
                else:
                    time.sleep(2e-5)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        lat = np.concatenate([np.array(r.step_ms) for r in requests])
        return {
            "requests": len(requests),
            "wall_s": wall,
            "pairs_per_s": units / 2 / wall,
            "p50_step_ms": float(np.percentile(lat, 50)),
            "p99_step_ms": float(np.percentile(lat, 99)),
            "mean_turnaround_s": float(np.mean([r.finished - r.arrival for r in requests])),
            "max_active": max_active,
        }
