"""Serving loop: independent exoplanet-ocean runs admitted continuously onto one GPU, each
advanced by dogcm's LF-AM3 dynamics step (DOGCM_TInt_LFAM3_advance_Dyn,
model/dogcm/src/tint/DOGCM_TInt_LFAM3_mod.f90:281-880) on the device (dsht_lfam3_*).

Each request is a run: truncation N, KM layers, rotation rate Omega (it enters only the
Coriolis term, xy_Cori = 2 Omega sin(lat)) and a number of steps.  Runs of one (N, KM) share
a pool of run slots whose whole state (three time levels of U, V, temperature, salinity and
surface pressure) stays resident in HBM.  A run is admitted when its pool has a free slot, the
total footprint of admitted runs stays within the memory budget and (optionally) the pool's
step time with one more run stays within a latency cap; requests that arrive while others are
on the GPU wait in a queue and are admitted as slots free up (continuous batching).  Active
runs occupy a pool's first slots: a finished run's slot is refilled by moving the pool's last
run into it (dsht_lfam3_copy_run).  One pool step is one hipGraph replay that advances all of
the pool's active runs, each transform of the step running over every layer of every run in
one call.  Every run is computed as independent columns, so its result does not depend on
what else is in the pool.

A request may start from a restart file (`resume_from`: its state and rotation rate replace
the seeded initial state) and may leave one when it finishes (`save_to`), in dogcm's restart
layout (dsht/restart.py; DOGCM_IO_Restart_mod.f90), so a served run can be stopped and resumed
later, in another pool or process.

A step makes (42 KM + 9) forward and (28 KM + 9) inverse transform units per run
(dsht_lfam3_units); pairs/s = units / 2 / s.
"""
import time

import numpy as np

from . import (DshtError, LF_B, LF_N, LF_SFCPRES, LF_TRC, LF_U, LF_V, LfAm3, Plan, VAL, lfam3_config)

EARTH_OMEGA = 7.292115e-5
RADIUS = 6.371e6
GRIDS = {21: (64, 32), 42: (128, 64), 85: (256, 128), 170: (512, 256)}
# time step per truncation (the explicit advection and internal-wave limits scale with the
# grid spacing)
DT = {21: 3600.0, 42: 2400.0, 85: 1200.0, 170: 600.0}
CAPACITY = {21: 16, 42: 16, 85: 8, 170: 6}


class Request:
    def __init__(self, rid, nmax, km, omega_scale, nsteps, arrival, seed, resume_from=None, save_to=None):
        self.rid, self.nmax, self.km = rid, nmax, km
        self.omega = EARTH_OMEGA * omega_scale
        self.nsteps, self.arrival, self.seed = nsteps, arrival, seed
        self.resume_from, self.save_to = resume_from, save_to
        self.model_time = 0.0  # seconds since the run's start (from the restart file if resumed)
        self.steps_done = 0
        self.admitted = None
        self.finished = None
        self.slot = None
        self.step_ms = []

    def units(self):
        return (42 * self.km + 9) + (28 * self.km + 9)


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


def initial_state(pool, slot, seed, depth=5000.0):
    """A seeded red-spectrum state written into `slot` of `pool` (levels B = N), on the
    current stream: velocities from random vorticity/divergence spectra (variance ~ (n+1)^-3)
    scaled to 5 cm/s RMS, temperature 25 -> 2 C [K] and salinity 35 -/+ 0.5 psu depth
    backgrounds (z at the layer centres of the Chebyshev-Lobatto sigma grid) with 1 K and
    0.05 psu red-spectrum perturbations; zero surface pressure."""
    import torch

    p = pool.plan
    km = pool.km
    dev = torch.device("cuda")
    g = torch.Generator(device=dev)
    g.manual_seed(int(seed))
    L = p.lm
    n = torch.floor(torch.sqrt(torch.arange(L, device=dev, dtype=torch.float64)))
    sd = torch.sqrt(0.5 * (n + 1.0) ** -3)[:, None]
    spec = torch.randn((L, 4 * km), generator=g, device=dev, dtype=torch.float64) * sd
    spec[:, km:2 * km] *= 0.1
    spec[0] = 0.0
    u = torch.empty((km, p.jm, p.im), dtype=torch.float64, device=dev)
    v = torch.empty_like(u)
    vor, div = spec[:, :km], spec[:, km:2 * km]
    p.vordiv2uv_d(vor.contiguous(), div.contiguous(), u, v)
    scale = 0.05 / torch.sqrt((u * u + v * v).mean())
    u *= scale
    v *= scale
    pert = torch.empty((2 * km, p.jm, p.im), dtype=torch.float64, device=dev)
    p.inv_d(VAL, spec[:, 2 * km:].contiguous(), pert)
    fk = -0.5 + 0.5 * torch.cos(torch.pi * torch.arange(km + 1, device=dev, dtype=torch.float64) / km)
    z = depth * 0.5 * (fk[:-1] + fk[1:])
    t = 273.15 + 2.0 + 23.0 * torch.exp(z / 1000.0)[:, None, None] + pert[:km] / pert[:km].std()
    s = 34.5 + 1.0 * torch.exp(z / 1000.0)[:, None, None] + 0.05 * pert[km:] / pert[km:].std()
    for lev in (LF_B, LF_N):
        pool.lf.run_view(LF_U, lev, slot).copy_(u)
        pool.lf.run_view(LF_V, lev, slot).copy_(v)
        trc = pool.lf.run_view(LF_TRC, lev, slot)
        trc[0].copy_(t)
        trc[1].copy_(s)
        pool.lf.run_view(LF_SFCPRES, lev, slot).zero_()


class Pool:
    """Run slots of one (truncation, layer count): an LF-AM3 pool on its own stream."""

    def __init__(self, plan, km, capacity, dt, dev, cfg_kw=None):
        import torch

        self.stream = torch.cuda.Stream(device=dev)
        self.graphs = {}
        self.plan, self.nmax, self.km, self.cap = plan, plan.nmax, km, capacity
        self.max_runs = capacity  # latency cap on the active runs (set by calibrate)
        self.lf = LfAm3(plan, lfam3_config(km=km, max_runs=capacity, dt=dt, **(cfg_kw or {})))
        # state, constants and work of one slot: about 41 grids per layer plus spectra
        self.bytes_per_run = 8 * (41 * km * plan.im * plan.jm + 12 * km * plan.lm)
        self.runs = []
        self.t_ref = None
        self.t_per_run = None
        torch.cuda.synchronize(dev)
        # one full-width step sizes every workspace of the pool's stream (captures allocate
        # nothing afterwards)
        self.lf.step(capacity, self.stream.cuda_stream)
        self.stream.synchronize()

    def calibrate(self, factor, reps=3):
        """Step times with one run and with every slot; caps the active runs so that a step
        stays within `factor` x the one-run step time (t(n) ~ a + b n)."""
        import torch
        ts = []
        for n in (1, self.cap):
            self.lf.step(n, self.stream.cuda_stream)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(self.stream)
            for _ in range(reps):
                self.lf.step(n, self.stream.cuda_stream)
            e1.record(self.stream)
            e1.synchronize()
            ts.append(e0.elapsed_time(e1) / reps)
        b = max((ts[1] - ts[0]) / max(self.cap - 1, 1), 1e-9)
        a = ts[0] - b
        self.t_ref = ts[0]
        self.t_per_run = b  # marginal step time of one more run in a full pool
        self.max_runs = max(1, min(self.cap, int((factor * ts[0] - a) / b)))
        return self.t_ref, self.max_runs

    def warm(self):
        """Capture the step graph of every active count up front, so that no capture lands
        inside a served step.  Replays nothing: the state is untouched."""
        for n in range(1, self.cap + 1):
            self._graph(n)
        self.stream.synchronize()

    def _graph(self, n):
        import torch
        g = self.graphs.get(n)
        if g is None:
            g = torch.cuda.CUDAGraph()
            with torch.cuda.stream(self.stream):
                with torch.cuda.graph(g, stream=self.stream):
                    self.lf.step(n, self.stream.cuda_stream)
            self.graphs[n] = g
        return g

    def launch(self):
        """Enqueue one step of the active runs on the pool's stream (a graph replay)."""
        import torch
        g = self._graph(len(self.runs))
        with torch.cuda.stream(self.stream):
            g.replay()

    def admit(self, req):
        import torch
        req.slot = len(self.runs)
        self.runs.append(req)
        if req.resume_from:
            from .restart import read_restart, load_run
            self.stream.synchronize()
            with torch.cuda.stream(self.stream):
                req.model_time = load_run(self.lf, req.slot, req.resume_from)
            attrs = read_restart(req.resume_from)[3]
            if "omega" in attrs:
                req.omega = float(attrs["omega"])
            else:
                self.lf.set_rotation(req.slot, req.omega)
            return
        with torch.cuda.stream(self.stream):
            initial_state(self, req.slot, req.seed)
        self.lf.set_rotation(req.slot, req.omega)

    def save(self, req):
        """Writes req's restart file (its state now, both time levels)."""
        from .restart import save_run
        self.stream.synchronize()
        save_run(self.lf, req.slot, req.save_to, time=req.model_time + req.steps_done * self.lf.cfg.dt,
                 omega=req.omega)

    def retire(self, req):
        """Frees req's slot; the pool's last run moves into it (state and rotation rate)."""
        last = self.runs[-1]
        if last is not req:
            self.lf.copy_run(req.slot, last.slot, self.stream.cuda_stream)
            last.slot = req.slot
            self.runs[req.slot] = last
        self.runs.pop()


class Server:
    """Continuous-batching request loop over (truncation, KM) pools, one stream each."""

    def __init__(self, capacity=None, budget_bytes=160 << 30, dev=None, latency_factor=None, serialize=True,
                 dt=None, cfg_kw=None):
        """serialize: one pool step on the GPU at a time (pools take turns round robin), so a
        step never shares the GPU with another pool's kernels and its latency is its own."""
        self.dev = dev
        self.capacity = capacity or dict(CAPACITY)
        self.budget = budget_bytes
        self.latency_factor = latency_factor
        self.serialize = serialize
        self.dt = dt or dict(DT)
        self.cfg_kw = cfg_kw
        self._turn = 0
        self.plans = {}
        self.pools = {}
        self.used = 0

    def pool(self, nmax, km):
        import torch

        key = (nmax, km)
        if key not in self.pools:
            dev = self.dev or torch.device("cuda:0")
            if nmax not in self.plans:
                self.plans[nmax] = Plan(nmax, *GRIDS[nmax], radius=RADIUS)
            p = Pool(self.plans[nmax], km, self.capacity[nmax], self.dt[nmax], dev, self.cfg_kw)
            if self.latency_factor:
                p.calibrate(self.latency_factor)
            p.warm()
            self.pools[key] = p
        return self.pools[key]

    def try_admit(self, req, now, max_runs=None):
        pool = self.pool(req.nmax, req.km)
        if self.used + pool.bytes_per_run > self.budget:
            return False
        cap = pool.max_runs if max_runs is None else max_runs
        if len(pool.runs) >= min(cap, pool.cap):
            return False
        req.admitted = now
        self.used += pool.bytes_per_run
        pool.admit(req)
        return True

    def retire(self, req, now):
        pool = self.pools[(req.nmax, req.km)]
        pool.retire(req)
        self.used -= pool.bytes_per_run
        req.finished = now

    def run(self, requests, batched=True, time_scale=1.0, on_start=None, on_finish=None):
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
        inflight = {}  # pool key -> (event, launch time, runs in the step)
        while pending or queue or active:
            tnow = now()
            while pending and pending[0].arrival <= tnow:
                queue.append(pending.pop(0))
            progressed = False
            # completions
            for key, (ev, tl, runs) in list(inflight.items()):
                if not ev.query():
                    continue
                del inflight[key]
                progressed = True
                ms = 1e3 * (time.perf_counter() - tl)
                tnow = now()
                for r in runs:
                    r.step_ms.append(ms)
                    r.steps_done += 1
                    units += r.units()
                    if r.steps_done >= r.nsteps:
                        if on_finish:
                            torch.cuda.current_stream().wait_stream(self.pools[key].stream)
                            on_finish(r, self.pools[key])
                        if r.save_to:
                            self.pools[key].save(r)
                        self.retire(r, tnow)
                        active.remove(r)
            # admission (first fit over the queue; one run at a time when unbatched)
            if batched:
                for r in list(queue):
                    if self.try_admit(r, tnow):
                        queue.remove(r)
                        active.append(r)
                        progressed = True
                        if on_start:
                            on_start(r, self.pools[(r.nmax, r.km)])
            elif not active and queue:
                r = queue.pop(0)
                if not self.try_admit(r, tnow, max_runs=1):
                    raise DshtError(2, "run does not fit in an empty pool")
                active.append(r)
                progressed = True
                if on_start:
                    on_start(r, self.pools[(r.nmax, r.km)])
            max_active = max(max_active, len(active))
            # launches: every pool with runs and no step in flight (serialized: the next pool
            # in turn, once nothing is in flight)
            order = list(self.pools.items())
            if self.serialize and order:
                k = self._turn % len(order)
                order = order[k:] + order[:k]
            for key, pool in order:
                if self.serialize and inflight:
                    break
                if key in inflight or not pool.runs:
                    continue
                if self.serialize:
                    self._turn = (list(self.pools).index(key) + 1) % len(self.pools)
                runs = list(pool.runs)
                tl = time.perf_counter()
                pool.launch()
                ev = torch.cuda.Event()
                ev.record(pool.stream)
                inflight[key] = (ev, tl, runs)
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
