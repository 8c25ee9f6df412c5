"""Serving loop over device LF-AM3 pools (dsht/serve.py): continuous batching of independent
runs; each result is the same alone and in the mix (every run is computed as independent
columns: relative L2 <= 1e-13 written here), a served run matches the CPU restatement of
the LF-AM3 step (relative L2 <= 1e-11, the bound of test_gpu_lfam3.py), and admission
respects the memory budget."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _reqs(nsteps=4):
    from dsht.serve import Request
    # overlapping arrivals over every truncation of the workload, uneven layer counts
    spec = [(21, 4, 0.5, 0.000), (42, 6, 2.0, 0.000), (21, 4, 1.0, 0.001), (85, 4, 4.0, 0.001), (170, 2, 0.25, 0.002),
            (42, 6, 1.5, 0.002), (21, 4, 3.0, 0.003), (170, 2, 1.0, 0.003), (85, 4, 0.75, 0.004)]
    return [Request(i, n, km, om, nsteps, t, 1000 + i) for i, (n, km, om, t) in enumerate(spec)]


def _keep(store, tag):
    def f(r, pool):
        import dsht
        s = pool.lf
        store[(tag, r.rid)] = [s.run_view(fld, dsht.LF_N, r.slot).cpu().numpy().copy()
                               for fld in (dsht.LF_U, dsht.LF_V, dsht.LF_TRC, dsht.LF_SFCPRES)]
    return f


CAP = {21: 3, 42: 2, 85: 2, 170: 2}


@pytest.mark.parametrize("serialize", [True, False])
def test_results_same_alone_and_in_mix(gpu, serialize):
    from dsht.serve import Server
    final = {}
    srv = Server(capacity=CAP, serialize=serialize)
    reqs = _reqs()
    stats = srv.run(reqs, batched=True, on_finish=_keep(final, "mix"))
    assert stats["max_active"] >= 4  # the mix really overlapped
    assert all(len(r.step_ms) == 4 for r in reqs)
    for r in _reqs():
        Server(capacity=CAP).run([r], batched=False, on_finish=_keep(final, "alone"))
    for r in reqs:
        for a, b in zip(final[("mix", r.rid)], final[("alone", r.rid)]):
            assert np.all(np.isfinite(a))
            assert rel_l2(a, b) <= 1e-13, (r.rid, rel_l2(a, b))


def test_served_run_matches_oracle(gpu):
    """A T21 run served beside others, against the CPU restatement started from the state
    the server loaded for it."""
    import torch
    import dsht
    import lfam3
    import oracle
    from dsht import serve
    start, final = {}, {}

    def on_start(r, pool):
        torch.cuda.current_stream().wait_stream(pool.stream)
        s = pool.lf
        st = {}
        for lev, tag in ((dsht.LF_B, "B"), (dsht.LF_N, "N")):
            trc = s.run_view(dsht.LF_TRC, lev, r.slot).cpu().numpy()
            st["U" + tag] = s.run_view(dsht.LF_U, lev, r.slot).cpu().numpy().copy()
            st["V" + tag] = s.run_view(dsht.LF_V, lev, r.slot).cpu().numpy().copy()
            st["T" + tag], st["S" + tag] = trc[0].copy(), trc[1].copy()
            st["P" + tag] = s.run_view(dsht.LF_SFCPRES, lev, r.slot).cpu().numpy().copy()
        start[r.rid] = st

    reqs = _reqs(nsteps=3)
    serve.Server(capacity=CAP).run(reqs, batched=True, on_start=on_start, on_finish=_keep(final, "mix"))
    o = oracle.Oracle(21, *serve.GRIDS[21], serve.RADIUS)
    for r in reqs:
        if r.nmax != 21:
            continue
        step = lfam3.Step(o, r.km, lfam3.Config(dt=serve.DT[21]), r.omega)
        st = start[r.rid]
        for _ in range(r.nsteps):
            st = step.advance(st)
        u, v, trc, p = final[("mix", r.rid)]
        for got, ref in ((u, st["UN"]), (v, st["VN"]), (trc[0], st["TN"]), (trc[1], st["SN"]), (p, st["PN"])):
            assert rel_l2(got, ref) < 1e-11, (r.rid, rel_l2(got, ref))


def test_admission_by_memory_budget(gpu):
    from dsht.serve import Request, Server
    srv = Server(capacity=CAP)
    per_run = srv.pool(21, 4).bytes_per_run
    srv.budget = per_run * 2  # room for two runs
    reqs = [Request(i, 21, 4, 1.0, 3, 0.0, i) for i in range(3)]
    stats = srv.run(reqs, batched=True)
    assert stats["max_active"] == 2
    admitted = sorted(r.admitted for r in reqs)
    assert admitted[2] > admitted[1]  # the third run waited for a slot
    assert all(r.steps_done == 3 for r in reqs)


def test_save_and_resume_through_the_server(gpu, tmp_path):
    """A run served for 3 steps that leaves a restart file, resumed by another server for 2
    more steps, ends where the same run served 5 steps at once ends (relative L2 <= 1e-13);
    the resumed request takes its rotation rate and model time from the file."""
    from dsht import serve
    from dsht.serve import Request, Server
    final = {}
    r0 = Request(0, 21, 4, 1.5, 5, 0.0, 77)
    Server(capacity=CAP).run([r0], batched=True, on_finish=_keep(final, "full"))
    path = str(tmp_path / "run.nc")
    r1 = Request(1, 21, 4, 1.5, 3, 0.0, 77, save_to=path)
    Server(capacity=CAP).run([r1], batched=True)
    r2 = Request(2, 21, 4, 0.3, 2, 0.0, 5, resume_from=path)
    Server(capacity=CAP).run([r2], batched=True, on_finish=_keep(final, "resumed"))
    assert r2.model_time == 3 * serve.DT[21] and r2.omega == r1.omega
    for a, b in zip(final[("resumed", 2)], final[("full", 0)]):
        assert np.all(np.isfinite(a))
        assert rel_l2(a, b) <= 1e-13, rel_l2(a, b)
