"""Serving loop: continuous batching of independent runs, each result the same alone
and in the mix (every column is computed independently, so the FP64 results agree bit
for bit; the bound written here is 1e-13 relative)."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu


def _reqs(nsteps=6):
    from dsht.serve import Request
    # overlapping arrivals, two truncations, uneven layer counts
    spec = [(21, 3, 0.5, 0.00), (42, 2, 2.0, 0.00), (21, 5, 1.0, 0.001), (42, 4, 4.0, 0.002), (21, 2, 0.25, 0.003)]
    return [Request(i, n, km, om, nsteps, t, 1000 + i) for i, (n, km, om, t) in enumerate(spec)]


@pytest.mark.parametrize("serialize", [True, False])
def test_results_same_alone_and_in_mix(gpu, serialize):
    from dsht.serve import Server
    final = {}

    def keep(tag):
        def f(r, pool):
            final[(tag, r.rid)] = pool.state[:, :, r.col0:r.col0 + r.km].cpu().numpy()
        return f

    cap = {21: 16, 42: 16, 85: 16, 170: 16}
    srv = Server(capacity=cap, dt=600.0, serialize=serialize)
    reqs = _reqs()
    stats = srv.run(reqs, batched=True, on_finish=keep("mix"))
    # the mix really overlapped: several runs were on the GPU in the same iteration
    assert stats["max_active"] >= 3
    assert all(len(r.step_ms) == 6 for r in reqs)
    for r in _reqs():
        srv_alone = Server(capacity=cap, dt=600.0)
        srv_alone.run([r], batched=False, on_finish=keep("alone"))
    for r in reqs:
        a, b = final[("mix", r.rid)], final[("alone", r.rid)]
        assert np.all(np.isfinite(a))
        assert rel_l2(a, b) <= 1e-13, (r.rid, rel_l2(a, b))


def test_admission_by_memory_budget(gpu):
    from dsht.serve import Request, Server
    cap = {21: 16, 42: 16, 85: 16, 170: 16}
    srv = Server(capacity=cap, dt=600.0)
    per_col = srv.pool(21).bytes_per_col
    srv.budget = per_col * 6  # room for 6 layer columns
    reqs = [Request(i, 21, 4, 1.0, 3, 0.0, i) for i in range(3)]
    srv.run(reqs, batched=True)
    # 4 + 4 > 6: the second and third run waited for slots
    admitted = sorted(r.admitted for r in reqs)
    assert admitted[1] > admitted[0]
    assert all(r.steps_done == 3 for r in reqs)
