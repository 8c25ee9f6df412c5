"""Batched instances (dsht_batch_*): uneven truncation and layer count in one step.

Checks: the batched step equals the per-instance sequential step to 1e-14 relative (each
column is computed independently; the per-instance calls use narrower column blocks, a
different kernel instantiation, so agreement is to rounding), matches the FP64 oracle,
and the captured hipGraph replays give the eager result bit for bit on every replay (with
eager work between replays).
"""
import numpy as np
import pytest

from conftest import red_spectrum, rel_l2

pytestmark = pytest.mark.gpu


def _setup(gpu, sizes, kms, seed=0):
    import torch
    import dsht
    plans = {s: dsht.Plan(*s, radius=6.371e6) for s in sizes}
    b = dsht.Batch()
    ids = []
    for i, km in enumerate(kms):
        s = sizes[i % len(sizes)]
        ids.append((b.add(plans[s], km), s, km))
    b.finalize()
    groups = {}
    for iid, s, km in ids:
        g, c0, gc, _ = b.info(iid)
        groups[g] = (s, gc)
    views = {}
    for g, (s, gc) in groups.items():
        p = plans[s]
        spec, grid, tend, spect = b.group_views(g, p, gc)
        rng = np.random.default_rng(seed + g)
        spec.copy_(torch.tensor(red_spectrum(p.nmax, spec.shape[1], seed + g).T.copy(), device=gpu))
        tend.copy_(torch.tensor(rng.standard_normal(tuple(tend.shape)), device=gpu))
        views[g] = (s, gc, spec, grid, tend, spect)
    torch.cuda.synchronize()
    return plans, b, ids, views


def test_batch_matches_sequential_and_oracle(gpu):
    import torch
    import oracle
    sizes = [(21, 64, 32), (42, 128, 64)]
    plans, b, ids, views = _setup(gpu, sizes, [2, 3, 4, 1, 2])
    b.step()
    torch.cuda.synchronize()
    batched = {g: [t.clone() for t in v[3:]] for g, v in views.items()}
    for v in views.values():
        v[3].zero_()
        v[5].zero_()
    b.step_sequential()
    torch.cuda.synchronize()
    for g, v in views.items():
        d3 = rel_l2(v[3].cpu().numpy(), batched[g][0].cpu().numpy())
        d5 = rel_l2(v[5].cpu().numpy(), batched[g][2].cpu().numpy())
        print(f"group {g}: batched vs sequential rel diff grid {d3:.2e}, spectra {d5:.2e}")
        assert d3 <= 1e-14 and d5 <= 1e-14
    # oracle on every field column of group 0 (values and derivatives)
    for g, (s, gc, spec, grid, tend, spect) in views.items():
        o = oracle.Oracle(*s, 6.371e6)
        w = spec.T.cpu().numpy()
        for k, kind in enumerate([oracle.VAL, oracle.GRADLON, oracle.GRADLAT]):
            assert rel_l2(grid[k].cpu().numpy(), o.inv(w, kind)) < 1e-12
        assert rel_l2(spect.T.cpu().numpy(), o.fwd(tend.cpu().numpy())) < 1e-12


def test_batch_graph_replays(gpu):
    import torch
    sizes = [(21, 64, 32), (42, 128, 64), (31, 96, 48)]
    plans, b, ids, views = _setup(gpu, sizes, [3, 2, 5, 1], seed=5)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        b.step(stream=s.cuda_stream)
    s.synchronize()
    eager = {g: [t.clone() for t in (v[3], v[5])] for g, v in views.items()}
    b.capture(s.cuda_stream)
    for rep in range(3):
        for v in views.values():
            v[3].zero_()
            v[5].fill_(float(rep))
        torch.cuda.synchronize()
        b.launch(stream=s.cuda_stream)
        s.synchronize()
        for g, v in views.items():
            assert torch.equal(v[3], eager[g][0]), rep
            assert torch.equal(v[5], eager[g][1]), rep
        # eager work between replays
        x = torch.randn(1000, device=gpu)
        (x * 2).sum().item()
    b.close()
