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


def test_batch_at_workload_grids(gpu):
    """The batched form at its workload's grids and layer counts: every truncation of
    {T21, T42, T85, T170} on the bench grids (72x36, 144x72, 270x136, 540x270), KM from
    {20, 30, 60}.  The FP64 batched step against the oracle on three columns of every
    output group; the split form bf16x3 (the batched workload's form) against the FP64
    batched step within its bound (6e-5 relative L2), and bf16x6 within 1e-6."""
    import torch
    import dsht
    import oracle
    grids = {21: (72, 36), 42: (144, 72), 85: (270, 136), 170: (540, 270)}
    inst = [(21, 20), (42, 30), (85, 60), (170, 20), (21, 60), (85, 30), (170, 30), (42, 20)]
    plans = {n: dsht.Plan(n, *grids[n], radius=6.371e6) for n in grids}
    b = dsht.Batch()
    for n, km in inst:
        b.add(plans[n], km)
    b.finalize()
    views = {}
    for iid in range(len(inst)):
        g, _, gc, _ = b.info(iid)
        n = inst[iid][0]
        if g in views:
            continue
        p = plans[n]
        spec, grid, tend, spect = b.group_views(g, p, gc)
        rng = np.random.default_rng(40 + g)
        spec.copy_(torch.tensor(red_spectrum(n, spec.shape[1], 40 + g).T.copy(), device=gpu))
        tend.copy_(torch.tensor(rng.standard_normal(tuple(tend.shape)), device=gpu))
        views[g] = (n, gc, spec, grid, tend, spect)
    assert sorted(v[0] for v in views.values()) == [21, 42, 85, 170]
    b.step()
    torch.cuda.synchronize()
    ref = {g: (v[3].clone(), v[5].clone()) for g, v in views.items()}
    for g, (n, gc, spec, grid, tend, spect) in views.items():
        o = oracle.Oracle(n, *grids[n], 6.371e6)
        cols = [0, gc + 3, spec.shape[1] - 1]
        w = spec[:, cols].T.cpu().numpy().copy()
        for k, kind in enumerate([oracle.VAL, oracle.GRADLON, oracle.GRADLAT]):
            err = rel_l2(grid[k][cols].cpu().numpy(), o.inv(w, kind))
            assert err < 1e-12, (n, kind, err)
        tcols = [1, 2 * gc + 5, tend.shape[0] - 1]
        err = rel_l2(spect[:, tcols].T.cpu().numpy(), o.fwd(tend[tcols].cpu().numpy()))
        assert err < 1e-12, (n, "forward", err)
    for form, bound in (("bf16x3", 6e-5), ("bf16x6", 1e-6)):
        pid = {v: k for k, v in dsht.PREC_NAMES.items()}[form]
        for p in plans.values():
            p.set_precision(pid)
        b.step()
        torch.cuda.synchronize()
        for g, (n, gc, spec, grid, tend, spect) in views.items():
            eg = rel_l2(grid.cpu().numpy(), ref[g][0].cpu().numpy())
            es = rel_l2(spect.cpu().numpy(), ref[g][1].cpu().numpy())
            print(f"{form} T{n}: grid {eg:.2e} spectral {es:.2e}")
            assert eg < bound and es < bound, (form, n, eg, es)
    b.close()
