"""The device LF-AM3 step (dsht_lfam3_*) against the CPU restatement of
DOGCM_TInt_LFAM3_advance_Dyn (oracle/lfam3.py, itself pinned in test_oracle_lfam3.py).

Several runs with different rotation rates and initial states share one pool; each run after
1-3 steps must match the oracle stepping that run alone, and a run's result must be the same
alone and in the mix (every transform and kernel treats runs as independent columns).

Tolerance: relative L2 over each whole field <= 1e-11 (measured 1e-15 to 2e-12), and for the
tracers relative to their perturbation (the field minus its layer means) <= 5e-11 (measured
4e-13 to 1.4e-11).  A step chains about 70 FP64 transform units, each agreeing with the
oracle to ~1e-15 relative to its whole input, and the step transforms fields whose layer
means exceed their evolving parts by 10^2-10^3 (280 K and 35 psu tracers against 1 K and
0.05 psu perturbations; the hydrostatic pressure in the Bernoulli term against its
horizontal variation), so the mean's round-off reaches the evolving part at that ratio.
The device step is also held to the Rossby-Haurwitz phase speed directly."""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-11
OMEGA = 7.292115e-5


def _pool(N, I, J, km, nruns, **cfg):
    import dsht
    p = dsht.Plan(N, I, J, radius=6.371e6)
    c = dsht.lfam3_config(km=km, max_runs=nruns, **cfg)
    return p, dsht.LfAm3(p, c)


def _load(pool, run, st):
    import torch
    import dsht
    for lev, tag in ((dsht.LF_B, "B"), (dsht.LF_N, "N")):
        pool.run_view(dsht.LF_U, lev, run).copy_(torch.tensor(st["U" + tag]))
        pool.run_view(dsht.LF_V, lev, run).copy_(torch.tensor(st["V" + tag]))
        trc = pool.run_view(dsht.LF_TRC, lev, run)
        trc[0].copy_(torch.tensor(st["T" + tag]))
        trc[1].copy_(torch.tensor(st["S" + tag]))
        pool.run_view(dsht.LF_SFCPRES, lev, run).copy_(torch.tensor(st["P" + tag]))


def _read(pool, run):
    import dsht
    trc = pool.run_view(dsht.LF_TRC, dsht.LF_N, run).cpu().numpy()
    return {"U": pool.run_view(dsht.LF_U, dsht.LF_N, run).cpu().numpy(),
            "V": pool.run_view(dsht.LF_V, dsht.LF_N, run).cpu().numpy(),
            "T": trc[0], "S": trc[1], "P": pool.run_view(dsht.LF_SFCPRES, dsht.LF_N, run).cpu().numpy(),
            "UB": pool.run_view(dsht.LF_U, dsht.LF_B, run).cpu().numpy()}


def _compare(got, st, tag):
    """(whole-field relative L2 of every field, tracer error relative to the perturbation)"""
    errs, pert = {}, {}
    for f in ("U", "V", "T", "S", "P"):
        errs[f] = rel_l2(got[f], st[f + "N"])
    for f in ("T", "S"):
        ref = st[f + "N"]
        pert[f] = float(np.linalg.norm(got[f] - ref) / np.linalg.norm(ref - ref.mean(axis=(1, 2), keepdims=True)))
    errs["UB"] = rel_l2(got["UB"], st["UB"])
    print(tag, {k: f"{v:.2e}" for k, v in errs.items()}, "perturbation", {k: f"{v:.2e}" for k, v in pert.items()},
          flush=True)
    return errs, pert


@pytest.mark.parametrize("N,I,J,km,nsteps", [(21, 64, 32, 4, 3), (42, 128, 64, 6, 2)])
def test_lfam3_matches_oracle(gpu, N, I, J, km, nsteps):
    import torch
    import lfam3
    import oracle
    rates = [0.5, 1.0, 2.5]
    p, pool = _pool(N, I, J, km, len(rates), dt=1800.0)
    o = oracle.Oracle(N, I, J, 6.371e6)
    cfg = lfam3.Config(dt=1800.0)
    states, steppers = [], []
    for r, fac in enumerate(rates):
        st = lfam3.initial_state(o, km, seed=10 * N + r)
        states.append(st)
        steppers.append(lfam3.Step(o, km, cfg, fac * OMEGA))
        _load(pool, r, st)
        pool.set_rotation(r, fac * OMEGA)
    torch.cuda.synchronize()
    for it in range(nsteps):
        pool.step(len(rates))
        torch.cuda.synchronize()
        for r in range(len(rates)):
            states[r] = steppers[r].advance(states[r])
            errs, pert = _compare(_read(pool, r), states[r], f"T{N} step {it + 1} run {r}")
            assert max(errs.values()) < TOL, (it, r, errs)
            assert max(pert.values()) < 5e-11, (it, r, pert)


def test_lfam3_alone_equals_mix(gpu):
    """Run 1 of a three-run pool against the same run as the only run of its own pool."""
    import torch
    import lfam3
    import oracle
    N, I, J, km = 21, 64, 32, 4
    o = oracle.Oracle(N, I, J, 6.371e6)
    sts = [lfam3.initial_state(o, km, seed=s) for s in (1, 2, 3)]
    _, mix = _pool(N, I, J, km, 3)
    for r in range(3):
        _load(mix, r, sts[r])
        mix.set_rotation(r, (r + 1) * OMEGA)
    _, alone = _pool(N, I, J, km, 1)
    _load(alone, 0, sts[1])
    alone.set_rotation(0, 2 * OMEGA)
    for _ in range(3):
        mix.step(3)
        alone.step(1)
    torch.cuda.synchronize()
    a, b = _read(mix, 1), _read(alone, 0)
    for f in a:
        assert rel_l2(a[f], b[f]) < 1e-13, (f, rel_l2(a[f], b[f]))


def test_lfam3_partial_pool_and_compaction(gpu):
    """Only the first nruns runs advance; copy_run moves a run (state and rotation) to another
    slot, after which it continues identically."""
    import torch
    import dsht
    import lfam3
    import oracle
    N, I, J, km = 21, 64, 32, 4
    o = oracle.Oracle(N, I, J, 6.371e6)
    _, pool = _pool(N, I, J, km, 3)
    sts = [lfam3.initial_state(o, km, seed=s) for s in (4, 5, 6)]
    for r in range(3):
        _load(pool, r, sts[r])
        pool.set_rotation(r, (0.5 + r) * OMEGA)
    u2 = pool.run_view(dsht.LF_U, dsht.LF_N, 2).clone()
    pool.step(2)
    torch.cuda.synchronize()
    assert torch.equal(pool.run_view(dsht.LF_U, dsht.LF_N, 2), u2)  # run 2 untouched
    pool.copy_run(0, 1)  # run 1 (rotation 1.5) now also in slot 0
    pool.step(2)
    torch.cuda.synchronize()
    a, b = _read(pool, 0), _read(pool, 1)
    for f in a:
        assert rel_l2(a[f], b[f]) < 1e-13, f
    with pytest.raises(dsht.DshtError):
        pool.step(4)
    f, i = pool.units(3)
    assert (f, i) == (3 * (42 * km + 9), 3 * (28 * km + 9))


def test_lfam3_rossby_haurwitz_on_device(gpu):
    """The device step's linear barotropic Rossby wave: a Y_4^3 streamfunction on a resting,
    unstratified ocean turns westward at omega = -2 Omega m/(n(n+1)); after 40 steps of
    1800 s the phase is within the scheme's O(dt^2) error (< 1e-2 rad) of the analytic one."""
    import torch
    import dsht
    import oracle
    N, I, J, km, n, m, dt = 21, 64, 32, 4, 4, 3, 1800.0
    p, pool = _pool(N, I, J, km, 1, dt=dt, hvisc=0.0, numdiff=0.0, vvisc=0.0, vdiff=0.0)
    o = oracle.Oracle(N, I, J, 6.371e6)
    vor = np.zeros((km, o.L))
    vor[:, o.l_nm(n, m)] = 1e-8
    u, v = o.vordiv2uv(vor, np.zeros_like(vor))
    T = np.full((km, J, I), 283.0)
    S = np.full((km, J, I), 35.0)
    st = {"UB": u, "VB": v, "TB": T, "SB": S, "PB": np.zeros((J, I))}
    st.update({k[0] + "N": val.copy() for k, val in st.items()})
    _load(pool, 0, st)
    pool.set_rotation(0, OMEGA)
    nst = 40
    for _ in range(nst):
        pool.step(1)
    torch.cuda.synchronize()
    un = pool.run_view(dsht.LF_U, dsht.LF_N, 0).cpu().numpy()
    vn = pool.run_view(dsht.LF_V, dsht.LF_N, 0).cpu().numpy()
    wv, wd = o.uv2vordiv(un * o.coslat[None, :, None], vn * o.coslat[None, :, None])
    c = wv[:, o.l_nm(n, m)] + 1j * wv[:, o.l_nm(n, -m)]
    want = 2.0 * OMEGA * m / (n * (n + 1)) * nst * dt
    err = np.abs(np.angle(c * np.exp(-1j * want)))
    print("phase error", err, "amplitude", np.abs(c) / 1e-8, flush=True)
    assert err.max() < 1e-2 and np.abs(np.abs(c) / 1e-8 - 1).max() < 2e-3
    assert np.abs(wd).max() < 1e-15
