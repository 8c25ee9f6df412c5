"""Pins the CPU restatement of the LF-AM3 dynamics step (oracle/lfam3.py) before it is trusted
as the oracle of the device step.

Known answers and invariants, each chosen so that a plausible mistake fails one of them:
  - Rossby-Haurwitz wave: a small Y_4^3 streamfunction on a resting, unstratified ocean
    propagates westward at omega = -2 Omega m / (n (n+1)) (the linear barotropic Rossby wave
    of the reference's BarotRossby experiment, exp/BarotRossby, and
    Exp_BarotRossbyWave_mod.org); the phase error falls as dt^2, and a Coriolis term of the
    wrong sign (a mistake planted in the oracle) fails;
  - solid-body advection: without rotation, u = U0 cos(phi) is steady and carries a tracer
    pattern Y_n^m at angular speed U0/R (pins the horizontal flux form and its sign);
  - a resting, horizontally uniform stratification stays at rest;
  - the global tracer content sum_k cdk H <T_k> is conserved to round-off;
  - the vertical grid is the Chebyshev-Lobatto one of get_SpmlGridInfo.
"""
import numpy as np
import pytest

import lfam3
import oracle as O

R, OMEGA = 6.371e6, 7.292115e-5


def _uniform_state(o, km, u, v, t0=283.0, s0=35.0):
    T = np.full((km, o.J, o.I), t0)
    S = np.full((km, o.J, o.I), s0)
    P = np.zeros((o.J, o.I))
    return {"UB": u, "VB": v, "TB": T, "SB": S, "PB": P, "UN": u.copy(), "VN": v.copy(), "TN": T.copy(),
            "SN": S.copy(), "PN": P.copy()}


def _rossby_phase_error(dt, flip_coriolis=False, n=4, m=3, t_end=72000.0):
    o = O.Oracle(21, 64, 32, R)
    km = 3
    cfg = lfam3.Config(dt=dt, hvisc=0.0, numdiff=0.0, vvisc=0.0, vdiff=0.0)
    vor = np.zeros((km, o.L))
    vor[:, o.l_nm(n, m)] = 1e-8  # |u| ~ 3 cm/s: linear regime
    u, v = o.vordiv2uv(vor, np.zeros_like(vor))
    st = _uniform_state(o, km, u, v)
    step = lfam3.Step(o, km, cfg, OMEGA)
    if flip_coriolis:
        step.cori = -step.cori
    for _ in range(int(round(t_end / dt))):
        st = step.advance(st)
    wv, wd = o.uv2vordiv(st["UN"] * o.coslat[None, :, None], st["VN"] * o.coslat[None, :, None])
    c = wv[:, o.l_nm(n, m)] + 1j * wv[:, o.l_nm(n, -m)]
    omega = -2.0 * OMEGA * m / (n * (n + 1))
    want = -omega * t_end  # s(t) = s(0) exp(-i omega t)
    err = np.abs(np.angle(c * np.exp(-1j * want)))
    return float(err.max()), float(np.abs(np.abs(c) / 1e-8 - 1).max()), float(np.abs(wd).max())


def test_rossby_haurwitz_phase_speed():
    e1, a1, d1 = _rossby_phase_error(1800.0)
    e2, a2, _ = _rossby_phase_error(900.0)
    assert e1 < 1e-2 and a1 < 2e-3, (e1, a1)
    assert 3.5 < e1 / e2 < 4.5, (e1, e2)   # second order in dt: the error is discretisation
    assert d1 < 1e-15                       # the rigid lid keeps the flow non-divergent


def test_rossby_haurwitz_detects_planted_coriolis_sign():
    err, _, _ = _rossby_phase_error(1800.0, flip_coriolis=True)
    assert err > 1.0, err


@pytest.mark.parametrize("m", [1, 3])
def test_solid_body_tracer_advection(m):
    """Without rotation, u = U0 cos(phi) is a steady flow; a tracer anomaly eps Y_n^m in every
    layer turns eastward at U0/R: its coefficient goes as exp(-i m U0 t / R).  The tracer is
    passive here (no thermal expansion), so the flow stays solid-body."""
    o = O.Oracle(21, 64, 32, R)
    km, n, U0, dt, nst = 3, 5, 5.0, 1800.0, 20
    cfg = lfam3.Config(dt=dt, hvisc=0.0, numdiff=0.0, vvisc=0.0, vdiff=0.0, hdiff=0.0, beta_t=0.0)
    cos = o.coslat[None, :, None] * np.ones((km, 1, o.I))
    st = _uniform_state(o, km, U0 * cos, 0.0 * cos)
    wt = np.zeros((km, o.L))
    wt[:, o.l_nm(n, m)] = 1e-3
    for lev in ("TB", "TN"):
        st[lev] = st[lev] + o.inv(wt)
    step = lfam3.Step(o, km, cfg, 0.0)
    for _ in range(nst):
        st = step.advance(st)
    w = o.fwd(st["TN"])
    c = w[:, o.l_nm(n, m)] + 1j * w[:, o.l_nm(n, -m)]
    want = -m * U0 * nst * dt / R
    err = np.abs(np.angle(c * np.exp(-1j * want)))
    assert err.max() < 2e-3 * abs(want), (err, want)
    assert np.abs(np.abs(c) / 1e-3 - 1).max() < 1e-3
    # the flow itself stays solid-body
    assert np.abs(st["UN"] - U0 * cos).max() < 1e-12 * U0
    # reversing the flow reverses the drift (the sign of the flux divergence)
    st = _uniform_state(o, km, -U0 * cos, 0.0 * cos)
    for lev in ("TB", "TN"):
        st[lev] = st[lev] + o.inv(wt)
    for _ in range(nst):
        st = step.advance(st)
    w = o.fwd(st["TN"])
    c = w[:, o.l_nm(n, m)] + 1j * w[:, o.l_nm(n, -m)]
    assert np.abs(np.angle(c * np.exp(1j * want))).max() < 2e-3 * abs(want)


def test_resting_stratified_ocean_stays_at_rest():
    o = O.Oracle(21, 64, 32, R)
    km = 6
    cfg = lfam3.Config(dt=3600.0)
    z = cfg.depth * lfam3.vertical_grid(km)[4]
    zero = np.zeros((km, o.J, o.I))
    st = _uniform_state(o, km, zero, zero.copy())
    for lev in ("T", "S"):
        prof = (273.15 + 2 + 23 * np.exp(z / 1000.0)) if lev == "T" else (34.5 + np.exp(z / 1000.0))
        st[lev + "B"] = prof[:, None, None] + zero
        st[lev + "N"] = st[lev + "B"].copy()
    step = lfam3.Step(o, km, cfg, OMEGA)
    t0 = st["TN"].copy()
    for _ in range(3):
        st = step.advance(st)
    assert np.abs(st["UN"]).max() < 1e-12 and np.abs(st["VN"]).max() < 1e-12
    assert np.abs(st["TN"] - t0).max() < 1e-10


def test_tracer_content_conserved():
    o = O.Oracle(21, 64, 32, R)
    km = 5
    cfg = lfam3.Config(dt=3600.0)
    st = lfam3.initial_state(o, km, seed=3)
    step = lfam3.Step(o, km, cfg, 2 * OMEGA)
    cdk = step.cdk

    def content(tr):
        return float(np.sum(cdk * cfg.depth * o.avrlonlat(tr)))

    c0 = content(st["TN"]), content(st["SN"])
    for _ in range(4):
        st = step.advance(st)
    c1 = content(st["TN"]), content(st["SN"])
    assert abs(c1[0] / c0[0] - 1) < 2e-14 and abs(c1[1] / c0[1] - 1) < 2e-14, (c0, c1)
    assert np.all(np.isfinite(st["UN"])) and np.abs(st["UN"]).max() < 1.0


def test_vertical_grid_is_chebyshev_lobatto():
    """Faces are the g_Sig levels get_SpmlGridInfo returns (dsht_sig_grid, itself pinned by
    Clenshaw-Curtis exactness in test_abi.py); the layer thicknesses tile [-1, 0]; the
    centre spacing at the top and bottom half levels reaches the ghost centres half a layer
    outside, so it equals the first and last layer thickness."""
    import dsht
    for km in (2, 6, 20, 60):
        cdk, fdk, rcdk, rfdk, zck = lfam3.vertical_grid(km)
        fk, _ = dsht.sig_grid(km, km)
        assert np.abs(np.cumsum(np.concatenate([[0.0], -cdk])) - fk).max() < 4e-16
        assert abs(cdk.sum() - 1.0) < 4e-16 and np.all(cdk > 0)
        assert np.abs(zck - 0.5 * (fk[:-1] + fk[1:])).max() < 2e-16
        assert abs(fdk[0] - cdk[0]) < 4e-16 and abs(fdk[-1] - cdk[-1]) < 4e-16
        assert np.abs(fdk.sum() - (cdk.sum() + 0.5 * (cdk[0] + cdk[-1]))) < 1e-15
        assert np.allclose(rcdk * cdk, 1.0, rtol=1e-15) and np.allclose(rfdk * fdk, 1.0, rtol=1e-15)
