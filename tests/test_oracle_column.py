"""Pins the oracle's vertical column diagnostics (oracle/dsht_oracle.c: orc_col_*), the
oracle of HBEDiagnose_OMG, _HydPres and _UVBarot
(model/dogcm/src/dynamics/hspm_vfvm/HBEDiagnose_hspm_vfvm_mod.f90:90-143, 221-255, 290-325)
and of the QUICK vertical advective flux (HBEBaroc_hspm_vfvm_mod.f90:233-317, pinned by the
polynomial closed forms further down).

Each check is one that a plausible mistake fails:
  - closed forms on uniform layers: omega falls linearly to 0 at the bottom half level,
    the pressure sits at the layer centres (k + 1/2), the mean of a linear profile is its
    middle value;
  - independent library routines on random columns: numpy's cumulative sums from the
    bottom (omega) and from the top (pressure), numpy's weighted average (barotropic mean);
  - invariants: omega at the top half level is minus the column integral of the
    divergence, the mean of a constant is that constant, the mean is odd in (u, v).
A scan in the wrong direction, an off-by-one half level, the pressure at the interfaces
instead of the centres, or a mean weighted without the thickness fails one of these.
"""
import numpy as np
import pytest

import oracle as O

GRAV = 9.8


def _cdk(km, seed):
    """Random positive layer thicknesses in sigma summing to 1 (the reference's z_CDK)."""
    c = np.random.default_rng(seed).uniform(0.2, 1.0, km)
    return c / c.sum()


@pytest.mark.parametrize("km", [1, 7, 8, 20, 60])
def test_omega_closed_form_and_cumsum(km):
    npts = 33
    cdk = np.full(km, 1.0 / km)
    D, H0 = 3e-6, 1.2
    omg = O.col_omega(np.full((km, npts), D), np.full((km, npts), H0), cdk)
    assert omg.shape == (km + 1, npts)
    assert np.all(omg[km] == 0.0)
    k = np.arange(km + 1)[:, None]
    np.testing.assert_allclose(omg, -D * H0 * (km - k) / km * np.ones((1, npts)), rtol=1e-14, atol=1e-22)
    # random columns against numpy's reversed cumulative sum
    rng = np.random.default_rng(km)
    div = rng.standard_normal((km, npts)) * 1e-6
    h = 1.0 + 0.1 * rng.standard_normal((km, npts))
    cdk = _cdk(km, km + 1)
    omg = O.col_omega(div, h, cdk)
    ref = -np.cumsum((div * (h * cdk[:, None]))[::-1], axis=0)[::-1]
    assert np.all(omg[km] == 0.0)
    np.testing.assert_allclose(omg[:km], ref, rtol=1e-12, atol=1e-20)
    # the top half level is minus the column integral of the divergence
    np.testing.assert_allclose(omg[0], -(div * h * cdk[:, None]).sum(0), rtol=1e-12, atol=1e-20)


@pytest.mark.parametrize("km", [1, 7, 8, 20, 60])
def test_hydpres_closed_form_and_cumsum(km):
    npts = 29
    cdk = np.full(km, 1.0 / km)
    rho, H0 = 1.5, 0.9
    p = O.col_hydpres(np.full((km, npts), rho), np.full((km, npts), H0), cdk, GRAV)
    k = np.arange(km)[:, None]
    np.testing.assert_allclose(p, GRAV * H0 * rho * (k + 0.5) / km * np.ones((1, npts)), rtol=1e-14)
    rng = np.random.default_rng(100 + km)
    dens = rng.standard_normal((km, npts))
    h = 1.0 + 0.1 * rng.standard_normal((km, npts))
    cdk = _cdk(km, km + 2)
    p = O.col_hydpres(dens, h, cdk, GRAV)
    d = GRAV * h * dens * cdk[:, None]
    ref = np.cumsum(d, axis=0) - 0.5 * d
    scale = np.abs(d).sum(0)
    assert np.all(np.abs(p - ref) <= 1e-13 * scale)


@pytest.mark.parametrize("km", [1, 7, 8, 20, 60])
def test_uvbarot_invariants(km):
    npts = 31
    rng = np.random.default_rng(200 + km)
    h = 1.0 + 0.1 * rng.standard_normal((km, npts))
    cdk = _cdk(km, km + 3)
    # a constant is its own mean (to rounding)
    ub, vb = O.col_uvbarot(np.full((km, npts), 0.25), np.full((km, npts), -2.0), h, cdk)
    np.testing.assert_allclose(ub, 0.25, rtol=4e-16 * km)
    np.testing.assert_allclose(vb, -2.0, rtol=4e-16 * km)
    # a linear profile on uniform layers averages to its middle value
    lin = np.arange(km, dtype=np.float64)[:, None] * np.ones((1, npts))
    ub, _ = O.col_uvbarot(lin, lin, np.ones((km, npts)), np.full(km, 1.0 / km))
    np.testing.assert_allclose(ub, (km - 1) / 2.0, rtol=1e-14, atol=1e-14)
    # random columns against numpy's weighted average; odd in (u, v)
    u = rng.standard_normal((km, npts))
    v = rng.standard_normal((km, npts))
    ub, vb = O.col_uvbarot(u, v, h, cdk)
    wts = h * cdk[:, None]
    np.testing.assert_allclose(ub, np.average(u, axis=0, weights=wts), rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(vb, np.average(v, axis=0, weights=wts), rtol=1e-12, atol=1e-15)
    ub2, vb2 = O.col_uvbarot(-u, -v, h, cdk)
    assert np.array_equal(ub2, -ub) and np.array_equal(vb2, -vb)


def test_planted_mistake_is_caught():
    """A top-down omega scan (a plausible mistake, planted here in a copy of the oracle's
    loop) fails the bottom-boundary and cumulative-sum checks above."""
    km, npts = 6, 5
    rng = np.random.default_rng(7)
    div = rng.standard_normal((km, npts))
    h = np.ones((km, npts))
    cdk = _cdk(km, 8)
    wrong = np.zeros((km + 1, npts))
    for k in range(km):  # integrates from the top instead of from the bottom
        wrong[k + 1] = wrong[k] - div[k] * (h[k] * cdk[k])
    good = O.col_omega(div, h, cdk)
    assert not np.all(wrong[km] == 0.0)
    assert not np.allclose(wrong, good)


# ---------------------------------------------------------------- QUICK vertical flux
# calc_ADVFlxZ_QUICK (HBEBaroc_hspm_vfvm_mod.f90:233-317).  On uniform layers (cdk = 1/km,
# reciprocal spacing rfdk = km) the reference's curvature term of layer l is
# 2 km^2 (T[l-1] - 2 T[l] + T[l+1]) and dfac = 1/(8 km^2), which fixes the flux of
# polynomial profiles in closed form:
#   constant or linear T: flux = omg (T[h-1] + T[h]) / 2 (the curvature term vanishes);
#   T = l^2: flux = omg ((T[h-1] + T[h])/2 - 1);
#   T = l^3: flux = omg ((T[h-1] + T[h])/2 - 1.5 (2h - 1)) - 1.5 |omg|  (the upwind term).
# Boundaries: flux[0] = 0 (omg[0] T[0] with a linear free surface), flux[km] = 0, and the
# centred interpolation at h = 1 and h = km - 1.

def _uniform(km):
    return np.full(km, 1.0 / km), np.full(km + 1, float(km))


@pytest.mark.parametrize("km", [5, 8, 20, 60])
@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_quick_polynomial_profiles(km, sign):
    npts = 7
    rng = np.random.default_rng(300 + km)
    omg = sign * rng.uniform(0.5, 2.0, (km + 1, npts))
    cdk, rfdk = _uniform(km)
    l = np.arange(km, dtype=np.float64)[:, None] * np.ones((1, npts))
    h = np.arange(1, km)[:, None]
    for T, corr in ((np.full((km, npts), 3.0), None), (2.0 - 0.5 * l, None), (l ** 2, "quad"), (l ** 3, "cubic")):
        f = O.col_quick(omg, T, cdk, rfdk)
        mid = 0.5 * (T[:-1] + T[1:])  # half levels 1..km-1
        ref = omg[1:km] * mid
        if corr == "quad":
            ref[1:-1] = omg[2:km - 1] * (mid[1:-1] - 1.0)
        elif corr == "cubic":
            hh = h[1:-1]
            ref[1:-1] = omg[2:km - 1] * (mid[1:-1] - 1.5 * (2 * hh - 1)) - 1.5 * np.abs(omg[2:km - 1])
        scale = np.abs(T).max() * np.abs(omg).max()
        assert np.abs(f[1:km] - ref).max() <= 1e-12 * scale, (km, corr)
        assert np.all(f[0] == 0.0) and np.all(f[km] == 0.0)
    # linear free surface: the top half level carries omg[0] T[0]
    T = rng.standard_normal((km, npts))
    f = O.col_quick(omg, T, cdk, rfdk, linfree=True)
    assert np.array_equal(f[0], omg[0] * T[0])


@pytest.mark.parametrize("km", [2, 3, 4])
def test_quick_short_columns(km):
    """Columns too short for the interior formula: only the centred interpolation."""
    npts = 4
    rng = np.random.default_rng(km)
    omg = rng.standard_normal((km + 1, npts))
    T = rng.standard_normal((km, npts))
    c = rng.uniform(0.2, 1.0, km)
    cdk = c / c.sum()
    rfdk = rng.uniform(1.0, 5.0, km + 1)
    f = O.col_quick(omg, T, cdk, rfdk)
    for hh in range(1, km):
        if 2 <= hh <= km - 2:
            continue
        m1 = cdk[hh] / (cdk[hh - 1] + cdk[hh])
        m2 = cdk[hh - 1] / (cdk[hh - 1] + cdk[hh])
        np.testing.assert_allclose(f[hh], omg[hh] * (m1 * T[hh - 1] + m2 * T[hh]), rtol=1e-15, atol=1e-15)
    assert np.all(f[0] == 0.0) and np.all(f[km] == 0.0)


# ---------------------------------------------------------------- implicit vertical diffusion
# calc_VDiffTRCEq (DOGCM_Phys_hspm_vfvm_mod.f90:752-861) solves its tridiagonal systems with
# LAPACK DGTSV, which is absent here; the oracle restates the published routine (LAPACK 3.x
# SRC/dgtsv.f).  The restatement is pinned against scipy's LAPACK dgtsv, bit for bit, on
# diagonally dominant systems and on systems that force row interchanges.  The assembly is
# pinned by closed forms: with kv = 0 the matrix is the identity (qa = q0 + dt rhs/e3), and
# multiplying row k by e3 cdk/dt and summing telescopes the flux terms, so
#   sum_k e3 cdk x = dt sum_k cdk rhs - dt alpha (2 r0 x0 [surface 'D'] + 2 rN xN [bottom 'D'])
# with r0 = kv0/(fdk0 e30) and rN = kv_{km-1}/(fdk_km e3_{km-1}); a wrong sign, index or
# boundary factor breaks it.

@pytest.mark.parametrize("n", [2, 3, 5, 60])
@pytest.mark.parametrize("kind", ["dominant", "general"])
def test_dgtsv_matches_lapack(n, kind):
    from scipy.linalg import lapack
    rng = np.random.default_rng(n + (0 if kind == "dominant" else 1000))
    for _ in range(5):
        dl, du = rng.standard_normal(n - 1), rng.standard_normal(n - 1)
        d = rng.standard_normal(n)
        if kind == "dominant":  # |d| > 7 >= |dl| + |du| + 1 after every elimination step
            dl, du = np.clip(dl, -3, 3), np.clip(du, -3, 3)
            d = np.sign(d) * (np.abs(d) + 7.0)
        b = rng.standard_normal(n)
        x, sw = O.dgtsv(dl, d, du, b)
        _, _, _, xl, info = lapack.dgtsv(dl.copy(), d.copy(), du.copy(), b.copy())
        assert info == 0
        assert np.array_equal(x, xl), (n, kind, np.abs(x - xl).max())
        if kind == "dominant":
            assert sw == 0


def _vdiff_inputs(km, npts, seed, pivot=False):
    rng = np.random.default_rng(seed)
    c = rng.uniform(0.2, 1.0, km)
    if pivot:  # thin layers under thick ones: DGTSV interchanges rows
        c[1::2] *= 1e-3
    cdk = c / c.sum()
    fdk = np.concatenate([[0.5 * cdk[0]], 0.5 * (cdk[:-1] + cdk[1:]), [0.5 * cdk[-1]]])
    kv = rng.uniform(1e-4, 1e-2, (km, npts)) * (30.0 if pivot else 1.0)
    e3 = 4000.0 * (1.0 + 0.05 * rng.standard_normal((km, npts)))
    q0 = 10.0 + rng.standard_normal((km, npts))
    rhs = rng.standard_normal((km, npts)) * 1e-5
    return q0, rhs, kv, e3, cdk, fdk


@pytest.mark.parametrize("km", [2, 3, 20, 60])
@pytest.mark.parametrize("bcs", [("N", "N"), ("D", "N"), ("N", "D"), ("D", "D")])
def test_vdiff_column_integral(km, bcs):
    npts = 6
    dt, alpha = 3600.0, 1.0
    for pivot in (False, True):
        q0, rhs, kv, e3, cdk, fdk = _vdiff_inputs(km, npts, 10 * km + len(bcs[0]), pivot)
        qa, sw = O.col_vdiff(q0, rhs, kv, e3, cdk, fdk, dt, alpha, *bcs)
        x = qa - q0
        lhs = (e3 * cdk[:, None] * x).sum(0)
        rhs_int = dt * (cdk[:, None] * rhs).sum(0)
        # tolerance relative to the magnitude of every term (the telescoped interior fluxes
        # cancel only to rounding of their own size)
        scale = (e3 * cdk[:, None] * np.abs(x)).sum(0) + dt * (cdk[:, None] * np.abs(rhs)).sum(0)
        for on, k, h in ((bcs[0] == "D", 0, 0), (bcs[1] == "D", -1, -1)):
            if on:
                t = dt * alpha * 2.0 * kv[k] / (fdk[h] * e3[k]) * x[k]
                rhs_int -= t
                scale += np.abs(t)
        assert np.all(np.abs(lhs - rhs_int) <= 1e-8 * scale), (km, bcs, pivot)
        if pivot and km >= 3:
            assert sw > 0  # the row-interchange path ran
    # no diffusivity: the identity system
    q0, rhs, kv, e3, cdk, fdk = _vdiff_inputs(km, npts, 5)
    qa, sw = O.col_vdiff(q0, rhs, np.zeros_like(kv), e3, cdk, fdk, dt, alpha, *bcs)
    assert np.array_equal(qa, q0 + dt * rhs / e3) and sw == 0


@pytest.mark.parametrize("km", [2, 20, 60])
@pytest.mark.parametrize("btm", ["N", "D"])
def test_vdiff_momentum_column_integral(km, btm):
    """calc_VDiffMOMEq: b = dt rhs, so sum_k e3 cdk x = dt sum_k e3 cdk rhs - dt alpha
    (2 r0 x0 [surface 'D'] + bfric x_bottom [bottom 'D']); bfric = 0 with a Dirichlet bottom
    is the no-flux bottom."""
    npts = 5
    dt, alpha = 1800.0, 1.0
    q0, rhs, kv, e3, cdk, fdk = _vdiff_inputs(km, npts, 40 + km)
    bf = np.random.default_rng(km).uniform(1e-4, 1e-2, npts)
    qa, _ = O.col_vdiff(q0, rhs, kv, e3, cdk, fdk, dt, alpha, "D", btm, btm_frict=bf)
    x = qa - q0
    lhs = (e3 * cdk[:, None] * x).sum(0)
    want = dt * (e3 * cdk[:, None] * rhs).sum(0)
    scale = (e3 * cdk[:, None] * np.abs(x)).sum(0) + dt * (e3 * cdk[:, None] * np.abs(rhs)).sum(0)
    t = dt * alpha * 2.0 * kv[0] / (fdk[0] * e3[0]) * x[0]
    want -= t
    scale += np.abs(t)
    if btm == "D":
        t = dt * alpha * bf * x[-1]
        want -= t
        scale += np.abs(t)
    assert np.all(np.abs(lhs - want) <= 1e-9 * scale), (km, btm)
    # the tracer form differs (rhs divided by e3)
    qt, _ = O.col_vdiff(q0, rhs, kv, e3, cdk, fdk, dt, alpha, "D", btm)
    assert not np.allclose(qt, qa)
