"""Pins the FP64 CPU oracle (oracle/dsht_oracle.c) before it is trusted as the parity oracle.

Each check is chosen so that a plausible mistake (dropped term, wrong sign, transposed
index, wrong normalisation or phase) fails at least one of them:
  - golden vectors from an independent 40-digit brute force (tests/golden/gen_golden.py)
  - library routines: numpy leggauss, numpy FFT, scipy lpmv
  - closed forms and invariants: orthonormality, round trip, Laplacian identity,
    solid-body rotation, mean identity, interpolation at nodes, implicit-diffusion decay
"""
import os

import numpy as np
import pytest
from scipy.special import lpmv, gammaln

import oracle as O
from conftest import red_spectrum, rel_l2

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("J", [8, 32, 64, 128, 256])
def test_gauss_nodes_weights(J):
    """Nodes against numpy's eigenvalue-based leggauss; weights against a 40-digit mpmath
    Newton (numpy's weights are only good to ~1e-11 relative at J = 256)."""
    import sys
    sys.path.insert(0, GOLD)
    from gen_golden import gauss

    o = O.Oracle(J // 2 - 1, 2 * J, J)
    x, w = np.polynomial.legendre.leggauss(J)
    assert np.abs(o.mu - x).max() < 4e-16
    assert np.abs(o.wq / (0.5 * w) - 1).max() < 1e-10
    xs, ws = gauss(J)
    assert np.abs(o.mu - np.array([float(v) for v in xs])).max() < 2e-16
    assert np.abs(o.wq / (0.5 * np.array([float(v) for v in ws])) - 1).max() < 1e-15
    assert abs(o.y_lat_weight.sum() - 2.0) < 1e-14


@pytest.mark.parametrize("I", [16, 64, 72, 96, 128, 270, 512])
def test_fft_vs_numpy(I):
    N = (I - 2) // 2 if I < 100 else (I - 1) // 3
    o = O.Oracle(N, I, 8 if N < 8 else 2 * ((N + 2) // 2))
    rng = np.random.default_rng(I)
    x = rng.standard_normal((6, I))
    G = o.fft_fwd(x)
    ref = np.fft.rfft(x, axis=1)[:, : N + 1] / I
    assert np.abs(G - ref).max() < 1e-15 * max(1, np.log2(I))
    Gb = ref.copy()
    Gb[:, 0] = Gb[:, 0].real
    full = np.zeros((6, I // 2 + 1), complex)
    full[:, : N + 1] = Gb
    assert np.abs(o.fft_inv(Gb) - np.fft.irfft(full * I, n=I, axis=1)).max() < 1e-13


def test_legendre_vs_scipy():
    N = 40
    o = O.Oracle(N, 128, 64)
    for m in [0, 1, 2, 7, 20, 40]:
        for n in range(m, N + 1, 3):
            p, _ = o.legendre(m, n)
            # scipy's lpmv carries the Condon-Shortley phase (-1)^m; ours does not
            norm = np.sqrt((2 * n + 1) * np.exp(gammaln(n - m + 1) - gammaln(n + m + 1)))
            ref = (-1) ** m * lpmv(m, n, o.mu) * norm
            assert np.abs(p - ref).max() < 1e-11 * max(1.0, np.abs(ref).max()), (m, n)
        assert np.all(o.legendre(m, m)[0] > 0)  # positive sectoral functions


def test_orthonormality():
    N = 42
    o = O.Oracle(N, 128, 64)
    for m in [0, 3, 21, 42]:
        P = np.array([o.legendre(m, n)[0] for n in range(m, N + 1)])
        gram = (P * o.wq) @ P.T
        assert np.abs(gram - np.eye(len(P))).max() < 1e-13


def test_H_is_derivative():
    """H_n^m = (1-mu^2) dP_n^m/dmu, checked by a centred difference of the recurrence."""
    N = 30
    o = O.Oracle(N, 64, 32)
    for m in [0, 1, 5, 30]:
        for n in range(m, N + 1, 4):
            _, H = o.legendre(m, n)
            # finite difference on a shifted oracle grid is awkward; use scipy's lpmv
            h = 1e-6
            norm = np.sqrt((2 * n + 1) * np.exp(gammaln(n - m + 1) - gammaln(n + m + 1)))
            d = ((-1) ** m * norm * (lpmv(m, n, o.mu + h) - lpmv(m, n, o.mu - h)) / (2 * h))
            ref = (1 - o.mu ** 2) * d
            assert np.abs(H - ref).max() < 1e-6 * max(1.0, np.abs(ref).max()), (m, n)


@pytest.mark.parametrize("T", [5, 10])
def test_golden_vectors(T):
    g = np.load(os.path.join(GOLD, f"sht_T{T}.npz"))
    N, I, J = int(g["N"]), int(g["I"]), int(g["J"])
    o = O.Oracle(N, I, J)
    assert np.abs(o.mu - g["mu"]).max() < 1e-15
    assert np.abs(o.wq - g["wq"]).max() < 1e-15
    w, xy = g["w_in"], g["xy_in"]
    tol = 1e-13
    assert rel_l2(o.inv(w)[0], g["inv_val"]) < tol
    assert rel_l2(o.inv(w, O.GRADLAMBDA)[0], g["inv_gradlambda"]) < tol
    assert rel_l2(o.inv(w, O.GRADMU)[0], g["inv_gradmu"]) < tol
    assert rel_l2(o.inv(w, O.GRADLON)[0], g["inv_gradlon"]) < tol
    assert rel_l2(o.inv(w, O.GRADLAT)[0], g["inv_gradlat"]) < tol
    assert rel_l2(o.fwd(xy)[0], g["fwd_val"]) < tol
    assert rel_l2(o.fwd(xy, O.DIVLAMBDA)[0], g["fwd_divlambda"]) < tol
    assert rel_l2(o.fwd(xy, O.DIVMU)[0], g["fwd_divmu"]) < tol


@pytest.mark.parametrize("N,I,J", [(21, 64, 32), (42, 128, 64), (21, 72, 36), (85, 256, 128)])
def test_roundtrip_and_projection(N, I, J):
    o = O.Oracle(N, I, J)
    w = red_spectrum(N, 3, N)
    assert rel_l2(o.fwd(o.inv(w)), w) < 1e-13
    rng = np.random.default_rng(1)
    xy = rng.standard_normal((2, J, I))
    q1 = o.inv(o.fwd(xy))
    assert rel_l2(o.inv(o.fwd(q1)), q1) < 1e-13  # inv o fwd is a projection


def _lapla_identity_error(o, w, gradmu_kind=O.GRADMU):
    a = o.fwd(o.inv(w, O.GRADLAMBDA), O.DIVLAMBDA)
    b = o.fwd(o.inv(w, gradmu_kind), O.DIVMU)
    return rel_l2(a + b, o.lapla(w))


@pytest.mark.parametrize("N,I,J", [(21, 64, 32), (42, 128, 64)])
def test_laplacian_identity(N, I, J):
    """DivLambda(GradLambda q) + DivMu(GradMu q) = -n(n+1) q on the unit sphere."""
    o = O.Oracle(N, I, J)
    w = red_spectrum(N, 2, 3)
    assert _lapla_identity_error(o, w) < 1e-12


def test_laplacian_identity_detects_planted_mistake():
    """The identity check is not vacuous: swapping GradMu for GradLambda fails it."""
    o = O.Oracle(21, 64, 32)
    w = red_spectrum(21, 1, 4)
    assert _lapla_identity_error(o, w, O.GRADLAMBDA) > 1e-2


def test_solid_body_rotation():
    """u = U0 cos(phi), v = 0: vor = 2 U0 sin(phi)/R, div = 0; s_1^0 = 2 U0/(sqrt(3) R)."""
    N, I, J, R, U0 = 21, 64, 32, 6.371e6, 10.0
    o = O.Oracle(N, I, J, R)
    cos = o.coslat[None, :, None] * np.ones((1, J, I))
    vor, div = o.uv2vordiv(U0 * cos * cos, 0 * cos)
    ref = np.zeros_like(vor)
    ref[0, o.l_nm(1, 0)] = 2 * U0 / (np.sqrt(3) * R)
    assert np.abs(vor - ref).max() < 1e-12 * abs(ref).max()
    assert np.abs(div).max() < 1e-12 * abs(ref).max()
    u, v = o.vordiv2uv(ref, 0 * ref)
    assert np.abs(u - U0 * cos).max() < 1e-11 * U0
    assert np.abs(v).max() < 1e-11 * U0


def test_vordiv_roundtrip():
    N, I, J, R = 21, 64, 32, 6.371e6
    o = O.Oracle(N, I, J, R)
    vor, div = red_spectrum(N, 2, 5), red_spectrum(N, 2, 6)
    vor[:, 0] = div[:, 0] = 0.0
    u, v = o.vordiv2uv(vor, div)
    cos = o.coslat[None, :, None]
    v2, d2 = o.uv2vordiv(u * cos, v * cos)
    assert rel_l2(v2, vor) < 1e-12 and rel_l2(d2, div) < 1e-12


def test_mean_and_interpolation():
    N, I, J = 21, 64, 32
    o = O.Oracle(N, I, J)
    w = red_spectrum(N, 3, 8)
    xy = o.inv(w)
    assert np.abs(o.avrlonlat(xy) - w[:, 0]).max() < 1e-14
    assert np.abs(o.intlonlat(xy) - 4 * np.pi * w[:, 0]).max() < 1e-13
    assert np.abs(o.ya_avrlon(xy) - xy.mean(axis=2)).max() < 1e-14
    assert np.abs(o.ya_intlon(xy) - 2 * np.pi * xy.mean(axis=2)).max() < 1e-13
    lon = np.repeat(o.lon[None, :], J, 0).ravel()
    lat = np.repeat(o.lat[:, None], I, 1).ravel()
    assert np.abs(o.interpolate(w[0], lon, lat) - xy[0].ravel()).max() < 1e-13


def test_implicit_diffusion_decay():
    """Each Y_n^m decays by exactly 1/(1 - dt c_n) (DOGCM_Phys_hspm_vfvm_mod.f90:640)."""
    N, I, J, R = 21, 64, 32, 6.371e6
    o = O.Oracle(N, I, J, R)
    dt, tau, ordh = 43200.0, 86400.0, 8
    numdiff = (N * (N + 1) / R ** 2) ** (-ordh // 2) / tau
    c = o.coef_hdiff(0.0, numdiff, ordh)
    for (n, m) in [(1, 0), (10, 3), (21, 21), (21, -5)]:
        w = np.zeros((1, o.L))
        w[0, o.l_nm(n, m)] = 1.0
        out = o.fwd(o.inv(w))
        out = o.diag(1.0 / (1.0 - dt * c), out)
        back = o.fwd(o.inv(out))
        assert abs(back[0, o.l_nm(n, m)] - 1.0 / (1.0 + dt * numdiff * (n * (n + 1) / R ** 2) ** 4)) < 1e-13
    # at n = N the default coefficient damps by 1/(1 + dt/tau)
    assert abs(c[o.l_nm(N, 0)] + 1.0 / tau) < 1e-18


def test_coefficient_tables():
    N, R = 42, 6.371e6
    o = O.Oracle(N, 128, 64, R)
    n = np.floor(np.sqrt(np.arange(o.L)))
    f = o.coef_advfilter()
    Nc = int(0.66 * N)
    assert np.all(f[n < Nc] == 1.0)
    assert abs(f[o.l_nm(N, 0)] - np.exp(-1.0)) < 1e-15
    h = o.coef_helmholtz(1e-12, 7.0)
    assert h[0] == 7.0
    assert abs(h[o.l_nm(3, 1)] - 1.0 / (12 / R ** 2 + 1e-12)) < 1e-18 * R ** 2
    v = o.coef_hvisc(1e3, 0.0, 8)
    assert abs(v[o.l_nm(1, 0)]) < 1e-20  # rigid rotation (n = 1) is not damped by viscosity


def _gradlonlat_modes(o):
    """Closed forms of single spherical harmonics on the oracle's grid, written in
    elementary functions of (lambda, phi) only.  The grid field of a coefficient pair is
    w(n,0) P_n^0 + 2 sum_m (w(n,m) cos m lambda - w(n,-m) sin m lambda) P_n^m, with P
    normalised to int P^2 dmu = 2 and no Condon-Shortley phase.  GradLon = (1/cos phi)
    d/dlambda and GradLat = d/dphi (the unit-sphere gradient; the reference's callers
    divide by the radius, HBEBaroc_hspm_vfvm_mod.f90:403,416)."""
    from math import factorial
    lam = o.lon[None, :]
    phi = o.lat[:, None]
    c, s = np.cos(phi), np.sin(phi)
    modes = []
    # Y_1^0 = sqrt(3) sin(phi)
    modes.append(((1, 0), 0 * lam * c, np.sqrt(3) * c * np.ones_like(lam)))
    # Y_1^1: sqrt(6) cos(phi) cos(lambda); the imaginary part: -sqrt(6) cos(phi) sin(lambda)
    r6 = np.sqrt(6.0)
    modes.append(((1, 1), -r6 * np.sin(lam) * np.ones_like(c), -r6 * s * np.cos(lam)))
    modes.append(((1, -1), -r6 * np.cos(lam) * np.ones_like(c), r6 * s * np.sin(lam)))
    # tesseral Y_2^1: sqrt(15/2) sin(2 phi) cos(lambda)
    k = np.sqrt(7.5)
    modes.append(((2, 1), -2 * k * s * np.sin(lam), 2 * k * np.cos(2 * phi) * np.cos(lam)))
    # sectoral Y_n^n: 2 c_n cos^n(phi) cos(n lambda), c_n^2 = (2n+1)!/(4^n (n!)^2)
    for n in (3, 5):
        cn = np.sqrt(factorial(2 * n + 1) / (4.0 ** n * factorial(n) ** 2))
        modes.append(((n, n), -2 * n * cn * c ** (n - 1) * np.sin(n * lam),
                      -2 * n * cn * c ** (n - 1) * s * np.cos(n * lam)))
    return modes


@pytest.mark.parametrize("N,I,J", [(5, 16, 8), (21, 64, 32), (42, 128, 64)])
def test_gradlonlat_closed_forms(N, I, J):
    """GRADLON and GRADLAT of Y_1^0, Y_1^{+-1}, a tesseral Y_2^1 and sectoral Y_3^3, Y_5^5
    against their closed forms."""
    o = O.Oracle(N, I, J)
    for (n, m), glon, glat in _gradlonlat_modes(o):
        w = np.zeros(o.L)
        w[o.l_nm(n, m)] = 1.0
        got_lon = o.inv(w, O.GRADLON)[0]
        got_lat = o.inv(w, O.GRADLAT)[0]
        scale = max(np.abs(glon).max(), np.abs(glat).max())
        assert np.abs(got_lon - glon).max() < 1e-13 * scale, (n, m)
        assert np.abs(got_lat - glat).max() < 1e-13 * scale, (n, m)


@pytest.mark.parametrize("mistake", ["cos", "1/cos^2", "none"])
def test_gradlonlat_checks_detect_planted_mistake(mistake):
    """The closed forms and the golden vectors are not vacuous for the cos(phi) factor:
    GradLon/GradLat formed with cos(phi) or 1/cos^2(phi) in place of 1/cos(phi) (a mistake
    planted in the oracle's output, on the CPU) fail both."""
    o = O.Oracle(21, 64, 32)
    c = o.coslat[:, None]
    fac = {"cos": c, "1/cos^2": 1.0 / c ** 2, "none": 1.0 / c}[mistake]
    worst = 0.0
    for (n, m), glon, glat in _gradlonlat_modes(o):
        w = np.zeros(o.L)
        w[o.l_nm(n, m)] = 1.0
        lon = o.inv(w, O.GRADLAMBDA)[0] * fac
        lat = o.inv(w, O.GRADMU)[0] * fac
        for got, ref in ((lon, glon), (lat, glat)):
            if np.abs(ref).max() > 0:
                worst = max(worst, rel_l2(got, ref))
    g = np.load(os.path.join(GOLD, "sht_T10.npz"))
    og = O.Oracle(int(g["N"]), int(g["I"]), int(g["J"]))
    cg = og.coslat[:, None]
    fg = {"cos": cg, "1/cos^2": 1.0 / cg ** 2, "none": 1.0 / cg}[mistake]
    gold = max(rel_l2(og.inv(g["w_in"], O.GRADLAMBDA)[0] * fg, g["inv_gradlon"]),
               rel_l2(og.inv(g["w_in"], O.GRADMU)[0] * fg, g["inv_gradlat"]))
    if mistake == "none":
        assert worst < 1e-13 and gold < 1e-13
    else:
        assert worst > 0.1 and gold > 0.1, (worst, gold)
