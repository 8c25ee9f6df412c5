"""Pins of the oracle's energy and enstrophy spectra (oracle.spec_energy; the cascaded SPMODEL
nma_/na_EnergyFromStreamfunc_wa and nma_/na_EnstrophyFromStreamfunc_wa, SpmlUtil_mod.f90:
131-132).  SPMODEL is absent from the reference, so the normalisation is fixed by Parseval's
identity on the Gaussian grid: the spectrum summed over (n, m) equals the global mean of
|grad psi|^2 / 2 (energy) or (lapla psi)^2 / 2 (enstrophy) of the streamfunction, evaluated
with the oracle's pinned inverse transforms on an alias-free grid; a single Y_1^0 mode has
energy a^2 (|grad (a sqrt(3) mu)|^2 / 2 averages to a^2).  A planted mistake (c_m = 1 for
every m) fails the identity."""
import numpy as np
import pytest

import oracle as O
from conftest import red_spectrum


def _parseval(o, w, nmax, enstrophy, spec=O.spec_energy):
    nma, na = spec(w, nmax, enstrophy)
    if enstrophy:
        n, _ = o.nm_l()
        zeta = o.inv(w * (-(n * (n + 1.0)))[None, :])
        grid = o.avrlonlat(0.5 * zeta * zeta)
    else:
        gl = o.inv(w, O.GRADLON)
        ga = o.inv(w, O.GRADLAT)
        grid = o.avrlonlat(0.5 * (gl * gl + ga * ga))
    return nma, na, grid


@pytest.mark.parametrize("enstrophy", [False, True])
@pytest.mark.parametrize("N,I,J", [(10, 32, 16), (21, 64, 32)])
def test_spectrum_sums_to_the_grid_mean(N, I, J, enstrophy):
    o = O.Oracle(N, I, J)
    w = red_spectrum(N, 3, seed=N + enstrophy)
    nma, na, grid = _parseval(o, w, N, enstrophy)
    np.testing.assert_allclose(na.sum(axis=1), grid, rtol=1e-13)
    np.testing.assert_allclose(nma.sum(axis=1), na, rtol=1e-14)
    for n in range(N + 1):  # zero where |m| > n
        assert np.all(nma[:, :N - n, n] == 0.0) and np.all(nma[:, N + n + 1:, n] == 0.0)


def test_single_mode_energy():
    N = 10
    w = np.zeros((1, (N + 1) ** 2))
    w[0, 2] = 0.7  # l_nm(1, 0) = 1*2 + 0
    nma, na = O.spec_energy(w, N)
    assert abs(na[0, 1] - 0.49) < 1e-16 and abs(na.sum() - 0.49) < 1e-16
    nma, na = O.spec_energy(w, N, enstrophy=True)
    assert abs(na[0, 1] - 0.49 * 2.0) < 1e-15  # n(n+1) = 2 once more


def test_planted_weight_fails():
    N, I, J = 10, 32, 16
    o = O.Oracle(N, I, J)
    w = red_spectrum(N, 2, seed=5)

    def planted(w, nmax, enstrophy=False):
        w = np.asarray(w).reshape(-1, (nmax + 1) ** 2)
        nma = np.zeros((w.shape[0], 2 * nmax + 1, nmax + 1))
        na = np.zeros((w.shape[0], nmax + 1))
        for n in range(nmax + 1):
            for m in range(-n, n + 1):
                v = (w[:, n * (n + 1) + m] ** 2) * (0.5 * n * (n + 1))
                nma[:, m + nmax, n] = v
                na[:, n] += v
        return nma, na

    _, na, grid = _parseval(o, w, N, False, spec=planted)
    assert np.max(np.abs(na.sum(axis=1) - grid) / grid) > 1e-2
