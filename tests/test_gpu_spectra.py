"""Energy and enstrophy spectra on the GPU (dsht_spec_energy, dsht_spec_energy_d; the cascaded
SPMODEL nma_/na_EnergyFromStreamfunc_wa and nma_/na_EnstrophyFromStreamfunc_wa,
SpmlUtil_mod.f90:131-132) against the oracle: the same products and the same summation
order with no fused multiply-adds, so bit for bit.  At T170 the spectrum's sum equals the
grid mean of |grad psi|^2/2 (Parseval on the library's own transforms).  Argument errors."""
import numpy as np
import pytest

from conftest import red_spectrum

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,I,J", [(21, 64, 32), (42, 128, 64), (170, 512, 256)])
@pytest.mark.parametrize("kind", [0, 1])
def test_spectra_match_oracle(gpu, N, I, J, kind):
    import torch
    import dsht
    import oracle
    p = dsht.Plan(N, I, J, radius=6.371e6)
    w = red_spectrum(N, 5, seed=N + kind)
    nma, na = p.spec_energy(w, kind)
    rnma, rna = oracle.spec_energy(w, N, enstrophy=bool(kind))
    assert np.array_equal(nma, rnma)
    assert np.array_equal(na, rna)
    # device form, coefficient-major spectra (sl = ncols, sc = 1), only the n spectrum
    wd = torch.tensor(np.ascontiguousarray(w.T), device=gpu)
    nad = torch.full((5, N + 1), np.nan, dtype=torch.float64, device=gpu)
    rc = dsht.load().dsht_spec_energy_d(p.h, kind, dsht._tptr(wd), 5, 1, 5, None, dsht._tptr(nad), None)
    assert rc == 0
    torch.cuda.synchronize()
    assert np.array_equal(nad.cpu().numpy(), rna)


def test_spectrum_parseval_t170(gpu):
    import dsht
    N, I, J = 170, 512, 256
    p = dsht.Plan(N, I, J, radius=6.371e6)
    w = red_spectrum(N, 2, seed=3)
    _, na = p.spec_energy(w, dsht.SPEC_ENERGY)
    xy, gl, ga = p.xy_GradLonLat_w(w)
    grid = np.array([float(np.sum(0.5 * (gl[c] ** 2 + ga[c] ** 2), axis=1) @ p.y_Lat_Weight) / (2.0 * I)
                     for c in range(2)])
    np.testing.assert_allclose(na.sum(axis=1), grid, rtol=1e-12)


def test_spectra_errors(gpu):
    import dsht
    p = dsht.Plan(21, 64, 32)
    w = red_spectrum(21, 1, seed=0)
    with pytest.raises(dsht.DshtError):
        p.spec_energy(w, 7)
    lib = dsht.load()
    assert lib.dsht_spec_energy(p.h, 0, dsht._ptr(np.ascontiguousarray(w)), 1, None, None) != 0
