"""GPU parity: libdsht (HIP, through the C ABI) against the FP64 CPU oracle.

Every SpmlUtil operator on the horizontal path is checked on the same seeded inputs at
sizes the oracle finishes in seconds.  Tolerance for the plain FP64 form: relative L2
<= 1e-12 (the bar of the reference's FP64 ISPACK transforms).
"""
import numpy as np
import pytest

from conftest import red_spectrum, rel_l2

pytestmark = pytest.mark.gpu
TOL = 1e-12

SIZES = [(5, 16, 8), (21, 64, 32), (42, 128, 64), (31, 96, 48), (21, 72, 36)]


def _pair(N, I, J, R=6.371e6):
    import dsht
    import oracle
    return dsht.Plan(N, I, J, radius=R), oracle.Oracle(N, I, J, R)


@pytest.mark.parametrize("N,I,J", SIZES)
def test_grid_matches_oracle(gpu, N, I, J):
    p, o = _pair(N, I, J)
    assert np.abs(np.sin(p.y_Lat) - o.mu).max() < 1e-15
    assert rel_l2(p.y_Lat_Weight, o.y_lat_weight) < 1e-14
    assert abs(p.y_Lat_Weight.sum() - 2.0) < 1e-13


@pytest.mark.parametrize("N,I,J", SIZES)
@pytest.mark.parametrize("kind", ["VAL", "GRADLAMBDA", "GRADMU", "GRADLON", "GRADLAT"])
def test_inverse(gpu, N, I, J, kind):
    import dsht
    p, o = _pair(N, I, J)
    k = getattr(dsht, kind)
    w = red_spectrum(N, 7, seed=N + k)
    got = p._inv(w, k)
    ref = o.inv(w, k)
    assert rel_l2(got, ref) < TOL, (kind, rel_l2(got, ref))


@pytest.mark.parametrize("N,I,J", SIZES)
@pytest.mark.parametrize("kind", ["VAL", "DIVLAMBDA", "DIVMU"])
def test_forward(gpu, N, I, J, kind):
    import dsht
    p, o = _pair(N, I, J)
    k = getattr(dsht, kind)
    rng = np.random.default_rng(100 + N + k)
    xy = rng.standard_normal((5, J, I))
    if k == dsht.DIVMU:  # a field vanishing at the poles like V cos(phi)
        xy *= o.coslat[None, :, None] ** 2
    got = p._fwd(xy, k)
    ref = o.fwd(xy, k)
    assert rel_l2(got, ref) < TOL, (kind, rel_l2(got, ref))


@pytest.mark.parametrize("N,I,J", SIZES[:3])
def test_alpha_and_vordiv(gpu, N, I, J):
    p, o = _pair(N, I, J)
    rng = np.random.default_rng(7)
    A = rng.standard_normal((3, J, I)) * o.coslat[None, :, None]
    B = rng.standard_normal((3, J, I)) * o.coslat[None, :, None]
    assert rel_l2(p.w_AlphaOptr_xy(A, B), o.alpha_fwd(A, B)) < TOL
    wa, wb = red_spectrum(N, 3, 1), red_spectrum(N, 3, 2)
    assert rel_l2(p.xy_AlphaOptr_w(wa, wb), o.alpha_inv(wa, wb)) < TOL
    vor, div = p.calc_UVCosLat2VorDiv(A, B)
    rv, rd = o.uv2vordiv(A, B)
    assert rel_l2(vor, rv) < TOL and rel_l2(div, rd) < TOL
    u, v = p.calc_VorDiv2UV(wa, wb)
    ru, rv2 = o.vordiv2uv(wa, wb)
    assert rel_l2(u, ru) < TOL and rel_l2(v, rv2) < TOL


@pytest.mark.parametrize("N,I,J", SIZES[1:3])
def test_means_diag_interp(gpu, N, I, J):
    import dsht
    p, o = _pair(N, I, J)
    rng = np.random.default_rng(3)
    xy = rng.standard_normal((4, J, I)) + 2.0
    assert np.allclose(p.AvrLonLat_xy(xy), o.avrlonlat(xy), rtol=1e-13, atol=0)
    assert np.allclose(p.IntLonLat_xy(xy), o.intlonlat(xy), rtol=1e-13, atol=0)
    assert np.allclose(p.ya_IntLon_xya(xy), o.ya_intlon(xy), rtol=1e-13, atol=0)
    assert np.allclose(p.ya_AvrLon_xya(xy), o.ya_avrlon(xy), rtol=1e-13, atol=0)
    w = red_spectrum(N, 2, 9)
    assert rel_l2(p.w_Lapla_w(w), o.lapla(w)) < 1e-15
    assert rel_l2(p.w_LaplaInv_w(w), o.laplainv(w)) < 1e-15
    lon = rng.uniform(0, 2 * np.pi, 11)
    lat = rng.uniform(-1.5, 1.5, 11)
    got = p.Interpolate_w(w[0], lon, lat)
    ref = o.interpolate(w[0], lon, lat)
    assert rel_l2(got, ref) < 1e-12
    R = p.radius
    assert rel_l2(p.coef(dsht.COEF_HDIFF, [1e3, 1e-30, 8]), o.coef_hdiff(1e3, 1e-30, 8)) < 1e-14
    assert rel_l2(p.coef(dsht.COEF_HVISC, [1e3, 1e-30, 8]), o.coef_hvisc(1e3, 1e-30, 8)) < 1e-14
    assert rel_l2(p.coef(dsht.COEF_ADVFILTER), o.coef_advfilter()) < 1e-15
    assert rel_l2(p.coef(dsht.COEF_HELMHOLTZ, [1e-13, 5.0]), o.coef_helmholtz(1e-13, 5.0)) < 1e-14
    # implicit diffusion round trip: xy_w( w_xy(q)/(1 - dt c) )
    dt = 43200.0
    d = p.coef(dsht.COEF_BEDIFF, [dt, 1e3, -1.0, 8, 0, 86400.0])
    c = o.coef_hdiff(1e3, (N * (N + 1) / R ** 2) ** -4 / 86400.0, 8)
    assert rel_l2(d, 1.0 / (1.0 - dt * c)) < 1e-14
    q = o.inv(red_spectrum(N, 3, 11))
    got = p.xy_diag_xy(d, q)
    ref = o.inv(o.diag(d, o.fwd(q)))
    assert rel_l2(got, ref) < TOL


@pytest.mark.parametrize("layout", ["LC", "CL"])
def test_device_api_layouts(gpu, layout):
    import torch
    import dsht
    N, I, J, nc = 42, 128, 64, 40
    p, o = _pair(N, I, J)
    w = red_spectrum(N, nc, 5)
    xy = torch.empty((nc, J, I), dtype=torch.float64, device=gpu)
    wt = torch.tensor(w.T.copy() if layout == "LC" else w, device=gpu)
    p.inv_d(dsht.VAL, wt, xy, layout=layout)
    torch.cuda.synchronize()
    ref = o.inv(w)
    assert rel_l2(xy.cpu().numpy(), ref) < TOL
    w2 = torch.zeros_like(wt)
    p.fwd_d(dsht.VAL, xy, w2, layout=layout)
    torch.cuda.synchronize()
    got = w2.cpu().numpy()
    got = got.T if layout == "LC" else got
    assert rel_l2(got, w) < TOL


@pytest.mark.parametrize("N,I,J", [(85, 256, 128), (170, 512, 256)])
def test_every_kind_at_full_size(gpu, N, I, J):
    """Every inverse kind (VAL, GRADLAMBDA, GRADMU, GRADLON, GRADLAT and the fused value +
    GradLon + GradLat call) and every forward kind (VAL, DIVLAMBDA, DIVMU, the 𝓗 operator,
    UV <-> vor/div) against the oracle at the benchmark grids T85 (256x128) and T170
    (512x256), through the device entry points on 36 columns (two full 16-column tiles and a
    ragged one), spectra in the fast "LC" layout the bench uses."""
    import torch
    import dsht
    p, o = _pair(N, I, J)
    nc = 36
    w = red_spectrum(N, nc, seed=300 + N)
    wt = torch.tensor(w.T.copy(), device=gpu)
    for kind in ("VAL", "GRADLAMBDA", "GRADMU", "GRADLON", "GRADLAT"):
        k = getattr(dsht, kind)
        xy = torch.empty((nc, J, I), dtype=torch.float64, device=gpu)
        p.inv_d(k, wt, xy)
        torch.cuda.synchronize()
        err = rel_l2(xy.cpu().numpy(), o.inv(w, k))
        assert err < TOL, (N, kind, err)
    outs = [torch.empty((nc, J, I), dtype=torch.float64, device=gpu) for _ in range(3)]
    p.inv_vgrad_d(wt, *outs)
    torch.cuda.synchronize()
    for kind, t in zip(("VAL", "GRADLON", "GRADLAT"), outs):
        err = rel_l2(t.cpu().numpy(), o.inv(w, getattr(dsht, kind)))
        assert err < TOL, (N, "fused", kind, err)
    rng = np.random.default_rng(400 + N)
    xy = rng.standard_normal((nc, J, I)) * o.coslat[None, :, None]
    xt = torch.tensor(xy, device=gpu)
    for kind in ("VAL", "DIVLAMBDA", "DIVMU"):
        k = getattr(dsht, kind)
        out = torch.zeros((p.lm, nc), dtype=torch.float64, device=gpu)
        p.fwd_d(k, xt, out)
        torch.cuda.synchronize()
        err = rel_l2(out.cpu().numpy().T, o.fwd(xy, k))
        assert err < TOL, (N, kind, err)
    A, B = xy[:12], xy[12:24]
    assert rel_l2(p.w_AlphaOptr_xy(A, B), o.alpha_fwd(A, B)) < TOL
    vor, div = p.calc_UVCosLat2VorDiv(A, B)
    rv, rd = o.uv2vordiv(A, B)
    assert rel_l2(vor, rv) < TOL and rel_l2(div, rd) < TOL
    wa, wb = w[:12], w[12:24]
    assert rel_l2(p.xy_AlphaOptr_w(wa, wb), o.alpha_inv(wa, wb)) < TOL
    u, v = p.calc_VorDiv2UV(wa, wb)
    ru, rv2 = o.vordiv2uv(wa, wb)
    assert rel_l2(u, ru) < TOL and rel_l2(v, rv2) < TOL


@pytest.mark.parametrize("chunk,nc", [(32, 80), (0, 800)])
def test_roundtrip_t170_chunked(gpu, chunk, nc):
    """T170 (I=512, J=256) round trip over several scratch chunks: a forced 32-column chunk
    with a ragged last chunk, and the automatic policy splitting 800 columns (over its
    512 MiB Fourier cap at FP64) into equal chunks."""
    import torch
    import dsht
    N, I, J = 170, 512, 256
    p = dsht.Plan(N, I, J, radius=6.371e6)
    p.set_tuning(dsht.TUNE_CHUNK_COLS, chunk)
    w = torch.tensor(red_spectrum(N, nc, 17).T.copy(), device=gpu)
    xy = torch.empty((nc, J, I), dtype=torch.float64, device=gpu)
    w2 = torch.zeros_like(w)
    p.inv_d(dsht.VAL, w, xy)
    p.fwd_d(dsht.VAL, xy, w2)
    torch.cuda.synchronize()
    assert rel_l2(w2.cpu().numpy(), w.cpu().numpy()) < TOL


# every FFT length with a register two-pass kernel (M = I/2 = R1*R2), plus the Stockham
# kernels of the same lengths: both against each other and against the oracle
FFT_SIZES = [(5, 16, 8), (10, 32, 16), (21, 64, 32), (23, 72, 36), (31, 96, 48), (42, 128, 64),
             (47, 144, 72), (63, 192, 96), (85, 256, 128), (85, 270, 136), (127, 384, 192),
             (170, 512, 256), (170, 540, 270)]


@pytest.mark.parametrize("N,I,J", FFT_SIZES)
def test_fft_kernels(gpu, N, I, J):
    import dsht
    p, o = _pair(N, I, J)
    nc = 3
    w = red_spectrum(N, nc, seed=I)
    xy = np.random.default_rng(I).standard_normal((nc, J, I))
    ref_i, ref_f = o.inv(w), o.fwd(xy)
    out = {}
    for kind in (0, 1):
        p.set_tuning(dsht.TUNE_FFT, kind)
        out[kind] = (p._inv(w, dsht.VAL), p._fwd(xy, dsht.VAL))
        assert rel_l2(out[kind][0], ref_i) < TOL, (kind, rel_l2(out[kind][0], ref_i))
        assert rel_l2(out[kind][1], ref_f) < TOL, (kind, rel_l2(out[kind][1], ref_f))
    assert rel_l2(out[0][0], out[1][0]) < 1e-14
    assert rel_l2(out[0][1], out[1][1]) < 1e-14
    # register kernels in both row-group orders: the same rows, only their grouping differs
    p.set_tuning(dsht.TUNE_FFT, 0)
    for order in (0, 1):
        p.set_tuning(dsht.TUNE_FFT_ORDER, order)
        got = (p._inv(w, dsht.VAL), p._fwd(xy, dsht.VAL))
        assert np.array_equal(got[0], out[0][0]) and np.array_equal(got[1], out[0][1]), order


def test_errors(gpu):
    import dsht
    with pytest.raises(dsht.DshtError):
        dsht.Plan(42, 64, 64)  # im < 2*nmax+2
    with pytest.raises(dsht.DshtError):
        dsht.Plan(42, 128, 63)  # odd jm
    with pytest.raises(dsht.DshtError):
        dsht.Plan(42, 2 * 7 * 11, 64)  # im/2 has factor 7
    p = dsht.Plan(5, 16, 8)
    with pytest.raises(dsht.DshtError):
        p._inv(np.zeros(36), 99)


def test_fortran_dropin(gpu):
    """The Fortran SpmlUtil_mod shim (amdflang, ISO_C_BINDING) run end to end on the GPU."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(__file__)), "dennou-ogcm_amd", "spml_dropin_test")
    if not os.path.exists(exe):
        pytest.skip("Fortran test program not built (amdflang missing at build time)")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "PASS" in r.stdout
    assert "threads that made per-layer calls: 4" in r.stdout


def test_fortran_dropin_refuses_halo_sized_vectors(gpu):
    """A KA-sized (halo) z_CDK passed to a column wrapper stops the program with the shim's
    shape error instead of being read as shifted layers."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(os.path.dirname(__file__)), "dennou-ogcm_amd", "spml_dropin_test")
    if not os.path.exists(exe):
        pytest.skip("Fortran test program not built (amdflang missing at build time)")
    r = subprocess.run([exe, "badsize"], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode != 0
    assert "z_CDK has 5 entries, expected 3" in r.stdout
    assert "BADSIZE NOT REFUSED" not in r.stdout


@pytest.mark.parametrize("nthreads", [4, 8])
def test_host_api_concurrent_threads(gpu, nthreads):
    """Several host threads call dsht_xy_w / dsht_w_xy / dsht_xy_vgrad_w on their own layers at
    once (the reference's OpenMP loops over k); every result equals the serial call bit for
    bit."""
    import threading
    import dsht
    N, I, J, K = 42, 128, 64, 24
    p = dsht.Plan(N, I, J, kmax=K, radius=6.371e6)
    w = red_spectrum(N, K, 77)
    ser = [(p._inv(w[k], dsht.GRADLAT), p.xy_GradLonLat_w(w[k])) for k in range(K)]
    ser_f = [p._fwd(ser[k][0], dsht.DIVLAMBDA) for k in range(K)]
    out = [None] * K
    errors = []

    def worker(t):
        try:
            for rep in range(3):
                for k in range(t, K, nthreads):
                    g = p._inv(w[k], dsht.GRADLAT)
                    vg = p.xy_GradLonLat_w(w[k])
                    f = p._fwd(g, dsht.DIVLAMBDA)
                    out[k] = (g, vg, f)
        except Exception as e:  # noqa: BLE001
            errors.append(e)

    ts = [threading.Thread(target=worker, args=(t,)) for t in range(nthreads)]
    for t in ts:
        t.start()
    for t in ts:
        t.join()
    assert not errors, errors
    for k in range(K):
        g, vg, f = out[k]
        assert np.array_equal(g, ser[k][0]), k
        for a, b in zip(vg, ser[k][1]):
            assert np.array_equal(a, b), k
        assert np.array_equal(f, ser_f[k]), k


@pytest.mark.parametrize("leg", [0, 1, 2])
@pytest.mark.parametrize("N,I,J", SIZES[1:3])
def test_fp64_families(gpu, N, I, J, leg):
    """Every FP64 contraction family (DSHT_TUNE_LEG: 0 = LDS-staged inverse with the
    fragment-stream forward, 1 = LDS-staged, 2 = fragment streams) against the oracle."""
    import dsht
    p, o = _pair(N, I, J)
    p.set_tuning(dsht.TUNE_LEG, leg)
    rng = np.random.default_rng(200 + N + leg)
    xy = rng.standard_normal((19, J, I)) * o.coslat[None, :, None]
    for kind in ("VAL", "DIVLAMBDA", "DIVMU"):
        k = getattr(dsht, kind)
        assert rel_l2(p._fwd(xy, k), o.fwd(xy, k)) < TOL, (leg, kind)
    A, B = xy[:10], xy[9:]
    assert rel_l2(p.w_AlphaOptr_xy(A, B), o.alpha_fwd(A, B)) < TOL
    vor, div = p.calc_UVCosLat2VorDiv(A, B)
    rv, rd = o.uv2vordiv(A, B)
    assert rel_l2(vor, rv) < TOL and rel_l2(div, rd) < TOL
    w = red_spectrum(N, 19, seed=N + leg)
    for kind in ("VAL", "GRADLON", "GRADLAT"):
        k = getattr(dsht, kind)
        assert rel_l2(p._inv(w, k), o.inv(w, k)) < TOL, (leg, kind)
    for k, got in zip((dsht.VAL, dsht.GRADLON, dsht.GRADLAT), p.xy_GradLonLat_w(w)):
        assert rel_l2(got, o.inv(w, k)) < TOL, (leg, "fused", k)
