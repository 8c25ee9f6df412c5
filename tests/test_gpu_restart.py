"""Snapshot and resume of an LF-AM3 run through a restart file in dogcm's layout
(dsht/restart.py; DOGCM_IO_Restart_mod.f90, DOGCM_Admin_Variable_mod.f90:264-300).

Checks: a run saved after 3 steps of a two-run pool and resumed alone in a fresh pool reaches,
after 2 more steps, the state the uninterrupted run reached (relative L2 <= 1e-13 per field:
runs are independent columns, the two pools differ only in launch widths, the bound of
test_gpu_serve.py); the loaded state equals the saved one bit for bit, and so does the rotation
rate; the diagnosed OMG and HydPres in the file match the oracle's calc_UVCosLat2VorDiv, xy_w
and column loops on the same state (OMG <= 1e-11 relative L2, the transforms' parity; HydPres
<= 1e-13, the column kernel being bit-exact); H, Z and SSH are the flat-bottom constants; a
history file gets one record per put."""
import types

import numpy as np
import pytest

from conftest import rel_l2

pytestmark = pytest.mark.gpu
OMEGA = 7.292115e-5


def _fields(lf, slot):
    import dsht
    return [lf.run_view(f, dsht.LF_N, slot).cpu().numpy().copy()
            for f in (dsht.LF_U, dsht.LF_V, dsht.LF_TRC, dsht.LF_SFCPRES)]


def test_snapshot_and_resume(gpu, tmp_path):
    import torch
    import dsht
    import oracle
    from dsht import restart, serve
    N, I, J, km, dt = 21, 64, 32, 6, 3600.0
    p = dsht.Plan(N, I, J, radius=serve.RADIUS)
    lf = dsht.LfAm3(p, dsht.lfam3_config(km=km, max_runs=2, dt=dt))
    holder = types.SimpleNamespace(plan=p, km=km, lf=lf)
    for slot, (seed, om) in enumerate(((11, 1.0), (12, 2.0))):
        serve.initial_state(holder, slot, seed)
        lf.set_rotation(slot, om * OMEGA)
    s = torch.cuda.current_stream().cuda_stream
    for _ in range(3):
        lf.step(2, s)
    path = str(tmp_path / "run1.nc")
    restart.save_run(lf, 1, path, time=3 * dt, omega=2.0 * OMEGA)
    saved = _fields(lf, 1)
    for _ in range(2):
        lf.step(2, s)
    torch.cuda.synchronize()
    ref = _fields(lf, 1)

    # the file's diagnostics against the oracle on the saved state
    _, f, t, attrs = restart.read_restart(path)
    assert t == 3 * dt and float(attrs["omega"]) == 2.0 * OMEGA
    o = oracle.Oracle(N, I, J, serve.RADIUS)
    cs = np.cos(p.y_Lat)[None, :, None]
    _, wdiv = o.uv2vordiv(f["U"] * cs, f["V"] * cs)
    fk, _ = dsht.sig_grid(km, km)
    cdk = fk[:-1] - fk[1:]
    zck = 0.5 * (fk[:-1] + fk[1:])
    cfg = lf.cfg
    h = np.full((km, J, I), cfg.depth)
    omg = oracle.col_omega(o.inv(wdiv), h, cdk)
    assert rel_l2(f["OMG"], omg[1:]) <= 1e-11, rel_l2(f["OMG"], omg[1:])
    geo = cfg.grav * (cfg.depth * zck)[:, None, None]
    press_dep = 1.0 / (cfg.ref_dens * cfg.cs0 ** 2) - cfg.ref_temp * cfg.beta_t ** 2 / (cfg.ref_dens * cfg.cp0)
    dens = cfg.ref_dens * (-cfg.beta_t * (f["PTemp"] - cfg.ref_temp) + press_dep * (-cfg.ref_dens * geo)
                           + cfg.beta_s * (f["Salt"] - cfg.ref_salt))
    hyd = oracle.col_hydpres(dens, h, cdk, cfg.grav)
    assert rel_l2(f["HydPres"], hyd) <= 1e-13, rel_l2(f["HydPres"], hyd)
    assert np.all(f["H"] == cfg.depth) and np.all(f["SSH"] == 0.0)
    assert np.array_equal(f["Z"], np.broadcast_to((cfg.depth * zck)[:, None, None], (km, J, I)))
    for a, b in zip((f["U"], f["V"], f["PTemp"], f["Salt"], f["SfcPres"]),
                    (saved[0], saved[1], saved[2][0], saved[2][1], saved[3])):
        assert np.array_equal(a, b)

    # resume alone in a fresh pool
    lf2 = dsht.LfAm3(p, dsht.lfam3_config(km=km, max_runs=1, dt=dt))
    assert restart.load_run(lf2, 0, path) == 3 * dt
    got0 = _fields(lf2, 0)
    for a, b in zip(got0, saved):
        assert np.array_equal(a, b)
    for _ in range(2):
        lf2.step(1, s)
    torch.cuda.synchronize()
    for a, b in zip(_fields(lf2, 0), ref):
        assert np.all(np.isfinite(a))
        assert rel_l2(a, b) <= 1e-13, rel_l2(a, b)

    # a pool of another grid refuses the file
    p2 = dsht.Plan(N, 96, 48, radius=serve.RADIUS)
    lf3 = dsht.LfAm3(p2, dsht.lfam3_config(km=km, max_runs=1, dt=dt))
    with pytest.raises(ValueError, match="does not match"):
        restart.load_run(lf3, 0, path)

    # history: one record per put
    hp = str(tmp_path / "hst.nc")
    hst = restart.History(hp, restart.axes_of(p, km), names=("U", "PTemp", "SfcPres"))
    for k in range(2):
        lf2.step(1, s)
        hst.put((6 + k) * dt, restart.run_fields(lf2, 0, dsht.LF_N, with_diagnostics=False))
    times, u = restart.read_history(hp, "U")
    assert list(times) == [6 * dt, 7 * dt] and u.shape == (2, km, J, I)
    assert np.array_equal(u[1], _fields(lf2, 0)[0])
    for x in (lf, lf2, lf3):
        x.close()
