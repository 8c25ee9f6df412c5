"""Restart and history files (dsht/restart.py) on the host: the layout of dogcm's restart file
(DOGCM_IO_Restart_Create, model/dogcm/src/io/DOGCM_IO_Restart_mod.f90:485-590; variables of
DOGCM_Admin_Variable_mod.f90:378-399), exact round trips of every field, and the errors a
reader must raise.  The file is read back with scipy's NetCDF reader directly, independently
of the module's own reader."""
import numpy as np
import pytest

from dsht import restart

IM, JM, KM = 8, 4, 3


def _axes():
    lon = np.arange(IM) * 360.0 / IM
    lat = np.degrees(np.arcsin(np.linspace(-0.9, 0.9, JM)))
    sig = -(np.arange(KM) + 0.5) / KM
    return lon, lat, sig


def _fields(seed=0):
    rng = np.random.default_rng(seed)
    f = {}
    for name, dims, _, _ in restart.RESTART_VARS:
        shape = (KM, JM, IM) if dims == restart.LAYER else (JM, IM)
        f[name] = rng.standard_normal(shape)
        f[name + "B"] = rng.standard_normal(shape)
    f["Z"] = rng.standard_normal((KM, JM, IM))
    return f


def test_restart_layout_matches_dogcm(tmp_path):
    from scipy.io import netcdf_file
    path = str(tmp_path / "rst.nc")
    f = _fields()
    restart.write_restart(path, _axes(), f, 7200.0, attrs={"omega": 1.5e-4, "km": KM})
    nc = netcdf_file(path, "r", mmap=False)
    try:
        assert nc.dimensions["lon"] == IM and nc.dimensions["lat"] == JM and nc.dimensions["sig"] == KM
        assert nc.dimensions["timestr"] == 64 and nc.dimensions["time"] is None  # unlimited
        want = {"U", "UB", "V", "VB", "OMG", "OMGB", "H", "HB", "SSH", "SSHB", "PTemp", "PTempB", "Salt",
                "SaltB", "HydPres", "HydPresB", "SfcPres", "SfcPresB", "Z"}
        assert want <= set(nc.variables)
        assert nc.variables["U"].dimensions == ("time", "sig", "lat", "lon")
        assert nc.variables["SfcPresB"].dimensions == ("time", "lat", "lon")
        assert nc.variables["PTemp"].units == b"K" and nc.variables["Salt"].units == b"psu"
        assert nc.variables["lon"].units == b"degree_east" and nc.variables["lat"].units == b"degree_north"
        assert nc.variables["time"].units == b"sec"
        assert nc.variables["datetime"][0].tobytes().decode().strip() == "7200.0 sec"
        assert np.array_equal(nc.variables["SaltB"][0], f["SaltB"])
        assert float(nc.omega) == 1.5e-4
    finally:
        nc.close()


def test_restart_round_trip_is_exact(tmp_path):
    path = str(tmp_path / "rst.nc")
    f = _fields(3)
    axes = _axes()
    restart.write_restart(path, axes, f, 3600.0)
    ax, g, t, _ = restart.read_restart(path)
    assert t == 3600.0
    for a, b in zip(ax, axes):
        assert np.array_equal(a, b)
    for k in f:
        assert np.array_equal(g[k], f[k]), k
        assert g[k].dtype.isnative, k  # device copies need native byte order
    assert restart.read_restart(path, time=3600.0)[2] == 3600.0
    with pytest.raises(ValueError, match="no record at time"):
        restart.read_restart(path, time=1800.0)


def test_restart_rejects_missing_and_misshaped_fields(tmp_path):
    f = _fields()
    del f["PTempB"]
    with pytest.raises(KeyError):
        restart.write_restart(str(tmp_path / "a.nc"), _axes(), f, 0.0)
    f = _fields()
    f["U"] = f["U"][:, :, :-1]
    with pytest.raises(ValueError, match="shape"):
        restart.write_restart(str(tmp_path / "b.nc"), _axes(), f, 0.0)


def test_history_appends_records(tmp_path):
    path = str(tmp_path / "hst.nc")
    h = restart.History(path, _axes(), names=("U", "SfcPres"))
    recs = [_fields(s) for s in range(3)]
    for i, f in enumerate(recs):
        h.put(600.0 * (i + 1), f)
    t, u = restart.read_history(path, "U")
    _, sp = restart.read_history(path, "SfcPres")
    assert list(t) == [600.0, 1200.0, 1800.0]
    assert u.shape == (3, KM, JM, IM) and sp.shape == (3, JM, IM)
    for i, f in enumerate(recs):
        assert np.array_equal(u[i], f["U"]) and np.array_equal(sp[i], f["SfcPres"])
    with pytest.raises(ValueError, match="unknown history variables"):
        restart.History(str(tmp_path / "x.nc"), _axes(), names=("U", "Vorticity"))
