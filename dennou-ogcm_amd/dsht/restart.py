"""Restart and history files of LF-AM3 runs in dogcm's gtool5/NetCDF layout.

A restart file holds what DOGCM_Admin_Variable_RestartPut writes for the ocean
(model/dogcm/src/admin/DOGCM_Admin_Variable_mod.f90:264-300; variables registered at
:378-399): U, V, OMG, H, PTemp, Salt, HydPres (layers) and SSH, SfcPres (one level), each at
time level N under its own name and at level B under the name with "B" appended, plus Z.  The
file's dimensions and axes are those of DOGCM_IO_Restart_Create
(model/dogcm/src/io/DOGCM_IO_Restart_mod.f90:485-590): lon, lat, sig, timestr, time (unlimited),
with lon/lat the cell centres in degrees (x_CI, y_CJ; axis names and units from
model/dogcm/src/util/grid/DOGCM_GaussSpmVFvmGrid_mod.f90:155-187), sig the layer centres
z_CK, time in seconds, and a `datetime` string per record.  Variable dimensions are the
Fortran order (lon, lat, sig, time), i.e. (time, sig, lat, lon) in the file.  Files are
NetCDF classic, written and read with scipy.io.netcdf_file (gtool5 writes NetCDF through
netcdf-fortran; the classic format is what it reads back).

A history file (DOGCM_IO_History, variables registered at DOGCM_Admin_Variable_mod.f90:356-375)
holds the same axes and a time series of the N level of chosen variables, one record per call.

On the device side, `save_run` diagnoses what the LF-AM3 state does not carry from it with
the library itself: OMG from the divergence of (U, V) (calc_UVCosLat2VorDiv, xy_w and
HBEDiagnose_OMG: half levels KS..KE, the rigid-lid surface level KS-1 being zero) and HydPres
from the linear equation of state (EOS_Linear_mod.f90) and HBEDiagnose_HydPres; H and Z are the
flat bottom's vertical scale factor (xyz_H = D) and the layer-centre depths D z_CK
(DOGCM_Admin_Grid_mod.f90:376-388), SSH is zero under the rigid lid.  `load_run`
reads U, V, PTemp, Salt and SfcPres at both levels back into a pool slot; the step diagnoses
the rest again, as the reference's LF-AM3 step does.
"""
import numpy as np

# (name, dims, long name, units): DOGCM_Admin_Variable_mod.f90:378-399 (restart) and :356-375
LAYER, LEVEL = "IJKT", "IJT"
RESTART_VARS = [
    ("U", LAYER, "velocity (i axis)", "m/s"),
    ("V", LAYER, "velocity (j axis)", "m/s"),
    ("OMG", LAYER, "velocity (k axis)", "m/s"),
    ("SSH", LEVEL, "sea surface height", "m"),
    ("H", LAYER, "vertical scale factor corresponding to layer thickness", "m"),
    ("PTemp", LAYER, "potential temperature", "K"),
    ("Salt", LAYER, "salinity", "psu"),
    ("HydPres", LAYER, "hydrostatic pressure", "Pa"),
    ("SfcPres", LEVEL, "sea surface pressure", "Pa"),
]
HISTORY_VARS = RESTART_VARS + [("Z", LAYER, "depth from sea surface", "m")]
PROGNOSTIC = ("U", "V", "PTemp", "Salt", "SfcPres")


def _netcdf():
    from scipy.io import netcdf_file
    return netcdf_file


def axes_of(plan, km):
    """(lon, lat, sig): cell centres in degrees east / north and the layer centres z_CK of the
    Chebyshev-Lobatto sigma grid on [-1, 0] (DOGCM_GaussSpmVFvmGrid_mod.f90:392-410)."""
    from . import sig_grid
    fk, _ = sig_grid(km, km)
    return (np.degrees(plan.x_Lon), np.degrees(plan.y_Lat), 0.5 * (fk[:-1] + fk[1:]))


def _create(path, axes, title):
    nc = _netcdf()(path, "w", version=2)
    nc.title = title
    nc.createDimension("time", None)  # the record dimension (scipy's writer defines it first)
    lon, lat, sig = axes
    for name, vals, long_name, units in (("lon", lon, "longitude", "degree_east"),
                                         ("lat", lat, "latitude", "degree_north"),
                                         ("sig", sig, "general vertical coordinate", "1")):
        nc.createDimension(name, len(vals))
        v = nc.createVariable(name, "d", (name,))
        v[:] = vals
        v.long_name = long_name
        v.standard_name = long_name
        v.units = units
    nc.createDimension("timestr", 64)
    t = nc.createVariable("time", "d", ("time",))
    t.long_name = "time"
    t.units = "sec"
    d = nc.createVariable("datetime", "c", ("time", "timestr"))
    d.long_name = "time represented as strings"
    d.units = "1"
    return nc


def _add_vars(nc, table, names):
    for name, dims, long_name, units in table:
        for nm in names(name):
            shape = ("time", "sig", "lat", "lon") if dims == LAYER else ("time", "lat", "lon")
            v = nc.createVariable(nm, "d", shape)
            v.long_name = long_name
            v.units = units


def _put_time(nc, rec, time, datetime):
    nc.variables["time"][rec] = float(time)
    s = (datetime or f"{float(time):.1f} sec").encode()[:64].ljust(64)
    nc.variables["datetime"][rec] = np.frombuffer(s, dtype="S1")


def write_restart(path, axes, fields, time, datetime=None, attrs=None):
    """Writes one restart record.  `fields` maps every RESTART_VARS name and its "B" twin, and
    "Z", to arrays (km, J, I) or (J, I); `attrs` become global attributes."""
    km, jm, im = len(axes[2]), len(axes[1]), len(axes[0])
    nc = _create(path, axes, "restart data of dennou-ogcm_amd LF-AM3 runs")
    try:
        _add_vars(nc, RESTART_VARS, lambda n: (n, n + "B"))
        _add_vars(nc, [("Z", LAYER, "depth from sea surface", "m")], lambda n: (n,))
        _put_time(nc, 0, time, datetime)
        for name, dims, _, _ in RESTART_VARS + [("Z", LAYER, "", "")]:
            for nm in ((name, name + "B") if name != "Z" else (name,)):
                a = np.asarray(fields[nm], dtype=np.float64)
                want = (km, jm, im) if dims == LAYER else (jm, im)
                if a.shape != want:
                    raise ValueError(f"restart field {nm} has shape {a.shape}, expected {want}")
                nc.variables[nm][0] = a
        for k, v in (attrs or {}).items():
            # floats as NC_DOUBLE (a plain Python float would be written as NC_FLOAT)
            setattr(nc, k, np.float64(v) if isinstance(v, float) else v)
    finally:
        nc.close()


def read_restart(path, time=None):
    """(axes, fields, time, attrs) of the record at `time` (the last record if None), like
    DOGCM_IO_Restart_HistGet with range='time=RestartTime'."""
    nc = _netcdf()(path, "r", mmap=False)
    try:
        times = np.array(nc.variables["time"][:], dtype=np.float64)  # native byte order
        if len(times) == 0:
            raise ValueError(f"{path}: no record")
        rec = len(times) - 1 if time is None else int(np.argmin(np.abs(times - time)))
        if time is not None and abs(times[rec] - time) > 1e-6 * max(1.0, abs(time)):
            raise ValueError(f"{path}: no record at time {time} (records at {list(times)})")
        axes = tuple(np.array(nc.variables[a][:], dtype=np.float64) for a in ("lon", "lat", "sig"))
        fields = {nm: np.array(v[rec], dtype=np.float64) for nm, v in nc.variables.items()
                  if nm not in ("lon", "lat", "sig", "time", "datetime")}
        attrs = {k: v for k, v in nc._attributes.items()}
        return axes, fields, float(times[rec]), attrs
    finally:
        nc.close()


class History:
    """A history file: the axes and one record per `put` of the chosen variables (N level)."""

    def __init__(self, path, axes, names=("U", "V", "PTemp", "Salt", "SfcPres")):
        table = [t for t in HISTORY_VARS if t[0] in names]
        if len(table) != len(names):
            raise ValueError(f"unknown history variables: {sorted(set(names) - {t[0] for t in table})}")
        self.path, self.names, self.nrec = path, tuple(names), 0
        nc = _create(path, axes, "history data of dennou-ogcm_amd LF-AM3 runs")
        _add_vars(nc, table, lambda n: (n,))
        nc.close()

    def put(self, time, fields, datetime=None):
        nc = _netcdf()(self.path, "a")
        try:
            _put_time(nc, self.nrec, time, datetime)
            for nm in self.names:
                nc.variables[nm][self.nrec] = np.asarray(fields[nm], dtype=np.float64)
        finally:
            nc.close()
        self.nrec += 1


def read_history(path, name):
    """(times, values[time, ...]) of one history variable."""
    nc = _netcdf()(path, "r", mmap=False)
    try:
        return (np.array(nc.variables["time"][:], dtype=np.float64),
                np.array(nc.variables[name][:], dtype=np.float64))
    finally:
        nc.close()


# ---------------------------------------------------------------- device side (LF-AM3 pools)
def _vertical(km, depth):
    """(z_CDK, z_CK, D z_CK) of the Chebyshev-Lobatto sigma grid."""
    from . import sig_grid
    fk, _ = sig_grid(km, km)
    cdk = fk[:-1] - fk[1:]
    zck = 0.5 * (fk[:-1] + fk[1:])
    return cdk, zck, depth * zck


def run_fields(lf, slot, level, with_diagnostics=True):
    """Host arrays of one run's state at one level (DSHT_LF_B or DSHT_LF_N), keyed by the
    restart names; with_diagnostics adds OMG and HydPres diagnosed on the device, H, Z, SSH."""
    import torch
    from . import LF_SFCPRES, LF_TRC, LF_U, LF_V, VAL

    plan, cfg, km = lf.plan, lf.cfg, lf.km
    u = lf.run_view(LF_U, level, slot)
    v = lf.run_view(LF_V, level, slot)
    trc = lf.run_view(LF_TRC, level, slot)
    sp = lf.run_view(LF_SFCPRES, level, slot)
    out = {"U": u, "V": v, "PTemp": trc[0], "Salt": trc[1], "SfcPres": sp}
    if with_diagnostics:
        cdk, zck, zk = _vertical(km, cfg.depth)
        dev = u.device
        h = torch.full((km, plan.jm, plan.im), float(cfg.depth), dtype=torch.float64, device=dev)  # xyz_H = D
        cdk_d = torch.tensor(cdk, device=dev)
        # OMG: divergence of (U, V) (calc_UVCosLat2VorDiv of U cos, V cos, then xy_w), then
        # HBEDiagnose_OMG; the restart holds half levels KS..KE
        cs = torch.tensor(np.cos(plan.y_Lat), device=dev)[None, :, None]
        vor = torch.empty((plan.lm, km), dtype=torch.float64, device=dev)
        div = torch.empty_like(vor)
        plan.uv2vordiv_d((u * cs).contiguous(), (v * cs).contiguous(), vor, div)
        dgrid = torch.empty((km, plan.jm, plan.im), dtype=torch.float64, device=dev)
        plan.inv_d(VAL, div, dgrid)
        omg = plan.col_omega_d(dgrid, h, cdk_d)
        # HydPres: density deviation of the linear equation of state at p = -RefDens Grav z
        # (EOS_Linear_Eval), then HBEDiagnose_HydPres
        geo = cfg.grav * torch.tensor(zk, device=dev)[:, None, None]
        press_dep = 1.0 / (cfg.ref_dens * cfg.cs0 ** 2) - cfg.ref_temp * cfg.beta_t ** 2 / (cfg.ref_dens * cfg.cp0)
        dens = cfg.ref_dens * (-cfg.beta_t * (trc[0] - cfg.ref_temp) + press_dep * (-cfg.ref_dens * geo)
                               + cfg.beta_s * (trc[1] - cfg.ref_salt))
        hyd = plan.col_hydpres_d(dens.contiguous(), h, cdk_d, cfg.grav)
        torch.cuda.synchronize()
        out.update({"OMG": omg[1:], "HydPres": hyd, "H": h, "Z": torch.tensor(zk)[:, None, None]
                    .expand(km, plan.jm, plan.im), "SSH": torch.zeros((plan.jm, plan.im))})
    torch.cuda.synchronize()
    return {k: t.cpu().numpy().copy() for k, t in out.items()}


def save_run(lf, slot, path, time, omega=None, datetime=None):
    """Writes the restart file of run `slot` of pool `lf` (levels N and B) at model time
    `time` [s]; `omega` (the run's rotation rate) is kept as a global attribute."""
    from . import LF_B, LF_N
    n = run_fields(lf, slot, LF_N)
    b = run_fields(lf, slot, LF_B)
    fields = dict(n)
    for k, a in b.items():
        if k != "Z":
            fields[k + "B"] = a
    attrs = {"nmax": lf.plan.nmax, "km": lf.km, "dt": lf.cfg.dt}
    if omega is not None:
        attrs["omega"] = float(omega)
    write_restart(path, axes_of(lf.plan, lf.km), fields, time, datetime, attrs)


def load_run(lf, slot, path, time=None):
    """Reads a restart file into run `slot` of pool `lf` (levels N and B; the rotation rate if
    the file holds one).  The grid must be the pool's.  Returns the record's time."""
    import torch
    from . import LF_B, LF_N, LF_SFCPRES, LF_TRC, LF_U, LF_V

    axes, f, t, attrs = read_restart(path, time)
    lon, lat, sig = axes_of(lf.plan, lf.km)
    if len(axes[0]) != len(lon) or len(axes[1]) != len(lat) or len(axes[2]) != len(sig):
        raise ValueError(f"{path}: grid {len(axes[0])}x{len(axes[1])}x{len(axes[2])} does not match the pool's "
                         f"{len(lon)}x{len(lat)}x{len(sig)}")
    if np.abs(axes[1] - lat).max() > 1e-9 or np.abs(axes[2] - sig).max() > 1e-12:
        raise ValueError(f"{path}: latitudes or sigma levels differ from the pool's grid")
    for lev, sfx in ((LF_N, ""), (LF_B, "B")):
        lf.run_view(LF_U, lev, slot).copy_(torch.from_numpy(f["U" + sfx]))
        lf.run_view(LF_V, lev, slot).copy_(torch.from_numpy(f["V" + sfx]))
        trc = lf.run_view(LF_TRC, lev, slot)
        trc[0].copy_(torch.from_numpy(f["PTemp" + sfx]))
        trc[1].copy_(torch.from_numpy(f["Salt" + sfx]))
        lf.run_view(LF_SFCPRES, lev, slot).copy_(torch.from_numpy(f["SfcPres" + sfx]))
    if "omega" in attrs:
        lf.set_rotation(slot, float(attrs["omega"]))
    torch.cuda.synchronize()
    return t
