"""Field snapshots and per-step analyzer profiles of a run, and the restart
from a snapshot: the content and the write schedule of the reference's
plot.h5 (WritePlots.jl:222-416), in a container this image can write without
an HDF5 library.

Container: a directory ``<prefix>plot.fdata/`` with
  meta.json          version "3.0", geometry, the unit set, the grid sizes,
                     the list of field datasets (name "000", "001", ... and
                     the attribute z in z units), the zdata row count and
                     column shapes;
  grid/<name>.npy    the grid axes (r, t, x, y as the geometry has them);
  field/<dset>.npy   Array(real.(field.E)) for RT and XYT, Array(field.E) for
                     T, R and XY (WritePlots.jl:310-331), in the reference's
                     array shape;
  zdata/<name>.bin   rows appended as raw little-endian float32 (FloatGPU):
                     RT  z, Fzr [Nr], rhozr [Nr], Fzt [Nt], iSzf [Nt/2+1]
                     XY  z, Izx [Nx] = I[:, Ny/2], Izy [Ny] = I[Nx/2, :]
                     XYT z, Fzxy [Nx, Ny], rhozxy [Nx, Ny], iSzf [Nt/2+1]
                     (WritePlots.jl:367-416, 422-459, 463-501).
Schedule (WritePlots.jl:263-283): a field dataset when z differs from the
last zdata z and z >= the next field distance (then + dz_field); a zdata row
when z >= the next zdata distance (then z + dz_zdata, dz_zdata = lam0/2/zu,
Filament.jl:94).

``read_field(path, dset)`` returns the stored field and its z, which is what
an initial-condition file of a restarted run reads
(examples/RT/gauss_restart.jl:1-22): ``initial_condition`` returns the stored
real field, the input's z is that z in metres, and the model then builds the
analytic signal and solves the plasma of that field as it does for any
initial field.
"""
import json
import os

import numpy as np

VERSION = "3.0"


class PlotField:
    def __init__(self, path, p, grid, z, dz_field, dz_zdata, ftype=np.float64):
        """z, dz_field (dz_plothdf) and dz_zdata in z units; the schedule's
        distances are kept in ``ftype``, the type of the run's z (Float32 for
        RT and XYT, Input.jl:107-110, :152-155)."""
        self.path = path
        self.geometry = p.get("geometry", "RT")
        self.zdata = self.geometry in ("RT", "XY", "XYT")
        os.makedirs(os.path.join(path, "field"), exist_ok=True)
        os.makedirs(os.path.join(path, "grid"), exist_ok=True)
        units = {k: float(p[k]) for k in ("ru", "xu", "yu", "zu", "tu", "Iu", "rhou") if k in p}
        axes = {}
        for name in ("r", "t", "x", "y"):
            if hasattr(grid, name):
                a = np.asarray(getattr(grid, name), dtype=np.float64)
                np.save(os.path.join(path, "grid", name + ".npy"), a)
                axes[name] = len(a)
        self.meta = dict(version=VERSION, geometry=self.geometry, units=units, grid=axes, field=[],
                         zdata=dict(rows=0, columns={}))
        self.ft = ftype
        self.ifield = 0
        self.dz_field = ftype(dz_field)
        self.zprev_field = ftype(-np.inf)
        self.znext_field = ftype(z)
        self.dz_zdata = ftype(dz_zdata)
        self.znext_zdata = ftype(z) if self.zdata else ftype(np.inf)
        if self.zdata:
            os.makedirs(os.path.join(path, "zdata"), exist_ok=True)
            for f in os.listdir(os.path.join(path, "zdata")):
                os.remove(os.path.join(path, "zdata", f))
        self._save_meta()

    def _save_meta(self):
        tmp = os.path.join(self.path, "meta.json.tmp")
        with open(tmp, "w") as f:
            json.dump(self.meta, f, indent=1)
        os.replace(tmp, os.path.join(self.path, "meta.json"))

    def write(self, model, d, z):
        """writeHDF (WritePlots.jl:263-283) after an analysis ``d`` (with
        arrays) of ``model`` at z (z units)."""
        z = self.ft(z)
        if z != self.zprev_field and z >= self.znext_field:
            E = model.field()
            cplx = self.geometry in ("T", "R", "XY")
            data = E if cplx else np.ascontiguousarray(E.real)
            if model.precision == 4:
                data = data.astype(np.complex64 if cplx else np.float32)
            dset = "%03d" % self.ifield
            np.save(os.path.join(self.path, "field", dset + ".npy"), data)
            self.meta["field"].append(dict(dset=dset, z=float(z)))
            self.ifield += 1
            self.znext_field = self.ft(self.znext_field + self.dz_field)
            self._save_meta()
        if self.zdata and z >= self.znext_zdata:
            if self.geometry == "RT":
                cols = dict(z=np.array([float(z)]), Fzr=d["Fr"], rhozr=d["rho_end"], Fzt=d["Ft"], iSzf=d["S"])
            elif self.geometry == "XY":
                cols = dict(z=np.array([float(z)]), Izx=d["Izx"], Izy=d["Izy"])
            else:
                cols = dict(z=np.array([float(z)]), Fzxy=d["Fxy"], rhozxy=d["rho_end"], iSzf=d["S"])
            for name, v in cols.items():
                a = np.asarray(v, dtype="<f4")
                with open(os.path.join(self.path, "zdata", name + ".bin"), "ab") as f:
                    f.write(a.tobytes(order="F"))
                self.meta["zdata"]["columns"][name] = list(a.shape)
            self.meta["zdata"]["rows"] += 1
            self.zprev_field = z
            self.znext_zdata = self.ft(z + self.dz_zdata)
            self._save_meta()


def meta(path):
    with open(os.path.join(path, "meta.json")) as f:
        return json.load(f)


def read_field(path, dset):
    """(field, z) of a stored dataset: the array as written and the attribute
    z in z units (gauss_restart.jl:9-16)."""
    m = meta(path)
    for e in m["field"]:
        if e["dset"] == dset:
            return np.load(os.path.join(path, "field", dset + ".npy"), allow_pickle=False), e["z"]
    raise KeyError(f"no field dataset {dset!r} in {path}")


def read_zdata(path, name):
    """zdata column ``name`` as [rows, ...] (float32)."""
    m = meta(path)
    shape = m["zdata"]["columns"][name]
    rows = m["zdata"]["rows"]
    a = np.fromfile(os.path.join(path, "zdata", name + ".bin"), dtype="<f4")
    per = int(np.prod(shape)) if shape else 1
    return np.stack([a[k * per:(k + 1) * per].reshape(shape, order="F") for k in range(rows)])


def restart(path, dset):
    """What an initial-condition file of a restarted run needs
    (gauss_restart.jl:1-22): the stored field and its z in metres."""
    E, z = read_field(path, dset)
    return E, z * meta(path)["units"]["zu"]
