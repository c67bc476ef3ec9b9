"""Write the HDF5 fixtures of the savepoint reader's tests with the HDF5 library.

Run with an interpreter that has h5py (here /opt/conda/bin/python3.9: h5py
3.3.0 on HDF5 1.10.6):

    /opt/conda/bin/python3.9 tests/golden/make_hdf5_savepoints.py

Each file holds a Microph-like savepoint laid out the way netCDF-4 lays out
a dataset in HDF5: variables [savepoint, rank, i*j, 1, k] attached to
dimension-scale datasets (NAME "This is a netCDF dimension but not a netCDF
variable.", _Netcdf4Dimid), a _FillValue per variable and _NCProperties on
the root group.  The variants cover the storage forms the reader supports:

* microph_earliest.nc: libver earliest (superblock 0, version 1 object
  headers, symbol-table group), contiguous data, one big-endian variable;
* microph_v108.nc: libver v108 with link creation order tracked on the root
  group (link info, dense link storage in a fractal heap with a name-index
  B-tree, as netCDF-4 creates groups), version 2 object headers, chunked data
  with deflate + shuffle (version 1 B-tree chunk index), a variable-length
  string attribute;
* microph_latest.nc: libver latest (superblock 3, layout message 4: fixed
  array and filtered single-chunk indexes), fletcher32, a compact dataset;
* many_links.nc: 150 scalar datasets in one group (a fractal heap with a
  root indirect block, a two-level name-index B-tree); scalar_i = i / 2.

expected.npz holds every variable's values (and "dims:<name>" entries with
the dimension names, comma-separated).
"""

import os

import h5py
import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hdf5")
NAMES = ["mph_qv1", "mph_ql1", "mph_qr1", "mph_qi1", "mph_qs1", "mph_qg1", "mph_qa1", "mph_uin", "mph_vin",
         "mph_delp", "mph_w", "mph_dz", "mph_pt", "mph_pt_dt", "mph_qv_dt"]
NSP, NRANK, NCELL, NK = 2, 1, 16, 8
DIM_NOTE = "This is a netCDF dimension but not a netCDF variable."


def values():
    rng = np.random.default_rng(20261204)
    out = {n: rng.standard_normal((NSP, NRANK, NCELL, 1, NK)) for n in NAMES}
    out["mph_land"] = (rng.random((NSP, NRANK, NCELL)) < 0.3).astype(np.float64)
    out["nstep"] = np.array([3, 4], dtype=np.int32)
    return out


def write(path, libver, track_order, chunked, filters, compact_name=None, big_endian=None):
    vals = values()
    with h5py.File(path, "w", libver=libver, track_order=track_order) as f:
        f.attrs["_NCProperties"] = np.bytes_(b"version=2,netcdf=4.9.2,hdf5=1.10.6")
        dims = {}
        for i, (dname, n) in enumerate((("savepoint", NSP), ("rank", NRANK), ("cell", NCELL), ("one", 1),
                                        ("level", NK))):
            d = f.create_dataset(dname, (n,), dtype="f4")
            d.make_scale(f"{DIM_NOTE}{n:10d}")
            d.attrs["_Netcdf4Dimid"] = np.int32(i)
            dims[dname] = d
        for name, v in vals.items():
            if name == "nstep":
                names = ("savepoint",)
            elif v.ndim == 3:
                names = ("savepoint", "rank", "cell")
            else:
                names = ("savepoint", "rank", "cell", "one", "level")
            kw = {}
            dtype = ">f8" if name == big_endian else v.dtype
            if name == compact_name:
                dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
                dcpl.set_layout(h5py.h5d.COMPACT)
                tid = h5py.h5t.py_create(np.dtype(dtype))
                sid = h5py.h5s.create_simple(v.shape)
                dsid = h5py.h5d.create(f.id, name.encode(), tid, sid, dcpl=dcpl)
                dsid.write(h5py.h5s.ALL, h5py.h5s.ALL, np.ascontiguousarray(v.astype(dtype)))
                ds = f[name]
            else:
                if chunked and v.ndim == 5:
                    kw["chunks"] = (1, 1, 5, 1, NK)  # ragged: 16 cells in chunks of 5
                    kw.update(filters)
                elif chunked:
                    kw["chunks"] = v.shape  # one chunk (single-chunk index in layout 4)
                    kw.update(filters)
                ds = f.create_dataset(name, data=v.astype(dtype), **kw)
            fill = v.dtype.type(-9.0e33) if v.dtype.kind == "f" else v.dtype.type(-2147483647)
            ds.attrs["_FillValue"] = fill
            for i, dn in enumerate(names):
                ds.dims[i].attach_scale(dims[dn])
        if filters:
            f["mph_pt"].attrs.create("units", "K", dtype=h5py.string_dtype())
    return vals


def main():
    os.makedirs(HERE, exist_ok=True)
    vals = write(os.path.join(HERE, "microph_earliest.nc"), "earliest", False, False, {}, big_endian="mph_pt")
    write(os.path.join(HERE, "microph_v108.nc"), "v108", True, True, {"compression": "gzip", "shuffle": True})
    write(os.path.join(HERE, "microph_latest.nc"), "latest", True, True, {"fletcher32": True},
          compact_name="mph_land")
    # a group of 150 links: the fractal heap of the dense link storage grows a
    # root indirect block and the name-index B-tree a second level
    with h5py.File(os.path.join(HERE, "many_links.nc"), "w", libver="latest", track_order=True) as f:
        for i in range(150):
            f.create_dataset(f"scalar_{i:03d}", data=np.float64(i) * 0.5)
    exp = dict(vals)
    for name, v in vals.items():
        if name == "nstep":
            exp["dims:" + name] = np.array("savepoint")
        elif v.ndim == 3:
            exp["dims:" + name] = np.array("savepoint,rank,cell")
        else:
            exp["dims:" + name] = np.array("savepoint,rank,cell,one,level")
    np.savez(os.path.join(HERE, "expected.npz"), **exp)
    for fn in sorted(os.listdir(HERE)):
        print(fn, os.path.getsize(os.path.join(HERE, fn)))


if __name__ == "__main__":
    main()
