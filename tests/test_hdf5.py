"""The HDF5 reader against files the HDF5 library wrote (tests/golden/hdf5,
made by tests/golden/make_hdf5_savepoints.py with h5py 3.3.0 on HDF5 1.10.6).

Each file is checked value for value against expected.npz, its dimension
names against the dimension scales attached to it, and the storage forms the
reader met against the ones the file was written with, so each path of the
reader is covered by a file the library itself laid out."""

import os
import shutil

import numpy as np
import pytest

from pyshield_amd import hdf5
from pyshield_amd import translate as T

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hdf5")
EXPECTED = os.path.join(HERE, "expected.npz")

# storage forms each file was written with (see the generator's docstring)
FORMS = {
    "microph_earliest.nc": {"superblock 0", "object header 1", "group symbol table", "layout 3 class 1"},
    "microph_v108.nc": {"superblock 2", "object header 2", "group dense links", "v2 B-tree depth 0",
                        "layout 3 class 2", "chunk index v1 B-tree", "filter 1", "filter 2"},
    "microph_latest.nc": {"superblock 3", "object header 2", "group dense links", "layout 4 class 0",
                          "layout 4 class 2", "chunk index 1", "chunk index 3", "filter 3"},
    "many_links.nc": {"superblock 3", "group dense links", "fractal heap root indirect block",
                      "v2 B-tree depth 1"},
}


@pytest.mark.parametrize("name", ["microph_earliest.nc", "microph_v108.nc", "microph_latest.nc"])
def test_values_and_dimensions(name):
    seen = set()
    got = hdf5.read(os.path.join(HERE, name), seen)
    assert FORMS[name] <= seen, FORMS[name] - seen
    with np.load(EXPECTED, allow_pickle=False) as z:
        for k in z.files:
            if k.startswith("dims:"):
                assert ",".join(got[k[5:]].dims) == str(z[k]), k
                continue
            a = got[k].data
            assert a.dtype == z[k].dtype and a.shape == z[k].shape, k
            assert np.array_equal(a, z[k]), k
    # attributes: the fill value, a variable-length string, netCDF's dimension ids
    assert got["mph_pt"].attrs["_FillValue"] == -9.0e33
    if name != "microph_earliest.nc":
        assert got["mph_pt"].attrs["units"] == "K"
    assert [int(got[d].attrs["_Netcdf4Dimid"]) for d in ("savepoint", "rank", "cell", "one", "level")] == \
        [0, 1, 2, 3, 4]


def test_big_endian_and_compact_storage():
    seen = set()
    a = hdf5.read(os.path.join(HERE, "microph_earliest.nc"), seen)["mph_pt"].data
    b = hdf5.read(os.path.join(HERE, "microph_latest.nc"))["mph_land"].data
    with np.load(EXPECTED, allow_pickle=False) as z:
        assert np.array_equal(a, z["mph_pt"])  # stored big-endian
        assert np.array_equal(b, z["mph_land"])  # compact layout
    assert a.dtype == np.float64 and a.dtype.isnative


def test_dense_links_with_indirect_block_and_two_level_index():
    seen = set()
    got = hdf5.read(os.path.join(HERE, "many_links.nc"), seen)
    assert FORMS["many_links.nc"] <= seen
    assert sorted(got) == [f"scalar_{i:03d}" for i in range(150)]
    for i in range(150):
        v = got[f"scalar_{i:03d}"].data
        assert v.shape == () and float(v) == 0.5 * i


def test_load_savepoint_reads_netcdf4(tmp_path):
    with np.load(EXPECTED, allow_pickle=False) as z:
        for name in ("microph_earliest.nc", "microph_v108.nc", "microph_latest.nc"):
            d = T.load_savepoint(os.path.join(HERE, name), savepoint=1, rank=0)
            # the savepoint and rank dimensions are indexed; dimension-only
            # datasets are not variables, a dimension without savepoint/rank
            # leading dimensions is returned whole
            assert "cell" not in d and "level" not in d
            assert np.array_equal(d["mph_pt"], z["mph_pt"][1, 0])
            assert np.array_equal(d["mph_land"], z["mph_land"][1, 0])
            assert np.array_equal(d["nstep"], z["nstep"])
            assert d["mph_qv1"].shape == (16, 1, 8)
    # an HDF5 file under another extension, and the layout transform of the harness
    p = tmp_path / "Microph-In.h5"
    shutil.copy(os.path.join(HERE, "microph_v108.nc"), p)
    d = T.load_savepoint(str(p), savepoint=0)
    with np.load(EXPECTED, allow_pickle=False) as z:
        q = T.microph_to_ijk(d["mph_qv1"])
        assert q.shape == (4, 4, 8)
        assert np.array_equal(q, z["mph_qv1"][0, 0, :, 0, :].reshape(4, 4, 8))


def test_errors(tmp_path):
    p = tmp_path / "not.nc"
    p.write_bytes(b"CDF\x01" + bytes(60))
    assert not hdf5.is_hdf5(str(p))
    with pytest.raises(hdf5.HDF5Error):
        hdf5.read(str(p))
    # a truncated signature block
    raw = open(os.path.join(HERE, "microph_latest.nc"), "rb").read()
    q = tmp_path / "broken.nc"
    q.write_bytes(raw[:8] + bytes([9]) + raw[9:])  # superblock version 9
    with pytest.raises(hdf5.HDF5Error, match="superblock version 9"):
        hdf5.read(str(q))
