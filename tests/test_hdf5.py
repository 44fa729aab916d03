"""HDF5 catalog ingestion without h5py (``cosmonpc_amd.hdf5``), checked
against files h5py itself wrote (tests/golden/hdf5, made by
tests/golden/make_hdf5_fixtures.py; the values are in expected.npz), and the
box branch of ``catalog_reader`` for .h5 catalogs
(``catalog_processor.py:16-88``, ``:388-449``)."""
import os

import numpy as np
import pytest

from cosmonpc_amd import hdf5

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "hdf5")


def _expected():
    return np.load(os.path.join(GOLD, "expected.npz"))


def test_every_fixture_dataset_bit_exact():
    """Superblocks 0 and 3, symbol-table / compact / dense groups, nested
    groups, contiguous / compact / chunked layouts (v1 B-tree, single chunk,
    fixed array over one and several pages), gzip / shuffle / fletcher32,
    big-endian floats and integers: every value equals what h5py wrote, and
    the dtype is the stored one in native byte order."""
    exp = _expected()
    assert len(exp.files) >= 130
    by_file = {}
    for key in exp.files:
        fn, ds = key.split("/", 1)
        by_file.setdefault(fn, []).append(ds)
    for fn, names in by_file.items():
        with hdf5.File(os.path.join(GOLD, fn + ".h5")) as f:
            for ds in names:
                got = f[ds].read()
                want = exp[f"{fn}/{ds}"]
                assert got.dtype == want.dtype.newbyteorder("="), (fn, ds, got.dtype, want.dtype)
                assert got.shape == want.shape, (fn, ds)
                np.testing.assert_array_equal(got, want, err_msg=f"{fn}/{ds}")


def test_group_listing_and_lookup():
    with hdf5.File(os.path.join(GOLD, "box_many_latest.h5")) as f:
        keys = f.keys()
        assert sorted(keys) == sorted([f"col{i:02d}" for i in range(40)] + ["x", "y", "z", "vx", "vy", "vz", "w",
                                                                             "id"])
        assert "x" in f and "missing" not in f
        with pytest.raises(KeyError):
            f["missing"]
    with hdf5.File(os.path.join(GOLD, "nested.h5")) as f:
        assert sorted(f.keys()) == ["cat", "top"]
        assert "cat/gal/x" in f and "cat/gal/nope" not in f
        assert len(f["cat/gal/w"]) == 20
        np.testing.assert_array_equal(f["cat/gal/x"][2:5], _expected()["nested/cat/gal/x"][2:5])


def test_unsupported_features_are_named():
    with hdf5.File(os.path.join(GOLD, "box_extensible.h5")) as f:
        with pytest.raises(NotImplementedError, match="chunk index type 4"):
            f["x"].read()


def test_not_an_hdf5_file(tmp_path):
    p = tmp_path / "junk.h5"
    p.write_bytes(b"\0" * 4096)
    with pytest.raises(ValueError, match="not an HDF5 file"):
        hdf5.File(str(p))


def test_read_columns_follows_h5_reader():
    """h5_reader: one float64 field per requested column; the row count is
    the length of the first requested column the file holds; columns the
    file lacks stay unset (NaN here); files are concatenated in order."""
    exp = _expected()
    a = os.path.join(GOLD, "box_default.h5")
    b = os.path.join(GOLD, "box_latest.h5")
    arr = hdf5.read_columns(a, ["nope", "x", "id", "w"])
    assert arr.dtype.names == ("nope", "x", "id", "w") and all(arr.dtype[n] == np.float64 for n in arr.dtype.names)
    assert len(arr) == 1000 and np.all(np.isnan(arr["nope"]))
    np.testing.assert_array_equal(arr["x"], exp["box_default/x"])
    np.testing.assert_array_equal(arr["id"], exp["box_default/id"].astype(np.float64))
    both = hdf5.read_columns([a, b], ["x", "vx"])
    assert len(both) == 2000
    np.testing.assert_array_equal(both["x"][1000:], exp["box_latest/x"])
    np.testing.assert_array_equal(both["vx"][:1000], exp["box_default/vx"].astype(np.float64))
    # the first requested column present sets the row count (10 here)
    short = hdf5.read_columns(b, ["id", "x"])
    assert len(short) == 10
    np.testing.assert_array_equal(short["x"], exp["box_latest/x"][:10])
    # no requested column present: no rows
    assert len(hdf5.read_columns(a, ["nope"])) == 0


def test_catalog_reader_box_hdf5_and_rsd():
    """Box .h5 catalogs go through the FITS/HDF5 branch
    (catalog_processor.py:388-449): x/y/z positions, weight 'w', RSD from
    vx/vy/vz."""
    from cosmonpc_amd import catalog

    exp = _expected()
    path = os.path.join(GOLD, "box_default.h5")
    pos = np.column_stack([exp[f"box_default/{a}"] for a in ("x", "y", "z")])
    cols = ["x", "y", "z", "vx", "vy", "vz", "w"]
    p, w = catalog.read_box_catalog(path, cols)
    np.testing.assert_array_equal(p, pos)
    np.testing.assert_array_equal(w, exp["box_default/w"])
    p, w = catalog.read_box_catalog(path, ["x", "y", "z"])
    np.testing.assert_array_equal(p, pos)
    assert np.all(w == 1.0)
    zb = 0.5
    p, _ = catalog.read_box_catalog(path, cols, apply_rsd=True, para_cosmo={"h": 0.7, "Omega0": 1.0},
                                    boxsize=[1000.0] * 3, redshift_box=zb, los=[0, 0, 1])
    fac = (1 + zb) / ((1 + zb) ** 1.5 * 100)
    vz = exp["box_default/vz"].astype(np.float64)
    np.testing.assert_allclose(p[:, 2], np.mod(pos[:, 2] + fac * vz, 1000.0), rtol=1e-13)
    np.testing.assert_array_equal(p[:, :2], pos[:, :2])
    # the .hdf5 extension too, and a list of files
    p2, _ = catalog.read_box_catalog([path, path], ["x", "y", "z"])
    assert len(p2) == 2000
