"""Writes the HDF5 catalog fixtures of tests/golden/hdf5 with h5py (the
library the reference's ``h5_reader`` uses, ``catalog_processor.py:16-88``),
and the values they hold as .npz, so the product's HDF5 reader is checked
against files the library itself wrote.

Run with an interpreter that has h5py (not needed by the tests):
    /opt/conda/bin/python3.9 tests/golden/make_hdf5_fixtures.py
Fixtures: the default file format (superblock 0, symbol-table groups,
contiguous datasets), chunked datasets with gzip / shuffle / fletcher32 and
big-endian and integer types, the newest format (superblock 3, link
messages, layout v4 with single-chunk and fixed-array chunk indexes, the
latter also over several pages), compact datasets, resizable datasets,
groups with more links than the compact limit (dense link storage), and a
nested group.
"""
import os

import h5py
import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "hdf5")


def columns(n, seed):
    rng = np.random.default_rng(seed)
    c = {a: rng.uniform(0.0, 1000.0, n) for a in ("x", "y", "z")}
    c.update({v: rng.normal(0.0, 300.0, n).astype(np.float32) for v in ("vx", "vy", "vz")})
    c["w"] = rng.uniform(0.5, 1.5, n)
    c["id"] = np.arange(n, dtype=np.int64) * 3 - 7
    return c


def main():
    os.makedirs(HERE, exist_ok=True)
    expected = {}
    # 1. defaults: superblock 0, old-style groups, contiguous datasets
    c = columns(1000, 1)
    with h5py.File(os.path.join(HERE, "box_default.h5"), "w") as f:
        for k, v in c.items():
            f.create_dataset(k, data=v)
    expected.update({f"box_default/{k}": v for k, v in c.items()})
    # 2. chunked + filters, big-endian and integer types
    c = columns(1500, 2)
    with h5py.File(os.path.join(HERE, "box_chunked.h5"), "w") as f:
        f.create_dataset("x", data=c["x"], chunks=(128,), compression="gzip", compression_opts=4, shuffle=True)
        f.create_dataset("y", data=c["y"], chunks=(200,), compression="gzip", fletcher32=True)
        f.create_dataset("z", data=c["z"], chunks=(1500,), shuffle=True)
        f.create_dataset("w", data=c["w"].astype(">f8"), chunks=(333,))
        f.create_dataset("vx", data=c["vx"].astype(">f4"), compression="gzip")
        f.create_dataset("id", data=c["id"], chunks=(100,), compression="gzip")
        f.create_dataset("i32", data=c["id"].astype(">i4"))
    expected.update({f"box_chunked/{k}": c[k] for k in ("x", "y", "z", "w", "vx", "id")})
    expected["box_chunked/i32"] = c["id"].astype(np.int32)
    # 3. newest format: superblock 3, link messages, layout v4
    c = columns(1000, 3)
    with h5py.File(os.path.join(HERE, "box_latest.h5"), "w", libver="latest") as f:
        f.create_dataset("x", data=c["x"])                                    # contiguous
        f.create_dataset("y", data=c["y"], chunks=(100,))                     # fixed array index
        f.create_dataset("z", data=c["z"], chunks=(1000,))                    # single chunk
        f.create_dataset("w", data=c["w"], chunks=(1000,), compression="gzip")  # single filtered chunk
        f.create_dataset("vx", data=c["vx"], chunks=(64,), compression="gzip", shuffle=True)  # filtered fixed array
        f.create_dataset("id", data=c["id"][:10])                             # small: compact below
        f.create_dataset("small", data=np.arange(5.0))
    expected.update({f"box_latest/{k}": c[k] for k in ("x", "y", "z", "w", "vx")})
    expected["box_latest/id"] = c["id"][:10]
    expected["box_latest/small"] = np.arange(5.0)
    # compact layout through the low-level API
    with h5py.File(os.path.join(HERE, "box_compact.h5"), "w") as f:
        dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
        dcpl.set_layout(h5py.h5d.COMPACT)
        data = np.linspace(-1.0, 1.0, 16)
        space = h5py.h5s.create_simple((16,))
        ds = h5py.h5d.create(f.id, b"x", h5py.h5t.IEEE_F64LE, space, dcpl=dcpl)
        ds.write(h5py.h5s.ALL, h5py.h5s.ALL, data)
    expected["box_compact/x"] = np.linspace(-1.0, 1.0, 16)
    # 4. many links: dense link storage (fractal heap + name index) in the
    # newest format, a big symbol table in the default one
    c = columns(50, 4)
    for tag, kw in (("many_latest", {"libver": "latest"}), ("many_default", {})):
        with h5py.File(os.path.join(HERE, f"box_{tag}.h5"), "w", **kw) as f:
            for i in range(40):
                f.create_dataset(f"col{i:02d}", data=c["x"] + i)
            for k, v in c.items():
                f.create_dataset(k, data=v)
        expected.update({f"box_{tag}/col{i:02d}": c["x"] + i for i in range(40)})
        expected.update({f"box_{tag}/{k}": v for k, v in c.items()})
    # 5. resizable (v1 B-tree index in the default format) and a fixed
    # array index of more than one page (1024 entries)
    c = columns(3000, 6)
    with h5py.File(os.path.join(HERE, "box_resizable.h5"), "w") as f:
        f.create_dataset("x", data=c["x"], maxshape=(None,), chunks=(256,))
        f.create_dataset("y", data=c["y"], maxshape=(None,), chunks=(7,), compression="gzip")
    with h5py.File(os.path.join(HERE, "box_paged.h5"), "w", libver="latest") as f:
        f.create_dataset("x", data=c["x"], chunks=(2,))
        f.create_dataset("y", data=c["y"], chunks=(2,), compression="gzip")
    # resizable in the newest format: an extensible-array chunk index, which
    # the reader refuses by name
    with h5py.File(os.path.join(HERE, "box_extensible.h5"), "w", libver="latest") as f:
        f.create_dataset("x", data=c["x"][:100], maxshape=(None,), chunks=(16,))
    # chunks allocated at creation, unfiltered: the implicit chunk index
    with h5py.File(os.path.join(HERE, "box_implicit.h5"), "w", libver="latest") as f:
        dcpl = h5py.h5p.create(h5py.h5p.DATASET_CREATE)
        dcpl.set_chunk((64,))
        dcpl.set_alloc_time(h5py.h5d.ALLOC_TIME_EARLY)
        space = h5py.h5s.create_simple((1000,))
        ds = h5py.h5d.create(f.id, b"x", h5py.h5t.IEEE_F64LE, space, dcpl=dcpl)
        ds.write(h5py.h5s.ALL, h5py.h5s.ALL, np.ascontiguousarray(c["x"][:1000]))
    expected["box_implicit/x"] = c["x"][:1000]
    expected.update({f"box_resizable/{k}": c[k] for k in ("x", "y")})
    expected.update({f"box_paged/{k}": c[k] for k in ("x", "y")})
    # 6. nested groups
    c = columns(20, 5)
    with h5py.File(os.path.join(HERE, "nested.h5"), "w") as f:
        g = f.create_group("cat").create_group("gal")
        for k, v in c.items():
            g.create_dataset(k, data=v)
        f.create_dataset("top", data=np.arange(3, dtype=np.int16))
    expected.update({f"nested/cat/gal/{k}": v for k, v in c.items()})
    expected["nested/top"] = np.arange(3, dtype=np.int16)
    np.savez(os.path.join(HERE, "expected.npz"), **expected)
    print(f"wrote {len(expected)} arrays in {HERE}")


if __name__ == "__main__":
    main()
