"""Model states on disk (somax_amd/io.py): the reference's IO tests
(tests/test_io_xarray.py) on this package's states, plus the zarr v3 layout
checked field by field against the format specification (the reference
writes zarr v3 through xarray; neither is installed here, so reading this
package's stores with xarray is not exercised: that interop is parity
unpinned beyond the specification)."""

from __future__ import annotations

import gzip
import json

import numpy as np
import pytest
import torch

from somax_amd import io
from somax_amd.qg import BaroclinicQGState, BarotropicQGState
from somax_amd.swm import MultilayerSW2DState, NonlinearSW2DState


def _swm(ny=8, nx=12):
    return NonlinearSW2DState(h=torch.ones((ny, nx)) * 100.0, u=torch.zeros((ny, nx)), v=torch.zeros((ny, nx)))


def test_single_state_round_trip_in_memory():
    """test_io_xarray.py:45-90."""
    st = _swm()
    ds = io.state_to_dataset(st)
    assert set(ds.data_vars) == {"h", "u", "v"} and ds["h"].dims == ("y", "x") and ds["h"].shape == (8, 12)
    assert ds.attrs["state_class"] == "NonlinearSW2DState"
    back = io.dataset_to_state(ds, NonlinearSW2DState, device="cpu")
    assert torch.equal(back.h, st.h) and torch.equal(back.u, st.u)
    q = BaroclinicQGState(q=torch.arange(3 * 6 * 8, dtype=torch.float32).reshape(3, 6, 8))
    ds = io.state_to_dataset(q)
    assert ds["q"].dims == ("layer", "y", "x") and ds["q"].shape == (3, 6, 8)
    assert torch.equal(io.dataset_to_state(ds, BaroclinicQGState, device="cpu").q, q.q)
    assert io.state_to_dataset(BarotropicQGState(q=torch.zeros((8, 12))))["q"].dims == ("y", "x")


def test_time_attrs_and_auto_class():
    """test_io_xarray.py:92-130."""
    ds = io.state_to_dataset(_swm(), time=42.0, attrs={"experiment": "smoke", "config_hash": "abc"})
    assert ds["h"].dims == ("time", "y", "x") and ds["h"].shape == (1, 8, 12)
    assert np.array_equal(ds["time"].values, [42.0]) and "time" in ds.coords
    assert ds.attrs["experiment"] == "smoke" and ds.attrs["state_class"] == "NonlinearSW2DState"
    back = io.dataset_to_state(ds, device="cpu")
    assert isinstance(back, NonlinearSW2DState) and back.h.shape == (8, 12)
    with pytest.raises(ValueError, match="state_class"):
        io.dataset_to_state(io.Dataset({"h": (("y", "x"), np.zeros((4, 4)))}), device="cpu")
    with pytest.raises(ValueError, match="missing variable 'u'"):
        io.dataset_to_state(io.Dataset({"h": (("y", "x"), np.zeros((4, 4)))}), NonlinearSW2DState, device="cpu")


def test_snapshots():
    """test_io_xarray.py:136-196."""
    nt, ny, nx = 5, 8, 12
    h = torch.arange(nt * ny * nx, dtype=torch.float32).reshape(nt, ny, nx)
    snaps = NonlinearSW2DState(h=h, u=torch.zeros_like(h), v=torch.zeros_like(h))
    ts = torch.linspace(0.0, 4.0, nt)
    ds = io.snapshots_to_dataset(snaps, ts)
    assert ds["h"].dims == ("time", "y", "x") and np.array_equal(ds["time"].values, ts.numpy())
    assert torch.equal(io.dataset_to_state(ds, NonlinearSW2DState, device="cpu").h, h[-1])
    q = BaroclinicQGState(q=torch.zeros((4, 2, 6, 8)))
    assert io.snapshots_to_dataset(q, torch.linspace(0, 3, 4))["q"].dims == ("time", "layer", "y", "x")
    bad = NonlinearSW2DState(h=torch.zeros((5, 4, 4)), u=torch.zeros((5, 4, 4)), v=torch.zeros((5, 4, 4)))
    with pytest.raises(ValueError, match="leading dim 5 but ts has length 3"):
        io.snapshots_to_dataset(bad, torch.tensor([0.0, 1.0, 2.0]))
    with pytest.raises(ValueError, match="ts must be 1-D"):
        io.snapshots_to_dataset(bad, torch.zeros((5, 1)))


def test_zarr_round_trip_and_append(tmp_path):
    """test_io_xarray.py:198-262."""
    st = MultilayerSW2DState(h=torch.rand((2, 6, 10)), u=torch.rand((2, 6, 10)), v=torch.rand((2, 6, 10)))
    store = tmp_path / "state.zarr"
    io.save_dataset(io.state_to_dataset(st, attrs={"experiment": "test"}), store)
    assert store.is_dir()
    loaded = io.load_dataset(store)
    back = io.dataset_to_state(loaded, device="cpu")
    assert isinstance(back, MultilayerSW2DState) and torch.equal(back.h, st.h) and torch.equal(back.v, st.v)
    assert loaded.attrs["experiment"] == "test"
    q = BaroclinicQGState(q=torch.arange(3 * 2 * 6 * 8, dtype=torch.float32).reshape(3, 2, 6, 8))
    ts = torch.linspace(0.0, 2.0, 3)
    io.save_dataset(io.snapshots_to_dataset(q, ts), tmp_path / "snap.zarr")
    got = io.load_dataset(tmp_path / "snap.zarr")
    assert np.array_equal(got["q"].values, q.q.numpy()) and np.array_equal(got["time"].values, ts.numpy())
    assert "time" in got.coords and "q" in got.data_vars
    first = NonlinearSW2DState(h=torch.ones((3, 4, 4)), u=torch.zeros((3, 4, 4)), v=torch.zeros((3, 4, 4)))
    io.save_dataset(io.snapshots_to_dataset(first, torch.arange(3, dtype=torch.float32)), tmp_path / "app.zarr")
    second = NonlinearSW2DState(h=torch.full((2, 4, 4), 2.0), u=torch.zeros((2, 4, 4)), v=torch.zeros((2, 4, 4)))
    io.append_to_dataset(io.snapshots_to_dataset(second, torch.arange(3, 5, dtype=torch.float32)),
                         tmp_path / "app.zarr", append_dim="time")
    got = io.load_dataset(tmp_path / "app.zarr")
    assert got["h"].shape == (5, 4, 4) and np.array_equal(got["time"].values, np.arange(5, dtype=np.float32))
    assert np.array_equal(got["h"].values[3:], np.full((2, 4, 4), 2.0, np.float32))
    with pytest.raises(FileExistsError):
        io.save_dataset(io.state_to_dataset(st), store, mode="w-")


def test_auto_import_allowlist(tmp_path):
    """test_io_xarray.py:270-311; a reference-written store (somax.*) maps
    to this package's class of the same name."""
    def tampered(module):
        ds = io.Dataset({"h": (("y", "x"), np.zeros((4, 4)))}, attrs={"state_class": "Whatever",
                                                                       "state_module": module})
        io.save_dataset(ds, tmp_path / "t.zarr")
        return io.load_dataset(tmp_path / "t.zarr")

    for module in ("os", "somaxxx.evil"):
        with pytest.raises(ValueError, match="refusing to auto-import"):
            io.dataset_to_state(tampered(module), device="cpu")
    with pytest.raises(ValueError, match="missing variable"):
        io.dataset_to_state(tampered("os"), state_class=NonlinearSW2DState, device="cpu")
    ds = io.state_to_dataset(_swm())
    ds.attrs["state_module"] = "somax._src.models.swm.nonlinear_2d"
    assert isinstance(io.dataset_to_state(ds, device="cpu"), NonlinearSW2DState)


def test_store_layout_follows_zarr_v3(tmp_path):
    """Group and array metadata and the chunk bytes, field by field, as the
    zarr v3 core specification defines them (bytes codec, little endian,
    default chunk keys 'c/0/0/0', dimension_names)."""
    q = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    ds = io.Dataset({"q": (("layer", "y", "x"), q)}, attrs={"state_class": "S", "n": np.int64(3)})
    io.save_dataset(ds, tmp_path / "s.zarr")
    group = json.loads((tmp_path / "s.zarr" / "zarr.json").read_text())
    assert group == {"zarr_format": 3, "node_type": "group", "attributes": {"state_class": "S", "n": 3}}
    meta = json.loads((tmp_path / "s.zarr" / "q" / "zarr.json").read_text())
    assert meta["zarr_format"] == 3 and meta["node_type"] == "array" and meta["shape"] == [2, 3, 4]
    assert meta["data_type"] == "float32" and meta["dimension_names"] == ["layer", "y", "x"]
    assert meta["chunk_grid"] == {"name": "regular", "configuration": {"chunk_shape": [2, 3, 4]}}
    assert meta["chunk_key_encoding"] == {"name": "default", "configuration": {"separator": "/"}}
    assert meta["codecs"] == [{"name": "bytes", "configuration": {"endian": "little"}}]
    raw = (tmp_path / "s.zarr" / "q" / "c" / "0" / "0" / "0").read_bytes()
    assert raw == q.astype("<f4").tobytes()


def test_reads_multichunk_gzip_and_missing_chunks(tmp_path):
    """A store as zarr-python writes it with several chunks, the gzip codec,
    a big-endian bytes codec and one chunk left out (fill value)."""
    root = tmp_path / "m.zarr"
    (root / "a" / "c").mkdir(parents=True)
    (root / "zarr.json").write_text(json.dumps({"zarr_format": 3, "node_type": "group", "attributes": {}}))
    a = np.arange(5 * 7, dtype=np.float64).reshape(5, 7)
    meta = {"zarr_format": 3, "node_type": "array", "shape": [5, 7], "data_type": "float64",
            "chunk_grid": {"name": "regular", "configuration": {"chunk_shape": [2, 4]}},
            "chunk_key_encoding": {"name": "default", "configuration": {"separator": "/"}}, "fill_value": "NaN",
            "codecs": [{"name": "bytes", "configuration": {"endian": "big"}}, {"name": "gzip",
                                                                                "configuration": {"level": 1}}],
            "attributes": {}, "dimension_names": ["y", "x"]}
    (root / "a" / "zarr.json").write_text(json.dumps(meta))
    want = a.copy()
    for i in range(3):
        for j in range(2):
            if (i, j) == (1, 1):
                want[2:4, 4:7] = np.nan
                continue
            chunk = np.zeros((2, 4))
            blk = a[2 * i:2 * i + 2, 4 * j:4 * j + 4]
            chunk[:blk.shape[0], :blk.shape[1]] = blk
            (root / "a" / "c" / str(i)).mkdir(exist_ok=True)
            (root / "a" / "c" / str(i) / str(j)).write_bytes(gzip.compress(chunk.astype(">f8").tobytes()))
    got = io.load_dataset(root)["a"]
    assert got.dims == ("y", "x") and np.array_equal(got.values, want, equal_nan=True)
    meta["codecs"] = [{"name": "bytes", "configuration": {"endian": "little"}}, {"name": "zstd"}]
    (root / "a" / "zarr.json").write_text(json.dumps(meta))
    with pytest.raises(ValueError, match="zstd"):
        io.load_dataset(root)


def test_failed_write_leaves_the_existing_store(tmp_path, monkeypatch):
    """save_dataset / append_to_dataset write into a sibling directory and swap
    it in only when complete: an error while writing keeps the old store."""
    from somax_amd import io

    path = tmp_path / "run.zarr"
    ds = io.Dataset({"q": io.DataArray(("time", "y"), np.arange(6.0).reshape(2, 3))},
                    {"time": io.DataArray(("time",), np.array([0.0, 1.0]))}, {"note": "first"})
    io.save_dataset(ds, path)
    before = io.load_dataset(path)

    def boom(*a, **k):
        raise OSError("disk full")

    monkeypatch.setattr(io, "_write_array", boom)
    new = io.Dataset({"q": io.DataArray(("time", "y"), np.ones((1, 3)))},
                     {"time": io.DataArray(("time",), np.array([2.0]))}, {})
    with pytest.raises(OSError, match="disk full"):
        io.append_to_dataset(new, path)
    monkeypatch.undo()
    after = io.load_dataset(path)
    np.testing.assert_array_equal(after.data_vars["q"].values, before.data_vars["q"].values)
    assert after.attrs["note"] == "first"
    assert sorted(p.name for p in tmp_path.iterdir()) == ["run.zarr"]  # no temporary left behind
    io.append_to_dataset(new, path)
    assert io.load_dataset(path).data_vars["q"].values.shape == (3, 3)
