"""Model states on disk: the zarr v3 layout of somax's xarray IO, without
xarray or zarr.

Mirror of ``somax/_src/io/xarray.py``: ``state_to_dataset`` (:71-114),
``snapshots_to_dataset`` (:117-169), ``dataset_to_state`` (:172-232),
``save_dataset`` / ``load_dataset`` / ``append_to_dataset`` (:235-286), with
the same names, dimension conventions (1-D ``x``, 2-D ``y x``, 3-D
``layer y x``, a leading ``time`` for snapshots), ``state_class`` /
``state_module`` attributes and error texts.  Neither xarray nor zarr is
installed here, so ``Dataset`` is a small container of NumPy arrays with the
parts of ``xarray.Dataset`` that these functions and their callers use
(``ds[name].dims / .shape / .values``, ``data_vars``, ``coords``,
``variables``, ``attrs``), and the store is written and read directly in the
zarr v3 format (https://zarr-specs.readthedocs.io/en/latest/v3/core/): a
group ``zarr.json`` carrying the attributes, one array per variable with
``dimension_names`` (what xarray reads as dimensions), one chunk per array,
the ``bytes`` codec (little-endian, uncompressed).  ``load_dataset`` also
reads multi-chunk arrays and the ``gzip`` codec; ``zstd``-compressed stores
(zarr-python's default) raise, as no zstd decoder is installed.

States are device tensors on the HIP path; ``dataset_to_state`` returns
tensors on ``device``.
"""

from __future__ import annotations

import dataclasses
import gzip
import importlib
import itertools
import json
import os
import shutil
from pathlib import Path

import numpy as np

_DIMS_BY_RANK = {1: ("x",), 2: ("y", "x"), 3: ("layer", "y", "x")}


def _infer_dims(rank: int, *, time_axis: bool) -> tuple[str, ...]:
    """xarray.py:46-60."""
    base = _DIMS_BY_RANK.get(rank) or tuple(f"dim{i}" for i in range(rank))
    return ("time", *base) if time_axis else base


def _state_attrs(state_class: type) -> dict[str, str]:
    return {"state_class": state_class.__name__, "state_module": state_class.__module__}


def _host(x) -> np.ndarray:
    if hasattr(x, "detach"):
        return x.detach().cpu().numpy()
    return np.asarray(x)


class DataArray:
    """One variable: dimension names and a NumPy array."""

    def __init__(self, dims, values):
        self.dims = tuple(dims)
        self.values = np.asarray(values)
        if self.values.ndim != len(self.dims):
            raise ValueError(f"{len(self.dims)} dimension names for an array of rank {self.values.ndim}")

    @property
    def shape(self):
        return self.values.shape

    def isel(self, **index):
        out, dims = self.values, list(self.dims)
        for name, i in index.items():
            ax = dims.index(name)
            out = np.take(out, i, axis=ax)
            dims.pop(ax)
        return DataArray(dims, out)


class Dataset:
    """Data variables, coordinates and attributes (the subset of
    ``xarray.Dataset`` the IO functions use)."""

    def __init__(self, data_vars=None, coords=None, attrs=None):
        self.data_vars = {k: v if isinstance(v, DataArray) else DataArray(*v) for k, v in (data_vars or {}).items()}
        self.coords = {k: v if isinstance(v, DataArray) else DataArray(*v) for k, v in (coords or {}).items()}
        self.attrs = dict(attrs or {})

    @property
    def variables(self):
        return {**self.coords, **self.data_vars}

    def __getitem__(self, name) -> DataArray:
        if name in self.data_vars:
            return self.data_vars[name]
        return self.coords[name]

    def __contains__(self, name) -> bool:
        return name in self.data_vars or name in self.coords


def state_to_dataset(state, *, time: float | None = None, attrs: dict | None = None) -> Dataset:
    """One state as a Dataset; ``time`` adds a singleton ``time`` axis and
    coordinate (xarray.py:71-114)."""
    data = {}
    for f in dataclasses.fields(state):
        leaf = _host(getattr(state, f.name))
        dims = _infer_dims(leaf.ndim, time_axis=time is not None)
        data[f.name] = (dims, leaf[None] if time is not None else leaf)
    coords = {"time": (("time",), np.asarray([time], dtype=np.float64))} if time is not None else {}
    ds = Dataset(data, coords, _state_attrs(type(state)))
    ds.attrs.update(attrs or {})
    return ds


def snapshots_to_dataset(snapshots, ts, *, state_class: type | None = None, attrs: dict | None = None) -> Dataset:
    """Time-stacked states (``Solution.ys``) as a Dataset with a ``time``
    coordinate (xarray.py:117-169)."""
    state_class = state_class or type(snapshots)
    ts_np = _host(ts)
    if ts_np.ndim != 1:
        raise ValueError(f"ts must be 1-D, got shape {ts_np.shape}")
    data = {}
    for f in dataclasses.fields(state_class):
        leaf = _host(getattr(snapshots, f.name))
        if leaf.shape[0] != ts_np.shape[0]:
            raise ValueError(f"snapshot leaf {f.name!r} has leading dim {leaf.shape[0]} "
                             f"but ts has length {ts_np.shape[0]}")
        data[f.name] = (_infer_dims(leaf.ndim - 1, time_axis=True), leaf)
    ds = Dataset(data, {"time": (("time",), ts_np)}, _state_attrs(state_class))
    ds.attrs.update(attrs or {})
    return ds


def _resolve_class(module_name: str, class_name: str) -> type:
    """Modules of this package import as themselves; a store written by the
    reference (``somax.*``) maps to this package's class of the same name,
    so it restarts a drop-in model.  Anything else is refused, as the
    reference refuses modules outside its namespace (xarray.py:203-217)."""
    if module_name == "somax_amd" or module_name.startswith("somax_amd."):
        return getattr(importlib.import_module(module_name), class_name)
    if module_name == "somax" or module_name.startswith("somax."):
        pkg = importlib.import_module("somax_amd")
        if hasattr(pkg, class_name):
            return getattr(pkg, class_name)
        raise ValueError(f"state class {class_name!r} of {module_name!r} has no counterpart in somax_amd; "
                         f"pass state_class explicitly")
    raise ValueError(
        f"refusing to auto-import state_module {module_name!r}: only modules under 'somax.' (or "
        f"'somax_amd.') are allowlisted for auto-recovery. Pass state_class explicitly to load this dataset."
    )


def dataset_to_state(ds: Dataset, state_class: type | None = None, *, time_index: int = -1, device="cuda"):
    """A state from a Dataset: the class from the arguments or from the
    ``state_class`` / ``state_module`` attributes, the ``time_index`` slice
    of time-stacked variables (xarray.py:172-232); fields are tensors on
    ``device``."""
    import torch

    if state_class is None:
        try:
            module_name, class_name = ds.attrs["state_module"], ds.attrs["state_class"]
        except KeyError as exc:
            raise ValueError("dataset_to_state called without state_class and the Dataset is missing "
                             "'state_class'/'state_module' attrs") from exc
        state_class = _resolve_class(module_name, class_name)
    kwargs = {}
    for f in dataclasses.fields(state_class):
        if f.name not in ds.variables:
            raise ValueError(f"Dataset is missing variable {f.name!r} required by {state_class.__name__}")
        var = ds[f.name]
        if "time" in var.dims:
            var = var.isel(time=time_index)
        kwargs[f.name] = torch.from_numpy(np.ascontiguousarray(var.values)).to(device)
    return state_class(**kwargs)


# ---------------------------------------------------------------------------
# zarr v3 directory store
# ---------------------------------------------------------------------------
_DTYPES = {"bool": "bool", "int8": "int8", "int16": "int16", "int32": "int32", "int64": "int64", "uint8": "uint8",
           "uint16": "uint16", "uint32": "uint32", "uint64": "uint64", "float32": "float32", "float64": "float64"}


def _json_attr(v):
    if isinstance(v, (np.integer,)):
        return int(v)
    if isinstance(v, (np.floating,)):
        return float(v)
    if isinstance(v, np.ndarray):
        return v.tolist()
    return v


def _fill_json(dtype: np.dtype):
    return 0.0 if dtype.kind == "f" else (False if dtype.kind == "b" else 0)


def _write_array(root: Path, name: str, var: DataArray) -> None:
    values = np.ascontiguousarray(var.values)
    dt = values.dtype.name
    if dt not in _DTYPES:
        raise TypeError(f"variable {name!r}: dtype {dt} is not stored")
    d = root / name
    d.mkdir(parents=True, exist_ok=True)
    meta = {
        "zarr_format": 3,
        "node_type": "array",
        "shape": list(values.shape),
        "data_type": _DTYPES[dt],
        "chunk_grid": {"name": "regular", "configuration": {"chunk_shape": [max(s, 1) for s in values.shape]}},
        "chunk_key_encoding": {"name": "default", "configuration": {"separator": "/"}},
        "fill_value": _fill_json(values.dtype),
        "codecs": [{"name": "bytes", "configuration": {"endian": "little"}}],
        "attributes": {},
        "dimension_names": list(var.dims),
    }
    (d / "zarr.json").write_text(json.dumps(meta, indent=2))
    if values.size == 0:
        return
    key = d / "c"
    for _ in range(values.ndim):
        key = key / "0"
    key.parent.mkdir(parents=True, exist_ok=True)
    key.write_bytes(values.astype(values.dtype.newbyteorder("<"), copy=False).tobytes(order="C"))


def save_dataset(ds: Dataset, path, *, mode: str = "w") -> None:
    """Writes ``ds`` as a zarr v3 directory store (xarray.py:235-249);
    ``mode`` ``"w"`` overwrites, ``"w-"`` fails if the store exists."""
    path = Path(path)
    if mode not in ("w", "w-"):
        raise ValueError(f"mode must be 'w' or 'w-', got {mode!r}")
    if path.exists() and mode == "w-":
        raise FileExistsError(f"{path} exists")
    # The store is written whole into a sibling directory first and swapped in
    # afterwards, so an error or interrupt while writing leaves any existing
    # store at `path` as it was.
    path.parent.mkdir(parents=True, exist_ok=True)
    tmp = path.with_name(f".{path.name}.tmp-{os.getpid()}")
    if tmp.exists():
        shutil.rmtree(tmp)
    tmp.mkdir()
    try:
        (tmp / "zarr.json").write_text(json.dumps(
            {"zarr_format": 3, "node_type": "group",
             "attributes": {k: _json_attr(v) for k, v in ds.attrs.items()}}, indent=2))
        for name, var in ds.variables.items():
            _write_array(tmp, name, var)
    except BaseException:
        shutil.rmtree(tmp, ignore_errors=True)
        raise
    if path.exists():
        old = path.with_name(f".{path.name}.old-{os.getpid()}")
        if old.exists():
            shutil.rmtree(old)
        os.replace(path, old)
        os.replace(tmp, path)
        shutil.rmtree(old)
    else:
        os.replace(tmp, path)


def _decode_chunk(raw: bytes, codecs, dtype: np.dtype, chunk_shape) -> np.ndarray:
    order = "C"
    for c in reversed(codecs):
        name, cfg = c["name"], c.get("configuration", {})
        if name == "gzip":
            raw = gzip.decompress(raw)
        elif name == "crc32c":
            raw = raw[:-4]
        elif name in ("zstd", "blosc"):
            raise ValueError(f"zarr codec {name!r} needs a decoder that is not installed here")
        elif name == "bytes":
            dtype = dtype.newbyteorder("<" if cfg.get("endian", "little") == "little" else ">")
        elif name == "transpose":
            order = tuple(cfg["order"])
        else:
            raise ValueError(f"unsupported zarr codec {name!r}")
    arr = np.frombuffer(raw, dtype=dtype).reshape(
        [chunk_shape[i] for i in order] if order != "C" else chunk_shape)
    if order != "C":
        arr = np.transpose(arr, np.argsort(order))
    return arr.astype(dtype.newbyteorder("="))


def _read_array(d: Path) -> DataArray:
    meta = json.loads((d / "zarr.json").read_text())
    if meta.get("zarr_format") != 3 or meta.get("node_type") != "array":
        raise ValueError(f"{d} is not a zarr v3 array")
    shape = tuple(meta["shape"])
    dtype = np.dtype(meta["data_type"])
    grid = meta["chunk_grid"]
    if grid["name"] != "regular":
        raise ValueError(f"unsupported chunk grid {grid['name']!r}")
    cs = tuple(grid["configuration"]["chunk_shape"])
    enc = meta.get("chunk_key_encoding", {"name": "default"})
    sep = enc.get("configuration", {}).get("separator", "/" if enc["name"] == "default" else ".")
    fill = meta.get("fill_value", 0)
    if isinstance(fill, str):
        fill = {"NaN": np.nan, "Infinity": np.inf, "-Infinity": -np.inf}.get(fill, 0)
    out = np.full(shape, fill, dtype=dtype)
    dims = meta.get("dimension_names") or meta.get("attributes", {}).get("_ARRAY_DIMENSIONS") or \
        [f"dim{i}" for i in range(len(shape))]
    if out.size == 0:
        return DataArray(dims, out)
    counts = [-(-s // c) for s, c in zip(shape, cs)]
    for idx in itertools.product(*[range(n) for n in counts]):
        parts = [str(i) for i in idx]
        key = ("c" + sep + sep.join(parts)) if enc["name"] == "default" else (sep.join(parts) or "0")
        f = d.joinpath(*key.split("/"))
        if not f.exists():
            continue  # an absent chunk holds the fill value
        chunk = _decode_chunk(f.read_bytes(), meta["codecs"], dtype, cs)
        sl = tuple(slice(i * c, min((i + 1) * c, s)) for i, c, s in zip(idx, cs, shape))
        out[sl] = chunk[tuple(slice(0, x.stop - x.start) for x in sl)]
    return DataArray(dims, out)


def load_dataset(path) -> Dataset:
    """Reads a zarr v3 store (xarray.py:252-266) into memory.  A variable
    whose only dimension is its own name is a coordinate, as in xarray."""
    path = Path(path)
    meta = json.loads((path / "zarr.json").read_text())
    if meta.get("zarr_format") != 3 or meta.get("node_type") != "group":
        raise ValueError(f"{path} is not a zarr v3 group")
    data, coords = {}, {}
    for d in sorted(p for p in path.iterdir() if (p / "zarr.json").is_file()):
        var = _read_array(d)
        (coords if var.dims == (d.name,) else data)[d.name] = var
    return Dataset(data, coords, meta.get("attributes", {}))


def append_to_dataset(ds: Dataset, path, *, append_dim: str = "time") -> None:
    """Extends a store along ``append_dim`` (xarray.py:269-286): every
    variable with that dimension is concatenated; the others must match.
    The merged store replaces the old one only once it is written whole
    (``save_dataset``), so a failed append leaves the old store intact."""
    old = load_dataset(path)
    merged_data, merged_coords = {}, {}
    for target, src_old, src_new in ((merged_data, old.data_vars, ds.data_vars),
                                     (merged_coords, old.coords, ds.coords)):
        for name, var in src_old.items():
            if name not in src_new:
                raise ValueError(f"variable {name!r} is missing from the appended dataset")
            new = src_new[name]
            if append_dim in var.dims:
                ax = var.dims.index(append_dim)
                target[name] = DataArray(var.dims, np.concatenate([var.values, new.values], axis=ax))
            else:
                target[name] = var
    save_dataset(Dataset(merged_data, merged_coords, {**old.attrs, **ds.attrs}), path)
