"""Plot-ready experiment artifacts in the reference's ``suetes-plot-data-v1``
layout (suetes/shared/experiment.py:16-230), so that outputs of this
implementation sit beside the reference's own runs and compare field by
field.

Kept from the reference: the bundle layout ``<root>/<kind>/<case>/<execution>``
with ``data/`` and ``figures/`` (and the same path checks), the provenance
(schema, experiment name, UTC time, git commit, caller metadata) in the
NetCDF global attributes and in a JSON sidecar, booleans stored as 0/1 and
other non-scalar attributes as JSON, and the write through a temporary file.

Different: this image has neither xarray nor netCDF4, so a dataset is the
small :class:`PlotDataset` container (an xarray ``Dataset`` is accepted
too when xarray is importable) and the file is NetCDF classic with 64-bit
offsets written by SciPy -- uncompressed, and readable by xarray's netCDF4
and scipy engines alike.  String coordinates are stored as character
arrays with a trailing ``string<N>`` dimension, the convention xarray
decodes back to strings.
"""
from __future__ import annotations

import argparse
import json
import os
import subprocess
from datetime import datetime, timezone
from pathlib import Path
from typing import Tuple

import numpy as np

SCHEMA = "suetes-plot-data-v1"
ARTIFACT_KINDS = frozenset({"benchmarks", "experiments", "runs", "verification"})


class ExperimentLayout:
    """Paths of one self-contained execution: ``<output_root>/<kind>/<case>/<execution>``
    with ``data/`` and ``figures/`` inside (experiment.py:20-63)."""

    def __init__(self, *, kind: str, case: str, execution: str = "default", output_root="output") -> None:
        if kind not in ARTIFACT_KINDS:
            raise ValueError(f"Unknown artifact kind {kind!r}; expected one of {', '.join(sorted(ARTIFACT_KINDS))}")
        for label, value in (("case", case), ("execution", execution)):
            p = Path(value)
            if not value or p.is_absolute() or len(p.parts) != 1 or value in {".", ".."}:
                raise ValueError(f"{label} must be a single non-empty path component")
        self.kind, self.case, self.execution = kind, case, execution
        self.output_root = Path(output_root)

    @property
    def root(self) -> Path:
        return self.output_root / self.kind / self.case / self.execution

    @property
    def data(self) -> Path:
        return self.root / "data"

    @property
    def figures(self) -> Path:
        return self.root / "figures"

    def create(self) -> "ExperimentLayout":
        self.data.mkdir(parents=True, exist_ok=True)
        self.figures.mkdir(parents=True, exist_ok=True)
        return self


def resolve_data_dir(*, kind, case, execution="default", output_root="output", output_dir=None) -> Path:
    """An explicit legacy directory, or the bundle's data directory (created)."""
    if output_dir is not None:
        p = Path(output_dir)
        p.mkdir(parents=True, exist_ok=True)
        return p
    return ExperimentLayout(kind=kind, case=case, execution=execution, output_root=output_root).create().data


def figure_dir_for(artifact) -> Path:
    """The bundle's figure directory for an artifact in its ``data/``, else the artifact's directory."""
    artifact = Path(artifact)
    return artifact.parent.parent / "figures" if artifact.parent.name == "data" else artifact.parent


def artifact_from_bundle(source, pattern: str = "artifact.nc") -> Path:
    """One artifact from a bundle root, a data directory or an explicit file."""
    source = Path(source)
    if source.is_file():
        return source
    data_dir = source / "data" if (source / "data").is_dir() else source
    matches = sorted(data_dir.glob(pattern))
    if len(matches) != 1:
        raise ValueError(f"Expected exactly one {pattern!r} artifact in {data_dir}, found {len(matches)}")
    return matches[0]


def add_experiment_args(parser: argparse.ArgumentParser, *, default_output_root=Path("output")) -> None:
    parser.add_argument("--output-root", type=Path, default=Path(default_output_root),
                        help="Root directory for the standard experiment bundle")
    parser.add_argument("--name", default="default", help="Execution name for the standard experiment bundle")
    parser.add_argument("--output-dir", type=Path, default=None,
                        help="Explicit legacy directory to bypass the standard bundle layout")


def setup_experiment_directories(args: argparse.Namespace, *, kind: str, case: str) -> Tuple[Path, Path]:
    """(data_dir, figure_dir) from the standard arguments."""
    if getattr(args, "output_dir", None) is not None:
        p = Path(args.output_dir)
        p.mkdir(parents=True, exist_ok=True)
        return p, p
    layout = ExperimentLayout(kind=kind, case=case, execution=getattr(args, "name", "default"),
                              output_root=getattr(args, "output_root", Path("output"))).create()
    return layout.data, layout.figures


def _git_commit() -> str:
    try:
        return subprocess.check_output(["git", "rev-parse", "HEAD"], text=True, stderr=subprocess.DEVNULL).strip()
    except (OSError, subprocess.SubprocessError):
        return "unknown"


class PlotDataset:
    """Named arrays over named dimensions: ``data_vars[name] = (dims, array)``,
    ``coords[name] = array`` (1-D over the dimension of that name) or
    ``(dims, array)``, ``attrs`` a dict of global attributes."""

    def __init__(self, data_vars=None, coords=None, attrs=None):
        self.data_vars = {k: (tuple(v[0]) if isinstance(v[0], (tuple, list)) else (v[0],), np.asarray(v[1]))
                          for k, v in (data_vars or {}).items()}
        self.coords = {}
        for k, v in (coords or {}).items():
            if isinstance(v, tuple):
                self.coords[k] = (tuple(v[0]) if isinstance(v[0], (tuple, list)) else (v[0],), np.asarray(v[1]))
            else:
                self.coords[k] = ((k,), np.asarray(v))
        self.attrs = dict(attrs or {})
        for name, (dims, arr) in list(self.data_vars.items()) + list(self.coords.items()):
            if len(dims) != arr.ndim:
                raise ValueError(f"{name}: {len(dims)} dimension names for a {arr.ndim}-d array")

    def __getitem__(self, name):
        return (self.data_vars.get(name) or self.coords[name])[1]


def _as_plot_dataset(ds) -> PlotDataset:
    if isinstance(ds, PlotDataset):
        return ds
    if hasattr(ds, "data_vars") and hasattr(ds, "coords"):  # an xarray Dataset
        return PlotDataset({k: (v.dims, v.values) for k, v in ds.data_vars.items()},
                           {k: (v.dims, v.values) for k, v in ds.coords.items()}, dict(ds.attrs))
    raise TypeError("expected a PlotDataset or an xarray Dataset")


def _attribute(value):
    """NetCDF attributes take numbers and strings: booleans become 0 / 1 and
    anything else JSON (experiment.py:141-149)."""
    if isinstance(value, (bool, np.bool_)):
        return int(value)
    if isinstance(value, (str, int, float, np.number)):
        return value
    return json.dumps(value)


def save_plot_dataset(dataset, path, *, experiment: str, metadata: dict | None = None) -> Path:
    """Write the NetCDF artifact and its JSON provenance sidecar
    (experiment.py:120-171)."""
    from scipy.io import netcdf_file
    ds = _as_plot_dataset(dataset)
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    provenance = {"artifact_schema": SCHEMA, "experiment": experiment,
                  "created_utc": datetime.now(timezone.utc).isoformat(), "git_commit": _git_commit(),
                  **dict(metadata or {})}
    tmp = path.with_name(f".{path.name}.tmp")
    try:
        f = netcdf_file(str(tmp), "w", version=2)
        try:
            dims = {}
            items = list(ds.coords.items()) + list(ds.data_vars.items())
            for name, (dn, arr) in items:
                for d, size in zip(dn, arr.shape):
                    if dims.setdefault(d, size) != size:
                        raise ValueError(f"dimension {d!r} has sizes {dims[d]} and {size}")
            for d, size in dims.items():
                f.createDimension(d, size)
            for name, (dn, arr) in items:
                if arr.dtype.kind in "US":  # strings: characters with a trailing string<N> dimension
                    strs = [str(s).encode() for s in arr.ravel()]
                    n = max(1, max((len(s) for s in strs), default=1))
                    sd = f"string{n}"
                    if sd not in f.dimensions:
                        f.createDimension(sd, n)
                    chars = np.array([list(s.ljust(n, b"\0")) for s in strs], dtype="u1").reshape(arr.shape + (n,))
                    v = f.createVariable(name, "c", dn + (sd,))
                    v[...] = chars.view("S1")
                    continue
                a = arr.astype(np.float64) if arr.dtype.kind == "f" else arr
                if a.dtype.kind == "b":
                    a = a.astype(np.int8)
                if a.dtype.kind in "iu" and a.dtype.itemsize > 4:
                    a = a.astype(np.int32)
                v = f.createVariable(name, a.dtype, dn)
                v[...] = a
            for k, val in list(ds.attrs.items()) + list(provenance.items()):
                setattr(f, k, _attribute(val))
        finally:
            f.close()
        os.replace(tmp, path)
    finally:
        if tmp.exists():
            tmp.unlink()
    with path.with_suffix(".json").open("w") as stream:
        json.dump(provenance, stream, indent=2)
    return path


def load_plot_dataset(path) -> PlotDataset:
    """Read an artifact written by :func:`save_plot_dataset` (or any NetCDF
    classic file) back into a :class:`PlotDataset`."""
    from scipy.io import netcdf_file
    f = netcdf_file(str(path), "r", mmap=False)
    try:
        data_vars, coords = {}, {}
        for name, v in f.variables.items():
            dn = tuple(v.dimensions)
            arr = np.array(v[...])
            if v.typecode() == "c":
                arr = np.array([b"".join(r).rstrip(b"\0").decode() for r in arr.reshape(-1, arr.shape[-1])])
                arr = arr.reshape(v.shape[:-1])
                dn = dn[:-1]
            (coords if dn == (name,) else data_vars)[name] = (dn, arr)
        attrs = {k: (v.decode() if isinstance(v, bytes) else v) for k, v in f._attributes.items()}
    finally:
        f.close()
    return PlotDataset(data_vars, coords, attrs)


STATE_DIMS = {"u": ("x_u", "y", "z"), "v": ("x", "y_v", "z"), "w": ("x", "y", "z_w"),
              "eta_dot": ("x", "y", "z_w")}


def state_dataset(state, grid, *, time=None, fields=("u", "v", "w", "eta_dot", "pi", "th_v", "rho"),
                  attrs=None) -> PlotDataset:
    """A SISL state (reference-shaped arrays or tensors, `step`'s output) as
    a dataset on the C grid's coordinates: mass points (x, y, z), u faces
    x_u, v faces y_v, w levels z_w (logical heights), plus the terrain
    height Z of the mass points; tracers and history keys are included when
    named in ``fields``."""
    def npy(a):
        return a.detach().cpu().numpy() if hasattr(a, "detach") else np.asarray(a)
    coords = {"x": npy(grid.x_m), "x_u": npy(grid.x_c), "y": npy(grid.y_m), "y_v": npy(grid.y_c),
              "z": npy(grid.z_m), "z_w": npy(grid.z_c)}
    data = {}
    for k in fields:
        if k not in state:
            continue
        arr = npy(state[k])
        data[k] = (STATE_DIMS.get(k, ("x", "y", "z")), arr[None] if time is not None else arr)
        if time is not None:
            data[k] = (("time",) + data[k][0], data[k][1])
    data["Z"] = (("x", "y", "z"), npy(grid.Z_m))
    if time is not None:
        coords["time"] = np.atleast_1d(np.asarray(time, dtype=np.float64))
    return PlotDataset(data, coords, dict(attrs or {}))
