"""Artifact writer in the reference's suetes-plot-data-v1 layout
(suetes/shared/experiment.py), following the reference's own tests of it
(tests/test_experiment_layout.py): bundle directories, unsafe path
components, legacy explicit directories, artifact resolution, boolean
attributes, the provenance sidecar; plus a SciPy read-back of the NetCDF
file (an independent reader) and a SISL state written on the C grid's
coordinates."""
import json
from pathlib import Path

import numpy as np
import pytest
import torch

from suetes_amd.experiment import (SCHEMA, ExperimentLayout, PlotDataset, artifact_from_bundle,
                                   figure_dir_for, load_plot_dataset, resolve_data_dir, save_plot_dataset,
                                   state_dataset)


def test_artifact_layout_creates_bundle_directories(tmp_path: Path):
    layout = ExperimentLayout(kind="benchmarks", case="rising_bubble_3d", execution="smoke",
                              output_root=tmp_path).create()
    assert layout.root == tmp_path / "benchmarks" / "rising_bubble_3d" / "smoke"
    assert layout.data.is_dir() and layout.figures.is_dir()


@pytest.mark.parametrize("value", ["", ".", "..", "nested/name"])
def test_artifact_layout_rejects_unsafe_components(value):
    with pytest.raises(ValueError):
        ExperimentLayout(kind="benchmarks", case=value)
    with pytest.raises(ValueError):
        ExperimentLayout(kind="nonsense", case="ok")


def test_resolve_data_dir_uses_bundle_or_explicit_legacy_path(tmp_path: Path):
    standard = resolve_data_dir(kind="experiments", case="example", execution="paper", output_root=tmp_path)
    assert standard == tmp_path / "experiments" / "example" / "paper" / "data"
    assert figure_dir_for(standard / "artifact.nc") == tmp_path / "experiments" / "example" / "paper" / "figures"
    explicit = resolve_data_dir(kind="experiments", case="ignored", output_dir=tmp_path / "legacy")
    assert explicit == tmp_path / "legacy"
    assert figure_dir_for(explicit / "artifact.nc") == explicit


def test_artifact_from_bundle_resolves_data_file(tmp_path: Path):
    layout = ExperimentLayout(kind="experiments", case="example", output_root=tmp_path).create()
    artifact = layout.data / "artifact.nc"
    artifact.touch()
    assert artifact_from_bundle(layout.root) == artifact
    assert artifact_from_bundle(layout.data) == artifact
    assert artifact_from_bundle(artifact) == artifact


def test_save_plot_dataset_round_trip_with_boolean_attribute_and_sidecar(tmp_path: Path):
    ds = PlotDataset(
        data_vars={"theta_perturbation": (("core", "time", "x"), np.arange(12.0).reshape(2, 2, 3)),
                   "mass_drift": ("core", [1.0e-10, 2.0e-10])},
        coords={"core": np.array(["sisl", "split-explicit"]), "time": [0.0, 5.0], "x": [-1.0, 0.0, 1.0]},
        attrs={"periodic_x": True, "dx_m": 50.0, "grid": [64, 64, 64]})
    artifact = save_plot_dataset(ds, tmp_path / "artifact.nc", experiment="round_trip",
                                 metadata={"workload": "bubble"})
    assert not (tmp_path / ".artifact.nc.tmp").exists()
    back = load_plot_dataset(artifact)
    assert back.attrs["periodic_x"] == 1 and back.attrs["dx_m"] == 50.0
    assert json.loads(back.attrs["grid"]) == [64, 64, 64]
    assert back.attrs["artifact_schema"] == SCHEMA and back.attrs["experiment"] == "round_trip"
    np.testing.assert_array_equal(back["theta_perturbation"], ds["theta_perturbation"])
    np.testing.assert_array_equal(back["mass_drift"], [1.0e-10, 2.0e-10])
    assert list(back["core"]) == ["sisl", "split-explicit"]
    side = json.loads(artifact.with_suffix(".json").read_text())
    assert side["artifact_schema"] == SCHEMA and side["workload"] == "bubble" and "created_utc" in side
    # an independent reader of the same file
    from scipy.io import netcdf_file
    f = netcdf_file(str(artifact), "r", mmap=False)
    try:
        assert f.dimensions["x"] == 3 and f.variables["theta_perturbation"].dimensions == ("core", "time", "x")
        np.testing.assert_array_equal(f.variables["x"][:], [-1.0, 0.0, 1.0])
    finally:
        f.close()


def test_state_dataset_on_the_c_grid(tmp_path: Path):
    import suetes_amd as S
    g = S.RegionalGrid3D(6, 5, 4, 1000.0, 1000.0, 500.0, 45.0, 0.0)
    rng = np.random.default_rng(0)
    st = {"u": torch.from_numpy(rng.standard_normal((7, 5, 4))), "v": torch.from_numpy(rng.standard_normal((6, 6, 4))),
          "w": torch.from_numpy(rng.standard_normal((6, 5, 5))), "pi": torch.from_numpy(rng.standard_normal((6, 5, 4))),
          "th_v": torch.from_numpy(rng.standard_normal((6, 5, 4))), "is_first_step": 0.0}
    ds = state_dataset(st, g, time=30.0, attrs={"dt_s": 10.0})
    assert ds.data_vars["u"][0] == ("time", "x_u", "y", "z") and ds["u"].shape == (1, 7, 5, 4)
    assert ds.data_vars["w"][0] == ("time", "x", "y", "z_w")
    assert "eta_dot" not in ds.data_vars and ds["Z"].shape == (6, 5, 4)
    path = save_plot_dataset(ds, tmp_path / "data" / "artifact.nc", experiment="sisl_state")
    back = load_plot_dataset(path)
    np.testing.assert_array_equal(back["v"][0], st["v"].numpy())
    np.testing.assert_array_equal(back["x_u"], g.x_c.numpy())
    assert back.attrs["dt_s"] == 10.0
