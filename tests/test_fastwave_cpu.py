"""fast-wave compatible runner: argument handling, prepared-directory parsing
and output files, checked on the CPU (the failing cases stop before any GPU
work; the GPU runs are in test_gpu_fastwave.py)."""

import importlib

import numpy as np
import pytest

from _prepared import write_prepared
from _rs import rs

FW = importlib.import_module("3d_ravesim_amd.fastwave")
MS = importlib.import_module("3d_ravesim_amd.multisim")


def test_cli_requires_source_index(capsys):
    assert FW.main(["/nonexistent"]) == 1
    assert "Source index is required" in capsys.readouterr().err


def test_cli_reports_failures_like_fast_wave(tmp_path, capsys):
    assert FW.main([str(tmp_path), "-s", "0"]) == 1
    err = capsys.readouterr().err
    assert err.startswith("uncaught exception in the fast-wave framework:")
    assert "config.yaml" in err


def test_sub_dir_is_zero_padded(tmp_path):
    assert MS.get_sub_dir(tmp_path, 7) == tmp_path / "00000007"
    assert MS.get_sub_dir(str(tmp_path), 12345678).name == "12345678"


def test_point_source_geometry_y_origin():
    g = MS.point_source_geometry_2d({"type": "point", "x": 1e-6, "y": -2e-6, "z": 0.5}, {})
    assert g == {"type": "point", "units": "m", "x": 1e-6, "y": -2e-6, "z": 0.5,
                 "y_origin": "explicit"}
    g = MS.point_source_geometry_2d({"type": "point", "x": 1e-6, "z": 0.5},
                                    {"y_range": [0.0, 0.0]})
    assert g["y"] == 0.0 and g["y_origin"] == "legacy_default_zero"
    with pytest.raises(ValueError, match="y_range is nonzero"):
        MS.point_source_geometry_2d({"type": "point", "x": 0.0, "z": 0.0}, {"y_range": [-1e-6, 1e-6]})
    with pytest.raises(ValueError, match="finite"):
        MS.point_source_geometry_2d({"type": "point", "x": float("nan"), "y": 0.0, "z": 0.0}, {})


def test_save_detector_outputs_stacks_phases(tmp_path):
    outs = [np.full((3, 5), k, dtype=np.float64) for k in range(4)]
    MS.save_detector_outputs(tmp_path / "detected.npy", outs)
    got = np.load(tmp_path / "detected.npy")
    assert got.shape == (4, 3, 5) and got.dtype == np.float64
    assert np.array_equal(got[2], outs[2])
    assert not (tmp_path / "detected.npy.part").exists()
    with pytest.raises(ValueError, match="same shape"):
        MS.save_detector_outputs(tmp_path / "x.npy", [np.zeros((2, 2)), np.zeros((2, 3))])
    with pytest.raises(ValueError, match="no detector outputs"):
        MS.save_detector_outputs(tmp_path / "x.npy", [])


def test_thin_fresnel_rejects_two_elements_before_gpu_work(tmp_path):
    grid = np.zeros((2, 24, 24), np.uint32)
    root, _ = write_prepared(tmp_path, n=16, nz=2, elements=2, fresnel="thin", grid=grid)
    with pytest.raises(ValueError, match="exactly one optical element"):
        MS.run_single_simulation(root, 0)
    assert FW.main([str(root), "-s", "0"]) == 1


def test_history_options_are_validated_before_gpu_work(tmp_path):
    root, _ = write_prepared(tmp_path, n=16, nz=2, grid=np.zeros((2, 24, 24), np.uint32))
    with pytest.raises(ValueError, match="unknown history options"):
        MS.run_single_simulation(root, 0, history_dz=1e-3, history_options={"bogus": 1})
    with pytest.raises(ValueError, match="downsample"):
        MS.run_single_simulation(root, 0, history_dz=1e-3, history_options={"downsample": [0, 1]})


def test_prepared_config_parses_like_the_reference(tmp_path):
    root, _ = write_prepared(tmp_path, n=16, nz=2, fresnel="cb",
                             grid=np.zeros((2, 24, 24), np.uint32))
    dct = rs.config.load(root / "config.yaml")
    p = rs.config.parse_sim_params(dct["sim_params"])
    assert p.use_cone_beam_bpm and p.use_fresnel_scaling and p.nx == 16 and p.ny == 16
    sub = rs.config.load(root / "00000000" / "subconfig.yaml")
    table = rs.config.parse_deltabeta_table(sub["deltabeta_table"])
    assert table[1][0] == rs.Material("B", 2.0) and table[1][1] == pytest.approx(3.3e-5 + 9e-8j)
