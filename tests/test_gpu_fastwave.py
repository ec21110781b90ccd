"""fast-wave compatible runner on the GPU versus the oracle.

A prepared directory (config.yaml, computed.yaml, 00000000/subconfig.yaml,
grid.npy) runs through run_single_simulation and through the command line;
detected.npy must hold (phases, ny_pix, nx_pix) images equal to the oracle's
end-to-end path: complex128 fields 1e-11, complex64 2e-5 (relative L2).
"""

import os
import subprocess
import sys
import importlib

import numpy as np
import pytest
import yaml

import oracle as O
from _prepared import DB, write_prepared
from _rs import ROOT

pytestmark = pytest.mark.gpu
MS = importlib.import_module("3d_ravesim_amd.multisim")


def say(*a):
    print("[case]", *a, flush=True)
    sys.stdout.flush()


def oracle_images(root, grid, fresnel=None):
    cfg = yaml.safe_load(open(root / "config.yaml"))
    sub = yaml.safe_load(open(root / "00000000" / "subconfig.yaml"))
    sp, el = cfg["sim_params"], cfg["elements"][0]
    wl = O.convert_energy_wavelength(sub["energy"])
    angle = yaml.safe_load(open(root / "computed.yaml"))["cutoff_angles"][0]
    cut = np.sin(angle) / wl
    db_list = np.array([0] + [complex(d, b) for _, (d, b) in DB])
    src = sub["source"]
    geom = None
    if fresnel is not None:
        geom = O.FresnelGeometry(z_detector=sp["z_detector"], cone_beam=(fresnel == "cb"))
        geom.configure(src["z"], el["z_start"])
    out = []
    for xp in el["x_positions"]:
        spec = O.SampleSpec(grid=grid, db_list=db_list, pixel_size_x=el["pixel_size_x"],
                            pixel_size_y=el["pixel_size_y"], pixel_size_z=el["pixel_size_z"],
                            x_pos=xp, y_pos=0.0, z_start=el["z_start"])
        out.append(O.run_point_source_sample(
            sp["nx"], sp["ny"], sp["dx"], sp["dy"], wl, (src["x"], src["y"], src["z"]), cut,
            sp["z_detector"], sp["detector_size_x"], sp["detector_size_y"],
            sp["detector_pixel_size_x"], sp["detector_pixel_size_y"], sample=spec,
            geometry=geom))
    return np.stack(out)


@pytest.mark.parametrize("dtype,tol", [("c16", 1e-11), ("c8", 2e-5)])
def test_runner_matches_oracle_with_phase_steps(tmp_path, dtype, tol):
    dx = 0.5e-6
    root, grid = write_prepared(tmp_path, dtype=dtype, x_positions=(0.0, 2 * dx, 5 * dx))
    say("run_single_simulation", dtype, "3 phase steps")
    sub = MS.run_single_simulation(root, 0)
    det = np.load(sub / "detected.npy")
    exp = oracle_images(root, grid)
    assert det.shape == exp.shape == (3, int(256 * dx * 0.8 // (1.7 * dx)),
                                      int(256 * dx * 0.9 // (2.3 * dx)))
    for k in range(3):
        assert O.relative_l2(det[k], exp[k]) < tol, k
    geo = yaml.safe_load(open(sub / "source_geometry.yaml"))
    assert geo == {"type": "point", "units": "m", "x": 0.7e-6, "y": -0.4e-6, "z": 0.0,
                   "y_origin": "explicit"}
    meta = yaml.safe_load(open(sub / "detector_metadata.yaml"))
    assert meta["phase_steps"] == 3 and meta["output_shape"] == list(det.shape[1:])


@pytest.mark.parametrize("fresnel", ["thin", "cb"])
def test_runner_fresnel_modes_match_oracle(tmp_path, fresnel):
    root, grid = write_prepared(tmp_path, dtype="c16", fresnel=fresnel, z_start=0.05,
                                z_detector=0.4,
                                source={"type": "point", "x": 0.3e-6, "y": 0.2e-6, "z": -0.01})
    say("run_single_simulation fresnel", fresnel)
    sub = MS.run_single_simulation(root, 0)
    det = np.load(sub / "detected.npy")
    exp = oracle_images(root, grid, fresnel)
    assert det.shape == exp.shape
    assert O.relative_l2(det[0], exp[0]) < 1e-11
    meta = yaml.safe_load(open(sub / "detector_metadata.yaml"))
    assert meta["effective_geometry"]["fresnel_scaled"]


def test_command_line_writes_the_same_images(tmp_path):
    root, grid = write_prepared(tmp_path, dtype="c8")
    say("fastwave command line (child process)")
    exe = os.path.join(ROOT, "3d_ravesim_amd", "bin", "fastwave")
    r = subprocess.run([sys.executable, exe, str(root), "-s", "0"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    det = np.load(root / "00000000" / "detected.npy")
    exp = oracle_images(root, grid)
    assert det.shape == exp.shape
    assert O.relative_l2(det[0], exp[0]) < 2e-5
