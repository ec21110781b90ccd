"""Write small prepared simulation directories (the layout big-wave's
multisim.setup_simulation produces and fast-wave's runner reads:
config.yaml, computed.yaml, <idx:08>/subconfig.yaml, the sample grid .npy)."""

from pathlib import Path

import numpy as np
import yaml

DB = [(["A", 1.0], [3.5e-7 * 30, 2e-10 * 30]), (["B", 2.0], [1.1e-6 * 30, 3e-9 * 30]),
      (["C", 1.5], [2.6e-7 * 30, 1.5e-10 * 30])]


def spheres(rng, nz, ny, nx, k):
    g = np.zeros((nz, ny, nx), dtype=np.uint32)
    zz, yy, xx = np.meshgrid(np.arange(nz) + 0.5, np.arange(ny) + 0.5, np.arange(nx) + 0.5,
                             indexing="ij")
    for _ in range(k):
        c = rng.uniform([0, 0, 0], [nz, ny, nx])
        r = rng.uniform(2, max(3, min(ny, nx) / 6))
        g[(zz - c[0]) ** 2 + (yy - c[1]) ** 2 + (xx - c[2]) ** 2 <= r * r] = rng.integers(1, 4)
    return g


def write_prepared(root: Path, *, n=256, nz=6, dx=0.5e-6, dtype="c16", energy=8000.0,
                   source=None, x_positions=(0.0,), z_start=0.6, z_detector=1.0, angle=None,
                   elements=1, fresnel=None, seed=5, grid=None):
    """Returns (sim_dir, grid).  ``fresnel``: None, "thin" or "cb"."""
    root = Path(root)
    root.mkdir(parents=True, exist_ok=True)
    rng = np.random.default_rng(seed)
    if grid is None:
        grid = spheres(rng, nz, n + 8, n + 8, 20)
    np.save(root / "grid.npy", grid)
    sim = {"N": n * n, "nx": n, "ny": n, "dx": dx, "dy": dx, "z_detector": z_detector,
           "detector_size": n * dx * 0.9, "detector_size_x": n * dx * 0.9,
           "detector_size_y": n * dx * 0.8, "detector_pixel_size_x": 2.3 * dx,
           "detector_pixel_size_y": 1.7 * dx, "chunk_size": n * 16}
    if fresnel == "thin":
        sim["use_fresnel_scaling"] = True
    elif fresnel == "cb":
        sim["use_cone_beam_bpm"] = "true"
    els = []
    for k in range(elements):
        els.append({"type": "sample", "z_start": z_start + k * 0.01, "pixel_size_x": dx,
                    "pixel_size_y": dx, "pixel_size_z": 0.5e-6, "grid_path": "grid.npy",
                    "materials": [["A", 1.0], ["B", 2.0], ["C", 1.5]],
                    "x_positions": [float(x) for x in x_positions]})
    cfg = {"dtype": dtype, "use_disk_vector": False, "save_final_u_vectors": False,
           "sim_params": sim,
           "multisource": {"type": "points", "nr_source_points": 1,
                           "energy_range": [energy, energy], "x_range": [0.0, 0.0],
                           "y_range": [0.0, 0.0], "z": 0.0, "seed": 1},
           "elements": els}
    with open(root / "config.yaml", "w") as f:
        yaml.safe_dump(cfg, f)
    wl = 6.62607004e-34 * 299792458 / (energy * 1.602176634e-19)
    a = angle if angle is not None else float(np.arcsin(0.4 / dx * wl))
    with open(root / "computed.yaml", "w") as f:
        yaml.safe_dump({"cutoff_angles": [a] * (elements + 1)}, f)
    sub = root / "00000000"
    sub.mkdir(exist_ok=True)
    src = source or {"type": "point", "x": 0.7e-6, "y": -0.4e-6, "z": 0.0}
    with open(sub / "subconfig.yaml", "w") as f:
        yaml.safe_dump({"energy": energy, "source": src,
                        "deltabeta_table": [[m, db] for m, db in DB]}, f)
    return root, grid
