"""Copy the reference's P5 end-to-end fingerprints into fixtures.

The reference keeps the results of its P5 convergence runs as JSON data
(tests/big_wave_2d/p5_results/{vacuum,sample}_convergence_<n>.json, inputs
in tests/big_wave_2d/p5_benchmark.py:69-185).  This script reads those data
files from a checkout of the reference and writes, per grid size, the
fields the GPU tests compare against (detector shape, centre line, energy,
peak) to p5_convergence_<n>.json beside it.  Run in a container that has
the reference checkout:

    python tests/golden/extract_p5_fingerprints.py /path/to/reference
"""

import json
import os
import sys

KEYS = ("mode", "nx", "ny", "dx", "dtype", "cutoff_frequency", "detected_energy",
        "detected_peak", "detected_shape", "centerline_256", "centerline_start_pixel",
        "simulation_fov", "integrator")


def extract(ref_root: str, n: int) -> dict:
    out = {}
    for mode in ("vacuum", "sample"):
        path = os.path.join(ref_root, "tests", "big_wave_2d", "p5_results",
                            f"{mode}_convergence_{n}.json")
        with open(path) as f:
            data = json.load(f)
        entry = {k: data[k] for k in KEYS if k in data}
        meta = data.get("detector_metadata", {})
        if "coverage_area" in meta:
            entry["detector_metadata_coverage_area"] = meta["coverage_area"]
        out[mode] = entry
    out["_source"] = (f"tests/big_wave_2d/p5_results/{{vacuum,sample}}_convergence_{n}.json of the "
                      "reference; inputs per tests/big_wave_2d/p5_benchmark.py:69-185")
    return out


def main() -> None:
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    here = os.path.dirname(os.path.abspath(__file__))
    for n in (4096, 8192, 16384):
        with open(os.path.join(here, f"p5_convergence_{n}.json"), "w") as f:
            json.dump(extract(ref, n), f, indent=1)
            f.write("\n")


if __name__ == "__main__":
    main()
