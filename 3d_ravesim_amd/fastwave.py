"""fast-wave compatible command line: ``fastwave SIM_DIR -s SOURCE_IDX``.

Same arguments, outputs and exit codes as fast-wave's ``main``
(fast-wave/main.cpp:18-53): reads the prepared simulation directory, runs one
source on the GPU and writes ``<SIM_DIR>/<idx:08>/detected.npy`` of shape
(phases, ny_pix, nx_pix) plus ``source_geometry.yaml`` for point sources.
Errors print ``uncaught exception in the fast-wave framework: ...`` and exit
with status 1.  The detector images are float64 (big-wave's dtype; fast-wave
writes its field's scalar type).  ``--history_dz`` records history frames
(history.npy (z, y, x) and history_z.npy, as fast-wave writes them, plus
big-wave's coordinate and metadata files).
"""

from __future__ import annotations

import argparse
import sys


def build_parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(prog="fastwave", description="RAVE-SIM single-source runner (HIP)")
    p.add_argument("sim_dir", help="Main simulation directory")
    p.add_argument("-s", "--source_idx", type=int, default=None,
                   help="The index of the source point to simulate (required)")
    p.add_argument("--history_dz", type=float, default=None,
                   help="Z-Spacing for history entries (optional)")
    return p


def main(argv: list[str] | None = None) -> int:
    parser = build_parser()
    args = parser.parse_args(argv)
    if args.source_idx is None:
        print("Source index is required\n", file=sys.stderr)
        parser.print_help()
        return 1
    try:
        from .multisim import run_single_simulation
        run_single_simulation(args.sim_dir, args.source_idx, history_dz=args.history_dz)
    except Exception as e:  # fast-wave reports every failure the same way
        print(f"uncaught exception in the fast-wave framework: {e}", file=sys.stderr)
        return 1
    return 0


if __name__ == "__main__":
    sys.exit(main())
