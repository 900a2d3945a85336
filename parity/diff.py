"""Parity harness: compare two output folders (a reference run and a run of
this library) frame by frame.

    python parity/diff.py REF_DIR OURS_DIR [--frames 0 5] [--tol position=1e-9 ...] [--exact]

Snapshot rules (cells and microenvironment .mat files): those of the
reference's parity harness (parity/diff.py:40-53,116-147,150-300 in the
reference) -- cells matched by ID, position / volume / phase tolerances,
voxel-centre gate, substrate L-infinity and RMSE.  --exact adds the integer
outputs of the step path, written beside the snapshots: cell->voxel keys,
mechanics keys, bin offsets, the hash-sorted cell list and the neighbour
lists, which must be equal value for value (neighbour lists as sorted sets
per cell and in visiting order).  Exit status 0 when every frame passes.
The rules live in physicell_amd.snapdiff; this file is its command line.
"""
import os
import sys

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "physicell-mac-metal_amd"))

from physicell_amd import snapdiff  # noqa: E402

if __name__ == "__main__":
    sys.exit(snapdiff.main())
