"""Writes ``<fixture>.full.task`` next to the four CP2K task fixtures: the
same task with the expected value of EVERY non-zero grid point.

The reference's dumps hold the x = 0 plane only (its dump loop advances the
wrong index, src/lib/grid_base_ref_c.cc:1160), so the other planes of the
fixtures are pinned nowhere in the reference. The values written here come
from the oracle (oracle/gpw_oracle.c), after checking that it reproduces the
fixture's own dumped values (x = 0 plane, exact count of non-zero points,
1e-13 relative) -- the oracle's other pins are in tests/test_oracle.py.
Run from the repository root: python tests/golden/make_full_tasks.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from libgpw_amd.taskfile import read_task, write_task  # noqa: E402
from oracle import oracle as O  # noqa: E402

NAMES = ["collocate_ortho_density", "collocate_ortho_tau", "collocate_general_density",
         "collocate_general_tau"]


def main():
    O.build()
    for name in NAMES:
        t = read_task(os.path.join(HERE, name + ".task"))
        grid = O.collocate(t, t["pab"])
        n = t["ngrid"]
        plane = np.zeros((n[2], n[1]))
        for z, y, x, v in t["grid_entries"]:
            assert x == 0
            plane[z, y] = v
        got = grid[:, :, 0]
        assert np.count_nonzero(got) == len(t["grid_entries"]), name
        scale = max(np.abs(plane).max(), 1e-300)
        assert np.abs(got - plane).max() <= 1e-13 * scale or not plane.any(), name
        out = os.path.join(HERE, name + ".full.task")
        write_task(out, t, grid)
        print(f"{name}: {np.count_nonzero(grid)} non-zero points -> {os.path.basename(out)}")


if __name__ == "__main__":
    main()
