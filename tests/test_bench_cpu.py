"""Host-side figures of bench.py (no GPU): the candidate count of the cell
walk that the VALU-per-candidate figure divides by, against a brute-force
count with the oracle's HashCell3D (pinned by its known answers)."""
import numpy as np
import pytest

import bench
from harness import ORC


def _brute_candidates(x, cell, n):
    c = np.floor(x / np.float32(cell)).astype(np.int64)
    key = lambda a, b, d: ORC.orc_hash_cell(int(a), int(b), int(d)) % n
    counts = np.bincount([key(*cc) for cc in c], minlength=n)
    total = 0
    for cx, cy, cz in c:
        for dz in (-1, 0, 1):
            for dy in (-1, 0, 1):
                for dx in (-1, 0, 1):
                    total += int(counts[key(cx + dx, cy + dy, cz + dz)])
    return total / len(x)


@pytest.mark.parametrize("n, lo, hi, cell", [
    (600, 0.0, 0.5, 0.05),     # one 50-block of cells
    (600, -2.0, 3.0, 0.05),    # cells below -25 (wrapping) and across 50-blocks
    (37, -0.2, 0.2, 0.04),     # tiny table: stencil cells share keys (buckets visited twice)
])
def test_candidates_per_particle_matches_brute_force(n, lo, hi, cell):
    rng = np.random.default_rng(n)
    x = rng.uniform(lo, hi, size=(n, 3)).astype(np.float32)
    assert bench.candidates_per_particle(x, cell, n) == pytest.approx(_brute_candidates(x, cell, n), rel=0, abs=0)
