"""The tile window's node-to-column map (TileWindow::col_of in
csrc/mpm_common.h) divides by the window height with a rounded float
reciprocal: col = trunc((i + 1/2) * fl(1/h)).  The claim that this equals
i // h for every node index of a window that fits a workgroup's 64 KiB of LDS
(at least 56 B per node, so i < 1171 and h < 391) is checked here
exhaustively in IEEE float32, as the device computes it."""
import numpy as np


def test_float_reciprocal_column_map_is_exact_for_all_windows():
    i = np.arange(1171, dtype=np.int64)
    fi = i.astype(np.float32) + np.float32(0.5)          # exact: i < 2^24
    for h in range(1, 391):
        inv = np.float32(1.0) / np.float32(h)            # correctly rounded
        col = (fi * inv).astype(np.int64)                # float32 product, truncation
        assert np.array_equal(col, i // h), h


def test_float_reciprocal_column_map_fails_without_the_half():
    # the half-node offset is what keeps (i/h) away from integers: without it
    # a rounded-down quotient at an exact multiple would lose a column
    i = np.arange(1171, dtype=np.int64)
    bad = 0
    for h in range(1, 391):
        inv = np.float32(1.0) / np.float32(h)
        col = (i.astype(np.float32) * inv).astype(np.int64)
        bad += int((col != i // h).sum())
    assert bad > 0
