"""Post-processing (clustering and final score): the oracle restatement pinned
by hand-built cases, and the library's host clustering against it."""
import numpy as np
import pytest

from oracle import post
from unidock_tools import _lib


def _pose(n, shift):
    rng = np.random.default_rng(42)
    base = rng.uniform(-5, 5, (n, 3)).astype(np.float32)
    return base + np.asarray(shift, dtype=np.float32)


def test_rmsd_bounds_closed_forms():
    a = _pose(10, (0, 0, 0))
    b = a + np.float32([0.6, 0.0, 0.0])
    assert post.rmsd_upper_bound(a, b) == pytest.approx(0.6, rel=1e-6)
    el = np.array([6, 6, 7, 8, 6, 6, 7, 8, 6, 6])
    perm = np.array([4, 1, 6, 7, 0, 5, 2, 3, 9, 8])  # swaps only among equal elements
    assert np.all(el[perm] == el)
    assert post.rmsd_lower_bound(a, a[perm], el) == pytest.approx(0.0, abs=1e-12)
    assert post.rmsd_upper_bound(a, a[perm]) > 1.0
    c = a + np.float32([0.3, -0.2, 0.1])
    assert post.rmsd_lower_bound(a, c, el) <= post.rmsd_upper_bound(a, c) + 1e-6


def test_container_merges_replaces_and_skips():
    xyz = [_pose(8, (0, 0, 0)), _pose(8, (0.2, 0, 0)), _pose(8, (5, 0, 0)), _pose(8, (10, 0, 0)),
           _pose(8, (20, 0, 0)), _pose(8, (30, 0, 0))]
    # 0 and 1 are 0.2 A apart: the better one stays; 2..5 are far from everything
    e = [-5.0, -6.0, -4.0, -3.0, 0.0, -3.5]
    kept = post.cluster(xyz, e, 1.0, 3)
    assert kept == [1, 2, 5]  # 1 replaced 0; 4 (energy exactly 0) skipped; 5 replaced the worst (3)
    kept = post.cluster(xyz, [-5.0, -4.0, -4.0, -3.0, -1.0, -2.0], 1.0, 10)
    assert kept == [0, 2, 3, 5, 4]  # 1 lost to 0; the others sorted by energy


def test_conf_independent_closed_form():
    w_int = 5 * 0.05846 / 0.1 - 1
    for n in (0, 3, 10):
        assert post.conf_independent(-8.0, n, w_int) == pytest.approx(-8.0 / (1 + 0.1 * (w_int + 1) * n / 5))


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_library_clustering_matches_oracle(seed):
    rng = np.random.default_rng(seed)
    nh = int(rng.integers(6, 30))
    centers = [rng.uniform(-6, 6, (nh, 3)) for _ in range(6)]
    xyz, e = [], []
    for t in range(96):
        c = centers[int(rng.integers(0, len(centers)))]
        xyz.append((c + rng.normal(0, rng.choice([0.05, 0.4, 1.2]), c.shape)).astype(np.float32))
        e.append(np.float32(rng.normal(-6, 1.5)))
    e[7] = np.float32(0.0)
    xyz = np.array(xyz)
    e = np.array(e, dtype=np.float32)
    for min_rmsd, size in ((1.0, 9), (0.5, 20), (2.0, 4)):
        ref = post.cluster(list(xyz), list(e), min_rmsd, size)
        got = _lib.cluster_modes(xyz, e, min_rmsd, size)
        assert got == ref, (min_rmsd, size)
