"""Shock-shell AMR generator: the leaves tile the domain, the level histogram
of the counting pass matches the built scene, the same seed gives the same
scene, and touching leaves differ by at most one level (checked directly on
a dense voxel grid over all 26 neighbour directions, and again through the
dual-mesh builder, which reports a conflict for any unbalanced vertex)."""
import numpy as np

from owlexastitcher_amd import AmrScene, DualMesh

ARGS = (24, 4, 80.0, 30.0, 0.7, 0.5, 1)


def dense_levels(cells, n):
    lev = np.full((n, n, n), -1, np.int8)
    for L in np.unique(cells[:, 3]):
        c = cells[cells[:, 3] == L]
        m = np.zeros((n >> L,) * 3, bool)
        m[c[:, 0] >> L, c[:, 1] >> L, c[:, 2] >> L] = True
        for ax in range(3):
            m = np.repeat(m, 1 << L, axis=ax)
        assert not np.any(lev[m] >= 0)  # leaves do not overlap
        lev[m] = L
    return lev


def test_shock_scene_tiles_and_is_balanced():
    a = AmrScene.shock(*ARGS)
    cells, scalars = a.arrays()
    n = ARGS[0] << ARGS[1]
    hist = np.bincount(cells[:, 3], minlength=ARGS[1] + 1)
    assert np.all(hist > 0)  # every level 0..4 present
    assert np.array_equal(hist, AmrScene.shock_counts(*ARGS))
    assert np.all(np.isfinite(scalars)) and scalars.min() >= 0.0 and scalars.max() <= 1.0
    assert np.all(cells[:, :3] % (1 << cells[:, 3:4]) == 0)  # aligned to their own width
    lev = dense_levels(cells, n)
    assert np.all(lev >= 0)  # no holes
    worst = 0
    for d in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (1, -1, 0), (1, 0, 1), (1, 0, -1), (0, 1, 1),
              (0, 1, -1), (1, 1, 1), (1, 1, -1), (1, -1, 1), (1, -1, -1)]:
        sl_a = tuple(slice(max(0, -k), n - max(0, k)) for k in d)
        sl_b = tuple(slice(max(0, k), n - max(0, -k)) for k in d)
        worst = max(worst, int(np.abs(lev[sl_a].astype(np.int16) - lev[sl_b]).max()))
    assert worst == 1
    c = DualMesh(a).counts()
    assert c["conflicts"] == 0 and c["degenerate"] == 0 and c["elements"] > 0


def test_shock_scene_is_deterministic():
    args = (8, 4, 40.0, 20.0, 0.6, 0.45, 3)
    c1, s1 = AmrScene.shock(*args).arrays()
    c2, s2 = AmrScene.shock(*args).arrays()
    assert np.array_equal(c1, c2) and np.array_equal(s1, s2)
    c3, _ = AmrScene.shock(*args[:-1], 4).arrays()
    assert len(c1) != len(c3) or not np.array_equal(c1, c3)  # the seed matters
