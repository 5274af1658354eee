"""Shared scene builders for the tests (product builders + oracle models)."""
import numpy as np

from owlexastitcher_amd import AmrScene, DualMesh, ExaStitchModel, tf_rasterize

# 5-point piecewise-linear RGBA transfer function (x, r, g, b, a)
TF5 = np.array([
    [0.00, 0.0, 0.0, 0.3, 0.00],
    [0.20, 0.1, 0.4, 0.9, 0.05],
    [0.45, 0.2, 0.9, 0.5, 0.30],
    [0.70, 0.9, 0.8, 0.1, 0.60],
    [1.00, 1.0, 0.2, 0.1, 1.00],
], np.float32)

# same shape with a zero-opacity band, so that whole primitives get culled
TF5_CULL = np.array([
    [0.00, 0.0, 0.0, 0.3, 0.00],
    [0.30, 0.1, 0.4, 0.9, 0.00],
    [0.45, 0.2, 0.9, 0.5, 0.40],
    [0.70, 0.9, 0.8, 0.1, 0.70],
    [1.00, 1.0, 0.2, 0.1, 1.00],
], np.float32)


class Scene:
    def __init__(self, amr):
        self.amr = amr
        self.cells, self.scalars = amr.arrays()
        self.dm = DualMesh(amr)
        self.idx, self.verts, self.gridlets, self.gscalars = self.dm.arrays()
        self.model = ExaStitchModel.from_dual_mesh(self.dm)
        self.cell_bounds, self.value_range = self.model.bounds()


def config1(seed=1):
    """2-level Gaussian scene: coarse 64^3 cells of width 2 over [0,128]^3,
    fine cells of width 1 filling [32,96]^3, 8 Gaussians."""
    return Scene(AmrScene.gaussians(64, 1, 32, 96, 8, seed))


def small(seed=3):
    """16^3 coarse cells of width 2, refined in [8,24]^3."""
    return Scene(AmrScene.gaussians(16, 1, 8, 24, 4, seed))


def noise_small(seed=5):
    """3-level fractal-noise scene on a 16^3 base grid of width 4."""
    return Scene(AmrScene.noise(16, 2, 0.55, 0.7, 6, 2.2, seed))


def octree(seed=1, split=0.4, affine=None):
    """Unbalanced AMR: 4^3 blocks of width-8 cells over [0,32]^3, each split
    recursively (probability `split` per cell, down to width 1) with no 2:1
    balance, so neighbouring leaves differ by up to three levels.  Its dual
    mesh has tetrahedra: dual cells whose top collapses to one vertex and
    whose base loses an edge (makeDual.cpp:512-528).  Scalars are uniform
    random, or the affine field a.xyz + b at the cell centre when
    `affine` = (a, b)."""
    rng = np.random.default_rng(seed)
    cells = []

    def rec(x, y, z, level):
        if level > 0 and rng.random() < split:
            h = 1 << (level - 1)
            for dz in (0, h):
                for dy in (0, h):
                    for dx in (0, h):
                        rec(x + dx, y + dy, z + dz, level - 1)
        else:
            cells.append((x, y, z, level))

    for i in range(4):
        for j in range(4):
            for k in range(4):
                rec(8 * i, 8 * j, 8 * k, 3)
    cells = np.array(cells, np.int32)
    if affine is None:
        scalars = rng.random(len(cells)).astype(np.float32)
    else:
        a, b = affine
        centre = cells[:, :3].astype(np.float64) + 0.5 * (1 << cells[:, 3])[:, None].astype(np.float64)
        scalars = (centre @ np.asarray(a, np.float64) + b).astype(np.float32)
    return Scene(AmrScene.from_cells(cells, scalars))


def tf128(points=TF5):
    return tf_rasterize(points, 128)
