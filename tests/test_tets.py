"""Tetrahedra: intersectTet (Plane.h:57-99) and the 4-vertex branch of the
element dispatch (programs/ExaStitchSampler.cu:133-161), on the CPU oracle and
on the GPU.

The scene is an unbalanced octree AMR (no 2:1 balance): its dual mesh holds
tetrahedra emitted where a dual cell's top collapses to one vertex and its
base loses an edge (makeDual.cpp:512-528), next to pyramids, wedges, hexes
and gridlets.  Tetrahedra come first in the element buffer
(ExaStitchModel.cpp:79-108), so element IDs below the tet count are tets.

Pins, each independent of the code under test:
- the dual-mesh elements tile space without overlap, so a tetrahedron's
  centroid lies in that tetrahedron and in no other primitive: point location
  must return exactly that element;
- barycentric interpolation reproduces an affine field a.x + b exactly up to
  float rounding (vertices carry the field at the AMR cell centres).
GPU parity: IDs and values bit-identical to the oracle (binned search, and a
brute force over every primitive on a subset), then whole frames of the
scene bit for bit in both launch schemes.
"""
import numpy as np
import pytest

from oracle import oracle as ora
from owlexastitcher_amd import ACCEL_BVH, ACCEL_TABLES, RENDER_MEGAKERNEL, RENDER_PERSISTENT, ExaStitchSampler
from tests import scenes

AFFINE = ((0.25, -0.5, 0.125), 3.0)


def _num_tets(sc):
    nv = np.count_nonzero(sc.idx >= 0, axis=1)
    nt = int(np.count_nonzero(nv == 4))
    assert np.all(nv[:nt] == 4)  # tets lead the combined buffer
    return nt


def _centroids(sc, nt):
    v = sc.verts[sc.idx[:nt, :4], :3].astype(np.float64)
    return v.mean(axis=1).astype(np.float32)


@pytest.fixture(scope="module")
def oct1():
    return scenes.octree(seed=1)


@pytest.fixture(scope="module")
def oct_affine():
    return scenes.octree(seed=2, affine=AFFINE)


def test_octree_dual_mesh_has_tets(oct1):
    c = oct1.dm.counts()
    assert c["tets"] > 300 and c["pyramids"] > 0 and c["wedges"] > 0 and c["hexes"] > 0 and c["gridlets"] > 0
    assert _num_tets(oct1) == c["tets"]


def test_oracle_tet_centroids_owned(oct1):
    nt = _num_tets(oct1)
    cent = _centroids(oct1, nt)
    om = ora.Model(oct1.verts, oct1.idx, oct1.gridlets, oct1.gscalars)
    ref, cnt = om.locate_brute(cent)
    assert np.array_equal(ref[:, 0], np.arange(nt)) and np.all(ref[:, 1] == -1)
    assert np.all(cnt == 1)
    assert np.array_equal(om.locate_binned(cent), ref)


def test_oracle_tets_reproduce_affine_field(oct_affine):
    sc = oct_affine
    nt = _num_tets(sc)
    om = ora.Model(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    rng = np.random.default_rng(3)
    b = sc.cell_bounds
    pos = np.concatenate([rng.uniform(b[:3], b[3:], size=(100000, 3)).astype(np.float32), _centroids(sc, nt)])
    out = om.locate_binned(pos)
    val = out[:, 2].copy().view(np.float32).astype(np.float64)
    aff = pos.astype(np.float64) @ np.array(AFFINE[0]) + AFFINE[1]
    tet = (out[:, 0] >= 0) & (out[:, 0] < nt)
    assert tet.sum() > 5000
    scale = np.abs(aff[tet]).max()
    assert np.abs(val[tet] - aff[tet]).max() <= 1e-6 * scale


# ---- GPU ----------------------------------------------------------------------


def _locate_gpu(sampler, pos, device):
    import torch
    d_pos = torch.from_numpy(np.ascontiguousarray(pos, np.float32)).to(device)
    d_out = torch.empty((len(pos), 3), dtype=torch.int32, device=device)
    sampler.locate(d_pos, d_out)
    torch.cuda.synchronize()
    return d_out.cpu().numpy()


@pytest.mark.gpu
@pytest.mark.parametrize("accel", [ACCEL_TABLES, ACCEL_BVH], ids=["tables", "bvh"])
def test_gpu_locate_tets_match_oracle(oct1, device, accel):
    print(f"case: locate octree scene with tets, accel={accel}", flush=True)
    sc = oct1
    nt = _num_tets(sc)
    sc.model.setNumGridCells((16, 16, 16))
    s = ExaStitchSampler()
    s.setAccel(accel)
    s.build(sc.model)
    rng = np.random.default_rng(9)
    b = sc.cell_bounds
    uni = rng.uniform(b[:3] - 1.0, b[3:] + 1.0, size=(1 << 18, 3)).astype(np.float32)
    cent = _centroids(sc, nt)
    pos = np.concatenate([uni, cent])
    got = _locate_gpu(s, pos, device)
    om = ora.Model(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    ref = om.locate_binned(pos)
    nbad = int(np.count_nonzero(np.any(got != ref, axis=1)))
    print(f"rows differing from the oracle: {nbad} of {len(pos)}; tet hits "
          f"{int(np.count_nonzero((got[:, 0] >= 0) & (got[:, 0] < nt)))}")
    assert nbad == 0
    # every centroid in its own tet (the 4-vertex branch answered)
    assert np.array_equal(got[len(uni):, 0], np.arange(nt))
    # brute force over every primitive on a subset of the uniform points
    sub = np.concatenate([uni[:4000], cent])
    ref_b, _ = om.locate_brute(sub)
    assert np.array_equal(_locate_gpu(s, sub, device), ref_b)


@pytest.mark.gpu
def test_gpu_tets_reproduce_affine_field(oct_affine, device):
    print("case: locate affine-field octree scene, tets against a.x + b", flush=True)
    sc = oct_affine
    nt = _num_tets(sc)
    sc.model.setNumGridCells((16, 16, 16))
    s = ExaStitchSampler()
    s.build(sc.model)
    rng = np.random.default_rng(4)
    b = sc.cell_bounds
    pos = np.concatenate([rng.uniform(b[:3], b[3:], size=(1 << 18, 3)).astype(np.float32), _centroids(sc, nt)])
    got = _locate_gpu(s, pos, device)
    val = got[:, 2].copy().view(np.float32).astype(np.float64)
    aff = pos.astype(np.float64) @ np.array(AFFINE[0]) + AFFINE[1]
    tet = (got[:, 0] >= 0) & (got[:, 0] < nt)
    assert tet.sum() > 10000
    assert np.abs(val[tet] - aff[tet]).max() <= 1e-6 * np.abs(aff[tet]).max()
    om = ora.Model(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    assert np.array_equal(got, om.locate_binned(pos))


@pytest.mark.gpu
@pytest.mark.parametrize("scheme", [RENDER_PERSISTENT, RENDER_MEGAKERNEL], ids=["persistent", "megakernel"])
@pytest.mark.parametrize("integrator", [0, 1])
def test_gpu_render_tet_scene_matches_oracle(oct1, device, scheme, integrator):
    from tests.test_gpu_render import _report, make_renderer, render_both
    print(f"case: render octree scene with tets, 96x64, integrator={integrator} scheme={scheme}", flush=True)
    r = make_renderer(oct1, 96, 64, num_mcs=(16, 16, 16), opacity=130.0)
    r.setScheme(scheme)
    r.setType(integrator)
    r.setMaxBounces(4)
    r.setDirectionalLight(0, (0.3, 1.0, 0.5), 2.0, True)
    acc, acc_o, fb, fb_o = render_both(oct1, r, 2, device)
    assert np.array_equal(acc, acc_o), _report(acc, acc_o)
    assert np.array_equal(fb, fb_o)
    r.accumID = 0
    cnt = r.render_counted()
    print(cnt)
    assert cnt["tet"] > 0  # the frame's point location answered in tets
