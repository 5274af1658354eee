"""GPU parity of the ExaStitch sampler kernels against the CPU oracle.

Bar: primitive/cell IDs and majorant-grid values bit-exact; interpolated
values within 1e-6 relative (they come out bit-identical when both sides
follow the reference's operation order).
"""
import numpy as np
import pytest

from oracle import oracle as ora
from owlexastitcher_amd import ACCEL_BVH, ACCEL_TABLES, ExaStitchSampler
from tests import scenes

pytestmark = pytest.mark.gpu

VALUE_RTOL = 1e-6


@pytest.fixture(scope="module")
def cfg1():
    return scenes.config1()


@pytest.fixture(scope="module", params=[ACCEL_TABLES, ACCEL_BVH], ids=["tables", "bvh"])
def sampler1(request, cfg1, device):
    cfg1.model.setNumGridCells((32, 32, 32))
    s = ExaStitchSampler()
    s.setAccel(request.param)
    s.build(cfg1.model)
    assert s.accel == request.param
    return s


def _locate_gpu(sampler, pos, device):
    import torch
    d_pos = torch.from_numpy(np.ascontiguousarray(pos, np.float32)).to(device)
    d_out = torch.empty((len(pos), 3), dtype=torch.int32, device=device)
    sampler.locate(d_pos, d_out)
    torch.cuda.synchronize()
    out = d_out.cpu().numpy()
    return out[:, :2].copy(), out[:, 2].copy().view(np.float32)


def _compare(gpu_ids, gpu_vals, ora_out):
    ora_ids = ora_out[:, :2]
    ora_vals = ora_out[:, 2].copy().view(np.float32)
    id_mismatch = np.count_nonzero(np.any(gpu_ids != ora_ids, axis=1))
    hit = ora_ids[:, 0] >= 0
    rel = np.abs(gpu_vals[hit] - ora_vals[hit]) / np.maximum(np.abs(ora_vals[hit]), 1e-30)
    return id_mismatch, (rel.max() if rel.size else 0.0), int(np.count_nonzero(gpu_vals[hit] != ora_vals[hit]))


def test_locate_1M_uniform_config1(cfg1, sampler1, device):
    """1M uniform points in [0,128]^3 (config 1 check) against the oracle."""
    print("case: locate config1 1M uniform, seed 7", flush=True)
    rng = np.random.default_rng(7)
    pos = rng.uniform(0.0, 128.0, size=(1 << 20, 3)).astype(np.float32)
    ids, vals = _locate_gpu(sampler1, pos, device)
    om = ora.Model(cfg1.verts, cfg1.idx, cfg1.gridlets, cfg1.gscalars)
    ref = om.locate_binned(pos)
    mism, rel, nonbit = _compare(ids, vals, ref)
    print(f"id mismatches {mism}, max rel value err {rel:.3g}, non-bit-identical values {nonbit}")
    assert mism == 0
    assert rel <= VALUE_RTOL
    miss = np.count_nonzero(ref[:, 0] < 0) / len(pos)
    assert 0.03 < miss < 0.07  # outside the dual-mesh extent [1,127]^3


def test_locate_brute_force_subset(cfg1, sampler1, device):
    """Brute force over every primitive on a subset, incl. exact lattice points."""
    rng = np.random.default_rng(11)
    pos = rng.uniform(0.0, 128.0, size=(3000, 3)).astype(np.float32)
    # points exactly on dual-cell faces/corners (k + 0.5 lattice) -- tie cases
    lat = (rng.integers(0, 128, size=(1000, 3)) + 0.5).astype(np.float32)
    pos = np.concatenate([pos, lat])
    ids, vals = _locate_gpu(sampler1, pos, device)
    om = ora.Model(cfg1.verts, cfg1.idx, cfg1.gridlets, cfg1.gscalars)
    ref, cnt = om.locate_brute(pos)
    mism, rel, _ = _compare(ids, vals, ref)
    print(f"brute: id mismatches {mism}, max rel {rel:.3g}, multi-containing {int((cnt > 1).sum())}")
    assert mism == 0 and rel <= VALUE_RTOL


def test_majorant_grid_bit_exact(cfg1, sampler1):
    dims, bounds = sampler1.grid_info()
    assert list(dims) == [32, 32, 32]
    assert np.array_equal(bounds, cfg1.cell_bounds)
    om = ora.Model(cfg1.verts, cfg1.idx, cfg1.gridlets, cfg1.gscalars)
    ref = om.grid_build(dims, bounds)
    got = sampler1.value_ranges()
    assert np.array_equal(got, ref), f"{np.count_nonzero(got != ref)} range entries differ"


def test_gridlet_value_ranges_bit_exact(cfg1, sampler1):
    om = ora.Model(cfg1.verts, cfg1.idx, cfg1.gridlets, cfg1.gscalars)
    ref = om.gridlet_value_ranges()
    got = sampler1.gridlet_value_ranges(len(cfg1.gridlets))
    assert np.array_equal(got, ref)


@pytest.mark.parametrize("tf", ["TF5", "TF5_CULL"])
def test_max_opacities_bit_exact(cfg1, sampler1, tf):
    rgba = scenes.tf128(getattr(scenes, tf))
    xr = cfg1.value_range
    sampler1.computeMaxOpacities(rgba, xr)
    om = ora.Model(cfg1.verts, cfg1.idx, cfg1.gridlets, cfg1.gscalars)
    dims, bounds = sampler1.grid_info()
    ref_mc = ora.map_max_opacity(om.grid_build(dims, bounds), rgba, xr)
    assert np.array_equal(sampler1.max_opacities(), ref_mc)
    g, e = sampler1.prim_max_opacities(len(cfg1.gridlets), len(cfg1.idx))
    assert np.array_equal(g, ora.map_max_opacity(om.gridlet_value_ranges(), rgba, xr))
    assert np.array_equal(e, om.element_max_opacity(rgba, xr))


def test_locate_with_culling(cfg1, sampler1, device):
    """Primitives with max opacity 0 are skipped (ExaStitchSampler.cu:35-37,81-83)."""
    rgba = scenes.tf128(scenes.TF5_CULL)
    xr = cfg1.value_range
    sampler1.computeMaxOpacities(rgba, xr)
    om = ora.Model(cfg1.verts, cfg1.idx, cfg1.gridlets, cfg1.gscalars)
    gmo = ora.map_max_opacity(om.gridlet_value_ranges(), rgba, xr)
    emo = om.element_max_opacity(rgba, xr)
    assert (gmo == 0).any() and (emo == 0).any(), "transfer function must cull something"
    om.set_culling(gmo, emo)
    rng = np.random.default_rng(13)
    pos = rng.uniform(0.0, 128.0, size=(1 << 18, 3)).astype(np.float32)
    ids, vals = _locate_gpu(sampler1, pos, device)
    ref = om.locate_binned(pos)
    mism, rel, _ = _compare(ids, vals, ref)
    assert mism == 0 and rel <= VALUE_RTOL


def test_test_sample(cfg1, sampler1, device):
    import torch
    rng = np.random.default_rng(17)
    n = 20000
    prim = rng.integers(-1, len(cfg1.gridlets), size=n).astype(np.int32)
    # points near the chosen gridlet so that about half hit
    g = cfg1.gridlets[np.maximum(prim, 0)]
    w = (1 << g[:, 3]).astype(np.float32)
    lo = (g[:, 0:3] + 0.5) * w[:, None]
    hi = (g[:, 0:3] + g[:, 4:7] + 0.5) * w[:, None]
    pos = (lo + (hi - lo) * rng.uniform(-0.2, 1.2, size=(n, 3))).astype(np.float32)
    d_pos = torch.from_numpy(pos).to(device)
    d_prim = torch.from_numpy(prim).to(device)
    d_out = torch.empty((n, 3), dtype=torch.int32, device=device)
    ExaStitchSampler.testSample(sampler1, d_pos, d_prim, d_out)
    out = d_out.cpu().numpy()
    sc = cfg1.gscalars
    for i in range(0, n, 7):
        v = np.zeros(1, np.float32)
        cid = np.zeros(1, np.int32)
        if prim[i] < 0:
            assert out[i, 0] == -1
            continue
        gg = np.ascontiguousarray(cfg1.gridlets[prim[i]])
        hit = ora.lib().ora_intersect_gridlet_f32(ora.ptr(v), ora.ptr(cid), ora.ptr(pos[i]), ora.ptr(gg),
                                                  ora.ptr(sc))
        if hit:
            assert out[i, 0] == prim[i] and out[i, 1] == cid[0]
            assert out[i, 2:3].view(np.float32)[0] == v[0]
        else:
            assert out[i, 0] == -1 and out[i, 1] == -1


@pytest.fixture(scope="module")
def noise_scene():
    return scenes.noise_small()


@pytest.mark.parametrize("accel", [ACCEL_TABLES, ACCEL_BVH], ids=["tables", "bvh"])
def test_locate_noise_scene(noise_scene, device, accel):
    """3-level scene (pyramids, wedges, hexes at two level transitions)."""
    sc = noise_scene
    sc.model.setNumGridCells((16, 16, 16))
    s = ExaStitchSampler()
    s.setAccel(accel)
    s.build(sc.model)
    lo, hi = sc.cell_bounds[:3], sc.cell_bounds[3:]
    rng = np.random.default_rng(23)
    pos = rng.uniform(lo - 1.0, hi + 1.0, size=(1 << 18, 3)).astype(np.float32)
    ids, vals = _locate_gpu(s, pos, device)
    om = ora.Model(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    ref = om.locate_binned(pos)
    mism, rel, _ = _compare(ids, vals, ref)
    assert mism == 0 and rel <= VALUE_RTOL
    # element hits of every kind present
    kinds = set(np.unique(np.count_nonzero(sc.idx[ref[ref[:, 1] < 0][:, 0][ref[ref[:, 1] < 0][:, 0] >= 0]] >= 0,
                                           axis=1)))
    assert {5, 6, 8} <= kinds


@pytest.mark.parametrize("accel", [ACCEL_TABLES, ACCEL_BVH], ids=["tables", "bvh"])
def test_elements_only_model(cfg1, device, accel):
    """A model with stitching elements and no gridlets (the gridlet stage has
    nothing to read): point location against brute force, and a rendered
    frame of both render schemes against the oracle."""
    import torch
    from owlexastitcher_amd import ExaStitchModel, OWLRenderer, RENDER_MEGAKERNEL, RENDER_PERSISTENT, camera_look_at
    print(f"case: elements-only model, accel {accel}", flush=True)
    model = ExaStitchModel.from_arrays(cfg1.verts, cfg1.idx, np.zeros((0, 8), np.int32), np.zeros(0, np.float32))
    model.setNumGridCells((16, 16, 16))
    s = ExaStitchSampler()
    s.setAccel(accel)
    s.build(model)
    rng = np.random.default_rng(5)
    cb, _ = model.bounds()
    pos = rng.uniform(cb[:3], cb[3:], (20000, 3)).astype(np.float32)
    ids, vals = _locate_gpu(s, pos, device)
    om = ora.Model(cfg1.verts, cfg1.idx, np.zeros((0, 8), np.int32), np.zeros(0, np.float32))
    ref, _ = om.locate_brute(pos)
    mism, rel, _ = _compare(ids, vals, ref)
    assert mism == 0 and rel <= VALUE_RTOL
    assert np.count_nonzero(ids[:, 0] >= 0) > 100 and np.all(ids[ids[:, 0] >= 0, 1] == -1)
    W, H = 40, 30
    accs = []
    for scheme in (RENDER_PERSISTENT, RENDER_MEGAKERNEL):
        r = OWLRenderer(model, (16, 16, 16))
        r.setAccel(accel)
        r.setScheme(scheme)
        b = r.modelBounds
        c = 0.5 * (b[:3] + b[3:])
        r.resize(W, H)
        r.setCamera(*camera_look_at(c + np.array([-0.3, 0.7, 1.0]) * (b[3:] - b[:3]), c, (0, 1, 0), 60.0, W, H))
        rgba = scenes.tf128()
        r.setColorMap(rgba)
        vr = r.valueRange
        r.setRange(vr[0], vr[1])
        r.setOpacityScale(120.0)
        r.setMaxBounces(4)
        st = r.state()
        fb = torch.zeros((H, W), dtype=torch.int32, device=device)
        r.render(fb)
        torch.cuda.synchronize()
        accs.append(r.accum())
    maj = r.sampler.max_opacities()
    om.set_culling(None, om.element_max_opacity(rgba, np.array(st.tfDomain, np.float32)))
    acc_o, _ = ora.render_f32(ora.render_params_from_state(st, maj, rgba), om)
    assert np.array_equal(accs[0], acc_o) and np.array_equal(accs[1], acc_o)


@pytest.mark.parametrize("accel", [ACCEL_TABLES, ACCEL_BVH], ids=["tables", "bvh"])
def test_locate_on_element_box_faces(noise_scene, device, accel):
    """Points on and one float step either side of stitching-element box
    faces, edges and corners (the element-bin lists store each box quantised
    to sub-cells of the bin, so these are the points where only the exact box
    decides) against the oracle, and against brute force on a subset."""
    print(f"case: locate on element box faces, noise scene, accel {accel}", flush=True)
    sc = noise_scene
    sc.model.setNumGridCells((16, 16, 16))
    s = ExaStitchSampler()
    s.setAccel(accel)
    s.build(sc.model)
    rng = np.random.default_rng(31)
    ne = len(sc.idx)
    e = rng.integers(0, ne, size=40000)
    v = sc.verts[:, :3]
    nv = np.count_nonzero(sc.idx >= 0, axis=1)
    corners = np.where((np.arange(8)[None, :] < nv[e, None])[..., None], v[np.maximum(sc.idx[e], 0)], np.nan)
    lo, hi = np.nanmin(corners, axis=1), np.nanmax(corners, axis=1)
    pos = rng.uniform(lo, hi).astype(np.float32)
    # snap one to three axes onto a face of the box, then nudge some of
    # those one float step inward or outward
    snap = rng.random((len(e), 3)) < 0.5
    side = rng.random((len(e), 3)) < 0.5
    face = np.where(side, hi, lo).astype(np.float32)
    pos = np.where(snap, face, pos)
    step = rng.integers(-1, 2, size=(len(e), 3))
    out = np.nextafter(pos, np.where(step > 0, np.inf, -np.inf).astype(np.float32))
    pos = np.where(snap & (step != 0), out, pos).astype(np.float32)
    ids, vals = _locate_gpu(s, pos, device)
    om = ora.Model(sc.verts, sc.idx, sc.gridlets, sc.gscalars)
    ref = om.locate_binned(pos)
    mism, rel, _ = _compare(ids, vals, ref)
    el_hits = int(np.count_nonzero((ref[:, 0] >= 0) & (ref[:, 1] < 0)))
    print(f"id mismatches {mism}, max rel {rel:.3g}, element hits {el_hits} of {len(pos)}")
    assert mism == 0 and rel <= VALUE_RTOL
    assert el_hits > len(pos) // 4
    ref_b, _ = om.locate_brute(pos[:3000])
    assert np.array_equal(ids[:3000], ref_b[:, :2])
