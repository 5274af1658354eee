"""Oracle checks of the ray marcher and the shade modes against what geometry
fixes, on a homogeneous single-level volume (every point of the box lies in a
gridlet, every sample has the same value):

- ray marcher (deviceCode.cu:1213-1260, 348-398): n samples of alpha a and
  colour c composite to alpha 1 - (1-a)^n, colour c * alpha, over the
  background; n is the number of lattice points dt * (ils + k) on the chord,
  floor(L/dt) or one more;
- Gridlets shade mode (deviceCode.cu:637-656): with the direct-lighting
  integrator, no lights and a huge extinction, a pixel shows the albedo at
  the ray's entry point, i.e. the checkerboard colour of the dual cell there
  (dark blue when the cell's x and y parities agree and z is even, or
  disagree and z is odd; white otherwise), checked where a pixel's whole
  jitter square enters one cell through a lower face;
- Teaser (657-695): equal to the default image where A u + B v <= C1 and to
  the Gridlets image where it is >= C2, in between a blend of the two.
"""
import numpy as np
import pytest

from oracle import oracle as ora

N, LEVEL = 8, 3          # 8^3 cells of width 8 over [0, 64]^3
W_IMG, H_IMG = 48, 40
EYE = np.array([-30.0, -25.0, -45.0])  # sees the three lower faces of the box


def scene():
    w = 1 << LEVEL
    g = np.stack(np.meshgrid(*(np.arange(N) * w,) * 3, indexing="ij"), -1).reshape(-1, 3)
    cells = np.concatenate([g, np.full((len(g), 1), LEVEL)], 1).astype(np.int32)
    counts, idx, verts, gl, gs = ora.dual_mesh(cells, np.full(len(cells), 0.5, np.float32))
    assert counts["elements"] == 0 and counts["gridlets"] > 0
    return ora.Model(verts, idx, gl, gs)


def camera(bounds):
    center = 0.5 * (bounds[:3] + bounds[3:])
    fwd = center - EYE
    fwd /= np.linalg.norm(fwd)
    right = np.cross(fwd, [0.0, 1.0, 0.0])
    right /= np.linalg.norm(right)
    up = np.cross(right, fwd)
    h = 2 * np.tan(np.radians(17.0))
    du = right * h * (W_IMG / H_IMG) / W_IMG
    dv = -up * h / H_IMG
    d00 = fwd - 0.5 * W_IMG * du - 0.5 * H_IMG * dv
    return d00.astype(np.float32), du.astype(np.float32), dv.astype(np.float32)


def params(om, rgba, integrator, shade=0, dt=1.0, os_=1.0):
    cb, _ = om.bounds()
    p = ora.OraRenderParams()
    p.W, p.H, p.spp, p.accumID = W_IMG, H_IMG, 1, 0
    p.integrator, p.maxBounces, p.shadeMode, p.dt = integrator, 4, shade, dt
    d00, du, dv = camera(cb.astype(np.float64))
    for name, v in (("camOrg", EYE.astype(np.float32)), ("camDir00", d00), ("camDu", du), ("camDv", dv)):
        for k in range(3):
            getattr(p, name)[k] = float(v[k])
    ident = [1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0]
    for k in range(12):
        p.voxelXfm[k] = p.voxelXfmInv[k] = p.lightXfm[k] = ident[k]
    dims = np.array([4, 4, 4], np.int32)
    for k in range(6):
        p.worldBounds[k] = p.gridBounds[k] = float(cb[k])
    for k in range(3):
        p.gridDims[k] = int(dims[k])
    xr = np.array([0.0, 1.0], np.float32)
    maj = ora.map_max_opacity(om.grid_build(dims, cb), rgba, xr)
    om.set_culling(ora.map_max_opacity(om.gridlet_value_ranges(), rgba, xr), None)
    p.tfN = len(rgba)
    p.tfDomain[0], p.tfDomain[1] = 0.0, 1.0
    p.opacityScale = os_
    p.debugCrosshair = 0
    p._keep = (maj, rgba)
    p.majorants = ora.ptr(maj).value
    p.tfRGBA = ora.ptr(rgba).value
    return p


def render(om, p):
    acc, _ = ora.render_f32(p, om)
    return acc.astype(np.float64)


def pixel_rays(p, sx, sy):
    d = np.array(p.camDir00)[None, None] + sx[..., None] * np.array(p.camDu) + sy[..., None] * np.array(p.camDv)
    return d / np.linalg.norm(d, axis=-1, keepdims=True)


def slab(org, d, lo, hi):
    t0 = (lo - org) / d
    t1 = (hi - org) / d
    tn = np.max(np.minimum(t0, t1), axis=-1)
    tf = np.min(np.maximum(t0, t1), axis=-1)
    return np.maximum(tn, 0.0), tf


def footprint(p, f):
    """f evaluated at the 4 corners and the centre of every pixel's jitter square."""
    ys, xs = np.mgrid[0:H_IMG, 0:W_IMG].astype(np.float64)
    return [f(xs + ox, ys + oy) for ox, oy in ((0, 0), (1, 0), (0, 1), (1, 1), (0.5, 0.5))]


def background(H):
    t = np.arange(H)[:, None] / H
    return np.stack([(1 - t) + 0.5 * t, (1 - t) + 0.7 * t, np.ones_like(t)], -1) * np.ones((1, W_IMG, 1))


def const_tf(rgb, alpha, n=64):
    out = np.zeros((n, 4), np.float32)
    out[:, :3] = rgb
    out[:, 3] = alpha
    return out


@pytest.mark.parametrize("dt", [1.0, 0.6])
def test_ray_marcher_compositing_closed_form(dt):
    om = scene()
    a, c = 0.02, np.array([0.2, 0.5, 0.8])
    p = params(om, const_tf(c, a), 2, dt=dt)
    acc = render(om, p)
    cb, _ = om.bounds()
    lo, hi = cb[:3].astype(np.float64), cb[3:].astype(np.float64)
    bg = background(H_IMG)
    checked = 0
    for y in range(H_IMG):
        for x in range(W_IMG):
            # the pixel's jitter and lattice offset: the first three numbers
            # of its LCG stream (pinned by the LCG golden-vector tests)
            rx, ry, ils = ora.lcg_stream(x + W_IMG * y, 0, 3).astype(np.float64)
            d = pixel_rays(p, np.array([[x + rx]]), np.array([[y + ry]]))
            t0, t1 = (v[0, 0] for v in slab(EYE, d, lo, hi))
            if t1 <= t0:
                assert np.allclose(acc[y, x, :3], bg[y, x]), (y, x)  # missed rays show the background
                continue
            k0 = np.ceil(t0 / dt - ils)
            n = max(0, int(np.floor(t1 / dt - ils) - k0 + 1))  # lattice points dt (ils + k) in [t0, t1]
            u0, u1 = t0 / dt - ils, t1 / dt - ils
            edge = min(abs(u0 - round(u0)), abs(u1 - round(u1))) < 1e-3
            ns = np.arange(max(0, n - 1), n + 2) if edge else np.array([n])  # float rounding at either end
            alpha = 1 - (1 - a) ** ns
            rgb = c[None] * alpha[:, None] + (1 - alpha[:, None]) * bg[y, x][None]
            err = np.abs(rgb - acc[y, x, :3][None]).max(axis=1)
            assert err.min() < 2e-5, (y, x, ns, acc[y, x])
            assert abs(acc[y, x, 3] - 1.0) < 1e-6
            checked += 1
    assert checked > 0.3 * W_IMG * H_IMG


def checker_colour(pts, lo):
    k = np.floor((pts - lo) / (1 << LEVEL)).astype(np.int64)
    k = np.clip(k, 0, N - 2)
    col = (k[..., 0] % 2) == (k[..., 1] % 2)
    col = np.where(k[..., 2] % 2 == 0, col, ~col)
    return np.where(col[..., None], [0.0, 0.0, 0.7], [1.0, 1.0, 1.0])


def test_gridlets_shade_mode_is_the_entry_cell_checkerboard():
    om = scene()
    p = params(om, const_tf([0.3, 0.6, 0.9], 1.0), 1, shade=1, os_=1e4)
    acc = render(om, p)
    cb, _ = om.bounds()
    lo, hi = cb[:3].astype(np.float64), cb[3:].astype(np.float64)

    def entry_cell(sx, sy):
        """(entry face, dual-cell index) of the ray through (sx, sy); a ray
        that misses the box gets -1s"""
        d = pixel_rays(p, sx, sy)
        tn, tf = slab(EYE, d, lo, hi)
        pts = EYE + d * tn[..., None]
        face = np.argmin(np.abs(np.concatenate([pts - lo, pts - hi], -1)), axis=-1)
        k = np.clip(np.floor((pts + 1e-3 * d - lo) / (1 << LEVEL)).astype(np.int64), 0, N - 2)
        key = np.concatenate([face[..., None], k], -1)
        return np.where((tf > tn + 1e-2)[..., None], key, -1)

    keys = footprint(p, entry_cell)
    # every corner of the jitter square enters through the same face into the
    # same dual cell: then the whole square does (the face is planar, the
    # cell convex)
    # (lower faces only: a sample within float32 rounding of an upper face
    # gets the cell index dims there, as in the reference)
    agree = np.all(np.stack([np.all(k == keys[0], axis=-1) for k in keys]), axis=0)
    agree &= (keys[0][..., 0] >= 0) & (keys[0][..., 0] < 3)
    assert agree.sum() > 0.3 * agree.size
    want = checker_colour(lo + (keys[4][..., 1:] + 0.5) * (1 << LEVEL), lo)
    assert np.allclose(acc[agree][:, :3], want[agree], rtol=0, atol=1e-6)
    cols = [want]
    whites = np.all(cols[0][agree] == 1.0, axis=-1).sum()
    assert 0 < whites < agree.sum()  # both colours appear


def test_teaser_blends_default_and_gridlets():
    om = scene()
    tf = const_tf([0.3, 0.6, 0.9], 1.0)
    img = {m: render(om, params(om, tf, 1, shade=m, os_=1e4)) for m in (0, 1, 2)}
    u = (np.arange(W_IMG, dtype=np.float32) / np.float32(W_IMG))[None, :]
    v = (np.arange(H_IMG, dtype=np.float32) / np.float32(H_IMG))[:, None]
    A, B = np.float32(0.6) - np.float32(0.4), np.float32(0.4) - np.float32(0.6)
    C1 = A * np.float32(0.4) + B * np.float32(0.6)
    C2 = A * np.float32(0.6) + B * np.float32(0.4)
    Cuv = A * u + B * v
    left, right = Cuv <= C1, Cuv >= C2
    assert left.any() and right.any() and (~left & ~right).any()
    assert np.array_equal(img[2][left], img[0][left])
    assert np.array_equal(img[2][right], img[1][right])
    mid = ~left & ~right
    lo = np.minimum(img[0], img[1])[mid]
    hi = np.maximum(img[0], img[1])[mid]
    assert np.all(img[2][mid] >= lo - 1e-6) and np.all(img[2][mid] <= hi + 1e-6)
    assert np.array_equal(img[2][..., 3], img[0][..., 3])


def test_sub_image_window_and_selection_box():
    """renderFrame_Impl's sub-image (deviceCode.cu:1285-1296): pixel (x, y)
    with jitter (rx, ry) traces the ray through image position
    ((1-u) lo + u hi) * W with u = (x + rx) / W (likewise y): with direct
    lighting, no lights and a huge extinction a pixel shows the TF colour
    where that ray crosses the box and the background where it misses.  The
    selection box (1337-1340) inverts exactly the frame-buffer pixels on its
    border."""
    om = scene()
    col = np.array([0.3, 0.6, 0.9], np.float32)
    p = params(om, const_tf(col, 1.0), 1, os_=1e4)
    p.subImageActive = 1
    lo, hi = (-0.2, -0.1), (0.45, 0.6)  # a window over the box's upper-left silhouette
    p.subImage[0], p.subImage[1], p.subImage[2], p.subImage[3] = lo[0], lo[1], hi[0], hi[1]
    acc, fb = ora.render_f32(p, om)
    cb, _ = om.bounds()
    blo, bhi = cb[:3].astype(np.float64), cb[3:].astype(np.float64)
    bg = background(H_IMG)
    inside = miss = 0
    for y in range(H_IMG):
        for x in range(W_IMG):
            rx, ry = ora.lcg_stream(x + W_IMG * y, 0, 2)
            u = np.float32(np.float32(x) + rx) / np.float32(W_IMG)
            v = np.float32(np.float32(y) + ry) / np.float32(H_IMG)
            sx = ((np.float32(1) - u) * np.float32(lo[0]) + u * np.float32(hi[0])) * np.float32(W_IMG)
            sy = ((np.float32(1) - v) * np.float32(lo[1]) + v * np.float32(hi[1])) * np.float32(H_IMG)
            d = pixel_rays(p, np.array([[float(sx)]]), np.array([[float(sy)]]))
            t0, t1 = (z[0, 0] for z in slab(EYE, d, blo, bhi))
            if t1 > t0 + 0.05:
                assert np.allclose(acc[y, x, :3], col, atol=1e-6), (y, x)
                inside += 1
            elif t1 < t0 - 0.05:
                assert np.allclose(acc[y, x, :3], bg[y, x], atol=1e-6), (y, x)
                miss += 1
    assert inside > 100 and miss > 100, (inside, miss)
    # the selection box inverts its border in the frame buffer only
    p.subImageSelecting = 1
    sel = (0.1, 0.25, 0.6, 0.75)
    for k in range(4):
        p.subImageSel[k] = sel[k]
    p.debugCrosshair = 0
    acc2, fb2 = ora.render_f32(p, om)
    assert np.array_equal(acc2, acc)
    lx, ly = int(np.float32(sel[0]) * np.float32(W_IMG)), int(np.float32(sel[1]) * np.float32(H_IMG))
    hx, hy = int(np.float32(sel[2]) * np.float32(W_IMG)), int(np.float32(sel[3]) * np.float32(H_IMG))
    ys, xs = np.mgrid[0:H_IMG, 0:W_IMG]
    border = (((xs == lx) | (xs == hx)) & (ys >= ly) & (ys <= hy)) | (((ys == ly) | (ys == hy)) & (xs >= lx) & (xs <= hx))
    val = acc.astype(np.float32)  # accumID 0, spp 1: the pixel values
    val = np.where(border[..., None], np.float32(1) - val, val)
    q = np.clip((val * np.float32(256)).astype(np.int64), 0, 255).astype(np.uint32)
    want = q[..., 0] | (q[..., 1] << 8) | (q[..., 2] << 16) | (q[..., 3] << 24)
    assert np.array_equal(fb2, want)
    assert border.sum() > 20 and not np.array_equal(fb2[border], fb[border])
