"""Pins for the oracle's integrators (test infrastructure checking test
infrastructure): the path tracer's control flow, Woodcock and ratio tracking,
HG phase-function sampling and sampleLight, each against a closed form or an
independent numpy random walk, so that a mistake shared by the GPU kernels
and the oracle would fail here.

Reference semantics restated (deviceCode.cu):
- free flights at rate majorant*os in every DDA cell; Woodcock accepts with
  probability sigma*os/majorant (736-776), ratio tracking multiplies Tr by
  1 - sigma*os/majorant (800-833): with sigma*os <= majorant both act as a
  homogeneous medium of extinction sigma*os^2;
- pathTracingIntegrate (991-1120): interaction -> `bounce++ >= maxBounces`
  zeroes the throughput; throughput *= albedo; Russian roulette when
  P = max(throughput) < 0.2 (survive with probability P, divide by P);
  Ld += throughput * sampleLight when lights are on; HG g = 0 scatter; with no
  lights Ld = 1; the sample is Ld * throughput, or the background when the
  camera ray met no interaction;
- directLightingIntegrate (1126-1207): one interaction, albedo * sampleLight;
- sampleLight (856-926): Tr of the shadow ray to min(light distance, domain
  exit) divided by ld^2, ld measured after lightSpaceTransform, times the
  intensity; the directional light (a build-side addition) is Tr * intensity;
- henyeyGreensteinSample with g = 0 (957-985) and make_orthonormal_basis
  (144-149).

The scene is one gridlet over [0.5, 8.5]^3 with a constant scalar, a 4^3
majorant grid, and a colour map whose alpha texels around the scalar make
sigma = 0.15 < majorant = 0.25.  The camera rays all run along +z through
(4.5, 4.5): chord length 8.
"""
import numpy as np
import pytest
from scipy import integrate, stats

from oracle import oracle as ora

LO, HI = 0.5, 8.5
W_IMG = H_IMG = 512
SIGMA, MAJ = 0.15, 0.25
Z_MAX = 5.0


# ---- HG and the orthonormal basis -------------------------------------------------


def test_orthonormal_basis():
    rng = np.random.default_rng(0)
    ws = rng.normal(size=(2000, 3))
    ws = np.concatenate([ws, np.eye(3), -np.eye(3), [[1, 1, 0], [0, 1, 1], [1e-4, 1, 0]]])
    for w in ws:
        w = (w / np.linalg.norm(w)).astype(np.float32)
        u, v = ora.orthonormal_basis(w)
        m = np.array([u, v, w], np.float64)
        assert np.abs(m @ m.T - np.eye(3)).max() < 3e-6
        assert np.linalg.det(m) > 0.999  # u = v x w: right-handed


@pytest.mark.parametrize("wo", [(0, 0, 1), (0.3, -0.4, np.sqrt(0.75)), (-1, 0, 0), (0.6, 0.8, 0)])
def test_hg_g0_uniform_on_sphere(wo):
    """chi-square: cos(theta) about wo uniform on [-1, 1] and the azimuth
    about wo uniform (Archimedes), and global z / azimuth uniform."""
    wo = np.asarray(wo, np.float32)
    n = 200000
    d = ora.hg_sample(wo, 12345, 7, n).astype(np.float64)
    assert np.abs(np.linalg.norm(d, axis=1) - 1).max() < 3e-6
    u, v = ora.orthonormal_basis(wo)
    pvals = []
    for c, phi in [(d @ wo.astype(np.float64), np.arctan2(d @ v.astype(np.float64), d @ u.astype(np.float64))),
                   (d[:, 2], np.arctan2(d[:, 1], d[:, 0]))]:
        for x, lo, hi in [(c, -1.0, 1.0), (phi, -np.pi, np.pi)]:
            h, _ = np.histogram(x, bins=32, range=(lo, hi))
            pvals.append(stats.chisquare(h).pvalue)
    assert min(pvals) > 1e-4, pvals
    assert np.abs(d.mean(axis=0)).max() < 0.01


# ---- the homogeneous box -------------------------------------------------------------


def _box_model(value=0.5):
    gridlets = np.array([[0, 0, 0, 0, 8, 8, 8, 0]], np.int32)
    scalars = np.full(9 ** 3, value, np.float32)
    return ora.Model(np.zeros((0, 4), np.float32), np.zeros((0, 8), np.int32), gridlets, scalars)


def _tf(rgb):
    """128 texels: alpha 0.05 below the texel pair around 0.5, 0.25 from
    texel 64 on; the filter at 0.5 blends texels 63 and 64 half and half."""
    rgba = np.zeros((128, 4), np.float32)
    rgba[:, :3] = rgb
    rgba[:, 3] = 0.05
    rgba[64:, 3] = 0.25
    return rgba


def _params(model, rgba, integrator, max_bounces=1024, os_=1.0, lights=(), light_scale=1.0):
    p = ora.OraRenderParams()
    p.W, p.H, p.spp, p.accumID = W_IMG, H_IMG, 1, 0
    p.integrator, p.maxBounces = integrator, max_bounces
    p.camOrg[:] = (4.5, 4.5, -10.0)
    p.camDir00[:] = (0.0, 0.0, 1.0)
    p.camDu[:] = (0.0, 0.0, 0.0)
    p.camDv[:] = (0.0, 0.0, 0.0)
    p.worldBounds[:] = (LO, LO, LO, HI, HI, HI)
    ident = (1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0)
    p.voxelXfm[:] = ident
    p.voxelXfmInv[:] = ident
    s = light_scale
    p.lightXfm[:] = (s, 0, 0, 0, s, 0, 0, 0, s, 0, 0, 0)
    p.gridDims[:] = (4, 4, 4)
    p.gridBounds[:] = (LO, LO, LO, HI, HI, HI)
    dom = np.array([0.0, 1.0], np.float32)
    maj = ora.map_max_opacity(model.grid_build(np.array([4, 4, 4], np.int32), np.array(p.gridBounds, np.float32)),
                              rgba, dom)
    p.tfN = len(rgba)
    p.tfDomain[:] = dom
    p.opacityScale = os_
    for i, (typ, pos, inten) in enumerate(lights):
        p.lights[i].pos[:] = pos
        p.lights[i].intensity = inten
        p.lights[i].on = 1
        p.lights[i].type = typ
    p.dt = 1.0
    p._keep = (maj, rgba)
    p.majorants = ora.ptr(maj).value
    p.tfRGBA = ora.ptr(rgba).value
    return p, maj


def _render(p, model):
    acc, _ = ora.render_f32(p, model)
    return acc.reshape(-1, 4).astype(np.float64)


def _bg():
    """background colour per pixel (deviceCode.cu:1320-1324), the row
    gradient evaluated in float32 as the renderer does"""
    f = np.float32
    t = (np.arange(H_IMG, dtype=f) / f(H_IMG))[:, None]
    one = np.ones_like(t)
    bg = np.concatenate([(f(1) - t) * f(1) + t * f(0.5), (f(1) - t) * f(1) + t * f(0.7), (f(1) - t) * f(1) + t, one],
                        axis=1)
    return np.repeat(bg, W_IMG, axis=0).astype(np.float64)


def _z(mean_a, var_a, n_a, mean_b, var_b=0.0, n_b=1):
    return np.abs(mean_a - mean_b) / np.sqrt(var_a / n_a + var_b / n_b)


def test_box_scene_sigma_and_majorant():
    m = _box_model()
    rgba = _tf(1.0)
    _, maj = _params(m, rgba, 0)
    assert np.all(maj == np.float32(MAJ))
    assert ora.tf_lookup(rgba, [0.0, 1.0], 0.5)[3] == np.float32(SIGMA)


def _walk(n, sigma, albedo, max_bounces, seed, light_y=None, rr_divide=True):
    """Independent restatement of pathTracingIntegrate's walk in the box:
    exponential free paths of rate sigma (the thinned Woodcock process),
    isotropic scattering, the bounce cap and Russian roulette with numpy's
    generator; a directional light along +y adds throughput * I *
    E[Tr] = exp(-sigma * (HI - y)) at every interaction.  Returns the
    interaction count K, the final throughput and Ld per walker.

    As in the reference, the box test of each scattered ray is clamped by
    the previous segment's end (boxTest takes min(ray.tmax, exit) and
    ray.tmax still holds the last t1, deviceCode.cu:1025, 1104; Sampler.h:
    114-115), so a segment never reaches further than the one before it."""
    rng = np.random.default_rng(seed)
    pos = np.tile([4.5, 4.5, LO], (n, 1))
    d = np.tile([0.0, 0.0, 1.0], (n, 1))
    alive = np.ones(n, bool)
    K = np.zeros(n, np.int64)
    thr = np.ones(n)
    Ld = np.zeros(n)
    tcap = np.full(n, np.inf)
    while alive.any():
        idx = np.nonzero(alive)[0]
        p, dd = pos[idx], d[idx]
        with np.errstate(divide="ignore", invalid="ignore"):
            tb = np.where(dd > 0, (HI - p) / dd, np.where(dd < 0, (LO - p) / dd, np.inf))
        limit = np.minimum(tcap[idx], tb.min(axis=1))
        tcap[idx] = limit
        t = rng.exponential(1.0 / sigma, len(idx))
        esc = t >= limit
        alive[idx[esc]] = False
        sc = idx[~esc]
        pos[sc] = p[~esc] + dd[~esc] * t[~esc, None]
        cap = K[sc] >= max_bounces
        K[sc] += 1
        thr[sc[cap]] = 0.0
        alive[sc[cap]] = False
        s2 = sc[~cap]
        thr[s2] *= albedo
        P = thr[s2].copy()
        low = P < 0.2
        kill = low & (rng.random(len(s2)) > P)
        thr[s2[kill]] = 0.0
        alive[s2[kill]] = False
        surv = low & ~kill
        if rr_divide:
            thr[s2[surv]] /= P[surv]
        s2 = s2[~kill]
        if light_y is not None:
            Ld[s2] += thr[s2] * light_y * np.exp(-sigma * (HI - pos[s2, 1]))
        v = rng.normal(size=(len(s2), 3))
        d[s2] = v / np.linalg.norm(v, axis=1, keepdims=True)
    return K, thr, Ld


@pytest.mark.parametrize("os_", [1.0, 0.5])
def test_path_tracer_albedo1_no_lights(os_):
    """Each sample is exactly the background (no interaction, probability
    exp(-sigma*os^2*L)) or exactly (1,1,1,1): Ld = 1 with no lights, no
    roulette at P = 1, `bounce ? color : bg`."""
    m = _box_model()
    p, _ = _params(m, _tf(1.0), 0, os_=os_)
    acc = _render(p, m)[W_IMG:]  # row 0's background equals (1,1,1,1)
    bg = _bg()[W_IMG:]
    is_bg = np.all(acc == bg.astype(np.float32), axis=1)
    one = np.all(acc == 1.0, axis=1)
    assert np.all(is_bg | one)
    q = np.exp(-SIGMA * os_ ** 2 * (HI - LO))
    frac = is_bg.mean()
    assert _z(frac, q * (1 - q), len(acc), q) < Z_MAX, (frac, q)


def test_path_tracer_bounce_cap():
    """maxBounces = 2: a third interaction zeroes the throughput, so
    P(colour 0) = P(K >= 3) of the walk; an off-by-one in the cap (K >= 2 or
    K >= 4) fails."""
    m = _box_model()
    p, _ = _params(m, _tf(1.0), 0, max_bounces=2)
    acc = _render(p, m)[W_IMG:]
    zero = np.all(acc[:, :3] == 0.0, axis=1) & (acc[:, 3] == 1.0)
    K, _, _ = _walk(1 << 21, SIGMA, 1.0, 1 << 30, 1)
    for k, expect_pass in [(3, True), (2, False), (4, False)]:
        q = (K >= k).mean()
        z = _z(zero.mean(), q * (1 - q), len(acc), q, q * (1 - q), len(K))
        assert (z < Z_MAX) == expect_pass, (k, zero.mean(), q, z)


def test_path_tracer_russian_roulette_unbiased():
    """albedo 0.3, no lights: the oracle's mean (Russian roulette from the
    second interaction on) equals E[0.3^K; K >= 1] + P(K = 0) * bg of the
    walk without roulette; roulette that forgets to divide by P fails."""
    m = _box_model()
    c = 0.3
    p, _ = _params(m, _tf(c), 0)
    acc = _render(p, m)
    bg = _bg()
    col = acc[:, 0]
    K, _, _ = _walk(1 << 21, SIGMA, 1.0, 1 << 30, 2)  # albedo 1: no roulette, the natural K
    p0 = (K == 0).mean()
    w = np.where(K > 0, c ** K.astype(np.float64), 0.0)
    expect = p0 * bg[:, 0].mean() + w.mean()
    assert _z(col.mean(), col.var(), len(col), expect, w.var(), len(K)) < Z_MAX, (col.mean(), expect)
    # a planted mistake in the restatement (survivors keep their throughput)
    # lies far from the oracle: the test has the power to see one there
    K2, thr2, _ = _walk(1 << 21, SIGMA, c, 1 << 30, 4, rr_divide=False)
    w2 = np.where(K2 > 0, thr2, 0.0)
    expect_wrong = (K2 == 0).mean() * bg[:, 0].mean() + w2.mean()
    assert _z(col.mean(), col.var(), len(col), expect_wrong, w2.var(), len(K2)) > 3 * Z_MAX


@pytest.mark.parametrize("albedo", [1.0, 0.5])
def test_path_tracer_directional_light(albedo):
    """A directional light along +y: Ld += throughput * Tr * I at every
    interaction, the sample is Ld * throughput (deviceCode.cu:1076-1115)."""
    m = _box_model()
    inten = 2.0
    p, _ = _params(m, _tf(albedo), 0, lights=[(1, (0.0, 1.0, 0.0), inten)])
    acc = _render(p, m)
    col = acc[:, 1]
    bg = _bg()
    K, thr, Ld = _walk(1 << 21, SIGMA, albedo, 1 << 30, 3, light_y=inten)
    samp = np.where(K > 0, Ld * thr, 0.0)
    expect = (K == 0).mean() * bg[:, 1].mean() + samp.mean()
    assert _z(col.mean(), col.var(), len(col), expect, samp.var(), len(K)) < Z_MAX, (col.mean(), expect)
    assert np.all(acc[:, 3] == 1.0)


@pytest.mark.parametrize("os_", [1.0, 0.5])
def test_direct_lighting_directional_light_closed_form(os_):
    """Single scatter + ratio-tracked shadow ray to the +y face, 4 units away
    from every point on the chord: E = q bg + (1-q) albedo I exp(-4 s) with
    s = sigma os^2, q = exp(-s L); the per-sample shadow transmittances are
    fractional products (1 - sigma os / majorant)^n."""
    m = _box_model()
    alb = np.array([0.8, 0.6, 0.4])
    inten = 2.0
    p, _ = _params(m, _tf(alb), 1, os_=os_, lights=[(1, (0.0, 1.0, 0.0), inten)])
    acc = _render(p, m)
    bg = _bg()
    s = SIGMA * os_ ** 2
    q = np.exp(-s * (HI - LO))
    expect = q * bg[:, :3].mean(axis=0) + (1 - q) * alb * inten * np.exp(-s * 4.0)
    for ch in range(3):
        assert _z(acc[:, ch].mean(), acc[:, ch].var(), len(acc), expect[ch]) < Z_MAX, (ch, acc[:, ch].mean(),
                                                                                      expect[ch])
    # the ratio tracker produced fractional transmittances, not only 0 / 1
    scat = ~np.all(acc == bg.astype(np.float32), axis=1)
    tr = acc[scat, 0] / (alb[0] * inten)
    assert np.count_nonzero((tr > 0.01) & (tr < 0.99)) > 1000


def test_direct_lighting_point_light_integral():
    """Point light outside the box with a light-space scale of 0.1
    (OWLRenderer.cpp:166-177): the mean equals the numerically integrated
    q bg + int_0^L s e^{-s t} albedo I e^{-s d(t)} / (0.1 |x_L - x(t)|)^2 dt,
    d(t) the shadow ray's path to the +y face."""
    m = _box_model()
    alb = np.array([0.7, 0.5, 0.9])
    inten = 5.0
    lp = np.array([4.5, 30.0, 4.5])
    p, _ = _params(m, _tf(alb), 1, lights=[(0, tuple(lp), inten)], light_scale=0.1)
    acc = _render(p, m)
    bg = _bg()
    s = SIGMA
    L = HI - LO

    def f(t):
        x = np.array([4.5, 4.5, LO + t])
        r = np.linalg.norm(lp - x)
        d = (HI - x[1]) * r / (lp[1] - x[1])
        return s * np.exp(-s * t) * np.exp(-s * d) / (0.1 * r) ** 2

    integral, _ = integrate.quad(f, 0.0, L, epsabs=1e-12, epsrel=1e-10)
    q = np.exp(-s * L)
    expect = q * bg[:, :3].mean(axis=0) + alb * inten * integral
    for ch in range(3):
        assert _z(acc[:, ch].mean(), acc[:, ch].var(), len(acc), expect[ch]) < Z_MAX, (ch, acc[:, ch].mean(),
                                                                                      expect[ch])
    # the same without the light-space scale is 100x brighter: the scale is applied
    assert abs(acc[:, 0].mean() - (q * bg[:, 0].mean() + alb[0] * inten * integral * 0.01)) > 0.01


def test_direct_lighting_two_lights_one_sample():
    """Two lights on: sampleLight picks one uniformly (uniformSampleOneLight,
    deviceCode.cu:82-86) and does not weight it by the light count, so the
    mean is the average of the two lights' contributions (both shadow rays
    run 4 units through the medium)."""
    m = _box_model()
    alb = np.array([0.8, 0.6, 0.4])
    i1, i2 = 1.0, 3.0
    p, _ = _params(m, _tf(alb), 1, lights=[(1, (0.0, 1.0, 0.0), i1), (1, (-1.0, 0.0, 0.0), i2)])
    acc = _render(p, m)
    bg = _bg()
    q = np.exp(-SIGMA * (HI - LO))
    expect = q * bg[:, :3].mean(axis=0) + (1 - q) * alb * np.exp(-SIGMA * 4.0) * 0.5 * (i1 + i2)
    for ch in range(3):
        assert _z(acc[:, ch].mean(), acc[:, ch].var(), len(acc), expect[ch]) < Z_MAX, (ch, acc[:, ch].mean(),
                                                                                      expect[ch])
    scat = ~np.all(acc == bg.astype(np.float32), axis=1)
    tr = acc[scat, 0] / alb[0]
    assert np.count_nonzero(tr > 1.5 * np.exp(-SIGMA * 4.0) * i1) > 1000  # the brighter light was picked
