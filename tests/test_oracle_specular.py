"""Pins of the oracle's camera rays, specular and transmissive transport,
envmap lookup and Russian roulette against closed forms.

Each check is chosen so that a plausible slip fails it:

* camera rays (camera.h:28-47): every ray passes through its pixel's point on
  the focus plane and starts on the lens disk, area-uniformly (a flipped row
  order, a missing sqrt or a disk not re-aimed at the focus point fail);
* refraction (bsdfs.h:101-114, shading.h:35-53,182-202): Snell's law on
  entering and leaving glass, total internal reflection exactly beyond the
  critical angle, Schlick frequencies, the IOR stack pushed on entry, popped
  on exit and indexed one below the top for the far side (an inverted η or a
  wrong stack index fail);
* event weights (shading.h:150-213): METALLIC = Schlick F(base colour)·G1 of
  the outgoing direction, FRESNEL = G1, TRANSMIT = base colour·G1;
* the GGX VNDF sampler (bsdfs.h:27-72): moments of the returned half vector
  against a quadrature of Heitz's visible-normal density with the reference's
  rejection rule (α = roughness², the stretch and its sign are fixed by it);
* specular NEE (light_sampling.h:358-390): a point light over a conductor and
  a dielectric floor, per pixel, against D·F·G1·G1/(4 n·l n·v) in float64;
  NEE from an emissive triangle pair (light_sampling.h:82-100,283-296)
  against a quadrature of the irradiance;
* the envmap (envmap.h:9-20): camera misses return the texel at the z-up
  equirectangular index of their direction (texels hold their own index);
* Russian roulette (emit_photons.cl:81-91, trace_rays.cl:150-156): photons in
  a closed box deposit P·4π·ρ^j on average at their j-th diffuse hit; with
  one channel of albedo 1 no photon is ever culled; camera paths in a closed
  gray mirror box score exactly S3, S3 + E or S3 + 2E with the survival
  frequencies q = max(throughput).
"""
import math

import numpy as np
import pytest

import oracle
from phosphor_amd import PassConfig, scenes
from phosphor_amd.abi import MATERIAL_DTYPE

L = oracle.lib
M32 = (1 << 32) - 1
EV_DIFFUSE, EV_METALLIC, EV_FRESNEL, EV_TRANSMIT = 0, 1, 2, 3


# ---------------- PCG32 from its definition (random.h:16-36) ----------------

def pcg_seed(seed: int) -> int:
    s = (0 * 747796405 + 2891336453) & M32
    s = (s + seed) & M32
    return (s * 747796405 + 2891336453) & M32


def pcg_float(state: int) -> tuple[int, float]:
    """Next state and random_float: (f32)u32 / 2^32 (round to nearest)."""
    old = state
    state = (state * 747796405 + 2891336453) & M32
    word = ((old >> ((old >> 28) + 4)) ^ old) & M32
    u = (word * 277803737) & M32
    return state, float(np.float32(u) / np.float32(4294967296.0))


def _material(**kw):
    m = np.zeros(1, MATERIAL_DTYPE)
    m[0] = scenes.default_material(**kw)
    return m


def _unit(v):
    v = np.asarray(v, np.float64)
    return v / np.linalg.norm(v, axis=-1, keepdims=True)


# ---------------- camera rays (camera.h:28-47, camera.cpp:15-56) ----------------

def _floor_scene(cam, mat=None, light=((0.3, 1.0, 0.2), (2.0, 2.0, 2.0)), half=2.0):
    b = scenes.SceneBuilder()
    m = b.material(mat if mat is not None else scenes.default_material((0.5, 0.5, 0.5)))
    b.add_quad((-half, 0, -half), (-half, 0, half), (half, 0, half), (half, 0, -half), m)
    if light is not None:
        b.add_point_light(*light)
    return b.build(cam, name="floor")


@pytest.mark.parametrize("defocus", [0.0, 4.0])
def test_camera_rays_pass_through_their_pixel_on_the_focus_plane(defocus):
    W, H, seed = 40, 24, 4242
    pos = np.array([0.3, 1.7, 2.0])
    look = np.array([0.0, 0.2, 0.0])
    focus = 2.5
    cam = scenes.make_camera(pos, look, (0.0, 1.0, 0.0), 50.0, W / H, defocus_angle=defocus, focus_distance=focus)
    sc = _floor_scene(cam)
    r = oracle.RefRenderer(PassConfig(width=W, height=H, photons_per_round=64, grid_res=8, seed=1), sc)
    c = cam[0]
    ll, hz, vt = (c[k][:3].astype(np.float64) for k in ("lower_left_corner", "horizontal", "vertical"))
    du, dv = c["defocus_disk_u"][:3].astype(np.float64), c["defocus_disk_v"][:3].astype(np.float64)
    w = _unit(pos - look)
    R = focus * math.tan(math.radians(defocus / 2))
    r2 = []
    for tid in range(W * H):
        px, py = tid % W, tid // W
        o, d = r.camera_ray(seed, tid)
        o, d = o[:3].astype(np.float64), d[:3].astype(np.float64)
        assert abs(np.linalg.norm(d) - 1) < 2e-6
        off = o - pos
        assert abs(np.dot(off, w)) < 1e-6                  # the lens disk is perpendicular to w
        # the ray meets the focus plane (distance `focus` along −w) inside its pixel
        hit = o + d * ((focus - np.dot(off, -w)) / np.dot(d, -w))
        rel = hit - ll
        s_ = np.dot(rel, hz) / np.dot(hz, hz)
        t_ = np.dot(rel, vt) / np.dot(vt, vt)
        assert px / W - 1e-5 <= s_ <= (px + 1) / W + 1e-5, (tid, s_)
        assert 1 - (py + 1) / H - 1e-5 <= t_ <= 1 - py / H + 1e-5, (tid, t_)   # row 0 at the top
        # exactly the pixel point of the first two draws and the lens point of the next two
        st = pcg_seed((seed + tid) & M32)
        st, u = pcg_float(st)
        st, v = pcg_float(st)
        s_exp = (px + 0.5 + (u - 0.5)) / W
        t_exp = 1 - (py + 0.5 + (v - 0.5)) / H
        assert abs(s_ - s_exp) < 2e-6 and abs(t_ - t_exp) < 2e-6
        if defocus:
            st, a = pcg_float(st)
            st, b = pcg_float(st)
            disk = math.sqrt(a) * (math.cos(2 * math.pi * b) * du + math.sin(2 * math.pi * b) * dv)
            assert np.allclose(off, disk, atol=2e-6), (tid, off, disk)
            assert np.linalg.norm(off) <= R * (1 + 1e-5)
            r2.append(float(np.dot(off, off)))
        else:
            assert np.array_equal(o, pos.astype(np.float32).astype(np.float64))
    if defocus:
        # area-uniform on the disk: E[r²] = R²/2, sd of r² is R²/√12
        m = np.mean(r2)
        assert abs(m - R * R / 2) < 5 * (R * R / math.sqrt(12)) / math.sqrt(len(r2))
    r.close()


# ---------------- BSDF sampling with an IOR stack ----------------

class Bsdf:
    """Draws sample_bsdf (shading.h:150-213) through the oracle with an explicit IOR stack."""

    def __init__(self, seed=17):
        self.state = np.array([L().ref_pcg_seed(seed)], np.uint32)

    def sample(self, mat, view, front=True, stack=(1.0,), normal=(0.0, 1.0, 0.0)):
        st = np.zeros(8, np.float32)
        st[:len(stack)] = stack
        cur = np.array([len(stack) - 1], np.uint32)
        v = np.zeros(4, np.float32)
        v[:3] = _unit(view)
        n = np.array([*normal, 0.0], np.float32)
        d = np.zeros(4, np.float32)
        t = np.zeros(4, np.float32)
        ev = L().ref_sample_bsdf_ior(self.state.ctypes.data, mat.ctypes.data, n.ctypes.data, v.ctypes.data,
                                      1 if front else 0, st.ctypes.data, cur.ctypes.data, d.ctypes.data,
                                      t.ctypes.data)
        return ev, d[:3].astype(np.float64), t[:3].astype(np.float64), list(st[:int(cur[0]) + 1])


def _schlick(r0, cos):
    return r0 + (1 - r0) * (1 - cos) ** 5


def _view(theta_deg, azimuth_deg=0.0):
    th, az = math.radians(theta_deg), math.radians(azimuth_deg)
    return np.array([math.sin(th) * math.cos(az), math.cos(th), math.sin(th) * math.sin(az)])


def _half_reflect(v, d):
    """Microfacet normal of a mirror event: h ∝ v + d."""
    return _unit(v + d)


def _half_refract(v, d, eta1, eta2):
    """Microfacet normal of a refraction from medium η1 into η2 along i = −v:
    Snell's law η1·i_t = η2·d_t about h makes h ∝ η1·i − η2·d (Walter 2007);
    oriented to the normal's side."""
    h = _unit(eta1 * (-v) - eta2 * d)
    return h if h[1] > 0 else -h


def _check_normals(hs, what):
    """At α = 10⁻⁶ the sampled microfacet normals lie within ~10⁻⁵ rad of n
    (GGX's tail leaves a few stragglers; an inverted η or a wrong stack
    index puts the recovered normal degrees away)."""
    if not hs:
        return
    ang = np.arccos(np.clip(np.array(hs)[:, 1], -1, 1))
    assert np.quantile(ang, 0.99) < 3e-5 and ang.max() < 0.02, (what, np.quantile(ang, 0.99), ang.max())


@pytest.mark.parametrize("theta", [0.0, 30.0, 60.0, 80.0])
def test_refraction_into_glass_follows_snell(theta):
    """Air → glass (η1 = 1, η2 = 1.5) at a near-smooth surface (α = 10⁻⁶):
    FRESNEL with Schlick's F (r0 = 0.04) about the normal, else TRANSMIT
    obeying Snell's law about the normal, going on forward, throughput = base
    colour (G1 = 1), and 1.5 pushed on the stack."""
    mat = _material(base_color=(0.9, 0.8, 0.7), transmission=1.0, ior=1.5, roughness=0.001)
    s = Bsdf(seed=int(theta) + 1)
    n = 6000
    v = _view(theta)
    fres = 0
    hr, ht = [], []
    for _ in range(n):
        ev, d, thr, stack = s.sample(mat, v)
        if ev == EV_FRESNEL:
            fres += 1
            assert stack == [1.0]
            hr.append(_half_reflect(v, d))
        else:
            assert ev == EV_TRANSMIT and d[1] < 0
            assert np.allclose(thr, [0.9, 0.8, 0.7], rtol=1e-6)
            assert stack == [1.0, 1.5]
            ht.append(_half_refract(v, d, 1.0, 1.5))
            if theta > 0:
                assert d[0] < 0                              # forward: the tangential part continues
    _check_normals(hr, "reflect")
    _check_normals(ht, "refract")
    p = _schlick(0.04, math.cos(math.radians(theta)))
    assert abs(fres / n - p) < 5 * math.sqrt(p * (1 - p) / n) + 1e-3, (fres / n, p)


@pytest.mark.parametrize("theta,tir", [(20.0, False), (41.5, False), (42.2, True), (60.0, True)])
def test_leaving_glass_refracts_or_reflects_totally(theta, tir):
    """Glass → air from inside (stack [1, 1.5], back face, η1 = 1.5, η2 = 1):
    the critical angle is asin(1/1.5) = 41.81°; below it TRANSMIT obeys Snell
    (sin θt = 1.5 sin θ) and pops the stack to air, beyond it every event is
    FRESNEL about the normal and the stack stays."""
    mat = _material(transmission=1.0, ior=1.5, roughness=0.001)
    s = Bsdf(seed=int(theta * 10))
    n = 4000
    th = math.radians(theta)
    v = _view(theta)
    trans = 0
    hr, ht = [], []
    for _ in range(n):
        ev, d, thr, stack = s.sample(mat, v, front=False, stack=(1.0, 1.5))
        if ev == EV_TRANSMIT:
            trans += 1
            assert stack == [1.0] and d[1] < 0
            ht.append(_half_refract(v, d, 1.5, 1.0))
        else:
            assert ev == EV_FRESNEL and stack == [1.0, 1.5]
            hr.append(_half_reflect(v, d))
    _check_normals(hr, "reflect")
    _check_normals(ht, "refract")
    if tir:
        assert trans == 0
    else:
        p = 1 - _schlick(0.04, math.cos(th))
        assert abs(trans / n - p) < 5 * math.sqrt(p * (1 - p) / n) + 1e-3


def test_ior_stack_nested_media_and_slab():
    """Entering water (1.333) from inside glass refracts with η1 = 1.5, η2 =
    1.333 and pushes 1.333; leaving the water again refracts with η1 = 1.333,
    η2 = 1.5 (the entry below the top) and pops back to glass; a glass slab
    entered and left returns the stack to air; a full stack (8 entries) is not
    pushed further."""
    water = _material(transmission=1.0, ior=1.333, roughness=0.001)
    glass = _material(transmission=1.0, ior=1.5, roughness=0.001)
    s = Bsdf(seed=5)
    v = _view(35.0)

    def transmits(mat, front, stack, n=300):
        out = []
        for _ in range(n):
            ev, d, _, st = s.sample(mat, v, front=front, stack=stack)
            if ev == EV_TRANSMIT:
                out.append((d, st))
        assert out, "no transmission"
        return out

    for front, stack, e1, e2, after in ((True, (1.0, 1.5), 1.5, 1.333, [1.0, 1.5, 1.333]),
                                        (False, (1.0, 1.5, 1.333), 1.333, 1.5, [1.0, 1.5])):
        got = transmits(water, front, stack)
        _check_normals([_half_refract(v, d, e1, e2) for d, _ in got], (front, stack))
        assert all(np.allclose(st, after) for _, st in got)
    _, st = transmits(glass, True, (1.0,))[0]
    assert st == [1.0, 1.5]
    _, st = transmits(glass, False, tuple(st))[0]
    assert st == [1.0]
    full = tuple([1.0] + [1.5] * 7)
    _, st = transmits(_material(transmission=1.0, ior=1.0, roughness=0.001), True, full)[0]
    assert len(st) == 8


def _g1(c, a):
    c = np.maximum(c, 0.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(c > 0, 2 * c / (c + np.sqrt(a * a + (1 - a * a) * c * c)), 0.0)


@pytest.mark.parametrize("rough,theta", [(0.3, 20.0), (0.7, 65.0)])
def test_metallic_weight_is_schlick_of_base_color_times_g1(rough, theta):
    f0 = np.array([0.9, 0.55, 0.2])
    mat = _material(base_color=tuple(f0), metallic=1.0, roughness=rough)
    s = Bsdf(seed=31)
    v = _view(theta)
    a = rough * rough
    for _ in range(3000):
        ev, d, thr, _ = s.sample(mat, v)
        assert ev == EV_METALLIC
        h = _unit(v + d)                                  # the mirror direction about h
        want = (f0 + (1 - f0) * (1 - abs(np.dot(v, h))) ** 5) * _g1(d[1], a)
        assert np.allclose(thr, want, rtol=2e-4, atol=1e-6), (thr, want)


def test_fresnel_and_transmit_weights_are_g1():
    rough = 0.5
    a = rough * rough
    mat = _material(base_color=(0.8, 0.6, 0.4), transmission=1.0, ior=1.5, roughness=rough)
    s = Bsdf(seed=77)
    v = _view(50.0)
    seen = set()
    for _ in range(3000):
        ev, d, thr, _ = s.sample(mat, v)
        seen.add(ev)
        if ev == EV_FRESNEL:
            assert np.allclose(thr, _g1(d[1], a), rtol=1e-5, atol=1e-7)
        elif ev == EV_TRANSMIT:
            assert d[1] < 0
            assert np.allclose(thr, np.array([0.8, 0.6, 0.4]) * _g1(abs(d[1]), a), rtol=1e-5, atol=1e-7)
    assert seen == {EV_FRESNEL, EV_TRANSMIT}


def _vndf_moments(alpha, theta_v, n_th=1500, n_ph=720):
    """Quadrature of Heitz's visible-normal density D_v(h) = G1(v)·max(0, v·h)·D(h)/(n·v)
    over the hemisphere, restricted to the reference's acceptance test
    (reflect(−v, h)·n ≥ 0), with up to 8 attempts and the normal as fallback.
    Returns E[h·n], E[h·x], E[h·z], E[(h·z)²] (x: the view's tangential
    direction, z ⟂ both; E[h·z] = 0 by symmetry)."""
    th = (np.arange(n_th) + 0.5) / n_th * (math.pi / 2)
    ph = (np.arange(n_ph) + 0.5) / n_ph * (2 * math.pi)
    T, P = np.meshgrid(th, ph, indexing="ij")
    hx, hy, hz = np.sin(T) * np.cos(P), np.cos(T), np.sin(T) * np.sin(P)
    vx, vy = math.sin(theta_v), math.cos(theta_v)
    a2 = alpha * alpha
    D = a2 / (math.pi * (hy * hy * (a2 - 1) + 1) ** 2)
    vh = hx * vx + hy * vy
    g1v = 2 * vy / (vy + math.sqrt(a2 + (1 - a2) * vy * vy))
    dens = g1v * np.maximum(vh, 0) * D / vy
    dw = np.sin(T) * (math.pi / 2 / n_th) * (2 * math.pi / n_ph)
    acc = (-vy + 2 * vh * hy) >= 0
    total = (dens * dw).sum()
    assert abs(total - 1) < 2e-3, total                      # the density integrates to one
    p_acc = (dens * dw * acc).sum()
    w = dens * dw * acc / p_acc
    p_fb = (1 - p_acc) ** 8
    m = [(w * f).sum() for f in (hy, hx, hz, hz * hz)]
    return [(1 - p_fb) * m[0] + p_fb, (1 - p_fb) * m[1], (1 - p_fb) * m[2], (1 - p_fb) * m[3]]


@pytest.mark.parametrize("rough,theta,az", [(0.5, 0.0, 0.0), (0.5, 60.0, 30.0), (0.3, 75.0, 0.0),
                                            (0.3, 75.0, 115.0), (0.8, 45.0, 250.0)])
def test_ggx_vndf_moments_match_quadrature(rough, theta, az):
    """The view is turned by an azimuth about the normal so that the sampler's
    tangent frame (make_tbn) is not aligned with the plane of incidence: a
    sampler mirrored inside that frame breaks E[h·z] = 0."""
    alpha = rough * rough
    mat = _material(base_color=(1.0, 1.0, 1.0), metallic=1.0, roughness=rough)
    s = Bsdf(seed=1000 + int(theta) + int(az))
    v = _view(theta, az)
    a = math.radians(az)
    xd = np.array([math.cos(a), 0.0, math.sin(a)])
    zd = np.array([-math.sin(a), 0.0, math.cos(a)])
    n = 20000
    hs = np.zeros((n, 3))
    for i in range(n):
        _, d, _, _ = s.sample(mat, v)
        hs[i] = _unit(v + d)
    hz = hs @ zd
    sample = [hs[:, 1], hs @ xd, hz, hz ** 2]
    want = _vndf_moments(alpha, math.radians(theta))
    for k, (x, m) in enumerate(zip(sample, want)):
        se = x.std() / math.sqrt(n)
        assert abs(x.mean() - m) < 5 * se + 1e-4, (k, x.mean(), m, se)


# ---------------- specular NEE (light_sampling.h:358-390) ----------------

def _spec_closed_form(pos, view, light, power, f0, alpha):
    n = np.array([0.0, 1.0, 0.0])
    lv = light[None, :] - pos
    d2 = (lv * lv).sum(1)
    l = lv / np.sqrt(d2)[:, None]
    ndl = l @ n
    ndv = view @ n
    h = _unit(l + view)
    ndh = np.maximum(h @ n, 0)
    a2 = alpha * alpha
    D = a2 / np.maximum(math.pi * (ndh * ndh * (a2 - 1) + 1) ** 2, 1e-5)
    F = f0[None, :] + (1 - f0[None, :]) * ((1 - np.abs((view * h).sum(1))) ** 5)[:, None]
    G = _g1(ndl, alpha) * _g1(ndv, alpha)
    rad = power * ndl / d2
    return (rad * D * G / (4 * ndl * ndv))[:, None] * F


@pytest.mark.parametrize("kind", ["conductor", "dielectric"])
def test_specular_nee_point_light_closed_form(kind):
    """Camera rays hit a GGX floor (roughness 0.5) lit by a point light; the
    first event is METALLIC (conductor, F0 = base colour) or, where it is
    FRESNEL, the dielectric's F0 = ((1 − 1.5)/(1 + 1.5))² = 0.04; the bounce
    then escapes, so the pixel's emission is exactly the specular NEE term."""
    W = H = 48
    seed = 777
    light, power = np.array([0.3, 1.0, 0.2]), 2.0
    cam = scenes.make_camera((0.0, 2.0, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 40.0, 1.0)
    if kind == "conductor":
        f0 = np.array([0.95, 0.64, 0.54])
        mat = scenes.default_material(tuple(f0), metallic=1.0, roughness=0.5)
    else:
        f0 = np.array([0.04, 0.04, 0.04])
        mat = scenes.default_material((0.5, 0.5, 0.5), metallic=0.0, roughness=0.5, ior=1.5)
    sc = _floor_scene(cam, mat, light=(tuple(light), (power,) * 3))
    r = oracle.RefRenderer(PassConfig(width=W, height=H, photons_per_round=64, grid_res=8, seed=1,
                                      direct_samples=16), sc)
    r.trace_range(seed, 0, W * H)
    hp = r.read_hitpoints()
    o = np.zeros((W * H, 3))
    d = np.zeros((W * H, 3))
    for tid in range(W * H):
        oo, dd = r.camera_ray(seed, tid)
        o[tid], d[tid] = oo[:3], dd[:3]
    t = -o[:, 1] / d[:, 1]
    pos = o + t[:, None] * d
    want = _spec_closed_form(pos, -d, light, power, f0, 0.25)
    em = hp[:, 5, :].astype(np.float64)
    sel = em[:, 3] == 0.0                                 # not a diffuse hitpoint
    if kind == "conductor":
        assert sel.all()
    else:
        assert 0.02 < sel.mean() < 0.3
    assert np.allclose(em[sel, :3], want[sel], rtol=2e-4, atol=1e-7), \
        np.abs(em[sel, :3] / want[sel] - 1).max()
    r.close()


def test_area_light_nee_matches_irradiance_quadrature():
    """A 2-triangle 0.006 × 0.006 emissive quad (E = 400, facing down; area
    above the 10⁻⁵ cutoff of light_sampling.h:279) at height 3 over a diffuse
    floor: each light sample is E·A·cosθ·cosθl/d² (the LIGHT_TEXTURED fragment
    with power π·E·A), the hitpoint's direct term ρ/π times their mean. The
    integrand varies by ~10⁻³ over the light, so the 16-sample mean equals a
    Gauss quadrature of the irradiance within 3·10⁻³ per pixel and 3·10⁻⁴ on
    average (a missing cosine, π or area factor fails by far more)."""
    W = H = 32
    seed = 99
    b = scenes.SceneBuilder()
    floor = b.material(scenes.default_material((0.6, 0.5, 0.4), ior=1.0, roughness=0.001))
    b.add_quad((-2, 0, -2), (-2, 0, 2), (2, 0, 2), (2, 0, -2), floor)
    E = 400.0
    lm = b.material(scenes.default_material((0.0, 0.0, 0.0), emissive=(E, E, E)))
    s = 0.003
    c = np.array([0.2, 3.0, -0.1])
    # p0..p3 wound so the face normal points down (−y)
    b.add_quad(c + [-s, 0, -s], c + [s, 0, -s], c + [s, 0, s], c + [-s, 0, s], lm)
    cam = scenes.make_camera((0.0, 2.5, 0.0), (0.0, 0.0, 0.0), (0.0, 0.0, -1.0), 60.0, 1.0)
    sc = b.build(cam, name="area")
    assert sc.lights["tangent"].shape[0] == 1
    r = oracle.RefRenderer(PassConfig(width=W, height=H, photons_per_round=64, grid_res=8, seed=1), sc)
    r.trace_range(seed, 0, W * H)
    hp = r.read_hitpoints()
    keep = (hp[:, 5, 3] == 1.0) & (np.abs(hp[:, 0, 1]) < 1e-6)
    assert keep.sum() > 0.9 * W * H
    pos = hp[keep, 0, :3].astype(np.float64)
    direct = hp[keep, 4, :3].astype(np.float64)
    g, wg = np.polynomial.legendre.leggauss(6)
    irr = np.zeros(len(pos))
    for xi, wx in zip(g, wg):
        for zi, wz in zip(g, wg):
            q = c + np.array([xi * s, 0.0, zi * s])
            lv = q[None, :] - pos
            d2 = (lv * lv).sum(1)
            cos_f = lv[:, 1] / np.sqrt(d2)
            irr += wx * wz * (s * s) * E * cos_f * cos_f / d2      # cosθl = cosθ for parallel planes
    want = irr[:, None] * np.array([0.6, 0.5, 0.4]) / math.pi
    ratio = direct / want
    assert np.abs(ratio - 1).max() < 3e-3, np.abs(ratio - 1).max()
    assert abs(ratio.mean() - 1) < 3e-4, ratio.mean()
    r.close()


# ---------------- envmap (envmap.h:9-20) ----------------

@pytest.mark.parametrize("look", [(1, 0.2, 0.1), (0.1, 0.3, 1), (-0.2, 0.1, -1), (-1, -0.7, 0.05)])
def test_envmap_lookup_is_z_up_equirectangular(look):
    EW, EH = 64, 32
    env = np.zeros((EH, EW, 3), np.float32)
    env[:, :, 0] = np.arange(EW)[None, :]
    env[:, :, 1] = np.arange(EH)[:, None]
    env[:, :, 2] = 1000 + np.arange(EW)[None, :] + EW * np.arange(EH)[:, None]
    W, H, seed = 48, 32, 5
    up = (0.0, 0.0, 1.0) if abs(look[2]) < 0.9 else (0.0, 1.0, 0.0)
    cam = scenes.make_camera((0.0, 0.0, 0.0), look, up, 150.0, W / H)
    b = scenes.SceneBuilder()
    m = b.material(scenes.default_material())
    far = 500.0                                         # one small triangle far outside the view
    b.add_mesh(np.array([[[far, far, far], [far + 1, far, far], [far, far + 1, far]]], np.float32), m)
    sc = b.build(cam, name="env", envmap=env)
    r = oracle.RefRenderer(PassConfig(width=W, height=H, photons_per_round=64, grid_res=8, seed=1), sc)
    r.trace_range(seed, 0, W * H)
    em = r.read_hitpoints()[:, 5, :3]
    exact = near = 0
    for tid in range(W * H):
        _, d = r.camera_ray(seed, tid)
        d = d[:3].astype(np.float64)
        theta = math.acos(max(-1.0, min(1.0, d[2])))
        phi = math.atan2(d[1], d[0])
        u = (phi + math.pi) / (2 * math.pi)
        v = min(max(1 - theta / math.pi, 0.0), 1.0)
        u -= math.floor(u)
        px = min(int(u * EW), EW - 1)
        py = min(int(v * EH), EH - 1)
        got = em[tid]
        if np.array_equal(got, env[py, px]):
            exact += 1
            continue
        # only where the index sits on a texel edge may float32 rounding pick the neighbour
        fu, fv = u * EW, v * EH
        assert min(fu - math.floor(fu), math.ceil(fu) - fu) < 2e-3 or \
            min(fv - math.floor(fv), math.ceil(fv) - fv) < 2e-3, (tid, got, env[py, px])
        assert abs(int(got[0]) - px) in (0, 1, EW - 1) and abs(int(got[1]) - py) <= 1
        near += 1
    assert exact >= 0.98 * W * H, (exact, near)
    r.close()


# ---------------- Russian roulette (emit_photons.cl:81-91, trace_rays.cl:150-156) ----------------

def _closed_box(albedo, light=(0.1, 0.05, -0.07), power=1.0):
    b = scenes.SceneBuilder()
    # ior 1 (r0 = 0) and roughness 0.001 (G1 = 1): FRESNEL events are lossless mirror bounces
    m = b.material(scenes.default_material(albedo, ior=1.0, roughness=0.001))
    b.add_box((-1, -1, -1), (1, 1, 1), m)
    b.add_point_light(light, (power, power, power))
    cam = scenes.make_camera((0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (0.0, 1.0, 0.0), 60.0, 1.0)
    return b.build(cam, name="closed box")


def _decode(pw):
    out = np.zeros((len(pw), 4), np.float32)
    for i, v in enumerate(pw):
        L().ref_decode_rgbe(int(v), out[i].ctypes.data)
    return out[:, :3].astype(np.float64)


def _deposit_table(counts, dec, jmax):
    """power of each photon's j-th deposit (0 where it made fewer)."""
    n = len(counts)
    c = counts.astype(np.int64)
    first = np.concatenate([[0], np.cumsum(c)[:-1]])
    tab = np.zeros((jmax, n, 3))
    for j in range(jmax):
        has = counts > j
        tab[j, has] = dec[first[has] + j]
    return tab


def test_russian_roulette_never_culls_at_albedo_one():
    """Albedo (1, 0.5, 0.25): the throughput's largest channel stays 1, so
    q = clamp(max, 0.05, 1) = 1 and no photon ends before max_bounces; every
    deposit carries P·4π·ρ^j exactly (no 1/q boost), RGBE-rounded down."""
    rho = np.array([1.0, 0.5, 0.25])
    sc = _closed_box(tuple(rho))
    n, B = 1 << 13, 16
    r = oracle.RefRenderer(PassConfig(width=4, height=4, photons_per_round=n, max_photons_per_batch=n,
                                      max_photon_bounces=B, grid_res=8, seed=1), sc)
    counts, _, pw, _ = r.emit_counted(4321, 0, n)
    assert counts.max() <= B
    # FRESNEL hits (probability (1 − cos)^5 at incidence cos) take bounces without a deposit
    assert counts.mean() > 0.9 * B and counts.min() >= B - 8
    tab = _deposit_table(counts, _decode(pw), 8)
    P = 4 * math.pi
    for j in range(8):
        want = P * rho ** j
        got = tab[j][counts > j]
        assert np.all(got[:, 0] <= want[0] * (1 + 1e-6)) and np.all(got[:, 0] >= want[0] * (1 - 2 ** -7))
        step = 2.0 ** (math.frexp(P)[1] - 8)             # channel step of the shared exponent
        assert np.all(np.abs(got[:, 1:] - want[1:]) <= step + 1e-6)
    r.close()


def test_russian_roulette_is_unbiased_per_deposit():
    """Albedo 0.5: roulette from the third hit on with q = max(throughput); the
    mean power of the j-th deposit over all photons stays P·4π·0.5^j."""
    sc = _closed_box((0.5, 0.5, 0.5))
    n, B = 1 << 15, 16
    r = oracle.RefRenderer(PassConfig(width=4, height=4, photons_per_round=n, max_photons_per_batch=n,
                                      max_photon_bounces=B, grid_res=8, seed=1), sc)
    counts, _, pw, _ = r.emit_counted(8765, 0, n)
    tab = _deposit_table(counts, _decode(pw), 7)
    P = 4 * math.pi
    frac_alive = [(counts > j).mean() for j in range(7)]
    assert frac_alive[0] > 0.97 and frac_alive[4] < 0.5     # roulette culls, deposits shrink
    for j in range(7):
        x = tab[j, :, 0]
        want = P * 0.5 ** j
        se = x.std() / math.sqrt(n)
        assert abs(x.mean() - want * (1 - 2 ** -9)) < 5 * se + want * 4e-3, (j, x.mean(), want, se)
    r.close()


def _mirror_path(o, d, n_hits, f0):
    """Mirror bounces inside the box [−1, 1]³: Schlick factor of each hit."""
    p = o.copy()
    d = d.copy()
    fs = []
    for _ in range(n_hits):
        t = np.where(d > 0, (1 - p) / np.where(d > 0, d, 1), np.where(d < 0, (-1 - p) / np.where(d < 0, d, 1), np.inf))
        a = int(np.argmin(t))
        p = p + t[a] * d
        fs.append(f0 + (1 - f0) * (1 - abs(d[a])) ** 5)
        d[a] = -d[a]
    return fs


def test_camera_path_roulette_scores_only_boosted_values():
    """Camera inside a closed gray mirror box (F0 = 0.8, emissive walls E, no
    lights, so NEE adds nothing): hits 0, 1 and 2 always score, then roulette at
    depths 2 and 3 with q = max(throughput) leaves the throughput at exactly 1,
    so a pixel's emission is S3 = E(1 + F0 + F0F1), S3 + E or S3 + 2E (five hits
    at most), with survival frequencies q2 = F0F1F2 and q3 = F3."""
    E = 0.5
    b = scenes.SceneBuilder()
    m = b.material(scenes.default_material((0.8, 0.8, 0.8), emissive=(E, E, E), metallic=1.0, roughness=0.001))
    b.add_box((-1, -1, -1), (1, 1, 1), m)
    W = H = 64
    seed = 2024
    cam = scenes.make_camera((0.2, -0.1, 0.3), (1.0, 0.3, -0.2), (0.0, 1.0, 0.0), 70.0, 1.0)
    sc = b.build(cam, name="mirror box")
    sc.lights = sc.lights[:0]                         # emissive walls, but no light to sample
    r = oracle.RefRenderer(PassConfig(width=W, height=H, photons_per_round=64, grid_res=8, seed=1), sc)
    r.trace_range(seed, 0, W * H)
    em = r.read_hitpoints()[:, 5, :3].astype(np.float64)
    assert np.allclose(em[:, 0], em[:, 1]) and np.allclose(em[:, 0], em[:, 2])
    surv2 = surv3 = q2s = q3s = 0.0
    for tid in range(W * H):
        o, d = r.camera_ray(seed, tid)
        f = _mirror_path(o[:3].astype(np.float64), d[:3].astype(np.float64), 4, 0.8)
        s3 = E * (1 + f[0] + f[0] * f[1])
        q2 = f[0] * f[1] * f[2]
        got = em[tid, 0]
        k = round((got - s3) / E)
        assert k in (0, 1, 2) and abs(got - (s3 + k * E)) < 2e-4 * got, (tid, got, s3)
        surv2 += k >= 1
        q2s += q2
        if k >= 1:
            surv3 += k == 2
            q3s += f[3]
    n = W * H
    assert abs(surv2 - q2s) < 5 * math.sqrt(q2s * (1 - q2s / n)), (surv2, q2s)
    assert abs(surv3 - q3s) < 5 * math.sqrt(max(q3s * (1 - q3s / max(surv2, 1)), 1.0)), (surv3, q3s)
    r.close()
