"""Closed-form pins of the oracle's shading, deep-shadow and raster
arithmetic, each derived by hand from the reference's WGSL (not from the
oracle's code), and each chosen so that a plausible transcription slip
fails it: a dropped absorption term, a wrong sign in Mp's exponent, a
missing p*pi in Phi, a wrong Legendre weight or 1/2 factor in Np, a wrong
DOM layer split, or a back-to-front instead of front-to-back blend.

WGSL quirks kept literally (strand_shading.wgsl): A passes the angle
acos(cos(theta_d) cos(gamma_i)) where fresnel_reflectance expects a cosine
(:398), the clamp(lo, hi, x) argument order makes Mp's angle clamps
one-sided (:491-492), and Mp's small-variance branch returns 0 at the peak
(:494-497).
"""
import ctypes as C
import math

import numpy as np
import pytest
from scipy import special

from bevy_strand_rasterizer_amd import abi, scene
from oracle import oracle

F32 = 2e-6  # relative, single precision through a few transcendental calls


@pytest.fixture(scope="module")
def lib():
    return oracle.load()


def vec3(x):
    return (C.c_float * 3)(*x)


def call3(fn, *args):
    out = (C.c_float * 3)()
    fn(*args, out)
    return np.array(out[:], np.float64)


def i0_series(x, terms):
    """Modified Bessel I0 as its power series sum_k (x/2)^(2k) / (k!)^2,
    truncated after `terms` terms."""
    return sum((x / 2.0) ** (2 * k) / math.factorial(k) ** 2 for k in range(terms))


def test_series_helper_is_scipy_i0():
    for x in (0.0, 0.5, 1.0, 4.0, 9.0):
        assert i0_series(x, 40) == pytest.approx(special.i0(x), rel=1e-12)


# ------------------------------------------------------------ Phi (:464-468)
@pytest.mark.parametrize("p", [0, 1, 2])
@pytest.mark.parametrize("eta_prime", [1.2, 1.55, 2.0])
def test_phi_at_the_axis_is_p_pi(lib, p, eta_prime):
    # h = 0: gamma_i = gamma_t = 0
    assert lib.oracle_phi(p, 0.0, eta_prime) == pytest.approx(p * math.pi, abs=1e-6)


@pytest.mark.parametrize("h", [-0.8, -0.3, 0.1, 0.6, 0.95])
def test_phi_without_refraction(lib, h):
    # eta' = 1: gamma_t = gamma_i, so Phi(p, h) = 2 (p - 1) asin(h) + p pi;
    # p = 0 is the mirror reflection -2 gamma_i, p = 1 passes straight through
    g = math.asin(h)
    for p in (0, 1, 2):
        assert lib.oracle_phi(p, h, 1.0) == pytest.approx(2 * (p - 1) * g + p * math.pi, abs=2e-6)


def test_phi_reflection_ignores_the_index(lib):
    for eta_prime in (1.1, 1.55, 3.0):
        assert lib.oracle_phi(0, 0.4, eta_prime) == pytest.approx(-2 * math.asin(0.4), abs=1e-6)


# ------------------------------------------------------------ A (:383-413)
def fresnel_normal(eta):
    return ((eta - 1.0) / (eta + 1.0)) ** 2


@pytest.mark.parametrize("eta", [1.3, 1.55, 2.0])
def test_attenuation_normal_incidence(lib, eta):
    """h = 0 and theta_d = 1 rad: A hands fresnel_reflectance the angle
    acos(cos 1) = 1 as its cosine, i.e. normal incidence: R0 = ((eta-1)/(eta+1))^2.
    Without absorption A_0 = R0, A_1 = (1 - R0)^2, A_2 = (1 - R0)^2 R0."""
    r0 = fresnel_normal(eta)
    zero = vec3((0.0, 0.0, 0.0))
    a = [call3(lib.oracle_a, p, 0.0, eta, eta, zero, 1.0) for p in range(3)]
    np.testing.assert_allclose(a[0], r0, rtol=1e-5)
    np.testing.assert_allclose(a[1], (1 - r0) ** 2, rtol=1e-5)
    np.testing.assert_allclose(a[2], (1 - r0) ** 2 * r0, rtol=1e-5)


def test_attenuation_grazing_argument(lib):
    """theta_d = 0, h = 0: the angle acos(1) = 0 read as a cosine is grazing
    incidence, where the Fresnel reflectance is 1: A_0 = 1, A_1 = A_2 = 0."""
    zero = vec3((0.0, 0.0, 0.0))
    np.testing.assert_allclose(call3(lib.oracle_a, 0, 0.0, 1.55, 1.55, zero, 0.0), 1.0, rtol=1e-6)
    for p in (1, 2):
        np.testing.assert_allclose(call3(lib.oracle_a, p, 0.0, 1.55, 1.55, zero, 0.0), 0.0, atol=1e-7)


def test_attenuation_absorption_per_path(lib):
    """h = 0: gamma_t = 0, so T = exp(-2 mu (1 + cos 0)) = exp(-4 mu); the R
    path carries T^2 (as written at :406), TT carries T, TRT carries T^2."""
    mu = np.array([0.1, 0.25, 0.4])
    zero = vec3((0.0, 0.0, 0.0))
    for p, power in ((0, 2), (1, 1), (2, 2)):
        clear = call3(lib.oracle_a, p, 0.0, 1.55, 1.7, zero, 1.0)
        dark = call3(lib.oracle_a, p, 0.0, 1.55, 1.7, vec3(mu), 1.0)
        np.testing.assert_allclose(dark / clear, np.exp(-4.0 * mu) ** power, rtol=2e-5)


def test_attenuation_even_in_h(lib):
    mu = vec3((0.2, 0.3, 0.5))
    for h in (0.15, 0.43, 0.87):
        for p in range(3):
            np.testing.assert_allclose(call3(lib.oracle_a, p, h, 1.55, 1.6, mu, 0.4),
                                       call3(lib.oracle_a, p, -h, 1.55, 1.6, mu, 0.4), rtol=1e-6)


# ------------------------------------------------------------ Mp (:470-523)
def test_mp_at_zero_angles_is_the_bessel_series(lib):
    # cos(0) cos(0) / v = 1 / v, sin terms 0; the WGSL sums 5 terms of I0
    for v in (1.0, 0.5, 0.25, 0.1225):
        assert lib.oracle_mp(v, 0.0, 0.0, 0.0) == pytest.approx(i0_series(1.0 / v, 5), rel=F32)
    assert lib.oracle_mp(1.0, 0.0, 0.0, 0.0) == pytest.approx(special.i0(1.0), rel=1e-6)


@pytest.mark.parametrize("theta", [0.2, 0.5, 1.0])
def test_mp_exponent_sign(lib, theta):
    """exp(sin(-theta_i) sin(theta_r) / v): negative for theta_r = theta_i,
    positive for theta_r = -theta_i (the specular cone)."""
    v = 0.5
    b = i0_series(math.cos(theta) ** 2 / v, 5)
    e = math.sin(theta) ** 2 / v
    assert lib.oracle_mp(v, theta, theta, 0.0) == pytest.approx(b * math.exp(-e), rel=F32)
    assert lib.oracle_mp(v, theta, -theta, 0.0) == pytest.approx(b * math.exp(e), rel=F32)


def test_mp_shift_moves_theta_r(lib):
    for ti, tr, s in ((0.3, -0.1, 0.1), (-0.4, 0.2, -0.05)):
        assert lib.oracle_mp(0.3, ti, tr, s) == lib.oracle_mp(0.3, ti, tr + s, 0.0)


def test_mp_one_sided_clamp(lib):
    """clamp(-pi/2 + 1e-6, pi/2 - 1e-6, x) = min(pi/2 - 1e-6, x): angles above
    pi/2 are clamped, angles below -pi/2 are not."""
    v = 0.4
    top = math.pi / 2 - 1e-6
    assert lib.oracle_mp(v, 2.0, 0.1, 0.0) == pytest.approx(lib.oracle_mp(v, top, 0.1, 0.0), rel=F32)
    want = i0_series(math.cos(-2.0) * math.cos(0.3) / v, 5) * math.exp(math.sin(2.0) * math.sin(0.3) / v)
    assert lib.oracle_mp(v, -2.0, 0.3, 0.0) == pytest.approx(want, rel=F32)


def test_mp_small_variance_branch(lib):
    tol = math.radians(0.1)
    assert lib.oracle_mp(1e-7, 0.2, -0.2, 0.0) == 0.0
    assert lib.oracle_mp(1e-7, 0.2, 0.5, 0.0) == pytest.approx(1e6 / tol, rel=1e-6)


# ------------------------------------------------------------ Np (:525-562)
LEG_H, LEG_W = np.polynomial.legendre.leggauss(10)


def np_inputs(theta_i, theta_r, eta, mu):
    """eta' and mu' of Np: theta_d = (theta_r - theta_i) / 2,
    eta' = sqrt(eta^2 - sin^2 theta_i) / cos theta_d, mu' = mu / cos theta_d."""
    td = (theta_r - theta_i) / 2
    ep = math.sqrt(eta * eta - math.sin(theta_i) ** 2) / math.cos(td)
    return td, ep, np.asarray(mu) / math.cos(td)


def test_np_wide_lobe_is_half_the_quadrature_over_2pi(lib):
    """beta = 5: the wrapped Gaussian is flat to 1e-5, D = 1/(2 pi), so
    Np = 1/2 sum_i w_i A(p, h_i) / (2 pi) at every phi."""
    ti, tr, eta, mu = 0.3, -0.2, 1.55, (0.2, 0.35, 0.6)
    td, ep, mup = np_inputs(ti, tr, eta, mu)
    for p in range(3):
        quad = sum(w * call3(lib.oracle_a, p, float(h), eta, ep, vec3(mup), td) for h, w in zip(LEG_H, LEG_W))
        want = 0.5 * quad / (2 * math.pi)
        for phi in (-2.0, 0.0, 1.3):
            np.testing.assert_allclose(call3(lib.oracle_np, p, phi, ti, tr, eta, vec3(mu), 5.0), want, rtol=2e-5)


@pytest.mark.parametrize("p", [0, 1, 2])
def test_np_narrow_lobe_peaks_at_phi(lib, p):
    """beta = 0.01: each quadrature node adds a narrow Gaussian at Phi(p, h_i);
    at phi = Phi(p, h_5) the node's own term dominates:
    Np ~ 1/2 w_5 A(p, h_5) / (sqrt(2 pi) beta).  Half a turn away (a Phi
    without its p pi) the lobe is far smaller."""
    ti, tr, eta, mu, beta = 0.1, 0.05, 1.55, (0.1, 0.2, 0.3), 0.01
    td, ep, mup = np_inputs(ti, tr, eta, mu)
    k = 5
    h = float(LEG_H[k])
    peak = lib.oracle_phi(p, h, ep)
    a = call3(lib.oracle_a, p, h, eta, ep, vec3(mup), td)
    want = 0.5 * LEG_W[k] * a / (math.sqrt(2 * math.pi) * beta)
    got = call3(lib.oracle_np, p, peak, ti, tr, eta, vec3(mu), beta)
    np.testing.assert_allclose(got, want, rtol=1e-3)
    off = call3(lib.oracle_np, p, peak - math.pi, ti, tr, eta, vec3(mu), beta)
    if p == 1:  # a pi shift lands between nodes only for odd p
        assert (off < 1e-3 * got).all()


# ------------------------------------------------------------ marschner (:565-605)
def material(**kw):
    m = np.zeros(1, abi.strand_material_dtype)
    for k, v in scene.default_material().items():
        m[k] = v
    for k, v in kw.items():
        m[k] = v
    return m


def angles():
    d = np.array([0.0, 0.0, 1.0])
    light = np.array([0.8, 0.1, 0.3])
    view = np.array([-0.2, 0.9, -0.25])
    return d, light / np.linalg.norm(light), view / np.linalg.norm(view)


def test_marschner_fully_occluded_is_the_dimmed_specular_lobe(lib):
    """occlusion 1: the scattering visibility is 0 (no Np lobes) and the direct
    lobe visibility mix(0.18, 1, 0) = 0.18, so
    BCSDF = 0.18 Mp(alpha^2, theta_i, theta_r, shift) specular.rgb specular.w."""
    d, light, view = angles()
    m = material()
    got = call3(lib.oracle_marschner, vec3(d), vec3(view), vec3(light), m.ctypes.data, 1.0)
    ti, tr = math.acos(d @ light), math.acos(d @ view)
    mp = lib.oracle_mp(float(m["alpha"][0]) ** 2, ti, tr, float(m["shift"][0]))
    spec = m["specular_color"][0]
    np.testing.assert_allclose(got, 0.18 * mp * spec[:3] * spec[3], rtol=1e-5)


def test_marschner_scattering_lobes_weighted_by_path(lib):
    """specular.w = 0, occlusion 0: BCSDF = sum_p Np(p) (p + 1) / 3 with
    sigma = 1 - absorption.rgb and beta_azim = sqrt(beta^2)."""
    d, light, view = angles()
    m = material(specular_color=(0.9, 0.72, 0.52, 0.0))
    got = call3(lib.oracle_marschner, vec3(d), vec3(view), vec3(light), m.ctypes.data, 0.0)
    ti, tr = math.acos(d @ light), math.acos(d @ view)
    lp = light - (light @ d) * d
    vp = view - (view @ d) * d
    lp, vp = lp / np.linalg.norm(lp), vp / np.linalg.norm(vp)
    phi = math.acos(lp @ vp) * np.sign(np.cross(lp, vp) @ d)
    sigma = 1.0 - m["absorption_color"][0][:3]
    want = sum(call3(lib.oracle_np, p, phi, ti, tr, float(m["eta"][0]), vec3(sigma), float(m["beta"][0]))
               * (p + 1) / 3.0 for p in range(3))
    np.testing.assert_allclose(got, want, rtol=1e-5)


# ------------------------------------------------------------ two strands, DOM and raster
PX = (20, 32)  # row, column of the pixel both strands cover fully


@pytest.mark.parametrize("layers", [4, 2])
def test_dom_layers_of_two_strands(layers):
    """Light view of the two-strand scene (strand_rasterizer.wgsl:527-711).
    Pass 1: z0 = the near strand's depth 0.8.  Pass 2 per strand, coverage 1:
    a = absorption.w / 2 (0.3 near, 0.35 far), u = (z0 - z) / z0 (0 near,
    0.375 far), t = u L split between slices floor(t) and floor(t) + 1 by
    fract(t); slice alphas composite by `over`, layers are their running
    `over`."""
    f = oracle.render(scene.config_two_strands(dom_layers=layers))
    depth = f.get("dom_depth").reshape(64, 64)
    op = f.get("dom_opacity").reshape(layers, 64, 64)
    assert depth[PX] == np.float32(0.8)
    a_near, a_far = 0.3, 0.35
    t = 0.375 * layers
    si, w = int(t), t - int(t)
    slices = np.zeros(layers)
    slices[0] = a_near
    for s, x in ((si, (1 - w) * a_far), (si + 1, w * a_far)):
        if s < layers:
            slices[s] = slices[s] + (1 - slices[s]) * x
    want = 1.0 - np.cumprod(1.0 - slices)
    np.testing.assert_allclose(op[:, PX[0], PX[1]], want, atol=2e-7)
    if layers == 4:
        np.testing.assert_allclose(want, [0.3, 0.4225, 0.5235625, 0.5235625])


def test_raster_blends_front_to_back():
    """Camera view with the light moved off the strands (visibility 1): the
    near strand (0.9, 0.1, 0.1, alpha 0.6) over the far one (0.1, 0.2, 0.9,
    alpha 0.7), premultiplied, front to back; the stored colour is
    unpremultiplied (strand_rasterizer.wgsl:156-162, 1036-1040) and the depth
    is the nearest fragment's."""
    f = oracle.render(scene.config_two_strands(light_offset=4.0))
    img = f.get("image").reshape(64, 64, 4)[PX]
    near, far = np.array([0.9, 0.1, 0.1, 0.6]), np.array([0.1, 0.2, 0.9, 0.7])
    rgb = near[:3] * near[3] + far[:3] * far[3] * (1 - near[3])
    a = near[3] + far[3] * (1 - near[3])
    np.testing.assert_allclose(img, np.r_[rgb / a, a], atol=2e-7)
    assert f.get("depth").reshape(64, 64)[PX] == np.float32(0.8)
    # back-to-front would give a different colour
    back = far[:3] * far[3] + near[:3] * near[3] * (1 - far[3])
    assert np.abs(back / a - img[:3]).max() > 0.05
