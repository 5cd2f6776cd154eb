"""The oracle's BSDFs, emitters and path integrator pinned by the
mathematics (tests/pins.py): each check passes on the oracle, and each
fails when a plausible mistake is planted in the oracle.

Reference functions covered: Diffuse (src/tracer/bsdf/diffuse.cpp:22-67),
Mirror (mirror.cpp:12-34), Dielectric (dielectric.cpp:40-58, fresnel /
refract src/core/math/math.cpp:8-47), Microfacet (microfacet.cpp:35-104),
the anisotropic GGX of the GPU path (src/optix/cuda/material/aniso_ggx.cu:
38-121), the conductor extension, AreaLight (src/tracer/emitter/area.cpp:
11-49, mesh.cpp:222-246), the point light (include/optix/light/point.h:
15-31), the environment map (include/optix/light/envmap.h:25-94), the
emitter pick (src/tracer/scene/scene.cpp:139-179) and PathIntegrator::Li
(src/tracer/integrator/path.cpp:13-108)."""
import numpy as np
import pytest

import oracle
import pins

D, MI, DI, MF, AG, CO = (oracle.BSDF_DIFFUSE, oracle.BSDF_MIRROR, oracle.BSDF_DIELECTRIC, oracle.BSDF_MICROFACET,
                         oracle.BSDF_ANISO_GGX, oracle.BSDF_CONDUCTOR)

# (type, params) of the non-delta BSDFs, with the configurations the scenes
# use: the 0.2-roughness conductor, the plates' GGX roughnesses (pd = 0,
# including the anisotropic (0.1, 0.01) one), Beckmann with the default and a
# zero diffuse part
SMOOTH = {
    "diffuse": (D, (0.5, 0.6, 0.7)),
    "microfacet_default": (MF, (0.1, 1.5046, 1.000277, 0.5, 0.5, 0.5)),
    "microfacet_kd0": (MF, (0.3, 1.5, 1.0, 0.0, 0.0, 0.0)),
    "microfacet_rough": (MF, (0.6, 1.7, 1.2, 0.2, 0.3, 0.1)),
    "aniso_pd0_0.1_0.01": (AG, (0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.1, 0.01)),
    "aniso_pd0_0.02": (AG, (0.0, 0.0, 0.0, 1.0, 1.0, 1.0, 0.02, 0.02)),
    "aniso_rough": (AG, (0.3, 0.2, 0.1, 0.6, 0.6, 0.6, 0.5, 0.3)),
    "conductor_0.2": (CO, (0.2, 0.2, 0.92, 1.1, 3.912, 2.452, 2.142)),
    "conductor_0.6": (CO, (0.6, 1.5, 0.4, 0.9, 2.0, 3.0, 4.0)),
}


@pytest.mark.parametrize("name", list(SMOOTH))
@pytest.mark.parametrize("theta", [12.0, 48.0, 77.0])
def test_bsdf_integrals(name, theta):
    btype, params = SMOOTH[name]
    pin = pins.bsdf_integrals(btype, params, pins.local_dir(theta, 35.0))
    print(pin.msg)
    assert pin, pin.msg


@pytest.mark.parametrize("name", ["diffuse", "microfacet_default", "microfacet_rough", "aniso_rough",
                                  "aniso_pd0_0.1_0.01", "conductor_0.2"])
def test_bsdf_reciprocity(name):
    pin = pins.reciprocity(*SMOOTH[name])
    assert pin, pin.msg


@pytest.mark.parametrize("alpha,eta,k", [(0.2, (0.2, 0.92, 1.1), (3.912, 2.452, 2.142)),
                                         (0.5, (1.5, 0.4, 0.9), (2.0, 3.0, 4.0))])
def test_conductor_matches_complex_fresnel(alpha, eta, k):
    pin = pins.conductor_closed_form(alpha, eta, k)
    assert pin, pin.msg


def test_conductor_fresnel_function_matches_complex_fresnel():
    cos = np.linspace(0.01, 1.0, 200)
    for eta, k in [(0.2, 3.912), (0.92, 2.452), (1.1, 2.142), (1.5, 0.0), (2.5, 7.0)]:
        inp = np.zeros((len(cos), 6), np.float32)
        inp[:, 0], inp[:, 1], inp[:, 2] = cos, eta, k
        got = oracle.optics(3, inp)[:, 0]
        ref = pins.fresnel_complex(cos.astype(np.float32), eta, k)
        assert np.abs(got - ref).max() < 2e-6 * 1e1, (eta, k, np.abs(got - ref).max())
    # normal incidence ((n - 1)^2 + k^2) / ((n + 1)^2 + k^2)
    inp = np.array([[1.0, 0.2, 3.912, 0, 0, 0]], np.float32)
    assert abs(oracle.optics(3, inp)[0, 0] - ((0.2 - 1) ** 2 + 3.912 ** 2) / ((1.2) ** 2 + 3.912 ** 2)) < 1e-6


@pytest.mark.parametrize("pd,ps,ax,ay", [((0.0,) * 3, (1.0,) * 3, 0.1, 0.01), ((0.3, 0.2, 0.1), (0.6, 0.5, 0.4), 0.5,
                                                                             0.3)])
def test_aniso_ggx_matches_closed_form(pd, ps, ax, ay):
    pin = pins.aniso_ggx_closed_form(pd, ps, ax, ay)
    assert pin, pin.msg


@pytest.mark.parametrize("alpha,int_ior,ext_ior,kd", [(0.1, 1.5046, 1.000277, (0.5, 0.5, 0.5)),
                                                      (0.3, 1.5, 1.0, (0.0, 0.0, 0.0)),
                                                      (0.6, 1.7, 1.2, (0.2, 0.3, 0.1))])
def test_microfacet_matches_closed_form(alpha, int_ior, ext_ior, kd):
    pin = pins.microfacet_closed_form(alpha, int_ior, ext_ior, kd)
    assert pin, pin.msg


@pytest.mark.parametrize("int_ior,ext_ior", [(1.5, 1.0), (1.5046, 1.000277), (1.33, 1.7)])
def test_dielectric_fresnel_snell_tir(int_ior, ext_ior):
    pin = pins.dielectric_optics(int_ior, ext_ior)
    assert pin, pin.msg


def test_mirror():
    pin = pins.mirror_optics()
    assert pin, pin.msg


def test_lambert_irradiance_formula_by_area_quadrature():
    """The polygon irradiance used as the area light's reference, checked by
    direct quadrature over the polygon's area."""
    _, V, _, _ = pins.area_light_scene()
    p = np.array([0.15, 0.0, 0.1])
    n = np.array([0.2, 1.0, -0.1])
    n /= np.linalg.norm(n)
    a = pins.polygon_irradiance(p, n, V[[0, 1, 2, 3]])
    b = pins.polygon_irradiance_quadrature(p, n, V[[0, 1, 2, 3]])
    assert abs(a - b) < 1e-4 * a, (a, b)


def test_area_light_sampling():
    pin = pins.area_light_sampling()
    assert pin, pin.msg


def test_emitter_pick_and_point_lights():
    pin = pins.multi_emitter_pick()
    assert pin, pin.msg


def test_point_light_irradiance():
    pin = pins.point_light_irradiance()
    assert pin, pin.msg


def test_env_map_sampling():
    pin = pins.env_sampling()
    assert pin, pin.msg


@pytest.mark.parametrize("max_depth", [1, 2, 4, -1])
def test_white_furnace_box(max_depth):
    pin = pins.furnace_box(max_depth)
    assert pin, pin.msg


@pytest.mark.parametrize("btype,params,expect", [(D, (0.5, 0.3, 0.8), (0.5, 0.3, 0.8)), (MI, (), (1, 1, 1)),
                                                 (DI, (1.5, 1.0), (1, 1, 1))])
def test_white_furnace_sphere_under_constant_sky(btype, params, expect):
    pin = pins.furnace_sphere(btype, params, expect)
    assert pin, pin.msg


def test_point_lit_plane_closed_form():
    pin = pins.point_lit_plane()
    assert pin, pin.msg


# ---------------------------------------------------------------------------
# planted mistakes: each pin above must catch at least the mistakes below
# ---------------------------------------------------------------------------
PLANTED = {
    "ggx_pdf_without_wh_z/aniso": (oracle.MUT_GGX_PDF_NO_WHZ, lambda: pins.bsdf_integrals(
        *SMOOTH["aniso_rough"], pins.local_dir(48.0, 35.0))),
    "ggx_pdf_without_wh_z/conductor": (oracle.MUT_GGX_PDF_NO_WHZ, lambda: pins.bsdf_integrals(
        *SMOOTH["conductor_0.6"], pins.local_dir(48.0, 35.0))),
    "dielectric_int_ext_swapped": (oracle.MUT_DIELECTRIC_SWAP_IOR, lambda: pins.dielectric_optics(1.5, 1.0)),
    "dielectric_without_fresnel": (oracle.MUT_DIELECTRIC_NO_FRESNEL, lambda: pins.dielectric_optics(1.5, 1.0)),
    "diffuse_uniform_pdf": (oracle.MUT_DIFFUSE_UNIFORM_PDF, lambda: pins.bsdf_integrals(
        *SMOOTH["diffuse"], pins.local_dir(48.0, 35.0))),
    "emitter_hit_without_mis/box": (oracle.MUT_EMITTER_HIT_NO_MIS, lambda: pins.furnace_box(4)),
    "emitter_hit_without_mis/sky": (oracle.MUT_EMITTER_HIT_NO_MIS, lambda: pins.furnace_sphere(
        D, (0.5, 0.3, 0.8), (0.5, 0.3, 0.8))),
    "area_pdf_without_dist2/emitter": (oracle.MUT_AREA_PDF_NO_DIST2, pins.area_light_sampling),
    "area_pdf_without_dist2/furnace": (oracle.MUT_AREA_PDF_NO_DIST2, lambda: pins.furnace_box(4)),
    "conductor_eta_k_swapped": (oracle.MUT_CONDUCTOR_SWAP_ETA_K, lambda: pins.conductor_closed_form(
        0.2, (0.2, 0.92, 1.1), (3.912, 2.452, 2.142))),
    "rr_without_divide": (oracle.MUT_RR_NO_DIVIDE, lambda: pins.furnace_box(-1)),
    "microfacet_g_of_wi_twice/reciprocity": (oracle.MUT_MICROFACET_G_WI_TWICE, lambda: pins.reciprocity(
        *SMOOTH["microfacet_rough"])),
    "microfacet_g_of_wi_twice/closed_form": (oracle.MUT_MICROFACET_G_WI_TWICE, lambda: pins.microfacet_closed_form(
        0.6, 1.7, 1.2, (0.2, 0.3, 0.1))),
    "point_light_1_over_d/irradiance": (oracle.MUT_POINT_INV_DIST, pins.point_light_irradiance),
    "point_light_1_over_d/plane": (oracle.MUT_POINT_INV_DIST, pins.point_lit_plane),
    "env_pdf_times_area/sampling": (oracle.MUT_ENV_PDF_TIMES_AREA, pins.env_sampling),
    "env_pdf_times_area/furnace": (oracle.MUT_ENV_PDF_TIMES_AREA, lambda: pins.furnace_sphere(
        D, (0.5, 0.3, 0.8), (0.5, 0.3, 0.8))),
}


@pytest.mark.parametrize("name", list(PLANTED))
def test_planted_mistake_is_caught(name):
    m, check = PLANTED[name]
    with oracle.mutation(m):
        pin = check()
    print(name, "->", pin.msg)
    assert not pin.ok, "planted mistake %s passed: %s" % (name, pin.msg)
