"""The FP32-shifted storage form against the FP64 oracle.

Storage and bulk arithmetic are float (g = f - w rho0); iolet values,
Windkessel sums and the output fields are double.  The bar (stated per
quantity) is the error relative to the field's largest magnitude:
    max |x32 - x64| / max |x64|  <= 1e-4   for velocity and WSS,
and for density the same measure taken on rho - 1.  Each test prints the
errors it measured.
"""
import numpy as np
import pytest

import hemepure_amd as H
from fixtures import (Case, IoletSpec, anisotropic_f, bifurcation_case, cylinder_case, four_cube,
                      windkessel_cylinder_case)

pytestmark = pytest.mark.gpu

TOL = 1e-4


def rel_err(got, want, mask=None):
    got = np.asarray(got, dtype=np.float64)
    want = np.asarray(want, dtype=np.float64)
    if mask is not None:
        got, want = got[mask], want[mask]
    scale = np.abs(want).max()
    return float(np.abs(got - want).max() / scale) if scale > 0 else float(np.abs(got - want).max())


def run_pair(case: Case, steps: int, f0=None, t0=1):
    lat = case.product(precision=H.FP32_SHIFTED)
    ref = case.oracle()
    if f0 is None:
        lat.init_equilibrium(1.0, t0)
        ref.init_equilibrium(1.0, t0)
    else:
        lat.upload_f(f0, t0)
        ref.set_f(f0, t0)
    lat.step(steps, properties=True)
    ref.run(steps, properties=True, threads=0)
    return lat, ref


def field_errors(lat, ref, tag):
    rho, u, wss = lat.fields()
    orho, ou, owss = ref.fields()
    wall = owss < 1e300
    e = dict(rho=rel_err(rho - 1.0, orho - 1.0), u=rel_err(u, ou),
             wss=rel_err(wss, owss, wall) if wall.any() else 0.0)
    np.testing.assert_array_equal(wss >= 1e300, ~wall)
    print(f"[fp32] {tag}: " + ", ".join(f"{k} {v:.2e}" for k, v in e.items()))
    return e


def test_cylinder_2000_steps():
    """The straight cylinder (r 10, L 100, LBGK tau 0.8, SBB, Nash 1.001 / 0.999)."""
    case = cylinder_case()
    lat, ref = run_pair(case, 2000)
    e = field_errors(lat, ref, "cylinder 2000")
    assert e["u"] <= TOL and e["wss"] <= TOL and e["rho"] <= TOL
    f32, f64 = lat.download_f(), ref.f()
    assert rel_err(f32 - H_W, f64 - H_W) <= TOL
    assert lat.stable()


H_W = np.array([1 / 3] + [1 / 18] * 6 + [1 / 36] * 12)


@pytest.mark.parametrize("kernel", [H.KERNEL_LBGK, H.KERNEL_TRT, H.KERNEL_LES])
@pytest.mark.parametrize("wall_bc", [H.WALL_SBB, H.WALL_BFL])
def test_four_cube_corners(kernel, wall_bc):
    g = four_cube(8, seed=3)
    case = Case(g, kernel=kernel, wall_bc=wall_bc, tau=0.7,
                inlets=[IoletSpec(H.COSINE, (3.5, 3.5, 0.5), (0, 0, 1), density_mean=1.01, density_amp=0.002,
                                  period=7.0)],
                outlets=[IoletSpec(H.COSINE, (3.5, 3.5, 6.5), (0, 0, -1), density_mean=0.99)])
    f0 = np.tile(H_W, (g.n_sites, 1)) * (1.0 + 0.02 * (anisotropic_f(g.n_sites) - 1.0))
    lat, ref = run_pair(case, 7, f0=f0, t0=3)
    e = field_errors(lat, ref, f"four_cube k{kernel} w{wall_bc}")
    assert max(e.values()) <= 1e-5


def test_bifurcation_small_trt_bfl_ladd():
    case = bifurcation_case(rp=8.0, lp=40.0, rd=6.5, ld=32.0, steps=400, peak=0.03, period=200.0)
    lat, ref = run_pair(case, 400)
    e = field_errors(lat, ref, "ybif small 400")
    assert e["u"] <= TOL and e["wss"] <= TOL and e["rho"] <= TOL


def test_windkessel_outlet():
    case = windkessel_cylinder_case(kind=H.WK3, length=24)
    lat, ref = run_pair(case, 600)
    e = field_errors(lat, ref, "wk3 600")
    st, ost = lat.iolet_state(True, 0), ref.iolet_state(True, 0)
    e_p = abs(st[0] - ost[0]) / abs(ost[0] - 1.0)
    e_q = abs(st[2] - ost[2]) / abs(ost[2])
    print(f"[fp32] wk3 600: density-1 {e_p:.2e}, flow rate {e_q:.2e}")
    assert e["u"] <= TOL and e_p <= TOL and e_q <= TOL


def test_upload_download_and_restart():
    case = cylinder_case(radius=6.0, length=32, box_xy=16)
    rng = np.random.default_rng(9)
    lat = case.product(precision=H.FP32_SHIFTED)
    f0 = np.tile(H_W, (lat.n, 1)) * (1.0 + 0.01 * rng.uniform(-1, 1, (lat.n, 19)))
    lat.upload_f(f0, 5)
    back = lat.download_f()
    # float of g = f - w: the error is the rounding of g plus of the sum
    np.testing.assert_allclose(back, f0, rtol=0, atol=1e-8)
    lat.step(30)
    mid = lat.download_f()
    lat.step(30, properties=True)
    rho_a, u_a, _ = lat.fields()
    again = case.product(precision=H.FP32_SHIFTED)
    again.upload_f(mid, 35)
    again.step(30, properties=True)
    rho_b, u_b, _ = again.fields()
    assert rel_err(u_b, u_a) <= 1e-5 and rel_err(rho_b - 1, rho_a - 1) <= 1e-5
    assert again.timestep() == lat.timestep() == 65


def test_stability_flag():
    case = cylinder_case(radius=6.0, length=16, box_xy=16)
    lat = case.product(precision=H.FP32_SHIFTED)
    f = np.tile(H_W, (lat.n, 1))
    lat.upload_f(f)
    assert lat.stable()
    f[lat.n // 2, 4] = -0.01
    lat.upload_f(f)
    assert not lat.stable()
