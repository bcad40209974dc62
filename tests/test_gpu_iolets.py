"""The remaining iolet providers on the GPU against the oracle: Womersley
velocity (Ladd, and behind Guo-Zheng-Shi walls), file pressure (Nash) with the
table built from the reference's own profile file, and the FP32-shifted form
with a Womersley inlet.  Bar: 1e-12 relative for FP64 (bit-identical where the
inputs are), 1e-4 for the FP32-shifted form."""
from pathlib import Path

import numpy as np
import pytest

import hemepure_amd as H
from fixtures import Case, IoletSpec
from test_gpu_parity import compare_all, run_both

pytestmark = pytest.mark.gpu

GOLDEN = Path(__file__).resolve().parent / "golden"


def womersley_case(kernel=H.KERNEL_LBGK, wall_bc=H.WALL_SBB, alpha=3.0, period=120.0, pga=2e-5):
    g = H.Geometry.cylinder(8.0, 48, 24)
    i0, o0 = g.inlets()[0], g.outlets()[0]
    return Case(g, kernel=kernel, wall_bc=wall_bc, inlet_bc=H.IOLET_LADD, outlet_bc=H.IOLET_NASH, tau=0.8,
                inlets=[IoletSpec(H.WOMERSLEY, i0["position"], i0["normal"], radius=8.5, period=period,
                                  pressure_gradient_amplitude=pga, womersley_number=alpha)],
                outlets=[IoletSpec(H.COSINE, o0["position"], o0["normal"], density_mean=1.0)])


@pytest.mark.parametrize("kernel,wall_bc", [(H.KERNEL_LBGK, H.WALL_SBB), (H.KERNEL_TRT, H.WALL_BFL),
                                            (H.KERNEL_LBGK, H.WALL_GZS)])
def test_womersley_inlet_matches_oracle(kernel, wall_bc):
    """1e-12 relative, not bit for bit: the profile's complex arithmetic
    (Bessel series, complex division and product) runs in host code that the
    product and the oracle build separately, and the compiler may round that
    expression differently at different optimisation levels (the reference
    builds at -O3).  The kernels' own arithmetic is exact as elsewhere."""
    case = womersley_case(kernel, wall_bc)
    lat, ref = run_both(case, 300)
    compare_all(lat, ref, f"womersley {kernel}/{wall_bc}", exact=False)
    _, u, _ = lat.fields()
    assert np.abs(u[:, 2]).max() > 1e-4  # the inlet drives a flow


def test_file_pressure_inlet_matches_oracle():
    """iolet.txt (78 -> 82 -> 78 mmHg over 4 s) as a Nash inlet; the step and
    voxel are chosen so the 4 s profile spans 400 steps and the pressure
    swing is a lattice density swing of order 1e-3."""
    total, dt, dx = 400, 0.01, 0.4
    table, dmin = H.iolet_file_table(GOLDEN / "iolet_pressure.txt", False, total, dt, dx)
    assert 1.0 < table.min() and table.max() - table.min() > 1e-4
    g = H.Geometry.cylinder(8.0, 48, 24)
    i0, o0 = g.inlets()[0], g.outlets()[0]
    rho_out = float(table.min())
    case = Case(g, tau=0.8,
                inlets=[IoletSpec(H.FILEPRES, i0["position"], i0["normal"], table=table, density_min=dmin)],
                outlets=[IoletSpec(H.COSINE, o0["position"], o0["normal"], density_mean=rho_out)])
    lat, ref = run_both(case, total)
    compare_all(lat, ref, "file pressure")
    assert lat.iolet_state(False, 0)[0] == ref.iolet_state(False, 0)[0]


def test_womersley_fp32_close_to_oracle():
    from test_gpu_fp32 import field_errors

    case = womersley_case(kernel=H.KERNEL_TRT, wall_bc=H.WALL_BFL)
    lat = case.product(precision=H.FP32_SHIFTED)
    ref = case.oracle()
    lat.init_equilibrium(1.0)
    ref.init_equilibrium(1.0)
    lat.step(300, properties=True)
    ref.run(300, properties=True, threads=0)
    e = field_errors(lat, ref, "womersley fp32")
    assert e["u"] <= 1e-4 and e["rho"] <= 1e-4


def test_small_vessel_tree_les_windkessel():
    """The vessel-tree configuration in miniature (7 tubes, root r 8, box
    96^3): LES around tau 0.505, SBB, cosine inlet, four WK3 outlets; FP64
    bit-exact against the oracle, FP32-shifted within 1e-4."""
    from fixtures import vessel_tree_case
    from test_gpu_fp32 import field_errors

    case = vessel_tree_case(generations=3, root_radius=8.0, box=(96, 96, 96), threads=4, period=200.0)
    lat, ref = run_both(case, 400)
    compare_all(lat, ref, "vessel tree")
    for k in range(len(case.outlets)):
        assert np.array_equal(lat.iolet_state(True, k), ref.iolet_state(True, k))
    lat32 = case.product(precision=H.FP32_SHIFTED)
    lat32.init_equilibrium(1.0)
    lat32.step(400, properties=True)
    e = field_errors(lat32, ref, "vessel tree fp32")
    assert e["u"] <= 1e-4 and e["rho"] <= 1e-4
