"""Pins the oracle's Windkessel iolets (InOutLetWK2 / InOutLetWK3 coupled
through the Nash delegate, src/lb/iolets/InOutLetWK3.cc:33-106,
NashZerothOrderPressureDelegate.h:36-116) by what is fixed independently of
its code:

  * the flow rate after a step is area / n * sum over the coupling sites of
    u . (-n), with u the velocity field the same step reports and the
    coupling sites read from the geometry (sites with an outlet link along
    the lattice direction closest to -n);
  * the outlet pressure P = rho c_s^2 follows the discretisations of the
    two- and three-element Windkessel ODEs,
        C dP/dt = Q - P/R                       (backward Euler)
        P + Rp Cp dP/dt = (Rc+Rp) Q + Rc Rp Cp dQ/dt
    with Q^n extrapolated as 2 Q^{n-1} - Q^{n-2} (InOutLetWK3.cc:78-89);
  * at steady state P = R Q (WK2) and P = (Rc + Rp) Q (WK3).
"""
import numpy as np
import pytest

import hemepure_amd as H
from fixtures import CX, CY, CZ, INV, windkessel_cylinder_case

CS2 = 1.0 / 3.0
OUTLET_TYPE = 3


def coupling_sites(m, normal):
    """Sites whose outlet link points along -(closest lattice direction to n)."""
    c = np.stack([CX, CY, CZ], axis=1).astype(np.float64)
    mag = np.linalg.norm(c, axis=1)
    mag[0] = 1.0
    closest = int(np.argmin(((c / mag[:, None]) - np.asarray(normal)) ** 2 @ np.ones(3)))
    link = INV[closest]
    _, iol, _, types = m.lat.site_masks()
    return np.nonzero(((iol >> (link - 1)) & 1).astype(bool) & (types == OUTLET_TYPE))[0]


@pytest.mark.parametrize("kind", [H.WK2, H.WK3])
def test_flow_rate_is_the_coupling_site_flux(kind):
    case = windkessel_cylinder_case(kind=kind, length=24)
    out = case.outlets[0]
    m = case.oracle()
    m.init_equilibrium(1.0)
    m.run(40, properties=True)
    _, u, _ = m.fields()
    sites = coupling_sites(m, out.normal)
    assert sites.size > 50
    m.run(1)  # DoComms at the start of this step reduces the previous step's flux
    expected = out.area / sites.size * np.sum(u[sites] @ -np.asarray(out.normal))
    q = m.iolet_state(True, 0)[2]
    assert abs(q) > 1e-3
    assert abs(q - expected) <= 1e-12 * abs(expected)


@pytest.mark.parametrize("kind", [H.WK2, H.WK3])
def test_pressure_follows_the_windkessel_discretisation(kind):
    case = windkessel_cylinder_case(kind=kind, length=24)
    out = case.outlets[0]
    m = case.oracle()
    m.init_equilibrium(1.0)
    m.run(30)
    states = []
    for _ in range(12):
        m.run(1)
        states.append(m.iolet_state(True, 0))
    states = np.array(states)
    # After each step the centre site has committed: density == densityNew, flowRateOld == flowRate.
    np.testing.assert_array_equal(states[:, 0], states[:, 1])
    np.testing.assert_array_equal(states[:, 2], states[:, 3])
    P = states[:, 0] * CS2
    Q = states[:, 2]
    assert np.ptp(P) > 1e-6 and np.ptp(Q) > 1e-4
    for m_ in range(1, len(states)):
        p0, p1, q0, q1 = P[m_ - 1], P[m_], Q[m_ - 1], Q[m_]
        if kind == H.WK2:
            R, Cp = out.resistance, out.capacitance
            resid = Cp * (p1 - p0) - (q1 - p1 / R)
            scale = Cp * abs(p1) + abs(q1) + abs(p1 / R)
        else:
            Rc, Rp, Cp = out.r_c, out.r_p, out.capacitance
            resid = p1 + Rp * Cp * (p1 - p0) - ((Rc + Rp) * (2 * q1 - q0) + Rc * Rp * Cp * (q1 - q0))
            scale = abs(p1) + Rp * Cp * abs(p1) + (Rc + Rp) * (abs(q1) + abs(q0)) * 3 + Rc * Rp * Cp * abs(q1)
        assert abs(resid) <= 1e-12 * scale, (m_, resid)


@pytest.mark.slow
@pytest.mark.parametrize("kind", [H.WK2, H.WK3])
def test_steady_state_pressure_is_resistance_times_flow(kind):
    case = windkessel_cylinder_case(kind=kind)
    out = case.outlets[0]
    m = case.oracle()
    m.init_equilibrium(1.0)
    m.run(1400, threads=8)
    p_prev = m.iolet_state(True, 0)[0] * CS2
    m.run(100, threads=8)
    st = m.iolet_state(True, 0)
    P, Q = st[0] * CS2, st[2]
    assert abs(P - p_prev) < 1e-5 * P
    R = out.resistance if kind == H.WK2 else out.r_c + out.r_p
    assert abs(P - R * Q) < 1e-4 * P
    assert m.stable()


def test_windkessel_needs_nash():
    case = windkessel_cylinder_case(kind=H.WK3)
    case.outlet_bc = H.IOLET_LADD
    with pytest.raises(H.HlbError, match="Nash"):
        case.product()


def test_windkessel_state_before_any_step():
    case = windkessel_cylinder_case(kind=H.WK2)
    lat = case.product()
    np.testing.assert_array_equal(lat.iolet_state(True, 0), [1.0, 1.0, 0.0, 0.0])
    with pytest.raises(H.HlbError, match="no such iolet"):
        lat.iolet_state(True, 1)
