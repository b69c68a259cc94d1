"""Pins of the oracle's unified phase: granular solid particles and
position-based friction (Macklin et al. 2014, eq. 24-25; beyond the reference,
which has neither).

Friction is checked through what it must do to a contact pair, not by
restating its formula: static friction leaves the pair's relative displacement
in the step purely normal; kinetic friction shortens the tangential part by
exactly mu_k times the penetration; both leave the pair's centre unchanged.
The phase masks are checked against closed forms (a solid has lambda 0; a fluid
particle whose neighbours are all solid has density W(0))."""
import numpy as np
import pytest

from conftest import small_dam


def _pair_params(F, mu_s=0.5, mu_k=0.3, fluid_terms=False):
    p = small_dam(F, 4)
    p.gravity = 0.0
    p.solver_iterations = 1
    p.stabilization_iterations = 1
    p.friction_static = mu_s
    p.friction_kinetic = mu_k
    p.active_update_positions = 0
    for k in ("active_compute_lambda", "active_compute_delta_positions", "active_compute_omegas",
              "active_compute_vorticity", "active_compute_viscosity"):
        setattr(p, k, 1 if fluid_terms else 0)
    return p


def _pair_step(F, O, p, slip, phases, gap=0.04):
    """Particle A at rest, B at distance `gap` along x, moving `slip` along y in the step."""
    d = O.derive(p)
    x = np.array([[0.55, 0.55, 0.55, 1.0], [0.55 + gap, 0.55, 0.55, 1.0]], np.float32)
    v = np.zeros((2, 4), np.float32)
    v[1, 1] = np.float32(slip / d.dt)
    st = O.State(x, v, phase=np.array(phases, np.uint32))
    # the predicted positions as the oracle's FP32 predict forms them (g = 0): p = x + v dt
    pstart = (x[:, :3] + (v[:, :3] * np.float32(d.dt_f)).astype(np.float32)).astype(np.float32).astype(np.float64)
    out = O.step(p, st)
    order = np.argsort(out["x4"][:, 0])  # A is the one with the smaller x
    P = out["p4"][order, :3]
    X = out["x4"][order, :3]
    u0 = (pstart[1] - x[1, :3].astype(np.float64)) - (pstart[0] - x[0, :3].astype(np.float64))
    return P, X, pstart, d, u0


def _slip_after(P, X):
    return (P[1] - X[1]) - (P[0] - X[0])


def test_static_friction_removes_tangential_slip(F, O):
    p = _pair_params(F)
    P, X, p0, d, _ = _pair_step(F, O, p, slip=0.002, phases=[1, 1])
    n = (p0[1] - p0[0]) / np.linalg.norm(p0[1] - p0[0])
    u = _slip_after(P, X)
    pen = d.diameter - np.linalg.norm(p0[1] - p0[0])
    assert 0.002 < 0.5 * pen  # below the static threshold
    assert np.linalg.norm(u - np.dot(u, n) * n) < 1e-9  # purely normal
    assert np.allclose(P.mean(axis=0), p0.mean(axis=0) + (X.mean(axis=0) - np.array([0.57, 0.55, 0.55])), atol=1e-9)


def test_kinetic_friction_shortens_slip_by_mu_k_times_penetration(F, O):
    p = _pair_params(F, mu_s=0.5, mu_k=0.3)
    slip = 0.02
    P, X, p0, d, u0 = _pair_step(F, O, p, slip=slip, phases=[1, 0])  # one solid: friction applies
    r = p0[1] - p0[0]
    n = r / np.linalg.norm(r)
    pen = d.diameter - np.linalg.norm(r)
    t0 = u0 - np.dot(u0, n) * n
    assert np.linalg.norm(t0) > 0.5 * pen  # above the static threshold
    u = _slip_after(P, X)
    t = u - np.dot(u, n) * n
    assert abs(np.linalg.norm(t) - (np.linalg.norm(t0) - 0.3 * pen)) < 1e-9
    assert np.dot(t, t0) > 0  # shortened, not reversed
    # friction is internal: the pair's centre moves only by the (also internal) pushes, i.e. not at all
    assert np.allclose(P.sum(axis=0), p0.sum(axis=0), atol=1e-9)


def test_no_friction_between_fluid_particles(F, O):
    p = _pair_params(F)
    P, X, p0, d, _ = _pair_step(F, O, p, slip=0.002, phases=[0, 0])
    u = _slip_after(P, X)
    assert abs(u[1] - 0.002) < 1e-6  # tangential slip untouched (the push moves both along x)


def test_phase_masks_density_and_lambda(F, O):
    """A fluid particle surrounded by solid ones: density W(0) (solids do not count)
    and lambda from C = W(0)/rho0 - 1; every solid particle has lambda 0."""
    p = small_dam(F, 4)
    p.gravity = 0.0
    p.box_jitter = 0.0
    pos, vel, col = F.generate_box(p)
    ph = np.ones(pos.shape[0], np.uint32)
    centre = np.argmin(np.linalg.norm(pos[:, :3] - pos[:, :3].mean(axis=0), axis=1))
    ph[centre] = 0
    st = O.State(pos, vel, col, phase=ph)
    out = O.step(p, st)
    fluid = np.where(st.phase == 0)[0]
    assert fluid.size == 1
    d = O.derive(p)
    w0 = O.w(d, 0.0)
    assert out["density"][fluid[0]] == pytest.approx(w0, rel=1e-12)
    assert np.all(out["lambda"][st.phase == 1] == 0.0)
    assert out["lambda"][fluid[0]] == pytest.approx(-(w0 / d.rho0 - 1.0) / d.eps, rel=1e-12)


def test_all_fluid_phases_equal_the_plain_step(F, O):
    p = small_dam(F, 6)
    p.friction_static = 0.5
    p.friction_kinetic = 0.3
    pos, vel, col = F.generate_box(p)
    a = O.State(pos, vel, col)
    b = O.State(pos, vel, col, phase=np.zeros(pos.shape[0], np.uint32))
    for _ in range(3):
        oa = O.step(p, a)
        ob = O.step(p, b)
    for k in ("x4", "v4", "p4", "f4", "density"):
        assert np.array_equal(oa[k], ob[k]), k


def test_generate_phases(F):
    p = small_dam(F, 16)
    p.box_solid_fraction = 0.3
    p.box_solid_radius = 0.1
    ph = F.capi.generate_phases(p)
    frac = ph.mean()
    assert 0.3 <= frac < 0.3 + 33 / ph.size + 1e-12  # stops at the first sphere reaching the share
    assert np.array_equal(ph, F.capi.generate_phases(p))  # seeded
    # spheres of radius 2 spacings: every solid lattice point has a solid lattice neighbour
    g = ph.reshape(16, 16, 16).astype(bool)
    nb = np.zeros_like(g)
    for ax in range(3):
        nb |= np.roll(g, 1, ax) | np.roll(g, -1, ax)
    assert np.all(nb[g])
    p.box_solid_fraction = 0.0
    assert not F.capi.generate_phases(p).any()
