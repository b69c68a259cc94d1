"""Pins of the oracle's FP64 reverse mode (oracle/pbf_adjoint.c).

orc64_gradient is checked against FP64 central finite differences of
orc64_forward (the same map), on horizons where the differences converge, and
against closed forms of free flight.  The step's tangent map grows quickly
under the reference's particle-collision pass (about x2 per step for a
jittered lattice at stiffness 0.3), so the horizons here stay at a few steps
with collisions and longer without them."""
import numpy as np
import pytest

from conftest import small_dam


def _scene(F, side, seed=1, vstd=0.1, **over):
    p = small_dam(F, side, **over)
    pos, _, _ = F.generate_box(p)
    x0 = pos[:, :3].astype(np.float64)
    rng = np.random.default_rng(seed)
    v0 = rng.normal(0.0, vstd, x0.shape)
    return p, x0, v0, rng


def _loss_setup(O, p, x0, v0, nsub, rng):
    xf, _, _, orig = O.forward64(p, x0, v0, nsub=nsub)
    n = x0.shape[0]
    target = xf[np.argsort(orig)] + rng.normal(0.0, 0.01, x0.shape)  # by initial index

    def loss(v=v0, x=x0, rho0=0.0, eps=-1.0):
        X, _, _, og = O.forward64(p, x, v, nsub=nsub, rho0=rho0, eps=eps)
        return np.sum((X - target[og]) ** 2) / n

    X, _, _, og = O.forward64(p, x0, v0, nsub=nsub)
    gx = 2.0 * (X - target[og]) / n
    g = O.gradient64(p, x0, v0, gx, np.zeros_like(gx), nsub=nsub)
    return loss, g


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("side,nsub,h,tol,over", [
    (6, 1, 1e-6, 1e-6, {}),
    (8, 4, 1e-7, 1e-5, {}),
    (8, 10, 1e-7, 1e-5, {"stabilization_iterations": 0}),
])
def test_adjoint_matches_central_differences(F, O, side, nsub, h, tol, over):
    p, x0, v0, rng = _scene(F, side, **over)
    loss, g = _loss_setup(O, p, x0, v0, nsub, rng)
    for _ in range(3):
        u = rng.normal(size=x0.shape)
        fd = (loss(v=v0 + h * u) - loss(v=v0 - h * u)) / (2 * h)
        assert _rel(np.sum(g["v0"] * u), fd) < tol
        hx = 0.1 * h  # positions act on the contacts directly: a smaller step keeps their active set
        fd = (loss(x=x0 + hx * u) - loss(x=x0 - hx * u)) / (2 * hx)
        assert _rel(np.sum(g["x0"] * u), fd) < tol
    d = O.derive(p)
    hr = 1e-6
    fd = (loss(rho0=d.rho0 * (1 + hr)) - loss(rho0=d.rho0 * (1 - hr))) / (2 * hr * d.rho0)
    assert _rel(g["rho0"], fd) < 10 * tol
    fd = (loss(eps=d.eps * (1 + hr)) - loss(eps=d.eps * (1 - hr))) / (2 * hr * d.eps)
    assert _rel(g["eps"], fd) < 10 * tol


def test_adjoint_free_flight_closed_form(F, O):
    """Particles farther apart than every kernel width and inside the box fly
    freely: x_T = x0 + N dt v0 + gravity terms, v_T = v0 + gravity terms, so
    dL/dv0 = N dt dL/dx_T + dL/dv_T and dL/dx0 = dL/dx_T, with no dependence
    on rho0 or eps."""
    p = small_dam(F, 4)
    p.box_spacing = 0.3
    p.box_origin[0], p.box_origin[1], p.box_origin[2] = 0.6, 0.5, 0.3
    pos, _, _ = F.generate_box(p)
    x0 = pos[:, :3].astype(np.float64)
    n = x0.shape[0]
    rng = np.random.default_rng(3)
    v0 = rng.normal(0.0, 0.05, x0.shape)
    nsub = 3
    X, V, _, og = O.forward64(p, x0, v0, nsub=nsub)
    gx = rng.normal(size=(n, 3))
    gv = rng.normal(size=(n, 3))
    g = O.gradient64(p, x0, v0, gx, gv, nsub=nsub)
    d = O.derive(p)
    inv = np.empty(n, np.int64)
    inv[og] = np.arange(n)  # final slot of each initial particle
    np.testing.assert_allclose(g["v0"], nsub * d.dt * gx[inv] + gv[inv], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g["x0"], gx[inv], rtol=1e-12, atol=1e-12)
    assert g["rho0"] == 0.0 and g["eps"] == 0.0


def test_adjoint_of_clamped_axis_is_zero(F, O):
    """A particle projected onto the floor in every step has no sensitivity of
    its final height to its initial vertical velocity (the clamp is constant)."""
    p = small_dam(F, 4)
    p.box_spacing = 0.3
    p.box_origin[0], p.box_origin[1], p.box_origin[2] = 0.6, 0.0015, 0.3
    p.box_jitter = 0.0
    pos, _, _ = F.generate_box(p)
    x0 = pos[:, :3].astype(np.float64)
    n = x0.shape[0]
    v0 = np.zeros_like(x0)
    v0[:, 1] = -1.0  # into the floor
    floor = np.abs(x0[:, 1] - x0[:, 1].min()) < 1e-9
    gx = np.zeros((n, 3))
    X, V, _, og = O.forward64(p, x0, v0, nsub=2)
    gx[:, 1] = 1.0
    g = O.gradient64(p, x0, v0, gx, np.zeros_like(gx), nsub=2)
    assert np.all(g["v0"][floor, 1] == 0.0)
    assert np.all(np.abs(g["v0"][~floor, 1]) > 0.0)


def test_adjoint_rejects_unsupported_steps(F, O):
    p, x0, v0, _ = _scene(F, 4)
    q = p.copy()
    q.kernel_variant = F.PBD_KERNEL_REFERENCE
    with pytest.raises(RuntimeError):
        O.forward64(q, x0, v0, nsub=1)
    q = p.copy()
    q.collision_type = 1
    with pytest.raises(RuntimeError):
        O.gradient64(q, x0, v0, np.zeros_like(x0), np.zeros_like(x0), nsub=1)
