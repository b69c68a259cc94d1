"""Pins of the oracle's FP64 reverse mode with phases: granular solid particles
and position-based friction (oracle/pbf_adjoint.c, fric64 / fric64_adj).

The FP64 forward with phases is checked against the FP32 step of the oracle
(orc_step), and its reverse mode against FP64 central finite differences: on a
scene with 30 % solid spheres and friction, and on two-particle contacts placed
in the static and in the kinetic friction branch (the derivative of the branch
the forward took; the steps stay far from the switch)."""
import numpy as np
import pytest

from conftest import small_dam


def _mixed(F, side=8, mu_s=0.5, mu_k=0.4, solid=0.3, seed=1, vstd=0.1):
    p = small_dam(F, side)
    p.friction_static = mu_s
    p.friction_kinetic = mu_k
    p.box_solid_fraction = solid
    p.box_solid_radius = 0.1
    pos, _, _ = F.generate_box(p)
    ph = F.capi.generate_phases(p)
    x0 = pos[:, :3].astype(np.float64)
    rng = np.random.default_rng(seed)
    v0 = rng.normal(0.0, vstd, x0.shape)
    return p, x0, v0, ph, rng


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def test_forward64_with_phases_matches_the_step(F, O):
    """One FP64 step with phases and friction against the oracle's FP32-input step."""
    p, x0, v0, ph, _ = _mixed(F, 10)
    assert 0.25 < ph.mean() < 0.45
    x4 = np.zeros((x0.shape[0], 4), np.float32)
    v4 = np.zeros_like(x4)
    x4[:, :3] = x0
    v4[:, :3] = v0
    x4[:, 3] = 1.0
    st = O.State(x4, v4, phase=ph)
    out = O.step(p, st)
    X, V, _, og = O.forward64(p, x4[:, :3].astype(np.float64), v4[:, :3].astype(np.float64), nsub=1, phase=ph)
    h = O.derive(p).h_spiky
    assert np.array_equal(og, out["ids"])
    assert np.abs(X - out["x4"][:, :3]).max() < 1e-5 * h
    # all-fluid phases reproduce the plain FP64 step exactly
    Xa, Va, _, _ = O.forward64(p, x0, v0, nsub=2, phase=np.zeros_like(ph))
    Xb, Vb, _, _ = O.forward64(p, x0, v0, nsub=2)
    assert np.array_equal(Xa, Xb) and np.array_equal(Va, Vb)


@pytest.mark.parametrize("nsub,h,tol", [(1, 1e-7, 1e-5), (3, 1e-7, 1e-5)])
def test_adjoint_with_phases_matches_central_differences(F, O, nsub, h, tol):
    p, x0, v0, ph, rng = _mixed(F, 8)
    xf, _, _, orig = O.forward64(p, x0, v0, nsub=nsub, phase=ph)
    n = x0.shape[0]
    target = xf[np.argsort(orig)] + rng.normal(0.0, 0.01, x0.shape)

    def loss(v=v0, x=x0, rho0=0.0, eps=-1.0):
        X, _, _, og = O.forward64(p, x, v, nsub=nsub, rho0=rho0, eps=eps, phase=ph)
        return np.sum((X - target[og]) ** 2) / n

    X, _, _, og = O.forward64(p, x0, v0, nsub=nsub, phase=ph)
    gx = 2.0 * (X - target[og]) / n
    g = O.gradient64(p, x0, v0, gx, np.zeros_like(gx), nsub=nsub, phase=ph)
    for _ in range(3):
        u = rng.normal(size=x0.shape)
        fd = (loss(v=v0 + h * u) - loss(v=v0 - h * u)) / (2 * h)
        assert _rel(np.sum(g["v0"] * u), fd) < tol
        hx = 0.1 * h
        fd = (loss(x=x0 + hx * u) - loss(x=x0 - hx * u)) / (2 * hx)
        assert _rel(np.sum(g["x0"] * u), fd) < tol
    d = O.derive(p)
    hr = 1e-6
    fd = (loss(rho0=d.rho0 * (1 + hr)) - loss(rho0=d.rho0 * (1 - hr))) / (2 * hr * d.rho0)
    assert _rel(g["rho0"], fd) < 10 * tol
    fd = (loss(eps=d.eps * (1 + hr)) - loss(eps=d.eps * (1 - hr))) / (2 * hr * d.eps)
    assert _rel(g["eps"], fd) < 10 * tol


@pytest.mark.parametrize("slip,branch", [(0.002, "static"), (0.02, "kinetic")])
def test_friction_adjoint_of_one_contact(F, O, slip, branch):
    """A solid particle at rest and a fluid one 0.04 away (penetration 0.01)
    sliding by `slip` in the step: below mu_s d the contact sticks, above it
    slides.  The gradient of a generic loss of the final state with respect to
    both particles' initial state matches central differences."""
    p = small_dam(F, 4)
    p.gravity = 0.0
    p.friction_static = 0.5
    p.friction_kinetic = 0.3
    d = O.derive(p)
    x0 = np.array([[0.55, 0.55, 0.55], [0.59, 0.551, 0.549]])
    v0 = np.array([[0.0, 0.0, 0.0], [0.001, slip / d.dt, 0.0005]])
    ph = np.array([1, 0], np.uint32)
    pen = d.diameter - 0.04
    assert (slip < 0.5 * pen) == (branch == "static")
    rng = np.random.default_rng(7)
    wx, wv = rng.normal(size=(2, 3)), rng.normal(size=(2, 3))

    def loss(x=x0, v=v0):
        X, V, _, og = O.forward64(p, x, v, nsub=1, phase=ph)
        inv = np.argsort(og)
        return float(np.sum(wx * X[inv]) + np.sum(wv * V[inv]))

    X, V, _, og = O.forward64(p, x0, v0, nsub=1, phase=ph)
    g = O.gradient64(p, x0, v0, wx[og], wv[og], nsub=1, phase=ph)
    h = 1e-7
    for _ in range(4):
        u = rng.normal(size=x0.shape)
        fd = (loss(v=v0 + h * u) - loss(v=v0 - h * u)) / (2 * h)
        assert _rel(np.sum(g["v0"] * u), fd) < 1e-6
        fd = (loss(x=x0 + h * u) - loss(x=x0 - h * u)) / (2 * h)
        assert _rel(np.sum(g["x0"] * u), fd) < 1e-6
