"""GPU unified phase (granular solids, position-based friction) against the
oracle, through the C-ABI.  Integer stages bit-exact; positions of one step
within 1e-5 h, except where the static/kinetic friction threshold or the
contact distance falls within rounding (a few pairs; bounded separately)."""
import numpy as np
import pytest

from conftest import small_dam

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


def _ctx(F, p, x4, v4=None, c4=None, phases=None):
    q = p.copy()
    q.box_spawn_in_the_beginning = 0
    n = x4.shape[0]
    ctx = F.Context(q, n)
    ctx.add_particles(x4, v4, c4)
    if phases is not None:
        ctx.upload("phases", phases)
    return ctx


@pytest.mark.parametrize("slip,phases", [(0.002, [1, 1]), (0.02, [1, 0]), (0.002, [0, 0])])
def test_gpu_friction_pair_matches_oracle(F, O, slip, phases):
    p = small_dam(F, 4)
    p.gravity = 0.0
    p.solver_iterations = 1
    p.stabilization_iterations = 1
    p.friction_static = 0.5
    p.friction_kinetic = 0.3
    p.active_update_positions = 0
    p.tile_bits = 0
    for k in ("active_compute_lambda", "active_compute_delta_positions", "active_compute_omegas",
              "active_compute_vorticity", "active_compute_viscosity"):
        setattr(p, k, 0)
    d = O.derive(p)
    x = np.array([[0.55, 0.55, 0.55, 1.0], [0.59, 0.55, 0.55, 1.0]], np.float32)
    v = np.zeros((2, 4), np.float32)
    v[1, 1] = np.float32(slip / d.dt)
    flush(f"friction pair: slip {slip}, phases {phases}")
    ctx = _ctx(F, p, x, v, phases=np.array(phases, np.uint32))
    ctx.compute()
    gp, gx = ctx.download("predicted"), ctx.download("positions")
    st = O.State(x, v, phase=np.array(phases, np.uint32))
    out = O.step(p, st)
    assert np.abs(gp[:, :3] - out["p4"][:, :3]).max() < 1e-6
    assert np.abs(gx[:, :3] - out["x4"][:, :3]).max() < 1e-6
    assert np.array_equal(ctx.download("phases"), st.phase)
    ctx.close()


@pytest.mark.parametrize("tile_bits", [5, 0])
def test_gpu_mixed_step_matches_oracle(F, O, tile_bits):
    """A 12^3 dam break with 30 % of its particles solid (seeded spheres of radius
    0.1, static mu 0.5, kinetic 0.4): one step against the oracle, on the LDS
    tiles (phases in the neighbour lists) and on the per-particle kernels."""
    p = small_dam(F, 12)
    p.tile_bits = tile_bits
    p.friction_static = 0.5
    p.friction_kinetic = 0.4
    p.box_solid_fraction = 0.3
    p.box_solid_radius = 0.1
    pos, vel, col = F.generate_box(p)
    ph = F.capi.generate_phases(p)
    assert 0.29 < ph.mean() < 0.4
    rng = np.random.default_rng(4)
    vel[:, :3] = rng.normal(0.0, 0.2, (pos.shape[0], 3)).astype(np.float32)
    flush(f"mixed step (tile bits {tile_bits}): n={pos.shape[0]}, solid share {ph.mean():.3f}")
    ctx = _ctx(F, p, pos, vel, col, ph)
    ctx.compute()
    keys, ids = ctx.download("cell_ids"), ctx.download("particle_ids")
    gx, gv, gph = ctx.download("positions"), ctx.download("velocities"), ctx.download("phases")
    st = O.State(pos, vel, col, phase=ph)
    out = O.step(p, st)
    assert np.array_equal(keys, out["keys"]) and np.array_equal(ids, out["ids"])
    assert np.array_equal(gph, st.phase)
    h = float(p.kernel_width_poly)
    err = np.abs(gx[:, :3] - out["x4"][:, :3]).max(axis=1)
    flush(f"  position error: max {err.max():.2e}, 99.9th pct {np.quantile(err, 0.999):.2e}")
    assert np.quantile(err, 0.999) < 1e-5 * h
    assert err.max() < 1e-3 * h
    solid = gph == 1
    d = O.derive(p)
    verr = np.abs(gv[:, :3] - out["v4"][:, :3]).max(axis=1)
    assert np.quantile(verr, 0.999) < 1e-5 * h / d.dt
    ctx.close()
    assert solid.any()


def test_gpu_phases_follow_particles(F, O):
    """With findNeighbours off the colours are carried through the sort: colour
    = initial index, so the phases after steps must be the initial phases of
    the indices the colours name."""
    p = small_dam(F, 10)
    p.active_find_neighbours = 0
    p.friction_static = 0.5
    p.friction_kinetic = 0.4
    p.box_solid_fraction = 0.3
    pos, vel, _ = F.generate_box(p)
    n = pos.shape[0]
    col = np.zeros((n, 4), np.float32)
    col[:, 0] = np.arange(n, dtype=np.float32)
    ph = F.capi.generate_phases(p)
    ctx = _ctx(F, p, pos, vel, col, ph)
    flush("phases through 3 frames")
    ctx.step(3)
    idx = ctx.download("colors")[:, 0].astype(np.int64)
    assert np.array_equal(np.sort(idx), np.arange(n))
    assert np.array_equal(ctx.download("phases"), ph[idx])
    assert ctx.check_finite() == 0
    ctx.close()
