"""GPU parity of the SPH config physics (cubic spline of support 2h, Tait,
artificial viscosity; pb_sph_set_wcsph) through the C ABI against the FP64
oracle: density, pressure and acceleration on the dam break at dx = 0.04,
on the lattice and jittered, with random velocities; ten steps of the
dam-break loop; and at the full 796k-particle size the size-independent
properties (sum m a = 0 without gravity, rest density rho0 for every
particle with full support on the initial lattice).  FP64 on both sides:
rel 1e-11 on densities, 1e-9 B on pressures, 1e-9 of max|a| on
accelerations."""
import numpy as np
import pytest

import particlebase_amd as pb
from particlebase_amd import workloads

pytestmark = pytest.mark.gpu
G = 9.81
ALPHA = 0.1


def say(msg):
    print(f"\n[case] {msg}", flush=True)


def make(ctx, w):
    s = pb.System(ctx, [w.n])
    s.set_cell(w.cell)
    s.import_particles(w.r, w.m, is_fixed=w.fixed)
    s.sph_set_wcsph(w.h, w.rho0, w.c[0], 7.0, ALPHA, 0.0)
    s.setup_cells(2 * w.h, periodic=False)
    return s


def oracle_state(orc, w, r, v):
    num, rhoinv = orc.wc_density(r, w.m, w.h, w.rho0, w.c[0], 7.0, 0.0)
    a = orc.wc_accel(r, v, w.m, num[:, 1], num[:, 2], w.h, w.c[0], ALPHA)
    return num, rhoinv, a


@pytest.mark.parametrize("jitter", [0.0, 0.3])
def test_density_and_acceleration(ctx, orc, jitter):
    say(f"SPH config physics, dam break dx=0.04 jitter={jitter}: density + acceleration")
    w = workloads.sph_dam_break_config(0.04, jitter=jitter, seed=5)
    s = make(ctx, w)
    rng = np.random.default_rng(4)
    v = rng.normal(scale=0.5, size=(w.n, 3)) * (1 - w.fixed)[:, None]
    s.set("v", v)
    s.bin()
    s.sph_density()
    num = s.get("sph_num")
    rhoinv = s.get("sph_rhoinv")
    s.sph_acceleration()
    a = s.get("a")
    no, ro, ao = oracle_state(orc, w, w.r, v)
    B = w.rho0 * w.c[0] ** 2 / 7
    assert np.allclose(num[:, :2], no[:, :2], rtol=1e-11, atol=0)
    assert np.abs(num[:, 2] - no[:, 2]).max() < 1e-9 * B
    assert np.allclose(rhoinv, ro, rtol=1e-11, atol=0)
    assert (no[:, 2] == 0).sum() > 100                 # the clamp is exercised
    if jitter:
        assert (no[:, 2] > 1e-6 * B).sum() > 100       # and real pressures
    assert np.abs(a - ao).max() < 1e-9 * np.abs(ao).max()
    assert s.sph_inspect_density() == pytest.approx(no[w.fixed == 0, 0].mean(), rel=1e-12)


def test_dam_loop(ctx, orc):
    say("SPH config physics, dam-break loop, 10 steps vs oracle")
    w = workloads.sph_dam_break_config(0.04, jitter=0.2, seed=9)
    s = make(ctx, w)
    dt = 0.25 * w.h / w.c[0]
    move = (1 - w.fixed).astype(np.uint8)
    r = w.r.copy()
    v = np.zeros_like(r)
    lo, hi = w.cell[0::2], w.cell[1::2]
    for _ in range(10):
        s.bin()
        s.sph_density()
        s.sph_acceleration()
        s.add_acceleration(2, -G)
        s.lf_step(dt)
        s.treat_absolute()
        _, _, a = oracle_state(orc, w, r, v)
        a[:, 2] -= G
        r, v = orc.lf_step(r, v, a, move, dt)
        r = np.clip(r, lo, hi)
    assert np.abs(s.get("r") - r).max() < 1e-12
    assert np.abs(s.get("v") - v).max() < 1e-9 * max(np.abs(v).max(), 1.0)
    assert np.abs(v).max() > 1e-3
    assert np.array_equal(s.get("r")[w.fixed == 1], w.r[w.fixed == 1])


def test_full_size_invariants(ctx):
    say("SPH config physics at dx=0.4/63 (796k particles): sum m a = 0, rest density")
    w = workloads.sph_dam_break_config()
    s = make(ctx, w)
    rng = np.random.default_rng(8)
    v = rng.normal(scale=0.2, size=(w.n, 3)) * (1 - w.fixed)[:, None]
    s.set("v", v)
    s.bin()
    s.sph_density()
    num = s.get("sph_num")
    s.sph_acceleration()
    ma = w.m[:, None] * s.get("a")
    assert np.abs(ma.sum(0)).max() < 1e-10 * np.abs(ma).sum()
    # particles at least 3 dx inside the fluid (and wall) lattice have the
    # full support of the initial lattice: density rho0
    dx = w.dx
    full = (w.r[:, 0] < 0.4 - 3 * dx) & (w.r[:, 2] < 0.8 - 3 * dx) & (w.fixed == 0)
    assert full.sum() > 300_000
    assert np.allclose(num[full, 1], w.rho0, rtol=1e-12, atol=0)
    assert np.abs(num[full, 2]).max() < 1e-9 * w.rho0 * w.c[0] ** 2 / 7


def test_switch_physics(ctx, orc):
    say("SPH: switching between config and reference physics on one system")
    w = workloads.sph_dam_break(0.08)
    s = pb.System(ctx, [w.n])
    s.set_cell(w.cell)
    s.import_particles(w.r, w.m, is_fixed=w.fixed)
    s.sph_set_wcsph(w.h, 1.0, w.c[0], 7.0, ALPHA, 0.0)
    s.setup_cells(w.h, periodic=False)
    s.bin()
    with pytest.raises(pb.PBError):
        s.sph_density()                      # cells shorter than the support 2h
    s.sph_set_properties(w.mu0, w.c, w.h)
    s.bin()
    s.sph_density()
    no, _ = orc.sph_density(w.r, w.m, w.mu0, w.h)
    assert np.allclose(s.get("sph_num"), no, rtol=1e-11, atol=0)
    with pytest.raises(pb.PBError):
        s.sph_set_wcsph(w.h, -1.0, w.c[0])   # rho0 must be positive


def test_rebin_between_density_and_acceleration(ctx, orc):
    say("SPH config physics: a new cell grid and pb_bin between density and acceleration")
    w = workloads.sph_dam_break_config(0.04, jitter=0.3, seed=12)
    s = make(ctx, w)
    rng = np.random.default_rng(6)
    v = rng.normal(scale=0.5, size=(w.n, 3)) * (1 - w.fixed)[:, None]
    s.set("v", v)
    s.bin()
    s.sph_density()
    s.setup_cells(2.6 * w.h, periodic=False)   # another grid: another cell order
    s.bin()
    s.sph_acceleration()
    _, _, ao = oracle_state(orc, w, w.r, v)
    assert np.abs(s.get("a") - ao).max() < 1e-9 * np.abs(ao).max()


@pytest.mark.parametrize("physics", ["reference", "config"])
def test_particles_outside_the_grid(ctx, orc, physics):
    say(f"SPH {physics} physics with particles just outside the cell grid (edge cells)")
    if physics == "config":
        w = workloads.sph_dam_break_config(0.04, jitter=0.2, seed=21)
        sup = 2 * w.h
    else:
        w = workloads.sph_dam_break(0.04, jitter=0.2, seed=21)
        sup = w.h
    r = w.r.copy()
    fluid = np.nonzero(w.fixed == 0)[0]
    top = fluid[np.argsort(r[fluid, 2])[-40:]]
    r[top, 2] = w.cell[5] + 0.3 * sup * np.linspace(0.1, 1.0, top.size)   # above the grid
    s = pb.System(ctx, [w.n])
    s.set_cell(w.cell)
    s.import_particles(r, w.m, is_fixed=w.fixed)
    if physics == "config":
        s.sph_set_wcsph(w.h, w.rho0, w.c[0], 7.0, ALPHA, 0.0)
    else:
        s.sph_set_properties(w.mu0, w.c, w.h)
    s.setup_cells(sup, periodic=False)
    s.bin()
    s.sph_density()
    num = s.get("sph_num")
    if physics == "config":
        no, _ = orc.wc_density(r, w.m, w.h, w.rho0, w.c[0], 7.0, 0.0)
        assert np.allclose(num[:, :2], no[:, :2], rtol=1e-11, atol=0)
    else:
        no, _ = orc.sph_density(r, w.m, w.mu0, w.h)
        assert np.allclose(num, no, rtol=1e-11, atol=0)
    # most moved particles have neighbours besides themselves (pairs across
    # the grid's top face and among the moved particles were found)
    w_self = 1 / (np.pi * w.h ** 3) if physics == "config" else 105 / (16 * np.pi * w.h ** 3)
    assert (no[top, 0] > 1.05 * w_self).sum() > top.size // 2
