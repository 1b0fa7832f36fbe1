"""GPU parity of the SPH path (cudaParticleSPH_NS, reference physics) through
the C ABI against the FP64 oracle: density, 1/rho and mu; the acceleration;
the reference driver loop of testSPH5.cu:193-258 (bin, density, acceleration,
gravity, leap-frog, absolute boundary) over several steps; inspectDensity.
FP64 on both sides, so the tolerance only covers summation order:
rel 1e-11 on densities, 1e-9 of max|a| on accelerations."""
import numpy as np
import pytest

import particlebase_amd as pb
from particlebase_amd import workloads

pytestmark = pytest.mark.gpu
G = 9.81


def say(msg):
    print(f"\n[case] {msg}", flush=True)


def make(ctx, w):
    s = pb.System(ctx, [w.n])
    s.set_cell(w.cell)
    s.import_particles(w.r, w.m, is_fixed=w.fixed)
    s.sph_set_properties(w.mu0, w.c, w.h)
    s.setup_cells(w.h, periodic=False)
    return s


def oracle_state(orc, w, r, v):
    num, rhoinv = orc.sph_density(r, w.m, w.mu0, w.h)
    a = orc.sph_accel(r, v, w.m, w.c ** 2, rhoinv, num[:, 2], w.h)
    return num, rhoinv, a


@pytest.mark.parametrize("jitter", [0.0, 0.3])
def test_density_and_acceleration(ctx, orc, jitter):
    say(f"SPH dam break dx=0.04 jitter={jitter}: density + acceleration vs oracle")
    w = workloads.sph_dam_break(0.04, jitter=jitter, seed=7)
    s = make(ctx, w)
    rng = np.random.default_rng(3)
    v = rng.normal(scale=0.1, size=(w.n, 3)) * (1 - w.fixed)[:, None]
    s.set("v", v)
    s.bin()
    s.sph_density()
    num = s.get("sph_num")
    rhoinv = s.get("sph_rhoinv")
    s.sph_acceleration()
    a = s.get("a")
    no, ro, ao = oracle_state(orc, w, w.r, v)
    assert np.allclose(num, no, rtol=1e-11, atol=0)
    assert np.allclose(rhoinv, ro, rtol=1e-11, atol=0)
    assert np.abs(a - ao).max() < 1e-9 * np.abs(ao).max()
    assert s.sph_inspect_density() == pytest.approx(no[w.fixed == 0, 0].mean(), rel=1e-12)


def test_momentum_free_fluid(ctx, orc):
    say("SPH free fluid block (no walls): sum m a = 0")
    rng = np.random.default_rng(11)
    n, h = 3000, 0.1
    r = rng.uniform(0.0, 1.0, (n, 3))
    w = workloads.SPHSystem(r=r, m=rng.uniform(0.5, 1.5, n) * 1e-3, fixed=np.zeros(n, np.uint8),
                            mu0=rng.uniform(0.01, 0.1, n), c=rng.uniform(1, 2, n), h=h,
                            cell=np.array([0.0, 1.0, 0.0, 1.0, 0.0, 1.0]), dx=0.0)
    s = make(ctx, w)
    v = rng.normal(size=(n, 3))
    s.set("v", v)
    s.bin()
    s.sph_density()
    s.sph_acceleration()
    a = s.get("a")
    tot = (w.m[:, None] * a).sum(0)
    assert np.abs(tot).max() < 1e-10 * np.abs(w.m[:, None] * a).sum()
    _, _, ao = oracle_state(orc, w, r, v)
    assert np.abs(a - ao).max() < 1e-9 * np.abs(ao).max()


def test_reference_loop(ctx, orc):
    say("SPH reference loop (testSPH5 order), 10 steps vs oracle")
    w = workloads.sph_dam_break(0.04, jitter=0.2, seed=9)
    s = make(ctx, w)
    dt = 2e-4
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
        res, _, _ = s.inspect_velocity(2 * w.h / dt, 0.04, 0.0025)
        s.treat_absolute()
        _, _, a = oracle_state(orc, w, r, v)
        a[:, 2] -= G
        r, v = orc.lf_step(r, v, a, move, dt)
        r = np.clip(r, lo, hi)
    assert np.abs(s.get("r") - r).max() < 1e-10
    assert np.abs(s.get("v") - v).max() < 1e-8 * max(np.abs(v).max(), 1.0)
    # the fluid fell and the walls did not move
    assert np.array_equal(s.get("r")[w.fixed == 1], w.r[w.fixed == 1])


def test_state_errors(ctx):
    say("SPH call-order errors")
    w = workloads.sph_dam_break(0.08)
    s = pb.System(ctx, [w.n])
    s.set_cell(w.cell)
    s.import_particles(w.r, w.m, is_fixed=w.fixed)
    with pytest.raises(pb.PBError):
        s.sph_density()                      # no properties
    s.sph_set_properties(w.mu0, w.c, w.h)
    s.setup_cells(w.h, periodic=False)
    with pytest.raises(pb.PBError):
        s.sph_density()                      # not binned
    s.bin()
    with pytest.raises(pb.PBError):
        s.sph_acceleration()                 # no density yet
