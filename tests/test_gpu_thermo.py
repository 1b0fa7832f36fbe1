"""GPU parity of the MD thermostat helpers through the C ABI against the FP64
oracle: GaussianThermo's step (closed form where the reference's Newton
loop converges, its one-pass result where the loop stops after one pass),
the testMD2 loop order (calcForce, TimeEvolution, treatPeriodicCondition)
over 40 steps, batched replicas against single systems, the isokinetic
property at 256k atoms, the no-solution count, adjustVelocities and
constTemp.  FP64 on both sides."""
import numpy as np
import pytest

import particlebase_amd as pb
from particlebase_amd import workloads

pytestmark = pytest.mark.gpu
RC2 = 6.25


def say(msg):
    print(f"\n[case] {msg}", flush=True)


def state(n, seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, 10, (n, 3))
    v = rng.normal(size=(n, 3))
    F = rng.normal(scale=3.0, size=(n, 3))
    Fold = F + rng.normal(scale=0.3, size=(n, 3))
    m = rng.uniform(0.5, 2.0, n)
    return r, v, F, Fold, m


def loaded(ctx, parts):
    s = pb.System(ctx, [p[0].shape[0] for p in parts])
    s.set_cell(np.tile([0, 10.0, 0, 10.0, 0, 10.0], len(parts)))
    cat = lambda k: np.concatenate([p[k] for p in parts])
    s.import_particles(cat(0), cat(4), v=cat(1), a=cat(3))
    s.set("F", cat(2))
    return s


@pytest.mark.parametrize("xi0,tol", [(0.07, 1e-12), (-0.2, 1e-8), (0.0, 1e-8), (1e-9, 1e-8)])
def test_gt_step_vs_oracle(ctx, orc, xi0, tol):
    say(f"Gaussian step xi0={xi0} tol={tol}, N=3000")
    r, v, F, Fold, m = state(3000, 1)
    dt = 0.01
    s = loaded(ctx, [(r, v, F, Fold, m)])
    s.gt_xi(set_xi=[xi0])
    s.gt_step(dt, tol)
    ro, vo, Fo, xio, it = orc.gt_step(r, v, F, Fold, 1 / m, dt, xi0, tol)
    assert it > 0
    xi, bad = s.gt_xi()
    assert bad == 0
    vg = s.get("v")
    # one-pass exits (xi0 = 0, tiny xi0 under tol 1e-8) are the same arithmetic
    # on both sides; converged loops sit within the loop's own tolerance of
    # the closed-form fixed point
    vtol = 1e-13 if it == 1 else max(1e-12, 10 * tol)
    assert np.abs(vg - vo).max() < vtol * np.abs(vo).max(), it
    assert xi[0] == pytest.approx(xio, rel=max(1e-12, 100 * tol), abs=1e-16)
    assert np.abs(s.get("r") - ro).max() < 1e-12 * np.abs(ro).max()
    assert np.array_equal(s.get("a"), F)


def lj_sys(ctx, w, counts=None):
    s = pb.System(ctx, [w.n] if counts is None else counts)
    cells = w.cell if counts is None else np.tile(w.cell, len(counts))
    s.set_cell(cells)
    return s


def test_gt_lj_loop_vs_oracle(ctx, orc):
    say("testMD2 loop order: LJ force, Gaussian step, wrap; 40 steps, N=4000")
    w = workloads.lj_config1()
    rng = np.random.default_rng(5)
    w.r = (w.r + rng.normal(scale=0.05, size=w.r.shape)) % w.L
    s = lj_sys(ctx, w)
    s.import_particles(w.r, w.m, v=w.v, type=w.type)
    s.lj_set_params([1.0, 1.0], 1)
    s.lj_set_cutoff(RC2, shifted=True)
    s.clear_force(True)
    p = orc.lj_params([1.0, 1.0], 1, RC2, True)
    dt, tol = 0.005, 1e-10
    x, u, Fold, xi = w.r.copy(), w.v.copy(), np.zeros_like(w.r), 0.0
    minv = 1 / w.m
    for k in range(40):
        s.lj_force(pb.PB_FP64, energies=False)
        s.gt_step(dt, tol)
        s.treat_periodic()
        F, _, _ = orc.lj_forces_cells(x, w.type, w.cell, RC2, p)
        x, u, Fold, xi, it = orc.gt_step(x, u, F, Fold, minv, dt, xi, tol)
        assert it > 0, k
        x = orc.wrap(x, w.cell)
    xg, ug = s.get("r"), s.get("v")
    d = xg - x
    d -= w.L * np.round(d / w.L)
    assert np.abs(d).max() < 1e-9
    assert np.abs(ug - u).max() < 1e-8 * np.abs(u).max()
    assert s.gt_xi()[0][0] == pytest.approx(xi, rel=1e-6, abs=1e-12)
    # isokinetic: sum m v^2 after the first step stays where the first step put it
    assert s.mv2()[0] == pytest.approx(orc.mv2(u, minv), rel=1e-10)


def test_gt_batched_equals_single(ctx):
    say("Gaussian step, 3 replicas in one system vs each alone")
    parts = [state(n, 10 + i) for i, n in enumerate((700, 1500, 64))]
    xis = [0.05, -0.01, 0.0]
    sb = loaded(ctx, parts)
    sb.gt_xi(set_xi=xis)
    for _ in range(3):
        sb.gt_step(0.01, 1e-9)
    vb, rb = sb.get("v"), sb.get("r")
    xb, _ = sb.gt_xi()
    off = 0
    for q, (p, x0) in enumerate(zip(parts, xis)):
        s = loaded(ctx, [p])
        s.gt_xi(set_xi=[x0])
        for _ in range(3):
            s.gt_step(0.01, 1e-9)
        n = p[0].shape[0]
        # per-replica sums run in the same fixed order in both layouts
        assert np.array_equal(s.get("v"), vb[off:off + n]), q
        assert np.array_equal(s.get("r"), rb[off:off + n]), q
        assert s.gt_xi()[0][0] == xb[q], q
        off += n


def test_gt_isokinetic_large(ctx):
    say("Gaussian thermostat holds sum m v^2 on the 256k-atom LJ liquid (100 steps)")
    w = workloads.lj_fcc(40, 0.8442, 0.72, seed=3)   # 256,000 atoms
    s = lj_sys(ctx, w)
    s.import_particles(w.r, w.m, v=w.v, type=w.type)
    s.lj_set_params([1.0, 1.0], 1)
    s.lj_set_cutoff(RC2, shifted=True)
    s.clear_force(True)
    k0 = None
    for k in range(100):
        s.lj_force(pb.PB_FP64, energies=False)
        s.gt_step(0.005, 1e-10)
        s.treat_periodic()
        if k == 1:
            k0 = s.mv2()[0]
    k1 = s.mv2()[0]
    # from a lattice start plain VV loses about half its kinetic energy here
    assert abs(k1 / k0 - 1) < 5e-3
    assert s.gt_xi()[1] == 0


def test_gt_no_solution_counted(ctx):
    say("Gaussian step whose implicit equation has no solution")
    n = 256
    rng = np.random.default_rng(3)
    F = rng.normal(size=(n, 3))
    # m = 1, dt = 1, Fold = F, v = -0.9 F, xi(t-dt) = 0.3: A = 0.235 F and
    # dt/2 g = 1/0.47 > 1, while the first Newton pass still moves v
    s = loaded(ctx, [(rng.uniform(0, 10, (n, 3)), -0.9 * F, F, F, np.ones(n))])
    s.gt_xi(set_xi=[0.3])
    s.gt_step(1.0, 0.0)
    xi, bad = s.gt_xi()
    assert bad == 1 and xi[0] == 0.0
    assert np.all(np.isfinite(s.get("v"))) and np.all(np.isfinite(s.get("r")))
    # restoring xi clears the count
    assert s.gt_xi(set_xi=[0.0])[1] == 1 and s.gt_xi()[1] == 0


def test_adjust_velocities_vs_oracle(ctx, orc):
    say("adjustVelocities, 2 replicas (5000, 777) vs oracle")
    rng = np.random.default_rng(8)
    vs = [rng.normal(loc=[1.0, -2.0, 3.0], scale=[0.5, 2.0, 1.0], size=(n, 3)) for n in (5000, 777)]
    s = pb.System(ctx, [5000, 777])
    s.set_cell(np.tile([0, 10.0, 0, 10.0, 0, 10.0], 2))
    s.import_particles(rng.uniform(0, 10, (5777, 3)), np.ones(5777), v=np.concatenate(vs))
    st = s.adjust_velocities(0.7)
    vg = s.get("v")
    off = 0
    for q, v in enumerate(vs):
        vo, sto = orc.adjust_velocities(v, 0.7)
        n = v.shape[0]
        assert np.allclose(st[q], sto, rtol=1e-11, atol=1e-14)
        assert np.abs(vg[off:off + n] - vo).max() < 1e-11
        off += n
    with pytest.raises(pb.PBError):
        z = pb.System(ctx, [64])
        z.set_cell(np.array([0, 10.0, 0, 10.0, 0, 10.0]))
        z.import_particles(rng.uniform(0, 10, (64, 3)), np.ones(64), v=np.zeros((64, 3)))
        z.adjust_velocities(1.0)


def test_const_temp_vs_oracle(ctx, orc):
    say("constTemp vs oracle, N=3000")
    r, v, F, Fold, m = state(3000, 4)
    s = loaded(ctx, [(r, v, F, Fold, m)])
    mv2, lam = s.const_temp()
    Fo, mv2o, lamo = orc.const_temp(v, F, 1 / m)
    assert mv2[0] == pytest.approx(mv2o, rel=1e-13)
    assert lam[0] == pytest.approx(lamo, rel=1e-12)
    assert np.abs(s.get("F") - Fo).max() < 1e-13 * np.abs(Fo).max()
