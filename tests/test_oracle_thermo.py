"""Pins of the thermostat oracle (oracle/pb_oracle_thermo.c) by mathematics
that does not repeat its steps:

* GaussianThermo's Newton-Raphson loop against the closed-form fixed point
  of its implicit equation, v = A (1 - dt/2 F.A / m A.A),
  xi = g / (1 - dt/2 g), and the positions that follow;
* the isokinetic property of the Gaussian thermostat in a field of
  harmonic wells (sum m v^2 is held where plain velocity Verlet swings);
* adjustVelocities: zero mean and variance v1^2 afterwards, statistics
  against numpy;
* constTemp: sum F.v = 0 after the correction, and a two-particle case
  worked by hand."""
import numpy as np
import pytest


def system(n, seed):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, 10, (n, 3))
    v = rng.normal(size=(n, 3))
    F = rng.normal(scale=3.0, size=(n, 3))
    Fold = F + rng.normal(scale=0.3, size=(n, 3))
    minv = 1.0 / rng.uniform(0.5, 2.0, n)
    return r, v, F, Fold, minv


def a_vector(v, F, Fold, minv, dt, xi0):
    return v + (F + Fold) * (0.5 * dt) * minv[:, None] - v * (xi0 * 0.5 * dt)


def xi_of(F, v, minv):
    return (F * v).sum() / ((v * v).sum(axis=1) / minv).sum()


@pytest.mark.parametrize("xi0", [0.07, -0.2, 1e-6])
def test_gt_newton_reaches_closed_form(orc, xi0):
    dt = 0.01
    r, v, F, Fold, minv = system(700, 3)
    r1, v1, Fold1, xi1, it = orc.gt_step(r, v, F, Fold, minv, dt, xi0, tol=1e-15)
    assert 2 <= it <= 10
    A = a_vector(v, F, Fold, minv, dt, xi0)
    g = (F * A).sum() / ((A * A).sum(axis=1) / minv).sum()
    vexp = A * (1.0 - 0.5 * dt * g)
    xiexp = g / (1.0 - 0.5 * dt * g)
    assert np.abs(v1 - vexp).max() < 1e-13 * np.abs(vexp).max()
    assert xi1 == pytest.approx(xiexp, rel=1e-11, abs=1e-14)
    rexp = r + vexp * (dt - xiexp * 0.5 * dt * dt) + F * (0.5 * dt * dt) * minv[:, None]
    assert np.abs(r1 - rexp).max() < 1e-13 * np.abs(rexp).max()
    assert np.array_equal(Fold1, F)
    # the converged xi is the definition's sum F.v / sum m v^2
    assert xi1 == pytest.approx(xi_of(F, v1, minv), rel=1e-12)


def test_gt_first_step_stops_after_one_pass(orc):
    # xi(t-dt) = 0: the first Newton pass has f = A - v' = 0, so the loop
    # stops at once with v(t) = A and xi(t) = xi(A), not at the fixed point
    dt = 0.01
    r, v, F, Fold, minv = system(700, 4)
    r1, v1, _, xi1, it = orc.gt_step(r, v, F, Fold, minv, dt, 0.0, tol=0.0)
    assert it == 1
    A = a_vector(v, F, Fold, minv, dt, 0.0)
    assert np.abs(v1 - A).max() <= 1e-15 * np.abs(A).max()
    assert xi1 == pytest.approx(xi_of(F, A, minv), rel=1e-13)
    rexp = r + A * (dt - xi1 * 0.5 * dt * dt) + F * (0.5 * dt * dt) * minv[:, None]
    assert np.abs(r1 - rexp).max() < 1e-13 * np.abs(rexp).max()
    # a small previous xi under a loose threshold also stops after one pass;
    # the result then sits off the fixed point by about dt/2 xi |v|
    _, v2, _, _, it2 = orc.gt_step(r, v, F, Fold, minv, dt, 1e-9, tol=1e-8)
    assert it2 == 1
    g = (F * A).sum() / ((A * A).sum(axis=1) / minv).sum()
    off = np.abs(v2 - A * (1.0 - 0.5 * dt * g)).max()
    assert 1e-7 < off < 1e-3


def test_gt_isokinetic_in_harmonic_wells(orc):
    n, k, dt = 200, 1.0, 0.02
    rng = np.random.default_rng(7)
    r0 = rng.uniform(0, 10, (n, 3))
    minv = 1.0 / rng.uniform(0.5, 2.0, n)
    v0 = rng.normal(size=(n, 3))
    r = r0 + rng.normal(scale=0.5, size=(n, 3))

    def force(x):
        return -k * (x - r0)

    def ke(v):
        return ((v * v).sum(axis=1) / minv).sum()

    # Gaussian thermostat
    x, v, Fold, xi = r.copy(), v0.copy(), np.zeros((n, 3)), 0.0
    kes = []
    for _ in range(300):
        x, v, Fold, xi, it = orc.gt_step(x, v, force(x), Fold, minv, dt, xi)
        assert it > 0
        kes.append(ke(v))
    kes = np.array(kes)
    # after the start-up step (Fold = 0), sum m v^2 is held to O(dt^2)
    assert np.abs(kes[1:] / kes[1] - 1).max() < 2e-3
    # plain velocity Verlet in the same wells swings by tens of percent
    x, v, Fold = r.copy(), v0.copy(), np.zeros((n, 3))
    kv = []
    for _ in range(300):
        x, v, Fold = orc.vv_step(x, v, force(x), Fold, minv, dt)
        kv.append(ke(v))
    assert np.abs(np.array(kv[1:]) / kv[1] - 1).max() > 0.1


def test_adjust_velocities(orc):
    rng = np.random.default_rng(11)
    n = 5000
    v = rng.normal(loc=[1.0, -2.0, 3.0], scale=[0.5, 2.0, 1.0], size=(n, 3))
    v1 = 0.7
    out, st = orc.adjust_velocities(v, v1)
    assert np.allclose(st[:3], v.mean(axis=0), rtol=1e-13, atol=1e-15)
    assert np.allclose(st[3:], v.var(axis=0), rtol=1e-11)
    assert np.abs(out.mean(axis=0)).max() < 1e-13
    assert np.allclose(out.var(axis=0), v1 * v1, rtol=1e-11)
    # an affine map per axis: the standardised values are unchanged
    assert np.allclose((out / v1), (v - v.mean(axis=0)) / v.std(axis=0), rtol=1e-11, atol=1e-12)


def test_const_temp(orc):
    # two particles worked by hand: m = 1, 2; v1 = (1,0,0), v2 = (0,1,0);
    # F1 = (2,0,0), F2 = (0,3,0): sum F.v = 5, sum m v^2 = 3, lambda = 5/3
    v = np.array([[1.0, 0, 0], [0, 1.0, 0]])
    F = np.array([[2.0, 0, 0], [0, 3.0, 0]])
    minv = np.array([1.0, 0.5])
    Fc, mv2, lam = orc.const_temp(v, F, minv)
    assert mv2 == 3.0 and lam == pytest.approx(5.0 / 3.0, rel=1e-15)
    assert np.allclose(Fc, [[1.0 / 3.0, 0, 0], [0, -1.0 / 3.0, 0]], rtol=1e-14, atol=1e-15)
    # random: the corrected force does no work on the velocities
    rng = np.random.default_rng(2)
    v = rng.normal(size=(1000, 3))
    F = rng.normal(size=(1000, 3)) + 0.3 * v
    minv = 1.0 / rng.uniform(0.5, 2.0, 1000)
    Fc, mv2, lam = orc.const_temp(v, F, minv)
    assert abs((Fc * v).sum()) < 1e-11 * np.abs(F * v).sum()
    assert mv2 == pytest.approx(((v * v).sum(axis=1) / minv).sum(), rel=1e-14)
