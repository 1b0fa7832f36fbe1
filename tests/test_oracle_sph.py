"""Pins of the SPH oracle (reference physics of cudaParticleSPH_NS) by the
mathematics: kernel normalisation, the kernel derivative against a finite
difference of the kernel, the self term, the SPH interpolation of a lattice,
momentum conservation and the signs of the pressure and viscous terms."""
import numpy as np
import pytest
from scipy import integrate

H = 0.5


def test_lucy_normalisation(orc):
    # int_0^h W(r) 4 pi r^2 dr = 1 for any h
    for h in (0.3, 0.5, 1.7):
        val, _ = integrate.quad(lambda r: orc.sph_lucy_w(r, h) * 4 * np.pi * r * r, 0, h)
        assert val == pytest.approx(1.0, rel=1e-12)
    assert orc.sph_lucy_w(H, H) == 0.0 and orc.sph_lucy_w(1.2 * H, H) == 0.0


def test_lucy_derivative_matches_kernel(orc):
    # (dW/dr)/r from the oracle against a central difference of W
    for r in np.linspace(0.05, 0.95, 10) * H:
        e = 1e-6 * H
        fd = (orc.sph_lucy_w(r + e, H) - orc.sph_lucy_w(r - e, H)) / (2 * e)
        assert orc.sph_lucy_dw(r, H) * r == pytest.approx(fd, rel=1e-7)
    assert orc.sph_lucy_dw(H * 1.01, H) == 0.0


def test_single_particle_density(orc):
    m, mu0 = 2.5, 0.7
    num, rhoinv = orc.sph_density(np.zeros((1, 3)), [m], [mu0], H)
    # self term: W(0) = 105 / (16 pi h^3)
    assert num[0, 1] == pytest.approx(105 * m / (16 * np.pi * H ** 3), rel=1e-14)
    assert num[0, 2] == pytest.approx(mu0, rel=1e-14)
    assert rhoinv[0] == pytest.approx(1 / num[0, 1], rel=1e-15)


def test_lattice_density_interpolates(orc):
    # a cubic lattice of spacing s and particle mass m has density m / s^3;
    # the SPH sum reproduces it at an interior particle when h >> s
    s, m = H / 3.0, 0.8
    g = np.arange(-6, 7) * s
    r = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = len(r)
    num, _ = orc.sph_density(r, np.full(n, m), np.ones(n), H)
    c = np.argmin((r ** 2).sum(1))
    assert num[c, 1] == pytest.approx(m / s ** 3, rel=0.02)
    assert num[c, 0] == pytest.approx(1 / s ** 3, rel=0.02)


def random_fluid(n, seed, box=1.2):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, box, (n, 3))
    v = rng.normal(size=(n, 3))
    m = rng.uniform(0.5, 1.5, n)
    mu0 = rng.uniform(0.01, 0.1, n)
    c2 = rng.uniform(1.0, 4.0, n)
    return r, v, m, mu0, c2


def test_momentum_conservation_and_translation(orc):
    r, v, m, mu0, c2 = random_fluid(400, 3)
    num, rhoinv = orc.sph_density(r, m, mu0, H)
    a = orc.sph_accel(r, v, m, c2, rhoinv, num[:, 2], H)
    # every pair term is antisymmetric: sum m_i a_i = 0 up to rounding
    tot = (m[:, None] * a).sum(0)
    assert np.abs(tot).max() < 1e-10 * np.abs(m[:, None] * a).sum()
    # translation and Galilean invariance
    num2, rhoinv2 = orc.sph_density(r + 3.1, m, mu0, H)
    a2 = orc.sph_accel(r + 3.1, v + 0.4, m, c2, rhoinv2, num2[:, 2], H)
    assert np.allclose(num2, num, rtol=1e-11) and np.allclose(a2, a, rtol=1e-8, atol=1e-8)


def test_pressure_repels_and_viscosity_damps(orc):
    # two particles, compressed (rho > rho0 = 1): the pressure term pushes
    # them apart; with the pressure off (c = 0) the viscous term reduces the
    # relative velocity
    r = np.array([[0.0, 0, 0], [0.2 * H, 0, 0]])
    m = np.array([3.0, 3.0])
    num, rhoinv = orc.sph_density(r, m, [0.0, 0.0], H)
    assert num[0, 1] > 1.0
    a = orc.sph_accel(r, np.zeros((2, 3)), m, [1.0, 1.0], rhoinv, [0.0, 0.0], H)
    assert a[0, 0] < 0 < a[1, 0] and np.allclose(a[:, 1:], 0)
    v = np.array([[1.0, 0.5, 0], [-1.0, 0.0, 0]])
    a = orc.sph_accel(r, v, m, [0.0, 0.0], rhoinv, [0.3, 0.3], H)
    dv = v[1] - v[0]
    da = a[1] - a[0]
    assert np.dot(dv, da) < 0
    # rarefied (rho < 1): p < 0 pulls the pair together
    num, rhoinv = orc.sph_density(r, [0.01, 0.01], [0.0, 0.0], H)
    assert num[0, 1] < 1.0
    a = orc.sph_accel(r, np.zeros((2, 3)), [0.01, 0.01], [1.0, 1.0], rhoinv, [0.0, 0.0], H)
    assert a[0, 0] > 0 > a[1, 0]


def test_lf_step(orc):
    rng = np.random.default_rng(1)
    r, v, a = rng.normal(size=(3, 50, 3))
    move = rng.integers(0, 2, 50)
    dt = 0.01
    r1, v1 = orc.lf_step(r, v, a, move, dt)
    mv = move[:, None].astype(float)
    assert np.allclose(v1, v + a * dt * mv, rtol=0, atol=1e-15)
    assert np.allclose(r1, r + v1 * dt * mv, rtol=0, atol=1e-15)


# ------------------------------------------- config physics (cubic spline)
HC = 0.4


def test_spline_normalisation_and_continuity(orc):
    # int_0^2h W 4 pi r^2 dr = 1; W and dW/dr continuous at q = 1 and q = 2
    for h in (0.2, 0.4, 1.3):
        val, _ = integrate.quad(lambda r: orc.wc_w(r, h) * 4 * np.pi * r * r, 0, 2 * h,
                                points=[h], epsabs=0, epsrel=1e-13)
        assert val == pytest.approx(1.0, rel=1e-12)
        e = 1e-9 * h
        for q in (1.0, 2.0):
            assert orc.wc_w(q * h - e, h) == pytest.approx(orc.wc_w(q * h + e, h), abs=1e-7 / h ** 3)
            assert orc.wc_dw(q * h - e, h) == pytest.approx(orc.wc_dw(q * h + e, h),
                                                            abs=1e-7 / h ** 5)
        assert orc.wc_w(2 * h, h) == 0.0 and orc.wc_dw(2.01 * h, h) == 0.0
    # the 3-D M4 value at the origin and at q = 1: 1/(pi h^3) and 1/(4 pi h^3)
    assert orc.wc_w(0.0, HC) == pytest.approx(1 / (np.pi * HC ** 3), rel=1e-15)
    assert orc.wc_w(HC, HC) == pytest.approx(0.25 / (np.pi * HC ** 3), rel=1e-15)


def test_spline_derivative_matches_kernel(orc):
    for q in np.linspace(0.05, 1.95, 20):
        r, e = q * HC, 1e-6 * HC
        fd = (orc.wc_w(r + e, HC) - orc.wc_w(r - e, HC)) / (2 * e)
        assert orc.wc_dw(r, HC) * r == pytest.approx(fd, rel=1e-7)


def test_spline_single_particle_and_lattice(orc):
    m = 1.7
    num, rhoinv = orc.wc_density(np.zeros((1, 3)), [m], HC, 1.0, 1.0, 7.0, -np.inf)
    assert num[0, 1] == pytest.approx(m / (np.pi * HC ** 3), rel=1e-15)
    assert num[0, 0] == pytest.approx(1 / (np.pi * HC ** 3), rel=1e-15)
    assert rhoinv[0] == pytest.approx(1 / num[0, 1], rel=1e-15)
    # a cubic lattice of spacing s interpolates to m / s^3 at an interior particle
    s = HC / 2.0
    g = np.arange(-6, 7) * s
    r = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    num, _ = orc.wc_density(r, 0.9, HC, 1.0, 1.0, 7.0, -np.inf)
    c = np.argmin((r ** 2).sum(1))
    assert num[c, 1] == pytest.approx(0.9 / s ** 3, rel=1e-3)
    assert num[c, 0] == pytest.approx(1 / s ** 3, rel=1e-3)


def wc_fluid(n, seed, box=1.0):
    rng = np.random.default_rng(seed)
    return (rng.uniform(0, box, (n, 3)), rng.normal(size=(n, 3)), rng.uniform(0.5, 1.5, n))


def test_tait_special_cases(orc):
    r, _, m = wc_fluid(300, 5)
    rho0, c0 = 250.0, 3.0     # interior densities above rho0, the faces' below
    # gamma = 1 reduces the Tait form to the reference's linear EOS:
    # p / rho^2 = c0^2 rhoinv (1 - rho0 rhoinv)
    num, rhoinv = orc.wc_density(r, m, HC, rho0, c0, 1.0, -np.inf)
    assert np.allclose(num[:, 2] * rhoinv ** 2, c0 ** 2 * rhoinv * (1 - rho0 * rhoinv),
                       rtol=1e-12, atol=1e-14 * c0 ** 2 / rho0)
    # gamma = 7: p + B scales as rho^7 between any two particles, p = 0 at rho0
    num7, _ = orc.wc_density(r, m, HC, rho0, c0, 7.0, -np.inf)
    B = rho0 * c0 ** 2 / 7
    pb_ = num7[:, 2] + B
    assert np.allclose(pb_ / pb_[0], (num7[:, 1] / num7[0, 1]) ** 7, rtol=1e-11)
    assert np.array_equal(num7[:, :2], num[:, :2])
    one, _ = orc.wc_density(np.zeros((1, 3)), [rho0 * np.pi * HC ** 3], HC, rho0, c0, 7.0, -np.inf)
    assert abs(one[0, 2]) < 1e-12 * B
    # the clamp
    numc, _ = orc.wc_density(r, m, HC, rho0, c0, 7.0, 0.0)
    assert np.array_equal(numc[:, 2], np.maximum(num7[:, 2], 0.0))
    assert (num7[:, 2] < 0).any() and (num7[:, 2] > 0).any()


def test_wc_momentum_angular_momentum_galilean(orc):
    r, v, m = wc_fluid(500, 6)
    rho0, c0 = 100.0, 2.0
    num, _ = orc.wc_density(r, m, HC, rho0, c0, 7.0, -np.inf)
    a = orc.wc_accel(r, v, m, num[:, 1], num[:, 2], HC, c0, 0.5)
    ma = m[:, None] * a
    scale = np.abs(ma).sum()
    assert np.abs(ma.sum(0)).max() < 1e-12 * scale
    # pair forces along r_ij: no torque about any point
    assert np.abs(np.cross(r - 0.3, ma).sum(0)).max() < 1e-12 * scale
    a2 = orc.wc_accel(r + 2.5, v - 0.7, m, num[:, 1], num[:, 2], HC, c0, 0.5)
    assert np.allclose(a2, a, rtol=1e-9, atol=1e-9 * np.abs(a).max())
    # the viscous part is linear in alpha
    a0 = orc.wc_accel(r, v, m, num[:, 1], num[:, 2], HC, c0, 0.0)
    a1 = orc.wc_accel(r, v, m, num[:, 1], num[:, 2], HC, c0, 1.0)
    assert np.allclose(a - a0, 0.5 * (a1 - a0), rtol=1e-9, atol=1e-10 * np.abs(a).max())
    assert np.abs(a1 - a0).max() > 1e-3 * np.abs(a0).max()


def test_wc_viscosity_dissipates_and_pressure_repels(orc):
    r, v, m = wc_fluid(400, 7)
    c0 = 2.0
    # rho0 far above every density and pmin = 0 switch the pressure off
    num, _ = orc.wc_density(r, m, HC, 1e9, c0, 7.0, 0.0)
    assert np.all(num[:, 2] == 0.0)
    a = orc.wc_accel(r, v, m, num[:, 1], num[:, 2], HC, c0, 0.3)
    assert (m[:, None] * v * a).sum() < 0
    # two receding particles feel nothing, two approaching ones are slowed
    x = np.array([[0.0, 0, 0], [0.5 * HC, 0.1 * HC, 0]])
    n2, _ = orc.wc_density(x, [1.0, 1.0], HC, 1e9, c0, 7.0, 0.0)
    away = np.array([[-1.0, 0.2, 0], [1.0, 0.0, 0.3]])
    assert np.all(orc.wc_accel(x, away, [1.0, 1.0], n2[:, 1], n2[:, 2], HC, c0, 0.3) == 0)
    a = orc.wc_accel(x, -away, [1.0, 1.0], n2[:, 1], n2[:, 2], HC, c0, 0.3)
    dv, da = -away[1] + away[0], a[1] - a[0]
    assert np.dot(dv, da) < 0 and np.dot(da, x[1] - x[0]) > 0
    # with positive pressure (rho0 below the pair density) the pair is pushed apart
    n3, _ = orc.wc_density(x, [1.0, 1.0], HC, 1.0, c0, 7.0, 0.0)
    assert n3[0, 2] > 0
    a = orc.wc_accel(x, np.zeros((2, 3)), [1.0, 1.0], n3[:, 1], n3[:, 2], HC, c0, 0.3)
    assert np.dot(a[0], x[1] - x[0]) < 0 < np.dot(a[1], x[1] - x[0])


def test_config_dam_rest_density(orc):
    # the config workload's mass (rho0 dx^3 over the lattice sum) gives
    # rho0 exactly to every particle with full support on the lattice,
    # and the lattice sum converges to 1 as h/dx grows
    from particlebase_amd import workloads
    w = workloads.sph_dam_break_config(0.08)
    num, _ = orc.wc_density(w.r, w.m, w.h, w.rho0, w.c[0], 7.0, 0.0)
    dx = w.dx
    full = (w.r[:, 0] < 0.4 - 3 * dx) & (w.r[:, 2] < 0.8 - 3 * dx) & (w.fixed == 0)
    assert full.sum() >= 30
    assert np.allclose(num[full, 1], w.rho0, rtol=1e-13, atol=0)
    assert np.all(num[full, 2] < 1e-6)
    assert workloads.m4_lattice_sum(1.3) == pytest.approx(1.0, abs=5e-3)
    assert workloads.m4_lattice_sum(4.0) == pytest.approx(1.0, abs=1e-5)
