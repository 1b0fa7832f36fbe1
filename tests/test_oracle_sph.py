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
