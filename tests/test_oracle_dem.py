"""Pins of the DEM oracle (cudaParticleDEM + Rotation) by geometry and
mechanics: contact sets against an independent numpy brute force, the
Hertz law's scaling, damping direction, restitution below 1, the Coulomb
cap of the tangential force, torque = Rij x Fs, Newton's third law for equal
spheres, history carry-over on a scripted contact sequence and the
rotation update."""
import numpy as np
import pytest

E, MU, SIG, MU_R = 2.0e5, 0.4, 0.29, 0.1
S2 = (2 - SIG) / 2 / (1 - SIG)


def gamma2(e):
    return -np.log(e) / np.sqrt(np.pi ** 2 + np.log(e) ** 2) * 2


def random_packing(n, seed, box=6.0, rmin=0.4, rmax=0.6):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0, box, (n, 3))
    rad = rng.uniform(rmin, rmax, n)
    move = (rng.uniform(size=n) > 0.2).astype(np.uint8)
    return r, rad, move


def test_contacts_brute_force(orc):
    r, rad, move = random_packing(300, 1)
    row, col = orc.dem_contacts(r, rad, move)
    d2 = ((r[:, None, :] - r[None, :, :]) ** 2).sum(-1)
    hit = d2 < (rad[:, None] + rad[None, :]) ** 2
    np.fill_diagonal(hit, False)
    hit &= move[:, None].astype(bool)
    ref_row = np.concatenate([[0], np.cumsum(hit.sum(1))])
    assert np.array_equal(row, ref_row)
    assert np.array_equal(col, np.nonzero(hit)[1])       # ascending j per row
    assert row[-1] > 50


def two_spheres(xi, r1=0.5, r2=0.5, v=(0, 0, 0), w1=(0, 0, 0), w2=(0, 0, 0), move2=1,
                d=(1.0, 0.0, 0.0)):
    d = np.asarray(d) / np.linalg.norm(d)
    r = np.array([[0.0, 0, 0], list((r1 + r2 - xi) * d)])
    vel = np.array([[0.0, 0, 0], list(v)])
    w = np.array([list(w1), list(w2)], dtype=float)
    rad = np.array([r1, r2])
    move = np.array([1, move2], dtype=np.uint8)
    return r, vel, w, rad, move


def force(orc, r, v, w, rad, move, val=None, minv=(1.0, 1.0), g2=0.0, dt=1e-4):
    row, col = orc.dem_contacts(r, rad, move)
    if val is None:
        val = np.zeros(len(col))
    F, T, val2, bad = orc.dem_force(r, v, w, rad, np.array(minv), move, row, col, val,
                                    E, MU, S2, g2, MU_R, dt)
    return F, T, val2, bad


def test_hertz_scaling(orc):
    # static normal force ~ xi^(3/2) sqrt(Rbar), repulsive along -x on sphere 0
    f = []
    for xi in (1e-3, 2e-3, 4e-3):
        F, T, _, _ = force(orc, *two_spheres(xi))
        assert F[0, 0] < 0 and np.allclose(F[0, 1:], 0) and np.allclose(T, 0)
        f.append(-F[0, 0])
    assert f[1] / f[0] == pytest.approx(2 ** 1.5, rel=1e-12)
    assert f[2] / f[1] == pytest.approx(2 ** 1.5, rel=1e-12)
    Fa, _, _, _ = force(orc, *two_spheres(2e-3, 0.5, 0.5))
    Fb, _, _, _ = force(orc, *two_spheres(2e-3, 1.0, 1.0))      # Rbar doubles
    assert Fb[0, 0] / Fa[0, 0] == pytest.approx(np.sqrt(2), rel=1e-12)
    # absolute value: 2/3 E sqrt(Rbar xi) xi with Rbar = r1 r2/(r1+r2)
    xi, rb = 2e-3, 0.25
    assert -Fa[0, 0] == pytest.approx(2 / 3 * E * np.sqrt(rb * xi) * xi, rel=1e-12)
    # a fixed partner doubles the force (Q = 2 - move_j)
    Fw, _, _, _ = force(orc, *two_spheres(2e-3, move2=0))
    assert Fw[0, 0] == pytest.approx(2 * Fa[0, 0], rel=1e-14)
    assert np.all(Fw[1] == 0)                       # fixed rows are empty


def test_damping_direction_and_restitution(orc):
    g2 = gamma2(0.8)
    Fs, _, _, _ = force(orc, *two_spheres(2e-3), g2=g2)
    Fc, _, _, _ = force(orc, *two_spheres(2e-3, v=(-0.1, 0, 0)), g2=g2)   # approaching
    Fo, _, _, _ = force(orc, *two_spheres(2e-3, v=(0.1, 0, 0)), g2=g2)    # separating
    assert -Fc[0, 0] > -Fs[0, 0] > -Fo[0, 0]
    # head-on collision integrated with the oracle forces: 0 < e < 1,
    # smaller for stronger damping
    def collide(g):
        x = np.array([[0.0, 0, 0], [1.0 + 1e-9, 0, 0]])
        v = np.array([[0.5, 0, 0], [-0.5, 0, 0]])
        rad, move, w = np.array([0.5, 0.5]), np.array([1, 1], np.uint8), np.zeros((2, 3))
        dt = 2e-6
        for _ in range(200000):
            row, col = orc.dem_contacts(x, rad, move)
            F, _, _, _ = orc.dem_force(x, v, w, rad, np.ones(2), move, row, col,
                                       np.zeros(len(col)), E, MU, S2, g, MU_R, dt)
            v = v + F * dt
            x = x + v * dt
            if x[1, 0] - x[0, 0] > 1.0 and v[1, 0] > v[0, 0]:
                break
        return (v[1, 0] - v[0, 0]) / 1.0
    e1, e2 = collide(gamma2(0.9)), collide(gamma2(0.6))
    assert 0 < e2 < e1 < 1.0
    assert collide(0.0) == pytest.approx(1.0, abs=1e-3)


def test_tangential_cap_torque_and_history(orc):
    # oblique contact: the tangential slip has three non-zero components (a
    # slip with a zero component gets no tangential force, as in the
    # reference functor, kernelinline.hh:296)
    d = np.array([1.0, 2.0, 2.0]) / 3.0
    base = two_spheres(2e-3, v=(0.0, 0.3, 0.2), w1=(0.1, 0.2, 0.3), d=d)
    r, v, w, rad, move = base
    F0, T0, val1, _ = force(orc, r, v, w, rad, move, val=np.zeros(2), dt=1e-3)
    fn = 2 / 3 * E * np.sqrt(0.25 * 2e-3) * 2e-3
    Fn = -fn
    # zero history: no tangential force, only the rolling friction torque
    assert np.allclose(F0[0], -fn * d, rtol=1e-12, atol=1e-12 * fn)
    wn = w[0] / np.linalg.norm(w[0])
    assert np.allclose(T0[0], -MU_R * rad[0] * abs(Fn) * wn, rtol=1e-12)
    # the history grew by |vt| dt (same for both rows here)
    assert np.all(val1 > 0)
    # history beyond xi_s_max: |Fs| = mu F0 exactly (Coulomb cap)
    big = np.full(2, 1.0)
    F1, T1, _, _ = force(orc, r, v, w, rad, move, val=big, dt=1e-3)
    Fs = F1[0] - F0[0]
    assert np.linalg.norm(Fs) == pytest.approx(MU * abs(Fn), rel=1e-12)
    assert abs(np.dot(Fs, d)) < 1e-12 * abs(Fn)                # tangential
    # torque of the tangential force: Rij x Fs, Rij = r_i/(r_i+r_j) (r_j - r_i)
    Rij = rad[0] / rad.sum() * (r[1] - r[0])
    assert np.allclose(T1[0] - T0[0], np.cross(Rij, Fs), rtol=1e-12, atol=1e-12)
    # half the cap: linear in the history
    xsm = MU * S2 * 2e-3
    F2, _, _, _ = force(orc, r, v, w, rad, move, val=np.full(2, 0.5 * xsm), dt=1e-3)
    assert np.linalg.norm(F2[0] - F0[0]) == pytest.approx(0.5 * MU * abs(Fn), rel=1e-12)


def test_newton_third_law_equal_spheres(orc):
    r, v, w, rad, move = two_spheres(3e-3, v=(0.01, 0.3, -0.2), w1=(0.2, -0.1, 0.4),
                                     w2=(-0.3, 0.2, 0.1), d=(2.0, -1.0, 2.0))
    w[:] = 0.0        # rolling friction is not a pair force
    F, T, _, _ = force(orc, r, v, w, rad, move, val=np.full(2, 1e-4), g2=gamma2(0.8))
    assert np.allclose(F[0], -F[1], rtol=1e-12, atol=1e-12)


def test_history_carry_scripted(orc):
    # step 1: (0,1), (0,2); step 2: (0,2), (0,3), (1,0)
    prow = np.array([0, 2, 2, 2, 2], np.uint32)
    pcol = np.array([1, 2], np.uint32)
    pval = np.array([0.5, 0.25])
    crow = np.array([0, 2, 3, 3, 3], np.uint32)
    ccol = np.array([2, 3, 0], np.uint32)
    cval = orc.dem_carry(prow, pcol, pval, crow, ccol)
    assert np.array_equal(cval, [0.25, 0.0, 0.0])
    # step 3: contact (0,2) persists again
    cval3 = orc.dem_carry(crow, ccol, cval, np.array([0, 1, 1, 1, 1], np.uint32),
                          np.array([2], np.uint32))
    assert np.array_equal(cval3, [0.25])


def test_rotation_update(orc):
    rng = np.random.default_rng(4)
    n = 20
    T = rng.normal(size=(n, 3))
    m = rng.uniform(1, 2, n)
    rad = rng.uniform(0.4, 0.6, n)
    move = rng.integers(0, 2, n).astype(np.uint8)
    m0inv = (1 / m) / (0.4 * rad ** 2)
    dt = 0.01
    w, L = orc.rot_step(np.zeros((n, 3)), np.zeros((n, 3)), T, m0inv, move, dt)
    mv = move[:, None]
    assert np.allclose(L, T * dt * mv, rtol=1e-15)
    # omega = L / I with the solid-sphere inertia I = 2/5 m r^2
    assert np.allclose(w, L / (0.4 * m * rad ** 2)[:, None] * mv, rtol=1e-14)


def test_lf_rollback_inverts_two_steps(orc):
    # under a constant acceleration two leap-frog steps from (r0, v0) and the
    # rollback give r0 back and v1 = v0 + a dt (the velocity after the first
    # step); fixed particles never move
    rng = np.random.default_rng(7)
    n, dt = 200, 0.01
    r0 = rng.uniform(0, 5, (n, 3))
    v0 = rng.normal(size=(n, 3))
    a = rng.normal(size=(n, 3))
    move = (rng.uniform(size=n) > 0.1).astype(np.uint8)
    r, v = orc.lf_step(r0, v0, a, move, dt)
    r, v = orc.lf_step(r, v, a, move, dt)
    rb, vb = orc.lf_rollback(r, v, a, move, dt)
    m = move[:, None].astype(float)
    assert np.abs(rb - r0).max() < 1e-13
    assert np.abs(vb - (v0 + a * dt * m)).max() < 1e-13
    assert np.array_equal(rb[move == 0], r0[move == 0])
    # a rollback is not a single backward step: it moves r by two steps
    r1, _ = orc.lf_step(r0, v0, a, move, dt)
    assert np.abs(rb - r1)[move == 1].max() > 1e-4


# ------------------------------------------------ linear spring-dashpot law
# Pins of the granular settle workload's contact law (linear normal spring +
# dashpot, tangential spring with a Coulomb cap; Cundall & Strack 1979) by
# the damped oscillator's closed forms and by contact-point kinematics.
KN, KT, MU_L = 1.0e4, 1.0e4 * 2.0 / 7.0, 0.5


def lsd(orc, r, v, w, rad, move, hist=None, minv=(1.0, 1.0), e=0.8, dt=1e-6):
    row, col = orc.dem_contacts(r, rad, move)
    if hist is None:
        hist = np.zeros((len(col), 3))
    return orc.dem_force_lsd(r, v, w, rad, np.array(minv, float), move, row, col, hist, KN,
                             orc.lsd_zeta(e), KT, MU_L, dt)


def collide(orc, e, move2=1, m=1e-3, v0=0.05, steps_per_tc=4000):
    """Head-on collision of two spheres of radius 1 mm (or one on a fixed
    sphere) integrated with leap-frog; returns (contact time, restitution,
    closed-form contact time)."""
    rad = np.array([1e-3, 1e-3])
    move = np.array([1, move2], dtype=np.uint8)
    minv = np.array([1.0 / m, 1.0 / m])
    meff = 1.0 / (minv[0] + (minv[1] if move2 else 0.0))
    z = orc.lsd_zeta(e)
    w0 = np.sqrt(KN / meff)
    tc = np.pi / (w0 * np.sqrt(1.0 - z * z))
    dt = tc / steps_per_tc
    x = np.array([[0.0, 0, 0], [2.0e-3 + 1e-7, 0, 0]])
    v = np.array([[v0, 0, 0], [0.0, 0, 0]]) if move2 == 0 else \
        np.array([[v0 / 2, 0, 0], [-v0 / 2, 0, 0]])
    w = np.zeros((2, 3))
    vrel0 = v[0, 0] - v[1, 0]
    prow = np.zeros(3, np.uint32)
    pcol = np.zeros(0, np.uint32)
    hist = np.zeros((0, 3))
    t_in = t_out = None
    for k in range(3 * steps_per_tc):
        row, col = orc.dem_contacts(x, rad, move)
        hist = orc.dem_carry_w(prow, pcol, hist, row, col, 3).reshape(-1, 3)
        F, T, hist, _ = orc.dem_force_lsd(x, v, w, rad, minv, move, row, col, hist, KN, z, KT,
                                          MU_L, dt)
        if len(col) and t_in is None:
            t_in = k
        if t_in is not None and not len(col):
            t_out = k
            break
        prow, pcol = row, col
        v = v + F * minv[:, None] * move[:, None] * dt
        x = x + v * move[:, None] * dt
    vrel1 = v[0, 0] - v[1, 0]
    return (t_out - t_in) * dt, -vrel1 / vrel0, tc, dt


@pytest.mark.parametrize("move2", [1, 0])
def test_lsd_collision_time_and_restitution(orc, move2):
    for e in (0.8, 0.5):
        t, eps, tc, dt = collide(orc, e, move2)
        assert abs(t - tc) <= 2 * dt
        assert abs(eps - e) < 2e-3 * e
    t, eps, tc, dt = collide(orc, 1.0, move2)      # no dashpot: elastic
    assert abs(eps - 1.0) < 1e-3
    assert abs(t - tc) <= 2 * dt


def test_lsd_static_spring_and_newton(orc, rng):
    for delta in (1e-6, 3e-5):
        r, v, w, rad, move = two_spheres(delta, 1e-3, 1e-3, d=rng.normal(size=3))
        F, T, h, bad = lsd(orc, r, v, w, rad, move)
        n = (r[0] - r[1]) / np.linalg.norm(r[0] - r[1])
        assert np.allclose(F[0], KN * delta * n, rtol=1e-9, atol=0)
        assert np.allclose(F[1], -F[0], rtol=1e-12, atol=0)
        assert np.all(T == 0) and np.all(h == 0) and bad == 0
    # moving, spinning, unequal spheres: F_0 = -F_1, T_0 / rad_0 = T_1 / rad_1
    r, v, w, rad, move = two_spheres(2e-5, 1.1e-3, 0.9e-3, d=rng.normal(size=3))
    v = rng.normal(size=(2, 3)) * 0.01
    w = rng.normal(size=(2, 3)) * 10.0
    F, T, h, _ = lsd(orc, r, v, w, rad, move, hist=rng.normal(size=(2, 3)) * 0, minv=(2e3, 3e3))
    assert np.allclose(F[0], -F[1], rtol=1e-12, atol=1e-18)
    assert np.allclose(T[0] / rad[0], T[1] / rad[1], rtol=1e-12, atol=1e-18)


def test_lsd_stick_slip_and_torque(orc, rng):
    """Tangential spring: with zero history the first step's force is
    -kt vt dt, where vt is the tangential part of the contact points'
    relative velocity computed from rigid-body kinematics; the spring then
    grows linearly until |Ft| reaches mu |Fn|, after which it slides at the
    cap against vt; the torque is the lever (-rad_i n) times Ft."""
    dt = 1e-5
    r, v, w, rad, move = two_spheres(1e-5, 1.0e-3, 0.8e-3, d=rng.normal(size=3))
    v = rng.normal(size=(2, 3)) * 1e-3
    w = rng.normal(size=(2, 3)) * 2.0
    n = (r[0] - r[1]) / np.linalg.norm(r[0] - r[1])
    pc0 = v[0] + np.cross(w[0], -rad[0] * n)        # contact point on sphere 0
    pc1 = v[1] + np.cross(w[1], rad[1] * n)         # and on sphere 1
    vc = pc0 - pc1
    vt = vc - (vc @ n) * n
    fn = KN * 1e-5 - 2 * orc.lsd_zeta(0.8) * np.sqrt(0.5e-3 * KN) * (vc @ n)
    F, T, h, _ = lsd(orc, r, v, w, rad, move, minv=(1e3, 1e3), dt=dt)
    Ft = F[0] - (F[0] @ n) * n
    assert np.allclose(F[0] @ n, fn, rtol=1e-10)
    assert np.allclose(Ft, -KT * vt * dt, rtol=1e-9, atol=1e-20)
    assert np.allclose(h[0], vt * dt, rtol=1e-9, atol=1e-20)
    assert np.allclose(T[0], np.cross(-rad[0] * n, Ft), rtol=1e-9, atol=1e-22)
    # stick: the history keeps adding vt dt while below the cap (static w, v)
    k_stick = int(MU_L * abs(fn) / (KT * np.linalg.norm(vt) * dt))
    hist = h
    for k in range(2, min(k_stick, 50)):
        F, T, hist, _ = lsd(orc, r, v, w, rad, move, hist=hist, minv=(1e3, 1e3), dt=dt)
        assert np.allclose(hist[0], k * vt * dt, rtol=1e-9, atol=1e-20)
    # slip: a history far beyond the cap gives |Ft| = mu |Fn| against the
    # spring, and the spring is reset to -Ft / kt
    big = -np.cross(n, np.cross(n, rng.normal(size=3)))     # a tangential vector
    big *= 10 * MU_L * abs(fn) / KT / np.linalg.norm(big)
    F, T, h, _ = lsd(orc, r, v, w, rad, move, hist=np.array([big, -big]), minv=(1e3, 1e3),
                     dt=dt)
    Ft = F[0] - (F[0] @ n) * n
    assert np.isclose(np.linalg.norm(Ft), MU_L * abs(F[0] @ n), rtol=1e-12)
    s = big + vt * dt
    assert np.allclose(Ft / np.linalg.norm(Ft), -s / np.linalg.norm(s), rtol=1e-10)
    assert np.allclose(h[0], -Ft / KT, rtol=1e-12)


def test_lsd_history_projection_and_bound(orc, rng):
    # no relative motion: the history loses its normal component only
    r, v, w, rad, move = two_spheres(1e-6, 1e-3, 1e-3, d=rng.normal(size=3))
    n = (r[0] - r[1]) / np.linalg.norm(r[0] - r[1])
    h0 = rng.normal(size=3) * 1e-9
    _, _, h, _ = lsd(orc, r, v, w, rad, move, hist=np.array([h0, -h0]))
    assert abs(h[0] @ n) < 1e-24
    assert np.allclose(h[0], h0 - (h0 @ n) * n, rtol=1e-12, atol=1e-25)
    # stability bound dt < 2 sqrt(meff / kn): meff = 0.5 g for two 1 g spheres
    lim = 2 * np.sqrt(0.5e-3 / KN)
    assert lsd(orc, r, v, w, rad, move, minv=(1e3, 1e3), dt=0.99 * lim)[3] == 0
    assert lsd(orc, r, v, w, rad, move, minv=(1e3, 1e3), dt=1.01 * lim)[3] == 2


def test_carry_width(orc, rng):
    prow = np.array([0, 2, 3, 3], np.uint32)
    pcol = np.array([1, 2, 0], np.uint32)
    pval = rng.normal(size=3)
    crow = np.array([0, 1, 3, 3], np.uint32)
    ccol = np.array([2, 0, 2], np.uint32)
    a = orc.dem_carry(prow, pcol, pval, crow, ccol)
    b = orc.dem_carry_w(prow, pcol, pval, crow, ccol, 1)
    assert np.array_equal(a, b) and np.array_equal(a, [pval[1], pval[2], 0.0])
    p3 = rng.normal(size=(3, 3))
    c3 = orc.dem_carry_w(prow, pcol, p3, crow, ccol, 3).reshape(-1, 3)
    assert np.array_equal(c3, np.array([p3[1], p3[2], [0, 0, 0]]))


def test_contacts_cells_equal_brute_force(orc):
    rng = np.random.default_rng(3)
    for n, box in ((3000, 12.0), (500, 5.0)):
        r = rng.uniform(0, box, (n, 3))
        r[:20] = rng.uniform(-0.3, box + 0.3, (20, 3))     # a few centres outside the box
        r[20:40, 0] = box                                   # on the upper face
        rad = rng.uniform(0.35, 0.6, n)
        move = (rng.uniform(size=n) > 0.1).astype(np.uint8)
        row, col = orc.dem_contacts(r, rad, move)
        cell = np.array([0.0, box, 0.0, box, 0.0, box])
        for cs in (1.2, 1.7):
            rowc, colc = orc.dem_contacts_cells(r, rad, move, cell, cs)
            assert np.array_equal(row, rowc) and np.array_equal(col, colc)
        assert len(col) > n
