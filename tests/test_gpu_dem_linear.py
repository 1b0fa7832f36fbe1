"""GPU parity of the granular settle workload's contact law (linear
spring-dashpot normal force, tangential spring with a Coulomb cap, rotation;
pb_dem_set_linear) through the C ABI against the FP64 oracle's restatement
of the published method: contact forces, torques and the 3-vector
tangential history over three calls with moving particles, the driver loop
of the settle (bin, contacts + force, acceleration, gravity, leap-frog,
rotation, reflecting walls) over 15 steps, and the full workload size
(200,000 spheres from random sequential insertion, moving as a granular
gas) with the contact CSR bit-exact against the oracle's cell-grid search.
FP64 on both sides: forces to 1e-11 of max|F|."""
import numpy as np
import pytest

import particlebase_amd as pb
from particlebase_amd import workloads

pytestmark = pytest.mark.gpu
KN, E_REST, MU = 1.0e4, 0.8, 0.5
KT = KN * 2.0 / 7.0


def say(msg):
    print(f"\n[case] {msg}", flush=True)


def packing(n, seed, box=8.0, rmin=0.4, rmax=0.6, fixed_frac=0.15):
    rng = np.random.default_rng(seed)
    r = rng.uniform(0.0, box, (n, 3))
    rad = rng.uniform(rmin, rmax, n)
    fixed = (rng.uniform(size=n) < fixed_frac).astype(np.uint8)
    m = rng.uniform(0.5, 2.0, n)
    v = rng.normal(scale=0.2, size=(n, 3)) * (1 - fixed)[:, None]
    w = rng.normal(scale=0.5, size=(n, 3)) * (1 - fixed)[:, None]
    cell = np.array([0.0, box, 0.0, box, 0.0, box])
    return r, rad, fixed, m, v, w, cell


def make(ctx, r, rad, fixed, m, v, cell, rcell):
    s = pb.System(ctx, [r.shape[0]])
    s.set_cell(cell)
    s.import_particles(r, m, v=v, is_fixed=fixed)
    s.dem_set_linear(KN, E_REST, KT, MU, rad)
    s.rot_set_inertia(rad)
    s.setup_cells(rcell, periodic=False)
    return s


def test_linear_law_three_calls(ctx, orc):
    say("DEM linear law: forces, torques, 3-vector history over three calls, N=2000")
    r, rad, fixed, m, v, w, cell = packing(2000, 21)
    s = make(ctx, r, rad, fixed, m, v, cell, 2 * rad.max())
    assert s.dem_history_width() == 3
    s.set("w", w)
    move = 1 - fixed
    zeta = orc.lsd_zeta(E_REST)
    rng = np.random.default_rng(8)
    x = r.copy()
    prow = pcol = pval = None
    dt = 1e-4
    for call in range(3):
        s.set("r", x)
        s.bin()
        nnz = s.dem_calc_force(dt)
        row, col, val = s.dem_contacts()
        orow, ocol = orc.dem_contacts(x, rad, move)
        assert nnz == len(ocol) and np.array_equal(row, orow) and np.array_equal(col, ocol), call
        cval = (np.zeros((len(ocol), 3)) if prow is None
                else orc.dem_carry_w(prow, pcol, pval, orow, ocol, 3).reshape(-1, 3))
        F, T, oval, bad = orc.dem_force_lsd(x, v, w, rad, 1 / m, move, orow, ocol, cval, KN, zeta,
                                            KT, MU, dt)
        assert bad == 0 and s.dem_violations() == 0
        Fg, Tg = s.get("F"), s.get("T")
        assert np.abs(Fg - F).max() < 1e-11 * np.abs(F).max(), call
        assert np.abs(Tg - T).max() < 1e-11 * np.abs(T).max(), call
        assert np.abs(val - oval).max() < 1e-11 * np.abs(oval).max(), call
        assert np.all(Fg[fixed == 1] == 0) and np.all(Tg[fixed == 1] == 0)
        prow, pcol, pval = orow, ocol, oval
        x = np.clip(x + rng.normal(scale=0.01, size=x.shape) * move[:, None], 0.0, 8.0 - 1e-9)
    # histories persisted and some contacts slid at the Coulomb cap
    assert np.count_nonzero(np.abs(pval).sum(1)) > 100


def test_linear_settle_loop(ctx, orc):
    say("DEM linear law: settle loop (bin, force, gravity, leap-frog, rotation, walls), 15 steps")
    rng = np.random.default_rng(4)
    nside, pitch = 8, 0.98
    g1 = (np.arange(nside) + 0.5) * pitch
    r = np.stack(np.meshgrid(g1, g1, g1, indexing="ij"), -1).reshape(-1, 3)
    n = r.shape[0]
    r += rng.uniform(-0.01, 0.01, r.shape)
    rad = rng.uniform(0.49, 0.51, n)
    fixed = (r[:, 2] < pitch).astype(np.uint8)
    m = rng.uniform(0.8, 1.2, n)
    v = rng.normal(scale=0.1, size=(n, 3)) * (1 - fixed)[:, None]
    box = nside * pitch
    cell = np.array([0.0, box, 0.0, box, 0.0, box])
    s = make(ctx, r, rad, fixed, m, v, cell, 2 * rad.max())
    dt, g = 2e-3, 9.8
    move = (1 - fixed).astype(np.uint8)
    minv = 1 / m
    m0inv = minv / (0.4 * rad ** 2)
    zeta = orc.lsd_zeta(E_REST)
    x, u = r.copy(), v.copy()
    om, L = np.zeros((n, 3)), np.zeros((n, 3))
    lo, hi = cell[0::2], cell[1::2]
    prow = pcol = pval = None
    for _ in range(15):
        s.bin()
        s.dem_calc_force(dt, autotune=True)
        s.calc_acceleration()
        s.add_acceleration(2, -g)
        s.lf_step(dt)
        s.rot_step(dt)
        s.treat_reflect()
        orow, ocol = orc.dem_contacts(x, rad, move)
        cval = (np.zeros((len(ocol), 3)) if prow is None
                else orc.dem_carry_w(prow, pcol, pval, orow, ocol, 3).reshape(-1, 3))
        F, T, pval, _ = orc.dem_force_lsd(x, u, om, rad, minv, move, orow, ocol, cval, KN, zeta,
                                          KT, MU, dt)
        prow, pcol = orow, ocol
        a = F * minv[:, None]
        a[:, 2] -= g
        x, u = orc.lf_step(x, u, a, move, dt)
        om, L = orc.rot_step(om, L, T, m0inv, move, dt)
        u = np.where(x < lo, np.abs(u), u)
        u = np.where(x > hi, -np.abs(u), u)
    assert len(pcol) > n                              # a dense bed: many contacts
    assert np.abs(s.get("r") - x).max() < 1e-11
    assert np.abs(s.get("v") - u).max() < 1e-9 * np.abs(u).max()
    assert np.abs(s.get("w") - om).max() < 1e-9 * max(np.abs(om).max(), 1e-30)
    assert np.abs(s.get("L") - L).max() < 1e-9 * max(np.abs(L).max(), 1e-30)


def test_config_size_200k(ctx, orc):
    """The workload at its full size: 200,000 spheres, radii U[0.9, 1.1] mm,
    random non-overlapping insertion in 0.1 x 0.1 x 0.4 m, given gas-like
    velocities (0.3 m/s) so that contacts form within a short run; after 150
    steps of the settle loop, one more force call is checked against the
    oracle: the contact CSR bit for bit (cell-grid search on the same FP64
    positions), forces, torques and histories carried from the GPU's
    previous matrix, and sum F = 0 (every contact is between two movable
    spheres)."""
    say("DEM linear law at 200,000 spheres (config size), 150 steps then parity")
    w = workloads.dem_settle_config(vscale=0.3)
    n = w.n
    s = make(ctx, w.r, w.rad, w.fixed, w.m, w.v, w.cell, 2 * w.rad.max())
    dt = 2e-5
    for _ in range(150):
        s.bin()
        s.dem_calc_force(dt)
        s.calc_acceleration()
        s.add_acceleration(2, -9.81)
        s.lf_step(dt)
        s.rot_step(dt)
        s.treat_reflect()
    prow, pcol, pval = s.dem_contacts()
    x, u, om = s.get("r"), s.get("v"), s.get("w")
    s.bin()
    nnz = s.dem_calc_force(dt)
    row, col, val = s.dem_contacts()
    move = np.ones(n, np.uint8)
    orow, ocol = orc.dem_contacts_cells(x, w.rad, move, w.cell, 2 * w.rad.max())
    assert nnz == len(ocol) and nnz > 1000
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    cval = orc.dem_carry_w(prow, pcol, pval, orow, ocol, 3).reshape(-1, 3)
    F, T, oval, bad = orc.dem_force_lsd(x, u, om, w.rad, 1 / w.m, move, orow, ocol, cval, KN,
                                        orc.lsd_zeta(E_REST), KT, MU, dt)
    assert bad == 0
    Fg, Tg = s.get("F"), s.get("T")
    assert np.abs(Fg - F).max() < 1e-11 * np.abs(F).max()
    assert np.abs(Tg - T).max() < 1e-11 * np.abs(T).max()
    assert np.abs(val - oval).max() < 1e-11 * np.abs(oval).max()
    assert np.count_nonzero(np.abs(cval).sum(1)) > 100        # histories were carried
    assert np.abs(Fg.sum(0)).max() < 1e-9 * np.abs(Fg).sum()
