"""GPU parity of the MD neighbour-list path (makeNList / calcForce with
useNList) and of initialAnnealing's pieces through the C ABI against the
oracle: the CSR list bit for bit (FP32 distance2p rules, rows ascending),
list forces (FP64 to 1e-12) for a complete list, for a stale list after
motion and against the cell-path forces, the soft-core force, the Euler
step, and 12 steps of the annealing loop (bin, soft core, constTemp,
acceleration, Euler, wrap) against the same loop on the oracle."""
import numpy as np
import pytest

import particlebase_amd as pb
from particlebase_amd import workloads

pytestmark = pytest.mark.gpu
RC2 = 6.25


def say(msg):
    print(f"\n[case] {msg}", flush=True)


def lj_system(ctx, seed=1, jitter=0.05):
    w = workloads.lj_config1()
    rng = np.random.default_rng(seed)
    w.r = (w.r + rng.normal(scale=jitter, size=w.r.shape)) % w.L
    s = pb.System(ctx, [w.n])
    s.set_cell(w.cell)
    s.import_particles(w.r, w.m, v=w.v, type=w.type)
    s.lj_set_params([1.0, 1.0], 1)
    s.lj_set_cutoff(RC2, shifted=True)
    return w, s


def test_nlist_bitexact_lj(ctx, orc):
    say("makeNList on the jittered N=4000 LJ lattice, list radius 2.8")
    w, s = lj_system(ctx)
    rt2 = np.float32(2.8 * 2.8)
    s.setup_cells(2.8, periodic=True)
    s.bin()
    nnz = s.nlist_build(rt2)
    row, col = s.nlist_get()
    orow, ocol = orc.nlist_f32(w.r, w.cell, 2.8, rt2)
    assert nnz == len(ocol) and nnz > 70 * w.n
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)


def test_nlist_bitexact_gas_two_types(ctx, orc):
    say("makeNList on a random two-type gas (N=6000), list radius 3.3")
    rng = np.random.default_rng(7)
    L = 30.0
    r = rng.uniform(0, L, (6000, 3))
    cell = np.array([0, L, 0, L, 0, L])
    s = pb.System(ctx, [6000])
    s.set_cell(cell)
    s.import_particles(r, np.ones(6000), type=(rng.uniform(size=6000) < 0.3).astype(np.uint32))
    s.setup_cells(3.3, periodic=True)
    s.bin()
    rt2 = np.float32(3.3 * 3.3)
    s.nlist_build(rt2)
    row, col = s.nlist_get()
    orow, ocol = orc.nlist_f32(r, cell, 3.3, rt2)
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)


def test_list_forces(ctx, orc):
    say("calcForce with useNList: fresh list, vs cell path, stale list after 10 VV steps")
    w, s = lj_system(ctx, seed=3)
    p = orc.lj_params([1.0, 1.0], 1, RC2, True)
    s.setup_cells(2.8, periodic=True)
    s.bin()
    s.nlist_build(np.float32(2.8 * 2.8))
    row, col = s.nlist_get()
    pe, vir = s.lj_force_nlist()
    F = s.get("F")
    Fo, peo, wo = orc.lj_forces_list(w.r, w.type, w.cell, RC2, p, row, col)
    assert np.abs(F - Fo).max() < 1e-12 * np.abs(Fo).max()
    assert pe == pytest.approx(peo, rel=1e-12) and vir == pytest.approx(wo, rel=1e-12)
    s.lj_force(pb.PB_FP64, energies=False)
    assert np.abs(s.get("F") - F).max() < 1e-12 * np.abs(F).max()
    # move without rebuilding the list: forces come from the old list
    s.clear_force(True)
    for _ in range(10):
        s.lj_force(pb.PB_FP64, energies=False)
        s.vv_step(0.005)
        s.treat_periodic()
    x = s.get("r")
    s.lj_force_nlist()
    Fo2, _, _ = orc.lj_forces_list(x, w.type, w.cell, RC2, p, row, col)
    assert np.abs(s.get("F") - Fo2).max() < 1e-12 * np.abs(Fo2).max()


def test_softcore_and_euler(ctx, orc):
    say("soft-core force and Euler step on a random gas (N=3000)")
    rng = np.random.default_rng(11)
    L = 60.0
    n = 3000
    r = rng.uniform(0, L, (n, 3))
    v = rng.normal(size=(n, 3))
    m = rng.uniform(28.0, 32.0, n)
    move = (rng.uniform(size=n) < 0.9).astype(np.uint8)
    cell = np.array([0, L, 0, L, 0, L])
    s = pb.System(ctx, [n])
    s.set_cell(cell)
    s.import_particles(r, m, v=v, is_fixed=1 - move)
    s.setup_cells(6.0, periodic=True)
    s.bin()
    s.softcore_force(5.0, 0.1, 40.0 ** 2)
    F = s.get("F")
    Fo = orc.softcore_forces(r, cell, 5.0, 0.1, 40.0 ** 2)
    assert np.abs(F - Fo).max() < 1e-13 * np.abs(Fo).max()
    assert np.count_nonzero(np.abs(Fo).sum(axis=1)) > n // 4
    s.calc_acceleration()
    s.euler_step(2.0)
    ro, vo = orc.euler_step(r, v, Fo / m[:, None], move, 2.0)
    assert np.abs(s.get("r") - ro).max() < 1e-12 * L
    assert np.abs(s.get("v") - vo).max() < 1e-13 * np.abs(vo).max()


def test_annealing_loop_vs_oracle(ctx, orc):
    say("initialAnnealing soft-core loop, 12 steps, two-type gas N=2000")
    rng = np.random.default_rng(13)
    L, n, dt, steps, rcf = 80.0, 2000, 1.0, 12, 5.0
    r = rng.uniform(0, L, (n, 3))
    m = np.where(rng.uniform(size=n) < 0.2, 32.0, 28.0)
    v = rng.normal(scale=0.003, size=(n, 3))
    cell = np.array([0, L, 0, L, 0, L])
    s = pb.System(ctx, [n])
    s.set_cell(cell)
    s.import_particles(r, m, v=v)
    s.setup_cells(8.0, periodic=True)
    x, u = r.copy(), v.copy()
    minv = 1 / m
    move = np.ones(n, np.uint8)
    for i in range(steps):
        rc = rcf * (i + 1) / steps
        s.bin()
        s.softcore_force(rc, 0.1, 40.0 ** 2)
        s.const_temp()
        s.calc_acceleration()
        s.euler_step(dt)
        s.treat_periodic()
        F = orc.softcore_forces(x, cell, rc, 0.1, 40.0 ** 2)
        F, _, _ = orc.const_temp(u, F, minv)
        x, u = orc.euler_step(x, u, F * minv[:, None], move, dt)
        x = orc.wrap(x, cell)
    d = s.get("r") - x
    d -= L * np.round(d / L)
    assert np.abs(d).max() < 1e-10
    assert np.abs(s.get("v") - u).max() < 1e-10 * np.abs(u).max()
