"""GPU parity of the DEM path (cudaParticleDEM + cudaParticleRotation) through
the C ABI against the FP64 oracle: the contact CSR bit for bit (rows and
sorted columns by original index) in many-particle and single-particle
cells, contact forces and torques, the tangential history carried over
three calls, the reference driver loop of testDEM5.cu:181-234 (bin, force,
acceleration, gravity, leap-frog, rotation, reflecting walls) over several
steps, and the Cundall-Strack error.  FP64 on both sides: forces to 1e-11
of max|F|."""
import os
import subprocess

import numpy as np
import pytest

import particlebase_amd as pb

pytestmark = pytest.mark.gpu
E, MU, SIG, MU_R = 2.0e5, 0.4, 0.29, 0.1
E_EFF = E * (1 - SIG ** 2)
S2 = (2 - SIG) / 2 / (1 - SIG)
G2 = -np.log(0.8) / np.sqrt(np.pi ** 2 + np.log(0.8) ** 2) * 2


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


def make(ctx, r, rad, fixed, m, v, cell, rcell, single=False):
    s = pb.System(ctx, [r.shape[0]])
    s.set_cell(cell)
    s.import_particles(r, m, v=v, is_fixed=fixed)
    s.dem_set_properties(E, MU, SIG, G2, MU_R, rad)
    s.rot_set_inertia(rad)
    s.setup_cells(rcell, periodic=False, single_particle=single)
    return s


def test_contacts_and_forces(ctx, orc):
    say("DEM contacts (many-particle cells) + forces vs oracle, N=2000")
    r, rad, fixed, m, v, w, cell = packing(2000, 1)
    s = make(ctx, r, rad, fixed, m, v, cell, 2 * rad.max())
    s.set("w", w)
    s.bin()
    nnz = s.dem_calc_force(1e-4)
    row, col, val = s.dem_contacts()
    orow, ocol = orc.dem_contacts(r, rad, 1 - fixed)
    assert nnz == len(ocol) and nnz > 1000
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    F, T, oval, bad = orc.dem_force(r, v, w, rad, 1 / m, 1 - fixed, orow, ocol, np.zeros(nnz),
                                    E_EFF, MU, S2, G2, MU_R, 1e-4)
    assert bad == 0 and s.dem_violations() == 0
    Fg, Tg = s.get("F"), s.get("T")
    assert np.abs(Fg - F).max() < 1e-11 * np.abs(F).max()
    assert np.abs(Tg - T).max() < 1e-11 * np.abs(T).max()
    assert np.allclose(val, oval, rtol=1e-12, atol=0)
    assert np.all(Fg[fixed == 1] == 0) and np.all(Tg[fixed == 1] == 0)


def test_history_carry_three_calls(ctx, orc):
    say("DEM tangential history over three calls with moving particles")
    r, rad, fixed, m, v, w, cell = packing(1500, 2)
    s = make(ctx, r, rad, fixed, m, v, cell, 2 * rad.max())
    s.set("w", w)
    move = 1 - fixed
    prow = pcol = pval = None
    rng = np.random.default_rng(5)
    x = r.copy()
    for call in range(3):
        s.set("r", x)
        s.bin()
        s.dem_calc_force(1e-3)
        row, col, val = s.dem_contacts()
        orow, ocol = orc.dem_contacts(x, rad, move)
        assert np.array_equal(row, orow) and np.array_equal(col, ocol), call
        cval = (np.zeros(len(ocol)) if prow is None
                else orc.dem_carry(prow, pcol, pval, orow, ocol))
        F, T, oval, _ = orc.dem_force(x, v, w, rad, 1 / m, move, orow, ocol, cval, E_EFF, MU, S2,
                                      G2, MU_R, 1e-3)
        assert np.allclose(val, oval, rtol=1e-11, atol=1e-15), call
        assert np.abs(s.get("F") - F).max() < 1e-11 * np.abs(F).max(), call
        prow, pcol, pval = orow, ocol, oval
        x = x + rng.normal(scale=0.01, size=x.shape) * move[:, None]
        x = np.clip(x, 0.0, 8.0 - 1e-9)
    # some contacts persisted across calls and carried a history
    assert np.count_nonzero(pval) > 100


def test_single_particle_cells(ctx, orc):
    say("DEM single-particle cells (125 neighbours) on a jittered lattice")
    R0, nside = 0.5, 10
    g = (np.arange(nside) + 0.5) * 0.98
    r = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    r += np.random.default_rng(3).uniform(-0.005, 0.005, r.shape)
    n = r.shape[0]
    rad = np.full(n, R0)
    box = nside * 0.98
    cell = np.array([0.0, box, 0.0, box, 0.0, box])
    # cells whose diagonal is at most one diameter (cudaCutoffBlock.cu:115-138)
    s = make(ctx, r, rad, np.zeros(n, np.uint8), np.ones(n), np.zeros((n, 3)), cell,
             R0 * 2 / np.sqrt(3.0), single=True)
    bn, bl = s.cell_info()
    # precondition of the 125-cell search: two cells span the contact distance
    assert 2 * bl[:3].min() >= 2 * R0
    s.bin()
    s.dem_calc_force(1e-4)
    row, col, _ = s.dem_contacts()
    orow, ocol = orc.dem_contacts(r, rad, np.ones(n, np.uint8))
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)
    assert len(ocol) > 4 * n


def lattice(nside, pitch, seed):
    rng = np.random.default_rng(seed)
    g = (np.arange(nside) + 0.5) * pitch
    r = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    n = r.shape[0]
    r += rng.uniform(-0.01, 0.01, r.shape)
    rad = rng.uniform(0.49, 0.51, n)
    fixed = (r[:, 2] < pitch).astype(np.uint8)           # the bottom layer is a floor
    m = rng.uniform(0.8, 1.2, n)
    v = rng.normal(scale=0.1, size=(n, 3)) * (1 - fixed)[:, None]
    box = nside * pitch
    return r, rad, fixed, m, v, np.array([0.0, box, 0.0, box, 0.0, box])


def test_reference_loop(ctx, orc):
    say("DEM reference loop (testDEM5 order), 15 steps vs oracle")
    r, rad, fixed, m, v, cell = lattice(8, 0.98, 4)
    s = make(ctx, r, rad, fixed, m, v, cell, 2 * rad.max())
    dt, g = 2e-4, 9.8
    move = (1 - fixed).astype(np.uint8)
    minv = 1 / m
    m0inv = minv / (0.4 * rad ** 2)
    x, u = r.copy(), v.copy()
    om, L = np.zeros((len(r), 3)), np.zeros((len(r), 3))
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
        cval = (np.zeros(len(ocol)) if prow is None
                else orc.dem_carry(prow, pcol, pval, orow, ocol))
        F, T, pval, _ = orc.dem_force(x, u, om, rad, minv, move, orow, ocol, cval, E_EFF, MU, S2,
                                      G2, MU_R, dt)
        prow, pcol = orow, ocol
        a = F * minv[:, None]
        a[:, 2] -= g
        x, u = orc.lf_step(x, u, a, move, dt)
        om, L = orc.rot_step(om, L, T, m0inv, move, dt)
        u = np.where(x < lo, np.abs(u), u)
        u = np.where(x > hi, -np.abs(u), u)
    assert np.abs(s.get("r") - x).max() < 1e-11
    assert np.abs(s.get("v") - u).max() < 1e-9 * np.abs(u).max()
    assert np.abs(s.get("w") - om).max() < 1e-9 * max(np.abs(om).max(), 1e-30)
    assert np.abs(s.get("L") - L).max() < 1e-9 * max(np.abs(L).max(), 1e-30)


def test_cundall_strack_error(ctx):
    say("DEM time step above the Cundall-Strack bound")
    r, rad, fixed, m, v, w, cell = packing(500, 6)
    s = make(ctx, r, rad, fixed, m, v, cell, 2 * rad.max())
    s.bin()
    with pytest.raises(pb.PBError) as e:
        s.dem_calc_force(10.0)
    assert e.value.code == pb._lib.PB_ERR_DT
    s.dem_calc_force(10.0, autotune=True)
    assert s.dem_violations() > 0


def test_long_rows(ctx, orc):
    say("DEM contacts with rows longer than the count kernel's row buffer (16), N=600")
    r, rad, fixed, m, v, w, cell = packing(600, 9, box=4.0)
    s = make(ctx, r, rad, fixed, m, v, cell, 2 * rad.max())
    s.bin()
    nnz = s.dem_calc_force(1e-9, autotune=True)
    row, col, _ = s.dem_contacts()
    orow, ocol = orc.dem_contacts(r, rad, 1 - fixed)
    lens = np.diff(orow)
    assert (lens > 16).sum() > 100 and (lens[lens > 0] <= 16).sum() > 0
    assert nnz == len(ocol)
    assert np.array_equal(row, orow) and np.array_equal(col, ocol)


BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "particlebase_amd", "bin")


def _read_iterations(path, n, niter):
    out = []
    with open(path, "rb") as f:
        for _ in range(niter):
            st = [np.frombuffer(f.read(24 * n), dtype=np.float64).reshape(n, 3) for _ in range(4)]
            nnz = int(np.frombuffer(f.read(4), dtype=np.uint32)[0])
            if nnz == 0xFFFFFFFF:
                out.append((st, None))
                continue
            row = np.frombuffer(f.read(4 * (n + 1)), dtype=np.uint32)
            col = np.frombuffer(f.read(4 * nnz), dtype=np.uint32)
            val = np.frombuffer(f.read(8 * nnz), dtype=np.float64)
            out.append((st, (row, col, val)))
    return out


@pytest.mark.parametrize("script", [[0, 0, 0, 1, 0, 0, 0, 0], [1, 0, 0, 0, 0]],
                         ids=["rollback_at_step3", "rollback_of_first_matrix"])
def test_rollback_through_adaptive_time(orc, tmp_path, script):
    """cudaParticleDEM::rollback / rollback2 (cudaParticleDEM.cu:302-310) as
    AdaptiveTime::Progress calls them (AdaptiveTime.hh:33-86): leap-frog
    rollback, rotation step of -dt and the contact matrix rewound (the next
    calcForce carries the history of the step before the discarded one),
    then rollback2 and dt / 4.  The class layer runs the loop on the GPU
    (tests/cpp/dem_rollback_check.cpp); the oracle runs the same sequence."""
    say(f"DEM rollback through AdaptiveTime, script {script}")
    r, rad, fixed, m, v, cell = lattice(7, 0.98, 11)
    n, niter = r.shape[0], len(script)
    f32 = lambda x: float(np.float32(x))
    dt0, g = f32(2e-4), f32(9.8)
    E_, MU_, SIG_, G2_, MUR_ = map(f32, (E, MU, SIG, G2, MU_R))
    rad32 = rad.astype(np.float32)
    rcell = f32(2 * rad32.max())
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    with open(inp, "wb") as f:
        f.write(np.array([n, niter], dtype=np.uint32).tobytes())
        for x in (r, v, m):
            f.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())
        f.write(rad32.tobytes())
        f.write(fixed.astype(np.uint8).tobytes())
        f.write(np.array(script, dtype=np.int32).tobytes())
        f.write(np.array([dt0, g, *cell, E_, MU_, SIG_, G2_, MUR_, rcell],
                         dtype=np.float32).tobytes())
    exe = os.path.join(BIN, "dem_rollback_check")
    p = subprocess.run([exe, str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert p.returncode == 0, p.stderr[-2000:]
    assert "rollbacks=1" in p.stdout, p.stdout
    gpu = _read_iterations(outp, n, niter)

    # the same loop on the oracle, in FP64 from the float-rounded inputs
    radd = rad32.astype(np.float64)
    e_eff = E_ * (1 - SIG_ ** 2)
    s2 = (2 - SIG_) / 2 / (1 - SIG_)
    move = (1 - fixed).astype(np.uint8)
    minv = 1 / m
    m0inv = minv * (1.0 / (0.4 * radd * radd))     # setInertia's arithmetic
    lo, hi = cell[0::2], cell[1::2]
    x, u = r.copy(), v.copy()
    om, L = np.zeros((n, 3)), np.zeros((n, 3))
    mats = []
    dt, rolling = dt0, False
    carried = 0
    for it, res in enumerate(script):
        row, col = orc.dem_contacts(x, radd, move)
        if mats:
            cval = orc.dem_carry(*mats[-1], row, col)
            carried += np.count_nonzero(cval)
        else:
            cval = np.zeros(len(col))
        F, T, val, _ = orc.dem_force(x, u, om, radd, minv, move, row, col, cval, e_eff, MU_, s2,
                                     G2_, MUR_, dt)
        mats.append((row, col, val))
        a = F * minv[:, None]
        a[:, 2] -= g
        x, u = orc.lf_step(x, u, a, move, dt)
        om, L = orc.rot_step(om, L, T, m0inv, move, dt)
        if not rolling:
            if res == 1:      # rollback(dt): LF rollback, rotation -dt, matrix rewound
                x, u = orc.lf_rollback(x, u, a, move, dt)
                om, L = orc.rot_step(om, L, T, m0inv, move, -dt)
                mats.pop()
                rolling = True
        else:                 # rollback2(dt): F = 0, v -= 0, rotation -dt; dt / 4
            om, L = orc.rot_step(om, L, T, m0inv, move, -dt)
            rolling = False
            dt = f32(dt / 4.0)
        u = np.where(x < lo, np.abs(u), u)
        u = np.where(x > hi, -np.abs(u), u)
        (gr, gv, gw, gL), csr = gpu[it]
        rel = lambda a, b: np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)
        errs = {"r": rel(gr, x), "v": rel(gv, u), "w": rel(gw, om), "L": rel(gL, L)}
        assert max(errs.values()) <= 1e-11, (it, errs)
        if not mats:
            assert csr is None, it
            continue
        orow, ocol, oval = mats[-1]
        assert csr is not None, it
        assert np.array_equal(csr[0], orow) and np.array_equal(csr[1], ocol), it
        assert rel(csr[2], oval) <= 1e-11, (it, rel(csr[2], oval))
    assert carried > 100      # persistent contacts carried a tangential history


def test_rewind_errors(ctx):
    say("pb_dem_rewind: before any contact matrix and twice in a row")
    r, rad, fixed, m, v, w, cell = packing(300, 12)
    s = make(ctx, r, rad, fixed, m, v, cell, 2 * rad.max())
    with pytest.raises(pb.PBError) as e:
        s.dem_rewind()
    assert e.value.code == pb._lib.PB_ERR_STATE
    s.bin()
    s.dem_calc_force(1e-4)
    s.dem_calc_force(1e-4)
    s.dem_rewind()
    with pytest.raises(pb.PBError) as e:
        s.dem_rewind()
    assert e.value.code == pb._lib.PB_ERR_STATE
    s.dem_calc_force(1e-4)
    s.dem_rewind()
