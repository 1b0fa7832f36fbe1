"""The C++ drivers (particlebase_amd/cpp/drivers) run the reference's driver
call sequences through the class layer and the C ABI on the GPU: LJ
(testMD1.cu order, per-call API and fused; and the testMD1 gas with
initialAnnealing and the neighbour-list path), liquid argon under the Gaussian
thermostat (testMD2.cu order), checkpoint / restart of MD, DEM and SPH
objects (bit-exact continuation), SPH dam break (testSPH5.cu order with
AdaptiveTime; reference and config physics), DEM settle (testDEM5.cu order
with AdaptiveTime) and the force / acceleration imports of cudaCutoffBlock.
Each must finish and report a physically sane end state."""
import os
import re
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "particlebase_amd", "bin")


def run(*args):
    exe = os.path.join(BIN, args[0])
    if not os.path.exists(exe):
        pytest.fail(f"{exe} not built (make -C particlebase_amd/cpp)")
    print(f"\n[case] driver {' '.join(map(str, args))}", flush=True)
    p = subprocess.run([exe, *map(str, args[1:])], capture_output=True, text=True, timeout=150)
    assert p.returncode == 0, p.stderr[-2000:]
    line = [x for x in p.stdout.splitlines() if x.startswith("done")][-1]
    return dict(re.findall(r"(\w+)=([-+0-9.eE]+)", line))


@pytest.mark.parametrize("mode", [("api", "fp64"), ("fused", "fp64"), ("fused", "mixed")])
def test_md_lj(mode):
    d = run("md_lj", 6, 400, *mode)
    assert int(d["N"]) == 864 and int(d["steps"]) == 400
    # total energy per particle is conserved to well below kT over 400 steps
    assert abs(float(d["dE"])) < 0.02


def test_sph_dam():
    d = run("sph_dam", 0.08, 60)
    assert int(d["steps"]) == 60
    # the column started at x <= 0.4 and starts to spread, walls hold it
    assert 0.3 < float(d["front_x"]) < 1.7
    assert float(d["mean_num"]) > 0


def test_sph_dam_config():
    # the config physics (cubic spline, Tait, artificial viscosity)
    d = run("sph_dam", 0.08, 60, "config")
    assert int(d["steps"]) == 60
    assert 0.3 < float(d["front_x"]) < 1.7
    assert float(d["mean_num"]) > 0


def test_dem_settle():
    d = run("dem_settle", 3000, 200)
    assert int(d["steps"]) == 200
    assert 0.0 < float(d["zmean"]) < 0.4
    # numSelectedBlocks: every sphere moves, so the occupied cells, at
    # least N / (spheres a cell can hold) of them and at most all cells
    sel, cells = int(d["selected_cells"]), int(d["cells"])
    assert 3000 // 8 <= sel <= cells


def test_dem_settle_linear_config(tmp_path):
    """The granular settle workload at its size through the driver: 200,000
    spheres from random non-overlapping insertion (moving as a gas, so that
    contacts form in a short run), the linear spring-dashpot law, adaptive
    time step with its cap."""
    import numpy as np
    from particlebase_amd import workloads
    w = workloads.dem_settle_config(vscale=0.3)
    path = tmp_path / "dem.bin"
    with open(path, "wb") as f:
        f.write(np.array([w.n], dtype=np.uint32).tobytes())
        for x in (w.r, w.rad, w.m, w.v):
            f.write(np.ascontiguousarray(x, dtype=np.float64).tobytes())
    d = run("dem_settle", w.n, 300, "linear", path)
    assert int(d["N"]) == 200_000 and int(d["steps"]) == 300
    assert int(d["max_contacts"]) > 100
    # the step grew from 2e-6 to its cap 0.2 sqrt(meff_min / kn) and no further
    mmin = w.m.min()
    assert 2e-6 < float(d["dt"]) <= 0.2 * np.sqrt(0.5 * mmin / 1e4) * 1.0001
    assert 0.0 < float(d["zmean"]) < 0.4


def test_md_gauss():
    # testMD2 order: liquid argon under the Gaussian thermostat
    d = run("md_gauss", 300)
    assert int(d["N"]) == 4924 and int(d["steps"]) == 300
    assert int(d["failed"]) == 0 and int(d["scaled"]) == 0
    # isokinetic: T after the first thermostatted step is held
    t0 = float(d["T0"])
    assert 70.0 < t0 < 100.0
    assert abs(float(d["T1"]) / t0 - 1) < 0.01
    assert float(d["Tmax"]) / float(d["Tmin"]) < 1.02


@pytest.mark.parametrize("mode", ["cells", "nlist"])
def test_md_gas(mode):
    # testMD1 order: random N2/O2 gas, adjustVelocities, initialAnnealing
    # (shortened to 200 + 200 steps), 100 main-loop steps
    d = run("md_gas", 100, 200, mode)
    assert int(d["N"]) == 10618 and int(d["steps"]) == 100
    # annealing separated the random start (LJ sigma ~ 3.2 A) and set T
    assert float(d["dmin"]) > 2.0
    assert abs(float(d["Tanneal"]) / 353.0 - 1) < 0.05
    assert 300.0 < float(d["Tmin"]) and float(d["Tmax"]) < 1.2 * 353.0


@pytest.mark.parametrize("mode", ["md", "dem", "sph", "sphc"])
def test_restart_exact(mode, tmp_path):
    # writeSerialization after k steps, readSerialization into a fresh
    # object, k more steps: the same end state as 2k uninterrupted steps
    d = run("restart", mode, 15, str(tmp_path / f"{mode}.ck"))
    assert int(d["k"]) == 15 and int(d["N"]) > 500
    assert float(d["maxdiff"]) == 0.0


def test_import_force_and_acceleration():
    # cudaCutoffBlock::importForce / importAcceleration between two objects:
    # the other object's device array and its TMP host array are added
    d = run("import_check")
    assert int(d["ok"]) == 1 and int(d["N"]) == 1000


def test_env_queries_and_stat_mv2():
    # CUDAenv::getShMem / getNumTh / getMaxGrid / getMemSize / getNumCpMP /
    # display and cudaParticleMD::statMV2 (per-particle v^2 listing)
    d = run("env_stat_check")
    assert int(d["ok"]) == 1 and int(d["rows"]) == int(d["N"]) == 500
    # 6 significant digits in the listing (the stream's default precision)
    assert float(d["relerr"]) < 1e-5 and int(d["cp"]) == 64


def test_put_tmp_selected(tmp_path):
    # putTMPselected on an SPH block: the walls through putUnSelected, the
    # fluid through getSelectedTypeID / getSelectedPosition / putTMP with
    # its n, rho, p columns, both parsed back and compared
    d = run("select_check", str(tmp_path))
    assert int(d["ok"]) == 1 and int(d["N1"]) == 12 * 12 * 8
    assert float(d["poserr"]) < 1e-5 and float(d["numerr"]) < 1e-7
    # calcDensityPost + calcAcceleration after a density import use the
    # imported rho and its reciprocal (testSPH5.cu:208-220)
    assert int(d["acc_from_import"]) == 1
