"""ctypes wrapper of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

Imported only by tests/, __graft_entry__.smoke() (as the checker) and
bench.py's cpu_baseline leg.  The product (particlebase_amd/) never imports
it.  See pb_oracle.h for the reference file:line each function restates.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
MAXT = 5

_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P = C.POINTER
        u32, d, f, i32 = C.c_uint32, C.c_double, C.c_float, C.c_int
        L.pbo_setup_cells.argtypes = [P(f), f, i32, P(u32), P(f)]
        L.pbo_setup_cells.restype = i32
        L.pbo_neighbor_table.argtypes = [P(u32), i32, i32, P(u32)]
        L.pbo_calc_bid.argtypes = [P(f), u32, P(f), P(u32), P(u32)]
        L.pbo_sort_bins.argtypes = [P(u32), u32, u32, P(u32), P(u32), P(u32)]
        L.pbo_lj_make_params.argtypes = [P(d), i32, d, i32, C.c_void_p]
        L.pbo_lj_forces_cells.argtypes = [P(d), P(C.c_uint16), u32, P(d), P(u32), P(d), d,
                                          C.c_void_p, P(d), P(d), P(d)]
        L.pbo_lj_forces_cells.restype = i32
        L.pbo_lj_forces_brute.argtypes = [P(d), P(C.c_uint16), u32, P(d), d, C.c_void_p,
                                          P(d), P(d), P(d)]
        L.pbo_vv_step.argtypes = [P(d), P(d), P(d), P(d), P(d), u32, d]
        L.pbo_wrap.argtypes = [P(d), u32, P(d)]
        L.pbo_mv2.argtypes = [P(d), P(d), u32]
        L.pbo_mv2.restype = d
        L.pbo_lj_run.argtypes = [P(d), P(d), P(d), P(d), P(C.c_uint16), u32, P(d), d, d,
                                 C.c_void_p, d, u32, P(d), i32]
        L.pbo_lj_run.restype = i32
        L.pbo_coord_counts_f32.argtypes = [P(f), u32, P(f), f, f, P(u32)]
        L.pbo_coord_counts_f32.restype = i32
        L.pbo_distance2p.argtypes = [f] * 9
        L.pbo_distance2p.restype = f
        L.pbo_gr_hist_f32.argtypes = [P(f), u32, P(f), f, u32, P(C.c_uint64)]
        L.pbo_gr_setup.argtypes = [f, f, P(f), P(u32)]
        L.pbo_gr_normalize.argtypes = [P(C.c_uint64), u32, f, P(f), u32, P(f)]
        _extra(L)
        _lib = L
    return _lib


def _extra(L):
    """Signatures of the SPH / DEM / thermostat restatements (pb_oracle_sph.c,
    _dem.c, _thermo.c)."""
    P = C.POINTER
    u32, d, u8 = C.c_uint32, C.c_double, C.c_uint8
    for name, res, args in [
        ("pbo_sph_w0", d, [d]),
        ("pbo_sph_lucy_w", d, [d, d, d]),
        ("pbo_sph_lucy_dw", d, [d, d, d]),
        ("pbo_sph_density", None, [P(d), P(d), P(d), u32, d, P(d), P(d)]),
        ("pbo_sph_accel", None, [P(d), P(d), P(d), P(d), P(d), P(d), u32, d, d, P(d)]),
        ("pbo_wc_w", d, [d, d]),
        ("pbo_wc_dw", d, [d, d]),
        ("pbo_wc_density", None, [P(d), P(d), u32, d, d, d, d, d, P(d), P(d)]),
        ("pbo_wc_accel", None, [P(d), P(d), P(d), P(d), P(d), u32, d, d, d, P(d)]),
        ("pbo_lf_step", None, [P(d), P(d), P(d), P(u8), u32, d]),
        ("pbo_lf_rollback", None, [P(d), P(d), P(d), P(u8), u32, d]),
        ("pbo_dem_contacts", u32, [P(d), P(d), P(u8), u32, P(u32), P(u32)]),
        ("pbo_dem_carry", None, [P(u32), P(u32), P(d), P(u32), P(u32), P(d), u32]),
        ("pbo_dem_force", u32, [P(d), P(d), P(d), P(d), P(d), P(u8), u32, P(u32), P(u32),
                                P(d), d, d, d, d, d, d, P(d), P(d)]),
        ("pbo_rot_step", None, [P(d), P(d), P(d), P(d), P(u8), u32, d]),
        ("pbo_dem_force_lsd", u32, [P(d), P(d), P(d), P(d), P(d), P(u8), u32, P(u32), P(u32),
                                    P(d), d, d, d, d, d, P(d), P(d)]),
        ("pbo_dem_contacts_cells", u32, [P(d), P(d), P(u8), u32, P(d), d, P(u32), P(u32)]),
        ("pbo_dem_carry_w", None, [P(u32), P(u32), P(d), P(u32), P(u32), P(d), u32, u32]),
        ("pbo_gt_step", C.c_int, [P(d), P(d), P(d), P(d), P(d), u32, d, P(d), d]),
        ("pbo_adjust_velocities", None, [P(d), u32, d, P(d)]),
        ("pbo_const_temp", d, [P(d), P(d), P(d), u32, P(d)]),
        ("pbo_nlist_f32", C.c_int, [P(C.c_float), u32, P(C.c_float), C.c_float, C.c_float,
                                    P(u32), P(u32), u32, P(u32)]),
        ("pbo_softcore_forces", None, [P(d), u32, P(d), d, d, d, P(d)]),
        ("pbo_lj_forces_list", None, [P(d), P(C.c_uint16), u32, P(d), d, C.c_void_p, P(u32),
                                      P(u32), P(d), P(d), P(d)]),
        ("pbo_euler_step", None, [P(d), P(d), P(d), P(u8), u32, d]),
    ]:
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class LJParams(C.Structure):
    _fields_ = [("ntypes", C.c_int), ("s2", C.c_double * 25), ("e4", C.c_double * 25),
                ("shift", C.c_double * 25)]


def setup_cells(cell6, rmax, periodic=True):
    bn = np.zeros(3, dtype=np.uint32)
    bl = np.zeros(6, dtype=np.float32)
    c = f32(cell6)
    rc = lib().pbo_setup_cells(_p(c, C.c_float), np.float32(rmax), int(periodic),
                               _p(bn, C.c_uint32), _p(bl, C.c_float))
    return rc, bn, bl


def neighbor_table(bn, w=1, periodic=True):
    bn = np.ascontiguousarray(bn, dtype=np.uint32)
    ncell = int(np.prod(bn.astype(np.int64)))
    t = np.empty(ncell * (2 * w + 1) ** 3, dtype=np.uint32)
    lib().pbo_neighbor_table(_p(bn, C.c_uint32), w, int(periodic), _p(t, C.c_uint32))
    return t


def calc_bid(r32, bl, bn):
    r = f32(r32).reshape(-1)
    n = r.size // 3
    out = np.empty(n, dtype=np.uint32)
    lib().pbo_calc_bid(_p(r, C.c_float), n, _p(f32(bl), C.c_float),
                       _p(np.ascontiguousarray(bn, dtype=np.uint32), C.c_uint32),
                       _p(out, C.c_uint32))
    return out


def sort_bins(bid, ncell):
    bid = np.ascontiguousarray(bid, dtype=np.uint32)
    n = bid.size
    pid = np.empty(n, dtype=np.uint32)
    bs = np.empty(n, dtype=np.uint32)
    bindex = np.empty(ncell + 1, dtype=np.uint32)
    lib().pbo_sort_bins(_p(bid, C.c_uint32), n, ncell, _p(pid, C.c_uint32), _p(bs, C.c_uint32),
                        _p(bindex, C.c_uint32))
    return pid, bs, bindex


def lj_params(sigma_eps, ntypes, rc2, shifted=True):
    p = LJParams()
    se = f64(sigma_eps)
    lib().pbo_lj_make_params(_p(se, C.c_double), ntypes, rc2, int(shifted), C.byref(p))
    return p


def lj_forces_cells(r, type_, cell6, rc2, params, rmax_cells=None):
    r = f64(r).reshape(-1)
    n = r.size // 3
    t = np.ascontiguousarray(type_, dtype=np.uint16)
    cell = f64(cell6)
    rmax = math.sqrt(rc2) if rmax_cells is None else rmax_cells
    bn = np.array([int((cell[2 * a + 1] - cell[2 * a]) / rmax) for a in range(3)], dtype=np.uint32)
    bl = np.array([(cell[2 * a + 1] - cell[2 * a]) / bn[a] for a in range(3)] +
                  [cell[0], cell[2], cell[4]])
    F = np.empty(3 * n)
    pe = C.c_double()
    w = C.c_double()
    rc = lib().pbo_lj_forces_cells(_p(r, C.c_double), _p(t, C.c_uint16), n, _p(cell, C.c_double),
                                   _p(bn, C.c_uint32), _p(bl, C.c_double), rc2, C.byref(params),
                                   _p(F, C.c_double), C.byref(pe), C.byref(w))
    if rc:
        raise ValueError("oracle cells invalid")
    return F.reshape(n, 3), pe.value, w.value


def lj_forces_brute(r, type_, cell6, rc2, params):
    r = f64(r).reshape(-1)
    n = r.size // 3
    t = np.ascontiguousarray(type_, dtype=np.uint16)
    cell = f64(cell6)
    F = np.empty(3 * n)
    pe = C.c_double()
    w = C.c_double()
    lib().pbo_lj_forces_brute(_p(r, C.c_double), _p(t, C.c_uint16), n, _p(cell, C.c_double), rc2,
                              C.byref(params), _p(F, C.c_double), C.byref(pe), C.byref(w))
    return F.reshape(n, 3), pe.value, w.value


def vv_step(r, v, F, Fold, minv, dt):
    r, v, F, Fold, minv = (f64(x).copy() for x in (r, v, F, Fold, minv))
    n = minv.size
    lib().pbo_vv_step(_p(r, C.c_double), _p(v, C.c_double), _p(F, C.c_double),
                      _p(Fold, C.c_double), _p(minv, C.c_double), n, dt)
    return r, v, Fold


def wrap(r, cell6):
    r = f64(r).copy()
    lib().pbo_wrap(_p(r, C.c_double), r.size // 3, _p(f64(cell6), C.c_double))
    return r


def mv2(v, minv):
    v = f64(v)
    minv = f64(minv)
    return lib().pbo_mv2(_p(v, C.c_double), _p(minv, C.c_double), minv.size)


def lj_run(r, v, Fold, minv, type_, cell6, rmax_cells, rc2, params, dt, nsteps, nthreads=0):
    r, v, Fold, minv = (f64(x).copy() for x in (r, v, Fold, minv))
    t = np.ascontiguousarray(type_, dtype=np.uint16)
    n = minv.size
    e = np.zeros(3 * max(nsteps, 1))
    rc = lib().pbo_lj_run(_p(r, C.c_double), _p(v, C.c_double), _p(Fold, C.c_double),
                          _p(minv, C.c_double), _p(t, C.c_uint16), n,
                          _p(f64(cell6), C.c_double), rmax_cells, rc2, C.byref(params), dt,
                          nsteps, _p(e, C.c_double), int(nthreads))
    if rc:
        raise ValueError("oracle run failed (cells)")
    return r.reshape(n, 3), v.reshape(n, 3), Fold.reshape(n, 3), e.reshape(-1, 3)[:nsteps]


def coord_counts_f32(r32, cell6, rmax_cells, rmax2t):
    r = f32(r32).reshape(-1)
    n = r.size // 3
    out = np.empty(n, dtype=np.uint32)
    rc = lib().pbo_coord_counts_f32(_p(r, C.c_float), n, _p(f32(cell6), C.c_float),
                                    np.float32(rmax_cells), np.float32(rmax2t),
                                    _p(out, C.c_uint32))
    if rc:
        raise ValueError("oracle cells invalid")
    return out


def distance2p(a, b, c3):
    return lib().pbo_distance2p(*(float(x) for x in a), *(float(x) for x in b),
                                *(float(x) for x in c3))


def gr_setup(rstep, rmax):
    ri = C.c_float()
    rn = C.c_uint32()
    lib().pbo_gr_setup(np.float32(rstep), np.float32(rmax), C.byref(ri), C.byref(rn))
    return ri.value, rn.value


def gr_hist_f32(r32, cell6, rstepinv, rnum):
    r = f32(r32).reshape(-1)
    n = r.size // 3
    h = np.zeros(rnum, dtype=np.uint64)
    lib().pbo_gr_hist_f32(_p(r, C.c_float), n, _p(f32(cell6), C.c_float), np.float32(rstepinv),
                          rnum, _p(h, C.c_uint64))
    return h


def gr_normalize(igr, rstepinv, cell6, n):
    igr = np.ascontiguousarray(igr, dtype=np.uint64)
    rnum = igr.size
    gr = np.zeros(rnum, dtype=np.float32)
    lib().pbo_gr_normalize(_p(igr, C.c_uint64), rnum, np.float32(rstepinv),
                           _p(f32(cell6), C.c_float), n, _p(gr, C.c_float))
    return gr


import math  # noqa: E402  (used by lj_forces_cells)


# ----------------------------------------------------------------- SPH
def sph_w0(h):
    return lib().pbo_sph_w0(h)


def sph_lucy_w(r, h):
    return lib().pbo_sph_lucy_w(r, h, sph_w0(h))


def sph_lucy_dw(r, h):
    return lib().pbo_sph_lucy_dw(r, h, -12.0 * sph_w0(h))


def sph_density(r, m, mu0, h):
    """-> num (n, 3: number density, rho, mu), rhoinv (n)."""
    r, m, mu0 = f64(r), f64(m), f64(mu0)
    n = r.shape[0]
    num = np.zeros((n, 3))
    rhoinv = np.zeros(n)
    d = C.c_double
    lib().pbo_sph_density(_p(r, d), _p(m, d), _p(mu0, d), n, h, _p(num, d), _p(rhoinv, d))
    return num, rhoinv


def sph_accel(r, v, m, c2, rhoinv, mu, h, rho0=1.0):
    r, v, m, c2, rhoinv, mu = map(f64, (r, v, m, c2, rhoinv, mu))
    n = r.shape[0]
    a = np.zeros((n, 3))
    d = C.c_double
    lib().pbo_sph_accel(_p(r, d), _p(v, d), _p(m, d), _p(c2, d), _p(rhoinv, d), _p(mu, d), n,
                        h, rho0, _p(a, d))
    return a


def wc_w(r, h):
    return lib().pbo_wc_w(r, h)


def wc_dw(r, h):
    return lib().pbo_wc_dw(r, h)


def wc_density(r, m, h, rho0, c0, gamma=7.0, pmin=0.0):
    """Config-physics density: -> num (n, 3: sum W, rho, p), rhoinv (n)."""
    r, m = f64(r), f64(np.broadcast_to(np.asarray(m, dtype=np.float64), (len(r),)))
    n = r.shape[0]
    num = np.zeros((n, 3))
    rhoinv = np.zeros(n)
    d = C.c_double
    lib().pbo_wc_density(_p(r, d), _p(m, d), n, h, rho0, c0, gamma, pmin, _p(num, d),
                         _p(rhoinv, d))
    return num, rhoinv


def wc_accel(r, v, m, rho, p, h, c0, alpha=0.1):
    """Config-physics pressure + artificial-viscosity acceleration."""
    r, v, rho, p = map(f64, (r, v, rho, p))
    m = f64(np.broadcast_to(np.asarray(m, dtype=np.float64), (len(r),)))
    n = r.shape[0]
    a = np.zeros((n, 3))
    d = C.c_double
    lib().pbo_wc_accel(_p(r, d), _p(v, d), _p(m, d), _p(rho, d), _p(p, d), n, h, c0, alpha,
                       _p(a, d))
    return a


def lf_step(r, v, a, move, dt):
    r, v, a = f64(r).copy(), f64(v).copy(), f64(a)
    mv = np.ascontiguousarray(move, dtype=np.uint8)
    d = C.c_double
    lib().pbo_lf_step(_p(r, d), _p(v, d), _p(a, d), _p(mv, C.c_uint8), r.shape[0], dt)
    return r, v


def lf_rollback(r, v, a, move, dt):
    r, v, a = f64(r).copy(), f64(v).copy(), f64(a)
    mv = np.ascontiguousarray(move, dtype=np.uint8)
    d = C.c_double
    lib().pbo_lf_rollback(_p(r, d), _p(v, d), _p(a, d), _p(mv, C.c_uint8), r.shape[0], dt)
    return r, v


# ----------------------------------------------------------------- DEM
def dem_contacts(r, rad, move):
    """-> rowptr (n+1), colidx (nnz): strict-overlap contacts of movable i."""
    r, rad = f64(r), f64(rad)
    mv = np.ascontiguousarray(move, dtype=np.uint8)
    n = r.shape[0]
    row = np.zeros(n + 1, dtype=np.uint32)
    d, u32 = C.c_double, C.c_uint32
    nnz = lib().pbo_dem_contacts(_p(r, d), _p(rad, d), _p(mv, C.c_uint8), n, _p(row, u32), None)
    col = np.zeros(max(nnz, 1), dtype=np.uint32)
    lib().pbo_dem_contacts(_p(r, d), _p(rad, d), _p(mv, C.c_uint8), n, _p(row, u32),
                           _p(col, u32))
    return row, col[:nnz]


def dem_contacts_cells(r, rad, move, cell, cellsize):
    """dem_contacts over a cell grid of edge >= cellsize (>= 2 max radius)."""
    r, rad, cell = f64(r), f64(rad), f64(cell)
    mv = np.ascontiguousarray(move, dtype=np.uint8)
    n = r.shape[0]
    row = np.zeros(n + 1, dtype=np.uint32)
    d, u32 = C.c_double, C.c_uint32
    nnz = lib().pbo_dem_contacts_cells(_p(r, d), _p(rad, d), _p(mv, C.c_uint8), n, _p(cell, d),
                                       cellsize, _p(row, u32), None)
    if nnz == 0xFFFFFFFF:
        raise ValueError("dem_contacts_cells: bad cell grid")
    col = np.zeros(max(nnz, 1), dtype=np.uint32)
    lib().pbo_dem_contacts_cells(_p(r, d), _p(rad, d), _p(mv, C.c_uint8), n, _p(cell, d),
                                 cellsize, _p(row, u32), _p(col, u32))
    return row, col[:nnz]


def dem_carry(prow, pcol, pval, crow, ccol):
    n = len(crow) - 1
    prow, pcol, crow, ccol = (np.ascontiguousarray(x, dtype=np.uint32) for x in
                              (prow, pcol, crow, ccol))
    pval = f64(pval)
    cval = np.zeros(max(len(ccol), 1))
    u32, d = C.c_uint32, C.c_double
    pc = pcol if len(pcol) else np.zeros(1, dtype=np.uint32)
    pv = pval if len(pval) else np.zeros(1)
    cc = ccol if len(ccol) else np.zeros(1, dtype=np.uint32)
    lib().pbo_dem_carry(_p(prow, u32), _p(pc, u32), _p(pv, d), _p(crow, u32), _p(cc, u32),
                        _p(cval, d), n)
    return cval[:len(ccol)]


def dem_force(r, v, w, rad, minv, move, row, col, val, E, mu, s2, gamma2, mu_r, dt):
    """-> F (n,3), T (n,3), updated history val, #Cundall-Strack violations."""
    r, v, w, rad, minv = map(f64, (r, v, w, rad, minv))
    mv = np.ascontiguousarray(move, dtype=np.uint8)
    row = np.ascontiguousarray(row, dtype=np.uint32)
    col = np.ascontiguousarray(col if len(col) else np.zeros(1), dtype=np.uint32)
    val = f64(val).copy() if len(val) else np.zeros(1)
    n = r.shape[0]
    F = np.zeros((n, 3))
    T = np.zeros((n, 3))
    d, u32 = C.c_double, C.c_uint32
    bad = lib().pbo_dem_force(_p(r, d), _p(v, d), _p(w, d), _p(rad, d), _p(minv, d),
                              _p(mv, C.c_uint8), n, _p(row, u32), _p(col, u32), _p(val, d),
                              E, mu, s2, gamma2, mu_r, dt, _p(F, d), _p(T, d))
    return F, T, val[:int(row[-1])], bad


def lsd_zeta(e):
    """damping ratio of a linear spring-dashpot with restitution e"""
    le = np.log(e)
    return -le / np.sqrt(np.pi ** 2 + le ** 2)


def dem_force_lsd(r, v, w, rad, minv, move, row, col, hist, kn, zeta, kt, mu, dt):
    """Linear spring-dashpot law -> F (n,3), T (n,3), updated history
    (nnz,3), #contacts above the stability bound."""
    r, v, w, rad, minv = map(f64, (r, v, w, rad, minv))
    mv = np.ascontiguousarray(move, dtype=np.uint8)
    row = np.ascontiguousarray(row, dtype=np.uint32)
    nnz = int(row[-1])
    col = np.ascontiguousarray(col if len(col) else np.zeros(1), dtype=np.uint32)
    h = f64(hist).reshape(-1).copy() if nnz else np.zeros(3)
    n = r.shape[0]
    F = np.zeros((n, 3))
    T = np.zeros((n, 3))
    d, u32 = C.c_double, C.c_uint32
    bad = lib().pbo_dem_force_lsd(_p(r, d), _p(v, d), _p(w, d), _p(rad, d), _p(minv, d),
                                  _p(mv, C.c_uint8), n, _p(row, u32), _p(col, u32), _p(h, d),
                                  kn, zeta, kt, mu, dt, _p(F, d), _p(T, d))
    return F, T, h[:3 * nnz].reshape(nnz, 3), bad


def dem_carry_w(prow, pcol, pval, crow, ccol, hw):
    n = len(crow) - 1
    prow, pcol, crow, ccol = (np.ascontiguousarray(x, dtype=np.uint32) for x in
                              (prow, pcol, crow, ccol))
    pval = f64(pval).reshape(-1)
    cval = np.zeros(max(len(ccol) * hw, 1))
    u32, d = C.c_uint32, C.c_double
    pc = pcol if len(pcol) else np.zeros(1, dtype=np.uint32)
    pv = pval if len(pval) else np.zeros(1)
    cc = ccol if len(ccol) else np.zeros(1, dtype=np.uint32)
    lib().pbo_dem_carry_w(_p(prow, u32), _p(pc, u32), _p(pv, d), _p(crow, u32), _p(cc, u32),
                          _p(cval, d), n, hw)
    return cval[:len(ccol) * hw]


def rot_step(w, L, T, m0inv, move, dt):
    w, L, T, m0inv = f64(w).copy(), f64(L).copy(), f64(T), f64(m0inv)
    mv = np.ascontiguousarray(move, dtype=np.uint8)
    d = C.c_double
    lib().pbo_rot_step(_p(w, d), _p(L, d), _p(T, d), _p(m0inv, d), _p(mv, C.c_uint8),
                       w.shape[0], dt)
    return w, L


# ----------------------------------------------------------- thermostats
def gt_step(r, v, F, Fold, minv, dt, xi, tol=1e-14):
    """GaussianThermo::TimeEvolution (Newton-Raphson form) -> r, v, Fold, xi,
    Newton passes (-1: not converged)."""
    r, v, F, Fold, minv = f64(r).copy(), f64(v).copy(), f64(F), f64(Fold).copy(), f64(minv)
    d = C.c_double
    x = d(xi)
    it = lib().pbo_gt_step(_p(r, d), _p(v, d), _p(F, d), _p(Fold, d), _p(minv, d), minv.size,
                           dt, C.byref(x), tol)
    return r, v, Fold, x.value, it


def adjust_velocities(v, v1):
    """adjustVelocity_F4 / _VD -> v, stats (means[3], variances[3])."""
    v = f64(v).copy()
    st = np.zeros(6)
    d = C.c_double
    lib().pbo_adjust_velocities(_p(v, d), v.shape[0], v1, _p(st, d))
    return v, st


def const_temp(v, F, minv):
    """cudaParticleMD::constTemp -> corrected F, sum m v^2, lambda."""
    v, F, minv = f64(v), f64(F).copy(), f64(minv)
    d = C.c_double
    lam = d()
    mv2 = lib().pbo_const_temp(_p(v, d), _p(F, d), _p(minv, d), minv.size, C.byref(lam))
    return F, mv2, lam.value


# ------------------------------------------ MD neighbour list, annealing
def nlist_f32(r32, cell6, rmax_cells, rmax2t):
    """makeNList -> rowptr (n+1), colidx (nnz) by original index, rows ascending."""
    r = f32(r32).reshape(-1)
    n = r.size // 3
    row = np.zeros(n + 1, dtype=np.uint32)
    z = C.c_uint32()
    c6 = f32(cell6)
    L = lib()
    args = (_p(r, C.c_float), n, _p(c6, C.c_float), np.float32(rmax_cells), np.float32(rmax2t),
            _p(row, C.c_uint32))
    if L.pbo_nlist_f32(*args, None, 0, C.byref(z)):
        raise ValueError("oracle cells invalid")
    col = np.zeros(max(z.value, 1), dtype=np.uint32)
    L.pbo_nlist_f32(*args, _p(col, C.c_uint32), col.size, C.byref(z))
    return row, col[:z.value]


def lj_forces_list(r, type_, cell6, rc2, params, rowptr, colidx):
    r = f64(r)
    n = r.shape[0]
    t = np.ascontiguousarray(type_, dtype=np.uint16)
    F = np.zeros((n, 3))
    pe, w = C.c_double(), C.c_double()
    rp = np.ascontiguousarray(rowptr, dtype=np.uint32)
    ci = np.ascontiguousarray(colidx, dtype=np.uint32)
    if ci.size == 0:
        ci = np.zeros(1, dtype=np.uint32)
    lib().pbo_lj_forces_list(_p(r, C.c_double), _p(t, C.c_uint16), n, _p(f64(cell6), C.c_double),
                             rc2, C.byref(params), _p(rp, C.c_uint32), _p(ci, C.c_uint32),
                             _p(F, C.c_double), C.byref(pe), C.byref(w))
    return F, pe.value, w.value


def softcore_forces(r, cell6, rc, f0, rmax2):
    r = f64(r)
    F = np.zeros_like(r)
    lib().pbo_softcore_forces(_p(r, C.c_double), r.shape[0], _p(f64(cell6), C.c_double), rc, f0,
                              rmax2, _p(F, C.c_double))
    return F


def euler_step(r, v, a, move, dt):
    r, v, a = f64(r).copy(), f64(v).copy(), f64(a)
    mv = np.ascontiguousarray(move, dtype=np.uint8)
    d = C.c_double
    lib().pbo_euler_step(_p(r, d), _p(v, d), _p(a, d), _p(mv, C.c_uint8), r.shape[0], dt)
    return r, v
