"""ctypes binding of libpbamd.so (include/pbamd.h).

This replaces the reference's SWIG/Python-2 module (cudaParticles.i:1-60):
Python drives the C ABI directly.  The library must have been built
(``__graft_entry__.build()`` or ``make -C particlebase_amd/csrc``); there is
no fallback of any kind -- a missing library raises at import of a system.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import sys
import weakref

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libpbamd.so")

PB_OK = 0
PB_ERR_ARG = -1
PB_ERR_HIP = -2
PB_ERR_STATE = -3
PB_ERR_CELLS = -4
PB_ERR_NLIST = -5
PB_ERR_SKIN = -6
PB_ERR_DT = -7

PB_FP64 = 0
PB_MIXED = 1

FIELD = {"r": 0, "v": 1, "a": 2, "F": 3, "minv": 4, "type": 5, "move": 6,
         "w": 7, "L": 8, "T": 9, "radius": 10, "sph_num": 11, "sph_rhoinv": 12}

# symbols declared in include/pbamd.h (checked by the CPU test suite)
EXPORTS = [
    "pb_ctx_create", "pb_ctx_destroy", "pb_ctx_last_error", "pb_ctx_sync", "pb_ctx_info",
    "pb_ctx_limits",
    "pb_sys_create", "pb_sys_destroy", "pb_sys_count", "pb_sys_replicas", "pb_sys_set_cell",
    "pb_sys_import", "pb_sys_get", "pb_sys_set", "pb_clear_force", "pb_calc_acceleration",
    "pb_add_force", "pb_add_acceleration", "pb_treat_periodic", "pb_treat_absolute",
    "pb_treat_reflect", "pb_vv_step", "pb_lf_vinit", "pb_lf_step", "pb_lf_rollback", "pb_mv2",
    "pb_scale_velocities", "pb_inspect_velocity", "pb_setup_cells", "pb_get_cell_info",
    "pb_bin", "pb_get_bins", "pb_get_neighbor_table", "pb_lj_set_params", "pb_lj_set_cutoff",
    "pb_lj_force", "pb_coord_counts", "pb_lj_run", "pb_lj_last_run_stats",
    "pb_lj_last_run_timing", "pb_lj_autotune", "pb_gr_hist",
    "pb_sph_set_properties", "pb_sph_set_wcsph", "pb_sph_density", "pb_sph_acceleration",
    "pb_sph_inspect_density",
    "pb_dem_set_properties", "pb_rot_set_inertia", "pb_rot_step", "pb_dem_calc_force",
    "pb_dem_get_contacts", "pb_dem_last_violations",
    "pb_gt_step", "pb_gt_xi", "pb_adjust_velocities", "pb_const_temp",
    "pb_nlist_build", "pb_nlist_get", "pb_lj_force_nlist", "pb_softcore_force", "pb_euler_step",
    "pb_dem_set_contacts", "pb_lj_set_tiles", "pb_dem_rewind", "pb_dem_set_linear",
    "pb_dem_history_width", "pb_gr_hist_noalloc", "pb_diag_valu_probe",
]


class pb_run_opts(C.Structure):
    _fields_ = [("dt", C.c_double), ("nsteps", C.c_uint32), ("precision", C.c_int),
                ("skin", C.c_double), ("rebuild_every", C.c_uint32),
                ("energy_every", C.c_uint32), ("use_graph", C.c_int),
                ("block_size", C.c_uint32), ("timing", C.c_int), ("persistent", C.c_int)]


class PBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"pbamd error {code}: {msg}")
        self.code = code


_lib = None


def load(path: str = LIB_PATH) -> C.CDLL:
    """Load libpbamd.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise FileNotFoundError(
            f"{path} not built: run __graft_entry__.build() (no CPU fallback exists)")
    lib = C.CDLL(path)
    vp, u32, i32, d, f = C.c_void_p, C.c_uint32, C.c_int, C.c_double, C.c_float
    P = C.POINTER
    sig = {
        "pb_ctx_create": (i32, [i32, P(vp)]),
        "pb_ctx_destroy": (i32, [vp]),
        "pb_ctx_last_error": (C.c_char_p, [vp]),
        "pb_ctx_sync": (i32, [vp]),
        "pb_ctx_info": (i32, [vp, C.c_char_p, C.c_size_t, P(i32)]),
        "pb_ctx_limits": (i32, [vp, P(C.c_uint64), P(u32), P(u32), P(u32), P(u32)]),
        "pb_sys_create": (i32, [vp, u32, P(u32), P(vp)]),
        "pb_sys_destroy": (None, [vp]),
        "pb_sys_count": (u32, [vp]),
        "pb_sys_replicas": (u32, [vp]),
        "pb_sys_set_cell": (i32, [vp, P(d)]),
        "pb_sys_import": (i32, [vp, P(d), P(d), P(d), P(d), P(u32), P(C.c_uint8)]),
        "pb_sys_get": (i32, [vp, i32, vp]),
        "pb_sys_set": (i32, [vp, i32, vp]),
        "pb_clear_force": (i32, [vp, i32]),
        "pb_calc_acceleration": (i32, [vp]),
        "pb_add_force": (i32, [vp, i32, d]),
        "pb_add_acceleration": (i32, [vp, i32, d]),
        "pb_treat_periodic": (i32, [vp]),
        "pb_treat_absolute": (i32, [vp]),
        "pb_treat_reflect": (i32, [vp]),
        "pb_vv_step": (i32, [vp, d]),
        "pb_lf_vinit": (i32, [vp, d]),
        "pb_lf_step": (i32, [vp, d]),
        "pb_lf_rollback": (i32, [vp, d]),
        "pb_mv2": (i32, [vp, P(d)]),
        "pb_scale_velocities": (i32, [vp, d]),
        "pb_inspect_velocity": (i32, [vp, d, d, d, P(i32), P(d), P(u32)]),
        "pb_setup_cells": (i32, [vp, f, i32, i32]),
        "pb_get_cell_info": (i32, [vp, P(u32), P(f)]),
        "pb_bin": (i32, [vp]),
        "pb_get_bins": (i32, [vp, P(u32), P(u32), P(u32)]),
        "pb_get_neighbor_table": (i32, [vp, P(u32)]),
        "pb_lj_set_params": (i32, [vp, P(d), i32]),
        "pb_lj_set_cutoff": (i32, [vp, d, i32]),
        "pb_lj_force": (i32, [vp, i32, P(d), P(d)]),
        "pb_coord_counts": (i32, [vp, f, P(u32)]),
        "pb_lj_run": (i32, [vp, P(pb_run_opts), P(d)]),
        "pb_lj_last_run_stats": (i32, [vp, P(u32), P(u32), P(i32)]),
        "pb_lj_last_run_timing": (i32, [vp, P(d), P(u32), P(d), P(u32)]),
        "pb_gr_hist": (i32, [vp, P(f), u32, P(C.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    for name, (res, args) in _extra_signatures().items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _extra_signatures():
    """Signatures of the SPH / DEM / thermostat entry points."""
    vp, u32, i32, d = C.c_void_p, C.c_uint32, C.c_int, C.c_double
    P = C.POINTER
    return {
        "pb_sph_set_properties": (i32, [vp, P(d), P(d), d]),
        "pb_sph_set_wcsph": (i32, [vp, d, d, d, d, d, d]),
        "pb_sph_density": (i32, [vp]),
        "pb_sph_acceleration": (i32, [vp]),
        "pb_sph_inspect_density": (i32, [vp, P(d)]),
        "pb_dem_set_properties": (i32, [vp, d, d, d, d, d, P(d), d]),
        "pb_rot_set_inertia": (i32, [vp, P(d), d]),
        "pb_rot_step": (i32, [vp, d]),
        "pb_gt_step": (i32, [vp, d, d]),
        "pb_gt_xi": (i32, [vp, P(d), P(d), P(u32)]),
        "pb_adjust_velocities": (i32, [vp, d, P(d)]),
        "pb_const_temp": (i32, [vp, P(d), P(d)]),
        "pb_nlist_build": (i32, [vp, C.c_float, P(u32)]),
        "pb_nlist_get": (i32, [vp, P(u32), P(u32), u32, P(u32)]),
        "pb_lj_force_nlist": (i32, [vp, P(d), P(d)]),
        "pb_softcore_force": (i32, [vp, d, d, d]),
        "pb_euler_step": (i32, [vp, d]),
        "pb_dem_set_contacts": (i32, [vp, P(u32), P(u32), P(d), u32]),
        "pb_lj_set_tiles": (i32, [vp, u32, d]),
        "pb_dem_calc_force": (i32, [vp, d, i32, P(u32)]),
        "pb_dem_get_contacts": (i32, [vp, P(u32), P(u32), P(d), u32, P(u32)]),
        "pb_dem_last_violations": (i32, [vp, P(u32)]),
        "pb_dem_rewind": (i32, [vp]),
        "pb_dem_set_linear": (i32, [vp, d, d, d, d, P(d), d]),
        "pb_dem_history_width": (i32, [vp, P(u32)]),
        "pb_gr_hist_noalloc": (i32, [vp, P(C.c_float), u32, P(C.c_uint64)]),
        "pb_diag_valu_probe": (i32, [vp, i32, u32, P(d), P(d)]),
        "pb_lj_autotune": (i32, [vp, P(pb_run_opts), u32, P(d), P(u32), P(d)]),
    }


def _p(arr, ctype):
    if arr is None:
        return None
    return arr.ctypes.data_as(C.POINTER(ctype))


_LIVE_CONTEXTS = weakref.WeakSet()


@atexit.register
def _close_all():
    # release device objects before the HIP runtime tears itself down
    for c in list(_LIVE_CONTEXTS):
        try:
            c.close()
        except Exception:
            pass


class Context:
    """One HIP device + one stream (pb_ctx)."""

    def __init__(self, device: int = 0):
        self.lib = load()
        h = C.c_void_p()
        rc = self.lib.pb_ctx_create(device, C.byref(h))
        if rc != PB_OK:
            raise PBError(rc, f"pb_ctx_create(device={device}) failed (is a GPU visible?)")
        self.h = h
        self.systems = weakref.WeakSet()
        _LIVE_CONTEXTS.add(self)

    def check(self, rc):
        if rc != PB_OK:
            raise PBError(rc, self.lib.pb_ctx_last_error(self.h).decode())
        return rc

    def info(self):
        buf = C.create_string_buffer(128)
        cu = C.c_int()
        self.check(self.lib.pb_ctx_info(self.h, buf, 128, C.byref(cu)))
        return buf.value.decode(), cu.value

    def limits(self):
        """Device limits: global memory bytes, threads per block, grid x,
        LDS bytes per block, lanes per compute unit."""
        mem = C.c_uint64()
        th, grid, lds, lanes = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint32()
        self.check(self.lib.pb_ctx_limits(self.h, C.byref(mem), C.byref(th), C.byref(grid),
                                          C.byref(lds), C.byref(lanes)))
        return {"global_mem": mem.value, "max_threads": th.value, "max_grid": grid.value,
                "lds_per_block": lds.value, "lanes_per_cu": lanes.value}

    def sync(self):
        self.check(self.lib.pb_ctx_sync(self.h))

    def valu_probe(self, kind=0, iters=4096):
        """pb_diag_valu_probe: (ms, wave64 VALU instructions) of a launch
        filling every SIMD with FP32 FMAs (kind 0), FP64 FMAs (1), integer
        xor-adds (2) or compare-select-adds (3): issue-rate calibration."""
        ms, wi = C.c_double(), C.c_double()
        self.check(self.lib.pb_diag_valu_probe(self.h, int(kind), int(iters), C.byref(ms),
                                               C.byref(wi)))
        return ms.value, wi.value

    def close(self):
        if getattr(self, "h", None):
            for s in list(self.systems):
                s.close()
            self.check(self.lib.pb_ctx_destroy(self.h))
            self.h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass


class System:
    """A particle system (cudaParticleBase object) or a batch of replicas."""

    def __init__(self, ctx: Context, counts):
        self.ctx = ctx
        self.lib = ctx.lib
        counts = np.ascontiguousarray(np.atleast_1d(counts), dtype=np.uint32)
        self.counts = counts
        self.nrep = int(counts.size)
        self.n = int(counts.sum())
        h = C.c_void_p()
        ctx.check(self.lib.pb_sys_create(ctx.h, self.nrep, _p(counts, C.c_uint32), C.byref(h)))
        self.h = h
        ctx.systems.add(self)

    # --- lifecycle
    def close(self):
        if getattr(self, "h", None):
            if self.ctx.h:
                self.lib.pb_sys_destroy(self.h)
            self.h = None

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def _c(self, rc):
        return self.ctx.check(rc)

    # --- state
    def set_cell(self, cells):
        cells = np.ascontiguousarray(np.asarray(cells, dtype=np.float64).reshape(-1))
        assert cells.size == 6 * self.nrep
        self._c(self.lib.pb_sys_set_cell(self.h, _p(cells, C.c_double)))

    def import_particles(self, r, m, v=None, a=None, type=None, is_fixed=None):
        f64 = lambda x: None if x is None else np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        r, v, a, m = f64(r), f64(v), f64(a), f64(m)
        t = None if type is None else np.ascontiguousarray(type, dtype=np.uint32)
        fx = None if is_fixed is None else np.ascontiguousarray(is_fixed, dtype=np.uint8)
        self._keep = (r, v, a, m, t, fx)
        self._c(self.lib.pb_sys_import(self.h, _p(r, C.c_double), _p(v, C.c_double),
                                       _p(a, C.c_double), _p(m, C.c_double),
                                       _p(t, C.c_uint32), _p(fx, C.c_uint8)))

    def get(self, field: str):
        if field in ("r", "v", "a", "F", "w", "L", "T", "sph_num"):
            out = np.empty((self.n, 3), dtype=np.float64)
        elif field in ("minv", "radius", "sph_rhoinv"):
            out = np.empty(self.n, dtype=np.float64)
        elif field == "type":
            out = np.empty(self.n, dtype=np.uint16)
        elif field == "move":
            out = np.empty(self.n, dtype=np.uint8)
        else:
            raise KeyError(field)
        self._c(self.lib.pb_sys_get(self.h, FIELD[field], out.ctypes.data_as(C.c_void_p)))
        return out

    def set(self, field: str, values):
        dt = {"type": np.uint16, "move": np.uint8}.get(field, np.float64)
        arr = np.ascontiguousarray(values, dtype=dt)
        self._c(self.lib.pb_sys_set(self.h, FIELD[field], arr.ctypes.data_as(C.c_void_p)))

    # --- cudaParticleBase / VV / LF surface
    def clear_force(self, with_fold=True):
        self._c(self.lib.pb_clear_force(self.h, int(with_fold)))

    def calc_acceleration(self):
        self._c(self.lib.pb_calc_acceleration(self.h))

    def add_force(self, axis, value):
        self._c(self.lib.pb_add_force(self.h, axis, value))

    def add_acceleration(self, axis, value):
        self._c(self.lib.pb_add_acceleration(self.h, axis, value))

    def treat_periodic(self):
        self._c(self.lib.pb_treat_periodic(self.h))

    def treat_absolute(self):
        self._c(self.lib.pb_treat_absolute(self.h))

    def treat_reflect(self):
        self._c(self.lib.pb_treat_reflect(self.h))

    def vv_step(self, dt):
        self._c(self.lib.pb_vv_step(self.h, dt))

    def lf_vinit(self, dt):
        self._c(self.lib.pb_lf_vinit(self.h, dt))

    def lf_step(self, dt):
        self._c(self.lib.pb_lf_step(self.h, dt))

    def lf_rollback(self, dt):
        self._c(self.lib.pb_lf_rollback(self.h, dt))

    def mv2(self):
        out = np.empty(self.nrep, dtype=np.float64)
        self._c(self.lib.pb_mv2(self.h, _p(out, C.c_double)))
        return out

    def scale_velocities(self, s):
        self._c(self.lib.pb_scale_velocities(self.h, s))

    # --- thermostats (GaussianThermo, adjustVelocities, constTemp)
    def gt_step(self, dt, tol=1e-8):
        self._c(self.lib.pb_gt_step(self.h, dt, tol))

    def gt_xi(self, set_xi=None):
        """-> (xi per replica, bad step count); set_xi replaces xi afterwards."""
        out = np.empty(self.nrep, dtype=np.float64)
        bad = C.c_uint32()
        sx = None
        if set_xi is not None:
            sx = np.ascontiguousarray(np.broadcast_to(np.asarray(set_xi, dtype=np.float64),
                                                      (self.nrep,)))
        self._c(self.lib.pb_gt_xi(self.h, None if sx is None else _p(sx, C.c_double),
                                  _p(out, C.c_double), C.byref(bad)))
        return out, bad.value

    def adjust_velocities(self, v1):
        """-> stats (nrep, 6): means and variances before the adjustment."""
        st = np.empty((self.nrep, 6), dtype=np.float64)
        self._c(self.lib.pb_adjust_velocities(self.h, v1, _p(st, C.c_double)))
        return st

    def const_temp(self):
        """-> (sum m v^2, lambda) per replica; F corrected in place."""
        mv2 = np.empty(self.nrep, dtype=np.float64)
        lam = np.empty(self.nrep, dtype=np.float64)
        self._c(self.lib.pb_const_temp(self.h, _p(mv2, C.c_double), _p(lam, C.c_double)))
        return mv2, lam

    # --- MD neighbour list (makeNList / calcForce with useNList), annealing
    def nlist_build(self, rmax2t):
        z = C.c_uint32()
        self._c(self.lib.pb_nlist_build(self.h, float(rmax2t), C.byref(z)))
        return z.value

    def nlist_get(self):
        """-> rowptr (n+1), colidx (nnz), by original index."""
        z = C.c_uint32()
        row = np.zeros(self.n + 1, dtype=np.uint32)
        self._c(self.lib.pb_nlist_get(self.h, _p(row, C.c_uint32), None, 0, C.byref(z)))
        col = np.zeros(max(z.value, 1), dtype=np.uint32)
        self._c(self.lib.pb_nlist_get(self.h, None, _p(col, C.c_uint32), col.size, C.byref(z)))
        return row, col[:z.value]

    def lj_force_nlist(self):
        pe, w = C.c_double(), C.c_double()
        self._c(self.lib.pb_lj_force_nlist(self.h, C.byref(pe), C.byref(w)))
        return pe.value, w.value

    def softcore_force(self, rc, f0, rmax2):
        self._c(self.lib.pb_softcore_force(self.h, rc, f0, rmax2))

    def euler_step(self, dt):
        self._c(self.lib.pb_euler_step(self.h, dt))

    def inspect_velocity(self, vlim, lim_u, lim_l):
        res = C.c_int()
        vr = C.c_double()
        im = C.c_uint32()
        self._c(self.lib.pb_inspect_velocity(self.h, vlim, lim_u, lim_l, C.byref(res),
                                             C.byref(vr), C.byref(im)))
        return res.value, vr.value, im.value

    # --- cells
    def setup_cells(self, rmax, periodic=True, single_particle=False):
        self._c(self.lib.pb_setup_cells(self.h, float(rmax), int(periodic), int(single_particle)))

    def cell_info(self):
        bn = np.zeros(3, dtype=np.uint32)
        bl = np.zeros(6, dtype=np.float32)
        self._c(self.lib.pb_get_cell_info(self.h, _p(bn, C.c_uint32), _p(bl, C.c_float)))
        return bn, bl

    def bin(self):
        self._c(self.lib.pb_bin(self.h))

    def get_bins(self):
        bn, _ = self.cell_info()
        ncell = int(np.prod(bn.astype(np.int64)))
        pid = np.empty(self.n, dtype=np.uint32)
        bid = np.empty(self.n, dtype=np.uint32)
        bindex = np.empty(ncell + 1, dtype=np.uint32)
        self._c(self.lib.pb_get_bins(self.h, _p(pid, C.c_uint32), _p(bid, C.c_uint32),
                                     _p(bindex, C.c_uint32)))
        return pid, bid, bindex

    def neighbor_table(self, w=1):
        bn, _ = self.cell_info()
        ncell = int(np.prod(bn.astype(np.int64)))
        t = np.empty(ncell * (2 * w + 1) ** 3, dtype=np.uint32)
        self._c(self.lib.pb_get_neighbor_table(self.h, _p(t, C.c_uint32)))
        return t

    # --- LJ
    def lj_set_params(self, sigma_eps, ntypes):
        se = np.ascontiguousarray(sigma_eps, dtype=np.float64).reshape(-1)
        assert se.size == 2 * ntypes * ntypes
        self._c(self.lib.pb_lj_set_params(self.h, _p(se, C.c_double), ntypes))

    def lj_set_cutoff(self, rc2, shifted=True):
        self._c(self.lib.pb_lj_set_cutoff(self.h, rc2, int(shifted)))

    def lj_force(self, precision=PB_FP64, energies=True):
        pe = np.zeros(self.nrep)
        w = np.zeros(self.nrep)
        self._c(self.lib.pb_lj_force(self.h, precision, _p(pe, C.c_double) if energies else None,
                                     _p(w, C.c_double) if energies else None))
        return pe, w

    def coord_counts(self, rmax2t):
        out = np.empty(self.n, dtype=np.uint32)
        self._c(self.lib.pb_coord_counts(self.h, np.float32(rmax2t), _p(out, C.c_uint32)))
        return out

    def lj_run(self, dt, nsteps, precision=PB_MIXED, skin=0.3, rebuild_every=10,
               energy_every=1, use_graph=False, block_size=0, timing=False, check=True,
               persistent=False):
        o = pb_run_opts(dt, nsteps, precision, skin, rebuild_every, energy_every,
                        int(use_graph), block_size, int(timing), int(persistent))
        nrec = (nsteps + energy_every - 1) // energy_every if energy_every else 0
        e = np.zeros((max(nrec, 1), self.nrep, 3))
        rc = self.lib.pb_lj_run(self.h, C.byref(o), _p(e, C.c_double))
        if check:
            self._c(rc)
        return e[:nrec], rc

    def lj_autotune(self, dt, steps=10, precision=PB_MIXED, skin=0.3, rebuild_every=20,
                    persistent=False):
        o = pb_run_opts(dt, steps, precision, skin, rebuild_every, 0, 0, 0, 1, int(persistent))
        ms, w, f = C.c_double(), C.c_uint32(), C.c_double()
        self._c(self.lib.pb_lj_autotune(self.h, C.byref(o), steps, C.byref(ms), C.byref(w),
                                        C.byref(f)))
        return {"ms_per_step": ms.value, "tile_w": w.value, "tile_fill": f.value}

    def lj_set_tiles(self, tile_w, tile_fill):
        self._c(self.lib.pb_lj_set_tiles(self.h, int(tile_w), float(tile_fill)))

    def lj_run_stats(self):
        rb = C.c_uint32()
        mx = C.c_uint32()
        sv = C.c_int()
        self._c(self.lib.pb_lj_last_run_stats(self.h, C.byref(rb), C.byref(mx), C.byref(sv)))
        return {"rebuilds": rb.value, "max_nn": mx.value, "skin_violation": sv.value}

    def lj_run_timing(self):
        sm, rm = C.c_double(), C.c_double()
        ns, nr = C.c_uint32(), C.c_uint32()
        self._c(self.lib.pb_lj_last_run_timing(self.h, C.byref(sm), C.byref(ns), C.byref(rm),
                                               C.byref(nr)))
        return {"step_ms": sm.value, "nstep": ns.value, "rebuild_ms": rm.value,
                "nrebuild": nr.value}

    # --- SPH (cudaParticleSPH_NS)
    def sph_set_properties(self, mu0, c, h):
        mu0 = np.ascontiguousarray(np.broadcast_to(np.asarray(mu0, dtype=np.float64), (self.n,)))
        c = np.ascontiguousarray(np.broadcast_to(np.asarray(c, dtype=np.float64), (self.n,)))
        self._c(self.lib.pb_sph_set_properties(self.h, _p(mu0, C.c_double), _p(c, C.c_double),
                                               float(h)))

    def sph_set_wcsph(self, h, rho0, c0, gamma=7.0, alpha=0.1, pmin=0.0):
        """Config physics (cubic spline, Tait, artificial viscosity) for the
        following sph_density / sph_acceleration calls."""
        self._c(self.lib.pb_sph_set_wcsph(self.h, float(h), float(rho0), float(c0), float(gamma),
                                          float(alpha), float(pmin)))

    def sph_density(self):
        self._c(self.lib.pb_sph_density(self.h))

    def sph_acceleration(self):
        self._c(self.lib.pb_sph_acceleration(self.h))

    def sph_inspect_density(self):
        x = C.c_double()
        self._c(self.lib.pb_sph_inspect_density(self.h, C.byref(x)))
        return x.value

    # --- DEM (cudaParticleDEM / cudaParticleRotation)
    def dem_set_properties(self, E, mu, sigma, gamma, mu_r, r0):
        if np.ndim(r0) == 0:
            self._c(self.lib.pb_dem_set_properties(self.h, E, mu, sigma, gamma, mu_r, None,
                                                   float(r0)))
        else:
            r = np.ascontiguousarray(r0, dtype=np.float64)
            self._c(self.lib.pb_dem_set_properties(self.h, E, mu, sigma, gamma, mu_r,
                                                   _p(r, C.c_double), 0.0))

    def dem_set_linear(self, kn, e, kt, mu, r0):
        """Linear spring-dashpot law (normal stiffness kn, restitution e,
        tangential stiffness kt, Coulomb friction mu) in place of Hertz."""
        if np.ndim(r0) == 0:
            self._c(self.lib.pb_dem_set_linear(self.h, kn, e, kt, mu, None, float(r0)))
        else:
            r = np.ascontiguousarray(r0, dtype=np.float64)
            self._c(self.lib.pb_dem_set_linear(self.h, kn, e, kt, mu, _p(r, C.c_double), 0.0))

    def dem_history_width(self):
        z = C.c_uint32()
        self._c(self.lib.pb_dem_history_width(self.h, C.byref(z)))
        return z.value

    def rot_set_inertia(self, r0):
        if np.ndim(r0) == 0:
            self._c(self.lib.pb_rot_set_inertia(self.h, None, float(r0)))
        else:
            r = np.ascontiguousarray(r0, dtype=np.float64)
            self._c(self.lib.pb_rot_set_inertia(self.h, _p(r, C.c_double), 0.0))

    def rot_step(self, dt):
        self._c(self.lib.pb_rot_step(self.h, dt))

    def dem_calc_force(self, dt, autotune=False):
        z = C.c_uint32()
        self._c(self.lib.pb_dem_calc_force(self.h, dt, int(autotune), C.byref(z)))
        return z.value

    def dem_contacts(self):
        z = C.c_uint32()
        row = np.zeros(self.n + 1, dtype=np.uint32)
        self._c(self.lib.pb_dem_get_contacts(self.h, _p(row, C.c_uint32), None, None, 0,
                                             C.byref(z)))
        hw = self.dem_history_width()
        col = np.zeros(max(z.value, 1), dtype=np.uint32)
        val = np.zeros(max(z.value, 1) * hw, dtype=np.float64)
        self._c(self.lib.pb_dem_get_contacts(self.h, _p(row, C.c_uint32), _p(col, C.c_uint32),
                                             _p(val, C.c_double), col.size, C.byref(z)))
        val = val[:z.value * hw]
        return row, col[:z.value], (val if hw == 1 else val.reshape(-1, hw))

    def dem_set_contacts(self, row, col, val):
        row = np.ascontiguousarray(row, dtype=np.uint32)
        col = np.ascontiguousarray(col, dtype=np.uint32)
        val = np.ascontiguousarray(val, dtype=np.float64).reshape(-1)
        if val.size != col.size * self.dem_history_width():
            raise ValueError("dem_set_contacts: val must hold the history width per contact")
        self._c(self.lib.pb_dem_set_contacts(self.h, _p(row, C.c_uint32), _p(col, C.c_uint32),
                                             _p(val, C.c_double), int(col.size)))

    def dem_rewind(self):
        self._c(self.lib.pb_dem_rewind(self.h))

    def dem_violations(self):
        z = C.c_uint32()
        self._c(self.lib.pb_dem_last_violations(self.h, C.byref(z)))
        return z.value

    # --- g(r)
    def gr_hist(self, rstep, rnum):
        rs = np.ascontiguousarray(np.broadcast_to(np.asarray(rstep, dtype=np.float32),
                                                  (self.nrep,)))
        h = np.zeros((self.nrep, rnum), dtype=np.uint64)
        self._c(self.lib.pb_gr_hist(self.h, _p(rs, C.c_float), rnum, _p(h, C.c_uint64)))
        return h
