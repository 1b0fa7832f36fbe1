"""pb_gr_hist_noalloc (g(r) histograms with no device allocation: the
replica table and histograms live in the library's static device array)
against pb_gr_hist through the C ABI: bit-identical histograms on two calls
in a row, the second on a smaller input (fewer replicas, fewer atoms, fewer
bins), so stale contents of the shared array would show; and the size
limits of the array are argument errors."""
import ctypes as C

import numpy as np
import pytest

import particlebase_amd as pb
from particlebase_amd import workloads

pytestmark = pytest.mark.gpu


def batch_system(ctx, nrep, nmin, nmax, seed):
    reps = workloads.lj_replicas(nrep, nmin, nmax, seed=seed)
    r, v, m, t, cells = reps.concat()
    s = pb.System(ctx, reps.counts)
    s.set_cell(cells)
    s.import_particles(r, m, v=v, type=t)
    rstep = np.array([np.float32(x.L / 400.0) for x in reps.systems], dtype=np.float32)
    return s, reps, rstep


def noalloc(s, rstep, rnum):
    h = np.zeros((s.nrep, rnum), dtype=np.uint64)
    rc = s.lib.pb_gr_hist_noalloc(s.h, rstep.ctypes.data_as(C.POINTER(C.c_float)), rnum,
                                  h.ctypes.data_as(C.POINTER(C.c_uint64)))
    return rc, h


def test_noalloc_matches_gr_hist_twice(ctx):
    print("\n[case] pb_gr_hist_noalloc vs pb_gr_hist: 12 replicas x 200 bins, then 3 x 150",
          flush=True)
    s1, reps1, rs1 = batch_system(ctx, 12, 500, 6000, seed=31)
    rc, h1 = noalloc(s1, rs1, 200)
    assert rc == pb.PB_OK, s1.lib.pb_ctx_last_error(ctx.h)
    ref1 = s1.gr_hist(rs1, 200)
    assert np.array_equal(h1, ref1)
    n1 = reps1.counts.astype(np.int64)
    assert np.array_equal(h1.sum(1).astype(np.int64), n1 * (n1 - 1) // 2)
    # a smaller second call: every slot it returns must be its own
    s2, reps2, rs2 = batch_system(ctx, 3, 500, 2000, seed=32)
    rc, h2 = noalloc(s2, rs2, 150)
    assert rc == pb.PB_OK
    ref2 = s2.gr_hist(rs2, 150)
    assert np.array_equal(h2, ref2)
    n2 = reps2.counts.astype(np.int64)
    assert np.array_equal(h2.sum(1).astype(np.int64), n2 * (n2 - 1) // 2)
    s1.close()
    s2.close()


def test_noalloc_limits(ctx):
    print("\n[case] pb_gr_hist_noalloc size limits", flush=True)
    s, reps, rs = batch_system(ctx, 9, 500, 800, seed=33)
    rc, _ = noalloc(s, rs, 16384)            # 9 x 16384 slots > 131072
    assert rc == pb.PB_ERR_ARG
    assert "131072" in s.lib.pb_ctx_last_error(ctx.h).decode()
    rc, _ = noalloc(s, rs, 16385)            # above the LDS histogram's bins
    assert rc == pb.PB_ERR_ARG
    rc, h = noalloc(s, rs, 14563)            # 9 x 14563 = 131067 slots: allowed
    assert rc == pb.PB_OK
    assert np.array_equal(h, s.gr_hist(rs, 14563))
    s.close()
