"""Pins of the oracle's cell grid, neighbour table, cell ids and sort
(restating cudaCutoffBlock.cu:17-141 and kernelfuncs.cu:807-932).

The reference ships no fixtures; these checks fix the functions by closed
forms, symmetry, library routines and edge cases."""
import numpy as np
import pytest

NONE = 0xFFFFFFFF


@pytest.mark.parametrize("L,rmax", [(16.795962, 2.5), (107.494156, 2.5), (800.0, 80.0),
                                    (10.0, 3.3)])
def test_setup_cells_closed_form(orc, L, rmax):
    rc, bn, bl = orc.setup_cells([0, L, -1, L - 1, 2, L + 2], rmax, periodic=True)
    for a in range(3):
        nb = int(bn[a])
        # largest integer with nb * rmax <= L (FP32 rounding slack)
        assert nb * rmax <= L * (1 + 1e-6) and (nb + 1) * rmax > L * (1 - 1e-6)
        assert abs(bl[a] * nb - L) <= 1e-5 * L
        assert bl[a] >= rmax * (1 - 1e-6)
    assert list(bl[3:]) == [0.0, -1.0, 2.0]
    assert rc == (0 if bn.min() > 2 else -1)


def test_setup_cells_nonperiodic_and_reject(orc):
    rc, bn, bl = orc.setup_cells([0, 10, 0, 10, 0, 10], 3.0, periodic=False)
    assert rc == 0 and list(bn) == [4, 4, 4] and np.allclose(bl[:3], 3.0)
    assert bn[0] * bl[0] >= 10.0           # the extra block covers the far face
    rc, bn, _ = orc.setup_cells([0, 5, 0, 10, 0, 10], 2.0, periodic=True)
    assert rc == -1 and bn[0] == 2         # the reference asserts nb > 2


def _table_closed_form(bn, w, periodic):
    nx, ny, nz = (int(x) for x in bn)
    out = []
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                for a in range(-w, w + 1):
                    for b in range(-w, w + 1):
                        for c in range(-w, w + 1):
                            ii, jj, kk = i + a, j + b, k + c
                            inside = 0 <= ii < nx and 0 <= jj < ny and 0 <= kk < nz
                            if periodic or inside:
                                out.append((kk % nz) + (jj % ny) * nz + (ii % nx) * ny * nz)
                            else:
                                out.append(NONE)
    return np.array(out, dtype=np.uint32)


@pytest.mark.parametrize("bn,w,periodic", [((3, 4, 5), 1, True), ((4, 3, 3), 1, False),
                                           ((5, 6, 5), 2, True), ((5, 5, 7), 2, False)])
def test_neighbor_table(orc, bn, w, periodic):
    t = orc.neighbor_table(np.array(bn), w, periodic)
    assert np.array_equal(t, _table_closed_form(bn, w, periodic))
    J = (2 * w + 1) ** 3
    tt = t.reshape(-1, J)
    ncell = tt.shape[0]
    # symmetry: J in N(I) <=> I in N(J)
    for I in range(0, ncell, max(1, ncell // 17)):
        for Jc in tt[I]:
            if Jc != NONE:
                assert I in tt[Jc]
    assert np.all(tt[:, J // 2] == np.arange(ncell))    # centre entry is the cell itself


def test_calc_bid_edges(orc):
    cell = np.array([0.0, 16.0, -4.0, 12.0, 1.0, 17.0], dtype=np.float32)
    rc, bn, bl = orc.setup_cells(cell, 2.5, True)
    assert rc == 0 and list(bn) == [6, 6, 6]
    xmax = np.float32(16.0)
    pts = np.array([
        [0.0, -4.0, 1.0],                                          # box minimum -> cell 0
        [np.nextafter(xmax, np.float32(-1)), 11.999999, 16.999998],  # just inside -> nb-1
        [16.0, 12.0, 17.0],                                        # on the max face -> nb-1
        [-0.001, -4.5, 0.0],                                       # below the box -> 0
        [-1e30, -1e30, -1e30],                                     # far below -> 0
        [16.0 + 2.7, 12.0 + 2.7, 17.0 + 2.7],                      # d > nb -> 0
        [np.nan, 5.0, 5.0],                                        # NaN saturates to 0
    ], dtype=np.float32)
    bid = orc.calc_bid(pts, bl, bn)
    nyz = 36
    assert bid[0] == 0
    assert bid[1] == 5 + 5 * 6 + 5 * nyz
    assert bid[2] == 5 + 5 * 6 + 5 * nyz
    assert bid[3] == 0 and bid[4] == 0 and bid[5] == 0
    assert bid[6] == (int((5.0 - 1.0) / bl[2])) + int((5.0 + 4.0) / bl[1]) * 6


def test_calc_bid_containment(orc, rng):
    cell = np.array([0.0, 16.795962, 0.0, 16.795962, 0.0, 16.795962], dtype=np.float32)
    rc, bn, bl = orc.setup_cells(cell, 2.5, True)
    x = rng.uniform(0, 16.795962, size=(20000, 3)).astype(np.float32)
    bid = orc.calc_bid(x, bl, bn)
    d = np.stack([bid // (bn[1] * bn[2]), (bid // bn[2]) % bn[1], bid % bn[2]], axis=1)
    lo = d * bl[:3].astype(np.float64)
    tol = 1e-5
    assert np.all(x >= lo - tol) and np.all(x <= lo + bl[:3] + tol)


def test_sort_bins_matches_stable_argsort(orc, rng):
    ncell = 97
    bid = rng.integers(0, ncell, size=5000).astype(np.uint32)
    bid[bid == 0] = 1                       # cell 0 empty: the reference's makeBindex quirk
    pid, bs, bindex = orc.sort_bins(bid, ncell)
    ref = np.argsort(bid, kind="stable")
    assert np.array_equal(pid, ref) and np.array_equal(bs, bid[ref])
    cnt = np.bincount(bid, minlength=ncell)
    starts = np.concatenate([[0], np.cumsum(cnt)[:-1]])
    exp = np.where(cnt > 0, starts, NONE).astype(np.uint32)
    assert np.array_equal(bindex[:ncell], exp) and bindex[ncell] == 5000
    assert bindex[0] == NONE and cnt.sum() == 5000
