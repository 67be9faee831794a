"""CPU checks of the lane maps of the matrix-core adjoint kernels (grid_mfma.hpp).

The accumulator layouts are the probed hardware ones:
v_mfma_f32_16x16x4f32 register i of lane l holds D[4 (l // 16) + i][l % 16],
v_mfma_f64_16x16x4f64 holds D[l // 16 + 4 i][l % 16], and
v_mfma_f32_32x32x2f32 holds D[i % 4 + 8 (i // 4) + 4 (l // 32)][l % 32].
Against those, the kernels' rules for which lane writes what must give:

* 3D (8^3 tiles, two 16-plane sets, the second starting at plane OFF1 = P - 16):
  every plane a in [0, P) of the slot receives set 0's row a (a < 16) plus set 1's
  row a - OFF1 (0 <= a - OFF1 < 16), exactly once per column, for FP32 and FP64 and
  every window width the kernel takes (W = 4..12);
* 2D (16^2 tiles, one 32-row product, two 32-column blocks): every (row, cell,
  re/im) of the P x P padded tile is written exactly once, at a distinct offset of
  the slot's destination-blocked order (the order the fold reads), W = 4..16.

A wrong lane offset in the set sum, a wrong register in the extra planes or a
missing block in the 2D store map fails these.
"""
import pytest


def acc_row_16x16(prec, g, i):
    """Row of register i in lane group g (lanes 16 g .. 16 g + 15)."""
    return 4 * g + i if prec == "single" else g + 4 * i


def set1_source(prec, g, i, off1):
    """(lane group, register) whose set-1 accumulator the kernel adds to the plane
    that register i of lane group g holds, or None for none (grid_mfma.hpp, put)."""
    if off1 == 0:
        return None
    if prec == "single":
        if i >= off1:
            return (g, i - off1)
        return (g - 1, 4 - off1 + i) if g > 0 else None
    # FP64: from group g - OFF1 register i, else group g - OFF1 + 4 register i - 1
    if g >= off1:
        return (g - off1, i)
    return (g - off1 + 4, i - 1) if i > 0 else None


def extra_planes(prec, off1):
    """(plane, lane group, register of set 1) of the planes 16 .. P - 1."""
    if prec == "single":
        return [(16 + i, 3, 4 - off1 + i) for i in range(off1)]
    return [(16 + g - (4 - off1), g, 3) for g in range(4 - off1, 4)]


@pytest.mark.parametrize("prec", ["single", "double"])
@pytest.mark.parametrize("W", range(4, 13))
def test_3d_set_sum_gives_every_plane_once(prec, W):
    P = 8 + W - 1
    off1 = P - 16 if P > 16 else 0
    got = {}
    for g in range(4):
        for i in range(4):
            a = acc_row_16x16(prec, g, i)
            if a >= P:
                continue
            terms = [("set0", a)]
            src = set1_source(prec, g, i, off1)
            if src is not None:
                gs, rs = src
                assert 0 <= gs < 4 and 0 <= rs < 4
                terms.append(("set1", acc_row_16x16(prec, gs, rs)))
            assert a not in got
            got[a] = sorted(terms)
    for a, gs, rs in extra_planes(prec, off1):
        assert a not in got
        got[a] = [("set1", acc_row_16x16(prec, gs, rs))]
    assert sorted(got) == list(range(P))
    for a in range(P):
        want = []
        if a < 16:
            want.append(("set0", a))
        if off1 and 0 <= a - off1 < 16:
            want.append(("set1", a - off1))
        assert got[a] == sorted(want), (a, got[a], want)


def ext(t, p, q):
    return p - q * t if p - q * t < t else t


def block_off_2d(P, q0, qL, T=16):
    # TileLayout::block_off with T1 = P1 = 1
    return q0 * T * P + ext(T, P, q0) * qL * T


@pytest.mark.parametrize("W", range(4, 17, 2))
def test_2d_store_map_covers_the_slot_once(W):
    P = 16 + W - 1
    assert P <= 32
    seen_cells, seen_offs = set(), set()
    for lane in range(64):
        nl, half = lane % 32, lane // 32
        for blk in range(2):
            c = 16 * blk + nl // 2
            ri = nl % 2
            if c >= P:
                continue
            EL = ext(16, P, blk)
            for i in range(16):
                a = i % 4 + 8 * (i // 4) + 4 * half
                if a >= P:
                    continue
                off = 2 * (block_off_2d(P, a // 16, blk) + (a % 16) * EL + (c - 16 * blk)) + ri
                assert (a, c, ri) not in seen_cells
                assert off not in seen_offs
                seen_cells.add((a, c, ri))
                seen_offs.add(off)
    assert len(seen_cells) == 2 * P * P
    assert seen_offs == set(range(2 * P * P))


def test_2d_operand_reads_stay_in_the_table():
    """Table rows of 100 floats: 32 row weights, then 64 columns; a step reads rows
    k + (lane // 32) for even k < nb <= 32, and lane % 32 of each part."""
    for nb in range(1, 33):
        for k in range(0, nb, 2):
            for lane in range(64):
                row = k + lane // 32
                assert row < 32
                for col in (lane % 32, 32 + lane % 32, 64 + lane % 32):
                    assert col < 100
