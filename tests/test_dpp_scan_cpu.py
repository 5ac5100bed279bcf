"""The wavefront scans of the blocked product, checked on the CPU: the DPP step
lists of seg_scan (segmented row sums) and tile_dot_store (the tile's share
of x.(Ax)) are read from elastic_amd/csrc/spmv_block.h and run on an
emulation of the gfx950 DPP lane moves (update_dpp with bound_ctrl off:
lanes without a source, or in rows the row mask leaves out, keep the `old`
operand).  Each list must give, in every lane, the exact sum of its run up to
that lane (keys ascending along the 64 lanes, as a tile's blocks are sorted by
row node), and lane 63 must hold the wavefront's total, on integer-valued
data so that no rounding is involved."""
import os
import re

import numpy as np

SRC = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))),
                   "elastic_amd", "csrc", "spmv_block.h")


def _steps(pattern):
    text = open(SRC).read()
    steps = [(int(c, 16), int(m, 16)) for c, m in re.findall(pattern, text)]
    assert len(steps) == 6
    return steps


def dpp_move(v, old, ctrl, row_mask):
    """update_dpp(old, v, ctrl, row_mask, bank_mask 0xF, bound_ctrl false)."""
    out = np.array(old, dtype=v.dtype, copy=True) if np.ndim(old) else np.full(64, old, v.dtype)
    for lane in range(64):
        row, pos = divmod(lane, 16)
        if not (row_mask >> row) & 1:
            continue
        if 0x111 <= ctrl <= 0x11F:  # row_shr:n
            n = ctrl - 0x110
            src = lane - n if pos >= n else None
        elif ctrl == 0x142:  # row_bcast:15 (lane 15 of the previous row)
            src = 16 * row - 1 if row > 0 else None
        elif ctrl == 0x143:  # row_bcast:31
            src = 31 if row >= 2 else None
        else:
            raise AssertionError("unexpected DPP control %#x" % ctrl)
        if src is not None:
            out[lane] = v[src]
    return out


def seg_scan(keys, y, steps):
    y = y.copy()
    for ctrl, mask in steps:
        ok = dpp_move(keys, -1, ctrl, mask)
        o = dpp_move(y, 0, ctrl, mask)
        y = np.where(ok == keys, y + o, y)
    return y


def test_seg_scan_sums_each_run_exactly():
    steps = _steps(r"seg_step<(0x[0-9A-Fa-f]+), (0x[0-9A-Fa-f]+)>\(key, y\)")
    rng = np.random.default_rng(7)
    for trial in range(400):
        # runs of 1..64 lanes (a node row holds at most 64 blocks), ascending
        # keys; the tail past the tile's blocks has the key 255
        lens = []
        while sum(lens) < 64:
            lens.append(int(rng.integers(1, 65 if trial % 4 == 0 else 17)))
        keys = np.repeat(np.arange(len(lens)), lens)[:64].astype(np.int64)
        if trial % 3 == 0:
            keys[int(rng.integers(1, 64)):] = 255
        y = rng.integers(-1000, 1000, 64).astype(np.float64)
        got = seg_scan(keys, y, steps)
        want = np.empty(64)
        for lane in range(64):
            start = lane
            while start > 0 and keys[start - 1] == keys[lane]:
                start -= 1
            want[lane] = y[start:lane + 1].sum()
        assert np.array_equal(got, want), (trial, keys, got - want)


def test_wavefront_sum_lands_in_lane_63():
    steps = _steps(r"v \+= dpp_move<(0x[0-9A-Fa-f]+), (0x[0-9A-Fa-f]+)>\(v\)")
    rng = np.random.default_rng(11)
    for _ in range(200):
        v0 = rng.integers(-1000, 1000, 64).astype(np.float64)
        v = v0.copy()
        for ctrl, mask in steps:
            v = v + dpp_move(v, 0.0, ctrl, mask)
        assert v[63] == v0.sum()
