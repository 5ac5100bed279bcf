"""bench.py's contract on the CPU: one GPU only (any other --gpus exits with
a message before anything touches a GPU), and the algorithmic byte count of
the symmetric product agrees with the CSRC sizes derived from first
principles for hex8 100^3 (neq 3,060,300, nad 119,965,791: 1,476 MB in FP64
and 984 MB with FP32 values, plus x and y), and the node-blocked layout's
bytes are built from its node blocks."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_product_bytes_of_the_hex8_100_system():
    import bench

    neq, nad = 3060300, 119965791
    csrc64 = 4 * (neq + 1) + 4 * nad + 8 * nad + 8 * neq
    csrc32 = 4 * (neq + 1) + 4 * nad + 4 * nad + 4 * neq
    assert round(csrc64 / 1e6) == 1476 and round(csrc32 / 1e6) == 984
    assert bench.spmv_bytes(neq, nad, 8) == csrc64 + 16 * neq
    assert bench.spmv_bytes(neq, nad, 4, 4) == csrc32 + 8 * neq == 1008691132


def test_blocked_bytes_of_the_hex8_100_system():
    """The node-blocked count: 12,989,499 off-diagonal 3x3 blocks (nad less the
    3 in-node lower entries of each of the 1,020,100 nodes, over 9), each one
    4-byte word and 9 values; 6 diagonal-block values per node; x and y; the
    per-tile terms grow with the tiles and their external nodes."""
    import bench

    neq, nad, nn = 3060300, 119965791, 1020100
    nblk = (nad - 3 * nn) // 9
    assert nblk * 9 + 3 * nn == nad and nblk == 12989499
    base32 = bench.blocked_bytes(neq, nad, 3, 0, 0, 4, 4)
    assert base32 == nblk * 40 + nn * 24 + 8 * neq + 4 * nn
    base64 = bench.blocked_bytes(neq, nad, 3, 0, 0, 8, 8)
    assert base64 == nblk * 76 + nn * 48 + 16 * neq + 4 * nn
    # one external node of one tile: x staged and one partial written (3
    # entries each), its list entry and segptr entry, and the tile's own terms
    assert bench.blocked_bytes(neq, nad, 3, 1, 1, 4, 4) - base32 == 24 + 4 + 2 + 8 + 36 + 4
