"""The node-blocked tile layout (elastic_amd/csrc/layout_block.cpp) on the
CPU, in both node orders: tests/layout_check.cpp builds the layout with the
library's own builder and applies it as the GPU product does, in FP64, and
the result must equal the symmetric CSRC product (D + L + L^T) x
(Matvec.f:528-621) to rounding; every CSRC block must be used exactly once,
from the right source and orientation, and d2r must be a permutation.  The
compact order must also cut the external column nodes of 3-D meshes."""
import os
import shutil
import subprocess

import numpy as np
import pytest

from elastic_amd import _lib
from elastic_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "..", "elastic_amd", "csrc")
GXX = shutil.which("g++")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if GXX is None:
        pytest.skip("g++ not installed")
    exe = str(tmp_path_factory.mktemp("layout") / "layout_check")
    subprocess.run([GXX, "-std=c++17", "-O2", "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include",
                    "-I" + CSRC, os.path.join(HERE, "layout_check.cpp"),
                    os.path.join(CSRC, "layout_block.cpp"), "-o", exe, "-lpthread"], check=True)
    return exe


def pattern(mesh):
    perm = _lib.reord(mesh)
    idn, neq = _lib.numeq(mesh, perm)
    return _lib.csrstruct(mesh, idn, neq)


def run(checker, tmp_path, ia, ja, compact, tile_blocks, seed=1):
    rng = np.random.default_rng(seed)
    neq, nad = ia.shape[0] - 1, ja.shape[0]
    f = str(tmp_path / ("sys_%d_%d.bin" % (compact, tile_blocks)))
    with open(f, "wb") as fh:
        np.array([neq, nad, compact, tile_blocks], np.int32).tofile(fh)
        ia.astype(np.int32).tofile(fh)
        ja.astype(np.int32).tofile(fh)
        (1.0 + rng.random(neq)).tofile(fh)
        rng.standard_normal(nad).tofile(fh)
        rng.standard_normal(neq).tofile(fh)
    r = subprocess.run([checker, f], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    nb, ntiles, next_nodes, comp, err = r.stdout.split()
    return int(nb), int(ntiles), int(next_nodes), int(comp), float(err)


MESHES = {
    "hex8_6": lambda: M.hex8_cube(6, jitter=0.1, seed=3),
    "hex8_14": lambda: M.hex8_cube(14, jitter=0.1, seed=3),
    "tet4_7": lambda: M.tet4_cube(7, jitter=0.2, seed=5),
    "quad4_20": lambda: M.quad4_square(20),
}


@pytest.mark.parametrize("name", sorted(MESHES))
@pytest.mark.parametrize("tile_blocks", [64, 1024])
def test_layout_product_equals_csrc_product(checker, tmp_path, name, tile_blocks):
    ia, ja = pattern(MESHES[name]())
    for compact in (0, 1):
        nb, ntiles, next_nodes, comp, err = run(checker, tmp_path, ia, ja, compact, tile_blocks)
        assert err < 1e-13, (compact, err)
        # (the builder takes either order for any node size; the library
        # asks for the compact one on 3-dof layouts)
        assert comp == compact


def test_compact_order_cuts_external_columns(checker, tmp_path):
    ia, ja = pattern(M.hex8_cube(24, jitter=0.1, seed=3))
    _, _, ext_csrc, _, _ = run(checker, tmp_path, ia, ja, 0, 1024)
    _, _, ext_compact, comp, err = run(checker, tmp_path, ia, ja, 1, 1024)
    assert comp == 1 and err < 1e-13
    assert ext_compact < 0.6 * ext_csrc
