"""Register and scratch budgets of the product kernels, read from the gfx950
assembly that hipcc writes for this library's own sources (same flags as the
build).  The products are bound by how many tiles are in flight: the FP32-sum
3-D single-system product must keep eight waves per SIMD (at most 64 VGPRs),
and no product kernel may keep values in scratch, since a spill of a loaded
value waits for every load in flight (the product then runs about 20% slower
with the same results, which no parity test would notice)."""
import os
import re
import shutil
import subprocess

import pytest

from elastic_amd import build as B


def _kernel_resources(src, tmp_path):
    if not shutil.which(B.HIPCC) and not os.path.exists(B.HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path / (os.path.basename(src) + ".s")
    cmd = [B.HIPCC] + B.COMMON + ["--cuda-device-only", "-S", "-x", "hip",
                                  os.path.join(B.CSRC, src), "-o", str(out)]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-2000:]
    res, cur = {}, None
    for line in open(out):
        m = re.match(r"^(_Z\S+):\s", line)
        if m:
            cur = m.group(1)
            res.setdefault(cur, {})
            continue
        if cur is None:
            continue
        for key in ("NumVgprs", "ScratchSize"):
            m = re.search(r"; %s: (\d+)" % key, line)
            if m and key not in res[cur]:
                res[cur][key] = int(m.group(1))
    return {k: v for k, v in res.items() if "NumVgprs" in v}


def test_single_system_products_fit_without_scratch(tmp_path):
    res = _kernel_resources("spmv_block.hip", tmp_path)
    prods = {k: v for k, v in res.items() if "k_spmv_blocks" in k}
    assert len(prods) >= 10
    for k, v in prods.items():
        assert v["ScratchSize"] == 0, (k, v)
    # k_spmv_blocks<float, float, 3, 4, true, 256, false>: the mixed solve's operator
    fp32 = [v for k, v in prods.items() if "k_spmv_blocksIffLi3ELi4ELb1ELi256ELb0E" in k]
    assert len(fp32) == 1
    assert fp32[0]["NumVgprs"] <= 64, fp32[0]


def test_batch_products_fit_without_scratch(tmp_path):
    res = _kernel_resources("batch.hip", tmp_path)
    prods = {k: v for k, v in res.items() if "k_bt_spmv" in k or "k_bt_dyn_spmv" in k}
    assert len(prods) >= 6
    for k, v in prods.items():
        assert v["ScratchSize"] == 0, (k, v)
