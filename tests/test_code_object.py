"""Register and LDS budgets of the built gfx950 kernels (CPU only).

The kernels' code objects are read out of libpsim_gpu.so's offload bundles
and their AMDGPU metadata printed with llvm-readelf.  The test holds the
properties the launch code and the measured throughput rely on: no kernel
spills to scratch memory, and the cell-per-workgroup pair kernels stay within
the VGPR and LDS budgets of the waves per SIMD they are compiled for (7, or 6
with LJ and AO both enabled or Calibration with LJ of a run-time exponent:
512 / waves VGPRs rounded down to a multiple of
8, and that many 4-wave workgroups in a CU's 160 KiB of LDS)."""
import os
import re
import shutil
import struct
import subprocess
import tempfile

import pytest

from proteinsim2015_amd import _lib

READELF = shutil.which("llvm-readelf", path="/opt/rocm/llvm/bin") or shutil.which("llvm-readelf")
BUNDLE_MAGIC = b"__CLANG_OFFLOAD_BUNDLE__"
KEYS = ("name", "private_segment_fixed_size", "vgpr_count", "vgpr_spill_count",
        "group_segment_fixed_size")


def gfx950_code_objects(path):
    """The gfx950 entries of every clang offload bundle in the file."""
    data = open(path, "rb").read()
    out = []
    pos = data.find(BUNDLE_MAGIC)
    while pos >= 0:
        (n,) = struct.unpack_from("<Q", data, pos + len(BUNDLE_MAGIC))
        q = pos + len(BUNDLE_MAGIC) + 8
        for _ in range(n):
            off, size, tlen = struct.unpack_from("<QQQ", data, q)
            q += 24
            triple = data[q:q + tlen].decode()
            q += tlen
            if triple.endswith("gfx950") and size:
                out.append(data[pos + off:pos + off + size])
        pos = data.find(BUNDLE_MAGIC, pos + len(BUNDLE_MAGIC))
    return out


def kernel_metadata(code_object):
    """{kernel symbol: {metadata key: value}} from the object's notes."""
    with tempfile.NamedTemporaryFile(suffix=".co") as f:
        f.write(code_object)
        f.flush()
        text = subprocess.run([READELF, "--notes", f.name], check=True, capture_output=True,
                              text=True).stdout
    kernels = {}
    for block in re.split(r"^  - (?=\.)", text, flags=re.M)[1:]:
        fields = {}
        for key in KEYS:
            m = re.search(r"^(?:    )?\." + key + r":\s+(\S+)", block, flags=re.M)
            if m:
                fields[key] = m.group(1)
        if "name" in fields and "vgpr_count" in fields:
            name = fields.pop("name")
            kernels[name] = {k: int(v) for k, v in fields.items()}
    return kernels


@pytest.fixture(scope="module")
def kernels():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("libpsim_gpu.so is not built")
    if READELF is None:
        pytest.skip("llvm-readelf not found")
    objs = gfx950_code_objects(_lib.LIB_PATH)
    assert objs, "no gfx950 code object in libpsim_gpu.so"
    found = {}
    for co in objs:
        found.update(kernel_metadata(co))
    return found


def test_every_kernel_is_found(kernels):
    names = " ".join(kernels)
    for k in ("k_step_dense", "k_forces_dense", "k_integrate", "k_move", "k_step",
              "k_ens_step", "k_adj_step", "k_step_brick"):
        assert k in names, k


def test_no_kernel_uses_scratch(kernels):
    spilled = {k: v for k, v in kernels.items()
               if v["private_segment_fixed_size"] != 0 or v.get("vgpr_spill_count", 0) != 0}
    assert not spilled


@pytest.mark.parametrize("kernel", ["k_step_dense", "k_forces_dense"])
def test_dense_kernels_fit_their_waves(kernels, kernel):
    seen = 0
    for name, v in kernels.items():
        m = re.search(kernel + r"ILi(\d+)ELi(\d+)E", name)
        if not m:
            continue
        seen += 1
        mask, ljn = int(m.group(1)), int(m.group(2))
        lj_ao = mask & 2 and mask & 4
        waves = 5 if (lj_ao and ljn == 0) else 6 if (lj_ao or (mask & 2 and ljn == 0)) else 7
        assert v["vgpr_count"] <= (512 // waves) // 8 * 8, (name, v)
        assert waves * v["group_segment_fixed_size"] <= 160 * 1024, (name, v)
    assert seen >= 16
