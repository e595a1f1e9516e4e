"""Resource usage of the built transport kernels, read from the gfx950 code
object inside liboptphot_hip.so (no GPU needed).

Two regressions are silent at run time and cost whole percents: a kernel that
starts using scratch memory (the compiler turned a select into a stack load),
and a kernel that crosses a VGPR granule and loses a wave per SIMD. Both show
in the code object's metadata, so the kernels on the measured paths are
checked here: no private segment, and the VGPR counts that keep their
occupancy (six waves per SIMD for the ring kernel of the long-histories tally,
which its LDS limits to six anyway, and for the batch kernel).
"""
import os
import re
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "gpu_optical_photons_amd", "liboptphot_hip.so")
LLVM = "/opt/rocm/llvm/bin"


def _kernels(tmp_path):
    if not os.path.exists(LIB):
        pytest.skip("liboptphot_hip.so is not built")
    tools = [os.path.join(LLVM, t) for t in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf")]
    if not all(shutil.which(t) for t in tools):
        pytest.skip("ROCm LLVM tools not found")
    fat, co = tmp_path / "fat.bin", tmp_path / "co.o"
    subprocess.run([tools[0], "--dump-section", f".hip_fatbin={fat}", LIB, str(tmp_path / "x.o")],
                   check=True, capture_output=True)
    subprocess.run([tools[1], "--unbundle", "--type=o", f"--input={fat}",
                    "--targets=hipv4-amdgcn-amd-amdhsa--gfx950", f"--output={co}"],
                   check=True, capture_output=True)
    notes = subprocess.run([tools[2], "--notes", str(co)], check=True, capture_output=True,
                           text=True).stdout
    out, name = {}, None
    for line in notes.splitlines():
        m = re.match(r"\s+\.name:\s+(\S+)", line)
        if m:
            name = m.group(1)
            out[name] = {}
            continue
        m = re.match(r"\s+\.(private_segment_fixed_size|vgpr_count|sgpr_spill_count|vgpr_spill_count):\s+(\d+)", line)
        if m and name:
            out[name][m.group(1)] = int(m.group(2))
    return out


def _match(kernels, pattern):
    found = {k: v for k, v in kernels.items() if re.search(pattern, k)}
    assert found, pattern
    return found


def test_measured_kernels_use_no_scratch(tmp_path):
    ks = _kernels(tmp_path)
    hot = [r"wavefront_ring_kernelILb0E", r"wavefront_kernelILb[01]ELb[01]ELb0E",
           r"history_kernelILb[01]ELb[01]ELb[01]ELb0E", r"batch_kernelILb0E",
           r"batch_ring_kernelILb0E", r"genstep_ring_kernelILb0E"]
    for pat in hot:
        for name, meta in _match(ks, pat).items():
            assert meta.get("private_segment_fixed_size", 0) == 0, name
            assert meta.get("vgpr_spill_count", 0) == 0, name


def test_occupancy_of_the_measured_kernels(tmp_path):
    ks = _kernels(tmp_path)
    # 512 VGPRs per SIMD lane, allocated in granules of 8: <= 80 keeps six waves
    for name, meta in _match(ks, r"wavefront_ring_kernelILb0ELi3ELb1E").items():
        assert meta["vgpr_count"] <= 80, (name, meta)
    for name, meta in _match(ks, r"batch_kernelILb0ELb0E").items():
        assert meta["vgpr_count"] <= 80, (name, meta)
