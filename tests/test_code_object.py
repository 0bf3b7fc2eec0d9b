"""The built gfx950 code objects keep the properties the kernel design depends on: the hot
kernels use no scratch memory (no spills, no dynamically indexed private arrays), the
float exact cost+gradient kernel fits four waves per SIMD (<= 128 VGPRs) and the two-wave
table relaxation kernel two (<= 256).  Read from the kernel metadata of the objects that
packing_amd/_build.py leaves in packing_amd/_obj (CPU only; skipped without a build)."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OBJ = os.path.join(ROOT, "packing_amd", "_obj")
LLVM = "/opt/rocm/llvm/bin"
TARGET = "hipv4-amdgcn-amd-amdhsa--gfx950"


def kernels(obj: str) -> dict:
    """name -> {vgpr, scratch, spill} from the device code object bundled in obj."""
    with tempfile.TemporaryDirectory() as d:
        fb, co = os.path.join(d, "fb"), os.path.join(d, "co")
        subprocess.run([os.path.join(LLVM, "llvm-objcopy"), f"--dump-section=.hip_fatbin={fb}", obj, os.devnull],
                       check=True, capture_output=True)
        subprocess.run([os.path.join(LLVM, "clang-offload-bundler"), "--unbundle", f"--input={fb}", "--type=o",
                        f"--targets={TARGET}", f"--output={co}"], check=True, capture_output=True)
        notes = subprocess.run([os.path.join(LLVM, "llvm-readelf"), "--notes", co], check=True,
                               capture_output=True, text=True).stdout
    out = {}
    for block in notes.split("- .agpr_count")[1:]:
        def field(f):
            return re.search(r"\." + f + r":\s+(\S+)", block).group(1)
        out[field("name")] = {"vgpr": int(field("vgpr_count")),
                              "scratch": int(field("private_segment_fixed_size")),
                              "spill": int(field("vgpr_spill_count"))}
    return out


def _objects():
    if not os.path.isdir(OBJ):
        pytest.skip("no build objects (run __graft_entry__.build())")
    for tool in ("llvm-objcopy", "clang-offload-bundler", "llvm-readelf"):
        if shutil.which(os.path.join(LLVM, tool)) is None:
            pytest.skip(f"{tool} not available")
    return OBJ


@pytest.mark.parametrize("src", ["pk_overlap", "pk_relax", "pk_lut", "pk_feasible", "pk_lap"])
def test_hot_kernels_use_no_scratch(src):
    obj = os.path.join(_objects(), src + ".o")
    if not os.path.exists(obj):
        pytest.skip(f"{src}.o not built")
    ks = kernels(obj)
    assert ks, src
    bad = {k: v for k, v in ks.items() if v["scratch"] or v["spill"]}
    assert not bad, bad


def test_occupancy_register_budgets():
    ov = kernels(os.path.join(_objects(), "pk_overlap.o"))
    # k_overlap<float, MODE_AREA, GRAD, !CRYSTAL, 256>: four 64-lane waves per SIMD
    k = "_ZN2pk9k_overlapIfLi0ELb1ELb0ELi256EEEvPKT_PKiiiPKNS_5GeomTIS1_EES3_iiPS1_SA_Py"
    assert ov[k]["vgpr"] <= 128, ov[k]
    rx = kernels(os.path.join(_objects(), "pk_relax.o"))
    # k_relax_w<MODE_LUT, 10, 8, 2> (mangled template arguments 2, 10, 8, 2): two waves per SIMD
    k = [n for n in rx if n.startswith("_ZN2pk9k_relax_wILi2ELi10ELi8ELi2EEEv")]
    assert len(k) == 1 and rx[k[0]]["vgpr"] <= 256, k
