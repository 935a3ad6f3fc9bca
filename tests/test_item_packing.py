"""CPU test of the pair kernels' work-item packing (kd_item.h): the tile
count and the two trim cuts share one 16-bit field, read and written only
through Item's accessors.  Compiles tests/native/item_packing.cpp as host
code with hipcc and runs it (no GPU involved)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_item_packing_round_trip(tmp_path):
    exe = str(tmp_path / "item_packing")
    subprocess.check_call([HIPCC, "-std=c++17", "-O1", "--cuda-host-only", "-x", "hip",
                           os.path.join(HERE, "native", "item_packing.cpp"), "-o", exe])
    out = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip() == "ok %d" % (127 * 16 * 16)


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_raw_read_of_the_packed_field_does_not_compile(tmp_path):
    """kd_item.h keeps the packed member private, so a kernel cannot index
    tile arrays with the packed value by mistake: a raw read is a compile
    error, the accessor compiles."""
    inc = os.path.join(HERE, "..", "corruscant_amd", "csrc", "kd_item.h")
    results = {}
    for name, expr in (("raw", "it.run"), ("accessor", "it.tiles()")):
        src = tmp_path / (name + ".cpp")
        src.write_text('#include "%s"\nint f(const kd::Item &it) { return %s; }\n' % (inc, expr))
        r = subprocess.run([HIPCC, "-std=c++17", "--cuda-host-only", "-x", "hip", "-fsyntax-only", str(src)],
                           capture_output=True, text=True, timeout=120)
        results[name] = r
    assert results["accessor"].returncode == 0, results["accessor"].stderr
    assert results["raw"].returncode != 0 and "private" in results["raw"].stderr
