"""CPU-side checks of the drop-in boundary: the C ABI library loads and exports
every function include/nicp_amd.h declares, the Python bindings mirror the
header's structs, and the product path has no CPU fallback."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from nicp_amd import capi, synth

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    text = open(os.path.join(ROOT, "include", "nicp_amd.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(nicp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 25
    out = subprocess.check_output(["nm", "-D", "--defined-only", capi.LIB_PATH]).decode()
    exported = set(re.findall(r" T (nicp_[a-z0-9_]+)", out))
    assert set(names) <= exported, set(names) - exported
    assert set(capi.EXPORTS) == set(names)
    L = capi.lib()
    for n in names:
        assert hasattr(L, n)


def test_struct_layouts_match_header():
    # sizes computed by the C compiler from the header
    src = r'''
#include <stdio.h>
#include <stddef.h>
#include "nicp_amd.h"
int main(void) {
  printf("%zu %zu %zu %zu %zu %zu %zu %zu %zu\n", sizeof(nicp_camera), sizeof(nicp_stats_params),
         sizeof(nicp_align_params), sizeof(nicp_align_result), offsetof(nicp_align_result, total_time_ms),
         offsetof(nicp_align_result, overlap_pixels), sizeof(nicp_gaussian), offsetof(nicp_gaussian, information_vector),
         sizeof(nicp_merger_params));
  return 0;
}
'''
    import tempfile
    with tempfile.TemporaryDirectory() as d:
        with open(os.path.join(d, "s.c"), "w") as f:
            f.write(src)
        subprocess.check_call(["gcc", "-I", os.path.join(ROOT, "include"), "-o", os.path.join(d, "s"),
                               os.path.join(d, "s.c")])
        sizes = list(map(int, subprocess.check_output([os.path.join(d, "s")]).split()))
    assert sizes == [C.sizeof(capi.Camera), C.sizeof(capi.StatsParams), C.sizeof(capi.AlignParams),
                     C.sizeof(capi.AlignResult), capi.AlignResult.total_time_ms.offset,
                     capi.AlignResult.overlap_pixels.offset, 28 * 4, 22 * 4, C.sizeof(capi.MergerParams)]


def test_no_cpu_fallback_without_device():
    # on a machine without a GPU every entry point that needs the device fails loudly
    if os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK):
        pytest.skip("a GPU is present")
    with pytest.raises(capi.NicpError):
        capi.Context(0)


def test_product_does_not_import_the_oracle():
    pkg = os.path.join(ROOT, "nicp_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".cpp", ".hip", ".h", ".sh")):
                text = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in text and "nicp_oracle" not in text, f
    assert "import oracle" not in open(os.path.join(ROOT, "include", "nicp_amd.h")).read()


def test_synthetic_pairs_are_seeded():
    a = synth.pair(11, 60, 80)
    b = synth.pair(11, 60, 80)
    c = synth.pair(12, 60, 80)
    assert all(np.array_equal(x, y) for x, y in zip(a, b))
    assert not np.array_equal(a[0], c[0])
    d, _, T = synth.pair(3, 240, 320)
    holes = (d == 0).mean()
    assert 0.08 < holes < 0.2
    R = T.reshape(4, 4, order="F")[:3, :3].astype(np.float64)
    assert np.allclose(R @ R.T, np.eye(3), atol=1e-6)
    ang, tr = synth.pose_error(T, np.eye(4, dtype=np.float32).flatten())
    assert ang <= 5.0 + 1e-4 and tr <= 0.05 + 1e-6
