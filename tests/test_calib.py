"""The VALU issue-rate calibration library (libmt_calib.so, the measured peak of
the closest-hit kernels' instruction roofline): it loads without a GPU, rejects
bad arguments before any HIP call, and each calibration kernel's loop issues
exactly the instruction count that bench.py divides by (read from the gfx950
assembly that the same compiler flags produce)."""
import ctypes
import os
import re
import shutil
import subprocess

import pytest

from mambatracer_amd import build as mbuild

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench_mix():
    import importlib.util
    spec = importlib.util.spec_from_file_location("bench_mod", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.VALU_CALIB_MIX


def test_calib_library_loads_and_checks_arguments():
    assert os.path.exists(mbuild.CALIB_LIB), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    lib = ctypes.CDLL(mbuild.CALIB_LIB)
    f = lib.mt_calib_valu
    f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
    f.restype = ctypes.c_int
    buf = (ctypes.c_float * 256)()
    # every one of these returns before touching the GPU
    assert f(-1, 1, 1, ctypes.cast(buf, ctypes.c_void_p), None) == -1
    assert f(8, 1, 1, ctypes.cast(buf, ctypes.c_void_p), None) == -1
    assert f(0, 0, 1, ctypes.cast(buf, ctypes.c_void_p), None) == -1
    assert f(0, 1, 0, ctypes.cast(buf, ctypes.c_void_p), None) == -1
    assert f(0, 1, 1, None, None) == -1


@pytest.mark.skipif(shutil.which("hipcc") is None and not os.path.exists("/opt/rocm/bin/hipcc"),
                    reason="needs hipcc")
def test_loop_instruction_counts_match_bench(tmp_path):
    out = tmp_path / "calib.s"
    cmd = [mbuild._hipcc(), *mbuild.HIPCC_FLAGS, "--cuda-device-only", "-S",
           os.path.join(mbuild.CSRC, "mt_calib.hip"), "-o", str(out)]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr
    asm = out.read_text()
    mix = _bench_mix()
    assert sorted(mix) == list(range(8))
    for kind, (_, per_iter) in mix.items():
        m = re.search(rf"^_ZN12_GLOBAL__N_112k_valu_issueILi{kind}EEEvPfif:(.*?)s_endpgm", asm, re.S | re.M)
        assert m, kind
        loop = re.search(r"Inner Loop Header.*?\n(.*?)s_cbranch_scc0", m.group(1), re.S)
        assert loop, kind
        valu = [ln for ln in loop.group(1).splitlines() if ln.strip().startswith("v_")]
        assert len(valu) == per_iter, (kind, len(valu), per_iter)
