"""The C ABI library loads and exports every entry point declared in
include/bh_hip.h; without a device, bh_create fails with a message.  No
compute call is made here."""
import ctypes
import os
import re

import pytest

from conftest import gpu_available
from barneshut.kernels import hiplib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "bh_hip.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:const\s+char\s*\*|int)\s+(bh_\w+)\s*\(", text, re.M)))


def test_header_declares_expected_entry_points():
    names = declared()
    assert "bh_step" in names and "bh_evaluate_boxes" in names and "bh_mlp_backward" in names
    assert sorted(hiplib.EXPORTED) == names


def test_library_exports_every_declared_symbol():
    L = hiplib.load()
    for name in declared():
        assert hasattr(L, name), name
        assert ctypes.cast(getattr(L, name), ctypes.c_void_p).value


def test_library_is_built_for_gfx950():
    path = hiplib.LIB_PATH
    assert os.path.exists(path)
    blob = open(path, "rb").read()
    assert b"gfx950" in blob


def test_version():
    assert hiplib.load().bh_version() == 1


@pytest.mark.skipif(gpu_available(), reason="checks the no-device error path")
def test_create_without_device_fails_cleanly():
    L = hiplib.load()
    h = ctypes.c_void_p()
    rc = L.bh_create(0, ctypes.byref(h))
    assert rc != 0
    assert L.bh_last_error()
    with pytest.raises(hiplib.BhError):
        hiplib.check(rc)


def test_null_arguments_rejected_without_device():
    L = hiplib.load()
    assert L.bh_create(0, None) == hiplib.BH_EINVAL
    assert L.bh_set_systems(None, 1, None, None) == hiplib.BH_EINVAL
    assert L.bh_destroy(None) == 0
    assert L.bh_mlp_forward(None, None, 0, None, None) == hiplib.BH_EINVAL
    q = ctypes.c_int64()
    assert L.bh_query(None, 0, ctypes.byref(q)) == hiplib.BH_EINVAL


def test_load_binds_the_library_to_torchs_hip_runtime():
    """hiplib.load() in a fresh process, torch imported only afterwards by the
    caller: the process must hold one HIP runtime (torch's), not /opt/rocm's
    beside it, or the library's calls cannot see the device."""
    import subprocess
    import sys
    code = ("import sys; sys.path[:0] = sys.argv[1:]\n"
            "from barneshut.kernels import hiplib\n"
            "hiplib.load()\n"
            "import torch\n"
            "libs = sorted({l.split()[-1] for l in open('/proc/self/maps') if 'libamdhip64' in l})\n"
            "print('\\n'.join(libs))\n")
    pkg = os.path.join(ROOT, "barnes-hut_amd")
    out = subprocess.run([sys.executable, "-c", code, pkg], capture_output=True, text=True,
                         timeout=300, check=True).stdout.split()
    assert len(out) == 1, out
