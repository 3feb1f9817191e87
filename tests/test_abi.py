"""The C-ABI library loads on a host without a GPU and exports every function
declared in include/*.h; device entry points fail cleanly without a device."""
import ctypes as C
import glob
import os
import re
import subprocess

import pytest

import forestfv_public_amd as F

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        src = open(h).read()
        src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
        for mt in re.finditer(r"^[A-Za-z_][\w \*]*?\b(fv_\w+)\s*\(", src, flags=re.M):
            names.add(mt.group(1))
    return sorted(names)


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert len(names) > 25
    out = subprocess.run(["nm", "-D", "--defined-only", F.LIB_PATH], capture_output=True, text=True, check=True).stdout
    exported = {ln.split()[-1] for ln in out.splitlines() if ln.strip()}
    missing = [n for n in names if n not in exported]
    assert not missing, missing
    L = F.lib()
    for n in names:
        assert getattr(L, n) is not None


def test_attach_without_device_is_a_status():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None and os.path.exists("/dev/kfd"):
        pytest.skip("a GPU may be present")
    p, m = F.make_params(), F.Mesh.uniform((1, 1, 1), (0, 0, 0), 1.0, 2, [4] * 6)
    with pytest.raises(F.FVError) as e:
        F.Device(p, m)
    assert e.value.code in (F.FV_ERR_HIP,)


@pytest.mark.parametrize("bad", [dict(prelax=4), dict(primtv=2), dict(nmat=4, gamma=(1.4,) * 4, pinf=(0.0,) * 4)])
def test_bad_configuration_rejected_before_the_device(bad):
    p = F.make_params(**dict(dict(nmat=2), **bad))
    m = F.Mesh.uniform((1, 1, 1), (0, 0, 0), 1.0, 2, [4] * 6)
    with pytest.raises(F.FVError) as e:
        F.Device(p, m)
    assert e.value.code == F.FV_ERR_ARG
