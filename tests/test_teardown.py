"""Failure handling of the context layer that needs no GPU: the shim closes
every open context before interpreter shutdown, and the library refuses a
context whose mesh install failed part way."""
import atexit
import subprocess
import sys

import forestfv_public_amd as F


class _Fake:
    def __init__(self):
        self.closed = 0

    def close(self):
        self.closed += 1


def test_atexit_hook_is_registered_when_the_library_loads():
    F.lib()
    assert F._atexit_registered
    # atexit.unregister returns nothing, so check registration through a child
    # interpreter: the hook must run (and close a tracked context) at exit
    code = (
        "import forestfv_public_amd as F\n"
        "F.lib()\n"
        "class X:\n"
        "    def close(self):\n"
        "        print('closed-at-exit', flush=True)\n"
        "x = X()\n"
        "F._track(x)\n"
    )
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, check=True,
                         cwd=F._HERE + "/..").stdout
    assert "closed-at-exit" in out


def test_close_all_closes_tracked_contexts_and_disables_del(monkeypatch):
    monkeypatch.setattr(F, "_live", F.weakref.WeakSet())
    monkeypatch.setattr(F, "_shutdown", False)
    a, b = _Fake(), _Fake()
    F._track(a)
    F._track(b)
    F._close_all()
    assert a.closed == 1 and b.closed == 1
    assert F._shutdown


def test_broken_context_is_refused_and_failed_reserve_keeps_buffers():
    rc = F.lib().fv_dev_selftest_guard()
    assert rc == 0, f"self-check {rc} failed: {F.lib().fv_dev_last_error()}"
