"""Failure handling that the host can force without touching a GPU: with no
visible device the library and the drop-in executable fail cleanly (a status
code and an exit code of 1, no output files), as the reference's process
fails on its own errors (src/main.zig:156-218; brain/evolve.py:439-442 raises
on a nonzero return code).  Every case runs in a child process whose HIP
runtime sees no device (HIP_VISIBLE_DEVICES=-1), so it holds on a GPU host
too."""
import os
import subprocess
import sys

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(REPO, "codec-arena_amd")
GOLDEN = os.path.join(REPO, "tests", "golden")


def _env():
    env = dict(os.environ)
    env["HIP_VISIBLE_DEVICES"] = "-1"
    env["PYTHONPATH"] = PKG + os.pathsep + env.get("PYTHONPATH", "")
    return env


def test_context_without_a_device_is_nodevice():
    code = ("import ctypes as C\n"
            "from csfunnel import _abi\n"
            "h = C.c_void_p()\n"
            "rc = _abi.lib().cs_ctx_create(0, C.byref(h))\n"
            "print(rc, bool(h.value), _abi.lib().cs_last_error().decode())\n")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=_env(), timeout=120)
    assert r.returncode == 0, r.stderr
    rc, handle, msg = r.stdout.split(" ", 2)
    assert int(rc) == -12 and handle == "False"  # CS_E_NODEV, no handle
    assert "device" in msg.lower()


def test_arena_screen_without_a_device_fails_cleanly(tmp_path):
    out, cnt = tmp_path / "promoted.txt", tmp_path / "screens.txt"
    argv = ["screen", "--model", os.path.join(GOLDEN, "c0_model.txt"), "--elites",
            os.path.join(GOLDEN, "c0_elites.txt"), "--out", str(out), "--count-out", str(cnt),
            "--lambda", "100000", "--keep", "16", "--seed", "1", "--threads", "4"]
    r = subprocess.run([sys.executable, os.path.join(PKG, "bin", "arena")] + argv, capture_output=True,
                       text=True, env=_env(), timeout=120)
    assert r.returncode == 1, (r.returncode, r.stderr)
    assert r.stderr.startswith("error: NoDevice"), r.stderr
    assert "Traceback" not in r.stderr
    assert not out.exists() and not cnt.exists()  # like the reference, nothing is written on failure
