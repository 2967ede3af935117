"""Host code under AddressSanitizer, UndefinedBehaviorSanitizer and
ThreadSanitizer: the CPU oracle, the codec label source and the ABI's host
code are compiled with gcc/g++ (host only, no GPU code) and exercised by the
drivers in tests/sanitize/; any sanitizer report fails the test."""
import os
import shutil
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer",
       "-ffp-contract=off", "-fno-fast-math", "-march=x86-64-v3", "-std=gnu11"]


TSAN = ["-fsanitize=thread" if x == "-fsanitize=address,undefined" else x for x in SAN
        if x != "-fno-sanitize-recover=undefined"]


def _build(tmp_path, name, sources, include, cxx=False, flags=SAN):
    cc = "g++" if cxx else "gcc"
    if shutil.which(cc) is None:
        pytest.skip(cc + " not available")
    exe = str(tmp_path / name)
    flags = [x if x != "-std=gnu11" or not cxx else "-std=c++17" for x in flags]
    inc = [x for d in ([include] if isinstance(include, str) else include) for x in ("-I", d)]
    r = subprocess.run([cc, *flags, *inc, "-o", exe, *sources, "-lpthread", "-lm"], capture_output=True,
                       text=True, timeout=300)
    if r.returncode != 0:
        if any(k in r.stderr.lower() for k in ("asan", "ubsan", "tsan")):
            pytest.skip("sanitizer runtime not available: " + r.stderr[-200:])
        raise AssertionError(r.stderr[-2000:])
    return exe


def _run(cmd):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert "runtime error" not in out and "AddressSanitizer" not in out and "ThreadSanitizer" not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    return r.stdout


def test_oracle_under_sanitizers(tmp_path, golden):
    exe = _build(tmp_path, "oracle_drv", [os.path.join(REPO, "tests", "sanitize", "oracle_drv.c"),
                                         os.path.join(REPO, "oracle", "screen_oracle.c")],
                 os.path.join(REPO, "oracle"))
    out = _run([exe, os.path.join(golden, "c0_model.txt"), os.path.join(golden, "c0_elites.txt")])
    assert out.strip().endswith("ok")


def test_codec_label_source_under_sanitizers(tmp_path):
    exe = _build(tmp_path, "codec_drv", [os.path.join(REPO, "tests", "sanitize", "codec_drv.c"),
                                        os.path.join(REPO, "codec-arena_amd", "csrc", "codec_eval.c")],
                 os.path.join(REPO, "include"))
    out = _run([exe, "1200"])
    ok = int(out.split()[1])
    assert ok > 300  # most clips encode and decode


def test_host_abi_code_under_sanitizers(tmp_path, golden):
    # funnel_host.cpp is host C++ only (no HIP): it builds on its own here
    exe = _build(tmp_path, "host_abi_drv", [os.path.join(REPO, "tests", "sanitize", "host_abi_drv.cpp"),
                                           os.path.join(REPO, "codec-arena_amd", "csrc", "funnel_host.cpp")],
                 [os.path.join(REPO, "include"), os.path.join(REPO, "codec-arena_amd", "csrc")], cxx=True)
    out = _run([exe, os.path.join(golden, "c0_model.txt"), os.path.join(golden, "c0_elites.txt"), str(tmp_path)])
    assert out.strip().endswith("ok")
    assert (tmp_path / "c.txt").read_text() == "2650000000\n"
    assert len((tmp_path / "p.txt").read_text().splitlines()) == 40


def test_threaded_code_under_thread_sanitizer(tmp_path, golden):
    # the oracle's screens run one thread per job; the codec label source is
    # called from several threads at once by the label and funnel code
    exe = _build(tmp_path, "oracle_tsan", [os.path.join(REPO, "tests", "sanitize", "oracle_drv.c"),
                                          os.path.join(REPO, "oracle", "screen_oracle.c")],
                 os.path.join(REPO, "oracle"), flags=TSAN)
    assert _run([exe, os.path.join(golden, "c0_model.txt"), os.path.join(golden, "c0_elites.txt")]).strip().endswith("ok")
    exe = _build(tmp_path, "codec_tsan", [os.path.join(REPO, "tests", "sanitize", "codec_threads_drv.c"),
                                         os.path.join(REPO, "codec-arena_amd", "csrc", "codec_eval.c")],
                 os.path.join(REPO, "include"), flags=TSAN)
    assert _run([exe]).strip().endswith("ok")
