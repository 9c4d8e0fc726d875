"""Host-side sanitizer runs (CPU only). The oracle is built with ASan and
UBSan, and separately with TSan for its threaded driver, and runs seeded
cases against its own brute force (tests/sanitize/oracle_check.c). The
library's host code (lc_api.hip, the engine's host side) is built with ASan
and UBSan on the host pass only and runs the reference's ABI contract with
no GPU (tests/sanitize/capi_check.c)."""
import os
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
SAN = ROOT / "tests" / "sanitize"
ORACLE = ROOT / "oracle" / "lc_oracle.c"


def _run(cmd, env=None, timeout=300):
    return subprocess.run(cmd, capture_output=True, text=True, timeout=timeout, env=env)


def _need(tool):
    if shutil.which(tool) is None:
        pytest.skip(f"{tool} not found")


def test_oracle_under_asan_and_ubsan(tmp_path):
    _need("gcc")
    exe = tmp_path / "oracle_asan"
    b = _run(["gcc", "-std=c11", "-O1", "-g", "-ffp-contract=off", "-fno-omit-frame-pointer",
              "-fsanitize=address,undefined", "-fno-sanitize-recover=all", str(ORACLE),
              str(SAN / "oracle_check.c"), "-lm", "-pthread", "-o", str(exe)])
    assert b.returncode == 0, b.stderr
    r = _run([str(exe)], env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(": ok") == 8 and "ERROR" not in r.stderr


def test_oracle_threads_under_tsan(tmp_path):
    _need("gcc")
    exe = tmp_path / "oracle_tsan"
    b = _run(["gcc", "-std=c11", "-O1", "-g", "-ffp-contract=off", "-fsanitize=thread",
              str(ORACLE), str(SAN / "oracle_check.c"), "-lm", "-pthread", "-o", str(exe)])
    assert b.returncode == 0, b.stderr
    r = _run([str(exe), "threads"])
    if "unexpected memory mapping" in r.stderr:
        pytest.skip("ThreadSanitizer cannot map its shadow memory here")
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(": ok") == 4 and "WARNING: ThreadSanitizer" not in r.stderr


def test_library_host_code_under_asan_and_ubsan(tmp_path):
    from linkcell_amd import _lib, build as b
    if _lib.load().lc_gpu_available() != 0:
        pytest.skip("a HIP device is present: the no-device contract is what this runs")
    clang = Path(b.hipcc()).resolve().parent.parent / "llvm" / "bin" / "clang"
    if not clang.exists():
        clang = Path("/opt/rocm/llvm/bin/clang")
    if not clang.exists():
        pytest.skip("ROCm clang not found")
    # the host pass only is sanitized (no device code is built with it)
    flags = [f for f in b.HIPFLAGS if f != "-O3"] + ["-O1", "-Xarch_host", "-fsanitize=address,undefined",
                                                     "-Xarch_host", "-fno-omit-frame-pointer"]
    procs, objs = [], []
    for src in b.SOURCES:
        obj = tmp_path / (Path(src).stem + ".o")
        objs.append(str(obj))
        procs.append(subprocess.Popen([b.hipcc(), *flags, "-I", str(ROOT / "include"), "-c",
                                       str(b.CSRC / src), "-o", str(obj)],
                                      stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True))
    for p in procs:
        _, err = p.communicate(timeout=600)
        assert p.returncode == 0, err[-2000:]
    so = tmp_path / "liblc_asan.so"
    rocm_lib = Path(b.hipcc()).resolve().parent.parent / "lib"
    link = _run([str(clang) + "++", "-shared", "-fPIC", "-fsanitize=address,undefined",
                 "-shared-libsan", *objs, f"-L{rocm_lib}", "-lamdhip64", "-o", str(so)])
    assert link.returncode == 0, link.stderr
    rt = _run([str(clang), "-print-resource-dir"]).stdout.strip() + "/lib/linux"
    exe = tmp_path / "capi_check"
    drv = _run([str(clang), "-std=c11", "-O1", "-g", "-fno-omit-frame-pointer",
                "-fsanitize=address,undefined", "-shared-libsan", "-fno-sanitize-recover=all",
                "-I", str(ROOT / "include"), str(SAN / "capi_check.c"), f"-L{tmp_path}",
                "-llc_asan", f"-Wl,-rpath,{tmp_path}", f"-Wl,-rpath,{rt}",
                f"-Wl,-rpath,{rocm_lib}", "-pthread", "-lm", "-o", str(exe)])
    assert drv.returncode == 0, drv.stderr
    r = _run([str(exe)], env={**os.environ, "ASAN_OPTIONS": "detect_leaks=1"})
    assert r.returncode == 0, r.stdout + r.stderr
    assert "capi host checks: ok" in r.stdout
