"""Build consumers of the library the way the reference's users do: through
the CMake package (find_package(linkcell) -> linkcell::linkcell / ::cpp /
::gpu, CMakeLists.txt:144-163) and through the pkg-config files
(cmake/linkcell.pc.in, cmake/linkcell-gpu.pc.in). The consumers only call
what runs without a GPU: lc_version, lc_gpu_available and an argument error
of the C waist."""
import os
import re
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
PKG = ROOT / "linkcell_amd"
LIB = PKG / "liblinkcell_amd.so"

C_MAIN = r"""
#include <stdio.h>
#include <string.h>
#include <linkcell.h>
#include <linkcell_gpu.h>
#ifndef LINKCELL_HAS_GPULITE
#error "linkcell::gpu must define LINKCELL_HAS_GPULITE"
#endif
int main(void) {
  const double xyz[6] = {0.0, 0.0, 0.0, 1.0, 1.0, 1.0};
  lc_cell box = lc_cell_ortho(10.0, 10.0, 10.0);
  int nn[2];
  int rc = lc_knearest(xyz, 2, &box, 0, NULL, 0.0, nn);
  printf("version=%s available=%d rc=%d err=%s\n", lc_version(), lc_gpu_available(),
         rc, lc_last_error());
  return rc == 1 && strcmp(lc_last_error(), "k must be at least 1") == 0 ? 0 : 3;
}
"""

CPP_MAIN = r"""
#include <cstdio>
#include <linkcell.hpp>
#include <linkcell_gpu.hpp>
int main() {
  std::printf("cpp available=%d version=%s\n", linkcell::gpu::available() ? 1 : 0,
              lc_version());
  return 0;
}
"""

CMAKE = """
cmake_minimum_required(VERSION 3.16)
project(consumer C CXX)
find_package(linkcell 0.3 CONFIG REQUIRED)
add_executable(c_consumer main.c)
target_link_libraries(c_consumer PRIVATE linkcell::gpu)
add_executable(cpp_consumer main.cpp)
target_link_libraries(cpp_consumer PRIVATE linkcell::gpu linkcell::cpp)
"""


def _need_lib():
    if not LIB.exists():
        pytest.skip("library not built")


def _run(cmd, cwd):
    res = subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert res.returncode == 0, res.stdout + res.stderr
    return res.stdout


@pytest.mark.skipif(shutil.which("cmake") is None, reason="cmake not installed")
def test_cmake_package_targets(tmp_path):
    _need_lib()
    (tmp_path / "main.c").write_text(C_MAIN)
    (tmp_path / "main.cpp").write_text(CPP_MAIN)
    (tmp_path / "CMakeLists.txt").write_text(CMAKE)
    b = tmp_path / "build"
    env_cmd = ["cmake", "-S", str(tmp_path), "-B", str(b),
               f"-Dlinkcell_DIR={PKG / 'cmake'}", "-DCMAKE_C_COMPILER=gcc",
               "-DCMAKE_CXX_COMPILER=g++"]
    _run(env_cmd, tmp_path)
    _run(["cmake", "--build", str(b), "-j", "4"], tmp_path)
    out = _run([str(b / "c_consumer")], tmp_path)
    assert "version=0.3.3" in out and "rc=1" in out
    out = _run([str(b / "cpp_consumer")], tmp_path)
    assert re.search(r"cpp available=[01] version=0\.3\.3", out)


def test_cmake_package_version_gate(tmp_path):
    """find_package(linkcell 0.4) must not accept this 0.3.3 package."""
    if shutil.which("cmake") is None:
        pytest.skip("cmake not installed")
    _need_lib()
    (tmp_path / "CMakeLists.txt").write_text(
        "cmake_minimum_required(VERSION 3.16)\nproject(v NONE)\n"
        "find_package(linkcell 0.4 CONFIG QUIET)\n"
        "if(linkcell_FOUND)\n  message(FATAL_ERROR \"0.4 accepted\")\nendif()\n")
    _run(["cmake", "-S", str(tmp_path), "-B", str(tmp_path / "b"),
          f"-Dlinkcell_DIR={PKG / 'cmake'}"], tmp_path)


def _pc_flags(name, field):
    """Cflags / Libs of a .pc file, with ${pcfiledir} and the file's own
    variables expanded (pkg-config itself when installed)."""
    pcdir = PKG / "pkgconfig"
    if shutil.which("pkg-config"):
        env = dict(os.environ, PKG_CONFIG_PATH=str(pcdir))
        opt = "--cflags" if field == "Cflags" else "--libs"
        res = subprocess.run(["pkg-config", opt, name], capture_output=True, text=True,
                             env=env, check=True)
        return res.stdout.split()
    text = (pcdir / f"{name}.pc").read_text()
    var = {"pcfiledir": str(pcdir)}
    val = None
    for line in text.splitlines():
        line = line.strip()
        if not line or line.startswith("#"):
            continue
        m = re.match(r"^(\w+)=(.*)$", line)
        if m:
            var[m.group(1)] = m.group(2)
        elif line.startswith(field + ":"):
            val = line.split(":", 1)[1]
    for _ in range(4):
        for k, v in var.items():
            val = val.replace("${" + k + "}", v)
            var = {kk: vv.replace("${" + k + "}", v) for kk, vv in var.items()}
    return val.split()


def test_pkg_config_files_build_a_consumer(tmp_path):
    _need_lib()
    if shutil.which("gcc") is None:
        pytest.skip("gcc not installed")
    (tmp_path / "main.c").write_text(C_MAIN)
    cflags = _pc_flags("linkcell-gpu", "Cflags")
    libs = _pc_flags("linkcell-gpu", "Libs")
    assert "-DLINKCELL_HAS_GPULITE=1" in cflags
    exe = tmp_path / "c_consumer"
    _run(["gcc", "main.c", "-o", str(exe), *cflags, *libs], tmp_path)
    out = _run([str(exe)], tmp_path)
    assert "version=0.3.3" in out
