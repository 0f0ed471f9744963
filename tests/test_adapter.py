"""The C++ drop-in adapter (include/pawan_gpu.hpp) over a test double of
pawan's class API: compiles on the host; on the GPU its driver replays
pawan's solve()/integrate() call sequence and checks against the oracle."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CXX = os.path.join(ROOT, "tests", "cxx")


def test_adapter_header_compiles():
    src = '#include "mini_pawan.hpp"\n#include "../../include/pawan_gpu.hpp"\n' \
          "template class pawan_gpu::interaction<mini::interaction, mini::wake, mini::vec>;\n" \
          "template class pawan_gpu::rk4<mini::rk4, mini::system, mini::vec>;\n" \
          "template class pawan_gpu::euler<mini::integration, mini::system, mini::vec>;\n"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Wextra", "-Werror", "-fsyntax-only",
                        "-x", "c++", "-I", CXX, "-"], input=src, text=True, cwd=CXX,
                       capture_output=True)
    assert r.returncode == 0, r.stderr


@pytest.mark.gpu
def test_adapter_driver_on_gpu(tmp_path):
    exe = os.path.join(CXX, "adapter_driver")
    assert os.path.exists(exe), "build it first: make -C tests/cxx"
    r = subprocess.run([exe, str(tmp_path / "adapter.wake")], capture_output=True, text=True,
                       timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    assert "driver: passed" in r.stdout
