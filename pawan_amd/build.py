"""Build the native parts in-tree: libpawan_hip.so (gfx950 HIP) and the C++
adapter/driver objects.  Compiles every time it is called."""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _make(path: str, jobs: int = 8) -> None:
    subprocess.run(["make", "-B", "-j", str(jobs), "-C", path], check=True)


def build_library() -> str:
    _make(os.path.join(HERE, "csrc"))
    return os.path.join(HERE, "libpawan_hip.so")


def build_oracle() -> str:
    _make(os.path.join(ROOT, "oracle"))
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_adapter_driver() -> str:
    """The C++ drop-in adapter's test driver (needs both libraries above)."""
    _make(os.path.join(ROOT, "tests", "cxx"))
    return os.path.join(ROOT, "tests", "cxx", "adapter_driver")


if __name__ == "__main__":
    build_library()
    build_oracle()
    build_adapter_driver()
