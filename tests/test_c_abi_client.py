"""The drop-in boundary from C: tests/c_abi_client.c includes every public header as C99, links
against libplens_gpu.so and calls the host-side entry points (no GPU)."""
import os
import shutil
import subprocess

import pytest

from tests.conftest import ROOT


def test_c_client_builds_links_and_runs(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    from plens_amd import _lib as L
    L.load()  # builds the library when it is missing
    libdir = os.path.dirname(L.LIB_PATH)
    exe = str(tmp_path / "c_abi_client")
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_abi_client.c"), "-o", exe, "-L" + libdir, "-lplens_gpu",
           "-Wl,-rpath," + libdir, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr
