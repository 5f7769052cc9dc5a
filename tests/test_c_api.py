"""The drop-in from C: tests/c/api_smoke.c is compiled with gcc against
include/vkhel/vkhel.h and linked to vkhel_amd/lib/libvkhel.so (CPU: build
only), then run on the GPU."""
import os
import subprocess

import pytest

from vkhel_amd import lib as L

SRC = os.path.join(L.REPO, "tests", "c", "api_smoke.c")


def build(out):
    libdir = os.path.dirname(L.LIB_PATH)
    subprocess.run(["gcc", "-std=c11", "-Wall", "-Werror", "-I",
                    os.path.join(L.REPO, "include"), SRC, "-o", str(out),
                    "-L", libdir, "-lvkhel", "-Wl,-rpath," + libdir],
                   check=True)


def test_c_caller_builds_against_public_header(tmp_path):
    build(tmp_path / "api_smoke")
    assert (tmp_path / "api_smoke").exists()


@pytest.mark.gpu
def test_c_caller_runs_reference_checks(tmp_path):
    exe = tmp_path / "api_smoke"
    build(exe)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    assert r.stdout.count("passed") == 9, r.stdout
