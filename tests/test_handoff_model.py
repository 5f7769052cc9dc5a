"""CPU model of the one-launch two-pass transform's protocol
(vkhel_amd/csrc/handoff.h, kernel k_ntt_handoff): tests/c/handoff_model.cpp
runs the kernel's own ticket schedule with host threads standing for a
bounded number of resident workgroups per XCD, producer tiles publishing
with a release increment and consumer tiles waiting with a bounded spin.
Every configuration must finish with no wait given up, every handoff
visible and every ticket done; a schedule with one consumer block moved
before its producers must be caught by the bounded spin."""
import os
import subprocess

from vkhel_amd import lib as L

SRC = os.path.join(L.REPO, "tests", "c", "handoff_model.cpp")


def _build(tmp_path):
    exe = tmp_path / "handoff_model"
    subprocess.run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-pthread",
                    "-I", os.path.join(L.REPO, "vkhel_amd", "csrc"), SRC,
                    "-o", str(exe)], check=True)
    return exe


def test_schedule_layout_and_protocol_model(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("model ok")
    assert "gave_up=0" in r.stdout and "stale=0" in r.stdout


def test_model_catches_a_consumer_before_its_producers(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([str(exe), "--planted"], capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "planted caught" in r.stdout
