"""The C ABI called from C: tests/c/abi_client.c, built with gcc against
include/bml_hip.h and linked with libbml_hip.so.

CPU: it compiles (with -Werror against the header's prototypes, so calls in
the reference's argument order type-check) and links; where bml's sources
are present it also compiles against the reference's struct definitions and
static_asserts the offsets of the members bml's dispatchers read.
GPU: the program's own checks (x2 on the three row engines, add, trace,
threshold, M overflow, ELLBLOCK x2, SP2) against dense loops in C.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "abi_client.c")
LIBDIR = os.path.join(ROOT, "bml_amd")
REF_SRC = "/root/reference/src/C-interface"


def build(out, extra=()):
    cmd = ["gcc", "-std=c11", "-O1", "-Wall", "-Werror", *extra, SRC, "-o", out,
           "-L" + LIBDIR, "-lbml_hip", "-Wl,-rpath," + LIBDIR, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def test_c_client_builds_and_links(tmp_path):
    exe = build(str(tmp_path / "abi_client"))
    r = subprocess.run([exe, "link"], capture_output=True, text=True)
    assert r.returncode == 0 and "linked" in r.stdout, r.stderr


@pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="bml's sources are not present")
def test_struct_prefix_matches_reference(tmp_path):
    build(str(tmp_path / "abi_client_ref"), ["-DBML_REF_INCLUDE", "-I" + REF_SRC])


@pytest.mark.gpu
def test_c_client_runs_on_gpu(tmp_path):
    exe = build(str(tmp_path / "abi_client"))
    r = subprocess.run([exe, "run"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "all checks passed" in r.stdout
