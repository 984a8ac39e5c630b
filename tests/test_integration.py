"""The reference-side binding (mav1-encoder_amd/host/backend_hip.c, the
Backend vtable of the encoder's src/backend.h) compiles against the
encoder's own headers and needs nothing but the C-ABI of libmav1hip and
libc; the HIP-less stub compiles too. Skipped where the encoder sources
are absent."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF_SRC = "/root/reference/src"
HOST = os.path.join(ROOT, "mav1-encoder_amd", "host")

needs_src = pytest.mark.skipif(not os.path.isdir(REF_SRC), reason="encoder sources not present")


@needs_src
def test_backend_hip_compiles_against_encoder_headers(tmp_path):
    obj = tmp_path / "backend_hip.o"
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-Wall", "-Werror", "-c", os.path.join(HOST, "backend_hip.c"),
                    "-I" + REF_SRC, "-I" + os.path.join(ROOT, "include"), "-o", str(obj)], check=True)
    undef = subprocess.run(["nm", "-u", str(obj)], check=True, capture_output=True, text=True).stdout.split()
    names = [u for u in undef if u not in ("U",)]
    allowed_libc = {"fprintf", "__fprintf_chk", "stderr", "calloc", "free", "__stack_chk_fail", "memset", "memcpy"}
    unexpected = [n for n in names if not n.startswith("mav1hip_") and n not in allowed_libc]
    assert not unexpected, unexpected
    defined = subprocess.run(["nm", "--defined-only", str(obj)], check=True, capture_output=True, text=True).stdout
    assert " T backend_hip_create" in defined


@needs_src
def test_stub_compiles(tmp_path):
    subprocess.run(["gcc", "-O2", "-std=gnu11", "-Wall", "-Werror", "-c", os.path.join(HOST, "backend_hip_stub.c"),
                    "-I" + REF_SRC, "-o", str(tmp_path / "stub.o")], check=True)
