"""The C ABI from plain C: tests/capi_check.c is compiled with gcc -std=c11 -Wall -Wextra
-Werror -pedantic against include/pentago_amd.h, linked against libpentago_amd.so, and run
(host-only entry points: tables, sections, slice sizes and groups, line state, board state)."""
import os
import subprocess
import tempfile

import pentago_amd as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_c_program_uses_the_abi():
    libdir = os.path.dirname(P.LIB_PATH)
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "capi_check")
        subprocess.check_call(["gcc", "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic",
                               "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "capi_check.c"),
                               "-o", exe, "-L", libdir, "-lpentago_amd", "-Wl,-rpath," + libdir])
        out = subprocess.run([exe], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0, out.stdout + out.stderr
        assert "capi_check ok" in out.stdout
