"""A C program compiled and linked against include/syndrome_ldpc.h and
libsyndrome_amd.so by a C compiler (the reference's CI does this for its C
ABI, .github/workflows/ci.yml:198-243).

CPU: the header compiles as strict C11 and as C++, and a program that takes
the address of every entry point the header declares links against the
library, so each declared symbol is exported with a prototype the header
gives.  GPU: the program runs and prints the encoder round trip and the
golden checksum a2448e8cf5272825 (bench/results/ldpc_rust.checksum).
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "c", "abi_check.c")
LIBDIR = os.path.join(ROOT, "syndrome_amd")
HIPLIB = "/opt/rocm/lib"


def _build(tmp_path):
    from syndrome_amd._lib import declared_symbols
    if not os.path.exists(os.path.join(LIBDIR, "libsyndrome_amd.so")):
        pytest.skip("libsyndrome_amd.so is not built")
    cc = shutil.which("gcc") or shutil.which("cc")
    if not cc:
        pytest.skip("no C compiler")
    table = tmp_path / "abi_symbols.inc"
    names = declared_symbols()
    assert len(names) > 40
    table.write_text("typedef void (*abi_fn)(void);\nconst abi_fn abi_symbols[] = {\n" +
                     "".join(f"    (abi_fn){n},\n" for n in names) + "};\n")
    exe = tmp_path / "abi_check"
    cmd = [cc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"),
           f'-DABI_SYMBOL_TABLE="{table}"', SRC, "-o", str(exe), "-L", LIBDIR, "-lsyndrome_amd",
           f"-Wl,-rpath,{LIBDIR}", f"-Wl,-rpath-link,{HIPLIB}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_header_is_valid_cxx(tmp_path):
    cxx = shutil.which("g++")
    if not cxx:
        pytest.skip("no C++ compiler")
    src = tmp_path / "h.cpp"
    src.write_text('#include "syndrome_ldpc.h"\nint main() { return SYNDROME_OK; }\n')
    r = subprocess.run([cxx, "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(ROOT, "include"),
                        str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_c_program_links_every_declared_symbol(tmp_path):
    exe = _build(tmp_path)
    assert exe.exists()


@pytest.mark.gpu
def test_c_program_runs_golden_checksum(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, r.stderr
    assert "golden checksum a2448e8cf5272825 after 10 iterations" in r.stdout
    assert "C ABI ok" in r.stdout
