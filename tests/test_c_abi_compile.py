"""The C ABI from C: include/claire_amd.h compiles as C99, and a C program
linked against libclaire_amd.so gets the reference's solver defaults and the
documented argument errors (no GPU is touched)."""
import os
import shutil
import subprocess

import pytest

import claire_amd

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "claire_amd.h"

int main(void) {
    clr_gn_options opt;
    if (clr_gn_options_default(&opt) != CLR_OK) return 10;
    if (opt.maxit != 50 || opt.krylov_maxit != 50 || opt.gttol != 5e-2 || opt.gatol != 1e-6) return 11;
    if (opt.ls_sigma != 1e-4 || opt.ls_stepinit != 1.0 || opt.krylov_block < 1) return 12;
    if (clr_create(NULL, NULL, 4, 0, NULL, NULL) != CLR_ERR_INVALID_ARGUMENT) return 13;
    const int big[3] = {256, 256, 256};
    float dummy = 0.0f;
    if (clr_grad_noalloc(big, &dummy, &dummy, NULL) != CLR_ERR_UNSUPPORTED_SIZE) return 14;
    if (strstr(clr_last_error(), "scratch") == NULL) return 15;
    const int odd[3] = {64, 64, 100};
    if (clr_grad_noalloc(odd, &dummy, &dummy, NULL) != CLR_ERR_UNSUPPORTED_SIZE) return 16;
    printf("ok\n");
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_header_is_c99_and_a_c_program_links(tmp_path):
    if not os.path.exists(claire_amd.LIB_PATH):
        pytest.fail("libclaire_amd.so not built: run __graft_entry__.build()")
    inc = os.path.join(ROOT, "include")
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-fsyntax-only", "-x", "c",
                    os.path.join(inc, "claire_amd.h")], check=True)
    src = tmp_path / "use_abi.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "use_abi"
    libdir = os.path.dirname(claire_amd.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", f"-I{inc}", str(src), "-o", str(exe), f"-L{libdir}",
                    "-lclaire_amd", f"-Wl,-rpath,{libdir}"], check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
    assert r.stdout.strip() == "ok"
