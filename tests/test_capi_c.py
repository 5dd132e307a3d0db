"""The C-ABI is usable from C: include/mpcf_hip.h compiles as C99 with
warnings as errors, a C program links against libmpcf_hip.so, and the calls
that need no device (argument errors, version, sizes, the operation count)
behave as declared."""
import os
import shutil
import subprocess

import pytest

from conftest import ROOT

PKG = os.path.join(ROOT, "cubism-jr-gpu-mpcf_amd")

C_SRC = r"""
#include <stdio.h>
#include "mpcf_hip.h"

int main(void)
{
    mpcf_ctx* ctx = 0;
    long long bytes = 0;
    double flops[16];
    printf("%s\n", mpcf_version());
    if (mpcf_create(NULL, 8, 8, 8, 0.1, 0) != MPCF_EINVAL) return 1;
    if (mpcf_create(&ctx, 4, 8, 8, 0.1, 0) != MPCF_EINVAL) return 2;
    if (ctx != NULL) return 3;
    if (mpcf_step(NULL, 0.3, 1e30, NULL) != MPCF_EINVAL) return 4;
    if (mpcf_set_pipeline(NULL, 4) != MPCF_EINVAL) return 5;
    if (mpcf_case_bytes(64, 64, 64, &bytes) != MPCF_OK || bytes <= 0) return 6;
    if (mpcf_flop_count(MPCF_FLAG_FAST, flops) != MPCF_OK || !(flops[0] > 0)) return 7;
    printf("ok %lld %.0f\n", bytes, flops[0]);
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="no C compiler")
def test_header_compiles_as_c_and_links(mpcf, tmp_path):
    lib = os.path.join(PKG, "libmpcf_hip.so")
    if not os.path.exists(lib):
        pytest.skip("library not built")
    src = tmp_path / "use_mpcf.c"
    exe = tmp_path / "use_mpcf"
    src.write_text(C_SRC)
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-I", os.path.join(ROOT, "include"), str(src),
           "-o", str(exe), "-L", PKG, "-lmpcf_hip", "-Wl,-rpath," + PKG, "-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    run = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, (run.returncode, run.stdout, run.stderr)
    assert "ok" in run.stdout
