"""A plain C caller of the C ABI: include/voxelproj_hip.h compiles as strict
C99 with gcc, links against the library, and its argument checks answer
before any device work (so this runs without a GPU).  This is the binding a
C or C++ host program would use in place of the reference's pybind module
(voxelproj/csrc/voxelproj.cpp:11-34)."""
import os
import shutil
import subprocess

import pytest

from voxelproj_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

PROGRAM = r"""
#include <stdio.h>
#include <string.h>
#include "voxelproj_hip.h"

int main(void)
{
    float dummy[4] = {0.f, 0.f, 0.f, 0.f};
    int rc;
    if (vp_abi_version() < 1) return 10;
    /* z_order must be 0 or 2 */
    rc = vp_forward(1, dummy, dummy, dummy, 1, 1, 1, 1, 1, NULL);
    if (rc != VP_ERR_ARG || strstr(vp_last_error(), "z_order") == NULL) return 11;
    /* negative extents */
    rc = vp_backward(0, dummy, dummy, dummy, 1, -1, 1, 1, 1, NULL);
    if (rc != VP_ERR_ARG) return 12;
    /* an empty problem is a successful no-op */
    rc = vp_forward(0, NULL, NULL, NULL, 0, 4, 4, 4, 4, NULL);
    if (rc != VP_OK) return 13;
    /* workspace sizes need no device */
    if (vp_workspace_elems(0, 2, 3, 4, 5, 6, 7, 0) != 4 * 5 * 6 + 3 * 7 * 6) return 14;
    if (vp_workspace_elems(0, 0, 3, 4, 5, 6, 7, 0) != 0) return 15;
    printf("abi %d ok\n", vp_abi_version());
    return 0;
}
"""


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_c99_program_compiles_links_and_validates(tmp_path):
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    src = tmp_path / "caller.c"
    src.write_text(PROGRAM)
    exe = tmp_path / "caller"
    libdir = os.path.dirname(_lib.LIB_PATH)
    cmd = ["gcc", "-std=c99", "-Wall", "-Wextra", "-Werror", "-pedantic",
           "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe),
           "-L", libdir, "-l:" + os.path.basename(_lib.LIB_PATH), "-Wl,-rpath," + libdir]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=60)
    assert out.returncode == 0, (out.returncode, out.stdout, out.stderr)
    assert "ok" in out.stdout
