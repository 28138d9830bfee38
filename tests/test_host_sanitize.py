"""The host C code of libbfamx under AddressSanitizer and UBSan: operators,
brick and vmaps, face connectivity, the VTK writer and the octree forest
with its mortars, built with gcc from the library's own sources and driven
over a range of orders and sizes (tests/host_sanitize.c)."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
HOST = os.path.join(ROOT, "bfam_amd", "csrc", "host")
SOURCES = ["common.c", "lsrk_pcg.c", "mesh.c", "octree.c", "ops.c", "vtk.c"]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not found")
def test_host_code_clean_under_sanitizers(tmp_path):
    exe = tmp_path / "host_sanitize"
    cmd = ["gcc", "-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-fno-omit-frame-pointer", "-std=gnu99",
           "-I" + os.path.join(ROOT, "include"), "-I" + os.path.join(ROOT, "bfam_amd", "csrc"),
           os.path.join(HERE, "host_sanitize.c")] + [os.path.join(HOST, s) for s in SOURCES] + \
          ["-lm", "-lz", "-o", str(exe)]
    b = subprocess.run(cmd, capture_output=True, text=True)
    if b.returncode != 0 and "asan" in (b.stderr or "").lower():
        pytest.skip("AddressSanitizer runtime not available: " + b.stderr[-300:])
    assert b.returncode == 0, b.stderr[-2000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([str(exe), str(tmp_path / "h.vtu")], capture_output=True, text=True,
                       env=env, timeout=300)
    assert r.returncode == 0, (r.stdout[-1000:], r.stderr[-3000:])
    assert "harness done" in r.stdout
