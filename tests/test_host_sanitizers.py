"""The host code that reads untrusted input, built with AddressSanitizer and UndefinedBehaviorSanitizer: the STL
parser and normaliser (aerosim_amd/csrc/stl.cpp) and the mesh statistics and frontal-area raster (geometry.cpp)
under tests/sanitize/host_fuzz.cpp's seeded random, truncated and mutated STL files and degenerate meshes. Any
sanitizer report aborts the harness. CPU only (g++)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_under_address_and_undefined_behaviour_sanitizers(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ is absent")
    exe = str(tmp_path / "host_fuzz")
    src = [os.path.join(ROOT, "tests", "sanitize", "host_fuzz.cpp"),
           os.path.join(ROOT, "aerosim_amd", "csrc", "stl.cpp"), os.path.join(ROOT, "aerosim_amd", "csrc", "geometry.cpp")]
    # geometry.cpp shares the lattice header with the kernels; its HIP qualifiers need only the runtime headers
    cmd = [gxx, "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-D__HIP_PLATFORM_AMD__",
           "-I" + os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "include"),
           "-fsanitize=address,undefined", "-fno-sanitize-recover=all", *src, "-o", exe]
    b = subprocess.run(cmd, capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe, "300"], capture_output=True, text=True, env=env, timeout=600)
    print(r.stdout[-500:])
    assert r.returncode == 0, (r.stdout[-2000:], r.stderr[-4000:])
    assert "host fuzz OK" in r.stdout
