"""The host parsers under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/cpp/sanitize_parsers.cpp drives the MSH 4.1 reader and node-data writer
(csrc/msh_io.cpp) and the YAML reader with the material loader
(epp/yaml_reader.cpp, epp/Material.cpp) over the reference's mesh and material
files and over thousands of truncated and corrupted copies.  It is compiled
here with g++ (host code only, no device code is involved) and any memory
error, leak or undefined behaviour aborts it with a non-zero status.
"""
from __future__ import annotations

import glob
import gzip
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "ultimateepm_amd")
SOURCES = [os.path.join(HERE, "cpp", "sanitize_parsers.cpp"), os.path.join(PKG, "csrc", "msh_io.cpp"),
           os.path.join(PKG, "epp", "yaml_reader.cpp"), os.path.join(PKG, "epp", "Material.cpp")]
SAN = "-fsanitize=address,undefined,float-cast-overflow"


@pytest.fixture(scope="module")
def sanitized_driver(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("san")
    exe = str(d / "sanitize_parsers")
    cmd = ["g++", "-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", SAN, "-fno-sanitize-recover=all",
           "-static-libasan",
           "-D__HIP_PLATFORM_AMD__", "-I/opt/rocm/include", f"-I{os.path.join(ROOT, 'include')}",
           f"-I{os.path.join(PKG, 'csrc')}", f"-I{os.path.join(PKG, 'epp')}", *SOURCES, "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    msh = str(d / "bz_mesh.msh")
    with gzip.open(os.path.join(HERE, "golden", "bz_mesh.msh.gz"), "rb") as src, open(msh, "wb") as dst:
        shutil.copyfileobj(src, dst)
    return exe, msh, str(d)


def test_parsers_clean_under_asan_ubsan(sanitized_driver):
    exe, msh, work = sanitized_driver
    yamls = sorted(glob.glob(os.path.join(PKG, "params", "*.yaml")))
    assert yamls
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([exe, msh, work, *yamls], capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ALL PASSED" in r.stdout
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out
    # the corruption sweeps exercise both outcomes
    for line in r.stdout.splitlines():
        if line.startswith(("msh:", "yaml:")):
            read = int(line.split(",")[1].split()[0])
            refused = int(line.split(",")[2].split()[0])
            assert read > 0 and refused > 0, line
