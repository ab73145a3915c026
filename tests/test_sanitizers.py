"""CPU builds of the oracle and of the modeling driver's host I/O with
AddressSanitizer and UndefinedBehaviorSanitizer (every report fatal), run on
small inputs that include the reference's undefined cases: a constant
stiffness component (0/0 in the quantiser), empty parameter-file lines, short
binary files, sources and receivers on the model's last row and column."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")
FLAGS = ["-std=c++17", "-O1", "-g", "-fno-omit-frame-pointer", "-ffp-contract=off", "-fsanitize=address,undefined,float-cast-overflow",
         "-fno-sanitize-recover=all"]


def _build_run(tmp, exe, srcs, extra, args=()):
    out = os.path.join(tmp, exe)
    subprocess.run(["g++", *FLAGS, *extra, "-o", out, *srcs], check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([out, *args], capture_output=True, text=True, env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    return r.stdout


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_oracle_under_asan_ubsan(tmp_path):
    out = _build_run(str(tmp_path), "oracle_san", [os.path.join(SAN, "oracle_main.cpp")], ["-fopenmp"])
    assert "ok" in out


@pytest.mark.skipif(shutil.which("g++") is None, reason="no g++")
def test_host_io_under_asan_ubsan(tmp_path):
    host = os.path.join(ROOT, "seisfat2d_amd", "host")
    srcs = [os.path.join(SAN, "host_main.cpp"), os.path.join(host, "admin.cpp"), os.path.join(host, "geometry.cpp")]
    out = _build_run(str(tmp_path), "host_san", srcs, [], (str(tmp_path),))
    assert "ok" in out
