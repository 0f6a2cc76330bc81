"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer.

- The CPU oracle (oracle/rtm_oracle.cpp), driven at the edges of its domain by
  tests/cpp/oracle_san.cpp, runs without a report.
- These builds do find the class of defect they are here for: the reference
  engine compiled the same way (oracle/Makefile, target ref-asan) reports its
  read of row -1 in the free-surface elastic step (wave2d.c:133-135), while
  the oracle's restatement of that step, which reads zeros there, does not.
- The engine's own host code, built with the sanitizers
  (python -m das_amd.build --sanitize), refuses every bad argument that
  tests/cpp/host_san.cpp passes it without a report.  The same driver runs
  the GPU paths in tests/test_gpu_sanitized_host.py.
The serving arena's check runs sanitized in tests/test_arena_cpu.py."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
REF_SRC = "/root/reference/tsw_rtm/c_engine/wave2d.c"
SAN = ["-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=all"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
           UBSAN_OPTIONS="print_stacktrace=1")

needs_gxx = pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")


def _gxx(tmp_path, name, sources, extra=()):
    exe = tmp_path / name
    subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", *SAN, *extra, *sources,
                    "-o", str(exe)], check=True)
    return str(exe)


@needs_gxx
def test_oracle_clean_under_asan_ubsan(tmp_path):
    exe = _gxx(tmp_path, "oracle_san", [os.path.join(HERE, "cpp", "oracle_san.cpp"),
                                        os.path.join(ROOT, "oracle", "rtm_oracle.cpp")])
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300, env=ENV)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stdout + r.stderr[-4000:]


@needs_gxx
@pytest.mark.skipif(not os.path.exists(REF_SRC), reason="reference source not present")
def test_sanitizer_reports_the_reference_row_minus_one_read(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "ref-asan"], check=True)
    ref_dir = os.path.join(ROOT, "oracle", "_ref")
    drv = os.path.join(HERE, "cpp", "elastic_rows_san.cpp")
    ref = _gxx(tmp_path, "rows_ref", [drv], ["-DREFERENCE_ENGINE", "-L", ref_dir,
                                             "-lwave2d_ref_asan", f"-Wl,-rpath,{ref_dir}"])
    r = subprocess.run([ref], capture_output=True, text=True, timeout=120, env=ENV)
    assert r.returncode != 0
    assert "heap-buffer-overflow" in r.stderr and "wave2d.c:135" in r.stderr, r.stderr[-4000:]
    ora = _gxx(tmp_path, "rows_oracle", [drv, os.path.join(ROOT, "oracle", "rtm_oracle.cpp")])
    r = subprocess.run([ora], capture_output=True, text=True, timeout=120, env=ENV)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr[-4000:]


def test_engine_host_refusals_under_asan_ubsan():
    from das_amd import build

    if not os.path.exists(build.SAN_EXE):
        build.build_sanitized()
    r = subprocess.run([build.SAN_EXE, "args"], capture_output=True, text=True, timeout=300,
                       env=ENV, cwd="/tmp")
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), r.stdout + r.stderr[-4000:]
