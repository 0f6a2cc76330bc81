"""The serving loop's arena (das_amd/csrc/arena.h) on the CPU: with the
request sizes of the serving GPU test and its budget, FIFO admission never
stalls in any completion order while results are held until the end, and
released blocks coalesce."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


def _up(b):
    b = b if b else 4
    return (b + 255) & ~255


def _sizes():
    """Working and result block bytes of the six requests of
    tests/test_gpu_serve.py (the layout of req_heavy / req_keep in serve.hip),
    and that test's budget."""
    specs = [("a", 64, 200, 130), ("e", 40, 96, 70), ("a", 96, 160, 90), ("a", 48, 256, 151),
             ("e", 56, 120, 61), ("a", 70, 300, 77)]
    k = 2
    out, need = [], []
    for kind, nz, nx, nt in specs:
        n = nx * nz
        nrec = len(range(1, nx - 1, 2))
        nb = _up(4 * n)
        nsnap = (nt - 1) // k + 1
        b = (_up(4 * nt) + _up(4 * nt * nrec) + _up(4 * (nx + 1)) + _up(4 * max(nrec, 1))
             + _up(4 * nsnap * n))
        b += 9 * nb if kind == "a" else 8 * _up(4 * n) + 10 * nb
        out.append((b, 2 * _up(4 * n)))
        need.append((10 if kind == "a" else 20) * 4 * n + 4 * nsnap * n)
    return out, int(2.2 * max(need))


SANITIZE = ["-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
            "-fno-sanitize-recover=all"]


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
@pytest.mark.parametrize("flags", [[], SANITIZE], ids=["plain", "asan_ubsan"])
def test_arena_never_stalls_and_coalesces(tmp_path, flags):
    exe = tmp_path / "arena_check"
    subprocess.run(["g++", "-std=c++17", "-O1", *flags, "-I",
                    os.path.join(ROOT, "das_amd", "csrc"),
                    os.path.join(HERE, "cpp", "arena_check.cpp"), "-o", str(exe)], check=True)
    sizes, budget = _sizes()
    args = [str(v) for hk in sizes for v in hk] + [str(budget)]
    r = subprocess.run([str(exe)] + args, capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("ok"), r.stdout + r.stderr
