"""The streaming blocked kernel (das_amd/csrc/ts.h) and the blocked kernel with
its levels in LDS (das_amd/csrc/tc.h) run on the CPU through the HIP stand-in
in tests/cpp/hip_emu, against the tiled blocked kernel (das_amd/csrc/tb.h) run
the same way: bit-identical illumination, image and snapshots over every block
of a set of awkward shots (tests/cpp/ts_emu.cpp), for a streaming
configuration with one depth segment per shot and wide strips and one with
several short segments and one-wave strips, and two tile shapes of k_tc.  The
GPU tests pin k_tb to the per-step kernel and the oracle, so this pins the
other two kernels' indexing, buffer rotation, barriers and edge handling
without a GPU."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)

# (T, WX, H): sub-steps, 64-column waves per strip, core rows per segment of
# k_ts; (T, R): sub-steps and rows per thread of k_tc
CONFIGS = [((5, 2, 100), (5, 6)), ((4, 1, 30), (4, 6))]


@pytest.mark.skipif(shutil.which("g++") is None, reason="g++ not found")
def test_streaming_kernel_equals_tiled_kernel_on_cpu(tmp_path):
    procs = []
    for (T, WX, H), (TC, RC) in CONFIGS:
        exe = str(tmp_path / f"ts_emu_{T}_{WX}_{H}")
        subprocess.run(["g++", "-O1", "-std=c++17", "-w",
                        "-I", os.path.join(HERE, "cpp", "hip_emu"),
                        "-I", os.path.join(ROOT, "include"),
                        "-I", os.path.join(ROOT, "das_amd", "csrc"),
                        f"-DTS_T={T}", f"-DTS_WX={WX}", f"-DTS_H={H}", f"-DTC_T={TC}", f"-DTC_R={RC}",
                        os.path.join(HERE, "cpp", "ts_emu.cpp"), "-o", exe], check=True)
        procs.append(((T, WX, H), subprocess.Popen([exe, "0.6"], stdout=subprocess.PIPE,
                                                   stderr=subprocess.STDOUT, text=True)))
    for cfg, p in procs:
        out, _ = p.communicate(timeout=600)
        assert p.returncode == 0, f"k_ts / k_tc {cfg} differ from k_tb:\n{out}"
        assert out.strip().splitlines()[-1].startswith("OK"), out
