"""One FP32 misfit gradient of the 16-shot, 300 x 800 gradient workload, for
running under a profiler (for example rocprofv3 --pmc, one counter per run):

    python -m das_amd.profile_gradient --nt 300

The data are zero, so the run is the forward, the checkpointed recompute and
the adjoint of the blocked kernels at their default configurations, and
nothing else of the library.
"""
import argparse

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nt", type=int, default=300, help="time steps (the workload has 5000)")
    args = ap.parse_args()
    import torch

    from das_amd.fwi import MisfitShot, WaveformMisfit
    from das_amd.workloads import GradientShots

    cfg = GradientShots()
    nt = min(args.nt, cfg.nt)
    src = cfg.source()[:nt]
    shots = [MisfitShot(src=src, six=a, siz=b, rix=r, riz=z,
                        d_obs=np.zeros((nt, len(r)), np.float32))
             for a, b, r, z in cfg.geometry()]
    op = WaveformMisfit(cfg.nx, cfg.nz, cfg.dx, cfg.dx, cfg.dt, shots, n_sponge=cfg.n_sponge)
    v = torch.tensor(cfg.start_vel, dtype=torch.float32, device="cuda")
    op.forward_loss(v)
    g = op.gradient(v)
    torch.cuda.synchronize()
    print(f"gradient of {len(shots)} shots, nt {nt}: max |g| {float(g.abs().max()):.3e}", flush=True)
    op.close()


if __name__ == "__main__":
    main()
