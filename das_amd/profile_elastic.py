"""One elastic P-SV RTM shot of the 400 x 1600 elastic workload through the
batched lane (one k_ela_fused launch per time step), for running under a
profiler (for example rocprofv3 --pmc, one counter per run):

    python -m das_amd.profile_elastic --nt 400

The gather is zero, so the run is the lane's initialisation, the forward
propagation with its snapshots and the back-propagation with its imaging,
and nothing else of the library: 2 nt + 1 launches of k_ela_fused, without
hipGraphs.
"""
import argparse

import numpy as np


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--nt", type=int, default=400, help="time steps (the workload has 8000)")
    args = ap.parse_args()
    from das_amd.batch import ShotBatch
    from das_amd.workloads import ElasticShot, PlainElasticShot

    cfg = PlainElasticShot(nt=args.nt)
    shot = ElasticShot(nx=cfg.nx, nz=cfg.nz, nt=cfg.nt, model_fwd=cfg.model_mig,
                       model_bwd=cfg.model_mig, src=cfg.src, six=cfg.six, siz=cfg.siz,
                       data=np.zeros((cfg.nt, len(cfg.rix)), np.float32), rix=cfg.rix,
                       riz=cfg.riz)
    b = ShotBatch([shot], cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge, k=cfg.k, graph_steps=0)
    b.run("batched")
    img, ill = b.shot_result(0)
    print(f"elastic shot {cfg.nz} x {cfg.nx}, nt {cfg.nt}: {2 * cfg.nt + 1} steps, "
          f"max illumination {float(np.abs(ill).max()):.3e}", flush=True)
    b.close()


if __name__ == "__main__":
    main()
