"""One run of the batched acoustic RTM workload's blocked kernel, for running
under a profiler (for example rocprofv3 --pmc, one counter per run):

    python -m das_amd.profile_batch --scale 0.05

The 64 windows of the batched workload (nz 300, nx 600..1600) are all
resident in 64 slots, as in the timed batch, so every launch covers the same
tiles as there; only nt is shortened (x `scale`, about 150..300 steps).  No
hipGraphs and no autotuning: the blocked kernel `--kernel` (configuration 15
by default) is launched directly, once per block of T steps, and nothing
else of the library runs.  Prints the launch count and the cell-updates, the
denominators of the counted bytes.
"""
import argparse
import json


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--scale", type=float, default=0.05, help="nt scale (the workload has 1.0)")
    ap.add_argument("--kernel", type=int, default=15)
    ap.add_argument("--shots", type=int, default=64)
    args = ap.parse_args()
    import torch

    from das_amd import _native as N
    from das_amd.batch import ShotBatch
    from das_amd.workloads import BatchedWindows

    N.check(N.lib().tsw_init(0))
    cfg = BatchedWindows(n_shots=args.shots).build(with_data=True, scale=args.scale)
    b = ShotBatch(cfg.shots, cfg.dx, cfg.dx, cfg.dt, n_sponge=cfg.n_sponge, k=cfg.k,
                  slots=args.shots, graph_steps=0, kernel=args.kernel)
    info = b.info()
    print(f"batch resident: {info}", flush=True)
    b.run("batched")
    torch.cuda.synchronize()
    ms, launches = b.time_kernels()
    b.close()
    print(json.dumps({"kernel": args.kernel, "shots": args.shots, "scale": args.scale,
                      "cell_updates_per_run": info["cell_updates"],
                      "launches_per_run": launches, "runs": 2,
                      "timed_run_kernel_ms": round(ms, 3)}), flush=True)


if __name__ == "__main__":
    main()
