"""LOS integrator benchmark on one MI355X.

Workload ("random_forward"): WMAP-LSA regular B, uniform thermal electrons,
analytic CRE; a global ES random B (rms 2 muG, k0 0.1, a0 1.7, seed 1) and a
DFT random TE (rms 0.001 cm^-3) on a 256x256x64 grid over
[-20,20]x[-20,20]x[-4,4] kpc, generated on the GPU; 3 shells at nside
64/128/256 (750/750/1500 steps of 0.01 kpc, 0-30 kpc); DM, FD and sync
I/Q/U at 23 GHz on nside 256.  FP64 throughout.

A step is one Integrator::write_grid pass (all shells, FD carry, shell
accumulation) with the grids already resident in HBM.  The value is LOS
samples per second: geometric samples sum_s npix_s * step_s, times the one
<sync> entry, per wall second.

roofline: the march kernel (los_fast, the specialised LSA + uniform TE +
analytic CRE march that the default schedule selects), algorithmic bytes = 256 B per
geometric sample (8 corners x {bx,by,bz,te} x 8 B), divided by its launch
durations measured with HIP events on the library stream.
cpu_baseline: the FP64 CPU oracle (OpenMP, schedule(dynamic) pixel loop) on
the same grids and radial sampling with 1/16 of the pixels.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
BYTES_PER_SAMPLE = 256


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="MIB",
                    help="only copy MIB MiB device-to-device three times (a plain copy "
                         "kernel of known size, to calibrate FETCH_SIZE / WRITE_SIZE)")
    return ap.parse_args()


def copy_calibrate(mib):
    import torch
    n = mib * (1 << 20) // 8
    a = torch.ones(n, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibrate_bytes": n * 8}), flush=True)


def traffic_from_profiles():
    """HBM bytes per march launch from the committed rocprofv3 PMC summary
    (profiles/los_fast_random_forward_pmc.txt), or None."""
    path = os.path.join(ROOT, "profiles", "los_fast_random_forward_pmc.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def cpu_baseline(par_text_fn, grids_host, threads):
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from hammurabix_amd import Param, los_samples
    from oracle_lib import write_grid
    par = Param(text=par_text_fn(nside_sim=16, nside_out=64))
    d = par.los_desc()
    bx, by, bz, te = grids_host
    box = par.grid_brnd.box()
    S = los_samples(d)
    t0 = time.perf_counter()
    write_grid(d, nthreads=threads, brnd=(bx, by, bz), brnd_box=box, ternd=te,
               ternd_box=par.grid_ternd.box())
    dt = time.perf_counter() - t0
    return dict(value=S / dt, unit="samples/s", cores=threads, kind="port",
                sample=f"random_forward with shells at nside 16/32/64 (1/16 of the "
                       f"pixels, {S} samples), same grids and radial sampling; "
                       f"{dt:.2f} s")


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the LOS path has no multi-GPU "
              "implementation in this build (pass --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibrate:
        copy_calibrate(args.copy_calibrate)
        return
    import numpy as np
    import torch
    from hammurabix_amd import Context, Param, _capi as capi, configs, los_samples

    par = Param(text=configs.random_forward())
    desc = par.los_desc(schedule=capi.HAMX_SCHED_AUTO)
    S_geo = los_samples(desc)
    k_sync = max(1, sum(1 for s in par.grid_obs.do_sync if s))
    S_ref = S_geo * k_sync
    dev = torch.device("cuda:0")
    g = par.grid_brnd
    n = g.full_size
    with Context(0) as ctx:
        bx, by, bz, te = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(4))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        ctx.brnd_es(par.es_desc(), 1, bx, by, bz)
        ctx.ternd_dft(par.dft_desc(), 1, te)
        gen_ms = (time.perf_counter() - t0) * 1e3
        npix = 12 * 256 * 256
        maps = {k: torch.empty(npix, dtype=torch.float64, device=dev)
                for k in ("dm", "fd", "is", "qs", "us")}
        kw = dict(brnd=(bx, by, bz), brnd_box=g.box(), ternd=te,
                  ternd_box=par.grid_ternd.box())
        for _ in range(args.warmup):
            ctx.los_forward(desc, maps, **kw)
        torch.cuda.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            ctx.los_forward(desc, maps, **kw)
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        times = ctx.kernel_times()
        ctx.profile(False)
        march = [t for t in times if t["kernel"].startswith("los_")]
        march_kernels = sorted({t["kernel"] for t in march})
        march_ms = sum(t["ms"] for t in march)
        march_samples = sum(t["work"] for t in march)
        acc_ms = sum(t["ms"] for t in times if t["kernel"] == "accumulate")
        finite = all(bool(torch.isfinite(m).all()) for m in maps.values())
        grids_host = None
        if not args.no_cpu:
            grids_host = tuple(np.ascontiguousarray(a.cpu().numpy()) for a in (bx, by, bz, te))

    ms_per_step = elapsed * 1e3 / args.steps
    value = S_ref * args.steps / elapsed
    achieved = march_samples * BYTES_PER_SAMPLE / (march_ms * 1e-3) / 1e9
    traffic = traffic_from_profiles()
    out = {
        "metric": "LOS samples/s and map+gradient evals/s at nside 256, vs plain GPU at "
                  "≤1e-5 rel",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "random_forward: LSA+unif TE+analytic CRE, ES random B and DFT "
                        "random TE on 256x256x64, shells nside 64/128/256, sync 23 GHz, "
                        "FP64 plain path",
            "samples_per_step": S_ref,
            "shells": [int(desc.nside_shell[i]) for i in range(desc.n_shell)],
            "grid": [g.nx, g.ny, g.nz],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
        },
        "detail": {
            "march_kernels": march_kernels,
            "march_ms_per_step": march_ms / args.steps,
            "accumulate_ms_per_step": acc_ms / args.steps,
            "random_field_generation_ms": gen_ms,
            "maps_finite": finite,
        },
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(configs.random_forward, grids_host,
                                           min(args.cpu_threads, os.cpu_count() or 1))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
