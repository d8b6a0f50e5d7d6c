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
# VALU issue peak of one MI355X: 256 CUs x 4 SIMDs, each issuing one wave64
# vector instruction per 4 cycles (16 lanes per cycle) at 2.4 GHz
VALU_PEAK_GINST = 256 * 4 * 2.4 / 4
ROOFLINE_FILE = os.path.join(ROOT, "profiles", "los_fast_random_forward_roofline.txt")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=16)
    ap.add_argument("--workload", choices=("random_forward", "batch", "large", "cre_grid"),
                    default="random_forward",
                    help="random_forward: the plain FP64 path on one realisation (default); "
                         "batch: 64 realisations x parameter sets x 1-4 frequencies in one "
                         "graph replay, against the plain path run pass by pass; large: 16 "
                         "realisations on 512x512x128 with 4 shells up to nside 1024 and 4 "
                         "frequencies, forward + adjoint in batched graphs")
    ap.add_argument("--instances", type=int, default=0,
                    help="batch size (batch: 64, large: 16 by default)")
    ap.add_argument("--check", type=int, default=0,
                    help="instances checked against the plain path (batch: all, large: 1)")
    ap.add_argument("--precision", choices=("f32", "f64"), default="f32",
                    help="batch sample arithmetic (batch workload)")
    ap.add_argument("--no-adjoint", action="store_true",
                    help="batch workload: forward only (no map+gradient evaluations)")
    ap.add_argument("--no-batch", action="store_true",
                    help="random_forward: skip the 64-instance batched map+gradient "
                         "measurement that the default run reports in its detail")
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


def roofline_inputs():
    """Per-pass counter figures of the march from the committed rocprofv3
    summary (key: value lines): VALU wave instructions per geometric sample,
    HBM bytes per write_grid pass (FETCH_SIZE + WRITE_SIZE, calibrated) and
    the L2 hit rate.  Missing keys are None."""
    keys = ("valu_wave_insts_per_sample", "hbm_bytes_per_pass", "l2_hit_rate")
    out = dict.fromkeys(keys)
    if os.path.exists(ROOFLINE_FILE):
        for line in open(ROOFLINE_FILE):
            k, _, v = line.partition(":")
            if k.strip() in out:
                out[k.strip()] = float(v)
    return out


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(par_text_fn, grids_host, threads):
    # threads pinned close to each other, as the reference's release runs
    os.environ.setdefault("OMP_PROC_BIND", "close")
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    from hammurabix_amd import Param, los_samples
    from oracle_lib import write_grid
    par = Param(text=par_text_fn(nside_sim=16, nside_out=64))
    d = par.los_desc()
    bx, by, bz, te = grids_host
    box = par.grid_brnd.box()
    S = los_samples(d)
    runs = []
    for _ in range(3):
        t0 = time.perf_counter()
        write_grid(d, nthreads=threads, brnd=(bx, by, bz), brnd_box=box, ternd=te,
                   ternd_box=par.grid_ternd.box())
        runs.append(time.perf_counter() - t0)
    dt = sorted(runs)[1]
    return dict(value=S / dt, unit="samples/s", cores=threads, kind="port",
                sample=f"random_forward with shells at nside 16/32/64 (1/16 of the "
                       f"pixels, {S} samples), same grids and radial sampling; median of "
                       f"3 runs {dt:.2f} s; {os.cpu_count()} host CPUs visible ({cpu_model()}), "
                       f"OMP_PROC_BIND={os.environ.get('OMP_PROC_BIND')}")


def plain_passes(d0, freqs):
    """The plain path's passes of one instance, in the reference pipeline's
    order: the last-listed frequency first with DM and FD, then the others
    sync only.  Yields (pass, frequency index, descriptor)."""
    import ctypes as C
    from hammurabix_amd import _capi as capi
    for p, fi in enumerate(reversed(range(len(freqs)))):
        d = capi.HamxLosDesc()
        C.memmove(C.byref(d), C.byref(d0), C.sizeof(d))
        d.sync_freq, d.do_sync = freqs[fi], 1
        if p > 0:
            d.do_dm = d.do_fd = 0
        yield p, fi, d


def pass_maps(d, dev):
    import torch
    z = lambda ns: torch.empty(12 * ns * ns, dtype=torch.float64, device=dev)  # noqa: E731
    m = {k: z(d.nside_sync) for k in ("is", "qs", "us")}
    m["fd"] = z(d.nside_fd if d.do_fd else d.nside_sync)
    if d.do_dm:
        m["dm"] = z(d.nside_dm)
    return m


def regenerate(ctx, par, sp, seed):
    """Draw the ES random B and DFT random TE of seed `seed` into the
    instance's grid tensors (the batch graph re-packs them on every replay)."""
    ctx.brnd_es(par.es_desc(), seed, *sp["brnd"])
    ctx.ternd_dft(par.dft_desc(), seed, sp["ternd"])


def likelihood_seeds(ctx, b, pars, specs, dev):
    """Fill every instance's sp["adj"] with dL/dmap of
    L = -1/2 sum (m - obs)^2 / sigma^2 over the Q/U maps of every frequency
    and the FD map: obs is the batch's forward on a held-out realisation
    (seed + 10000) plus Gaussian noise of sigma = 0.1 x rms(obs), m the
    batch's forward on the instance itself.  The I and DM seeds are zero.
    Both the batch and the plain adjoints are then given these same seeds."""
    import torch
    keys = lambda sp: [(k, f) for k in ("qs", "us") for f in range(len(sp["freqs"]))] + [("fd", None)]  # noqa: E731
    pick = lambda m, k, f: m[k][f] if f is not None else m[k]  # noqa: E731
    for par, sp in zip(pars, specs):
        regenerate(ctx, par, sp, par.brnd_seed + 10000)
    b.run()
    torch.cuda.synchronize()
    obs = [{(k, f): pick(sp["maps"], k, f).clone() for k, f in keys(sp)} for sp in specs]
    for par, sp in zip(pars, specs):
        regenerate(ctx, par, sp, par.brnd_seed)
    b.run()
    torch.cuda.synchronize()
    gen = torch.Generator(device=dev).manual_seed(20001)
    for sp, ob in zip(specs, obs):
        for t in sp["adj"]["is"]:
            t.zero_()
        sp["adj"]["dm"].zero_()
        for (k, f), o in ob.items():
            sigma = 0.1 * float(o.square().mean().sqrt())
            o += sigma * torch.randn(o.shape[0], dtype=torch.float64, device=dev, generator=gen)
            pick(sp["adj"], k, f).copy_((o - pick(sp["maps"], k, f)) / (sigma * sigma))


def progress(msg):
    """One line per finished phase on stderr, so long runs show they are alive."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def run_batch(args):
    """Print the bench line of the batch workloads (see measure_batch)."""
    large = args.workload == "large"
    n_inst = args.instances if args.instances else (16 if large else 64)
    prec = "f32" if args.precision == "f32" else "f64"
    print(json.dumps(measure_batch(large, n_inst, args.check, prec, not args.no_adjoint,
                                   args.steps, args.warmup)), flush=True)


def rms_scaled(pairs):
    """max_i |batch_i - plain_i| / rms_i(plain_i) per parameter, over the
    (batch, plain) gradient pairs of the checked instances."""
    import torch
    from hammurabix_amd import _capi as capi
    if not pairs:
        return {}
    got = torch.stack([p[0] for p in pairs])
    ref = torch.stack([p[1] for p in pairs])
    rms = ref.square().mean(0).sqrt().clamp_min(1e-300)
    worst = ((got - ref).abs().max(0).values / rms).tolist()
    return dict(zip(capi.PARAM_NAMES, worst))


def measure_batch(large, n_inst, check, precision, with_adj, steps, warmup):
    """The batch workloads; returns the bench line as a dict.

    batch: instance i has random fields of seed i + 1 on the 256x256x64 box,
    top nside (64, 128, 256)[i % 3] over 3 shells, the first 1 + i % 4 of
    (0.408, 1.4, 23, 30) GHz and a frozen (b0, psi0) draw; 64 instances.
    large: 16 realisations (seeds 1-16) on 512x512x128 grids, 4 shells at
    nside 128/256/512/1024 with 0.005 kpc steps, DM, FD and sync at all four
    frequencies on nside 1024.

    One step is one replay of the captured batch graph.  The plain path is an
    instance's passes run one by one through hamx_los_forward (FP64), the way
    the reference pipeline runs them; it gives the maps (and, with the
    adjoint, the gradients) the batch is checked against, for every instance
    (batch) or the first --check instances (large).

    Map+gradient evaluations differentiate a Gaussian log-likelihood of the
    sync Q/U maps of every frequency plus the FD map against 'observed' maps:
    the same instance on a held-out realisation with 10% Gaussian noise (see
    likelihood_seeds)."""
    import ctypes as C
    import torch
    from hammurabix_amd import Context, Param, _capi as capi, configs, los_samples
    from hammurabix_amd.autograd import desc_params
    dev = torch.device("cuda:0")
    prec = capi.HAMX_PREC_F32 if precision == "f32" else capi.HAMX_PREC_F64
    n_check = min(n_inst, check if check else (1 if large else n_inst))
    make = configs.large_instance if large else configs.batch_instance
    specs, pars = [], []
    with Context(0) as ctx:
        t0 = time.perf_counter()
        for i in range(n_inst):
            par = Param(text=make(i))
            pars.append(par)
            n = par.grid_brnd.full_size
            bx, by, bz, te = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(4))
            ctx.brnd_es(par.es_desc(), par.brnd_seed, bx, by, bz)
            ctx.ternd_dft(par.dft_desc(), par.brnd_seed, te)
            d = par.los_desc()
            g = par.grid_obs
            z = lambda ns: torch.zeros(12 * ns * ns, dtype=torch.float64, device=dev)  # noqa
            freqs = list(g.sim_sync_freq)
            maps = dict(dm=z(d.nside_dm), fd=z(d.nside_fd),
                        **{k: [z(d.nside_sync) for _ in freqs] for k in ("is", "qs", "us")})
            spec = dict(desc=d, freqs=freqs, maps=maps, brnd=(bx, by, bz), ternd=te,
                        box=par.grid_brnd.box())
            if with_adj:
                spec["adj"] = {"dm": z(d.nside_dm), "fd": z(d.nside_fd),
                               **{k: [z(d.nside_sync) for _ in freqs]
                                  for k in ("is", "qs", "us")}}
                spec["grads"] = {
                    "dparams": torch.zeros(capi.HAMX_NPARAM, dtype=torch.float64, device=dev),
                    "dbrnd": tuple(torch.zeros(n, dtype=torch.float64, device=dev)
                                   for _ in range(3)),
                    "dternd": torch.zeros(n, dtype=torch.float64, device=dev)}
            specs.append(spec)
        torch.cuda.synchronize()
        gen_s = time.perf_counter() - t0
        progress(f"{n_inst} realisations generated in {gen_s:.1f} s")
        t0 = time.perf_counter()
        b = ctx.batch(specs, precision=prec, tune=True, adjoint=with_adj)
        plan_s = time.perf_counter() - t0
        progress(f"batch planned in {plan_s:.1f} s")
        if with_adj:
            likelihood_seeds(ctx, b, pars, specs, dev)
        info = b.info()
        for _ in range(warmup):
            b.run()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            b.run()
        torch.cuda.synchronize()
        elapsed = time.perf_counter() - t0
        progress(f"forward replays timed: {elapsed * 1e3 / steps:.1f} ms each")
        mg = {}
        if with_adj:
            # map+gradient evaluations: forward replay + reverse replay
            b.run()
            b.run_adjoint()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(steps):
                b.run()
                b.run_adjoint()
            torch.cuda.synchronize()
            mg_s = (time.perf_counter() - t0) / steps
            progress(f"forward + reverse replays timed: {mg_s * 1e3:.1f} ms each")
            mg = dict(batch_map_gradient_ms=mg_s * 1e3,
                      batch_map_gradient_evals_per_s=n_inst / mg_s)
        # the plain path, pass by pass, and the maps it gives
        worst, plain_s, plain_samples = 0.0, 0.0, 0
        plain_mg_s, worst_dp, worst_grid = 0.0, 0.0, 0.0
        if with_adj:
            n = specs[0]["brnd"][0].shape[0]
            pdp = torch.zeros(capi.HAMX_NPARAM, dtype=torch.float64, device=dev)
            pgb = tuple(torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3))
            pgt = torch.zeros(n, dtype=torch.float64, device=dev)
        map_err = {}  # worst rel. L2 per map kind and frequency
        worst_dln = 0.0
        dp_err = dict.fromkeys(capi.PARAM_NAMES, 0.0)
        dp_pairs = []  # (batch, plain FP64) parameter gradients per checked instance
        for ci, sp in enumerate(specs[:n_check]):
            progress(f"plain-path comparison, instance {ci + 1} of {n_check}")
            freqs = sp["freqs"]
            kw = dict(brnd=sp["brnd"], brnd_box=sp["box"], ternd=sp["ternd"],
                      ternd_box=sp["box"])
            if with_adj:
                acc_dp = torch.zeros_like(pdp)
                acc_gb = [torch.zeros_like(pgt) for _ in range(3)]
                acc_gt = torch.zeros_like(pgt)
            for p, fi, d in plain_passes(sp["desc"], freqs):
                m = pass_maps(d, dev)
                torch.cuda.synchronize()
                t1 = time.perf_counter()
                ctx.los_forward(d, m, keep_tape=with_adj, **kw)
                plain_s += time.perf_counter() - t1
                plain_samples += los_samples(d)
                if with_adj:
                    a = {k: sp["adj"][k][fi] for k in ("is", "qs", "us")}
                    if p == 0:
                        a["dm"], a["fd"] = sp["adj"]["dm"], sp["adj"]["fd"]
                    t1 = time.perf_counter()
                    ctx.los_adjoint(d, a, pdp, dbrnd=pgb, dternd=pgt, use_tape=True, **kw)
                    torch.cuda.synchronize()
                    plain_mg_s += time.perf_counter() - t1
                    acc_dp += pdp
                    for t, g in zip(acc_gb, pgb):
                        t += g
                    acc_gt += pgt
                pairs = [(f"{k}@{freqs[fi]:g}GHz", sp["maps"][k][fi], m[k])
                         for k in ("is", "qs", "us")]
                if p == 0:
                    pairs += [("dm", sp["maps"]["dm"], m["dm"]),
                              ("fd", sp["maps"]["fd"], m["fd"])]
                for name, got, ref in pairs:
                    nr = float(torch.linalg.vector_norm(ref))
                    err = float(torch.linalg.vector_norm(got - ref)) / nr if nr else 0.0
                    worst = max(worst, err)
                    map_err[name] = max(map_err.get(name, 0.0), err)
            if with_adj:
                g = sp["grads"]
                dp_pairs.append((g["dparams"].clone(), acc_dp.clone()))
                rel = (g["dparams"] - acc_dp).abs() / acc_dp.abs().clamp_min(1e-300)
                worst_dp = max(worst_dp, float(rel.max()))
                for name, e in zip(capi.PARAM_NAMES, rel.tolist()):
                    dp_err[name] = max(dp_err[name], e)
                # the whole parameter gradient in log-parameter units (p dL/dp)
                pv = desc_params(sp["desc"]).to(dev)
                ln_err = float(torch.linalg.vector_norm(pv * (g["dparams"] - acc_dp)) /
                               torch.linalg.vector_norm(pv * acc_dp).clamp_min(1e-300))
                worst_dln = max(worst_dln, ln_err)
                for got, ref in zip(list(g["dbrnd"]) + [g["dternd"]], acc_gb + [acc_gt]):
                    nr = float(torch.linalg.vector_norm(ref))
                    err = float(torch.linalg.vector_norm(got - ref)) / nr if nr else 0.0
                    worst_grid = max(worst_grid, err)
        b.close()
    ms = elapsed * 1e3 / steps
    value = info["samples_ref"] * steps / elapsed
    plain_rate = plain_samples / plain_s
    grid_bytes = 128 if prec == capi.HAMX_PREC_F32 else 256
    achieved = info["samples_geo"] * grid_bytes / (ms * 1e-3) / 1e9
    out = {
        "metric": "LOS samples/s and map+gradient evals/s at nside 256, vs plain GPU at "
                  "≤1e-5 rel",
        "value": value,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == capi.HAMX_PREC_F32 else "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"large: {n_inst} ES/DFT realisations (seeds 1-{n_inst}) on "
                         "512x512x128, 4 shells at nside 128/256/512/1024 with 0.005 kpc steps, "
                         "DM, FD and sync at 0.408/1.4/23/30 GHz on nside 1024; one step = one "
                         f"batch graph replay; {n_check} instance(s) checked against the plain path"
                         if large else
                         f"batch: {n_inst} ES/DFT realisations (seeds 1-{n_inst}) on "
                         "256x256x64, top nside (64,128,256)[i%3] over 3 shells, first "
                         "1+i%4 of (0.408,1.4,23,30) GHz, frozen b0/psi0 draws; one step = "
                         "one batch graph replay"),
            "instances": n_inst,
            "samples_ref_per_step": info["samples_ref"],
            "samples_geo_per_step": info["samples_geo"],
            "launch_shape": info["shape"],
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": None,
        },
        "detail": {
            "plain_path_samples_per_s": plain_rate,
            "speedup_vs_plain_pass_by_pass": value / plain_rate,
            "max_rel_l2_vs_plain_fp64": worst,
            "max_rel_l2_by_map": map_err,
            "autotune_ms": info["variant_ms"],
            **mg,
            **({"plain_map_gradient_evals_per_s": n_check / (plain_s + plain_mg_s),
                "map_gradient_speedup_vs_plain":
                    (mg["batch_map_gradient_evals_per_s"] * (plain_s + plain_mg_s) / n_check),
                "max_rel_err_dparams_vs_plain_fp64": worst_dp,
                "max_rel_err_by_param": dp_err,
                # each parameter's worst absolute error over the RMS of its
                # gradient across the checked instances: the error at the
                # scale an optimiser or sampler sees, where the relative error
                # of an instance whose gradient happens to be near zero is not
                "max_err_by_param_over_rms_gradient": rms_scaled(dp_pairs),
                "max_rel_l2_dlnparams_vs_plain_fp64": worst_dln,
                "max_rel_l2_grid_gradient_vs_plain_fp64": worst_grid} if with_adj else {}),
            "plan_s": plan_s,
            "realisation_generation_s": gen_s,
        },
    }
    return out


CRE_GRID = dict(shape=(80, 80, 80), nE=80, E_GeV=(0.1, 1000.0))


def run_cre_grid(args):
    """Workload "cre_grid": the random_forward pass with the CRE taken from an
    imported flux grid (CRE_num, integrator.cc:281-401) instead of the analytic
    model: the analytic model's flux written on the device on an 80x80x80 box
    over [-20,20]x[-20,20]x[-4,4] kpc with 80 energies from 0.1 to 1000 GeV
    (the template's box_cre sizes).  Every sample integrates F and G over the
    79 energy bins, so the march is FP64 arithmetic bound, not HBM bound."""
    import math

    import numpy as np
    import torch
    from hammurabix_amd import Context, Param, _capi as capi, configs, los_samples
    from hammurabix_amd.units import CGS

    par = Param(text=configs.random_forward())
    desc = par.los_desc(schedule=capi.HAMX_SCHED_AUTO)
    cre_desc = par.cre_desc()
    desc.cre_model = capi.HAMX_CRE_GRID
    S = los_samples(desc)
    nE = CRE_GRID["nE"]
    E_min, E_max = (e * CGS.GeV for e in CRE_GRID["E_GeV"])
    E_fact = math.log(E_max / E_min) / (nE - 1)
    cbox = capi.HamxBox(*CRE_GRID["shape"], -20 * CGS.kpc, 20 * CGS.kpc, -20 * CGS.kpc,
                        20 * CGS.kpc, -4 * CGS.kpc, 4 * CGS.kpc)
    dev = torch.device("cuda:0")
    g = par.grid_brnd
    n = g.full_size
    with Context(0) as ctx:
        bx, by, bz, te = (torch.empty(n, dtype=torch.float64, device=dev) for _ in range(4))
        ctx.brnd_es(par.es_desc(), 1, bx, by, bz)
        ctx.ternd_dft(par.dft_desc(), 1, te)
        cre = torch.empty(cbox.nx * cbox.ny * cbox.nz * nE, dtype=torch.float64, device=dev)
        ctx.cre_write_grid(cre_desc, cbox, E_min, E_fact, nE, cre)
        npix = 12 * 256 * 256
        maps = {k: torch.empty(npix, dtype=torch.float64, device=dev)
                for k in ("dm", "fd", "is", "qs", "us")}
        kw = dict(brnd=(bx, by, bz), brnd_box=g.box(), ternd=te,
                  ternd_box=par.grid_ternd.box(), cre=cre, cre_box=cbox,
                  cre_E=(E_min, E_fact, nE))
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
        march_ms = sum(t["ms"] for t in march)
        finite = all(bool(torch.isfinite(m).all()) for m in maps.values())
        host = None
        if not args.no_cpu:
            host = tuple(np.ascontiguousarray(a.cpu().numpy()) for a in (bx, by, bz, te, cre))
    out = {
        "metric": "LOS samples/s with an imported CRE grid (CRE_num) at nside 256",
        "value": S * args.steps / elapsed,
        "unit": "samples/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "cre_grid: random_forward (LSA, unif TE, ES random B and DFT random "
                        "TE on 256x256x64, shells nside 64/128/256, sync 23 GHz) with the CRE "
                        "from an 80x80x80 grid of 80 energies (0.1-1000 GeV)",
            "samples_per_step": S,
            "energy_bins_per_sample": nE - 1,
        },
        "roofline": None,
        "detail": {
            "march_kernels": sorted({t["kernel"] for t in march}),
            "march_ms_per_step": march_ms / args.steps,
            "energy_bin_evaluations_per_s": S * (nE - 1) * args.steps / (march_ms * 1e-3),
            "maps_finite": finite,
            "bound": "FP64 VALU: two logarithms, two exponentials and two Chebyshev sums "
                     "per energy bin and sample",
        },
    }
    if host is not None:
        sys.path.insert(0, os.path.join(ROOT, "tests"))
        from oracle_lib import write_grid
        threads = min(args.cpu_threads, os.cpu_count() or 1)
        p2 = Param(text=configs.random_forward(nside_sim=2, nside_out=8))
        d2 = p2.los_desc()
        d2.cre_model = capi.HAMX_CRE_GRID
        S2 = los_samples(d2)
        t0 = time.perf_counter()
        write_grid(d2, nthreads=threads, brnd=host[:3], brnd_box=g.box(), ternd=host[3],
                   ternd_box=par.grid_ternd.box(), cre=host[4], cre_box=cbox,
                   cre_E=(E_min, E_fact, nE))
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = dict(
            value=S2 / dt, unit="samples/s", cores=threads, kind="port",
            sample=f"the same pass with shells at nside 2/4/8 (1/1024 of the pixels, "
                   f"{S2} samples), same grids; {dt:.2f} s")
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py runs on one GPU only: the LOS path has no multi-GPU "
              "implementation in this build (pass --gpus 1)", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibrate:
        copy_calibrate(args.copy_calibrate)
        return
    if args.workload in ("batch", "large"):
        run_batch(args)
        return
    if args.workload == "cre_grid":
        run_cre_grid(args)
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
        # the ES generator alone, warm: algorithmic 800 B per vertex (draw 32,
        # 3 x 2 FFTs of 3 axis passes 576, variance 8, rescale 64, Gram-Schmidt
        # with the reassembly in place 64, Hermitian extraction 56)
        t0 = time.perf_counter()
        for _ in range(3):
            ctx.brnd_es(par.es_desc(), 1, bx, by, bz)
        es_ms = (time.perf_counter() - t0) * 1e3 / 3
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
        # map+gradient evaluations: forward keeping its tape, then the adjoint
        # of L = sum(w * maps) into the six parameters and the full B / TE grids
        gen = torch.Generator(device=dev).manual_seed(7)
        adj = {k: torch.randn(v.shape[0], dtype=torch.float64, device=dev, generator=gen)
               for k, v in maps.items()}
        dparams = torch.zeros(capi.HAMX_NPARAM, dtype=torch.float64, device=dev)
        dgrid = tuple(torch.zeros(n, dtype=torch.float64, device=dev) for _ in range(3))
        dte = torch.zeros(n, dtype=torch.float64, device=dev)
        ctx.los_forward(desc, maps, keep_tape=True, **kw)
        ctx.los_adjoint(desc, adj, dparams, dbrnd=dgrid, dternd=dte, use_tape=True, **kw)
        torch.cuda.synchronize()
        ctx.profile(True)
        t0 = time.perf_counter()
        n_mg = max(1, args.steps // 2)
        for _ in range(n_mg):
            ctx.los_forward(desc, maps, keep_tape=True, **kw)
            ctx.los_adjoint(desc, adj, dparams, dbrnd=dgrid, dternd=dte, use_tape=True, **kw)
        torch.cuda.synchronize()
        mg_s = (time.perf_counter() - t0) / n_mg
        adj_times = ctx.kernel_times()
        ctx.profile(False)
        adj_ms = sum(t["ms"] for t in adj_times if t["kernel"] == "los_adjoint") / n_mg
        grad_finite = bool(torch.isfinite(dgrid[0]).all()) and bool(torch.isfinite(dparams).all())
        grids_host = None
        if not args.no_cpu:
            grids_host = tuple(np.ascontiguousarray(a.cpu().numpy()) for a in (bx, by, bz, te))

    ms_per_step = elapsed * 1e3 / args.steps
    value = S_ref * args.steps / elapsed
    # the roofline, per write_grid pass (3 k_los_fast launches, one per shell)
    march_s_pass = march_ms * 1e-3 / args.steps
    samples_pass = march_samples / args.steps
    rf = roofline_inputs()
    roofline = None
    if rf["valu_wave_insts_per_sample"] is not None:
        achieved = rf["valu_wave_insts_per_sample"] * samples_pass / march_s_pass / 1e9
        hbm = rf["hbm_bytes_per_pass"]
        roofline = {
            "bound": "valu",
            "achieved": achieved,
            "peak": VALU_PEAK_GINST,
            "unit": "Ginst/s",
            "frac": achieved / VALU_PEAK_GINST,
            "traffic": hbm,
            "basis": "one write_grid pass = 3 k_los_fast launches; wave64 VALU instructions "
                     "per pass (SQ_INSTS_VALU per geometric sample x samples) / live march "
                     "time; traffic = FETCH_SIZE + WRITE_SIZE bytes per pass",
            "hbm_GBps": hbm / march_s_pass / 1e9 if hbm else None,
            "hbm_frac": hbm / march_s_pass / 1e9 / HBM_PEAK_GBS if hbm else None,
            "l2_hit_rate": rf["l2_hit_rate"],
            "algorithmic_GBps": samples_pass * BYTES_PER_SAMPLE / march_s_pass / 1e9,
        }
    batch = None
    if not args.no_batch:
        progress("batch map+gradient (64 instances)")
        batch = measure_batch(False, 64, 0, "f32", True, max(1, min(args.steps, 5)), 1)
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
        "roofline": roofline,
        "detail": {
            "march_kernels": march_kernels,
            "march_ms_per_step": march_ms / args.steps,
            "accumulate_ms_per_step": acc_ms / args.steps,
            "random_field_generation_ms": gen_ms,
            "es_generator_ms": es_ms,
            "es_generator_algorithmic_GBps": 800 * n / (es_ms * 1e-3) / 1e9,
            "maps_finite": finite,
            "map_gradient_evals_per_s": 1.0 / mg_s,
            "map_gradient_ms": mg_s * 1e3,
            "adjoint_march_ms": adj_ms,
            "gradients_finite": grad_finite,
        },
    }
    if batch is not None:
        bd = batch["detail"]
        out["detail"]["batch"] = {
            "workload": batch["config"]["workload"],
            "precision": batch["dtype"],
            "forward_ms": batch["ms_per_step"],
            "forward_samples_per_s": batch["value"],
            **{k: bd[k] for k in (
                "batch_map_gradient_ms", "batch_map_gradient_evals_per_s",
                "plain_map_gradient_evals_per_s", "map_gradient_speedup_vs_plain",
                "speedup_vs_plain_pass_by_pass", "max_rel_l2_vs_plain_fp64",
                "max_rel_err_by_param", "max_err_by_param_over_rms_gradient",
                "max_rel_err_dparams_vs_plain_fp64", "max_rel_l2_dlnparams_vs_plain_fp64",
                "max_rel_l2_grid_gradient_vs_plain_fp64") if k in bd},
            "launch_shape": batch["config"]["launch_shape"],
        }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(configs.random_forward, grids_host,
                                           min(args.cpu_threads, os.cpu_count() or 1))
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
