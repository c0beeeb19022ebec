#!/usr/bin/env python3
"""Benchmark: Nside-4096 T+P full MASTER wall time on one MI355X.

One step = the whole pseudo-C_ell pipeline on inputs already resident in HBM:
``NmtField(mask, [T], n_iter=3)``, ``NmtField(mask, [Q, U], n_iter=3)``,
``NmtBin.from_nside_linear(nside, 16)``, the 00 / 02 / 22 workspaces and three
``decouple_cell(compute_coupled_cell(...))`` calls (the reference's timed region,
``benchmarks/scaling/bench_act_warm.py:83-112``).  Every cache (recurrence tables,
coupling blocks) is dropped before each step, so each step recomputes everything.

Workload (synthetic, seeded): mask = 1 except 20 random 5-degree disc holes, C1-apodized
2 degrees; T, Q, U Gaussian with C_TT = 400/(l+10)^2, C_EE = 0.1 C_TT, C_TE = 0.3
sqrt(TT EE), C_BB = 0.01 C_EE; lmax = 3 nside - 1; bins of 16.

Two forms are timed in the same run: plain FP64 and the chosen form (``--form``, default
``fast``): T, the mask and Q/U of the realisation through one shared FP64 recurrence with
the contraction on the matrix unit (FP16 hi/lo products in the analyses, FP32 in the
syntheses, FP64 near the poles), the Jacobi refinement on corrections, and the coupling
from the Toeplitz approximation with a wide exact region checked by exact spot checks.
``value`` is the chosen form's time; ``fp64_s`` the plain form's; ``bandpower_err_vs_fp64``
and ``alm_err_vs_fp64`` the chosen form's largest relative errors.

Prints one JSON line (value = seconds per pipeline, lower is better) with a roofline
entry for the chosen form's dominant kernel (the batched analysis) and per-stage entries,
measured with HIP events on the launching stream during the timed steps, and a CPU
baseline (the project's C/OpenMP FP64 restatement of the same Legendre and 3j work, timed
on a bounded sample of orders / rows at the same Nside and extrapolated to the whole
pipeline, beside unsampled runs at Nside 64 / 256 / 1024).

Other modes: ``--sim-batch N`` (config 3), ``--coupling-forms`` (config 4), ``--sims N``,
``--sht-batch K``.
"""

import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12  # MI355X dense FP64 vector / matrix, FLOP/s
PEAK_FP32 = 157.3e12  # MI355X FP32 vector (unpacked) / matrix, FLOP/s


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def legendre_flops(nside, lmax):
    """Algorithmic FLOP of one Legendre pass, full triangle, no polar skip.

    spin-0 pair (field + mask, folded rings): 2 (2 + 2k) T0, k = 2, T0 = L(L+1)/2 * 2 nside
    spin 2 (direct + mirror channels, all rings): 2 (2 + 4) T2, T2 = L(L+1)/2 * (4 nside - 1)
    """
    L = lmax + 1
    tri = L * (L + 1) / 2
    T0 = tri * 2 * nside
    T2 = tri * (4 * nside - 1)
    return 2 * (2 + 4) * T0, 2 * (2 + 4) * T2


def _n3j(lmax):
    """l3 terms of the exact coupling: sum over l1 <= l2 of the triangle-rule count."""
    import numpy as np

    L = lmax + 1
    l1 = np.arange(L, dtype=np.int64)[:, None]
    l2 = np.arange(L, dtype=np.int64)[None, :]
    cnt = np.minimum(l1 + l2, lmax) - (l2 - l1) + 1
    return int(np.sum(np.where(l2 >= l1, np.maximum(cnt, 0), 0)))


def _measured_traffic(nside, name):
    """HBM bytes per launch of kernel ``name`` from the committed rocprofv3 FETCH_SIZE +
    WRITE_SIZE summary at this Nside (copy-kernel scaled), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"{name}_traffic_nside{nside}.txt")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return float(line.split(":", 1)[1])
    return None


def run_pipeline(nmt, mask, tqu, nside, nlb, form="fp64", info=None, keep=None):
    from gmaster_amd import batched, field, sht, workspaces

    nmt.set_form(form)
    sht.clear_plans()
    batched.clear_plans()
    workspaces.clear_coupling_cache()
    field.clear_mask_cache()
    fT = nmt.NmtField(mask, tqu[0:1], n_iter=3)
    fP = nmt.NmtField(mask, tqu[1:3], n_iter=3)
    b = nmt.NmtBin.from_nside_linear(nside, nlb)
    out = []
    for f1, f2 in ((fT, fT), (fT, fP), (fP, fP)):
        w = nmt.NmtWorkspace.from_fields(f1, f2, b)
        out.append(w.decouple_cell(nmt.compute_coupled_cell(f1, f2)))
        if info is not None:
            info.append(w.coupling_info)
    nmt.set_form("fp64")
    if keep is not None:
        keep["alms"] = {"T": fT.alm[0], "E": fP.alm[0], "B": fP.alm[1], "mask": fT.alm_mask[0]}
    return out


def alm_errors(got, ref):
    """Largest |a - a_fp64| relative to the largest |a_fp64|, per field component."""
    import torch

    return {k: float(torch.max(torch.abs(got[k] - ref[k])) / torch.max(torch.abs(ref[k]))) for k in ref}


def bandpower_errors(bp, ref):
    """Relative error per spectrum of decoupled bandpowers against the FP64 form: auto and TE
    spectra relative to themselves, the null spectra (TB, EB, BE) relative to
    sqrt(XX YY) of the band."""
    import torch

    tt, te, tb = ref[0][0], ref[1][0], ref[1][1]
    ee, eb, be, bb = ref[2]
    out = {}

    def rel(a, r, norm=None):
        norm = torch.abs(r) if norm is None else norm
        return float(torch.max(torch.abs(a - r) / norm))

    out["TT"] = rel(bp[0][0], tt)
    out["TE"] = rel(bp[1][0], te)
    out["EE"] = rel(bp[2][0], ee)
    out["BB"] = rel(bp[2][3], bb)
    out["TB"] = rel(bp[1][1], tb, torch.sqrt(torch.abs(tt * bb)))
    out["EB"] = rel(bp[2][1], eb, torch.sqrt(torch.abs(ee * bb)))
    out["BE"] = rel(bp[2][2], be, torch.sqrt(torch.abs(ee * bb)))
    out["max"] = max(out.values())
    return out


def sht_batch(args):
    """Per-map throughput of spin-0 map2alm (n_iter 3) + alm2map: ``--sht-batch K`` maps
    sharing one recurrence (mixed precision) against the same maps one by one in FP64."""
    import torch

    from gmaster_amd import sht, workload

    nside = args.nside
    lmax = 3 * nside - 1
    K = args.sht_batch
    dev = torch.device("cuda", 0)
    mask = workload.hole_mask(nside, device=dev)
    maps = torch.stack([workload.tqu_maps(nside, lmax, seed=s)[0] * mask for s in range(K)])
    torch.cuda.synchronize()

    def run(batched):
        if batched:
            a = sht.map2alm_s0(maps, lmax, n_iter=3, precision="mixed", max_group=args.max_group)
            return a, sht.alm2map_s0(a, nside, lmax, precision="mixed", max_group=args.max_group)
        a = torch.cat([sht.map2alm_s0(maps[i:i + 1], lmax, n_iter=3) for i in range(K)])
        return a, torch.cat([sht.alm2map_s0(a[i:i + 1], nside, lmax) for i in range(K)])

    out = {}
    for batched in (False, True):
        for _ in range(args.warmup):
            run(batched)
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(args.steps):
            res = run(batched)
        torch.cuda.synchronize()
        out[batched] = ((time.perf_counter() - t) / args.steps, res)
        log(f"{'batched mixed' if batched else 'single FP64'}: {out[batched][0]:.3f}s for {K} maps")
    a_s, a_b = out[False][1][0], out[True][1][0]
    err = float(torch.max(torch.abs(a_b - a_s)) / torch.max(torch.abs(a_s)))
    print(json.dumps({
        "metric": f"spin-0 map2alm(n_iter 3) + alm2map maps/s, HEALPix Nside {nside}",
        "value": K / out[True][0], "unit": "maps/s", "higher_is_better": True, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "dtype": "f64 recurrence, f32 contraction",
        "data": "synthetic", "config": {"workload": f"{K} Gaussian maps x hole mask, Nside {nside}, lmax {lmax}",
                                         "batch": K, "groups": [k for _, k in sht._groups(K, True, args.max_group)]},
        "single_fp64_maps_per_s": K / out[False][0], "speedup": out[False][0] / out[True][0],
        "alm_rel_err_vs_fp64": err,
    }))


def sims(args):
    """Simulations sharing one workspace trio (00, 02, 22) per Nside: per-simulation T+P
    fields, coupled spectra and decoupling, timed in plain FP64 and in the chosen form; the
    workspaces are built once per form outside the timed region."""
    import torch

    import gmaster_amd as nmt
    from gmaster_amd import field, sht, workload, workspaces

    nside, n = args.nside, args.sims
    dev = torch.device("cuda", 0)
    mask = workload.hole_mask(nside, device=dev)
    maps = [workload.tqu_maps(nside, seed=1000 + s) for s in range(n)]
    b = nmt.NmtBin.from_nside_linear(nside, args.nlb)

    def fields(tqu):
        return nmt.NmtField(mask, tqu[0:1], n_iter=3), nmt.NmtField(mask, tqu[1:3], n_iter=3)

    def one(ws, tqu):
        fT, fP = fields(tqu)
        return [w.decouple_cell(nmt.compute_coupled_cell(f1, f2)).as_subclass(torch.Tensor)
                for w, (f1, f2) in zip(ws, ((fT, fT), (fT, fP), (fP, fP)))]

    out = {}
    for form in ("fp64", "fast"):
        nmt.set_form(form)
        sht.clear_plans()
        workspaces.clear_coupling_cache()
        field.clear_mask_cache()
        fT, fP = fields(maps[0])
        ws = [nmt.NmtWorkspace.from_fields(f1, f2, b) for f1, f2 in ((fT, fT), (fT, fP), (fP, fP))]
        one(ws, maps[0])
        torch.cuda.synchronize()
        t = time.perf_counter()
        res = [one(ws, tqu) for tqu in maps]
        torch.cuda.synchronize()
        out[form] = ((time.perf_counter() - t) / n, res)
        log(f"{form}: {out[form][0]:.3f} s per simulation over {n}")
    nmt.set_form("fp64")
    worst = max(max(bandpower_errors(g, r).values()) for g, r in zip(out["fast"][1], out["fp64"][1]))
    print(json.dumps({
        "metric": f"T+P MASTER simulations/s sharing one workspace trio, HEALPix Nside {nside}",
        "value": 1.0 / out["fast"][0], "unit": "sims/s", "higher_is_better": True, "n_gpus": 1,
        "steps": n, "warmup": 1, "dtype": "f64 recurrence, f32 contraction", "data": "synthetic",
        "config": {"workload": f"{n} Gaussian T+P skies x 20-hole mask, Nside {nside}, lmax {3 * nside - 1}, "
                               f"bins of {args.nlb}"},
        "fp64_sims_per_s": 1.0 / out["fp64"][0], "speedup_vs_fp64": out["fp64"][0] / out["fast"][0],
        "max_bandpower_err_vs_fp64": worst,
    }))


def coupling_forms(args):
    """Config 4: the mode coupling at the benchmark Nside by exact 3j, by the chosen form's
    wide-band Toeplitz (with its spot check and fallback, and without), by the committed
    surrogate's narrow Toeplitz alone and with its learned residual.  Per mask (the benchmark mask and ``--fresh-masks`` random ones): the
    time of the three workspaces' coupling, the relative Frobenius error of each block
    against the exact one, and the bandpower errors of T+P fields decoupled with each
    coupling against the exact decoupling of the same coupled spectra."""
    import numpy as np
    import torch

    import gmaster_amd as nmt
    from gmaster_amd import learned, workload, workspaces

    nside, nlb = args.nside, args.nlb
    lmax = 3 * nside - 1
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    sur = learned.committed_surrogate(dev)
    narrow = sur.grid.toeplitz()
    rng = np.random.default_rng(2024)
    tqu = workload.tqu_maps(nside, lmax, seed=4321, device=dev)
    b = nmt.NmtBin.from_nside_linear(nside, nlb)
    masks = [("benchmark: 20 holes of 5 deg, C1 2 deg", workload.hole_mask(nside, nholes=20, radius_deg=5.0,
                                                                          apo_deg=2.0, kind="C1", seed=1234,
                                                                          device=dev))]
    for i in range(args.fresh_masks):
        masks.append((f"fresh random mask {i}", learned.random_mask(nside, rng, dev)))
    wide = workspaces.nmt_params.fast_toeplitz
    variants = [("exact", "fp64", {}), ("toeplitz_wide", "fast", {}),
                ("toeplitz_wide_unchecked", "fp64", {"l_toeplitz": wide[0], "l_exact": wide[1], "dl_band": wide[2]}),
                ("toeplitz_narrow", "fp64", {"l_toeplitz": narrow[0], "l_exact": narrow[1], "dl_band": narrow[2]}),
                ("learned", "learned", {})]
    out = []
    for name, mask in masks:
        nmt.set_form("fast")
        fT = nmt.NmtField(mask, tqu[0:1], n_iter=3)
        fP = nmt.NmtField(mask, tqu[1:3], n_iter=3)
        cls = [nmt.compute_coupled_cell(f1, f2) for f1, f2 in ((fT, fT), (fT, fP), (fP, fP))]
        rec = {"mask": name, "fsky": float(mask.mean())}
        ref_bp = ref_blocks = None
        for vname, form, kw in variants:
            nmt.set_form(form)
            workspaces.clear_coupling_cache()
            torch.cuda.synchronize()
            t = time.perf_counter()
            ws = [nmt.NmtWorkspace.from_fields(f1, f2, b, **kw) for f1, f2 in ((fT, fT), (fT, fP), (fP, fP))]
            torch.cuda.synchronize()
            dt = time.perf_counter() - t
            bp = [w.decouple_cell(c) for w, c in zip(ws, cls)]
            blocks = {k: ws[2].blocks[k] for k in ("00", "02", "22p", "22m")}
            entry = {"coupling_s": dt, "info": ws[2].coupling_info}
            if vname == "exact":
                ref_bp, ref_blocks = bp, {k: v.clone() for k, v in blocks.items()}
            else:
                entry["block_frob_err"] = {k: float(torch.linalg.norm(blocks[k] - ref_blocks[k]) /
                                                    torch.linalg.norm(ref_blocks[k])) for k in ref_blocks}
                entry["bandpower_err"] = bandpower_errors(bp, ref_bp)
                if vname == "learned":
                    # the correction is zero above the residual grid (l > l_top): errors of the
                    # learned 22 block by row range, against the narrow Toeplitz alone
                    top = int(sur.grid.l_grid[-1])
                    col = 2 * torch.arange(lmax + 1, dtype=torch.float64, device=dev) + 1
                    for part, sl in (("l<=" + str(top), slice(0, top + 1)), ("l>" + str(top), slice(top + 1, None))):
                        d = (blocks["22p"] - ref_blocks["22p"])[sl] / col
                        r = ref_blocks["22p"][sl] / col
                        entry.setdefault("rows", {})[part] = float(torch.linalg.norm(d) / torch.linalg.norm(r))
            rec[vname] = entry
            log(f"{name}: {vname}: coupling {dt:.3f}s " + (f"bandpower max err {entry['bandpower_err']['max']:.2e}"
                                                           if vname != "exact" else ""))
            del ws, bp
        nmt.set_form("fp64")
        workspaces.clear_coupling_cache()
        out.append(rec)
        del fT, fP, cls
    bench_mask = out[0]
    print(json.dumps({
        "metric": f"Mode coupling at Nside {nside} (lmax {lmax}): exact 3j vs Toeplitz vs learned residual",
        "value": bench_mask["learned"]["coupling_s"], "unit": "s", "higher_is_better": False, "n_gpus": 1,
        "dtype": "f64 coupling, bf16-weight / fp32-accumulate network", "data": "synthetic",
        "config": {"workload": f"00/02/22 workspaces of T+P fields at Nside {nside}, bins of {nlb}",
                   "narrow_toeplitz": list(narrow), "wide_toeplitz": list(workspaces.nmt_params.fast_toeplitz),
                   "surrogate_grid_l_top": int(sur.grid.l_grid[-1])},
        "masks": out,
    }))


def sim_batch(args):
    """Config 3: N T+P realisations of mixed Nside (split evenly over ``--sim-nsides``),
    map2alm (n_iter 3) + alm2map of every map through the batched transforms, the whole job
    captured as one HIP graph and replayed; a sample of realisations per Nside is also run
    through the per-spin FP64 transforms for the error and the FP64 time per realisation."""
    import torch

    from gmaster_amd import sht, simbatch

    nsides = [int(x) for x in args.sim_nsides.split(",")]
    n = args.sim_batch
    counts = [n // len(nsides) + (1 if i < n % len(nsides) else 0) for i in range(len(nsides))]
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    t = time.perf_counter()
    job = simbatch.SimJob(nsides, counts, n_iter=3, device=dev, log=log)
    log(f"batch sizes tuned and buffers allocated in {time.perf_counter() - t:.1f}s; "
        f"{len(job.batches)} batches: " + ", ".join(
            f"Nside {p['nside']}: {p['n']} sims" for p in job.parts))
    t = time.perf_counter()
    job.fill(seed=1000)
    torch.cuda.synchronize()
    log(f"{job.nsims} realisations drawn in {time.perf_counter() - t:.1f}s")
    t = time.perf_counter()
    job.run()
    torch.cuda.synchronize()
    t_eager = time.perf_counter() - t
    log(f"eager run {t_eager:.3f}s")
    t = time.perf_counter()
    job.capture()
    torch.cuda.synchronize()
    log(f"captured one graph in {time.perf_counter() - t:.1f}s (includes one eager run)")
    for _ in range(args.warmup):
        job.replay()
    torch.cuda.synchronize()
    times = []
    for i in range(args.steps):
        t = time.perf_counter()
        job.replay()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        log(f"graph replay {i}: {times[-1]:.3f}s")
    step = sum(times) / len(times)
    # per-spin FP64 transforms of a sample of realisations per Nside
    nsample = 2
    err = {"T": 0.0, "E": 0.0, "B": 0.0}
    fp64_job_s = 0.0
    for part in job.parts:
        lmax, nside, k = part["lmax"], part["nside"], min(nsample, part["n"])
        torch.cuda.synchronize()
        t = time.perf_counter()
        for i in range(k):
            at = sht.map2alm_s0(part["maps_t"][i:i + 1], lmax, n_iter=3, precision="fp64")
            ap = sht.map2alm_spin(part["maps_p"][i], lmax, 2, n_iter=3, precision="fp64")
            mt = sht.alm2map_s0(at, nside, lmax)
            mp = sht.alm2map_spin(ap, nside, lmax, 2)
        torch.cuda.synchronize()
        fp64_job_s += (time.perf_counter() - t) / k * part["n"]
        for i in range(k):
            at = sht.map2alm_s0(part["maps_t"][i:i + 1], lmax, n_iter=3, precision="fp64")[0]
            ap = sht.map2alm_spin(part["maps_p"][i], lmax, 2, n_iter=3, precision="fp64")
            for name, got, ref in (("T", part["alm_t"][i], at), ("E", part["alm_p"][i, 0], ap[0]),
                                   ("B", part["alm_p"][i, 1], ap[1])):
                err[name] = max(err[name], float((got - ref).abs().max() / ref.abs().max()))
    log(f"fp64 per realisation, extrapolated to the job: {fp64_job_s:.1f}s; errors vs fp64 (sample): {err}")
    print(json.dumps({
        "metric": f"Batched SHT: {job.nsims} T+P realisations of Nside {'/'.join(map(str, nsides))} "
                  f"(map2alm n_iter 3 + alm2map), one HIP graph",
        "value": job.nsims / step, "unit": "sims/s", "higher_is_better": True, "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": step * 1e3,
        "dtype": "f64 recurrence; f32 / f16 hi-lo matrix-unit contraction (f64 near the poles)",
        "data": "synthetic",
        "config": {"workload": f"{job.nsims} Gaussian T+P skies (benchmark spectra) of Nside "
                               f"{', '.join(map(str, nsides))}, lmax 3 Nside - 1, n_iter 3",
                   "counts": counts, "batches": len(job.batches),
                   "batch_sizes": {str(p["nside"]): list(simbatch.tune_batches(p["nside"], 3, dev))
                                   for p in job.parts}},
        "eager_s": t_eager, "graph_replay_s": times,
        "fp64_per_spin_s_extrapolated": fp64_job_s, "speedup_vs_fp64": fp64_job_s / step,
        "alm_err_vs_fp64_sample": err, "fp64_sample_per_nside": nsample,
        "peak_device_memory_gb": torch.cuda.max_memory_allocated(dev) / 1e9,
    }))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--nside", type=int, default=4096)
    ap.add_argument("--nlb", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="target CPU sample time")
    ap.add_argument("--form", default="fast", choices=["fp64", "fast", "learned"],
                    help="form timed for the metric (fp64 is always timed beside it)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run one 1 GiB device-to-device copy first (scales PMC byte counters)")
    ap.add_argument("--sht-batch", type=int, default=0,
                    help="instead of the MASTER pipeline, time K spin-0 maps through the batched transforms")
    ap.add_argument("--max-group", type=int, default=2, help="maps per shared recurrence for --sht-batch")
    ap.add_argument("--coupling-forms", action="store_true",
                    help="config 4: exact vs Toeplitz vs learned coupling at --nside instead")
    ap.add_argument("--fresh-masks", type=int, default=2, help="random masks beside the benchmark mask")
    ap.add_argument("--sim-batch", type=int, default=0,
                    help="config 3: time N T+P realisations of mixed Nside as one HIP graph instead")
    ap.add_argument("--sim-nsides", default="256,512,1024", help="Nside values of --sim-batch")
    ap.add_argument("--sims", type=int, default=0,
                    help="time N T+P simulations sharing one workspace trio (FP64 vs chosen form) instead")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; multi-GPU runs are not part of this build "
              f"(got --gpus {args.gpus})", file=sys.stderr)
        sys.exit(2)

    if args.sht_batch > 0:
        return sht_batch(args)
    if args.sims > 0:
        return sims(args)
    if args.sim_batch > 0:
        return sim_batch(args)
    if args.coupling_forms:
        return coupling_forms(args)

    import torch

    import gmaster_amd as nmt
    from gmaster_amd import _lib, workload

    nside, nlb = args.nside, args.nlb
    lmax = 3 * nside - 1
    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    log(f"inputs: nside={nside} lmax={lmax}")
    t0 = time.time()
    mask = workload.hole_mask(nside, nholes=20, radius_deg=5.0, apo_deg=2.0, kind="C1", seed=1234, device=dev)
    tqu = workload.tqu_maps(nside, lmax, seed=4321, device=dev)
    torch.cuda.synchronize()
    log(f"inputs ready in {time.time() - t0:.1f}s (fsky={float(mask.mean()):.3f})")

    if args.calibrate_copy:
        src = torch.empty(1 << 27, dtype=torch.float64, device=dev).fill_(1.0)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        log("calibration copy: 1 GiB read + 1 GiB written")

    def timed(form, with_timers):
        for i in range(args.warmup):
            t = time.time()
            run_pipeline(nmt, mask, tqu, nside, nlb, form)
            torch.cuda.synchronize()
            log(f"{form} warmup {i}: {time.time() - t:.3f}s")
        torch.cuda.reset_peak_memory_stats(dev)
        timers = _lib.enable_timers(True) if with_timers else None
        times, info = [], []
        bp = None
        for i in range(args.steps):
            info.clear()
            keep.pop(form, None)
            kept = keep.setdefault(form, {}) if i == args.steps - 1 else None
            torch.cuda.synchronize()
            t = time.perf_counter()
            bp = run_pipeline(nmt, mask, tqu, nside, nlb, form, info, kept)
            torch.cuda.synchronize()
            times.append(time.perf_counter() - t)
            log(f"{form} step {i}: {times[-1]:.3f}s")
        if with_timers:
            _lib.enable_timers(False)
        peak_mem[form] = torch.cuda.max_memory_allocated(dev) / 1e9
        return times, bp, timers, list(info)

    forms = ["fp64"] if args.form == "fp64" else ["fp64", args.form]
    keep, peak_mem = {}, {}
    runs = {f: timed(f, True) for f in forms}
    times_fp64, bp_fp64, timers_fp64, _ = runs["fp64"]
    times, bp, timers, coupling_info = runs[args.form]
    total = sum(times)
    bp_err = bandpower_errors(bp, bp_fp64) if args.form != "fp64" else None
    alm_err = None
    if bp_err is not None:
        log(f"{args.form} vs fp64 bandpower errors: {bp_err}")
        alm_err = alm_errors(keep[args.form]["alms"], keep["fp64"]["alms"])
        log(f"{args.form} vs fp64 alm errors (max |da| / max |a|): {alm_err}")
    keep.clear()

    def per_call(tm):
        # per entry point: mean duration per launch (ms) over the timed steps
        out = {}
        for name, evs in tm.items():
            ms = [s_.elapsed_time(e) for s_, e in evs]
            out[name] = {"calls": len(ms), "mean_ms": sum(ms) / len(ms), "total_ms_per_step": sum(ms) / args.steps}
        return out

    stages, stages_fp64 = per_call(timers), per_call(timers_fp64)
    fp64_s = sum(times_fp64) / len(times_fp64)
    f0, f2 = legendre_flops(nside, lmax)

    def roof_batched(st):
        # the batched analysis: one spin-0 recurrence (2 FP64 FMA per (l, m, ring pair))
        # shared by 12 real channels (T, mask, Q, U, Q/sin^2, U/sin^2: 12 FP32 FMA on the
        # matrix unit), i.e. 2 (2 + 2k) T0 FLOP with k = 6 complex channels
        dom_ms = st.get("gm_bat_ana", {}).get("mean_ms", float("nan"))
        L = lmax + 1
        T0 = L * (L + 1) / 2 * 2 * nside
        flop = 2 * (2 + 12) * T0
        achieved = flop / (dom_ms * 1e-3)
        peak = 28.0 / (4.0 / PEAK_FP64 + 24.0 / PEAK_FP32)
        return {
            "kernel": "k_bat_ana (T + mask + Q/U analysis on one recurrence, gm_bat_ana: FP32 ring blocks on the "
                      "matrix unit + FP64 polar blocks)",
            "bound": "valu",
            "ops": "fp64 valu + fp32 mfma",
            "pipe": "FP64 VALU recurrence (78.6 TFLOP/s) + FP32 MFMA contraction (157.3 TFLOP/s): op-mix peak",
            "achieved": achieved / 1e12,
            "peak": peak / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": _measured_traffic(nside, "k_bat_ana"),
            "flop_per_launch": flop,
            "mean_launch_ms": dom_ms,
            "replaces_ms": None,
        }

    def roof(st, mixed):
        name = "gm_lat_ana_s_mixed" if mixed else "gm_lat_ana_s"
        dom_ms = st.get(name, {}).get("mean_ms", float("nan"))
        achieved = f2 / (dom_ms * 1e-3)
        # 2 (2 + 4) FLOP per (l, m, ring): 4 of the recurrence (FP64), 8 of the contraction
        # (FP64, or FP32 in the mixed kernel); the peak is the time-weighted one of that mix
        peak = PEAK_FP64 if not mixed else 12.0 / (4.0 / PEAK_FP64 + 8.0 / PEAK_FP32)
        return {
            "kernel": f"k_ana_s{'<mixed>' if mixed else ''} (spin-2 Legendre analysis, {name})",
            "bound": "valu",
            "pipe": ("FP64 VALU (MI355X FP64 vector and matrix peaks are both 78.6 TFLOP/s)" if not mixed else
                     "FP64 VALU recurrence (78.6 TFLOP/s) + FP32 VALU contraction (157.3 TFLOP/s): op-mix peak"),
            "achieved": achieved / 1e12,
            "peak": peak / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": _measured_traffic(nside, "k_ana_s_mixed" if mixed else "k_ana_s"),
            "flop_per_launch": f2,
            "mean_launch_ms": dom_ms,
        }

    L = lmax + 1
    T0 = L * (L + 1) / 2 * 2 * nside  # (l, m, folded ring pair) triples of a batched pass
    nring = 4 * nside - 1
    npix = 12 * nside * nside
    PEAK_F16 = 2.5e15  # MI355X dense FP16 matrix, FLOP/s
    PEAK_HBM = 8.0e12

    def hbm_entry(name, kernel, nbytes, what):
        ms = stages.get(name, {}).get("mean_ms", float("nan"))
        ach = nbytes / (ms * 1e-3)
        return {"kernel": kernel, "bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": ach / PEAK_HBM, "bytes_per_launch": nbytes, "bytes": what, "mean_launch_ms": ms}

    def flop_entry(name, kernel, flop, peak, ops, st=None):
        # bound "valu": the FP64 recurrence on the vector ALU sets these kernels' time (matrix
        # and vector instructions issue one after the other on a SIMD); "ops" names the mix
        ms = (st or stages).get(name, {}).get("mean_ms", float("nan"))
        ach = flop / (ms * 1e-3)
        return {"kernel": kernel, "bound": "valu", "ops": ops, "achieved": ach / 1e12, "peak": peak / 1e12,
                "unit": "TFLOP/s", "frac": ach / peak, "flop_per_launch": flop, "mean_launch_ms": ms}

    def roof_residual(st):
        # the analyses on FP16 hi/lo products (the largest share of the step): per (l, m, ring
        # pair) 2 FP64 FMA of the recurrence + 12 channel FMA, the latter as 3 FP16 hi/lo
        # products on the matrix unit; op-mix peak of FP64 VALU and FP16 MFMA
        e = flop_entry("gm_bat_ana_res", "k_bat_ana_h (T + mask + Q/U analysis, gm_bat_ana_res: FP16 "
                       "hi/lo ring blocks on the matrix unit + FP64 polar blocks)", 2 * (2 + 12) * T0,
                       28.0 / (4.0 / PEAK_FP64 + 3 * 24.0 / PEAK_F16), "fp64 valu + f16 mfma (3 hi/lo products)", st)
        e["pipe"] = ("FP64 VALU recurrence (78.6 TFLOP/s) + FP16 MFMA contraction (2.5 PFLOP/s, 3 products per "
                     "FP32 product): op-mix peak; the FP16 split (3 VALU ops per mu: FP64 -> FP32 convert, FP16 "
                     "convert, mixed fma) is not in the FLOP count")
        e["traffic"] = _measured_traffic(nside, "k_bat_ana_h")
        return e

    if "gm_bat_ana_res" in stages:
        roofline = roof_residual(stages)
    elif "gm_bat_ana" in stages:
        roofline = roof_batched(stages)
    else:
        roofline = roof(stages, args.form != "fp64")
    stage_roofs = {}
    if "gm_bat_ana" in stages:
        stage_roofs["first_analysis"] = roof_batched(stages)
    if "gm_bat_syn" in stages:
        stage_roofs["synthesis"] = flop_entry(
            "gm_bat_syn", "k_bat_syn (T + mask + Q/U synthesis, gm_bat_syn: FP32 4x4x1 MFMA + FP64 polar pairs)",
            2 * (2 + 8) * T0, 20.0 / (4.0 / PEAK_FP64 + 16.0 / PEAK_FP32), "fp64 valu + fp32 mfma")
        spec_b = nring * L * 16
        stage_roofs["fft_fwd_t_mask"] = hbm_entry(
            "gm_ring_fft_fwd_s0", "k_ring_fwd_s0 (T and mask)", 2 * (npix * 8 + spec_b), "2 maps read + 2 spectra written")
        stage_roofs["fft_fwd_qu"] = hbm_entry(
            "gm_ring_fft_fwd_s", "k_ring_fwd_s (Q + iU)", 2 * npix * 8 + nring * (2 * L - 1) * 16,
            "Q, U read + F+ spectrum written")
        stage_roofs["fold_t_mask"] = hbm_entry(
            "gm_ring_fold_residual", "k_fold_s0 (T and mask)", 3 * 2 * spec_b, "synthesis + residual read, residual written")
        stage_roofs["fold_qu"] = hbm_entry(
            "gm_ring_fold_residual_c", "k_fold_s (Q + iU)", 3 * nring * (2 * L - 1) * 16,
            "synthesis + residual read, residual written")
    stage_roofs["coupling_fp64"] = flop_entry(
        "gm_coupling", "k_coupling (exact 0-0, 0-2, 2-2 3j rows, plain FP64 form)", 24.0 * _n3j(lmax), PEAK_FP64,
        "fp64 valu", stages_fp64)
    roofline_fp64 = roof(stages_fp64, False)
    ana0_ms = stages_fp64.get("gm_lat_ana_s0", {}).get("mean_ms", float("nan"))
    cpl_ms = stages_fp64.get("gm_coupling", {}).get("mean_ms", float("nan"))

    cpu = None
    if not args.no_cpu:
        try:
            from oracle.c import cpu_baseline

            cpu = cpu_baseline.estimate_pipeline(nside, lmax, target_seconds=args.cpu_seconds, log=log)
            cpu["ladder"] = cpu_baseline.ladder((64, 256, 1024), log=log)
        except Exception as exc:  # the baseline is reported, never required
            log(f"cpu baseline unavailable: {exc!r}")

    res = {
        "metric": "Nside-4096 T+P MASTER s: chosen vs plain FP64 (1 GPU, bandpower err<=1e-4)",
        "value": total / args.steps,
        "unit": "s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "median_s": statistics.median(times),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": ("f64" if args.form == "fp64" else
                  "f64 recurrence; f16 hi/lo matrix products in the analyses (f64 within 5.9 deg of the poles), "
                  "f32 in the syntheses (f64 within 14.5 deg of the poles); f64 coupling and LU"),
        "data": "synthetic",
        "form": args.form,
        "fp64_s": fp64_s,
        "speedup_vs_fp64": fp64_s / (total / args.steps),
        "bandpower_err_vs_fp64": bp_err,
        "coupling": coupling_info[-1] if coupling_info else None,
        "config": {"workload": f"T+P full MASTER, HEALPix Nside {nside}, lmax {lmax}, n_iter 3, "
                               f"20 x 5deg holes C1 2deg, bins of {nlb}", "nside": nside, "lmax": lmax,
                   "n_iter": 3, "nlb": nlb, "parallelism": "single GPU"},
        "alm_err_vs_fp64": alm_err,
        "peak_device_memory_gb": peak_mem,
        "roofline": roofline,
        "roofline_fp64": roofline_fp64,
        "roofline_stages": stage_roofs,
        "stages": {
            "ana_s0_pair_fp64_frac": (f0 / (ana0_ms * 1e-3)) / PEAK_FP64,
            "coupling_fp64_ms": cpl_ms,
            "per_call": stages,
            "per_call_fp64": stages_fp64,
        },
        "bandpowers_finite": bool(all(torch.isfinite(x).all() for x in bp)),
        "timed_steps_s": times,
        "fp64_steps_s": times_fp64,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res))


if __name__ == "__main__":
    main()
