#!/usr/bin/env python
"""bench.py -- GLIMPSE2_phase compute core on one MI355X.

Workload (the batched configuration): a synthetic panel of 10,000 haplotypes x
60,000 biallelic SNPs (5,000 SNPs per Mb at 1 cM/Mb, 50-founder Li-Stephens
mosaics, 1/x frequency spectrum, mutation 1e-3), 200 diploid targets with read
coverage U(0.1, 2)x, and 5 chunks of U[20k, 60k] sites (windows of the panel,
each with its own sparse PBWT); every chunk is phased for all 200 targets,
Kpbwt per target U[500, 2000], default Gibbs schedule (1 INIT + 5 burn-in + 15
main).  One step = one complete phase run of every chunk (all targets, all
iterations) with the inputs already resident in HBM; chunks run concurrently
on their own HIP streams.

Two paths over the same inputs:
  plain PBWT  -- every target conditions on its Kpbwt PBWT states;
  learned     -- from iteration 1 on, the selector network keeps
                 min(round(0.75 Kpbwt), 1024) states of each target's PBWT
                 pool, falling back to the whole pool when the kept states'
                 predicted mass is below 0.95 (INIT is unchanged).
The network is first trained on the GPU (outside the timed region) on rows a
plain run records: a separate panel of 4,000 haplotypes x 10,000 SNPs (2 Mb),
48 targets at 1x, Kpbwt 2,000, every PBWT iteration; labels are the HMM's own
posterior masses.

metric value: target-site imputations/s = sum over chunks of targets x chunk
sites / wall time of one step, for the learned path when its dosage r2 is
within 0.001 of the plain path's (the "equal r2" condition), else for the
plain path (config.path says which).  Both paths' figures, dosage r2 / NRC
against the simulated truth, and the learned path's mean |DS - DS(plain)| (its
dosage error measured against the plain run at Kpbwt) are in the same line.
The roofline object reports the dominant kernel (haploid forward-backward)
against HBM: algorithmic bytes per launch / average launch duration (HIP
events on the launch stream); `traffic` is the HBM bytes per launch measured
with rocprofv3 counters for this workload
(profiles/fb_kernel_batch200x5_<path>_traffic.md).
cpu_baseline: the CPU oracle on a bounded sample of the same workload (targets
of the smallest chunk, INIT + 1 burn-in + 1 main iteration each, scaled to the
21 iterations of a full run), one target per thread on up to 16 host threads,
with the one-core figure beside it.
"""
import argparse
import json
import os
import sys
import threading
import time

# One hardware queue per concurrently running stream (HIP's default is 4; with
# 5 chunks, each with a second stream for its smaller set-size class, streams
# would share queues and run one after the other).
# Must be set before the HIP runtime starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "target-site imputations/s at equal dosage r² (learned vs PBWT K); HBM %"


def slice_panel(P, o, l):
    """Window [o, o+l) of a panel dict as its own chunk panel (cM rebased)."""
    s = slice(o, o + l)
    common = P["common"][s]
    crow0 = int(P["common"][:o].sum())
    nc = int(common.sum())
    rp = P["rare_ptr"][o:o + l + 1]
    sub = dict(L=l, N=P["N"], bp=P["bp"][s].copy(), cm=P["cm"][s] - P["cm"][o],
               cref=P["cref"][s].copy(), calt=P["calt"][s].copy(), common=common.copy(),
               major=P["major"][s].copy(), common_words=P["common_words"][crow0:crow0 + nc].copy(),
               rare_ptr=(rp - rp[0]).astype(np.int32), rare_idx=P["rare_idx"][rp[0]:rp[-1]].copy(),
               lq=np.zeros(l, np.uint8))
    return sub


def build_workload(quick, seed=15052011):
    from glimpse_amd import synth
    if quick:
        N, Lp, T, n_chunks, lo, hi, kmin, kmax = 2000, 12000, 24, 3, 4000, 12000, 200, 600
    else:
        N, Lp, T, n_chunks, lo, hi, kmin, kmax = 10000, 60000, 200, 5, 20000, 60000, 500, 2000
    P = synth.make_panel(N, Lp, T, length_mb=Lp / 5000.0, seed=seed)
    rng = np.random.default_rng(seed)
    cov = rng.uniform(0.1, 2.0, T)
    K = rng.integers(kmin, kmax + 1, T)
    lens = rng.integers(lo, hi + 1, n_chunks)
    offs = [int(rng.integers(0, Lp - ln + 1)) for ln in lens]
    gl, flat = synth.make_reads(P["tgt_haps"], cov, seed=seed + 1)
    truth = synth.truth_genotypes(P["tgt_haps"])
    chunks = []
    for c in range(n_chunks):
        tg = np.arange(T)   # every chunk (a genomic region) is phased for every target sample
        o, ln = offs[c], int(lens[c])
        chunks.append(dict(panel=slice_panel(P, o, ln), targets=tg,
                           gl=np.ascontiguousarray(gl[tg][:, o:o + ln]),
                           flat=np.ascontiguousarray(flat[tg][:, o:o + ln]),
                           K=K[tg].astype(np.int32), truth=truth[tg][:, o:o + ln]))
    desc = ("batched plain PBWT: %d haps x %d-site panel, %d targets at U(0.1,2)x, each phased on %d chunks of "
            "U[%d,%d] sites, Kpbwt U[%d,%d] per target, 1+5+15 Gibbs iterations" % (N, Lp, T, n_chunks, lo, hi, kmin, kmax))
    return chunks, desc, dict(N=N, panel_sites=Lp, targets=T, chunks=n_chunks, chunk_sites=[int(x) for x in lens])


def dosage(gp0, gp1):
    return gp1 + 2.0 * np.clip(1.0 - gp0 - gp1, 0.0, 1.0)


def accuracy(outs, chunks):
    ds = np.concatenate([dosage(o["gp0"], o["gp1"]).ravel() for o in outs])
    tr = np.concatenate([c["truth"].ravel() for c in chunks]).astype(np.float64)
    r = np.corrcoef(ds, tr)[0, 1]
    g = np.concatenate([(o["h0"].astype(int) + o["h1"]).ravel() for o in outs])
    # non-reference concordance (concordance/src/containers/call_set_writing.cpp:46-55)
    mism = np.sum((g != tr) & ((tr > 0) | (g > 0)))
    match_nonref = np.sum((g == tr) & (tr > 0))
    nrc = 1.0 - mism / max(1, mism + match_nonref)
    return dict(dosage_r2=float(r * r), nrc=float(nrc))


def cpu_baseline(chunks):
    """The CPU oracle timed on the host: one target alone (one core), then one
    target per thread on as many threads as the host share allows (up to
    16), each an INIT + 1 burn-in + 1 main run of the smallest chunk,
    scaled by 21/3 to the full 1 + 5 + 15 schedule.  The oracle's C calls
    release the interpreter lock, so the threads run in parallel."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    c = min(chunks, key=lambda x: x["panel"]["L"])
    OP = orc.Panel(c["panel"])
    L = c["panel"]["L"]

    def one(t):
        prm = orc.default_params(burnin=1, main_iter=1, kpbwt=int(c["K"][t]), rng_mode=orc.RNG_REFERENCE)
        OP.phase_run(c["gl"][t:t + 1], c["flat"][t:t + 1], prm)

    t0 = time.perf_counter()
    one(0)
    dt1 = time.perf_counter() - t0
    n = max(1, min(16, os.cpu_count() or 1, c["gl"].shape[0]))
    th = [threading.Thread(target=one, args=(t,)) for t in range(n)]
    t0 = time.perf_counter()
    for x in th:
        x.start()
    for x in th:
        x.join()
    dtn = time.perf_counter() - t0
    v1 = L / (dt1 * 21.0 / 3.0)
    vn = n * L / (dtn * 21.0 / 3.0)
    return dict(value=vn, unit="target-site imputations/s", cores=n, kind="port",
                one_core=v1,
                sample="oracle (C++ restatement) on %d targets x %d-site chunk, Kpbwt per target (%d..%d), INIT+1 "
                       "burn-in+1 main iterations, one target per thread on %d threads in %.2f s (one target on one "
                       "thread: %.2f s), scaled x21/3 to a full run"
                       % (n, L, int(c["K"][:n].min()), int(c["K"][:n].max()), n, dtn, dt1))


def train_selector(seed=31337, epochs=8):
    """Trains the selector network on rows a plain K=2000 run records."""
    from glimpse_amd import hip, synth
    t0 = time.perf_counter()
    P = synth.make_panel(4000, 10000, 48, length_mb=2.0, seed=seed)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=seed + 1)
    panel = hip.Panel(P)
    run = hip.Run(panel, gl, flat, hip.default_params(kpbwt=2000))
    run.collect(1)
    run.execute()
    rows = int(run.stats()["collected_rows"])
    t1 = time.perf_counter()
    net = hip.Mlp(seed=seed)
    losses = net.train_run(run, epochs=epochs, batch=16384, lr=2e-3, seed=seed)
    t2 = time.perf_counter()
    run.close()
    info = dict(rows=rows, epochs=epochs, label_run_s=round(t1 - t0, 2), train_s=round(t2 - t1, 3),
                loss_first=float(losses[0]), loss_last=float(losses[-1]),
                data="4000 haps x 10000 SNPs (2 Mb), 48 targets at 1x, Kpbwt 2000, rows from iterations 1-20")
    return net, info


def timed(runs, a, label):
    import torch

    def step():
        if a.serial:
            for r in runs:
                r.execute()
            return
        errs = []

        def go(r):
            try:
                r.execute()
            except Exception as e:  # surfaced below
                errs.append(e)
        th = [threading.Thread(target=go, args=(r,)) for r in runs]
        for t in th:
            t.start()
        for t in th:
            t.join()
        if errs:
            raise errs[0]

    for i in range(a.warmup):
        step()
        print("[bench] %s warmup step %d done" % (label, i), file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for i in range(a.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / max(1, a.steps)
    print("[bench] %s: %.2f s per step" % (label, dt), file=sys.stderr, flush=True)
    return dt, [r.fetch() for r in runs]


# workload tag of the committed rocprofv3 summaries (200 targets x 5 chunks, full size)
PROFILE_WORKLOAD = "batch200x5"


def profiled_traffic(path):
    """HBM bytes per forward-backward launch from the committed rocprofv3 counter
    summary of this workload and path (FETCH_SIZE and WRITE_SIZE runs, scaled by
    a 1 GiB copy's counter readings), or None when there is none."""
    f = os.path.join(ROOT, "profiles", "fb_kernel_%s_%s_traffic.md" % (PROFILE_WORKLOAD, path))
    if not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith("traffic_bytes_per_launch"):
            return float(line.split("=")[1])
    return None


def fb_roofline(outs, path=None):
    fb_bytes = sum(o["stats"]["fb_bytes"] for o in outs)
    fb_ms = sum(o["stats"]["fb_kernel_ms"] for o in outs)
    fb_n = sum(o["stats"]["fb_launches"] for o in outs)
    per_launch = fb_bytes / fb_n
    avg_ms = fb_ms / fb_n
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": profiled_traffic(path) if path else None,
            "kernel": "fb_kernel (haploid forward-backward)",
            "bytes_per_launch": per_launch, "avg_launch_ms": avg_ms, "launches_per_step": fb_n}


def stage_ms(outs):
    return {k: round(sum(o["stats"][k] for o in outs), 1)
            for k in ("ms_select", "ms_learn", "ms_condset", "ms_fb", "ms_phase", "ms_other")}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--quick", action="store_true", help="small workload (profiling / smoke)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--serial", action="store_true", help="run chunks one after another")
    ap.add_argument("--path", choices=("both", "plain", "learned"), default="both")
    ap.add_argument("--k-frac", type=float, default=0.75,
                    help="learned path: kept fraction of the PBWT pool (0.65 was the smallest tried whose dosage r2 "
                         "is within 0.001 of the plain path's on this workload; 0.75 keeps a margin at equal speed)")
    ap.add_argument("--k-max", type=int, default=1024,
                    help="learned path: at most this many kept states per target (0: no cap); 1024 keeps the "
                         "learned sets in the 1024-state launch class")
    ap.add_argument("--min-mass", type=float, default=0.95,
                    help="learned path: a target whose kept states hold less than this share of the predicted mass "
                         "keeps its whole PBWT pool (the plain selection) for that iteration")
    ap.add_argument("--latency-runs", type=int, default=0,
                    help="the N chunks with the most work (sites x states) use latency launch shapes (more waves "
                         "per target); with every chunk in flight the throughput defaults were faster for all, so 0")
    ap.add_argument("--tune", action="store_true",
                    help="autotune the launch shapes during the warm-up steps (each chunk's tuner times its own "
                         "launches; with all chunks in flight the measured defaults are faster)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of 1 GiB three times (HBM counter calibration)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build runs on exactly one GPU (the Gibbs chains are independent per target and "
              "chunk; multi-GPU sharding is design only); run with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    torch.cuda.synchronize()   # bring torch's HIP runtime up before the library's first call
    if a.copy_calibration:
        n = 1 << 28   # 2^28 floats = 1 GiB read + 1 GiB written per copy
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n, "copies": 3}), flush=True)
        return
    from glimpse_amd import hip

    t_setup = time.perf_counter()
    chunks, desc, meta = build_workload(a.quick)
    panels = [hip.Panel(c["panel"]) for c in chunks]
    prm = hip.default_params()
    runs = [hip.Run(p, c["gl"], c["flat"], prm, kpbwt_per_target=c["K"]) for p, c in zip(panels, chunks)]
    if a.tune:
        for r in runs:
            r.set_tuning(1)   # the warm-up executions time each launch shape and keep the fastest
    else:
        work = [c["panel"]["L"] * int(np.sum(c["K"])) for c in chunks]
        for ci in sorted(range(len(chunks)), key=lambda i: -work[i])[:max(0, a.latency_runs)]:
            runs[ci].set_tuning("latency")
    setup_s = time.perf_counter() - t_setup
    print("[bench] setup %.1f s (%s)" % (setup_s, desc), file=sys.stderr, flush=True)
    tsites = sum(len(c["targets"]) * c["panel"]["L"] for c in chunks)

    res = {}
    if a.path in ("both", "plain"):
        dt, outs = timed(runs, a, "plain")
        res["plain"] = dict(value=tsites / dt, ms_per_step=dt * 1e3, outs=outs)
    free_b, total_b = torch.cuda.mem_get_info()
    hbm_used_gb = (total_b - free_b) / 1e9   # device memory in use with every chunk's run resident
    train = None
    if a.path in ("both", "learned"):
        net, train = train_selector()
        print("[bench] selector trained: %s" % json.dumps(train), file=sys.stderr, flush=True)
        for r, c in zip(runs, chunks):
            kl = np.maximum(1, np.round(c["K"] * a.k_frac))
            if a.k_max > 0:
                kl = np.minimum(kl, a.k_max)
            r.set_selector(net, kl.astype(np.int32), a.min_mass)
        dt, outs = timed(runs, a, "learned")
        res["learned"] = dict(value=tsites / dt, ms_per_step=dt * 1e3, outs=outs)
    for k, v in res.items():
        v["accuracy"] = accuracy(v["outs"], chunks)
        v["stage_ms_summed_over_chunks"] = stage_ms(v["outs"])
    # the headline path: the learned selector when its dosage r2 is within 0.001
    # of the plain path's, else the plain path
    main_path = "learned" if "learned" in res else "plain"
    if "learned" in res and "plain" in res and \
            abs(res["learned"]["accuracy"]["dosage_r2"] - res["plain"]["accuracy"]["dosage_r2"]) > 0.001:
        main_path = "plain"
    M = res[main_path]
    line = {
        "metric": METRIC,
        "value": M["value"],
        "unit": "target-site imputations/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": M["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded simulator: founder mosaics + Poisson reads)",
        "config": dict(workload=desc, path=main_path,
                       concurrency="serial" if a.serial else "chunks on concurrent streams", **meta),
        "roofline": fb_roofline(M["outs"], main_path if a.quick is False else None),
        "accuracy": M["accuracy"],
        "stage_ms_summed_over_chunks": M["stage_ms_summed_over_chunks"],
        "launch_shapes": [dict(fb="%dx%d" % (o["stats"]["fb_nw"], o["stats"]["fb_mwl"]),
                               phase="%dx%d" % (o["stats"]["ph_nw"], o["stats"]["ph_kpt"])) for o in M["outs"]],
        "setup_s": setup_s,
        "hbm_used_gb": round(hbm_used_gb, 1),
        "chunk_wall_ms": [round(o["stats"]["ms_total"], 1) for o in M["outs"]],
    }
    if "learned" in res:
        L_ = res["learned"]
        st = [o["stats"] for o in L_["outs"]]
        line["learned"] = dict(value=L_["value"], ms_per_step=L_["ms_per_step"], accuracy=L_["accuracy"],
                               roofline=fb_roofline(L_["outs"], None if a.quick else "learned"),
                               k_frac=a.k_frac, k_max=a.k_max, min_mass=a.min_mass, training=train,
                               target_iterations_learned=sum(s["learned_sel"] for s in st),
                               target_iterations_pool_kept=sum(s["pool_kept"] for s in st),
                               mean_states=float(np.mean(np.concatenate(
                                   [o["nstates"][:, 1:].ravel() for o in L_["outs"]]))))
    if "plain" in res:
        Pn = res["plain"]
        line["plain_pbwt"] = dict(value=Pn["value"], ms_per_step=Pn["ms_per_step"], accuracy=Pn["accuracy"],
                                  roofline=fb_roofline(Pn["outs"], None if a.quick else "plain"),
                                  stage_ms_summed_over_chunks=Pn["stage_ms_summed_over_chunks"],
                                  mean_states=float(np.mean(np.concatenate(
                                      [o["nstates"][:, 1:].ravel() for o in Pn["outs"]]))))
        if "learned" in res:
            dsl = np.concatenate([dosage(o["gp0"], o["gp1"]).ravel() for o in res["learned"]["outs"]])
            dsp = np.concatenate([dosage(o["gp0"], o["gp1"]).ravel() for o in Pn["outs"]])
            line["learned"]["mean_abs_ds_vs_plain"] = float(np.mean(np.abs(dsl - dsp)))
            line["learned"]["r2_minus_plain"] = res["learned"]["accuracy"]["dosage_r2"] - Pn["accuracy"]["dosage_r2"]
            line["learned"]["speedup_vs_plain"] = res["learned"]["value"] / Pn["value"]
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(chunks)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
