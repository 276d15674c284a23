#!/usr/bin/env python
"""bench.py -- GLIMPSE2_phase compute core on one MI355X.

Default workload: the mixed production batch (C5).  A reference panel of
100,000 haplotypes split into 10 chunks of U[40k, 150k] biallelic SNPs
(about 6,700 SNPs per cM, 75% of candidate sites injected rare and the 1/x
spectrum adding more, so about 80% of sites have MAF < 0.1%; 50-founder
Li-Stephens mosaics, mutation 1e-3), 1,000 diploid targets with read coverage
log-uniform in 0.1-4x; every chunk is phased for all 1,000 targets with the
default Gibbs schedule (1 INIT + 5 burn-in + 15 main), Kpbwt 1,000.  The
chunks are independent synthetic regions (one panel, one sparse PBWT and one
set of held-out target haplotypes each).

One step = the whole job: every (chunk, block of 125 targets) unit runs one
complete phase run (all iterations) with its inputs already resident in HBM.
The 80 units are served longest chunk first by 16 host threads, each executing
its units one after another in its own device workspace (which owns the
execution streams), so 16 runs (2,000 Gibbs chains) are in flight at a time
and only their workspaces are held (about 240 GB of HBM in all; 8 units of
250 targets: 4% slower on the plain path, 2% on the learned one).
Within a unit, iterations 0 (INIT) and 1 are launched from the host and
iteration 2 is captured as a HIP graph that makes every per-iteration decision
on the device; the graph is replayed for iterations 3-20 (gh_run_set_graph).
Before the timed step each worker executes its largest unit once on both
paths (outputs discarded) so the workspaces are allocated; --warmup adds full
untimed steps.

Two paths over the same resident inputs:
  plain PBWT  -- every target conditions on its Kpbwt PBWT states;
  learned     -- from iteration 1 on, the selector network keeps
                 round(keep x Kpbwt) states of each target's PBWT pool, or the
                 whole pool when the kept states' predicted mass is below
                 0.95 (INIT is unchanged).
The network is trained on the GPU (outside the timed region) on rows a plain
run records: 4,000 haplotypes x 10,000 SNPs (2 Mb), 48 targets at 1x, Kpbwt
2,000; labels are the HMM's own posterior masses.

metric value: target-site imputations/s = sum over chunks of targets x chunk
sites / wall time of one step, for the learned path when its dosage r2 is
within 0.001 of the plain path's (equal r2), else for the plain path
(config.path says which).  Both paths' figures, per-chunk throughput, dosage
r2 / NRC against the simulated truth, and the learned path's distance to the
plain run are in the same line.  large_k_judge: the smallest chunk phased
again (after the timed steps) with plain PBWT selection at Kpbwt 2,000, the
large-K plain HMM run that judges the paths' dosage error (mean |DS - DS_2000|
and r2 on that chunk).

roofline: the dominant kernel (haploid forward-backward) against HBM,
algorithmic bytes per dispatch / average dispatch duration (HIP events on the
launch stream), measured on one run alone at full occupancy (all 1,000 targets
of the largest chunk in one run, plain path: 2,000 haplotype items per dispatch
keep every CU's workgroup slots filled; every iteration launched from the host
so each dispatch is timed; --probe-targets sets the count), after the step's
runs and workspaces are released, so no other dispatch shares its time;
`step_level_frac` is the whole step's algorithmic forward-backward bytes over
the step time; `traffic` is the measured HBM bytes
per dispatch from the committed rocprofv3 counter summary of the same probe.
cpu_baseline: the CPU oracle (a C++ restatement of the reference path) timed
on this host on a bounded sample of the same workload: full 1+5+15 runs of as
many targets as threads (up to 16) on the smallest chunk, the PBWT sweep of
an iteration serial and the targets spread over the threads as the
reference's --threads does.

--workload c2 runs the batched configuration instead (10,000 haplotypes x
60,000 SNPs, 200 targets at U(0.1, 2)x, 5 chunks of U[20k, 60k] sites, Kpbwt
U[500, 2000], every chunk one unit, all 5 in flight).
"""
import argparse
import json
import os
import queue
import sys
import threading
import time

# Two HIP streams per run in flight (K-class batching); HIP's default of 4
# hardware queues would serialise them.  Must be set before the runtime starts.
if int(os.environ.get("GPU_MAX_HW_QUEUES", "4") or 4) < 16:
    os.environ["GPU_MAX_HW_QUEUES"] = "16"

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "target-site imputations/s at equal dosage r² (learned vs PBWT K); HBM %"


def log(msg):
    print("[bench] " + msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------ workloads
def build_c5(quick, seed):
    """C5 chunks: independent synthetic regions of one 100k-haplotype panel."""
    from glimpse_amd import synth
    if quick:
        N, T, n_chunks, lo, hi, block = 20000, 200, 3, 10000, 30000, 100
    else:
        N, T, n_chunks, lo, hi, block = 100000, 1000, 10, 40000, 150000, 125
    rng = np.random.default_rng(seed)
    lens = [int(x) for x in rng.integers(lo, hi + 1, n_chunks)]
    cov = np.exp(rng.uniform(np.log(0.1), np.log(4.0), T))
    chunks = []
    for c, L in enumerate(lens):
        P = synth.make_panel(N, L, T, length_mb=L / 4167.0, rare_frac=0.75, seed=seed + 101 * (c + 1), fast=True)
        gl, flat = synth.make_reads(P["tgt_haps"], cov, seed=seed + 103 * (c + 1), threads=0)
        truth = synth.truth_genotypes(P["tgt_haps"])
        del P["tgt_haps"]
        chunks.append(dict(panel=P, gl=gl, flat=flat, truth=truth, K=np.full(T, 1000, np.int32), L=L))
        log("chunk %d: %d sites, %.1f cM, common %.3f" % (c, L, P["cm"][-1], P["common"].mean()))
    desc = ("C5 mixed production batch: %d haplotypes, %d chunks of U[%d,%d] sites (rare-heavy, ~80%% MAF<0.1%%), "
            "%d targets at log-uniform 0.1-4x, Kpbwt 1000, 1+5+15 Gibbs iterations" % (N, n_chunks, lo, hi, T))
    meta = dict(N=N, targets=T, chunks=n_chunks, chunk_sites=lens, kpbwt=1000, block=block)
    return chunks, desc, meta


def slice_panel(P, o, l):
    """Window [o, o+l) of a panel dict as its own chunk panel (cM rebased)."""
    s = slice(o, o + l)
    common = P["common"][s]
    crow0 = int(P["common"][:o].sum())
    nc = int(common.sum())
    rp = P["rare_ptr"][o:o + l + 1]
    return dict(L=l, N=P["N"], bp=P["bp"][s].copy(), cm=P["cm"][s] - P["cm"][o],
                cref=P["cref"][s].copy(), calt=P["calt"][s].copy(), common=common.copy(),
                major=P["major"][s].copy(), common_words=P["common_words"][crow0:crow0 + nc].copy(),
                rare_ptr=(rp - rp[0]).astype(np.int32), rare_idx=P["rare_idx"][rp[0]:rp[-1]].copy(),
                lq=np.zeros(l, np.uint8))


def build_c2(quick, seed):
    from glimpse_amd import synth
    if quick:
        N, Lp, T, n_chunks, lo, hi, kmin, kmax = 2000, 12000, 24, 3, 4000, 12000, 200, 600
    else:
        N, Lp, T, n_chunks, lo, hi, kmin, kmax = 10000, 60000, 200, 5, 20000, 60000, 500, 2000
    P = synth.make_panel(N, Lp, T, length_mb=Lp / 5000.0, seed=seed)
    rng = np.random.default_rng(seed)
    cov = rng.uniform(0.1, 2.0, T)
    K = rng.integers(kmin, kmax + 1, T).astype(np.int32)
    lens = [int(x) for x in rng.integers(lo, hi + 1, n_chunks)]
    offs = [int(rng.integers(0, Lp - ln + 1)) for ln in lens]
    gl, flat = synth.make_reads(P["tgt_haps"], cov, seed=seed + 1)
    truth = synth.truth_genotypes(P["tgt_haps"])
    chunks = []
    for o, ln in zip(offs, lens):
        chunks.append(dict(panel=slice_panel(P, o, ln), gl=np.ascontiguousarray(gl[:, o:o + ln]),
                           flat=np.ascontiguousarray(flat[:, o:o + ln]), K=K, truth=truth[:, o:o + ln], L=ln))
    desc = ("C2 batched: %d haps x %d-site panel, %d targets at U(0.1,2)x, each phased on %d chunks of "
            "U[%d,%d] sites, Kpbwt U[%d,%d] per target, 1+5+15 Gibbs iterations" % (N, Lp, T, n_chunks, lo, hi,
                                                                                      kmin, kmax))
    meta = dict(N=N, targets=T, chunks=n_chunks, chunk_sites=lens, kpbwt="U[%d,%d]" % (kmin, kmax), block=T)
    return chunks, desc, meta


def make_panels(chunks):
    """Device panels (sparse-PBWT build + upload), one host thread per chunk."""
    from glimpse_amd import hip
    panels = [None] * len(chunks)
    errs = []

    def one(i):
        try:
            panels[i] = hip.Panel(chunks[i]["panel"])
        except Exception as e:   # surfaced below
            errs.append(e)
    th = [threading.Thread(target=one, args=(i,)) for i in range(len(chunks))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    if errs:
        raise errs[0]
    return panels


# ------------------------------------------------------------------ accuracy
def dosage(gp0, gp1):
    return gp1 + 2.0 * np.clip(1.0 - gp0 - gp1, 0.0, 1.0)


def accuracy(outs, chunks):
    """Dosage r2 (basic_stats.h:61-125) and NRC = 1 - NRD
    (concordance/src/containers/call_set_writing.cpp:46-55) against the truth,
    over every chunk, target and site."""
    ds = np.concatenate([dosage(o["gp0"], o["gp1"]).ravel() for o in outs])
    tr = np.concatenate([c["truth"].ravel() for c in chunks]).astype(np.float64)
    r = np.corrcoef(ds, tr)[0, 1]
    g = np.concatenate([(o["h0"].astype(int) + o["h1"]).ravel() for o in outs])
    mism = np.sum(g != tr)
    match_nonref = np.sum((g == tr) & (tr > 0))
    nrc = 1.0 - mism / max(1, mism + match_nonref)
    return dict(dosage_r2=float(r * r), nrc=float(nrc))


def large_k_judge(chunks, panels, workspaces, res, units, k_judge=2000):
    """The large-K plain HMM run as the judge of the learned path's dosage
    error: the smallest chunk phased again with plain PBWT selection at
    Kpbwt = k_judge for every target (outside the timed region); the mean
    absolute DS difference of each timed path's output from it, and each one's
    dosage r2 on that chunk."""
    ci = min(range(len(chunks)), key=lambda c: chunks[c]["L"])
    c = chunks[ci]
    T = c["gl"].shape[0]
    jc = dict(c, K=np.full(T, k_judge, np.int32))
    block = max(u.hi - u.lo for u in units)
    ju = [Unit(jc, ci, lo, min(lo + block, T), panels[ci]) for lo in range(0, T, block)]
    run_units(ju, workspaces[:len(ju)], "large-K judge (chunk %d, K=%d)" % (ci, k_judge))
    fs = [u.run.fetch() for u in sorted(ju, key=lambda u: u.lo)]
    ds_j = np.concatenate([dosage(f["gp0"], f["gp1"]) for f in fs])
    for u in ju:
        u.run.close()
    tr = c["truth"].astype(np.float64).ravel()
    out = dict(chunk_sites=c["L"], kpbwt=k_judge,
               dosage_r2=float(np.corrcoef(ds_j.ravel(), tr)[0, 1] ** 2))
    for p, R in res.items():
        ds_p = R["ds_small"]
        out[p] = dict(mean_abs_ds_vs_judge=float(np.mean(np.abs(ds_p - ds_j))),
                      dosage_r2=float(np.corrcoef(ds_p.ravel(), tr)[0, 1] ** 2))
    return out


# ------------------------------------------------------------------ CPU baseline
def cpu_baseline(chunks):
    """The oracle on this host: full 1 + 5 + 15 runs of `cores` targets on the
    smallest chunk, the targets of an iteration spread over `cores` threads
    after the serial PBWT sweep (the reference's --threads model; counter
    draws, so the threaded run is reproducible)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc
    c = min(chunks, key=lambda x: x["L"])
    cores = max(1, min(16, os.cpu_count() or 1, c["gl"].shape[0]))
    t0 = time.perf_counter()
    OP = orc.Panel(c["panel"])
    prep = time.perf_counter() - t0
    kp = np.ascontiguousarray(c["K"][:cores], np.int32)
    prm = orc.default_params(rng_mode=orc.RNG_COUNTER, n_threads=cores)
    t0 = time.perf_counter()
    OP.phase_run(c["gl"][:cores], c["flat"][:cores], prm, kpbwt_per_target=kp)
    dt = time.perf_counter() - t0
    return dict(value=cores * c["L"] / dt, unit="target-site imputations/s", cores=cores, kind="port",
                sample="oracle (C++ restatement of the reference path) on %d targets x the %d-site chunk "
                       "(N=%d, Kpbwt %d): one full 1+5+15 run, %.1f s on %d threads (panel and PBWT build %.1f s "
                       "not counted)" % (cores, c["L"], c["panel"]["N"], int(kp[0]), dt, cores, prep))


# ------------------------------------------------------------------ selector training
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
    return net, dict(rows=rows, epochs=epochs, label_run_s=round(t1 - t0, 2), train_s=round(t2 - t1, 3),
                     loss_first=float(losses[0]), loss_last=float(losses[-1]),
                     data="4000 haps x 10000 SNPs (2 Mb), 48 targets at 1x, Kpbwt 2000, rows from iterations 1-20")


# ------------------------------------------------------------------ units and workers
class Unit:
    def __init__(self, chunk, ci, lo, hi, panel):
        from glimpse_amd import hip
        self.ci, self.lo, self.hi, self.L = ci, lo, hi, chunk["L"]
        self.run = hip.Run(panel, chunk["gl"][lo:hi], chunk["flat"][lo:hi], hip.default_params(),
                           kpbwt_per_target=chunk["K"][lo:hi])
        self.t0 = self.t1 = 0.0

    @property
    def work(self):
        return self.L * (self.hi - self.lo)


def probe_unit(chunks, panels, n):
    """The roofline probe's unit: the first n targets of the largest chunk,
    whatever the step's block size (2n haplotype items per forward-backward
    dispatch)."""
    ci = max(range(len(chunks)), key=lambda c: chunks[c]["L"])
    c = chunks[ci]
    return Unit(c, ci, 0, min(n, c["gl"].shape[0]), panels[ci])


def run_units(units, workspaces, label):
    """Serves the units (longest first) to one host thread per workspace;
    returns the wall time.  Each thread executes its units one after another."""
    q = queue.Queue()
    for u in sorted(units, key=lambda u: -u.work):
        q.put(u)
    errs = []
    T0 = time.perf_counter()

    def worker(ws):
        while True:
            try:
                u = q.get_nowait()
            except queue.Empty:
                return
            try:
                u.run.set_workspace(ws)
                u.t0 = time.perf_counter() - T0
                u.run.execute()
                u.t1 = time.perf_counter() - T0
            except Exception as e:   # surfaced below
                errs.append(e)
                return
    th = [threading.Thread(target=worker, args=(ws,)) for ws in workspaces]
    for t in th:
        t.start()
    for t in th:
        t.join()
    dt = time.perf_counter() - T0
    if errs:
        raise errs[0]
    log("%s: %.2f s for %d units" % (label, dt, len(units)))
    return dt


def set_path(units, path, net, keep, min_mass, chunks, mode=1):
    for u in units:
        if path == "learned":
            kl = np.maximum(1, np.round(chunks[u.ci]["K"][u.lo:u.hi] * keep)).astype(np.int32)
            u.run.set_selector(net, kl, min_mass, mode=mode)
        else:
            u.run.set_selector(None, None)


def gather(units, chunks):
    """Outputs per chunk, targets in order."""
    outs = []
    for ci in range(len(chunks)):
        us = sorted([u for u in units if u.ci == ci], key=lambda u: u.lo)
        fs = [u.run.fetch() for u in us]
        o = {k: np.concatenate([f[k] for f in fs]) for k in ("gp0", "gp1", "h0", "h1", "nstates")}
        o["stats"] = [u.run.stats() for u in us]
        outs.append(o)
    return outs


def fb_roofline(stats_list):
    fb_bytes = sum(s["fb_bytes"] for s in stats_list)
    fb_ms = sum(s["fb_kernel_ms"] for s in stats_list)
    fb_n = sum(s["fb_launches"] for s in stats_list)
    per_launch = fb_bytes / fb_n
    avg_ms = fb_ms / fb_n
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    return dict(achieved=achieved, bytes_per_launch=per_launch, avg_launch_ms=avg_ms, launches=fb_n,
                fb_bytes=fb_bytes)


def profile_value(name, key):
    """A figure from a committed rocprofv3 summary (profiles/<name>): the line `key = value`."""
    f = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(f):
        return None
    for line in open(f):
        if line.startswith(key):
            try:
                return float(line.split("=")[1].split()[0])
            except ValueError:
                return None
    return None


def per_chunk(units, chunks):
    rows = []
    for ci, c in enumerate(chunks):
        us = [u for u in units if u.ci == ci]
        t0, t1 = min(u.t0 for u in us), max(u.t1 for u in us)
        n = sum(u.hi - u.lo for u in us)
        rows.append(dict(sites=c["L"], window_s=round(t1 - t0, 2),
                         unit_s=round(sum(u.t1 - u.t0 for u in us), 2),
                         imputations_per_s=round(n * c["L"] / (t1 - t0))))
    return rows


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1)
    ap.add_argument("--warmup", type=int, default=0)
    ap.add_argument("--workload", choices=("c5", "c2"), default="c5")
    ap.add_argument("--quick", action="store_true", help="small workload (smoke / profiling of the plumbing)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-judge", action="store_true", help="skip the large-K plain run that judges dosage error")
    ap.add_argument("--path", choices=("both", "plain", "learned"), default="both")
    ap.add_argument("--workers", type=int, default=0, help="units in flight (default: 16 for C5, all chunks for C2)")
    ap.add_argument("--block", type=int, default=0, help="targets per unit (default: 125 for C5, all for C2)")
    ap.add_argument("--keep", type=float, default=0.6,
                    help="learned path: kept fraction of each target's PBWT pool")
    ap.add_argument("--min-mass", type=float, default=0.95,
                    help="learned path: a target whose kept states hold less than this share of the predicted mass "
                         "keeps its whole PBWT pool (the plain selection) for that iteration")
    ap.add_argument("--probe-targets", type=int, default=1000,
                    help="targets of the roofline probe's unit (2 haplotype items each per dispatch)")
    ap.add_argument("--probe", action="store_true",
                    help="only the roofline probe: the largest chunk's first targets alone, plain path (for rocprofv3)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a device-to-device copy of 1 GiB three times (HBM counter calibration)")
    ap.add_argument("--graph", type=int, choices=(0, 1, 2), default=1,
                    help="1: iterations from 2 on replay a captured HIP graph; 0: every iteration launched from the "
                         "host (gh_run_set_graph)")
    ap.add_argument("--sel-mode", type=int, default=1, choices=(0, 1),
                    help="learned selection mode (gh_run_set_selector): 0 best scores of the pool, 1 groups served "
                         "round robin")
    ap.add_argument("--seed", type=int, default=15052011)
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
    build = build_c5 if a.workload == "c5" else build_c2
    chunks, desc, meta = build(a.quick, a.seed)
    log("inputs %.1f s" % (time.perf_counter() - t_setup))
    panels = make_panels(chunks)
    log("panels %.1f s" % (time.perf_counter() - t_setup))
    block = a.block or meta["block"]
    meta["block"] = block
    units = [Unit(c, ci, lo, min(lo + block, c["gl"].shape[0]), panels[ci])
             for ci, c in enumerate(chunks) for lo in range(0, c["gl"].shape[0], block)]
    for u in units:
        u.run.set_graph(a.graph)
    n_workers = a.workers or (16 if a.workload == "c5" else len(chunks))
    workspaces = [hip.Workspace() for _ in range(n_workers)]
    tsites = sum(c["gl"].shape[0] * c["L"] for c in chunks)
    setup_s = time.perf_counter() - t_setup
    log("setup %.1f s: %s; %d units, %d in flight" % (setup_s, desc, len(units), n_workers))

    if a.probe:
        u = probe_unit(chunks, panels, a.probe_targets)
        set_path([u], "plain", None, 0, 0, chunks)
        u.run.set_graph(0)
        u.run.set_workspace(workspaces[0])
        u.run.execute()
        st = u.run.stats()
        rf = fb_roofline([st])
        free_b, total_b = torch.cuda.mem_get_info()
        print(json.dumps(dict(hbm_used_gb=round((total_b - free_b) / 1e9, 1), probe="largest chunk's first %d targets alone, plain path" % (u.hi - u.lo), chunk_sites=u.L, targets=u.hi - u.lo,
                              fb_state_sites_per_launch=st["fb_state_sites"] / st["fb_launches"],
                              stage_ms={k: round(st[k], 1) for k in ("ms_select", "ms_condset", "ms_fb", "ms_phase")},
                              **rf)), flush=True)
        return

    paths = ["learned", "plain"] if a.path == "both" else [a.path]
    net, train = (None, None)
    if "learned" in paths:
        net, train = train_selector()
        log("selector trained: %s" % json.dumps(train))
    # workspace sizing: each worker executes its largest unit once per path (outputs discarded)
    big = sorted(units, key=lambda u: -u.work)[:n_workers]
    for p in paths:
        set_path(big, p, net, a.keep, a.min_mass, chunks, a.sel_mode)
        run_units(big, workspaces, "workspace sizing (%s)" % p)

    res = {}
    small = min(range(len(chunks)), key=lambda c: chunks[c]["L"])   # the large-K judge's chunk
    for p in paths:
        set_path(units, p, net, a.keep, a.min_mass, chunks, a.sel_mode)
        for i in range(a.warmup):
            run_units(units, workspaces, "%s warmup step %d" % (p, i))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(a.steps):
            run_units(units, workspaces, "%s step %d" % (p, i))
        torch.cuda.synchronize()
        dt = (time.perf_counter() - t0) / max(1, a.steps)
        outs = gather(units, chunks)
        stats = [s for o in outs for s in o["stats"]]
        rf = fb_roofline(stats)
        res[p] = dict(value=tsites / dt, ms_per_step=dt * 1e3, accuracy=accuracy(outs, chunks),
                      per_chunk=per_chunk(units, chunks), step_fb_bytes=rf["fb_bytes"],
                      step_level_frac=rf["fb_bytes"] / dt / 1e9 / HBM_PEAK_GBS,
                      mean_states=float(np.mean(np.concatenate([o["nstates"][:, 1:].ravel() for o in outs]))),
                      stage_ms_summed=({k: round(sum(s[k] for s in stats), 1)
                                        for k in ("ms_select", "ms_learn", "ms_condset", "ms_fb", "ms_phase",
                                                  "ms_graph", "ms_other")}),
                      graph=dict(replayed_iterations=int(sum(s["graph_iters"] for s in stats)),
                                 iterations=int(sum(o["nstates"].shape[1] for o in outs for _ in o["stats"])),
                                 reruns=int(sum(s["graph_reruns"] for s in stats))),
                      ds=np.concatenate([dosage(o["gp0"], o["gp1"]).ravel() for o in outs]),
                      ds_small=dosage(outs[small]["gp0"], outs[small]["gp1"]))
        if p == "learned":
            res[p]["target_iterations_learned"] = sum(s["learned_sel"] for s in stats)
            res[p]["target_iterations_pool_kept"] = sum(s["pool_kept"] for s in stats)
        log("%s: %.3fM target-site imputations/s, r2 %.5f, NRC %.5f" % (
            p, res[p]["value"] / 1e6, res[p]["accuracy"]["dosage_r2"], res[p]["accuracy"]["nrc"]))
    free_b, total_b = torch.cuda.mem_get_info()
    hbm_used_gb = (total_b - free_b) / 1e9

    judge = None
    if not a.no_judge:
        judge = large_k_judge(chunks, panels, workspaces, {("learned" if p == "learned" else "plain_pbwt"): res[p]
                                                            for p in paths}, units)
        log("large-K judge: %s" % json.dumps(judge))

    # roofline probe: the largest chunk's first targets alone on the plain path,
    # after the step's runs and workspaces have released their device memory
    for x in units:
        x.run.close()
    for w in workspaces:
        w.close()
    u = probe_unit(chunks, panels, a.probe_targets)
    set_path([u], "plain", None, 0, 0, chunks)
    u.run.set_graph(0)
    probe_ws = hip.Workspace()
    u.run.set_workspace(probe_ws)
    u.run.execute()
    probe = fb_roofline([u.run.stats()])

    main_path = paths[0]
    if len(paths) == 2 and abs(res["learned"]["accuracy"]["dosage_r2"] - res["plain"]["accuracy"]["dosage_r2"]) > 0.001:
        main_path = "plain"
    M = res[main_path]
    wl = a.workload + ("-quick" if a.quick else "")
    traffic = profile_value("fb_kernel_%s_probe_traffic.md" % wl, "traffic_bytes_per_launch")
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
        "config": dict(workload=desc, path=main_path, units_in_flight=n_workers, **meta),
        "roofline": {"bound": "hbm", "achieved": probe["achieved"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": probe["achieved"] / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "fb_kernel (haploid forward-backward), largest chunk's first %d targets alone" % (u.hi - u.lo),
                     "bytes_per_launch": probe["bytes_per_launch"], "avg_launch_ms": probe["avg_launch_ms"],
                     "step_level_frac": M["step_level_frac"],
                     "valu_issue_frac": profile_value("fb_kernel_%s_probe_valu.md" % wl, "valu_issue_frac"),
                     "valu_lane_insts_per_state_site": profile_value("fb_kernel_%s_probe_valu.md" % wl,
                                                                     "valu_lane_insts_per_state_site")},
        "accuracy": M["accuracy"],
        "per_chunk": M["per_chunk"],
        "large_k_judge": judge,
        "setup_s": round(setup_s, 1),
        "hbm_used_gb": round(hbm_used_gb, 1),
    }
    for p in paths:
        R = res[p]
        d = {k: R[k] for k in ("value", "ms_per_step", "accuracy", "mean_states", "step_level_frac",
                               "stage_ms_summed", "graph")}
        if p == "learned":
            d.update(keep=a.keep, min_mass=a.min_mass, training=train,
                     target_iterations_learned=R["target_iterations_learned"],
                     target_iterations_pool_kept=R["target_iterations_pool_kept"])
        line["learned" if p == "learned" else "plain_pbwt"] = d
    if len(paths) == 2:
        line["learned"]["mean_abs_ds_vs_plain"] = float(np.mean(np.abs(res["learned"]["ds"] - res["plain"]["ds"])))
        line["learned"]["r2_minus_plain"] = (res["learned"]["accuracy"]["dosage_r2"] -
                                             res["plain"]["accuracy"]["dosage_r2"])
        line["learned"]["speedup_vs_plain"] = res["learned"]["value"] / res["plain"]["value"]
    if not a.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline(chunks)
        log("cpu baseline: %s" % json.dumps(line["cpu_baseline"]))
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
