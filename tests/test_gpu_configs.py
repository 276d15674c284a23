"""The workload configurations at their full sizes, on the GPU, checked
against the oracle (bit-exact state sets, forward-backward within 1e-4,
Gibbs runs draw for draw) and through size-independent properties.

  C1  1,000 haplotypes x 10,000 SNPs, 1 target at 1x, Kpbwt 500, 1+5+15 iterations
  C2  one full-size chunk of the batched workload: 10,000 haplotypes, 60,000
      sites, 200 targets at U(0.1, 2)x, Kpbwt U[500, 2000]
  C3  100,000 haplotypes x 150,000 sites, about 80% of sites at MAF < 0.1%,
      a chunk of more than 21 cM so PBWT selection runs every 0.5 cM, K 1,000
Each test prints the figures it checks."""
import time

import numpy as np
import pytest

import orc
from conftest import dosage_r2, nrc
from glimpse_amd import synth

pytestmark = pytest.mark.gpu


def _stored(OP, depth, modulo, kpbwt, rng):
    eff, grp = OP.groups(depth, modulo, kpbwt)
    st = np.zeros(OP.n_common_hq, np.uint8)
    lo = 0
    for g in grp:
        st[rng.integers(lo, g)] = 1
        lo = g
    return eff, st


@pytest.fixture(scope="module")
def c3():
    """C3 panel: 100,000 haplotypes x 150,000 sites over about 22 cM, 75% of
    candidate sites injected rare (the 1/x spectrum adds more), 4 targets at 1x."""
    t0 = time.time()
    P = synth.make_panel(100000, 150000, 4, length_mb=36.0, rare_frac=0.75, seed=303, fast=True)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=304, threads=0)
    print("[c3] panel %.1f s: %d sites over %.2f cM, common %.3f, MAF < 0.1%%: %.3f" % (
        time.time() - t0, P["L"], P["cm"][-1], P["common"].mean(),
        np.mean(np.minimum(P["cref"], P["calt"]) / P["N"] < 0.001)), flush=True)
    return P, gl, flat


@pytest.mark.timeout(900)
def test_c3_full_size_selection_and_posteriors(c3):
    from glimpse_amd import hip
    P, gl, flat = c3
    t0 = time.time()
    OP = orc.Panel(P)
    t1 = time.time()
    GP = hip.Panel(P, with_pbwt=True)
    t2 = time.time()
    info = GP.info()
    print("[c3] oracle panel %.1f s, GPU panel %.1f s (%.2f GB on the device)" % (
        t1 - t0, t2 - t1, info["device_bytes"] / 1e9), flush=True)
    assert P["cm"][-1] >= 21.0, "chunk shorter than 21 cM"
    assert np.mean(np.minimum(P["cref"], P["calt"]) / P["N"] < 0.001) >= 0.75
    depth, kpbwt = 12, 1000
    rng = np.random.default_rng(305)
    eff, stored = _stored(OP, depth, 0.5, kpbwt, rng)
    assert eff == pytest.approx(0.5), "PBWT modulo adjusted to %g" % eff
    haps = (P["tgt_haps"] ^ (rng.random(P["tgt_haps"].shape) < 0.01)).astype(np.uint8)
    start = rng.integers(0, P["N"], haps.shape[0]).astype(np.int32)
    t0 = time.time()
    ref = OP.pbwt_select(haps, stored, start, depth=depth, modulo=eff)
    t1 = time.time()
    got, states = GP.pbwt_select(haps, stored, start, depth=depth, modulo=eff, kpbwt=kpbwt)
    t2 = time.time()
    T = haps.shape[0] // 2
    sets = []
    for t in range(T):
        for d in range(depth):
            assert np.array_equal(got[t][d], ref[t][d]), "target %d rank %d list differs" % (t, d)
        want = OP.condset(ref[t], kpbwt)["states"]
        assert np.array_equal(states[t], want), "target %d compacted state set differs" % t
        assert len(want) == kpbwt
        sets.append(want)
    print("[c3] PBWT selection: oracle %.1f s, GPU call %.2f s; %d targets bit-exact" % (t1 - t0, t2 - t1, T),
          flush=True)
    # forward-backward of each target's two haplotypes on its selected set
    HLs, fl, items = [], [], []
    for t in range(T):
        for j in range(2):
            other = P["tgt_haps"][2 * t + 1 - j] if j else None
            HLs.append(orc.hap_likelihoods(gl[t], flat[t], other))
            fl.append(flat[t])
            items.append(sets[t])
    HP = GP.impute_batch(items, np.stack(HLs), np.stack(fl))
    worst = 0.0
    for i, s in enumerate(items):
        r = OP.impute_hap(s, HLs[i], fl[i])
        worst = max(worst, float(np.abs(HP[i] - r).max()))
    print("[c3] forward-backward at K=%d on %d haplotypes: max |HP - oracle| = %.2e" % (kpbwt, len(items), worst))
    assert worst <= 1e-4


# ------------------------------------------------------------------ C1
@pytest.fixture(scope="module")
def c1():
    """C1 panel: 1,000 haplotypes x 10,000 SNPs over 2 Mb (2 cM), 32 held-out
    targets at 1x (target 0 is the configuration's one target; all 32 resolve
    accuracy differences of 0.001)."""
    P = synth.make_panel(1000, 10000, 32, length_mb=2.0, seed=101)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=102, threads=0)
    return P, gl, flat


def test_c1_exact_run_matches_oracle(c1):
    """C1 as configured: one target, Kpbwt 500 (modulo auto-halved to 0.05 cM),
    Kinit min(1000, N), the default 1 + 5 + 15 schedule and seed, against the
    oracle with the same counter draws."""
    from glimpse_amd import hip
    P, gl, flat = c1
    kw = dict(kpbwt=500)
    ref = orc.Panel(P).phase_run(gl[:1], flat[:1], orc.default_params(rng_mode=orc.RNG_COUNTER, **kw))
    got = hip.Panel(P).phase_run(gl[:1], flat[:1], hip.default_params(**kw))
    truth = synth.truth_genotypes(P["tgt_haps"][:2])
    same_k = int(np.sum(got["nstates"] == ref["nstates"]))
    d = np.abs(got["gp0"] - ref["gp0"]) + np.abs(got["gp1"] - ref["gp1"])
    agree = float(np.mean(d < 1e-4))
    a, b = dosage_r2(got["gp0"], got["gp1"], truth), dosage_r2(ref["gp0"], ref["gp1"], truth)
    na = nrc(got["h0"].astype(int) + got["h1"], truth)
    nb = nrc(ref["h0"].astype(int) + ref["h1"], truth)
    print("[c1] nstates equal in %d of %d iterations; GP within 1e-4 at %.4f of sites; r2 %.5f vs %.5f; "
          "NRC %.5f vs %.5f" % (same_k, got["nstates"].size, agree, a, b, na, nb))
    assert got["nstates"][0, 0] == 1000
    assert same_k == got["nstates"].size
    assert agree >= 0.995
    assert abs(a - b) <= 0.001 and abs(na - nb) <= 0.001


@pytest.mark.timeout(600)
def test_c1_accuracy_matches_reference_mode_oracle(c1):
    """The accuracy check against the reference's own random streams: 32 C1
    runs (one target each, seed 15052011 + t) of the oracle with std::mt19937
    and libstdc++'s distributions as the reference draws them, against the GPU
    run of the same 32 targets (counter draws).  Pooled dosage r2 and NRC
    against the truth must agree within 0.001."""
    import threading
    from glimpse_amd import hip
    P, gl, flat = c1
    T = gl.shape[0]
    got = hip.Panel(P).phase_run(gl, flat, hip.default_params(kpbwt=500))
    OP = orc.Panel(P)
    outs = [None] * T

    def one(t):
        outs[t] = OP.phase_run(gl[t:t + 1], flat[t:t + 1],
                               orc.default_params(rng_mode=orc.RNG_REFERENCE, kpbwt=500, seed=15052011 + t))
    t0 = time.time()
    for lo in range(0, T, 16):
        th = [threading.Thread(target=one, args=(t,)) for t in range(lo, min(T, lo + 16))]
        for x in th:
            x.start()
        for x in th:
            x.join()
    ref = {k: np.concatenate([o[k] for o in outs]) for k in ("gp0", "gp1", "h0", "h1")}
    truth = synth.truth_genotypes(P["tgt_haps"])
    a, b = dosage_r2(got["gp0"], got["gp1"], truth), dosage_r2(ref["gp0"], ref["gp1"], truth)
    na = nrc(got["h0"].astype(int) + got["h1"], truth)
    nb = nrc(ref["h0"].astype(int) + ref["h1"], truth)
    print("[c1] 32 targets: GPU r2 %.5f NRC %.5f; reference-mode oracle r2 %.5f NRC %.5f (oracle %.1f s)" % (
        a, na, b, nb, time.time() - t0))
    assert abs(a - b) <= 0.001, "dosage r2 GPU %.5f vs reference-mode oracle %.5f" % (a, b)
    assert abs(na - nb) <= 0.001, "NRC GPU %.5f vs reference-mode oracle %.5f" % (na, nb)


# ------------------------------------------------------------------ C2
def _check_run_properties(out, kp, kinit, truth, label):
    gp0, gp1, ns = out["gp0"], out["gp1"], out["nstates"]
    assert np.all((gp0 >= 0) & (gp0 <= 1) & (gp1 >= 0) & (gp1 <= 1)), label + ": GP outside [0, 1]"
    assert np.all(gp0 + gp1 <= 1.0 + 1e-5), label + ": GP sums above 1"
    assert np.all(ns[:, 0] == kinit), label + ": INIT set sizes"
    assert np.all((ns[:, 1:] > 0) & (ns[:, 1:] <= kp[:, None])), label + ": PBWT set sizes"
    g = out["h0"].astype(int) + out["h1"]
    assert np.all(out["h0"] <= out["h1"]), label + ": best-guess genotype encoding"
    r2 = dosage_r2(gp0, gp1, truth)
    return r2, nrc(g, truth)


@pytest.mark.timeout(600)
def test_c2_full_chunk():
    """One full-size chunk of the batched workload: 10,000 haplotypes x 60,000
    sites, 200 targets at U(0.1, 2)x, Kpbwt U[500, 2000], full schedule, every
    target in one run.  Properties of the whole run, then the first two
    targets against the oracle draw for draw (INIT + 1 + 1 iterations)."""
    from glimpse_amd import hip
    rng = np.random.default_rng(2002)
    T = 200
    P = synth.make_panel(10000, 60000, T, length_mb=12.0, seed=2003, fast=True)
    cov = rng.uniform(0.1, 2.0, T)
    kp = rng.integers(500, 2001, T).astype(np.int32)
    gl, flat = synth.make_reads(P["tgt_haps"], cov, seed=2004, threads=0)
    truth = synth.truth_genotypes(P["tgt_haps"])
    GP = hip.Panel(P)
    t0 = time.time()
    out = GP.phase_run(gl, flat, hip.default_params(), kpbwt_per_target=kp)
    dt = time.time() - t0
    r2, nc = _check_run_properties(out, kp, 1000, truth, "C2")
    print("[c2] 200 targets x 60,000 sites: %.2f s, %.2fM target-site imputations/s, r2 %.4f NRC %.4f" % (
        dt, T * P["L"] / dt / 1e6, r2, nc))
    assert r2 > 0.8
    # the first two targets against the oracle (their own run, same keys)
    kw = dict(burnin=1, main_iter=1, seed=99)
    sub = slice(0, 2)
    ref = orc.Panel(P).phase_run(gl[sub], flat[sub], orc.default_params(rng_mode=orc.RNG_COUNTER, **kw),
                                 kpbwt_per_target=kp[sub])
    got = GP.phase_run(gl[sub], flat[sub], hip.default_params(**kw), kpbwt_per_target=kp[sub])
    assert np.array_equal(got["nstates"], ref["nstates"]), "conditioning-set sizes differ"
    d = np.abs(got["gp0"] - ref["gp0"]) + np.abs(got["gp1"] - ref["gp1"])
    print("[c2] 2 targets vs oracle: GP within 1e-4 at %.5f of sites" % np.mean(d < 1e-4))
    assert np.mean(d < 1e-4) > 0.995


# ------------------------------------------------------------------ learned selector (C4, C5)
@pytest.fixture(scope="module")
def trained_net():
    """The selector network trained on the GPU on rows a plain run records:
    4,000 haplotypes x 10,000 SNPs (2 Mb), 48 targets at 1x, Kpbwt 2,000 (the
    large-K plain HMM's posterior masses are the labels)."""
    from glimpse_amd import hip
    P = synth.make_panel(4000, 10000, 48, length_mb=2.0, seed=31337)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=31338)
    panel = hip.Panel(P)
    run = hip.Run(panel, gl, flat, hip.default_params(kpbwt=2000))
    run.collect(1)
    run.execute()
    net = hip.Mlp(seed=31337)
    losses = net.train_run(run, epochs=8, batch=16384, lr=2e-3, seed=31337)
    print("[learn] %d rows, loss %.3f -> %.3f" % (int(run.stats()["collected_rows"]), losses[0], losses[-1]))
    run.close()
    return net


def _run(panel, gl, flat, params, kp=None, net=None, k_learn=None, min_mass=0.95, prior=0.0):
    from glimpse_amd import hip
    r = hip.Run(panel, gl, flat, params, kpbwt_per_target=kp)
    if net is not None:
        r.set_selector(net, k_learn, min_mass=min_mass)
        r.set_selector_prior(prior)
    t0 = time.time()
    r.execute()
    dt = time.time() - t0
    out = r.fetch()
    out["stats"] = r.stats()
    out["s"] = dt
    r.close()
    return out


@pytest.mark.timeout(600)
@pytest.mark.parametrize("seed", [4040, 5050, 6060])
def test_c4_learned_k200_against_pbwt(trained_net, seed):
    """C4 on held-out seeds: a 4,000-haplotype C1-shaped panel (the 1,000-haplotype
    C1 panel collapses K = 1,000 and 2,000 to the full panel), 24 targets at 1x.
    The network keeps 200 states (no fallback; groups served round robin,
    PBWT-order prior 8) of a Kpbwt = 1,000 pool walked at depth 48, which
    keeps the 0.1 cM selection interval (at depth 12 allocatePBWT halves the
    interval of a 1,000-state pool on a 2 cM chunk twice, haplotype_set.cpp:
    281-286, and the pool holds shallow neighbours at 80 positions).  Against
    plain PBWT at K = 1,000 (default depth 12): at matched dosage r2 (within
    0.001) or better with a fifth of the states.  Plain PBWT at K = 200 and at
    depth 48 are printed beside it."""
    from glimpse_amd import hip
    P = synth.make_panel(4000, 10000, 24, length_mb=2.0, seed=seed)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=seed + 1)
    truth = synth.truth_genotypes(P["tgt_haps"])
    panel = hip.Panel(P)
    T = gl.shape[0]
    big = _run(panel, gl, flat, hip.default_params(kpbwt=1000))
    small = _run(panel, gl, flat, hip.default_params(kpbwt=200))
    deep = _run(panel, gl, flat, hip.default_params(kpbwt=200, pbwt_depth=48))
    lrn = _run(panel, gl, flat, hip.default_params(kpbwt=1000, pbwt_depth=48), net=trained_net,
               k_learn=np.full(T, 200, np.int32), min_mass=0.0, prior=8.0)
    r_big, r_small, r_deep, r_lrn = (dosage_r2(o["gp0"], o["gp1"], truth) for o in (big, small, deep, lrn))
    ks = lrn["nstates"][:, 1:]
    print("[c4] seed %d r2: PBWT K=1000 %.5f (%.2f s), PBWT K=200 %.5f (%.2f s), PBWT K=200 depth 48 %.5f, "
          "learned K=200 of 1000 %.5f (%.2f s); learned minus K=1000: %+.5f; learned set sizes %d..%d" % (
              seed, r_big, big["s"], r_small, small["s"], r_deep, r_lrn, lrn["s"], r_lrn - r_big, ks.min(),
              ks.max()))
    assert ks.max() <= 200
    assert r_lrn >= r_big - 0.001, "learned K=200 r2 %.5f vs PBWT K=1000 %.5f" % (r_lrn, r_big)


@pytest.mark.timeout(900)
def test_c5_full_chunk_learned_with_fallback(trained_net):
    """One chunk of the mixed production batch at its smallest length: 100,000
    haplotypes x 40,000 sites, 1,000 targets at log-uniform 0.1-4x.  Plain PBWT
    (Kpbwt 1,000) and the learned selector (keeps 500, whole pool when the
    predicted mass is below 0.95) on the same resident inputs; properties of
    both runs, their accuracy, and two targets against the oracle."""
    from glimpse_amd import hip
    rng = np.random.default_rng(5005)
    T = 1000
    t0 = time.time()
    P = synth.make_panel(100000, 40000, T, length_mb=9.6, rare_frac=0.75, seed=5006, fast=True)
    cov = np.exp(rng.uniform(np.log(0.1), np.log(4.0), T))
    gl, flat = synth.make_reads(P["tgt_haps"], cov, seed=5007, threads=0)
    truth = synth.truth_genotypes(P["tgt_haps"])
    GP = hip.Panel(P)
    print("[c5] inputs and panel %.1f s: %d sites, %.2f cM, common %.3f" % (
        time.time() - t0, P["L"], P["cm"][-1], P["common"].mean()), flush=True)
    kp = np.full(T, 1000, np.int32)
    plain = _run(GP, gl, flat, hip.default_params(kpbwt=1000))
    lrn = _run(GP, gl, flat, hip.default_params(kpbwt=1000), net=trained_net, k_learn=np.full(T, 500, np.int32))
    rp = _check_run_properties(plain, kp, 1000, truth, "C5 plain")
    rl = _check_run_properties(lrn, kp, 1000, truth, "C5 learned")
    st = lrn["stats"]
    print("[c5] plain: %.2f s, %.2fM/s, r2 %.4f NRC %.4f; learned: %.2f s, %.2fM/s, r2 %.4f NRC %.4f; "
          "fallbacks %d of %d target-iterations" % (
              plain["s"], T * P["L"] / plain["s"] / 1e6, rp[0], rp[1], lrn["s"], T * P["L"] / lrn["s"] / 1e6,
              rl[0], rl[1], int(st["pool_kept"]), int(st["learned_sel"])))
    assert rp[0] > 0.8 and rl[0] > 0.8
    kw = dict(burnin=1, main_iter=1, kpbwt=1000, seed=77)
    sub = slice(0, 2)
    ref = orc.Panel(P).phase_run(gl[sub], flat[sub], orc.default_params(rng_mode=orc.RNG_COUNTER, **kw))
    got = GP.phase_run(gl[sub], flat[sub], hip.default_params(**kw))
    assert np.array_equal(got["nstates"], ref["nstates"]), "conditioning-set sizes differ"
    d = np.abs(got["gp0"] - ref["gp0"]) + np.abs(got["gp1"] - ref["gp1"])
    print("[c5] 2 targets vs oracle: GP within 1e-4 at %.5f of sites" % np.mean(d < 1e-4))
    assert np.mean(d < 1e-4) > 0.995
