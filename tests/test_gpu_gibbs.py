"""End-to-end GPU Gibbs runs (gh_phase_run) against the oracle run in counter
RNG mode (the same Philox draws).  Short runs must agree nearly exactly; long
runs agree statistically (the chains may split after a draw lands within
float rounding of a posterior).  The panels carry low-quality sites (no
emission, excluded from the PBWT, flat hets in phasing)."""
import numpy as np
import pytest

import orc
from conftest import dosage_r2, nrc
from glimpse_amd import synth

pytestmark = pytest.mark.gpu


def _with_lq(P, every, phase):
    P["lq"] = (np.arange(P["L"]) % every == phase).astype(np.uint8)
    return P


@pytest.fixture(scope="module")
def case():
    from glimpse_amd import hip
    P = _with_lq(synth.make_panel(1000, 3000, 8, length_mb=0.6, seed=21), 11, 5)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=4)
    return P, gl, flat, orc.Panel(P), hip.Panel(P)


@pytest.mark.parametrize("kpbwt", [300, 1000])
def test_short_run_matches_oracle(case, kpbwt):
    from glimpse_amd import hip
    P, gl, flat, OP, GP = case
    kw = dict(burnin=0, main_iter=1, kpbwt=kpbwt, kinit=600, seed=99)
    ref = OP.phase_run(gl, flat, orc.default_params(rng_mode=orc.RNG_COUNTER, **kw))
    got = GP.phase_run(gl, flat, hip.default_params(**kw))
    assert np.array_equal(got["nstates"], ref["nstates"]), "conditioning-set sizes differ"
    d = np.abs(got["gp0"] - ref["gp0"]) + np.abs(got["gp1"] - ref["gp1"])
    assert np.mean(d < 1e-4) > 0.995, "GP differs at %.4f of sites" % np.mean(d >= 1e-4)


def test_long_run_statistical_parity():
    """32 targets x 3000 sites, INIT + 3 burn-in + 5 main iterations: dosage r2
    and NRC against the truth within 0.001 of the oracle's run."""
    from glimpse_amd import hip
    P = _with_lq(synth.make_panel(1000, 3000, 32, length_mb=0.6, seed=23), 13, 2)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=6)
    kw = dict(burnin=3, main_iter=5, kpbwt=300, kinit=600, seed=7)
    ref = orc.Panel(P).phase_run(gl, flat, orc.default_params(rng_mode=orc.RNG_COUNTER, **kw))
    got = hip.Panel(P).phase_run(gl, flat, hip.default_params(**kw))
    truth = synth.truth_genotypes(P["tgt_haps"])
    a, b = dosage_r2(got["gp0"], got["gp1"], truth), dosage_r2(ref["gp0"], ref["gp1"], truth)
    assert abs(a - b) <= 0.001, "dosage r2 GPU %.5f vs oracle %.5f" % (a, b)
    gg = got["h0"].astype(int) + got["h1"]
    gr = ref["h0"].astype(int) + ref["h1"]
    na, nb = nrc(gg, truth), nrc(gr, truth)
    assert abs(na - nb) <= 0.001, "NRC GPU %.5f vs oracle %.5f" % (na, nb)
    assert np.mean(gg == gr) > 0.98
    assert got["stats"]["fb_state_sites"] > 0


@pytest.mark.parametrize("mode", [0, 1, 2])
def test_mixed_k_run_matches_oracle(case, mode):
    """Targets whose PBWT K spans three set-size classes.  With shape modes 0
    (throughput defaults) and 2 (latency defaults) every forward-backward and
    phasing pass splits into a launch for the largest class and one beside it
    on a second stream; mode 1 (autotuning) keeps one launch per pass.  All
    must give the oracle's run."""
    from glimpse_amd import hip
    P, gl, flat, OP, GP = case
    kpt = np.array([100, 250, 600, 950, 100, 400, 1000, 200], np.int32)
    kw = dict(burnin=0, main_iter=1, kpbwt=int(kpt.max()), kinit=600, seed=99)
    ref = OP.phase_run(gl, flat, orc.default_params(rng_mode=orc.RNG_COUNTER, **kw), kpbwt_per_target=kpt)
    run = hip.Run(GP, gl, flat, hip.default_params(**kw), kpbwt_per_target=kpt)
    run.set_tuning(mode)
    run.execute()
    got = run.fetch()
    run.close()
    assert np.array_equal(got["nstates"], ref["nstates"]), "conditioning-set sizes differ"
    d = np.abs(got["gp0"] - ref["gp0"]) + np.abs(got["gp1"] - ref["gp1"])
    assert np.mean(d < 1e-4) > 0.995, "GP differs at %.4f of sites" % np.mean(d >= 1e-4)


def test_runs_sharing_a_workspace(case):
    """Runs executed one after another in one shared workspace (a larger run on
    another panel first, then the case's run, then the larger one again) give
    exactly the results of runs with their own workspaces."""
    from glimpse_amd import hip
    P, gl, flat, OP, GP = case
    kw = dict(burnin=1, main_iter=2, kpbwt=300, kinit=600, seed=5)
    P2 = _with_lq(synth.make_panel(1500, 4000, 10, length_mb=0.8, seed=33), 7, 3)
    gl2, flat2 = synth.make_reads(P2["tgt_haps"], 1.5, seed=34)
    GP2 = hip.Panel(P2)
    own1 = GP.phase_run(gl, flat, hip.default_params(**kw))
    own2 = GP2.phase_run(gl2, flat2, hip.default_params(**kw))
    ws = hip.Workspace()
    runs = [(hip.Run(GP2, gl2, flat2, hip.default_params(**kw)), own2),
            (hip.Run(GP, gl, flat, hip.default_params(**kw)), own1)]
    for r, _ in runs:
        r.set_workspace(ws)
    for r, own in runs + runs[:1]:
        r.execute()
        got = r.fetch()
        for k in ("gp0", "gp1", "h0", "h1", "nstates"):
            assert np.array_equal(got[k], own[k]), k
    assert ws.bytes > 0
    for r, _ in runs:
        r.close()


def test_run_with_sets_above_2048_states():
    """A Gibbs run whose PBWT sets reach 3,000 states (and a full-panel target of
    3,500 haplotypes) against the counter-mode oracle."""
    from glimpse_amd import hip
    P = synth.make_panel(3500, 2500, 3, length_mb=0.5, seed=41)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=42)
    kpt = np.array([3000, 1200, 3500], np.int32)
    kw = dict(burnin=0, main_iter=1, kpbwt=3500, kinit=1000, seed=43)
    ref = orc.Panel(P).phase_run(gl, flat, orc.default_params(rng_mode=orc.RNG_COUNTER, **kw), kpbwt_per_target=kpt)
    got = hip.Panel(P).phase_run(gl, flat, hip.default_params(**kw), kpbwt_per_target=kpt)
    assert np.array_equal(got["nstates"], ref["nstates"]), "conditioning-set sizes differ"
    assert got["nstates"][2, 1] == 3500
    d = np.abs(got["gp0"] - ref["gp0"]) + np.abs(got["gp1"] - ref["gp1"])
    assert np.mean(d < 1e-4) > 0.995, "GP differs at %.4f of sites" % np.mean(d >= 1e-4)
