"""Iterations replayed from a captured HIP graph (gh_run_set_graph 1, the
default) against the same runs with every iteration launched from the host
(mode 0).  The replayed iterations make every per-iteration decision on the
device (stored PBWT sites and start indices, placement of Hvar rows,
checkpoints, loci and phasing messages, launch groups and shapes), so every
output must be bit-identical: GP sums, best-guess haplotypes, conditioning-set
sizes of every iteration, and the run's counters.  The panel carries rare and
low-quality sites; the cases cover one set-size class, several classes and a
full-panel target (two launch groups on two streams), a handful of targets
(latency shapes), the latency mode, the learned selector, a second execution
of the same run, and the forced repeat when the captured buffers are too
small."""
import numpy as np
import pytest

from glimpse_amd import synth

pytestmark = pytest.mark.gpu

T = 40


@pytest.fixture(scope="module")
def case():
    from glimpse_amd import hip
    P = synth.make_panel(2000, 4000, T, length_mb=0.8, rare_frac=0.3, flip_frac=0.05, seed=41)
    P["lq"] = (np.arange(P["L"]) % 17 == 3).astype(np.uint8)
    gl, flat = synth.make_reads(P["tgt_haps"], 1.0, seed=42)
    return P, gl, flat, hip.Panel(P)


def _params(**kw):
    from glimpse_amd import hip
    base = dict(burnin=3, main_iter=4, kpbwt=1000, kinit=600, seed=5)
    base.update(kw)
    return hip.default_params(**base)


def _exec(panel, gl, flat, params, mode, kp=None, latency=False, net=None, k_learn=None, repeat=1):
    from glimpse_amd import hip
    r = hip.Run(panel, gl, flat, params, kpbwt_per_target=kp)
    r.set_graph(mode)
    if latency:
        r.set_tuning("latency")
    if net is not None:
        r.set_selector(net, k_learn)
    outs = []
    for _ in range(repeat):
        r.execute()
        outs.append(r.fetch())
    r.close()
    return outs if repeat > 1 else outs[0]


COUNTERS = ("fb_state_sites", "fb_bytes", "fb_poly_sites", "ph_state_sites", "ph_segments", "pbwt_hap_sites")


def _same(a, b):
    for k in ("gp0", "gp1", "hds", "h0", "h1", "nstates"):
        assert np.array_equal(a[k], b[k]), "%s differs (%d entries)" % (k, int(np.sum(a[k] != b[k])))
    for k in COUNTERS:
        assert a["stats"][k] == b["stats"][k], "counter %s: %r vs %r" % (k, a["stats"][k], b["stats"][k])


@pytest.mark.parametrize("kind", ["one_class", "mixed", "few", "latency"])
def test_graph_matches_host_launched(case, kind):
    P, gl, flat, panel = case
    params = _params()
    n_iter = 1 + params.burnin + params.main_iter
    kp, lat = None, kind == "latency"
    if kind == "mixed":
        # 2, 5, 10, 16, 24 and 32 words (a full-panel target): several forward-backward and phasing classes
        kp = np.resize(np.array([100, 300, 600, 1000, 1500, 2000], np.int32), T)
    if kind == "few":
        gl, flat = gl[:6], flat[:6]
    g = _exec(panel, gl, flat, params, 1, kp=kp, latency=lat)
    h = _exec(panel, gl, flat, params, 0, kp=kp, latency=lat)
    print("[graph] %s: replayed %d of %d iterations, %.1f ms (host-launched run %.1f ms)" % (
        kind, g["stats"]["graph_iters"], n_iter, g["stats"]["ms_total"], h["stats"]["ms_total"]))
    assert g["stats"]["graph_iters"] == n_iter - 2 and g["stats"]["graph_reruns"] == 0
    assert h["stats"]["graph_iters"] == 0
    _same(g, h)


def test_graph_learned_selector_matches_host_launched(case):
    """The learned selection (features, network, top-k with the fallback) in
    the replayed iterations; an untrained network still ranks the pools."""
    from glimpse_amd import hip
    P, gl, flat, panel = case
    net = hip.Mlp(seed=7)
    kl = np.full(T, 400, np.int32)
    params = _params()
    g = _exec(panel, gl, flat, params, 1, net=net, k_learn=kl)
    h = _exec(panel, gl, flat, params, 0, net=net, k_learn=kl)
    assert g["stats"]["graph_iters"] > 0
    _same(g, h)
    assert g["stats"]["learned_sel"] == h["stats"]["learned_sel"]
    assert g["stats"]["pool_kept"] == h["stats"]["pool_kept"]


def test_graph_second_execution_is_identical(case):
    P, gl, flat, panel = case
    a, b = _exec(panel, gl, flat, _params(seed=8), 1, repeat=2)
    _same(a, b)
    assert b["stats"]["graph_reruns"] == 0


def test_graph_changed_launch_groups_repeat_without_graph(case, monkeypatch):
    """A graph captured with other launch shapes than the replays' groups need
    (forced: group A captured with the other latency shape): the first replay
    places nothing and the execution is repeated host-launched, with the
    host-launched run's results."""
    P, gl, flat, panel = case
    params = _params(seed=10)
    monkeypatch.setenv("GH_GRAPH_FLIP_LATENCY", "1")
    g = _exec(panel, gl, flat, params, 1)
    monkeypatch.delenv("GH_GRAPH_FLIP_LATENCY")
    h = _exec(panel, gl, flat, params, 0)
    assert g["stats"]["graph_reruns"] == 2 and g["stats"]["graph_iters"] == 0
    _same(g, h)


def test_device_driven_stream_matches_host_launched(case):
    """Mode 2: the device-decided iterations enqueued without a capture."""
    P, gl, flat, panel = case
    params = _params(seed=11)
    kp = np.resize(np.array([100, 600, 1500], np.int32), T)
    g = _exec(panel, gl, flat, params, 2, kp=kp)
    h = _exec(panel, gl, flat, params, 0, kp=kp)
    assert g["stats"]["graph_iters"] == 1 + params.burnin + params.main_iter - 2
    _same(g, h)


def test_graph_outgrown_buffers_repeat_without_graph(case, monkeypatch):
    """Captured buffers capped at 5% of what the host-launched iterations
    needed: the first replay's placement overflows, nothing is placed past
    the caps, and the execution is repeated with every iteration launched
    from the host; the results are the host-launched run's."""
    P, gl, flat, panel = case
    params = _params(seed=9)
    monkeypatch.setenv("GH_GRAPH_HEADROOM", "0.05")
    g = _exec(panel, gl, flat, params, 1)
    monkeypatch.delenv("GH_GRAPH_HEADROOM")
    h = _exec(panel, gl, flat, params, 0)
    assert g["stats"]["graph_reruns"] == 1 and g["stats"]["graph_iters"] == 0
    _same(g, h)
