"""GPU parity: the INIT iteration's state selection on the device
(gh_init_states; haplotype_set::initRareTar + performSelection_RARE_INIT_GL,
haplotype_set.cpp:108-169, 775-833) against the oracle's restatement with the
same counter draws: identical sets, target by target.  The oracle's selection
is pinned in test_oracle_pins.py (uniform subsets, the rare-evidence rule).

Cases: rare-heavy panels with targets at low and high coverage (more
rare-evidence sites than 0.8 Kinit: the minor-allele-count order; fewer:
floor(0.8 Kinit / sites) carriers per site), all-flat targets (the fill alone),
low-quality sites, Kinit above the panel size, and a panel of 20,000
haplotypes."""
import numpy as np
import pytest

import orc
from glimpse_amd import synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("N,L,T,cov,kinit,seed", [
    (2000, 3000, 12, 1.0, 600, 3),     # few rare-evidence sites per target
    (2000, 3000, 12, 8.0, 100, 5),     # more than 0.8 Kinit sites: sorted by minor-allele count
    (1500, 2500, 8, 0.3, 2000, 7),     # Kinit clipped to N
    (20000, 4000, 6, 4.0, 1000, 9),    # a larger panel (fill over 20,000 haplotypes)
])
def test_init_sets_match_oracle(N, L, T, cov, kinit, seed):
    from glimpse_amd import hip
    P = synth.make_panel(N, L, T, length_mb=L / 4000.0, rare_frac=0.5, flip_frac=0.05, seed=seed, fast=N > 5000)
    P["lq"] = (np.arange(P["L"]) % 13 == 4).astype(np.uint8)
    gl, flat = synth.make_reads(P["tgt_haps"], cov, seed=seed + 1)
    flat[0] = 1   # an all-flat target: the fill alone
    OP, GP = orc.Panel(P), hip.Panel(P)
    got = GP.init_states(gl, flat, kinit, seed=1234 + seed)
    k = min(kinit, N)
    for t in range(T):
        ref = OP.init_states(gl[t], flat[t], kinit, orc.RNG_COUNTER, 1234 + seed, tgt=t)
        assert len(got[t]) == len(ref) == k, "target %d: %d vs %d states" % (t, len(got[t]), len(ref))
        assert np.array_equal(got[t], ref), "target %d: sets differ" % t
        assert np.all(np.diff(got[t]) > 0)
    print("[init] N=%d L=%d T=%d cov=%.1f Kinit=%d: %d targets identical" % (N, L, T, cov, kinit, T))
