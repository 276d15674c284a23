"""Output fields (genotype_writer::writeGenotypes, phase/src/io/genotype_writer.cpp:109-190):
the oracle pinned on closed forms, and the GPU kernel bit-exact against it."""
import numpy as np
import pytest

import orc


def test_oracle_closed_forms():
    gp0 = np.array([[1.0, 0.3334, 0.1, 0.0, 0.2]], np.float32)
    gp1 = np.array([[0.0, 0.3333, 0.8, 0.0, 0.3]], np.float32)
    hds = np.array([[0, 0, 1, 0, 0]], np.uint8)
    o = orc.output_fields(gp0, gp1, hds, np.array([90, 50, 99, 0, 10], np.uint32), np.array([10, 50, 1, 100, 10],
                                                                                              np.uint32))
    # RR call: DS 0, GP (1,0,0), GT 0|0
    assert o["ds"][0, 0] == 0 and list(o["gp"][0, 0]) == [1, 0, 0] and list(o["gt"][0, 0]) == [0, 0]
    # het: the alt allele goes on haplotype hds
    assert list(o["gt"][0, 2]) == [0, 1] and abs(o["ds"][0, 2] - 1.0) < 1e-6   # 0.8 + 2 * 0.1
    # AA call
    assert list(o["gt"][0, 3]) == [1, 1] and o["ds"][0, 3] == 2 and list(o["gp"][0, 3]) == [0, 0, 1]
    # RAF = calt / (calt + cref); one target: AF = (unrounded) DS / 2
    assert np.allclose(o["raf"], [0.1, 0.5, 0.01, 1.0, 0.5]) and np.allclose(o["af"], o["ds"][0] / 2, atol=6e-4)
    # INFO = 1 when every dosage is a hard call
    hard = orc.output_fields(np.array([[1, 0], [0, 0], [0, 1]], np.float32), np.array([[0, 1], [1, 0], [0, 0]],
                                                                                         np.float32),
                             np.zeros((3, 2), np.uint8), np.array([5, 5], np.uint32), np.array([5, 5], np.uint32))
    assert np.all(hard["info"] == 1.0)
    # rounding invariants on random posteriors: every GP is within 1e-3 of its value and on the 1e-3 grid,
    # the three sum to >= 0.9999, and DS is round(gp1 + 2 gp2, 1e-3)
    rng = np.random.default_rng(3)
    a = rng.dirichlet([0.5, 0.5, 0.5], (50, 200)).astype(np.float32)
    r = orc.output_fields(a[..., 0], a[..., 1], np.zeros((50, 200), np.uint8), np.ones(200, np.uint32),
                          np.ones(200, np.uint32))
    gp2 = np.clip(1 - a[..., 1] - a[..., 0], 0, 1)
    truth = np.stack([a[..., 0], a[..., 1], gp2], -1).astype(np.float64)
    # GP0, GP1 are floored (and maybe raised once); GP2 is floored from 1 - GP0 - GP1, so it can sit up to
    # 2e-3 above gp2 before a raise
    assert np.all(np.abs(r["gp"][..., :2] - truth[..., :2]) <= 1.0011e-3)
    assert np.all(np.abs(r["gp"][..., 2] - truth[..., 2]) <= 2.0011e-3)
    assert np.all(np.abs(r["gp"] * 1000 - np.round(r["gp"] * 1000)) < 1e-3)
    tot = r["gp"].astype(np.float32).sum(-1)
    assert np.all(tot >= 0.9999 - 1e-6) and np.all(tot <= 1.002)
    assert np.all(np.abs(r["ds"] - np.round((a[..., 1] + 2 * gp2).astype(np.float64), 3)) <= 1.0001e-3)
    # and below 1 when dosages are uncertain: 1 - (sum(gp1 + 4 gp2) - sum(ds^2)) / (2T f (1-f))
    g0 = np.array([[0.5], [0.2]], np.float32)
    g1 = np.array([[0.4], [0.5]], np.float32)
    u = orc.output_fields(g0, g1, np.zeros((2, 1), np.uint8), np.array([1], np.uint32), np.array([1], np.uint32))
    g2 = 1 - g0 - g1
    ds = g1 + 2 * g2
    f = ds.sum() / 4
    want = 1 - ((g1 + 4 * g2).sum() - (ds ** 2).sum()) / (4 * f * (1 - f))
    assert abs(u["info"][0] - round(want, 3)) < 1e-6


@pytest.mark.gpu
def test_gpu_output_fields_bit_exact():
    from glimpse_amd import hip
    rng = np.random.default_rng(9)
    T, L = 37, 3001
    a = rng.dirichlet([0.3, 0.3, 0.3], (T, L)).astype(np.float32)
    gp0, gp1 = a[..., 0].copy(), a[..., 1].copy()
    gp0[:, :50], gp1[:, :50] = 1.0, 0.0                      # not stored
    gp0[:, 50:100] = np.round(gp0[:, 50:100], 3)             # remainders on the 1e-3 grid
    gp1[:, 50:100] = np.round(np.minimum(gp1[:, 50:100], 1 - gp0[:, 50:100]), 3)
    gp0[:, 100:150], gp1[:, 100:150] = 1 / 3, 1 / 3          # ties
    hds = rng.integers(0, 2, (T, L)).astype(np.uint8)
    cref = rng.integers(0, 5000, L).astype(np.uint32)
    calt = (5000 - cref).astype(np.uint32)
    calt[:3] = 0
    got = hip.output_fields(gp0, gp1, hds, cref, calt)
    ref = orc.output_fields(gp0, gp1, hds, cref, calt)
    for k in ("ds", "gp", "gt", "info", "af", "raf"):
        assert np.array_equal(got[k].view(np.uint8), ref[k].view(np.uint8)), k
