"""Synthetic panels and low-coverage targets (ctypes front-end of csrc/synth.cpp)."""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_lib = None


class SynthCfg(ctypes.Structure):
    _fields_ = [
        ("n_haps", ctypes.c_int32), ("n_sites", ctypes.c_int32), ("n_targets", ctypes.c_int32),
        ("n_founders", ctypes.c_int32), ("length_mb", ctypes.c_double), ("start_bp", ctypes.c_int32),
        ("rho_per_cm", ctypes.c_double), ("mu", ctypes.c_double), ("rare_frac", ctypes.c_double),
        ("flip_frac", ctypes.c_double), ("sparse_maf", ctypes.c_float), ("keep_dense", ctypes.c_int32),
        ("seed", ctypes.c_uint64),
    ]


def lib():
    global _lib
    if _lib is None:
        path = os.path.join(_HERE, "libgh_synth.so")
        if not os.path.exists(path):
            raise RuntimeError("libgh_synth.so is not built: run __graft_entry__.build()")
        L = ctypes.CDLL(path)
        L.gh_synth_panel_new.restype = ctypes.c_void_p
        L.gh_synth_panel_new.argtypes = [ctypes.POINTER(SynthCfg)]
        L.gh_synth_panel_free.argtypes = [ctypes.c_void_p]
        L.gh_synth_error.restype = ctypes.c_char_p
        L.gh_synth_error.argtypes = [ctypes.c_void_p]
        L.gh_synth_panel_sizes.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 4
        L.gh_synth_panel_copy.argtypes = [ctypes.c_void_p] + [ctypes.c_void_p] * 11
        L.gh_synth_reads.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                     ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p]
        L.gh_synth_panel_fast.restype = ctypes.c_void_p
        L.gh_synth_panel_fast.argtypes = [ctypes.POINTER(SynthCfg), ctypes.c_int32]
        L.gh_synth_reads_mt.argtypes = [ctypes.c_int32, ctypes.c_int32, ctypes.c_void_p, ctypes.c_void_p,
                                        ctypes.c_uint64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int32]
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _threads(n):
    return int(n) if n else max(1, min(16, os.cpu_count() or 1))


def make_panel(n_haps, n_sites, n_targets, n_founders=50, length_mb=2.0, rho_per_cm=None,
               mu=1e-3, rare_frac=0.0, flip_frac=0.05, sparse_maf=0.001, keep_dense=False,
               seed=15052011, start_bp=1_000_000, ne=100000, fast=False, threads=0):
    """Returns a dict describing one chunk panel plus 2*n_targets held-out haplotypes.

    rho_per_cm defaults to 0.04*Ne/F switches per cM (the reference's own
    recombination scale with F founders).  fast=True uses the founder-segment
    generator (same model, per-site and per-haplotype random streams, threads
    host threads; no dense matrix), which large panels need."""
    if rho_per_cm is None:
        rho_per_cm = 0.04 * ne / max(1, n_founders)
    cfg = SynthCfg(n_haps, n_sites, n_targets, n_founders, float(length_mb), start_bp,
                   float(rho_per_cm), float(mu), float(rare_frac), float(flip_frac),
                   float(sparse_maf), int(bool(keep_dense)), int(seed))
    L_ = lib()
    h = L_.gh_synth_panel_fast(ctypes.byref(cfg), _threads(threads)) if fast else \
        L_.gh_synth_panel_new(ctypes.byref(cfg))
    try:
        err = L_.gh_synth_error(h).decode()
        if err:
            raise RuntimeError(err)
        Lr, Nr, nc = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int32()
        nr = ctypes.c_int64()
        L_.gh_synth_panel_sizes(h, ctypes.byref(Lr), ctypes.byref(Nr), ctypes.byref(nc), ctypes.byref(nr))
        L, N, NC, NR = Lr.value, Nr.value, nc.value, nr.value
        W = (N + 63) // 64
        P = dict(
            L=L, N=N,
            bp=np.zeros(L, np.int32), cm=np.zeros(L, np.float64),
            cref=np.zeros(L, np.uint32), calt=np.zeros(L, np.uint32),
            common=np.zeros(L, np.uint8), major=np.zeros(L, np.uint8),
            common_words=np.zeros((NC, W), np.uint64),
            rare_ptr=np.zeros(L + 1, np.int32), rare_idx=np.zeros(max(NR, 1), np.int32),
            tgt_haps=np.zeros((2 * n_targets, L), np.uint8),
            dense=np.zeros((N, L), np.uint8) if keep_dense else None,
            lq=np.zeros(L, np.uint8),
        )
        L_.gh_synth_panel_copy(h, _p(P["bp"]), _p(P["cm"]), _p(P["cref"]), _p(P["calt"]),
                               _p(P["common"]), _p(P["major"]), _p(P["common_words"]),
                               _p(P["rare_ptr"]), _p(P["rare_idx"]), _p(P["tgt_haps"]),
                               _p(P["dense"]))
        P["rare_idx"] = P["rare_idx"][:NR]
        return P
    finally:
        L_.gh_synth_panel_free(h)


def make_reads(tgt_haps, coverage, seed=1, threads=None):
    """tgt_haps: [2T][L] uint8; coverage: scalar or [T].  Returns (gl [T][L][3], flat [T][L]).
    threads=None: one random stream for all targets; else one stream per target
    on that many host threads (0: all, up to 16)."""
    haps = np.ascontiguousarray(tgt_haps, dtype=np.uint8)
    T = haps.shape[0] // 2
    L = haps.shape[1]
    cov = np.broadcast_to(np.asarray(coverage, np.float32), (T,)).copy()
    gl = np.zeros((T, L, 3), np.uint8)
    flat = np.zeros((T, L), np.uint8)
    if threads is None:
        lib().gh_synth_reads(L, T, _p(haps), _p(cov), int(seed), _p(gl), _p(flat))
    else:
        lib().gh_synth_reads_mt(L, T, _p(haps), _p(cov), int(seed), _p(gl), _p(flat), _threads(threads))
    return gl, flat


def truth_genotypes(tgt_haps):
    h = np.asarray(tgt_haps)
    return (h[0::2].astype(np.int32) + h[1::2].astype(np.int32)).astype(np.uint8)
