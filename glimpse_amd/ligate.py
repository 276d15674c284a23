"""Host mirror of the ligate tool (ligate/src/ligater/ligater_algorithm.cpp)
over gh_ligate / gh_ligate_scan.  A chunk is the record list of one phased
output file: contig id, position, allele key and biallelic flag per record and
its GT values in htslib's encoding ((allele + 1) << 1 | phased, 0 missing,
INT32_MIN + 1 vector end), [records][samples][2] int32.  The VCF/BCF reading
and writing around it (ligater::ligate, :216-330) are outside this module.
No CPU fallback: every call goes to libglimpse_hip.so."""
import ctypes

import numpy as np

from . import hip

vp = ctypes.c_void_p
VECTOR_END = -2147483647


def gt_phased(allele):
    return ((np.asarray(allele, np.int32) + 1) << 1) | 1


def gt_unphased(allele):
    return (np.asarray(allele, np.int32) + 1) << 1


def gt_allele(v):
    return (np.asarray(v, np.int32) >> 1) - 1


class LigChunk(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_int32), ("contig", vp), ("pos", vp), ("key", vp), ("biallelic", vp),
                ("gt", vp)]


class LigOverlap(ctypes.Structure):
    _fields_ = [("n_common", ctypes.c_int32), ("n_scanned", ctypes.c_int32), ("n_swapped", ctypes.c_int32),
                ("mean_hets", ctypes.c_float), ("mean_phaseq", ctypes.c_float)]


_bound = False


def _lib():
    global _bound
    L = hip.lib()
    if not _bound:
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        L.gh_ligate.argtypes = [i32, i32, ctypes.POINTER(LigChunk), i64, vp, vp, vp, vp, ctypes.POINTER(LigOverlap)]
        L.gh_ligate_scan.argtypes = [i32, i32, vp, vp, vp, vp, vp, vp, vp]
        _bound = True
    return L


def scan(gt_a, gt_b, swap_a=None):
    """gh_ligate_scan over aligned overlap records: per sample nmatch, nmism,
    the swap of chunk B and the phase quality."""
    gt_a = np.ascontiguousarray(gt_a, np.int32)
    gt_b = np.ascontiguousarray(gt_b, np.int32)
    S, N = gt_a.shape[0], gt_a.shape[1]
    assert gt_a.shape == (S, N, 2) and gt_b.shape == (S, N, 2)
    swap_a = np.zeros(N, np.uint8) if swap_a is None else np.ascontiguousarray(swap_a, np.uint8)
    nmatch, nmism = np.zeros(N, np.int32), np.zeros(N, np.int32)
    swap_b, q = np.zeros(N, np.uint8), np.zeros(N, np.float32)
    hip._check(_lib().gh_ligate_scan(S, N, hip._p(gt_a), hip._p(gt_b), hip._p(swap_a), hip._p(nmatch),
                                     hip._p(nmism), hip._p(swap_b), hip._p(q)))
    return nmatch, nmism, swap_b, q


def ligate(chunks, n_samples):
    """gh_ligate.  chunks: list of dicts with "contig", "pos", "key",
    optional "biallelic", and "gt" ([n][N][2] int32, updated in place).
    Returns (order [n_out][2] of (chunk, record), swap [n_chunks][N],
    per-overlap statistics as dicts)."""
    keep, descs = [], []
    total = 0
    for c in chunks:
        n = len(c["pos"])
        arrs = {"contig": np.ascontiguousarray(c["contig"], np.int32),
                "pos": np.ascontiguousarray(c["pos"], np.int32),
                "key": np.ascontiguousarray(c["key"], np.uint64),
                "biallelic": None if c.get("biallelic") is None else np.ascontiguousarray(c["biallelic"], np.uint8)}
        gt = c["gt"]
        if not (isinstance(gt, np.ndarray) and gt.dtype == np.int32 and gt.flags.c_contiguous):
            raise ValueError("gt must be a C-contiguous int32 array (it is updated in place)")
        assert gt.shape == (n, n_samples, 2)
        keep.append(arrs)
        descs.append(LigChunk(n, hip._p(arrs["contig"]), hip._p(arrs["pos"]), hip._p(arrs["key"]),
                              hip._p(arrs["biallelic"]), hip._p(gt)))
        total += n
    nc = len(chunks)
    arr = (LigChunk * nc)(*descs)
    out_chunk, out_site = np.zeros(max(total, 1), np.int32), np.zeros(max(total, 1), np.int32)
    n_out = np.zeros(1, np.int64)
    swap = np.zeros((nc, n_samples), np.uint8)
    ov = (LigOverlap * max(nc - 1, 1))()
    hip._check(_lib().gh_ligate(nc, n_samples, arr, total, hip._p(out_chunk), hip._p(out_site), hip._p(n_out),
                                hip._p(swap), ov))
    n = int(n_out[0])
    order = np.stack([out_chunk[:n], out_site[:n]], axis=1)
    stats = [{f: getattr(ov[k], f) for f, _ in LigOverlap._fields_} for k in range(nc - 1)]
    return order, swap, stats
