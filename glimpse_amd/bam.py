"""Genotype likelihoods from a sample's BAM file through libglimpse_hip.so
(gh_bam_call_gls): the phase tool's BAM caller (genotype_bam_caller) with
the standard model.  No CPU fallback."""
import ctypes

import numpy as np

from . import hip

vp = ctypes.c_void_p
cp = ctypes.c_char_p


class BamParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int32) for n in ("min_mapq", "min_baseq", "max_depth", "fflag", "keep_orphan",
                                              "check_orientation", "check_proper_pair")]


class BamRegion(ctypes.Structure):
    _fields_ = [("contig", cp), ("beg", ctypes.c_int32), ("end", ctypes.c_int32), ("n_sites", ctypes.c_int32),
                ("pos", vp), ("ref", ctypes.POINTER(cp)), ("alt", ctypes.POINTER(cp))]


class BamStats(ctypes.Structure):
    _fields_ = [("n_records", ctypes.c_int64), ("n_kept", ctypes.c_int64), ("sites_over_max_depth", ctypes.c_int64)]


_bound = False


def _lib():
    global _bound
    L = hip.lib()
    if not _bound:
        L.gh_bam_default_params.argtypes = [ctypes.POINTER(BamParams)]
        L.gh_bam_default_params.restype = None
        L.gh_bam_tables.argtypes = [vp]
        L.gh_bam_call_gls.argtypes = [cp, ctypes.c_int32, ctypes.POINTER(BamRegion), ctypes.POINTER(BamParams), vp,
                                      vp, vp, ctypes.POINTER(BamStats)]
        _bound = True
    return L


def default_params(**kw):
    p = BamParams()
    _lib().gh_bam_default_params(ctypes.byref(p))
    for k, v in kw.items():
        setattr(p, k, int(v))
    return p


def tables():
    t = np.zeros(1024, np.float32)
    hip._check(_lib().gh_bam_tables(hip._p(t)))
    return t.reshape(4, 256)


def call_gls(path, contig, pos, ref, alt, ploidy=2, beg=0, end=0, **params):
    """(gl [L][ploidy+1], flat [L], depth [L], stats dict) for one BAM file."""
    pos = np.ascontiguousarray(pos, np.int32)
    L = pos.shape[0]
    ra = (cp * max(L, 1))(*[x.encode() for x in ref])
    aa = (cp * max(L, 1))(*[x.encode() for x in alt])
    rg = BamRegion(contig.encode(), beg, end, L, hip._p(pos), ra, aa)
    prm = default_params(**params)
    gl = np.zeros((L, ploidy + 1), np.uint8)
    flat = np.zeros(L, np.uint8)
    depth = np.zeros(L, np.int32)
    st = BamStats()
    hip._check(_lib().gh_bam_call_gls(str(path).encode(), ploidy, ctypes.byref(rg), ctypes.byref(prm), hip._p(gl),
                                      hip._p(flat), hip._p(depth), ctypes.byref(st)))
    return gl, flat, depth, {f: getattr(st, f) for f, _ in BamStats._fields_}
