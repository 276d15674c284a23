"""The split_reference binary panel (boost binary archive of ref_haplotype_set
and variant_map) through libglimpse_hip.so: read into the panel arrays that
hip.Panel takes, and written in the same layout.  Host code only."""
import ctypes

import numpy as np

from . import hip

vp = ctypes.c_void_p
cp = ctypes.c_char_p


class BinInfo(ctypes.Structure):
    _fields_ = [("chrid", cp), ("input_start", ctypes.c_int32), ("input_stop", ctypes.c_int32),
                ("output_start", ctypes.c_int32), ("output_stop", ctypes.c_int32), ("input_gregion", cp),
                ("output_gregion", cp), ("n_sites", ctypes.c_int32), ("n_haps", ctypes.c_int32),
                ("n_common", ctypes.c_int32), ("n_common_hq", ctypes.c_int32), ("library_version", ctypes.c_int32),
                ("sparse_maf", ctypes.c_float)]


class BinMeta(ctypes.Structure):
    _fields_ = [("chrid", cp), ("input_start", ctypes.c_int32), ("input_stop", ctypes.c_int32),
                ("output_start", ctypes.c_int32), ("output_stop", ctypes.c_int32), ("input_gregion", cp),
                ("output_gregion", cp), ("pos", vp), ("id", ctypes.POINTER(cp)), ("ref", ctypes.POINTER(cp)),
                ("alt", ctypes.POINTER(cp)), ("type", vp), ("library_version", ctypes.c_int32)]


_bound = False


def _lib():
    global _bound
    L = hip.lib()
    if not _bound:
        L.gh_bin_open.argtypes = [cp, ctypes.POINTER(vp)]
        L.gh_bin_close.argtypes = [vp]
        L.gh_bin_info.argtypes = [vp, ctypes.POINTER(BinInfo)]
        L.gh_bin_desc.argtypes = [vp, ctypes.POINTER(hip.PanelDesc)]
        L.gh_bin_sites.argtypes = [vp, ctypes.POINTER(ctypes.POINTER(ctypes.c_int32)),
                                   ctypes.POINTER(ctypes.POINTER(cp)), ctypes.POINTER(ctypes.POINTER(cp)),
                                   ctypes.POINTER(ctypes.POINTER(cp)), ctypes.POINTER(ctypes.POINTER(ctypes.c_uint8))]
        L.gh_bin_write.argtypes = [cp, ctypes.POINTER(hip.PanelDesc), ctypes.POINTER(BinMeta)]
        _bound = True
    return L


def _arr(ptr, n, dt):
    if n == 0:
        return np.zeros(0, dt)
    return np.ctypeslib.as_array(ctypes.cast(ptr, ctypes.POINTER(np.ctypeslib.as_ctypes_type(dt))), (n,)).copy()


def read(path):
    """Returns (P, info): P holds the panel arrays (L, N, cm, cref, calt, lq,
    common_words, rare_ptr, rare_idx, ypacked, asmall_ptr, asmall_idx,
    sparse_maf, pos, id, ref, alt, type); info the archive's header fields."""
    L = _lib()
    h = vp()
    hip._check(L.gh_bin_open(str(path).encode(), ctypes.byref(h)))
    try:
        inf = BinInfo()
        hip._check(L.gh_bin_info(h, ctypes.byref(inf)))
        d = hip.PanelDesc()
        hip._check(L.gh_bin_desc(h, ctypes.byref(d)))
        n, N, nc, nhq = inf.n_sites, inf.n_haps, inf.n_common, inf.n_common_hq
        W = (N + 63) // 64
        rp = _arr(d.rare_ptr, n + 1, np.int32)
        ap = _arr(d.asmall_ptr, nhq + 2, np.int32)
        P = {"L": n, "N": N, "cm": _arr(d.cm, n, np.float64), "cref": _arr(d.cref, n, np.uint32),
             "calt": _arr(d.calt, n, np.uint32), "lq": _arr(d.lq, n, np.uint8),
             "common_words": _arr(d.common_words, nc * W, np.uint64).reshape(nc, W),
             "rare_ptr": rp, "rare_idx": _arr(d.rare_idx, max(int(rp[-1]), 1), np.int32),
             "ypacked": _arr(d.ypacked, d.ypacked_len, np.uint8), "asmall_ptr": ap,
             "asmall_idx": _arr(d.asmall_idx, max(int(ap[-1]), 1), np.int32), "sparse_maf": d.sparse_maf}
        pos = ctypes.POINTER(ctypes.c_int32)()
        ids, refs, alts = ctypes.POINTER(cp)(), ctypes.POINTER(cp)(), ctypes.POINTER(cp)()
        ty = ctypes.POINTER(ctypes.c_uint8)()
        hip._check(L.gh_bin_sites(h, ctypes.byref(pos), ctypes.byref(ids), ctypes.byref(refs), ctypes.byref(alts),
                                  ctypes.byref(ty)))
        P["pos"] = np.array(pos[:n], np.int32)
        P["id"] = [ids[i].decode() for i in range(n)]
        P["ref"] = [refs[i].decode() for i in range(n)]
        P["alt"] = [alts[i].decode() for i in range(n)]
        P["type"] = np.array(ty[:n], np.uint8)
        info = {f: (getattr(inf, f).decode() if isinstance(getattr(inf, f), bytes) else getattr(inf, f))
                for f, _ in BinInfo._fields_}
        return P, info
    finally:
        L.gh_bin_close(h)


def write(path, P, pos, ref, alt, ids=None, chrid="chr1", region=(0, 0, 0, 0), gregion=("", ""),
          library_version=18):
    """Writes panel P (arrays as hip.Panel takes them) with its sparse PBWT
    (built with hip.build_sparse_pbwt) and site metadata in the split_reference
    layout."""
    y, ap, ai = hip.build_sparse_pbwt(P)
    k = hip._desc_arrays(P)
    d = hip._desc(P, k, y, (ap, ai))
    if "sparse_maf" in P:
        d.sparse_maf = P["sparse_maf"]
    pos = np.ascontiguousarray(pos, np.int32)
    n = pos.shape[0]
    mk = lambda xs: (cp * max(n, 1))(*[x.encode() for x in xs])
    ra, aa = mk(ref), mk(alt)
    ia = mk(ids) if ids is not None else None
    m = BinMeta(chrid.encode(), region[0], region[1], region[2], region[3], gregion[0].encode(), gregion[1].encode(),
                hip._p(pos), ia, ra, aa, None, library_version)
    hip._check(_lib().gh_bin_write(str(path).encode(), ctypes.byref(d), ctypes.byref(m)))
    return y, ap, ai
