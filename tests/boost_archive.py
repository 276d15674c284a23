"""A Python writer/reader of the split_reference binary panel layout for the
binary-panel tests: boost.serialization's binary archive (header with the
"serialization::archive" signature, library version, native type sizes and
endianness word; class preambles of tracking byte + version; collection sizes,
item versions, pointers as class id + object id) applied to the
ref_haplotype_set and variant_map members in their serialize() order
(ref_haplotype_set.h:144-160, variant_map.h:71-84, variant.h:59-73,
bitmatrix.h:63-77).  No boost library is in the image, so neither this nor
the library's reader is checked against a file boost wrote."""
import struct

import numpy as np


class W:
    def __init__(self, ver_w=4, iv_w=4, cs_w=8):
        self.b = bytearray()
        self.ver_w, self.iv_w, self.cs_w = ver_w, iv_w, cs_w

    def u(self, v, w):
        self.b += int(v).to_bytes(w, "little", signed=False)

    def i32(self, v):
        self.b += struct.pack("<i", int(v))

    def s(self, x):
        x = x.encode()
        self.u(len(x), 8)
        self.b += x

    def pre(self, tracking=0):
        self.u(tracking, 1)
        self.u(0, self.ver_w)

    def count(self, n):
        self.u(n, self.cs_w)


def encode(H, V, lib=18, ver_w=4, iv_w=4, cs_w=8):
    """H: dict n_tot, n_rar, n_com, n_com_hq, n_haps, flag_common, major,
    common2tot, shap (list of lists), hv (bytes), hv_rows, hv_cols, ypacked,
    asmall (list of lists).  V: dict chrid, starts (4 ints), gregions (2),
    variants (dicts bp, id, ref, alt, type, idx, cref, calt, cm, lq)."""
    w = W(ver_w, iv_w, cs_w)
    w.s("serialization::archive")
    w.u(lib, 2)
    w.b += bytes([4, 8, 4, 8])
    w.i32(1)
    w.pre()
    for k in ("n_tot", "n_rar", "n_com", "n_com_hq", "n_haps"):
        w.u(H[k], 4)
    for k in ("flag_common", "major"):
        w.count(len(H[k]))
        w.b += bytes(int(x) for x in H[k])
    w.count(len(H["common2tot"]))
    w.b += np.asarray(H["common2tot"], "<i4").tobytes()

    def vv(rows, first):
        if first:
            w.pre()
        w.count(len(rows))
        w.u(0, w.iv_w)
        for r in rows:
            w.count(len(r))
            w.b += np.asarray(r, "<i4").tobytes()

    vv(H["shap"], True)
    w.pre()
    w.u(len(H["hv"]), 8)
    w.u(H["hv_cols"], 8)
    w.u(H["hv_rows"], 8)
    w.b += bytes(H["hv"])
    w.count(len(H["ypacked"]))
    w.b += bytes(H["ypacked"])
    vv(H["asmall"], False)
    w.pre()
    w.s(V["chrid"])
    for x in V["starts"]:
        w.i32(x)
    for g in V["gregions"]:
        w.s(g)
    w.pre()
    vs = V["variants"]
    w.count(len(vs))
    w.u(0, w.iv_w)
    for j, v in enumerate(vs):
        w.u(7, 2)                      # class id of variant
        if j == 0:
            w.pre(tracking=1)
        w.u(j, 4)                      # object id
        w.i32(v["bp"])
        w.s(v["id"])
        w.s(v["ref"])
        w.s(v["alt"])
        w.u(v["type"], 1)
        w.i32(v["idx"])
        w.u(v["cref"], 4)
        w.u(v["calt"], 4)
        w.b += struct.pack("<d", v["cm"])
        w.u(v["lq"], 1)
    # map_pos: pairs of (bp, reference to the saved variant)
    w.pre()
    w.count(len(vs))
    w.u(0, w.iv_w)
    for j, v in enumerate(vs):
        if j == 0:
            w.pre()
        w.i32(v["bp"])
        w.u(7, 2)
        w.u(j, 4)
    return bytes(w.b)


def from_panel(P, pos, ref, alt, ypacked, ap, ai, chrid="chr20", sparse_maf=0.001):
    """H and V dicts for a panel in the arrays hip.Panel takes."""
    L, N = P["L"], P["N"]
    cref, calt = np.asarray(P["cref"]), np.asarray(P["calt"])
    tot = (cref + calt).astype(np.float32)
    maf = np.minimum(cref.astype(np.float32) / tot, calt.astype(np.float32) / tot)
    common = maf >= np.float32(sparse_maf)
    c2t = np.nonzero(common)[0]
    lq = np.asarray(P.get("lq", np.zeros(L, np.uint8)))
    shap = [[] for _ in range(N)]
    rp, ri = P["rare_ptr"], P["rare_idx"]
    for l in range(L):
        if not common[l]:
            for k in range(rp[l], rp[l + 1]):
                shap[ri[k]].append(l)
    rows, cols = (len(c2t) + 7) // 8 * 8, (N + 7) // 8 * 8
    bits = np.zeros((rows, cols), np.uint8)
    words = np.asarray(P["common_words"])
    for c in range(len(c2t)):
        bits[c, :N] = np.unpackbits(words[c].view(np.uint8), bitorder="little")[:N]
    hv = np.packbits(bits, axis=1, bitorder="big").tobytes()
    nhq = int((common & (lq == 0)).sum())
    H = {"n_tot": L, "n_rar": int(L - common.sum()), "n_com": int(common.sum()), "n_com_hq": nhq, "n_haps": N,
         "flag_common": common.astype(int), "major": (calt > cref).astype(int), "common2tot": c2t, "shap": shap,
         "hv": hv, "hv_rows": rows, "hv_cols": cols, "ypacked": bytes(ypacked),
         "asmall": [list(ai[ap[k]:ap[k + 1]]) for k in range(len(ap) - 1)]}
    V = {"chrid": chrid, "starts": [1, int(pos[-1]), 1, int(pos[-1])],
         "gregions": ["%s:1-%d" % (chrid, pos[-1])] * 2,
         "variants": [{"bp": int(pos[l]), "id": ".", "ref": ref[l], "alt": alt[l], "type": 1, "idx": l,
                       "cref": int(cref[l]), "calt": int(calt[l]), "cm": float(P["cm"][l]), "lq": int(lq[l])}
                      for l in range(L)]}
    return H, V
