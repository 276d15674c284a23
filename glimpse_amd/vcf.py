"""VCF / BCF input and output through libglimpse_hip.so (include/glimpse_hip.h,
host code on zlib): the target GL reader of the phase tool
(genotype_reader::scanTarGenotypes / parseTarGenotypes), its output writer
(genotype_writer::writeGenotypes) and a record reader with htslib's value
conventions for the concordance and ligate front ends."""
import ctypes

import numpy as np

from . import hip

vp = ctypes.c_void_p
cp = ctypes.c_char_p
I32_MISSING = -2147483648
I32_END = -2147483647


class VcfRecord(ctypes.Structure):
    _fields_ = [("chrom", cp), ("pos", ctypes.c_int32), ("id", cp), ("n_allele", ctypes.c_int32), ("ref", cp),
                ("alt", cp), ("is_snp", ctypes.c_int32)]


class GlQuery(ctypes.Structure):
    _fields_ = [("n_sites", ctypes.c_int32), ("pos", vp), ("ref", ctypes.POINTER(cp)), ("alt", ctypes.POINTER(cp)),
                ("use_gl", ctypes.c_int32), ("use_indels", ctypes.c_int32)]


class VcfOut(ctypes.Structure):
    _fields_ = [("contig", cp), ("n_samples", ctypes.c_int32), ("sample_names", ctypes.POINTER(cp)),
                ("n_sites", ctypes.c_int32), ("pos", vp), ("id", ctypes.POINTER(cp)), ("ref", ctypes.POINTER(cp)),
                ("alt", ctypes.POINTER(cp)), ("gt", vp), ("ds", vp), ("gp", vp), ("info", vp), ("af", vp),
                ("raf", vp), ("n_main", ctypes.c_int32), ("fploidy", ctypes.c_int32), ("source", cp),
                ("file_date", cp), ("ploidy", vp)]


_bound = False


def _lib():
    global _bound
    L = hip.lib()
    if not _bound:
        i32, i64 = ctypes.c_int32, ctypes.c_int64
        L.gh_vcf_open.argtypes = [cp, ctypes.POINTER(vp)]
        L.gh_vcf_close.argtypes = [vp]
        L.gh_vcf_samples.argtypes = [vp, vp, vp, i64]
        L.gh_vcf_next.argtypes = [vp, ctypes.POINTER(VcfRecord)]
        L.gh_vcf_format.argtypes = [vp, cp, vp, vp, i64]
        L.gh_vcf_info.argtypes = [vp, cp, vp, vp, i64]
        L.gh_vcf_read_gls.argtypes = [vp, ctypes.POINTER(GlQuery), vp, vp, vp]
        L.gh_write_genotypes.argtypes = [cp, ctypes.POINTER(VcfOut)]
        _bound = True
    return L


def _strs(xs):
    arr = (cp * max(len(xs), 1))()
    keep = [x if isinstance(x, bytes) else str(x).encode() for x in xs]
    for i, x in enumerate(keep):
        arr[i] = x
    return arr, keep


class Reader:
    """Record reader.  for rec in Reader(path): rec is a dict with chrom,
    pos (1-based), id, n_allele, ref, alt, is_snp; format(key) / info(key)
    return the current record's field as htslib hands it out."""

    def __init__(self, path):
        h = vp()
        hip._check(_lib().gh_vcf_open(str(path).encode(), ctypes.byref(h)))
        self.h = h
        n = np.zeros(1, np.int32)
        hip._check(_lib().gh_vcf_samples(self.h, hip._p(n), None, 0))
        self.n_samples = int(n[0])
        cap = 1 << 20
        while True:
            buf = ctypes.create_string_buffer(cap)
            rc = _lib().gh_vcf_samples(self.h, hip._p(n), buf, cap)
            if rc == -4:
                cap *= 4
                continue
            hip._check(rc)
            break
        raw = buf.raw.split(b"\0")
        self.samples = [x.decode() for x in raw[: self.n_samples]]
        self._rec = VcfRecord()

    def close(self):
        if self.h:
            _lib().gh_vcf_close(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __iter__(self):
        while True:
            rc = _lib().gh_vcf_next(self.h, ctypes.byref(self._rec))
            if rc == 0:
                return
            hip._check(0 if rc == 1 else rc)
            r = self._rec
            yield {"chrom": r.chrom.decode(), "pos": r.pos, "id": r.id.decode(), "n_allele": r.n_allele,
                   "ref": r.ref.decode(), "alt": r.alt.decode(), "is_snp": bool(r.is_snp)}

    def _field(self, fn, key, is_float, cap):
        key = key.encode()
        while True:
            buf = np.zeros(cap, np.float32 if is_float else np.int32)
            rc = fn(self.h, key, None if is_float else hip._p(buf), hip._p(buf) if is_float else None, cap)
            if rc == -4:
                cap *= 4
                continue
            hip._check(min(rc, 0))
            return buf[:rc]

    def format(self, key, is_float=False):
        """Per-sample values, shape [n_samples][count] (count 0: field absent)."""
        v = self._field(_lib().gh_vcf_format, key, is_float, max(4 * self.n_samples, 64))
        return v.reshape(self.n_samples, -1) if v.size else v.reshape(self.n_samples, 0)

    def info(self, key, is_float=False):
        return self._field(_lib().gh_vcf_info, key, is_float, 64)

    def read_gls(self, pos, ref, alt, use_gl=False, use_indels=False):
        """gh_vcf_read_gls over the remaining records: (gl [T][L][3] uint8,
        flat [T][L] uint8, number of records used)."""
        pos = np.ascontiguousarray(pos, np.int32)
        L = pos.shape[0]
        ra, keep_r = _strs(ref)
        aa, keep_a = _strs(alt)
        q = GlQuery(L, hip._p(pos), ra, aa, int(use_gl), int(use_indels))
        gl = np.zeros((self.n_samples, L, 3), np.uint8)
        flat = np.zeros((self.n_samples, L), np.uint8)
        n = np.zeros(1, np.int64)
        hip._check(_lib().gh_vcf_read_gls(self.h, ctypes.byref(q), hip._p(gl), hip._p(flat), hip._p(n)))
        return gl, flat, int(n[0])


def read_target_gls(path, pos, ref, alt, use_gl=False, use_indels=False):
    """Target samples and their GLs at the panel's sites (pos 1-based)."""
    r = Reader(path)
    try:
        gl, flat, n = r.read_gls(pos, ref, alt, use_gl, use_indels)
        return r.samples, gl, flat, n
    finally:
        r.close()


def write_genotypes(path, contig, samples, pos, ref, alt, fields, ids=None, n_main=15, fploidy=None,
                    source="GLIMPSE_phase (MI355X)", file_date="", ploidy=None):
    """gh_write_genotypes from the dict gh_output_fields returns
    (gt [T][L][2], ds [T][L], gp [T][L][3], info / af / raf [L]); ploidy [T]
    (None: all diploid) also sets the format ploidy (2, 1 or -2 when mixed)."""
    pl = None if ploidy is None else np.ascontiguousarray(ploidy, np.int32)
    if fploidy is None:
        fploidy = 2 if pl is None or (pl == 2).all() else (1 if (pl == 1).all() else -2)
    pos = np.ascontiguousarray(pos, np.int32)
    T, L = len(samples), pos.shape[0]
    gt = np.ascontiguousarray(fields["gt"], np.uint8)
    ds = np.ascontiguousarray(fields["ds"], np.float32)
    gp = np.ascontiguousarray(fields["gp"], np.float32)
    info = np.ascontiguousarray(fields["info"], np.float32)
    af = np.ascontiguousarray(fields["af"], np.float32)
    raf = np.ascontiguousarray(fields["raf"], np.float32)
    assert gt.shape == (T, L, 2) and ds.shape == (T, L) and gp.shape == (T, L, 3)
    sa, k1 = _strs(samples)
    ra, k2 = _strs(ref)
    aa, k3 = _strs(alt)
    ia, k4 = _strs(ids) if ids is not None else (None, None)
    d = VcfOut(contig.encode(), T, sa, L, hip._p(pos), ia, ra, aa, hip._p(gt), hip._p(ds), hip._p(gp),
               hip._p(info), hip._p(af), hip._p(raf), n_main, fploidy, source.encode(), file_date.encode(),
               hip._p(pl))
    hip._check(_lib().gh_write_genotypes(str(path).encode(), ctypes.byref(d)))
