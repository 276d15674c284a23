"""A small, independent BCF 2.2 / BGZF codec in Python for the I/O tests,
written from the published specifications (VCF v4.3 section 6, "BCF
specification"; SAM v1 section 4.1, "BGZF"), not from the library's C++:
it writes the BCF files the readers are tested on and decodes the BCF files
the writer produces."""
import gzip
import struct
import zlib

BGZF_EOF = bytes.fromhex("1f8b08040000000000ff0600424302001b0003000000000000000000")
INT8_MISSING, INT8_END = -128, -127
INT32_MISSING, INT32_END = -2147483648, -2147483647
F32_MISSING, F32_END = 0x7F800001, 0x7F800002


def bgzf_compress(data, block=0xFF00):
    out = bytearray()
    for i in range(0, len(data), block) if data else []:
        chunk = data[i:i + block]
        c = zlib.compressobj(6, zlib.DEFLATED, -15)
        cdata = c.compress(chunk) + c.flush()
        bsize = len(cdata) + 25
        out += struct.pack("<BBBBIBBHBBHH", 31, 139, 8, 4, 0, 0, 255, 6, 66, 67, 2, bsize)
        out += cdata + struct.pack("<II", zlib.crc32(chunk) & 0xFFFFFFFF, len(chunk))
    return bytes(out + BGZF_EOF)


def bgzf_blocks(raw):
    """Walks BGZF blocks: list of (block size, uncompressed size); checks each block's framing."""
    res, p = [], 0
    while p < len(raw):
        id1, id2, cm, flg, _mt, _xfl, _os, xlen = struct.unpack_from("<BBBBIBBH", raw, p)
        assert (id1, id2, cm, flg) == (31, 139, 8, 4)
        si1, si2, slen, bsize = struct.unpack_from("<BBHH", raw, p + 12)
        assert (si1, si2, slen, xlen) == (66, 67, 2, 6)
        total = bsize + 1
        crc, isize = struct.unpack_from("<II", raw, p + total - 8)
        data = zlib.decompress(raw[p + 18:p + total - 8], -15)
        assert len(data) == isize and (zlib.crc32(data) & 0xFFFFFFFF) == crc
        res.append((total, isize))
        p += total
    return res


# ---------------------------------------------------------------- typed values
def _desc(t, n):
    if n < 15:
        return bytes([n << 4 | t])
    return bytes([15 << 4 | t]) + _typed_int(n)


def _typed_int(v):
    if -120 <= v <= 127:
        return bytes([0x11]) + struct.pack("<b", v)
    if -32760 <= v <= 32767:
        return bytes([0x12]) + struct.pack("<h", v)
    return bytes([0x13]) + struct.pack("<i", v)


def _typed_str(s):
    b = s.encode()
    return _desc(7, len(b)) + b


def _ints(vals, width):
    fmt = {1: "<b", 2: "<h", 4: "<i"}[width]
    miss = {1: (-128, -127), 2: (-32768, -32767), 4: (INT32_MISSING, INT32_END)}[width]
    out = b""
    for v in vals:
        v = miss[0] if v is None else (miss[1] if v == "end" else v)
        out += struct.pack(fmt, v)
    return out


def encode_bcf(header_text, samples, records, dictionary, contigs):
    """records: dicts with chrom, pos (1-based), id, alleles, info {key: (kind, values)},
    fmt [(key, kind, per_sample_values)] with kind "i8", "i16", "i32", "f", and
    per-sample values lists (None missing, "end" vector end)."""
    hdr = header_text.encode() + b"\0"
    body = b"BCF\x02\x02" + struct.pack("<I", len(hdr)) + hdr
    for r in records:
        sh = struct.pack("<iiiI", contigs.index(r["chrom"]), r["pos"] - 1, len(r["alleles"][0]), F32_MISSING)
        sh += struct.pack("<I", len(r["alleles"]) << 16 | len(r.get("info", {})))
        sh += struct.pack("<I", len(r.get("fmt", [])) << 24 | len(samples))
        sh += _typed_str("" if r.get("id", ".") == "." else r["id"])
        for a in r["alleles"]:
            sh += _typed_str(a)
        sh += _desc(0, 0)
        for key, (kind, vals) in r.get("info", {}).items():
            sh += _typed_int(dictionary[key])
            if kind == "f":
                sh += _desc(5, len(vals)) + b"".join(struct.pack("<f", v) for v in vals)
            else:
                sh += _desc(3, len(vals)) + _ints(vals, 4)
        ind = b""
        for key, kind, per in r.get("fmt", []):
            n = max(len(v) for v in per)
            ind += _typed_int(dictionary[key])
            if kind == "f":
                ind += _desc(5, n)
                for v in per:
                    for k in range(n):
                        x = v[k] if k < len(v) else "end"
                        bits = F32_MISSING if x is None else F32_END if x == "end" else None
                        ind += struct.pack("<I", bits) if bits is not None else struct.pack("<f", x)
            else:
                w = {"i8": 1, "i16": 2, "i32": 4}[kind]
                ind += _desc({1: 1, 2: 2, 4: 3}[w], n)
                for v in per:
                    ind += _ints([v[k] if k < len(v) else "end" for k in range(n)], w)
        body += struct.pack("<II", len(sh), len(ind)) + sh + ind
    return bgzf_compress(body)


# ---------------------------------------------------------------- decoding
class _Buf:
    def __init__(self, b):
        self.b, self.p = b, 0

    def take(self, n):
        v = self.b[self.p:self.p + n]
        assert len(v) == n
        self.p += n
        return v

    def desc(self):
        d = self.take(1)[0]
        t, n = d & 15, d >> 4
        if n == 15:
            n = self.typed_int()
        return t, n

    def typed_int(self):
        t, n = self.desc()
        assert n == 1
        return self.scalar(t)

    def scalar(self, t):
        return {1: lambda: struct.unpack("<b", self.take(1))[0], 2: lambda: struct.unpack("<h", self.take(2))[0],
                3: lambda: struct.unpack("<i", self.take(4))[0], 5: lambda: struct.unpack("<f", self.take(4))[0],
                7: lambda: self.take(1).decode()}[t]()

    def vector(self, t, n):
        if t == 7:
            return self.take(n).decode().rstrip("\0")
        if t == 0:
            return []
        return [self.scalar(t) for _ in range(n)]


def decode_bcf(raw):
    """(header text, list of records with raw typed values), from BGZF bytes."""
    data = gzip.decompress(raw)
    assert data[:5] == b"BCF\x02\x02"
    lt = struct.unpack_from("<I", data, 5)[0]
    text = data[9:9 + lt].rstrip(b"\0").decode()
    p = 9 + lt
    recs = []
    while p < len(data):
        ls, li = struct.unpack_from("<II", data, p)
        p += 8
        sh, ind = _Buf(data[p:p + ls]), _Buf(data[p + ls:p + ls + li])
        p += ls + li
        rid, pos, rlen, qual_bits, nai, nfs = struct.unpack("<iiiIII", sh.take(24))
        r = {"rid": rid, "pos": pos + 1, "rlen": rlen, "qual_bits": qual_bits,
             "n_info": nai & 0xFFFF, "n_allele": nai >> 16, "n_fmt": nfs >> 24, "n_sample": nfs & 0xFFFFFF}
        r["id"] = sh.vector(*sh.desc())
        r["alleles"] = [sh.vector(*sh.desc()) for _ in range(r["n_allele"])]
        r["filter"] = sh.vector(*sh.desc())
        r["info"] = {}
        for _ in range(r["n_info"]):
            k = sh.typed_int()
            r["info"][k] = sh.vector(*sh.desc())
        r["fmt"] = {}
        for _ in range(r["n_fmt"]):
            k = ind.typed_int()
            t, n = ind.desc()
            r["fmt"][k] = (t, n, [ind.vector(t, n) for _ in range(r["n_sample"])])
        recs.append(r)
    return text, recs
