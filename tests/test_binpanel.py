"""The split_reference binary panel reader / writer (host code; no GPU):
* a panel written by an independent Python encoder of the layout
  (tests/boost_archive.py) decodes into exactly the arrays hip.Panel takes,
  including the stored sparse-PBWT stream and A_small_idx;
* the library's writer and reader round-trip a synthetic panel, and the
  Python encoder's bytes equal the writer's;
* the older field widths decode; bad signature, foreign platform, truncation
  and inconsistent counts are GH_ERR_IO.
Parity against files boost itself wrote is unpinned: boost is not in the image
and the reference ships no panel."""
import numpy as np
import pytest

import boost_archive as ba


@pytest.fixture(scope="module")
def panel():
    from glimpse_amd import hip, synth
    P = synth.make_panel(600, 900, 2, length_mb=0.5, rare_frac=0.4, flip_frac=0.05, seed=21)
    P["lq"] = (np.arange(P["L"]) % 11 == 5).astype(np.uint8)
    y, ap, ai = hip.build_sparse_pbwt(P)
    pos = np.arange(1, P["L"] + 1, dtype=np.int32) * 53 + 1000
    ref = ["ACGT"[i % 4] for i in range(P["L"])]
    alt = ["CGTA"[i % 4] for i in range(P["L"])]
    return P, y, ap, ai, pos, ref, alt


def _check(Q, P, y, ap, ai, pos, ref, alt):
    assert Q["L"] == P["L"] and Q["N"] == P["N"]
    for k in ("cm", "cref", "calt", "lq", "common_words", "rare_ptr"):
        assert np.array_equal(np.asarray(Q[k]), np.asarray(P[k])), k
    n = int(P["rare_ptr"][-1])
    assert np.array_equal(Q["rare_idx"][:n], np.asarray(P["rare_idx"])[:n])
    assert np.array_equal(Q["ypacked"], y)
    assert np.array_equal(Q["asmall_ptr"], ap) and np.array_equal(Q["asmall_idx"][:ap[-1]], ai[:ap[-1]])
    assert np.array_equal(Q["pos"], pos) and Q["ref"] == ref and Q["alt"] == alt


def test_reads_python_encoded_panel(tmp_path, panel):
    from glimpse_amd import binpanel
    P, y, ap, ai, pos, ref, alt = panel
    H, V = ba.from_panel(P, pos, ref, alt, y, ap, ai)
    p = tmp_path / "chunk.bin"
    p.write_bytes(ba.encode(H, V))
    Q, info = binpanel.read(p)
    _check(Q, P, y, ap, ai, pos, ref, alt)
    assert info["chrid"] == "chr20" and info["library_version"] == 18
    assert info["n_common"] == H["n_com"] and info["n_common_hq"] == H["n_com_hq"]
    # the threshold recovered from the flags reproduces them
    tot = (P["cref"] + P["calt"]).astype(np.float32)
    maf = np.minimum(P["cref"].astype(np.float32) / tot, P["calt"].astype(np.float32) / tot)
    assert np.array_equal(maf >= np.float32(info["sparse_maf"]), np.asarray(H["flag_common"], bool))


def test_writer_round_trip_and_bytes(tmp_path, panel):
    from glimpse_amd import binpanel
    P, y, ap, ai, pos, ref, alt = panel
    p = tmp_path / "w.bin"
    binpanel.write(p, P, pos, ref, alt, chrid="chr20", region=(1, int(pos[-1]), 1, int(pos[-1])),
                   gregion=("chr20:1-%d" % pos[-1], "chr20:1-%d" % pos[-1]))
    Q, info = binpanel.read(p)
    _check(Q, P, y, ap, ai, pos, ref, alt)
    H, V = ba.from_panel(P, pos, ref, alt, y, ap, ai)
    want = ba.encode(H, V)
    got = p.read_bytes()
    # identical except the class id the writer gives variant (it counts the
    # archive's registered classes) -- compare with that field masked out
    assert len(got) == len(want)
    diff = [i for i in range(len(got)) if got[i] != want[i]]
    assert all(got[i] in (7, 8) and want[i] in (7, 8) for i in diff)


@pytest.mark.parametrize("widths", [(1, 4, 8), (4, 1, 8), (1, 1, 4)])
def test_older_field_widths_decode(tmp_path, panel, widths):
    from glimpse_amd import binpanel
    P, y, ap, ai, pos, ref, alt = panel
    H, V = ba.from_panel(P, pos, ref, alt, y, ap, ai)
    p = tmp_path / "o.bin"
    p.write_bytes(ba.encode(H, V, lib=7, ver_w=widths[0], iv_w=widths[1], cs_w=widths[2]))
    Q, _ = binpanel.read(p)
    _check(Q, P, y, ap, ai, pos, ref, alt)


def test_bad_archives_are_io_errors(tmp_path, panel):
    from glimpse_amd import binpanel, hip
    P, y, ap, ai, pos, ref, alt = panel
    H, V = ba.from_panel(P, pos, ref, alt, y, ap, ai)
    good = ba.encode(H, V)
    cases = {
        "sig": good.replace(b"serialization::archive", b"serialization::archivX", 1),
        "native": good[:32] + bytes([4, 4, 4, 8]) + good[36:],
        "trunc": good[: len(good) // 2],
    }
    bad = dict(H, n_rar=H["n_rar"] + 1)
    cases["counts"] = ba.encode(bad, V)
    for name, data in cases.items():
        p = tmp_path / (name + ".bin")
        p.write_bytes(data)
        with pytest.raises(hip.GhError) as e:
            binpanel.read(p)
        assert "error -6" in str(e.value), name
