"""The pre-packed DB file (ssa_amd_db_pack) and the built-in provider reading it.

Host-only: packing maps on the CPU and the provider only memory-maps the file,
so these run without a GPU.  The file is parsed here from its documented
layout and every mapped code is compared with the oracle's DB mapping
(us_map_sequence, src/util/util_sequence.c:298-318) and reverse complement
(src/util/util_sequence.c:427-437).  Searches on packed files are compared
with searches on the FASTA file in test_gpu_batch.py.
"""
import ctypes as C
import os
import random
import struct

import pytest

import _oracle as O

ssa = pytest.importorskip("libssa_amd")

HEADER = struct.Struct("<8sIIii" + "Q" * 13)
assert HEADER.size == 128


def parse(path):
    raw = open(path, "rb").read()
    f = HEADER.unpack_from(raw, 0)
    h = dict(zip(["magic", "version", "header_bytes", "symtype", "strands", "seq_count", "entry_count",
                  "ascii_bytes", "code_bytes", "unknown_count", "off_seqs", "off_ascii", "off_entries", "off_first",
                  "off_bylen", "off_codes", "off_unknown", "file_bytes"], f))
    n, e = h["seq_count"], h["entry_count"]
    seqs = [struct.unpack_from("<QQQ", raw, h["off_seqs"] + 24 * i) for i in range(n)]
    ascii_ = raw[h["off_ascii"]:h["off_ascii"] + h["ascii_bytes"]]
    ents = [struct.unpack_from("<QIBBHq", raw, h["off_entries"] + 24 * k) for k in range(e)]
    first = list(struct.unpack_from("<%dI" % (n + 1), raw, h["off_first"]))
    bylen = list(struct.unpack_from("<%dI" % e, raw, h["off_bylen"]))
    codes = raw[h["off_codes"]:h["off_codes"] + h["code_bytes"]]
    unk = [struct.unpack_from("<QQ", raw, h["off_unknown"] + 16 * u) for u in range(h["unknown_count"])]
    return raw, h, seqs, ascii_, ents, first, bylen, codes, unk


def write_fasta(path, seqs):
    with open(path, "w") as f:
        for i, s in enumerate(seqs):
            f.write(">s%d\n" % i)
            for k in range(0, len(s), 50):
                f.write(s[k:k + 50] + "\n")
            if not s:
                f.write("\n")


def random_db(rnd, symtype, n):
    # upper and lower case, empty records, and symbols outside the alphabet
    alpha = "ACGTNacgtRY" if symtype == O.NUCLEOTIDE else "ARNDCQEGHILKMFPSTWYVXBZarndc*"
    out = []
    for i in range(n):
        if i % 17 == 5:
            out.append("")
            continue
        s = "".join(rnd.choice(alpha) for _ in range(rnd.randint(1, 300)))
        if i % 11 == 3:
            s = s[:len(s) // 2] + "#1" + s[len(s) // 2:]
        out.append(s)
    return out


def unknown_count(symtype, s):
    f = O.lib.orc_map_nt if symtype == O.NUCLEOTIDE else O.lib.orc_map_aa
    return sum(1 for ch in s.encode() if f(ch) < 0)


def setup(symtype, strands, tmp_path, seqs):
    ssa.reset_settings()
    ssa.set_output_mode(ssa.OUTPUT_ERROR)
    ssa.init_symbol_translation(symtype, strands, 1, 1)
    fas = str(tmp_path / "db.fas")
    write_fasta(fas, seqs)
    ssa.init_db(fas)
    return fas


@pytest.mark.parametrize("symtype,strands", [(O.AMINOACID, 1), (O.NUCLEOTIDE, 3), (O.NUCLEOTIDE, 1)])
def test_packed_layout_matches_oracle_mapping(tmp_path, symtype, strands):
    rnd = random.Random(7 + symtype * 10 + strands)
    seqs = random_db(rnd, symtype, 120)
    setup(symtype, strands, tmp_path, seqs)
    out = str(tmp_path / "db.ssapack")
    ssa.pack_db(out)
    raw, h, recs, ascii_, ents, first, bylen, codes, unk = parse(out)
    nuc = symtype == O.NUCLEOTIDE
    assert h["magic"] == b"SSAPACK1" and h["version"] == 1 and h["header_bytes"] == 128
    assert h["symtype"] == symtype and h["strands"] == (3 if nuc else 1)
    assert h["file_bytes"] == len(raw) and h["seq_count"] == len(seqs)
    for off in ("off_seqs", "off_ascii", "off_entries", "off_first", "off_bylen", "off_codes", "off_unknown"):
        assert h[off] % 64 == 0
    # sequences: provider IDs (FASTA: the record index) and the residues as read
    assert [r[0] for r in recs] == list(range(len(seqs)))
    assert [ascii_[o:o + n].decode() for _, o, n in recs] == seqs
    # entries: one per non-empty sequence, its reverse complement after it for nucleotides
    want = []
    want_first = []
    for i, s in enumerate(seqs):
        want_first.append(len(want))
        if not s:
            continue
        fwd = O.map_db(symtype, s)
        want.append((i, len(s), 0, fwd))
        if nuc:
            want.append((i, len(s), 1, O.revcomp(fwd)))
    want_first.append(len(want))
    assert first == want_first and h["entry_count"] == len(want)
    for (eid, elen, strand, frame, pad, off), (wid, wlen, wstrand, wcodes) in zip(ents, want):
        assert (eid, elen, strand, frame, pad) == (wid, wlen, wstrand, 0, 0)
        assert off % 4 == 0
        assert codes[off:off + elen] == wcodes
        assert codes[off + elen:off + (elen + 3) // 4 * 4] == bytes((4 - elen % 4) % 4)
    # the length order the search schedules from: longest first, ties in entry order
    assert bylen == sorted(range(len(want)), key=lambda k: -want[k][1])
    assert unk == [(i, unknown_count(symtype, s)) for i, s in enumerate(seqs) if s and unknown_count(symtype, s)]


def test_packed_provider_serves_the_same_sequences(tmp_path):
    rnd = random.Random(11)
    seqs = random_db(rnd, O.AMINOACID, 80)
    setup(O.AMINOACID, 1, tmp_path, seqs)
    want = [ssa.db_sequence(i) for i in range(len(seqs))]
    out = str(tmp_path / "db.ssapack")
    ssa.pack_db(out)
    ssa.init_db(out)
    assert ssa.lib.ssa_db_get_sequence_count() == len(seqs)
    assert [ssa.db_sequence(i) for i in range(len(seqs))] == want
    assert ssa.db_sequence(len(seqs)) is None
    # a packed file packs again to the same bytes
    again = str(tmp_path / "again.ssapack")
    ssa.pack_db(again)
    assert open(again, "rb").read() == open(out, "rb").read()


def test_packed_golden_db(tmp_path):
    # the reference's own test DB (tests/testdata/AF091148.fas, 1403 records)
    ssa.set_output_mode(ssa.OUTPUT_ERROR)
    ssa.init_symbol_translation(ssa.NUCLEOTIDE, 3, 1, 1)
    ssa.init_db(O.data("AF091148.fas"))
    want = [ssa.db_sequence(i) for i in range(1403)]
    out = str(tmp_path / "af.ssapack")
    ssa.pack_db(out)
    _, h, recs, ascii_, ents, first, bylen, codes, unk = parse(out)
    assert h["seq_count"] == 1403 and h["entry_count"] == 2 * sum(1 for _, s in want if s)
    ssa.init_db(out)
    assert [ssa.db_sequence(i) for i in range(1403)] == want


def corrupt(src, dst, at, data):
    raw = bytearray(open(src, "rb").read())
    raw[at:at + len(data)] = data
    open(dst, "wb").write(bytes(raw))


def test_corrupt_packed_files_are_rejected(tmp_path):
    rnd = random.Random(5)
    seqs = random_db(rnd, O.AMINOACID, 40)
    setup(O.AMINOACID, 1, tmp_path, seqs)
    ssa.set_output_mode(ssa.OUTPUT_SILENT)
    good = str(tmp_path / "good.ssapack")
    ssa.pack_db(good)
    _, h, recs, ascii_, ents, first, bylen, codes, unk = parse(good)
    bad = str(tmp_path / "bad.ssapack")
    cases = [
        (8, struct.pack("<I", 2)),                                          # version
        (48, struct.pack("<Q", h["code_bytes"] + 10 ** 9)),                 # codes past the end
        (h["off_entries"] + 16, struct.pack("<q", h["code_bytes"] - 2)),    # an entry's codes past the end
        (h["off_entries"] + 16, struct.pack("<q", 2)),                      # misaligned entry
        (h["off_first"] + 4, struct.pack("<I", 10 ** 6)),                   # entry table out of order
        (h["off_bylen"], struct.pack("<I", h["entry_count"])),              # length order out of range
        (h["off_bylen"], struct.pack("<I", bylen[1])),                      # not a permutation
        (h["off_seqs"] + 16, struct.pack("<Q", 10 ** 9)),                   # residues past the end
    ]
    # residue codes: the first entry's first code, a code in the middle of a
    # long entry, and a padding byte after an entry whose length is not a
    # multiple of 4 (a code of 64 or more would forge the kernels' flag bits)
    e0 = ents[0]
    long_e = max(ents, key=lambda e: e[1])
    odd = next(e for e in ents if e[1] % 4)
    cases += [
        (h["off_codes"] + e0[5], bytes([32])),                              # first code out of range
        (h["off_codes"] + long_e[5] + long_e[1] // 2, bytes([0x40])),       # a code with a flag bit
        (h["off_codes"] + odd[5] + odd[1], bytes([1])),                     # non-zero padding
    ]
    for at, data in cases:
        corrupt(good, bad, at, data)
        assert ssa.lib.ssa_db_init(bad.encode()) == -1, (at, data)
        assert ssa.lib.ssa_db_get_sequence_count() == 0
    # a nucleotide pack: codes of 16 or more are outside the nucleotide alphabet
    setup(O.NUCLEOTIDE, 3, tmp_path, random_db(rnd, O.NUCLEOTIDE, 30))
    ssa.set_output_mode(ssa.OUTPUT_SILENT)
    nt_good = str(tmp_path / "nt.ssapack")
    ssa.pack_db(nt_good)
    _, hn, _, _, ents_n, _, _, _, _ = parse(nt_good)
    corrupt(nt_good, bad, hn["off_codes"] + ents_n[1][5], bytes([16]))
    assert ssa.lib.ssa_db_init(bad.encode()) == -1
    corrupt(nt_good, bad, hn["off_codes"] + ents_n[1][5], bytes([15]))    # the largest valid code
    assert ssa.lib.ssa_db_init(bad.encode()) == 0
    ssa.init_symbol_translation(O.AMINOACID, 1, 1, 1)
    open(bad, "wb").write(open(good, "rb").read()[:-8])                     # truncated
    assert ssa.lib.ssa_db_init(bad.encode()) == -1
    assert ssa.lib.ssa_db_init(good.encode()) == 0
    assert ssa.lib.ssa_db_get_sequence_count() == len(seqs)
    ssa.set_output_mode(ssa.OUTPUT_ERROR)


@pytest.mark.parametrize("symtype", [O.TRANS_DB, O.TRANS_BOTH])
def test_pack_refuses_translated_db(tmp_path, symtype):
    setup(O.AMINOACID, 1, tmp_path, ["ACGTAC", "GGGTTT"])
    ssa.set_output_mode(ssa.OUTPUT_SILENT)
    ssa.init_symbol_translation(symtype, 3, 1, 1)
    with pytest.raises(RuntimeError):
        ssa.pack_db(str(tmp_path / "t.ssapack"))
    ssa.reset_settings()  # init_symbol_translation exits once TRANS_DB / TRANS_BOTH is set
    ssa.set_output_mode(ssa.OUTPUT_ERROR)
    assert not os.path.exists(str(tmp_path / "t.ssapack"))


def test_pack_translated_query_db_is_amino_acid(tmp_path):
    # a translated query searches the amino-acid DB: its pack is that of AMINOACID
    seqs = ["ACDEKLMN", "MKV*X", "WWYY"]
    setup(O.AMINOACID, 1, tmp_path, seqs)
    ssa.pack_db(str(tmp_path / "a.ssapack"))
    ssa.init_symbol_translation(ssa.TRANS_QUERY, 3, 1, 1)
    ssa.pack_db(str(tmp_path / "t.ssapack"))
    ssa.init_symbol_translation(ssa.AMINOACID, 1, 1, 1)
    assert open(str(tmp_path / "t.ssapack"), "rb").read() == open(str(tmp_path / "a.ssapack"), "rb").read()
    assert parse(str(tmp_path / "t.ssapack"))[1]["symtype"] == O.AMINOACID
