"""The oracle's translated modes (TRANS_QUERY, TRANS_DB, TRANS_BOTH).

Pinned three ways: the golden values of the reference's own tests (the
adapter's six frames of one_seq.fas, the TRANS_QUERY search scores), textbook
facts of the genetic codes that a wrong table index, base order or ambiguity
rule would break, and every entry of every genetic code against the
reference's compiled util_sequence.c (oracle/_ref) where it was built.
"""
import random

import pytest

import _oracle as O

ONE_SEQ = "ATGCCCAAGCTGAATAGCGTAGAGGGGTTTTCATCATTTGAGGACGATGTATAA"
AA = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"
VALID = [1, 2, 3, 4, 5, 6, 9, 10, 11, 12, 13, 14, 15, 16, 21, 22, 23]
REF = O.LazyRef()


def nt(s):
    return O.map_db(O.NUCLEOTIDE, s)


def prot(codes):
    return "".join(AA[c] for c in codes)


def tr(s, gencode=1, strand=0, frame=0):
    return prot(O.translate(O.translate_table(gencode), nt(s), strand, frame))


def test_gencode_numbers():
    # src/libssa.c:139-145 accepts the codes that src/util/util_sequence.c:95-123 names
    assert [c for c in range(0, 30) if O.lib.orc_gencode_valid(c)] == VALID


def test_db_six_frames_golden():
    # tests/test_db_adapter.c:177-290: one_seq.fas under DB genetic code 3
    t = O.translate_table(3)
    d = nt(ONE_SEQ)
    got = [prot(O.translate(t, d, s, f)) for s in range(2) for f in range(3)]
    assert got == ["MPKTNSVEGFSSFEDDV*", "CPSWMA*RGFHHLRTMY", "AQAE*RRGVFIIWGRCM",
                   "LYIVTKWWKPTYAIQTGH", "YTSSSNDENPSTTFSLG", "MHRPQMMKTPTRYSAWA"]


@pytest.mark.parametrize("codon,gencode,aa", [
    ("ATG", 1, "M"), ("TGG", 1, "W"), ("TAA", 1, "*"), ("TAG", 1, "*"), ("TGA", 1, "*"),
    ("TTT", 1, "F"), ("GGG", 1, "G"), ("CTG", 1, "L"), ("AGA", 1, "R"), ("ATA", 1, "I"),
    ("TGA", 2, "W"), ("AGA", 2, "*"), ("ATA", 2, "M"),         # vertebrate mitochondrial
    ("CTG", 3, "T"), ("CTA", 3, "T"), ("ATA", 3, "M"),         # yeast mitochondrial
    ("TAA", 6, "Q"), ("TAG", 6, "Q"), ("TGA", 6, "*"),         # ciliate nuclear
    ("AAA", 9, "N"), ("TGA", 10, "C"), ("CTG", 12, "S"),       # echinoderm, euplotid, alt. yeast
    ("AGA", 13, "G"), ("TAA", 14, "Y"), ("TAG", 15, "Q"),      # ascidian, alt. flatworm, blepharisma
    ("TAG", 16, "L"), ("TCA", 22, "*"), ("TTA", 23, "*"),      # chlorophycean, scenedesmus, thraustochytrium
])
def test_textbook_codons(codon, gencode, aa):
    assert tr(codon, gencode) == aa


@pytest.mark.parametrize("codon,gencode,aa", [
    ("GCN", 1, "A"), ("GGN", 1, "G"), ("RAY", 1, "B"), ("SAR", 1, "Z"), ("NNN", 1, "X"),
    ("YTR", 1, "L"), ("YTR", 3, "X"), ("MGR", 1, "R"), ("TAR", 1, "*"), ("TRA", 1, "*"),
    ("TRA", 2, "X"), ("AAY", 1, "N"), ("AAN", 1, "X"), ("---", 1, "X"), ("ATH", 1, "I"),
])
def test_ambiguity_codes(codon, gencode, aa):
    # an IUPAC codon gives the amino acid shared by all its codons, B for D/N,
    # Z for E/Q, X otherwise; '-' (no base) gives X
    assert tr(codon, gencode) == aa


def test_frames_and_reverse_strand_by_construction():
    rnd = random.Random(11)
    for _ in range(50):
        s = "".join(rnd.choice("ACGT") for _ in range(rnd.randint(3, 60)))
        rc = "".join({"A": "T", "C": "G", "G": "C", "T": "A"}[c] for c in reversed(s))
        for f in range(3):
            assert tr(s, strand=0, frame=f) == tr(s[f:])
            assert tr(s, strand=1, frame=f) == tr(rc[f:])
            assert len(tr(s, frame=f)) == (len(s) - f) // 3


def test_sequence_shorter_than_frame():
    # (len - frame) / 3 in size_t: the reference asks for ~6e18 bytes and exits
    t = O.translate_table(1)
    assert O.translate(t, nt("A"), 0, 2) is None
    assert O.translate(t, nt("A"), 0, 1) == b""
    assert O.translate(t, nt("AC"), 1, 2) == b""


# tests/algo/test_searcher.c:248-330 (setup :32-54): TRANS_QUERY, both strands,
# genetic code 3, constant 1/-1, gaps -1/-1, six hits against short_AA.fas
TQ = "ATGCCCAAAATTGACTTGAATAGGGGCGTGAGGAGGGGTTTTCAATACTATTTTTTG"


@pytest.mark.parametrize("bw,lanes", [(64, (32, 16)), (16, (16, 8)), (8, (16, 8))])
@pytest.mark.parametrize("t,expect", [(O.SW, [2, 2, 2, 2, 2, 1]), (O.NW, [-109, -109, -110, -110, -110, -113])])
def test_searcher_translate_query_golden(bw, lanes, t, expect):
    qb = O.query_buffers(O.TRANS_QUERY, 3, TQ, q_gencode=3)
    assert len(qb) == 6
    db = [s for _, s in O.read_fasta(O.data("short_AA.fas"))]
    hits, _ = O.search(t, bw, O.TRANS_QUERY, 3, qb, db, 6, O.matrix_constant(1, -1), -1, -1, chunk=6,
                       lanes8=lanes[0], lanes16=lanes[1])
    assert [h[0] for h in hits] == expect and all(h[1] == 0 for h in hits)


def test_search_trans_db_entries():
    # TRANS_DB both strands: six entries per DB sequence with strand 0/1 and
    # frames 0..2; one strand: three entries whose strand field is the strand
    # setting (src/db_adapter.c:72-104); a DB hit is the best frame's score
    db = [ONE_SEQ, "ACGTACGT", "GGGCCCAAATTT"]
    q = O.map_query(O.TRANS_DB, "MPKTNSVEGF")
    m = O.matrix_builtin("blosum62")
    hits, st = O.search(O.SW, 64, O.TRANS_DB, 3, [q], db, 18, m, -10, -1, d_gencode=3, frames=True)
    assert st.seq_count == 18 and len(hits) == 18
    assert sorted((h[2], h[4]) for h in hits if h[1] == 0) == [(s, f) for s in range(2) for f in range(3)]
    t = O.translate_table(3)
    for score, i, s, _, f in hits:
        assert score == O.full(O.SW, O.translate(t, nt(db[i]), s, f), q, m, -10, -1)
    for strands in (1, 2):
        hits, st = O.search(O.SW, 64, O.TRANS_DB, strands, [q], db, 9, m, -10, -1, d_gencode=3, frames=True)
        assert st.seq_count == 9 and {h[2] for h in hits} == {strands}
        for score, i, s, _, f in hits:
            assert score == O.full(O.SW, O.translate(t, nt(db[i]), strands - 1, f), q, m, -10, -1)


def test_search_trans_db_short_sequence_exits():
    q = O.map_query(O.TRANS_DB, "MPK")
    with pytest.raises(ValueError):
        O.search(O.SW, 16, O.TRANS_DB, 3, [q], ["ACGTA", "T"], 5, O.matrix_constant(1, -1), -1, -1)


@pytest.mark.parametrize("bw", [8, 16, 64])
@pytest.mark.parametrize("t", [O.SW, O.NW])
def test_search_trans_db_empty_frames(bw, t):
    # DB sequences of 2..4 nucleotides leave frames with no codon: those
    # entries still take a lane for one block of padding (scores SW 0; NW the
    # padding block's last column), and 64-bit scores them as empty sequences
    q = O.map_query(O.TRANS_DB, "MKWV")
    db = ["ACG", "AC", "ACGT", "ATGAAATGG"]
    hits, st = O.search(t, bw, O.TRANS_DB, 3, [q], db, 24, O.matrix_builtin("blosum62"), -3, -1,
                        frames=True)
    assert st.seq_count == 24 and len(hits) == 24


# ---------------------------------------------------------------- vs oracle/_ref
ALL_CODONS = "".join(a + b + c for a in "-ACMGRSVTWYHKDBN" for b in "-ACMGRSVTWYHKDBN" for c in "-ACMGRSVTWYHKDBN")


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("gencode", VALID)
def test_every_codon_matches_reference(gencode):
    # all 16^3 nt16 codons of the table in frame 0, both tables (query, DB)
    d = nt(ALL_CODONS)
    mine = O.translate(O.translate_table(gencode), d, 0, 0)
    assert O.ref_translate(REF, 0, gencode, 1, d, 0, 0) == mine
    assert O.ref_translate(REF, 1, 1, gencode, d, 0, 0) == mine


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("seed", range(10))
def test_random_frames_match_reference(seed):
    rnd = random.Random(seed)
    gq, gd = rnd.choice(VALID), rnd.choice(VALID)
    for _ in range(20):
        s = "".join(rnd.choice("ACGTACGTNRYKMSW") for _ in range(rnd.randint(2, 200)))
        d = nt(s)
        for strand in range(2):
            for f in range(3):
                assert O.ref_translate(REF, 0, gq, gd, d, strand, f) == O.translate(O.translate_table(gq), d, strand, f)
                assert O.ref_translate(REF, 1, gq, gd, d, strand, f) == O.translate(O.translate_table(gd), d, strand, f)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
@pytest.mark.parametrize("seed", range(20))
def test_lane_kernels_with_empty_entries_match_reference(seed):
    # translated DBs hand the lane kernels entries with no residue; the
    # reference runs one padding block for them
    rnd = random.Random(7000 + seed)
    if seed % 2:
        REF.ref_matrix_builtin(b"blosum62")
        m = O.matrix_builtin("blosum62")
    else:
        match, mismatch = rnd.choice([(1, -1), (5, -4), (127, -1)])
        REF.ref_matrix_constant(match, mismatch)
        m = O.matrix_constant(match, mismatch)
    go, ge = rnd.choice([(-1, -1), (-10, -1), (-3, -1), (-128, -1), (-60, -60)])
    REF.ref_set_gaps(go, ge)
    alpha = "ARNDCQEGHILKMFPSTWYV"
    seqs = [O.map_db(O.AMINOACID, "".join(rnd.choice(alpha) for _ in range(rnd.choice([0, 0, 1, 3, 9, 40]))))
            for _ in range(rnd.randint(1, 60))]
    q = O.map_db(O.AMINOACID, "".join(rnd.choice(alpha) for _ in range(rnd.randint(1, 90))))
    for t in (O.SW, O.NW):
        for width in (8, 16):
            for avx2 in (True, False):
                lanes = {(8, True): 32, (8, False): 16, (16, True): 16, (16, False): 8}[(width, avx2)]
                ref_ins, ref_ovf = O.ref_lane_kernel(REF, t, width, avx2, seqs, q)
                ev, _ = O.lane_kernel(t, width, lanes, seqs, q, m, go, ge)
                assert [(i, s) for i, k, s in ev if k == 0] == ref_ins, (t, width, avx2)
                assert [i for i, k, s in ev if k == 1] == ref_ovf, (t, width, avx2)


@pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")
def test_full64_empty_db_matches_reference():
    m = O.matrix_builtin("blosum62")
    REF.ref_matrix_builtin(b"blosum62")
    REF.ref_set_gaps(-10, -1)
    q = O.map_db(O.AMINOACID, "MKWVTFISLL")
    for t in (O.SW, O.NW):
        assert O.full(t, b"", q, m, -10, -1) == REF.ref_full(t, b"", 0, q, len(q))
