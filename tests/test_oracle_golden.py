"""The oracle against the golden values held by the reference's own tests.

Each test restates one reference test case (file:line cited) with its inputs
and expected outputs; the oracle must reproduce them exactly.
"""
import pytest

import _oracle as O

AF_Q = "ATGCCCAAGCTGAATAGCGTAGAGGGGTTTTCATCATTTGAGGACGATGTATAA"
LONG_AA = ("MVQRWLYSTNAKDIAVLYFMLAIFSGMAGTAMSLIIRLELAAPGSQYLHGNSQLFNVLVVGHAVLMIFFLVMPALIGGFG"
           "NYLLPLMIGATDTAFPRINNIAFWVLPMGLVCLVTSTLVESGAGTGWTVYPPLSSIQAHSGPSVDLAIFALHLTSISSLL"
           "GAINFIVTTLNMRTNGMTMHKLPLFVWSIFITAFLLLLSLPVLSAGITMLLLDRNFNTSFFEVSGGGDPILYEHLFWFFG"
           "HPEVYILIIPGFGIISHVVSTYSKKPVFGEISMVYAMASIGLLGFLVWSHHMYIVGLDADTRAYFTSATMIIAIPTGIKI"
           "FSWLATIHGGSIRLATPMLYAIAFLFLFTMGGLTGVALANASLDVAFHDTYYVVGHFHYVLSMGAIFSLFAGYYYWSPQI"
           "LGLNYNEKLAQIQFWLIFIGANVIFFPMHFLGINGMPRRIPDYPDAFAGWNYVASIGSFIATLSLFLFIYILYDQLVNGL"
           "NNKVNNKSVIYNKAPDFVESNTIFNLNTVKSSSIEFLLTSPPAVHSFNTPAVQS")
HPEVY = "HPEVYILIIPGFGIISHVVSTYSKKPVFGEISMVYAMASIGLLGFLVWSHHMYIVGLDADTRAYFTSATMIIAIPTGIKI"

LANES = {"avx2": (32, 16), "sse": (16, 8)}


def db(name):
    return [s for _, s in O.read_fasta(O.data(name))]


# ------------------------------------------------------------------ matrices
# tests/test_matrices.c:46-111: entries [33] (A,A) and [48] (A,R)
@pytest.mark.parametrize("name,aa,ar", [("blosum45", 5, -2), ("blosum50", 5, -2), ("blosum62", 4, -1),
                                        ("blosum80", 5, -2), ("blosum90", 5, -2), ("pam30", 6, -7),
                                        ("pam70", 5, -4), ("pam250", 2, -2)])
def test_matrix_builtin_spots(name, aa, ar):
    m = O.matrix_builtin(name)
    assert m.s[33] == aa and m.s[48] == ar


def test_matrix_from_file_blosum90():
    # tests/test_matrices.c:113-121
    m = O.matrix_text(open(O.data("blosum90.txt")).read())
    assert m.s[33] == 5 and m.s[48] == -2


def test_matrix_constant():
    # tests/test_matrices.c:123-135
    m = O.matrix_constant(4, -2)
    assert [m.s[i] for i in (32, 33, 34, 64, 66)] == [-1, 4, -2, -1, 4]


# ------------------------------------------------------------------ heap
def test_heap_add_size_elements():
    # tests/util/test_minheap.c:33-61
    r = O.heap_run(5, [(3, 11, 1), (4, 22, 2), (2, 33, 3), (5, 44, 4), (1, 55, 5)])
    assert len(r) == 5 and r[0] == (5, 44, 4)


def test_heap_add_more():
    # tests/util/test_minheap.c:63-90
    r = O.heap_run(5, [(i, i, i) for i in range(20)])
    assert [x[0] for x in r] == [19, 18, 17, 16, 15] and r[0] == (19, 19, 19)


# ------------------------------------------------------------------ 64 bit
def test_sw64_simple():
    # tests/algo/64/test_smith_waterman_63.c:33-52
    m = O.matrix_constant(1, -1)
    q = O.map_query(O.NUCLEOTIDE, "AT")
    d = O.map_db(O.NUCLEOTIDE, "AATG")
    assert O.full(O.SW, d, q, m, -1, -1) == 2


def test_sw64_simple_2():
    # tests/algo/64/test_smith_waterman_63.c:54-73: the DB sequence is cut to 4 symbols
    m = O.matrix_constant(1, -1)
    q = O.map_query(O.NUCLEOTIDE, "ATGCAAA")
    d = O.map_db(O.NUCLEOTIDE, "ATGC")
    assert O.full(O.SW, d, q, m, -1, -1) == 4


@pytest.mark.parametrize("t,expect", [(O.SW, [8, 8, 8, 8, 8]), (O.NW, [-43, -50, -52, -52, -147])])
def test_search64_more_sequences(t, expect):
    # tests/algo/64/test_search_64.c:68-114 (chunk size = hit count = 5)
    m = O.matrix_constant(1, -1)
    hits, st = O.search(t, 64, O.NUCLEOTIDE, 1, [O.map_query(O.NUCLEOTIDE, AF_Q)], db("test.fas"), 5, m, -1, -1, chunk=5)
    assert [h[0] for h in hits] == expect
    assert hits[0][1] == 4 and hits[0][3] == 0
    assert st.chunk_count == 1 and st.seq_count == 5


# ------------------------------------------------------------------ searcher matrix
@pytest.mark.parametrize("bw", [8, 16, 64])
@pytest.mark.parametrize("isa", ["avx2", "sse"])
@pytest.mark.parametrize("t,score", [(O.SW, 2), (O.NW, -2)])
def test_searcher_simple(bw, isa, t, score):
    # tests/algo/test_searcher.c:64-137
    m = O.matrix_constant(1, -1)
    l8, l16 = LANES[isa]
    hits, st = O.search(t, bw, O.NUCLEOTIDE, 1, [O.map_query(O.NUCLEOTIDE, "AT")], db("short_db.fas"), 1, m, -1, -1,
                        chunk=1, lanes8=l8, lanes16=l16)
    assert st.chunk_count == 1 and st.seq_count == 1
    assert hits == [(score, 0, 0, 0)]


MULTI_NW = [(-1, 0), (-3, 3), (-4, 7), (-4, 6), (-4, 4), (-5, 2), (-6, 5), (-6, 1)]
MULTI_SW = [(4, 3), (4, 1), (4, 0), (3, 7), (3, 5), (3, 4), (3, 2), (2, 6)]


@pytest.mark.parametrize("bw", [8, 16, 64])
@pytest.mark.parametrize("isa", ["avx2", "sse"])
@pytest.mark.parametrize("t,expect", [(O.SW, MULTI_SW), (O.NW, MULTI_NW)])
def test_searcher_multiple_sequences(bw, isa, t, expect):
    # tests/algo/test_searcher.c:139-227 (chunk size = hit count = 8)
    m = O.matrix_constant(1, -1)
    l8, l16 = LANES[isa]
    hits, _ = O.search(t, bw, O.NUCLEOTIDE, 1, [O.map_query(O.NUCLEOTIDE, "ATGCAAATTT")], db("short_nuc_db.fas"), 8, m,
                       -1, -1, chunk=8, lanes8=l8, lanes16=l16)
    assert [(h[0], h[1]) for h in hits] == expect


@pytest.mark.parametrize("bw", [8, 16, 64])
@pytest.mark.parametrize("t,score", [(O.SW, 3), (O.NW, -87)])
def test_searcher_aa_constant(bw, t, score):
    # tests/algo/test_searcher.c:351-431
    m = O.matrix_constant(1, -1)
    hits, _ = O.search(t, bw, O.AMINOACID, 1, [O.map_query(O.AMINOACID, HPEVY)], db("short_AA.fas"), 1, m, -1, -1,
                       chunk=1, lanes8=16, lanes16=8)
    assert hits == [(score, 0, 0, 0)]


@pytest.mark.parametrize("bw,isa", [(64, "avx2"), (16, "sse"), (8, "sse"), (8, "avx2")])
@pytest.mark.parametrize("t,score", [(O.SW, 103), (O.NW, 82)])
def test_searcher_aa_blosum62(bw, isa, t, score):
    # tests/algo/test_searcher.c:433-505
    m = O.matrix_builtin("blosum62")
    l8, l16 = LANES[isa]
    hits, _ = O.search(t, bw, O.AMINOACID, 1, [O.map_query(O.AMINOACID, HPEVY)], db("short_AA.fas"), 1, m, -1, -1,
                       chunk=1, lanes8=l8, lanes16=l16)
    assert hits == [(score, 0, 0, 0)]


@pytest.mark.parametrize("t", [O.SW, O.NW])
def test_searcher_overflow_to_64bit(t):
    # tests/algo/test_searcher.c:507-546
    m = O.matrix_constant(127, -1)
    qraw = O.read_fasta(O.data("NP_009305.1.fas"))[0][1]
    hits, st = O.search(t, 8, O.AMINOACID, 1, [O.map_query(O.AMINOACID, qraw)], db("NP_009305.1.fas"), 1, m, -1, -1,
                        chunk=1)
    assert st.overflow_8 == 1 and st.overflow_16 == 1
    assert hits == [(67818, 0, 0, 0)]


# ------------------------------------------------------------------ 8-bit kernel files
@pytest.mark.parametrize("isa", ["avx2", "sse"])
def test_sw8_simple(isa):
    # tests/algo/8/test_8_simd_avx2_sw.c:59-95, test_8_simd_sse41_sw.c:67-93
    m = O.matrix_constant(1, -1)
    l8, l16 = LANES[isa]
    for qs, f, score in [("AT", "short_db.fas", 2), ("ATGCAAA", "tmp.fas", 4)]:
        hits, _ = O.search(O.SW, 8, O.NUCLEOTIDE, 1, [O.map_query(O.NUCLEOTIDE, qs)], db(f), 1, m, -1, -1, chunk=1,
                           lanes8=l8, lanes16=l16)
        assert hits[0][0] == score


def test_sw8_avx2_overflow():
    # tests/algo/8/test_8_simd_avx2_sw.c:97-124
    m = O.matrix_constant(1, -1)
    hits, st = O.search(O.SW, 8, O.AMINOACID, 1, [O.map_query(O.AMINOACID, LONG_AA)], db("NP_009305.1.fas"), 1, m,
                        -1, -1, chunk=1)
    assert hits[0][0] == 534 and st.overflow_8 == 1 and st.overflow_16 == 0


def test_sw8_sse41_overflow():
    # tests/algo/8/test_8_simd_sse41_sw.c:106-116: match score 30
    m = O.matrix_constant(30, -1)
    hits, st = O.search(O.SW, 8, O.AMINOACID, 1, [O.map_query(O.AMINOACID, "MVQRWLYSTN")], db("short_AA.fas"), 1, m,
                        -1, -1, chunk=1, lanes8=16, lanes16=8)
    assert hits[0][0] == 300 and st.overflow_8 == 1 and st.overflow_16 == 0


@pytest.mark.parametrize("isa", ["avx2", "sse"])
def test_sw8_more_sequences(isa):
    # tests/algo/8/test_8_simd_avx2_sw.c:126-142
    m = O.matrix_constant(1, -1)
    l8, l16 = LANES[isa]
    hits, _ = O.search(O.SW, 8, O.NUCLEOTIDE, 1, [O.map_query(O.NUCLEOTIDE, AF_Q)], db("test.fas"), 5, m, -1, -1,
                       chunk=5, lanes8=l8, lanes16=l16)
    assert hits[0][1] == 4 and hits[0][3] == 0 and [h[0] for h in hits] == [8] * 5


@pytest.mark.parametrize("isa", ["avx2", "sse"])
def test_nw8(isa):
    # tests/algo/8/test_8_simd_avx2_nw.c:59-95
    m = O.matrix_constant(1, -1)
    l8, l16 = LANES[isa]
    hits, _ = O.search(O.NW, 8, O.NUCLEOTIDE, 1, [O.map_query(O.NUCLEOTIDE, "AT")], db("short_db.fas"), 1, m, -1, -1,
                       chunk=1, lanes8=l8, lanes16=l16)
    assert hits[0][0] == -2
    hits, st = O.search(O.NW, 8, O.NUCLEOTIDE, 1, [O.map_query(O.NUCLEOTIDE, AF_Q)], db("test.fas"), 5, m, -1, -1,
                        chunk=5, lanes8=l8, lanes16=l16)
    assert st.overflow_8 == 4 and st.overflow_16 == 0
    assert [h[0] for h in hits] == [-43, -50, -52, -52, -147]


# ------------------------------------------------------------------ bigger database
AF_NW = [(-88, 908), (-92, 1050), (-94, 938), (-98, 378), (-102, 75), (-104, 23), (-106, 1229), (-110, 1148),
         (-110, 612), (-112, 1016)]
AF_SW = [(20, 1177), (18, 384), (18, 277), (18, 215), (18, 214), (18, 210), (18, 151), (18, 126), (18, 114),
         (18, 101)]


@pytest.mark.parametrize("bw", [64, 16, 8])
@pytest.mark.parametrize("t,expect", [(O.NW, AF_NW), (O.SW, AF_SW)])
def test_af091148(bw, t, expect):
    # tests/test_bigger_databases.c:70-92: constant 2/-2, gaps -4/-2, one thread, chunk 1000
    m = O.matrix_constant(2, -2)
    q = O.map_query(O.NUCLEOTIDE, O.read_fasta(O.data("one_seq.fas"))[0][1])
    hits, _ = O.search(t, bw, O.NUCLEOTIDE, 1, [q], db("AF091148.fas"), 10, m, -4, -2)
    assert [(h[0], h[1]) for h in hits] == expect


# ------------------------------------------------------------------ CIGAR / region
def nt(s):
    return O.map_db(O.NUCLEOTIDE, s)


@pytest.mark.parametrize("a,b,expect", [("AATG", "AT", "MDMD"), ("AT", "AATG", "MIMI"),
                                        ("TACGGGTAT", "GGACGTACG", "MI3M2D3MI")])
def test_cigar_nw(a, b, expect):
    # tests/algo/test_cigar.c:55-95
    m = O.matrix_constant(1, -1)
    c, _ = O.cigar(O.NW, nt(a), nt(b), m, -1, -1)
    assert c == expect


@pytest.mark.parametrize("a,b,expect", [("AATG", "AT", "2M"), ("AT", "AATG", "2M"), ("TACGGGTAT", "GGACGTACG", "4M")])
def test_cigar_sw(a, b, expect):
    # tests/algo/test_cigar.c:97-134
    m = O.matrix_constant(1, -1)
    c, _ = O.cigar(O.SW, nt(a), nt(b), m, -1, -1)
    assert c == expect


def test_region_local():
    # tests/algo/test_align.c:49-61
    m = O.matrix_constant(1, -1)
    assert O.region_local(nt("AATC"), nt("AT"), m, -1, -1) == (1, 2, 0, 1)


def test_aligner_cases():
    # tests/algo/test_aligner.c:85-218 (region printed as q_start,q_end,d_start,d_end)
    m = O.matrix_constant(1, -1)
    q = O.map_query(O.NUCLEOTIDE, "AT")
    c, r = O.cigar(O.SW, q, nt("AATG"), m, -1, -1)
    assert c == "2M" and r == (0, 1, 1, 2)
    c, r = O.cigar(O.NW, q, nt("AATG"), m, -1, -1)
    assert c == "MIMI" and r == (0, 1, 0, 3)
    q = O.map_query(O.NUCLEOTIDE, "ATGCAAAAA")
    c, _ = O.cigar(O.SW, q, nt(db("one_seq_db.fas")[0]), m, -1, -1)
    assert c == "9M"
    q = O.map_query(O.NUCLEOTIDE, AF_Q)
    c, r = O.cigar(O.SW, q, nt(db("test.fas")[0]), m, -1, -1)
    assert c == "20MD4M" and r == (10, 34, 31, 54)


LIBSSA_SW = [(877, 91, "2MD2MIM2I3MD3MD4M2D8M5I3M5IM3I7MI7M3I5MI3M"),
             (847, 91, "2MD2MIM2I3MD3MD4M2D8M5I3M5IM3I7MI7M3I4MI4M"),
             (753, 91, "2MD2MIM2I3MD3MD3M2D9M5I3M5IM3I7MI7M3I5MI3M"),
             (565, 91, "2MD2MIM2I3MD3MD4M2D8M5I3M5IM3I7MI7M3I5MI3M"),
             (398, 91, "2MD2MIM2I3MD3MD3M2D9M5I3M5IM3I7MI7M3I5MI3M")]
LIBSSA_NW = [(1050, 33, "2MI2MI6M3I6MD3M19I4M3IMI3M5I3M5IM3I7MI7M4IMIMI4MI2M7I"),
             (908, 28, "2MI2MI6M3I6MD3M5IM2I2M3I2M5IMI9MD2M3I3M9I2M3I2M4I2M2IMI4MI2M7I"),
             (938, 24, "2MI2MI6M3I6MD3M5IM2I2M3I2M5IMI9MD2M3I3M9I2M3I2M4IM3IMIMIM2I5M10I"),
             (378, 21, "2MI2MI6M3I5M9I2M7IM6IMI3M3I6M5I3M5IM3I7MI7M4IMIMI4MI2M7I"),
             (75, 12, "2MI2MI6M3I5M9I2M7IM6IMI3M3I6M5I3M5IM3I7MI7M4IMIMI4MI2M7I")]


@pytest.mark.parametrize("t,expect", [(O.SW, LIBSSA_SW), (O.NW, LIBSSA_NW)])
def test_libssa_one_thread_with_cigars(t, expect):
    # tests/test_libssa.c:48-93: constant 5/-4, gaps -4/-2, 64 bit, COMPUTE_ALIGNMENT
    m = O.matrix_constant(5, -4)
    q = O.map_query(O.NUCLEOTIDE, O.read_fasta(O.data("one_seq.fas"))[0][1])
    dbs = db("AF091148.fas")
    hits, _ = O.search(t, 64, O.NUCLEOTIDE, 1, [q], dbs, 5, m, -4, -2)
    got = []
    for score, did, strand, qid in hits:
        c, _ = O.cigar(t, q, nt(dbs[did]), m, -4, -2)
        got.append((did, score, c))
    assert got == expect
