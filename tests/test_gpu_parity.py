"""GPU search (through the libssa.h drop-in API) against the oracle.

Scores, hit selection and order, strands and CIGAR strings must be identical
to the reference run with one thread (the oracle restates it and is pinned by
the reference's goldens and its own compiled kernels).
"""
import os
import random
import tempfile

import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

ssa = pytest.importorskip("libssa_amd")

LANES = {"avx2": (32, 16, 2), "sse": (16, 8, 1)}  # lanes8, lanes16, COMPUTE_ON_*


def write_fasta(path, seqs):
    with open(path, "w") as f:
        for i, s in enumerate(seqs):
            f.write(">s%d\n" % i)
            for k in range(0, len(s), 60):
                f.write(s[k:k + 60] + "\n")
            if not s:
                f.write("\n")


def configure(symtype, strands, gaps, matrix, isa="avx2", chunk=1000):
    ssa.reset_settings()  # init_symbol_translation tests the type in force (src/libssa.c:146-151)
    ssa.set_output_mode(ssa.OUTPUT_ERROR)
    ssa.init_symbol_translation(symtype, strands, 1, 1)
    ssa.set_simd_compute_mode(LANES[isa][2])
    ssa.set_chunk_size(chunk)
    ssa.init_gap_penalties(*gaps)
    kind, arg = matrix
    if kind == "const":
        ssa.init_constant_scores(*arg)
    elif kind == "builtin":
        ssa.init_score_matrix(ssa.MATRIX_BUILDIN, arg)
    else:
        ssa.init_score_matrix(ssa.READ_FROM_STRING, arg)


def oracle_matrix(matrix):
    kind, arg = matrix
    if kind == "const":
        return O.matrix_constant(*arg)
    if kind == "builtin":
        return O.matrix_builtin(arg)
    return O.matrix_text(arg)


def gpu_search(t, q, k, bw, align):
    f = ssa.sw_align if t == O.SW else ssa.nw_align
    return f(q, k, bw, align)


def check(t, bw, symtype, strands, qraw, db, k, matrix, gaps, isa="avx2", chunk=1000, align=False):
    """Runs the GPU search and the oracle on the same inputs and compares everything."""
    configure(symtype, strands, gaps, matrix, isa, chunk)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, qraw)
        got = gpu_search(t, q, k, bw, ssa.COMPUTE_ALIGNMENT if align else ssa.COMPUTE_SCORE)
        stats = ssa.last_stats()
        ssa.free_sequence(q)
    m = oracle_matrix(matrix)
    qbufs = O.query_buffers(symtype, strands, qraw)
    l8, l16, _ = LANES[isa]
    want, st = O.search(t, bw, symtype, strands, qbufs, db, k, m, gaps[0], gaps[1], chunk=chunk, lanes8=l8,
                        lanes16=l16)
    qstrands = [s for s in range(2) if (s + 1) & strands] if symtype == O.NUCLEOTIDE else [0]
    assert [(a.score, a.db_id, a.db_frame, qstrands.index(a.query_strand)) for a in got] == \
        [(s, i, strand, qi) for s, i, strand, qi in want]
    assert (stats.overflow_8_bit_count, stats.overflow_16_bit_count) == (st.overflow_8, st.overflow_16)
    assert (stats.chunk_count, stats.seq_count) == (st.chunk_count, st.seq_count)
    if align:
        for a, (s, i, strand, qi) in zip(got, want):
            dseq = O.map_db(symtype, db[i])
            if t == O.SW and s <= 0:
                continue
            if t == O.SW and O.region_local(qbufs[qi], dseq, m, gaps[0], gaps[1]) is None:
                # no positive cell against the attached sequence: the
                # reference's region is undefined, this build reports none
                assert (a.alignment or "", a.align_q_start, a.align_q_end, a.align_d_start, a.align_d_end) == \
                    ("", 0, 0, 0, 0)
                continue
            c, r = O.cigar(t, qbufs[qi], dseq, m, gaps[0], gaps[1])
            assert a.alignment == c
            assert (a.align_q_start, a.align_q_end, a.align_d_start, a.align_d_end) == r
            assert a.db_seq == dseq
    return got


def golden_db(name):
    return [s for _, s in O.read_fasta(O.data(name))]


# ------------------------------------------------------------------ reference goldens
AF_NW = [(-88, 908), (-92, 1050), (-94, 938), (-98, 378), (-102, 75), (-104, 23), (-106, 1229), (-110, 1148),
         (-110, 612), (-112, 1016)]
AF_SW = [(20, 1177), (18, 384), (18, 277), (18, 215), (18, 214), (18, 210), (18, 151), (18, 126), (18, 114),
         (18, 101)]


@pytest.mark.parametrize("bw", [64, 16, 8])
@pytest.mark.parametrize("t,expect", [(O.NW, AF_NW), (O.SW, AF_SW)])
def test_af091148_bigger_database(bw, t, expect):
    # tests/test_bigger_databases.c:70-92
    qraw = O.read_fasta(O.data("one_seq.fas"))[0][1]
    got = check(t, bw, O.NUCLEOTIDE, 1, qraw, golden_db("AF091148.fas"), 10, ("const", (2, -2)), (-4, -2))
    assert [(a.score, a.db_id) for a in got] == expect


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
def test_public_api_cigars(t, expect):
    # tests/test_libssa.c:48-93 (one thread, 64 bit, COMPUTE_ALIGNMENT)
    configure(O.NUCLEOTIDE, 1, (-4, -2), ("const", (5, -4)))
    ssa.init_db(O.data("AF091148.fas"))
    q = ssa.init_sequence_fasta(ssa.READ_FROM_FILE, O.data("one_seq.fas"))
    got = gpu_search(t, q, 5, 64, ssa.COMPUTE_ALIGNMENT)
    ssa.free_sequence(q)
    assert [(a.db_id, a.score, a.alignment) for a in got] == expect


@pytest.mark.parametrize("t", [O.SW, O.NW])
def test_overflow_to_64bit(t):
    # tests/algo/test_searcher.c:507-546
    qraw = O.read_fasta(O.data("NP_009305.1.fas"))[0][1]
    got = check(t, 8, O.AMINOACID, 1, qraw, golden_db("NP_009305.1.fas"), 1, ("const", (127, -1)), (-1, -1), chunk=1)
    assert [(a.score, a.db_id) for a in got] == [(67818, 0)]
    st = ssa.last_stats()
    assert st.overflow_8_bit_count == 1 and st.overflow_16_bit_count == 1


MULTI_NW = [(-1, 0), (-3, 3), (-4, 7), (-4, 6), (-4, 4), (-5, 2), (-6, 5), (-6, 1)]
MULTI_SW = [(4, 3), (4, 1), (4, 0), (3, 7), (3, 5), (3, 4), (3, 2), (2, 6)]


@pytest.mark.parametrize("bw", [8, 16, 64])
@pytest.mark.parametrize("isa", ["avx2", "sse"])
@pytest.mark.parametrize("t,expect", [(O.SW, MULTI_SW), (O.NW, MULTI_NW)])
def test_searcher_multiple_sequences(bw, isa, t, expect):
    # tests/algo/test_searcher.c:139-227
    got = check(t, bw, O.NUCLEOTIDE, 1, "ATGCAAATTT", golden_db("short_nuc_db.fas"), 8, ("const", (1, -1)),
                (-1, -1), isa=isa, chunk=8)
    assert [(a.score, a.db_id) for a in got] == expect


@pytest.mark.parametrize("bw", [8, 16, 64])
@pytest.mark.parametrize("t,score", [(O.SW, 103), (O.NW, 82)])
def test_blosum62_short_aa(bw, t, score):
    # tests/algo/test_searcher.c:433-505
    got = check(t, bw, O.AMINOACID, 1, "HPEVYILIIPGFGIISHVVSTYSKKPVFGEISMVYAMASIGLLGFLVWSHHMYIVGLDADTRAYFTSATMIIAIPTGIKI",
                golden_db("short_AA.fas"), 1, ("builtin", "blosum62"), (-1, -1), chunk=1)
    assert got[0].score == score


def test_nw8_test_fas_overflow_counts():
    # tests/algo/8/test_8_simd_avx2_nw.c:78-95
    got = check(O.NW, 8, O.NUCLEOTIDE, 1, "ATGCCCAAGCTGAATAGCGTAGAGGGGTTTTCATCATTTGAGGACGATGTATAA",
                golden_db("test.fas"), 5, ("const", (1, -1)), (-1, -1), chunk=5)
    assert [a.score for a in got] == [-43, -50, -52, -52, -147]
    assert ssa.last_stats().overflow_8_bit_count == 4


# ------------------------------------------------------------------ randomized parity
def rand_db(rnd, symtype, n, lo, hi, homolog_of=None, frac=0.0):
    alpha = "ACGT" if symtype == O.NUCLEOTIDE else "ARNDCQEGHILKMFPSTWYV"
    out = []
    for _ in range(n):
        if homolog_of and rnd.random() < frac:
            s = list(homolog_of[rnd.randint(0, len(homolog_of) // 3):])
            for k in range(len(s)):
                if rnd.random() < 0.15:
                    s[k] = rnd.choice(alpha)
            out.append("".join(s[:rnd.randint(1, len(s))]))
        else:
            out.append("".join(rnd.choice(alpha) for _ in range(rnd.randint(lo, hi))))
    return out


# SSA_PARITY_SEEDS=N runs N random cases (default 24) from seed
# SSA_PARITY_SEED0 (default 0); cases past the first 24 also draw longer
# queries, which take several passes of the score kernel, and force a random
# kernel shape
_SEED0 = int(os.environ.get("SSA_PARITY_SEED0", "0"))
CASES = list(range(_SEED0, _SEED0 + int(os.environ.get("SSA_PARITY_SEEDS", "24"))))


@pytest.mark.parametrize("seed", CASES)
def test_random_parity(seed):
    rnd = random.Random(9000 + seed)
    symtype = O.AMINOACID if seed % 3 else O.NUCLEOTIDE
    strands = 1 if symtype == O.AMINOACID else rnd.choice([1, 2, 3])
    t = O.SW if seed % 2 == 0 else O.NW
    bw = [8, 16, 64][seed % 3 if seed < 12 else rnd.randrange(3)]
    isa = "avx2" if seed % 4 else "sse"
    alpha = "ACGT" if symtype == O.NUCLEOTIDE else "ARNDCQEGHILKMFPSTWYV"
    qlens = [5, 17, 33, 64, 100, 150] if seed < 24 else [5, 17, 33, 64, 100, 150, 300, 600, 1100]
    qraw = "".join(rnd.choice(alpha) for _ in range(rnd.choice(qlens)))
    db = rand_db(rnd, symtype, rnd.randint(50, 700), 1, rnd.choice([40, 120, 400]), homolog_of=qraw, frac=0.1)
    if symtype == O.NUCLEOTIDE:
        matrix = ("const", rnd.choice([(1, -1), (2, -4), (5, -4), (2, -2)]))
    else:
        matrix = ("builtin", rnd.choice(["blosum62", "blosum50", "pam30"])) if seed % 5 else ("const", (3, -2))
    gaps = rnd.choice([(-1, -1), (-10, -1), (-4, -2), (-20, -2), (-3, -1)])
    k = rnd.choice([1, 5, 10, 50, 200])
    chunk = rnd.choice([1000, 37, 250])
    # past the default cases, a forced kernel shape (NW takes 16 rows where 32 are forced)
    shapes = [(8, 4), (16, 4), (32, 4), (64, 4), (8, 8), (16, 8), (32, 8), (64, 8), (8, 16), (16, 16), (32, 16),
              (8, 32), (16, 32)]
    ssa.set_kernel_shape(*(rnd.choice(shapes) if seed >= 24 else (0, 0)))
    # and a forced column tile (shapes without 8-column tiles keep 4)
    ssa.set_kernel_variant(rnd.choice([0, 4, 8]) if seed >= 24 else 0, 0)
    # the SW first tier in int16 instead of fp16 windows: three of the default
    # cases, a random half past them
    ssa.set_cell_arithmetic((seed % 8 == 6) if seed < 24 else rnd.randrange(2))
    try:
        check(t, bw, symtype, strands, qraw, db, k, matrix, gaps, isa=isa, chunk=chunk, align=(seed % 4 == 1))
    finally:
        ssa.set_kernel_shape(0, 0)
        ssa.set_kernel_variant(0, 0)
        ssa.set_cell_arithmetic(0)


@pytest.mark.parametrize("t", [O.SW, O.NW])
@pytest.mark.parametrize("bw", [8, 16, 64])
def test_edge_inputs(t, bw):
    """Empty records among the sequences (the adapter skips them,
    src/db_adapter.c:231-233, and the IDs keep counting them), one-residue
    query and entries, a hit count above the number of entries, k = 1, and a
    DB of empty records only."""
    rnd = random.Random(61 + bw + 7 * t)
    db = ["", "A", "", "MKV", "".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(70)), "", "W"]
    db += rand_db(rnd, O.AMINOACID, 20, 1, 60)
    blosum = ("builtin", "blosum62")
    check(t, bw, O.AMINOACID, 1, "M", db, 500, blosum, (-10, -1), align=True)
    check(t, bw, O.AMINOACID, 1, "MKVW" * 10, db, 1, blosum, (-10, -1), align=True)
    assert check(t, bw, O.AMINOACID, 1, "MKV", ["", "", ""], 5, blosum, (-10, -1)) == []


def test_traceback_tie_across_query_passes():
    """Two equal best local alignments whose end cells sit on rows 512 apart
    (the same lane of the traceback kernel, in its first and second pass of
    512 query rows), the second pass's at the smaller DB column: the region
    is the first maximum in (DB column, query row) order, the second pass's."""
    rnd = random.Random(515)
    m1 = "".join(rnd.choice("GT") for _ in range(30))
    m2 = "".join(rnd.choice("GT") for _ in range(30))
    q = "A" * 100 + m1 + "A" * 482 + m2 + "A" * 258
    db = ["C" * 50 + m2 + "C" * 220 + m1 + "C" * 70]
    db += ["".join(rnd.choice("AC") for _ in range(rnd.randint(10, 300))) for _ in range(20)]
    got = check(O.SW, 16, O.NUCLEOTIDE, 1, q, db, 3, ("const", (1, -1)), (-3, -1), align=True)
    assert (got[0].db_id, got[0].score, got[0].alignment) == (0, 30, "30M")
    assert (got[0].align_q_end, got[0].align_d_end) == (641, 79)


@pytest.mark.parametrize("shape", [(8, 4), (8, 16), (16, 8), (32, 4), (32, 16), (64, 4), (64, 8),
                                   (8, 32), (16, 32)])
def test_kernel_shapes_do_not_change_results(shape):
    rnd = random.Random(77)
    qraw = "".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(131))
    db = rand_db(rnd, O.AMINOACID, 400, 1, 300, homolog_of=qraw, frac=0.2)
    ssa.set_kernel_shape(*shape)
    try:
        for t, bw in [(O.SW, 8), (O.NW, 16), (O.NW, 8)]:
            check(t, bw, O.AMINOACID, 1, qraw, db, 20, ("builtin", "blosum62"), (-10, -1))
    finally:
        ssa.set_kernel_shape(0, 0)


@pytest.mark.parametrize("shape", [(8, 32), (16, 32)])
def test_32_row_shapes_over_passes(shape):
    # 32 rows per lane (SW only; NW runs 16): several passes, pair boundaries
    # handed between them, 8-bit saturation and the 16-bit tier
    rnd = random.Random(91)
    qraw = "".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(1100))
    db = rand_db(rnd, O.AMINOACID, 300, 1, 900, homolog_of=qraw, frac=0.15)
    ssa.set_kernel_shape(*shape)
    try:
        for t, bw in [(O.SW, 8), (O.SW, 16), (O.NW, 16)]:
            check(t, bw, O.AMINOACID, 1, qraw, db, 30, ("builtin", "blosum62"), (-10, -1))
        check(O.SW, 8, O.AMINOACID, 1, qraw, db, 30, ("const", (5, -4)), (-3, -1))
    finally:
        ssa.set_kernel_shape(0, 0)


@pytest.mark.parametrize("shape", [(8, 4), (8, 8), (8, 16), (16, 4), (16, 8), (16, 16)])
def test_column_tile_of_8(shape):
    """Two 4-column blocks per step: pair starts, ends and the first flagged
    block fall on either block of a step, passes of a 1,100-residue query
    hand their bottom rows on per block, one-block entries start and end
    inside one step; SW, NW with and without the overflow trackers."""
    rnd = random.Random(808 + shape[0] * shape[1])
    qraw = "".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(1100))
    db = rand_db(rnd, O.AMINOACID, 300, 1, 500, homolog_of=qraw, frac=0.15)
    db += ["".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(rnd.randint(1, 9))) for _ in range(40)]
    short = qraw[:150]
    ssa.set_kernel_shape(*shape)
    ssa.set_kernel_variant(8, 0)
    try:
        for t, bw in [(O.SW, 8), (O.SW, 16), (O.NW, 16), (O.NW, 8)]:
            check(t, bw, O.AMINOACID, 1, qraw, db, 25, ("builtin", "blosum62"), (-10, -1))
            st = ssa.last_stats()
            if (t, bw) != (O.NW, 8):  # (NW 8-bit skips its first tier here: every pair overflows at once)
                assert (st.lanes_per_stream, st.rows_per_lane, st.columns_per_step) == shape + (8,)
            check(t, bw, O.AMINOACID, 1, short, db, 25, ("builtin", "blosum62"), (-10, -1), align=(bw == 16))
        check(O.SW, 8, O.AMINOACID, 1, qraw, db, 30, ("const", (5, -4)), (-3, -1))
        check(O.NW, 16, O.NUCLEOTIDE, 2, "".join(rnd.choice("ACGT") for _ in range(200)),
              ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(1, 300))) for _ in range(300)], 40,
              ("const", (2, -4)), (-20, -2))
    finally:
        ssa.set_kernel_shape(0, 0)
        ssa.set_kernel_variant(0, 0)


@pytest.mark.parametrize("shape,waves", [((8, 4), 4), ((16, 8), 37), ((64, 4), 9), ((32, 8), 1), ((8, 16), 16)])
def test_passes_run_as_one_pipeline(shape, waves):
    # a 1,300-residue query takes 6 to 41 passes at these shapes; each lane
    # moves to the next pass on its own while the lanes below finish the
    # last, the waves of a workgroup differ in length, and short DB entries
    # give waves shorter than a pass's minimum length
    rnd = random.Random(131 + waves)
    qraw = "".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(1300))
    db = rand_db(rnd, O.AMINOACID, 260, 1, 700, homolog_of=qraw, frac=0.12)
    ssa.set_kernel_shape(*shape)
    ssa.set_stream_waves(waves)
    try:
        for t, bw in [(O.SW, 8), (O.NW, 16), (O.NW, 8), (O.SW, 16)]:
            check(t, bw, O.AMINOACID, 1, qraw, db, 25, ("builtin", "blosum62"), (-10, -1))
    finally:
        ssa.set_kernel_shape(0, 0)
        ssa.set_stream_waves(0)


@pytest.mark.parametrize("waves", [1, 4, 37])
def test_stream_packing_does_not_change_results(waves):
    """Few waves put many entries end to end in each stream."""
    rnd = random.Random(78 + waves)
    qraw = "".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(300))
    db = rand_db(rnd, O.AMINOACID, 500, 1, 400, homolog_of=qraw, frac=0.2)
    ssa.set_stream_waves(waves)
    ssa.set_kernel_shape(32, 4)
    try:
        for t, bw in [(O.SW, 8), (O.SW, 16), (O.NW, 16), (O.NW, 8)]:
            check(t, bw, O.AMINOACID, 1, qraw, db, 20, ("builtin", "blosum62"), (-10, -1), align=(bw == 16))
    finally:
        ssa.set_stream_waves(0)
        ssa.set_kernel_shape(0, 0)


def _revcomp_nt(s):
    return "".join({"A": "T", "C": "G", "G": "C", "T": "A"}[c] for c in reversed(s))


@pytest.mark.parametrize("t,bw", [(O.NW, 16), (O.NW, 64), (O.SW, 16), (O.SW, 64)])
def test_both_strands_reverse_from_forward_twins(t, bw):
    """Both-strand DNA at 16/64 bit where no int16 value can reach either end
    of the range: only the forward query strand is scored and the reverse
    strand's results are read from each entry's reverse-complement twin.  Hit
    lists (scores, IDs, strands, query strands), counters and CIGARs equal the
    oracle's, which scores both strands in full; the scored cells show that
    only one strand ran."""
    rnd = random.Random(3100 + bw + 7 * t)
    q = "".join(rnd.choice("ACGT") for _ in range(250))
    db = []
    for i in range(400):
        n = rnd.randint(150, 250)
        s = "".join(rnd.choice("ACGT") for _ in range(n))
        if i % 40 == 0:  # planted copies of the query, half reverse-complemented
            s = q[rnd.randint(0, 60):] if i % 80 else _revcomp_nt(q)[rnd.randint(0, 60):]
        db.append(s)
    check(t, bw, O.NUCLEOTIDE, 3, q, db, 25, ("const", (2, -4)), (-20, -2), align=True)
    st = ssa.last_stats()
    assert 2 * st.cells_scored == st.cells


@pytest.mark.parametrize("t", [O.NW, O.SW])
def test_both_strands_both_scored_past_int16_bound(t):
    """The same search with a query and entries long enough that the int16
    bound fails: both query strands are scored."""
    rnd = random.Random(3200 + t)
    q = "".join(rnd.choice("ACGT") for _ in range(700))
    db = ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(700, 900))) for _ in range(120)]
    db[7] = _revcomp_nt(q)
    check(t, 16, O.NUCLEOTIDE, 3, q, db, 10, ("const", (2, -4)), (-20, -2))
    st = ssa.last_stats()
    assert st.cells_scored == st.cells


UNTRACKED_NW_SHAPES = {(8, 8), (8, 16), (16, 8), (16, 16), (32, 8)}


@pytest.mark.parametrize("shape", [(8, 8), (16, 16), (32, 8), (16, 4), (64, 4)])
@pytest.mark.parametrize("long_case", [False, True])
def test_nw16_without_overflow_tracking(shape, long_case):
    """NW at 16 bit where the int16 bound (query length, longest entry, gap
    and matrix extremes) rules out every overflow flag: the kernel variant
    without h_min / h_max trackers runs on the shapes that have it; results,
    overflow counts and CIGARs must equal the oracle's either way.  The long
    case fails the bound and keeps the trackers."""
    rnd = random.Random(4600 + shape[0] * 7 + shape[1] + (1000 if long_case else 0))
    qlen, lo, hi, n = (1700, 1200, 1500, 40) if long_case else (230, 1, 400, 300)
    q = "".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(qlen))
    db = rand_db(rnd, O.AMINOACID, n, lo, hi, homolog_of=q, frac=0.1)
    ssa.set_kernel_shape(*shape)
    try:
        check(O.NW, 16, O.AMINOACID, 1, q, db, 15, ("builtin", "blosum62"), (-10, -1), align=True)
        st = ssa.last_stats()
    finally:
        ssa.set_kernel_shape(0, 0)
    assert (st.lanes_per_stream, st.rows_per_lane) == shape
    assert st.overflow_tracking == (1 if long_case or shape not in UNTRACKED_NW_SHAPES else 0)


@pytest.mark.parametrize("qlen,skipped", [(60, False), (140, True), (420, True)])
def test_nw8_first_block_overflow_without_the_8bit_pass(qlen, skipped):
    """NW at 8 bit: once the query is long enough, a bound on column 0 shows
    that every pair overflows in its first block, and the 8-bit launch is
    skipped (no cells scored by a first tier); hits, overflow counts and
    CIGARs must still equal the oracle's, which runs the 8-bit tier."""
    rnd = random.Random(4400 + qlen)
    base = "".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(qlen))
    db = rand_db(rnd, O.AMINOACID, 300, 1, 500, homolog_of=base, frac=0.1)
    check(O.NW, 8, O.AMINOACID, 1, base, db, 20, ("builtin", "blosum62"), (-10, -1), align=True)
    st = ssa.last_stats()
    assert (st.cells_scored == 0) == skipped


def test_nw8_first_block_overflow_both_strands():
    rnd = random.Random(4500)
    q = "".join(rnd.choice("ACGT") for _ in range(180))
    db = rand_db(rnd, O.NUCLEOTIDE, 200, 20, 300, homolog_of=q, frac=0.1)
    check(O.NW, 8, O.NUCLEOTIDE, 3, q, db, 15, ("const", (2, -4)), (-20, -2))
    assert ssa.last_stats().cells_scored == 0
