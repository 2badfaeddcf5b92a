"""The SW first tier in fp16 windows (score kernel mode 3) against the oracle
and against the same searches in packed int16.

Mode 3 keeps each lane's values relative to an offset of its own and moves
them when the window's maximum passes its trigger, every 128 columns, when
values are handed down from a lane with another offset, and per half of the
packed registers; these cases drive those paths with high-scoring homologs
(offsets of thousands, pairs past the 16-bit tier's 65,535), every SW kernel
shape and column tile, DNA and protein, and check that the search falls back
to int16 where the matrix and gaps would leave the window.
"""
import os
import random
import tempfile

import pytest

import _oracle as O
from test_gpu_parity import check, configure, write_fasta

pytestmark = pytest.mark.gpu

ssa = pytest.importorskip("libssa_amd")

AA = "ARNDCQEGHILKMFPSTWYV"

SW_SHAPES = [(8, 4), (16, 4), (32, 4), (64, 4), (8, 8), (16, 8), (32, 8), (64, 8), (8, 16), (16, 16), (32, 16),
             (8, 32), (16, 32)]


def homolog_db(rnd, qraw, n, lo, hi, copies, identity=0.9):
    """Random entries of lo..hi residues, `copies` of them replaced by mutated
    fragments of the query (long local alignments with high scores)."""
    db = ["".join(rnd.choice(AA) for _ in range(rnd.randint(lo, hi))) for _ in range(n)]
    for c in range(copies):
        a = rnd.randint(0, len(qraw) // 4)
        b = rnd.randint(3 * len(qraw) // 4, len(qraw))
        frag = [x if rnd.random() < identity else rnd.choice(AA) for x in qraw[a:b]]
        if c % 3 == 1:  # an indel in the middle
            m = len(frag) // 2
            frag = frag[:m] + frag[m + rnd.randint(1, 6):]
        db[rnd.randrange(n)] = "".join(frag)
    return db


def run_both(fn):
    """fn() under fp16 (the default) and under int16; returns the cell
    arithmetic each ran with."""
    seen = []
    try:
        for arith in (0, 1):
            ssa.set_cell_arithmetic(arith)
            fn()
            seen.append(ssa.last_stats().cell_arithmetic)
    finally:
        ssa.set_cell_arithmetic(0)
    return seen


@pytest.mark.parametrize("shape", SW_SHAPES)
def test_fp16_windows_every_shape(shape):
    """Homologs scoring up to ~8,000 (offsets past 6,000, many re-centrings)
    over every SW shape, with the 8-column tile where the shape has one."""
    rnd = random.Random(1700 + 13 * shape[0] + shape[1])
    qraw = "".join(rnd.choice(AA) for _ in range(1500))
    db = homolog_db(rnd, qraw, 160, 1, 400, 12)
    db += ["".join(rnd.choice(AA) for _ in range(rnd.randint(1, 9))) for _ in range(20)]
    ssa.set_kernel_shape(*shape)
    try:
        for cols in (4, 8):
            ssa.set_kernel_variant(cols, 0)
            seen = run_both(lambda: check(O.SW, 16, O.AMINOACID, 1, qraw, db, 25, ("builtin", "blosum62"),
                                          (-10, -1)))
            assert seen == [2, 1]
            seen = run_both(lambda: check(O.SW, 8, O.AMINOACID, 1, qraw, db, 25, ("builtin", "blosum62"),
                                          (-10, -1)))
            assert seen == [2, 1]
    finally:
        ssa.set_kernel_shape(0, 0)
        ssa.set_kernel_variant(0, 0)


@pytest.mark.parametrize("gaps", [(-10, -1), (-5, -3), (-11, -4)])
def test_fp16_long_entries_column_shift(gaps):
    """Mode 3 holds the values of DB column j less j gap extend, so over long
    entries the held values rise by one extend per column even where the real
    ones stay low: unrelated entries of 3,000-9,000 residues re-centre their
    windows on that rise alone, and homologs that start after thousands of
    such columns (or end thousands of columns before the entry does) score
    across many re-centrings and 128-column re-centrings down; extends of 1,
    3 and 4, in the 16- and 8-bit cascades, against the oracle."""
    rnd = random.Random(4100 - 7 * gaps[1])
    qraw = "".join(rnd.choice(AA) for _ in range(400))
    db = ["".join(rnd.choice(AA) for _ in range(rnd.randint(3000, 9000))) for _ in range(24)]
    for c in range(6):
        a, b = rnd.randint(0, 100), rnd.randint(250, 400)
        frag = "".join(x if rnd.random() < 0.85 else rnd.choice(AA) for x in qraw[a:b])
        pre = "".join(rnd.choice(AA) for _ in range(rnd.randint(1500, 6000)))
        post = "".join(rnd.choice(AA) for _ in range(rnd.randint(0, 3000) if c % 2 else 0))
        db.insert(rnd.randrange(len(db)), pre + frag + post)
    db += ["".join(rnd.choice(AA) for _ in range(rnd.randint(20, 300))) for _ in range(30)]
    for bw in (16, 8):
        seen = run_both(lambda: check(O.SW, bw, O.AMINOACID, 1, qraw, db, 12, ("builtin", "blosum62"), gaps))
        assert seen == [2, 1]


@pytest.mark.parametrize("gaps", [(-5, 0), (0, -1), (-1, 0), (0, 0)])
def test_fp16_shifted_values_edge_gaps(gaps):
    """Gap extend 0 (the shifted kernels add -0 per column and hold unshifted
    values) and gap open 0 (E and F take H itself): SW and NW, protein and
    DNA, fp16 and int16, against the oracle."""
    rnd = random.Random(515 + 10 * gaps[0] + gaps[1])
    qraw = "".join(rnd.choice(AA) for _ in range(300))
    db = homolog_db(rnd, qraw, 150, 1, 700, 8)
    for t in (O.SW, O.NW):
        run_both(lambda: check(t, 16, O.AMINOACID, 1, qraw, db, 15, ("builtin", "blosum62"), gaps))
    dq = "".join(rnd.choice("ACGT") for _ in range(200))
    ddb = ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(50, 300))) for _ in range(200)]
    ddb[17] = dq[20:180]
    for t in (O.SW, O.NW):
        run_both(lambda: check(t, 16, O.NUCLEOTIDE, 3, dq, ddb, 15, ("const", (2, -3)), gaps))


def test_fp16_scores_past_the_16_bit_tier():
    """A 13,000-residue query against its own copy (75,255: past 65,535, so
    the 16-bit overflow flag sends it to the exact tier), a prefix scoring just
    below 65,535 (offsets near the 16-bit threshold) and a middle piece."""
    rnd = random.Random(77)
    qraw = "".join(rnd.choice(AA) for _ in range(13000))
    db = ["".join(rnd.choice(AA) for _ in range(rnd.randint(50, 400))) for _ in range(40)]
    db += [qraw, qraw[:11200], qraw[500:9000]]
    m = O.matrix_builtin("blosum62")
    # (the lane schedule, and so the order of ties at the 10th score, depends
    # on the bit width)
    wants = {bw: O.search(O.SW, bw, O.AMINOACID, 1, [O.map_query(O.AMINOACID, qraw)], db, 10, m, -10, -1)[0]
             for bw in (8, 16)}
    want = wants[16]
    assert want[0][0] > 65535 and 60000 < max(s for s, i, _, _ in want if i == 41) < 65535
    ssa.reset_settings()
    ssa.set_output_mode(ssa.OUTPUT_ERROR)
    ssa.init_symbol_translation(ssa.AMINOACID, ssa.FORWARD_STRAND, 1, 1)
    ssa.init_score_matrix(ssa.MATRIX_BUILDIN, "blosum62")
    ssa.init_gap_penalties(-10, -1)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, qraw)
        try:
            for arith, expect in ((0, 2), (1, 1)):
                ssa.set_cell_arithmetic(arith)
                for bw in (8, 16):
                    got = ssa.sw_align(q, 10, bw, ssa.COMPUTE_SCORE)
                    st = ssa.last_stats()
                    assert st.cell_arithmetic == expect
                    assert st.overflow_16_bit_count >= 1
                    assert [(a.score, a.db_id) for a in got] == [(s, i) for s, i, _, _ in wants[bw]]
        finally:
            ssa.set_cell_arithmetic(0)
            ssa.free_sequence(q)


@pytest.mark.parametrize("strands", [1, 3])
def test_fp16_dna(strands):
    """Constant DNA scoring, one and both strands (the reverse strand read
    from the entries' reverse-complement twins), high-identity copies."""
    rnd = random.Random(31 + strands)
    qraw = "".join(rnd.choice("ACGT") for _ in range(900))
    db = ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(1, 600))) for _ in range(200)]
    comp = str.maketrans("ACGT", "TGCA")
    for c in range(10):
        a, b = rnd.randint(0, 200), rnd.randint(600, 900)
        frag = "".join(x if rnd.random() < 0.95 else rnd.choice("ACGT") for x in qraw[a:b])
        db[rnd.randrange(200)] = frag.translate(comp)[::-1] if c % 2 else frag
    seen = run_both(lambda: check(O.SW, 16, O.NUCLEOTIDE, strands, qraw, db, 30, ("const", (5, -4)), (-10, -2)))
    assert seen == [2, 1]


def test_fp16_falls_back_where_the_window_does_not_hold():
    """Scores that move a lane's maximum by more than the window allows per
    block (match 100) run in int16 and stay exact."""
    rnd = random.Random(5)
    qraw = "".join(rnd.choice("ACGT") for _ in range(300))
    db = ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(1, 300))) for _ in range(100)] + [qraw[20:280]]
    seen = run_both(lambda: check(O.SW, 16, O.NUCLEOTIDE, 1, qraw, db, 10, ("const", (100, -4)), (-10, -2)))
    assert seen == [1, 1]


def test_fp16_packed_and_paired_queries():
    """A batch of SW searches (packed queries, two by two on the paired
    kernel) with high-scoring homologs of several queries."""
    rnd = random.Random(404)
    qraws = ["".join(rnd.choice(AA) for _ in range(L)) for L in (40, 300, 700, 1200, 90, 500)]
    db = ["".join(rnd.choice(AA) for _ in range(rnd.randint(1, 400))) for _ in range(300)]
    for i, q in enumerate(qraws):
        db[7 + 11 * i] = q[len(q) // 8:]
    m = O.matrix_builtin("blosum62")
    configure(O.AMINOACID, 1, (-10, -1), ("builtin", "blosum62"))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        qs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, q) for q in qraws]
        try:
            wants = []
            for q in qraws:
                want, _ = O.search(O.SW, 16, O.AMINOACID, 1, [O.map_query(O.AMINOACID, q)], db, 15, m, -10, -1)
                wants.append([(s, i) for s, i, _, _ in want])
            for arith in (0, 1):
                ssa.set_cell_arithmetic(arith)
                lists, _ = ssa.batch_align([(ssa.SEARCH_SW, q, 15, 16, ssa.COMPUTE_SCORE) for q in qs])
                for got, want in zip(lists, wants):
                    assert [(a.score, a.db_id) for a in got] == want
        finally:
            ssa.set_cell_arithmetic(0)
            for q in qs:
                ssa.free_sequence(q)


@pytest.mark.parametrize("strands", [1, 3])
def test_fp16_nw_dna(strands):
    """NW without overflow trackers in fp16 (mode 4): DNA, constant +2/-4 with
    gaps 20/2 as the DNA workload, entries of 100-300 (every value within
    +-2047), one and both strands, scores and CIGARs."""
    rnd = random.Random(61 + strands)
    qraw = "".join(rnd.choice("ACGT") for _ in range(250))
    db = ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(100, 300))) for _ in range(400)]
    comp = str.maketrans("ACGT", "TGCA")
    for c in range(8):
        frag = "".join(x if rnd.random() < 0.9 else rnd.choice("ACGT") for x in qraw)
        db[rnd.randrange(400)] = frag.translate(comp)[::-1] if c % 2 else frag
    # the shapes with an untracked NW kernel (the cost model picks <32, 4> for
    # a DB this small; the DNA workload runs <16, 16>)
    try:
        for shape in ((16, 16), (8, 16), (16, 8), (8, 8), (32, 8)):
            ssa.set_kernel_shape(*shape)
            seen = run_both(lambda: check(O.NW, 16, O.NUCLEOTIDE, strands, qraw, db, 40, ("const", (2, -4)),
                                          (-20, -2), align=shape == (16, 16)))
            assert seen == [2, 1]
    finally:
        ssa.set_kernel_shape(0, 0)


def test_fp16_nw_protein_and_fallback():
    """NW protein with short entries runs in fp16; entries long enough that a
    two-gap path passes -2047 run in int16; both against the oracle."""
    rnd = random.Random(12)
    qraw = "".join(rnd.choice(AA) for _ in range(120))
    short = ["".join(rnd.choice(AA) for _ in range(rnd.randint(1, 200))) for _ in range(300)]
    short[5] = qraw[10:110]
    ssa.set_kernel_shape(16, 16)
    try:
        seen = run_both(lambda: check(O.NW, 16, O.AMINOACID, 1, qraw, short, 20, ("builtin", "blosum62"), (-10, -1)))
        assert seen == [2, 1]
        long_db = short[:200] + ["".join(rnd.choice(AA) for _ in range(2500))]
        seen = run_both(lambda: check(O.NW, 16, O.AMINOACID, 1, qraw, long_db, 20, ("builtin", "blosum62"),
                                      (-10, -1)))
        assert seen == [1, 1]
    finally:
        ssa.set_kernel_shape(0, 0)
