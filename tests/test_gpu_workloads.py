"""The benchmark workloads themselves, on the GPU, against the oracle.

Each test draws its data from libssa_amd/synthetic.py -- the generator the
benchmark uses (Robinson & Robinson background, lognormal or normal lengths,
planted fragments at 60-95 % identity with indels) -- at a size the oracle
finishes in seconds, and compares every hit (score, ID, strand, query strand),
the overflow counters and, where the workload asks for them, every region and
CIGAR with the oracle, which restates the reference run with one thread
(src/libssa.c:229-251).

- C1: the plain 16-bit SW search at its full size (2,000 sequences, q = 120).
- C2: the 8-bit SW cascade with the 50-, 500- and 5,000-residue queries; the
  5,000-residue query takes ten 512-row passes of the score kernel, run with
  the modelled shape and with the 16-lane x 32-row shape the full-size
  workload runs, and with streams of many entries each.
- C3: NW DNA, both strands, top 100 with CIGARs, 16 bit (the benchmark's
  width) and 8 bit (the NW 8-bit skip).
- C4: one batch of 32 log-uniform queries, against the oracle query by query.
- C5: the served SW/NW mix with CIGARs through a server of two workers, each
  served list against the oracle.

Oracle searches run on host threads (the oracle is C; ctypes releases the
GIL), so the larger cases take a few seconds of wall time.
"""
import concurrent.futures as cf
import os
import tempfile

import numpy as np
import pytest

import _oracle as O
from test_gpu_parity import configure

pytestmark = pytest.mark.gpu

ssa = pytest.importorskip("libssa_amd")
S = pytest.importorskip("libssa_amd.synthetic")

BLOSUM62 = ("builtin", "blosum62")
GAPS = (-10, -1)


def pool_map(fn, items):
    """fn over items on host threads (the oracle's C code runs without the GIL)."""
    with cf.ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 4)) as ex:
        return list(ex.map(fn, items))


def load_db(db):
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        S.write_fasta(path, [s.encode() for s in db])
        ssa.init_db(path)


def oracle_hits(t, bw, symtype, strands, qraw, db, k, m, gaps):
    qbufs = O.query_buffers(symtype, strands, qraw)
    want, st = O.search(t, bw, symtype, strands, qbufs, db, k, m, gaps[0], gaps[1])
    return qbufs, want, st


def assert_same(got, stats, oracle, symtype, strands, db, m, gaps, t, align, check_counts=True):
    """got: the library's list (and stats, or None); oracle: (qbufs, hits, stats)."""
    qbufs, want, st = oracle
    qstrands = [s for s in range(2) if (s + 1) & strands] if symtype == O.NUCLEOTIDE else [0]
    assert [(a.score, a.db_id, a.db_frame, qstrands.index(a.query_strand)) for a in got] == \
        [(s, i, strand, qi) for s, i, strand, qi in want]
    if stats is not None and check_counts:
        assert (stats.overflow_8_bit_count, stats.overflow_16_bit_count) == (st.overflow_8, st.overflow_16)
        assert (stats.chunk_count, stats.seq_count) == (st.chunk_count, st.seq_count)
    if not align:
        return
    for a, (s, i, strand, qi) in zip(got, want):
        dseq = O.map_db(symtype, db[i])
        assert a.db_seq == dseq
        if t == O.SW and O.region_local(qbufs[qi], dseq, m, gaps[0], gaps[1]) is None:
            assert (a.alignment or "", a.align_q_start, a.align_q_end, a.align_d_start, a.align_d_end) == \
                ("", 0, 0, 0, 0)
            continue
        c, r = O.cigar(t, qbufs[qi], dseq, m, gaps[0], gaps[1])
        assert a.alignment == c, (i, s)
        assert (a.align_q_start, a.align_q_end, a.align_d_start, a.align_d_end) == r


def text(db):
    return [s.decode() for s in db]


# ---------------------------------------------------------------------- C1
def test_c1_full_size():
    # 2,000 sequences of uniform length 40-600, q = 120, sw_align(q, 10, 16, scores)
    qs, db, _ = S.config_c1()
    db = text(db)
    configure(O.AMINOACID, 1, GAPS, BLOSUM62)
    load_db(db)
    q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, qs[0].decode())
    got = ssa.sw_align(q, 10, 16, ssa.COMPUTE_SCORE)
    stats = ssa.last_stats()
    ssa.free_sequence(q)
    m = O.matrix_builtin("blosum62")
    ora = oracle_hits(O.SW, 16, O.AMINOACID, 1, qs[0].decode(), db, 10, m, GAPS)
    assert len(ora[1]) == 10
    assert_same(got, stats, ora, O.AMINOACID, 1, db, m, GAPS, O.SW, False)
    assert stats.seq_count == 2000


# ---------------------------------------------------------------------- C2
@pytest.fixture(scope="module")
def c2():
    """1,000 lognormal sequences, 5 % planted fragments of the three queries,
    and the oracle's top 50 of each query."""
    qs, db, planted = S.config_c2(scale=0.005)
    db = text(db)
    m = O.matrix_builtin("blosum62")
    ora = pool_map(lambda q: oracle_hits(O.SW, 8, O.AMINOACID, 1, q.decode(), db, 50, m, GAPS), qs)
    return qs, db, planted, m, ora


@pytest.mark.parametrize("shape", [None, (16, 32), (16, 4), (64, 8)])
@pytest.mark.parametrize("qi", [0, 1, 2])
def test_c2_cascade(c2, qi, shape):
    qs, db, planted, m, ora = c2
    configure(O.AMINOACID, 1, GAPS, BLOSUM62)
    load_db(db)
    if shape:
        ssa.set_kernel_shape(*shape)
    try:
        q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, qs[qi].decode())
        got = ssa.sw_align(q, 50, 8, ssa.COMPUTE_SCORE)
        stats = ssa.last_stats()
        ssa.free_sequence(q)
    finally:
        ssa.set_kernel_shape(0, 0)
    assert_same(got, stats, ora[qi], O.AMINOACID, 1, db, m, GAPS, O.SW, False)
    if shape:
        assert (stats.lanes_per_stream, stats.rows_per_lane) == shape
    if qi == 2:
        # the 5,000-residue query: planted fragments saturate 8 bit
        assert stats.overflow_8_bit_count > 0
        if shape == (16, 32):
            assert -(-5000 // (16 * 32)) == 10  # ten passes of 512 rows


def test_c2_long_query_many_entries_per_stream(c2):
    # the 5,000-residue query over streams of many entries each (8 waves)
    qs, db, planted, m, ora = c2
    configure(O.AMINOACID, 1, GAPS, BLOSUM62)
    load_db(db)
    ssa.set_stream_waves(8)
    ssa.set_kernel_shape(16, 32)
    try:
        q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, qs[2].decode())
        got = ssa.sw_align(q, 50, 8, ssa.COMPUTE_SCORE)
        stats = ssa.last_stats()
        ssa.free_sequence(q)
    finally:
        ssa.set_kernel_shape(0, 0)
        ssa.set_stream_waves(0)
    assert_same(got, stats, ora[2], O.AMINOACID, 1, db, m, GAPS, O.SW, False)


def test_c2_batch_of_the_three_queries(c2):
    # the three queries as one batch (packed SW queries in one launch)
    qs, db, planted, m, ora = c2
    configure(O.AMINOACID, 1, GAPS, BLOSUM62)
    load_db(db)
    hs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, q.decode()) for q in qs]
    lists, stats = ssa.batch_align([(ssa.SEARCH_SW, h, 50, 8, ssa.COMPUTE_SCORE) for h in hs])
    for h in hs:
        ssa.free_sequence(h)
    for got, st, o in zip(lists, stats, ora):
        assert_same(got, st, o, O.AMINOACID, 1, db, m, GAPS, O.SW, False)


# ---------------------------------------------------------------------- C3
@pytest.mark.parametrize("bw,scale", [(16, 0.005), (8, 0.002)])
def test_c3_nw_dna_both_strands_cigars(bw, scale):
    # q = 250 nt; normal lengths (200, 50) clipped to 100-300; 1 % planted
    # copies of the query, half of them reverse-complemented; +2/-4, gaps
    # 20/2; nw_align(q, 100, bw, alignment)
    qs, db, planted = S.config_c3(scale=scale)
    db = text(db)
    gaps = (-20, -2)
    configure(O.NUCLEOTIDE, 3, gaps, ("const", (2, -4)))
    load_db(db)
    q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, qs[0].decode())
    got = ssa.nw_align(q, 100, bw, ssa.COMPUTE_ALIGNMENT)
    stats = ssa.last_stats()
    ssa.free_sequence(q)
    m = O.matrix_constant(2, -4)
    ora = oracle_hits(O.NW, bw, O.NUCLEOTIDE, 3, qs[0].decode(), db, 100, m, gaps)
    assert len(ora[1]) == 100
    assert_same(got, stats, ora, O.NUCLEOTIDE, 3, db, m, gaps, O.NW, True)
    # hits on both DB strands and from both query strands
    assert {a.db_frame for a in got} == {0, 1} and {a.query_strand for a in got} == {0, 1}


# ---------------------------------------------------------------------- C4
def test_c4_batch():
    # 32 queries log-uniform 30-2,000 in one batch against 2,000 lognormal
    # sequences with 2 % planted homologs; sw_align(q, 20, 8, scores) each
    qs, db, planted = S.config_c4(scale=0.004, nq=32)
    db = text(db)
    configure(O.AMINOACID, 1, GAPS, BLOSUM62)
    load_db(db)
    hs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, q.decode()) for q in qs]
    lists, stats = ssa.batch_align([(ssa.SEARCH_SW, h, 20, 8, ssa.COMPUTE_SCORE) for h in hs])
    alone = [ssa.sw_align(h, 20, 8, ssa.COMPUTE_SCORE) for h in hs]
    for h in hs:
        ssa.free_sequence(h)
    m = O.matrix_builtin("blosum62")
    ora = pool_map(lambda q: oracle_hits(O.SW, 8, O.AMINOACID, 1, q.decode(), db, 20, m, GAPS), qs)
    for got, st, one, o in zip(lists, stats, alone, ora):
        assert_same(got, st, o, O.AMINOACID, 1, db, m, GAPS, O.SW, False)
        assert [(a.score, a.db_id) for a in one] == [(a.score, a.db_id) for a in got]


# ---------------------------------------------------------------------- C5
def test_c5_served_mix():
    # the served mix: 80 % SW / 20 % NW, top 32 with CIGARs, 8-bit start,
    # query lengths 70 % U(50,400) + 30 % log-U(400,5000) (capped at 1,000
    # residues here so that the oracle finishes), against 1,000 lognormal
    # sequences, through a server of two workers with the arrivals of a
    # Poisson stream
    db = text(S.config_c5(scale=0.001))
    reqs = S.c5_requests(32, rate=200.0)
    for r in reqs:
        r["query"] = r["query"][:1000]
    assert {r["kind"] for r in reqs} == {"sw", "nw"}
    configure(O.AMINOACID, 1, GAPS, BLOSUM62)
    load_db(db)
    hs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, r["query"].decode()) for r in reqs]
    srv = ssa.Server(slots=2, max_batch=16)
    try:
        import time
        t0 = time.perf_counter()
        tickets = []
        for r, h in zip(reqs, hs):
            d = t0 + r["t"] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            kind = ssa.SEARCH_SW if r["kind"] == "sw" else ssa.SEARCH_NW
            tickets.append(srv.submit(kind, h, 32, 8, ssa.COMPUTE_ALIGNMENT))
        served = [srv.wait(t) for t in tickets]
        counters = srv.counters()
    finally:
        srv.stop()
    for h in hs:
        ssa.free_sequence(h)
    assert counters.requests == len(reqs)
    m = O.matrix_builtin("blosum62")

    def want(r):
        t = O.SW if r["kind"] == "sw" else O.NW
        return oracle_hits(t, 8, O.AMINOACID, 1, r["query"].decode(), db, 32, m, GAPS)

    ora = pool_map(want, reqs)
    for r, (got, st, lat, qms), o in zip(reqs, served, ora):
        t = O.SW if r["kind"] == "sw" else O.NW
        assert len(got) == 32
        assert_same(got, st, o, O.AMINOACID, 1, db, m, GAPS, t, True)
        assert lat >= qms >= 0


# ------------------------------------------------- DB rebuilt between searches
def test_symbol_type_changes_without_init_db():
    # one DB load, then searches under NUCLEOTIDE both strands, forward only,
    # both again and TRANS_DB, with no init_db in between: every cache derived
    # from the DB entries follows the entries (src/libssa.c:139-151 lets the
    # caller change the symbol type between searches)
    rng = np.random.default_rng(77)
    base = S.random_residues(rng, "nt", 400).tobytes().decode()
    db, _ = S.make_db(78, 300, "nt", "normal", [base.encode()], 0.05, normal=(150, 60, 4, 400))
    db = text(db)
    qraw = base[30:250]
    m = O.matrix_constant(2, -3)
    gaps = (-5, -2)
    configure(O.NUCLEOTIDE, 3, gaps, ("const", (2, -3)))
    load_db(db)
    q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, qraw)
    for strands in (3, 1, 3, 2):
        ssa.init_symbol_translation(ssa.NUCLEOTIDE, strands, 1, 1)
        for t, bw in ((O.SW, 8), (O.NW, 16)):
            f = ssa.sw_align if t == O.SW else ssa.nw_align
            got = f(q, 15, bw, ssa.COMPUTE_SCORE)
            stats = ssa.last_stats()
            ora = oracle_hits(t, bw, O.NUCLEOTIDE, strands, qraw, db, 15, m, gaps)
            assert_same(got, stats, ora, O.NUCLEOTIDE, strands, db, m, gaps, t, False)
    ssa.free_sequence(q)
    # TRANS_DB: the DB entries become 6 translated frames per sequence
    ssa.init_symbol_translation(ssa.TRANS_DB, 3, 1, 1)
    ssa.init_score_matrix(ssa.MATRIX_BUILDIN, "blosum62")
    pq = "MKVLAAGIVGLLLAHSTQA"
    q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, pq)
    got = ssa.sw_align(q, 10, 16, ssa.COMPUTE_SCORE)
    ssa.free_sequence(q)
    mb = O.matrix_builtin("blosum62")
    qb = O.query_buffers(O.TRANS_DB, 3, pq)
    want, _ = O.search(O.SW, 16, O.TRANS_DB, 3, qb, db, 10, mb, gaps[0], gaps[1], frames=True)
    assert [(a.score, a.db_id, a.db_frame, a.db_strand) for a in got] == \
        [(s, i, strand, frame) for s, i, strand, qi, frame in want]
    ssa.reset_settings()  # init_symbol_translation exits once TRANS_DB is set (src/libssa.c:146-151)
