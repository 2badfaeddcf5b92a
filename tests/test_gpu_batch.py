"""Batched searches (ssa_amd_batch_align) and the serving loop (ssa_amd_server_*).

Every result of a batch or of the server must be identical to the same search
run alone through sw_align / nw_align (which test_gpu_parity pins against the
oracle); a subset is also compared with the oracle directly.
"""
import os
import random
import tempfile
import threading
import time

import pytest

import _oracle as O
from test_gpu_parity import check, configure, rand_db, write_fasta

pytestmark = pytest.mark.gpu

ssa = pytest.importorskip("libssa_amd")

AA = "ARNDCQEGHILKMFPSTWYV"


def key(hits):
    return [(a.score, a.db_id, a.db_strand, a.db_frame, a.query_strand, a.alignment, a.align_q_start,
             a.align_q_end, a.align_d_start, a.align_d_end, a.db_seq) for a in hits]


def alone(req):
    t, q, k, bw, al = req
    f = ssa.sw_align if t == ssa.SEARCH_SW else ssa.nw_align
    return f(q, k, bw, al)


def make_requests(rnd, queries, types=(ssa.SEARCH_SW, ssa.SEARCH_NW), widths=(8, 16, 64)):
    reqs = []
    for i, q in enumerate(queries):
        t = types[i % len(types)]
        bw = widths[(i // len(types)) % len(widths)]
        al = ssa.COMPUTE_ALIGNMENT if i % 3 == 0 else ssa.COMPUTE_SCORE
        reqs.append((t, q, rnd.choice([1, 5, 20]), bw, al))
    return reqs


@pytest.fixture
def protein_db():
    rnd = random.Random(2024)
    base = "".join(rnd.choice(AA) for _ in range(700))
    db = rand_db(rnd, O.AMINOACID, 350, 1, 450, homolog_of=base, frac=0.15)
    configure(O.AMINOACID, 1, (-10, -1), ("builtin", "blosum62"))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        qraws = [base[:7], base[10:43], "".join(rnd.choice(AA) for _ in range(64)), base[100:220], base[:250],
                 base[150:551], "".join(rnd.choice(AA) for _ in range(512)), base[:700], base[300:330],
                 base[5:600], "".join(rnd.choice(AA) for _ in range(90)), base[400:480]]
        qs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, s) for s in qraws]
        yield rnd, db, qraws, qs
        for q in qs:
            ssa.free_sequence(q)


def test_batch_equals_alone(protein_db):
    rnd, db, qraws, qs = protein_db
    reqs = make_requests(rnd, qs)
    want = [key(alone(r)) for r in reqs]
    lists, stats = ssa.batch_align(reqs)
    assert [key(x) for x in lists] == want
    assert all(s.seq_count == len(db) for s in stats)


def test_batch_matches_oracle(protein_db):
    rnd, db, qraws, qs = protein_db
    reqs = [(ssa.SEARCH_SW, qs[3], 20, 8, ssa.COMPUTE_ALIGNMENT), (ssa.SEARCH_NW, qs[5], 10, 16, ssa.COMPUTE_SCORE),
            (ssa.SEARCH_SW, qs[7], 10, 16, ssa.COMPUTE_SCORE)]
    lists, _ = ssa.batch_align(reqs)
    m = O.matrix_builtin("blosum62")
    for (t, _, k, bw, al), got, qi in zip(reqs, lists, (3, 5, 7)):
        qb = [O.map_query(O.AMINOACID, qraws[qi])]
        want, _ = O.search(O.SW if t == ssa.SEARCH_SW else O.NW, bw, O.AMINOACID, 1, qb, db, k, m, -10, -1)
        assert [(a.score, a.db_id) for a in got] == [(s, i) for s, i, _, _ in want]
        if al == ssa.COMPUTE_ALIGNMENT:
            for a, (s, i, _, _) in zip(got, want):
                c, r = O.cigar(O.SW, qb[0], O.map_db(O.AMINOACID, db[i]), m, -10, -1)
                assert a.alignment == c


def test_batch_nucleotide_both_strands():
    rnd = random.Random(99)
    base = "".join(rnd.choice("ACGT") for _ in range(300))
    db = rand_db(rnd, O.NUCLEOTIDE, 250, 20, 320, homolog_of=base, frac=0.2)
    configure(O.NUCLEOTIDE, 3, (-20, -2), ("const", (2, -4)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        qraws = [base[:250], base[40:140], base, base[::-1][:77]]
        qs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, s) for s in qraws]
        reqs = [(ssa.SEARCH_NW, qs[0], 20, 16, ssa.COMPUTE_ALIGNMENT), (ssa.SEARCH_SW, qs[1], 10, 8, ssa.COMPUTE_SCORE),
                (ssa.SEARCH_NW, qs[2], 5, 8, ssa.COMPUTE_SCORE), (ssa.SEARCH_SW, qs[3], 15, 16, ssa.COMPUTE_ALIGNMENT)]
        want = [key(alone(r)) for r in reqs]
        lists, _ = ssa.batch_align(reqs)
        assert [key(x) for x in lists] == want
        for q in qs:
            ssa.free_sequence(q)
    # the alone path is the oracle-checked one: spot-check one request against the oracle as well
    check(O.NW, 16, O.NUCLEOTIDE, 3, qraws[0], db, 20, ("const", (2, -4)), (-20, -2), align=True)


@pytest.mark.parametrize("slots,max_batch", [(1, 1), (2, 3), (3, 16)])
def test_server_results_equal_alone(protein_db, slots, max_batch):
    rnd, db, qraws, qs = protein_db
    reqs = make_requests(rnd, qs) * 2
    want = [key(alone(r)) for r in reqs]
    srv = ssa.Server(slots=slots, max_batch=max_batch)
    try:
        tickets = [None] * len(reqs)

        def submit():
            for i, r in enumerate(reqs):
                tickets[i] = srv.submit(*r)
                time.sleep(rnd.choice([0.0, 0.0, 0.002]))

        th = threading.Thread(target=submit)
        th.start()
        th.join()
        got = []
        for t in tickets:
            hits, st, lat, qms = srv.wait(t)
            assert lat >= qms >= 0.0
            got.append(key(hits))
        c = srv.counters()
        assert c.requests == len(reqs) and c.queued == 0
        assert c.max_batch_seen <= max_batch
    finally:
        srv.stop()
    assert got == want


def test_server_rejects_invalid_requests(protein_db):
    rnd, db, qraws, qs = protein_db
    srv = ssa.Server(slots=1, max_batch=4)
    try:
        with pytest.raises(ValueError):
            srv.submit(7, qs[0], 5, 8, ssa.COMPUTE_SCORE)
        with pytest.raises(ValueError):
            srv.submit(ssa.SEARCH_SW, qs[0], 5, 12, ssa.COMPUTE_SCORE)
        with pytest.raises(KeyError):
            srv.wait(12345)
    finally:
        srv.stop()


def test_autotune_keeps_results(protein_db):
    rnd, db, qraws, qs = protein_db
    reqs = make_requests(rnd, qs)
    want = [key(alone(r)) for r in reqs]
    try:
        tuned = ssa.autotune(ssa.SEARCH_SW, [40, 300, 700], copies=2)
        assert [t[0] for t in tuned] == [40, 300, 700]
        for _, lanes, rows, ms in tuned:
            assert lanes in (8, 16, 32, 64) and rows in (4, 8, 16, 32) and ms > 0
        ssa.autotune(ssa.SEARCH_NW, [100])
        assert [key(alone(r)) for r in reqs] == want
        lists, _ = ssa.batch_align(reqs)
        assert [key(x) for x in lists] == want
    finally:
        ssa.autotune_reset()


@pytest.mark.parametrize("symtype,strands", [(O.AMINOACID, 1), (O.NUCLEOTIDE, 3), (O.NUCLEOTIDE, 1)])
def test_packed_db_searches_equal_fasta(tmp_path, symtype, strands):
    # the same searches on the FASTA file and on its packed file (packed with
    # both strands' entries; a forward-only search takes strand 0's only)
    rnd = random.Random(31 + strands)
    nuc = symtype == O.NUCLEOTIDE
    alpha = "ACGT" if nuc else AA
    base = "".join(rnd.choice(alpha) for _ in range(400))
    db = rand_db(rnd, symtype, 300, 0, 380, homolog_of=base, frac=0.2)
    matrix = ("const", (2, -4)) if nuc else ("builtin", "blosum62")
    gaps = (-20, -2) if nuc else (-10, -1)
    configure(symtype, 3 if nuc else 1, gaps, matrix)
    fas = str(tmp_path / "db.fas")
    write_fasta(fas, db)
    ssa.init_db(fas)
    packed = str(tmp_path / "db.ssapack")
    ssa.pack_db(packed)
    configure(symtype, strands, gaps, matrix)
    qraws = [base[:300], base[50:120], "".join(rnd.choice(alpha) for _ in range(150))]
    qs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, s) for s in qraws]
    reqs = [(ssa.SEARCH_SW, qs[0], 20, 8, ssa.COMPUTE_ALIGNMENT), (ssa.SEARCH_NW, qs[1], 10, 16, ssa.COMPUTE_ALIGNMENT),
            (ssa.SEARCH_SW, qs[2], 15, 64, ssa.COMPUTE_SCORE), (ssa.SEARCH_NW, qs[0], 5, 8, ssa.COMPUTE_SCORE)]
    try:
        ssa.init_db(fas)
        want = [key(alone(r)) for r in reqs]
        ssa.init_db(packed)
        assert [key(alone(r)) for r in reqs] == want
        lists, _ = ssa.batch_align(reqs)
        assert [key(x) for x in lists] == want
    finally:
        for q in qs:
            ssa.free_sequence(q)
    # and the FASTA path is the oracle-checked one
    check(O.SW, 8, symtype, strands, qraws[0], db, 20, matrix, gaps, align=True)


def test_packed_db_other_symbol_type_maps_from_residues(tmp_path):
    # a protein-packed file searched as nucleotides: the codes are mapped
    # again from the stored residues, as for a FASTA file
    rnd = random.Random(77)
    db = ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(20, 200))) for _ in range(150)]
    configure(O.AMINOACID, 1, (-10, -1), ("builtin", "blosum62"))
    fas = str(tmp_path / "db.fas")
    write_fasta(fas, db)
    ssa.init_db(fas)
    packed = str(tmp_path / "db.ssapack")
    ssa.pack_db(packed)
    configure(O.NUCLEOTIDE, 3, (-20, -2), ("const", (2, -4)))
    q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, db[3][10:90])
    try:
        ssa.init_db(fas)
        want = key(ssa.sw_align(q, 10, 16, ssa.COMPUTE_ALIGNMENT))
        ssa.init_db(packed)
        assert key(ssa.sw_align(q, 10, 16, ssa.COMPUTE_ALIGNMENT)) == want
    finally:
        ssa.free_sequence(q)


SW_SHAPES = [(8, 4), (16, 4), (64, 4), (8, 8), (16, 8), (32, 8), (64, 8), (8, 16), (16, 16), (32, 16), (8, 32), (16, 32)]


@pytest.mark.parametrize("shape", SW_SHAPES)
def test_packed_sw_batch_matches_oracle(shape):
    """The SW searches of a batch run as packed queries: their rows end to
    end, each query from a lane boundary, and, for shapes of at most 256 rows
    per pass, two packed queries sharing every stream (one per half).  Queries
    shorter than a lane, of exactly one lane, ending one row into a lane,
    filling a pass exactly and crossing pass boundaries, an odd number of
    them, at 8/16/64 bit, every SW kernel shape; each list and its overflow
    counters must be the oracle's."""
    lanes, rows = shape
    pr = lanes * rows
    rnd = random.Random(7000 + lanes * 100 + rows)
    base = "".join(rnd.choice(AA) for _ in range(900))
    db = rand_db(rnd, O.AMINOACID, 160, 1, 300, homolog_of=base, frac=0.25)
    configure(O.AMINOACID, 1, (-10, -1), ("builtin", "blosum62"))
    lens = sorted({1, rows - 1 or 1, rows, rows + 1, pr - 1, pr, pr + 3, min(2 * pr + 5, 900), rnd.randint(1, 300)})
    qraws = []
    for n in lens:
        o = rnd.randint(0, 900 - n)
        qraws.append(base[o:o + n] if rnd.random() < 0.7 else "".join(rnd.choice(AA) for _ in range(n)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        qs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, s) for s in qraws]
        widths = [(8, 16, 64)[i % 3] for i in range(len(qs))]
        reqs = [(ssa.SEARCH_SW, q, rnd.choice([3, 10, 40]), bw, ssa.COMPUTE_SCORE) for q, bw in zip(qs, widths)]
        ssa.set_kernel_shape(lanes, rows)
        try:
            lists, stats = ssa.batch_align(reqs)
        finally:
            ssa.set_kernel_shape(0, 0)
        for q in qs:
            ssa.free_sequence(q)
    m = O.matrix_builtin("blosum62")
    for (_, _, k, bw, _), raw, got, st in zip(reqs, qraws, lists, stats):
        want, ost = O.search(O.SW, bw, O.AMINOACID, 1, [O.map_query(O.AMINOACID, raw)], db, k, m, -10, -1)
        assert [(a.score, a.db_id) for a in got] == [(s, i) for s, i, _, _ in want], (len(raw), bw)
        assert (st.overflow_8_bit_count, st.overflow_16_bit_count) == (ost.overflow_8, ost.overflow_16)
        assert (st.lanes_per_stream, st.rows_per_lane) == (lanes, rows)


# SSA_BATCH_SEEDS=N runs N random batches (default 4) from seed SSA_BATCH_SEED0 (default 0)
_BSEED0 = int(os.environ.get("SSA_BATCH_SEED0", "0"))
BATCH_CASES = list(range(_BSEED0, _BSEED0 + int(os.environ.get("SSA_BATCH_SEEDS", "4"))))


@pytest.mark.parametrize("seed", BATCH_CASES)
def test_random_batch_matches_oracle(seed):
    """A random batch — SW and NW, 8/16/64 bit, scores and CIGARs, protein or
    DNA with one or both strands, queries from 1 residue to several passes —
    against the oracle request by request: hits, strands, overflow and chunk
    counters, CIGARs and regions."""
    rnd = random.Random(31000 + seed)
    symtype = O.AMINOACID if seed % 2 == 0 else O.NUCLEOTIDE
    strands = 1 if symtype == O.AMINOACID else rnd.choice([1, 2, 3])
    alpha = AA if symtype == O.AMINOACID else "ACGT"
    base = "".join(rnd.choice(alpha) for _ in range(1200))
    db = rand_db(rnd, symtype, rnd.randint(60, 300), 1, rnd.choice([40, 200, 500]), homolog_of=base, frac=0.2)
    if symtype == O.AMINOACID:
        matrix = ("builtin", rnd.choice(["blosum62", "blosum50", "pam30"]))
    else:
        matrix = ("const", rnd.choice([(1, -1), (2, -4), (5, -4)]))
    gaps = rnd.choice([(-10, -1), (-4, -2), (-20, -2), (-3, -1)])
    chunk = rnd.choice([1000, 37, 250])
    configure(symtype, strands, gaps, matrix, chunk=chunk)
    qraws = []
    for _ in range(rnd.randint(1, 12)):
        n = rnd.choice([1, 5, 30, 64, 100, 257, 513, 900])
        o = rnd.randint(0, 1200 - n)
        qraws.append(base[o:o + n] if rnd.random() < 0.6 else "".join(rnd.choice(alpha) for _ in range(n)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        qs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, s) for s in qraws]
        reqs = [(rnd.choice([ssa.SEARCH_SW, ssa.SEARCH_NW]), q, rnd.choice([1, 5, 20, 60]), rnd.choice([8, 16, 64]),
                 ssa.COMPUTE_ALIGNMENT if rnd.random() < 0.3 else ssa.COMPUTE_SCORE) for q in qs]
        # past the default cases, a forced kernel shape (NW takes 16 rows where 32 are forced)
        shape = rnd.choice(SW_SHAPES) if seed >= 4 else (0, 0)
        ssa.set_kernel_shape(*shape)
        try:
            lists, stats = ssa.batch_align(reqs)
        finally:
            ssa.set_kernel_shape(0, 0)
            for q in qs:
                ssa.free_sequence(q)
    m = O.matrix_builtin(matrix[1]) if matrix[0] == "builtin" else O.matrix_constant(*matrix[1])
    qstrands = [s for s in range(2) if (s + 1) & strands] if symtype == O.NUCLEOTIDE else [0]
    for (t, _, k, bw, al), raw, got, st in zip(reqs, qraws, lists, stats):
        ot = O.SW if t == ssa.SEARCH_SW else O.NW
        qbufs = O.query_buffers(symtype, strands, raw)
        want, ost = O.search(ot, bw, symtype, strands, qbufs, db, k, m, gaps[0], gaps[1], chunk=chunk)
        case = (seed, shape, len(raw), ot, bw, k)
        assert [(a.score, a.db_id, a.db_frame, qstrands.index(a.query_strand)) for a in got] == \
            [(s, i, strand, qi) for s, i, strand, qi in want], case
        assert (st.overflow_8_bit_count, st.overflow_16_bit_count) == (ost.overflow_8, ost.overflow_16), case
        assert (st.chunk_count, st.seq_count) == (ost.chunk_count, ost.seq_count), case
        if al == ssa.COMPUTE_ALIGNMENT:
            for a, (s, i, strand, qi) in zip(got, want):
                if ot == O.SW and s <= 0:
                    continue
                dseq = O.map_db(symtype, db[i])
                if ot == O.SW and O.region_local(qbufs[qi], dseq, m, gaps[0], gaps[1]) is None:
                    # the attached sequence (a reverse-strand hit carries the
                    # forward one) has no positive cell: the reference's region
                    # is undefined there, this build reports an empty alignment
                    assert (a.alignment or "", a.align_q_start, a.align_q_end, a.align_d_start, a.align_d_end) == \
                        ("", 0, 0, 0, 0), case
                    continue
                c, r = O.cigar(ot, qbufs[qi], dseq, m, gaps[0], gaps[1])
                assert (a.alignment, a.align_q_start, a.align_q_end, a.align_d_start, a.align_d_end) == (c,) + tuple(r), case


# SSA_SERVER_SEEDS=N runs N random serving rounds (default 2)
SERVER_CASES = list(range(int(os.environ.get("SSA_SERVER_SEEDS", "2"))))


@pytest.mark.parametrize("seed", SERVER_CASES)
def test_random_server_results_equal_alone(protein_db, seed):
    """Random requests from several submitting threads, with random gaps,
    into a server of 1-4 workers and random batch limits: every result is
    the one the search gives alone."""
    _, db, qraws, qs = protein_db
    rnd = random.Random(52000 + seed)
    reqs = [(rnd.choice([ssa.SEARCH_SW, ssa.SEARCH_NW]), rnd.choice(qs), rnd.choice([1, 5, 20, 50]),
             rnd.choice([8, 16, 64]), ssa.COMPUTE_ALIGNMENT if rnd.random() < 0.4 else ssa.COMPUTE_SCORE)
            for _ in range(rnd.randint(5, 40))]
    want = [key(alone(r)) for r in reqs]
    srv = ssa.Server(slots=rnd.randint(1, 4), max_batch=rnd.choice([1, 2, 5, 16]))
    try:
        tickets = [None] * len(reqs)
        nthreads = rnd.randint(1, 4)
        delays = [rnd.choice([0.0, 0.0, 0.001, 0.01]) for _ in reqs]

        def submit(part):
            for i in range(part, len(reqs), nthreads):
                tickets[i] = srv.submit(*reqs[i])
                time.sleep(delays[i])

        ths = [threading.Thread(target=submit, args=(p,)) for p in range(nthreads)]
        for th in ths:
            th.start()
        for th in ths:
            th.join()
        got = [key(srv.wait(t)[0]) for t in tickets]
    finally:
        srv.stop()
    assert got == want, seed
