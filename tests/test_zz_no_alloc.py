"""ssa_amd_align_noalloc (the search that allocates no device memory) against
sw_align / nw_align with COMPUTE_SCORE, through the C ABI.

Each case runs the allocating entry point first (which also leaves the DB and
a stream layout resident, the no-allocation call's precondition), then two
no-allocation calls in a row, the second with a smaller input (a shorter
query, fewer hits), and requires the same hits in the same order: score, DB
ID, the strand and frame fields, the query buffer's strand and frame.  The
scratch array keeps the first call's contents when the second runs.
"""
import os
import random
import tempfile

import pytest

import _oracle as O
from test_gpu_parity import configure, rand_db, write_fasta

pytestmark = pytest.mark.gpu

ssa = pytest.importorskip("libssa_amd")

AA = "ARNDCQEGHILKMFPSTWYV"


def key_list(hits):
    return [(a.score, a.db_id, a.db_strand, a.db_frame, a.query_strand, a.query_frame) for a in hits]


def key_noalloc(hits):
    return [(h.score, h.db_id, h.db_strand, h.db_frame, h.query_strand, h.query_frame) for h in hits]


def run_case(t, bw, big, small):
    """big / small: (query handle, hitcount)."""
    f = ssa.sw_align if t == ssa.SEARCH_SW else ssa.nw_align
    want = [key_list(f(q, k, bw, ssa.COMPUTE_SCORE)) for q, k in (big, small)]
    for (q, k), w in zip((big, small), want):
        rc, hits = ssa.align_noalloc(t, q, k, bw)
        assert rc == 0
        assert key_noalloc(hits) == w
        assert len(hits) == len(w)


@pytest.fixture(scope="module")
def protein():
    rnd = random.Random(4711)
    base = "".join(rnd.choice(AA) for _ in range(1400))
    db = rand_db(rnd, O.AMINOACID, 900, 1, 600, homolog_of=base, frac=0.1)
    db[5] = ""  # an empty record
    configure(O.AMINOACID, 1, (-10, -1), ("builtin", "blosum62"))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
    raws = [base[:1300], base[200:260], base[100:700], base[650:700]]
    qs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, s) for s in raws]
    yield qs
    for q in qs:
        ssa.free_sequence(q)


@pytest.mark.parametrize("bw", [8, 16, 64])
@pytest.mark.parametrize("t", [ssa.SEARCH_SW, ssa.SEARCH_NW])
def test_noalloc_equals_align(protein, t, bw):
    configure(O.AMINOACID, 1, (-10, -1), ("builtin", "blosum62"))
    q_long, q_short, q_mid, q_tiny = protein
    # a query of three 512-row passes (pass boundaries in the scratch), then a
    # one-pass query and fewer hits
    run_case(t, bw, (q_long, 30), (q_short, 7))
    run_case(t, bw, (q_mid, 12), (q_tiny, 1))


def test_noalloc_nucleotide_both_strands():
    # two query buffers (a packed SW launch; NW per strand)
    rnd = random.Random(99)
    base = "".join(rnd.choice("ACGT") for _ in range(500))
    db = rand_db(rnd, O.NUCLEOTIDE, 400, 20, 400, homolog_of=base, frac=0.1)
    configure(O.NUCLEOTIDE, 3, (-5, -2), ("const", (2, -3)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
    q1 = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, base[:420])
    q2 = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, base[100:180])
    for t, bw in ((ssa.SEARCH_SW, 8), (ssa.SEARCH_NW, 16), (ssa.SEARCH_SW, 16)):
        run_case(t, bw, (q1, 25), (q2, 4))
    ssa.free_sequence(q1)
    ssa.free_sequence(q2)


def test_noalloc_needs_a_resident_db():
    # a DB loaded but never searched: no device state, so the call says so
    # (-3) instead of allocating; after one search it runs
    rnd = random.Random(5)
    db = ["".join(rnd.choice(AA) for _ in range(rnd.randint(10, 200))) for _ in range(50)]
    configure(O.AMINOACID, 1, (-10, -1), ("builtin", "blosum62"))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
    q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, db[3][:80])
    ssa.set_output_mode(ssa.OUTPUT_SILENT)
    rc, hits = ssa.align_noalloc(ssa.SEARCH_SW, q, 5, 16)
    ssa.set_output_mode(ssa.OUTPUT_ERROR)
    assert rc == -3 and hits == []
    want = key_list(ssa.sw_align(q, 5, 16, ssa.COMPUTE_SCORE))
    rc, hits = ssa.align_noalloc(ssa.SEARCH_SW, q, 5, 16)
    assert rc == 0 and key_noalloc(hits) == want
    assert ssa.align_noalloc(7, q, 5, 16)[0] == -1
    assert ssa.align_noalloc(ssa.SEARCH_SW, q, 5, 12)[0] == -1
    ssa.free_sequence(q)


def test_noalloc_long_query_on_a_capped_grid():
    # 50,000 lognormal sequences and the 5,000-residue query of the SW 8-bit
    # workload: at the full persistent grid its ten passes would keep ~0.25 GB
    # of pass boundaries, twice the scratch, so fewer workgroups run
    S = pytest.importorskip("libssa_amd.synthetic")
    qs, db, _ = S.config_c2(scale=0.25)
    configure(O.AMINOACID, 1, (-10, -1), ("builtin", "blosum62"))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        S.write_fasta(path, db)
        ssa.init_db(path)
    hs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, q.decode()) for q in qs]
    run_case(ssa.SEARCH_SW, 8, (hs[2], 50), (hs[0], 10))
    for h in hs:
        ssa.free_sequence(h)
