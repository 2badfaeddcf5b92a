"""Translated searches (TRANS_QUERY, TRANS_DB, TRANS_BOTH) on the GPU against
the oracle, through the libssa.h drop-in API.

Scores, hit selection and order, the strand and frame fields of every hit
(with the reference's swap of the two in db_seq, src/algo/aligner.c:76-77),
the query frame, the attached DB sequence and the CIGAR strings must be
identical to the reference run with one thread.
"""
import os
import random
import subprocess
import sys
import tempfile

import pytest

import _oracle as O

pytestmark = pytest.mark.gpu

ssa = pytest.importorskip("libssa_amd")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LANES = {"avx2": (32, 16, 2), "sse": (16, 8, 1)}  # lanes8, lanes16, COMPUTE_ON_*
VALID = [1, 2, 3, 4, 5, 6, 9, 10, 11, 12, 13, 14, 15, 16, 21, 22, 23]


def write_fasta(path, seqs):
    with open(path, "w") as f:
        for i, s in enumerate(seqs):
            f.write(">s%d\n" % i)
            for k in range(0, len(s), 60):
                f.write(s[k:k + 60] + "\n")


def configure(symtype, strands, gencodes, gaps, matrix, isa, chunk):
    # init_symbol_translation exits once a translated DB type is in force
    # (src/libssa.c:146-151): start from the initial settings
    ssa.reset_settings()
    ssa.set_output_mode(ssa.OUTPUT_ERROR)
    ssa.init_symbol_translation(symtype, strands, gencodes[1], gencodes[0])  # (type, strands, db, query)
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


def prepared_db(symtype, d_gencode, raw, strand_field, frame_field):
    """The DB sequence the reference attaches to a hit (us_prepare_sequence,
    src/util/util_sequence.c:384-417): a translated DB's frame of the strand the
    field names, 0 forward and any other value the reverse complement."""
    if symtype in (O.TRANS_DB, O.TRANS_BOTH):
        return O.translate(O.translate_table(d_gencode), O.map_db(O.NUCLEOTIDE, raw), strand_field, frame_field)
    return O.map_db(O.AMINOACID, raw)


def check(t, bw, symtype, strands, gencodes, qraw, db, k, matrix, gaps, isa="avx2", chunk=1000, align=False):
    configure(symtype, strands, gencodes, gaps, matrix, isa, chunk)
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, qraw)
        f = ssa.sw_align if t == O.SW else ssa.nw_align
        got = f(q, k, bw, ssa.COMPUTE_ALIGNMENT if align else ssa.COMPUTE_SCORE)
        stats = ssa.last_stats()
        ssa.free_sequence(q)
    m = oracle_matrix(matrix)
    qbufs = O.query_buffers(symtype, strands, qraw, q_gencode=gencodes[0])
    qtags = [(s, f) for s in range(2) if (s + 1) & strands for f in range(3)] \
        if symtype in (O.TRANS_QUERY, O.TRANS_BOTH) else [(0, 0)]
    l8, l16, _ = LANES[isa]
    want, st = O.search(t, bw, symtype, strands, qbufs, db, k, m, gaps[0], gaps[1], chunk=chunk, lanes8=l8,
                        lanes16=l16, d_gencode=gencodes[1], frames=True)
    # db_seq.strand holds the entry's frame and db_seq.frame its strand (src/algo/aligner.c:76-77)
    assert [(a.score, a.db_id, a.db_frame, a.db_strand, (a.query_strand, a.query_frame)) for a in got] == \
        [(s, i, strand, frame, qtags[qi]) for s, i, strand, qi, frame in want]
    assert (stats.overflow_8_bit_count, stats.overflow_16_bit_count) == (st.overflow_8, st.overflow_16)
    assert (stats.chunk_count, stats.seq_count) == (st.chunk_count, st.seq_count)
    for a, (s, i, strand, qi, frame) in zip(got, want):
        dseq = prepared_db(symtype, gencodes[1], db[i], strand, frame)
        assert a.db_seq == dseq and a.query == qbufs[qi]
        if align and not (t == O.SW and s <= 0) and len(dseq) and len(qbufs[qi]):
            if t == O.SW and O.region_local(qbufs[qi], dseq, m, gaps[0], gaps[1]) is None:
                # the attached sequence (re-translated from the strand field)
                # has no positive cell: the reference's region is undefined
                # there, this build reports an empty alignment
                assert (a.alignment or "", a.align_q_start, a.align_q_end, a.align_d_start, a.align_d_end) == \
                    ("", 0, 0, 0, 0)
                continue
            c, r = O.cigar(t, qbufs[qi], dseq, m, gaps[0], gaps[1])
            assert a.alignment == c
            assert (a.align_q_start, a.align_q_end, a.align_d_start, a.align_d_end) == r
    return got


# tests/algo/test_searcher.c:248-330 (setup :32-54): TRANS_QUERY, both strands,
# genetic code 3, constant 1/-1, gaps -1/-1, six hits against short_AA.fas
TQ = "ATGCCCAAAATTGACTTGAATAGGGGCGTGAGGAGGGGTTTTCAATACTATTTTTTG"


@pytest.mark.parametrize("bw,isa", [(64, "avx2"), (16, "sse"), (8, "sse")])
@pytest.mark.parametrize("t,expect", [(O.SW, [2, 2, 2, 2, 2, 1]), (O.NW, [-109, -109, -110, -110, -110, -113])])
def test_translate_query_golden(bw, isa, t, expect):
    db = [s for _, s in O.read_fasta(O.data("short_AA.fas"))]
    got = check(t, bw, O.TRANS_QUERY, 3, (3, 3), TQ, db, 6, ("const", (1, -1)), (-1, -1), isa=isa, chunk=6)
    assert [a.score for a in got] == expect


def rand_dna(rnd, n):
    return "".join(rnd.choice("ACGT") for _ in range(n))


def rand_aa(rnd, n):
    return "".join(rnd.choice("ARNDCQEGHILKMFPSTWYV") for _ in range(n))


def revcomp(s):
    return "".join({"A": "T", "C": "G", "G": "C", "T": "A"}[c] for c in reversed(s))


def back_translate(rnd, prot, gencode):
    """A DNA sequence whose frame 0 translates to prot under the genetic code."""
    table = O.translate_table(gencode)
    codons = {}
    for a in "ACGT":
        for b in "ACGT":
            for c in "ACGT":
                x = O.translate(table, O.map_db(O.NUCLEOTIDE, a + b + c), 0, 0)[0]
                codons.setdefault("-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"[x], []).append(a + b + c)
    return "".join(rnd.choice(codons[x]) for x in prot)


CASES = [(sym, strands, t, bw) for sym in (O.TRANS_QUERY, O.TRANS_DB, O.TRANS_BOTH) for strands in (1, 2, 3)
         for t in (O.SW, O.NW) for bw in (8, 16)]
CASES += [(sym, 3, t, 64) for sym in (O.TRANS_QUERY, O.TRANS_DB, O.TRANS_BOTH) for t in (O.SW, O.NW)]
# SSA_TRANSLATE_ROUNDS=N repeats every configuration N times with other seeds (default 1)
CASES = CASES * int(os.environ.get("SSA_TRANSLATE_ROUNDS", "1"))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_random_translated_parity(case):
    sym, strands, t, bw = CASES[case]
    rnd = random.Random(900 + case)
    gencodes = (rnd.choice(VALID), rnd.choice(VALID))
    matrix = rnd.choice([("builtin", "BLOSUM62"), ("builtin", "BLOSUM45"), ("builtin", "PAM30"),
                         ("const", (5, -4))])
    gaps = rnd.choice([(-10, -1), (-3, -1), (-11, -2), (-5, -5)])
    if sym == O.TRANS_QUERY:
        prot = [rand_aa(rnd, rnd.randint(1, 120)) for _ in range(rnd.randint(40, 300))]
        qraw = rand_dna(rnd, rnd.randint(12, 240))
        # homologs: translations of the query's frames
        tq = O.query_buffers(O.TRANS_QUERY, 3, qraw, q_gencode=gencodes[0])
        for _ in range(5):
            frame = rnd.choice(tq)
            prot[rnd.randrange(len(prot))] = "".join("-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"[c] for c in frame)
        db = prot
    else:
        # DNA DB with sequences of 2..4 nucleotides (empty frames) and planted
        # copies of the query (its back-translation for a protein query) on
        # either strand, shifted into any frame
        db = [rand_dna(rnd, rnd.choice([2, 3, 4, 5, rnd.randint(6, 60), rnd.randint(60, 400)]))
              for _ in range(rnd.randint(40, 240))]
        if sym == O.TRANS_DB:
            qraw = rand_aa(rnd, rnd.randint(5, 80))
        else:
            qraw = rand_dna(rnd, rnd.randint(9, 180))
        for _ in range(4):
            i = rnd.randrange(len(db))
            src = qraw if sym == O.TRANS_BOTH else back_translate(rnd, qraw, gencodes[1])
            db[i] = rand_dna(rnd, rnd.randint(0, 4)) + (src if rnd.random() < 0.5 else revcomp(src))
    align = case % 3 != 1
    isa = "sse" if case % 4 == 3 else "avx2"
    k = rnd.choice([5, 17, 40])
    check(t, bw, sym, strands, gencodes, qraw, db, k, matrix, gaps, isa=isa, chunk=rnd.choice([7, 50, 1000]),
          align=align)


def test_translated_db_too_short_exits():
    # a 1-nt DB sequence has no frame 2: (len - frame) / 3 wraps in the
    # reference and its allocation fails, which exits with this message
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, ["ACGTACGTAC", "T", "GGGCCC"])
        code = (
            "import sys; sys.path.insert(0, %r)\n"
            "import libssa_amd as ssa\n"
            "ssa.init_symbol_translation(ssa.TRANS_DB, ssa.BOTH_STRANDS, 1, 1)\n"
            "ssa.init_score_matrix(ssa.MATRIX_BUILDIN, 'BLOSUM62')\n"
            "ssa.init_gap_penalties(-10, -1)\n"
            "ssa.init_db(%r)\n"
            "q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, 'MKV')\n"
            "ssa.sw_align(q, 5, ssa.BIT_WIDTH_16, ssa.COMPUTE_SCORE)\n" % (ROOT, path))
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
        assert r.returncode == 1
        assert "Unable to re-allocate enough memory" in r.stderr + r.stdout


@pytest.mark.parametrize("bw", [8, 16, 64])
def test_translated_query_with_empty_frames(bw):
    # a 4-nt query: frames of 1, 1 and 0 residues on each strand; the six
    # buffers run as one packed query, the empty ones score 0 against every entry
    rnd = random.Random(4242 + bw)
    db = [rand_aa(rnd, rnd.randint(1, 60)) for _ in range(120)]
    check(O.SW, bw, O.TRANS_QUERY, 3, (1, 1), "ACGT", db, 30, ("builtin", "BLOSUM62"), (-10, -1))
