"""The query profile in VGPRs (profile placement 2): each lane holds its rows'
scores for A, C, G and T and picks a column's pair with v_perm, for
nucleotide DBs of A, C, G and T only.  Checked against the oracle over the
shapes and modes that have it (SW in int16 and fp16 windows, NW without
overflow trackers in int16 and fp16), the fallback to the LDS profile where
the DB or the search lacks it, and the autotuner's sweep of the placement.
"""
import os
import random
import tempfile

import pytest

import _oracle as O
from test_gpu_parity import check, configure, write_fasta

pytestmark = pytest.mark.gpu

ssa = pytest.importorskip("libssa_amd")

VP_SW_SHAPES = [(8, 4), (16, 4), (32, 4), (64, 4), (8, 8), (16, 8), (32, 8), (64, 8), (8, 16), (16, 16), (32, 16)]
VP_NW_SHAPES = [(8, 8), (16, 8), (32, 8), (8, 16), (16, 16)]


def acgt_db(rnd, qraw, n=250, lo=1, hi=400, copies=8):
    comp = str.maketrans("ACGT", "TGCA")
    db = ["".join(rnd.choice("ACGT") for _ in range(rnd.randint(lo, hi))) for _ in range(n)]
    for c in range(copies):
        a, b = rnd.randint(0, len(qraw) // 3), rnd.randint(2 * len(qraw) // 3, len(qraw))
        frag = "".join(x if rnd.random() < 0.9 and x != "N" else rnd.choice("ACGT") for x in qraw[a:b])
        db[rnd.randrange(n)] = frag.translate(comp)[::-1] if c % 2 else frag
    return db


def run_placed(fn, shape):
    """fn() with the profile forced into VGPRs (and LDS), fp16 and int16
    arithmetic; returns the placements the searches reported."""
    seen = []
    ssa.set_kernel_shape(*shape)
    try:
        for arith in (0, 1):
            ssa.set_cell_arithmetic(arith)
            for place in (2, 1):
                ssa.set_kernel_variant(4, place)
                fn()
                seen.append(ssa.last_stats().profile_placement)
    finally:
        ssa.set_kernel_shape(0, 0)
        ssa.set_kernel_variant(0, 0)
        ssa.set_cell_arithmetic(0)
    return seen


@pytest.mark.parametrize("shape", VP_SW_SHAPES)
def test_vgpr_profile_sw(shape):
    rnd = random.Random(300 + 7 * shape[0] + shape[1])
    qraw = "".join(rnd.choice("ACGT") for _ in range(rnd.choice([90, 400, 1100])))
    qraw = qraw[:40] + "N" + qraw[41:]  # a query code without a slot of its own reads the slots like any other
    db = acgt_db(rnd, qraw)
    # (one strand: both strands may run as paired packed queries, whose
    # kernel keeps the profile in LDS)
    seen = run_placed(lambda: check(O.SW, rnd.choice([8, 16]), O.NUCLEOTIDE, 1, qraw, db, 25,
                                    ("const", (5, -4)), (-10, -2)), shape)
    assert seen == [2, 1, 2, 1]


@pytest.mark.parametrize("shape", VP_NW_SHAPES)
def test_vgpr_profile_nw(shape):
    rnd = random.Random(500 + 7 * shape[0] + shape[1])
    qraw = "".join(rnd.choice("ACGT") for _ in range(250))
    db = acgt_db(rnd, qraw, lo=100, hi=300)
    seen = run_placed(lambda: check(O.NW, 16, O.NUCLEOTIDE, 3, qraw, db, 30, ("const", (2, -4)), (-20, -2),
                                    align=shape == (16, 16)), shape)
    assert seen == [2, 1, 2, 1]


def test_vgpr_profile_falls_back():
    """An N in the DB, NW with the overflow trackers, or a 32-row shape: the
    profile stays in LDS, and the results are the oracle's."""
    rnd = random.Random(9)
    qraw = "".join(rnd.choice("ACGT") for _ in range(200))
    db = acgt_db(rnd, qraw)
    with_n = list(db)
    with_n[3] = with_n[3][:5] + "N" + with_n[3][6:] if len(with_n[3]) > 6 else "ACGTNACGT"
    assert run_placed(lambda: check(O.SW, 16, O.NUCLEOTIDE, 1, qraw, with_n, 20, ("const", (5, -4)), (-10, -2)),
                      (16, 8)) == [1, 1, 1, 1]
    # NW at 8 bit keeps its trackers (or skips its 8-bit launch)
    assert 2 not in run_placed(lambda: check(O.NW, 8, O.NUCLEOTIDE, 1, qraw[:60], [s[:80] for s in db], 20,
                                             ("const", (2, -4)), (-20, -2)), (16, 8))
    assert run_placed(lambda: check(O.SW, 16, O.NUCLEOTIDE, 1, qraw, db, 20, ("const", (5, -4)), (-10, -2)),
                      (16, 32)) == [1, 1, 1, 1]


def test_autotune_sweeps_the_placement():
    """On a DNA DB of A, C, G and T the autotuner measures both placements and
    a search of a tuned length runs the shape, tile and placement it chose."""
    rnd = random.Random(21)
    qraw = "".join(rnd.choice("ACGT") for _ in range(300))
    db = acgt_db(rnd, qraw, n=600)
    configure(O.NUCLEOTIDE, 1, (-10, -2), ("const", (5, -4)))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        try:
            res = ssa.autotune(ssa.SEARCH_SW, [60, 300])
            choices = ssa.autotune_choices()
            assert len(choices) == len(res) == 2
            for (L, lanes, rows, ms), ch in zip(res, choices):
                assert (lanes, rows) == ch[:2] and ch[2] in (4, 8) and ch[3] in (1, 2) and ch[4] == ms
            q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, qraw)
            got = ssa.sw_align(q, 10, 16, ssa.COMPUTE_SCORE)
            st = ssa.last_stats()
            assert (st.lanes_per_stream, st.rows_per_lane, st.columns_per_step, st.profile_placement) == choices[1][:4]
            m = O.matrix_constant(5, -4)
            want, _ = O.search(O.SW, 16, O.NUCLEOTIDE, 1, O.query_buffers(O.NUCLEOTIDE, 1, qraw), db, 10, m, -10, -2)
            assert [(a.score, a.db_id) for a in got] == [(s, i) for s, i, _, _ in want]
            ssa.free_sequence(q)
        finally:
            ssa.autotune_reset()
