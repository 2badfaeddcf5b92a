"""The score kernels' task trace (ssa_amd_set_task_trace / ssa_amd_task_trace):
tracing leaves the results the oracle's, and the records describe the launch
(every task once, on a workgroup of the grid, the first ones by index, end
after start, steps for the passes of the query)."""
import os
import random
import tempfile

import pytest

import _oracle as O
from test_gpu_parity import check, configure, write_fasta

pytestmark = pytest.mark.gpu

ssa = pytest.importorskip("libssa_amd")

AA = "ARNDCQEGHILKMFPSTWYV"


def traced(fn):
    ssa.set_task_trace(True)
    try:
        fn()
        return ssa.task_trace()
    finally:
        ssa.set_task_trace(False)


def check_records(recs, passes, rows):
    assert recs
    grid = max(r[2] for r in recs) + 1
    assert grid <= len(recs)
    for i, (start, end, wg, steps, launch) in enumerate(recs):
        assert end >= start
        assert 0 <= wg < grid
        if i < grid:  # each workgroup's first task is its index
            assert wg == i
        assert steps == 0 or steps % passes == 0
    assert min(r[0] for r in recs) == 0


@pytest.mark.parametrize("qlen,shape", [(300, (16, 8)), (1100, (32, 16)), (1100, (16, 16))])
def test_trace_single_search(qlen, shape):
    rnd = random.Random(qlen + shape[0])
    qraw = "".join(rnd.choice(AA) for _ in range(qlen))
    db = ["".join(rnd.choice(AA) for _ in range(rnd.randint(1, 500))) for _ in range(400)]
    db[17] = qraw[100:700]
    ssa.set_kernel_shape(*shape)
    try:
        recs = traced(lambda: check(O.SW, 16, O.AMINOACID, 1, qraw, db, 20, ("builtin", "blosum62"), (-10, -1)))
    finally:
        ssa.set_kernel_shape(0, 0)
    rows = shape[0] * shape[1]
    check_records(recs, (qlen + rows - 1) // rows, rows)
    assert {r[4] for r in recs} == {0}
    # tracing off again: no records are replaced
    check(O.SW, 16, O.AMINOACID, 1, qraw, db, 20, ("builtin", "blosum62"), (-10, -1))
    assert ssa.task_trace() == recs


def test_trace_batch_launches():
    """A batch of SW and NW searches: one launch per first-tier group, each
    with its own records, and every result the oracle's."""
    rnd = random.Random(8)
    qraws = ["".join(rnd.choice(AA) for _ in range(L)) for L in (60, 400, 900)]
    db = ["".join(rnd.choice(AA) for _ in range(rnd.randint(1, 300))) for _ in range(300)]
    m = O.matrix_builtin("blosum62")
    configure(O.AMINOACID, 1, (-10, -1), ("builtin", "blosum62"))
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "db.fas")
        write_fasta(path, db)
        ssa.init_db(path)
        qs = [ssa.init_sequence_fasta(ssa.READ_FROM_STRING, q) for q in qraws]
        try:
            reqs = [(ssa.SEARCH_SW, qs[0], 10, 16, ssa.COMPUTE_SCORE), (ssa.SEARCH_SW, qs[1], 10, 16, ssa.COMPUTE_SCORE),
                    (ssa.SEARCH_NW, qs[2], 10, 16, ssa.COMPUTE_SCORE)]
            out = {}
            recs = traced(lambda: out.setdefault("lists", ssa.batch_align(reqs)[0]))
            for (t, _, k, bw, _), qraw, got in zip(reqs, qraws, out["lists"]):
                want, _ = O.search(O.SW if t == ssa.SEARCH_SW else O.NW, bw, O.AMINOACID, 1,
                                   [O.map_query(O.AMINOACID, qraw)], db, k, m, -10, -1)
                assert [(a.score, a.db_id) for a in got] == [(s, i) for s, i, _, _ in want]
            launches = sorted({r[4] for r in recs})
            assert len(launches) >= 2
            for g in launches:
                sub = [r for r in recs if r[4] == g]
                grid = max(r[2] for r in sub) + 1
                assert all(sub[i][2] == i for i in range(grid))
                assert all(r[1] >= r[0] for r in sub)
        finally:
            for q in qs:
                ssa.free_sequence(q)
