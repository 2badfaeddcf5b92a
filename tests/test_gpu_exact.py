"""The exact tier (the reference's 64-bit full_sw / full_nw,
src/algo/64/smith_waterman_63.c:32-98, needleman_wunsch_64.c:32-98) on the GPU
against the oracle, where it carries the results.

- Constant 127/-1 scoring (the reference's own 64-bit trigger,
  tests/algo/test_searcher.c:507-546) over a DB where a thousand mutated
  copies of the query overflow the 16-bit tier: every one reaches the exact
  tier, which runs one wave per pair in int32.
- A matrix with entries in the millions: the scores leave int32, so the exact
  tier computes in int64 (the reference's width) -- a query of three 512-row
  passes, so the wave's pass boundaries are exercised -- where an int32 tier
  would have to give up.
"""
import random

import pytest

import _oracle as O
from test_gpu_parity import check

pytestmark = pytest.mark.gpu

ssa = pytest.importorskip("libssa_amd")

AA = "ARNDCQEGHILKMFPSTWYV"


def mutated_copies(rnd, base, n, identity):
    out = []
    for _ in range(n):
        s = [c if rnd.random() < identity else rnd.choice(AA) for c in base]
        cut = rnd.randint(0, len(s) // 10)
        out.append("".join(s[cut:]))
    return out


@pytest.mark.parametrize("bw", [8, 16, 64])
@pytest.mark.parametrize("t", [O.SW, O.NW])
def test_many_pairs_overflow_16_bit(t, bw):
    rnd = random.Random(127 + bw + 10 * t)
    base = "".join(rnd.choice(AA) for _ in range(700))
    db = mutated_copies(rnd, base, 1000, 0.9)
    db += ["".join(rnd.choice(AA) for _ in range(rnd.randint(20, 700))) for _ in range(500)]
    rnd.shuffle(db)
    got = check(t, bw, O.AMINOACID, 1, base, db, 40, ("const", (127, -1)), (-1, -1))
    st = ssa.last_stats()
    if bw != 64:
        # the copies score ~0.9 * 700 * 127 > 65,535: they overflow 16 bit
        assert st.overflow_16_bit_count >= 1000 if t == O.SW else st.overflow_16_bit_count > 0
    if t == O.SW:
        assert got[0].score > 65535


def big_matrix_text(scale, offset):
    """BLOSUM62 with every entry v written as v * scale + offset."""
    lines = open(O.data("matrices/blosum62.txt")).read().splitlines()
    out = []
    for line in lines:
        if line.startswith("#") or not line.strip() or line[0] == " ":
            out.append(line)
            continue
        parts = line.split()
        out.append(" ".join([parts[0]] + [str(int(v) * scale + offset) for v in parts[1:]]))
    return "\n".join(out) + "\n"


@pytest.mark.parametrize("bw", [16, 64])
@pytest.mark.parametrize("t", [O.SW, O.NW])
def test_int64_scores(t, bw):
    rnd = random.Random(64 + bw + t)
    base = "".join(rnd.choice(AA) for _ in range(1100))
    db = mutated_copies(rnd, base, 40, 0.85) + \
        ["".join(rnd.choice(AA) for _ in range(rnd.randint(50, 1200))) for _ in range(160)]
    rnd.shuffle(db)
    text = big_matrix_text(1000003, 7)
    got = check(t, bw, O.AMINOACID, 1, base, db, 25, ("text", text), (-10, -1))
    if t == O.SW:
        # a copy scores ~0.85 * 1100 * 5e6: far past int32
        assert got[0].score > 2 ** 31
