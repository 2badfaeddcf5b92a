"""The oracle against the reference's own compiled sources (oracle/_ref).

oracle/_ref is built from the reference's kernels, heap, matrices and
traceback (see oracle/Makefile); it exists only where the reference tree was
mounted at build time.  Inputs are random but seeded, and chosen to reach the
corner cases: 8-bit and 16-bit saturation, sequence lengths that leave partial
4-column blocks, large gap penalties whose sums wrap in the 8-bit lane type,
matrices with entries outside int8, and many tied scores.
"""
import ctypes as C
import random

import pytest

import _oracle as O

REF = O.LazyRef()
pytestmark = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built (reference sources absent)")

NAMES = ["blosum45", "blosum50", "blosum62", "blosum80", "blosum90", "pam30", "pam70", "pam250"]


def ref_matrix(name=None, text=None, const=None):
    if const is not None:
        REF.ref_matrix_constant(*const)
    elif name is not None:
        REF.ref_matrix_builtin(name.encode())
    else:
        REF.ref_matrix_string(text.encode())
    m64 = (C.c_int64 * 1024)()
    m8 = (C.c_int8 * 1024)()
    m16 = (C.c_int16 * 1024)()
    REF.ref_matrix_get(m64, m8, m16)
    return list(m64), list(m8), list(m16)


@pytest.mark.parametrize("name", NAMES)
def test_builtin_tables_match_reference(name):
    m64, m8, m16 = ref_matrix(name=name)
    mine = O.matrix_builtin(name)
    assert list(mine.s) == m64
    assert m8 == [((v + 128) % 256) - 128 for v in m64]
    assert m16 == m64


def test_symbol_maps_match_reference():
    aa = (C.c_int8 * 256)()
    nt = (C.c_int8 * 256)()
    REF.ref_maps(aa, nt)
    for c in range(128):
        assert O.lib.orc_map_aa(c) == aa[c], chr(c)
        assert O.lib.orc_map_nt(c) == nt[c], chr(c)


def test_revcomp_matches_reference():
    rnd = random.Random(7)
    s = bytes(rnd.randrange(16) for _ in range(300))
    out = C.create_string_buffer(len(s) + 1)
    REF.ref_revcomp(s, len(s), out)
    assert out.raw[:len(s)] == O.revcomp(s)


def random_matrix_text(rnd, lo, hi, symbols="ARNDCQEGHILKMFPSTWYVBJZX*"):
    rows = ["   " + "  ".join(symbols)]
    for a in symbols:
        rows.append(a + " " + " ".join(str(rnd.randint(lo, hi)) for _ in symbols))
    return "\n".join(rows) + "\n"


def scoring(rnd, kind):
    if kind == "negative":  # mismatches far below two gap opens: E/F boundaries decide cells
        text = random_matrix_text(rnd, -60, 8)
        REF.ref_matrix_string(text.encode())
        return O.matrix_text(text), O.AMINOACID
    if kind == "const":
        match, mismatch = rnd.choice([(1, -1), (2, -3), (5, -4), (30, -1), (127, -1), (11, -2)])
        REF.ref_matrix_constant(match, mismatch)
        return O.matrix_constant(match, mismatch), O.NUCLEOTIDE
    if kind == "blosum62":
        REF.ref_matrix_builtin(b"blosum62")
        return O.matrix_builtin("blosum62"), O.AMINOACID
    text = random_matrix_text(rnd, -20, 40) if kind == "wide" else random_matrix_text(rnd, -6, 12)
    REF.ref_matrix_string(text.encode())
    return O.matrix_text(text), O.AMINOACID


def rand_seq(rnd, symtype, n):
    alpha = "ACGT" if symtype == O.NUCLEOTIDE else "ARNDCQEGHILKMFPSTWYV"
    if rnd.random() < 0.2:  # low-complexity runs drive scores to saturation
        return O.map_db(symtype, rnd.choice(alpha) * n)
    return O.map_db(symtype, "".join(rnd.choice(alpha) for _ in range(n)))


GAPS = [(-1, -1), (-10, -1), (-4, -2), (-20, -2), (-3, -1), (-100, -50), (-128, -1), (-60, -60)]


@pytest.mark.parametrize("seed", range(100))
def test_lane_kernels_match_reference(seed):
    rnd = random.Random(1000 + seed)
    kind = ["const", "blosum62", "narrow", "wide", "negative"][seed % 5]
    m, symtype = scoring(rnd, kind)
    go, ge = GAPS[seed % len(GAPS)] if seed % 3 else (rnd.randint(-8, 0), rnd.randint(-4, -1))
    REF.ref_set_gaps(go, ge)
    n = rnd.randint(1, 70)
    seqs = [rand_seq(rnd, symtype, rnd.choice([1, 2, 3, 4, 5, 7, 8, 13, rnd.randint(1, 60), rnd.randint(60, 200)]))
            for _ in range(n)]
    q = rand_seq(rnd, symtype, rnd.choice([1, 2, 5, rnd.randint(1, 40), rnd.randint(40, 140)]))
    for t in (O.SW, O.NW):
        for width in (8, 16):
            for avx2 in (True, False):
                lanes = {(8, True): 32, (8, False): 16, (16, True): 16, (16, False): 8}[(width, avx2)]
                ref_ins, ref_ovf = O.ref_lane_kernel(REF, t, width, avx2, seqs, q)
                ev, _ = O.lane_kernel(t, width, lanes, seqs, q, m, go, ge)
                mine_ins = [(i, s) for i, k, s in ev if k == 0]
                mine_ovf = [i for i, k, s in ev if k == 1]
                assert mine_ins == ref_ins, (t, width, avx2)
                assert mine_ovf == ref_ovf, (t, width, avx2)


@pytest.mark.parametrize("seed", range(20))
def test_full64_matches_reference(seed):
    rnd = random.Random(2000 + seed)
    m, symtype = scoring(rnd, ["const", "blosum62", "wide"][seed % 3])
    go, ge = GAPS[seed % len(GAPS)]
    REF.ref_set_gaps(go, ge)
    for _ in range(20):
        d = rand_seq(rnd, symtype, rnd.randint(1, 80))
        q = rand_seq(rnd, symtype, rnd.randint(1, 80))
        for t in (O.SW, O.NW):
            assert O.full(t, d, q, m, go, ge) == REF.ref_full(t, d, len(d), q, len(q))


@pytest.mark.parametrize("seed", range(30))
def test_heap_matches_reference(seed):
    """Scores drift upwards with heavy ties, so equal-score elements are evicted
    from the root and the heap layout decides which of them survive."""
    rnd = random.Random(3000 + seed)
    k = rnd.randint(1, 20)
    n = rnd.randint(0, 600)
    drift = rnd.choice([0, 0.01, 0.03, 0.1])
    elems = [(int(i * drift) + rnd.randint(0, 6), rnd.randint(0, 50), rnd.randint(0, 5)) for i in range(n)]
    ids = (C.c_int64 * max(n, 1))(*[e[1] for e in elems])
    sc = (C.c_int64 * max(n, 1))(*[e[0] for e in elems])
    qs = (C.c_int32 * max(n, 1))(*[e[2] for e in elems])
    oi = (C.c_int64 * k)()
    os_ = (C.c_int64 * k)()
    oq = (C.c_int32 * k)()
    c = REF.ref_heap_run(k, ids, sc, qs, n, oi, os_, oq)
    assert O.heap_run(k, elems) == [(os_[i], oi[i], oq[i]) for i in range(c)]


@pytest.mark.parametrize("seed", range(20))
def test_region_and_cigar_match_reference(seed):
    rnd = random.Random(4000 + seed)
    m, symtype = scoring(rnd, ["const", "blosum62", "narrow"][seed % 3])
    go, ge = GAPS[seed % 5]
    REF.ref_set_gaps(go, ge)
    done = 0
    while done < 15:
        a = rand_seq(rnd, symtype, rnd.randint(1, 60))
        b = rand_seq(rnd, symtype, rnd.randint(1, 60))
        for t in (O.SW, O.NW):
            if t == O.SW and O.full(O.SW, b, a, m, go, ge) <= 0:
                continue  # the reference leaves the end cell uninitialised for score 0
            reg = (C.c_int64 * 4)()
            buf = C.create_string_buffer(4 * (len(a) + len(b)) + 64)
            REF.ref_align(t, a, len(a), b, len(b), reg, buf, len(buf))
            c, r = O.cigar(t, a, b, m, go, ge)
            assert r == tuple(reg), (t, a, b)
            assert c == buf.value.decode(), (t, a, b)
        done += 1


@pytest.mark.parametrize("seed", range(8))
def test_lane_kernels_16bit_saturation_match_reference(seed):
    """Inputs that overflow the 16-bit lanes: long near-identical pairs with a
    match score of 127 (top saturation, SW and NW) and a matrix of large
    negative entries (NW bottom saturation)."""
    rnd = random.Random(5000 + seed)
    if seed % 2 == 0:
        go, ge = GAPS[seed % len(GAPS)]
    else:  # expensive gaps so that negative cells reach the 16-bit floor
        go, ge = [(-100, -50), (-60, -60), (-128, -120), (-1, -127)][(seed // 2) % 4]
    REF.ref_set_gaps(go, ge)
    q = rand_seq(rnd, O.AMINOACID, rnd.randint(260, 560))
    seqs = []
    for _ in range(rnd.randint(3, 40)):
        r = rnd.random()
        if r < 0.4:
            s = bytearray(q)
            for _ in range(rnd.randint(0, 20)):
                s[rnd.randrange(len(s))] = rnd.randrange(1, 21)
            cut = rnd.randint(0, 40)
            seqs.append(bytes(s[cut:len(s) - rnd.randint(0, 40)]))
        else:
            seqs.append(rand_seq(rnd, O.AMINOACID, rnd.randint(1, 600)))
    if seed % 2 == 0:
        REF.ref_matrix_constant(127, -1)
        m = O.matrix_constant(127, -1)
    else:
        text = random_matrix_text(rnd, -127, -90)
        REF.ref_matrix_string(text.encode())
        m = O.matrix_text(text)
    seen16 = 0
    for t in (O.SW, O.NW):
        for width in (8, 16):
            for avx2 in (True, False):
                lanes = {(8, True): 32, (8, False): 16, (16, True): 16, (16, False): 8}[(width, avx2)]
                ref_ins, ref_ovf = O.ref_lane_kernel(REF, t, width, avx2, seqs, q)
                ev, _ = O.lane_kernel(t, width, lanes, seqs, q, m, go, ge)
                assert [(i, s) for i, k, s in ev if k == 0] == ref_ins, (t, width, avx2)
                assert [i for i, k, s in ev if k == 1] == ref_ovf, (t, width, avx2)
                if width == 16:
                    seen16 += len(ref_ovf)
    assert seen16 > 0


@pytest.mark.parametrize("seed", range(30))
def test_nw8_boundary_wrap_sweep(seed):
    """Gap penalties whose sums leave the int8 range: the 8-bit NW kernel stores
    the converted (wrapped) boundary constants.  Short queries and sequences keep
    some lanes unflagged so the wrapped values reach accepted scores."""
    rnd = random.Random(6000 + seed)
    go, ge = rnd.randint(-70, 0), rnd.randint(-70, -20)
    text = random_matrix_text(rnd, -60, 40)
    REF.ref_matrix_string(text.encode())
    m = O.matrix_text(text)
    REF.ref_set_gaps(go, ge)
    seqs = [rand_seq(rnd, O.AMINOACID, rnd.randint(1, 9)) for _ in range(rnd.randint(5, 60))]
    q = rand_seq(rnd, O.AMINOACID, rnd.randint(1, 3))
    for width in (8, 16):
        for avx2 in (True, False):
            lanes = {(8, True): 32, (8, False): 16, (16, True): 16, (16, False): 8}[(width, avx2)]
            ref_ins, ref_ovf = O.ref_lane_kernel(REF, O.NW, width, avx2, seqs, q)
            ev, _ = O.lane_kernel(O.NW, width, lanes, seqs, q, m, go, ge)
            assert [(i, s) for i, k, s in ev if k == 0] == ref_ins, (width, avx2)
            assert [i for i, k, s in ev if k == 1] == ref_ovf, (width, avx2)
