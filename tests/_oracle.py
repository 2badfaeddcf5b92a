"""ctypes access to the oracle (oracle/liboracle.so) and, where it was built,
to the reference's own compiled sources (oracle/_ref/libssaref.so).

Test infrastructure only: the product never imports this module.
"""
import ctypes as C
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DATA = os.path.join(ROOT, "tests", "golden", "data")
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libssaref.so")

SW, NW = 0, 1
NUCLEOTIDE, AMINOACID, TRANS_QUERY, TRANS_DB, TRANS_BOTH = 0, 1, 2, 3, 4


def _ensure_oracle():
    if not os.path.exists(ORACLE_SO):
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle"), "oracle"])
    return C.CDLL(ORACLE_SO)


lib = _ensure_oracle()


class Matrix(C.Structure):
    _fields_ = [("s", C.c_int64 * 1024), ("constant", C.c_int)]


class Event(C.Structure):
    _fields_ = [("item", C.c_int32), ("kind", C.c_int32), ("score", C.c_int64)]


class Elem(C.Structure):
    _fields_ = [("db_id", C.c_uint64), ("db_frame", C.c_uint8), ("db_strand", C.c_uint8),
                ("query_id", C.c_uint8), ("score", C.c_int64)]


class Heap(C.Structure):
    _fields_ = [("alloc", C.c_size_t), ("count", C.c_size_t), ("a", C.POINTER(Elem))]


class Params(C.Structure):
    _fields_ = [("type", C.c_int), ("bit_width", C.c_int), ("symtype", C.c_int), ("strands", C.c_int),
                ("lanes8", C.c_int), ("lanes16", C.c_int), ("chunk_size", C.c_size_t),
                ("hit_count", C.c_size_t), ("gap_open", C.c_int), ("gap_extend", C.c_int),
                ("matrix", C.POINTER(Matrix)), ("q_count", C.c_int),
                ("q_seq", C.c_void_p * 6), ("q_len", C.c_size_t * 6),
                ("db_count", C.c_size_t), ("db_raw", C.POINTER(C.c_char_p)),
                ("db_len", C.POINTER(C.c_size_t)), ("d_gencode", C.c_int)]


class Stats(C.Structure):
    _fields_ = [("chunk_count", C.c_size_t), ("seq_count", C.c_size_t),
                ("overflow_8", C.c_size_t), ("overflow_16", C.c_size_t)]


class Region(C.Structure):
    _fields_ = [("a_begin", C.c_int64), ("a_end", C.c_int64), ("b_begin", C.c_int64), ("b_end", C.c_int64)]


_P, _S, _I = C.c_void_p, C.c_size_t, C.c_int


def _sig(l, name, res, *args):
    f = getattr(l, name)
    f.restype = res
    f.argtypes = list(args)


_sig(lib, "orc_matrix_constant", None, _P, _I, _I)
_sig(lib, "orc_matrix_from_text", _I, _P, C.c_char_p)
_sig(lib, "orc_map_aa", _I, _I)
_sig(lib, "orc_map_nt", _I, _I)
_sig(lib, "orc_map_db", None, _I, C.c_char_p, _S, _P)
_sig(lib, "orc_map_query", _S, _I, C.c_char_p, _S, _P)
_sig(lib, "orc_revcomp", None, C.c_char_p, _S, _P)
_sig(lib, "orc_gencode_valid", _I, _I)
_sig(lib, "orc_translate_table", None, _I, _P)
_sig(lib, "orc_translate", _S, _P, C.c_char_p, _S, _I, _I, _P)
_sig(lib, "orc_full_sw", C.c_int64, C.c_char_p, _S, C.c_char_p, _S, _P, _I, _I)
_sig(lib, "orc_full_nw", C.c_int64, C.c_char_p, _S, C.c_char_p, _S, _P, _I, _I)
_sig(lib, "orc_lane_kernel", _S, _I, _I, _I, _P, _P, _S, C.c_char_p, _S, _P, _I, _I, _P, _P)
_sig(lib, "orc_heap_init", None, _P, _P, _S)
_sig(lib, "orc_heap_add", None, _P, _P)
_sig(lib, "orc_heap_sort", None, _P)
_sig(lib, "orc_search", _S, _P, _P, _P)
_sig(lib, "orc_region_local", _I, C.c_char_p, _S, C.c_char_p, _S, _P, _I, _I, _P)
_sig(lib, "orc_cigar", _S, _I, C.c_char_p, _S, C.c_char_p, _S, _P, _P, _I, _I, _P, _S)


# ---------------------------------------------------------------- inputs
def read_fasta(path):
    """Records of a FASTA file: (header, sequence) with the lines of a record joined."""
    recs = []
    head, parts = None, []
    with open(path) as f:
        for line in f:
            line = line.rstrip("\r\n")
            if line.startswith(">"):
                if head is not None:
                    recs.append((head, "".join(parts)))
                head, parts = line[1:], []
            elif head is not None:
                parts.append(line)
    if head is not None:
        recs.append((head, "".join(parts)))
    return recs


def data(name):
    return os.path.join(DATA, name)


def matrix_constant(match, mismatch):
    m = Matrix()
    lib.orc_matrix_constant(C.byref(m), match, mismatch)
    return m


def matrix_text(text):
    m = Matrix()
    rc = lib.orc_matrix_from_text(C.byref(m), text.encode())
    assert rc == 0, rc
    return m


def matrix_builtin(name):
    with open(os.path.join(DATA, "matrices", name.lower() + ".txt")) as f:
        return matrix_text(f.read())


def map_query(symtype, raw):
    out = (C.c_uint8 * (len(raw) + 1))()
    n = lib.orc_map_query(symtype, raw.encode(), len(raw), out)
    return bytes(out[:n])


def map_db(symtype, raw):
    out = (C.c_uint8 * (len(raw) + 1))()
    lib.orc_map_db(symtype, raw.encode(), len(raw), out)
    return bytes(out[:len(raw)])


def revcomp(codes):
    out = (C.c_uint8 * (len(codes) + 1))()
    lib.orc_revcomp(codes, len(codes), out)
    return bytes(out[:len(codes)])


def translate_table(gencode):
    t = (C.c_int8 * 4096)()
    lib.orc_translate_table(gencode, t)
    return t


def translate(table, dna_codes, strand, frame):
    """Protein codes of one strand and frame of mapped nucleotide codes, or None
    where the sequence is shorter than the frame (the reference exits there)."""
    out = (C.c_uint8 * (len(dna_codes) // 3 + 1))()
    n = lib.orc_translate(table, dna_codes, len(dna_codes), strand, frame, out)
    if n == 2 ** 64 - 1:
        return None
    return bytes(out[:n])


def query_buffers(symtype, strands, raw, q_gencode=1):
    """The query buffers s_create_searchdata builds (src/algo/searcher.c:42-90)."""
    q = map_query(symtype, raw)
    if symtype == NUCLEOTIDE:
        nt = [q, revcomp(q)]
        return [nt[s] for s in range(2) if (s + 1) & strands]
    if symtype in (TRANS_QUERY, TRANS_BOTH):
        t = translate_table(q_gencode)
        return [translate(t, q, s, f) for s in range(2) if (s + 1) & strands for f in range(3)]
    return [q]


# ---------------------------------------------------------------- kernels
def full(type_, d, q, m, go, ge):
    f = lib.orc_full_sw if type_ == SW else lib.orc_full_nw
    return f(d, len(d), q, len(q), C.byref(m), go, ge)


def lane_kernel(type_, width, lanes, seqs, q, m, go, ge):
    n = len(seqs)
    arr = (C.c_char_p * max(n, 1))(*seqs)
    lens = (C.c_size_t * max(n, 1))(*[len(s) for s in seqs])
    ev = (Event * max(n, 1))()
    fb = (C.c_int32 * max(n, 1))()
    ne = lib.orc_lane_kernel(type_, width, lanes, arr, lens, n, q, len(q), C.byref(m), go, ge, ev, fb)
    return [(ev[i].item, ev[i].kind, ev[i].score) for i in range(ne)], list(fb[:n])


def heap_run(k, elems):
    """elems: list of (score, db_id, query_id); returns sorted (score, db_id, query_id)."""
    store = (Elem * max(k, 1))()
    h = Heap()
    lib.orc_heap_init(C.byref(h), store, k)
    for s, i, q in elems:
        e = Elem(db_id=i, query_id=q, score=s)
        lib.orc_heap_add(C.byref(h), C.byref(e))
    lib.orc_heap_sort(C.byref(h))
    return [(store[i].score, store[i].db_id, store[i].query_id) for i in range(h.count)]


def search(type_, bit_width, symtype, strands, queries, db, k, m, go, ge, chunk=1000, lanes8=32, lanes16=16,
           d_gencode=1, frames=False):
    """queries: mapped query buffers; db: list of raw DB strings in ID order.
    Returns ([(score, db_id, strand, query_id)], stats), or with frames=True
    ([(score, db_id, strand, query_id, frame)], stats)."""
    p = Params()
    p.type, p.bit_width, p.symtype, p.strands = type_, bit_width, symtype, strands
    p.lanes8, p.lanes16, p.chunk_size, p.hit_count = lanes8, lanes16, chunk, k
    p.gap_open, p.gap_extend = go, ge
    p.matrix = C.pointer(m)
    p.q_count = len(queries)
    keep = []
    for i, q in enumerate(queries):
        buf = C.create_string_buffer(q, len(q) + 1)
        keep.append(buf)
        p.q_seq[i] = C.cast(buf, C.c_void_p)
        p.q_len[i] = len(q)
    raw = [s.encode() for s in db]
    p.db_count = len(raw)
    p.db_raw = (C.c_char_p * max(len(raw), 1))(*raw)
    p.db_len = (C.c_size_t * max(len(raw), 1))(*[len(s) for s in raw])
    out = (Elem * max(k, 1))()
    st = Stats()
    p.d_gencode = d_gencode
    n = lib.orc_search(C.byref(p), out, C.byref(st))
    if n == 2 ** 64 - 1:
        raise ValueError("a DB sequence is shorter than a frame it is translated in")
    if frames:
        return [(out[i].score, out[i].db_id, out[i].db_strand, out[i].query_id, out[i].db_frame)
                for i in range(n)], st
    hits = [(out[i].score, out[i].db_id, out[i].db_strand, out[i].query_id) for i in range(n)]
    return hits, st


def region_local(a, b, m, go, ge):
    r = Region()
    rc = lib.orc_region_local(a, len(a), b, len(b), C.byref(m), go, ge, C.byref(r))
    return None if rc else (r.a_begin, r.a_end, r.b_begin, r.b_end)


def cigar(type_, a, b, m, go, ge, region=None):
    if region is None:
        region = region_local(a, b, m, go, ge) if type_ == SW else (0, len(a) - 1, 0, len(b) - 1)
    r = Region(*region)
    buf = C.create_string_buffer(3 * (len(a) + len(b)) + 64)
    lib.orc_cigar(type_, a, len(a), b, len(b), C.byref(r), C.byref(m), go, ge, buf, len(buf))
    return buf.value.decode(), region


# ---------------------------------------------------------------- reference
def ref_available():
    """True where oracle/_ref was built (decided without opening it)."""
    return os.path.exists(REF_SO)


class LazyRef:
    """The reference's compiled library, opened on the first call through it:
    collecting a test module opens nothing, so a run that selects none of its
    tests (a GPU run, for one) never maps the library."""

    def __init__(self):
        self._lib = None

    def __getattr__(self, name):
        if self._lib is None:
            self._lib = ref_lib()
            if self._lib is None:
                raise RuntimeError("oracle/_ref not built")
        return getattr(self._lib, name)


def ref_lib():
    """The reference's own compiled kernels, or None when oracle/_ref was not built."""
    if not os.path.exists(REF_SO):
        return None
    r = C.CDLL(REF_SO)
    _sig(r, "ref_matrix_constant", None, _I, _I)
    _sig(r, "ref_matrix_builtin", None, C.c_char_p)
    _sig(r, "ref_matrix_string", None, C.c_char_p)
    _sig(r, "ref_matrix_get", None, _P, _P, _P)
    _sig(r, "ref_set_gaps", None, _I, _I)
    _sig(r, "ref_maps", None, _P, _P)
    _sig(r, "ref_revcomp", None, C.c_char_p, _S, _P)
    _sig(r, "ref_full", C.c_int64, _I, C.c_char_p, _S, C.c_char_p, _S)
    _sig(r, "ref_lane_kernel", _S, _I, _I, _I, _P, _P, _S, C.c_char_p, _S, _P, _P, _P, _P)
    _sig(r, "ref_heap_run", _S, _S, _P, _P, _P, _S, _P, _P, _P)
    _sig(r, "ref_align", _I, _I, C.c_char_p, _S, C.c_char_p, _S, _P, _P, _S)
    _sig(r, "ref_translate", None, _I, _I, _I, C.c_char_p, _S, _I, _I, _P, _P)
    return r


def ref_translate(r, db, q_gencode, d_gencode, dna_codes, strand, frame):
    out = (C.c_char * (len(dna_codes) // 3 + 1))()
    n = C.c_size_t()
    r.ref_translate(db, q_gencode, d_gencode, dna_codes, len(dna_codes), strand, frame, out, C.byref(n))
    return bytes(out[:n.value])


def ref_lane_kernel(r, type_, width, avx2, seqs, q):
    n = len(seqs)
    arr = (C.c_char_p * max(n, 1))(*seqs)
    lens = (C.c_size_t * max(n, 1))(*[len(s) for s in seqs])
    ii = (C.c_int64 * max(n, 1))()
    sc = (C.c_int64 * max(n, 1))()
    oi = (C.c_int64 * max(n, 1))()
    no = C.c_size_t()
    ni = r.ref_lane_kernel(type_, width, 1 if avx2 else 0, arr, lens, n, q, len(q), ii, sc, oi, C.byref(no))
    return [(ii[i], sc[i]) for i in range(ni)], list(oi[:no.value])
