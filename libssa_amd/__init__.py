"""Python mirror of the libssa.h API over libssa_amd.so (the MI355X build).

The functions keep the reference's names, argument meaning and error behaviour
(src/libssa.h:122-263): misuse ends the process through the library's fatal
path, file errors print and return None.  Searches run on the GPU only; if the
shared library is missing, importing this package raises.

    import libssa_amd as ssa
    ssa.init_score_matrix(ssa.MATRIX_BUILDIN, ssa.BLOSUM62)
    ssa.init_gap_penalties(-10, -1)
    ssa.init_db("db.fasta")
    q = ssa.init_sequence_fasta(ssa.READ_FROM_STRING, "MKVLAT...")
    hits = ssa.sw_align(q, 10, ssa.BIT_WIDTH_8, ssa.COMPUTE_SCORE)
    ssa.free_sequence(q)
"""
import ctypes as _C
import os as _os
from dataclasses import dataclass as _dataclass

_HERE = _os.path.dirname(_os.path.abspath(__file__))
# SSA_AMD_LIB names another build of the library to load (the sanitizer
# build of the host code, for instance)
LIB_PATH = _os.environ.get("SSA_AMD_LIB") or _os.path.join(_HERE, "libssa_amd.so")

# ---- constants (src/libssa.h:29-72) ----
BLOSUM45, BLOSUM50, BLOSUM62, BLOSUM80, BLOSUM90 = "blosum45", "blosum50", "blosum62", "blosum80", "blosum90"
PAM30, PAM70, PAM250 = "pam30", "pam70", "pam250"
NUCLEOTIDE, AMINOACID, TRANS_QUERY, TRANS_DB, TRANS_BOTH = 0, 1, 2, 3, 4
FORWARD_STRAND, COMPLEMENTARY_STRAND, BOTH_STRANDS = 1, 2, 3
BIT_WIDTH_8, BIT_WIDTH_16, BIT_WIDTH_64 = 8, 16, 64
OUTPUT_SILENT, OUTPUT_ERROR, OUTPUT_WARNING, OUTPUT_INFO = 0, 1, 2, 3
COMPUTE_SCORE, COMPUTE_ALIGNMENT = 0, 1
READ_FROM_FILE, READ_FROM_STRING, MATRIX_BUILDIN = 0, 1, 2
COMPUTE_ON_SSE2, COMPUTE_ON_SSE41, COMPUTE_ON_AVX2 = 0, 1, 2


class _DbSeq(_C.Structure):
    _fields_ = [("seq", _C.POINTER(_C.c_uint8)), ("len", _C.c_size_t), ("ID", _C.c_size_t),
                ("strand", _C.c_int), ("frame", _C.c_int)]


class _QSeq(_C.Structure):
    _fields_ = [("seq", _C.POINTER(_C.c_uint8)), ("len", _C.c_size_t), ("strand", _C.c_int), ("frame", _C.c_int)]


class _Alignment(_C.Structure):
    _fields_ = [("db_seq", _DbSeq), ("query", _QSeq), ("alignment", _C.c_char_p), ("alignment_len", _C.c_size_t),
                ("score", _C.c_long), ("align_q_start", _C.c_size_t), ("align_q_end", _C.c_size_t),
                ("align_d_start", _C.c_size_t), ("align_d_end", _C.c_size_t)]


class _AlignmentList(_C.Structure):
    _fields_ = [("alignments", _C.POINTER(_C.POINTER(_Alignment))), ("len", _C.c_size_t)]


class _SeqInfo(_C.Structure):
    _fields_ = [("ID", _C.c_size_t), ("seqlen", _C.c_size_t), ("seq", _C.POINTER(_C.c_char))]


class _Request(_C.Structure):
    """ssa_amd_request_t (include/libssa_amd.h)."""
    _fields_ = [("search_type", _C.c_int), ("query", _C.c_void_p), ("hitcount", _C.c_size_t),
                ("bit_width", _C.c_int), ("align_type", _C.c_int)]


SEARCH_SW, SEARCH_NW = 0, 1


class Hit(_C.Structure):
    """One hit of align_noalloc (ssa_amd_hit_t)."""
    _fields_ = [("db_id", _C.c_size_t), ("db_strand", _C.c_int), ("db_frame", _C.c_int),
                ("query_strand", _C.c_int), ("query_frame", _C.c_int), ("score", _C.c_long)]


class AdmissionRequest(_C.Structure):
    _fields_ = [("wait_ms", _C.c_double), ("cells", _C.c_double), ("device_bytes", _C.c_size_t),
                ("occupancy", _C.c_double)]


class AdmissionState(_C.Structure):
    _fields_ = [("max_batch", _C.c_size_t), ("batch_cells", _C.c_double), ("aging_ms", _C.c_double),
                ("assist_ms", _C.c_double), ("device_budget", _C.c_size_t), ("device_in_use", _C.c_size_t),
                ("occupancy_in_use", _C.c_double), ("running", _C.c_int)]


class ServerCounters(_C.Structure):
    """ssa_amd_server_counters_t (include/libssa_amd.h)."""
    _fields_ = [("batches", _C.c_size_t), ("requests", _C.c_size_t), ("max_batch_seen", _C.c_size_t),
                ("queued", _C.c_size_t), ("device_budget", _C.c_size_t), ("deferrals", _C.c_size_t),
                ("device_in_use", _C.c_size_t)]


class Stats(_C.Structure):
    """ssa_amd_stats_t (include/libssa_amd.h)."""
    _fields_ = [("chunk_count", _C.c_size_t), ("seq_count", _C.c_size_t),
                ("overflow_8_bit_count", _C.c_size_t), ("overflow_16_bit_count", _C.c_size_t),
                ("cells", _C.c_uint64), ("cells_computed", _C.c_uint64), ("cells_scored", _C.c_uint64),
                ("lanes_per_stream", _C.c_int), ("rows_per_lane", _C.c_int),
                ("overflow_tracking", _C.c_int),
                ("ms_gpu_first", _C.c_double), ("ms_gpu_requeue", _C.c_double), ("ms_replay", _C.c_double),
                ("ms_traceback", _C.c_double), ("ms_total", _C.c_double), ("d2h_bytes", _C.c_uint64),
                ("replayed_chunks", _C.c_uint64), ("ms_graph", _C.c_double), ("graph_reused", _C.c_int),
                ("columns_per_step", _C.c_int), ("profile_placement", _C.c_int),
                ("cell_arithmetic", _C.c_int), ("ms_setup", _C.c_double)]


class TaskRecord(_C.Structure):
    """ssa_amd_task_record_t (include/libssa_amd.h)."""
    _fields_ = [("start_ns", _C.c_uint64), ("end_ns", _C.c_uint64), ("workgroup", _C.c_int32),
                ("steps", _C.c_int32), ("launch", _C.c_int32), ("pad", _C.c_int32)]


@_dataclass
class Alignment:
    """One entry of the returned alignment list (alignment_t, src/libssa.h:100-110)."""
    db_id: int
    db_seq: bytes
    db_strand: int
    db_frame: int
    query: bytes
    query_strand: int
    query_frame: int
    score: int
    alignment: str
    align_q_start: int
    align_q_end: int
    align_d_start: int
    align_d_end: int


def _load():
    if not _os.path.exists(LIB_PATH):
        raise ImportError("libssa_amd.so is not built; run __graft_entry__.build() (make -C libssa_amd)")
    lib = _C.CDLL(LIB_PATH)
    P, S, I, cp = _C.c_void_p, _C.c_size_t, _C.c_int, _C.c_char_p
    sig = {
        "set_output_mode": (None, [I]), "set_simd_compute_mode": (None, [I]), "set_chunk_size": (None, [S]),
        "set_thread_count": (None, [S]), "init_score_matrix": (None, [I, cp]),
        "init_constant_scores": (None, [_C.c_int8, _C.c_int8]), "init_gap_penalties": (None, [_C.c_int8, _C.c_int8]),
        "init_symbol_translation": (None, [I, I, I, I]), "init_db": (None, [cp]),
        "init_sequence_fasta": (P, [I, cp]), "free_sequence": (None, [P]),
        "sw_align": (_C.POINTER(_AlignmentList), [P, S, I, I]), "nw_align": (_C.POINTER(_AlignmentList), [P, S, I, I]),
        "free_alignment": (None, [_C.POINTER(_AlignmentList)]), "ssa_exit": (None, []),
        "ssa_db_init": (I, [cp]), "ssa_db_get_sequence_count": (S, []), "ssa_db_close": (None, []),
        "ssa_db_get_sequence": (_C.POINTER(_SeqInfo), [S]), "ssa_amd_db_pack": (I, [cp]),
        "ssa_amd_last_stats": (None, [_C.POINTER(Stats)]), "ssa_amd_reset_settings": (None, []),
        "ssa_amd_set_kernel_shape": (None, [I, I]), "ssa_amd_set_stream_waves": (None, [I]),
        "ssa_amd_set_kernel_variant": (None, [I, I]),
        "ssa_amd_set_cell_arithmetic": (None, [I]),
        "ssa_amd_set_task_trace": (None, [I]), "ssa_amd_task_trace": (S, [_C.POINTER(TaskRecord), S]),
        "ssa_amd_autotune_choice": (I, [_C.c_size_t, _C.POINTER(I), _C.POINTER(I), _C.POINTER(I), _C.POINTER(I),
                                        _C.POINTER(_C.c_double)]),
        "ssa_amd_get_matrix": (None, [P]), "ssa_amd_device_arch": (cp, []), "ssa_amd_device_cus": (I, []),
        "ssa_amd_batch_align": (I, [_C.POINTER(_Request), S, _C.POINTER(_C.POINTER(_AlignmentList)),
                                    _C.POINTER(Stats)]),
        "ssa_amd_autotune": (I, [I, _C.POINTER(_C.c_size_t), S, I, _C.POINTER(I), _C.POINTER(I),
                                 _C.POINTER(_C.c_double)]),
        "ssa_amd_autotune_reset": (None, []),
        "ssa_amd_selfcheck_replay": (I, [_C.c_uint64, I]),
        "ssa_amd_server_start": (P, [I, S, S]),
        "ssa_amd_server_submit": (_C.c_int64, [P, _C.POINTER(_Request)]),
        "ssa_amd_server_wait": (_C.POINTER(_AlignmentList), [P, _C.c_int64, _C.POINTER(Stats),
                                                             _C.POINTER(_C.c_double), _C.POINTER(_C.c_double)]),
        "ssa_amd_server_counters": (None, [P, _C.POINTER(ServerCounters)]),
        "ssa_amd_server_tune": (None, [P, _C.c_double, _C.c_double]),
        "ssa_amd_server_stop": (None, [P]),
        "ssa_amd_align_noalloc": (I, [I, P, S, I, _C.POINTER(Hit), _C.POINTER(_C.c_size_t)]),
        "ssa_amd_admission_plan": (S, [_C.POINTER(AdmissionRequest), S, _C.POINTER(AdmissionState),
                                       _C.POINTER(_C.c_size_t)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib


lib = _load()


def _b(s):
    return s.encode() if isinstance(s, str) else s


# ---- technical initialisation ----
def set_output_mode(mode): lib.set_output_mode(mode)
def set_simd_compute_mode(mode): lib.set_simd_compute_mode(mode)
def set_chunk_size(size): lib.set_chunk_size(size)
def set_thread_count(count): lib.set_thread_count(count)


# ---- initialisation ----
def init_score_matrix(mode, matrix): lib.init_score_matrix(mode, _b(matrix))
def init_constant_scores(p, m): lib.init_constant_scores(p, m)
def init_gap_penalties(gap_open, gap_extend): lib.init_gap_penalties(gap_open, gap_extend)
def init_symbol_translation(type_, strands, db_gencode, q_gencode):
    lib.init_symbol_translation(type_, strands, db_gencode, q_gencode)
def init_db(path): lib.init_db(_b(path))


def init_sequence_fasta(mode, text):
    """Returns an opaque query handle, or None if the file could not be read."""
    h = lib.init_sequence_fasta(mode, _b(text))
    return h if h else None


def free_sequence(q): lib.free_sequence(q)
def ssa_exit(): lib.ssa_exit()


def _convert(ptr):
    out = []
    lst = ptr.contents
    for i in range(lst.len):
        a = lst.alignments[i].contents
        out.append(Alignment(
            db_id=a.db_seq.ID, db_seq=_C.string_at(a.db_seq.seq, a.db_seq.len), db_strand=a.db_seq.strand,
            db_frame=a.db_seq.frame, query=_C.string_at(a.query.seq, a.query.len), query_strand=a.query.strand,
            query_frame=a.query.frame, score=a.score,
            alignment=a.alignment.decode() if a.alignment else None,
            align_q_start=a.align_q_start, align_q_end=a.align_q_end,
            align_d_start=a.align_d_start, align_d_end=a.align_d_end))
    return out


def sw_align(query, hitcount, bit_width, align_type):
    """Smith-Waterman search; returns the alignment list as Alignment objects."""
    p = lib.sw_align(query, hitcount, bit_width, align_type)
    try:
        return _convert(p)
    finally:
        lib.free_alignment(p)


def nw_align(query, hitcount, bit_width, align_type):
    """Needleman-Wunsch search; returns the alignment list as Alignment objects."""
    p = lib.nw_align(query, hitcount, bit_width, align_type)
    try:
        return _convert(p)
    finally:
        lib.free_alignment(p)


# ---- extensions (include/libssa_amd.h) ----
def last_stats():
    s = Stats()
    lib.ssa_amd_last_stats(_C.byref(s))
    return s


def reset_settings():
    """Every libssa.h setting back to its initial value (ssa_amd_reset_settings);
    the matrix and the DB stay."""
    lib.ssa_amd_reset_settings()


def set_kernel_shape(lanes_per_stream, rows_per_lane): lib.ssa_amd_set_kernel_shape(lanes_per_stream, rows_per_lane)


def set_kernel_variant(columns_per_step=0, profile_placement=0):
    """Column tile (4 or 8 DB columns per step) and profile placement (1 LDS,
    2 VGPRs) of the score kernel; 0 = chosen per search."""
    lib.ssa_amd_set_kernel_variant(columns_per_step, profile_placement)


def set_stream_waves(waves): lib.ssa_amd_set_stream_waves(waves)


def set_cell_arithmetic(arithmetic):
    """0: the SW first tier in fp16 windows where the matrix and gaps allow
    (the default), 1: packed int16 everywhere (ssa_amd_set_cell_arithmetic)."""
    lib.ssa_amd_set_cell_arithmetic(arithmetic)


def get_matrix():
    buf = (_C.c_int64 * 1024)()
    lib.ssa_amd_get_matrix(buf)
    return list(buf)


def device_cus():
    """Compute units of the GPU (ssa_amd_device_cus)."""
    return lib.ssa_amd_device_cus()


def device_arch():
    return lib.ssa_amd_device_arch().decode()


def db_sequence(index):
    """(ID, residues as bytes) of the plugin DB's sequence, or None past the end."""
    p = lib.ssa_db_get_sequence(index)
    if not p:
        return None
    s = p.contents
    return s.ID, _C.string_at(s.seq, s.seqlen) if s.seqlen else b""


def pack_db(path):
    """Writes the DB currently open to a pre-packed file (ssa_amd_db_pack),
    mapped for the symbol type in force; init_db(path) loads it."""
    if lib.ssa_amd_db_pack(_b(path)) != 0:
        raise RuntimeError("cannot write packed database %s" % path)


def batch_align(requests):
    """Runs searches as one batch.  requests: (search_type, query, hitcount,
    bit_width, align_type) tuples with search_type SEARCH_SW or SEARCH_NW.
    Returns (alignment lists, Stats records), one of each per request."""
    n = len(requests)
    if n == 0:
        return [], []
    arr = (_Request * n)(*[_Request(t, q, k, bw, al) for (t, q, k, bw, al) in requests])
    outs = (_C.POINTER(_AlignmentList) * n)()
    stats = (Stats * n)()
    rc = lib.ssa_amd_batch_align(arr, n, outs, stats)
    if rc != 0:
        raise ValueError("ssa_amd_batch_align rejected the batch (code %d)" % rc)
    lists = []
    for i in range(n):
        try:
            lists.append(_convert(outs[i]))
        finally:
            lib.free_alignment(outs[i])
    return lists, list(stats)


class Server:
    """Continuous-batching request loop (ssa_amd_server_*): submit() queues a
    search and returns a ticket at once; wait(ticket) returns (alignments,
    Stats, latency_ms, queue_ms).  Queries must outlive their results."""

    def __init__(self, slots=2, max_batch=64, device_bytes=0):
        self.h = lib.ssa_amd_server_start(slots, max_batch, device_bytes)
        if not self.h:
            raise ValueError("ssa_amd_server_start failed (arguments or settings invalid)")

    def submit(self, search_type, query, hitcount, bit_width, align_type):
        r = _Request(search_type, query, hitcount, bit_width, align_type)
        t = lib.ssa_amd_server_submit(self.h, _C.byref(r))
        if t < 0:
            raise ValueError("invalid request")
        return t

    def wait(self, ticket):
        st = Stats()
        lat, q = _C.c_double(), _C.c_double()
        p = lib.ssa_amd_server_wait(self.h, ticket, _C.byref(st), _C.byref(lat), _C.byref(q))
        if not p:
            raise KeyError(ticket)
        try:
            return _convert(p), st, lat.value, q.value
        finally:
            lib.free_alignment(p)

    def tune(self, batch_cells=0.0, aging_ms=0.0):
        lib.ssa_amd_server_tune(self.h, batch_cells, aging_ms)

    def counters(self):
        c = ServerCounters()
        lib.ssa_amd_server_counters(self.h, _C.byref(c))
        return c

    def stop(self):
        if self.h:
            lib.ssa_amd_server_stop(self.h)
            self.h = None


def autotune(search_type, lengths, copies=1):
    """Measures the kernel shapes on the loaded DB (ssa_amd_autotune); returns
    [(length, lanes_per_stream, rows_per_lane, ms per query)]."""
    n = len(lengths)
    L = (_C.c_size_t * n)(*lengths)
    lanes, rows, ms = (_C.c_int * n)(), (_C.c_int * n)(), (_C.c_double * n)()
    k = lib.ssa_amd_autotune(search_type, L, n, copies, lanes, rows, ms)
    if k < 0:
        raise ValueError("ssa_amd_autotune failed (code %d)" % k)
    return [(lengths[i], lanes[i], rows[i], ms[i]) for i in range(k)]


def autotune_choices():
    """The choices of the last autotune (ssa_amd_autotune_choice): [(lanes
    per stream, rows per lane, columns per step, profile placement, ms)]."""
    out, i = [], 0
    while True:
        v = [_C.c_int() for _ in range(4)]
        ms = _C.c_double()
        if lib.ssa_amd_autotune_choice(i, *[_C.byref(x) for x in v], _C.byref(ms)) != 0:
            return out
        out.append(tuple(x.value for x in v) + (ms.value,))
        i += 1


def autotune_reset(): lib.ssa_amd_autotune_reset()


def set_task_trace(on):
    """Record the tasks of the following searches' first-tier score launches
    (ssa_amd_set_task_trace)."""
    lib.ssa_amd_set_task_trace(1 if on else 0)


def task_trace():
    """The last traced search's tasks (ssa_amd_task_trace): [(start ns, end ns,
    workgroup, steps, launch)]."""
    n = lib.ssa_amd_task_trace(None, 0)
    buf = (TaskRecord * max(n, 1))()
    n = lib.ssa_amd_task_trace(buf, n)
    return [(r.start_ns, r.end_ns, r.workgroup, r.steps, r.launch) for r in buf[:n]]


def align_noalloc(search_type, query, hitcount, bit_width):
    """sw_align / nw_align with COMPUTE_SCORE through ssa_amd_align_noalloc
    (no device allocation).  Returns (code, [Hit]); code 0 on success."""
    hits = (Hit * max(hitcount, 1))()
    n = _C.c_size_t()
    rc = lib.ssa_amd_align_noalloc(search_type, query, hitcount, bit_width, hits, _C.byref(n))
    return rc, [hits[i] for i in range(n.value)]


def admission_plan(queue, state):
    """The server's admission policy (ssa_amd_admission_plan): queue is a list
    of (wait_ms, cells, device_bytes, occupancy), oldest first; state a dict of
    AdmissionState fields.  Returns the indices of the requests chosen (empty:
    wait)."""
    n = len(queue)
    arr = (AdmissionRequest * max(n, 1))(*[AdmissionRequest(*r) for r in queue])
    st = AdmissionState(**state)
    out = (_C.c_size_t * max(n, 1))()
    k = lib.ssa_amd_admission_plan(arr, n, _C.byref(st), out)
    return [out[i] for i in range(k)]


def selfcheck_replay(seed, cases):
    """Host-only check of the replay's chunk filter against the full
    simulation; returns the number of differing cases (0 = pass)."""
    return lib.ssa_amd_selfcheck_replay(seed, cases)
