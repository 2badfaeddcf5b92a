"""Host-side checks that need no GPU: the C-ABI library loads and exports
every function the public headers declare, and the host logic behind the API
(matrices, FASTA provider, query reading) behaves as the reference's tests pin."""
import ctypes as C
import os
import re
import subprocess
import sys

import pytest

import _oracle as O

ROOT = O.ROOT
ssa = pytest.importorskip("libssa_amd")


def declared_functions():
    names = []
    for h in ("libssa.h", "libssa_extern_db.h", "libssa_amd.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\([^;{]*\)\s*;", text, re.M):
            if not m.group(0).startswith("typedef"):
                names.append(m.group(1))
    return names


def test_library_exports_every_declared_function():
    names = declared_functions()
    assert "sw_align" in names and "ssa_db_get_sequence" in names and "ssa_amd_last_stats" in names
    lib = C.CDLL(ssa.LIB_PATH)
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


@pytest.mark.parametrize("name", ["blosum45", "blosum50", "blosum62", "blosum80", "blosum90", "pam30", "pam70", "pam250"])
def test_builtin_matrices_equal_reference_tables(name):
    ssa.init_score_matrix(ssa.MATRIX_BUILDIN, name)
    assert ssa.get_matrix() == list(O.matrix_builtin(name).s)


def test_matrix_from_file_and_string():
    # tests/test_matrices.c:113-131
    ssa.init_score_matrix(ssa.READ_FROM_FILE, O.data("blosum90.txt"))
    m = ssa.get_matrix()
    assert m[33] == 5 and m[48] == -2
    text = open(O.data("matrices/blosum80.txt")).read()
    ssa.init_score_matrix(ssa.READ_FROM_STRING, text)
    assert ssa.get_matrix() == list(O.matrix_text(text).s)


def test_constant_scores_first_argument_is_match():
    # tests/test_matrices.c:123-135 via init_constant_scores (src/libssa.c:115-117)
    ssa.init_constant_scores(4, -2)
    m = ssa.get_matrix()
    assert [m[i] for i in (32, 33, 34, 64, 66)] == [-1, 4, -2, -1, 4]


@pytest.mark.parametrize("name,count", [("one_seq.fas", 1), ("AF091148.fas", 1403), ("test.fas", 5)])
def test_fasta_provider_counts_and_ids(name, count):
    # tests/test_libssa_extern_db.c:24-58
    lib = C.CDLL(ssa.LIB_PATH)  # its own function objects: the package's keep their types
    lib.ssa_db_get_sequence.restype = C.c_void_p
    lib.ssa_db_get_sequence.argtypes = [C.c_size_t]
    assert lib.ssa_db_init(O.data(name).encode()) == 0
    assert lib.ssa_db_get_sequence_count() == count
    for i in (0, count - 1):
        p = lib.ssa_db_get_sequence(i)
        assert p and C.cast(p, C.POINTER(C.c_size_t))[0] == i
    assert not lib.ssa_db_get_sequence(count)
    recs = O.read_fasta(O.data(name))
    p = lib.ssa_db_get_sequence(count - 1)
    seqlen = C.cast(p, C.POINTER(C.c_size_t))[1]
    seqptr = C.cast(p, C.POINTER(C.c_char_p))[2]
    assert seqlen == len(recs[-1][1]) and seqptr == recs[-1][1].encode()
    lib.ssa_db_close()


def run_child(code):
    return subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, cwd=ROOT, timeout=120)


def test_fatal_paths_exit_with_status_1():
    # src/libssa.c:196-219: NUC without constant scoring, missing query, bad modes
    r = run_child("import libssa_amd as s\n"
                  "s.init_score_matrix(s.MATRIX_BUILDIN, 'blosum62')\n"
                  "s.init_symbol_translation(s.NUCLEOTIDE, s.FORWARD_STRAND, 1, 1)\n"
                  "q = s.init_sequence_fasta(s.READ_FROM_STRING, 'ACGT')\n"
                  "s.sw_align(q, 5, 16, 0)\n")
    assert r.returncode == 1 and "Nucleotide sequences can only be aligned using constant scores." in r.stderr
    r = run_child("import libssa_amd as s\ns.init_score_matrix(7, 'x')\n")
    assert r.returncode == 1 and "Unknown mode for reading score matrices: 7" in r.stderr
    r = run_child("import libssa_amd as s\ns.init_score_matrix(s.MATRIX_BUILDIN, 'nope')\n")
    assert r.returncode == 1 and "Unknown matrix: nope" in r.stderr
    r = run_child("import libssa_amd as s\ns.init_symbol_translation(1, 1, 7, 1)\n")
    assert r.returncode == 1 and "Illegal database genetic code specified." in r.stderr
    r = run_child("import libssa_amd as s\ns.init_constant_scores(1, -1)\ns.sw_align(None, 5, 16, 0)\n")
    assert r.returncode == 1 and "Query not initialized." in r.stderr


def test_query_reading():
    # tests/test_query.c:210-235: header and mapped length; unknown symbols dropped
    ssa.reset_settings()
    ssa.init_symbol_translation(ssa.AMINOACID, ssa.FORWARD_STRAND, 1, 1)
    q = ssa.init_sequence_fasta(ssa.READ_FROM_FILE, O.data("NP_009305.1.fas"))
    assert q
    ssa.free_sequence(q)
    assert ssa.init_sequence_fasta(ssa.READ_FROM_FILE, "/nonexistent/file.fas") is None


def test_chunk_size_zero_falls_back_to_default():
    r = run_child("import libssa_amd as s\ns.set_chunk_size(0)\n")
    assert r.returncode == 0 and "Only non zero chunk sizes are allowed" in r.stdout
