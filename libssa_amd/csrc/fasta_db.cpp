// fasta_db.cpp -- default implementation of the DB plugin interface
// (include/libssa_extern_db.h; reference interface src/libssa_extern_db.h:28-56).
//
// The reference links an external library (libsdb) for this; this provider
// reads a FASTA file: records in file order get IDs 0, 1, ...; a record's
// sequence is its lines joined without line terminators, as a NUL-terminated
// string; NULL past the end.
// A pre-packed file (dbpack.h) is served from its memory map instead, with
// its stored IDs.  Programs that define these four functions themselves
// replace this provider.
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "common.h"
#include "dbpack.h"
#include "search.h"

namespace {

struct FastaDB {
    std::vector<char> residues;   // all sequences back to back, each NUL-terminated
    std::vector<size_t> offset;   // record i occupies [offset[i], offset[i+1] - 1)
    std::vector<seqinfo_t> infos;
    bool open = false;
    // a packed file: infos point into the map
    void* map = nullptr;
    size_t map_bytes = 0;
    ssa::PackedView view;
};

FastaDB& db() {
    static FastaDB d;
    return d;
}

std::mutex g_mu;

bool load(const char* path, FastaDB& d) {
    FILE* fp = fopen(path, "rb");
    if (!fp) return false;
    std::vector<char> buf(1 << 20);
    bool in_header = false, at_line_start = true, have_record = false;
    size_t n;
    while ((n = fread(buf.data(), 1, buf.size(), fp)) > 0) {
        for (size_t k = 0; k < n; ++k) {
            const char c = buf[k];
            if (at_line_start && c == '>') {
                if (have_record) d.residues.push_back('\0');
                d.offset.push_back(d.residues.size());
                have_record = true;
                in_header = true;
            } else if (c == '\n') {
                in_header = false;
            } else if (!in_header && have_record && c != '\r') {
                d.residues.push_back(c);
            }
            at_line_start = (c == '\n');
        }
    }
    fclose(fp);
    if (have_record) d.residues.push_back('\0');
    d.offset.push_back(d.residues.size());
    const size_t count = d.offset.size() - 1;
    d.infos.resize(count);
    for (size_t i = 0; i < count; ++i) {
        d.infos[i].ID = i;
        d.infos[i].seqlen = d.offset[i + 1] - d.offset[i] - 1;
        d.infos[i].seq = d.residues.data() + d.offset[i];
    }
    return true;
}

}  // namespace

extern "C" {

int ssa_db_init(const char* db_name) {
    std::lock_guard<std::mutex> lk(g_mu);
    ssa::db_generation_bump();
    FastaDB& d = db();
    ssa::packed_close(d.map, d.map_bytes);
    d = FastaDB();
    if (db_name && ssa::is_packed_file(db_name)) {
        const char* why = "";
        if (!ssa::packed_open(db_name, &d.view, &d.map, &d.map_bytes, &why)) {
            d = FastaDB();
            d.open = true;
            ssa::print_error("Invalid packed database file: %s (%s)", db_name, why);
            return -1;
        }
        const ssa::PackedView& v = d.view;
        d.infos.resize(v.h->seq_count);
        for (size_t i = 0; i < d.infos.size(); ++i) {
            d.infos[i].ID = v.seqs[i].id;
            d.infos[i].seqlen = v.seqs[i].len;
            d.infos[i].seq = (char*)v.ascii + v.seqs[i].ascii_off;
        }
        d.open = true;
        return 0;
    }
    if (!db_name || !load(db_name, d)) {
        d = FastaDB();
        d.open = true;
        ssa::print_error("Cannot open database file: %s", db_name ? db_name : "(null)");
        return -1;
    }
    d.open = true;
    return 0;
}

size_t ssa_db_get_sequence_count(void) { return db().infos.size(); }

p_seqinfo ssa_db_get_sequence(size_t id) {
    FastaDB& d = db();
    if (id >= d.infos.size()) return nullptr;
    return &d.infos[id];
}

void ssa_db_close(void) {
    std::lock_guard<std::mutex> lk(g_mu);
    ssa::db_generation_bump();
    ssa::packed_close(db().map, db().map_bytes);
    db() = FastaDB();
}

}  // extern "C"

namespace ssa {

const PackedView* provider_packed_view() {
    FastaDB& d = db();
    return d.map ? &d.view : nullptr;
}

const seqinfo_t* provider_first_info() {
    FastaDB& d = db();
    return d.infos.empty() ? nullptr : &d.infos[0];
}

}  // namespace ssa
