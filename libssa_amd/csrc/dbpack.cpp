// dbpack.cpp -- writing and mapping pre-packed DB files (layout in dbpack.h).
//
// The reference maps every DB sequence to codes on every search
// (src/db_adapter.c:47-110, 212-239).  This library maps once per init_db; a
// packed file moves that step offline, so a process that opens one only reads
// the file and uploads the codes.
#include "dbpack.h"

#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace ssa {

namespace {

uint64_t align64(uint64_t x) { return (x + 63) & ~(uint64_t)63; }

bool section_fits(uint64_t off, uint64_t n, uint64_t size, uint64_t file_bytes) {
    if (off % 8 || off < sizeof(PackHeader) || off > file_bytes) return false;
    if (size && n > (file_bytes - off) / size) return false;
    return true;
}

bool write_all(FILE* fp, const void* p, size_t n) { return n == 0 || fwrite(p, 1, n, fp) == n; }

bool write_padded(FILE* fp, uint64_t& at, uint64_t off, const void* p, size_t n) {
    static const char zeros[64] = {0};
    while (at < off) {
        const size_t k = (size_t)std::min<uint64_t>(64, off - at);
        if (!write_all(fp, zeros, k)) return false;
        at += k;
    }
    if (!write_all(fp, p, n)) return false;
    at += n;
    return true;
}

}  // namespace

bool is_packed_file(const char* path) {
    if (!path) return false;
    FILE* fp = fopen(path, "rb");
    if (!fp) return false;
    char m[8];
    const bool yes = fread(m, 1, 8, fp) == 8 && memcmp(m, kPackMagic, 8) == 0;
    fclose(fp);
    return yes;
}

bool packed_open(const char* path, PackedView* v, void** map, size_t* map_bytes, const char** why) {
    *map = nullptr;
    *map_bytes = 0;
    const int fd = open(path, O_RDONLY);
    if (fd < 0) {
        *why = "cannot open";
        return false;
    }
    struct stat st;
    if (fstat(fd, &st) != 0 || (size_t)st.st_size < sizeof(PackHeader)) {
        close(fd);
        *why = "truncated header";
        return false;
    }
    const size_t bytes = (size_t)st.st_size;
    // private and writable: seqinfo_t.seq is a char* the caller may write through
    void* m = mmap(nullptr, bytes, PROT_READ | PROT_WRITE, MAP_PRIVATE, fd, 0);
    close(fd);
    if (m == MAP_FAILED) {
        *why = "cannot map";
        return false;
    }
    madvise(m, bytes, MADV_WILLNEED);
    const char* base = (const char*)m;
    const PackHeader& h = *(const PackHeader*)base;
    const char* bad = nullptr;
    if (memcmp(h.magic, kPackMagic, 8) != 0)
        bad = "bad magic";
    else if (h.version != kPackVersion || h.header_bytes != sizeof(PackHeader))
        bad = "unsupported version";
    else if (h.file_bytes != bytes)
        bad = "file size does not match the header";
    else if ((h.symtype != NUCLEOTIDE || h.strands != 3) && (h.symtype != AMINOACID || h.strands != 1))
        bad = "unknown symbol type";
    else if (h.seq_count >= UINT32_MAX || h.entry_count >= UINT32_MAX ||
             !section_fits(h.off_seqs, h.seq_count, sizeof(SeqRec), bytes) ||
             !section_fits(h.off_ascii, h.ascii_bytes, 1, bytes) ||
             !section_fits(h.off_entries, h.entry_count, sizeof(EntryRec), bytes) ||
             !section_fits(h.off_first, h.seq_count + 1, 4, bytes) ||
             !section_fits(h.off_bylen, h.entry_count, 4, bytes) ||
             !section_fits(h.off_codes, h.code_bytes, 1, bytes) ||
             !section_fits(h.off_unknown, h.unknown_count, sizeof(UnknownRec), bytes))
        bad = "section outside the file";
    if (!bad) {
        v->h = &h;
        v->seqs = (const SeqRec*)(base + h.off_seqs);
        v->ascii = base + h.off_ascii;
        v->entries = (const EntryRec*)(base + h.off_entries);
        v->first = (const uint32_t*)(base + h.off_first);
        v->by_length = (const uint32_t*)(base + h.off_bylen);
        v->codes = (const uint8_t*)(base + h.off_codes);
        v->unknown = (const UnknownRec*)(base + h.off_unknown);
        // every index and range the search follows
        for (uint64_t i = 0; i < h.seq_count && !bad; ++i) {
            const SeqRec& s = v->seqs[i];
            if (s.ascii_off >= h.ascii_bytes || s.len >= h.ascii_bytes - s.ascii_off) bad = "sequence outside the residues";
            else if (v->ascii[s.ascii_off + s.len] != 0) bad = "sequence not terminated";
            if (v->first[i] > v->first[i + 1]) bad = "entry table out of order";
        }
        if (!bad && (v->first[0] != 0 || v->first[h.seq_count] != h.entry_count)) bad = "entry table out of order";
        for (uint64_t i = 0; i < h.seq_count && !bad; ++i)
            for (uint32_t k = v->first[i]; k < v->first[i + 1] && !bad; ++k) {
                const EntryRec& e = v->entries[k];
                const SeqRec& s = v->seqs[i];
                if (e.len != s.len || e.len == 0 || e.id != s.id || e.strand > 1 ||
                    (e.strand == 1 && h.symtype != NUCLEOTIDE) || e.code_off < 0 || e.code_off % 4 ||
                    (uint64_t)e.code_off > h.code_bytes || ((uint64_t)e.len + 3) / 4 * 4 > h.code_bytes - e.code_off)
                    bad = "entry outside the codes";
            }
        // a permutation of the entries, longest first
        std::vector<bool> seen(bad ? 0 : h.entry_count, false);
        for (uint64_t k = 0; k < h.entry_count && !bad; ++k) {
            const uint32_t e = v->by_length[k];
            if (e >= h.entry_count || seen[e] || (k && v->entries[v->by_length[k - 1]].len < v->entries[e].len))
                bad = "length order is not a longest-first permutation";
            else
                seen[e] = true;
        }
        for (uint64_t u = 0; u < h.unknown_count && !bad; ++u)
            if (v->unknown[u].id >= h.seq_count) bad = "unknown-symbol record out of range";
        // every residue code below the alphabet's size (the kernels index
        // 32-code profiles with them and keep flag bits above a doubled
        // code), and the padding to whole 4-residue blocks zero
        const unsigned limit = h.symtype == NUCLEOTIDE ? 16u : 32u;
        for (uint64_t k = 0; k < h.entry_count && !bad; ++k) {
            const EntryRec& e = v->entries[k];
            const uint8_t* c = v->codes + e.code_off;
            const uint64_t padded = ((uint64_t)e.len + 3) / 4 * 4;
            for (uint64_t i = 0; i < e.len && !bad; ++i)
                if (c[i] >= limit) bad = "residue code out of range";
            for (uint64_t i = e.len; i < padded && !bad; ++i)
                if (c[i] != 0) bad = "non-zero padding after an entry's codes";
        }
    }
    if (bad) {
        munmap(m, bytes);
        *why = bad;
        return false;
    }
    *map = m;
    *map_bytes = bytes;
    return true;
}

void packed_close(void* map, size_t map_bytes) {
    if (map) munmap(map, map_bytes);
}

int write_packed_db(const char* path) {
    const Config& cf = config();
    if (!path) {
        print_error("Cannot write packed database: no file name");
        return -1;
    }
    if (db_is_translated(cf.symtype)) {
        print_error("Cannot write packed database: translated DB sequences are not packed");
        return -1;
    }
    // a translated query searches an amino-acid DB
    const bool nuc = cf.symtype == NUCLEOTIDE;
    HostDb d;
    map_provider_db(nuc ? NUCLEOTIDE : AMINOACID, nuc ? BOTH_STRANDS : FORWARD_STRAND, 0, d);
    std::vector<SeqRec> seqs(d.count);
    uint64_t ascii_bytes = 0;
    for (size_t i = 0; i < d.count; ++i) {
        p_seqinfo s = ssa_db_get_sequence(i);
        seqs[i] = SeqRec{(uint64_t)s->ID, ascii_bytes, (uint64_t)s->seqlen};
        ascii_bytes += s->seqlen + 1;  // NUL-terminated, like the FASTA provider's records
    }
    std::vector<EntryRec> ents(d.entries.size());
    for (size_t k = 0; k < ents.size(); ++k) {
        const DbEntry& e = d.entries[k];
        ents[k] = EntryRec{e.id, e.len, e.strand, e.frame, 0, e.code_off};
    }
    std::vector<UnknownRec> unk(d.unknown.size());
    for (size_t u = 0; u < unk.size(); ++u) unk[u] = UnknownRec{d.unknown[u].first, d.unknown[u].second};

    PackHeader h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, kPackMagic, 8);
    h.version = kPackVersion;
    h.header_bytes = sizeof(PackHeader);
    h.symtype = nuc ? NUCLEOTIDE : AMINOACID;
    h.strands = nuc ? 3 : 1;
    h.seq_count = d.count;
    h.entry_count = ents.size();
    h.ascii_bytes = ascii_bytes;
    h.code_bytes = d.codes.size();
    h.unknown_count = unk.size();
    h.off_seqs = align64(sizeof(PackHeader));
    h.off_ascii = align64(h.off_seqs + seqs.size() * sizeof(SeqRec));
    h.off_entries = align64(h.off_ascii + ascii_bytes);
    h.off_first = align64(h.off_entries + ents.size() * sizeof(EntryRec));
    h.off_bylen = align64(h.off_first + d.first_entry.size() * 4);
    h.off_codes = align64(h.off_bylen + d.by_length.size() * 4);
    h.off_unknown = align64(h.off_codes + d.codes.size());
    h.file_bytes = h.off_unknown + unk.size() * sizeof(UnknownRec);

    // written beside the target and renamed, so a reader never sees half a file
    const std::string tmp = std::string(path) + ".tmp";
    FILE* fp = fopen(tmp.c_str(), "wb");
    if (!fp) {
        print_error("Cannot write packed database: %s", path);
        return -1;
    }
    uint64_t at = 0;
    bool ok = write_padded(fp, at, 0, &h, sizeof h) &&
              write_padded(fp, at, h.off_seqs, seqs.data(), seqs.size() * sizeof(SeqRec));
    for (size_t i = 0; ok && i < d.count; ++i) {
        p_seqinfo s = ssa_db_get_sequence(i);
        ok = write_padded(fp, at, i == 0 ? h.off_ascii : at, s->seq, s->seqlen) &&
             write_padded(fp, at, at, "", 1);
    }
    if (ok && d.count == 0) ok = write_padded(fp, at, h.off_ascii, nullptr, 0);
    ok = ok && write_padded(fp, at, h.off_entries, ents.data(), ents.size() * sizeof(EntryRec)) &&
         write_padded(fp, at, h.off_first, d.first_entry.data(), d.first_entry.size() * 4) &&
         write_padded(fp, at, h.off_bylen, d.by_length.data(), d.by_length.size() * 4) &&
         write_padded(fp, at, h.off_codes, d.codes.data(), d.codes.size()) &&
         write_padded(fp, at, h.off_unknown, unk.data(), unk.size() * sizeof(UnknownRec)) && at == h.file_bytes;
    ok = (fclose(fp) == 0) && ok;
    if (!ok || rename(tmp.c_str(), path) != 0) {
        unlink(tmp.c_str());
        print_error("Cannot write packed database: %s", path);
        return -1;
    }
    return 0;
}

}  // namespace ssa
