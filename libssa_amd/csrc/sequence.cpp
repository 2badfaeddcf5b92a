// sequence.cpp -- symbol codes, reverse complement, query reading.
//
// Codes follow the reference's tables (src/util/util_sequence.c:40-88):
// amino acids are numbered by their position in "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"
// ('-' itself is not an amino-acid symbol, '*' is); nucleotides by their
// position in "-ACMGRSVTWYHKDBN" with U read as T, which makes a nucleotide
// code the IUPAC bit set A=1, C=2, G=4, T=8.
#include <cstdio>
#include <cstring>

#include "common.h"

namespace ssa {

int map_aa(int c) {
    static const char* sym = "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ";
    if (c == '*') return 25;
    if (c >= 'a' && c <= 'z') c -= 32;
    if (c < 'A' || c > 'Z') return -1;
    const char* f = strchr(sym + 1, c);
    return f ? (int)(f - sym) : -1;
}

int map_nt(int c) {
    static const char* sym = "-ACMGRSVTWYHKDBN";
    if (c == '-') return 0;
    if (c >= 'a' && c <= 'z') c -= 32;
    if (c == 'U') return 8;
    if (c < 'A' || c > 'Z') return -1;
    const char* f = strchr(sym + 1, c);
    return f ? (int)(f - sym) : -1;
}

// complement swaps the A/T and C/G bits (src/util/util_sequence.c:174)
int complement_nt(int code) {
    return ((code & 1) << 3) | ((code & 2) << 1) | ((code & 4) >> 1) | ((code & 8) >> 3);
}

// src/util/util_sequence.c:427-437
void revcomp(const uint8_t* in, size_t len, uint8_t* out) {
    for (size_t i = 0; i < len; ++i) out[i] = (uint8_t)complement_nt(in[len - 1 - i]);
}

// genetic codes 1..23 that exist (src/util/util_sequence.c:93-118)
bool gencode_valid(int code) {
    if (code < 1 || code > 23) return false;
    return !(code == 7 || code == 8 || (code >= 17 && code <= 20));
}

// map_sequence (src/query.c:102-130): unknown symbols are dropped; anything but
// line breaks and blanks is reported
static std::vector<uint8_t> map_query_text(const char* s, size_t len, bool nucleotide) {
    std::vector<uint8_t> out;
    out.reserve(len);
    std::string unknown;
    for (size_t i = 0; i < len; ++i) {
        const int c = (unsigned char)s[i];
        const int m = nucleotide ? map_nt(c) : map_aa(c);
        if (m >= 0) out.push_back((uint8_t)m);
        else if (c != '\n' && c != ' ' && c != '\t') unknown.push_back((char)c);
    }
    if (!unknown.empty())
        print_warning("%ld unknown symbols found and removed: '%s'", (long)unknown.size(), unknown.c_str());
    return out;
}

// fill_and_map_query (src/query.c:132-163) for the symbol types the search
// supports; the translated types are refused when a search starts.
static p_query make_query(const char* text, size_t len) {
    p_query q = new _query();
    const int st = config().symtype;
    const bool nuc = !(st == AMINOACID || st == TRANS_DB);
    std::vector<uint8_t> mapped = map_query_text(text, len, nuc);
    if (st == NUCLEOTIDE) {
        q->nt[0] = mapped;
        q->has_nt[0] = true;
        if (config().strands & 2) {
            q->nt[1].resize(mapped.size());
            revcomp(mapped.data(), mapped.size(), q->nt[1].data());
            q->has_nt[1] = true;
        }
    } else if (st == AMINOACID || st == TRANS_DB) {
        q->aa[0] = mapped;
        q->has_aa[0] = true;
    } else {
        q->nt[0] = mapped;  // kept untranslated
        q->has_nt[0] = true;
    }
    return q;
}

p_query query_from_string(const char* s) { return make_query(s, strlen(s)); }

// query_read_from_file (src/query.c:187-264): an optional '>' header line,
// then every following line up to the next '>' line.
p_query query_from_file(const char* path) {
    if (strcmp(path, "-") == 0) {
        print_error("Query not specified");
        return nullptr;
    }
    FILE* fp = fopen(path, "r");
    if (!fp) {
        print_error("Cannot open query file: %s", path);
        return nullptr;
    }
    char line[2048];
    line[0] = 0;
    if (!fgets(line, sizeof line, fp)) {
        print_error("Could not initialise from query sequence");
        fclose(fp);
        return nullptr;
    }
    std::string header;
    size_t len = strlen(line);
    if (len && line[len - 1] == '\n') line[--len] = 0;
    if (line[0] == '>') {
        header.assign(line + 1);
        line[0] = 0;
        if (!fgets(line, sizeof line, fp)) {
            print_error("Could not read first line from query sequence");
            fclose(fp);
            return nullptr;
        }
    }
    std::string seq;
    while (line[0] && line[0] != '>') {
        seq.append(line);
        if (!fgets(line, sizeof line, fp)) break;
    }
    fclose(fp);
    p_query q = make_query(seq.data(), seq.size());
    q->header = header;
    return q;
}

}  // namespace ssa
