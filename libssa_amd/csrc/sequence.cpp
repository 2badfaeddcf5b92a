// sequence.cpp -- symbol codes, reverse complement, translation, query reading.
//
// Codes follow the reference's tables (src/util/util_sequence.c:40-88):
// amino acids are numbered by their position in "-ABCDEFGHIKLMNPQRSTVWXYZU*OJ"
// ('-' itself is not an amino-acid symbol, '*' is); nucleotides by their
// position in "-ACMGRSVTWYHKDBN" with U read as T, which makes a nucleotide
// code the IUPAC bit set A=1, C=2, G=4, T=8.
#include <algorithm>
#include <cstdio>
#include <cstring>
#include <mutex>

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

// NCBI's genetic codes 1..23 as published (the ones the reference names,
// src/util/util_sequence.c:95-123): the 64 amino acids of the codons in TCAG
// order, first base slowest; nullptr where NCBI has no code of that number
static const char* const kGeneticCodes[23] = {
    "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 1 standard
    "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIMMTTTTNNKKSS**VVVVAAAADDEEGGGG",  // 2 vertebrate mitochondrial
    "FFLLSSSSYY**CCWWTTTTPPPPHHQQRRRRIIMMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 3 yeast mitochondrial
    "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 4 mold, protozoan, coelenterate mt.
    "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIMMTTTTNNKKSSSSVVVVAAAADDEEGGGG",  // 5 invertebrate mitochondrial
    "FFLLSSSSYYQQCC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 6 ciliate, dasycladacean, hexamita
    nullptr, nullptr,
    "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIIMTTTTNNNKSSSSVVVVAAAADDEEGGGG",  // 9 echinoderm, flatworm mt.
    "FFLLSSSSYY**CCCWLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 10 euplotid nuclear
    "FFLLSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 11 bacterial, archaeal, plastid
    "FFLLSSSSYY**CC*WLLLSPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 12 alternative yeast nuclear
    "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIMMTTTTNNKKSSGGVVVVAAAADDEEGGGG",  // 13 ascidian mitochondrial
    "FFLLSSSSYYY*CCWWLLLLPPPPHHQQRRRRIIIMTTTTNNNKSSSSVVVVAAAADDEEGGGG",  // 14 alternative flatworm mt.
    "FFLLSSSSYY*QCC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 15 blepharisma nuclear
    "FFLLSSSSYY*LCC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 16 chlorophycean mt.
    nullptr, nullptr, nullptr, nullptr,
    "FFLLSSSSYY**CCWWLLLLPPPPHHQQRRRRIIMMTTTTNNNKSSSSVVVVAAAADDEEGGGG",  // 21 trematode mitochondrial
    "FFLLSS*SYY*LCC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 22 scenedesmus obliquus mt.
    "FF*LSSSSYY**CC*WLLLLPPPPHHQQRRRRIIIMTTTTNNKKSSRRVVVVAAAADDEEGGGG",  // 23 thraustochytrium mt.
};

bool gencode_valid(int code) { return code >= 1 && code <= 23 && kGeneticCodes[code - 1] != nullptr; }

// An nt16 code is a set of bases (bit 0 A, 1 C, 2 G, 3 T).  A codon of base
// sets translates to the amino acid that every codon in it gives; a mix of
// only D and N is B, of only E and Q is Z; anything else, or no base, is X.
const int8_t* translation_table(int gencode) {
    static std::vector<int8_t> tables[23];
    static std::mutex mu;
    std::lock_guard<std::mutex> lk(mu);
    std::vector<int8_t>& t = tables[gencode - 1];
    if (!t.empty()) return t.data();
    const char* code = kGeneticCodes[gencode - 1];
    static const int kTcag[4] = {2, 1, 3, 0};  // base bit (A, C, G, T) -> TCAG position
    t.resize(4096);
    for (int a = 0; a < 16; ++a)
        for (int b = 0; b < 16; ++b)
            for (int c = 0; c < 16; ++c) {
                std::string set;
                for (int i = 0; i < 4; ++i)
                    for (int j = 0; j < 4; ++j)
                        for (int k = 0; k < 4; ++k)
                            if ((a >> i & 1) && (b >> j & 1) && (c >> k & 1)) {
                                const char x = code[16 * kTcag[i] + 4 * kTcag[j] + kTcag[k]];
                                if (set.find(x) == std::string::npos) set.push_back(x);
                            }
                std::sort(set.begin(), set.end());
                char aa = 'X';
                if (set.size() == 1) aa = set[0];
                else if (set == "DN") aa = 'B';
                else if (set == "EQ") aa = 'Z';
                t[256 * a + 16 * b + c] = (int8_t)map_aa(aa);
            }
    return t.data();
}

void translate(const int8_t* table, const uint8_t* dna, size_t len, int strand, int frame, std::vector<uint8_t>& out) {
    if (len < (size_t)frame)  // the reference's (len - frame) / 3 wraps and the allocation fails
        fatal("Unable to re-allocate enough memory: %ld.", (long)(((size_t)len - (size_t)frame) / 3 + 1));
    const size_t plen = (len - (size_t)frame) / 3;
    out.resize(plen);
    for (size_t p = 0; p < plen; ++p) {
        int c[3];
        for (int k = 0; k < 3; ++k) {
            const size_t i = (size_t)frame + 3 * p + (size_t)k;  // position on the strand read
            c[k] = strand == 0 ? dna[i] : complement_nt(dna[len - 1 - i]);
        }
        out[p] = (uint8_t)table[256 * c[0] + 16 * c[1] + c[2]];
    }
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

// fill_and_map_query (src/query.c:132-163): a nucleotide query keeps its
// strands; a translated one becomes the three frames of each strand selected,
// with the query genetic code in force when it is read.
static p_query make_query(const char* text, size_t len) {
    p_query q = new _query();
    const Config& cf = config();
    const int st = cf.symtype;
    const bool nuc = !(st == AMINOACID || st == TRANS_DB);
    std::vector<uint8_t> mapped = map_query_text(text, len, nuc);
    if (st == NUCLEOTIDE) {
        q->nt[0] = mapped;
        q->has_nt[0] = true;
        if (cf.strands & 2) {
            q->nt[1].resize(mapped.size());
            revcomp(mapped.data(), mapped.size(), q->nt[1].data());
            q->has_nt[1] = true;
        }
    } else if (query_is_translated(st)) {
        const int8_t* table = translation_table(cf.q_gencode);
        for (int s = 0; s < 2; ++s)
            if ((s + 1) & cf.strands)
                for (int f = 0; f < 3; ++f) {
                    translate(table, mapped.data(), mapped.size(), s, f, q->aa[3 * s + f]);
                    q->has_aa[3 * s + f] = true;
                }
    } else {
        q->aa[0] = mapped;
        q->has_aa[0] = true;
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
