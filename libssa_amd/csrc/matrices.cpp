// matrices.cpp -- substitution score matrices (reference: src/matrices.c).
//
// A matrix is 32 x 32 int64 indexed [db code][query code]; entries that the
// source does not define stay -1 (src/matrices.c:335-345).  The 8- and 16-bit
// tiers use the entries converted to their lane type (src/matrices.c:351-379),
// which the search derives from this table.
#include <cstdio>
#include <cstring>
#include <strings.h>

#include "common.h"

namespace ssa {

struct ScoreTable {
    const char* name;
    const char* rows;  // row symbols in table order
    const char* cols;  // column symbols in table order
    const int8_t* values;
};

#include "builtin_matrices.inc"

Matrix& matrix() {
    static Matrix m;
    return m;
}

static void reset(Matrix& m) {
    for (int i = 0; i < kDim * kDim; ++i) m.s[i] = -1;
    m.constant = false;
    m.valid = true;
}

void matrix_free() { matrix().valid = false; }

// src/matrices.c:447-460: codes 1..31 get match/mismatch, code 0 keeps -1
void matrix_init_constant(int match, int mismatch) {
    Matrix& m = matrix();
    reset(m);
    m.constant = true;
    for (int a = 1; a < kDim; ++a)
        for (int b = 1; b < kDim; ++b) m.s[a * kDim + b] = (a == b) ? match : mismatch;
}

// Line parser with the behaviour of read_line (src/matrices.c:388-437): '#'
// and empty lines are skipped; a line starting with a blank or tab lists the
// column symbols (the symbol count accumulates over such lines, as in the
// reference); any other line starts with its row symbol followed by one
// integer per column.  Symbols are mapped with the amino-acid code table.
struct TextParser {
    int order[2048];
    int symbols = 0;

    void line(const char* line) {
        const char* p = line;
        const char c = *p++;
        if (c == '\n' || c == '#') return;
        if (c == ' ' || c == '\t') {
            int k = 0;
            char x;
            while ((x = *p++)) {
                if (!strchr(" \t\n", x)) {
                    order[k++] = map_aa((unsigned char)x);
                    symbols++;
                }
            }
            return;
        }
        const int a = map_aa((unsigned char)c);
        for (int i = 0; i < symbols; ++i) {
            long sc;
            int used = 0;
            if (sscanf(p, "%ld%n", &sc, &used) != 1) fatal("Problem parsing score matrix file.");
            const int b = order[i];
            if (a >= 0 && b >= 0 && a < kDim && b < kDim) matrix().s[a * kDim + b] = sc;
            p += used;
        }
    }
};

// src/matrices.c:497-531
void matrix_init_text(const char* text) {
    if (!text) fatal("Cannot read score matrix string.");
    reset(matrix());
    TextParser tp;
    const char* s = text;
    char buf[2048];
    while (*s) {
        const char* nl = strchr(s, '\n');
        size_t len = nl ? (size_t)(nl - s) : strlen(s);
        if (len >= sizeof buf) fatal("Problem parsing score matrix file.");
        memcpy(buf, s, len);
        buf[len] = 0;
        tp.line(buf);
        s = nl ? nl + 1 : s + len;
    }
}

// src/matrices.c:466-495
void matrix_init_file(const char* path) {
    reset(matrix());
    FILE* fp = fopen(path, "r");
    if (!fp) fatal("Cannot open score matrix file.");
    TextParser tp;
    char line[2048];
    while (fgets(line, sizeof line, fp)) tp.line(line);
    fclose(fp);
}

// src/matrices.c:549-568
void matrix_init_builtin(const char* name) {
    for (const ScoreTable* t : k_builtin_tables) {
        if (strcasecmp(name, t->name) != 0) continue;
        Matrix& m = matrix();
        reset(m);
        const size_t nc = strlen(t->cols);
        for (size_t r = 0; t->rows[r]; ++r) {
            const int a = map_aa((unsigned char)t->rows[r]);
            for (size_t c = 0; c < nc; ++c) {
                const int b = map_aa((unsigned char)t->cols[c]);
                if (a >= 0 && b >= 0) m.s[a * kDim + b] = t->values[r * nc + c];
            }
        }
        return;
    }
    fatal("Unknown matrix: %s", name);
}

}  // namespace ssa
