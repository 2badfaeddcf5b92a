// search.cpp -- one database search on the GPU, with the reference's results.
//
// The reference (src/algo/manager.c:120-185 with one thread) walks the DB in
// chunks; per chunk it runs the 8-bit (or 16-bit) lane kernel for each query,
// re-runs overflowed entries at 16 bit for every query, then at 64 bit, and
// inserts each accepted score into a bounded min-heap whose tie handling
// depends on the insertion order.  Here every (query, entry) pair of a tier is
// scored at once on the GPU; the kernels also report the block at which each
// pair would have been flagged as overflowed.  From those results the host
// replays the reference's lane schedule (which lane retires which entry in
// which iteration) and feeds the heap in exactly the reference's order.
#include "search.h"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <numeric>
#include <functional>
#include <list>
#include <memory>
#include <mutex>
#include <queue>
#include <random>
#include <thread>

#include "dbpack.h"
#include "gpu.h"
#include "kernels.h"

namespace ssa {

// ============================================================================
// device-resident DB
// ============================================================================
namespace {

uint64_t g_generation = 1;
uint64_t g_build = 0;  // incremented at every rebuild of the device DB

struct DeviceDB {
    uint64_t generation = 0;  // the provider DB it was built from (init_db bumps it)
    uint64_t build = 0;       // this build: changes whenever the entries are rebuilt, for
                              // a new DB or new symbol type / strands / genetic code, so
                              // every cache derived from the entries is keyed on it
    int symtype = -1, strands = -1, gencode = -1;
    size_t count = 0;
    std::vector<DbEntry> entries;
    std::vector<uint32_t> first_entry;  // per DB ID (count + 1 values)
    uint8_t* d_codes = nullptr;         // all entries' codes, each padded to a multiple of 4
    size_t d_bytes = 0;
    std::vector<uint32_t> by_length;    // all entries, longest first (stable)
    bool packed = false;                // loaded from a packed file
    bool acgt = false;                  // nucleotide entries of A, C, G and T only (the query
                                        // profile may sit in VGPRs)
};

DeviceDB g_db;

}  // namespace

void db_generation_bump() { ++g_generation; }

void release_layouts();

void release_device_db() {
    release_layouts();
    if (g_db.d_codes) (void)hipFree(g_db.d_codes);
    g_db = DeviceDB();
}

static size_t map_db_codes(const char* s, size_t n, bool nucleotide, uint8_t* out) {
    size_t unknown = 0;
    for (size_t i = 0; i < n; ++i) {
        const int m = nucleotide ? map_nt((unsigned char)s[i]) : map_aa((unsigned char)s[i]);
        if (m < 0) {
            out[i] = 0;
            ++unknown;
        } else {
            out[i] = (uint8_t)m;
        }
    }
    return unknown;
}

// True when every residue of every entry is A, C, G or T (codes 1, 2, 4, 8;
// the padding after an entry is not looked at)
static bool entries_acgt(const uint8_t* codes, const std::vector<DbEntry>& entries) {
    for (const DbEntry& e : entries) {
        const uint8_t* c = codes + e.code_off;
        for (uint32_t i = 0; i < e.len; ++i)
            if (c[i] != 1 && c[i] != 2 && c[i] != 4 && c[i] != 8) return false;
    }
    return true;
}

static void warn_unknown(size_t unknown) {
    if (unknown) print_warning("%ld unknown symbols found and set to zero", (long)unknown);
}

// codes of an entry of len residues in the DB code buffer: a multiple of 4,
// and 4 for an empty entry (a translated frame without a codon), which the
// score kernels read as one block of padding
static size_t entry_code_bytes(size_t len) { return len ? (len + 3) & ~(size_t)3 : 4; }

// set_translated_sequences for TRANS_DB / TRANS_BOTH (src/db_adapter.c:72-104):
// frames 0..2 of both strands, strand field 0 / 1; or of the one strand
// selected, whose strand field is then the strand setting itself (1 or 2).
static void map_translated_db(int strands, int gencode, size_t count, HostDb& d) {
    const int8_t* table = translation_table(gencode);
    const int per_seq = strands == BOTH_STRANDS ? 6 : 3;
    std::vector<uint8_t> nt, prot;
    d.first_entry.assign(count + 1, 0);
    d.codes.clear();
    for (size_t id = 0; id < count; ++id) {
        d.first_entry[id] = (uint32_t)d.entries.size();
        p_seqinfo s = ssa_db_get_sequence(id);
        if (s->seqlen == 0) continue;  // src/db_adapter.c:231-233
        nt.resize(s->seqlen);
        const size_t unknown = map_db_codes(s->seq, s->seqlen, true, nt.data());
        if (unknown) d.unknown.emplace_back(id, unknown);
        for (int k = 0; k < per_seq; ++k) {
            const int strand = per_seq == 6 ? k / 3 : strands - 1, frame = k % 3;
            translate(table, nt.data(), nt.size(), strand, frame, prot);
            DbEntry e;
            e.id = s->ID;
            e.len = (uint32_t)prot.size();
            e.strand = (uint8_t)(per_seq == 6 ? strand : strands);
            e.frame = (uint8_t)frame;
            e.code_off = (int64_t)d.codes.size();
            d.codes.insert(d.codes.end(), prot.begin(), prot.end());
            d.codes.resize(e.code_off + entry_code_bytes(prot.size()), 0);
            d.entries.push_back(e);
        }
    }
    d.first_entry[count] = (uint32_t)d.entries.size();
    d.codes.resize(d.codes.size() + 4, 0);
}

// one entry per sequence (amino acids), or the forward strand and, with
// both, the reverse complement (nucleotides; src/db_adapter.c:47-70)
static void map_plain_db(bool nuc, bool both, size_t count, HostDb& d) {
    size_t total = 0;
    for (size_t id = 0; id < count; ++id) total += ((ssa_db_get_sequence(id)->seqlen + 3) & ~(size_t)3) * (both ? 2 : 1);
    d.first_entry.assign(count + 1, 0);
    d.codes.assign(total + 4, 0);
    size_t off = 0;
    for (size_t id = 0; id < count; ++id) {
        d.first_entry[id] = (uint32_t)d.entries.size();
        p_seqinfo s = ssa_db_get_sequence(id);
        if (s->seqlen == 0) continue;  // src/db_adapter.c:231-233
        const size_t len = s->seqlen, padded = (len + 3) & ~(size_t)3;
        DbEntry e;
        e.id = s->ID;
        e.len = (uint32_t)len;
        e.strand = 0;
        e.frame = 0;
        e.code_off = (int64_t)off;
        const size_t unknown = map_db_codes(s->seq, len, nuc, d.codes.data() + off);
        if (unknown) d.unknown.emplace_back(id, unknown);
        d.entries.push_back(e);
        off += padded;
        if (both) {
            DbEntry r = e;
            r.strand = 1;
            r.code_off = (int64_t)off;
            revcomp(d.codes.data() + e.code_off, len, d.codes.data() + off);
            d.entries.push_back(r);
            off += padded;
        }
    }
    d.first_entry[count] = (uint32_t)d.entries.size();
}

void map_provider_db(int symtype, int strands, int gencode, HostDb& d) {
    d = HostDb();
    size_t count = ssa_db_get_sequence_count();
    for (size_t id = 0; id < count; ++id)
        if (!ssa_db_get_sequence(id)) {
            count = id;
            break;
        }
    d.count = count;
    if (db_is_translated(symtype))
        map_translated_db(strands, gencode, count, d);
    else
        map_plain_db(symtype == NUCLEOTIDE, symtype == NUCLEOTIDE && (strands & 2), count, d);
    d.by_length.resize(d.entries.size());
    std::iota(d.by_length.begin(), d.by_length.end(), 0u);
    std::stable_sort(d.by_length.begin(), d.by_length.end(),
                     [&](uint32_t a, uint32_t b) { return d.entries[a].len > d.entries[b].len; });
}

// Takes the entries of a packed file: all of them, or for a nucleotide search
// of the forward strand only those of strand 0 (the packed file holds both).
static void load_packed(const PackedView& pv, bool keep_reverse, DeviceDB& d) {
    const PackHeader& h = *pv.h;
    const size_t count = h.seq_count;
    std::vector<uint32_t> renum(h.entry_count, UINT32_MAX);
    d.entries.clear();
    d.entries.reserve(keep_reverse ? h.entry_count : (h.entry_count + 1) / 2);
    d.first_entry.assign(count + 1, 0);
    size_t seq = 0;
    for (uint64_t k = 0; k < h.entry_count; ++k) {
        while (seq < count && pv.first[seq + 1] <= k) d.first_entry[++seq] = (uint32_t)d.entries.size();
        const EntryRec& r = pv.entries[k];
        if (r.strand == 1 && !keep_reverse) continue;
        renum[k] = (uint32_t)d.entries.size();
        DbEntry e;
        e.id = r.id;
        e.len = r.len;
        e.strand = r.strand;
        e.frame = r.frame;
        e.code_off = r.code_off;
        d.entries.push_back(e);
    }
    while (seq < count) d.first_entry[++seq] = (uint32_t)d.entries.size();
    d.by_length.clear();
    d.by_length.reserve(d.entries.size());
    for (uint64_t k = 0; k < h.entry_count; ++k)
        if (renum[pv.by_length[k]] != UINT32_MAX) d.by_length.push_back(renum[pv.by_length[k]]);
    for (uint64_t u = 0; u < h.unknown_count; ++u) warn_unknown(pv.unknown[u].count);
}

// Reads the plugin DB once, maps it (src/db_adapter.c:47-110) and keeps the
// codes resident on the GPU until the DB or the symbol settings change.  A
// packed file opened by the built-in provider is uploaded as it is.  With
// may_build false (the no-allocation path) a DB that is not resident for the
// current settings is an error instead.
static DeviceDB& device_db(bool may_build = true) {
    const Config& cf = config();
    size_t count = ssa_db_get_sequence_count();
    // the DB entries depend on the symbol type of the DB sequences, the
    // strands (nucleotide and translated DBs) and the DB genetic code
    // (translated DBs); a translated query searches an amino-acid DB
    const int db_type = cf.symtype == TRANS_QUERY ? AMINOACID : cf.symtype;
    const int strands = db_type == AMINOACID ? 0 : cf.strands;
    const int gencode = db_is_translated(db_type) ? cf.d_gencode : 0;
    if (g_db.generation == g_generation && g_db.symtype == db_type && g_db.strands == strands &&
        g_db.gencode == gencode && g_db.count == count && g_db.d_codes)
        return g_db;
    if (!may_build)
        throw FixedRegionError{-3, "the DB is not resident on the GPU with the current settings; run sw_align or "
                                   "nw_align once first"};
    release_device_db();
    DeviceDB& d = g_db;
    d.generation = g_generation;
    d.build = ++g_build;
    d.symtype = db_type;
    d.strands = strands;
    d.gencode = gencode;
    const bool nuc = db_type == NUCLEOTIDE;
    const bool both = nuc && (cf.strands & 2);
    const PackedView* pv = provider_packed_view();
    // the built-in provider's packed file, mapped for this symbol type, and
    // not hidden by a caller's own provider
    if (pv && pv->h->symtype == db_type && pv->h->seq_count == count &&
        (count == 0 || ssa_db_get_sequence(0) == provider_first_info())) {
        d.count = count;
        load_packed(*pv, both, d);
        d.d_bytes = pv->h->code_bytes + 4;
        SSA_HIP(hipMalloc(&d.d_codes, d.d_bytes));
        SSA_HIP(hipMemset(d.d_codes + pv->h->code_bytes, 0, 4));
        SSA_HIP(hipMemcpy(d.d_codes, pv->codes, pv->h->code_bytes, hipMemcpyHostToDevice));
        d.packed = true;
        d.acgt = nuc && entries_acgt(pv->codes, d.entries);
        return d;
    }
    HostDb hd;
    map_provider_db(db_type, cf.strands, cf.d_gencode, hd);
    for (auto& u : hd.unknown) warn_unknown(u.second);
    d.count = hd.count;
    d.entries.swap(hd.entries);
    d.first_entry.swap(hd.first_entry);
    d.by_length.swap(hd.by_length);
    d.d_bytes = hd.codes.size();
    SSA_HIP(hipMalloc(&d.d_codes, d.d_bytes));
    SSA_HIP(hipMemcpy(d.d_codes, hd.codes.data(), d.d_bytes, hipMemcpyHostToDevice));
    d.acgt = nuc && entries_acgt(hd.codes.data(), d.entries);
    return d;
}

// ============================================================================
// tier constants
// ============================================================================
namespace {

struct Tier {
    int bits;   // 8 or 16
    int shift;  // 8 for the 8-bit tier (values held shifted in int16), 0 for 16
    int64_t imin() const { return bits == 8 ? -128 : -32768; }
    int64_t imax() const { return bits == 8 ? 127 : 32767; }
    int64_t wrap(int64_t v) const { return bits == 8 ? (int64_t)(int8_t)v : (int64_t)(int16_t)v; }
    int64_t sat(int64_t v) const { return v < imin() ? imin() : (v > imax() ? imax() : v); }
    int32_t repr(int64_t v) const { return (int32_t)(v * (1 << shift)); }
    int64_t decode(int32_t r) const { return (int64_t)(r >> shift); }
};

const Tier kTier8{8, 8}, kTier16{16, 0};

uint32_t pack2(int32_t v) { return ((uint32_t)(uint16_t)v << 16) | (uint32_t)(uint16_t)v; }

// IEEE binary16 bits of an integer that binary16 holds exactly (at most 11
// significant bits, magnitude below 2^16)
int32_t f16_bits(int64_t v) {
    if (v == 0) return 0;
    const uint32_t sign = v < 0 ? 0x8000u : 0u;
    const uint64_t a = (uint64_t)(v < 0 ? -v : v);
    const int e = 63 - __builtin_clzll(a);  // a = 1.m * 2^e
    if (e > 15 || (e > 10 && (a & ((1ull << (e - 10)) - 1)) != 0)) fatal("libssa_amd: %lld is not an fp16 integer", (long long)v);
    const uint32_t mant = (uint32_t)(e <= 10 ? (a << (10 - e)) : (a >> (e - 10))) & 0x3ffu;
    return (int32_t)(sign | ((uint32_t)(e + 15) << 10) | mant);
}

// f16_bits of a 16-bit value through a process-wide table (each value
// converted on first use): a profile holds few distinct values but up to
// millions of entries.
int16_t f16_of_int16(int16_t v) {
    static std::vector<int32_t> table(65536, -1);
    static std::mutex mu;
    const uint16_t k = (uint16_t)v;
    int32_t b = __atomic_load_n(&table[k], __ATOMIC_ACQUIRE);
    if (b < 0) {
        std::lock_guard<std::mutex> lk(mu);
        b = f16_bits(v);
        __atomic_store_n(&table[k], b, __ATOMIC_RELEASE);
    }
    return (int16_t)b;
}

int64_t now_us() {
    return std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct MatrixInfo {
    bool fits8 = true, fits16 = true, fits32 = true;
    int64_t maxv = INT64_MIN, minv = INT64_MAX;
};

MatrixInfo matrix_info() {
    MatrixInfo mi;
    const Matrix& m = matrix();
    for (int i = 0; i < kDim * kDim; ++i) {
        const int64_t v = m.s[i];
        mi.maxv = std::max(mi.maxv, v);
        mi.minv = std::min(mi.minv, v);
        if (v < -128 || v > 127) mi.fits8 = false;
        if (v < -32768 || v > 32767) mi.fits16 = false;
        if (v < INT32_MIN / 4 || v > INT32_MAX / 4) mi.fits32 = false;
    }
    return mi;
}

// Mode 3 of the score kernel (score_kernel.h) keeps each lane's values in an
// fp16 window around an offset of its own, re-centred (maximum near 1024)
// once the window's maximum passes 1536 and every 128 columns while the
// offset is above 0.  True when the matrix and gaps keep every value a lane
// of `rows` query rows holds within +-2047 of its offset: a cell differs
// from its neighbours by at most max score + |gap open + extend| (step), so
// a lane's values lie within (rows - 1) steps (E and F up to |open + extend|
// more) below its maximum, a block of four columns raises the maximum by at
// most four steps (one more for the values handed down), and 128 columns
// lower it by at most |open + extend| + 127 |extend|.
bool f16_window_ok(int rows, int64_t qr, int64_t r) {
    const MatrixInfo mi = matrix_info();
    if (mi.maxv > 1024 || mi.minv < -1024) return false;
    const int64_t smax = std::max<int64_t>(0, mi.maxv);
    const int64_t aqr = qr < 0 ? -qr : qr, ar = r < 0 ? -r : r;
    const int64_t step = smax + aqr;
    const int64_t spread = (int64_t)(rows - 1) * step + aqr;
    const int64_t grow = 5 * (step + ar);  // the kernel's values of column j less j gap extend
    const int64_t fall = aqr + 127 * ar;
    return 1536 + grow <= 2047 && 1024 - spread - fall >= -2047;
}

// True when no H, E or F value of an alignment of a query of qlen residues
// with a DB entry of at most dmax residues can reach either end of the int16
// range (with room for one more gap step): the 16-bit tier is then exact and
// no overflow predicate can fire.
bool int16_exact(size_t qlen, int64_t dmax, const MatrixInfo& mi) {
    const Config& cf = config();
    const int64_t gap = std::abs((int64_t)cf.gap_open) + std::abs((int64_t)cf.gap_extend);
    const int64_t ml = std::min<int64_t>((int64_t)qlen, dmax);
    const int64_t span = 3 * gap + gap * ((int64_t)qlen + dmax + 4);
    const int64_t up = std::max<int64_t>(mi.maxv, 0) * ml;
    const int64_t lowest = -span - std::max<int64_t>(-mi.minv, 0) * ml;
    return lowest > -32768 + gap + 1 && up < 32767;
}

// True when every value the NW fp16 kernel (mode 4) holds for a query of
// qlen residues and a DB entry of at most dmax residues (padding columns
// included) lies within +-2047, so that fp16 holds it exactly.  That kernel
// holds each H, E and F of cell (i, j) less (i + j) * gap extend (g <= 0):
// the extension adds vanish from the recurrence (E and F take H + gap open,
// the diagonal takes the score - 2 g, the borders are constants).  A shifted
// H is at least the path of two gaps to its cell, 2 gap open + (i + j) g,
// less the shift: 2 gap open; E, F and the gap candidates one gap open
// below that, the diagonal candidate the lowest score below it; and at most
// max score * min(i, j) - (i + j) g.
bool nw_f16_exact(size_t qlen, int64_t dmax, const MatrixInfo& mi) {
    const Config& cf = config();
    if (cf.gap_open > 0 || cf.gap_extend > 0) return false;  // the bounds below take penalties <= 0
    const int64_t go = -(int64_t)cf.gap_open, ge = -(int64_t)cf.gap_extend;
    const int64_t smax = std::max<int64_t>(mi.maxv, 0), sneg = std::max<int64_t>(-mi.minv, 0);
    const int64_t lowest = -(3 * go) - sneg;
    const int64_t up = smax * std::min<int64_t>((int64_t)qlen, dmax) + ((int64_t)qlen + dmax) * ge;
    return mi.maxv <= 1024 && mi.minv >= -1024 && smax + 2 * ge <= 2047 && lowest >= -2047 && up <= 2047;
}

// result of one pair at one tier, decoded: d > 0 is a heap insert of score
// after d blocks in a reference lane, d < 0 an overflow after -d blocks
// (blocks >= 1 either way)
struct TierOut {
    int32_t d;
    int32_t score;
    bool accept() const { return d > 0; }
    int32_t dur() const { return d > 0 ? d : -d; }
};

}  // namespace

// ============================================================================
// GPU score tier
// ============================================================================
namespace {

struct Shape {
    int lanes;      // G: lanes per stream pair
    int rows;       // R: query rows per lane
    int cols = 4;   // columns per step (the column tile): 4 or 8
    int place = 1;  // query profile: 1 LDS, 2 VGPRs (nucleotide DBs of A, C, G, T)
};

}  // namespace

Shape& shape_override_ref();
int& cols_override_ref();
int& arith_override_ref();
int& placement_override_ref();
bool& task_trace_ref();
void release_layouts();

namespace {

Shape shape_override() { return shape_override_ref(); }

int& target_waves() {
    static int w = 10240;
    return w;
}

// Kernel shape from a cost model in wave-step units (one wave-step = one
// block of 4 columns on every lane of a wave, R rows per lane).  The chip
// works through passes * sum(wave blocks + G - 1) wave-steps on 1024 SIMDs;
// the wave holding the longest stream needs passes * (longest + G - 1) steps
// on its own and shares its SIMD for part of that time (factor 2).  A
// wave-step costs R rows times a per-row factor that charges the fixed
// per-step work (DPP moves, addressing, pair bookkeeping) to short lanes.
// shape measurements (ssa_amd_autotune): per search type and tuned query
// length, the time per query of every shape
struct TunedShape {
    int type;
    size_t qlen;
    Shape sh;
    double ms;
};

std::vector<TunedShape>& tuned_shapes() {
    static std::vector<TunedShape> t;
    return t;
}

std::mutex& tuned_mutex() {
    static std::mutex m;
    return m;
}

// every shape of the score kernel for a search type (0 SW, 1 NW)
const std::vector<Shape>& kernel_shapes(int type) {
    static const std::vector<Shape> sw = {{16, 32}, {8, 32}, {32, 16}, {16, 16}, {8, 16}, {64, 8}, {32, 8},
                                          {16, 8},  {8, 8},  {64, 4},  {32, 4},  {16, 4}, {8, 4}};
    static const std::vector<Shape> nw = {{32, 16}, {16, 16}, {8, 16}, {64, 8}, {32, 8}, {16, 8},
                                          {8, 8},   {64, 4},  {32, 4}, {16, 4}, {8, 4}};
    return type == 0 ? sw : nw;
}

// true when the whole-DB stream layout of this lane width (paired or not) is
// cached for the current device DB build
bool layout_cached(int lanes, bool paired);

// pair: packed SW queries of several buffers, which run paired (without the
// v_perm per cell pair) on the shapes with a paired kernel; only those are
// considered.  fixed (the no-allocation path): only shapes whose whole-DB
// layout is cached, since no layout may be built -- the unpaired one, or for
// packed queries the paired one of a shape with a paired kernel (packed
// queries also run unpaired, so there any cached shape will do).
Shape choose_shape_base(size_t qlen_max, int64_t total_blocks, int64_t max_blocks, int type, bool pair, bool fixed) {
    const Shape ov = shape_override();
    // a forced shape of 32 rows per lane runs NW with 16 (the NW kernel has no 32-row shape)
    if (ov.lanes && ov.rows) return (type == 1 && ov.rows == 32) ? Shape{ov.lanes, 16} : ov;
    auto allowed = [&](const Shape& sh) {
        return !fixed || layout_cached(sh.lanes, false) ||
               (pair && score_pair_supported(sh.lanes, sh.rows) && layout_cached(sh.lanes, true));
    };
    if (fixed) {
        bool any = false;
        for (const Shape& sh : kernel_shapes(type == 0 ? 0 : 1)) any |= allowed(sh);
        if (!any)
            throw FixedRegionError{-3, "no stream layout of the DB is resident; run sw_align or nw_align once first"};
        // packed queries may run unpaired here: the paired filter below
        // gives way to what is cached
        if (pair) {
            bool any_pair = false;
            for (const Shape& sh : kernel_shapes(0))
                any_pair |= score_pair_supported(sh.lanes, sh.rows) && allowed(sh);
            if (!any_pair) pair = false;
        }
    }
    if (type >= 0) {
        // measured: each shape's time at the nearest tuned length (log scale),
        // scaled by the passes the shape needs for this length
        std::lock_guard<std::mutex> lk(tuned_mutex());
        size_t near = 0;
        double bd = 1e300;
        for (const TunedShape& t : tuned_shapes()) {
            if (t.type != type) continue;
            const double d = std::fabs(std::log((double)std::max<size_t>(qlen_max, 1)) - std::log((double)t.qlen));
            if (d < bd) {
                bd = d;
                near = t.qlen;
            }
        }
        if (near) {
            Shape best{0, 0};
            double bt = 1e300;
            for (const TunedShape& t : tuned_shapes()) {
                if (t.type != type || t.qlen != near || !allowed(t.sh)) continue;
                if (pair && !score_pair_supported(t.sh.lanes, t.sh.rows)) continue;
                const size_t pr = (size_t)t.sh.lanes * t.sh.rows;
                const double p_q = (double)std::max<size_t>(1, (qlen_max + pr - 1) / pr);
                const double p_t = (double)std::max<size_t>(1, (t.qlen + pr - 1) / pr);
                const double est = t.ms * p_q / p_t;
                if (est < bt) {
                    bt = est;
                    best = t.sh;
                }
            }
            if (best.lanes) return best;
        }
    }
    Shape best = pair ? Shape{8, 32} : kernel_shapes(type == 0 ? 0 : 1)[0];
    double best_cost = 1e300;
    for (const Shape& sh : kernel_shapes(type == 0 ? 0 : 1)) {
        if (pair && !score_pair_supported(sh.lanes, sh.rows)) continue;
        if (!allowed(sh)) continue;
        const int64_t pr = (int64_t)sh.lanes * sh.rows;
        const int64_t passes = std::max<int64_t>(1, ((int64_t)qlen_max + pr - 1) / pr);
        // measured per-row costs: short lanes pay the fixed per-step work (32
        // rows run at two waves per SIMD, 16 at up to three); G = 8 and 32 fix
        // up their top lanes after the DPP shift
        double per_row = sh.rows == 4 ? 1.32 : (sh.rows == 8 ? 1.10 : (sh.rows == 16 ? 1.0 : 0.955));
        if (sh.lanes == 8 || sh.lanes == 32) per_row *= 1.04;
        // 512-row passes double-buffer 64 KiB of profile: two workgroups per
        // CU, which costs the 16-row shape its third wave per SIMD
        if (sh.lanes * sh.rows == 512 && sh.rows == 16) per_row *= 1.06;
        // paired: no v_perm (one of the ten packed ops of a SW cell pair)
        if (pair) per_row *= 0.9;
        const double waves = std::max<double>(target_waves(), 1.0);
        // the passes of a task run as one pipeline: G - 1 fill steps per task
        const double work =
            ((double)passes * (double)total_blocks * sh.lanes / 128.0 + waves * (sh.lanes - 1)) / 1024.0;
        const double tail = 2.0 * ((double)passes * (double)max_blocks + sh.lanes - 1);
        const double cost = std::max(work, tail) * sh.rows * per_row;
        if (cost < best_cost) {
            best = sh;
            best_cost = cost;
        }
    }
    return best;
}

// The shape with the column tile forced by ssa_amd_set_kernel_variant, where
// the shape has it.
Shape choose_shape(size_t qlen_max, int64_t total_blocks, int64_t max_blocks, int type, bool pair,
                   bool fixed = false) {
    Shape sh = choose_shape_base(qlen_max, total_blocks, max_blocks, type, pair, fixed);
    const int cols = cols_override_ref();
    if (cols && score_shape_supported(sh.lanes, sh.rows, type == 0 ? 0 : 1, cols)) sh.cols = cols;
    return sh;
}

struct PairSet {
    // pairs (q, e) for all queries q and the entries listed (DB entry
    // indices): the set's own list, or a shared one (ext)
    std::vector<uint32_t> entries;
    const std::vector<uint32_t>* ext = nullptr;
    size_t nq = 0;
    const std::vector<uint32_t>& list() const { return ext ? *ext : entries; }
    size_t index(size_t q, size_t k) const { return q * list().size() + k; }
    size_t size() const { return nq * list().size(); }
};

// What a search needs of the loaded DB beyond its entries, for one chunk
// size: the chunks' entry ranges (whole chunks of DB IDs until the first
// chunk without entries, as search_8/16/64 stop there,
// src/algo/8/search_8.c:135-143), the entries they hold (0 .. N-1) and their
// totals.  Computed once per device DB build and chunk size.
struct DbChunks {
    uint64_t build = ~0ull;
    size_t chunk_size = 0;
    std::vector<std::pair<uint32_t, uint32_t>> chunks;
    uint32_t N = 0;
    std::vector<uint32_t> all;  // 0 .. N-1
    int64_t total_blocks = 0, max_blocks = 0;
    uint64_t total_residues = 0;
    bool twins = false;  // entries 2k, 2k+1 are a sequence and its reverse complement
    std::vector<int32_t> nb;  // blocks per entry (entry_blocks)
};

// Streams of one entry set for one lane-group width: entries are dealt
// longest first to the least loaded stream (LPT), so streams come out of
// nearly equal length; streams are then ordered longest first and taken in
// runs of 2*64/G per wave, so the waves that start first are the longest.
struct StreamLayout {
    int lanes = 0;
    bool paired = false;  // the lo and hi streams of each lane group hold the same entries
    int32_t max_entry_blocks = 0;  // blocks of the longest entry
    int nwaves = 0;                      // a multiple of kWavesPerWG
    std::vector<int32_t> wave_nblocks;   // per wave
    std::vector<int64_t> wave_res_off;   // per wave, uint2 units
    std::vector<int32_t> slot_first;     // per (wave, slot) + 1: range in slot_entries
    std::vector<uint32_t> slot_entries;  // positions into PairSet::entries, stream by stream
    int64_t res_words = 0;               // uint2 words of the residue array
    uint2* d_res = nullptr;              // device residue array
    ScoreWave* d_waves = nullptr;        // device copies of the wave table and the stream CSR
    int32_t* d_slot_first = nullptr;
    int32_t* d_slot_entries = nullptr;
    uint64_t build = 0;                  // cached layouts: the device DB build they belong to

    void release() {
        if (d_res) (void)hipFree(d_res);
        if (d_waves) (void)hipFree(d_waves);
        if (d_slot_first) (void)hipFree(d_slot_first);
        if (d_slot_entries) (void)hipFree(d_slot_entries);
        d_res = nullptr;
        d_waves = nullptr;
        d_slot_first = nullptr;
        d_slot_entries = nullptr;
    }
};

void build_layout(Exec& ex, StreamLayout& L, int lanes, const PairSet& ps, const std::vector<uint32_t>& order,
                  bool paired) {
    DeviceDB& db = g_db;
    const int ng = 64 / lanes, spw = 2 * ng;
    // distinct streams per wave: both halves of a lane group, or (paired) one
    const int gsz = paired ? ng : spw;
    const std::vector<uint32_t>& pe_list = ps.list();
    const size_t n = pe_list.size();
    L.lanes = lanes;
    L.paired = paired;
    std::vector<int32_t> nbk(n);
    int64_t total = 0;
    L.max_entry_blocks = 0;
    for (size_t k = 0; k < n; ++k) {
        nbk[k] = entry_blocks(db.entries[pe_list[k]].len);
        total += nbk[k];
        L.max_entry_blocks = std::max(L.max_entry_blocks, nbk[k]);
    }
    const int64_t per_wave = std::max<int64_t>(1, ((int64_t)target_waves()));
    const int64_t B = std::max<int64_t>(16 * (int64_t)lanes, (total + gsz * per_wave - 1) / (gsz * per_wave));
    int64_t ns = std::max<int64_t>(1, (total + B - 1) / B);
    const int64_t unit = (int64_t)gsz * kWavesPerWG;
    ns = (ns + unit - 1) / unit * unit;
    const int nw = (int)(ns / gsz);

    // Lockstep groups: the entries in length order (longest first) are cut
    // into groups of spw; a wave runs its groups one after the other, stream
    // s of the wave taking entry s of every group, each entry padded to the
    // group's longest.  All pairs of a group start at the same block, so the
    // lanes of a wave reset their state at a pair start in the same steps
    // (one run of G steps per group) instead of whenever any of the wave's
    // streams starts a pair; consecutive entries in length order differ by
    // little, so the padding is small.  Groups go longest first to the wave
    // least loaded for its capacity (LPT): the last fifth of the waves hold
    // half and the last tenth a quarter of a full wave's share, so that the
    // workgroup tasks that end a launch's queue (costliest first) are small
    // and the launch's tail is short.
    const size_t ngroups = (n + gsz - 1) / gsz;
    std::vector<int64_t> wload(nw, 0);
    std::vector<std::vector<uint32_t>> wgroups(nw);
    {
        auto quarters = [nw](int w) { return w >= nw - nw / 10 ? 1 : w >= nw - nw / 5 ? 2 : 4; };
        typedef std::pair<double, int32_t> Item;  // (load per capacity, wave)
        std::priority_queue<Item, std::vector<Item>, std::greater<Item>> pq;
        for (int w = 0; w < nw; ++w) pq.push(Item(0.0, w));
        for (size_t gi = 0; gi < ngroups; ++gi) {
            Item it = pq.top();
            pq.pop();
            const int w = it.second;
            wgroups[w].push_back((uint32_t)gi);
            wload[w] += nbk[order[gi * gsz]];  // the group's first entry is its longest
            pq.push(Item((double)wload[w] / quarters(w), w));
        }
    }
    // waves longest first, so that the four waves of a workgroup task are alike
    std::vector<int32_t> word(nw);
    std::iota(word.begin(), word.end(), 0);
    std::stable_sort(word.begin(), word.end(), [&](int32_t a, int32_t b) { return wload[a] > wload[b]; });
    L.nwaves = nw;
    L.wave_nblocks.assign(nw, 0);
    L.wave_res_off.assign(nw, 0);
    L.slot_first.assign((size_t)nw * spw + 1, 0);
    L.slot_entries.clear();
    L.slot_entries.reserve(n);
    std::vector<PackEntry> pe;
    pe.reserve(n);
    int64_t off = 0;
    for (int w = 0; w < nw; ++w) {
        const std::vector<uint32_t>& gs = wgroups[word[w]];
        L.wave_nblocks[w] = (int32_t)wload[word[w]];
        L.wave_res_off[w] = off;
        for (int sl = 0; sl < spw; ++sl) {
            const int k = sl / 2, h = sl % 2;
            int64_t pos = 0;
            for (uint32_t gi : gs) {
                const size_t i = gi * gsz + (paired ? k : sl);
                if (i < n) {
                    const uint32_t at = order[i];
                    L.slot_entries.push_back(at);
                    const DbEntry& de = db.entries[pe_list[at]];
                    PackEntry p;
                    p.code_off = de.code_off;
                    p.dst = (off + pos * ng + k) * 2 + h;
                    p.nblocks = entry_blocks(de.len);
                    p.last_col = (int32_t)((de.len + 3) % 4);  // 3 for an empty entry, as (len + 3) % 4 in the reference
                    p.stride = 2 * ng;
                    p.pad = 0;
                    pe.push_back(p);
                }
                pos += nbk[order[gi * gsz]];
            }
            L.slot_first[(size_t)w * spw + sl + 1] = (int32_t)L.slot_entries.size();
        }
        off += (int64_t)L.wave_nblocks[w] * ng;
    }
    L.res_words = off;

    L.release();
    std::vector<ScoreWave> wt(L.nwaves);
    for (int w = 0; w < L.nwaves; ++w) wt[w] = ScoreWave{L.wave_nblocks[w], w * spw, L.wave_res_off[w]};
    SSA_HIP(hipMalloc(&L.d_waves, std::max<size_t>(wt.size(), 1) * sizeof(ScoreWave)));
    SSA_HIP(hipMalloc(&L.d_slot_first, L.slot_first.size() * sizeof(int32_t)));
    SSA_HIP(hipMalloc(&L.d_slot_entries, std::max<size_t>(L.slot_entries.size(), 1) * sizeof(int32_t)));
    ex.upload(L.d_waves, wt.data(), wt.size() * sizeof(ScoreWave));
    ex.upload(L.d_slot_first, L.slot_first.data(), L.slot_first.size() * sizeof(int32_t));
    ex.upload(L.d_slot_entries, L.slot_entries.data(), L.slot_entries.size() * sizeof(int32_t));
    SSA_HIP(hipMalloc(&L.d_res, (size_t)std::max<int64_t>(L.res_words, 1) * sizeof(uint2)));
    auto* d_pe = (PackEntry*)ex.arena("pack", std::max<size_t>(pe.size(), 1) * sizeof(PackEntry));
    ex.upload(d_pe, pe.data(), pe.size() * sizeof(PackEntry));
    SSA_HIP(launch_pack((int)pe.size(), ex.stream(), d_pe, db.d_codes, reinterpret_cast<uint32_t*>(L.d_res),
                        L.res_words * 2));
    ex.sync();  // pe is staged from pageable memory
}

// layouts of the whole DB, kept while the DB stays loaded (one per lane width);
// std::list keeps the addresses of cached layouts stable
std::list<StreamLayout>& layout_cache() {
    static std::list<StreamLayout> c;
    return c;
}

std::mutex& layout_mutex() {
    static std::mutex m;
    return m;
}

}  // namespace

void release_layouts() {
    std::lock_guard<std::mutex> lk(layout_mutex());
    for (StreamLayout& L : layout_cache()) L.release();
    layout_cache().clear();
}

namespace {

bool layout_cached(int lanes, bool paired) {
    std::lock_guard<std::mutex> lk(layout_mutex());
    for (const StreamLayout& c : layout_cache())
        if (c.lanes == lanes && c.paired == paired && c.build == g_db.build) return true;
    return false;
}

}  // namespace

namespace {

// One launch of the score kernel over the pairs (q, e) for all queries q of
// qs and the entries of ps, in three parts: prepare() builds and uploads the
// inputs, launch() issues the kernels (no host synchronisation, so a graph
// can capture it), finish() queues the copy of the results back.  Results are
// in PairSet::index order over qs.
struct TierRun {
    SearchType type = kSW;
    Tier T = kTier16;
    std::vector<QueryBuf> qs;
    const PairSet* ps = nullptr;
    const std::vector<uint32_t>* sorted_all = nullptr;  // ps is the whole DB in this length order
    Shape sh{0, 0};       // 0, 0 = chosen by the cost model
    std::string pfx;      // arena names of this run
    PairResult* out = nullptr;  // results in pinned host memory (PairSet::index order)
    size_t nout = 0;
    uint64_t computed = 0;

    StreamLayout local;
    StreamLayout* L = nullptr;
    int grid = 0;
    ScoreParams P;
    ScoreLaunch K{};
    PairResult* d_out = nullptr;
    // host staging, kept until finish
    std::vector<ScoreTask> tasks;
    std::vector<QueryDesc> qd;
    std::vector<int16_t> prof;
    std::vector<uint32_t> lbv;
    std::vector<SegLane> segs;
    bool packed = false;  // SW with several queries: scored as packed queries
    bool paired = false;  // packed queries two by two, one per half of every lane group (score_kernel_pair)
    bool untracked = false;  // NW 16-bit where no pair can overflow: the kernel without trackers
    bool vp = false;         // the query profile in VGPRs
    int32_t prio_steps = 0;  // the score kernels' wave priority unit (steps)
    bool f16 = false;        // SW first tier in windowed packed fp16 (kernel mode 3), or the
                             // untracked NW kernel in packed fp16 (mode 4)
    int mode() const { return type == kSW ? (f16 ? 3 : 0) : (untracked ? (f16 ? 4 : 2) : 1); }

    ~TierRun() { local.release(); }
    bool empty() const { return ps == nullptr || ps->list().empty() || qs.empty(); }
};

void tier_prepare(Exec& ex, TierRun& R) {
    DeviceDB& db = g_db;
    const Config& cf = config();
    const Matrix& M = matrix();
    const Tier& T = R.T;
    const int go = cf.gap_open, ge = cf.gap_extend;
    const PairSet& ps = *R.ps;
    const size_t n = ps.list().size(), nq = R.qs.size();
    R.nout = ps.size();
    if (R.empty()) return;
    R.out = (PairResult*)ex.pinned(R.pfx + "results", R.nout * sizeof(PairResult));

    // the query length the shape is chosen for: a packed query's total
    size_t qmax = 0, qsum = 0;
    for (const QueryBuf& q : R.qs) qmax = std::max(qmax, q.len), qsum += q.len;
    if (R.type == kSW && nq > 1) qmax = qsum;
    if (R.sh.lanes == 0) {
        int64_t total_blocks = 0, max_blocks = 0;
        for (uint32_t e : ps.list()) {
            const int64_t nb = entry_blocks(db.entries[e].len);
            total_blocks += nb;
            max_blocks = std::max(max_blocks, nb);
        }
        R.sh = choose_shape(qmax, total_blocks, max_blocks, R.type == kSW ? 0 : 1, R.type == kSW && nq > 1,
                            ex.fixed());
    }
    const Shape sh = R.sh;
    if (!score_shape_supported(sh.lanes, sh.rows, R.type == kSW ? 0 : 1, sh.cols))
        fatal("libssa_amd: unsupported kernel shape");
    R.f16 = R.type == kSW && T.bits == 16 && arith_override_ref() != 1 &&
            f16_window_ok(sh.rows, T.wrap(cf.gap_open + cf.gap_extend), T.wrap(cf.gap_extend));
    const int PR = sh.lanes * sh.rows, ng = 64 / sh.lanes, spw = 2 * ng;

    // ---- profiles [pass][code][row], left boundaries, per-query descriptors ----
    R.qd.assign(nq, QueryDesc{});
    for (QueryDesc& d : R.qd) d.seg_off = d.seg_off_hi = -1;
    R.prof.clear();
    R.lbv.clear();
    R.segs.clear();
    std::vector<int> npass(nq);
    const int64_t QRw = T.wrap(go + ge), Rw = T.wrap(ge);
    // SW with several queries: packed queries (below), two by two where the
    // shape has a paired kernel
    R.packed = R.type == kSW && nq > 1;
    R.paired = R.packed && sh.cols == 4 && score_pair_supported(sh.lanes, sh.rows);
    // no layout may be built on the no-allocation path: pairing only where
    // the paired layout is already there, else unpaired packed queries
    if (ex.fixed() && R.paired && !layout_cached(sh.lanes, true)) R.paired = false;
    for (size_t qi = 0; qi < nq && !R.packed; ++qi) {
        const QueryBuf& q = R.qs[qi];
        npass[qi] = (int)((q.len + PR - 1) / PR);
        const size_t rows = (size_t)npass[qi] * PR;
        R.qd[qi].qlen = (int32_t)q.len;
        R.qd[qi].qp_off = (int32_t)R.prof.size();
        R.qd[qi].lb_off = (int32_t)R.lbv.size();
        // the tier's value of every (DB code, query code), then the rows
        int16_t val[kDim][32];
        for (int d = 0; d < kDim; ++d)
            for (int c = 0; c < 32; ++c) val[d][c] = (int16_t)T.repr(T.wrap(M.at(d, c)));
        const int16_t pad = R.type == kSW ? (int16_t)-32768 : (int16_t)0;
        const size_t base = R.prof.size();
        R.prof.resize(base + rows * kDim);
        int16_t* out = R.prof.data() + base;
        for (int p = 0; p < npass[qi]; ++p)
            for (int d = 0; d < kDim; ++d)
                for (int r = 0; r < PR; ++r) {
                    const size_t i = (size_t)p * PR + r;
                    *out++ = i < q.len ? val[d][q.seq[i] & 31] : pad;
                }
        int64_t c = QRw;
        for (size_t i = 0; i < rows; ++i) {
            if (i < q.len) {
                R.lbv.push_back(pack2(T.repr(c)));
                R.lbv.push_back(pack2(T.repr(T.sat(c + QRw))));
                c = T.sat(c + Rw);
            } else {
                R.lbv.push_back(0u);
                R.lbv.push_back(0u);
            }
        }
    }
    ScoreParams& P = R.P;
    P.qr = pack2(T.repr(QRw));
    P.r = pack2(T.repr(Rw));
    P.zero = pack2(-32768);
    if (R.type == kSW) {
        if (T.bits == 16) {
            P.thr_a = 255 - 32768;  // the 8-bit tier's overflow (true maximum >= 255)
            P.thr_b = 32767;        // the 16-bit tier's overflow
        } else {
            P.thr_a = T.repr(127);
            P.thr_b = 1 << 20;
        }
    } else {
        P.thr_a = T.repr(T.wrap(T.imin() - go - ge - 1));
        P.thr_b = T.repr(T.imax());
    }
    if (R.f16) {  // real scores in fp16 windows: real 0 at offset 0, thresholds as real maxima
        // values of DB column j held less j ge: E and F take H + gap open
        // (P.qr), F adds ge once more (P.r), the profile holds score - ge
        P.qr = pack2(f16_bits(go));
        P.r = pack2(f16_bits(Rw));
        P.zero = 0;
        P.thr_a = 255;    // the 8-bit tier's overflow
        P.thr_b = 65535;  // the 16-bit tier's overflow
    }
    P.hb_init[0] = 0;
    for (int k = 1; k < 4; ++k) P.hb_init[k] = pack2(T.repr(T.wrap(go + k * ge)));
    for (int k = 0; k < 4; ++k) P.fb_init[k] = pack2(T.repr(T.wrap(go + (k + 1) * ge)));
    P.shift = R.f16 ? (int32_t)Rw : 0;
    P.e4 = 0;
    P.thr_start = 0x7c007c00u;
    if (R.f16) {  // the fire threshold a half starts a pair with (no flag, offset 0, column 4)
        const int64_t t = std::min<int64_t>(P.thr_a, P.thr_b) - 1 - 4 * Rw;
        P.e4 = pack2(f16_bits(-4 * Rw));
        if (t <= 2047) P.thr_start = pack2(f16_bits(t));  // (above, the trigger at 1,536 fires first)
    }

    // ---- streams: cached for the whole DB, built per call for subsets ----
    if (R.sorted_all) {
        std::lock_guard<std::mutex> lk(layout_mutex());
        for (StreamLayout& c : layout_cache())
            if (c.lanes == sh.lanes && c.paired == R.paired && c.build == db.build) R.L = &c;
        if (!R.L) {
            if (ex.fixed())
                throw FixedRegionError{-3, "the DB's stream layout for this kernel shape is not resident; run "
                                           "sw_align or nw_align once first"};
            layout_cache().emplace_back();
            R.L = &layout_cache().back();
            build_layout(ex, *R.L, sh.lanes, ps, *R.sorted_all, R.paired);
            R.L->build = db.build;
        }
    } else {
        if (ex.fixed()) throw FixedRegionError{-3, "a search over part of the DB builds a stream layout"};
        std::vector<uint32_t> order(n);
        std::iota(order.begin(), order.end(), 0u);
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            return db.entries[ps.list()[a]].len > db.entries[ps.list()[b]].len;
        });
        build_layout(ex, R.local, sh.lanes, ps, order, R.paired);
        R.L = &R.local;
    }
    const StreamLayout* L = R.L;

    const int nw = L->nwaves;
    if ((uint64_t)nq * n >= (uint64_t)INT32_MAX) fatal("libssa_amd: too many (query, entry) pairs in one launch");
    const int nt_q = nw / kWavesPerWG;
    // NW at 16 bit where no value of any pair can reach either end of the
    // range: no block can be flagged, so the kernel drops the trackers
    {
        const MatrixInfo mi = matrix_info();
        R.untracked = R.type == kNW && T.bits == 16 && score_nw_untracked_supported(sh.lanes, sh.rows);
        for (size_t qi = 0; qi < nq && R.untracked; ++qi)
            R.untracked = int16_exact(R.qs[qi].len, 4 * (int64_t)L->max_entry_blocks, mi);
        // and in fp16 where every value lies within +-2047
        if (R.untracked && arith_override_ref() != 1) {
            R.f16 = true;
            for (size_t qi = 0; qi < nq && R.f16; ++qi)
                R.f16 = nw_f16_exact(R.qs[qi].len, 4 * (int64_t)L->max_entry_blocks, mi);
        }
        if (R.type == kNW && R.f16) {
            // every value of cell (i, j) less (i + j) ge (nw_f16_exact): the
            // extension adds drop out (E and F take H + gap open, P.r = 0),
            // the diagonal takes the score - 2 ge, and the borders are
            // constants: H(0, j) = H(i, 0) = gap open, E(i, 1) = F(1, j) =
            // 2 gap open, the corner 0
            P.qr = pack2(f16_bits(go));
            P.r = 0;
            P.zero = 0;
            P.hb_init[0] = 0;
            for (int c = 1; c < 4; ++c) P.hb_init[c] = pack2(f16_bits(go));
            for (int c = 0; c < 4; ++c) P.fb_init[c] = pack2(f16_bits(go));
            P.shift = (int32_t)ge;
            for (size_t qi = 0; qi < nq; ++qi) {
                const size_t len = R.qs[qi].len, rows = (size_t)npass[qi] * PR;
                uint32_t* lb = R.lbv.data() + R.qd[qi].lb_off;
                for (size_t i = 0; i < rows; ++i) {
                    lb[2 * i] = i < len ? pack2(f16_bits(go)) : 0u;
                    lb[2 * i + 1] = i < len ? pack2(f16_bits(2 * go)) : 0u;
                }
                // the rows of the query (the padding rows keep 0); the
                // profile becomes fp16 at upload
                int16_t* pr = R.prof.data() + R.qd[qi].qp_off;
                for (int p = 0; p < npass[qi]; ++p)
                    for (int d = 0; d < kDim; ++d)
                        for (int r = 0; r < PR; ++r)
                            if ((size_t)p * PR + r < len) pr[((size_t)p * kDim + d) * PR + r] -= (int16_t)(2 * ge);
            }
        }
    }
    // the query profile in VGPRs: forced, or the autotuner's choice for the
    // shape, where the DB is A, C, G and T only and the kernel has it
    {
        const int po = placement_override_ref();
        R.vp = !R.paired && db.acgt && score_vp_supported(sh.lanes, sh.rows, R.mode(), sh.cols) &&
               (po == 2 || (po == 0 && sh.place == 2));
    }
    const int cap = R.paired ? score_pair_resident_workgroups(sh.lanes, sh.rows, R.f16)
                             : score_resident_workgroups(sh.lanes, sh.rows, R.mode(), sh.cols, R.vp);

    // ---- packed queries (SW, several queries) ----
    // Each query starts at a lane boundary and is padded to whole lanes (the
    // padding rows score the lane minimum and never raise a maximum), so that
    // short queries fill passes.  The queries are taken in order into packed
    // queries of at most sq_max passes: a task runs all passes of its waves
    // one after the other, so its length is passes x its longest wave, and
    // the tasks over the longest waves (the longest DB entries) must stay a
    // small part of a workgroup's share of the launch, or they run on alone
    // after the rest has finished.
    std::vector<int> sq_npass;
    if (R.packed) {
        int64_t per_pass = 0, longest = 1;  // task steps per pass: all tasks, the longest task
        for (int t = 0; t < nt_q; ++t) {
            per_pass += L->wave_nblocks[t * kWavesPerWG] + sh.lanes;
            longest = std::max<int64_t>(longest, L->wave_nblocks[t * kWavesPerWG] + sh.lanes);
        }
        std::vector<int64_t> qrows(nq);
        int64_t total = 0;
        for (size_t qi = 0; qi < nq; ++qi) {
            qrows[qi] = (int64_t)((R.qs[qi].len + sh.rows - 1) / sh.rows) * sh.rows;
            total += qrows[qi];
        }
        const int halves = R.paired ? 2 : 1;
        const int64_t total_passes = std::max<int64_t>(1, (total + (int64_t)halves * PR - 1) / ((int64_t)halves * PR));
        const int64_t sq_max =
            std::max<int64_t>(1, total_passes * per_pass / (4 * (int64_t)std::max(cap, 1) * longest));
        // a packed query: the rows of the listed queries end to end
        struct Rows {
            std::vector<uint32_t> qs;
            std::vector<int64_t> start;
            int64_t rows = 0;
        };
        auto rows_of = [&](const std::vector<uint32_t>& list) {
            Rows r;
            r.qs = list;
            for (uint32_t qi : list) {
                r.start.push_back(r.rows);
                r.rows += qrows[qi];
            }
            return r;
        };
        // cuts a list of queries (in order) into packed queries of at most sq_max passes
        auto cut = [&](const std::vector<uint32_t>& list) {
            std::vector<std::vector<uint32_t>> out;
            std::vector<uint32_t> cur;
            int64_t rows = 0;
            for (uint32_t qi : list) {
                if (!cur.empty() && rows + qrows[qi] > sq_max * PR) {
                    out.push_back(cur);
                    cur.clear();
                    rows = 0;
                }
                cur.push_back(qi);
                rows += qrows[qi];
            }
            if (!cur.empty()) out.push_back(cur);
            return out;
        };
        // the profile values and segments of one half of a packed query
        auto half = [&](const Rows& rw, int np, int h, int16_t* pf, std::vector<SegLane>& segs) {
            std::vector<int32_t> row_q((size_t)np * PR + sh.rows, -1);  // list position of each row
            for (size_t x = 0; x < rw.qs.size(); ++x)
                for (int64_t i = rw.start[x]; i < rw.start[x] + qrows[rw.qs[x]]; ++i) row_q[i] = (int32_t)x;
            for (int p = 0; p < np; ++p)
                for (int c = 0; c < kDim; ++c)
                    for (int r = 0; r < PR; ++r) {
                        const int64_t i = (int64_t)p * PR + r;
                        const int32_t x = row_q[i];
                        int32_t v = -32768;
                        if (x >= 0) {
                            const QueryBuf& q = R.qs[rw.qs[x]];
                            const int64_t k = i - rw.start[x];
                            if (k < (int64_t)q.len) v = T.repr(T.wrap(M.at(c, q.seq[k])));
                        }
                        pf[(((size_t)p * 32 + c) * PR + r) * halves + h] = (int16_t)v;
                    }
            for (int p = 0; p < np; ++p)
                for (int g = 0; g < sh.lanes; ++g) {
                    const int64_t i = (int64_t)p * PR + (int64_t)g * sh.rows;
                    const int32_t x = row_q[i];
                    if (x < 0) {
                        segs.push_back(SegLane{0, kSegNone});
                        continue;
                    }
                    const bool top = i == rw.start[x];
                    const bool bottom = g == sh.lanes - 1 || row_q[i + sh.rows] != x;
                    segs.push_back(
                        SegLane{(int32_t)(rw.qs[x] * n), (top ? kSegTop : 0) | (bottom ? kSegBottom : 0)});
                }
        };
        // one packed query (lo), or two sharing the streams (lo, hi)
        auto build = [&](const std::vector<uint32_t>& lo, const std::vector<uint32_t>& hi) {
            const Rows rl = rows_of(lo), rh = rows_of(hi);
            const int64_t rows = std::max(rl.rows, rh.rows);
            const int np = (int)((rows + PR - 1) / PR);
            if (np == 0) return;  // empty queries only: their results keep the initial values
            if ((int64_t)R.prof.size() + (int64_t)np * 32 * PR * halves >= (int64_t)INT32_MAX)
                fatal("libssa_amd: packed queries too long");
            QueryDesc d{};
            d.qlen = (int32_t)std::min<int64_t>(rows, INT32_MAX);
            d.qp_off = (int32_t)R.prof.size();
            R.prof.resize(R.prof.size() + (size_t)np * 32 * PR * halves);
            int16_t* pf = R.prof.data() + d.qp_off;
            d.seg_off = (int32_t)R.segs.size();
            half(rl, np, 0, pf, R.segs);
            d.seg_off_hi = (int32_t)R.segs.size();
            if (R.paired) half(rh, np, 1, pf, R.segs);
            R.qd.push_back(d);
            sq_npass.push_back(np);
        };
        std::vector<uint32_t> all_q(nq);
        std::iota(all_q.begin(), all_q.end(), 0u);
        if (!R.paired) {
            for (const auto& sq : cut(all_q)) build(sq, {});
        } else {
            // two halves of nearly equal rows (longest queries first, each to
            // the half with fewer rows), cut alike and paired in order
            std::stable_sort(all_q.begin(), all_q.end(), [&](uint32_t a, uint32_t b) { return qrows[a] > qrows[b]; });
            std::vector<uint32_t> lo, hi;
            int64_t rlo = 0, rhi = 0;
            for (uint32_t qi : all_q) {
                if (rlo <= rhi) {
                    lo.push_back(qi);
                    rlo += qrows[qi];
                } else {
                    hi.push_back(qi);
                    rhi += qrows[qi];
                }
            }
            const auto slo = cut(lo), shi = cut(hi);
            for (size_t x = 0; x < std::max(slo.size(), shi.size()); ++x)
                build(x < slo.size() ? slo[x] : std::vector<uint32_t>(), x < shi.size() ? shi[x] : std::vector<uint32_t>());
        }
    }

    // ---- tasks: every (packed) query against every group of 4 waves, costliest first ----
    R.tasks.clear();
    std::vector<int64_t> cost;
    bool multipass = false;
    for (size_t qi = 0; qi < nq; ++qi) R.qd[qi].pair_off = (int32_t)(qi * n);
    // a query of no rows gets no tasks: its results keep the initial values
    // (SW: score 0, no flag)
    auto add_tasks = [&](int32_t query, int np) {
        if (np <= 0) return;
        multipass |= np > 1;
        for (int t = 0; t < nt_q; ++t) {
            R.tasks.push_back(ScoreTask{query, t * kWavesPerWG});
            cost.push_back((int64_t)L->wave_nblocks[t * kWavesPerWG] * np + sh.lanes);
        }
        for (int w = 0; w < nw; ++w) R.computed += (uint64_t)L->wave_nblocks[w] * 4 * spw * (uint64_t)np * PR;
    };
    if (R.packed) {
        for (size_t s = 0; s < sq_npass.size(); ++s) add_tasks((int32_t)(nq + s), sq_npass[s]);
    } else {
        for (size_t qi = 0; qi < nq; ++qi) add_tasks((int32_t)qi, npass[qi]);
    }
    if (R.tasks.size() > (size_t)nt_q) {
        std::vector<uint32_t> ord(R.tasks.size());
        std::iota(ord.begin(), ord.end(), 0u);
        std::stable_sort(ord.begin(), ord.end(), [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
        std::vector<ScoreTask> sorted(R.tasks.size());
        for (size_t i = 0; i < ord.size(); ++i) sorted[i] = R.tasks[ord[i]];
        R.tasks.swap(sorted);
    }
    R.grid = std::max(1, std::min<int>((int)R.tasks.size(), cap > 0 ? cap : (int)R.tasks.size()));
    // the wave priority unit: half the median task's steps
    if (!cost.empty()) {
        std::vector<int64_t> c(cost);
        std::nth_element(c.begin(), c.begin() + c.size() / 2, c.end());
        R.prio_steps = (int32_t)std::max<int64_t>(1, c[c.size() / 2] / (sh.cols / 4) / 2);
    }
    int32_t max_nb = 0;
    for (int w = 0; w < nw; ++w) max_nb = std::max(max_nb, L->wave_nblocks[w]);
    // the kernel's pass length: the longest wave of a task, at least G + 2
    // steps, an even number of steps
    const int64_t nb_step = sh.cols / 4;
    const int64_t max_nbp =
        (std::max<int64_t>(max_nb, nb_step * (sh.lanes + 2)) + 2 * nb_step - 1) / (2 * nb_step) * (2 * nb_step);
    // (mode 3 keeps each block's window offsets, one uint32 per lane group,
    // after the bottom rows).  Both parts are whole 128-byte lines, so that
    // every wave's slice starts on a line and a block's bottom rows of the
    // wave's lane groups (128 bytes for 16-lane groups) fill one line instead
    // of straddling two.
    auto lines = [](int64_t u4) { return (u4 + 7) / 8 * 8; };
    const int64_t bnd_rows = multipass ? lines(max_nbp * ng * 2) : 0;
    const int64_t bnd_per_wave =
        bnd_rows + (multipass && R.type == kSW && R.f16 ? lines((max_nbp * ng + 3) / 4) : 0);

    // ---- device buffers ----
    auto name = [&](const char* s) { return R.pfx + s; };
    if (ex.fixed() && bnd_per_wave > 0) {
        // the persistent grid keeps one slice of pass boundaries per resident
        // workgroup: on the no-allocation path as many workgroups run as the
        // scratch left after this launch's other buffers holds slices
        // (fewer workgroups take more tasks each; the results are the same)
        const size_t other = 256 * 8 + R.tasks.size() * sizeof(ScoreTask) + R.qd.size() * sizeof(QueryDesc) +
                             R.prof.size() * sizeof(int16_t) + R.lbv.size() * sizeof(uint32_t) +
                             R.segs.size() * sizeof(SegLane) + R.nout * sizeof(PairResult);
        const size_t per_wg = (size_t)kWavesPerWG * (size_t)bnd_per_wave * sizeof(uint4);
        const size_t left = ex.region_left() > other ? ex.region_left() - other : 0;
        const int fit = (int)std::min<size_t>(left / per_wg, (size_t)INT32_MAX);
        if (fit < 1) throw FixedRegionError{-4, "the query's pass boundaries do not fit the no-allocation scratch"};
        R.grid = std::min(R.grid, fit);
    }
    auto* d_tasks = (ScoreTask*)ex.arena(name("tasks"), std::max<size_t>(R.tasks.size(), 1) * sizeof(ScoreTask));
    auto* d_qd = (QueryDesc*)ex.arena(name("qdesc"), R.qd.size() * sizeof(QueryDesc));
    auto* d_prof = (int16_t*)ex.arena(name("prof"), std::max<size_t>(R.prof.size(), 8) * sizeof(int16_t));
    auto* d_lb = (uint32_t*)ex.arena(name("lb"), std::max<size_t>(R.lbv.size(), 1) * sizeof(uint32_t));
    auto* d_segs = (SegLane*)ex.arena(name("segs"), std::max<size_t>(R.segs.size(), 1) * sizeof(SegLane));
    auto* d_bnd = (uint4*)ex.arena(
        name("bnd"), (size_t)std::max<int64_t>((int64_t)R.grid * kWavesPerWG * bnd_per_wave, 1) * sizeof(uint4));
    auto* d_counter = (int32_t*)ex.arena(name("counter"), sizeof(int32_t));
    R.d_out = (PairResult*)ex.arena(name("results"), R.nout * sizeof(PairResult));
    ex.upload(d_tasks, R.tasks.data(), R.tasks.size() * sizeof(ScoreTask));
    ex.upload(d_qd, R.qd.data(), R.qd.size() * sizeof(QueryDesc));
    // the fp16 kernels' profile: the 16-bit tier's scores as fp16 values (the
    // padding rows' -32768 (SW) and 0 (NW) as well); SW's scores less the gap
    // extend (the column shift of its values; NW's shift is in already)
    if (R.f16 && R.type == kSW) {
        const int16_t sh = (int16_t)-R.P.shift;
        for (int16_t& x : R.prof) x = f16_of_int16(x == -32768 ? x : (int16_t)(x + sh));
    } else if (R.f16) {
        for (int16_t& x : R.prof) x = f16_of_int16(x);
    }
    ex.upload(d_prof, R.prof.data(), R.prof.size() * sizeof(int16_t));
    ex.upload(d_lb, R.lbv.data(), R.lbv.size() * sizeof(uint32_t));
    ex.upload(d_segs, R.segs.data(), R.segs.size() * sizeof(SegLane));
    ScoreLaunch& K = R.K;
    K.waves = L->d_waves;
    K.tasks = d_tasks;
    K.ntasks = (int32_t)R.tasks.size();
    K.prio_steps = R.prio_steps;
    K.counter = d_counter;
    K.queries = d_qd;
    K.qp = d_prof;
    K.lb = d_lb;
    K.res = L->d_res;
    K.bnd = d_bnd;
    K.bnd_per_wave = bnd_per_wave;
    K.bnd_rows = bnd_rows;
    K.slot_first = L->d_slot_first;
    K.slot_entries = L->d_slot_entries;
    K.results = R.d_out;
    K.segs = d_segs;
    K.trace = task_trace_ref() && !R.tasks.empty()
                  ? (TaskTrace*)ex.arena(name("trace"), R.tasks.size() * sizeof(TaskTrace))
                  : nullptr;
}

void tier_launch(hipStream_t st, TierRun& R) {
    if (R.empty()) return;
    SSA_HIP(launch_init_results(st, R.d_out, (int64_t)R.nout, -32768));
    if (R.paired) SSA_HIP(launch_score_pair(R.sh.lanes, R.sh.rows, R.f16, st, R.K, R.P, R.grid));
    else SSA_HIP(launch_score(R.sh.lanes, R.sh.rows, R.mode(), R.sh.cols, R.vp, st, R.K, R.P, R.grid));
}

void tier_finish(Exec& ex, TierRun& R) {
    if (R.empty()) return;
    ex.download(R.out, R.d_out, R.nout * sizeof(PairResult));
}

// One tier run on its own, timed.
void gpu_score_tier(Exec& ex, TierRun& R, double* ms) {
    tier_prepare(ex, R);
    GpuTimer tm(ex.stream());
    tm.start();
    tier_launch(ex.stream(), R);
    tm.stop();
    tier_finish(ex, R);
    ex.sync();
    *ms += tm.ms();
}

// Exact scores of the pairs (q, e) listed (the reference's 64-bit tier,
// full_sw / full_nw): one wave per pair, in int32 where every value of the
// pair provably fits, else in int64 as the reference computes.
void gpu_exact(Exec& gpu, SearchType type, const std::vector<QueryBuf>& qs,
               const std::vector<std::pair<uint32_t, uint32_t>>& pairs, std::vector<int64_t>& out, double* ms,
               uint64_t* cells_computed) {
    out.assign(pairs.size(), 0);
    if (pairs.empty()) return;
    DeviceDB& db = g_db;
    const Config& cf = config();
    const MatrixInfo mi = matrix_info();
    std::vector<QueryDesc> qd(qs.size());
    std::vector<uint8_t> qcodes;
    for (size_t qi = 0; qi < qs.size(); ++qi) {
        qd[qi] = QueryDesc{};
        qd[qi].qlen = (int32_t)qs[qi].len;
        qd[qi].qp_off = (int32_t)qcodes.size();
        qcodes.insert(qcodes.end(), qs[qi].seq, qs[qi].seq + qs[qi].len);
    }
    // every value of a pair stays within +-(|go|+|ge|)*(qlen+dlen+4)*3 + max|s|*min(qlen,dlen)
    const int64_t gap = std::abs((int64_t)cf.gap_open) + std::abs((int64_t)cf.gap_extend);
    const int64_t smax = std::max(std::abs(mi.maxv), std::abs(mi.minv));
    std::vector<ExactPair> ep[2];  // int32, int64
    int64_t bnd_elems[2] = {0, 0};
    for (size_t k = 0; k < pairs.size(); ++k) {
        const DbEntry& e = db.entries[pairs[k].second];
        const size_t ql = qs[pairs[k].first].len;
        const int64_t bound = gap * (int64_t)(ql + e.len + 4) * 3 + smax * (int64_t)std::min<size_t>(ql, e.len);
        const int w = (mi.fits32 && bound < INT32_MAX / 2) ? 0 : 1;
        ExactPair x{};
        x.res_off = e.code_off;
        x.dlen = (int32_t)e.len;
        x.query = (int32_t)pairs[k].first;
        x.out = (int32_t)k;
        x.bnd_off = bnd_elems[w];
        if (ql > (size_t)kExactRows) bnd_elems[w] += 2 * (int64_t)e.len;
        ep[w].push_back(x);
        *cells_computed += (uint64_t)ql * e.len;
    }
    auto* d_qd = (QueryDesc*)gpu.arena("exact_qdesc", qd.size() * sizeof(QueryDesc));
    auto* d_qc = (uint8_t*)gpu.arena("exact_qcodes", std::max<size_t>(qcodes.size(), 1));
    auto* d_mat = (int64_t*)gpu.arena("exact_mat", kDim * kDim * sizeof(int64_t));
    auto* d_out = (int64_t*)gpu.arena("exact_out", pairs.size() * sizeof(int64_t));
    gpu.upload(d_qd, qd.data(), qd.size() * sizeof(QueryDesc));
    gpu.upload(d_qc, qcodes.data(), qcodes.size());
    gpu.upload(d_mat, matrix().s, kDim * kDim * sizeof(int64_t));
    ExactPair* d_ep[2];
    void* d_bnd[2];
    for (int w = 0; w < 2; ++w) {
        const std::string tag = w ? "exact64_" : "exact32_";
        d_ep[w] = (ExactPair*)gpu.arena(tag + "pairs", std::max<size_t>(ep[w].size(), 1) * sizeof(ExactPair));
        d_bnd[w] = gpu.arena(tag + "bnd", (size_t)std::max<int64_t>(bnd_elems[w], 1) * (w ? 8 : 4));
        gpu.upload(d_ep[w], ep[w].data(), ep[w].size() * sizeof(ExactPair));
    }
    GpuTimer tm(gpu.stream());
    tm.start();
    for (int w = 0; w < 2; ++w)
        SSA_HIP(launch_exact(w == 1, (int)ep[w].size(), gpu.stream(), d_ep[w], d_qd, d_qc, d_mat, db.d_codes,
                             type == kSW ? 0 : 1, cf.gap_open, cf.gap_extend, d_bnd[w], d_out));
    tm.stop();
    gpu.download(out.data(), d_out, out.size() * sizeof(int64_t));
    gpu.sync();
    *ms += tm.ms();
}

// Decodes one tier result into (duration, accept, score).
TierOut decode(SearchType type, const Tier& T, const PairResult& r, int32_t nb, bool sw16_as_8) {
    int32_t flag, score;
    bool accept;
    if (type == kSW) {
        flag = sw16_as_8 ? r.flag_a : (T.bits == 16 ? r.flag_b : r.flag_a);
        score = (int32_t)(T.decode(r.score) - T.imin());
        accept = flag == kNoFlag;
    } else {
        flag = r.flag_a;
        score = (int32_t)T.decode(r.score);
        accept = flag == kNoFlag && score > T.imin() && score < T.imax();
    }
    const int32_t dur = flag == kNoFlag ? nb : std::min(nb, flag + 1);
    return TierOut{accept ? dur : -dur, score};
}

}  // namespace

// ============================================================================
// replay of the reference's lane schedule and heap
// ============================================================================
namespace {

struct Elem {
    uint64_t db_id;
    uint8_t db_frame, db_strand, query_id;
    int64_t score;
};

// bounded min-heap with the reference's rules (src/util/minheap.c:60-91)
class MinHeap {
public:
    explicit MinHeap(size_t k) : alloc_(k) { a_.reserve(k); }
    // true when add(e) would leave the heap as it is: full, and the root's
    // score is not below e's
    bool rejects(int64_t score) const { return a_.size() >= alloc_ && (a_.empty() || !(a_[0].score < score)); }
    void add(const Elem& e) {
        if (a_.size() < alloc_) {
            size_t i = a_.size();
            a_.push_back(e);
            while (i > 0) {
                const size_t p = (i - 1) / 2;
                if (!(e.score < a_[p].score)) break;
                a_[i] = a_[p];
                i = p;
            }
            a_[i] = e;
        } else if (!a_.empty() && a_[0].score < e.score) {
            size_t p = 0, c = 1;
            const size_t cnt = a_.size();
            while (c < cnt) {
                if (c + 1 < cnt && a_[c + 1].score < a_[c].score) ++c;
                if (a_[c].score < e.score) a_[p] = a_[c];
                else break;
                p = c;
                c = 2 * p + 1;
            }
            a_[p] = e;
        }
    }
    // heap order, then a stable sort by (score desc, db_id desc) as minheap_sort
    // under a merge-sort qsort (src/util/minheap.c:93-106)
    std::vector<Elem> sorted() const {
        std::vector<Elem> v = a_;
        std::stable_sort(v.begin(), v.end(), [](const Elem& x, const Elem& y) {
            if (x.score != y.score) return x.score > y.score;
            return x.db_id > y.db_id;
        });
        return v;
    }

private:
    size_t alloc_;
    std::vector<Elem> a_;
};

// One reference kernel call: W lanes take the items in order; a lane that
// takes an item in iteration t retires it in iteration t + dur; all lanes
// retiring in one iteration are served in lane order and immediately take the
// next item (src/algo/simd/search_simd_sw.c:338-428).  Calls emit(k) for the
// items (positions into the list) in retirement order.  The lanes sit in a
// binary min-heap of keys (retire iteration << 6 | lane), so one integer
// compare orders them by iteration, then lane (W <= 64).
template <class Dur, class Emit>
void lane_schedule(int W, size_t n, const Dur& dur, const Emit& emit) {
    uint64_t key[64];
    size_t item[64];
    int m = 0;
    size_t next = 0;
    for (int c = 0; c < W && next < n; ++c, ++next) {
        key[m++] = ((uint64_t)dur(next) << 6) | (uint64_t)c;
        item[c] = next;
    }
    auto sift_down = [&](int i) {
        const uint64_t k = key[i];
        for (;;) {
            int c = 2 * i + 1;
            if (c >= m) break;
            if (c + 1 < m && key[c + 1] < key[c]) ++c;
            if (key[c] >= k) break;
            key[i] = key[c];
            i = c;
        }
        key[i] = k;
    };
    for (int i = m / 2 - 1; i >= 0; --i) sift_down(i);
    while (m > 0) {
        const uint64_t k = key[0];
        const int lane = (int)(k & 63);
        emit(item[lane]);
        if (next < n) {
            key[0] = (((k >> 6) + (uint64_t)dur(next)) << 6) | (uint64_t)lane;
            item[lane] = next++;
        } else {
            key[0] = key[--m];
        }
        if (m > 0) sift_down(0);
    }
}

}  // namespace

// ============================================================================
// results list and traceback
// ============================================================================
namespace {

// the sequence the reference attaches to a hit (us_prepare_sequence,
// src/util/util_sequence.c:384-417), unknown symbols mapped to 0: for a
// nucleotide DB the forward strand unless strand == 2, which the adapter never
// produces; for a translated DB the frame of the strand the field names (0
// forward, otherwise the reverse complement), so a one-strand forward search,
// whose entries carry strand 1, gets the reverse strand's frame here.  (The
// reference translates into an uninitialised buffer there; this is the
// result when that buffer is a fresh allocation.)
std::vector<uint8_t> prepared_db_sequence(p_seqinfo s, int strand_field, int frame_field) {
    std::vector<uint8_t> v(s->seqlen);
    const int st = config().symtype;
    const bool nuc = st == NUCLEOTIDE || db_is_translated(st);
    warn_unknown(map_db_codes(s->seq, s->seqlen, nuc, v.data()));
    if (db_is_translated(st)) {
        std::vector<uint8_t> p;
        translate(translation_table(config().d_gencode), v.data(), v.size(), strand_field, frame_field, p);
        v.swap(p);
    } else if (nuc && strand_field == 2) {
        std::vector<uint8_t> r(v.size());
        revcomp(v.data(), v.size(), r.data());
        v.swap(r);
    }
    return v;
}

// Direction-matrix bytes one traceback launch may use: the hits of a batch
// run in consecutive launches whose matrices fit this budget (a hit larger
// than the budget runs alone), so the arena stays bounded however many long
// hits a batch asks for.
constexpr int64_t kTraceDirBudget = (int64_t)1 << 30;

// Shape of the workgroup traceback for a hit of qlen query rows and dlen DB
// columns: rows per lane (1, 2, 4, 8) and waves (1-16), the one of least
// modelled time -- phases of the pass schedule (passes start kTgL phases
// apart, a wave takes its next pass once it has finished the one before) times
// a step's cost (a fixed part for the carries and loads, a part per row).
struct TbShape {
    int rows = 8, waves = 1;
};

TbShape traceback_shape(int qlen, int dlen) {
    constexpr int kS = 16, kL = 5;  // as the kernel's phase width and pass lag
    const int nblk = (dlen + 63 + kS - 1) / kS;
    TbShape best;
    double best_cost = 1e300;
    std::vector<int> start;
    for (int R : {8, 4, 2, 1}) {
        const int P = std::max(1, (qlen + 64 * R - 1) / (64 * R));
        if (P > traceback_wg_max_passes()) continue;
        const int W = std::min(16, P);
        start.assign(P, 0);
        for (int p = 1; p < P; ++p) {
            start[p] = start[p - 1] + kL;
            if (p >= W) start[p] = std::max(start[p], start[p - W] + nblk);
        }
        const double cost = (double)(start[P - 1] + nblk) * kS * (12.0 + 10.0 * R);
        if (cost < best_cost) {
            best_cost = cost;
            best = TbShape{R, W};
        }
    }
    return best;
}

// Regions and CIGARs of the hits listed (types[k] = the search type of al[k]).
void gpu_traceback(Exec& gpu, const std::vector<SearchType>& types, std::vector<p_alignment>& al_in, double* ms) {
    if (al_in.empty()) return;
    const Config& cf = config();
    // hits in order of their workgroup shape, so that each run of one shape is one launch
    std::vector<TbShape> shape_of(al_in.size());
    std::vector<size_t> order(al_in.size());
    for (size_t k = 0; k < al_in.size(); ++k) {
        shape_of[k] = traceback_shape((int)al_in[k]->query.len, (int)al_in[k]->db_seq.len);
        order[k] = k;
    }
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return shape_of[x].rows != shape_of[y].rows ? shape_of[x].rows > shape_of[y].rows
                                                     : shape_of[x].waves > shape_of[y].waves;
    });
    std::vector<p_alignment> al(al_in.size());
    std::vector<TbShape> shapes(al_in.size());
    std::vector<SearchType> ty(al_in.size());
    for (size_t k = 0; k < order.size(); ++k) {
        al[k] = al_in[order[k]];
        shapes[k] = shape_of[order[k]];
        ty[k] = types[order[k]];
    }
    std::vector<uint8_t> codes;
    std::vector<TraceTask> tasks(al.size());
    std::vector<int64_t> dir_bytes(al.size());
    int64_t hh_total = 0, ops_total = 0;
    bool int32_ok = true;
    int64_t smax = 0;
    for (int i = 0; i < kDim * kDim; ++i) smax = std::max<int64_t>(smax, std::abs(matrix().s[i]));
    for (size_t k = 0; k < al.size(); ++k) {
        p_alignment a = al[k];
        TraceTask& t = tasks[k];
        t.q_off = (int64_t)codes.size();
        codes.insert(codes.end(), (const uint8_t*)a->query.seq, (const uint8_t*)a->query.seq + a->query.len);
        t.d_off = (int64_t)codes.size();
        codes.insert(codes.end(), (const uint8_t*)a->db_seq.seq, (const uint8_t*)a->db_seq.seq + a->db_seq.len);
        t.qlen = (int32_t)a->query.len;
        t.dlen = (int32_t)a->db_seq.len;
        // 8-byte aligned per task (packed stores)
        dir_bytes[k] = ((((int64_t)t.qlen + 7) & ~(int64_t)7) * t.dlen + 8 + 7) & ~(int64_t)7;
        t.hh_off = hh_total;
        // int64 scratch of the one-thread kernel, or the int32 carries and
        // last-row copies of the workgroup kernel, whichever is larger
        hh_total += std::max<int64_t>(2 * (int64_t)std::max(t.qlen, t.dlen) + 2,
                                      (3 * (int64_t)t.dlen + 4 * (int64_t)t.qlen) / 2 + 2);
        // the workgroup kernel computes in int32: every value of the hit must fit
        const int64_t bound = ((int64_t)t.qlen + t.dlen + 4) * (std::abs((int64_t)cf.gap_open) +
                                                                 std::abs((int64_t)cf.gap_extend) + smax);
        if (bound >= (int64_t)1 << 30) int32_ok = false;
        t.ops_off = ops_total;
        t.ops_cap = t.qlen + t.dlen + 1;
        ops_total += t.ops_cap;
        t.type = ty[k] == kSW ? 0 : 1;
    }
    // launches: consecutive hits whose direction matrices fit the budget
    std::vector<std::pair<size_t, size_t>> launches;  // [first, last) hit
    int64_t dir_max = 8;
    for (size_t k = 0; k < al.size();) {
        size_t e = k;
        int64_t used = 0;
        while (e < al.size() && (e == k || used + dir_bytes[e] <= kTraceDirBudget)) {
            tasks[e].dir_off = used;
            used += dir_bytes[e];
            ++e;
        }
        dir_max = std::max(dir_max, used);
        launches.push_back({k, e});
        k = e;
    }
    const int64_t t0 = now_us();
    auto* d_codes = (uint8_t*)gpu.arena("tb_codes", codes.size() + 1);
    auto* d_tasks = (TraceTask*)gpu.arena("tb_tasks", tasks.size() * sizeof(TraceTask));
    auto* d_mat = (int64_t*)gpu.arena("tb_mat", kDim * kDim * sizeof(int64_t));
    auto* d_dirs = (uint8_t*)gpu.arena("tb_dirs", (size_t)dir_max);
    auto* d_hh = (int64_t*)gpu.arena("tb_hh", (size_t)hh_total * sizeof(int64_t));
    auto* d_ops = (int32_t*)gpu.arena("tb_ops", (size_t)ops_total * 2 * sizeof(int32_t));
    auto* d_res = (TraceResult*)gpu.arena("tb_res", tasks.size() * sizeof(TraceResult));
    gpu.upload(d_codes, codes.data(), codes.size());
    gpu.upload(d_tasks, tasks.data(), tasks.size() * sizeof(TraceTask));
    gpu.upload(d_mat, matrix().s, kDim * kDim * sizeof(int64_t));
    // launches in stream order share the direction arena; within one, a
    // workgroup launch per run of hits of one shape
    for (const auto& lr : launches) {
        if (!int32_ok) {
            SSA_HIP(launch_traceback((int)(lr.second - lr.first), gpu.stream(), d_tasks + lr.first, d_codes, d_mat,
                                     cf.gap_open, cf.gap_extend, d_dirs, d_hh, d_ops, d_res + lr.first));
            continue;
        }
        for (size_t k = lr.first; k < lr.second;) {
            size_t e = k;
            while (e < lr.second && shapes[e].rows == shapes[k].rows && shapes[e].waves == shapes[k].waves) ++e;
            SSA_HIP(launch_traceback_wg(shapes[k].rows, shapes[k].waves, (int)(e - k), gpu.stream(), d_tasks + k,
                                        d_codes, d_mat, cf.gap_open, cf.gap_extend, d_dirs, d_hh, d_ops, d_res + k));
            k = e;
        }
    }
    std::vector<TraceResult> res(tasks.size());
    std::vector<int32_t> ops((size_t)ops_total * 2);
    gpu.download(res.data(), d_res, res.size() * sizeof(TraceResult));
    gpu.download(ops.data(), d_ops, ops.size() * sizeof(int32_t));
    gpu.sync();
    *ms += (now_us() - t0) / 1000.0;

    for (size_t k = 0; k < al.size(); ++k) {
        p_alignment a = al[k];
        const TraceResult& r = res[k];
        if (r.status == 2) fatal("libssa_amd: traceback buffer too small");
        std::string cig;
        if (r.status == 1) {
            // no positive cell: the reference's region is undefined here
            print_warning("no local alignment with a positive score for DB sequence %ld", (long)a->db_seq.ID);
            a->align_q_start = a->align_q_end = a->align_d_start = a->align_d_end = 0;
        } else {
            const int32_t* o = ops.data() + tasks[k].ops_off * 2;
            for (int32_t x = r.nops - 1; x >= 0; --x) {
                const int32_t cnt = o[2 * x + 1];
                if (cnt > 1) cig += std::to_string(cnt);
                cig.push_back((char)o[2 * x]);
            }
            a->align_q_start = (size_t)r.a_begin;
            a->align_q_end = (size_t)r.a_end;
            a->align_d_start = (size_t)r.b_begin;
            a->align_d_end = (size_t)r.b_end;
        }
        a->alignment = (char*)malloc(cig.size() + 1);
        memcpy(a->alignment, cig.c_str(), cig.size() + 1);
        a->alignment_len = cig.size();
    }
}

}  // namespace

// ============================================================================
// the search: one or many sw_align / nw_align calls
// ============================================================================
ssa_amd_stats_t& last_stats() {
    static ssa_amd_stats_t s;
    return s;
}

Shape& shape_override_ref() {
    static Shape s{0, 0};
    return s;
}

void set_shape_override(int lanes, int rows) { shape_override_ref() = Shape{lanes, rows}; }

int& cols_override_ref() {
    static int c = 0;  // 0: chosen per search
    return c;
}

void set_cols_override(int cols) { cols_override_ref() = cols; }

int& arith_override_ref() {
    static int a = 0;  // 0: fp16 windows for the SW first tier where they hold; 1: int16 everywhere
    return a;
}

void set_arith_override(int arith) { arith_override_ref() = arith; }

int& placement_override_ref() {
    static int p = 0;  // 0: chosen per search
    return p;
}

void set_placement_override(int placement) { placement_override_ref() = placement; }

bool& task_trace_ref() {
    static bool on = false;
    return on;
}

void set_task_trace(bool on) { task_trace_ref() = on; }

namespace {
std::mutex g_trace_mu;
std::vector<ssa_amd_task_record_t> g_trace;  // the last traced search's (or batch's) tasks
}  // namespace

std::vector<ssa_amd_task_record_t> last_task_trace() {
    std::lock_guard<std::mutex> lk(g_trace_mu);
    return g_trace;
}

// The first tiers' task records of a batch, in ns from its first task start.
static void collect_task_trace(Exec& ex, const std::vector<std::unique_ptr<TierRun>>& runs) {
    std::vector<ssa_amd_task_record_t> recs;
    std::vector<TaskTrace> buf;
    uint64_t t0 = UINT64_MAX;
    for (size_t g = 0; g < runs.size(); ++g) {
        const TierRun& R = *runs[g];
        if (R.empty() || !R.K.trace) continue;
        buf.resize((size_t)R.K.ntasks);
        ex.download(buf.data(), R.K.trace, buf.size() * sizeof(TaskTrace));
        ex.sync();
        for (const TaskTrace& t : buf) {
            recs.push_back(ssa_amd_task_record_t{t.start, t.end, t.wg, t.steps, (int32_t)g, 0});
            if (t.start) t0 = std::min(t0, t.start);
        }
    }
    int dev = 0, khz = 0;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess)
        khz = 0;
    const double ns_per_tick = khz > 0 ? 1e6 / khz : 10.0;
    for (ssa_amd_task_record_t& r : recs) {
        r.start_ns = r.start_ns >= t0 ? (uint64_t)((double)(r.start_ns - t0) * ns_per_tick) : 0;
        r.end_ns = r.end_ns >= t0 ? (uint64_t)((double)(r.end_ns - t0) * ns_per_tick) : 0;
    }
    std::lock_guard<std::mutex> lk(g_trace_mu);
    g_trace.swap(recs);
}

void set_target_waves(int waves) {
    target_waves() = waves > 0 ? waves : 10240;
    release_layouts();
}

namespace {

std::mutex& db_mutex() {
    static std::mutex m;
    return m;
}

// Each chunk's m largest inserts (descending; m = min(k + 1, kChunkTop)) and
// its overflow counts: the order-free first pass of the replay, computed on
// the host from the decoded tier results or on the GPU (chunk_top_kernel).
constexpr size_t kChunkTop = 16;
struct ChunkFirst {
    std::vector<int64_t> top;  // kChunkTop per chunk
    std::vector<uint8_t> ntop;
    std::vector<size_t> ov8, ov16;
};

// per-request state of a batch
struct Job {
    SearchRequest rq;
    std::vector<QueryBuf> qs;
    bool sw_single = false;  // SW 8-bit whose 8-bit results come from one int16 pass
    int first_bits = 0;      // first tier: 8, 16, or 0 (exact tier only)
    HostArray<TierOut> t8, t16;      // per pair (q * N + e)
    HostArray<uint8_t> in16, need64;  // per entry: reaches the 16-bit / exact tier
    // exact scores: the entries that reach the exact tier, ascending, and
    // their scores, one per query each
    std::vector<uint32_t> x_entries;
    std::vector<int64_t> x64;
    ssa_amd_stats_t st;
    Shape shape{0, 0};
    size_t q_off = 0;  // first query of this job in its first-tier group
    int group = -1;
    bool requeue_full = false;  // 16-bit re-run over the whole DB (most entries overflowed at 8 bit)
    bool twin = false;          // both strands: the reverse strand's first tier read from the forward's
    bool skip8 = false;         // NW at 8 bit whose every pair provably overflows in its first block
    size_t nrun = 0;            // query buffers the first tier scores
    bool requeue_sub = false;   // 16-bit re-run over the overflowed entries only
    size_t rq_off = 0;          // first query of this job in its re-run group
    int rgroup = -1;
    std::vector<Elem> hits;
    p_alignment_list list = nullptr;
    size_t st_replayed_chunks = 0;  // chunks whose lane schedules the replay ran
    // device top-k (chunk_top_kernel): the first pass runs on the GPU and the
    // host reads tier results only for the chunks it simulates
    bool dev_top = false;
    int top_mode = 0;  // kTopMode*
    const PairResult* dev16[6] = {};  // device results the 16-bit decode reads, per query buffer
    uint32_t dev_flip = 0;            // bit q: query q reads entry e ^ 1 (strand twins)
    ChunkFirst first;                 // the first pass as the GPU computed it
};

// Decodes the tier results of entries [e0, e1) of a search whose first pass
// ran on the GPU (kTopMode*), as the host path decodes them: res(q, e) is the
// PairResult the 16-bit decode reads for query buffer q and entry e.
template <class Res>
void decode_top_range(Job& J, int mode, const Res& res, const int32_t* nbk, uint32_t N, uint32_t e0, uint32_t e1) {
    const SearchType type = J.rq.type;
    const size_t nq = J.qs.size();
    for (size_t q = 0; q < nq; ++q) {
        TierOut* t16 = J.t16.data() + q * N;
        TierOut* t8 = mode != kTopMode16 ? J.t8.data() + q * N : nullptr;
        for (uint32_t e = e0; e < e1; ++e) {
            const PairResult r = res(q, e);
            if (mode == kTopModeSwSingle) {
                const TierOut o8 = decode(type, kTier16, r, nbk[e], true);
                t8[e] = o8;
                if (!o8.accept()) J.in16[e] = 1;
            } else if (mode == kTopModeAll8) {
                t8[e] = TierOut{-1, 0};
                J.in16[e] = 1;
            }
            const TierOut o = decode(type, kTier16, r, nbk[e], false);
            t16[e] = o;
            if (mode != kTopModeSwSingle && !o.accept()) J.need64[e] = 1;
        }
    }
    if (mode == kTopModeSwSingle)
        for (size_t q = 0; q < nq; ++q)
            for (uint32_t e = e0; e < e1; ++e)
                if (J.in16[e] && !J.t16[q * N + e].accept()) J.need64[e] = 1;
}

// Requests that share the search type, the first tier and the kernel shape
// share one first-tier launch.
struct Group {
    SearchType type;
    int bits;
    Shape sh;
    std::vector<size_t> jobs;
};

void job_queries(Job& J) {
    const Config& cf = config();
    p_query query = J.rq.query;
    // query buffers (src/algo/searcher.c:42-90)
    if (cf.symtype == NUCLEOTIDE) {
        for (int s = 0; s < 2; ++s)
            if (((s + 1) & cf.strands) && query->has_nt[s])
                J.qs.push_back(QueryBuf{query->nt[s].data(), query->nt[s].size(), s, 0});
    } else if (query_is_translated(cf.symtype)) {
        for (int s = 0; s < 2; ++s)
            if ((s + 1) & cf.strands)
                for (int f = 0; f < 3; ++f)
                    J.qs.push_back(QueryBuf{query->aa[3 * s + f].data(), query->aa[3 * s + f].size(), s, f});
    } else {
        J.qs.push_back(QueryBuf{query->aa[0].data(), query->aa[0].size(), 0, 0});
    }
    for (const QueryBuf& q : J.qs)
        if (q.len == 0 && J.rq.type == kNW) fatal("libssa_amd: NW search with an empty query");
}

// The score an exact-tier insert of (q, e) carries: the exact tier's where
// the entry reached it, else the 16-bit tier's (a 64-bit search whose 16-bit
// result is provably exact).
int64_t score64(const Job& J, size_t q, uint32_t e, uint32_t N) {
    if (J.need64[e]) {
        const size_t k = std::lower_bound(J.x_entries.begin(), J.x_entries.end(), e) - J.x_entries.begin();
        return J.x64[k * J.qs.size() + q];
    }
    return J.t16[q * N + e].score;
}

// The host replay of one search: the reference's chunk loop, lane schedules,
// cascade and heap (src/algo/8/search_8.c:94-146, src/algo/16/search_16.c:92-134,
// src/algo/64/search_64.c:44-79, src/algo/manager.c:120-185).
//
// The heap's scores are always the k largest inserted so far, so when chunk c
// starts with a full heap, its root is at least the k-th largest of any k
// inserts of the chunks before, and an insert of chunk c at or below that is
// rejected wherever the lane schedule puts it.  The replay therefore takes a
// threshold per chunk from the earlier chunks' largest inserts (which do not
// depend on order), runs the lane schedules only for chunks with two or more
// inserts above it, and feeds the heap only those inserts, in the reference's
// order.  filter = false simulates every chunk in full (the self-check's
// reference).
// given: the first pass, already computed (then the tier results are read
// only for the chunks the replay simulates); prepare(work) runs before the
// simulation with the chunks it will read, in ascending order.
void job_replay(Job& J, const std::vector<DbEntry>& entries, const std::vector<std::pair<uint32_t, uint32_t>>& chunks,
                uint32_t N, bool threads, bool filter = true, const ChunkFirst* given = nullptr,
                const std::function<void(const std::vector<uint32_t>&)>& prepare = nullptr) {
    const int64_t t_replay = now_us();
    const int W8 = lanes8(), W16 = lanes16();
    const size_t nq = J.qs.size();
    const int bit_width = J.rq.bit_width;
    const size_t k = J.rq.hitcount;
    const size_t nc = chunks.size();
    auto at = [N](size_t q, size_t e) { return q * N + e; };
    auto s64 = [&](size_t q, uint32_t e) { return score64(J, q, e, N); };
    auto over_chunks = [&](size_t n, const std::function<void(size_t)>& fn) {
        if (threads && n > 8) {
            parallel_for(n, fn);
        } else {
            for (size_t i = 0; i < n; ++i) fn(i);
        }
    };
    struct Ins {
        uint32_t e;
        uint8_t q;
        int64_t score;
    };

    // Every insert of chunk c in no particular order, as f(q, e, score,
    // copies): an entry that overflows for m8 queries at 8 bit stands m8 times
    // in the chunk's 16-bit list, which every query scans; each 16-bit
    // overflow met in that scan adds the entry to the exact list once more,
    // which every query scans again.  Also counts the overflows.
    const TierOut* p8[6] = {};
    const TierOut* p16[6] = {};
    for (size_t q = 0; q < nq && q < 6; ++q) {
        if (!J.t8.empty()) p8[q] = J.t8.data() + at(q, 0);
        if (!J.t16.empty()) p16[q] = J.t16.data() + at(q, 0);
    }
    auto inserts = [&](size_t c, size_t& ov8, size_t& ov16, auto&& f) {
        for (uint32_t e = chunks[c].first; e < chunks[c].second; ++e) {
            if (bit_width == BIT_WIDTH_64) {
                for (size_t q = 0; q < nq; ++q) f(q, e, s64(q, e), (size_t)1);
                continue;
            }
            size_t m8 = 1;
            if (bit_width == BIT_WIDTH_8) {
                m8 = 0;
                for (size_t q = 0; q < nq; ++q) {
                    const TierOut o = p8[q][e];
                    if (o.accept()) f(q, e, (int64_t)o.score, (size_t)1);
                    else ++m8;
                }
                ov8 += m8;
                if (m8 == 0) continue;
            }
            size_t m16 = 0;
            for (size_t q = 0; q < nq; ++q) {
                const TierOut o = p16[q][e];
                if (o.accept()) f(q, e, (int64_t)o.score, m8);
                else ++m16;
            }
            ov16 += m8 * m16;
            if (m16)
                for (size_t q = 0; q < nq; ++q) f(q, e, s64(q, e), m8 * m16);
        }
    };

    // each chunk's m largest inserts (descending) and its overflow counts
    constexpr size_t kTop = kChunkTop;
    const size_t m = std::min<size_t>(k + 1, kTop);
    ChunkFirst own;
    if (!given) {
        own.top.assign(nc * kTop, 0);
        own.ntop.assign(nc, 0);
        own.ov8.assign(nc, 0);
        own.ov16.assign(nc, 0);
        over_chunks(nc, [&](size_t c) {
            int64_t tv[kTop];
            size_t tn = 0, n8 = 0, n16 = 0;
            inserts(c, n8, n16, [&](size_t, uint32_t, int64_t score, size_t copies) {
                for (size_t r = 0; r < copies; ++r) {
                    if (tn == m && !(score > tv[m - 1])) break;
                    size_t i = tn < m ? tn++ : m - 1;
                    while (i > 0 && tv[i - 1] < score) {
                        tv[i] = tv[i - 1];
                        --i;
                    }
                    tv[i] = score;
                }
            });
            std::copy(tv, tv + tn, own.top.begin() + c * kTop);
            own.ntop[c] = (uint8_t)tn;
            own.ov8[c] = n8;
            own.ov16[c] = n16;
        });
    }
    const ChunkFirst& F = given ? *given : own;
    const std::vector<int64_t>& top = F.top;
    const std::vector<uint8_t>& ntop = F.ntop;
    const std::vector<size_t>&ov8 = F.ov8, &ov16 = F.ov16;
    for (size_t c = 0; c < nc; ++c) {
        J.st.overflow_8_bit_count += ov8[c];
        J.st.overflow_16_bit_count += ov16[c];
    }
    if (k == 0) {  // a heap of no elements keeps nothing
        J.st.ms_replay = (now_us() - t_replay) / 1000.0;
        return;
    }

    // Thresholds: the k-th largest of the chunks' m largest inserts before
    // each chunk, a lower bound of the heap's root there (exact for m = k+1).
    // above[c] counts the chunk's inserts over it, exactly while below m.
    std::vector<uint8_t> full(nc, 0);
    std::vector<int64_t> thr(nc, 0);
    std::vector<size_t> above(nc, 0);
    {
        std::priority_queue<int64_t, std::vector<int64_t>, std::greater<int64_t>> best;
        for (size_t c = 0; c < nc; ++c) {
            const int64_t* tv = &top[c * kTop];
            full[c] = filter && best.size() >= k;
            if (full[c]) {
                thr[c] = best.top();
                for (size_t i = 0; i < ntop[c]; ++i) above[c] += tv[i] > thr[c];
            }
            for (size_t i = 0; i < ntop[c]; ++i) {
                if (best.size() < k) {
                    best.push(tv[i]);
                } else if (best.top() < tv[i]) {
                    best.pop();
                    best.push(tv[i]);
                } else {
                    break;  // descending: the rest are no larger
                }
            }
        }
    }

    // the inserts that can change the heap, chunk by chunk in the reference's order
    std::vector<std::vector<Ins>> per_chunk(nc);
    std::vector<uint32_t> work;
    for (size_t c = 0; c < nc; ++c)
        if (!full[c] || above[c] > 0) work.push_back((uint32_t)c);
    J.st_replayed_chunks = work.size();
    if (prepare) prepare(work);
    over_chunks(work.size(), [&](size_t w) {
        const size_t c = work[w];
        std::vector<Ins>& out = per_chunk[c];
        const bool filter = full[c];
        const int64_t t = thr[c];
        auto keep = [&](int64_t score) { return !filter || score > t; };
        if (filter && above[c] == 1) {
            // one insert above the threshold (m >= 2 here, so the count is
            // exact): its place in the order does not matter
            size_t d8 = 0, d16 = 0;
            inserts(c, d8, d16, [&](size_t q, uint32_t e, int64_t score, size_t) {
                if (score > t) out.push_back(Ins{e, (uint8_t)q, score});
            });
            return;
        }
        const uint32_t e0 = chunks[c].first, e1 = chunks[c].second;
        std::vector<uint32_t> items(e1 - e0), ovf, ovf2;
        std::iota(items.begin(), items.end(), e0);
        auto tier64 = [&](const std::vector<uint32_t>& list) {
            for (size_t q = 0; q < nq; ++q)
                for (uint32_t e : list) {
                    const int64_t score = s64(q, e);
                    if (keep(score)) out.push_back(Ins{e, (uint8_t)q, score});
                }
        };
        auto tier16 = [&](const std::vector<uint32_t>& list) {
            ovf2.clear();
            for (size_t q = 0; q < nq; ++q)
                lane_schedule(
                    W16, list.size(), [&](size_t i) { return (int64_t)J.t16[at(q, list[i])].dur(); },
                    [&](size_t i) {
                        const TierOut o = J.t16[at(q, list[i])];
                        if (!o.accept()) ovf2.push_back(list[i]);
                        else if (keep(o.score)) out.push_back(Ins{list[i], (uint8_t)q, o.score});
                    });
            if (!ovf2.empty()) tier64(ovf2);
        };
        if (bit_width == BIT_WIDTH_8) {
            for (size_t q = 0; q < nq; ++q)
                lane_schedule(
                    W8, items.size(), [&](size_t i) { return (int64_t)J.t8[at(q, items[i])].dur(); },
                    [&](size_t i) {
                        const TierOut o = J.t8[at(q, items[i])];
                        if (!o.accept()) ovf.push_back(items[i]);
                        else if (keep(o.score)) out.push_back(Ins{items[i], (uint8_t)q, o.score});
                    });
            if (!ovf.empty()) tier16(ovf);
        } else if (bit_width == BIT_WIDTH_16) {
            tier16(items);
        } else {
            tier64(items);
        }
    });

    MinHeap heap(k);
    for (size_t c = 0; c < nc; ++c)
        for (const Ins& x : per_chunk[c]) {
            if (heap.rejects(x.score)) continue;
            const DbEntry& e = entries[x.e];
            heap.add(Elem{e.id, e.frame, e.strand, x.q, x.score});
        }
    J.hits = heap.sorted();
    J.st.replayed_chunks = J.st_replayed_chunks;
    J.st.ms_replay = (now_us() - t_replay) / 1000.0;
}

// the alignment list of one search (src/algo/manager.c:83-113, src/algo/aligner.c:62-99)
void job_list(Job& J, std::vector<p_alignment>& al_out, std::vector<SearchType>& types_out) {
    const std::vector<Elem>& hits = J.hits;
    p_alignment_list alist = (p_alignment_list)malloc(sizeof(alignment_list_t));
    alist->len = hits.size();
    alist->alignments = (p_alignment*)malloc(std::max<size_t>(hits.size(), 1) * sizeof(p_alignment));
    for (size_t k = 0; k < hits.size(); ++k) {
        const Elem& h = hits[k];
        p_seqinfo info = ssa_db_get_sequence(h.db_id);
        if (!info) fatal("Could not get sequence from DB: %ld", (long)h.db_id);
        std::vector<uint8_t> dseq = prepared_db_sequence(info, h.db_strand, h.db_frame);
        p_alignment a = (p_alignment)calloc(1, sizeof(alignment_t));
        a->db_seq.seq = (char*)malloc(dseq.size() + 1);
        memcpy(a->db_seq.seq, dseq.data(), dseq.size());
        a->db_seq.seq[dseq.size()] = 0;
        a->db_seq.len = dseq.size();
        a->db_seq.strand = h.db_frame;  // the reference stores these two swapped (src/algo/aligner.c:76-77)
        a->db_seq.frame = h.db_strand;
        a->db_seq.ID = info->ID;
        const QueryBuf& q = J.qs[h.query_id];
        a->query.seq = (char*)q.seq;
        a->query.len = q.len;
        a->query.strand = q.strand;
        a->query.frame = q.frame;
        a->score = h.score;
        alist->alignments[k] = a;
        if (J.rq.align_type != COMPUTE_SCORE) {
            al_out.push_back(a);
            types_out.push_back(J.rq.type);
        }
    }
    J.list = alist;
}

}  // namespace

// The chunk view of the loaded DB for one chunk size, cached per device DB
// build (a new DB, or new symbol type, strands or genetic code, rebuilds the
// entries); the last few chunk sizes are kept, and entries are never freed
// while a search may hold a reference, only when the DB is rebuilt.
const DbChunks& db_chunks(const DeviceDB& db, size_t cs) {
    static std::mutex mu;
    static std::list<DbChunks> cache;
    std::lock_guard<std::mutex> lk(mu);
    for (auto it = cache.begin(); it != cache.end();) {
        if (it->build == db.build && it->chunk_size == cs) return *it;
        // a DB that was rebuilt cannot be in use by a search still running
        // (the DB is rebuilt only between searches)
        if (it->build != db.build) it = cache.erase(it);
        else ++it;
    }
    cache.emplace_back();
    DbChunks& d = cache.back();
    d.build = db.build;
    d.chunk_size = cs;
    for (size_t start = 0; start < db.count; start += cs) {
        const size_t stop = std::min(db.count, start + cs);
        const uint32_t e0 = db.first_entry[start], e1 = db.first_entry[stop];
        if (e0 == e1) break;
        d.chunks.push_back({e0, e1});
    }
    d.N = d.chunks.empty() ? 0 : d.chunks.back().second;
    d.all.resize(d.N);
    std::iota(d.all.begin(), d.all.end(), 0u);
    d.nb.resize(d.N);
    for (uint32_t e = 0; e < d.N; ++e) {
        const int64_t nb = entry_blocks(db.entries[e].len);
        d.nb[e] = (int32_t)nb;
        d.total_blocks += nb;
        d.max_blocks = std::max(d.max_blocks, nb);
        d.total_residues += db.entries[e].len;
    }
    d.twins = db.symtype == NUCLEOTIDE && d.N % 2 == 0 && d.N > 0;
    for (uint32_t e = 0; d.twins && e < d.N; e += 2) {
        const DbEntry &f = db.entries[e], &r = db.entries[e + 1];
        d.twins = f.strand == 0 && r.strand == 1 && f.id == r.id && f.len == r.len;
    }
    return d;
}

// NW at 8 bit flags every pair in its first block once the query is long
// enough: down column 0 the left border and the vertical gap from row 0 fall
// by a gap extension per row, and soon a cell lies below the overflow
// threshold (I_MIN - gapO - gapE - 1, src/algo/simd/search_simd_nw.c:372).
// The score kernel's column 0 is one chain of saturating max-plus steps down
// the rows (the left border H and E staged per row, the top border's F
// entering row 0, the diagonal from the row above's left border), monotone in
// the profile values; run on the host with each row's best score over every
// DB code, it bounds every entry's column 0 from above.  If the bound falls
// below the threshold on some real row, every pair's first block is flagged
// (a lane spends one block on it), which is all the 8-bit tier yields for an
// overflowed pair, so the search records that and skips the 8-bit launch.
bool nw8_first_block_flags(const QueryBuf& q) {
    const Config& cf = config();
    const Matrix& M = matrix();
    const Tier& T = kTier8;
    const int go = cf.gap_open, ge = cf.gap_extend;
    const int64_t QR = T.wrap(go + ge), RR = T.wrap(ge);
    const int64_t thr = T.wrap(T.imin() - go - ge - 1);
    int64_t f = T.sat(T.wrap(go + ge) + QR);  // F entering row 0: the top border's, plus open
    int64_t c = QR;                           // left border H of row i
    int64_t hprev = 0;                        // left border H of row i - 1 (the corner for row 0)
    for (size_t i = 0; i < q.len; ++i) {
        int64_t v = INT64_MIN;
        for (int d = 0; d < kDim; ++d) v = std::max(v, T.wrap(M.at(d, q.seq[i])));
        int64_t h = std::max(T.sat(hprev + v), f);
        h = std::max(h, T.sat(c + QR));  // E entering column 0: the left border's, plus open
        if (h < thr) return true;
        f = std::max(T.sat(f + RR), T.sat(h + QR));
        hprev = c;
        c = T.sat(c + RR);
    }
    return false;
}

// A both-strand nucleotide search whose reverse query strand need not be
// scored: the alignment score is unchanged when both sequences are reversed,
// and, with a matrix that scores complements alike, when both are
// complemented, so (reverse query, entry) scores as (forward query, the
// entry's reverse-complement twin).  The replay also uses each pair's
// overflow block, which differs between the two directions, so this holds
// only where no int16 value of any pair can reach either end of the range
// (the bound the 64-bit shortcut uses): then no pair is flagged and the
// 16-bit results are exact.  8-bit searches keep both strands (their flags
// are the point).
bool strand_twins(const Job& J, const DbChunks& dc, const MatrixInfo& mi) {
    const Config& cf = config();
    if (!dc.twins || cf.symtype != NUCLEOTIDE || J.qs.size() != 2 || J.first_bits != 16 ||
        J.rq.bit_width == BIT_WIDTH_8)
        return false;
    const QueryBuf &f = J.qs[0], &r = J.qs[1];
    if (f.strand != 0 || r.strand != 1 || f.len != r.len || f.len == 0) return false;
    for (size_t i = 0; i < f.len; ++i)
        if (r.seq[i] != (uint8_t)complement_nt(f.seq[f.len - 1 - i])) return false;
    const Matrix& M = matrix();
    for (int a = 0; a < 16; ++a)
        for (int b = 0; b < 16; ++b)
            if (M.at(complement_nt(a), complement_nt(b)) != M.at(a, b)) return false;
    return int16_exact(f.len, 4 * dc.max_blocks, mi);
}

// every entry of the loaded DB, 0 .. entries - 1
const std::vector<uint32_t>& db_all_entries(const DeviceDB& db) {
    static std::mutex mu;
    static uint64_t build = ~0ull;
    static std::vector<uint32_t> all;
    std::lock_guard<std::mutex> lk(mu);
    if (build != db.build || all.size() != db.entries.size()) {
        all.resize(db.entries.size());
        std::iota(all.begin(), all.end(), 0u);
        build = db.build;
    }
    return all;
}

void run_batch(Exec& ex, const std::vector<SearchRequest>& reqs, std::vector<SearchResult>& results,
               bool one_graph) {
    const int64_t t_start = now_us();
    const Config& cf = config();
    results.assign(reqs.size(), SearchResult{});
    if (reqs.empty()) return;

    std::vector<Job> jobs(reqs.size());
    for (size_t j = 0; j < reqs.size(); ++j) {
        Job& J = jobs[j];
        J.rq = reqs[j];
        memset(&J.st, 0, sizeof J.st);
        const int bw = J.rq.bit_width;
        if (bw != BIT_WIDTH_8 && bw != BIT_WIDTH_16 && bw != BIT_WIDTH_64)
            fatal("\nunknown bit width provided: %d\n\n", bw);  // src/algo/searcher.c:115-117
        if (bw == BIT_WIDTH_8 && cf.simd_mode < COMPUTE_ON_SSE41)
            fatal("libssa_amd: 8-bit searches replay SSE4.1 or AVX2 lanes; COMPUTE_ON_SSE2 selects neither");
        job_queries(J);
    }

    // the no-allocation path (a context with a fixed region) builds no device
    // state: the DB and a stream layout must be resident
    const bool fixed = ex.fixed();
    DeviceDB* dbp;
    {
        std::lock_guard<std::mutex> lk(db_mutex());
        dbp = &device_db(!fixed);
    }
    DeviceDB& db = *dbp;

    const DbChunks& dc = db_chunks(db, cf.chunk_size ? cf.chunk_size : kDefaultChunkSize);
    const int32_t* nbk = dc.nb.data();  // blocks per entry
    const std::vector<std::pair<uint32_t, uint32_t>>& chunks = dc.chunks;  // entry ranges
    const uint32_t N = dc.N;
    auto at = [N](size_t q, size_t e) { return q * N + e; };
    PairSet all;
    all.ext = &dc.all;
    bool full_db = N == db.entries.size();
    if (fixed && !full_db) {
        // the chunks stop before the last entry (a chunk of empty records):
        // score every entry on the cached whole-DB layout and read the first N
        all.ext = &db_all_entries(db);
        full_db = true;
    }
    const int64_t total_blocks = dc.total_blocks, max_blocks = dc.max_blocks;
    const uint64_t total_residues = dc.total_residues;

    const MatrixInfo mi = matrix_info();
    const int go = cf.gap_open, ge = cf.gap_extend;

    // ---- first tiers: grouped by (type, tier, shape), one launch per group ----
    std::vector<Group> groups;
    for (size_t j = 0; j < jobs.size(); ++j) {
        Job& J = jobs[j];
        J.st.chunk_count = chunks.size();
        J.st.seq_count = N;
        for (const QueryBuf& q : J.qs) J.st.cells += (uint64_t)q.len * total_residues;
        const size_t nq = J.qs.size();
        J.in16.reset(N);
        J.in16.fill_zero();
        J.need64.reset(N);
        J.need64.fill_zero();
        J.sw_single = J.rq.type == kSW && mi.fits8 && (go + ge) >= -128;
        if (J.rq.bit_width == BIT_WIDTH_8) J.first_bits = J.sw_single ? 16 : 8;
        else if (J.rq.bit_width == BIT_WIDTH_16) J.first_bits = 16;
        else J.first_bits = mi.fits16 ? 16 : 0;
        J.twin = strand_twins(J, dc, mi);
        J.nrun = J.twin ? 1 : nq;
        J.skip8 = J.rq.type == kNW && J.first_bits == 8 && nq > 0 &&
                  std::all_of(J.qs.begin(), J.qs.end(), [](const QueryBuf& q) { return nw8_first_block_flags(q); });
        if (J.first_bits != 0 && !J.skip8)
            for (size_t q = 0; q < J.nrun; ++q) J.st.cells_scored += (uint64_t)J.qs[q].len * total_residues;
        // the replay's first pass on the GPU (chunk_top_kernel) where every
        // insert follows from one 16-bit result per pair: 16-bit searches, SW
        // 8-bit searches from one int16 pass, NW 8-bit searches whose every
        // pair overflows in its first block (the 16-bit re-run's results)
        if (!fixed && J.rq.hitcount > 0 && N > 0 && nq > 0 && nq <= 6) {
            if (J.rq.bit_width == BIT_WIDTH_16) J.dev_top = true, J.top_mode = kTopMode16;
            else if (J.rq.bit_width == BIT_WIDTH_8 && J.sw_single) J.dev_top = true, J.top_mode = kTopModeSwSingle;
            else if (J.rq.bit_width == BIT_WIDTH_8 && J.skip8) J.dev_top = true, J.top_mode = kTopModeAll8;
        }
        if (J.first_bits == 0 || N == 0 || nq == 0 || J.skip8) continue;
        size_t qmax = 0;
        for (const QueryBuf& q : J.qs) qmax = std::max(qmax, q.len);
        // (SW with several query buffers runs packed: the group's choice below
        // replaces this one)
        J.shape = choose_shape(qmax, total_blocks, max_blocks, J.rq.type == kSW ? 0 : 1,
                               J.rq.type == kSW && J.nrun > 1, fixed);
        // SW searches of one tier share one launch as a packed query whatever
        // their lengths; NW searches share a launch per kernel shape
        int gi = -1;
        for (size_t g = 0; g < groups.size(); ++g)
            if (groups[g].type == J.rq.type && groups[g].bits == J.first_bits &&
                (J.rq.type == kSW || (groups[g].sh.lanes == J.shape.lanes && groups[g].sh.rows == J.shape.rows)))
                gi = (int)g;
        if (gi < 0) {
            groups.push_back(Group{J.rq.type, J.first_bits, J.shape, {}});
            gi = (int)groups.size() - 1;
        }
        J.group = gi;
        groups[gi].jobs.push_back(j);
    }
    for (Group& G : groups) {
        if (G.type != kSW) continue;
        size_t nqg = 0, qsum = 0;
        for (size_t j : G.jobs)
            for (size_t q = 0; q < jobs[j].nrun; ++q) ++nqg, qsum += jobs[j].qs[q].len;
        if (nqg > 1) G.sh = choose_shape(qsum, total_blocks, max_blocks, 0, true, fixed);
        for (size_t j : G.jobs) jobs[j].shape = G.sh;
    }
    std::vector<std::unique_ptr<TierRun>> runs(groups.size());
    std::vector<std::unique_ptr<PairSet>> run_sets(groups.size());
    for (size_t g = 0; g < groups.size(); ++g) {
        runs[g].reset(new TierRun());
        TierRun& R = *runs[g];
        R.type = groups[g].type;
        R.T = groups[g].bits == 8 ? kTier8 : kTier16;
        R.sh = groups[g].sh;
        R.pfx = "g" + std::to_string(g) + "_";
        R.sorted_all = full_db ? &db.by_length : nullptr;
        for (size_t j : groups[g].jobs) {
            jobs[j].q_off = R.qs.size();
            R.qs.insert(R.qs.end(), jobs[j].qs.begin(), jobs[j].qs.begin() + jobs[j].nrun);
        }
        run_sets[g].reset(new PairSet(all));
        run_sets[g]->nq = R.qs.size();
        R.ps = run_sets[g].get();
        tier_prepare(ex, R);
    }
    // the first tiers' results the host decodes: every search's but those
    // whose replay takes its first pass from the GPU
    auto download_first = [&](size_t g) {
        TierRun& R = *runs[g];
        if (R.empty()) return;
        for (size_t j : groups[g].jobs) {
            Job& J = jobs[j];
            if (J.dev_top) continue;
            const size_t a = R.ps->index(J.q_off, 0), b = R.ps->index(J.q_off + J.nrun, 0);
            ex.download(R.out + a, R.d_out + a, (b - a) * sizeof(PairResult));
            J.st.d2h_bytes += (b - a) * sizeof(PairResult);
        }
    };
    const double ms_setup = (now_us() - t_start) / 1000.0;
    double ms_first = 0, ms_graph = 0;
    bool graph_reused = false;
    if (!groups.empty()) {
        GpuTimer tm(ex.stream());
        if (one_graph) {
            // every first-tier launch of the batch in one graph; groups run on
            // parallel branches
            hipGraph_t graph;
            hipGraphExec_t gexec;
            const hipStream_t st = ex.stream();
            const int64_t t_graph = now_us();
            SSA_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            hipEvent_t fork;
            SSA_HIP(hipEventCreateWithFlags(&fork, hipEventDisableTiming));
            SSA_HIP(hipEventRecord(fork, st));
            std::vector<hipEvent_t> joins(groups.size());
            for (size_t g = 0; g < groups.size(); ++g) {
                const hipStream_t br = ex.aux(g);
                SSA_HIP(hipStreamWaitEvent(br, fork, 0));
                tier_launch(br, *runs[g]);
                SSA_HIP(hipEventCreateWithFlags(&joins[g], hipEventDisableTiming));
                SSA_HIP(hipEventRecord(joins[g], br));
                SSA_HIP(hipStreamWaitEvent(st, joins[g], 0));
            }
            SSA_HIP(hipStreamEndCapture(st, &graph));
            // the context's last instantiated graph, updated to this batch's
            // kernel arguments when the topology is the same, else a new one
            gexec = ex.graph_exec();
            graph_reused = false;
            if (gexec) {
                hipGraphNode_t bad = nullptr;
                hipGraphExecUpdateResult ur;
                graph_reused = hipGraphExecUpdate(gexec, graph, &bad, &ur) == hipSuccess && ur == hipGraphExecUpdateSuccess;
                if (!graph_reused) {
                    (void)hipGetLastError();
                    SSA_HIP(hipGraphExecDestroy(gexec));
                    ex.set_graph_exec(nullptr);
                }
            }
            if (!graph_reused) {
                SSA_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
                ex.set_graph_exec(gexec);
            }
            ms_graph = (now_us() - t_graph) / 1000.0;
            tm.start();
            SSA_HIP(hipGraphLaunch(gexec, st));
            tm.stop();
            for (size_t g = 0; g < groups.size(); ++g) download_first(g);
            ex.sync();
            ms_first = tm.ms();
            SSA_HIP(hipGraphDestroy(graph));
            SSA_HIP(hipEventDestroy(fork));
            for (hipEvent_t e : joins) SSA_HIP(hipEventDestroy(e));
        } else {
            tm.start();
            for (size_t g = 0; g < groups.size(); ++g) tier_launch(ex.stream(), *runs[g]);
            tm.stop();
            for (size_t g = 0; g < groups.size(); ++g) download_first(g);
            ex.sync();
            ms_first = tm.ms();
        }
    }
    if (task_trace_ref()) collect_task_trace(ex, runs);
    ex.recycle();  // the first tiers' results are on the host

    // ---- decode the first tiers; re-queued and exact tiers per search ----
    std::vector<double> ms_re_job(jobs.size(), 0.0);
    std::vector<uint64_t> computed_job(jobs.size(), 0);
    const size_t nthreads = parallel_width();
    // host work per search: searches spread over threads, or, for fewer
    // searches than threads, each search's entries spread over threads
    const bool inner = jobs.size() < nthreads;
    auto over_entries = [&](const std::function<void(uint32_t, uint32_t)>& fn) {
        if (!inner || N < 65536) {
            fn(0, N);
            return;
        }
        parallel_for(nthreads, [&](size_t t) {
            fn((uint32_t)((uint64_t)N * t / nthreads), (uint32_t)((uint64_t)N * (t + 1) / nthreads));
        });
    };
    auto parallel_jobs = [&](const std::function<void(size_t)>& fn) {
        if (inner) {
            for (size_t j = 0; j < jobs.size(); ++j) fn(j);
            return;
        }
        parallel_for(jobs.size(), fn);
    };
    parallel_jobs([&](size_t j) {
        Job& J = jobs[j];
        const SearchType type = J.rq.type;
        const size_t nq = J.qs.size();
        J.st.ms_gpu_first = J.group >= 0 ? ms_first : 0.0;
        J.st.ms_setup = ms_setup;
        J.st.ms_graph = ms_graph;
        J.st.graph_reused = graph_reused ? 1 : 0;
        uint64_t& computed = computed_job[j];
        if (J.group >= 0) {
            TierRun& R = *runs[J.group];
            computed += R.computed * J.nrun / std::max<size_t>(R.qs.size(), 1);
            J.st.lanes_per_stream = R.sh.lanes;
            J.st.rows_per_lane = R.sh.rows;
            J.st.columns_per_step = R.sh.cols;
            J.st.profile_placement = R.vp ? 2 : 1;
            J.st.overflow_tracking = !R.untracked;
            J.st.cell_arithmetic = R.f16 ? 2 : 1;
        }
        if (J.dev_top) {
            // decoded later, only for the chunks the replay simulates; here the
            // device results each query buffer's 16-bit decode reads
            if (J.top_mode != kTopModeAll8) {
                const TierRun& R = *runs[J.group];
                for (size_t q = 0; q < nq; ++q) {
                    const bool fl = J.twin && q == 1;
                    J.dev16[q] = R.d_out + R.ps->index(J.q_off + (fl ? 0 : q), 0);
                    if (fl) J.dev_flip |= 1u << q;
                }
            } else {
                J.requeue_full = true;  // the 16-bit pass the cascade runs next, over the whole DB
            }
            if (J.rq.bit_width == BIT_WIDTH_8) J.t8.reset(nq * N);
            J.t16.reset(nq * N);
            return;
        }
        // the first tier's results of query buffer q: entry e at base[e ^ flip]
        // (a twin search reads its reverse strand from the forward strand's
        // reverse-complement twins)
        auto src = [&](size_t q, uint32_t& flip) -> const PairResult* {
            const TierRun& R = *runs[J.group];
            flip = (J.twin && q == 1) ? 1u : 0u;
            return R.out + R.ps->index(J.q_off + (flip ? 0 : q), 0);
        };
        const int bw = J.rq.bit_width;
        uint8_t* in16 = J.in16.data();
        uint8_t* need64 = J.need64.data();
        if (bw == BIT_WIDTH_8) {
            J.t8.reset(nq * N);
            if (J.sw_single) {
                J.t16.reset(nq * N);
                over_entries([&](uint32_t e0, uint32_t e1) {
                    if (e1 <= e0) return;
                    for (size_t q = 0; q < nq; ++q) {
                        uint32_t fl;
                        const PairResult* base = src(q, fl);
                        TierOut* t8 = J.t8.data() + at(q, 0);
                        TierOut* t16 = J.t16.data() + at(q, 0);
                        for (uint32_t e = e0; e < e1; ++e) {
                            const PairResult r = base[e ^ fl];
                            const TierOut o8 = decode(type, kTier16, r, nbk[e], true);
                            t8[e] = o8;
                            t16[e] = decode(type, kTier16, r, nbk[e], false);
                            if (!o8.accept()) in16[e] = 1;
                        }
                    }
                    // an entry reaches the exact tier when some query overflows
                    // both tiers
                    for (size_t q = 0; q < nq; ++q)
                        for (uint32_t e = e0; e < e1; ++e)
                            if (J.in16[e] && !J.t16[at(q, e)].accept()) J.need64[e] = 1;
                });
            } else if (J.skip8) {
                // every pair overflows in its first block (nw8_first_block_flags)
                TierOut* t8 = J.t8.data();
                for (size_t i = 0; i < nq * N; ++i) t8[i] = TierOut{-1, 0};
                if (N) memset(in16, 1, N);
                J.t16.reset(nq * N);
                J.requeue_full = N > 0;
            } else {
                size_t n16 = 0;
                std::mutex mu;
                over_entries([&](uint32_t e0, uint32_t e1) {
                    if (e1 <= e0) return;
                    for (size_t q = 0; q < nq; ++q) {
                        uint32_t fl;
                        const PairResult* base = src(q, fl);
                        TierOut* t8 = J.t8.data() + at(q, 0);
                        for (uint32_t e = e0; e < e1; ++e) {
                            const TierOut o8 = decode(type, kTier8, base[e ^ fl], nbk[e], false);
                            t8[e] = o8;
                            if (!o8.accept()) in16[e] = 1;
                        }
                    }
                    size_t n = 0;
                    for (uint32_t e = e0; e < e1; ++e) n += J.in16[e];
                    std::lock_guard<std::mutex> lk(mu);
                    n16 += n;
                });
                J.t16.reset(nq * N);
                if ((n16 * 2 >= N || (fixed && n16 > 0)) && N > 0) {
                    // most entries overflowed (or no layout of a subset may be
                    // built): re-run the whole DB on its cached streams
                    // together with the batch's other such searches
                    J.requeue_full = true;
                } else if (n16 > 0) {
                    J.requeue_sub = true;
                }
            }
        } else if (bw == BIT_WIDTH_16) {
            J.t16.reset(nq * N);
            if (N) memset(J.in16.data(), 1, N);
            over_entries([&](uint32_t e0, uint32_t e1) {
                if (e1 <= e0) return;
                for (size_t q = 0; q < nq; ++q) {
                    uint32_t fl;
                    const PairResult* base = src(q, fl);
                    TierOut* t16 = J.t16.data() + at(q, 0);
                    for (uint32_t e = e0; e < e1; ++e) {
                        const TierOut o = decode(type, kTier16, base[e ^ fl], nbk[e], false);
                        t16[e] = o;
                        if (!o.accept()) need64[e] = 1;
                    }
                }
            });
        }
    });

    // ---- 16-bit re-runs over the 8-bit overflows of single searches ----
    for (size_t j = 0; j < jobs.size(); ++j) {
        Job& J = jobs[j];
        if (!J.requeue_sub) continue;
        const size_t nq = J.qs.size();
        PairSet sub;
        sub.nq = nq;
        for (uint32_t e = 0; e < N; ++e)
            if (J.in16[e]) sub.entries.push_back(e);
        TierRun R2;
        R2.type = J.rq.type;
        R2.T = kTier16;
        R2.qs = J.qs;
        R2.ps = &sub;
        R2.pfx = "re_";
        ex.recycle();
        gpu_score_tier(ex, R2, &ms_re_job[j]);
        computed_job[j] += R2.computed;
        J.st.d2h_bytes += R2.nout * sizeof(PairResult);
        for (size_t q = 0; q < nq; ++q)
            for (size_t k = 0; k < sub.entries.size(); ++k) {
                const uint32_t e = sub.entries[k];
                const TierOut o = decode(J.rq.type, kTier16, R2.out[sub.index(q, k)], nbk[e], false);
                J.t16[at(q, e)] = o;
                if (!o.accept()) J.need64[e] = 1;
            }
    }

    // ---- 16-bit re-runs over the whole DB, grouped like the first tiers ----
    {
        std::vector<Group> rgroups;
        for (size_t j = 0; j < jobs.size(); ++j) {
            Job& J = jobs[j];
            if (!J.requeue_full) continue;
            size_t qmax = 0;
            for (const QueryBuf& q : J.qs) qmax = std::max(qmax, q.len);
            const Shape sh = choose_shape(qmax, total_blocks, max_blocks, J.rq.type == kSW ? 0 : 1,
                                          J.rq.type == kSW && J.qs.size() > 1, fixed);
            int gi = -1;
            for (size_t g = 0; g < rgroups.size(); ++g)
                if (rgroups[g].type == J.rq.type &&
                    (J.rq.type == kSW || (rgroups[g].sh.lanes == sh.lanes && rgroups[g].sh.rows == sh.rows)))
                    gi = (int)g;
            if (gi < 0) {
                rgroups.push_back(Group{J.rq.type, 16, sh, {}});
                gi = (int)rgroups.size() - 1;
            }
            J.rgroup = gi;
            rgroups[gi].jobs.push_back(j);
        }
        for (Group& G : rgroups) {
            if (G.type != kSW) continue;
            size_t nqg = 0, qsum = 0;
            for (size_t j : G.jobs)
                for (const QueryBuf& q : jobs[j].qs) ++nqg, qsum += q.len;
            if (nqg > 1) G.sh = choose_shape(qsum, total_blocks, max_blocks, 0, true, fixed);
        }
        for (size_t g = 0; g < rgroups.size(); ++g) {
            TierRun R;
            R.type = rgroups[g].type;
            R.T = kTier16;
            R.sh = rgroups[g].sh;
            // a group's own arenas: the device top-k reads them after every group ran
            R.pfx = "rg" + std::to_string(g) + "_";
            R.sorted_all = full_db ? &db.by_length : nullptr;
            for (size_t j : rgroups[g].jobs) {
                jobs[j].rq_off = R.qs.size();
                R.qs.insert(R.qs.end(), jobs[j].qs.begin(), jobs[j].qs.end());
            }
            PairSet ps(all);
            ps.nq = R.qs.size();
            R.ps = &ps;
            double ms = 0;
            ex.recycle();
            tier_prepare(ex, R);
            {
                GpuTimer tm(ex.stream());
                tm.start();
                tier_launch(ex.stream(), R);
                tm.stop();
                if (!R.empty())
                    for (size_t j : rgroups[g].jobs) {
                        Job& J = jobs[j];
                        if (J.dev_top) continue;
                        const size_t a = ps.index(J.rq_off, 0), b = ps.index(J.rq_off + J.qs.size(), 0);
                        ex.download(R.out + a, R.d_out + a, (b - a) * sizeof(PairResult));
                        J.st.d2h_bytes += (b - a) * sizeof(PairResult);
                    }
                ex.sync();
                ms += tm.ms();
            }
            for (size_t j : rgroups[g].jobs) {
                Job& J = jobs[j];
                const size_t nq = J.qs.size();
                ms_re_job[j] += ms;
                computed_job[j] += R.computed * nq / std::max<size_t>(R.qs.size(), 1);
                if (J.dev_top) {
                    for (size_t q = 0; q < nq; ++q) J.dev16[q] = R.d_out + ps.index(J.rq_off + q, 0);
                    J.dev_flip = 0;
                    continue;
                }
                for (size_t q = 0; q < nq; ++q)
                    for (uint32_t e = 0; e < N; ++e)
                        if (J.in16[e]) {
                            const TierOut o =
                                decode(J.rq.type, kTier16, R.out[ps.index(J.rq_off + q, e)], nbk[e], false);
                            J.t16[at(q, e)] = o;
                            if (!o.accept()) J.need64[e] = 1;
                        }
            }
        }
    }

    // ---- the replay's first pass on the GPU: per chunk, its largest inserts ----
    {
        std::vector<size_t> dj;
        for (size_t j = 0; j < jobs.size(); ++j)
            if (jobs[j].dev_top) dj.push_back(j);
        const size_t nc = chunks.size();
        if (!dj.empty() && nc > 0) {
            // The selection runs on a side stream after the tiers' work: on
            // the score kernels' own stream, this kernel was measured to slow
            // the next searches' score kernels by ~20 % whenever consecutive
            // searches differ (query length, kernel variant), an effect that
            // no kernel of the same grid, and none of this one's results,
            // shows from a second stream.
            const hipStream_t side = ex.aux(0);
            hipEvent_t ready;
            SSA_HIP(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
            SSA_HIP(hipEventRecord(ready, ex.stream()));
            SSA_HIP(hipStreamWaitEvent(side, ready, 0));
            std::vector<uint2> hc(nc);
            for (size_t c = 0; c < nc; ++c) hc[c] = make_uint2(chunks[c].first, chunks[c].second);
            auto* d_chunks = (uint2*)ex.arena("top_chunks", nc * sizeof(uint2));
            auto* d_top = (ChunkTop*)ex.arena("top_out", dj.size() * nc * sizeof(ChunkTop));
            auto* h_top = (ChunkTop*)ex.pinned("top_out", dj.size() * nc * sizeof(ChunkTop));
            auto* h_chunks = (uint2*)ex.pinned("top_chunks", nc * sizeof(uint2));
            memcpy(h_chunks, hc.data(), nc * sizeof(uint2));
            SSA_HIP(hipMemcpyAsync(d_chunks, h_chunks, nc * sizeof(uint2), hipMemcpyHostToDevice, side));
            for (size_t i = 0; i < dj.size(); ++i) {
                const Job& J = jobs[dj[i]];
                TopSource S{};
                for (size_t q = 0; q < J.qs.size(); ++q) S.res16[q] = J.dev16[q];
                S.flip = J.dev_flip;
                S.nq = (int32_t)J.qs.size();
                S.mode = J.top_mode;
                S.type = J.rq.type == kSW ? 0 : 1;
                S.m = (int32_t)std::min<size_t>(J.rq.hitcount + 1, kChunkTop);
                SSA_HIP(launch_chunk_top((int)nc, side, S, d_chunks, d_top + i * nc));
            }
            SSA_HIP(hipMemcpyAsync(h_top, d_top, dj.size() * nc * sizeof(ChunkTop), hipMemcpyDeviceToHost, side));
            SSA_HIP(hipStreamSynchronize(side));
            SSA_HIP(hipEventDestroy(ready));
            for (size_t i = 0; i < dj.size(); ++i) {
                Job& J = jobs[dj[i]];
                const ChunkTop* t = h_top + i * nc;
                J.st.d2h_bytes += nc * sizeof(ChunkTop);
                bool exact = false;
                for (size_t c = 0; c < nc && !exact; ++c) exact = t[c].need_exact != 0;
                if (exact) {
                    // some pair overflowed 16 bit: the exact tier's scores enter
                    // the inserts, so this search takes the host path (all of
                    // its results decoded on the host, then the exact tier)
                    const size_t nq = J.qs.size();
                    std::vector<std::vector<PairResult>> h(nq, std::vector<PairResult>(N));
                    for (size_t q = 0; q < nq; ++q) ex.download(h[q].data(), J.dev16[q], N * sizeof(PairResult));
                    ex.sync();
                    J.st.d2h_bytes += nq * N * sizeof(PairResult);
                    decode_top_range(
                        J, J.top_mode, [&](size_t q, uint32_t e) { return h[q][e ^ ((J.dev_flip >> q) & 1u)]; }, nbk,
                        N, 0, N);
                    J.dev_top = false;
                    continue;
                }
                ChunkFirst& F = J.first;
                F.top.assign(nc * kChunkTop, 0);
                F.ntop.assign(nc, 0);
                F.ov8.assign(nc, 0);
                F.ov16.assign(nc, 0);
                for (size_t c = 0; c < nc; ++c) {
                    F.ntop[c] = (uint8_t)t[c].ntop;
                    for (int x = 0; x < t[c].ntop; ++x) F.top[c * kChunkTop + x] = t[c].top[x];
                    F.ov8[c] = (size_t)t[c].ov8;
                    F.ov16[c] = (size_t)t[c].ov16;
                }
            }
        }
    }

    parallel_jobs([&](size_t j) {
        Job& J = jobs[j];
        const SearchType type = J.rq.type;
        const size_t nq = J.qs.size();
        const int bw = J.rq.bit_width;
        auto raw = [&](size_t q, uint32_t e) -> const PairResult& {
            const TierRun& R = *runs[J.group];
            if (J.twin && q == 1) return R.out[R.ps->index(J.q_off, e ^ 1u)];
            return R.out[R.ps->index(J.q_off + q, e)];
        };
        // a 64-bit search scores with the 16-bit pass where it is provably
        // exact; the other entries reach the exact tier
        if (bw != BIT_WIDTH_64) return;
        if (J.group < 0) {
            if (N) memset(J.need64.data(), 1, N);
            return;
        }
        J.t16.reset(nq * N);
        const int64_t gap = std::abs((int64_t)go) + std::abs((int64_t)ge);
        over_entries([&](uint32_t e0, uint32_t e1) {
            for (size_t q = 0; q < nq; ++q)
                for (uint32_t e = e0; e < e1; ++e) {
                    TierOut o = decode(type, kTier16, raw(q, e), nbk[e], false);
                    // an empty entry's lane scores a padding block; 64 bit scores the empty sequence
                    bool exact = o.accept() && db.entries[e].len != 0;
                    if (type == kNW) {
                        // exact when no value can reach either end of the int16 range
                        const int64_t ml = (int64_t)std::min<size_t>(J.qs[q].len, db.entries[e].len);
                        const int64_t span = 3 * gap + gap * (int64_t)(J.qs[q].len + db.entries[e].len + 4);
                        const int64_t up = std::max<int64_t>(mi.maxv, 0) * ml;
                        const int64_t lowest = -span - std::max<int64_t>(-mi.minv, 0) * ml;
                        if (!(lowest > -32768 + gap + 1 && up < 32767)) exact = false;
                    }
                    if (!exact) {
                        o.d = -o.dur();
                        J.need64[e] = 1;
                    }
                    J.t16[at(q, e)] = o;
                }
        });
    });
    for (size_t j = 0; j < jobs.size(); ++j) {
        Job& J = jobs[j];
        const size_t nq = J.qs.size();
        double& ms_re = ms_re_job[j];
        uint64_t& computed = computed_job[j];
        // pairs of the exact tier: entry by entry, every query of each
        J.x_entries.clear();
        J.x64.clear();
        const uint8_t* nd = J.need64.data();
        for (uint32_t e = 0; e < N;) {
            uint64_t w;
            if (e + 8 <= N && (memcpy(&w, nd + e, 8), w == 0)) {
                e += 8;
                continue;
            }
            if (nd[e]) J.x_entries.push_back(e);
            ++e;
        }
        if (!J.x_entries.empty()) {
            std::vector<std::pair<uint32_t, uint32_t>> p64;
            p64.reserve(J.x_entries.size() * nq);
            for (uint32_t e : J.x_entries)
                for (size_t q = 0; q < nq; ++q) p64.push_back({(uint32_t)q, e});
            ex.recycle();
            gpu_exact(ex, J.rq.type, J.qs, p64, J.x64, &ms_re, &computed);
        }
        J.st.ms_gpu_requeue = ms_re;
        J.st.cells_computed = computed;
    }
    runs.clear();
    run_sets.clear();

    // ---- replay: searches in parallel, or chunks in parallel for few searches ----
    // A search whose first pass ran on the GPU reads the tier results of the
    // chunks it simulates only: gathered into one compact buffer on the
    // device, downloaded and decoded (one context, so one search at a time).
    std::mutex gather_mu;
    auto replay = [&](Job& J, bool threads) {
        if (!J.dev_top) {
            job_replay(J, db.entries, chunks, N, threads);
            return;
        }
        job_replay(J, db.entries, chunks, N, threads, true, &J.first, [&](const std::vector<uint32_t>& work) {
            if (work.empty()) return;
            const size_t nq = J.qs.size();
            // query buffers reading the same device results (strand twins)
            // share one gathered range
            size_t src_of[6];
            for (size_t q = 0; q < nq; ++q) {
                src_of[q] = q;
                for (size_t p = 0; p < q; ++p)
                    if (J.dev16[p] == J.dev16[q]) {
                        src_of[q] = src_of[p];
                        break;
                    }
            }
            std::vector<GatherSeg> segs;
            std::vector<int64_t> at(work.size() * nq);
            segs.reserve(work.size() * nq);
            int64_t total = 0;
            for (size_t w = 0; w < work.size(); ++w)
                for (size_t q = 0; q < nq; ++q) {
                    const uint32_t e0 = chunks[work[w]].first, e1 = chunks[work[w]].second;
                    if (src_of[q] != q) {
                        at[w * nq + q] = at[w * nq + src_of[q]];
                        continue;
                    }
                    at[w * nq + q] = total;
                    segs.push_back(GatherSeg{J.dev16[q] + e0, total, (int32_t)(e1 - e0), 0});
                    total += e1 - e0;
                }
            // the context's stream and staging buffers serve one search at a
            // time; the decode runs outside the lock on the search's own copy
            HostArray<PairResult> got;
            got.reset((size_t)total);
            {
                std::lock_guard<std::mutex> lk(gather_mu);
                auto* h_segs = (GatherSeg*)ex.pinned("gather_segs", segs.size() * sizeof(GatherSeg));
                auto* d_segs = (GatherSeg*)ex.arena("gather_segs", segs.size() * sizeof(GatherSeg));
                auto* d_out = (PairResult*)ex.arena("gather_out", (size_t)total * sizeof(PairResult));
                auto* h_out = (PairResult*)ex.pinned("gather_out", (size_t)total * sizeof(PairResult));
                memcpy(h_segs, segs.data(), segs.size() * sizeof(GatherSeg));
                ex.upload(d_segs, h_segs, segs.size() * sizeof(GatherSeg));
                SSA_HIP(launch_gather((int)segs.size(), ex.stream(), d_segs, d_out));
                ex.download(h_out, d_out, (size_t)total * sizeof(PairResult));
                ex.sync();
                memcpy(got.data(), h_out, (size_t)total * sizeof(PairResult));
            }
            J.st.d2h_bytes += (uint64_t)total * sizeof(PairResult);
            auto decode_chunk = [&](size_t w) {
                const uint32_t e0 = chunks[work[w]].first, e1 = chunks[work[w]].second;
                const PairResult* base[6];
                for (size_t q = 0; q < nq; ++q) base[q] = got.data() + at[w * nq + q];
                decode_top_range(
                    J, J.top_mode,
                    [&](size_t q, uint32_t e) { return base[q][(e ^ ((J.dev_flip >> q) & 1u)) - e0]; }, nbk, N, e0,
                    e1);
            };
            if (threads && work.size() > 8) {
                parallel_for(work.size(), decode_chunk);
            } else {
                for (size_t w = 0; w < work.size(); ++w) decode_chunk(w);
            }
        });
    };
    if (jobs.size() >= 4 && nthreads > 1) {
        parallel_for(jobs.size(), [&](size_t j) { replay(jobs[j], false); });
    } else {
        for (Job& J : jobs) replay(J, true);
    }
    for (Job& J : jobs)
        if (J.st.overflow_8_bit_count || J.st.overflow_16_bit_count)
            print_info("Overflow occurred: %ld sequences were re-aligned with 16 bit, and %ld sequences with 64 bit\n",
                       (long)J.st.overflow_8_bit_count, (long)J.st.overflow_16_bit_count);

    // ---- the no-allocation path: the hits themselves, no lists ----
    if (fixed) {
        for (size_t j = 0; j < jobs.size(); ++j) {
            Job& J = jobs[j];
            std::vector<ssa_amd_hit_t>& out = results[j].hits;
            for (const Elem& h : J.hits) {
                const QueryBuf& q = J.qs[h.query_id];
                // the fields alignment_t carries, with the reference's swap of
                // the entry's strand and frame (src/algo/aligner.c:76-77)
                out.push_back(ssa_amd_hit_t{(size_t)h.db_id, (int)h.db_frame, (int)h.db_strand, q.strand, q.frame,
                                            (long)h.score});
            }
            J.st.ms_total = (now_us() - t_start) / 1000.0;
            results[j].stats = J.st;
        }
        return;
    }

    // ---- alignment lists; one traceback launch for every hit of the batch ----
    std::vector<p_alignment> al;
    std::vector<SearchType> types;
    std::vector<size_t> al_job;
    for (size_t j = 0; j < jobs.size(); ++j) {
        const size_t before = al.size();
        job_list(jobs[j], al, types);
        for (size_t k = before; k < al.size(); ++k) al_job.push_back(j);
    }
    double ms_tb = 0;
    gpu_traceback(ex, types, al, &ms_tb);
    for (size_t j = 0; j < jobs.size(); ++j) {
        Job& J = jobs[j];
        if (J.rq.align_type != COMPUTE_SCORE) J.st.ms_traceback = ms_tb;
        J.st.ms_total = (now_us() - t_start) / 1000.0;
        results[j].list = J.list;
        results[j].stats = J.st;
    }
}

// Times every kernel shape on the whole loaded DB for `copies` random queries
// of each length in one launch (the first 16-bit tier, a search's main
// launch; copies > 1 measures the shape as a batch runs it) and keeps the
// fastest per (type, length).
int autotune(int type, const std::vector<size_t>& lengths, int copies, std::vector<AutotuneResult>& out) {
    const std::vector<Shape>& kShapes = kernel_shapes(type);
    Exec& ex = Gpu::get().exec();
    DeviceDB* dbp;
    {
        std::lock_guard<std::mutex> lk(db_mutex());
        dbp = &device_db();
    }
    DeviceDB& db = *dbp;
    const Config& cf = config();
    const uint32_t N = (uint32_t)db.entries.size();
    out.clear();
    if (N == 0) return 0;
    copies = std::max(1, copies);
    PairSet all;
    all.ext = &db_all_entries(db);
    // query symbols of the current alphabet (codes of the standard residues)
    std::vector<uint8_t> alphabet;
    const char* syms = cf.symtype == NUCLEOTIDE ? "ACGT" : "ARNDCQEGHILKMFPSTWYV";
    for (const char* c = syms; *c; ++c)
        alphabet.push_back((uint8_t)(cf.symtype == NUCLEOTIDE ? map_nt(*c) : map_aa(*c)));
    uint64_t x = 0x9E3779B97F4A7C15ull;
    auto run = [&](const Shape& sh, const std::vector<std::vector<uint8_t>>& qs) {
        TierRun R;
        R.type = type == 0 ? kSW : kNW;
        R.T = kTier16;
        for (const auto& q : qs) R.qs.push_back(QueryBuf{q.data(), q.size(), 0, 0});
        PairSet ps(all);
        ps.nq = R.qs.size();
        R.ps = &ps;
        R.sorted_all = &db.by_length;
        R.sh = sh;
        R.pfx = "tune_";
        double ms = 0;
        gpu_score_tier(ex, R, &ms);
        return ms;
    };
    // build every lane width's streams first, outside the timed runs
    {
        std::vector<std::vector<uint8_t>> tiny(1, std::vector<uint8_t>(8, alphabet[0]));
        for (int lanes : {8, 16, 32, 64}) run(Shape{lanes, 4}, tiny);
    }
    for (size_t L : lengths) {
        if (L == 0) continue;
        std::vector<std::vector<uint8_t>> qs(copies, std::vector<uint8_t>(L));
        for (auto& q : qs)
            for (size_t i = 0; i < L; ++i) {
                x ^= x << 13;
                x ^= x >> 7;
                x ^= x << 17;
                q[i] = alphabet[x % alphabet.size()];
            }
        AutotuneResult best{L, 0, 0, 4, 1, 1e300};
        std::vector<TunedShape> meas;
        for (const Shape& base : kShapes)
            for (int cols : {4, 8})
                for (int place : {1, 2}) {
                    if (!score_shape_supported(base.lanes, base.rows, type, cols)) continue;
                    // the profile in VGPRs where the DB allows it (NW: the kernel
                    // without overflow trackers has it, which the run decides)
                    if (place == 2 && !(db.acgt && score_vp_supported(base.lanes, base.rows, type == 0 ? 0 : 2, cols)))
                        continue;
                    Shape sh = base;
                    sh.cols = cols;
                    sh.place = place;
                    const double ms = run(sh, qs) / copies;
                    meas.push_back(TunedShape{type, L, sh, ms});
                    if (ms < best.ms) best = AutotuneResult{L, sh.lanes, sh.rows, sh.cols, sh.place, ms};
                }
        out.push_back(best);
        std::lock_guard<std::mutex> lk(tuned_mutex());
        auto& t = tuned_shapes();
        t.erase(std::remove_if(t.begin(), t.end(), [&](const TunedShape& e) { return e.type == type && e.qlen == L; }),
                t.end());
        t.insert(t.end(), meas.begin(), meas.end());
    }
    return (int)out.size();
}

void autotune_reset() {
    std::lock_guard<std::mutex> lk(tuned_mutex());
    tuned_shapes().clear();
}

int replay_selfcheck(uint64_t seed, int cases) {
    std::mt19937_64 rng(seed);
    auto R = [&](uint64_t n) { return (uint64_t)(rng() % n); };
    int bad = 0;
    for (int t = 0; t < cases; ++t) {
        static const int kWidths[3] = {BIT_WIDTH_8, BIT_WIDTH_16, BIT_WIDTH_64};
        static const size_t kQueries[3] = {1, 2, 6};
        static const size_t kHits[7] = {1, 2, 3, 5, 50, 100, 5000};
        Job J;
        memset(&J.st, 0, sizeof J.st);
        J.rq = SearchRequest{kSW, nullptr, kHits[R(7)], kWidths[R(3)], COMPUTE_SCORE};
        const size_t nq = kQueries[R(3)];
        J.qs.assign(nq, QueryBuf{nullptr, 1, 0, 0});
        const uint32_t N = 1 + (uint32_t)R(3000);
        const int range = R(2) ? (int)(2 + R(40)) : 1000;  // narrow ranges: many equal scores
        std::vector<DbEntry> entries(N);
        for (uint32_t e = 0; e < N; ++e)
            entries[e] = DbEntry{e / (1 + R(2)), 1, (uint8_t)R(2), (uint8_t)R(3), 0};
        J.t8.reset(nq * N);
        J.t16.reset(nq * N);
        J.in16.reset(N);
        J.in16.fill_zero();
        J.need64.reset(N);
        J.need64.fill_zero();
        const uint64_t p8 = R(2) ? 20 : 300, p16 = R(2) ? 50 : 500;  // overflow rates per 1000
        for (uint32_t e = 0; e < N; ++e)
            for (size_t q = 0; q < nq; ++q) {
                const int32_t d = 1 + (int32_t)R(12);
                const TierOut o{R(1000) < p8 ? -d : d, (int32_t)R(range)};
                J.t8[q * N + e] = o;
                if (J.rq.bit_width == BIT_WIDTH_8 && !o.accept()) J.in16[e] = 1;
            }
        if (J.rq.bit_width != BIT_WIDTH_8) memset(J.in16.data(), 1, N);
        for (uint32_t e = 0; e < N; ++e)
            for (size_t q = 0; q < nq; ++q) {
                const int32_t d = 1 + (int32_t)R(12);
                const TierOut o{R(1000) < p16 ? -d : d, (int32_t)R(range) + 50};
                J.t16[q * N + e] = o;
                if (J.in16[e] && !o.accept()) J.need64[e] = 1;
            }
        for (uint32_t e = 0; e < N; ++e)
            if (J.need64[e]) {
                J.x_entries.push_back(e);
                for (size_t q = 0; q < nq; ++q) J.x64.push_back((int64_t)R(range) + 100);
            }
        const size_t cs = 1 + R(400);
        std::vector<std::pair<uint32_t, uint32_t>> chunks;
        for (uint32_t c0 = 0; c0 < N; c0 += (uint32_t)cs) chunks.push_back({c0, (uint32_t)std::min<size_t>(N, c0 + cs)});
        const bool threads = R(2);
        job_replay(J, entries, chunks, N, threads, false);
        const std::vector<Elem> full = J.hits;
        const ssa_amd_stats_t st_full = J.st;
        memset(&J.st, 0, sizeof J.st);
        job_replay(J, entries, chunks, N, !threads, true);
        bool same = full.size() == J.hits.size() && st_full.overflow_8_bit_count == J.st.overflow_8_bit_count &&
                    st_full.overflow_16_bit_count == J.st.overflow_16_bit_count;
        for (size_t i = 0; same && i < full.size(); ++i) {
            const Elem &a = full[i], &b = J.hits[i];
            same = a.score == b.score && a.db_id == b.db_id && a.query_id == b.query_id && a.db_frame == b.db_frame &&
                   a.db_strand == b.db_strand;
        }
        bad += !same;
    }
    return bad;
}

// The device DB for the current settings, built if needed (the server makes
// it resident once, at start, before its workers run).
void ensure_device_db() {
    std::lock_guard<std::mutex> lk(db_mutex());
    (void)device_db();
}

// What one request will take on the device, from the kernel shape the search
// would choose over the resident DB: its first-tier cells, the HBM its
// buffers hold while it runs (pair results of the first tier and of a 16-bit
// re-run, pass boundaries of a multi-pass query, traceback matrices), and
// the share of the chip's resident workgroup slots its persistent first-tier
// grid fills (the shape's LDS and VGPR use set how many workgroups a CU holds).
RequestPlan plan_request(const SearchRequest& rq, const std::vector<size_t>& qlens) {
    RequestPlan pl;
    DeviceDB* dbp;
    {
        std::lock_guard<std::mutex> lk(db_mutex());
        dbp = &device_db(false);
    }
    const DeviceDB& db = *dbp;
    const Config& cf = config();
    const DbChunks& dc = db_chunks(db, cf.chunk_size ? cf.chunk_size : kDefaultChunkSize);
    const int type = rq.type == kSW ? 0 : 1;
    size_t qmax = 0, qsum = 0;
    for (size_t l : qlens) qmax = std::max(qmax, l), qsum += l;
    const bool packed = type == 0 && qlens.size() > 1;
    const Shape sh = choose_shape(packed ? qsum : qmax, dc.total_blocks, dc.max_blocks, type, packed);
    pl.lanes = sh.lanes;
    pl.rows = sh.rows;
    const double nw8 = (rq.type == kNW && rq.bit_width == BIT_WIDTH_8) ? 2.0 : 1.0;  // its 16-bit re-run
    pl.cells = (double)qsum * (double)dc.total_residues * nw8;
    const size_t nq = qlens.size(), N = dc.N;
    size_t bytes = (size_t)(nq * N * sizeof(PairResult) * nw8);
    // waves of the shape's stream layout (cached, else the layout's target)
    int nwaves = 0, max_nb = 0;
    {
        std::lock_guard<std::mutex> lk(layout_mutex());
        for (const StreamLayout& c : layout_cache())
            if (c.lanes == sh.lanes && c.build == db.build && !c.wave_nblocks.empty()) {
                nwaves = c.nwaves;
                max_nb = *std::max_element(c.wave_nblocks.begin(), c.wave_nblocks.end());
            }
    }
    if (nwaves == 0) {
        nwaves = std::max(kWavesPerWG, target_waves());
        max_nb = (int)std::max<int64_t>(dc.max_blocks, 16 * (int64_t)sh.lanes);
    }
    const int PR = sh.lanes * sh.rows;
    const int passes = (int)((std::max<size_t>(packed ? qsum : qmax, 1) + PR - 1) / PR);
    const int cap = std::max(1, score_resident_workgroups(sh.lanes, type == 0 ? sh.rows : std::min(sh.rows, 16),
                                                          type == 0 ? 0 : 1, sh.cols, false));
    const int tasks = std::max(1, nwaves / kWavesPerWG) * (packed ? 1 : (int)nq);
    const int grid = std::min(tasks, cap);
    pl.occupancy = (double)grid / (double)cap;
    if (passes > 1)
        bytes += (size_t)grid * kWavesPerWG * (size_t)(max_nb + 2) * (size_t)(64 / sh.lanes) *
                 (2 * sizeof(uint4) + (type == 0 ? sizeof(uint32_t) : 0));
    if (rq.align_type != COMPUTE_SCORE) {
        int64_t dmax = 0;
        for (uint32_t e = 0; e < N; ++e) dmax = std::max<int64_t>(dmax, db.entries[e].len);
        bytes += (size_t)std::min<int64_t>((int64_t)rq.hitcount * (int64_t)((qmax + 7) & ~(size_t)7) * dmax,
                                           kTraceDirBudget);
    }
    pl.hbm_bytes = bytes;
    return pl;
}

int run_search_noalloc(SearchType type, p_query query, size_t hitcount, int bit_width,
                       std::vector<ssa_amd_hit_t>& hits) {
    static std::mutex mu;  // calls share the device's scratch array
    std::lock_guard<std::mutex> lk(mu);
    Exec& ex = Gpu::get().fixed_exec();
    ex.begin_call();
    std::vector<SearchRequest> rq(1, SearchRequest{type, query, hitcount, bit_width, COMPUTE_SCORE});
    std::vector<SearchResult> res;
    try {
        run_batch(ex, rq, res, false);
    } catch (const FixedRegionError& e) {
        (void)hipStreamSynchronize(ex.stream());  // work already queued on the scratch ends first
        ex.begin_call();
        print_error("ssa_amd_align_noalloc: %s", e.what);
        return e.code;
    }
    ex.begin_call();
    last_stats() = res[0].stats;
    hits.swap(res[0].hits);
    return 0;
}

p_alignment_list run_search(SearchType type, p_query query, size_t hitcount, int bit_width, int align_type) {
    std::vector<SearchRequest> rq(1, SearchRequest{type, query, hitcount, bit_width, align_type});
    std::vector<SearchResult> res;
    run_batch(Gpu::get().exec(), rq, res, false);
    last_stats() = res[0].stats;
    return res[0].list;
}

}  // namespace ssa
