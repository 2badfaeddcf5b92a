// dbpack.h -- the pre-packed DB file: the plugin DB as the search holds it.
//
// A packed file carries, next to the sequences themselves (so the built-in
// provider serves ssa_db_get_sequence from it), the mapped symbol codes of
// every search entry (reverse complements included for nucleotide DBs), the
// entry table and the entries in length order.  Loading it is a read and one
// upload; no sequence is mapped again.
//
// Layout (little-endian; every section starts on a 64-byte boundary):
//   PackHeader (128 bytes)
//   SeqRec[seq_count]          provider ID, offset and length of the residues
//   char   ascii[ascii_bytes]  the residues as the provider returned them, each
//                              followed by a NUL
//   EntryRec[entry_count]      one per search entry, in DB order (strand 0
//                              then strand 1 for nucleotide DBs)
//   u32    first[seq_count+1]  first entry of each sequence (empty ones own none)
//   u32    by_length[entry_count]  entries longest first, ties in entry order
//   u8     codes[code_bytes]   each entry's codes, padded to a multiple of 4
//   UnknownRec[unknown_count]  sequences holding symbols that mapped to 0
#pragma once
#include <cstddef>
#include <cstdint>

#include "search.h"

extern "C" {
#include "libssa_extern_db.h"
}

namespace ssa {

constexpr char kPackMagic[8] = {'S', 'S', 'A', 'P', 'A', 'C', 'K', '1'};
constexpr uint32_t kPackVersion = 1;

struct PackHeader {
    char magic[8];
    uint32_t version;
    uint32_t header_bytes;
    int32_t symtype;  // NUCLEOTIDE or AMINOACID
    int32_t strands;  // 3: nucleotide, both strands' entries; 1: amino acid
    uint64_t seq_count, entry_count, ascii_bytes, code_bytes, unknown_count;
    uint64_t off_seqs, off_ascii, off_entries, off_first, off_bylen, off_codes, off_unknown;
    uint64_t file_bytes;
};
static_assert(sizeof(PackHeader) == 128, "packed header is 128 bytes");

struct SeqRec {
    uint64_t id, ascii_off, len;
};

struct EntryRec {
    uint64_t id;
    uint32_t len;
    uint8_t strand, frame;
    uint16_t pad;
    int64_t code_off;
};
static_assert(sizeof(EntryRec) == 24, "packed entry is 24 bytes");

struct UnknownRec {
    uint64_t id, count;
};

// A validated, memory-mapped packed file.
struct PackedView {
    const PackHeader* h = nullptr;
    const SeqRec* seqs = nullptr;
    const char* ascii = nullptr;
    const EntryRec* entries = nullptr;
    const uint32_t* first = nullptr;
    const uint32_t* by_length = nullptr;
    const uint8_t* codes = nullptr;
    const UnknownRec* unknown = nullptr;
};

// true when path starts with the packed magic
bool is_packed_file(const char* path);

// Maps and validates a packed file; on failure returns false with a reason.
bool packed_open(const char* path, PackedView* view, void** map, size_t* map_bytes, const char** why);
void packed_close(void* map, size_t map_bytes);

// The packed file the built-in provider has open, or nullptr (FASTA open,
// nothing open).  Defined by the provider.
const PackedView* provider_packed_view();
// provider's seqinfo of sequence 0, to tell the built-in provider from a
// caller's own (interposed) one
const seqinfo_t* provider_first_info();

// Writes the DB the plugin currently serves, mapped for the symbol type in
// force, to path.  Returns 0, or -1 after an error message.
int write_packed_db(const char* path);

}  // namespace ssa
