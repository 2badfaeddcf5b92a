/*
 * libssa_extern_db.h -- the database plugin interface the library imports.
 *
 * Same interface as the reference's src/libssa_extern_db.h:28-56.  The caller
 * may link its own implementation; libssa_amd.so ships a default provider that
 * reads FASTA files (0-based IDs in file order, NULL past the end), which a
 * caller's own definitions override by ordinary symbol interposition.
 */
#ifndef LIBSSA_EXTERN_DB_H_
#define LIBSSA_EXTERN_DB_H_

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    size_t ID;
    size_t seqlen;
    char* seq;
} seqinfo_t; /* src/libssa_extern_db.h:28-34 */

typedef seqinfo_t* p_seqinfo;

/* src/libssa_extern_db.h:39 */
int ssa_db_init(const char* db_name);
/* src/libssa_extern_db.h:44 */
size_t ssa_db_get_sequence_count(void);
/* src/libssa_extern_db.h:49 */
p_seqinfo ssa_db_get_sequence(size_t id);
/* src/libssa_extern_db.h:56 */
void ssa_db_close(void);

#ifdef __cplusplus
}
#endif

#endif /* LIBSSA_EXTERN_DB_H_ */
