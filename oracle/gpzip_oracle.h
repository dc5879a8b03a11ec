/*
 * gpzip_oracle.h -- CPU restatement of gpzip's GPU gzip compress path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (gpzip_amd/, the HIP
 * library, include/gpzip_hip.h) links, loads or calls this code.  Only the
 * tests, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it, for
 * the expected outputs and as the reported CPU baseline.
 *
 * Every function names the reference file:line it restates (paths relative
 * to the reference repository, crates/gpzip- prefix dropped).
 *
 * Token format (reference lz77_hash.rs:50-62, walk.rs:492-497): one u32 per
 * token, length << 16 | (distance or literal byte); length == 0 => literal.
 */
#ifndef GPZIP_ORACLE_H
#define GPZIP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_HASH_BITS 16u     /* codec-gpu/src/gpu/lz77_hash.rs:43 */
#define ORC_SEG_LOG2 12u      /* codec-gpu/src/gpu/lz77_hash.rs:68 */
#define ORC_WINDOW 32768u     /* codec-gpu/src/gpu/lz77_hash.rs:48 */
#define ORC_MIN_MATCH 3u      /* codec-gpu/src/gpu/lz77_hash.rs:45 */
#define ORC_MAX_MATCH 258u    /* codec-gpu/src/gpu/lz77_hash.rs:46 */
#define ORC_OLDEST_SKIP 16u   /* codec-gpu/src/gpu/lz77_hash_lookup.wgsl:102 */
#define ORC_SUB_CHUNK 131072u /* codec-gpu/src/gpu.rs:160 */

/* hash3 of lz77_hash_build.wgsl:36-43; requires p + 2 < n. */
uint32_t orc_hash3(const uint8_t *in, size_t p);

/* Segmented-hash match finder: build (lz77_hash_build.wgsl:45-54) then
 * lookup (lz77_hash_lookup.wgsl:45-119) over a dense [hash][segment] table
 * exactly as the reference lays it out (lz77_hash.rs:120-181).  Writes n
 * per-position tokens.  Returns 0, or -1 on allocation failure. */
int orc_match_find(const uint8_t *in, size_t n, uint32_t window, uint32_t *tok);

/* Second, table-free statement of the same candidate rule: for every
 * segment the newest/oldest candidate is found by scanning the segment for
 * positions with the same hash.  O(n^2); tiny inputs only. */
int orc_match_find_brute(const uint8_t *in, size_t n, uint32_t window, uint32_t *tok);

/* Lazy walk, codec-gpu/src/gpu/lz77.rs:260-284.  Returns token count. */
size_t orc_greedy_walk(const uint32_t *per_pos, const uint8_t *in, size_t n, uint32_t *out);

/* Strict walk, codec-gpu/src/gpu/walk.rs:512-525.  Returns token count. */
size_t orc_strict_walk(const uint32_t *per_pos, size_t n, uint32_t *out);

/* The opt-in ratio mode's walk (no reference counterpart): demote a match
 * also when the match two positions ahead is longer by two or more. */
size_t orc_lookahead_walk(const uint32_t *per_pos, const uint8_t *in, size_t n, uint32_t *out);

/* reconstruct, codec-gpu/src/gpu/lz77.rs:288-304.  Returns bytes written or
 * (size_t)-1 when a distance reaches before the start / out_cap overflows. */
size_t orc_reconstruct(const uint32_t *tok, size_t ntok, uint8_t *out, size_t out_cap);

/* try_build_code_lengths, codec-gpu/src/gpu/huffman.rs:26-60 (+63-117).
 * Returns 1 on success, 0 when the reference returns None. */
int orc_build_code_lengths(const uint32_t *freq, size_t n, uint32_t max_bits, uint8_t *lens);

/* canonical_codes, codec-gpu/src/gpu/huffman.rs:122-152. */
void orc_canonical_codes(const uint8_t *lens, size_t n, uint32_t *codes);

/* encode_block_fast, codec-gpu/src/gpu/deflate.rs:752-959.  Returns the
 * number of bytes written or (size_t)-1 if out_cap is too small or a token
 * is outside the DEFLATE alphabet. */
size_t orc_encode_block_fast(const uint32_t *tok, size_t ntok, uint8_t *out, size_t out_cap);

/* CRC-32 (IEEE, reflected 0xEDB88320), as crc32fast::hash at deflate.rs:974. */
uint32_t orc_crc32(const uint8_t *in, size_t n);

/* gzip_wrap, codec-gpu/src/gpu/deflate.rs:962-980. */
size_t orc_gzip_wrap(const uint8_t *deflate, size_t dlen, const uint8_t *orig, size_t n,
                     uint8_t *out, size_t out_cap);

/* encode_one (gpu.rs:271-287) on one piece: match_find -> greedy_walk ->
 * encode_block_fast -> gzip_wrap.  Returns member bytes or (size_t)-1. */
size_t orc_gzip_member(const uint8_t *in, size_t n, uint8_t *out, size_t out_cap);
/* rule 0: the reference's lazy walk; rule 2: orc_lookahead_walk (ratio mode) */
size_t orc_gzip_member_rule(const uint8_t *in, size_t n, int rule, uint8_t *out, size_t out_cap);
size_t orc_gzip_chunk_rule(const uint8_t *in, size_t n, size_t sub_chunk, int rule, uint8_t *out,
                           size_t out_cap);

/* gzip_chunk_fn (gpu.rs:193-208): split into sub_chunk pieces, one member
 * each, concatenated.  Returns total bytes or (size_t)-1. */
size_t orc_gzip_chunk(const uint8_t *in, size_t n, size_t sub_chunk, uint8_t *out,
                      size_t out_cap);

/* Safe output capacity for orc_gzip_chunk / orc_gzip_member. */
size_t orc_chunk_bound(size_t n, size_t sub_chunk);

#ifdef __cplusplus
}
#endif
#endif
