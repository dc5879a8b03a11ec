// gpz_internal.h -- device data layout shared by the kernels and the host
// runtime of libgpzip_hip.so.
//
// A "batch" is a list of chunks; each chunk is cut into members of at most
// sub_chunk bytes (one gzip member each, reference codec-gpu/src/gpu.rs:193-208),
// and each member into 4 KiB segments (the reference's SEG_LOG2 = 12,
// codec-gpu/src/gpu/lz77_hash.rs:68).  Kernels are gridded over segments
// (match finding, walk) or members (Huffman, emission).
#pragma once
#include <stdint.h>

namespace gpz {

constexpr uint32_t SEG_LOG2 = 12;
constexpr uint32_t SEG = 1u << SEG_LOG2;
constexpr uint32_t MAX_MATCH = 258;
constexpr uint32_t OLDEST_SKIP = 16;   // lz77_hash_lookup.wgsl:102
constexpr uint32_t SPAN = 258;         // entry offsets into a segment: 0..257
constexpr uint32_t NSYM = 320;         // 286 litlen + 30 dist, padded
constexpr uint32_t DSYM0 = 286;        // first dist slot in the 320-entry tables
constexpr uint32_t PREFIX_WORDS = 160; // gzip header + dynamic block header bits
constexpr uint32_t POS_ALIGN = 64;     // member token regions start on 64-position boundaries
constexpr uint32_t IN_ALIGN = 256;     // chunk data alignment in device input buffers
// Per-segment hash table (k_seg_table -> k_match_find), 7168 words:
//   [0, 2048)      presence bitmap of the segment's 16-bit hashes
//   [2048, 3072)   2048 x u16: set bits before each bitmap word (rank)
//   [3072, 7168)   per hash, in hash order: first << 12 | last position
constexpr uint32_t TAB_BITS = 0, TAB_RANK = 2048, TAB_ENT = 3072, TAB_WORDS = 7168;

// One gzip member: n input bytes at in_off in the device input buffer.
struct MemberDesc {
    uint64_t in_off;    // byte offset of the member's first byte (16-B aligned)
    uint64_t pos_base;  // first index of its region in tok[] / walked[] (64-aligned)
    uint32_t n;         // member length in bytes
    uint32_t seg_base;  // global index of its first segment
    uint32_t nseg;      // ceil(n / 4096)
    uint32_t chunk;     // owning chunk
};

// Batch sizes as read by persistent kernels: [0] segments, [1] members, [2] chunks.
struct Counts {
    uint32_t nseg, nmem, nchunk, pad;
};

// Pointers to every device array a batch run uses.
struct Work {
    const uint8_t *in;          // input bytes
    const MemberDesc *mem;      // [M]
    const uint32_t *seg_mem;    // [G] member of each segment
    const uint32_t *chunk_mem0; // [C] first member of each chunk
    const Counts *counts;
    uint32_t *tok;              // [P] per-position tokens
    uint32_t *segtab;           // [G * TAB_WORDS] per-segment hash tables (layout above)
    uint32_t *seg_ntab;         // [G] distinct hashes of each segment
    uint32_t *wsum;             // [G * SPAN] walk summary: exit | count << 16
    uint32_t *seg_crc;          // [G] raw CRC-32 (init 0, no final xor) of the segment
    uint32_t *seg_entry;        // [G] walk entry offset into the segment
    uint32_t *seg_tokbase;      // [G] tokens emitted before the segment (member-relative)
    uint32_t *walked;           // [P] walked tokens (member region starts at pos_base)
    uint32_t *hist;             // [M * NSYM]
    uint32_t *ntok;             // [M]
    uint32_t *codes;            // [M * NSYM] reversed code | length << 16
    uint32_t *prefix;           // [M * PREFIX_WORDS]
    uint32_t *prefix_bits;      // [M]
    uint64_t *mem_bytes;        // [M] bytes of the member's output
    uint32_t *mem_crc;          // [M]
    uint64_t *mem_off;          // [M] output offset
    uint64_t *chunk_off;        // [C]
    uint64_t *chunk_len;        // [C]
    uint64_t *total;            // [1]
    uint64_t *em_state;         // [G] k_emit_tiled look-back: flag << 62 | bits
    uint64_t *em_join;          // [G] k_emit_tiled: parts added << 32 | bits of a word two slots share
    uint32_t *em_tick;          // [M] k_emit_tiled: token slots taken of each member
    uint8_t *out;               // output bytes (4-byte aligned)
    uint32_t strict;            // walk rule: 0 lazy (production), 1 strict, 2 lazy with a two-token
                                // lookahead (the opt-in ratio mode)
    uint32_t gzip;              // 1: gzip members, 0: raw DEFLATE (encode_block)
    uint32_t window;            // 32768
    uint32_t sub;               // member size (k_plan_batch)
    // chunk descriptors the device builds the tables above from (k_plan_batch),
    // or null when the host uploads the tables: [0] the chunk count C, [1]
    // unused, then C pairs {byte offset of the chunk in `in`, length}
    const uint64_t *cdesc;
};

}  // namespace gpz
