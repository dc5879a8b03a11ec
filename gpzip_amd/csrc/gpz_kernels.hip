// gpz_kernels.hip -- gfx950 kernels of the gzip compress path.
//
// Pipeline per batch (one launch each, all gridded so one instance can be
// replayed from a HIP graph for any batch that fits the workspace):
//   k_plan_batch  (device-built tables) member and segment tables from the
//                 chunk descriptors.
//   k_seg_table   per 4 KiB segment: the hash set as a bitmap with ranks, the
//                 first/last position of each hash, and the raw CRC-32.
//   k_match_find  one 1024-thread workgroup per segment: exact segmented-hash
//                 candidates for the segment's positions from an LDS copy of
//                 its 36 KiB window and of each back segment's table,
//                 candidate verification and extension, per-position tokens
//                 and, for the per-segment walk, the segment's walk summary
//                 (exit/count for every possible entry offset).
//   walk          per segment below 512 members: k_chain (one wave per member
//                 chains the summaries into each segment's entry and token
//                 base) and k_walk_emit (one workgroup per segment marks the
//                 walked positions, compacts them, counts the histogram); per
//                 member from 512 members up: k_walk_member (each member cut
//                 into parts walked at once by the waves of a workgroup).
//   k_huffman     one wave per member: exact heap-order Huffman code lengths,
//                 canonical codes, dynamic block header, member size and CRC.
//   k_scan        member output offsets and per-chunk offsets/lengths.
//   emission      k_emit (one workgroup per member: per-token bit lengths,
//                 workgroup scan, LDS word assembly, coalesced stores) from 512
//                 members up; k_emit_tiled (one workgroup per 4,096-token slot
//                 of a member, all at once) below.
//
// Reference semantics restated here (paths relative to the reference repo,
// crates/gpzip- prefix dropped):
//   hash3/build   codec-gpu/src/gpu/lz77_hash_build.wgsl:36-54
//   lookup        codec-gpu/src/gpu/lz77_hash_lookup.wgsl:45-119
//   lazy walk     codec-gpu/src/gpu/lz77.rs:260-284
//   strict walk   codec-gpu/src/gpu/walk.rs:512-525
//   Huffman       codec-gpu/src/gpu/huffman.rs:26-152
//   block writer  codec-gpu/src/gpu/deflate.rs:254-335, 752-959
//   gzip framing  codec-gpu/src/gpu/deflate.rs:962-980
#include <hip/hip_runtime.h>

#include "gpz_internal.h"

namespace gpz {

// ---------------------------------------------------------------------------
// constants

__constant__ uint32_t c_crc_tab[1024];  // reflected CRC-32 slicing-by-4 tables T0 .. T3
__constant__ uint32_t c_xpowb[4096];    // x^(8*b) mod P, reflected, b < 4096 (bytes within a segment)
__constant__ uint32_t c_segpow[4097];   // x^(8*4096*j) mod P, reflected (segments of 4 KiB)

__constant__ uint16_t c_len_base[29] = {3,  4,  5,  6,  7,  8,  9,  10, 11,  13,  15,  17,  19,  23, 27,
                                        31, 35, 43, 51, 59, 67, 83, 99, 115, 131, 163, 195, 227, 258};
__constant__ uint16_t c_dist_base[30] = {1,    2,    3,    4,    5,    7,    9,    13,    17,    25,
                                         33,   49,   65,   97,   129,  193,  257,  385,   513,   769,
                                         1025, 1537, 2049, 3073, 4097, 6145, 8193, 12289, 16385, 24577};
__constant__ uint8_t c_meta_order[19] = {16, 17, 18, 0, 8, 7, 9, 6, 10, 5, 11, 4, 12, 3, 13, 2, 14, 1, 15};

// ---------------------------------------------------------------------------
// small helpers

// Length 3..258 -> index 0..28 of the RFC 1951 length rows (symbol 257+idx).
__device__ __forceinline__ uint32_t len_idx(uint32_t len) {
    if (len <= 10) return len - 3;
    if (len == 258) return 28;
    uint32_t l = len - 3;
    uint32_t k = 31 - __clz(l);
    return 4 * (k - 1) + ((l >> (k - 2)) & 3);
}
__device__ __forceinline__ uint32_t len_extra(uint32_t idx) {
    return (idx < 8 || idx == 28) ? 0 : (idx - 4) >> 2;
}
// Distance 1..32768 -> symbol 0..29.
__device__ __forceinline__ uint32_t dist_sym(uint32_t d) {
    if (d <= 4) return d - 1;
    uint32_t x = d - 1;
    uint32_t k = 31 - __clz(x);
    return 2 * k + ((x >> (k - 1)) & 1);
}
__device__ __forceinline__ uint32_t dist_extra(uint32_t sym) { return sym < 4 ? 0 : (sym >> 1) - 1; }

// Carry-less multiply modulo the CRC-32 polynomial, reflected representation
// (bit 31 holds the x^0 coefficient).
__device__ __forceinline__ uint32_t gf2_mul(uint32_t a, uint32_t b) {
    uint32_t r = 0;
#pragma unroll
    for (int i = 0; i < 32; i++) {
        r ^= b & (0u - ((a >> (31 - i)) & 1u));
        b = (b >> 1) ^ (0xEDB88320u & (0u - (b & 1u)));
    }
    return r;
}

// Workgroup exclusive scan (u32) for blocks of BLOCK threads (multiple of 64).
template <int BLOCK>
__device__ __forceinline__ uint32_t block_scan(uint32_t v, uint32_t *tmp, uint32_t *total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    if (wid == 0) {
        uint32_t t = lane < BLOCK / 64 ? tmp[lane] : 0;
#pragma unroll
        for (int d = 1; d < BLOCK / 64; d <<= 1) {
            uint32_t y = __shfl_up(t, d, 64);
            if (lane >= d) t += y;
        }
        if (lane < BLOCK / 64) tmp[lane] = t;
    }
    __syncthreads();
    uint32_t base = wid ? tmp[wid - 1] : 0;
    *total = tmp[BLOCK / 64 - 1];
    __syncthreads();
    return base + x - v;
}

// Inclusive scan over the 64 lanes of a wave with DPP row shifts and row
// broadcasts (no lane-index registers, unlike a shuffle-based scan): rows of
// 16 lanes first, then lane 15 into row 1 and row 3, then lane 31 into rows 2
// and 3.  Lanes whose source is outside the row read the `old` operand, 0.
__device__ __forceinline__ uint32_t wave_scan_incl(uint32_t v) {
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);  // row_shr:1
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);  // row_shr:2
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);  // row_shr:4
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);  // row_shr:8
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);  // row_bcast:15
    v += (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);  // row_bcast:31
    return v;
}

// Workgroup exclusive scan built on wave_scan_incl; tmp holds BLOCK / 64
// words.  Ends with a barrier, so tmp may be reused at once.
template <int BLOCK>
__device__ __forceinline__ uint32_t block_scan_dpp(uint32_t v, uint32_t *tmp, uint32_t *total) {
    const uint32_t wid = threadIdx.x >> 6;
    const uint32_t inc = wave_scan_incl(v);
    if ((threadIdx.x & 63) == 63) tmp[wid] = inc;
    __syncthreads();
    uint32_t before = 0, tot = 0;
#pragma unroll
    for (uint32_t j = 0; j < BLOCK / 64; j++) {
        const uint32_t t = tmp[j];
        before += j < wid ? t : 0;
        tot += t;
    }
    *total = tot;
    __syncthreads();
    return before + inc - v;
}

template <int BLOCK>
__device__ __forceinline__ uint64_t block_scan64(uint64_t v, uint64_t *tmp, uint64_t *total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint64_t x = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        uint64_t y = __shfl_up(x, d, 64);
        if (lane >= d) x += y;
    }
    if (lane == 63) tmp[wid] = x;
    __syncthreads();
    if (wid == 0) {
        uint64_t t = lane < BLOCK / 64 ? tmp[lane] : 0;
#pragma unroll
        for (int d = 1; d < BLOCK / 64; d <<= 1) {
            uint64_t y = __shfl_up(t, d, 64);
            if (lane >= d) t += y;
        }
        if (lane < BLOCK / 64) tmp[lane] = t;
    }
    __syncthreads();
    uint64_t base = wid ? tmp[wid - 1] : 0;
    *total = tmp[BLOCK / 64 - 1];
    __syncthreads();
    return base + x - v;
}

// Values that are the same on every lane (workgroup-uniform loads) moved to
// scalar registers.
__device__ __forceinline__ uint32_t uni(uint32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ uint64_t uni64(uint64_t v) {
    return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}
// One dword at a wave-uniform index through the scalar cache.  The wait is on
// lgkmcnt, so vector loads still in flight (prefetches) are not waited for.
__device__ __forceinline__ uint32_t sload_u32(const uint32_t *base, uint32_t i) {
    const uint32_t *p = reinterpret_cast<const uint32_t *>(uni64(reinterpret_cast<uint64_t>(base + i)));
    uint32_t r;
    asm volatile("s_load_dword %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(r) : "s"(p) : "memory");
    return r;
}
// Member descriptor m, read by a whole workgroup (m must be uniform), with
// one scalar load into scalar registers.  The compiler cannot prove the
// descriptor array unwritten inside kernels that store to global memory, so
// it would otherwise load it into vector registers (and spill them).  The
// wait is inside the asm block: the outputs are complete when it ends.
typedef uint32_t u32x8 __attribute__((ext_vector_type(8)));
__device__ __forceinline__ MemberDesc load_member(const MemberDesc *mem, uint32_t m) {
    const MemberDesc *p = mem + m;
    u32x8 r;
    asm volatile("s_load_dwordx8 %0, %1, 0x0\n\ts_waitcnt lgkmcnt(0)" : "=s"(r) : "s"(p) : "memory");
    MemberDesc d;
    d.in_off = (uint64_t)r[0] | ((uint64_t)r[1] << 32);
    d.pos_base = (uint64_t)r[2] | ((uint64_t)r[3] << 32);
    d.n = r[4];
    d.seg_base = r[5];
    d.nseg = r[6];
    d.chunk = r[7];
    return d;
}

// Spread consecutive work items so that each XCD (blocks are dealt
// round-robin to the 8 XCDs) walks a contiguous range: neighbouring segments
// share 8/9 of their window, so they should share an L2.  XCD x takes the
// P / 8 (+1 for x < P % 8) items from x * (P / 8) + min(x, P % 8): a bijection
// of [0, P) for any grid size P.
__device__ __forceinline__ uint32_t xcd_block() {
    const uint32_t b = blockIdx.x, P = gridDim.x;
    const uint32_t x = b & 7u, q = P >> 3, r = P & 7u;
    return x * q + min(x, r) + (b >> 3);
}

// Bytes x..x+3 of an LDS byte array stored as u32 words.
__device__ __forceinline__ uint32_t ld4(const uint32_t *win, uint32_t x) {
    uint32_t w = x >> 2;
    return __builtin_amdgcn_alignbyte(win[w + 1], win[w], x & 3u);
}

// ---------------------------------------------------------------------------
// walk helpers
//
// The lazy rule (lz77.rs:264-282) gives every position a successor:
//   literal                                  -> x + 1
//   match followed by a strictly longer match -> x + 1 (demoted to a literal)
//   otherwise                                 -> x + len
// (strict rule, walk.rs:512-525: x + len for every match).  The walk is the
// orbit of the member's position 0.  Two levels: inside each 64-position
// block a wave doubles successors with lane shuffles (no barriers) to get,
// for every position, the first orbit point past the block and the 64-bit
// set of positions it visits inside the block; across blocks the workgroup
// follows those block exits (a segment has 64 blocks: at most 64 hops).

// Orders one wave's LDS stores before its later LDS loads (wave-private
// data: no workgroup barrier).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Successor of segment position x.  next_known: tokL[L] (and tokL[L+1] for
// rule 2) hold the next segment's first tokens, 0 past the member end.
// absorb_last: the last position's successor depends on the next segment and
// is left to k_chain (self-loop).  Rule 2 is the opt-in ratio mode (no
// reference counterpart): a match is demoted when the next position holds a
// longer match or the one after it a match longer by two or more; it is only
// walked with the next tokens known.
__device__ __forceinline__ uint32_t succ_of(const uint32_t *tokL, uint32_t x, uint32_t L, bool next_known,
                                            uint32_t strict, bool absorb_last, bool &dem) {
    dem = false;
    const uint32_t t = tokL[x], len = t >> 16;
    if (len == 0) return x + 1;
    if (strict == 2) {
        const uint32_t n1 = (x + 1 < L || next_known) ? tokL[x + 1] >> 16 : 0;
        const uint32_t n2 = (x + 2 < L || next_known) ? tokL[x + 2] >> 16 : 0;
        if (n1 > len || n2 > len + 1) {
            dem = true;
            return x + 1;
        }
        return x + len;
    }
    if (strict) return x + len;
    if (x + 1 < L || next_known) {
        const uint32_t nl = tokL[x + 1] >> 16;
        if (nl != 0 && nl > len) {
            dem = true;
            return x + 1;
        }
        return x + len;
    }
    return absorb_last ? x : x + len;
}

// Wave-level doubling inside the 64-position block starting at bs.  J is the
// lane's successor (segment-relative).  Returns the first orbit point at or
// past the block end (or the self-loop position) and the visited set.
__device__ __forceinline__ uint32_t block_walk(uint32_t J, uint32_t bs, uint32_t L, bool valid, bool selfloop,
                                               uint64_t &mask) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t lim = min(64u, L - bs);
    uint32_t j = valid ? J - bs : 64u;
    uint64_t M = (valid && !selfloop) ? (1ull << lane) : 0ull;
#pragma unroll
    for (int r = 0; r < 6; r++) {
        const bool inside = j < lim;
        if (!__ballot(inside)) break;  // every lane has left the block (long matches)
        const int src = inside ? (int)j : (int)lane;
        const uint32_t jj = (uint32_t)__shfl((int)j, src, 64);
        const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)M, src, 64);
        const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(M >> 32), src, 64);
        if (inside) {
            j = jj;
            M |= ((uint64_t)hi << 32) | lo;
        }
    }
    mask = M;
    return bs + j;
}

// The same doubling with the number of positions visited inside the block in
// place of the set: the sets joined in a round are disjoint stretches of the
// orbit (a self-loop adds nothing), so counts add.
__device__ __forceinline__ uint32_t block_walk_count(uint32_t J, uint32_t bs, uint32_t L, bool valid, bool selfloop,
                                                     uint32_t &cnt) {
    const uint32_t lane = threadIdx.x & 63;
    const uint32_t lim = min(64u, L - bs);
    uint32_t j = valid ? J - bs : 64u;
    uint32_t c = (valid && !selfloop) ? 1u : 0u;
#pragma unroll
    for (int r = 0; r < 6; r++) {
        const bool inside = j < lim;
        const int src = inside ? (int)j : (int)lane;
        const uint32_t jj = (uint32_t)__shfl((int)j, src, 64);
        const uint32_t cc = (uint32_t)__shfl((int)c, src, 64);
        if (inside) {
            j = jj;
            c += cc;
        }
    }
    cnt = c;
    return bs + j;
}

// ---------------------------------------------------------------------------
// k_seg_table: per 4 KiB segment, the oldest and newest position of every
// hash3 value in the segment -- the reference's seg_oldest/seg_newest pair
// (lz77_hash_build.wgsl:45-54) restricted to the hashes that occur -- as a
// compact list of (hash | first << 16, last) entries (offsets in the
// segment), built once per segment and streamed by the 9 segments whose
// window covers it.  Also computes the segment's raw CRC-32.

constexpr uint32_t EMPTY = 0xFFFFFFFFu;

// hash3 of lz77_hash_build.wgsl:36-43 on three bytes
__device__ __forceinline__ uint32_t hash3(uint32_t b0, uint32_t b1, uint32_t b2) {
    return ((b0 << 16 | b1 << 8 | b2) * 0x9E3779B1u) >> 16;
}

// k_seg_table: the reference's build pass (lz77_hash_build.wgsl:36-54,
// oldest/newest p+1 per (hash, segment) by global atomics) restated per
// segment in LDS: the set of 16-bit hashes of the segment as a bitmap, their
// ranks, and per hash (in hash order) its first and last position.  Also the
// raw CRC-32 of the segment's bytes.  BT threads (256, 512 or 1024; the
// autotuner picks), PPT = 4096 / BT positions each.
template <int BT>
__global__ __launch_bounds__(BT) void k_seg_table(Work w) {
    constexpr uint32_t PPT = SEG / BT;           // positions per thread
    constexpr uint32_t NW = (PPT + 2 + 3) / 4;   // words holding them and the two bytes after
    constexpr uint32_t WPT = 2048 / BT;          // bitmap words per thread
    __shared__ __align__(16) uint32_t bm[2048];
    __shared__ uint16_t rk[2048];
    __shared__ uint32_t fst[SEG];
    __shared__ uint32_t lst[SEG];
    __shared__ uint32_t crc_tab[4][256];
    __shared__ uint32_t tmp[16];
    const uint32_t tid = threadIdx.x;
    const uint32_t G = uni(w.counts->nseg);
    for (uint32_t i = tid; i < 1024; i += BT) crc_tab[i >> 8][i & 255] = c_crc_tab[i];
    // bytes x0 .. x0 + 4*NW - 1 of a segment (buffers are padded past every
    // member), loaded one segment ahead so the loads are in flight while the
    // current one is processed; lanes of wave 0 also load the 64 bytes of
    // their CRC run (64 lane .. 64 lane + 63)
    const uint32_t x0 = tid * PPT;
    auto fetch = [&](uint32_t gg, uint32_t (&d)[NW], uint4 (&cq)[4]) {
#pragma unroll
        for (uint32_t j = 0; j < NW; j++) d[j] = 0;
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) cq[j] = make_uint4(0, 0, 0, 0);
        if (gg >= G) return;
        const MemberDesc mm = load_member(w.mem, sload_u32(w.seg_mem, gg));
        const uint32_t SS = (gg - mm.seg_base) << SEG_LOG2, LL = min(SS + SEG, mm.n) - SS;
        if (x0 < LL) {
            const uint32_t *src = reinterpret_cast<const uint32_t *>(w.in + mm.in_off + SS + x0);
#pragma unroll
            for (uint32_t j = 0; j < NW; j++) d[j] = src[j];
        }
        if (tid < 64 && 64u * tid < LL) {
            // member starts are 16-byte aligned, and at most 63 bytes past a
            // member's end are read (the input's 256-byte read slack)
            const uint4 *src = reinterpret_cast<const uint4 *>(w.in + mm.in_off + SS + 64u * tid);
#pragma unroll
            for (uint32_t j = 0; j < 4; j++) cq[j] = src[j];
        }
    };
    uint32_t nx[NW];
    uint4 ncq[4];
    fetch(xcd_block(), nx, ncq);
    for (uint32_t g = xcd_block(); g < G; g += gridDim.x) {
        const MemberDesc md = load_member(w.mem, sload_u32(w.seg_mem, g));
        const uint32_t n = md.n, s = g - md.seg_base;
        const uint32_t S = s << SEG_LOG2, E = min(S + SEG, n), L = E - S;
        uint32_t *T = w.segtab + (uint64_t)g * TAB_WORDS;
        uint32_t d[NW];
#pragma unroll
        for (uint32_t j = 0; j < NW; j++) d[j] = nx[j];
        uint32_t cw[16];
#pragma unroll
        for (uint32_t j = 0; j < 4; j++) {
            cw[4 * j] = ncq[j].x;
            cw[4 * j + 1] = ncq[j].y;
            cw[4 * j + 2] = ncq[j].z;
            cw[4 * j + 3] = ncq[j].w;
        }
        auto byte_at = [&](uint32_t i) { return (d[i >> 2] >> (8 * (i & 3))) & 0xFFu; };
#pragma unroll
        for (uint32_t r = 0; r < WPT; r++) bm[tid + r * BT] = 0;
        __syncthreads();
        // raw CRC-32 of the segment on wave 0 alone (the other waves skip its
        // instructions): lane l takes bytes 64 l .. 64 l + 63, four at a time
        // through the slicing-by-4 tables (Kounavis and Berry), then folds its
        // CRC by x^(8 * bytes after them); reduced over the wave at the end
        uint32_t crc = 0;
        if (tid < 64 && 64u * tid < L) {
            const uint32_t cnt = min(64u, L - 64u * tid);
#pragma unroll
            for (uint32_t j = 0; j < 16; j++) {
                if (4 * j + 4 <= cnt) {
                    const uint32_t c = crc ^ cw[j];
                    crc = crc_tab[3][c & 0xFF] ^ crc_tab[2][(c >> 8) & 0xFF] ^ crc_tab[1][(c >> 16) & 0xFF] ^
                          crc_tab[0][c >> 24];
                } else {
#pragma unroll
                    for (uint32_t b = 0; b < 3; b++)
                        if (4 * j + b < cnt) crc = crc_tab[0][(crc ^ (cw[j] >> (8 * b))) & 0xFF] ^ (crc >> 8);
                }
            }
            crc = gf2_mul(crc, c_xpowb[L - 64u * tid - cnt]);
        }
        uint32_t hh[PPT];
        uint32_t hvm = 0;
#pragma unroll
        for (uint32_t i = 0; i < PPT; i++) {
            const uint32_t x = x0 + i;
            const bool hv = x < L && S + x + 2 < n;
            hh[i] = hash3(byte_at(i), byte_at(i + 1), byte_at(i + 2));
            hvm |= (uint32_t)hv << i;
            if (hv) atomicOr(&bm[hh[i] >> 5], 1u << (hh[i] & 31));
        }
        // the bytes are in registers now: the next segment's loads go out
        // without a wait for them
        asm volatile("" ::: "memory");
        fetch(g + gridDim.x, nx, ncq);
        __syncthreads();
        uint32_t D;  // distinct hashes in the segment
        {
            // ranks of words WPT*t .. WPT*t + WPT - 1; bitmap and ranks go out as they are
            uint32_t q[WPT], c[WPT], tot = 0;
#pragma unroll
            for (uint32_t j = 0; j < WPT; j++) {
                q[j] = bm[WPT * tid + j];
                c[j] = __popc(q[j]);
                tot += c[j];
            }
            const uint32_t ex = block_scan_dpp<BT>(tot, tmp, &D);
            uint32_t r = ex;
#pragma unroll
            for (uint32_t j = 0; j < WPT; j++) {
                rk[WPT * tid + j] = (uint16_t)r;
                T[TAB_BITS + WPT * tid + j] = q[j];
                if (j & 1) T[TAB_RANK + (WPT * tid + j) / 2] = (r << 16) | (r - c[j - 1]);
                r += c[j];
            }
            if (tid == 0) w.seg_ntab[g] = D;
            for (uint32_t id = tid; id < D; id += BT) {
                fst[id] = EMPTY;
                lst[id] = 0;
            }
        }
        __syncthreads();
#pragma unroll
        for (uint32_t i = 0; i < PPT; i++) {
            if (!((hvm >> i) & 1u)) continue;
            const uint32_t h = hh[i];
            const uint32_t id = rk[h >> 5] + __popc(bm[h >> 5] & ((1u << (h & 31)) - 1u));
            atomicMin(&fst[id], x0 + i);
            atomicMax(&lst[id], x0 + i);
        }
        __syncthreads();
        for (uint32_t id = tid; id < D; id += BT) T[TAB_ENT + id] = (fst[id] << 12) | lst[id];
        if (tid < 64) {
            uint32_t c = crc;
#pragma unroll
            for (int dd = 32; dd >= 1; dd >>= 1) c ^= __shfl_xor(c, dd, 64);
            if (tid == 0) w.seg_crc[g] = c;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_match_find: per-position tokens for one 4 KiB segment.
//
// The segment's window (32 KiB back, the segment, 258+ bytes ahead) is staged
// in LDS.  Then for k = 0..8 (the reference's segment order,
// lz77_hash_lookup.wgsl:85-112; older segments only give distances > 32768,
// rejected at :50) the whole table of segment s-k -- 65536-bit presence map,
// ranks, rank-ordered (first, last) entries -- is copied to LDS, each position
// finds its hash's entry there (the reference's slot for (hash, segment)), and
// tries newest, then oldest while best < 16, stopping at 258 (:96-110).  The
// copy for s-k-1 travels while s-k is searched.  69 KB of LDS: two workgroups
// per CU.

// The walk summary of segment g from its tokens in LDS (tokL[0..L)); 1024
// threads.  P is a ping-pong pair of 4096 words exit | count << 16.  Ends
// before the caller's closing barrier.
// next_known: tokL[L], tokL[L+1] hold the next segment's first tokens (0 past
// the member end), so no position is absorbed.
__device__ __forceinline__ void seg_summary(const Work &w, uint32_t g, uint32_t L, bool last_seg, bool any_match,
                                            const uint32_t *tokL, uint32_t *P0, uint32_t *P1, bool next_known = false) {
    const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const bool absorb_last = w.strict != 1 && !last_seg && !next_known;
    uint32_t *out = w.wsum + (uint64_t)g * SPAN;
    // the last position is a self-loop only when it holds a match
    const uint32_t absorb = (absorb_last && (tokL[L - 1] >> 16) != 0) ? L - 1 : 0xFFFFu;
    if (!any_match) {
        // no matches: every position is walked
        for (uint32_t x = tid; x < L && x < SPAN; x += 1024) out[x] = L | ((L - x) << 16);
        return;
    }
    // level 1: 16 waves x 4 blocks; per position the exit from its 64-position
    // block and the tokens emitted inside it
    for (uint32_t b = wid; b < 64; b += 16) {
        const uint32_t bs = b * 64;
        if (bs >= L) break;
        const uint32_t x = bs + lane;
        const bool valid = x < L;
        bool dem;
        const uint32_t J = valid ? succ_of(tokL, x, L, next_known, w.strict, absorb_last, dem) : 0;
        uint32_t c;
        const uint32_t ex = block_walk_count(J, bs, L, valid, valid && x == absorb, c);
        if (valid) P0[x] = ex | (c << 16);
    }
    __syncthreads();
    // level 2: pointer doubling over the block exits, three rounds (8 hops
    // per pointer); then each entry x < 258 follows the remaining pointers.
    // Every exit lies in a later block, so at most 8 more hops reach the
    // segment end or the absorbing last position (its own exit, count 0).
    // Three rounds over all positions and a short chase on 258 lanes cost
    // fewer instructions than six rounds.
    uint32_t *cur = P0, *nxt = P1;
    for (int r = 0; r < 3; r++) {
        for (uint32_t x = tid; x < L; x += 1024) {
            const uint32_t v = cur[x], y = v & 0xFFFFu;
            uint32_t nv = v;
            if (y < L) {
                const uint32_t v2 = cur[y];
                nv = (v2 & 0xFFFFu) | (((v >> 16) + (v2 >> 16)) << 16);
            }
            nxt[x] = nv;
        }
        __syncthreads();
        uint32_t *t = cur;
        cur = nxt;
        nxt = t;
    }
    for (uint32_t x = tid; x < L && x < SPAN; x += 1024) {
        const uint32_t v = cur[x];
        uint32_t y = v & 0xFFFFu, c = v >> 16;
        for (int h = 0; h < 64 && y < L; h++) {
            const uint32_t v2 = cur[y], y2 = v2 & 0xFFFFu;
            if (y2 == y) break;  // the absorbing last position
            c += v2 >> 16;
            y = y2;
        }
        out[x] = y | (c << 16);
    }
}

constexpr int K1_T = 1024;
constexpr int WIN_WORDS = 9344;  // 37,376 bytes: 32 KiB back + 4 KiB segment + 258 + slack

// A candidate resolved over its first 16 bytes (lz77_hash_lookup.wgsl:45-67:
// rejected -- 0 -- when cand >= p, dist > window or the first three bytes
// differ).  Returns the exact length, capped at lim = min(258, n - p), when it
// is below 16 or lim <= 16; otherwise 16 | PEND: the length is at least 16 and
// its exact value is left to ext_len.  own0 / own1 hold bytes p .. p+7.
// FAST: the caller knows the candidate lies before p within the window and
// that lim is 258, so the range checks and the limit compares fold away.
constexpr uint32_t PEND = 0x10000u;
template <bool FAST>
__device__ __forceinline__ uint32_t cand_short(const uint32_t *win, uint32_t WB, uint32_t lim_in, uint32_t p,
                                               uint32_t cand, uint32_t window, uint32_t own0, uint32_t own1) {
    const uint32_t lim = FAST ? MAX_MATCH : lim_in;
    if (!FAST && (cand >= p || p - cand > window)) return 0;
    const uint32_t c = cand - WB, ic = c >> 2, sc = c & 3u;
    const uint32_t w0 = win[ic], w1 = win[ic + 1], w2 = win[ic + 2];
    const uint32_t x0 = own0 ^ __builtin_amdgcn_alignbyte(w1, w0, sc);
    if (x0 & 0x00FFFFFFu) return 0;
    if (x0) return 3;
    const uint32_t x1 = own1 ^ __builtin_amdgcn_alignbyte(w2, w1, sc);
    if (x1) return min(4 + ((uint32_t)__builtin_ctz(x1) >> 3), lim);
    if (lim <= 8) return lim;
    // bytes 8 .. 15: own from the window, the candidate's dwords continue
    const uint32_t a = p - WB + 8, ia = a >> 2, sa = a & 3u;
    const uint32_t w3 = win[ic + 3], w4 = win[ic + 4];
    const uint32_t a0 = win[ia], a1 = win[ia + 1], a2 = win[ia + 2];
    const uint32_t y0 = __builtin_amdgcn_alignbyte(a1, a0, sa) ^ __builtin_amdgcn_alignbyte(w3, w2, sc);
    const uint32_t y1 = __builtin_amdgcn_alignbyte(a2, a1, sa) ^ __builtin_amdgcn_alignbyte(w4, w3, sc);
    if (y0 | y1)
        return min(8 + (y0 ? (uint32_t)__builtin_ctz(y0) >> 3 : 4 + ((uint32_t)__builtin_ctz(y1) >> 3)), lim);
    return lim <= 16 ? lim : (16u | PEND);
}

// Exact match length between window offsets a (position) and c (candidate)
// that are known to agree over `len` bytes, capped at lim: 8 bytes per step,
// three dwords of each side realigned with alignbyte, one exit.
__device__ __forceinline__ uint32_t ext_len(const uint32_t *win, uint32_t a, uint32_t c, uint32_t len,
                                            uint32_t lim) {
    a += len;
    c += len;
    const uint32_t sa = a & 3u, sc = c & 3u;
    const uint32_t *pa = win + (a >> 2), *pc = win + (c >> 2);
    uint32_t j = 0, x0, x1;
    do {
        x0 = __builtin_amdgcn_alignbyte(pa[j + 1], pa[j], sa) ^ __builtin_amdgcn_alignbyte(pc[j + 1], pc[j], sc);
        x1 = __builtin_amdgcn_alignbyte(pa[j + 2], pa[j + 1], sa) ^
             __builtin_amdgcn_alignbyte(pc[j + 2], pc[j + 1], sc);
        j += 2;
    } while ((x0 | x1) == 0 && len + 4 * j < lim);
    len += 4 * j - 8;
    len += x0 ? (uint32_t)__builtin_ctz(x0) >> 3 : (x1 ? 4 + ((uint32_t)__builtin_ctz(x1) >> 3) : 8);
    return len > lim ? lim : len;
}

// MINB: 8 gives two resident workgroups per CU (64 VGPRs, some spills), 4
// gives one (128 VGPRs); the autotuner picks.
template <int MINB>
__global__ __launch_bounds__(K1_T, MINB) void k_match_find(Work w, int do_summary) {
    __shared__ __align__(16) uint32_t win[WIN_WORDS];
    // the hash table of the segment being searched (own or back), copied whole:
    // bitmap, ranks and its entries (layout of k_seg_table), +4 words so an
    // absent hash's entry index (at most 4096) reads inside the array
    __shared__ __align__(16) uint32_t tab[TAB_WORDS + 4];
    __shared__ uint32_t kq[K1_T];    // per wave: 64-entry queue of pending candidates
    // Workgroup votes, one word each: [k] back segment k, [15] the walk summary.  A
    // wave with a set predicate writes 1 before a barrier; everyone reads after it.
    // (__syncthreads_or builds a flat work-item id from the y/z ids and the dispatch
    // packet's sizes; at 64 VGPRs that state spilled and cost 4 % of this kernel.)
    __shared__ uint32_t kvote[16];
    const uint32_t tid = threadIdx.x;
    const uint32_t G = uni(w.counts->nseg);
    // segment g-k's table -> tab through the direct global-to-LDS path (no data
    // registers): the bitmap and ranks (TAB_ENT words) and the segment's D
    // entries, 16 bytes per lane, lane i of a wave at its base + 16 i
    auto copy_table = [&](uint32_t gk) {
        uint32_t t = tid;
        asm volatile("" : "+v"(t));  // lane offsets recomputed per copy, not held
        const uint32_t *T = w.segtab + (uint64_t)gk * TAB_WORDS;
        const uint32_t n16 = (TAB_ENT + uni(w.seg_ntab[gk]) + 3u) >> 2;
        static_assert(TAB_WORDS <= 8 * K1_T, "two 16-byte copies per lane");
#pragma unroll
        for (uint32_t r = 0; r < 2; r++) {
            const uint32_t u = r * K1_T + t;
            const uint32_t wb = __builtin_amdgcn_readfirstlane(r * K1_T + (t & ~63u));
            if (u < n16)
                __builtin_amdgcn_global_load_lds((const void __attribute__((address_space(1))) *)(T + 4u * u),
                                                 (void __attribute__((address_space(3))) *)(tab + 4u * wb), 16, 0, 0);
        }
    };
    for (uint32_t g = xcd_block(); g < G; g += gridDim.x) {
        const MemberDesc md = load_member(w.mem, uni(w.seg_mem[g]));
        const uint32_t n = md.n, s = g - md.seg_base;
        const uint32_t S = s << SEG_LOG2;
        const uint32_t E = min(S + SEG, n), L = E - S;
        const uint32_t WB = S >= 32768u ? S - 32768u : 0u;
        const uint32_t WE = min(n, E + 272u);
        const uint32_t nb = WE - WB;
        const uint32_t kmax = min(s, 8u);

        copy_table(g);  // the own table (k = 0) travels while the window loads
        // ---- window -> LDS (zero beyond the member end)
        {
            const uint8_t *src = w.in + md.in_off + WB;
            for (uint32_t v = tid; v < WIN_WORDS / 4; v += K1_T) {
                const uint32_t b0 = v * 16u;
                uint4 q = make_uint4(0, 0, 0, 0);
                if (b0 < nb) {
                    q = *reinterpret_cast<const uint4 *>(src + b0);
                    if (b0 + 16u > nb) {
                        const uint32_t rem = nb - b0;
                        uint32_t *qq = reinterpret_cast<uint32_t *>(&q);
#pragma unroll
                        for (int j = 0; j < 4; j++) {
                            const uint32_t lo = 4u * j;
                            if (lo >= rem) qq[j] = 0;
                            else if (lo + 4 > rem) qq[j] &= (1u << (8 * (rem - lo))) - 1u;
                        }
                    }
                }
                reinterpret_cast<uint4 *>(win)[v] = q;
            }
            if (tid < 16) kvote[tid] = 0;
        }
        __syncthreads();

        // ---- own positions: their hashes, two 16-bit hashes per register
        const uint32_t x0 = tid * 4u;
        const uint32_t lx = S - WB + x0;
        uint32_t hh0 = 0, hh1 = 0;
        uint32_t hvm = 0;  // bit i: position x0+i can start a match (p + 2 < n)
        const uint32_t own_lo0 = win[lx >> 2], own_hi0 = win[(lx >> 2) + 1];  // bytes x0 .. x0+7
        const uint32_t own_w20 = win[(lx >> 2) + 2];                           // bytes x0+8 .. x0+11
        {
            const uint64_t vv = (uint64_t)own_lo0 | ((uint64_t)own_hi0 << 32);
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const uint32_t x = x0 + i;
                const bool hv = x < L && S + x + 2u < n;
                const uint32_t h = hash3((uint32_t)(vv >> (8 * i)) & 0xFF, (uint32_t)(vv >> (8 * i + 8)) & 0xFF,
                                         (uint32_t)(vv >> (8 * i + 16)) & 0xFF);
                if (i < 2) hh0 |= h << (16 * i);
                else hh1 |= h << (16 * (i - 2));
                hvm |= (uint32_t)hv << i;
            }
        }

        uint32_t st[4] = {0, 0, 0, 0};  // best len << 16 | dist per position
        for (uint32_t k = 0; k <= kmax; k++) {
            // tab holds segment s-k's table once this wave's copies have landed
            // and every wave has passed the barrier; from k = 1 every wave has
            // also voted on k - 1
            __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0)
            __syncthreads();
            if (k && !kvote[k - 1]) break;
            // Values the loop does not change pass through empty asm operands
            // each step, so what is derived from them (lookup addresses and
            // masks from the packed hashes; byte-aligned own words and limits
            // per position of the unrolled passes) is recomputed per step:
            // hoisted out of the loop, it would be held across it and spill.
            asm volatile("" : "+v"(hh0), "+v"(hh1));
            uint32_t own_lo = own_lo0, own_hi = own_hi0, own_w2 = own_w20;
            asm volatile("" : "+v"(own_lo), "+v"(own_hi), "+v"(own_w2));
            uint32_t xk = x0;
            asm volatile("" : "+v"(xk));
            // each position's (first, last) in segment s-k, or EMPTY when its
            // hash does not occur there (the reference's empty slot)
            uint32_t fv[4], pd[4] = {0, 0, 0, 0};
            {
                const uint8_t *tb = reinterpret_cast<const uint8_t *>(tab);
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    // hash h = bits 16 (i & 1) .. +15 of the packed pair: bitmap
                    // word h >> 5; the bit-field builtins take h & 31 from the
                    // low five bits of `hb`
                    const uint32_t hw = i < 2 ? hh0 : hh1;
                    const uint32_t wi = __builtin_amdgcn_ubfe(hw, 16 * (i & 1) + 5, 11);
                    const uint32_t hb = (i & 1) ? hw >> 16 : hw;
                    uint32_t e = EMPTY;
                    if (((hvm >> i) & 1u) && (st[i] >> 16) < MAX_MATCH) {
                        const uint32_t wd = tab[TAB_BITS + wi];
                        if (__builtin_amdgcn_ubfe(wd, hb, 1)) {
                            const uint32_t r = *reinterpret_cast<const uint16_t *>(tb + 4 * TAB_RANK + 2 * wi);
                            e = *reinterpret_cast<const uint32_t *>(
                                tb + 4 * TAB_ENT + 4 * (r + __popc(__builtin_amdgcn_ubfe(wd, 0, hb))));
                        }
                    }
                    fv[i] = e;
                }
            }
            // every wave has its entries: tab is free for the next table
            __syncthreads();
            // evaluate segment s-k in the reference order per position: newest,
            // then oldest while best < 16.  Pass 1 resolves every candidate over
            // its first 16 bytes, position by position.  A candidate whose 16 bytes agree is
            // left pending: best is then >= 16 whatever its length, so the
            // oldest rule is already decided, and a position has at most one
            // pending candidate.  Pass 2 extends the pending candidates of the
            // wave densely, 64 at a time, through a wave-private LDS queue.
            const uint32_t base = (s - k) << SEG_LOG2;
            auto pass1 = [&](auto fast) {
            constexpr bool FAST = decltype(fast)::value;
#pragma unroll
            for (uint32_t i = 0; i < 4; i++) {
                const uint32_t v = fv[i];
                uint32_t cs = st[i], pend = 0;
                if (v != EMPTY) {
                    const uint32_t p = S + xk + i;
                    const uint32_t lim = FAST ? MAX_MATCH : min(n - p, MAX_MATCH);
                    const uint32_t own0 = i ? __builtin_amdgcn_alignbyte(own_hi, own_lo, i) : own_lo;
                    const uint32_t own1 = i ? __builtin_amdgcn_alignbyte(own_w2, own_hi, i) : own_hi;
                    const uint32_t cn = base + (v & 0xFFFu), co = base + ((v >> 12) & 0xFFFu);
                    const uint32_t ln = cand_short<FAST>(win, WB, lim, p, cn, w.window, own0, own1);
                    if (ln & PEND) {
                        pend = 0x80000000u | (cn - WB);
                    } else {
                        if (ln > (cs >> 16)) cs = (ln << 16) | (p - cn);
                        // the oldest entry is tried only while best < 16
                        // (lz77_hash_lookup.wgsl:102); equal to the newest it
                        // cannot give a strictly longer match
                        if ((cs >> 16) < OLDEST_SKIP && co != cn) {
                            const uint32_t lo = cand_short<FAST>(win, WB, lim, p, co, w.window, own0, own1);
                            if (lo & PEND) pend = 0xC0000000u | (co - WB);
                            else if (lo > (cs >> 16)) cs = (lo << 16) | (p - co);
                        }
                    }
                }
                st[i] = cs;
                pd[i] = pend;
            }
            };
            // interior segments (every match limit is 258) and back segments 1 .. 7
            // (candidates before the position, within the window) take the
            // check-free pass
            if (E + MAX_MATCH <= n && k >= 1 && 4095u + k * SEG <= w.window)
                pass1(std::integral_constant<bool, true>{});
            else
                pass1(std::integral_constant<bool, false>{});
            {
                // pass 2 (wave-local: no workgroup barrier, the trip count differs per wave)
                const uint32_t lane = tid & 63u;
                const uint64_t lt = (1ull << lane) - 1;
                uint32_t *q = kq + (tid & ~63u);
                uint64_t bm[4];
                uint32_t bs[4], T = 0;
#pragma unroll
                for (int i = 0; i < 4; i++) {
                    bm[i] = __ballot(pd[i] != 0);
                    bs[i] = T;
                    T += __popcll(bm[i]);
                }
                for (uint32_t r0 = 0; r0 < T; r0 += 64) {
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const uint32_t idx = bs[i] + __popcll(bm[i] & lt) - r0;
                        if (pd[i] && idx < 64) q[idx] = (x0 + i) | ((pd[i] & 0xFFFFu) << 12);
                    }
                    wave_lds_sync();
                    if (r0 + lane < T) {
                        const uint32_t it = q[lane];
                        const uint32_t x = it & 0xFFFu;
                        const uint32_t lim = min(n - (S + x), MAX_MATCH);
                        q[lane] = ext_len(win, S - WB + x, it >> 12, 16, lim);
                    }
                    wave_lds_sync();
#pragma unroll
                    for (int i = 0; i < 4; i++) {
                        const uint32_t idx = bs[i] + __popcll(bm[i] & lt) - r0;
                        if (pd[i] && idx < 64) {
                            const uint32_t len = q[idx];
                            const uint32_t d = S + x0 + i - (WB + (pd[i] & 0xFFFFu));
                            // a pending oldest is strictly longer than best (< 16)
                            if (((pd[i] >> 30) & 1u) || len > (st[i] >> 16)) st[i] = (len << 16) | d;
                        }
                    }
                    wave_lds_sync();
                }
            }
            // the next table travels while the waves vote and meet (issued here,
            // after both passes, it holds no registers through them)
            if (k < kmax) copy_table(g - k - 1);
            int active = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) active |= ((hvm >> i) & 1u) && (st[i] >> 16) < MAX_MATCH;
            if (__ballot(active) && (tid & 63u) == 0) kvote[k] = 1u;
        }

        // ---- tokens
        uint32_t t4[4];
        {
            const uint32_t ob = win[lx >> 2];  // re-read: keeping own_lo alive costs a spill
#pragma unroll
            for (int i = 0; i < 4; i++) t4[i] = st[i] ? st[i] : (ob >> (8 * i)) & 0xFF;
        }
        uint32_t *tg = w.tok + md.pos_base + S + x0;
        if (x0 + 4 <= L) {
            *reinterpret_cast<uint4 *>(tg) = make_uint4(t4[0], t4[1], t4[2], t4[3]);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++)
                if (x0 + i < L) tg[i] = t4[i];
        }
        if (do_summary) {
            // the segment's walk summary from the same tokens, in the LDS the
            // match search no longer needs: tokens over the window, one
            // exit | count array after them, the other over tab
            __syncthreads();
            uint32_t *tokL = win;
            int anym = 0;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                tokL[x0 + i] = t4[i];
                anym |= (x0 + i < L) && (t4[i] >> 16) != 0;
            }
            if (__ballot(anym) && (tid & 63u) == 0) kvote[15] = 1u;
            __syncthreads();
            seg_summary(w, g, L, E == n, kvote[15] != 0, tokL, win + SEG, tab);
        }
        __syncthreads();
    }
}

// k_walk_summary: for every entry offset x < 258 of a segment, the first walk
// position at or past the segment end (relative to the segment start) and the
// number of tokens emitted before it.  When the segment is not the member's
// last, the lazy rule at its last position depends on the next segment's
// first token; that position is a self-loop here and k_chain resolves it.
constexpr int KW_T = 1024;


__global__ __launch_bounds__(KW_T) void k_walk_summary(Work w) {
    __shared__ uint32_t tokL[SEG + 2];
    __shared__ uint32_t P[2][SEG];
    const uint32_t tid = threadIdx.x;
    const uint32_t G = uni(w.counts->nseg);
    // rule 2 reads the next segment's first two tokens (all tokens exist when
    // this kernel runs), so no position is left to k_chain
    const bool lookahead = w.strict == 2;
    for (uint32_t g = xcd_block(); g < G; g += gridDim.x) {
        const MemberDesc md = load_member(w.mem, uni(w.seg_mem[g]));
        const uint32_t n = md.n, s = g - md.seg_base;
        const uint32_t S = s << SEG_LOG2, E = min(S + SEG, n), L = E - S;
        int anym = 0;
        for (uint32_t x = tid; x < L + 2; x += KW_T) {
            const uint32_t t = S + x < n ? w.tok[md.pos_base + S + x] : 0;
            tokL[x] = t;
            anym |= x < L && (t >> 16) != 0;
        }
        const int any_match = __syncthreads_or(anym);
        seg_summary(w, g, L, E == n, any_match != 0, tokL, P[0], P[1], lookahead);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_chain: one wave per member

__global__ __launch_bounds__(64) void k_chain(Work w) {
    __shared__ __align__(16) uint32_t sum[32 * SPAN];
    __shared__ uint32_t btok[64];
    const uint32_t M = uni(w.counts->nmem);
    const uint32_t lane = threadIdx.x;
    for (uint32_t m = blockIdx.x; m < M; m += gridDim.x) {
        const MemberDesc md = load_member(w.mem, m);
        const uint32_t nseg = md.nseg;
        const bool cached = nseg <= 32;
        if (cached) {
            // the member's summaries (8-byte aligned: SPAN is even), eight
            // loads in flight per lane before the first wait
            const uint2 *src = reinterpret_cast<const uint2 *>(w.wsum + (uint64_t)md.seg_base * SPAN);
            uint2 *dst = reinterpret_cast<uint2 *>(sum);
            const uint32_t n2 = nseg * (SPAN / 2);
            for (uint32_t i0 = lane; i0 < n2; i0 += 64 * 8) {
                uint2 v[8];
#pragma unroll
                for (int j = 0; j < 8; j++) v[j] = src[min(i0 + 64 * j, n2 - 1)];
                asm volatile("" ::: "memory");
#pragma unroll
                for (int j = 0; j < 8; j++)
                    if (i0 + 64 * j < n2) dst[i0 + 64 * j] = v[j];
            }
            if (lane + 1 < nseg) {
                uint64_t e = md.pos_base + ((uint64_t)(lane + 1) << SEG_LOG2);
                btok[2 * lane] = w.tok[e - 1];
                btok[2 * lane + 1] = w.tok[e];
            }
        }
        for (uint32_t i = lane; i < NSYM; i += 64) w.hist[(uint64_t)m * NSYM + i] = 0;
        __syncthreads();
        if (lane == 0) {
            uint32_t x = 0, base = 0;
            for (uint32_t s = 0; s < nseg; s++) {
                const uint32_t g = md.seg_base + s;
                const uint32_t S = s << SEG_LOG2;
                const uint32_t L = min(SEG, md.n - S);
                w.seg_entry[g] = x;
                w.seg_tokbase[g] = base;
                if (x >= L) {
                    x -= L;
                    continue;
                }
                // entry offsets stay below SPAN for token lengths <= 258
                const uint32_t xe = x < SPAN ? x : SPAN - 1;
                uint32_t v = cached ? sum[s * SPAN + xe] : w.wsum[(uint64_t)g * SPAN + xe];
                uint32_t jx = v & 0xFFFFu, c = v >> 16;
                if (jx < L) {  // absorbed at the segment's last position
                    uint32_t t, t1;
                    if (cached) {
                        t = btok[2 * s];
                        t1 = btok[2 * s + 1];
                    } else {
                        uint64_t e = md.pos_base + S + L;
                        t = w.tok[e - 1];
                        t1 = w.tok[e];
                    }
                    uint32_t len = t >> 16, nl = t1 >> 16;
                    uint32_t succ = (nl != 0 && nl > len) ? L : (L - 1 + len);
                    base += c + 1;
                    x = succ - L;
                } else {
                    base += c;
                    x = jx - L;
                }
            }
            w.ntok[m] = base;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_walk_emit: the walked positions of a segment from its known entry, their
// tokens compacted into the member's token stream, and the histogram.
// One lane follows the block exits (level 1) from the entry, marking the
// orbit's block entries (level 2); each marked entry's visited set comes from
// a second level-1 pass.

constexpr int K3_T = 1024;

// Tokens of segment g for this thread (positions 4t .. 4t+3; 0 past the
// segment end), the input bytes at those positions and, on thread 0, the next
// segment's first token.  Issued one segment ahead so the loads are in flight
// while the current one is walked; every wait before their use is a scalar
// (lgkmcnt) one, which leaves them outstanding.
struct SegPrefetch {
    uint4 q;
    uint32_t raw, tn, tn2;
};
__device__ __forceinline__ void tok_prefetch(const Work &w, uint32_t G, uint32_t g, SegPrefetch &f) {
    f.q = make_uint4(0, 0, 0, 0);
    f.raw = 0;
    f.tn = 0;
    f.tn2 = 0;
    if (g >= G) return;
    const MemberDesc md = load_member(w.mem, sload_u32(w.seg_mem, g));
    const uint32_t S = (g - md.seg_base) << SEG_LOG2, E = min(S + SEG, md.n), L = E - S;
    const uint32_t *tg = w.tok + md.pos_base + S;
    const uint8_t *ib = w.in + md.in_off + S;
    const uint32_t x = threadIdx.x * 4u;
    if (x + 4 <= L) {
        f.q = *reinterpret_cast<const uint4 *>(tg + x);
        // member starts are 16-byte aligned, segments 4 KiB apart
        f.raw = *reinterpret_cast<const uint32_t *>(ib + x);
    } else {
        if (x < L) f.q.x = tg[x];
        if (x + 1 < L) f.q.y = tg[x + 1];
        if (x + 2 < L) f.q.z = tg[x + 2];
        if (x < L) f.raw = ib[x];
        if (x + 1 < L) f.raw |= (uint32_t)ib[x + 1] << 8;
        if (x + 2 < L) f.raw |= (uint32_t)ib[x + 2] << 16;
    }
    if (threadIdx.x == 0 && E < md.n) f.tn = tg[L];
    if (threadIdx.x == 0 && E + 1 < md.n) f.tn2 = tg[L + 1];
}

__global__ __launch_bounds__(K3_T) void k_walk_emit(Work w) {
    __shared__ __align__(16) uint32_t tokL[SEG + 4];
    __shared__ uint64_t Mv[SEG];     // per position: positions walked inside its block from it
    __shared__ uint16_t Ja[SEG];     // per position: exit from its block
    __shared__ uint64_t vb[64];      // walked positions of each block
    __shared__ uint64_t demb[64];    // demoted matches of each block (lazy rule)
    __shared__ uint32_t cntb[64];
    __shared__ uint32_t hist[NSYM];
    __shared__ uint32_t rawL[SEG / 4];  // the segment's input bytes
    const uint32_t tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    const uint64_t lt = (1ull << lane) - 1;
    const uint32_t G = uni(w.counts->nseg);
    SegPrefetch pf;
    tok_prefetch(w, G, xcd_block(), pf);
    for (uint32_t g = xcd_block(); g < G; g += gridDim.x) {
        const uint32_t m = sload_u32(w.seg_mem, g);
        const MemberDesc md = load_member(w.mem, m);
        const uint32_t n = md.n, s = g - md.seg_base;
        const uint32_t S = s << SEG_LOG2, E = min(S + SEG, n), L = E - S;
        const uint32_t e = sload_u32(w.seg_entry, g);
        if (e >= L) {  // the walk jumps over this segment (uniform)
            tok_prefetch(w, G, g + gridDim.x, pf);
            continue;
        }
        const bool next_known = E < n;
        const uint32_t tokbase = sload_u32(w.seg_tokbase, g);
        reinterpret_cast<uint4 *>(tokL)[tid] = pf.q;
        rawL[tid] = pf.raw;
        if (tid == 0 && L == SEG) {  // (a shorter segment is its member's last)
            tokL[SEG] = pf.tn;
            tokL[SEG + 1] = pf.tn2;
        }
        // the stores above wait for this segment's loads only if the next
        // segment's are not yet issued: keep them apart
        asm volatile("" ::: "memory");
        tok_prefetch(w, G, g + gridDim.x, pf);
        for (uint32_t i = tid; i < NSYM; i += K3_T) hist[i] = 0;
        if (tid < 64) vb[tid] = 0;
        __syncthreads();
        // level 1: per position its block exit and the positions walked
        // inside the block from it
        for (uint32_t b = wid; b < 64; b += K3_T / 64) {
            const uint32_t bs = b * 64;
            if (bs >= L) break;
            const uint32_t x = bs + lane;
            const bool valid = x < L;
            bool dem = false;
            const uint32_t J = valid ? succ_of(tokL, x, L, next_known, w.strict, false, dem) : 0;
            uint64_t M;
            const uint32_t ex = block_walk(J, bs, L, valid, false, M);
            if (valid) {
                Ja[x] = (uint16_t)ex;
                Mv[x] = M;
            }
            const uint64_t dm = __ballot(valid && dem);
            if (lane == 0) demb[b] = dm;
        }
        __syncthreads();
        // level 2: one lane follows the block exits from the entry (at most
        // 64 hops); the entry of each block it meets selects the block's set
        if (tid == 0) {
            uint32_t y = e;
            while (y < L) {
                vb[y >> 6] = Mv[y];
                y = Ja[y];
            }
        }
        __syncthreads();
        if (wid == 0) {  // exclusive scan of the 64 block counts
            const uint32_t c = __popcll(vb[lane]);
            cntb[lane] = wave_scan_incl(c) - c;
        }
        __syncthreads();
        uint32_t *outp = w.walked + md.pos_base + tokbase;
        for (uint32_t b = wid; b < 64; b += K3_T / 64) {
            const uint32_t bs = b * 64;
            if (bs >= L) break;
            const uint64_t V = vb[b];
            if (!((V >> lane) & 1)) continue;
            const uint32_t x = bs + lane;
            uint32_t t = tokL[x];
            if ((demb[b] >> lane) & 1) t = (rawL[x >> 2] >> (8 * (x & 3))) & 0xFFu;
            outp[cntb[b] + __popcll(V & lt)] = t;
            const uint32_t len = t >> 16, v = t & 0xFFFFu;
            if (len == 0) {
                if (v < 256) atomicAdd(&hist[v], 1u);
            } else if (len >= 3 && len <= 258 && v >= 1 && v <= 32768) {
                atomicAdd(&hist[257 + len_idx(len)], 1u);
                atomicAdd(&hist[DSYM0 + dist_sym(v)], 1u);
            }
        }
        __syncthreads();
        for (uint32_t i = tid; i < NSYM; i += K3_T)
            if (hist[i]) atomicAdd(&w.hist[(uint64_t)m * NSYM + i], hist[i]);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_walk_member: a wave walks a member (or a part of one) in position order, 64
// positions (one block) at a time, from position 0 -- the reference's walk
// (lz77.rs:260-284; strict rule walk.rs:512-525; rule 2 the opt-in ratio
// mode).  Per block the lanes compute their successors; the orbit through the
// block is then followed on the scalar unit, one step per match taken: the
// next position whose successor is not the next position is the first set
// bit of a ballot, its successor one readlane.  The walked set is rebuilt on
// the vector unit from the matches taken (a position is walked when it is at
// or past the entry and at or past the end of the last match taken before
// it: a wave max-scan).  The walked tokens are compacted into the member's
// token stream and counted into a wave-private histogram that is stored
// whole at the end (no global atomics, no clearing pass).  Without
// per-segment walk summaries there is no chaining pass either.  Tokens and
// input bytes are loaded one group of KM_G blocks ahead.  This is the walk of
// large batches: a part's blocks run in sequence, so a batch needs many
// members to fill the chip.

constexpr int KM_T = 256;     // four waves per workgroup (four members on one wave each, two on two, one on four)
constexpr uint32_t KM_G = 8;  // blocks of 64 positions per group
// a group's input bytes are 8 per lane (WalkGroup::r), 512 bytes: at most 8 blocks
static_assert(KM_G <= 8, "a group's input bytes cover at most 8 blocks");
constexpr uint32_t HIST_USED = 316;  // litlen 0..285, distance 286..315; 319 takes invalid tokens

// One group of KM_G blocks in registers: lane l's token of each block, of the
// first block of the next group (the walk rule reads ahead), and 8 input
// bytes (the group's 512 bytes spread 8 per lane, for demoted matches).
struct WalkGroup {
    uint32_t t[KM_G + 1];
    uint2 r;
};

// Lane i receives lane i + 1's v, lane 63 receives `last` (DPP wave_shl:1).
__device__ __forceinline__ uint32_t lane_next(uint32_t v, uint32_t last) {
    return (uint32_t)__builtin_amdgcn_update_dpp((int)last, (int)v, 0x130, 0xf, 0xf, false);
}

// Maximum of v over the lanes below this one (0 on lane 0): DPP row shifts and
// row broadcasts as in wave_scan_incl, then wave_shr:1.
__device__ __forceinline__ uint32_t wave_max_excl(uint32_t v) {
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false));  // row_shr:1
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false));  // row_shr:2
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false));  // row_shr:4
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false));  // row_shr:8
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false));  // row_bcast:15
    v = max(v, (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false));  // row_bcast:31
    return (uint32_t)__builtin_amdgcn_update_dpp(0, (int)v, 0x138, 0xf, 0xf, false);        // wave_shr:1
}

// Histogram slots of a walked token, without branches: the literal or length
// symbol (HIST_USED + 3 for a token outside the alphabet) and, for a match,
// the distance symbol.
__device__ __forceinline__ uint32_t tok_bin1(uint32_t to, bool &match) {
    const uint32_t tl = to >> 16, tv = to & 0xFFFFu;
    const bool okl = tl == 0 && tv < 256;
    match = tl >= 3 && tl <= 258 && tv >= 1 && tv <= 32768;
    const uint32_t l = tl - 3, kl = 31 - __clz(max(l, 8u));
    const uint32_t li = tl <= 10 ? l : (tl == 258 ? 28u : 4 * (kl - 1) + ((l >> (kl - 2)) & 3u));
    return okl ? tv : (match ? 257 + li : HIST_USED + 3);
}
__device__ __forceinline__ uint32_t tok_bin2(uint32_t to) {
    const uint32_t d = to & 0xFFFFu, x = max(d, 5u) - 1, k = 31 - __clz(x);
    return DSYM0 + (d <= 4 ? d - 1 : 2 * k + ((x >> (k - 1)) & 1u));
}

// The walk over blocks [b0, b1) of member md from entry e0 (member-relative, at
// or past block b0's start).  Returns the first walked position at or past
// block b1's start.  EMIT: the walked tokens are stored at out (in order; their
// count in *cnt) and counted into hist; without it the walk only follows the
// orbit.
template <int RULE, bool EMIT>
__device__ __forceinline__ uint32_t walk_span(const Work &w, const MemberDesc &md, uint32_t e0, uint32_t b0,
                                              uint32_t b1, uint32_t *out, uint32_t *cnt, uint32_t *hist,
                                              uint32_t *stg) {
    const uint32_t lane = threadIdx.x & 63u;
    const uint32_t n = md.n;
    if (b1 <= b0 || e0 >= (b1 << 6)) {  // (uniform) nothing of the span is walked
        if (EMIT) *cnt = 0;
        return e0;
    }
    // the span's addresses are computed here, not hoisted over an earlier
    // walk (where they would hold registers through its whole loop)
    b0 = uni(b0);
    asm volatile("" : "+s"(b0));
    const uint8_t *tk = reinterpret_cast<const uint8_t *>(w.tok + md.pos_base);
    const uint8_t *ib = w.in + md.in_off;
    const uint32_t ng = (b1 - b0 + KM_G - 1) / KM_G;
    // Loads are unconditional, their addresses clamped into the member (token
    // at the last position, input bytes at the last 8-byte boundary); what
    // lies past the member end is zeroed or ignored where it is used.  Token
    // offsets are 32-bit byte offsets (a member holds at most 2^24 tokens), so
    // each load takes a scalar base and one offset register.
    const uint32_t xl = n - 1, rl = (n - 1) & ~7u;
    auto load = [&](WalkGroup &G, uint32_t gi) {
        const uint32_t p0 = (b0 + gi * KM_G) << 6;
#pragma unroll
        for (uint32_t k = 0; k <= KM_G; k++)
            G.t[k] = *reinterpret_cast<const uint32_t *>(tk + (min(p0 + 64u * k + lane, xl) << 2));
        if (EMIT && RULE != 1) G.r = *reinterpret_cast<const uint2 *>(ib + min(p0 + 8u * lane, rl));
    };
    // Loads and stores share one counter on this target, so the compiler's
    // waits before the use of a loaded register drain it; there is one
    // explicit drain per group instead, with the group's loads issued a whole
    // group earlier and the previous group's stores issued just as early:
    // walked tokens are staged in LDS (stg, 64 * KM_G words per wave) and
    // stored at the start of the next group.
    // next walked position (member-relative); uniform, kept in a scalar register
    uint32_t e = (uint32_t)__builtin_amdgcn_readfirstlane((int)e0);
    uint32_t outc = 0;   // tokens stored so far
    uint32_t scnt = 0;   // tokens staged
    auto flush = [&]() {
        if (!EMIT) return;
        wave_lds_sync();
        for (uint32_t i = lane; i < scnt; i += 64) out[outc + i] = stg[i];
        outc += scnt;
        scnt = 0;
        wave_lds_sync();  // the reads are done before the next group's writes
    };
    const uint32_t pend = min(b1 << 6, n);  // the span's end (positions)
    WalkGroup A, B;  // A: the group being walked (loaded), B: the next (in flight)
    load(B, 0);
    for (uint32_t gi = 0; gi < ng; gi++) {
        __builtin_amdgcn_s_waitcnt(0x0F70);  // vmcnt(0): group gi and the stores before it
        A = B;
        if (gi + 1 < ng) load(B, gi + 1);
        flush();
#pragma unroll
        for (uint32_t k = 0; k < KM_G; k++) {
            const uint32_t bs = (b0 + gi * KM_G + k) << 6;
            // (uniform) past the span end, or a match jumps over the block
            if (e >= bs + 64u || bs >= pend) continue;
            const bool in = bs + lane < n;
            const uint32_t t = in ? A.t[k] : 0u, len = t >> 16;
            // tokens at x + 1 (and x + 2): the next lane, past lane 63 the next
            // block's first lanes; 0 past the member end
            const uint32_t rn0 = (uint32_t)__builtin_amdgcn_readlane((int)A.t[k + 1], 0);
            const uint32_t tn0 = bs + 64u < n ? rn0 : 0u;
            const uint32_t t1 = lane_next(t, tn0);
            bool dem = false;
            if (RULE == 0) {
                dem = len != 0 && (t1 >> 16) > len;
            } else if (RULE == 2) {
                const uint32_t rn1 = (uint32_t)__builtin_amdgcn_readlane((int)A.t[k + 1], 1);
                const uint32_t t2 = lane_next(t1, bs + 65u < n ? rn1 : 0u);
                dem = len != 0 && ((t1 >> 16) > len || (t2 >> 16) > len + 1);
            }
            const bool lit = len == 0 || dem;
            const uint32_t jt = lane + (lit ? 1u : len);  // successor, block-relative
            // where a run of next-position successors stops: a match taken, or
            // lane 63 (its successor leaves the block either way)
            const uint64_t stop = ~__ballot(lit) | (1ull << 63);
            // the walk's next stop after a match taken here: the first stop at
            // or past the successor, or the successor itself once it leaves
            // the block (every stop mask has bit 63, so one exists)
            const uint32_t kt = jt >= 64u ? jt : jt + (uint32_t)__builtin_ctzll(stop >> jt);
            const uint32_t c0 = e - bs;
            uint32_t cur = c0 + (uint32_t)__builtin_ctzll(stop >> c0);
            uint64_t jm = 0;  // positions where the walk takes a match
            do {
                asm("s_bitset1_b64 %0, %1" : "+s"(jm) : "s"(cur));
                cur = (uint32_t)__builtin_amdgcn_readlane((int)kt, (int)cur);
            } while (cur < 64u);
            e = bs + cur;
            if (!EMIT) continue;
            uint64_t V = __ballot(in) & (~0ull << c0);
            if (jm & (~0ull >> 1)) {
                // past a match taken at m the walk resumes at its successor
                const uint32_t r = wave_max_excl(((jm >> lane) & 1ull) ? jt : 0u);
                V &= __ballot(lane >= r);
            }
            uint32_t to = t;
            if (RULE != 1 && (__ballot(dem) & V)) {
                // a demoted match is its literal byte: position 64 k + lane of
                // the group is byte lane & 7 of lane 8 k + lane / 8
                const int src = (int)(8u * k + (lane >> 3));
                const uint32_t lo = (uint32_t)__shfl((int)A.r.x, src, 64);
                const uint32_t hi = (uint32_t)__shfl((int)A.r.y, src, 64);
                const uint32_t rb = (((lane & 4u) ? hi : lo) >> (8u * (lane & 3u))) & 0xFFu;
                if (dem) to = rb;
            }
            bool match;
            const uint32_t b1s = tok_bin1(to, match);
            const uint32_t b2s = tok_bin2(to);
            const uint32_t rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(V >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)V, 0u));
            if ((V >> lane) & 1ull) {
                stg[scnt + rank] = to;
                atomicAdd(&hist[b1s], 1u);
                if (match) atomicAdd(&hist[b2s], 1u);
            }
            scnt += (uint32_t)__popcll(V);
        }
    }
    flush();
    if (EMIT) *cnt = outc;
    return e;
}

// Walk state of one part of a member, shared by the waves of a workgroup.
struct PartState {
    uint32_t guess;  // the part's entry as found by its warm-up walk
    uint32_t exit;   // first walked position past the part
    uint32_t count;  // tokens the part emits
};

// One member on P waves (P = 1, 2 or 4; 4 / P members per 256-thread
// workgroup).  With P = 1 the member is walked in order by one wave.  With P >
// 1 its blocks are cut into P parts, walked at once: wave q first walks the
// KM_G blocks before its part without emitting, from that stretch's first
// position, which gives a guess of the part's entry (walks from different
// entries merge within a few positions on every corpus measured), then walks
// and emits the part from the guess into a staging area (the member's stretch
// of the segment tables, which the match finder no longer needs).  After a
// barrier the guesses are checked in part order against the true entries --
// part q's true entry is part q-1's exit -- and a part whose guess differs is
// walked again from the true entry, so the result is exact.  The staged tokens
// are then moved behind the previous parts' tokens, and the parts' histograms
// summed.  Every wave runs the same barriers, whether or not its member slot
// holds a member.
template <int RULE, int P>
__device__ __forceinline__ void walk_member(const Work &w, uint32_t M, uint32_t (*hs)[NSYM], uint32_t (*stg)[64 * KM_G],
                                            PartState (*ps)[4]) {
    constexpr uint32_t MPW = 4 / P;  // members per workgroup
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wv = uni(tid >> 6);
    const uint32_t slot = wv / P, q = wv % P;
    uint32_t *hist = hs[wv];
    for (uint32_t g0 = blockIdx.x * MPW; g0 < M; g0 += gridDim.x * MPW) {
        const uint32_t m = uni(g0 + slot);
        const bool has = m < M;
        MemberDesc md{};
        if (has) md = load_member(w.mem, m);
        const uint32_t n = has ? md.n : 0u;
        const uint32_t nb = (n + 63u) >> 6;
        const uint32_t pb = nb * q / P, pe = nb * (q + 1) / P;  // part q: blocks [pb, pe)
        for (uint32_t i = lane; i < NSYM; i += 64) hist[i] = 0;
        wave_lds_sync();
        uint32_t *dst = has ? w.walked + md.pos_base : nullptr;
        uint32_t *stage = (P > 1 && q > 0 && has) ? w.segtab + (uint64_t)md.seg_base * TAB_WORDS + (pb << 6) : dst;
        uint32_t cnt = 0, entry = pb << 6;
        if (P > 1 && n && q > 0) {
            const uint32_t wb = pb - min(KM_G, pb - nb * (q - 1) / P);
            entry = max(entry, walk_span<RULE, false>(w, md, wb << 6, wb, pb, nullptr, nullptr, hist, stg[wv]));
        }
        // round 0: every part walks from its guess; round r > 0: part r walks
        // again from its true entry (part r-1's exit) if its guess was wrong
        for (uint32_t r = 0; r < (uint32_t)P; r++) {
            bool go = r == 0;
            if (P > 1 && r > 0) {
                const uint32_t truth = uni(ps[slot][r - 1].exit);
                go = q == r && truth != entry;
                if (go) {
                    entry = truth;
                    for (uint32_t i = lane; i < NSYM; i += 64) hist[i] = 0;
                    wave_lds_sync();
                }
            }
            if (go && n) {
                const uint32_t ex = walk_span<RULE, true>(w, md, entry, pb, pe, stage, &cnt, hist, stg[wv]);
                if (P > 1 && lane == 0) ps[slot][q] = PartState{entry, ex, cnt};
            } else if (P > 1 && go && lane == 0) {
                ps[slot][q] = PartState{entry, entry, 0};
            }
            if (P > 1) __syncthreads();
        }
        if (P > 1) {
            // move the staged tokens behind the previous parts'
            if (n && q > 0) {
                uint32_t off = 0;
                for (uint32_t r = 0; r < q; r++) off += uni(ps[slot][r].count);
                for (uint32_t i = lane; i < cnt; i += 64) dst[off + i] = stage[i];
            }
            __syncthreads();  // every part's histogram is complete
            if (has) {
                for (uint32_t i = q * 64 + lane; i < NSYM; i += 64 * P) {
                    uint32_t v = 0;
#pragma unroll
                    for (uint32_t r = 0; r < (uint32_t)P; r++) v += hs[slot * P + r][i];
                    w.hist[(uint64_t)m * NSYM + i] = i < HIST_USED ? v : 0u;
                }
                if (q == 0 && lane == 0) {
                    uint32_t tot = 0;
                    for (uint32_t r = 0; r < (uint32_t)P; r++) tot += ps[slot][r].count;
                    w.ntok[m] = tot;
                }
            }
            __syncthreads();  // the histograms and part states are reused
        } else if (has) {
            wave_lds_sync();
            for (uint32_t i = lane; i < NSYM; i += 64) w.hist[(uint64_t)m * NSYM + i] = i < HIST_USED ? hist[i] : 0u;
            if (lane == 0) w.ntok[m] = cnt;
            wave_lds_sync();  // the histogram is cleared again for the next member
        }
    }
}

// Eight waves per SIMD (at most 64 registers of each kind; a few scalar values
// move to vector-register lanes): the walk is latency-bound, and 2,048 members
// on four waves each then run in one round on the 1,024 SIMDs.
template <int P>
__global__ __launch_bounds__(KM_T) __attribute__((amdgpu_waves_per_eu(8, 8))) void k_walk_member(Work w) {
    __shared__ uint32_t hs[KM_T / 64][NSYM];
    __shared__ uint32_t stg[KM_T / 64][64 * KM_G];
    __shared__ PartState ps[4][4];
    const uint32_t M = uni(w.counts->nmem);
    const uint32_t rule = uni(w.strict);
    if (rule == 0) walk_member<0, P>(w, M, hs, stg, ps);
    else if (rule == 1) walk_member<1, P>(w, M, hs, stg, ps);
    else walk_member<2, P>(w, M, hs, stg, ps);
}

// Histogram of caller-supplied walked tokens (stage-level gpz_encode_block).
__global__ __launch_bounds__(1024) void k_hist_tokens(Work w) {
    __shared__ uint32_t hist[NSYM];
    const uint32_t M = uni(w.counts->nmem);
    for (uint32_t m = blockIdx.x; m < M; m += gridDim.x) {
        const MemberDesc md = load_member(w.mem, m);
        for (uint32_t i = threadIdx.x; i < NSYM; i += 1024) hist[i] = 0;
        __syncthreads();
        const uint32_t nt = w.ntok[m];
        for (uint32_t i = threadIdx.x; i < nt; i += 1024) {
            uint32_t t = w.walked[md.pos_base + i], len = t >> 16, v = t & 0xFFFFu;
            if (len == 0) {
                atomicAdd(&hist[v & 0xFF], 1u);
            } else {
                atomicAdd(&hist[257 + len_idx(len)], 1u);
                atomicAdd(&hist[DSYM0 + dist_sym(v)], 1u);
            }
        }
        __syncthreads();
        for (uint32_t i = threadIdx.x; i < NSYM; i += 1024) w.hist[(uint64_t)m * NSYM + i] = hist[i];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_huffman: one wave per member

struct HuffScratch {
    // keys of the used leaves, the sorted leaves and the internal nodes: 32- or
    // 64-bit (build_lengths<K>), so sized for 64
    __align__(16) uint64_t ck[292];  // used leaves' keys (freq << 10 | id), then 4 x ~0
    uint64_t sl[288];                 // the same keys sorted
    uint64_t ik[288];                 // internal nodes' keys (freq << 10 | 1 << 9 | index)
    uint16_t hg[288];                 // internal nodes' heights
    uint16_t par[576];                // parent: leaf id -> par[id], internal j -> par[288 + j]
    uint16_t rb[290];                 // first internal node of each merge round
    uint8_t dep[288];                 // internal nodes' depths
};

__device__ __forceinline__ uint32_t uni_k(uint32_t v) { return uni(v); }
__device__ __forceinline__ uint64_t uni_k(uint64_t v) { return uni64(v); }
__device__ __forceinline__ uint32_t lane1_k(uint32_t v) { return __builtin_amdgcn_readlane(v, 1); }
__device__ __forceinline__ uint64_t lane1_k(uint64_t v) {
    return ((uint64_t)__builtin_amdgcn_readlane((uint32_t)(v >> 32), 1) << 32) |
           __builtin_amdgcn_readlane((uint32_t)v, 1);
}
__device__ __forceinline__ uint32_t shfl_k(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, 64); }
__device__ __forceinline__ uint64_t shfl_k(uint64_t v, int src) {
    return ((uint64_t)(uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64) << 32) | (uint32_t)__shfl((int)(uint32_t)v, src, 64);
}
__device__ __forceinline__ uint32_t shfl_xor_k(uint32_t v, int d) { return (uint32_t)__shfl_xor((int)v, d, 64); }
__device__ __forceinline__ uint64_t shfl_xor_k(uint64_t v, int d) {
    return ((uint64_t)(uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), d, 64) << 32) |
           (uint32_t)__shfl_xor((int)(uint32_t)v, d, 64);
}

// try_build_code_lengths (huffman.rs:26-60) for n <= 288 symbols on one
// wave.  The reference pops the minimum (freq, id) from a binary heap whose
// internal nodes get ids n, n+1, ...; merged frequencies never decrease and
// internal ids exceed every leaf id, so every pop takes the least item under
// the order (freq, leaf before internal, leaf id / creation index).  Items
// are keys freq << 10 | internal << 9 | index in that order, of type K: 32
// bits while every frequency sum stays below 2^22 (a member of fewer than
// 2^22 - 1 tokens), 64 bits otherwise.
//
// Merge rounds: the 64 least remaining items are the least 64 of the 64
// least leaves and the 64 least queued internal nodes, sorted by one bitonic
// merge across the wave.  Let N = S0 + S1 be the next internal node.  Every
// node made after it is larger, and an item is popped before N exactly when
// its frequency is at most N's, so the heap pops S0, S1, S2, ... up to the
// last such item before it ever takes a new node: those pairs are merged at
// once, one per lane.  A round merges at least one pair; a node's parent is
// made in a later round, so depths follow from the rounds in reverse.
// Frequencies are halved while the longest code exceeds `limit`
// (huffman.rs:44-58).
template <typename K>
__device__ void build_lengths(const uint32_t *f, int n, int limit, uint8_t *lens, HuffScratch &hs) {
    const uint32_t lane = threadIdx.x;
    const uint64_t lt = (1ull << lane) - 1;
    constexpr K NONE = ~(K)0;
    K *ck = reinterpret_cast<K *>(hs.ck);
    K *sl = reinterpret_cast<K *>(hs.sl);
    K *ik = reinterpret_cast<K *>(hs.ik);
    uint32_t fr[5];
#pragma unroll
    for (int j = 0; j < 5; j++) fr[j] = (int)(lane + 64 * j) < n ? f[lane + 64 * j] : 0u;
    for (;;) {
        // the used leaves' keys, compacted
        uint32_t u = 0;
#pragma unroll
        for (int j = 0; j < 5; j++) {
            const uint64_t m = __ballot(fr[j] != 0);
            if (fr[j]) ck[u + __popcll(m & lt)] = ((K)fr[j] << 10) | (lane + 64 * j);
            u += __popcll(m);
        }
        u = uni(u);
        if (u <= 1) {
#pragma unroll
            for (int j = 0; j < 5; j++)
                if ((int)(lane + 64 * j) < n) lens[lane + 64 * j] = fr[j] ? 1 : 0;
            __syncthreads();
            return;
        }
        if (lane < 4) ck[u + lane] = NONE;
        __syncthreads();
        // sorted by rank (keys are distinct)
        {
            K kk[5];
            uint32_t r[5];
#pragma unroll
            for (int j = 0; j < 5; j++) {
                kk[j] = lane + 64 * j < u ? ck[lane + 64 * j] : NONE;
                r[j] = 0;
            }
            for (uint32_t i = 0; i < u; i += 4) {
                const K a0 = ck[i], a1 = ck[i + 1], a2 = ck[i + 2], a3 = ck[i + 3];
#pragma unroll
                for (int j = 0; j < 5; j++) r[j] += (a0 < kk[j]) + (a1 < kk[j]) + (a2 < kk[j]) + (a3 < kk[j]);
            }
#pragma unroll
            for (int j = 0; j < 5; j++)
                if (kk[j] != NONE) sl[r[j]] = kk[j];
        }
        __syncthreads();
        // merge rounds
        uint32_t nl = 0, qh = 0, k = 0, nr = 0;
        while (k + 1 < u) {
            // leaves ascending, queued internal nodes descending: one bitonic
            // sequence of 128 whose lower half after the first exchange holds
            // the 64 least; half-cleaners sort it
            const K a = nl + lane < u ? sl[nl + lane] : NONE;
            const uint32_t bi = qh + 63 - lane;
            const K b = bi < k ? ik[bi] : NONE;
            K v = a < b ? a : b;
#pragma unroll
            for (int d = 32; d >= 1; d >>= 1) {
                const K o = shfl_xor_k(v, d);
                v = (lane & d) ? (v > o ? v : o) : (v < o ? v : o);
            }
            const K thr = (uni_k(v) >> 10) + (lane1_k(v) >> 10);
            const uint32_t T = __popcll(__ballot((v >> 10) <= thr)) >> 1;  // >= 1
            const K x = shfl_k(v, (int)(2 * lane) & 63);
            const K y = shfl_k(v, (int)(2 * lane + 1) & 63);
            if (lane == 0) hs.rb[nr] = (uint16_t)k;
            if (lane < T) {
                const uint32_t node = k + lane;
                const bool xi = (x >> 9) & 1u, yi = (y >> 9) & 1u;
                const uint32_t xs = (uint32_t)x & 0x1FFu, ys = (uint32_t)y & 0x1FFu;
                const uint32_t hx = xi ? hs.hg[xs] : 0u, hy = yi ? hs.hg[ys] : 0u;
                ik[node] = (((x >> 10) + (y >> 10)) << 10) | 0x200u | node;
                hs.hg[node] = (uint16_t)(max(hx, hy) + 1);
                hs.par[(xi ? 288u : 0u) + xs] = (uint16_t)node;
                hs.par[(yi ? 288u : 0u) + ys] = (uint16_t)node;
            }
            const uint32_t lu = __popcll(__ballot(lane < 2 * T && !((v >> 9) & 1u)));
            nl += lu;
            qh += 2 * T - lu;
            k += T;
            nr++;
            __syncthreads();
        }
        const uint32_t root = u - 2;
        if (hs.hg[root] > (uint32_t)limit) {
            // too long: halve every frequency > 1 and rebuild (huffman.rs:44-58)
#pragma unroll
            for (int j = 0; j < 5; j++)
                if (fr[j] > 1) fr[j] >>= 1;
            __syncthreads();
            continue;
        }
        // depths: the root's round last, then back to the first
        if (lane == 0) {
            hs.rb[nr] = (uint16_t)k;
            hs.dep[root] = 0;
        }
        __syncthreads();
        for (int r = (int)nr - 2; r >= 0; r--) {
            const uint32_t node = hs.rb[r] + lane;
            if (node < hs.rb[r + 1]) hs.dep[node] = hs.dep[hs.par[288 + node]] + 1;
            __syncthreads();
        }
#pragma unroll
        for (int j = 0; j < 5; j++)
            if ((int)(lane + 64 * j) < n) lens[lane + 64 * j] = fr[j] ? (uint8_t)(hs.dep[hs.par[lane + 64 * j]] + 1) : 0;
        __syncthreads();
        return;
    }
}

// canonical codes (huffman.rs:122-152), stored bit-reversed with the length
// (deflate.rs:876-884).  All 64 lanes: per-length counts and in-length ranks
// from ballots.
__device__ void canonical_rev(const uint8_t *lens, int n, uint32_t *codes) {
    const int lane = threadIdx.x;
    const uint64_t lt = (1ull << lane) - 1;
    uint32_t cnt[16];
#pragma unroll
    for (int b = 0; b < 16; b++) cnt[b] = 0;
    for (int c0 = 0; c0 < n; c0 += 64) {
        const uint32_t l = c0 + lane < n ? lens[c0 + lane] : 0;
#pragma unroll
        for (int b = 1; b < 16; b++) cnt[b] += __popcll(__ballot(l == (uint32_t)b));
    }
    uint32_t next[16];
    uint32_t code = 0;
    next[0] = 0;
#pragma unroll
    for (int b = 1; b < 16; b++) {
        code = (code + (b > 1 ? cnt[b - 1] : 0)) << 1;
        next[b] = code;
    }
    for (int c0 = 0; c0 < n; c0 += 64) {
        const uint32_t l = c0 + lane < n ? lens[c0 + lane] : 0;
        uint32_t mine = 0;
#pragma unroll
        for (int b = 1; b < 16; b++) {
            const uint64_t m = __ballot(l == (uint32_t)b);
            if (l == (uint32_t)b) mine = next[b] + __popcll(m & lt);
            next[b] += __popcll(m);
        }
        if (c0 + lane < n) codes[c0 + lane] = l ? ((__builtin_bitreverse32(mine) >> (32 - l)) | (l << 16)) : 0;
    }
    __syncthreads();
}

// Codes of one run of equal code lengths (rle_encode_lengths,
// deflate.rs:254-326): a zero run becomes 18s (11-138 zeros), then one 17
// (3-10), then single zeros; a run of v becomes v, then 16s (3-6 copies),
// then single v.  Entries are sym | extra bits << 8 | extra value << 16.
__device__ __forceinline__ uint32_t rle_count(uint32_t v, uint32_t R) {
    if (v == 0) {
        const uint32_t a = R / 138, b = R % 138;
        const uint32_t n18 = a + (b >= 11), r1 = b >= 11 ? 0 : b;
        return n18 + (r1 >= 3 ? 1 : r1);
    }
    const uint32_t r = R - 1;
    return 1 + r / 6 + ((r % 6) >= 3 ? 1 : r % 6);
}
__device__ __forceinline__ void rle_write(uint32_t v, uint32_t R, uint32_t *out) {
    uint32_t o = 0;
    if (v == 0) {
        const uint32_t a = R / 138, b = R % 138;
        for (uint32_t i = 0; i < a; i++) out[o++] = 18u | (7u << 8) | (127u << 16);
        uint32_t r1 = b;
        if (b >= 11) {
            out[o++] = 18u | (7u << 8) | ((b - 11) << 16);
            r1 = 0;
        }
        if (r1 >= 3) {
            out[o++] = 17u | (3u << 8) | ((r1 - 3) << 16);
        } else {
            for (uint32_t i = 0; i < r1; i++) out[o++] = 0;
        }
        return;
    }
    out[o++] = v;
    const uint32_t r = R - 1;
    for (uint32_t i = 0; i < r / 6; i++) out[o++] = 16u | (2u << 8) | (3u << 16);
    const uint32_t t = r % 6;
    if (t >= 3) {
        out[o++] = 16u | (2u << 8) | ((t - 3) << 16);
    } else {
        for (uint32_t i = 0; i < t; i++) out[o++] = v;
    }
}

// OR the low nb bits of v into an LSB-first bit buffer at bit position pos.
__device__ __forceinline__ void put_bits(uint32_t *buf, uint32_t pos, uint32_t v, uint32_t nb) {
    if (!nb) return;
    const uint32_t w = pos >> 5, sh = pos & 31u;
    atomicOr(&buf[w], v << sh);
    if (sh + nb > 32) atomicOr(&buf[w + 1], v >> (32 - sh));
}

__global__ __launch_bounds__(64) void k_huffman(Work w) {
    __shared__ uint32_t freq[NSYM];
    __shared__ uint8_t lens[NSYM];
    __shared__ uint32_t codes[NSYM];
    __shared__ uint8_t comb[320];
    __shared__ uint32_t rle[320];
    __shared__ uint32_t mfreq[20];
    __shared__ uint8_t mlens[20];
    __shared__ uint32_t mcodes[20];
    __shared__ uint32_t pbuf[PREFIX_WORDS + 2];
    __shared__ HuffScratch hs;
    const uint32_t M = uni(w.counts->nmem);
    const uint32_t lane = threadIdx.x;
    for (uint32_t m = blockIdx.x; m < M; m += gridDim.x) {
        const MemberDesc md = load_member(w.mem, m);
        for (uint32_t i = lane; i < NSYM; i += 64) freq[i] = w.hist[(uint64_t)m * NSYM + i];
        for (uint32_t s = lane; s < md.nseg; s += 64) {  // k_emit_tiled's slot state
            w.em_state[md.seg_base + s] = 0;
            w.em_join[md.seg_base + s] = 0;
        }
        if (lane == 0) w.em_tick[m] = 0;
        for (uint32_t i = lane; i < 20; i += 64) mfreq[i] = 0;
        for (uint32_t i = lane; i < PREFIX_WORDS + 2; i += 64) pbuf[i] = 0;
        __syncthreads();
        if (lane == 0) freq[256] += 1;  // end of block
        __syncthreads();
        // 32-bit heap keys hold frequency sums below 2^22: members of fewer
        // than 2^22 - 1 tokens (every member of the 128 KiB split)
        if (md.n + 1u < (1u << 22)) {
            build_lengths<uint32_t>(freq, 286, 15, lens, hs);
            build_lengths<uint32_t>(freq + DSYM0, 30, 15, lens + DSYM0, hs);
        } else {
            build_lengths<uint64_t>(freq, 286, 15, lens, hs);
            build_lengths<uint64_t>(freq + DSYM0, 30, 15, lens + DSYM0, hs);
        }
        {
            // single/zero distance-code fix-up and litlen fix-up (deflate.rs:855-867)
            const uint64_t dm = __ballot(lane < 30 && lens[DSYM0 + lane] != 0);
            uint32_t used_ll = 0;
            for (uint32_t i = lane; i < 286; i += 64) used_ll += lens[i] != 0;
            for (int d = 32; d >= 1; d >>= 1) used_ll += __shfl_xor(used_ll, d, 64);
            __syncthreads();
            if (lane == 0) {
                const uint32_t used_d = __popcll(dm);
                if (used_d == 0) {
                    lens[DSYM0] = 1;
                    lens[DSYM0 + 1] = 1;
                } else if (used_d == 1) {
                    const uint32_t only = __builtin_ctzll(dm);
                    lens[DSYM0 + only] = 1;
                    lens[DSYM0 + 1 - min(only, 1u)] = 1;
                }
                if (used_ll < 2) {
                    lens[0] = 1;
                    lens[256] = 1;
                }
            }
            __syncthreads();
        }
        canonical_rev(lens, 286, codes);
        canonical_rev(lens + DSYM0, 30, codes + DSYM0);
        // trim_to_min (deflate.rs:329-335): HLIT >= 257, HDIST >= 1
        uint32_t nlit, ndist;
        {
            uint32_t hl = 0, hd = 0;
            for (uint32_t i = lane; i < 286; i += 64)
                if (lens[i]) hl = max(hl, i + 1);
            if (lane < 30 && lens[DSYM0 + lane]) hd = lane + 1;
            for (int d = 32; d >= 1; d >>= 1) {
                hl = max(hl, (uint32_t)__shfl_xor(hl, d, 64));
                hd = max(hd, (uint32_t)__shfl_xor(hd, d, 64));
            }
            nlit = uni(max(hl, 257u));
            ndist = uni(max(hd, 1u));
        }
        const uint32_t nc = nlit + ndist;
        for (uint32_t i = lane; i < nc; i += 64) comb[i] = i < nlit ? lens[i] : lens[DSYM0 + i - nlit];
        __syncthreads();
        // RLE of the combined lengths, one run per starting lane: run starts
        // by ballot, run lengths from the next start, output slots by scan
        uint32_t K = 0;
        {
            uint64_t B[5];
#pragma unroll
            for (int j = 0; j < 5; j++) {
                const uint32_t p = lane + 64 * j;
                B[j] = __ballot(p < nc && (p == 0 || comb[p] != comb[p - 1]));
            }
            uint32_t carry = 0;
#pragma unroll
            for (int j = 0; j < 5; j++) {
                const uint32_t p = lane + 64 * j;
                const bool st = (B[j] >> lane) & 1;
                uint32_t R = 0, cnt = 0, v = 0;
                if (st) {
                    // next start after p
                    uint32_t nx = nc;
                    const uint64_t above = lane == 63 ? 0ull : B[j] & ~((2ull << lane) - 1);
                    if (above) {
                        nx = 64 * j + __builtin_ctzll(above);
                    } else {
#pragma unroll
                        for (int j2 = j + 1; j2 < 5; j2++)
                            if (nx == nc && B[j2]) nx = 64 * j2 + __builtin_ctzll(B[j2]);
                    }
                    R = nx - p;
                    v = comb[p];
                    cnt = rle_count(v, R);
                }
                const uint32_t inc = wave_scan_incl(cnt);
                if (st) rle_write(v, R, rle + carry + inc - cnt);
                carry += __builtin_amdgcn_readlane(inc, 63);
            }
            K = uni(carry);
        }
        __syncthreads();
        for (uint32_t r = lane; r < K; r += 64) atomicAdd(&mfreq[rle[r] & 0xFF], 1u);
        __syncthreads();
        build_lengths<uint32_t>(mfreq, 19, 7, mlens, hs);
        canonical_rev(mlens, 19, mcodes);
        // dynamic block header (deflate.rs:886-930): all fields at known
        // offsets, the RLE codes placed by a scan of their bit counts
        uint32_t hclen;
        {
            uint32_t h = 0;
            if (lane < 19 && mlens[c_meta_order[lane]]) h = lane + 1;
            for (int d = 32; d >= 1; d >>= 1) h = max(h, (uint32_t)__shfl_xor(h, d, 64));
            hclen = uni(max(h, 4u));
        }
        const uint32_t b0 = w.gzip ? 80u : 0u;
        if (lane == 0) {
            if (w.gzip) {
                // gzip header 1f 8b 08 00 00000000 00 ff (deflate.rs:965-972)
                pbuf[0] = 0x00088b1fu;
                pbuf[1] = 0;
                pbuf[2] = 0xff00u;
            }
            put_bits(pbuf, b0, 5u | ((nlit - 257) << 3) | ((ndist - 1) << 8) | ((hclen - 4) << 13), 17);
        }
        if (lane < hclen) put_bits(pbuf, b0 + 17 + 3 * lane, mlens[c_meta_order[lane]], 3);
        uint32_t pbits;
        {
            uint32_t pos = b0 + 17 + 3 * hclen;
            for (uint32_t r0 = 0; r0 < K; r0 += 64) {
                const uint32_t r = r0 + lane;
                uint32_t nb = 0, val = 0;
                if (r < K) {
                    const uint32_t e = rle[r], c = e & 0xFF, eb = (e >> 8) & 0xFF, ev = e >> 16;
                    const uint32_t cl = mlens[c];
                    nb = cl + eb;
                    val = (mcodes[c] & 0xFFFFu) | (ev << cl);
                }
                const uint32_t inc = wave_scan_incl(nb);
                if (r < K) put_bits(pbuf, pos + inc - nb, val, nb);
                pos += __builtin_amdgcn_readlane(inc, 63);
            }
            pbits = uni(pos);
        }
        __syncthreads();
        // publish codes / prefix, count body bits
        uint64_t body = 0;
        for (uint32_t i = lane; i < NSYM; i += 64) {
            uint32_t c = 0;
            if (i < 286 || (i >= DSYM0 && i < DSYM0 + 30)) c = codes[i];
            w.codes[(uint64_t)m * NSYM + i] = c;
            const uint32_t f = freq[i];
            if (!f) continue;
            uint32_t extra = 0;
            if (i >= 265 && i < 285) extra = len_extra(i - 257);
            if (i >= DSYM0 && i < DSYM0 + 30) extra = dist_extra(i - DSYM0);
            body += (uint64_t)f * ((c >> 16) + extra);
        }
        for (int d = 32; d >= 1; d >>= 1) body += __shfl_xor(body, d, 64);
        for (uint32_t i = lane; i < (pbits + 31) / 32; i += 64) w.prefix[(uint64_t)m * PREFIX_WORDS + i] = pbuf[i];
        // member CRC-32 from the segments' raw CRCs
        // (crc(A||B) = crc(A) * x^(8|B|) ^ crc(B) on raw CRCs; segment s is
        // followed by nseg-2-s full segments and the Lb bytes of the last)
        uint32_t crc = 0;
        if (w.gzip && md.n) {
            const uint32_t Lb = md.n - ((md.nseg - 1) << SEG_LOG2);  // 1 .. 4096
            const uint32_t XL = Lb < SEG ? c_xpowb[Lb] : c_segpow[1];  // x^(8 Lb)
            for (uint32_t s = lane; s < md.nseg; s += 64) {
                const uint32_t raw = w.seg_crc[md.seg_base + s];
                crc ^= s + 1 == md.nseg ? raw : gf2_mul(raw, gf2_mul(c_segpow[md.nseg - 2 - s], XL));
            }
            for (int d = 32; d >= 1; d >>= 1) crc ^= __shfl_xor(crc, d, 64);
            crc ^= gf2_mul(0xFFFFFFFFu, gf2_mul(c_segpow[md.nseg - 1], XL)) ^ 0xFFFFFFFFu;
        } else if (w.gzip) {
            crc = 0;  // empty member: CRC-32 of nothing
        }
        if (lane == 0) {
            w.prefix_bits[m] = pbits;
            const uint64_t bits = pbits + body;
            w.mem_bytes[m] = (bits + 7) / 8 + (w.gzip ? 8 : 0);
            w.mem_crc[m] = crc;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_scan: member and chunk output offsets (single workgroup)

__global__ __launch_bounds__(1024) void k_scan(Work w) {
    __shared__ uint64_t tmp[16];
    const uint32_t M = uni(w.counts->nmem), C = uni(w.counts->nchunk);
    uint64_t carry = 0;
    for (uint32_t b = 0; b < M; b += 1024) {
        uint32_t i = b + threadIdx.x;
        uint64_t v = i < M ? w.mem_bytes[i] : 0, tot;
        uint64_t ex = block_scan64<1024>(v, tmp, &tot);
        if (i < M) w.mem_off[i] = carry + ex;
        carry += tot;
    }
    if (threadIdx.x == 0) *w.total = carry;
    __syncthreads();
    for (uint32_t c = threadIdx.x; c < C; c += 1024) {
        uint32_t m0 = w.chunk_mem0[c], m1 = w.chunk_mem0[c + 1];
        uint64_t o0 = w.mem_off[m0];
        uint64_t o1 = m1 < M ? w.mem_off[m1] : carry;
        w.chunk_off[c] = o0;
        w.chunk_len[c] = o1 - o0;
    }
}

// ---------------------------------------------------------------------------
// k_emit: one workgroup per member

// Block size BT: 1024 threads (least latency per member) or 256 (eight
// members per CU, for batches).  A tile is 4 tokens per thread.
template <int BT>
struct EmitShape {
    static constexpr int BUF_WORDS = (4 * BT * 48) / 32 + 16;  // tile x 48 bits + carry
};

template <int BT>
struct Emitter {
    uint32_t *buf;
    uint8_t *out;
    uint64_t A, B;      // member byte range
    uint64_t cursor;    // physical bit position
    uint32_t carry;     // bits below cursor%32 of word cursor/32

    __device__ __forceinline__ void or_bits(uint32_t rel, uint32_t v, uint32_t nb) const {
        if (!nb) return;
        uint32_t pos = (uint32_t)(cursor & 31) + rel;
        uint32_t wi = pos >> 5, sh = pos & 31;
        atomicOr(&buf[wi], v << sh);
        if (sh + nb > 32) atomicOr(&buf[wi + 1], v >> (32 - sh));
    }
    __device__ __forceinline__ void store_word(uint64_t wi, uint32_t v) const {
        uint64_t b0 = wi * 4;
        if (b0 >= A && b0 + 4 <= B) {
            reinterpret_cast<uint32_t *>(out)[wi] = v;
        } else {
            for (int k = 0; k < 4; k++) {
                uint64_t a = b0 + k;
                if (a >= A && a < B) out[a] = (uint8_t)(v >> (8 * k));
            }
        }
    }
    // before contributions: clear the words the tile touches, seed the carry
    __device__ __forceinline__ void begin(uint32_t tile_bits) {
        uint32_t nw = ((uint32_t)(cursor & 31) + tile_bits + 31) >> 5;
        for (uint32_t i = threadIdx.x; i < nw + 1; i += BT) buf[i] = i == 0 ? carry : 0;
        __syncthreads();
    }
    // after contributions: store complete words, keep the partial one
    __device__ __forceinline__ void end(uint32_t tile_bits) {
        __syncthreads();
        uint32_t off0 = (uint32_t)(cursor & 31);
        uint32_t full = (off0 + tile_bits) >> 5;
        uint64_t w0 = cursor >> 5;
        for (uint32_t i = threadIdx.x; i < full; i += BT) store_word(w0 + i, buf[i]);
        carry = ((off0 + tile_bits) & 31) ? buf[full] : 0;
        cursor += tile_bits;
        __syncthreads();
    }
};

template <int BT>
__global__ __launch_bounds__(BT) void k_emit(Work w) {
    __shared__ uint32_t ctab[NSYM];
    __shared__ uint32_t buf[EmitShape<BT>::BUF_WORDS];
    __shared__ uint32_t tmp[32];
    const uint32_t tid = threadIdx.x;
    const uint32_t M = uni(w.counts->nmem);
    for (uint32_t m = blockIdx.x; m < M; m += gridDim.x) {
        const MemberDesc md = load_member(w.mem, m);
        for (uint32_t i = tid; i < NSYM; i += BT) ctab[i] = w.codes[(uint64_t)m * NSYM + i];
        Emitter<BT> em;
        em.buf = buf;
        em.out = w.out;
        em.A = w.mem_off[m];
        em.B = em.A + w.mem_bytes[m];
        em.cursor = em.A * 8;
        em.carry = 0;
        __syncthreads();

        // prefix: gzip header + dynamic block header
        {
            const uint32_t pbits = w.prefix_bits[m];
            em.begin(pbits);
            const uint32_t *pw = w.prefix + (uint64_t)m * PREFIX_WORDS;
            for (uint32_t i = tid; i * 32 < pbits; i += BT) em.or_bits(i * 32, pw[i], min(32u, pbits - i * 32));
            em.end(pbits);
        }
        // tokens, 4096 per tile
        const uint32_t nt = w.ntok[m];
        const uint32_t *tk = w.walked + md.pos_base;
        // this thread's 4 tokens of a tile (member regions are 64-position
        // aligned, so a full group is one 16-byte load), loaded one tile ahead
        auto load4 = [&](uint32_t t0) {
            const uint32_t ti = t0 + tid * 4;
            uint4 q = make_uint4(0, 0, 0, 0);
            if (ti + 4 <= nt) {
                q = *reinterpret_cast<const uint4 *>(tk + ti);
            } else {
                if (ti < nt) q.x = tk[ti];
                if (ti + 1 < nt) q.y = tk[ti + 1];
                if (ti + 2 < nt) q.z = tk[ti + 2];
            }
            return q;
        };
        uint4 nxt = load4(0);
        for (uint32_t t0 = 0; t0 < nt; t0 += 4 * BT) {
            const uint4 cur = nxt;
            nxt = load4(t0 + 4 * BT);
            const uint32_t tq[4] = {cur.x, cur.y, cur.z, cur.w};
            uint32_t v[4][4], nb[4][4];
            uint32_t mybits = 0;
#pragma unroll
            for (int k = 0; k < 4; k++) {
                uint32_t ti = t0 + tid * 4 + k;
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    v[k][f] = 0;
                    nb[k][f] = 0;
                }
                if (ti >= nt) continue;
                uint32_t t = tq[k], len = t >> 16, x = t & 0xFFFFu;
                if (len == 0) {
                    uint32_t c = ctab[x & 0xFF];
                    v[k][0] = c & 0xFFFF;
                    nb[k][0] = c >> 16;
                } else {
                    uint32_t li = len_idx(len);
                    uint32_t c = ctab[257 + li];
                    v[k][0] = c & 0xFFFF;
                    nb[k][0] = c >> 16;
                    v[k][1] = len - c_len_base[li];
                    nb[k][1] = len_extra(li);
                    uint32_t ds = dist_sym(x);
                    uint32_t dc = ctab[DSYM0 + ds];
                    v[k][2] = dc & 0xFFFF;
                    nb[k][2] = dc >> 16;
                    v[k][3] = x - c_dist_base[ds];
                    nb[k][3] = dist_extra(ds);
                }
                mybits += nb[k][0] + nb[k][1] + nb[k][2] + nb[k][3];
            }
            uint32_t tile_bits;
            uint32_t rel = block_scan<BT>(mybits, tmp, &tile_bits);
            em.begin(tile_bits);
#pragma unroll
            for (int k = 0; k < 4; k++)
#pragma unroll
                for (int f = 0; f < 4; f++) {
                    em.or_bits(rel, v[k][f], nb[k][f]);
                    rel += nb[k][f];
                }
            em.end(tile_bits);
        }
        // end of block, zero pad to a byte, gzip trailer
        {
            uint32_t eob = ctab[256];
            uint32_t eb = eob >> 16;
            uint32_t padded = (uint32_t)(((em.cursor + eb + 7) & ~7ull) - em.cursor);
            uint32_t tile_bits = padded + (w.gzip ? 64 : 0);
            em.begin(tile_bits);
            if (tid == 0) {
                em.or_bits(0, eob & 0xFFFF, eb);
                if (w.gzip) {
                    em.or_bits(padded, w.mem_crc[m], 32);
                    em.or_bits(padded + 32, md.n, 32);
                }
            }
            em.end(tile_bits);
            if (tid == 0 && (em.cursor & 31)) em.store_word(em.cursor >> 5, em.carry);
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// k_emit_tiled: bit emission with one workgroup per slot of a member, every
// slot at once.  Member m has a header slot (the gzip header and dynamic
// block header bits that k_huffman assembled) and one token slot per segment.
// Token slot t = 1 .. nseg holds tokens [ET (t-1), ET t) while that range is
// full (t <= nfull = min(ntok / ET, nseg - 1)); the last slot (t = nseg) holds
// the tokens from ET nfull to ntok, the end-of-block code, the pad to a byte
// and the gzip trailer; the slots between are empty.  A token slot's bit
// offset is the header's bits plus those of the member's full slots before
// it, each of which publishes its count as soon as it has it.  Every nonempty
// slot spans at least 32 bits (the header at least 80, a full slot at least
// ET, the last at least 64), so a 32-bit output word is shared by at most two
// slots, neighbours among the nonempty ones: both add their parts into one
// atomic word, and the second to arrive writes the output word.  Words at the
// member's two ends are written byte by byte within the member.  A member's
// token-slot workgroups take their slots from a per-member ticket in the
// order they start, so a slot only ever waits on slots whose workgroups run.
// k_huffman clears the tickets and the slot state.
constexpr uint32_t ET = 4096;                   // tokens per token slot
constexpr uint32_t ET_BUF = ET * 48 / 32 + 16;  // slot bits (<= 48 per token) + offset + trailer
constexpr uint64_t ES_AGG = 1ull << 62, ES_VAL = (1ull << 62) - 1;

// Word wi of the output, only the bytes inside [A, B).
__device__ __forceinline__ void store_word_in(uint8_t *out, uint64_t wi, uint32_t v, uint64_t A, uint64_t B) {
    const uint64_t b0 = wi * 4;
    if (b0 >= A && b0 + 4 <= B) {
        reinterpret_cast<uint32_t *>(out)[wi] = v;
    } else {
        for (int k = 0; k < 4; k++) {
            const uint64_t a = b0 + k;
            if (a >= A && a < B) out[a] = (uint8_t)(v >> (8 * k));
        }
    }
}

// The word shared with a neighbouring slot.  The two slots' parts occupy
// disjoint bits, so adding a part is or-ing it: each slot adds 1 << 32 plus
// its part to join[key] in one atomic, and the one that finds the other's
// part there (count 1) writes the word.  The data travels in the atomic, so
// no fence is needed.
__device__ __forceinline__ void combine_word(const Work &w, uint32_t key, uint32_t part, uint64_t wi, uint64_t A,
                                             uint64_t B) {
    const uint64_t old = atomicAdd(reinterpret_cast<unsigned long long *>(&w.em_join[key]),
                                   (unsigned long long)((1ull << 32) | part));
    if ((old >> 32) == 1) store_word_in(w.out, wi, (uint32_t)old | part, A, B);
}

// Code bits of one token: up to four fields (code, extra, distance code, extra).
__device__ __forceinline__ uint32_t token_fields(const uint32_t *ctab, uint32_t t, uint32_t v[4], uint32_t nb[4]) {
    const uint32_t len = t >> 16, x = t & 0xFFFFu;
    v[1] = v[2] = v[3] = 0;
    nb[1] = nb[2] = nb[3] = 0;
    if (len == 0) {
        const uint32_t c = ctab[x & 0xFF];
        v[0] = c & 0xFFFF;
        nb[0] = c >> 16;
    } else {
        const uint32_t li = len_idx(len);
        const uint32_t c = ctab[257 + li];
        v[0] = c & 0xFFFF;
        nb[0] = c >> 16;
        v[1] = len - c_len_base[li];
        nb[1] = len_extra(li);
        const uint32_t ds = dist_sym(x);
        const uint32_t dc = ctab[DSYM0 + ds];
        v[2] = dc & 0xFFFF;
        nb[2] = dc >> 16;
        v[3] = x - c_dist_base[ds];
        nb[3] = dist_extra(ds);
    }
    return nb[0] + nb[1] + nb[2] + nb[3];
}

__device__ __forceinline__ void or_bits_at(uint32_t *buf, uint32_t pos, uint32_t v, uint32_t nb) {
    if (!nb) return;
    const uint32_t wi = pos >> 5, sh = pos & 31;
    atomicOr(&buf[wi], v << sh);
    if (sh + nb > 32) atomicOr(&buf[wi + 1], v >> (32 - sh));
}

template <int BT>
__global__ __launch_bounds__(BT) void k_emit_tiled(Work w) {
    constexpr uint32_t TPT = ET / BT;  // tokens per thread
    __shared__ uint32_t ctab[NSYM];
    __shared__ uint32_t buf[ET_BUF];
    __shared__ uint32_t tmp[32];
    __shared__ uint32_t s_slot;
    __shared__ uint64_t s_pre;
    const uint32_t tid = threadIdx.x;
    const uint32_t G = uni(w.counts->nseg), M = uni(w.counts->nmem);
    const uint32_t b = blockIdx.x;
    if (b >= G + M) return;
    // workgroups [0, G) take the token slots, one per segment, in the order
    // they start (a ticket per member); workgroups [G, G + M) the headers
    const bool hdr = b >= G;
    const uint32_t m = hdr ? b - G : uni(w.seg_mem[b]);
    const MemberDesc md = load_member(w.mem, m);
    const uint32_t nseg = md.nseg, nt = uni(w.ntok[m]);
    if (!hdr) {
        if (tid == 0) s_slot = atomicAdd(&w.em_tick[m], 1u);
        __syncthreads();
    }
    const uint32_t t = hdr ? 0u : uni(s_slot) + 1;
    const uint32_t u = md.seg_base + t - 1;  // a token slot's segment (its state index)
    const uint32_t nfull = nseg ? min(nt / ET, nseg - 1) : 0u;
    const bool last = t == nseg;  // the member's final slot (the header of an empty member)
    if (!hdr && !last && t > nfull) return;  // an empty slot
    const uint64_t A = w.mem_off[m], B = A + w.mem_bytes[m];
    for (uint32_t i = tid; i < NSYM; i += BT) ctab[i] = w.codes[(uint64_t)m * NSYM + i];
    const uint32_t pbits = uni(w.prefix_bits[m]);

    // this thread's tokens and their bit count
    const uint32_t t0 = hdr ? 0u : (last ? nfull * ET : (t - 1) * ET);
    const uint32_t t1 = hdr ? 0u : (last ? nt : t * ET);
    const uint32_t *tk = w.walked + md.pos_base;
    uint32_t tq[TPT];
#pragma unroll
    for (uint32_t k = 0; k < TPT; k += 4) {
        const uint32_t i = t0 + tid * TPT + k;
        uint4 q = make_uint4(0, 0, 0, 0);
        if (i + 4 <= t1) {
            q = *reinterpret_cast<const uint4 *>(tk + i);
        } else {
            if (i < t1) q.x = tk[i];
            if (i + 1 < t1) q.y = tk[i + 1];
            if (i + 2 < t1) q.z = tk[i + 2];
        }
        tq[k] = q.x;
        tq[k + 1] = q.y;
        tq[k + 2] = q.z;
        tq[k + 3] = q.w;
    }
    __syncthreads();  // ctab
    uint32_t mybits = 0;
    {
        uint32_t v[4], nb[4];
#pragma unroll
        for (uint32_t k = 0; k < TPT; k++)
            if (t0 + tid * TPT + k < t1) mybits += token_fields(ctab, tq[k], v, nb);
    }
    uint32_t tile_bits;
    const uint32_t rel = block_scan<BT>(mybits, tmp, &tile_bits);
    if (hdr) tile_bits = pbits;

    // bits of the member before this slot: the header's and those of the full
    // slots before it, which each publish their count as soon as they have
    // it; the first wave reads up to 64 of them at once
    uint64_t pre = 0;
    if (!hdr) {
        if (tid < 64) {
            uint64_t *st = w.em_state + md.seg_base;  // slot t at st[t - 1]
            if (!last && tid == 0)
                __hip_atomic_store(&st[t - 1], ES_AGG | tile_bits, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            uint64_t acc = 0;
            for (uint32_t j0 = 1, top = last ? nfull : t - 1; j0 <= top; j0 += 64) {
                const uint32_t j = j0 + tid;
                uint64_t s = 0;
                if (j <= top) {
                    // slot j's workgroup has started (tickets), so it
                    // publishes; the bound only keeps a broken state from
                    // hanging the queue
                    for (uint32_t spin = 0; spin < (1u << 22); spin++) {
                        s = __hip_atomic_load(&st[j - 1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        if (s >> 62) break;
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                uint64_t v = s & ES_VAL;
                for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
                acc += v;
            }
            if (tid == 0) s_pre = acc + pbits;
        }
        __syncthreads();
        pre = s_pre;
    }

    const uint64_t S = A * 8 + pre;  // absolute bit position of the slot's first bit
    const uint32_t off0 = (uint32_t)(S & 31);
    uint32_t all_bits = tile_bits, padded = 0;
    const uint32_t eob = ctab[256], eb = eob >> 16;
    if (last) {
        const uint64_t e = S + tile_bits;
        padded = (uint32_t)(((e + eb + 7) & ~7ull) - e);
        all_bits += padded + (w.gzip ? 64 : 0);
    }
    const uint32_t nw = (off0 + all_bits + 31) >> 5;
    for (uint32_t i = tid; i <= nw; i += BT) buf[i] = 0;
    __syncthreads();
    if (hdr) {
        const uint32_t *pw = w.prefix + (uint64_t)m * PREFIX_WORDS;
        for (uint32_t i = tid; i * 32 < pbits; i += BT) or_bits_at(buf, off0 + i * 32, pw[i], min(32u, pbits - i * 32));
    } else {
        uint32_t r = off0 + rel;
#pragma unroll
        for (uint32_t k = 0; k < TPT; k++) {
            if (t0 + tid * TPT + k >= t1) break;
            uint32_t v[4], nb[4];
            token_fields(ctab, tq[k], v, nb);
#pragma unroll
            for (int f = 0; f < 4; f++) {
                or_bits_at(buf, r, v[f], nb[f]);
                r += nb[f];
            }
        }
    }
    if (last && tid == 0) {
        const uint32_t e = off0 + tile_bits;
        or_bits_at(buf, e, eob & 0xFFFF, eb);
        if (w.gzip) {
            or_bits_at(buf, e + padded, w.mem_crc[m], 32);
            or_bits_at(buf, e + padded + 32, md.n, 32);
        }
    }
    __syncthreads();
    // words: the member's first and last words byte by byte, a word shared
    // with a neighbouring slot through combine_word, the rest whole
    const uint64_t w0 = S >> 5;
    const bool head_shared = !hdr && off0 != 0;
    const bool tail_shared = !last && ((off0 + all_bits) & 31) != 0;
    for (uint32_t i = tid; i < nw; i += BT) {
        if ((i == 0 && head_shared) || (i + 1 == nw && tail_shared)) continue;
        store_word_in(w.out, w0 + i, buf[i], A, B);
    }
    if (tid == 0) {
        if (head_shared) combine_word(w, u, buf[0], w0, A, B);  // (never the header)
        if (tail_shared) {
            // the next nonempty slot of the member
            const uint32_t tn = (t < nfull) ? t + 1 : nseg;
            combine_word(w, md.seg_base + tn - 1, buf[nw - 1], w0 + nw - 1, A, B);
        }
    }
}

// ---------------------------------------------------------------------------
// k_plan_one: the member and segment tables of one chunk of len bytes cut into
// members of `sub` bytes, written on the device (the host's build_plan, in
// closed form: every member but the last is full).  For callers whose
// descriptor tables live in memory the host does not stage.

__global__ __launch_bounds__(256) void k_plan_one(Work w, uint32_t len, uint32_t sub) {
    MemberDesc *mem = const_cast<MemberDesc *>(w.mem);
    uint32_t *seg_mem = const_cast<uint32_t *>(w.seg_mem);
    uint32_t *chunk_mem0 = const_cast<uint32_t *>(w.chunk_mem0);
    Counts *counts = const_cast<Counts *>(w.counts);
    const uint32_t M = len <= sub ? 1u : (len + sub - 1) / sub;
    const uint32_t spm = (sub + SEG - 1) >> SEG_LOG2;  // segments of a full member
    const uint32_t pstep = (sub + POS_ALIGN - 1) / POS_ALIGN * POS_ALIGN;
    const uint32_t nlast = len <= sub ? len : len - (M - 1) * sub;
    const uint32_t G = (M - 1) * spm + ((nlast + SEG - 1) >> SEG_LOG2);
    for (uint32_t i = threadIdx.x; i < M; i += blockDim.x) {
        MemberDesc d;
        d.n = i + 1 == M ? nlast : sub;
        d.in_off = (uint64_t)i * sub;
        d.pos_base = (uint64_t)i * pstep;
        d.seg_base = i * spm;
        d.nseg = (d.n + SEG - 1) >> SEG_LOG2;
        d.chunk = 0;
        mem[i] = d;
    }
    for (uint32_t g = threadIdx.x; g < G; g += blockDim.x) seg_mem[g] = g / spm;
    if (threadIdx.x == 0) {
        chunk_mem0[0] = 0;
        chunk_mem0[1] = M;
        *counts = Counts{G, M, 1, 0};
    }
}

// k_plan_batch: the member and segment tables of a batch, built on the device
// from its chunk descriptors (w.cdesc) -- the host's build_plan in closed form
// per chunk: ceil(len / sub) members (one when len <= sub), every member but
// the last full.  The descriptors are the only per-batch upload, so a
// captured graph that starts with this kernel serves any batch within the
// workspace's capacity.  One workgroup: chunks in tiles of 1024, a scan of
// their member, segment and position counts per tile, then one thread per
// member of the tile.
__global__ __launch_bounds__(1024) void k_plan_batch(Work w) {
    __shared__ uint32_t tmp[16];
    __shared__ uint32_t t_m0[1024], t_g0[1024], t_p0[1024];  // tile-relative bases per chunk
    MemberDesc *mem = const_cast<MemberDesc *>(w.mem);
    uint32_t *seg_mem = const_cast<uint32_t *>(w.seg_mem);
    uint32_t *chunk_mem0 = const_cast<uint32_t *>(w.chunk_mem0);
    Counts *counts = const_cast<Counts *>(w.counts);
    const uint64_t *cd = w.cdesc;
    const uint32_t C = (uint32_t)cd[0];
    const uint32_t sub = w.sub, tid = threadIdx.x;
    const uint32_t spm = (sub + SEG - 1) >> SEG_LOG2;        // segments of a full member
    const uint32_t pstep = (sub + POS_ALIGN - 1) / POS_ALIGN;  // positions of a full member, / 64
    uint32_t cm = 0, cg = 0;
    uint64_t cp = 0;  // members, segments and positions / 64 before the tile
    for (uint32_t c0 = 0; c0 < C; c0 += 1024) {
        const uint32_t nc = min(1024u, C - c0), c = c0 + tid;
        uint32_t nm = 0, ng = 0, np = 0;
        if (tid < nc) {
            const uint64_t len = cd[3 + 2 * (uint64_t)c];
            const uint64_t m = len <= sub ? 1 : (len + sub - 1) / sub;
            const uint32_t nlast = (uint32_t)(len - (m - 1) * sub);
            nm = (uint32_t)m;
            ng = (nm - 1) * spm + ((nlast + SEG - 1) >> SEG_LOG2);
            np = (nm - 1) * pstep + (nlast + POS_ALIGN - 1) / POS_ALIGN;
        }
        uint32_t tm, tg, tp;
        t_m0[tid] = block_scan_dpp<1024>(nm, tmp, &tm);
        t_g0[tid] = block_scan_dpp<1024>(ng, tmp, &tg);
        t_p0[tid] = block_scan_dpp<1024>(np, tmp, &tp);
        if (tid < nc) chunk_mem0[c] = cm + t_m0[tid];
        __syncthreads();
        for (uint32_t j = tid; j < tm; j += 1024) {
            // the tile's chunk of member j: the last t with t_m0[t] <= j (every
            // chunk has a member, so the bases rise strictly)
            uint32_t lo = 0, hi = nc;
            while (hi - lo > 1) {
                const uint32_t mid = (lo + hi) >> 1;
                if (t_m0[mid] <= j) lo = mid;
                else hi = mid;
            }
            const uint32_t cc = c0 + lo, i = j - t_m0[lo];
            const uint64_t off = cd[2 + 2 * (uint64_t)cc], len = cd[3 + 2 * (uint64_t)cc];
            const uint64_t m = len <= sub ? 1 : (len + sub - 1) / sub;
            MemberDesc d;
            d.n = i + 1 == m ? (uint32_t)(len - (m - 1) * sub) : sub;
            d.in_off = off + (uint64_t)i * sub;
            d.pos_base = (cp + t_p0[lo] + (uint64_t)i * pstep) * POS_ALIGN;
            d.seg_base = cg + t_g0[lo] + i * spm;
            d.nseg = (d.n + SEG - 1) >> SEG_LOG2;
            d.chunk = cc;
            const uint32_t mi = cm + j;
            mem[mi] = d;
            for (uint32_t q = 0; q < d.nseg; q++) seg_mem[d.seg_base + q] = mi;
        }
        cm += tm;
        cg += tg;
        cp += tp;
        __syncthreads();
    }
    if (tid == 0) {
        chunk_mem0[C] = cm;
        *counts = Counts{cg, cm, C, 0};
    }
}

}  // namespace gpz

// ---------------------------------------------------------------------------
// launch helpers used by the host runtime

namespace gpz {

hipError_t upload_constants() {
    uint32_t tab[256];
    for (uint32_t i = 0; i < 256; i++) {
        uint32_t c = i;
        for (int k = 0; k < 8; k++) c = (c & 1u) ? (c >> 1) ^ 0xEDB88320u : c >> 1;
        tab[i] = c;
    }
    auto mul = [](uint32_t a, uint32_t b) {
        uint32_t r = 0;
        for (int i = 0; i < 32; i++) {
            if (a & (0x80000000u >> i)) r ^= b;
            b = (b & 1u) ? (b >> 1) ^ 0xEDB88320u : b >> 1;
        }
        return r;
    };
    // x^(8b) for b < 4096 by repeated multiplication with x^8, then whole segments
    static uint32_t xpowb[4096], segpow[4097];
    const uint32_t x8 = 0x00800000u;  // x^8, reflected
    xpowb[0] = 0x80000000u;           // 1
    for (uint32_t b = 1; b < 4096; b++) xpowb[b] = mul(xpowb[b - 1], x8);
    const uint32_t seg1 = mul(xpowb[4095], x8);  // x^(8 * 4096)
    segpow[0] = 0x80000000u;
    for (uint32_t j = 1; j <= 4096; j++) segpow[j] = mul(segpow[j - 1], seg1);
    // slicing-by-4: T_k[i] is the CRC of byte i followed by k zero bytes
    uint32_t tab4[1024];
    for (uint32_t i = 0; i < 256; i++) tab4[i] = tab[i];
    for (uint32_t k = 1; k < 4; k++)
        for (uint32_t i = 0; i < 256; i++) {
            const uint32_t p = tab4[(k - 1) * 256 + i];
            tab4[k * 256 + i] = (p >> 8) ^ tab[p & 0xFF];
        }
    hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_crc_tab), tab4, sizeof(tab4));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_segpow), segpow, sizeof(segpow));
    if (e == hipSuccess) e = hipMemcpyToSymbol(HIP_SYMBOL(c_xpowb), xpowb, sizeof(xpowb));
    return e;
}

void launch_seg_table(const Work &w, uint32_t grid, int variant, hipStream_t st) {
    if (variant == 1)
        hipLaunchKernelGGL(k_seg_table<256>, dim3(grid), dim3(256), 0, st, w);
    else if (variant == 2)
        hipLaunchKernelGGL(k_seg_table<1024>, dim3(grid), dim3(1024), 0, st, w);
    else
        hipLaunchKernelGGL(k_seg_table<512>, dim3(grid), dim3(512), 0, st, w);
}
void launch_match_find(const Work &w, uint32_t grid, int summary, int variant, hipStream_t st) {
    if (variant == 1)
        hipLaunchKernelGGL(k_match_find<4>, dim3(grid), dim3(K1_T), 0, st, w, summary);
    else
        hipLaunchKernelGGL(k_match_find<8>, dim3(grid), dim3(K1_T), 0, st, w, summary);
}
void launch_walk_summary(const Work &w, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(k_walk_summary, dim3(grid), dim3(KW_T), 0, st, w);
}
void launch_chain(const Work &w, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(k_chain, dim3(grid), dim3(64), 0, st, w);
}
void launch_walk_emit(const Work &w, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(k_walk_emit, dim3(grid), dim3(K3_T), 0, st, w);
}
void launch_walk_member(const Work &w, uint32_t grid, int parts, hipStream_t st) {
    if (parts == 4)
        hipLaunchKernelGGL(k_walk_member<4>, dim3(grid), dim3(KM_T), 0, st, w);
    else if (parts == 2)
        hipLaunchKernelGGL(k_walk_member<2>, dim3(grid), dim3(KM_T), 0, st, w);
    else
        hipLaunchKernelGGL(k_walk_member<1>, dim3(grid), dim3(KM_T), 0, st, w);
}
void launch_hist_tokens(const Work &w, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(k_hist_tokens, dim3(grid), dim3(1024), 0, st, w);
}
void launch_huffman(const Work &w, uint32_t grid, hipStream_t st) {
    hipLaunchKernelGGL(k_huffman, dim3(grid), dim3(64), 0, st, w);
}
void launch_scan(const Work &w, hipStream_t st) { hipLaunchKernelGGL(k_scan, dim3(1), dim3(1024), 0, st, w); }
void launch_plan_one(const Work &w, uint32_t len, uint32_t sub, hipStream_t st) {
    hipLaunchKernelGGL(k_plan_one, dim3(1), dim3(256), 0, st, w, len, sub);
}
void launch_plan_batch(const Work &w, hipStream_t st) {
    hipLaunchKernelGGL(k_plan_batch, dim3(1), dim3(1024), 0, st, w);
}
void launch_emit(const Work &w, uint32_t grid, uint32_t slots, int variant, hipStream_t st) {
    if (variant == 3)
        hipLaunchKernelGGL(k_emit_tiled<256>, dim3(slots), dim3(256), 0, st, w);
    else if (variant == 2)
        hipLaunchKernelGGL(k_emit_tiled<1024>, dim3(slots), dim3(1024), 0, st, w);
    else if (variant == 1)
        hipLaunchKernelGGL(k_emit<256>, dim3(grid), dim3(256), 0, st, w);
    else
        hipLaunchKernelGGL(k_emit<1024>, dim3(grid), dim3(1024), 0, st, w);
}

}  // namespace gpz
