// gpz_runtime.h -- host-side structures shared by the runtime translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <memory>
#include <vector>

#include "gpz_internal.h"
#include "gpzip_hip.h"

struct gpz_ctx;

namespace gpz {

enum { MODE_FULL = 0, MODE_MATCH = 1, MODE_WALK = 2, MODE_ENCODE = 3 };

int fail(int code, const char *fmt, ...);
void clear_error();
const char *last_error();

struct DevBuf;
struct HostBuf;

// Chunk list expanded to members and segments.  With device_tables the
// member and segment tables are not built here: only the totals (which size
// the workspace) and the chunk descriptors, from which k_plan_batch builds
// the tables on the device.
struct Plan {
    std::vector<MemberDesc> mem;
    std::vector<uint32_t> seg_mem;
    std::vector<uint32_t> chunk_mem0;     // C + 1 entries
    std::vector<uint64_t> chunk_dev_off;  // chunk offsets in the device input buffer
    std::vector<uint64_t> chunk_len;
    uint64_t P = 0, in_end = 0, in_bytes = 0, out_cap = 0;
    uint32_t G = 0, M = 0, C = 0, sub = 0;
    int gzip = 1;
    bool device_tables = false;
};

int build_plan(Plan &pl, const uint64_t *lens, const uint64_t *dev_off, size_t C, uint32_t sub_chunk,
               int gzip, bool device_tables = false);

struct Caps {
    uint64_t P = 0, G = 0, M = 0, C = 0, out = 0;
};

struct Layout {
    uint64_t mem, seg_mem, chunk_mem0, counts, cdesc, desc_end;
    uint64_t tok, segtab, seg_ntab, wsum, seg_crc, seg_entry, seg_tokbase, walked, hist, ntok, codes, prefix, prefix_bits;
    uint64_t mem_bytes, mem_crc, mem_off, chunk_off, chunk_len, total, em_state, em_join, em_tick, out, end;
};

struct Grids {
    uint32_t k0 = 1, k1 = 1, k3 = 1, chain = 1, huff = 1, emit = 1;
    uint32_t k1v = 0;    // match-find variant: 0 two workgroups per CU, 1 one
    uint32_t emitv = 0;  // bit-emission variant: 0 1024 threads per member, 1 256; 2 / 3 one
                         // workgroup of 1024 / 256 threads per slot (k_emit_tiled)
    uint32_t eslots = 1; // k_emit_tiled grid: segments + members
    uint32_t k0v = 0;    // segment-table workgroup: 0 512 threads, 1 256, 2 1024
    uint32_t km = 1;     // grid of the per-member walk (four waves per workgroup)
    uint32_t walkv = 0;  // walk: 0 per segment (summaries, chain, k_walk_emit), 1 per member
    uint32_t kmp = 1;    // per-member walk: waves (parts) per member, 1, 2 or 4
    bool operator==(const Grids &o) const {
        return k0 == o.k0 && k1 == o.k1 && k3 == o.k3 && chain == o.chain && huff == o.huff && emit == o.emit &&
               k1v == o.k1v && emitv == o.emitv && eslots == o.eslots && k0v == o.k0v && km == o.km && walkv == o.walkv && kmp == o.kmp;
    }
};

Grids grids_for(const Caps &c, int n_cu);
// Override the persistent grids with the non-zero fields of a tuning record,
// then cap the segment grids at G and the member grid at M.
void apply_tune(Grids &g, const gpz_tune &t, uint64_t G, uint64_t M);
void enqueue(const Work &w, const Grids &g, int mode, hipStream_t st, hipEvent_t *ev);

}  // namespace gpz

// DevBuf / HostBuf are defined in gpz_runtime.hip; the structs below hold them
// by value, so the definitions are repeated here for the other units.
namespace gpz {

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    ~DevBuf();
};

struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes);
    ~HostBuf();
};

Layout layout_for(const Caps &c);
Work work_at(const Layout &L, uint8_t *b);

struct Workspace {
    DevBuf ws;
    HostBuf hdesc;
    Caps caps;
    Layout lay{};
    Work work{};
    uint8_t *base = nullptr;  // device memory the layout starts at
    bool external = false;    // laid out in memory it does not own (bind): fixed capacity
    uint64_t generation = 0;
    // Grow (own memory) to hold `want`.
    int reserve(const Caps &want);
    // Lay out capacity c in caller-owned device memory of `bytes` bytes.
    int bind(const Caps &c, void *mem, uint64_t bytes);
    int upload_plan(const Plan &pl, hipStream_t st);
    void point(uint8_t *b);
};

struct Lane {
    hipStream_t st = nullptr;
    hipEvent_t ev[6] = {};
    Workspace wsp;
    DevBuf din;
    HostBuf hin, hout;
    hipGraphExec_t graph = nullptr;
    uint64_t graph_gen = 0;       // workspace generation the graph was captured for
    const void *graph_in = nullptr;
    Grids graph_grids;            // launch shapes the graph was captured with
    uint32_t strict = 0;
    int init(int dev);
    ~Lane();
};

int ctx_device(gpz_ctx *c);
int ctx_cus(gpz_ctx *c);
uint32_t ctx_sub_chunk(gpz_ctx *c);
// walk rule of the context's compress paths: 0 the reference's lazy rule,
// 2 the opt-in ratio mode (GPZ_CFG_RATIO)
uint32_t ctx_walk_rule(gpz_ctx *c);
const gpz_tune &ctx_tune(gpz_ctx *c);
Lane *ctx_acquire(gpz_ctx *c);
void ctx_release(gpz_ctx *c, Lane *l);
int lane_run(gpz_ctx *ctx, Lane *ln, const Plan &pl, int mode, const uint8_t *host_in, const uint64_t *host_lens);
int lane_fetch_chunk(Lane *ln, uint32_t c, uint8_t *out, size_t out_cap, size_t *out_len);

// Batched step on a lane: pack the chunks into pinned staging, upload, replay
// the lane's graph.  Returns once enqueued.
int lane_batch_launch(gpz_ctx *ctx, Lane *ln, const Plan &pl, const uint8_t *const *ptrs, const uint64_t *lens,
                      unsigned copy_threads);
// Batched step over chunks already in device memory (plan offsets relative to
// d_in): upload the plan, replay the lane's graph.  Returns once enqueued.
int lane_step_resident(gpz_ctx *ctx, Lane *ln, const Plan &pl, const void *d_in);
// Wait for the step and copy every chunk's members to host.
int lane_batch_collect(Lane *ln, const Plan &pl, std::unique_ptr<uint8_t[]> &out, std::vector<uint64_t> &off,
                       std::vector<uint64_t> &len);
// Same, leaving the members in the lane's pinned staging (valid until the lane's next step).
int lane_batch_collect_pinned(Lane *ln, const Plan &pl, const uint8_t **data, std::vector<uint64_t> &off,
                              std::vector<uint64_t> &len);

}  // namespace gpz
