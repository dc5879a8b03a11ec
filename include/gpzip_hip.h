/*
 * gpzip_hip.h -- C ABI of libgpzip_hip.so, the MI355X (gfx950) gzip compress
 * path that stands in for gpzip's GPU backend.
 *
 * Every entry point is extern "C", takes plain pointers and sizes, never
 * unwinds, and reports an int status (GPZ_OK or a GPZ_E_* code).  A
 * human-readable message for the most recent failure on the calling thread
 * is available from gpz_last_error().
 *
 * Reference interfaces replaced (paths relative to the reference repo,
 * crates/gpzip- prefix dropped):
 *   gpz_init / gpz_destroy    GpuBackend::try_init + GpuContext::try_init
 *                             (codec-gpu/src/gpu.rs:125-186,
 *                              codec-gpu/src/gpu/context.rs:65-106)
 *   gpz_compress_chunk        GpuBackend::gzip_chunk_fn closure
 *                             (codec-gpu/src/gpu.rs:193-208)
 *   gpz_gzip_member           encode_one (codec-gpu/src/gpu.rs:271-287)
 *   gpz_match_find            BatchedLz77::submit / Lz77HashPipeline::match_find
 *                             (codec-gpu/src/gpu/batch.rs:74-85,
 *                              codec-gpu/src/gpu/lz77_hash.rs:306-453)
 *   gpz_walk                  greedy_walk (codec-gpu/src/gpu/lz77.rs:260-284);
 *                             strict: BlockWalkPipeline::walk /
 *                             host_strict_greedy (codec-gpu/src/gpu/walk.rs:291-462, 512-525)
 *   gpz_encode_block          encode_block_fast (codec-gpu/src/gpu/deflate.rs:752-959)
 *   gpz_batch_*               BatchedLz77 submitter/completer pair
 *                             (codec-gpu/src/gpu/batch.rs:48-141) -- here a whole
 *                             batch of uneven chunks runs as one graph launch
 *   gpz_plan_*                device-resident form of gpz_batch_* (inputs and
 *                             outputs already in HBM); no reference counterpart
 *                             beyond Lz77HashPipeline::submit_batch_async/collect_async
 *                             (codec-gpu/src/gpu/lz77_hash.rs:473-664)
 *   gpz_server_*              no reference counterpart (serving form); its
 *                             in-order member stream follows ParallelChunkedWriter's
 *                             ordering contract (codec-cpu/src/parallel.rs:144-150)
 *
 * Token format (reference lz77_hash.rs:50-62): one uint32 per token,
 * length << 16 | (distance or literal byte); length == 0 => literal.
 */
#ifndef GPZIP_HIP_H
#define GPZIP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GPZ_OK = 0,
    GPZ_E_NODEV = 1,   /* no usable GPU / bad device index */
    GPZ_E_OOM = 2,     /* device or pinned-host allocation failed */
    GPZ_E_INVAL = 3,   /* bad argument or token outside the DEFLATE alphabet */
    GPZ_E_DEVICE = 4,  /* a HIP call failed */
    GPZ_E_NOSPACE = 5, /* caller's output buffer too small; *out_len holds the need */
    GPZ_E_BUSY = 6,    /* serving form: admission refused (budget) */
    GPZ_E_AGAIN = 7    /* serving form: nothing ready yet */
};

typedef struct gpz_ctx gpz_ctx;

typedef struct gpz_config {
    int device;            /* HIP device index; -1 = take GPZIP_GPU env or 0 */
    uint32_t sub_chunk;    /* member size; reference fixes 128 KiB (gpu.rs:160) */
    uint32_t window;       /* DEFLATE window; reference fixes 32 KiB (lz77_hash.rs:48) */
    uint32_t lanes;        /* concurrent per-chunk pipelines (streams) */
    uint32_t max_in_flight;/* advertised to ParallelChunkedWriter (gpu.rs:165) */
    uint32_t flags;        /* GPZ_CFG_* bits, 0 by default */
    uint64_t lane_bytes;   /* per-lane input capacity reserved at init (0 = 2 MiB) */
} gpz_config;

/* GPZ_CFG_RATIO: the opt-in ratio mode, outside reference parity (no reference
 * counterpart).  The walk demotes a match to a literal when the next position
 * holds a longer match (the reference's lazy rule) or the position after it a
 * match longer by two or more; 0.4-1.7 % smaller output on the synthetic
 * corpora (binmix 0.4 %, log 1.4 %, text 1.7 %), same match finder.  Applies to every compress path of the context
 * (plain, no-alloc, batched, plans, serving); gpz_walk takes the rule as
 * strict = 2 in any context. */
enum { GPZ_CFG_RATIO = 1u };

/* Fill *cfg with the reference's defaults: 128 KiB members, 32 KiB window,
 * 16 in flight. */
void gpz_config_default(gpz_config *cfg);

/* Select the device, check it is gfx950, load the kernels and build the
 * lane pool.  GPZ_E_NODEV when no GPU is present (the caller then falls back
 * to its CPU path, as LazyGpuBackend::try_get does at gpu.rs:53-57). */
int gpz_init(const gpz_config *cfg, gpz_ctx **out);
void gpz_destroy(gpz_ctx *ctx);

/* Message of the most recent failure on this thread ("" if none). */
const char *gpz_last_error(void);
const char *gpz_status_string(int status);

/* Number of HIP devices visible (0 on a host without a GPU). */
int gpz_device_count(int *count);

/* Device label, e.g. "AMD Instinct MI355X (gfx950)". */
int gpz_device_label(gpz_ctx *ctx, char *buf, size_t cap);

/* Upper bound of compressed bytes for a chunk of len bytes. */
size_t gpz_compress_bound(const gpz_ctx *ctx, size_t len);

/* ---- plain per-chunk form --------------------------------------------- */

/* Compress one chunk into ceil(len / sub_chunk) concatenated gzip members
 * (one member when len <= sub_chunk, including len == 0).  Thread-safe and
 * blocking; concurrent callers run on separate lanes (streams).  On
 * GPZ_E_NOSPACE *out_len is set to the required size. */
int gpz_compress_chunk(gpz_ctx *ctx, const uint8_t *in, size_t len, uint8_t *out, size_t out_cap,
                       size_t *out_len);

/* The same compression with no allocation inside the call: no hipMalloc,
 * hipHostMalloc or free, for callers that cannot allocate (tight serving
 * loops, harnesses that forbid it).  No counterpart in the reference.  The
 * workspace, input staging and output staging come from a fixed 40 MiB
 * device array that each GPU holds until the process ends, which limits a
 * call to len <= 2 MiB (GPZ_E_INVAL above that, or when the context's member
 * size needs more scratch).  in and out may be host or device memory.  Calls
 * through one context are serialized; calls on one GPU from different
 * contexts share the array and race.  Output bytes equal
 * gpz_compress_chunk's. */
int gpz_compress_chunk_noalloc(gpz_ctx *ctx, const uint8_t *in, size_t len, uint8_t *out, size_t out_cap,
                               size_t *out_len);

/* ---- stage-level entry points (parity / debugging) --------------------- */

/* Per-position packed tokens for one member (len <= 16 MiB). */
int gpz_match_find(gpz_ctx *ctx, const uint8_t *in, size_t len, uint32_t *tokens);

/* Walk per-position tokens into the emitted token stream.  strict = 0: lazy
 * rule of greedy_walk (needs in for demoted literals); strict = 1: the
 * strict rule of host_strict_greedy (in may be NULL); strict = 2: the ratio
 * mode's lookahead rule (GPZ_CFG_RATIO; needs in).  Token lengths must be
 * 0..258.  walked must hold len entries; *n_walked receives the count. */
int gpz_walk(gpz_ctx *ctx, const uint8_t *in, const uint32_t *per_pos, size_t len, int strict,
             uint32_t *walked, size_t *n_walked);

/* Raw DEFLATE (one BFINAL dynamic block) for a walked token stream. */
int gpz_encode_block(gpz_ctx *ctx, const uint32_t *walked, size_t n, uint8_t *out, size_t out_cap,
                     size_t *out_len);

/* One gzip member for len bytes (any len <= 16 MiB, no sub-chunk split). */
int gpz_gzip_member(gpz_ctx *ctx, const uint8_t *in, size_t len, uint8_t *out, size_t out_cap,
                    size_t *out_len);

/* ---- decode side (host cores) --------------------------------------------
 * The reference decodes on the CPU: gpzip x scans a gzip stream for the fixed
 * 10-byte member header and inflates the members in parallel, falling back
 * to a serial multi-member decode (codec-cpu/src/parallel_decompress.rs:35-113,
 * codec-cpu/src/decompress.rs:28-44).  Host code on system zlib; no GPU. */

/* Offsets of the fixed gzip member headers in [in, in + len).  *count gets
 * the number found; GPZ_E_NOSPACE when more than cap. */
int gpz_gzip_member_starts(const uint8_t *in, size_t len, uint64_t *starts, size_t cap, size_t *count);
/* Decode a gzip stream (any number of members) into out.  threads = 0: one
 * per hardware thread.  GPZ_E_NOSPACE with *out_len = the decoded size when
 * out_cap is too small; GPZ_E_INVAL for a corrupt stream. */
int gpz_gunzip(const uint8_t *in, size_t len, uint8_t *out, size_t out_cap, size_t *out_len, unsigned threads);

/* ---- batched form (host buffers) ---------------------------------------- */

typedef struct gpz_desc {
    uint64_t offset; /* byte offset of the chunk in base */
    uint64_t len;    /* chunk length */
} gpz_desc;

typedef struct gpz_batch gpz_batch;

/* Launch all chunks as one batch (one graph replay on one lane).  Returns
 * as soon as the work is enqueued; inputs are copied before return. */
int gpz_batch_submit(gpz_ctx *ctx, const uint8_t *base, const gpz_desc *descs, size_t n,
                     gpz_batch **out);
int gpz_batch_wait(gpz_batch *b);
/* Result of chunk i: pointer valid until gpz_batch_free. */
int gpz_batch_result(gpz_batch *b, size_t i, const uint8_t **data, size_t *len);
void gpz_batch_free(gpz_batch *b);

/* ---- device-resident form (inputs and outputs in HBM) ------------------- */

typedef struct gpz_plan gpz_plan;

/* GPZ_PLAN_TIMED records HIP events around the kernels of every run, for
 * gpz_plan_timing.  The events are not free: on 2 MiB chunks they cost about
 * 15 % of the pipeline, so plans that are not being profiled leave it off. */
enum { GPZ_PLAN_GRAPH = 1u, GPZ_PLAN_PERSISTENT = 2u, GPZ_PLAN_TIMED = 4u };

/* Launch shape of the batched form.  Zero fields keep the built-in choice.
 * Grids are absolute; segments-per-workgroup fields give the grid relative
 * to each batch (ceil(segments / n)) and take precedence over the grid. */
typedef struct gpz_tune {
    uint32_t k0_grid;    /* persistent grid of the segment-table kernel */
    uint32_t k1_grid;    /* persistent grid of the match-find kernel */
    uint32_t k3_grid;    /* persistent grid of the walk/emit-tokens kernel */
    uint32_t emit_grid;  /* persistent grid of the bit-emission kernel */
    uint32_t k1_variant; /* match-find occupancy: 1 = two workgroups per CU (64 VGPRs),
                            2 = one workgroup per CU (128 VGPRs); 0 = built-in */
    uint32_t emit_variant; /* bit emission: 1 = one 1024-thread workgroup per member,
                              2 = one 256-thread workgroup per member (eight members per CU),
                              3 / 4 = one 1024 / 256-thread workgroup per 4,096-token slot of
                              a member, every slot at once (emit_grid unused); 0 = built-in (3
                              below 512 members, else 2; the autotuner records 0 when the
                              built-in choice is the faster one) */
    uint32_t k0_variant;   /* segment-table workgroup: 1 = 256 threads (16 positions each),
                              2 = 512 (8), 3 = 1024 (4); 0 = built-in (2) */
    uint32_t k1_seg_per_wg; /* match-find segments per workgroup; 0 = use k1_grid */
    uint32_t k3_seg_per_wg; /* walk/emit-tokens segments per workgroup; 0 = use k3_grid */
    uint32_t walk_variant;  /* walk: 1 = per segment (walk summaries in the match-find kernel,
                               a chaining pass, one workgroup per segment), 2 = per member
                               (each member walked in order by one to four waves); 0 = built-in (2 for
                               batches of 512 members or more, else 1; the autotuner records 0
                               when the built-in choice is the faster one) */
    uint32_t walk_parts;    /* per-member walk: waves per member, 1, 2 or 4 (a member's blocks
                               cut into that many parts walked at once); 0 = built-in (4; the
                               autotuner records 0 when 4 is the fastest) */
    uint32_t reserved[1];
} gpz_tune;

/* Tuning record used by every batched launch of this context (gpz_batch_*,
 * the serving form, and gpz_plan_create with GPZ_PLAN_PERSISTENT and a NULL
 * tune); its walk_variant also applies to the single-chunk and stage-level
 * calls.  NULL restores the built-in grids.  No reference counterpart: the
 * reference's batching constants are fixed (codec-gpu/src/gpu/batch.rs:31-35). */
int gpz_set_tune(gpz_ctx *ctx, const gpz_tune *tune);
int gpz_get_tune(gpz_ctx *ctx, gpz_tune *tune);

/* Autotuner: times candidate launch shapes of every kernel group on the given
 * device-resident batch (same arguments as gpz_plan_create + gpz_plan_run)
 * and returns the fastest in *out.  ms[0] receives the pipeline time with the
 * chosen shape, ms[1] with the built-in one (device time, HIP events).  Does
 * not change the context's tuning record. */
int gpz_autotune(gpz_ctx *ctx, const gpz_desc *descs, size_t n, const void *d_in, void *d_out, gpz_tune *out,
                 float *ms);

/* Plan for a fixed list of chunks laid out in a device buffer: workspace is
 * allocated once; gpz_plan_run may then be called repeatedly.  Chunk offsets
 * must be multiples of 16.  d_in must have 256 readable bytes past the last
 * chunk.  With GPZ_PLAN_GRAPH the run is one graph replay whose first kernel
 * builds the member and segment tables on the device from the chunk
 * descriptors (as gpz_batch_* and the serving form do); without it the
 * tables are built here once and the kernels are launched directly. */
int gpz_plan_create(gpz_ctx *ctx, const gpz_desc *descs, size_t n, uint32_t flags,
                    const gpz_tune *tune, gpz_plan **out);
/* Output capacity the plan needs in d_out. */
size_t gpz_plan_out_capacity(const gpz_plan *p);
/* Enqueue one run on the plan's stream (graph replay with GPZ_PLAN_GRAPH). */
int gpz_plan_run(gpz_plan *p, const void *d_in, void *d_out);
int gpz_plan_sync(gpz_plan *p);
/* After sync: per-chunk output offset/length in d_out and the total. */
int gpz_plan_result(gpz_plan *p, uint64_t *chunk_off, uint64_t *chunk_len, uint64_t *total);
/* Device time (ms) of the last run, from HIP events on the plan's stream:
 * [0] whole run, [1] match-find kernel, [2] emit kernel, [3] segment-table
 * kernel ([1]-[3] only for runs without GPZ_PLAN_GRAPH).  GPZ_E_INVAL unless
 * the plan was created with GPZ_PLAN_TIMED. */
int gpz_plan_timing(gpz_plan *p, float *ms, size_t n);
void gpz_plan_destroy(gpz_plan *p);

/* ---- serving form -------------------------------------------------------- */

typedef struct gpz_server gpz_server;

/* Device memory of a server is one arena of budget_bytes, allocated by
 * gpz_server_start: `streams` lane workspaces of fixed capacity (hash tables,
 * tokens, Huffman tables, step output) and an input pool.  Admitted requests
 * wait in the pool, in HBM, until compressed. */
typedef struct gpz_server_config {
    uint64_t budget_bytes;     /* device memory of the arena (workspaces + input pool) */
    uint32_t max_batch_chunks; /* chunks gathered into one device step */
    uint32_t max_batch_bytes;  /* input bytes gathered into one device step (8 MiB) */
    uint32_t chunk_bytes;      /* request split size (2 MiB; a multiple of 256) */
    uint32_t streams;          /* device steps in flight (worker lanes, at most 16) */
} gpz_server_config;

void gpz_server_config_default(gpz_server_config *cfg);
/* GPZ_E_INVAL if budget_bytes cannot hold the lane workspaces plus one chunk
 * (gpz_server_stats [6] of a started server gives the workspace bytes). */
int gpz_server_start(gpz_ctx *ctx, const gpz_server_config *cfg, gpz_server **out);
/* Admit a request: its bytes are copied into the arena's input pool before
 * return.  GPZ_E_BUSY if it can never fit the pool; otherwise blocks until it
 * is admitted (first come, first served).  A tag may not be in use (submitted
 * and not yet released). */
int gpz_request_submit(gpz_server *s, const uint8_t *in, size_t len, uint64_t tag);
/* Next finished request, in completion order; each request's members are
 * concatenated in input order.  timeout_ms < 0 waits forever.  The buffer is
 * owned by the server until gpz_request_release(tag).  *latency_s runs from
 * the gpz_request_submit call (any wait for pool space included) to the
 * request's completion. */
int gpz_request_next(gpz_server *s, int timeout_ms, uint64_t *tag, const uint8_t **data,
                     size_t *len, double *latency_s);
/* Free a request that gpz_request_next returned; GPZ_E_INVAL for a tag it has
 * not returned yet (the server may still be writing that request's output). */
int gpz_request_release(gpz_server *s, uint64_t tag);
/* Counters: [0] requests done, [1] input bytes done, [2] device steps,
 * [3] peak arena bytes in use (workspaces + admitted inputs), [4] arena bytes,
 * [5] admissions that had to wait for pool space, [6] lane workspace bytes. */
int gpz_server_stats(gpz_server *s, uint64_t *out, size_t n);
/* Stops after every admitted request is compressed.  No other call on s may
 * overlap it. */
void gpz_server_stop(gpz_server *s);

#ifdef __cplusplus
}
#endif
#endif
