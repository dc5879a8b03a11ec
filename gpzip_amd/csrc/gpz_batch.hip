// gpz_batch.hip -- batched form (host buffers) and device-resident plans.
//
// A batch of uneven chunks runs as one replay of a captured HIP graph on one
// lane: the kernels are gridded persistently over the segment/member tables,
// whose sizes they read from device memory, so one instantiated graph serves
// every batch that fits the lane's workspace.  This replaces the reference's
// submitter/completer thread pair that packs up to 16 chunks per wgpu
// submission (codec-gpu/src/gpu/batch.rs:101-141).
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <memory>
#include <thread>
#include <vector>

#include "gpz_internal.h"
#include "gpz_runtime.h"
#include "gpzip_hip.h"

using namespace gpz;

#define HIPCHK(x)                                                                              \
    do {                                                                                       \
        hipError_t e_ = (x);                                                                   \
        if (e_ != hipSuccess) return fail(GPZ_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

namespace gpz {

// Host copies of a batch: one memcpy per chunk, split over up to max_threads
// threads (contiguous runs of chunks of about equal bytes) once the batch is
// large enough for a single core's copy rate to matter.  Callers that already
// copy from several threads at once (the server's workers) pass 1.
static void copy_pieces(uint8_t *dst, const uint64_t *dst_off, const uint8_t *const *src, const uint64_t *len,
                        size_t n, unsigned max_threads) {
    uint64_t total = 0;
    for (size_t i = 0; i < n; i++) total += len[i];
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const size_t T = total < (16ull << 20) ? 1 : std::min<size_t>({(size_t)max_threads, (size_t)hw, n});
    auto run = [&](size_t a, size_t b) {
        for (size_t i = a; i < b; i++)
            if (len[i]) memcpy(dst + dst_off[i], src[i], len[i]);
    };
    if (T <= 1) {
        run(0, n);
        return;
    }
    std::vector<size_t> cut{0};
    uint64_t acc = 0;
    for (size_t i = 0; i < n && cut.size() < T; i++) {
        acc += len[i];
        if (acc * T >= total * cut.size()) cut.push_back(i + 1);
    }
    if (cut.back() != n) cut.push_back(n);
    std::vector<std::thread> th;
    for (size_t k = 1; k + 1 < cut.size(); k++) th.emplace_back(run, cut[k], cut[k + 1]);
    run(cut[0], cut[1]);
    for (auto &t : th) t.join();
}

// One large buffer copied by up to 8 threads.
// Host copies and transfers of a batch go in pieces of this size, each
// overlapping the next piece's.
constexpr uint64_t PIPE_BYTES = 32ull << 20;

static void copy_bytes(uint8_t *dst, const uint8_t *src, uint64_t n) {
    constexpr uint64_t PIECE = 4ull << 20;
    const uint64_t np = (n + PIECE - 1) / PIECE;
    std::vector<uint64_t> off(np), len(np);
    std::vector<const uint8_t *> ptr(np);
    for (uint64_t i = 0; i < np; i++) {
        off[i] = i * PIECE;
        len[i] = std::min(PIECE, n - off[i]);
        ptr[i] = src + off[i];
    }
    copy_pieces(dst, off.data(), ptr.data(), len.data(), np, 8);
}

// Capture the FULL-mode kernel sequence for w on st into *exec.
static int capture_graph(const Work &w, const Grids &g, hipStream_t st, hipGraphExec_t *exec) {
    if (*exec) {
        (void)hipGraphExecDestroy(*exec);
        *exec = nullptr;
    }
    hipGraph_t graph = nullptr;
    HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
    enqueue(w, g, MODE_FULL, st, nullptr);
    hipError_t e = hipStreamEndCapture(st, &graph);
    if (e != hipSuccess) return fail(GPZ_E_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) {
        *exec = nullptr;
        return fail(GPZ_E_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    return GPZ_OK;
}

int lane_step_resident(gpz_ctx *ctx, Lane *ln, const Plan &pl, const void *d_in) {
    int rc;
    if ((rc = ln->wsp.reserve(Caps{pl.P, pl.G, pl.M, pl.C, pl.out_cap}))) return rc;
    HIPCHK(hipSetDevice(ctx_device(ctx)));
    if ((rc = ln->wsp.upload_plan(pl, ln->st))) return rc;
    Work w = ln->wsp.work;
    if (pl.device_tables) {
        w.cdesc = (const uint64_t *)(ln->wsp.base + ln->wsp.lay.cdesc);
        w.sub = pl.sub;
    }
    w.in = (const uint8_t *)d_in;
    w.gzip = 1;
    w.strict = ctx_walk_rule(ctx);
    Grids g = grids_for(ln->wsp.caps, ctx_cus(ctx));
    apply_tune(g, ctx_tune(ctx), ln->wsp.caps.G, ln->wsp.caps.M);  // capacity: stable, so the graph is kept
    if (!ln->graph || ln->graph_gen != ln->wsp.generation || ln->graph_in != d_in || !(ln->graph_grids == g)) {
        if ((rc = capture_graph(w, g, ln->st, &ln->graph))) return rc;
        ln->graph_gen = ln->wsp.generation;
        ln->graph_in = d_in;
        ln->graph_grids = g;
    }
    HIPCHK(hipGraphLaunch(ln->graph, ln->st));
    return GPZ_OK;
}

int lane_batch_launch(gpz_ctx *ctx, Lane *ln, const Plan &pl, const uint8_t *const *ptrs, const uint64_t *lens,
                      unsigned copy_threads) {
    int rc;
    if ((rc = ln->din.ensure(pl.in_bytes))) return rc;
    if ((rc = ln->hin.ensure(pl.in_bytes))) return rc;
    HIPCHK(hipSetDevice(ctx_device(ctx)));
    // the host copy in pieces of about PIPE_BYTES of chunks, each piece's
    // upload issued as soon as it is in the pinned buffer, so that the copy of
    // the next piece overlaps it
    uint8_t *hin = (uint8_t *)ln->hin.p, *din = (uint8_t *)ln->din.p;
    for (size_t c0 = 0; c0 < pl.C;) {
        size_t c1 = c0 + 1;
        while (c1 < pl.C && pl.chunk_dev_off[c1] + lens[c1] - pl.chunk_dev_off[c0] <= PIPE_BYTES) c1++;
        copy_pieces(hin, pl.chunk_dev_off.data() + c0, ptrs + c0, lens + c0, c1 - c0, copy_threads);
        const uint64_t a = c0 ? pl.chunk_dev_off[c0] : 0, b = c1 < pl.C ? pl.chunk_dev_off[c1] : pl.in_end;
        if (b > a) HIPCHK(hipMemcpyAsync(din + a, hin + a, b - a, hipMemcpyHostToDevice, ln->st));
        c0 = c1;
    }
    return lane_step_resident(ctx, ln, pl, ln->din.p);
}

int lane_batch_collect_pinned(Lane *ln, const Plan &pl, const uint8_t **data, std::vector<uint64_t> &off,
                              std::vector<uint64_t> &len) {
    off.resize(pl.C);
    len.resize(pl.C);
    uint64_t total = 0;
    const Work &w = ln->wsp.work;
    HIPCHK(hipMemcpyAsync(off.data(), w.chunk_off, pl.C * 8, hipMemcpyDeviceToHost, ln->st));
    HIPCHK(hipMemcpyAsync(len.data(), w.chunk_len, pl.C * 8, hipMemcpyDeviceToHost, ln->st));
    HIPCHK(hipMemcpyAsync(&total, w.total, 8, hipMemcpyDeviceToHost, ln->st));
    HIPCHK(hipStreamSynchronize(ln->st));
    int rc = ln->hout.ensure(total + 8);
    if (rc) return rc;
    if (total) {
        HIPCHK(hipMemcpyAsync(ln->hout.p, w.out, total, hipMemcpyDeviceToHost, ln->st));
        HIPCHK(hipStreamSynchronize(ln->st));
    }
    *data = (const uint8_t *)ln->hout.p;
    return GPZ_OK;
}

// Results into caller-owned host memory: the download in pieces of
// PIPE_BYTES through the pinned buffer, each piece copied out while the next
// ones are still in flight.
int lane_batch_collect(Lane *ln, const Plan &pl, std::unique_ptr<uint8_t[]> &out, std::vector<uint64_t> &off,
                       std::vector<uint64_t> &len) {
    off.resize(pl.C);
    len.resize(pl.C);
    uint64_t total = 0;
    const Work &w = ln->wsp.work;
    HIPCHK(hipMemcpyAsync(off.data(), w.chunk_off, pl.C * 8, hipMemcpyDeviceToHost, ln->st));
    HIPCHK(hipMemcpyAsync(len.data(), w.chunk_len, pl.C * 8, hipMemcpyDeviceToHost, ln->st));
    HIPCHK(hipMemcpyAsync(&total, w.total, 8, hipMemcpyDeviceToHost, ln->st));
    HIPCHK(hipStreamSynchronize(ln->st));
    int rc = ln->hout.ensure(total + 8);
    if (rc) return rc;
    out.reset(new uint8_t[std::max<uint64_t>(total, 1)]);  // no zero fill: every byte is copied
    uint8_t *h = (uint8_t *)ln->hout.p;
    const uint64_t np = (total + PIPE_BYTES - 1) / PIPE_BYTES;
    std::vector<hipEvent_t> ev(np, nullptr);
    for (uint64_t k = 0; k < np && !rc; k++) {
        const uint64_t a = k * PIPE_BYTES, n = std::min(PIPE_BYTES, total - a);
        if (hipEventCreateWithFlags(&ev[k], hipEventDisableTiming) != hipSuccess ||
            hipMemcpyAsync(h + a, w.out + a, n, hipMemcpyDeviceToHost, ln->st) != hipSuccess ||
            hipEventRecord(ev[k], ln->st) != hipSuccess)
            rc = fail(GPZ_E_DEVICE, "result download failed");
    }
    for (uint64_t k = 0; k < np && !rc; k++) {
        const uint64_t a = k * PIPE_BYTES, n = std::min(PIPE_BYTES, total - a);
        if (hipEventSynchronize(ev[k]) != hipSuccess) {
            rc = fail(GPZ_E_DEVICE, "result download failed");
            break;
        }
        copy_bytes(out.get() + a, h + a, n);
    }
    if (rc) (void)hipStreamSynchronize(ln->st);
    for (hipEvent_t e : ev)
        if (e) (void)hipEventDestroy(e);
    return rc;
}

}  // namespace gpz

struct gpz_batch {
    gpz_ctx *ctx = nullptr;
    Lane *lane = nullptr;
    Plan pl;
    std::unique_ptr<uint8_t[]> out;
    std::vector<uint64_t> off, len;
    int status = GPZ_OK;
    bool done = false;
};

extern "C" {

int gpz_batch_submit(gpz_ctx *ctx, const uint8_t *base, const gpz_desc *descs, size_t n, gpz_batch **out) {
    clear_error();
    if (!ctx || !out || (n && (!descs || !base))) return fail(GPZ_E_INVAL, "bad argument");
    *out = nullptr;
    std::vector<uint64_t> lens(n);
    for (size_t i = 0; i < n; i++) lens[i] = descs[i].len;
    auto b = new gpz_batch();
    b->ctx = ctx;
    int rc = build_plan(b->pl, lens.data(), nullptr, n, ctx_sub_chunk(ctx), 1, true);
    if (rc) {
        delete b;
        return rc;
    }
    if (n == 0) {
        b->done = true;
        *out = b;
        return GPZ_OK;
    }
    Lane *ln = ctx_acquire(ctx);
    b->lane = ln;
    std::vector<const uint8_t *> ptrs(n);
    for (size_t c = 0; c < n; c++) ptrs[c] = base + descs[c].offset;
    rc = lane_batch_launch(ctx, ln, b->pl, ptrs.data(), lens.data(), 8);
    if (rc) {
        (void)hipStreamSynchronize(ln->st);
        ctx_release(ctx, ln);
        b->lane = nullptr;
        delete b;
        return rc;
    }
    *out = b;
    return GPZ_OK;
}

int gpz_batch_wait(gpz_batch *b) {
    if (!b) return fail(GPZ_E_INVAL, "null batch");
    if (b->done) return b->status;
    Lane *ln = b->lane;
    const Plan &pl = b->pl;
    int rc = lane_batch_collect(ln, pl, b->out, b->off, b->len);
    ctx_release(b->ctx, ln);
    b->lane = nullptr;
    b->done = true;
    b->status = rc;
    return rc;
}

int gpz_batch_result(gpz_batch *b, size_t i, const uint8_t **data, size_t *len) {
    if (!b || !data || !len) return fail(GPZ_E_INVAL, "bad argument");
    if (!b->done) return fail(GPZ_E_INVAL, "batch not waited");
    if (b->status) return b->status;
    if (i >= b->pl.C) return fail(GPZ_E_INVAL, "chunk index %zu out of range", i);
    *data = b->out.get() + b->off[i];
    *len = b->len[i];
    return GPZ_OK;
}

void gpz_batch_free(gpz_batch *b) {
    if (!b) return;
    if (!b->done) gpz_batch_wait(b);
    delete b;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// device-resident plans

struct gpz_plan {
    gpz_ctx *ctx = nullptr;
    Plan pl;
    Lane lane;
    uint32_t flags = 0;
    Grids grids;
    const void *bound_in = nullptr;
    void *bound_out = nullptr;
    bool timed = false;
};

extern "C" {

int gpz_plan_create(gpz_ctx *ctx, const gpz_desc *descs, size_t n, uint32_t flags, const gpz_tune *tune,
                    gpz_plan **out) {
    clear_error();
    if (!ctx || !out || (n && !descs)) return fail(GPZ_E_INVAL, "bad argument");
    *out = nullptr;
    if (n == 0) return fail(GPZ_E_INVAL, "empty plan");
    if (flags & ~(uint32_t)(GPZ_PLAN_GRAPH | GPZ_PLAN_PERSISTENT | GPZ_PLAN_TIMED))
        return fail(GPZ_E_INVAL, "unknown plan flag");
    std::vector<uint64_t> lens(n), offs(n);
    for (size_t i = 0; i < n; i++) {
        lens[i] = descs[i].len;
        offs[i] = descs[i].offset;
    }
    auto p = new gpz_plan();
    p->ctx = ctx;
    p->flags = flags;
    // graph plans (the batched form) build their tables on the device from the
    // descriptors at each replay; direct-launch plans (the plain form's shape)
    // keep the tables uploaded once here, as gpz_compress_chunk uploads its own
    int rc = build_plan(p->pl, lens.data(), offs.data(), n, ctx_sub_chunk(ctx), 1, (flags & GPZ_PLAN_GRAPH) != 0);
    if (!rc) rc = p->lane.init(ctx_device(ctx));
    if (!rc) {
        const Plan &pl = p->pl;
        Caps c{pl.P, pl.G, pl.M, pl.C, 256};  // output goes to the caller's buffer
        rc = p->lane.wsp.reserve(c);
    }
    if (!rc) rc = p->lane.wsp.upload_plan(p->pl, p->lane.st);
    if (!rc && hipStreamSynchronize(p->lane.st) != hipSuccess) rc = fail(GPZ_E_DEVICE, "plan upload failed");
    if (rc) {
        delete p;
        return rc;
    }
    p->grids = grids_for(p->lane.wsp.caps, ctx_cus(ctx));
    if (!(flags & GPZ_PLAN_PERSISTENT)) {
        // one workgroup per segment / member
        p->grids.k0 = p->grids.k1 = p->grids.k3 = std::max<uint32_t>(1, p->pl.G);
        p->grids.emit = p->grids.huff = p->grids.chain = std::max<uint32_t>(1, p->pl.M);
    }
    if (tune) {
        if (tune->k1_variant > 2 || tune->emit_variant > 4 || tune->k0_variant > 3 || tune->walk_variant > 2 ||
            tune->walk_parts == 3 || tune->walk_parts > 4) {
            delete p;
            return fail(GPZ_E_INVAL,
                        "k1_variant and walk_variant must be 0-2, k0_variant 0-3, emit_variant 0-4, "
                    "walk_parts 0, 1, 2 or 4");
        }
        apply_tune(p->grids, *tune, p->pl.G, p->pl.M);
    } else if (flags & GPZ_PLAN_PERSISTENT) {
        apply_tune(p->grids, ctx_tune(ctx), p->pl.G, p->pl.M);
    }
    *out = p;
    return GPZ_OK;
}

// Autotuner.  Kernel groups are timed one at a time on the caller's batch
// (their times are independent: each runs alone on the stream), each over a
// short candidate list, keeping the others at the best found so far.
int gpz_autotune(gpz_ctx *ctx, const gpz_desc *descs, size_t n, const void *d_in, void *d_out, gpz_tune *out,
                 float *ms) {
    clear_error();
    if (!ctx || !out || !d_in || !d_out || (n && !descs)) return fail(GPZ_E_INVAL, "bad argument");
    gpz_plan *p = nullptr;
    int rc = gpz_plan_create(ctx, descs, n, GPZ_PLAN_PERSISTENT | GPZ_PLAN_TIMED, nullptr, &p);
    if (rc) return rc;
    const uint32_t cu = (uint32_t)ctx_cus(ctx);
    const uint32_t G = std::max<uint32_t>(1, p->pl.G), M = std::max<uint32_t>(1, p->pl.M);
    auto cap = [](uint64_t v, uint32_t hi) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(v, hi)); };
    const Grids base = p->grids;
    // time one configuration: median of three runs of the event interval [a, b)
    auto timed = [&](const Grids &g, int a, int b, float *med) -> int {
        p->grids = g;
        float v[3];
        for (int r = 0; r < 4; r++) {
            int e = gpz_plan_run(p, d_in, d_out);
            if (!e) e = gpz_plan_sync(p);
            if (e) return e;
            if (r == 0) continue;  // warm-up
            float t = 0;
            if (hipEventElapsedTime(&t, p->lane.ev[a], p->lane.ev[b]) != hipSuccess)
                return fail(GPZ_E_DEVICE, "event timing failed");
            v[r - 1] = t;
        }
        std::sort(v, v + 3);
        *med = v[1];
        return GPZ_OK;
    };
    // event order in enqueue: 0 start, 5 after seg_table, 1 after match-find,
    // 3 before emit, 4 after emit, 2 end
    Grids best = base;
    float t0 = 0;
    if ((rc = timed(base, 0, 2, &t0))) {
        gpz_plan_destroy(p);
        return rc;
    }
    struct Cand {
        uint32_t grid, variant;
    };
    // segments per workgroup -> grid on this batch
    auto spw = [&](uint32_t n) { return (G + n - 1) / n; };
    uint32_t best_k1_spw = 0, best_k3_spw = 0;
    {  // walk: per segment (summaries in match-find, chain, k_walk_emit) or per member
       // on 1, 2 or 4 waves; the interval holds match-find, the walk, Huffman and scan
        const uint32_t cs[][2] = {{0, 1}, {1, 1}, {1, 2}, {1, 4}};
        float bt = 1e30f;
        for (const auto &c : cs) {
            Grids g = best;
            g.walkv = c[0];
            g.kmp = c[1];
            g.km = (uint32_t)std::max<uint64_t>(1, ((uint64_t)M * c[1] + 3) / 4);
            float t;
            if ((rc = timed(g, 5, 3, &t))) break;
            if (t < bt) {
                bt = t;
                best.walkv = g.walkv;
                best.kmp = g.kmp;
                best.km = g.km;
            }
        }
    }
    if (!rc) {  // match-find: occupancy variant x segments per workgroup (1 = one workgroup per segment)
        const uint32_t spws[] = {1, 2, 4, 8, 16, 32, 64};
        float bt = 1e30f;
        for (uint32_t v = 0; v < 2 && !rc; v++) {
            for (uint32_t n : spws) {
                if (n > 1 && spw(n) < cu / 2) continue;  // too few workgroups to fill the chip
                Grids g = best;
                g.k1 = spw(n);
                g.k1v = v;
                float t;
                if ((rc = timed(g, 5, 1, &t))) break;
                if (t < bt) {
                    bt = t;
                    best.k1 = g.k1;
                    best.k1v = g.k1v;
                    best_k1_spw = n;
                }
            }
        }
    }
    if (!rc) {  // segment tables: workgroup size x persistent grid
        const Cand cs[] = {{cap((uint64_t)cu * 6, G), 0}, {cap((uint64_t)cu * 12, G), 0}, {G, 0},
                           {cap((uint64_t)cu * 12, G), 1}, {cap((uint64_t)cu * 24, G), 1}, {G, 1},
                           {cap((uint64_t)cu * 3, G), 2}, {cap((uint64_t)cu * 6, G), 2}, {G, 2}};
        float bt = 1e30f;
        for (const Cand &c : cs) {
            Grids g = best;
            g.k0 = c.grid;
            g.k0v = c.variant;
            float t;
            if ((rc = timed(g, 0, 5, &t))) break;
            if (t < bt) {
                bt = t;
                best.k0 = g.k0;
                best.k0v = g.k0v;
            }
        }
    }
    if (!rc && !best.walkv) {  // per-segment walk / token emission (with chain, Huffman and scan in the interval): segments per workgroup
        const uint32_t spws[] = {1, 2, 4, 8, 16, 32};
        float bt = 1e30f;
        for (uint32_t n : spws) {
            if (n > 1 && spw(n) < cu / 2) continue;
            Grids g = best;
            g.k3 = spw(n);
            float t;
            if ((rc = timed(g, 1, 3, &t))) break;
            if (t < bt) {
                bt = t;
                best.k3 = g.k3;
                best_k3_spw = n;
            }
        }
    }
    if (!rc) {  // bit emission: workgroup size x persistent grid
        const Cand cs[] = {{cap(cu, M), 0},
                           {cap((uint64_t)cu * 2, M), 0},
                           {cap((uint64_t)cu * 4, M), 0},
                           {M, 0},
                           {cap((uint64_t)cu * 8, M), 1},
                           {cap((uint64_t)cu * 16, M), 1},
                           {M, 1},
                           {M, 2},  // one workgroup per slot (grid: segments + members)
                           {M, 3}};
        float bt = 1e30f;
        for (const Cand &c : cs) {
            Grids g = best;
            g.emit = c.grid;
            g.emitv = c.variant;
            float t;
            if ((rc = timed(g, 3, 4, &t))) break;
            if (t < bt) {
                bt = t;
                best.emit = g.emit;
                best.emitv = g.emitv;
            }
        }
    }
    float t1 = 0;
    if (!rc) rc = timed(best, 0, 2, &t1);
    gpz_plan_destroy(p);
    if (rc) return rc;
    *out = gpz_tune{};
    out->k0_grid = best.k0;
    out->k1_grid = best.k1;
    out->k3_grid = best.k3;
    out->emit_grid = best.emit;
    out->k1_variant = best.k1v + 1;
    // recorded only where it differs from the built-in choice by member count,
    // which keeps a server's small steps on the per-slot emission
    out->emit_variant = best.emitv == base.emitv ? 0 : best.emitv + 1;
    out->k0_variant = best.k0v == 0 ? 2 : (best.k0v == 1 ? 1 : 3);
    out->k1_seg_per_wg = best_k1_spw;
    out->k3_seg_per_wg = best_k3_spw;
    // the walk is recorded only where the built-in rule (by member count)
    // would not pick it: the same record then keeps the per-segment walk on
    // the small steps of a server and the per-member walk on large batches
    out->walk_variant = best.walkv == base.walkv ? 0 : best.walkv + 1;
    out->walk_parts = best.walkv && best.kmp != base.kmp ? best.kmp : 0;
    if (ms) {
        ms[0] = t1;
        ms[1] = t0;
    }
    return GPZ_OK;
}

size_t gpz_plan_out_capacity(const gpz_plan *p) { return p ? p->pl.out_cap : 0; }

int gpz_plan_run(gpz_plan *p, const void *d_in, void *d_out) {
    clear_error();
    if (!p || !d_in || !d_out) return fail(GPZ_E_INVAL, "bad argument");
    if (((uintptr_t)d_in & 15) || ((uintptr_t)d_out & 3)) return fail(GPZ_E_INVAL, "misaligned device buffer");
    Lane &ln = p->lane;
    HIPCHK(hipSetDevice(ctx_device(p->ctx)));
    Work w = ln.wsp.work;
    if (p->pl.device_tables) {
        w.cdesc = (const uint64_t *)(ln.wsp.base + ln.wsp.lay.cdesc);  // tables built per run on the device
        w.sub = p->pl.sub;
    }
    w.in = (const uint8_t *)d_in;
    w.out = (uint8_t *)d_out;
    w.gzip = 1;
    w.strict = ctx_walk_rule(p->ctx);
    if (p->flags & GPZ_PLAN_GRAPH) {
        if (!ln.graph || p->bound_in != d_in || p->bound_out != d_out) {
            int rc = capture_graph(w, p->grids, ln.st, &ln.graph);
            if (rc) return rc;
            p->bound_in = d_in;
            p->bound_out = d_out;
        }
        const bool ev = p->flags & GPZ_PLAN_TIMED;
        if (ev) HIPCHK(hipEventRecord(ln.ev[0], ln.st));
        HIPCHK(hipGraphLaunch(ln.graph, ln.st));
        if (ev) HIPCHK(hipEventRecord(ln.ev[2], ln.st));
        p->timed = false;
    } else {
        // events only when asked for: each one is a marker on the queue
        enqueue(w, p->grids, MODE_FULL, ln.st, (p->flags & GPZ_PLAN_TIMED) ? ln.ev : nullptr);
        HIPCHK(hipGetLastError());
        p->timed = true;
    }
    return GPZ_OK;
}

int gpz_plan_sync(gpz_plan *p) {
    if (!p) return fail(GPZ_E_INVAL, "null plan");
    HIPCHK(hipStreamSynchronize(p->lane.st));
    return GPZ_OK;
}

int gpz_plan_result(gpz_plan *p, uint64_t *chunk_off, uint64_t *chunk_len, uint64_t *total) {
    if (!p) return fail(GPZ_E_INVAL, "null plan");
    const Work &w = p->lane.wsp.work;
    hipStream_t st = p->lane.st;
    if (chunk_off) HIPCHK(hipMemcpyAsync(chunk_off, w.chunk_off, p->pl.C * 8, hipMemcpyDeviceToHost, st));
    if (chunk_len) HIPCHK(hipMemcpyAsync(chunk_len, w.chunk_len, p->pl.C * 8, hipMemcpyDeviceToHost, st));
    if (total) HIPCHK(hipMemcpyAsync(total, w.total, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return GPZ_OK;
}

int gpz_plan_timing(gpz_plan *p, float *ms, size_t n) {
    if (!p || !ms) return fail(GPZ_E_INVAL, "bad argument");
    if (!(p->flags & GPZ_PLAN_TIMED)) return fail(GPZ_E_INVAL, "plan created without GPZ_PLAN_TIMED");
    Lane &ln = p->lane;
    float v[4] = {0, 0, 0, 0};
    HIPCHK(hipEventElapsedTime(&v[0], ln.ev[0], ln.ev[2]));
    if (p->timed) {
        HIPCHK(hipEventElapsedTime(&v[1], ln.ev[5], ln.ev[1]));
        HIPCHK(hipEventElapsedTime(&v[2], ln.ev[3], ln.ev[4]));
        HIPCHK(hipEventElapsedTime(&v[3], ln.ev[0], ln.ev[5]));
    }
    for (size_t i = 0; i < n && i < 4; i++) ms[i] = v[i];
    return GPZ_OK;
}

void gpz_plan_destroy(gpz_plan *p) {
    if (!p) return;
    (void)hipSetDevice(ctx_device(p->ctx));
    (void)hipStreamSynchronize(p->lane.st);
    delete p;
}

}  // extern "C"
