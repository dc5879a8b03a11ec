// gpz_server.hip -- serving form: continuous batching of compress requests.
//
// Device memory is one arena of budget_bytes, allocated at start:
//   * `streams` lane workspaces of fixed capacity (a step of up to
//     max(max_batch_bytes, chunk_bytes) input bytes in max_batch_chunks
//     chunks): hash tables, per-position and walked tokens, Huffman tables and
//     the step's output, reused for the server's lifetime;
//   * the rest is the input pool.  A request is admitted when its input
//     (rounded to 256 B, plus 256 B the kernels may read past it) fits a free
//     range of the pool; admission is first come first served, so a large
//     request is not starved by small ones behind it.  The admitted bytes are
//     copied straight from the caller's buffer into the pool and stay there,
//     resident in HBM, until the request's last chunk is compressed.  A
//     request larger than the whole pool is refused (GPZ_E_BUSY).
// Admitted requests are split into chunks of at most chunk_bytes and queued.
// `streams` worker threads each own a lane; whenever a worker is free it
// gathers queued chunks -- from any number of requests, oldest first -- into
// one device step (a graph replay over the pool, whose chunks the plan
// addresses in place), so requests that arrive while others are on the GPU
// join the next step and steps of different workers overlap on the device.
// Steps hold at most max_batch_bytes (8 MiB by default), so a large request
// is spread over several workers and finishes sooner.  A request completes
// when all its chunks have; its members are concatenated in input order,
// which is the ordering contract of the reference's ParallelChunkedWriter
// (codec-cpu/src/parallel.rs:144-150).  A chunk's member bytes depend only on
// the chunk, so a request's output is the same alone or in any mix.
#include <hip/hip_runtime.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "gpz_internal.h"
#include "gpz_runtime.h"
#include "gpzip_hip.h"

using namespace gpz;

namespace {

using clk = std::chrono::steady_clock;

constexpr uint64_t ARENA_ALIGN = 256;

uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// First-fit allocator over [0, cap) with coalescing free ranges.
struct RangeAlloc {
    std::map<uint64_t, uint64_t> free_;  // offset -> length
    uint64_t cap = 0, used = 0;

    void reset(uint64_t c) {
        free_.clear();
        cap = c;
        used = 0;
        if (c) free_[0] = c;
    }
    bool alloc(uint64_t len, uint64_t *off) {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < len) continue;
            *off = it->first;
            const uint64_t rest = it->second - len;
            const uint64_t at = it->first + len;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used += len;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t len) {
        used -= len;
        auto next = free_.lower_bound(off);
        if (next != free_.end() && off + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += len;
                return;
            }
        }
        free_[off] = len;
    }
};

struct Request {
    uint64_t tag = 0;
    uint64_t nbytes = 0;
    uint64_t dev_off = 0, dev_len = 0;  // input range in the pool
    bool admitted = false;
    uint32_t nchunks = 0, done_chunks = 0;
    std::vector<std::vector<uint8_t>> chunk_out;
    std::vector<uint8_t> out;
    clk::time_point t_submit, t_done;
    int status = GPZ_OK;
    bool handed_out = false;  // returned by gpz_request_next: only then may it be released
};

struct ChunkRef {
    Request *req;
    uint32_t idx;
    uint64_t off, len;  // off: byte offset in the arena's input pool
};

}  // namespace

struct gpz_server {
    gpz_ctx *ctx = nullptr;
    gpz_server_config cfg{};
    void *arena = nullptr;  // budget_bytes of device memory
    uint64_t ws_bytes = 0;  // lane workspaces at the front of the arena
    uint8_t *pool = nullptr;
    RangeAlloc inputs;
    std::mutex mu;
    std::condition_variable cv_work, cv_admit, cv_done;
    std::deque<ChunkRef> pending;
    std::map<uint64_t, std::unique_ptr<Request>> live;  // tag reserved at submit, until released
    std::deque<Request *> finished;
    uint64_t next_ticket = 0, admit_ticket = 0;  // admission in submission order
    uint64_t peak = 0, waited = 0;
    uint64_t n_done = 0, bytes_done = 0, steps = 0;
    bool stopping = false;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::vector<std::thread> workers;
    std::vector<hipStream_t> copy_streams;
    std::vector<hipStream_t> free_copy;

    ~gpz_server() {
        lanes.clear();
        for (hipStream_t s : copy_streams) (void)hipStreamDestroy(s);
        if (arena) (void)hipFree(arena);
    }

    hipStream_t take_copy_stream() {
        std::unique_lock<std::mutex> lk(mu);
        cv_admit.wait(lk, [&] { return !free_copy.empty(); });
        hipStream_t s = free_copy.back();
        free_copy.pop_back();
        return s;
    }
    void give_copy_stream(hipStream_t s) {
        {
            std::lock_guard<std::mutex> lk(mu);
            free_copy.push_back(s);
        }
        cv_admit.notify_all();
    }

    void worker(Lane *ln) {
        std::vector<ChunkRef> step;
        std::vector<uint64_t> lens, offs, off, len;
        for (;;) {
            step.clear();
            {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait(lk, [&] { return stopping || !pending.empty(); });
                if (stopping && pending.empty()) return;
                uint64_t bytes = 0;
                while (!pending.empty() && step.size() < cfg.max_batch_chunks) {
                    const ChunkRef &c = pending.front();
                    if (!step.empty() && bytes + c.len > cfg.max_batch_bytes) break;
                    bytes += c.len;
                    step.push_back(c);
                    pending.pop_front();
                }
                steps++;
            }
            lens.resize(step.size());
            offs.resize(step.size());
            for (size_t i = 0; i < step.size(); i++) {
                lens[i] = step[i].len;
                offs[i] = step[i].off;
            }
            Plan pl;
            int rc = build_plan(pl, lens.data(), offs.data(), step.size(), ctx_sub_chunk(ctx), 1, true);
            if (!rc) rc = lane_step_resident(ctx, ln, pl, pool);
            const uint8_t *res = nullptr;
            if (!rc) rc = lane_batch_collect_pinned(ln, pl, &res, off, len);
            if (rc) (void)hipStreamSynchronize(ln->st);
            // copy each chunk's members out without the lock (chunk_out slots
            // are distinct per chunk and never resized while in flight)
            if (!rc)
                for (size_t i = 0; i < step.size(); i++)
                    step[i].req->chunk_out[step[i].idx].assign(res + off[i], res + off[i] + len[i]);
            std::vector<Request *> completed;
            {
                std::lock_guard<std::mutex> lk(mu);
                for (size_t i = 0; i < step.size(); i++) {
                    Request *r = step[i].req;
                    if (rc) r->status = rc;
                    if (++r->done_chunks == r->nchunks) completed.push_back(r);
                }
            }
            // this worker finished these requests: concatenate outside the lock
            for (Request *r : completed) {
                if (r->chunk_out.size() == 1) {
                    r->out = std::move(r->chunk_out[0]);  // one chunk: no concatenation
                } else {
                    size_t total = 0;
                    for (auto &c : r->chunk_out) total += c.size();
                    r->out.reserve(total);
                    for (auto &c : r->chunk_out) r->out.insert(r->out.end(), c.begin(), c.end());
                }
                r->chunk_out.clear();
                r->chunk_out.shrink_to_fit();
            }
            if (!completed.empty()) {
                std::lock_guard<std::mutex> lk(mu);
                for (Request *r : completed) {
                    r->t_done = clk::now();
                    inputs.release(r->dev_off, r->dev_len);  // the pool range is free again
                    n_done++;
                    bytes_done += r->nbytes;
                    finished.push_back(r);
                }
            }
            if (!completed.empty()) {
                cv_done.notify_all();
                cv_admit.notify_all();
            }
        }
    }
};

extern "C" {

void gpz_server_config_default(gpz_server_config *cfg) {
    if (!cfg) return;
    cfg->budget_bytes = 64ull << 30;
    // steps of up to 8 MiB (~2k segments each, 12 in flight keep the GPU full)
    // spread a large request over several workers and bound the time a request
    // waits behind a step; 32 MiB steps gave twice the p99 at equal throughput
    cfg->max_batch_chunks = 64;
    cfg->max_batch_bytes = 8u << 20;
    cfg->chunk_bytes = 2u << 20;
    cfg->streams = 12;
}

// Workspace capacity of a step of at most B input bytes in at most K chunks
// (members of at most `sub` bytes).
static Caps step_caps(uint64_t B, uint64_t K, uint64_t sub) {
    Caps c;
    c.M = B / sub + K;                    // each chunk adds at most one partial member
    c.G = B / SEG + c.M;                  // each member adds at most one partial segment
    c.P = B + c.M * POS_ALIGN + POS_ALIGN;
    c.C = K;
    c.out = round_up(2 * B + c.M * (1024 + 18), 256) + 256;
    return c;
}

int gpz_server_start(gpz_ctx *ctx, const gpz_server_config *cfg_in, gpz_server **out) {
    clear_error();
    if (!ctx || !out) return fail(GPZ_E_INVAL, "bad argument");
    *out = nullptr;
    gpz_server_config cfg;
    if (cfg_in)
        cfg = *cfg_in;
    else
        gpz_server_config_default(&cfg);
    if (!cfg.max_batch_chunks || !cfg.max_batch_bytes || !cfg.chunk_bytes || !cfg.streams || !cfg.budget_bytes)
        return fail(GPZ_E_INVAL, "server config fields must be non-zero");
    if (cfg.chunk_bytes % ARENA_ALIGN) return fail(GPZ_E_INVAL, "chunk_bytes must be a multiple of 256");
    if (cfg.streams > 16) cfg.streams = 16;
    auto s = std::make_unique<gpz_server>();
    s->ctx = ctx;
    s->cfg = cfg;
    const uint64_t B = std::max<uint64_t>(cfg.max_batch_bytes, cfg.chunk_bytes);
    const Caps caps = step_caps(B, cfg.max_batch_chunks, ctx_sub_chunk(ctx));
    const uint64_t per_lane = round_up(layout_for(caps).end, ARENA_ALIGN);
    s->ws_bytes = per_lane * cfg.streams;
    // the pool must hold at least one full chunk (plus the kernels' read slack)
    if (cfg.budget_bytes < s->ws_bytes + cfg.chunk_bytes + 2 * ARENA_ALIGN)
        return fail(GPZ_E_INVAL, "budget of %llu bytes is below the %llu bytes of lane workspaces plus one chunk",
                    (unsigned long long)cfg.budget_bytes, (unsigned long long)(s->ws_bytes + cfg.chunk_bytes));
    if (hipSetDevice(ctx_device(ctx)) != hipSuccess) return fail(GPZ_E_DEVICE, "hipSetDevice failed");
    const uint64_t arena_bytes = round_up(cfg.budget_bytes, ARENA_ALIGN);
    if (hipMalloc(&s->arena, arena_bytes) != hipSuccess) {
        s->arena = nullptr;
        return fail(GPZ_E_OOM, "hipMalloc(%llu) for the serving arena failed", (unsigned long long)arena_bytes);
    }
    s->pool = (uint8_t *)s->arena + s->ws_bytes;
    s->inputs.reset(arena_bytes - s->ws_bytes);
    s->peak = s->ws_bytes;
    for (uint32_t i = 0; i < cfg.streams; i++) {
        auto ln = std::make_unique<Lane>();
        int rc = ln->init(ctx_device(ctx));
        if (!rc) rc = ln->wsp.bind(caps, (uint8_t *)s->arena + (uint64_t)i * per_lane, per_lane);
        if (rc) return rc;
        s->lanes.push_back(std::move(ln));
    }
    for (int i = 0; i < 4; i++) {
        hipStream_t cs;
        if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess)
            return fail(GPZ_E_DEVICE, "copy stream creation failed");
        s->copy_streams.push_back(cs);
        s->free_copy.push_back(cs);
    }
    gpz_server *raw = s.get();
    for (auto &ln : s->lanes) {
        Lane *l = ln.get();
        int dev = ctx_device(ctx);
        s->workers.emplace_back([raw, l, dev] {
            (void)hipSetDevice(dev);
            raw->worker(l);
        });
    }
    *out = s.release();
    return GPZ_OK;
}

int gpz_request_submit(gpz_server *s, const uint8_t *in, size_t len, uint64_t tag) {
    clear_error();
    if (!s || (len && !in)) return fail(GPZ_E_INVAL, "bad argument");
    const uint64_t need = round_up(std::max<uint64_t>(len, 1), ARENA_ALIGN) + ARENA_ALIGN;
    if (need > s->inputs.cap)
        return fail(GPZ_E_BUSY, "request of %zu bytes exceeds the %llu-byte input pool", len,
                    (unsigned long long)s->inputs.cap);
    auto owned = std::make_unique<Request>();
    Request *r = owned.get();
    r->tag = tag;
    r->nbytes = len;
    r->t_submit = clk::now();  // latency includes any wait for pool space
    const uint64_t cb = s->cfg.chunk_bytes;
    r->nchunks = len == 0 ? 1 : (uint32_t)((len + cb - 1) / cb);
    r->chunk_out.resize(r->nchunks);
    std::unique_lock<std::mutex> lk(s->mu);
    // the tag is reserved before any wait, so a second submit of it fails
    if (s->live.count(tag)) return fail(GPZ_E_INVAL, "tag %llu already in use", (unsigned long long)tag);
    s->live[tag] = std::move(owned);
    const uint64_t ticket = s->next_ticket++;
    bool blocked = false;
    s->cv_admit.wait(lk, [&] {
        if (s->stopping) return true;
        if (ticket == s->admit_ticket && s->inputs.alloc(need, &r->dev_off)) return true;
        blocked = true;
        return false;
    });
    if (s->stopping) {
        if (ticket == s->admit_ticket) s->admit_ticket++;
        s->live.erase(tag);
        return fail(GPZ_E_INVAL, "server stopping");
    }
    s->admit_ticket++;
    r->dev_len = need;
    r->admitted = true;
    if (blocked) s->waited++;
    s->peak = std::max(s->peak, s->ws_bytes + s->inputs.used);
    lk.unlock();
    s->cv_admit.notify_all();  // the next ticket may be admitted now
    // the input goes to its pool range and stays there until compressed
    int rc = GPZ_OK;
    if (len) {
        hipStream_t cs = s->take_copy_stream();
        if (hipSetDevice(ctx_device(s->ctx)) != hipSuccess ||
            hipMemcpyAsync(s->pool + r->dev_off, in, len, hipMemcpyHostToDevice, cs) != hipSuccess ||
            hipStreamSynchronize(cs) != hipSuccess)
            rc = fail(GPZ_E_DEVICE, "input upload of %zu bytes failed", len);
        s->give_copy_stream(cs);
    }
    lk.lock();
    if (rc) {
        s->inputs.release(r->dev_off, r->dev_len);
        s->live.erase(tag);
        lk.unlock();
        s->cv_admit.notify_all();
        return rc;
    }
    for (uint32_t i = 0; i < r->nchunks; i++) {
        const uint64_t off = (uint64_t)i * cb;
        s->pending.push_back(ChunkRef{r, i, r->dev_off + off, std::min<uint64_t>(cb, len - std::min<uint64_t>(len, off))});
    }
    lk.unlock();
    s->cv_work.notify_all();
    return GPZ_OK;
}

int gpz_request_next(gpz_server *s, int timeout_ms, uint64_t *tag, const uint8_t **data, size_t *len,
                     double *latency_s) {
    if (!s || !tag || !data || !len) return fail(GPZ_E_INVAL, "bad argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto ready = [&] { return !s->finished.empty(); };
    if (timeout_ms < 0)
        s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return GPZ_E_AGAIN;
    Request *r = s->finished.front();
    s->finished.pop_front();
    r->handed_out = true;
    *tag = r->tag;
    *data = r->out.data();
    *len = r->out.size();
    if (latency_s) *latency_s = std::chrono::duration<double>(r->t_done - r->t_submit).count();
    return r->status;
}

int gpz_request_release(gpz_server *s, uint64_t tag) {
    if (!s) return fail(GPZ_E_INVAL, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->live.find(tag);
    if (it == s->live.end()) return fail(GPZ_E_INVAL, "unknown tag");
    if (!it->second->handed_out) return fail(GPZ_E_INVAL, "request not yet returned by gpz_request_next");
    s->live.erase(it);
    return GPZ_OK;
}

int gpz_server_stats(gpz_server *s, uint64_t *out, size_t n) {
    if (!s || !out) return fail(GPZ_E_INVAL, "bad argument");
    std::lock_guard<std::mutex> lk(s->mu);
    const uint64_t v[7] = {s->n_done, s->bytes_done, s->steps, s->peak,
                           s->ws_bytes + s->inputs.cap, s->waited, s->ws_bytes};
    for (size_t i = 0; i < n && i < 7; i++) out[i] = v[i];
    return GPZ_OK;
}

void gpz_server_stop(gpz_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stopping = true;
    }
    s->cv_work.notify_all();
    s->cv_admit.notify_all();
    for (auto &t : s->workers) t.join();
    (void)hipSetDevice(ctx_device(s->ctx));
    delete s;
}

}  // extern "C"
