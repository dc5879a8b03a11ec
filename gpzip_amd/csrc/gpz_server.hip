// gpz_server.hip -- serving form: continuous batching of compress requests.
//
// Requests are admitted against a device-memory budget (each admitted byte of
// input accounts for the device arrays it occupies while in flight: input,
// per-position tokens, walked tokens, output bound, i.e. DEVICE_BYTES_PER_INPUT
// bytes per input byte), split into chunks of at most chunk_bytes and queued.
// `streams` worker threads each own a lane (stream + pooled workspace +
// pinned staging, reused for the server's lifetime); whenever a worker is
// free it gathers queued chunks -- from any number of requests, oldest first
// -- into one device step (a graph replay), so requests that arrive while
// others are on the GPU join the next step and steps of different workers
// overlap on the device.  Steps hold at most max_batch_bytes (8 MiB by
// default), so a large request is spread over several workers and finishes
// sooner; larger steps cost less per byte but were not faster in total.  A request completes when all its
// chunks have; its members are concatenated in input order, which is the
// ordering contract of the reference's ParallelChunkedWriter
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

constexpr uint64_t DEVICE_BYTES_PER_INPUT = 12;

using clk = std::chrono::steady_clock;

struct Request {
    uint64_t tag = 0;
    std::vector<uint8_t> data;
    uint64_t nbytes = 0;
    uint32_t nchunks = 0, done_chunks = 0;
    std::vector<std::vector<uint8_t>> chunk_out;
    std::vector<uint8_t> out;
    clk::time_point t_submit, t_done;
    uint64_t charge = 0;
    int status = GPZ_OK;
    bool handed_out = false;  // returned by gpz_request_next: only then may it be released
};

struct ChunkRef {
    Request *req;
    uint32_t idx;
    uint64_t off, len;
};

}  // namespace

struct gpz_server {
    gpz_ctx *ctx = nullptr;
    gpz_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_admit, cv_done;
    std::deque<ChunkRef> pending;
    std::map<uint64_t, std::unique_ptr<Request>> live;  // admitted, not yet released
    std::deque<Request *> finished;
    uint64_t admitted = 0, peak = 0;
    uint64_t n_done = 0, bytes_done = 0, steps = 0;
    bool stopping = false;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::vector<std::thread> workers;

    void worker(Lane *ln) {
        std::vector<ChunkRef> step;
        std::vector<const uint8_t *> ptrs;
        std::vector<uint64_t> lens, off, len;
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
            ptrs.resize(step.size());
            lens.resize(step.size());
            for (size_t i = 0; i < step.size(); i++) {
                ptrs[i] = step[i].req->data.data() + step[i].off;
                lens[i] = step[i].len;
            }
            Plan pl;
            int rc = build_plan(pl, lens.data(), nullptr, step.size(), ctx_sub_chunk(ctx), 1);
            if (!rc) rc = lane_batch_launch(ctx, ln, pl, ptrs.data(), lens.data(), 1);  // workers copy in parallel already
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
                r->data.clear();
                r->data.shrink_to_fit();
            }
            if (!completed.empty()) {
                std::lock_guard<std::mutex> lk(mu);
                for (Request *r : completed) {
                    r->t_done = clk::now();
                    admitted -= r->charge;
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
    if (cfg.streams > 16) cfg.streams = 16;
    auto s = std::make_unique<gpz_server>();
    s->ctx = ctx;
    s->cfg = cfg;
    for (uint32_t i = 0; i < cfg.streams; i++) {
        auto ln = std::make_unique<Lane>();
        int rc = ln->init(ctx_device(ctx));
        if (rc) return rc;
        s->lanes.push_back(std::move(ln));
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
    uint64_t charge = std::max<uint64_t>(len, 1) * DEVICE_BYTES_PER_INPUT;
    if (charge > s->cfg.budget_bytes) return fail(GPZ_E_BUSY, "request of %zu bytes exceeds the budget", len);
    auto r = std::make_unique<Request>();
    r->tag = tag;
    r->data.assign(in, in + len);
    r->nbytes = len;
    r->charge = charge;
    uint64_t cb = s->cfg.chunk_bytes;
    r->nchunks = len == 0 ? 1 : (uint32_t)((len + cb - 1) / cb);
    r->chunk_out.resize(r->nchunks);
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->live.count(tag)) return fail(GPZ_E_INVAL, "tag %llu already in use", (unsigned long long)tag);
    s->cv_admit.wait(lk, [&] { return s->stopping || s->admitted + charge <= s->cfg.budget_bytes; });
    if (s->stopping) return fail(GPZ_E_INVAL, "server stopping");
    r->t_submit = clk::now();
    s->admitted += charge;
    s->peak = std::max(s->peak, s->admitted);
    Request *rp = r.get();
    s->live[tag] = std::move(r);
    for (uint32_t i = 0; i < rp->nchunks; i++) {
        uint64_t off = (uint64_t)i * cb;
        s->pending.push_back(ChunkRef{rp, i, off, std::min<uint64_t>(cb, len - std::min<uint64_t>(len, off))});
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
    uint64_t v[4] = {s->n_done, s->bytes_done, s->steps, s->peak};
    for (size_t i = 0; i < n && i < 4; i++) out[i] = v[i];
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
    delete s;
}

}  // extern "C"
