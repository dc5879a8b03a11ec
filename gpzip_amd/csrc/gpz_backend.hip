// gpz_backend.hip -- C++ mirror of the reference GPU backend surface
// (include/gpzip_backend.hpp) over the C ABI.  Host code only.
#include <string.h>

#include "gpzip_backend.hpp"
#include "gpzip_hip.h"

namespace gpzip {

// ---------------------------------------------------------------------------
// ParallelChunkedWriter (codec-cpu/src/parallel.rs:29-229)

ParallelChunkedWriter::ParallelChunkedWriter(std::unique_ptr<Writer> out, size_t chunk_size, size_t max_in_flight,
                                             ChunkFn chunk_fn)
    : out_(std::move(out)), chunk_size_(chunk_size), max_in_flight_(max_in_flight), fn_(std::move(chunk_fn)) {
    if (chunk_size_ == 0) throw std::invalid_argument("chunk_size must be > 0");
    if (max_in_flight_ == 0) throw std::invalid_argument("max_in_flight must be > 0");
    current_.reserve(chunk_size_);
}

ParallelChunkedWriter::~ParallelChunkedWriter() {
    try {
        finalize();
    } catch (...) {
        // Drop cannot report errors (parallel.rs:222-229)
    }
}

void ParallelChunkedWriter::check_sticky() {
    std::lock_guard<std::mutex> lk(mu_);
    if (sticky_) std::rethrow_exception(sticky_);
}

// Write every result that is next in index order.  Called by the producer
// (the thread that calls write_all / finalize) with mu_ held; the writes
// themselves run with mu_ released, so workers can post results meanwhile.
void ParallelChunkedWriter::write_ready_locked(std::unique_lock<std::mutex> &lk) {
    for (;;) {
        auto it = pending_.find(write_index_);
        if (it == pending_.end() || sticky_) return;
        std::vector<uint8_t> bytes = std::move(it->second);
        pending_.erase(it);
        write_index_++;
        lk.unlock();
        std::exception_ptr err;
        try {
            out_->write_all(bytes.data(), bytes.size());
        } catch (...) {
            err = std::current_exception();
        }
        lk.lock();
        if (err) {
            if (!sticky_) sticky_ = err;
            return;
        }
    }
}

// A pool of up to max_in_flight workers takes dispatched chunks in order, as
// rayon's pool runs the spawned closures (parallel.rs:101-104).
void ParallelChunkedWriter::worker() {
    std::unique_lock<std::mutex> lk(mu_);
    for (;;) {
        cv_jobs_.wait(lk, [&] { return stop_ || !jobs_.empty(); });
        if (jobs_.empty()) return;  // stop_ and drained
        auto job = std::move(jobs_.front());
        jobs_.pop_front();
        lk.unlock();
        std::vector<uint8_t> res;
        std::exception_ptr err;
        try {
            res = fn_(job.second->data(), job.second->size());
        } catch (...) {
            err = std::current_exception();
        }
        lk.lock();
        if (spare_.size() <= max_in_flight_) {
            job.second->clear();
            spare_.push_back(std::move(*job.second));
        }
        job.second.reset();
        if (err) {
            if (!sticky_) sticky_ = err;
        } else {
            pending_.emplace(job.first, std::move(res));
        }
        in_flight_--;
        cv_.notify_all();
    }
}

void ParallelChunkedWriter::dispatch_current() {
    check_sticky();
    if (current_.empty()) return;
    uint32_t idx = next_index_++;
    auto bytes = std::make_shared<std::vector<uint8_t>>(std::move(current_));
    std::unique_lock<std::mutex> lk(mu_);
    if (!spare_.empty()) {
        current_ = std::move(spare_.back());
        spare_.pop_back();
    } else {
        current_ = std::vector<uint8_t>();
        current_.reserve(chunk_size_);
    }
    in_flight_++;
    jobs_.emplace_back(idx, std::move(bytes));
    if (workers_.size() < max_in_flight_ && workers_.size() < in_flight_) workers_.emplace_back([this] { worker(); });
    cv_jobs_.notify_one();
    write_ready_locked(lk);
    while (in_flight_ >= max_in_flight_ && !sticky_) {
        cv_.wait(lk);
        write_ready_locked(lk);
    }
    if (sticky_) std::rethrow_exception(sticky_);
}

void ParallelChunkedWriter::write_all(const uint8_t *data, size_t len) {
    check_sticky();
    size_t written = 0;
    while (written < len) {
        size_t space = chunk_size_ - current_.size();
        size_t take = std::min(len - written, space);
        current_.insert(current_.end(), data + written, data + written + take);
        written += take;
        if (current_.size() >= chunk_size_) dispatch_current();
    }
}

void ParallelChunkedWriter::flush() {
    // tar and other writers flush between entries; cutting a chunk here would
    // fragment compression (parallel.rs:213-219)
}

void ParallelChunkedWriter::finalize() {
    if (finalized_) {
        check_sticky();
        return;
    }
    finalized_ = true;
    try {
        dispatch_current();
    } catch (...) {
    }
    {
        std::unique_lock<std::mutex> lk(mu_);
        for (;;) {
            write_ready_locked(lk);
            if (in_flight_ == 0) break;
            cv_.wait(lk);
        }
        write_ready_locked(lk);
        stop_ = true;
    }
    cv_jobs_.notify_all();
    for (auto &t : workers_) t.join();
    workers_.clear();
    try {
        out_->flush();
    } catch (...) {
        std::lock_guard<std::mutex> lk(mu_);
        if (!sticky_) sticky_ = std::current_exception();
    }
    check_sticky();
}

// ---------------------------------------------------------------------------
// GpuBackend (codec-gpu/src/gpu.rs:98-243)

namespace {

class GpuGzipCompressor : public Compressor {
  public:
    GpuGzipCompressor(ChunkFn fn, size_t chunk, size_t in_flight) : fn_(std::move(fn)), chunk_(chunk), inflight_(in_flight) {}
    Algorithm algorithm() const override { return Algorithm::Gzip; }
    std::unique_ptr<Writer> wrap_writer(std::unique_ptr<Writer> w) override {
        return std::make_unique<ParallelChunkedWriter>(std::move(w), chunk_, inflight_, fn_);
    }

  private:
    ChunkFn fn_;
    size_t chunk_, inflight_;
};

}  // namespace

std::shared_ptr<GpuBackend> GpuBackend::try_init(uint32_t lanes) {
    gpz_config cfg;
    gpz_config_default(&cfg);
    cfg.lanes = lanes ? lanes : cfg.max_in_flight;
    gpz_ctx *ctx = nullptr;
    int rc = gpz_init(&cfg, &ctx);
    if (rc != GPZ_OK) throw Error(Error::Kind::Codec, std::string("gpu init failed: ") + gpz_last_error());
    std::shared_ptr<GpuBackend> b(new GpuBackend());
    b->ctx_ = ctx;
    b->chunk_size_ = cfg.sub_chunk;
    b->max_in_flight_ = cfg.max_in_flight;
    b->writer_chunk_ = (size_t)cfg.sub_chunk * 16;
    return b;
}

GpuBackend::~GpuBackend() {
    if (ctx_) gpz_destroy(ctx_);
}

ChunkFn GpuBackend::gzip_chunk_fn() const {
    std::shared_ptr<const GpuBackend> self = shared_from_this();
    return [self](const uint8_t *data, size_t len) {
        // the bound is twice the chunk: a per-thread buffer of that size,
        // reused across calls (no zero fill, no fresh pages per chunk), and
        // one copy of what is used into the result
        thread_local std::unique_ptr<uint8_t[]> buf;
        thread_local size_t buf_cap = 0;
        const size_t cap = gpz_compress_bound(self->ctx_, len);
        if (cap > buf_cap) {
            buf.reset(new uint8_t[cap]);
            buf_cap = cap;
        }
        size_t n = 0;
        int rc = gpz_compress_chunk(self->ctx_, data, len, buf.get(), cap, &n);
        if (rc != GPZ_OK) throw Error(Error::Kind::Io, std::string("gpu compress failed: ") + gpz_last_error());
        return std::vector<uint8_t>(buf.get(), buf.get() + n);
    };
}

Capability GpuBackend::supports(Algorithm a) const {
    return a == Algorithm::Gzip ? Capability::CompressOnly : Capability::None;
}

std::unique_ptr<Compressor> GpuBackend::compressor(Algorithm a, Level) const {
    if (a != Algorithm::Gzip) throw Error(Error::Kind::UnsupportedAlgorithm, "backend `gpu` does not support this algorithm");
    return std::make_unique<GpuGzipCompressor>(gzip_chunk_fn(), writer_chunk_, max_in_flight_);
}

const std::shared_ptr<GpuBackend> *LazyGpuBackend::try_get() const {
    std::call_once(once_, [this] {
        try {
            inner_ = GpuBackend::try_init(lanes_);
        } catch (...) {
            inner_.reset();
        }
    });
    return inner_ ? &inner_ : nullptr;
}

Capability LazyGpuBackend::supports(Algorithm a) const {
    return a == Algorithm::Gzip ? Capability::CompressOnly : Capability::None;
}

std::unique_ptr<Compressor> LazyGpuBackend::compressor(Algorithm a, Level level) const {
    auto g = try_get();
    if (!g) throw Error(Error::Kind::NoBackend, "no backend available for gzip");
    return (*g)->compressor(a, level);
}

}  // namespace gpzip
