// gpzip_backend.hpp -- C++ host interface over libgpzip_hip.so that mirrors
// the reference's GPU backend surface, so a caller written against
// gpzip-codec-gpu finds the same names, argument meaning and errors:
//
//   ChunkFn                codec-cpu/src/parallel.rs:21
//   ParallelChunkedWriter  codec-cpu/src/parallel.rs:29-229 (in-order members,
//                          bounded in-flight, sticky first error, finalize on
//                          destruction; flush() does not cut a chunk)
//   CodecBackend           core/src/codec.rs:23-27
//   Algorithm/Capability   core/src/algorithm.rs:3-58
//   GpuBackend             codec-gpu/src/gpu.rs:98-243 (name "gpu", Gzip
//                          compress-only, 128 KiB members, 16 in flight)
//   LazyGpuBackend         codec-gpu/src/gpu.rs:28-96 (init once on first use;
//                          try_get() == nullptr when no GPU, so callers fall
//                          back to CPU as cli/src/hybrid.rs:128-136 does)
//   CpuBackend,            codec-cpu/src/backend.rs:14-60, compress.rs:52-59
//   GzipCompressor         (gzip only: one zlib gzip member per 2 MiB chunk)
//   parallel_decompress,   codec-cpu/src/parallel_decompress.rs:35-113,
//   GzipDecompressor       codec-cpu/src/decompress.rs:28-44 (the decode side)
//   HybridBackend          cli/src/hybrid.rs:26-163 (`--backend auto`)
//   TarBuilder,            core/src/archive/pack.rs:103-176 (`gpzip a x.tar.gz`),
//   BatchedChunkedWriter,  with the chunks of the stream submitted to the GPU
//   pack_tar_gz            as batches of (offset, len) descriptors
//
// Errors are thrown as gpzip::Error with the reference's error kinds.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <atomic>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

struct gpz_ctx;
struct gpz_batch;

namespace gpzip {

enum class Algorithm { Deflate, Gzip, Zstd, Lzma, Bzip2, Rar };
enum class Capability { None, DecompressOnly, CompressOnly, Both };
struct Level {
    uint8_t v = 5;
};

class Error : public std::runtime_error {
  public:
    enum class Kind { Io, UnsupportedAlgorithm, CompressionUnsupported, DecompressionUnsupported, NoBackend, Codec };
    Error(Kind k, const std::string &msg) : std::runtime_error(msg), kind(k) {}
    Kind kind;
};

// Output sink (Box<dyn Write + Send> in the reference).
class Writer {
  public:
    virtual ~Writer() = default;
    virtual void write_all(const uint8_t *data, size_t len) = 0;  // throws Error(Io)
    virtual void flush() {}
};

// Compress one chunk into one self-contained member; throws Error on failure.
using ChunkFn = std::function<std::vector<uint8_t>(const uint8_t *, size_t)>;

class ParallelChunkedWriter : public Writer {
  public:
    ParallelChunkedWriter(std::unique_ptr<Writer> out, size_t chunk_size, size_t max_in_flight, ChunkFn chunk_fn);
    ~ParallelChunkedWriter() override;  // finalize(), errors swallowed (reference logs them)
    void write_all(const uint8_t *data, size_t len) override;
    void flush() override;  // intentionally does not dispatch a partial chunk
    void finalize();        // idempotent; rethrows the sticky error
    Writer *inner() { return out_.get(); }

  private:
    void dispatch_current();
    void write_ready_locked(std::unique_lock<std::mutex> &lk);
    void check_sticky();
    void worker();
    std::unique_ptr<Writer> out_;
    size_t chunk_size_, max_in_flight_;
    ChunkFn fn_;
    std::vector<uint8_t> current_;
    uint32_t next_index_ = 0, write_index_ = 0;
    size_t in_flight_ = 0;  // dispatched, result not yet posted
    std::map<uint32_t, std::vector<uint8_t>> pending_;
    std::deque<std::pair<uint32_t, std::shared_ptr<std::vector<uint8_t>>>> jobs_;
    // input buffers whose chunk is done, kept for the next chunks: a fresh
    // chunk-sized allocation per chunk costs the producer a page fault per
    // 4 KiB while it copies, which caps a single producer near 2 GB/s
    std::vector<std::vector<uint8_t>> spare_;
    std::mutex mu_;
    std::condition_variable cv_, cv_jobs_;
    std::vector<std::thread> workers_;
    std::exception_ptr sticky_;
    bool finalized_ = false, stop_ = false;
};

// Input source (Box<dyn Read + Send> in the reference); read() returns 0 at the end.
class Reader {
  public:
    virtual ~Reader() = default;
    virtual size_t read(uint8_t *buf, size_t cap) = 0;  // throws Error(Io)
};

class Compressor {
  public:
    virtual ~Compressor() = default;
    virtual Algorithm algorithm() const = 0;
    virtual std::unique_ptr<Writer> wrap_writer(std::unique_ptr<Writer> w) = 0;
};

class Decompressor {
  public:
    virtual ~Decompressor() = default;
    virtual Algorithm algorithm() const = 0;
    virtual std::unique_ptr<Reader> wrap_reader(std::unique_ptr<Reader> r) = 0;
};

class CodecBackend {
  public:
    virtual ~CodecBackend() = default;
    virtual const char *name() const = 0;
    virtual Capability supports(Algorithm a) const = 0;
    virtual std::unique_ptr<Compressor> compressor(Algorithm a, Level level) const = 0;
    // compress-only backends (the GPU one, codec-gpu/src/gpu.rs:237-242)
    virtual std::unique_ptr<Decompressor> decompressor(Algorithm) const {
        throw Error(Error::Kind::DecompressionUnsupported, std::string("backend `") + name() + "` cannot decompress");
    }
};

// ---- CPU side (host cores, system zlib) -------------------------------------

// One gzip member of a chunk at `level` (flate2 GzEncoder's fixed header).
std::vector<uint8_t> cpu_gzip_member(const uint8_t *data, size_t len, Level level);

struct GzipCompressor {
    static ChunkFn chunk_fn(Level level);  // codec-cpu/src/compress.rs:52-59
};

class CpuBackend : public CodecBackend {
  public:
    static constexpr const char *NAME = "cpu";
    static constexpr size_t DEFAULT_CHUNK_SIZE = 2u << 20;  // codec-cpu/src/backend.rs:13
    // max_in_flight 0: one per hardware thread (num_cpus)
    explicit CpuBackend(size_t chunk_size = DEFAULT_CHUNK_SIZE, size_t max_in_flight = 0);
    size_t chunk_size() const { return chunk_size_; }
    size_t max_in_flight() const { return max_in_flight_; }
    const char *name() const override { return NAME; }
    Capability supports(Algorithm a) const override;
    std::unique_ptr<Compressor> compressor(Algorithm a, Level level) const override;
    std::unique_ptr<Decompressor> decompressor(Algorithm a) const override;

  private:
    size_t chunk_size_, max_in_flight_;
};

// Offsets of every fixed 10-byte gzip member header (1f 8b 08 00 00000000 00 ff).
std::vector<size_t> find_member_starts(const uint8_t *data, size_t len);
// Decode a whole gzip stream: members in parallel on `threads` workers (0: one
// per hardware thread) when the scan finds two or more, else (or when any
// member fails) a serial multi-member decode.  Throws Error(Io) on a corrupt stream.
std::vector<uint8_t> parallel_decompress(const uint8_t *data, size_t len, unsigned threads = 0);

class GzipDecompressor : public Decompressor {
  public:
    Algorithm algorithm() const override { return Algorithm::Gzip; }
    std::unique_ptr<Reader> wrap_reader(std::unique_ptr<Reader> r) override;  // buffers the whole input
};

class GpuBackend : public CodecBackend, public std::enable_shared_from_this<GpuBackend> {
  public:
    static constexpr const char *NAME = "gpu";
    // lanes: concurrent GPU pipelines (0: max_in_flight, one per chunk in flight)
    static std::shared_ptr<GpuBackend> try_init(uint32_t lanes = 0);  // throws Error(Codec)
    ~GpuBackend() override;
    // the member size (gpu.rs:160): the chunk function cuts its input into
    // members of this size
    size_t chunk_size() const { return chunk_size_; }
    size_t max_in_flight() const { return max_in_flight_; }
    // the compressor's writer chunk: a multiple of chunk_size, so the members
    // (and bytes) are those of chunk_size chunks while each GPU call carries
    // 16 members instead of one
    size_t writer_chunk() const { return writer_chunk_; }
    gpz_ctx *context() const { return ctx_; }
    ChunkFn gzip_chunk_fn() const;
    const char *name() const override { return NAME; }
    Capability supports(Algorithm a) const override;
    std::unique_ptr<Compressor> compressor(Algorithm a, Level level) const override;

  private:
    GpuBackend() = default;
    gpz_ctx *ctx_ = nullptr;
    size_t chunk_size_ = 128 * 1024;
    size_t max_in_flight_ = 16;
    size_t writer_chunk_ = 16 * 128 * 1024;
};

class LazyGpuBackend : public CodecBackend {
  public:
    LazyGpuBackend() = default;
    // lanes: concurrent GPU pipelines of the context (0: the library default)
    explicit LazyGpuBackend(uint32_t lanes) : lanes_(lanes) {}
    const std::shared_ptr<GpuBackend> *try_get() const;  // nullptr when no GPU
    const char *name() const override { return GpuBackend::NAME; }
    Capability supports(Algorithm a) const override;
    std::unique_ptr<Compressor> compressor(Algorithm a, Level level) const override;

  private:
    uint32_t lanes_ = 0;
    mutable std::once_flag once_;
    mutable std::shared_ptr<GpuBackend> inner_;
};

// ---- `gpzip a` onto the batched GPU form --------------------------------------

// ParallelChunkedWriter's contract -- chunk_size chunks, members in chunk
// order, flush() does not cut a chunk, sticky first error -- with every
// batch_chunks chunks submitted as one device batch (gpz_batch_submit, one
// graph replay); up to max_batches batches in flight.  The members equal the
// per-chunk GPU writer's for the same chunk size.
class BatchedChunkedWriter : public Writer {
  public:
    BatchedChunkedWriter(std::unique_ptr<Writer> out, gpz_ctx *ctx, size_t chunk_size = 2u << 20,
                         size_t batch_chunks = 64, size_t max_batches = 2);
    ~BatchedChunkedWriter() override;
    void write_all(const uint8_t *data, size_t len) override;
    void flush() override;
    void finalize();  // idempotent; throws the sticky error
    size_t batches_submitted() const { return batches_submitted_; }
    size_t chunks_submitted() const { return chunks_submitted_; }

  private:
    void submit();
    void drain_one();
    void check_sticky();
    std::unique_ptr<Writer> out_;
    gpz_ctx *ctx_;
    size_t chunk_size_, batch_chunks_, max_batches_;
    std::vector<uint8_t> buf_;
    std::deque<gpz_batch *> inflight_;
    std::deque<size_t> counts_;
    std::string sticky_;
    bool finalized_ = false;
    size_t batches_submitted_ = 0, chunks_submitted_ = 0;
};

// The tar stream of pack_tar (GNU headers; directories, then their entries in
// name order; symlinks stored, not followed; two zero blocks at the end).
class TarBuilder {
  public:
    explicit TarBuilder(std::unique_ptr<Writer> out);
    ~TarBuilder();  // finish(), errors swallowed
    void append_path(const std::string &src, const std::string &arc_path);
    void append_data(const std::string &arc_path, const uint8_t *data, size_t len, uint32_t mode = 0644,
                     uint64_t mtime = 0);
    void finish();
    Writer *inner() { return out_.get(); }

  private:
    void header(const std::string &path, char type, uint64_t size, uint32_t mode, uint64_t mtime,
                const std::string &link);
    static void finish_header(uint8_t *h);
    void pad(uint64_t size);
    std::unique_ptr<Writer> out_;
    bool finished_ = false;
};

// `gpzip a archive.tar.gz inputs...` on the GPU's batched form.
void pack_tar_gz(const std::string &archive, const std::vector<std::string> &inputs, gpz_ctx *ctx,
                 size_t chunk_size = 2u << 20, size_t batch_chunks = 64);

// ---- `--backend auto` (cli/src/hybrid.rs) -----------------------------------

struct HybridStats {
    std::atomic<uint64_t> gpu_chunks{0}, cpu_chunks{0};
};

struct HybridTuning {
    size_t gpu_workers = 1;    // GPU permits (hybrid.rs:72)
    size_t warmup_chunks = 4;  // first chunks always on the CPU (hybrid.rs:118)
    // with no permit free, wait this long for one before the chunk goes to
    // the CPU (the reference falls back at once: 0)
    uint32_t gpu_wait_ms = 0;
    std::shared_ptr<HybridStats> stats;  // optional counters
    static HybridTuning reference() { return HybridTuning{}; }
    static HybridTuning retuned();
};

class HybridBackend : public CodecBackend {
  public:
    static constexpr const char *NAME = "hybrid";
    HybridBackend(std::shared_ptr<CpuBackend> cpu, std::shared_ptr<LazyGpuBackend> gpu,
                  HybridTuning t = HybridTuning::reference());
    const char *name() const override { return NAME; }
    Capability supports(Algorithm a) const override;
    std::unique_ptr<Compressor> compressor(Algorithm a, Level level) const override;
    std::unique_ptr<Decompressor> decompressor(Algorithm a) const override;

  private:
    std::shared_ptr<CpuBackend> cpu_;
    std::shared_ptr<LazyGpuBackend> gpu_;
    HybridTuning t_;
};

}  // namespace gpzip
