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
//
// Errors are thrown as gpzip::Error with the reference's error kinds.
#pragma once
#include <stddef.h>
#include <stdint.h>

#include <condition_variable>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

struct gpz_ctx;

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
    std::unique_ptr<Writer> out_;
    size_t chunk_size_, max_in_flight_;
    ChunkFn fn_;
    std::vector<uint8_t> current_;
    uint32_t next_index_ = 0, write_index_ = 0;
    size_t in_flight_ = 0;
    std::map<uint32_t, std::vector<uint8_t>> pending_;
    std::mutex mu_;
    std::condition_variable cv_;
    std::vector<std::thread> threads_;
    std::exception_ptr sticky_;
    bool finalized_ = false;
};

class Compressor {
  public:
    virtual ~Compressor() = default;
    virtual Algorithm algorithm() const = 0;
    virtual std::unique_ptr<Writer> wrap_writer(std::unique_ptr<Writer> w) = 0;
};

class CodecBackend {
  public:
    virtual ~CodecBackend() = default;
    virtual const char *name() const = 0;
    virtual Capability supports(Algorithm a) const = 0;
    virtual std::unique_ptr<Compressor> compressor(Algorithm a, Level level) const = 0;
};

class GpuBackend : public CodecBackend, public std::enable_shared_from_this<GpuBackend> {
  public:
    static constexpr const char *NAME = "gpu";
    static std::shared_ptr<GpuBackend> try_init();  // throws Error(Codec)
    ~GpuBackend() override;
    size_t chunk_size() const { return chunk_size_; }
    size_t max_in_flight() const { return max_in_flight_; }
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
};

class LazyGpuBackend : public CodecBackend {
  public:
    LazyGpuBackend() = default;
    const std::shared_ptr<GpuBackend> *try_get() const;  // nullptr when no GPU
    const char *name() const override { return GpuBackend::NAME; }
    Capability supports(Algorithm a) const override;
    std::unique_ptr<Compressor> compressor(Algorithm a, Level level) const override;

  private:
    mutable std::once_flag once_;
    mutable std::shared_ptr<GpuBackend> inner_;
};

}  // namespace gpzip
