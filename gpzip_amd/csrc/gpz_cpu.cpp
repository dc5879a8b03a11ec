// gpz_cpu.cpp -- the CPU side of gpzip that sits next to the GPU compress
// path, mirrored on system zlib (host code only):
//
//   cpu_gzip_member / CpuBackend   gpzip-codec-cpu's gzip chunk function and
//                                  backend (codec-cpu/src/compress.rs:52-59,
//                                  codec-cpu/src/backend.rs:14-60): one gzip
//                                  member per chunk, flate2's fixed header
//   parallel_decompress            codec-cpu/src/parallel_decompress.rs:35-113:
//                                  scan for the fixed 10-byte member header,
//                                  inflate the members on worker threads,
//                                  fall back to a serial multi-member inflate
//                                  when fewer than two members are found or
//                                  any member fails to decode
//   GzipDecompressor               codec-cpu/src/decompress.rs:28-44
//   gpz_gzip_member_starts,        C ABI of the scan and of the decoder
//   gpz_gunzip
//
// The reference runs these on flate2 / zlib-rs 0.6.3; system zlib produces
// different compressed bytes for the same level, and the same decoded bytes.
#include <string.h>
#include <zlib.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <vector>

#include "gpzip_backend.hpp"
#include "gpzip_hip.h"

namespace gpz {
int fail(int code, const char *fmt, ...);  // gpz_runtime.hip: sets gpz_last_error()
void clear_error();
}  // namespace gpz

namespace gpzip {

namespace {

// flate2 GzEncoder's header at levels 2-8: no name, mtime 0, XFL 0, OS 255
// (the GPU path writes the same bytes, codec-gpu/src/gpu/deflate.rs:965-972)
constexpr uint8_t GZ_HEADER[10] = {0x1f, 0x8b, 0x08, 0x00, 0, 0, 0, 0, 0x00, 0xff};

uint32_t le32(const uint8_t *p) { return (uint32_t)p[0] | (uint32_t)p[1] << 8 | (uint32_t)p[2] << 16 | (uint32_t)p[3] << 24; }

// Raw-inflate exactly [in, in + len) into out[0, cap); true when the DEFLATE
// stream ends at the last input byte and produced exactly cap bytes.
bool inflate_exact(const uint8_t *in, size_t len, uint8_t *out, size_t cap) {
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (inflateInit2(&zs, -15) != Z_OK) return false;
    bool ok = true;
    size_t in_done = 0, out_done = 0;
    int rc = Z_OK;
    // zlib counts in uInt: feed at most 1 GiB per call
    while (rc == Z_OK) {
        const size_t ci = std::min<size_t>(len - in_done, 1u << 30), co = std::min<size_t>(cap - out_done, 1u << 30);
        zs.next_in = const_cast<Bytef *>(in + in_done);
        zs.avail_in = (uInt)ci;
        zs.next_out = out + out_done;
        zs.avail_out = (uInt)co;
        rc = inflate(&zs, Z_NO_FLUSH);
        in_done += ci - zs.avail_in;
        out_done += co - zs.avail_out;
        if (rc == Z_BUF_ERROR) {
            ok = false;  // out of input or output before the end of the stream
            break;
        }
    }
    if (rc != Z_STREAM_END || in_done != len || out_done != cap) ok = false;
    inflateEnd(&zs);
    return ok;
}

// Serial decoder of any concatenation of gzip members (flate2's
// MultiGzDecoder): zlib's gzip wrapper, reset after each member.
bool serial_gunzip(const uint8_t *in, size_t len, std::vector<uint8_t> &out) {
    out.clear();
    if (!len) return true;
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    if (inflateInit2(&zs, 16 + 15) != Z_OK) return false;
    size_t pos = 0;
    std::vector<uint8_t> buf(1 << 20);
    bool ok = true, ended = false;
    while (pos < len || (!ended && zs.avail_out == 0)) {
        const size_t ci = std::min<size_t>(len - pos, 1u << 30);
        zs.next_in = const_cast<Bytef *>(in + pos);
        zs.avail_in = (uInt)ci;
        zs.next_out = buf.data();
        zs.avail_out = (uInt)buf.size();
        const int rc = inflate(&zs, Z_NO_FLUSH);
        out.insert(out.end(), buf.data(), buf.data() + (buf.size() - zs.avail_out));
        pos += ci - zs.avail_in;
        ended = rc == Z_STREAM_END;
        if (ended) {
            if (pos < len && inflateReset(&zs) != Z_OK) {
                ok = false;
                break;
            }
        } else if (rc != Z_OK && !(rc == Z_BUF_ERROR && zs.avail_out == 0)) {
            ok = false;
            break;
        }
    }
    inflateEnd(&zs);
    return ok && ended;  // a truncated last member is an error
}

}  // namespace

std::vector<uint8_t> cpu_gzip_member(const uint8_t *data, size_t len, Level level) {
    z_stream zs;
    memset(&zs, 0, sizeof(zs));
    const int lvl = std::min<int>(level.v, 9);
    if (deflateInit2(&zs, lvl, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK)
        throw Error(Error::Kind::Io, "deflateInit2 failed");
    std::vector<uint8_t> out(10 + deflateBound(&zs, (uLong)len) + 8 + 64);
    memcpy(out.data(), GZ_HEADER, 10);
    zs.next_in = const_cast<Bytef *>(data);
    zs.avail_in = (uInt)len;
    zs.next_out = out.data() + 10;
    zs.avail_out = (uInt)(out.size() - 18);
    const int rc = deflate(&zs, Z_FINISH);
    const size_t dl = zs.total_out;
    deflateEnd(&zs);
    if (rc != Z_STREAM_END) throw Error(Error::Kind::Io, "deflate failed");
    const uint32_t crc = (uint32_t)crc32(0L, data, (uInt)len), isz = (uint32_t)len;
    for (int i = 0; i < 4; i++) out[10 + dl + i] = (uint8_t)(crc >> (8 * i));
    for (int i = 0; i < 4; i++) out[14 + dl + i] = (uint8_t)(isz >> (8 * i));
    out.resize(dl + 18);
    return out;
}

ChunkFn GzipCompressor::chunk_fn(Level level) {
    return [level](const uint8_t *d, size_t n) { return cpu_gzip_member(d, n, level); };
}

namespace {

class CpuGzipCompressor : public Compressor {
  public:
    CpuGzipCompressor(Level level, size_t chunk, size_t in_flight) : level_(level), chunk_(chunk), inflight_(in_flight) {}
    Algorithm algorithm() const override { return Algorithm::Gzip; }
    std::unique_ptr<Writer> wrap_writer(std::unique_ptr<Writer> w) override {
        return std::make_unique<ParallelChunkedWriter>(std::move(w), chunk_, inflight_, GzipCompressor::chunk_fn(level_));
    }

  private:
    Level level_;
    size_t chunk_, inflight_;
};

class MemReader : public Reader {
  public:
    explicit MemReader(std::vector<uint8_t> b) : b_(std::move(b)) {}
    size_t read(uint8_t *buf, size_t cap) override {
        const size_t n = std::min(cap, b_.size() - pos_);
        memcpy(buf, b_.data() + pos_, n);
        pos_ += n;
        return n;
    }

  private:
    std::vector<uint8_t> b_;
    size_t pos_ = 0;
};

}  // namespace

CpuBackend::CpuBackend(size_t chunk_size, size_t max_in_flight) : chunk_size_(chunk_size), max_in_flight_(max_in_flight) {
    if (!chunk_size_) throw std::invalid_argument("chunk_size must be > 0");
    if (!max_in_flight_) max_in_flight_ = std::max(1u, std::thread::hardware_concurrency());
}

Capability CpuBackend::supports(Algorithm a) const {
    return a == Algorithm::Gzip ? Capability::Both : Capability::None;
}

std::unique_ptr<Compressor> CpuBackend::compressor(Algorithm a, Level level) const {
    if (a != Algorithm::Gzip) throw Error(Error::Kind::UnsupportedAlgorithm, "backend `cpu` mirrors gzip only");
    return std::make_unique<CpuGzipCompressor>(level, chunk_size_, max_in_flight_);
}

std::unique_ptr<Decompressor> CpuBackend::decompressor(Algorithm a) const {
    if (a != Algorithm::Gzip) throw Error(Error::Kind::UnsupportedAlgorithm, "backend `cpu` mirrors gzip only");
    return std::make_unique<GzipDecompressor>();
}

std::vector<size_t> find_member_starts(const uint8_t *data, size_t len) {
    std::vector<size_t> starts;
    size_t i = 0;
    while (i + sizeof(GZ_HEADER) <= len) {
        const uint8_t *hit = (const uint8_t *)memchr(data + i, 0x1f, len - sizeof(GZ_HEADER) + 1 - i);
        if (!hit) break;
        i = (size_t)(hit - data);
        if (memcmp(hit, GZ_HEADER, sizeof(GZ_HEADER)) == 0) {
            starts.push_back(i);
            i += sizeof(GZ_HEADER);  // a header cannot start inside a header
        } else {
            i++;
        }
    }
    return starts;
}

// Members decoded on `threads` workers straight into their places in out.
// False if a member does not decode to its trailer's length and CRC.
static bool parallel_members(const uint8_t *in, size_t len, const std::vector<size_t> &starts, uint8_t *out,
                             const std::vector<uint64_t> &off, unsigned threads) {
    const size_t n = starts.size();
    std::atomic<size_t> next{0};
    std::atomic<bool> ok{true};
    auto work = [&] {
        for (size_t i = next++; i < n && ok; i = next++) {
            const size_t a = starts[i], b = i + 1 < n ? starts[i + 1] : len;
            if (b - a < 18) {
                ok = false;
                break;
            }
            const uint64_t isz = off[i + 1] - off[i];
            if (!inflate_exact(in + a + 10, b - a - 18, out + off[i], isz) ||
                (uint32_t)crc32(0L, out + off[i], (uInt)isz) != le32(in + b - 8))
                ok = false;
        }
    };
    std::vector<std::thread> pool;
    for (unsigned t = 1; t < threads; t++) pool.emplace_back(work);
    work();
    for (auto &t : pool) t.join();
    return ok;
}

std::vector<uint8_t> parallel_decompress(const uint8_t *in, size_t len, unsigned threads) {
    std::vector<uint8_t> out;
    if (!len) return out;
    if (!threads) threads = std::max(1u, std::thread::hardware_concurrency());
    const std::vector<size_t> starts = find_member_starts(in, len);
    // at least two members (and the stream starting with one) for the parallel path
    if (starts.size() >= 2 && starts[0] == 0) {
        std::vector<uint64_t> off(starts.size() + 1, 0);
        bool sane = true;
        for (size_t i = 0; i < starts.size(); i++) {
            const size_t b = i + 1 < starts.size() ? starts[i + 1] : len;
            if (b - starts[i] < 18) {
                sane = false;
                break;
            }
            off[i + 1] = off[i] + le32(in + b - 4);  // ISIZE (members here are < 4 GiB)
        }
        if (sane) {
            out.resize(off.back());
            if (parallel_members(in, len, starts, out.data(), off, (unsigned)std::min<size_t>(threads, starts.size())))
                return out;
        }
    }
    // one member, a stream not written by gpzip, or a boundary mis-detected
    // inside DEFLATE bytes: the serial decoder over the whole stream
    if (!serial_gunzip(in, len, out)) throw Error(Error::Kind::Io, "corrupt gzip stream");
    return out;
}

std::unique_ptr<Reader> GzipDecompressor::wrap_reader(std::unique_ptr<Reader> r) {
    std::vector<uint8_t> compressed;
    uint8_t buf[1 << 16];
    for (;;) {
        const size_t n = r->read(buf, sizeof(buf));
        if (!n) break;
        compressed.insert(compressed.end(), buf, buf + n);
    }
    return std::make_unique<MemReader>(parallel_decompress(compressed.data(), compressed.size(), 0));
}

}  // namespace gpzip

extern "C" {

int gpz_gzip_member_starts(const uint8_t *in, size_t len, uint64_t *starts, size_t cap, size_t *count) {
    gpz::clear_error();
    if ((!in && len) || !count || (cap && !starts)) return gpz::fail(GPZ_E_INVAL, "bad argument");
    const std::vector<size_t> s = gpzip::find_member_starts(in, len);
    *count = s.size();
    for (size_t i = 0; i < s.size() && i < cap; i++) starts[i] = s[i];
    return s.size() > cap ? gpz::fail(GPZ_E_NOSPACE, "%zu member starts", s.size()) : GPZ_OK;
}

int gpz_gunzip(const uint8_t *in, size_t len, uint8_t *out, size_t out_cap, size_t *out_len, unsigned threads) {
    gpz::clear_error();
    if ((!in && len) || !out_len || (out_cap && !out)) return gpz::fail(GPZ_E_INVAL, "bad argument");
    try {
        const std::vector<uint8_t> d = gpzip::parallel_decompress(in, len, threads);
        *out_len = d.size();
        if (d.size() > out_cap) return gpz::fail(GPZ_E_NOSPACE, "output needs %zu bytes", d.size());
        if (!d.empty()) memcpy(out, d.data(), d.size());
        return GPZ_OK;
    } catch (const std::exception &e) {
        *out_len = 0;
        return gpz::fail(GPZ_E_INVAL, "%s", e.what());
    }
}

}  // extern "C"
