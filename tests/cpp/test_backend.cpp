// Tests of the C++ backend mirror (include/gpzip_backend.hpp), modelled on the
// reference's scheduler tests (codec-cpu/src/parallel.rs:231-424).
//   test_backend cpu        -- ParallelChunkedWriter with an identity framer, no GPU
//   test_backend gpu FILE   -- GpuBackend compressor over FILE, output to stdout
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <memory>
#include <string>
#include <thread>
#include <vector>

#include "gpzip_backend.hpp"

using namespace gpzip;

#define CHECK(c)                                                             \
    do {                                                                     \
        if (!(c)) {                                                          \
            fprintf(stderr, "CHECK failed %s:%d: %s\n", __FILE__, __LINE__, #c); \
            exit(1);                                                         \
        }                                                                    \
    } while (0)

struct Sink : Writer {
    std::shared_ptr<std::vector<uint8_t>> buf = std::make_shared<std::vector<uint8_t>>();
    int fail_after = -1;
    int writes = 0;
    void write_all(const uint8_t *d, size_t n) override {
        if (fail_after >= 0 && writes >= fail_after) throw Error(Error::Kind::Io, "sink failed");
        writes++;
        buf->insert(buf->end(), d, d + n);
    }
};

// identity framer: [u32 len][bytes]
static ChunkFn identity_framer(int delay_us = 0) {
    return [delay_us](const uint8_t *d, size_t n) {
        if (delay_us) std::this_thread::sleep_for(std::chrono::microseconds(delay_us * (int)(n % 7)));
        std::vector<uint8_t> out(4 + n);
        uint32_t l = (uint32_t)n;
        memcpy(out.data(), &l, 4);
        memcpy(out.data() + 4, d, n);
        return out;
    };
}

static std::vector<uint8_t> unframe(const std::vector<uint8_t> &b) {
    std::vector<uint8_t> out;
    size_t o = 0;
    while (o < b.size()) {
        uint32_t l;
        memcpy(&l, b.data() + o, 4);
        out.insert(out.end(), b.begin() + o + 4, b.begin() + o + 4 + l);
        o += 4 + l;
    }
    return out;
}

static void order_preserved() {  // parallel.rs:314-328
    auto sink = std::make_unique<Sink>();
    auto buf = sink->buf;
    std::vector<uint8_t> data(64 * 1000 + 123);
    for (size_t i = 0; i < data.size(); i++) data[i] = (uint8_t)(i * 131 + 7);
    {
        ParallelChunkedWriter w(std::move(sink), 1000, 4, identity_framer(50));
        size_t o = 0, step = 1;
        while (o < data.size()) {
            size_t n = std::min(step, data.size() - o);
            w.write_all(data.data() + o, n);
            o += n;
            step = step * 3 % 997 + 1;
        }
        w.flush();  // does not cut a chunk
        w.finalize();
    }
    CHECK(unframe(*buf) == data);
    CHECK(buf->size() == data.size() + 4 * 65);
}

static void backpressure() {  // parallel.rs:330-343
    std::atomic<int> live{0}, peak{0};
    ChunkFn fn = [&](const uint8_t *d, size_t n) {
        int v = ++live;
        int p = peak.load();
        while (v > p && !peak.compare_exchange_weak(p, v)) {
        }
        std::this_thread::sleep_for(std::chrono::milliseconds(2));
        --live;
        return std::vector<uint8_t>(d, d + n);
    };
    auto sink = std::make_unique<Sink>();
    {
        ParallelChunkedWriter w(std::move(sink), 16, 2, fn);
        std::vector<uint8_t> data(16 * 40, 1);
        w.write_all(data.data(), data.size());
        w.finalize();
    }
    CHECK(peak.load() <= 2);
}

static void sticky_error() {  // parallel.rs:345-369
    std::atomic<int> calls{0};
    ChunkFn fn = [&](const uint8_t *d, size_t n) -> std::vector<uint8_t> {
        if (++calls == 3) throw Error(Error::Kind::Io, "chunk failed");
        return std::vector<uint8_t>(d, d + n);
    };
    auto sink = std::make_unique<Sink>();
    ParallelChunkedWriter w(std::move(sink), 10, 1, fn);
    std::vector<uint8_t> data(100, 2);
    bool threw = false;
    try {
        w.write_all(data.data(), data.size());
    } catch (const Error &) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        w.write_all(data.data(), 5);
    } catch (const Error &) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        w.finalize();
    } catch (const Error &) {
        threw = true;
    }
    CHECK(threw);
    threw = false;
    try {
        w.finalize();  // idempotent, still failing
    } catch (const Error &) {
        threw = true;
    }
    CHECK(threw);
}

static void lazy_backend_without_gpu() {
    LazyGpuBackend lazy;
    CHECK(lazy.supports(Algorithm::Gzip) == Capability::CompressOnly);
    CHECK(lazy.supports(Algorithm::Zstd) == Capability::None);
    CHECK(std::string(lazy.name()) == "gpu");
    if (!lazy.try_get()) {
        bool threw = false;
        try {
            lazy.compressor(Algorithm::Gzip, Level{});
        } catch (const Error &e) {
            threw = e.kind == Error::Kind::NoBackend;
        }
        CHECK(threw);
        CHECK(lazy.try_get() == nullptr);  // init attempted once, cached
    }
}

struct FileSink : Writer {
    FILE *f;
    explicit FileSink(FILE *f) : f(f) {}
    void write_all(const uint8_t *d, size_t n) override {
        if (fwrite(d, 1, n, f) != n) throw Error(Error::Kind::Io, "write failed");
    }
    void flush() override { fflush(f); }
};

int main(int argc, char **argv) {
    if (argc >= 2 && !strcmp(argv[1], "cpu")) {
        order_preserved();
        backpressure();
        sticky_error();
        lazy_backend_without_gpu();
        printf("cpu ok\n");
        return 0;
    }
    if (argc >= 3 && !strcmp(argv[1], "gpu")) {
        FILE *in = fopen(argv[2], "rb");
        CHECK(in);
        std::vector<uint8_t> data;
        uint8_t tmp[1 << 16];
        size_t n;
        while ((n = fread(tmp, 1, sizeof tmp, in)) > 0) data.insert(data.end(), tmp, tmp + n);
        fclose(in);
        LazyGpuBackend lazy;
        auto g = lazy.try_get();
        CHECK(g != nullptr);
        CHECK((*g)->chunk_size() == 131072 && (*g)->max_in_flight() == 16);
        auto comp = lazy.compressor(Algorithm::Gzip, Level{5});
        auto w = comp->wrap_writer(std::make_unique<FileSink>(stdout));
        // odd write sizes, like a tar writer
        size_t o = 0, step = 4099;
        while (o < data.size()) {
            size_t k = std::min(step, data.size() - o);
            w->write_all(data.data() + o, k);
            o += k;
            step = step * 7 % 65521 + 1;
        }
        static_cast<ParallelChunkedWriter *>(w.get())->finalize();
        return 0;
    }
    fprintf(stderr, "usage: test_backend cpu | gpu FILE\n");
    return 2;
}
