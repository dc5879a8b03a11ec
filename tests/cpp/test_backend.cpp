// Tests of the C++ backend mirror (include/gpzip_backend.hpp), modelled on the
// reference's scheduler tests (codec-cpu/src/parallel.rs:231-424).
//   test_backend cpu        -- ParallelChunkedWriter with an identity framer, no GPU
//   test_backend gpu FILE   -- GpuBackend compressor over FILE, output to stdout,
//                              "seconds S" to stderr
//   test_backend hybrid FILE GPU_WORKERS WARMUP [GPU_WAIT_MS]
//                           -- HybridBackend (cli/src/hybrid.rs) over FILE,
//                              stream to stdout, "gpu_chunks N cpu_chunks M
//                              seconds S cold_seconds C" to stderr (two
//                              passes: C includes first-use allocations)
//   test_backend decode FILE -- GzipDecompressor over FILE, output to stdout
//   test_backend tar OUT INPUT...    -- TarBuilder stream (uncompressed) to OUT
//   test_backend targz OUT CHUNK BATCH INPUT...
//                           -- pack_tar_gz on the batched GPU form; prints
//                              "batches B chunks C" to stderr
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
#include "gpzip_hip.h"

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

struct VecReader : Reader {
    std::vector<uint8_t> b;
    size_t pos = 0;
    size_t read(uint8_t *buf, size_t cap) override {
        size_t n = std::min(cap, b.size() - pos);
        memcpy(buf, b.data() + pos, n);
        pos += n;
        return n;
    }
};

static std::vector<uint8_t> read_all(Reader &r) {
    std::vector<uint8_t> out;
    uint8_t tmp[4096];
    size_t n;
    while ((n = r.read(tmp, sizeof tmp)) > 0) out.insert(out.end(), tmp, tmp + n);
    return out;
}

static std::vector<uint8_t> pattern(size_t n, uint32_t seed) {
    std::vector<uint8_t> d(n);
    uint32_t x = seed;
    for (size_t i = 0; i < n; i++) {
        x = x * 1664525u + 1013904223u;
        d[i] = (i / 97) % 3 ? (uint8_t)("abcdefgh"[(x >> 24) & 7]) : (uint8_t)(x >> 24);
    }
    return d;
}

// CpuBackend compress (codec-cpu/src/compress.rs) -> GzipDecompressor
// (codec-cpu/src/decompress.rs): the round trip, one member per 2 MiB chunk.
static void cpu_codec_round_trip() {
    CpuBackend cpu(1 << 20, 3);
    CHECK(cpu.supports(Algorithm::Gzip) == Capability::Both);
    const std::vector<uint8_t> data = pattern((5 << 20) + 777, 7);
    auto sink = std::make_unique<Sink>();
    auto buf = sink->buf;
    {
        auto w = cpu.compressor(Algorithm::Gzip, Level{5})->wrap_writer(std::move(sink));
        w->write_all(data.data(), data.size());
    }
    CHECK(find_member_starts(buf->data(), buf->size()).size() == 6);
    auto rd = std::make_unique<VecReader>();
    rd->b = *buf;
    auto r = cpu.decompressor(Algorithm::Gzip)->wrap_reader(std::move(rd));
    CHECK(read_all(*r) == data);
    CHECK(parallel_decompress(buf->data(), buf->size(), 1) == data);
    bool threw = false;
    try {
        LazyGpuBackend().decompressor(Algorithm::Gzip);
    } catch (const Error &e) {
        threw = e.kind == Error::Kind::DecompressionUnsupported;
    }
    CHECK(threw);
}

// `--backend auto` on a host without a GPU: every chunk on the CPU path.
static void hybrid_without_gpu() {
    auto stats = std::make_shared<HybridStats>();
    HybridTuning t = HybridTuning::retuned();
    t.stats = stats;
    HybridBackend hy(std::make_shared<CpuBackend>(1 << 20, 2), std::make_shared<LazyGpuBackend>(), t);
    CHECK(std::string(hy.name()) == "hybrid");
    const std::vector<uint8_t> data = pattern((7 << 20) + 5, 9);
    auto sink = std::make_unique<Sink>();
    auto buf = sink->buf;
    {
        auto w = hy.compressor(Algorithm::Gzip, Level{5})->wrap_writer(std::move(sink));
        w->write_all(data.data(), data.size());
    }
    CHECK(parallel_decompress(buf->data(), buf->size()) == data);
    if (!LazyGpuBackend().try_get()) CHECK(stats->gpu_chunks == 0 && stats->cpu_chunks == 8);
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
        cpu_codec_round_trip();
        hybrid_without_gpu();
        printf("cpu ok\n");
        return 0;
    }
    if (argc >= 5 && !strcmp(argv[1], "hybrid")) {
        FILE *in = fopen(argv[2], "rb");
        CHECK(in);
        std::vector<uint8_t> data;
        uint8_t tmp[1 << 16];
        size_t n;
        while ((n = fread(tmp, 1, sizeof tmp, in)) > 0) data.insert(data.end(), tmp, tmp + n);
        fclose(in);
        auto stats = std::make_shared<HybridStats>();
        HybridTuning t;
        t.gpu_workers = (size_t)atoi(argv[3]);
        t.warmup_chunks = (size_t)atoi(argv[4]);
        t.gpu_wait_ms = argc >= 6 ? (uint32_t)atoi(argv[5]) : 0;
        t.stats = stats;
        auto lazy = std::make_shared<LazyGpuBackend>((uint32_t)std::max<size_t>(4, t.gpu_workers));
        CHECK(lazy->try_get() != nullptr);  // initialised before the clock starts
        HybridBackend hy(std::make_shared<CpuBackend>(), lazy, t);
        // two passes: the first pays the lanes' first-use allocations (device
        // workspace, pinned staging); the second is the steady state
        double sec[2];
        std::shared_ptr<std::vector<uint8_t>> buf;
        for (int pass = 0; pass < 2; pass++) {
            stats->gpu_chunks = 0;
            stats->cpu_chunks = 0;
            auto sink = std::make_unique<Sink>();
            buf = sink->buf;
            buf->reserve(data.size() + (data.size() >> 4) + (1 << 20));
            auto t0 = std::chrono::steady_clock::now();
            {
                auto w = hy.compressor(Algorithm::Gzip, Level{5})->wrap_writer(std::move(sink));
                w->write_all(data.data(), data.size());
            }
            sec[pass] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        }
        fwrite(buf->data(), 1, buf->size(), stdout);
        fprintf(stderr, "gpu_chunks %llu cpu_chunks %llu seconds %.6f cold_seconds %.6f\n",
                (unsigned long long)stats->gpu_chunks.load(), (unsigned long long)stats->cpu_chunks.load(), sec[1], sec[0]);
        return 0;
    }
    if (argc >= 4 && !strcmp(argv[1], "tar")) {
        FILE *f = fopen(argv[2], "wb");
        CHECK(f);
        {
            TarBuilder tb(std::make_unique<FileSink>(f));
            for (int i = 3; i < argc; i++) {
                std::string in = argv[i];
                while (in.size() > 1 && in.back() == '/') in.pop_back();
                const size_t k = in.rfind('/');
                tb.append_path(in, k == std::string::npos ? in : in.substr(k + 1));
            }
            tb.finish();
        }
        fclose(f);
        return 0;
    }
    if (argc >= 6 && !strcmp(argv[1], "targz")) {
        gpz_config cfg;
        gpz_config_default(&cfg);
        gpz_ctx *ctx = nullptr;
        CHECK(gpz_init(&cfg, &ctx) == GPZ_OK);
        std::vector<std::string> inputs(argv + 5, argv + argc);
        {
            // the same composition as pack_tar_gz, keeping the writer's counters
            FILE *f = fopen(argv[2], "wb");
            CHECK(f);
            auto gz = std::make_unique<BatchedChunkedWriter>(std::make_unique<FileSink>(f), ctx,
                                                             (size_t)atol(argv[3]), (size_t)atol(argv[4]));
            BatchedChunkedWriter *gzp = gz.get();
            TarBuilder tb(std::move(gz));
            for (const std::string &in0 : inputs) {
                std::string in = in0;
                while (in.size() > 1 && in.back() == '/') in.pop_back();
                const size_t k = in.rfind('/');
                tb.append_path(in, k == std::string::npos ? in : in.substr(k + 1));
            }
            tb.finish();
            gzp->finalize();
            fprintf(stderr, "batches %zu chunks %zu\n", gzp->batches_submitted(), gzp->chunks_submitted());
            fclose(f);
        }
        pack_tar_gz(std::string(argv[2]) + ".2", inputs, ctx, (size_t)atol(argv[3]), (size_t)atol(argv[4]));
        gpz_destroy(ctx);
        return 0;
    }
    if (argc >= 3 && !strcmp(argv[1], "decode")) {
        FILE *in = fopen(argv[2], "rb");
        CHECK(in);
        auto rd = std::make_unique<VecReader>();
        uint8_t tmp[1 << 16];
        size_t n;
        while ((n = fread(tmp, 1, sizeof tmp, in)) > 0) rd->b.insert(rd->b.end(), tmp, tmp + n);
        fclose(in);
        auto r = CpuBackend().decompressor(Algorithm::Gzip)->wrap_reader(std::move(rd));
        std::vector<uint8_t> out = read_all(*r);
        fwrite(out.data(), 1, out.size(), stdout);
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
        // two passes (the first pays the lanes' first-use allocations); the
        // second, timed, goes to stdout
        double sec[2];
        for (int pass = 0; pass < 2; pass++) {
            auto sink = std::make_unique<Sink>();
            auto buf = sink->buf;
            buf->reserve(data.size() + (data.size() >> 4) + (1 << 20));
            auto t0 = std::chrono::steady_clock::now();
            auto w = comp->wrap_writer(std::move(sink));
            // odd write sizes, like a tar writer
            size_t o = 0, step = 4099;
            while (o < data.size()) {
                size_t k = std::min(step, data.size() - o);
                w->write_all(data.data() + o, k);
                o += k;
                step = step * 7 % 65521 + 1;
            }
            static_cast<ParallelChunkedWriter *>(w.get())->finalize();
            sec[pass] = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
            if (pass == 1) fwrite(buf->data(), 1, buf->size(), stdout);
        }
        fprintf(stderr, "seconds %.6f cold_seconds %.6f\n", sec[1], sec[0]);
        return 0;
    }
    fprintf(stderr, "usage: test_backend cpu | gpu FILE\n");
    return 2;
}
