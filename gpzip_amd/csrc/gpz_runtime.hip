// gpz_runtime.hip -- host runtime behind include/gpzip_hip.h.
//
// One process drives one GPU.  A context owns a pool of lanes; a lane is a
// HIP stream plus a workspace (device arrays laid out by capacity, so the
// same kernel arguments -- and hence the same captured HIP graph -- serve
// every batch that fits) and pinned staging buffers.  The plain per-chunk
// entry point takes a free lane per call, so up to `lanes` callers overlap on
// the device, as the reference's 16 in-flight ChunkFn calls share one wgpu
// queue (codec-gpu/src/gpu.rs:161-165, codec-gpu/src/gpu/batch.rs:36-41).
#include <hip/hip_runtime.h>

#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <condition_variable>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "gpz_internal.h"
#include "gpz_runtime.h"
#include "gpzip_hip.h"

namespace gpz {

hipError_t upload_constants();
void launch_seg_table(const Work &w, uint32_t grid, int variant, hipStream_t st);
void launch_match_find(const Work &w, uint32_t grid, int summary, int variant, hipStream_t st);
void launch_walk_summary(const Work &w, uint32_t grid, hipStream_t st);
void launch_chain(const Work &w, uint32_t grid, hipStream_t st);
void launch_walk_emit(const Work &w, uint32_t grid, hipStream_t st);
void launch_walk_member(const Work &w, uint32_t grid, int parts, hipStream_t st);
void launch_hist_tokens(const Work &w, uint32_t grid, hipStream_t st);
void launch_huffman(const Work &w, uint32_t grid, hipStream_t st);
void launch_scan(const Work &w, hipStream_t st);
void launch_emit(const Work &w, uint32_t grid, uint32_t slots, int variant, hipStream_t st);
void launch_plan_one(const Work &w, uint32_t len, uint32_t sub, hipStream_t st);
void launch_plan_batch(const Work &w, hipStream_t st);

static thread_local char g_err[512];

int fail(int code, const char *fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return code;
}
void clear_error() { g_err[0] = 0; }
const char *last_error() { return g_err; }

#define HIPCHK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) return fail(GPZ_E_DEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

static inline uint64_t round_up(uint64_t x, uint64_t a) { return (x + a - 1) / a * a; }

// ---------------------------------------------------------------------------
// buffers

int DevBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return GPZ_OK;
    if (p) hipFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(round_up(bytes, 1 << 20), 1 << 20);
    if (hipMalloc(&p, want) != hipSuccess) {
        p = nullptr;
        return fail(GPZ_E_OOM, "hipMalloc(%zu) failed", want);
    }
    cap = want;
    return GPZ_OK;
}
DevBuf::~DevBuf() {
    if (p) hipFree(p);
}

int HostBuf::ensure(size_t bytes) {
    if (bytes <= cap && p) return GPZ_OK;
    if (p) hipHostFree(p);
    p = nullptr;
    cap = 0;
    size_t want = std::max<size_t>(round_up(bytes, 1 << 20), 1 << 20);
    if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) {
        p = nullptr;
        return fail(GPZ_E_OOM, "hipHostMalloc(%zu) failed", want);
    }
    cap = want;
    return GPZ_OK;
}
HostBuf::~HostBuf() {
    if (p) hipHostFree(p);
}

// ---------------------------------------------------------------------------
// plan: chunk list -> members -> segments

int build_plan(Plan &pl, const uint64_t *lens, const uint64_t *dev_off, size_t C, uint32_t sub_chunk,
               int gzip, bool device_tables) {
    pl = Plan();
    if (sub_chunk == 0) return fail(GPZ_E_INVAL, "sub_chunk must be > 0");
    if (sub_chunk > (16u << 20)) return fail(GPZ_E_INVAL, "members above 16 MiB");
    pl.C = (uint32_t)C;
    pl.sub = sub_chunk;
    pl.device_tables = device_tables;
    pl.chunk_dev_off.resize(C);
    pl.chunk_len.assign(lens, lens + C);
    uint64_t in_off = 0, pos = 0, out = 0;
    if (device_tables) {
        // totals in closed form per chunk; k_plan_batch writes the tables
        const uint64_t spm = (sub_chunk + SEG - 1) >> SEG_LOG2, pstep = round_up(sub_chunk, POS_ALIGN);
        uint64_t M = 0, G = 0;
        for (size_t c = 0; c < C; c++) {
            const uint64_t len = lens[c];
            if (dev_off) {
                if (dev_off[c] % 16) return fail(GPZ_E_INVAL, "chunk %zu offset not 16-byte aligned", c);
                pl.chunk_dev_off[c] = dev_off[c];
            } else {
                pl.chunk_dev_off[c] = in_off;
                in_off += round_up(len, IN_ALIGN);
            }
            pl.in_end = std::max<uint64_t>(pl.in_end, pl.chunk_dev_off[c] + len);
            const uint64_t nm = len <= sub_chunk ? 1 : (len + sub_chunk - 1) / sub_chunk;
            const uint64_t nlast = len - (nm - 1) * sub_chunk;
            M += nm;
            G += (nm - 1) * spm + ((nlast + SEG - 1) >> SEG_LOG2);
            pos += (nm - 1) * pstep + round_up(nlast, POS_ALIGN);
            out += 2 * len + nm * (1024 + (gzip ? 18 : 0));
        }
        if (M >= (1ull << 32) || G >= (1ull << 32)) return fail(GPZ_E_INVAL, "batch too large");
        pl.M = (uint32_t)M;
        pl.G = (uint32_t)G;
        pl.P = pos + POS_ALIGN;
        pl.out_cap = round_up(out, 256) + 256;
        pl.in_bytes = round_up(pl.in_end, IN_ALIGN) + IN_ALIGN;
        pl.gzip = gzip;
        return GPZ_OK;
    }
    pl.chunk_mem0.resize(C + 1);
    for (size_t c = 0; c < C; c++) {
        uint64_t len = lens[c];
        if (dev_off) {
            if (dev_off[c] % 16) return fail(GPZ_E_INVAL, "chunk %zu offset not 16-byte aligned", c);
            pl.chunk_dev_off[c] = dev_off[c];
        } else {
            pl.chunk_dev_off[c] = in_off;
            in_off += round_up(len, IN_ALIGN);
        }
        pl.in_end = std::max<uint64_t>(pl.in_end, pl.chunk_dev_off[c] + len);
        pl.chunk_mem0[c] = (uint32_t)pl.mem.size();
        uint64_t nm = len <= sub_chunk ? 1 : (len + sub_chunk - 1) / sub_chunk;
        for (uint64_t i = 0; i < nm; i++) {
            MemberDesc d{};
            uint64_t off = i * sub_chunk;
            uint64_t n = std::min<uint64_t>(sub_chunk, len - off);
            if (len <= sub_chunk) n = len;
            if (n > (16u << 20)) return fail(GPZ_E_INVAL, "member of %llu bytes exceeds 16 MiB", (unsigned long long)n);
            d.in_off = pl.chunk_dev_off[c] + off;
            d.pos_base = pos;
            d.n = (uint32_t)n;
            d.nseg = (uint32_t)((n + SEG - 1) >> SEG_LOG2);
            d.seg_base = (uint32_t)pl.seg_mem.size();
            d.chunk = (uint32_t)c;
            for (uint32_t s = 0; s < d.nseg; s++) pl.seg_mem.push_back((uint32_t)pl.mem.size());
            pl.mem.push_back(d);
            pos += round_up(n, POS_ALIGN);
            out += 2 * n + 1024 + (gzip ? 18 : 0);
        }
    }
    pl.chunk_mem0[C] = (uint32_t)pl.mem.size();
    pl.M = (uint32_t)pl.mem.size();
    pl.G = (uint32_t)pl.seg_mem.size();
    pl.P = pos + POS_ALIGN;
    pl.out_cap = round_up(out, 256) + 256;
    pl.in_bytes = round_up(pl.in_end, IN_ALIGN) + IN_ALIGN;
    pl.gzip = gzip;
    return GPZ_OK;
}

// ---------------------------------------------------------------------------
// workspace laid out by capacity

static uint64_t carve(uint64_t &cur, uint64_t bytes) {
    uint64_t o = round_up(cur, 256);
    cur = o + round_up(bytes, 256);
    return o;
}

Layout layout_for(const Caps &c) {
    uint64_t cur = 0;
    Layout L;
    L.mem = carve(cur, c.M * sizeof(MemberDesc));
    L.seg_mem = carve(cur, c.G * 4);
    L.chunk_mem0 = carve(cur, (c.C + 1) * 4);
    L.counts = carve(cur, sizeof(Counts));
    L.cdesc = carve(cur, 16 + c.C * 16);
    L.desc_end = cur;
    L.tok = carve(cur, c.P * 4);
    L.segtab = carve(cur, (uint64_t)c.G * TAB_WORDS * 4);
    L.seg_ntab = carve(cur, c.G * 4);
    L.wsum = carve(cur, (uint64_t)c.G * SPAN * 4);
    L.seg_crc = carve(cur, c.G * 4);
    L.seg_entry = carve(cur, c.G * 4);
    L.seg_tokbase = carve(cur, c.G * 4);
    L.walked = carve(cur, c.P * 4);
    L.hist = carve(cur, (uint64_t)c.M * NSYM * 4);
    L.ntok = carve(cur, c.M * 4);
    L.codes = carve(cur, (uint64_t)c.M * NSYM * 4);
    L.prefix = carve(cur, (uint64_t)c.M * PREFIX_WORDS * 4);
    L.prefix_bits = carve(cur, c.M * 4);
    L.mem_bytes = carve(cur, c.M * 8);
    L.mem_crc = carve(cur, c.M * 4);
    L.mem_off = carve(cur, c.M * 8);
    L.chunk_off = carve(cur, c.C * 8);
    L.chunk_len = carve(cur, c.C * 8);
    L.total = carve(cur, 8);
    L.em_state = carve(cur, c.G * 8);
    L.em_join = carve(cur, c.G * 8);
    L.em_tick = carve(cur, c.M * 4);
    L.out = carve(cur, c.out);
    L.end = cur;
    return L;
}

// Kernel arguments for a workspace laid out by L from device address b.
Work work_at(const Layout &L, uint8_t *b) {
    Work w{};
    w.mem = (const MemberDesc *)(b + L.mem);
    w.seg_mem = (const uint32_t *)(b + L.seg_mem);
    w.chunk_mem0 = (const uint32_t *)(b + L.chunk_mem0);
    w.counts = (const Counts *)(b + L.counts);
    w.tok = (uint32_t *)(b + L.tok);
    w.segtab = (uint32_t *)(b + L.segtab);
    w.seg_ntab = (uint32_t *)(b + L.seg_ntab);
    w.wsum = (uint32_t *)(b + L.wsum);
    w.seg_crc = (uint32_t *)(b + L.seg_crc);
    w.seg_entry = (uint32_t *)(b + L.seg_entry);
    w.seg_tokbase = (uint32_t *)(b + L.seg_tokbase);
    w.walked = (uint32_t *)(b + L.walked);
    w.hist = (uint32_t *)(b + L.hist);
    w.ntok = (uint32_t *)(b + L.ntok);
    w.codes = (uint32_t *)(b + L.codes);
    w.prefix = (uint32_t *)(b + L.prefix);
    w.prefix_bits = (uint32_t *)(b + L.prefix_bits);
    w.mem_bytes = (uint64_t *)(b + L.mem_bytes);
    w.mem_crc = (uint32_t *)(b + L.mem_crc);
    w.mem_off = (uint64_t *)(b + L.mem_off);
    w.chunk_off = (uint64_t *)(b + L.chunk_off);
    w.chunk_len = (uint64_t *)(b + L.chunk_len);
    w.total = (uint64_t *)(b + L.total);
    w.em_state = (uint64_t *)(b + L.em_state);
    w.em_join = (uint64_t *)(b + L.em_join);
    w.em_tick = (uint32_t *)(b + L.em_tick);
    w.out = b + L.out;
    w.window = 32768;
    return w;
}

void Workspace::point(uint8_t *b) {
    work = work_at(lay, b);
    base = b;
}

int Workspace::reserve(const Caps &want) {
    Caps c = caps;
    bool grow = !base || want.P > c.P || want.G > c.G || want.M > c.M || want.C > c.C || want.out > c.out;
    if (!grow) return GPZ_OK;
    if (external) return fail(GPZ_E_INVAL, "step exceeds the fixed workspace of a serving lane");
    c.P = std::max(c.P, want.P);
    c.G = std::max(c.G, want.G);
    c.M = std::max(c.M, want.M);
    c.C = std::max(c.C, want.C);
    c.out = std::max(c.out, want.out);
    const Layout L = layout_for(c);
    int rc = ws.ensure(L.end);
    if (rc) return rc;
    rc = hdesc.ensure(L.desc_end);
    if (rc) return rc;
    caps = c;
    lay = L;
    generation++;
    point((uint8_t *)ws.p);
    return GPZ_OK;
}

int Workspace::bind(const Caps &c, void *mem, uint64_t bytes) {
    const Layout L = layout_for(c);
    if (L.end > bytes) return fail(GPZ_E_INVAL, "workspace needs %llu bytes", (unsigned long long)L.end);
    int rc = hdesc.ensure(L.desc_end);
    if (rc) return rc;
    caps = c;
    lay = L;
    external = true;
    generation++;
    point((uint8_t *)mem);
    return GPZ_OK;
}

// Copy the plan's descriptor tables into pinned staging and enqueue the upload.
int Workspace::upload_plan(const Plan &pl, hipStream_t st) {
    uint8_t *h = (uint8_t *)hdesc.p;
    if (pl.device_tables) {
        // the chunk descriptors only; k_plan_batch derives the tables
        uint64_t *d = (uint64_t *)(h + lay.cdesc);
        d[0] = pl.C;
        d[1] = 0;
        for (uint32_t c = 0; c < pl.C; c++) {
            d[2 + 2 * (uint64_t)c] = pl.chunk_dev_off[c];
            d[3 + 2 * (uint64_t)c] = pl.chunk_len[c];
        }
        HIPCHK(hipMemcpyAsync(base + lay.cdesc, d, 16 + (uint64_t)pl.C * 16, hipMemcpyHostToDevice, st));
        return GPZ_OK;
    }
    memcpy(h + lay.mem, pl.mem.data(), pl.mem.size() * sizeof(MemberDesc));
    memcpy(h + lay.seg_mem, pl.seg_mem.data(), pl.seg_mem.size() * 4);
    memcpy(h + lay.chunk_mem0, pl.chunk_mem0.data(), pl.chunk_mem0.size() * 4);
    Counts cn{pl.G, pl.M, pl.C, 0};
    memcpy(h + lay.counts, &cn, sizeof(cn));
    HIPCHK(hipMemcpyAsync(base, hdesc.p, lay.desc_end, hipMemcpyHostToDevice, st));
    return GPZ_OK;
}

// Built-in launch shapes (what the autotuner picked on the binmix corpus):
// the segment kernels with one workgroup per segment (the dispatcher balances
// better than a persistent loop, and no registers stay live across
// segments), the segment-table kernel persistent, member kernels one
// workgroup per member.  Kernels still loop over the device-side counts, so a
// graph captured for a capacity serves any smaller batch.
Grids grids_for(const Caps &c, int n_cu) {
    Grids g;
    g.k0 = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(c.G, (uint64_t)n_cu * 12));
    g.k1 = (uint32_t)std::max<uint64_t>(1, c.G);
    g.k3 = (uint32_t)std::max<uint64_t>(1, c.G);
    g.chain = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(c.M, (uint64_t)n_cu * 16));
    g.huff = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(c.M, (uint64_t)n_cu * 16));
    g.emit = (uint32_t)std::max<uint64_t>(1, c.M);
    g.eslots = (uint32_t)std::max<uint64_t>(1, c.G + c.M);
    // bit emission: few members (a plain chunk's 16) leave one workgroup per
    // member latency-bound, so every 4,096-token slot gets its own; large
    // batches keep one 256-thread workgroup per member
    g.emitv = c.M < 512 ? 2 : 1;
    // the per-member walk runs a member's blocks in sequence: it pays once a
    // batch has enough members to fill the chip, the per-segment walk below
    // that; each member is cut into four parts walked by four waves (faster
    // than one or two parts on 2,048 and on 21,444 members)
    g.walkv = c.M >= 512 ? 1 : 0;
    g.kmp = 4;
    g.km = (uint32_t)std::max<uint64_t>(1, (c.M * g.kmp + 3) / 4);
    return g;
}

void apply_tune(Grids &g, const gpz_tune &t, uint64_t G, uint64_t M) {
    if (t.k0_grid) g.k0 = t.k0_grid;
    if (t.k1_grid) g.k1 = t.k1_grid;
    if (t.k3_grid) g.k3 = t.k3_grid;
    if (t.emit_grid) g.emit = t.emit_grid;
    // segments per workgroup: the grid follows the batch
    if (t.k1_seg_per_wg) g.k1 = (uint32_t)((G + t.k1_seg_per_wg - 1) / t.k1_seg_per_wg);
    if (t.k3_seg_per_wg) g.k3 = (uint32_t)((G + t.k3_seg_per_wg - 1) / t.k3_seg_per_wg);
    if (t.k1_variant == 1 || t.k1_variant == 2) g.k1v = t.k1_variant - 1;
    if (t.emit_variant >= 1 && t.emit_variant <= 4) g.emitv = t.emit_variant - 1;
    if (t.k0_variant >= 1 && t.k0_variant <= 3) g.k0v = t.k0_variant == 2 ? 0 : (t.k0_variant == 1 ? 1 : 2);
    if (t.walk_variant == 1 || t.walk_variant == 2) g.walkv = t.walk_variant - 1;
    if (t.walk_parts == 1 || t.walk_parts == 2 || t.walk_parts == 4) {
        g.kmp = t.walk_parts;
        g.km = (uint32_t)std::max<uint64_t>(1, (M * g.kmp + 3) / 4);
    }
    // a record tuned on a larger batch must not launch workgroups with no
    // segment or member to take
    auto cap = [](uint32_t v, uint64_t n) { return (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(v, n)); };
    g.k0 = cap(g.k0, G);
    g.k1 = cap(g.k1, G);
    g.k3 = cap(g.k3, G);
    g.emit = cap(g.emit, M);
}

// Enqueue the kernels of one run.  Events (if given) bracket the match-find
// and emit kernels for timing.
void enqueue(const Work &w, const Grids &g, int mode, hipStream_t st, hipEvent_t *ev) {
    if (ev) hipEventRecord(ev[0], st);
    switch (mode) {
        case MODE_MATCH:
            launch_seg_table(w, g.k0, g.k0v, st);
            launch_match_find(w, g.k1, 0, g.k1v, st);
            break;
        case MODE_WALK:
            if (g.walkv) {
                launch_walk_member(w, g.km, (int)g.kmp, st);
            } else {
                launch_walk_summary(w, g.k1, st);
                launch_chain(w, g.chain, st);
                launch_walk_emit(w, g.k3, st);
            }
            break;
        case MODE_ENCODE:
            launch_hist_tokens(w, g.huff, st);
            launch_huffman(w, g.huff, st);
            launch_scan(w, st);
            if (ev) hipEventRecord(ev[3], st);
            launch_emit(w, g.emit, g.eslots, g.emitv, st);
            if (ev) hipEventRecord(ev[4], st);
            break;
        default:  // MODE_FULL
            if (w.cdesc) launch_plan_batch(w, st);  // tables from the chunk descriptors
            launch_seg_table(w, g.k0, g.k0v, st);
            if (ev) hipEventRecord(ev[5], st);
            if (g.walkv) {
                // per-member walk: no summaries, no chaining pass
                launch_match_find(w, g.k1, 0, g.k1v, st);
                if (ev) hipEventRecord(ev[1], st);
                launch_walk_member(w, g.km, (int)g.kmp, st);
            } else {
                if (w.strict == 2) {
                    // ratio mode: the walk rule reads two tokens ahead, so the
                    // summary runs once every token exists
                    launch_match_find(w, g.k1, 0, g.k1v, st);
                    if (ev) hipEventRecord(ev[1], st);
                    launch_walk_summary(w, g.k1, st);
                } else {
                    launch_match_find(w, g.k1, 1, g.k1v, st);  // with the walk summary
                    if (ev) hipEventRecord(ev[1], st);
                }
                launch_chain(w, g.chain, st);
                launch_walk_emit(w, g.k3, st);
            }
            launch_huffman(w, g.huff, st);
            launch_scan(w, st);
            if (ev) hipEventRecord(ev[3], st);
            launch_emit(w, g.emit, g.eslots, g.emitv, st);
            if (ev) hipEventRecord(ev[4], st);
            break;
    }
    if (ev) hipEventRecord(ev[2], st);
}

// ---------------------------------------------------------------------------
// lanes

Lane::~Lane() {
    if (graph) hipGraphExecDestroy(graph);
    for (auto &e : ev)
        if (e) hipEventDestroy(e);
    if (st) hipStreamDestroy(st);
}

int Lane::init(int dev) {
    HIPCHK(hipSetDevice(dev));
    HIPCHK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
    for (auto &e : ev) HIPCHK(hipEventCreate(&e));
    return GPZ_OK;
}

// ---------------------------------------------------------------------------
// Scratch of the allocation-free entry point (gpz_compress_chunk_noalloc):
// the workspace of one chunk of at most NOALLOC_MAX_BYTES plus its input and
// output staging, in a file-scope device array.  Every GPU the library runs
// on holds its own copy until the process ends.  In host code `scratch`
// names a host-side stand-in, not a device address: the device address of
// each GPU's copy comes from hipGetSymbolAddress with that GPU current (done
// once per device at gpz_init) and is the only one ever used.
constexpr uint64_t NOALLOC_MAX_BYTES = 2u << 20;
constexpr uint64_t NOALLOC_SCRATCH_BYTES = 40u << 20;
static __device__ __align__(256) uint8_t scratch[NOALLOC_SCRATCH_BYTES];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[MAX_DEVICES];

// Device address of `scratch` on device dev (current).  GPZ_OK or GPZ_E_DEVICE.
static int scratch_address(int dev, void **out) {
    if (dev < 0 || dev >= MAX_DEVICES) return fail(GPZ_E_NODEV, "device index %d out of range", dev);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(GPZ_E_DEVICE, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *out = g_scratch_ptr[dev];
    return GPZ_OK;
}

}  // namespace gpz

using namespace gpz;

struct gpz_ctx {
    int device = 0;
    int n_cu = 256;
    gpz_config cfg{};
    gpz_tune tune{};  // batched-form launch shapes (zero fields: built-in)
    char label[256] = {0};
    std::mutex mu;
    std::condition_variable cv;
    std::vector<std::unique_ptr<Lane>> lanes;
    std::vector<Lane *> free_lanes;
    void *scratch = nullptr;               // this device's copy of the no-alloc scratch
    hipStream_t noalloc_st = nullptr;      // stream of gpz_compress_chunk_noalloc
    std::mutex noalloc_mu;                 // one no-alloc call at a time per context

    ~gpz_ctx() {
        if (noalloc_st) (void)hipStreamDestroy(noalloc_st);
    }

    Lane *acquire() {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return !free_lanes.empty(); });
        Lane *l = free_lanes.back();
        free_lanes.pop_back();
        return l;
    }
    void release(Lane *l) {
        {
            std::lock_guard<std::mutex> lk(mu);
            free_lanes.push_back(l);
        }
        cv.notify_one();
    }
};

namespace gpz {
int ctx_device(gpz_ctx *c) { return c->device; }
int ctx_cus(gpz_ctx *c) { return c->n_cu; }
uint32_t ctx_sub_chunk(gpz_ctx *c) { return c->cfg.sub_chunk; }
uint32_t ctx_walk_rule(gpz_ctx *c) { return (c->cfg.flags & GPZ_CFG_RATIO) ? 2u : 0u; }
const gpz_tune &ctx_tune(gpz_ctx *c) { return c->tune; }

// Grids of the per-lane launches (single chunks, stage-level calls): the
// built-in shapes, with the walk of the context's tuning record when it names
// one.
static Grids lane_grids(gpz_ctx *ctx, const Caps &c) {
    Grids g = grids_for(c, ctx->n_cu);
    const uint32_t wv = ctx->tune.walk_variant, wp = ctx->tune.walk_parts;
    if (wv == 1 || wv == 2) g.walkv = wv - 1;
    if (wp == 1 || wp == 2 || wp == 4) {
        g.kmp = wp;
        g.km = (uint32_t)std::max<uint64_t>(1, (c.M * wp + 3) / 4);
    }
    return g;
}
Lane *ctx_acquire(gpz_ctx *c) { return c->acquire(); }
void ctx_release(gpz_ctx *c, Lane *l) { c->release(l); }

// Run one plan on a lane: reserve, upload descriptors, (optionally) upload the
// host input, enqueue kernels.  Input is already in lane->din when host_in is null.
int lane_run(gpz_ctx *ctx, Lane *ln, const Plan &pl, int mode, const uint8_t *host_in,
             const uint64_t *host_lens) {
    HIPCHK(hipSetDevice(ctx->device));
    Caps want{pl.P, pl.G, pl.M, pl.C, pl.out_cap};
    int rc = ln->wsp.reserve(want);
    if (rc) return rc;
    rc = ln->din.ensure(pl.in_bytes);
    if (rc) return rc;
    if (host_in && pl.C == 1) {
        // one chunk at offset 0: straight from the caller's memory (the runtime
        // stages pageable memory itself; a copy through our pinned buffer first
        // costs a host pass over the chunk)
        if (pl.in_end) HIPCHK(hipMemcpyAsync(ln->din.p, host_in, pl.in_end, hipMemcpyHostToDevice, ln->st));
    } else if (host_in) {
        rc = ln->hin.ensure(pl.in_bytes);
        if (rc) return rc;
        uint8_t *h = (uint8_t *)ln->hin.p;
        uint64_t src = 0;
        for (uint32_t c = 0; c < pl.C; c++) {
            memcpy(h + pl.chunk_dev_off[c], host_in + src, host_lens[c]);
            src += host_lens[c];
        }
        HIPCHK(hipMemcpyAsync(ln->din.p, ln->hin.p, pl.in_end, hipMemcpyHostToDevice, ln->st));
    }
    rc = ln->wsp.upload_plan(pl, ln->st);
    if (rc) return rc;
    Work w = ln->wsp.work;
    w.in = (const uint8_t *)ln->din.p;
    w.gzip = pl.gzip;
    w.strict = ln->strict;
    Grids g = lane_grids(ctx, ln->wsp.caps);
    enqueue(w, g, mode, ln->st, nullptr);
    HIPCHK(hipGetLastError());
    return GPZ_OK;
}

// Copy one chunk's result (after the run completed its kernels) into out.
int lane_fetch_chunk(Lane *ln, uint32_t c, uint8_t *out, size_t out_cap, size_t *out_len) {
    Workspace &W = ln->wsp;
    uint64_t meta[2];
    HIPCHK(hipMemcpyAsync(meta, W.work.chunk_off + c, 8, hipMemcpyDeviceToHost, ln->st));
    HIPCHK(hipMemcpyAsync(meta + 1, W.work.chunk_len + c, 8, hipMemcpyDeviceToHost, ln->st));
    HIPCHK(hipStreamSynchronize(ln->st));
    *out_len = meta[1];
    if (meta[1] > out_cap) return fail(GPZ_E_NOSPACE, "output needs %llu bytes", (unsigned long long)meta[1]);
    if (meta[1]) {
        // straight into the caller's buffer (the runtime stages pageable memory)
        HIPCHK(hipMemcpyAsync(out, W.work.out + meta[0], meta[1], hipMemcpyDeviceToHost, ln->st));
        HIPCHK(hipStreamSynchronize(ln->st));
    }
    return GPZ_OK;
}

}  // namespace gpz

// ---------------------------------------------------------------------------
// C ABI

extern "C" {

int gpz_set_tune(gpz_ctx *ctx, const gpz_tune *tune) {
    clear_error();
    if (!ctx) return fail(GPZ_E_INVAL, "null context");
    if (tune && (tune->k1_variant > 2 || tune->emit_variant > 4 || tune->k0_variant > 3 || tune->walk_variant > 2 ||
                 tune->walk_parts == 3 || tune->walk_parts > 4))
        return fail(GPZ_E_INVAL,
                    "k1_variant and walk_variant must be 0-2, k0_variant 0-3, emit_variant 0-4, "
                    "walk_parts 0, 1, 2 or 4");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (tune) ctx->tune = *tune;
    else ctx->tune = gpz_tune{};
    return GPZ_OK;
}

int gpz_get_tune(gpz_ctx *ctx, gpz_tune *tune) {
    clear_error();
    if (!ctx || !tune) return fail(GPZ_E_INVAL, "bad argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    *tune = ctx->tune;
    return GPZ_OK;
}

void gpz_config_default(gpz_config *cfg) {
    if (!cfg) return;
    memset(cfg, 0, sizeof(*cfg));
    cfg->device = -1;
    cfg->sub_chunk = 128 * 1024;
    cfg->window = 32 * 1024;
    cfg->lanes = 4;
    cfg->max_in_flight = 16;
    cfg->lane_bytes = 2u << 20;
}

const char *gpz_last_error(void) { return gpz::last_error(); }

const char *gpz_status_string(int s) {
    switch (s) {
        case GPZ_OK: return "ok";
        case GPZ_E_NODEV: return "no usable GPU";
        case GPZ_E_OOM: return "out of memory";
        case GPZ_E_INVAL: return "invalid argument";
        case GPZ_E_DEVICE: return "device error";
        case GPZ_E_NOSPACE: return "output buffer too small";
        case GPZ_E_BUSY: return "admission refused";
        case GPZ_E_AGAIN: return "nothing ready";
        default: return "unknown status";
    }
}

int gpz_device_count(int *count) {
    if (!count) return fail(GPZ_E_INVAL, "count is null");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
    *count = n;
    return GPZ_OK;
}

int gpz_init(const gpz_config *cfg_in, gpz_ctx **out) {
    clear_error();
    if (!out) return fail(GPZ_E_INVAL, "out is null");
    *out = nullptr;
    gpz_config cfg;
    if (cfg_in)
        cfg = *cfg_in;
    else
        gpz_config_default(&cfg);
    if (cfg.sub_chunk == 0 || cfg.sub_chunk > (16u << 20)) return fail(GPZ_E_INVAL, "sub_chunk out of range");
    if (cfg.window != 32768) return fail(GPZ_E_INVAL, "only the 32 KiB DEFLATE window is supported");
    if (cfg.flags & ~(uint32_t)GPZ_CFG_RATIO) return fail(GPZ_E_INVAL, "unknown gpz_config flag");
    if (cfg.lanes == 0) cfg.lanes = 1;
    if (cfg.lanes > 32) cfg.lanes = 32;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n == 0) return fail(GPZ_E_NODEV, "no HIP device");
    int dev = cfg.device;
    if (dev < 0) {
        // GPZIP_GPU index semantics of the reference (codec-gpu/src/gpu/context.rs:119-155);
        // "high"/"low" pick device 0 (one GPU per process here).
        const char *e = getenv("GPZIP_GPU");
        dev = 0;
        if (e && *e && strcmp(e, "high") && strcmp(e, "low")) {
            char *end = nullptr;
            long v = strtol(e, &end, 10);
            if (end && *end == 0 && v >= 0) dev = (int)v;
        }
    }
    if (dev >= n) return fail(GPZ_E_NODEV, "GPZIP_GPU=%d but only %d device(s) available", dev, n);
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return fail(GPZ_E_NODEV, "cannot query device %d", dev);
    if (strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return fail(GPZ_E_NODEV, "device %d is %s; this library is built for gfx950", dev, prop.gcnArchName);
    if (hipSetDevice(dev) != hipSuccess) return fail(GPZ_E_NODEV, "hipSetDevice(%d) failed", dev);
    if (gpz::upload_constants() != hipSuccess) return fail(GPZ_E_DEVICE, "constant upload failed");
    auto ctx = std::make_unique<gpz_ctx>();
    ctx->device = dev;
    ctx->n_cu = prop.multiProcessorCount;
    ctx->cfg = cfg;
    {
        int rc = scratch_address(dev, &ctx->scratch);
        if (rc) return rc;
        if (hipStreamCreateWithFlags(&ctx->noalloc_st, hipStreamNonBlocking) != hipSuccess)
            return fail(GPZ_E_DEVICE, "stream creation failed");
    }
    snprintf(ctx->label, sizeof(ctx->label), "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    for (uint32_t i = 0; i < cfg.lanes; i++) {
        auto ln = std::make_unique<Lane>();
        int rc = ln->init(dev);
        if (rc) return rc;
        ctx->free_lanes.push_back(ln.get());
        ctx->lanes.push_back(std::move(ln));
    }
    *out = ctx.release();
    return GPZ_OK;
}

void gpz_destroy(gpz_ctx *ctx) {
    if (!ctx) return;
    hipSetDevice(ctx->device);
    hipDeviceSynchronize();
    delete ctx;
}

int gpz_device_label(gpz_ctx *ctx, char *buf, size_t cap) {
    if (!ctx || !buf || !cap) return fail(GPZ_E_INVAL, "bad argument");
    snprintf(buf, cap, "%s", ctx->label);
    return GPZ_OK;
}

size_t gpz_compress_bound(const gpz_ctx *ctx, size_t len) {
    uint64_t sub = ctx ? ctx->cfg.sub_chunk : 131072;
    uint64_t members = len <= sub ? 1 : (len + sub - 1) / sub;
    return 2 * len + members * (1024 + 18);
}

int gpz_compress_chunk(gpz_ctx *ctx, const uint8_t *in, size_t len, uint8_t *out, size_t out_cap,
                       size_t *out_len) {
    clear_error();
    if (!ctx || (!in && len) || !out_len) return fail(GPZ_E_INVAL, "bad argument");
    Plan pl;
    uint64_t l = len;
    int rc = build_plan(pl, &l, nullptr, 1, ctx->cfg.sub_chunk, 1);
    if (rc) return rc;
    Lane *ln = ctx->acquire();
    ln->strict = ctx_walk_rule(ctx);
    rc = lane_run(ctx, ln, pl, MODE_FULL, in, &l);
    if (!rc) rc = lane_fetch_chunk(ln, 0, out, out_cap, out_len);
    if (rc) hipStreamSynchronize(ln->st);
    ln->strict = 0;
    ctx->release(ln);
    return rc;
}

int gpz_compress_chunk_noalloc(gpz_ctx *ctx, const uint8_t *in, size_t len, uint8_t *out, size_t out_cap,
                               size_t *out_len) {
    clear_error();
    if (!ctx || (!in && len) || !out_len) return fail(GPZ_E_INVAL, "bad argument");
    if (len > NOALLOC_MAX_BYTES)
        return fail(GPZ_E_INVAL, "chunk of %zu bytes exceeds the no-alloc limit of %llu bytes", len,
                    (unsigned long long)NOALLOC_MAX_BYTES);
    // the plan in closed form (k_plan_one writes the same tables on the device)
    const uint64_t sub = ctx->cfg.sub_chunk;
    const uint64_t M = len <= sub ? 1 : (len + sub - 1) / sub;
    const uint64_t nlast = len <= sub ? len : len - (M - 1) * sub;
    const uint64_t spm = (sub + SEG - 1) >> SEG_LOG2, pstep = round_up(sub, POS_ALIGN);
    Caps c;
    c.M = M;
    c.G = (M - 1) * spm + ((nlast + SEG - 1) >> SEG_LOG2);
    c.P = (M - 1) * pstep + round_up(nlast, POS_ALIGN) + POS_ALIGN;
    c.C = 1;
    c.out = round_up(2 * (uint64_t)len + M * (1024 + 18), 256) + 256;
    const Layout L = layout_for(c);
    const uint64_t in_at = round_up(L.end, 256), in_bytes = round_up(len, 256) + 256;
    if (in_at + in_bytes > NOALLOC_SCRATCH_BYTES)
        return fail(GPZ_E_INVAL, "chunk of %zu bytes in %llu-byte members needs %llu bytes of scratch (%llu)", len,
                    (unsigned long long)sub, (unsigned long long)(in_at + in_bytes),
                    (unsigned long long)NOALLOC_SCRATCH_BYTES);
    std::lock_guard<std::mutex> lk(ctx->noalloc_mu);
    HIPCHK(hipSetDevice(ctx->device));
    hipStream_t st = ctx->noalloc_st;
    uint8_t *base = (uint8_t *)ctx->scratch;
    Work w = work_at(L, base);
    w.in = base + in_at;
    w.gzip = 1;
    w.strict = ctx_walk_rule(ctx);
    // no array needs to start at zero: every kernel writes what it later
    // reads (the per-member walk stores the histograms whole, k_chain clears
    // them for the per-segment walk), as on reused lane workspaces
    if (len) HIPCHK(hipMemcpyAsync(base + in_at, in, len, hipMemcpyDefault, st));
    launch_plan_one(w, (uint32_t)len, (uint32_t)sub, st);
    enqueue(w, lane_grids(ctx, c), MODE_FULL, st, nullptr);
    HIPCHK(hipGetLastError());
    uint64_t meta[2] = {0, 0};
    HIPCHK(hipMemcpyAsync(&meta[0], w.chunk_off, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&meta[1], w.chunk_len, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    *out_len = meta[1];
    if (meta[1] > out_cap) return fail(GPZ_E_NOSPACE, "output needs %llu bytes", (unsigned long long)meta[1]);
    if (meta[1]) {
        HIPCHK(hipMemcpyAsync(out, w.out + meta[0], meta[1], hipMemcpyDefault, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    return GPZ_OK;
}

int gpz_gzip_member(gpz_ctx *ctx, const uint8_t *in, size_t len, uint8_t *out, size_t out_cap,
                    size_t *out_len) {
    clear_error();
    if (!ctx || (!in && len) || !out_len) return fail(GPZ_E_INVAL, "bad argument");
    if (len > (16u << 20)) return fail(GPZ_E_INVAL, "member larger than 16 MiB");
    Plan pl;
    uint64_t l = len;
    int rc = build_plan(pl, &l, nullptr, 1, (uint32_t)std::max<size_t>(len, 1), 1);
    if (rc) return rc;
    Lane *ln = ctx->acquire();
    ln->strict = ctx_walk_rule(ctx);
    rc = lane_run(ctx, ln, pl, MODE_FULL, in, &l);
    if (!rc) rc = lane_fetch_chunk(ln, 0, out, out_cap, out_len);
    if (rc) hipStreamSynchronize(ln->st);
    ln->strict = 0;
    ctx->release(ln);
    return rc;
}

int gpz_match_find(gpz_ctx *ctx, const uint8_t *in, size_t len, uint32_t *tokens) {
    clear_error();
    if (!ctx || (!in && len) || (!tokens && len)) return fail(GPZ_E_INVAL, "bad argument");
    if (len == 0) return GPZ_OK;
    if (len > (16u << 20)) return fail(GPZ_E_INVAL, "member larger than 16 MiB");
    Plan pl;
    uint64_t l = len;
    int rc = build_plan(pl, &l, nullptr, 1, (uint32_t)len, 1);
    if (rc) return rc;
    Lane *ln = ctx->acquire();
    ln->strict = 0;
    rc = lane_run(ctx, ln, pl, MODE_MATCH, in, &l);
    if (!rc) {
        int r2 = ln->hout.ensure(len * 4);
        if (r2) {
            rc = r2;
        } else if (hipMemcpyAsync(ln->hout.p, ln->wsp.work.tok, len * 4, hipMemcpyDeviceToHost, ln->st) != hipSuccess ||
                   hipStreamSynchronize(ln->st) != hipSuccess) {
            rc = fail(GPZ_E_DEVICE, "token read-back failed");
        } else {
            memcpy(tokens, ln->hout.p, len * 4);
        }
    }
    if (rc) hipStreamSynchronize(ln->st);
    ctx->release(ln);
    return rc;
}

int gpz_walk(gpz_ctx *ctx, const uint8_t *in, const uint32_t *per_pos, size_t len, int strict,
             uint32_t *walked, size_t *n_walked) {
    clear_error();
    if (!ctx || !n_walked || strict < 0 || strict > 2 || (len && (!per_pos || !walked || (strict != 1 && !in))))
        return fail(GPZ_E_INVAL, "bad argument");
    *n_walked = 0;
    if (len == 0) return GPZ_OK;
    if (len > (16u << 20)) return fail(GPZ_E_INVAL, "more than 16 Mi positions");
    for (size_t i = 0; i < len; i++)
        if ((per_pos[i] >> 16) > 258) return fail(GPZ_E_INVAL, "token %zu: length %u > 258", i, per_pos[i] >> 16);
    Plan pl;
    uint64_t l = len;
    int rc = build_plan(pl, &l, nullptr, 1, (uint32_t)len, 1);
    if (rc) return rc;
    Lane *ln = ctx->acquire();
    ln->strict = (uint32_t)strict;
    do {
        Caps want{pl.P, pl.G, pl.M, pl.C, pl.out_cap};
        if ((rc = ln->wsp.reserve(want))) break;
        if ((rc = ln->din.ensure(pl.in_bytes))) break;
        if ((rc = ln->hin.ensure(std::max<uint64_t>(pl.in_bytes, len * 4)))) break;
        if (hipSetDevice(ctx->device) != hipSuccess) {
            rc = fail(GPZ_E_DEVICE, "hipSetDevice failed");
            break;
        }
        // tokens -> w.tok, input bytes -> din
        memcpy(ln->hin.p, per_pos, len * 4);
        if (hipMemcpyAsync(ln->wsp.work.tok, ln->hin.p, len * 4, hipMemcpyHostToDevice, ln->st) != hipSuccess ||
            hipStreamSynchronize(ln->st) != hipSuccess) {
            rc = fail(GPZ_E_DEVICE, "token upload failed");
            break;
        }
        if (in) {
            memcpy(ln->hin.p, in, len);
            if (hipMemcpyAsync(ln->din.p, ln->hin.p, len, hipMemcpyHostToDevice, ln->st) != hipSuccess) {
                rc = fail(GPZ_E_DEVICE, "input upload failed");
                break;
            }
        }
        if ((rc = ln->wsp.upload_plan(pl, ln->st))) break;
        Work w = ln->wsp.work;
        w.in = (const uint8_t *)ln->din.p;
        w.strict = ln->strict;
        w.gzip = 1;
        enqueue(w, lane_grids(ctx, ln->wsp.caps), MODE_WALK, ln->st, nullptr);
        uint32_t nt = 0;
        if (hipMemcpyAsync(&nt, w.ntok, 4, hipMemcpyDeviceToHost, ln->st) != hipSuccess ||
            hipStreamSynchronize(ln->st) != hipSuccess) {
            rc = fail(GPZ_E_DEVICE, "walk failed");
            break;
        }
        if ((rc = ln->hout.ensure((uint64_t)nt * 4 + 4))) break;
        if (nt && (hipMemcpyAsync(ln->hout.p, w.walked, (uint64_t)nt * 4, hipMemcpyDeviceToHost, ln->st) != hipSuccess ||
                   hipStreamSynchronize(ln->st) != hipSuccess)) {
            rc = fail(GPZ_E_DEVICE, "walk read-back failed");
            break;
        }
        memcpy(walked, ln->hout.p, (uint64_t)nt * 4);
        *n_walked = nt;
    } while (0);
    if (rc) hipStreamSynchronize(ln->st);
    ln->strict = 0;
    ctx->release(ln);
    return rc;
}

int gpz_encode_block(gpz_ctx *ctx, const uint32_t *walked, size_t n, uint8_t *out, size_t out_cap,
                     size_t *out_len) {
    clear_error();
    if (!ctx || !out_len || (n && !walked)) return fail(GPZ_E_INVAL, "bad argument");
    if (n > (16u << 20)) return fail(GPZ_E_INVAL, "more than 16 Mi tokens");
    for (size_t i = 0; i < n; i++) {
        uint32_t len = walked[i] >> 16, v = walked[i] & 0xFFFFu;
        bool ok = len == 0 ? v <= 255 : (len >= 3 && len <= 258 && v >= 1 && v <= 32768);
        if (!ok) return fail(GPZ_E_INVAL, "token %zu outside the DEFLATE alphabet", i);
    }
    Plan pl;
    uint64_t l = n;
    int rc = build_plan(pl, &l, nullptr, 1, (uint32_t)std::max<size_t>(n, 1), 0);
    if (rc) return rc;
    // build_plan bounds byte inputs (2 B per byte); a token costs up to 48 bits
    // (15-bit length code + 5 extra + 15-bit distance code + 13 extra), and the
    // block header at most PREFIX_WORDS words
    pl.out_cap = round_up(6 * (uint64_t)n + PREFIX_WORDS * 4 + 64, 256) + 256;
    Lane *ln = ctx->acquire();
    do {
        Caps want{pl.P, pl.G, pl.M, pl.C, pl.out_cap};
        if ((rc = ln->wsp.reserve(want))) break;
        if ((rc = ln->din.ensure(pl.in_bytes))) break;
        if ((rc = ln->hin.ensure(n * 4 + 64))) break;
        if (hipSetDevice(ctx->device) != hipSuccess) {
            rc = fail(GPZ_E_DEVICE, "hipSetDevice failed");
            break;
        }
        memcpy(ln->hin.p, walked, n * 4);
        uint32_t nt = (uint32_t)n;
        Work w = ln->wsp.work;
        if ((n && hipMemcpyAsync(w.walked, ln->hin.p, n * 4, hipMemcpyHostToDevice, ln->st) != hipSuccess) ||
            hipMemcpyAsync(w.ntok, &nt, 4, hipMemcpyHostToDevice, ln->st) != hipSuccess ||
            hipStreamSynchronize(ln->st) != hipSuccess) {
            rc = fail(GPZ_E_DEVICE, "token upload failed");
            break;
        }
        if ((rc = ln->wsp.upload_plan(pl, ln->st))) break;
        w.in = (const uint8_t *)ln->din.p;
        w.gzip = 0;
        enqueue(w, lane_grids(ctx, ln->wsp.caps), MODE_ENCODE, ln->st, nullptr);
        rc = lane_fetch_chunk(ln, 0, out, out_cap, out_len);
    } while (0);
    if (rc) hipStreamSynchronize(ln->st);
    ctx->release(ln);
    return rc;
}

}  // extern "C"
