// gpz_synth.cpp -- seeded synthetic workloads (bench and test data, not part
// of the compress path).
//
// PRNG: SplitMix64 expands the seed into the state of xoshiro256**.
//   text    Zipf(s=1.1) over a 4,000-word vocabulary (word lengths U{2..10},
//           letters U{a..z}, de-duplicated), words separated by spaces, a
//           newline after every U{8..16} words.
//   binmix  1 MiB segments in seeded random order: 40 % uniform random bytes,
//           30 % 64-byte records (u32 LE counter +1 per record, u32 drawn from a
//           fixed table of 256 values, 56 bytes over a 16-symbol alphabet),
//           30 % text.
//   log     lines from 50 fixed templates of 60-120 ASCII bytes with 2-4 slots
//           each: monotonically increasing timestamp, uniform u32 decimal, or
//           64-bit hex ID.
//   sizes   log-uniform or log-normal chunk/request sizes.
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <set>
#include <string>
#include <vector>

namespace {

struct Rng {
    uint64_t s[4];
    explicit Rng(uint64_t seed) {
        uint64_t x = seed;
        for (auto &v : s) {
            x += 0x9E3779B97F4A7C15ull;
            uint64_t z = x;
            z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
            z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
            v = z ^ (z >> 31);
        }
    }
    static uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    uint64_t next() {
        uint64_t r = rotl(s[1] * 5, 7) * 9;
        uint64_t t = s[1] << 17;
        s[2] ^= s[0];
        s[3] ^= s[1];
        s[1] ^= s[2];
        s[0] ^= s[3];
        s[2] ^= t;
        s[3] = rotl(s[3], 45);
        return r;
    }
    uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    uint32_t range(uint32_t lo, uint32_t hi) { return lo + below(hi - lo + 1); }  // inclusive
    double unit() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
};

struct TextGen {
    std::vector<std::string> vocab;
    std::vector<double> cdf;
    Rng rng;
    uint32_t words_left = 0;
    explicit TextGen(uint64_t seed) : rng(seed) {
        Rng vr(seed ^ 0x5EEDF00Dull);
        std::set<std::string> seen;
        while (vocab.size() < 4000) {
            uint32_t len = vr.range(2, 10);
            std::string w;
            for (uint32_t i = 0; i < len; i++) w.push_back((char)('a' + vr.below(26)));
            if (seen.insert(w).second) vocab.push_back(w);
        }
        cdf.resize(vocab.size());
        double acc = 0;
        for (size_t r = 0; r < vocab.size(); r++) {
            acc += 1.0 / pow((double)(r + 1), 1.1);
            cdf[r] = acc;
        }
        for (auto &c : cdf) c /= acc;
        words_left = rng.range(8, 16);
    }
    const std::string &word() {
        double u = rng.unit();
        size_t r = std::lower_bound(cdf.begin(), cdf.end(), u) - cdf.begin();
        if (r >= vocab.size()) r = vocab.size() - 1;
        return vocab[r];
    }
    void fill(uint8_t *out, size_t n) {
        size_t o = 0;
        while (o < n) {
            const std::string &w = word();
            for (char c : w) {
                if (o >= n) return;
                out[o++] = (uint8_t)c;
            }
            if (o >= n) return;
            if (--words_left == 0) {
                out[o++] = '\n';
                words_left = rng.range(8, 16);
            } else {
                out[o++] = ' ';
            }
        }
    }
};

void fill_binmix(uint64_t seed, uint8_t *out, size_t n) {
    const size_t SEGB = 1u << 20;
    size_t nseg = (n + SEGB - 1) / SEGB;
    size_t n_rand = (size_t)llround(nseg * 0.4), n_rec = (size_t)llround(nseg * 0.3);
    if (n_rand + n_rec > nseg) n_rec = nseg - n_rand;
    std::vector<int> kind(nseg, 2);
    for (size_t i = 0; i < n_rand; i++) kind[i] = 0;
    for (size_t i = 0; i < n_rec; i++) kind[n_rand + i] = 1;
    Rng rng(seed);
    for (size_t i = nseg; i > 1; i--) std::swap(kind[i - 1], kind[rng.below((uint32_t)i)]);
    uint32_t table[256];
    for (auto &t : table) t = (uint32_t)rng.next();
    static const char alpha[17] = "0123456789abcdef";
    uint32_t counter = 0;
    TextGen text(seed * 31 + 7);
    for (size_t sgi = 0; sgi < nseg; sgi++) {
        uint8_t *p = out + sgi * SEGB;
        size_t len = std::min(SEGB, n - sgi * SEGB);
        if (kind[sgi] == 0) {
            size_t i = 0;
            for (; i + 8 <= len; i += 8) {
                uint64_t v = rng.next();
                memcpy(p + i, &v, 8);
            }
            for (; i < len; i++) p[i] = (uint8_t)rng.next();
        } else if (kind[sgi] == 1) {
            uint8_t rec[64];
            for (size_t i = 0; i < len; i += 64) {
                uint32_t v = table[rng.below(256)];
                for (int b = 0; b < 4; b++) rec[b] = (uint8_t)(counter >> (8 * b));
                for (int b = 0; b < 4; b++) rec[4 + b] = (uint8_t)(v >> (8 * b));
                for (int b = 8; b < 64; b++) rec[b] = (uint8_t)alpha[rng.below(16)];
                counter++;
                memcpy(p + i, rec, std::min<size_t>(64, len - i));
            }
        } else {
            text.fill(p, len);
        }
    }
}

struct LogGen {
    struct Piece {
        std::string lit;
        int slot;  // -1 literal, 0 timestamp, 1 decimal, 2 hex
    };
    std::vector<std::vector<Piece>> templates;
    Rng rng;
    uint64_t ms = 1764763200000ull;  // 2025-12-03T12:00:00Z
    explicit LogGen(uint64_t seed) : rng(seed) {
        Rng tr(seed ^ 0x10C10C10Cull);
        TextGen words(seed ^ 0xABCDEFull);
        static const char *levels[] = {"INFO", "WARN", "DEBUG", "ERROR", "TRACE"};
        static const char *keys[] = {"id=", "req=", "user=", "span=", "latency_us=", "bytes=", "node="};
        for (int t = 0; t < 50; t++) {
            uint32_t target = tr.range(60, 120);
            uint32_t nslots = tr.range(2, 4);
            std::vector<Piece> pcs;
            std::vector<int> kinds;
            for (uint32_t k = 0; k < nslots; k++) kinds.push_back((int)tr.below(3));
            // literal budget: target minus typical slot widths (24 ts, ~10 dec, 16 hex)
            int slot_w = 0;
            for (int k : kinds) slot_w += k == 0 ? 24 : (k == 1 ? 10 : 16);
            int lit_budget = std::max<int>(16, (int)target - slot_w);
            std::string head = std::string(levels[tr.below(5)]) + " ";
            std::vector<std::string> lits;
            for (uint32_t k = 0; k <= nslots; k++) {
                std::string s;
                int want = lit_budget / (int)(nslots + 1);
                while ((int)s.size() < want) {
                    s += words.word();
                    s.push_back(' ');
                }
                if (k < nslots) s += keys[tr.below(7)];
                lits.push_back(s);
            }
            lits[0] = head + lits[0];
            for (uint32_t k = 0; k < nslots; k++) {
                pcs.push_back({lits[k], -1});
                pcs.push_back({"", kinds[k]});
            }
            pcs.push_back({lits[nslots], -1});
            templates.push_back(pcs);
        }
    }
    void fill(uint8_t *out, size_t n) {
        size_t o = 0;
        char buf[64];
        while (o < n) {
            const auto &tp = templates[rng.below(50)];
            ms += rng.below(40);
            std::string line;
            for (const auto &pc : tp) {
                if (pc.slot < 0) {
                    line += pc.lit;
                } else if (pc.slot == 0) {
                    uint64_t s = ms / 1000;
                    unsigned msr = (unsigned)(ms % 1000);
                    unsigned hh = (unsigned)((s / 3600) % 24), mi = (unsigned)((s / 60) % 60), ss = (unsigned)(s % 60);
                    unsigned day = (unsigned)(s / 86400 % 28) + 1;
                    snprintf(buf, sizeof(buf), "2025-12-%02uT%02u:%02u:%02u.%03uZ", day, hh, mi, ss, msr);
                    line += buf;
                } else if (pc.slot == 1) {
                    snprintf(buf, sizeof(buf), "%u", (unsigned)(rng.next() >> 32));
                    line += buf;
                } else {
                    snprintf(buf, sizeof(buf), "%016llx", (unsigned long long)rng.next());
                    line += buf;
                }
            }
            line.push_back('\n');
            size_t take = std::min(line.size(), n - o);
            memcpy(out + o, line.data(), take);
            o += take;
        }
    }
};

}  // namespace

extern "C" {

// kind: 0 text, 1 binmix, 2 log
void gpz_synth_fill(int kind, uint64_t seed, uint8_t *out, size_t n) {
    if (!out || !n) return;
    if (kind == 0) {
        TextGen g(seed);
        g.fill(out, n);
    } else if (kind == 1) {
        fill_binmix(seed, out, n);
    } else {
        LogGen g(seed);
        g.fill(out, n);
    }
}

// Log-uniform sizes in [lo, hi].
void gpz_synth_loguniform_sizes(uint64_t seed, size_t count, uint64_t lo, uint64_t hi, uint64_t *out) {
    Rng r(seed);
    double a = log((double)lo), b = log((double)hi);
    for (size_t i = 0; i < count; i++) {
        uint64_t v = (uint64_t)llround(exp(a + (b - a) * r.unit()));
        out[i] = std::min(hi, std::max(lo, v));
    }
}

// Log-normal sizes (median, sigma) clamped to [lo, hi].
void gpz_synth_lognormal_sizes(uint64_t seed, size_t count, double median, double sigma, uint64_t lo,
                               uint64_t hi, uint64_t *out) {
    Rng r(seed);
    for (size_t i = 0; i < count; i++) {
        double u1 = std::max(r.unit(), 1e-300), u2 = r.unit();
        double z = sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
        uint64_t v = (uint64_t)llround(median * exp(sigma * z));
        out[i] = std::min(hi, std::max(lo, v));
    }
}

// Uniform doubles in [0, 1) (arrival processes, content choice).
void gpz_synth_uniform(uint64_t seed, size_t count, double *out) {
    Rng r(seed);
    for (size_t i = 0; i < count; i++) out[i] = r.unit();
}

}  // extern "C"
