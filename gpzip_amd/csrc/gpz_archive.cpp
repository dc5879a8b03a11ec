// gpz_archive.cpp -- `gpzip a` for .tar.gz on the batched GPU form (host code).
//
//   TarBuilder            the tar stream pack_tar writes (core/src/archive/
//                         pack.rs:103-176 over the tar crate's Builder): GNU
//                         headers, directories before their entries in name
//                         order, symlinks stored as links (not followed), two
//                         zero blocks at the end
//   BatchedChunkedWriter  ParallelChunkedWriter's contract (in-order members,
//                         one member group per chunk, flush does not cut a
//                         chunk, sticky first error) with the chunks of a
//                         stretch of the stream submitted as one device batch
//                         (gpz_batch_submit: one graph replay over a table of
//                         (offset, len) descriptors) instead of one pipeline
//                         per chunk (codec-cpu/src/parallel.rs:91-129)
//   pack_tar_gz           the two composed: the members equal the per-chunk
//                         GPU writer's for the same chunk size
#include <dirent.h>
#include <stdio.h>
#include <string.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <deque>
#include <string>
#include <vector>

#include "gpzip_backend.hpp"
#include "gpzip_hip.h"

namespace gpzip {

// ---------------------------------------------------------------------------
// BatchedChunkedWriter

BatchedChunkedWriter::BatchedChunkedWriter(std::unique_ptr<Writer> out, gpz_ctx *ctx, size_t chunk_size,
                                           size_t batch_chunks, size_t max_batches)
    : out_(std::move(out)), ctx_(ctx), chunk_size_(chunk_size), batch_chunks_(batch_chunks),
      max_batches_(std::max<size_t>(1, max_batches)) {
    if (!ctx_) throw Error(Error::Kind::NoBackend, "no GPU context");
    if (!chunk_size_ || !batch_chunks_) throw std::invalid_argument("chunk_size and batch_chunks must be > 0");
    buf_.reserve(chunk_size_ * batch_chunks_);
}

BatchedChunkedWriter::~BatchedChunkedWriter() {
    try {
        finalize();
    } catch (...) {
    }
    for (gpz_batch *b : inflight_) gpz_batch_free(b);
}

void BatchedChunkedWriter::check_sticky() {
    if (!sticky_.empty()) throw Error(Error::Kind::Io, sticky_);
}

// Wait for the oldest batch and write its chunks' members in order.
void BatchedChunkedWriter::drain_one() {
    gpz_batch *b = inflight_.front();
    inflight_.pop_front();
    const size_t n = counts_.front();
    counts_.pop_front();
    int rc = gpz_batch_wait(b);
    for (size_t i = 0; rc == GPZ_OK && i < n; i++) {
        const uint8_t *d = nullptr;
        size_t len = 0;
        rc = gpz_batch_result(b, i, &d, &len);
        if (rc == GPZ_OK) {
            try {
                out_->write_all(d, len);
            } catch (const std::exception &e) {
                if (sticky_.empty()) sticky_ = e.what();
                rc = GPZ_E_INVAL;
            }
        }
    }
    if (rc != GPZ_OK && sticky_.empty()) sticky_ = std::string("gpu batch failed: ") + gpz_last_error();
    gpz_batch_free(b);
}

// Submit the buffered chunks (the last may be partial) as one batch.
void BatchedChunkedWriter::submit() {
    if (buf_.empty()) return;
    const size_t n = (buf_.size() + chunk_size_ - 1) / chunk_size_;
    std::vector<gpz_desc> descs(n);
    for (size_t i = 0; i < n; i++) {
        descs[i].offset = i * chunk_size_;
        descs[i].len = std::min(chunk_size_, buf_.size() - i * chunk_size_);
    }
    while (inflight_.size() >= max_batches_) drain_one();
    check_sticky();
    gpz_batch *b = nullptr;
    // the input is copied before return, so the buffer is free again at once
    if (gpz_batch_submit(ctx_, buf_.data(), descs.data(), n, &b) != GPZ_OK) {
        sticky_ = std::string("gpu batch submit failed: ") + gpz_last_error();
        check_sticky();
    }
    inflight_.push_back(b);
    counts_.push_back(n);
    buf_.clear();
    chunks_submitted_ += n;
    batches_submitted_++;
}

void BatchedChunkedWriter::write_all(const uint8_t *data, size_t len) {
    check_sticky();
    const size_t cap = chunk_size_ * batch_chunks_;
    while (len) {
        const size_t take = std::min(len, cap - buf_.size());
        buf_.insert(buf_.end(), data, data + take);
        data += take;
        len -= take;
        if (buf_.size() == cap) submit();
    }
}

void BatchedChunkedWriter::flush() {
    // like ParallelChunkedWriter::flush: a tar writer flushes between entries,
    // and cutting a chunk there would change the members
}

void BatchedChunkedWriter::finalize() {
    if (finalized_) {
        check_sticky();
        return;
    }
    finalized_ = true;
    if (sticky_.empty()) submit();
    while (!inflight_.empty()) drain_one();
    if (sticky_.empty()) {
        try {
            out_->flush();
        } catch (const std::exception &e) {
            sticky_ = e.what();
        }
    }
    check_sticky();
}

// ---------------------------------------------------------------------------
// TarBuilder (GNU format, as tar::Header::new_gnu)

namespace {

void octal(char *dst, size_t width, uint64_t v) {
    // width - 1 digits and a NUL, zero padded
    for (size_t i = width - 1; i-- > 0;) {
        dst[i] = (char)('0' + (v & 7));
        v >>= 3;
    }
    dst[width - 1] = 0;
}

}  // namespace

TarBuilder::TarBuilder(std::unique_ptr<Writer> out) : out_(std::move(out)) {}

TarBuilder::~TarBuilder() {
    try {
        finish();
    } catch (...) {
    }
}

void TarBuilder::pad(uint64_t size) {
    static const uint8_t zero[512] = {0};
    const size_t r = (size_t)(size % 512);
    if (r) out_->write_all(zero, 512 - r);
}

void TarBuilder::header(const std::string &path, char type, uint64_t size, uint32_t mode, uint64_t mtime,
                        const std::string &link) {
    // names and link targets longer than the 100-byte fields go first as GNU
    // long-name / long-link records (type L / K) holding the full string
    auto long_record = [&](char t, const std::string &s) {
        uint8_t h[512] = {0};
        memcpy(h, "././@LongLink", 13);
        octal((char *)h + 100, 8, 0644);
        octal((char *)h + 108, 8, 0);
        octal((char *)h + 116, 8, 0);
        octal((char *)h + 124, 12, s.size() + 1);
        octal((char *)h + 136, 12, 0);
        h[156] = (uint8_t)t;
        memcpy(h + 257, "ustar  ", 8);
        finish_header(h);
        out_->write_all(h, 512);
        out_->write_all((const uint8_t *)s.c_str(), s.size() + 1);
        pad(s.size() + 1);
    };
    if (path.size() > 100) long_record('L', path);
    if (link.size() > 100) long_record('K', link);
    uint8_t h[512] = {0};
    memcpy(h, path.data(), std::min<size_t>(path.size(), 100));
    octal((char *)h + 100, 8, mode & 07777);
    octal((char *)h + 108, 8, 0);  // uid
    octal((char *)h + 116, 8, 0);  // gid
    octal((char *)h + 124, 12, size);
    octal((char *)h + 136, 12, mtime);
    h[156] = (uint8_t)type;
    memcpy(h + 157, link.data(), std::min<size_t>(link.size(), 100));
    memcpy(h + 257, "ustar  ", 8);  // GNU magic and version
    finish_header(h);
    out_->write_all(h, 512);
}

void TarBuilder::finish_header(uint8_t *h) {
    memset(h + 148, ' ', 8);
    uint32_t sum = 0;
    for (int i = 0; i < 512; i++) sum += h[i];
    octal((char *)h + 148, 7, sum);  // six digits, NUL, then the space already there
    h[155] = ' ';
}

void TarBuilder::append_data(const std::string &arc_path, const uint8_t *data, size_t len, uint32_t mode,
                             uint64_t mtime) {
    header(arc_path, '0', len, mode, mtime, "");
    out_->write_all(data, len);
    pad(len);
}

void TarBuilder::append_path(const std::string &src, const std::string &arc_path) {
    struct stat st;
    if (lstat(src.c_str(), &st) != 0) throw Error(Error::Kind::Io, "cannot stat " + src);
    if (S_ISDIR(st.st_mode)) {
        header(arc_path + "/", '5', 0, st.st_mode, (uint64_t)st.st_mtime, "");
        std::vector<std::string> names;
        if (DIR *d = opendir(src.c_str())) {
            while (struct dirent *e = readdir(d))
                if (strcmp(e->d_name, ".") && strcmp(e->d_name, "..")) names.push_back(e->d_name);
            closedir(d);
        } else {
            throw Error(Error::Kind::Io, "cannot read directory " + src);
        }
        std::sort(names.begin(), names.end());  // stable order (pack.rs:111-113)
        for (const std::string &nm : names) append_path(src + "/" + nm, arc_path + "/" + nm);
    } else if (S_ISREG(st.st_mode)) {
        FILE *f = fopen(src.c_str(), "rb");
        if (!f) throw Error(Error::Kind::Io, "cannot open " + src);
        header(arc_path, '0', (uint64_t)st.st_size, st.st_mode, (uint64_t)st.st_mtime, "");
        std::vector<uint8_t> buf(1 << 20);
        uint64_t left = (uint64_t)st.st_size;
        while (left) {
            const size_t want = (size_t)std::min<uint64_t>(left, buf.size());
            const size_t got = fread(buf.data(), 1, want, f);
            if (got != want) {
                fclose(f);
                throw Error(Error::Kind::Io, "short read of " + src);
            }
            out_->write_all(buf.data(), got);
            left -= got;
        }
        fclose(f);
        pad((uint64_t)st.st_size);
    } else if (S_ISLNK(st.st_mode)) {
        std::string target(4096, '\0');
        const ssize_t n = readlink(src.c_str(), &target[0], target.size());
        if (n < 0) throw Error(Error::Kind::Io, "cannot read link " + src);
        target.resize((size_t)n);
        header(arc_path, '2', 0, st.st_mode, (uint64_t)st.st_mtime, target);
    }
    // other file types are skipped, as in pack.rs:103-141
}

void TarBuilder::finish() {
    if (finished_) return;
    finished_ = true;
    static const uint8_t zero[1024] = {0};
    out_->write_all(zero, sizeof(zero));
    out_->flush();
}

// ---------------------------------------------------------------------------

namespace {

struct FileWriter : Writer {
    FILE *f;
    explicit FileWriter(FILE *f) : f(f) {}
    ~FileWriter() override {
        if (f) fclose(f);
    }
    void write_all(const uint8_t *d, size_t n) override {
        if (fwrite(d, 1, n, f) != n) throw Error(Error::Kind::Io, "archive write failed");
    }
    void flush() override {
        if (fflush(f) != 0) throw Error(Error::Kind::Io, "archive flush failed");
    }
};

std::string base_name(const std::string &p) {
    std::string s = p;
    while (s.size() > 1 && s.back() == '/') s.pop_back();
    const size_t k = s.rfind('/');
    return k == std::string::npos ? s : s.substr(k + 1);
}

}  // namespace

void pack_tar_gz(const std::string &archive, const std::vector<std::string> &inputs, gpz_ctx *ctx,
                 size_t chunk_size, size_t batch_chunks) {
    FILE *f = fopen(archive.c_str(), "wb");
    if (!f) throw Error(Error::Kind::Io, "cannot create " + archive);
    auto gz = std::make_unique<BatchedChunkedWriter>(std::make_unique<FileWriter>(f), ctx, chunk_size, batch_chunks);
    BatchedChunkedWriter *gzp = gz.get();
    TarBuilder tb(std::move(gz));
    for (const std::string &in : inputs) {
        const std::string name = base_name(in);
        if (name.empty() || name == "/" || name == "." || name == "..")
            throw Error(Error::Kind::Io, "input has no file name: " + in);
        tb.append_path(in, name);
    }
    tb.finish();
    gzp->finalize();
}

}  // namespace gpzip
