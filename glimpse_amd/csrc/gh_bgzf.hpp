// gh_bgzf.hpp -- BGZF (blocked gzip, SAM/BAM/VCF specification section 4.1)
// input and output on zlib, for the VCF/BCF and BAM readers and the BCF/VCF
// writer.  Host code only.
//
// Input accepts BGZF, plain (multi-member) gzip and uncompressed files: the
// stream is inflated member after member with zlib's gzip decoder.  Output
// writes BGZF blocks of at most 0xff00 input bytes (each a gzip member with
// the BC extra field holding the block size) and the 28-byte end-of-file
// block.
#pragma once
#include <zlib.h>

#include <cstdint>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

namespace gh {
namespace io {

struct IoError : std::runtime_error {
    explicit IoError(const std::string &m) : std::runtime_error(m) {}
};

class InStream {
public:
    explicit InStream(const std::string &path) : path_(path) {
        f_ = std::fopen(path.c_str(), "rb");
        if (!f_) throw IoError("cannot open " + path);
        unsigned char m[2] = {0, 0};
        const size_t n = std::fread(m, 1, 2, f_);
        gz_ = n == 2 && m[0] == 0x1f && m[1] == 0x8b;
        std::rewind(f_);
        in_.resize(1 << 16);
        if (gz_) {
            std::memset(&z_, 0, sizeof(z_));
            if (inflateInit2(&z_, 15 + 16) != Z_OK) throw IoError("inflateInit2 failed");
            zinit_ = true;
        }
    }
    ~InStream() {
        if (zinit_) inflateEnd(&z_);
        if (f_) std::fclose(f_);
    }
    InStream(const InStream &) = delete;
    InStream &operator=(const InStream &) = delete;

    // Reads up to n bytes; returns the count (0 at the end).
    size_t read(void *dst, size_t n) {
        size_t got = 0;
        unsigned char *d = (unsigned char *)dst;
        while (got < n) {
            if (pos_ == buf_.size()) {
                if (!fill()) break;
                continue;
            }
            const size_t k = std::min(n - got, buf_.size() - pos_);
            std::memcpy(d + got, buf_.data() + pos_, k);
            pos_ += k;
            got += k;
        }
        return got;
    }
    void read_exact(void *dst, size_t n) {
        if (read(dst, n) != n) throw IoError("truncated file " + path_);
    }
    // One line without its end-of-line; false at the end of the stream.
    bool getline(std::string &line) {
        line.clear();
        for (;;) {
            if (pos_ == buf_.size()) {
                if (!fill()) return !line.empty();
                continue;
            }
            const unsigned char *s = buf_.data() + pos_;
            const void *nl = std::memchr(s, '\n', buf_.size() - pos_);
            if (nl) {
                const size_t k = (const unsigned char *)nl - s;
                line.append((const char *)s, k);
                pos_ += k + 1;
                if (!line.empty() && line.back() == '\r') line.pop_back();
                return true;
            }
            line.append((const char *)s, buf_.size() - pos_);
            pos_ = buf_.size();
        }
    }
    bool eof() {
        while (pos_ == buf_.size())
            if (!fill()) return true;
        return false;
    }

private:
    bool fill() {
        buf_.clear();
        pos_ = 0;
        if (!gz_) {
            buf_.resize(1 << 16);
            const size_t n = std::fread(buf_.data(), 1, buf_.size(), f_);
            buf_.resize(n);
            return n > 0;
        }
        buf_.resize(1 << 17);
        for (;;) {
            if (z_.avail_in == 0) {
                const size_t n = std::fread(in_.data(), 1, in_.size(), f_);
                if (n == 0) {
                    if (!member_done_) throw IoError("truncated gzip stream in " + path_);
                    buf_.clear();
                    return false;
                }
                z_.next_in = in_.data();
                z_.avail_in = (uInt)n;
            }
            if (member_done_) {   // a new member starts: BGZF blocks are one member each
                if (inflateReset(&z_) != Z_OK) throw IoError("inflateReset failed");
                member_done_ = false;
            }
            z_.next_out = buf_.data();
            z_.avail_out = (uInt)buf_.size();
            const int rc = inflate(&z_, Z_NO_FLUSH);
            if (rc == Z_STREAM_END) member_done_ = true;
            else if (rc != Z_OK && rc != Z_BUF_ERROR) throw IoError("corrupt gzip data in " + path_);
            const size_t produced = buf_.size() - z_.avail_out;
            if (produced) {
                buf_.resize(produced);
                return true;
            }
        }
    }

    std::string path_;
    FILE *f_ = nullptr;
    bool gz_ = false, zinit_ = false, member_done_ = true;
    z_stream z_;
    std::vector<unsigned char> in_, buf_;
    size_t pos_ = 0;
};

class OutStream {
public:
    // bgzf: BGZF blocks, else plain bytes.
    OutStream(const std::string &path, bool bgzf, int level = 6) : path_(path), bgzf_(bgzf), level_(level) {
        f_ = std::fopen(path.c_str(), "wb");
        if (!f_) throw IoError("cannot create " + path);
    }
    ~OutStream() {
        if (f_) {
            try {
                close();
            } catch (...) {
            }
        }
    }
    OutStream(const OutStream &) = delete;
    OutStream &operator=(const OutStream &) = delete;

    void write(const void *src, size_t n) {
        const unsigned char *s = (const unsigned char *)src;
        if (!bgzf_) {
            if (n && std::fwrite(s, 1, n, f_) != n) throw IoError("write failed: " + path_);
            return;
        }
        while (n) {
            const size_t k = std::min(n, kBlockIn - pend_.size());
            pend_.insert(pend_.end(), s, s + k);
            s += k;
            n -= k;
            if (pend_.size() == kBlockIn) flush_block();
        }
    }
    void write(const std::string &s) { write(s.data(), s.size()); }
    void close() {
        if (!f_) return;
        if (bgzf_) {
            if (!pend_.empty()) flush_block();
            static const unsigned char eof[28] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 0x42, 0x43,
                                                  2,    0,    0x1b, 0, 3, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            if (std::fwrite(eof, 1, 28, f_) != 28) throw IoError("write failed: " + path_);
        }
        const int rc = std::fclose(f_);
        f_ = nullptr;
        if (rc != 0) throw IoError("close failed: " + path_);
    }

private:
    static constexpr size_t kBlockIn = 0xff00;
    void flush_block() {
        std::vector<unsigned char> out(compressBound((uLong)pend_.size()) + 64);
        z_stream z;
        std::memset(&z, 0, sizeof(z));
        if (deflateInit2(&z, level_, Z_DEFLATED, -15, 8, Z_DEFAULT_STRATEGY) != Z_OK) throw IoError("deflateInit2");
        z.next_in = pend_.data();
        z.avail_in = (uInt)pend_.size();
        z.next_out = out.data() + 18;
        z.avail_out = (uInt)(out.size() - 26);
        const int rc = deflate(&z, Z_FINISH);
        const size_t clen = z.total_out;
        deflateEnd(&z);
        if (rc != Z_STREAM_END) throw IoError("deflate failed");
        const size_t bsize = clen + 26;   // header 18 + data + crc 4 + isize 4
        if (bsize > 65536) throw IoError("BGZF block too large");
        unsigned char *h = out.data();
        const unsigned char hdr[18] = {0x1f, 0x8b, 8, 4, 0, 0, 0, 0, 0, 0xff, 6, 0, 0x42, 0x43, 2, 0, 0, 0};
        std::memcpy(h, hdr, 18);
        h[16] = (unsigned char)((bsize - 1) & 0xff);
        h[17] = (unsigned char)((bsize - 1) >> 8);
        const uint32_t crc = (uint32_t)crc32(crc32(0L, Z_NULL, 0), pend_.data(), (uInt)pend_.size());
        const uint32_t isize = (uint32_t)pend_.size();
        unsigned char *t = out.data() + 18 + clen;
        for (int k = 0; k < 4; k++) t[k] = (unsigned char)(crc >> (8 * k));
        for (int k = 0; k < 4; k++) t[4 + k] = (unsigned char)(isize >> (8 * k));
        if (std::fwrite(out.data(), 1, bsize, f_) != bsize) throw IoError("write failed: " + path_);
        pend_.clear();
    }

    std::string path_;
    FILE *f_ = nullptr;
    bool bgzf_;
    int level_;
    std::vector<unsigned char> pend_;
};

}  // namespace io
}  // namespace gh
