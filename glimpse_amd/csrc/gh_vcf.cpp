// gh_vcf.cpp -- VCF / BCF input and output of the phase tool (host code).
//
// Reader: VCF text (plain or BGZF) and BCF 2.2 (BGZF), record by record, with
// htslib's conventions for what it hands to callers (bcf_get_format_int32 /
// _float and bcf_get_genotypes: per-sample vectors padded to the record's
// largest count with the vector-end value; missing values as htslib's missing
// sentinels; GT as (allele + 1) << 1 | phased).  On top of it the target GL
// reader of genotype_reader::scanTarGenotypes / parseTarGenotypes
// (phase/src/io/genotype_reader.cpp:352-521): biallelic records matched to the
// panel by position, REF and ALT; FORMAT/PL, or FORMAT/GL as lroundf(-10 GL),
// capped at 255; a site is flat for a sample unless its values differ.
//
// Writer: the output of genotype_writer::writeGenotypes
// (phase/src/io/genotype_writer.cpp:43-220): its header lines, and per site
// RAF, AF, INFO and per sample GT (phased), DS and GP; as BCF 2.2 in BGZF
// blocks (".bcf"), BGZF-compressed VCF (".vcf.gz") or plain VCF.
#include <algorithm>
#include <cctype>
#include <cerrno>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <unordered_map>
#include <vector>

#include "gh_bgzf.hpp"
#include "gh_error.hpp"

namespace gh {
namespace io {

constexpr int32_t I32_MISSING = INT32_MIN;
constexpr int32_t I32_END = INT32_MIN + 1;
constexpr uint32_t F32_MISSING = 0x7F800001u;
constexpr uint32_t F32_END = 0x7F800002u;

static inline float f32_bits(uint32_t b) {
    float f;
    std::memcpy(&f, &b, 4);
    return f;
}
static inline uint32_t bits_f32(float f) {
    uint32_t b;
    std::memcpy(&b, &f, 4);
    return b;
}

// One FORMAT or INFO field of a record, decoded: n_per values per sample
// (n_per * n_sample values for FORMAT, n_per for INFO), as int32 or float.
struct Field {
    bool is_float = false, is_string = false;
    int n_per = 0;
    std::vector<int32_t> iv;
    std::vector<float> fv;
    std::string sv;
};

struct Record {
    int32_t rid = 0, pos = 0, rlen = 0;
    std::string chrom, id;
    std::vector<std::string> alleles;
    std::map<std::string, Field> info, fmt;
};

// Header: sample names, contigs, the shared string dictionary of BCF.
struct Header {
    std::vector<std::string> samples, contigs, lines;
    std::map<std::string, std::string> fmt_type, info_type;   // ID -> Type
    std::unordered_map<int, std::string> dict;                 // BCF string index -> ID
    std::unordered_map<std::string, int> contig_idx;

    static std::string attr(const std::string &body, const std::string &key) {
        // body: the part inside <...>; values may be quoted
        size_t i = 0;
        while (i < body.size()) {
            size_t eq = body.find('=', i);
            if (eq == std::string::npos) return "";
            const std::string k = body.substr(i, eq - i);
            size_t j = eq + 1;
            std::string v;
            if (j < body.size() && body[j] == '"') {
                size_t e = j + 1;
                while (e < body.size() && body[e] != '"') e += (body[e] == '\\') ? 2 : 1;
                v = body.substr(j + 1, e - j - 1);
                j = e + 1;
            } else {
                size_t e = body.find(',', j);
                if (e == std::string::npos) e = body.size();
                v = body.substr(j, e - j);
                j = e;
            }
            if (k == key) return v;
            i = (j < body.size() && body[j] == ',') ? j + 1 : j;
        }
        return "";
    }

    void parse(const std::string &text) {
        std::unordered_map<std::string, int> ids;
        ids["PASS"] = 0;
        dict[0] = "PASS";
        int next = 1, next_ctg = 0;
        size_t p = 0;
        while (p < text.size()) {
            size_t e = text.find('\n', p);
            if (e == std::string::npos) e = text.size();
            std::string line = text.substr(p, e - p);
            p = e + 1;
            if (!line.empty() && line.back() == '\r') line.pop_back();
            if (line.empty() || line[0] == '\0') continue;
            if (line.compare(0, 6, "#CHROM") == 0) {
                size_t k = 0;
                int col = 0;
                while (k <= line.size()) {
                    size_t t = line.find('\t', k);
                    if (t == std::string::npos) t = line.size();
                    if (col >= 9) samples.push_back(line.substr(k, t - k));
                    col++;
                    k = t + 1;
                }
                continue;
            }
            lines.push_back(line);
            auto body_of = [&](const char *tag) -> std::string {
                const size_t n = std::strlen(tag);
                if (line.compare(0, n, tag) != 0) return std::string();
                const size_t a = line.find('<'), b = line.rfind('>');
                if (a == std::string::npos || b == std::string::npos || b < a) return std::string();
                return line.substr(a + 1, b - a - 1);
            };
            std::string body;
            if (!(body = body_of("##contig=")).empty()) {
                const std::string id = attr(body, "ID"), idx = attr(body, "IDX");
                const int ci = idx.empty() ? next_ctg : std::atoi(idx.c_str());
                next_ctg = std::max(next_ctg, ci + 1);
                if ((int)contigs.size() <= ci) contigs.resize(ci + 1);
                contigs[ci] = id;
                contig_idx[id] = ci;
                continue;
            }
            const bool is_info = !(body = body_of("##INFO=")).empty();
            const bool is_fmt = !is_info && !(body = body_of("##FORMAT=")).empty();
            const bool is_filter = !is_info && !is_fmt && !(body = body_of("##FILTER=")).empty();
            if (!(is_info || is_fmt || is_filter)) continue;
            const std::string id = attr(body, "ID"), idx = attr(body, "IDX");
            if (is_info) info_type[id] = attr(body, "Type");
            if (is_fmt) fmt_type[id] = attr(body, "Type");
            if (!idx.empty()) {
                const int v = std::atoi(idx.c_str());
                ids[id] = v;
                dict[v] = id;
                next = std::max(next, v + 1);
            } else if (!ids.count(id)) {
                ids[id] = next;
                dict[next] = id;
                next++;
            }
        }
    }
};

// ------------------------------------------------------------------ reader
class Reader {
public:
    explicit Reader(const std::string &path) : in_(path), path_(path) {
        char magic[5] = {0, 0, 0, 0, 0};
        // BCF starts "BCF\2\2"; VCF starts "##"
        const size_t n = in_.read(magic, 5);
        if (n == 5 && std::memcmp(magic, "BCF", 3) == 0) {
            bcf_ = true;
            if (magic[3] != 2) throw IoError("unsupported BCF major version in " + path);
            uint32_t lt = 0;
            in_.read_exact(&lt, 4);
            std::string text(lt, '\0');
            in_.read_exact(&text[0], lt);
            hdr_.parse(text);
        } else {
            std::string text(magic, n), line;
            // finish the first line, then read header lines up to #CHROM
            std::string first;
            in_.getline(first);
            text += first;
            text += '\n';
            bool chrom = text.compare(0, 6, "#CHROM") == 0;
            if (text.compare(0, 2, "##") != 0 && !chrom) throw IoError("not a VCF or BCF file: " + path);
            while (!chrom && in_.getline(line)) {
                text += line;
                text += '\n';
                if (line.compare(0, 6, "#CHROM") == 0) chrom = true;
            }
            if (!chrom) throw IoError("VCF header without #CHROM line: " + path);
            hdr_.parse(text);
        }
    }
    const Header &header() const { return hdr_; }
    bool next(Record &r) { return bcf_ ? next_bcf(r) : next_vcf(r); }

private:
    // ---- BCF typed values
    struct Typed {
        int type = 0, count = 0;
    };
    Typed read_desc(const unsigned char *&p, const unsigned char *end) {
        if (p >= end) throw IoError("truncated BCF record in " + path_);
        Typed t;
        t.type = *p & 0xf;
        t.count = *p >> 4;
        p++;
        if (t.count == 15) {
            Typed c = read_desc(p, end);
            t.count = (int)read_int(p, end, c.type);
        }
        return t;
    }
    static int type_size(int type) {
        switch (type) {
            case 0: return 0;
            case 1: return 1;
            case 2: return 2;
            case 3: return 4;
            case 5: return 4;
            case 7: return 1;
        }
        return -1;
    }
    int64_t read_int(const unsigned char *&p, const unsigned char *end, int type) {
        const int sz = type_size(type);
        if (sz <= 0 || p + sz > end) throw IoError("bad BCF integer in " + path_);
        int64_t v;
        if (type == 1) v = (int8_t)p[0];
        else if (type == 2) {
            int16_t x;
            std::memcpy(&x, p, 2);
            v = x;
        } else {
            int32_t x;
            std::memcpy(&x, p, 4);
            v = x;
        }
        p += sz;
        return v;
    }
    // an int value of a typed vector, mapped to int32 with htslib's sentinels
    static int32_t int_value(const unsigned char *p, int type) {
        if (type == 1) {
            const int8_t x = (int8_t)p[0];
            return x == INT8_MIN ? I32_MISSING : x == INT8_MIN + 1 ? I32_END : x;
        }
        if (type == 2) {
            int16_t x;
            std::memcpy(&x, p, 2);
            return x == INT16_MIN ? I32_MISSING : x == INT16_MIN + 1 ? I32_END : x;
        }
        int32_t x;
        std::memcpy(&x, p, 4);
        return x;
    }
    void read_vector(const unsigned char *&p, const unsigned char *end, const Typed &t, int n, Field &f) {
        const int sz = type_size(t.type);
        if (sz < 0) throw IoError("unknown BCF type in " + path_);
        const size_t bytes = (size_t)sz * t.count * n;
        if (p + bytes > end) throw IoError("truncated BCF field in " + path_);
        f.n_per = t.count;
        if (t.type == 7) {
            f.is_string = true;
            f.sv.assign((const char *)p, bytes);
        } else if (t.type == 5) {
            f.is_float = true;
            f.fv.resize((size_t)t.count * n);
            std::memcpy(f.fv.data(), p, bytes);
        } else if (t.type != 0) {
            f.iv.resize((size_t)t.count * n);
            for (size_t k = 0; k < f.iv.size(); k++) f.iv[k] = int_value(p + k * sz, t.type);
        }
        p += bytes;
    }
    bool next_bcf(Record &r) {
        uint32_t ls = 0, li = 0;
        const size_t n = in_.read(&ls, 4);
        if (n == 0) return false;
        if (n != 4) throw IoError("truncated BCF record in " + path_);
        in_.read_exact(&li, 4);
        rec_.resize((size_t)ls + li);
        in_.read_exact(rec_.data(), rec_.size());
        const unsigned char *p = rec_.data(), *end = p + ls;
        if (ls < 24) throw IoError("short BCF record in " + path_);
        int32_t w[6];
        std::memcpy(w, p, 24);
        p += 24;
        r = Record();
        r.rid = w[0];
        r.pos = w[1];
        r.rlen = w[2];
        const int n_info = (uint32_t)w[4] & 0xffff, n_allele = (uint32_t)w[4] >> 16;
        const int n_fmt = (uint32_t)w[5] >> 24, n_sample = (uint32_t)w[5] & 0xffffff;
        r.chrom = (r.rid >= 0 && r.rid < (int)hdr_.contigs.size()) ? hdr_.contigs[r.rid] : std::to_string(r.rid);
        {
            Typed t = read_desc(p, end);
            Field f;
            read_vector(p, end, t, 1, f);
            r.id = f.sv.empty() ? "." : std::string(f.sv.c_str());
        }
        for (int a = 0; a < n_allele; a++) {
            Typed t = read_desc(p, end);
            Field f;
            read_vector(p, end, t, 1, f);
            r.alleles.push_back(std::string(f.sv.c_str()));
        }
        {
            Typed t = read_desc(p, end);   // FILTER
            Field f;
            read_vector(p, end, t, 1, f);
        }
        for (int k = 0; k < n_info; k++) {
            Typed kt = read_desc(p, end);
            const int key = (int)read_int(p, end, kt.type);
            Typed t = read_desc(p, end);
            Field f;
            read_vector(p, end, t, 1, f);
            auto it = hdr_.dict.find(key);
            if (it != hdr_.dict.end()) r.info[it->second] = std::move(f);
        }
        p = rec_.data() + ls;
        end = p + li;
        if (n_sample != (int)hdr_.samples.size() && n_fmt > 0)
            throw IoError("BCF record sample count differs from the header in " + path_);
        for (int k = 0; k < n_fmt; k++) {
            Typed kt = read_desc(p, end);
            const int key = (int)read_int(p, end, kt.type);
            Typed t = read_desc(p, end);
            Field f;
            read_vector(p, end, t, n_sample, f);
            auto it = hdr_.dict.find(key);
            if (it != hdr_.dict.end()) r.fmt[it->second] = std::move(f);
        }
        return true;
    }

    // ---- VCF text
    static void split(const std::string &s, char d, std::vector<std::string> &out) {
        out.clear();
        size_t k = 0;
        for (;;) {
            const size_t t = s.find(d, k);
            if (t == std::string::npos) {
                out.push_back(s.substr(k));
                return;
            }
            out.push_back(s.substr(k, t - k));
            k = t + 1;
        }
    }
    static int32_t parse_int(const std::string &v) {
        if (v.empty() || v == ".") return I32_MISSING;
        char *e = nullptr;
        const long x = std::strtol(v.c_str(), &e, 10);
        return (e && *e == '\0') ? (int32_t)x : I32_MISSING;
    }
    static float parse_float(const std::string &v) {
        if (v.empty() || v == ".") return f32_bits(F32_MISSING);
        return std::strtof(v.c_str(), nullptr);
    }
    // GT text -> htslib values: (allele + 1) << 1 | phased, '.' -> 0
    static void parse_gt(const std::string &v, std::vector<int32_t> &out) {
        out.clear();
        size_t k = 0;
        bool phased = false;
        while (k <= v.size()) {
            size_t e = k;
            while (e < v.size() && v[e] != '/' && v[e] != '|') e++;
            const std::string a = v.substr(k, e - k);
            const int32_t al = (a == "." || a.empty()) ? -1 : std::atoi(a.c_str());
            out.push_back(((al + 1) << 1) | (phased ? 1 : 0));
            if (e >= v.size()) break;
            phased = v[e] == '|';
            k = e + 1;
        }
    }
    bool next_vcf(Record &r) {
        std::string line;
        do {
            if (!in_.getline(line)) return false;
        } while (line.empty() || line[0] == '#');
        split(line, '\t', cols_);
        if (cols_.size() < 8) throw IoError("VCF record with fewer than 8 columns in " + path_);
        r = Record();
        r.chrom = cols_[0];
        auto ci = hdr_.contig_idx.find(r.chrom);
        r.rid = ci == hdr_.contig_idx.end() ? -1 : ci->second;
        r.pos = std::atoi(cols_[1].c_str()) - 1;
        r.id = cols_[2];
        r.alleles.push_back(cols_[3]);
        if (cols_[4] != ".") {
            split(cols_[4], ',', parts_);
            for (auto &a : parts_) r.alleles.push_back(a);
        }
        r.rlen = (int32_t)cols_[3].size();
        if (cols_[7] != ".") {
            split(cols_[7], ';', parts_);
            for (auto &kv : parts_) {
                const size_t eq = kv.find('=');
                Field f;
                const std::string key = kv.substr(0, eq);
                if (eq == std::string::npos) {
                    r.info[key] = f;
                    continue;
                }
                const std::string val = kv.substr(eq + 1);
                const std::string ty = hdr_.info_type.count(key) ? hdr_.info_type.at(key) : "String";
                std::vector<std::string> vs;
                split(val, ',', vs);
                f.n_per = (int)vs.size();
                if (ty == "Integer") {
                    for (auto &x : vs) f.iv.push_back(parse_int(x));
                } else if (ty == "Float") {
                    f.is_float = true;
                    for (auto &x : vs) f.fv.push_back(parse_float(x));
                } else {
                    f.is_string = true;
                    f.sv = val;
                }
                r.info[key] = std::move(f);
            }
        }
        const int ns = (int)hdr_.samples.size();
        if (cols_.size() < 9 || ns == 0) return true;
        if ((int)cols_.size() < 9 + ns) throw IoError("VCF record with fewer sample columns than the header in " + path_);
        std::vector<std::string> keys;
        split(cols_[8], ':', keys);
        // per sample, per key: value strings
        std::vector<std::vector<std::string>> sv(keys.size(), std::vector<std::string>(ns));
        for (int i = 0; i < ns; i++) {
            split(cols_[9 + i], ':', parts_);
            for (size_t k = 0; k < keys.size(); k++) sv[k][i] = k < parts_.size() ? parts_[k] : ".";
        }
        std::vector<int32_t> gtv;
        for (size_t k = 0; k < keys.size(); k++) {
            Field f;
            const std::string &key = keys[k];
            if (key == "GT") {
                std::vector<std::vector<int32_t>> g(ns);
                int mx = 0;
                for (int i = 0; i < ns; i++) {
                    parse_gt(sv[k][i], g[i]);
                    mx = std::max(mx, (int)g[i].size());
                }
                f.n_per = mx;
                f.iv.assign((size_t)mx * ns, I32_END);
                for (int i = 0; i < ns; i++)
                    for (size_t a = 0; a < g[i].size(); a++) f.iv[(size_t)i * mx + a] = g[i][a];
                r.fmt[key] = std::move(f);
                continue;
            }
            const std::string ty = hdr_.fmt_type.count(key) ? hdr_.fmt_type.at(key) : "String";
            std::vector<std::vector<std::string>> vals(ns);
            int mx = 0;
            for (int i = 0; i < ns; i++) {
                split(sv[k][i], ',', vals[i]);
                mx = std::max(mx, (int)vals[i].size());
            }
            f.n_per = mx;
            if (ty == "Integer") {
                f.iv.assign((size_t)mx * ns, I32_END);
                for (int i = 0; i < ns; i++)
                    for (size_t a = 0; a < vals[i].size(); a++) f.iv[(size_t)i * mx + a] = parse_int(vals[i][a]);
            } else if (ty == "Float") {
                f.is_float = true;
                f.fv.assign((size_t)mx * ns, f32_bits(F32_END));
                for (int i = 0; i < ns; i++)
                    for (size_t a = 0; a < vals[i].size(); a++) f.fv[(size_t)i * mx + a] = parse_float(vals[i][a]);
            } else {
                f.is_string = true;
            }
            r.fmt[key] = std::move(f);
        }
        return true;
    }

    InStream in_;
    std::string path_;
    bool bcf_ = false;
    Header hdr_;
    std::vector<unsigned char> rec_;
    std::vector<std::string> cols_, parts_;
};

// htslib's bcf_get_variant_types for a biallelic record, reduced to the SNP test
// the GL reader needs: REF and ALT one base each, ALT a base (not '*', '<...>').
static bool is_snp(const Record &r) {
    if (r.alleles.size() != 2) return false;
    const std::string &a = r.alleles[0], &b = r.alleles[1];
    auto base = [](char c) {
        c = (char)std::toupper((unsigned char)c);
        return c == 'A' || c == 'C' || c == 'G' || c == 'T' || c == 'N';
    };
    return a.size() == 1 && b.size() == 1 && base(a[0]) && base(b[0]) && std::toupper(a[0]) != std::toupper(b[0]);
}

// ------------------------------------------------------------------ writer
struct TypedWriter {
    std::string s;
    void u8(uint8_t v) { s.push_back((char)v); }
    void i32(int32_t v) { s.append((const char *)&v, 4); }
    void f32(float v) { s.append((const char *)&v, 4); }
    void desc(int type, int count) {
        if (count < 15) {
            u8((uint8_t)(count << 4 | type));
        } else {
            u8((uint8_t)(15 << 4 | type));
            typed_int(count);
        }
    }
    void typed_int(int32_t v) {
        if (v >= -120 && v <= 127) {
            u8(0x11);
            u8((uint8_t)(int8_t)v);
        } else if (v >= -32760 && v <= 32767) {
            u8(0x12);
            const int16_t x = (int16_t)v;
            s.append((const char *)&x, 2);
        } else {
            u8(0x13);
            i32(v);
        }
    }
    void typed_str(const std::string &v) {
        desc(7, (int)v.size());
        s += v;
    }
};

static std::string fmt_float(float v) {
    if (std::isnan(v)) return ".";
    char b[32];
    std::snprintf(b, sizeof(b), "%g", (double)v);
    return b;
}

struct OutDesc {
    std::string contig, source, date;
    std::vector<std::string> samples;
    int n_main = 15, fploidy = 2;
};

static std::string header_text(const OutDesc &d) {
    std::string h = "##fileformat=VCFv4.2\n";
    h += "##FILTER=<ID=PASS,Description=\"All filters passed\">\n";
    h += "##fileDate=" + d.date + "\n";
    h += "##source=" + d.source + "\n";
    h += "##contig=<ID=" + d.contig + ">\n";
    h += "##INFO=<ID=RAF,Number=A,Type=Float,Description=\"ALT allele frequency in the reference panel\">\n";
    h += "##INFO=<ID=AF,Number=A,Type=Float,Description=\"ALT allele frequency computed from DS/GP field across "
         "target samples\">\n";
    if (d.fploidy > 0)
        h += "##INFO=<ID=INFO,Number=A,Type=Float,Description=\"IMPUTE info quality score for diploid samples\">\n";
    h += "##FORMAT=<ID=GT,Number=1,Type=String,Description=\"Phased genotypes\">\n";
    h += "##FORMAT=<ID=DS,Number=1,Type=Float,Description=\"Genotype dosage\">\n";
    h += "##FORMAT=<ID=GP,Number=3,Type=Float,Description=\"Genotype posteriors\">\n";
    h += "##NMAIN=" + std::to_string(d.n_main) + "\n";
    h += "##FPLOIDY=" + std::to_string(d.fploidy) + "\n";
    h += "#CHROM\tPOS\tID\tREF\tALT\tQUAL\tFILTER\tINFO\tFORMAT";
    for (auto &s : d.samples) h += "\t" + s;
    h += "\n";
    return h;
}

}  // namespace io

// ============================================================ C-ABI glue
struct VcfReader {
    io::Reader r;
    io::Record rec;
    bool have = false;
    explicit VcfReader(const std::string &p) : r(p) {}
};

}  // namespace gh

struct gh_vcf_reader {
    std::unique_ptr<gh::VcfReader> v;
};

namespace gh {

void vcf_open(const char *path, gh_vcf_reader **out) {
    GH_CHECK(path && out, GH_ERR_ARG, "gh_vcf_open: null argument");
    try {
        auto *h = new gh_vcf_reader;
        h->v.reset(new VcfReader(path));
        *out = h;
    } catch (const io::IoError &e) {
        throw Error(GH_ERR_IO, e.what());
    }
}

void vcf_close(gh_vcf_reader *r) { delete r; }

void vcf_samples(gh_vcf_reader *r, int32_t *n, char *names, int64_t cap) {
    GH_CHECK(r && n, GH_ERR_ARG, "gh_vcf_samples: null argument");
    const auto &S = r->v->r.header().samples;
    *n = (int32_t)S.size();
    if (!names) return;
    int64_t need = 0;
    for (auto &s : S) need += (int64_t)s.size() + 1;
    GH_CHECK(need <= cap, GH_ERR_CAPACITY, "gh_vcf_samples: names buffer too small");
    char *p = names;
    for (auto &s : S) {
        std::memcpy(p, s.c_str(), s.size() + 1);
        p += s.size() + 1;
    }
}

int vcf_next(gh_vcf_reader *r, gh_vcf_record *out) {
    GH_CHECK(r && out, GH_ERR_ARG, "gh_vcf_next: null argument");
    VcfReader &v = *r->v;
    try {
        v.have = v.r.next(v.rec);
    } catch (const io::IoError &e) {
        throw Error(GH_ERR_IO, e.what());
    }
    if (!v.have) return 0;
    out->chrom = v.rec.chrom.c_str();
    out->pos = v.rec.pos + 1;
    out->id = v.rec.id.c_str();
    out->n_allele = (int32_t)v.rec.alleles.size();
    out->ref = v.rec.alleles.empty() ? "" : v.rec.alleles[0].c_str();
    out->alt = v.rec.alleles.size() > 1 ? v.rec.alleles[1].c_str() : ".";
    out->is_snp = io::is_snp(v.rec) ? 1 : 0;
    return 1;
}

static const io::Field *find_field(const std::map<std::string, io::Field> &m, const char *key) {
    auto it = m.find(key);
    return it == m.end() ? nullptr : &it->second;
}

int32_t vcf_format(gh_vcf_reader *r, const char *key, int32_t *iv, float *fv, int64_t cap) {
    GH_CHECK(r && key && r->v->have, GH_ERR_ARG, "gh_vcf_format: no current record");
    const io::Field *f = find_field(r->v->rec.fmt, key);
    if (!f || f->is_string) return 0;
    const int64_t n = f->is_float ? (int64_t)f->fv.size() : (int64_t)f->iv.size();
    GH_CHECK(n <= cap, GH_ERR_CAPACITY, "gh_vcf_format: buffer too small");
    if (f->is_float) {
        GH_CHECK(fv != nullptr, GH_ERR_ARG, std::string("gh_vcf_format: ") + key + " is a Float field");
        std::memcpy(fv, f->fv.data(), sizeof(float) * n);
    } else {
        GH_CHECK(iv != nullptr, GH_ERR_ARG, std::string("gh_vcf_format: ") + key + " is an Integer field");
        std::memcpy(iv, f->iv.data(), sizeof(int32_t) * n);
    }
    return (int32_t)n;
}

int32_t vcf_info(gh_vcf_reader *r, const char *key, int32_t *iv, float *fv, int64_t cap) {
    GH_CHECK(r && key && r->v->have, GH_ERR_ARG, "gh_vcf_info: no current record");
    const io::Field *f = find_field(r->v->rec.info, key);
    if (!f || f->is_string) return 0;
    const int64_t n = f->is_float ? (int64_t)f->fv.size() : (int64_t)f->iv.size();
    GH_CHECK(n <= cap, GH_ERR_CAPACITY, "gh_vcf_info: buffer too small");
    if (f->is_float) {
        GH_CHECK(fv != nullptr, GH_ERR_ARG, std::string("gh_vcf_info: ") + key + " is a Float field");
        std::memcpy(fv, f->fv.data(), sizeof(float) * n);
    } else {
        GH_CHECK(iv != nullptr, GH_ERR_ARG, std::string("gh_vcf_info: ") + key + " is an Integer field");
        std::memcpy(iv, f->iv.data(), sizeof(int32_t) * n);
    }
    return (int32_t)n;
}

// genotype_reader::scanTarGenotypes + parseTarGenotypes for diploid targets.
void vcf_read_gls(gh_vcf_reader *r, const gh_gl_query *q, uint8_t *gl, uint8_t *flat, int64_t *n_matched) {
    GH_CHECK(r && q && gl && flat, GH_ERR_ARG, "gh_vcf_read_gls: null argument");
    GH_CHECK(q->n_sites >= 0 && q->pos && q->ref && q->alt, GH_ERR_ARG, "gh_vcf_read_gls: panel sites");
    VcfReader &v = *r->v;
    const int T = (int)v.r.header().samples.size();
    const int L = q->n_sites;
    // V.getByRef: panel sites by position, then REF and ALT; the first match wins
    std::unordered_multimap<int32_t, int32_t> by_pos;
    by_pos.reserve((size_t)L * 2);
    for (int l = L - 1; l >= 0; l--) by_pos.emplace(q->pos[l], l);
    std::memset(gl, 0, (size_t)T * L * 3);
    std::memset(flat, 1, (size_t)T * L);
    int64_t matched = 0;
    const char *key = q->use_gl ? "GL" : "PL";
    try {
        while (v.r.next(v.rec)) {
            const io::Record &rec = v.rec;
            if (rec.alleles.size() != 2) continue;
            int32_t site = -1;
            auto range = by_pos.equal_range(rec.pos + 1);
            for (auto it = range.first; it != range.second; ++it) {
                const int32_t l = it->second;
                if (rec.alleles[0] == q->ref[l] && rec.alleles[1] == q->alt[l] && (site < 0 || l < site)) site = l;
            }
            if (site < 0) continue;
            if (!q->use_indels && !io::is_snp(rec)) continue;
            const io::Field *f = find_field(rec.fmt, key);
            if (!f || f->n_per != 3 || (q->use_gl ? !f->is_float : f->is_float)) continue;   // max ploidy + 1 values
            matched++;
            for (int i = 0; i < T; i++) {
                long llk[3];
                if (q->use_gl) {
                    const float *p = f->fv.data() + (size_t)i * 3;
                    bool skip = false;
                    for (int k = 0; k < 3; k++) {
                        const uint32_t b = io::bits_f32(p[k]);
                        skip = skip || b == io::F32_MISSING || b == io::F32_END;
                    }
                    if (skip) continue;
                    for (int k = 0; k < 3; k++) llk[k] = lroundf(-10 * p[k]);
                    if (llk[0] < 0 || llk[1] < 0 || llk[2] < 0)
                        throw Error(GH_ERR_IO, "Found positive value in FORMAT/GL for individual [" +
                                                   v.r.header().samples[i] + "] at position " +
                                                   std::to_string(rec.pos + 1));
                } else {
                    const int32_t *p = f->iv.data() + (size_t)i * 3;
                    bool skip = false;
                    for (int k = 0; k < 3; k++) skip = skip || p[k] == io::I32_MISSING || p[k] == io::I32_END;
                    if (skip) continue;
                    if (p[0] < 0 || p[1] < 0 || p[2] < 0)
                        throw Error(GH_ERR_IO, "Found negative value in FORMAT/PL for individual [" +
                                                   v.r.header().samples[i] + "] at position " +
                                                   std::to_string(rec.pos + 1));
                    for (int k = 0; k < 3; k++) llk[k] = p[k];
                }
                uint8_t *g = gl + ((size_t)i * L + site) * 3;
                for (int k = 0; k < 3; k++) g[k] = (uint8_t)std::min(llk[k], 255l);
                if (!(g[0] == g[1] && g[0] == g[2])) flat[(size_t)i * L + site] = 0;
            }
        }
    } catch (const io::IoError &e) {
        throw Error(GH_ERR_IO, e.what());
    }
    v.have = false;
    if (n_matched) *n_matched = matched;
}

// genotype_writer::writeGenotypes from the output fields of gh_output_fields.
void write_genotypes(const char *path, const gh_vcf_out *d) {
    GH_CHECK(path && d && d->contig && d->n_sites >= 0 && d->n_samples > 0 && d->pos && d->ref && d->alt &&
                 d->gt && d->ds && d->gp && d->af && d->raf,
             GH_ERR_ARG, "gh_write_genotypes: bad arguments");
    GH_CHECK(d->fploidy <= 0 || d->info != nullptr, GH_ERR_ARG, "gh_write_genotypes: INFO scores missing");
    const std::string p(path);
    auto ends = [&](const char *s) {
        const size_t n = std::strlen(s);
        return p.size() >= n && p.compare(p.size() - n, n, s) == 0;
    };
    const bool bcf = ends(".bcf");
    const bool gz = bcf || ends(".gz");
    io::OutDesc od;
    od.contig = d->contig;
    od.source = d->source ? d->source : "GLIMPSE_phase (MI355X)";
    od.date = d->file_date ? d->file_date : "";
    od.n_main = d->n_main;
    od.fploidy = d->fploidy;
    for (int i = 0; i < d->n_samples; i++) od.samples.push_back(d->sample_names ? d->sample_names[i] : ("S" + std::to_string(i)));
    const int T = d->n_samples, L = d->n_sites;
    auto hap = [&](int i) { return d->ploidy != nullptr && d->ploidy[i] == 1; };
    try {
        io::OutStream out(p, gz);
        const std::string htext = io::header_text(od);
        if (bcf) {
            out.write("BCF\2\2", 5);
            const uint32_t lt = (uint32_t)htext.size() + 1;
            out.write(&lt, 4);
            out.write(htext.c_str(), lt);
        } else {
            out.write(htext);
        }
        // string dictionary: PASS 0, RAF 1, AF 2, INFO 3 (when written), GT, DS, GP
        const int kRAF = 1, kAF = 2, kINFO = 3;
        const int kGT = d->fploidy > 0 ? 4 : 3, kDS = kGT + 1, kGP = kGT + 2;
        io::TypedWriter sh, ind;
        std::string line;
        for (int l = 0; l < L; l++) {
            const std::string id = d->id && d->id[l] ? d->id[l] : ".";
            if (bcf) {
                sh.s.clear();
                ind.s.clear();
                sh.i32(0);
                sh.i32(d->pos[l] - 1);
                sh.i32((int32_t)std::strlen(d->ref[l]));
                sh.f32(io::f32_bits(io::F32_MISSING));
                const int n_info = d->fploidy > 0 ? 3 : 2;
                sh.i32((int32_t)(2u << 16 | (uint32_t)n_info));
                sh.i32((int32_t)(3u << 24 | (uint32_t)T));
                sh.typed_str(id == "." ? std::string() : id);
                sh.typed_str(d->ref[l]);
                sh.typed_str(d->alt[l]);
                sh.desc(0, 0);   // no FILTER
                sh.typed_int(kRAF);
                sh.desc(5, 1);
                sh.f32(d->raf[l]);
                sh.typed_int(kAF);
                sh.desc(5, 1);
                sh.f32(d->af[l]);
                if (d->fploidy > 0) {
                    sh.typed_int(kINFO);
                    sh.desc(5, 1);
                    sh.f32(d->info[l]);
                }
                ind.typed_int(kGT);
                ind.desc(1, 2);
                for (int i = 0; i < T; i++) {
                    const uint8_t *g = d->gt + ((size_t)i * L + l) * 2;
                    // bcf_gt_phased on both alleles, as the writer fills them; a
                    // haploid sample's second value is the vector end
                    ind.u8((uint8_t)(((g[0] + 1) << 1) | 1));
                    ind.u8(hap(i) ? (uint8_t)0x81 : (uint8_t)(((g[1] + 1) << 1) | 1));
                }
                ind.typed_int(kDS);
                ind.desc(5, 1);
                for (int i = 0; i < T; i++) ind.f32(d->ds[(size_t)i * L + l]);
                ind.typed_int(kGP);
                ind.desc(5, 3);
                for (int i = 0; i < T; i++)
                    for (int k = 0; k < 3; k++)
                        ind.f32(hap(i) && k == 2 ? io::f32_bits(io::F32_END) : d->gp[((size_t)i * L + l) * 3 + k]);
                const uint32_t ls = (uint32_t)sh.s.size(), li = (uint32_t)ind.s.size();
                out.write(&ls, 4);
                out.write(&li, 4);
                out.write(sh.s);
                out.write(ind.s);
            } else {
                line.clear();
                line += od.contig + "\t" + std::to_string(d->pos[l]) + "\t" + id + "\t" + d->ref[l] + "\t" + d->alt[l] +
                        "\t.\t.\tRAF=" + io::fmt_float(d->raf[l]) + ";AF=" + io::fmt_float(d->af[l]);
                if (d->fploidy > 0) line += ";INFO=" + io::fmt_float(d->info[l]);
                line += "\tGT:DS:GP";
                for (int i = 0; i < T; i++) {
                    const size_t o = (size_t)i * L + l;
                    const uint8_t *g = d->gt + o * 2;
                    line += "\t";
                    if (hap(i))
                        line += std::to_string(g[0]) + ":" + io::fmt_float(d->ds[o]) + ":" + io::fmt_float(d->gp[o * 3]) +
                                "," + io::fmt_float(d->gp[o * 3 + 1]);
                    else
                        line += std::to_string(g[0]) + "|" + std::to_string(g[1]) + ":" + io::fmt_float(d->ds[o]) + ":" +
                                io::fmt_float(d->gp[o * 3]) + "," + io::fmt_float(d->gp[o * 3 + 1]) + "," +
                                io::fmt_float(d->gp[o * 3 + 2]);
                }
                line += "\n";
                out.write(line);
            }
        }
        out.close();
    } catch (const io::IoError &e) {
        throw Error(GH_ERR_IO, e.what());
    }
}

}  // namespace gh

// ------------------------------------------------------------------ C-ABI
namespace {
template <typename F>
int io_guard(F &&f) {
    try {
        return f();
    } catch (const gh::Error &e) {
        gh::set_error(e.what());
        return e.code;
    } catch (const gh::io::IoError &e) {
        gh::set_error(e.what());
        return GH_ERR_IO;
    } catch (const std::exception &e) {
        gh::set_error(e.what());
        return GH_ERR_ARG;
    }
}
}  // namespace

extern "C" {

int gh_vcf_open(const char *path, gh_vcf_reader **out) {
    return io_guard([&] {
        gh::vcf_open(path, out);
        return 0;
    });
}
int gh_vcf_close(gh_vcf_reader *r) {
    return io_guard([&] {
        gh::vcf_close(r);
        return 0;
    });
}
int gh_vcf_samples(gh_vcf_reader *r, int32_t *n_samples, char *names, int64_t cap) {
    return io_guard([&] {
        gh::vcf_samples(r, n_samples, names, cap);
        return 0;
    });
}
int gh_vcf_next(gh_vcf_reader *r, gh_vcf_record *rec) {
    return io_guard([&] { return gh::vcf_next(r, rec); });
}
int gh_vcf_format(gh_vcf_reader *r, const char *key, int32_t *iv, float *fv, int64_t cap) {
    return io_guard([&] { return gh::vcf_format(r, key, iv, fv, cap); });
}
int gh_vcf_info(gh_vcf_reader *r, const char *key, int32_t *iv, float *fv, int64_t cap) {
    return io_guard([&] { return gh::vcf_info(r, key, iv, fv, cap); });
}
int gh_vcf_read_gls(gh_vcf_reader *r, const gh_gl_query *q, uint8_t *gl, uint8_t *flat, int64_t *n_matched) {
    return io_guard([&] {
        gh::vcf_read_gls(r, q, gl, flat, n_matched);
        return 0;
    });
}
int gh_write_genotypes(const char *path, const gh_vcf_out *d) {
    return io_guard([&] {
        gh::write_genotypes(path, d);
        return 0;
    });
}

}  // extern "C"
