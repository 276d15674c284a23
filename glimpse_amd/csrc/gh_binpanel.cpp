// gh_binpanel.cpp -- the split_reference binary panel (host code): the boost
// binary archive that split_reference writes (split_reference/src/caller/
// caller_initialise.cpp:86-89, "oa << H; oa << V;") and the phase tool reads
// (phase/src/caller/caller_initialise.cpp:178-246), decoded into the arrays
// gh_panel_create takes, and written back in the same layout.
//
// Archive layout (boost.serialization binary archive, 64-bit little endian):
//   header   string "serialization::archive" (u64 length + bytes), u16 library
//            version, u8 sizeof(int), sizeof(long), sizeof(float),
//            sizeof(double), i32 1;
//   H        ref_haplotype_set (common/src/containers/ref_haplotype_set.h:
//            144-160): class preamble (u8 tracking, u32 version); u32
//            n_tot_sites, n_rar_sites, n_com_sites, n_com_sites_hq, n_ref_haps;
//            vector<bool> flag_common, major_alleles (u64 count, one byte
//            each); vector<int> common2tot (u64 count, i32 each);
//            vector<vector<int>> ShapRef (preamble, u64 count, u32 item
//            version, per row u64 count + i32s); bitmatrix HvarRef (preamble,
//            u64 n_bytes, n_cols, n_rows, raw bytes; bitmatrix.h:63-77);
//            vector<uchar> Ypacked; vector<vector<int>> A_small_idx;
//   V        variant_map (variant_map.h:71-84): preamble; string chrid; i32
//            input_start, input_stop, output_start, output_stop; strings
//            input_gregion, output_gregion; vector<variant*> vec_pos
//            (preamble, u64 count, u32 item version; per pointer u16 class
//            id, for the first one u8 tracking + u32 version, u32 object id;
//            then the variant: i32 bp, strings id, ref, alt, u8 type, i32 idx,
//            u32 cref, calt, f64 cm, u8 LQ; variant.h:59-73); the multimap
//            map_pos that follows repeats vec_pos and is not read.
// The widths of the archive's bookkeeping fields changed across boost
// releases; the reader tries the current layout first and the older widths
// after it, and keeps the first decoding whose counts agree with each other.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "gh_error.hpp"

namespace gh {
namespace bin {

struct Layout {
    int ver_w;    // class version field: 4 (library version > 7) or 1
    int iv_w;     // item version of collections of classes: 4 or 1
    int cs_w;     // collection sizes: 8 (library version > 5) or 4
    bool oid;     // object ids of tracked pointers
};

struct Cursor {
    const uint8_t *p, *end;
    const char *what = "";
    void need(size_t n) {
        if ((size_t)(end - p) < n) throw Error(GH_ERR_IO, std::string("binary panel truncated while reading ") + what);
    }
    template <typename T>
    T get() {
        need(sizeof(T));
        T v;
        std::memcpy(&v, p, sizeof(T));
        p += sizeof(T);
        return v;
    }
    uint64_t uint(int w) {
        if (w == 8) return get<uint64_t>();
        if (w == 4) return get<uint32_t>();
        if (w == 2) return get<uint16_t>();
        return get<uint8_t>();
    }
    std::string str() {
        const uint64_t n = get<uint64_t>();
        need(n);
        std::string s((const char *)p, (size_t)n);
        p += n;
        return s;
    }
};

struct Panel {
    uint16_t lib = 0;
    uint32_t n_tot = 0, n_rar = 0, n_com = 0, n_com_hq = 0, n_haps = 0;
    std::vector<uint8_t> flag_common, major;
    std::vector<int32_t> common2tot;
    std::vector<std::vector<int32_t>> shap, asmall;
    uint64_t hv_bytes = 0, hv_cols = 0, hv_rows = 0;
    std::vector<uint8_t> hv, ypacked;
    std::string chrid, in_greg, out_greg;
    int32_t in_start = 0, in_stop = 0, out_start = 0, out_stop = 0;
    std::vector<int32_t> bp, idx;
    std::vector<std::string> id, ref, alt;
    std::vector<uint8_t> type, lq;
    std::vector<uint32_t> cref, calt;
    std::vector<double> cm;
    // derived for gh_panel_desc
    std::vector<uint64_t> words;
    std::vector<int32_t> rare_ptr, rare_idx, asmall_ptr, asmall_idx;
    std::vector<const char *> id_c, ref_c, alt_c;
    float sparse_maf = 0.001f;
};

static void vec_bool(Cursor &c, const Layout &L, std::vector<uint8_t> &v) {
    const uint64_t n = c.uint(L.cs_w);
    c.need(n);
    v.assign(c.p, c.p + n);
    c.p += n;
}
static void vec_int(Cursor &c, const Layout &L, std::vector<int32_t> &v) {
    const uint64_t n = c.uint(L.cs_w);
    c.need(n * 4);
    v.resize(n);
    std::memcpy(v.data(), c.p, n * 4);
    c.p += n * 4;
}
static void vec_vec_int(Cursor &c, const Layout &L, std::vector<std::vector<int32_t>> &v, uint64_t max_rows) {
    const uint64_t n = c.uint(L.cs_w);
    if (n > max_rows) throw Error(GH_ERR_IO, std::string("implausible row count in ") + c.what);
    (void)c.uint(L.iv_w);
    v.resize(n);
    for (auto &r : v) vec_int(c, L, r);
}
static void preamble(Cursor &c, const Layout &L) {
    const uint8_t tracking = c.get<uint8_t>();
    (void)c.uint(L.ver_w);
    if (tracking > 1) throw Error(GH_ERR_IO, std::string("bad class preamble before ") + c.what);
}

static void decode(const std::vector<uint8_t> &buf, const Layout &L, Panel &P) {
    Cursor c{buf.data(), buf.data() + buf.size()};
    c.what = "the archive header";
    const std::string sig = c.str();
    if (sig != "serialization::archive") throw Error(GH_ERR_IO, "not a boost binary archive (invalid signature)");
    P.lib = c.get<uint16_t>();
    const uint8_t si = c.get<uint8_t>(), sl = c.get<uint8_t>(), sf = c.get<uint8_t>(), sd = c.get<uint8_t>();
    if (si != 4 || sl != 8 || sf != 4 || sd != 8 || c.get<int32_t>() != 1)
        throw Error(GH_ERR_IO, "binary panel written on an incompatible platform (incompatible native format)");
    c.what = "ref_haplotype_set";
    preamble(c, L);
    P.n_tot = c.get<uint32_t>();
    P.n_rar = c.get<uint32_t>();
    P.n_com = c.get<uint32_t>();
    P.n_com_hq = c.get<uint32_t>();
    P.n_haps = c.get<uint32_t>();
    if (P.n_rar + P.n_com != P.n_tot || P.n_com_hq > P.n_com || P.n_haps == 0)
        throw Error(GH_ERR_IO, "inconsistent site counts in the binary panel");
    c.what = "flag_common";
    vec_bool(c, L, P.flag_common);
    c.what = "major_alleles";
    vec_bool(c, L, P.major);
    c.what = "common2tot";
    vec_int(c, L, P.common2tot);
    if (P.flag_common.size() != P.n_tot || P.major.size() != P.n_tot || P.common2tot.size() != P.n_com)
        throw Error(GH_ERR_IO, "binary panel site vectors disagree with the site counts");
    c.what = "ShapRef";
    preamble(c, L);
    vec_vec_int(c, L, P.shap, P.n_haps);
    if (P.shap.size() != P.n_haps) throw Error(GH_ERR_IO, "ShapRef rows disagree with the haplotype count");
    c.what = "HvarRef";
    preamble(c, L);
    P.hv_bytes = c.get<uint64_t>();
    P.hv_cols = c.get<uint64_t>();
    P.hv_rows = c.get<uint64_t>();
    if (P.hv_rows != ((uint64_t)P.n_com + 7) / 8 * 8 && P.hv_rows != P.n_com)
        throw Error(GH_ERR_IO, "HvarRef rows disagree with the common site count");
    if (P.hv_cols < P.n_haps || P.hv_bytes != P.hv_rows * P.hv_cols / 8)
        throw Error(GH_ERR_IO, "HvarRef shape disagrees with the haplotype count");
    c.need(P.hv_bytes);
    P.hv.assign(c.p, c.p + P.hv_bytes);
    c.p += P.hv_bytes;
    c.what = "Ypacked";
    vec_bool(c, L, P.ypacked);
    c.what = "A_small_idx";
    vec_vec_int(c, L, P.asmall, (uint64_t)P.n_com_hq + 1);
    if (P.asmall.size() != (size_t)P.n_com_hq + 1) throw Error(GH_ERR_IO, "A_small_idx rows disagree with the HQ site count");
    c.what = "variant_map";
    preamble(c, L);
    P.chrid = c.str();
    P.in_start = c.get<int32_t>();
    P.in_stop = c.get<int32_t>();
    P.out_start = c.get<int32_t>();
    P.out_stop = c.get<int32_t>();
    P.in_greg = c.str();
    P.out_greg = c.str();
    c.what = "vec_pos";
    preamble(c, L);
    const uint64_t nv = c.uint(L.cs_w);
    if (nv != P.n_tot) throw Error(GH_ERR_IO, "variant count disagrees with the site count");
    (void)c.uint(L.iv_w);
    for (uint64_t v = 0; v < nv; v++) {
        c.what = "a variant";
        (void)c.get<int16_t>();   // class id
        if (v == 0) {
            const uint8_t tracking = c.get<uint8_t>();
            (void)c.uint(L.ver_w);
            if (tracking > 1) throw Error(GH_ERR_IO, "bad class preamble of variant");
        }
        if (L.oid) {
            const uint32_t oid = c.get<uint32_t>();
            if (v == 0 && oid > 64) throw Error(GH_ERR_IO, "implausible object id");
        }
        P.bp.push_back(c.get<int32_t>());
        P.id.push_back(c.str());
        P.ref.push_back(c.str());
        P.alt.push_back(c.str());
        P.type.push_back(c.get<uint8_t>());
        P.idx.push_back(c.get<int32_t>());
        P.cref.push_back(c.get<uint32_t>());
        P.calt.push_back(c.get<uint32_t>());
        P.cm.push_back(c.get<double>());
        P.lq.push_back(c.get<uint8_t>());
        if (P.idx.back() != (int32_t)v || P.lq.back() > 1) throw Error(GH_ERR_IO, "variant record out of sequence");
    }
}

// What gh_panel_create takes, from the decoded archive.
static void derive(Panel &P) {
    const uint32_t N = P.n_haps, Lc = P.n_com, L = P.n_tot;
    const size_t W = (N + 63) / 64, row_bytes = P.hv_cols / 8;
    P.words.assign((size_t)Lc * W, 0);
    for (uint32_t r = 0; r < Lc; r++)
        for (uint32_t h = 0; h < N; h++) {
            // bitmatrix::get: bit 7 - col % 8 of byte row * (n_cols / 8) + col / 8
            const uint8_t b = (P.hv[(size_t)r * row_bytes + h / 8] >> (7 - h % 8)) & 1;
            if (b) P.words[(size_t)r * W + h / 64] |= 1ull << (h % 64);
        }
    std::vector<int32_t> cnt(L + 1, 0);
    for (uint32_t h = 0; h < N; h++)
        for (int32_t s : P.shap[h]) {
            if (s < 0 || (uint32_t)s >= L || P.flag_common[s]) throw Error(GH_ERR_IO, "ShapRef names a site that is not rare");
            cnt[s + 1]++;
        }
    P.rare_ptr.assign(L + 1, 0);
    for (uint32_t l = 0; l < L; l++) P.rare_ptr[l + 1] = P.rare_ptr[l] + cnt[l + 1];
    P.rare_idx.assign(std::max<int32_t>(P.rare_ptr[L], 1), 0);
    std::vector<int32_t> fill(P.rare_ptr.begin(), P.rare_ptr.end() - 1);
    for (uint32_t h = 0; h < N; h++)   // haplotype order, as transposeRareRef pushes them
        for (int32_t s : P.shap[h]) P.rare_idx[fill[s]++] = (int32_t)h;
    P.asmall_ptr.assign(P.asmall.size() + 1, 0);
    for (size_t k = 0; k < P.asmall.size(); k++) P.asmall_ptr[k + 1] = P.asmall_ptr[k] + (int32_t)P.asmall[k].size();
    P.asmall_idx.clear();
    for (auto &r : P.asmall) P.asmall_idx.insert(P.asmall_idx.end(), r.begin(), r.end());
    if (P.asmall_idx.empty()) P.asmall_idx.push_back(0);
    // the sparse-MAF threshold the flags were made with: commonness is MAF >=
    // threshold (ref_genotype_reader.cpp:171-172), so the smallest common MAF works
    float thr = 1.0f;
    float max_rare = -1.0f;
    for (uint32_t l = 0; l < L; l++) {
        const float tot = (float)(P.cref[l] + P.calt[l]);
        const float maf = std::min(P.cref[l] * 1.0f / tot, P.calt[l] * 1.0f / tot);
        if (P.flag_common[l]) thr = std::min(thr, maf);
        else max_rare = std::max(max_rare, maf);
        if (P.major[l] != (P.calt[l] > P.cref[l] ? 1 : 0)) throw Error(GH_ERR_IO, "major allele flags disagree with the allele counts");
    }
    if (!(max_rare < thr)) throw Error(GH_ERR_IO, "common-site flags are not a MAF threshold of the allele counts");
    P.sparse_maf = thr;
    for (uint32_t c = 0; c < Lc; c++)
        if (P.common2tot[c] < 0 || (uint32_t)P.common2tot[c] >= L || !P.flag_common[P.common2tot[c]])
            throw Error(GH_ERR_IO, "common2tot names a site that is not common");
    for (uint32_t l = 0; l < L; l++) {
        P.id_c.push_back(P.id[l].c_str());
        P.ref_c.push_back(P.ref[l].c_str());
        P.alt_c.push_back(P.alt[l].c_str());
    }
}

static const Layout kLayouts[] = {
    {4, 4, 8, true},    // boost >= 1.43 archives (library versions 8 and later)
    {1, 4, 8, true},
    {4, 1, 8, true},
    {4, 4, 8, false},
    {1, 1, 4, true},
};

// ------------------------------------------------------------------ writer
struct Out {
    std::vector<uint8_t> b;
    template <typename T>
    void put(T v) {
        const uint8_t *p = (const uint8_t *)&v;
        b.insert(b.end(), p, p + sizeof(T));
    }
    void str(const std::string &s) {
        put<uint64_t>(s.size());
        b.insert(b.end(), s.begin(), s.end());
    }
    void pre() {
        put<uint8_t>(0);
        put<uint32_t>(0);
    }
};

}  // namespace bin
}  // namespace gh

struct gh_bin_panel {
    gh::bin::Panel P;
};

namespace gh {

void bin_open(const char *path, gh_bin_panel **out) {
    GH_CHECK(path && out, GH_ERR_ARG, "gh_bin_open: null argument");
    FILE *f = std::fopen(path, "rb");
    GH_CHECK(f != nullptr, GH_ERR_IO, std::string("could not open binary panel ") + path);
    std::vector<uint8_t> buf;
    uint8_t tmp[1 << 16];
    size_t n;
    while ((n = std::fread(tmp, 1, sizeof(tmp), f)) > 0) buf.insert(buf.end(), tmp, tmp + n);
    std::fclose(f);
    std::string first_err;
    for (const auto &L : bin::kLayouts) {
        auto h = std::make_unique<gh_bin_panel>();
        try {
            bin::decode(buf, L, h->P);
            bin::derive(h->P);
            *out = h.release();
            return;
        } catch (const Error &e) {
            if (first_err.empty()) first_err = e.what();
            if (std::string(e.what()).find("signature") != std::string::npos ||
                std::string(e.what()).find("native format") != std::string::npos)
                throw;
        }
    }
    throw Error(GH_ERR_IO, "could not decode binary panel " + std::string(path) + ": " + first_err);
}

void bin_close(gh_bin_panel *b) { delete b; }

void bin_info(const gh_bin_panel *b, gh_bin_panel_info *info) {
    GH_CHECK(b && info, GH_ERR_ARG, "gh_bin_info: null argument");
    const bin::Panel &P = b->P;
    info->chrid = P.chrid.c_str();
    info->input_start = P.in_start;
    info->input_stop = P.in_stop;
    info->output_start = P.out_start;
    info->output_stop = P.out_stop;
    info->input_gregion = P.in_greg.c_str();
    info->output_gregion = P.out_greg.c_str();
    info->n_sites = (int32_t)P.n_tot;
    info->n_haps = (int32_t)P.n_haps;
    info->n_common = (int32_t)P.n_com;
    info->n_common_hq = (int32_t)P.n_com_hq;
    info->library_version = P.lib;
    info->sparse_maf = P.sparse_maf;
}

void bin_desc(const gh_bin_panel *b, gh_panel_desc *d) {
    GH_CHECK(b && d, GH_ERR_ARG, "gh_bin_desc: null argument");
    const bin::Panel &P = b->P;
    std::memset(d, 0, sizeof(*d));
    d->n_sites = (int32_t)P.n_tot;
    d->n_haps = (int32_t)P.n_haps;
    d->cm = P.cm.data();
    d->cref = P.cref.data();
    d->calt = P.calt.data();
    d->lq = P.lq.data();
    d->common_words = P.words.data();
    d->rare_ptr = P.rare_ptr.data();
    d->rare_idx = P.rare_idx.data();
    d->ypacked = P.ypacked.data();
    d->ypacked_len = (int64_t)P.ypacked.size();
    d->asmall_ptr = P.asmall_ptr.data();
    d->asmall_idx = P.asmall_idx.data();
    d->sparse_maf = P.sparse_maf;
}

void bin_sites(const gh_bin_panel *b, const int32_t **pos, const char *const **id, const char *const **ref,
               const char *const **alt, const uint8_t **type) {
    GH_CHECK(b, GH_ERR_ARG, "gh_bin_sites: null argument");
    const bin::Panel &P = b->P;
    if (pos) *pos = P.bp.data();
    if (id) *id = P.id_c.data();
    if (ref) *ref = P.ref_c.data();
    if (alt) *alt = P.alt_c.data();
    if (type) *type = P.type.data();
}

void bin_write(const char *path, const gh_panel_desc *d, const gh_bin_meta *m) {
    GH_CHECK(path && d && m && m->pos && m->ref && m->alt && m->chrid, GH_ERR_ARG, "gh_bin_write: null argument");
    GH_CHECK(d->ypacked && d->asmall_ptr && d->asmall_idx && d->common_words && d->rare_ptr && d->cm, GH_ERR_ARG,
             "gh_bin_write: the description needs the sparse PBWT (gh_build_sparse_pbwt)");
    const int L = d->n_sites, N = d->n_haps;
    const float smaf = d->sparse_maf > 0 ? d->sparse_maf : 0.001f;
    std::vector<uint8_t> common(L), major(L);
    std::vector<int32_t> c2t;
    for (int l = 0; l < L; l++) {
        const float tot = (float)(d->cref[l] + d->calt[l]);
        const double maf = std::min(d->cref[l] * 1.0f / tot, d->calt[l] * 1.0f / tot);
        common[l] = maf >= smaf;
        major[l] = d->calt[l] > d->cref[l];
        if (common[l]) c2t.push_back(l);
    }
    const int Lc = (int)c2t.size();
    int n_hq = 0;   // HQ common sites: common and not LQ (ref_genotype_reader.cpp:290)
    for (int c = 0; c < Lc; c++) n_hq += !(d->lq && d->lq[c2t[c]]);
    bin::Out o;
    o.str("serialization::archive");
    o.put<uint16_t>((uint16_t)(m->library_version ? m->library_version : 18));
    o.put<uint8_t>(4);
    o.put<uint8_t>(8);
    o.put<uint8_t>(4);
    o.put<uint8_t>(8);
    o.put<int32_t>(1);
    o.pre();
    o.put<uint32_t>((uint32_t)L);
    o.put<uint32_t>((uint32_t)(L - Lc));
    o.put<uint32_t>((uint32_t)Lc);
    o.put<uint32_t>((uint32_t)n_hq);
    o.put<uint32_t>((uint32_t)N);
    o.put<uint64_t>(L);
    o.b.insert(o.b.end(), common.begin(), common.end());
    o.put<uint64_t>(L);
    o.b.insert(o.b.end(), major.begin(), major.end());
    o.put<uint64_t>(Lc);
    for (int c = 0; c < Lc; c++) o.put<int32_t>(c2t[c]);
    // ShapRef: rare sites each haplotype carries the non-major allele at
    std::vector<std::vector<int32_t>> shap(N);
    for (int l = 0; l < L; l++)
        if (!common[l])
            for (int k = d->rare_ptr[l]; k < d->rare_ptr[l + 1]; k++) shap[d->rare_idx[k]].push_back(l);
    o.pre();
    o.put<uint64_t>(N);
    o.put<uint32_t>(0);
    for (auto &r : shap) {
        o.put<uint64_t>(r.size());
        for (int32_t v : r) o.put<int32_t>(v);
    }
    // HvarRef: rows and columns padded to multiples of 8, MSB-first bytes
    const uint64_t rows = ((uint64_t)Lc + 7) / 8 * 8, cols = ((uint64_t)N + 7) / 8 * 8;
    const size_t W = ((size_t)N + 63) / 64;
    std::vector<uint8_t> hv(rows * cols / 8, 0);
    for (int c = 0; c < Lc; c++)
        for (int h = 0; h < N; h++)
            if ((d->common_words[(size_t)c * W + h / 64] >> (h % 64)) & 1) hv[(size_t)c * (cols / 8) + h / 8] |= 1 << (7 - h % 8);
    o.pre();
    o.put<uint64_t>(rows * cols / 8);
    o.put<uint64_t>(cols);
    o.put<uint64_t>(rows);
    o.b.insert(o.b.end(), hv.begin(), hv.end());
    o.put<uint64_t>((uint64_t)d->ypacked_len);
    o.b.insert(o.b.end(), d->ypacked, d->ypacked + d->ypacked_len);
    GH_CHECK(d->asmall_ptr[0] == 0, GH_ERR_ARG, "gh_bin_write: A_small_idx pointers");
    o.put<uint64_t>((uint64_t)n_hq + 1);
    o.put<uint32_t>(0);
    for (int k = 0; k <= n_hq; k++) {
        o.put<uint64_t>((uint64_t)(d->asmall_ptr[k + 1] - d->asmall_ptr[k]));
        for (int j = d->asmall_ptr[k]; j < d->asmall_ptr[k + 1]; j++) o.put<int32_t>(d->asmall_idx[j]);
    }
    // variant_map
    o.pre();
    o.str(m->chrid);
    o.put<int32_t>(m->input_start);
    o.put<int32_t>(m->input_stop);
    o.put<int32_t>(m->output_start);
    o.put<int32_t>(m->output_stop);
    o.str(m->input_gregion ? m->input_gregion : "");
    o.str(m->output_gregion ? m->output_gregion : "");
    o.pre();
    o.put<uint64_t>(L);
    o.put<uint32_t>(0);
    const int16_t variant_class = 8;
    for (int l = 0; l < L; l++) {
        o.put<int16_t>(variant_class);
        if (l == 0) {
            o.put<uint8_t>(1);     // tracked: variants are saved through pointers
            o.put<uint32_t>(0);
        }
        o.put<uint32_t>((uint32_t)l);
        o.put<int32_t>(m->pos[l]);
        o.str(m->id && m->id[l] ? m->id[l] : ".");
        o.str(m->ref[l]);
        o.str(m->alt[l]);
        const bool snp = std::strlen(m->ref[l]) == 1 && std::strlen(m->alt[l]) == 1;
        o.put<uint8_t>(m->type ? m->type[l] : (uint8_t)(snp ? 1 : 4));   // htslib VCF_SNP / VCF_INDEL
        o.put<int32_t>(l);
        o.put<uint32_t>(d->cref[l]);
        o.put<uint32_t>(d->calt[l]);
        o.put<double>(d->cm[l]);
        o.put<uint8_t>(d->lq && d->lq[l] ? 1 : 0);
    }
    // map_pos: multimap<int, variant*> (every entry a reference to a saved variant)
    o.pre();
    o.put<uint64_t>(L);
    o.put<uint32_t>(0);
    for (int l = 0; l < L; l++) {
        if (l == 0) o.pre();   // std::pair preamble
        o.put<int32_t>(m->pos[l]);
        o.put<int16_t>(variant_class);
        o.put<uint32_t>((uint32_t)l);
    }
    FILE *f = std::fopen(path, "wb");
    GH_CHECK(f != nullptr, GH_ERR_IO, std::string("could not create ") + path);
    const size_t w = std::fwrite(o.b.data(), 1, o.b.size(), f);
    const int rc = std::fclose(f);
    GH_CHECK(w == o.b.size() && rc == 0, GH_ERR_IO, std::string("write failed: ") + path);
}

}  // namespace gh

namespace {
template <typename F>
int bin_guard(F &&f) {
    try {
        f();
        return 0;
    } catch (const gh::Error &e) {
        gh::set_error(e.what());
        return e.code;
    } catch (const std::exception &e) {
        gh::set_error(e.what());
        return GH_ERR_ARG;
    }
}
}  // namespace

extern "C" {
int gh_bin_open(const char *path, gh_bin_panel **out) {
    return bin_guard([&] { gh::bin_open(path, out); });
}
int gh_bin_close(gh_bin_panel *b) {
    return bin_guard([&] { gh::bin_close(b); });
}
int gh_bin_info(const gh_bin_panel *b, gh_bin_panel_info *info) {
    return bin_guard([&] { gh::bin_info(b, info); });
}
int gh_bin_desc(const gh_bin_panel *b, gh_panel_desc *d) {
    return bin_guard([&] { gh::bin_desc(b, d); });
}
int gh_bin_sites(const gh_bin_panel *b, const int32_t **pos, const char *const **id, const char *const **ref,
                 const char *const **alt, const uint8_t **type) {
    return bin_guard([&] { gh::bin_sites(b, pos, id, ref, alt, type); });
}
int gh_bin_write(const char *path, const gh_panel_desc *d, const gh_bin_meta *m) {
    return bin_guard([&] { gh::bin_write(path, d, m); });
}
}
