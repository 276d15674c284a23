// gh_bam.hip -- genotype likelihoods of one sample from its BAM file at the
// panel's SNP sites (genotype_bam_caller::glimpse_mpileup_reg with the
// "standard" model, phase/src/io/genotype_bam_caller.cpp:225-488).
//
// The host decodes the BGZF-compressed BAM records (SAM specification 4.2),
// applies read_bam's filters (:30-59) and cuts every kept read's alignment
// into its M / = / X segments.  The device does the per-site work:
//   bam_count_kernel / bam_fill_kernel  one thread per aligned segment finds
//                      the panel sites inside the segment (binary search in
//                      the sorted positions) and emits the read's key
//                      quality << 5 | reverse << 4 | base for each site
//                      (bcf_call_glfgen_min, :270-290), counted, then placed
//                      into per-site lists;
//   bam_site_kernel    one thread per site sorts its keys, keeps the first
//                      max_depth of them (the reference shuffles, then sorts
//                      the whole list and reads only its head, so the kept
//                      keys are the smallest), accumulates the log10
//                      likelihoods of the three genotypes over groups of
//                      identical keys in float (standard_errmod, :406-488) and
//                      writes GL as the reference quantises it (:240-245).
// Indel sites are not called (the reference's default without
// --call-indels): they stay flat.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gh_bgzf.hpp"
#include "gh_internal.hpp"

namespace gh {

constexpr uint16_t BAM_FPAIRED = 0x1, BAM_FPROPER = 0x2, BAM_FMUNMAP = 0x8, BAM_FREVERSE = 0x10,
                   BAM_FMREVERSE = 0x20;

// calling_utils tables of genotype_bam_caller.h:36-42 for qualities 0..255:
// log10(1 - e), log10(e / 3), log10(1/2 - e / 3), log10(e), e = 10^(-q/10),
// each printed to six significant digits; where the reference's printed
// table departs from that rule, its entry.
static void bam_tables(float *one_m, float *div3, float *half, float *err) {
    char b[64];
    auto g = [&](double v) {
        std::snprintf(b, sizeof(b), "%g", v);
        return std::strtof(b, nullptr);
    };
    for (int q = 0; q < 256; q++) {
        const double e = std::pow(10.0, -q / 10.0);
        one_m[q] = q == 0 ? 0.99999f : q < 163 ? g(std::log10(1.0 - e)) : -4.82164e-17f;
        div3[q] = g(std::log10(e / 3.0));
        half[q] = g(std::log10(0.5 - e / 3.0));
        err[q] = g(-q / 10.0);
    }
    one_m[61] = -3.44972e-07f;
    one_m[97] = -8.66531e-11f;
    half[23] = -0.302483f;
    half[43] = -0.301044f;
}

// seq_nt16_int: A C G T -> 0..3, every other 4-bit code -> 4
__host__ __device__ __forceinline__ int nt16_int(int c) {
    return c == 1 ? 0 : c == 2 ? 1 : c == 4 ? 2 : c == 8 ? 3 : 4;
}

struct BamSeg {
    int32_t ref;    // first reference position (0-based)
    int32_t len;
    int64_t base;   // index of the segment's first base in the concatenated bases
    int32_t rev;    // read on the reverse strand
    int32_t pad;
};

__device__ __forceinline__ int lower_pos(const int32_t *__restrict__ pos, int L, int32_t x) {
    int lo = 0, hi = L;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (pos[mid] < x) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// FILL = false: count the keys of each site; true: write them.
template <bool FILL>
__global__ __launch_bounds__(256) void bam_emit_kernel(int n_seg, const BamSeg *__restrict__ segs, int L,
                                                       const int32_t *__restrict__ pos,
                                                       const uint8_t *__restrict__ callable,
                                                       const uint8_t *__restrict__ bases,
                                                       const uint8_t *__restrict__ quals, int min_bq,
                                                       int32_t *__restrict__ cnt, const int64_t *__restrict__ off,
                                                       uint16_t *__restrict__ keys) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n_seg) return;
    const BamSeg g = segs[s];
    for (int l = lower_pos(pos, L, g.ref); l < L && pos[l] < g.ref + g.len; l++) {
        if (!callable[l]) continue;
        const int64_t qp = g.base + (pos[l] - g.ref);
        const int b = nt16_int(bases[qp]);
        const int q = quals[qp];
        if (b > 3 || q < min_bq) continue;
        const int slot = atomicAdd(&cnt[l], 1);
        if (FILL) keys[off[l] + slot] = (uint16_t)(q << 5 | g.rev << 4 | b);
    }
}

__global__ __launch_bounds__(256) void bam_site_kernel(int L, int ploidy, int max_dp, const int64_t *__restrict__ off,
                                                       const int32_t *__restrict__ cnt,
                                                       const uint8_t *__restrict__ ref_code,
                                                       const uint8_t *__restrict__ alt_code,
                                                       uint16_t *__restrict__ keys, const float *__restrict__ tab,
                                                       uint8_t *__restrict__ gl, uint8_t *__restrict__ flat,
                                                       int32_t *__restrict__ depth) {
#pragma clang fp contract(off)
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    const int n = cnt[l];
    const int P1 = ploidy + 1;
    depth[l] = 0;
    flat[l] = 1;
    for (int j = 0; j < P1; j++) gl[(int64_t)l * P1 + j] = 0;
    if (n == 0) return;
    uint16_t *k = keys + off[l];
    for (int i = 1; i < n; i++) {   // insertion sort: lists hold about the coverage
        const uint16_t v = k[i];
        int j = i - 1;
        while (j >= 0 && k[j] > v) {
            k[j + 1] = k[j];
            j--;
        }
        k[j + 1] = v;
    }
    const int m = min(n, max_dp);
    const float *one_m = tab, *div3 = tab + 256, *half = tab + 512, *err = tab + 768;
    const int r = ref_code[l], a = alt_code[l];
    float prr = 0.0f, pra = 0.0f, paa = 0.0f;
    int dp = 0;
    for (int j = 0; j < m;) {
        const uint16_t b = k[j];
        int g = 1;
        while (j + g < m && k[j + g] == b) g++;
        const int q = (b >> 5) < 4 ? 4 : (b >> 5);
        const int base = b & 0xf;
        if (ploidy > 1) {
            const float fn = (float)g;
            const float em = one_m[q] * fn, d3 = div3[q] * fn, h = half[q] * fn;
            prr += base == r ? em : d3;
            paa += base == a ? em : d3;
            pra += (base == r || base == a) ? h : d3;
        } else {   // the haploid model adds its terms once per group
            prr += base == r ? one_m[q] : err[q];
            paa += base == a ? one_m[q] : err[q];
        }
        dp += g;
        j += g;
    }
    depth[l] = dp;
    if (dp <= 0) return;
    const float llk[3] = {prr, ploidy > 1 ? pra : paa, paa};
    int mx = 0;
    for (int j = 1; j < P1; j++)
        if (llk[j] > llk[mx]) mx = j;
    uint8_t v[3];
    for (int j = 0; j < P1; j++) {
        const float d = -10.0f * (llk[j] - (float)mx);   // the reference subtracts the index
        v[j] = (uint8_t)min((int)d, 255);
        gl[(int64_t)l * P1 + j] = v[j];
    }
    if (!(v[0] == v[1] && v[0] == v[ploidy])) flat[l] = 0;
}

namespace {

struct BamData {
    std::vector<BamSeg> segs;
    std::vector<uint8_t> bases, quals;
    int64_t n_read = 0, n_kept = 0;
};

// read_bam (genotype_bam_caller.cpp:30-59)
bool keep_read(int32_t tid, int32_t pos, uint8_t mapq, uint16_t flag, int32_t mtid, int32_t mpos,
               const gh_bam_params &p) {
    if (tid < 0 || (flag & p.fflag) || mapq < p.min_mapq) return false;
    if (flag & BAM_FPAIRED) {
        if (flag & BAM_FMUNMAP) {
            if (!p.keep_orphan) return false;
        } else {
            if (p.check_orientation &&
                (tid != mtid || ((flag & BAM_FREVERSE) && (flag & BAM_FMREVERSE)) ||
                 (pos < mpos && (flag & BAM_FREVERSE)) || (pos > mpos && !(flag & BAM_FREVERSE))))
                return false;
            if (p.check_proper_pair && !(flag & BAM_FPROPER)) return false;
        }
    }
    return true;
}

void read_bam(const std::string &path, const std::string &contig, int32_t beg0, int32_t end0,
              const gh_bam_params &p, BamData &D) {
    io::InStream in(path);
    char magic[4];
    in.read_exact(magic, 4);
    if (std::memcmp(magic, "BAM\1", 4) != 0) throw io::IoError("not a BAM file: " + path);
    int32_t l_text = 0, n_ref = 0;
    in.read_exact(&l_text, 4);
    if (l_text < 0) throw io::IoError("bad BAM header in " + path);
    std::string text((size_t)l_text, '\0');
    if (l_text) in.read_exact(&text[0], (size_t)l_text);
    in.read_exact(&n_ref, 4);
    int32_t want = -1;
    for (int32_t r = 0; r < n_ref; r++) {
        int32_t l_name = 0, l_ref = 0;
        in.read_exact(&l_name, 4);
        if (l_name <= 0) throw io::IoError("bad BAM reference name in " + path);
        std::string name((size_t)l_name, '\0');
        in.read_exact(&name[0], (size_t)l_name);
        in.read_exact(&l_ref, 4);
        if (std::string(name.c_str()) == contig) want = r;
    }
    if (want < 0) throw io::IoError("contig " + contig + " is not in the header of " + path);
    std::vector<uint8_t> rec;
    for (;;) {
        int32_t bs = 0;
        const size_t n = in.read(&bs, 4);
        if (n == 0) break;
        if (n != 4 || bs < 32) throw io::IoError("truncated BAM record in " + path);
        rec.resize((size_t)bs);
        in.read_exact(rec.data(), rec.size());
        D.n_read++;
        const uint8_t *b = rec.data();
        int32_t tid, pos, l_seq, mtid, mpos;
        std::memcpy(&tid, b, 4);
        std::memcpy(&pos, b + 4, 4);
        const uint8_t l_name = b[8], mapq = b[9];
        uint16_t n_cig, flag;
        std::memcpy(&n_cig, b + 12, 2);
        std::memcpy(&flag, b + 14, 2);
        std::memcpy(&l_seq, b + 16, 4);
        std::memcpy(&mtid, b + 20, 4);
        std::memcpy(&mpos, b + 24, 4);
        const size_t cig_at = 32 + (size_t)l_name, seq_at = cig_at + 4 * (size_t)n_cig;
        const size_t qual_at = seq_at + ((size_t)l_seq + 1) / 2;
        if (l_seq < 0 || qual_at + (size_t)l_seq > rec.size()) throw io::IoError("malformed BAM record in " + path);
        if (tid != want || !keep_read(tid, pos, mapq, flag, mtid, mpos, p)) continue;
        // reference span, then the aligned segments inside the region
        int32_t ref = pos, qpos = 0;
        bool any = false;
        const size_t first = D.segs.size();
        const int64_t base0 = (int64_t)D.bases.size();
        for (int c = 0; c < n_cig; c++) {
            uint32_t op;
            std::memcpy(&op, b + cig_at + 4 * (size_t)c, 4);
            const int32_t len = (int32_t)(op >> 4);
            switch (op & 0xf) {
                case 0: case 7: case 8:   // M = X
                    if (ref + len > beg0 && ref < end0) {
                        D.segs.push_back(BamSeg{ref, len, base0 + qpos, (flag & BAM_FREVERSE) ? 1 : 0, 0});
                        any = true;
                    }
                    ref += len;
                    qpos += len;
                    break;
                case 1: case 4: qpos += len; break;   // I S
                case 2: case 3: ref += len; break;    // D N
                default: break;                       // H P
            }
        }
        if (qpos > l_seq) {
            D.segs.resize(first);
            throw io::IoError("CIGAR longer than the sequence in a record of " + path);
        }
        if (!any) continue;
        D.n_kept++;
        for (int32_t i = 0; i < l_seq; i++) {
            const uint8_t byte = b[seq_at + i / 2];
            D.bases.push_back((i & 1) ? (byte & 0xf) : (byte >> 4));
        }
        D.quals.insert(D.quals.end(), b + qual_at, b + qual_at + l_seq);
    }
}

int base_code(char c) {
    switch (c) {
        case 'A': case 'a': return 0;
        case 'C': case 'c': return 1;
        case 'G': case 'g': return 2;
        case 'T': case 't': return 3;
    }
    return 4;
}

}  // namespace

void bam_default_params(gh_bam_params *p) {
    GH_CHECK(p != nullptr, GH_ERR_ARG, "gh_bam_default_params: null argument");
    p->min_mapq = 10;
    p->min_baseq = 10;
    p->max_depth = 40;
    p->fflag = 0x4 | 0x100 | 0x200 | 0x400;   // unmapped, secondary, QC fail, duplicate
    p->keep_orphan = 0;
    p->check_orientation = 1;
    p->check_proper_pair = 0;
}

void bam_tables_out(float *t) {
    GH_CHECK(t != nullptr, GH_ERR_ARG, "gh_bam_tables: null argument");
    bam_tables(t, t + 256, t + 512, t + 768);
}

void bam_call_gls(const char *path, int32_t ploidy, const gh_bam_region *rg, const gh_bam_params *prm, uint8_t *gl,
                  uint8_t *flat, int32_t *depth, gh_bam_stats *st) {
    GH_CHECK(path && rg && rg->contig && gl && flat && depth, GH_ERR_ARG, "gh_bam_call_gls: null argument");
    GH_CHECK(ploidy == 1 || ploidy == 2, GH_ERR_ARG, "gh_bam_call_gls: ploidy must be 1 or 2");
    const int L = rg->n_sites;
    GH_CHECK(L >= 0 && (L == 0 || (rg->pos && rg->ref && rg->alt)), GH_ERR_ARG, "gh_bam_call_gls: sites");
    gh_bam_params p;
    if (prm) p = *prm;
    else bam_default_params(&p);
    GH_CHECK(p.max_depth > 0 && p.max_depth <= 65535, GH_ERR_ARG, "gh_bam_call_gls: max_depth");
    for (int l = 1; l < L; l++)
        GH_CHECK(rg->pos[l] >= rg->pos[l - 1], GH_ERR_ARG, "gh_bam_call_gls: site positions must ascend");
    const int P1 = ploidy + 1;
    std::memset(gl, 0, (size_t)L * P1);
    std::memset(flat, 1, (size_t)L);
    std::memset(depth, 0, sizeof(int32_t) * L);
    if (st) std::memset(st, 0, sizeof(*st));
    if (L == 0) return;
    // sites: 0-based positions; SNPs inside [beg, end] are called
    std::vector<int32_t> pos0(L);
    std::vector<uint8_t> callable(L, 0), rc(L, 4), ac(L, 4);
    const int32_t beg = rg->beg > 0 ? rg->beg : 1, end = rg->end > 0 ? rg->end : INT32_MAX;
    for (int l = 0; l < L; l++) {
        pos0[l] = rg->pos[l] - 1;
        const bool snp = std::strlen(rg->ref[l]) == 1 && std::strlen(rg->alt[l]) == 1;
        if (!snp || rg->pos[l] < beg || rg->pos[l] > end) continue;
        rc[l] = (uint8_t)base_code(rg->ref[l][0]);
        ac[l] = (uint8_t)base_code(rg->alt[l][0]);
        GH_CHECK(rc[l] < 4 && ac[l] < 4, GH_ERR_ARG, "ERROR base sequences.");   // standard_errmod, :411
        callable[l] = 1;
    }
    BamData D;
    try {
        read_bam(path, rg->contig, beg - 1, end == INT32_MAX ? INT32_MAX : end, p, D);
    } catch (const io::IoError &e) {
        throw Error(GH_ERR_IO, e.what());
    }
    if (st) {
        st->n_records = D.n_read;
        st->n_kept = D.n_kept;
    }
    if (D.segs.empty()) return;
    float tab[1024];
    bam_tables(tab, tab + 256, tab + 512, tab + 768);

    hipStream_t s = nullptr;
    GH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    DBuf d_seg, d_pos, d_call, d_b, d_q, d_cnt, d_off, d_keys, d_rc, d_ac, d_tab, d_gl, d_flat, d_depth;
    const int n_seg = (int)D.segs.size();
    GH_CHECK(D.segs.size() < (size_t)INT32_MAX, GH_ERR_ARG, "gh_bam_call_gls: too many aligned segments");
    const BamSeg *segs = dalloc_copy(d_seg, D.segs.data(), D.segs.size(), s);
    const int32_t *dpos = dalloc_copy(d_pos, pos0.data(), (size_t)L, s);
    const uint8_t *dcall = dalloc_copy(d_call, callable.data(), (size_t)L, s);
    const uint8_t *db = dalloc_copy(d_b, D.bases.data(), D.bases.size(), s);
    const uint8_t *dq = dalloc_copy(d_q, D.quals.data(), D.quals.size(), s);
    d_cnt.reserve(sizeof(int32_t) * L);
    GH_HIP(hipMemsetAsync(d_cnt.p, 0, sizeof(int32_t) * L, s));
    const dim3 gs((n_seg + 255) / 256), bs(256);
    hipLaunchKernelGGL(bam_emit_kernel<false>, gs, bs, 0, s, n_seg, segs, L, dpos, dcall, db, dq, p.min_baseq,
                       d_cnt.as<int32_t>(), (const int64_t *)nullptr, (uint16_t *)nullptr);
    GH_HIP(hipGetLastError());
    std::vector<int32_t> cnt(L);
    GH_HIP(hipMemcpyAsync(cnt.data(), d_cnt.p, sizeof(int32_t) * L, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
    std::vector<int64_t> off(L + 1, 0);
    for (int l = 0; l < L; l++) off[l + 1] = off[l] + cnt[l];
    const int64_t *doff = dalloc_copy(d_off, off.data(), (size_t)L + 1, s);
    d_keys.reserve(sizeof(uint16_t) * (size_t)std::max<int64_t>(off[L], 1));
    GH_HIP(hipMemsetAsync(d_cnt.p, 0, sizeof(int32_t) * L, s));
    hipLaunchKernelGGL(bam_emit_kernel<true>, gs, bs, 0, s, n_seg, segs, L, dpos, dcall, db, dq, p.min_baseq,
                       d_cnt.as<int32_t>(), doff, d_keys.as<uint16_t>());
    GH_HIP(hipGetLastError());
    const uint8_t *drc = dalloc_copy(d_rc, rc.data(), (size_t)L, s), *dac = dalloc_copy(d_ac, ac.data(), (size_t)L, s);
    const float *dtab = dalloc_copy(d_tab, tab, 1024, s);
    d_gl.reserve((size_t)L * P1);
    d_flat.reserve((size_t)L);
    d_depth.reserve(sizeof(int32_t) * L);
    hipLaunchKernelGGL(bam_site_kernel, dim3((L + 255) / 256), dim3(256), 0, s, L, ploidy, p.max_depth, doff,
                       d_cnt.as<int32_t>(), drc, dac, d_keys.as<uint16_t>(), dtab, d_gl.as<uint8_t>(),
                       d_flat.as<uint8_t>(), d_depth.as<int32_t>());
    GH_HIP(hipGetLastError());
    GH_HIP(hipMemcpyAsync(gl, d_gl.p, (size_t)L * P1, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(flat, d_flat.p, (size_t)L, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(depth, d_depth.p, sizeof(int32_t) * L, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
    if (st)
        for (int l = 0; l < L; l++) st->sites_over_max_depth += cnt[l] > p.max_depth;
}

}  // namespace gh
