/*
 * glimpse_oracle.cpp -- CPU oracle: a from-scratch C++ restatement of the
 * GLIMPSE2_phase compute core.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP path and the
 * CPU baseline that bench.py times.  Nothing in the product library links or
 * calls it.  Each routine cites the reference function it follows
 * (paths relative to the GLIMPSE repository root).
 *
 * Arithmetic order: the reference's AVX2 loops keep 8 independent float lanes
 * (state k accumulates into lane k % 8), use one FMA where the reference calls
 * _mm256_fmadd_ps, reduce with horizontal_add = ((l0+l4)+(l1+l5))+((l2+l6)+(l3+l7)),
 * and handle the K % 8 tail with scalar mul-then-add (the reference is built
 * with -std=c++17, so no contraction).  The oracle is compiled with
 * -ffp-contract=off and reproduces that order with explicit 8-lane arrays.
 */
#include "glimpse_oracle.h"

#include <algorithm>
#include <array>
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <iterator>
#include <limits>
#include <map>
#include <memory>
#include <numeric>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <thread>
#include <atomic>
#include <vector>

namespace orc {

static thread_local std::string g_err;

struct OrcError : std::runtime_error {
    explicit OrcError(const std::string &m) : std::runtime_error(m) {}
};

/* ------------------------------------------------------------------ */
/* unphred table: common/src/utils/otools.h:97-98.  The reference ships the
 * literal table 10^(-i/10) printed with 7 significant digits; regenerate it
 * the same way (printf "%.6e" then parse as float). */
static float UNPHRED[256];
static bool unphred_init = [] {
    char buf[64];
    for (int i = 0; i < 256; i++) {
        std::snprintf(buf, sizeof(buf), "%.6e", std::pow(10.0, -i / 10.0));
        UNPHRED[i] = std::strtof(buf, nullptr);
    }
    return true;
}();

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Counter-mode draws. */
static inline void philox(const uint32_t c_in[4], const uint32_t k_in[2], uint32_t out[4]) {
    uint32_t c0 = c_in[0], c1 = c_in[1], c2 = c_in[2], c3 = c_in[3];
    uint32_t k0 = k_in[0], k1 = k_in[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

enum Purpose {
    P_H0 = 0, P_H1 = 1, P_DIP_ROOT = 2, P_DIP_SEG = 3, P_FLAT_HET = 4,
    P_MONO_HET = 5, P_PBWT_STORED = 6, P_PBWT_START = 7, P_INIT_SEL = 8
};

static inline uint32_t cdraw(uint32_t seed, int it, int purpose, uint32_t a, uint32_t b) {
    uint32_t c[4] = {((uint32_t)it << 8) | (uint32_t)purpose, a, b, 0u};
    uint32_t k[2] = {seed, 0u};
    uint32_t o[4];
    philox(c, k, o);
    return o[0];
}
static inline float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
static inline uint32_t urange(uint32_t x, uint32_t lo, uint32_t hi) {
    return lo + (uint32_t)(((uint64_t)x * (uint64_t)(hi - lo + 1u)) >> 32);
}

/* The run's random stream.  Reference mode: common/src/utils/random_number.h:35-125
 * (one global mt19937; getInt via uniform_int_distribution<unsigned>, getFloat via
 * uniform_real_distribution<float>, sample() = linear CDF scan with u <= csum). */
struct Rng {
    int mode = ORC_RNG_REFERENCE;
    uint32_t seed = 15052011u;
    int it = 0;
    std::mt19937 eng;
    std::uniform_int_distribution<unsigned int> di{0, 32768};
    std::uniform_real_distribution<float> df{0.0f, 1.0f};

    void set_seed(uint32_t s) { seed = s; eng.seed(s); }
    unsigned getInt(unsigned lo, unsigned hi, int purpose, uint32_t a, uint32_t b) {
        if (mode == ORC_RNG_REFERENCE)
            return di(eng, std::uniform_int_distribution<unsigned int>{lo, hi}.param());
        return urange(cdraw(seed, it, purpose, a, b), lo, hi);
    }
    /* returns double like random_number_generator::getFloat */
    double getFloat(int purpose, uint32_t a, uint32_t b) {
        if (mode == ORC_RNG_REFERENCE)
            return df(eng, std::uniform_real_distribution<float>{0.0f, 1.0f}.param());
        return (double)u01(cdraw(seed, it, purpose, a, b));
    }
    /* random_number.h:89-97 */
    int sample8(const float *vec, float sum, int purpose, uint32_t a, uint32_t b) {
        float csum = vec[0];
        float u = getFloat(purpose, a, b) * sum;
        for (int i = 0; i < 7; ++i) {
            if (u <= csum) return i;
            csum += vec[i + 1];
        }
        return 7;
    }
};

/* ------------------------------------------------------------------ */
/* bitmatrix: common/src/containers/bitmatrix.h:40-119 (MSB-first bytes,
 * rows and columns padded to multiples of 8). */
struct BitMat {
    uint64_t n_rows = 0, n_cols = 0, n_bytes = 0;
    std::vector<uint8_t> bytes;
    void allocate(unsigned nrow, unsigned ncol) {
        n_rows = nrow + ((nrow % 8) ? (8 - (nrow % 8)) : 0);
        n_cols = ncol + ((ncol % 8) ? (8 - (ncol % 8)) : 0);
        n_bytes = (n_cols / 8) * n_rows;
        bytes.assign(n_bytes, 0);
    }
    void reallocate(unsigned nrow, unsigned ncol) {
        n_rows = nrow + ((nrow % 8) ? (8 - (nrow % 8)) : 0);
        n_cols = ncol + ((ncol % 8) ? (8 - (ncol % 8)) : 0);
        n_bytes = (n_cols / 8) * n_rows;
        if (bytes.size() < n_bytes) bytes.resize(n_bytes, 0);
    }
    inline void set(unsigned row, unsigned col, unsigned char bit) {
        unsigned bitcol = col % 8;
        uint64_t a = (uint64_t)row * (n_cols / 8) + col / 8;
        bytes[a] &= (uint8_t)~(1u << (7 - bitcol));
        bytes[a] |= (uint8_t)(bit << (7 - bitcol));
    }
    inline void setrow(unsigned row, unsigned char bit) {
        std::memset(&bytes[(uint64_t)row * (n_cols / 8)], bit * 255, n_cols / 8);
    }
    inline unsigned char get(unsigned row, unsigned col) const {
        return (bytes[(uint64_t)row * (n_cols >> 3) + (col >> 3)] >> (7 - (col % 8))) & 1;
    }
    inline unsigned char getByte(unsigned row, unsigned col) const {
        return bytes[(uint64_t)row * (n_cols >> 3) + (col >> 3)];
    }
    inline void setByte(unsigned row, unsigned col, unsigned char b) {
        bytes[(uint64_t)row * (n_cols >> 3) + (col >> 3)] = b;
    }
};

/* ------------------------------------------------------------------ */
/* pack3 run-length code: common/src/containers/ref_haplotype_set.h:45-85
 * (R. Durbin's pbwt pack3: bit7 = symbol; 0x40 clear -> n = low 6 bits;
 *  0x40 set, 0x20 clear -> n = (low5) << 6; both set -> n = (low5) << 11). */
static int P3DECODE[128];
static bool p3_init = [] {
    for (int n = 0; n < 64; ++n) P3DECODE[n] = n;
    for (int n = 64; n < 96; ++n) P3DECODE[n] = (n - 64) << 6;
    for (int n = 96; n < 128; ++n) P3DECODE[n] = (n - 96) << 11;
    return true;
}();
static const uint32_t ENC_MAX1 = 64, ENC_MAX2 = (95 - 63) << 6, ENC_MAX3 = (127 - 96) << 11;

static void pack3_add(unsigned char y, uint32_t n, std::vector<uint8_t> &out) {
    y <<= 7;
    while (n >= ENC_MAX3) { out.push_back(y | 0x7f); n -= ENC_MAX3; }
    if (n >= ENC_MAX2) { out.push_back(y | 0x60 | (n >> 11)); n &= 0x7ff; }
    if (n >= ENC_MAX1) { out.push_back(y | 0x40 | (n >> 6)); n &= 0x3f; }
    if (n) out.push_back(y | n);
}

/* ------------------------------------------------------------------ */
/* Panel: the fields of ref_haplotype_set + variant_map the path reads. */
struct Panel {
    int L = 0, N = 0;
    std::vector<double> cm;
    std::vector<float> cm_pos;      /* haplotype_set.cpp:277-279 */
    std::vector<uint32_t> cref, calt;
    std::vector<char> lq, common, major;
    std::vector<int> common2tot;
    int n_com = 0, n_rar = 0, n_com_hq = 0;
    BitMat HvarRef;                  /* n_com x N */
    std::vector<std::vector<int>> ShapRef, SvarRef;
    std::vector<uint8_t> Ypacked;
    std::vector<std::vector<int>> A_small_idx;

    /* build state (ref_haplotype_set.cpp:47-193) */
    std::vector<int> A, B, sA, sB;

    void update_full(int ref_rac_l, int l, std::vector<int> &ref_idx);
    void update_small(int ref_rac_l, const std::vector<bool> &small_haps);
    void build_init_common(int l);
    void init_small_rare(int k, int l, const std::vector<int> &ref_idx,
                         std::vector<bool> &small_hap, std::vector<int> &map_big_small);
    void build_sparse_pbwt();
};

/* ref_haplotype_set.cpp:81-102 */
void Panel::update_full(int ref_rac_l, int l, std::vector<int> &ref_idx) {
    int occ[2] = {0, ref_rac_l};
    int m = 0;
    while (m < N) {
        unsigned char z = HvarRef.get(l, A[m]);
        int m0 = m;
        while (++m < N && HvarRef.get(l, A[m]) == z) {}
        std::copy(A.begin() + m0, A.begin() + m, B.begin() + occ[z]);
        pack3_add(z, m - m0, Ypacked);
        occ[z] += m - m0;
    }
    A.swap(B);
    for (int h = 0; h < N; h++) ref_idx[A[h]] = h;
}

/* ref_haplotype_set.cpp:104-124 */
void Panel::update_small(int ref_rac_l, const std::vector<bool> &small_haps) {
    int occ[2] = {0, ref_rac_l};
    int m = 0;
    const int size_a = (int)sA.size();
    while (m < size_a) {
        unsigned char z = small_haps[sA[m]];
        int m0 = m;
        while (++m < size_a && small_haps[sA[m]] == z) {}
        std::copy(sA.begin() + m0, sA.begin() + m, sB.begin() + occ[z]);
        pack3_add(z, m - m0, Ypacked);
        occ[z] += m - m0;
    }
    sA.swap(sB);
}

/* ref_haplotype_set.cpp:126-140 */
void Panel::build_init_common(int l) {
    std::vector<bool> big(N, true);
    const std::vector<int> &small_idx = A_small_idx[l];
    for (size_t h = 0; h < small_idx.size(); ++h) big[small_idx[h]] = false;
    int nz = 0;
    for (int h = 0; h < N; ++h) if (big[h]) B[nz++] = A[h];
    for (size_t h = 0; h < sA.size(); ++h, ++nz) B[nz] = A[small_idx[sA[h]]];
    A.swap(B);
}

/* ref_haplotype_set.cpp:142-169 */
void Panel::init_small_rare(int k, int l, const std::vector<int> &ref_idx,
                            std::vector<bool> &small_hap, std::vector<int> &map_big_small) {
    std::vector<int> rare_haps;
    std::map<int, int> pos2hap;
    if (l > 0) for (int h : A_small_idx[l - 1]) map_big_small[h] = -1;
    for (int kk = k; kk < L; ++kk) {
        if (lq[kk]) continue;
        if (common[kk]) break;
        rare_haps.insert(rare_haps.end(), SvarRef[kk].begin(), SvarRef[kk].end());
    }
    std::sort(rare_haps.begin(), rare_haps.end());
    rare_haps.erase(std::unique(rare_haps.begin(), rare_haps.end()), rare_haps.end());
    for (int h : rare_haps) pos2hap.insert(std::make_pair(ref_idx[h], h));
    A_small_idx[l] = std::vector<int>(pos2hap.size());
    int kk2 = 0;
    for (auto it = pos2hap.begin(); it != pos2hap.end(); ++it, ++kk2) {
        map_big_small[it->second] = kk2;
        A_small_idx[l][kk2] = it->first;
    }
    sA.resize(A_small_idx[l].size());
    std::iota(sA.begin(), sA.end(), 0);
    sB.resize(A_small_idx[l].size());
    small_hap.resize(A_small_idx[l].size());
}

/* ref_haplotype_set.cpp:47-79 */
void Panel::build_sparse_pbwt() {
    std::vector<bool> small_hap(N, false);
    std::vector<int> map_big_small(N, -1);
    std::vector<int> ref_idx(N);
    A.assign(N, 0); B.assign(N, 0);
    std::iota(A.begin(), A.end(), 0);
    std::iota(ref_idx.begin(), ref_idx.end(), 0);
    sA.clear(); sB.clear();
    Ypacked.clear();
    A_small_idx.assign(n_com_hq + 1, std::vector<int>());
    int l_hq = 0, last_k = -1, l_all = 0;
    for (int k = 0; k < L; k++) {
        if (lq[k]) { l_all += common[k]; continue; }
        if (common[k]) {
            if (sA.size() > 0 && !(last_k >= 0 && common[last_k])) build_init_common(l_hq);
            update_full((int)cref[k], l_all, ref_idx);
            ++l_hq; ++l_all;
        } else {
            if (last_k < 0 || common[last_k]) init_small_rare(k, l_hq, ref_idx, small_hap, map_big_small);
            std::fill(small_hap.begin(), small_hap.end(), (bool)major[k]);
            for (int h : SvarRef[k]) small_hap[map_big_small[h]] = !major[k];
            int ref_rac_l = major[k] ? N - (int)calt[k] : (int)sA.size() - (int)calt[k];
            update_small(ref_rac_l, small_hap);
        }
        last_k = k;
    }
}

/* Site classification: common/src/io/ref_genotype_reader.cpp:165-180
 * (MAF computed in float, compared as double against sparse_maf). */
static void classify(uint32_t cref, uint32_t calt, float sparse_maf, bool &is_common, bool &maj) {
    double MAF = std::min(cref * 1.0f / (cref + calt), calt * 1.0f / (cref + calt));
    is_common = MAF >= sparse_maf;
    maj = calt > cref;
}

static Panel *make_panel(const orc_panel_desc *d, float sparse_maf) {
    Panel *P = new Panel();
    P->L = d->n_sites; P->N = d->n_haps;
    const int L = P->L, N = P->N;
    P->cm.assign(d->cm, d->cm + L);
    P->cref.assign(d->cref, d->cref + L);
    P->calt.assign(d->calt, d->calt + L);
    P->lq.resize(L); P->common.resize(L); P->major.resize(L);
    P->cm_pos.resize(L);
    for (int l = 0; l < L; l++) {
        bool c, m;
        classify(P->cref[l], P->calt[l], sparse_maf, c, m);
        P->common[l] = c; P->major[l] = m;
        P->lq[l] = d->lq ? d->lq[l] : 0;
        if (c) P->common2tot.push_back(l);
        P->n_com += c; P->n_rar += !c;
        P->n_com_hq += c && !P->lq[l];
        P->cm_pos[l] = std::max(0.0f, (float)P->cm[l]);
    }
    const int W = (N + 63) / 64;
    P->HvarRef.allocate(P->n_com, N);
    for (int lc = 0; lc < P->n_com; lc++) {
        const uint64_t *row = d->common_words + (size_t)lc * W;
        uint32_t ones = 0;
        for (int h = 0; h < N; h++) {
            unsigned a = (row[h >> 6] >> (h & 63)) & 1ull;
            if (a) { P->HvarRef.set(lc, h, 1); ones++; }
        }
        int l = P->common2tot[lc];
        if (ones != P->calt[l] || (uint32_t)N - ones != P->cref[l])
            throw OrcError("panel: AC/AN inconsistent with common-site bits at site " + std::to_string(l));
    }
    P->ShapRef.assign(N, {});
    P->SvarRef.assign(L, {});
    for (int l = 0; l < L; l++) {
        if (P->common[l]) {
            if (d->rare_ptr[l + 1] != d->rare_ptr[l]) throw OrcError("panel: carriers listed at a common site");
            continue;
        }
        uint32_t nmin = P->major[l] ? P->cref[l] : P->calt[l];
        if ((uint32_t)(d->rare_ptr[l + 1] - d->rare_ptr[l]) != nmin)
            throw OrcError("panel: AC/AN inconsistent with rare carriers at site " + std::to_string(l));
        for (int j = d->rare_ptr[l]; j < d->rare_ptr[l + 1]; j++) {
            int h = d->rare_idx[j];
            P->SvarRef[l].push_back(h);
        }
    }
    /* haplotype_set::transposeRareRef order: ShapRef[h] sites ascending; SvarRef[l] haps ascending */
    for (int l = 0; l < L; l++) {
        std::sort(P->SvarRef[l].begin(), P->SvarRef[l].end());
        for (int h : P->SvarRef[l]) P->ShapRef[h].push_back(l);
    }
    P->build_sparse_pbwt();
    return P;
}

/* ------------------------------------------------------------------ */
/* Target-side sparse-PBWT sweep: phase/src/containers/haplotype_set.cpp:317-728.
 * Operates on a list of target haplotypes (alleles given per site). */
struct Sweep {
    const Panel &P;
    int n_tar_haps;
    int depth;
    float modulo;
    const std::vector<std::vector<uint8_t>> &hap;   /* [htr][L] alleles */
    std::vector<int> hap2sample;                      /* tar_hapid2ind */
    int n_samples;
    std::vector<int> A, B, V, sA, sB, sV;
    std::vector<int> idx, f, g, sidx, sf, sg, last_reset, last_rare;
    std::vector<unsigned char> tar_hap;
    std::vector<int> rareTar;
    std::vector<std::vector<int>> SvarTar;
    /* trace for pinning tests: after every HQ common site, per target haplotype
     * (site, hap, idx, f, g, site of the last reset of [f, g) or -1) */
    std::vector<int32_t> *trace = nullptr;
    std::vector<int> creset;
    std::vector<bool> stored;
    std::vector<std::vector<std::vector<int>>> states;  /* [sample][depth] */

    Sweep(const Panel &p, const std::vector<std::vector<uint8_t>> &h, int ns, int d, float mod)
        : P(p), n_tar_haps((int)h.size()), depth(d), modulo(mod), hap(h), n_samples(ns) {
        hap2sample.resize(n_tar_haps);
        for (int i = 0; i < n_tar_haps; i++) hap2sample[i] = i / 2;
    }

    void read_full(const uint8_t *&pY, int ref_rac_l);
    void read_small(const uint8_t *&pY, int ref_rac_l, bool update_v);
    void select_common(int k, int l_hq, int l_all, int ref_rac_l, int prev_ref_rac_l);
    void select_rare(int k, int ref_rac_l);
    void init_common(int k, int l, int prev_ref_rac_l);
    void init_rare(int k, int l);
    void selectK(int htr, int ref_rac_l, const std::vector<int> &arr, unsigned char a);
    void run(const std::vector<bool> &stored_flags, const std::vector<int> &start);
};

/* haplotype_set.cpp:396-419 */
void Sweep::read_full(const uint8_t *&pY, int ref_rac_l) {
    int occ[2] = {0, ref_rac_l};
    int m = 0;
    const int size_a = P.N;
    V[0] = 0;
    while (m < size_a) {
        unsigned char z = *pY++;
        int n = P3DECODE[z & 0x7f];
        int mm = m + n;
        z >>= 7;
        std::copy(A.begin() + m, A.begin() + mm, B.begin() + occ[z]);
        if (z) std::iota(V.begin() + m + 1, V.begin() + mm + 1, occ[1] - ref_rac_l + 1);
        else std::fill(V.begin() + m + 1, V.begin() + mm + 1, occ[1] - ref_rac_l);
        m = mm;
        occ[z] += n;
    }
    B.swap(A);
}

/* haplotype_set.cpp:421-444 */
void Sweep::read_small(const uint8_t *&pY, int ref_rac_l, bool update_v) {
    int occ[2] = {0, ref_rac_l};
    int m = 0;
    const int size_a = (int)sA.size();
    sV[0] = 0;
    while (m < size_a) {
        unsigned char z = *pY++;
        int n = P3DECODE[z & 0x7f];
        z >>= 7;
        int mm = m + n;
        std::copy(sA.begin() + m, sA.begin() + mm, sB.begin() + occ[z]);
        if (update_v) {
            if (z) std::iota(sV.begin() + m + 1, sV.begin() + mm + 1, occ[1] - ref_rac_l + 1);
            else std::fill(sV.begin() + m + 1, sV.begin() + mm + 1, occ[1] - ref_rac_l);
        }
        m = mm;
        occ[z] += n;
    }
    sB.swap(sA);
}

/* haplotype_set.cpp:682-728 (k0 = pbwt_depth) */
void Sweep::selectK(int htr, int ref_rac_l, const std::vector<int> &arr, unsigned char a) {
    const int sample = hap2sample[htr];
    const int pidx = idx[htr];
    int nh_up = 0, nh_down = 0;
    int d_up = idx[htr] - f[htr];
    int d_down = g[htr] - idx[htr];
    const int k1 = depth;
    if (d_up < k1 && d_down < k1) { nh_up = d_up; nh_down = d_down; }
    else if (d_up >= k1 && d_down >= k1) { nh_up = k1; nh_down = k1; }
    else if (d_up < k1) nh_down += k1 - d_up;
    else if (d_down < k1) nh_up += k1 - d_down;
    const int max_hh = std::max(nh_up, nh_down);
    for (int o = 0, od = 0, ou = 0; o < max_hh; ++o) {
        if (od < nh_down) {
            const int i2 = pidx + od;
            if (i2 < P.N && ((i2 >= ref_rac_l) == (bool)a)) { states[sample][o].push_back(arr[i2]); ++od; }
        }
        if (ou < nh_up) {
            const int i2 = pidx - (ou + 1);
            if (i2 >= 0 && ((i2 >= ref_rac_l) == (bool)a)) { states[sample][o].push_back(arr[i2]); ++ou; }
        }
    }
}

/* haplotype_set.cpp:445-490 */
void Sweep::select_common(int k, int l_hq, int l_all, int ref_rac_l, int prev_ref_rac_l) {
    (void)l_all;
    for (int htr = 0; htr < n_tar_haps; htr++) {
        bool reset = false;
        const unsigned char prev = tar_hap[htr];
        const unsigned char a = hap[htr][k];
        tar_hap[htr] = a;
        int id = a * (ref_rac_l + V[idx[htr]]) + (1 - a) * (idx[htr] - V[idx[htr]]);
        int fd = a * (ref_rac_l + V[f[htr]]) + (1 - a) * (f[htr] - V[f[htr]]);
        int gd = a * (ref_rac_l + V[g[htr]]) + (1 - a) * (g[htr] - V[g[htr]]);
        if (gd <= fd) {
            if (stored[l_hq] || P.cm_pos[k - 1] - P.cm_pos[last_reset[htr]] > modulo / 2.0f)
                selectK(htr, prev_ref_rac_l, B, prev);
            fd = a * ref_rac_l;
            gd = a * P.N + (1 - a) * ref_rac_l;
            last_reset[htr] = k;
            reset = true;
            if (trace) creset[htr] = k;
        }
        idx[htr] = id; f[htr] = fd; g[htr] = gd;
        if (!reset && stored[l_hq]) {
            selectK(htr, ref_rac_l, A, a);
            f[htr] = a * ref_rac_l;
            g[htr] = a * P.N + (1 - a) * ref_rac_l;
            last_reset[htr] = k;
            if (trace) creset[htr] = k;
        }
        if (trace) {
            const int32_t row[6] = {k, htr, idx[htr], f[htr], g[htr], creset[htr]};
            trace->insert(trace->end(), row, row + 6);
        }
    }
}

/* haplotype_set.cpp:492-560 */
void Sweep::select_rare(int k, int ref_rac_l) {
    if (rareTar.empty()) return;
    const int n_small = (int)sA.size();
    const int n_rare = (int)rareTar.size();
    int id_rare = 0;
    for (int htr : SvarTar[k]) {
        last_rare[htr] = k;
        const int si = ref_rac_l + sV[sidx[htr]];
        int fd = ref_rac_l + sV[sf[htr]];
        int gd = ref_rac_l + sV[sg[htr]];
        if (gd <= fd) { fd = ref_rac_l; gd = n_small; last_reset[htr] = k; }
        sidx[htr] = si; sf[htr] = fd; sg[htr] = gd;
        while (rareTar[id_rare] < htr) {
            const int h2 = rareTar[id_rare];
            const int si2 = sidx[h2] - sV[sidx[h2]];
            int fd2 = sf[h2] - sV[sf[h2]];
            int gd2 = sg[h2] - sV[sg[h2]];
            if (gd2 <= fd2) { fd2 = 0; gd2 = ref_rac_l; last_reset[h2] = k; }
            sidx[h2] = si2; sf[h2] = fd2; sg[h2] = gd2;
            ++id_rare;
        }
        ++id_rare;
    }
    while (id_rare < n_rare) {
        const int h2 = rareTar[id_rare];
        const int si2 = sidx[h2] - sV[sidx[h2]];
        int fd2 = sf[h2] - sV[sf[h2]];
        int gd2 = sg[h2] - sV[sg[h2]];
        if (gd2 <= fd2) { fd2 = 0; gd2 = ref_rac_l; last_reset[h2] = k; }
        sidx[h2] = si2; sf[h2] = fd2; sg[h2] = gd2;
        ++id_rare;
    }
}

/* haplotype_set.cpp:562-630.  common2tot[l-1] is indexed by the HQ-common
 * index l exactly as in the reference. */
void Sweep::init_common(int k, int l, int prev_ref_rac_l) {
    const std::vector<int> &small_idx = P.A_small_idx[l];
    if (small_idx.size() <= 0) return;
    const int ns = (int)small_idx.size();
    const int prev_site = (l >= 1) ? P.common2tot[l - 1] : 0;
    size_t j = 0;
    for (int htr = 0; htr < n_tar_haps; ++htr) {
        int fd, gd, id;
        if (j < rareTar.size() && htr == rareTar[j]) {
            id = P.N - ns + sidx[htr];
            fd = last_reset[htr] >= last_rare[htr] ? 0 : P.N - ns + sf[htr];
            gd = P.N - ns + sg[htr];
            if (gd <= fd) {
                selectK(htr, prev_ref_rac_l, A, tar_hap[htr]);
                fd = 0; gd = P.N;
                last_reset[htr] = k - 1;
            }
            ++j;
        } else {
            int c0 = (int)(std::lower_bound(small_idx.begin(), small_idx.end(), f[htr]) - small_idx.begin());
            int c1 = (int)(std::lower_bound(small_idx.begin(), small_idx.end(), g[htr]) - small_idx.begin());
            int c2 = (int)(std::lower_bound(small_idx.begin(), small_idx.end(), idx[htr]) - small_idx.begin());
            id = idx[htr] - c2;
            fd = f[htr] - c0;
            gd = g[htr] - c1;
            if (gd <= fd) {
                if (l < 1) throw OrcError("sweep: common2tot[-1] read (rare run before the first common site)");
                if (P.cm_pos[prev_site] - P.cm_pos[last_reset[htr]] > modulo / 5.0f)
                    selectK(htr, prev_ref_rac_l, A, tar_hap[htr]);
                fd = 0;
                gd = P.N - ns;
                last_reset[htr] = prev_site;
            }
        }
        idx[htr] = id; f[htr] = fd; g[htr] = gd;
    }
    std::vector<bool> big(P.N, true);
    for (int h : small_idx) big[h] = false;
    int nz = small_idx[0];
    for (int h = nz + 1; h < P.N; ++h) if (big[h]) A[nz++] = A[h];
    for (size_t h = 0; h < sA.size(); ++h, ++nz) A[nz] = sA[h];
}

/* haplotype_set.cpp:632-680 */
void Sweep::init_rare(int k, int l) {
    const std::vector<int> &small_idx = P.A_small_idx[l];
    const int ns = (int)small_idx.size();
    sA.resize(ns);
    for (int i = 0; i < ns; ++i) sA[i] = A[small_idx[i]];
    sB.resize(ns);
    sV.resize(ns + 1);
    rareTar.clear();
    if (ns <= 0) return;
    for (int kk = k; kk < P.L; ++kk) {
        if (P.lq[kk]) continue;
        if (P.common[kk]) break;
        rareTar.insert(rareTar.end(), SvarTar[kk].begin(), SvarTar[kk].end());
    }
    std::sort(rareTar.begin(), rareTar.end());
    rareTar.erase(std::unique(rareTar.begin(), rareTar.end()), rareTar.end());
    for (int htr : rareTar) {
        int c0 = (int)(std::lower_bound(small_idx.begin(), small_idx.end(), f[htr]) - small_idx.begin());
        int c1 = (int)(std::lower_bound(small_idx.begin(), small_idx.end(), g[htr]) - small_idx.begin());
        int c2 = (int)(std::lower_bound(small_idx.begin(), small_idx.end(), idx[htr]) - small_idx.begin());
        int fd = c0, gd = c1;
        if (gd <= fd) { fd = 0; gd = ns; last_reset[htr] = k; }
        sidx[htr] = c2; sf[htr] = fd; sg[htr] = gd;
    }
}

/* haplotype_set.cpp:317-394 (stored flags and start indices supplied). */
void Sweep::run(const std::vector<bool> &stored_flags, const std::vector<int> &start) {
    const int N = P.N;
    A.assign(N, 0); B.assign(N, 0); V.assign(N + 1, 0);
    std::iota(A.begin(), A.end(), 0);
    sA.clear(); sB.clear(); sV.clear();
    idx.assign(n_tar_haps, 0); f.assign(n_tar_haps, 0); g.assign(n_tar_haps, N);
    sidx.assign(n_tar_haps, 0); sf.assign(n_tar_haps, 0); sg.assign(n_tar_haps, 0);
    last_reset.assign(n_tar_haps, 0); last_rare.assign(n_tar_haps, 0);
    tar_hap.assign(n_tar_haps, 0);
    stored = stored_flags;
    for (int e = 0; e < n_tar_haps; e++) { f[e] = 0; g[e] = N; idx[e] = start[e]; }
    creset.assign(n_tar_haps, -1);
    states.assign(n_samples, std::vector<std::vector<int>>(depth));
    /* transposeRareTar: target haps carrying the non-major allele at rare sites */
    SvarTar.assign(P.L, {});
    for (int h = 0; h < n_tar_haps; h++)
        for (int l = 0; l < P.L; l++)
            if (!P.common[l] && hap[h][l] != (uint8_t)P.major[l]) SvarTar[l].push_back(h);

    const uint8_t *pY = P.Ypacked.data();
    int ref_rac_com = 0, prev_ref_rac_com = 0;
    int l_hq = 0, last_k = -1, l_all = 0;
    for (int k = 0; k < P.L; k++) {
        if (P.lq[k]) { l_all += P.common[k]; continue; }
        if (P.common[k]) {
            if (P.A_small_idx[l_hq].size() > 0 && !(last_k >= 0 && P.common[last_k]))
                init_common(k, l_hq, prev_ref_rac_com);
            ref_rac_com = (int)P.cref[k];
            read_full(pY, ref_rac_com);
            select_common(k, l_hq, l_all, ref_rac_com, prev_ref_rac_com);
            prev_ref_rac_com = ref_rac_com;
            ++l_all; ++l_hq;
        } else {
            if (last_k < 0 || P.common[last_k]) init_rare(k, l_hq);
            int ref_rac_rare = P.major[k] ? (int)P.cref[k] : (int)P.A_small_idx[l_hq].size() - (int)P.calt[k];
            read_small(pY, ref_rac_rare, rareTar.size() != 0);
            select_rare(k, ref_rac_rare);
        }
        last_k = k;
    }
    if (pY != P.Ypacked.data() + P.Ypacked.size()) throw OrcError("sweep: pack3 stream not fully consumed");
}

/* haplotype_set.cpp:229-303: group boundaries and modulo auto-adjust. */
static void pbwt_groups(const Panel &P, int depth, float modulo_in, int kpbwt,
                        float &modulo, std::vector<int> &grp) {
    modulo = modulo_in;
    float length = P.cm_pos.back() - P.cm_pos[0];
    if ((length / modulo) * 2 * depth < kpbwt)
        while ((length / modulo) * 2 * depth < kpbwt && modulo > 0.02) modulo /= 2;
    grp.clear();
    for (int l_hq = 0, l_all = 0, src = 0; l_all < P.n_com; l_all++) {
        const int k = P.common2tot[l_all];
        if (P.lq[k]) continue;
        const int tmp = (int)std::round(P.cm_pos[k] / modulo);
        if (src != tmp) { src = tmp; if (l_hq > 0) grp.push_back(l_hq); }
        ++l_hq;
    }
    if (grp.empty() || grp.back() < P.n_com_hq) grp.push_back(P.n_com_hq);
}

/* ------------------------------------------------------------------ */
/* Conditioning set: phase/src/containers/conditioning_set.cpp:75-196. */
enum { TYPE_COMMON = 0, TYPE_RARE = 1, TYPE_MONO = 2 };

struct CondSet {
    const Panel &P;
    double nrho;
    const double one_l = 1.0 - 1e-7;
    float ed_phs, ee_phs, ed_imp, ee_imp;
    unsigned n_states = 0;
    std::vector<int> idx;
    std::vector<std::vector<unsigned>> Svar;
    std::vector<unsigned char> var_type;
    std::vector<unsigned> poly, mono;
    BitMat Hvar;
    std::vector<float> t, nt;

    CondSet(const Panel &p, int ne, float err_imp, float err_phs)
        : P(p), nrho((-0.04 * ne) / p.N), ed_phs(err_phs), ee_phs((float)(1.0 - err_phs)),
          ed_imp(err_imp), ee_imp((float)(1.0 - err_imp)) {
        Svar.resize(p.L);
        var_type.resize(p.L);
    }
    /* conditioning_set.h:101-105 */
    float getTransition(int prev_abs, int next_abs) const {
        return (float)std::clamp(-std::expm1(nrho * (P.cm[next_abs] - P.cm[prev_abs])), 1e-7, one_l);
    }
    void build(const std::vector<int> &states);
    void transitions();
};

void CondSet::build(const std::vector<int> &states) {
    idx = states;
    n_states = (unsigned)idx.size();
    if (n_states == 0) throw OrcError("condset: zero states");
    for (int l = 0; l < P.L; l++) Svar[l].clear();
    for (unsigned k = 0; k < n_states; k++)
        for (int r : P.ShapRef[idx[k]]) Svar[r].push_back(k);
    poly.clear(); mono.clear();
    for (int l = 0; l < P.L; l++) {
        if (P.common[l]) var_type[l] = TYPE_COMMON;
        else if (!Svar[l].empty()) var_type[l] = TYPE_RARE;
        else var_type[l] = TYPE_MONO;
        if (var_type[l] == TYPE_MONO) mono.push_back(l); else poly.push_back(l);
    }
    Hvar.allocate((unsigned)poly.size(), n_states);
    const int nfull = (n_states / 8) * 8;
    for (int labs = 0, lrel = 0, lcom = 0; labs < P.L; labs++) {
        if (var_type[labs] == TYPE_COMMON) {
            for (int k = 0; k < nfull; k += 8) {
                unsigned char b = 0;
                for (int i = 0; i < 8; i++) b |= (unsigned char)(P.HvarRef.get(lcom, idx[k + i]) << (7 - i));
                Hvar.setByte(lrel, k, b);
            }
            for (unsigned k = nfull; k < n_states; k++) Hvar.set(lrel, k, P.HvarRef.get(lcom, idx[k]));
            lrel++; lcom++;
        } else if (var_type[labs] == TYPE_RARE) {
            Hvar.setrow(lrel, P.major[labs]);
            for (unsigned r : Svar[labs]) Hvar.set(lrel, r, !P.major[labs]);
            lrel++;
        }
    }
    transitions();
}

void CondSet::transitions() {
    t.clear(); nt.clear();
    if (poly.empty()) return;
    t.resize(poly.size() - 1);
    nt.resize(poly.size() - 1);
    for (size_t l = 1; l < poly.size(); l++) {
        t[l - 1] = (float)std::clamp(-std::expm1(nrho * (P.cm[poly[l]] - P.cm[poly[l - 1]])), 1e-7, one_l);
        nt[l - 1] = 1.0f - t[l - 1];
    }
}

/* conditioning_set::compactSelection, Kpbwt branch (conditioning_set.cpp:96-104) */
static std::vector<int> compact_states(const std::vector<std::vector<int>> &lists, int kpbwt) {
    std::set<int> s;
    for (size_t i = 0; i < lists.size() && (int)s.size() < kpbwt; ++i)
        for (int j = (int)lists[i].size() - 1; j >= 0 && (int)s.size() < kpbwt; --j)
            s.insert(lists[i][j]);
    return std::vector<int>(s.begin(), s.end());
}

/* ------------------------------------------------------------------ */
/* 8-lane float vector with the AVX2 semantics the reference relies on. */
struct V8 {
    float v[8];
    static V8 set1(float x) { V8 r; for (int i = 0; i < 8; i++) r.v[i] = x; return r; }
};
static inline V8 add8(const V8 &a, const V8 &b) { V8 r; for (int i = 0; i < 8; i++) r.v[i] = a.v[i] + b.v[i]; return r; }
static inline V8 mul8(const V8 &a, const V8 &b) { V8 r; for (int i = 0; i < 8; i++) r.v[i] = a.v[i] * b.v[i]; return r; }
static inline V8 fma8(const V8 &a, const V8 &b, const V8 &c) { V8 r; for (int i = 0; i < 8; i++) r.v[i] = std::fma(a.v[i], b.v[i], c.v[i]); return r; }
static inline V8 load8(const float *p) { V8 r; std::memcpy(r.v, p, 32); return r; }
static inline void store8(float *p, const V8 &a) { std::memcpy(p, a.v, 32); }
/* blendv(a, b, mask from byte): lane i takes b when bit (7-i) of byte is set */
static inline V8 blend8(const V8 &a, const V8 &b, unsigned char byte) {
    V8 r; for (int i = 0; i < 8; i++) r.v[i] = ((byte >> (7 - i)) & 1) ? b.v[i] : a.v[i]; return r;
}
/* horizontal_add: imputation_hmm.cpp:28-39 / phasing_hmm.h:47-57 */
static inline float hadd8(const V8 &a) {
    float s0 = a.v[0] + a.v[4], s1 = a.v[1] + a.v[5], s2 = a.v[2] + a.v[6], s3 = a.v[3] + a.v[7];
    return (s0 + s1) + (s2 + s3);
}

/* ------------------------------------------------------------------ */
/* Haploid imputation HMM: phase/src/models/imputation_hmm.cpp:61-367. */
struct ImputeHMM {
    CondSet *C;
    unsigned modK = 0;
    std::vector<float> Em, Alpha, AlphaSum, Beta;
    explicit ImputeHMM(CondSet *c) : C(c) { Em.resize(2 * c->P.L); }

    void run(const std::vector<float> &HL, const std::vector<char> &flat, std::vector<float> &HP) {
        const unsigned K = C->n_states;
        modK = ((K / 8) + (K % 8 ? 1 : 0)) * 8;
        AlphaSum.assign(C->poly.size(), 0.0f);
        Alpha.assign(C->poly.size() * modK, 0.0f);
        Beta.assign(modK, 0.0f);
        for (int l = 0; l < C->P.L; l++) {
            float p0 = HL[2 * l + 0] * C->ee_imp + HL[2 * l + 1] * C->ed_imp;
            float p1 = HL[2 * l + 0] * C->ed_imp + HL[2 * l + 1] * C->ee_imp;
            Em[2 * l + 0] = p0 / (p0 + p1);
            Em[2 * l + 1] = p1 / (p0 + p1);
        }
        forward(flat);
        backward(HL, flat, HP);
    }

    void forward(const std::vector<char> &flat) {
        const unsigned K = C->n_states, K8 = (K / 8) * 8;
        const std::vector<unsigned> &poly = C->poly;
        for (size_t l = 0; l < poly.size(); l++) {
            AlphaSum[l] = 0.0f;
            const unsigned site = poly[l];
            float *cur = &Alpha[l * modK];
            if (flat[site] || C->P.lq[site]) {
                if (l == 0) {
                    std::fill(cur, cur + modK, 1.0f / K);
                    AlphaSum[l] = 1.0f;
                } else {
                    const float *prv = &Alpha[(l - 1) * modK];
                    const float fact1 = C->t[l - 1] / K;
                    const float fact2 = C->nt[l - 1] / AlphaSum[l - 1];
                    V8 F1 = V8::set1(fact1), F2 = V8::set1(fact2), S = V8::set1(0.0f);
                    unsigned k = 0;
                    for (; k < K8; k += 8) {
                        V8 pc = fma8(load8(prv + k), F2, F1);
                        S = add8(S, pc);
                        store8(cur + k, pc);
                    }
                    if (k) AlphaSum[l] = hadd8(S);
                    for (unsigned o = K8; o < K; o++) {
                        float m = prv[o] * fact2; cur[o] = m + fact1;
                        AlphaSum[l] += cur[o];
                    }
                }
            } else {
                const float e0 = Em[2 * site], e1 = Em[2 * site + 1];
                const float emit[2] = {e0, e1};
                V8 E0 = V8::set1(e0), E1 = V8::set1(e1);
                if (l == 0) {
                    const float fact1 = 1.0f / K;
                    V8 F1 = V8::set1(fact1), S = V8::set1(0.0f);
                    unsigned k = 0;
                    for (; k < K8; k += 8) {
                        V8 em = blend8(E0, E1, C->Hvar.getByte(l, k));
                        V8 pc = mul8(em, F1);
                        S = add8(S, pc);
                        store8(cur + k, pc);
                    }
                    if (k) AlphaSum[l] = hadd8(S);
                    for (unsigned o = K8; o < K; o++) {
                        cur[o] = emit[C->Hvar.get(l, o)] * fact1;
                        AlphaSum[l] += cur[o];
                    }
                } else {
                    const float *prv = &Alpha[(l - 1) * modK];
                    const float fact1 = C->t[l - 1] / K;
                    const float fact2 = C->nt[l - 1] / AlphaSum[l - 1];
                    V8 F1 = V8::set1(fact1), F2 = V8::set1(fact2), S = V8::set1(0.0f);
                    unsigned k = 0;
                    for (; k < K8; k += 8) {
                        V8 em = blend8(E0, E1, C->Hvar.getByte(l, k));
                        V8 tmp = fma8(load8(prv + k), F2, F1);
                        V8 pc = mul8(tmp, em);
                        S = add8(S, pc);
                        store8(cur + k, pc);
                    }
                    if (k) AlphaSum[l] = hadd8(S);
                    for (unsigned o = K8; o < K; o++) {
                        float m = prv[o] * fact2; float a = m + fact1;
                        cur[o] = a * emit[C->Hvar.get(l, o)];
                        AlphaSum[l] += cur[o];
                    }
                }
            }
        }
    }

    void backward(const std::vector<float> &HL, const std::vector<char> &flat, std::vector<float> &HP) {
        const unsigned K = C->n_states, K8 = (K / 8) * 8;
        const std::vector<unsigned> &poly = C->poly;
        const float ee = C->ee_imp, ed = C->ed_imp;
        float betaSum = 0.0f, betaSumNext = 0.0f;
        float ph[2], po[2];
        const V8 ZERO = V8::set1(0.0f), ONE = V8::set1(1.0f);
        std::fill(Beta.begin(), Beta.end(), 1.0f);
        const int Lp = (int)poly.size();
        for (int l = Lp - 1; l >= 0; l--) {
            betaSum = 0.0f; ph[0] = 0.0f; ph[1] = 0.0f;
            V8 S = V8::set1(0.0f), P0 = V8::set1(0.0f), P1 = V8::set1(0.0f);
            const unsigned site = poly[l];
            const float *al = &Alpha[(size_t)l * modK];
            if (flat[site] || C->P.lq[site]) {
                if (l == Lp - 1) {
                    const float fact1 = 1.0f / K;
                    V8 F1 = V8::set1(fact1);
                    unsigned k = 0;
                    for (; k < K8; k += 8) {
                        unsigned char byte = C->Hvar.getByte(l, k);
                        V8 m0 = blend8(ONE, ZERO, byte), m1 = blend8(ZERO, ONE, byte);
                        V8 a = load8(al + k);
                        P0 = add8(P0, mul8(a, m0));
                        P1 = add8(P1, mul8(a, m1));
                        store8(&Beta[k], F1);
                    }
                    if (k) { ph[0] = hadd8(P0); ph[1] = hadd8(P1); }
                    for (unsigned o = K8; o < K; o++) {
                        Beta[o] = fact1;
                        ph[C->Hvar.get(l, o)] += al[o];
                    }
                    betaSum = 1.0f;
                } else {
                    const float fact1 = C->t[l] / K;
                    const float fact2 = C->nt[l] / betaSumNext;
                    V8 F1 = V8::set1(fact1), F2 = V8::set1(fact2);
                    unsigned k = 0;
                    for (; k < K8; k += 8) {
                        unsigned char byte = C->Hvar.getByte(l, k);
                        V8 m0 = blend8(ONE, ZERO, byte), m1 = blend8(ZERO, ONE, byte);
                        V8 pc = fma8(load8(&Beta[k]), F2, F1);
                        V8 dp = mul8(load8(al + k), pc);
                        P0 = add8(P0, mul8(dp, m0));
                        P1 = add8(P1, mul8(dp, m1));
                        S = add8(S, pc);
                        store8(&Beta[k], pc);
                    }
                    if (k) { ph[0] = hadd8(P0); ph[1] = hadd8(P1); betaSum = hadd8(S); }
                    for (unsigned o = K8; o < K; o++) {
                        float m = Beta[o] * fact2; Beta[o] = m + fact1;
                        float q = al[o] * Beta[o];
                        ph[C->Hvar.get(l, o)] += q;
                        betaSum += Beta[o];
                    }
                }
                float x0 = ph[0] * ee, x1 = ph[1] * ed; po[0] = x0 + x1;
                float y0 = ph[0] * ed, y1 = ph[1] * ee; po[1] = y0 + y1;
                if (!flat[site]) { po[0] *= HL[2 * site + 0]; po[1] *= HL[2 * site + 1]; }
            } else {
                const float emit[2] = {Em[2 * site], Em[2 * site + 1]};
                V8 E0 = V8::set1(emit[0]), E1 = V8::set1(emit[1]);
                if (l == Lp - 1) {
                    const float fact1 = 1.0f / K;
                    V8 F1 = V8::set1(fact1);
                    unsigned k = 0;
                    for (; k < K8; k += 8) {
                        unsigned char byte = C->Hvar.getByte(l, k);
                        V8 em = blend8(E0, E1, byte);
                        V8 m0 = blend8(ONE, ZERO, byte), m1 = blend8(ZERO, ONE, byte);
                        V8 a = load8(al + k);
                        V8 pn = mul8(em, F1);
                        P0 = add8(P0, mul8(a, m0));
                        P1 = add8(P1, mul8(a, m1));
                        S = add8(S, pn);
                        store8(&Beta[k], pn);
                    }
                    if (k) { ph[0] = hadd8(P0); ph[1] = hadd8(P1); betaSum = hadd8(S); }
                    for (unsigned o = K8; o < K; o++) {
                        ph[C->Hvar.get(l, o)] += al[o];
                        Beta[o] = emit[C->Hvar.get(l, o)] * fact1;
                        betaSum += Beta[o];
                    }
                } else {
                    const float fact1 = C->t[l] / K;
                    const float fact2 = C->nt[l] / betaSumNext;
                    V8 F1 = V8::set1(fact1), F2 = V8::set1(fact2);
                    unsigned k = 0;
                    for (; k < K8; k += 8) {
                        unsigned char byte = C->Hvar.getByte(l, k);
                        V8 em = blend8(E0, E1, byte);
                        V8 m0 = blend8(ONE, ZERO, byte), m1 = blend8(ZERO, ONE, byte);
                        V8 pc = fma8(load8(&Beta[k]), F2, F1);
                        V8 dp = mul8(load8(al + k), pc);
                        V8 pn = mul8(pc, em);
                        P0 = add8(P0, mul8(dp, m0));
                        P1 = add8(P1, mul8(dp, m1));
                        S = add8(S, pn);
                        store8(&Beta[k], pn);
                    }
                    if (k) { ph[0] = hadd8(P0); ph[1] = hadd8(P1); betaSum = hadd8(S); }
                    for (unsigned o = K8; o < K; o++) {
                        float m = Beta[o] * fact2; Beta[o] = m + fact1;
                        float q = al[o] * Beta[o];
                        ph[C->Hvar.get(l, o)] += q;
                        Beta[o] *= emit[C->Hvar.get(l, o)];
                        betaSum += Beta[o];
                    }
                }
                ph[0] /= emit[0];
                ph[1] /= emit[1];
                float x0 = ph[0] * ee, x1 = ph[1] * ed; po[0] = (x0 + x1) * HL[2 * site + 0];
                float y0 = ph[0] * ed, y1 = ph[1] * ee; po[1] = (y0 + y1) * HL[2 * site + 1];
            }
            HP[2 * site + 0] = po[0] / (po[0] + po[1]);
            HP[2 * site + 1] = po[1] / (po[0] + po[1]);
            betaSumNext = betaSum;
        }
        for (unsigned site : C->mono) {
            po[C->P.major[site]] = ee;
            po[!C->P.major[site]] = ed;
            if (!flat[site]) { po[0] *= HL[2 * site + 0]; po[1] *= HL[2 * site + 1]; }
            HP[2 * site + 0] = po[0] / (po[0] + po[1]);
            HP[2 * site + 1] = po[1] / (po[0] + po[1]);
        }
    }
};

/* ------------------------------------------------------------------ */
/* Diploid phasing HMM: phase/src/models/phasing_hmm.h:39-447, phasing_hmm.cpp:28-302. */
static const int HAPN = 8;
static const int VAR_PEAK_HOM = -1, VAR_FLAT_HET = -2;

struct PhaseHMM {
    CondSet *C;
    std::vector<signed char> VT;
    std::vector<char> VALT;
    std::vector<int> VABS, VREL;
    unsigned n_segs = 0, n_miss = 0;
    std::vector<int> segments;
    int cidx = 0, crel = 0, cseg = 0, cseg_loc = 0, cmiss = 0;
    float probSumT = 0.0f;
    std::vector<float> prob, probSumK;
    V8 probSumH;
    std::vector<float> phasingProb, phasingProbSum, phasingProbSumSum;
    std::vector<float> imputeProb, imputeProbSum, imputeProbSumSum, imputeProbOf1s;
    std::vector<float> hp_trace;   /* TRANS_HAP's HProbs / sumHProbs at each boundary (for pinning tests) */
    std::vector<int> dip;
    float EMIT0[3][8], EMIT1[3][8];
    float HProbs[64], DProbs[8];
    float sumHProbs = 0.0f, sumDProbs = 0.0f, nt = 0.0f, yt = 0.0f;

    explicit PhaseHMM(CondSet *c) : C(c) {
        const float d = C->ed_phs, e = C->ee_phs;
        /* phasing_hmm.cpp:38-56: allele of het j under hypothesis h is 1-bit(2-j) of h */
        for (int j = 0; j < 3; j++)
            for (int h = 0; h < 8; h++) {
                bool hyp_alt0 = (h >> (2 - j)) & 1;  /* hypothesis carries allele 0 */
                EMIT0[j][h] = hyp_alt0 ? e : d;
                EMIT1[j][h] = hyp_alt0 ? d : e;
            }
    }

    void reallocate(const std::vector<uint8_t> &H0, const std::vector<uint8_t> &H1, const std::vector<char> &flat);
    void INIT_PEAK_HET(int het);
    void INIT_PEAK_HOM(bool ag);
    void INIT_FLAT_HET();
    void RUN_PEAK_HET(int het);
    void RUN_PEAK_HOM(bool ag);
    void RUN_FLAT_HET();
    void COLLAPSE_PEAK_HET(int het);
    void COLLAPSE_PEAK_HOM(bool ag);
    void COLLAPSE_FLAT_HET();
    void SUMK();
    bool TRANS_HAP();
    bool SAMPLE_DIP(Rng &rng, int tgt);
    void IMPUTE_FLAT_HET();
    void forward(Rng &rng, int tgt);
    void backward();
    void rephase(std::vector<uint8_t> &H0, std::vector<uint8_t> &H1, const std::vector<char> &flat, Rng &rng, int tgt);
};

void PhaseHMM::reallocate(const std::vector<uint8_t> &H0, const std::vector<uint8_t> &H1, const std::vector<char> &flat) {
    VT.clear(); VALT.clear(); VABS.clear(); VREL.clear();
    const std::vector<unsigned> &poly = C->poly;
    for (int l = 0, n_het = 0; l < (int)poly.size(); l++) {
        bool a0 = H0[poly[l]], a1 = H1[poly[l]];
        if (!flat[poly[l]] && !C->P.lq[poly[l]]) {
            if (a0 != a1) { VT.push_back(n_het % 3); VALT.push_back(a0); VABS.push_back(poly[l]); VREL.push_back(l); n_het++; }
            else { VT.push_back(VAR_PEAK_HOM); VALT.push_back(a0); VABS.push_back(poly[l]); VREL.push_back(l); }
        } else if (a0 != a1) {
            VT.push_back(VAR_FLAT_HET); VALT.push_back(a0); VABS.push_back(poly[l]); VREL.push_back(l);
        }
    }
    int nv = 0;
    segments.clear();
    n_miss = 0;
    for (int l = 0, n_hets = 0; l < (int)VT.size();) {
        n_hets += (VT[l] >= 0);
        n_miss += (VT[l] == VAR_FLAT_HET);
        if (n_hets == 4) { segments.push_back(nv); n_hets = 0; nv = 0; }
        else { nv++; l++; }
    }
    segments.push_back(nv);
    n_segs = (unsigned)segments.size();
    dip.assign(n_segs, -1);
    const size_t KH = (size_t)C->n_states * HAPN;
    prob.assign(KH, 0.0f);
    probSumK.assign(C->n_states, 0.0f);
    phasingProb.assign(n_segs * KH, 0.0f);
    phasingProbSum.assign(n_segs * HAPN, 0.0f);
    phasingProbSumSum.assign(n_segs, 0.0f);
    imputeProb.assign(n_miss * KH, 0.0f);
    imputeProbSum.assign(n_miss * HAPN, 0.0f);
    imputeProbSumSum.assign(n_miss, 0.0f);
    imputeProbOf1s.assign(n_miss * HAPN, 0.0f);
    hp_trace.assign((size_t)(n_segs > 0 ? n_segs - 1 : 0) * HAPN * HAPN, 0.0f);
}

void PhaseHMM::INIT_PEAK_HET(int het) {
    V8 e[2] = {load8(EMIT0[het]), load8(EMIT1[het])};
    V8 S = V8::set1(0.0f);
    for (unsigned k = 0; k < C->n_states; ++k) {
        bool ah = C->Hvar.get(crel, k);
        S = add8(S, e[ah]);
        store8(&prob[k * HAPN], e[ah]);
    }
    probSumH = S; probSumT = hadd8(S);
}
void PhaseHMM::INIT_PEAK_HOM(bool ag) {
    V8 e[2] = {V8::set1(1.0f), V8::set1(C->ed_phs / C->ee_phs)};
    V8 S = V8::set1(0.0f);
    for (unsigned k = 0; k < C->n_states; ++k) {
        bool x = C->Hvar.get(crel, k) != ag;
        S = add8(S, e[x]);
        store8(&prob[k * HAPN], e[x]);
    }
    probSumH = S; probSumT = hadd8(S);
}
void PhaseHMM::INIT_FLAT_HET() {
    std::fill(prob.begin(), prob.end(), 1.0f / (HAPN * C->n_states));
    probSumH = V8::set1(1.0f / HAPN);
    probSumT = 1.0f;
}
void PhaseHMM::RUN_PEAK_HET(int het) {
    V8 tF = mul8(probSumH, V8::set1(yt / (C->n_states * probSumT)));
    V8 NT = V8::set1(nt / probSumT);
    V8 e[2] = {load8(EMIT0[het]), load8(EMIT1[het])};
    V8 S = V8::set1(0.0f);
    for (unsigned k = 0; k < C->n_states; ++k) {
        bool ah = C->Hvar.get(crel, k);
        V8 pc = mul8(fma8(load8(&prob[k * HAPN]), NT, tF), e[ah]);
        S = add8(S, pc);
        store8(&prob[k * HAPN], pc);
    }
    probSumH = S; probSumT = hadd8(S);
}
void PhaseHMM::RUN_PEAK_HOM(bool ag) {
    V8 tF = mul8(probSumH, V8::set1(yt / (C->n_states * probSumT)));
    V8 NT = V8::set1(nt / probSumT);
    V8 MI = V8::set1(C->ed_phs / C->ee_phs);
    V8 S = V8::set1(0.0f);
    for (unsigned k = 0; k < C->n_states; ++k) {
        bool ah = C->Hvar.get(crel, k);
        V8 pc = fma8(load8(&prob[k * HAPN]), NT, tF);
        if (ag != ah) pc = mul8(pc, MI);
        S = add8(S, pc);
        store8(&prob[k * HAPN], pc);
    }
    probSumH = S; probSumT = hadd8(S);
}
void PhaseHMM::RUN_FLAT_HET() {
    V8 tF = mul8(probSumH, V8::set1(yt / (C->n_states * probSumT)));
    V8 NT = V8::set1(nt / probSumT);
    V8 S = V8::set1(0.0f);
    for (unsigned k = 0; k < C->n_states; ++k) {
        V8 pc = fma8(load8(&prob[k * HAPN]), NT, tF);
        S = add8(S, pc);
        store8(&prob[k * HAPN], pc);
    }
    probSumH = S; probSumT = hadd8(S);
}
void PhaseHMM::COLLAPSE_PEAK_HET(int het) {
    V8 tF = mul8(probSumH, V8::set1(yt / (C->n_states * probSumT)));
    V8 NT = V8::set1(nt / probSumT);
    V8 e[2] = {load8(EMIT0[het]), load8(EMIT1[het])};
    V8 S = V8::set1(0.0f);
    for (unsigned k = 0; k < C->n_states; ++k) {
        bool ah = C->Hvar.get(crel, k);
        V8 pc = mul8(fma8(V8::set1(probSumK[k]), NT, tF), e[ah]);
        S = add8(S, pc);
        store8(&prob[k * HAPN], pc);
    }
    probSumH = S; probSumT = hadd8(S);
}
void PhaseHMM::COLLAPSE_PEAK_HOM(bool ag) {
    V8 tF = mul8(probSumH, V8::set1(yt / (C->n_states * probSumT)));
    V8 NT = V8::set1(nt / probSumT);
    V8 MI = V8::set1(C->ed_phs / C->ee_phs);
    V8 S = V8::set1(0.0f);
    for (unsigned k = 0; k < C->n_states; ++k) {
        bool ah = C->Hvar.get(crel, k);
        V8 pc = fma8(V8::set1(probSumK[k]), NT, tF);
        if (ag != ah) pc = mul8(pc, MI);
        S = add8(S, pc);
        store8(&prob[k * HAPN], pc);
    }
    probSumH = S; probSumT = hadd8(S);
}
void PhaseHMM::COLLAPSE_FLAT_HET() {
    V8 tF = mul8(probSumH, V8::set1(yt / (C->n_states * probSumT)));
    V8 NT = V8::set1(nt / probSumT);
    V8 S = V8::set1(0.0f);
    for (unsigned k = 0; k < C->n_states; ++k) {
        V8 pc = fma8(V8::set1(probSumK[k]), NT, tF);
        S = add8(S, pc);
        store8(&prob[k * HAPN], pc);
    }
    probSumH = S; probSumT = hadd8(S);
}
void PhaseHMM::SUMK() {
    for (unsigned k = 0; k < C->n_states; ++k) probSumK[k] = hadd8(load8(&prob[k * HAPN]));
}
bool PhaseHMM::TRANS_HAP() {
    const size_t KH = (size_t)C->n_states * HAPN;
    sumHProbs = 0.0f;
    yt = C->getTransition(VABS[cidx], VABS[cidx + 1]);
    nt = 1.0f - yt;
    V8 F2 = V8::set1(nt / probSumT);
    for (int h1 = 0, j = 0; h1 < HAPN; h1++, j += HAPN) {
        V8 F1 = V8::set1((probSumH.v[h1] / probSumT) * yt / C->n_states);
        V8 S = V8::set1(0.0f);
        const float *nxt = &phasingProb[(size_t)(cseg + 1) * KH];
        for (unsigned k = 0; k < C->n_states; ++k) {
            V8 p0 = fma8(V8::set1(prob[k * HAPN + h1]), F2, F1);
            S = add8(S, mul8(p0, load8(nxt + k * HAPN)));
        }
        store8(&HProbs[j], S);
        sumHProbs += hadd8(S);
    }
    for (int i = 0; i < HAPN * HAPN; i++) hp_trace[(size_t)cseg * HAPN * HAPN + i] = HProbs[i] / sumHProbs;
    return std::isnan(sumHProbs) || std::isinf(sumHProbs) || sumHProbs < std::numeric_limits<float>::min();
}
bool PhaseHMM::SAMPLE_DIP(Rng &rng, int tgt) {
    sumDProbs = 0.0f;
    for (int d = 0; d < HAPN; d++) {
        int p0 = dip[cseg], p1 = HAPN - dip[cseg] - 1;
        DProbs[d] = (HProbs[p0 * HAPN + d] / sumHProbs) * (HProbs[p1 * HAPN + (HAPN - d - 1)] / sumHProbs);
        sumDProbs += DProbs[d];
    }
    if (std::isnan(sumDProbs) || std::isinf(sumDProbs) || sumDProbs < std::numeric_limits<float>::min()) return true;
    dip[cseg + 1] = rng.sample8(DProbs, sumDProbs, P_DIP_SEG, (uint32_t)tgt, (uint32_t)(cseg + 1));
    return false;
}
void PhaseHMM::IMPUTE_FLAT_HET() {
    const size_t KH = (size_t)C->n_states * HAPN;
    V8 sR = load8(&imputeProbSum[cmiss * HAPN]);
    V8 sL = probSumH;
    for (int i = 0; i < 8; i++) { sR.v[i] = 1.0f / sR.v[i]; sL.v[i] = 1.0f / sL.v[i]; }
    V8 sums[2] = {V8::set1(0.0f), V8::set1(0.0f)};
    for (unsigned k = 0; k < C->n_states; ++k) {
        bool ah = C->Hvar.get(crel, k);
        V8 p1 = mul8(load8(&imputeProb[cmiss * KH + k * HAPN]), sR);
        V8 p2 = mul8(load8(&prob[k * HAPN]), sL);
        sums[ah] = add8(sums[ah], mul8(p1, p2));
    }
    for (int i = 0; i < 8; i++) {
        float n = sums[1].v[i] / (sums[0].v[i] + sums[1].v[i]);
        imputeProbOf1s[cmiss * HAPN + i] = std::max(0.0f, std::min(1.0f, n));
    }
}
void PhaseHMM::forward(Rng &rng, int tgt) {
    cseg = 0; cseg_loc = 0; cmiss = 0;
    for (cidx = 0; cidx < (int)VT.size(); cidx++) {
        crel = VREL[cidx];
        yt = 0.0f;
        if (cidx) yt = C->getTransition(VABS[cidx - 1], VABS[cidx]);
        nt = 1.0f - yt;
        if (VT[cidx] >= 0) {
            if (cidx == 0) INIT_PEAK_HET(VT[cidx]);
            else if (cseg_loc != 0) RUN_PEAK_HET(VT[cidx]);
            else COLLAPSE_PEAK_HET(VT[cidx]);
        } else if (VT[cidx] == VAR_PEAK_HOM) {
            if (cidx == 0) INIT_PEAK_HOM(VALT[cidx]);
            else if (cseg_loc != 0) RUN_PEAK_HOM(VALT[cidx]);
            else COLLAPSE_PEAK_HOM(VALT[cidx]);
        } else {
            if (cidx == 0) INIT_FLAT_HET();
            else if (cseg_loc != 0) RUN_FLAT_HET();
            else COLLAPSE_FLAT_HET();
        }
        if (cseg_loc == segments[cseg] - 1) SUMK();
        if (cseg_loc == segments[cseg] - 1 && cidx != (int)VT.size() - 1) {
            if (TRANS_HAP()) throw OrcError("phasing: underflow in TRANS_HAP");
            if (SAMPLE_DIP(rng, tgt)) throw OrcError("phasing: underflow in SAMPLE_DIP");
        }
        if (VT[cidx] == VAR_FLAT_HET) { IMPUTE_FLAT_HET(); cmiss++; }
        cseg_loc++;
        if (cseg_loc >= segments[cseg]) { cseg++; cseg_loc = 0; }
    }
}
void PhaseHMM::backward() {
    const size_t KH = (size_t)C->n_states * HAPN;
    cseg = n_segs - 1;
    cseg_loc = segments.back() - 1;
    cmiss = (int)n_miss - 1;
    for (cidx = (int)VT.size() - 1; cidx >= 0; cidx--) {
        crel = VREL[cidx];
        yt = 0.0f;
        if (cidx < (int)VT.size() - 1) yt = C->getTransition(VABS[cidx], VABS[cidx + 1]);
        nt = 1.0f - yt;
        const bool last = cidx == (int)VT.size() - 1;
        const bool collapse = cseg_loc == segments[cseg] - 1;
        if (VT[cidx] >= 0) {
            if (last) INIT_PEAK_HET(VT[cidx]);
            else if (!collapse) RUN_PEAK_HET(VT[cidx]);
            else COLLAPSE_PEAK_HET(VT[cidx]);
        } else if (VT[cidx] == VAR_PEAK_HOM) {
            if (last) INIT_PEAK_HOM(VALT[cidx]);
            else if (!collapse) RUN_PEAK_HOM(VALT[cidx]);
            else COLLAPSE_PEAK_HOM(VALT[cidx]);
        } else {
            if (last) INIT_FLAT_HET();
            else if (!collapse) RUN_FLAT_HET();
            else COLLAPSE_FLAT_HET();
        }
        if (cseg_loc == 0) {
            SUMK();
            std::copy(prob.begin(), prob.end(), phasingProb.begin() + (size_t)cseg * KH);
            store8(&phasingProbSum[cseg * HAPN], probSumH);
            phasingProbSumSum[cseg] = probSumT;
        }
        if (VT[cidx] == VAR_FLAT_HET) {
            std::copy(prob.begin(), prob.end(), imputeProb.begin() + (size_t)cmiss * KH);
            store8(&imputeProbSum[cmiss * HAPN], probSumH);
            imputeProbSumSum[cmiss] = probSumT;
            cmiss--;
        }
        cseg_loc--;
        if (cseg_loc < 0 && cseg > 0) { cseg--; cseg_loc = segments[cseg] - 1; }
    }
}
/* phasing_hmm.cpp:247-302 (including the FLAT_HET write-back that assigns H0 twice) */
void PhaseHMM::rephase(std::vector<uint8_t> &H0, std::vector<uint8_t> &H1, const std::vector<char> &flat, Rng &rng, int tgt) {
    reallocate(H0, H1, flat);
    backward();
    sumDProbs = 0.0f;
    for (int d = 0; d < HAPN; d++) {
        DProbs[d] = (phasingProbSum[d] / phasingProbSumSum[0]) * (phasingProbSum[HAPN - d - 1] / phasingProbSumSum[0]);
        sumDProbs += DProbs[d];
    }
    dip[0] = rng.sample8(DProbs, sumDProbs, P_DIP_ROOT, (uint32_t)tgt, 0u);
    forward(rng, tgt);
    cseg = 0; cseg_loc = 0; cmiss = 0;
    for (cidx = 0; cidx < (int)VT.size(); cidx++) {
        if (VT[cidx] >= 0) {
            const int i0 = dip[cseg], i1 = HAPN - dip[cseg] - 1;
            const bool a0 = ((HAPN - i0 - 1) >> (2 - VT[cidx])) & 1;
            const bool a1 = ((HAPN - i1 - 1) >> (2 - VT[cidx])) & 1;
            H0[VABS[cidx]] = a0;
            H1[VABS[cidx]] = a1;
        } else if (VT[cidx] == VAR_FLAT_HET) {
            const int i0 = dip[cseg], i1 = HAPN - dip[cseg] - 1;
            const float h0a0 = 1.0f - imputeProbOf1s[cmiss * HAPN + i0];
            const float h0a1 = imputeProbOf1s[cmiss * HAPN + i0];
            const float h1a0 = 1.0f - imputeProbOf1s[cmiss * HAPN + i1];
            const float h1a1 = imputeProbOf1s[cmiss * HAPN + i1];
            const float ee = C->ee_phs, ed = C->ed_phs;
            float p01 = (h0a0 * ee + h0a1 * ed) * (h1a1 * ee + h1a0 * ed);
            float p10 = (h0a1 * ee + h0a0 * ed) * (h1a0 * ee + h1a1 * ed);
            float sum = p01 + p10;
            p01 = std::clamp(p01 / sum, 0.0f, 1.0f);
            p10 = std::clamp(p10 / sum, 0.0f, 1.0f);
            sum = p01 + p10;
            const bool rf = (rng.getFloat(P_FLAT_HET, (uint32_t)tgt, (uint32_t)cmiss) * sum) < p01;
            H0[VABS[cidx]] = rf;
            H0[VABS[cidx]] = !rf;
            cmiss++;
        }
        cseg_loc++;
        if (cseg_loc >= segments[cseg]) { cseg++; cseg_loc = 0; }
    }
    for (unsigned site : C->mono) {
        if (H0[site] != H1[site]) {
            const bool rf = rng.getFloat(P_MONO_HET, (uint32_t)tgt, site) < 0.5f;
            H0[site] = rf;
            H1[site] = !rf;
        }
    }
}

/* ------------------------------------------------------------------ */
/* genotype: common/src/objects/genotype.cpp:56-244 (diploid). */
static void hl_init(int L, const uint8_t *gl, const std::vector<char> &flat, float min_gl, std::vector<float> &HL) {
    for (int l = 0; l < L; l++) {
        if (!flat[l]) {
            float t0 = UNPHRED[gl[3 * l + 0]], t1 = UNPHRED[gl[3 * l + 1]], t2 = UNPHRED[gl[3 * l + 2]];
            float s = t0 + t1 + t2;
            t0 /= s; t1 /= s; t2 /= s;
            HL[2 * l + 0] = t0;
            HL[2 * l + 1] = t1;
            HL[2 * l + 0] = (float)((double)HL[2 * l + 0] + 0.5 * (double)t1);
            HL[2 * l + 1] = (float)(0.5 * (double)t1 + (double)t2);
        } else {
            HL[2 * l + 0] = 0.5f;
            HL[2 * l + 1] = 0.5f;
        }
        if (HL[2 * l + 0] < min_gl) { HL[2 * l + 0] = min_gl; HL[2 * l + 1] = 1.0f - min_gl; }
        if (HL[2 * l + 1] < min_gl) { HL[2 * l + 0] = 1.0f - min_gl; HL[2 * l + 1] = min_gl; }
    }
}
/* initHaplotypeLikelihoods for ploidy 1 (genotype.cpp:56-90): the pair of GL values, normalised */
static void hl_init_hap(int L, const uint8_t *gl, const std::vector<char> &flat, float min_gl, std::vector<float> &HL) {
    for (int l = 0; l < L; l++) {
        if (!flat[l]) {
            float t0 = UNPHRED[gl[3 * l + 0]], t1 = UNPHRED[gl[3 * l + 1]], t2 = 0.0f;
            float s = t0 + t1 + t2;
            t0 /= s; t1 /= s;
            HL[2 * l + 0] = t0;
            HL[2 * l + 1] = t1;
        } else {
            HL[2 * l + 0] = 0.5f;
            HL[2 * l + 1] = 0.5f;
        }
        if (HL[2 * l + 0] < min_gl) { HL[2 * l + 0] = min_gl; HL[2 * l + 1] = 1.0f - min_gl; }
        if (HL[2 * l + 1] < min_gl) { HL[2 * l + 0] = 1.0f - min_gl; HL[2 * l + 1] = min_gl; }
    }
}
static void hl_make(int L, const uint8_t *gl, const std::vector<char> &flat, const std::vector<uint8_t> &other,
                    float min_gl, std::vector<float> &HL) {
    for (int l = 0; l < L; l++) {
        if (flat[l]) continue;
        float t[3] = {UNPHRED[gl[3 * l + 0]], UNPHRED[gl[3 * l + 1]], UNPHRED[gl[3 * l + 2]]};
        float s = t[0] + t[1] + t[2];
        t[0] /= s; t[1] /= s; t[2] /= s;
        const int c = other[l];
        HL[2 * l + 0] = t[0 + c] / (t[0 + c] + t[1 + c]);
        HL[2 * l + 1] = t[1 + c] / (t[0 + c] + t[1 + c]);
        if (HL[2 * l + 0] < min_gl) { HL[2 * l + 0] = min_gl; HL[2 * l + 1] = 1.0f - min_gl; }
        if (HL[2 * l + 1] < min_gl) { HL[2 * l + 0] = 1.0f - min_gl; HL[2 * l + 1] = min_gl; }
    }
}

struct Stored { float gp0, gp1; bool hds; int32_t idx; };

struct Target {
    int L;
    int ploidy = 2;
    const uint8_t *gl;
    std::vector<char> flat;
    std::vector<uint8_t> H0, H1;
    int stored_cnt = 0;
    std::vector<Stored> stored;
    std::vector<int> sindgl;         /* SindTarGL */
    std::set<int> init_states;

    /* genotype.cpp:168-209 */
    void store(const std::vector<float> &HP0, const std::vector<float> &HP1) {
        std::vector<bool> flag(L, false);
        for (auto &e : stored) {
            int v = e.idx;
            float p0 = std::clamp(HP0[2 * v] * HP1[2 * v], 0.0f, 1.0f);
            float p1 = std::clamp(HP0[2 * v] * HP1[2 * v + 1] + HP0[2 * v + 1] * HP1[2 * v], 0.0f, 1.0f);
            float p2 = std::clamp(HP0[2 * v + 1] * HP1[2 * v + 1], 0.0f, 1.0f);
            e.gp0 += p0 / (p0 + p1 + p2);
            e.gp1 += p1 / (p0 + p1 + p2);
            e.hds = HP0[2 * v + 1] < HP1[2 * v + 1];
            flag[v] = true;
        }
        for (int l = 0; l < L; l++) {
            if (flag[l]) continue;
            float p0 = std::clamp(HP0[2 * l] * HP1[2 * l], 0.0f, 1.0f);
            float p1 = std::clamp(HP0[2 * l] * HP1[2 * l + 1] + HP0[2 * l + 1] * HP1[2 * l], 0.0f, 1.0f);
            float p2 = std::clamp(HP0[2 * l + 1] * HP1[2 * l + 1], 0.0f, 1.0f);
            if (p0 / (p0 + p1 + p2) < 0.99999f)
                stored.push_back({p0 / (p0 + p1 + p2) + stored_cnt * 1.0f, p1 / (p0 + p1 + p2), HP0[2 * l + 1] < HP1[2 * l + 1], l});
        }
        stored_cnt++;
    }

    /* genotype.cpp:127-166, haploid */
    void store_hap(const std::vector<float> &HP0) {
        std::vector<bool> flag(L, false);
        for (auto &e : stored) {
            const int v = e.idx;
            const float p0 = HP0[2 * v], p1 = HP0[2 * v + 1];
            const float sc = 1.0f / (p0 + p1);
            e.gp0 += p0 * sc;
            e.gp1 += p1 * sc;
            e.hds = false;
            flag[v] = true;
        }
        for (int l = 0; l < L; l++) {
            if (flag[l]) continue;
            const float p0 = HP0[2 * l], p1 = HP0[2 * l + 1];
            const float sc = 1.0f / (p0 + p1);
            if (p0 * sc < 0.99999f) stored.push_back({p0 * sc + (stored_cnt % 16) * 1.0f, p1 * sc, false, l});
        }
        stored_cnt++;
    }
};

/* haplotype_set.cpp:108-169 */
static void init_rare_tar(const Panel &P, Target &G) {
    G.sindgl.clear();
    for (int v = 0; v < P.L; v++) {
        if (G.flat[v]) continue;
        const int mg = 2 * P.major[v];
        if (!P.common[v] && !P.lq[v]) {
            /* a haploid target's GL array holds two values per site: the
             * reference's offsets v*2 + {0, 1, 2} reach the next site's first
             * value (0 past the last site); its third probability is 0 */
            uint8_t gl[3] = {G.gl[3 * v], G.gl[3 * v + 1], G.gl[3 * v + 2]};
            if (G.ploidy == 1) gl[2] = v + 1 < P.L ? G.gl[3 * (v + 1)] : 0;
            if (gl[1] <= gl[mg] || gl[2 - mg] <= gl[mg]) {
                float t[3] = {UNPHRED[gl[0]], UNPHRED[gl[1]], G.ploidy == 1 ? 0.0f : UNPHRED[gl[2]]};
                float s = t[0] + t[1] + t[2];
                t[0] /= s; t[1] /= s; t[2] /= s;
                const float af = P.calt[v] * 1.0 / P.N;
                t[0] *= (1.0f - af) * (1.0f - af);
                t[1] *= 2.0f * af * (1.0f - af);
                t[2] *= (af * af);
                s = t[0] + t[1] + t[2];
                t[0] /= s; t[1] /= s; t[2] /= s;
                if (t[1] + t[2 - mg] > t[mg]) G.sindgl.push_back(v);
            }
        }
    }
}

/* Selection sampling (Knuth, TAOCP vol. 2, Algorithm S) with counter draws. */
static void counter_sample(const std::vector<int> &pop, int n, std::vector<int> &out,
                           Rng &rng, uint32_t tgt, uint32_t &ctr) {
    int N = (int)pop.size(), need = std::min(n, N);
    for (int i = 0; i < N && need > 0; i++) {
        float u = (float)rng.getFloat(P_INIT_SEL, tgt, ctr++);
        if (u * (float)(N - i) < (float)need) { out.push_back(pop[i]); need--; }
    }
}

/* haplotype_set.cpp:775-833, one target (index ind keys its counter draws) */
static void init_selection_one(const Panel &P, Target &T, size_t ind, int kinit, Rng &rng) {
    std::vector<int> iota_all(P.N);
    std::iota(iota_all.begin(), iota_all.end(), 0);
    const int k_init = kinit;
    const int k_init8 = (int)std::floor(k_init * 0.8f);
    std::vector<int> sampled;
    {
        std::set<int> &S = T.init_states;
        uint32_t ctr = 0;
        auto do_sample = [&](const std::vector<int> &pop, int n) {
            sampled.clear();
            if (rng.mode == ORC_RNG_REFERENCE)
                std::sample(pop.begin(), pop.end(), std::back_inserter(sampled), n, rng.eng);
            else
                counter_sample(pop, n, sampled, rng, (uint32_t)ind, ctr);
        };
        if (!T.sindgl.empty()) {
            if ((int)T.sindgl.size() > k_init8) {
                std::vector<std::pair<int, int>> mac;
                for (int v : T.sindgl) mac.push_back({(int)std::min(P.cref[v], P.calt[v]), v});
                std::sort(mac.begin(), mac.end());
                for (size_t r0 = 0; r0 < T.sindgl.size() && (int)S.size() < k_init8; r0++) {
                    const int v = mac[r0].second;
                    std::vector<int> pop(iota_all.begin(), iota_all.begin() + P.SvarRef[v].size());
                    do_sample(pop, 5);
                    for (size_t i = 0; i < sampled.size() && (int)S.size() < k_init8; ++i) S.insert(P.SvarRef[v][sampled[i]]);
                }
            } else {
                const int size_max = std::max((int)std::floor(k_init8 * 1.0 / T.sindgl.size()), 1);
                for (int v : T.sindgl) {
                    do_sample(P.SvarRef[v], size_max);
                    for (size_t i = 0; i < sampled.size() && (int)S.size() < k_init8; ++i) S.insert(sampled[i]);
                }
            }
        }
        if ((int)S.size() < k_init) {
            do_sample(iota_all, k_init);
            for (size_t i = 0; i < sampled.size() && (int)S.size() < k_init; ++i) S.insert(sampled[i]);
        }
    }
}

static void init_selection(const Panel &P, std::vector<Target> &G, int kinit, Rng &rng) {
    for (size_t ind = 0; ind < G.size(); ++ind) init_selection_one(P, G[ind], ind, kinit, rng);
}

/* genotype::sortAndNormAndInferGenotype (genotype.cpp:211-244) for one target */
static void finalize_target(Target &g, int L, float *gp0, float *gp1, uint8_t *hds, uint8_t *h0, uint8_t *h1) {
    std::sort(g.stored.begin(), g.stored.end(), [](const Stored &a, const Stored &b) { return a.idx < b.idx; });
    for (auto &e : g.stored) { e.gp0 /= g.stored_cnt; e.gp1 /= g.stored_cnt; }
    for (int l = 0; l < L; l++) { gp0[l] = 1.0f; gp1[l] = 0.0f; hds[l] = 0; h0[l] = 0; h1[l] = 0; }
    for (auto &e : g.stored) {
        const size_t o = e.idx;
        gp0[o] = e.gp0; gp1[o] = e.gp1; hds[o] = e.hds;
        if (g.ploidy == 1) { h0[o] = e.gp1 > e.gp0; h1[o] = 0; continue; }   /* infer_haploid */
        const float g2 = 1.0f - e.gp1 - e.gp0;
        int gt = 0;
        if (e.gp0 > e.gp1 && e.gp0 > g2) gt = 0;
        else if (e.gp1 > e.gp0 && e.gp1 > g2) gt = 1;
        else if (g2 > e.gp0 && g2 > e.gp1) gt = 2;
        h0[o] = gt == 2; h1[o] = gt >= 1;
    }
}

/* ------------------------------------------------------------------ */
/* Gibbs driver: phase/src/caller/caller_algorithm.cpp:53-243 and
 * caller_initialise.cpp:142-164 (threads = 1). */
static void phase_run(const Panel &P, int T, const uint8_t *gl, const uint8_t *flat, const orc_params &prm,
                      const int32_t *kpt, float *gp0, float *gp1, uint8_t *hds, uint8_t *h0, uint8_t *h1,
                      int32_t *nstates_out) {
    const int L = P.L, N = P.N;
    if (prm.rng_mode == ORC_RNG_REFERENCE && kpt) throw OrcError("reference RNG mode needs one Kpbwt for all targets");
    Rng rng; rng.mode = prm.rng_mode; rng.set_seed(prm.seed);
    const int kinit = std::min(prm.kinit, N);
    const float err_imp = std::clamp(prm.err_imp, 1e-12f, 1e-3f);
    std::vector<int> kp(T);
    for (int t = 0; t < T; t++) kp[t] = std::min(kpt ? kpt[t] : prm.kpbwt, N);

    std::vector<Target> G(T);
    for (int t = 0; t < T; t++) {
        G[t].L = L; G[t].gl = gl + (size_t)t * L * 3;
        G[t].flat.assign(flat + (size_t)t * L, flat + (size_t)(t + 1) * L);
        G[t].H0.assign(L, 0); G[t].H1.assign(L, 0);
        G[t].ploidy = prm.ploidy ? prm.ploidy[t] : 2;
        if (G[t].ploidy != 1 && G[t].ploidy != 2) throw OrcError("ploidy must be 1 or 2");
    }
    /* PBWT classes: targets sharing an effective modulo share groups and stored draws */
    std::map<float, std::vector<int>> classes;
    std::map<float, std::vector<int>> class_grp;
    std::map<float, int> class_id;
    for (int t = 0; t < T; t++) {
        if (kp[t] == 0 || kp[t] >= N) continue;
        float em; std::vector<int> grp;
        pbwt_groups(P, prm.pbwt_depth, prm.pbwt_modulo_cm, kp[t], em, grp);
        classes[em].push_back(t);
        class_grp[em] = grp;
    }
    { int cid = 0; for (auto &c : classes) class_id[c.first] = cid++; }

    const int n_threads = prm.rng_mode == ORC_RNG_COUNTER ? std::max(1, std::min(prm.n_threads, T)) : 1;
    /* per-worker models and buffers (caller_initialise.cpp:143-161) */
    struct Worker {
        CondSet C;
        ImputeHMM HMM;
        PhaseHMM DMM;
        std::vector<float> HL, HP0, HP1;
        Worker(const Panel &P, int ne, float ei, float ep, int L)
            : C(P, ne, ei, ep), HMM(&C), DMM(&C), HL(2 * L, 0.0f), HP0(2 * L, 0.0f), HP1(2 * L, 0.0f) {}
    };
    std::vector<std::unique_ptr<Worker>> workers;
    for (int w = 0; w < n_threads; w++)
        workers.emplace_back(new Worker(P, prm.ne, err_imp, prm.err_phase, L));
    const int n_iter = 1 + prm.burnin + prm.main_iter;
    std::vector<std::vector<std::vector<int>>> pbwt_states(T);

    for (int it = 0; it < n_iter; it++) {
        rng.it = it;
        const int stage = it == 0 ? 0 : (it <= prm.burnin ? 1 : 2);
        if (stage == 0) {
            if (kinit > 0) {
                for (int t = 0; t < T; t++) init_rare_tar(P, G[t]);
                init_selection(P, G, kinit, rng);
            }
        } else {
            for (auto &cls : classes) {
                const float em = cls.first;
                const std::vector<int> &tg = cls.second;
                const std::vector<int> &grp = class_grp[em];
                std::vector<bool> stored(P.n_com_hq, false);
                for (int gi = 0, lo = 0; gi < (int)grp.size(); ++gi) {
                    int s = (int)rng.getInt(lo, grp[gi] - 1, P_PBWT_STORED, (uint32_t)gi, (uint32_t)class_id[em]);
                    stored[s] = true;
                    lo = grp[gi];
                }
                /* target haplotypes: two per diploid target, one per haploid one
                 * (tar_hapid2ind, haplotype_set.cpp:263-270) */
                std::vector<std::vector<uint8_t>> haps;
                std::vector<int> start, h2s, hid;
                for (size_t i = 0; i < tg.size(); i++) {
                    const int t = tg[i];
                    haps.push_back(G[t].H0); h2s.push_back((int)i); hid.push_back(2 * t);
                    if (G[t].ploidy > 1) { haps.push_back(G[t].H1); h2s.push_back((int)i); hid.push_back(2 * t + 1); }
                }
                for (size_t e = 0; e < haps.size(); e++)
                    start.push_back((int)rng.getInt(0, N - 1, P_PBWT_START, (uint32_t)hid[e], 0u));
                Sweep S(P, haps, (int)tg.size(), prm.pbwt_depth, em);
                S.hap2sample = h2s;
                S.run(stored, start);
                for (size_t i = 0; i < tg.size(); i++) pbwt_states[tg[i]] = S.states[i];
            }
        }
        /* caller::phase_individual per target (caller_algorithm.cpp:53-84) */
        auto individual = [&](int t, Worker &w, Rng &r) {
            Target &g = G[t];
            std::vector<int> states;
            if (stage == 0 && kinit > 0) states.assign(g.init_states.begin(), g.init_states.end());
            else if (kp[t] > 0 && kp[t] < N) states = compact_states(pbwt_states[t], kp[t]);
            else if (kp[t] >= N) { states.resize(N); std::iota(states.begin(), states.end(), 0); }
            w.C.build(states);
            if (nstates_out) nstates_out[t * n_iter + it] = (int32_t)w.C.n_states;
            if (g.ploidy == 1) {   /* caller_algorithm.cpp:58-77 with ploidy 1 */
                hl_init_hap(L, g.gl, g.flat, prm.min_gl, w.HL);
                w.HMM.run(w.HL, g.flat, w.HP0);
                for (int l = 0; l < L; l++) g.H0[l] = r.getFloat(P_H0, (uint32_t)t, (uint32_t)l) > w.HP0[2 * l];
                if (stage == 2) g.store_hap(w.HP0);
                return;
            }
            if (stage == 0) hl_init(L, g.gl, g.flat, prm.min_gl, w.HL);
            else hl_make(L, g.gl, g.flat, g.H1, prm.min_gl, w.HL);
            w.HMM.run(w.HL, g.flat, w.HP0);
            for (int l = 0; l < L; l++) g.H0[l] = r.getFloat(P_H0, (uint32_t)t, (uint32_t)l) > w.HP0[2 * l];
            hl_make(L, g.gl, g.flat, g.H0, prm.min_gl, w.HL);
            w.HMM.run(w.HL, g.flat, w.HP1);
            for (int l = 0; l < L; l++) g.H1[l] = r.getFloat(P_H1, (uint32_t)t, (uint32_t)l) > w.HP1[2 * l];
            w.DMM.rephase(g.H0, g.H1, g.flat, r, t);
            if (stage == 2) g.store(w.HP0, w.HP1);
        };
        if (n_threads == 1) {
            for (int t = 0; t < T; t++) individual(t, *workers[0], rng);
        } else {
            /* counter draws are keyed by (iteration, target, site): any thread order gives the same run */
            std::atomic<int> next(0);
            std::vector<std::thread> th;
            std::vector<std::string> errs(n_threads);
            for (int w = 0; w < n_threads; w++)
                th.emplace_back([&, w]() {
                    Rng r = rng;
                    try {
                        for (int t = next++; t < T; t = next++) individual(t, *workers[w], r);
                    } catch (const std::exception &e) {
                        errs[w] = e.what();
                    }
                });
            for (auto &x : th) x.join();
            for (auto &e : errs) if (!e.empty()) throw OrcError(e);
        }
        if (stage == 0) for (auto &g : G) g.init_states.clear();
    }
    for (int t = 0; t < T; t++) {
        const size_t o = (size_t)t * L;
        finalize_target(G[t], L, gp0 + o, gp1 + o, hds + o, h0 + o, h1 + o);
    }
}

} // namespace orc

/* ------------------------------------------------------------------ */
/* C interface */
using namespace orc;

#define ORC_TRY(...) try { __VA_ARGS__; return 0; } catch (const std::exception &e) { g_err = e.what(); return -1; }

extern "C" {

void orc_default_params(orc_params *p) {
    /* phase/src/caller/caller_parameters.cpp:34-98 */
    p->burnin = 5; p->main_iter = 15; p->ne = 100000;
    p->min_gl = 1e-10f; p->err_imp = 1e-12f; p->err_phase = 0.0001f;
    p->pbwt_depth = 12; p->pbwt_modulo_cm = 0.1f;
    p->kinit = 1000; p->kpbwt = 2000;
    p->sparse_maf = 0.001f; p->seed = 15052011u; p->rng_mode = ORC_RNG_REFERENCE; p->n_threads = 1;
}

const char *orc_last_error(void) { return g_err.c_str(); }

void *orc_panel_create(const orc_panel_desc *d, float sparse_maf) {
    try { return make_panel(d, sparse_maf); }
    catch (const std::exception &e) { g_err = e.what(); return nullptr; }
}
void orc_panel_free(void *p) { delete (Panel *)p; }
int32_t orc_panel_n_common(void *p) { return ((Panel *)p)->n_com; }
int32_t orc_panel_n_common_hq(void *p) { return ((Panel *)p)->n_com_hq; }
void orc_panel_site_class(void *p, uint8_t *fc, uint8_t *mj) {
    Panel *P = (Panel *)p;
    for (int l = 0; l < P->L; l++) { fc[l] = P->common[l]; mj[l] = P->major[l]; }
}
int64_t orc_panel_ypacked(void *p, uint8_t *out, int64_t cap) {
    Panel *P = (Panel *)p;
    int64_t n = (int64_t)P->Ypacked.size();
    if (out && cap >= n) std::memcpy(out, P->Ypacked.data(), n);
    return n;
}
int64_t orc_panel_asmall(void *p, int32_t *ptr, int32_t *idx, int64_t cap) {
    Panel *P = (Panel *)p;
    int64_t tot = 0;
    for (auto &v : P->A_small_idx) tot += v.size();
    if (!ptr || !idx || cap < tot) return tot;
    int64_t o = 0;
    for (size_t l = 0; l < P->A_small_idx.size(); l++) {
        ptr[l] = (int32_t)o;
        for (int x : P->A_small_idx[l]) idx[o++] = x;
    }
    ptr[P->A_small_idx.size()] = (int32_t)o;
    return tot;
}

int orc_phase_run(void *panel, int32_t T, const uint8_t *gl, const uint8_t *flat, const orc_params *p,
                  const int32_t *kpt, float *gp0, float *gp1, uint8_t *hds, uint8_t *h0, uint8_t *h1,
                  int32_t *nstates) {
    ORC_TRY(phase_run(*(Panel *)panel, T, gl, flat, *p, kpt, gp0, gp1, hds, h0, h1, nstates));
}

int orc_pbwt_select(void *panel, int32_t T, const uint8_t *hap_alleles, const uint8_t *stored,
                    const int32_t *start_idx, int32_t depth, float modulo_cm,
                    int32_t *out_off, int32_t *out_cnt, int32_t *out_ids, int64_t cap) {
    ORC_TRY({
        Panel &P = *(Panel *)panel;
        std::vector<std::vector<uint8_t>> haps(2 * T);
        for (int h = 0; h < 2 * T; h++) haps[h].assign(hap_alleles + (size_t)h * P.L, hap_alleles + (size_t)(h + 1) * P.L);
        std::vector<bool> st(P.n_com_hq);
        for (int i = 0; i < P.n_com_hq; i++) st[i] = stored[i];
        std::vector<int> start(start_idx, start_idx + 2 * T);
        Sweep S(P, haps, T, depth, modulo_cm);
        S.run(st, start);
        int64_t o = 0;
        for (int t = 0; t < T; t++)
            for (int d = 0; d < depth; d++) {
                const auto &v = S.states[t][d];
                out_off[t * depth + d] = (int32_t)o;
                out_cnt[t * depth + d] = (int32_t)v.size();
                if (o + (int64_t)v.size() > cap) throw OrcError("pbwt_select: output capacity exceeded");
                for (int x : v) out_ids[o++] = x;
            }
    });
}

int orc_pbwt_trace(void *panel, int32_t T, const uint8_t *hap_alleles, const uint8_t *stored,
                   const int32_t *start_idx, int32_t depth, float modulo_cm, int32_t *rows, int64_t cap,
                   int64_t *n_rows) {
    ORC_TRY({
        Panel &P = *(Panel *)panel;
        std::vector<std::vector<uint8_t>> haps(2 * T);
        for (int h = 0; h < 2 * T; h++) haps[h].assign(hap_alleles + (size_t)h * P.L, hap_alleles + (size_t)(h + 1) * P.L);
        std::vector<bool> st(P.n_com_hq);
        for (int i = 0; i < P.n_com_hq; i++) st[i] = stored[i];
        std::vector<int> start(start_idx, start_idx + 2 * T);
        std::vector<int32_t> tr;
        Sweep S(P, haps, T, depth, modulo_cm);
        S.trace = &tr;
        S.run(st, start);
        *n_rows = (int64_t)tr.size() / 6;
        if ((int64_t)tr.size() > cap * 6) throw OrcError("pbwt_trace: capacity");
        std::copy(tr.begin(), tr.end(), rows);
    });
}

int orc_pbwt_groups(void *panel, int32_t depth, float modulo_cm, int32_t kpbwt, float *eff, int32_t *grp,
                    int32_t cap, int32_t *n_grp) {
    ORC_TRY({
        std::vector<int> g; float em;
        pbwt_groups(*(Panel *)panel, depth, modulo_cm, kpbwt, em, g);
        *eff = em; *n_grp = (int32_t)g.size();
        if ((int)g.size() > cap) throw OrcError("pbwt_groups: capacity");
        for (size_t i = 0; i < g.size(); i++) grp[i] = g[i];
    });
}

static void condset_out(CondSet &C, int32_t *out_poly, int32_t *out_npoly, uint8_t *out_hvar, int64_t cap, float *out_t) {
    *out_npoly = (int32_t)C.poly.size();
    for (size_t i = 0; i < C.poly.size(); i++) out_poly[i] = (int32_t)C.poly[i];
    const int64_t stride = (int64_t)C.Hvar.n_cols / 8;
    if ((int64_t)C.poly.size() * stride > cap) throw OrcError("condset: hvar capacity");
    for (size_t l = 0; l < C.poly.size(); l++)
        std::memcpy(out_hvar + l * stride, &C.Hvar.bytes[l * stride], stride);
    for (size_t i = 0; i < C.t.size(); i++) out_t[i] = C.t[i];
}

int orc_condset(void *panel, int32_t depth, const int32_t *off, const int32_t *cnt, const int32_t *ids,
                int32_t kpbwt, int32_t ne, int32_t *out_states, int32_t *out_K, int32_t *out_poly,
                int32_t *out_npoly, uint8_t *out_hvar, int64_t cap, float *out_t) {
    ORC_TRY({
        Panel &P = *(Panel *)panel;
        std::vector<std::vector<int>> lists(depth);
        for (int d = 0; d < depth; d++) lists[d].assign(ids + off[d], ids + off[d] + cnt[d]);
        std::vector<int> st = compact_states(lists, kpbwt);
        *out_K = (int32_t)st.size();
        for (size_t i = 0; i < st.size(); i++) out_states[i] = st[i];
        CondSet C(P, ne, 1e-12f, 1e-4f);
        C.build(st);
        condset_out(C, out_poly, out_npoly, out_hvar, cap, out_t);
    });
}

int orc_condset_from_states(void *panel, const int32_t *states, int32_t K, int32_t ne, int32_t *out_poly,
                            int32_t *out_npoly, uint8_t *out_hvar, int64_t cap, float *out_t) {
    ORC_TRY({
        Panel &P = *(Panel *)panel;
        CondSet C(P, ne, 1e-12f, 1e-4f);
        C.build(std::vector<int>(states, states + K));
        condset_out(C, out_poly, out_npoly, out_hvar, cap, out_t);
    });
}

int orc_impute_hap(void *panel, const int32_t *states, int32_t K, int32_t ne, float err_imp, const float *HL,
                   const uint8_t *flat, float *HP) {
    ORC_TRY({
        Panel &P = *(Panel *)panel;
        CondSet C(P, ne, std::clamp(err_imp, 1e-12f, 1e-3f), 1e-4f);
        C.build(std::vector<int>(states, states + K));
        ImputeHMM H(&C);
        std::vector<float> hl(HL, HL + 2 * P.L), hp(2 * P.L, 0.0f);
        std::vector<char> fl(flat, flat + P.L);
        H.run(hl, fl, hp);
        std::memcpy(HP, hp.data(), sizeof(float) * 2 * P.L);
    });
}

int orc_hap_likelihoods(int32_t L, const uint8_t *gl, const uint8_t *flat, const uint8_t *other, int32_t mode,
                        float min_gl, float *HL) {
    ORC_TRY({
        std::vector<char> fl(flat, flat + L);
        std::vector<float> hl(HL, HL + 2 * L);
        if (mode == 0) hl_init(L, gl, fl, min_gl, hl);
        else hl_make(L, gl, fl, std::vector<uint8_t>(other, other + L), min_gl, hl);
        std::memcpy(HL, hl.data(), sizeof(float) * 2 * L);
    });
}

int orc_rephase(void *panel, const int32_t *states, int32_t K, int32_t ne, float err_phase, const uint8_t *flat,
                uint8_t *h0, uint8_t *h1, uint32_t seed, int32_t it, int32_t tgt) {
    ORC_TRY({
        Panel &P = *(Panel *)panel;
        CondSet C(P, ne, 1e-12f, err_phase);
        C.build(std::vector<int>(states, states + K));
        PhaseHMM D(&C);
        Rng rng; rng.mode = ORC_RNG_COUNTER; rng.seed = seed; rng.it = it;
        std::vector<uint8_t> H0(h0, h0 + P.L), H1(h1, h1 + P.L);
        std::vector<char> fl(flat, flat + P.L);
        D.rephase(H0, H1, fl, rng, tgt);
        std::memcpy(h0, H0.data(), P.L);
        std::memcpy(h1, H1.data(), P.L);
    });
}

int orc_phase_trace(void *panel, const int32_t *states, int32_t K, int32_t ne, float err_phase, const uint8_t *flat,
                    const uint8_t *h0, const uint8_t *h1, uint32_t seed, int32_t it, int32_t tgt, int32_t cap_segs,
                    int32_t cap_miss, int32_t *n_segs, int32_t *n_miss, float *seg_sums, float *trans_hap,
                    float *of1) {
    ORC_TRY({
        Panel &P = *(Panel *)panel;
        CondSet C(P, ne, 1e-12f, err_phase);
        C.build(std::vector<int>(states, states + K));
        PhaseHMM D(&C);
        Rng rng; rng.mode = ORC_RNG_COUNTER; rng.seed = seed; rng.it = it;
        std::vector<uint8_t> H0(h0, h0 + P.L), H1(h1, h1 + P.L);
        std::vector<char> fl(flat, flat + P.L);
        D.rephase(H0, H1, fl, rng, tgt);
        *n_segs = (int32_t)D.n_segs;
        *n_miss = (int32_t)D.n_miss;
        if ((int32_t)D.n_segs > cap_segs || (int32_t)D.n_miss > cap_miss) throw OrcError("orc_phase_trace: capacity");
        for (unsigned s = 0; s < D.n_segs; s++)
            for (int h = 0; h < HAPN; h++)
                seg_sums[s * HAPN + h] = D.phasingProbSum[s * HAPN + h] / D.phasingProbSumSum[s];
        std::copy(D.hp_trace.begin(), D.hp_trace.end(), trans_hap);
        std::copy(D.imputeProbOf1s.begin(), D.imputeProbOf1s.end(), of1);
    });
}

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) { philox(ctr, key, out); }

uint32_t orc_mt19937_nth(uint32_t seed, int64_t n) {
    std::mt19937 e(seed);
    uint32_t x = 0;
    for (int64_t i = 0; i < n; i++) x = e();
    return x;
}

float orc_unphred(int32_t i) { return UNPHRED[i & 255]; }

/* One sampling call of the INIT selection over the population 0..n_pop-1:
 * std::sample driven by std::mt19937(seed) (reference mode) or Algorithm S
 * with counter draws keyed (seed, iteration 0, INIT purpose, tgt, 0, 1, ...). */
int orc_sample(int32_t n_pop, int32_t k, int32_t rng_mode, uint32_t seed, int32_t tgt, int32_t *out, int32_t *n_out) {
    ORC_TRY({
        Rng rng; rng.mode = rng_mode; rng.set_seed(seed); rng.it = 0;
        std::vector<int> pop(n_pop), sampled;
        std::iota(pop.begin(), pop.end(), 0);
        if (rng_mode == ORC_RNG_REFERENCE) std::sample(pop.begin(), pop.end(), std::back_inserter(sampled), k, rng.eng);
        else { uint32_t ctr = 0; counter_sample(pop, k, sampled, rng, (uint32_t)tgt, ctr); }
        *n_out = (int32_t)sampled.size();
        std::copy(sampled.begin(), sampled.end(), out);
    });
}

/* The INIT-iteration state set of one target (haplotype_set.cpp:108-169, 775-833). */
int orc_init_states(void *panel, const uint8_t *gl, const uint8_t *flat, int32_t kinit, int32_t rng_mode,
                    uint32_t seed, int32_t tgt, int32_t *out, int32_t cap, int32_t *n_out) {
    ORC_TRY({
        const Panel &P = *(const Panel *)panel;
        Target g;
        g.L = P.L; g.gl = gl;
        g.flat.assign(flat, flat + P.L);
        Rng rng; rng.mode = rng_mode; rng.set_seed(seed); rng.it = 0;
        init_rare_tar(P, g);
        init_selection_one(P, g, (size_t)tgt, std::min(kinit, P.N), rng);
        if ((int)g.init_states.size() > cap) throw OrcError("init_states: capacity");
        *n_out = (int32_t)g.init_states.size();
        std::copy(g.init_states.begin(), g.init_states.end(), out);
    });
}

/* GP storage over n_iter main iterations of one target, then the final
 * normalisation and best guess (genotype.cpp:168-244): HP0/HP1 [n_iter][2L]. */
int orc_gp_run(int32_t L, int32_t n_iter, const float *HP0, const float *HP1, float *gp0, float *gp1, uint8_t *hds,
               uint8_t *h0, uint8_t *h1) {
    ORC_TRY({
        Target g;
        g.L = L;
        for (int i = 0; i < n_iter; i++) {
            std::vector<float> a(HP0 + (size_t)i * 2 * L, HP0 + (size_t)(i + 1) * 2 * L);
            std::vector<float> b(HP1 + (size_t)i * 2 * L, HP1 + (size_t)(i + 1) * 2 * L);
            g.store(a, b);
        }
        finalize_target(g, L, gp0, gp1, hds, h0, h1);
    });
}

/* getFloat draws: counter mode, n draws keyed (seed, it, purpose, a, b = 0..n-1);
 * reference mode, the first n draws of std::mt19937(seed). */
int orc_draws(uint32_t seed, int32_t it, int32_t purpose, int32_t a, int32_t n, int32_t rng_mode, float *out) {
    ORC_TRY({
        Rng rng; rng.mode = rng_mode; rng.set_seed(seed); rng.it = it;
        for (int i = 0; i < n; i++) out[i] = (float)rng.getFloat(purpose, (uint32_t)a, (uint32_t)i);
    });
}

/* genotype_writer::writeGenotypes per-site fields for diploid and haploid targets
 * (phase/src/io/genotype_writer.cpp:109-190): DS, GP with the writer's
 * floor-then-raise rounding, phased GT, INFO / AF / RAF.  A target whose
 * genotype is not stored at a site enters as gp = (1, 0). */
int orc_output_fields_ploidy(int32_t T, int32_t L, const int32_t *ploidy, const float *gp0s, const float *gp1s,
                             const uint8_t *hdss, const uint32_t *cref, const uint32_t *calt, float *ds_out,
                             float *gp_out, uint8_t *gt_out, float *info_out, float *af_out, float *raf_out) {
    ORC_TRY(
        int n_tar_haps = 0, n_hap = 0;
        for (int i = 0; i < T; i++) {
            const int pl = ploidy ? ploidy[i] : 2;
            n_tar_haps += pl;
            n_hap += pl == 1;
        }
        /* format ploidy (genotype_reader.cpp:553-554): 2, 1, or -2 when mixed */
        const int fploidy = n_hap == 0 ? 2 : (n_hap == T ? 1 : -2);
        for (int l = 0; l < L; l++) {
            float ds_sum = 0.0, ds2_sum = 0.0, ds4_sum = 0.0;
            for (int i = 0; i < T; i++) {
                const size_t o = (size_t)i * L + l;
                const float gp0 = gp0s[o], gp1 = gp1s[o];
                const bool hap = ploidy && ploidy[i] == 1;
                const float gp2 = hap ? 0.0f : std::clamp(1.0f - gp1 - gp0, 0.0f, 1.0f);
                const float ds = hap ? gp1 : gp1 + 2.0f * gp2;
                uint8_t g[2];
                if (hap) {
                    g[0] = gp0 < gp1;
                    g[1] = 255;   /* bcf_int32_vector_end */
                } else if (gp1 > gp0 && gp1 > gp2) {
                    const bool hds = hdss[o] != 0;
                    g[hds] = 1;
                    g[1 - hds] = 0;
                } else {
                    g[0] = g[1] = gp0 < gp2;
                }
                float P[3];
                ds_out[o] = std::roundf(ds * 1000.0) / 1000.0;
                P[0] = floorf(gp0 * 1000.0) / 1000.0;
                P[1] = floorf(gp1 * 1000.0) / 1000.0;
                std::map<float, float *> mp;
                mp.insert(std::make_pair(1.0f - (gp0 - P[0]), &P[0]));
                mp.insert(std::make_pair(1.0f - (gp1 - P[1]), &P[1]));
                const int n_gp = hap ? 2 : 3;
                if (!hap) {
                    P[2] = floorf(std::max(1.0f - (P[0] + P[1]), 0.0f) * 1000.0) / 1000.0;
                    mp.insert(std::make_pair(1.0f - (gp2 - P[2]), &P[2]));
                } else {
                    P[2] = std::numeric_limits<float>::quiet_NaN();   /* bcf_float_vector_end */
                }
                for (auto it = mp.begin(); it != mp.end() && std::accumulate(P, P + n_gp, 0.0f) < 0.9999f; ++it)
                    *it->second += 0.001f;
                std::memcpy(gp_out + 3 * o, P, sizeof(P));
                gt_out[2 * o] = g[0];
                gt_out[2 * o + 1] = g[1];
                ds_sum += ds;
                if (fploidy == 2) {
                    ds2_sum += ds * ds;
                    ds4_sum += gp1 + 4.0 * gp2;
                }
            }
            const float raf = calt[l] * 1.0f / (calt[l] + cref[l]);
            const float af = ds_sum / n_tar_haps;
            float info = (fploidy == 2 && af > 0.0 && af < 1.0)
                             ? (float)(1.0 - (ds4_sum - ds2_sum) / (n_tar_haps * af * (1.0 - af)))
                             : 1.0f;
            info = (info < 0.0f) ? 0.0f : info;
            info = roundf(info * 1000.0) / 1000.0;
            info_out[l] = info;
            af_out[l] = af;
            raf_out[l] = raf;
        })
}

int orc_output_fields(int32_t T, int32_t L, const float *gp0s, const float *gp1s, const uint8_t *hdss,
                      const uint32_t *cref, const uint32_t *calt, float *ds_out, float *gp_out, uint8_t *gt_out,
                      float *info_out, float *af_out, float *raf_out) {
    return orc_output_fields_ploidy(T, L, nullptr, gp0s, gp1s, hdss, cref, calt, ds_out, gp_out, gt_out, info_out,
                                    af_out, raf_out);
}

} // extern "C"
