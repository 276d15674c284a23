// gh_select.hip -- learned conditioning-state selection (no counterpart in the
// reference): per-candidate features of a target's PBWT pool, top-K by the
// network's predicted posterior mass with a fallback to the whole pool, and the
// training labels taken from the forward-backward posterior mass.
//
//   target_mask_kernel  once per run: per target, bit masks over the common
//                       sites (words of 64 common rows, the transposed panel's
//                       layout): has data, confident RR, AA, RA genotype.
//   feature_kernel      one workgroup per target: packs the current H0/H1 at the
//                       common sites into LDS (ballot), marks panel haplotypes
//                       that carry a rare allele the target carries, then one
//                       wave per candidate streams the candidate's transposed
//                       panel row (coalesced 64-bit words) against them.
//   topk_kernel         one workgroup per target: bitonic sort of the pool by
//                       predicted score, mass coverage of the top k, then the
//                       kept ids sorted ascending (the order the conditioning
//                       set builder expects).
//   label_kernel        log posterior mass of every pool candidate.
#include <algorithm>
#include <cstring>
#include <vector>

#include "gh_device.hpp"
#include "gh_internal.hpp"
#include "gh_select.hpp"

namespace gh {

/* ------------------------------------------------------------------ masks */
__global__ __launch_bounds__(64) void target_mask_kernel(const uint8_t *__restrict__ gl,
                                                         const uint8_t *__restrict__ flat,
                                                         const int32_t *__restrict__ c2t, int L, int Lc, int Wc,
                                                         uint64_t *__restrict__ tmask) {
    const int wi = blockIdx.x, t = blockIdx.y, lane = threadIdx.x;
    const int c = wi * 64 + lane;
    bool d = false, rr = false, aa = false, ra = false;
    if (c < Lc) {
        const int site = c2t[c];
        const size_t o = (size_t)t * L + site;
        if (!flat[o]) {
            d = true;
            const uint8_t *g = gl + 3 * o;   // phred-scaled, normalised to min 0
            rr = g[0] == 0 && g[1] >= FEAT_CONF_PHRED && g[2] >= FEAT_CONF_PHRED;
            ra = g[1] == 0 && g[0] >= FEAT_CONF_PHRED && g[2] >= FEAT_CONF_PHRED;
            aa = g[2] == 0 && g[0] >= FEAT_CONF_PHRED && g[1] >= FEAT_CONF_PHRED;
        }
    }
    const uint64_t md = __ballot(d), mr = __ballot(rr), ma = __ballot(aa), mh = __ballot(ra);
    if (lane == 0) {
        uint64_t *m = tmask + (size_t)t * 4 * Wc;
        m[wi] = md;
        m[Wc + wi] = mr;
        m[2 * Wc + wi] = ma;
        m[3 * Wc + wi] = mh;
    }
}

/* ------------------------------------------------------------------ features */
// Longest run of set bits in z (z != all ones is not required).
__device__ __forceinline__ int longest_ones(uint64_t z) {
    if (z == ~0ull) return 64;
    // A_k = bits that start a run of >= k ones; A_{a+b} = A_a & (A_b >> a)
    const uint64_t p1 = z, p2 = p1 & (p1 >> 1), p4 = p2 & (p2 >> 2), p8 = p4 & (p4 >> 4), p16 = p8 & (p8 >> 8),
                   p32 = p16 & (p16 >> 16);
    uint64_t acc = ~0ull;
    int len = 0;
    const uint64_t P[6] = {p32, p16, p8, p4, p2, p1};
#pragma unroll
    for (int i = 0; i < 6; i++) {
        const int k = 32 >> i;
        const uint64_t cand = acc & (P[i] >> len);
        if (cand) { acc = cand; len += k; }
    }
    return len;
}

// Run state over consecutive chunks of n <= 64 words; z = words that match.
struct RunState {
    int cur = 0, best = 0;
    __device__ __forceinline__ void push(uint64_t z, int n) {
        const uint64_t vm = n == 64 ? ~0ull : ((1ull << n) - 1ull);
        z &= vm;
        if (z == vm) { cur += n; best = max(best, cur); return; }
        const int lead = __builtin_ctzll(~z);
        best = max(best, cur + lead);
        best = max(best, longest_ones(z));
        const int hz = 63 - __builtin_clzll(~z & vm);
        cur = n - 1 - hz;
    }
};

__device__ __forceinline__ float pick32(int lane, const float *f) {
    float v = 0.0f;
#pragma unroll
    for (int i = 0; i < MLP_FEAT; i++) v = lane == i ? f[i] : v;
    return v;
}

__global__ __launch_bounds__(FEAT_NT) void feature_kernel(FeatArgs a) {
    extern __shared__ uint64_t fs[];
    const int Wc = a.Wc, Wn = (a.N + 63) / 64;
    uint64_t *s_h0 = fs, *s_h1 = fs + Wc, *s_d = fs + 2 * Wc, *s_rr = fs + 3 * Wc, *s_aa = fs + 4 * Wc,
             *s_ra = fs + 5 * Wc, *bm0 = fs + 6 * Wc, *bm1 = bm0 + Wn;
    __shared__ float s_cnt[4];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = FEAT_NT / 64;
    const int i = blockIdx.x, t = a.targets[i];
    const uint8_t *H0 = a.H + (size_t)(2 * t) * a.L, *H1 = a.H + (size_t)(2 * t + 1) * a.L;
    const uint64_t *tm = a.tmask + (size_t)t * 4 * Wc;
    // current haplotypes at the common sites, one ballot per 64 common rows
    for (int c0 = wv * 64; c0 < Wc * 64; c0 += nw * 64) {
        const int c = c0 + lane;
        bool b0 = false, b1 = false;
        if (c < a.Lc) {
            const int site = a.c2t[c];
            b0 = H0[site] != 0;
            b1 = H1[site] != 0;
        }
        const uint64_t m0 = __ballot(b0), m1 = __ballot(b1);
        if (lane == 0) { s_h0[c0 >> 6] = m0; s_h1[c0 >> 6] = m1; }
    }
    for (int w = tid; w < Wc; w += FEAT_NT) {
        s_d[w] = tm[w];
        s_rr[w] = tm[Wc + w];
        s_aa[w] = tm[2 * Wc + w];
        s_ra[w] = tm[3 * Wc + w];
    }
    for (int w = tid; w < 2 * Wn; w += FEAT_NT) bm0[w] = 0ull;
    __syncthreads();
    // panel haplotypes sharing a rare allele with H0 / H1
    for (int l = tid; l < a.L; l += FEAT_NT) {
        const uint8_t f = a.flags[l];
        if (f & SF_COMMON) continue;
        const uint8_t maj = (f & SF_MAJOR) ? 1 : 0;
        const bool r0 = H0[l] != maj, r1 = H1[l] != maj;
        if (!r0 && !r1) continue;
        for (int p = a.rare_ptr[l]; p < a.rare_ptr[l + 1]; p++) {
            const int c = a.rare_idx[p];
            const uint64_t m = 1ull << (c & 63);
            if (r0) atomicOr((unsigned long long *)&bm0[c >> 6], (unsigned long long)m);
            if (r1) atomicOr((unsigned long long *)&bm1[c >> 6], (unsigned long long)m);
        }
    }
    if (wv == 0) {
        float nd = 0, nrr = 0, naa = 0, nra = 0;
        for (int w = lane; w < Wc; w += 64) {
            nd += __popcll(s_d[w]);
            nrr += __popcll(s_rr[w]);
            naa += __popcll(s_aa[w]);
            nra += __popcll(s_ra[w]);
        }
        nd = wave_sum(nd); nrr = wave_sum(nrr); naa = wave_sum(naa); nra = wave_sum(nra);
        if (lane == 0) { s_cnt[0] = nd; s_cnt[1] = nrr; s_cnt[2] = naa; s_cnt[3] = nra; }
    }
    __syncthreads();
    const float inv_d = 1.0f / fmaxf(s_cnt[0], 1.0f), inv_rr = 1.0f / fmaxf(s_cnt[1], 1.0f),
                inv_aa = 1.0f / fmaxf(s_cnt[2], 1.0f), inv_ra = 1.0f / fmaxf(s_cnt[3], 1.0f);
    const float inv_lc = 1.0f / (float)max(a.Lc, 1), inv_wc = 1.0f / (float)max(Wc, 1);
    const int K = a.Kpool[t];
    const int tail = a.Lc - 64 * (Wc - 1);   // valid bits of the last word
    const uint64_t lastm = tail >= 64 ? ~0ull : ((1ull << tail) - 1ull);
    const int32_t *ins = a.ins + (size_t)t * a.kstride;
    const int32_t *isite = a.ins_site + (size_t)t * a.kstride;
    float *Xo = a.X + ((size_t)a.row0 + (size_t)i * a.rstride) * MLP_FEAT;
    const float inv_k = 1.0f / (float)max(K, 1);
    // sites of common words [lo, hi) clipped to the chunk
    auto wsites = [&](int lo, int hi) -> float {
        lo = max(lo, 0);
        hi = min(hi, Wc);
        return (float)max(min(64 * hi, a.Lc) - 64 * lo, 0);
    };
    for (int j = wv; j < K; j += nw) {
        const int v = ins[j];
        const int id = v & 0xFFFFF;
        const int cw = max(a.crow[isite[j]], 0) >> 6;
        const uint64_t *row = a.hm + (size_t)id * Wc;
        float f[MLP_FEAT];
        f[0] = (float)((v >> 20) & 63) / (float)a.depth;
        f[1] = (float)j * inv_k;
        f[2] = (float)((v >> 26) & 1);
        f[3] = (float)((v >> 27) & 1);
        float g0 = 0, g1 = 0, gD0 = 0, gD1 = 0;
        float nb0 = 0, nb1 = 0, nf0 = 0, nf1 = 0, fb0 = 0, fb1 = 0, ff0 = 0, ff1 = 0;
        float lD0 = 0, lD1 = 0, lRR = 0, lAA = 0, lRA = 0, lalt = 0, nD = 0, nRR = 0, nAA = 0, nRA = 0;
        RunState run0, run1;
        for (int c0 = 0; c0 < Wc; c0 += 64) {
            const int w = c0 + lane;
            const bool valid = w < Wc;
            uint64_t x0 = 1, x1 = 1;
            if (valid) {
                const uint64_t vm = w == Wc - 1 ? lastm : ~0ull;
                const uint64_t R = row[w] & vm;
                x0 = (R ^ s_h0[w]) & vm;
                x1 = (R ^ s_h1[w]) & vm;
                const uint64_t D = s_d[w];
                const float p0 = (float)__popcll(x0), p1 = (float)__popcll(x1);
                g0 += p0;
                g1 += p1;
                gD0 += __popcll(x0 & D);
                gD1 += __popcll(x1 & D);
                const int d = w - cw;
                if (d >= -FEAT_FAR && d < FEAT_FAR) {
                    lD0 += __popcll(x0 & D);
                    lD1 += __popcll(x1 & D);
                    lRR += __popcll(R & s_rr[w]);
                    lAA += __popcll(~R & s_aa[w] & vm);
                    lRA += __popcll(R & s_ra[w]);
                    lalt += __popcll(R);
                    nD += __popcll(D);
                    nRR += __popcll(s_rr[w]);
                    nAA += __popcll(s_aa[w]);
                    nRA += __popcll(s_ra[w]);
                    if (d < 0) {
                        fb0 += p0; fb1 += p1;
                        if (d >= -FEAT_NEAR) { nb0 += p0; nb1 += p1; }
                    } else {
                        ff0 += p0; ff1 += p1;
                        if (d < FEAT_NEAR) { nf0 += p0; nf1 += p1; }
                    }
                }
            }
            const int n = min(64, Wc - c0);
            run0.push(__ballot(valid && x0 == 0ull), n);
            run1.push(__ballot(valid && x1 == 0ull), n);
        }
        // identical words running back from cw-1 and forward from cw (up to FEAT_RUN)
        int back0, back1, fwd0, fwd1;
        {
            int w = cw - FEAT_RUN + lane;
            bool c0 = false, c1 = false;
            if (w >= 0 && w < Wc) {
                const uint64_t vm = w == Wc - 1 ? lastm : ~0ull;
                const uint64_t R = row[w] & vm;
                c0 = ((R ^ s_h0[w]) & vm) == 0ull;
                c1 = ((R ^ s_h1[w]) & vm) == 0ull;
            }
            uint64_t z0 = __ballot(c0), z1 = __ballot(c1);
            back0 = z0 == ~0ull ? 64 : __builtin_clzll(~z0);
            back1 = z1 == ~0ull ? 64 : __builtin_clzll(~z1);
            w = cw + lane;
            c0 = c1 = false;
            if (w < Wc) {
                const uint64_t vm = w == Wc - 1 ? lastm : ~0ull;
                const uint64_t R = row[w] & vm;
                c0 = ((R ^ s_h0[w]) & vm) == 0ull;
                c1 = ((R ^ s_h1[w]) & vm) == 0ull;
            }
            z0 = __ballot(c0);
            z1 = __ballot(c1);
            fwd0 = z0 == ~0ull ? 64 : __builtin_ctzll(~z0);
            fwd1 = z1 == ~0ull ? 64 : __builtin_ctzll(~z1);
        }
        wave_sum3(g0, g1, gD0);
        wave_sum3(gD1, nb0, nb1);
        wave_sum3(nf0, nf1, fb0);
        wave_sum3(fb1, ff0, ff1);
        wave_sum3(lD0, lD1, lRR);
        wave_sum3(lAA, lRA, lalt);
        wave_sum3(nD, nRR, nAA);
        nRA = wave_sum(nRA);
        const float s_nb = wsites(cw - FEAT_NEAR, cw), s_nf = wsites(cw, cw + FEAT_NEAR);
        const float s_fb = wsites(cw - FEAT_FAR, cw), s_ff = wsites(cw, cw + FEAT_FAR);
        const float s_seg = wsites(cw - FEAT_FAR, cw + FEAT_FAR);
        f[4] = g0 * inv_lc;
        f[5] = g1 * inv_lc;
        f[6] = gD0 * inv_d;
        f[7] = gD1 * inv_d;
        f[8] = (float)run0.best * inv_wc;
        f[9] = (float)run1.best * inv_wc;
        f[10] = nb0 / fmaxf(s_nb, 1.0f);
        f[11] = nb1 / fmaxf(s_nb, 1.0f);
        f[12] = nf0 / fmaxf(s_nf, 1.0f);
        f[13] = nf1 / fmaxf(s_nf, 1.0f);
        f[14] = fb0 / fmaxf(s_fb, 1.0f);
        f[15] = fb1 / fmaxf(s_fb, 1.0f);
        f[16] = ff0 / fmaxf(s_ff, 1.0f);
        f[17] = ff1 / fmaxf(s_ff, 1.0f);
        f[18] = lD0 / fmaxf(nD, 1.0f);
        f[19] = lD1 / fmaxf(nD, 1.0f);
        f[20] = (float)back0 * (1.0f / FEAT_RUN);
        f[21] = (float)back1 * (1.0f / FEAT_RUN);
        f[22] = (float)fwd0 * (1.0f / FEAT_RUN);
        f[23] = (float)fwd1 * (1.0f / FEAT_RUN);
        f[24] = lRR / fmaxf(nRR, 1.0f);
        f[25] = lAA / fmaxf(nAA, 1.0f);
        f[26] = lRA / fmaxf(nRA, 1.0f);
        f[27] = lalt / fmaxf(s_seg, 1.0f);
        f[28] = (float)((bm0[id >> 6] >> (id & 63)) & 1ull);
        f[29] = (float)((bm1[id >> 6] >> (id & 63)) & 1ull);
        f[30] = (float)cw * inv_wc;
        f[31] = s_seg * (1.0f / (64.0f * 2 * FEAT_FAR));
        const float out = pick32(lane, f);
        if (lane < MLP_FEAT) Xo[(size_t)j * MLP_FEAT + lane] = out;
    }
}

/* ------------------------------------------------------------------ top-k */
// Bitonic sort of s[0..n2) (n2 a power of two), descending when DESC.
template <bool DESC, typename T>
__device__ void bitonic(T *s, int n2) {
    for (int size = 2; size <= n2; size <<= 1)
        for (int stride = size >> 1; stride > 0; stride >>= 1) {
            for (int q = threadIdx.x; q < (n2 >> 1); q += blockDim.x) {
                const int p = 2 * q - (q & (stride - 1));
                const int r = p + stride;
                const bool up = (p & size) == 0;
                const T x = s[p], y = s[r];
                const bool sw = DESC ? ((x < y) == up) : ((x > y) == up);
                if (sw && x != y) { s[p] = y; s[r] = x; }
            }
            __syncthreads();
        }
}

__device__ __forceinline__ uint32_t ord_key(float y) {
    if (!(y == y)) y = -INFINITY;   // NaN sorts last
    const uint32_t u = __float_as_uint(y);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float key_score(uint32_t k) {
    const uint32_t u = (k & 0x80000000u) ? (k & 0x7FFFFFFFu) : ~k;
    return __uint_as_float(u);
}

__device__ __forceinline__ float block_sum(float v, float *red) {
    v = wave_sum(v);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    float s = 0.0f;
    for (int k = 0; k < nw; k++) s += red[k];
    return s;
}

// The ranking score of pool candidate j: the predicted log mass, less the
// PBWT-order prior (j / K is the candidate's insertion position: the PBWT's
// own selection of k states is the first k of the pool).
__device__ __forceinline__ float rank_score(const TopkArgs &a, const float *y, int j, int K) {
    return y[j] - a.prior * ((float)j / (float)K);
}

__global__ __launch_bounds__(TOPK_NT) void topk_kernel(TopkArgs a) {
    __shared__ uint64_t s_key[TOPK_MAX];
    __shared__ float red[TOPK_NT / 64];
    const int i = blockIdx.x, t = a.targets[i];
    const int K = a.Kpool[t];
    const int k = min(a.klearn[t], K);
    if (k <= 0 || k >= K) return;   // nothing to drop
    const float *y = a.y + (size_t)a.row0 + (size_t)i * a.rstride;
    int n2 = 1;
    while (n2 < K) n2 <<= 1;
    const bool quota = a.ins_site != nullptr;
    if (!quota) {
        // global ranking: (score desc, insertion order asc)
        for (int j = threadIdx.x; j < n2; j += TOPK_NT)
            s_key[j] = j < K ? (((uint64_t)ord_key(rank_score(a, y, j, K)) << 32) | (uint64_t)(0xFFFFFFFFu - (uint32_t)j))
                             : 0ull;
        __syncthreads();
        bitonic<true>(s_key, n2);
        for (int p = threadIdx.x; p < n2; p += TOPK_NT)
            s_key[p] = p < K ? (uint64_t)(0xFFFFFFFFu - (uint32_t)(s_key[p] & 0xFFFFFFFFull)) : ~0ull;
        __syncthreads();
    } else {
        // rank within the listing group by (score desc, insertion asc) ...
        const int32_t *sg = a.site_grp + (size_t)a.tcls[t] * a.L;
        const int32_t *isite = a.ins_site + (size_t)t * a.kstride;
        for (int j = threadIdx.x; j < n2; j += TOPK_NT) {
            uint64_t key = ~0ull;
            if (j < K) {
                const uint64_t g = (uint64_t)max(sg[isite[j]], 0);
                key = (g << 43) | ((uint64_t)(~ord_key(rank_score(a, y, j, K))) << 11) | (uint64_t)j;
            }
            s_key[j] = key;
        }
        __syncthreads();
        bitonic<false>(s_key, n2);
        // ... then serve the groups round robin: (rank in group, group, insertion)
        uint64_t nk[TOPK_MAX / TOPK_NT];
#pragma unroll
        for (int q = 0; q < TOPK_MAX / TOPK_NT; q++) {
            const int p = threadIdx.x + q * TOPK_NT;
            nk[q] = ~0ull;
            if (p < K) {
                const uint64_t x = s_key[p], g = x >> 43;
                int lo = 0, hi = p;   // first position of group g
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if ((s_key[mid] >> 43) < g) lo = mid + 1; else hi = mid;
                }
                nk[q] = ((uint64_t)(p - lo) << 43) | (g << 11) | (x & 0x7FFull);
            }
        }
        __syncthreads();
#pragma unroll
        for (int q = 0; q < TOPK_MAX / TOPK_NT; q++) {
            const int p = threadIdx.x + q * TOPK_NT;
            if (p < n2) s_key[p] = nk[q];
        }
        __syncthreads();
        bitonic<false>(s_key, n2);
        for (int p = threadIdx.x; p < n2; p += TOPK_NT) s_key[p] = p < K ? (s_key[p] & 0x7FFull) : ~0ull;
        __syncthreads();
    }
    // s_key[p] = insertion index of the p-th kept candidate (p < K)
    float top = 0.0f, all = 0.0f;
    for (int p = threadIdx.x; p < K; p += TOPK_NT) {
        const float m = fmaxf(expf(y[(int)s_key[p]]) - LABEL_EPS, 0.0f);
        all += m;
        top += p < k ? m : 0.0f;
    }
    top = block_sum(top, red);
    all = block_sum(all, red);
    const bool keep_pool = !(all > 0.0f) || top < a.min_mass * all;
    if (threadIdx.x == 0) {
        a.coverage[t] = all > 0.0f ? top / all : 0.0f;
        atomicAdd(keep_pool ? &a.counters[1] : &a.counters[0], 1);
    }
    if (keep_pool) return;   // plain PBWT selection stays
    const int32_t *ins = a.ins + (size_t)t * a.kstride;
    __syncthreads();
    for (int p = threadIdx.x; p < n2; p += TOPK_NT)
        s_key[p] = p < k ? (uint64_t)(ins[(int)s_key[p]] & 0xFFFFF) : ~0ull;
    __syncthreads();
    bitonic<false>(s_key, n2);
    int32_t *st = a.states + (size_t)t * a.kstride;
    for (int j = threadIdx.x; j < k; j += TOPK_NT) st[j] = (int32_t)s_key[j];
    if (threadIdx.x == 0) {
        a.Kdev[t] = k;
        a.items[t].K = k;
        a.items[t].W = (k + 63) / 64;
    }
}

/* ------------------------------------------------------------------ labels */
// Segment [lo, hi) in sites of a candidate listed at site s.
__device__ __forceinline__ void seg_sites(int s, const int32_t *crow, const int32_t *c2t, int Lc, int Wc, int &lo,
                                          int &hi) {
    const int cw = max(crow[s], 0) >> 6;
    const int wlo = max(cw - FEAT_FAR, 0), whi = min(cw + FEAT_FAR, Wc);
    lo = c2t[64 * wlo];
    hi = c2t[min(64 * whi, Lc) - 1] + 1;
}

__global__ __launch_bounds__(256) void win_kernel(WinArgs a) {
    const int t = a.targets[blockIdx.x];
    const int K = a.Kpool[t];
    const int32_t *st = a.states + (size_t)t * a.kstride;
    const int32_t *ins = a.ins + (size_t)t * a.kstride, *isite = a.ins_site + (size_t)t * a.kstride;
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        const int id = ins[j] & 0xFFFFF;
        int lo = 0, hi = K;   // position of id in the sorted pool
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (st[mid] < id) lo = mid + 1; else hi = mid;
        }
        int s0, s1;
        seg_sites(isite[j], a.crow, a.c2t, a.Lc, a.Wc, s0, s1);
        a.win_lo[(size_t)t * a.kstride + lo] = s0;
        a.win_hi[(size_t)t * a.kstride + lo] = s1;
    }
}

__global__ __launch_bounds__(256) void label_kernel(LabelArgs a) {
    const int i = blockIdx.x, t = a.targets[i];
    const CondItem c = a.items[t];
    const int K = c.K;
    const int32_t *st = a.states + c.state_off;
    const int32_t *poly = a.poly + c.site_off;
    const int32_t *ins = a.ins + (size_t)t * a.kstride;
    const float *m0 = a.mass0 + c.state_off, *m1 = a.mass1 + c.state_off;
    const size_t r0 = (size_t)a.row0 + (size_t)i * a.rstride;
    auto count_below = [&](int v) {   // polymorphic sites < v
        int lo = 0, hi = c.Lp;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (poly[mid] < v) lo = mid + 1; else hi = mid;
        }
        return lo;
    };
    for (int j = threadIdx.x; j < a.rstride; j += blockDim.x) {
        float lab = 0.0f, wt = 0.0f;
        if (j < K) {
            const int id = ins[j] & 0xFFFFF;
            int lo = 0, hi = K;   // lower_bound in the sorted state list
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (st[mid] < id) lo = mid + 1; else hi = mid;
            }
            const int np = count_below(a.win_hi[c.state_off + lo]) - count_below(a.win_lo[c.state_off + lo]);
            if (np > 0) {
                const float m = (m0[lo] + m1[lo]) / (2.0f * (float)np);
                lab = __logf((float)K * m + LABEL_EPS);
                wt = 1.0f;
            }
        }
        if (wt == 0.0f)
            for (int f = 0; f < MLP_FEAT; f++) a.X[(r0 + j) * MLP_FEAT + f] = 0.0f;
        a.Y[r0 + j] = lab;
        a.W[r0 + j] = wt;
    }
}

/* ------------------------------------------------------------------ launchers */
void target_mask_launch(gh_panel *P, int T, const uint8_t *gl, const uint8_t *flat, uint64_t *tmask,
                        hipStream_t s) {
    if (P->Wc == 0) return;
    hipLaunchKernelGGL(target_mask_kernel, dim3(P->Wc, T), dim3(64), 0, s, gl, flat, P->d_c2t.as<int32_t>(), P->L,
                       P->Lc, P->Wc, tmask);
    GH_HIP(hipGetLastError());
}

size_t feature_lds_bytes(const gh_panel *P) {
    return sizeof(uint64_t) * (6 * (size_t)P->Wc + 2 * (size_t)((P->N + 63) / 64));
}

void feature_launch(gh_panel *P, int n, const FeatArgs &a, hipStream_t s) {
    const size_t lds = feature_lds_bytes(P);
    GH_CHECK(lds <= FEAT_LDS_MAX, GH_ERR_CAPACITY,
             "learned selection: chunk too large for the feature kernel (common sites x 6 + haplotypes x 2 bits "
             "must fit 144 KiB of LDS)");
    static const bool attr = [] {
        GH_HIP(hipFuncSetAttribute((const void *)feature_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)FEAT_LDS_MAX));
        return true;
    }();
    (void)attr;
    if (n <= 0 || P->Wc == 0) return;
    hipLaunchKernelGGL(feature_kernel, dim3(n), dim3(FEAT_NT), lds, s, a);
    GH_HIP(hipGetLastError());
}

void topk_launch(int n, const TopkArgs &a, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(topk_kernel, dim3(n), dim3(TOPK_NT), 0, s, a);
    GH_HIP(hipGetLastError());
}

void label_launch(int n, const LabelArgs &a, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(label_kernel, dim3(n), dim3(256), 0, s, a);
    GH_HIP(hipGetLastError());
}

void win_launch(int n, const WinArgs &a, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(win_kernel, dim3(n), dim3(256), 0, s, a);
    GH_HIP(hipGetLastError());
}

/* ------------------------------------------------------------------ batch entry points */
void select_features_batch(gh_panel *P, int T, const uint8_t *H_h, const uint8_t *gl_h, const uint8_t *flat_h,
                           const int32_t *ins_h, const int32_t *site_h, const int32_t *K_h, int kstride, int depth,
                           float *X_h) {
    GH_CHECK(P && T > 0 && H_h && gl_h && flat_h && ins_h && K_h && X_h && kstride > 0 && depth > 0, GH_ERR_ARG,
             "gh_select_features: bad arguments");
    GH_CHECK(P->N < (1 << 20), GH_ERR_ARG, "gh_select_features: at most 2^20 panel haplotypes");
    for (int t = 0; t < T; t++) {
        GH_CHECK(K_h[t] >= 0 && K_h[t] <= kstride, GH_ERR_ARG, "gh_select_features: K out of range");
        for (int j = 0; j < K_h[t]; j++) {
            GH_CHECK((ins_h[(size_t)t * kstride + j] & 0xFFFFF) < P->N, GH_ERR_ARG, "gh_select_features: bad id");
            const int s = site_h[(size_t)t * kstride + j];
            GH_CHECK(s >= 0 && s < P->L && P->flags[s] & SF_COMMON, GH_ERR_ARG,
                     "gh_select_features: listing sites must be common sites");
        }
    }
    hipStream_t s = P->stream;
    const int L = P->L;
    DBuf H, gl, flat, ins, isite, K, X, tm, tg;
    std::vector<int32_t> tv(T);
    for (int t = 0; t < T; t++) tv[t] = t;
    FeatArgs a;
    a.hm = P->d_hm.as<uint64_t>(); a.Wc = P->Wc; a.Lc = P->Lc; a.N = P->N; a.L = L; a.depth = depth;
    a.kstride = kstride; a.rstride = kstride; a.c2t = P->d_c2t.as<int32_t>(); a.flags = P->d_flags.as<uint8_t>();
    a.rare_ptr = P->d_rare_ptr.as<int32_t>(); a.rare_idx = P->d_rare_idx.as<int32_t>();
    a.H = dalloc_copy(H, H_h, (size_t)2 * T * L, s);
    const uint8_t *gld = dalloc_copy(gl, gl_h, (size_t)3 * T * L, s);
    const uint8_t *fld = dalloc_copy(flat, flat_h, (size_t)T * L, s);
    tm.reserve(sizeof(uint64_t) * (size_t)T * 4 * std::max(P->Wc, 1));
    target_mask_launch(P, T, gld, fld, tm.as<uint64_t>(), s);
    a.tmask = tm.as<uint64_t>();
    a.targets = dalloc_copy(tg, tv.data(), (size_t)T, s);
    a.Kpool = dalloc_copy(K, K_h, (size_t)T, s);
    a.ins = dalloc_copy(ins, ins_h, (size_t)T * kstride, s);
    a.ins_site = dalloc_copy(isite, site_h, (size_t)T * kstride, s);
    a.crow = P->d_crow.as<int32_t>();
    X.reserve(sizeof(float) * (size_t)T * kstride * MLP_FEAT);
    GH_HIP(hipMemsetAsync(X.p, 0, sizeof(float) * (size_t)T * kstride * MLP_FEAT, s));
    a.X = X.as<float>();
    a.row0 = 0;
    feature_launch(P, T, a, s);
    GH_HIP(hipMemcpyAsync(X_h, X.p, sizeof(float) * (size_t)T * kstride * MLP_FEAT, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
}

void select_topk_batch(int T, const float *y_h, const int32_t *ins_h, const int32_t *K_h, const int32_t *klearn_h,
                       int kstride, float min_mass, int32_t *states_h, int32_t *Kout_h, float *cov_h) {
    GH_CHECK(T > 0 && y_h && ins_h && K_h && klearn_h && states_h && Kout_h && kstride > 0 && kstride <= TOPK_MAX,
             GH_ERR_ARG, "gh_select_topk: bad arguments");
    for (int t = 0; t < T; t++)
        GH_CHECK(K_h[t] >= 0 && K_h[t] <= kstride && klearn_h[t] >= 0, GH_ERR_ARG, "gh_select_topk: bad sizes");
    hipStream_t s = nullptr;
    GH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    DBuf y, ins, K, kl, st, Kd, it, cov, cnt, tg;
    std::vector<int32_t> tv(T);
    std::vector<CondItem> items(T);
    for (int t = 0; t < T; t++) {
        tv[t] = t;
        std::memset(&items[t], 0, sizeof(CondItem));
        items[t].K = K_h[t];
        items[t].W = (K_h[t] + 63) / 64;
    }
    TopkArgs a;
    a.prior = 0.0f;   // the batch call ranks by the scores alone
    a.targets = dalloc_copy(tg, tv.data(), (size_t)T, s);
    a.y = dalloc_copy(y, y_h, (size_t)T * kstride, s);
    a.ins = dalloc_copy(ins, ins_h, (size_t)T * kstride, s);
    a.Kpool = dalloc_copy(K, K_h, (size_t)T, s);
    a.Kdev = (int32_t *)a.Kpool;
    a.klearn = dalloc_copy(kl, klearn_h, (size_t)T, s);
    a.row0 = 0;
    a.kstride = kstride;
    a.rstride = kstride;
    a.min_mass = min_mass;
    a.states = dalloc_copy(st, states_h, (size_t)T * kstride, s);
    a.items = dalloc_copy(it, items.data(), (size_t)T, s);
    a.coverage = (float *)cov.reserve(sizeof(float) * T);
    GH_HIP(hipMemsetAsync(cov.p, 0, sizeof(float) * T, s));
    a.counters = (int32_t *)cnt.reserve(sizeof(int32_t) * 2);
    GH_HIP(hipMemsetAsync(cnt.p, 0, sizeof(int32_t) * 2, s));
    a.ins_site = nullptr; a.site_grp = nullptr; a.tcls = nullptr; a.L = 0;
    topk_launch(T, a, s);
    GH_HIP(hipMemcpyAsync(states_h, st.p, sizeof(int32_t) * (size_t)T * kstride, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(Kout_h, K.p, sizeof(int32_t) * (size_t)T, hipMemcpyDeviceToHost, s));
    if (cov_h) GH_HIP(hipMemcpyAsync(cov_h, cov.p, sizeof(float) * (size_t)T, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
}

}  // namespace gh
