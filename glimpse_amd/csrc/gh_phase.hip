// gh_phase.hip -- diploid phasing HMM and its Gibbs draws
// (phasing_hmm::reallocate / backward / forward / rephaseHaplotypes,
//  phase/src/models/phasing_hmm.h:115-423, phasing_hmm.cpp:65-302).
//
// phase_prep_kernel (one workgroup per target): classifies the polymorphic
// loci (peak het j = het ordinal mod 3, peak hom, flat het), assigns segments
// of three hets (a segment starts at het ordinals 0, 3, 6, ...), the flat-het
// ordinal, and counts segments / flat hets.
//
// phase_kernel<NW, KPT> (one workgroup of NW waves per target): thread t owns
// states t + 64*NW*i (i < KPT) and keeps their 8 hypothesis values in
// registers.  Per locus the 8 per-hypothesis sums are reduce-scattered with DPP
// inside each wave (lane l ends with hypothesis l & 7) and combined across the
// NW waves through one LDS exchange; TRANS_HAP's 64 sums are reduce-scattered
// the same way, in two halves of 32.  The locus update is specialised by mode
// (RUN / COLLAPSE) and emission kind, so a state costs about three VALU
// operations per hypothesis.  More waves per target shorten each wave's share
// of a locus; fewer let more targets share a CU within its register file,
// which wins when many targets are in flight; the launch shape comes from
// phase_shape (measured defaults) or the run's tuner.  The backward pass stores
// segment-start and flat-het messages (K x 8 floats, [state][hyp] with the
// state index interleaved as in registers) in HBM; the forward pass samples
// the segment diplotypes (TRANS_HAP + SAMPLE_DIP) and the flat hets, then the
// write-back updates H0/H1 (including the reference's FLAT_HET write-back
// that assigns H0 twice, phasing_hmm.cpp:281-284).
#include "gh_device.hpp"
#include "gh_internal.hpp"
#include "gh_phase.hpp"

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <type_traits>
#include <vector>

namespace gh {

// four waves per target (a workgroup that needs a whole CU's wave slots waits
// behind the other runs' long-lived workgroups when several runs share the GPU)
constexpr int PREP_NT = 256;

template <int NT>
__device__ __forceinline__ void block_scan3(int a, int b, int c, int &ea, int &eb, int &ec, int &ta, int &tb,
                                            int &tc) {
    __shared__ int ws[3][NT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = a, y = b, z = c;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int xa = __shfl_up(x, o), ya = __shfl_up(y, o), za = __shfl_up(z, o);
        if (lane >= o) { x += xa; y += ya; z += za; }
    }
    if (lane == 63) { ws[0][wid] = x; ws[1][wid] = y; ws[2][wid] = z; }
    __syncthreads();
    if (wid == 0) {
        int u = lane < NT / 64 ? ws[0][lane] : 0, v = lane < NT / 64 ? ws[1][lane] : 0,
            w = lane < NT / 64 ? ws[2][lane] : 0;
#pragma unroll
        for (int o = 1; o < NT / 64; o <<= 1) {
            int ua = __shfl_up(u, o), va = __shfl_up(v, o), wa = __shfl_up(w, o);
            if (lane >= o) { u += ua; v += va; w += wa; }
        }
        if (lane < NT / 64) { ws[0][lane] = u; ws[1][lane] = v; ws[2][lane] = w; }
    }
    __syncthreads();
    ta = ws[0][NT / 64 - 1]; tb = ws[1][NT / 64 - 1]; tc = ws[2][NT / 64 - 1];
    ea = (wid ? ws[0][wid - 1] : 0) + x - a;
    eb = (wid ? ws[1][wid - 1] : 0) + y - b;
    ec = (wid ? ws[2][wid - 1] : 0) + z - c;
    __syncthreads();
}

// Locus classes per polymorphic site: 0 = not a phasing locus, 1 = peak het,
// 2 = peak hom, 3 = flat het  (phasing_hmm.cpp:71-96).
__device__ __forceinline__ int locus_class(const uint8_t *H0, const uint8_t *H1, const uint8_t *flat,
                                           const uint8_t *flags, int site) {
    const bool a0 = H0[site], a1 = H1[site];
    const bool data = !flat[site] && !(flags[site] & SF_LQ);
    if (data) return a0 != a1 ? 1 : 2;
    return a0 != a1 ? 3 : 0;
}

__global__ __launch_bounds__(PREP_NT) void phase_prep_kernel(PhItem *__restrict__ items,
                                                            const CondItem *__restrict__ cond,
                                                            const int32_t *__restrict__ poly_all,
                                                            const uint8_t *__restrict__ H0_all,
                                                            const uint8_t *__restrict__ H1_all,
                                                            const uint8_t *__restrict__ flat_all,
                                                            const uint8_t *__restrict__ flags,
                                                            Locus *__restrict__ loci_all) {
    PhItem &it = items[blockIdx.x];
    const CondItem c = cond[it.cond];
    const int Lp = c.Lp, tid = threadIdx.x;
    const int32_t *poly = poly_all + c.site_off;
    const uint8_t *H0 = H0_all + it.h0_off, *H1 = H1_all + it.h1_off, *flat = flat_all + it.flat_off;
    Locus *loci = loci_all + it.locus_off;
    const int C = (Lp + PREP_NT - 1) / PREP_NT;
    const int q0 = min(Lp, tid * C), q1 = min(Lp, q0 + C);
    int nv = 0, nh = 0, nm = 0;
    for (int q = q0; q < q1; q++) {
        const int cl = locus_class(H0, H1, flat, flags, poly[q]);
        nv += cl != 0; nh += cl == 1; nm += cl == 3;
    }
    int ev, eh, em, tv, th, tm;
    block_scan3<PREP_NT>(nv, nh, nm, ev, eh, em, tv, th, tm);
    for (int q = q0; q < q1; q++) {
        const int site = poly[q];
        const int cl = locus_class(H0, H1, flat, flags, site);
        if (!cl) continue;
        Locus L;
        L.rel = q;
        L.abs = site;
        L.alt = H0[site];
        L.miss = cl == 3 ? em : -1;
        if (cl == 1) {
            L.type = (int8_t)(eh % 3);
            L.seg = eh / 3;
        } else {
            L.type = cl == 2 ? (int8_t)-1 : (int8_t)-2;
            L.seg = eh == 0 ? 0 : (eh - 1) / 3;
        }
        loci[ev] = L;
        ev++;
        eh += cl == 1;
        em += cl == 3;
    }
    if (tid == 0) {
        it.n_loci = tv;
        it.n_segs = th == 0 ? 1 : (th - 1) / 3 + 1;
        it.n_miss = tm;
    }
}

// EMIT tables (phasing_hmm.cpp:38-56): hypothesis h carries allele 0 at het j
// when bit (2-j) of h is set.
__device__ __forceinline__ float emit_het(int j, int h, bool allele, float ee, float ed) {
    const bool hyp0 = (h >> (2 - j)) & 1;
    return (hyp0 != allele) ? ee : ed;
}

// ((v0+v4)+(v1+v5)) + ((v2+v6)+(v3+v7)) -- the reference's horizontal_add order
__device__ __forceinline__ float hadd8(const float *v) {
    return ((v[0] + v[4]) + (v[1] + v[5])) + ((v[2] + v[6]) + (v[3] + v[7]));
}

__device__ __forceinline__ float transition(const double *cm, int prev, int next, double nrho) {
    double x = -expm1(nrho * (cm[next] - cm[prev]));
    const double one_l = 1.0 - 1e-7;
    x = x < 1e-7 ? 1e-7 : (x > one_l ? one_l : x);
    return (float)x;
}

constexpr int PH_TILE = 64;

template <int NW, int KPT>
// shapes of up to four slots keep four waves per SIMD (128 VGPRs; the 4-slot
// shapes spill a few registers of the per-target set-up and segment-end code,
// not of the per-locus update: C5 phasing 8% faster than at three waves)
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(KPT <= 4 ? 4 : 1))) void phase_kernel(const PhItem *__restrict__ items, const CondItem *__restrict__ cond,
                                                       const uint64_t *__restrict__ hvar_all,
                                                       const Locus *__restrict__ loci_all,
                                                       const int32_t *__restrict__ pos_all,
                                                       const double *__restrict__ cm, double nrho,
                                                       float *__restrict__ seg_all, float *__restrict__ segsum_all,
                                                       float *__restrict__ imp_all, float *__restrict__ impsum_all,
                                                       float *__restrict__ of1_all, int32_t *__restrict__ dip_all,
                                                       uint8_t *__restrict__ H0_all, uint8_t *__restrict__ H1_all,
                                                       int L, float ee, float ed, uint32_t seed, uint32_t iter,
                                                       int32_t *__restrict__ status,
                                                       const int32_t *__restrict__ order, GraphGate gate) {
    constexpr int PH_NT = 64 * NW;
    extern __shared__ __align__(16) uint64_t s_rows[];   // [PH_TILE][W] Hvar rows of the current tile
    __shared__ float lds8[2][NW * 8];
    __shared__ float red64[NW][64];
    __shared__ float hprobs[64];
    __shared__ int dip_sh;
    __shared__ int s_rel[PH_TILE], s_seg[PH_TILE], s_miss[PH_TILE];
    __shared__ signed char s_type[PH_TILE];
    __shared__ unsigned char s_alt[PH_TILE], s_first[PH_TILE], s_last[PH_TILE];
    __shared__ float s_yin[PH_TILE], s_yout[PH_TILE];
    if (gate.key) {
        if (*gate.key != gate.want || (int)blockIdx.x >= *gate.n) return;
        iter = *gate.iter;
    }
    const PhItem it = wave_uniform(items[order ? order[blockIdx.x] : (int)blockIdx.x]);   // optional item order
    const CondItem c = wave_uniform(cond[it.cond]);
    const int K = c.K, W = c.W, tid = threadIdx.x, lane = tid & 63, wid = __builtin_amdgcn_readfirstlane(tid >> 6);
    const uint64_t *hv = hvar_all + c.hvar_off;
    const Locus *loci = loci_all + it.locus_off;
    const int nloc = it.n_loci;
    float *segp = seg_all + it.seg_off;
    float *segs = segsum_all + (size_t)it.segsum_off;
    float *impp = imp_all + it.imp_off;
    float *imps = impsum_all + (size_t)it.impsum_off;
    float *of1 = of1_all + (size_t)it.impsum_off;
    int32_t *dip = dip_all + it.dip_off;
    uint8_t *H0 = H0_all + it.h0_off, *H1 = H1_all + it.h1_off;
    const float Kf = (float)K, invK = 1.0f / Kf;
    const float mism = ed / ee;
    const size_t KH = (size_t)it.msg;           // stored message stride (states rounded up to 64)
    const int Kr = it.msg >> 3;

    bool valid[KPT];
    int kk[KPT];
#pragma unroll
    for (int i = 0; i < KPT; i++) { kk[i] = tid + PH_NT * i; valid[i] = kk[i] < K; }

    // slots past K hold exact zeros throughout: set here, never written again
    float p[KPT][8];
#pragma unroll
    for (int i = 0; i < KPT; i++)
#pragma unroll
        for (int h = 0; h < 8; h++) p[i][h] = 0.0f;
    float sumK[KPT];
    float sumH[8];
    float sumT = 1.0f;
    int par = 0;
#pragma unroll
    for (int i = 0; i < KPT; i++) sumK[i] = 0.0f;

    // per-hypothesis sums over the workgroup: DPP reduce-scatter in each wave,
    // one LDS exchange and barrier across the waves (the slot alternates), then
    // the 8 totals are read out of lanes 0..7
    auto bsum8 = [&](float v[8]) {
        float x = wave_rs8(v);
        if (NW > 1) {
            par ^= 1;
            if (lane < 8) lds8[par][wid * 8 + lane] = x;
            __syncthreads();
            float t = 0.0f;
#pragma unroll
            for (int q = 0; q < NW; q++) t += lds8[par][q * 8 + (lane & 7)];
            x = t;
        }
#pragma unroll
        for (int h = 0; h < 8; h++) v[h] = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), h));
    };

    // tile of loci [q0, q0+n): scalars and Hvar rows into LDS
    auto load_tile = [&](int q0, int n) {
        __syncthreads();
        if (tid < n) {
            const int q = q0 + tid;
            const Locus Lc = loci[q];
            s_rel[tid] = Lc.rel;
            s_seg[tid] = Lc.seg;
            s_miss[tid] = Lc.miss;
            s_type[tid] = Lc.type;
            s_alt[tid] = Lc.alt;
            const bool has_prev = q > 0, has_next = q < nloc - 1;
            const Locus Lp_ = has_prev ? loci[q - 1] : Lc;
            const Locus Ln_ = has_next ? loci[q + 1] : Lc;
            s_yin[tid] = has_prev ? transition(cm, Lp_.abs, Lc.abs, nrho) : 0.0f;
            s_yout[tid] = has_next ? transition(cm, Lc.abs, Ln_.abs, nrho) : 0.0f;
            s_first[tid] = !has_prev || Lp_.seg != Lc.seg;
            s_last[tid] = !has_next || Ln_.seg != Lc.seg;
        }
        __syncthreads();
        for (int e = tid; e < n * W; e += PH_NT) {
            const int r = e / W, w = e - r * W;
            s_rows[e] = hv[(size_t)s_rel[r] * W + w];
        }
        __syncthreads();
    };
    // per-locus metadata is the same for every lane: read it into SGPRs
    auto ui = [](int v) { return __builtin_amdgcn_readfirstlane(v); };
    auto uf = [](float v) { return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v))); };
    // allele of state kk[i] at tile locus j: bit (lane & 63) of its Hvar word
    // (wid + NW*i).  Each lane reads the 32-bit half that holds its bit and
    // widens the bit to an all-ones / all-zeros mask (v_bfe_i32); the selects
    // are then bitwise.  The words of all slots are read together before a slot
    // loop and its skips, so their LDS latencies overlap; one base per locus and
    // constant slot offsets.  Slots past K read past the row's end (the next
    // row, or the unused rest of the tile, which holds PH_TILE rows of
    // NW*KPT >= W words) and ignore the word.
    const unsigned lane_sh = lane & 31;
    auto rwords = [&](int j, uint32_t (&rw)[KPT]) {
        const uint32_t *rb = (const uint32_t *)(s_rows + j * W + wid) + (lane >> 5);
#pragma unroll
        for (int i = 0; i < KPT; i++) rw[i] = rb[2 * NW * i];
    };
    auto bit_of = [&](const uint32_t (&rw)[KPT], int i) { return __builtin_amdgcn_sbfe((int)rw[i], lane_sh, 1u); };
    auto sel = [](int m, float v1, float v0) {
        return __int_as_float((m & __float_as_int(v1)) | (~m & __float_as_int(v0)));
    };
    auto keep = [](int m, float v) { return __int_as_float(m & __float_as_int(v)); };

    // One locus update (INIT / RUN / COLLAPSE), tile offset j.  A state's
    // update is base(h) * e(h) with base = fma(src, NT, tFreq(h)) (src = its own
    // value for RUN, its hypothesis sum for COLLAPSE; INIT is NT = 0, tFreq = 1,
    // so base is exactly 1) and e(h) the locus emission.  At a peak het
    // hypothesis h carries allele 0 when bit (2 - type) of h is set; its
    // emission is ee if that matches the state's allele, else ed, so a state
    // needs two selects (eA for hypotheses with the bit set, eB for the rest)
    // and the bit of each h is a compile-time constant in the specialised
    // update.  A peak hom's emission does not depend on h; a flat het has none
    // (a multiply by exactly 1 in the reference).  States past K get emission
    // 0.  The values are those of the reference's per-kind kernels.
    auto update = [&](auto TYc, auto RUNc, int j, float NT, const float tF[8], float eh0, float eh1,
                      float acc[8]) {
        constexpr int ty = decltype(TYc)::value;    // 0..2: het bit position, 3: peak hom, 4: flat het
        constexpr bool run = decltype(RUNc)::value;
        uint32_t rw[KPT];
        rwords(j, rw);
#pragma unroll
        for (int i = 0; i < KPT; i++) {
            if (PH_NT * i + 64 * wid >= K) continue;   // p[i] stays 0
            const bool vi = valid[i];
            const int ah = bit_of(rw, i);
            float eA = 1.0f, eB = 1.0f;
            if constexpr (ty <= 2) {
                eA = vi ? sel(ah, ed, ee) : 0.0f;
                eB = vi ? sel(ah, ee, ed) : 0.0f;
            } else if constexpr (ty == 3) {
                eA = vi ? sel(ah, eh1, eh0) : 0.0f;
            } else {
                eA = vi ? 1.0f : 0.0f;
            }
            const float sk = sumK[i];
#pragma unroll
            for (int h = 0; h < 8; h++) {
                float e = eA;
                if constexpr (ty <= 2) e = ((h >> ty) & 1) ? eA : eB;
                const float base = fmaf(run ? p[i][h] : sk, NT, tF[h]);
                const float x = base * e;
                p[i][h] = x;
                acc[h] += x;
            }
        }
    };
    auto step = [&](int j, int mode, float yt) {
        const int type = __builtin_amdgcn_readfirstlane(s_type[j]);
        const bool alt = __builtin_amdgcn_readfirstlane(s_alt[j]);
        if (type == -2 && mode == 0) {
            const float v = 1.0f / (8 * (unsigned)K);
#pragma unroll
            for (int i = 0; i < KPT; i++)
#pragma unroll
                for (int h = 0; h < 8; h++) p[i][h] = valid[i] ? v : 0.0f;
#pragma unroll
            for (int h = 0; h < 8; h++) sumH[h] = 1.0f / 8;
            sumT = 1.0f;
            return;
        }
        const float eh0 = alt ? mism : 1.0f, eh1 = alt ? 1.0f : mism;   // peak hom emissions
        const float nt = 1.0f - yt;
        float tF[8];
        // one reciprocal of the running total serves both scale factors
        const float rS = __builtin_amdgcn_rcpf(sumT);
        const float NT = (mode != 0) ? nt * rS : 0.0f;
        const float sc = yt * rS * invK;
#pragma unroll
        for (int h = 0; h < 8; h++) tF[h] = (mode != 0) ? sumH[h] * sc : 1.0f;
        float acc[8];
#pragma unroll
        for (int h = 0; h < 8; h++) acc[h] = 0.0f;
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        using I2 = std::integral_constant<int, 2>;
        using I3 = std::integral_constant<int, 3>;
        using I4 = std::integral_constant<int, 4>;
        using RN = std::integral_constant<bool, true>;
        using CL = std::integral_constant<bool, false>;
        const int ty = type >= 0 ? 2 - type : (type == -1 ? 3 : 4);
        if (mode == 1) {
            switch (ty) {
                case 0: update(I0{}, RN{}, j, NT, tF, eh0, eh1, acc); break;
                case 1: update(I1{}, RN{}, j, NT, tF, eh0, eh1, acc); break;
                case 2: update(I2{}, RN{}, j, NT, tF, eh0, eh1, acc); break;
                case 3: update(I3{}, RN{}, j, NT, tF, eh0, eh1, acc); break;
                default: update(I4{}, RN{}, j, NT, tF, eh0, eh1, acc); break;
            }
        } else {
            switch (ty) {
                case 0: update(I0{}, CL{}, j, NT, tF, eh0, eh1, acc); break;
                case 1: update(I1{}, CL{}, j, NT, tF, eh0, eh1, acc); break;
                case 2: update(I2{}, CL{}, j, NT, tF, eh0, eh1, acc); break;
                case 3: update(I3{}, CL{}, j, NT, tF, eh0, eh1, acc); break;
                default: update(I4{}, CL{}, j, NT, tF, eh0, eh1, acc); break;
            }
        }
        bsum8(acc);
#pragma unroll
        for (int h = 0; h < 8; h++) sumH[h] = acc[h];
        sumT = hadd8(sumH);
    };
    auto sumk = [&]() {
#pragma unroll
        for (int i = 0; i < KPT; i++) sumK[i] = hadd8(p[i]);
    };
    // a wave whose 64 states all lie beyond K (rounded up to 64) stores nothing
    auto store_msg = [&](float *dst, float *dsum, float *dsumT) {
#pragma unroll
        for (int i = 0; i < KPT; i++) {
            if (PH_NT * i + 64 * wid >= Kr) continue;
            float4 *d4 = (float4 *)(dst + ((size_t)i * PH_NT + tid) * 8);
            d4[0] = make_float4(p[i][0], p[i][1], p[i][2], p[i][3]);
            d4[1] = make_float4(p[i][4], p[i][5], p[i][6], p[i][7]);
        }
        if (tid < 8) dsum[tid] = sumH[tid];
        if (tid == 0 && dsumT) *dsumT = sumT;
    };

    // ---------------- backward (phasing_hmm.cpp:193-245)
    if (nloc > 0) {
        const int last_tile = ((nloc - 1) / PH_TILE) * PH_TILE;
        for (int q0 = last_tile; q0 >= 0; q0 -= PH_TILE) {
            const int n = min(PH_TILE, nloc - q0);
            load_tile(q0, n);
            for (int j = n - 1; j >= 0; j--) {
                const int q = q0 + j;
                const bool last = q == nloc - 1;
                step(j, last ? 0 : (ui(s_last[j]) ? 2 : 1), uf(s_yout[j]));
                if (ui(s_first[j])) {
                    sumk();
                    const int sg = ui(s_seg[j]);
                    store_msg(segp + (size_t)sg * KH, segs + sg * 9, segs + sg * 9 + 8);
                }
                if (ui(s_type[j]) == -2)
                    store_msg(impp + (size_t)ui(s_miss[j]) * KH, imps + ui(s_miss[j]) * 8, nullptr);
            }
        }
    }
    __syncthreads();
    // ---------------- root diplotype (phasing_hmm.cpp:250-256)
    if (tid == 0) {
        float D[8], S = 0.0f;
        if (nloc > 0) {
            const float s0 = segs[8];
            for (int d = 0; d < 8; d++) { D[d] = (segs[d] / s0) * (segs[7 - d] / s0); S += D[d]; }
        } else {
            for (int d = 0; d < 8; d++) D[d] = 0.0f;
        }
        const float u = u01(philox_draw(seed, iter, P_DIP_ROOT, it.tgt, 0u)) * S;
        float cs = D[0];
        int r = 7;
        for (int i = 0; nloc > 0 && i < 7; i++) {
            if (u <= cs) { r = i; break; }
            cs += D[i + 1];
        }
        dip[0] = r;
        dip_sh = r;
    }
    __syncthreads();
    int cur_dip = dip_sh;
    // ---------------- forward (phasing_hmm.cpp:142-191)
    for (int q0 = 0; q0 < nloc; q0 += PH_TILE) {
        const int n = min(PH_TILE, nloc - q0);
        load_tile(q0, n);
        for (int j = 0; j < n; j++) {
            const int q = q0 + j;
            step(j, q == 0 ? 0 : (ui(s_first[j]) ? 2 : 1), uf(s_yin[j]));
            const bool seg_last = ui(s_last[j]);
            if (seg_last) sumk();
            if (seg_last && q != nloc - 1) {
                // TRANS_HAP (phasing_hmm.h:351-373)
                const float yt2 = uf(s_yout[j]);
                const float nt2 = 1.0f - yt2;
                const float F2 = nt2 / sumT;
                // the next segment's stored backward message (segment-start loci
                // are rare, so it is loaded here rather than held in registers)
                const float *nxt = segp + (size_t)(ui(s_seg[j]) + 1) * KH;
                float nv[KPT][8];
#pragma unroll
                for (int i = 0; i < KPT; i++) {
                    float4 a = make_float4(0, 0, 0, 0), b = a;
                    if (valid[i]) {
                        const float4 *n4 = (const float4 *)(nxt + ((size_t)i * PH_NT + tid) * 8);
                        a = n4[0];
                        b = n4[1];
                    }
                    nv[i][0] = a.x; nv[i][1] = a.y; nv[i][2] = a.z; nv[i][3] = a.w;
                    nv[i][4] = b.x; nv[i][5] = b.y; nv[i][6] = b.z; nv[i][7] = b.w;
                }
                // the 64 sums (h1, h2) in two halves of 32 (h1 < 4, h1 >= 4)
#pragma unroll
                for (int hh = 0; hh < 2; hh++) {
                    float part[32];
#pragma unroll
                    for (int e = 0; e < 32; e++) part[e] = 0.0f;
#pragma unroll
                    for (int i = 0; i < KPT; i++) {
#pragma unroll
                        for (int u = 0; u < 4; u++) {
                            const int h1 = 4 * hh + u;
                            const float F1 = (sumH[h1] / sumT) * yt2 / Kf;
                            const float pv = fmaf(p[i][h1], F2, F1);
#pragma unroll
                            for (int h2 = 0; h2 < 8; h2++) part[u * 8 + h2] += pv * nv[i][h2];
                        }
                    }
                    const float r = wave_rs32(part);
                    if (lane < 32) red64[wid][32 * hh + rs32_index(lane)] = r;
                }
                __syncthreads();
                if (tid < 64) {
                    float t = 0.0f;
#pragma unroll
                    for (int q = 0; q < NW; q++) t += red64[q][tid];
                    hprobs[tid] = t;
                }
                __syncthreads();
                if (tid == 0) {
                    float sumHP = 0.0f;
                    for (int h1 = 0; h1 < 8; h1++) sumHP += hadd8(&hprobs[h1 * 8]);
                    float D[8], S = 0.0f;
                    const int d0 = cur_dip, d1 = 7 - cur_dip;
                    for (int d = 0; d < 8; d++) {
                        D[d] = (hprobs[d0 * 8 + d] / sumHP) * (hprobs[d1 * 8 + (7 - d)] / sumHP);
                        S += D[d];
                    }
                    int r = 7;
                    if (!(S >= 1.17549435e-38f) || !(sumHP >= 1.17549435e-38f) || isinf(S) || isinf(sumHP)) {
                        atomicExch(status, 1);
                    } else {
                        const float u =
                            u01(philox_draw(seed, iter, P_DIP_SEG, it.tgt, (uint32_t)(ui(s_seg[j]) + 1))) * S;
                        float cs = D[0];
                        for (int i = 0; i < 7; i++) {
                            if (u <= cs) { r = i; break; }
                            cs += D[i + 1];
                        }
                    }
                    dip[ui(s_seg[j]) + 1] = r;
                    dip_sh = r;
                }
                __syncthreads();
                cur_dip = dip_sh;
            }
            if (ui(s_type[j]) == -2) {
                // IMPUTE_FLAT_HET (phasing_hmm.h:389-423)
                const int mi = ui(s_miss[j]);
                const float *ip = impp + (size_t)mi * KH;
                float sR[8], sL[8], s0[8], s1[8];
#pragma unroll
                for (int h = 0; h < 8; h++) {
                    sR[h] = 1.0f / imps[mi * 8 + h];
                    sL[h] = 1.0f / sumH[h];
                    s0[h] = 0.0f;
                    s1[h] = 0.0f;
                }
                // slots past K are skipped; lanes past K hold 0 in both the
                // forward values and the stored message, so they add exact zeros
                uint32_t rw[KPT];
                rwords(j, rw);
#pragma unroll
                for (int i = 0; i < KPT; i++) {
                    if (PH_NT * i + 64 * wid >= K) continue;
                    const int ah = bit_of(rw, i);
                    const float4 *i4 = (const float4 *)(ip + ((size_t)i * PH_NT + tid) * 8);
                    const float4 a = i4[0], b = i4[1];
                    const float iv[8] = {a.x, a.y, a.z, a.w, b.x, b.y, b.z, b.w};
#pragma unroll
                    for (int h = 0; h < 8; h++) {
                        const float v = (iv[h] * sR[h]) * (p[i][h] * sL[h]);
                        s1[h] += keep(ah, v);
                        s0[h] += keep(~ah, v);
                    }
                }
                bsum8(s0);
                bsum8(s1);
                if (tid < 8) {
                    const float nn = s1[tid] / (s0[tid] + s1[tid]);
                    of1[mi * 8 + tid] = fmaxf(0.0f, fminf(1.0f, nn));
                }
            }
        }
    }
    __syncthreads();
    // ---------------- write-back (phasing_hmm.cpp:257-301)
    for (int q = tid; q < nloc; q += PH_NT) {
        const Locus Lc = loci[q];
        const int d0 = dip[Lc.seg], d1 = 7 - d0;
        if (Lc.type >= 0) {
            H0[Lc.abs] = ((7 - d0) >> (2 - Lc.type)) & 1;
            H1[Lc.abs] = ((7 - d1) >> (2 - Lc.type)) & 1;
        } else if (Lc.type == -2) {
            const float h0a1 = of1[Lc.miss * 8 + d0], h0a0 = 1.0f - h0a1;
            const float h1a1 = of1[Lc.miss * 8 + d1], h1a0 = 1.0f - h1a1;
            float p01 = (h0a0 * ee + h0a1 * ed) * (h1a1 * ee + h1a0 * ed);
            float p10 = (h0a1 * ee + h0a0 * ed) * (h1a0 * ee + h1a1 * ed);
            float sum = p01 + p10;
            p01 = fminf(fmaxf(p01 / sum, 0.0f), 1.0f);
            p10 = fminf(fmaxf(p10 / sum, 0.0f), 1.0f);
            sum = p01 + p10;
            const float u = u01(philox_draw(seed, iter, P_FLAT_HET, it.tgt, (uint32_t)Lc.miss));
            const bool rf = ((double)u * (double)sum) < (double)p01;
            H0[Lc.abs] = !rf;
        }
    }
    // heterozygous monomorphic sites get a random phase (phasing_hmm.cpp:294-301)
    const int32_t *pos = pos_all + c.site_off;
    for (int s = tid; s < L; s += PH_NT) {
        if (pos[s] >= 0) continue;
        if (H0[s] != H1[s]) {
            const bool rf = u01(philox_draw(seed, iter, P_MONO_HET, it.tgt, (uint32_t)s)) < 0.5f;
            H0[s] = rf;
            H1[s] = !rf;
        }
    }
}

void phase_prep_launch(int n, PhItem *d_items, const CondItem *d_cond, const int32_t *d_poly, const uint8_t *d_H0,
                       const uint8_t *d_H1, const uint8_t *d_flat, const uint8_t *d_flags, Locus *d_loci,
                       hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(phase_prep_kernel, dim3(n), dim3(PREP_NT), 0, s, d_items, d_cond, d_poly, d_H0, d_H1, d_flat,
                       d_flags, d_loci);
    GH_HIP(hipGetLastError());
}

template <int NW, int KPT>
static void phase_go(gh_panel *P, int n, const PhItem *d_items, const CondItem *d_cond, const uint64_t *d_hvar,
                     const Locus *d_loci, const int32_t *d_pos, double nrho, PhaseBufs &B, uint8_t *d_H0,
                     uint8_t *d_H1, float ee, float ed, uint32_t seed, uint32_t iter, int32_t *d_status,
                     hipStream_t s, const int32_t *d_order, GraphGate gate) {
    const size_t smem = (size_t)PH_TILE * ((KPT * 64 * NW + 63) / 64) * 8;
    hipLaunchKernelGGL((phase_kernel<NW, KPT>), dim3(n), dim3(64 * NW), smem, s, d_items, d_cond, d_hvar, d_loci,
                       d_pos, P->d_cm.as<double>(), nrho, B.seg, B.segsum, B.imp, B.impsum, B.of1, B.dip, d_H0, d_H1,
                       P->L, ee, ed, seed, iter, d_status, d_order, gate);
}

// Launch shapes (waves per target, states per thread) the kernel is built for,
// by state-count class; the tuning table picks one per class.
static const PhaseShape kPhaseShapes[] = {
    {1, 1}, {1, 2}, {1, 4}, {2, 1}, {2, 2}, {4, 1},                 // K <= 256
    {1, 8}, {2, 4}, {4, 2}, {8, 1},                                 // K <= 512
    {2, 8}, {4, 4}, {8, 2}, {16, 1},                                // K <= 1024
    {4, 8}, {8, 4}, {16, 2},                                        // K <= 2048
    {8, 8},                                                         // K <= 4096 (16 x 4 would spill:
                                                                    // 1024 threads leave 128 VGPRs)
};

std::vector<PhaseShape> phase_shapes(int maxK) {
    std::vector<PhaseShape> v;
    for (const PhaseShape &sh : kPhaseShapes) {
        const int cap = 64 * sh.NW * sh.KPT;
        if (cap >= maxK && cap < 2 * std::max(maxK, 128)) v.push_back(sh);
    }
    return v;
}

PhaseShape phase_shape(int maxK, bool latency) {
    GH_CHECK(maxK <= 4096, GH_ERR_ARG, "phasing: more than 4096 conditioning states is not supported");
    // GH_PHASE_SHAPE=<NW>x<KPT> pins a shape wherever it fits (testing, reproducing a tuned run)
    if (const char *e = std::getenv("GH_PHASE_SHAPE")) {
        int nw = 0, kpt = 0;
        if (std::sscanf(e, "%dx%d", &nw, &kpt) == 2)
            for (const PhaseShape &sh : phase_shapes(maxK))
                if (sh.NW == nw && sh.KPT == kpt) return sh;
    }
    // Defaults measured with 1000 targets in flight (five times the CUs): the
    // fewest waves per target whose register budget still lets two or more
    // targets share a CU (4 states per thread below 2048 states, 8 at 2048).
    // Latency defaults (a run on the critical path): more waves per target.
    const int cls = maxK <= 256 ? 0 : maxK <= 512 ? 1 : maxK <= 1024 ? 2 : maxK <= 2048 ? 3 : 4;
    static const PhaseShape dflt[5] = {{1, 4}, {2, 4}, {4, 4}, {4, 8}, {8, 8}};
    static const PhaseShape fast[5] = {{4, 1}, {4, 2}, {8, 2}, {8, 4}, {8, 8}};
    return latency ? fast[cls] : dflt[cls];
}

void phase_launch(gh_panel *P, int n, int maxK, const PhItem *d_items, const CondItem *d_cond,
                  const uint64_t *d_hvar, const Locus *d_loci, const int32_t *d_pos, int ne, PhaseBufs &B,
                  uint8_t *d_H0, uint8_t *d_H1, float err_phase, uint32_t seed, uint32_t iter, int32_t *d_status,
                  hipStream_t s, PhaseShape sh, const int32_t *d_order, GraphGate gate) {
    if (n <= 0) return;
    const double nrho = (-0.04 * ne) / P->N;
    const float ed = err_phase, ee = (float)(1.0 - (double)err_phase);
    if (sh.NW == 0) sh = phase_shape(maxK);
    GH_CHECK(64 * sh.NW * sh.KPT >= maxK, GH_ERR_ARG, "phasing: launch shape smaller than the conditioning set");
#define GH_PH(nw, kpt)                                                                                          \
    if (sh.NW == nw && sh.KPT == kpt) {                                                                         \
        phase_go<nw, kpt>(P, n, d_items, d_cond, d_hvar, d_loci, d_pos, nrho, B, d_H0, d_H1, ee, ed, seed, iter, \
                          d_status, s, d_order, gate);                                                          \
        GH_HIP(hipGetLastError());                                                                              \
        return;                                                                                                 \
    }
    GH_PH(1, 1) GH_PH(1, 2) GH_PH(1, 4) GH_PH(2, 1) GH_PH(2, 2) GH_PH(4, 1) GH_PH(2, 4) GH_PH(4, 2) GH_PH(8, 1)
    GH_PH(1, 8) GH_PH(2, 8) GH_PH(4, 4) GH_PH(8, 2) GH_PH(16, 1) GH_PH(4, 8) GH_PH(8, 4) GH_PH(16, 2)
    GH_PH(8, 8)
#undef GH_PH
    throw Error(GH_ERR_ARG, "phasing: no kernel for the requested launch shape");
}

int phase_kpt_states(PhaseShape sh) { return 64 * sh.NW * sh.KPT; }

}  // namespace gh
