// gh_impute.hip -- haploid Li-Stephens forward-backward
// (imputation_hmm::forward / backward, phase/src/models/imputation_hmm.cpp:143-430).
//
// One workgroup of NW wavefronts per haplotype item.  Word w of a Hvar row (64
// states) belongs to wave w % NW; lane j of that wave owns state 64w + j, so a
// row word gives every lane its state's allele as bit j.
//
// Latency: the per-site metadata (site id, flat/LQ flags, emissions,
// transitions, likelihoods) of a 32-site tile is gathered once, one site per
// lane, and stored in LDS as three 16-byte records per site, which the serial
// site loop reads with broadcast loads; the tile's Hvar rows are staged in LDS
// too.  Per-site sums use DPP row reductions and permlane swaps inside a wave,
// then one LDS exchange (double-buffered by site parity, one barrier) across
// the NW waves.
//
// Memory: alpha never round-trips through HBM.  The forward pass keeps alpha
// in registers and writes a checkpoint (the full alpha vector) every C sites;
// the backward pass walks blocks of C sites from the end, recomputes the
// block's alpha from its checkpoint into LDS (thread-private slots, the same
// operations in the same order, so bit-identical to the forward values) and
// runs the backward recursion against it.  HBM traffic per site is the Hvar row
// (three reads) plus 1/C of an alpha vector; the arithmetic per state-site is
// about 1.5x a plain forward-backward.
//
// Arithmetic per state follows the reference: alpha_l = (alpha_{l-1}*nt/S +
// t/K)*e (one FMA, one multiply), beta~ = beta*nt/S + t/K, the site's own
// emission divided out of the allele posteriors.  The order of the sums over
// states differs from the reference's 8-lane AVX2 order, and the posterior
// sums are accumulated as the total and the allele-1 part with FMAs (the
// allele-0 part is their difference, never negative because rounding is
// monotone and both sums take the same order); HP stays within float rounding
// of the reference's.
#include "gh_device.hpp"
#include "gh_internal.hpp"

#include <cstdio>
#include <type_traits>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace gh {

struct FbArgs {
    const FbItem *fb;
    const CondItem *cond;
    const int32_t *poly_all;
    const float *t_all;
    const uint64_t *hvar_all;
    const float *HL_all;
    const uint8_t *flat_all;
    const uint8_t *flags;
    float *ckpt_all;
    float *HP_all;
    float *mass_all;   // optional: per-state posterior mass (summed over sites)
    const int32_t *win_lo, *win_hi;   // optional with mass_all: per-state site window [lo, hi) of the sum
    const int32_t *order;             // optional: workgroup b runs item order[b]
    GraphGate gate;                   // optional: launch gated by the device (replayed iterations)
    float ee, ed;
    int C;             // sites per checkpoint block (divides FB_TILE)
    int maxW;          // LDS row stride bound
};


template <int NW, int MWL>
// three waves per SIMD (FB_WG_PER_CU workgroups of the 4-wave shapes) must stay
// possible; the 16-word shapes (tuner candidates only) keep their registers
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(MWL <= 8 ? 3 : 1))) void fb_kernel(FbArgs A) {
    extern __shared__ __align__(16) unsigned char smem[];
    if (A.gate.key && (*A.gate.key != A.gate.want || (int)blockIdx.x >= *A.gate.n)) return;
    const FbItem f = wave_uniform(A.fb[A.order ? A.order[blockIdx.x] : (int)blockIdx.x]);
    const CondItem c = wave_uniform(A.cond[f.cond]);
    const int K = c.K, W = c.W, Lp = c.Lp, C = A.C;
    const int tid = threadIdx.x, lane = tid & 63, wv = __builtin_amdgcn_readfirstlane(tid >> 6);
    if (Lp <= 0) return;
    static_assert(NW * MWL <= FB_ROW_PAD, "row pad covers every shape");
    uint64_t *rows = (uint64_t *)smem + FB_ROW_PAD;                       // [FB_TILE][W]
    float *ablk = (float *)(smem + (size_t)(FB_ROW_PAD + FB_TILE * A.maxW) * 8);   // [C][NW*MWL][64]
    float *red = ablk + (size_t)C * NW * MWL * 64;                        // [2][NW][4]
    float4 *smeta = (float4 *)(red + 2 * 16 * 4);                         // [FB_TILE][3] site records
    const int32_t *poly = A.poly_all + c.site_off;
    const float *tr = A.t_all + c.site_off;
    const uint64_t *hv = A.hvar_all + c.hvar_off;
    const float *HL = A.HL_all + f.hl_off;
    const uint8_t *flat = A.flat_all + f.flat_off;
    float *ck = A.ckpt_all + f.alpha_off;
    float *HP = A.HP_all + f.hl_off;
    const float ee = A.ee, ed = A.ed;
    const float Kf = (float)K, invK = 1.0f / Kf;
    const size_t vstride = (size_t)64 * W;   // one alpha vector (checkpoint) in floats
    // the forward pass's per-site scaling sums, after the checkpoints: the
    // recomputation reads them instead of reducing across the waves again
    float *ssum = ck + (size_t)((Lp + C - 1) / C) * vstride;

    // this wave owns words wv, wv + NW, wv + 2 NW, ... of the item, nI of them;
    // slots i >= nI hold no word of this item and are skipped (a wave-uniform
    // bound: items with fewer states than the launch capacity do proportionally
    // less work)
    const int nI = min(MWL, max(0, (W - wv + NW - 1) / NW));
    // Slots run from the wave's highest word down: slot i holds word
    // (nI-1-i)*NW + wv.  When K is not a multiple of 64 the item's last word
    // W-1 is partial, and it is slot 0 of wave (W-1) % NW.  Slot 0 therefore
    // always uses emissions that are 0 on lanes past K (pv; equal to the
    // site's emissions elsewhere), which keeps those lanes' alpha and beta
    // exactly 0, so no other step masks lanes.
    const bool partW = (K & 63) != 0 && ((W - 1) % NW) == wv;
    const bool pv = !partW || (64 * (W - 1) + lane < K);
    int wid[MWL];
#pragma unroll
    for (int i = 0; i < MWL; i++) wid[i] = i < nI ? (nI - 1 - i) * NW + wv : 0;   // clamped: reads stay in the tile
    auto vlane = [&](int i) { return i < nI && 64 * wid[i] + lane < K; };
    // The row words of all word slots at tile site j are read together, before
    // the slot loop and its wave-uniform skips, so their LDS latencies overlap
    // (slots past nI read a word of the tile or of the pad below it and ignore it).
    // Slot i reads word (nI-1-i)*NW + wv of the row, i.e. rb[(MWL-1-i)*NW] from
    // one base per site: constant offsets, one address computation per site.
    // Each lane reads the 32-bit half of the word that holds its own state's
    // bit and widens the bit to an all-ones / all-zeros mask (v_bfe_i32); the
    // selects are then bitwise (one v_bitop3_b32 or v_and_b32), with no trip
    // through scalar registers.
    const unsigned lane_sh = lane & 31;
    auto rwords = [&](int j, uint32_t (&rw)[MWL]) {
        const uint32_t *rb = (const uint32_t *)(rows + j * W + (nI - MWL) * NW + wv) + (lane >> 5);
#pragma unroll
        for (int i = 0; i < MWL; i++) rw[i] = rb[2 * (MWL - 1 - i) * NW];
    };
    auto rmask = [&](const uint32_t (&rw)[MWL], int i) { return __builtin_amdgcn_sbfe((int)rw[i], lane_sh, 1u); };
    auto sel = [](int m, float v1, float v0) {
        return __int_as_float((m & __float_as_int(v1)) | (~m & __float_as_int(v0)));
    };
    auto keep = [](int m, float v) { return __int_as_float(m & __float_as_int(v)); };

    // cross-wave sums in a fixed order (the recomputation takes the forward's
    // scaling sums from ssum, so only the forward pass reduces them).  The LDS
    // slot alternates with every reduction, so one barrier per reduction suffices.
    int par = 0;
    auto xsum1 = [&](float v) -> float {
        v = wave_allsum(v);
        if (NW == 1) return v;
        par ^= 1;
        if (lane == 0) red[(par * NW + wv) * 4] = v;
        __syncthreads();
        float s = 0.0f;
#pragma unroll
        for (int q = 0; q < NW; q++) s += red[(par * NW + q) * 4];
        return s;
    };
    // the backward site's three sums: the beta sum z (every wave), and the
    // posterior total x and allele-1 part y, read back only where need_p
    // (wave 0, which writes HP; every wave when posterior masses are summed)
    auto xsum_bwd = [&](float &x, float &y, float &z, bool need_p) {
        z = wave_allsum(z);
        const float xy = wave_sum_pair(x, y);   // x in lanes 0-31, y in lanes 32-63
        if (NW == 1) {
            x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xy), 0));
            y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(xy), 32));
            return;
        }
        par ^= 1;
        float *r = red + (par * NW + wv) * 4;
        if (lane == 0) { r[0] = z; r[1] = xy; }
        if (lane == 32) r[2] = xy;
        __syncthreads();
        z = 0.0f;
#pragma unroll
        for (int q = 0; q < NW; q++) z += red[(par * NW + q) * 4 + 0];
        if (need_p) {
            x = 0.0f; y = 0.0f;
#pragma unroll
            for (int q = 0; q < NW; q++) {
                x += red[(par * NW + q) * 4 + 1];
                y += red[(par * NW + q) * 4 + 2];
            }
        }
    };

    // Tile metadata: three 16-byte records per site in LDS, read with one
    // broadcast load each (every lane the same address):
    //   [0] emissions of allele 0 / 1 (1 at flat and LQ sites), t/K and 1-t into the site
    //   [1] t/K and 1-t out of the site, the haplotype likelihoods h0 / h1
    //   [2] site index, flags (bit 0 flat or LQ, bit 1 flat), 1 / emission 0, 1 / emission 1
    // The next tile is fetched into registers while the current one is processed:
    // stage 1 (site ids, transitions, Hvar rows) right after a tile starts,
    // stage 2 (flags and likelihoods of those sites) once the ids have arrived.
    // A tile then only costs the LDS stores and two barriers.
    int pf_site = 0, pf_l0 = 0, pf_n = 0;
    float pf_tin = 0.0f, pf_tout = 0.0f, pf_h0 = 0.5f, pf_h1 = 0.5f;
    uint8_t pf_flat = 0, pf_flag = 0;
    uint64_t pf_rows[MWL];
    auto prefetch1 = [&](int l0, int n) {
        pf_l0 = l0;
        pf_n = n;
        if (lane < n) {
            const int l = l0 + lane;
            pf_site = poly[l];
            pf_tin = l > 0 ? tr[l - 1] : 0.0f;
            pf_tout = l < Lp - 1 ? tr[l] : 0.0f;
        }
        const uint64_t *src = hv + (size_t)l0 * W;
#pragma unroll
        for (int k = 0; k < MWL; k++) {
            const int e = tid + k * 64 * NW;
            pf_rows[k] = e < n * W ? src[e] : 0ull;
        }
    };
    auto prefetch2 = [&]() {
        if (lane < pf_n) {
            pf_flat = flat[pf_site];
            pf_flag = A.flags[pf_site];
            pf_h0 = HL[2 * pf_site];
            pf_h1 = HL[2 * pf_site + 1];
        }
    };
    // make the fetched tile current (every thread calls it)
    auto commit_tile = [&]() {
        __syncthreads();   // previous tile's LDS rows and records are no longer read
        if (wv == 0 && lane < pf_n) {
            const int site = pf_site;
            const bool isflat = pf_flat != 0;
            const bool fl = isflat || (pf_flag & SF_LQ);
            const float h0 = pf_h0, h1 = pf_h1;
            const float p0 = h0 * ee + h1 * ed, p1 = h0 * ed + h1 * ee;
            const float e0 = fl ? 1.0f : p0 / (p0 + p1), e1 = fl ? 1.0f : p1 / (p0 + p1);
            smeta[3 * lane + 0] = make_float4(e0, e1, pf_tin / Kf, 1.0f - pf_tin);
            smeta[3 * lane + 1] = make_float4(pf_tout / Kf, 1.0f - pf_tout, h0, h1);
            smeta[3 * lane + 2] = make_float4(__int_as_float(site), __int_as_float((fl ? 1 : 0) | (isflat ? 2 : 0)),
                                              1.0f / e0, 1.0f / e1);
        }
#pragma unroll
        for (int k = 0; k < MWL; k++) {
            const int e = tid + k * 64 * NW;
            if (e < pf_n * W) rows[e] = pf_rows[k];
        }
        __syncthreads();
    };
    auto load_tile = [&](int l0, int n) {
        prefetch1(l0, n);
        prefetch2();
        commit_tile();
    };

    // A slot loop runs as straight-line code for a wave whose slots all hold
    // words (nI == MWL: every wave of the launch's largest items, whose chains
    // are the longest), so the compiler can interleave the slots' instruction
    // chains; other waves skip their empty slots one by one.
    auto by_fill = [&](auto &&body) {
        if (nI == MWL) body(std::true_type{});
        else body(std::false_type{});
    };
    float a[MWL];
#pragma unroll
    for (int i = 0; i < MWL; i++) a[i] = 0.0f;
    // forward step at site l (tile offset j): alpha in a[], scaling sum in sp
    // (flat and LQ sites have emission 1 for both alleles)
    auto fwd_site = [&](int j, int l, float &sp) {
        const float4 mf = smeta[3 * j];
        uint32_t rw[MWL];
        rwords(j, rw);
        const float e0 = mf.x, e1 = mf.y;
        const float e0p = pv ? e0 : 0.0f, e1p = pv ? e1 : 0.0f;   // the partial word's emissions
        float s = 0.0f;
        if (l == 0) {
#pragma unroll
            for (int i = 0; i < MWL; i++) {
                if (i >= nI) continue;
                const int m = rmask(rw, i);
                const float x = (i > 0 ? sel(m, e1, e0) : sel(m, e1p, e0p)) * invK;
                a[i] = x;
                s += x;
            }
            const bool fl = __builtin_amdgcn_readfirstlane(__float_as_int(smeta[3 * j + 2].y)) & 1;
            sp = fl ? 1.0f : xsum1(s);
        } else {
            const float f1 = mf.z, f2 = mf.w * __builtin_amdgcn_rcpf(sp);
            by_fill([&](auto FULLc) {
                constexpr bool full = decltype(FULLc)::value;
#pragma unroll
                for (int i = 0; i < MWL; i++) {
                    if (!full && i >= nI) continue;
                    const int m = rmask(rw, i);
                    const float x = fmaf(a[i], f2, f1) * (i > 0 ? sel(m, e1, e0) : sel(m, e1p, e0p));
                    a[i] = x;
                    s += x;
                }
            });
            sp = xsum1(s);
        }
    };

    // the same update at site l > 0 without the scaling sum (the recomputation
    // takes it from ssum): identical alpha, no cross-wave reduction or barrier
    auto fwd_site_nosum = [&](int j, float sp) {
        const float4 mf = smeta[3 * j];
        uint32_t rw[MWL];
        rwords(j, rw);
        const float e0 = mf.x, e1 = mf.y;
        const float e0p = pv ? e0 : 0.0f, e1p = pv ? e1 : 0.0f;
        const float f1 = mf.z, f2 = mf.w * __builtin_amdgcn_rcpf(sp);
        by_fill([&](auto FULLc) {
            constexpr bool full = decltype(FULLc)::value;
#pragma unroll
            for (int i = 0; i < MWL; i++) {
                if (!full && i >= nI) continue;
                const int m = rmask(rw, i);
                a[i] = fmaf(a[i], f2, f1) * (i > 0 ? sel(m, e1, e0) : sel(m, e1p, e0p));
            }
        });
    };

    // ---------------- forward: checkpoints every C sites
    {
        float sp = 1.0f;
        load_tile(0, min(FB_TILE, Lp));
        for (int l0 = 0; l0 < Lp; l0 += FB_TILE) {
            const int n = min(FB_TILE, Lp - l0);
            if (l0 > 0) commit_tile();
            const bool more = l0 + FB_TILE < Lp;
            if (more) prefetch1(l0 + FB_TILE, min(FB_TILE, Lp - l0 - FB_TILE));
            for (int j = 0; j < n; j++) {
                const int l = l0 + j;
                if (more && j == FB_TILE / 2) prefetch2();
                fwd_site(j, l, sp);
                if (tid == 0) ssum[l] = sp;
                if ((l & (C - 1)) == 0) {
                    float *dst = ck + (size_t)(l / C) * vstride + lane;
#pragma unroll
                    for (int i = 0; i < MWL; i++) {
                        if (i >= nI) continue;
                        dst[64 * wid[i]] = a[i];
                    }
                }
            }
            if (more && n <= FB_TILE / 2) prefetch2();
        }
    }

    // ---------------- backward over checkpoint blocks, alpha recomputed into LDS
    float b[MWL], ms[MWL];
    int wlo[MWL], whi[MWL];
#pragma unroll
    for (int i = 0; i < MWL; i++) {
        wlo[i] = 0;
        whi[i] = 0x7FFFFFFF;
        if (A.win_lo && vlane(i)) {
            wlo[i] = A.win_lo[c.state_off + 64 * wid[i] + lane];
            whi[i] = A.win_hi[c.state_off + 64 * wid[i] + lane];
        }
    }
#pragma unroll
    for (int i = 0; i < MWL; i++) { b[i] = 0.0f; ms[i] = 0.0f; }
    float bsum_next = 1.0f;
    const int last_tile = ((Lp - 1) / FB_TILE) * FB_TILE;
    // checkpoint of the next block to process, fetched one block ahead
    float nck[MWL];
#pragma unroll
    for (int i = 0; i < MWL; i++) nck[i] = 0.0f;
    auto fetch_ck = [&](int bs) {
        const float *srcv = ck + (size_t)(bs / C) * vstride + lane;
#pragma unroll
        for (int i = 0; i < MWL; i++) {
            if (i >= nI) continue;
            nck[i] = srcv[64 * wid[i]];
        }
    };
    fetch_ck(((Lp - 1) / C) * C);
    load_tile(last_tile, Lp - last_tile);
    for (int l0 = last_tile; l0 >= 0; l0 -= FB_TILE) {
        const int n = min(FB_TILE, Lp - l0);
        if (l0 < last_tile) commit_tile();
        const bool more = l0 > 0;
        if (more) prefetch1(l0 - FB_TILE, FB_TILE);
        for (int q = (n - 1) / C; q >= 0; q--) {
            const int bs = l0 + q * C, be = min(bs + C, l0 + n);
            // recompute alpha on [bs, be) from the checkpoint at bs
#pragma unroll
            for (int i = 0; i < MWL; i++) a[i] = nck[i];
            if (bs > 0) fetch_ck(bs - C);
            float sp = ssum[bs];
#pragma unroll
            for (int i = 0; i < MWL; i++) {
                if (i >= nI) continue;
                ablk[((size_t)0 * NW * MWL + i * NW + wv) * 64 + lane] = a[i];
            }
            for (int l = bs + 1; l < be; l++) {
                const float spn = ssum[l];
                fwd_site_nosum(l - l0, sp);
                sp = spn;
                by_fill([&](auto FULLc) {
                    constexpr bool full = decltype(FULLc)::value;
#pragma unroll
                    for (int i = 0; i < MWL; i++) {
                        if (!full && i >= nI) continue;
                        ablk[((size_t)(l - bs) * NW * MWL + i * NW + wv) * 64 + lane] = a[i];
                    }
                });
            }
            if (more && q == (n - 1) / C) prefetch2();   // the next tile's site ids have arrived by now
            // backward recursion on the block
            for (int l = be - 1; l >= bs; l--) {
                const int j = l - l0;
                const float4 mf = smeta[3 * j], mb = smeta[3 * j + 1], mi = smeta[3 * j + 2];
                const float e0 = mf.x, e1 = mf.y;   // 1 at flat and LQ sites
                const float e0p = pv ? e0 : 0.0f, e1p = pv ? e1 : 0.0f;
                const int code = __builtin_amdgcn_readfirstlane(__float_as_int(mi.y));
                const bool fl = code & 1, isflat = code & 2;
                const int site = __float_as_int(mi.x);
                // posterior sums of this wave's states: total pt and allele-1 part p1
                float pt = 0.0f, p1 = 0.0f, bs_ = 0.0f;
                float btv[MWL];   // beta~ per slot (1 at the last site), for the posterior masses
                // row words and recomputed alpha of every slot, read before the slot loop
                uint32_t rw[MWL];
                rwords(j, rw);
                float avs[MWL];
#pragma unroll
                for (int i = 0; i < MWL; i++) avs[i] = ablk[((size_t)(l - bs) * NW * MWL + i * NW + wv) * 64 + lane];
                if (l == Lp - 1) {
#pragma unroll
                    for (int i = 0; i < MWL; i++) {
                        if (i >= nI) continue;
                        const int m = rmask(rw, i);
                        const float av = avs[i];
                        const float bv = (i > 0 ? sel(m, e1, e0) : sel(m, e1p, e0p)) * invK;
                        b[i] = bv;
                        btv[i] = 1.0f;
                        pt += av;
                        p1 += keep(m, av);
                        bs_ += b[i];
                    }
                } else {
                    const float f1 = mb.x, f2 = mb.y * __builtin_amdgcn_rcpf(bsum_next);
                    by_fill([&](auto FULLc) {
                        constexpr bool full = decltype(FULLc)::value;
#pragma unroll
                        for (int i = 0; i < MWL; i++) {
                            if (!full && i >= nI) continue;
                            const int m = rmask(rw, i);
                            const float av = avs[i];
                            const float bt = fmaf(b[i], f2, f1);   // lanes past K: b = 0 and alpha = 0
                            btv[i] = bt;
                            pt = fmaf(av, bt, pt);
                            p1 = fmaf(keep(m, av), bt, p1);
                            b[i] = bt * (i > 0 ? sel(m, e1, e0) : sel(m, e1p, e0p));
                            bs_ += b[i];
                        }
                    });
                }
                xsum_bwd(pt, p1, bs_, wv == 0 || A.mass_all != nullptr);
                if (fl && l == Lp - 1) bs_ = 1.0f;
                if (A.mass_all) {
                    const float inv = 1.0f / pt;
#pragma unroll
                    for (int i = 0; i < MWL; i++) {
                        if (i >= nI) continue;
                        const bool in = site >= wlo[i] && site < whi[i];
                        ms[i] = in ? fmaf(avs[i] * btv[i], inv, ms[i]) : ms[i];
                    }
                }
                if (tid == 0) {
                    // the site's own emission divided out (the record's reciprocals, 1 at
                    // flat and LQ sites), the genotyping error applied, the likelihoods
                    // multiplied in away from flat sites, normalised
                    const float q0 = (pt - p1) * mi.z, q1 = p1 * mi.w;
                    float o0 = q0 * ee + q1 * ed, o1 = q0 * ed + q1 * ee;
                    if (!isflat) { o0 *= mb.z; o1 *= mb.w; }
                    const float r = __builtin_amdgcn_rcpf(o0 + o1);
                    HP[2 * site] = o0 * r;
                    HP[2 * site + 1] = o1 * r;
                }
                bsum_next = bs_;
            }
        }
    }
    if (A.mass_all) {
        float *mo = A.mass_all + (size_t)c.state_off;
#pragma unroll
        for (int i = 0; i < MWL; i++)
            if (vlane(i)) mo[64 * wid[i] + lane] = ms[i];
    }
}

// imputation_hmm.cpp:417-429: monomorphic sites take the major allele's closed form.
__global__ void fb_mono_kernel(const FbItem *__restrict__ fb, const CondItem *__restrict__ cond, int L,
                               const int32_t *__restrict__ pos_all, const uint8_t *__restrict__ flags,
                               const float *__restrict__ HL_all, const uint8_t *__restrict__ flat_all,
                               float *__restrict__ HP_all, float ee, float ed) {
    const FbItem f = fb[blockIdx.y];
    const CondItem c = cond[f.cond];
    const int site = blockIdx.x * blockDim.x + threadIdx.x;
    if (site >= L) return;
    if (pos_all[c.site_off + site] >= 0) return;
    const bool maj = flags[site] & SF_MAJOR;
    float o[2];
    o[maj] = ee;
    o[!maj] = ed;
    if (!flat_all[f.flat_off + site]) {
        o[0] *= HL_all[f.hl_off + 2 * site];
        o[1] *= HL_all[f.hl_off + 2 * site + 1];
    }
    HP_all[f.hl_off + 2 * site] = o[0] / (o[0] + o[1]);
    HP_all[f.hl_off + 2 * site + 1] = o[1] / (o[0] + o[1]);
}

template <int NW, int MWL>
static void fb_go(const FbArgs &A, int n, size_t smem, hipStream_t s) {
    hipLaunchKernelGGL((fb_kernel<NW, MWL>), dim3(n), dim3(64 * NW), smem, s, A);
}

template <int NW, int MWL>
static void fb_attr() {
    GH_HIP(hipFuncSetAttribute((const void *)fb_kernel<NW, MWL>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               FB_LDS_BYTES));
}

// The LDS limit of every shape is raised once, before any launch (a launch
// captured into a graph must not be the first use of its kernel).
static void fb_set_attributes() {
    fb_attr<1, 4>(); fb_attr<2, 2>(); fb_attr<4, 1>(); fb_attr<1, 8>(); fb_attr<2, 4>(); fb_attr<4, 2>();
    fb_attr<8, 1>(); fb_attr<1, 16>(); fb_attr<2, 8>(); fb_attr<4, 4>(); fb_attr<8, 2>(); fb_attr<16, 1>();
    fb_attr<2, 16>(); fb_attr<4, 8>(); fb_attr<8, 4>(); fb_attr<16, 2>(); fb_attr<8, 8>(); fb_attr<16, 4>();
}

void fb_launch(gh_panel *P, int n, const FbItem *d_fb, const CondItem *d_cond, int maxW, int ckpt_maxW,
               const int32_t *d_poly,
               const float *d_t, const uint64_t *d_hvar, const float *d_HL, const uint8_t *d_flat, float *d_ckpt,
               float *d_HP, float ee, float ed, hipStream_t s, float *d_mass, const int32_t *d_win_lo,
               const int32_t *d_win_hi, FbPick pick, const int32_t *d_order, GraphGate gate) {
    if (n <= 0) return;
    static std::once_flag attr_once;
    std::call_once(attr_once, fb_set_attributes);
    GH_CHECK(maxW <= FB_MAX_WORDS, GH_ERR_ARG, "fb: more than 4096 conditioning states per haplotype is not supported");
    const FbShape sh = fb_shape(maxW, pick);
    // the launch writes a checkpoint every sh.C sites into buffers sized for
    // blocks of fb_block_sites(ckpt_maxW): longer blocks write fewer, shorter
    // ones would overrun them
    GH_CHECK(fb_launch_fits(maxW, ckpt_maxW, sh), GH_ERR_ARG,
             "fb: launch block length shorter than the checkpoint layout");
    FbArgs A;
    A.fb = d_fb; A.cond = d_cond; A.poly_all = d_poly; A.t_all = d_t; A.hvar_all = d_hvar; A.HL_all = d_HL;
    A.flat_all = d_flat; A.flags = P->d_flags.as<uint8_t>(); A.ckpt_all = d_ckpt; A.HP_all = d_HP;
    A.mass_all = d_mass; A.win_lo = d_win_lo; A.win_hi = d_win_hi; A.order = d_order; A.gate = gate;
    A.ee = ee; A.ed = ed; A.C = sh.C; A.maxW = maxW;
#define GH_FB(nw, mwl)                    \
    if (sh.NW == nw && sh.MWL == mwl) {   \
        fb_go<nw, mwl>(A, n, sh.smem, s); \
        GH_HIP(hipGetLastError());        \
        return;                           \
    }
    GH_FB(1, 4) GH_FB(2, 2) GH_FB(4, 1) GH_FB(1, 8) GH_FB(2, 4) GH_FB(4, 2) GH_FB(8, 1)
    GH_FB(1, 16) GH_FB(2, 8) GH_FB(4, 4) GH_FB(8, 2) GH_FB(16, 1)
    GH_FB(2, 16) GH_FB(4, 8) GH_FB(8, 4) GH_FB(16, 2) GH_FB(8, 8) GH_FB(16, 4)
#undef GH_FB
    throw Error(GH_ERR_ARG, "fb: no kernel for the requested launch shape");
}

void fb_mono_launch(gh_panel *P, int n, const FbItem *d_fb, const CondItem *d_cond, const int32_t *d_pos,
                    const float *d_HL, const uint8_t *d_flat, float *d_HP, float ee, float ed, hipStream_t s) {
    if (n <= 0) return;
    dim3 grid((P->L + 255) / 256, n);
    hipLaunchKernelGGL(fb_mono_kernel, grid, dim3(256), 0, s, d_fb, d_cond, P->L, d_pos, P->d_flags.as<uint8_t>(),
                       d_HL, d_flat, d_HP, ee, ed);
    GH_HIP(hipGetLastError());
}

}  // namespace gh
