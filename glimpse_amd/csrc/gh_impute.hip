// gh_impute.hip -- haploid Li-Stephens forward-backward
// (imputation_hmm::forward / backward, phase/src/models/imputation_hmm.cpp:143-430).
//
// One wavefront per haplotype item.  Lane j owns states j, 64+j, 128+j, ...
// (MW registers), so word w of a Hvar row gives lane j its state's allele as
// bit j: every row word is a wave-uniform (scalar) load and alpha rows are
// written/read as contiguous 256-byte segments per word.  Per-site scaling
// sums use DPP row reductions plus v_readlane (no LDS, no barriers).
//
// Arithmetic follows the reference per state: alpha_l = (alpha_{l-1} * nt/S +
// t/K) * e (one FMA then a multiply), beta~ = beta * nt/S + t/K, posteriors
// divide out the site's own emission.  Only the order of the cross-state sums
// differs from the reference's 8-lane AVX2 order.
#include "gh_device.hpp"
#include "gh_internal.hpp"

namespace gh {

template <int MW>
__global__ __launch_bounds__(64) void fb_kernel(const FbItem *__restrict__ fb, const CondItem *__restrict__ cond,
                                               const int32_t *__restrict__ poly_all, const float *__restrict__ t_all,
                                               const uint64_t *__restrict__ hvar_all, const float *__restrict__ HL_all,
                                               const uint8_t *__restrict__ flat_all,
                                               const uint8_t *__restrict__ flags, float *__restrict__ alpha_all,
                                               float *__restrict__ HP_all, float ee, float ed) {
    const FbItem f = fb[blockIdx.x];
    const CondItem c = cond[f.cond];
    const int K = c.K, W = c.W, Lp = c.Lp;
    const int lane = threadIdx.x;
    const int32_t *poly = poly_all + c.site_off;
    const float *t = t_all + c.site_off;
    const uint64_t *hv = hvar_all + c.hvar_off;
    const float *HL = HL_all + f.hl_off;
    const uint8_t *flat = flat_all + f.flat_off;
    float *al = alpha_all + f.alpha_off;
    float *HP = HP_all + f.hl_off;
    const float Kf = (float)K;
    const float invK = 1.0f / Kf;
    const size_t rstride = (size_t)64 * W;

    bool valid[MW];
#pragma unroll
    for (int i = 0; i < MW; i++) valid[i] = (i < W) && (64 * i + lane < K);

    // ---------------- forward
    float a[MW];
    float sum_prev = 1.0f;
    for (int l = 0; l < Lp; l++) {
        const int site = poly[l];
        const bool fl = flat[site] || (flags[site] & SF_LQ);
        const uint64_t *row = hv + (size_t)l * W;
        float s = 0.0f;
        if (fl) {
            if (l == 0) {
#pragma unroll
                for (int i = 0; i < MW; i++) a[i] = valid[i] ? invK : 0.0f;
            } else {
                const float tt = t[l - 1];
                const float f1 = tt / Kf, f2 = (1.0f - tt) / sum_prev;
#pragma unroll
                for (int i = 0; i < MW; i++) {
                    a[i] = valid[i] ? fmaf(a[i], f2, f1) : 0.0f;
                    s += a[i];
                }
            }
        } else {
            const float h0 = HL[2 * site], h1 = HL[2 * site + 1];
            const float p0 = h0 * ee + h1 * ed, p1 = h0 * ed + h1 * ee;
            const float e0 = p0 / (p0 + p1), e1 = p1 / (p0 + p1);
            if (l == 0) {
#pragma unroll
                for (int i = 0; i < MW; i++) {
                    if (i < W) {
                        const bool b = (row[i] >> lane) & 1ull;
                        a[i] = valid[i] ? (b ? e1 : e0) * invK : 0.0f;
                        s += a[i];
                    }
                }
            } else {
                const float tt = t[l - 1];
                const float f1 = tt / Kf, f2 = (1.0f - tt) / sum_prev;
#pragma unroll
                for (int i = 0; i < MW; i++) {
                    if (i < W) {
                        const bool b = (row[i] >> lane) & 1ull;
                        a[i] = valid[i] ? fmaf(a[i], f2, f1) * (b ? e1 : e0) : 0.0f;
                        s += a[i];
                    }
                }
            }
        }
        sum_prev = (fl && l == 0) ? 1.0f : wave_sum(s);
        float *dst = al + (size_t)l * rstride + lane;
#pragma unroll
        for (int i = 0; i < MW; i++)
            if (i < W) dst[64 * i] = a[i];
    }

    // ---------------- backward + posteriors
    float b[MW];
#pragma unroll
    for (int i = 0; i < MW; i++) b[i] = 0.0f;
    float bsum_next = 1.0f;
    for (int l = Lp - 1; l >= 0; l--) {
        const int site = poly[l];
        const bool isflat = flat[site] != 0;
        const bool fl = isflat || (flags[site] & SF_LQ);
        const uint64_t *row = hv + (size_t)l * W;
        const float *src = al + (size_t)l * rstride + lane;
        float p0 = 0.0f, p1 = 0.0f, bs = 0.0f;
        float e0 = 1.0f, e1 = 1.0f;
        const float h0 = HL[2 * site], h1 = HL[2 * site + 1];
        if (!fl) {
            const float q0 = h0 * ee + h1 * ed, q1 = h0 * ed + h1 * ee;
            e0 = q0 / (q0 + q1);
            e1 = q1 / (q0 + q1);
        }
        if (l == Lp - 1) {
#pragma unroll
            for (int i = 0; i < MW; i++) {
                if (i < W) {
                    const bool bit = (row[i] >> lane) & 1ull;
                    const float av = src[64 * i];
                    b[i] = valid[i] ? (fl ? invK : (bit ? e1 : e0) * invK) : 0.0f;
                    if (bit) p1 += av; else p0 += av;
                    bs += b[i];
                }
            }
        } else {
            const float tt = t[l];
            const float f1 = tt / Kf, f2 = (1.0f - tt) / bsum_next;
#pragma unroll
            for (int i = 0; i < MW; i++) {
                if (i < W) {
                    const bool bit = (row[i] >> lane) & 1ull;
                    const float av = src[64 * i];
                    const float bt = valid[i] ? fmaf(b[i], f2, f1) : 0.0f;
                    const float dp = av * bt;
                    if (bit) p1 += dp; else p0 += dp;
                    b[i] = fl ? bt : bt * (bit ? e1 : e0);
                    bs += b[i];
                }
            }
        }
        wave_sum3(p0, p1, bs);
        if (fl && l == Lp - 1) bs = 1.0f;
        if (!fl) { p0 /= e0; p1 /= e1; }
        float o0 = p0 * ee + p1 * ed, o1 = p0 * ed + p1 * ee;
        if (!isflat) { o0 *= h0; o1 *= h1; }
        if (lane == 0) {
            HP[2 * site] = o0 / (o0 + o1);
            HP[2 * site + 1] = o1 / (o0 + o1);
        }
        bsum_next = bs;
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

template <int MW>
static void fb_go(int n, const FbItem *d_fb, const CondItem *d_cond, const int32_t *d_poly, const float *d_t,
                  const uint64_t *d_hvar, const float *d_HL, const uint8_t *d_flat, const uint8_t *d_flags,
                  float *d_alpha, float *d_HP, float ee, float ed, hipStream_t s) {
    hipLaunchKernelGGL(fb_kernel<MW>, dim3(n), dim3(64), 0, s, d_fb, d_cond, d_poly, d_t, d_hvar, d_HL, d_flat,
                       d_flags, d_alpha, d_HP, ee, ed);
}

void fb_launch(gh_panel *P, int n, const FbItem *d_fb, const CondItem *d_cond, int maxW, const int32_t *d_poly,
               const float *d_t, const uint64_t *d_hvar, const float *d_HL, const uint8_t *d_flat, float *d_alpha,
               float *d_HP, float ee, float ed, hipStream_t s) {
    if (n <= 0) return;
    const uint8_t *fl = P->d_flags.as<uint8_t>();
    if (maxW <= 2) fb_go<2>(n, d_fb, d_cond, d_poly, d_t, d_hvar, d_HL, d_flat, fl, d_alpha, d_HP, ee, ed, s);
    else if (maxW <= 4) fb_go<4>(n, d_fb, d_cond, d_poly, d_t, d_hvar, d_HL, d_flat, fl, d_alpha, d_HP, ee, ed, s);
    else if (maxW <= 8) fb_go<8>(n, d_fb, d_cond, d_poly, d_t, d_hvar, d_HL, d_flat, fl, d_alpha, d_HP, ee, ed, s);
    else if (maxW <= 16) fb_go<16>(n, d_fb, d_cond, d_poly, d_t, d_hvar, d_HL, d_flat, fl, d_alpha, d_HP, ee, ed, s);
    else if (maxW <= 32) fb_go<32>(n, d_fb, d_cond, d_poly, d_t, d_hvar, d_HL, d_flat, fl, d_alpha, d_HP, ee, ed, s);
    else throw Error(GH_ERR_ARG, "fb: more than 2048 conditioning states per haplotype is not supported");
    GH_HIP(hipGetLastError());
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
