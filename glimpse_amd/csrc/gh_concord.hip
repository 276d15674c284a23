// gh_concord.hip -- genotype concordance metrics of the concordance tool
// (concordance/src/containers/call_set_reading.cpp:476-760; the truth and
// estimate rules of call_set_header.h:162-259).
//
// Every (site, sample) genotype is classified independently, so the work is a
// streaming pass over the site-major inputs with reductions in three
// directions: per site (and from there per frequency bin / site group), per
// sample, and per calibration bin.  All reductions are in a fixed order:
//   site_kernel    one wave per site; lanes stride the samples, wave
//                  reductions give the site's record; calibration and
//                  rejection counts go through LDS integer histograms.
//   sample_kernel  one thread per sample over a range of sites; partial
//                  records per (range, sample), summed by sample_final_kernel
//                  over the ranges in order.
//   group_kernel   one workgroup per bin / group over the site records.
// Integer counts are exact.  The DS sums are double sums in a fixed order; the
// reference accumulates the same quantities with Welford updates
// (basic_stats.h:53-75), so correlations agree to double rounding.
#include <cmath>
#include <cstring>

#include "gh_device.hpp"
#include "gh_internal.hpp"

namespace gh {

constexpr int CC_NCNT = GH_CONC_NCNT;   // 9 table cells, 3 usable truths, filtered
constexpr int CC_NSUM = GH_CONC_NSUM;   // sum x, sum x^2, sum x*y, min x, max x
constexpr int CC_NCAL = GH_CONC_NCAL;
constexpr int CC_NREJ = GH_CONC_NREJ;
constexpr int CC_MAX_GROUPS = 1 << 20;

struct ConcDev {
    int S, N, P;
    int n_bins, n_groups;
    float T;
    int D;
    float gp_filter;
    int use_alt_af, truth_gt, esti_gt;
    const double *bins;       // n_bins + 1 edges
    const int32_t *site_group;
    const float *unph;        // 256
    const float *af;
    const uint8_t *type;
    const int32_t *truth_pl;
    const int8_t *truth_gtv;
    const int32_t *truth_dp;
    const float *esti_gp;
    const float *esti_ds;
    const int8_t *esti_gtv;
};

// A site's classification: its group (bin) and type, and whether its alleles flip.
struct SiteInfo {
    int tag;     // 2 * group + type index, -1: skipped
    int rej;     // rejection slot for a skipped site, -1 otherwise
    bool flip;
    float maf;
};

__device__ __forceinline__ SiteInfo site_info(const ConcDev &c, int s) {
    SiteInfo si;
    si.tag = -1;
    si.rej = -1;
    const float af = c.af[s];
    const int ty = c.type[s];
    si.flip = c.use_alt_af ? false : af > 0.5f;
    si.maf = c.use_alt_af ? af : ((1.0f - af) < af ? 1.0f - af : af);   // std::min(af, 1 - af)
    if (isnan(af) || (ty != 1 && ty != 2)) {
        si.rej = 9;
        return si;
    }
    int grp = -1;
    if (c.n_bins > 0) {
        // getFrequencyBin (call_set_header.h:207-212): the float maf against double edges
        const double m = (double)si.maf;
        for (int b = 1; b <= c.n_bins; b++)
            if (m > c.bins[b - 1] && m <= c.bins[b]) {
                grp = b - 1;
                break;
            }
    } else {
        grp = c.site_group[s];
        if (grp >= c.n_groups) grp = -1;
    }
    if (grp < 0) {
        si.rej = 10;
        return si;
    }
    si.tag = 2 * grp + (ty == 2 ? 1 : 0);
    return si;
}

struct Cell {
    int truth;   // >= 0 class, else the reference's rejection code
    int esti;    // >= 0 class, else code (meaningful when truth >= 0)
    int cal;     // calibration bin when esti >= 0
    float ds;
};

// getTruth (call_set_header.h:175-205)
__device__ __forceinline__ int truth_class(float g0, float g1, float g2, int P, float T) {
#pragma clang fp contract(off)
    if (P == 1) {
        if (g0 < 0.0f || g1 < 0.0f) return -3;
        const float sc = 1.0f / (g0 + g1);
        const float p0 = g0 * sc, p1 = g1 * sc;
        if (p0 < T && p1 < T) return -4;
        if (p0 > p1) return 0;
        if (p1 > p0) return 1;
        return -1;
    }
    if (g0 < 0.0f || g1 < 0.0f || g2 < 0.0f) return -3;
    const float sc = (float)(1.0 / (double)((g0 + g1) + g2));
    const float p0 = g0 * sc, p1 = g1 * sc, p2 = g2 * sc;
    if (p0 < T && p1 < T && p2 < T) return -4;
    if (p0 > p1 && p0 > p2) return 0;
    if (p1 > p0 && p1 > p2) return 1;
    if (p2 > p0 && p2 > p1) return 2;
    return -1;
}

// getMostLikely (call_set_header.h:214-246).  Its DS range test can never be
// true and is left out.
__device__ __forceinline__ int esti_class(float g0, float g1, float g2, int P, float gp_filter) {
    const double hi = 1.0 + 1e-7;   // 1.0f + 1e-7 is a double
    if (g0 < 0.0f || (double)g0 > hi) return -3;
    if (g1 < 0.0f || (double)g1 > hi) return -3;
    if (P == 1) {
        if (fmaxf(g0, g1) < gp_filter) return -4;
        if (g0 > g1) return 0;
        if (g1 > g0) return 1;
        return -1;
    }
    if (g2 < 0.0f || (double)g2 > hi) return -3;
    if (fmaxf(g0, fmaxf(g1, g2)) < gp_filter) return -4;
    if (g0 > g1 && g0 > g2) return 0;
    if (g1 > g0 && g1 > g2) return 1;
    if (g2 > g0 && g2 > g1) return 2;
    return 0;   // ties: the reference returns 0 here
}

__device__ __forceinline__ Cell classify(const ConcDev &c, int s, int i, bool flip) {
#pragma clang fp contract(off)
    Cell r;
    const int P = c.P;
    const size_t o = (size_t)s * c.N + i;
    // truth (call_set_reading.cpp:519-563, getTruth)
    int dp = 0;
    if (c.D > 0) dp = c.truth_dp[o];
    float g[3];
    if (c.truth_gt) {
        const int gt = c.truth_gtv[o];
        if (gt < 0) {
            g[0] = g[1] = g[2] = -1.0f;
        } else {
            g[0] = gt == 0 ? 1.0f : 0.0f;
            g[1] = gt == 1 ? 1.0f : 0.0f;
            g[2] = gt == 2 ? 1.0f : 0.0f;
            if (P == 1) g[2] = g[1];
        }
    } else {
        const int32_t *pl = c.truth_pl + o * (P + 1);
        bool miss = false;
        int v[3];
        for (int k = 0; k <= P; k++) {
            v[k] = pl[k];
            miss = miss || v[k] < 0;   // htslib's missing and vector-end values are negative
        }
        if (miss) {
            g[0] = g[1] = g[2] = -1.0f;
        } else {
            g[0] = c.unph[min(v[0], 255)];
            g[1] = c.unph[min(v[1], 255)];
            g[2] = c.unph[min(v[P], 255)];
        }
    }
    if (flip) {
        const float t = g[2];
        g[2] = g[0];
        g[0] = t;
        if (P == 1) g[1] = g[2];
    }
    if (c.D > 0 && dp == INT32_MIN) r.truth = -5;
    else if (dp < c.D) r.truth = -2;
    else r.truth = truth_class(g[0], g[1], g[2], P, c.T);
    r.esti = -3;
    r.cal = 0;
    r.ds = 0.0f;
    if (r.truth < 0) return r;
    // estimate (call_set_reading.cpp:567-642, getMostLikely)
    float e[3], ds;
    bool miss;
    if (c.esti_gt) {
        const int gt = c.esti_gtv[o];
        miss = gt < 0;
        ds = (float)gt;
        e[0] = gt == 0 ? 1.0f : 0.0f;
        e[1] = gt == 1 ? 1.0f : 0.0f;
        e[2] = gt == P ? 1.0f : 0.0f;
    } else {
        miss = c.esti_gtv != nullptr && c.esti_gtv[o] < 0;
        const float *gp = c.esti_gp + o * (P + 1);
        e[0] = gp[0];
        e[1] = gp[1];
        e[2] = gp[P];
        miss = miss || isnan(e[0]) || isnan(e[1]) || isnan(e[2]);
        ds = c.esti_ds[o];
    }
    if (miss) {
        r.esti = -3;
        return r;
    }
    if (flip) {
        const float t = e[2];
        e[2] = e[0];
        e[0] = t;
        if (P == 1) e[1] = e[2];
        ds = (float)P - ds;
    }
    r.esti = esti_class(e[0], e[1], e[2], P, c.gp_filter);
    r.ds = ds;
    if (r.esti >= 0) r.cal = (int)truncf(fmaxf(e[0], fmaxf(e[1], e[2])) * (float)(CC_NCAL - 1));
    return r;
}

// Register accumulator of one record (counts kept in 32 bits per thread).
struct Acc {
    uint32_t n[CC_NCNT];
    double sx, sxx, sxy;
    float mn, mx;
    __device__ __forceinline__ void clear() {
#pragma unroll
        for (int k = 0; k < CC_NCNT; k++) n[k] = 0;
        sx = sxx = sxy = 0.0;
        mn = INFINITY;
        mx = -INFINITY;
    }
    // Adds a cell with a usable truth (no dynamic register indexing).
    __device__ __forceinline__ void add(const Cell &c) {
        const int cell = c.esti >= 0 ? 3 * c.truth + c.esti : -1;
        const int filt = c.esti == -4 ? 12 : -1;
#pragma unroll
        for (int k = 0; k < CC_NCNT; k++) n[k] += (k == cell) + (k == 9 + c.truth) + (k == filt);
        if (c.esti >= 0) {
            const double x = (double)c.ds, y = (double)c.truth;
            sx += x;
            sxx += x * x;
            sxy += x * y;
            mn = fminf(mn, c.ds);
            mx = fmaxf(mx, c.ds);
        }
    }
};

__device__ __forceinline__ uint32_t wsum_u(uint32_t v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ __forceinline__ double wsum_d(double v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}
__device__ __forceinline__ float wmin_f(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fminf(v, __shfl_xor(v, m, 64));
    return v;
}
__device__ __forceinline__ float wmax_f(float v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v = fmaxf(v, __shfl_xor(v, m, 64));
    return v;
}

constexpr int SITE_WAVES = 4;   // waves (sites in flight) per workgroup of site_kernel

// One wave per site, grid-stride over sites.  Writes the site's record and tag;
// calibration and rejection counts accumulate in LDS and are added to the
// global 64-bit counters once per workgroup.
__global__ __launch_bounds__(64 * SITE_WAVES) void site_kernel(ConcDev c, uint32_t *__restrict__ site_cnt,
                                                               double *__restrict__ site_sum,
                                                               int32_t *__restrict__ site_tag,
                                                               unsigned long long *__restrict__ cal_cnt,
                                                               unsigned long long *__restrict__ rej_cnt) {
    __shared__ uint32_t cal[CC_NCAL * 2 * 9];
    __shared__ uint32_t rej[CC_NREJ];
    for (int k = threadIdx.x; k < CC_NCAL * 2 * 9; k += blockDim.x) cal[k] = 0;
    for (int k = threadIdx.x; k < CC_NREJ; k += blockDim.x) rej[k] = 0;
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int s = blockIdx.x * SITE_WAVES + wave; s < c.S; s += gridDim.x * SITE_WAVES) {
        const SiteInfo si = site_info(c, s);
        if (lane == 0) site_tag[s] = si.tag;
        Acc a;
        a.clear();
        if (si.tag >= 0) {
            const int ty = si.tag & 1;
            for (int i = lane; i < c.N; i += 64) {
                const Cell x = classify(c, s, i, si.flip);
                if (x.truth < 0) {
                    atomicAdd(&rej[-x.truth - 1], 1u);
                    continue;
                }
                a.add(x);
                if (x.esti >= 0) atomicAdd(&cal[(x.cal * 2 + ty) * 9 + 3 * x.truth + x.esti], 1u);
                else atomicAdd(&rej[5 + (-x.esti - 1)], 1u);
            }
        } else if (lane == 0) {
            atomicAdd(&rej[si.rej], 1u);
        }
#pragma unroll
        for (int k = 0; k < CC_NCNT; k++) a.n[k] = wsum_u(a.n[k]);
        a.sx = wsum_d(a.sx);
        a.sxx = wsum_d(a.sxx);
        a.sxy = wsum_d(a.sxy);
        a.mn = wmin_f(a.mn);
        a.mx = wmax_f(a.mx);
        if (lane < CC_NCNT) {
            uint32_t v = 0;
#pragma unroll
            for (int k = 0; k < CC_NCNT; k++) v = lane == k ? a.n[k] : v;
            site_cnt[(size_t)s * CC_NCNT + lane] = v;
        }
        if (lane == 0) {
            double *o = site_sum + (size_t)s * CC_NSUM;
            o[0] = a.sx;
            o[1] = a.sxx;
            o[2] = a.sxy;
            o[3] = (double)a.mn;
            o[4] = (double)a.mx;
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < CC_NCAL * 2 * 9; k += blockDim.x)
        if (cal[k]) atomicAdd(&cal_cnt[k], (unsigned long long)cal[k]);
    for (int k = threadIdx.x; k < CC_NREJ; k += blockDim.x)
        if (rej[k]) atomicAdd(&rej_cnt[k], (unsigned long long)rej[k]);
}

// Partial per-sample records over a range of sites.  Layout (coalesced over
// samples): pc [range][type][NCNT][N] counts, ps [range][type][3][N] sums,
// pm [range][type][2][N] min / max.
__global__ __launch_bounds__(256) void sample_kernel(ConcDev c, const int32_t *__restrict__ site_tag, int per_range,
                                                     uint32_t *__restrict__ pc, double *__restrict__ ps,
                                                     float *__restrict__ pm) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int r = blockIdx.y;
    if (i >= c.N) return;
    const int s0 = r * per_range, s1 = min(c.S, s0 + per_range);
    Acc a0, a1;
    a0.clear();
    a1.clear();
    for (int s = s0; s < s1; s++) {
        const int tag = site_tag[s];
        if (tag < 0) continue;
        const float af = c.af[s];
        const bool flip = c.use_alt_af ? false : af > 0.5f;
        const Cell x = classify(c, s, i, flip);
        if (x.truth < 0) continue;
        if (tag & 1) a1.add(x);
        else a0.add(x);
    }
    const size_t N = c.N;
    for (int t = 0; t < 2; t++) {
        const Acc &a = t ? a1 : a0;
        const size_t base = (size_t)r * 2 + t;
#pragma unroll
        for (int k = 0; k < CC_NCNT; k++) pc[(base * CC_NCNT + k) * N + i] = a.n[k];
        ps[(base * 3 + 0) * N + i] = a.sx;
        ps[(base * 3 + 1) * N + i] = a.sxx;
        ps[(base * 3 + 2) * N + i] = a.sxy;
        pm[(base * 2 + 0) * N + i] = a.mn;
        pm[(base * 2 + 1) * N + i] = a.mx;
    }
}

// Sums the ranges in order: one thread per (sample, type).
__global__ void sample_final_kernel(int N, int R, const uint32_t *__restrict__ pc, const double *__restrict__ ps,
                                    const float *__restrict__ pm, uint64_t *__restrict__ spl_cnt,
                                    double *__restrict__ spl_sum) {
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= 2 * N) return;
    const int i = g >> 1, t = g & 1;
    uint64_t n[CC_NCNT];
    for (int k = 0; k < CC_NCNT; k++) n[k] = 0;
    double sx = 0, sxx = 0, sxy = 0;
    float mn = INFINITY, mx = -INFINITY;
    for (int r = 0; r < R; r++) {
        const size_t base = (size_t)r * 2 + t;
        for (int k = 0; k < CC_NCNT; k++) n[k] += pc[(base * CC_NCNT + k) * N + i];
        sx += ps[(base * 3 + 0) * N + i];
        sxx += ps[(base * 3 + 1) * N + i];
        sxy += ps[(base * 3 + 2) * N + i];
        mn = fminf(mn, pm[(base * 2 + 0) * N + i]);
        mx = fmaxf(mx, pm[(base * 2 + 1) * N + i]);
    }
    for (int k = 0; k < CC_NCNT; k++) spl_cnt[((size_t)i * 2 + t) * CC_NCNT + k] = n[k];
    double *o = spl_sum + ((size_t)i * 2 + t) * CC_NSUM;
    o[0] = sx;
    o[1] = sxx;
    o[2] = sxy;
    o[3] = (double)mn;
    o[4] = (double)mx;
}

constexpr int GRP_THREADS = 256;
constexpr int GRP_VALS = CC_NCNT + 4;   // counts, sum x, sum x^2, sum x*y, sum maf (min / max apart)

// One workgroup per group and type: threads stride the sites, then a fixed
// tree over the workgroup.
__global__ __launch_bounds__(GRP_THREADS) void group_kernel(ConcDev c, const int32_t *__restrict__ site_tag,
                                                            const uint32_t *__restrict__ site_cnt,
                                                            const double *__restrict__ site_sum,
                                                            uint64_t *__restrict__ grp_cnt,
                                                            double *__restrict__ grp_sum) {
    __shared__ double red[GRP_THREADS / 64][GRP_VALS];
    __shared__ float redm[GRP_THREADS / 64][2];
    const int want = blockIdx.x;   // 2 * group + type
    uint64_t n[CC_NCNT];
    for (int k = 0; k < CC_NCNT; k++) n[k] = 0;
    double sx = 0, sxx = 0, sxy = 0, smaf = 0;
    float mn = INFINITY, mx = -INFINITY;
    for (int s = threadIdx.x; s < c.S; s += GRP_THREADS) {
        if (site_tag[s] != want) continue;
        const uint32_t *sc = site_cnt + (size_t)s * CC_NCNT;
        uint32_t ne = 0;
        for (int k = 0; k < CC_NCNT; k++) n[k] += sc[k];
        for (int k = 0; k < 9; k++) ne += sc[k];
        const double *ss = site_sum + (size_t)s * CC_NSUM;
        sx += ss[0];
        sxx += ss[1];
        sxy += ss[2];
        mn = fminf(mn, (float)ss[3]);
        mx = fmaxf(mx, (float)ss[4]);
        const float af = c.af[s];
        const float maf = c.use_alt_af ? af : ((1.0f - af) < af ? 1.0f - af : af);
        smaf += (double)ne * (double)maf;
    }
    double v[GRP_VALS];
    for (int k = 0; k < CC_NCNT; k++) v[k] = (double)n[k];   // exact: counts < 2^53
    v[CC_NCNT + 0] = sx;
    v[CC_NCNT + 1] = sxx;
    v[CC_NCNT + 2] = sxy;
    v[CC_NCNT + 3] = smaf;
    for (int k = 0; k < GRP_VALS; k++) v[k] = wsum_d(v[k]);
    mn = wmin_f(mn);
    mx = wmax_f(mx);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    if (lane == 0) {
        for (int k = 0; k < GRP_VALS; k++) red[wave][k] = v[k];
        redm[wave][0] = mn;
        redm[wave][1] = mx;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < GRP_THREADS / 64; w++) {
            for (int k = 0; k < GRP_VALS; k++) red[0][k] += red[w][k];
            redm[0][0] = fminf(redm[0][0], redm[w][0]);
            redm[0][1] = fmaxf(redm[0][1], redm[w][1]);
        }
        for (int k = 0; k < CC_NCNT; k++) grp_cnt[(size_t)want * CC_NCNT + k] = (uint64_t)red[0][k];
        double *o = grp_sum + (size_t)want * (CC_NSUM + 1);
        o[0] = red[0][CC_NCNT + 0];
        o[1] = red[0][CC_NCNT + 1];
        o[2] = red[0][CC_NCNT + 2];
        o[3] = (double)redm[0][0];
        o[4] = (double)redm[0][1];
        o[5] = red[0][CC_NCNT + 3];
    }
}

// common/src/utils/otools.h:97-98 (the concordance tool's copy is the same
// table): 10^(-i/10) printed with 7 significant digits.
static void conc_unphred(float *t) {
    char buf[64];
    for (int i = 0; i < 256; i++) {
        std::snprintf(buf, sizeof(buf), "%.6e", std::pow(10.0, -i / 10.0));
        t[i] = std::strtof(buf, nullptr);
    }
}

void concordance(const gh_conc_params *prm, const gh_conc_input *in, gh_conc_output *out) {
    GH_CHECK(prm && in && out, GH_ERR_ARG, "gh_concordance: null argument");
    const int S = in->n_sites, N = in->n_samples, P = prm->ploidy;
    GH_CHECK(S >= 0 && N > 0 && (P == 1 || P == 2), GH_ERR_ARG, "gh_concordance: bad sizes or ploidy");
    GH_CHECK(in->af && in->type, GH_ERR_ARG, "gh_concordance: af and type are required");
    GH_CHECK(prm->truth_from_gt ? in->truth_gt != nullptr : in->truth_pl != nullptr, GH_ERR_ARG,
             "gh_concordance: truth input missing");
    GH_CHECK(prm->min_val_dp <= 0 || in->truth_dp != nullptr, GH_ERR_ARG,
             "gh_concordance: min_val_dp > 0 needs truth_dp");
    GH_CHECK(prm->esti_from_gt ? in->esti_gt != nullptr : (in->esti_gp && in->esti_ds), GH_ERR_ARG,
             "gh_concordance: estimate input missing");
    int G;
    if (prm->n_bins > 0) {
        GH_CHECK(prm->bins && prm->n_bins <= 4096, GH_ERR_ARG, "gh_concordance: bins");
        G = prm->n_bins;
    } else {
        GH_CHECK(prm->site_group && prm->n_groups > 0 && prm->n_groups <= CC_MAX_GROUPS, GH_ERR_ARG,
                 "gh_concordance: neither frequency bins nor site groups");
        G = prm->n_groups;
    }
    GH_CHECK(out->spl_cnt && out->spl_sum && out->grp_cnt && out->grp_sum && out->cal_cnt && out->rejected,
             GH_ERR_ARG, "gh_concordance: output arrays");
    std::memset(out->spl_cnt, 0, sizeof(uint64_t) * N * 2 * CC_NCNT);
    std::memset(out->spl_sum, 0, sizeof(double) * N * 2 * CC_NSUM);
    std::memset(out->grp_cnt, 0, sizeof(uint64_t) * (size_t)G * 2 * CC_NCNT);
    std::memset(out->grp_sum, 0, sizeof(double) * (size_t)G * 2 * (CC_NSUM + 1));
    std::memset(out->cal_cnt, 0, sizeof(uint64_t) * CC_NCAL * 2 * 9);
    std::memset(out->rejected, 0, sizeof(uint64_t) * CC_NREJ);
    if (S == 0) return;

    hipStream_t s = nullptr;
    GH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    const size_t SN = (size_t)S * N;
    float unph[256];
    conc_unphred(unph);
    DBuf d_unph, d_bins, d_grp, d_af, d_type, d_tpl, d_tgt, d_tdp, d_egp, d_eds, d_egt;
    ConcDev c{};
    c.S = S;
    c.N = N;
    c.P = P;
    c.n_bins = prm->n_bins > 0 ? prm->n_bins : 0;
    c.n_groups = G;
    c.T = prm->min_val_gl;
    c.D = prm->min_val_dp > 0 ? prm->min_val_dp : 0;
    c.gp_filter = prm->min_tar_gp;
    c.use_alt_af = prm->use_alt_af != 0;
    c.truth_gt = prm->truth_from_gt != 0;
    c.esti_gt = prm->esti_from_gt != 0;
    c.unph = dalloc_copy(d_unph, unph, 256, s);
    if (c.n_bins) c.bins = dalloc_copy(d_bins, prm->bins, (size_t)c.n_bins + 1, s);
    else c.site_group = dalloc_copy(d_grp, prm->site_group, (size_t)S, s);
    c.af = dalloc_copy(d_af, in->af, (size_t)S, s);
    c.type = dalloc_copy(d_type, in->type, (size_t)S, s);
    if (c.truth_gt) c.truth_gtv = dalloc_copy(d_tgt, in->truth_gt, SN, s);
    else c.truth_pl = dalloc_copy(d_tpl, in->truth_pl, SN * (P + 1), s);
    if (c.D > 0) c.truth_dp = dalloc_copy(d_tdp, in->truth_dp, SN, s);
    if (c.esti_gt) {
        c.esti_gtv = dalloc_copy(d_egt, in->esti_gt, SN, s);
    } else {
        c.esti_gp = dalloc_copy(d_egp, in->esti_gp, SN * (P + 1), s);
        c.esti_ds = dalloc_copy(d_eds, in->esti_ds, SN, s);
        if (in->esti_gt) c.esti_gtv = dalloc_copy(d_egt, in->esti_gt, SN, s);
    }

    // per-site records, calibration and rejections
    DBuf d_scnt, d_ssum, d_stag, d_cal, d_rej;
    d_scnt.reserve(sizeof(uint32_t) * (size_t)S * CC_NCNT);
    d_ssum.reserve(sizeof(double) * (size_t)S * CC_NSUM);
    d_stag.reserve(sizeof(int32_t) * (size_t)S);
    d_cal.reserve(sizeof(uint64_t) * CC_NCAL * 2 * 9);
    d_rej.reserve(sizeof(uint64_t) * CC_NREJ);
    GH_HIP(hipMemsetAsync(d_cal.p, 0, d_cal.bytes, s));
    GH_HIP(hipMemsetAsync(d_rej.p, 0, d_rej.bytes, s));
    const int site_blocks = (int)std::min<int64_t>(((int64_t)S + SITE_WAVES - 1) / SITE_WAVES, 4096);
    hipLaunchKernelGGL(site_kernel, dim3(site_blocks), dim3(64 * SITE_WAVES), 0, s, c, d_scnt.as<uint32_t>(),
                       d_ssum.as<double>(), d_stag.as<int32_t>(), d_cal.as<unsigned long long>(),
                       d_rej.as<unsigned long long>());
    GH_HIP(hipGetLastError());

    // per-sample records: about 2048 workgroups of 256 samples x a range of sites
    const int nb = (N + 255) / 256;
    const int R = (int)std::max<int64_t>(1, std::min<int64_t>(S, 2048 / nb));
    const int per_range = (S + R - 1) / R;
    DBuf d_pc, d_ps, d_pm, d_scnt_out, d_ssum_out;
    d_pc.reserve(sizeof(uint32_t) * (size_t)R * 2 * CC_NCNT * N);
    d_ps.reserve(sizeof(double) * (size_t)R * 2 * 3 * N);
    d_pm.reserve(sizeof(float) * (size_t)R * 2 * 2 * N);
    hipLaunchKernelGGL(sample_kernel, dim3(nb, R), dim3(256), 0, s, c, d_stag.as<int32_t>(), per_range,
                       d_pc.as<uint32_t>(), d_ps.as<double>(), d_pm.as<float>());
    GH_HIP(hipGetLastError());
    d_scnt_out.reserve(sizeof(uint64_t) * (size_t)N * 2 * CC_NCNT);
    d_ssum_out.reserve(sizeof(double) * (size_t)N * 2 * CC_NSUM);
    hipLaunchKernelGGL(sample_final_kernel, dim3((2 * N + 255) / 256), dim3(256), 0, s, N, R, d_pc.as<uint32_t>(),
                       d_ps.as<double>(), d_pm.as<float>(), d_scnt_out.as<uint64_t>(), d_ssum_out.as<double>());
    GH_HIP(hipGetLastError());

    // per group
    DBuf d_gcnt, d_gsum;
    d_gcnt.reserve(sizeof(uint64_t) * (size_t)G * 2 * CC_NCNT);
    d_gsum.reserve(sizeof(double) * (size_t)G * 2 * (CC_NSUM + 1));
    hipLaunchKernelGGL(group_kernel, dim3(2 * G), dim3(GRP_THREADS), 0, s, c, d_stag.as<int32_t>(),
                       d_scnt.as<uint32_t>(), d_ssum.as<double>(), d_gcnt.as<uint64_t>(), d_gsum.as<double>());
    GH_HIP(hipGetLastError());

    GH_HIP(hipMemcpyAsync(out->spl_cnt, d_scnt_out.p, d_scnt_out.bytes, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(out->spl_sum, d_ssum_out.p, d_ssum_out.bytes, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(out->grp_cnt, d_gcnt.p, d_gcnt.bytes, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(out->grp_sum, d_gsum.p, d_gsum.bytes, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(out->cal_cnt, d_cal.p, d_cal.bytes, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(out->rejected, d_rej.p, d_rej.bytes, hipMemcpyDeviceToHost, s));
    if (out->site_cnt)
        GH_HIP(hipMemcpyAsync(out->site_cnt, d_scnt.p, sizeof(uint32_t) * (size_t)S * CC_NCNT,
                              hipMemcpyDeviceToHost, s));
    if (out->site_sum)
        GH_HIP(hipMemcpyAsync(out->site_sum, d_ssum.p, sizeof(double) * (size_t)S * CC_NSUM, hipMemcpyDeviceToHost,
                              s));
    if (out->site_tag)
        GH_HIP(hipMemcpyAsync(out->site_tag, d_stag.p, sizeof(int32_t) * (size_t)S, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
}

double conc_r2(uint64_t n, double sx, double sxx, double sxy, double sy, double syy, bool xc, bool yc) {
    if (n == 0) return 0.0;
    if (n == 1) return NAN;
    if (xc || yc) return 0.0;
    const long double N = (long double)n;
    const long double cxy = N * sxy - (long double)sx * sy;
    const long double vx = N * sxx - (long double)sx * sx;
    const long double vy = N * syy - (long double)sy * sy;
    if (vx <= 0 || vy <= 0) return 0.0;
    return (double)(cxy * cxy / (vx * vy));
}

}  // namespace gh
