// gh_graph.hip -- device-side placement and draws of a replayed Gibbs
// iteration (see gh_graph.hpp).  Each kernel is one workgroup of PL_NT
// threads walking the targets in chunks, with exclusive scans for the
// offsets and the launch orders, so every decision is the one the host makes
// for an iteration it launches itself (gh_gibbs.hip, run_execute).
#include "gh_device.hpp"
#include "gh_graph.hpp"

namespace gh {

constexpr int PL_NT = 1024;

// Exclusive scan of one value per thread over the workgroup; sh holds PL_NT
// values.  Returns the thread's prefix and sets total to the workgroup's sum.
template <typename V>
__device__ V block_scan(V v, V *sh, V &total) {
    const int tid = threadIdx.x;
    sh[tid] = v;
    __syncthreads();
    for (int o = 1; o < PL_NT; o <<= 1) {
        const V y = tid >= o ? sh[tid - o] : V(0);
        __syncthreads();
        sh[tid] += y;
        __syncthreads();
    }
    total = sh[PL_NT - 1];
    const V ex = sh[tid] - v;
    __syncthreads();
    return ex;
}

// haplotype_set.cpp:335-348 with counter draws: one stored site per PBWT group
// and class, one start index per walked haplotype.  The same draws as the
// host's (philox_draw and urange are shared host/device code).
__global__ __launch_bounds__(PL_NT) void pbwt_draw_kernel(const GraphCtl *__restrict__ ctl, uint32_t seed,
                                                          const int32_t *__restrict__ grp_ptr,
                                                          const int32_t *__restrict__ grp_end, int ncls, int Lhq,
                                                          uint8_t *__restrict__ stored,
                                                          const int32_t *__restrict__ targets, int npt, int N,
                                                          int32_t *__restrict__ start) {
    const uint32_t it = ctl->iter;
    const int tid = threadIdx.x;
    const int64_t n = (int64_t)ncls * Lhq;
    for (int64_t i = tid; i < n; i += PL_NT) stored[i] = 0;
    __syncthreads();
    for (int c = 0; c < ncls; c++) {
        const int g0 = grp_ptr[c], ng = grp_ptr[c + 1] - g0;
        for (int gi = tid; gi < ng; gi += PL_NT) {
            const uint32_t lo = gi == 0 ? 0u : (uint32_t)grp_end[g0 + gi - 1];
            const uint32_t hi = (uint32_t)grp_end[g0 + gi] - 1u;
            const uint32_t x = philox_draw(seed, it, P_PBWT_STORED, (uint32_t)gi, (uint32_t)c);
            const uint32_t l = urange(x, lo, hi);
            if (l < (uint32_t)Lhq) stored[(int64_t)c * Lhq + l] = 1;
        }
    }
    for (int i = tid; i < 2 * npt; i += PL_NT) {
        const uint32_t x = philox_draw(seed, it, P_PBWT_START, (uint32_t)(2 * targets[i >> 1] + (i & 1)), 0u);
        start[i] = (int32_t)urange(x, 0u, (uint32_t)(N - 1));
    }
}

// Places every target's Hvar rows, alpha checkpoints and phasing loci by its
// polymorphic-site count (in target order, as the host does), splits the
// targets into the forward-backward launch groups and records the iteration's
// set sizes and counters.  On overflow nothing is placed: every item gets
// Lp = W = 0, so the conditioning-set kernels write no rows, and both groups
// are empty.
__global__ __launch_bounds__(PL_NT) void fb_place_kernel(GraphCtl *__restrict__ ctl, CondItem *__restrict__ items,
                                                         FbItem *__restrict__ fb, PhItem *__restrict__ ph, int T,
                                                         int L, int C, int few, int latency, GraphCaps caps,
                                                         int32_t *__restrict__ ford, int32_t *__restrict__ nstates,
                                                         int n_iter) {
    __shared__ long long sh[PL_NT];
    __shared__ unsigned long long tot[3], st[3];
    __shared__ int s_top, s_capB;
    const int tid = threadIdx.x;
    const uint32_t iter = ctl->iter;
    if (tid == 0) {
        tot[0] = tot[1] = tot[2] = 0;
        st[0] = st[1] = st[2] = 0;
        s_top = 0;
        s_capB = 0;
    }
    __syncthreads();
    for (int t = tid; t < T; t += PL_NT) {
        const CondItem c = items[t];
        atomicAdd(&tot[0], (unsigned long long)((long long)c.Lp * c.W));
        atomicAdd(&tot[1], (unsigned long long)((long long)((c.Lp + C - 1) / C) * 64 * c.W + c.Lp));
        atomicAdd(&tot[2], (unsigned long long)c.Lp);
        atomicMax(&s_top, fb_cap_class(c.W));
    }
    __syncthreads();
    if ((long long)tot[0] > caps.hvar_words || (long long)tot[1] > caps.alpha_floats ||
        (long long)tot[2] > caps.loci) {
        for (int t = tid; t < T; t += PL_NT) {
            items[t].Lp = 0;
            items[t].W = 0;
        }
        if (tid == 0) {
            ctl->overflow = 1;
            ctl->fb_n[0] = ctl->fb_n[1] = 0;
        }
        return;
    }
    const int top = s_top;
    for (int t = tid; t < T; t += PL_NT) {
        const int cw = fb_cap_class(items[t].W);
        if (cw < top) atomicMax(&s_capB, cw);
    }
    __syncthreads();
    long long ch = 0, ca = 0, cl = 0, cA = 0, cB = 0;
    for (int base = 0; base < T; base += PL_NT) {
        const int t = base + tid;
        const bool live = t < T;
        CondItem c{};
        if (live) c = items[t];
        const long long nh = live ? (long long)c.Lp * c.W : 0;
        const long long na = live ? (long long)((c.Lp + C - 1) / C) * 64 * c.W + c.Lp : 0;
        const long long nl = live ? c.Lp : 0;
        const long long inA = live && fb_cap_class(c.W) == top, inB = live && !inA;
        long long th, ta, tl, tA, tB;
        const long long oh = block_scan(nh, sh, th);
        const long long oa = block_scan(na, sh, ta);
        const long long ol = block_scan(nl, sh, tl);
        const long long pA = block_scan(inA, sh, tA);
        const long long pB = block_scan(inB, sh, tB);
        if (live) {
            items[t].hvar_off = ch + oh;
            fb[t].alpha_off = ca + oa;
            ph[t].locus_off = cl + ol;
            if (inA) ford[cA + pA] = t;
            else ford[T + cB + pB] = t;
            nstates[(size_t)t * n_iter + iter] = c.K;
            // two forward-backward passes per diploid target (run_execute's counters)
            atomicAdd(&st[0], (unsigned long long)(2LL * c.K * c.Lp));
            atomicAdd(&st[1], (unsigned long long)(2LL * ((long long)c.Lp * (16LL * c.W + 8LL * c.K + 16) + 8LL * L)));
            atomicAdd(&st[2], (unsigned long long)(2LL * c.Lp));
        }
        ch += th; ca += ta; cl += tl; cA += tA; cB += tB;
    }
    __syncthreads();
    if (tid == 0) {
        ctl->fb_n[0] = (int32_t)cA;
        ctl->fb_n[1] = (int32_t)cB;
        ctl->fb_key[0] = graph_key(top, latency || cA <= few);
        ctl->fb_key[1] = cB > 0 ? graph_key(s_capB, latency || cB <= few) : -1;
        ctl->stats[GS_FB_STATE_SITES] += st[0];
        ctl->stats[GS_FB_BYTES] += st[1];
        ctl->stats[GS_FB_POLY_SITES] += st[2];
        ctl->stats[GS_FB_LAUNCHES] += cB > 0 ? 4 : 2;
    }
}

// Places the phasing messages by the segment and flat-het counts of the
// prepass and splits the targets into the phasing launch groups.  On overflow
// both groups are empty (no phasing kernel runs).
__global__ __launch_bounds__(PL_NT) void ph_place_kernel(GraphCtl *__restrict__ ctl, PhItem *__restrict__ ph,
                                                         const CondItem *__restrict__ items, int T, int few,
                                                         int latency, GraphCaps caps, int32_t *__restrict__ pord) {
    __shared__ long long sh[PL_NT];
    __shared__ unsigned long long tot[5], st[2];
    __shared__ int s_maxK, s_maxKB;
    const int tid = threadIdx.x;
    if (tid == 0) {
        for (int i = 0; i < 5; i++) tot[i] = 0;
        st[0] = st[1] = 0;
        s_maxK = 1;
        s_maxKB = 1;
    }
    __syncthreads();
    for (int t = tid; t < T; t += PL_NT) {
        const int K = items[t].K;
        const long long msg = (long long)((K + 63) / 64) * 64 * 8;
        const PhItem p = ph[t];
        atomicAdd(&tot[0], (unsigned long long)(p.n_segs * msg));
        atomicAdd(&tot[1], (unsigned long long)(p.n_segs * 9LL));
        atomicAdd(&tot[2], (unsigned long long)(p.n_miss * msg));
        atomicAdd(&tot[3], (unsigned long long)(p.n_miss * 8LL));
        atomicAdd(&tot[4], (unsigned long long)p.n_segs);
        atomicMax(&s_maxK, K);
    }
    __syncthreads();
    if ((long long)tot[0] > caps.seg || (long long)tot[1] > caps.segsum || (long long)tot[2] > caps.imp ||
        (long long)tot[3] > caps.impsum || (long long)tot[4] > caps.dip) {
        if (tid == 0) {
            ctl->overflow = 1;
            ctl->ph_n[0] = ctl->ph_n[1] = 0;
        }
        return;
    }
    const int top = ph_cap_states(s_maxK);
    for (int t = tid; t < T; t += PL_NT) {
        const int K = items[t].K;
        if (ph_cap_states(K) < top) atomicMax(&s_maxKB, K);
    }
    __syncthreads();
    long long cs = 0, css = 0, ci = 0, cis = 0, cd = 0, cA = 0, cB = 0;
    for (int base = 0; base < T; base += PL_NT) {
        const int t = base + tid;
        const bool live = t < T;
        int K = 0;
        PhItem p{};
        if (live) {
            K = items[t].K;
            p = ph[t];
        }
        const long long msg = (long long)((K + 63) / 64) * 64 * 8;
        const long long inA = live && ph_cap_states(K) == top, inB = live && !inA;
        long long t0, t1, t2, t3, t4, tA, tB;
        const long long o0 = block_scan(live ? p.n_segs * msg : 0LL, sh, t0);
        const long long o1 = block_scan(live ? p.n_segs * 9LL : 0LL, sh, t1);
        const long long o2 = block_scan(live ? p.n_miss * msg : 0LL, sh, t2);
        const long long o3 = block_scan(live ? p.n_miss * 8LL : 0LL, sh, t3);
        const long long o4 = block_scan(live ? (long long)p.n_segs : 0LL, sh, t4);
        const long long pA = block_scan(inA, sh, tA);
        const long long pB = block_scan(inB, sh, tB);
        if (live) {
            ph[t].msg = (int32_t)msg;
            ph[t].seg_off = cs + o0;
            ph[t].segsum_off = css + o1;
            ph[t].imp_off = ci + o2;
            ph[t].impsum_off = cis + o3;
            ph[t].dip_off = cd + o4;
            if (inA) pord[cA + pA] = t;
            else pord[T + cB + pB] = t;
            atomicAdd(&st[0], (unsigned long long)((long long)K * p.n_loci));
            atomicAdd(&st[1], (unsigned long long)(p.n_segs + p.n_miss));
        }
        cs += t0; css += t1; ci += t2; cis += t3; cd += t4; cA += tA; cB += tB;
    }
    __syncthreads();
    if (tid == 0) {
        ctl->ph_n[0] = (int32_t)cA;
        ctl->ph_n[1] = (int32_t)cB;
        ctl->ph_key[0] = graph_key(top, latency || cA <= few);
        ctl->ph_key[1] = cB > 0 ? graph_key(ph_cap_states(s_maxKB), latency || cB <= few) : -1;
        ctl->stats[GS_PH_STATE_SITES] += st[0];
        ctl->stats[GS_PH_SEGMENTS] += st[1];
    }
}

__global__ void graph_advance_kernel(GraphCtl *ctl) {
    if (threadIdx.x == 0) ctl->iter += 1;
}

void pbwt_draw_launch(const GraphCtl *ctl, uint32_t seed, const int32_t *grp_ptr, const int32_t *grp_end, int ncls,
                      int Lhq, uint8_t *stored, const int32_t *targets, int npt, int N, int32_t *start,
                      hipStream_t s) {
    hipLaunchKernelGGL(pbwt_draw_kernel, dim3(1), dim3(PL_NT), 0, s, ctl, seed, grp_ptr, grp_end, ncls, Lhq, stored,
                       targets, npt, N, start);
    GH_HIP(hipGetLastError());
}

void fb_place_launch(GraphCtl *ctl, CondItem *items, FbItem *fb, PhItem *ph, int T, int L, int C, int few,
                     bool latency, const GraphCaps &caps, int32_t *ford, int32_t *nstates, int n_iter,
                     hipStream_t s) {
    hipLaunchKernelGGL(fb_place_kernel, dim3(1), dim3(PL_NT), 0, s, ctl, items, fb, ph, T, L, C, few, (int)latency,
                       caps, ford, nstates, n_iter);
    GH_HIP(hipGetLastError());
}

void ph_place_launch(GraphCtl *ctl, PhItem *ph, const CondItem *items, int T, int few, bool latency,
                     const GraphCaps &caps, int32_t *pord, hipStream_t s) {
    hipLaunchKernelGGL(ph_place_kernel, dim3(1), dim3(PL_NT), 0, s, ctl, ph, items, T, few, (int)latency, caps, pord);
    GH_HIP(hipGetLastError());
}

void graph_advance_launch(GraphCtl *ctl, hipStream_t s) {
    hipLaunchKernelGGL(graph_advance_kernel, dim3(1), dim3(64), 0, s, ctl);
    GH_HIP(hipGetLastError());
}

}  // namespace gh
