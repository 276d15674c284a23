// gh_graph.hip -- device-side placement and draws of a replayed Gibbs
// iteration (see gh_graph.hpp).  Each kernel is one wave walking the targets
// in chunks of 64, with wave prefix sums for the offsets and the launch
// orders, so every decision is the one the host makes for an iteration it
// launches itself (gh_gibbs.hip, run_execute).
#include "gh_device.hpp"
#include "gh_graph.hpp"

namespace gh {

// One wave per kernel: a launch of a single small wave finds room on a busy
// GPU at once, where a workgroup of many waves waits for a whole CU to drain
// behind the long-lived workgroups of other runs.
constexpr int PL_NT = 64;

template <typename V>
__device__ __forceinline__ V wave_incl_scan(V v) {
    const int lane = threadIdx.x;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const V y = __shfl_up(v, o, 64);
        if (lane >= o) v += y;
    }
    return v;
}
template <typename V>
__device__ __forceinline__ V wave_sum(V v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ int wave_max(int v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}
// Exclusive prefix of v over the wave; total gets the wave's sum.
template <typename V>
__device__ __forceinline__ V wave_excl(V v, V &total) {
    const V inc = wave_incl_scan(v);
    total = __shfl(inc, 63, 64);
    return inc - v;
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
    const int lane = threadIdx.x;
    const int64_t n = (int64_t)ncls * Lhq;
    const int64_t n4 = n >> 2;
    uint32_t *s4 = reinterpret_cast<uint32_t *>(stored);   // the buffer is 256-byte aligned (hipMalloc)
    for (int64_t i = lane; i < n4; i += PL_NT) s4[i] = 0u;
    for (int64_t i = (n4 << 2) + lane; i < n; i += PL_NT) stored[i] = 0;
    __syncthreads();
    for (int c = 0; c < ncls; c++) {
        const int g0 = grp_ptr[c], ng = grp_ptr[c + 1] - g0;
        for (int gi = lane; gi < ng; gi += PL_NT) {
            const uint32_t lo = gi == 0 ? 0u : (uint32_t)grp_end[g0 + gi - 1];
            const uint32_t hi = (uint32_t)grp_end[g0 + gi] - 1u;
            const uint32_t x = philox_draw(seed, it, P_PBWT_STORED, (uint32_t)gi, (uint32_t)c);
            const uint32_t l = urange(x, lo, hi);
            if (l < (uint32_t)Lhq) stored[(int64_t)c * Lhq + l] = 1;
        }
    }
    for (int i = lane; i < 2 * npt; i += PL_NT) {
        const uint32_t x = philox_draw(seed, it, P_PBWT_START, (uint32_t)(2 * targets[i >> 1] + (i & 1)), 0u);
        start[i] = (int32_t)urange(x, 0u, (uint32_t)(N - 1));
    }
}

// Places every target's Hvar rows, alpha checkpoints and phasing loci by its
// polymorphic-site count (in target order, as the host does), splits the
// targets into the forward-backward launch groups and records the iteration's
// set sizes and counters.  The graph holds the launches of the group keys
// want_a / want_b (-1: no group B) only; when this iteration's keys differ, or
// the placement exceeds the captured buffers, nothing is placed: every item
// gets Lp = W = 0, so the conditioning-set kernels write no rows, both groups
// are empty, and the overflow flag records why (GO_*).
__global__ __launch_bounds__(PL_NT) void fb_place_kernel(GraphCtl *__restrict__ ctl, CondItem *__restrict__ items,
                                                         FbItem *__restrict__ fb, PhItem *__restrict__ ph, int T,
                                                         int L, int C, int few, int latency, GraphCaps caps,
                                                         int want_a, int want_b, int32_t *__restrict__ ford,
                                                         int32_t *__restrict__ nstates, int n_iter) {
    const int lane = threadIdx.x;
    const uint32_t iter = ctl->iter;
    long long th = 0, ta = 0, tl = 0;
    int top = 0;
    for (int t = lane; t < T; t += PL_NT) {
        const CondItem c = items[t];
        th += (long long)c.Lp * c.W;
        ta += (long long)((c.Lp + C - 1) / C) * 64 * c.W + c.Lp;
        tl += c.Lp;
        top = max(top, fb_cap_class(c.W));
    }
    th = wave_sum(th);
    ta = wave_sum(ta);
    tl = wave_sum(tl);
    top = wave_max(top);
    int nA = 0, nB = 0, capB = 0;
    for (int t = lane; t < T; t += PL_NT) {
        const int cw = fb_cap_class(items[t].W);
        if (cw < top) {
            capB = max(capB, cw);
            nB++;
        } else {
            nA++;
        }
    }
    nA = wave_sum(nA);
    nB = wave_sum(nB);
    capB = wave_max(capB);
    const int key_a = graph_key(top, latency || nA <= few);
    const int key_b = nB > 0 ? graph_key(capB, latency || nB <= few) : -1;
    const int bad = (th > caps.hvar_words || ta > caps.alpha_floats || tl > caps.loci ? GO_CAPACITY : 0) |
                    (key_a != want_a || key_b != want_b ? GO_SHAPE : 0);
    if (bad) {
        for (int t = lane; t < T; t += PL_NT) {
            items[t].Lp = 0;
            items[t].W = 0;
        }
        if (lane == 0) {
            ctl->overflow |= bad;
            ctl->fb_n[0] = ctl->fb_n[1] = 0;
        }
        return;
    }
    long long ch = 0, ca = 0, cl = 0;
    int cA = 0, cB = 0;
    unsigned long long s0 = 0, s1 = 0, s2 = 0;
    for (int base = 0; base < T; base += PL_NT) {
        const int t = base + lane;
        const bool live = t < T;
        CondItem c{};
        if (live) c = items[t];
        const long long nh = live ? (long long)c.Lp * c.W : 0;
        const long long na = live ? (long long)((c.Lp + C - 1) / C) * 64 * c.W + c.Lp : 0;
        const long long nl = live ? c.Lp : 0;
        const int inA = live && fb_cap_class(c.W) == top, inB = live && !inA;
        long long zh, za, zl;
        int zA, zB;
        const long long oh = wave_excl(nh, zh), oa = wave_excl(na, za), ol = wave_excl(nl, zl);
        const int pA = wave_excl(inA, zA), pB = wave_excl(inB, zB);
        if (live) {
            items[t].hvar_off = ch + oh;
            fb[t].alpha_off = ca + oa;
            ph[t].locus_off = cl + ol;
            if (inA) ford[cA + pA] = t;
            else ford[T + cB + pB] = t;
            nstates[(size_t)t * n_iter + iter] = c.K;
            // two forward-backward passes per diploid target (run_execute's counters)
            s0 += (unsigned long long)(2LL * c.K * c.Lp);
            s1 += (unsigned long long)(2LL * ((long long)c.Lp * (16LL * c.W + 8LL * c.K + 16) + 8LL * L));
            s2 += (unsigned long long)(2LL * c.Lp);
        }
        ch += zh; ca += za; cl += zl; cA += zA; cB += zB;
    }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    s2 = wave_sum(s2);
    if (lane == 0) {
        ctl->fb_n[0] = cA;
        ctl->fb_n[1] = cB;
        ctl->fb_key[0] = key_a;
        ctl->fb_key[1] = key_b;
        ctl->stats[GS_FB_STATE_SITES] += s0;
        ctl->stats[GS_FB_BYTES] += s1;
        ctl->stats[GS_FB_POLY_SITES] += s2;
        ctl->stats[GS_FB_LAUNCHES] += cB > 0 ? 4 : 2;
    }
}

// Places the phasing messages by the segment and flat-het counts of the
// prepass and splits the targets into the phasing launch groups.  When the
// groups' keys differ from the captured ones (want_a, want_b) or the messages
// exceed the captured buffers, both groups are empty (no phasing kernel runs)
// and the overflow flag records why.
__global__ __launch_bounds__(PL_NT) void ph_place_kernel(GraphCtl *__restrict__ ctl, PhItem *__restrict__ ph,
                                                         const CondItem *__restrict__ items, int T, int few,
                                                         int latency, GraphCaps caps, int want_a, int want_b,
                                                         int32_t *__restrict__ pord) {
    const int lane = threadIdx.x;
    long long t0 = 0, t1 = 0, t2 = 0, t3 = 0, t4 = 0;
    int maxK = 1;
    for (int t = lane; t < T; t += PL_NT) {
        const int K = items[t].K;
        const long long msg = (long long)((K + 63) / 64) * 64 * 8;
        const PhItem p = ph[t];
        t0 += p.n_segs * msg;
        t1 += p.n_segs * 9LL;
        t2 += p.n_miss * msg;
        t3 += p.n_miss * 8LL;
        t4 += p.n_segs;
        maxK = max(maxK, K);
    }
    t0 = wave_sum(t0); t1 = wave_sum(t1); t2 = wave_sum(t2); t3 = wave_sum(t3); t4 = wave_sum(t4);
    const int top = ph_cap_states(wave_max(maxK));
    int nA = 0, nB = 0, maxKB = 1;
    for (int t = lane; t < T; t += PL_NT) {
        const int K = items[t].K;
        if (ph_cap_states(K) < top) {
            maxKB = max(maxKB, K);
            nB++;
        } else {
            nA++;
        }
    }
    nA = wave_sum(nA);
    nB = wave_sum(nB);
    maxKB = wave_max(maxKB);
    const int key_a = graph_key(top, latency || nA <= few);
    const int key_b = nB > 0 ? graph_key(ph_cap_states(maxKB), latency || nB <= few) : -1;
    const int bad = (t0 > caps.seg || t1 > caps.segsum || t2 > caps.imp || t3 > caps.impsum || t4 > caps.dip
                         ? GO_CAPACITY : 0) |
                    (key_a != want_a || key_b != want_b ? GO_SHAPE : 0);
    if (bad) {
        if (lane == 0) {
            ctl->overflow |= bad;
            ctl->ph_n[0] = ctl->ph_n[1] = 0;
        }
        return;
    }
    long long cs = 0, css = 0, ci = 0, cis = 0, cd = 0;
    int cA = 0, cB = 0;
    unsigned long long u0 = 0, u1 = 0;
    for (int base = 0; base < T; base += PL_NT) {
        const int t = base + lane;
        const bool live = t < T;
        int K = 0;
        PhItem p{};
        if (live) {
            K = items[t].K;
            p = ph[t];
        }
        const long long msg = (long long)((K + 63) / 64) * 64 * 8;
        const int inA = live && ph_cap_states(K) == top, inB = live && !inA;
        long long z0, z1, z2, z3, z4;
        int zA, zB;
        const long long o0 = wave_excl(live ? p.n_segs * msg : 0LL, z0);
        const long long o1 = wave_excl(live ? p.n_segs * 9LL : 0LL, z1);
        const long long o2 = wave_excl(live ? p.n_miss * msg : 0LL, z2);
        const long long o3 = wave_excl(live ? p.n_miss * 8LL : 0LL, z3);
        const long long o4 = wave_excl(live ? (long long)p.n_segs : 0LL, z4);
        const int pA = wave_excl(inA, zA), pB = wave_excl(inB, zB);
        if (live) {
            ph[t].msg = (int32_t)msg;
            ph[t].seg_off = cs + o0;
            ph[t].segsum_off = css + o1;
            ph[t].imp_off = ci + o2;
            ph[t].impsum_off = cis + o3;
            ph[t].dip_off = cd + o4;
            if (inA) pord[cA + pA] = t;
            else pord[T + cB + pB] = t;
            u0 += (unsigned long long)((long long)K * p.n_loci);
            u1 += (unsigned long long)(p.n_segs + p.n_miss);
        }
        cs += z0; css += z1; ci += z2; cis += z3; cd += z4; cA += zA; cB += zB;
    }
    u0 = wave_sum(u0);
    u1 = wave_sum(u1);
    if (lane == 0) {
        ctl->ph_n[0] = cA;
        ctl->ph_n[1] = cB;
        ctl->ph_key[0] = key_a;
        ctl->ph_key[1] = key_b;
        ctl->stats[GS_PH_STATE_SITES] += u0;
        ctl->stats[GS_PH_SEGMENTS] += u1;
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
                     bool latency, const GraphCaps &caps, const int want[2], int32_t *ford, int32_t *nstates,
                     int n_iter, hipStream_t s) {
    hipLaunchKernelGGL(fb_place_kernel, dim3(1), dim3(PL_NT), 0, s, ctl, items, fb, ph, T, L, C, few, (int)latency,
                       caps, want[0], want[1], ford, nstates, n_iter);
    GH_HIP(hipGetLastError());
}

void ph_place_launch(GraphCtl *ctl, PhItem *ph, const CondItem *items, int T, int few, bool latency,
                     const GraphCaps &caps, const int want[2], int32_t *pord, hipStream_t s) {
    hipLaunchKernelGGL(ph_place_kernel, dim3(1), dim3(PL_NT), 0, s, ctl, ph, items, T, few, (int)latency, caps,
                       want[0], want[1], pord);
    GH_HIP(hipGetLastError());
}

void graph_advance_launch(GraphCtl *ctl, hipStream_t s) {
    hipLaunchKernelGGL(graph_advance_kernel, dim3(1), dim3(PL_NT), 0, s, ctl);
    GH_HIP(hipGetLastError());
}

}  // namespace gh
