// gh_init.hip -- the INIT iteration's state selection on the device
// (haplotype_set::initRareTar + performSelection_RARE_INIT_GL,
//  phase/src/containers/haplotype_set.cpp:108-169, 775-833).
//
// One workgroup per target:
//   1. the rare sites where the target's GLs favour the rare allele over the
//      frequency prior (initRareTar), compacted in site order by block scans;
//   2. with more such sites than 0.8 Kinit, the sites ordered by minor-allele
//      count (a stable counting sort: (count, site) ascending, as std::sort of
//      the pairs) and 5 carriers sampled from each until 0.8 Kinit states are
//      set; else floor(0.8 Kinit / sites) carriers sampled from every site;
//   3. the set filled to Kinit with a sample of all N haplotypes;
//   4. the set written in ascending order (std::set's order).
// Sampling is selection sampling (Knuth's Algorithm S) with one counter draw
// per examined candidate, keyed (seed, target, draw number): the sequence of
// decisions is sequential and runs on one thread; the draws of the final fill
// are computed by the whole workgroup ahead of it.  The set is a bitmap over
// the panel haplotypes in LDS.
#include <algorithm>
#include <string>

#include "gh_device.hpp"
#include "gh_init.hpp"

namespace gh {

constexpr int INIT_NT = 256;

__device__ __forceinline__ float init_draw(uint32_t seed, uint32_t tgt, uint32_t ctr) {
    return u01(philox_draw(seed, 0u, P_INIT_SEL, tgt, ctr));
}

// Exclusive prefix of v over the workgroup; total gets the sum.
__device__ __forceinline__ int init_scan(int v, int &total, int *ws) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        const int y = __shfl_up(x, o, 64);
        if (lane >= o) x += y;
    }
    if (lane == 63) ws[wid] = x;
    __syncthreads();
    int before = 0;
    total = 0;
#pragma unroll
    for (int w = 0; w < INIT_NT / 64; w++) {
        before += w < wid ? ws[w] : 0;
        total += ws[w];
    }
    __syncthreads();
    return before + x - v;
}

__global__ __launch_bounds__(INIT_NT) void init_select_kernel(InitArgs a) {
#pragma clang fp contract(off)
    extern __shared__ uint64_t bm[];                 // [Wn] the state set
    __shared__ int ws[INIT_NT / 64];
    __shared__ int s_cnt[INIT_MAC_MAX + 1];          // minor-allele-count buckets
    __shared__ float s_u[INIT_NT];                   // draws of the final fill
    __shared__ int s_size, s_need, s_ctr, s_ns;
    const int tid = threadIdx.x;
    const int t = a.targets ? a.targets[blockIdx.x] : (int)blockIdx.x;
    const int L = a.L, N = a.N, Wn = (N + 63) / 64;
    const int kinit = a.kinit, k8 = (int)floorf((float)kinit * 0.8f);
    const uint8_t *gl = a.gl + (size_t)t * L * 3, *flat = a.flat + (size_t)t * L;
    const bool hap = a.ploidy && a.ploidy[t] == 1;
    int32_t *sind = a.sind + (size_t)t * L, *sorted = a.sorted + (size_t)t * L;
    for (int w = tid; w < Wn; w += INIT_NT) bm[w] = 0ull;
    // 1. rare sites with evidence for the rare allele (initRareTar, :108-169)
    int ns = 0;
    for (int v0 = 0; v0 < L; v0 += INIT_NT) {
        const int v = v0 + tid;
        int take = 0;
        if (v < L && !flat[v]) {
            const uint8_t f = a.flags[v];
            const int mg = 2 * ((f & SF_MAJOR) ? 1 : 0);
            if (!(f & SF_COMMON) && !(f & SF_LQ)) {
                // a haploid target's GL array holds two values per site; the
                // reference reads it at 3v + 0..2, so its third value is the
                // next site's first (0 past the last site)
                uint8_t g[3] = {gl[3 * v], gl[3 * v + 1], gl[3 * v + 2]};
                if (hap) g[2] = v + 1 < L ? gl[3 * (v + 1)] : 0;
                if (g[1] <= g[mg] || g[2 - mg] <= g[mg]) {
                    float x0 = a.unph[g[0]], x1 = a.unph[g[1]], x2 = hap ? 0.0f : a.unph[g[2]];
                    float s = x0 + x1 + x2;
                    x0 /= s; x1 /= s; x2 /= s;
                    const float af = (float)((double)a.calt[v] * 1.0 / (double)N);
                    x0 *= (1.0f - af) * (1.0f - af);
                    x1 *= 2.0f * af * (1.0f - af);
                    x2 *= af * af;
                    s = x0 + x1 + x2;
                    x0 /= s; x1 /= s; x2 /= s;
                    const float xm = mg ? x2 : x0, xo = mg ? x0 : x2;
                    take = x1 + xo > xm;
                }
            }
        }
        int tot;
        const int p = init_scan(take, tot, ws);
        if (take) sind[ns + p] = v;
        ns += tot;
    }
    if (tid == 0) { s_size = 0; s_ctr = 0; s_ns = ns; }
    __syncthreads();
    // 2. carriers of those sites (performSelection_RARE_INIT_GL, :775-833)
    if (ns > k8) {
        for (int b = tid; b <= a.max_mac; b += INIT_NT) s_cnt[b] = 0;
        __syncthreads();
        for (int i = tid; i < ns; i += INIT_NT) {
            const int v = sind[i];
            atomicAdd(&s_cnt[min(a.cref[v], a.calt[v])], 1);
        }
        __syncthreads();
        if (tid == 0) {
            int run = 0;
            for (int b = 0; b <= a.max_mac; b++) { const int c = s_cnt[b]; s_cnt[b] = run; run += c; }
            for (int i = 0; i < ns; i++) {   // stable: sites keep their order within a count
                const int v = sind[i];
                sorted[s_cnt[min(a.cref[v], a.calt[v])]++] = v;
            }
            uint32_t ctr = 0;
            int size = 0;
            for (int r0 = 0; r0 < ns && size < k8; r0++) {
                const int v = sorted[r0];
                const int c0 = a.rare_ptr[v], nc = a.rare_ptr[v + 1] - c0;
                int need = min(5, nc);
                for (int i = 0; i < nc && need > 0; i++) {
                    const float u = init_draw(a.seed, (uint32_t)t, ctr++);
                    if (u * (float)(nc - i) < (float)need) {
                        need--;
                        const int id = a.rare_idx[c0 + i];
                        if (size < k8 && !((bm[id >> 6] >> (id & 63)) & 1ull)) {
                            bm[id >> 6] |= 1ull << (id & 63);
                            size++;
                        }
                    }
                }
            }
            s_size = size;
            s_ctr = (int)ctr;
        }
    } else if (ns > 0) {
        if (tid == 0) {
            const int size_max = max((int)floor((double)k8 * 1.0 / (double)ns), 1);
            uint32_t ctr = 0;
            int size = 0;
            for (int r0 = 0; r0 < ns; r0++) {   // every site is sampled, also once the set is full
                const int v = sind[r0];
                const int c0 = a.rare_ptr[v], nc = a.rare_ptr[v + 1] - c0;
                int need = min(size_max, nc);
                for (int i = 0; i < nc && need > 0; i++) {
                    const float u = init_draw(a.seed, (uint32_t)t, ctr++);
                    if (u * (float)(nc - i) < (float)need) {
                        need--;
                        const int id = a.rare_idx[c0 + i];
                        if (size < k8 && !((bm[id >> 6] >> (id & 63)) & 1ull)) {
                            bm[id >> 6] |= 1ull << (id & 63);
                            size++;
                        }
                    }
                }
            }
            s_size = size;
            s_ctr = (int)ctr;
        }
    }
    __syncthreads();
    // 3. fill to Kinit with a sample of all N haplotypes (population order)
    if (s_size < kinit) {
        const uint32_t ctr0 = (uint32_t)s_ctr;
        if (tid == 0) s_need = min(kinit, N);
        __syncthreads();
        for (int base = 0; base < N; base += INIT_NT) {
            if (s_need <= 0) break;   // uniform: read after the barrier that ends the previous chunk
            const int c = base + tid;
            s_u[tid] = c < N ? init_draw(a.seed, (uint32_t)t, ctr0 + (uint32_t)c) : 0.0f;
            __syncthreads();
            if (tid == 0) {
                int need = s_need, size = s_size;
                for (int j = 0; j < INIT_NT && base + j < N && need > 0; j++) {
                    const int cand = base + j;
                    if (s_u[j] * (float)(N - cand) < (float)need) {
                        need--;
                        if (size < kinit && !((bm[cand >> 6] >> (cand & 63)) & 1ull)) {
                            bm[cand >> 6] |= 1ull << (cand & 63);
                            size++;
                        }
                    }
                }
                s_need = need;
                s_size = size;
            }
            __syncthreads();
        }
    }
    __syncthreads();
    // 4. the set in ascending order
    int32_t *out = a.states + (size_t)t * a.kstride;
    int done = 0;
    for (int w0 = 0; w0 < Wn; w0 += INIT_NT) {
        const int w = w0 + tid;
        const uint64_t x = w < Wn ? bm[w] : 0ull;
        int tot;
        int p = done + init_scan(__popcll(x), tot, ws);
        uint64_t y = x;
        while (y) {
            const int b = __builtin_ctzll(y);
            out[p++] = 64 * w + b;
            y &= y - 1;
        }
        done += tot;
    }
    if (tid == 0) {
        a.K[t] = done;
        if (a.items) {
            a.items[t].K = done;
            a.items[t].W = (done + 63) / 64;
        }
    }
}

// Full-panel targets (Kpbwt >= N, conditioning_set.cpp:105-113): states 0..N-1.
__global__ void full_panel_kernel(const int32_t *__restrict__ targets, int n, int N, int kstride,
                                  int32_t *__restrict__ states, int32_t *__restrict__ K, CondItem *__restrict__ items) {
    const int i = blockIdx.y;
    if (i >= n) return;
    const int t = targets[i];
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < N; j += gridDim.x * blockDim.x)
        states[(size_t)t * kstride + j] = j;
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        K[t] = N;
        if (items) {
            items[t].K = N;
            items[t].W = (N + 63) / 64;
        }
    }
}

size_t init_lds_bytes(int N) { return sizeof(uint64_t) * (size_t)((N + 63) / 64); }

void init_select_launch(const InitArgs &a, int n, hipStream_t s) {
    if (n <= 0) return;
    GH_CHECK(a.max_mac <= INIT_MAC_MAX, GH_ERR_ARG,
             "INIT selection: rare sites with a minor-allele count above " + std::to_string(INIT_MAC_MAX));
    const size_t lds = init_lds_bytes(a.N);
    GH_CHECK(lds <= INIT_LDS_MAX, GH_ERR_ARG, "INIT selection: panel too large for the LDS state set");
    static const bool attr = [] {
        GH_HIP(hipFuncSetAttribute((const void *)init_select_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)INIT_LDS_MAX));
        return true;
    }();
    (void)attr;
    hipLaunchKernelGGL(init_select_kernel, dim3(n), dim3(INIT_NT), lds, s, a);
    GH_HIP(hipGetLastError());
}

void init_states_batch(gh_panel *P, int T, const uint8_t *gl_h, const uint8_t *flat_h, int kinit, uint32_t seed,
                       int kstride, int32_t *states_h, int32_t *K_h) {
    GH_CHECK(P && T > 0 && gl_h && flat_h && states_h && K_h, GH_ERR_ARG, "gh_init_states: bad arguments");
    const int L = P->L, N = P->N;
    kinit = std::min(kinit, N);
    GH_CHECK(kinit > 0 && kstride >= kinit, GH_ERR_ARG, "gh_init_states: need 0 < min(kinit, N) <= kstride");
    hipStream_t s = P->stream;
    float unph[256];
    unphred_table(unph);
    DBuf gl, flat, sind, sorted, st, K, up;
    InitArgs a;
    a.targets = nullptr;
    a.gl = dalloc_copy(gl, gl_h, (size_t)T * L * 3, s);
    a.flat = dalloc_copy(flat, flat_h, (size_t)T * L, s);
    a.ploidy = nullptr;
    a.flags = P->d_flags.as<uint8_t>();
    a.cref = P->d_cref.as<uint32_t>(); a.calt = P->d_calt.as<uint32_t>();
    a.rare_ptr = P->d_rare_ptr.as<int32_t>(); a.rare_idx = P->d_rare_idx.as<int32_t>();
    a.unph = dalloc_copy(up, unph, 256, s);
    a.L = L; a.N = N; a.kinit = kinit; a.kstride = kstride; a.max_mac = P->max_rare_mac; a.seed = seed;
    a.sind = (int32_t *)sind.reserve(sizeof(int32_t) * (size_t)T * L);
    a.sorted = (int32_t *)sorted.reserve(sizeof(int32_t) * (size_t)T * L);
    a.states = (int32_t *)st.reserve(sizeof(int32_t) * (size_t)T * kstride);
    a.K = (int32_t *)K.reserve(sizeof(int32_t) * (size_t)T);
    a.items = nullptr;
    init_select_launch(a, T, s);
    GH_HIP(hipMemcpyAsync(states_h, a.states, sizeof(int32_t) * (size_t)T * kstride, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(K_h, a.K, sizeof(int32_t) * (size_t)T, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
}

void full_panel_launch(const int32_t *targets, int n, int N, int kstride, int32_t *states, int32_t *K,
                       CondItem *items, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(full_panel_kernel, dim3((N + 255) / 256 < 64 ? (N + 255) / 256 : 64, n), dim3(256), 0, s,
                       targets, n, N, kstride, states, K, items);
    GH_HIP(hipGetLastError());
}

}  // namespace gh
