// gh_condset.hip -- conditioning sets on the GPU
// (conditioning_set::compactSelection + updateTransitions,
//  phase/src/containers/conditioning_set.cpp:75-196).
//
// Per item (one target's state set, sorted state ids):
//   cond_index_kernel  : site classes (common / rare carried by a state / mono),
//                        polymorphic list, site -> row map, float transitions
//   cond_rare_kernel   : rare rows (major fill, carriers flipped with atomics),
//                        once the item's rows have their place (hvar_off)
//   cond_common_kernel : common rows of Hvar gathered from the haplotype-major
//                        panel with a 64x64 ballot transpose per wave
// Hvar row layout: ceil(K/64) uint64 words per polymorphic site, bit j of
// word w is state 64w+j.
#include "gh_device.hpp"
#include "gh_internal.hpp"

namespace gh {

template <int NT>
__device__ __forceinline__ int block_excl_scan(int v, int &total) {
    __shared__ int wsum[NT / 64];
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    int x = v;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        int y = __shfl_up(x, o);
        if (lane >= o) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (wid == 0) {
        int w = lane < NT / 64 ? wsum[lane] : 0;
#pragma unroll
        for (int o = 1; o < NT / 64; o <<= 1) {
            int y = __shfl_up(w, o);
            if (lane >= o) w += y;
        }
        if (lane < NT / 64) wsum[lane] = w;
    }
    __syncthreads();
    total = wsum[NT / 64 - 1];
    const int before = wid ? wsum[wid - 1] : 0;
    __syncthreads();
    return before + x - v;
}

// 256 threads (four waves) per item: a workgroup that needs a whole CU's wave
// slots waits behind the other runs' long-lived workgroups when several runs
// share the GPU
constexpr int COND_NT = 256;

__global__ __launch_bounds__(COND_NT) void cond_index_kernel(CondItem *__restrict__ items,
                                                             const int32_t *__restrict__ states, int L,
                                                             const uint8_t *__restrict__ flags,
                                                             const int32_t *__restrict__ shap_ptr,
                                                             const int32_t *__restrict__ shap_idx,
                                                             const double *__restrict__ cm, double nrho,
                                                             int32_t *__restrict__ poly_all,
                                                             int32_t *__restrict__ pos_all, float *__restrict__ t_all,
                                                             uint8_t *__restrict__ mark_all) {
    const CondItem it = items[blockIdx.x];
    const int tid = threadIdx.x;
    uint8_t *mark = mark_all + it.site_off;
    int32_t *poly = poly_all + it.site_off;
    int32_t *pos = pos_all + it.site_off;
    float *t = t_all + it.site_off;
    const int32_t *st = states + it.state_off;
    const int K = it.K;

    for (int s = tid; s < L; s += COND_NT) mark[s] = flags[s] & SF_COMMON;
    __syncthreads();
    for (int k = tid; k < K; k += COND_NT) {
        const int h = st[k];
        for (int j = shap_ptr[h]; j < shap_ptr[h + 1]; j++) mark[shap_idx[j]] = 1;
    }
    __syncthreads();
    const int C = (L + COND_NT - 1) / COND_NT;
    const int s0 = min(L, tid * C), s1 = min(L, s0 + C);
    int cnt = 0;
    for (int s = s0; s < s1; s++) cnt += mark[s];
    int total;
    int p = block_excl_scan<COND_NT>(cnt, total);
    for (int s = s0; s < s1; s++) {
        if (mark[s]) { pos[s] = p; poly[p] = s; p++; }
        else pos[s] = -(s - p) - 1;   // -(mono index) - 1
    }
    if (tid == 0) items[blockIdx.x].Lp = total;
    __syncthreads();
    const int Lp = total;
    const double one_l = 1.0 - 1e-7;
    for (int q = tid + 1; q < Lp; q += COND_NT) {
        double x = -expm1(nrho * (cm[poly[q]] - cm[poly[q - 1]]));
        x = x < 1e-7 ? 1e-7 : (x > one_l ? one_l : x);
        t[q - 1] = (float)x;
    }
}

// Rare rows of Hvar (major allele everywhere, the states' carriers flipped);
// runs once the item's row offset (hvar_off, room for Lp rows) is set.
__global__ __launch_bounds__(COND_NT) void cond_rare_kernel(const CondItem *__restrict__ items,
                                                            const int32_t *__restrict__ states,
                                                            const uint8_t *__restrict__ flags,
                                                            const int32_t *__restrict__ shap_ptr,
                                                            const int32_t *__restrict__ shap_idx,
                                                            const int32_t *__restrict__ poly_all,
                                                            const int32_t *__restrict__ pos_all,
                                                            uint64_t *__restrict__ hvar_all) {
    const CondItem it = items[blockIdx.x];
    const int tid = threadIdx.x;
    const int32_t *poly = poly_all + it.site_off;
    const int32_t *pos = pos_all + it.site_off;
    uint64_t *hv = hvar_all + it.hvar_off;
    const int32_t *st = states + it.state_off;
    const int K = it.K, W = it.W, Lp = it.Lp;
    for (int64_t e = tid; e < (int64_t)Lp * W; e += COND_NT) {
        const int q = (int)(e / W);
        const uint8_t f = flags[poly[q]];
        if (!(f & SF_COMMON)) hv[e] = (f & SF_MAJOR) ? ~0ull : 0ull;
    }
    __syncthreads();
    for (int k = tid; k < K; k += COND_NT) {
        const int h = st[k];
        const uint64_t bit = 1ull << (k & 63);
        for (int j = shap_ptr[h]; j < shap_ptr[h + 1]; j++) {
            const int q = pos[shap_idx[j]];
            atomicXor((unsigned long long *)&hv[(size_t)q * W + (k >> 6)], (unsigned long long)bit);
        }
    }
}

__global__ __launch_bounds__(64) void cond_common_kernel(const CondItem *__restrict__ items,
                                                        const int32_t *__restrict__ states,
                                                        const uint64_t *__restrict__ hm, int Wc, int Lc,
                                                        const int32_t *__restrict__ c2t,
                                                        const int32_t *__restrict__ pos_all,
                                                        uint64_t *__restrict__ hvar_all) {
    const int rb = blockIdx.x, w = blockIdx.y;
    const CondItem it = items[blockIdx.z];
    if (w >= it.W) return;
    const int lane = threadIdx.x;
    const int k = 64 * w + lane;
    uint64_t x = 0;
    if (k < it.K) x = hm[(size_t)states[it.state_off + k] * Wc + rb];
    uint64_t mine = 0;
#pragma unroll 8
    for (int s = 0; s < 64; s++) {
        uint64_t m = __ballot((int)((x >> s) & 1ull));
        if (lane == s) mine = m;
    }
    const int lc = rb * 64 + lane;
    if (lc < Lc) {
        const int q = pos_all[it.site_off + c2t[lc]];
        hvar_all[it.hvar_off + (size_t)q * it.W + w] = mine;
    }
}

void condset_index(gh_panel *P, int n, CondItem *d_items, const int32_t *d_states, int32_t *d_poly, int32_t *d_pos,
                   float *d_t, uint8_t *d_mark, int ne, hipStream_t s) {
    if (n <= 0) return;
    const double nrho = (-0.04 * ne) / P->N;   // conditioning_set.cpp:34 (no state list)
    hipLaunchKernelGGL(cond_index_kernel, dim3(n), dim3(COND_NT), 0, s, d_items, d_states, P->L,
                       P->d_flags.as<uint8_t>(), P->d_shap_ptr.as<int32_t>(), P->d_shap_idx.as<int32_t>(),
                       P->d_cm.as<double>(), nrho, d_poly, d_pos, d_t, d_mark);
    GH_HIP(hipGetLastError());
}

void condset_fill(gh_panel *P, int n, const CondItem *d_items, const int32_t *d_states, const int32_t *d_poly,
                  const int32_t *d_pos, uint64_t *d_hvar, int maxW, hipStream_t s) {
    if (n <= 0) return;
    hipLaunchKernelGGL(cond_rare_kernel, dim3(n), dim3(COND_NT), 0, s, d_items, d_states, P->d_flags.as<uint8_t>(),
                       P->d_shap_ptr.as<int32_t>(), P->d_shap_idx.as<int32_t>(), d_poly, d_pos, d_hvar);
    GH_HIP(hipGetLastError());
    if (P->Lc > 0) {
        hipLaunchKernelGGL(cond_common_kernel, dim3(P->Wc, maxW, n), dim3(64), 0, s, d_items, d_states,
                           P->d_hm.as<uint64_t>(), P->Wc, P->Lc, P->d_c2t.as<int32_t>(), d_pos, d_hvar);
        GH_HIP(hipGetLastError());
    }
}

void condset_build(gh_panel *P, int n, CondItem *d_items, const CondItem *h_items, const int32_t *d_states,
                   int32_t *d_poly, int32_t *d_pos, float *d_t, uint64_t *d_hvar, uint8_t *d_mark, int ne,
                   int maxW, hipStream_t s) {
    (void)h_items;
    condset_index(P, n, d_items, d_states, d_poly, d_pos, d_t, d_mark, ne, s);
    condset_fill(P, n, d_items, d_states, d_poly, d_pos, d_hvar, maxW, s);
}

}  // namespace gh
