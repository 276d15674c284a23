// gh_mlp.hip -- the learned state selector's network: 32 -> 128 -> 128 -> 1 MLP
// with ReLU.  BF16 operands on the matrix cores (v_mfma_f32_16x16x16_bf16),
// FP32 accumulation, FP32 master weights, Adam.  No counterpart in the reference.
//
// A workgroup is 4 wavefronts working on tiles of 64 rows (16 per wave).  Each
// matrix product is built from 16x16 output tiles, each accumulated over K in
// steps of 16 by one wave from LDS-resident BF16 operands (mma_tile).  Operands
// whose K index is contiguous in LDS are fetched 4 halves at a time; LDS rows
// are padded by 8 halves so the 16 rows a wave touches fall in different banks.
//
// Layer 3 (128 -> 1) is a dot product in FP32 (DPP row reduction), and so are
// the bias and W3 gradients.  Training: mlp_grad_kernel is persistent over the
// batch's row tiles; every workgroup keeps its weight-gradient tiles in MFMA
// accumulators and writes them once to its own slice of a partial buffer, and
// mlp_reduce_kernel sums the slices in a fixed order, so gradients are
// deterministic.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gh_device.hpp"
#include "gh_internal.hpp"
#include "gh_mlp.hpp"

namespace gh {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

// float -> bf16, round to nearest even (inputs are finite).
__device__ __forceinline__ uint16_t f2bf(float x) {
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t b) { return __uint_as_float((uint32_t)b << 16); }

constexpr int MLP_NT = 256;        // 4 waves
constexpr int TR = 64;             // rows per tile
constexpr int LDX = MLP_IN + 8;    // padded row strides (halves)
constexpr int LDH = MLP_H + 8;
constexpr int GRAD_GRID = 256;     // workgroups of the gradient kernel (partial slices)

// D += A[16 x K] * B[K x 16] from LDS.  A(m, k) = A[m*sam + k*sak], B(k, n) =
// B[k*sbk + n*sbn].  v_mfma_f32_16x16x16_bf16 operand layout: lane l supplies
// A row l%16 and B column l%16 at k = 4*(l/16) .. +3; it receives D column l%16,
// rows 4*(l/16) .. +3.
template <bool ACONT, bool BCONT>
__device__ __forceinline__ f32x4 mma_tile(f32x4 acc, const uint16_t *A, int sam, int sak, const uint16_t *B, int sbk,
                                          int sbn, int K) {
    const int lane = threadIdx.x & 63, r = lane & 15, kg = lane >> 4;
    for (int k0 = 0; k0 < K; k0 += 16) {
        bf16x4 a, b;
        if (ACONT) {
            a = *reinterpret_cast<const bf16x4 *>(A + r * sam + k0 + 4 * kg);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) a[i] = (short)A[r * sam + (k0 + 4 * kg + i) * sak];
        }
        if (BCONT) {
            b = *reinterpret_cast<const bf16x4 *>(B + r * sbn + k0 + 4 * kg);
        } else {
#pragma unroll
            for (int i = 0; i < 4; i++) b[i] = (short)B[(k0 + 4 * kg + i) * sbk + r * sbn];
        }
        acc = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, acc, 0, 0, 0);
    }
    return acc;
}

// Sum over the 16 lanes of a DPP row (lanes with equal l/16).
__device__ __forceinline__ float row16_sum(float v) {
    v += dppf<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dppf<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dppf<0x141>(v);   // row_half_mirror
    v += dppf<0x140>(v);   // row_mirror
    return v;
}

struct FwdLds {
    __align__(16) uint16_t X[TR * LDX];         // bf16 inputs of the tile
    __align__(16) uint16_t H1[TR * LDH];        // bf16 relu(layer 1)
    __align__(16) uint16_t W1T[MLP_H * LDX];    // W1T[n][i] = W1[i][n]
    __align__(16) uint16_t W2T[MLP_H * LDH];    // W2T[n][c] = W2[c][n]
    float W3[MLP_H], b1[MLP_H], b2[MLP_H];
    float y[TR];
    float b3;
};
struct GradLds {
    FwdLds f;
    __align__(16) uint16_t D2[TR * LDH];        // bf16 dL/d(pre-activation 2)
    __align__(16) uint16_t D1[TR * LDH];        // bf16 dL/d(pre-activation 1)
    __align__(16) uint16_t W2[MLP_H * LDH];     // W2[j][c] row-major (backward operand)
    float g[TR];                                // dL/dy per row
    float vpart[3][4][MLP_H];                   // per-wave gb1, gb2, gW3 partials
};

__device__ void load_weights(FwdLds &S, const float *w, const uint16_t *wb, uint16_t *W2rm) {
    const uint16_t *W1b = wb, *W2b = wb + MLP_IN * MLP_H;
    for (int e = threadIdx.x; e < MLP_IN * MLP_H; e += MLP_NT) S.W1T[(e % MLP_H) * LDX + e / MLP_H] = W1b[e];
    for (int e = threadIdx.x; e < MLP_H * MLP_H; e += MLP_NT) {
        const uint16_t v = W2b[e];
        S.W2T[(e % MLP_H) * LDH + e / MLP_H] = v;
        if (W2rm) W2rm[(e / MLP_H) * LDH + e % MLP_H] = v;
    }
    for (int e = threadIdx.x; e < MLP_H; e += MLP_NT) {
        S.W3[e] = w[MLP_OFF_W3 + e];
        S.b1[e] = w[MLP_OFF_B1 + e];
        S.b2[e] = w[MLP_OFF_B2 + e];
    }
    if (threadIdx.x == 0) S.b3 = w[MLP_OFF_B3];
}

// Forward pass of one tile whose rows are already in S.X (all threads call it).
// Leaves relu(layer 1) in S.H1 (bf16), relu(layer 2) of the wave's 16 rows in
// h2 (FP32, MFMA D layout: h2[nt][i] = row 16w+4(l/16)+i, column 16nt+l%16)
// and y in S.y.
__device__ __forceinline__ void forward_tile(FwdLds &S, f32x4 h2[8]) {
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, col = lane & 15, kg = lane >> 4;
#pragma unroll
    for (int nt = 0; nt < 8; nt++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile<true, true>(acc, S.X + 16 * w * LDX, LDX, 1, S.W1T + 16 * nt * LDX, 1, LDX, MLP_IN);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int c = 16 * nt + col;
            S.H1[(16 * w + 4 * kg + i) * LDH + c] = f2bf(fmaxf(acc[i] + S.b1[c], 0.0f));
        }
    }
    __syncthreads();
    float part[4] = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int nt = 0; nt < 8; nt++) {
        f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
        acc = mma_tile<true, true>(acc, S.H1 + 16 * w * LDH, LDH, 1, S.W2T + 16 * nt * LDH, 1, LDH, MLP_H);
        const int c = 16 * nt + col;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            acc[i] = fmaxf(acc[i] + S.b2[c], 0.0f);
            part[i] = fmaf(acc[i], S.W3[c], part[i]);
        }
        h2[nt] = acc;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const float v = row16_sum(part[i]);
        if (col == 0) S.y[16 * w + 4 * kg + i] = v + S.b3;
    }
}

__global__ __launch_bounds__(MLP_NT) void mlp_forward_kernel(const float *__restrict__ w,
                                                            const uint16_t *__restrict__ wb,
                                                            const float *__restrict__ X, int64_t n,
                                                            float *__restrict__ y) {
    extern __shared__ __align__(16) unsigned char smem[];
    FwdLds &S = *reinterpret_cast<FwdLds *>(smem);
    load_weights(S, w, wb, nullptr);
    for (int64_t r0 = (int64_t)blockIdx.x * TR; r0 < n; r0 += (int64_t)gridDim.x * TR) {
        for (int e = threadIdx.x; e < TR * MLP_IN; e += MLP_NT) {
            const int r = e / MLP_IN;
            S.X[r * LDX + e % MLP_IN] = f2bf(r0 + r < n ? X[(r0 + r) * MLP_IN + e % MLP_IN] : 0.0f);
        }
        __syncthreads();
        f32x4 h2[8];
        forward_tile(S, h2);
        __syncthreads();
        if (threadIdx.x < TR && r0 + threadIdx.x < n) y[r0 + threadIdx.x] = S.y[threadIdx.x];
    }
}

__global__ __launch_bounds__(MLP_NT) void mlp_grad_kernel(const float *__restrict__ wp,
                                                         const uint16_t *__restrict__ wb, MlpBatch B,
                                                         float inv_wsum, float *__restrict__ part) {
    extern __shared__ __align__(16) unsigned char smem[];
    GradLds &S = *reinterpret_cast<GradLds *>(smem);
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6, col = lane & 15, kg = lane >> 4;
    load_weights(S.f, wp, wb, S.W2);
    // weight-gradient tiles owned by this wave: gW2 row tiles 2w, 2w+1 with all 8
    // column tiles; gW1 row tile w/2 with column tiles 4(w%2) .. +3
    f32x4 gW2[2][8], gW1[4];
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 8; b++) gW2[a][b] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
    for (int b = 0; b < 4; b++) gW1[b] = (f32x4){0.0f, 0.0f, 0.0f, 0.0f};
    float vb1[8], vb2[8], vw3[8];
#pragma unroll
    for (int nt = 0; nt < 8; nt++) vb1[nt] = vb2[nt] = vw3[nt] = 0.0f;
    float gb3 = 0.0f, loss = 0.0f;   // kept by thread 0
    for (int r0 = blockIdx.x * TR; r0 < B.n; r0 += gridDim.x * TR) {
        for (int e = tid; e < TR * MLP_IN; e += MLP_NT) {
            const int r = e / MLP_IN;
            float v = 0.0f;
            if (r0 + r < B.n) {
                const int64_t row = (int64_t)((B.a * (uint64_t)(B.i0 + r0 + r) + B.b) % (uint64_t)B.n_total);
                v = B.X[row * MLP_IN + e % MLP_IN];
            }
            S.f.X[r * LDX + e % MLP_IN] = f2bf(v);
        }
        __syncthreads();
        f32x4 h2[8];
        forward_tile(S.f, h2);
        __syncthreads();
        if (tid < TR) {
            float g = 0.0f, l2 = 0.0f;
            if (r0 + tid < B.n) {
                const int64_t row = (int64_t)((B.a * (uint64_t)(B.i0 + r0 + tid) + B.b) % (uint64_t)B.n_total);
                const float wt = B.w ? B.w[row] : 1.0f;
                const float d = S.f.y[tid] - B.t[row];
                g = 2.0f * wt * d * inv_wsum;
                l2 = wt * d * d * inv_wsum;
            }
            S.g[tid] = g;
            S.f.y[tid] = l2;
        }
        __syncthreads();
        if (tid == 0)
            for (int r = 0; r < TR; r++) { gb3 += S.g[r]; loss += S.f.y[r]; }
        // layer 3 and relu 2: D2 = g * W3 * [h2 > 0]
#pragma unroll
        for (int nt = 0; nt < 8; nt++) {
            const int c = 16 * nt + col;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = 16 * w + 4 * kg + i;
                const float g = S.g[row];
                const float d = h2[nt][i] > 0.0f ? g * S.f.W3[c] : 0.0f;
                S.D2[row * LDH + c] = f2bf(d);
                vw3[nt] = fmaf(h2[nt][i], g, vw3[nt]);
                vb2[nt] += d;
            }
        }
        __syncthreads();
        // gW2[k][n] += sum_r H1[r][k] D2[r][n]
#pragma unroll
        for (int a = 0; a < 2; a++)
#pragma unroll
            for (int nt = 0; nt < 8; nt++)
                gW2[a][nt] = mma_tile<false, false>(gW2[a][nt], S.f.H1 + 16 * (2 * w + a), 1, LDH, S.D2 + 16 * nt,
                                                    LDH, 1, TR);
        // D1 = (D2 W2^T) * [h1 > 0]   (B(k=c, n=j) = W2[j][c], contiguous in c)
#pragma unroll
        for (int nt = 0; nt < 8; nt++) {
            f32x4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            acc = mma_tile<true, true>(acc, S.D2 + 16 * w * LDH, LDH, 1, S.W2 + 16 * nt * LDH, 1, LDH, MLP_H);
            const int c = 16 * nt + col;
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const int row = 16 * w + 4 * kg + i;
                const float d = bf2f(S.f.H1[row * LDH + c]) > 0.0f ? acc[i] : 0.0f;
                S.D1[row * LDH + c] = f2bf(d);
                vb1[nt] += d;
            }
        }
        __syncthreads();
        // gW1[i][k] += sum_r X[r][i] D1[r][k]
#pragma unroll
        for (int b = 0; b < 4; b++)
            gW1[b] = mma_tile<false, false>(gW1[b], S.f.X + 16 * (w >> 1), 1, LDX, S.D1 + 16 * (4 * (w & 1) + b),
                                            LDH, 1, TR);
        __syncthreads();
    }
    // this workgroup's slice of the partial buffer
    float *P = part + (size_t)blockIdx.x * (MLP_NPARAM + 1);
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
        for (int nt = 0; nt < 8; nt++)
#pragma unroll
            for (int i = 0; i < 4; i++)
                P[MLP_OFF_W2 + (16 * (2 * w + a) + 4 * kg + i) * MLP_H + 16 * nt + col] = gW2[a][nt][i];
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
        for (int i = 0; i < 4; i++)
            P[MLP_OFF_W1 + (16 * (w >> 1) + 4 * kg + i) * MLP_H + 16 * (4 * (w & 1) + b) + col] = gW1[b][i];
    // vector gradients: lanes col, col+16, col+32, col+48 hold the same column
#pragma unroll
    for (int nt = 0; nt < 8; nt++) {
        float x1 = vb1[nt], x2 = vb2[nt], x3 = vw3[nt];
        x1 += __shfl_xor(x1, 16); x1 += __shfl_xor(x1, 32);
        x2 += __shfl_xor(x2, 16); x2 += __shfl_xor(x2, 32);
        x3 += __shfl_xor(x3, 16); x3 += __shfl_xor(x3, 32);
        if (kg == 0) {
            S.vpart[0][w][16 * nt + col] = x1;
            S.vpart[1][w][16 * nt + col] = x2;
            S.vpart[2][w][16 * nt + col] = x3;
        }
    }
    __syncthreads();
    if (tid < MLP_H) {
        const int c = tid;
        P[MLP_OFF_B1 + c] = ((S.vpart[0][0][c] + S.vpart[0][1][c]) + S.vpart[0][2][c]) + S.vpart[0][3][c];
        P[MLP_OFF_B2 + c] = ((S.vpart[1][0][c] + S.vpart[1][1][c]) + S.vpart[1][2][c]) + S.vpart[1][3][c];
        P[MLP_OFF_W3 + c] = ((S.vpart[2][0][c] + S.vpart[2][1][c]) + S.vpart[2][2][c]) + S.vpart[2][3][c];
    }
    if (tid == 0) {
        P[MLP_OFF_B3] = gb3;
        P[MLP_NPARAM] = loss;
    }
}

__global__ void mlp_reduce_kernel(const float *__restrict__ part, int nparts, float *__restrict__ g,
                                  float *__restrict__ loss_acc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i > MLP_NPARAM) return;
    float s = 0.0f;
    for (int p = 0; p < nparts; p++) s += part[(size_t)p * (MLP_NPARAM + 1) + i];
    g[i] = s;
    if (i == MLP_NPARAM && loss_acc) *loss_acc += s;
}

// Deterministic sum of x[0..n): SUM_PARTS strips, then one wave over the strip sums.
constexpr int SUM_PARTS = 1024;
__global__ void strip_sum_kernel(const float *__restrict__ x, int64_t n, float *__restrict__ part) {
    const int64_t per = (n + SUM_PARTS - 1) / SUM_PARTS;
    const int64_t lo = (int64_t)blockIdx.x * per, hi = min(n, lo + per);
    float s = 0.0f;
    for (int64_t i = lo + threadIdx.x; i < hi; i += 64) s += x[i];
    s = wave_sum(s);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}
__global__ void parts_sum_kernel(const float *__restrict__ part, float *__restrict__ out) {
    float s = 0.0f;
    for (int i = threadIdx.x; i < SUM_PARTS; i += 64) s += part[i];
    s = wave_sum(s);
    if (threadIdx.x == 0) *out = s;
}

__global__ void adam_kernel(float *__restrict__ w, const float *__restrict__ g, float *__restrict__ m,
                            float *__restrict__ v, int n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
}

__global__ void to_bf16_kernel(const float *__restrict__ w, uint16_t *__restrict__ wb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < MLP_IN * MLP_H) wb[i] = f2bf(w[MLP_OFF_W1 + i]);
    else if (i < MLP_NBF16) wb[i] = f2bf(w[MLP_OFF_W2 + i - MLP_IN * MLP_H]);
}

size_t mlp_fwd_lds_bytes() { return sizeof(FwdLds); }
size_t mlp_grad_lds_bytes() { return sizeof(GradLds); }
size_t mlp_grad_part_floats() { return (size_t)GRAD_GRID * (MLP_NPARAM + 1); }

void mlp_refresh_bf16(const float *w, uint16_t *wb, hipStream_t s) {
    hipLaunchKernelGGL(to_bf16_kernel, dim3((MLP_NBF16 + 255) / 256), dim3(256), 0, s, w, wb);
    GH_HIP(hipGetLastError());
}

void mlp_forward_launch(const float *w, const uint16_t *wb, const float *X, int64_t n, float *y, hipStream_t s) {
    static const bool attr = [] {
        GH_HIP(hipFuncSetAttribute((const void *)mlp_forward_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)sizeof(FwdLds)));
        return true;
    }();
    (void)attr;
    if (n <= 0) return;
    // at most 256 workgroups, each loading the weights once and looping over
    // tiles: a workgroup holds 69 KB of LDS, and one per tile waited for a CU
    // with that much room behind the other runs' workgroups
    const int64_t tiles = (n + TR - 1) / TR;
    const int grid = (int)std::min<int64_t>(tiles, 256);
    hipLaunchKernelGGL(mlp_forward_kernel, dim3(grid), dim3(MLP_NT), sizeof(FwdLds), s, w, wb, X, n, y);
    GH_HIP(hipGetLastError());
}

static void grad_launch(const float *w, const uint16_t *wb, const MlpBatch &B, float inv_wsum, float *part, float *g,
                        float *loss_acc, hipStream_t s) {
    static const bool attr = [] {
        GH_HIP(hipFuncSetAttribute((const void *)mlp_grad_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)sizeof(GradLds)));
        return true;
    }();
    (void)attr;
    GH_CHECK(B.n > 0 && B.n_total > 0 && B.a > 0, GH_ERR_ARG, "mlp gradient: empty batch");
    // every slice is written, also by workgroups that get no rows
    hipLaunchKernelGGL(mlp_grad_kernel, dim3(GRAD_GRID), dim3(MLP_NT), sizeof(GradLds), s, w, wb, B, inv_wsum, part);
    hipLaunchKernelGGL(mlp_reduce_kernel, dim3((MLP_NPARAM + 1 + 255) / 256), dim3(256), 0, s, part, GRAD_GRID, g,
                       loss_acc);
    GH_HIP(hipGetLastError());
}

void mlp_grad_launch(const float *w, const uint16_t *wb, const MlpBatch &B, float inv_wsum, float *part, float *g,
                     hipStream_t s) {
    grad_launch(w, wb, B, inv_wsum, part, g, nullptr, s);
}

void adam_launch(float *w, const float *g, float *m, float *v, float lr, int step, hipStream_t s) {
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float c1 = (float)(1.0 - std::pow(0.9, step)), c2 = (float)(1.0 - std::pow(0.999, step));
    hipLaunchKernelGGL(adam_kernel, dim3((MLP_NPARAM + 255) / 256), dim3(256), 0, s, w, g, m, v, MLP_NPARAM, lr, b1,
                       b2, eps, c1, c2);
    GH_HIP(hipGetLastError());
}

/* ------------------------------------------------------------------ host objects */
static float host_normal(uint32_t seed, uint32_t i) {
    // Box-Muller on two counter draws
    const float u1 = ((float)(philox_draw(seed, 0u, P_MLP_INIT, i, 0u) >> 8) + 0.5f) * (1.0f / 16777216.0f);
    const float u2 = u01(philox_draw(seed, 0u, P_MLP_INIT, i, 1u));
    return std::sqrt(-2.0f * std::log(u1)) * std::cos(6.2831853f * u2);
}

static void reset_opt(gh_mlp *M) {
    GH_HIP(hipMemsetAsync(M->m.p, 0, sizeof(float) * MLP_NPARAM, M->s));
    GH_HIP(hipMemsetAsync(M->v.p, 0, sizeof(float) * MLP_NPARAM, M->s));
    M->step = 0;
}

void mlp_set(gh_mlp *M, const float *w) {
    GH_CHECK(M && w, GH_ERR_ARG, "gh_mlp: bad arguments");
    GH_HIP(hipMemcpyAsync(M->w.p, w, sizeof(float) * MLP_NPARAM, hipMemcpyHostToDevice, M->s));
    mlp_refresh_bf16(M->w.as<float>(), M->wb.as<uint16_t>(), M->s);
    reset_opt(M);
    GH_HIP(hipStreamSynchronize(M->s));
}

void mlp_get(gh_mlp *M, float *w) {
    GH_CHECK(M && w, GH_ERR_ARG, "gh_mlp: bad arguments");
    GH_HIP(hipMemcpyAsync(w, M->w.p, sizeof(float) * MLP_NPARAM, hipMemcpyDeviceToHost, M->s));
    GH_HIP(hipStreamSynchronize(M->s));
}

void mlp_create(uint32_t seed, gh_mlp **out) {
    GH_CHECK(out, GH_ERR_ARG, "gh_mlp_create: null output");
    auto *M = new gh_mlp();
    try {
        GH_HIP(hipStreamCreateWithFlags(&M->s, hipStreamNonBlocking));
        M->w.reserve(sizeof(float) * MLP_NPARAM);
        M->wb.reserve(sizeof(uint16_t) * MLP_NBF16);
        M->m.reserve(sizeof(float) * MLP_NPARAM);
        M->v.reserve(sizeof(float) * MLP_NPARAM);
        M->g.reserve(sizeof(float) * (MLP_NPARAM + 1));
        M->part.reserve(sizeof(float) * std::max(mlp_grad_part_floats(), (size_t)SUM_PARTS));
        // He initialisation for the ReLU layers, zero biases
        std::vector<float> w(MLP_NPARAM, 0.0f);
        const float s1 = std::sqrt(2.0f / MLP_IN), s2 = std::sqrt(2.0f / MLP_H), s3 = std::sqrt(1.0f / MLP_H);
        for (int i = 0; i < MLP_IN * MLP_H; i++) w[MLP_OFF_W1 + i] = s1 * host_normal(seed, (uint32_t)(MLP_OFF_W1 + i));
        for (int i = 0; i < MLP_H * MLP_H; i++) w[MLP_OFF_W2 + i] = s2 * host_normal(seed, (uint32_t)(MLP_OFF_W2 + i));
        for (int i = 0; i < MLP_H; i++) w[MLP_OFF_W3 + i] = s3 * host_normal(seed, (uint32_t)(MLP_OFF_W3 + i));
        mlp_set(M, w.data());
        *out = M;
    } catch (...) {
        mlp_destroy(M);
        throw;
    }
}

void mlp_destroy(gh_mlp *M) {
    if (!M) return;
    hipStream_t s = M->s;
    delete M;
    if (s) (void)hipStreamDestroy(s);
}

static void stage_rows(gh_mlp *M, int64_t n, const float *X, const float *t, const float *w) {
    GH_CHECK(n > 0 && n < (int64_t)1 << 40 && X, GH_ERR_ARG, "gh_mlp: empty input");
    dalloc_copy(M->X, X, (size_t)n * MLP_IN, M->s);
    if (t) dalloc_copy(M->t, t, (size_t)n, M->s);
    if (w) dalloc_copy(M->wt, w, (size_t)n, M->s);
}

void mlp_forward_host(gh_mlp *M, int64_t n, const float *X, float *y) {
    GH_CHECK(M && y, GH_ERR_ARG, "gh_mlp_forward: bad arguments");
    stage_rows(M, n, X, nullptr, nullptr);
    M->y.reserve(sizeof(float) * (size_t)n);
    mlp_forward_launch(M->w.as<float>(), M->wb.as<uint16_t>(), M->X.as<float>(), n, M->y.as<float>(), M->s);
    GH_HIP(hipMemcpyAsync(y, M->y.p, sizeof(float) * (size_t)n, hipMemcpyDeviceToHost, M->s));
    GH_HIP(hipStreamSynchronize(M->s));
}

static float host_inv_wsum(int n, const float *w) {
    double s = 0.0;
    for (int i = 0; i < n; i++) s += w ? w[i] : 1.0f;
    GH_CHECK(s > 0.0, GH_ERR_ARG, "gh_mlp: row weights sum to zero");
    return (float)(1.0 / s);
}

void mlp_backward_host(gh_mlp *M, int n, const float *X, const float *t, const float *w, float *grad, float *loss) {
    GH_CHECK(M && t, GH_ERR_ARG, "gh_mlp_backward: bad arguments");
    stage_rows(M, n, X, t, w);
    MlpBatch B{M->X.as<float>(), M->t.as<float>(), w ? M->wt.as<float>() : nullptr, n, 0, n, 1, 0};
    mlp_grad_launch(M->w.as<float>(), M->wb.as<uint16_t>(), B, host_inv_wsum(n, w), M->part.as<float>(),
                    M->g.as<float>(), M->s);
    std::vector<float> g(MLP_NPARAM + 1);
    GH_HIP(hipMemcpyAsync(g.data(), M->g.p, sizeof(float) * g.size(), hipMemcpyDeviceToHost, M->s));
    GH_HIP(hipStreamSynchronize(M->s));
    if (grad) std::memcpy(grad, g.data(), sizeof(float) * MLP_NPARAM);
    if (loss) *loss = g[MLP_NPARAM];
}

void mlp_step_host(gh_mlp *M, int n, const float *X, const float *t, const float *w, float lr, float *loss) {
    mlp_backward_host(M, n, X, t, w, nullptr, loss);
    adam_launch(M->w.as<float>(), M->g.as<float>(), M->m.as<float>(), M->v.as<float>(), lr, ++M->step, M->s);
    mlp_refresh_bf16(M->w.as<float>(), M->wb.as<uint16_t>(), M->s);
    GH_HIP(hipStreamSynchronize(M->s));
}

static uint64_t gcd64(uint64_t a, uint64_t b) {
    while (b) { const uint64_t r = a % b; a = b; b = r; }
    return a;
}

void mlp_train_device(gh_mlp *M, const float *X, const float *Y, const float *W, int64_t n_total, int epochs,
                      int batch, float lr, uint32_t seed, hipStream_t data_stream, float *loss_per_epoch) {
    GH_CHECK(M && X && Y && n_total > 0 && epochs > 0 && batch > 0, GH_ERR_ARG, "gh_mlp_train: bad arguments");
    if (data_stream) GH_HIP(hipStreamSynchronize(data_stream));
    hipStream_t s = M->s;
    float *wsum = (float *)M->acc.reserve(sizeof(float) * (size_t)(epochs + 1));
    float frac = 1.0f;
    if (W) {
        hipLaunchKernelGGL(strip_sum_kernel, dim3(SUM_PARTS), dim3(64), 0, s, W, n_total, M->part.as<float>());
        hipLaunchKernelGGL(parts_sum_kernel, dim3(1), dim3(64), 0, s, M->part.as<float>(), wsum + epochs);
        GH_HIP(hipGetLastError());
        float ws = 0.0f;
        GH_HIP(hipMemcpyAsync(&ws, wsum + epochs, sizeof(float), hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
        GH_CHECK(ws > 0.0f, GH_ERR_ARG, "gh_mlp_train: no labelled rows");
        frac = ws / (float)n_total;
    }
    GH_HIP(hipMemsetAsync(wsum, 0, sizeof(float) * (size_t)epochs, s));
    const int64_t spe = (n_total + batch - 1) / batch;
    const int64_t total = spe * epochs;
    int64_t done = 0;
    for (int e = 0; e < epochs; e++) {
        uint64_t a = 1, b = 0;
        if (n_total > 1) {
            const uint32_t r0 = philox_draw(seed, (uint32_t)e, P_MLP_PERM, 0u, 0u);
            const uint32_t r1 = philox_draw(seed, (uint32_t)e, P_MLP_PERM, 1u, 0u);
            a = 1 + (uint64_t)r0 % (uint64_t)(n_total - 1);
            while (gcd64(a, (uint64_t)n_total) != 1) a = a % (uint64_t)(n_total - 1) + 1;
            b = (uint64_t)r1 % (uint64_t)n_total;
        }
        for (int64_t st = 0; st < spe; st++, done++) {
            const int64_t i0 = st * batch;
            const int n = (int)std::min<int64_t>(batch, n_total - i0);
            MlpBatch B{X, Y, W, n_total, i0, n, a, b};
            grad_launch(M->w.as<float>(), M->wb.as<uint16_t>(), B, 1.0f / ((float)n * frac), M->part.as<float>(),
                        M->g.as<float>(), wsum + e, s);
            const float lr_s = lr * (1.0f - 0.9f * (float)done / (float)std::max<int64_t>(total, 1));
            adam_launch(M->w.as<float>(), M->g.as<float>(), M->m.as<float>(), M->v.as<float>(), lr_s, ++M->step, s);
            mlp_refresh_bf16(M->w.as<float>(), M->wb.as<uint16_t>(), s);
        }
    }
    std::vector<float> le(epochs);
    GH_HIP(hipMemcpyAsync(le.data(), wsum, sizeof(float) * epochs, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
    if (loss_per_epoch)
        for (int e = 0; e < epochs; e++) loss_per_epoch[e] = le[e] / (float)spe;
}

}  // namespace gh
