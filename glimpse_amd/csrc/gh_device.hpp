// gh_device.hpp -- device-side helpers shared by the HIP kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define GH_WAVE 64

namespace gh {

// ---------------------------------------------------------------- DPP / wave
template <int CTRL>
__device__ __forceinline__ float dppf(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, false));
}
// Value of lane l ^ 16 / l ^ 32 (v_permlane16_swap / v_permlane32_swap with
// both operands v: the first result holds the even rows' (low half's) values in
// both rows (halves) of a pair, the second the odd rows' (high half's)).
__device__ __forceinline__ float partner16(float v) {
    const auto r = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float((threadIdx.x & 16) ? r[0] : r[1]);
}
__device__ __forceinline__ float partner32(float v) {
    const auto r = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float((threadIdx.x & 32) ? r[0] : r[1]);
}

// Sum of v over the 64 lanes, returned wave-uniform.  Four DPP adds reduce each
// 16-lane row (xor1, xor2, half-mirror, mirror), so every lane holds its row's
// sum R0..R3; row_bcast:15 adds row 0 into row 1 and row 2 into row 3, then
// row_bcast:31 adds row 1 into rows 2 and 3, and lane 63 holds
// (R3 + R2) + (R1 + R0), the same value as (R0 + R1) + (R2 + R3).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dppf_rows(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, false));
}
__device__ __forceinline__ float wave_sum(float v) {
    v += dppf<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dppf<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dppf<0x141>(v);  // row_half_mirror
    v += dppf<0x140>(v);  // row_mirror
    v += dppf_rows<0x142, 0xA>(v);   // row_bcast:15 into rows 1 and 3
    v += dppf_rows<0x143, 0xC>(v);   // row_bcast:31 into rows 2 and 3
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}
// Three independent sums at once (interleaved for ILP).
__device__ __forceinline__ void wave_sum3(float &x, float &y, float &z) {
    x += dppf<0xB1>(x); y += dppf<0xB1>(y); z += dppf<0xB1>(z);
    x += dppf<0x4E>(x); y += dppf<0x4E>(y); z += dppf<0x4E>(z);
    x += dppf<0x141>(x); y += dppf<0x141>(y); z += dppf<0x141>(z);
    x += dppf<0x140>(x); y += dppf<0x140>(y); z += dppf<0x140>(z);
    x += dppf_rows<0x142, 0xA>(x); y += dppf_rows<0x142, 0xA>(y); z += dppf_rows<0x142, 0xA>(z);
    x += dppf_rows<0x143, 0xC>(x); y += dppf_rows<0x143, 0xC>(y); z += dppf_rows<0x143, 0xC>(z);
    x = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), 63));
    y = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(y), 63));
    z = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(z), 63));
}

// Sum of v over the 64 lanes, the same value in every lane, in 8 VALU
// instructions: the four row DPP adds of wave_sum, then the xor-32 and xor-16
// partners.  A permlane swap of v with itself returns the low half's (even
// rows') values and the high half's (odd rows') values of each pair, so their
// sum is the pair's sum in every lane, with the same operands in the same order.
__device__ __forceinline__ float wave_allsum(float v) {
    v += dppf<0xB1>(v);
    v += dppf<0x4E>(v);
    v += dppf<0x141>(v);
    v += dppf<0x140>(v);
    const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    const auto b = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}
// Two sums at once: returns the wave total of x in lanes 0-31 and of y in
// lanes 32-63 (12 VALU instructions).  v_permlane32_swap of (x, y) gives
// [x low half, y low half] and [x high half, y high half]; their sum holds each
// value's two halves combined, and the xor-16 step then adds the row pairs.
__device__ __forceinline__ float wave_sum_pair(float x, float y) {
    x += dppf<0xB1>(x); y += dppf<0xB1>(y);
    x += dppf<0x4E>(x); y += dppf<0x4E>(y);
    x += dppf<0x141>(x); y += dppf<0x141>(y);
    x += dppf<0x140>(x); y += dppf<0x140>(y);
    const auto a = __builtin_amdgcn_permlane32_swap(__float_as_uint(x), __float_as_uint(y), false, false);
    const float v = __uint_as_float(a[0]) + __uint_as_float(a[1]);
    const auto b = __builtin_amdgcn_permlane16_swap(__float_as_uint(v), __float_as_uint(v), false, false);
    return __uint_as_float(b[0]) + __uint_as_float(b[1]);
}

// Reduce-scatter of 8 per-lane values over the wave: returns, in every lane l,
// the wave total of value (l & 7).  Three DPP halving steps inside each group
// of 8 lanes (partner 7-i, i^2, i^1; each lane keeps the half its lane bit
// selects), then the 8 groups are combined (row_ror:8, then the xor-16 and
// xor-32 partners by permlane swaps).
__device__ __forceinline__ float wave_rs8(const float v[8]) {
    const int i = threadIdx.x & 7;
    const bool b2 = i & 4, b1 = i & 2, b0 = i & 1;
    float u[4], w[2];
#pragma unroll
    for (int k = 0; k < 4; k++) u[k] = (b2 ? v[4 + k] : v[k]) + dppf<0x141>(b2 ? v[k] : v[4 + k]);
#pragma unroll
    for (int k = 0; k < 2; k++) w[k] = (b1 ? u[2 + k] : u[k]) + dppf<0x4E>(b1 ? u[k] : u[2 + k]);
    float x = (b0 ? w[1] : w[0]) + dppf<0xB1>(b0 ? w[0] : w[1]);
    x += dppf<0x128>(x);
    x += partner16(x);
    x += partner32(x);
    return x;
}

// The same for 32 values: five halving steps (lane bits 2, 1, 0, 3, 4), then the
// two wave halves are added, so lanes l and l ^ 32 both hold the wave total of
// value rs32_index(l) = 16*b2 + 8*b1 + 4*b0 + 2*b3 + b4.
__device__ __forceinline__ int rs32_index(int l) {
    return 16 * ((l >> 2) & 1) + 8 * ((l >> 1) & 1) + 4 * (l & 1) + 2 * ((l >> 3) & 1) + ((l >> 4) & 1);
}
__device__ __forceinline__ float wave_rs32(const float v[32]) {
    const int l = threadIdx.x & 63;
    const bool b2 = l & 4, b1 = l & 2, b0 = l & 1, b3 = l & 8, b4 = l & 16;
    float b[16], c[8], d[4], e[2];
#pragma unroll
    for (int k = 0; k < 16; k++) b[k] = (b2 ? v[16 + k] : v[k]) + dppf<0x141>(b2 ? v[k] : v[16 + k]);
#pragma unroll
    for (int k = 0; k < 8; k++) c[k] = (b1 ? b[8 + k] : b[k]) + dppf<0x4E>(b1 ? b[k] : b[8 + k]);
#pragma unroll
    for (int k = 0; k < 4; k++) d[k] = (b0 ? c[4 + k] : c[k]) + dppf<0xB1>(b0 ? c[k] : c[4 + k]);
#pragma unroll
    for (int k = 0; k < 2; k++) e[k] = (b3 ? d[2 + k] : d[k]) + dppf<0x128>(b3 ? d[k] : d[2 + k]);
    const float x = (b4 ? e[1] : e[0]) + partner16(b4 ? e[0] : e[1]);
    return x + partner32(x);
}

// A wave-uniform copy of a value every lane holds (a work-item descriptor, an
// LDS broadcast): each 32-bit word goes through v_readfirstlane, so the
// compiler keeps it in SGPRs and branches on it with scalar branches.
template <typename T>
__device__ __forceinline__ T wave_uniform(const T &v) {
    static_assert(sizeof(T) % 4 == 0, "wave_uniform: size must be a multiple of 4 bytes");
    T r;
    const int *src = reinterpret_cast<const int *>(&v);
    int *dst = reinterpret_cast<int *>(&r);
#pragma unroll
    for (int k = 0; k < (int)(sizeof(T) / 4); k++) dst[k] = __builtin_amdgcn_readfirstlane(src[k]);
    return r;
}

// ---------------------------------------------------------------- counter RNG
// Philox4x32-10 (Salmon et al., SC'11).  Every draw of the Gibbs sampler is
// philox({it<<8 | purpose, a, b, 0}, {seed, 0})[0].
__host__ __device__ __forceinline__ uint32_t philox_draw(uint32_t seed, uint32_t it, uint32_t purpose,
                                                         uint32_t a, uint32_t b) {
    uint32_t c0 = (it << 8) | purpose, c1 = a, c2 = b, c3 = 0u;
    uint32_t k0 = seed, k1 = 0u;
#pragma unroll
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        c0 = hi1 ^ c1 ^ k0;
        c1 = lo1;
        c2 = hi0 ^ c3 ^ k1;
        c3 = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    return c0;
}
__host__ __device__ __forceinline__ float u01(uint32_t x) { return (float)(x >> 8) * (1.0f / 16777216.0f); }
__host__ __device__ __forceinline__ uint32_t urange(uint32_t x, uint32_t lo, uint32_t hi) {
    return lo + (uint32_t)(((uint64_t)x * (uint64_t)(hi - lo + 1u)) >> 32);
}

enum Purpose {
    P_H0 = 0, P_H1 = 1, P_DIP_ROOT = 2, P_DIP_SEG = 3, P_FLAT_HET = 4,
    P_MONO_HET = 5, P_PBWT_STORED = 6, P_PBWT_START = 7, P_INIT_SEL = 8, P_MLP_INIT = 9, P_MLP_PERM = 10
};

// site flag bits in the resident panel
enum { SF_COMMON = 1, SF_MAJOR = 2, SF_LQ = 4 };

}  // namespace gh
