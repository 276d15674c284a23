// gh_internal.hpp -- host-side internals shared by the HIP translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>
#include <vector>

#include "glimpse_hip.h"
#include "gh_fb_layout.hpp"

#include "gh_error.hpp"

namespace gh {

#define GH_HIP(x)                                                                               \
    do {                                                                                        \
        hipError_t _e = (x);                                                                    \
        if (_e != hipSuccess)                                                                   \
            throw ::gh::Error(GH_ERR_HIP, std::string(#x) + ": " + hipGetErrorString(_e));      \
    } while (0)

// Device buffer with RAII; grows, never shrinks.
struct DBuf {
    void *p = nullptr;
    size_t bytes = 0;
    DBuf() = default;
    DBuf(const DBuf &) = delete;
    DBuf &operator=(const DBuf &) = delete;
    ~DBuf() { release(); }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    void *reserve(size_t b) {
        if (b > bytes) {
            release();
            GH_HIP(hipMalloc(&p, b ? b : 1));
            bytes = b;
        }
        return p;
    }
    template <typename T>
    T *as() const { return (T *)p; }
};

template <typename T>
inline T *dalloc_copy(DBuf &buf, const T *h, size_t n, hipStream_t s) {
    buf.reserve(sizeof(T) * (n ? n : 1));
    if (n) GH_HIP(hipMemcpyAsync(buf.p, h, sizeof(T) * n, hipMemcpyHostToDevice, s));
    return buf.as<T>();
}

}  // namespace gh

// Resident panel (one chunk).
struct gh_panel {
    int L = 0, N = 0, Lc = 0, Lhq = 0;
    int max_rare_mac = 0;             // largest minor-allele count of a rare, high-quality site
    int W64 = 0;   // ceil(N/64): words per common site row
    int Wc = 0;    // ceil(Lc/64): words per haplotype row of the transposed panel
    std::vector<double> cm;
    std::vector<float> cm_pos;
    std::vector<uint8_t> flags;       // SF_* bits
    std::vector<int32_t> common2tot;
    std::vector<uint32_t> cref, calt;
    std::vector<int32_t> rare_ptr_h, rare_idx_h;   // carriers of the non-major allele per site
    int64_t device_bytes = 0;
    int device = 0;                   // the HIP device the panel lives on
    hipStream_t stream = nullptr;

    gh::DBuf d_cm, d_cmpos, d_flags, d_crow, d_c2t, d_sm, d_hm, d_shap_ptr, d_shap_idx, d_rare_ptr,
        d_rare_idx, d_cref, d_calt;
    // sparse PBWT (per-site structures for the selection kernel)
    std::vector<uint8_t> ypacked;
    std::vector<int32_t> asmall_ptr, asmall_idx;
    struct PbwtDev;
    PbwtDev *pbwt = nullptr;
};

namespace gh {

// A launch replayed from a captured graph (gh_graph.hpp) is gated by the
// device: it returns at once unless *key == want, workgroups at or beyond *n
// return, and the iteration that keys its draws is read from *iter.  A
// default gate (key == nullptr) gates nothing.
struct GraphGate {
    const int32_t *key = nullptr;
    int32_t want = 0;
    const int32_t *n = nullptr;
    const uint32_t *iter = nullptr;
};

// Per-item description of a conditioning set living on the device.
struct CondItem {
    int32_t K;          // states
    int32_t W;          // ceil(K/64)
    int32_t Lp;         // polymorphic sites (filled by the index kernel)
    int32_t pad;
    int64_t state_off;  // into states
    int64_t site_off;   // into per-site scratch (L per item): poly list, poly_of_site
    int64_t hvar_off;   // words, rows of W words, L rows reserved
};

// Builds conditioning sets for n items whose sorted states are on the device.
// items/poly/hvar/t are device arrays sized by the caller:
//   poly, pos: L per item; t: L per item; hvar: L*W per item (words).
// Returns with Lp of each item written to items[i].Lp (device).
// The same in two steps, for callers that place each item's rows by its Lp:
// condset_index writes Lp (and poly, pos, t), condset_fill the Hvar rows at
// the items' hvar_off.
void condset_index(gh_panel *P, int n, CondItem *d_items, const int32_t *d_states, int32_t *d_poly, int32_t *d_pos,
                   float *d_t, uint8_t *d_mark, int ne, hipStream_t s);
void condset_fill(gh_panel *P, int n, const CondItem *d_items, const int32_t *d_states, const int32_t *d_poly,
                  const int32_t *d_pos, uint64_t *d_hvar, int maxW, hipStream_t s);
void condset_build(gh_panel *P, int n, CondItem *d_items, const CondItem *h_items, const int32_t *d_states,
                   int32_t *d_poly, int32_t *d_pos, float *d_t, uint64_t *d_hvar, uint8_t *d_mark,
                   int ne, int maxW, hipStream_t s);

// FB over n haploid items.  HL [item][2L], flat [item][L] (device); alpha
// checkpoints per item: fb_ckpt_floats(Lp, W, maxW) floats at alpha_off[i].
// HP out [item][2L]; optional per-state posterior mass at mass + state_off,
// summed over the polymorphic sites inside the state's [win_lo, win_hi) when
// those per-state windows are given (at win + state_off), else over all sites.
struct FbItem {
    int32_t K, W, Lp, cond;   // cond = index of the CondItem
    int64_t alpha_off;        // floats
    int64_t hl_off;           // into HL / HP ([2L] per item)
    int64_t flat_off;         // into flat ([L] per item)
};
// ckpt_maxW: the word count the checkpoint buffers were sized for (fb_ckpt_floats)
void fb_launch(gh_panel *P, int n, const FbItem *d_fb, const CondItem *d_cond, int maxW, int ckpt_maxW,
               const int32_t *d_poly,
               const float *d_t, const uint64_t *d_hvar, const float *d_HL, const uint8_t *d_flat, float *d_ckpt,
               float *d_HP, float ee, float ed, hipStream_t s, float *d_mass = nullptr,
               const int32_t *d_win_lo = nullptr, const int32_t *d_win_hi = nullptr, FbPick pick = FbPick{0, 0},
               const int32_t *d_order = nullptr, GraphGate gate = GraphGate{});
// HP at monomorphic sites (closed form) for the same items.
void fb_mono_launch(gh_panel *P, int n, const FbItem *d_fb, const CondItem *d_cond, const int32_t *d_pos,
                    const float *d_HL, const uint8_t *d_flat, float *d_HP, float ee, float ed, hipStream_t s);

}  // namespace gh
