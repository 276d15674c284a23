// gh_noalloc.hip -- gh_impute_batch_noalloc: the haploid forward-backward batch
// (the kernels of gh_impute_batch) with every device buffer it needs carved
// out of one file-scope device array, for callers that cannot allocate device
// memory (code under graph capture, tight serving loops, harnesses that forbid
// hipMalloc).  The array is GH_NOALLOC_SCRATCH_BYTES per device and lives
// until the process ends; its device address is looked up once per device
// with hipGetSymbolAddress and only that address is handed to kernels and
// copies.  Every call on a device shares the array and leaves the last call's
// contents in it; calls that overlap on one device race.  Nothing in it has to
// start at zero: every kernel writes what it later reads.
#include <algorithm>
#include <cstring>
#include <map>
#include <mutex>
#include <vector>

#include "gh_internal.hpp"

namespace gh {

static __device__ __align__(256) unsigned char scratch[GH_NOALLOC_SCRATCH_BYTES];

// device address of `scratch` on the current device (looked up on its first use there)
static unsigned char *scratch_ptr(int device) {
    static std::mutex mu;
    static std::map<int, unsigned char *> by_device;
    std::lock_guard<std::mutex> lk(mu);
    auto it = by_device.find(device);
    if (it != by_device.end()) return it->second;
    unsigned char *ptr = nullptr;
    GH_HIP(hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch)));
    GH_CHECK(ptr != nullptr, GH_ERR_HIP, "gh_impute_batch_noalloc: no device address for the scratch array");
    by_device[device] = ptr;
    return ptr;
}

// Bump allocation inside the scratch array, 256-byte aligned.
struct Carve {
    size_t used = 0;
    size_t take(size_t bytes) {
        const size_t at = used;
        used += (bytes + 255) & ~(size_t)255;
        return at;
    }
};

void impute_batch_noalloc(gh_panel *P, int n, const int32_t *state_ptr, const int32_t *states, const float *HL,
                          const uint8_t *flat, float err_imp, int ne, float *HP) {
    GH_CHECK(P && n > 0 && state_ptr && states && HL && flat && HP, GH_ERR_ARG,
             "gh_impute_batch_noalloc: bad arguments");
    int dev = -1;
    GH_HIP(hipGetDevice(&dev));
    GH_CHECK(dev == P->device, GH_ERR_ARG, "gh_impute_batch_noalloc: the panel lives on another device");
    const int L = P->L;
    // item tables (host), the same validation as gh_impute_batch
    std::vector<CondItem> items(n);
    std::vector<FbItem> fb(n);
    int maxW = 0;
    int64_t words = 0;
    for (int i = 0; i < n; i++) {
        const int K = state_ptr[i + 1] - state_ptr[i];
        GH_CHECK(K > 0, GH_ERR_STATE, "conditioning set " + std::to_string(i) + " has zero states");
        GH_CHECK(K <= P->N && K <= 4096, GH_ERR_ARG, "gh_impute_batch_noalloc: at most min(N, 4096) states per set");
        for (int k = state_ptr[i]; k < state_ptr[i + 1]; k++) {
            GH_CHECK(states[k] >= 0 && states[k] < P->N, GH_ERR_ARG, "state id out of range");
            GH_CHECK(k == state_ptr[i] || states[k] > states[k - 1], GH_ERR_ARG, "states must be sorted and unique");
        }
        CondItem &c = items[i];
        c.K = K;
        c.W = (K + 63) / 64;
        c.Lp = 0;
        c.pad = 0;
        c.state_off = state_ptr[i] - state_ptr[0];
        c.site_off = (int64_t)i * L;
        c.hvar_off = words;
        words += (int64_t)L * c.W;
        maxW = std::max(maxW, c.W);
    }
    // checkpoints sized by L (a bound on the polymorphic sites), so the layout
    // is known before anything runs
    int64_t aoff = 0;
    for (int i = 0; i < n; i++) {
        fb[i] = FbItem{items[i].K, items[i].W, 0, i, aoff, (int64_t)i * 2 * L, (int64_t)i * L};
        aoff += fb_ckpt_floats(L, items[i].W, maxW);
    }
    const int64_t n_states = state_ptr[n] - state_ptr[0];
    Carve cv;
    const size_t o_items = cv.take(sizeof(CondItem) * n), o_fb = cv.take(sizeof(FbItem) * n),
                 o_states = cv.take(sizeof(int32_t) * n_states), o_HL = cv.take(sizeof(float) * (size_t)n * 2 * L),
                 o_flat = cv.take((size_t)n * L), o_HP = cv.take(sizeof(float) * (size_t)n * 2 * L),
                 o_poly = cv.take(sizeof(int32_t) * (size_t)n * L), o_pos = cv.take(sizeof(int32_t) * (size_t)n * L),
                 o_t = cv.take(sizeof(float) * (size_t)n * L), o_mark = cv.take((size_t)n * L),
                 o_hvar = cv.take(sizeof(uint64_t) * (size_t)std::max<int64_t>(words, 1)),
                 o_alpha = cv.take(sizeof(float) * (size_t)std::max<int64_t>(aoff, 1));
    GH_CHECK(cv.used <= (size_t)GH_NOALLOC_SCRATCH_BYTES, GH_ERR_CAPACITY,
             "gh_impute_batch_noalloc: the batch needs " + std::to_string(cv.used >> 20) +
                 " MiB of scratch, more than its " + std::to_string(GH_NOALLOC_SCRATCH_BYTES >> 20) +
                 " MiB (split the batch)");
    unsigned char *base = scratch_ptr(dev);
    auto at = [&](size_t off) { return (void *)(base + off); };
    hipStream_t s = P->stream;
    GH_HIP(hipMemcpyAsync(at(o_items), items.data(), sizeof(CondItem) * n, hipMemcpyHostToDevice, s));
    GH_HIP(hipMemcpyAsync(at(o_fb), fb.data(), sizeof(FbItem) * n, hipMemcpyHostToDevice, s));
    GH_HIP(hipMemcpyAsync(at(o_states), states + state_ptr[0], sizeof(int32_t) * n_states, hipMemcpyHostToDevice, s));
    GH_HIP(hipMemcpyAsync(at(o_HL), HL, sizeof(float) * (size_t)n * 2 * L, hipMemcpyHostToDevice, s));
    GH_HIP(hipMemcpyAsync(at(o_flat), flat, (size_t)n * L, hipMemcpyHostToDevice, s));
    CondItem *d_items = (CondItem *)at(o_items);
    condset_build(P, n, d_items, items.data(), (const int32_t *)at(o_states), (int32_t *)at(o_poly),
                  (int32_t *)at(o_pos), (float *)at(o_t), (uint64_t *)at(o_hvar), (uint8_t *)at(o_mark), ne, maxW, s);
    const float ei = std::min(std::max(err_imp, 1e-12f), 1e-3f);   // caller_initialise.cpp:152
    const float ed = ei, ee = (float)(1.0 - (double)ei);
    fb_launch(P, n, (const FbItem *)at(o_fb), d_items, maxW, maxW, (const int32_t *)at(o_poly), (const float *)at(o_t),
              (const uint64_t *)at(o_hvar), (const float *)at(o_HL), (const uint8_t *)at(o_flat), (float *)at(o_alpha),
              (float *)at(o_HP), ee, ed, s);
    fb_mono_launch(P, n, (const FbItem *)at(o_fb), d_items, (const int32_t *)at(o_pos), (const float *)at(o_HL),
                   (const uint8_t *)at(o_flat), (float *)at(o_HP), ee, ed, s);
    GH_HIP(hipMemcpyAsync(HP, at(o_HP), sizeof(float) * (size_t)n * 2 * L, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
}

}  // namespace gh
