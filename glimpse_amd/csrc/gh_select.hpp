// gh_select.hpp -- learned conditioning-state selection (features, top-k, labels).
#pragma once
#include <stdint.h>

#include "gh_internal.hpp"
#include "gh_mlp.hpp"

namespace gh {

constexpr int MLP_FEAT = MLP_IN;
// 4-wave workgroups (one candidate per wave at a time): larger ones wait for a
// whole CU's wave slots behind other runs' workgroups when runs share the GPU
constexpr int FEAT_NT = 256;
constexpr int FEAT_WIN = 8;                   // windows of the per-window mismatch features
constexpr int FEAT_CONF_PHRED = 10;           // a genotype is "confident" when both others are >= 10 phred
constexpr size_t FEAT_LDS_MAX = 144 * 1024;
constexpr int TOPK_NT = 256;
constexpr int TOPK_MAX = 2048;                // largest pool (conditioning sets are capped at 2048)
constexpr float LABEL_EPS = 1e-3f;            // label = ln(K * mass + LABEL_EPS)

// A candidate's "segment" is the stretch of the chunk around the stored site
// whose PBWT record listed it: common-site words [cw - FEAT_FAR, cw + FEAT_FAR)
// with cw = (common row of that site) / 64, clipped to the chunk.  Its label is
// its posterior mass averaged over the polymorphic sites of that segment.
constexpr int FEAT_NEAR = 4;     // words
constexpr int FEAT_FAR = 16;     // words
constexpr int FEAT_RUN = 64;     // longest clean run looked at on each side (words)

// Feature layout (MLP_FEAT = 32 per candidate), for candidate j of a pool of K
// in PBWT insertion order, and the target's current haplotypes H0/H1:
//   0 first PBWT rank / depth   1 j / K   2 listed by H1's walk   3 from the upper neighbours
//   4,5   mismatch rate vs H0, H1 over all common sites
//   6,7   mismatch rate vs H0, H1 over common sites with read data
//   8,9   longest run of identical 64-site words vs H0, H1 / words
//   10,11 mismatch rate vs H0, H1 in words [cw-4, cw)     12,13 in [cw, cw+4)
//   14,15 mismatch rate vs H0, H1 in words [cw-16, cw)    16,17 in [cw, cw+16)
//   18,19 mismatch rate vs H0, H1 at sites with data in the segment
//   20,21 identical words running back from cw-1 (H0, H1) / 64
//   22,23 identical words running forward from cw (H0, H1) / 64
//   24 alt where confidently RR, 25 ref where confidently AA, 26 alt where
//      confidently RA (each over such sites of the segment)   27 alt rate in the segment
//   28,29 shares a rare allele with H0, H1
//   30 cw / words   31 segment sites / (64 * 2 * FEAT_FAR)
struct FeatArgs {
    const uint64_t *hm;        // transposed common panel [N][Wc]
    int Wc, Lc, N, L, depth, kstride;
    const int32_t *c2t;        // common row -> site
    const int32_t *crow;       // site -> common row (-1 if not common)
    const uint8_t *flags;
    const int32_t *rare_ptr, *rare_idx;
    const uint8_t *H;          // [2T][L] current haplotypes
    const uint64_t *tmask;     // [T][4][Wc]
    const int32_t *targets;    // [n] target per workgroup
    const int32_t *Kpool;      // [T]
    const int32_t *ins;        // [T][kstride] pool in insertion order
    const int32_t *ins_site;   // [T][kstride] site of the listing record
    float *X;                  // rows row0 + i*rstride + j
    int64_t row0;
    int rstride;               // rows per target (>= the largest pool)
};

struct TopkArgs {
    const int32_t *targets;    // [n]
    const int32_t *Kpool;      // [T] (pool size = current K)
    const int32_t *klearn;     // [T]
    const float *y;            // scores, rows row0 + i*rstride + j
    int64_t row0;
    int kstride, rstride;
    float min_mass;
    float prior;               // weight of the PBWT-order prior in the ranking (0: scores alone)
    const int32_t *ins;        // [T][kstride]
    int32_t *states;           // [T][kstride] sorted state lists (pool in, selection out)
    int32_t *Kdev;             // [T]
    CondItem *items;           // [T]
    float *coverage;           // [T] predicted mass share of the kept states
    int32_t *counters;         // [0] learned selections, [1] kept pools
    // group quota (when ins_site != NULL): candidates compete within the PBWT
    // group of the site that listed them; groups are served round robin
    const int32_t *ins_site;   // [T][kstride]
    const int32_t *site_grp;   // [n_classes][L] PBWT group of a site (-1 if none)
    const int32_t *tcls;       // [T] class of each target
    int L;
};

// Per-state segment windows in sorted-state order (for the FB mass output).
struct WinArgs {
    const int32_t *targets;    // [n]
    const int32_t *Kpool;      // [T]
    const int32_t *states;     // [T][kstride] sorted pool
    const int32_t *ins, *ins_site;
    const int32_t *c2t, *crow;
    int Lc, Wc, kstride;
    int32_t *win_lo, *win_hi;  // [T][kstride]
};

struct LabelArgs {
    const int32_t *targets;    // [n]
    const CondItem *items;     // [T] (K, Lp, state_off, site_off)
    const int32_t *states;
    const int32_t *poly;       // polymorphic site lists (site_off per target)
    const int32_t *win_lo, *win_hi;
    const int32_t *ins;        // [T][kstride]
    const float *mass0, *mass1;   // per-state posterior mass summed over the state's segment (both haplotypes)
    int kstride, rstride;
    float *X, *Y, *W;          // rows row0 + i*rstride + j
    int64_t row0;
};

void target_mask_launch(gh_panel *P, int T, const uint8_t *gl, const uint8_t *flat, uint64_t *tmask, hipStream_t s);
size_t feature_lds_bytes(const gh_panel *P);
void feature_launch(gh_panel *P, int n, const FeatArgs &a, hipStream_t s);
void topk_launch(int n, const TopkArgs &a, hipStream_t s);
void label_launch(int n, const LabelArgs &a, hipStream_t s);
void win_launch(int n, const WinArgs &a, hipStream_t s);
void select_features_batch(gh_panel *P, int T, const uint8_t *H_h, const uint8_t *gl_h, const uint8_t *flat_h,
                           const int32_t *ins_h, const int32_t *site_h, const int32_t *K_h, int kstride, int depth,
                           float *X_h);
void select_topk_batch(int T, const float *y_h, const int32_t *ins_h, const int32_t *K_h, const int32_t *klearn_h,
                       int kstride, float min_mass, int32_t *states_h, int32_t *Kout_h, float *cov_h);

}  // namespace gh
