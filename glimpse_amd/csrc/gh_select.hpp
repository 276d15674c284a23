// gh_select.hpp -- learned conditioning-state selection (features, top-k, labels).
#pragma once
#include <stdint.h>

#include "gh_internal.hpp"
#include "gh_mlp.hpp"

namespace gh {

constexpr int MLP_FEAT = MLP_IN;
constexpr int FEAT_NT = 512;                  // 8 waves, one candidate per wave at a time
constexpr int FEAT_WIN = 8;                   // windows of the per-window mismatch features
constexpr int FEAT_CONF_PHRED = 10;           // a genotype is "confident" when both others are >= 10 phred
constexpr size_t FEAT_LDS_MAX = 144 * 1024;
constexpr int TOPK_NT = 1024;
constexpr int TOPK_MAX = 2048;                // largest pool (conditioning sets are capped at 2048)
constexpr float LABEL_EPS = 1e-3f;            // label = ln(K * mass + LABEL_EPS)

// Feature layout (MLP_FEAT = 32 per candidate), for candidate j of a pool of K
// in PBWT insertion order, and the target's current haplotypes H0/H1:
//   0 first PBWT rank / depth      1 j / K     2 listed first by H1's walk
//   3 listed from the upper neighbours
//   4,5 mismatch rate vs H0, H1 over common sites with read data
//   6,7 mismatch rate vs H0, H1 over all common sites
//   8,9 longest run of identical 64-common-site words vs H0, H1 / words
//   10 alt allele where the target is confidently RR (/ such sites)
//   11 ref allele where confidently AA   12 alt allele where confidently RA
//   13 alt allele rate   14,15 shares a rare allele with H0, H1
//   16..23 mismatch rate vs H0 in 8 equal windows of common sites; 24..31 vs H1
struct FeatArgs {
    const uint64_t *hm;        // transposed common panel [N][Wc]
    int Wc, Lc, N, L, depth, kstride;
    const int32_t *c2t;        // common row -> site
    const uint8_t *flags;
    const int32_t *rare_ptr, *rare_idx;
    const uint8_t *H;          // [2T][L] current haplotypes
    const uint64_t *tmask;     // [T][4][Wc]
    const int32_t *targets;    // [n] target per workgroup
    const int32_t *Kpool;      // [T]
    const int32_t *ins;        // [T][kstride] pool in insertion order
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

struct LabelArgs {
    const int32_t *targets;    // [n]
    const CondItem *items;     // [T] (K, Lp, state_off)
    const int32_t *states;
    const int32_t *ins;        // [T][kstride]
    const float *mass0, *mass1;   // per-state posterior mass summed over sites (both haplotypes)
    int kstride, rstride;
    float *X, *Y, *W;          // rows row0 + i*rstride + j
    int64_t row0;
};

void target_mask_launch(gh_panel *P, int T, const uint8_t *gl, const uint8_t *flat, uint64_t *tmask, hipStream_t s);
size_t feature_lds_bytes(const gh_panel *P);
void feature_launch(gh_panel *P, int n, const FeatArgs &a, hipStream_t s);
void topk_launch(int n, const TopkArgs &a, hipStream_t s);
void label_launch(int n, const LabelArgs &a, hipStream_t s);
void select_features_batch(gh_panel *P, int T, const uint8_t *H_h, const uint8_t *gl_h, const uint8_t *flat_h,
                           const int32_t *ins_h, const int32_t *K_h, int kstride, int depth, float *X_h);
void select_topk_batch(int T, const float *y_h, const int32_t *ins_h, const int32_t *K_h, const int32_t *klearn_h,
                       int kstride, float min_mass, int32_t *states_h, int32_t *Kout_h, float *cov_h);

}  // namespace gh
