// gh_pbwt.hpp -- sparse-PBWT structures shared by the host decoder and kernels.
#pragma once
#include <stdint.h>

#include <vector>

#include "gh_internal.hpp"

namespace gh {

// One non-LQ site of the sweep, in site order.
struct SiteOp {
    int32_t k;         // site index
    int32_t l;         // HQ common index (common) / index of the run's A_small_idx (rare)
    int32_t ref;       // ref_rac_l: count of 0 symbols (full or small PBWT)
    int32_t prev_ref;  // ref of the previous HQ common site
    int32_t nbits;     // size of the small set of A_small_idx at l: of the rare run ending here (common), of this run (rare)
    int32_t si_loc;    // init ops: offset of that small set in its walk tile's copy, -1 when not staged
    int32_t next_common;   // rare ops: index of the next common op (the end of the run), n_ops if none
    int32_t pad1;
    int64_t roff;      // offset of this site's column words / popcount prefixes (common or rare arrays)
    uint8_t kind;      // 0 common, 1 rare
    uint8_t init;      // common: rare run just ended (init_common); rare: first site of a run (init_rare)
    uint8_t major;     // rare: major allele
    uint8_t pad0;
    // genetic positions the walk's reset tests compare (so the walk reads no
    // map from global memory): cM of site k, of site k-1, and of site ps, the
    // site common2tot[l-1] an init_common reset moves last_reset to
    int32_t ps;
    float cm_k, cm_km1, cm_ps;
    int32_t a_off;     // offset of that small set in A_small_idx (A_small_ptr[l])
};
static_assert(sizeof(SiteOp) == 64, "SiteOp is one 64-byte record");

// Prefix arrays of the sweep, named by code: 2*(l+1) is the order after HQ
// common site l (code 0: the identity order before the first site), 2*(l+1)+1
// the order entering site l (after the rearrangement of a rare run that ends
// there).  Only every S-th "after" array is stored; any other entry is found
// by stepping back through the sites' columns (inverse LF mapping: the p-th 0
// or 1 of a column in prefix order) and the rare-run rearrangements.
__host__ __device__ inline int32_t code_after(int l) { return 2 * (l + 1); }
__host__ __device__ inline int32_t code_entering(int l) { return 2 * (l + 1) + 1; }

// Per HQ common site: where its column lives and how its entering order was made.
struct ColInfo {
    int64_t roff;    // common column words / popcount prefixes (prefix order entering the site)
    int32_t ref;     // zeros in the column (cref)
    int32_t ns;      // carriers of the rare run that ends here (0: none)
    int32_t a_off;   // their positions in A_small_idx and their final small order in sfin
    int32_t ck;      // index of the stored "after" array, -1 if not stored
};
static_assert(sizeof(ColInfo) == 24, "ColInfo is 24 bytes");

// Arguments of one selection launch.
struct PbwtRun {
    int n_haps, n_samples, depth, cap;
    const uint8_t *haps;        // current alleles, one row of L per haplotype
    const int32_t *hap_row;     // [n_haps] row of each walked haplotype in haps (NULL: identity)
    const int32_t *start;       // [n_haps]
    const uint8_t *stored;      // [n_classes][Lhq]
    const int32_t *hap_class;   // [n_haps]
    const float *modulo;        // [n_classes]
    int32_t *ev;                // [n_haps][cap][2 + 2*depth]
    int32_t *ev_ci;             // [n_haps][cap][2]: prefix-array code and insertion index of each record
    int32_t *ev_cnt;            // [n_haps]
    int32_t *status;
    // optional compaction
    const int32_t *kpbwt;       // [n_samples]
    const int32_t *sample_hap;  // [n_samples] first walked haplotype of the sample
    const uint8_t *sample_one;  // [n_samples] 1: haploid sample, its second walker's lists are not used (NULL: none)
    const int32_t *sample_row;  // [n_samples] output row (NULL: identity)
    int32_t *out_states;        // [n_samples][out_stride]
    int out_stride;
    int32_t *out_K;             // [n_samples]
    // optional: the same states in insertion order, id | rank << 20 | second
    // haplotype << 26 | from the upper neighbours << 27  ([n_samples][out_stride])
    int32_t *ins;
    int32_t *ins_site;          // optional: site of the record that listed each of them
};

void build_sparse_pbwt(const gh_panel_desc *d, std::vector<uint8_t> &Y, std::vector<int32_t> &aptr,
                       std::vector<int32_t> &aidx, int &n_hq);
void pbwt_decode(gh_panel *P);
void pbwt_select_launch(gh_panel *P, const PbwtRun &R, hipStream_t s);
int pbwt_event_cap(gh_panel *P, float modulo, int n_groups);
float pbwt_groups(gh_panel *P, int depth, float modulo_in, int kpbwt, std::vector<int> &grp);

}  // namespace gh

struct gh_panel::PbwtDev {
    gh::DBuf ops, bits, cnt, cbits, ccnt, A, aidx, sfin, cols, tiles, tsi;
    int n_ops = 0, n_runs = 0, n_arrays = 0, n_tiles = 0;
    int ck_every = 1;            // S: one stored prefix array every S HQ common sites
    bool stage_common = true;    // common columns staged in the walk's LDS tiles (else read globally)
    int64_t total_words = 0, total_cwords = 0;   // rare / common column words
};

namespace gh {
// Sites between stored prefix arrays for a panel of Lhq HQ common sites and N
// haplotypes: every site when they fit in budget_mb (GH_PBWT_CKPT_MB, default
// 1024 MiB), else the smallest S that does; GH_PBWT_CKPT_EVERY pins S.
int pbwt_ckpt_every(int Lhq, int N);
}  // namespace gh
