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
    int32_t a_out;     // prefix array after this common site
    int32_t a_in;      // prefix array before it (after any rare-run rearrangement)
    int32_t a_prev;    // prefix array after the previous common site (iota at the start)
    int64_t roff;      // offset of this site's column words / popcount prefixes
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
    int32_t *ev_cnt;            // [n_haps]
    int32_t *status;
    // optional compaction
    const int32_t *kpbwt;       // [n_samples]
    const int32_t *sample_hap;  // [n_samples] first walked haplotype of the sample
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
    gh::DBuf ops, bits, cnt, A, aidx, tiles;
    int n_ops = 0, n_runs = 0, n_arrays = 0, n_tiles = 0;
    int64_t total_words = 0;
};
