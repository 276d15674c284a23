/*
 * glimpse_hip.h -- C-ABI of the MI355X-native GLIMPSE2_phase compute core.
 *
 * Plain C types only: caller-owned host buffers, library-owned device state
 * behind opaque handles, int status codes (GH_OK = 0, negative on error,
 * message via gh_last_error).  Nothing here exits the process.  One HIP
 * stream per handle; calls on different handles may run from different
 * threads.  All randomness is counter based (Philox4x32-10 keyed by the
 * seed, counters = iteration, purpose, target, site), so results do not
 * depend on thread scheduling.
 *
 * Each entry point names the reference interface it replaces
 * (paths relative to the GLIMPSE repository root).
 */
#ifndef GLIMPSE_HIP_H
#define GLIMPSE_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
    GH_OK = 0,
    GH_ERR_ARG = -1,      /* invalid argument / shape */
    GH_ERR_HIP = -2,      /* HIP runtime error (message holds hipGetErrorString) */
    GH_ERR_NUMERIC = -3,  /* underflow the reference asserts on (phasing_hmm.cpp:174-175) */
    GH_ERR_CAPACITY = -4, /* an output buffer is too small */
    GH_ERR_STATE = -5     /* zero conditioning states (conditioning_set.cpp:125) */
};

typedef struct gh_panel gh_panel;

/* One chunk of the reference panel, as split_reference serialises it
 * (common/src/containers/ref_haplotype_set.h:144-160, variant_map.h:71-84).
 * common_words: [n_common][ceil(N/64)] uint64, bit (h % 64) of word h/64 is
 * the allele of haplotype h at that common site (site order).
 * rare_ptr/rare_idx: CSR over all sites, haplotypes carrying the non-major
 * allele at rare sites (empty rows at common sites).
 * ypacked / asmall_*: the sparse-PBWT stream and A_small_idx (CSR with
 * n_common_hq + 2 pointers) produced by gh_build_sparse_pbwt. */
typedef struct {
    int32_t n_sites;
    int32_t n_haps;
    const double *cm;
    const uint32_t *cref;
    const uint32_t *calt;
    const uint8_t *lq;              /* may be NULL (all high quality) */
    const uint64_t *common_words;
    const int32_t *rare_ptr;
    const int32_t *rare_idx;
    const uint8_t *ypacked;
    int64_t ypacked_len;
    const int32_t *asmall_ptr;
    const int32_t *asmall_idx;
    float sparse_maf;               /* 0.001 */
} gh_panel_desc;

/* Model and schedule parameters (phase/src/caller/caller_parameters.cpp:34-98). */
typedef struct {
    int32_t burnin;        /* 5 */
    int32_t main_iter;     /* 15 */
    int32_t ne;            /* 100000 */
    float min_gl;          /* 1e-10 */
    float err_imp;         /* 1e-12, clamped to [1e-12, 1e-3] */
    float err_phase;       /* 1e-4 */
    int32_t pbwt_depth;    /* 12 */
    float pbwt_modulo_cm;  /* 0.1 */
    int32_t kinit;         /* 1000 */
    int32_t kpbwt;         /* 2000 */
    uint32_t seed;         /* 15052011 */
} gh_params;

/* Per-run counters (for the roofline): sums over targets and iterations. */
typedef struct {
    double fb_state_sites;     /* sum of K * L_p over haploid FB runs */
    double fb_poly_sites;      /* sum of L_p over haploid FB runs */
    double ph_state_sites;     /* sum of K * (phasing loci) */
    double ph_segments;        /* sum of (segments + flat hets) */
    double pbwt_hap_sites;     /* target haplotypes x sites swept */
    double ms_select, ms_condset, ms_fb, ms_phase, ms_other, ms_total;
    double fb_kernel_ms;       /* summed HIP-event duration of the forward-backward kernel launches */
    double fb_launches;        /* number of those launches */
    double fb_bytes;           /* their algorithmic HBM bytes: per haplotype Lp*(2*8*ceil(K/64) + 8*K + 16) + 8*L */
} gh_run_stats;

const char *gh_last_error(void);
void gh_default_params(gh_params *p);

/* split_reference producer step: ref_haplotype_set::build_sparsePBWT
 * (common/src/containers/ref_haplotype_set.cpp:47-79).  Two calls: first with
 * NULL outputs to get sizes, then with buffers. */
int gh_build_sparse_pbwt(const gh_panel_desc *d, uint8_t *ypacked, int64_t *ypacked_len,
                         int32_t *asmall_ptr, int32_t *asmall_idx, int64_t *asmall_len,
                         int32_t *n_common_hq);

/* Replaces read_binary_reference_panel + transposeRareRef + allocatePBWT's
 * panel part (phase/src/caller/caller_initialise.cpp:99-164): uploads the
 * panel once; it stays resident in HBM as bit-packed words. */
int gh_panel_create(const gh_panel_desc *d, gh_panel **out);
int gh_panel_destroy(gh_panel *p);
int gh_panel_info(const gh_panel *p, int32_t *n_common, int32_t *n_common_hq, int64_t *device_bytes);

/* conditioning_set::select for explicit state sets, batched
 * (phase/src/containers/conditioning_set.cpp:75-196).  Item i uses
 * states[state_ptr[i] .. state_ptr[i+1]) (sorted, unique).  Outputs per item:
 * poly[poly_ptr[i] ..] polymorphic site list, t (float transitions, one fewer
 * than sites), hvar rows of ceil(K/64) uint64 words (bit j of word w = state
 * 64w+j).  poly_ptr has n_items+1 entries; hvar_ptr (in words) too. */
int gh_condset_batch(gh_panel *p, int32_t n_items, const int32_t *state_ptr, const int32_t *states,
                     int32_t ne, int32_t *poly_ptr, int32_t *poly, int64_t poly_cap,
                     int64_t *hvar_ptr, uint64_t *hvar, int64_t hvar_cap, float *t);

/* imputation_hmm::computePosteriors batched over haplotypes
 * (phase/src/models/imputation_hmm.cpp:73-367).  HL: [n_items][2L],
 * flat: [n_items][L], HP out: [n_items][2L]. */
int gh_impute_batch(gh_panel *p, int32_t n_items, const int32_t *state_ptr, const int32_t *states,
                    const float *HL, const uint8_t *flat, float err_imp, int32_t ne, float *HP);

/* phasing_hmm::rephaseHaplotypes batched over diploid targets
 * (phase/src/models/phasing_hmm.cpp:247-302) with counter draws keyed by
 * (seed, it, target index tgt[i]).  h0/h1: [n_items][L] updated in place. */
int gh_rephase_batch(gh_panel *p, int32_t n_items, const int32_t *state_ptr, const int32_t *states,
                     const uint8_t *flat, uint8_t *h0, uint8_t *h1, float err_phase, int32_t ne,
                     uint32_t seed, int32_t it, const int32_t *tgt);

/* haplotype_set::matchHapsFromCompressedPBWTSmall
 * (phase/src/containers/haplotype_set.cpp:317-394) for 2T target haplotypes:
 * hap_alleles [2T][L], stored [n_common_hq], start_idx [2T].  Output: the
 * pbwt_states[T][depth] lists (off/cnt [T*depth], ids).  When out_states is
 * not NULL, also the conditioning set conditioning_set::compactSelection
 * takes from those lists (phase/src/containers/conditioning_set.cpp:96-104):
 * out_K[t] sorted ids in out_states[t*kpbwt ..]. */
int gh_pbwt_select(gh_panel *p, int32_t T, const uint8_t *hap_alleles, const uint8_t *stored,
                   const int32_t *start_idx, int32_t depth, float modulo_cm,
                   int32_t *out_off, int32_t *out_cnt, int32_t *out_ids, int64_t cap,
                   int32_t kpbwt, int32_t *out_states, int32_t *out_K);

/* One complete GLIMPSE2_phase Gibbs run on diploid targets
 * (phase/src/caller/caller_algorithm.cpp:53-243): INIT + burnin + main with
 * selection, conditioning sets, both haploid FB passes, sampling, phasing and
 * GP accumulation on the GPU.  gl [T][L][3] phred, flat [T][L];
 * kpbwt_per_target optional [T].  Outputs [T][L]: gp0, gp1, hds, best-guess
 * h0/h1; nstates [T][n_iter] optional; stats optional. */
int gh_phase_run(gh_panel *p, int32_t T, const uint8_t *gl, const uint8_t *flat, const gh_params *prm,
                 const int32_t *kpbwt_per_target, float *gp0, float *gp1, uint8_t *hds,
                 uint8_t *h0, uint8_t *h1, int32_t *nstates, gh_run_stats *stats);

/* The same run split in three so a caller can time or repeat the device part
 * with its inputs resident: create uploads the inputs and allocates the
 * workspace (everything gh_phase_run allocates), execute runs the full Gibbs
 * schedule on the device (a fresh chain each call, same seed, same result),
 * fetch copies the outputs back.  These have no single reference counterpart:
 * together they are caller::phase_loop split at its I/O boundaries
 * (phase/src/caller/caller_management.cpp:47-56). */
typedef struct gh_run gh_run;
int gh_run_create(gh_panel *p, int32_t T, const uint8_t *gl, const uint8_t *flat, const gh_params *prm,
                  const int32_t *kpbwt_per_target, gh_run **out);
int gh_run_execute(gh_run *r);
int gh_run_fetch(gh_run *r, float *gp0, float *gp1, uint8_t *hds, uint8_t *h0, uint8_t *h1, int32_t *nstates,
                 gh_run_stats *stats);
int gh_run_destroy(gh_run *r);

#ifdef __cplusplus
}
#endif
#endif
