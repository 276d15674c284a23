/*
 * glimpse_oracle.h -- C interface of the CPU oracle.
 *
 * TEST INFRASTRUCTURE ONLY.  The oracle is a from-scratch C++ restatement of
 * the GLIMPSE2_phase compute core (sparse-PBWT state selection, conditioning
 * set, haploid Li-Stephens forward-backward, diploid phasing HMM, Gibbs
 * schedule, GP accumulation).  It is the checker for the HIP path and the
 * timed CPU baseline.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product library never does.
 *
 * Every buffer is caller-owned; sizes are given explicitly.  Functions return
 * 0 on success and a negative code on error (message via orc_last_error).
 */
#ifndef GLIMPSE_ORACLE_H
#define GLIMPSE_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Random stream used for every draw of a run.
 * ORC_RNG_REFERENCE : one global std::mt19937 consumed in the reference's
 *                     order with libstdc++ distributions and std::sample
 *                     (common/src/utils/random_number.h:35-125).
 * ORC_RNG_COUNTER   : Philox4x32-10 keyed by (seed) with counters
 *                     (iteration|purpose, a, b, 0); identical to the HIP path. */
enum { ORC_RNG_REFERENCE = 0, ORC_RNG_COUNTER = 1 };

typedef struct {
    int32_t n_sites;            /* L */
    int32_t n_haps;             /* N */
    const double *cm;           /* [L] genetic position, baseline subtracted */
    const uint32_t *cref;       /* [L] REF allele count in the panel */
    const uint32_t *calt;       /* [L] ALT allele count in the panel */
    const uint8_t *lq;          /* [L] low-quality flag (excluded from PBWT) */
    const uint64_t *common_words; /* [n_common][ceil(N/64)], bit h%64 of word h/64 = allele of hap h */
    const int32_t *rare_ptr;    /* [L+1] CSR: haps carrying the non-major allele at rare sites */
    const int32_t *rare_idx;
} orc_panel_desc;

typedef struct {
    int32_t burnin, main_iter;  /* --burnin, --main */
    int32_t ne;                 /* --ne */
    float min_gl, err_imp, err_phase;
    int32_t pbwt_depth;
    float pbwt_modulo_cm;
    int32_t kinit, kpbwt;
    float sparse_maf;
    uint32_t seed;
    int32_t rng_mode;
    int32_t n_threads;          /* --threads: targets of an iteration spread over threads after the
                                   serial PBWT sweep (caller_algorithm.cpp:94-116); counter mode only,
                                   the reference's shared mt19937 makes threaded runs irreproducible */
    const int32_t *ploidy;      /* [T] 1 or 2 per target (the --samples-file ploidy); NULL: all diploid */
} orc_params;

void orc_default_params(orc_params *p);
const char *orc_last_error(void);

/* Panel: classifies sites (common/major), builds the sparse PBWT (pack3
 * stream + A_small_idx) exactly as split_reference does. */
void *orc_panel_create(const orc_panel_desc *d, float sparse_maf);
void orc_panel_free(void *panel);
int32_t orc_panel_n_common(void *panel);
int32_t orc_panel_n_common_hq(void *panel);
void orc_panel_site_class(void *panel, uint8_t *flag_common, uint8_t *major);
int64_t orc_panel_ypacked(void *panel, uint8_t *out, int64_t cap);
/* A_small_idx as CSR over HQ common index 0..n_common_hq (inclusive) */
int64_t orc_panel_asmall(void *panel, int32_t *ptr, int32_t *idx, int64_t cap);

/* One complete GLIMPSE2_phase run (INIT + burnin + main) on diploid targets.
 * gl    : [T][L][3] phred PL (uint8)
 * flat  : [T][L]
 * kpbwt_per_target : optional [T] (NULL = params->kpbwt for all); a target's
 *        value changes only its own conditioning-set size.
 * out_gp0/out_gp1 : [T][L] averaged posteriors (1,0 where never stored)
 * out_hds : [T][L]; out_h0/out_h1 : [T][L] best-guess haplotypes
 * out_nstates : [T][n_iter] states used per iteration (optional) */
int orc_phase_run(void *panel, int32_t T, const uint8_t *gl, const uint8_t *flat,
                  const orc_params *p, const int32_t *kpbwt_per_target,
                  float *out_gp0, float *out_gp1, uint8_t *out_hds,
                  uint8_t *out_h0, uint8_t *out_h1, int32_t *out_nstates);

/* ---- fine-grained entry points for parity tests ---- */

/* Sparse-PBWT sweep (haplotype_set::matchHapsFromCompressedPBWTSmall) for all
 * target haplotypes at once.  hap_alleles: [2T][L] current target alleles;
 * stored: [n_common_hq] selection flags; start_idx: [2T] initial insertion
 * index.  Output: the reference's pbwt_states[T][depth] lists flattened:
 * out_cnt[T*depth], entries written in list order to out_ids with
 * out_off[T*depth] offsets (cap = capacity of out_ids). */
int orc_pbwt_select(void *panel, int32_t T, const uint8_t *hap_alleles,
                    const uint8_t *stored, const int32_t *start_idx,
                    int32_t depth, float modulo_cm,
                    int32_t *out_off, int32_t *out_cnt, int32_t *out_ids, int64_t cap);

/* allocatePBWT group boundaries + effective modulo for a given Kpbwt. */
int orc_pbwt_groups(void *panel, int32_t depth, float modulo_cm, int32_t kpbwt,
                    float *eff_modulo, int32_t *grp, int32_t cap, int32_t *n_grp);

/* conditioning_set::compactSelection (Kpbwt branch) + updateTransitions for
 * one target given its pbwt_states lists.  Outputs sorted state ids, the
 * polymorphic site list, Hvar in the reference's MSB-first byte layout
 * (row stride ceil(K/8) bytes) and t (float). */
int orc_condset(void *panel, int32_t depth, const int32_t *off, const int32_t *cnt,
                const int32_t *ids, int32_t kpbwt, int32_t ne,
                int32_t *out_states, int32_t *out_K,
                int32_t *out_poly, int32_t *out_npoly,
                uint8_t *out_hvar, int64_t hvar_cap, float *out_t);

/* Same, for an explicit sorted state list (INIT / full-panel paths). */
int orc_condset_from_states(void *panel, const int32_t *states, int32_t K, int32_t ne,
                            int32_t *out_poly, int32_t *out_npoly,
                            uint8_t *out_hvar, int64_t hvar_cap, float *out_t);

/* imputation_hmm::computePosteriors for one haplotype.  HL [2L], flat [L],
 * states/K define the conditioning set.  Output HP [2L]. */
int orc_impute_hap(void *panel, const int32_t *states, int32_t K, int32_t ne,
                   float err_imp, const float *HL, const uint8_t *flat, float *HP);

/* genotype::initHaplotypeLikelihoods / makeHaplotypeLikelihoods for one target. */
int orc_hap_likelihoods(int32_t L, const uint8_t *gl, const uint8_t *flat,
                        const uint8_t *other_hap, int32_t mode, float min_gl, float *HL);

/* phasing_hmm::rephaseHaplotypes for one target with counter-mode draws
 * (iteration it, target index tgt, seed).  h0/h1 are updated in place. */
int orc_rephase(void *panel, const int32_t *states, int32_t K, int32_t ne,
                float err_phase, const uint8_t *flat, uint8_t *h0, uint8_t *h1,
                uint32_t seed, int32_t it, int32_t tgt);

/* The deterministic quantities of one rephase (for pinning tests): per
 * segment start the backward pass's per-hypothesis sums over the states
 * divided by their total (n_segs x 8), per segment boundary TRANS_HAP's 8x8
 * joint hypothesis probabilities divided by their total ((n_segs-1) x 64), and
 * per flat het the probability of allele 1 under each hypothesis (n_miss x 8). */
int orc_phase_trace(void *panel, const int32_t *states, int32_t K, int32_t ne, float err_phase, const uint8_t *flat,
                    const uint8_t *h0, const uint8_t *h1, uint32_t seed, int32_t it, int32_t tgt, int32_t cap_segs,
                    int32_t cap_miss, int32_t *n_segs, int32_t *n_miss, float *seg_sums, float *trans_hap,
                    float *of1);

/* The sparse-PBWT sweep of orc_pbwt_select with a trace (for pinning tests):
 * after every HQ common site, per target haplotype, one row (site, haplotype,
 * idx, f, g, site of the last reset of [f, g) or -1). */
int orc_pbwt_trace(void *panel, int32_t T, const uint8_t *hap_alleles, const uint8_t *stored,
                   const int32_t *start_idx, int32_t depth, float modulo_cm, int32_t *rows, int64_t cap,
                   int64_t *n_rows);

/* Philox4x32-10 block (for known-answer tests). */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* The n-th output (1-based) of a default-seeded std::mt19937. */
uint32_t orc_mt19937_nth(uint32_t seed, int64_t n);
float orc_unphred(int32_t i);
int orc_output_fields_ploidy(int32_t T, int32_t L, const int32_t *ploidy, const float *gp0, const float *gp1,
                             const uint8_t *hds, const uint32_t *cref, const uint32_t *calt, float *ds, float *gp,
                             uint8_t *gt, float *info, float *af, float *raf);
int orc_output_fields(int32_t T, int32_t L, const float *gp0, const float *gp1, const uint8_t *hds,
                      const uint32_t *cref, const uint32_t *calt, float *ds, float *gp, uint8_t *gt, float *info,
                      float *af, float *raf);

#ifdef __cplusplus
}
#endif
#endif
