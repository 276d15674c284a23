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
    GH_ERR_STATE = -5,    /* zero conditioning states (conditioning_set.cpp:125) */
    GH_ERR_IO = -6        /* unreadable, malformed or unwritable file */
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
    double ms_learn;           /* learned selection: features + network + top-k (included in ms_select) */
    double learned_sel;        /* target-iterations that used the learned selection */
    double pool_kept;          /* target-iterations that fell back to the whole PBWT pool */
    double collected_rows;     /* training rows recorded by gh_run_collect */
    double fb_nw, fb_mwl;      /* forward-backward launch shape of the last iteration (waves, words per lane) */
    double ph_nw, ph_kpt;      /* phasing launch shape of the last iteration (waves, states per thread) */
    double graph_iters;        /* iterations replayed from the captured graph (gh_run_set_graph) */
    double graph_reruns;       /* why the execution left the graph (bits): 1 captured buffers outgrown, 2 launch
                                  groups changed (1, 2: repeated host-launched), 4 capture refused by the
                                  runtime (the same iterations enqueued without a graph) */
    double ms_graph;           /* wall time of the capture and the replays (not in ms_select .. ms_phase) */
} gh_run_stats;

const char *gh_last_error(void);
void gh_default_params(gh_params *p);

/* split_reference producer step: ref_haplotype_set::build_sparsePBWT
 * (common/src/containers/ref_haplotype_set.cpp:47-79).  Two calls: first with
 * NULL outputs to get sizes, then with buffers; the second call on the same
 * thread with the same (unchanged) description reuses the first one's build. */
int gh_build_sparse_pbwt(const gh_panel_desc *d, uint8_t *ypacked, int64_t *ypacked_len,
                         int32_t *asmall_ptr, int32_t *asmall_idx, int64_t *asmall_len,
                         int32_t *n_common_hq);

/* Replaces read_binary_reference_panel + transposeRareRef + allocatePBWT's
 * panel part (phase/src/caller/caller_initialise.cpp:99-164): uploads the
 * panel once; it stays resident in HBM as bit-packed words. */
int gh_panel_create(const gh_panel_desc *d, gh_panel **out);
int gh_panel_destroy(gh_panel *p);
int gh_panel_info(const gh_panel *p, int32_t *n_common, int32_t *n_common_hq, int64_t *device_bytes);

/* The INIT iteration's state selection: haplotype_set::initRareTar +
 * performSelection_RARE_INIT_GL (phase/src/containers/haplotype_set.cpp:
 * 108-169, 775-833), for T diploid targets at once, with the counter draws of
 * gh_run_execute's INIT iteration (selection sampling in place of std::sample).
 * gl [T][L][3], flat [T][L]; out: states [T][kstride] (ascending ids), K [T].
 * kinit is clipped to N. */
int gh_init_states(gh_panel *p, int32_t T, const uint8_t *gl, const uint8_t *flat, int32_t kinit, uint32_t seed,
                   int32_t kstride, int32_t *states, int32_t *K);

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

/* gh_impute_batch with no device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * same arguments, checks, kernels and results as gh_impute_batch; every
 * device buffer (inputs, conditioning sets, checkpoints, outputs) is carved
 * from one file-scope device array of GH_NOALLOC_SCRATCH_BYTES per device,
 * held until the process ends.  A batch that needs more returns
 * GH_ERR_CAPACITY (split it).  The panel's device must be current.  Calls on
 * one device share the array: calls that overlap on one device race. */
#define GH_NOALLOC_SCRATCH_BYTES (256u << 20)
int gh_impute_batch_noalloc(gh_panel *p, int32_t n_items, const int32_t *state_ptr, const int32_t *states,
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
 * outputs (the workspace comes with the first execute, see gh_workspace
 * below), execute runs the full Gibbs
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
/* Execution workspace (no reference counterpart: memory management for
 * batches).  A run's execution builds and discards device buffers: the
 * current haplotypes, likelihoods and posteriors, state sets, conditioning
 * sets (placed by each target's polymorphic-site count), alpha checkpoints,
 * phasing loci and messages, selection records.  By default each run keeps
 * its own, allocated by its first execution and grown as needed.  Runs that
 * execute one after another (never at the same time) can share one
 * workspace, so a batch of many runs holds only the workspaces of the runs
 * executing at once; the run's inputs and outputs stay its own.  The
 * workspace must outlive the runs that use it. */
typedef struct gh_workspace gh_workspace;
int gh_workspace_create(gh_workspace **out);
int gh_workspace_destroy(gh_workspace *w);
int gh_workspace_bytes(const gh_workspace *w, int64_t *bytes);
/* NULL: the run's own workspace again. */
int gh_run_set_workspace(gh_run *r, gh_workspace *w);
/* Launch shapes (waves per target and work per lane of the forward-backward
 * and phasing kernels).  mode 0 (the default): throughput defaults, the
 * fewest waves per target that fill a CU, right when many runs share the GPU.
 * mode 1: autotuning; the PBWT iterations of the first executions each run
 * the next candidate shape of the current set-size class and time it; once all
 * candidates are timed the fastest is kept for the rest of the run's
 * executions.  mode 2: latency defaults, more waves per target, for the run
 * whose chain is the critical path of a batch (the longest chunk).  Shapes
 * change only the summation order, so results agree to float rounding; a run
 * is bit-reproducible with modes 0 and 2 (or with GH_FB_SHAPE=<NW>x<MWL> and
 * GH_PHASE_SHAPE=<NW>x<KPT> pinning the shapes in mode 0). */
int gh_run_set_tuning(gh_run *r, int32_t mode);

/* How the iterations are launched (no reference counterpart: the schedule of
 * caller::phase_loop, caller_algorithm.cpp:223-243, on the device).  mode 1
 * (the default): iterations 0 (INIT) and 1 are launched from the host; the
 * next one is captured once as a HIP graph in which every per-iteration
 * decision (stored PBWT sites and start indices, the placement of Hvar rows,
 * checkpoints, loci and phasing messages by each target's counts, the launch
 * groups' membership and counts) is made by device kernels, and the graph is
 * replayed for every later iteration with no host round trip.  The graph holds
 * the launch shapes of iteration 1's groups; the placed buffers are sized
 * before the capture from the host-launched iterations with headroom.  If a
 * replay needs more memory or other shapes, it places nothing and the
 * execution is repeated with mode 0 (same results; gh_run_stats.graph_reruns).
 * mode 2: the same device-decided iterations enqueued on the run's stream
 * without a capture.  mode 0: every iteration launched from the host, with
 * three host round trips each (per-stage timings in gh_run_stats).  Runs that
 * tune shapes (gh_run_set_tuning 1), record training rows or have haploid
 * targets launch every iteration from the host.  Results are identical in all
 * modes. */
int gh_run_set_graph(gh_run *r, int32_t mode);

/* Per-target ploidy (1 or 2; default 2), before the first execution: a haploid
 * target (genotype with ploidy 1, caller_algorithm.cpp:53-84) takes one
 * forward-backward pass per iteration with the marginal haplotype
 * likelihoods of its two GL values (gl[t][l][0..1]), walks one haplotype in
 * the PBWT selection, is not phased, and stores GP from its one posterior
 * (genotype.cpp:127-166); its gp0 / gp1 are the two genotype probabilities,
 * h0 the allele of gp1 > gp0.  The learned selector needs diploid targets. */
int gh_run_set_ploidy(gh_run *r, const int32_t *ploidy);

/* Output fields as genotype_writer::writeGenotypes computes them
 * (phase/src/io/genotype_writer.cpp:109-190), bit for bit, for T diploid
 * targets at L sites from the finalized run outputs gp0, gp1, hds [T][L] and
 * the panel's allele counts cref, calt [L]:
 *   ds [T][L]   = round(gp1 + 2 gp2, 1e-3)
 *   gp [T][L][3] floored to 1e-3, then the largest remainders raised by 1e-3
 *                 until the three sum to >= 0.9999
 *   gt [T][L][2] phased alleles (het: the alt allele on haplotype hds)
 *   info, af, raf [L]: INFO score (rounded to 1e-3), alt frequency of the
 *                 targets' dosages, alt frequency of the panel. */
int gh_output_fields(int32_t T, int32_t L, const float *gp0, const float *gp1, const uint8_t *hds,
                     const uint32_t *cref, const uint32_t *calt, float *ds, float *gp, uint8_t *gt, float *info,
                     float *af, float *raf);
/* The same for targets of ploidy[t] 1 or 2 (genotype_writer.cpp:138-190): a
 * haploid target gets DS = gp1, GP (gp0, gp1) with the third value NaN (the
 * writer's vector end), gt (gp0 < gp1, 255 = vector end); AF divides by the
 * number of target haplotypes; INFO is computed only when every target is
 * diploid, else 1. */
int gh_output_fields_ploidy(int32_t T, int32_t L, const int32_t *ploidy, const float *gp0, const float *gp1,
                            const uint8_t *hds, const uint32_t *cref, const uint32_t *calt, float *ds, float *gp,
                            uint8_t *gt, float *info, float *af, float *raf);

/* ---------------------------------------------------------------------------
 * Learned state selector.  No counterpart in the reference: it is an addition
 * beside conditioning_set::compactSelection's Kpbwt branch
 * (phase/src/containers/conditioning_set.cpp:96-104), which stays the fallback.
 *
 * A 32 -> 128 -> 128 -> 1 ReLU network scores every candidate of a target's
 * PBWT pool (the Kpbwt states the plain method would use) from 32 features of
 * the candidate against the target's current haplotypes and reads, predicting
 * y = ln(K * posterior mass + 1e-3).  The selector keeps the k highest-scoring
 * candidates; when their predicted share of the pool's mass is below min_mass
 * the whole pool is kept.  BF16 operands on the matrix cores, FP32
 * accumulation, FP32 master weights, Adam.
 */
#define GH_MLP_IN 32
#define GH_MLP_HIDDEN 128
#define GH_MLP_NPARAM 20865   /* W1[32][128], b1[128], W2[128][128], b2[128], W3[128], b3 */
typedef struct gh_mlp gh_mlp;
/* He-initialised network on the current device. */
int gh_mlp_create(uint32_t seed, gh_mlp **out);
int gh_mlp_destroy(gh_mlp *m);
/* Parameters as one FP32 block of GH_MLP_NPARAM (layout above); set also
 * resets the optimiser. */
int gh_mlp_get_params(gh_mlp *m, float *w);
int gh_mlp_set_params(gh_mlp *m, const float *w);
/* y[n] = net(X[n][32]) (host arrays). */
int gh_mlp_forward(gh_mlp *m, int64_t n, const float *X, float *y);
/* Gradient of sum_r w_r (net(X_r) - t_r)^2 / sum_r w_r (w may be NULL: all
 * ones) into grad[GH_MLP_NPARAM]; the loss into *loss. */
int gh_mlp_backward(gh_mlp *m, int32_t n, const float *X, const float *t, const float *w, float *grad, float *loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) on that gradient. */
int gh_mlp_step(gh_mlp *m, int32_t n, const float *X, const float *t, const float *w, float lr, float *loss);
/* Trains on the rows a run recorded (gh_run_collect), on the device:
 * shuffled batches, learning rate decaying linearly from lr to lr/10;
 * loss_per_epoch optional [epochs]. */
int gh_mlp_train_run(gh_mlp *m, gh_run *r, int32_t epochs, int32_t batch, float lr, uint32_t seed,
                     float *loss_per_epoch);

/* Before gh_run_execute: record a training row (features, label, weight) for
 * every PBWT pool candidate of every target at iterations >= from_iter (>= 1);
 * the label is ln(K * m + 1e-3) with m the candidate's posterior mass from both
 * haplotypes' forward-backward, averaged over the polymorphic sites of its
 * segment (the 2048 common sites around the stored site that listed it). */
int gh_run_collect(gh_run *r, int32_t from_iter);
/* Copies the recorded rows (NULL buffers: only *n_rows). */
int gh_run_collected(gh_run *r, int64_t *n_rows, float *X, float *y, float *w);
/* Before gh_run_execute: use the learned selector from iteration 1 on.  Target
 * t keeps k_learn[t] states of its pool, or the whole pool when the predicted
 * mass share is below min_mass (0.95).  mode 0 keeps the k best scores of the
 * pool; mode 1 ranks candidates within the PBWT group of the stored site that
 * listed them and serves the groups round robin (rank 1 of every group, then
 * rank 2, ...), which keeps the plain method's coverage along the chunk.
 * m = NULL switches it off.  The network must not be trained while the run
 * executes. */
int gh_run_set_selector(gh_run *r, gh_mlp *m, const int32_t *k_learn, float min_mass, int32_t mode);
/* The learned ranking with a PBWT-order prior (no reference counterpart): a
 * candidate's rank score is its predicted log mass less weight x (its position
 * in the pool / pool size).  The pool is in the PBWT's insertion order
 * (conditioning_set.cpp:96-104: rank by rank, newest group first), so the plain
 * method's selection of k states is the pool's first k: weight 0 ranks by the
 * network alone, a large weight gives back the plain selection. */
int gh_run_set_selector_prior(gh_run *r, float weight);

/* The selector's two steps as batch calls (what a run does between the PBWT
 * compaction and the conditioning sets).  pool [T][kstride]: candidates in PBWT
 * insertion order, each id | rank << 20 | (listed by H1's walk) << 26 |
 * (upper neighbour) << 27, K[t] of them; pool_site [T][kstride]: the common
 * site whose PBWT record listed each candidate.  H [2T][L], gl [T][L][3],
 * flat [T][L].  Features X [T][kstride][32] (rows >= K[t] are zero). */
int gh_select_features(gh_panel *p, int32_t T, const uint8_t *H, const uint8_t *gl, const uint8_t *flat,
                       const int32_t *pool, const int32_t *pool_site, const int32_t *K, int32_t kstride,
                       int32_t depth, float *X);
/* Top k_learn[t] of the pool by score (ties: earlier insertion first); states
 * [T][kstride] holds the sorted pool on input and the sorted selection on
 * output, K_out[t] its size; when k_learn[t] >= K[t] or the kept candidates'
 * share of sum(max(exp(score) - 1e-3, 0)) is below min_mass, the pool is left
 * as it is.  coverage [T] optional (0 where nothing was dropped). */
int gh_select_topk(int32_t T, const float *scores, const int32_t *pool, const int32_t *K, const int32_t *k_learn,
                   int32_t kstride, float min_mass, int32_t *states, int32_t *K_out, float *coverage);

/* ---------------------------------------------------------------------------
 * Concordance: the genotype metrics of the concordance tool's call_set
 * (concordance/src/containers/call_set_reading.cpp:476-760 per site and
 * sample, call_set_header.h:162-259 for the truth / estimate rules), for
 * n_sites biallelic records of n_samples samples of one ploidy (1 or 2).
 *
 * Per site: the truth and estimate of every sample are classified as the
 * reference does (getTruth with the probability threshold min_val_gl and depth
 * threshold min_val_dp, getMostLikely with min_tar_gp; alleles flipped at
 * af > 0.5 unless use_alt_af), and every validated genotype is counted into
 * its sample, its frequency bin or site group, and its calibration bin.
 *
 * A record of counts holds, for truth t and estimate e in 0..ploidy:
 *   [3 t + e]   genotypes with truth t estimated as e (a 3x3 table; errors of
 *               truth t = its row minus the diagonal),
 *   [9 + t]     genotypes with a usable truth t (#val_gt columns),
 *   [12]        genotypes with a usable truth whose estimate failed min_tar_gp.
 * A record of sums holds, over the estimated genotypes: [0] sum DS, [1] sum
 * DS^2, [2] sum DS * truth, [3] min DS, [4] max DS (with the counts they give
 * the squared correlations the reference's stats2D computes; gh_conc_r2).
 * Records come per type: index 0 SNPs, 1 indels (the reference's "all" rows
 * are their sums).
 */
#define GH_CONC_NCNT 13
#define GH_CONC_NSUM 5
#define GH_CONC_NCAL 100   /* calibration bins: trunc(max GP * 99) (N_BIN_CAL, call_set_header.h:30) */
#define GH_CONC_NREJ 11
typedef struct {
    int32_t ploidy;          /* 1 or 2 for every sample */
    int32_t n_bins;          /* > 0: frequency bins, edges bins[0..n_bins]; bin b holds bins[b] < maf <= bins[b+1] */
    const double *bins;
    int32_t n_groups;        /* n_bins == 0: site groups; site_group[s] in [0, n_groups) or -1 (site not used) */
    const int32_t *site_group;
    float min_val_gl;        /* T: a truth class needs probability >= T */
    int32_t min_val_dp;      /* D: truth depth threshold; 0: depth not used */
    float min_tar_gp;        /* estimates whose largest GP is below this are filtered */
    int32_t use_alt_af;      /* 1: no allele flip, maf = af */
    int32_t truth_from_gt;   /* 1: truth from truth_gt (--gt-val), else from truth_pl */
    int32_t esti_from_gt;    /* 1: estimates from esti_gt alone (--gt-tar), else from esti_gp and esti_ds */
} gh_conc_params;

typedef struct {
    int32_t n_sites, n_samples;
    const float *af;         /* [S] allele frequency of the site (AF, or AC/AN); NaN: none, the site is skipped */
    const uint8_t *type;     /* [S] 1 SNP, 2 indel, anything else: skipped */
    const int32_t *truth_pl; /* [S][N][ploidy + 1] PL; negative (htslib's missing / vector end): missing */
    const int8_t *truth_gt;  /* [S][N] alt allele count of the truth GT, -1 missing */
    const int32_t *truth_dp; /* [S][N] DP (htslib's missing value INT32_MIN allowed), NULL when min_val_dp == 0 */
    const float *esti_gp;    /* [S][N][ploidy + 1] GP (NaN: missing) */
    const float *esti_ds;    /* [S][N] DS */
    const int8_t *esti_gt;   /* [S][N] alt allele count of the estimated GT, -1 missing; optional with GP input */
} gh_conc_input;

typedef struct {
    uint64_t *spl_cnt;   /* [N][2][GH_CONC_NCNT] per sample and type */
    double *spl_sum;     /* [N][2][GH_CONC_NSUM] */
    uint64_t *grp_cnt;   /* [G][2][GH_CONC_NCNT] per frequency bin / site group (G = n_bins or n_groups) */
    double *grp_sum;     /* [G][2][GH_CONC_NSUM + 1]: + sum of the site maf over the estimated genotypes */
    uint64_t *cal_cnt;   /* [GH_CONC_NCAL][2][9] truth x estimate per calibration bin */
    uint64_t *rejected;  /* [GH_CONC_NREJ]: truth codes -1..-5, estimate codes -1..-4 (of usable truths),
                            sites without AF or of another type, sites in no bin */
    uint32_t *site_cnt;  /* [S][GH_CONC_NCNT] per site, optional (NULL) */
    double *site_sum;    /* [S][GH_CONC_NSUM] per site (r2 per variant), optional */
    int32_t *site_tag;   /* [S] 2 * group + type index, -1 for a skipped site, optional */
} gh_conc_output;

int gh_concordance(const gh_conc_params *prm, const gh_conc_input *in, gh_conc_output *out);

/* Squared correlation of n (x, y) pairs as stats2D::corrXY gives it
 * (concordance/src/utils/basic_stats.h:110-115): 0 for n = 0, NaN for n = 1,
 * 0 when x or y is constant, else (n sxy - sx sy)^2 / ((n sxx - sx^2)(n syy - sy^2)). */
double gh_conc_r2(uint64_t n, double sx, double sxx, double sxy, double sy, double syy, int32_t x_constant,
                  int32_t y_constant);

/* ---------------------------------------------------------------------------
 * Ligation of overlapping phased chunks (ligate/src/ligater/ligater_algorithm.cpp).
 * GT values are htslib's (bcf_get_genotypes): (allele + 1) << 1 | phased,
 * 0 missing, INT32_MIN + 1 vector end (haploid), two per sample.
 *
 * The overlap of chunk f and f + 1 (same contig, f + 1 starting at or before
 * f's last site) is scanned as scan_overlap / update_distances do (:108-127,
 * :160-216): over the records present in both (same contig, position and
 * allele key) and biallelic in f, each sample's phased heterozygous genotypes
 * count as matches or mismatches against chunk f's phase as already swapped;
 * chunk f + 1 is swapped for the samples with fewer matches than mismatches.
 * A chunk without an overlap keeps the previous chunk's swaps (the reference's
 * swap state carries over).  The first half of the overlap's common records is
 * emitted from chunk f, the rest from f + 1 (:434-444); records in one chunk
 * only are emitted from it.  phase_update (:90-106) is applied in place to
 * every row of a swapped chunk.
 * Outputs: the emission order (out_chunk, out_site), swap [n_chunks][N],
 * and per overlap (n_chunks - 1 entries) its statistics. */
typedef struct {
    int32_t n_sites;
    const int32_t *contig;    /* [n_sites] contig id */
    const int32_t *pos;       /* [n_sites] position, ascending within a contig */
    const uint64_t *key;      /* [n_sites] allele key (records match on contig, position and key) */
    const uint8_t *biallelic; /* [n_sites] or NULL (all biallelic) */
    int32_t *gt;              /* [n_sites][N][2], updated in place */
} gh_lig_chunk;

typedef struct {
    int32_t n_common;      /* records present in both chunks (0: no overlap) */
    int32_t n_scanned;     /* of them biallelic in the first chunk (L_isec) */
    int32_t n_swapped;     /* samples whose phase was swapped */
    float mean_hets;       /* mean over samples of the phased hets compared (Avg #hets) */
    float mean_phaseq;     /* mean phase quality (Avg phaseQ) */
} gh_lig_overlap;

int gh_ligate(int32_t n_chunks, int32_t n_samples, const gh_lig_chunk *chunks, int64_t cap, int32_t *out_chunk,
              int32_t *out_site, int64_t *n_out, uint8_t *swap, gh_lig_overlap *ov);

/* The overlap scan alone: n_sites aligned records of the two chunks
 * (gt_a [n][N][2], gt_b [n][N][2]) and chunk A's swaps; per sample the match
 * and mismatch counts, the swap of chunk B and the phase quality
 * 99 (0.7 + f ln f + (1 - f) ln(1 - f)) / 0.7, f = matches / (matches + mismatches)
 * (99 when either count is 0). */
int gh_ligate_scan(int32_t n_sites, int32_t n_samples, const int32_t *gt_a, const int32_t *gt_b,
                   const uint8_t *swap_a, int32_t *nmatch, int32_t *nmism, uint8_t *swap_b, float *phaseq);

/* ---------------------------------------------------------------------------
 * VCF / BCF input and output (host code; zlib for BGZF).  Reads VCF text
 * (plain or BGZF-compressed) and BCF 2.2; values are handed out as htslib's
 * bcf_get_format_* / bcf_get_info_* / bcf_get_genotypes do (per-sample
 * vectors padded with the vector-end value, htslib's missing sentinels,
 * GT = (allele + 1) << 1 | phased).
 */
typedef struct gh_vcf_reader gh_vcf_reader;
typedef struct {
    const char *chrom;   /* valid until the next gh_vcf_next */
    int32_t pos;         /* 1-based */
    const char *id;
    int32_t n_allele;
    const char *ref, *alt;   /* first ALT ("." when none) */
    int32_t is_snp;      /* biallelic, one base each */
} gh_vcf_record;
int gh_vcf_open(const char *path, gh_vcf_reader **out);
int gh_vcf_close(gh_vcf_reader *r);
/* Sample names, NUL-separated into names (cap bytes); names NULL: count only. */
int gh_vcf_samples(gh_vcf_reader *r, int32_t *n_samples, char *names, int64_t cap);
/* Next record: 1, 0 at the end of the file, negative on error. */
int gh_vcf_next(gh_vcf_reader *r, gh_vcf_record *rec);
/* A FORMAT / INFO field of the current record into iv (Integer) or fv (Float):
 * returns the number of values (n_samples x per-sample count for FORMAT; the
 * per-sample count is the return value / n_samples), 0 when absent. */
int gh_vcf_format(gh_vcf_reader *r, const char *key, int32_t *iv, float *fv, int64_t cap);
int gh_vcf_info(gh_vcf_reader *r, const char *key, int32_t *iv, float *fv, int64_t cap);

/* Target genotype likelihoods for the panel's sites
 * (genotype_reader::scanTarGenotypes + parseTarGenotypes,
 * phase/src/io/genotype_reader.cpp:352-521), for diploid targets: reads the
 * remaining records of r; a biallelic record counts for the first panel site
 * with the same position, REF and ALT (SNPs only unless use_indels), with
 * FORMAT/PL, or FORMAT/GL as lroundf(-10 GL), capped at 255; records whose
 * field does not hold 3 values per sample are skipped, as are samples with a
 * missing value.  gl [T][L][3] (0 where nothing was read), flat [T][L] (0
 * where the sample's three values differ).  A negative PL or positive GL is
 * an error (GH_ERR_IO), as in the reference. */
typedef struct {
    int32_t n_sites;
    const int32_t *pos;            /* 1-based */
    const char *const *ref, *const *alt;
    int32_t use_gl;                /* 1: FORMAT/GL, 0: FORMAT/PL */
    int32_t use_indels;
} gh_gl_query;
int gh_vcf_read_gls(gh_vcf_reader *r, const gh_gl_query *q, uint8_t *gl, uint8_t *flat, int64_t *n_matched);

/* The phase tool's output file (genotype_writer::writeGenotypes,
 * phase/src/io/genotype_writer.cpp:43-190): the writer's header lines (with
 * ##NMAIN, ##FPLOIDY), per site ID, REF, ALT, INFO/RAF, AF and INFO, and per
 * sample GT (both alleles phased), DS and GP, from gh_output_fields' arrays
 * (gt [T][L][2], ds [T][L], gp [T][L][3], info / af / raf [L]).  BCF 2.2 in
 * BGZF blocks for a ".bcf" path, BGZF-compressed VCF for ".gz", else VCF. */
typedef struct {
    const char *contig;
    int32_t n_samples;
    const char *const *sample_names;
    int32_t n_sites;
    const int32_t *pos;
    const char *const *id;         /* NULL: "." */
    const char *const *ref, *const *alt;
    const uint8_t *gt;
    const float *ds, *gp, *info, *af, *raf;
    int32_t n_main, fploidy;
    const char *source, *file_date;
    const int32_t *ploidy;         /* [T] 1 or 2; NULL: every sample diploid.  Haploid: one allele, two GP values */
} gh_vcf_out;
int gh_write_genotypes(const char *path, const gh_vcf_out *d);

/* ---------------------------------------------------------------------------
 * The split_reference binary panel: the boost binary archive of
 * ref_haplotype_set and variant_map that split_reference writes
 * (split_reference/src/caller/caller_initialise.cpp:86-89) and the phase tool
 * reads (read_binary_reference_panel, phase/src/caller/caller_initialise.cpp:
 * 178-246).  gh_bin_open decodes it (a bad signature, another platform's
 * native format or an archive whose counts disagree is GH_ERR_IO, where the
 * reference stops with its version / signature errors); gh_bin_desc fills a
 * gh_panel_desc for gh_panel_create (arrays owned by the handle: the panel's
 * bit matrix as site-major words, ShapRef transposed into per-site carrier
 * lists as transposeRareRef does, the stored sparse-PBWT stream and
 * A_small_idx, and the sparse-MAF threshold the stored common flags imply);
 * gh_bin_write writes a panel in the same layout. */
typedef struct gh_bin_panel gh_bin_panel;
typedef struct {
    const char *chrid;
    int32_t input_start, input_stop, output_start, output_stop;
    const char *input_gregion, *output_gregion;
    int32_t n_sites, n_haps, n_common, n_common_hq;
    int32_t library_version;   /* boost archive library version */
    float sparse_maf;
} gh_bin_panel_info;
typedef struct {
    const char *chrid;
    int32_t input_start, input_stop, output_start, output_stop;
    const char *input_gregion, *output_gregion;
    const int32_t *pos;              /* 1-based */
    const char *const *id, *const *ref, *const *alt;
    const uint8_t *type;             /* htslib variant type per site; NULL: SNP / indel from the alleles */
    int32_t library_version;         /* 0: 18 */
} gh_bin_meta;
int gh_bin_open(const char *path, gh_bin_panel **out);
int gh_bin_close(gh_bin_panel *b);
int gh_bin_info(const gh_bin_panel *b, gh_bin_panel_info *info);
int gh_bin_desc(const gh_bin_panel *b, gh_panel_desc *d);
int gh_bin_sites(const gh_bin_panel *b, const int32_t **pos, const char *const **id, const char *const **ref,
                 const char *const **alt, const uint8_t **type);
int gh_bin_write(const char *path, const gh_panel_desc *d, const gh_bin_meta *m);

/* ---------------------------------------------------------------------------
 * Genotype likelihoods from sequencing reads: genotype_bam_caller with the
 * "standard" model (phase/src/io/genotype_bam_caller.cpp:225-488), for one
 * sample's coordinate-sorted BAM file over one contig.  Reads pass read_bam's
 * filters (:30-59); at every SNP site of the panel inside [beg, end] the
 * bases of the reads' M / = / X operations with base quality >= min_baseq
 * give the site's log10 likelihoods (at most max_depth reads: the lowest
 * keys quality << 5 | strand << 4 | base, as the reference's sort leaves
 * them), quantised as the reference does: GL[j] = min(int(-10 (llk[j] - i)),
 * 255), i the index of the largest llk.  Indel sites stay flat (the
 * reference's default).  Outputs: gl [L][ploidy + 1], flat [L], depth [L]. */
typedef struct {
    int32_t min_mapq;           /* --mapq 10 */
    int32_t min_baseq;          /* --baseq 10 */
    int32_t max_depth;          /* --max-depth 40 */
    int32_t fflag;              /* excluded flags: unmapped | secondary | QC fail | duplicate */
    int32_t keep_orphan;        /* --keep-orphan-reads */
    int32_t check_orientation;  /* 1 unless --ignore-orientation */
    int32_t check_proper_pair;  /* --check-proper-pairing */
} gh_bam_params;
typedef struct {
    const char *contig;
    int32_t beg, end;           /* 1-based, inclusive; <= 0: unbounded */
    int32_t n_sites;
    const int32_t *pos;         /* 1-based, ascending */
    const char *const *ref, *const *alt;
} gh_bam_region;
typedef struct {
    int64_t n_records;          /* BAM records read */
    int64_t n_kept;             /* kept reads aligned inside the region */
    int64_t sites_over_max_depth;
} gh_bam_stats;
void gh_bam_default_params(gh_bam_params *p);
int gh_bam_call_gls(const char *path, int32_t ploidy, const gh_bam_region *r, const gh_bam_params *p, uint8_t *gl,
                    uint8_t *flat, int32_t *depth, gh_bam_stats *stats);
/* The four error tables of the model (calling_utils, genotype_bam_caller.h:36-42):
 * t[0..255] log10(1 - e), t[256..] log10(e / 3), t[512..] log10(1/2 - e / 3),
 * t[768..] log10(e), e = 10^(-q/10). */
int gh_bam_tables(float *t);

#ifdef __cplusplus
}
#endif
#endif
