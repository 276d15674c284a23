// gh_gibbs.hip -- the device-resident Gibbs schedule of GLIMPSE2_phase
// (caller::phase_loop / phase_iteration / phase_individual,
//  phase/src/caller/caller_algorithm.cpp:53-243).
//
// Every target is processed in the same launches; per iteration:
//   selection   : INIT -> rare-allele seeded random states (host, once);
//                 BURN/MAIN -> sparse-PBWT walk + compaction (gh_pbwt.hip)
//   condset     : conditioning sets (gh_condset.hip)
//   H0          : haploid likelihoods -> forward-backward -> monomorphic closed form -> draw
//   H1          : same, conditioned on the new H0
//   phasing     : loci/segments prepass -> phasing HMM with segment draws
//   MAIN only   : GP accumulation (dense form of genotype::storeGenotypePosteriorsAndHaplotypes)
// then normalisation and best-guess genotypes (genotype::sortAndNormAndInferGenotype).
//
// Haplotypes live as bytes in H[2T][L] (row 2t = H0 of target t, 2t+1 = H1).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <numeric>
#include <atomic>
#include <mutex>
#include <string>
#include <thread>

#include "gh_device.hpp"
#include "gh_graph.hpp"
#include "gh_init.hpp"
#include "gh_internal.hpp"
#include "gh_pbwt.hpp"
#include "gh_phase.hpp"
#include "gh_select.hpp"

namespace gh {

__constant__ float c_unphred[256];

// common/src/utils/otools.h:97-98: 10^(-i/10) printed with 7 significant digits.
void unphred_table(float *t) {
    char buf[64];
    for (int i = 0; i < 256; i++) {
        std::snprintf(buf, sizeof(buf), "%.6e", std::pow(10.0, -i / 10.0));
        t[i] = std::strtof(buf, nullptr);
    }
}

// genotype::initHaplotypeLikelihoods / makeHaplotypeLikelihoods (genotype.cpp:56-116).
// A haploid target (ploidy[t] == 1) always takes initHaplotypeLikelihoods with
// its two GL values (caller_algorithm.cpp:58-62): HL = the normalised pair.
__global__ void hl_kernel(const uint8_t *__restrict__ gl, const uint8_t *__restrict__ flat,
                          const uint8_t *__restrict__ H, int L, int T, int mode, int which, float min_gl,
                          const uint8_t *__restrict__ ploidy, float *__restrict__ HL) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (l >= L) return;
    const size_t o = (size_t)t * L + l;
    const bool hap = ploidy && ploidy[t] == 1;
    float h0, h1;
    if (flat[o]) {
        h0 = 0.5f; h1 = 0.5f;
    } else if (hap) {
        const uint8_t *g = gl + o * 3;
        float t0 = c_unphred[g[0]], t1 = c_unphred[g[1]];
        const float s = (t0 + t1) + 0.0f;
        t0 /= s; t1 /= s;
        h0 = t0;
        h1 = t1;
    } else {
        const uint8_t *g = gl + o * 3;
        float t0 = c_unphred[g[0]], t1 = c_unphred[g[1]], t2 = c_unphred[g[2]];
        const float s = t0 + t1 + t2;
        t0 /= s; t1 /= s; t2 /= s;
        if (mode == 0) {
            h0 = (float)((double)t0 + 0.5 * (double)t1);
            h1 = (float)(0.5 * (double)t1 + (double)t2);
        } else {
            // condition on the other haplotype of the target
            const int c = H[(size_t)(2 * t + (which ^ 1)) * L + l];
            const float a = c ? t1 : t0, b = c ? t2 : t1;
            h0 = a / (a + b);
            h1 = b / (a + b);
        }
    }
    if (h0 < min_gl) { h0 = min_gl; h1 = 1.0f - min_gl; }
    if (h1 < min_gl) { h0 = 1.0f - min_gl; h1 = min_gl; }
    HL[2 * o] = h0;
    HL[2 * o + 1] = h1;
}

// genotype::sampleHaplotypeH0/H1 (genotype.cpp:118-125) with counter draws
// (iter_dev, when set: the iteration is read from the device, for replayed iterations)
__global__ void sample_kernel(const float *__restrict__ HP, uint8_t *__restrict__ H, int L, int which, uint32_t seed,
                              uint32_t it, const uint32_t *__restrict__ iter_dev) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (l >= L) return;
    if (iter_dev) it = *iter_dev;
    const float u = u01(philox_draw(seed, it, which ? P_H1 : P_H0, (uint32_t)t, (uint32_t)l));
    H[(size_t)(2 * t + which) * L + l] = u > HP[((size_t)t * L + l) * 2];
}

// genotype::storeGenotypePosteriorsAndHaplotypes, diploid (genotype.cpp:168-209),
// in dense form: a site not yet stored whose P(RR) >= 0.99999 contributes the
// implicit (1, 0); from its first storage on, the actual normalised GP.
// A haploid target stores the normalised HP0 pair (genotype.cpp:127-166; its
// first-storage offset is stored_cnt % 16, equal to stored_cnt for --main <= 15).
// (iter_dev, when set: a replayed iteration, which stores only from iteration first_main on)
__global__ void gp_kernel(const float *__restrict__ HP0, const float *__restrict__ HP1, int L,
                          const uint8_t *__restrict__ ploidy, float *__restrict__ gp0, float *__restrict__ gp1,
                          uint8_t *__restrict__ stored, uint8_t *__restrict__ hds,
                          const uint32_t *__restrict__ iter_dev, uint32_t first_main) {
#pragma clang fp contract(off)
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (l >= L) return;
    if (iter_dev && *iter_dev < first_main) return;
    const size_t o = (size_t)t * L + l;
    if (ploidy && ploidy[t] == 1) {
        const float p0 = HP0[2 * o], p1 = HP0[2 * o + 1];
        const float sc = 1.0f / (p0 + p1);
        const float n0 = p0 * sc, n1 = p1 * sc;
        if (!stored[o] && !(n0 < 0.99999f)) {
            gp0[o] += 1.0f;
            return;
        }
        stored[o] = 1;
        gp0[o] += n0;
        gp1[o] += n1;
        hds[o] = 0;
        return;
    }
    const float a0 = HP0[2 * o], a1 = HP0[2 * o + 1], b0 = HP1[2 * o], b1 = HP1[2 * o + 1];
    const float p0 = fminf(fmaxf(a0 * b0, 0.0f), 1.0f);
    const float p1 = fminf(fmaxf(a0 * b1 + a1 * b0, 0.0f), 1.0f);
    const float p2 = fminf(fmaxf(a1 * b1, 0.0f), 1.0f);
    const float n0 = p0 / (p0 + p1 + p2), n1 = p1 / (p0 + p1 + p2);
    if (!stored[o] && !(n0 < 0.99999f)) {
        gp0[o] += 1.0f;
        return;
    }
    stored[o] = 1;
    gp0[o] += n0;
    gp1[o] += n1;
    hds[o] = a1 < b1;
}

// genotype::sortAndNormAndInferGenotype + inferred_genotype::infer (genotype.h:30-36, genotype.cpp:211-244)
__global__ void finalize_kernel(float *__restrict__ gp0, float *__restrict__ gp1, uint8_t *__restrict__ stored,
                                uint8_t *__restrict__ hds, uint8_t *__restrict__ h0, uint8_t *__restrict__ h1,
                                int L, int cnt, const uint8_t *__restrict__ ploidy) {
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    const int t = blockIdx.y;
    if (l >= L) return;
    const size_t o = (size_t)t * L + l;
    if (!stored[o] || cnt == 0) {
        gp0[o] = 1.0f; gp1[o] = 0.0f; hds[o] = 0; h0[o] = 0; h1[o] = 0;
        return;
    }
    const float a = gp0[o] / cnt, b = gp1[o] / cnt;
    gp0[o] = a; gp1[o] = b;
    if (ploidy && ploidy[t] == 1) {   // inferred_genotype::infer_haploid (genotype.h:66-68)
        h0[o] = b > a;
        h1[o] = 0;
        hds[o] = 0;
        return;
    }
    const float c = 1.0f - b - a;
    int g = 0;
    if (a > b && a > c) g = 0;
    else if (b > a && b > c) g = 1;
    else if (c > a && c > b) g = 2;
    h0[o] = g == 2;
    h1[o] = g >= 1;
}

__global__ void set_k_kernel(CondItem *__restrict__ items, const int32_t *__restrict__ Kdev,
                             const int32_t *__restrict__ targets, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int t = targets[i];
    const int K = Kdev[t];
    items[t].K = K;
    items[t].W = (K + 63) / 64;
}

/* ------------------------------------------------------------------ driver */
// Online launch-shape tuner: while a class (set-size bucket) is being tuned,
// each iteration runs the next untimed candidate shape and records its kernel
// time; once every candidate has a time, the fastest is used from then on.
template <typename S>
struct ShapeTuner {
    int cls = -1, cur = -1, best = -1;
    std::vector<S> cand;
    std::vector<float> ms;
    S pick(int c, const std::vector<S> &cands, S dflt) {
        if (c != cls) {
            cls = c;
            cand = cands;
            ms.assign(cand.size(), -1.0f);
            best = -1;
        }
        cur = -1;
        if (best >= 0) return cand[best];
        for (size_t i = 0; i < cand.size(); i++)
            if (ms[i] < 0.0f) { cur = (int)i; return cand[i]; }
        return dflt;
    }
    void record(float t) {
        if (cur < 0) return;
        ms[cur] = t;
        cur = -1;
        int b = 0;
        for (size_t i = 0; i < ms.size(); i++) {
            if (ms[i] < 0.0f) return;
            if (ms[i] < ms[b]) b = (int)i;
        }
        best = b;
    }
};

// Grow-only reservation with a quarter of headroom, so buffers placed by
// per-iteration counts settle after the first iterations.
static inline void grow(DBuf &b, size_t bytes) {
    if (bytes > b.bytes) b.reserve(bytes + bytes / 4);
}

// A launch of at most this many items uses the latency launch shapes.
constexpr int FEW_ITEMS = 32;

struct EvPair {
    hipEvent_t a = nullptr, b = nullptr;
    void make() { GH_HIP(hipEventCreate(&a)); GH_HIP(hipEventCreate(&b)); }
    ~EvPair() { if (a) (void)hipEventDestroy(a); if (b) (void)hipEventDestroy(b); }
    float ms() { float x = 0.0f; (void)hipEventElapsedTime(&x, a, b); return x; }
};

}  // namespace gh

// Device workspace of a run's execution: everything an execution builds and
// discards (current haplotypes, likelihoods and posteriors, state sets,
// conditioning sets, checkpoints, phasing loci and messages, selection
// records, selector rows).  Buffers grow to the largest execution they serve
// and are kept; runs executed one after another can share one workspace, so
// only the executing runs hold this memory.
// workspaces whose run stream exists (executions may use a second stream each)
static std::atomic<int> g_live_workspaces{0};

struct gh_workspace {
    gh::DBuf H, HL, HP0, HP1, states, poly, pos, t, mark, hvar, alpha, loci, seg, segsum, imp, impsum, of1, dip, ev,
        evci, ins, isite, fx, fy, mass0, mass1, wlo, whi;
    // The executions' streams belong to the workspace, not the run: a batch of
    // many resident runs has as many streams in use as workspaces, so runs
    // executing together get distinct hardware queues (streams beyond
    // GPU_MAX_HW_QUEUES share queues, and a stream holding a whole run's
    // replayed iterations would block the other).  The second stream (K-class
    // batching: the other classes beside the largest, while the streams fit the
    // hardware queues) is created the first time such a pass has a second class.
    hipStream_t s = nullptr, s2 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    void streams() {
        if (s) return;
        GH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
        g_live_workspaces.fetch_add(1);
        GH_HIP(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming));
        GH_HIP(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming));
    }
    hipStream_t side() {
        if (!s2) GH_HIP(hipStreamCreateWithFlags(&s2, hipStreamNonBlocking));
        return s2;
    }
    ~gh_workspace() {
        if (ev_fork) (void)hipEventDestroy(ev_fork);
        if (ev_join) (void)hipEventDestroy(ev_join);
        if (s2) (void)hipStreamDestroy(s2);
        if (s) {
            (void)hipStreamDestroy(s);
            g_live_workspaces.fetch_sub(1);
        }
    }
};

// One prepared Gibbs run: inputs, outputs and tables resident on the device;
// the workspace is allocated by the first execution (or shared, gh_run_set_workspace).
struct gh_run {
    gh_panel *P = nullptr;
    int T = 0, L = 0, N = 0, n_iter = 0, kinit = 0, kstride = 0, maxW = 0, npt = 0;
    int pstride = 1;   // selector rows per PBWT target: the largest Kpbwt
    gh_params prm;
    float ee_i = 0, ed_i = 0;
    float unph[256];
    std::vector<int> kp, kmax, hap_cls, pbwt_targets;
    std::vector<uint8_t> ploidy;         // [T] 1 or 2 (genotype::ploidy)
    bool any_haploid = false;
    std::vector<std::vector<int>> cls_grp;
    std::vector<float> cls_mod;
    std::vector<gh::CondItem> items;
    std::vector<gh::FbItem> fb;
    std::vector<gh::PhItem> ph;
    std::vector<uint8_t> stored_h;
    std::vector<int32_t> start_h, nstates;   // nstates [T][n_iter]
    gh::DBuf d_gl, d_flat, d_K, d_items, d_fb, d_gp0, d_gp1, d_stored, d_hds, d_h0, d_h1, d_ph, d_status, d_pst,
        d_pstart, d_pcls, d_pmod, d_evcnt, d_pkp, d_psh, d_prow, d_ptg, d_hrow, d_ploidy, d_pone, d_dord;
    std::vector<int32_t> dord;   // diploid targets in launch order (H1 forward-backward, phasing)
    gh_workspace own;
    gh_workspace *W = &own;
    hipStream_t stream = nullptr;   // the run's own stream: uploads, fetches (executions use the workspace's)
    gh::PbwtRun R;
    gh_run_stats st;
    bool done = false;
    // learned selection / training-row collection
    bool tune = false;
    bool latency = false;   // latency default shapes (a run on the critical path)
    // K-class batching: the items of the largest set-size class run in one
    // launch on the execution stream, the others in a launch sized for their
    // own class on the workspace's second stream beside it
    gh::DBuf d_ford, d_pord;
    std::vector<int32_t> ford, pord;
    gh::ShapeTuner<gh::FbPick> fb_tuner;
    gh::ShapeTuner<gh::PhaseShape> ph_tuner;
    gh_mlp *mlp = nullptr;
    float min_mass = 0.95f;
    float sel_prior = 0.0f;    // PBWT-order prior of the learned ranking
    int sel_mode = 1;          // 0: global top-k, 1: top-k per PBWT group, groups served round robin
    int collect_from = -1;
    int64_t col_rows = 0;
    bool masks_ready = false;
    gh::DBuf d_klearn, d_cov, d_cnt, d_tmask, d_cx, d_cy, d_cw, d_sgrp, d_tcls;
    // replayed iterations (gh_graph.hpp)
    int graph_mode = 1;        // 1: iterations from 2 on replay a captured graph; 0: all launched from the host
    gh::GraphCaps need{};      // largest placement needs of the iterations launched from the host
    int gk_fb[2] = {-1, -1}, gk_ph[2] = {-1, -1};   // launch keys of the last host-launched iteration
    int graph_fail = 0;        // GO_* bits of the last execution whose replays placed nothing
    gh::DBuf d_ctl, d_nst, d_gptr, d_gend;
    gh::DBuf d_unph, d_fullp;  // the unphred table; full-panel targets (Kpbwt >= N)
    int n_full = 0;
    ~gh_run() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace gh {

// Walked haplotype rows of the PBWT targets: a diploid target walks H0 and H1;
// a haploid one walks H0 only (its second walker repeats H0 and its lists are
// not used), haplotype_set.cpp:263-270.
static void upload_pbwt_rows(gh_run *G, hipStream_t s) {
    const int npt = G->npt;
    if (npt == 0) return;
    std::vector<int32_t> hrow;
    std::vector<uint8_t> one(npt);
    for (int i = 0; i < npt; i++) {
        const int t = G->pbwt_targets[i];
        const bool hap = G->ploidy[t] == 1;
        hrow.push_back(2 * t);
        hrow.push_back(hap ? 2 * t : 2 * t + 1);
        one[i] = hap;
    }
    G->R.hap_row = dalloc_copy(G->d_hrow, hrow.data(), hrow.size(), s);
    G->R.sample_one = G->any_haploid ? dalloc_copy(G->d_pone, one.data(), one.size(), s) : nullptr;
}

void run_set_ploidy(gh_run *G, const int32_t *ploidy) {
    GH_CHECK(G && ploidy, GH_ERR_ARG, "gh_run_set_ploidy: null argument");
    GH_CHECK(!G->done, GH_ERR_ARG, "gh_run_set_ploidy: set the ploidy before the first execution");
    bool any = false;
    for (int t = 0; t < G->T; t++) {
        GH_CHECK(ploidy[t] == 1 || ploidy[t] == 2, GH_ERR_ARG, "gh_run_set_ploidy: ploidy must be 1 or 2");
        G->ploidy[t] = (uint8_t)ploidy[t];
        any = any || ploidy[t] == 1;
    }
    G->any_haploid = any;
    dalloc_copy(G->d_ploidy, G->ploidy.data(), (size_t)G->T, G->stream);
    upload_pbwt_rows(G, G->stream);
    GH_HIP(hipStreamSynchronize(G->stream));
}

void run_create(gh_panel *P, int T, const uint8_t *gl_h, const uint8_t *flat_h, const gh_params *prm,
                const int32_t *kpt, gh_run **out) {
    GH_CHECK(P && T > 0 && gl_h && flat_h && prm && out, GH_ERR_ARG, "gh_run_create: bad arguments");
    GH_CHECK(prm->burnin >= 0 && prm->main_iter >= 0 && prm->main_iter <= 15, GH_ERR_ARG,
             "gh_run_create: --main must be in 0..15 (caller_parameters.cpp:139-140)");
    GH_CHECK(prm->pbwt_depth > 0 && prm->pbwt_depth <= 64, GH_ERR_ARG, "gh_run_create: pbwt depth must be in 1..64");
    auto *G = new gh_run();
    try {
        GH_HIP(hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking));
        hipStream_t s = G->stream;
        G->P = P; G->T = T; G->L = P->L; G->N = P->N; G->prm = *prm;
        const int L = P->L, N = P->N;
        G->kinit = std::min(prm->kinit, N);
        const float err_imp = std::min(std::max(prm->err_imp, 1e-12f), 1e-3f);   // caller_initialise.cpp:152
        G->ed_i = err_imp;
        G->ee_i = (float)(1.0 - (double)err_imp);
        G->n_iter = 1 + prm->burnin + prm->main_iter;
        unphred_table(G->unph);
        G->ploidy.assign(T, 2);
        // per-target K (caller_initialise.cpp:150) and PBWT classes by effective modulo
        G->kp.resize(T);
        for (int t = 0; t < T; t++) G->kp[t] = std::min(kpt ? kpt[t] : prm->kpbwt, N);
        // Class ids follow ascending effective modulo: the id keys the class's
        // stored-site draws, so it must not depend on the order of the targets.
        std::map<float, std::vector<int>> grp_of_mod;
        std::vector<float> em_t(T, 0.0f);
        G->hap_cls.assign(2 * T, 0);
        for (int t = 0; t < T; t++) {
            GH_CHECK(G->kp[t] > 0, GH_ERR_STATE, "target " + std::to_string(t) + " has Kpbwt = 0 and no state list");
            GH_CHECK(G->kinit > 0 || G->kp[t] >= N, GH_ERR_STATE,
                     "target " + std::to_string(t) + ": Kinit = 0 leaves no INIT states");
            if (G->kp[t] >= N) continue;
            GH_CHECK(P->pbwt, GH_ERR_ARG, "Kpbwt < N needs a panel created with its sparse-PBWT stream");
            std::vector<int> grp;
            em_t[t] = pbwt_groups(P, prm->pbwt_depth, prm->pbwt_modulo_cm, G->kp[t], grp);
            grp_of_mod[em_t[t]] = grp;
        }
        std::map<float, int> cls_of_mod;
        for (auto &g : grp_of_mod) {
            cls_of_mod[g.first] = (int)G->cls_mod.size();
            G->cls_mod.push_back(g.first);
            G->cls_grp.push_back(g.second);
        }
        for (int t = 0; t < T; t++) {
            if (G->kp[t] >= N) continue;
            const int c = cls_of_mod[em_t[t]];
            G->hap_cls[2 * t] = G->hap_cls[2 * t + 1] = c;
            G->pbwt_targets.push_back(t);
        }
        for (int t : G->pbwt_targets) G->pstride = std::max(G->pstride, G->kp[t]);
        int kmax_t = 0;
        G->kmax.resize(T);
        for (int t = 0; t < T; t++) {
            G->kmax[t] = std::max(G->kinit, G->kp[t]);
            kmax_t = std::max(kmax_t, G->kmax[t]);
        }
        GH_CHECK(kmax_t <= 4096, GH_ERR_ARG, "more than 4096 conditioning states is not supported");
        G->kstride = kmax_t;
        // ---------------- device workspace
        dalloc_copy(G->d_gl, gl_h, (size_t)T * L * 3, s);
        dalloc_copy(G->d_flat, flat_h, (size_t)T * L, s);
        G->d_K.reserve(sizeof(int32_t) * T);
        G->items.resize(T);
        G->fb.resize(T);
        for (int t = 0; t < T; t++) G->maxW = std::max(G->maxW, (G->kmax[t] + 63) / 64);
        // Hvar rows, checkpoints and phasing loci are placed per iteration by
        // each target's polymorphic-site count (offsets set after the index step)
        for (int t = 0; t < T; t++) {
            const int Wt = (G->kmax[t] + 63) / 64;
            G->items[t] = CondItem{G->kmax[t], Wt, 0, 0, (int64_t)t * G->kstride, (int64_t)t * L, 0};
            G->fb[t] = FbItem{0, 0, 0, t, 0, (int64_t)t * 2 * L, (int64_t)t * L};
        }
        dalloc_copy(G->d_items, G->items.data(), T, s);
        dalloc_copy(G->d_fb, G->fb.data(), T, s);
        G->d_gp0.reserve(sizeof(float) * (size_t)T * L);
        G->d_gp1.reserve(sizeof(float) * (size_t)T * L);
        G->d_stored.reserve((size_t)T * L);
        G->d_hds.reserve((size_t)T * L);
        G->d_h0.reserve((size_t)T * L);
        G->d_h1.reserve((size_t)T * L);
        G->d_status.reserve(sizeof(int32_t) * 2);
        G->ph.resize(T);
        for (int t = 0; t < T; t++) {
            std::memset(&G->ph[t], 0, sizeof(PhItem));
            G->ph[t].cond = t;
            G->ph[t].tgt = t;
            G->ph[t].h0_off = (int64_t)(2 * t) * L;
            G->ph[t].h1_off = (int64_t)(2 * t + 1) * L;
            G->ph[t].flat_off = (int64_t)t * L;
            G->ph[t].locus_off = 0;   // placed per iteration by Lp
        }
        PbwtRun &R = G->R;
        std::memset(&R, 0, sizeof(R));
        G->npt = (int)G->pbwt_targets.size();
        const int npt = G->npt;
        if (npt > 0) {
            int ngrp = 0;
            for (auto &g : G->cls_grp) ngrp = std::max(ngrp, (int)g.size());
            const float minmod = *std::min_element(G->cls_mod.begin(), G->cls_mod.end());
            R.n_haps = 2 * npt; R.n_samples = npt; R.depth = prm->pbwt_depth;
            R.cap = pbwt_event_cap(P, minmod, ngrp);
            std::vector<int32_t> srow, pc(2 * npt), kpv(npt), sh(npt);
            for (int i = 0; i < npt; i++) {
                const int t = G->pbwt_targets[i];
                srow.push_back(t);
                pc[2 * i] = pc[2 * i + 1] = G->hap_cls[2 * t];
                kpv[i] = G->kp[t];
                sh[i] = 2 * i;
            }
            R.hap_class = dalloc_copy(G->d_pcls, pc.data(), pc.size(), s);
            R.modulo = dalloc_copy(G->d_pmod, G->cls_mod.data(), G->cls_mod.size(), s);
            R.kpbwt = dalloc_copy(G->d_pkp, kpv.data(), kpv.size(), s);
            R.sample_hap = dalloc_copy(G->d_psh, sh.data(), sh.size(), s);
            R.sample_row = dalloc_copy(G->d_prow, srow.data(), srow.size(), s);
            upload_pbwt_rows(G, s);
            R.ev_cnt = (int32_t *)G->d_evcnt.reserve(sizeof(int32_t) * 2 * npt);
        }
        dalloc_copy(G->d_ptg, G->pbwt_targets.data(), G->pbwt_targets.size(), s);
        G->stored_h.assign(std::max<size_t>(G->cls_mod.size(), 1) * std::max(P->Lhq, 1), 0);
        G->start_h.assign(std::max(2 * npt, 1), 0);
        G->d_pst.reserve(G->stored_h.size());
        G->d_pstart.reserve(sizeof(int32_t) * G->start_h.size());
        G->nstates.assign((size_t)T * G->n_iter, 0);
        // per-iteration tables, allocated once here (an execution frees nothing:
        // a free waits for the whole device)
        G->d_ph.reserve(sizeof(PhItem) * T);
        dalloc_copy(G->d_unph, G->unph, 256, s);
        {
            std::vector<int32_t> fullp;
            for (int t = 0; t < T; t++)
                if (G->kp[t] >= N) fullp.push_back(t);
            G->n_full = (int)fullp.size();
            dalloc_copy(G->d_fullp, fullp.data(), fullp.size(), s);
        }
        G->d_ford.reserve(sizeof(int32_t) * 2 * (size_t)T);
        G->d_pord.reserve(sizeof(int32_t) * 2 * (size_t)T);
        G->d_dord.reserve(sizeof(int32_t) * (size_t)T);
        G->d_ctl.reserve(sizeof(GraphCtl));
        G->d_nst.reserve(sizeof(int32_t) * (size_t)T * G->n_iter);
        {
            std::vector<int32_t> gptr(1, 0), gend;
            for (auto &g : G->cls_grp) {
                gend.insert(gend.end(), g.begin(), g.end());
                gptr.push_back((int32_t)gend.size());
            }
            dalloc_copy(G->d_gptr, gptr.data(), gptr.size(), s);
            dalloc_copy(G->d_gend, gend.data(), gend.size(), s);
        }
        GH_HIP(hipStreamSynchronize(s));
        *out = G;
    } catch (...) {
        delete G;
        throw;
    }
}

// One execution of the schedule.  With use_graph, iterations 0 and 1 are
// launched from the host and iteration 2 is captured as a HIP graph, which is
// replayed for every later iteration (gh_graph.hpp).  Returns false when a
// replayed iteration needed more placed memory than the captured buffers hold
// (its results are then invalid and the caller executes again without the graph).
// Whether the second set-size class of a pass goes beside the first on the
// workspace's second stream: while every live workspace's two streams fit the
// process's hardware queues (GPU_MAX_HW_QUEUES, HIP's default 4); beyond that,
// streams share queues and a queue holding one run's replayed iterations blocks
// the other run's.  GH_SIDE_STREAM=0 / 1 forces either.
static bool side_stream_enabled() {
    static const int forced = [] {
        const char *e = std::getenv("GH_SIDE_STREAM");
        return e ? (std::atoi(e) != 0 ? 1 : 0) : -1;
    }();
    static const int queues = [] {
        const char *e = std::getenv("GPU_MAX_HW_QUEUES");
        const int q = e ? std::atoi(e) : 0;
        return q > 0 ? q : 4;
    }();
    if (forced >= 0) return forced == 1;
    return 2 * g_live_workspaces.load() <= queues;
}

static bool run_execute_once(gh_run *G, bool use_graph) {
    const auto wall0 = std::chrono::steady_clock::now();
    gh_panel *P = G->P;
    const int T = G->T, L = G->L, N = G->N, n_iter = G->n_iter, npt = G->npt;
    const gh_params *prm = &G->prm;
    const uint32_t seed = prm->seed;
    gh_workspace *Wk = G->W;
    Wk->streams();
    hipStream_t s = Wk->s;
    // The second set-size class of a pass (K-class batching) runs beside the
    // first on the workspace's second stream when the streams fit the hardware
    // queues, else after the first on the run's stream (runs executing together
    // then supply the concurrency: sixteen C5 runs whose kept sets fell in a
    // smaller class than the fallbacks ran 2.3 times faster so).
    const bool side_on = side_stream_enabled();
    auto fork = [&]() -> hipStream_t {
        if (!side_on) return s;
        GH_HIP(hipEventRecord(Wk->ev_fork, s));
        GH_HIP(hipStreamWaitEvent(Wk->side(), Wk->ev_fork, 0));
        return Wk->side();
    };
    auto join = [&]() {
        if (!side_on) return;
        GH_HIP(hipEventRecord(Wk->ev_join, Wk->side()));
        GH_HIP(hipStreamWaitEvent(s, Wk->ev_join, 0));
    };
    const int kst = G->kstride, prs = G->pstride;
    const bool learned = G->mlp != nullptr, collecting = G->collect_from >= 1;
    GH_CHECK(!(learned && collecting), GH_ERR_ARG, "a run either records training rows or uses the learned selector");
    GH_CHECK(!(learned || collecting) || !G->any_haploid, GH_ERR_ARG,
             "the learned selector and its training rows need diploid targets");
    // workspace of this execution (grows to the largest execution it serves)
    Wk->H.reserve((size_t)2 * T * L);
    Wk->HL.reserve(sizeof(float) * (size_t)T * 2 * L);
    Wk->HP0.reserve(sizeof(float) * (size_t)T * 2 * L);
    Wk->HP1.reserve(sizeof(float) * (size_t)T * 2 * L);
    Wk->states.reserve(sizeof(int32_t) * (size_t)T * kst);
    Wk->poly.reserve(sizeof(int32_t) * (size_t)T * L);
    Wk->pos.reserve(sizeof(int32_t) * (size_t)T * L);
    Wk->t.reserve(sizeof(float) * (size_t)T * L);
    Wk->mark.reserve((size_t)T * L);
    if (npt > 0) {
        Wk->ev.reserve(sizeof(int32_t) * (size_t)2 * npt * G->R.cap * (2 + 2 * G->R.depth));
        Wk->evci.reserve(sizeof(int32_t) * (size_t)2 * npt * G->R.cap * 2);
    }
    GH_HIP(hipMemcpyToSymbolAsync(HIP_SYMBOL(c_unphred), G->unph, sizeof(G->unph), 0, hipMemcpyHostToDevice, s));
    uint8_t *H = Wk->H.as<uint8_t>();
    float *HL = Wk->HL.as<float>(), *HP0 = Wk->HP0.as<float>(), *HP1 = Wk->HP1.as<float>();
    int32_t *states = Wk->states.as<int32_t>(), *Kdev = G->d_K.as<int32_t>();
    CondItem *items_d = G->d_items.as<CondItem>();
    FbItem *fb_d = G->d_fb.as<FbItem>();
    int32_t *poly = Wk->poly.as<int32_t>(), *pos = Wk->pos.as<int32_t>();
    float *tr = Wk->t.as<float>();
    uint8_t *mark = Wk->mark.as<uint8_t>();
    float *gp0 = G->d_gp0.as<float>(), *gp1 = G->d_gp1.as<float>();
    uint8_t *sto = G->d_stored.as<uint8_t>(), *hds = G->d_hds.as<uint8_t>();
    int32_t *status = G->d_status.as<int32_t>();
    const uint8_t *gl_d = G->d_gl.as<uint8_t>(), *flat_d = G->d_flat.as<uint8_t>();
    int32_t *ptg = G->d_ptg.as<int32_t>();
    const uint8_t *ploidy_d = G->any_haploid ? G->d_ploidy.as<uint8_t>() : nullptr;
    int n_dip = T;
    if (G->any_haploid) {
        n_dip = 0;
        for (int t = 0; t < T; t++) n_dip += G->ploidy[t] == 2;
    }
    uint8_t *pst = G->d_pst.as<uint8_t>();
    int32_t *pstart = G->d_pstart.as<int32_t>();
    if ((learned || collecting) && npt > 0) {
        GH_CHECK(prs <= TOPK_MAX, GH_ERR_ARG, "learned selection: PBWT pools of at most 2048 states");
        GH_CHECK(feature_lds_bytes(P) <= FEAT_LDS_MAX, GH_ERR_CAPACITY,
                 "learned selection: chunk too large for the feature kernel");
        GH_CHECK(N < (1 << 20), GH_ERR_ARG, "learned selection supports at most 2^20 panel haplotypes");
        Wk->ins.reserve(sizeof(int32_t) * (size_t)T * kst);
        if (!G->masks_ready) {
            G->d_tmask.reserve(sizeof(uint64_t) * (size_t)T * 4 * std::max(P->Wc, 1));
            target_mask_launch(P, T, gl_d, flat_d, G->d_tmask.as<uint64_t>(), s);
            G->masks_ready = true;
        }
        G->d_cnt.reserve(sizeof(int32_t) * 2);
        GH_HIP(hipMemsetAsync(G->d_cnt.p, 0, sizeof(int32_t) * 2, s));
        if (!G->d_sgrp.p) {
            // PBWT group of every HQ common site, per class; class of every target
            const int nc = (int)G->cls_grp.size();
            std::vector<int32_t> sg((size_t)std::max(nc, 1) * L, -1), tc(T, 0);
            for (int c = 0; c < nc; c++) {
                int g = 0;
                for (int l_all = 0, l_hq = 0; l_all < P->Lc; l_all++) {
                    const int k = P->common2tot[l_all];
                    if (P->flags[k] & SF_LQ) continue;
                    while (g < (int)G->cls_grp[c].size() && l_hq >= G->cls_grp[c][g]) g++;
                    sg[(size_t)c * L + k] = g;
                    l_hq++;
                }
            }
            for (int t = 0; t < T; t++) tc[t] = G->hap_cls[2 * t];
            dalloc_copy(G->d_sgrp, sg.data(), sg.size(), s);
            dalloc_copy(G->d_tcls, tc.data(), tc.size(), s);
            Wk->isite.reserve(sizeof(int32_t) * (size_t)T * kst);
        }
    }
    if (learned) {
        Wk->fx.reserve(sizeof(float) * (size_t)npt * prs * MLP_FEAT);
        Wk->fy.reserve(sizeof(float) * (size_t)npt * prs);
        G->d_cov.reserve(sizeof(float) * (size_t)T);
    }
    if (collecting) {
        Wk->mass0.reserve(sizeof(float) * (size_t)T * kst);
        Wk->mass1.reserve(sizeof(float) * (size_t)T * kst);
        Wk->wlo.reserve(sizeof(int32_t) * (size_t)T * kst);
        Wk->whi.reserve(sizeof(int32_t) * (size_t)T * kst);
    }
    G->col_rows = 0;
    PbwtRun R = G->R;
    R.ins = (learned || collecting) ? Wk->ins.as<int32_t>() : nullptr;
    R.ins_site = (learned || collecting) ? Wk->isite.as<int32_t>() : nullptr;
    R.haps = H;
    R.ev = Wk->ev.as<int32_t>();
    R.ev_ci = Wk->evci.as<int32_t>();
    R.status = status + 1;
    R.out_states = states;
    R.out_stride = G->kstride;
    R.out_K = Kdev;
    R.stored = pst;
    R.start = pstart;
    // fresh chain: haplotypes, accumulators, item sizes
    GH_HIP(hipMemsetAsync(H, 0, (size_t)2 * T * L, s));
    GH_HIP(hipMemsetAsync(gp0, 0, sizeof(float) * (size_t)T * L, s));
    GH_HIP(hipMemsetAsync(gp1, 0, sizeof(float) * (size_t)T * L, s));
    GH_HIP(hipMemsetAsync(sto, 0, (size_t)T * L, s));
    GH_HIP(hipMemsetAsync(hds, 0, (size_t)T * L, s));
    GH_HIP(hipMemsetAsync(status, 0, sizeof(int32_t) * 2, s));
    gh_run_stats st;
    std::memset(&st, 0, sizeof(st));
    EvPair e_sel, e_cond, e_fb, e_ph, e_k0, e_k1, e_k0b, e_k1b, e_learn, e_pk;
    e_sel.make(); e_cond.make(); e_fb.make(); e_ph.make(); e_k0.make(); e_k1.make(); e_learn.make(); e_pk.make();
    e_k0b.make(); e_k1b.make();
    const dim3 gsite((L + 255) / 256, T), bsite(256);
    int main_done = 0;
    std::vector<CondItem> &items = G->items;
    std::vector<PhItem> &ph = G->ph;

    // learned selection (features, network, top-k) and training-row windows
    // of one PBWT iteration, on s
    auto learned_select = [&](int it, bool col_it) {
        FeatArgs fa;
        fa.hm = P->d_hm.as<uint64_t>(); fa.Wc = P->Wc; fa.Lc = P->Lc; fa.N = N; fa.L = L;
        fa.depth = prm->pbwt_depth; fa.kstride = kst; fa.rstride = prs;
        fa.c2t = P->d_c2t.as<int32_t>(); fa.crow = P->d_crow.as<int32_t>(); fa.flags = P->d_flags.as<uint8_t>();
        fa.ins_site = Wk->isite.as<int32_t>();
        fa.rare_ptr = P->d_rare_ptr.as<int32_t>(); fa.rare_idx = P->d_rare_idx.as<int32_t>();
        fa.H = H; fa.tmask = G->d_tmask.as<uint64_t>(); fa.targets = ptg; fa.Kpool = Kdev;
        fa.ins = Wk->ins.as<int32_t>();
        if (learned) {
            fa.X = Wk->fx.as<float>();
            fa.row0 = 0;
        } else {
            const int64_t need = G->col_rows + (int64_t)npt * prs;
            if ((int64_t)(G->d_cx.bytes / (sizeof(float) * MLP_FEAT)) < need) {
                // grow the row store (rare: sized for all collected iterations up front)
                const int64_t cap = (int64_t)npt * prs * std::max(1, n_iter - G->collect_from);
                GH_CHECK(G->col_rows == 0, GH_ERR_CAPACITY, "training-row store overflow");
                G->d_cx.reserve(sizeof(float) * (size_t)cap * MLP_FEAT);
                G->d_cy.reserve(sizeof(float) * (size_t)cap);
                G->d_cw.reserve(sizeof(float) * (size_t)cap);
            }
            fa.X = G->d_cx.as<float>();
            fa.row0 = G->col_rows;
        }
        feature_launch(P, npt, fa, s);
        if (col_it) {
            WinArgs wa;
            wa.targets = ptg; wa.Kpool = Kdev; wa.states = states; wa.ins = Wk->ins.as<int32_t>();
            wa.ins_site = Wk->isite.as<int32_t>(); wa.c2t = P->d_c2t.as<int32_t>();
            wa.crow = P->d_crow.as<int32_t>(); wa.Lc = P->Lc; wa.Wc = P->Wc; wa.kstride = kst;
            wa.win_lo = Wk->wlo.as<int32_t>(); wa.win_hi = Wk->whi.as<int32_t>();
            win_launch(npt, wa, s);
        }
        if (learned) {
            mlp_forward_launch(G->mlp->w.as<float>(), G->mlp->wb.as<uint16_t>(), Wk->fx.as<float>(),
                               (int64_t)npt * prs, Wk->fy.as<float>(), s);
            TopkArgs ta;
            ta.targets = ptg; ta.Kpool = Kdev; ta.klearn = G->d_klearn.as<int32_t>();
            ta.y = Wk->fy.as<float>(); ta.row0 = 0; ta.kstride = kst; ta.rstride = prs;
            ta.min_mass = G->min_mass;
            ta.prior = G->sel_prior;
            ta.ins = Wk->ins.as<int32_t>(); ta.states = states; ta.Kdev = Kdev; ta.items = items_d;
            ta.coverage = G->d_cov.as<float>(); ta.counters = G->d_cnt.as<int32_t>();
            ta.ins_site = G->sel_mode == 1 ? Wk->isite.as<int32_t>() : nullptr;
            ta.site_grp = G->d_sgrp.as<int32_t>(); ta.tcls = G->d_tcls.as<int32_t>(); ta.L = L;
            topk_launch(npt, ta, s);
        }
    };

    const int n_host = use_graph ? std::min(n_iter, 2) : n_iter;   // iterations launched from the host
    for (int it = 0; it < n_host; it++) {
        const int stage = it == 0 ? 0 : (it <= prm->burnin ? 1 : 2);
        bool learn_ev = false;
        GH_HIP(hipEventRecord(e_sel.a, s));
        if (stage == 0 && G->kinit > 0) {
            // INIT selection on the device (haplotype_set.cpp:108-169, 775-833); the
            // rare-site lists use the poly / pos scratch before the conditioning sets do
            InitArgs ia;
            ia.targets = nullptr;
            ia.gl = gl_d; ia.flat = flat_d; ia.ploidy = ploidy_d; ia.flags = P->d_flags.as<uint8_t>();
            ia.cref = P->d_cref.as<uint32_t>(); ia.calt = P->d_calt.as<uint32_t>();
            ia.rare_ptr = P->d_rare_ptr.as<int32_t>(); ia.rare_idx = P->d_rare_idx.as<int32_t>();
            ia.unph = G->d_unph.as<float>();
            ia.L = L; ia.N = N; ia.kinit = G->kinit; ia.kstride = kst; ia.max_mac = P->max_rare_mac; ia.seed = seed;
            ia.sind = poly; ia.sorted = pos; ia.states = states; ia.K = Kdev; ia.items = items_d;
            init_select_launch(ia, T, s);
        } else if ((stage == 0 || it == 1) && G->n_full > 0) {
            // full-panel targets (conditioning_set.cpp:105-113); PBWT targets are set by the walk
            full_panel_launch(G->d_fullp.as<int32_t>(), G->n_full, N, kst, states, Kdev, items_d, s);
        }
        if (stage > 0 && npt > 0) {
            // haplotype_set.cpp:335-348: one stored site per group, random start per haplotype
            std::fill(G->stored_h.begin(), G->stored_h.end(), 0);
            for (size_t c = 0; c < G->cls_mod.size(); c++) {
                int lo = 0;
                for (size_t gi = 0; gi < G->cls_grp[c].size(); gi++) {
                    const int hi = G->cls_grp[c][gi] - 1;
                    const uint32_t x = philox_draw(seed, (uint32_t)it, P_PBWT_STORED, (uint32_t)gi, (uint32_t)c);
                    G->stored_h[c * P->Lhq + urange(x, (uint32_t)lo, (uint32_t)hi)] = 1;
                    lo = G->cls_grp[c][gi];
                }
            }
            for (int i = 0; i < npt; i++)
                for (int j = 0; j < 2; j++) {
                    const uint32_t x =
                        philox_draw(seed, (uint32_t)it, P_PBWT_START, (uint32_t)(2 * G->pbwt_targets[i] + j), 0u);
                    G->start_h[2 * i + j] = (int32_t)urange(x, 0u, (uint32_t)(N - 1));
                }
            GH_HIP(hipMemcpyAsync(pst, G->stored_h.data(), G->stored_h.size(), hipMemcpyHostToDevice, s));
            GH_HIP(hipMemcpyAsync(pstart, G->start_h.data(), sizeof(int32_t) * 2 * npt, hipMemcpyHostToDevice, s));
            pbwt_select_launch(P, R, s);
            hipLaunchKernelGGL(set_k_kernel, dim3((npt + 127) / 128), dim3(128), 0, s, items_d, Kdev, ptg, npt);
            GH_HIP(hipGetLastError());
            st.pbwt_hap_sites += (double)2 * npt * L;
            const bool col_it = collecting && it >= G->collect_from;
            if (learned || col_it) {
                GH_HIP(hipEventRecord(e_learn.a, s));
                learned_select(it, col_it);
                GH_HIP(hipEventRecord(e_learn.b, s));
                learn_ev = true;
            }
        }
        GH_HIP(hipEventRecord(e_sel.b, s));
        GH_HIP(hipEventRecord(e_cond.a, s));
        // Conditioning sets.  The index step gives each item's polymorphic-site
        // count Lp; its Hvar rows, alpha checkpoints and phasing loci are placed
        // by Lp (not by the chunk's L), then the rows are filled.
        condset_index(P, T, items_d, states, poly, pos, tr, mark, prm->ne, s);
        GH_HIP(hipMemcpyAsync(items.data(), items_d, sizeof(CondItem) * T, hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
        int maxW_it = 1;
        {
            int64_t words = 0, aoff = 0, loff = 0;
            for (int t = 0; t < T; t++) {
                maxW_it = std::max(maxW_it, items[t].W);
                items[t].hvar_off = words;
                words += (int64_t)items[t].Lp * items[t].W;
                G->fb[t].alpha_off = aoff;
                aoff += fb_ckpt_floats(items[t].Lp, items[t].W, G->maxW);
                ph[t].locus_off = loff;
                loff += items[t].Lp;
            }
            grow(Wk->hvar, sizeof(uint64_t) * (size_t)std::max<int64_t>(words, 1));
            grow(Wk->alpha, sizeof(float) * (size_t)std::max<int64_t>(aoff, 1));
            grow(Wk->loci, sizeof(Locus) * (size_t)std::max<int64_t>(loff, 1));
            G->need.hvar_words = std::max(G->need.hvar_words, words);
            G->need.alpha_floats = std::max(G->need.alpha_floats, aoff);
            G->need.loci = std::max(G->need.loci, loff);
        }
        uint64_t *hvar = Wk->hvar.as<uint64_t>();
        float *alpha = Wk->alpha.as<float>();
        Locus *loci = Wk->loci.as<Locus>();
        GH_HIP(hipMemcpyAsync(items_d, items.data(), sizeof(CondItem) * T, hipMemcpyHostToDevice, s));
        GH_HIP(hipMemcpyAsync(fb_d, G->fb.data(), sizeof(FbItem) * T, hipMemcpyHostToDevice, s));
        condset_fill(P, T, items_d, states, poly, pos, hvar, maxW_it, s);
        // K-class batching (not while tuning): items of the largest word-count
        // class first in G->ford, the rest after them
        int fbA = T, fbB = 0, fbWA = maxW_it, fbWB = 0;
        if (!G->tune) {
            auto capw = [](int w) { return fb_cap_words(w); };
            int top = 0;
            for (int t = 0; t < T; t++) top = std::max(top, capw(items[t].W));
            G->ford.clear();
            fbWA = 1; fbWB = 1;
            for (int t = 0; t < T; t++)
                if (capw(items[t].W) == top) { G->ford.push_back(t); fbWA = std::max(fbWA, items[t].W); }
            fbA = (int)G->ford.size();
            for (int t = 0; t < T; t++)
                if (capw(items[t].W) < top) { G->ford.push_back(t); fbWB = std::max(fbWB, items[t].W); }
            fbB = T - fbA;
            maxW_it = fbWA;
            if (fbB > 0) dalloc_copy(G->d_ford, G->ford.data(), (size_t)T, s);
        }
        const int wcls = maxW_it <= 4 ? 0 : maxW_it <= 8 ? 1 : maxW_it <= 16 ? 2 : maxW_it <= 32 ? 3 : 4;
        // a launch of a few items cannot fill the GPU: its chains are the
        // critical path, so it takes the latency shapes
        const bool fewA = fbA <= FEW_ITEMS, fewB = fbB <= FEW_ITEMS;
        // the launch keys of this iteration's groups: the only ones a graph captured after it holds
        G->gk_fb[0] = graph_key(fb_cap_words(maxW_it), G->latency || fewA);
        G->gk_fb[1] = fbB > 0 ? graph_key(fb_cap_words(fbWB), G->latency || fewB) : -1;
        const FbPick fbp = (G->tune && stage > 0) ? G->fb_tuner.pick(wcls, fb_shapes(maxW_it), FbPick{0, 0})
                           : (G->latency || fewA) ? fb_default_pick(maxW_it, true)
                                                  : FbPick{0, 0};
        const bool label_it = collecting && stage > 0 && npt > 0 && it >= G->collect_from;
        float *mass0 = label_it ? Wk->mass0.as<float>() : nullptr, *mass1 = label_it ? Wk->mass1.as<float>() : nullptr;
        const int32_t *wlo = label_it ? Wk->wlo.as<int32_t>() : nullptr, *whi = label_it ? Wk->whi.as<int32_t>() : nullptr;
        GH_HIP(hipEventRecord(e_cond.b, s));
        GH_HIP(hipEventRecord(e_fb.a, s));
        // one forward-backward pass over every item: the top class, then the rest
        // (sized for its own class) beside it on the second stream or after it on
        // the run's stream (each dispatch timed by its own event pair)
        auto fb_all = [&](float *HPx, float *massx, EvPair &ea, EvPair &eb) {
            if (fbB == 0) {
                GH_HIP(hipEventRecord(ea.a, s));
                fb_launch(P, T, fb_d, items_d, maxW_it, G->maxW, poly, tr, hvar, HL, flat_d, alpha, HPx, G->ee_i, G->ed_i, s,
                          massx, wlo, whi, fbp);
                GH_HIP(hipEventRecord(ea.b, s));
                return;
            }
            const int32_t *ord = G->d_ford.as<int32_t>();
            const hipStream_t sB = fork();
            GH_HIP(hipEventRecord(ea.a, s));
            fb_launch(P, fbA, fb_d, items_d, fbWA, G->maxW, poly, tr, hvar, HL, flat_d, alpha, HPx, G->ee_i, G->ed_i, s,
                      massx, wlo, whi, fbp, ord);
            GH_HIP(hipEventRecord(ea.b, s));
            GH_HIP(hipEventRecord(eb.a, sB));
            fb_launch(P, fbB, fb_d, items_d, fbWB, G->maxW, poly, tr, hvar, HL, flat_d, alpha, HPx, G->ee_i, G->ed_i, sB,
                      massx, wlo, whi, (G->latency || fewB) ? fb_default_pick(fbWB, true) : FbPick{0, 0}, ord + fbA);
            GH_HIP(hipEventRecord(eb.b, sB));
            join();
        };
        // H0 (caller_algorithm.cpp:58-65)
        hipLaunchKernelGGL(hl_kernel, gsite, bsite, 0, s, gl_d, flat_d, H, L, T, stage == 0 ? 0 : 1, 0, prm->min_gl,
                           ploidy_d, HL);
        fb_all(HP0, mass0, e_k0, e_k0b);
        fb_mono_launch(P, T, fb_d, items_d, pos, HL, flat_d, HP0, G->ee_i, G->ed_i, s);
        hipLaunchKernelGGL(sample_kernel, gsite, bsite, 0, s, HP0, H, L, 0, seed, (uint32_t)it, nullptr);
        // H1 (caller_algorithm.cpp:66-71), diploid targets only
        hipLaunchKernelGGL(hl_kernel, gsite, bsite, 0, s, gl_d, flat_d, H, L, T, 1, 1, prm->min_gl, ploidy_d, HL);
        if (!G->any_haploid) {
            fb_all(HP1, mass1, e_k1, e_k1b);
        } else if (n_dip == 0) {
            GH_HIP(hipEventRecord(e_k1.a, s));   // an empty interval keeps the timing code uniform
            GH_HIP(hipEventRecord(e_k1.b, s));
        } else {
            // the diploid targets in pass-0 launch order, split where pass 0 split its classes
            std::vector<int32_t> &d1 = G->dord;
            d1.clear();
            int a1 = 0;
            for (int i = 0; i < T; i++) {
                const int t = G->tune ? i : G->ford[i];
                if (G->ploidy[t] != 2) continue;
                d1.push_back(t);
                a1 += i < fbA;
            }
            const int32_t *ord = dalloc_copy(G->d_dord, d1.data(), d1.size(), s);
            GH_HIP(hipEventRecord(e_k1.a, s));
            if (a1 > 0)
                fb_launch(P, a1, fb_d, items_d, maxW_it, G->maxW, poly, tr, hvar, HL, flat_d, alpha, HP1, G->ee_i,
                          G->ed_i, s, mass1, wlo, whi, fbp, ord);
            if (n_dip - a1 > 0)
                fb_launch(P, n_dip - a1, fb_d, items_d, fbB > 0 ? fbWB : maxW_it, G->maxW, poly, tr, hvar, HL, flat_d,
                          alpha, HP1, G->ee_i, G->ed_i, s, mass1, wlo, whi,
                          (G->latency || n_dip - a1 <= FEW_ITEMS) ? fb_default_pick(fbB > 0 ? fbWB : maxW_it, true)
                                                                  : FbPick{0, 0},
                          ord + a1);
            GH_HIP(hipEventRecord(e_k1.b, s));
        }
        fb_mono_launch(P, T, fb_d, items_d, pos, HL, flat_d, HP1, G->ee_i, G->ed_i, s);
        hipLaunchKernelGGL(sample_kernel, gsite, bsite, 0, s, HP1, H, L, 1, seed, (uint32_t)it, nullptr);
        GH_HIP(hipGetLastError());
        GH_HIP(hipEventRecord(e_fb.b, s));
        if (label_it) {
            LabelArgs la;
            la.targets = ptg; la.items = items_d; la.states = states; la.ins = Wk->ins.as<int32_t>();
            la.poly = poly; la.win_lo = wlo; la.win_hi = whi;
            la.mass0 = mass0; la.mass1 = mass1; la.kstride = kst; la.rstride = prs;
            la.X = G->d_cx.as<float>(); la.Y = G->d_cy.as<float>(); la.W = G->d_cw.as<float>();
            la.row0 = G->col_rows;
            label_launch(npt, la, s);
            G->col_rows += (int64_t)npt * prs;
        }
        // phasing (phasing_hmm::rephaseHaplotypes)
        GH_HIP(hipEventRecord(e_ph.a, s));
        PhItem *ph_d = dalloc_copy(G->d_ph, ph.data(), T, s);
        phase_prep_launch(T, ph_d, items_d, poly, H, H, flat_d, P->d_flags.as<uint8_t>(), loci, s);
        GH_HIP(hipMemcpyAsync(ph.data(), ph_d, sizeof(PhItem) * T, hipMemcpyDeviceToHost, s));
        GH_HIP(hipMemcpyAsync(items.data(), items_d, sizeof(CondItem) * T, hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
        if (G->any_haploid)
            for (int t = 0; t < T; t++)
                if (G->ploidy[t] == 1) { ph[t].n_segs = 0; ph[t].n_miss = 0; ph[t].n_loci = 0; }   // no phasing
        int maxK = 1;
        for (int t = 0; t < T; t++)
            if (G->ploidy[t] == 2) maxK = std::max(maxK, items[t].K);
        // K-class batching of the phasing launch, as for forward-backward; haploid targets are left out
        int phA = n_dip, phB = 0, phKB = 1;
        const bool ph_order = G->any_haploid || !G->tune;
        if (ph_order) {
            auto capk = [](int k) { return k <= 256 ? 256 : k <= 512 ? 512 : k <= 1024 ? 1024 : k <= 2048 ? 2048 : 4096; };
            const int top = capk(maxK);
            G->pord.clear();
            for (int t = 0; t < T; t++)
                if (G->ploidy[t] == 2 && (G->tune || capk(items[t].K) == top)) G->pord.push_back(t);
            phA = (int)G->pord.size();
            for (int t = 0; t < T; t++)
                if (G->ploidy[t] == 2 && !G->tune && capk(items[t].K) < top) {
                    G->pord.push_back(t);
                    phKB = std::max(phKB, items[t].K);
                }
            phB = n_dip - phA;
            if (!G->pord.empty()) dalloc_copy(G->d_pord, G->pord.data(), G->pord.size(), s);
        }
        G->gk_ph[0] = graph_key(ph_cap_states(maxK), G->latency || phA <= FEW_ITEMS);
        G->gk_ph[1] = phB > 0 ? graph_key(ph_cap_states(phKB), G->latency || phB <= FEW_ITEMS) : -1;
        const int kcls = maxK <= 256 ? 0 : maxK <= 512 ? 1 : maxK <= 1024 ? 2 : maxK <= 2048 ? 3 : 4;
        const PhaseShape phs = (G->tune && stage > 0)
                                   ? G->ph_tuner.pick(kcls, phase_shapes(maxK), phase_shape(maxK, G->latency))
                                   : phase_shape(maxK, G->latency || phA <= FEW_ITEMS);
        int64_t so = 0, ss = 0, io = 0, is = 0, dd = 0;
        for (int t = 0; t < T; t++) {
            ph[t].msg = phase_msg_floats(items[t].K);
            ph[t].seg_off = so; so += (int64_t)ph[t].n_segs * ph[t].msg;
            ph[t].segsum_off = ss; ss += ph[t].n_segs * 9;
            ph[t].imp_off = io; io += (int64_t)ph[t].n_miss * ph[t].msg;
            ph[t].impsum_off = is; is += ph[t].n_miss * 8;
            ph[t].dip_off = dd; dd += ph[t].n_segs;
            const double np = G->ploidy[t];   // forward-backward passes of the target
            st.fb_state_sites += np * items[t].K * items[t].Lp;
            st.fb_bytes += np * ((double)items[t].Lp * (2.0 * 8 * items[t].W + 8.0 * items[t].K + 16) + 8.0 * L);
            st.fb_poly_sites += np * items[t].Lp;
            st.ph_state_sites += (double)items[t].K * ph[t].n_loci;
            st.ph_segments += ph[t].n_segs + ph[t].n_miss;
            G->nstates[(size_t)t * n_iter + it] = items[t].K;
        }
        GH_HIP(hipMemcpyAsync(ph_d, ph.data(), sizeof(PhItem) * T, hipMemcpyHostToDevice, s));
        G->need.seg = std::max(G->need.seg, so);
        G->need.segsum = std::max(G->need.segsum, ss);
        G->need.imp = std::max(G->need.imp, io);
        G->need.impsum = std::max(G->need.impsum, is);
        G->need.dip = std::max(G->need.dip, dd);
        PhaseBufs PB;
        grow(Wk->seg, sizeof(float) * std::max<int64_t>(so, 1));
        grow(Wk->segsum, sizeof(float) * std::max<int64_t>(ss, 1));
        grow(Wk->imp, sizeof(float) * std::max<int64_t>(io, 1));
        grow(Wk->impsum, sizeof(float) * std::max<int64_t>(is, 1));
        grow(Wk->of1, sizeof(float) * std::max<int64_t>(is, 1));
        grow(Wk->dip, sizeof(int32_t) * std::max<int64_t>(dd, 1));
        PB.seg = Wk->seg.as<float>();
        PB.segsum = Wk->segsum.as<float>();
        PB.imp = Wk->imp.as<float>();
        PB.impsum = Wk->impsum.as<float>();
        PB.of1 = Wk->of1.as<float>();
        PB.dip = Wk->dip.as<int32_t>();
        GH_HIP(hipEventRecord(e_pk.a, s));
        if (!ph_order) {
            phase_launch(P, T, maxK, ph_d, items_d, hvar, loci, pos, prm->ne, PB, H, H, prm->err_phase, seed,
                         (uint32_t)it, status, s, phs);
        } else if (phB == 0) {
            if (phA > 0)
                phase_launch(P, phA, maxK, ph_d, items_d, hvar, loci, pos, prm->ne, PB, H, H, prm->err_phase, seed,
                             (uint32_t)it, status, s, phs, G->d_pord.as<int32_t>());
        } else {
            const int32_t *ord = G->d_pord.as<int32_t>();
            const hipStream_t sB = fork();
            if (phA > 0)
                phase_launch(P, phA, maxK, ph_d, items_d, hvar, loci, pos, prm->ne, PB, H, H, prm->err_phase, seed,
                             (uint32_t)it, status, s, phs, ord);
            phase_launch(P, phB, phKB, ph_d, items_d, hvar, loci, pos, prm->ne, PB, H, H, prm->err_phase, seed,
                         (uint32_t)it, status, sB, phase_shape(phKB, G->latency || phB <= FEW_ITEMS), ord + phA);
            join();
        }
        GH_HIP(hipEventRecord(e_pk.b, s));
        GH_HIP(hipEventRecord(e_ph.b, s));
        if (stage == 2) {
            hipLaunchKernelGGL(gp_kernel, gsite, bsite, 0, s, HP0, HP1, L, ploidy_d, gp0, gp1, sto, hds, nullptr, 0u);
            GH_HIP(hipGetLastError());
            main_done++;
        }
        int32_t sflags[2] = {0, 0};
        GH_HIP(hipMemcpyAsync(sflags, status, sizeof(sflags), hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
        GH_CHECK(sflags[0] == 0, GH_ERR_NUMERIC, "phasing HMM underflow (TRANS_HAP / SAMPLE_DIP)");
        GH_CHECK(sflags[1] == 0, GH_ERR_CAPACITY, "PBWT selection record capacity exceeded");
        st.ms_select += e_sel.ms();
        st.ms_condset += e_cond.ms();
        st.ms_fb += e_fb.ms();
        st.ms_phase += e_ph.ms();
        const bool k1b = fbB > 0 && !G->any_haploid;   // the H1 pass used the second stream
        st.fb_kernel_ms += e_k0.ms() + e_k1.ms() + (fbB > 0 ? e_k0b.ms() : 0.0f) + (k1b ? e_k1b.ms() : 0.0f);
        st.fb_launches += (fbB > 0 ? 2 : 1) + (G->any_haploid ? (n_dip > 0 ? 1 : 0) : (fbB > 0 ? 2 : 1));
        if (G->tune && stage > 0) {
            G->fb_tuner.record(e_k0.ms() + e_k1.ms());
            G->ph_tuner.record(e_pk.ms());
        }
        {
            const FbPick fu = fbp.NW ? fbp : fb_default_pick(maxW_it, G->latency);
            st.fb_nw = fu.NW; st.fb_mwl = fu.MWL; st.ph_nw = phs.NW; st.ph_kpt = phs.KPT;
        }
        if (learn_ev) st.ms_learn += e_learn.ms();
    }
    if (n_host < n_iter) {
        const auto g0 = std::chrono::steady_clock::now();
        // Buffers placed per iteration: the replays use what the host-launched
        // iterations reserved (at least a quarter above their largest need,
        // often more in a shared workspace).  Nothing is reallocated here: a
        // free waits for the whole device, every run in flight included.
        // GH_GRAPH_HEADROOM=f < 1 (testing) caps the replays at f x the need,
        // which forces the repeated execution.
        const char *eh = std::getenv("GH_GRAPH_HEADROOM");
        const double head = eh ? std::max(0.01, std::atof(eh)) : 1.0;
        auto cap_of = [&](DBuf &b, int64_t need, size_t elem) -> int64_t {
            const int64_t have = (int64_t)(b.bytes / elem);
            return head >= 1.0 ? have : std::min(have, (int64_t)std::max(1.0, std::ceil((double)need * head)));
        };
        GraphCaps caps;
        caps.hvar_words = cap_of(Wk->hvar, G->need.hvar_words, sizeof(uint64_t));
        caps.alpha_floats = cap_of(Wk->alpha, G->need.alpha_floats, sizeof(float));
        caps.loci = cap_of(Wk->loci, G->need.loci, sizeof(Locus));
        caps.seg = cap_of(Wk->seg, G->need.seg, sizeof(float));
        caps.segsum = cap_of(Wk->segsum, G->need.segsum, sizeof(float));
        caps.imp = cap_of(Wk->imp, G->need.imp, sizeof(float));
        caps.impsum = std::min(cap_of(Wk->impsum, G->need.impsum, sizeof(float)),
                               cap_of(Wk->of1, G->need.impsum, sizeof(float)));
        caps.dip = cap_of(Wk->dip, G->need.dip, sizeof(int32_t));
        // device tables of the replays
        GraphCtl c0;
        std::memset(&c0, 0, sizeof(c0));
        c0.iter = (uint32_t)n_host;
        // (per-run tables, allocated by run_create)
        GraphCtl *ctl = G->d_ctl.as<GraphCtl>();
        GH_HIP(hipMemcpyAsync(ctl, &c0, sizeof(c0), hipMemcpyHostToDevice, s));
        int32_t *nst = G->d_nst.as<int32_t>();
        int32_t *ford = G->d_ford.as<int32_t>();
        int32_t *pord = G->d_pord.as<int32_t>();
        const int32_t *d_gptr = G->d_gptr.as<int32_t>(), *d_gend = G->d_gend.as<int32_t>();
        PhItem *ph_d = G->d_ph.as<PhItem>();
        uint64_t *hvar = Wk->hvar.as<uint64_t>();
        float *alpha = Wk->alpha.as<float>();
        Locus *loci = Wk->loci.as<Locus>();
        PhaseBufs PB;
        PB.seg = Wk->seg.as<float>(); PB.segsum = Wk->segsum.as<float>(); PB.imp = Wk->imp.as<float>();
        PB.impsum = Wk->impsum.as<float>(); PB.of1 = Wk->of1.as<float>(); PB.dip = Wk->dip.as<int32_t>();
        GH_HIP(hipStreamSynchronize(s));   // host tables above are copied
        const uint32_t *itp = &ctl->iter;
        // the launch groups of the last host-launched iteration (keys checked by the placement kernels)
        // GH_GRAPH_FLIP_LATENCY=1 (testing): capture group A of both passes with
        // the other latency shape, so every replay's keys differ (the repeat path)
        if (const char *fl = std::getenv("GH_GRAPH_FLIP_LATENCY"))
            if (std::atoi(fl) == 1) {
                G->gk_fb[0] ^= 1;
                G->gk_ph[0] ^= 1;
            }
        auto fb_key_launch = [&](int grp, hipStream_t st, float *HPx) {
            const int key = G->gk_fb[grp], cap = key / 2;
            const int mw = std::min(cap, G->maxW);
            const FbPick pick = (key & 1) ? fb_default_pick(mw, true) : FbPick{0, 0};
            GraphGate g{&ctl->fb_key[grp], key, &ctl->fb_n[grp], itp};
            fb_launch(P, T, fb_d, items_d, mw, G->maxW, poly, tr, hvar, HL, flat_d, alpha, HPx, G->ee_i, G->ed_i, st,
                      nullptr, nullptr, nullptr, pick, ford + (size_t)grp * T, g);
        };
        auto ph_key_launch = [&](int grp, hipStream_t st) {
            const int key = G->gk_ph[grp], cap = key / 2;
            GraphGate g{&ctl->ph_key[grp], key, &ctl->ph_n[grp], itp};
            phase_launch(P, T, cap, ph_d, items_d, hvar, loci, pos, prm->ne, PB, H, H, prm->err_phase, seed, 0u,
                         status, st, phase_shape(cap, key & 1), pord + (size_t)grp * T, g);
        };
        auto fb_groups = [&](float *HPx) {
            if (G->gk_fb[1] < 0) {
                fb_key_launch(0, s, HPx);
                return;
            }
            const hipStream_t sB = fork();
            fb_key_launch(0, s, HPx);
            fb_key_launch(1, sB, HPx);
            join();
        };
        auto ph_groups = [&]() {
            if (G->gk_ph[1] < 0) {
                ph_key_launch(0, s);
                return;
            }
            const hipStream_t sB = fork();
            ph_key_launch(0, s);
            ph_key_launch(1, sB);
            join();
        };
        // one PBWT iteration, every decision on the device
        auto iteration = [&]() {
            if (npt > 0) {
                pbwt_draw_launch(ctl, seed, d_gptr, d_gend, (int)G->cls_grp.size(), P->Lhq, pst, ptg, npt, N, pstart,
                                 s);
                pbwt_select_launch(P, R, s);
                hipLaunchKernelGGL(set_k_kernel, dim3((npt + 127) / 128), dim3(128), 0, s, items_d, Kdev, ptg, npt);
                GH_HIP(hipGetLastError());
                if (learned) learned_select(-1, false);
            }
            condset_index(P, T, items_d, states, poly, pos, tr, mark, prm->ne, s);
            fb_place_launch(ctl, items_d, fb_d, ph_d, T, L, fb_block_sites(G->maxW), FEW_ITEMS, G->latency, caps,
                            G->gk_fb, ford, nst, n_iter, s);
            condset_fill(P, T, items_d, states, poly, pos, hvar, G->maxW, s);
            hipLaunchKernelGGL(hl_kernel, gsite, bsite, 0, s, gl_d, flat_d, H, L, T, 1, 0, prm->min_gl, ploidy_d, HL);
            fb_groups(HP0);
            fb_mono_launch(P, T, fb_d, items_d, pos, HL, flat_d, HP0, G->ee_i, G->ed_i, s);
            hipLaunchKernelGGL(sample_kernel, gsite, bsite, 0, s, HP0, H, L, 0, seed, 0u, itp);
            hipLaunchKernelGGL(hl_kernel, gsite, bsite, 0, s, gl_d, flat_d, H, L, T, 1, 1, prm->min_gl, ploidy_d, HL);
            fb_groups(HP1);
            fb_mono_launch(P, T, fb_d, items_d, pos, HL, flat_d, HP1, G->ee_i, G->ed_i, s);
            hipLaunchKernelGGL(sample_kernel, gsite, bsite, 0, s, HP1, H, L, 1, seed, 0u, itp);
            phase_prep_launch(T, ph_d, items_d, poly, H, H, flat_d, P->d_flags.as<uint8_t>(), loci, s);
            ph_place_launch(ctl, ph_d, items_d, T, FEW_ITEMS, G->latency, caps, G->gk_ph, pord, s);
            ph_groups();
            hipLaunchKernelGGL(gp_kernel, gsite, bsite, 0, s, HP0, HP1, L, ploidy_d, gp0, gp1, sto, hds, itp,
                               (uint32_t)(1 + prm->burnin));
            graph_advance_launch(ctl, s);
            GH_HIP(hipGetLastError());
        };
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        struct GraphHold {   // released after the replays have finished (the sync below)
            hipGraph_t &g;
            hipGraphExec_t &e;
            ~GraphHold() {
                if (e) (void)hipGraphExecDestroy(e);
                if (g) (void)hipGraphDestroy(g);
            }
        } hold{graph, exec};
        // Capture and instantiation are serialised over the process (runs
        // executing on several host threads capture one at a time).  A capture
        // the runtime refuses leaves the stream usable and the iterations are
        // enqueued without it (GO_UNCAPTURED; the same launches and results).
        bool captured = false;
        if (side_on && (G->gk_fb[1] >= 0 || G->gk_ph[1] >= 0)) (void)Wk->side();   // created outside the capture
        if (G->graph_mode == 1) {
            static std::mutex capture_mu;
            std::lock_guard<std::mutex> lk(capture_mu);
            (void)hipGetLastError();   // an earlier unchecked status must not be taken for a capture error
            GH_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
            std::string why;
            try {
                iteration();
            } catch (const std::exception &e) {
                why = e.what();
                hipGraph_t g = nullptr;
                (void)hipStreamEndCapture(s, &g);
                if (g) (void)hipGraphDestroy(g);
            }
            hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
            if (why.empty() && hipStreamIsCapturing(s, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
                hipError_t e = hipStreamEndCapture(s, &graph);
                if (e == hipSuccess && graph) e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
                if (e == hipSuccess && graph) captured = true;
                else why = std::string("end/instantiate: ") + hipGetErrorString(e);
            }
            (void)hipGetLastError();
            if (!captured) {
                G->graph_fail |= GO_UNCAPTURED;
                if (std::getenv("GH_GRAPH_VERBOSE")) std::fprintf(stderr, "[gh] graph capture refused: %s\n", why.c_str());
            }
        }
        for (int it = n_host; it < n_iter; it++) {
            if (captured)
                GH_HIP(hipGraphLaunch(exec, s));
            else
                iteration();
            if (it > prm->burnin) main_done++;
        }
        GraphCtl cend;
        std::vector<int32_t> nst_h((size_t)T * n_iter);
        GH_HIP(hipMemcpyAsync(&cend, ctl, sizeof(cend), hipMemcpyDeviceToHost, s));
        GH_HIP(hipMemcpyAsync(nst_h.data(), nst, sizeof(int32_t) * nst_h.size(), hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
        if (cend.overflow) {
            G->graph_fail |= cend.overflow;
            return false;
        }
        int32_t sflags[2] = {0, 0};
        GH_HIP(hipMemcpyAsync(sflags, status, sizeof(sflags), hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
        GH_CHECK(sflags[0] == 0, GH_ERR_NUMERIC, "phasing HMM underflow (TRANS_HAP / SAMPLE_DIP)");
        GH_CHECK(sflags[1] == 0, GH_ERR_CAPACITY, "PBWT selection record capacity exceeded");
        for (int t = 0; t < T; t++)
            for (int it = n_host; it < n_iter; it++)
                G->nstates[(size_t)t * n_iter + it] = nst_h[(size_t)t * n_iter + it];
        st.fb_state_sites += (double)cend.stats[GS_FB_STATE_SITES];
        st.fb_bytes += (double)cend.stats[GS_FB_BYTES];
        st.fb_poly_sites += (double)cend.stats[GS_FB_POLY_SITES];
        st.ph_state_sites += (double)cend.stats[GS_PH_STATE_SITES];
        st.ph_segments += (double)cend.stats[GS_PH_SEGMENTS];
        st.pbwt_hap_sites += (double)(n_iter - n_host) * 2 * npt * L;
        st.graph_iters = n_iter - n_host;
        st.graph_reruns = G->graph_fail;   // GO_UNCAPTURED only: the replays ran, enqueued without the graph
        st.ms_graph = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - g0).count();
    }
    if ((learned || collecting) && npt > 0) {
        int32_t cnt[2] = {0, 0};
        GH_HIP(hipMemcpyAsync(cnt, G->d_cnt.p, sizeof(cnt), hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
        st.learned_sel = cnt[0];
        st.pool_kept = cnt[1];
    }
    st.collected_rows = (double)G->col_rows;
    hipLaunchKernelGGL(finalize_kernel, gsite, bsite, 0, s, gp0, gp1, sto, hds, G->d_h0.as<uint8_t>(),
                       G->d_h1.as<uint8_t>(), L, main_done, ploidy_d);
    GH_HIP(hipGetLastError());
    GH_HIP(hipStreamSynchronize(s));
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - wall0).count();
    st.ms_other = st.ms_total - st.ms_select - st.ms_condset - st.ms_fb - st.ms_phase - st.ms_graph;
    G->st = st;
    G->done = true;
    return true;
}

void run_execute(gh_run *G) {
    GH_CHECK(G, GH_ERR_ARG, "gh_run_execute: null run");
    G->graph_fail = 0;
    const bool graph = G->graph_mode >= 1 && !G->tune && G->collect_from < 1 && !G->any_haploid && G->n_iter > 2;
    if (run_execute_once(G, graph)) return;
    // a replayed iteration outgrew its captured buffers: execute again with
    // every iteration launched from the host (the needs it records size the
    // next capture)
    run_execute_once(G, false);
    G->st.graph_reruns = G->graph_fail;   // GO_* bits
}

void run_set_selector_prior(gh_run *G, float w) {
    GH_CHECK(G && w >= 0.0f && w < 1e6f, GH_ERR_ARG, "gh_run_set_selector_prior: weight must be in [0, 1e6)");
    G->sel_prior = w;
}

void run_set_graph(gh_run *G, int mode) {
    GH_CHECK(G && mode >= 0 && mode <= 2, GH_ERR_ARG, "gh_run_set_graph: mode must be 0, 1 or 2");
    G->graph_mode = mode;
}

void run_collect(gh_run *G, int from_iter) {
    GH_CHECK(G, GH_ERR_ARG, "gh_run_collect: null run");
    GH_CHECK(from_iter >= 1 || from_iter < 0, GH_ERR_ARG, "gh_run_collect: rows come from PBWT iterations (>= 1)");
    GH_CHECK(from_iter < 0 || !G->mlp, GH_ERR_ARG, "gh_run_collect: the run uses the learned selector");
    G->collect_from = from_iter < 0 ? -1 : from_iter;
    if (G->collect_from >= 1 && G->npt > 0) {
        const int64_t cap = (int64_t)G->npt * G->pstride * std::max(1, G->n_iter - G->collect_from);
        G->d_cx.reserve(sizeof(float) * (size_t)cap * MLP_FEAT);
        G->d_cy.reserve(sizeof(float) * (size_t)cap);
        G->d_cw.reserve(sizeof(float) * (size_t)cap);
    }
}

void run_collected(gh_run *G, int64_t *n_rows, float *X, float *y, float *w) {
    GH_CHECK(G && n_rows, GH_ERR_ARG, "gh_run_collected: bad arguments");
    *n_rows = G->col_rows;
    hipStream_t s = G->stream;
    const size_t n = (size_t)G->col_rows;
    if (n == 0) return;
    if (X) GH_HIP(hipMemcpyAsync(X, G->d_cx.p, sizeof(float) * n * MLP_FEAT, hipMemcpyDeviceToHost, s));
    if (y) GH_HIP(hipMemcpyAsync(y, G->d_cy.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    if (w) GH_HIP(hipMemcpyAsync(w, G->d_cw.p, sizeof(float) * n, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
}

void run_train(gh_mlp *M, gh_run *G, int epochs, int batch, float lr, uint32_t seed, float *loss) {
    GH_CHECK(M && G, GH_ERR_ARG, "gh_mlp_train_run: bad arguments");
    GH_CHECK(G->col_rows > 0, GH_ERR_STATE, "gh_mlp_train_run: the run recorded no training rows");
    GH_CHECK(G->mlp != M, GH_ERR_ARG, "gh_mlp_train_run: the network is this run's selector");
    mlp_train_device(M, G->d_cx.as<float>(), G->d_cy.as<float>(), G->d_cw.as<float>(), G->col_rows, epochs, batch,
                     lr, seed, G->stream, loss);
}

void run_set_selector(gh_run *G, gh_mlp *M, const int32_t *klearn, float min_mass, int mode) {
    GH_CHECK(mode == 0 || mode == 1, GH_ERR_ARG, "gh_run_set_selector: mode must be 0 or 1");
    if (G) G->sel_mode = mode;
    GH_CHECK(G, GH_ERR_ARG, "gh_run_set_selector: null run");
    if (!M) { G->mlp = nullptr; return; }
    GH_CHECK(klearn, GH_ERR_ARG, "gh_run_set_selector: k_learn is required");
    GH_CHECK(G->collect_from < 1, GH_ERR_ARG, "gh_run_set_selector: the run records training rows");
    GH_CHECK(min_mass >= 0.0f && min_mass <= 1.0f, GH_ERR_ARG, "gh_run_set_selector: min_mass must be in [0, 1]");
    for (int t = 0; t < G->T; t++) GH_CHECK(klearn[t] > 0, GH_ERR_ARG, "gh_run_set_selector: k_learn must be > 0");
    dalloc_copy(G->d_klearn, klearn, (size_t)G->T, G->stream);
    GH_HIP(hipStreamSynchronize(G->stream));
    G->mlp = M;
    G->min_mass = min_mass;
}

void workspace_create(gh_workspace **out) {
    GH_CHECK(out, GH_ERR_ARG, "gh_workspace_create: null output");
    *out = new gh_workspace();
}

void workspace_destroy(gh_workspace *w) { delete w; }

int64_t workspace_bytes(const gh_workspace *w) {
    const gh::DBuf *b[] = {&w->H, &w->HL, &w->HP0, &w->HP1, &w->states, &w->poly, &w->pos, &w->t, &w->mark,
                           &w->hvar, &w->alpha, &w->loci, &w->seg, &w->segsum, &w->imp, &w->impsum, &w->of1,
                           &w->dip, &w->ev, &w->evci, &w->ins, &w->isite, &w->fx, &w->fy, &w->mass0, &w->mass1,
                           &w->wlo, &w->whi};
    int64_t n = 0;
    for (const gh::DBuf *x : b) n += (int64_t)x->bytes;
    return n;
}

void run_set_workspace(gh_run *G, gh_workspace *w) {
    GH_CHECK(G, GH_ERR_ARG, "gh_run_set_workspace: null run");
    G->W = w ? w : &G->own;
}

void run_set_tuning(gh_run *G, int mode) {
    GH_CHECK(G, GH_ERR_ARG, "gh_run_set_tuning: null run");
    GH_CHECK(mode >= 0 && mode <= 2, GH_ERR_ARG, "gh_run_set_tuning: mode must be 0, 1 or 2");
    G->tune = mode == 1;
    G->latency = mode == 2;
}

void run_fetch(gh_run *G, float *gp0_h, float *gp1_h, uint8_t *hds_h, uint8_t *h0_h, uint8_t *h1_h,
               int32_t *nstates_h, gh_run_stats *stats) {
    GH_CHECK(G && G->done, GH_ERR_ARG, "gh_run_fetch: the run has not been executed");
    hipStream_t s = G->stream;
    const size_t TL = (size_t)G->T * G->L;
    if (gp0_h) GH_HIP(hipMemcpyAsync(gp0_h, G->d_gp0.p, sizeof(float) * TL, hipMemcpyDeviceToHost, s));
    if (gp1_h) GH_HIP(hipMemcpyAsync(gp1_h, G->d_gp1.p, sizeof(float) * TL, hipMemcpyDeviceToHost, s));
    if (hds_h) GH_HIP(hipMemcpyAsync(hds_h, G->d_hds.p, TL, hipMemcpyDeviceToHost, s));
    if (h0_h) GH_HIP(hipMemcpyAsync(h0_h, G->d_h0.p, TL, hipMemcpyDeviceToHost, s));
    if (h1_h) GH_HIP(hipMemcpyAsync(h1_h, G->d_h1.p, TL, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
    if (nstates_h) std::memcpy(nstates_h, G->nstates.data(), sizeof(int32_t) * G->nstates.size());
    if (stats) *stats = G->st;
}

void run_destroy(gh_run *G) { delete G; }

void phase_run(gh_panel *P, int T, const uint8_t *gl_h, const uint8_t *flat_h, const gh_params *prm,
               const int32_t *kpt, float *gp0_h, float *gp1_h, uint8_t *hds_h, uint8_t *h0_h, uint8_t *h1_h,
               int32_t *nstates_h, gh_run_stats *stats) {
    gh_run *G = nullptr;
    run_create(P, T, gl_h, flat_h, prm, kpt, &G);
    try {
        run_execute(G);
        run_fetch(G, gp0_h, gp1_h, hds_h, h0_h, h1_h, nstates_h, stats);
    } catch (...) {
        run_destroy(G);
        throw;
    }
    run_destroy(G);
}

}  // namespace gh
