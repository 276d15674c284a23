// gh_capi.hip -- the C-ABI (include/glimpse_hip.h): argument checks, host<->device
// staging for the batch entry points, error capture.  No entry point exits.
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "gh_device.hpp"
#include "gh_internal.hpp"
#include "gh_mlp.hpp"
#include "gh_select.hpp"
#include "gh_phase.hpp"
#include "gh_pbwt.hpp"

namespace gh {
void panel_create(const gh_panel_desc *d, gh_panel **out);
void panel_destroy(gh_panel *p);
void phase_run(gh_panel *P, int T, const uint8_t *gl_h, const uint8_t *flat_h, const gh_params *prm,
               const int32_t *kpt, float *gp0_h, float *gp1_h, uint8_t *hds_h, uint8_t *h0_h, uint8_t *h1_h,
               int32_t *nstates_h, gh_run_stats *stats);
void run_create(gh_panel *P, int T, const uint8_t *gl_h, const uint8_t *flat_h, const gh_params *prm,
                const int32_t *kpt, gh_run **out);
void run_execute(gh_run *G);
void run_fetch(gh_run *G, float *gp0_h, float *gp1_h, uint8_t *hds_h, uint8_t *h0_h, uint8_t *h1_h,
               int32_t *nstates_h, gh_run_stats *stats);
void run_destroy(gh_run *G);
void run_collect(gh_run *G, int from_iter);
void run_set_tuning(gh_run *G, int mode);
void run_set_graph(gh_run *G, int mode);
void init_states_batch(gh_panel *P, int T, const uint8_t *gl, const uint8_t *flat, int kinit, uint32_t seed,
                       int kstride, int32_t *states, int32_t *K);
void run_set_selector_prior(gh_run *G, float w);
void workspace_create(gh_workspace **out);
void workspace_destroy(gh_workspace *w);
int64_t workspace_bytes(const gh_workspace *w);
void run_set_workspace(gh_run *G, gh_workspace *w);
void run_collected(gh_run *G, int64_t *n_rows, float *X, float *y, float *w);
void run_train(gh_mlp *M, gh_run *G, int epochs, int batch, float lr, uint32_t seed, float *loss);
void run_set_selector(gh_run *G, gh_mlp *M, const int32_t *klearn, float min_mass, int mode);
void impute_batch_noalloc(gh_panel *P, int n, const int32_t *state_ptr, const int32_t *states, const float *HL,
                          const uint8_t *flat, float err_imp, int ne, float *HP);
void output_fields(int T, int L, const float *gp0, const float *gp1, const uint8_t *hds, const uint32_t *cref,
                   const uint32_t *calt, float *ds, float *gp, uint8_t *gt, float *info, float *af, float *raf,
                   const int32_t *ploidy);
void run_set_ploidy(gh_run *G, const int32_t *ploidy);
void concordance(const gh_conc_params *prm, const gh_conc_input *in, gh_conc_output *out);
double conc_r2(uint64_t n, double sx, double sxx, double sxy, double sy, double syy, bool xc, bool yc);
void ligate_scan(int S, int N, const int32_t *gt_a, const int32_t *gt_b, const uint8_t *swap_a, int32_t *nmatch,
                 int32_t *nmism, uint8_t *swap_b, float *phaseq);
void ligate(int n_chunks, int N, const gh_lig_chunk *chunks, int64_t cap, int32_t *out_chunk, int32_t *out_site,
            int64_t *n_out, uint8_t *swap, gh_lig_overlap *ov);
void bam_default_params(gh_bam_params *p);
void bam_tables_out(float *t);
void bam_call_gls(const char *path, int32_t ploidy, const gh_bam_region *rg, const gh_bam_params *prm, uint8_t *gl,
                  uint8_t *flat, int32_t *depth, gh_bam_stats *st);

static thread_local std::string g_err;
void set_error(const std::string &m) { g_err = m; }

template <typename F>
static int guard(F &&f) {
    try {
        f();
        return GH_OK;
    } catch (const Error &e) {
        g_err = e.what();
        return e.code;
    } catch (const std::exception &e) {
        g_err = e.what();
        return GH_ERR_ARG;
    }
}

// Device staging of conditioning sets for a batch of explicit state lists.
struct CondBatch {
    DBuf items, states, poly, pos, t, hvar, mark;
    std::vector<CondItem> h;
    int maxW = 0;
    int64_t total_words = 0;

    void build(gh_panel *P, int n, const int32_t *state_ptr, const int32_t *states_h, int ne, hipStream_t s) {
        h.resize(n);
        const int L = P->L;
        maxW = 0;
        int64_t words = 0;
        for (int i = 0; i < n; i++) {
            const int K = state_ptr[i + 1] - state_ptr[i];
            GH_CHECK(K > 0, GH_ERR_STATE, "conditioning set " + std::to_string(i) + " has zero states");
            GH_CHECK(K <= P->N, GH_ERR_ARG, "more states than panel haplotypes");
            for (int k = state_ptr[i]; k < state_ptr[i + 1]; k++) {
                GH_CHECK(states_h[k] >= 0 && states_h[k] < P->N, GH_ERR_ARG, "state id out of range");
                GH_CHECK(k == state_ptr[i] || states_h[k] > states_h[k - 1], GH_ERR_ARG, "states must be sorted and unique");
            }
            CondItem &c = h[i];
            c.K = K;
            c.W = (K + 63) / 64;
            c.Lp = 0;
            c.pad = 0;
            c.state_off = state_ptr[i];
            c.site_off = (int64_t)i * L;
            c.hvar_off = words;
            words += (int64_t)L * c.W;
            maxW = std::max(maxW, c.W);
        }
        total_words = words;
        dalloc_copy(items, h.data(), n, s);
        dalloc_copy(states, states_h, state_ptr[n], s);
        poly.reserve(sizeof(int32_t) * (size_t)n * L);
        pos.reserve(sizeof(int32_t) * (size_t)n * L);
        t.reserve(sizeof(float) * (size_t)n * L);
        mark.reserve((size_t)n * L);
        hvar.reserve(sizeof(uint64_t) * std::max<int64_t>(words, 1));
        condset_build(P, n, items.as<CondItem>(), h.data(), states.as<int32_t>(), poly.as<int32_t>(),
                      pos.as<int32_t>(), t.as<float>(), hvar.as<uint64_t>(), mark.as<uint8_t>(), ne, maxW, s);
        GH_HIP(hipMemcpyAsync(h.data(), items.p, sizeof(CondItem) * n, hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
    }
};

}  // namespace gh

using namespace gh;

extern "C" {

const char *gh_last_error(void) { return g_err.c_str(); }

void gh_default_params(gh_params *p) {
    // phase/src/caller/caller_parameters.cpp:34-98
    p->burnin = 5;
    p->main_iter = 15;
    p->ne = 100000;
    p->min_gl = 1e-10f;
    p->err_imp = 1e-12f;
    p->err_phase = 0.0001f;
    p->pbwt_depth = 12;
    p->pbwt_modulo_cm = 0.1f;
    p->kinit = 1000;
    p->kpbwt = 2000;
    p->seed = 15052011u;
}

int gh_panel_create(const gh_panel_desc *d, gh_panel **out) {
    return guard([&] { panel_create(d, out); });
}

int gh_panel_destroy(gh_panel *p) {
    return guard([&] {
        if (!p) return;
        panel_destroy(p);
    });
}

int gh_panel_info(const gh_panel *p, int32_t *n_common, int32_t *n_common_hq, int64_t *device_bytes) {
    return guard([&] {
        GH_CHECK(p, GH_ERR_ARG, "null panel");
        if (n_common) *n_common = p->Lc;
        if (n_common_hq) *n_common_hq = p->Lhq;
        if (device_bytes) *device_bytes = p->device_bytes;
    });
}

int gh_condset_batch(gh_panel *P, int32_t n, const int32_t *state_ptr, const int32_t *states, int32_t ne,
                     int32_t *poly_ptr, int32_t *poly, int64_t poly_cap, int64_t *hvar_ptr, uint64_t *hvar,
                     int64_t hvar_cap, float *t) {
    return guard([&] {
        GH_CHECK(P && n > 0 && state_ptr && states, GH_ERR_ARG, "gh_condset_batch: bad arguments");
        hipStream_t s = P->stream;
        CondBatch B;
        B.build(P, n, state_ptr, states, ne, s);
        int64_t np = 0, nw = 0;
        for (int i = 0; i < n; i++) { np += B.h[i].Lp; nw += (int64_t)B.h[i].Lp * B.h[i].W; }
        GH_CHECK(np <= poly_cap && nw <= hvar_cap, GH_ERR_CAPACITY, "gh_condset_batch: output capacity");
        int64_t po = 0, wo = 0;
        for (int i = 0; i < n; i++) {
            const CondItem &c = B.h[i];
            poly_ptr[i] = (int32_t)po;
            hvar_ptr[i] = wo;
            GH_HIP(hipMemcpyAsync(poly + po, B.poly.as<int32_t>() + c.site_off, sizeof(int32_t) * c.Lp,
                                  hipMemcpyDeviceToHost, s));
            if (c.Lp > 1)
                GH_HIP(hipMemcpyAsync(t + po - i, B.t.as<float>() + c.site_off, sizeof(float) * (c.Lp - 1),
                                      hipMemcpyDeviceToHost, s));
            GH_HIP(hipMemcpyAsync(hvar + wo, B.hvar.as<uint64_t>() + c.hvar_off, sizeof(uint64_t) * c.Lp * c.W,
                                  hipMemcpyDeviceToHost, s));
            po += c.Lp;
            wo += (int64_t)c.Lp * c.W;
        }
        poly_ptr[n] = (int32_t)po;
        hvar_ptr[n] = wo;
        GH_HIP(hipStreamSynchronize(s));
    });
}

int gh_impute_batch(gh_panel *P, int32_t n, const int32_t *state_ptr, const int32_t *states, const float *HL,
                    const uint8_t *flat, float err_imp, int32_t ne, float *HP) {
    return guard([&] {
        GH_CHECK(P && n > 0 && state_ptr && states && HL && flat && HP, GH_ERR_ARG, "gh_impute_batch: bad arguments");
        hipStream_t s = P->stream;
        const int L = P->L;
        CondBatch B;
        B.build(P, n, state_ptr, states, ne, s);
        std::vector<FbItem> fb(n);
        int64_t aoff = 0;
        const int C = fb_block_sites(B.maxW);
        for (int i = 0; i < n; i++) {
            fb[i].K = B.h[i].K;
            fb[i].W = B.h[i].W;
            fb[i].Lp = B.h[i].Lp;
            fb[i].cond = i;
            fb[i].alpha_off = aoff;
            fb[i].hl_off = (int64_t)i * 2 * L;
            fb[i].flat_off = (int64_t)i * L;
            aoff += fb_ckpt_floats(B.h[i].Lp, B.h[i].W, B.maxW);
        }
        DBuf d_fb, d_HL, d_flat, d_alpha, d_HP;
        dalloc_copy(d_fb, fb.data(), n, s);
        dalloc_copy(d_HL, HL, (size_t)n * 2 * L, s);
        dalloc_copy(d_flat, flat, (size_t)n * L, s);
        d_alpha.reserve(sizeof(float) * std::max<int64_t>(aoff, 1));
        d_HP.reserve(sizeof(float) * (size_t)n * 2 * L);
        const float ei = std::min(std::max(err_imp, 1e-12f), 1e-3f);   // caller_initialise.cpp:152
        const float ed = ei, ee = (float)(1.0 - (double)ei);
        fb_launch(P, n, d_fb.as<FbItem>(), B.items.as<CondItem>(), B.maxW, B.maxW, B.poly.as<int32_t>(), B.t.as<float>(),
                  B.hvar.as<uint64_t>(), d_HL.as<float>(), d_flat.as<uint8_t>(), d_alpha.as<float>(),
                  d_HP.as<float>(), ee, ed, s);
        fb_mono_launch(P, n, d_fb.as<FbItem>(), B.items.as<CondItem>(), B.pos.as<int32_t>(), d_HL.as<float>(),
                       d_flat.as<uint8_t>(), d_HP.as<float>(), ee, ed, s);
        GH_HIP(hipMemcpyAsync(HP, d_HP.p, sizeof(float) * (size_t)n * 2 * L, hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
    });
}

int gh_impute_batch_noalloc(gh_panel *P, int32_t n, const int32_t *state_ptr, const int32_t *states, const float *HL,
                            const uint8_t *flat, float err_imp, int32_t ne, float *HP) {
    return guard([&] { impute_batch_noalloc(P, n, state_ptr, states, HL, flat, err_imp, ne, HP); });
}

int gh_build_sparse_pbwt(const gh_panel_desc *d, uint8_t *ypacked, int64_t *ypacked_len, int32_t *asmall_ptr,
                         int32_t *asmall_idx, int64_t *asmall_len, int32_t *n_common_hq) {
    return guard([&] {
        GH_CHECK(d && ypacked_len && asmall_len && n_common_hq, GH_ERR_ARG, "gh_build_sparse_pbwt: bad arguments");
        // the size query keeps its result for the fill call that follows it on
        // the same thread with the same description (the build is O(N x sites))
        struct Built {
            const gh_panel_desc *d = nullptr;
            const uint64_t *words = nullptr;
            int32_t L = 0, N = 0;
            std::vector<uint8_t> Y;
            std::vector<int32_t> ap, ai;
            int nhq = 0;
        };
        static thread_local Built last;
        const bool fill = ypacked && asmall_ptr && asmall_idx;
        const bool same = last.d == d && last.words == d->common_words && last.L == d->n_sites && last.N == d->n_haps;
        if (!same) {
            last = Built();
            build_sparse_pbwt(d, last.Y, last.ap, last.ai, last.nhq);
            last.d = d; last.words = d->common_words; last.L = d->n_sites; last.N = d->n_haps;
        }
        const std::vector<uint8_t> &Y = last.Y;
        const std::vector<int32_t> &ap = last.ap, &ai = last.ai;
        const int nhq = last.nhq;
        if (fill) {
            GH_CHECK(*ypacked_len >= (int64_t)Y.size() && *asmall_len >= (int64_t)ai.size(), GH_ERR_CAPACITY,
                     "gh_build_sparse_pbwt: output capacity");
            std::memcpy(ypacked, Y.data(), Y.size());
            std::memcpy(asmall_ptr, ap.data(), sizeof(int32_t) * ap.size());
            std::memcpy(asmall_idx, ai.data(), sizeof(int32_t) * ai.size());
        }
        *ypacked_len = (int64_t)Y.size();
        *asmall_len = (int64_t)ai.size();
        *n_common_hq = nhq;
        if (fill) last = Built();
    });
}

int gh_rephase_batch(gh_panel *P, int32_t n, const int32_t *state_ptr, const int32_t *states, const uint8_t *flat,
                     uint8_t *h0, uint8_t *h1, float err_phase, int32_t ne, uint32_t seed, int32_t it,
                     const int32_t *tgt) {
    return guard([&] {
        GH_CHECK(P && n > 0 && state_ptr && states && flat && h0 && h1 && tgt, GH_ERR_ARG,
                 "gh_rephase_batch: bad arguments");
        hipStream_t s = P->stream;
        const int L = P->L;
        CondBatch B;
        B.build(P, n, state_ptr, states, ne, s);
        int maxK = 0;
        std::vector<PhItem> ph(n);
        for (int i = 0; i < n; i++) {
            std::memset(&ph[i], 0, sizeof(PhItem));
            ph[i].cond = i;
            ph[i].tgt = tgt[i];
            ph[i].h0_off = ph[i].h1_off = ph[i].flat_off = (int64_t)i * L;
            ph[i].locus_off = (int64_t)i * L;
            maxK = std::max(maxK, B.h[i].K);
        }
        DBuf d_ph, d_H0, d_H1, d_flat, d_loci, d_seg, d_segsum, d_imp, d_impsum, d_of1, d_dip, d_status;
        dalloc_copy(d_ph, ph.data(), n, s);
        dalloc_copy(d_H0, h0, (size_t)n * L, s);
        dalloc_copy(d_H1, h1, (size_t)n * L, s);
        dalloc_copy(d_flat, flat, (size_t)n * L, s);
        d_loci.reserve(sizeof(Locus) * (size_t)n * L);
        phase_prep_launch(n, d_ph.as<PhItem>(), B.items.as<CondItem>(), B.poly.as<int32_t>(), d_H0.as<uint8_t>(),
                          d_H1.as<uint8_t>(), d_flat.as<uint8_t>(), P->d_flags.as<uint8_t>(), d_loci.as<Locus>(), s);
        GH_HIP(hipMemcpyAsync(ph.data(), d_ph.p, sizeof(PhItem) * n, hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
        const PhaseShape psh = phase_shape(maxK);
        int64_t so = 0, ss = 0, io = 0, is = 0, dd = 0;
        for (int i = 0; i < n; i++) {
            ph[i].msg = phase_msg_floats(B.h[i].K);
            ph[i].seg_off = so; so += (int64_t)ph[i].n_segs * ph[i].msg;
            ph[i].segsum_off = ss; ss += ph[i].n_segs * 9;
            ph[i].imp_off = io; io += (int64_t)ph[i].n_miss * ph[i].msg;
            ph[i].impsum_off = is; is += ph[i].n_miss * 8;
            ph[i].dip_off = dd; dd += ph[i].n_segs;
        }
        dalloc_copy(d_ph, ph.data(), n, s);
        PhaseBufs PB;
        PB.seg = (float *)d_seg.reserve(sizeof(float) * std::max<int64_t>(so, 1));
        PB.segsum = (float *)d_segsum.reserve(sizeof(float) * std::max<int64_t>(ss, 1));
        PB.imp = (float *)d_imp.reserve(sizeof(float) * std::max<int64_t>(io, 1));
        PB.impsum = (float *)d_impsum.reserve(sizeof(float) * std::max<int64_t>(is, 1));
        PB.of1 = (float *)d_of1.reserve(sizeof(float) * std::max<int64_t>(is, 1));
        PB.dip = (int32_t *)d_dip.reserve(sizeof(int32_t) * std::max<int64_t>(dd, 1));
        int32_t *st = (int32_t *)d_status.reserve(sizeof(int32_t));
        GH_HIP(hipMemsetAsync(st, 0, sizeof(int32_t), s));
        phase_launch(P, n, maxK, d_ph.as<PhItem>(), B.items.as<CondItem>(), B.hvar.as<uint64_t>(), d_loci.as<Locus>(),
                     B.pos.as<int32_t>(), ne, PB, d_H0.as<uint8_t>(), d_H1.as<uint8_t>(), err_phase, seed,
                     (uint32_t)it, st, s, psh);
        int32_t status = 0;
        GH_HIP(hipMemcpyAsync(&status, st, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GH_HIP(hipMemcpyAsync(h0, d_H0.p, (size_t)n * L, hipMemcpyDeviceToHost, s));
        GH_HIP(hipMemcpyAsync(h1, d_H1.p, (size_t)n * L, hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
        GH_CHECK(status == 0, GH_ERR_NUMERIC, "phasing HMM underflow (TRANS_HAP / SAMPLE_DIP)");
    });
}

int gh_pbwt_select(gh_panel *P, int32_t T, const uint8_t *hap_alleles, const uint8_t *stored,
                   const int32_t *start_idx, int32_t depth, float modulo_cm, int32_t *out_off, int32_t *out_cnt,
                   int32_t *out_ids, int64_t cap, int32_t kpbwt, int32_t *out_states, int32_t *out_K) {
    return guard([&] {
        GH_CHECK(P && T > 0 && hap_alleles && stored && start_idx && out_off && out_cnt && out_ids, GH_ERR_ARG,
                 "gh_pbwt_select: bad arguments");
        GH_CHECK(depth > 0 && depth <= 64, GH_ERR_ARG, "gh_pbwt_select: depth must be in 1..64");
        GH_CHECK(P->pbwt, GH_ERR_ARG, "gh_pbwt_select: panel has no sparse-PBWT stream");
        hipStream_t s = P->stream;
        const int L = P->L, nh = 2 * T, rec = 2 + 2 * depth;
        int ngrp = 0;
        for (int i = 0; i < P->Lhq; i++) ngrp += stored[i] != 0;
        for (int h = 0; h < nh; h++)
            GH_CHECK(start_idx[h] >= 0 && start_idx[h] < P->N, GH_ERR_ARG, "gh_pbwt_select: start index out of range");
        const int evcap = pbwt_event_cap(P, modulo_cm, ngrp);
        DBuf d_h, d_st, d_stored, d_cls, d_mod, d_ev, d_evci, d_cnt, d_status, d_kp, d_sh, d_out, d_K;
        PbwtRun R;
        std::memset(&R, 0, sizeof(R));
        R.n_haps = nh; R.n_samples = T; R.depth = depth; R.cap = evcap;
        R.haps = dalloc_copy(d_h, hap_alleles, (size_t)nh * L, s);
        R.start = dalloc_copy(d_st, start_idx, nh, s);
        R.stored = dalloc_copy(d_stored, stored, P->Lhq, s);
        std::vector<int32_t> cls(nh, 0);
        R.hap_class = dalloc_copy(d_cls, cls.data(), nh, s);
        R.modulo = dalloc_copy(d_mod, &modulo_cm, 1, s);
        R.ev = (int32_t *)d_ev.reserve(sizeof(int32_t) * (size_t)nh * evcap * rec);
        R.ev_ci = (int32_t *)d_evci.reserve(sizeof(int32_t) * (size_t)nh * evcap * 2);
        R.ev_cnt = (int32_t *)d_cnt.reserve(sizeof(int32_t) * nh);
        R.status = (int32_t *)d_status.reserve(sizeof(int32_t));
        GH_HIP(hipMemsetAsync(R.status, 0, sizeof(int32_t), s));
        std::vector<int32_t> kp(T, kpbwt), sh(T);
        for (int t = 0; t < T; t++) sh[t] = 2 * t;
        if (out_states && out_K) {
            GH_CHECK(kpbwt > 0, GH_ERR_ARG, "gh_pbwt_select: kpbwt must be positive");
            R.kpbwt = dalloc_copy(d_kp, kp.data(), T, s);
            R.sample_hap = dalloc_copy(d_sh, sh.data(), T, s);
            R.out_stride = std::min(kpbwt, P->N);
            R.out_states = (int32_t *)d_out.reserve(sizeof(int32_t) * (size_t)T * R.out_stride);
            R.out_K = (int32_t *)d_K.reserve(sizeof(int32_t) * T);
        }
        pbwt_select_launch(P, R, s);
        int32_t status = 0;
        std::vector<int32_t> cnt(nh), ev((size_t)nh * evcap * rec);
        GH_HIP(hipMemcpyAsync(&status, R.status, sizeof(int32_t), hipMemcpyDeviceToHost, s));
        GH_HIP(hipMemcpyAsync(cnt.data(), R.ev_cnt, sizeof(int32_t) * nh, hipMemcpyDeviceToHost, s));
        GH_HIP(hipMemcpyAsync(ev.data(), R.ev, sizeof(int32_t) * ev.size(), hipMemcpyDeviceToHost, s));
        if (out_states && out_K) {
            std::vector<int32_t> ob((size_t)T * R.out_stride);
            GH_HIP(hipMemcpyAsync(ob.data(), R.out_states, sizeof(int32_t) * ob.size(), hipMemcpyDeviceToHost, s));
            GH_HIP(hipMemcpyAsync(out_K, R.out_K, sizeof(int32_t) * T, hipMemcpyDeviceToHost, s));
            GH_HIP(hipStreamSynchronize(s));
            for (int t = 0; t < T; t++)
                std::memcpy(out_states + (size_t)t * kpbwt, ob.data() + (size_t)t * R.out_stride,
                            sizeof(int32_t) * out_K[t]);
        }
        GH_HIP(hipStreamSynchronize(s));
        GH_CHECK(status == 0, GH_ERR_CAPACITY, "gh_pbwt_select: selection record capacity exceeded");
        // rebuild pbwt_states[t][o] in the reference's push order: records ascending by
        // (site, phase), first haplotype before second; per rank, down entry then up entry
        int64_t o_ids = 0;
        for (int t = 0; t < T; t++) {
            std::vector<std::vector<int32_t>> lists(depth);
            const int32_t *A = ev.data() + (size_t)(2 * t) * evcap * rec;
            const int32_t *B = ev.data() + (size_t)(2 * t + 1) * evcap * rec;
            int ia = 0, ib = 0;
            const int na = cnt[2 * t], nb = cnt[2 * t + 1];
            while (ia < na || ib < nb) {
                const int32_t *r;
                if (ia < na && (ib >= nb || A[(size_t)ia * rec] <= B[(size_t)ib * rec])) r = A + (size_t)ia++ * rec;
                else r = B + (size_t)ib++ * rec;
                const int nd = r[1] & 0xffff, nu = r[1] >> 16;
                for (int o = 0; o < depth; o++) {
                    if (o < nd) lists[o].push_back(r[2 + o]);
                    if (o < nu) lists[o].push_back(r[2 + depth + o]);
                }
            }
            for (int o = 0; o < depth; o++) {
                out_off[t * depth + o] = (int32_t)o_ids;
                out_cnt[t * depth + o] = (int32_t)lists[o].size();
                GH_CHECK(o_ids + (int64_t)lists[o].size() <= cap, GH_ERR_CAPACITY, "gh_pbwt_select: ids capacity");
                std::memcpy(out_ids + o_ids, lists[o].data(), sizeof(int32_t) * lists[o].size());
                o_ids += lists[o].size();
            }
        }
    });
}

int gh_phase_run(gh_panel *P, int32_t T, const uint8_t *gl, const uint8_t *flat, const gh_params *prm,
                 const int32_t *kpbwt_per_target, float *gp0, float *gp1, uint8_t *hds, uint8_t *h0, uint8_t *h1,
                 int32_t *nstates, gh_run_stats *stats) {
    return guard([&] { phase_run(P, T, gl, flat, prm, kpbwt_per_target, gp0, gp1, hds, h0, h1, nstates, stats); });
}

int gh_run_create(gh_panel *P, int32_t T, const uint8_t *gl, const uint8_t *flat, const gh_params *prm,
                  const int32_t *kpbwt_per_target, gh_run **out) {
    return guard([&] { run_create(P, T, gl, flat, prm, kpbwt_per_target, out); });
}
int gh_run_execute(gh_run *r) {
    return guard([&] { run_execute(r); });
}
int gh_run_fetch(gh_run *r, float *gp0, float *gp1, uint8_t *hds, uint8_t *h0, uint8_t *h1, int32_t *nstates,
                 gh_run_stats *stats) {
    return guard([&] { run_fetch(r, gp0, gp1, hds, h0, h1, nstates, stats); });
}
int gh_run_destroy(gh_run *r) {
    return guard([&] { run_destroy(r); });
}

int gh_mlp_create(uint32_t seed, gh_mlp **out) {
    return guard([&] { mlp_create(seed, out); });
}
int gh_mlp_destroy(gh_mlp *m) {
    return guard([&] { mlp_destroy(m); });
}
int gh_mlp_get_params(gh_mlp *m, float *w) {
    return guard([&] { mlp_get(m, w); });
}
int gh_mlp_set_params(gh_mlp *m, const float *w) {
    return guard([&] { mlp_set(m, w); });
}
int gh_mlp_forward(gh_mlp *m, int64_t n, const float *X, float *y) {
    return guard([&] { mlp_forward_host(m, n, X, y); });
}
int gh_mlp_backward(gh_mlp *m, int32_t n, const float *X, const float *t, const float *w, float *grad,
                    float *loss) {
    return guard([&] { mlp_backward_host(m, n, X, t, w, grad, loss); });
}
int gh_mlp_step(gh_mlp *m, int32_t n, const float *X, const float *t, const float *w, float lr, float *loss) {
    return guard([&] { mlp_step_host(m, n, X, t, w, lr, loss); });
}
int gh_mlp_train_run(gh_mlp *m, gh_run *r, int32_t epochs, int32_t batch, float lr, uint32_t seed,
                     float *loss_per_epoch) {
    return guard([&] { run_train(m, r, epochs, batch, lr, seed, loss_per_epoch); });
}
int gh_workspace_create(gh_workspace **out) {
    return guard([&] { workspace_create(out); });
}

int gh_workspace_destroy(gh_workspace *w) {
    return guard([&] { workspace_destroy(w); });
}

int gh_workspace_bytes(const gh_workspace *w, int64_t *bytes) {
    return guard([&] {
        GH_CHECK(w && bytes, GH_ERR_ARG, "gh_workspace_bytes: bad arguments");
        *bytes = workspace_bytes(w);
    });
}

int gh_run_set_workspace(gh_run *r, gh_workspace *w) {
    return guard([&] { run_set_workspace(r, w); });
}

int gh_run_set_tuning(gh_run *r, int32_t mode) {
    return guard([&] { run_set_tuning(r, mode); });
}

int gh_init_states(gh_panel *p, int32_t T, const uint8_t *gl, const uint8_t *flat, int32_t kinit, uint32_t seed,
                   int32_t kstride, int32_t *states, int32_t *K) {
    return guard([&] { init_states_batch(p, T, gl, flat, kinit, seed, kstride, states, K); });
}

int gh_run_set_graph(gh_run *r, int32_t mode) {
    return guard([&] { run_set_graph(r, mode); });
}

int gh_run_set_selector_prior(gh_run *r, float weight) {
    return guard([&] { run_set_selector_prior(r, weight); });
}
int gh_run_collect(gh_run *r, int32_t from_iter) {
    return guard([&] { run_collect(r, from_iter); });
}
int gh_run_collected(gh_run *r, int64_t *n_rows, float *X, float *y, float *w) {
    return guard([&] { run_collected(r, n_rows, X, y, w); });
}
int gh_run_set_selector(gh_run *r, gh_mlp *m, const int32_t *k_learn, float min_mass, int32_t mode) {
    return guard([&] { run_set_selector(r, m, k_learn, min_mass, mode); });
}
int gh_output_fields(int32_t T, int32_t L, const float *gp0, const float *gp1, const uint8_t *hds,
                     const uint32_t *cref, const uint32_t *calt, float *ds, float *gp, uint8_t *gt, float *info,
                     float *af, float *raf) {
    return guard([&] { output_fields(T, L, gp0, gp1, hds, cref, calt, ds, gp, gt, info, af, raf, nullptr); });
}
int gh_output_fields_ploidy(int32_t T, int32_t L, const int32_t *ploidy, const float *gp0, const float *gp1,
                            const uint8_t *hds, const uint32_t *cref, const uint32_t *calt, float *ds, float *gp,
                            uint8_t *gt, float *info, float *af, float *raf) {
    return guard([&] { output_fields(T, L, gp0, gp1, hds, cref, calt, ds, gp, gt, info, af, raf, ploidy); });
}
int gh_run_set_ploidy(gh_run *r, const int32_t *ploidy) {
    return guard([&] { run_set_ploidy(r, ploidy); });
}
int gh_select_features(gh_panel *p, int32_t T, const uint8_t *H, const uint8_t *gl, const uint8_t *flat,
                       const int32_t *pool, const int32_t *pool_site, const int32_t *K, int32_t kstride, int32_t depth,
                       float *X) {
    return guard([&] { select_features_batch(p, T, H, gl, flat, pool, pool_site, K, kstride, depth, X); });
}
int gh_select_topk(int32_t T, const float *scores, const int32_t *pool, const int32_t *K, const int32_t *k_learn,
                   int32_t kstride, float min_mass, int32_t *states, int32_t *K_out, float *coverage) {
    return guard([&] { select_topk_batch(T, scores, pool, K, k_learn, kstride, min_mass, states, K_out, coverage); });
}

int gh_concordance(const gh_conc_params *prm, const gh_conc_input *in, gh_conc_output *out) {
    return guard([&] { concordance(prm, in, out); });
}
double gh_conc_r2(uint64_t n, double sx, double sxx, double sxy, double sy, double syy, int32_t x_constant,
                  int32_t y_constant) {
    return conc_r2(n, sx, sxx, sxy, sy, syy, x_constant != 0, y_constant != 0);
}
int gh_ligate_scan(int32_t n_sites, int32_t n_samples, const int32_t *gt_a, const int32_t *gt_b,
                   const uint8_t *swap_a, int32_t *nmatch, int32_t *nmism, uint8_t *swap_b, float *phaseq) {
    return guard([&] { ligate_scan(n_sites, n_samples, gt_a, gt_b, swap_a, nmatch, nmism, swap_b, phaseq); });
}
int gh_ligate(int32_t n_chunks, int32_t n_samples, const gh_lig_chunk *chunks, int64_t cap, int32_t *out_chunk,
              int32_t *out_site, int64_t *n_out, uint8_t *swap, gh_lig_overlap *ov) {
    return guard([&] { ligate(n_chunks, n_samples, chunks, cap, out_chunk, out_site, n_out, swap, ov); });
}

void gh_bam_default_params(gh_bam_params *p) {
    if (p) bam_default_params(p);
}
int gh_bam_tables(float *t) {
    return guard([&] { bam_tables_out(t); });
}
int gh_bam_call_gls(const char *path, int32_t ploidy, const gh_bam_region *r, const gh_bam_params *p, uint8_t *gl,
                    uint8_t *flat, int32_t *depth, gh_bam_stats *stats) {
    return guard([&] { bam_call_gls(path, ploidy, r, p, gl, flat, depth, stats); });
}

}  // extern "C"

namespace gh {
void panel_destroy(gh_panel *p) {
    hipStream_t s = p->stream;
    delete p->pbwt;
    p->pbwt = nullptr;
    delete p;
    if (s) (void)hipStreamDestroy(s);
}
}  // namespace gh
