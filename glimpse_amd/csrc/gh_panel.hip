// gh_panel.hip -- resident panel: upload, site classification, bit transposition.
//
// HBM layout of one chunk panel (N haplotypes, L sites, Lc common sites):
//   d_sm   [Lc][W64]  uint64  site-major common rows, bit h%64 of word h/64 = hap h
//   d_hm   [N][Wc]    uint64  haplotype-major copy, bit j of word w = common row 64w+j
//   d_flags[L] uint8 (SF_COMMON | SF_MAJOR | SF_LQ), d_cm[L] double, d_cmpos[L] float
//   d_crow [L] int32 common row of a site (-1 if rare), d_c2t [Lc] site of a common row
//   d_shap_ptr/idx  CSR per haplotype of rare sites where it carries the non-major allele
//   d_rare_ptr/idx  CSR per site of haplotypes carrying the non-major allele
#include <algorithm>
#include <cstring>

#include "gh_device.hpp"
#include "gh_internal.hpp"
#include "gh_pbwt.hpp"

namespace gh {

// 64x64 bit-tile transpose with ballots: lane j holds the word of common row
// 64*rb + j for haplotypes 64*hb..64*hb+63; ballot over bit s gives haplotype
// (64*hb + s)'s word for rows 64*rb..64*rb+63.
__global__ __launch_bounds__(64) void transpose_panel_kernel(const uint64_t *__restrict__ sm, uint64_t *__restrict__ hm,
                                                            int Lc, int N, int W64, int Wc) {
    const int rb = blockIdx.x, hb = blockIdx.y, lane = threadIdx.x;
    const int row = rb * 64 + lane;
    uint64_t x = (row < Lc) ? sm[(size_t)row * W64 + hb] : 0ull;
    uint64_t mine = 0;
#pragma unroll 8
    for (int s = 0; s < 64; s++) {
        uint64_t m = __ballot((int)((x >> s) & 1ull));
        if (lane == s) mine = m;
    }
    const int h = hb * 64 + lane;
    if (h < N) hm[(size_t)h * Wc + rb] = mine;
}

static void classify(uint32_t cref, uint32_t calt, float sparse_maf, bool &is_common, bool &maj) {
    // common/src/io/ref_genotype_reader.cpp:171-175 (MAF in float, compared as double)
    double MAF = std::min(cref * 1.0f / (cref + calt), calt * 1.0f / (cref + calt));
    is_common = MAF >= sparse_maf;
    maj = calt > cref;
}

void panel_create(const gh_panel_desc *d, gh_panel **out) {
    GH_CHECK(d && out && d->n_sites > 0 && d->n_haps > 1, GH_ERR_ARG, "gh_panel_create: bad descriptor");
    gh_panel *P = new gh_panel();
    try {
        const int L = d->n_sites, N = d->n_haps;
        P->L = L; P->N = N;
        P->W64 = (N + 63) / 64;
        GH_HIP(hipGetDevice(&P->device));
        GH_HIP(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
        P->cm.assign(d->cm, d->cm + L);
        P->cref.assign(d->cref, d->cref + L);
        P->calt.assign(d->calt, d->calt + L);
        P->flags.assign(L, 0);
        P->cm_pos.resize(L);
        std::vector<int32_t> crow(L, -1);
        for (int l = 0; l < L; l++) {
            GH_CHECK(d->cref[l] + d->calt[l] == (uint32_t)N && d->cref[l] > 0 && d->calt[l] > 0, GH_ERR_ARG,
                     "gh_panel_create: site " + std::to_string(l) + " is monomorphic or AN != N");
            bool c, m;
            classify(d->cref[l], d->calt[l], d->sparse_maf > 0 ? d->sparse_maf : 0.001f, c, m);
            const bool lq = d->lq ? d->lq[l] != 0 : false;
            P->flags[l] = (c ? SF_COMMON : 0) | (m ? SF_MAJOR : 0) | (lq ? SF_LQ : 0);
            if (c) { crow[l] = (int32_t)P->common2tot.size(); P->common2tot.push_back(l); }
            P->Lhq += c && !lq;
            if (!c && !lq) P->max_rare_mac = std::max(P->max_rare_mac, (int)std::min(d->cref[l], d->calt[l]));
            P->cm_pos[l] = std::max(0.0f, (float)P->cm[l]);
        }
        P->Lc = (int)P->common2tot.size();
        P->Wc = (P->Lc + 63) / 64;
        // rare carriers: per site CSR, and per haplotype (ShapRef) by counting sort
        const int64_t nr = d->rare_ptr[L];
        std::vector<int32_t> shap_ptr(N + 1, 0), shap_idx(std::max<int64_t>(nr, 1));
        for (int l = 0; l < L; l++) {
            const bool com = P->flags[l] & SF_COMMON;
            const uint32_t nmin = (P->flags[l] & SF_MAJOR) ? d->cref[l] : d->calt[l];
            const int cnt = d->rare_ptr[l + 1] - d->rare_ptr[l];
            GH_CHECK(com ? cnt == 0 : cnt == (int)nmin, GH_ERR_ARG,
                     "gh_panel_create: carrier list inconsistent with AC/AN at site " + std::to_string(l));
            for (int j = d->rare_ptr[l]; j < d->rare_ptr[l + 1]; j++) {
                GH_CHECK(d->rare_idx[j] >= 0 && d->rare_idx[j] < N, GH_ERR_ARG, "gh_panel_create: carrier id");
                shap_ptr[d->rare_idx[j] + 1]++;
            }
        }
        for (int h = 0; h < N; h++) shap_ptr[h + 1] += shap_ptr[h];
        {
            std::vector<int32_t> fill(shap_ptr.begin(), shap_ptr.end() - 1);
            for (int l = 0; l < L; l++)
                for (int j = d->rare_ptr[l]; j < d->rare_ptr[l + 1]; j++) shap_idx[fill[d->rare_idx[j]]++] = l;
        }
        P->rare_ptr_h.assign(d->rare_ptr, d->rare_ptr + L + 1);
        P->rare_idx_h.assign(d->rare_idx, d->rare_idx + nr);
        hipStream_t s = P->stream;
        dalloc_copy(P->d_cm, P->cm.data(), L, s);
        dalloc_copy(P->d_cmpos, P->cm_pos.data(), L, s);
        dalloc_copy(P->d_flags, P->flags.data(), L, s);
        dalloc_copy(P->d_crow, crow.data(), L, s);
        dalloc_copy(P->d_c2t, P->common2tot.data(), P->Lc, s);
        dalloc_copy(P->d_sm, d->common_words, (size_t)P->Lc * P->W64, s);
        dalloc_copy(P->d_shap_ptr, shap_ptr.data(), N + 1, s);
        dalloc_copy(P->d_shap_idx, shap_idx.data(), (size_t)nr, s);
        dalloc_copy(P->d_rare_ptr, d->rare_ptr, L + 1, s);
        dalloc_copy(P->d_rare_idx, d->rare_idx, (size_t)nr, s);
        dalloc_copy(P->d_cref, d->cref, L, s);
        dalloc_copy(P->d_calt, d->calt, L, s);
        P->d_hm.reserve(sizeof(uint64_t) * std::max<size_t>((size_t)N * P->Wc, 1));
        if (P->Lc > 0) {
            dim3 grid(P->Wc, P->W64);
            hipLaunchKernelGGL(transpose_panel_kernel, grid, dim3(64), 0, s, P->d_sm.as<uint64_t>(),
                               P->d_hm.as<uint64_t>(), P->Lc, N, P->W64, P->Wc);
            GH_HIP(hipGetLastError());
        }
        if (d->ypacked && d->ypacked_len > 0) {
            P->ypacked.assign(d->ypacked, d->ypacked + d->ypacked_len);
            P->asmall_ptr.assign(d->asmall_ptr, d->asmall_ptr + P->Lhq + 2);
            P->asmall_idx.assign(d->asmall_idx, d->asmall_idx + P->asmall_ptr.back());
        }
        GH_HIP(hipStreamSynchronize(s));
        if (!P->ypacked.empty()) pbwt_decode(P);
        P->device_bytes += (int64_t)(P->d_cm.bytes + P->d_cmpos.bytes + P->d_flags.bytes + P->d_crow.bytes +
                                    P->d_c2t.bytes + P->d_sm.bytes + P->d_hm.bytes + P->d_shap_ptr.bytes +
                                    P->d_shap_idx.bytes + P->d_rare_ptr.bytes + P->d_rare_idx.bytes +
                                    P->d_cref.bytes + P->d_calt.bytes);
        *out = P;
    } catch (...) {
        if (P->stream) (void)hipStreamDestroy(P->stream);
        delete P->pbwt;
        delete P;
        throw;
    }
}

}  // namespace gh
