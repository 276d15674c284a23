// gh_ligate.hip -- ligation of overlapping phased chunks
// (ligate/src/ligater/ligater_algorithm.cpp).
//
// The host plans the record order of the ligated output from the chunks' site
// lists (contig, position, allele key), as the reference's synced reader walks
// them (:339-444): records of chunk f alone, then the overlap with chunk f + 1
// (records in both: the first half from f, the rest from f + 1), then f + 1.
// The device does the per-sample work: one launch counts, for every overlap at
// once, each sample's phased heterozygous genotypes that agree with chunk f's
// phase or with its swap (update_distances, :108-127); the host chains the
// swap decisions from chunk to chunk (scan_overlap, :196-213) and one launch
// per swapped chunk rewrites its genotypes (phase_update, :90-106).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "gh_internal.hpp"

namespace gh {

constexpr int32_t GT_VECTOR_END = INT32_MIN + 1;   // htslib bcf_int32_vector_end

__device__ __host__ __forceinline__ bool gt_missing(int32_t v) { return (v >> 1) == 0; }
__device__ __host__ __forceinline__ bool gt_phased(int32_t v) { return v & 1; }
__device__ __host__ __forceinline__ int32_t gt_allele(int32_t v) { return (v >> 1) - 1; }

// Per overlap o and sample i: m[o][i][0] phased hets with the same orientation
// in both chunks, m[o][i][1] with the opposite one.  pairs[k] = (element offset
// of the row in chunk f, of the row in chunk f + 1) in the concatenated GT buffer.
__global__ __launch_bounds__(256) void lig_count_kernel(int N, const int32_t *__restrict__ gt,
                                                        const int64_t *__restrict__ pairs,
                                                        const int64_t *__restrict__ ov_ptr,
                                                        int32_t *__restrict__ m) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int o = blockIdx.y;
    if (i >= N) return;
    int32_t same = 0, cross = 0;
    for (int64_t k = ov_ptr[o]; k < ov_ptr[o + 1]; k++) {
        const int2 ga = *(const int2 *)(gt + pairs[2 * k] + 2 * (int64_t)i);
        const int2 gb = *(const int2 *)(gt + pairs[2 * k + 1] + 2 * (int64_t)i);
        if (ga.y == GT_VECTOR_END || gb.y == GT_VECTOR_END) continue;
        if (gt_missing(ga.x) || gt_missing(ga.y) || gt_missing(gb.x) || gt_missing(gb.y)) continue;
        if (!gt_phased(ga.y) || !gt_phased(gb.y)) continue;
        const int32_t a0 = gt_allele(ga.x), a1 = gt_allele(ga.y), b0 = gt_allele(gb.x), b1 = gt_allele(gb.y);
        if (a0 == a1 || b0 == b1) continue;
        same += (a0 == b0 && a1 == b1);
        cross += (a0 == b1 && a1 == b0);
    }
    m[((int64_t)o * N + i) * 2] = same;
    m[((int64_t)o * N + i) * 2 + 1] = cross;
}

// phase_update on every row of one chunk: a swapped sample's phased diploid
// genotype becomes (allele of the second == 1) | (allele of the first == 1).
__global__ __launch_bounds__(256) void lig_apply_kernel(int64_t rows, int N, int32_t *__restrict__ gt,
                                                        const uint8_t *__restrict__ swap) {
    const int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= rows * N) return;
    const int i = (int)(g % N);
    if (!swap[i]) return;
    int2 *p = (int2 *)(gt + 2 * g);
    const int2 v = *p;
    if (gt_missing(v.x) || v.y == GT_VECTOR_END) return;
    if (!gt_phased(v.y)) return;
    int2 w;
    w.x = ((gt_allele(v.y) == 1 ? 1 : 0) + 1) << 1 | 1;
    w.y = ((gt_allele(v.x) == 1 ? 1 : 0) + 1) << 1 | 1;
    *p = w;
}

// The quality scan_overlap reports per sample (:200-206).
static float phase_quality(int32_t nmatch, int32_t nmism) {
    float f = 99;
    if (nmatch && nmism) {
        const float f0 = (float)nmatch / (nmatch + nmism);
        f = (float)(99 * (0.7 + f0 * logf(f0) + (1 - f0) * logf(1 - f0)) / 0.7);
    }
    return f;
}

namespace {

struct Emit {
    int32_t chunk, site;
};

// The merged walk over the overlap of chunks a and b.
struct OverlapPlan {
    bool present = false;
    int32_t a_from = 0;                    // first row of a inside the overlap
    int32_t b_from = 0;                    // first row of b read (rows before are skipped, :369)
    int32_t b_next = 0;                    // first row of b after the overlap
    std::vector<int32_t> rec_a, rec_b;     // merged records: row in a / b, -1 when absent
    std::vector<int32_t> pair_a, pair_b;   // records in both and biallelic in a (scanned)
    int32_t n_common = 0;
};

OverlapPlan plan_overlap(const gh_lig_chunk &A, const gh_lig_chunk &B) {
    OverlapPlan p;
    if (A.n_sites == 0 || B.n_sites == 0) return p;
    const int32_t ctg = B.contig[0], start = B.pos[0];
    if (A.contig[0] != ctg) return p;   // another contig: start_pos = -1 (:250-252)
    int32_t ia = 0;
    while (ia < A.n_sites && !(A.contig[ia] == ctg && A.pos[ia] >= start)) ia++;
    if (ia == A.n_sites) return p;
    p.present = true;
    p.a_from = ia;
    const int32_t seek = A.pos[ia];
    int32_t ib = 0;
    while (ib < B.n_sites && B.contig[ib] == ctg && B.pos[ib] < seek) ib++;
    p.b_from = ib;
    // merge by position; at one position, records of a pair with the first
    // unpaired record of b with the same allele key, then b's unpaired ones
    while (ia < A.n_sites) {
        const int32_t pos = A.pos[ia];
        int32_t ea = ia;
        while (ea < A.n_sites && A.pos[ea] == pos && A.contig[ea] == A.contig[ia]) ea++;
        while (ib < B.n_sites && B.contig[ib] == A.contig[ia] && B.pos[ib] < pos) {
            p.rec_a.push_back(-1);
            p.rec_b.push_back(ib++);
        }
        int32_t eb = ib;
        while (eb < B.n_sites && B.contig[eb] == A.contig[ia] && B.pos[eb] == pos) eb++;
        std::vector<char> used(eb - ib, 0);
        for (int32_t r = ia; r < ea; r++) {
            int32_t hit = -1;
            for (int32_t q = ib; q < eb; q++)
                if (!used[q - ib] && B.key[q] == A.key[r]) {
                    hit = q;
                    break;
                }
            p.rec_a.push_back(r);
            p.rec_b.push_back(hit);
            if (hit >= 0) {
                used[hit - ib] = 1;
                p.n_common++;
                if (!A.biallelic || A.biallelic[r]) {
                    p.pair_a.push_back(r);
                    p.pair_b.push_back(hit);
                }
            }
        }
        for (int32_t q = ib; q < eb; q++)
            if (!used[q - ib]) {
                p.rec_a.push_back(-1);
                p.rec_b.push_back(q);
            }
        ib = eb;
        ia = ea;
    }
    p.b_next = ib;
    return p;
}

}  // namespace

void ligate_scan(int S, int N, const int32_t *gt_a, const int32_t *gt_b, const uint8_t *swap_a, int32_t *nmatch,
                 int32_t *nmism, uint8_t *swap_b, float *phaseq) {
    GH_CHECK(S >= 0 && N > 0 && gt_a && gt_b && swap_a && nmatch && nmism && swap_b, GH_ERR_ARG,
             "gh_ligate_scan: bad arguments");
    hipStream_t s = nullptr;
    GH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    const size_t row = (size_t)2 * N;
    std::vector<int64_t> pairs(2 * (size_t)S), ov_ptr = {0, S};
    for (int k = 0; k < S; k++) {
        pairs[2 * k] = (int64_t)k * row;
        pairs[2 * k + 1] = (int64_t)(S + k) * row;
    }
    DBuf d_gt, d_pairs, d_ptr, d_m;
    d_gt.reserve(sizeof(int32_t) * 2 * row * (S ? S : 1));
    if (S) {
        GH_HIP(hipMemcpyAsync(d_gt.p, gt_a, sizeof(int32_t) * row * S, hipMemcpyHostToDevice, s));
        GH_HIP(hipMemcpyAsync(d_gt.as<int32_t>() + row * S, gt_b, sizeof(int32_t) * row * S, hipMemcpyHostToDevice,
                              s));
    }
    const int64_t *pp = dalloc_copy(d_pairs, pairs.data(), pairs.size(), s);
    const int64_t *op = dalloc_copy(d_ptr, ov_ptr.data(), ov_ptr.size(), s);
    d_m.reserve(sizeof(int32_t) * 2 * N);
    hipLaunchKernelGGL(lig_count_kernel, dim3((N + 255) / 256, 1), dim3(256), 0, s, N, d_gt.as<int32_t>(), pp, op,
                       d_m.as<int32_t>());
    GH_HIP(hipGetLastError());
    std::vector<int32_t> m(2 * (size_t)N);
    GH_HIP(hipMemcpyAsync(m.data(), d_m.p, sizeof(int32_t) * 2 * N, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
    for (int i = 0; i < N; i++) {
        nmatch[i] = swap_a[i] ? m[2 * i + 1] : m[2 * i];
        nmism[i] = swap_a[i] ? m[2 * i] : m[2 * i + 1];
        swap_b[i] = nmatch[i] < nmism[i];
        if (phaseq) phaseq[i] = phase_quality(nmatch[i], nmism[i]);
    }
}

void ligate(int n_chunks, int N, const gh_lig_chunk *chunks, int64_t cap, int32_t *out_chunk, int32_t *out_site,
            int64_t *n_out, uint8_t *swap, gh_lig_overlap *ov) {
    GH_CHECK(n_chunks > 0 && N > 0 && chunks && out_chunk && out_site && n_out && swap, GH_ERR_ARG,
             "gh_ligate: bad arguments");
    int64_t total = 0;
    for (int f = 0; f < n_chunks; f++) {
        const gh_lig_chunk &c = chunks[f];
        GH_CHECK(c.n_sites >= 0 && (c.n_sites == 0 || (c.contig && c.pos && c.key && c.gt)), GH_ERR_ARG,
                 "gh_ligate: chunk arrays");
        for (int32_t r = 1; r < c.n_sites; r++)
            GH_CHECK(c.contig[r] != c.contig[r - 1] || c.pos[r] >= c.pos[r - 1], GH_ERR_ARG,
                     "gh_ligate: positions of a chunk must ascend within a contig");
        total += c.n_sites;
    }
    // the files must come in ascending order (:262)
    for (int f = 1; f < n_chunks; f++)
        if (chunks[f].n_sites && chunks[f - 1].n_sites && chunks[f].contig[0] == chunks[f - 1].contig[0])
            GH_CHECK(chunks[f].pos[0] >= chunks[f - 1].pos[0], GH_ERR_ARG, "gh_ligate: chunks not in ascending order");

    std::vector<OverlapPlan> plans(n_chunks > 1 ? n_chunks - 1 : 0);
    for (int f = 0; f + 1 < n_chunks; f++) {
        plans[f] = plan_overlap(chunks[f], chunks[f + 1]);
        if (plans[f].present && f + 2 < n_chunks) {
            // three files open at one position (:386) is refused, as the reference does
            const gh_lig_chunk &B = chunks[f + 1], &C = chunks[f + 2];
            const gh_lig_chunk &A = chunks[f];
            if (C.n_sites && A.n_sites && C.contig[0] == A.contig[A.n_sites - 1])
                GH_CHECK(C.pos[0] > A.pos[A.n_sites - 1], GH_ERR_ARG, "gh_ligate: three chunks overlap");
            (void)B;
        }
    }

    // device copy of every chunk's GT rows
    hipStream_t s = nullptr;
    GH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    const int64_t row = 2 * (int64_t)N;
    std::vector<int64_t> base(n_chunks + 1, 0);
    for (int f = 0; f < n_chunks; f++) base[f + 1] = base[f] + (int64_t)chunks[f].n_sites * row;
    DBuf d_gt;
    d_gt.reserve(sizeof(int32_t) * (base[n_chunks] ? base[n_chunks] : 1));
    for (int f = 0; f < n_chunks; f++)
        if (chunks[f].n_sites)
            GH_HIP(hipMemcpyAsync(d_gt.as<int32_t>() + base[f], chunks[f].gt,
                                  sizeof(int32_t) * (size_t)chunks[f].n_sites * row, hipMemcpyHostToDevice, s));

    // one count launch over all overlaps
    const int n_ov = (int)plans.size();
    std::vector<int64_t> pairs, ov_ptr(n_ov + 1, 0);
    for (int f = 0; f < n_ov; f++) {
        for (size_t k = 0; k < plans[f].pair_a.size(); k++) {
            pairs.push_back(base[f] + (int64_t)plans[f].pair_a[k] * row);
            pairs.push_back(base[f + 1] + (int64_t)plans[f].pair_b[k] * row);
        }
        ov_ptr[f + 1] = (int64_t)pairs.size() / 2;
    }
    std::vector<int32_t> m((size_t)std::max(n_ov, 1) * N * 2, 0);
    if (n_ov > 0 && ov_ptr[n_ov] > 0) {
        DBuf d_pairs, d_ptr, d_m;
        const int64_t *pp = dalloc_copy(d_pairs, pairs.data(), pairs.size(), s);
        const int64_t *op = dalloc_copy(d_ptr, ov_ptr.data(), ov_ptr.size(), s);
        d_m.reserve(sizeof(int32_t) * m.size());
        hipLaunchKernelGGL(lig_count_kernel, dim3((N + 255) / 256, n_ov), dim3(256), 0, s, N, d_gt.as<int32_t>(), pp,
                           op, d_m.as<int32_t>());
        GH_HIP(hipGetLastError());
        GH_HIP(hipMemcpyAsync(m.data(), d_m.p, sizeof(int32_t) * m.size(), hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));
    }

    // chain the swaps: chunk 0 as it is; chunk f + 1 from f's overlap, else f's swaps
    std::memset(swap, 0, (size_t)N);
    for (int f = 0; f < n_ov; f++) {
        const uint8_t *sa = swap + (size_t)f * N;
        uint8_t *sb = swap + (size_t)(f + 1) * N;
        const OverlapPlan &p = plans[f];
        gh_lig_overlap st{};
        st.n_common = p.n_common;
        st.n_scanned = (int32_t)p.pair_a.size();
        if (p.n_common == 0) {
            std::memcpy(sb, sa, (size_t)N);
            st.mean_phaseq = 0.0f;
        } else {
            GH_CHECK(!p.pair_a.empty(), GH_ERR_ARG, "gh_ligate: overlap is empty");
            double hets = 0.0, q = 0.0;
            for (int i = 0; i < N; i++) {
                const int32_t same = m[((size_t)f * N + i) * 2], cross = m[((size_t)f * N + i) * 2 + 1];
                const int32_t nmatch = sa[i] ? cross : same, nmism = sa[i] ? same : cross;
                sb[i] = nmatch < nmism;
                st.n_swapped += sb[i];
                hets += nmatch + nmism;
                q += phase_quality(nmatch, nmism);
            }
            st.mean_hets = (float)(hets / N);
            st.mean_phaseq = (float)(q / N);
        }
        if (ov) ov[f] = st;
    }

    // emission order
    std::vector<Emit> out;
    out.reserve((size_t)total);
    int32_t from = 0;   // first row of the current chunk not yet emitted
    for (int f = 0; f < n_chunks; f++) {
        const gh_lig_chunk &A = chunks[f];
        const OverlapPlan *p = f < n_ov && plans[f].present ? &plans[f] : nullptr;
        const int32_t alone_end = p ? p->a_from : A.n_sites;
        for (int32_t r = from; r < alone_end; r++) out.push_back({f, r});
        if (!p) {
            from = 0;
            continue;
        }
        const int32_t half = (int32_t)p->pair_a.size() / 2;
        int32_t n_both = 0;
        for (size_t k = 0; k < p->rec_a.size(); k++) {
            const int32_t ra = p->rec_a[k], rb = p->rec_b[k];
            if (ra >= 0 && rb >= 0) {
                if (n_both < half) out.push_back({f, ra});
                else out.push_back({f + 1, rb});
                n_both++;
            } else if (ra >= 0) {
                out.push_back({f, ra});
            } else {
                out.push_back({f + 1, rb});
            }
        }
        from = p->b_next;
    }
    GH_CHECK((int64_t)out.size() <= cap, GH_ERR_CAPACITY, "gh_ligate: output capacity");
    for (size_t k = 0; k < out.size(); k++) {
        out_chunk[k] = out[k].chunk;
        out_site[k] = out[k].site;
    }
    *n_out = (int64_t)out.size();

    // phase_update on the swapped chunks, then the rows back to the caller
    DBuf d_sw;
    d_sw.reserve((size_t)N);
    for (int f = 0; f < n_chunks; f++) {
        const gh_lig_chunk &c = chunks[f];
        const uint8_t *sw = swap + (size_t)f * N;
        if (!c.n_sites || std::none_of(sw, sw + N, [](uint8_t v) { return v != 0; })) continue;
        GH_HIP(hipMemcpyAsync(d_sw.p, sw, (size_t)N, hipMemcpyHostToDevice, s));
        const int64_t cells = (int64_t)c.n_sites * N;
        hipLaunchKernelGGL(lig_apply_kernel, dim3((unsigned)((cells + 255) / 256)), dim3(256), 0, s,
                           (int64_t)c.n_sites, N, d_gt.as<int32_t>() + base[f], d_sw.as<uint8_t>());
        GH_HIP(hipGetLastError());
        GH_HIP(hipMemcpyAsync(c.gt, d_gt.as<int32_t>() + base[f], sizeof(int32_t) * (size_t)c.n_sites * row,
                              hipMemcpyDeviceToHost, s));
        GH_HIP(hipStreamSynchronize(s));   // d_sw is reused by the next chunk
    }
}

}  // namespace gh
