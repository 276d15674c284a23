// gh_pbwt.hip -- sparse-PBWT conditioning-state selection on the GPU.
//
// split_reference side (host, once per panel):
//   build_sparse_pbwt  -- ref_haplotype_set::build_sparsePBWT
//                         (common/src/containers/ref_haplotype_set.cpp:47-193): pack3
//                         run-length stream of every non-LQ site in prefix order,
//                         with a "small" PBWT over the carriers of each rare run.
//   pbwt_decode        -- replays the stream once and keeps, per non-LQ site, the
//                         column bits in prefix order with per-word popcount prefixes
//                         (so the reference's rank array V[x] is one popcount), the
//                         prefix array A after every S-th HQ common site (S = 1 when
//                         all of them fit the memory budget) and the final small
//                         order of every rare run.  All of it stays resident in HBM;
//                         the per-iteration sweep then needs no decode at all.
// Per Gibbs iteration (device):
//   pbwt_walk_kernel   -- one lane per target haplotype replays
//                         haplotype_set::matchHapsFromCompressedPBWTSmall
//                         (phase/src/containers/haplotype_set.cpp:317-728) for that
//                         haplotype alone: LF-mapping of the insertion index and
//                         match interval, resets, rare-run re-indexing, and selectK
//                         events recorded as (site, phase) keyed records of neighbour
//                         positions in a named prefix order.  A
//                         workgroup is one walker wave and three loader waves that
//                         stage the next tile of rank data, site ops and alleles
//                         into the second of two LDS buffers.
//   pbwt_resolve_kernel -- panel ids of those positions: a stored prefix array, or
//                         steps back through the columns (select) and rare-run
//                         rearrangements to the nearest stored one.
//   pbwt_compact_kernel -- one wave per target merges its two haplotypes' records
//                         newest first, rank by rank, into the first Kpbwt distinct
//                         panel haplotypes (conditioning_set.cpp:96-104), then
//                         writes them sorted.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <numeric>

#include "gh_device.hpp"
#include "gh_internal.hpp"
#include "gh_pbwt.hpp"

namespace gh {

// LDS tiles of the walk: the rank data (column words + popcount prefixes) of
// consecutive site ops is contiguous, so a tile is one copy of a word range.
// A workgroup of the staged walk is one wave of walkers (one target haplotype
// per lane) and three loader waves; tiles are double buffered, so the loaders
// copy tile t+1 while the walkers walk tile t.
constexpr int WALK_WALKERS = 64;
constexpr int WALK_LOADERS = 192;
constexpr int WALK_NT = WALK_WALKERS + WALK_LOADERS;
// The tile sizes keep a walk workgroup at 80 KB of LDS (two buffers), so it
// shares a CU with forward-backward and phasing workgroups of other runs: with
// 4096-word tiles (147 KB) it could start only on an otherwise idle CU and kept
// that CU's compute workgroups off it for the whole walk.
constexpr int TILE_WORDS = 2048;
constexpr int TILE_OPS = 64;
// carrier positions (A_small_idx) of the tile's init ops, staged so the walk's
// lower-bound scans over them read LDS (at large N these sets hold hundreds)
constexpr int TILE_SI = 2048;
constexpr size_t WALK_BUF = (size_t)TILE_WORDS * 12 + sizeof(SiteOp) * TILE_OPS + (size_t)TILE_OPS * WALK_WALKERS +
                            (size_t)TILE_SI * 4;
constexpr size_t WALK_LDS = 2 * WALK_BUF;
static_assert(WALK_BUF % 16 == 0, "tile buffers keep 16-byte alignment");

/* ------------------------------------------------------------------ host: pack3 */
// pack3 code (ref_haplotype_set.h:45-85): bit 7 = symbol; 0x40 clear -> count in
// low 6 bits; 0x40 set and 0x20 clear -> (low 5 bits) << 6; both set -> << 11.
static int p3dec(uint8_t z) {
    const int c = z & 0x7f;
    return c < 64 ? c : (c < 96 ? (c - 64) << 6 : (c - 96) << 11);
}
static void p3add(uint8_t sym, uint32_t n, std::vector<uint8_t> &out) {
    const uint8_t y = (uint8_t)(sym << 7);
    const uint32_t M1 = 64, M2 = (95 - 63) << 6, M3 = (127 - 96) << 11;
    while (n >= M3) { out.push_back(y | 0x7f); n -= M3; }
    if (n >= M2) { out.push_back((uint8_t)(y | 0x60 | (n >> 11))); n &= 0x7ff; }
    if (n >= M1) { out.push_back((uint8_t)(y | 0x40 | (n >> 6))); n &= 0x3f; }
    if (n) out.push_back((uint8_t)(y | n));
}

struct HostSites {
    int L, N, Lc, Lhq;
    std::vector<uint8_t> common, major, lq;
    std::vector<uint32_t> cref, calt;
    std::vector<std::vector<int32_t>> carriers;  // per site, ascending
    const uint64_t *words;                        // [Lc][W64]
    int W64;
    std::vector<int32_t> crow;
};

static HostSites host_sites(const gh_panel_desc *d) {
    HostSites H;
    H.L = d->n_sites; H.N = d->n_haps; H.W64 = (H.N + 63) / 64;
    H.words = d->common_words;
    H.common.resize(H.L); H.major.resize(H.L); H.lq.resize(H.L); H.crow.assign(H.L, -1);
    H.cref.assign(d->cref, d->cref + H.L); H.calt.assign(d->calt, d->calt + H.L);
    H.carriers.resize(H.L);
    H.Lc = H.Lhq = 0;
    const float maf = d->sparse_maf > 0 ? d->sparse_maf : 0.001f;
    for (int l = 0; l < H.L; l++) {
        const uint32_t cr = d->cref[l], ca = d->calt[l];
        const double MAF = std::min(cr * 1.0f / (cr + ca), ca * 1.0f / (cr + ca));
        H.common[l] = MAF >= maf;
        H.major[l] = ca > cr;
        H.lq[l] = d->lq ? (d->lq[l] != 0) : 0;
        if (H.common[l]) { H.crow[l] = H.Lc++; H.Lhq += !H.lq[l]; }
        H.carriers[l].assign(d->rare_idx + d->rare_ptr[l], d->rare_idx + d->rare_ptr[l + 1]);
        std::sort(H.carriers[l].begin(), H.carriers[l].end());
    }
    return H;
}
static inline int hbit(const HostSites &H, int row, int h) {
    return (int)((H.words[(size_t)row * H.W64 + (h >> 6)] >> (h & 63)) & 1ull);
}

void build_sparse_pbwt(const gh_panel_desc *d, std::vector<uint8_t> &Y, std::vector<int32_t> &aptr,
                       std::vector<int32_t> &aidx, int &n_hq) {
    const HostSites H = host_sites(d);
    const int N = H.N;
    n_hq = H.Lhq;
    std::vector<std::vector<int32_t>> asmall(H.Lhq + 1);
    std::vector<int32_t> A(N), B(N), map_big_small(N, -1), sA, sB;
    std::vector<uint8_t> in_small(N, 0);        // positions of the run's carriers (cleared after use)
    std::vector<uint64_t> carrier((N + 63) / 64, 0ull);   // haplotype bitmap of the run's carriers
    std::iota(A.begin(), A.end(), 0);
    std::vector<uint8_t> small_hap;
    std::vector<std::pair<int, int>> by_pos;
    Y.clear();
    Y.reserve((size_t)H.Lc * 64 + (size_t)H.L * 4);
    int l_hq = 0, last_k = -1;
    for (int k = 0; k < H.L; k++) {
        if (H.lq[k]) continue;
        if (H.common[k]) {
            if (!sA.empty() && !(last_k >= 0 && H.common[last_k])) {
                // rare run -> common: non-carriers keep their order, carriers follow in small order
                for (int p : asmall[l_hq]) in_small[p] = 1;
                int nz = 0;
                for (int p = 0; p < N; p++) if (!in_small[p]) B[nz++] = A[p];
                for (size_t j = 0; j < sA.size(); j++) B[nz++] = A[asmall[l_hq][sA[j]]];
                for (int p : asmall[l_hq]) in_small[p] = 0;
                A.swap(B);
            }
            // one pass: the column in prefix order as runs of equal symbols, and
            // the stable partition of the prefix array by allele
            const uint64_t *row = H.words + (size_t)H.crow[k] * H.W64;
            int occ0 = 0, occ1 = (int)H.cref[k];
            int cur = -1;
            uint32_t run = 0;
            for (int m = 0; m < N; m++) {
                const int h = A[m];
                const int z = (int)((row[h >> 6] >> (h & 63)) & 1ull);
                if (z != cur) {
                    if (run) p3add((uint8_t)cur, run, Y);
                    cur = z;
                    run = 0;
                }
                run++;
                if (z) B[occ1++] = h; else B[occ0++] = h;
            }
            if (run) p3add((uint8_t)cur, run, Y);
            A.swap(B);
            l_hq++;
        } else {
            if (last_k < 0 || H.common[last_k]) {
                // new rare run: carriers of any non-LQ rare site until the next non-LQ common site
                std::vector<int32_t> haps;
                for (int kk = k; kk < H.L; kk++) {
                    if (H.lq[kk]) continue;
                    if (H.common[kk]) break;
                    haps.insert(haps.end(), H.carriers[kk].begin(), H.carriers[kk].end());
                }
                std::sort(haps.begin(), haps.end());
                haps.erase(std::unique(haps.begin(), haps.end()), haps.end());
                // their positions in the current prefix order, ascending
                by_pos.clear();
                for (int h : haps) carrier[h >> 6] |= 1ull << (h & 63);
                for (int p = 0; p < N && by_pos.size() < haps.size(); p++) {
                    const int h = A[p];
                    if ((carrier[h >> 6] >> (h & 63)) & 1ull) by_pos.emplace_back(p, h);
                }
                for (int h : haps) carrier[h >> 6] = 0ull;
                asmall[l_hq].clear();
                int j = 0;
                for (auto &e : by_pos) { map_big_small[e.second] = j++; asmall[l_hq].push_back(e.first); }
                sA.resize(asmall[l_hq].size());
                std::iota(sA.begin(), sA.end(), 0);
                sB.resize(sA.size());
                small_hap.resize(sA.size());
            }
            std::fill(small_hap.begin(), small_hap.end(), H.major[k]);
            for (int h : H.carriers[k]) small_hap[map_big_small[h]] = !H.major[k];
            const int ns = (int)sA.size();
            const int ref = H.major[k] ? N - (int)H.calt[k] : ns - (int)H.calt[k];
            int occ[2] = {0, ref};
            int m = 0;
            while (m < ns) {
                const int z = small_hap[sA[m]];
                const int m0 = m;
                while (++m < ns && small_hap[sA[m]] == z) {}
                for (int jj = m0; jj < m; jj++) sB[occ[z]++] = sA[jj];
                p3add((uint8_t)z, (uint32_t)(m - m0), Y);
            }
            sA.swap(sB);
        }
        last_k = k;
    }
    aptr.assign(H.Lhq + 2, 0);
    for (int l = 0; l <= H.Lhq; l++) aptr[l + 1] = aptr[l] + (int32_t)asmall[l].size();
    aidx.clear();
    for (auto &v : asmall) aidx.insert(aidx.end(), v.begin(), v.end());
}

/* ------------------------------------------------------------------ host: decode */
int pbwt_ckpt_every(int Lhq, int N) {
    if (const char *e = std::getenv("GH_PBWT_CKPT_EVERY")) {
        const int v = std::atoi(e);
        if (v >= 1) return v;
    }
    double mb = 1024.0;
    if (const char *e = std::getenv("GH_PBWT_CKPT_MB")) {
        const double v = std::atof(e);
        if (v > 0) mb = v;
    }
    const double bytes = (double)std::max(Lhq, 1) * N * 4.0, budget = mb * 1048576.0;
    return bytes <= budget ? 1 : (int)std::ceil(bytes / budget);
}

void pbwt_decode(gh_panel *P) {
    GH_CHECK(!P->ypacked.empty(), GH_ERR_ARG, "panel has no sparse-PBWT stream");
    const int N = P->N, L = P->L;
    auto *D = new gh_panel::PbwtDev();
    P->pbwt = D;
    std::vector<SiteOp> ops;
    std::vector<uint64_t> bits, cbits;   // column words of rare ops / of common sites
    std::vector<uint32_t> cnt, ccnt;
    const int Lhq = P->Lhq;
    const int S = pbwt_ckpt_every(Lhq, N);
    D->ck_every = S;
    std::vector<ColInfo> cols(std::max(Lhq, 1));
    std::vector<int32_t> sfin(std::max<size_t>(P->asmall_idx.size(), 1), 0);
    std::vector<int32_t> Astore;   // [stored arrays][N]: the order after every S-th HQ common site
    Astore.reserve((size_t)(Lhq / S + 1) * N);
    int n_ck = 0;
    std::vector<int32_t> A(N), B(N), sA, sB;
    std::vector<uint8_t> in_small(N, 0);
    std::iota(A.begin(), A.end(), 0);
    const uint8_t *y = P->ypacked.data(), *yend = y + P->ypacked.size();
    int l_hq = 0, last_k = -1, prev_ref = 0, n_runs = 0;
    auto decode_col = [&](int n, int ref, std::vector<int32_t> &a, std::vector<int32_t> &b,
                          std::vector<uint64_t> &bits, std::vector<uint32_t> &cnt) {
        // appends column bits (prefix order) + popcount prefixes, permutes a -> b -> a
        const int W = (n + 63) / 64;
        const size_t base = bits.size();
        bits.resize(base + W, 0ull);
        cnt.resize(base + W + 1, 0u);
        int occ[2] = {0, ref};
        int m = 0;
        while (m < n) {
            GH_CHECK(y < yend, GH_ERR_ARG, "sparse-PBWT stream truncated");
            const uint8_t z = *y++;
            const int c = p3dec(z), sym = z >> 7;
            GH_CHECK(c > 0 && m + c <= n, GH_ERR_ARG, "sparse-PBWT stream inconsistent with the panel");
            if (sym) {
                for (int j = m; j < m + c; j++) bits[base + (j >> 6)] |= 1ull << (j & 63);
                std::memcpy(b.data() + occ[1], a.data() + m, sizeof(int32_t) * c);
                occ[1] += c;
            } else {
                std::memcpy(b.data() + occ[0], a.data() + m, sizeof(int32_t) * c);
                occ[0] += c;
            }
            m += c;
        }
        GH_CHECK(occ[0] == ref && occ[1] == n, GH_ERR_ARG, "sparse-PBWT symbol counts inconsistent with AC/AN");
        uint32_t acc = 0;
        for (int w = 0; w < W; w++) { cnt[base + w] = acc; acc += (uint32_t)__builtin_popcountll(bits[base + w]); }
        cnt[base + W] = acc;
        // pad so the count array stays aligned with the word array
        bits.push_back(0ull);
        a.swap(b);
        return (int64_t)base;
    };
    for (int k = 0; k < L; k++) {
        const uint8_t fl = P->flags[k];
        if (fl & SF_LQ) continue;
        SiteOp op;
        std::memset(&op, 0, sizeof(op));
        op.k = k;
        op.cm_k = P->cm_pos[k];
        op.cm_km1 = P->cm_pos[k > 0 ? k - 1 : 0];
        op.cm_ps = P->cm_pos[0];
        if (fl & SF_COMMON) {
            op.kind = 0;
            op.l = l_hq;
            op.ps = l_hq >= 1 ? P->common2tot[l_hq - 1] : 0;
            op.cm_ps = P->cm_pos[op.ps];
            const int ns = P->asmall_ptr[l_hq + 1] - P->asmall_ptr[l_hq];
            ColInfo &ci = cols[l_hq];
            ci.ns = 0;
            ci.a_off = P->asmall_ptr[l_hq];
            if (ns > 0 && last_k >= 0 && !(P->flags[last_k] & SF_COMMON)) {
                // rare run -> common: the carriers leave their places and follow
                // the others in the run's final small order (kept in sfin)
                op.init = 1;
                GH_CHECK((int)sA.size() == ns, GH_ERR_ARG, "sparse-PBWT small set inconsistent with A_small_idx");
                const int32_t *si = P->asmall_idx.data() + P->asmall_ptr[l_hq];
                for (int j = 0; j < ns; j++) in_small[si[j]] = 1;
                int nz = 0;
                for (int p = 0; p < N; p++) if (!in_small[p]) B[nz++] = A[p];
                for (int j = 0; j < ns; j++) { B[nz++] = sA[j]; sfin[ci.a_off + j] = sA[j]; in_small[si[j]] = 0; }
                A.swap(B);
                ci.ns = ns;
            }
            op.ref = (int)P->cref[k];
            op.prev_ref = prev_ref;
            op.nbits = ns;
            op.a_off = P->asmall_ptr[l_hq];
            op.roff = decode_col(N, op.ref, A, B, cbits, ccnt);
            ci.roff = op.roff;
            ci.ref = op.ref;
            if ((l_hq + 1) % S == 0) {
                ci.ck = n_ck++;
                Astore.insert(Astore.end(), A.begin(), A.end());
            } else {
                ci.ck = -1;
            }
            prev_ref = op.ref;
            l_hq++;
        } else {
            op.kind = 1;
            op.l = l_hq;
            op.major = (fl & SF_MAJOR) ? 1 : 0;
            const int ns = P->asmall_ptr[l_hq + 1] - P->asmall_ptr[l_hq];
            if (last_k < 0 || (P->flags[last_k] & SF_COMMON)) {
                op.init = 1;
                n_runs++;
                const int32_t *si = P->asmall_idx.data() + P->asmall_ptr[l_hq];
                sA.resize(ns);
                for (int j = 0; j < ns; j++) sA[j] = A[si[j]];
                sB.resize(ns);
            }
            op.ref = op.major ? (int)P->cref[k] : ns - (int)P->calt[k];
            op.nbits = ns;
            op.a_off = P->asmall_ptr[l_hq];
            op.roff = decode_col(ns, op.ref, sA, sB, bits, cnt);
        }
        ops.push_back(op);
        last_k = k;
    }
    GH_CHECK(y == yend, GH_ERR_ARG, "sparse-PBWT stream longer than the panel");
    D->n_ops = (int)ops.size();
    for (int o = (int)ops.size() - 1, nxt = (int)ops.size(); o >= 0; o--) {
        ops[o].next_common = nxt;
        if (ops[o].kind == 0) nxt = o;
    }
    D->n_runs = n_runs;
    // LDS tiles for the walk: greedy runs of at most TILE_OPS ops whose rank
    // data fits TILE_WORDS.  A tile holds the rare ops' columns, and the common
    // sites' columns too when one is small (N <= 16,320: eight columns to a tile); larger common columns
    // stay in global memory and the walkers look them up there (a tile of two
    // 100,000-haplotype columns would cost a full copy per two sites).
    D->total_words = (int64_t)bits.size();
    D->total_cwords = (int64_t)cbits.size();
    const int64_t ccol = (N + 63) / 64 + 1;   // words of one common column (with its pad)
    D->stage_common = ccol <= TILE_WORDS / 8;
    {
        std::vector<int64_t> t_op, t_r, t_c, t_s;
        std::vector<int32_t> tsi;   // each tile's staged carrier sets, concatenated
        const int n = (int)ops.size();
        auto si_of = [&](const SiteOp &o) -> int64_t { return (o.init && o.nbits <= TILE_SI) ? o.nbits : 0; };
        auto words_of = [&](const SiteOp &o) -> int64_t {
            return o.kind == 0 ? (D->stage_common ? ccol : 0) : (int64_t)((o.nbits + 63) / 64 + 1);
        };
        bool ok = true;
        int64_t r = 0, c = 0;
        for (int o = 0; o < n && ok;) {
            const int o0 = o;
            if (words_of(ops[o0]) > TILE_WORDS) { ok = false; break; }
            t_op.push_back(o0); t_r.push_back(r); t_c.push_back(c); t_s.push_back((int64_t)tsi.size());
            int64_t wsum = 0, ssum = 0;
            while (o < n && o - o0 < TILE_OPS && wsum + words_of(ops[o]) <= TILE_WORDS &&
                   ssum + si_of(ops[o]) <= TILE_SI) {
                wsum += words_of(ops[o]);
                SiteOp &op = ops[o];
                op.si_loc = -1;
                if (si_of(op) > 0) {
                    op.si_loc = (int32_t)ssum;
                    tsi.insert(tsi.end(), P->asmall_idx.begin() + op.a_off, P->asmall_idx.begin() + op.a_off + op.nbits);
                    ssum += op.nbits;
                }
                if (op.kind == 0) c += ccol; else r += (op.nbits + 63) / 64 + 1;
                o++;
            }
        }
        if (ok && n > 0) {
            t_op.push_back(n); t_r.push_back(r); t_c.push_back(c); t_s.push_back((int64_t)tsi.size());
            GH_CHECK(r == D->total_words && c == D->total_cwords, GH_ERR_ARG, "pbwt tiles: rank data out of step");
            D->n_tiles = (int)t_op.size() - 1;
            std::vector<int64_t> tiles;
            tiles.insert(tiles.end(), t_op.begin(), t_op.end());
            tiles.insert(tiles.end(), t_r.begin(), t_r.end());
            tiles.insert(tiles.end(), t_c.begin(), t_c.end());
            tiles.insert(tiles.end(), t_s.begin(), t_s.end());
            dalloc_copy(D->tiles, tiles.data(), tiles.size(), P->stream);
            dalloc_copy(D->tsi, tsi.data(), tsi.size(), P->stream);
        } else {
            D->n_tiles = 0;   // a single op's rank data exceeds a tile: global lookups
            for (auto &op : ops) op.si_loc = -1;
        }
    }
    D->n_arrays = n_ck;
    hipStream_t s = P->stream;
    dalloc_copy(D->ops, ops.data(), ops.size(), s);
    dalloc_copy(D->bits, bits.data(), bits.size(), s);
    dalloc_copy(D->cnt, cnt.data(), cnt.size(), s);
    dalloc_copy(D->cbits, cbits.data(), cbits.size(), s);
    dalloc_copy(D->ccnt, ccnt.data(), ccnt.size(), s);
    dalloc_copy(D->A, Astore.data(), Astore.size(), s);
    dalloc_copy(D->aidx, P->asmall_idx.data(), P->asmall_idx.size(), s);
    dalloc_copy(D->sfin, sfin.data(), sfin.size(), s);
    dalloc_copy(D->cols, cols.data(), cols.size(), s);
    GH_HIP(hipStreamSynchronize(s));
    P->device_bytes += (int64_t)(D->ops.bytes + D->bits.bytes + D->cnt.bytes + D->cbits.bytes + D->ccnt.bytes +
                                 D->A.bytes + D->aidx.bytes +
                                 D->sfin.bytes + D->cols.bytes);
}

/* ------------------------------------------------------------------ device: walk */
struct WalkArgs {
    const SiteOp *ops;
    int n_ops;
    const uint64_t *bits;    // rare ops' columns
    const uint32_t *cnt;
    const uint64_t *cbits;   // common sites' columns
    const uint32_t *ccnt;
    int N;
    const int32_t *aidx;
    float cm0;             // cM of site 0 (the walk's first reset position)
    const uint8_t *haps;   // rows of L
    const int32_t *hap_row;
    int L;
    int n_haps;
    const int32_t *start;  // [n_haps]
    const uint8_t *stored; // [n_classes][Lhq]
    const int32_t *hap_class;  // [n_haps]
    const float *modulo;   // [n_classes]
    int Lhq;
    int depth;
    int32_t *ev;           // [n_haps][cap][rec]
    int32_t *ev_ci;        // [n_haps][cap][2]: prefix-array code and insertion index
    int cap;
    int32_t *ev_cnt;       // [n_haps]
    int32_t *status;
    const int64_t *tile_op;  // [n_tiles + 1] first op of each LDS tile
    const int64_t *tile_r;   // [n_tiles + 1] its first rare column word
    const int64_t *tile_c;   // [n_tiles + 1] its first common column word
    const int64_t *tile_s;   // [n_tiles + 1] its first staged carrier position in tsi
    const int32_t *tsi;
    int n_tiles;
};

__device__ __forceinline__ int rank1(const uint64_t *bits, const uint32_t *cnt, int64_t off, int x) {
    const int w = x >> 6, b = x & 63;
    int r = (int)cnt[off + w];
    if (b) r += __popcll(bits[off + w] & ((1ull << b) - 1ull));
    return r;
}
// Counts of the entries of the sorted list a[0, n) below v0, v1 and v2 (three
// lower bounds at once): branchless binary searches whose step count depends
// only on n, so the lanes of a wave, which search the same carrier list of a
// rare run, stay together; the three searches' loads overlap.
__device__ __forceinline__ void lower_count3(const int32_t *a, int n, int v0, int v1, int v2, int &c0, int &c1,
                                             int &c2) {
    if (n <= 0) { c0 = c1 = c2 = 0; return; }
    int b0 = 0, b1 = 0, b2 = 0;
    int m = n;
    while (m > 1) {
        const int half = m >> 1;
        const int x0 = a[b0 + half], x1 = a[b1 + half], x2 = a[b2 + half];
        b0 = x0 < v0 ? b0 + half : b0;
        b1 = x1 < v1 ? b1 + half : b1;
        b2 = x2 < v2 ? b2 + half : b2;
        m -= half;
    }
    c0 = b0 + (a[b0] < v0);
    c1 = b1 + (a[b1] < v1);
    c2 = b2 + (a[b2] < v2);
}

// STAGED: the rank data, site ops and this tile's target alleles / stored
// flags are copied to LDS tile by tile, so the per-site dependency chain (rank
// lookup -> new insertion index -> next lookup) runs at LDS latency.  Without
// STAGED every lookup goes to global memory (panels whose single site exceeds
// a tile).
// CGLOBAL (with STAGED): the common sites' columns are not staged; their rank
// lookups read global memory (large panels).
template <bool STAGED, bool CGLOBAL>
__global__ __launch_bounds__(STAGED ? WALK_NT : 64) void pbwt_walk_kernel(WalkArgs w) {
    extern __shared__ uint64_t wsm[];
    constexpr int NWALK = STAGED ? WALK_WALKERS : 64;
    const bool loader = STAGED && (int)threadIdx.x >= NWALK;
    const int h0 = blockIdx.x * NWALK + (loader ? 0 : (int)threadIdx.x);
    const bool act = !loader && h0 < w.n_haps;
    if (!STAGED && !act) return;
    const int h = act ? h0 : 0;   // loaders and idle walkers only move tiles
    const int N = w.N, depth = w.depth, rec = 2 + 2 * depth;
    const uint8_t *hap = w.haps + (size_t)(w.hap_row ? w.hap_row[h] : h) * w.L;
    const uint8_t *stored = w.stored + (size_t)w.hap_class[h] * w.Lhq;
    const float modulo = w.modulo[w.hap_class[h]];
    int32_t *ev = w.ev + (size_t)h * w.cap * rec;
    int32_t *ev_ci = w.ev_ci + (size_t)h * w.cap * 2;
    int nev = 0;
    int idx = w.start[h], f = 0, g = N, last_reset = 0, last_rare = 0, tar = 0;
    float cm_last = w.cm0;   // cM of site last_reset
    int sidx = 0, sf = 0, sg = 0;
    bool in_rare = false;
    bool overflow = false;

    // haplotype_set::selectK (haplotype_set.cpp:682-728) as one record: the
    // neighbour positions idx, idx+1, ... (down) and idx-1, idx-2, ... (up) that
    // carry the target allele; their panel ids are looked up afterwards in the
    // prefix array named by code (pbwt_resolve_kernel)
    auto selectK = [&](int key, int ref, int code, int a) {
        const int d_up = idx - f, d_down = g - idx, k1 = depth;
        int nh_up = 0, nh_down = 0;
        if (d_up < k1 && d_down < k1) { nh_up = d_up; nh_down = d_down; }
        else if (d_up >= k1 && d_down >= k1) { nh_up = k1; nh_down = k1; }
        else if (d_up < k1) nh_down = k1 - d_up;
        else if (d_down < k1) nh_up = k1 - d_down;
        if (nev >= w.cap) { overflow = true; return; }
        int32_t *r = ev + (size_t)nev * rec;
        int nd = 0, nu = 0;
        for (int j = 0; j < nh_down; j++) {
            const int i2 = idx + j;
            if (i2 < N && ((i2 >= ref) == (a != 0))) nd++; else break;
        }
        for (int j = 0; j < nh_up; j++) {
            const int i2 = idx - (j + 1);
            if (i2 >= 0 && ((i2 >= ref) == (a != 0))) nu++; else break;
        }
        if (nd == 0 && nu == 0) return;   // nothing pushed: no record needed
        r[0] = key;
        r[1] = nd | (nu << 16);
        ev_ci[2 * nev] = code;
        ev_ci[2 * nev + 1] = idx;
        nev++;
    };

    // Copies tile ti into buffer b: rank words and counts, the site ops, and
    // for every walker of this workgroup its allele at each op (bit 0) and, at
    // common ops, its class's stored flag (bit 1).  Threads t0, t0+nt, ... take
    // part; batches of independent loads, so a tile costs a few latencies.
    auto stage = [&](int ti, int b, int t0, int nt) __attribute__((always_inline)) {
        uint64_t *sb = (uint64_t *)((uint8_t *)wsm + (size_t)b * WALK_BUF);
        uint32_t *sc = (uint32_t *)(sb + TILE_WORDS);
        SiteOp *so = (SiteOp *)(sc + TILE_WORDS);
        uint8_t *sa = (uint8_t *)(so + TILE_OPS);
        int32_t *ss = (int32_t *)(sa + TILE_OPS * NWALK);
        const int o0 = (int)w.tile_op[ti], o1 = (int)w.tile_op[ti + 1], no = o1 - o0;
        {
            const int64_t s0 = w.tile_s[ti];
            const int nsi = (int)(w.tile_s[ti + 1] - s0);
            for (int i0 = t0; i0 < nsi; i0 += nt * 8) {
                int32_t v[8];
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int i = i0 + u * nt;
                    if (i < nsi) v[u] = w.tsi[s0 + i];
                }
#pragma unroll
                for (int u = 0; u < 8; u++) {
                    const int i = i0 + u * nt;
                    if (i < nsi) ss[i] = v[u];
                }
            }
        }
        // rare columns first, then (when staged) the common ones
        const int64_t r0 = w.tile_r[ti], c0 = w.tile_c[ti];
        const int nr = (int)(w.tile_r[ti + 1] - r0), nc = CGLOBAL ? 0 : (int)(w.tile_c[ti + 1] - c0);
        const int nwd = nr + nc;
        for (int i0 = t0; i0 < nwd; i0 += nt * 8) {
            uint64_t vb[8];
            uint32_t vc[8];
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = i0 + u * nt;
                if (i < nr) { vb[u] = w.bits[r0 + i]; vc[u] = w.cnt[r0 + i]; }
                else if (i < nwd) { vb[u] = w.cbits[c0 + i - nr]; vc[u] = w.ccnt[c0 + i - nr]; }
            }
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const int i = i0 + u * nt;
                if (i < nwd) { sb[i] = vb[u]; sc[i] = vc[u]; }
            }
        }
        for (int i = t0; i < no; i += nt) so[i] = w.ops[o0 + i];
        // pairs (walker, op), ops fastest: neighbouring lanes read neighbouring
        // sites of one haplotype row
        const int npair = no * NWALK;
        for (int p0 = t0; p0 < npair; p0 += nt * 16) {
            uint8_t va[16];
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int p = p0 + u * nt;
                if (p < npair) {
                    const int wk = p / no, j = p - wk * no;
                    const int hw0 = blockIdx.x * NWALK + wk;
                    const int hw = hw0 < w.n_haps ? hw0 : 0;
                    const SiteOp &oj = w.ops[o0 + j];
                    const uint8_t *hp = w.haps + (size_t)(w.hap_row ? w.hap_row[hw] : hw) * w.L;
                    const uint8_t *st = w.stored + (size_t)w.hap_class[hw] * w.Lhq;
                    va[u] = (uint8_t)(hp[oj.k] | (oj.kind == 0 ? (st[oj.l] != 0) << 1 : 0));
                }
            }
#pragma unroll
            for (int u = 0; u < 16; u++) {
                const int p = p0 + u * nt;
                if (p < npair) {
                    const int wk = p / no, j = p - wk * no;
                    sa[j * NWALK + wk] = va[u];
                }
            }
        }
    };

    const int n_tiles = STAGED ? w.n_tiles : 1;
    if (STAGED) {
        stage(0, 0, threadIdx.x, WALK_NT);
        __syncthreads();
    }
    for (int ti = 0; ti < n_tiles; ti++) {
    int o0 = 0, o1 = w.n_ops;
    int64_t r0 = 0, c0 = 0, nr = 0;
    const uint64_t *Rb = w.bits;
    const uint32_t *Rc = w.cnt;
    const SiteOp *s_ops = nullptr;
    const uint8_t *s_a = nullptr;
    const int32_t *s_si = nullptr;
    if (STAGED) {
        o0 = (int)w.tile_op[ti];
        o1 = (int)w.tile_op[ti + 1];
        r0 = w.tile_r[ti];
        c0 = w.tile_c[ti];
        nr = w.tile_r[ti + 1] - r0;
        const int b = ti & 1;
        Rb = (const uint64_t *)((const uint8_t *)wsm + (size_t)b * WALK_BUF);
        Rc = (const uint32_t *)(Rb + TILE_WORDS);
        s_ops = (const SiteOp *)(Rc + TILE_WORDS);
        s_a = (const uint8_t *)(s_ops + TILE_OPS);
        s_si = (const int32_t *)(s_a + TILE_OPS * NWALK);
        // buffer b ^ 1 was last read during tile ti - 1, before the barrier
        if (loader && ti + 1 < n_tiles) stage(ti + 1, b ^ 1, (int)threadIdx.x - NWALK, WALK_LOADERS);
    }
    if (act)
    for (int o = o0; o < o1; o++) {
        const SiteOp op = STAGED ? s_ops[o - o0] : w.ops[o];
        const int k = op.k;
        // where this op's column is: the tile (rare, then staged common), or global memory
        const int64_t roff = !STAGED ? op.roff : op.kind ? op.roff - r0 : CGLOBAL ? op.roff : nr + (op.roff - c0);
        const uint64_t *Cb = (STAGED && !CGLOBAL) ? Rb : w.cbits;
        const uint32_t *Cc = (STAGED && !CGLOBAL) ? Rc : w.ccnt;
        const int sa = STAGED ? s_a[(o - o0) * NWALK + threadIdx.x] : 0;
        const int a_here = STAGED ? (sa & 1) : hap[k];
        if (op.kind == 0) {
            if (op.init) {
                // haplotype_set::init_common (haplotype_set.cpp:562-630)
                const int ns = op.nbits;
                const int32_t *si = (STAGED && op.si_loc >= 0) ? s_si + op.si_loc : w.aidx + op.a_off;
                const int Aprev = code_after(op.l - 1);   // the order after the previous HQ common site
                int id, fd, gd;
                if (in_rare) {
                    id = N - ns + sidx;
                    fd = last_reset >= last_rare ? 0 : N - ns + sf;
                    gd = N - ns + sg;
                    if (gd <= fd) {
                        selectK(2 * k, op.prev_ref, Aprev, tar);
                        fd = 0; gd = N;
                        last_reset = k - 1;
                        cm_last = op.cm_km1;
                    }
                } else {
                    int c0, c1, c2;
                    lower_count3(si, ns, f, g, idx, c0, c1, c2);
                    id = idx - c2; fd = f - c0; gd = g - c1;
                    if (gd <= fd) {
                        const int ps = op.ps;
                        if (op.cm_ps - cm_last > modulo / 5.0f) selectK(2 * k, op.prev_ref, Aprev, tar);
                        fd = 0; gd = N - ns;
                        last_reset = ps;
                        cm_last = op.cm_ps;
                    }
                }
                idx = id; f = fd; g = gd;
                in_rare = false;
            }
            // haplotype_set::select_common_pd_fg (haplotype_set.cpp:445-490)
            const int a = a_here;
            const int prev = tar;
            tar = a;
            const int ref = op.ref;
            const int vi = rank1(Cb, Cc, roff, idx);
            const int vf = rank1(Cb, Cc, roff, f);
            const int vg = rank1(Cb, Cc, roff, g);
            const int id = a ? ref + vi : idx - vi;
            int fd = a ? ref + vf : f - vf;
            int gd = a ? ref + vg : g - vg;
            bool reset = false;
            const bool st = STAGED ? ((sa >> 1) & 1) != 0 : stored[op.l] != 0;
            if (gd <= fd) {
                if (st || op.cm_km1 - cm_last > modulo / 2.0f)
                    selectK(2 * k + 1, op.prev_ref, code_entering(op.l), prev);
                fd = a * ref;
                gd = a * N + (1 - a) * ref;
                last_reset = k; cm_last = op.cm_k;
                reset = true;
            }
            idx = id; f = fd; g = gd;
            if (!reset && st) {
                selectK(2 * k + 1, ref, code_after(op.l), a);
                f = a * ref;
                g = a * N + (1 - a) * ref;
                last_reset = k; cm_last = op.cm_k;
            }
        } else {
            const int ns = op.nbits;
            if (op.init) {
                // haplotype_set::init_rare (haplotype_set.cpp:632-680)
                in_rare = false;
                if (ns > 0) {
                    // the run's ops and this haplotype's alleles are in the tile up to o1
                    for (int j = o; j < w.n_ops; j++) {
                        int kind, major, aj;
                        if (STAGED && j < o1) {
                            kind = s_ops[j - o0].kind;
                            major = s_ops[j - o0].major;
                            aj = s_a[(j - o0) * NWALK + threadIdx.x] & 1;
                        } else {
                            const SiteOp &oj = w.ops[j];
                            kind = oj.kind;
                            major = oj.major;
                            aj = kind == 0 ? 0 : hap[oj.k];
                        }
                        if (kind == 0) break;
                        if (aj != major) { in_rare = true; break; }
                    }
                }
                if (in_rare) {
                    const int32_t *si = (STAGED && op.si_loc >= 0) ? s_si + op.si_loc : w.aidx + op.a_off;
                    int fd, gd, c2;
                    lower_count3(si, ns, f, g, idx, fd, gd, c2);
                    if (gd <= fd) { fd = 0; gd = ns; last_reset = k; cm_last = op.cm_k; }
                    sidx = c2; sf = fd; sg = gd;
                }
            }
            if (in_rare) {
                // haplotype_set::select_rare_pd_fg (haplotype_set.cpp:492-560)
                const int ref = op.ref;
                const int vi = rank1(Rb, Rc, roff, sidx);
                const int vf = rank1(Rb, Rc, roff, sf);
                const int vg = rank1(Rb, Rc, roff, sg);
                int si, fd, gd;
                if (a_here != op.major) {
                    last_rare = k;
                    si = ref + vi; fd = ref + vf; gd = ref + vg;
                    if (gd <= fd) { fd = ref; gd = ns; last_reset = k; cm_last = op.cm_k; }
                } else {
                    si = sidx - vi; fd = sf - vf; gd = sg - vg;
                    if (gd <= fd) { fd = 0; gd = ref; last_reset = k; cm_last = op.cm_k; }
                }
                sidx = si; sf = fd; sg = gd;
            }
            // no walker of this wave is inside the rare run: its remaining ops
            // change nothing, go to the run's end (or the tile's)
            if (!__any(in_rare)) o = min(op.next_common, o1) - 1;
        }
    }
    if (STAGED) __syncthreads();   // tile ti + 1 is in place and tile ti is no longer read
    }   // tiles
    if (!act) return;
    w.ev_cnt[h] = nev;
    if (overflow) atomicExch(w.status, 2);
}

/* ------------------------------------------------------------------ device: resolve */
// Panel ids of the neighbour positions the walk recorded.  A stored prefix
// array answers directly; otherwise a position is stepped back site by site
// (an "after" order to the entering order of the same site by the inverse LF
// mapping, select on the column; an entering order that a rare run
// rearranged to the previous "after" order, or to the run's final small
// order for its carriers) until a stored array or the identity order.
struct ResolveArgs {
    const ColInfo *cols;
    const uint64_t *bits;
    const uint32_t *cnt;
    const int32_t *A;      // [stored][N]
    const int32_t *aidx;   // A_small_idx (positions of the carriers, sorted)
    const int32_t *sfin;   // final small order of each rare run (aligned with aidx)
    int N, W;              // haplotypes, column words
    int32_t *ev;
    const int32_t *ev_ci;
    const int32_t *ev_cnt;
    int cap, depth, n_haps;
};

// position of the r-th (0-based) set bit of v
__device__ __forceinline__ int select_in_word(uint64_t v, int r) {
    int pos = 0;
#pragma unroll
    for (int sh = 32; sh > 0; sh >>= 1) {
        const uint64_t lo = v & ((1ull << sh) - 1ull);
        const int c = __popcll(lo);
        if (r >= c) { r -= c; v >>= sh; pos += sh; } else v = lo;
    }
    return pos;
}

// position of the x-th (0-based) 1 (one) or 0 of a column of W words with
// popcount prefixes cnt (ones before each word)
__device__ int select_col(const uint64_t *bits, const uint32_t *cnt, int64_t off, int W, int x, bool one) {
    int lo = 0, hi = W - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        const int c = one ? (int)cnt[off + mid] : 64 * mid - (int)cnt[off + mid];
        if (c <= x) lo = mid; else hi = mid - 1;
    }
    const uint64_t v = bits[off + lo];
    const int before = one ? (int)cnt[off + lo] : 64 * lo - (int)cnt[off + lo];
    return 64 * lo + select_in_word(one ? v : ~v, x - before);
}

__device__ int resolve_id(const ResolveArgs &a, int code, int p) {
    int l = code / 2 - 1;
    bool entering = code & 1;
    for (;;) {
        if (!entering) {
            if (l < 0) return p;   // identity order before the first site
            const ColInfo c = a.cols[l];
            if (c.ck >= 0) return a.A[(size_t)c.ck * a.N + p];
            p = p < c.ref ? select_col(a.bits, a.cnt, c.roff, a.W, p, false)
                          : select_col(a.bits, a.cnt, c.roff, a.W, p - c.ref, true);
            entering = true;
        } else {
            const ColInfo c = a.cols[l];
            if (c.ns > 0) {
                if (p >= a.N - c.ns) return a.sfin[c.a_off + p - (a.N - c.ns)];
                // the p-th position that is not a carrier's
                const int32_t *si = a.aidx + c.a_off;
                int q = p;
                for (int j = 0; j < c.ns && si[j] <= q; j++) q++;
                p = q;
            }
            l--;
            entering = false;
        }
    }
}

__global__ __launch_bounds__(256) void pbwt_resolve_kernel(ResolveArgs a) {
    const int h = blockIdx.x;
    if (h >= a.n_haps) return;
    const int rec = 2 + 2 * a.depth, per = 2 * a.depth;
    const int nev = a.ev_cnt[h];
    int32_t *ev = a.ev + (size_t)h * a.cap * rec;
    const int32_t *ci = a.ev_ci + (size_t)h * a.cap * 2;
    for (int i = threadIdx.x; i < nev * per; i += blockDim.x) {
        const int r = i / per, j = i - r * per;
        int32_t *rr = ev + (size_t)r * rec;
        const int nd = rr[1] & 0xffff, nu = rr[1] >> 16;
        const int code = ci[2 * r], idx = ci[2 * r + 1];
        if (j < a.depth) {
            if (j < nd) rr[2 + j] = resolve_id(a, code, idx + j);
        } else {
            const int u = j - a.depth;
            if (u < nu) rr[2 + a.depth + u] = resolve_id(a, code, idx - 1 - u);
        }
    }
}

/* ------------------------------------------------------------------ device: compact */
struct CompactArgs {
    const int32_t *ev;
    const int32_t *ev_cnt;
    int cap, depth, N;
    const int32_t *kpbwt;       // [n_samples]
    const int32_t *sample_hap;  // [n_samples]: first haplotype index (second is +1)
    const uint8_t *sample_one;  // [n_samples]: haploid sample (only the first haplotype's lists), or NULL
    const int32_t *sample_row;  // [n_samples]: output row
    int32_t *out_states;        // [n_samples][out_stride]
    int out_stride;
    int32_t *out_K;             // [n_samples]
    int32_t *ins;               // [n_samples][out_stride] or NULL
    int32_t *ins_site;          // [n_samples][out_stride] or NULL
};

__global__ __launch_bounds__(64) void pbwt_compact_kernel(CompactArgs c) {
    extern __shared__ uint64_t bm[];
    const int t = blockIdx.x, lane = threadIdx.x;
    const int Wn = (c.N + 63) / 64;
    for (int i = lane; i < Wn; i += 64) bm[i] = 0ull;
    __syncthreads();
    const int rec = 2 + 2 * c.depth;
    const int h0 = c.sample_hap[t], h1 = h0 + 1;
    const int32_t *eA = c.ev + (size_t)h0 * c.cap * rec, *eB = c.ev + (size_t)h1 * c.cap * rec;
    // a haploid target has one haplotype (haplotype_set.cpp:263-270): its second walker is not read
    const int nA = c.ev_cnt[h0], nB = (c.sample_one && c.sample_one[t]) ? 0 : c.ev_cnt[h1];
    const int kp = c.kpbwt[t];
    const int orow = c.sample_row ? c.sample_row[t] : t;
    int32_t *ins = c.ins ? c.ins + (size_t)orow * c.out_stride : nullptr;
    int32_t *ins_site = c.ins_site ? c.ins_site + (size_t)orow * c.out_stride : nullptr;
    if (lane == 0) {
        int count = 0;
        for (int o = 0; o < c.depth && count < kp; o++) {
            int ia = nA - 1, ib = nB - 1;
            while ((ia >= 0 || ib >= 0) && count < kp) {
                const int32_t *r;
                int fromB;
                // newest first; at equal key the second haplotype pushed last
                if (ib >= 0 && (ia < 0 || eB[(size_t)ib * rec] >= eA[(size_t)ia * rec])) {
                    r = eB + (size_t)ib-- * rec;
                    fromB = 1;
                } else {
                    r = eA + (size_t)ia-- * rec;
                    fromB = 0;
                }
                const int nd = r[1] & 0xffff, nu = r[1] >> 16;
                if (o < nu) {
                    const int x = r[2 + c.depth + o];
                    const uint64_t m = 1ull << (x & 63);
                    if (!(bm[x >> 6] & m)) {
                        bm[x >> 6] |= m;
                        if (ins) ins[count] = x | (o << 20) | (fromB << 26) | (1 << 27);
                        if (ins_site) ins_site[count] = r[0] >> 1;
                        count++;
                    }
                }
                if (o < nd && count < kp) {
                    const int x = r[2 + o];
                    const uint64_t m = 1ull << (x & 63);
                    if (!(bm[x >> 6] & m)) {
                        bm[x >> 6] |= m;
                        if (ins) ins[count] = x | (o << 20) | (fromB << 26);
                        if (ins_site) ins_site[count] = r[0] >> 1;
                        count++;
                    }
                }
            }
        }
    }
    __syncthreads();
    // sorted emission: wave prefix over per-word popcounts
    int base = 0;
    const int row = c.sample_row ? c.sample_row[t] : t;
    int32_t *out = c.out_states + (size_t)row * c.out_stride;
    for (int w0 = 0; w0 < Wn; w0 += 64) {
        const int wi = w0 + lane;
        const uint64_t x = wi < Wn ? bm[wi] : 0ull;
        const int pc = __popcll(x);
        int incl = pc;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int y = __shfl_up(incl, d);
            if (lane >= d) incl += y;
        }
        int p = base + incl - pc;
        uint64_t v = x;
        while (v) {
            const int b = __ffsll((long long)v) - 1;
            out[p++] = wi * 64 + b;
            v &= v - 1;
        }
        base += __shfl(incl, 63);
    }
    if (lane == 0) c.out_K[row] = base;
}

void pbwt_select_launch(gh_panel *P, const PbwtRun &R, hipStream_t s) {
    auto *D = P->pbwt;
    GH_CHECK(D, GH_ERR_ARG, "panel was created without a sparse-PBWT stream");
    WalkArgs w;
    w.ops = D->ops.as<SiteOp>(); w.n_ops = D->n_ops;
    w.bits = D->bits.as<uint64_t>(); w.cnt = D->cnt.as<uint32_t>();
    w.cbits = D->cbits.as<uint64_t>(); w.ccnt = D->ccnt.as<uint32_t>();
    w.N = P->N;
    w.aidx = D->aidx.as<int32_t>();
    w.cm0 = P->cm_pos[0];
    w.haps = R.haps; w.hap_row = R.hap_row; w.L = P->L; w.n_haps = R.n_haps; w.start = R.start;
    w.stored = R.stored; w.hap_class = R.hap_class; w.modulo = R.modulo; w.Lhq = P->Lhq;
    GH_CHECK(R.ev_ci, GH_ERR_ARG, "pbwt selection: no record index buffer");
    w.depth = R.depth; w.ev = R.ev; w.ev_ci = R.ev_ci; w.cap = R.cap; w.ev_cnt = R.ev_cnt; w.status = R.status;
    w.n_tiles = D->n_tiles;
    if (D->n_tiles > 0) {
        const int64_t *t = D->tiles.as<int64_t>();
        w.tile_op = t;
        w.tile_r = t + D->n_tiles + 1;
        w.tile_c = t + 2 * (D->n_tiles + 1);
        w.tile_s = t + 3 * (D->n_tiles + 1);
        w.tsi = D->tsi.as<int32_t>();
    }
    const dim3 gw((R.n_haps + WALK_WALKERS - 1) / WALK_WALKERS);
    if (D->n_tiles > 0 && D->stage_common) {
        static const bool attr = [] {
            GH_HIP(hipFuncSetAttribute((const void *)pbwt_walk_kernel<true, false>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)WALK_LDS));
            return true;
        }();
        (void)attr;
        hipLaunchKernelGGL((pbwt_walk_kernel<true, false>), gw, dim3(WALK_NT), WALK_LDS, s, w);
    } else if (D->n_tiles > 0) {
        static const bool attr = [] {
            GH_HIP(hipFuncSetAttribute((const void *)pbwt_walk_kernel<true, true>,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)WALK_LDS));
            return true;
        }();
        (void)attr;
        hipLaunchKernelGGL((pbwt_walk_kernel<true, true>), gw, dim3(WALK_NT), WALK_LDS, s, w);
    } else {
        hipLaunchKernelGGL((pbwt_walk_kernel<false, false>), dim3((R.n_haps + 63) / 64), dim3(64), 0, s, w);
    }
    GH_HIP(hipGetLastError());
    {
        ResolveArgs ra;
        ra.cols = D->cols.as<ColInfo>(); ra.bits = D->cbits.as<uint64_t>(); ra.cnt = D->ccnt.as<uint32_t>();
        ra.A = D->A.as<int32_t>(); ra.aidx = D->aidx.as<int32_t>(); ra.sfin = D->sfin.as<int32_t>();
        ra.N = P->N; ra.W = (P->N + 63) / 64;
        ra.ev = R.ev; ra.ev_ci = R.ev_ci; ra.ev_cnt = R.ev_cnt; ra.cap = R.cap; ra.depth = R.depth;
        ra.n_haps = R.n_haps;
        hipLaunchKernelGGL(pbwt_resolve_kernel, dim3(R.n_haps), dim3(256), 0, s, ra);
        GH_HIP(hipGetLastError());
    }
    if (R.out_states) {
        CompactArgs c;
        c.ev = R.ev; c.ev_cnt = R.ev_cnt; c.cap = R.cap; c.depth = R.depth; c.N = P->N;
        c.kpbwt = R.kpbwt; c.sample_hap = R.sample_hap; c.sample_one = R.sample_one; c.sample_row = R.sample_row; c.out_states = R.out_states; c.out_stride = R.out_stride;
        c.out_K = R.out_K;
        c.ins = R.ins;
        c.ins_site = R.ins_site;
        const size_t shm = sizeof(uint64_t) * ((P->N + 63) / 64);
        hipLaunchKernelGGL(pbwt_compact_kernel, dim3(R.n_samples), dim3(64), shm, s, c);
        GH_HIP(hipGetLastError());
    }
}

int pbwt_event_cap(gh_panel *P, float modulo, int n_groups) {
    const float len = P->cm_pos.back() - P->cm_pos.front();
    const double c = (double)n_groups + (P->pbwt ? P->pbwt->n_runs : 0) + 8.0 * (len / modulo) + 64.0;
    return (int)std::min(2.0 * c, 1.0e7);
}

// allocatePBWT group boundaries and modulo auto-adjust (haplotype_set.cpp:229-303)
float pbwt_groups(gh_panel *P, int depth, float modulo_in, int kpbwt, std::vector<int> &grp) {
    float modulo = modulo_in;
    const float length = P->cm_pos.back() - P->cm_pos[0];
    if ((length / modulo) * 2 * depth < kpbwt)
        while ((length / modulo) * 2 * depth < kpbwt && modulo > 0.02) modulo /= 2;
    grp.clear();
    for (int l_hq = 0, l_all = 0, src = 0; l_all < P->Lc; l_all++) {
        const int k = P->common2tot[l_all];
        if (P->flags[k] & SF_LQ) continue;
        const int tmp = (int)std::round(P->cm_pos[k] / modulo);
        if (src != tmp) { src = tmp; if (l_hq > 0) grp.push_back(l_hq); }
        ++l_hq;
    }
    if (grp.empty() || grp.back() < P->Lhq) grp.push_back(P->Lhq);
    return modulo;
}

}  // namespace gh
