/*
 * synth.cpp -- synthetic reference panels and low-coverage targets.
 *
 * Panel model: F founder haplotypes; per candidate SNP the founder allele
 * frequency f has density proportional to 1/f on [1/(2N), 1/2].  Every panel
 * and target haplotype is a Li-Stephens mosaic of founders (switch when the
 * genetic distance since the last switch exceeds an Exp(rho) draw, rho in
 * switches per cM) with per-allele mutation rate mu.  A fraction of candidate
 * sites can instead be "injected rare" variants whose minor allele count is
 * drawn log-uniformly in [1, N*sparse_maf), which gives rare-variant-heavy
 * panels.  Positions are uniform over the chunk at 1 cM/Mb
 * (variant_map::setGeneticMap, common/src/containers/variant_map.cpp:151-158),
 * and monomorphic panel sites are dropped as GLIMPSE's panel reader does
 * (common/src/io/ref_genotype_reader.cpp:168).
 *
 * Targets: held-out mosaics.  Per site read depth ~ Poisson(coverage); each
 * read picks one of the two haplotypes, has a base error e ~ U(1e-3, 1e-2)
 * and flips the allele with probability e.  Genotype likelihoods are
 * log10 P(reads | g), normalised to max 0, converted to phred like
 * genotype_reader.cpp:446-461 (lroundf(-10*GL), cap 255); a site with three
 * equal phred values is flat.
 */
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include <sys/mman.h>

namespace {

struct Xoshiro {
    uint64_t s[4];
    static uint64_t splitmix(uint64_t &x) {
        uint64_t z = (x += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    explicit Xoshiro(uint64_t seed) { for (int i = 0; i < 4; i++) s[i] = splitmix(seed); }
    static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }
    inline uint64_t next() {
        const uint64_t r = rotl(s[1] * 5, 7) * 9;
        const uint64_t t = s[1] << 17;
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
        s[2] ^= t; s[3] = rotl(s[3], 45);
        return r;
    }
    inline double u01() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    inline uint32_t below(uint32_t n) { return (uint32_t)(((next() >> 32) * (uint64_t)n) >> 32); }
    inline double expo(double rate) { return -std::log(1.0 - u01()) / rate; }
    int poisson(double lam) {
        if (lam < 30.0) {
            double L = std::exp(-lam), p = 1.0; int k = 0;
            do { k++; p *= u01(); } while (p > L);
            return k - 1;
        }
        double x = lam + std::sqrt(lam) * std::sqrt(-2.0 * std::log(1.0 - u01())) * std::cos(6.283185307179586 * u01());
        return std::max(0, (int)std::lround(x));
    }
};

/* common/src/io/ref_genotype_reader.cpp:171-175 */
static bool is_common_site(uint32_t cref, uint32_t calt, float sparse_maf) {
    double MAF = std::min(cref * 1.0f / (cref + calt), calt * 1.0f / (cref + calt));
    return MAF >= sparse_maf;
}

struct SynthPanel {
    int L = 0, N = 0, T = 0, W = 0;
    std::vector<int32_t> bp;
    std::vector<double> cm;
    std::vector<uint32_t> cref, calt;
    std::vector<uint8_t> common, major;
    std::vector<uint64_t> common_words;   /* [n_common][W] */
    std::vector<int32_t> rare_ptr, rare_idx;
    std::vector<uint8_t> tgt;             /* [2T][L] */
    std::vector<uint8_t> ref_alleles;     /* optional dense [N][L] for small panels */
    std::string err;
};

static inline uint64_t mix64(uint64_t a, uint64_t b) {
    uint64_t x = a ^ (b + 0x9E3779B97F4A7C15ull + (a << 6) + (a >> 2));
    return Xoshiro::splitmix(x);
}

template <typename F>
static void parallel_for(int n, int n_threads, F &&f) {
    n_threads = std::max(1, std::min(n_threads, n));
    if (n_threads == 1) { for (int i = 0; i < n; i++) f(i); return; }
    std::vector<std::thread> th;
    std::atomic<int> next(0);
    for (int t = 0; t < n_threads; t++)
        th.emplace_back([&]() {
            for (;;) {
                const int i = next.fetch_add(1);
                if (i >= n) break;
                f(i);
            }
        });
    for (auto &x : th) x.join();
}

// Large zeroed scratch arrays, mapped with their pages populated up front:
// touching gigabytes of fresh pages one fault at a time costs far more than
// the generator's own work on some hosts.
template <typename T>
struct Populated {
    T *p = nullptr;
    size_t n = 0;
    explicit Populated(size_t count) : n(count) {
        void *m = mmap(nullptr, std::max<size_t>(n, 1) * sizeof(T), PROT_READ | PROT_WRITE,
                       MAP_PRIVATE | MAP_ANONYMOUS | MAP_POPULATE, -1, 0);
        if (m == MAP_FAILED) throw std::bad_alloc();
        p = (T *)m;
    }
    ~Populated() { if (p) munmap(p, std::max<size_t>(n, 1) * sizeof(T)); }
    Populated(const Populated &) = delete;
    Populated &operator=(const Populated &) = delete;
    T *data() { return p; }
    T &operator[](size_t i) { return p[i]; }
};

// 64x64 bit transpose (Hacker's Delight 7-3, recursive block swaps): bit j of
// a[i] moves to bit i of a[j].
static void transpose64(uint64_t a[64]) {
    uint64_t m = 0x00000000FFFFFFFFull;
    for (int j = 32; j != 0; j >>= 1, m ^= (m << j)) {
        for (int k = 0; k < 64; k = ((k | j) + 1) & ~j) {
            const uint64_t t = ((a[k] >> j) ^ a[k | j]) & m;
            a[k] ^= t << j;
            a[k | j] ^= t;
        }
    }
}

} // namespace

extern "C" {

typedef struct {
    int32_t n_haps;        /* N panel haplotypes */
    int32_t n_sites;       /* L kept (polymorphic) sites */
    int32_t n_targets;     /* T diploid targets (2T held-out haplotypes) */
    int32_t n_founders;    /* F */
    double length_mb;      /* chunk length; 1 cM per Mb */
    int32_t start_bp;
    double rho_per_cm;     /* founder switches per cM */
    double mu;             /* per-allele mutation rate */
    double rare_frac;      /* fraction of candidate sites injected as rare variants */
    double flip_frac;      /* fraction of sites whose ALT is the major allele */
    float sparse_maf;      /* common/rare threshold (0.001) */
    int32_t keep_dense;    /* also keep a dense [N][L] allele matrix (small panels only) */
    uint64_t seed;
} gh_synth_cfg;

void *gh_synth_panel_new(const gh_synth_cfg *c) {
    SynthPanel *P = new SynthPanel();
    const int N = c->n_haps, L = c->n_sites, T = c->n_targets, F = std::max(1, c->n_founders);
    const int H = N + 2 * T;
    P->L = L; P->N = N; P->T = T; P->W = (N + 63) / 64;
    Xoshiro rng(c->seed);
    /* candidate positions: generous oversampling, sorted, distinct */
    const double len_bp = c->length_mb * 1e6;
    int M = (int)(L * 1.6) + 64;
    std::vector<int32_t> cand(M);
    for (int i = 0; i < M; i++) cand[i] = c->start_bp + (int32_t)(rng.u01() * len_bp);
    std::sort(cand.begin(), cand.end());
    for (int i = 1; i < M; i++) if (cand[i] <= cand[i - 1]) cand[i] = cand[i - 1] + 1;

    std::vector<int> founder(H);
    std::vector<double> next_sw(H);
    for (int h = 0; h < H; h++) { founder[h] = (int)rng.below(F); next_sw[h] = rng.expo(c->rho_per_cm); }
    std::vector<uint8_t> fa(F), col(H);
    P->rare_ptr.push_back(0);
    P->tgt.assign((size_t)2 * T * L, 0);
    if (c->keep_dense) P->ref_alleles.assign((size_t)N * L, 0);
    const double lo = std::log(1.0 / (2.0 * N)), hi = std::log(0.5);
    const double first_cm = cand[0] / 1e6;
    int kept = 0;
    for (int i = 0; i < M && kept < L; i++) {
        const double cmv = cand[i] / 1e6 - first_cm;
        for (int h = 0; h < H; h++) {
            if (cmv >= next_sw[h]) { founder[h] = (int)rng.below(F); next_sw[h] = cmv + rng.expo(c->rho_per_cm); }
        }
        const bool injected = rng.u01() < c->rare_frac;
        if (!injected) {
            const double f = std::exp(lo + rng.u01() * (hi - lo));
            for (int j = 0; j < F; j++) fa[j] = rng.u01() < f;
            for (int h = 0; h < H; h++) col[h] = fa[founder[h]];
            /* mutations: Binomial(H, mu) flips by geometric skipping */
            if (c->mu > 0) {
                double lq = std::log(1.0 - c->mu);
                long pos = -1;
                for (;;) {
                    pos += 1 + (long)std::floor(std::log(1.0 - rng.u01()) / lq);
                    if (pos >= H) break;
                    col[pos] ^= 1;
                }
            }
        } else {
            std::fill(col.begin(), col.end(), 0);
            const double mmax = std::max(1.0, N * (double)c->sparse_maf);
            int mac = (int)std::floor(std::exp(rng.u01() * std::log(mmax)));
            mac = std::max(1, std::min(mac, H - 1));
            for (int j = 0; j < mac; j++) col[rng.below(H)] = 1;
        }
        if (rng.u01() < c->flip_frac) for (int h = 0; h < H; h++) col[h] ^= 1;
        uint32_t ones = 0;
        for (int h = 0; h < N; h++) ones += col[h];
        if (ones == 0 || ones == (uint32_t)N) continue;  /* monomorphic in the panel: dropped */
        const uint32_t ca = ones, cr = (uint32_t)N - ones;
        const bool com = is_common_site(cr, ca, c->sparse_maf);
        const bool maj = ca > cr;
        P->bp.push_back(cand[i]);
        P->cref.push_back(cr); P->calt.push_back(ca);
        P->common.push_back(com); P->major.push_back(maj);
        if (com) {
            size_t base = P->common_words.size();
            P->common_words.resize(base + P->W, 0ull);
            for (int h = 0; h < N; h++) if (col[h]) P->common_words[base + (h >> 6)] |= 1ull << (h & 63);
        } else {
            for (int h = 0; h < N; h++) if (col[h] != (uint8_t)maj) P->rare_idx.push_back(h);
        }
        P->rare_ptr.push_back((int32_t)P->rare_idx.size());
        for (int t = 0; t < 2 * T; t++) P->tgt[(size_t)t * L + kept] = col[N + t];
        if (c->keep_dense) for (int h = 0; h < N; h++) P->ref_alleles[(size_t)h * L + kept] = col[h];
        kept++;
    }
    if (kept < L) { P->err = "synth: not enough polymorphic sites"; P->L = kept; }
    /* genetic positions: cM = bp / 1e6, baseline subtracted (variant_map.cpp:151-158) */
    P->cm.resize(P->L);
    for (int l = 0; l < P->L; l++) P->cm[l] = P->bp[l] * 1.0 / 1e6;
    const double base = P->L ? P->cm[0] : 0.0;
    for (int l = 0; l < P->L; l++) P->cm[l] -= base;
    return P;
}

/* Founder-segment generator for large panels (the same model as
 * gh_synth_panel_new, organised for N in the 1e5 range).  Site-level draws
 * (frequency spectrum, founder alleles, injected rare variants, mutations,
 * ALT flips) come from a per-site stream and haplotype switch points from a
 * per-haplotype stream, so the work splits over n_threads threads and the
 * result does not depend on their number.  Each haplotype's alleles are built
 * 64 sites at a time from its founder segments (OR of the founder's allele
 * word, masked to the segment), mutations flip single bits, and a 64x64 bit
 * transpose gives the site-major rows. */
void *gh_synth_panel_fast(const gh_synth_cfg *c, int32_t n_threads) {
    SynthPanel *P = new SynthPanel();
    const int N = c->n_haps, L = c->n_sites, T = c->n_targets, F = std::max(1, c->n_founders);
    const int H = N + 2 * T;
    P->L = L; P->N = N; P->T = T; P->W = (N + 63) / 64;
    if (F > 64) { P->err = "synth: the fast generator takes at most 64 founders"; P->L = 0; return P; }
    if (c->keep_dense) { P->err = "synth: the fast generator keeps no dense matrix"; P->L = 0; return P; }
    Xoshiro rng(c->seed);
    const double len_bp = c->length_mb * 1e6;
    const int M = (int)(L * 1.6) + 64;
    std::vector<int32_t> cand(M);
    for (int i = 0; i < M; i++) cand[i] = c->start_bp + (int32_t)(rng.u01() * len_bp);
    std::sort(cand.begin(), cand.end());
    for (int i = 1; i < M; i++) if (cand[i] <= cand[i - 1]) cand[i] = cand[i - 1] + 1;
    const double first_cm = cand[0] / 1e6;
    std::vector<double> cmv(M);
    for (int i = 0; i < M; i++) cmv[i] = cand[i] / 1e6 - first_cm;
    const int MB = (M + 63) / 64, HW = (H + 63) / 64;
    // ---- per-site draws
    std::vector<uint64_t> famask(M, 0ull);
    std::vector<uint8_t> injected(M, 0), flip(M, 0);
    std::vector<std::vector<int32_t>> hits(M);   // mutated haplotypes, or the carriers of an injected site
    const double lo = std::log(1.0 / (2.0 * N)), hi = std::log(0.5);
    const double lq = c->mu > 0 ? std::log(1.0 - c->mu) : 0.0;
    parallel_for(M, n_threads, [&](int i) {
        Xoshiro r(mix64(c->seed, 0x5173ull + (uint64_t)i));
        injected[i] = r.u01() < c->rare_frac;
        if (!injected[i]) {
            const double f = std::exp(lo + r.u01() * (hi - lo));
            uint64_t m = 0;
            for (int j = 0; j < F; j++) if (r.u01() < f) m |= 1ull << j;
            famask[i] = m;
            if (c->mu > 0) {
                long pos = -1;
                for (;;) {
                    pos += 1 + (long)std::floor(std::log(1.0 - r.u01()) / lq);
                    if (pos >= H) break;
                    hits[i].push_back((int32_t)pos);
                }
            }
        } else {
            const double mmax = std::max(1.0, N * (double)c->sparse_maf);
            int mac = (int)std::floor(std::exp(r.u01() * std::log(mmax)));
            mac = std::max(1, std::min(mac, H - 1));
            for (int j = 0; j < mac; j++) hits[i].push_back((int32_t)r.below(H));
        }
        flip[i] = r.u01() < c->flip_frac;
    });
    // ---- founder allele words: fbits[f][b] bit j = allele of founder f at site 64b+j
    std::vector<uint64_t> fbits((size_t)F * MB, 0ull);
    for (int i = 0; i < M; i++) {
        uint64_t m = famask[i];
        while (m) {
            const int f = __builtin_ctzll(m);
            m &= m - 1;
            fbits[(size_t)f * MB + (i >> 6)] |= 1ull << (i & 63);
        }
    }
    // ---- haplotype-major words from founder segments
    auto *hwp = new Populated<uint64_t>((size_t)H * MB);
    Populated<uint64_t> &hw = *hwp;
    parallel_for((H + 255) / 256, n_threads, [&](int hb) {
        for (int h = hb * 256; h < std::min(H, hb * 256 + 256); h++) {
            Xoshiro r(mix64(c->seed, 0xA11CE00000ull + (uint64_t)h));
            uint64_t *row = hw.data() + (size_t)h * MB;
            int fo = (int)r.below(F);
            double next_sw = r.expo(c->rho_per_cm);
            int a = 0;
            while (a < M) {
                // first site at or after the switch point (one switch per site at
                // most): galloping search from a, segments are short
                int b;
                {
                    int lo2 = a + 1, step = 1;
                    while (lo2 + step < M && cmv[lo2 + step] < next_sw) { lo2 += step; step <<= 1; }
                    if (lo2 < M && cmv[lo2] >= next_sw) b = lo2;
                    else b = (int)(std::lower_bound(cmv.begin() + lo2, cmv.begin() + std::min(M, lo2 + step + 1),
                                                    next_sw) - cmv.begin());
                }
                // sites [a, b) copy founder fo
                const uint64_t *src = fbits.data() + (size_t)fo * MB;
                for (int blk = a >> 6; blk <= (b - 1) >> 6; blk++) {
                    const int s0 = std::max(a, blk * 64) - blk * 64, s1 = std::min(b, blk * 64 + 64) - blk * 64;
                    const uint64_t mk = (s1 - s0 == 64) ? ~0ull : (((1ull << (s1 - s0)) - 1ull) << s0);
                    row[blk] |= src[blk] & mk;
                }
                if (b >= M) break;
                fo = (int)r.below(F);
                next_sw = cmv[b] + r.expo(c->rho_per_cm);
                a = b;
            }
        }
    });
    // ---- mutations and injected carriers (each thread owns a range of 64-site blocks)
    parallel_for(MB, n_threads, [&](int blk) {
        for (int i = blk * 64; i < std::min(M, blk * 64 + 64); i++)
            for (int32_t h : hits[i]) {
                uint64_t &w = hw[(size_t)h * MB + (i >> 6)];
                if (injected[i]) w |= 1ull << (i & 63); else w ^= 1ull << (i & 63);
            }
    });
    // ---- transpose to site-major rows of HW words
    Populated<uint64_t> sm((size_t)M * HW);
    // tiles of 8 x 8 blocks, so every cache line read or written is used whole
    const int TB = 8, nsb = (MB + TB - 1) / TB, nhb = (HW + TB - 1) / TB;
    parallel_for(nsb * nhb, n_threads, [&](int tile) {
        const int sb0 = (tile / nhb) * TB, hb0 = (tile % nhb) * TB;
        uint64_t a[64];
        for (int sb = sb0; sb < std::min(MB, sb0 + TB); sb++)
            for (int hb = hb0; hb < std::min(HW, hb0 + TB); hb++) {
                for (int k = 0; k < 64; k++) {
                    const int h = hb * 64 + k;
                    a[k] = h < H ? hw[(size_t)h * MB + sb] : 0ull;
                }
                transpose64(a);
                for (int k = 0; k < 64; k++) {
                    const int i = sb * 64 + k;
                    if (i < M) sm[(size_t)i * HW + hb] = a[k];
                }
            }
    });
    delete hwp;
    // ---- flips, monomorphic filter, classification
    P->rare_ptr.push_back(0);
    P->tgt.assign((size_t)2 * T * L, 0);
    const int W = P->W;
    const uint64_t lastN = (N & 63) ? ((1ull << (N & 63)) - 1ull) : ~0ull;
    int kept = 0;
    for (int i = 0; i < M && kept < L; i++) {
        uint64_t *row = sm.data() + (size_t)i * HW;
        if (flip[i]) {
            for (int w = 0; w < HW; w++) row[w] = ~row[w];
            if (H & 63) row[HW - 1] &= (1ull << (H & 63)) - 1ull;
        }
        uint32_t ones = 0;
        for (int w = 0; w < W - 1; w++) ones += (uint32_t)__builtin_popcountll(row[w]);
        ones += (uint32_t)__builtin_popcountll(row[W - 1] & lastN);
        if (ones == 0 || ones == (uint32_t)N) continue;
        const uint32_t ca = ones, cr = (uint32_t)N - ones;
        const bool com = is_common_site(cr, ca, c->sparse_maf);
        const bool maj = ca > cr;
        P->bp.push_back(cand[i]);
        P->cref.push_back(cr); P->calt.push_back(ca);
        P->common.push_back(com); P->major.push_back(maj);
        if (com) {
            const size_t base = P->common_words.size();
            P->common_words.insert(P->common_words.end(), row, row + W);
            P->common_words[base + W - 1] &= lastN;
        } else {
            for (int w = 0; w < W; w++) {
                uint64_t x = maj ? ~row[w] : row[w];
                if (w == W - 1) x &= lastN;
                while (x) {
                    const int b = __builtin_ctzll(x);
                    x &= x - 1;
                    P->rare_idx.push_back(w * 64 + b);
                }
            }
        }
        P->rare_ptr.push_back((int32_t)P->rare_idx.size());
        for (int t = 0; t < 2 * T; t++) {
            const int h = N + t;
            P->tgt[(size_t)t * L + kept] = (uint8_t)((row[h >> 6] >> (h & 63)) & 1ull);
        }
        kept++;
    }
    if (kept < L) { P->err = "synth: not enough polymorphic sites"; P->L = kept; }
    P->cm.resize(P->L);
    for (int l = 0; l < P->L; l++) P->cm[l] = P->bp[l] * 1.0 / 1e6;
    const double base = P->L ? P->cm[0] : 0.0;
    for (int l = 0; l < P->L; l++) P->cm[l] -= base;
    return P;
}

void gh_synth_panel_free(void *p) { delete (SynthPanel *)p; }
const char *gh_synth_error(void *p) { return ((SynthPanel *)p)->err.c_str(); }

void gh_synth_panel_sizes(void *p, int32_t *L, int32_t *N, int32_t *n_common, int64_t *n_rare_idx) {
    SynthPanel *P = (SynthPanel *)p;
    *L = P->L; *N = P->N;
    *n_common = (int32_t)(P->common_words.size() / (P->W ? P->W : 1));
    *n_rare_idx = (int64_t)P->rare_idx.size();
}

void gh_synth_panel_copy(void *p, int32_t *bp, double *cm, uint32_t *cref, uint32_t *calt,
                         uint8_t *common, uint8_t *major, uint64_t *common_words,
                         int32_t *rare_ptr, int32_t *rare_idx, uint8_t *tgt_haps, uint8_t *dense) {
    SynthPanel *P = (SynthPanel *)p;
    const int L = P->L;
    std::memcpy(bp, P->bp.data(), sizeof(int32_t) * L);
    std::memcpy(cm, P->cm.data(), sizeof(double) * L);
    std::memcpy(cref, P->cref.data(), sizeof(uint32_t) * L);
    std::memcpy(calt, P->calt.data(), sizeof(uint32_t) * L);
    std::memcpy(common, P->common.data(), L);
    std::memcpy(major, P->major.data(), L);
    std::memcpy(common_words, P->common_words.data(), sizeof(uint64_t) * P->common_words.size());
    std::memcpy(rare_ptr, P->rare_ptr.data(), sizeof(int32_t) * (L + 1));
    std::memcpy(rare_idx, P->rare_idx.data(), sizeof(int32_t) * P->rare_idx.size());
    /* target haplotypes were laid out with stride n_sites requested; compact to L */
    for (int t = 0; t < 2 * P->T; t++)
        std::memcpy(tgt_haps + (size_t)t * L, P->tgt.data() + (size_t)t * (P->tgt.size() / (2 * P->T ? 2 * P->T : 1)), L);
    if (dense && !P->ref_alleles.empty())
        for (int h = 0; h < P->N; h++)
            std::memcpy(dense + (size_t)h * L, P->ref_alleles.data() + (size_t)h * (P->ref_alleles.size() / P->N), L);
}

/* Reads and genotype likelihoods for T diploid targets.
 * haps: [2T][L]; coverage: [T]; out gl [T][L][3], flat [T][L]. */
int gh_synth_reads(int32_t L, int32_t T, const uint8_t *haps, const float *coverage, uint64_t seed,
                   uint8_t *gl, uint8_t *flat) {
    Xoshiro rng(seed);
    for (int t = 0; t < T; t++) {
        const uint8_t *a = haps + (size_t)(2 * t) * L, *b = haps + (size_t)(2 * t + 1) * L;
        for (int l = 0; l < L; l++) {
            const int d = rng.poisson(coverage[t]);
            double ll[3] = {0.0, 0.0, 0.0};
            for (int r = 0; r < d; r++) {
                const uint8_t tru = (rng.u01() < 0.5) ? a[l] : b[l];
                const double e = 1e-3 + rng.u01() * (1e-2 - 1e-3);
                const uint8_t obs = tru ^ (uint8_t)(rng.u01() < e);
                const double q1 = obs ? 1.0 - e : e;   /* P(obs | allele 1) */
                const double q0 = obs ? e : 1.0 - e;   /* P(obs | allele 0) */
                for (int g = 0; g < 3; g++) ll[g] += std::log10(0.5 * g * q1 + (1.0 - 0.5 * g) * q0);
            }
            const double mx = std::max(ll[0], std::max(ll[1], ll[2]));
            uint8_t *o = gl + ((size_t)t * L + l) * 3;
            for (int g = 0; g < 3; g++) {
                const float glv = (float)(ll[g] - mx);
                long v = lroundf(-10.0f * glv);
                o[g] = (uint8_t)std::min(std::max(v, 0l), 255l);
            }
            flat[(size_t)t * L + l] = (o[0] == o[1] && o[0] == o[2]) ? 1 : 0;
        }
    }
    return 0;
}

/* The same reads with one random stream per target (seeded by seed and the
 * target index), spread over n_threads threads. */
int gh_synth_reads_mt(int32_t L, int32_t T, const uint8_t *haps, const float *coverage, uint64_t seed,
                      uint8_t *gl, uint8_t *flat, int32_t n_threads) {
    parallel_for(T, n_threads, [&](int t) {
        gh_synth_reads(L, 1, haps + (size_t)(2 * t) * L, coverage + t, mix64(seed, 0xBEEF0000ull + (uint64_t)t),
                       gl + (size_t)t * L * 3, flat + (size_t)t * L);
    });
    return 0;
}

} // extern "C"
