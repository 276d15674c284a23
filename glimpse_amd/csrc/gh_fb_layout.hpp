// gh_fb_layout.hpp -- launch shapes and checkpoint layout of the haploid
// forward-backward kernel (gh_impute.hip).  Plain C++ with no HIP types, so
// the host logic that sizes checkpoint buffers and picks launch shapes can be
// compiled and checked on a machine without a GPU.
//
// A launch runs NW waves per haplotype item with MWL 64-state words per lane;
// NW x MWL words (the capacity class: 4, 8, 16 or 32 words) cover the largest
// conditioning set of the launch.  The forward pass writes one alpha vector
// (64 W floats) every C polymorphic sites; C is the longest block that fits
// the LDS budget of the shape.  Buffers are sized with fb_ckpt_floats for the
// largest word count a run can launch (ckpt_maxW); a launch is valid only when
// its block length is at least the one the buffers were sized for, because a
// shorter block writes more checkpoints than the buffer holds.
#pragma once
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

namespace gh {

constexpr int FB_TILE = 32;               // sites per metadata / Hvar tile
constexpr int FB_LDS_BYTES = 156 * 1024;  // LDS a workgroup may use
constexpr int FB_WG_PER_CU = 3;           // workgroups per CU the block length leaves room for
// LDS words before the Hvar tile: the row-word reads of word slots a wave does
// not fill may point up to NW*MWL (<= 64) words below a row
constexpr int FB_ROW_PAD = 64;
constexpr int FB_MAX_WORDS = 64;          // largest conditioning set: 64 words, 4096 states

// Launch shape: NW waves per haplotype, MWL 64-state words per lane.  {0, 0}: default.
struct FbPick {
    int NW, MWL;
};

struct FbShape {
    int NW, MWL, C;
    size_t smem;
};

// Shapes the kernel is built for (NW waves x MWL words per lane covers 64*NW*MWL states).
static const int kFbShapes[][2] = {
    {1, 4}, {2, 2}, {4, 1},                     // W <= 4
    {1, 8}, {2, 4}, {4, 2}, {8, 1},             // W <= 8
    {1, 16}, {2, 8}, {4, 4}, {8, 2}, {16, 1},   // W <= 16
    {2, 16}, {4, 8}, {8, 4}, {16, 2},           // W <= 32 (one wave of 32 words spills)
    {8, 8}, {16, 4},                            // W <= 64
};
constexpr int kNumFbShapes = (int)(sizeof(kFbShapes) / sizeof(kFbShapes[0]));

inline FbShape fb_shape_of(int maxW, int NW, int MWL) {
    FbShape s;
    s.NW = NW;
    s.MWL = MWL;
    const size_t rows = (size_t)(FB_ROW_PAD + FB_TILE * maxW) * 8;
    const size_t per_site = (size_t)s.NW * s.MWL * 64 * 4;
    const size_t red = 2 * 16 * 4 * 4 + FB_TILE * 3 * 16;   // [2][NW <= 16][4] floats, then the site records
    // The longest block (C divides FB_TILE) that leaves room for FB_WG_PER_CU
    // workgroups per CU (what the register budget of the 4-wave shapes allows),
    // unless that would make blocks shorter than 4 sites; else the longest that fits.
    int C = FB_TILE;
    while (C > 4 && rows + C * per_site + red > (size_t)FB_LDS_BYTES / FB_WG_PER_CU) C >>= 1;
    while (C > 1 && rows + C * per_site + red > (size_t)FB_LDS_BYTES) C >>= 1;
    s.C = C;
    s.smem = rows + C * per_site + red;
    return s;
}

inline int fb_cap_words(int maxW) { return maxW <= 4 ? 4 : maxW <= 8 ? 8 : maxW <= 16 ? 16 : maxW <= 32 ? 32 : 64; }

// The shape of a launch whose largest set has maxW words: an explicit pick of
// the capacity class, else a pinned "<NW>x<MWL>" string (pin, may be NULL) of
// the class, else the default: the throughput shape (the fewest waves per item
// that fill a CU, measured with 1000 items in flight) or, with latency, the
// one that shortens a single item's chain.
inline FbShape fb_shape_pinned(int maxW, FbPick pick, bool latency, const char *pin) {
    const int cap = fb_cap_words(maxW);
    if (pick.NW > 0 && pick.NW * pick.MWL == cap) return fb_shape_of(maxW, pick.NW, pick.MWL);
    if (pin) {
        int nw = 0, mwl = 0;
        if (std::sscanf(pin, "%dx%d", &nw, &mwl) == 2 && nw * mwl == cap)
            for (const auto &sh : kFbShapes)
                if (sh[0] == nw && sh[1] == mwl) return fb_shape_of(maxW, nw, mwl);
    }
    if (cap == 4) return fb_shape_of(maxW, 1, 4);
    if (cap == 8) return fb_shape_of(maxW, 2, 4);
    if (cap == 16) return fb_shape_of(maxW, 4, 4);
    if (cap == 32) return latency ? fb_shape_of(maxW, 8, 4) : fb_shape_of(maxW, 4, 8);
    return latency ? fb_shape_of(maxW, 16, 4) : fb_shape_of(maxW, 8, 8);
}

// GH_FB_SHAPE=<NW>x<MWL> pins a shape wherever it fits (testing, reproducing a tuned run).
inline FbShape fb_shape(int maxW, FbPick pick = FbPick{0, 0}, bool latency = false) {
    return fb_shape_pinned(maxW, pick, latency, std::getenv("GH_FB_SHAPE"));
}

inline FbPick fb_default_pick(int maxW, bool latency = false) {
    const FbShape sh = fb_shape(maxW, FbPick{0, 0}, latency);
    return FbPick{sh.NW, sh.MWL};
}

inline std::vector<FbPick> fb_shapes(int maxW) {
    std::vector<FbPick> v;
    for (const auto &sh : kFbShapes)
        if (sh[0] * sh[1] == fb_cap_words(maxW)) v.push_back(FbPick{sh[0], sh[1]});
    return v;
}

// The checkpoint interval only depends on the words a shape covers (NW x MWL),
// which every shape of a capacity class shares, and on maxW (the LDS row tile).
inline int fb_block_sites(int maxW) { return fb_shape_pinned(maxW, FbPick{0, 0}, false, nullptr).C; }

// Checkpoint floats of one haplotype item with up to Lp polymorphic sites and W
// words: every C-th alpha vector (64 W floats), then the forward pass's
// scaling sum of every site (Lp floats).
inline int64_t fb_ckpt_floats(int Lp, int W, int maxW) {
    const int C = fb_block_sites(maxW);
    return (int64_t)((Lp + C - 1) / C) * 64 * W + Lp;
}

// The host-side guard of a launch: a launch of up to maxW words with shape sh
// may write into checkpoint buffers sized for ckpt_maxW words.
inline bool fb_launch_fits(int maxW, int ckpt_maxW, const FbShape &sh) {
    return maxW <= ckpt_maxW && sh.C >= fb_block_sites(ckpt_maxW);
}

// Floats a launch with block length C writes for an item of Lp sites and W
// words (its checkpoints, then its Lp scaling sums right after them).
inline int64_t fb_ckpt_written(int Lp, int W, int C) { return (int64_t)((Lp + C - 1) / C) * 64 * W + Lp; }

}  // namespace gh
