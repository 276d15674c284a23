// CPU check of the forward-backward checkpoint layout (glimpse_amd/csrc/gh_fb_layout.hpp).
// Enumerates every launch the run driver and the batch entry point can make --
// the word count of the launch (maxW) up to the count the checkpoint buffers
// were sized for (ckpt_maxW), default / latency / each tuner candidate shape,
// and every GH_FB_SHAPE pin -- and checks that the launch passes the host guard
// and never writes more checkpoint floats than the buffers hold.  Then it shows
// that the guard refuses the launches that would overrun.
#include <cstdio>
#include <string>

#include "gh_fb_layout.hpp"

using namespace gh;

static int fails = 0;
#define CHECK(c, ...)                                       \
    do {                                                    \
        if (!(c)) {                                         \
            if (fails < 20) std::printf("FAIL: " __VA_ARGS__); \
            fails++;                                        \
        }                                                   \
    } while (0)

int main() {
    long launches = 0;
    std::vector<std::string> pins = {""};
    for (int i = 0; i < kNumFbShapes; i++)
        pins.push_back(std::to_string(kFbShapes[i][0]) + "x" + std::to_string(kFbShapes[i][1]));
    const int Lps[] = {1, 2, 3, 4, 5, 7, 8, 9, 31, 32, 33, 63, 64, 65, 1000, 58604};
    for (int ckW = 1; ckW <= FB_MAX_WORDS; ckW++) {
        const int Cl = fb_block_sites(ckW);
        CHECK(Cl >= 1 && FB_TILE % Cl == 0, "layout block %d of ckpt_maxW %d does not divide the tile\n", Cl, ckW);
        for (int maxW = 1; maxW <= ckW; maxW++) {
            std::vector<FbPick> picks = fb_shapes(maxW);
            picks.push_back(FbPick{0, 0});
            for (const FbPick &pk : picks)
                for (int lat = 0; lat < 2; lat++)
                    for (const std::string &pin : pins) {
                        const FbShape sh = fb_shape_pinned(maxW, pk, lat != 0, pin.empty() ? nullptr : pin.c_str());
                        launches++;
                        CHECK(sh.NW * sh.MWL == fb_cap_words(maxW), "shape %dx%d does not cover %d words\n", sh.NW,
                              sh.MWL, maxW);
                        CHECK(sh.smem <= (size_t)FB_LDS_BYTES, "shape %dx%d needs %zu B of LDS\n", sh.NW, sh.MWL,
                              sh.smem);
                        CHECK(FB_TILE % sh.C == 0, "block %d does not divide the tile\n", sh.C);
                        CHECK(fb_launch_fits(maxW, ckW, sh),
                              "launch maxW %d shape %dx%d (C %d) refused for buffers of ckpt_maxW %d (C %d)\n", maxW,
                              sh.NW, sh.MWL, sh.C, ckW, Cl);
                        for (int W = 1; W <= maxW; W++)
                            for (int Lp : Lps)
                                CHECK(fb_ckpt_written(Lp, W, sh.C) <= fb_ckpt_floats(Lp, W, ckW),
                                      "launch C %d writes %lld floats past %lld (Lp %d W %d ckpt_maxW %d)\n", sh.C,
                                      (long long)fb_ckpt_written(Lp, W, sh.C), (long long)fb_ckpt_floats(Lp, W, ckW),
                                      Lp, W, ckW);
                    }
        }
        // launches the guard must refuse: more words than the buffers, or a
        // shorter block than the layout (a shape built for a smaller LDS share)
        for (int maxW = ckW + 1; maxW <= FB_MAX_WORDS; maxW++)
            CHECK(!fb_launch_fits(maxW, ckW, fb_shape_pinned(maxW, FbPick{0, 0}, false, nullptr)),
                  "guard accepts maxW %d > ckpt_maxW %d\n", maxW, ckW);
        if (Cl > 1) {
            FbShape shorter = fb_shape_pinned(ckW, FbPick{0, 0}, false, nullptr);
            shorter.C = Cl / 2;
            CHECK(!fb_launch_fits(ckW, ckW, shorter), "guard accepts block %d under layout block %d\n", shorter.C, Cl);
            CHECK(fb_ckpt_written(Cl + 1, 1, shorter.C) > fb_ckpt_floats(Cl + 1, 1, ckW),
                  "a shorter block would not overrun (Cl %d)\n", Cl);
        }
    }
    std::printf("launches %ld failures %d\n", launches, fails);
    return fails ? 1 : 0;
}
