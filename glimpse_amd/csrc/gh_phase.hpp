// gh_phase.hpp -- host/device structures of the phasing HMM.
#pragma once
#include <stdint.h>

#include <vector>

#include "gh_internal.hpp"

namespace gh {

// One phasing locus (phasing_hmm.cpp:71-96): rel = row in the conditioning
// set, abs = site, type = het ordinal mod 3 (0..2), -1 peak hom, -2 flat het.
struct Locus {
    int32_t rel;
    int32_t abs;
    int32_t seg;
    int32_t miss;
    int8_t type;
    uint8_t alt;
    uint8_t pad[2];
};

// Per-target phasing work item.
struct PhItem {
    int32_t cond;        // CondItem index
    int32_t tgt;         // target index (keys the counter draws)
    int32_t n_loci, n_segs, n_miss;
    int32_t msg;         // floats per stored message: 8 per state, states rounded up to 64
    int64_t h0_off;      // row of H0 in the H0 array
    int64_t h1_off;      // row of H1 in the H1 array
    int64_t flat_off;    // row of the flat flags
    int64_t locus_off;   // into loci (Lp reserved)
    int64_t seg_off;     // floats: n_segs * KH
    int64_t segsum_off;  // floats: n_segs * 9
    int64_t imp_off;     // floats: n_miss * KH
    int64_t impsum_off;  // floats: n_miss * 8 (also imputeProbOf1s)
    int64_t dip_off;     // ints: n_segs
};

struct PhaseBufs {
    float *seg, *segsum, *imp, *impsum, *of1;
    int32_t *dip;
};

void phase_prep_launch(int n, PhItem *d_items, const CondItem *d_cond, const int32_t *d_poly, const uint8_t *d_H0,
                       const uint8_t *d_H1, const uint8_t *d_flat, const uint8_t *d_flags, Locus *d_loci,
                       hipStream_t s);
// Launch shape of the phasing kernel: NW waves per target, KPT states per thread.
struct PhaseShape {
    int NW, KPT;
};
// Shapes the kernel is built for that fit maxK states (the tuner's candidates).
std::vector<PhaseShape> phase_shapes(int maxK);
// The shape for maxK states: a GH_PHASE_SHAPE pin if it fits, else the
// throughput default of the class, or with latency the latency default.
PhaseShape phase_shape(int maxK, bool latency = false);
// sh.NW == 0: phase_shape(maxK).  d_order (optional): workgroup b runs item d_order[b].
void phase_launch(gh_panel *P, int n, int maxK, const PhItem *d_items, const CondItem *d_cond,
                  const uint64_t *d_hvar, const Locus *d_loci, const int32_t *d_pos, int ne, PhaseBufs &B,
                  uint8_t *d_H0, uint8_t *d_H1, float err_phase, uint32_t seed, uint32_t iter, int32_t *d_status,
                  hipStream_t s, PhaseShape sh = PhaseShape{0, 0}, const int32_t *d_order = nullptr,
                  GraphGate gate = GraphGate{});
// Padded state count of a shape.
int phase_kpt_states(PhaseShape sh);
// Floats per stored message of a K-state target (8 per state, K rounded up to 64).
inline int32_t phase_msg_floats(int K) { return ((K + 63) / 64) * 64 * 8; }

}  // namespace gh
