// gh_graph.hpp -- the per-iteration host decisions of the Gibbs schedule moved
// onto the device, so that an iteration can be captured once as a HIP graph
// and replayed (gh_gibbs.hip, run_execute).
//
// An iteration launched from the host reads three things back: each target's
// polymorphic-site count (to place its Hvar rows, alpha checkpoints and
// phasing loci), the phasing segment counts (to place the messages) and the
// set-size classes (to pick the launch groups and shapes).  Replayed from a
// graph, the same decisions are made by single-workgroup kernels:
//   pbwt_draw_kernel  the stored site of every PBWT group and the start index
//                     of every walked haplotype (the draws the host makes);
//   fb_place_kernel   Hvar / checkpoint / locus offsets from Lp, the two
//                     forward-backward launch groups and their shape keys;
//   ph_place_kernel   message offsets from the segment counts, the two
//                     phasing launch groups and their shape keys.
// Buffers cannot grow inside a graph: they are sized before the capture from
// the iterations launched from the host, with headroom, and a placement that
// needs more sets a sticky overflow flag instead of writing past them (the run
// is then executed again with every iteration launched from the host; the
// counter-keyed draws make that run identical).
//
// A launch group's shape key (capacity class, latency shape) is data: the
// graph holds only the launches of the keys the last host-launched iteration
// used (HIP graphs have no device-side conditional nodes, and a launch held
// only to return at once still waits for a CU with room for its workgroups,
// behind the long-lived workgroups of other runs).  A replay whose keys differ
// is treated like an overflow.
#pragma once
#include <stdint.h>

#include "gh_internal.hpp"
#include "gh_phase.hpp"

namespace gh {

// why an execution left the graph: placement exceeded the captured buffers,
// launch keys changed (both: executed again host-launched), or the runtime
// refused the capture (the iterations were enqueued without it)
enum { GO_CAPACITY = 1, GO_SHAPE = 2, GO_UNCAPTURED = 4 };
enum { GS_FB_STATE_SITES, GS_FB_BYTES, GS_FB_POLY_SITES, GS_PH_STATE_SITES, GS_PH_SEGMENTS, GS_FB_LAUNCHES, GS_N };

// Device control block of a replayed schedule.
struct GraphCtl {
    uint32_t iter;         // iteration the next replay runs (advanced by the last kernel of the graph)
    int32_t overflow;      // sticky GO_* bits: why a replay placed nothing
    int32_t fb_key[2];     // launch group A (largest capacity class) and B (the rest): shape key
    int32_t fb_n[2];       //   and item count
    int32_t ph_key[2];
    int32_t ph_n[2];
    int32_t pad[2];
    unsigned long long stats[GS_N];   // counters of the replayed iterations (GS_*)
};

// Capacities (elements) of the placed buffers a captured graph writes.
struct GraphCaps {
    int64_t hvar_words, alpha_floats, loci, seg, segsum, imp, impsum, dip;
};

// Launch keys: capacity class x 2 + latency shape.  The forward-backward class
// is fb_cap_words of the group's largest word count, the phasing class the
// state count rounded up to 256 / 512 / 1024 / 2048 / 4096.
__host__ __device__ inline int ph_cap_states(int K) {
    return K <= 256 ? 256 : K <= 512 ? 512 : K <= 1024 ? 1024 : K <= 2048 ? 2048 : 4096;
}
__host__ __device__ inline int fb_cap_class(int W) { return W <= 4 ? 4 : W <= 8 ? 8 : W <= 16 ? 16 : W <= 32 ? 32 : 64; }
__host__ __device__ inline int graph_key(int cap, bool latency) { return cap * 2 + (latency ? 1 : 0); }

void pbwt_draw_launch(const GraphCtl *ctl, uint32_t seed, const int32_t *grp_ptr, const int32_t *grp_end, int ncls,
                      int Lhq, uint8_t *stored, const int32_t *targets, int npt, int N, int32_t *start, hipStream_t s);
// C: the checkpoint block length of the run's layout; few: a group of at most
// this many items takes the latency shape; ford: [2][T] item orders of A and B.
// want: the captured launch keys of groups A and B (-1: no group B).
void fb_place_launch(GraphCtl *ctl, CondItem *items, FbItem *fb, PhItem *ph, int T, int L, int C, int few,
                     bool latency, const GraphCaps &caps, const int want[2], int32_t *ford, int32_t *nstates,
                     int n_iter, hipStream_t s);
void ph_place_launch(GraphCtl *ctl, PhItem *ph, const CondItem *items, int T, int few, bool latency,
                     const GraphCaps &caps, const int want[2], int32_t *pord, hipStream_t s);
void graph_advance_launch(GraphCtl *ctl, hipStream_t s);

}  // namespace gh
