// gh_init.hpp -- the INIT iteration's state selection on the device (gh_init.hip).
#pragma once
#include <stdint.h>

#include "gh_internal.hpp"

namespace gh {

constexpr int INIT_MAC_MAX = 2047;              // largest minor-allele count of a rare site (bucket table)
constexpr size_t INIT_LDS_MAX = 96 * 1024;      // the state set's bitmap: N <= 786,432

struct InitArgs {
    const int32_t *targets;    // optional: workgroup b selects for target targets[b] (else b)
    const uint8_t *gl;         // [T][L][3] phred GLs
    const uint8_t *flat;       // [T][L]
    const uint8_t *ploidy;     // [T] or NULL (all diploid)
    const uint8_t *flags;      // [L] SF_* bits
    const uint32_t *cref, *calt;
    const int32_t *rare_ptr, *rare_idx;   // carriers of the non-major allele per site
    const float *unph;         // [256] the unphred table
    int L, N, kinit, kstride, max_mac;
    uint32_t seed;
    int32_t *sind, *sorted;    // scratch [T][L] each
    int32_t *states;           // [T][kstride] out: sorted state ids
    int32_t *K;                // [T] out: set sizes
    CondItem *items;           // optional [T]: K and W written
};

// the unphred table (common/src/utils/otools.h:97-98), gh_gibbs.hip
void unphred_table(float *t);

size_t init_lds_bytes(int N);
void init_select_launch(const InitArgs &a, int n, hipStream_t s);
// gh_init_states: INIT sets of T diploid targets (host arrays in and out).
void init_states_batch(gh_panel *P, int T, const uint8_t *gl, const uint8_t *flat, int kinit, uint32_t seed,
                       int kstride, int32_t *states, int32_t *K);
void full_panel_launch(const int32_t *targets, int n, int N, int kstride, int32_t *states, int32_t *K,
                       CondItem *items, hipStream_t s);

}  // namespace gh
