// gh_output.hip -- per-site output fields of the phased, imputed targets:
// DS, GP (floored to 1e-3, then the largest remainders raised by 1e-3 until
// they sum to >= 0.9999), phased GT, and the site's INFO score, AF and RAF
// (genotype_writer::writeGenotypes, phase/src/io/genotype_writer.cpp:109-190).
//
// A target without a stored genotype at a site (GP taken as (1, 0, 0)) and one
// whose stored GP is (1, 0) give the same fields, so the finalized dense GP
// arrays are all the writer needs.  One thread per site walks the targets in
// order, so the INFO sums accumulate in the writer's order.  Float and double
// operations follow the writer's expression types (its literals 1000.0 and 4.0
// are doubles) and nothing is contracted into an FMA, so every field is
// bit-identical to what it would write.
#include "gh_device.hpp"
#include "gh_internal.hpp"

namespace gh {

// The writer keys a std::map<float, float*> by 1 - (gp - floored) and walks it
// in ascending key order; a key that is already present is not inserted again.
__device__ __forceinline__ void gp_fixup(float gp0, float gp1, float gp2, float q[3]) {
#pragma clang fp contract(off)
    q[0] = (float)((double)floorf((float)((double)gp0 * 1000.0)) / 1000.0);
    q[1] = (float)((double)floorf((float)((double)gp1 * 1000.0)) / 1000.0);
    q[2] = (float)((double)floorf((float)((double)fmaxf(1.0f - (q[0] + q[1]), 0.0f) * 1000.0)) / 1000.0);
    const float key[3] = {1.0f - (gp0 - q[0]), 1.0f - (gp1 - q[1]), 1.0f - (gp2 - q[2])};
    int ord[3], n = 0;
    for (int a = 0; a < 3; a++) {
        bool dup = false;
        for (int b = 0; b < a; b++) dup = dup || key[b] == key[a];
        if (!dup) ord[n++] = a;
    }
    // ascending key (n <= 3: insertion sort)
    for (int a = 1; a < n; a++)
        for (int b = a; b > 0 && key[ord[b]] < key[ord[b - 1]]; b--) {
            const int tmp = ord[b];
            ord[b] = ord[b - 1];
            ord[b - 1] = tmp;
        }
    for (int a = 0; a < n && ((q[0] + q[1]) + q[2]) < 0.9999f; a++) q[ord[a]] += 0.001f;
}

__global__ void output_kernel(int T, int L, const float *__restrict__ gp0_all, const float *__restrict__ gp1_all,
                              const uint8_t *__restrict__ hds_all, const uint32_t *__restrict__ cref,
                              const uint32_t *__restrict__ calt, float *__restrict__ ds_out,
                              float *__restrict__ gp_out, uint8_t *__restrict__ gt_out, float *__restrict__ info_out,
                              float *__restrict__ af_out, float *__restrict__ raf_out) {
#pragma clang fp contract(off)
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    float ds_sum = 0.0f, ds2_sum = 0.0f, ds4_sum = 0.0f;
    for (int i = 0; i < T; i++) {
        const size_t o = (size_t)i * L + l;
        const float gp0 = gp0_all[o], gp1 = gp1_all[o];
        const float gp2 = fminf(fmaxf(1.0f - gp1 - gp0, 0.0f), 1.0f);
        const float ds = gp1 + 2.0f * gp2;
        uint8_t a0, a1;
        if (gp1 > gp0 && gp1 > gp2) {   // het: the alt allele on haplotype hds
            const int h = hds_all[o] ? 1 : 0;
            a0 = h == 0;
            a1 = h == 1;
        } else {
            a0 = a1 = gp0 < gp2;
        }
        ds_out[o] = (float)((double)roundf((float)((double)ds * 1000.0)) / 1000.0);
        float q[3];
        gp_fixup(gp0, gp1, gp2, q);
        gp_out[3 * o] = q[0];
        gp_out[3 * o + 1] = q[1];
        gp_out[3 * o + 2] = q[2];
        gt_out[2 * o] = a0;
        gt_out[2 * o + 1] = a1;
        ds_sum += ds;
        ds2_sum += ds * ds;
        ds4_sum = (float)((double)ds4_sum + ((double)gp1 + 4.0 * (double)gp2));
    }
    const int n_haps = 2 * T;
    const float raf = (float)calt[l] * 1.0f / (float)(calt[l] + cref[l]);
    const float af = ds_sum / (float)n_haps;
    float info = 1.0f;
    if (af > 0.0f && af < 1.0f)
        info = (float)(1.0 - (double)(ds4_sum - ds2_sum) / ((double)((float)n_haps * af) * (1.0 - (double)af)));
    info = info < 0.0f ? 0.0f : info;
    info = (float)((double)roundf((float)((double)info * 1000.0)) / 1000.0);
    info_out[l] = info;
    af_out[l] = af;
    raf_out[l] = raf;
}

void output_fields(int T, int L, const float *gp0, const float *gp1, const uint8_t *hds, const uint32_t *cref,
                   const uint32_t *calt, float *ds, float *gp, uint8_t *gt, float *info, float *af, float *raf) {
    GH_CHECK(T > 0 && L > 0 && gp0 && gp1 && hds && cref && calt && ds && gp && gt && info && af && raf,
             GH_ERR_ARG, "gh_output_fields: bad arguments");
    hipStream_t s = nullptr;
    GH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    const size_t TL = (size_t)T * L;
    DBuf d0, d1, dh, dcr, dca, dds, dgp, dgt, dinfo, daf, draf;
    const float *g0 = dalloc_copy(d0, gp0, TL, s), *g1 = dalloc_copy(d1, gp1, TL, s);
    const uint8_t *h = dalloc_copy(dh, hds, TL, s);
    const uint32_t *cr = dalloc_copy(dcr, cref, (size_t)L, s), *ca = dalloc_copy(dca, calt, (size_t)L, s);
    dds.reserve(sizeof(float) * TL);
    dgp.reserve(sizeof(float) * 3 * TL);
    dgt.reserve(2 * TL);
    dinfo.reserve(sizeof(float) * L);
    daf.reserve(sizeof(float) * L);
    draf.reserve(sizeof(float) * L);
    hipLaunchKernelGGL(output_kernel, dim3((L + 255) / 256), dim3(256), 0, s, T, L, g0, g1, h, cr, ca,
                       dds.as<float>(), dgp.as<float>(), dgt.as<uint8_t>(), dinfo.as<float>(), daf.as<float>(),
                       draf.as<float>());
    GH_HIP(hipGetLastError());
    GH_HIP(hipMemcpyAsync(ds, dds.p, sizeof(float) * TL, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(gp, dgp.p, sizeof(float) * 3 * TL, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(gt, dgt.p, 2 * TL, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(info, dinfo.p, sizeof(float) * L, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(af, daf.p, sizeof(float) * L, hipMemcpyDeviceToHost, s));
    GH_HIP(hipMemcpyAsync(raf, draf.p, sizeof(float) * L, hipMemcpyDeviceToHost, s));
    GH_HIP(hipStreamSynchronize(s));
}

}  // namespace gh
