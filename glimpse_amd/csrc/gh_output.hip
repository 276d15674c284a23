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
#include <vector>

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

// The haploid form: two posteriors, keys of the two remainders.
__device__ __forceinline__ void gp_fixup2(float gp0, float gp1, float q[2]) {
#pragma clang fp contract(off)
    q[0] = (float)((double)floorf((float)((double)gp0 * 1000.0)) / 1000.0);
    q[1] = (float)((double)floorf((float)((double)gp1 * 1000.0)) / 1000.0);
    const float key[2] = {1.0f - (gp0 - q[0]), 1.0f - (gp1 - q[1])};
    int ord[2] = {0, 1}, n = 2;
    if (key[1] == key[0]) n = 1;
    else if (key[1] < key[0]) { ord[0] = 1; ord[1] = 0; }
    for (int a = 0; a < n && (q[0] + q[1]) < 0.9999f; a++) q[ord[a]] += 0.001f;
}

// ploidy (NULL: every target diploid): a haploid target writes GP (gp0, gp1),
// DS = gp1, one allele (gp0 < gp1) and a vector end (gt = 255, GP[2] = NaN);
// INFO is computed only when every target is diploid (fploidy 2), else 1
// (genotype_writer.cpp:138-190).
__global__ void output_kernel(int T, int L, const float *__restrict__ gp0_all, const float *__restrict__ gp1_all,
                              const uint8_t *__restrict__ hds_all, const uint32_t *__restrict__ cref,
                              const uint32_t *__restrict__ calt, const uint8_t *__restrict__ ploidy, int n_haps,
                              int fploidy, float *__restrict__ ds_out,
                              float *__restrict__ gp_out, uint8_t *__restrict__ gt_out, float *__restrict__ info_out,
                              float *__restrict__ af_out, float *__restrict__ raf_out) {
#pragma clang fp contract(off)
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    float ds_sum = 0.0f, ds2_sum = 0.0f, ds4_sum = 0.0f;
    for (int i = 0; i < T; i++) {
        const size_t o = (size_t)i * L + l;
        const float gp0 = gp0_all[o], gp1 = gp1_all[o];
        if (ploidy && ploidy[i] == 1) {
            const float ds = gp1;
            ds_out[o] = (float)((double)roundf((float)((double)ds * 1000.0)) / 1000.0);
            float q[2];
            gp_fixup2(gp0, gp1, q);
            gp_out[3 * o] = q[0];
            gp_out[3 * o + 1] = q[1];
            gp_out[3 * o + 2] = __int_as_float(0x7fc00000);
            gt_out[2 * o] = gp0 < gp1;
            gt_out[2 * o + 1] = 255;
            ds_sum += ds;
            continue;
        }
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
        if (fploidy == 2) {
            ds2_sum += ds * ds;
            ds4_sum = (float)((double)ds4_sum + ((double)gp1 + 4.0 * (double)gp2));
        }
    }
    const float raf = (float)calt[l] * 1.0f / (float)(calt[l] + cref[l]);
    const float af = ds_sum / (float)n_haps;
    float info = 1.0f;
    if (fploidy == 2 && af > 0.0f && af < 1.0f)
        info = (float)(1.0 - (double)(ds4_sum - ds2_sum) / ((double)((float)n_haps * af) * (1.0 - (double)af)));
    info = info < 0.0f ? 0.0f : info;
    info = (float)((double)roundf((float)((double)info * 1000.0)) / 1000.0);
    info_out[l] = info;
    af_out[l] = af;
    raf_out[l] = raf;
}

void output_fields(int T, int L, const float *gp0, const float *gp1, const uint8_t *hds, const uint32_t *cref,
                   const uint32_t *calt, float *ds, float *gp, uint8_t *gt, float *info, float *af, float *raf,
                   const int32_t *ploidy_h) {
    GH_CHECK(T > 0 && L > 0 && gp0 && gp1 && hds && cref && calt && ds && gp && gt && info && af && raf,
             GH_ERR_ARG, "gh_output_fields: bad arguments");
    // H.n_tar_haps and the format ploidy (genotype_reader.cpp:553-560): 2 all
    // diploid, 1 all haploid, -2 mixed
    int n_haps = 2 * T, n1 = 0;
    std::vector<uint8_t> pl;
    if (ploidy_h) {
        pl.resize(T);
        n_haps = 0;
        for (int i = 0; i < T; i++) {
            GH_CHECK(ploidy_h[i] == 1 || ploidy_h[i] == 2, GH_ERR_ARG, "gh_output_fields: ploidy must be 1 or 2");
            pl[i] = (uint8_t)ploidy_h[i];
            n_haps += ploidy_h[i];
            n1 += ploidy_h[i] == 1;
        }
    }
    const int fploidy = n1 == 0 ? 2 : n1 == T ? 1 : -2;
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
    DBuf dpl;
    const uint8_t *dp = pl.empty() ? nullptr : dalloc_copy(dpl, pl.data(), (size_t)T, s);
    dds.reserve(sizeof(float) * TL);
    dgp.reserve(sizeof(float) * 3 * TL);
    dgt.reserve(2 * TL);
    dinfo.reserve(sizeof(float) * L);
    daf.reserve(sizeof(float) * L);
    draf.reserve(sizeof(float) * L);
    hipLaunchKernelGGL(output_kernel, dim3((L + 255) / 256), dim3(256), 0, s, T, L, g0, g1, h, cr, ca, dp, n_haps,
                       fploidy, dds.as<float>(), dgp.as<float>(), dgt.as<uint8_t>(), dinfo.as<float>(), daf.as<float>(),
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
