// pt_env.hip -- environment importance-sampling tables built on the GPU in FP64.
//
// Weight of texel (x, y) = r + g + b (environment_light.cu:44-47), optionally times
// sin(theta).  One thread per row builds that row's alias table with the sweep
// construction (heavy items absorb the deficits of light items in index order; a heavy item
// whose residual drops below one becomes a light bucket aliased to the next heavy item);
// one thread builds the marginal table over the row sums.  The operation order is fixed
// and fused multiply-adds are disabled here, so the tables are reproducible bit for bit.
#include <hip/hip_runtime.h>

#include "pt_gpu.h"

#pragma clang fp contract(off)

namespace pt {

namespace {

__global__ void k_env_weights(const float4 *rgba, int w, int h, int sinWeight, double *wts) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= w * h) return;
    const int y = i / w;
    const float4 v = rgba[i];
    const double st = sinWeight ? sin((y + 0.5) / h * 3.14159265358979323846) : 1.0;
    wts[i] = ((double)v.x + (double)v.y + (double)v.z) * st;
}

__device__ void aliasSweep(const double *w, int n, double *prob, int *alias, double *sumOut) {
    double sum = 0.0;
    for (int i = 0; i < n; i++) sum += w[i];
    *sumOut = sum;
    for (int i = 0; i < n; i++) {
        prob[i] = 1.0;
        alias[i] = i;
    }
    if (!(sum > 0.0)) return;
    const double scale = (double)n / sum;
    int i = 0;
    while (i < n && !(w[i] * scale < 1.0)) i++;
    int j = 0;
    while (j < n && w[j] * scale < 1.0) j++;
    if (j >= n) return;
    double wj = w[j] * scale;
    // every step advances i or j, so the loop ends after at most 2n steps
    while (true) {
        if (wj >= 1.0) {
            if (i >= n) break;
            const double qi = w[i] * scale;
            prob[i] = qi;
            alias[i] = j;
            wj = wj + (qi - 1.0);
            i++;
            while (i < n && !(w[i] * scale < 1.0)) i++;
        } else {
            int jn = j + 1;
            while (jn < n && w[jn] * scale < 1.0) jn++;
            if (jn >= n) break;
            prob[j] = wj;
            alias[j] = jn;
            wj = w[jn] * scale + (wj - 1.0);
            j = jn;
        }
    }
}

__global__ void k_env_rows(const double *wts, int w, int h, double *prob, int *alias, double *rowSum) {
    const int y = blockIdx.x * blockDim.x + threadIdx.x;
    if (y >= h) return;
    aliasSweep(wts + (size_t)y * w, w, prob + (size_t)y * w, alias + (size_t)y * w, &rowSum[y]);
}

__global__ void k_env_marginal(const double *rowSum, int h, double *prob, int *alias, double *total) {
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    aliasSweep(rowSum, h, prob, alias, total);
}

__global__ void k_env_pack(const double *wts, const double *rowProb, const int *rowAlias, const double *margProb, const int *margAlias,
                           const double *total, int w, int h, float2 *rows, float2 *marginal, float *texelPdf) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < w * h) {
        rows[i] = make_float2((float)rowProb[i], asFloat(rowAlias[i]));
        const double t = *total;
        const double norm = t > 0.0 ? (double)w * h / (2.0 * 3.14159265358979323846 * 3.14159265358979323846) / t : 0.0;
        texelPdf[i] = (float)(wts[i] * norm);
    }
    if (i < h) marginal[i] = make_float2((float)margProb[i], asFloat(margAlias[i]));
}

}  // namespace

__global__ void k_env_pack_pdf(float4 *rgba, const float *texelPdf, int n) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) rgba[i].w = texelPdf[i];
}

void packEnvPdf(float4 *rgba, const float *texelPdf, int n, hipStream_t stream) {
    if (n <= 0) return;
    k_env_pack_pdf<<<(n + 255) / 256, 256, 0, stream>>>(rgba, texelPdf, n);
    PT_HIP_TRY(hipGetLastError());
}

void buildEnvTables(const float4 *rgba, int w, int h, int sinWeight, float2 *rows, float2 *marginal, float *texelPdf, double *rowProbOut,
                    int *rowAliasOut, double *margProbOut, int *margAliasOut, hipStream_t stream) {
    const size_t n = (size_t)w * h;
    DevBuf wts, rp, ra, rs, mp, ma, tot;
    wts.alloc(n * sizeof(double));
    rp.alloc(n * sizeof(double));
    ra.alloc(n * sizeof(int));
    rs.alloc(h * sizeof(double));
    mp.alloc(h * sizeof(double));
    ma.alloc(h * sizeof(int));
    tot.alloc(sizeof(double));
    const int nb = (int)((n + 255) / 256);
    k_env_weights<<<nb, 256, 0, stream>>>(rgba, w, h, sinWeight, wts.as<double>());
    k_env_rows<<<(h + 63) / 64, 64, 0, stream>>>(wts.as<double>(), w, h, rp.as<double>(), ra.as<int>(), rs.as<double>());
    k_env_marginal<<<1, 64, 0, stream>>>(rs.as<double>(), h, mp.as<double>(), ma.as<int>(), tot.as<double>());
    const int nb2 = (int)((std::max(n, (size_t)h) + 255) / 256);
    k_env_pack<<<nb2, 256, 0, stream>>>(wts.as<double>(), rp.as<double>(), ra.as<int>(), mp.as<double>(), ma.as<int>(), tot.as<double>(), w, h,
                                        rows, marginal, texelPdf);
    PT_HIP_TRY(hipGetLastError());
    if (rowProbOut) PT_HIP_TRY(hipMemcpyAsync(rowProbOut, rp.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (rowAliasOut) PT_HIP_TRY(hipMemcpyAsync(rowAliasOut, ra.p, n * sizeof(int), hipMemcpyDeviceToHost, stream));
    if (margProbOut) PT_HIP_TRY(hipMemcpyAsync(margProbOut, mp.p, h * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (margAliasOut) PT_HIP_TRY(hipMemcpyAsync(margAliasOut, ma.p, h * sizeof(int), hipMemcpyDeviceToHost, stream));
    // temporaries are freed on return: wait for the kernels that read them
    PT_HIP_TRY(hipStreamSynchronize(stream));
}

}  // namespace pt
