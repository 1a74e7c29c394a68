// pt_env.hip -- environment importance-sampling tables built on the GPU in FP64.
//
// Weight of texel (x, y) = r + g + b (environment_light.cu:44-47), optionally times
// sin(theta).  One thread per row builds that row's alias table with the sweep
// construction (heavy items absorb the deficits of light items in index order; a heavy item
// whose residual drops below one becomes a light bucket aliased to the next heavy item);
// one thread builds the marginal table over the row sums.  The operation order is fixed
// and fused multiply-adds are disabled here, so the tables are reproducible bit for bit.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "../../include/pathed.h"
#include "pt_gpu.h"
#include "pt_internal.h"

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

namespace {

// The FP64 working arrays of one table build.
struct EnvScratch {
    double *wts, *rowProb, *rowSum, *margProb, *total;
    int *rowAlias, *margAlias;
};

void launchEnvTables(const float4 *rgba, int w, int h, int sinWeight, const EnvScratch &s, float2 *rows, float2 *marginal,
                     float *texelPdf, hipStream_t stream) {
    const size_t n = (size_t)w * h;
    const int nb = (int)((n + 255) / 256);
    k_env_weights<<<nb, 256, 0, stream>>>(rgba, w, h, sinWeight, s.wts);
    k_env_rows<<<(h + 63) / 64, 64, 0, stream>>>(s.wts, w, h, s.rowProb, s.rowAlias, s.rowSum);
    k_env_marginal<<<1, 64, 0, stream>>>(s.rowSum, h, s.margProb, s.margAlias, s.total);
    const int nb2 = (int)((std::max(n, (size_t)h) + 255) / 256);
    k_env_pack<<<nb2, 256, 0, stream>>>(s.wts, s.rowProb, s.rowAlias, s.margProb, s.margAlias, s.total, w, h, rows, marginal, texelPdf);
    PT_HIP_TRY(hipGetLastError());
}

}  // namespace

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
    const EnvScratch sc{wts.as<double>(), rp.as<double>(), rs.as<double>(), mp.as<double>(), tot.as<double>(), ra.as<int>(), ma.as<int>()};
    launchEnvTables(rgba, w, h, sinWeight, sc, rows, marginal, texelPdf, stream);
    if (rowProbOut) PT_HIP_TRY(hipMemcpyAsync(rowProbOut, rp.p, n * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (rowAliasOut) PT_HIP_TRY(hipMemcpyAsync(rowAliasOut, ra.p, n * sizeof(int), hipMemcpyDeviceToHost, stream));
    if (margProbOut) PT_HIP_TRY(hipMemcpyAsync(margProbOut, mp.p, h * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (margAliasOut) PT_HIP_TRY(hipMemcpyAsync(margAliasOut, ma.p, h * sizeof(int), hipMemcpyDeviceToHost, stream));
    // temporaries are freed on return: wait for the kernels that read them
    PT_HIP_TRY(hipStreamSynchronize(stream));
}

// ---------------------------------------------------------------------------------------------
// Table build without allocation: the working arrays live in one fixed device array per GPU.
// ---------------------------------------------------------------------------------------------
namespace {

constexpr size_t kAlign = 256;
constexpr size_t alignUp(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }
constexpr size_t scratchBytes(size_t texels, size_t rows) {
    return alignUp(texels * sizeof(double)) * 2 + alignUp(texels * sizeof(int)) + alignUp(rows * sizeof(double)) * 2 +
           alignUp(rows * sizeof(int)) + alignUp(sizeof(double));
}
constexpr size_t kScratchBytes = scratchBytes(PT_ENV_NOALLOC_MAX_TEXELS, PT_ENV_NOALLOC_MAX_ROWS);

// Device-side working memory of pt_env_tables_noalloc.  Host code never uses this name as an
// address (on the host it is a stand-in symbol): it uses the per-device pointer that
// hipGetSymbolAddress returns.
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex gScratchMutex;
void *gScratchPtr[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up once per device.
void *scratchOnDevice(int dev) {
    std::lock_guard<std::mutex> lock(gScratchMutex);
    if (!gScratchPtr[dev]) {
        void *ptr = nullptr;
        PT_HIP_TRY(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) throw HipFailure("hipGetSymbolAddress(scratch) returned NULL");
        gScratchPtr[dev] = ptr;
    }
    return gScratchPtr[dev];
}

}  // namespace

}  // namespace pt

extern "C" int pt_env_tables_noalloc(const float *rgba, int width, int height, int sin_theta_weight, float *rows, float *marginal,
                                     float *texel_pdf, void *stream) {
    using namespace pt;
    if (!rgba || !rows || !marginal || !texel_pdf || width <= 0 || height <= 0)
        return fail(PT_ERR_INVALID, "pt_env_tables_noalloc: bad arguments");
    if ((int64_t)width * height > PT_ENV_NOALLOC_MAX_TEXELS || height > PT_ENV_NOALLOC_MAX_ROWS)
        return fail(PT_ERR_INVALID, "pt_env_tables_noalloc: map of " + std::to_string(width) + "x" + std::to_string(height) +
                                        " exceeds the fixed scratch (at most " + std::to_string(PT_ENV_NOALLOC_MAX_TEXELS) +
                                        " texels and " + std::to_string(PT_ENV_NOALLOC_MAX_ROWS) + " rows)");
    try {
        int dev = 0;
        PT_HIP_TRY(hipGetDevice(&dev));
        if (dev < 0 || dev >= kMaxDevices) return fail(PT_ERR_DEVICE, "pt_env_tables_noalloc: device ordinal out of range");
        unsigned char *base = (unsigned char *)scratchOnDevice(dev);
        const size_t n = (size_t)width * height, hh = (size_t)height;
        EnvScratch sc;
        size_t at = 0;
        auto carve = [&](size_t bytes) {
            unsigned char *p = base + at;
            at += alignUp(bytes);
            return p;
        };
        sc.wts = (double *)carve(n * sizeof(double));
        sc.rowProb = (double *)carve(n * sizeof(double));
        sc.rowAlias = (int *)carve(n * sizeof(int));
        sc.rowSum = (double *)carve(hh * sizeof(double));
        sc.margProb = (double *)carve(hh * sizeof(double));
        sc.margAlias = (int *)carve(hh * sizeof(int));
        sc.total = (double *)carve(sizeof(double));
        if (at > kScratchBytes) return fail(PT_ERR_INVALID, "pt_env_tables_noalloc: scratch layout exceeds its array");
        // every working array is written in full before it is read, so nothing is cleared
        launchEnvTables((const float4 *)rgba, width, height, sin_theta_weight, sc, (float2 *)rows, (float2 *)marginal, texel_pdf,
                        (hipStream_t)stream);
        return PT_OK;
    } catch (const HipFailure &e) {
        return fail(PT_ERR_DEVICE, std::string("pt_env_tables_noalloc: ") + e.msg);
    }
}
