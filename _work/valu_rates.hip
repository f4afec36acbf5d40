// Issue rates of the FP32 VALU instruction forms that the engine's pair and
// translation kernels are built from, on one MI355X (gfx950).
//
// Each kernel runs 16 independent chains of one instruction form per lane
// (8192 workgroups of 256 lanes, 4096 iterations) and reports lane operations
// per second.  The forms differ only in where their operands live: the
// engine's kernels keep wave-uniform operands (softening, translation
// coefficients) either in SGPRs or in VGPRs, and that choice sets the rate.
//
// Build and run (packed FP32 disabled, as for the engine):
//   hipcc --offload-arch=gfx950 -O3 -Xclang -target-feature -Xclang -packed-fp32-ops \
//       -o valu_rates valu_rates.hip && ./valu_rates
#include <hip/hip_runtime.h>
#include <cstdio>

#define CK(x)                                                                       \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::printf("%s failed: %s\n", #x, hipGetErrorString(e_));             \
            return 1;                                                               \
        }                                                                           \
    } while (0)

enum Form { kFmaSgpr, kFmacSgpr, kMulSgpr, kFmaVgpr, kFmacVgpr, kMulVgpr, kSubVgpr, kFmaSquare, kRsq, kNForms };

static const char* kFormName[kNForms] = {
    "v_fma_f32   a*s+v (one SGPR operand)", "v_fmac_f32  s*v+a (SGPR operand)",
    "v_mul_f32   a*s   (SGPR operand)",     "v_fma_f32   a*v+v (VGPR operands)",
    "v_fmac_f32  v*v+a (VGPR operands)",    "v_mul_f32   a*v   (VGPR operands)",
    "v_sub_f32   v-a   (VGPR operands)",    "v_fma_f32   a*a+v (r^2 form, VGPRs)",
    "v_rsq_f32   (with one v_add_f32)"};

template <int F>
__global__ __launch_bounds__(256) void k_form(float* out, int iters, float b, float c) {
    float a[16];
    // lane-dependent copies of b and c live in VGPRs; b and c themselves are
    // kernel arguments (SGPRs)
    const float bv = b + threadIdx.x * 1e-9f, cv = c + threadIdx.x * 1e-9f;
#pragma unroll
    for (int i = 0; i < 16; i++) a[i] = 1.0f + threadIdx.x * 1e-3f + i;
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int i = 0; i < 16; i++) {
            if constexpr (F == kFmaSgpr) a[i] = fmaf(a[i], b, cv);
            else if constexpr (F == kFmacSgpr) a[i] = fmaf(b, cv, a[i]);
            else if constexpr (F == kMulSgpr) a[i] = a[i] * b;
            else if constexpr (F == kFmaVgpr) a[i] = fmaf(a[i], bv, cv);
            else if constexpr (F == kFmacVgpr) a[i] = fmaf(bv, cv, a[i]);
            else if constexpr (F == kMulVgpr) a[i] = a[i] * bv;
            else if constexpr (F == kSubVgpr) a[i] = bv - a[i];
            else if constexpr (F == kFmaSquare) a[i] = fmaf(a[i], a[i], cv);
            else a[i] = __builtin_amdgcn_rsqf(a[i] + bv);
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int i = 0; i < 16; i++) s += a[i];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

template <int F>
static int run(float* out, int blocks, int threads, int iters) {
    hipEvent_t e0, e1;
    CK(hipEventCreate(&e0));
    CK(hipEventCreate(&e1));
    float ms = 0.0f;
    for (int rep = 0; rep < 2; rep++) {  // the first launch warms up
        CK(hipEventRecord(e0));
        hipLaunchKernelGGL(k_form<F>, dim3(blocks), dim3(threads), 0, 0, out, iters, 0.999f, 1e-3f);
        CK(hipGetLastError());
        CK(hipEventRecord(e1));
        CK(hipEventSynchronize(e1));
        CK(hipEventElapsedTime(&ms, e0, e1));
    }
    const double ops = (double)blocks * threads * iters * 16;
    std::printf("%-40s %8.3f ms  %.3e lane-ops/s\n", kFormName[F], ms, ops / (ms * 1e-3));
    CK(hipEventDestroy(e0));
    CK(hipEventDestroy(e1));
    return 0;
}

int main() {
    const int blocks = 8192, threads = 256, iters = 4096;
    float* out = nullptr;
    CK(hipMalloc(&out, sizeof(float) * blocks * threads));
    int rc = run<kFmaSgpr>(out, blocks, threads, iters) || run<kFmacSgpr>(out, blocks, threads, iters) ||
             run<kMulSgpr>(out, blocks, threads, iters) || run<kFmaVgpr>(out, blocks, threads, iters) ||
             run<kFmacVgpr>(out, blocks, threads, iters) || run<kMulVgpr>(out, blocks, threads, iters) ||
             run<kSubVgpr>(out, blocks, threads, iters) || run<kFmaSquare>(out, blocks, threads, iters) ||
             run<kRsq>(out, blocks, threads, iters / 4);
    CK(hipFree(out));
    return rc;
}
