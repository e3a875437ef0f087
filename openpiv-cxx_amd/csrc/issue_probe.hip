// issue_probe.hip -- VALU issue cost of single opcodes on this GPU.
//
// Each probe kernel keeps 16 independent accumulators per lane and applies
// one opcode to each of them in turn (written as inline asm so the compiler
// cannot change the encoding), 8 waves per SIMD, so that neither dependency
// latency nor occupancy limits the rate: the kernel time then gives the
// issue cost of the opcode.  The result is printed in cycles per wave64
// instruction per SIMD at the nominal 2.4 GHz clock (the unit the kernels'
// issue model uses; the clock under load may be lower, which scales every
// row alike).  The asm here is arithmetic on registers only: no memory
// access.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <utility>

#define R16(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7) X(8) X(9) X(10) X(11) X(12) X(13) X(14) X(15)
#define R8(X) X(0) X(1) X(2) X(3) X(4) X(5) X(6) X(7)

namespace {

enum Op {
    ADD_F32, SUB_F32, MUL_F32, FMAC_F32, FMAMK_F32, FMA_F32_VOP3, ADD_F32_VOP3, MOV_B32, ADD_U32, AND_B32,
    XOR_B32, ASHR_I32, MAX_F32, MIN_F32, MAX3_F32, MED3_F32, MAX_I32, MAX_U32, BFI_B32, MUL_LO_U32, CVT_F32_I32,
    MOV_DPP, MAX_DPP, CMP_CNDMASK, ADD_F64, FMA_F64, N_OPS
};

const char* kNames[N_OPS] = {
    "v_add_f32_e32", "v_sub_f32_e32", "v_mul_f32_e32", "v_fmac_f32_e32", "v_fmamk_f32", "v_fma_f32 (VOP3)",
    "v_add_f32_e64 (VOP3)", "v_mov_b32_e32", "v_add_u32_e32", "v_and_b32_e32", "v_xor_b32_e32", "v_ashrrev_i32_e32",
    "v_max_f32_e32", "v_min_f32_e32", "v_max3_f32", "v_med3_f32", "v_max_i32_e32", "v_max_u32_e32", "v_bfi_b32",
    "v_mul_lo_u32", "v_cvt_f32_i32_e32", "v_mov_b32_dpp", "v_max_f32_dpp", "v_cmp_gt_f32 + v_cndmask_b32",
    "v_add_f64", "v_fma_f64"};

template <int OP>
__global__ __launch_bounds__(256) void probe_kernel(float* out, int iters, float s)
{
    float a[16];
    double d[8];
#define INIT(i) a[i] = (float)threadIdx.x * 1e-3f + (float)i;
    R16(INIT)
#define INITD(i) d[i] = (double)a[i];
    R8(INITD)
    const float b = s, c = s * 0.5f;
    const double bd = b, cd = c;
    const int bi = (int)threadIdx.x, ci = 7;
    for (int it = 0; it < iters; ++it) {
#define OPS(i)                                                                                                    \
        if constexpr (OP == ADD_F32) asm volatile("v_add_f32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(b));            \
        else if constexpr (OP == SUB_F32) asm volatile("v_sub_f32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(b));       \
        else if constexpr (OP == MUL_F32) asm volatile("v_mul_f32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(b));       \
        else if constexpr (OP == FMAC_F32) asm volatile("v_fmac_f32_e32 %0, %1, %2" : "+v"(a[i]) : "v"(b), "v"(c)); \
        else if constexpr (OP == FMAMK_F32) asm volatile("v_fmamk_f32 %0, %0, 0x3f3504f3, %1" : "+v"(a[i]) : "v"(b)); \
        else if constexpr (OP == FMA_F32_VOP3) asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a[i]) : "v"(b), "v"(c)); \
        else if constexpr (OP == ADD_F32_VOP3) asm volatile("v_add_f32_e64 %0, %1, %0" : "+v"(a[i]) : "v"(b));  \
        else if constexpr (OP == MOV_B32) asm volatile("v_mov_b32_e32 %0, %1" : "=v"(a[i]) : "v"(a[(i + 15) & 15])); \
        else if constexpr (OP == ADD_U32) asm volatile("v_add_u32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(bi));      \
        else if constexpr (OP == AND_B32) asm volatile("v_and_b32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(bi));      \
        else if constexpr (OP == XOR_B32) asm volatile("v_xor_b32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(bi));      \
        else if constexpr (OP == ASHR_I32) asm volatile("v_ashrrev_i32_e32 %0, 31, %0" : "+v"(a[i]));          \
        else if constexpr (OP == MAX_F32) asm volatile("v_max_f32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(b));       \
        else if constexpr (OP == MIN_F32) asm volatile("v_min_f32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(b));       \
        else if constexpr (OP == MAX3_F32) asm volatile("v_max3_f32 %0, %0, %1, %2" : "+v"(a[i]) : "v"(b), "v"(c)); \
        else if constexpr (OP == MED3_F32) asm volatile("v_med3_f32 %0, %0, %1, %2" : "+v"(a[i]) : "v"(b), "v"(c)); \
        else if constexpr (OP == MAX_I32) asm volatile("v_max_i32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(bi));      \
        else if constexpr (OP == MAX_U32) asm volatile("v_max_u32_e32 %0, %1, %0" : "+v"(a[i]) : "v"(bi));      \
        else if constexpr (OP == BFI_B32) asm volatile("v_bfi_b32 %0, %1, %0, %2" : "+v"(a[i]) : "v"(bi), "v"(ci)); \
        else if constexpr (OP == MUL_LO_U32) asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(a[i]) : "v"(bi));    \
        else if constexpr (OP == CVT_F32_I32) asm volatile("v_cvt_f32_i32_e32 %0, %0" : "+v"(a[i]));           \
        else if constexpr (OP == MOV_DPP)                                                                        \
            asm volatile("v_mov_b32_dpp %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(a[i]));   \
        else if constexpr (OP == MAX_DPP)                                                                        \
            asm volatile("v_max_f32_dpp %0, %0, %0 quad_perm:[1,0,3,2] row_mask:0xf bank_mask:0xf" : "+v"(a[i]));
        R16(OPS)
#define PAIRS(i)                                                                                                 \
        if constexpr (OP == CMP_CNDMASK) {                                                                       \
            asm volatile("v_cmp_gt_f32_e32 vcc, %1, %0\n\tv_cndmask_b32_e32 %0, %1, %0, vcc" : "+v"(a[i]) : "v"(b) : "vcc"); \
            asm volatile("v_cmp_gt_f32_e32 vcc, %1, %0\n\tv_cndmask_b32_e32 %0, %1, %0, vcc" : "+v"(a[i + 8]) : "v"(c) : "vcc"); \
        } else if constexpr (OP == ADD_F64) {                                                                    \
            asm volatile("v_add_f64 %0, %0, %1" : "+v"(d[i]) : "v"(bd));                                         \
            asm volatile("v_add_f64 %0, %0, %1" : "+v"(d[i]) : "v"(cd));                                         \
        } else if constexpr (OP == FMA_F64) {                                                                    \
            asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(d[i]) : "v"(bd), "v"(cd));                            \
            asm volatile("v_fma_f64 %0, %0, %2, %1" : "+v"(d[i]) : "v"(bd), "v"(cd));                            \
        }
        R8(PAIRS)
    }
    float r = 0.f;
#define SUM(i) r += a[i];
    R16(SUM)
#define SUMD(i) r += (float)d[i];
    R8(SUMD)
    out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

// VALU instructions per loop iteration of probe OP
constexpr int per_iter(int op) { return op == CMP_CNDMASK ? 32 : 16; }

void check(hipError_t e, const char* what)
{
    if (e != hipSuccess) {
        std::fprintf(stderr, "%s: %s\n", what, hipGetErrorString(e));
        std::exit(1);
    }
}

template <int OP>
void run(float* out, int blocks, int iters)
{
    hipLaunchKernelGGL(probe_kernel<OP>, dim3(blocks), dim3(256), 0, 0, out, iters, 1.0001f);
    check(hipDeviceSynchronize(), "warm-up");
    hipEvent_t e0, e1;
    check(hipEventCreate(&e0), "event");
    check(hipEventCreate(&e1), "event");
    float best = 1e30f;
    for (int rep = 0; rep < 3; ++rep) {
        check(hipEventRecord(e0, 0), "record");
        hipLaunchKernelGGL(probe_kernel<OP>, dim3(blocks), dim3(256), 0, 0, out, iters, 1.0001f);
        check(hipEventRecord(e1, 0), "record");
        check(hipEventSynchronize(e1), "sync");
        float ms = 0.f;
        check(hipEventElapsedTime(&ms, e0, e1), "elapsed");
        best = std::min(best, ms);
    }
    check(hipEventDestroy(e0), "event");
    check(hipEventDestroy(e1), "event");
    int dev = 0, cus = 0;
    check(hipGetDevice(&dev), "device");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
    const double waves = blocks * 4.0, simds = cus * 4.0;
    const double instr_per_simd = (double)iters * per_iter(OP) * waves / simds;
    const double cycles = best * 1e-3 * 2.4e9 / instr_per_simd;
    std::printf("%-30s %8.3f ms  %5.2f cycles per wave64 instruction\n", kNames[OP], best, cycles);
}

template <int... OPS>
void run_all(float* out, int blocks, int iters, std::integer_sequence<int, OPS...>)
{
    (run<OPS>(out, blocks, iters), ...);
}

} // namespace

int main()
{
    int dev = 0, cus = 0;
    check(hipGetDevice(&dev), "device");
    check(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev), "attribute");
    const int blocks = cus * 8;   // 256 threads each: 8 waves per SIMD
    const int iters = 8192;
    float* out = nullptr;
    check(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)), "hipMalloc");
    std::printf("%d CUs, %d blocks of 256 threads (8 waves per SIMD), %d iterations\n", cus, blocks, iters);
    run_all(out, blocks, iters, std::make_integer_sequence<int, N_OPS>{});
    check(hipFree(out), "hipFree");
    return 0;
}
