// Instruction issue ceilings of the device: every wave runs eight
// independent chains of f32 adds, of u32 adds (vector), then of scalar u32
// adds, written as inline asm, 8 waves per SIMD, so that nothing but issue
// bounds the kernel. Prints the wave instructions per SIMD per nanosecond;
// rocprofv3's SQ_INSTS_VALU / SQ_INSTS_SALU and GRBM_GUI_ACTIVE over the same
// dispatches give them per cycle. The trace's counters summary divides by
// these rates.
#include <hip/hip_runtime.h>

#include <cstdio>

#define CHECK(x)                                                                  \
    do {                                                                          \
        hipError_t e_ = (x);                                                      \
        if (e_ != hipSuccess) {                                                   \
            std::fprintf(stderr, "%s: %s\n", #x, hipGetErrorString(e_));          \
            return 1;                                                             \
        }                                                                         \
    } while (0)

template <int KIND>
__global__ void __launch_bounds__(256) k_valu_peak(float* out, int iters) {
    float a[8];
    unsigned u[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        a[k] = (float)(threadIdx.x + k);
        u[k] = threadIdx.x + k;
    }
    const float c = 1.0001f;
    const unsigned d = blockIdx.x;
    unsigned sc[8];
#pragma unroll
    for (int k = 0; k < 8; ++k) sc[k] = __builtin_amdgcn_readfirstlane(d + k);
    for (int i = 0; i < iters; ++i) {
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            if (KIND == 2)
                asm volatile("s_add_u32 %0, %0, %8\n s_add_u32 %1, %1, %8\n s_add_u32 %2, %2, %8\n s_add_u32 %3, %3, %8\n"
                             "s_add_u32 %4, %4, %8\n s_add_u32 %5, %5, %8\n s_add_u32 %6, %6, %8\n s_add_u32 %7, %7, %8"
                             : "+s"(sc[0]), "+s"(sc[1]), "+s"(sc[2]), "+s"(sc[3]), "+s"(sc[4]), "+s"(sc[5]), "+s"(sc[6]),
                               "+s"(sc[7])
                             : "s"(d)
                             : "scc");
            else if (KIND == 1)
                asm volatile("v_add_u32 %0, %0, %8\n v_add_u32 %1, %1, %8\n v_add_u32 %2, %2, %8\n v_add_u32 %3, %3, %8\n"
                             "v_add_u32 %4, %4, %8\n v_add_u32 %5, %5, %8\n v_add_u32 %6, %6, %8\n v_add_u32 %7, %7, %8"
                             : "+v"(u[0]), "+v"(u[1]), "+v"(u[2]), "+v"(u[3]), "+v"(u[4]), "+v"(u[5]), "+v"(u[6]),
                               "+v"(u[7])
                             : "v"(d));
            else
                asm volatile("v_add_f32 %0, %0, %8\n v_add_f32 %1, %1, %8\n v_add_f32 %2, %2, %8\n v_add_f32 %3, %3, %8\n"
                             "v_add_f32 %4, %4, %8\n v_add_f32 %5, %5, %8\n v_add_f32 %6, %6, %8\n v_add_f32 %7, %7, %8"
                             : "+v"(a[0]), "+v"(a[1]), "+v"(a[2]), "+v"(a[3]), "+v"(a[4]), "+v"(a[5]), "+v"(a[6]),
                               "+v"(a[7])
                             : "v"(c));
        }
    }
    float s = 0.0f;
#pragma unroll
    for (int k = 0; k < 8; ++k) s += a[k] + (float)u[k] + (float)sc[k];
    out[blockIdx.x * blockDim.x + threadIdx.x] = s;
}

int main() {
    int cus = 0;
    CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, 0));
    const int blocks = cus * 8;   // 8 blocks of 4 waves per CU: 8 waves per SIMD
    const int iters = 2000;
    float* out = nullptr;
    CHECK(hipMalloc(&out, (size_t)blocks * 256 * sizeof(float)));
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    const double insts = (double)blocks * 4 * iters * 128;   // wave instructions of the timed loops
    for (int kind = 0; kind < 3; ++kind) {
        for (int rep = 0; rep < 3; ++rep) {
            CHECK(hipEventRecord(e0));
            if (kind == 0) hipLaunchKernelGGL(k_valu_peak<0>, dim3(blocks), dim3(256), 0, 0, out, iters);
            else if (kind == 1) hipLaunchKernelGGL(k_valu_peak<1>, dim3(blocks), dim3(256), 0, 0, out, iters);
            else hipLaunchKernelGGL(k_valu_peak<2>, dim3(blocks), dim3(256), 0, 0, out, iters);
            CHECK(hipGetLastError());
            CHECK(hipEventRecord(e1));
            CHECK(hipEventSynchronize(e1));
            float ms = 0.0f;
            CHECK(hipEventElapsedTime(&ms, e0, e1));
            std::printf("{\"ops\": \"%s\", \"rep\": %d, \"ms\": %.4f, \"wave_insts\": %.0f, \"per_simd_per_ns\": %.4f}\n",
                        kind == 2 ? "s_add_u32" : (kind ? "v_add_u32" : "v_add_f32"), rep, ms, insts,
                        insts / (cus * 4.0) / (ms * 1e6));
        }
    }
    CHECK(hipFree(out));
    return 0;
}
