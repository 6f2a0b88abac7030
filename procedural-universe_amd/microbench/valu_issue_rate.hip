// VALU issue rates on gfx950: register-only loops of exact instructions
// (inline asm, so the compiler neither reorders nor substitutes them), timed by
// the kernel duration (HIP events) and the shader clock measured per workgroup
// against the 100 MHz real-time clock.  It prints lane operations per clock per
// CU for each instruction pattern at 1, 2, 4 and 8 waves per SIMD: the numbers
// behind the direct-sum kernel's issue model (the pair's 12 add/sub/mul/fma and
// one v_rsq_f32, and the rates of 3-VGPR, SGPR, inline-constant and
// transcendental operands).  Every thread writes one float to out[gid]; the
// grid is exactly sized to `out` and `clk`.
//
//   make -C procedural-universe_amd valu_issue_rate && procedural-universe_amd/valu_issue_rate
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <vector>
#include <string>

#define CHECK(x) do { hipError_t e = (x); if (e != hipSuccess) { printf("HIP error %s at %d\n", hipGetErrorString(e), __LINE__); exit(1);} } while (0)

// MODE: instruction pattern per loop body (all on 8 accumulators a0..a7)
//  0: 8 x v_fma_f32 vgpr, independent (distance 8)
//  1: 8 x v_fma_f32 vgpr, chains of 2 (distance 2: a0,a1,a0,a1...)
//  2: 8 x v_fma_f32 vgpr, one chain (distance 1)
//  3: 8 x v_fma_f32 with one SGPR operand, independent
//  4: 8 x v_mul_f32 vgpr, independent
//  5: 8 x v_rsq_f32, independent
//  6: 8 x v_sub_f32 vgpr, independent
//  7: pair-like mix: 4 pairs x (12 vgpr fma/mul/sub) + 4 grouped rsq
//  8: 8 x v_fma_f32 vgpr, chains of 4 (distance 4)
//  9: 8 x v_add_f32 vgpr, independent
// 10: 4 x v_rsq_f32 grouped then 24 vgpr fma (ratio 1:6)
// 11: 8 x v_fma_f32 with a literal constant, independent
template <int MODE>
__global__ void bench(float* out, int iters, float sv, unsigned long long* clk) {
    float a0 = threadIdx.x * 1e-3f + 1.0f, a1 = a0 + 0.1f, a2 = a0 + 0.2f, a3 = a0 + 0.3f;
    float a4 = a0 + 0.4f, a5 = a0 + 0.5f, a6 = a0 + 0.6f, a7 = a0 + 0.7f;
    float b = 0.999f + threadIdx.x * 1e-7f, c = 1e-4f;
    float d0 = 1.f, d1 = 1.f, d2 = 1.f, d3 = 1.f;
    __syncthreads();
    const unsigned long long t0 = clock64();
    const unsigned long long w0 = wall_clock64();
    for (int i = 0; i < iters; ++i) {
        if constexpr (MODE == 0) {
            asm volatile(
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %2, %2, %8, %9\n v_fma_f32 %3, %3, %8, %9\n"
                "v_fma_f32 %4, %4, %8, %9\n v_fma_f32 %5, %5, %8, %9\n v_fma_f32 %6, %6, %8, %9\n v_fma_f32 %7, %7, %8, %9\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(b), "v"(c));
        } else if constexpr (MODE == 1) {
            asm volatile(
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n"
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(b), "v"(c));
        } else if constexpr (MODE == 2) {
            asm volatile(
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %0, %0, %8, %9\n"
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %0, %0, %8, %9\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(b), "v"(c));
        } else if constexpr (MODE == 3) {
            asm volatile(
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %2, %2, %8, %9\n v_fma_f32 %3, %3, %8, %9\n"
                "v_fma_f32 %4, %4, %8, %9\n v_fma_f32 %5, %5, %8, %9\n v_fma_f32 %6, %6, %8, %9\n v_fma_f32 %7, %7, %8, %9\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(b), "s"(sv));
        } else if constexpr (MODE == 4) {
            asm volatile(
                "v_mul_f32 %0, %0, %8\n v_mul_f32 %1, %1, %8\n v_mul_f32 %2, %2, %8\n v_mul_f32 %3, %3, %8\n"
                "v_mul_f32 %4, %4, %8\n v_mul_f32 %5, %5, %8\n v_mul_f32 %6, %6, %8\n v_mul_f32 %7, %7, %8\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(b));
        } else if constexpr (MODE == 5) {
            asm volatile(
                "v_rsq_f32 %0, %0\n v_rsq_f32 %1, %1\n v_rsq_f32 %2, %2\n v_rsq_f32 %3, %3\n"
                "v_rsq_f32 %4, %4\n v_rsq_f32 %5, %5\n v_rsq_f32 %6, %6\n v_rsq_f32 %7, %7\n s_nop 1\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7));
        } else if constexpr (MODE == 6) {
            asm volatile(
                "v_sub_f32 %0, %0, %8\n v_sub_f32 %1, %1, %8\n v_sub_f32 %2, %2, %8\n v_sub_f32 %3, %3, %8\n"
                "v_sub_f32 %4, %4, %8\n v_sub_f32 %5, %5, %8\n v_sub_f32 %6, %6, %8\n v_sub_f32 %7, %7, %8\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(c));
        } else if constexpr (MODE == 7) {
            // 4 "pairs": 12 arithmetic each (interleaved over the 4), then 4 grouped rsq
            asm volatile(
                "v_sub_f32 %0, %8, %4\n v_sub_f32 %1, %8, %5\n v_sub_f32 %2, %8, %6\n v_sub_f32 %3, %8, %7\n"
                "v_sub_f32 %0, %0, %9\n v_sub_f32 %1, %1, %9\n v_sub_f32 %2, %2, %9\n v_sub_f32 %3, %3, %9\n"
                "v_sub_f32 %0, %0, %9\n v_sub_f32 %1, %1, %9\n v_sub_f32 %2, %2, %9\n v_sub_f32 %3, %3, %9\n"
                "v_fma_f32 %0, %0, %0, %9\n v_fma_f32 %1, %1, %1, %9\n v_fma_f32 %2, %2, %2, %9\n v_fma_f32 %3, %3, %3, %9\n"
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %2, %2, %8, %9\n v_fma_f32 %3, %3, %8, %9\n"
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %2, %2, %8, %9\n v_fma_f32 %3, %3, %8, %9\n"
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %2, %2, %8, %9\n v_fma_f32 %3, %3, %8, %9\n"
                "v_mul_f32 %0, %0, %8\n v_mul_f32 %1, %1, %8\n v_mul_f32 %2, %2, %8\n v_mul_f32 %3, %3, %8\n"
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %2, %2, %8, %9\n v_fma_f32 %3, %3, %8, %9\n"
                "v_rsq_f32 %0, %0\n v_rsq_f32 %1, %1\n v_rsq_f32 %2, %2\n v_rsq_f32 %3, %3\n s_nop 1\n"
                "v_fma_f32 %4, %0, %8, %4\n v_fma_f32 %5, %1, %8, %5\n v_fma_f32 %6, %2, %8, %6\n v_fma_f32 %7, %3, %8, %7\n"
                "v_fma_f32 %4, %0, %9, %4\n v_fma_f32 %5, %1, %9, %5\n v_fma_f32 %6, %2, %9, %6\n v_fma_f32 %7, %3, %9, %7\n"
                "v_fma_f32 %4, %0, %8, %4\n v_fma_f32 %5, %1, %8, %5\n v_fma_f32 %6, %2, %8, %6\n v_fma_f32 %7, %3, %8, %7\n"
                : "+v"(d0), "+v"(d1), "+v"(d2), "+v"(d3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(b), "v"(c));
        } else if constexpr (MODE == 8) {
            asm volatile(
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %2, %2, %8, %9\n v_fma_f32 %3, %3, %8, %9\n"
                "v_fma_f32 %0, %0, %8, %9\n v_fma_f32 %1, %1, %8, %9\n v_fma_f32 %2, %2, %8, %9\n v_fma_f32 %3, %3, %8, %9\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(b), "v"(c));
        } else if constexpr (MODE == 9) {
            asm volatile(
                "v_add_f32 %0, %0, %8\n v_add_f32 %1, %1, %8\n v_add_f32 %2, %2, %8\n v_add_f32 %3, %3, %8\n"
                "v_add_f32 %4, %4, %8\n v_add_f32 %5, %5, %8\n v_add_f32 %6, %6, %8\n v_add_f32 %7, %7, %8\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(c));
        } else if constexpr (MODE == 10) {
            asm volatile(
                "v_rsq_f32 %0, %0\n v_rsq_f32 %1, %1\n v_rsq_f32 %2, %2\n v_rsq_f32 %3, %3\n s_nop 1\n"
                "v_fma_f32 %4, %4, %8, %9\n v_fma_f32 %5, %5, %8, %9\n v_fma_f32 %6, %6, %8, %9\n v_fma_f32 %7, %7, %8, %9\n"
                "v_fma_f32 %4, %4, %8, %9\n v_fma_f32 %5, %5, %8, %9\n v_fma_f32 %6, %6, %8, %9\n v_fma_f32 %7, %7, %8, %9\n"
                "v_fma_f32 %4, %4, %8, %9\n v_fma_f32 %5, %5, %8, %9\n v_fma_f32 %6, %6, %8, %9\n v_fma_f32 %7, %7, %8, %9\n"
                "v_fma_f32 %4, %4, %8, %9\n v_fma_f32 %5, %5, %8, %9\n v_fma_f32 %6, %6, %8, %9\n v_fma_f32 %7, %7, %8, %9\n"
                "v_fma_f32 %4, %4, %8, %9\n v_fma_f32 %5, %5, %8, %9\n v_fma_f32 %6, %6, %8, %9\n v_fma_f32 %7, %7, %8, %9\n"
                "v_fma_f32 %4, %4, %8, %9\n v_fma_f32 %5, %5, %8, %9\n v_fma_f32 %6, %6, %8, %9\n v_fma_f32 %7, %7, %8, %9\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(b), "v"(c));
        } else if constexpr (MODE == 11) {
            asm volatile(
                "v_fma_f32 %0, %0, %8, 0.5\n v_fma_f32 %1, %1, %8, 0.5\n v_fma_f32 %2, %2, %8, 0.5\n v_fma_f32 %3, %3, %8, 0.5\n"
                "v_fma_f32 %4, %4, %8, 0.5\n v_fma_f32 %5, %5, %8, 0.5\n v_fma_f32 %6, %6, %8, 0.5\n v_fma_f32 %7, %7, %8, 0.5\n"
                : "+v"(a0), "+v"(a1), "+v"(a2), "+v"(a3), "+v"(a4), "+v"(a5), "+v"(a6), "+v"(a7) : "v"(b));
        }
    }
    const unsigned long long t1 = clock64();
    const unsigned long long w1 = wall_clock64();
    const int gid = blockIdx.x * blockDim.x + threadIdx.x;
    out[gid] = a0 + a1 + a2 + a3 + a4 + a5 + a6 + a7 + d0 + d1 + d2 + d3;
    if (threadIdx.x == 0) {
        clk[2 * blockIdx.x] = t1 - t0;
        clk[2 * blockIdx.x + 1] = w1 - w0;
    }
}

// instructions per loop body, of which rsq
static const int kOps[12] = {8, 8, 8, 8, 8, 8, 8, 52, 8, 8, 28, 8};
static const int kRsq[12] = {0, 0, 0, 0, 0, 8, 0, 4, 0, 0, 4, 0};
static const char* kName[12] = {"fma vgpr dist8", "fma vgpr dist2", "fma vgpr dist1", "fma sgpr dist8", "mul vgpr dist8",
                                "rsq dist8", "sub vgpr dist8", "pair mix 12+rsq x4", "fma vgpr dist4", "add vgpr dist8",
                                "rsq4 + 24 fma", "fma literal dist8"};

template <int MODE>
void run(int waves_per_simd, int cus, float* out, unsigned long long* clk) {
    const int block = 256;  // 4 waves: one per SIMD
    const int blocks = cus * waves_per_simd;
    const int iters = 20000;
    hipEvent_t e0, e1;
    CHECK(hipEventCreate(&e0));
    CHECK(hipEventCreate(&e1));
    hipLaunchKernelGGL(bench<MODE>, dim3(blocks), dim3(block), 0, 0, out, 100, 1.0f, clk);  // warm
    CHECK(hipDeviceSynchronize());
    CHECK(hipEventRecord(e0));
    hipLaunchKernelGGL(bench<MODE>, dim3(blocks), dim3(block), 0, 0, out, iters, 1.0f, clk);
    CHECK(hipEventRecord(e1));
    CHECK(hipDeviceSynchronize());
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, e0, e1));
    std::vector<unsigned long long> h(2 * blocks);
    CHECK(hipMemcpy(h.data(), clk, h.size() * sizeof(unsigned long long), hipMemcpyDeviceToHost));
    double cyc = 0, wall = 0;
    for (int b = 0; b < blocks; ++b) { cyc += h[2 * b]; wall += h[2 * b + 1]; }
    cyc /= blocks; wall /= blocks;
    const double ghz = cyc / (wall / 100e6) / 1e9;  // shader clock over the 100 MHz wall clock
    const double lane_ops = double(blocks) * block * iters * kOps[MODE];
    // per CU per clock, from the kernel duration and the measured clock
    const double rate = lane_ops / (ms * 1e-3 * ghz * 1e9 * cus);
    // per wave: cycles per instruction (wave64) from the wave's own clock
    const double cpi = cyc / (double(iters) * kOps[MODE]);
    printf("mode %2d %-20s waves/SIMD %2d  %8.3f ms  %.3f GHz  %7.1f lane-ops/clk/CU  %6.2f clk/instr/wave  (rsq %d of %d)\n",
           MODE, kName[MODE], waves_per_simd, ms, ghz, rate, cpi, kRsq[MODE], kOps[MODE]);
    fflush(stdout);
    CHECK(hipEventDestroy(e0));
    CHECK(hipEventDestroy(e1));
}

int main() {
    hipDeviceProp_t p;
    CHECK(hipGetDeviceProperties(&p, 0));
    const int cus = p.multiProcessorCount;
    printf("device %s, %d CUs, clock %d kHz\n", p.gcnArchName, cus, p.clockRate);
    const int max_w = 8;
    float* out = nullptr;
    unsigned long long* clk = nullptr;
    CHECK(hipMalloc(&out, size_t(cus) * max_w * 256 * sizeof(float)));
    CHECK(hipMalloc(&clk, size_t(cus) * max_w * 2 * sizeof(unsigned long long)));
    for (int w : {1, 2, 4, 8}) {
        run<0>(w, cus, out, clk);
        run<1>(w, cus, out, clk);
        run<2>(w, cus, out, clk);
        run<8>(w, cus, out, clk);
        run<3>(w, cus, out, clk);
        run<11>(w, cus, out, clk);
        run<4>(w, cus, out, clk);
        run<6>(w, cus, out, clk);
        run<9>(w, cus, out, clk);
        run<5>(w, cus, out, clk);
        run<10>(w, cus, out, clk);
        run<7>(w, cus, out, clk);
    }
    CHECK(hipFree(out));
    CHECK(hipFree(clk));
    return 0;
}
