// Issue-rate probe of the gfx950 VALU instructions the search kernel is built from: the
// denominator of its roofline. Each kernel runs, per lane, 16 independent dependency
// chains of one instruction written in inline assembly (so the compiler can neither fuse
// nor simplify them), interleaved so that consecutive instructions never depend on each
// other. The grid fills every SIMD with `waves` waves. The result per instruction:
//   lane-ops/s              = lanes x iterations x 16 / kernel time (HIP events)
//   cycles per instruction  = shader cycles (s_memtime, per wave, averaged) x SIMDs /
//                             wave-instructions issued on the chip, i.e. how many clocks
//                             one SIMD needs per wave64 instruction
//   clock                   = s_memtime cycles / s_memrealtime time (100 MHz)
// Output: one JSON line per instruction. Register-only work; each lane writes one word of
// `out` and wave 0 lane 0 of each workgroup one timing record, both sized for the grid.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

typedef uint32_t u32;
typedef uint64_t u64;

struct timing {
    u64 t0, t1, r0, r1;
};

enum op_kind {
    OP_MAD_U64_U32,
    OP_MUL_LO_U32,
    OP_MUL_HI_U32,
    OP_MAD_U32_U24,
    OP_ADD_U32,
    OP_ADD3_U32,
    OP_ADD_CO_ADDC,
    OP_LSHL_ADD_U64,
    OP_ALIGNBIT,
    OP_BITOP3,
    OP_XOR,
    OP_LSHLREV_B64,
    OP_COUNT
};

static const char* kNames[OP_COUNT] = {"v_mad_u64_u32",  "v_mul_lo_u32",   "v_mul_hi_u32",
                                       "v_mad_u32_u24",  "v_add_u32",      "v_add3_u32",
                                       "v_add_co_u32+v_addc_co_u32", "v_lshl_add_u64", "v_alignbit_b32",
                                       "v_bitop3_b32",   "v_xor_b32",      "v_lshlrev_b64"};
// VALU instructions per chain step
static const int kInsts[OP_COUNT] = {1, 1, 1, 1, 1, 1, 2, 1, 1, 1, 1, 1};

template <int OP>
__device__ __forceinline__ void step(u32& x, u32& y, u64& w, u32 c) {
    if constexpr (OP == OP_MAD_U64_U32) {
        u64 cc;
        asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(w), "=s"(cc) : "v"(x), "v"(c));
    } else if constexpr (OP == OP_MUL_LO_U32) {
        asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(x) : "v"(c));
    } else if constexpr (OP == OP_MUL_HI_U32) {
        asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(x) : "v"(c));
    } else if constexpr (OP == OP_MAD_U32_U24) {
        asm volatile("v_mad_u32_u24 %0, %0, %1, %2" : "+v"(x) : "v"(c), "v"(y));
    } else if constexpr (OP == OP_ADD_U32) {
        asm volatile("v_add_u32 %0, %0, %1" : "+v"(x) : "v"(c));
    } else if constexpr (OP == OP_ADD3_U32) {
        asm volatile("v_add3_u32 %0, %0, %1, %2" : "+v"(x) : "v"(c), "v"(y));
    } else if constexpr (OP == OP_ADD_CO_ADDC) {
        u64 cc;
        asm volatile("v_add_co_u32 %0, %2, %0, %3\n\tv_addc_co_u32 %1, %2, %1, %4, %2"
                     : "+v"(x), "+v"(y), "=s"(cc)
                     : "v"(c), "v"(c));
    } else if constexpr (OP == OP_LSHL_ADD_U64) {
        asm volatile("v_lshl_add_u64 %0, %0, 0, %1" : "+v"(w) : "v"((u64)c));
    } else if constexpr (OP == OP_ALIGNBIT) {
        asm volatile("v_alignbit_b32 %0, %0, %1, 7" : "+v"(x) : "v"(c));
    } else if constexpr (OP == OP_BITOP3) {
        asm volatile("v_bitop3_b32 %0, %0, %1, %2 bitop3:0xca" : "+v"(x) : "v"(c), "v"(y));
    } else if constexpr (OP == OP_XOR) {
        asm volatile("v_xor_b32 %0, %0, %1" : "+v"(x) : "v"(c));
    } else {
        asm volatile("v_lshlrev_b64 %0, 3, %0" : "+v"(w));
    }
}

template <int OP>
__global__ void __launch_bounds__(64) probe(u32* __restrict__ out, timing* __restrict__ tm, u32 iters, u32 seed) {
    const u32 gid = blockIdx.x * 64 + threadIdx.x;
    u32 x[16], y[16];
    u64 w[16];
    for (int k = 0; k < 16; k++) {
        x[k] = gid * 2654435761u + seed + k;
        y[k] = x[k] ^ 0x9e3779b9u;
        w[k] = ((u64)y[k] << 32) | x[k];
    }
    const u32 c = seed | 1u;
    const u64 t0 = __builtin_amdgcn_s_memtime();
    const u64 r0 = __builtin_amdgcn_s_memrealtime();
    for (u32 it = 0; it < iters; it++) {
#pragma unroll
        for (int k = 0; k < 16; k++) step<OP>(x[k], y[k], w[k], c);
    }
    const u64 t1 = __builtin_amdgcn_s_memtime();
    const u64 r1 = __builtin_amdgcn_s_memrealtime();
    u32 acc = 0;
    for (int k = 0; k < 16; k++) acc ^= x[k] ^ y[k] ^ (u32)w[k] ^ (u32)(w[k] >> 32);
    out[gid] = acc;
    if (threadIdx.x == 0) {
        tm[blockIdx.x].t0 = t0;
        tm[blockIdx.x].t1 = t1;
        tm[blockIdx.x].r0 = r0;
        tm[blockIdx.x].r1 = r1;
    }
}

#define CHECK(call)                                                                        \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess) {                                                            \
            fprintf(stderr, "%s: %s (line %d)\n", #call, hipGetErrorString(e_), __LINE__); \
            exit(1);                                                                       \
        }                                                                                  \
    } while (0)

template <int OP>
static void run(u32 grid, u32* out, timing* tm, u32 iters, u32 simds, double* ms_out, double* cyc_out,
                double* real_out) {
    hipEvent_t a, b;
    CHECK(hipEventCreate(&a));
    CHECK(hipEventCreate(&b));
    hipLaunchKernelGGL(probe<OP>, dim3(grid), dim3(64), 0, 0, out, tm, iters / 8, 12345u);  // warm-up
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(a, 0));
    hipLaunchKernelGGL(probe<OP>, dim3(grid), dim3(64), 0, 0, out, tm, iters, 12345u);
    CHECK(hipGetLastError());
    CHECK(hipEventRecord(b, 0));
    CHECK(hipEventSynchronize(b));
    float ms = 0;
    CHECK(hipEventElapsedTime(&ms, a, b));
    std::vector<timing> h(grid);
    CHECK(hipMemcpy(h.data(), tm, sizeof(timing) * grid, hipMemcpyDeviceToHost));
    double cyc = 0, real = 0;
    for (const timing& t : h) {
        cyc += (double)(t.t1 - t.t0);
        real += (double)(t.r1 - t.r0);
    }
    *ms_out = ms;
    *cyc_out = cyc / grid;
    *real_out = real / grid;
    CHECK(hipEventDestroy(a));
    CHECK(hipEventDestroy(b));
    (void)simds;
}

int main(int argc, char** argv) {
    const u32 waves_per_simd = argc > 1 ? (u32)atoi(argv[1]) : 2;
    const u32 iters = argc > 2 ? (u32)atoi(argv[2]) : 500000;
    hipDeviceProp_t prop;
    CHECK(hipGetDeviceProperties(&prop, 0));
    const u32 simds = (u32)prop.multiProcessorCount * 4;
    const u32 grid = simds * waves_per_simd;
    u32* out = nullptr;
    timing* tm = nullptr;
    CHECK(hipMalloc(&out, sizeof(u32) * 64 * (size_t)grid));
    CHECK(hipMalloc(&tm, sizeof(timing) * (size_t)grid));
    fprintf(stderr, "# %s CUs=%d SIMDs=%u waves/SIMD=%u grid=%u iters=%u\n", prop.gcnArchName,
            prop.multiProcessorCount, simds, waves_per_simd, grid, iters);
    for (int op = 0; op < OP_COUNT; op++) {
        double ms = 0, cyc = 0, real = 0;
        switch (op) {
            case OP_MAD_U64_U32: run<OP_MAD_U64_U32>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_MUL_LO_U32: run<OP_MUL_LO_U32>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_MUL_HI_U32: run<OP_MUL_HI_U32>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_MAD_U32_U24: run<OP_MAD_U32_U24>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_ADD_U32: run<OP_ADD_U32>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_ADD3_U32: run<OP_ADD3_U32>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_ADD_CO_ADDC: run<OP_ADD_CO_ADDC>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_LSHL_ADD_U64: run<OP_LSHL_ADD_U64>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_ALIGNBIT: run<OP_ALIGNBIT>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_BITOP3: run<OP_BITOP3>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            case OP_XOR: run<OP_XOR>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
            default: run<OP_LSHLREV_B64>(grid, out, tm, iters, simds, &ms, &cyc, &real); break;
        }
        const double steps = (double)grid * 64 * iters * 16;  // chain steps, all lanes
        const double insts = steps * kInsts[op];                // lane-instructions
        const double wave_insts_per_simd = (double)waves_per_simd * iters * 16 * kInsts[op];
        const double clock_ghz = cyc / (real / 100e6) / 1e9;  // s_memrealtime runs at 100 MHz
        printf("{\"inst\": \"%s\", \"lane_ops_per_s\": %.4e, \"lane_insts_per_s\": %.4e, \"ms\": %.3f, "
               "\"cycles_per_wave_inst\": %.3f, \"clock_ghz\": %.3f, \"waves_per_simd\": %u}\n",
               kNames[op], steps / (ms * 1e-3), insts / (ms * 1e-3), ms, cyc / wave_insts_per_simd, clock_ghz,
               waves_per_simd);
        fflush(stdout);
    }
    CHECK(hipFree(out));
    CHECK(hipFree(tm));
    return 0;
}
