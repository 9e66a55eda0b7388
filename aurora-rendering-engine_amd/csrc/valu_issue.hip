// valu_issue — issue rate of gfx950's vector ALU for the instruction mixes
// of the path kernel: how many wave64 VALU instructions per second the chip
// retires for one instruction kind alone, and for pairs of kinds
// interleaved.  It is the denominator of the path kernel's issue roofline.
//
// Each kernel runs kChains independent dependency chains per lane, every
// instruction written in inline assembly so that the count is exact: per
// loop iteration kChains * kReps VALU instructions (the loop's own counter
// and branch are scalar instructions).  The grid holds 8 waves per SIMD on
// every CU, as the path kernel does.  No kernel reads memory inside its loop;
// each writes one value per thread at the end, so the compiler keeps the
// chains.
//
//   valu_issue            prints one line per mix: ms per launch, G wave64
//                         instructions per second, cycles per instruction per SIMD
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

#define HIP_OK(x)                                                                          \
    do {                                                                                   \
        hipError_t e_ = (x);                                                               \
        if (e_ != hipSuccess) {                                                            \
            std::fprintf(stderr, "valu_issue: %s at line %d\n", hipGetErrorString(e_), __LINE__); \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

namespace {

constexpr int kChains = 8;
constexpr int kReps = 8;
constexpr int kIters = 2048;
constexpr int kBlock = 256;

enum Mix {
    FMA_F32,
    ADD_F32,
    MUL_F32,
    MAX_F32,
    MIN_MAX_F32,
    FMA_MAX,
    ADD_MAX,
    FMA_MIN_MAX,
    FMA_CNDMASK,
    FMA_CMP,
    FMA_XOR,
    XOR_MAX,
    MULHI_MAX,
    FMA_F64,
    ADD_S,      // v_add_f32 with a scalar-register operand
    ADD_S_ADD,  // the same beside v_add_f32 on vector registers only
    ADD_S_MAX,
    SUB_S_MUL,  // a slab test's (scalar box - origin) * reciprocal
    MULLO,      // v_mul_lo_u32
    MULHI_MULLO,
    MAD_U64,    // v_mad_u64_u32: a Philox round's 32x32 -> 64-bit product
    MAD_U64_FMA,
};

// one instruction on chain j; kind k of the mix's pair
template <Mix M, int J>
__device__ __forceinline__ void step(float (&a)[kChains], uint32_t (&x)[kChains], double (&d)[kChains], float b,
                                     float c, uint64_t mask, uint64_t& cm, uint64_t (&w)[kChains]) {
    constexpr bool even = (J % 2) == 0;
    // a wave-uniform value in a scalar register
    const float sb = __builtin_bit_cast(float, __builtin_amdgcn_readfirstlane(__builtin_bit_cast(uint32_t, b)));
    if constexpr (M == ADD_S || ((M == ADD_S_ADD || M == ADD_S_MAX) && even)) {
        asm volatile("v_add_f32 %0, %1, %0" : "+v"(a[J]) : "s"(sb));
    } else if constexpr (M == ADD_S_ADD) {
        asm volatile("v_add_f32 %0, %0, %1" : "+v"(a[J]) : "v"(c));
    } else if constexpr (M == ADD_S_MAX) {
        asm volatile("v_max_f32 %0, %0, %1" : "+v"(a[J]) : "v"(c));
    } else if constexpr (M == SUB_S_MUL) {
        if constexpr (even)
            asm volatile("v_sub_f32 %0, %1, %0" : "+v"(a[J]) : "s"(sb));
        else
            asm volatile("v_mul_f32 %0, %0, %1" : "+v"(a[J]) : "v"(c));
    } else if constexpr (M == FMA_F32 || ((M == FMA_MAX || M == FMA_CNDMASK || M == FMA_CMP || M == FMA_XOR) && even) ||
                  (M == FMA_MIN_MAX && J % 3 == 0)) {
        asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a[J]) : "v"(b), "v"(c));
    } else if constexpr (M == ADD_F32 || (M == ADD_MAX && even)) {
        asm volatile("v_add_f32 %0, %0, %1" : "+v"(a[J]) : "v"(b));
    } else if constexpr (M == MUL_F32) {
        asm volatile("v_mul_f32 %0, %0, %1" : "+v"(a[J]) : "v"(b));
    } else if constexpr (M == MAX_F32 || ((M == FMA_MAX || M == ADD_MAX || M == XOR_MAX || M == MULHI_MAX) && !even) ||
                         (M == MIN_MAX_F32 && !even) || (M == FMA_MIN_MAX && J % 3 == 1)) {
        asm volatile("v_max_f32 %0, %0, %1" : "+v"(a[J]) : "v"(c));
    } else if constexpr ((M == MIN_MAX_F32 && even) || (M == FMA_MIN_MAX && J % 3 == 2)) {
        asm volatile("v_min_f32 %0, %0, %1" : "+v"(a[J]) : "v"(b));
    } else if constexpr (M == FMA_CNDMASK) {
        asm volatile("v_cndmask_b32_e64 %0, %0, %1, %2" : "+v"(a[J]) : "v"(b), "s"(mask));
    } else if constexpr (M == FMA_CMP) {
        uint64_t r;
        asm volatile("v_cmp_lt_f32_e64 %0, %1, %2" : "=s"(r) : "v"(a[J]), "v"(c));
        cm ^= r;
    } else if constexpr ((M == FMA_XOR && !even) || (M == XOR_MAX && even)) {
        asm volatile("v_xor_b32 %0, %0, %1" : "+v"(x[J]) : "v"(x[(J + 1) % kChains]));
    } else if constexpr (M == MULHI_MAX) {
        asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(x[J]) : "v"(x[(J + 1) % kChains]));
    } else if constexpr (M == MULLO || (M == MULHI_MULLO && !even)) {
        asm volatile("v_mul_lo_u32 %0, %0, %1" : "+v"(x[J]) : "v"(x[(J + 1) % kChains]));
    } else if constexpr (M == MULHI_MULLO) {
        asm volatile("v_mul_hi_u32 %0, %0, %1" : "+v"(x[J]) : "v"(x[(J + 1) % kChains]));
    } else if constexpr (M == MAD_U64 || (M == MAD_U64_FMA && even)) {
        uint64_t carry;
        asm volatile("v_mad_u64_u32 %0, %1, %2, %3, %0" : "+v"(w[J]), "=s"(carry) : "v"(x[J]), "v"(x[(J + 1) % kChains]));
    } else if constexpr (M == MAD_U64_FMA) {
        asm volatile("v_fma_f32 %0, %0, %1, %2" : "+v"(a[J]) : "v"(b), "v"(c));
    } else if constexpr (M == FMA_F64) {
        const double db = b, dc = c;
        asm volatile("v_fma_f64 %0, %0, %1, %2" : "+v"(d[J]) : "v"(db), "v"(dc));
    }
}

template <Mix M, int J>
__device__ __forceinline__ void chains(float (&a)[kChains], uint32_t (&x)[kChains], double (&d)[kChains], float b,
                                       float c, uint64_t mask, uint64_t& cm, uint64_t (&w)[kChains]) {
    step<M, J>(a, x, d, b, c, mask, cm, w);
    if constexpr (J + 1 < kChains) chains<M, J + 1>(a, x, d, b, c, mask, cm, w);
}

template <Mix M>
__global__ __launch_bounds__(kBlock) void k_issue(float* out, float b, float c) {
    float a[kChains];
    uint32_t x[kChains];
    double d[kChains];
    uint64_t w[kChains];
    for (int j = 0; j < kChains; ++j) {
        a[j] = (float)threadIdx.x * 1e-3f + (float)j;
        x[j] = threadIdx.x * 2654435761u + (uint32_t)j;
        d[j] = (double)a[j];
        w[j] = x[j];
    }
    const uint64_t mask = 0x5555555555555555ull;
    uint64_t cm = 0;
    for (int it = 0; it < kIters; ++it) {
#pragma unroll
        for (int r = 0; r < kReps; ++r) chains<M, 0>(a, x, d, b, c, mask, cm, w);
    }
    float s = (float)(cm & 1u);
    for (int j = 0; j < kChains; ++j) s += a[j] + (float)(x[j] & 255u) + (float)d[j] + (float)(w[j] & 255u);
    out[(size_t)blockIdx.x * kBlock + threadIdx.x] = s;
}

struct Case {
    const char* name;
    void (*kernel)(float*, float, float);
};

}  // namespace

int main() {
    hipDeviceProp_t p;
    HIP_OK(hipGetDeviceProperties(&p, 0));
    const int cus = p.multiProcessorCount;
    const int grid = cus * 8;  // 8 blocks of 4 waves per CU: 8 waves per SIMD
    float* out = nullptr;
    HIP_OK(hipMalloc(&out, sizeof(float) * (size_t)grid * kBlock));
    const Case cases[] = {
        {"v_fma_f32", k_issue<FMA_F32>},
        {"v_add_f32", k_issue<ADD_F32>},
        {"v_mul_f32", k_issue<MUL_F32>},
        {"v_max_f32", k_issue<MAX_F32>},
        {"v_min_f32 + v_max_f32", k_issue<MIN_MAX_F32>},
        {"v_fma_f32 + v_max_f32", k_issue<FMA_MAX>},
        {"v_add_f32 + v_max_f32", k_issue<ADD_MAX>},
        {"v_fma_f32 + v_min/max_f32 (1:2)", k_issue<FMA_MIN_MAX>},
        {"v_fma_f32 + v_cndmask_b32", k_issue<FMA_CNDMASK>},
        {"v_fma_f32 + v_cmp_lt_f32", k_issue<FMA_CMP>},
        {"v_fma_f32 + v_xor_b32", k_issue<FMA_XOR>},
        {"v_xor_b32 + v_max_f32", k_issue<XOR_MAX>},
        {"v_mul_hi_u32 + v_max_f32", k_issue<MULHI_MAX>},
        {"v_fma_f64", k_issue<FMA_F64>},
        {"v_add_f32 (scalar operand)", k_issue<ADD_S>},
        {"v_add_f32 (scalar) + v_add_f32", k_issue<ADD_S_ADD>},
        {"v_add_f32 (scalar) + v_max_f32", k_issue<ADD_S_MAX>},
        {"v_sub_f32 (scalar) + v_mul_f32", k_issue<SUB_S_MUL>},
        {"v_mul_lo_u32", k_issue<MULLO>},
        {"v_mul_hi_u32 + v_mul_lo_u32", k_issue<MULHI_MULLO>},
        {"v_mad_u64_u32", k_issue<MAD_U64>},
        {"v_mad_u64_u32 + v_fma_f32", k_issue<MAD_U64_FMA>},
    };
    hipEvent_t e0, e1;
    HIP_OK(hipEventCreate(&e0));
    HIP_OK(hipEventCreate(&e1));
    const double ghz = p.clockRate * 1e-6;
    std::printf("%d CUs, peak engine clock %.2f GHz, %d waves per SIMD, %d VALU instructions per wave\n", cus, ghz, 8,
                kIters * kReps * kChains);
    constexpr int kLaunches = 5;
    for (const Case& k : cases) {
        hipLaunchKernelGGL(k.kernel, dim3(grid), dim3(kBlock), 0, 0, out, 1.0001f, 0.5f);
        HIP_OK(hipGetLastError());
        HIP_OK(hipDeviceSynchronize());
        HIP_OK(hipEventRecord(e0));
        for (int r = 0; r < kLaunches; ++r) hipLaunchKernelGGL(k.kernel, dim3(grid), dim3(kBlock), 0, 0, out, 1.0001f, 0.5f);
        HIP_OK(hipEventRecord(e1));
        HIP_OK(hipEventSynchronize(e1));
        float ms = 0.0f;
        HIP_OK(hipEventElapsedTime(&ms, e0, e1));
        const double insts = (double)grid * (kBlock / 64) * kIters * kReps * kChains * kLaunches;
        const double rate = insts / (ms * 1e-3);
        std::printf("%-34s %7.3f ms  %7.1f G wave64 instructions/s  %.2f cycles per instruction per SIMD\n", k.name,
                    ms / kLaunches, rate * 1e-9, cus * 4.0 * ghz * 1e9 / rate);
    }
    HIP_OK(hipFree(out));
    return 0;
}
