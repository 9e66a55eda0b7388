// unet_kernels.hip — denoiser kernels for gfx950: 3x3 convolution as an
// implicit GEMM on MFMA (v_mfma_f32_16x16x32_bf16: BF16 operands, FP32
// accumulation) for the forward pass, the data gradient and the weight
// gradient, plus the elementwise pieces of the U-Net (average pooling,
// concat/upsample gradient split, ReLU masks, input preparation, loss, Adam).
//
// Tensors are NHWC.  Activations and the gradients fed to MFMA are BF16;
// gradients that are accumulated are FP32.  A convolution input is the
// concatenation of a full-resolution tensor A (Ca channels) and a
// half-resolution tensor B (Cb channels) read through nearest upsampling, so
// the decoder's upsample + skip concat is never materialised.  Ca, Cb and
// Cout are multiples of 16.  Weights of a layer are [Cout][9][Cin] (k =
// tap*Cin + c, tap = 3*(dy+1) + (dx+1)), padded to Kpad = ceil(9*Cin/32)*32.
//
// MFMA operand layout (16x16x32, wave64): lane l holds A[l%16][8*(l/16)+j]
// and B[8*(l/16)+j][l%16] for j = 0..7 and returns D[4*(l/16)+r][l%16] for
// r = 0..3 (checked on the device by aurora_selftest_mfma).
//
// No reference counterpart: the reference has no learned component.
#include "unet.hpp"

#include <algorithm>

namespace aur {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef uint16_t u16;

__device__ __forceinline__ float bf2f(u16 h) { return __uint_as_float((uint32_t)h << 16); }
// round to nearest even (gfx950's v_cvt_pk_bf16_f32; NaN stays NaN)
__device__ __forceinline__ u16 f2bf(float f) { return __builtin_bit_cast(u16, (__bf16)f); }

union Pack8 {
    uint4 u;
    bf16x8 v;
    u16 h[8];
};

__device__ __forceinline__ bf16x8 zero8() {
    Pack8 p;
    p.u = make_uint4(0, 0, 0, 0);
    return p.v;
}

// 8 consecutive channels [c, c+8) of the concatenated input at (n, y, x);
// zero outside the image.
__device__ __forceinline__ bf16x8 gather8(const ConvIn& in, int n, int y, int x, int H, int W, int c) {
    if (y < 0 || y >= H || x < 0 || x >= W) return zero8();
    Pack8 p;
    if (c < in.ca) {
        const u16* src = in.a + (((size_t)n * H + y) * W + x) * in.ca + c;
        p.u = *reinterpret_cast<const uint4*>(src);
    } else {
        const int h2 = H >> 1, w2 = W >> 1;
        const u16* src = in.b + (((size_t)n * h2 + (y >> 1)) * w2 + (x >> 1)) * in.cb + (c - in.ca);
        p.u = *reinterpret_cast<const uint4*>(src);
    }
    return p.v;
}

// one channel of the concatenated input (weight-gradient operand gather)
__device__ __forceinline__ u16 gather1(const ConvIn& in, int n, int y, int x, int H, int W, int c) {
    if (y < 0 || y >= H || x < 0 || x >= W) return 0;
    if (c < in.ca) return in.a[(((size_t)n * H + y) * W + x) * in.ca + c];
    const int h2 = H >> 1, w2 = W >> 1;
    return in.b[(((size_t)n * h2 + (y >> 1)) * w2 + (x >> 1)) * in.cb + (c - in.ca)];
}

// ------------------------------------------------------------ forward ---
// Block: 4 waves; each wave a 64-pixel segment of one image row (4 MFMA row
// tiles of 16 pixels) x Cout (NT column tiles of 16 channels), so every
// weight fragment feeds 4 MFMAs and every input fragment NT.  K = 9*Cin in
// chunks of 32; the next chunk's fragments are loaded before this chunk's
// MFMAs issue, so the loads overlap the matrix work.
template <int NT, bool RELU, bool OUT_F32, bool ACCUM>
__global__ __launch_bounds__(256) void k_conv3x3(ConvIn in, const u16* __restrict__ wk, int kpad,
                                                 const float* __restrict__ bias, void* out, int N, int H, int W,
                                                 int out_stride, int co_base) {
    // blockIdx.y: the group of 16*NT output channels
    co_base += blockIdx.y * 16 * NT;
    wk += (size_t)blockIdx.y * 16 * NT * kpad;
    if (bias) bias += blockIdx.y * 16 * NT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int segs = (W + 63) / 64;
    const int seg = blockIdx.x * 4 + wave;
    const int row = seg / segs;
    const int x0 = (seg % segs) * 64;
    const int n = row / H, y = row % H;
    if (n >= N) return;  // the whole wave: seg is wave-uniform
    const int cin = in.ca + in.cb;
    const int K = 9 * cin;
    const int m = lane & 15;   // pixel of the A fragment within its row tile
    const int kg = lane >> 4;  // k group of 8
    floatx4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    bf16x8 a[4], b[NT];
    auto load = [&](int k0, bf16x8* fa, bf16x8* fb) {
        const int k = k0 + 8 * kg;
        if (k < K) {
            const int tap = k / cin, c = k - tap * cin;
            const int yy = y + tap / 3 - 1, dx = tap % 3 - 1;
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) fa[mt] = gather8(in, n, yy, x0 + 16 * mt + m + dx, H, W, c);
        } else {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) fa[mt] = zero8();
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            Pack8 w;
            w.u = *reinterpret_cast<const uint4*>(wk + (size_t)(t * 16 + m) * kpad + k);
            fb[t] = w.v;
        }
    };
    load(0, a, b);
    for (int k0 = 0; k0 < kpad; k0 += 32) {
        bf16x8 na[4], nb[NT];
        const bool more = k0 + 32 < kpad;
        if (more) load(k0 + 32, na, nb);
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int t = 0; t < NT; ++t) acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(b[t], a[mt], acc[mt][t], 0, 0, 0);
        if (more) {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) a[mt] = na[mt];
#pragma unroll
            for (int t = 0; t < NT; ++t) b[t] = nb[t];
        }
    }
    // the MFMAs computed each tile transposed (weights as the A operand),
    // so this lane holds channels co .. co+3 of pixel x: one 8-byte (BF16)
    // or 16-byte (FP32) store each; out_stride, co_base and co are multiples
    // of 4, so the stores are aligned
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + 4 * kg;
        float bv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) bv[r] = bias ? bias[co + r] : 0.0f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int x = x0 + 16 * mt + m;
            if (x >= W) continue;
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = acc[mt][t][r] + bv[r];
                if (RELU) v[r] = fmaxf(v[r], 0.0f);
            }
            const size_t o = (((size_t)n * H + y) * W + x) * out_stride + co_base + co;
            if (OUT_F32) {
                float4* of = reinterpret_cast<float4*>(static_cast<float*>(out) + o);
                float4 w4 = make_float4(v[0], v[1], v[2], v[3]);
                if (ACCUM) {
                    const float4 c4 = *of;
                    w4 = make_float4(c4.x + v[0], c4.y + v[1], c4.z + v[2], c4.w + v[3]);
                }
                *of = w4;
            } else {
                const uint2 h = make_uint2((uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                                           (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16));
                *reinterpret_cast<uint2*>(static_cast<u16*>(out) + o) = h;
            }
        }
    }
}

// Same convolution with the input tile staged in LDS: a block of 4 waves
// computes 4 rows x 64 pixels x (16*NT) channels.  It first copies its input
// tile into LDS with coalesced 16-byte loads, once, instead of every wave
// gathering each of its 9 taps from L1/L2: the full-resolution part (ca
// channels) as 6 x 66 pixels, the halo included, and the half-resolution
// part of a concat (cb channels, read through nearest upsampling) at its own
// resolution, 4 x 34 pixels; zeros outside the image.  A pixel's channels sit
// at a stride of 2*C + 16 bytes: 16 consecutive pixels then start in
// distinct LDS banks, so an A fragment (16 pixels x 8 channels per k group)
// reads without bank conflicts.
__host__ __device__ constexpr int lds_pix_stride(int c) { return 2 * c + 16; }

// Index i in a mixed radix of three digits, the first fastest:
// i = (r * radix_px + px) * radix_g + g.  Decomposed once for i = t, then
// advanced by 256 (the block's threads) with two carries per step instead of
// three integer divisions by runtime radices.
struct Radix256 {
    int g, px, r, dg, dpx, dr, rg, rpx;
    __device__ Radix256(int t, int radix_g, int radix_px) : rg(radix_g), rpx(radix_px) {
        const int q0 = t / radix_g;
        g = t - q0 * radix_g;
        px = q0 % radix_px;
        r = q0 / radix_px;
        const int qs = 256 / radix_g;  // 256 = qs * radix_g + dg, qs = dr * radix_px + dpx
        dg = 256 - qs * radix_g;
        dpx = qs % radix_px;
        dr = qs / radix_px;
    }
    __device__ void step() {
        g += dg;
        const int c1 = g >= rg ? 1 : 0;
        g -= c1 * rg;
        px += dpx + c1;
        const int c2 = px >= rpx ? 1 : 0;
        px -= c2 * rpx;
        r += dr + c2;
    }
};
__host__ __device__ constexpr int lds_tile_bytes(int ca, int cb) {
    return 6 * 66 * lds_pix_stride(ca) + (cb ? 4 * 34 * lds_pix_stride(cb) : 0);
}

// HB: the input has a half-resolution part (cb > 0)
template <int NT, bool RELU, bool OUT_F32, bool ACCUM, bool HB>
__global__ __launch_bounds__(256) void k_conv3x3_lds(ConvIn in, const u16* __restrict__ wk, int kpad,
                                                     const float* __restrict__ bias, void* out, int N, int H, int W,
                                                     int out_stride, int co_base) {
    extern __shared__ __align__(16) unsigned char tile[];
    // blockIdx.y: the group of 16*NT output channels
    co_base += blockIdx.y * 16 * NT;
    wk += (size_t)blockIdx.y * 16 * NT * kpad;
    if (bias) bias += blockIdx.y * 16 * NT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int segs = (W + 63) / 64, bands = (H + 3) / 4;
    const int x0 = (blockIdx.x % segs) * 64;
    const int band = (blockIdx.x / segs) % bands;
    const int n = blockIdx.x / (segs * bands);
    const int y0 = band * 4;
    const int ca = in.ca, cb = HB ? in.cb : 0, cin = ca + cb;
    const int psa = lds_pix_stride(ca), psb = lds_pix_stride(cb);
    unsigned char* const tb = tile + 6 * 66 * psa;  // the half-resolution part
    // full-resolution part: rows y0-1 .. y0+4, pixels x0-1 .. x0+64.  Each
    // thread issues all its 16-byte loads before its first LDS store, so a
    // block waits for memory once, not once per load (kStageLoads per
    // thread cover 6 x 66 pixels of up to 64 channels: 3168 pieces)
    const int ga = ca >> 3;  // 16-byte channel groups per pixel
    constexpr int kStageLoads = 13;
    // Pieces outside the image load from a valid clamped address and are
    // zeroed by a select, without branches; element offsets fit 32 bits (the
    // host checks N*H*W*C < 2^31); the unrolled loops stop at the
    // wave-uniform piece count.
    {
        const int total = 6 * 66 * ga;
        // pixel fastest: consecutive lanes write one channel group of
        // consecutive pixels, whose addresses (stride 2*C + 16 bytes) fall in
        // distinct LDS banks
        Radix256 it((int)threadIdx.x, 66, 6);  // (pixel, row, group)
        uint4 v[kStageLoads];
        int dst[kStageLoads];
#pragma unroll
        for (int j = 0; j < kStageLoads; ++j) {
            v[j] = make_uint4(0, 0, 0, 0);
            dst[j] = -1;
            if (j * 256 < total) {  // wave-uniform
                const int px = it.g, r = it.px, g = it.r;
                const int y = y0 - 1 + r, x = x0 - 1 + px;
                const bool in_img = (unsigned)y < (unsigned)H && (unsigned)x < (unsigned)W;
                const int off = ((n * H + (in_img ? y : 0)) * W + (in_img ? x : 0)) * ca + 8 * (g < ga ? g : 0);
                const uint4 l = *reinterpret_cast<const uint4*>(in.a + off);
                v[j] = in_img ? l : make_uint4(0, 0, 0, 0);
                dst[j] = g < ga ? (r * 66 + px) * psa + 16 * g : -1;
                it.step();
            }
        }
#pragma unroll
        for (int j = 0; j < kStageLoads; ++j)
            if (dst[j] >= 0) *reinterpret_cast<uint4*>(tile + dst[j]) = v[j];
    }
    // half-resolution part: the half-rows and half-pixels the full-resolution
    // rows y0-1 .. y0+4 and pixels x0-1 .. x0+64 fall in (arithmetic shifts:
    // -1 maps to -1, outside); 4 x 34 pixels of up to 64 channels: 1088 pieces
    const int hy0 = (y0 - 1) >> 1, hx0 = (x0 - 1) >> 1, h2 = H >> 1, w2 = W >> 1;
    const int gb = cb >> 3;
    if (HB) {
        constexpr int kHalfLoads = 5;
        const int total = 4 * 34 * gb;
        Radix256 it((int)threadIdx.x, 34, 4);  // (pixel, row, group)
        uint4 v[kHalfLoads];
        int dst[kHalfLoads];
#pragma unroll
        for (int j = 0; j < kHalfLoads; ++j) {
            v[j] = make_uint4(0, 0, 0, 0);
            dst[j] = -1;
            if (j * 256 < total) {  // wave-uniform
                const int px = it.g, r = it.px, g = it.r;
                const int y = hy0 + r, x = hx0 + px;
                const bool in_img = (unsigned)y < (unsigned)h2 && (unsigned)x < (unsigned)w2;
                const int off = ((n * h2 + (in_img ? y : 0)) * w2 + (in_img ? x : 0)) * cb + 8 * (g < gb ? g : 0);
                const uint4 l = *reinterpret_cast<const uint4*>(in.b + off);
                v[j] = in_img ? l : make_uint4(0, 0, 0, 0);
                dst[j] = g < gb ? (r * 34 + px) * psb + 16 * g : -1;
                it.step();
            }
        }
#pragma unroll
        for (int j = 0; j < kHalfLoads; ++j)
            if (dst[j] >= 0) *reinterpret_cast<uint4*>(tb + dst[j]) = v[j];
    }
    __syncthreads();
    const int y = y0 + wave;
    if (y >= H) return;  // no barrier follows
    const int m = lane & 15;   // pixel of the A fragment within its row tile
    const int kg = lane >> 4;  // k group of 8
    floatx4 acc[4][NT];
#pragma unroll
    for (int mt = 0; mt < 4; ++mt)
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[mt][t] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // This lane's k group (k = k0 + 8 kg) as tap row ty, tap column tx and
    // channel c, advanced by 32 per chunk with a few integer adds (a chunk
    // crosses at most two taps: cin >= 16); ty == 3 is past the last tap.
    int ty = 0, tx = 0, c = 8 * kg;
    auto advance = [&](int step) {
        c += step;
        while (c >= cin) {
            c -= cin;
            if (++tx == 3) {
                tx = 0;
                ++ty;
            }
        }
    };
    advance(0);
    const u16* wrow[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) wrow[t] = wk + (size_t)(t * 16 + m) * kpad + 8 * kg;
    auto load = [&](bf16x8* fa, bf16x8* fb) {
        if (ty < 3) {
            if (!HB || c < ca) {
                // tile row of input row y + ty - 1, tile pixel of x0 + m + tx - 1
                const unsigned char* src = tile + ((wave + ty) * 66 + m + tx) * psa + 2 * c;
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    Pack8 v;
                    v.u = *reinterpret_cast<const uint4*>(src + 16 * mt * psa);
                    fa[mt] = v.v;
                }
            } else {
                // input pixel (y + ty - 1, x) read at half resolution
                const int hr = ((y + ty - 1) >> 1) - hy0;
#pragma unroll
                for (int mt = 0; mt < 4; ++mt) {
                    const int hp = ((x0 + 16 * mt + m + tx - 1) >> 1) - hx0;
                    Pack8 v;
                    v.u = *reinterpret_cast<const uint4*>(tb + (hr * 34 + hp) * psb + 2 * (c - ca));
                    fa[mt] = v.v;
                }
            }
        } else {
#pragma unroll
            for (int mt = 0; mt < 4; ++mt) fa[mt] = zero8();
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
            Pack8 w;
            w.u = *reinterpret_cast<const uint4*>(wrow[t]);
            fb[t] = w.v;
            wrow[t] += 32;
        }
        advance(32);
    };
    auto mma = [&](const bf16x8* fa, const bf16x8* fb) {
#pragma unroll
        for (int mt = 0; mt < 4; ++mt)
#pragma unroll
            for (int t = 0; t < NT; ++t)
                acc[mt][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fb[t], fa[mt], acc[mt][t], 0, 0, 0);
    };
    // two fragment sets in turn: chunk q + 1 loads while chunk q multiplies.
    // A trip takes two chunks and has no exit in its middle (an odd last
    // chunk follows the loop): with an exit after each chunk the compiler
    // copied all 16 * NT accumulators out of the MFMA registers and back on
    // every trip.  kpad is a multiple of 32.
    bf16x8 a0[4], b0[NT], a1[4], b1[NT];
    load(a0, b0);
    int rest = kpad / 32 - 1;  // chunks after the first
    for (; rest >= 2; rest -= 2) {
        load(a1, b1);
        mma(a0, b0);
        load(a0, b0);
        mma(a1, b1);
    }
    if (rest == 1) {
        load(a1, b1);
        mma(a0, b0);
        mma(a1, b1);
    } else {
        mma(a0, b0);
    }
    // the MFMAs computed each tile transposed (weights as the A operand),
    // so this lane holds channels co .. co+3 of pixel x: one 8-byte (BF16)
    // or 16-byte (FP32) store each; out_stride, co_base and co are multiples
    // of 4, so the stores are aligned
#pragma unroll
    for (int t = 0; t < NT; ++t) {
        const int co = t * 16 + 4 * kg;
        float bv[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) bv[r] = bias ? bias[co + r] : 0.0f;
#pragma unroll
        for (int mt = 0; mt < 4; ++mt) {
            const int x = x0 + 16 * mt + m;
            if (x >= W) continue;
            float v[4];
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                v[r] = acc[mt][t][r] + bv[r];
                if (RELU) v[r] = fmaxf(v[r], 0.0f);
            }
            const size_t o = (((size_t)n * H + y) * W + x) * out_stride + co_base + co;
            if (OUT_F32) {
                float4* of = reinterpret_cast<float4*>(static_cast<float*>(out) + o);
                float4 w4 = make_float4(v[0], v[1], v[2], v[3]);
                if (ACCUM) {
                    const float4 c4 = *of;
                    w4 = make_float4(c4.x + v[0], c4.y + v[1], c4.z + v[2], c4.w + v[3]);
                }
                *of = w4;
            } else {
                const uint2 h = make_uint2((uint32_t)f2bf(v[0]) | ((uint32_t)f2bf(v[1]) << 16),
                                           (uint32_t)f2bf(v[2]) | ((uint32_t)f2bf(v[3]) << 16));
                *reinterpret_cast<uint2*>(static_cast<u16*>(out) + o) = h;
            }
        }
    }
}

// ---------------------------------------------------- weight gradient ---
// dW[co][tap][ci] = sum_p g[p][co] * X[p + tap][ci] over the rows of this
// block: MFMA with M = 16 output channels, N = 16 input channels, K = 32
// pixels of a row.  A block owns one (co tile, ci tile) pair and all 9 taps:
// per 32-pixel chunk a wave loads its g fragment once and, for each of the 3
// input rows, a 10-pixel window of X from which the 3 column shifts are
// taken in registers, then issues 9 MFMAs.  Each of the 4 waves walks every
// 4th row; the partial tiles are summed in LDS and written per block and tap
// (reduced afterwards in a fixed order, so the result is deterministic).
__global__ __launch_bounds__(256) void k_conv3x3_dw(ConvIn in, const u16* __restrict__ g, int cout, int N, int H,
                                                    int W, int rows_per_block, float* __restrict__ partial) {
    __shared__ float red[4][256];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int cin = in.ca + in.cb;
    const int n_ci = cin / 16, n_co = cout / 16;
    const int pair = blockIdx.y;  // (co tile, ci tile)
    const int co0 = (pair / n_ci) * 16, ci0 = (pair % n_ci) * 16;
    const int m = lane & 15, kg = lane >> 4;
    floatx4 acc[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) acc[k] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    const int r0 = blockIdx.x * rows_per_block;
    const int r1 = min(N * H, r0 + rows_per_block);
    for (int row = r0 + wave; row < r1; row += 4) {
        const int n = row / H, y = row % H;
        for (int x0 = 0; x0 < W; x0 += 32) {
            const int xl = x0 + 8 * kg;  // this lane's first pixel of the chunk
            Pack8 a;
#pragma unroll
            for (int j = 0; j < 8; ++j) {
                const int x = xl + j;
                a.h[j] = x < W ? g[(((size_t)n * H + y) * W + x) * cout + co0 + m] : (u16)0;
            }
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy) {
                // X at pixels xl - 1 .. xl + 8 of row y + dy (zero outside the
                // image; pixels past W meet a zero g value)
                u16 win[10];
#pragma unroll
                for (int i = 0; i < 10; ++i) win[i] = gather1(in, n, y + dy, xl - 1 + i, H, W, ci0 + m);
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) {
                    Pack8 b;
#pragma unroll
                    for (int j = 0; j < 8; ++j) b.h[j] = win[j + dx + 1];
                    acc[(dy + 1) * 3 + dx + 1] =
                        __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc[(dy + 1) * 3 + dx + 1], 0, 0, 0);
                }
            }
        }
    }
    // D[4*kg + r][m]: row = co offset, col = ci offset
    const int ntiles = 9 * n_co * n_ci;
#pragma unroll
    for (int tap = 0; tap < 9; ++tap) {
#pragma unroll
        for (int r = 0; r < 4; ++r) red[wave][(4 * kg + r) * 16 + m] = acc[tap][r];
        __syncthreads();
        const int t = threadIdx.x;
        const float s = (red[0][t] + red[1][t]) + (red[2][t] + red[3][t]);
        partial[((size_t)blockIdx.x * ntiles + tap * (n_co * n_ci) + pair) * 256 + t] = s;
        __syncthreads();  // red is rewritten for the next tap
    }
}

// dW = sum over row blocks; scatter tiles into [Cout][9][Cin].  Eight
// interleaved partial sums (block b into sum b mod 8) keep eight loads in
// flight; they are combined in a fixed order, so the result is deterministic.
__global__ void k_dw_reduce(const float* __restrict__ partial, int nblk, int ntiles, int cout, int cin,
                            float* __restrict__ dw) {
    const int idx = blockIdx.x * 256 + threadIdx.x;
    if (idx >= ntiles * 256) return;
    const int tile = idx / 256, e = idx % 256;
    const size_t stride = (size_t)ntiles * 256;
    const float* p = partial + (size_t)tile * 256 + e;
    float q[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    int b = 0;
    for (; b + 8 <= nblk; b += 8) {
#pragma unroll
        for (int k = 0; k < 8; ++k) q[k] += p[(size_t)(b + k) * stride];
    }
    for (int k = 0; b < nblk; ++b, ++k) q[k] += p[(size_t)b * stride];
    const float s = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
    const int n_ci = cin / 16, n_co = cout / 16;
    const int tap = tile / (n_co * n_ci), rem = tile % (n_co * n_ci);
    const int co = (rem / n_ci) * 16 + e / 16, ci = (rem % n_ci) * 16 + e % 16;
    dw[((size_t)co * 9 + tap) * cin + ci] = s;
}

// db[co] = sum_p g[p][co] (per-block partials, then an ordered sum).  A
// thread sums every groups-th pixel of the block's range into eight
// interleaved partial sums (eight loads in flight), combined in a fixed order.
__global__ void k_db_partial(const u16* __restrict__ g, int cout, size_t npix, size_t per_block,
                             float* __restrict__ partial) {
    __shared__ float red[256];
    const int co = threadIdx.x % cout, lane_group = threadIdx.x / cout, groups = 256 / cout;
    float q[8] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
    const size_t p0 = blockIdx.x * per_block, p1 = min(npix, p0 + per_block);
    if (lane_group < groups) {
        size_t p = p0 + lane_group;
        for (; p + 7 * (size_t)groups < p1; p += 8 * (size_t)groups) {
#pragma unroll
            for (int k = 0; k < 8; ++k) q[k] += bf2f(g[(p + (size_t)k * groups) * cout + co]);
        }
        for (int k = 0; p < p1; p += groups, ++k) q[k] += bf2f(g[p * cout + co]);
    }
    const float s = ((q[0] + q[1]) + (q[2] + q[3])) + ((q[4] + q[5]) + (q[6] + q[7]));
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x < cout) {
        float t = 0.0f;
        for (int k = 0; k < groups; ++k) t += red[k * cout + threadIdx.x];
        partial[(size_t)blockIdx.x * cout + threadIdx.x] = t;
    }
}
// one block per channel: strided partial sums, then a fixed-shape tree (the
// order of additions does not depend on timing, so the result is
// deterministic)
__global__ void k_db_reduce(const float* __restrict__ partial, int nblk, int cout, float* __restrict__ db) {
    __shared__ float red[256];
    const int co = blockIdx.x;
    float s = 0.0f;
    for (int b = threadIdx.x; b < nblk; b += 256) s += partial[(size_t)b * cout + co];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) db[co] = red[0];
}

// ------------------------------------------------------- elementwise ---
// g = bf16(grad * (act > 0))  (act null: no mask)
__global__ void k_mask_bf16(const float* __restrict__ grad, const u16* __restrict__ act, size_t n,
                            u16* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    float v = grad[i];
    if (act && !(bf2f(act[i]) > 0.0f)) v = 0.0f;
    out[i] = f2bf(v);
}

// 2x2 average of BF16 activations, eight channels (one 16-byte load of
// each of the four pixels) per thread; C is a multiple of 8
__global__ void k_avgpool(const u16* __restrict__ in, int N, int H, int W, int C, u16* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;  // group of 8 output channels
    const int h2 = H / 2, w2 = W / 2, cg = C / 8;
    const size_t total = (size_t)N * h2 * w2 * cg;
    if (i >= total) return;
    const int g = (int)(i % cg);
    size_t q = i / cg;
    const int x = (int)(q % w2);
    q /= w2;
    const int y = (int)(q % h2);
    const int n = (int)(q / h2);
    auto at = [&](int yy, int xx) {
        Pack8 p;
        p.u = *reinterpret_cast<const uint4*>(in + (((size_t)n * H + yy) * W + xx) * C + 8 * g);
        return p;
    };
    const Pack8 a = at(2 * y, 2 * x), b = at(2 * y, 2 * x + 1), c = at(2 * y + 1, 2 * x), d = at(2 * y + 1, 2 * x + 1);
    Pack8 o;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
        const float s = (bf2f(a.h[k]) + bf2f(b.h[k])) + (bf2f(c.h[k]) + bf2f(d.h[k]));
        o.h[k] = f2bf(0.25f * s);
    }
    *reinterpret_cast<uint4*>(out + (((size_t)n * h2 + y) * w2 + x) * C + 8 * g) = o.u;
}

// gin[p] += 0.25 * gout[parent(p)]
__global__ void k_avgpool_bwd(const float* __restrict__ gout, int N, int H, int W, int C, float* __restrict__ gin) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t total = (size_t)N * H * W * C;
    if (i >= total) return;
    const int c = i % C;
    size_t q = i / C;
    const int x = q % W;
    q /= W;
    const int y = q % H;
    const int n = q / H;
    gin[i] += 0.25f * gout[(((size_t)n * (H / 2) + y / 2) * (W / 2) + x / 2) * C + c];
}

// split the gradient of a concatenated input [A | up(B)]: A part added to
// ga, B part summed over each 2x2 block into gb (written, or added)
__global__ void k_split_grad(const float* __restrict__ d, int N, int H, int W, int ca, int cb, float* __restrict__ ga,
                             float* __restrict__ gb, int gb_accum) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    const size_t na = (size_t)N * H * W * ca;
    const size_t nb = (size_t)N * (H / 2) * (W / 2) * cb;
    const int C = ca + cb;
    if (i < na) {
        const size_t p = i / ca;
        ga[i] += d[p * C + i % ca];
    } else if (i < na + nb) {
        const size_t j = i - na;
        const int c = j % cb;
        size_t q = j / cb;
        const int x = q % (W / 2);
        q /= (W / 2);
        const int y = q % (H / 2);
        const int n = q / (H / 2);
        auto at = [&](int yy, int xx) { return d[(((size_t)n * H + yy) * W + xx) * C + ca + c]; };
        const float s = (at(2 * y, 2 * x) + at(2 * y, 2 * x + 1)) + (at(2 * y + 1, 2 * x) + at(2 * y + 1, 2 * x + 1));
        gb[j] = gb_accum ? gb[j] + s : s;
    }
}

// network input: 16 channels = log(1 + radiance mean), albedo mean, normal
// mean, zeros; also keeps log(1 + radiance) for the residual output
__global__ void k_prep(const float* __restrict__ rgb, const float* __restrict__ alb, const float* __restrict__ nrm,
                       size_t npix, float inv_spp, u16* __restrict__ x, float* __restrict__ noisy_log) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
    Pack8 lo, hi;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        const float l = log1pf(fmaxf(rgb[3 * p + c] * inv_spp, 0.0f));
        noisy_log[3 * p + c] = l;
        lo.h[c] = f2bf(l);
        lo.h[3 + c] = f2bf(alb ? alb[3 * p + c] * inv_spp : 0.0f);
    }
    lo.h[6] = f2bf(nrm ? nrm[3 * p] * inv_spp : 0.0f);
    lo.h[7] = f2bf(nrm ? nrm[3 * p + 1] * inv_spp : 0.0f);
    hi.u = make_uint4(0, 0, 0, 0);
    hi.h[0] = f2bf(nrm ? nrm[3 * p + 2] * inv_spp : 0.0f);
    reinterpret_cast<uint4*>(x + 16 * p)[0] = lo.u;
    reinterpret_cast<uint4*>(x + 16 * p)[1] = hi.u;
}

// denoised linear radiance = expm1(log(1 + noisy) + out[0..2])
__global__ void k_output(const float* __restrict__ out16, const float* __restrict__ noisy_log, size_t npix,
                         float* __restrict__ img) {
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (p >= npix) return;
#pragma unroll
    for (int c = 0; c < 3; ++c) img[3 * p + c] = fmaxf(expm1f(noisy_log[3 * p + c] + out16[16 * p + c]), 0.0f);
}

// kind 0: loss = mean over pixels and channels of (z - log(1 + target))^2,
//   z = log(1 + noisy) + out, the prediction in log space;
// kind 1: loss = mean of (D - T)^2 / (D^2 + 0.01), D = expm1(z), with the
//   denominator held constant in the gradient: the relative MSE of the
//   error metric, unbiased under noisy targets (Lehtinen et al. 2018,
//   Noise2Noise); dD/dz = D + 1.
// g16 = dloss/dout (FP32, 16 channels), per-block partial loss
__global__ void k_loss(const float* __restrict__ out16, const float* __restrict__ noisy_log,
                       const float* __restrict__ target, size_t npix, int kind, float* __restrict__ g16f,
                       float* __restrict__ loss_partial) {
    __shared__ float red[256];
    const size_t p = (size_t)blockIdx.x * 256 + threadIdx.x;
    float l = 0.0f;
    if (p < npix) {
        const float scale = 2.0f / (float)(3 * npix);
#pragma unroll
        for (int c = 0; c < 16; ++c) {
            float gv = 0.0f;
            if (c < 3) {
                const float z = noisy_log[3 * p + c] + out16[16 * p + c];
                const float t = fmaxf(target[3 * p + c], 0.0f);
                if (kind == 0) {
                    const float d = z - log1pf(t);
                    l += d * d;
                    gv = scale * d;
                } else {
                    const float D = expm1f(z);
                    const float e = D - t;
                    const float inv = 1.0f / (D * D + 0.01f);
                    l += e * e * inv;
                    gv = scale * e * inv * (D + 1.0f);
                }
            }
            g16f[16 * p + c] = gv;
        }
    }
    red[threadIdx.x] = l;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0];
}

// Adam on the FP32 master parameters
__global__ void k_adam(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m, float* __restrict__ v,
                       size_t n, float lr, float b1, float b2, float eps, float c1, float c2) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.0f - b1) * gi;
    const float vi = b2 * v[i] + (1.0f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    w[i] -= lr * (mi * c1) / (sqrtf(vi * c2) + eps);
}

// BF16 operand copies of one layer: forward [Cout][Kpad] (k = tap*Cin+ci)
// and data-gradient [Cin][Kpad'] (k' = tap'*Cout+co, tap' = 8 - tap)
__global__ void k_pack_weights(const float* __restrict__ w, int cout, int cin, int kpad, int kpad_t,
                               u16* __restrict__ wk, u16* __restrict__ wt) {
    const int i = blockIdx.x * 256 + threadIdx.x;
    const int nf = cout * kpad, nt = cin * kpad_t;
    if (i < nf) {
        const int co = i / kpad, k = i % kpad;
        wk[i] = k < 9 * cin ? f2bf(w[(size_t)co * 9 * cin + k]) : (u16)0;
    } else if (i < nf + nt) {
        const int j = i - nf;
        const int ci = j / kpad_t, k = j % kpad_t;
        u16 v = 0;
        if (k < 9 * cout) {
            const int tap = k / cout, co = k % cout;
            v = f2bf(w[((size_t)co * 9 + (8 - tap)) * cin + ci]);
        }
        wt[j] = v;
    }
}

__global__ void k_f32_to_bf16(const float* __restrict__ in, size_t n, u16* __restrict__ out) {
    const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) out[i] = f2bf(in[i]);
}

// 16x16x32 MFMA on one tile: the lane layout self-test
__global__ void k_mfma_selftest(const u16* A, const u16* B, float* D) {
    const int l = threadIdx.x;
    Pack8 a, b;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        a.h[j] = A[(l % 16) * 32 + 8 * (l / 16) + j];  // A is 16 x 32 row-major
        b.h[j] = B[(8 * (l / 16) + j) * 16 + l % 16];  // B is 32 x 16 row-major
    }
    floatx4 acc = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a.v, b.v, acc, 0, 0, 0);
#pragma unroll
    for (int r = 0; r < 4; ++r) D[(4 * (l / 16) + r) * 16 + l % 16] = acc[r];
}

inline unsigned nblocks(size_t n) { return (unsigned)((n + 255) / 256); }

}  // namespace

// ---------------------------------------------------------------- host ---
Status conv3x3_forward(const ConvIn& in, const uint16_t* wk_all, int kpad, const float* bias_all, void* out,
                       bool out_f32, bool accum, bool relu, int N, int H, int W, int cout, hipStream_t s) {
    if ((in.ca % 16) || (in.cb % 16) || (cout % 16) || cout < 16 || (in.ca + in.cb) < 16)
        return Status{AURORA_ERR_INVALID, "conv3x3: channel counts must be multiples of 16"};
    // the LDS-tiled kernel where its tiles fit 80 KB (of the 160 KB of LDS per
    // CU; every layer of the network: the largest, the level-2 decoder input
    // of 64 + 64 channels, takes 76.6 KB) and each part has at most 64
    // channels, else the gather form
    // (its staging loads cover up to 64 channels per part in 256 threads)
    const bool small = (size_t)N * H * W * std::max(in.ca, in.cb) < (1ull << 31);  // 32-bit staging offsets
    const int tr = lds_tile_bytes(in.ca, in.cb) <= 81920 && in.ca <= 64 && in.cb <= 64 && small ? 4 : 0;
    const unsigned grid = (unsigned)(((size_t)N * H * ((W + 63) / 64) + 3) / 4);  // 4 row segments per block
    const unsigned grid_lds = tr ? (unsigned)((size_t)N * ((H + 3) / 4) * ((W + 63) / 64)) : 0u;
    const size_t lds = tr ? (size_t)lds_tile_bytes(in.ca, in.cb) : 0;
    // Output channels: waves of up to 64 channels (4 MFMA tiles, every A
    // fragment feeding four MFMAs), one launch per 64.  A layer of fewer than
    // 128 pixel blocks in the LDS form (the coarsest level: 64 blocks at
    // 1024^2) instead takes one 16-channel group per grid row (y), so its one
    // launch has cout/16 times the blocks: 22.8 -> 16.8 us per layer.  At 256
    // blocks the wide waves measured faster (21 vs 29 us), and in the gather
    // form far faster (44 vs 81 us: the L1 gathers are no longer shared by
    // four tiles).
    const unsigned pix_blocks = tr ? grid_lds : grid;
    const bool narrow = tr && pix_blocks < 128 && cout > 16;
    const int step = narrow ? cout : 64;
    for (int co_base = 0; co_base < cout; co_base += step) {
    const int nt = narrow ? 1 : std::min(64, cout - co_base) / 16;
    const unsigned groups = narrow ? (unsigned)(cout / 16) : 1u;
    const uint16_t* wk = wk_all + (size_t)co_base * kpad;
    const float* bias = bias_all ? bias_all + co_base : nullptr;
#define AUR_CONV(NTv, R, F, A)                                                                                  \
    if (tr && in.cb)                                                                                           \
        k_conv3x3_lds<NTv, R, F, A, true><<<dim3(grid_lds, groups), 64 * tr, lds, s>>>(in, wk, kpad, bias, out, N, \
                                                                                      H, W, cout, co_base);     \
    else if (tr)                                                                                               \
        k_conv3x3_lds<NTv, R, F, A, false><<<dim3(grid_lds, groups), 64 * tr, lds, s>>>(in, wk, kpad, bias, out,   \
                                                                                       N, H, W, cout, co_base); \
    else                                                                                                       \
        k_conv3x3<NTv, R, F, A><<<dim3(grid, groups), 256, 0, s>>>(in, wk, kpad, bias, out, N, H, W, cout, co_base)
#define AUR_CONV_NT(R, F, A)          \
    switch (nt) {                     \
        case 1: AUR_CONV(1, R, F, A); break; \
        case 2: AUR_CONV(2, R, F, A); break; \
        case 3: AUR_CONV(3, R, F, A); break; \
        default: AUR_CONV(4, R, F, A); break; \
    }
    if (!out_f32) {
        if (relu) { AUR_CONV_NT(true, false, false) } else { AUR_CONV_NT(false, false, false) }
    } else if (accum) {
        AUR_CONV_NT(false, true, true)
    } else if (relu) {
        AUR_CONV_NT(true, true, false)
    } else {
        AUR_CONV_NT(false, true, false)
    }
#undef AUR_CONV_NT
#undef AUR_CONV
    }
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

Status conv3x3_weight_grad(const ConvIn& in, const uint16_t* g, int cout, int N, int H, int W, float* dw, float* db,
                           DevBuf& scratch, hipStream_t s) {
    const int cin = in.ca + in.cb;
    const int ntiles = 9 * (cout / 16) * (cin / 16);
    const int rows = N * H;
    const int rows_per_block = 4;  // one row per wave
    const int nblk = (rows + rows_per_block - 1) / rows_per_block;
    const size_t npix = (size_t)N * H * W;
    // bias gradient: 512 pixels per block (a 128x128x8 batch is 256 blocks)
    const size_t per_block = 512;
    const int nbb = (int)((npix + per_block - 1) / per_block);
    AUR_TRY(scratch.alloc(((size_t)nblk * ntiles * 256 + (size_t)nbb * cout) * 4));
    float* part = scratch.as<float>();
    float* dbpart = part + (size_t)nblk * ntiles * 256;
    k_conv3x3_dw<<<dim3(nblk, ntiles / 9), 256, 0, s>>>(in, g, cout, N, H, W, rows_per_block, part);
    k_dw_reduce<<<nblocks((size_t)ntiles * 256), 256, 0, s>>>(part, nblk, ntiles, cout, cin, dw);
    k_db_partial<<<nbb, 256, 0, s>>>(g, cout, npix, per_block, dbpart);
    k_db_reduce<<<cout, 256, 0, s>>>(dbpart, nbb, cout, db);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

Status mask_to_bf16(const float* grad, const uint16_t* act, size_t n, uint16_t* out, hipStream_t s) {
    k_mask_bf16<<<nblocks(n), 256, 0, s>>>(grad, act, n, out);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status avgpool2(const uint16_t* in, int N, int H, int W, int C, uint16_t* out, hipStream_t s) {
    if (C % 8) return Status{AURORA_ERR_INVALID, "avgpool: channel count must be a multiple of 8"};
    k_avgpool<<<nblocks((size_t)N * (H / 2) * (W / 2) * (C / 8)), 256, 0, s>>>(in, N, H, W, C, out);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status avgpool2_backward(const float* gout, int N, int H, int W, int C, float* gin, hipStream_t s) {
    k_avgpool_bwd<<<nblocks((size_t)N * H * W * C), 256, 0, s>>>(gout, N, H, W, C, gin);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status split_concat_grad(const float* d, int N, int H, int W, int ca, int cb, float* ga, float* gb, bool gb_accum,
                         hipStream_t s) {
    const size_t total = (size_t)N * H * W * ca + (size_t)N * (H / 2) * (W / 2) * cb;
    k_split_grad<<<nblocks(total), 256, 0, s>>>(d, N, H, W, ca, cb, ga, gb, gb_accum ? 1 : 0);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status prep_input(const float* rgb, const float* alb, const float* nrm, size_t npix, float inv_spp, uint16_t* x,
                  float* noisy_log, hipStream_t s) {
    k_prep<<<nblocks(npix), 256, 0, s>>>(rgb, alb, nrm, npix, inv_spp, x, noisy_log);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status output_image(const float* out16, const float* noisy_log, size_t npix, float* img, hipStream_t s) {
    k_output<<<nblocks(npix), 256, 0, s>>>(out16, noisy_log, npix, img);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status loss_grad(const float* out16, const float* noisy_log, const float* target, size_t npix, int kind,
                 float* g16f, float* loss_partial, hipStream_t s) {
    if (kind != 0 && kind != 1) return Status{AURORA_ERR_INVALID, "unknown loss kind"};
    k_loss<<<nblocks(npix), 256, 0, s>>>(out16, noisy_log, target, npix, kind, g16f, loss_partial);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status adam_step(float* w, const float* g, float* m, float* v, size_t n, float lr, float b1, float b2, float eps,
                 int t, hipStream_t s) {
    const float c1 = 1.0f / (1.0f - powf(b1, (float)t)), c2 = 1.0f / (1.0f - powf(b2, (float)t));
    k_adam<<<nblocks(n), 256, 0, s>>>(w, g, m, v, n, lr, b1, b2, eps, c1, c2);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status pack_weights(const float* w, int cout, int cin, int kpad, int kpad_t, uint16_t* wk, uint16_t* wt,
                    hipStream_t s) {
    k_pack_weights<<<nblocks((size_t)cout * kpad + (size_t)cin * kpad_t), 256, 0, s>>>(w, cout, cin, kpad, kpad_t,
                                                                                       wk, wt);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status f32_to_bf16(const float* in, size_t n, uint16_t* out, hipStream_t s) {
    k_f32_to_bf16<<<nblocks(n), 256, 0, s>>>(in, n, out);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}
Status mfma_selftest(const uint16_t* A, const uint16_t* B, float* D, hipStream_t s) {
    k_mfma_selftest<<<1, 64, 0, s>>>(A, B, D);
    AUR_HIP(hipGetLastError());
    return Status::ok();
}

}  // namespace aur
