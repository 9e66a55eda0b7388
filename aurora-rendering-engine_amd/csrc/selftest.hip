// selftest.hip — device self-test of the fast correctly rounded reciprocal
// (dev_common.hpp rcp_fast): every float32 bit pattern of a range through
// rcp_fast and through IEEE division, compared bit for bit.  The path kernels
// use rcp_fast only for 2^-126 <= |x| < 2^126 (rcp_rn); the test over those
// patterns is what licenses it.
#include <string>

#include "internal.hpp"

using namespace aur;

struct aurora_ctx;
// the context's stream, device and error slot (capi.hip)
hipStream_t ctx_stream(aurora_ctx* c);
int ctx_device(aurora_ctx* c);
int ctx_fail(aurora_ctx* c, int code, const std::string& msg);

namespace {

constexpr int kBlock = 256;
constexpr int kGrid = 4096;
constexpr int kFirst = 16;  // mismatching patterns recorded

// counts[0]: mismatches; first[0..kFirst): the first ones found (any order)
__global__ __launch_bounds__(kBlock) void k_rcp_selftest(uint32_t lo, uint32_t hi, unsigned long long* count,
                                                         uint32_t* first) {
    const uint64_t stride = (uint64_t)gridDim.x * kBlock;
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x + lo; i < hi; i += stride) {
        const float x = __uint_as_float((uint32_t)i);
        const float ieee = 1.0f / x;
        const float fast = rcp_fast(x);
        if (__float_as_uint(fast) != __float_as_uint(ieee)) {
            const unsigned long long k = atomicAdd(count, 1ull);
            if (k < (unsigned long long)kFirst) first[k] = (uint32_t)i;
        }
    }
}

}  // namespace

extern "C" {

int aurora_selftest_rcp(aurora_ctx* ctx, uint32_t lo, uint32_t hi, uint64_t* mismatches, uint32_t* first) {
    if (!ctx || !mismatches) return ctx_fail(ctx, AURORA_ERR_INVALID, "NULL argument");
    if (hi < lo) return ctx_fail(ctx, AURORA_ERR_INVALID, "selftest_rcp: hi < lo");
    hipStream_t s = ctx_stream(ctx);
    auto hs = [](hipError_t e) { return e == hipSuccess ? Status::ok() : Status{AURORA_ERR_RUNTIME, hipGetErrorString(e)}; };
    Status st = hs(hipSetDevice(ctx_device(ctx)));
    DevBuf buf;
    if (!st.bad()) st = buf.alloc(8 + 4 * kFirst);
    unsigned long long* d_count = reinterpret_cast<unsigned long long*>(buf.p);
    uint32_t* d_first = reinterpret_cast<uint32_t*>(static_cast<char*>(buf.p) + 8);
    if (!st.bad()) st = hs(hipMemsetAsync(buf.p, 0, 8 + 4 * kFirst, s));
    if (!st.bad() && hi > lo) {
        k_rcp_selftest<<<kGrid, kBlock, 0, s>>>(lo, hi, d_count, d_first);
        st = hs(hipGetLastError());
    }
    unsigned long long count = 0;
    uint32_t host_first[kFirst] = {};
    if (!st.bad()) st = hs(hipMemcpyAsync(&count, d_count, 8, hipMemcpyDeviceToHost, s));
    if (!st.bad()) st = hs(hipMemcpyAsync(host_first, d_first, 4 * kFirst, hipMemcpyDeviceToHost, s));
    if (!st.bad()) st = hs(hipStreamSynchronize(s));
    (void)hipStreamSynchronize(s);
    buf.release();
    if (st.bad()) return ctx_fail(ctx, st.code, st.msg);
    *mismatches = count;
    if (first)
        for (int k = 0; k < kFirst; ++k) first[k] = host_first[k];
    return AURORA_OK;
}

}  // extern "C"
