// denoise.hip — C ABI of the learned denoiser (include/aurora.h): create,
// parameters in and out, inference on a rendered image, loss + gradients,
// and one Adam training step; plus the MFMA lane-layout self-test.
#include <vector>

#include "unet.hpp"

using namespace aur;

struct aurora_ctx;
// the context's stream and error slot (capi.hip)
hipStream_t ctx_stream(aurora_ctx* c);
int ctx_device(aurora_ctx* c);
int ctx_fail(aurora_ctx* c, int code, const std::string& msg);

struct aurora_denoiser {
    aurora_ctx* ctx = nullptr;
    Unet* net = nullptr;
};

namespace {
int fail(aurora_denoiser* d, const Status& s) { return ctx_fail(d ? d->ctx : nullptr, s.code, s.msg); }
#define DCK(d, expr)                          \
    do {                                      \
        Status st_ = (expr);                  \
        if (st_.bad()) return fail(d, st_);   \
    } while (0)
Status set_device(aurora_denoiser* d) {
    AUR_HIP(hipSetDevice(ctx_device(d->ctx)));
    return Status::ok();
}
}  // namespace

extern "C" {

int aurora_denoiser_create(aurora_ctx* ctx, uint32_t seed, aurora_denoiser** out) {
    if (!ctx || !out) return ctx_fail(ctx, AURORA_ERR_INVALID, "ctx or out is NULL");
    *out = nullptr;
    auto* d = new aurora_denoiser();
    d->ctx = ctx;
    d->net = new Unet(ctx_stream(ctx));
    Status s = set_device(d);
    if (!s.bad()) s = d->net->init(seed);
    if (s.bad()) {
        d->net->release();
        delete d->net;
        delete d;
        return ctx_fail(ctx, s.code, s.msg);
    }
    *out = d;
    return AURORA_OK;
}

void aurora_denoiser_destroy(aurora_denoiser* d) {
    if (!d) return;
    (void)hipStreamSynchronize(ctx_stream(d->ctx));
    d->net->release();
    delete d->net;
    delete d;
}

uint64_t aurora_denoiser_param_count(const aurora_denoiser* d) { return d ? d->net->param_count() : 0; }

int aurora_denoiser_get_params(aurora_denoiser* d, float* host, uint64_t n) {
    if (!d || !host || n != d->net->param_count()) return ctx_fail(d ? d->ctx : nullptr, AURORA_ERR_INVALID, "bad parameter buffer");
    DCK(d, set_device(d));
    DCK(d, d->net->get_params(host));
    return AURORA_OK;
}

int aurora_denoiser_set_params(aurora_denoiser* d, const float* host, uint64_t n) {
    if (!d || !host || n != d->net->param_count()) return ctx_fail(d ? d->ctx : nullptr, AURORA_ERR_INVALID, "bad parameter buffer");
    DCK(d, set_device(d));
    DCK(d, d->net->set_params(host));
    return AURORA_OK;
}

int aurora_denoise(aurora_denoiser* d, const float* rgb, const float* alb, const float* nrm, uint32_t spp,
                   uint32_t width, uint32_t height, float* out) {
    if (!d) return ctx_fail(nullptr, AURORA_ERR_INVALID, "denoiser is NULL");
    if (!rgb || !alb || !nrm || !out || spp == 0)
        return ctx_fail(d->ctx, AURORA_ERR_INVALID, "NULL buffer or spp == 0");
    DCK(d, set_device(d));
    DCK(d, d->net->forward(rgb, alb, nrm, 1.0f / (float)spp, 1, (int)height, (int)width));
    DCK(d, d->net->write_output(out));
    return AURORA_OK;
}

int aurora_denoiser_set_loss(aurora_denoiser* d, int kind) {
    if (!d) return ctx_fail(nullptr, AURORA_ERR_INVALID, "denoiser is NULL");
    if (kind != AURORA_LOSS_LOG_L2 && kind != AURORA_LOSS_REL_L2)
        return ctx_fail(d->ctx, AURORA_ERR_INVALID, "loss kind must be AURORA_LOSS_LOG_L2 or AURORA_LOSS_REL_L2");
    d->net->set_loss(kind);
    return AURORA_OK;
}

int aurora_denoiser_loss_grad(aurora_denoiser* d, const float* rgb, const float* alb, const float* nrm,
                              const float* target, uint32_t batch, uint32_t spp, uint32_t width, uint32_t height,
                              float* loss, float* host_grads) {
    if (!d) return ctx_fail(nullptr, AURORA_ERR_INVALID, "denoiser is NULL");
    if (!rgb || !alb || !nrm || !target || spp == 0 || batch == 0)
        return ctx_fail(d->ctx, AURORA_ERR_INVALID, "NULL buffer, spp == 0 or batch == 0");
    DCK(d, set_device(d));
    DCK(d, d->net->forward(rgb, alb, nrm, 1.0f / (float)spp, (int)batch, (int)height, (int)width));
    DCK(d, d->net->backward(target, loss));
    if (host_grads) DCK(d, d->net->grads_to_host(host_grads));
    return AURORA_OK;
}

int aurora_denoiser_train_step(aurora_denoiser* d, const float* rgb, const float* alb, const float* nrm,
                               const float* target, uint32_t batch, uint32_t spp, uint32_t width, uint32_t height,
                               float lr, float* loss) {
    int rc = aurora_denoiser_loss_grad(d, rgb, alb, nrm, target, batch, spp, width, height, loss, nullptr);
    if (rc != AURORA_OK) return rc;
    DCK(d, d->net->adam(lr));
    return AURORA_OK;
}

int aurora_selftest_mfma(aurora_ctx* ctx, const uint16_t* A, const uint16_t* B, float* D) {
    if (!ctx || !A || !B || !D) return ctx_fail(ctx, AURORA_ERR_INVALID, "NULL argument");
    hipStream_t s = ctx_stream(ctx);
    DevBuf a, b, dd;
    Status st = a.alloc(16 * 32 * 2);
    if (!st.bad()) st = b.alloc(32 * 16 * 2);
    if (!st.bad()) st = dd.alloc(16 * 16 * 4);
    auto hs = [](hipError_t e) { return e == hipSuccess ? Status::ok() : Status{AURORA_ERR_RUNTIME, hipGetErrorString(e)}; };
    if (!st.bad()) st = hs(hipMemcpyAsync(a.p, A, 16 * 32 * 2, hipMemcpyHostToDevice, s));
    if (!st.bad()) st = hs(hipMemcpyAsync(b.p, B, 32 * 16 * 2, hipMemcpyHostToDevice, s));
    if (!st.bad()) st = mfma_selftest(a.as<uint16_t>(), b.as<uint16_t>(), dd.as<float>(), s);
    if (!st.bad()) st = hs(hipMemcpyAsync(D, dd.p, 16 * 16 * 4, hipMemcpyDeviceToHost, s));
    if (!st.bad()) st = hs(hipStreamSynchronize(s));
    (void)hipStreamSynchronize(s);
    a.release();
    b.release();
    dd.release();
    if (st.bad()) return ctx_fail(ctx, st.code, st.msg);
    return AURORA_OK;
}

}  // extern "C"
