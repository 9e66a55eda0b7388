// unet.hip — the denoiser network on top of the kernels of unet_kernels.hip:
// a 4-level U-Net with 32/48/64/64 channels (15 3x3 convolutions, average
// pooling, nearest upsampling with skip concatenation fused into the
// decoder convolutions) predicting a residual in log(1 + radiance) from the
// noisy radiance mean, albedo and normal AOVs.  Forward, backward and Adam
// all run on the device; parameters are FP32 masters with BF16 copies for
// the MFMA operands.
#include <cmath>
#include <vector>

#include "unet.hpp"

namespace aur {

namespace {
// activation slots: channels and resolution level
constexpr int kSlotC[19] = {16, 32, 32, 32, 48, 48, 48, 64, 64, 64, 64, 64, 64, 64, 48, 48, 32, 32, 16};
constexpr int kSlotL[19] = {0, 0, 0, 1, 1, 1, 2, 2, 2, 3, 3, 3, 2, 2, 1, 1, 0, 0, 0};
constexpr int kOut = 18;
// (A slot, B slot, out slot, relu)
constexpr int kLayers[15][4] = {{0, -1, 1, 1},  {1, -1, 2, 1},   {3, -1, 4, 1},  {4, -1, 5, 1},  {6, -1, 7, 1},
                                {7, -1, 8, 1},  {9, -1, 10, 1},  {10, -1, 11, 1}, {8, 11, 12, 1}, {12, -1, 13, 1},
                                {5, 13, 14, 1}, {14, -1, 15, 1}, {2, 15, 16, 1},  {16, -1, 17, 1}, {17, -1, 18, 0}};
// pooling: (source slot, pooled slot)
constexpr int kPoolSlots[3][2] = {{2, 3}, {5, 6}, {8, 9}};
inline int pad32(int k) { return (k + 31) / 32 * 32; }
}  // namespace

Status Unet::init(uint32_t seed) {
    layers_.clear();
    size_t off = 0;
    for (int i = 0; i < 15; ++i) {
        Layer L;
        L.a = kLayers[i][0];
        L.b = kLayers[i][1];
        L.out = kLayers[i][2];
        L.ca = kSlotC[L.a];
        L.cb = L.b >= 0 ? kSlotC[L.b] : 0;
        L.cout = kSlotC[L.out];
        L.level = kSlotL[L.out];
        L.relu = kLayers[i][3] != 0;
        L.kpad = pad32(9 * (L.ca + L.cb));
        L.kpad_t = pad32(9 * L.cout);
        L.w_off = off;
        off += (size_t)L.cout * 9 * (L.ca + L.cb);
        L.b_off = off;
        off += L.cout;
        layers_.push_back(std::move(L));
    }
    n_params_ = off;
    for (int s = 0; s < kSlots; ++s) {
        slot_c_[s] = kSlotC[s];
        slot_lv_[s] = kSlotL[s];
    }
    // He-normal weights from Philox(seed, stream 3) by Box–Muller; zero
    // biases; the output layer starts at zero (denoised = noisy)
    std::vector<float> h(n_params_, 0.0f);
    for (size_t li = 0; li + 1 < layers_.size(); ++li) {
        const Layer& L = layers_[li];
        const int cin = L.ca + L.cb;
        const int fan_in = 9 * (li == 0 ? 9 : cin);
        const double sd = std::sqrt(2.0 / fan_in);
        for (int co = 0; co < L.cout; ++co)
            for (int k = 0; k < 9 * cin; ++k) {
                const size_t idx = L.w_off + (size_t)co * 9 * cin + k;
                const int ci = k % cin;
                if (li == 0 && ci >= 9) continue;  // padding input channels
                const u4 r = philox(u4{(uint32_t)idx, (uint32_t)(idx >> 32), 0u, 0u}, seed, 3u);
                const double u1 = 1.0 - (double)(r.x >> 8) / 16777216.0, u2 = (double)(r.y >> 8) / 16777216.0;
                h[idx] = (float)(sd * std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2));
            }
    }
    AUR_TRY(params_.alloc(n_params_ * 4));
    AUR_TRY(m_.alloc(n_params_ * 4));
    AUR_TRY(v_.alloc(n_params_ * 4));
    AUR_HIP(hipMemsetAsync(m_.p, 0, n_params_ * 4, s_));
    AUR_HIP(hipMemsetAsync(v_.p, 0, n_params_ * 4, s_));
    t_ = 0;
    return set_params(h.data());
}

Status Unet::repack() {
    for (Layer& L : layers_) {
        const int cin = L.ca + L.cb;
        AUR_TRY(L.wk.alloc((size_t)L.cout * L.kpad * 2));
        AUR_TRY(L.wt.alloc((size_t)cin * L.kpad_t * 2));
        AUR_TRY(pack_weights(params_.as<float>() + L.w_off, L.cout, cin, L.kpad, L.kpad_t, L.wk.as<uint16_t>(),
                             L.wt.as<uint16_t>(), s_));
    }
    return Status::ok();
}

Status Unet::get_params(float* host) {
    AUR_HIP(hipMemcpyAsync(host, params_.p, n_params_ * 4, hipMemcpyDeviceToHost, s_));
    AUR_HIP(hipStreamSynchronize(s_));
    return Status::ok();
}

Status Unet::set_params(const float* host) {
    AUR_HIP(hipMemcpyAsync(params_.p, host, n_params_ * 4, hipMemcpyHostToDevice, s_));
    AUR_TRY(repack());
    AUR_HIP(hipStreamSynchronize(s_));
    return Status::ok();
}

Status Unet::alloc(int N, int H, int W) {
    if (N == N_ && H == H_ && W == W_) return Status::ok();
    for (int s = 0; s < kSlots; ++s) {
        const size_t cnt = (size_t)N * (H >> slot_lv_[s]) * (W >> slot_lv_[s]) * slot_c_[s];
        AUR_TRY(act_[s].alloc(cnt * (s == kOut ? 4 : 2)));
    }
    AUR_TRY(noisy_log_.alloc((size_t)N * H * W * 3 * 4));
    AUR_TRY(loss_part_.alloc(((size_t)N * H * W + 255) / 256 * 4 + 16));
    N_ = N;
    H_ = H;
    W_ = W;
    return Status::ok();
}

Status Unet::forward(const float* rgb, const float* alb, const float* nrm, float inv_spp, int N, int H, int W) {
    if (N < 1 || H < 8 || W < 8 || (H % 8) || (W % 8))
        return Status{AURORA_ERR_INVALID, "denoiser: image height and width must be positive multiples of 8"};
    AUR_TRY(alloc(N, H, W));
    const size_t npix = (size_t)N * H * W;
    AUR_TRY(prep_input(rgb, alb, nrm, npix, inv_spp, act_[0].as<uint16_t>(), noisy_log_.as<float>(), s_));
    for (size_t li = 0; li < layers_.size(); ++li) {
        const Layer& L = layers_[li];
        // pooled inputs are produced just before the first layer that reads them
        for (const auto& pl : kPoolSlots)
            if (pl[1] == L.a)
                AUR_TRY(avgpool2(act_[pl[0]].as<uint16_t>(), N, H >> slot_lv_[pl[0]], W >> slot_lv_[pl[0]],
                                 slot_c_[pl[0]], act_[pl[1]].as<uint16_t>(), s_));
        ConvIn in{act_[L.a].as<uint16_t>(), L.b >= 0 ? act_[L.b].as<uint16_t>() : nullptr, L.ca, L.cb};
        const bool f32 = L.out == kOut;
        AUR_TRY(conv3x3_forward(in, L.wk.as<uint16_t>(), L.kpad, params_.as<float>() + L.b_off, act_[L.out].p, f32,
                                false, L.relu, N, H >> L.level, W >> L.level, L.cout, s_));
    }
    return Status::ok();
}

Status Unet::write_output(float* img) {
    return output_image(act_[kOut].as<float>(), noisy_log_.as<float>(), (size_t)N_ * H_ * W_, img, s_);
}

Status Unet::backward(const float* target, float* loss) {
    const int N = N_, H = H_, W = W_;
    const size_t npix = (size_t)N * H * W;
    size_t max_cat = 0, max_act = 0;
    for (int s = 0; s < kSlots; ++s) {
        const size_t cnt = (size_t)N * (H >> slot_lv_[s]) * (W >> slot_lv_[s]) * slot_c_[s];
        AUR_TRY(grad_[s].alloc(cnt * 4));
        AUR_HIP(hipMemsetAsync(grad_[s].p, 0, cnt * 4, s_));
        max_act = std::max(max_act, cnt);
    }
    for (const Layer& L : layers_)
        if (L.b >= 0) max_cat = std::max(max_cat, (size_t)N * (H >> L.level) * (W >> L.level) * (L.ca + L.cb));
    AUR_TRY(gbf_.alloc(max_act * 2));
    AUR_TRY(dcat_.alloc(max_cat * 4));
    AUR_TRY(grads_.alloc(n_params_ * 4));
    const unsigned nb = (unsigned)((npix + 255) / 256);
    AUR_TRY(loss_grad(act_[kOut].as<float>(), noisy_log_.as<float>(), target, npix, loss_kind_,
                      grad_[kOut].as<float>(), loss_part_.as<float>(), s_));
    for (int li = (int)layers_.size() - 1; li >= 0; --li) {
        const Layer& L = layers_[li];
        const int h = H >> L.level, w = W >> L.level;
        const size_t cnt = (size_t)N * h * w * L.cout;
        AUR_TRY(mask_to_bf16(grad_[L.out].as<float>(), L.relu ? act_[L.out].as<uint16_t>() : nullptr, cnt,
                             gbf_.as<uint16_t>(), s_));
        ConvIn in{act_[L.a].as<uint16_t>(), L.b >= 0 ? act_[L.b].as<uint16_t>() : nullptr, L.ca, L.cb};
        AUR_TRY(conv3x3_weight_grad(in, gbf_.as<uint16_t>(), L.cout, N, h, w, grads_.as<float>() + L.w_off,
                                    grads_.as<float>() + L.b_off, scratch_, s_));
        if (li == 0) break;  // no gradient for the network input
        ConvIn gin{gbf_.as<uint16_t>(), nullptr, L.cout, 0};
        const int cin = L.ca + L.cb;
        if (L.b < 0) {
            AUR_TRY(conv3x3_forward(gin, L.wt.as<uint16_t>(), L.kpad_t, nullptr, grad_[L.a].p, true, true, false, N, h,
                                    w, cin, s_));
        } else {
            AUR_TRY(conv3x3_forward(gin, L.wt.as<uint16_t>(), L.kpad_t, nullptr, dcat_.p, true, false, false, N, h, w,
                                    cin, s_));
            AUR_TRY(split_concat_grad(dcat_.as<float>(), N, h, w, L.ca, L.cb, grad_[L.a].as<float>(),
                                      grad_[L.b].as<float>(), true, s_));
        }
        // a pooled slot's gradient is complete once its only reader is done
        for (const auto& pl : kPoolSlots)
            if (pl[1] == L.a)
                AUR_TRY(avgpool2_backward(grad_[pl[1]].as<float>(), N, H >> slot_lv_[pl[0]], W >> slot_lv_[pl[0]],
                                          slot_c_[pl[0]], grad_[pl[0]].as<float>(), s_));
    }
    if (loss) {
        std::vector<float> part(nb);
        AUR_HIP(hipMemcpyAsync(part.data(), loss_part_.p, nb * 4, hipMemcpyDeviceToHost, s_));
        AUR_HIP(hipStreamSynchronize(s_));
        double sum = 0.0;
        for (float v : part) sum += v;
        *loss = (float)(sum / (3.0 * (double)npix));
    }
    return Status::ok();
}

Status Unet::adam(float lr) {
    ++t_;
    AUR_TRY(adam_step(params_.as<float>(), grads_.as<float>(), m_.as<float>(), v_.as<float>(), n_params_, lr, 0.9f,
                      0.999f, 1e-8f, t_, s_));
    return repack();
}

Status Unet::grads_to_host(float* host) {
    AUR_HIP(hipMemcpyAsync(host, grads_.p, n_params_ * 4, hipMemcpyDeviceToHost, s_));
    AUR_HIP(hipStreamSynchronize(s_));
    return Status::ok();
}

void Unet::release() {
    params_.release();
    grads_.release();
    m_.release();
    v_.release();
    for (auto& L : layers_) {
        L.wk.release();
        L.wt.release();
    }
    for (int s = 0; s < kSlots; ++s) {
        act_[s].release();
        grad_[s].release();
    }
    noisy_log_.release();
    gbf_.release();
    dcat_.release();
    scratch_.release();
    loss_part_.release();
}

}  // namespace aur
