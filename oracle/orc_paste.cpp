// oracle/orc_paste.cpp — CPU restatement of are::Texture::paste
// (src/texture.cpp:85-360), the reference's homography warp.  TEST
// INFRASTRUCTURE ONLY (loaded by tests/ to check libaurora's paste kernel).
//
// Steps, each as the reference states them (texture.cpp line ranges):
//   early returns: empty source :93-95, quad bounding box outside the
//     destination :123-126, singular system :252-255 or homography :290-292;
//   the quad LT, RT, RB, LB and its bounding box clipped to the image
//     :111-133;
//   the homography: two rows per source corner (0,0), (w-1,0), (0,h-1),
//     (w-1,h-1) -> LT, RT, LB, RB :145-206, solved by Gauss-Jordan
//     elimination with partial pivoting and pivots below 1e-12 rejected
//     :208-248; its inverse from the cofactors :258-292;
//   per pixel centre (x+0.5, y+0.5): ray-casting inside test :136-145,
//     inverse mapping with |denominator| < 1e-12 skipped :339-346, source
//     bounds test :348-351, bilinear sample with clamped neighbours
//     :294-316.
// Source and destination texels are float; the arithmetic is double (the
// reference's Color3), rounded once to float when stored.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <utility>

#include "oracle.h"

namespace {

const double kEps = 1e-12;

bool gauss_jordan8(double M[8][9], double h[8]) {
    for (int col = 0; col < 8; ++col) {
        int piv = col;
        double best = std::fabs(M[col][col]);
        for (int r = col + 1; r < 8; ++r) {
            const double a = std::fabs(M[r][col]);
            if (a > best) { best = a; piv = r; }
        }
        if (best < kEps) return false;
        if (piv != col)
            for (int c = col; c < 9; ++c) std::swap(M[col][c], M[piv][c]);
        const double d = M[col][col];
        for (int c = col; c < 9; ++c) M[col][c] /= d;
        for (int r = 0; r < 8; ++r) {
            if (r == col) continue;
            const double f = M[r][col];
            if (std::fabs(f) < kEps) continue;
            for (int c = col; c < 9; ++c) M[r][c] -= f * M[col][c];
        }
    }
    for (int i = 0; i < 8; ++i) h[i] = M[i][8];
    return true;
}

bool inverse3(const double H[3][3], double R[3][3]) {
    const double a = H[0][0], b = H[0][1], c = H[0][2];
    const double d = H[1][0], e = H[1][1], f = H[1][2];
    const double g = H[2][0], h = H[2][1], i = H[2][2];
    const double k00 = (e * i - f * h), k01 = -(d * i - f * g), k02 = (d * h - e * g);
    const double k10 = -(b * i - c * h), k11 = (a * i - c * g), k12 = -(a * h - b * g);
    const double k20 = (b * f - c * e), k21 = -(a * f - c * d), k22 = (a * e - b * d);
    const double det = a * k00 + b * k01 + c * k02;
    if (std::fabs(det) < kEps) return false;
    const double s = 1.0 / det;
    R[0][0] = k00 * s; R[0][1] = k10 * s; R[0][2] = k20 * s;
    R[1][0] = k01 * s; R[1][1] = k11 * s; R[1][2] = k21 * s;
    R[2][0] = k02 * s; R[2][1] = k12 * s; R[2][2] = k22 * s;
    return true;
}

}  // namespace

extern "C" int orc_paste(const float* src, int sw, int sh, float* dst, int dw, int dh, const int32_t* corners) {
    if (sw <= 0 || sh <= 0) return 0;
    const double px[4] = {double(corners[0]), double(corners[2]), double(corners[6]), double(corners[4])};
    const double py[4] = {double(corners[1]), double(corners[3]), double(corners[7]), double(corners[5])};
    const double lx = std::min({px[0], px[1], px[2], px[3]}), hx = std::max({px[0], px[1], px[2], px[3]});
    const double ly = std::min({py[0], py[1], py[2], py[3]}), hy = std::max({py[0], py[1], py[2], py[3]});
    if (hx < 0.0 || hy < 0.0 || lx > double(dw - 1) || ly > double(dh - 1)) return 0;
    const int x0 = std::max(0, int(std::floor(lx))), x1 = std::min(dw - 1, int(std::ceil(hx)));
    const int y0 = std::max(0, int(std::floor(ly))), y1 = std::min(dh - 1, int(std::ceil(hy)));
    const double W1 = double(sw) - 1.0, H1 = double(sh) - 1.0;
    const double sxs[4] = {0.0, W1, 0.0, W1}, sys[4] = {0.0, 0.0, H1, H1};
    double M[8][9] = {};
    for (int k = 0; k < 4; ++k) {
        const double x = sxs[k], y = sys[k], u = double(corners[2 * k]), v = double(corners[2 * k + 1]);
        M[2 * k][0] = x; M[2 * k][1] = y; M[2 * k][2] = 1.0;
        M[2 * k][6] = -u * x; M[2 * k][7] = -u * y; M[2 * k][8] = u;
        M[2 * k + 1][3] = x; M[2 * k + 1][4] = y; M[2 * k + 1][5] = 1.0;
        M[2 * k + 1][6] = -v * x; M[2 * k + 1][7] = -v * y; M[2 * k + 1][8] = v;
    }
    double h[8];
    if (!gauss_jordan8(M, h)) return 0;
    const double H[3][3] = {{h[0], h[1], h[2]}, {h[3], h[4], h[5]}, {h[6], h[7], 1.0}};
    double R[3][3];
    if (!inverse3(H, R)) return 0;
    int painted = 0;
    for (int y = y0; y <= y1; ++y)
        for (int x = x0; x <= x1; ++x) {
            const double X = double(x) + 0.5, Y = double(y) + 0.5;
            bool in = false;
            for (int i = 0, j = 3; i < 4; j = i++) {
                const double dy = py[j] - py[i];
                if (((py[i] > Y) != (py[j] > Y)) && (X < (px[j] - px[i]) * (Y - py[i]) / (dy == 0.0 ? 1e-30 : dy) + px[i]))
                    in = !in;
            }
            if (!in) continue;
            const double den = R[2][0] * X + R[2][1] * Y + R[2][2];
            if (std::fabs(den) < kEps) continue;
            const double u = (R[0][0] * X + R[0][1] * Y + R[0][2]) / den;
            const double v = (R[1][0] * X + R[1][1] * Y + R[1][2]) / den;
            if (u < 0.0 || v < 0.0 || u > W1 || v > H1) continue;
            const double cu = std::clamp(u, 0.0, W1), cv = std::clamp(v, 0.0, H1);
            const int ix = int(std::floor(cu)), iy = int(std::floor(cv));
            const int ix1 = std::min(ix + 1, sw - 1), iy1 = std::min(iy + 1, sh - 1);
            const double tx = cu - double(ix), ty = cv - double(iy);
            const float* a = src + 3 * (size_t(iy) * sw + ix);
            const float* b = src + 3 * (size_t(iy) * sw + ix1);
            const float* c = src + 3 * (size_t(iy1) * sw + ix);
            const float* d = src + 3 * (size_t(iy1) * sw + ix1);
            float* o = dst + 3 * (size_t(y) * dw + x);
            for (int k = 0; k < 3; ++k)
                o[k] = float(double(a[k]) * ((1.0 - tx) * (1.0 - ty)) + double(b[k]) * (tx * (1.0 - ty)) +
                             double(c[k]) * ((1.0 - tx) * ty) + double(d[k]) * (tx * ty));
            ++painted;
        }
    return painted;
}
