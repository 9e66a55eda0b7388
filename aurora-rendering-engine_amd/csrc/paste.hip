// paste.hip — are::Texture::paste (src/texture.cpp:85-360) on gfx950: a
// source RGB image warped into a quadrilateral of a destination image by the
// homography that sends the source's corner pixels to the quad's corners.
//
// Host: the homography (an 8x8 linear system for its eight free
// coefficients, Gauss-Jordan elimination with partial pivoting) and its
// inverse (3x3 adjugate over the determinant), in double.
// Device: one thread per pixel of the quad's bounding box clipped to the
// destination: ray-casting point-in-polygon test of the pixel centre, the
// inverse mapping into the source, the source bounds test and the bilinear
// sample with clamped neighbours, in double, stored as float.
//
// Which pixels are painted is a set of discrete double-precision decisions,
// so the solve, the inverse and the per-pixel formulas keep the reference's
// operation order: every decision and every sample then equals the
// reference's bit for bit (its double result rounded once to float).
#include <algorithm>
#include <cmath>

#include "internal.hpp"

namespace aur {

namespace {

constexpr double kEps = 1e-12;  // the reference's GEOMETRY_EPSILON (include/basic/math.h:7)

struct PastePlan {
    double qx[4], qy[4];  // quad in boundary order: left-top, right-top, right-bottom, left-bottom
    double inv[9];        // destination -> source homography, row-major
    int x0, y0, nx, ny;   // clipped bounding box of the quad
    int sw, sh, dw;
};

// Solve M h = rhs (M 8x8, augmented as column 8) by Gauss-Jordan elimination
// with partial pivoting; false when a pivot falls below kEps.
bool solve8(double M[8][9], double h[8]) {
    for (int col = 0; col < 8; ++col) {
        int piv = col;
        double mag = std::fabs(M[col][col]);
        for (int r = col + 1; r < 8; ++r)
            if (std::fabs(M[r][col]) > mag) {
                mag = std::fabs(M[r][col]);
                piv = r;
            }
        if (mag < kEps) return false;
        if (piv != col)
            for (int c = col; c < 9; ++c) std::swap(M[col][c], M[piv][c]);
        const double p = M[col][col];
        for (int c = col; c < 9; ++c) M[col][c] /= p;
        for (int r = 0; r < 8; ++r) {
            if (r == col) continue;
            const double f = M[r][col];
            if (std::fabs(f) < kEps) continue;
            for (int c = col; c < 9; ++c) M[r][c] -= f * M[col][c];
        }
    }
    for (int k = 0; k < 8; ++k) h[k] = M[k][8];
    return true;
}

// Inverse of a 3x3 matrix from its cofactors; false for |det| < kEps.
bool invert3(const double m[9], double out[9]) {
    const double a = m[0], b = m[1], c = m[2], d = m[3], e = m[4], f = m[5], g = m[6], h = m[7], i = m[8];
    const double c00 = (e * i - f * h), c01 = -(d * i - f * g), c02 = (d * h - e * g);
    const double c10 = -(b * i - c * h), c11 = (a * i - c * g), c12 = -(a * h - b * g);
    const double c20 = (b * f - c * e), c21 = -(a * f - c * d), c22 = (a * e - b * d);
    const double det = a * c00 + b * c01 + c * c02;
    if (std::fabs(det) < kEps) return false;
    const double s = 1.0 / det;
    out[0] = c00 * s;
    out[1] = c10 * s;
    out[2] = c20 * s;
    out[3] = c01 * s;
    out[4] = c11 * s;
    out[5] = c21 * s;
    out[6] = c02 * s;
    out[7] = c12 * s;
    out[8] = c22 * s;
    return true;
}

// false: nothing to paste (empty source, quad outside the destination, or a
// degenerate mapping), as the reference returns early in those cases
bool plan_paste(int sw, int sh, int dw, int dh, const int32_t* corners, PastePlan& P) {
    if (sw <= 0 || sh <= 0) return false;
    // corners: LT, RT, LB, RB as (x, y); the polygon runs LT, RT, RB, LB
    const int order[4] = {0, 1, 3, 2};
    for (int k = 0; k < 4; ++k) {
        P.qx[k] = (double)corners[2 * order[k]];
        P.qy[k] = (double)corners[2 * order[k] + 1];
    }
    const double lx = std::min({P.qx[0], P.qx[1], P.qx[2], P.qx[3]});
    const double hx = std::max({P.qx[0], P.qx[1], P.qx[2], P.qx[3]});
    const double ly = std::min({P.qy[0], P.qy[1], P.qy[2], P.qy[3]});
    const double hy = std::max({P.qy[0], P.qy[1], P.qy[2], P.qy[3]});
    if (hx < 0.0 || hy < 0.0 || lx > (double)(dw - 1) || ly > (double)(dh - 1)) return false;
    P.x0 = std::max(0, (int)std::floor(lx));
    P.y0 = std::max(0, (int)std::floor(ly));
    P.nx = std::min(dw - 1, (int)std::ceil(hx)) - P.x0 + 1;
    P.ny = std::min(dh - 1, (int)std::ceil(hy)) - P.y0 + 1;
    // x' = (h0 x + h1 y + h2) / (h6 x + h7 y + 1), y' = (h3 x + h4 y + h5) / (...):
    // two rows per source corner (0,0), (w-1,0), (0,h-1), (w-1,h-1)
    const double W1 = (double)sw - 1.0, H1 = (double)sh - 1.0;
    const double sx[4] = {0.0, W1, 0.0, W1}, sy[4] = {0.0, 0.0, H1, H1};
    double M[8][9] = {};
    for (int k = 0; k < 4; ++k) {
        const double x = sx[k], y = sy[k];
        const double u = (double)corners[2 * k], v = (double)corners[2 * k + 1];
        double* r0 = M[2 * k];
        double* r1 = M[2 * k + 1];
        r0[0] = x; r0[1] = y; r0[2] = 1.0; r0[6] = -u * x; r0[7] = -u * y; r0[8] = u;
        r1[3] = x; r1[4] = y; r1[5] = 1.0; r1[6] = -v * x; r1[7] = -v * y; r1[8] = v;
    }
    double h[8];
    if (!solve8(M, h)) return false;
    const double H[9] = {h[0], h[1], h[2], h[3], h[4], h[5], h[6], h[7], 1.0};
    if (!invert3(H, P.inv)) return false;
    P.sw = sw;
    P.sh = sh;
    P.dw = dw;
    return true;
}

__global__ __launch_bounds__(256) void k_paste(const float* __restrict__ src, float* __restrict__ dst, PastePlan P) {
    const int x = P.x0 + (int)(blockIdx.x * 16 + threadIdx.x % 16);
    const int y = P.y0 + (int)(blockIdx.y * 16 + threadIdx.x / 16);
    if (x >= P.x0 + P.nx || y >= P.y0 + P.ny) return;
    const double X = (double)x + 0.5, Y = (double)y + 0.5;
    // ray casting across the four edges (j, i) = (3,0), (0,1), (1,2), (2,3)
    bool inside = false;
    for (int i = 0, j = 3; i < 4; j = i++) {
        const double xi = P.qx[i], yi = P.qy[i], xj = P.qx[j], yj = P.qy[j];
        const double dy = yj - yi;
        const bool cross = ((yi > Y) != (yj > Y)) && (X < (xj - xi) * (Y - yi) / (dy == 0.0 ? 1e-30 : dy) + xi);
        inside = inside != cross;
    }
    if (!inside) return;
    const double* m = P.inv;
    const double den = m[6] * X + m[7] * Y + m[8];
    if (fabs(den) < kEps) return;
    const double u = (m[0] * X + m[1] * Y + m[2]) / den;
    const double v = (m[3] * X + m[4] * Y + m[5]) / den;
    const double W1 = (double)P.sw - 1.0, H1 = (double)P.sh - 1.0;
    if (u < 0.0 || v < 0.0 || u > W1 || v > H1) return;
    const double cu = fmin(fmax(u, 0.0), W1), cv = fmin(fmax(v, 0.0), H1);
    const int ix = (int)floor(cu), iy = (int)floor(cv);
    const int ix1 = min(ix + 1, P.sw - 1), iy1 = min(iy + 1, P.sh - 1);
    const double tx = cu - (double)ix, ty = cv - (double)iy;
    const double w00 = (1.0 - tx) * (1.0 - ty), w10 = tx * (1.0 - ty), w01 = (1.0 - tx) * ty, w11 = tx * ty;
    const float* c00 = src + 3 * ((size_t)iy * P.sw + ix);
    const float* c10 = src + 3 * ((size_t)iy * P.sw + ix1);
    const float* c01 = src + 3 * ((size_t)iy1 * P.sw + ix);
    const float* c11 = src + 3 * ((size_t)iy1 * P.sw + ix1);
    float* o = dst + 3 * ((size_t)y * P.dw + x);
#pragma unroll
    for (int k = 0; k < 3; ++k)
        o[k] = (float)((((double)c00[k] * w00 + (double)c10[k] * w10) + (double)c01[k] * w01) + (double)c11[k] * w11);
}

}  // namespace

Status paste_device(const float* src, uint32_t sw, uint32_t sh, float* dst, uint32_t dw, uint32_t dh,
                    const int32_t* corners, hipStream_t s, bool* painted) {
    PastePlan P{};
    *painted = false;
    if (!plan_paste((int)sw, (int)sh, (int)dw, (int)dh, corners, P)) return Status::ok();
    const dim3 grid((unsigned)((P.nx + 15) / 16), (unsigned)((P.ny + 15) / 16));
    k_paste<<<grid, 256, 0, s>>>(src, dst, P);
    AUR_HIP(hipGetLastError());
    *painted = true;
    return Status::ok();
}

}  // namespace aur
