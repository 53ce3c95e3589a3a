// The reference's fixed-step integrator on gfx950 (method='numeric').
//
// Forward (libcpab/core/cpab_ops.cpp:224-260): nstep applications of
// Trel_c = expm(A~_c / nstep), c the cell found at the start of each step.
// Backward (libcpab/core/cpab_ops.cpp:262-372): per basis direction k a
// midpoint-RK2 integration of (p, q = dp/dtheta_k), q' = B_k,c p~ + A_c q.
// Arithmetic follows the reference operation by operation (FP32 state, FP64
// cell lookup and RK2 half steps, no fused multiply-adds: contraction is off
// in this file), so results match the reference's compiled CPU op bit for bit
// on the same Trels.  One thread per (point, theta) in the forward and per
// (point, theta, k) in the backward, as the work decomposes in the reference.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

namespace cpab_ref {

__device__ __forceinline__ int mymin(int a, double b) { return !(b < a) ? a : (int)round(b); }

__device__ __forceinline__ int find1(const float* p, int nx) {
#pragma clang fp contract(off)
    int idx = (int)floor(p[0] * nx);
    idx = idx > nx - 1 ? nx - 1 : idx;
    return idx < 0 ? 0 : idx;
}

__device__ __forceinline__ int find2(const float* p, int nx, int ny) {
#pragma clang fp contract(off)
    const double px = p[0], py = p[1];
    const float inc_x = (float)(1.0 / nx), inc_y = (float)(1.0 / ny);
    const double p0 = fmin(nx * inc_x - 0.000000001, fmax(0.0, px));
    const double p1 = fmin(ny * inc_y - 0.000000001, fmax(0.0, py));
    const double xmod = fmod(p0, (double)inc_x), ymod = fmod(p1, (double)inc_y);
    const double x = xmod / inc_x, y = ymod / inc_y;
    int cell = mymin(nx - 1, (p0 - xmod) / inc_x) + mymin(ny - 1, (p1 - ymod) / inc_y) * nx;
    cell *= 4;
    if (px <= 0) {
        if (py <= 0 && py / inc_y < px / inc_x) return cell;
        if (py >= ny * inc_y && py / inc_y - ny > -px / inc_x) return cell + 2;
        return cell + 3;
    }
    if (px >= nx * inc_x) {
        if (py <= 0 && -py / inc_y > px / inc_x - nx) return cell;
        if (py >= ny * inc_y && py / inc_y - ny > px / inc_x - nx) return cell + 2;
        return cell + 1;
    }
    if (py <= 0) return cell;
    if (py >= ny * inc_y) return cell + 2;
    if (x < y) return cell + ((1 - x < y) ? 2 : 3);
    return cell + ((1 - x < y) ? 1 : 0);
}

__device__ __forceinline__ int find3(const float* pin, int nx, int ny, int nz) {
#pragma clang fp contract(off)
    const float inc_x = (float)(1.0 / nx), inc_y = (float)(1.0 / ny), inc_z = (float)(1.0 / nz);
    float point[3] = {pin[0], pin[1], pin[2]};
    if (point[0] < 0.0 || point[0] > 1.0 || point[1] < 0.0 || point[1] > 1.0) {
        const float half = 0.5f;
        point[0] -= half;
        point[1] -= half;
        point[2] -= half;
        const float ax = fabsf(point[0]), ay = fabsf(point[1]), az = fabsf(point[2]);
        const float px = (ax < ay && ax < az) ? half * inc_x : 0.0f;
        const float py = (ay < ax && ax < az) ? half * inc_y : 0.0f;
        const float pz = (az < ax && ax < ay) ? half * inc_z : 0.0f;
        if (ax > half) point[0] = copysignf(half - px, point[0]);
        if (ay > half) point[1] = copysignf(half - py, point[1]);
        if (az > half) point[2] = copysignf(half - pz, point[2]);
        point[0] += half;
        point[1] += half;
        point[2] += half;
    }
    const float p0 = fminf((float)(nx * inc_x - 1e-8), fmaxf(0.0f, point[0]));
    const float p1 = fminf((float)(ny * inc_y - 1e-8), fmaxf(0.0f, point[1]));
    const float p2 = fminf((float)(nz * inc_x - 1e-8), fmaxf(0.0f, point[2]));  // inc_x as in the reference
    const double xmod = fmodf(p0, inc_x), ymod = fmodf(p1, inc_y), zmod = fmodf(p2, inc_z);
    const int i = mymin(nx - 1, ((p0 - xmod) / inc_x));
    const int j = mymin(ny - 1, ((p1 - ymod) / inc_y));
    const int k = mymin(nz - 1, ((p2 - zmod) / inc_z));
    int cell = 5 * (i + j * nx + k * nx * ny);
    double x = xmod / inc_x, y = ymod / inc_y;
    const double z = zmod / inc_z;
    bool tf = false;
    if (k % 2 == 0) {
        if ((i % 2 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 2 == 0)) tf = true;
    } else if ((i % 2 == 0 && j % 2 == 0) || (i % 2 == 1 && j % 2 == 1)) {
        tf = true;
    }
    if (tf) {
        const double t = x;
        x = y;
        y = 1 - t;
    }
    if (-x - y + z >= 0) return cell + 1;
    if (x + y + z - 2 >= 0) return cell + 2;
    if (-x + y - z >= 0) return cell + 3;
    if (x - y - z >= 0) return cell + 4;
    return cell;
}

template <int D>
__device__ __forceinline__ int find(const float* p, const int* nc) {
    if constexpr (D == 1) return find1(p, nc[0]);
    else if constexpr (D == 2) return find2(p, nc[0], nc[1]);
    else return find3(p, nc[0], nc[1], nc[2]);
}

// x = A b~ (A row-major D x (D+1)), evaluated left to right like the reference
template <int D>
__device__ __forceinline__ void affine(float* x, const float* A, const float* b) {
#pragma clang fp contract(off)
#pragma unroll
    for (int r = 0; r < D; ++r) {
        float acc = A[r * (D + 1)] * b[0];
#pragma unroll
        for (int a = 1; a < D; ++a) acc = acc + A[r * (D + 1) + a] * b[a];
        x[r] = acc + A[r * (D + 1) + D];
    }
}

template <int D>
__device__ __forceinline__ void linear(float* x, const float* A, const float* b) {
#pragma clang fp contract(off)
#pragma unroll
    for (int r = 0; r < D; ++r) {
        float acc = A[r * (D + 1)] * b[0];
#pragma unroll
        for (int a = 1; a < D; ++a) acc = acc + A[r * (D + 1) + a] * b[a];
        x[r] = acc;
    }
}

}  // namespace cpab_ref
