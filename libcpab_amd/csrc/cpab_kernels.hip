// HIP kernels and the extern "C" boundary of the CPA integration path.
//
// Kernel map (one block = 256 threads = 4 waves of 64; a block owns one
// theta and a tile of 256*ppt points, its theta's per-cell parameters and the
// facet table are staged in LDS; block ids are remapped so that all tiles of a
// theta land on the same XCD and share its L2):
//   k_theta_to_A        A = B theta, converted to q-space        (per theta)
//   k_transform         closed-form forward of explicit points    (transform_grid)
//   k_transform_data    meshgrid -> closed-form -> interpolation  (transform_data)
//   k_transform_bwd     adjoint of k_transform, dL/dA partials per tile (+ dL/dpoints)
//   k_transform_data_bwd  adjoint of k_transform_data (+ dL/ddata scatter)
//   k_interp_fwd / k_interp_bwd   linear interpolation of data at a grid
//   k_reduce_theta      per-theta sum of the tile partials, then B^T (many thetas)
//   k_reduce_cols       column slices of that sum (few thetas), then
//   k_reduce_bt         B^T per theta (all in a fixed order: deterministic)
#include <hip/hip_runtime.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include <mutex>

#include "cpab_device.hpp"
#include "cpab_numeric.hpp"
#include "../../include/cpab_hip.h"

using namespace cpab;

namespace {

constexpr int BLOCK = 256;
constexpr int WAVES = BLOCK / 64;

// ---------------------------------------------------------------------------
// argument blocks
// ---------------------------------------------------------------------------
struct Geo {
    int ndim, nc0, nc1, nc2, nC;
    const float4* pl;
    const int* nb;
    float tol0;
    int* unfinished;  // nullable: count of points that used up the chunk budget
    int nV;           // virtual cells after the nC simplices (3D outside the domain)
    const int* amap;  // [nV] simplex whose map each virtual cell uses
};


struct Src {  // where points come from
    const float* pts;  // explicit [ndim, nP] or [n_theta, ndim, nP]
    int per_theta;
    const float* lin;  // meshgrid: concatenated axis coordinates
    int g0, g1, g2;    // meshgrid / output sizes per axis
};

struct Img {  // data for interpolation, NC[W(,H(,D))] layout
    const float* data;
    float* gdata;
    int C, s0, s1, s2;
};

struct Launch {
    int n_theta, nP, tiles, ppt;
    int ty;  // 2D output tiles TX x ty (TX = ppt * BLOCK / ty); 0: runs of ppt*BLOCK consecutive points
    float t_end;
};

// Tiling of a meshgrid output.  2D grids with a second axis of at least 8 go
// in compact tiles (their warped footprint in the data is compact too, which
// the backward accumulates in LDS); 3D grids whose last two axes are at least
// 8 go in tiles of 8 x 8 lanes per wave (last two axes) stacked along the
// first axis; everything else in consecutive runs.  The returned tile height
// is 0 for runs.
__host__ __device__ inline int tile_ty(int ndim, int g1, int g2 = 1) {
    if (ndim == 2) return g1 >= 16 ? 16 : (g1 >= 8 ? 8 : 0);  // 16: a wave is 4 x 16 pixels (32 and 8 measured slower)
    if (ndim == 3) return (g1 >= 8 && g2 >= 8) ? 8 : 0;
    return 0;
}

__host__ __device__ inline int tile_count(int ndim, int nP, int g0, int g1, int ppt, int g2 = 1) {
    const int ty = tile_ty(ndim, g1, g2);
    if (ty == 0) {
        const int n = (nP + BLOCK * ppt - 1) / (BLOCK * ppt);
        return n > 0 ? n : 1;
    }
    if (ndim == 3) {  // 8 x 8 lanes in (g1, g2), WAVES * ppt rows of g0
        const int tx = ppt * WAVES;
        return ((g0 + tx - 1) / tx) * ((g1 + 7) / 8) * ((g2 + 7) / 8);
    }
    const int tx = ppt * (BLOCK / ty);
    return ((g0 + tx - 1) / tx) * ((g1 + ty - 1) / ty);
}

// dL/ddata bins of a 2D tile: at most BIN_CELLS footprint cells (11-bit ids)
constexpr int BIN_CELLS = 2048;

// in-place exclusive prefix sum of v[0, n) by the whole block; v[n-1] becomes
// the sum of the first n-1 entries
__device__ __forceinline__ void block_exclusive_scan(int* v, int n) {
    __shared__ int s_wsum[WAVES];
    const int per = (n + BLOCK - 1) / BLOCK;
    const int b0 = (int)threadIdx.x * per, b1 = min(b0 + per, n);
    int loc = 0;
    for (int k = b0; k < b1; ++k) loc += v[k];
    const int inc = wave_inclusive_scan(loc);
    const int w = threadIdx.x >> 6;
    if (__lane_id() == 63) s_wsum[w] = inc;
    __syncthreads();
    int run = inc - loc;
    for (int w2 = 0; w2 < w; ++w2) run += s_wsum[w2];
    for (int k = b0; k < b1; ++k) {
        const int x = v[k];
        v[k] = run;
        run += x;
    }
    __syncthreads();
}

// q = n / d, rem = n - q d for 0 <= n < 2^20 and 0 < d < 2^20, rd = 1.0f / d:
// the float product is within 1/8 of n / d, so truncating it is off by at
// most one, which the remainder test corrects
__device__ __forceinline__ int div_small(int n, int d, float rd, int& rem) {
    int q = (int)((float)n * rd);
    int r = n - q * d;
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
    rem = r;
    return q;
}

// output-order index t of round r of this thread in tile `tile`, -1 if outside
template <int D>
__device__ __forceinline__ int tile_t(const Launch& L, const Src& s, int tile, int r) {
    if (L.ty == 0) {
        const int t = (tile * L.ppt + r) * BLOCK + (int)threadIdx.x;
        return t < L.nP ? t : -1;
    }
    if constexpr (D == 3) {
        const int tiles_z = (s.g2 + 7) >> 3, tiles_y = (s.g1 + 7) >> 3;
        const int tz = tile % tiles_z, ty = (tile / tiles_z) % tiles_y, tx = tile / (tiles_z * tiles_y);
        const int iz = tz * 8 + ((int)threadIdx.x & 7);
        const int iy = ty * 8 + (((int)threadIdx.x >> 3) & 7);
        const int ix = (tx * L.ppt + r) * WAVES + ((int)threadIdx.x >> 6);
        return (ix < s.g0 && iy < s.g1 && iz < s.g2) ? (ix * s.g1 + iy) * s.g2 + iz : -1;
    }
    const int rows = BLOCK / L.ty;
    const int tiles_y = (s.g1 + L.ty - 1) / L.ty;
    const int ix = (tile / tiles_y) * (L.ppt * rows) + r * rows + (int)threadIdx.x / L.ty;
    const int iy = (tile % tiles_y) * L.ty + (int)threadIdx.x % L.ty;
    return (ix < s.g0 && iy < s.g1) ? ix * s.g1 + iy : -1;
}

__device__ __forceinline__ int xcd_remap(int b, int nb) {
    // blocks are dispatched round-robin over 8 XCDs; give each XCD a contiguous
    // range of logical tiles so one theta's tiles share an L2
    const int xcd = b & 7, k = b >> 3;
    const int per = nb >> 3, rem = nb & 7;
    return xcd * per + (xcd < rem ? xcd : rem) + k;
}

template <int D>
__device__ __forceinline__ int nc_at(const Geo& g, int a) {
    return a == 0 ? g.nc0 : (a == 1 ? g.nc1 : g.nc2);
}

// point index p of output-order index t (last output axis fastest)
template <int D>
__device__ __forceinline__ int out_to_point(const Src& s, int t, int (&ix)[3]) {
    if constexpr (D == 1) {
        ix[0] = t;
        return t;
    } else if constexpr (D == 2) {
        ix[1] = t % s.g1;
        ix[0] = t / s.g1;
        return ix[0] + s.g0 * ix[1];
    } else {
        ix[2] = t % s.g2;
        const int r = t / s.g2;
        ix[1] = r % s.g1;
        ix[0] = r / s.g1;
        return ix[0] + s.g0 * (ix[1] + s.g1 * ix[2]);
    }
}

// LDS layout of one theta's tables: A_q [nC*P], 1/|M|_F [nC (padded to 4)],
// facet planes [nC*NF] float4, neighbours [nC*NF] int
template <int D>
__device__ __forceinline__ CellTable<D> table_view(float* smem, int nC) {
    constexpr int P = D * (D + 1);
    CellTable<D> t;
    float* sA = smem;
    float* sInv = sA + nC * P;
    float4* sPl = reinterpret_cast<float4*>(sInv + ((nC + 3) & ~3));
    t.A = sA;
    t.inv_n = sInv;
    t.pl = sPl;
    t.nb = reinterpret_cast<int*>(sPl + (D > 1 ? nC * NF : 0));
    return t;
}

// stage theta's q-space affine parameters, their inverse norms and the facet
// table; a virtual cell gets a copy of the parameters of the simplex it uses
template <int D>
__device__ __forceinline__ CellTable<D> stage_tables(float* smem, const float* Aq, const Geo& g) {
    constexpr int P = D * (D + 1);
    const int nC = g.nC;
    const int nCt = nC + g.nV;
    CellTable<D> t = table_view<D>(smem, nCt);
    float* sA = const_cast<float*>(t.A);
    float* sInv = const_cast<float*>(t.inv_n);
    float4* sPl = const_cast<float4*>(t.pl);
    int* sNb = const_cast<int*>(t.nb);
    for (int i = threadIdx.x; i < nC * P; i += BLOCK) sA[i] = Aq[i];
    if constexpr (D == 3) {
        for (int i = threadIdx.x; i < g.nV * P; i += BLOCK) sA[nC * P + i] = Aq[g.amap[i / P] * P + i % P];
    }
    if constexpr (D > 1) {
        for (int i = threadIdx.x; i < nCt * NF; i += BLOCK) {
            sPl[i] = g.pl[i];
            sNb[i] = g.nb[i];
        }
    }
    __syncthreads();
    for (int c = threadIdx.x; c < nCt; c += BLOCK) {
        float s = 0.f;
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int k = 0; k < D; ++k) {
                const float m = sA[c * P + r * (D + 1) + k];
                s = fmaf(m, m, s);
            }
        sInv[c] = s > 0.f ? rsqrtf(s) : INFINITY;
    }
    __syncthreads();
    return t;
}

// Rows of the fused forward's saved state per point: warped point, q-space
// Jacobian and, in 2D/3D, the chunk log (NREC durations + the facet code).
__host__ __device__ constexpr int saved_rows(int D) { return D + D * D + (D > 1 ? NREC + 1 : 0); }
// Rows of the grid transform's saved state (the jac buffer): q-space Jacobian
// and, in 2D/3D, the chunk log.
__host__ __device__ constexpr int jac_rows(int D) { return D * D + (D > 1 ? NREC + 1 : 0); }

template <int D>
__host__ __device__ constexpr size_t table_floats(int nC) {
    return (size_t)nC * D * (D + 1) + ((nC + 3) & ~3) + (D > 1 ? (size_t)nC * NF * 4 + (size_t)nC * NF : 0);
}

// Start cell of a point: the lookup, except that a 3D point outside the
// domain of a tessellation with virtual cells starts in the region that
// contains it -- the virtual cell whose largest facet value is the smallest
// (its simplex is then the one the lookup returns for the pushed point).
template <int D>
__device__ __forceinline__ int start_cell(const float (&x)[D], const int* nc, const CellTable<D>& tab, const Geo& g) {
    int c = find_cell<D>(x, nc);
    if constexpr (D == 3) {
        if (g.nV > 0 && (x[0] < 0.f || x[0] > 1.f || x[1] < 0.f || x[1] > 1.f || x[2] < 0.f || x[2] > 1.f)) {
            const float q[3] = {x[0] * (float)nc[0], x[1] * (float)nc[1], x[2] * (float)nc[2]};
            float best = INFINITY;
            for (int v = g.nC; v < g.nC + g.nV; ++v) {
                float worst = -INFINITY;
#pragma unroll
                for (int f = 0; f < NF; ++f) {
                    if (tab.nb[v * NF + f] == -2) continue;
                    const float4 P = tab.pl[v * NF + f];
                    worst = fmaxf(worst, fmaf(P.x, q[0], fmaf(P.y, q[1], fmaf(P.z, q[2], P.w))));
                }
                if (worst < best) {
                    best = worst;
                    c = v;
                }
            }
        }
    }
    return c;
}

// The simplex whose dL/dA a chunk in `cell` adds to (-1 stays -1)
template <int D>
__device__ __forceinline__ int grad_cell(int cell, const Geo& g) {
    if constexpr (D == 3) {
        if (cell >= g.nC) return g.amap[cell - g.nC];
    }
    return cell;
}

// Pass 2 over one point per lane (wave-convergent); accumulates into acc.
// The chunks recorded in lg by the forward are replayed (no facet search);
// the rest are found again.
template <int D>
__device__ __forceinline__ void pass2(bool valid, const float (&q0)[D], int c0, float T, float (&mu)[D],
                                      const CellTable<D>& tab, const Geo& g, float* acc, const ChunkLog& lg) {
    constexpr int P = D * (D + 1);
    PointState<D> st;
#pragma unroll
    for (int a = 0; a < D; ++a) st.q[a] = q0[a];
    st.T = valid ? T : 0.f;
    st.c = c0;
    st.came_from = -1;
    st.zero_hops = 0;
    st.iters = 0;
    int nrec = lg.code & 3;
    while (true) {
        const bool active = st.T > 0.f && st.iters < MAX_ITERS;
        if (__ballot(active) == 0) break;
        int cell = -1;
        float G[P];
#pragma unroll
        for (int p = 0; p < P; ++p) G[p] = 0.f;
        if (active) {
            if constexpr (D == 1) {
                Chunk1 ch;
                advance1(st, tab, g.nc0, ch);
                float G1[2];
                chunk_grad1(ch, mu[0], G1);
                G[0] = G1[0];
                G[1] = G1[1];
                cell = ch.cell;
                if (ch.facet >= 0) {
                    const float vn = fmaf(tab.A[2 * ch.next], ch.q1, tab.A[2 * ch.next + 1]);
                    const float vp = fmaf(ch.a, ch.q1, ch.b);
                    if (fabsf(vn) > 1e-6f * fabsf(vp) && vn != 0.f) mu[0] *= vp / vn;
                }
            } else {
                Chunk<D> ch;
                float h = 0.f;
                int bf = -1;
                if (st.iters < nrec) {
                    logged_chunk(lg, st.iters, h, bf);
                    // only a chunk the forward can have made is replayed (a
                    // saved state from another call is searched again instead)
                    if (!(h >= 0.f && h <= st.T) || bf >= NF || (bf >= 0 && tab.nb[st.c * NF + bf] < 0)) nrec = 0;
                }
                if (st.iters < nrec)
                    replay<D>(st, tab, h, bf, ch);
                else
                    advance<D>(st, tab, g.tol0, ch);
                chunk_grad<D>(ch, mu, G);
                cell = ch.cell;
                if (ch.facet >= 0) saltation_adj<D>(tab, ch, mu);
            }
        }
        wave_accumulate<P>(acc, grad_cell<D>(cell, g), G);
    }
}

// ---------------------------------------------------------------------------
// interpolation helpers (libcpab/pytorch/interpolation.py semantics)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void axis_corners(float x, int size, int& i0, int& i1, float& fd) {
    // the rounded FP32 product, as the reference's float32 tensors form it:
    // no contraction of x * (size - 1) into fd's subtraction (an FMA would
    // keep the unrounded product, up to half an ulp of x * (size - 1) away)
#pragma clang fp contract(off)
    const float xs = x * (float)(size - 1);
    float f0 = floorf(xs);
    f0 = fminf(fmaxf(f0, -1.f), (float)size);
    int a0 = (int)f0;
    int a1 = a0 + 1;
    a0 = clampi(a0, 0, size - 1);
    a1 = clampi(a1, 0, size - 1);
    i0 = a0;
    i1 = a1;
    fd = xs - (float)a0;
}

template <int D>
__device__ __forceinline__ void corners(const float (&y)[D], const Img& im, int (&i0)[3], int (&i1)[3], float (&fd)[3]) {
    const int sz[3] = {im.s0, im.s1, im.s2};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        i0[a] = 0;
        i1[a] = 0;
        fd[a] = 0.f;
    }
#pragma unroll
    for (int a = 0; a < D; ++a) axis_corners(y[a], sz[a], i0[a], i1[a], fd[a]);
}

// Offsets of the 2^D interpolation corners in one channel's data, corner
// j = sum_a bit_a 2^(D-1-a) taking i1 on axis a where bit_a is set (2D:
// c00, c01, c10, c11 with the second index the contiguous axis).
template <int D>
__device__ __forceinline__ void corner_offsets(const Img& im, const int (&i0)[3], const int (&i1)[3],
                                               int (&off)[1 << D]) {
    if constexpr (D == 1) {
        off[0] = i0[0];
        off[1] = i1[0];
    } else if constexpr (D == 2) {
        const int H = im.s1;
        off[0] = i0[0] * H + i0[1];
        off[1] = i0[0] * H + i1[1];
        off[2] = i1[0] * H + i0[1];
        off[3] = i1[0] * H + i1[1];
    } else {
        const int H = im.s1, Z = im.s2;
        const int ii[3][2] = {{i0[0], i1[0]}, {i0[1], i1[1]}, {i0[2], i1[2]}};
#pragma unroll
        for (int bx = 0; bx < 2; ++bx)
#pragma unroll
            for (int by = 0; by < 2; ++by)
#pragma unroll
                for (int bz = 0; bz < 2; ++bz) off[bx * 4 + by * 2 + bz] = (ii[0][bx] * H + ii[1][by]) * Z + ii[2][bz];
    }
}

// the interpolated value from the corner values v (corner_offsets order)
template <int D>
__device__ __forceinline__ float blend(const float (&v)[1 << D], const float (&fd)[3]) {
    if constexpr (D == 1) {
        return v[0] * (1.f - fd[0]) + v[1] * fd[0];
    } else if constexpr (D == 2) {
        const float c0 = v[0] * (1.f - fd[0]) + v[2] * fd[0];
        const float c1 = v[1] * (1.f - fd[0]) + v[3] * fd[0];
        return c0 * (1.f - fd[1]) + c1 * fd[1];
    } else {
        const float c00 = v[0] * (1.f - fd[0]) + v[4] * fd[0];
        const float c01 = v[1] * (1.f - fd[0]) + v[5] * fd[0];
        const float c10 = v[2] * (1.f - fd[0]) + v[6] * fd[0];
        const float c11 = v[3] * (1.f - fd[0]) + v[7] * fd[0];
        const float c0 = c00 * (1.f - fd[1]) + c10 * fd[1];
        const float c1 = c01 * (1.f - fd[1]) + c11 * fd[1];
        return c0 * (1.f - fd[2]) + c1 * fd[2];
    }
}

// value of channel data dp at the corners (i0, i1, fd)
template <int D>
__device__ __forceinline__ float interp_value(const float* __restrict__ dp, const Img& im, const int (&i0)[3],
                                              const int (&i1)[3], const float (&fd)[3]) {
    int off[1 << D];
    corner_offsets<D>(im, i0, i1, off);
    float v[1 << D];
#pragma unroll
    for (int j = 0; j < (1 << D); ++j) v[j] = dp[off[j]];
    return blend<D>(v, fd);
}

// Channels whose loads one group issues together.  The forward keeps one (a
// group of four, measured on the 2D image workload, gained nothing and cost
// it registers); the 2D backward takes two (the others keep their own loop).
template <int D>
constexpr int chan_group() { return 1; }
template <int D>
constexpr int chan_group_bwd() { return 2; }

// out[c * ostride] = channel c (data + c * dstride) at the corners, for every
// channel.  The loads of a group of channels are all issued before the first
// of them is used, so their latencies overlap instead of adding up.
template <int D>
__device__ __forceinline__ void interp_channels(const float* __restrict__ data, size_t dstride, const Img& im,
                                                const int (&i0)[3], const int (&i1)[3], const float (&fd)[3],
                                                float* __restrict__ out, size_t ostride) {
    constexpr int NK = 1 << D, CG = chan_group<D>();
    int off[NK];
    corner_offsets<D>(im, i0, i1, off);
    for (int c0 = 0; c0 < im.C; c0 += CG) {
        float v[CG][NK];
#pragma unroll
        for (int k = 0; k < CG; ++k)
            if (c0 + k < im.C) {
                const float* dp = data + (size_t)(c0 + k) * dstride;
#pragma unroll
                for (int j = 0; j < NK; ++j) v[k][j] = dp[off[j]];
            }
#pragma unroll
        for (int k = 0; k < CG; ++k)
            if (c0 + k < im.C) out[(size_t)(c0 + k) * ostride] = blend<D>(v[k], fd);
    }
}

// backward of blend for upstream g at corner values v: adds d value / d fd[a]
// into dfd and, with gp (nullable), scatters g times each corner's weight
template <int D>
__device__ __forceinline__ void blend_backward(const float (&v)[1 << D], const int (&off)[1 << D], float* gp,
                                               const float (&fd)[3], float g, float (&dfd)[3]) {
    if constexpr (D == 1) {
        if (gp) {
            atomicAdd(gp + off[0], g * (1.f - fd[0]));
            atomicAdd(gp + off[1], g * fd[0]);
        }
        dfd[0] = fmaf(g, v[1] - v[0], dfd[0]);
    } else if constexpr (D == 2) {
        const float c00 = v[0], c01 = v[1], c10 = v[2], c11 = v[3];
        const float wx0 = 1.f - fd[0], wx1 = fd[0], wy0 = 1.f - fd[1], wy1 = fd[1];
        if (gp) {
            atomicAdd(gp + off[0], g * wx0 * wy0);
            atomicAdd(gp + off[2], g * wx1 * wy0);
            atomicAdd(gp + off[1], g * wx0 * wy1);
            atomicAdd(gp + off[3], g * wx1 * wy1);
        }
        const float c0 = c00 * wx0 + c10 * wx1;
        const float c1 = c01 * wx0 + c11 * wx1;
        dfd[0] = fmaf(g, (c10 - c00) * wy0 + (c11 - c01) * wy1, dfd[0]);
        dfd[1] = fmaf(g, c1 - c0, dfd[1]);
    } else {
        const float w[3][2] = {{1.f - fd[0], fd[0]}, {1.f - fd[1], fd[1]}, {1.f - fd[2], fd[2]}};
        float cv[2][2][2];
#pragma unroll
        for (int bx = 0; bx < 2; ++bx)
#pragma unroll
            for (int by = 0; by < 2; ++by)
#pragma unroll
                for (int bz = 0; bz < 2; ++bz) {
                    const int j = bx * 4 + by * 2 + bz;
                    cv[bx][by][bz] = v[j];
                    if (gp) atomicAdd(gp + off[j], g * w[0][bx] * w[1][by] * w[2][bz]);
                }
        // derivative of the trilinear blend w.r.t. fd[a]
        float dx = 0.f, dy = 0.f, dz = 0.f;
#pragma unroll
        for (int by = 0; by < 2; ++by)
#pragma unroll
            for (int bz = 0; bz < 2; ++bz) dx += (cv[1][by][bz] - cv[0][by][bz]) * w[1][by] * w[2][bz];
#pragma unroll
        for (int bx = 0; bx < 2; ++bx)
#pragma unroll
            for (int bz = 0; bz < 2; ++bz) dy += (cv[bx][1][bz] - cv[bx][0][bz]) * w[0][bx] * w[2][bz];
#pragma unroll
        for (int bx = 0; bx < 2; ++bx)
#pragma unroll
            for (int by = 0; by < 2; ++by) dz += (cv[bx][by][1] - cv[bx][by][0]) * w[0][bx] * w[1][by];
        dfd[0] = fmaf(g, dx, dfd[0]);
        dfd[1] = fmaf(g, dy, dfd[1]);
        dfd[2] = fmaf(g, dz, dfd[2]);
    }
}

// backward of interp_value for channel data dp with upstream g: scatters into
// gp (nullable) and returns d value / d fd[a] accumulated into dfd
template <int D>
__device__ __forceinline__ void interp_backward(const float* __restrict__ dp, float* gp, const Img& im, const int (&i0)[3],
                                                const int (&i1)[3], const float (&fd)[3], float g, float (&dfd)[3]) {
    int off[1 << D];
    corner_offsets<D>(im, i0, i1, off);
    float v[1 << D];
#pragma unroll
    for (int j = 0; j < (1 << D); ++j) v[j] = dp[off[j]];
    blend_backward<D>(v, off, gp, fd, g, dfd);
}

// interp_backward over every channel c, upstream gout[c * gstride] (the first
// four also returned in gk): the loads of a group of channels' upstream values
// and corners are issued before any of them is used; dfd accumulates in
// channel order.  gdata (nullable) receives the scatter.
template <int D>
__device__ __forceinline__ void interp_backward_channels(const float* __restrict__ data, float* gdata, size_t dstride,
                                                         const Img& im, const int (&i0)[3], const int (&i1)[3],
                                                         const float (&fd)[3], const float* __restrict__ gout,
                                                         size_t gstride, float (&gk)[4], float (&dfd)[3]) {
    constexpr int NK = 1 << D, CG = chan_group_bwd<D>();
    int off[NK];
    corner_offsets<D>(im, i0, i1, off);
    for (int c0 = 0; c0 < im.C; c0 += CG) {
        float v[CG][NK], gv[CG];
#pragma unroll
        for (int k = 0; k < CG; ++k)
            if (c0 + k < im.C) {
                const float* dp = data + (size_t)(c0 + k) * dstride;
                gv[k] = gout[(size_t)(c0 + k) * gstride];
#pragma unroll
                for (int j = 0; j < NK; ++j) v[k][j] = dp[off[j]];
            }
#pragma unroll
        for (int k = 0; k < CG; ++k)
            if (c0 + k < im.C) {
                blend_backward<D>(v[k], off, gdata ? gdata + (size_t)(c0 + k) * dstride : nullptr, fd, gv[k], dfd);
                if (c0 + k < 4) {
#pragma unroll
                    for (int j = 0; j < 4; ++j)
                        if (j == c0 + k) gk[j] = gv[k];
                }
            }
    }
}

// ---------------------------------------------------------------------------
// kernels
// ---------------------------------------------------------------------------
// A = B theta in q-space: grid (ceil(Dt/256), n_theta)
__global__ __launch_bounds__(BLOCK) void k_theta_to_A(const float* __restrict__ theta, const float* __restrict__ B,
                                                      int d, int Dt, int ndim, int nc0, int nc1, int nc2,
                                                      float* __restrict__ Aq) {
    extern __shared__ float sth[];
    const int th = blockIdx.y;
    for (int k = threadIdx.x; k < d; k += BLOCK) sth[k] = theta[(size_t)th * d + k];
    __syncthreads();
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= Dt) return;
    const float* b = B + (size_t)i * d;
    float acc = 0.f;
    for (int k = 0; k < d; ++k) acc = fmaf(b[k], sth[k], acc);
    const int P = ndim * (ndim + 1);
    const int rem = i % P, r = rem / (ndim + 1), col = rem % (ndim + 1);
    const int nc[3] = {nc0, nc1, nc2};
    const float sc = (col < ndim) ? (float)nc[r] / (float)nc[col] : (float)nc[r];
    Aq[(size_t)th * Dt + i] = acc * sc;
}

template <int D, bool WITHJ>
__global__ __launch_bounds__(BLOCK) void k_transform(Launch L, Geo g, Src s, const float* __restrict__ Aq,
                                                     float* __restrict__ out, float* __restrict__ jac) {
    extern __shared__ float smem[];
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int th = lb / L.tiles, tile = lb % L.tiles;
    if (th >= L.n_theta) return;
    const int nC = g.nC;
    const CellTable<D> tab = stage_tables<D>(smem, Aq + (size_t)th * nC * D * (D + 1), g);
    const int nc[3] = {g.nc0, g.nc1, g.nc2};
    const float rnc[3] = {1.0f / (float)g.nc0, 1.0f / (float)g.nc1, 1.0f / (float)g.nc2};  // q -> x
    const float* pts = s.pts + (s.per_theta ? (size_t)th * D * L.nP : 0);
    for (int r = 0; r < L.ppt; ++r) {
        const int p = tile_t<D>(L, s, tile, r);  // point index (2D meshgrid tiles when L.ty > 0)
        if (p < 0) continue;
        float x[D], q[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            x[a] = pts[(size_t)a * L.nP + p];
            q[a] = x[a] * (float)nc[a];
        }
        const int c0 = start_cell<D>(x, nc, tab, g);
        float J[D][D];
        float* jo = jac + (size_t)th * jac_rows(D) * L.nP;
        LogSink lg;
        if constexpr (WITHJ && D > 1) lg.init(jo + (size_t)D * D * L.nP, L.nP, p);
        integrate<D, WITHJ, WITHJ && (D > 1), LogSink>(q, c0, L.t_end, tab, g.nc0, g.tol0, J, lg, g.unfinished);
        float* o = out + (size_t)th * D * L.nP;
#pragma unroll
        for (int a = 0; a < D; ++a) o[(size_t)a * L.nP + p] = q[a] * rnc[a];
        if constexpr (WITHJ) {
            if constexpr (D > 1) jo[(size_t)(D * D + NREC) * L.nP + p] = __int_as_float(lg.code);
#pragma unroll
            for (int r2 = 0; r2 < D; ++r2)
#pragma unroll
                for (int k = 0; k < D; ++k) jo[(size_t)(r2 * D + k) * L.nP + p] = J[r2][k];
        }
    }
}

// Validation: the cells each point's trajectory passes through, as segments
// of positive duration (consecutive chunks in one cell merge; zero-length
// visits at vertices are not segments).  cells [n_theta, nP, max_segs] gets
// the first max_segs (-1 padded), nseg [n_theta, nP] the full count.  Same
// integration as k_transform.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_transform_cells(Launch L, Geo g, Src s, const float* __restrict__ Aq,
                                                           int max_segs, int* __restrict__ cells,
                                                           int* __restrict__ nseg) {
    extern __shared__ float smem[];
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int th = lb / L.tiles, tile = lb % L.tiles;
    if (th >= L.n_theta) return;
    const CellTable<D> tab = stage_tables<D>(smem, Aq + (size_t)th * g.nC * D * (D + 1), g);
    const int nc[3] = {g.nc0, g.nc1, g.nc2};
    const float* pts = s.pts + (s.per_theta ? (size_t)th * D * L.nP : 0);
    for (int r = 0; r < L.ppt; ++r) {
        const int p = (tile * L.ppt + r) * BLOCK + threadIdx.x;
        if (p >= L.nP) break;
        float x[D], q[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            x[a] = pts[(size_t)a * L.nP + p];
            q[a] = x[a] * (float)nc[a];
        }
        PointState<D> st;
        init_point<D>(st, q, start_cell<D>(x, nc, tab, g), L.t_end);
        int* out = cells + ((size_t)th * L.nP + p) * max_segs;
        int n = 0;
        float seg_t = 0.f;
        while (point_active<D>(st)) {
            int cell, facet;
            float h;
            if constexpr (D == 1) {
                Chunk1 ch;
                advance1(st, tab, g.nc0, ch);
                cell = ch.cell;
                facet = ch.facet;
                h = ch.h;
            } else {
                Chunk<D> ch;
                advance<D>(st, tab, g.tol0, ch);
                cell = ch.cell;
                facet = ch.facet;
                h = ch.h;
            }
            seg_t += h;
            if (facet >= 0) {
                if (seg_t > 0.f) {
                    if (n < max_segs) out[n] = cell;
                    ++n;
                }
                seg_t = 0.f;
            } else if (!point_active<D>(st) && seg_t > 0.f) {
                if (n < max_segs) out[n] = cell;
                ++n;
            }
        }
        for (int k = n; k < max_segs; ++k) out[k] = -1;
        nseg[(size_t)th * L.nP + p] = n;
    }
}

// meshgrid -> integrate -> interpolate.  saved (nullable): [n_theta, rows, nP]
// holds the warped point (x-space), the q-space Jacobian and, in 2D/3D, the
// log of the first NREC chunks (durations, then the packed facet code) that
// the backward replays instead of searching again; rows = D + D*D (+ NREC + 1).
// One tile (tile_t) of one instance's fused forward; data / out / sv point at
// the instance's slices.
template <int D, bool SAVE>
__device__ __forceinline__ void fused_fwd_tile(const Launch& L, const Geo& g, const Src& s, const Img& im,
                                               const CellTable<D>& tab, int tile, const float* __restrict__ data,
                                               float* __restrict__ out, float* __restrict__ sv) {
    const int nc[3] = {g.nc0, g.nc1, g.nc2};
    const float rnc[3] = {1.0f / (float)g.nc0, 1.0f / (float)g.nc1, 1.0f / (float)g.nc2};  // q -> x
    const int goff[3] = {0, s.g0, s.g0 + s.g1};
    const size_t dstride = (size_t)im.s0 * (D > 1 ? im.s1 : 1) * (D > 2 ? im.s2 : 1);
    // points that used up the chunk budget (NaN) in this wave, counted once
    // at the end of the tile: a wave-uniform sum, no per-lane state
    int wave_bad = 0;
    for (int r = 0; r < L.ppt; ++r) {
        const int t = tile_t<D>(L, s, tile, r);
        if (t < 0) continue;
        int ix[3];
        out_to_point<D>(s, t, ix);
        float x[D], q[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            x[a] = s.lin[goff[a] + ix[a]];
            q[a] = x[a] * (float)nc[a];
        }
        const int c0 = start_cell<D>(x, nc, tab, g);
        float J[D][D];
        // with saved state (2D/3D) the chunk durations go straight to their rows
        LogSink lg;
        if constexpr (SAVE && D > 1) lg.init(sv + (size_t)(D + D * D) * L.nP, L.nP, t);
        integrate<D, SAVE, SAVE && (D > 1), LogSink>(q, c0, L.t_end, tab, g.nc0, g.tol0, J, lg);
        float y[D];
#pragma unroll
        for (int a = 0; a < D; ++a) y[a] = q[a] * rnc[a];
        if constexpr (SAVE) {
            // indexed in output order t (the order threads run in) for coalesced stores
#pragma unroll
            for (int a = 0; a < D; ++a) sv[(size_t)a * L.nP + t] = y[a];
#pragma unroll
            for (int r2 = 0; r2 < D; ++r2)
#pragma unroll
                for (int k = 0; k < D; ++k) sv[(size_t)(D + r2 * D + k) * L.nP + t] = J[r2][k];
            if constexpr (D > 1) sv[(size_t)(D + D * D + NREC) * L.nP + t] = __int_as_float(lg.code);
        }
        int i0[3], i1[3];
        float fd[3];
        corners<D>(y, im, i0, i1, fd);
        interp_channels<D>(data, dstride, im, i0, i1, fd, out + t, L.nP);
        wave_bad += __popcll(__ballot(y[0] != y[0]));
    }
    if (wave_bad && g.unfinished && __lane_id() == 0) atomicAdd(g.unfinished, wave_bad);
}

template <int D, bool SAVE>
__global__ __launch_bounds__(BLOCK) void k_transform_data(Launch L, Geo g, Src s, Img im, const float* __restrict__ Aq,
                                                          float* __restrict__ out, float* __restrict__ saved) {
    extern __shared__ float smem[];
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int th = lb / L.tiles, tile = lb % L.tiles;
    if (th >= L.n_theta) return;
    const CellTable<D> tab = stage_tables<D>(smem, Aq + (size_t)th * g.nC * D * (D + 1), g);
    const size_t dstride = (size_t)im.s0 * (D > 1 ? im.s1 : 1) * (D > 2 ? im.s2 : 1);
    fused_fwd_tile<D, SAVE>(L, g, s, im, tab, tile, im.data + (size_t)th * im.C * dstride,
                            out + (size_t)th * im.C * L.nP, SAVE ? saved + (size_t)th * saved_rows(D) * L.nP : nullptr);
}

template <int D>
__device__ __forceinline__ void write_partials(float* smem_acc, float* __restrict__ part, int nCP) {
    __syncthreads();
    for (int i = threadIdx.x; i < nCP; i += BLOCK) {
        float s = 0.f;
#pragma unroll
        for (int w = 0; w < WAVES; ++w) s += smem_acc[w * nCP + i];
        part[i] = s;
    }
}

// adjoint of k_transform: grad_out [n_theta, D, nP] (x-space).  jac (nullable)
// is the q-space Jacobian stored by the forward.  Writes per-tile partials of
// dL/dA_q [n_theta, tiles, nC*P] and, if gpts != null, dL/dpoints.
template <int D, bool HAVEJ>
__global__ __launch_bounds__(BLOCK) void k_transform_bwd(Launch L, Geo g, Src s, const float* __restrict__ Aq,
                                                         const float* __restrict__ gout, const float* __restrict__ jac,
                                                         float* __restrict__ part, float* __restrict__ gpts) {
    extern __shared__ float smem[];
    constexpr int P = D * (D + 1);
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int th = lb / L.tiles, tile = lb % L.tiles;
    if (th >= L.n_theta) return;
    const int nC = g.nC;
    float* acc = smem + table_floats<D>(nC + g.nV);
    for (int i = threadIdx.x; i < WAVES * nC * P; i += BLOCK) acc[i] = 0.f;
    const CellTable<D> tab = stage_tables<D>(smem, Aq + (size_t)th * nC * P, g);
    float* wacc = acc + (threadIdx.x >> 6) * nC * P;
    const int nc[3] = {g.nc0, g.nc1, g.nc2};
    const float rnc[3] = {1.0f / (float)g.nc0, 1.0f / (float)g.nc1, 1.0f / (float)g.nc2};  // q -> x
    const float* pts = s.pts + (s.per_theta ? (size_t)th * D * L.nP : 0);
    for (int r = 0; r < L.ppt; ++r) {
        const int p = tile_t<D>(L, s, tile, r);  // point index (2D meshgrid tiles when L.ty > 0)
        const bool valid = p >= 0;
        float q0[D], mu[D];
        int c0 = 0;
        ChunkLog lg;
        lg.code = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            q0[a] = 0.f;
            mu[a] = 0.f;
        }
        if (valid) {
            float x[D];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                x[a] = pts[(size_t)a * L.nP + p];
                q0[a] = x[a] * (float)nc[a];
            }
            c0 = start_cell<D>(x, nc, tab, g);
            float J[D][D];
            if constexpr (HAVEJ) {
                const float* jo = jac + (size_t)th * jac_rows(D) * L.nP;
#pragma unroll
                for (int r2 = 0; r2 < D; ++r2)
#pragma unroll
                    for (int k = 0; k < D; ++k) J[r2][k] = jo[(size_t)(r2 * D + k) * L.nP + p];
                if constexpr (D > 1) {
#pragma unroll
                    for (int k = 0; k < NREC; ++k) lg.h[k] = jo[(size_t)(D * D + k) * L.nP + p];
                    lg.code = __float_as_int(jo[(size_t)(D * D + NREC) * L.nP + p]);
                }
            } else {
                float q[D];
#pragma unroll
                for (int a = 0; a < D; ++a) q[a] = q0[a];
                integrate<D, true, true>(q, c0, L.t_end, tab, g.nc0, g.tol0, J, lg);
            }
            float w[D];
            const float* go = gout + (size_t)th * D * L.nP;
#pragma unroll
            for (int a = 0; a < D; ++a) w[a] = go[(size_t)a * L.nP + p] * rnc[a];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                float acc2 = 0.f;
#pragma unroll
                for (int a = 0; a < D; ++a) acc2 = fmaf(J[a][k], w[a], acc2);
                mu[k] = acc2;
            }
            if (gpts) {
                float* gp = gpts + (size_t)th * D * L.nP;
#pragma unroll
                for (int a = 0; a < D; ++a) gp[(size_t)a * L.nP + p] = mu[a] * (float)nc[a];
            }
        }
        pass2<D>(valid, q0, c0, L.t_end, mu, tab, g, wacc, lg);
    }
    write_partials<D>(acc, part + ((size_t)th * L.tiles + tile) * nC * P, nC * P);
}

// adjoint of k_transform_data.  gout [n_theta, C, outsize] in output order;
// saved (nullable) from the forward.  im.gdata (nullable, zero-filled by the
// caller) receives dL/ddata by atomic scatter.
// One tile of one instance's fused backward; adds dL/dA_q of the tile's points
// into this wave's LDS accumulator wacc (all lanes of the block take part).
//
// dL/ddata of a 2D tile with saved state (C <= 4) goes through LDS when box
// (box_cap words, nullable) holds the tile's footprint bins and point slots:
// each point is binned by the data cell of its lower corner (one integer LDS
// atomic in a counting pass before the integration), the bin counts are
// scanned into offsets, every point writes its corner weights and grad_out
// straight to its slot in bin order, and every cell of the footprint's
// bounding box gathers the weighted grad_out of the points in the four bins
// that touch it and adds it to dL/ddata with one global atomic per channel --
// instead of four atomics per point and channel.
template <int D, bool HAVES>
__device__ __forceinline__ void fused_bwd_tile(const Launch& L, const Geo& g, const Src& s, const Img& im,
                                               const CellTable<D>& tab, int tile, const float* __restrict__ data,
                                               float* __restrict__ gdata, const float* __restrict__ gout,
                                               const float* __restrict__ sv, float* wacc, float* box, int box_cap) {
    const int nc[3] = {g.nc0, g.nc1, g.nc2};
    const float rnc[3] = {1.0f / (float)g.nc0, 1.0f / (float)g.nc1, 1.0f / (float)g.nc2};  // q -> x
    const int goff[3] = {0, s.g0, s.g0 + s.g1};
    const int sz[3] = {im.s0, im.s1, im.s2};
    const size_t dstride = (size_t)im.s0 * (D > 1 ? im.s1 : 1) * (D > 2 ? im.s2 : 1);
    // footprint box [bb0, bb1] x [bb2, bb3] of the tile (2D, saved state only)
    __shared__ int s_bb[4];
    bool use_box = false;
    int bx0 = 0, by0 = 0, bh = 1, cells = 0;
    const int nslot = L.ppt * BLOCK;
    int* code = reinterpret_cast<int*>(box);  // [nslot] lower corner i0 << 15 | j0, then bin | rank << 11; -1: none
    int* bins = code + nslot;                 // [cells + 1] counts, then offsets
    float2* fd2 = nullptr;                    // [nslot] corner weights (0 where a corner is clamped), bin order
    float4* gv4 = nullptr;                    // [nslot] grad_out of the channels, bin order
    if constexpr (D == 2 && HAVES) {
        // uniform over the block; corner indices must fit the 15-bit fields of code
        if (box && gdata && nslot <= box_cap && im.s0 <= 32768 && im.s1 <= 32768) {
            if (threadIdx.x == 0) {
                s_bb[0] = INT_MAX;
                s_bb[1] = INT_MIN;
                s_bb[2] = INT_MAX;
                s_bb[3] = INT_MIN;
            }
            __syncthreads();
            int lo0 = INT_MAX, hi0 = INT_MIN, lo1 = INT_MAX, hi1 = INT_MIN;
            for (int r = 0; r < L.ppt; ++r) {
                const int t = tile_t<D>(L, s, tile, r);
                code[r * BLOCK + threadIdx.x] = -1;
                if (t < 0) continue;
                const float y[2] = {sv[t], sv[(size_t)L.nP + t]};
                int i0[3], i1[3];
                float fd[3];
                corners<2>(y, im, i0, i1, fd);
                code[r * BLOCK + threadIdx.x] = i0[0] << 15 | i0[1];
                lo0 = min(lo0, i0[0]);
                hi0 = max(hi0, i1[0]);
                lo1 = min(lo1, i0[1]);
                hi1 = max(hi1, i1[1]);
            }
            lo0 = wave_min(lo0);
            hi0 = wave_max(hi0);
            lo1 = wave_min(lo1);
            hi1 = wave_max(hi1);
            if (__lane_id() == 0 && lo0 != INT_MAX) {
                atomicMin(&s_bb[0], lo0);
                atomicMax(&s_bb[1], hi0);
                atomicMin(&s_bb[2], lo1);
                atomicMax(&s_bb[3], hi1);
            }
            __syncthreads();
            bx0 = s_bb[0];
            by0 = s_bb[2];
            bh = s_bb[3] - by0 + 1;
            cells = (s_bb[1] - bx0 + 1) * bh;
            const long long slot0 = ((long long)nslot + cells + 1 + 3) & ~3LL;  // 16-byte aligned
            use_box = bx0 != INT_MAX && im.C <= 4 && cells < BIN_CELLS && slot0 + 6LL * nslot <= box_cap;
            __syncthreads();  // s_bb is reused by the next tile of a persistent block
            if (use_box) {
                fd2 = reinterpret_cast<float2*>(box + slot0);
                gv4 = reinterpret_cast<float4*>(fd2 + nslot);
                for (int i = threadIdx.x; i <= cells; i += BLOCK) bins[i] = 0;
                __syncthreads();
                // count the bins; each thread re-codes its own slots
                for (int sl = threadIdx.x; sl < nslot; sl += BLOCK) {
                    const int cd = code[sl];
                    if (cd < 0) continue;
                    const int k = ((cd >> 15) - bx0) * bh + ((cd & 0x7FFF) - by0);
                    const int rank = atomicAdd(&bins[k], 1);
                    code[sl] = k | (rank << 11);
                }
                __syncthreads();
                block_exclusive_scan(bins, cells + 1);
            }
        }
    }
    for (int r = 0; r < L.ppt; ++r) {
        const int t = tile_t<D>(L, s, tile, r);
        const bool valid = t >= 0;
        float q0[D], mu[D];
        int c0 = 0;
        ChunkLog lg;
        lg.code = 0;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            q0[a] = 0.f;
            mu[a] = 0.f;
        }
        if (valid) {
            int ix[3];
            out_to_point<D>(s, t, ix);
            float x[D];
#pragma unroll
            for (int a = 0; a < D; ++a) {
                x[a] = s.lin[goff[a] + ix[a]];
                q0[a] = x[a] * (float)nc[a];
            }
            c0 = start_cell<D>(x, nc, tab, g);
            float J[D][D], y[D];
            if constexpr (HAVES) {
#pragma unroll
                for (int a = 0; a < D; ++a) y[a] = sv[(size_t)a * L.nP + t];
#pragma unroll
                for (int r2 = 0; r2 < D; ++r2)
#pragma unroll
                    for (int k = 0; k < D; ++k) J[r2][k] = sv[(size_t)(D + r2 * D + k) * L.nP + t];
                if constexpr (D > 1) {
#pragma unroll
                    for (int k = 0; k < NREC; ++k) lg.h[k] = sv[(size_t)(D + D * D + k) * L.nP + t];
                    lg.code = __float_as_int(sv[(size_t)(D + D * D + NREC) * L.nP + t]);
                }
            } else {
                float q[D];
#pragma unroll
                for (int a = 0; a < D; ++a) q[a] = q0[a];
                integrate<D, true, true>(q, c0, L.t_end, tab, g.nc0, g.tol0, J, lg);
#pragma unroll
                for (int a = 0; a < D; ++a) y[a] = q[a] * rnc[a];
            }
            int i0[3], i1[3];
            float fd[3], dfd[3] = {0.f, 0.f, 0.f};
            corners<D>(y, im, i0, i1, fd);
            const int slot = r * BLOCK + threadIdx.x;
            float gk[4] = {0.f, 0.f, 0.f, 0.f};
            if constexpr (D == 2) {
                interp_backward_channels<D>(data, (gdata && !use_box) ? gdata : nullptr, dstride, im, i0, i1, fd,
                                            gout + t, L.nP, gk, dfd);
            } else {
                for (int c = 0; c < im.C; ++c) {
                    const size_t base = (size_t)c * dstride;
                    const float gv = gout[(size_t)c * L.nP + t];
                    interp_backward<D>(data + base, gdata ? gdata + base : nullptr, im, i0, i1, fd, gv, dfd);
                    if (c < 4) gk[c] = gv;
                }
            }
            if constexpr (D == 2) {
                if (use_box) {
                    const int cd = code[slot];
                    const int pos = bins[cd & 2047] + (cd >> 11);
                    // a clamped corner pair (i1 == i0) puts its whole weight on i0
                    fd2[pos] = make_float2(i1[0] != i0[0] ? fd[0] : 0.f, i1[1] != i0[1] ? fd[1] : 0.f);
                    gv4[pos] = make_float4(gk[0], gk[1], gk[2], gk[3]);
                }
            }
            float w[D];
#pragma unroll
            for (int a = 0; a < D; ++a) w[a] = dfd[a] * (float)(sz[a] - 1) * rnc[a];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                float acc2 = 0.f;
#pragma unroll
                for (int a = 0; a < D; ++a) acc2 = fmaf(J[a][k], w[a], acc2);
                mu[k] = acc2;
            }
        }
        pass2<D>(valid, q0, c0, L.t_end, mu, tab, g, wacc, lg);
    }
    if (use_box) {
        __syncthreads();
        // Cell (i, j) is a corner of the points whose lower corner is in bins
        // (i - a, j - b), a, b in {0, 1}.  Bins (i - a, j - 1) and (i - a, j)
        // are adjacent in bin order, so the four bins are two ranges, walked
        // as one loop (its trip count is the wave's largest total, not the
        // sum of four per-bin maxima): row a = 0 is [s0, e0) with bin
        // (i, j - 1) below m0, row a = 1 is [s1, e1) split at m1.
        const float rbh = 1.0f / (float)bh;
        for (int k = threadIdx.x; k < cells; k += BLOCK) {
            int j;
            const int i = div_small(k, bh, rbh, j);
            const int jb = j > 0 ? 1 : 0;
            const int s0 = bins[k - jb], m0 = bins[k], e0 = bins[k + 1];
            int s1 = 0, m1 = 0, e1 = 0;
            if (i > 0) {
                s1 = bins[k - bh - jb];
                m1 = bins[k - bh];
                e1 = bins[k - bh + 1];
            }
            const int n0 = e0 - s0, n = n0 + (e1 - s1);
            float acc[4] = {0.f, 0.f, 0.f, 0.f};
            for (int u = 0; u < n; ++u) {
                const bool row0 = u < n0;
                const int e = row0 ? s0 + u : s1 + (u - n0);
                const float2 f = fd2[e];
                const float4 gq = gv4[e];
                const float wx = row0 ? 1.f - f.x : f.x;
                const float wy = e < (row0 ? m0 : m1) ? f.y : 1.f - f.y;
                const float w = wx * wy;
                acc[0] = fmaf(w, gq.x, acc[0]);
                acc[1] = fmaf(w, gq.y, acc[1]);
                acc[2] = fmaf(w, gq.z, acc[2]);
                acc[3] = fmaf(w, gq.w, acc[3]);
            }
#pragma unroll
            for (int c = 0; c < 4; ++c)
                if (c < im.C && acc[c] != 0.f)
                    atomicAdd(gdata + (size_t)c * dstride + (size_t)(bx0 + i) * im.s1 + (by0 + j), acc[c]);
        }
    }
}

// Blocks per theta of the fused backward with tpb tiles per block (tpb > 1
// only on 2D tiled grids): a block takes tpb vertically adjacent tiles of
// one tile column, so horizontally adjacent tiles -- which share the cache
// lines of the output-order rows (16 columns = half a 128-byte line) -- run
// in neighbouring blocks at the same time instead of one after the other in
// one block (which fetched every such line twice).
__host__ __device__ inline int bwd_blocks_per_theta(int tiles, int g1, int ty, int tpb) {
    if (tpb == 1) return tiles;
    const int tiles_y = (g1 + ty - 1) / ty, tiles_x = tiles / tiles_y;
    return ((tiles_x + tpb - 1) / tpb) * tiles_y;
}

// adjoint of k_transform_data.  gout [n_theta, C, outsize] in output order;
// saved (nullable) from the forward.  im.gdata (nullable, zero-filled by the
// caller) receives dL/ddata by atomic scatter.  Each block runs tpb tiles of
// one theta (bwd_blocks_per_theta) into one row of partials, so the tables
// are staged and the partials written once per tpb tiles.
template <int D, bool HAVES>
__global__ __launch_bounds__(BLOCK, (D == 3 && HAVES) ? 4 : 1) void k_transform_data_bwd(Launch L, Geo g, Src s, Img im, const float* __restrict__ Aq,
                                                              const float* __restrict__ gout,
                                                              const float* __restrict__ saved, float* __restrict__ part,
                                                              int box_cap, int tpb) {
    extern __shared__ float smem[];
    constexpr int P = D * (D + 1);
    const int nbt = bwd_blocks_per_theta(L.tiles, s.g1, L.ty, tpb);
    const int lb = xcd_remap(blockIdx.x, gridDim.x);
    const int th = lb / nbt, bt = lb % nbt;
    if (th >= L.n_theta) return;
    const int tiles_y = tpb > 1 ? (s.g1 + L.ty - 1) / L.ty : 1, tiles_x = L.tiles / tiles_y;
    const int nC = g.nC;
    float* acc = smem + table_floats<D>(nC + g.nV);
    for (int i = threadIdx.x; i < WAVES * nC * P; i += BLOCK) acc[i] = 0.f;
    const CellTable<D> tab = stage_tables<D>(smem, Aq + (size_t)th * nC * P, g);
    const size_t dstride = (size_t)im.s0 * (D > 1 ? im.s1 : 1) * (D > 2 ? im.s2 : 1);
    for (int k = 0; k < tpb; ++k) {
        // tile row (bt / tiles_y) * tpb + k of tile column bt % tiles_y
        const int trow = (bt / tiles_y) * tpb + k;
        if (tpb > 1 && trow >= tiles_x) break;
        const int tile = tpb > 1 ? trow * tiles_y + bt % tiles_y : bt;
        fused_bwd_tile<D, HAVES>(L, g, s, im, tab, tile, im.data + (size_t)th * im.C * dstride,
                                 im.gdata ? im.gdata + (size_t)th * im.C * dstride : nullptr,
                                 gout + (size_t)th * im.C * L.nP,
                                 HAVES ? saved + (size_t)th * saved_rows(D) * L.nP : nullptr,
                                 acc + (threadIdx.x >> 6) * nC * P, box_cap ? acc + WAVES * nC * P : nullptr, box_cap);
    }
    write_partials<D>(acc, part + ((size_t)th * nbt + bt) * nC * P, nC * P);
}

// interpolation at an explicit grid [N, D, nP]; out [N, C, outsize], output order
// The (instance n, output index t, grid index p) of this thread.  2D outputs
// run in 16 x 16 blocks of 8 x 8 wave tiles, so the output accesses (last
// axis fastest) and the grid accesses (first axis fastest) both come in runs
// of 8 lanes; blocks: N * ceil(g0 / 16) * ceil(g1 / 16).  Other
// dimensions run in output order.
template <int D>
__device__ __forceinline__ bool interp_item(const Src& s, int N, int nP, int& n, int& t, int& p) {
    if constexpr (D == 2) {
        const int t0 = (s.g0 + 15) >> 4, t1 = (s.g1 + 15) >> 4;
        const int b = blockIdx.x;
        n = b / (t0 * t1);
        const int rem = b - n * t0 * t1;
        const int tid = threadIdx.x;
        const int i0 = (rem / t1) * 16 + ((tid & 63) >> 3) + ((tid >> 7) << 3);
        const int i1 = (rem % t1) * 16 + (tid & 7) + (((tid >> 6) & 1) << 3);
        if (n >= N || i0 >= s.g0 || i1 >= s.g1) return false;
        t = i0 * s.g1 + i1;
        p = i0 + s.g0 * i1;
        return true;
    } else {
        const size_t gid = (size_t)blockIdx.x * BLOCK + threadIdx.x;
        if (gid >= (size_t)N * nP) return false;
        n = (int)(gid / nP);
        t = (int)(gid % nP);
        int ix[3];
        p = out_to_point<D>(s, t, ix);
        return true;
    }
}

template <int D>
__global__ __launch_bounds__(BLOCK) void k_interp_fwd(int N, int nP, Src s, Img im, float* __restrict__ out) {
    int n, t, p;
    if (!interp_item<D>(s, N, nP, n, t, p)) return;
    float y[D];
#pragma unroll
    for (int a = 0; a < D; ++a) y[a] = s.pts[((size_t)n * D + a) * nP + p];
    int i0[3], i1[3];
    float fd[3];
    corners<D>(y, im, i0, i1, fd);
    const size_t dstride = (size_t)im.s0 * (D > 1 ? im.s1 : 1) * (D > 2 ? im.s2 : 1);
    for (int c = 0; c < im.C; ++c)
        out[((size_t)n * im.C + c) * nP + t] = interp_value<D>(im.data + ((size_t)n * im.C + c) * dstride, im, i0, i1, fd);
}

template <int D>
__global__ __launch_bounds__(BLOCK) void k_interp_bwd(int N, int nP, Src s, Img im, const float* __restrict__ gout,
                                                      float* __restrict__ ggrid) {
    int n, t, p;
    if (!interp_item<D>(s, N, nP, n, t, p)) return;
    float y[D];
#pragma unroll
    for (int a = 0; a < D; ++a) y[a] = s.pts[((size_t)n * D + a) * nP + p];
    int i0[3], i1[3];
    float fd[3], dfd[3] = {0.f, 0.f, 0.f};
    corners<D>(y, im, i0, i1, fd);
    const size_t dstride = (size_t)im.s0 * (D > 1 ? im.s1 : 1) * (D > 2 ? im.s2 : 1);
    for (int c = 0; c < im.C; ++c) {
        const size_t base = ((size_t)n * im.C + c) * dstride;
        interp_backward<D>(im.data + base, im.gdata ? im.gdata + base : nullptr, im, i0, i1, fd,
                           gout[((size_t)n * im.C + c) * nP + t], dfd);
    }
    if (ggrid) {
        const int sz[3] = {im.s0, im.s1, im.s2};
#pragma unroll
        for (int a = 0; a < D; ++a) ggrid[((size_t)n * D + a) * nP + p] = dfd[a] * (float)(sz[a] - 1);
    }
}

// per theta: dL/dA (x-space) = sum over tiles of the q-space partials, scaled
// back; dL/dtheta = B^T dL/dA.  One block of nw = blockDim.x / 64 waves per
// theta (memory-bound: many waves keep loads in flight); every sum runs in a
// fixed order.
constexpr int RBLOCK = 1024;
// dL/dtheta of one theta from its tile partials pp [tiles, Dt] (q-space
// dL/dA): the block's waves sum the tiles, the wave sums are added in wave
// order, scaled to x-space (gA, nullable, receives dL/dA), and B^T dL/dA is
// split over the waves by rows of B and added in wave order.  sg holds
// [nw + 1, Dt] floats.
// Row j of the partials that reduce_theta sums (j < its row count).
struct AllRows {
    __device__ __forceinline__ int operator()(int j) const { return j; }
};

template <class Rows = AllRows>
__device__ __forceinline__ void reduce_theta(float* sg, const float* __restrict__ pp, int tiles, int Dt, int ndim,
                                             const int (&nc)[3], const float* __restrict__ B, int d,
                                             float* __restrict__ gth, float* __restrict__ gA, Rows rows = Rows()) {
    const int nw = blockDim.x >> 6;
    const int P = ndim * (ndim + 1);
    // wave w sums tiles w, w + nw, ... (two chains for overlap)
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int i = lane; i < Dt; i += 64) {
        float s0 = 0.f, s1 = 0.f;
        int t = w;
        for (; t + nw < tiles; t += 2 * nw) {
            s0 += pp[(size_t)rows(t) * Dt + i];
            s1 += pp[(size_t)rows(t + nw) * Dt + i];
        }
        if (t < tiles) s0 += pp[(size_t)rows(t) * Dt + i];
        sg[w * Dt + i] = s0 + s1;
    }
    __syncthreads();
    float* sA = sg + nw * Dt;
    for (int i = threadIdx.x; i < Dt; i += blockDim.x) {
        float s = sg[i];
        for (int k = 1; k < nw; ++k) s += sg[k * Dt + i];
        const int rem = i % P, r = rem / (ndim + 1), col = rem % (ndim + 1);
        const float sc = (col < ndim) ? (float)nc[r] / (float)nc[col] : (float)nc[r];
        s *= sc;
        sA[i] = s;
        if (gA) gA[i] = s;
    }
    __syncthreads();
    // B^T dL/dA: wave w takes rows [i0, i1) of B, the wave partials (reusing
    // sg[0, nw*d), d <= Dt) are added in wave order
    const int i0 = (int)((long long)Dt * w / nw), i1 = (int)((long long)Dt * (w + 1) / nw);
    for (int k = lane; k < d; k += 64) {
        float s = 0.f;
#pragma unroll 8
        for (int i = i0; i < i1; ++i) s = fmaf(B[(size_t)i * d + k], sA[i], s);
        sg[w * d + k] = s;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < d; k += blockDim.x) {
        float s = sg[k];
        for (int w2 = 1; w2 < nw; ++w2) s += sg[w2 * d + k];
        gth[k] = s;
    }
}

// Both steps in one block per theta, for launches with enough thetas to fill
// the chip by themselves.
__global__ __launch_bounds__(RBLOCK) void k_reduce_theta(const float* __restrict__ part, int tiles, int Dt, int ndim,
                                                         int nc0, int nc1, int nc2, const float* __restrict__ B, int d,
                                                         float* __restrict__ gtheta, float* __restrict__ gA) {
    extern __shared__ float sg[];  // [nw + 1, Dt]
    const int th = blockIdx.x;
    const int nc[3] = {nc0, nc1, nc2};
    reduce_theta(sg, part + (size_t)th * tiles * Dt, tiles, Dt, ndim, nc, B, d, gtheta + (size_t)th * d,
                 gA ? gA + (size_t)th * Dt : nullptr);
}

// Column slice blockIdx.y (columns [c0, c0 + cw)) of theta blockIdx.x's
// dL/dA: the block's waves sum the tiles (wave w: tiles w, w + nw, ...), the
// wave sums are added in wave order and scaled to x-space -- per column the
// same arithmetic as reduce_theta.  The result overwrites the tile-0 row of
// the partials (each block owns its columns and has read them all first) and
// goes to gA when given.  Slicing the columns gives a few thetas enough
// blocks to fill the chip.
__global__ __launch_bounds__(RBLOCK) void k_reduce_cols(float* __restrict__ part, int tiles, int Dt, int cw, int ndim,
                                                        int nc0, int nc1, int nc2, float* __restrict__ gA) {
    extern __shared__ float sg[];  // [nw, cw]
    const int th = blockIdx.x;
    const int c0 = blockIdx.y * cw, c1 = (c0 + cw < Dt) ? c0 + cw : Dt;
    const int nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int P = ndim * (ndim + 1);
    const int nc[3] = {nc0, nc1, nc2};
    float* pp = part + (size_t)th * tiles * Dt;
    for (int i = c0 + lane; i < c1; i += 64) {
        float s0 = 0.f, s1 = 0.f;
        int t = w;
        for (; t + nw < tiles; t += 2 * nw) {
            s0 += pp[(size_t)t * Dt + i];
            s1 += pp[(size_t)(t + nw) * Dt + i];
        }
        if (t < tiles) s0 += pp[(size_t)t * Dt + i];
        sg[w * cw + (i - c0)] = s0 + s1;
    }
    __syncthreads();
    for (int i = c0 + threadIdx.x; i < c1; i += blockDim.x) {
        float s = sg[i - c0];
        for (int k = 1; k < nw; ++k) s += sg[k * cw + (i - c0)];
        const int rem = i % P, r = rem / (ndim + 1), col = rem % (ndim + 1);
        const float sc = (col < ndim) ? (float)nc[r] / (float)nc[col] : (float)nc[r];
        s *= sc;
        pp[i] = s;
        if (gA) gA[(size_t)th * Dt + i] = s;
    }
}

// dL/dtheta = B^T dL/dA for theta blockIdx.x, dL/dA (x-space) in the tile-0
// row of its partials: wave w takes rows [i0, i1) of B, the wave partials are
// added in wave order.
__global__ __launch_bounds__(RBLOCK) void k_reduce_bt(const float* __restrict__ part, int tiles, int Dt,
                                                      const float* __restrict__ B, int d, float* __restrict__ gtheta) {
    extern __shared__ float sg[];  // [Dt + nw * d]
    const int th = blockIdx.x;
    const int nw = blockDim.x >> 6, w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const float* pp = part + (size_t)th * tiles * Dt;
    float* sA = sg;
    float* sw = sg + Dt;
    for (int i = threadIdx.x; i < Dt; i += blockDim.x) sA[i] = pp[i];
    __syncthreads();
    const int i0 = (int)((long long)Dt * w / nw), i1 = (int)((long long)Dt * (w + 1) / nw);
    for (int k = lane; k < d; k += 64) {
        float s = 0.f;
#pragma unroll 8
        for (int i = i0; i < i1; ++i) s = fmaf(B[(size_t)i * d + k], sA[i], s);
        sw[w * d + k] = s;
    }
    __syncthreads();
    for (int k = threadIdx.x; k < d; k += blockDim.x) {
        float s = sw[k];
        for (int w2 = 1; w2 < nw; ++w2) s += sw[w2 * d + k];
        gtheta[(size_t)th * d + k] = s;
    }
}

// cell index of each point (libcpab/numpy/findcellidx.py semantics, FP32)
template <int D>
__global__ __launch_bounds__(BLOCK) void k_findcell(int nc0, int nc1, int nc2, const float* __restrict__ pts, int nP,
                                                    int* __restrict__ idx) {
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= nP) return;
    const int nc[3] = {nc0, nc1, nc2};
    float x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = pts[(size_t)a * nP + p];
    idx[p] = find_cell<D>(x, nc);
}

// velocity of one theta's field at each point, x-space (libcpab/pytorch/functions.py:111-129)
template <int D>
__global__ __launch_bounds__(BLOCK) void k_vectorfield(Geo g, const float* __restrict__ Aq, const float* __restrict__ pts,
                                                       int nP, float* __restrict__ v) {
    const int p = blockIdx.x * BLOCK + threadIdx.x;
    if (p >= nP) return;
    const int nc[3] = {g.nc0, g.nc1, g.nc2};
    const float rnc[3] = {1.0f / (float)g.nc0, 1.0f / (float)g.nc1, 1.0f / (float)g.nc2};  // q -> x
    float x[D], q[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        x[a] = pts[(size_t)a * nP + p];
        q[a] = x[a] * (float)nc[a];
    }
    const int c = find_cell<D>(x, nc);
    const float* A = Aq + c * D * (D + 1);
#pragma unroll
    for (int r = 0; r < D; ++r) {
        float acc = A[r * (D + 1) + D];
#pragma unroll
        for (int k = 0; k < D; ++k) acc = fmaf(A[r * (D + 1) + k], q[k], acc);
        v[(size_t)r * nP + p] = acc * rnc[r];
    }
}

// ---------------------------------------------------------------------------
// ragged batches: persistent workers over (instance, tile) items
// ---------------------------------------------------------------------------
__host__ __device__ inline size_t align256_hd(size_t b) { return (b + 255) & ~(size_t)255; }

__device__ __forceinline__ void item_args(const cpab_instance_t* I, int ppt, float t_end, Launch& L, Geo& g, Src& s,
                                          Img& im) {
    const int nd = I->tess.ndim;
    L.n_theta = 1;
    L.nP = I->nP;
    L.tiles = I->tiles;
    L.ppt = ppt;
    L.ty = tile_ty(nd, nd > 1 ? I->outsize[1] : 1, nd > 2 ? I->outsize[2] : 1);
    L.t_end = t_end;
    g.ndim = nd;
    g.nc0 = I->tess.nc[0];
    g.nc1 = nd > 1 ? I->tess.nc[1] : 1;
    g.nc2 = nd > 2 ? I->tess.nc[2] : 1;
    g.nC = I->tess.nC;
    g.pl = reinterpret_cast<const float4*>(I->tess.planes);
    g.nb = I->tess.nbr;
    g.unfinished = I->tess.unfinished;
    g.nV = I->tess.nV;
    g.amap = I->tess.amap;
    int m = g.nc0 > g.nc1 ? g.nc0 : g.nc1;
    m = m > g.nc2 ? m : g.nc2;
    g.tol0 = 2e-6f * (float)m;
    s.pts = nullptr;
    s.per_theta = 0;
    s.lin = I->lin;
    s.g0 = I->outsize[0];
    s.g1 = nd > 1 ? I->outsize[1] : 1;
    s.g2 = nd > 2 ? I->outsize[2] : 1;
    im.data = I->data;
    im.gdata = I->grad_data;
    im.C = I->C;
    im.s0 = I->dsize[0];
    im.s1 = nd > 1 ? I->dsize[1] : 1;
    im.s2 = nd > 2 ? I->dsize[2] : 1;
}

// A block takes GRAB consecutive work items per atomic: consecutive items
// mostly belong to one instance, whose cell tables are then staged once for
// all of them.
constexpr int GRAB = 4;
// The backward's LDS (tables + per-wave accumulators + dL/ddata box) is sized
// for the largest tessellation of its launch; instances are grouped into
// classes that fit 4, 3 and 2 blocks per CU, one launch per class.
constexpr int BATCH_CLASSES = 3;

__device__ __forceinline__ int next_grab(int* counter, int* s_item) {
    if (threadIdx.x == 0) *s_item = atomicAdd(counter, GRAB);
    __syncthreads();
    const int item = __builtin_amdgcn_readfirstlane(*s_item);
    __syncthreads();
    return item;
}

// the work queues at counter[0], counter[16] and counter[32] (one per LDS
// class of the backward, a cache line apart)
__global__ void k_counter_reset(int* counter) {
    if (blockIdx.x == 0 && threadIdx.x < BATCH_CLASSES) counter[threadIdx.x * 16] = 0;
}

__global__ __launch_bounds__(BLOCK) void k_fill_zero(float* p, size_t n) {
    for (size_t i = (size_t)blockIdx.x * BLOCK + threadIdx.x; i < n; i += (size_t)gridDim.x * BLOCK) p[i] = 0.f;
}

// A_q = B theta for every instance: grid (ceil(maxDt/BLOCK), n_inst)
__global__ __launch_bounds__(BLOCK) void k_batch_A(const cpab_instance_t* __restrict__ inst) {
    extern __shared__ float sth[];
    const cpab_instance_t* I = inst + blockIdx.y;
    const int nd = I->tess.ndim, d = I->tess.d;
    const int P = nd * (nd + 1);
    const int Dt = I->tess.nC * P;
    if ((int)(blockIdx.x * BLOCK) >= Dt) return;
    for (int k = threadIdx.x; k < d; k += BLOCK) sth[k] = I->theta[k];
    __syncthreads();
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= Dt) return;
    const float* b = I->tess.B + (size_t)i * d;
    float acc = 0.f;
    for (int k = 0; k < d; ++k) acc = fmaf(b[k], sth[k], acc);
    const int rem = i % P, r = rem / (nd + 1), col = rem % (nd + 1);
    const float sc = (col < nd) ? (float)I->tess.nc[r] / (float)I->tess.nc[col] : (float)I->tess.nc[r];
    static_cast<float*>(I->ws)[i] = acc * sc;
}

// MAXD: the largest ndim among the instances (2 or 3); without 3D instances
// the 3D code is not compiled in, which keeps the register count (and the
// occupancy) of the 1D/2D paths.
template <bool SAVE, int MAXD>
__global__ __launch_bounds__(BLOCK) void k_batch_fwd(const cpab_instance_t* __restrict__ inst,
                                                     const int* __restrict__ item_inst, int n_items, int ppt,
                                                     float t_end, int* counter) {
    extern __shared__ float smem[];
    __shared__ int s_item;
    int cur = -1;
    for (;;) {
        const int first = next_grab(counter, &s_item);
        if (first >= n_items) break;
        const int last = first + GRAB < n_items ? first + GRAB : n_items;
        for (int item = first; item < last; ++item) {
            // the previous item's waves are done with the tables and accumulators
            if (item != first) __syncthreads();
            const int ii = __builtin_amdgcn_readfirstlane(item_inst[item]);
            const cpab_instance_t* I = inst + ii;
            Launch L;
            Geo g;
            Src s;
            Img im;
            item_args(I, ppt, t_end, L, g, s, im);
            const int tile = item - I->item0;
            const bool restage = ii != cur;
            cur = ii;
            const float* Aq = static_cast<const float*>(I->ws);
            switch (g.ndim) {
                case 1: {
                    const CellTable<1> tab = restage ? stage_tables<1>(smem, Aq, g) : table_view<1>(smem, g.nC + g.nV);
                    fused_fwd_tile<1, SAVE>(L, g, s, im, tab, tile, im.data, I->out, I->saved);
                } break;
                case 2: {
                    const CellTable<2> tab = restage ? stage_tables<2>(smem, Aq, g) : table_view<2>(smem, g.nC + g.nV);
                    fused_fwd_tile<2, SAVE>(L, g, s, im, tab, tile, im.data, I->out, I->saved);
                } break;
                default:
                    if constexpr (MAXD == 3) {
                        const CellTable<3> tab = restage ? stage_tables<3>(smem, Aq, g) : table_view<3>(smem, g.nC + g.nV);
                        fused_fwd_tile<3, SAVE>(L, g, s, im, tab, tile, im.data, I->out, I->saved);
                    }
                    break;
            }
        }
    }
}

// One work item of the batched backward.  A run of consecutive items of one
// instance inside one grab accumulates into the same LDS accumulators
// (cleared at the run's first item) and writes one row of partials, at the
// run's first tile, after its last item.
template <int D, bool HAVES>
__device__ __forceinline__ void batch_bwd_item(float* smem, float* acc_base, float* box, int box_cap,
                                               const cpab_instance_t* I, bool restage, bool run_start, bool run_end,
                                               int row, const Launch& L, const Geo& g, const Src& s, const Img& im,
                                               int tile) {
    constexpr int P = D * (D + 1);
    const int nCP = g.nC * P;
    float* acc = acc_base;
    if (run_start)
        for (int i = threadIdx.x; i < WAVES * nCP; i += BLOCK) acc[i] = 0.f;
    const CellTable<D> tab = restage ? stage_tables<D>(smem, static_cast<const float*>(I->ws), g) : table_view<D>(smem, g.nC + g.nV);
    if (!restage) __syncthreads();
    fused_bwd_tile<D, HAVES>(L, g, s, im, tab, tile, im.data, im.gdata, I->grad_out, I->saved,
                             acc + (threadIdx.x >> 6) * nCP, box, box_cap);
    if (run_end) {
        float* part = reinterpret_cast<float*>(static_cast<char*>(I->ws) + align256_hd((size_t)nCP * 4));
        write_partials<D>(acc, part + (size_t)row * nCP, nCP);
    }
}

template <bool HAVES, int MAXD>
// Work items [item_lo, item_hi) (one LDS class of instances, see
// cpab_batch_plan), pulled from their own queue at counter.
__global__ __launch_bounds__(BLOCK, MAXD == 2 ? 4 : 1) void k_batch_bwd(const cpab_instance_t* __restrict__ inst,
                                                     const int* __restrict__ item_inst, int item_lo, int item_hi,
                                                     int ppt, float t_end, int* counter, int acc_offset,
                                                     int box_offset, int box_cap) {
    extern __shared__ float smem[];
    __shared__ int s_item;
    int cur = -1;
    // grabs are the GRAB-aligned item groups [k GRAB, (k + 1) GRAB) that meet
    // [item_lo, item_hi), so the rows of partials an instance gets written
    // (batch_rows) do not depend on the class ranges
    const int base = item_lo & ~(GRAB - 1);
    for (;;) {
        const int g0 = base + next_grab(counter, &s_item);
        if (g0 >= item_hi) break;
        const int first = g0 > item_lo ? g0 : item_lo;
        const int last = g0 + GRAB < item_hi ? g0 + GRAB : item_hi;
        int row = 0;
        for (int item = first; item < last; ++item) {
            // the previous item's waves are done with the tables and accumulators
            if (item != first) __syncthreads();
            const int ii = __builtin_amdgcn_readfirstlane(item_inst[item]);
            const cpab_instance_t* I = inst + ii;
            Launch L;
            Geo g;
            Src s;
            Img im;
            item_args(I, ppt, t_end, L, g, s, im);
            const int tile = item - I->item0;
            const bool restage = ii != cur;
            cur = ii;
            const bool run_start = item == first || tile == 0;
            const bool run_end = item == last - 1 || tile == I->tiles - 1;
            if (run_start) row = tile;
            float* acc = smem + acc_offset;  // after the largest table of the batch
            float* box = box_cap ? smem + box_offset : nullptr;
            switch (g.ndim) {
                case 1:
                    batch_bwd_item<1, HAVES>(smem, acc, box, box_cap, I, restage, run_start, run_end, row, L, g, s, im,
                                             tile);
                    break;
                case 2:
                    batch_bwd_item<2, HAVES>(smem, acc, box, box_cap, I, restage, run_start, run_end, row, L, g, s, im,
                                             tile);
                    break;
                default:
                    if constexpr (MAXD == 3)
                        batch_bwd_item<3, HAVES>(smem, acc, box, box_cap, I, restage, run_start, run_end, row, L, g, s,
                                                 im, tile);
                    break;
            }
        }
    }
}

// The rows of partials the batched backward writes for an instance whose
// items start at item0: the first item's, then one per GRAB-aligned item
// after it -- j = 0, 1, ... maps to tile rows 0, a, a + GRAB, ... with a the
// distance to the next multiple of GRAB (GRAB when item0 is one).
struct RunRows {
    int a;
    __device__ __forceinline__ int operator()(int j) const { return j == 0 ? 0 : a + GRAB * (j - 1); }
};
__device__ __forceinline__ int batch_rows(int item0, int tiles, RunRows& rr) {
    const int r = item0 & (GRAB - 1);
    rr.a = r ? GRAB - r : GRAB;
    return 1 + (rr.a < tiles ? (tiles - 1 - rr.a) / GRAB + 1 : 0);
}

// per instance: dL/dtheta = B^T (sum over its items of the partials, in x-space)
__global__ __launch_bounds__(RBLOCK) void k_batch_reduce(const cpab_instance_t* __restrict__ inst) {
    extern __shared__ float sg[];  // [nw + 1, largest Dt of the batch]
    const cpab_instance_t* I = inst + blockIdx.x;
    const int nd = I->tess.ndim;
    const int Dt = I->tess.nC * nd * (nd + 1);
    const float* part = reinterpret_cast<const float*>(static_cast<const char*>(I->ws) + align256_hd((size_t)Dt * 4));
    const int nc[3] = {I->tess.nc[0], I->tess.nc[1], I->tess.nc[2]};
    RunRows rr;
    const int nrows = batch_rows(I->item0, I->tiles, rr);
    reduce_theta(sg, part, nrows, Dt, nd, nc, I->tess.B, I->tess.d, I->grad_theta, nullptr, rr);
}

// ---------------------------------------------------------------------------
// method='numeric': the reference's fixed-step scheme
// ---------------------------------------------------------------------------
// expm(scale * A~) for n affine matrices A [D, D+1] (homogeneous last row 0),
// FP64 scaling and squaring with a degree-12 Taylor polynomial; top D rows out.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_expm_affine(const float* __restrict__ A, int n, float scale,
                                                       float* __restrict__ T) {
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= n) return;
    constexpr int E = D + 1;
    double X[E][E], R[E][E], Pw[E][E], tmp[E][E];
    double nrm = 0.0;
#pragma unroll
    for (int r = 0; r < E; ++r)
#pragma unroll
        for (int c = 0; c < E; ++c) {
            X[r][c] = (r < D) ? (double)scale * (double)A[(size_t)i * D * E + r * E + c] : 0.0;
            nrm += fabs(X[r][c]);
        }
    int sq = 0;
    while (nrm > 0.25 && sq < 60) {
        nrm *= 0.5;
        ++sq;
    }
    const double f = ldexp(1.0, -sq);
#pragma unroll
    for (int r = 0; r < E; ++r)
#pragma unroll
        for (int c = 0; c < E; ++c) {
            X[r][c] *= f;
            R[r][c] = (r == c) ? 1.0 : 0.0;
            Pw[r][c] = (r == c) ? 1.0 : 0.0;
        }
    for (int k = 1; k <= 12; ++k) {
#pragma unroll
        for (int r = 0; r < E; ++r)
#pragma unroll
            for (int c = 0; c < E; ++c) {
                double acc = 0.0;
#pragma unroll
                for (int m = 0; m < E; ++m) acc += Pw[r][m] * X[m][c];
                tmp[r][c] = acc / k;
            }
#pragma unroll
        for (int r = 0; r < E; ++r)
#pragma unroll
            for (int c = 0; c < E; ++c) {
                Pw[r][c] = tmp[r][c];
                R[r][c] += tmp[r][c];
            }
    }
    for (int k = 0; k < sq; ++k) {
#pragma unroll
        for (int r = 0; r < E; ++r)
#pragma unroll
            for (int c = 0; c < E; ++c) {
                double acc = 0.0;
#pragma unroll
                for (int m = 0; m < E; ++m) acc += R[r][m] * R[m][c];
                tmp[r][c] = acc;
            }
#pragma unroll
        for (int r = 0; r < E; ++r)
#pragma unroll
            for (int c = 0; c < E; ++c) R[r][c] = tmp[r][c];
    }
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < E; ++c) T[(size_t)i * D * E + r * E + c] = (float)R[r][c];
}

// libcpab/core/cpab_ops.cpp:224-260: grid (ceil(nP/BLOCK), batch); Trels of the
// block's theta staged in LDS
template <int D>
__global__ __launch_bounds__(BLOCK) void k_ref_forward(float* __restrict__ newpoints, const float* __restrict__ points,
                                                       const float* __restrict__ trels, int nstep, int nc0, int nc1,
                                                       int nc2, int nC, int nP, int broadcast) {
    extern __shared__ float sT[];
    constexpr int P = D * (D + 1);
    const int t = blockIdx.y;
    for (int k = threadIdx.x; k < nC * P; k += BLOCK) sT[k] = trels[(size_t)t * nC * P + k];
    __syncthreads();
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nP) return;
    const int nc[3] = {nc0, nc1, nc2};
    float pt[D];
#pragma unroll
    for (int a = 0; a < D; ++a) pt[a] = points[(size_t)broadcast * t * nP * D + i + (size_t)a * nP];
    for (int n = 0; n < nstep; ++n) {
        const int idx = cpab_ref::find<D>(pt, nc);
        float np_[D];
        cpab_ref::affine<D>(np_, sT + P * idx, pt);
#pragma unroll
        for (int a = 0; a < D; ++a) pt[a] = np_[a];
    }
#pragma unroll
    for (int a = 0; a < D; ++a) newpoints[(size_t)t * D * nP + i + (size_t)a * nP] = pt[a];
}

// libcpab/core/cpab_ops.cpp:262-372: one thread per (point, theta, k);
// grid (ceil(nP/BLOCK), n_theta, d); grad [d, n_theta, D, nP]
template <int D>
__global__ __launch_bounds__(BLOCK) void k_ref_backward(float* __restrict__ grad, const float* __restrict__ points,
                                                        const float* __restrict__ As, const float* __restrict__ Bs,
                                                        int nstep, int nc0, int nc1, int nc2, int n_theta, int nC,
                                                        int nP, int broadcast) {
#pragma clang fp contract(off)
    extern __shared__ float sAB[];
    constexpr int P = D * (D + 1);
    const int b = blockIdx.y, k = blockIdx.z;
    float* sA = sAB;
    float* sB = sAB + nC * P;
    for (int j = threadIdx.x; j < nC * P; j += BLOCK) {
        sA[j] = As[(size_t)b * nC * P + j];
        sB[j] = Bs[(size_t)k * nC * P + j];
    }
    __syncthreads();
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nP) return;
    const int nc[3] = {nc0, nc1, nc2};
    float p[D], q[D], v[D], pm[D], vm[D], qm[D], bt[D], aq[D], u[D], um[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        p[a] = points[(size_t)broadcast * b * D * nP + i + (size_t)a * nP];
        q[a] = 0.f;
    }
    const double h = 1.0 / nstep;
    for (int s = 0; s < nstep; ++s) {
        const int cell = cpab_ref::find<D>(p, nc);
        const float* Al = sA + P * cell;
        const float* Bl = sB + P * cell;
        cpab_ref::affine<D>(v, Al, p);
#pragma unroll
        for (int a = 0; a < D; ++a) pm[a] = (float)(p[a] + h * v[a] / 2.0);
        cpab_ref::affine<D>(vm, Al, pm);
        cpab_ref::affine<D>(bt, Bl, p);
        cpab_ref::linear<D>(aq, Al, q);
#pragma unroll
        for (int a = 0; a < D; ++a) u[a] = bt[a] + aq[a];
#pragma unroll
        for (int a = 0; a < D; ++a) qm[a] = (float)(q[a] + h * u[a] / 2.0);
        cpab_ref::affine<D>(bt, Bl, pm);
        cpab_ref::linear<D>(aq, Al, qm);
#pragma unroll
        for (int a = 0; a < D; ++a) um[a] = bt[a] + aq[a];
#pragma unroll
        for (int a = 0; a < D; ++a) q[a] = (float)(q[a] + um[a] * h);
#pragma unroll
        for (int a = 0; a < D; ++a) p[a] = (float)(p[a] + vm[a] * h);
    }
    const size_t box = (size_t)D * nP * n_theta;
#pragma unroll
    for (int a = 0; a < D; ++a) grad[(size_t)k * box + (size_t)D * nP * b + i + (size_t)a * nP] = q[a];
}

// ---------------------------------------------------------------------------
// use_slow: the reference's pure-torch stepping (libcpab/pytorch/transformer.py:
// 90-128) -- nstep applications of Trel = expm(A~/nstep) of the cell the torch
// lookup finds -- with the exact adjoint of that discrete map as its gradient
// (what torch autograd gives the reference), w.r.t. the points and the Trels
// ---------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ int slow_cell(const float (&x)[D], const int* nc) {
    if constexpr (D == 3) return find_cell3(x[0], x[1], x[2], nc[0], nc[1], nc[2], 1.f - 1e-4f);
    else return find_cell<D>(x, nc);
}

// x <- T[:, :D] x + T[:, D] for a [D, D+1] row-major affine map
template <int D>
__device__ __forceinline__ void slow_step(const float* __restrict__ T, float (&x)[D]) {
    float y[D];
#pragma unroll
    for (int r = 0; r < D; ++r) {
        float acc = T[r * (D + 1) + D];
#pragma unroll
        for (int k = 0; k < D; ++k) acc = fmaf(T[r * (D + 1) + k], x[k], acc);
        y[r] = acc;
    }
#pragma unroll
    for (int r = 0; r < D; ++r) x[r] = y[r];
}

// grid (ceil(nP/BLOCK), n_theta); trels [n_theta, nC, D, D+1] staged in LDS
template <int D>
__global__ __launch_bounds__(BLOCK) void k_slow_forward(const float* __restrict__ points, const float* __restrict__ trels,
                                                        int nstep, int nc0, int nc1, int nc2, int nC, int nP,
                                                        int per_theta, float* __restrict__ out) {
    extern __shared__ float sT[];
    constexpr int P = D * (D + 1);
    const int th = blockIdx.y;
    for (int k = threadIdx.x; k < nC * P; k += BLOCK) sT[k] = trels[(size_t)th * nC * P + k];
    __syncthreads();
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i >= nP) return;
    const int nc[3] = {nc0, nc1, nc2};
    const float* pts = points + (per_theta ? (size_t)th * D * nP : 0);
    float x[D];
#pragma unroll
    for (int a = 0; a < D; ++a) x[a] = pts[(size_t)a * nP + i];
    for (int n = 0; n < nstep; ++n) slow_step<D>(sT + P * slow_cell<D>(x, nc), x);
#pragma unroll
    for (int a = 0; a < D; ++a) out[((size_t)th * D + a) * nP + i] = x[a];
}

// Exact adjoint of k_slow_forward for x_{k+1} = M_k x_k + b_k, L(x_n):
// lambda_k = M_k^T lambda_{k+1}, dL/dM_k += lambda_{k+1} x_k^T, dL/db_k +=
// lambda_{k+1}, dL/dx_0 = lambda_0.  No trajectory storage: pass 1 carries
// J = M_{n-1}..M_0 (lambda_0 = J^T g), pass 2 steps forward again carrying
// lambda_{k+1} = M_k^-T lambda_k (tinv holds M^-1 per cell) and adds each
// step's outer products to the block's LDS accumulator, which goes to
// gtrels (zero-filled by the caller) by global float atomics.
template <int D>
__global__ __launch_bounds__(BLOCK) void k_slow_backward(const float* __restrict__ points,
                                                         const float* __restrict__ trels, const float* __restrict__ tinv,
                                                         int nstep, int nc0, int nc1, int nc2, int nC, int nP,
                                                         int per_theta, const float* __restrict__ gout,
                                                         float* __restrict__ gpts, float* __restrict__ gtrels) {
    extern __shared__ float smem[];
    constexpr int P = D * (D + 1);
    const int th = blockIdx.y;
    float* sT = smem;
    float* sI = sT + nC * P;
    float* sG = sI + nC * D * D;
    for (int k = threadIdx.x; k < nC * P; k += BLOCK) {
        sT[k] = trels[(size_t)th * nC * P + k];
        sG[k] = 0.f;
    }
    for (int k = threadIdx.x; k < nC * D * D; k += BLOCK) sI[k] = tinv[(size_t)th * nC * D * D + k];
    __syncthreads();
    const int i = blockIdx.x * BLOCK + threadIdx.x;
    if (i < nP) {
        const int nc[3] = {nc0, nc1, nc2};
        const float* pts = points + (per_theta ? (size_t)th * D * nP : 0);
        float x0[D], x[D], J[D][D], lam[D];
#pragma unroll
        for (int a = 0; a < D; ++a) {
            x0[a] = pts[(size_t)a * nP + i];
            x[a] = x0[a];
        }
#pragma unroll
        for (int r = 0; r < D; ++r)
#pragma unroll
            for (int k = 0; k < D; ++k) J[r][k] = (r == k) ? 1.f : 0.f;
        for (int n = 0; n < nstep; ++n) {
            const float* T = sT + P * slow_cell<D>(x, nc);
            float Jn[D][D];
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int k = 0; k < D; ++k) {
                    float acc = 0.f;
#pragma unroll
                    for (int a = 0; a < D; ++a) acc = fmaf(T[r * (D + 1) + a], J[a][k], acc);
                    Jn[r][k] = acc;
                }
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int k = 0; k < D; ++k) J[r][k] = Jn[r][k];
            slow_step<D>(T, x);
        }
#pragma unroll
        for (int k = 0; k < D; ++k) {
            float acc = 0.f;
#pragma unroll
            for (int r = 0; r < D; ++r) acc = fmaf(J[r][k], gout[((size_t)th * D + r) * nP + i], acc);
            lam[k] = acc;
        }
        if (gpts) {
#pragma unroll
            for (int k = 0; k < D; ++k) gpts[((size_t)th * D + k) * nP + i] = lam[k];
        }
#pragma unroll
        for (int a = 0; a < D; ++a) x[a] = x0[a];
        for (int n = 0; n < nstep; ++n) {
            const int c = slow_cell<D>(x, nc);
            const float* Mi = sI + D * D * c;
            float ln[D];
#pragma unroll
            for (int k = 0; k < D; ++k) {
                float acc = 0.f;
#pragma unroll
                for (int r = 0; r < D; ++r) acc = fmaf(Mi[r * D + k], lam[r], acc);
                ln[k] = acc;
            }
#pragma unroll
            for (int r = 0; r < D; ++r) {
                lam[r] = ln[r];
#pragma unroll
                for (int k = 0; k < D; ++k) atomicAdd(sG + P * c + r * (D + 1) + k, lam[r] * x[k]);
                atomicAdd(sG + P * c + r * (D + 1) + D, lam[r]);
            }
            slow_step<D>(sT + P * c, x);
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < nC * P; k += BLOCK)
        if (sG[k] != 0.f) atomicAdd(gtrels + (size_t)th * nC * P + k, sG[k]);
}

// ---------------------------------------------------------------------------
// host helpers
// ---------------------------------------------------------------------------
int cells_of(int ndim, const int* nc) {
    int s = ndim == 1 ? 1 : (ndim == 2 ? 4 : 5);
    for (int a = 0; a < ndim; ++a) s *= nc[a];
    return s;
}

int check_tess(const cpab_tess_t* t) {
    if (!t) return CPAB_ERR_ARG;
    if (t->ndim < 1 || t->ndim > 3) return CPAB_ERR_NDIM;
    for (int a = 0; a < t->ndim; ++a)
        if (t->nc[a] < 1) return CPAB_ERR_ARG;
    if (t->nC != cells_of(t->ndim, t->nc)) return CPAB_ERR_ARG;
    if (t->nV < 0 || (t->nV > 0 && (t->ndim != 3 || !t->amap))) return CPAB_ERR_ARG;
    if (t->ndim > 1 && (!t->planes || !t->nbr)) return CPAB_ERR_ARG;
    return CPAB_OK;
}

Geo make_geo(const cpab_tess_t* t) {
    Geo g;
    g.ndim = t->ndim;
    g.nc0 = t->nc[0];
    g.nc1 = t->ndim > 1 ? t->nc[1] : 1;
    g.nc2 = t->ndim > 2 ? t->nc[2] : 1;
    g.nC = t->nC;
    g.pl = reinterpret_cast<const float4*>(t->planes);
    g.nb = t->nbr;
    g.unfinished = t->unfinished;
    g.nV = t->nV;
    g.amap = t->amap;
    int m = g.nc0 > g.nc1 ? g.nc0 : g.nc1;
    m = m > g.nc2 ? m : g.nc2;
    g.tol0 = 2e-6f * (float)(m > 1 ? m : 1);
    return g;
}

// points per thread: the largest of 8, 4, 2, 1 that still gives >= 16 blocks
// per CU: with a few rounds of blocks per CU the last round drains the chip
// (seq2d 1.56 -> 1.47 ms/step, vol3d 0.593 -> 0.574 against >= 4 per CU)
int pick_ppt(int n_theta, int nP) {
    const long long want = 16LL * 256;
    for (int ppt = 8; ppt > 1; ppt >>= 1) {
        const long long blocks = (long long)n_theta * ((nP + BLOCK * ppt - 1) / (BLOCK * ppt));
        if (blocks >= want) return ppt;
    }
    return 1;
}

// the fused forward of 2D tiled grids: at most 4 points per thread (images2d
// forward 0.508 ms at 8, 0.505 at 4: more, shorter blocks drain better at
// the end of the launch)
int pick_ppt_fwd(int ndim, int n_theta, int nP, int g1) {
    const int ppt = pick_ppt(n_theta, nP);
    return (ndim == 2 && tile_ty(2, g1) > 0 && ppt > 4) ? 4 : ppt;
}

// the fused backward of 2D tiled grids runs 512-point tiles (32 x 16 for the
// usual grids) whose data footprint fits the LDS box
int pick_ppt_bwd(int ndim, int n_theta, int nP, int g1) {
    const int ppt = pick_ppt(n_theta, nP);
    return (ndim == 2 && tile_ty(2, g1) > 0 && ppt > 2) ? 2 : ppt;
}

// tiles per block of the fused 2D tiled backward: the most of 4, 2 that
// still leaves 16 blocks per CU (256 CUs), else 1 (images2d backward, tiles
// down a column: 0.81 ms at 1, 0.766 at 2, 0.741 at 4, 0.745 at 8)
int bwd_tiles_per_block(int ndim, int ty, int n_theta, int tiles) {
    if (ndim != 2 || ty == 0) return 1;
    for (int tpb = 4; tpb > 1; tpb >>= 1)
        if ((long long)n_theta * ((tiles + tpb - 1) / tpb) >= 16LL * 256) return tpb;
    return 1;
}

// LDS words for the dL/ddata bins and point slots of the fused 2D backward
// (a 512-point tile needs 3584 + footprint cells + 1 words: up to 2047
// cells, a 32 x 16 tile at up to ~2x magnification)
constexpr int BOX_FLOATS = 6144;

size_t lds_tables(int ndim, int nC) {  // nC: every cell of the facet table
    switch (ndim) {
        case 1: return table_floats<1>(nC) * 4;
        case 2: return table_floats<2>(nC) * 4;
        default: return table_floats<3>(nC) * 4;
    }
}

// tables of all nCt cells, per-wave dL/dA accumulators of the nC simplices
size_t lds_bwd(int ndim, int nC, int nCt) { return lds_tables(ndim, nCt) + (size_t)WAVES * nC * ndim * (ndim + 1) * 4; }

// LDS one workgroup may allocate on gfx950 (160 KB per CU, all of it per block)
constexpr size_t LDS_LIMIT = 160 * 1024;

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

int status_from_hip() {
    hipError_t e = hipGetLastError();
    return e == hipSuccess ? CPAB_OK : CPAB_ERR_HIP;
}

int launch_theta_to_A(const cpab_tess_t* t, const float* theta, int n_theta, float* Aq, hipStream_t st) {
    const int Dt = t->nC * t->ndim * (t->ndim + 1);
    dim3 grid((Dt + BLOCK - 1) / BLOCK, n_theta);
    Geo g = make_geo(t);
    hipLaunchKernelGGL(k_theta_to_A, grid, dim3(BLOCK), t->d * sizeof(float), st, theta, t->B, t->d, Dt, t->ndim, g.nc0,
                       g.nc1, g.nc2, Aq);
    return status_from_hip();
}

// dL/dtheta from the per-tile partials (overwritten): column-sliced tile sums,
// then B^T per theta.  Deterministic: fixed summation order throughout.
int launch_reduce(const cpab_tess_t* t, float* part, int n_theta, int tiles, float* gtheta, float* gA,
                  hipStream_t st) {
    const int Dt = t->nC * t->ndim * (t->ndim + 1);
    Geo g = make_geo(t);
    // column slices of at least 64 columns, enough of them for ~1024 blocks
    int nslice = (1024 + n_theta - 1) / n_theta;
    const int max_slices = (Dt + 63) / 64;
    if (nslice > max_slices) nslice = max_slices;
    if (nslice < 1) nslice = 1;
    if (nslice == 1 || tiles <= 64) {
        // enough thetas, or few rows per theta: one block per theta, as many
        // waves (up to 16) as the LDS holds [nw + 1, Dt] floats for
        int nw = RBLOCK / 64;
        while (nw > 1 && (size_t)(nw + 1) * Dt * sizeof(float) > LDS_LIMIT) nw >>= 1;
        if ((size_t)(nw + 1) * Dt * sizeof(float) <= LDS_LIMIT) {
            hipLaunchKernelGGL(k_reduce_theta, dim3(n_theta), dim3(nw * 64), (nw + 1) * Dt * sizeof(float), st, part,
                               tiles, Dt, t->ndim, g.nc0, g.nc1, g.nc2, t->B, t->d, gtheta, gA);
            return status_from_hip();
        }
    }
    // k_reduce_cols keeps [16 waves, cw] floats in LDS: at most cols_fit columns per slice
    const int nw = RBLOCK / 64;
    const int cols_fit = (int)(LDS_LIMIT / ((size_t)nw * sizeof(float)));
    if (nslice < (Dt + cols_fit - 1) / cols_fit) nslice = (Dt + cols_fit - 1) / cols_fit;
    const int cw = (Dt + nslice - 1) / nslice;
    nslice = (Dt + cw - 1) / cw;
    // B^T: as many waves (up to 16) as the LDS holds Dt + nw * d floats for;
    // checked before the first launch
    int nwb = RBLOCK / 64;
    while (nwb > 1 && (size_t)(Dt + nwb * t->d) * sizeof(float) > LDS_LIMIT) nwb >>= 1;
    if ((size_t)(Dt + nwb * t->d) * sizeof(float) > LDS_LIMIT) return CPAB_ERR_LDS;
    hipLaunchKernelGGL(k_reduce_cols, dim3(n_theta, nslice), dim3(nw * 64), (size_t)nw * cw * sizeof(float), st,
                       part, tiles, Dt, cw, t->ndim, g.nc0, g.nc1, g.nc2, gA);
    hipLaunchKernelGGL(k_reduce_bt, dim3(n_theta), dim3(nwb * 64), (size_t)(Dt + nwb * t->d) * sizeof(float), st,
                       part, tiles, Dt, t->B, t->d, gtheta);
    return status_from_hip();
}

}  // namespace

// ===========================================================================
// static scratch of cpab_transform_data_fwd_noalloc
// ===========================================================================
// A_q of the allocation-free forward lives here instead of in a workspace.
// Each device that loads the library holds its own copy until the process
// ends.  In host code the array's name is not a device address: kernels get
// the address hipGetSymbolAddress returns on the device (kept per device).
static __device__ __align__(256) float scratch[CPAB_NOALLOC_SCRATCH_BYTES / sizeof(float)];

namespace {

constexpr int SCRATCH_MAX_DEVICES = 64;
std::mutex g_scratch_mu;
float* g_scratch_ptr[SCRATCH_MAX_DEVICES];

// Device address of `scratch` on the stream's device, looked up once per
// device with that device current.  The stream must belong to the current
// device (the null stream always does).
int scratch_on_device(hipStream_t stream, float** out) {
    int dev = -1;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= SCRATCH_MAX_DEVICES) {
        (void)hipGetLastError();
        return CPAB_ERR_HIP;
    }
    if (stream) {
        hipDevice_t sdev = -1;
        if (hipStreamGetDevice(stream, &sdev) != hipSuccess) {
            (void)hipGetLastError();
            return CPAB_ERR_HIP;
        }
        if (sdev != dev) return CPAB_ERR_ARG;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
            (void)hipGetLastError();
            return CPAB_ERR_HIP;
        }
        g_scratch_ptr[dev] = static_cast<float*>(p);
    }
    *out = g_scratch_ptr[dev];
    return CPAB_OK;
}

}  // namespace

// ===========================================================================
// extern "C" boundary
// ===========================================================================
extern "C" {

const char* cpab_status_string(int status) {
    switch (status) {
        case CPAB_OK: return "ok";
        case CPAB_ERR_ARG: return "invalid argument";
        case CPAB_ERR_NDIM: return "ndim must be 1, 2 or 3";
        case CPAB_ERR_WORKSPACE: return "workspace too small";
        case CPAB_ERR_HIP: return "HIP launch failed";
        case CPAB_ERR_LDS: return "tessellation too large for the LDS budget of one block";
        case CPAB_ERR_UNSUPPORTED: return "unsupported configuration";
        case CPAB_ERR_SCRATCH:
            return "problem needs more than the CPAB_NOALLOC_SCRATCH_BYTES of static scratch of the allocation-free "
                   "entry point";
        default: return "unknown status";
    }
}

int cpab_abi_version(void) { return CPAB_ABI_VERSION; }

// Launch shape of a grid transform.  2D points that form a meshgrid of
// shape[0] x shape[1] (x fastest, shape[0] * shape[1] == nP) run in 2D tiles --
// tile_t with the axes swapped, so its index i1 * shape[0] + i0 is the point
// index -- when at most max_tiles tiles result (the partials the caller's
// workspace holds; 0: no limit); everything else in runs of consecutive points.
// The order changes which points share a tile, not any point's result.
static void grid_launch(const cpab_tess_t* t, int n_theta, int nP, const int* shape, size_t max_tiles, Launch& L,
                        Src& s) {
    L.n_theta = n_theta;
    L.nP = nP;
    L.ppt = pick_ppt(n_theta, nP);
    L.ty = 0;
    L.tiles = nP > 0 ? (nP + BLOCK * L.ppt - 1) / (BLOCK * L.ppt) : 1;
    if (shape && t->ndim == 2 && shape[0] > 0 && shape[1] > 0 && (long long)shape[0] * shape[1] == nP) {
        const int ty = tile_ty(2, shape[0]);
        const int tiles = tile_count(2, nP, shape[1], shape[0], L.ppt);
        if (ty > 0 && (max_tiles == 0 || (size_t)tiles <= max_tiles)) {
            L.ty = ty;
            L.tiles = tiles;
            s.g0 = shape[1];
            s.g1 = shape[0];
        }
    }
}

size_t cpab_workspace_bytes(const cpab_tess_t* t, int n_theta, int nP) {
    if (check_tess(t) != CPAB_OK || n_theta < 0 || nP < 0) return 0;
    const size_t Dt = (size_t)t->nC * t->ndim * (t->ndim + 1);
    const int ppt = pick_ppt(n_theta, nP);
    size_t tiles = (nP + (size_t)BLOCK * ppt - 1) / ((size_t)BLOCK * ppt);
    if (tiles == 0) tiles = 1;
    return align256((size_t)n_theta * Dt * 4) + align256((size_t)n_theta * tiles * Dt * 4);
}

int cpab_theta_to_A(const cpab_tess_t* t, const float* theta, int n_theta, float* Aq, hipStream_t stream) {
    int rc = check_tess(t);
    if (rc) return rc;
    if (!theta || !Aq || !t->B || n_theta < 0) return CPAB_ERR_ARG;
    if (n_theta == 0) return CPAB_OK;
    return launch_theta_to_A(t, theta, n_theta, Aq, stream);
}

#define CPAB_DISPATCH_D(ndim, FN) \
    switch (ndim) {               \
        case 1: FN(1); break;     \
        case 2: FN(2); break;     \
        default: FN(3); break;    \
    }

int cpab_transform_fwd(const cpab_tess_t* t, const float* theta, int n_theta, const float* points, int nP,
                       int per_theta_grid, const int* grid_shape, float t_end, float* out, float* jac,
                       void* workspace, size_t ws_bytes, hipStream_t stream) {
    int rc = check_tess(t);
    if (rc) return rc;
    if (n_theta < 0 || nP < 0) return CPAB_ERR_ARG;
    if (n_theta == 0 || nP == 0) return CPAB_OK;  // arrays with no elements may be NULL
    if (!theta || !points || !out) return CPAB_ERR_ARG;
    if (ws_bytes < cpab_workspace_bytes(t, n_theta, nP) || !workspace) return CPAB_ERR_WORKSPACE;
    const size_t lds = lds_tables(t->ndim, t->nC + t->nV);
    if (lds > LDS_LIMIT) return CPAB_ERR_LDS;
    float* Aq = static_cast<float*>(workspace);
    if ((rc = launch_theta_to_A(t, theta, n_theta, Aq, stream))) return rc;
    Launch L;
    Src s = {};
    grid_launch(t, n_theta, nP, grid_shape, 0, L, s);
    L.t_end = t_end;
    Geo g = make_geo(t);
    s.pts = points;
    s.per_theta = per_theta_grid;
    dim3 grid(L.tiles * n_theta);
#define LAUNCH_TF(DD)                                                                                                \
    if (jac)                                                                                                          \
        hipLaunchKernelGGL((k_transform<DD, true>), grid, dim3(BLOCK), lds, stream, L, g, s, Aq, out, jac);           \
    else                                                                                                              \
        hipLaunchKernelGGL((k_transform<DD, false>), grid, dim3(BLOCK), lds, stream, L, g, s, Aq, out, jac);
    CPAB_DISPATCH_D(t->ndim, LAUNCH_TF)
#undef LAUNCH_TF
    return status_from_hip();
}

int cpab_transform_cells(const cpab_tess_t* t, const float* theta, int n_theta, const float* points, int nP,
                         int per_theta_grid, float t_end, int max_segs, int* cells, int* nseg, void* workspace,
                         size_t ws_bytes, hipStream_t stream) {
    int rc = check_tess(t);
    if (rc) return rc;
    if (n_theta < 0 || nP < 0 || max_segs < 0) return CPAB_ERR_ARG;
    if (n_theta == 0 || nP == 0) return CPAB_OK;
    if (!theta || !points || !nseg || (max_segs > 0 && !cells)) return CPAB_ERR_ARG;
    if (ws_bytes < cpab_workspace_bytes(t, n_theta, nP) || !workspace) return CPAB_ERR_WORKSPACE;
    const size_t lds = lds_tables(t->ndim, t->nC + t->nV);
    if (lds > LDS_LIMIT) return CPAB_ERR_LDS;
    float* Aq = static_cast<float*>(workspace);
    if ((rc = launch_theta_to_A(t, theta, n_theta, Aq, stream))) return rc;
    Launch L;
    L.n_theta = n_theta;
    L.nP = nP;
    L.ppt = 1;
    L.ty = 0;
    L.tiles = (nP + BLOCK - 1) / BLOCK;
    L.t_end = t_end;
    Geo g = make_geo(t);
    Src s = {};
    s.pts = points;
    s.per_theta = per_theta_grid;
    dim3 grid(L.tiles * n_theta);
#define LAUNCH_TC(DD) \
    hipLaunchKernelGGL((k_transform_cells<DD>), grid, dim3(BLOCK), lds, stream, L, g, s, Aq, max_segs, cells, nseg);
    CPAB_DISPATCH_D(t->ndim, LAUNCH_TC)
#undef LAUNCH_TC
    return status_from_hip();
}

int cpab_transform_bwd(const cpab_tess_t* t, const float* theta, int n_theta, const float* points, int nP,
                       int per_theta_grid, const int* grid_shape, float t_end, const float* grad_out, const float* jac,
                       float* grad_theta, float* grad_points, void* workspace, size_t ws_bytes, hipStream_t stream) {
    int rc = check_tess(t);
    if (rc) return rc;
    if (n_theta < 0 || nP < 0) return CPAB_ERR_ARG;
    if (n_theta == 0) return CPAB_OK;
    // no points: grad_theta is written with zeros; points and grad_out may be NULL
    if (!theta || !grad_theta || (nP > 0 && (!points || !grad_out))) return CPAB_ERR_ARG;
    if (ws_bytes < cpab_workspace_bytes(t, n_theta, nP) || !workspace) return CPAB_ERR_WORKSPACE;
    const size_t lds = lds_bwd(t->ndim, t->nC, t->nC + t->nV);
    if (lds > LDS_LIMIT) return CPAB_ERR_LDS;
    const int Dt = t->nC * t->ndim * (t->ndim + 1);
    float* Aq = static_cast<float*>(workspace);
    float* part = reinterpret_cast<float*>(static_cast<char*>(workspace) + align256((size_t)n_theta * Dt * 4));
    if ((rc = launch_theta_to_A(t, theta, n_theta, Aq, stream))) return rc;
    Launch L;
    Src s = {};
    // 2D tiles only as far as the workspace holds their partials
    const size_t head = align256((size_t)n_theta * Dt * 4);
    grid_launch(t, n_theta, nP, grid_shape, (ws_bytes - head) / ((size_t)n_theta * Dt * 4), L, s);
    L.t_end = t_end;
    Geo g = make_geo(t);
    s.pts = points;
    s.per_theta = per_theta_grid;
    dim3 grid(L.tiles * n_theta);
#define LAUNCH_TB(DD)                                                                                                   \
    if (jac)                                                                                                             \
        hipLaunchKernelGGL((k_transform_bwd<DD, true>), grid, dim3(BLOCK), lds, stream, L, g, s, Aq, grad_out, jac, part, \
                           grad_points);                                                                                 \
    else                                                                                                                 \
        hipLaunchKernelGGL((k_transform_bwd<DD, false>), grid, dim3(BLOCK), lds, stream, L, g, s, Aq, grad_out, jac,     \
                           part, grad_points);
    CPAB_DISPATCH_D(t->ndim, LAUNCH_TB)
#undef LAUNCH_TB
    if ((rc = status_from_hip())) return rc;
    return launch_reduce(t, part, n_theta, L.tiles, grad_theta, nullptr, stream);
}

static int img_from(int ndim, const float* data, int C, const int* dsize, Img& im) {
    if (!data || !dsize || C < 1) return CPAB_ERR_ARG;
    im.data = data;
    im.gdata = nullptr;
    im.C = C;
    im.s0 = dsize[0];
    im.s1 = ndim > 1 ? dsize[1] : 1;
    im.s2 = ndim > 2 ? dsize[2] : 1;
    if (im.s0 < 1 || im.s1 < 1 || im.s2 < 1) return CPAB_ERR_ARG;
    return CPAB_OK;
}

static int src_out(int ndim, const int* outsize, Src& s, int& nP) {
    if (!outsize) return CPAB_ERR_ARG;
    s.g0 = outsize[0];
    s.g1 = ndim > 1 ? outsize[1] : 1;
    s.g2 = ndim > 2 ? outsize[2] : 1;
    if (s.g0 < 1 || s.g1 < 1 || s.g2 < 1) return CPAB_ERR_ARG;
    nP = s.g0 * s.g1 * s.g2;
    return CPAB_OK;
}

int cpab_interpolate_fwd(int ndim, const float* data, int N, int C, const int* dsize, const float* grid,
                         const int* outsize, float* out, hipStream_t stream) {
    if (ndim < 1 || ndim > 3) return CPAB_ERR_NDIM;
    Img im;
    Src s = {};
    int nP, rc;
    if ((rc = img_from(ndim, data, C, dsize, im))) return rc;
    if ((rc = src_out(ndim, outsize, s, nP))) return rc;
    if (N < 0) return CPAB_ERR_ARG;
    if (N == 0) return CPAB_OK;
    if (!grid || !out) return CPAB_ERR_ARG;
    s.pts = grid;
    const size_t total = (size_t)N * nP;
    const size_t blocks = ndim == 2 ? (size_t)N * ((s.g0 + 15) / 16) * ((s.g1 + 15) / 16) : (total + BLOCK - 1) / BLOCK;
    dim3 g((unsigned)blocks);
#define LAUNCH_IF(DD) hipLaunchKernelGGL((k_interp_fwd<DD>), g, dim3(BLOCK), 0, stream, N, nP, s, im, out);
    CPAB_DISPATCH_D(ndim, LAUNCH_IF)
#undef LAUNCH_IF
    return status_from_hip();
}

int cpab_interpolate_bwd(int ndim, const float* data, int N, int C, const int* dsize, const float* grid,
                         const int* outsize, const float* grad_out, float* grad_data, float* grad_grid,
                         hipStream_t stream) {
    if (ndim < 1 || ndim > 3) return CPAB_ERR_NDIM;
    Img im;
    Src s = {};
    int nP, rc;
    if ((rc = img_from(ndim, data, C, dsize, im))) return rc;
    if ((rc = src_out(ndim, outsize, s, nP))) return rc;
    if (N < 0) return CPAB_ERR_ARG;
    if (N == 0) return CPAB_OK;
    if (!grid || !grad_out) return CPAB_ERR_ARG;
    im.gdata = grad_data;
    s.pts = grid;
    const size_t total = (size_t)N * nP;
    const size_t blocks = ndim == 2 ? (size_t)N * ((s.g0 + 15) / 16) * ((s.g1 + 15) / 16) : (total + BLOCK - 1) / BLOCK;
    dim3 g((unsigned)blocks);
#define LAUNCH_IB(DD) hipLaunchKernelGGL((k_interp_bwd<DD>), g, dim3(BLOCK), 0, stream, N, nP, s, im, grad_out, grad_grid);
    CPAB_DISPATCH_D(ndim, LAUNCH_IB)
#undef LAUNCH_IB
    return status_from_hip();
}

size_t cpab_saved_floats(int ndim, int n_theta, int nP) { return (size_t)n_theta * saved_rows(ndim) * nP; }

size_t cpab_jac_floats(int ndim, int n_theta, int nP) { return (size_t)n_theta * jac_rows(ndim) * nP; }

size_t cpab_data_workspace_bytes(const cpab_tess_t* t, int n_theta, const int* outsize) {
    if (check_tess(t) != CPAB_OK || n_theta < 0 || !outsize) return 0;
    Src s = {};
    int nP;
    if (src_out(t->ndim, outsize, s, nP)) return 0;
    const size_t Dt = (size_t)t->nC * t->ndim * (t->ndim + 1);
    const size_t tiles = tile_count(t->ndim, nP, s.g0, s.g1, pick_ppt_bwd(t->ndim, n_theta, nP, s.g1), s.g2);
    return align256((size_t)n_theta * Dt * 4) + align256((size_t)n_theta * tiles * Dt * 4);
}

// The fused forward with A_q either in the caller's workspace or (workspace
// NULL, ws_bytes 0, from_scratch) in this device's static scratch array.
static int data_fwd(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                    const int* dsize, const int* outsize, const float* lin, float t_end, float* out, float* saved,
                    void* workspace, size_t ws_bytes, bool from_scratch, hipStream_t stream) {
    int rc = check_tess(t);
    if (rc) return rc;
    Img im;
    Src s = {};
    int nP;
    if ((rc = img_from(t->ndim, data, C, dsize, im))) return rc;
    if ((rc = src_out(t->ndim, outsize, s, nP))) return rc;
    if (!theta || !lin || !out || n_theta < 0) return CPAB_ERR_ARG;
    if (n_theta == 0) return CPAB_OK;
    float* Aq;
    if (from_scratch) {
        const size_t Dt = (size_t)t->nC * t->ndim * (t->ndim + 1);
        if ((size_t)n_theta * Dt * sizeof(float) > (size_t)CPAB_NOALLOC_SCRATCH_BYTES) return CPAB_ERR_SCRATCH;
        if ((rc = scratch_on_device(stream, &Aq))) return rc;
    } else {
        if (ws_bytes < cpab_data_workspace_bytes(t, n_theta, outsize) || !workspace) return CPAB_ERR_WORKSPACE;
        Aq = static_cast<float*>(workspace);
    }
    Launch L;
    L.n_theta = n_theta;
    L.nP = nP;
    L.ppt = pick_ppt_fwd(t->ndim, n_theta, nP, s.g1);
    L.ty = tile_ty(t->ndim, s.g1, s.g2);
    const size_t lds = lds_tables(t->ndim, t->nC + t->nV);
    if (lds > LDS_LIMIT) return CPAB_ERR_LDS;
    if ((rc = launch_theta_to_A(t, theta, n_theta, Aq, stream))) return rc;
    s.lin = lin;
    L.tiles = tile_count(t->ndim, nP, s.g0, s.g1, L.ppt, s.g2);
    L.t_end = t_end;
    Geo g = make_geo(t);
    dim3 grid(L.tiles * n_theta);
#define LAUNCH_DF(DD)                                                                                                    \
    if (saved)                                                                                                            \
        hipLaunchKernelGGL((k_transform_data<DD, true>), grid, dim3(BLOCK), lds, stream, L, g, s, im, Aq, out, saved);    \
    else                                                                                                                  \
        hipLaunchKernelGGL((k_transform_data<DD, false>), grid, dim3(BLOCK), lds, stream, L, g, s, im, Aq, out, saved);
    CPAB_DISPATCH_D(t->ndim, LAUNCH_DF)
#undef LAUNCH_DF
    return status_from_hip();
}

int cpab_transform_data_fwd(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                            const int* dsize, const int* outsize, const float* lin, float t_end, float* out,
                            float* saved, void* workspace, size_t ws_bytes, hipStream_t stream) {
    return data_fwd(t, theta, n_theta, data, C, dsize, outsize, lin, t_end, out, saved, workspace, ws_bytes, false,
                    stream);
}

int cpab_transform_data_fwd_noalloc(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                                    const int* dsize, const int* outsize, const float* lin, float t_end, float* out,
                                    float* saved, hipStream_t stream) {
    return data_fwd(t, theta, n_theta, data, C, dsize, outsize, lin, t_end, out, saved, nullptr, 0, true, stream);
}

int cpab_transform_data_bwd(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                            const int* dsize, const int* outsize, const float* lin, float t_end, const float* grad_out,
                            const float* saved, float* grad_data, float* grad_theta, void* workspace, size_t ws_bytes,
                            hipStream_t stream) {
    int rc = check_tess(t);
    if (rc) return rc;
    Img im;
    Src s = {};
    int nP;
    if ((rc = img_from(t->ndim, data, C, dsize, im))) return rc;
    if ((rc = src_out(t->ndim, outsize, s, nP))) return rc;
    if (!theta || !lin || !grad_out || !grad_theta || n_theta < 0) return CPAB_ERR_ARG;
    if (n_theta == 0) return CPAB_OK;
    if (ws_bytes < cpab_data_workspace_bytes(t, n_theta, outsize) || !workspace) return CPAB_ERR_WORKSPACE;
    Launch L;
    L.n_theta = n_theta;
    L.nP = nP;
    L.ppt = pick_ppt_bwd(t->ndim, n_theta, nP, s.g1);
    L.ty = tile_ty(t->ndim, s.g1, s.g2);
    // 2D with saved state and dL/ddata: accumulate dL/ddata per tile in LDS
    const int box_cap = (t->ndim == 2 && L.ty > 0 && saved && grad_data &&
                         lds_bwd(t->ndim, t->nC, t->nC + t->nV) + (size_t)BOX_FLOATS * 4 <= LDS_LIMIT)
                            ? BOX_FLOATS
                            : 0;
    const size_t lds = lds_bwd(t->ndim, t->nC, t->nC + t->nV) + (size_t)box_cap * 4;
    if (lds > LDS_LIMIT) return CPAB_ERR_LDS;
    const int Dt = t->nC * t->ndim * (t->ndim + 1);
    float* Aq = static_cast<float*>(workspace);
    float* part = reinterpret_cast<float*>(static_cast<char*>(workspace) + align256((size_t)n_theta * Dt * 4));
    if ((rc = launch_theta_to_A(t, theta, n_theta, Aq, stream))) return rc;
    im.gdata = grad_data;
    s.lin = lin;
    L.tiles = tile_count(t->ndim, nP, s.g0, s.g1, L.ppt, s.g2);
    L.t_end = t_end;
    Geo g = make_geo(t);
    const int tpb = bwd_tiles_per_block(t->ndim, L.ty, n_theta, L.tiles);
    const int nbt = bwd_blocks_per_theta(L.tiles, s.g1, L.ty, tpb);
    dim3 grid(nbt * n_theta);
#define LAUNCH_DB(DD)                                                                                                     \
    if (saved)                                                                                                             \
        hipLaunchKernelGGL((k_transform_data_bwd<DD, true>), grid, dim3(BLOCK), lds, stream, L, g, s, im, Aq, grad_out,    \
                           saved, part, box_cap, tpb);                                                                     \
    else                                                                                                                   \
        hipLaunchKernelGGL((k_transform_data_bwd<DD, false>), grid, dim3(BLOCK), lds, stream, L, g, s, im, Aq, grad_out,   \
                           saved, part, 0, tpb);
    CPAB_DISPATCH_D(t->ndim, LAUNCH_DB)
#undef LAUNCH_DB
    if ((rc = status_from_hip())) return rc;
    return launch_reduce(t, part, n_theta, nbt, grad_theta, nullptr, stream);
}

int cpab_findcell(int ndim, const int* nc, const float* points, int nP, int* idx, hipStream_t stream) {
    if (ndim < 1 || ndim > 3) return CPAB_ERR_NDIM;
    if (!nc || nP < 0) return CPAB_ERR_ARG;
    if (nP == 0) return CPAB_OK;
    if (!points || !idx) return CPAB_ERR_ARG;
    const int n0 = nc[0], n1 = ndim > 1 ? nc[1] : 1, n2 = ndim > 2 ? nc[2] : 1;
    dim3 g((nP + BLOCK - 1) / BLOCK);
#define LAUNCH_FC(DD) hipLaunchKernelGGL((k_findcell<DD>), g, dim3(BLOCK), 0, stream, n0, n1, n2, points, nP, idx);
    CPAB_DISPATCH_D(ndim, LAUNCH_FC)
#undef LAUNCH_FC
    return status_from_hip();
}

int cpab_vectorfield(const cpab_tess_t* t, const float* theta, const float* points, int nP, float* v, void* workspace,
                     size_t ws_bytes, hipStream_t stream) {
    int rc = check_tess(t);
    if (rc) return rc;
    if (nP < 0) return CPAB_ERR_ARG;
    if (nP == 0) return CPAB_OK;
    if (!theta || !points || !v) return CPAB_ERR_ARG;
    if (ws_bytes < cpab_workspace_bytes(t, 1, nP) || !workspace) return CPAB_ERR_WORKSPACE;
    float* Aq = static_cast<float*>(workspace);
    if ((rc = launch_theta_to_A(t, theta, 1, Aq, stream))) return rc;
    Geo g = make_geo(t);
    dim3 grid((nP + BLOCK - 1) / BLOCK);
#define LAUNCH_VF(DD) hipLaunchKernelGGL((k_vectorfield<DD>), grid, dim3(BLOCK), 0, stream, g, Aq, points, nP, v);
    CPAB_DISPATCH_D(t->ndim, LAUNCH_VF)
#undef LAUNCH_VF
    return status_from_hip();
}

/* One launch sequence per instance: the plain method the batched launches are
   measured against.  Same arguments and results as the batched calls. */
int cpab_transform_data_fwd_plain(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                                  const int* dsize, const int* outsize, const float* lin, float t_end, float* out,
                                  float* saved, void* workspace, size_t ws_bytes, hipStream_t stream) {
    int rc = check_tess(t);
    if (rc) return rc;
    Img im;
    Src s = {};
    int nP;
    if ((rc = img_from(t->ndim, data, C, dsize, im))) return rc;
    if ((rc = src_out(t->ndim, outsize, s, nP))) return rc;
    const size_t dstride = (size_t)im.s0 * im.s1 * im.s2;
    const size_t sv = (size_t)saved_rows(t->ndim) * nP;
    if (ws_bytes < cpab_data_workspace_bytes(t, 1, outsize) || !workspace) return CPAB_ERR_WORKSPACE;
    for (int i = 0; i < n_theta; ++i) {
        rc = cpab_transform_data_fwd(t, theta + (size_t)i * t->d, 1, data + (size_t)i * C * dstride, C, dsize, outsize,
                                     lin, t_end, out + (size_t)i * C * nP, saved ? saved + i * sv : nullptr, workspace,
                                     ws_bytes, stream);
        if (rc) return rc;
    }
    return CPAB_OK;
}

int cpab_transform_data_bwd_plain(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                                  const int* dsize, const int* outsize, const float* lin, float t_end,
                                  const float* grad_out, const float* saved, float* grad_data, float* grad_theta,
                                  void* workspace, size_t ws_bytes, hipStream_t stream) {
    int rc = check_tess(t);
    if (rc) return rc;
    Img im;
    Src s = {};
    int nP;
    if ((rc = img_from(t->ndim, data, C, dsize, im))) return rc;
    if ((rc = src_out(t->ndim, outsize, s, nP))) return rc;
    const size_t dstride = (size_t)im.s0 * im.s1 * im.s2;
    const size_t sv = (size_t)saved_rows(t->ndim) * nP;
    if (ws_bytes < cpab_data_workspace_bytes(t, 1, outsize) || !workspace) return CPAB_ERR_WORKSPACE;
    for (int i = 0; i < n_theta; ++i) {
        rc = cpab_transform_data_bwd(t, theta + (size_t)i * t->d, 1, data + (size_t)i * C * dstride, C, dsize, outsize,
                                     lin, t_end, grad_out + (size_t)i * C * nP, saved ? saved + i * sv : nullptr,
                                     grad_data ? grad_data + (size_t)i * C * dstride : nullptr,
                                     grad_theta + (size_t)i * t->d, workspace, ws_bytes, stream);
        if (rc) return rc;
    }
    return CPAB_OK;
}

size_t cpab_batch_workspace_bytes(const cpab_tess_t* t, const int* outsize, int ppt) {
    if (check_tess(t) != CPAB_OK || !outsize || ppt < 1) return 0;
    Src s = {};
    int nP;
    if (src_out(t->ndim, outsize, s, nP)) return 0;
    const size_t Dt = (size_t)t->nC * t->ndim * (t->ndim + 1);
    const size_t tiles = tile_count(t->ndim, nP, s.g0, s.g1, ppt, s.g2);
    const size_t batched = align256(Dt * 4) + align256(tiles * Dt * 4);
    const size_t plain = cpab_data_workspace_bytes(t, 1, outsize);
    return batched > plain ? batched : plain;
}

// LDS class of an instance in the batched backward: its tables and
// per-wave accumulators next to a full dL/ddata box fit 4 (class 0), 3 (1)
// or 2 or fewer (2) blocks per CU
static int batch_class(const cpab_instance_t& I) {
    const int nd = I.tess.ndim;
    const size_t need = align256(lds_tables(nd, I.tess.nC + I.tess.nV)) +
                        (size_t)WAVES * I.tess.nC * nd * (nd + 1) * 4 + (size_t)BOX_FLOATS * 4 + 256;
    return need <= LDS_LIMIT / 4 ? 0 : (need <= LDS_LIMIT / 3 ? 1 : 2);
}

int cpab_batch_plan(cpab_instance_t* inst, int n_inst, int ppt, int* item_inst, int max_items) {
    if (!inst || n_inst < 0 || ppt < 1) return -CPAB_ERR_ARG;
    long long item = 0;
    for (int i = 0; i < n_inst; ++i) {
        const cpab_instance_t& I = inst[i];
        int rc = check_tess(&I.tess);
        if (rc) return -rc;
        const int nd = I.tess.ndim;
        item += tile_count(nd, I.nP, I.outsize[0], nd > 1 ? I.outsize[1] : 1, ppt, nd > 2 ? I.outsize[2] : 1);
        if (item > INT_MAX) return -CPAB_ERR_UNSUPPORTED;
    }
    if (!item_inst) return (int)item;  // count only
    if (item > max_items) return -CPAB_ERR_WORKSPACE;
    // the items of each backward LDS class together, classes in order
    item = 0;
    for (int cls = 0; cls < BATCH_CLASSES; ++cls)
        for (int i = 0; i < n_inst; ++i) {
            cpab_instance_t& I = inst[i];
            if (batch_class(I) != cls) continue;
            const int nd = I.tess.ndim;
            const int tiles =
                tile_count(nd, I.nP, I.outsize[0], nd > 1 ? I.outsize[1] : 1, ppt, nd > 2 ? I.outsize[2] : 1);
            I.item0 = (int)item;
            I.tiles = tiles;
            for (int k = 0; k < tiles; ++k) item_inst[item + k] = i;
            item += tiles;
        }
    return (int)item;
}

static int batch_check(const cpab_instance_t* h, int n_inst, int ppt, size_t& lds_tab, size_t& lds_acc, int& maxd,
                       int& maxDt, bool bwd) {
    lds_tab = 0;
    lds_acc = 0;
    maxd = 1;
    maxDt = 1;
    for (int i = 0; i < n_inst; ++i) {
        const cpab_instance_t& I = h[i];
        int rc = check_tess(&I.tess);
        if (rc) return rc;
        if (!I.theta || !I.data || !I.lin || !I.out || !I.ws) return CPAB_ERR_ARG;
        if (bwd && (!I.grad_out || !I.grad_theta)) return CPAB_ERR_ARG;
        if (I.ws_bytes < cpab_batch_workspace_bytes(&I.tess, I.outsize, ppt)) return CPAB_ERR_WORKSPACE;
        const int nd = I.tess.ndim;
        const size_t Dt = (size_t)I.tess.nC * nd * (nd + 1);
        const size_t tb = lds_tables(nd, I.tess.nC + I.tess.nV);
        lds_tab = tb > lds_tab ? tb : lds_tab;
        lds_acc = (size_t)WAVES * Dt * 4 > lds_acc ? (size_t)WAVES * Dt * 4 : lds_acc;
        maxd = I.tess.d > maxd ? I.tess.d : maxd;
        maxDt = (int)Dt > maxDt ? (int)Dt : maxDt;
    }
    lds_tab = align256(lds_tab);
    if (lds_tab + (bwd ? lds_acc : 0) > LDS_LIMIT) return CPAB_ERR_LDS;
    return CPAB_OK;
}

int cpab_batch_fwd(const cpab_instance_t* host_inst, const cpab_instance_t* dev_inst, int n_inst,
                   const int* dev_item_inst, int n_items, int ppt, int blocks, float t_end, int* counter,
                   hipStream_t stream) {
    if (!host_inst || !dev_inst || !dev_item_inst || !counter || n_inst < 0 || n_items < 0 || ppt < 1 || blocks < 1)
        return CPAB_ERR_ARG;
    if (n_inst == 0) return CPAB_OK;
    size_t lds_tab, lds_acc;
    int maxd, maxDt;
    int rc = batch_check(host_inst, n_inst, ppt, lds_tab, lds_acc, maxd, maxDt, false);
    if (rc) return rc;
    bool save = true;
    for (int i = 0; i < n_inst; ++i) save = save && host_inst[i].saved != nullptr;
    hipLaunchKernelGGL(k_batch_A, dim3((maxDt + BLOCK - 1) / BLOCK, n_inst), dim3(BLOCK), maxd * sizeof(float), stream,
                       dev_inst);
    hipLaunchKernelGGL(k_counter_reset, dim3(1), dim3(64), 0, stream, counter);
    int maxnd = 1;
    for (int i = 0; i < n_inst; ++i) maxnd = host_inst[i].tess.ndim > maxnd ? host_inst[i].tess.ndim : maxnd;
#define LAUNCH_BF(S, M)                                                                                           \
    hipLaunchKernelGGL((k_batch_fwd<S, M>), dim3(blocks), dim3(BLOCK), lds_tab, stream, dev_inst, dev_item_inst, \
                       n_items, ppt, t_end, counter)
    if (save) {
        if (maxnd == 3) LAUNCH_BF(true, 3); else LAUNCH_BF(true, 2);
    } else {
        if (maxnd == 3) LAUNCH_BF(false, 3); else LAUNCH_BF(false, 2);
    }
#undef LAUNCH_BF
    return status_from_hip();
}

int cpab_batch_bwd(const cpab_instance_t* host_inst, const cpab_instance_t* dev_inst, int n_inst,
                   const int* dev_item_inst, int n_items, int ppt, int blocks, float t_end, int* counter,
                   hipStream_t stream) {
    if (!host_inst || !dev_inst || !dev_item_inst || !counter || n_inst < 0 || n_items < 0 || ppt < 1 || blocks < 1)
        return CPAB_ERR_ARG;
    if (n_inst == 0) return CPAB_OK;
    size_t lds_tab, lds_acc;
    int maxd, maxDt;
    int rc = batch_check(host_inst, n_inst, ppt, lds_tab, lds_acc, maxd, maxDt, true);
    if (rc) return rc;
    bool have = true;
    for (int i = 0; i < n_inst; ++i) have = have && host_inst[i].saved != nullptr;
    // with the forward's saved state, its A_q is still in every instance's
    // workspace (the backward only writes the partials after it)
    if (!have)
        hipLaunchKernelGGL(k_batch_A, dim3((maxDt + BLOCK - 1) / BLOCK, n_inst), dim3(BLOCK), maxd * sizeof(float),
                           stream, dev_inst);
    hipLaunchKernelGGL(k_counter_reset, dim3(1), dim3(64), 0, stream, counter);
    // the item range of each LDS class (contiguous, as cpab_batch_plan lays
    // them out); a plan in another order runs as one class
    int lo[BATCH_CLASSES], hi[BATCH_CLASSES];
    long long cnt[BATCH_CLASSES] = {0, 0, 0};
    for (int c = 0; c < BATCH_CLASSES; ++c) {
        lo[c] = INT_MAX;
        hi[c] = 0;
    }
    for (int i = 0; i < n_inst; ++i) {
        const cpab_instance_t& I = host_inst[i];
        const int c = batch_class(I);
        lo[c] = I.item0 < lo[c] ? I.item0 : lo[c];
        hi[c] = I.item0 + I.tiles > hi[c] ? I.item0 + I.tiles : hi[c];
        cnt[c] += I.tiles;
    }
    bool split = true;
    for (int c = 0; c < BATCH_CLASSES; ++c)
        split = split && (cnt[c] == 0 || (hi[c] - lo[c] == cnt[c] && hi[c] <= n_items));
    for (int c = 0; c < BATCH_CLASSES; ++c) {
        if (split && cnt[c] == 0) continue;
        if (!split && c > 0) break;
        const int item_lo = split ? lo[c] : 0, item_hi = split ? hi[c] : n_items;
        // LDS of this launch: the largest tables and accumulators of its instances
        size_t ltab = 0, lacc = 0;
        bool want_box = false;
        int maxnd = 1;
        for (int i = 0; i < n_inst; ++i) {
            const cpab_instance_t& I = host_inst[i];
            if (split && batch_class(I) != c) continue;
            const int nd = I.tess.ndim;
            maxnd = nd > maxnd ? nd : maxnd;
            const size_t tb = lds_tables(nd, I.tess.nC + I.tess.nV);
            const size_t ab = (size_t)WAVES * I.tess.nC * nd * (nd + 1) * 4;
            ltab = tb > ltab ? tb : ltab;
            lacc = ab > lacc ? ab : lacc;
            want_box = want_box || (nd == 2 && I.grad_data && tile_ty(2, I.outsize[1]) > 0);
        }
        ltab = align256(ltab);
        const int acc_off = (int)(ltab / 4);
        // LDS box for dL/ddata when some 2D instance has a tiled grid and wants it
        const int box_cap = (have && want_box && ltab + lacc + (size_t)BOX_FLOATS * 4 <= LDS_LIMIT) ? BOX_FLOATS : 0;
        const int box_off = (int)((ltab + lacc) / 4);
        const size_t lds = ltab + lacc + (size_t)box_cap * 4;
        int* q = counter + 16 * c;
#define LAUNCH_BB(S, M, CAP)                                                                                      \
    hipLaunchKernelGGL((k_batch_bwd<S, M>), dim3(blocks), dim3(BLOCK), lds, stream, dev_inst, dev_item_inst,     \
                       item_lo, item_hi, ppt, t_end, q, acc_off, box_off, CAP)
        if (have) {
            if (maxnd == 3) LAUNCH_BB(true, 3, box_cap); else LAUNCH_BB(true, 2, box_cap);
        } else {
            if (maxnd == 3) LAUNCH_BB(false, 3, 0); else LAUNCH_BB(false, 2, 0);
        }
    }
#undef LAUNCH_BB
    if ((rc = status_from_hip())) return rc;
    // as many waves (up to 16) as the LDS holds [nw + 1, maxDt] floats for
    int nw = RBLOCK / 64;
    while (nw > 1 && (size_t)(nw + 1) * maxDt * sizeof(float) > LDS_LIMIT) nw >>= 1;
    if ((size_t)(nw + 1) * maxDt * sizeof(float) > LDS_LIMIT) return CPAB_ERR_LDS;
    hipLaunchKernelGGL(k_batch_reduce, dim3(n_inst), dim3(nw * 64), (nw + 1) * maxDt * sizeof(float), stream,
                       dev_inst);
    return status_from_hip();
}

int cpab_batch_fwd_plain(const cpab_instance_t* host_inst, int n_inst, float t_end, hipStream_t stream) {
    if (!host_inst || n_inst < 0) return CPAB_ERR_ARG;
    for (int i = 0; i < n_inst; ++i) {
        const cpab_instance_t& I = host_inst[i];
        int rc = cpab_transform_data_fwd(&I.tess, I.theta, 1, I.data, I.C, I.dsize, I.outsize, I.lin, t_end, I.out,
                                         I.saved, I.ws, I.ws_bytes, stream);
        if (rc) return rc;
    }
    return CPAB_OK;
}

int cpab_batch_bwd_plain(const cpab_instance_t* host_inst, int n_inst, float t_end, hipStream_t stream) {
    if (!host_inst || n_inst < 0) return CPAB_ERR_ARG;
    for (int i = 0; i < n_inst; ++i) {
        const cpab_instance_t& I = host_inst[i];
        int rc = cpab_transform_data_bwd(&I.tess, I.theta, 1, I.data, I.C, I.dsize, I.outsize, I.lin, t_end,
                                         I.grad_out, I.saved, I.grad_data, I.grad_theta, I.ws, I.ws_bytes, stream);
        if (rc) return rc;
    }
    return CPAB_OK;
}

int cpab_fill_zero(float* p, size_t n, hipStream_t stream) {
    if (!p && n) return CPAB_ERR_ARG;
    if (n == 0) return CPAB_OK;
    size_t blocks = (n + BLOCK - 1) / BLOCK;
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)blocks), dim3(BLOCK), 0, stream, p, n);
    return status_from_hip();
}

int cpab_expm_affine(const float* A, int n, int ndim, float scale, float* T, hipStream_t stream) {
    if (ndim < 1 || ndim > 3) return CPAB_ERR_NDIM;
    if (!A || !T || n < 0) return CPAB_ERR_ARG;
    if (n == 0) return CPAB_OK;
    dim3 g((n + BLOCK - 1) / BLOCK);
#define LAUNCH_EX(DD) hipLaunchKernelGGL((k_expm_affine<DD>), g, dim3(BLOCK), 0, stream, A, n, scale, T);
    CPAB_DISPATCH_D(ndim, LAUNCH_EX)
#undef LAUNCH_EX
    return status_from_hip();
}

int cpab_ref_forward(float* newpoints, const float* points, const float* trels, int nstepsolver, const int* nc,
                     int ndim, int nP, int batch_size, int broadcast, hipStream_t stream) {
    if (ndim < 1 || ndim > 3) return CPAB_ERR_NDIM;
    if (!newpoints || !points || !trels || !nc || nstepsolver < 0 || nP < 0 || batch_size < 0) return CPAB_ERR_ARG;
    if (nP == 0 || batch_size == 0) return CPAB_OK;
    const int n0 = nc[0], n1 = ndim > 1 ? nc[1] : 1, n2 = ndim > 2 ? nc[2] : 1;
    const int nC = (ndim == 1 ? 1 : (ndim == 2 ? 4 : 5)) * n0 * n1 * n2;
    const size_t lds = (size_t)nC * ndim * (ndim + 1) * 4;
    if (lds > LDS_LIMIT) return CPAB_ERR_LDS;
    dim3 g((nP + BLOCK - 1) / BLOCK, batch_size);
#define LAUNCH_RF(DD) \
    hipLaunchKernelGGL((k_ref_forward<DD>), g, dim3(BLOCK), lds, stream, newpoints, points, trels, nstepsolver, n0, n1, \
                       n2, nC, nP, broadcast ? 1 : 0);
    CPAB_DISPATCH_D(ndim, LAUNCH_RF)
#undef LAUNCH_RF
    return status_from_hip();
}

int cpab_ref_backward(float* grad, const float* points, const float* As, const float* Bs, int nstepsolver,
                      const int* nc, int n_theta, int d, int ndim, int nP, int nC, int broadcast, hipStream_t stream) {
    if (ndim < 1 || ndim > 3) return CPAB_ERR_NDIM;
    if (!grad || !points || !As || !Bs || !nc || nstepsolver < 1 || n_theta < 0 || d < 0 || nP < 0) return CPAB_ERR_ARG;
    if (nP == 0 || n_theta == 0 || d == 0) return CPAB_OK;
    const int n0 = nc[0], n1 = ndim > 1 ? nc[1] : 1, n2 = ndim > 2 ? nc[2] : 1;
    if (nC != (ndim == 1 ? 1 : (ndim == 2 ? 4 : 5)) * n0 * n1 * n2) return CPAB_ERR_ARG;
    const size_t lds = 2 * (size_t)nC * ndim * (ndim + 1) * 4;
    if (lds > LDS_LIMIT) return CPAB_ERR_LDS;
    if (n_theta > 65535 || d > 65535) return CPAB_ERR_UNSUPPORTED;
    dim3 g((nP + BLOCK - 1) / BLOCK, n_theta, d);
#define LAUNCH_RB(DD) \
    hipLaunchKernelGGL((k_ref_backward<DD>), g, dim3(BLOCK), lds, stream, grad, points, As, Bs, nstepsolver, n0, n1, n2, \
                       n_theta, nC, nP, broadcast ? 1 : 0);
    CPAB_DISPATCH_D(ndim, LAUNCH_RB)
#undef LAUNCH_RB
    return status_from_hip();
}

int cpab_slow_forward(const float* points, const float* trels, int nstepsolver, const int* nc, int ndim, int nP,
                      int n_theta, int per_theta_grid, float* out, hipStream_t stream) {
    if (ndim < 1 || ndim > 3) return CPAB_ERR_NDIM;
    if (!nc || nstepsolver < 0 || nP < 0 || n_theta < 0) return CPAB_ERR_ARG;
    if (nP == 0 || n_theta == 0) return CPAB_OK;
    if (!points || !trels || !out) return CPAB_ERR_ARG;
    for (int a = 0; a < ndim; ++a)
        if (nc[a] < 1) return CPAB_ERR_ARG;
    if (n_theta > 65535) return CPAB_ERR_UNSUPPORTED;
    const int n0 = nc[0], n1 = ndim > 1 ? nc[1] : 1, n2 = ndim > 2 ? nc[2] : 1;
    const int nC = cells_of(ndim, nc);
    const size_t lds = (size_t)nC * ndim * (ndim + 1) * 4;
    if (lds > LDS_LIMIT) return CPAB_ERR_LDS;
    dim3 g((nP + BLOCK - 1) / BLOCK, n_theta);
#define LAUNCH_SF(DD)                                                                                                \
    hipLaunchKernelGGL((k_slow_forward<DD>), g, dim3(BLOCK), lds, stream, points, trels, nstepsolver, n0, n1, n2, nC, \
                       nP, per_theta_grid ? 1 : 0, out);
    CPAB_DISPATCH_D(ndim, LAUNCH_SF)
#undef LAUNCH_SF
    return status_from_hip();
}

int cpab_slow_backward(const float* points, const float* trels, const float* trels_inv, int nstepsolver,
                       const int* nc, int ndim, int nP, int n_theta, int per_theta_grid, const float* grad_out,
                       float* grad_points, float* grad_trels, hipStream_t stream) {
    if (ndim < 1 || ndim > 3) return CPAB_ERR_NDIM;
    if (!nc || nstepsolver < 0 || nP < 0 || n_theta < 0) return CPAB_ERR_ARG;
    if (nP == 0 || n_theta == 0) return CPAB_OK;
    if (!points || !trels || !trels_inv || !grad_out || !grad_trels) return CPAB_ERR_ARG;
    for (int a = 0; a < ndim; ++a)
        if (nc[a] < 1) return CPAB_ERR_ARG;
    if (n_theta > 65535) return CPAB_ERR_UNSUPPORTED;
    const int n0 = nc[0], n1 = ndim > 1 ? nc[1] : 1, n2 = ndim > 2 ? nc[2] : 1;
    const int nC = cells_of(ndim, nc);
    const size_t lds = (size_t)nC * (2 * ndim * (ndim + 1) + ndim * ndim) * 4;
    if (lds > LDS_LIMIT) return CPAB_ERR_LDS;
    dim3 g((nP + BLOCK - 1) / BLOCK, n_theta);
#define LAUNCH_SB(DD)                                                                                                  \
    hipLaunchKernelGGL((k_slow_backward<DD>), g, dim3(BLOCK), lds, stream, points, trels, trels_inv, nstepsolver, n0, \
                       n1, n2, nC, nP, per_theta_grid ? 1 : 0, grad_out, grad_points, grad_trels);
    CPAB_DISPATCH_D(ndim, LAUNCH_SB)
#undef LAUNCH_SB
    return status_from_hip();
}

}  // extern "C"
