// Device-side integrator for continuous piecewise-affine (CPA) velocity
// fields on gfx950.
//
// Coordinates.  All integration runs in scaled coordinates q = x * nc
// (x in [0,1]^D, nc = cells per axis), where the field of cell c is
// q' = Mq q + bq with Mq = S M S^-1, bq = S b, S = diag(nc).  Facets are unit
// planes n.q + e <= 0 (inside), stored per cell with the cell across them.
//
// In-cell flow.  For v = M q + b the exact flow is
//     q(s) = q0 + sum_{k>=1} s^k / k! M^(k-1) v0 ,   v0 = M q0 + b,
// an entire series.  It is evaluated on chunks with |M|_F h <= RHO = 1/2 and
// KS = 7 terms, so the truncation (< RHO^7/8! = 1.9e-7 relative to the chunk
// displacement; 1e-9 at the typical |M|_F h ~ 0.25) is at FP32 resolution; no
// eigenvalue case analysis (real / complex / repeated) and no transcendental
// calls are needed in 2D/3D.
// 1D uses the closed form q(s) = q0 + s*phi1(a s)*v0 and the analytic hit time
// s = (qb-q0)/v0 * log1p(r)/r, r = a (qb-q0)/v0.
//
// Boundary hits.  For each facet G(s) = n.q(s) + e is the same series.  A
// facet is skipped when a quadratic bound on G over the chunk stays negative
// (for a start on the facet moving inwards: when G' stays negative).  Where a
// bound on the series tail certifies the sign of G'' over the chunk, G(h)
// decides the crossing and [0, h] brackets a unique root.  Starts on a facet,
// uncertified curvature and interior maxima are bracketed by sampling G and G'
// at NSAMP points (a local maximum with G > 0 inside a sample interval
// counts).  The candidate with the earliest regula-falsi estimate is solved
// by Newton steps safeguarded by its bracket; another candidate is solved
// only when it is positive at (or its bracket ends before) that root.
//
// Gradients.  The field is continuous across facets, so the sensitivity of the
// final point obeys the variational equation with the saltation matrix
// S = I + (v_next - v_prev) n^T / (n.v_prev) at every crossing (S = I for a
// continuous field; it is applied anyway so the hit-time and cell-change terms
// are part of the derivative).  Pass 1 carries J = dq(T)/dq0; the adjoint at
// t = 0 is mu0 = J^T dL/dq(T).  Pass 2 re-integrates forward carrying
// mu(s) = exp(-s M^T) mu0 and adds int_0^h mu(s) [q(s);1]^T ds to dL/dA_q of
// the cell, an exact double series.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#define CPAB_HD __host__ __device__
#include <stdint.h>
#include <limits.h>
#include <type_traits>

namespace cpab {

constexpr int NF = 4;           // facet slots per cell
constexpr int KS = 7;           // series terms
constexpr float RHO = 0.5f;     // chunk bound |M|_F * h
constexpr float GROW = 1.6488f;   // e^RHO: speed growth bound over a chunk
constexpr int NSAMP = 4;        // samples per chunk for root isolation
constexpr int MAX_ITERS = 512;  // chunks + crossings per point (termination bound; a point that
                                // has time left after them is unfinished: NaN, and counted)
constexpr float ZERO_HOP = 1e-7f;
constexpr int MAX_ZERO_HOPS = 8;
constexpr int NREC = 2;         // chunks per point the forward records for the backward (2D/3D)

template <int D>
struct CellTable {
    const float* A;      // [nC, D, D+1]  q-space affine
    const float* inv_n;  // [nC]          1 / |Mq|_F (inf for Mq = 0)
    const float4* pl;    // [nC, NF]      (n0, n1, n2, e)
    const int* nb;       // [nC, NF]      neighbour, <0: none
};

template <int D>
struct PointState {
    float q[D];
    float T;        // remaining time
    int c;          // current cell
    int came_from;  // cell left by the last crossing
    int zero_hops;  // consecutive crossings of (near) zero length
    int iters;
};

// ---------------------------------------------------------------------------
// small helpers
// ---------------------------------------------------------------------------
// 1/x to ~1 ulp for the safeguarded iterations (their steps are clamped to a
// bracket, so the correctly rounded quotient is not needed)
CPAB_HD __forceinline__ float fast_rcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}
// sqrt(x) to ~1 ulp for the chunk bounds and thresholds (their constants
// carry the margin; the correctly rounded square root costs a dozen more
// instructions per chunk)
CPAB_HD __forceinline__ float fast_sqrt(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}
template <int D>
CPAB_HD __forceinline__ void load_affine(const CellTable<D>& t, int c, float (&M)[D][D], float (&b)[D]) {
    const float* a = t.A + c * (D * (D + 1));
#pragma unroll
    for (int r = 0; r < D; ++r) {
#pragma unroll
        for (int k = 0; k < D; ++k) M[r][k] = a[r * (D + 1) + k];
        b[r] = a[r * (D + 1) + D];
    }
}

template <int D>
CPAB_HD __forceinline__ void affine_apply(const float (&M)[D][D], const float (&b)[D], const float (&q)[D], float (&v)[D]) {
#pragma unroll
    for (int r = 0; r < D; ++r) {
        float acc = b[r];
#pragma unroll
        for (int k = 0; k < D; ++k) acc = fmaf(M[r][k], q[k], acc);
        v[r] = acc;
    }
}

// G(s) and G'(s) of sum_k cf[k] s^k by simultaneous Horner.
CPAB_HD __forceinline__ void horner2(const float (&cf)[KS + 1], float s, float& g, float& d) {
    float gg = cf[KS], dd = 0.f;
#pragma unroll
    for (int k = KS - 1; k >= 0; --k) {
        dd = fmaf(dd, s, gg);
        gg = fmaf(gg, s, cf[k]);
    }
    g = gg;
    d = dd;
}

CPAB_HD __forceinline__ float horner1(const float (&cf)[KS + 1], float s) {
    float gg = cf[KS];
#pragma unroll
    for (int k = KS - 1; k >= 0; --k) gg = fmaf(gg, s, cf[k]);
    return gg;
}

// Root of G in [lo, hi] with G(lo) <= 0 < G(hi): Newton steps from the
// regula-falsi point, safeguarded by bisection.  A lane stops after a Newton
// step below 1e-5 of its bracket (quadratic convergence: the error is then
// ~1e-10) or after 6 steps; its result depends on its own iterates only, not
// on the other lanes of the wave.
CPAB_HD __forceinline__ float solve_bracket(const float (&cf)[KS + 1], float lo, float hi, float glo, float ghi) {
    float s = lo + (hi - lo) * (glo * fast_rcp(glo - ghi));
    if (!(s >= lo && s <= hi)) s = 0.5f * (lo + hi);
    const float tol = 1e-5f * (hi - lo);
    bool fin = false;
#pragma unroll
    for (int it = 0; it < 6; ++it) {
        if (!fin) {
            float g, d;
            horner2(cf, s, g, d);
            if (g > 0.f) hi = s; else lo = s;
            float sn = s - g * fast_rcp(d);
            const bool newton = sn >= lo && sn <= hi;
            if (!newton) sn = 0.5f * (lo + hi);
            fin = newton && fabsf(sn - s) <= tol;
            s = sn;
        }
#if defined(__HIP_DEVICE_COMPILE__)
        if (__ballot(!fin) == 0) break;
#else
        if (fin) break;
#endif
    }
    return s;
}

// Zero of G' in [lo, hi] given G'(lo) and G'(hi) of opposite signs:
// Illinois-modified secant steps inside the shrinking bracket; returns the
// bracket end with the smaller |G'|.
CPAB_HD __forceinline__ float crit_point(const float (&cf)[KS + 1], float lo, float hi, float dlo, float dhi) {
    int side = 0;
#pragma unroll 1
    for (int it = 0; it < 4; ++it) {
        const float m = (lo * dhi - hi * dlo) * fast_rcp(dhi - dlo);
        if (!(m > lo && m < hi)) break;
        float gm, dm;
        horner2(cf, m, gm, dm);
        if ((dm > 0.f) == (dlo > 0.f)) {
            lo = m;
            dlo = dm;
            if (side == -1) dhi *= 0.5f;
            side = -1;
        } else {
            hi = m;
            dhi = dm;
            if (side == 1) dlo *= 0.5f;
            side = 1;
        }
    }
    return (fabsf(dlo) < fabsf(dhi)) ? lo : hi;
}

// Bracket of the first exit through one facet within a chunk.
struct Bracket {
    float lo, hi, glo, ghi;
    int idx;  // 0: exits at the start; i: inside sample interval i; NSAMP+1: none
};

// Isolates the first time in [0, h] at which the trajectory leaves through the
// facet with coefficients cf (G(s) = sum_k cf[k] s^k, G <= 0 inside).
//  * A start on the facet (G(0) >= -tol0, taken as G(0) = 0) moving out
//    exits at 0 when allow_start: G'(0) > tol1, or 0 <= G'(0) <= tol1 (an
//    outward speed within the rounding of the field) with G''(0) > 0.
//  * A start on the facet with G'(0) < 0, however small, first dips below 0;
//    a later positive sample is bracketed from the dip's minimum (zero of
//    G').  Two cells round the normal speed of the same point differently, so
//    a point handed over by a zero-length crossing may see a tangential start
//    here that its previous cell saw as a clear exit: the dip still has to be
//    followed out.
//  * Any other start on the facet (G'(0) >= 0 with the start exit
//    disallowed) is ignored until G has gone negative.
//  * Between two negative samples a local maximum above 0 is a crossing.
CPAB_HD __forceinline__ void first_bracket(const float (&cf)[KS + 1], float h, float tol0, float tol1, bool allow_start,
                                           Bracket& br) {
    br.idx = NSAMP + 1;
    const bool on = cf[0] >= -tol0;
    if (on && allow_start && (cf[1] > tol1 || (cf[1] >= 0.f && cf[2] > 0.f))) {
        br.idx = 0;
        br.lo = br.hi = 0.f;
        return;
    }
    bool skip = on && !(cf[1] < 0.f);
    bool dip = on && !skip;
    float ps = 0.f, pg = cf[0], pd = cf[1];
#pragma unroll 1
    for (int i = 1; i <= NSAMP; ++i) {
        const float s = h * ((float)i * (1.0f / NSAMP));
        float g, d;
        horner2(cf, s, g, d);
        // 0: nothing, 1: dip minimum (G comes back above 0), 2: graze maximum
        int crit = 0;
        if (skip) {
            if (g <= 0.f) skip = false;
        } else if (dip) {
            if (g <= 0.f) dip = false;
            else if (pd < 0.f && d > 0.f) crit = 1;
        } else {
            if (g > 0.f && pg <= 0.f) {
                br.idx = i; br.lo = ps; br.hi = s; br.glo = pg; br.ghi = g;
                return;
            }
            if (pd > 0.f && d < 0.f) crit = 2;
        }
        if (crit) {
            const float m = crit_point(cf, ps, s, pd, d);
            const float gm = horner1(cf, m);
            if (crit == 1 && gm < 0.f) {
                br.idx = i; br.lo = m; br.hi = s; br.glo = gm; br.ghi = g;
                return;
            }
            if (crit == 2 && gm > 0.f) {
                br.idx = i; br.lo = ps; br.hi = m; br.glo = pg; br.ghi = gm;
                return;
            }
            if (crit == 1) {  // never went inside: treat as outside until G < 0
                skip = true;
                dip = false;
            }
        }
        ps = s;
        pg = g;
        pd = d;
    }
}

// Series vectors W[k] = M^k v0 / (k+1)!, so q(s) = q0 + sum_k s^(k+1) W[k].
template <int D>
CPAB_HD __forceinline__ void series(const float (&M)[D][D], const float (&v0)[D], float (&W)[KS][D]) {
#pragma unroll
    for (int a = 0; a < D; ++a) W[0][a] = v0[a];
#pragma unroll
    for (int k = 1; k < KS; ++k) {
        const float inv = 1.0f / (float)(k + 1);
#pragma unroll
        for (int r = 0; r < D; ++r) {
            float acc = 0.f;
#pragma unroll
            for (int a = 0; a < D; ++a) acc = fmaf(M[r][a], W[k - 1][a], acc);
            W[k][r] = acc * inv;
        }
    }
}

template <int D>
CPAB_HD __forceinline__ void series_eval(const float (&q0)[D], const float (&W)[KS][D], float s, float (&q)[D]) {
#pragma unroll
    for (int a = 0; a < D; ++a) {
        float acc = W[KS - 1][a];
#pragma unroll
        for (int k = KS - 2; k >= 0; --k) acc = fmaf(acc, s, W[k][a]);
        q[a] = fmaf(acc, s, q0[a]);
    }
}

// E = exp(N), N = s M, by the Taylor series to N^KS with every power reduced
// by Cayley-Hamilton to the span of {I, N} (2D: N^2 = tr N - det I) or
// {I, N, N^2} (3D: N^3 = tr N^2 - sigma N + det I): scalar recurrences for the
// coordinates of N^k, then one combination.
template <int D>
CPAB_HD __forceinline__ void exp_mat(const float (&M)[D][D], float s, float (&E)[D][D]) {
    float N[D][D];
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int c = 0; c < D; ++c) N[r][c] = s * M[r][c];
    if constexpr (D == 2) {
        const float tr = N[0][0] + N[1][1];
        const float det = fmaf(N[0][0], N[1][1], -N[0][1] * N[1][0]);
        // N^k = a N + b I;  exp(N) = al N + be I
        float a = 1.f, b = 0.f, al = 1.f, be = 1.f, fac = 1.f;
#pragma unroll
        for (int k = 2; k <= KS; ++k) {
            const float an = fmaf(tr, a, b);
            b = -det * a;
            a = an;
            fac *= 1.0f / (float)k;
            al = fmaf(a, fac, al);
            be = fmaf(b, fac, be);
        }
#pragma unroll
        for (int r = 0; r < 2; ++r)
#pragma unroll
            for (int c = 0; c < 2; ++c) E[r][c] = fmaf(al, N[r][c], (r == c) ? be : 0.f);
    } else {
        float N2[3][3];
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                float acc = 0.f;
#pragma unroll
                for (int k = 0; k < 3; ++k) acc = fmaf(N[r][k], N[k][c], acc);
                N2[r][c] = acc;
            }
        const float tr = N[0][0] + N[1][1] + N[2][2];
        const float sg = fmaf(N[0][0], N[1][1], -N[0][1] * N[1][0]) + fmaf(N[0][0], N[2][2], -N[0][2] * N[2][0]) +
                         fmaf(N[1][1], N[2][2], -N[1][2] * N[2][1]);
        const float det = N[0][0] * fmaf(N[1][1], N[2][2], -N[1][2] * N[2][1]) -
                          N[0][1] * fmaf(N[1][0], N[2][2], -N[1][2] * N[2][0]) +
                          N[0][2] * fmaf(N[1][0], N[2][1], -N[1][1] * N[2][0]);
        // N^k = a N^2 + b N + c I, from k = 2;  exp(N) = ga N^2 + be N + al I
        float a = 1.f, b = 0.f, c = 0.f, ga = 0.5f, be = 1.f, al = 1.f, fac = 0.5f;
#pragma unroll
        for (int k = 3; k <= KS; ++k) {
            const float an = fmaf(tr, a, b);
            const float bn = fmaf(-sg, a, c);
            c = det * a;
            a = an;
            b = bn;
            fac *= 1.0f / (float)k;
            ga = fmaf(a, fac, ga);
            be = fmaf(b, fac, be);
            al = fmaf(c, fac, al);
        }
#pragma unroll
        for (int r = 0; r < 3; ++r)
#pragma unroll
            for (int cc = 0; cc < 3; ++cc) E[r][cc] = fmaf(ga, N2[r][cc], fmaf(be, N[r][cc], (r == cc) ? al : 0.f));
    }
}

// Series coefficients of G(s) = n.q(s) + e for facet f of cell c.  A start
// within tol0 of the facet is a start on it: G(0) = 0 exactly, so the sign of
// G just after the start is that of G'(0) (and, for a shallow dip into the
// cell, of the dip) rather than that of the rounding in n.q + e.
template <int D>
CPAB_HD __forceinline__ void facet_coeffs(const CellTable<D>& tab, int c, int f, const float (&q)[D],
                                          const float (&W)[KS][D], float tol0, float (&cf)[KS + 1]) {
    const float4 P = tab.pl[c * NF + f];
    const float n[3] = {P.x, P.y, P.z};
    float g0 = P.w;
#pragma unroll
    for (int a = 0; a < D; ++a) g0 = fmaf(n[a], q[a], g0);
    cf[0] = (g0 >= -tol0) ? 0.f : g0;
#pragma unroll
    for (int k = 0; k < KS; ++k) {
        float acc = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) acc = fmaf(n[a], W[k][a], acc);
        cf[k + 1] = acc;
    }
}

// What one chunk did; handed to the pass callbacks.
template <int D>
struct Chunk {
    float M[D][D];
    float b[D];
    float q0[D];
    float W[KS][D];  // W[0] = v0
    float h;         // time advanced
    int cell;
    int facet;       // crossed facet (-1: none)
    int next;        // cell entered (== cell if no crossing)
    float q1[D];     // end point
};

// End of a chunk of duration step that crosses facet bf (-1: none): end point,
// chunk record and the point's state (ch.M, ch.b and ch.W are set).
template <int D>
CPAB_HD __forceinline__ void finish_chunk(PointState<D>& st, const CellTable<D>& tab, float step, int bf,
                                          Chunk<D>& ch) {
    const int c = st.c;
#pragma unroll
    for (int a = 0; a < D; ++a) ch.q0[a] = st.q[a];
    series_eval<D>(ch.q0, ch.W, step, ch.q1);
    ch.h = step;
    ch.cell = c;
    ch.facet = bf;
    ch.next = (bf >= 0) ? tab.nb[c * NF + bf] : c;
#pragma unroll
    for (int a = 0; a < D; ++a) st.q[a] = ch.q1[a];
    st.T = (bf >= 0) ? fmaxf(st.T - step, 0.f) : st.T - step;
    if (bf < 0 && st.T < 0.f) st.T = 0.f;
    if (bf >= 0) {
        st.zero_hops = (step <= ZERO_HOP) ? st.zero_hops + 1 : 0;
        st.came_from = c;
        st.c = ch.next;
    } else if (step > ZERO_HOP) {
        // time went by inside the cell: the guard against crossing straight
        // back only covers a run of zero-length crossings
        st.zero_hops = 0;
    }
    st.iters++;
}

// The first NREC chunks of a point's trajectory: durations and crossed facets.
// code: bits 0-1 chunks recorded; bits 2+3i..4+3i facet crossed by chunk i, + 1
// (0: none).
struct ChunkLog {
    float h[NREC];
    int code;
};

CPAB_HD __forceinline__ void log_chunk(ChunkLog& lg, float h, int facet) {
    const int n = lg.code & 3;
#pragma unroll
    for (int k = 0; k < NREC; ++k)
        if (k == n) lg.h[k] = h;
    if (n < NREC) lg.code = ((lg.code & ~3) | (n + 1)) | ((facet + 1) << (2 + 3 * n));
}

// Sink for the forward's chunk log: durations go straight to memory (h[k] of
// point i at base[k * stride + i], base common to the wave) as the chunks end;
// only the facet code stays in a register.  On the device the store is a raw
// buffer store with a 32-bit lane offset, bounded by the log rows.
struct LogSink {
    float* base;
    int stride;
    int i;
    int code;
#if defined(__HIP_DEVICE_COMPILE__)
    __amdgpu_buffer_rsrc_t rsrc;
#endif
    CPAB_HD void init(float* b, int s, int idx) {
        base = b;
        stride = s;
        i = idx;
        code = 0;
#if defined(__HIP_DEVICE_COMPILE__)
        rsrc = __builtin_amdgcn_make_buffer_rsrc(b, 0, NREC * s * 4, 0x00020000);
#endif
    }
};

// n: index of the chunk just finished (the point's chunk count - 1).
CPAB_HD __forceinline__ void sink_chunk(LogSink& lg, int n, float h, int facet) {
    if (n < NREC) {
#if defined(__HIP_DEVICE_COMPILE__)
        // the builtin's data operand is an integer: pass the bits, not the value
        __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(h), lg.rsrc, (n * lg.stride + lg.i) * 4, 0, 0);
#else
        lg.base[(size_t)n * lg.stride + lg.i] = h;
#endif
        lg.code = ((lg.code & ~3) | (n + 1)) | ((facet + 1) << (2 + 3 * n));
    }
}

CPAB_HD __forceinline__ void logged_chunk(const ChunkLog& lg, int k, float& h, int& facet) {
    h = lg.h[0];
#pragma unroll
    for (int j = 1; j < NREC; ++j)
        if (k == j) h = lg.h[j];
    facet = ((lg.code >> (2 + 3 * k)) & 7) - 1;
}

// ---------------------------------------------------------------------------
// one chunk of integration, D = 2, 3
// ---------------------------------------------------------------------------
template <int D>
CPAB_HD __forceinline__ void advance(PointState<D>& st, const CellTable<D>& tab, float tol0, Chunk<D>& ch) {
    const int c = st.c;
    load_affine<D>(tab, c, ch.M, ch.b);
    float v0[D];
    affine_apply<D>(ch.M, ch.b, st.q, v0);
    float h = fminf(st.T, RHO * tab.inv_n[c]);
    series<D>(ch.M, v0, ch.W);
    float vn2 = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) vn2 = fmaf(v0[a], v0[a], vn2);
    const float vn = fast_sqrt(vn2);
    // Tangential band for an outward normal speed G'(0): the neighbouring
    // cells' FP32 fields agree on a shared facet only up to the rounding of
    // M q + b, whose terms can be much larger than |v0| where they cancel, so
    // both may round a vanishing normal speed to a small exit.  Inside the
    // band a start exit also needs G'' > 0 (the same side from both cells for
    // a tangent velocity); an inward G'(0) always follows the dip.
    float vabs = 0.f;
#pragma unroll
    for (int r = 0; r < D; ++r) {
        float t = fabsf(ch.b[r]);
#pragma unroll
        for (int a = 0; a < D; ++a) t = fmaf(fabsf(ch.M[r][a]), fabsf(st.q[a]), t);
        vabs += t;
    }
    const float tol1 = 1e-6f * vabs;
    // A velocity inside that band is rounding noise: the point sits on a
    // fixed point of the field at FP32 resolution (e.g. a domain corner under
    // zero_boundary) and moves by at most e^RHO |v0| h over the chunk.  It
    // crosses no facet in this chunk (otherwise the noise could carry it
    // back and forth across a facet through the fixed point, a hop per
    // chunk, until the chunk budget ran out).
    const bool still = vn <= tol1;
    // Bounds on the chunk (|n| = 1, |M| <= |M|_F, |M|_F h <= RHO, |cf_k| <=
    // |M|^(k-1) |v0| / k!):  the terms of G from s^3 on sum to at most
    // K3 s^3, K3 = |M|_F^2 |v0| e^RHO / 6, their derivative to 3 K3 s^2, and
    // |G''(s) - 2 cf2 - 6 cf3 s| <= K4 = |M|_F^3 |v0| h^2 e^RHO / 2.
    const float nrm = fast_rcp(tab.inv_n[c]);  // |M|_F (the bounds below carry margin for its last bit)
    const float mh = h * nrm;
    const float K = GROW * vn * nrm;
    const float K4 = 0.5f * K * mh * mh;
    const float K3 = K * nrm * (1.0f / 6.0f);

    // phase 1: classify every facet.  A facet whose first crossing lies in a
    // bracket with a unique root becomes a candidate; starts on a facet, and
    // facets whose curvature sign is not certified, take the sampling path.
    int start_f = -1;
    unsigned cand = 0u, slow = 0u;
    // bracket of each candidate facet (set only where cand has its bit)
    float clo[NF], chi[NF], cglo[NF], cghi[NF];
#pragma unroll
    for (int f = 0; f < NF; ++f) {
        const int nbf = tab.nb[c * NF + f];
        if (nbf < 0 || still) continue;
        const float4 P = tab.pl[c * NF + f];
        const float n[3] = {P.x, P.y, P.z};
        float g0 = P.w, g1 = 0.f, g2 = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) {
            g0 = fmaf(n[a], st.q[a], g0);
            g1 = fmaf(n[a], v0[a], g1);
        }
#pragma unroll
        for (int a = 0; a < D; ++a) g2 = fmaf(n[a], ch.W[1][a], g2);
        const bool on = g0 >= -tol0;
        // G(s) <= g0 + g1 s + max(cf2, 0) s^2 + K3 s^3 on [0, h], a convex bound
        // (its s^2 and s^3 coefficients are >= 0), so no crossing when it is
        // below 0 at h
        const float c2p = fmaxf(g2, 0.f);
        if (!on && fmaf(h, fmaf(h, fmaf(h, K3, c2p), g1), g0) < 0.f) continue;
        // start on the facet moving inwards with G' <= g1 + 2 max(cf2, 0) s +
        // 3 K3 s^2 < 0 on all of [0, h]: no return
        if (on && g1 < -tol1 && fmaf(h, fmaf(h, 3.f * K3, 2.f * c2p), g1) < 0.f) continue;
        if (on) {
            const bool allow_start = !(st.zero_hops > 0 && nbf == st.came_from) && st.zero_hops < MAX_ZERO_HOPS;
            if (allow_start && (g1 > tol1 || (g1 >= 0.f && g2 > 0.f))) {
                if (start_f < 0) start_f = f;
            } else {
                slow |= 1u << f;
            }
            continue;
        }
        float cf[KS + 1];
        cf[0] = g0;
        cf[1] = g1;
        cf[2] = g2;
#pragma unroll
        for (int k = 2; k < KS; ++k) {
            float acc = 0.f;
#pragma unroll
            for (int a = 0; a < D; ++a) acc = fmaf(n[a], ch.W[k][a], acc);
            cf[k + 1] = acc;
        }
        // G''(s) = 2 cf2 + 6 cf3 s + R(s), |R| <= K4: the sign of G'' is
        // certified when the linear part keeps one sign with margin K4
        const float l0 = 2.f * g2, l1 = fmaf(6.f * cf[3], h, l0);
        if (!((l0 > K4 && l1 > K4) || (l0 < -K4 && l1 < -K4))) {
            slow |= 1u << f;
            continue;
        }
        // G convex (g2 > 0) or concave (g2 < 0) on [0, h], G(0) < 0: a positive
        // G(h) means exactly one crossing; otherwise only a concave G rising
        // and falling inside the chunk can cross (sampling path).
        float gh, dh;
        horner2(cf, h, gh, dh);
        if (gh > 0.f) {
            cand |= 1u << f;
            clo[f] = 0.f;
            chi[f] = h;
            cglo[f] = g0;
            cghi[f] = gh;
        } else if (g2 < 0.f && g1 > 0.f && dh < 0.f) {
            slow |= 1u << f;
        }
    }
    // sampling path (starts on a facet, uncertified curvature, interior maxima)
    if (start_f < 0) {
#pragma unroll 1
        while (slow) {
            const int f = __builtin_ctz(slow);
            slow &= slow - 1u;
            float cf[KS + 1];
            facet_coeffs<D>(tab, c, f, st.q, ch.W, tol0, cf);
            const int nbf = tab.nb[c * NF + f];
            const bool allow_start = !(st.zero_hops > 0 && nbf == st.came_from) && st.zero_hops < MAX_ZERO_HOPS;
            Bracket br;
            first_bracket(cf, h, tol0, tol1, allow_start, br);
            if (br.idx >= 1 && br.idx <= NSAMP) {
                cand |= 1u << f;
#pragma unroll
                for (int k = 0; k < NF; ++k)
                    if (k == f) {
                        clo[k] = br.lo;
                        chi[k] = br.hi;
                        cglo[k] = br.glo;
                        cghi[k] = br.ghi;
                    }
            }
        }
    }
    // phase 2: solve the candidate with the earliest regula-falsi estimate, then
    // any other candidate that is still positive at (or ends before) that root
    float best = h;
    int bf = -1;
    if (start_f >= 0) {
        best = 0.f;
        bf = start_f;
    } else if (cand) {
        float est = INFINITY, lo = 0.f, hi = 0.f, glo = 0.f, ghi = 0.f;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
            if (!((cand >> f) & 1u)) continue;
            const float e = fmaf(chi[f] - clo[f], cglo[f] * fast_rcp(cglo[f] - cghi[f]), clo[f]);
            if (e < est) {
                est = e;
                bf = f;
                lo = clo[f];
                hi = chi[f];
                glo = cglo[f];
                ghi = cghi[f];
            }
        }
        float cf[KS + 1];
        facet_coeffs<D>(tab, c, bf, st.q, ch.W, tol0, cf);
        best = solve_bracket(cf, lo, hi, glo, ghi);
        unsigned others = cand & ~(1u << bf);
#pragma unroll 1
        while (others) {
            const int f = __builtin_ctz(others);
            others &= others - 1u;
#pragma unroll
            for (int k = 0; k < NF; ++k)
                if (k == f) {
                    lo = clo[k];
                    hi = chi[k];
                    glo = cglo[k];
                    ghi = cghi[k];
                }
            if (!(lo < best)) continue;
            facet_coeffs<D>(tab, c, f, st.q, ch.W, tol0, cf);
            if (best < hi) {
                hi = best;
                ghi = horner1(cf, best);
            }
            if (ghi > 0.f) {
                const float s2 = solve_bracket(cf, lo, hi, glo, ghi);
                if (s2 < best) {
                    best = s2;
                    bf = f;
                }
            }
        }
    }
    finish_chunk<D>(st, tab, best, bf, ch);
}

// Replays a chunk whose duration h and crossed facet bf (-1: none) are known
// (recorded by the forward): the same series and the same end-of-chunk update
// as advance, without the facet search, so the state comes out bit for bit
// as in the forward.
template <int D>
CPAB_HD __forceinline__ void replay(PointState<D>& st, const CellTable<D>& tab, float h, int bf, Chunk<D>& ch) {
    load_affine<D>(tab, st.c, ch.M, ch.b);
    float v0[D];
    affine_apply<D>(ch.M, ch.b, st.q, v0);
    series<D>(ch.M, v0, ch.W);
    finish_chunk<D>(st, tab, h, bf, ch);
}

// ---------------------------------------------------------------------------
// 1D: closed form
// ---------------------------------------------------------------------------
CPAB_HD __forceinline__ float phi1(float z) {  // (e^z - 1)/z
    return (fabsf(z) < 1e-4f) ? fmaf(z, fmaf(z, 1.0f / 6.0f, 0.5f), 1.0f) : expm1f(z) / z;
}
CPAB_HD __forceinline__ float phi2(float z) {  // (e^z - 1 - z)/z^2
    if (fabsf(z) < 0.5f) {
        float acc = 1.0f / 3628800.0f;  // 1/10!
        acc = fmaf(acc, z, 1.0f / 362880.0f);
        acc = fmaf(acc, z, 1.0f / 40320.0f);
        acc = fmaf(acc, z, 1.0f / 5040.0f);
        acc = fmaf(acc, z, 1.0f / 720.0f);
        acc = fmaf(acc, z, 1.0f / 120.0f);
        acc = fmaf(acc, z, 1.0f / 24.0f);
        acc = fmaf(acc, z, 1.0f / 6.0f);
        acc = fmaf(acc, z, 0.5f);
        return acc;
    }
    return (expm1f(z) - z) / (z * z);
}
CPAB_HD __forceinline__ float log1p_over(float r) {  // log1p(r)/r
    return (fabsf(r) < 1e-4f) ? fmaf(r, fmaf(r, 1.0f / 3.0f, -0.5f), 1.0f) : log1pf(r) / r;
}

struct Chunk1 {
    float a, b, q0, v0, h;
    int cell, facet, next;
    float q1;
};

CPAB_HD __forceinline__ void advance1(PointState<1>& st, const CellTable<1>& tab, int nx, Chunk1& ch) {
    const int c = st.c;
    const float a = tab.A[2 * c], b = tab.A[2 * c + 1];
    const float q0 = st.q[0];
    const float v0 = fmaf(a, q0, b);
    float step = st.T;
    int facet = -1, next = c;
    if (v0 != 0.f) {
        const bool right = v0 > 0.f;
        const int nbc = right ? c + 1 : c - 1;
        if (nbc >= 0 && nbc < nx) {
            const float qb = right ? (float)(c + 1) : (float)c;
            const float dq = qb - q0;
            const float r = a * dq / v0;
            if (1.0f + r > 0.f) {
                float th = (dq / v0) * log1p_over(r);
                th = fmaxf(th, 0.f);
                bool back = (st.zero_hops > 0 && nbc == st.came_from && th <= ZERO_HOP);
                if (th < st.T && !back) {
                    step = th;
                    facet = right ? 1 : 0;
                    next = nbc;
                }
            }
        }
    }
    ch.a = a; ch.b = b; ch.q0 = q0; ch.v0 = v0; ch.h = step;
    ch.cell = c; ch.facet = facet; ch.next = next;
    if (facet >= 0) {
        ch.q1 = (facet == 1) ? (float)(c + 1) : (float)c;
        st.T = fmaxf(st.T - step, 0.f);
        st.zero_hops = (step <= ZERO_HOP) ? st.zero_hops + 1 : 0;
        st.came_from = c;
        st.c = next;
    } else {
        ch.q1 = fmaf(step * phi1(a * step), v0, q0);
        st.T = 0.f;
    }
    st.q[0] = ch.q1;
    st.iters++;
}

// ---------------------------------------------------------------------------
// initial cell (the reference's lookup, libcpab/numpy/findcellidx.py, in FP32)
// ---------------------------------------------------------------------------
CPAB_HD __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

CPAB_HD __forceinline__ int find_cell1(float x, int nx) {
    float f = floorf(x * (float)nx);
    f = fminf(fmaxf(f, 0.f), (float)(nx - 1));
    return (int)f;
}

CPAB_HD __forceinline__ int find_cell2(float x, float y, int nx, int ny) {
    const float qx = x * (float)nx, qy = y * (float)ny;
    const int i = clampi((int)floorf(fminf(fmaxf(qx, 0.f), (float)nx)), 0, nx - 1);
    const int j = clampi((int)floorf(fminf(fmaxf(qy, 0.f), (float)ny)), 0, ny - 1);
    const int base = 4 * (i + j * nx);
    if (x <= 0.f) {
        if (y <= 0.f && qy < qx) return base;
        if (y >= 1.f && qy - (float)ny > -qx) return base + 2;
        return base + 3;
    }
    if (x >= 1.f) {
        if (y <= 0.f && -qy > qx - (float)nx) return base;
        if (y >= 1.f && qy - (float)ny > qx - (float)nx) return base + 2;
        return base + 1;
    }
    if (y <= 0.f) return base;
    if (y >= 1.f) return base + 2;
    const float fx = qx - (float)i, fy = qy - (float)j;
    if (fx < fy) return base + ((1.f - fx < fy) ? 2 : 3);
    return base + ((1.f - fx < fy) ? 1 : 0);
}

// top: largest coordinate the lookup keeps (libcpab/numpy/findcellidx.py
// clamps to 1 - 1e-8, which is 1 in FP32; the torch lookup of the reference's
// use_slow path clamps to 1 - 1e-4, libcpab/pytorch/findcellidx.py:112-114)
CPAB_HD __forceinline__ int find_cell3(float x, float y, float z, int nx, int ny, int nz, float top = 1.f) {
    float p[3] = {x, y, z};
    const int n[3] = {nx, ny, nz};
    const bool outside = x < 0.f || x > 1.f || y < 0.f || y > 1.f || z < 0.f || z > 1.f;
    if (outside) {
        // libcpab/numpy/findcellidx.py:85-108: push the point onto the boundary
        const float half = 0.5f;
        float d[3] = {x - half, y - half, z - half};
        const float ax = fabsf(d[0]), ay = fabsf(d[1]), az = fabsf(d[2]);
        const float push[3] = {(ax < ay && ax < az) ? half / nx : 0.f, (ay < ax && ax < az) ? half / ny : 0.f,
                               (az < ax && ax < ay) ? half / nz : 0.f};
        const float ab[3] = {ax, ay, az};
#pragma unroll
        for (int a = 0; a < 3; ++a)
            if (ab[a] > half) d[a] = copysignf(half - push[a], d[a]);
#pragma unroll
        for (int a = 0; a < 3; ++a) p[a] = d[a] + half;
    }
    int ijk[3];
    float loc[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float q = fminf(fmaxf(p[a], 0.f), top) * (float)n[a];
        ijk[a] = clampi((int)floorf(q), 0, n[a] - 1);
        loc[a] = q - (float)ijk[a];
    }
    int idx = 5 * (ijk[0] + ijk[1] * nx + ijk[2] * nx * ny);
    float lx = loc[0], ly = loc[1], lz = loc[2];
    if ((ijk[0] + ijk[1] + ijk[2]) & 1) {
        const float t = lx;
        lx = ly;
        ly = 1.f - t;
    }
    if (-lx - ly + lz >= 0.f) idx += 1;
    else if (lx + ly + lz - 2.f >= 0.f) idx += 2;
    else if (-lx + ly - lz >= 0.f) idx += 3;
    else if (lx - ly - lz >= 0.f) idx += 4;
    return idx;
}

template <int D>
CPAB_HD __forceinline__ int find_cell(const float (&x)[D], const int* nc) {
    if constexpr (D == 1) return find_cell1(x[0], nc[0]);
    else if constexpr (D == 2) return find_cell2(x[0], x[1], nc[0], nc[1]);
    else return find_cell3(x[0], x[1], x[2], nc[0], nc[1], nc[2]);
}

// ---------------------------------------------------------------------------
// per-point integration with optional Jacobian (pass 1)
// ---------------------------------------------------------------------------
template <int D>
CPAB_HD __forceinline__ void saltation_fwd(const CellTable<D>& tab, const Chunk<D>& ch, float (&J)[D][D]) {
    // J <- S J,  S = I + (v_next - v_prev) n^T / (n . v_prev)
    float Mn[D][D], bn[D], vp[D], vn[D];
    load_affine<D>(tab, ch.next, Mn, bn);
    affine_apply<D>(ch.M, ch.b, ch.q1, vp);
    affine_apply<D>(Mn, bn, ch.q1, vn);
    const float4 P = tab.pl[ch.cell * NF + ch.facet];
    const float n[3] = {P.x, P.y, P.z};
    float nv = 0.f, vv = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        nv = fmaf(n[a], vp[a], nv);
        vv = fmaf(vp[a], vp[a], vv);
    }
    if (fabsf(nv) <= 1e-6f * fast_sqrt(vv)) return;
    const float inv = fast_rcp(nv);
#pragma unroll
    for (int j = 0; j < D; ++j) {
        float nj = 0.f;
#pragma unroll
        for (int a = 0; a < D; ++a) nj = fmaf(n[a], J[a][j], nj);
#pragma unroll
        for (int r = 0; r < D; ++r) J[r][j] = fmaf((vn[r] - vp[r]) * inv, nj, J[r][j]);
    }
}

template <int D>
CPAB_HD __forceinline__ void saltation_adj(const CellTable<D>& tab, const Chunk<D>& ch, float (&mu)[D]) {
    // mu <- S^-T mu = mu - n ((v_next - v_prev) . mu) / (n . v_next)
    float Mn[D][D], bn[D], vp[D], vn[D];
    load_affine<D>(tab, ch.next, Mn, bn);
    affine_apply<D>(ch.M, ch.b, ch.q1, vp);
    affine_apply<D>(Mn, bn, ch.q1, vn);
    const float4 P = tab.pl[ch.cell * NF + ch.facet];
    const float n[3] = {P.x, P.y, P.z};
    float nv = 0.f, vv = 0.f, um = 0.f;
#pragma unroll
    for (int a = 0; a < D; ++a) {
        nv = fmaf(n[a], vn[a], nv);
        vv = fmaf(vn[a], vn[a], vv);
        um = fmaf(vn[a] - vp[a], mu[a], um);
    }
    if (fabsf(nv) <= 1e-6f * fast_sqrt(vv)) return;
    const float f = um * fast_rcp(nv);
#pragma unroll
    for (int a = 0; a < D; ++a) mu[a] = fmaf(-n[a], f, mu[a]);
}

template <int D>
CPAB_HD __forceinline__ void init_point(PointState<D>& st, const float (&q)[D], int c0, float T) {
#pragma unroll
    for (int a = 0; a < D; ++a) st.q[a] = q[a];
    st.T = T;
    st.c = c0;
    st.came_from = -1;
    st.zero_hops = 0;
    st.iters = 0;
}

template <int D>
CPAB_HD __forceinline__ bool point_active(const PointState<D>& st) {
    return st.T > 0.f && st.iters < MAX_ITERS;
}

// One chunk of pass 1: advance the point and carry J = dq(T)/dq0; with LOG
// (2D/3D) the chunk goes into lg (a ChunkLog or a LogSink).
template <int D, bool WITHJ, bool LOG = false, class Log = ChunkLog>
CPAB_HD __forceinline__ void step_point(PointState<D>& st, const CellTable<D>& tab, int nx, float tol0,
                                        float (&J)[D][D], Log& lg) {
    if constexpr (D == 1) {
        Chunk1 ch;
        advance1(st, tab, nx, ch);
        if constexpr (WITHJ) {
            J[0][0] *= expf(ch.a * ch.h);
            if (ch.facet >= 0) {
                const float vn = fmaf(tab.A[2 * ch.next], ch.q1, tab.A[2 * ch.next + 1]);
                const float vp = fmaf(ch.a, ch.q1, ch.b);
                if (fabsf(vp) > 1e-6f * fabsf(vn) && vp != 0.f) J[0][0] *= vn / vp;
            }
        }
    } else {
        Chunk<D> ch;
        advance<D>(st, tab, tol0, ch);
        if constexpr (WITHJ) {
            float E[D][D], Jn[D][D];
            exp_mat<D>(ch.M, ch.h, E);
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int j = 0; j < D; ++j) {
                    float acc = 0.f;
#pragma unroll
                    for (int a = 0; a < D; ++a) acc = fmaf(E[r][a], J[a][j], acc);
                    Jn[r][j] = acc;
                }
#pragma unroll
            for (int r = 0; r < D; ++r)
#pragma unroll
                for (int j = 0; j < D; ++j) J[r][j] = Jn[r][j];
            if (ch.facet >= 0) saltation_fwd<D>(tab, ch, J);
        }
        // logged last, when the chunk's temporaries are dead
        if constexpr (LOG) {
            if constexpr (std::is_same<Log, LogSink>::value)
                sink_chunk(lg, st.iters - 1, ch.h, ch.facet);
            else
                log_chunk(lg, ch.h, ch.facet);
        }
    }
}

template <int D>
CPAB_HD __forceinline__ void identity(float (&J)[D][D]) {
#pragma unroll
    for (int r = 0; r < D; ++r)
#pragma unroll
        for (int k = 0; k < D; ++k) J[r][k] = (r == k) ? 1.f : 0.f;
}

// Integrate one point for time T (pass 1 if WITHJ); with LOG its first NREC
// chunks go into lg for the backward to replay.  A point that uses up
// MAX_ITERS with time left ends as NaN (q and J) rather than short of its end,
// and adds 1 to *unfinished (nullable).  The check sits in the rare branch
// only, so the common path carries no extra live state.
template <int D, bool WITHJ, bool LOG = false, class Log = ChunkLog>
CPAB_HD __forceinline__ void integrate(float (&q)[D], int c0, float T, const CellTable<D>& tab, int nx, float tol0,
                                       float (&J)[D][D], Log& lg, int* unfinished = nullptr) {
    PointState<D> st;
    init_point<D>(st, q, c0, T);
    if constexpr (WITHJ) identity<D>(J);
    lg.code = 0;
    while (st.T > 0.f) {
        step_point<D, WITHJ, LOG, Log>(st, tab, nx, tol0, J, lg);
        if (st.iters >= MAX_ITERS && st.T > 0.f) {
            const float nan = __builtin_nanf("");
            st.T = 0.f;
#pragma unroll
            for (int a = 0; a < D; ++a) st.q[a] = nan;
            if constexpr (WITHJ) {
#pragma unroll
                for (int r = 0; r < D; ++r)
#pragma unroll
                    for (int k = 0; k < D; ++k) J[r][k] = nan;
            }
            if (unfinished) {
#if defined(__HIP_DEVICE_COMPILE__)
                // the operand is made opaque so that its register is set here,
                // on the rare path, and not hoisted out of the loop
                int one = 1;
                asm volatile("" : "+v"(one));
                atomicAdd(unfinished, one);
#else
                ++*unfinished;
#endif
            }
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) q[a] = st.q[a];
}

template <int D, bool WITHJ>
CPAB_HD __forceinline__ void integrate(float (&q)[D], int c0, float T, const CellTable<D>& tab, int nx, float tol0,
                                       float (&J)[D][D], int* unfinished = nullptr) {
    ChunkLog lg;
    integrate<D, WITHJ, false>(q, c0, T, tab, nx, tol0, J, lg, unfinished);
}

// ---------------------------------------------------------------------------
// pass 2: gradient contribution of one chunk
// ---------------------------------------------------------------------------
template <int D>
CPAB_HD __forceinline__ void chunk_grad(const Chunk<D>& ch, float (&mu)[D], float (&G)[D * (D + 1)]) {
    const float h = ch.h;
    float hp[KS + 2];
    hp[0] = 1.f;
#pragma unroll
    for (int r = 1; r < KS + 2; ++r) hp[r] = hp[r - 1] * h;
    float m[D], mend[D];
#pragma unroll
    for (int a = 0; a < D; ++a) {
        m[a] = mu[a];
        mend[a] = 0.f;
    }
#pragma unroll
    for (int p = 0; p < D * (D + 1); ++p) G[p] = 0.f;
#pragma unroll
    for (int j = 0; j <= KS; ++j) {
        // Y_j = sum_{k=0}^{KS-j} h^(j+k+1)/(j+k+1) X_k,  X_0 = [q0;1], X_k = [W[k-1];0]
        const float w0 = hp[j + 1] * (1.0f / (float)(j + 1));
        float y[D];
#pragma unroll
        for (int a = 0; a < D; ++a) y[a] = w0 * ch.q0[a];
#pragma unroll
        for (int k = 1; k <= KS - j; ++k) {
            const float wk = hp[j + k + 1] * (1.0f / (float)(j + k + 1));
#pragma unroll
            for (int a = 0; a < D; ++a) y[a] = fmaf(wk, ch.W[k - 1][a], y[a]);
        }
        const float sg = (j & 1) ? -1.f : 1.f;
#pragma unroll
        for (int r = 0; r < D; ++r) {
            const float mr = sg * m[r];
#pragma unroll
            for (int a = 0; a < D; ++a) G[r * (D + 1) + a] = fmaf(mr, y[a], G[r * (D + 1) + a]);
            G[r * (D + 1) + D] = fmaf(mr, w0, G[r * (D + 1) + D]);
            mend[r] = fmaf(sg * hp[j], m[r], mend[r]);
        }
        if (j < KS) {
            const float inv = 1.0f / (float)(j + 1);
            float t[D];
#pragma unroll
            for (int r = 0; r < D; ++r) {
                float acc = 0.f;
#pragma unroll
                for (int a = 0; a < D; ++a) acc = fmaf(ch.M[a][r], m[a], acc);
                t[r] = acc * inv;
            }
#pragma unroll
            for (int r = 0; r < D; ++r) m[r] = t[r];
        }
    }
#pragma unroll
    for (int a = 0; a < D; ++a) mu[a] = mend[a];
}

CPAB_HD __forceinline__ void chunk_grad1(const Chunk1& ch, float& mu, float (&G)[2]) {
    const float z = -ch.a * ch.h;
    G[0] = mu * fmaf(ch.h * ch.h, ch.b * phi2(z), ch.q0 * ch.h);
    G[1] = mu * ch.h * phi1(z);
    mu *= expf(z);
}

// ---------------------------------------------------------------------------
// wave-level deterministic accumulation of per-lane cell contributions
// ---------------------------------------------------------------------------
// One DPP step: every lane receives v of the lane CTRL selects (0 where there
// is none or where ROWS masks the row out).
template <int CTRL, int ROWS>
__device__ __forceinline__ float dpp_take(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xF, true));
}

// Sum over the 64 lanes in a fixed order: inclusive scans inside each row of
// 16 (row_shr 1, 2, 4, 8), then the row totals carried by row_bcast:15 and
// row_bcast:31; lane 63 ends with the total, read into a scalar.
__device__ __forceinline__ float wave_total(float v) {
    v += dpp_take<0x111, 0xF>(v);
    v += dpp_take<0x112, 0xF>(v);
    v += dpp_take<0x114, 0xF>(v);
    v += dpp_take<0x118, 0xF>(v);
    v += dpp_take<0x142, 0xA>(v);
    v += dpp_take<0x143, 0xC>(v);
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), 63));
}

// Inclusive prefix sum over the 64 lanes (lane i gets lanes 0..i), same DPP
// pattern as wave_total.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_take_int(int v) {
    return __builtin_amdgcn_update_dpp(0, v, CTRL, ROWS, 0xF, true);
}
__device__ __forceinline__ int wave_inclusive_scan(int v) {
    v += dpp_take_int<0x111, 0xF>(v);
    v += dpp_take_int<0x112, 0xF>(v);
    v += dpp_take_int<0x114, 0xF>(v);
    v += dpp_take_int<0x118, 0xF>(v);
    v += dpp_take_int<0x142, 0xA>(v);
    v += dpp_take_int<0x143, 0xC>(v);
    return v;
}

// Minimum / maximum over the 64 lanes (same DPP pattern; lanes without a
// source keep the identity), read into a scalar.
template <int CTRL, int ROWS>
__device__ __forceinline__ int dpp_keep(int ident, int v) {
    return __builtin_amdgcn_update_dpp(ident, v, CTRL, ROWS, 0xF, false);
}
__device__ __forceinline__ int wave_min(int v) {
    v = min(v, dpp_keep<0x111, 0xF>(INT_MAX, v));
    v = min(v, dpp_keep<0x112, 0xF>(INT_MAX, v));
    v = min(v, dpp_keep<0x114, 0xF>(INT_MAX, v));
    v = min(v, dpp_keep<0x118, 0xF>(INT_MAX, v));
    v = min(v, dpp_keep<0x142, 0xA>(INT_MAX, v));
    v = min(v, dpp_keep<0x143, 0xC>(INT_MAX, v));
    return __builtin_amdgcn_readlane(v, 63);
}
__device__ __forceinline__ int wave_max(int v) {
    v = max(v, dpp_keep<0x111, 0xF>(INT_MIN, v));
    v = max(v, dpp_keep<0x112, 0xF>(INT_MIN, v));
    v = max(v, dpp_keep<0x114, 0xF>(INT_MIN, v));
    v = max(v, dpp_keep<0x118, 0xF>(INT_MIN, v));
    v = max(v, dpp_keep<0x142, 0xA>(INT_MIN, v));
    v = max(v, dpp_keep<0x143, 0xC>(INT_MIN, v));
    return __builtin_amdgcn_readlane(v, 63);
}

// One DPP step on every row: adds the value of the lane CTRL selects (0
// where there is none), a single add with a DPP operand.
template <int CTRL>
__device__ __forceinline__ float dpp_add(float v) {
    return v + __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, 0xF, 0xF, true));
}

// wave_total of P values at once, left in lane 63: every step is taken for
// all P before the next, so the steps of one value do not wait on each
// other.  The two row broadcasts run on all rows; the rows wave_total masks
// out (0 and 2, then 0 and 1) end with other sums, but lane 63 -- which
// reads only the rows those steps keep -- gets wave_total's sum bit for bit.
template <int P>
__device__ __forceinline__ void wave_totals(float (&v)[P]) {
#pragma unroll
    for (int p = 0; p < P; ++p) v[p] = dpp_add<0x111>(v[p]);
#pragma unroll
    for (int p = 0; p < P; ++p) v[p] = dpp_add<0x112>(v[p]);
#pragma unroll
    for (int p = 0; p < P; ++p) v[p] = dpp_add<0x114>(v[p]);
#pragma unroll
    for (int p = 0; p < P; ++p) v[p] = dpp_add<0x118>(v[p]);
#pragma unroll
    for (int p = 0; p < P; ++p) v[p] = dpp_add<0x142>(v[p]);
#pragma unroll
    for (int p = 0; p < P; ++p) v[p] = dpp_add<0x143>(v[p]);
}

// Adds g (P values) of every lane with cell >= 0 into acc[cell*P + p] of this
// wave's LDS region; the lanes sharing a cell are summed across the wave in a
// fixed order and lane 63 adds the sums (LDS float adds, issued in program
// order by the one lane).  Must be reached by all lanes of the wave.
template <int P>
__device__ __forceinline__ void wave_accumulate(float* acc, int cell, const float (&g)[P]) {
    const int lane = __lane_id();
    uint64_t pending = __ballot(cell >= 0);
    while (pending) {
        const int leader = __ffsll((unsigned long long)pending) - 1;
        const int c0 = __builtin_amdgcn_readlane(cell, leader);
        const bool mine = (cell == c0);
        pending &= ~__ballot(mine);
        float tot[P];
#pragma unroll
        for (int p = 0; p < P; ++p) tot[p] = mine ? g[p] : 0.f;
        wave_totals<P>(tot);
        if (lane == 63) {
#pragma unroll
            for (int p = 0; p < P; ++p) atomicAdd(&acc[c0 * P + p], tot[p]);
        }
    }
}

}  // namespace cpab
