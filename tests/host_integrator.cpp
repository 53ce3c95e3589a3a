// Host build of the device integrator (libcpab_amd/csrc/cpab_device.hpp) for
// CPU tests: the same per-point code path as k_transform, compiled for x86 so
// its logic can be compared with the FP64 oracle without a GPU.  Test-only;
// the library never loads this.
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "../libcpab_amd/csrc/cpab_device.hpp"

using namespace cpab;

// returns the number of points that used up the chunk budget (NaN outputs)
template <int D>
static int run(const int* nc, int nC, const float* Aq, const float* planes, const int* nbr, const float* pts,
               int nP, int n_theta, int per_theta, float t_end, float tol0, float* out, float* jac) {
    constexpr int P = D * (D + 1);
    int unfinished = 0;
    float* inv = (float*)malloc(sizeof(float) * nC);
    for (int t = 0; t < n_theta; ++t) {
        const float* A = Aq + (size_t)t * nC * P;
        for (int c = 0; c < nC; ++c) {
            float s = 0.f;
            for (int r = 0; r < D; ++r)
                for (int k = 0; k < D; ++k) s += A[c * P + r * (D + 1) + k] * A[c * P + r * (D + 1) + k];
            inv[c] = s > 0.f ? 1.0f / sqrtf(s) : INFINITY;
        }
        CellTable<D> tab;
        tab.A = A;
        tab.inv_n = inv;
        tab.pl = reinterpret_cast<const float4*>(planes);
        tab.nb = nbr;
        const float* pp = pts + (per_theta ? (size_t)t * D * nP : 0);
        for (int p = 0; p < nP; ++p) {
            float x[D], q[D], J[D][D];
            for (int a = 0; a < D; ++a) {
                x[a] = pp[(size_t)a * nP + p];
                q[a] = x[a] * (float)nc[a];
            }
            const int c0 = find_cell<D>(x, nc);
            integrate<D, true>(q, c0, t_end, tab, nc[0], tol0, J, &unfinished);
            for (int a = 0; a < D; ++a) out[((size_t)t * D + a) * nP + p] = q[a] / (float)nc[a];
            if (jac)
                for (int r = 0; r < D; ++r)
                    for (int k = 0; k < D; ++k) jac[((size_t)t * D * D + r * D + k) * nP + p] = J[r][k];
        }
    }
    free(inv);
    return unfinished;
}

extern "C" int host_transform(int ndim, const int* nc, int nC, const float* Aq, const float* planes, const int* nbr,
                              const float* pts, int nP, int n_theta, int per_theta, float t_end, float tol0,
                              float* out, float* jac) {
    if (ndim == 1) return run<1>(nc, nC, Aq, planes, nbr, pts, nP, n_theta, per_theta, t_end, tol0, out, jac);
    if (ndim == 2) return run<2>(nc, nC, Aq, planes, nbr, pts, nP, n_theta, per_theta, t_end, tol0, out, jac);
    return run<3>(nc, nC, Aq, planes, nbr, pts, nP, n_theta, per_theta, t_end, tol0, out, jac);
}

extern "C" int host_findcell(int ndim, const int* nc, const float* pts, int nP, int* idx) {
    for (int p = 0; p < nP; ++p) {
        if (ndim == 1) {
            float x[1] = {pts[p]};
            idx[p] = find_cell<1>(x, nc);
        } else if (ndim == 2) {
            float x[2] = {pts[p], pts[nP + p]};
            idx[p] = find_cell<2>(x, nc);
        } else {
            float x[3] = {pts[p], pts[nP + p], pts[2 * nP + p]};
            idx[p] = find_cell<3>(x, nc);
        }
    }
    return 0;
}

// Integration from a given state (2D, q-space): cell c, remaining time T and
// the record of the last crossing (came_from, zero_hops) as the state machine
// holds them after a crossing.  q is overwritten with the end point.
extern "C" int host_integrate_state2(const int* nc, int nC, const float* Aq, const float* planes, const int* nbr,
                                     float* q, int c, float T, int came_from, int zero_hops, float tol0) {
    constexpr int P = 6;
    float* inv = (float*)malloc(sizeof(float) * nC);
    for (int k = 0; k < nC; ++k) {
        float s = 0.f;
        for (int r = 0; r < 2; ++r)
            for (int j = 0; j < 2; ++j) s += Aq[k * P + r * 3 + j] * Aq[k * P + r * 3 + j];
        inv[k] = s > 0.f ? 1.0f / sqrtf(s) : INFINITY;
    }
    CellTable<2> tab;
    tab.A = Aq;
    tab.inv_n = inv;
    tab.pl = reinterpret_cast<const float4*>(planes);
    tab.nb = nbr;
    PointState<2> st;
    const float q0[2] = {q[0], q[1]};
    init_point<2>(st, q0, c, T);
    st.came_from = came_from;
    st.zero_hops = zero_hops;
    float J[2][2];
    ChunkLog lg;
    while (point_active<2>(st)) step_point<2, false>(st, tab, nc[0], tol0, J, lg);
    q[0] = st.q[0];
    q[1] = st.q[1];
    free(inv);
    return 0;
}

// The backward's replay of the forward's chunk log: integrates every point
// with the log on, then again replaying the logged chunks (the rest by the
// facet search) and counts the points whose chunks (cell, duration, facet,
// end point) or end state differ in any bit.  Returns the mismatch count.
template <int D>
static int replay_mismatches(const int* nc, int nC, const float* Aq, const float* planes, const int* nbr,
                             const float* pts, int nP, float t_end, float tol0, int* n_logged) {
    constexpr int P = D * (D + 1);
    float* inv = (float*)malloc(sizeof(float) * nC);
    for (int k = 0; k < nC; ++k) {
        float s = 0.f;
        for (int r = 0; r < D; ++r)
            for (int j = 0; j < D; ++j) s += Aq[k * P + r * (D + 1) + j] * Aq[k * P + r * (D + 1) + j];
        inv[k] = s > 0.f ? 1.0f / sqrtf(s) : INFINITY;
    }
    CellTable<D> tab;
    tab.A = Aq;
    tab.inv_n = inv;
    tab.pl = reinterpret_cast<const float4*>(planes);
    tab.nb = nbr;
    int bad = 0, logged = 0;
    for (int p = 0; p < nP; ++p) {
        float x[D], q[D];
        for (int a = 0; a < D; ++a) {
            x[a] = pts[(size_t)a * nP + p];
            q[a] = x[a] * (float)nc[a];
        }
        const int c0 = find_cell<D>(x, nc);
        // forward: search, keep every chunk
        PointState<D> s1, s2;
        init_point<D>(s1, q, c0, t_end);
        init_point<D>(s2, q, c0, t_end);
        ChunkLog lg;
        lg.code = 0;
        Chunk<D> c1[16];
        int n1 = 0;
        while (point_active<D>(s1)) {
            Chunk<D> ch;
            advance<D>(s1, tab, tol0, ch);
            log_chunk(lg, ch.h, ch.facet);
            if (n1 < 16) c1[n1] = ch;
            n1++;
        }
        const int nrec = lg.code & 3;
        logged += nrec;
        // the forward's memory sink records the same log
        {
            float hs[NREC];
            for (int k = 0; k < NREC; ++k) hs[k] = -1.f;
            LogSink sk;
            sk.init(hs, 1, 0);
            float qs[D], Js[D][D];
            for (int a = 0; a < D; ++a) qs[a] = q[a];
            integrate<D, true, true, LogSink>(qs, c0, t_end, tab, nc[0], tol0, Js, sk);
            bool ok = sk.code == lg.code;
            for (int k = 0; k < nrec; ++k) ok = ok && memcmp(&hs[k], &lg.h[k], sizeof(float)) == 0;
            bad += !ok;
        }
        int n2 = 0;
        bool same = true;
        while (point_active<D>(s2)) {
            Chunk<D> ch;
            if (s2.iters < nrec) {
                float h;
                int bf;
                logged_chunk(lg, s2.iters, h, bf);
                replay<D>(s2, tab, h, bf, ch);
            } else {
                advance<D>(s2, tab, tol0, ch);
            }
            if (n2 < 16 && n2 < n1) {
                same = same && ch.cell == c1[n2].cell && ch.facet == c1[n2].facet && ch.next == c1[n2].next &&
                       memcmp(&ch.h, &c1[n2].h, sizeof(float)) == 0 && memcmp(ch.q1, c1[n2].q1, sizeof(ch.q1)) == 0;
            }
            n2++;
        }
        same = same && n1 == n2 && memcmp(s1.q, s2.q, sizeof(s1.q)) == 0 && s1.c == s2.c &&
               memcmp(&s1.T, &s2.T, sizeof(float)) == 0;
        bad += !same;
    }
    free(inv);
    *n_logged = logged;
    return bad;
}

extern "C" int host_replay_mismatches(int ndim, const int* nc, int nC, const float* Aq, const float* planes,
                                      const int* nbr, const float* pts, int nP, float t_end, float tol0,
                                      int* n_logged) {
    if (ndim == 2) return replay_mismatches<2>(nc, nC, Aq, planes, nbr, pts, nP, t_end, tol0, n_logged);
    return replay_mismatches<3>(nc, nC, Aq, planes, nbr, pts, nP, t_end, tol0, n_logged);
}
