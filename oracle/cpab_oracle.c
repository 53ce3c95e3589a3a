/*
 * CPU oracle for CPA integration -- TEST INFRASTRUCTURE ONLY.
 *
 * Only tests/, the package smoke check and bench.py's cpu_baseline leg may load
 * this library, and only as the checker: the product path never calls it.
 *
 * Contents
 *   1. Cell lookup, restating libcpab/numpy/findcellidx.py:16-149 in FP64
 *      (the authority for cell ids) and libcpab/core/cpab_ops.cpp:26-190
 *      (FP32 input, the arithmetic the reference's fixed-step op uses).
 *   2. The reference's fixed-step integrator, restated from
 *      libcpab/core/cpab_ops.cpp:224-260 (forward: nstep applications of
 *      Trel = expm(A~/nstep) of the cell found at the start of each step) and
 *      libcpab/core/cpab_ops.cpp:262-372 (backward: midpoint-RK2 of the
 *      sensitivity ODE q' = B_k x~ + A q, once per basis direction k).
 *      Both run in FP32 like the reference ("f32" entry points) or FP64.
 *   3. The exact (closed-form) integrator in FP64: in every cell the flow of
 *      v = M x + b is the convergent series x(s) = x0 + sum_k s^k/k! M^(k-1) v0
 *      (evaluated on chunks with |M| h <= 1/4 and 24 terms, far below FP64
 *      resolution); boundary-hit times are the first roots of
 *      G(s) = n.x(s) + e for each facet, isolated by dense sampling of G and G'
 *      and refined by bisection/Newton to 1e-15.  Cell changes follow the
 *      facet table (neighbour across the facet hit first).
 *   4. Its gradient by a reverse sweep over the stored segments: the adjoint
 *      lambda runs backwards with exp(h M^T), each chunk adds
 *      int_0^h exp((h-s)M^T) lambda x~(s)^T ds to dL/dA_c (exact double
 *      series), and each facet crossing applies the transposed saltation
 *      matrix S = I + (v_next - v_prev) n^T / (n . v_prev) -- the hit-time and
 *      cell-change terms, identically I when the field is continuous.
 *
 * What pins each function (closed forms, invariants, the reference's compiled
 * op, finite differences) is tested in tests/test_oracle.py.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define OR_MAXD 3
#define OR_NF 4
#define OR_K 24          /* series terms */
#define OR_RHO 0.25      /* chunk size: |M|_F * h <= OR_RHO */
#define OR_NS 48         /* samples per chunk for root isolation */

/* ------------------------------------------------------------------------ */
/* 1. cell lookup                                                            */
/* ------------------------------------------------------------------------ */

static double np_mymin(double a, double b) { return (a < b) ? a : nearbyint(b); }

/* libcpab/numpy/findcellidx.py:22-27 */
static int np_find1(double p, int nx) {
    double idx = floor(p * nx);
    idx = fmax(0.0, fmin(idx, nx - 1));
    return (int)idx;
}

/* libcpab/numpy/findcellidx.py:30-80 */
static int np_find2(const double* p, int nx, int ny) {
    double inc_x = 1.0 / nx, inc_y = 1.0 / ny;
    double p0 = fmin(nx * inc_x - 1e-8, fmax(0.0, p[0]));
    double p1 = fmin(ny * inc_y - 1e-8, fmax(0.0, p[1]));
    double xmod = fmod(p0, inc_x), ymod = fmod(p1, inc_y);
    double x = xmod / inc_x, y = ymod / inc_y;
    double idx = np_mymin(nx - 1, (p0 - xmod) / inc_x) + np_mymin(ny - 1, (p1 - ymod) / inc_y) * nx;
    idx *= 4;
    int c1 = (p[0] <= 0) && (p[1] <= 0) && (p[1] / inc_y < p[0] / inc_x);
    int c2 = !c1 && (p[0] <= 0) && (p[1] >= ny * inc_y) && (p[1] / inc_y - ny > -p[0] / inc_x);
    int c3 = !c1 && !c2 && (p[0] <= 0);
    if (c2) idx += 2;
    if (c3) idx += 3;
    int out = c1 || c2 || c3;
    int c4 = !out && (p[0] >= nx * inc_x) && (p[1] <= 0) && (-p[1] / inc_y > p[0] / inc_x - nx);
    int c5 = !out && !c4 && (p[0] >= nx * inc_x) && (p[1] >= ny * inc_y) && (p[1] / inc_y - ny > p[0] / inc_x - nx);
    int c6 = !out && !c4 && !c5 && (p[0] >= nx * inc_x);
    if (c5) idx += 2;
    if (c6) idx += 1;
    out = out || c4 || c5 || c6;
    int c7 = !out && (p[1] >= ny * inc_y);
    if (c7) idx += 2;
    out = out || c7;
    int c8 = !out && (x < y) && (1 - x < y);
    int c9 = !out && !c8 && (x < y);
    int c10 = !out && !c8 && !c9 && (x >= y) && (1 - x < y);
    if (c8) idx += 2;
    if (c9) idx += 3;
    if (c10) idx += 1;
    return (int)idx;
}

/* libcpab/numpy/findcellidx.py:83-149 */
static int np_find3(const double* pin, int nx, int ny, int nz) {
    double p[3] = {pin[0], pin[1], pin[2]};
    double inc[3] = {1.0 / nx, 1.0 / ny, 1.0 / nz};
    int outside = p[0] < 0.0 || p[0] > 1.0 || p[1] < 0.0 || p[1] > 1.0 || p[2] < 0.0 || p[2] > 1.0;
    if (outside) {
        const double half = 0.5;
        double q[3] = {p[0] - half, p[1] - half, p[2] - half};
        double ax = fabs(q[0]), ay = fabs(q[1]), az = fabs(q[2]);
        double push_x = (half * inc[0]) * ((ax < ay) && (ax < az));
        double push_y = (half * inc[1]) * ((ay < ax) && (ax < az));
        double push_z = (half * inc[2]) * ((az < ax) && (ax < ay));
        if (ax > half) q[0] = copysign(half - push_x, q[0]);
        if (ay > half) q[1] = copysign(half - push_y, q[1]);
        if (az > half) q[2] = copysign(half - push_z, q[2]);
        p[0] = q[0] + half; p[1] = q[1] + half; p[2] = q[2] + half;
    }
    int n[3] = {nx, ny, nz};
    double pc[3], md[3], ijk[3], loc[3];
    for (int a = 0; a < 3; ++a) {
        pc[a] = fmin(n[a] * inc[a] - 1e-8, fmax(0.0, p[a]));
        md[a] = pc[a] - floor(pc[a] / inc[a]) * inc[a]; /* np.mod for positive operands */
        if (md[a] < 0) md[a] += inc[a];
        ijk[a] = np_mymin(n[a] - 1, (pc[a] - md[a]) / inc[a]);
        loc[a] = md[a] / inc[a];
    }
    int i = (int)ijk[0], j = (int)ijk[1], k = (int)ijk[2];
    int idx = 5 * (i + j * nx + k * nx * ny);
    double x = loc[0], y = loc[1], z = loc[2];
    int flip = ((k % 2 == 0) && (i % 2 == 0) && (j % 2 == 1)) || ((k % 2 == 0) && (i % 2 == 1) && (j % 2 == 0)) ||
               ((k % 2 == 1) && (i % 2 == 0) && (j % 2 == 0)) || ((k % 2 == 1) && (i % 2 == 1) && (j % 2 == 1));
    if (flip) { double t = x; x = y; y = 1 - t; }
    if (-x - y + z >= 0) idx += 1;
    else if (x + y + z - 2 >= 0) idx += 2;
    else if (-x + y - z >= 0) idx += 3;
    else if (x - y - z >= 0) idx += 4;
    return idx;
}

static int np_find(int ndim, const int* nc, const double* p) {
    if (ndim == 1) return np_find1(p[0], nc[0]);
    if (ndim == 2) return np_find2(p, nc[0], nc[1]);
    return np_find3(p, nc[0], nc[1], nc[2]);
}

/* points SoA [ndim, nP] */
int oracle_findcell(int ndim, const int* nc, const double* pts, int nP, int* idx) {
    for (int i = 0; i < nP; ++i) {
        double p[3];
        for (int a = 0; a < ndim; ++a) p[a] = pts[a * nP + i];
        idx[i] = np_find(ndim, nc, p);
    }
    return 0;
}

/* ---- FP32-input lookup of libcpab/core/cpab_ops.cpp:22-190 ---- */
static int cc_mymin(int a, double b) { return !(b < a) ? a : (int)round(b); }

static int cc_find1(const float* p, int nx) {
    int idx = (int)floor(p[0] * nx);
    if (idx > nx - 1) idx = nx - 1;
    if (idx < 0) idx = 0;
    return idx;
}

static int cc_find2(const float* p, int nx, int ny) {
    double point[2] = {p[0], p[1]};
    const float inc_x = (float)(1.0 / nx), inc_y = (float)(1.0 / ny);
    double p0 = fmin(nx * inc_x - 0.000000001, fmax(0.0, point[0]));
    double p1 = fmin(ny * inc_y - 0.000000001, fmax(0.0, point[1]));
    double xmod = fmod(p0, (double)inc_x), ymod = fmod(p1, (double)inc_y);
    double x = xmod / inc_x, y = ymod / inc_y;
    int cell = cc_mymin(nx - 1, (p0 - xmod) / inc_x) + cc_mymin(ny - 1, (p1 - ymod) / inc_y) * nx;
    cell *= 4;
    if (point[0] <= 0) {
        if (point[1] <= 0 && point[1] / inc_y < point[0] / inc_x) {
        } else if (point[1] >= ny * inc_y && point[1] / inc_y - ny > -point[0] / inc_x) {
            cell += 2;
        } else {
            cell += 3;
        }
        return cell;
    }
    if (point[0] >= nx * inc_x) {
        if (point[1] <= 0 && -point[1] / inc_y > point[0] / inc_x - nx) {
        } else if (point[1] >= ny * inc_y && point[1] / inc_y - ny > point[0] / inc_x - nx) {
            cell += 2;
        } else {
            cell += 1;
        }
        return cell;
    }
    if (point[1] <= 0) return cell;
    if (point[1] >= ny * inc_y) return cell + 2;
    if (x < y) {
        if (1 - x < y) cell += 2; else cell += 3;
    } else if (1 - x < y) {
        cell += 1;
    }
    return cell;
}

static int cc_find3(const float* pin, int nx, int ny, int nz) {
    const float inc_x = (float)(1.0 / nx), inc_y = (float)(1.0 / ny), inc_z = (float)(1.0 / nz);
    float point[3] = {pin[0], pin[1], pin[2]};
    if (point[0] < 0.0 || point[0] > 1.0 || point[1] < 0.0 || point[1] > 1.0) {
        const float half = 0.5f;
        point[0] -= half; point[1] -= half; point[2] -= half;
        const float ax = fabsf(point[0]), ay = fabsf(point[1]), az = fabsf(point[2]);
        const float px = (ax < ay && ax < az) ? half * inc_x : 0.0f;
        const float py = (ay < ax && ax < az) ? half * inc_y : 0.0f;
        const float pz = (az < ax && ax < ay) ? half * inc_z : 0.0f;
        if (ax > half) point[0] = copysignf(half - px, point[0]);
        if (ay > half) point[1] = copysignf(half - py, point[1]);
        if (az > half) point[2] = copysignf(half - pz, point[2]);
        point[0] += half; point[1] += half; point[2] += half;
    }
    float zero = 0.0f;
    float p0 = fminf((float)(nx * inc_x - 1e-8), fmaxf(zero, point[0]));
    float p1 = fminf((float)(ny * inc_y - 1e-8), fmaxf(zero, point[1]));
    float p2 = fminf((float)(nz * inc_x - 1e-8), fmaxf(zero, point[2]));
    double xmod = fmodf(p0, inc_x), ymod = fmodf(p1, inc_y), zmod = fmodf(p2, inc_z);
    int i = cc_mymin(nx - 1, ((p0 - xmod) / inc_x));
    int j = cc_mymin(ny - 1, ((p1 - ymod) / inc_y));
    int k = cc_mymin(nz - 1, ((p2 - zmod) / inc_z));
    int cell = 5 * (i + j * nx + k * nx * ny);
    double x = xmod / inc_x, y = ymod / inc_y, z = zmod / inc_z;
    int tf = 0;
    if (k % 2 == 0) {
        if ((i % 2 == 0 && j % 2 == 1) || (i % 2 == 1 && j % 2 == 0)) tf = 1;
    } else if ((i % 2 == 0 && j % 2 == 0) || (i % 2 == 1 && j % 2 == 1)) {
        tf = 1;
    }
    if (tf) { double t = x; x = y; y = 1 - t; }
    if (-x - y + z >= 0) cell += 1;
    else if (x + y + z - 2 >= 0) cell += 2;
    else if (-x + y - z >= 0) cell += 3;
    else if (x - y - z >= 0) cell += 4;
    return cell;
}

static int cc_find(int ndim, const float* p, const int* nc) {
    if (ndim == 1) return cc_find1(p, nc[0]);
    if (ndim == 2) return cc_find2(p, nc[0], nc[1]);
    return cc_find3(p, nc[0], nc[1], nc[2]);
}

int oracle_findcell_f32(int ndim, const int* nc, const float* pts, int nP, int* idx) {
    for (int i = 0; i < nP; ++i) {
        float p[3];
        for (int a = 0; a < ndim; ++a) p[a] = pts[a * nP + i];
        idx[i] = cc_find(ndim, p, nc);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 2. reference fixed-step integrator (restated)                             */
/* ------------------------------------------------------------------------ */

static int n_cells(int ndim, const int* nc) {
    int s = (ndim == 1) ? 1 : (ndim == 2 ? 4 : 5);
    for (int a = 0; a < ndim; ++a) s *= nc[a];
    return s;
}

/* libcpab/core/cpab_ops.cpp:224-260, FP32 state like the reference.
 * points [ndim, nP] (broadcast=0) or [n_theta, ndim, nP]; trels [n_theta, nC, ndim, ndim+1] */
int oracle_step_forward_f32(int ndim, const int* nc, int nstep, const float* points, const float* trels,
                            int nP, int n_theta, int broadcast, float* out) {
    const int P = ndim * (ndim + 1);
    const int stride = n_cells(ndim, nc) * P;
    for (int t = 0; t < n_theta; ++t) {
        for (int i = 0; i < nP; ++i) {
            float pt[3], np_[3];
            for (int a = 0; a < ndim; ++a) pt[a] = points[broadcast * t * nP * ndim + i + a * nP];
            for (int n = 0; n < nstep; ++n) {
                int idx = cc_find(ndim, pt, nc);
                const float* T = trels + t * stride + P * idx;
                for (int r = 0; r < ndim; ++r) {
                    float acc = T[r * (ndim + 1)] * pt[0];
                    for (int a = 1; a < ndim; ++a) acc = acc + T[r * (ndim + 1) + a] * pt[a];
                    np_[r] = acc + T[r * (ndim + 1) + ndim];
                }
                for (int a = 0; a < ndim; ++a) pt[a] = np_[a];
            }
            for (int a = 0; a < ndim; ++a) out[t * ndim * nP + i + a * nP] = pt[a];
        }
    }
    return 0;
}

/* FP64 variant of the same stepping with the numpy lookup
 * (libcpab/numpy/transformer.py:22-61). */
int oracle_step_forward_f64(int ndim, const int* nc, int nstep, const double* points, const double* trels,
                            int nP, int n_theta, int broadcast, double* out) {
    const int P = ndim * (ndim + 1);
    const int stride = n_cells(ndim, nc) * P;
    for (int t = 0; t < n_theta; ++t) {
        for (int i = 0; i < nP; ++i) {
            double pt[3], np_[3];
            for (int a = 0; a < ndim; ++a) pt[a] = points[broadcast * t * nP * ndim + i + a * nP];
            for (int n = 0; n < nstep; ++n) {
                int idx = np_find(ndim, nc, pt);
                const double* T = trels + t * stride + P * idx;
                for (int r = 0; r < ndim; ++r) {
                    double acc = T[r * (ndim + 1) + ndim];
                    for (int a = 0; a < ndim; ++a) acc += T[r * (ndim + 1) + a] * pt[a];
                    np_[r] = acc;
                }
                for (int a = 0; a < ndim; ++a) pt[a] = np_[a];
            }
            for (int a = 0; a < ndim; ++a) out[t * ndim * nP + i + a * nP] = pt[a];
        }
    }
    return 0;
}

static void affine_f(int ndim, float* x, const float* A, const float* b) {
    for (int r = 0; r < ndim; ++r) {
        float acc = A[r * (ndim + 1)] * b[0];
        for (int a = 1; a < ndim; ++a) acc = acc + A[r * (ndim + 1) + a] * b[a];
        x[r] = acc + A[r * (ndim + 1) + ndim];
    }
}

static void linear_f(int ndim, float* x, const float* A, const float* b) {
    for (int r = 0; r < ndim; ++r) {
        float acc = A[r * (ndim + 1)] * b[0];
        for (int a = 1; a < ndim; ++a) acc = acc + A[r * (ndim + 1) + a] * b[a];
        x[r] = acc;
    }
}

/* libcpab/core/cpab_ops.cpp:262-372: Jacobian grad [d, n_theta, ndim, nP] of the
 * final point w.r.t. theta_k by midpoint RK2 of (p, q). As [n_theta, nC, ndim, ndim+1],
 * Bs [d, nC, ndim, ndim+1]. */
int oracle_step_backward_f32(int ndim, const int* nc, int nstep, const float* points, const float* As,
                             const float* Bs, int nP, int n_theta, int d, int broadcast, float* grad) {
    const int P = ndim * (ndim + 1);
    const int nC = n_cells(ndim, nc);
    const int stride = nC * P;
    memset(grad, 0, sizeof(float) * (size_t)d * n_theta * ndim * nP);
    for (int bi = 0; bi < n_theta; ++bi) {
        for (int pi = 0; pi < nP; ++pi) {
            for (int k = 0; k < d; ++k) {
                int index = ndim * nP * bi + pi;
                int box = ndim * nP * n_theta;
                float p[3], v[3], pm[3], vm[3], q[3], qm[3], bt[3], aq[3], u[3], um[3];
                for (int a = 0; a < ndim; ++a) p[a] = points[broadcast * bi * ndim * nP + pi + a * nP];
                double h = 1.0 / nstep;
                for (int s = 0; s < nstep; ++s) {
                    int cell = cc_find(ndim, p, nc);
                    const float* Al = As + bi * stride + P * cell;
                    affine_f(ndim, v, Al, p);
                    for (int a = 0; a < ndim; ++a) pm[a] = (float)(p[a] + h * v[a] / 2.0);
                    affine_f(ndim, vm, Al, pm);
                    const float* Bl = Bs + (size_t)P * k * nC + P * cell;
                    for (int a = 0; a < ndim; ++a) q[a] = grad[(size_t)k * box + index + a * nP];
                    affine_f(ndim, bt, Bl, p);
                    linear_f(ndim, aq, Al, q);
                    for (int a = 0; a < ndim; ++a) u[a] = bt[a] + aq[a];
                    for (int a = 0; a < ndim; ++a) qm[a] = (float)(q[a] + h * u[a] / 2.0);
                    affine_f(ndim, bt, Bl, pm);
                    linear_f(ndim, aq, Al, qm);
                    for (int a = 0; a < ndim; ++a) um[a] = bt[a] + aq[a];
                    for (int a = 0; a < ndim; ++a) q[a] = (float)(q[a] + um[a] * h);
                    for (int a = 0; a < ndim; ++a) grad[(size_t)k * box + index + a * nP] = q[a];
                    for (int a = 0; a < ndim; ++a) p[a] = (float)(p[a] + vm[a] * h);
                }
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 3. exact integrator (FP64)                                                */
/* ------------------------------------------------------------------------ */

typedef struct {
    int ndim;
    int nc[3];
    int nC;
    const double* A;      /* [nC, ndim, ndim+1], field in x-space */
    const double* planes; /* [nC, OR_NF, 4] in q = x*nc space */
    const int* nbr;       /* [nC, OR_NF] */
    int max_seg;
} geom_t;

typedef struct {
    int cell;
    double h;          /* chunk duration */
    double x0[3];      /* chunk start */
    int cross_facet;   /* facet crossed at the chunk end, -1 if none */
} chunk_t;

static void cell_M(const geom_t* g, int c, double M[3][3], double b[3]) {
    int n = g->ndim;
    const double* A = g->A + (size_t)c * n * (n + 1);
    for (int r = 0; r < n; ++r) {
        for (int a = 0; a < n; ++a) M[r][a] = A[r * (n + 1) + a];
        b[r] = A[r * (n + 1) + n];
    }
}

static double frob(int n, double M[3][3]) {
    double s = 0;
    for (int r = 0; r < n; ++r)
        for (int a = 0; a < n; ++a) s += M[r][a] * M[r][a];
    return sqrt(s);
}

/* facet plane in x-space: G(x) = nx . x + e */
static void facet_x(const geom_t* g, int c, int f, double nx[3], double* e) {
    const double* P = g->planes + ((size_t)c * OR_NF + f) * 4;
    for (int a = 0; a < g->ndim; ++a) nx[a] = P[a] * g->nc[a];
    *e = P[3];
}

/* series vectors W[k] = M^k v0 / k!  (k = 0..OR_K-1): x(s) = x0 + sum_k s^(k+1)/(k+1) W[k] */
static void series_vectors(int n, double M[3][3], const double* v0, double W[OR_K][3]) {
    for (int a = 0; a < n; ++a) W[0][a] = v0[a];
    for (int k = 1; k < OR_K; ++k)
        for (int r = 0; r < n; ++r) {
            double acc = 0;
            for (int a = 0; a < n; ++a) acc += M[r][a] * W[k - 1][a];
            W[k][r] = acc / k;
        }
}

static void flow_eval(int n, const double* x0, double W[OR_K][3], double s, double* x) {
    for (int a = 0; a < n; ++a) {
        double acc = 0;
        for (int k = OR_K - 1; k >= 0; --k) acc = acc * s + W[k][a] / (k + 1);
        x[a] = x0[a] + s * acc;
    }
}

/* polynomial G(s) = c[0] + c[1] s + ... + c[OR_K] s^OR_K */
static double poly(const double* c, double s) {
    double acc = 0;
    for (int k = OR_K; k >= 0; --k) acc = acc * s + c[k];
    return acc;
}
static double dpoly(const double* c, double s) {
    double acc = 0;
    for (int k = OR_K; k >= 1; --k) acc = acc * s + k * c[k];
    return acc;
}

static double solve_bracket(const double* c, double lo, double hi) {
    /* G(lo) <= 0 < G(hi) */
    double s = 0.5 * (lo + hi);
    for (int it = 0; it < 200; ++it) {
        double gs = poly(c, s);
        if (gs > 0) hi = s; else lo = s;
        double d = dpoly(c, s);
        double sn = (d != 0) ? s - gs / d : 0.5 * (lo + hi);
        if (!(sn > lo && sn < hi)) sn = 0.5 * (lo + hi);
        if (fabs(sn - s) < 1e-16 * (1 + fabs(s)) || hi - lo < 1e-16) { s = sn; break; }
        s = sn;
    }
    return s;
}

/* zero of G' in (lo, hi) with G'(lo), G'(hi) of opposite signs (bisection) */
static double crit_point(const double* c, double lo, double hi) {
    const int up = dpoly(c, lo) > 0;
    for (int it = 0; it < 200 && hi - lo > 1e-17; ++it) {
        double m = 0.5 * (lo + hi);
        if ((dpoly(c, m) > 0) == up) lo = m; else hi = m;
    }
    return 0.5 * (lo + hi);
}

/* first exit time in [0, h] through a facet with coefficients c, or -1.
 * A start on the facet (G(0) >= -tol) moving out exits at s = 0 when
 * allow_start_exit; a start on the facet moving in dips below 0 first, so a
 * positive sample is bracketed from the dip minimum; any other start on the
 * facet is ignored until G has gone negative. */
static double first_exit(const double* c, double h, double vscale, int allow_start_exit) {
    const double tol0 = 1e-12, tol1 = 1e-13 * vscale;
    int on = c[0] >= -tol0;
    if (on && allow_start_exit) {
        if (c[1] > tol1) return 0.0;
        if (fabs(c[1]) <= tol1 && c[2] > 0) return 0.0;
    }
    int skip = on && !(c[1] < -tol1);
    int dip = on && !skip;
    double prev_s = 0.0, prev_g = c[0], prev_d = dpoly(c, 0.0);
    for (int i = 1; i <= OR_NS; ++i) {
        double s = h * i / OR_NS;
        double gs = poly(c, s), ds = dpoly(c, s);
        if (skip) {
            if (gs <= 0) skip = 0;
        } else if (dip) {
            if (gs <= 0) {
                dip = 0;
            } else if (prev_d < 0 && ds > 0) {
                double m = crit_point(c, prev_s, s);
                double gm = poly(c, m);
                if (gm < 0) return solve_bracket(c, m, s);
                skip = 1;
                dip = 0;
            }
        } else {
            if (gs > 0 && prev_g <= 0) return solve_bracket(c, prev_s, s);
            if (prev_d > 0 && ds < 0) {
                double m = crit_point(c, prev_s, s);
                if (poly(c, m) > 0) return solve_bracket(c, prev_s, m);
            }
        }
        prev_s = s; prev_g = gs; prev_d = ds;
    }
    return -1.0;
}

/* integrate one point for time T; returns number of chunks written (if chunks != NULL) */
static int integrate_point(const geom_t* g, const double* xin, double T, double* xout,
                           chunk_t* chunks, int max_chunks, int* seg_cells, double* seg_dur, int max_segs,
                           int* n_segs_out, int c_start) {
    int n = g->ndim;
    double x[3] = {0.0, 0.0, 0.0};
    for (int a = 0; a < n; ++a) x[a] = xin[a];
    /* c_start >= 0: the caller's start cell (a region outside the domain) */
    int c = c_start >= 0 ? c_start : np_find(n, g->nc, x);
    int nch = 0, nseg = 0;
    double seg_t = 0;
    int came_from = -1, zero_hops = 0;
    double M[3][3], b[3];
    int guard = 0;
    while (T > 0 && guard++ < 100000) {
        cell_M(g, c, M, b);
        double nm = frob(n, M);
        double h = T;
        if (nm * h > OR_RHO) h = OR_RHO / nm;
        double v0[3], W[OR_K][3];
        for (int r = 0; r < n; ++r) {
            double acc = b[r];
            for (int a = 0; a < n; ++a) acc += M[r][a] * x[a];
            v0[r] = acc;
        }
        double vs = 0;
        for (int a = 0; a < n; ++a) vs += v0[a] * v0[a];
        vs = sqrt(vs);
        series_vectors(n, M, v0, W);
        double best = -1; int bf = -1;
        for (int f = 0; f < OR_NF; ++f) {
            int nb = g->nbr[c * OR_NF + f];
            if (nb < 0) continue;
            double nx[3], e;
            facet_x(g, c, f, nx, &e);
            double cf[OR_K + 1];
            cf[0] = e;
            for (int a = 0; a < n; ++a) cf[0] += nx[a] * x[a];
            for (int k = 0; k < OR_K; ++k) {
                double acc = 0;
                for (int a = 0; a < n; ++a) acc += nx[a] * W[k][a];
                cf[k + 1] = acc / (k + 1);
            }
            double nn = 0;
            for (int a = 0; a < n; ++a) nn += nx[a] * nx[a];
            int allow_start = !(zero_hops > 0 && nb == came_from) && zero_hops < 16;
            double s = first_exit(cf, h, vs * sqrt(nn), allow_start);
            if (s >= 0 && (best < 0 || s < best)) { best = s; bf = f; }
        }
        double step = (bf >= 0) ? best : h;
        if (chunks && nch < max_chunks) {
            chunks[nch].cell = c;
            chunks[nch].h = step;
            for (int a = 0; a < n; ++a) chunks[nch].x0[a] = x[a];
            chunks[nch].cross_facet = bf;
        }
        nch++;
        double xn[3];
        flow_eval(n, x, W, step, xn);
        for (int a = 0; a < n; ++a) x[a] = xn[a];
        seg_t += step;
        T -= step;
        if (bf >= 0) {
            if (seg_t > 0) {
                if (seg_cells && nseg < max_segs) { seg_cells[nseg] = c; seg_dur[nseg] = seg_t; }
                nseg++;
            }
            seg_t = 0;
            zero_hops = (step <= 1e-14) ? zero_hops + 1 : 0;
            came_from = c;
            c = g->nbr[c * OR_NF + bf];
        }
        if (T < 1e-300) T = 0;
    }
    if (seg_t > 0) {
        if (seg_cells && nseg < max_segs) { seg_cells[nseg] = c; seg_dur[nseg] = seg_t; }
        nseg++;
    }
    for (int a = 0; a < n; ++a) xout[a] = x[a];
    if (n_segs_out) *n_segs_out = nseg;
    return nch;
}

static void setup_geom(geom_t* g, int ndim, const int* nc, const double* A, const double* planes, const int* nbr) {
    g->ndim = ndim;
    for (int a = 0; a < 3; ++a) g->nc[a] = a < ndim ? nc[a] : 1;
    g->nC = n_cells(ndim, nc);
    g->A = A;
    g->planes = planes;
    g->nbr = nbr;
}

/* Exact forward. A: [n_theta, nC_all, ndim, ndim+1] (nC_all = 0: the nC simplices; more:
 * virtual cells after them, each with its own copy of the map it uses); points [ndim, nP] or
 * [n_theta, ndim, nP]; c0 (nullable, laid out like the points without the ndim axis): start
 * cells, -1 for the lookup.  out [n_theta, ndim, nP]. Optional cell sequences: seq [n_theta,
 * nP, max_segs] (-1 padded), nseg [n_theta, nP], durations dur [n_theta, nP, max_segs]. */
int oracle_cf_forward(int ndim, const int* nc, const double* A, const double* planes, const int* nbr,
                      const double* points, int nP, int n_theta, int broadcast, double T,
                      double* out, int max_segs, int* seq, double* dur, int* nseg, int nC_all, const int* c0) {
    geom_t g;
    if (nC_all <= 0) nC_all = n_cells(ndim, nc);
    for (int t = 0; t < n_theta; ++t) {
        setup_geom(&g, ndim, nc, A + (size_t)t * nC_all * ndim * (ndim + 1), planes, nbr);
        for (int i = 0; i < nP; ++i) {
            double x[3], y[3];
            for (int a = 0; a < ndim; ++a) x[a] = points[(size_t)broadcast * t * nP * ndim + a * nP + i];
            int ns = 0;
            int* sq = seq ? seq + ((size_t)t * nP + i) * max_segs : NULL;
            double* du = dur ? dur + ((size_t)t * nP + i) * max_segs : NULL;
            if (sq) for (int s = 0; s < max_segs; ++s) { sq[s] = -1; du[s] = 0; }
            const int cs = c0 ? c0[(size_t)broadcast * t * nP + i] : -1;
            integrate_point(&g, x, T, y, NULL, 0, sq, du, sq ? max_segs : 0, &ns, cs);
            if (nseg) nseg[(size_t)t * nP + i] = ns;
            for (int a = 0; a < ndim; ++a) out[(size_t)t * ndim * nP + a * nP + i] = y[a];
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* 4. exact gradient (FP64 reverse sweep)                                     */
/* ------------------------------------------------------------------------ */

#define OR_MAXCH 20000

/* lambda <- exp(h M^T) lambda, and G += int_0^h exp((h-s)M^T) lambda x~(s)^T ds */
static void chunk_adjoint(int n, double M[3][3], const double* x0, const double* v0, double h,
                          double* lam, double* G /* n x (n+1) */) {
    /* m_j = (M^T)^j lam / j!, X_0 = [x0;1], X_k = [M^(k-1) v0 / k! ; 0] */
    double m[OR_K + 1][3], X[OR_K + 1][4];
    for (int a = 0; a < n; ++a) m[0][a] = lam[a];
    for (int j = 1; j <= OR_K; ++j)
        for (int r = 0; r < n; ++r) {
            double acc = 0;
            for (int a = 0; a < n; ++a) acc += M[a][r] * m[j - 1][a];
            m[j][r] = acc / j;
        }
    for (int a = 0; a < n; ++a) X[0][a] = x0[a];
    X[0][n] = 1.0;
    for (int a = 0; a < n; ++a) X[1][a] = v0[a];
    X[1][n] = 0.0;
    for (int k = 2; k <= OR_K; ++k) {
        for (int r = 0; r < n; ++r) {
            double acc = 0;
            for (int a = 0; a < n; ++a) acc += M[r][a] * X[k - 1][a];
            X[k][r] = acc / k;
        }
        X[k][n] = 0.0;
    }
    /* int_0^h (h-s)^j/j! s^k/k! ds = h^(j+k+1)/(j+k+1)!  (m, X already carry 1/j!, 1/k!) */
    double hp[2 * OR_K + 3];
    hp[0] = 1.0;
    for (int r = 1; r < 2 * OR_K + 3; ++r) hp[r] = hp[r - 1] * h / r; /* h^r / r! */
    for (int j = 0; j <= OR_K; ++j)
        for (int k = 0; j + k <= OR_K; ++k) {
            /* (h-s)^j s^k integral = h^(j+k+1) j! k! /(j+k+1)!; with m_j,X_k carrying 1/j!,1/k! the
               weight is j! k! * h^(j+k+1)/(j+k+1)! ; multiply back j! k! */
            double w = hp[j + k + 1];
            for (int q = 1; q <= j; ++q) w *= q;
            for (int q = 1; q <= k; ++q) w *= q;
            for (int r = 0; r < n; ++r)
                for (int a = 0; a <= n; ++a) G[r * (n + 1) + a] += w * m[j][r] * X[k][a];
        }
    /* lambda(start) = exp(h M^T) lambda = sum_j h^j m_j (m_j has 1/j!) */
    double nl[3] = {0, 0, 0};
    for (int j = OR_K; j >= 0; --j)
        for (int a = 0; a < n; ++a) nl[a] = nl[a] * h + m[j][a];
    for (int a = 0; a < n; ++a) lam[a] = nl[a];
}

/* Exact gradient. grad_out [n_theta, ndim, nP]; grad_A [n_theta, nC_all, ndim, ndim+1]
 * (overwritten; nC_all, c0 as in oracle_cf_forward); grad_pts [n_theta, ndim, nP] (nullable).
 * saltation: apply the facet-crossing terms. */
int oracle_cf_backward(int ndim, const int* nc, const double* A, const double* planes, const int* nbr,
                       const double* points, int nP, int n_theta, int broadcast, double T,
                       const double* grad_out, double* grad_A, double* grad_pts, int saltation, int nC_all,
                       const int* c0) {
    geom_t g;
    int nC = nC_all > 0 ? nC_all : n_cells(ndim, nc);
    int P = ndim * (ndim + 1);
    chunk_t* ch = (chunk_t*)malloc(sizeof(chunk_t) * OR_MAXCH);
    if (!ch) return -1;
    memset(grad_A, 0, sizeof(double) * (size_t)n_theta * nC * P);
    for (int t = 0; t < n_theta; ++t) {
        const double* At = A + (size_t)t * nC * P;
        double* GA = grad_A + (size_t)t * nC * P;
        setup_geom(&g, ndim, nc, At, planes, nbr);
        for (int i = 0; i < nP; ++i) {
            double x[3], y[3];
            for (int a = 0; a < ndim; ++a) x[a] = points[(size_t)broadcast * t * nP * ndim + a * nP + i];
            const int cs = c0 ? c0[(size_t)broadcast * t * nP + i] : -1;
            int nch = integrate_point(&g, x, T, y, ch, OR_MAXCH, NULL, NULL, 0, NULL, cs);
            if (nch > OR_MAXCH) { free(ch); return -2; }
            double lam[3];
            for (int a = 0; a < ndim; ++a) lam[a] = grad_out[(size_t)t * ndim * nP + a * nP + i];
            for (int k = nch - 1; k >= 0; --k) {
                int c = ch[k].cell;
                double M[3][3], b[3], v0[3];
                cell_M(&g, c, M, b);
                for (int r = 0; r < ndim; ++r) {
                    double acc = b[r];
                    for (int a = 0; a < ndim; ++a) acc += M[r][a] * ch[k].x0[a];
                    v0[r] = acc;
                }
                if (saltation && ch[k].cross_facet >= 0) {
                    /* crossing at the end of this chunk into nb: lam_before = S^T lam_after,
                       S = I + (v_nb - v_c) n^T / (n . v_c) evaluated at the crossing point */
                    int f = ch[k].cross_facet;
                    int nb = g.nbr[c * OR_NF + f];
                    double W[OR_K][3], xe[3];
                    series_vectors(ndim, M, v0, W);
                    flow_eval(ndim, ch[k].x0, W, ch[k].h, xe);
                    double M2[3][3], b2[3], vc[3], vn[3], nx[3], e;
                    cell_M(&g, nb, M2, b2);
                    for (int r = 0; r < ndim; ++r) {
                        double a1 = b[r], a2 = b2[r];
                        for (int a = 0; a < ndim; ++a) { a1 += M[r][a] * xe[a]; a2 += M2[r][a] * xe[a]; }
                        vc[r] = a1; vn[r] = a2;
                    }
                    facet_x(&g, c, f, nx, &e);
                    double nv = 0, vsz = 0;
                    for (int a = 0; a < ndim; ++a) { nv += nx[a] * vc[a]; vsz += vc[a] * vc[a]; }
                    if (fabs(nv) > 1e-9 * sqrt(vsz)) {
                        double ul = 0; /* (v_nb - v_c) . lam */
                        for (int a = 0; a < ndim; ++a) ul += (vn[a] - vc[a]) * lam[a];
                        for (int a = 0; a < ndim; ++a) lam[a] += nx[a] * ul / nv;
                    }
                }
                chunk_adjoint(ndim, M, ch[k].x0, v0, ch[k].h, lam, GA + (size_t)c * P);
            }
            if (grad_pts)
                for (int a = 0; a < ndim; ++a) grad_pts[(size_t)t * ndim * nP + a * nP + i] = lam[a];
        }
    }
    free(ch);
    return 0;
}

/* Velocity field at points (libcpab/numpy/functions.py:107-122): v = A_c x~ with c from the lookup.
 * A [nC, ndim, ndim+1] for one theta. */
int oracle_vectorfield(int ndim, const int* nc, const double* A, const double* pts, int nP, double* v) {
    int P = ndim * (ndim + 1);
    for (int i = 0; i < nP; ++i) {
        double p[3];
        for (int a = 0; a < ndim; ++a) p[a] = pts[a * nP + i];
        int c = np_find(ndim, nc, p);
        const double* Ac = A + (size_t)c * P;
        for (int r = 0; r < ndim; ++r) {
            double acc = Ac[r * (ndim + 1) + ndim];
            for (int a = 0; a < ndim; ++a) acc += Ac[r * (ndim + 1) + a] * p[a];
            v[r * nP + i] = acc;
        }
    }
    return 0;
}
