/*
 * C ABI of the MI355X (gfx950) CPA-integration library, libcpab_hip.so.
 *
 * Every pointer argument except cpab_tess_t itself and the small host arrays
 * dsize / outsize / nc is a device pointer.  The library never allocates:
 * outputs, gradients and the workspace (size from cpab_workspace_bytes,
 * cpab_data_workspace_bytes or cpab_batch_workspace_bytes) are
 * owned by the caller, and nothing is zero-filled inside the library except
 * where stated (grad_data must arrive zero-filled).  A pointer to an array
 * with no elements (nP, N or n_theta of 0) may be NULL; such a call does
 * nothing except write the zeros of a gradient whose shape is not empty.
 * Every launch goes to the given stream; no call synchronises the host.
 * Return value: CPAB_OK (0) or a CPAB_ERR_* code, text from
 * cpab_status_string().
 *
 * Layouts follow the reference's pytorch backend:
 *   theta  [n_theta, d] f32; B [D, d] f32 with D = nC * ndim * (ndim+1)
 *   points [ndim, nP] (shared grid) or [n_theta, ndim, nP] (per_theta_grid=1)
 *   out    [n_theta, ndim, nP]
 *   data   [N, C, s0(, s1(, s2))], interpolated output [N, C, out0(, out1(, out2))]
 *   grid point p of an outsize grid is x-fastest: p = i0 + out0*(i1 + out1*i2)
 */
#ifndef CPAB_HIP_H
#define CPAB_HIP_H

#include <stddef.h>
#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define CPAB_ABI_VERSION 7

#define CPAB_OK 0
#define CPAB_ERR_ARG 1
#define CPAB_ERR_NDIM 2
#define CPAB_ERR_WORKSPACE 3
#define CPAB_ERR_HIP 4
#define CPAB_ERR_LDS 5
#define CPAB_ERR_UNSUPPORTED 6
#define CPAB_ERR_SCRATCH 7

/* Tessellation on the device.  planes [nC + nV, 4, 4] f32 (n0, n1, n2, e)
 * unit outward facet planes in q = x*nc space, nbr [nC + nV, 4] i32 cell
 * across each facet (-1: wall, -2: unused slot).  Both may be NULL for
 * ndim == 1.  nV virtual cells follow the nC simplices (3D without zero
 * boundary: the convex regions outside the domain); virtual cell nC + v uses
 * the affine map of simplex amap[v] (device int [nV]), the simplex the
 * reference's lookup gives its points after pushing them onto the boundary
 * (libcpab/numpy/findcellidx.py:84-109).  Replaces the (params.nc,
 * params.basis) pair that libcpab/pytorch/transformer.py:146-159 ships to the
 * device on every call.
 * unfinished (nullable, device int): the closed-form integration gives every
 * point at most 512 chunks and facet crossings (fields with |M|_F * t_end up
 * to ~256 per cell fit); a point that uses them up with time left gets NaN
 * outputs (point, Jacobian, so also its gradients), and each forward call
 * (cpab_transform_fwd, the cpab_transform_data_fwd family, cpab_batch_fwd)
 * adds the number of such points here.  The caller zeroes it. */
typedef struct cpab_tess {
    int ndim;
    int nc[3];
    int nC;
    int d;
    const float* B;
    const float* planes;
    const int* nbr;
    int* unfinished;
    int nV;
    const int* amap;
} cpab_tess_t;

const char* cpab_status_string(int status);
int cpab_abi_version(void);

/* Workspace for one transform call (cpab_transform_*, cpab_vectorfield) on
 * n_theta x nP points. */
size_t cpab_workspace_bytes(const cpab_tess_t* t, int n_theta, int nP);

/* Workspace for one transform_data call (fused forward / backward, also the
 * _plain variants with n_theta = 1) onto a meshgrid of outsize[ndim]. */
size_t cpab_data_workspace_bytes(const cpab_tess_t* t, int n_theta, const int* outsize);

/* Floats in the forward's saved state for the fused backward:
 * n_theta * (ndim + ndim^2 + (ndim > 1 ? 3 : 0)) * nP -- warped points,
 * Jacobians and, in 2D/3D, the durations and crossed facets of each point's
 * first two chunks, which the backward replays without searching again. */
size_t cpab_saved_floats(int ndim, int n_theta, int nP);

/* Floats in cpab_transform_fwd's saved state (jac) for cpab_transform_bwd:
 * n_theta * (ndim^2 + (ndim > 1 ? 3 : 0)) * nP -- the q-space Jacobians and,
 * in 2D/3D, each point's chunk log. */
size_t cpab_jac_floats(int ndim, int n_theta, int nP);

/* A_q = B theta per cell, in q-space.  Replaces the basis expansion of
 * libcpab/pytorch/transformer.py:146-150. */
int cpab_theta_to_A(const cpab_tess_t* t, const float* theta, int n_theta, float* Aq, hipStream_t stream);

/* Closed-form integration of the points for time t_end (1 in the Cpab API).
 * grid_shape (nullable, host int[2]): the 2D points form a meshgrid of
 * grid_shape[0] x grid_shape[1] points, x fastest (as from
 * libcpab/pytorch/functions.py:102-108); they are then walked in compact 2D
 * tiles, which changes no point's result.  Pass the same value to
 * cpab_transform_bwd.
 * jac (nullable, cpab_jac_floats) receives per theta the q-space Jacobian
 * [ndim*ndim, nP] and (2D/3D) the chunk log [3, nP], which let
 * cpab_transform_bwd skip its first pass and replay the chunks of its second.  Replaces cpab_gpu.forward,
 * libcpab/pytorch/transformer_cuda.cpp:17-41 (+ the Trels expm of
 * libcpab/pytorch/transformer.py:152-155). */
int cpab_transform_fwd(const cpab_tess_t* t, const float* theta, int n_theta, const float* points, int nP,
                       int per_theta_grid, const int* grid_shape, float t_end, float* out, float* jac,
                       void* workspace, size_t ws_bytes, hipStream_t stream);

/* grad_theta [n_theta, d] = d<grad_out, transform(points)>/d theta, and (if
 * grad_points != NULL) its gradient w.r.t. the points [n_theta, ndim, nP].
 * Replaces cpab_gpu.backward (libcpab/pytorch/transformer_cuda.cpp:43-70) plus
 * the host-side reduction of libcpab/pytorch/transformer.py:201; the
 * reference returns no point gradient (transformer.py:202). */
int cpab_transform_bwd(const cpab_tess_t* t, const float* theta, int n_theta, const float* points, int nP,
                       int per_theta_grid, const int* grid_shape, float t_end, const float* grad_out, const float* jac,
                       float* grad_theta, float* grad_points, void* workspace, size_t ws_bytes, hipStream_t stream);

/* Validation: the cell sequence of each point's trajectory under the same
 * integration as cpab_transform_fwd, as segments of positive duration
 * (consecutive chunks in one cell merge, zero-length visits at vertices are
 * dropped).  cells [n_theta, nP, max_segs] receives the first max_segs cells
 * (-1 padded), nseg [n_theta, nP] the number of segments.  The reference has
 * no counterpart (its numpy path's cell indices are what it is compared
 * against, libcpab/numpy/transformer.py:22-61). */
int cpab_transform_cells(const cpab_tess_t* t, const float* theta, int n_theta, const float* points, int nP,
                         int per_theta_grid, float t_end, int max_segs, int* cells, int* nseg, void* workspace,
                         size_t ws_bytes, hipStream_t stream);

/* Linear interpolation of data at grid [N, ndim, nP], nP = prod(outsize).
 * Replaces libcpab/pytorch/interpolation.py:11-171 (interpolate1D/2D/3D). */
int cpab_interpolate_fwd(int ndim, const float* data, int N, int C, const int* dsize, const float* grid,
                         const int* outsize, float* out, hipStream_t stream);

/* Backward of cpab_interpolate_fwd: grad_data (nullable, zero-filled by the
 * caller, accumulated atomically) and grad_grid (nullable).  The reference
 * gets these from torch autograd through interpolation.py. */
int cpab_interpolate_bwd(int ndim, const float* data, int N, int C, const int* dsize, const float* grid,
                         const int* outsize, const float* grad_out, float* grad_data, float* grad_grid,
                         hipStream_t stream);

/* Fused uniform meshgrid -> closed-form integration -> interpolation
 * (Cpab.transform_data, libcpab/cpab.py:304-331).  lin holds the meshgrid axis
 * coordinates concatenated (outsize[0] + outsize[1] + ... floats; the values
 * of libcpab/pytorch/functions.py:102-108).  saved (nullable,
 * cpab_saved_floats) keeps the warped points, Jacobians and chunk logs for the backward. */
int cpab_transform_data_fwd(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                            const int* dsize, const int* outsize, const float* lin, float t_end, float* out,
                            float* saved, void* workspace, size_t ws_bytes, hipStream_t stream);

/* cpab_transform_data_fwd without a workspace: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  A_q
 * lives in a static device array of CPAB_NOALLOC_SCRATCH_BYTES per GPU, held
 * until the process ends; n_theta * nC * ndim * (ndim+1) * 4 bytes beyond it
 * return CPAB_ERR_SCRATCH.  The call allocates nothing.  All calls on one GPU
 * share that array: calls that overlap on one GPU (concurrent streams or
 * host threads) race.  The stream must belong to the current device.  Same
 * results as cpab_transform_data_fwd. */
#define CPAB_NOALLOC_SCRATCH_BYTES (4u << 20)
int cpab_transform_data_fwd_noalloc(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                                    const int* dsize, const int* outsize, const float* lin, float t_end, float* out,
                                    float* saved, hipStream_t stream);

/* Backward of cpab_transform_data_fwd: grad_data (nullable, zero-filled,
 * [n_theta, C, dsize...]) and grad_theta [n_theta, d].  saved may be NULL
 * (the trajectory is then recomputed). */
int cpab_transform_data_bwd(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                            const int* dsize, const int* outsize, const float* lin, float t_end, const float* grad_out,
                            const float* saved, float* grad_data, float* grad_theta, void* workspace, size_t ws_bytes,
                            hipStream_t stream);

/* Cell index of each point [ndim, nP] -> idx [nP] (int32).  Replaces
 * findcellidx of libcpab/pytorch/findcellidx.py:16-148 (numpy semantics of
 * libcpab/numpy/findcellidx.py). */
int cpab_findcell(int ndim, const int* nc, const float* points, int nP, int* idx, hipStream_t stream);

/* Velocity field of one theta ([1, d]) at points [ndim, nP] -> v [ndim, nP].
 * Replaces calc_vectorfield, libcpab/pytorch/functions.py:111-129. */
int cpab_vectorfield(const cpab_tess_t* t, const float* theta, const float* points, int nP, float* v, void* workspace,
                     size_t ws_bytes, hipStream_t stream);

/* The plain method: the same work as cpab_transform_data_fwd/_bwd issued as
 * one launch sequence per instance (theta).  It is the baseline the batched
 * single-launch path is measured against; cpab_data_workspace_bytes(t, 1,
 * outsize) bytes of workspace suffice.  No reference counterpart (the
 * reference launches once per call, libcpab/pytorch/transformer_cuda.cu:18-64). */
int cpab_transform_data_fwd_plain(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                                  const int* dsize, const int* outsize, const float* lin, float t_end, float* out,
                                  float* saved, void* workspace, size_t ws_bytes, hipStream_t stream);
int cpab_transform_data_bwd_plain(const cpab_tess_t* t, const float* theta, int n_theta, const float* data, int C,
                                  const int* dsize, const int* outsize, const float* lin, float t_end,
                                  const float* grad_out, const float* saved, float* grad_data, float* grad_theta,
                                  void* workspace, size_t ws_bytes, hipStream_t stream);

/* ---- ragged batches --------------------------------------------------------
 * One transform_data problem per instance; instances may differ in ndim,
 * tessellation, theta size, channels, data and output sizes.  All pointers
 * are device pointers except that the descriptor array itself is passed both
 * as a host copy (validation, launch sizing) and a device copy (read by the
 * kernels).  cpab_batch_plan fills item0/tiles and the item -> instance map. */
typedef struct cpab_instance {
    cpab_tess_t tess;
    int C;
    int dsize[3];
    int outsize[3];
    int nP;
    int item0;           /* first work item (tile of BLOCK*ppt points); set by cpab_batch_plan */
    int tiles;           /* work items of this instance; set by cpab_batch_plan */
    const float* theta;  /* [d] */
    const float* data;   /* [C, dsize...] */
    const float* lin;    /* outsize[0] + outsize[1] + outsize[2] meshgrid axis coordinates */
    float* out;          /* [C, outsize...] */
    float* saved;        /* cpab_saved_floats(ndim, 1, nP) floats or NULL */
    const float* grad_out;
    float* grad_data;    /* NULL or zero-filled [C, dsize...] */
    float* grad_theta;   /* [d] */
    void* ws;            /* cpab_batch_workspace_bytes(&tess, outsize, ppt) bytes */
    size_t ws_bytes;
} cpab_instance_t;

/* Per-instance workspace that serves both the batched and the plain calls. */
size_t cpab_batch_workspace_bytes(const cpab_tess_t* t, const int* outsize, int ppt);

/* Splits every instance into tiles of 256*ppt points (compact 2D tiles for 2D
 * grids whose second axis has at least 8 points): writes inst[i].item0 /
 * .tiles and item_inst[item] (host array of max_items ints).  The items of
 * one instance are consecutive; instances are grouped by the LDS their
 * backward needs (three classes, in class order).  Returns the number of
 * work items, or -CPAB_ERR_* .  With item_inst == NULL it only counts the
 * items. */
int cpab_batch_plan(cpab_instance_t* inst, int n_inst, int ppt, int* item_inst, int max_items);

/* One persistent launch over all work items (blocks = grid size; workers pull
 * items from an atomic queue at counter[0], reset inside the call by a
 * kernel, so the sequence can be captured into a graph; counter holds at
 * least 33 ints), preceded by the
 * per-instance A = B theta kernel.  No reference counterpart (the reference
 * handles one tessellation and one size per call,
 * libcpab/pytorch/transformer_cuda.cu:18-119). */
int cpab_batch_fwd(const cpab_instance_t* host_inst, const cpab_instance_t* dev_inst, int n_inst,
                   const int* dev_item_inst, int n_items, int ppt, int blocks, float t_end, int* counter,
                   hipStream_t stream);

/* Backward of cpab_batch_fwd: dL/ddata (if grad_data), dL/dtheta of every
 * instance (per-item partials reduced per instance in a fixed order).  One
 * persistent launch per LDS class of the plan, each with the LDS of its
 * largest tessellation and its own queue (counter[0], [16], [32]).  When
 * every instance has saved state, it is the state of the preceding
 * cpab_batch_fwd on the same instances and thetas, and the backward takes
 * A = B theta from the workspaces where that forward left it. */
int cpab_batch_bwd(const cpab_instance_t* host_inst, const cpab_instance_t* dev_inst, int n_inst,
                   const int* dev_item_inst, int n_items, int ppt, int blocks, float t_end, int* counter,
                   hipStream_t stream);

/* The same work as one launch sequence per instance (the plain method). */
int cpab_batch_fwd_plain(const cpab_instance_t* host_inst, int n_inst, float t_end, hipStream_t stream);
int cpab_batch_bwd_plain(const cpab_instance_t* host_inst, int n_inst, float t_end, hipStream_t stream);

/* p[0:n] = 0 with a kernel (safe inside graph capture). */
int cpab_fill_zero(float* p, size_t n, hipStream_t stream);

/* ---- method='numeric': the reference's fixed-step scheme ------------------ */

/* T = expm(scale * A~)[:ndim] for n affine matrices A [n, ndim, ndim+1]
 * (Trels = expm(A~/nstepsolver), libcpab/pytorch/transformer.py:152-155,
 * which computes it with the Pade-13 of libcpab/pytorch/expm.py). */
int cpab_expm_affine(const float* A, int n, int ndim, float scale, float* T, hipStream_t stream);

/* Same arguments and results as the reference's cpab_forward_op
 * (libcpab/core/cpab_ops.h:13-16, libcpab/core/cpab_ops.cpp:224-260) on device
 * pointers; nstepsolver by value, nc a host array.  Bit-exact with the
 * reference's op on the same Trels. */
int cpab_ref_forward(float* newpoints, const float* points, const float* trels, int nstepsolver, const int* nc,
                     int ndim, int nP, int batch_size, int broadcast, hipStream_t stream);

/* Same as the reference's cpab_backward_op (libcpab/core/cpab_ops.h:17-21,
 * libcpab/core/cpab_ops.cpp:262-372): grad [d, n_theta, ndim, nP] is written
 * completely (no zero-fill needed). */
int cpab_ref_backward(float* grad, const float* points, const float* As, const float* Bs, int nstepsolver,
                      const int* nc, int n_theta, int d, int ndim, int nP, int nC, int broadcast, hipStream_t stream);

/* ---- use_slow: the reference's pure-torch stepping path ------------------ */

/* out [n_theta, ndim, nP] = nstepsolver steps x <- Trel[cell(x)] [x; 1] of
 * each point under its theta's trels [n_theta, nC, ndim, ndim+1] (the first
 * ndim rows of expm(A~ / nstepsolver)), the cell from the torch lookup
 * (libcpab/pytorch/findcellidx.py; numpy rules, 3D clamp 1 - 1e-4).  Replaces
 * CPAB_transformer_slow, libcpab/pytorch/transformer.py:90-128. */
int cpab_slow_forward(const float* points, const float* trels, int nstepsolver, const int* nc, int ndim, int nP,
                      int n_theta, int per_theta_grid, float* out, hipStream_t stream);

/* Exact gradient of <grad_out, cpab_slow_forward(points, trels)> w.r.t. the
 * points (grad_points [n_theta, ndim, nP], nullable) and the trels
 * (grad_trels, same shape as trels, zero-filled by the caller, accumulated
 * with atomics).  trels_inv [n_theta, nC, ndim, ndim] holds the inverse of
 * each Trel's linear part.  The reference gets this gradient from torch
 * autograd through the loop of libcpab/pytorch/transformer.py:117-121. */
int cpab_slow_backward(const float* points, const float* trels, const float* trels_inv, int nstepsolver,
                       const int* nc, int ndim, int nP, int n_theta, int per_theta_grid, const float* grad_out,
                       float* grad_points, float* grad_trels, hipStream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* CPAB_HIP_H */
