// C-callable wrappers around the reference's CPU op, built only into
// oracle/_ref/ (test infrastructure: golden vectors and the CPU baseline).
// The wrapped functions are libcpab/core/cpab_ops.cpp:224-260 (forward) and
// libcpab/core/cpab_ops.cpp:262-372 (backward); the prototypes come from the
// reference's own header, libcpab/core/cpab_ops.h:13-21.
#include "cpab_ops.h"

extern "C" {

int ref_cpab_forward(float* newpoints, const float* points, const float* trels,
                     int nstepsolver, const int* nc, int ndim, int nP,
                     int batch_size, int broadcast) {
    cpab_forward_op(newpoints, points, trels, &nstepsolver, nc, ndim, nP,
                    batch_size, broadcast);
    return 0;
}

// grad must be zero-filled by the caller ([d, n_theta, ndim, nP]), as the
// reference's binding does with torch::zeros (libcpab/pytorch/transformer.cpp:45).
int ref_cpab_backward(float* grad, const float* points, const float* As,
                      const float* Bs, int nstepsolver, const int* nc,
                      int n_theta, int d, int ndim, int nP, int nC,
                      int broadcast) {
    cpab_backward_op(grad, points, As, Bs, &nstepsolver, nc, n_theta, d, ndim,
                     nP, nC, broadcast);
    return 0;
}

int ref_findcellidx(int ndim, const float* p, const int* nc) {
    return findcellidx(ndim, p, nc);
}

}
