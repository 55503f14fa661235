/* Plain-C descriptor types shared by libcntrx.so and its callers. */
#ifndef CNTRX_TYPES_H
#define CNTRX_TYPES_H

#ifdef __cplusplus
extern "C" {
#endif

/* One cntr::herm_matrix<double> (libcntr/cntr/cntr_herm_matrix_decl.hpp:44-301).
 * sig = -1 fermion, +1 boson.  nt = -1 means Matsubara only. */
typedef struct cntrx_herm {
    int nt, ntau, size1, sig;
    double *ret, *les, *tv, *mat;
} cntrx_herm;

/* One cntr::function<double> (libcntr/cntr/cntr_function_decl.hpp:26-107). */
typedef struct cntrx_func {
    int nt, size1;
    double *data;
} cntrx_func;

/* One propagation request for the continuous-batching server (cntrx_server_*): the
 * reference's dyson(...) sequence (libcntr/cntr/cntr_dyson_impl.hpp:1725-1734) on
 * G of shape (nt, ntau, size1, sig) with H(t) and Sigma = V^2 g, where g is the free
 * contour GF of the constant bath Hamiltonian Hbath (cntr::green_from_H,
 * libcntr/cntr/cntr_equilibrium_impl.hpp:873-953) at the same mu and beta.
 * H and Hbath are host arrays of interleaved complex doubles, copied at submission:
 * H(t) for t = -1..nt at (t+1)*size1^2, Hbath one size1 x size1 block. */
typedef struct cntrx_request {
    int size1, nt, ntau, sig, SolveOrder, mat_method;
    double beta, h, mu, V;
    const double *H;
    const double *Hbath;
} cntrx_request;

enum {
    CNTRX_OK = 0,
    CNTRX_EINVAL = -1,
    CNTRX_ENOMEM = -2,
    CNTRX_EHIP = -3,
    CNTRX_EUNSUPPORTED = -4
};

#define CNTRX_MAT_FOURIER 0
#define CNTRX_MAT_CG 1
#define CNTRX_MAT_FIXPOINT 2

#ifdef __cplusplus
}
#endif
#endif
