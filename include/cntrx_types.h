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
