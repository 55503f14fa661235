/*
 * CPU oracle for the Kadanoff-Baym time-stepping path (TEST INFRASTRUCTURE).
 *
 * This is a serial C++ restatement of the reference libcntr algorithms, used only
 * as the checker for the HIP implementation in nessi_amd/.  Nothing in the product
 * library links or calls it.  Every function cites the reference file:line it
 * follows; paths are relative to the reference root (libcntr/cntr/...).
 *
 * Storage descriptors use the same packed layout as cntr::herm_matrix
 * (libcntr/cntr/cntr_herm_matrix_impl.hpp:456-478):
 *   ret(i,j<=i)  at ((i*(i+1))/2 + j) * es
 *   les(i<=j,j)  at ((j*(j+1))/2 + i) * es
 *   tv(i,m)      at (i*(ntau+1) + m) * es
 *   mat(m)       at m * es
 * with es = size1*size1 and every element a row-major size1 x size1 block of
 * interleaved complex doubles.  A function H(t), t=-1..nt, stores H(t) at (t+1)*es
 * (libcntr/cntr/cntr_function_decl.hpp:45-48).
 */
#ifndef CNTR_ORACLE_H
#define CNTR_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_herm {
    int nt, ntau, size1, sig;
    double *ret, *les, *tv, *mat;
} orc_herm;

typedef struct orc_func {
    int nt, size1;
    double *data;
} orc_func;

/* Integrator weights (libcntr/cntr/integration.hpp:128-514). out arrays are sized by
 * the caller: greg 4*(k+1)^2, omega k+1, bd k+2, P (k+1)^2, D (k+1)^2, W (k+1)^3,
 * rcorr max(k-1,0)*(k+1)^2. */
int orc_integrator_tables(int k, double *greg, double *omega, double *bd, double *P,
                          double *D, double *W, double *rcorr);
double orc_gregory_weight(int k, int n, int j);

/* contour operations */
int orc_convolution_matsubara(orc_herm *C, orc_herm *A, orc_herm *B, double beta, int k);
int orc_convolution_timestep(int n, orc_herm *C, orc_herm *A, orc_herm *Acc, orc_herm *B,
                             orc_herm *Bcc, double beta, double h, int k);
int orc_incr_convolution(int tstp, const unsigned char *mask, double alpha_re, double alpha_im,
                         orc_herm *C, orc_herm *A, orc_herm *Acc, const double *f0,
                         const double *ft, orc_herm *B, orc_herm *Bcc, double beta, double h,
                         int k);
int orc_dyson_mat(orc_herm *G, double mu, orc_func *H, orc_herm *Sigma, double beta, int k,
                  int method, int force_hermitian);
int orc_dyson_start(orc_herm *G, double mu, orc_func *H, orc_herm *Sigma, double beta,
                    double h, int k);
int orc_dyson_timestep(int n, orc_herm *G, double mu, orc_func *H, orc_herm *Sigma,
                       double beta, double h, int k);
int orc_vie2_mat(orc_herm *G, orc_herm *F, orc_herm *Fcc, orc_herm *Q, double beta, int k,
                 int method);
int orc_vie2_start(orc_herm *G, orc_herm *F, orc_herm *Fcc, orc_herm *Q, double beta,
                   double h, int k);
int orc_vie2_timestep(int n, orc_herm *G, orc_herm *F, orc_herm *Fcc, orc_herm *Q,
                      double beta, double h, int k, int method);
int orc_bubble1(int tstp, orc_herm *C, int c1, int c2, orc_herm *A, orc_herm *Acc, int a1,
                int a2, orc_herm *B, orc_herm *Bcc, int b1, int b2);
int orc_bubble2(int tstp, orc_herm *C, int c1, int c2, orc_herm *A, orc_herm *Acc, int a1,
                int a2, orc_herm *B, orc_herm *Bcc, int b1, int b2);
int orc_extrapolate_timestep(int n, orc_herm *G, int k);
double orc_distance_norm2(int tstp, orc_herm *g1, orc_herm *g2);
int orc_green_from_H_const(orc_herm *G, double mu, const double *H0, double beta, double h);
/* time-dependent H (cntr_equilibrium_impl.hpp:754-870, 955-1040): tstp < -1 fills every
 * component, tstp >= -1 only that timestep (the reference's timestep form); fixHam = 0
 * overwrites H(t > SolveOrder) with its extrapolation as the reference does */
int orc_green_from_H_td(orc_herm *G, double mu, orc_func *H, double beta, double h, int SolveOrder,
                        int cf_order, int fixHam, int tstp);
int orc_dyson_solve(orc_herm *G, orc_herm *Sigma, double mu, orc_func *H, const double *Hbath,
                    double V, double beta, double h, int k, int method, int nt_stop);
int orc_force_matsubara_hermitian(orc_herm *G);

/* bounded CPU baseline for the k-batched lattice workload (OpenMP over k): the setup
   (equilibrium, bootstrap) on setup_threads, the timed stepping on nthreads; G_loc and
   the G_k listed in kidx[0..nk_out-1] copied out when requested */
int orc_lattice_run(int nk, const double *epsk, double U, double beta, int ntau, int nt,
                    double h, int k, int nt_stop, double tol, int maxiter, int nthreads,
                    int setup_threads, double *out_density, int *out_iters, double *out_seconds,
                    orc_herm *gloc_out, int nk_out, const int *kidx, orc_herm *gk_out);

#ifdef __cplusplus
}
#endif
#endif
