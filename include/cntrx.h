/*
 * cntrx -- MI355X-native (gfx950, HIP, complex FP64) Kadanoff-Baym time stepping.
 *
 * C ABI of libcntrx.so.  Every entry point replaces one public function of the
 * reference libcntr (namespace cntr, explicitly instantiated for double); the
 * reference declaration it stands in for is cited next to each prototype.  Paths
 * are relative to the reference root.
 *
 * Storage: a cntrx_herm describes one cntr::herm_matrix whose four components live
 * in device memory with exactly the reference's packed layout
 * (libcntr/cntr/cntr_herm_matrix_impl.hpp:456-478):
 *     ret(i,j<=i)  at ((i*(i+1))/2 + j) * es      (row i contiguous)
 *     les(i<=j,j)  at ((j*(j+1))/2 + i) * es      (column j contiguous)
 *     tv(i,m)      at (i*(ntau+1) + m) * es
 *     mat(m)       at m * es
 * es = size1*size1; every element is a row-major size1 x size1 block of interleaved
 * complex doubles.  A cntrx_func describes a cntr::function: H(t), t=-1..nt, at
 * (t+1)*es (libcntr/cntr/cntr_function_decl.hpp:45-48).
 *
 * Errors: every call returns CNTRX_OK (0) or a negative code and never aborts;
 * cntrx_last_error() gives a message.  The reference asserts/aborts instead
 * (libcntr/cntr/cntr_dyson_impl.hpp:1625-1629).
 *
 * Streams: the trailing `stream` argument is a hipStream_t (NULL = default stream).
 * Calls are asynchronous on that stream unless stated otherwise.  Synchronise a stream
 * (or the device) before destroying it once it has carried cntrx calls.  Per-call scratch
 * lives in one grow-only arena per device: calls that overlap on different streams of
 * one device race on it (issue them on one stream); a cntrx_lattice owns its scratch.
 *
 * Block sizes: the solvers take size1 1..8.  Their kernels are built for 1, 2, 4 and 8;
 * a call at size1 3 or 5..7 runs on copies of its arguments embedded in the next of those
 * sizes (each s x s block A becomes [[A, 0], [0, d 1]], d = mu + 1 for the Hamiltonian
 * and 0 otherwise; the padding never feeds the s x s part, which is the size-s
 * solution).  Such a call copies timesteps 0..n of every argument in and the timesteps it
 * writes back, so descriptors with only a window of timesteps resident cannot be used at
 * these sizes, and cntrx_dyson_timestep_noalloc and the server take 1, 2, 4 and 8 only.
 * The bubbles, extrapolate_timestep, distance_norm2 and timestep_lincomb take any
 * size1 >= 1.  Larger sizes return CNTRX_EUNSUPPORTED from the solvers.
 * SolveOrder k in 1..5 (0 only where the reference uses it).
 */
#ifndef CNTRX_H
#define CNTRX_H

#include <stddef.h>

#include "cntrx_types.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- library state ---------------------------------------------------------- */
/* Selects the device and uploads the Integrator tables for k = 0..5.  Replaces the
 * lazily built, unsynchronised integration::I<T>(k) cache
 * (libcntr/cntr/integration.hpp:506-514).  Called implicitly by every entry point. */
int cntrx_init(int device);
const char *cntrx_last_error(void);
/* Number of bounded inter-workgroup waits (the size1 >= 2 recurrences) that gave up on
 * the current device since the library was loaded.  Always 0 unless a launch went
 * wrong; results of such a launch are invalid.  Synchronises the device. */
int cntrx_wait_timeouts(long *count);
/* Host copy of the generated weights (integration.hpp:128-514 accessors). Arrays as
 * in the oracle: greg 4(k+1)^2, omega k+1, bd k+2, P,D (k+1)^2, W (k+1)^3,
 * rcorr (k-1)(k+1)^2.  Any pointer may be NULL. */
int cntrx_integrator_tables(int k, double *greg, double *omega, double *bd, double *P,
                            double *D, double *W, double *rcorr);

/* ---- Dyson equation  [i d/dt + mu - H(t)] G - Sigma*G = delta_C ----------------- */
/* cntr::dyson_mat (libcntr/cntr/cntr_dyson_decl.hpp:78-80); method 0 = Fourier,
 * 1 = conjugate gradients (CNTR_MAT_CG: at most 40 iterations to a residual of 1e-12,
 * cntr_dyson_impl.hpp:1403-1406; it reads the residual on the host once per iteration,
 * so it synchronises the stream), 2 = fixpoint. */
int cntrx_dyson_mat(const cntrx_herm *G, double mu, const cntrx_func *H,
                    const cntrx_herm *Sigma, double beta, int SolveOrder, int method,
                    int force_hermitian, void *stream);
/* cntr::dyson_start (libcntr/cntr/cntr_dyson_decl.hpp:88) */
int cntrx_dyson_start(const cntrx_herm *G, double mu, const cntrx_func *H,
                      const cntrx_herm *Sigma, double beta, double h, int SolveOrder,
                      void *stream);
/* cntr::dyson_timestep (libcntr/cntr/cntr_dyson_decl.hpp:92-93); requires n > SolveOrder */
int cntrx_dyson_timestep(int n, const cntrx_herm *G, double mu, const cntrx_func *H,
                         const cntrx_herm *Sigma, double beta, double h, int SolveOrder,
                         void *stream);
/* cntrx_dyson_timestep for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  No counterpart
 * in the reference.  Same arguments, kernels and results (bit for bit) as
 * cntrx_dyson_timestep, but it calls no hipMalloc, hipFree or other allocating routine:
 * its scratch (the job descriptor, tile lists and per-solve workspace) is a fixed 16 MiB
 * device array per device, set up by cntrx_init(device), which must have been called
 * on the current device first (else CNTRX_EINVAL).  A timestep whose scratch does not fit
 * returns CNTRX_EINVAL with the size it needs: (3*(max(nt,k)+1)+8)*size1^2*16 bytes of
 * workspace, plus (ntau+2+n)*256 bytes at size1 = 1, plus its tile lists (for example
 * nt <= 5400 at size1 = 8).  Every call on a device shares that array: calls that overlap
 * on one device (on different streams, or threads) race; issue them on one stream. */
int cntrx_dyson_timestep_noalloc(int n, const cntrx_herm *G, double mu, const cntrx_func *H,
                                 const cntrx_herm *Sigma, double beta, double h, int SolveOrder,
                                 void *stream);
/* Batched form: count independent dyson_timestep calls (own n, size1, G, H, mu, beta, h;
 * Sigma may be shared) in one call: one staging upload, then one launch sequence per size1
 * class.  Solves with their own Sigma give the same result bit for bit as alone; solves of
 * size1 1 or 8 that share Sigma run their lesser convolution (and for size1 8 the Matsubara
 * integral of the mixed component) as products over the batch.  No reference counterpart: replaces
 * the per-k loop over dyson_timestep inside the MPI k-parallel driver
 * (examples/programs/gw.cpp:401-411, libcntr/cntr/cntr_distributed_timestep_array_impl.hpp:87-126). */
int cntrx_dyson_timestep_batched(int count, const int *n, const cntrx_herm *G,
                                 const double *mu, const cntrx_func *H,
                                 const cntrx_herm *Sigma, const double *beta,
                                 const double *h, int SolveOrder, void *stream);

/* Autotuner of the batched calls (no reference counterpart).  With it on, the first
 * cntrx_dyson_timestep_batched call (or server tick) that meets an untuned (size1,
 * history-length bucket: n < 64, then one per power of two) times the candidate launch
 * shapes on its own live solves, keeps the fastest for the rest of the process and records
 * the decision: the panel size of the size1 2 and 4 recurrences (two candidates), and the
 * workgroup target of the size1 8 shared-Sigma GEMMs (2048, 1024, 4096).  Candidates sum
 * in different orders, so results may differ between them at rounding level.  Off by
 * default (the first candidate everywhere); decisions made while it was on stay in use
 * after it is turned off.  A batched call looks the panel size up per solve, by the
 * solve's own n, so with the table fixed a solve's result does not depend on its
 * batch-mates. */
int cntrx_autotune(int enable);
/* decisions so far, 8 doubles each: size1, bucket, what (0 panel, 1 GEMM split), number of
 * candidates, chosen index, ms of candidates 0..2; *count = number of decisions */
int cntrx_autotune_decisions(int cap, double *rows8, int *count);

/* ---- integral form  [1+F]G = Q ----------------------------------------------- */
/* cntr::vie2_mat (libcntr/cntr/cntr_vie2_decl.hpp), method 0 Fourier (F alone),
 * 1 conjugate gradients on [1 + (F* + *Fcc)/2] G = Q (synchronises the stream),
 * 2 fixpoint */
int cntrx_vie2_mat(const cntrx_herm *G, const cntrx_herm *F, const cntrx_herm *Fcc,
                   const cntrx_herm *Q, double beta, int SolveOrder, int method, void *stream);
/* cntr::vie2_start (libcntr/cntr/cntr_vie2_decl.hpp:67) */
int cntrx_vie2_start(const cntrx_herm *G, const cntrx_herm *F, const cntrx_herm *Fcc,
                     const cntrx_herm *Q, double beta, double h, int SolveOrder, void *stream);
/* cntr::vie2_timestep (libcntr/cntr/cntr_vie2_decl.hpp:72-73): n=-1 -> vie2_mat,
 * n <= SolveOrder -> vie2_start with order n (libcntr/cntr/cntr_vie2_impl.hpp:1604-1611) */
int cntrx_vie2_timestep(int n, const cntrx_herm *G, const cntrx_herm *F, const cntrx_herm *Fcc,
                        const cntrx_herm *Q, double beta, double h, int SolveOrder, int method,
                        void *stream);

/* ---- convolution C = A*B ----------------------------------------------------- */
/* cntr::convolution_timestep (libcntr/cntr/cntr_convolution_decl.hpp:61-66); n=-1 is
 * the Matsubara convolution (cntr::convolution_matsubara, :545-571) */
int cntrx_convolution_timestep(int n, const cntrx_herm *C, const cntrx_herm *A,
                               const cntrx_herm *Acc, const cntrx_herm *B,
                               const cntrx_herm *Bcc, double beta, double h, int SolveOrder,
                               void *stream);
/* cntr::incr_convolution and its mat/ret/tv/les parts
 * (libcntr/cntr/cntr_convolution_decl.hpp:297-318, cntr_convolution_impl.hpp:4163-4805):
 * C += alpha * A*f*B at timestep tstp.  mask_host: host bytes, one per output (nonzero =
 * computed): ntau+1 Matsubara points for tstp = -1, else tstp+1 ret, ntau+1 tv and
 * tstp+1 les outputs in that order; NULL selects every output, as the public
 * incr_convolution does.  f0 = f(-i beta) (size1^2 complex) and ft = f(t) for
 * t = 0..max(tstp, SolveOrder) are device pointers, or NULL for f = 1.  The tstp-less
 * overloads (convolution_timestep_new, :4810-4983) are set_timestep_zero followed by
 * this call with alpha = 1. */
int cntrx_incr_convolution(int tstp, const unsigned char *mask_host, double alpha_re,
                           double alpha_im, const cntrx_herm *C, const cntrx_herm *A,
                           const cntrx_herm *Acc, const double *f0, const double *ft,
                           const cntrx_herm *B, const cntrx_herm *Bcc, double beta, double h,
                           int SolveOrder, void *stream);

/* ---- diagrams ----------------------------------------------------------------- */
/* cntr::Bubble1 (libcntr/cntr/cntr_bubble_decl.hpp:62-180): C_{c1c2}(t,t') =
 * i A_{a1a2}(t,t') B_{b2b1}(t',t) at timestep tstp (tstp=-1: Matsubara) */
int cntrx_bubble1(int tstp, const cntrx_herm *C, int c1, int c2, const cntrx_herm *A,
                  const cntrx_herm *Acc, int a1, int a2, const cntrx_herm *B,
                  const cntrx_herm *Bcc, int b1, int b2, void *stream);
/* cntr::Bubble2 (libcntr/cntr/cntr_bubble_decl.hpp:181-301): C_{c1c2}(t,t') =
 * i A_{a1a2}(t,t') B_{b1b2}(t,t') */
int cntrx_bubble2(int tstp, const cntrx_herm *C, int c1, int c2, const cntrx_herm *A,
                  const cntrx_herm *Acc, int a1, int a2, const cntrx_herm *B,
                  const cntrx_herm *Bcc, int b1, int b2, void *stream);

/* ---- utilities ---------------------------------------------------------------- */
/* cntr::extrapolate_timestep (libcntr/cntr/cntr_utilities_decl.hpp:23) */
int cntrx_extrapolate_timestep(int n, const cntrx_herm *G, int SolveOrder, void *stream);
/* cntr::distance_norm2 (libcntr/cntr/cntr_utilities_impl.hpp:1621); synchronises
 * the stream and writes the result to *out (host). */
int cntrx_distance_norm2(int tstp, const cntrx_herm *g1, const cntrx_herm *g2, double *out,
                         void *stream);
/* cntr::green_from_H for constant H (libcntr/cntr/cntr_equilibrium_impl.hpp:873-953,
 * 1152); H0 is a host size1 x size1 complex matrix. */
int cntrx_green_from_H(const cntrx_herm *G, double mu, const double *H0_host, double beta,
                       double h, void *stream);
/* cntr::green_from_H for a time-dependent H (libcntr/cntr/cntr_equilibrium_impl.hpp:754-870,
 * 955-1040; declarations cntr_equilibrium_decl.hpp:79-85, 93-98).  H is a function whose
 * data pointer is HOST memory, t = -1..nt (H(-1) the equilibrium Hamiltonian).  The
 * propagators U(t) are commutator-free exponentials of order cf_order (2: CF2, 4: CF4:2)
 * with H interpolated at order SolveOrder; they are formed on the host and G is filled on
 * the device.  tstp = CNTRX_ALL_TIMESTEPS fills every component; tstp >= -1 is the
 * reference's timestep form: only timestep tstp (mat for -1), with the reference's phase
 * convention for mu != 0 (U(tstp) multiplied by e^{i mu h n} for n = 0..tstp).  fixHam = 0
 * replaces H(t) for t > SolveOrder by its extrapolation, writing it into H as the reference
 * does; the public whole form of the reference passes fixHam = 1. */
#define CNTRX_ALL_TIMESTEPS (-2)
int cntrx_green_from_H_td(const cntrx_herm *G, double mu, const cntrx_func *H_host, double beta,
                          double h, int SolveOrder, int cf_order, int fixHam, int tstp, void *stream);
/* cntr::force_matsubara_hermitian (libcntr/cntr/cntr_utilities_impl.hpp:2077-2087) */
int cntrx_force_matsubara_hermitian(const cntrx_herm *G, void *stream);
/* Timestep algebra: C(tstp) = sum_i w_i X_i(tstp) over ret row, tv row, les column
 * (tstp=-1: mat).  Replaces herm_matrix::set_timestep/incr_timestep/smul
 * (libcntr/cntr/cntr_herm_matrix_impl.hpp:2064-2400) and the k-sum get_loc /
 * Reduce_timestep (examples/programs/gw_selfene_impl.cpp:54-62,
 * libcntr/cntr/cntr_mpitools_impl.hpp:99-106). */
int cntrx_timestep_lincomb(int tstp, const cntrx_herm *C, int count, const cntrx_herm *X,
                           const double *w, void *stream);
/* herm_matrix_timestep::left_multiply / right_multiply with a scalar contour
 * function (libcntr/cntr/cntr_herm_matrix_timestep_impl.hpp:1570-1823), size1 = 1. */
int cntrx_timestep_scale_by_function(int tstp, const cntrx_herm *C, const cntrx_func *F_host,
                                     double weight, int left, void *stream);

/* Host<->device copy of timestep slices or whole components (the "Host<->device
 * slice copy" a drop-in must export; the reference keeps everything in host memory).
 * kind: 1 host->device, 2 device->host, 3 device->device.  Synchronises the stream. */
int cntrx_copy(void *dst, const void *src, size_t bytes, int kind, void *stream);
/* Device memory for callers that include no HIP header (the C++ drop-in cntr_amd.hpp):
 * storage of the same packed layout as the host arrays that herm_matrix and function
 * allocate with new[] (libcntr/cntr/cntr_herm_matrix_impl.hpp:62-109,
 * libcntr/cntr/cntr_function_impl.hpp).  Free synchronises the device. */
int cntrx_device_alloc(size_t bytes, void **out);
int cntrx_device_free(void *p);

/* ---- text checkpoints ---------------------------------------------------------- */
/* The reference's plain-text files, byte for byte ("%.*g" at the given precision; 17
 * digits reproduce every double exactly).  The descriptors point at device storage; each
 * call copies through host memory and synchronises the stream.  Where the reference would
 * resize G (a file with more timesteps, another ntau or size1), the read fails with
 * CNTRX_EINVAL: read the header, reallocate, call again.  A short or malformed file gives
 * CNTRX_EINVAL and leaves G unchanged. */
/* herm_matrix<T>::print_to_file (libcntr/cntr/cntr_herm_matrix_impl.hpp:1308-1359) */
int cntrx_herm_print_to_file(const cntrx_herm *G, const char *file, int precision, void *stream);
/* the header line "# nt ntau size1 sig" of such a file (host only) */
int cntrx_herm_file_header(const char *file, int *nt, int *ntau, int *size1, int *sig);
/* herm_matrix<T>::read_from_file(file) (libcntr/cntr/cntr_herm_matrix_impl.hpp:1376-1448):
 * mat and timesteps 0..n of the file's n <= G->nt; *sig (if not NULL) gets the file's sig */
int cntrx_herm_read_from_file(const cntrx_herm *G, const char *file, int *sig, void *stream);
/* herm_matrix<T>::read_from_file(nt1, file) (libcntr/cntr/cntr_herm_matrix_impl.hpp:1469-1566):
 * mat and timesteps 0..nt1; needs nt1 <= G->nt, nt1 <= the file's nt, equal ntau and size1 */
int cntrx_herm_read_from_file_nt1(const cntrx_herm *G, int nt1, const char *file, void *stream);
/* function<T>::print_to_file (libcntr/cntr/cntr_function_impl.hpp:719-732) */
int cntrx_function_print_to_file(const cntrx_func *F, const char *file, int precision, void *stream);
/* the header line "# nt size1" (host only) */
int cntrx_function_file_header(const char *file, int *nt, int *size1);
/* function<T>::read_from_file(file) (libcntr/cntr/cntr_function_impl.hpp:750-777) */
int cntrx_function_read_from_file(const cntrx_func *F, const char *file, void *stream);
/* function<T>::read_from_file(nt1, file) (libcntr/cntr/cntr_function_impl.hpp:795-824) */
int cntrx_function_read_from_file_nt1(const cntrx_func *F, int nt1, const char *file, void *stream);

/* ---- k-batched lattice driver ------------------------------------------------- */
/* One GPU holds every k-point of a size-1 lattice with a local second-order
 * self-energy Sigma(t,t') = U(t)U(t') G_loc(t,t')^2 G_loc(t',t) built with
 * Bubble1/Bubble2 (U(-1) = 0, U(t>=0) = U) and G_loc = (1/Nk) sum_k G_k.  Replaces the
 * caller loop of examples/programs/hubbard_chain_2b.cpp:310-322 and the MPI
 * k-parallel loop of examples/programs/gw.cpp:386-430 (block distribution
 * libcntr/cntr/cntr_distributed_array_impl.hpp:88-152) on a single device. */
typedef struct cntrx_lattice cntrx_lattice;
int cntrx_lattice_create(int nk, const double *epsk, double U, double mu, double beta, int ntau,
                         int nt, double h, int SolveOrder, cntrx_lattice **out);
int cntrx_lattice_destroy(cntrx_lattice *L);
/* which >= 0: G_k; -1: G_loc; -2: Sigma (device descriptors) */
int cntrx_lattice_matrix(cntrx_lattice *L, int which, cntrx_herm *out);
/* dyson_mat for every k (Sigma^M = 0) */
int cntrx_lattice_equilibrium(cntrx_lattice *L, void *stream);
/* self-consistent dyson_start for steps 0..SolveOrder */
int cntrx_lattice_bootstrap(cntrx_lattice *L, double tol, int maxiter, int *iters, void *stream);
/* one physical timestep: predictor + corrector iterations until
 * distance_norm2(G_loc_new, G_loc_old) < tol; batched != 0 advances all k per launch,
 * batched == 0 calls cntrx_dyson_timestep once per k. */
int cntrx_lattice_step(cntrx_lattice *L, int tstp, int batched, double tol, int maxiter, int *iters,
                       double *err, void *stream);
/* Batched timesteps t0..t1 driven from the device: the same predictor and corrector
 * iterations as cntrx_lattice_step(L, t, 1, ...) for t = t0..t1 in turn, bit for bit, but
 * the host never waits for a timestep to end.  Each timestep's kernels begin on the
 * device once the previous timestep has converged (or reached maxiter); the host keeps
 * one chunk of iterations queued ahead and reads the iteration counts afterwards.
 * iters[t - t0] and err[t - t0] receive each timestep's corrector iterations and last
 * distance.  Profile modes 0 and 6 only (6 times one kernel in every fourth chunk, in its
 * first iteration).  A chunk's kernels are launched one by one; with CNTRX_LATTICE_GRAPHS=1
 * in the environment at cntrx_lattice_create, each chunk is one captured graph instead
 * (the same kernels and results).  No reference counterpart: it replaces the caller's timestep loop
 * (examples/programs/hubbard_chain_2b.cpp:310-325). */
int cntrx_lattice_run(cntrx_lattice *L, int t0, int t1, double tol, int maxiter, int *iters,
                      double *err, void *stream);
/* Time the batched Dyson kernels (0 k_tv, 1 k_retrec, 2 lesser convolution, 3 k_lesrec)
 * with HIP events on the lattice stream.  mode 0 off; 1 all four in every iteration
 * (direct launches); 2+q kernel q only (the rest of the iteration replayed as graphs);
 * 6 sampled: every 8th iteration times one kernel, the four in turn, and the other
 * iterations run as the full graph.  Setting a mode resets the counters; stats give
 * the summed milliseconds and the number of timed launches per kernel, and the
 * algorithmic HBM bytes of the timed k_tv launches. */
int cntrx_lattice_profile(cntrx_lattice *L, int mode);
int cntrx_lattice_kernel_stats(cntrx_lattice *L, double *ms4, long *launches4, double *tv_bytes);
/* The timed launches one by one (kernel 0..3, timestep, milliseconds), up to cap of
 * them; *count gets the number recorded since the mode was set. */
int cntrx_lattice_kernel_samples(cntrx_lattice *L, long cap, int *kernel, int *tstp, float *ms,
                                 long *count);
/* 1 when the lattice runs the mixed component and the retarded row of an iteration as one
 * launch (k_tvret1: then kernel 0 of the timing above is that launch and kernel 1 times
 * nothing), 0 when they run as two kernels on two streams.  No reference counterpart:
 * the lattice driver replaces the caller loops of examples/programs/hubbard_chain_2b.cpp. */
int cntrx_lattice_fused(cntrx_lattice *L, int *fused);

/* ---- continuous batching of independent solves ------------------------------------ */
/* Requests are independent Dyson propagations of uneven size1, nt, ntau and beta.  Each
 * runs the reference's dyson(...) sequence (libcntr/cntr/cntr_dyson_impl.hpp:1725-1734:
 * dyson_mat, dyson_start, dyson_timestep for n = SolveOrder+1..nt) with its own H(t) and
 * a bath self-energy Sigma = V^2 g[H_bath] (cntr::green_from_H).  No reference counterpart:
 * it replaces a caller's loop over solves.  One cntrx_server_step retires solves whose last
 * timestep has completed, admits queued requests first come first served while their
 * footprint fits the server's device pool, and advances every live solve by one timestep
 * in one cntrx_dyson_timestep_batched call.  With the autotuner's table fixed, a solve's
 * result does not depend on what else is live, bit for bit.  Calls on one server must come
 * from one thread; the server uses its own stream. */
typedef struct cntrx_server cntrx_server;
/* device bytes one request holds in the pool (G, Sigma, H); 0 if the request is invalid */
size_t cntrx_request_footprint(const cntrx_request *r);
/* a server with a device pool of budget_bytes (one allocation) */
int cntrx_server_create(size_t budget_bytes, cntrx_server **out);
int cntrx_server_destroy(cntrx_server *S);
/* queue one request (its arrays are copied); *id identifies it from now on.  A request
 * whose footprint exceeds the pool returns CNTRX_ENOMEM. */
int cntrx_server_submit(cntrx_server *S, const cntrx_request *r, long *id);
/* one tick: retire, admit, advance; *live and *queued after it (either may be NULL) */
int cntrx_server_step(cntrx_server *S, int *live, int *queued);
/* requests completed since the last poll: ids and latency (seconds from submission to
 * the tick that saw the last timestep complete) */
int cntrx_server_poll(cntrx_server *S, long cap, long *ids, double *latency_s, long *count);
/* step until every queued and live request has completed */
int cntrx_server_drain(cntrx_server *S);
/* device descriptor of a completed request's G (valid until cntrx_server_release) */
int cntrx_server_result(cntrx_server *S, long id, cntrx_herm *G);
int cntrx_server_release(cntrx_server *S, long id);
/* ticks, timesteps, admitted, completed, pool bytes used, pool peak, pool capacity, live */
int cntrx_server_stats(cntrx_server *S, double *out8);

#ifdef __cplusplus
}
#endif
#endif
