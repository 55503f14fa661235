// Work descriptors shared by the launch layer (api.hip) and the kernels.
#pragma once
#include "cntrx_device.h"

namespace cntrx {

// Kernel modes for the shared time-step kernels.
enum StepMode : int {
    MODE_CONV = 0,   // C = A*B                 (convolution_timestep)
    MODE_DYSON = 1,  // G from Dyson, A=Sigma   (dyson_timestep)
    MODE_VIE2 = 2    // G from [1+F]G=Q, A=F    (vie2_timestep)
};

// One timestep of one contour problem.
//   MODE_DYSON: out = B = Bcc = G, A = Acc = Sigma, H set.
//   MODE_VIE2 : out = B = Bcc = G, A = F, Acc = Fcc, Q set.  The retarded kernel is
//               launched with A/Acc swapped, as the reference calls
//               vie2_timestep_ret(n, G, Fcc, F, Q) (libcntr/cntr/cntr_vie2_impl.hpp:1615).
//   MODE_CONV : out = C and A, Acc, B, Bcc as given.
struct Job {
    HermView out, A, Acc, B, Bcc, Q;
    GPtr<const cplx> H;                // cntr::function data, H(t) at (t+1)*es
    GPtr<cplx> gles, qqret, qqles, small;  // per-job device workspace
    GPtr<const int> gate;              // optional: the kernels return at once when *gate == 0
    GPtr<cplx> bm;                     // size 1, own Sigma in a batch: K x 16 lesser-convolution operand
    int n;                             // timestep
    int k;                             // integration order for this job
    double mu, beta, h;
};

// device-side iteration control: a job whose gate reads 0 is skipped by every kernel
__device__ inline bool job_gated_off(const Job &J) {
    const int *g = J.gate;
    return g != nullptr && *g == 0;
}

// workspace per job, in units of es complex numbers (ntmax = max(nt, k))
inline int64_t job_workspace_blocks(int ntmax) { return 3 * (int64_t)(ntmax + 1) + 8; }

}  // namespace cntrx
