// libcntrx: host launch layer and C ABI (single translation unit).
//
// Builds the Integrator tables, validates descriptors, lays out per-call device
// workspace, and sequences the kernels of kernels_step.inc / kernels_misc.inc.
// Every exported function returns a status code (cntrx.h) instead of asserting.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "cntrx_jobs.h"

#include "kernels_step.inc"
#include "kernels_misc.inc"
namespace cntrx {
#include "kernels_rec1.inc"
#include "kernels_rec1p.inc"
#include "kernels_recmw.inc"
#include "kernels_mm.inc"
#include "kernels_zgemm.inc"
#include "kernels_incr.inc"
}  // namespace cntrx
namespace cntrx {
#include "lattice.inc"
}  // namespace cntrx

using namespace cntrx;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

#define HIPCHK(x)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess)                                                               \
            return fail(CNTRX_EHIP, "%s failed: %s (%s:%d)", #x, hipGetErrorString(e_),     \
                        __FILE__, __LINE__);                                                \
    } while (0)

#define LAUNCHCHK()                                                                         \
    do {                                                                                    \
        hipError_t e_ = hipGetLastError();                                                  \
        if (e_ != hipSuccess)                                                               \
            return fail(CNTRX_EHIP, "kernel launch failed: %s (%s:%d)", hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                                \
    } while (0)

// ------------------------------------------------------------------------------------
// Integrator tables by exact rational arithmetic (integration.hpp:128-514):
// interpolation = inverse Vandermonde on nodes 0..k; differentiation/integration of
// the interpolant; BDF(k+1) = derivative at node k+1 of the order-(k+1) interpolant;
// Gregory end weights fixed by exactness for degree <= k (Euler-Maclaurin end terms);
// start table = poly_integration(0,n,.) for n<=k and overlapping end corrections for
// k<n<2k+2; rcorr(m,j,l) = int_0^m L_j(m-x) L_l(x) dx.
// ------------------------------------------------------------------------------------
struct Q {
    __int128 p = 0, q = 1;
    Q() {}
    Q(long long a) : p(a), q(1) {}
    Q(__int128 a, __int128 b) : p(a), q(b) { norm(); }
    static __int128 gcd(__int128 a, __int128 b) {
        if (a < 0) a = -a;
        if (b < 0) b = -b;
        while (b) {
            __int128 t = a % b;
            a = b;
            b = t;
        }
        return a;
    }
    void norm() {
        if (q < 0) {
            p = -p;
            q = -q;
        }
        __int128 d = gcd(p, q);
        if (d > 1) {
            p /= d;
            q /= d;
        }
    }
    Q operator+(const Q &o) const { return Q(p * o.q + o.p * q, q * o.q); }
    Q operator-(const Q &o) const { return Q(p * o.q - o.p * q, q * o.q); }
    Q operator*(const Q &o) const { return Q(p * o.p, q * o.q); }
    Q operator/(const Q &o) const { return Q(p * o.q, q * o.p); }
    double d() const { return (double)((long double)p / (long double)q); }
};

Q qpow(long long x, int e) {
    Q r(1);
    for (int i = 0; i < e; i++) r = r * Q(x);
    return r;
}

std::vector<Q> qinverse(int n, std::vector<Q> a) {
    std::vector<Q> inv(n * n, Q(0));
    for (int i = 0; i < n; i++) inv[i * n + i] = Q(1);
    for (int c = 0; c < n; c++) {
        int p = c;
        while (a[p * n + c].p == 0) p++;
        for (int j = 0; j < n; j++) {
            std::swap(a[c * n + j], a[p * n + j]);
            std::swap(inv[c * n + j], inv[p * n + j]);
        }
        Q d = a[c * n + c];
        for (int j = 0; j < n; j++) {
            a[c * n + j] = a[c * n + j] / d;
            inv[c * n + j] = inv[c * n + j] / d;
        }
        for (int r = 0; r < n; r++) {
            if (r == c || a[r * n + c].p == 0) continue;
            Q f = a[r * n + c];
            for (int j = 0; j < n; j++) {
                a[r * n + j] = a[r * n + j] - f * a[c * n + j];
                inv[r * n + j] = inv[r * n + j] - f * inv[c * n + j];
            }
        }
    }
    return inv;
}

std::vector<Q> lagrange_coeffs(int k) {  // c[alpha*(k+1)+l]
    int k1 = k + 1;
    std::vector<Q> V(k1 * k1);
    for (int i = 0; i <= k; i++)
        for (int n = 0; n <= k; n++) V[i * k1 + n] = qpow(i, n);
    return qinverse(k1, V);
}

void build_order(int k, WeightTable &T) {
    int k1 = k + 1, k2 = 2 * k1;
    std::vector<Q> c = lagrange_coeffs(k);
    for (int i = 0; i < k1 * k1; i++) T.P[k][i] = c[i].d();
    for (int i = 0; i <= k; i++)
        for (int l = 0; l <= k; l++) {
            Q s;
            for (int n = 1; n <= k; n++) s = s + Q(n) * qpow(i, n - 1) * c[n * k1 + l];
            T.D[k][i * k1 + l] = s.d();
        }
    std::vector<Q> W(k1 * k1 * k1);
    for (int i = 0; i <= k; i++)
        for (int j = 0; j <= k; j++)
            for (int l = 0; l <= k; l++) {
                Q s;
                for (int r = 0; r <= k; r++)
                    s = s + c[r * k1 + l] * (qpow(j, r + 1) - qpow(i, r + 1)) / Q(r + 1);
                W[(i * k1 + j) * k1 + l] = s;
                T.W[k][(i * k1 + j) * k1 + l] = s.d();
            }
    {
        int q = k + 1, q1 = q + 1;
        std::vector<Q> cq = lagrange_coeffs(q);
        for (int l = 0; l <= q; l++) {
            Q s;
            for (int n = 1; n <= q; n++) s = s + Q(n) * qpow(q, n - 1) * cq[n * q1 + (q - l)];
            T.bd[k][l] = s.d();
        }
    }
    const Q bern[] = {Q(1), Q(-1, 2), Q(1, 6), Q(0), Q(-1, 30), Q(0), Q(1, 42)};
    std::vector<Q> V(k1 * k1);
    for (int p = 0; p <= k; p++)
        for (int j = 0; j <= k; j++) V[p * k1 + j] = qpow(j, p);
    std::vector<Q> Vi = qinverse(k1, V), R(k1), om(k1);
    for (int p = 0; p <= k; p++) {
        if (p == 0) R[p] = Q(-1, 2);
        else if (p % 2 == 1) R[p] = bern[p + 1] / Q(p + 1);
    }
    for (int j = 0; j <= k; j++) {
        Q a;
        for (int p = 0; p <= k; p++) a = a + Vi[j * k1 + p] * R[p];
        om[j] = Q(1) + a;
        T.omega[k][j] = om[j].d();
    }
    for (int i = 0; i < k2 * k2; i++) T.greg[k][i] = 0.0;
    for (int n = 0; n <= k; n++)
        for (int i = 0; i <= k; i++) T.greg[k][n * k2 + i] = W[(0 * k1 + n) * k1 + i].d();
    for (int n = k1; n < k2; n++)
        for (int i = 0; i <= n; i++) {
            Q w(1);
            if (i <= k) w = w + om[i] - Q(1);
            if (i >= n - k) w = w + om[n - i] - Q(1);
            T.greg[k][n * k2 + i] = w.d();
        }
    for (int m = 1; m < k; m++)
        for (int j = 0; j <= k; j++) {
            std::vector<Q> a(k1);
            for (int e = 0; e <= k; e++) {
                long long binom = 1;
                for (int q = 0; q <= e; q++) {
                    Q term = c[e * k1 + j] * Q(binom) * qpow(m, e - q);
                    a[q] = (q % 2) ? a[q] - term : a[q] + term;
                    binom = binom * (e - q) / (q + 1);
                }
            }
            for (int l = 0; l <= k; l++) {
                std::vector<Q> prod(2 * k1);
                for (int p = 0; p <= k; p++)
                    for (int q = 0; q <= k; q++) prod[p + q] = prod[p + q] + a[p] * c[q * k1 + l];
                Q s;
                for (int e = 0; e < 2 * k1; e++) s = s + prod[e] * qpow(m, e + 1) / Q(e + 1);
                T.rc[k][((m - 1) * k1 + j) * k1 + l] = s.d();
            }
        }
}

const WeightTable &host_weights() {
    static WeightTable T;
    static std::once_flag once;
    std::call_once(once, [] {
        memset(&T, 0, sizeof T);
        for (int k = 0; k <= 5; k++) build_order(k, T);
    });
    return T;
}

// ------------------------------------------------------------------------------------
// per-device state: uploaded weights and a grow-only workspace
// ------------------------------------------------------------------------------------
// Scratch of the no-allocation entry point (cntrx_dyson_timestep_noalloc): one fixed
// array in the code object, so each device that loads the library holds its own copy
// until the process ends.  Host code must never pass the array's name to a kernel or a
// copy (on the host it is only a stand-in symbol): it uses the device address that
// ensure_init() takes with hipGetSymbolAddress on each device (DevState::scratch).
constexpr size_t NOALLOC_SCRATCH_BYTES = size_t(16) << 20;
static __device__ __align__(256) char scratch[NOALLOC_SCRATCH_BYTES];

struct DevState {
    bool ready = false;
    void *ws = nullptr;
    size_t ws_bytes = 0;
    double *red = nullptr;      // reduction scalar
    double *red_host = nullptr; // pinned host mirror
    char *scratch = nullptr;    // this device's address of `scratch`
};
std::mutex g_mu;
DevState g_dev[64];

int ensure_init() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lk(g_mu);
    DevState &D = g_dev[dev];
    if (D.ready) return CNTRX_OK;
    const WeightTable &T = host_weights();
    HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_w), &T, sizeof(WeightTable)));
    HIPCHK(hipMalloc(&D.red, 64));
    HIPCHK(hipHostMalloc(&D.red_host, 64));
    void *sp = nullptr;
    const hipError_t se = hipGetSymbolAddress(&sp, HIP_SYMBOL(scratch));
    if (se != hipSuccess || !sp)
        return fail(CNTRX_EHIP, "no device address for the no-alloc scratch on device %d: %s", dev,
                    hipGetErrorString(se));
    D.scratch = (char *)sp;
    D.ready = true;
    return CNTRX_OK;
}

// Workspace: one grow-only allocation per device, shared by all calls on that device.
// Growing synchronises the device first.  Calls that overlap on different streams of
// one device must not share it (documented in cntrx.h).
int workspace(size_t bytes, void **out) {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    DevState &D = g_dev[dev];
    if (bytes > D.ws_bytes) {
        HIPCHK(hipDeviceSynchronize());
        if (D.ws) HIPCHK(hipFree(D.ws));
        D.ws = nullptr;
        size_t nb = std::max(bytes, (size_t)(1 << 20));
        nb = (nb + 4095) & ~(size_t)4095;
        if (hipMalloc(&D.ws, nb) != hipSuccess) {
            D.ws_bytes = 0;
            (void)hipGetLastError();
            return fail(CNTRX_ENOMEM, "workspace allocation of %zu bytes failed", nb);
        }
        D.ws_bytes = nb;
    }
    *out = D.ws;
    return CNTRX_OK;
}

// simple bump allocator over the workspace
struct Carve {
    char *base;
    size_t off = 0;
    explicit Carve(void *b) : base((char *)b) {}
    template <class T>
    T *take(size_t n) {
        off = (off + 255) & ~(size_t)255;
        T *p = (T *)(base + off);
        off += n * sizeof(T);
        return p;
    }
    static size_t round(size_t b) { return (b + 255) & ~(size_t)255; }
};

bool supported_size(int s) { return s == 1 || s == 2 || s == 4 || s == 8; }

// any_size: the caller's kernels take size1 at run time (every size1 >= 1)
int check_herm(const cntrx_herm *G, const char *name, int need_nt, bool need_rt = true, bool any_size = false) {
    if (!G) return fail(CNTRX_EINVAL, "%s is NULL", name);
    if (any_size ? G->size1 < 1 : !supported_size(G->size1))
        return fail(CNTRX_EUNSUPPORTED, "%s: size1=%d not supported (1..8)", name, G->size1);
    if (G->sig != 1 && G->sig != -1) return fail(CNTRX_EINVAL, "%s: sig must be +-1", name);
    if (G->ntau < 1 || !G->mat) return fail(CNTRX_EINVAL, "%s: bad ntau/mat", name);
    if (need_rt) {
        if (G->nt < need_nt) return fail(CNTRX_EINVAL, "%s: nt=%d < %d", name, G->nt, need_nt);
        if (!G->ret || !G->les || !G->tv) return fail(CNTRX_EINVAL, "%s: NULL component", name);
    }
    return CNTRX_OK;
}

int same_shape(const cntrx_herm *A, const cntrx_herm *B, const char *na, const char *nb) {
    if (A->size1 != B->size1 || A->ntau != B->ntau)
        return fail(CNTRX_EINVAL, "%s and %s differ in size1/ntau", na, nb);
    return CNTRX_OK;
}

#define RC(x)                         \
    do {                              \
        int rc_ = (x);                \
        if (rc_ != CNTRX_OK) return rc_; \
    } while (0)

#define DISPATCH_S(s, ...)                 \
    switch (s) {                           \
        case 1: {                          \
            constexpr int S = 1;           \
            __VA_ARGS__;                   \
        } break;                           \
        case 2: {                          \
            constexpr int S = 2;           \
            __VA_ARGS__;                   \
        } break;                           \
        case 4: {                          \
            constexpr int S = 4;           \
            __VA_ARGS__;                   \
        } break;                           \
        case 8: {                          \
            constexpr int S = 8;           \
            __VA_ARGS__;                   \
        } break;                           \
        default:                           \
            return fail(CNTRX_EUNSUPPORTED, "size1 %d", s); \
    }


inline int nblocks(int64_t n, int per) { return (int)((n + per - 1) / per); }

// ------------------------------------------------------------------------------------
// Staging of per-call device data: jobs, index lists (copied from a pinned host
// image) and per-job scratch (not copied), carved from the device workspace.
// ------------------------------------------------------------------------------------
struct Stage {
    std::vector<Job> *jobs;
    std::vector<int> ntmax;          // per job: max(nt, k)
    std::vector<std::vector<int2>> lists;
    // results
    Job *djobs = nullptr;
    std::vector<int2 *> dlists;
    cplx *extra = nullptr;           // optional extra device scratch (complex elements)
    size_t extra_elems = 0;
    std::vector<int64_t> bm_elems;   // optional, per job: elements of its own Job::bm operand
};

struct Pinned {
    char *buf = nullptr;
    size_t bytes = 0;
    hipEvent_t ev = nullptr;
    bool pending = false;
};
Pinned g_pin[64];

// The staging image's event may have been recorded on a stream that the library is
// about to destroy (a lattice's or engine's private stream): wait for it while the
// stream still exists and replace it by a fresh event, so that no later call touches
// an event whose stream is gone.
int pin_quiesce() {
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    Pinned &P = g_pin[dev];
    if (!P.ev) return CNTRX_OK;
    if (P.pending) HIPCHK(hipEventSynchronize(P.ev));
    P.pending = false;
    HIPCHK(hipEventDestroy(P.ev));
    P.ev = nullptr;
    HIPCHK(hipEventCreateWithFlags(&P.ev, hipEventDisableTiming));
    return CNTRX_OK;
}

int stage_upload(Stage &sg, hipStream_t st) {
    std::vector<Job> &jobs = *sg.jobs;
    auto es_of = [&](size_t i) { return (int64_t)jobs[i].out.s * jobs[i].out.s; };
    size_t img = Carve::round(std::max<size_t>(jobs.size(), 1) * sizeof(Job));
    for (auto &l : sg.lists) img += Carve::round((l.size() + 1) * sizeof(int2));
    size_t scratch = 0;
    for (size_t i = 0; i < jobs.size(); i++)
        scratch += Carve::round(job_workspace_blocks(sg.ntmax[i]) * es_of(i) * sizeof(cplx));
    scratch += Carve::round(sg.extra_elems * sizeof(cplx));
    for (int64_t e : sg.bm_elems) scratch += Carve::round(e * sizeof(cplx));
    void *base;
    RC(workspace(img + scratch + 512, &base));
    Carve cv(base);
    sg.djobs = cv.take<Job>(std::max<size_t>(jobs.size(), 1));
    sg.dlists.clear();
    for (auto &l : sg.lists) sg.dlists.push_back(cv.take<int2>(l.size() + 1));
    size_t img_end = cv.off;
    for (size_t i = 0; i < jobs.size(); i++) {
        const int64_t nb = sg.ntmax[i] + 1, ES = es_of(i);
        cplx *sc = cv.take<cplx>(job_workspace_blocks(sg.ntmax[i]) * ES);
        jobs[i].gles = sc;
        jobs[i].qqret = sc + nb * ES;
        jobs[i].qqles = sc + 2 * nb * ES;
        jobs[i].small = sc + 3 * nb * ES;
    }
    if (sg.extra_elems) sg.extra = cv.take<cplx>(sg.extra_elems);
    for (size_t i = 0; i < sg.bm_elems.size() && i < jobs.size(); i++)
        if (sg.bm_elems[i] > 0) jobs[i].bm = cv.take<cplx>(sg.bm_elems[i]);
    // host image in pinned memory
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    Pinned &P = g_pin[dev];
    if (P.pending) {
        HIPCHK(hipEventSynchronize(P.ev));
        P.pending = false;
    }
    if (P.bytes < img_end) {
        if (P.buf) HIPCHK(hipHostFree(P.buf));
        P.bytes = std::max(img_end, (size_t)(1 << 16));
        HIPCHK(hipHostMalloc((void **)&P.buf, P.bytes));
        if (!P.ev) HIPCHK(hipEventCreateWithFlags(&P.ev, hipEventDisableTiming));
    }
    if (!jobs.empty())
        memcpy(P.buf + ((char *)sg.djobs - (char *)base), jobs.data(), jobs.size() * sizeof(Job));
    for (size_t l = 0; l < sg.lists.size(); l++)
        if (!sg.lists[l].empty())
            memcpy(P.buf + ((char *)sg.dlists[l] - (char *)base), sg.lists[l].data(),
                   sg.lists[l].size() * sizeof(int2));
    HIPCHK(hipMemcpyAsync(base, P.buf, img_end, hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(P.ev, st));
    P.pending = true;
    return CNTRX_OK;
}

// tau points per tile of the mixed-component kernel (k_tv1 for size 1, else k_tv)
inline int tv_tile_points(int S, int ntau) { return (S == 1 && ntau < NT1) ? TV1P : 256 / (S * S); }

void tiles_over(std::vector<int2> &out, int job, int last, int per) {
    for (int j0 = 0; j0 <= last; j0 += per) out.push_back(make_int2(job, j0));
}

// tv tiles of `count` jobs of the same ntau in an XCD-aware order: workgroups are dealt to
// the eight XCDs in turn (workgroup b to XCD b % 8), so job q's tiles go to workgroups with
// b % 8 == q % 8 and share that XCD's L2 (the cache lines a row's tile boundaries split,
// and the job's B^M).  Slots past a group's end hold (0, -1), which k_tv1 skips.
// short_last: within each XCD's list the tiles that hold fewer than `per` points (the
// last of every job) come after all full ones, so that the round of tiles left when
// the retarded rows of the fused launch end is not set by a full tile dealt after them.
std::vector<int2> xcd_tiles(int count, int ntau, int per, bool short_last = false) {
    constexpr int NX = 8;
    std::vector<int2> grp[NX];
    for (int q = 0; q < count; q++) tiles_over(grp[q % NX], q, ntau, per);
    if (short_last)
        for (auto &g : grp)
            std::stable_partition(g.begin(), g.end(), [&](const int2 &t) { return t.y + per <= ntau + 1; });
    size_t len = 0;
    for (auto &g : grp) len = std::max(len, g.size());
    std::vector<int2> out(NX * len, make_int2(0, -1));
    for (int x = 0; x < NX; x++)
        for (size_t i = 0; i < grp[x].size(); i++) out[NX * i + x] = grp[x][i];
    return out;
}

Job make_job(const cntrx_herm *out, const cntrx_herm *A, const cntrx_herm *Acc,
             const cntrx_herm *B, const cntrx_herm *Bcc, const cntrx_herm *Qm, const cntrx_func *H,
             int n, int k, double mu, double beta, double h) {
    Job J;
    memset((void *)&J, 0, sizeof J);
    J.out = HermView(*out);
    if (A) J.A = HermView(*A);
    if (Acc) J.Acc = HermView(*Acc);
    if (B) J.B = HermView(*B);
    if (Bcc) J.Bcc = HermView(*Bcc);
    if (Qm) J.Q = HermView(*Qm);
    J.H = H ? (const cplx *)H->data : nullptr;
    J.n = n;
    J.k = k;
    J.mu = mu;
    J.beta = beta;
    J.h = h;
    return J;
}

// ------------------------------------------------------------------------------------
// Time-step sequence: tv | ret  ->  les convolution -> les recurrence
// lists: 0 = tv tiles, 1 = les tiles.  Job arrays: [0,count) main jobs; for VIE2 the
// retarded kernel uses jobs [count, 2count) with A/Acc swapped.
// ------------------------------------------------------------------------------------
// the recurrences: wavefront-synchronous scalar kernels when size1 = 1 and n fits LDS
// size1 >= 2: panel recurrences on many workgroups (kernels_recmw.inc), after the
// reset of their ticket/flag words; nmax bounds n over the jobs (the grid covers the
// panels of the largest one, the others' extra workgroups exit at once).  The panel
// size comes from the tuning table: the autotuner's choice for (size1, history-length
// bucket), else the default RecShape<S>::P0.  Batched calls look it up per solve, by
// the solve's own n (RecSplit), so that a solve gets the same panel size, and the same
// result bit for bit, alone and in any batch once the table is fixed.
int g_panel_choice[4][16];  // [size class][bucket]: 0 untuned (P0 used), 1 + choice

// Autotuner (cntrx_autotune): with it on, the first batched timestep call that meets an
// untuned (size1, history-length bucket) times the candidate launch shapes on its own live
// jobs, keeps the fastest for the rest of the process and records the decision.
//   size1 2, 4: the panel size of the recurrences (RecShape<S>::P0, P1);
//   size1 8 with a shared Sigma: the workgroup target of the split-K GEMMs (1024, 2048, 4096).
// The candidates sum in different orders, so results can differ between them at rounding
// level.  With the tuner off (the default) the table stays as it is (untuned: the first
// candidate everywhere); decisions made while it was on stay in use after it is turned off.
bool g_autotune = false;
int g_split_choice[16];  // size1 8 GEMM split target by bucket: 0 untuned, 1 + index
const int kSplitTargets[3] = {2048, 1024, 4096};
struct TuneRec {
    int S, bucket, what, ncand, choice;  // what: 0 recurrence panel, 1 GEMM split target
    float ms[3];
};
std::vector<TuneRec> g_tuned;
inline int size_class(int S) { return S == 1 ? 0 : S == 2 ? 1 : S == 4 ? 2 : 3; }
inline int hist_bucket(int nmax) {  // 0: n < 64, then one bucket per power of two
    int b = 0;
    while (b < 15 && (64 << b) <= nmax) b++;
    return b;
}
inline int &panel_slot(int S, int nmax) { return g_panel_choice[size_class(S)][hist_bucket(nmax)]; }
inline int panel_choice(int S, int nmax) { return std::max(0, panel_slot(S, nmax) - 1); }

template <int S, int P, int MODE>
void launch_recmw(bool les, const Job *jobs, int count, int nmax, hipStream_t st) {
    const int gx = (nmax + P - 1) / P + 2;
    hipLaunchKernelGGL(k_rmw_reset, dim3(count), dim3(256), 0, st, jobs, les ? 1 : 0, gx + 2);
    if (les) hipLaunchKernelGGL((k_lesmw<S, P, MODE>), dim3(gx * count), dim3(RMW_T), 0, st, jobs, count);
    else hipLaunchKernelGGL((k_retmw<S, P, MODE>), dim3(gx * count), dim3(RMW_T), 0, st, jobs, count);
}
template <int S, int MODE>
void launch_rec_choice(bool les, int choice, const Job *jobs, int count, int nmax, hipStream_t st) {
    if (choice == 1) launch_recmw<S, RecShape<S>::P1, MODE>(les, jobs, count, nmax, st);
    else launch_recmw<S, RecShape<S>::P0, MODE>(les, jobs, count, nmax, st);
}

// Size1 2, 4: the jobs of a launch ordered by the panel size of their own history-length
// bucket, jobs [0, cnt0) with the first candidate, the rest with the second, each part
// launched on its own, so that a solve's panel size, and with it its order of summation,
// does not depend on the other solves of the launch.
struct RecSplit {
    int cnt0 = 0, nmax0 = 0, nmax1 = 0;
};
template <int S, int MODE>
void launch_rec_split(bool les, const RecSplit &rs, const Job *jobs, int count, hipStream_t st) {
    if (rs.cnt0 > 0) launch_rec_choice<S, MODE>(les, 0, jobs, rs.cnt0, rs.nmax0, st);
    if (count > rs.cnt0) launch_rec_choice<S, MODE>(les, 1, jobs + rs.cnt0, count - rs.cnt0, rs.nmax1, st);
}

// FAST (the Dyson timesteps): the pipelined size-1 retarded recurrence (the two
// instantiations differ only in the size of the history in LDS); otherwise k_retrec1
template <int S, int MODE, bool FAST = false>
void launch_ret(const Job *jobs, int count, int nmax, hipStream_t st, const RecSplit *rs = nullptr) {
    if constexpr (S == 1) {
        if (nmax < NQ1) {
            if constexpr (FAST && MODE == MODE_DYSON) {
                if (nmax < 512)
                    hipLaunchKernelGGL((k_retrec1p<512>), dim3(count), dim3(64 * PL_NW), 0, st, jobs);
                else
                    hipLaunchKernelGGL((k_retrec1p<NQ1>), dim3(count), dim3(64 * PL_NW), 0, st, jobs);
            }
            else
                hipLaunchKernelGGL((k_retrec1<MODE>), dim3(count), dim3(RT1), 0, st, jobs);
        }
        else hipLaunchKernelGGL((k_retrec<S, MODE>), dim3(count), dim3(256), 0, st, jobs);
    } else if (rs) {
        launch_rec_split<S, MODE>(false, *rs, jobs, count, st);
    } else {
        launch_rec_choice<S, MODE>(false, panel_choice(S, nmax), jobs, count, nmax, st);
    }
}
// FAST (the Dyson timesteps): the pipelined size-1 lesser recurrence; otherwise k_lesrec1
template <int S, int MODE, bool FAST = false>
void launch_les(const Job *jobs, int count, int nmax, hipStream_t st, const RecSplit *rs = nullptr) {
    if constexpr (S == 1) {
        if (nmax < NQ1) {
            if constexpr (FAST && MODE == MODE_DYSON)
                hipLaunchKernelGGL(k_lesrec1p, dim3(count), dim3(64 * PLL_NW), 0, st, jobs);
            else
                hipLaunchKernelGGL((k_lesrec1<MODE>), dim3(count), dim3(RT1), 0, st, jobs);
        }
        else hipLaunchKernelGGL((k_lesrec<S, MODE>), dim3(count), dim3(256), 0, st, jobs);
    } else if (rs) {
        launch_rec_split<S, MODE>(true, *rs, jobs, count, st);
    } else {
        launch_rec_choice<S, MODE>(true, panel_choice(S, nmax), jobs, count, nmax, st);
    }
}

// tv: tiles are (job, first tau) with 256/size1^2 points per tile; the scalar kernel
// needs ntau+1 <= NT1
template <int S, int MODE>
void launch_tv(const Job *jobs, const int2 *tiles, int ntiles, int ntau_max, hipStream_t st) {
    if (S == 1 && ntau_max < NT1)
        hipLaunchKernelGGL((k_tv1<MODE>), dim3(ntiles), dim3(256), 2 * (ntau_max + 1) * sizeof(cplx),
                           st, jobs, tiles);
    else
        hipLaunchKernelGGL((k_tv<S, MODE>), dim3(ntiles), dim3(256), 0, st, jobs, tiles);
}

// Size1 8 Dyson batches whose solves share Sigma (the same Sigma^tv/Sigma^</Sigma^R
// storage) and n, ntau, k, beta and sig run the Matsubara integral of the mixed component
// and the lesser convolution as complex GEMMs on the matrix cores (kernels_zgemm.inc).
// The plan gives the workspace layout: the packed left operand (K, then P, in one region
// reused in turn) and the solves' lesser-convolution column blocks after it.
struct ZgemmPlan {
    int64_t mi2_dim = 0, les_rows = 0, les_k = 0, bm_off = 0, part_off = 0, elems = 0;
    int mi2_split = 1, mi2_kc = 0, les_split = 1, les_kc = 0;
    bool tune = false;  // the autotuner times the split targets on this call first
    double dtau = 0.0;
};
// K is split into chunks (multiples of 16, at least 256 wide) until the product has about
// `target` workgroups (2048 unless the autotuner chose another); the choice depends only on
// the shapes and the target, so a batch always sums alike
inline void zgemm_split(int64_t M, int64_t N, int64_t K, int &nsplit, int &kc, int target = 2048) {
    const int64_t tiles = ((M + ZG_BM - 1) / ZG_BM) * ((N + ZG_BN - 1) / ZG_BN);
    int64_t want = std::max<int64_t>(1, (target + tiles - 1) / tiles);
    want = std::min<int64_t>(want, std::max<int64_t>(1, K / 256));
    kc = (int)((((K + want - 1) / want) + ZG_BK - 1) / ZG_BK * ZG_BK);
    nsplit = (int)((K + kc - 1) / kc);
}
bool zgemm_plan(const std::vector<Job> &jobs, int count, ZgemmPlan &zp) {
    if (count < 2) return false;
    const Job &a = jobs[0];
    if (a.out.s != 8) return false;
    for (int i = 0; i < count; i++) {
        const Job &b = jobs[i];
        if (b.n != a.n || b.k != a.k || b.out.ntau != a.out.ntau || b.beta != a.beta ||
            b.B.sig != a.B.sig || (const cplx *)b.A.tv != (const cplx *)a.A.tv ||
            (const cplx *)b.A.les != (const cplx *)a.A.les || (const cplx *)b.Acc.les != (const cplx *)a.Acc.les ||
            (const cplx *)b.A.ret != (const cplx *)a.A.ret || b.A.sig != a.A.sig)
            return false;
    }
    const int ntau = a.out.ntau, n1 = std::max(a.n, a.k);
    if (ntau < 2 * a.k + 2 || a.n <= a.k) return false;
    zp.mi2_dim = 8 * (int64_t)(ntau + 1);
    zp.les_rows = 8 * (int64_t)(n1 + 1);
    zp.les_k = 8 * (int64_t)(ntau + 2 + n1);
    zp.bm_off = std::max(zp.mi2_dim * zp.mi2_dim, zp.les_rows * zp.les_k);
    zp.bm_off = (zp.bm_off + 15) / 16 * 16;
    zp.part_off = zp.bm_off + ((int64_t)count * zp.les_k * 8 + 15) / 16 * 16;
    const int64_t N = 8 * (int64_t)count;
    const int slot = g_split_choice[hist_bucket(n1)];
    const int target = kSplitTargets[slot > 0 ? slot - 1 : 0];
    zp.tune = g_autotune && slot == 0;
    zgemm_split(zp.mi2_dim, N, zp.mi2_dim, zp.mi2_split, zp.mi2_kc, target);
    zgemm_split(zp.les_rows, N, zp.les_k, zp.les_split, zp.les_kc, target);
    int64_t part = 0;
    for (int c = 0; c < (zp.tune ? 3 : 1); c++) {
        const int tg = zp.tune ? kSplitTargets[c] : target;
        int ms, mk, ls, lk;
        zgemm_split(zp.mi2_dim, N, zp.mi2_dim, ms, mk, tg);
        zgemm_split(zp.les_rows, N, zp.les_k, ls, lk, tg);
        if (ms > 1) part = std::max(part, ms * zp.mi2_dim * N);
        if (ls > 1) part = std::max(part, ls * zp.les_rows * N);
    }
    zp.elems = zp.part_off + part;
    zp.dtau = a.beta / ntau;
    return true;
}

// The lesser-convolution sums of size-1 solves that share Sigma (same A/Acc, n, ntau,
// k, h, beta) form one GEMM over the solves (k_lesconv_mm); returns the leading
// dimension of its K x ldb operand, or 0 when the solves do not qualify.
int lesconv_mm_ldb(const std::vector<Job> &jobs, int count) {
    const Job &a = jobs[0];
    for (int i = 0; i < count; i++) {
        const Job &b = jobs[i];
        if (b.out.s != 1 || b.n != a.n || b.k != a.k || b.n <= b.k || b.out.ntau != a.out.ntau ||
            b.h != a.h || b.beta != a.beta || (const cplx *)b.A.tv != (const cplx *)a.A.tv ||
            (const cplx *)b.A.les != (const cplx *)a.A.les ||
            (const cplx *)b.Acc.les != (const cplx *)a.Acc.les || b.A.sig != a.A.sig)
            return 0;
    }
    return (count + 15) / 16 * 16;
}
inline int64_t lesconv_mm_elems(const Job &a, int ldb) {
    return (int64_t)(a.out.ntau + 2 + a.n) * ldb;
}

template <int MODE>
void launch_lesconv_mm(const Job *djobs, int count, int nmax, int ntau, cplx *Bm, int ldb,
                       hipStream_t st) {
    (void)ntau;
    (void)Bm;
    (void)ldb;
    hipLaunchKernelGGL((k_lesconv_mm<MODE>), dim3(nblocks(nmax + 1, 16), nblocks(count, 16)),
                       dim3(64 * MMW), 0, st, djobs, count);
}

// One timestep of `count` jobs: djobs[0..count) (VIE2: the retarded kernel reads the
// swapped jobs djobs[count..2count)), tv tiles tv[0..ntv), lesser-convolution tiles
// les[0..nles), and for the size-1 GEMM form of the lesser convolution (mm_ldb > 0) its
// K x mm_ldb operand in `bm`.
template <int S, int MODE>
int launch_timestep_raw(const Job *djobs, const int2 *tv, int ntv, const int2 *les, int nles,
                        int ntau_max, cplx *bm, int count, int nmax, hipStream_t st, int mm_ldb,
                        const ZgemmPlan *zp = nullptr, bool each_mm = false, const RecSplit *rs = nullptr) {
    const Job *ret_jobs = (MODE == MODE_VIE2) ? djobs + count : djobs;
    const bool zg = (S == 8 && MODE == MODE_DYSON && zp != nullptr);
    if constexpr (S == 8 && MODE == MODE_DYSON) {
        if (zg) {
            // Matsubara integral of every solve's tv row as one GEMM over the solves, then
            // the history sum, BDF terms and solve per tau point (k_tv8)
            const int64_t dim = zp->mi2_dim;
            cplx *part = bm + zp->part_off;
            hipLaunchKernelGGL((k_pack_mi2<8>), dim3(nblocks(dim * dim, 256)), dim3(256), 0, st, djobs, bm);
            hipLaunchKernelGGL((k_zgemm_cols<8, ZG_MI2>),
                               dim3(nblocks(8 * count, ZG_BN), nblocks(dim, ZG_BM), zp->mi2_split), dim3(256), 0, st,
                               (int)dim, (int)dim, count, (const cplx *)bm, dim, djobs, (const cplx *)nullptr,
                               (int64_t)0, zp->dtau, zp->mi2_kc, part);
            if (zp->mi2_split > 1)
                hipLaunchKernelGGL((k_zgemm_reduce<8, ZG_MI2>), dim3(nblocks(dim * 8 * count, 256)), dim3(256), 0, st,
                                   (int)dim, count, zp->mi2_split, (const cplx *)part, djobs, zp->dtau);
            LAUNCHCHK();
            hipLaunchKernelGGL(k_tv8, dim3(ntv), dim3(256), 0, st, djobs, tv);
        }
    }
    if (!zg) launch_tv<S, MODE>(djobs, tv, ntv, ntau_max, st);
    LAUNCHCHK();
    launch_ret<S, MODE, true>(ret_jobs, count, nmax, st, rs);
    LAUNCHCHK();
    if constexpr (S == 8 && MODE == MODE_DYSON) {
        if (zg) {
            // lesser convolution (tv*vt + les*adv) of every solve as one GEMM
            const int64_t rows = zp->les_rows, kdim = zp->les_k;
            cplx *bmq = bm + zp->bm_off;
            hipLaunchKernelGGL((k_pack_lesp<8>), dim3(nblocks(rows * kdim, 256)), dim3(256), 0, st, djobs, bm, kdim);
            hipLaunchKernelGGL((k_pack_lesb<8>), dim3(nblocks(kdim * 8, 256), count), dim3(256), 0, st, djobs, bmq,
                               kdim * 8);
            cplx *part = bm + zp->part_off;
            hipLaunchKernelGGL((k_zgemm_cols<8, ZG_LES>),
                               dim3(nblocks(8 * count, ZG_BN), nblocks(rows, ZG_BM), zp->les_split), dim3(256), 0, st,
                               (int)rows, (int)kdim, count, (const cplx *)bm, kdim, djobs, (const cplx *)bmq,
                               kdim * 8, 1.0, zp->les_kc, part);
            if (zp->les_split > 1)
                hipLaunchKernelGGL((k_zgemm_reduce<8, ZG_LES>), dim3(nblocks(rows * 8 * count, 256)), dim3(256), 0,
                                   st, (int)rows, count, zp->les_split, (const cplx *)part, djobs, 1.0);
        }
    }
    if (zg) {
    } else if (S == 1 && MODE == MODE_DYSON && each_mm) {
        // size-1 solves with their own Sigma: each its one-column product (k_lesconv_mm_each)
        hipLaunchKernelGGL((k_lesconv_bmat_each<MODE_DYSON>), dim3(nblocks(ntau_max + 2 + nmax, 256), count),
                           dim3(256), 0, st, djobs);
        hipLaunchKernelGGL((k_lesconv_mm_each<MODE_DYSON>), dim3(nblocks(nmax + 1, 16), count), dim3(64 * MMW), 0,
                           st, djobs);
    } else if (S == 1 && MODE != MODE_CONV && mm_ldb > 0)
        launch_lesconv_mm<MODE == MODE_CONV ? MODE_DYSON : MODE>(djobs, count, nmax, ntau_max, bm,
                                                                  mm_ldb, st);
    else
        hipLaunchKernelGGL((k_lesconv<S, MODE>), dim3((unsigned)nles), dim3(256), 0, st, djobs, les);
    LAUNCHCHK();
    launch_les<S, MODE, true>(djobs, count, nmax, st, rs);
    LAUNCHCHK();
    return CNTRX_OK;
}

template <int S, int MODE>
int launch_timestep(const Stage &sg, int count, int nmax, hipStream_t st, int mm_ldb = 0,
                    const ZgemmPlan *zp = nullptr) {
    int ntau_max = 0;
    for (auto &J : *sg.jobs) ntau_max = std::max(ntau_max, J.out.ntau);
    return launch_timestep_raw<S, MODE>(sg.djobs, sg.dlists[0], (int)sg.lists[0].size(), sg.dlists[1],
                                        (int)sg.lists[1].size(), ntau_max, sg.extra, count, nmax, st,
                                        mm_ldb, zp);
}

// The job and tile lists of a single-solve timestep, written on the device from kernel
// arguments (the no-allocation entry point has no host staging): tile i of a list is
// (job 0, first index i * per), as tiles_over lays them out.
__global__ void k_single_job_setup(Job J, Job *job, int2 *tv, int ntv, int tvper, int2 *les, int nles,
                                   int lesper) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i == 0) *job = J;
    if (i < ntv) tv[i] = make_int2(0, i * tvper);
    if (i < nles) les[i] = make_int2(0, i * lesper);
}

// best of three timed runs of fn on st after one untimed run (milliseconds)
template <class F>
int time_best(hipStream_t st, F &&fn, float *best) {
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    *best = 1e30f;
    int rc = CNTRX_OK;
    for (int rep = 0; rep < 4 && rc == CNTRX_OK; rep++) {
        if (hipEventRecord(e0, st) != hipSuccess) rc = fail(CNTRX_EHIP, "autotuner: event record");
        if (rc == CNTRX_OK) rc = fn();
        if (rc == CNTRX_OK && (hipEventRecord(e1, st) != hipSuccess || hipEventSynchronize(e1) != hipSuccess))
            rc = fail(CNTRX_EHIP, "autotuner: event wait");
        float ms = 0.f;
        if (rc == CNTRX_OK && rep > 0 && hipEventElapsedTime(&ms, e0, e1) == hipSuccess) *best = std::min(*best, ms);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    return rc;
}

// the recurrence panel size of size1 S for this history bucket, timed on the live jobs
// (both recurrences of each candidate; they write only timestep n of G, which the call
// then computes again)
template <int S>
int tune_panels(const Job *dj, int cnt, int nmax, hipStream_t st) {
    float ms[2];
    for (int cand = 0; cand < 2; cand++)
        RC(time_best(st, [&]() -> int {
            launch_rec_choice<S, MODE_DYSON>(false, cand, dj, cnt, nmax, st);
            launch_rec_choice<S, MODE_DYSON>(true, cand, dj, cnt, nmax, st);
            LAUNCHCHK();
            return CNTRX_OK;
        }, &ms[cand]));
    const int choice = ms[1] < ms[0] ? 1 : 0;
    panel_slot(S, nmax) = 1 + choice;
    g_tuned.push_back({S, hist_bucket(nmax), 0, 2, choice, {ms[0], ms[1], 0.f}});
    return CNTRX_OK;
}

// the workgroup target of the size1 8 split-K GEMMs for this history bucket, timed on the
// Matsubara-integral product of the live jobs; updates the plan to the choice
int tune_split(ZgemmPlan &zp, const Job *dj, int cnt, cplx *bm, hipStream_t st) {
    const int64_t dim = zp.mi2_dim, N = 8 * (int64_t)cnt;
    cplx *part = bm + zp.part_off;
    hipLaunchKernelGGL((k_pack_mi2<8>), dim3(nblocks(dim * dim, 256)), dim3(256), 0, st, dj, bm);
    LAUNCHCHK();
    float ms[3];
    for (int c = 0; c < 3; c++) {
        int ns, kc;
        zgemm_split(dim, N, dim, ns, kc, kSplitTargets[c]);
        RC(time_best(st, [&]() -> int {
            hipLaunchKernelGGL((k_zgemm_cols<8, ZG_MI2>), dim3(nblocks(N, ZG_BN), nblocks(dim, ZG_BM), ns), dim3(256),
                               0, st, (int)dim, (int)dim, cnt, (const cplx *)bm, dim, dj, (const cplx *)nullptr,
                               (int64_t)0, zp.dtau, kc, part);
            if (ns > 1)
                hipLaunchKernelGGL((k_zgemm_reduce<8, ZG_MI2>), dim3(nblocks(dim * N, 256)), dim3(256), 0, st,
                                   (int)dim, cnt, ns, (const cplx *)part, dj, zp.dtau);
            LAUNCHCHK();
            return CNTRX_OK;
        }, &ms[c]));
    }
    int choice = 0;
    for (int c = 1; c < 3; c++)
        if (ms[c] < ms[choice]) choice = c;
    const int b = hist_bucket((int)(zp.les_rows / 8 - 1));
    g_split_choice[b] = 1 + choice;
    zgemm_split(dim, N, dim, zp.mi2_split, zp.mi2_kc, kSplitTargets[choice]);
    zgemm_split(zp.les_rows, N, zp.les_k, zp.les_split, zp.les_kc, kSplitTargets[choice]);
    zp.tune = false;
    g_tuned.push_back({8, b, 1, 3, choice, {ms[0], ms[1], ms[2]}});
    return CNTRX_OK;
}

// count timesteps of independent solves in one call: the solves are grouped by size1
// (classes 1, 2, 4, 8, each in the caller's order), share one staging upload, and each
// class runs its launch sequence on the stream.  A solve's kernels read only its own data,
// so its result does not depend on the rest of the batch, except where solves of size1 1
// or 8 share one Sigma and its sums become products over the batch (the GEMM forms).
int dyson_timestep_impl(int count, const int *n, const cntrx_herm *G, const double *mu,
                        const cntrx_func *H, const cntrx_herm *Sigma, const double *beta,
                        const double *h, int k, hipStream_t st) {
    RC(ensure_init());
    if (count <= 0) return CNTRX_OK;
    if (k < 1 || k > 5) return fail(CNTRX_EINVAL, "SolveOrder %d out of range 1..5", k);
    if (!G || !n || !mu || !H || !Sigma || !beta || !h) return fail(CNTRX_EINVAL, "NULL argument");
    for (int i = 0; i < count; i++) {
        RC(check_herm(&G[i], "G", n[i]));
        RC(check_herm(&Sigma[i], "Sigma", n[i]));
        RC(same_shape(&G[i], &Sigma[i], "G", "Sigma"));
        if (n[i] <= k) return fail(CNTRX_EINVAL, "dyson_timestep needs n > SolveOrder (n=%d)", n[i]);
        if (!H[i].data || H[i].nt < n[i] || H[i].size1 != G[i].size1)
            return fail(CNTRX_EINVAL, "H: bad function descriptor");
        if (G[i].sig != Sigma[i].sig) return fail(CNTRX_EINVAL, "G and Sigma differ in sig");
    }
    constexpr int NC = 4;
    struct Cls {
        int S = 0, first = 0, cnt = 0, ntau_max = 0, nmax = 0, mm_ldb = 0;
        bool each = false, zg = false;
        ZgemmPlan zp;
        RecSplit rs;
        int64_t extra_off = 0;
    } cl[NC];
    std::vector<Job> jobs;
    jobs.reserve(count);
    Stage sg;
    sg.jobs = &jobs;
    sg.lists.resize(2 * NC);
    for (int c = 0; c < NC; c++) {
        Cls &C = cl[c];
        C.S = 1 << c;
        C.first = (int)jobs.size();
        // size1 2, 4: the solves whose own bucket uses the first panel candidate come first
        for (int part = 0; part < 2; part++) {
            for (int i = 0; i < count; i++) {
                if (G[i].size1 != C.S) continue;
                const int ni = std::max(n[i], k);
                const int choice = (C.S == 2 || C.S == 4) ? (panel_choice(C.S, ni) != 0) : 0;
                if (choice != part) continue;
                jobs.push_back(make_job(&G[i], &Sigma[i], &Sigma[i], &G[i], &G[i], nullptr, &H[i], n[i], k, mu[i],
                                        beta[i], h[i]));
                sg.ntmax.push_back(std::max(G[i].nt, k));
                C.ntau_max = std::max(C.ntau_max, G[i].ntau);
                C.nmax = std::max(C.nmax, ni);
                if (part == 0) {
                    C.rs.cnt0++;
                    C.rs.nmax0 = std::max(C.rs.nmax0, ni);
                } else {
                    C.rs.nmax1 = std::max(C.rs.nmax1, ni);
                }
            }
        }
        C.cnt = (int)jobs.size() - C.first;
    }
    if ((int)jobs.size() != count) return fail(CNTRX_EUNSUPPORTED, "size1 not in {1, 2, 4, 8}");
    // GEMM forms and their workspace (one extra region, a slice per class)
    int64_t extra = 0;
    sg.bm_elems.assign(count, 0);
    for (int c = 0; c < NC; c++) {
        Cls &C = cl[c];
        if (!C.cnt) continue;
        std::vector<Job> cj(jobs.begin() + C.first, jobs.begin() + C.first + C.cnt);
        if (C.S == 1) {
            C.mm_ldb = lesconv_mm_ldb(cj, C.cnt);
            if (C.mm_ldb) {
                C.extra_off = extra;
                extra += (lesconv_mm_elems(cj[0], C.mm_ldb) + 15) / 16 * 16;
            } else {
                C.each = true;
                for (int i = 0; i < C.cnt; i++) sg.bm_elems[C.first + i] = lesconv_mm_elems(cj[i], 16);
            }
        } else if (C.S == 8) {
            C.zg = zgemm_plan(cj, C.cnt, C.zp);
            if (C.zg) {
                C.extra_off = extra;
                extra += (C.zp.elems + 15) / 16 * 16;
            }
        }
        const int NP = 256 / (C.S * C.S);
        for (int i = 0; i < C.cnt; i++) {
            const Job &J = cj[i];
            tiles_over(sg.lists[2 * c], i, J.out.ntau, C.zg ? TV8P : tv_tile_points(C.S, J.out.ntau));
            tiles_over(sg.lists[2 * c + 1], i, std::max(J.n, k), NP);
        }
    }
    sg.extra_elems = extra;
    RC(stage_upload(sg, st));
    for (int c = 0; c < NC; c++) {
        const Cls &C = cl[c];
        if (!C.cnt) continue;
        const Job *dj = sg.djobs + C.first;
        const int2 *tv = sg.dlists[2 * c], *les = sg.dlists[2 * c + 1];
        const int ntv = (int)sg.lists[2 * c].size(), nles = (int)sg.lists[2 * c + 1].size();
        cplx *ex = sg.extra ? sg.extra + C.extra_off : nullptr;
        if (g_autotune && C.S == 2 && panel_slot(2, C.nmax) == 0) RC(tune_panels<2>(dj, C.cnt, C.nmax, st));
        if (g_autotune && C.S == 4 && panel_slot(4, C.nmax) == 0) RC(tune_panels<4>(dj, C.cnt, C.nmax, st));
        ZgemmPlan zp = C.zp;
        if (C.zg && zp.tune) RC(tune_split(zp, dj, C.cnt, ex, st));
        DISPATCH_S(C.S, RC((launch_timestep_raw<S, MODE_DYSON>(dj, tv, ntv, les, nles, C.ntau_max, ex, C.cnt, C.nmax,
                                                               st, C.mm_ldb, C.zg ? &zp : nullptr, C.each,
                                                               (C.S == 2 || C.S == 4) ? &C.rs : nullptr))));
    }
    return CNTRX_OK;
}

int vie2_timestep_impl(int n, const cntrx_herm *G, const cntrx_herm *F, const cntrx_herm *Fcc,
                       const cntrx_herm *Qm, double beta, double h, int k, hipStream_t st) {
    const int S = G->size1, NP = 256 / (S * S);
    std::vector<Job> jobs(2);
    jobs[0] = make_job(G, F, Fcc, G, G, Qm, nullptr, n, k, 0.0, beta, h);
    jobs[1] = make_job(G, Fcc, F, G, G, Qm, nullptr, n, k, 0.0, beta, h);
    Stage sg;
    sg.jobs = &jobs;
    sg.ntmax = {std::max(G->nt, k), std::max(G->nt, k)};
    sg.lists.resize(2);
    tiles_over(sg.lists[0], 0, G->ntau, tv_tile_points(S, G->ntau));
    tiles_over(sg.lists[1], 0, std::max(n, k), NP);
    const int mm_ldb = (S == 1) ? lesconv_mm_ldb(jobs, 1) : 0;
    if (mm_ldb) sg.extra_elems = lesconv_mm_elems(jobs[0], mm_ldb);
    RC(stage_upload(sg, st));
    // the swapped retarded job must use the same scratch as job 0 for nothing shared;
    // it owns its own qq_ret region, which is all k_retrec touches.
    DISPATCH_S(S, RC((launch_timestep<S, MODE_VIE2>(sg, 1, std::max(n, k), st, mm_ldb))));
    return CNTRX_OK;
}

// Start phase (dyson_start / vie2_start): job 0 drives the retarded and mixed start
// blocks, jobs 1..k+1 are the lesser columns n = 0..k.
// lists: 0 = (job 0, m) items for k_start_tv, 1 = les tiles of jobs 1..k+1.
template <int S, int MODE>
int launch_start(const Stage &sg, int k, hipStream_t st) {
    hipLaunchKernelGGL((k_start_ret<S, MODE>), dim3(1), dim3(256), 0, st, sg.djobs);
    LAUNCHCHK();
    hipLaunchKernelGGL((k_start_tv<S, MODE>), dim3((unsigned)sg.lists[0].size()), dim3(256), 0, st,
                       sg.djobs, sg.dlists[0]);
    LAUNCHCHK();
    hipLaunchKernelGGL((k_lesconv<S, MODE>), dim3((unsigned)sg.lists[1].size()), dim3(256), 0, st,
                       sg.djobs, sg.dlists[1]);
    LAUNCHCHK();
    launch_les<S, MODE>(sg.djobs + 1, k + 1, k, st);
    LAUNCHCHK();
    return CNTRX_OK;
}

int start_impl(int mode, const cntrx_herm *G, const cntrx_herm *A, const cntrx_herm *Acc,
               const cntrx_herm *Qm, const cntrx_func *H, double mu, double beta, double h, int k,
               hipStream_t st) {
    const int S = G->size1, NP = 256 / (S * S);
    std::vector<Job> jobs(k + 2);
    Stage sg;
    sg.jobs = &jobs;
    sg.ntmax.assign(k + 2, std::max(G->nt, k));
    sg.lists.resize(2);
    jobs[0] = make_job(G, A, Acc, G, G, Qm, H, 0, k, mu, beta, h);
    for (int m = 0; m <= G->ntau; m++) sg.lists[0].push_back(make_int2(0, m));
    for (int n = 0; n <= k; n++) {
        jobs[n + 1] = make_job(G, A, Acc, G, G, Qm, H, n, k, mu, beta, h);
        tiles_over(sg.lists[1], n + 1, k, NP);
    }
    RC(stage_upload(sg, st));
    if (mode == MODE_DYSON) {
        DISPATCH_S(S, RC((launch_start<S, MODE_DYSON>(sg, k, st))));
    } else {
        DISPATCH_S(S, RC((launch_start<S, MODE_VIE2>(sg, k, st))));
    }
    return CNTRX_OK;
}

template <int S>
int launch_conv(const Stage &sg, hipStream_t st) {
    hipLaunchKernelGGL((k_retconv<S>), dim3((unsigned)sg.lists[2].size()), dim3(256), 0, st, sg.djobs,
                       sg.dlists[2]);
    LAUNCHCHK();
    launch_tv<S, MODE_CONV>(sg.djobs, sg.dlists[0], (int)sg.lists[0].size(), (*sg.jobs)[0].out.ntau, st);
    LAUNCHCHK();
    hipLaunchKernelGGL((k_lesconv<S, MODE_CONV>), dim3((unsigned)sg.lists[1].size()), dim3(256), 0,
                       st, sg.djobs, sg.dlists[1]);
    LAUNCHCHK();
    return CNTRX_OK;
}

// ------------------------------------------------------------------------------------
// Matsubara helpers
// ------------------------------------------------------------------------------------
template <int S>
int conv_mat_launch(cplx *C, const cplx *A, const cplx *B, int ntau, int k, int sig,
                    double beta, hipStream_t st) {
    hipLaunchKernelGGL((k_conv_mat<S>), dim3(nblocks((int64_t)(ntau + 1) * S * S, 4)), dim3(256), 0,
                       st, C, A, B, ntau, k, sig, beta / ntau);
    LAUNCHCHK();
    return CNTRX_OK;
}

// G = vie2 Fourier solution of [1+F]G = Q on the Matsubara axis; scratch holds
// 2(ntau+1)+nf blocks plus 2 small blocks.
template <int S>
int vie2_fourier_launch(cplx *Gm, const cplx *Fm, const cplx *Qm, int ntau, int sig, double beta,
                        int pcf, cplx *scr, hipStream_t st) {
    constexpr int ES = S * S;
    const int nf = (2 * pcf + 1) * (ntau + 1);
    cplx *fdft = scr, *qdft = fdft + (int64_t)(ntau + 1) * ES, *Z = qdft + (int64_t)(ntau + 1) * ES;
    cplx *qmasy = Z + (int64_t)nf * ES;
    hipLaunchKernelGGL(k_qmasy, dim3(1), dim3(64), 0, st, qmasy, Qm, ntau, ES, sig);
    LAUNCHCHK();
    const size_t rb = (size_t)2 * ntau * sizeof(cplx);
    hipLaunchKernelGGL(k_mat_dft, dim3(nblocks((int64_t)(ntau + 1) * ES, 4)), dim3(256), rb, st, fdft,
                       Fm, ntau, ES, sig);
    hipLaunchKernelGGL(k_mat_dft, dim3(nblocks((int64_t)(ntau + 1) * ES, 4)), dim3(256), rb, st, qdft,
                       Qm, ntau, ES, sig);
    LAUNCHCHK();
    hipLaunchKernelGGL((k_freq_vie2<S>), dim3(nblocks(nf, 256)), dim3(256), 0, st, Z, Fm, fdft, Qm,
                       qdft, qmasy, ntau, sig, beta, pcf);
    LAUNCHCHK();
    hipLaunchKernelGGL((k_freq_sum_vie2<S>), dim3(nblocks((int64_t)(ntau + 1) * ES, 4)), dim3(256),
                       rb, st, Gm, Z, qmasy, ntau, sig, beta, pcf);
    LAUNCHCHK();
    return CNTRX_OK;
}
inline int64_t vie2_fourier_scratch(int ntau, int pcf, int ES) {
    return ((int64_t)2 * (ntau + 1) + (int64_t)(2 * pcf + 1) * (ntau + 1) + 2) * ES;
}

// [1+F]G = Q, fix-point refinement (cntr_vie2_impl.hpp:459-482)
template <int S>
int vie2_fixpoint_launch(cplx *Gm, const cplx *Fm, const cplx *Qm, int ntau, int sig, double beta,
                         int k, int nfix, cplx *scr, hipStream_t st) {
    constexpr int ES = S * S;
    const int pcf = 5;
    const int64_t nm = (int64_t)(ntau + 1) * ES;
    cplx *Qn = scr, *dG = Qn + nm, *fs = dG + nm;
    RC(vie2_fourier_launch<S>(Gm, Fm, Qm, ntau, sig, beta, pcf, fs, st));
    for (int it = 0; it < nfix; it++) {
        RC(conv_mat_launch<S>(Qn, Fm, Gm, ntau, k, sig, beta, st));
        hipLaunchKernelGGL(k_add_sub, dim3(nblocks(nm, 256)), dim3(256), 0, st, Qn, Gm, Qm, nm);
        LAUNCHCHK();
        RC(vie2_fourier_launch<S>(dG, Fm, Qn, ntau, sig, beta, pcf, fs, st));
        hipLaunchKernelGGL(k_axpby, dim3(nblocks(nm, 256)), dim3(256), 0, st, Gm, dG, 1.0, -1.0, nm);
        LAUNCHCHK();
    }
    return CNTRX_OK;
}

// [1 + (F* + *Fcc)/2] G = Q by conjugate gradients (cntr_vie2_impl.hpp:485-602): the
// Fourier solution with 3 periods of correction as the start, then at most maxiter
// iterations, none if <R,R> <= tol at the start, stopping once sqrt(<R,R>) < tol.  The
// stopping test needs <R,R> on the host: one 8-byte read (and stream synchronisation) per
// iteration, against two Matsubara convolutions of O(ntau^2 S^3) each.  active < S: the
// arguments are copies of size-`active` matrices embedded in S; the residual and direction
// start at zero outside the active block and stay there, so the inner products, and with
// them every step, are those of the size-`active` problem.
inline int64_t vie2_cg_scratch(int ntau, int ES) {
    return 4 * (int64_t)(ntau + 1) * ES + vie2_fourier_scratch(ntau, 3, ES) + 1;
}
template <int S>
int vie2_cg_launch(cplx *Gm, const cplx *Fm, const cplx *Fccm, const cplx *Qm, int ntau, int sig, double beta,
                   int k, int maxiter, double tol, int active, cplx *scr, hipStream_t st) {
    constexpr int ES = S * S;
    const int64_t nm = (int64_t)(ntau + 1) * ES;
    cplx *C1 = scr, *C2 = C1 + nm, *R = C2 + nm, *P = R + nm, *fs = P + nm;
    double *sc = (double *)(fs + vie2_fourier_scratch(ntau, 3, ES));
    RC(vie2_fourier_launch<S>(Gm, Fm, Qm, ntau, sig, beta, 3, fs, st));
    if (maxiter <= 0) return CNTRX_OK;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    double *host = g_dev[dev].red_host;
    auto read = [&](double *v) -> int {
        HIPCHK(hipMemcpyAsync(host, sc, sizeof(double), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
        *v = host[0];
        return CNTRX_OK;
    };
    RC(conv_mat_launch<S>(C1, Fm, Gm, ntau, k, sig, beta, st));
    RC(conv_mat_launch<S>(C2, Gm, Fccm, ntau, k, sig, beta, st));
    hipLaunchKernelGGL((k_cg_init<S>), dim3(1), dim3(CG_T), 0, st, (const cplx *)Gm, Qm, (const cplx *)C1,
                       (const cplx *)C2, R, P, ntau, k, active, sc);
    LAUNCHCHK();
    double rs;
    RC(read(&rs));
    if (!(rs > tol)) return CNTRX_OK;
    for (int it = 0; it < maxiter; it++) {
        RC(conv_mat_launch<S>(C1, Fm, P, ntau, k, sig, beta, st));
        RC(conv_mat_launch<S>(C2, P, Fccm, ntau, k, sig, beta, st));
        hipLaunchKernelGGL((k_cg_step<S>), dim3(1), dim3(CG_T), 0, st, Gm, R, P, C1, (const cplx *)C2, ntau, k, sc);
        LAUNCHCHK();
        RC(read(&rs));
        if (std::sqrt(rs) < tol) break;
    }
    return CNTRX_OK;
}

// free GF of constant H into G (all components present in G).  mu - H0 goes to the
// device from `stage` (pinned host memory of ES entries that stays untouched until the
// stream has passed this call) or, with stage NULL, from the host stack followed by a
// stream synchronisation.  scr: (nt+3) ES complex entries plus S doubles.
template <int S>
int green_launch(const cntrx_herm *G, double mu, const double *H0_host, double beta, double h,
                 cplx *scr, hipStream_t st, cplx *stage = nullptr) {
    constexpr int ES = S * S;
    cplx Hloc[ES];
    cplx *Hmu = stage ? stage : Hloc;
    for (int i = 0; i < ES; i++) Hmu[i] = mk(-H0_host[2 * i], -H0_host[2 * i + 1]);
    for (int i = 0; i < S; i++) Hmu[i * S + i].x += mu;
    cplx *dH = scr, *V = dH + ES, *U = V + ES;
    double *ev = (double *)(U + (int64_t)(std::max(G->nt, 0) + 1) * ES);
    HIPCHK(hipMemcpyAsync(dH, Hmu, ES * sizeof(cplx), hipMemcpyHostToDevice, st));
    if (!stage) HIPCHK(hipStreamSynchronize(st));  // Hloc lives on the host stack
    hipLaunchKernelGGL((k_eig<S>), dim3(1), dim3(64), 0, st, dH, V, ev);
    LAUNCHCHK();
    if (G->nt >= 0) {
        hipLaunchKernelGGL((k_green_U<S>), dim3(nblocks((int64_t)(G->nt + 1) * ES, 256)), dim3(256), 0,
                           st, U, V, ev, G->nt, h);
        LAUNCHCHK();
    }
    int64_t tot = (int64_t)(G->ntau + 1) * ES;
    if (G->nt >= 0)
        tot += (int64_t)(G->nt + 1) * (G->ntau + 1) * ES + (int64_t)(G->nt + 1) * (G->nt + 2) / 2 * ES;
    hipLaunchKernelGGL((k_green_fill<S>), dim3(nblocks(tot, 256)), dim3(256), 0, st, HermView(*G), V,
                       ev, U, beta, 0);
    LAUNCHCHK();
    return CNTRX_OK;
}

// ---------------------------------------------------------------------------------------
// green_from_H with a time-dependent H (libcntr/cntr/cntr_equilibrium_impl.hpp:754-870,
// 955-1040).  U(-1) = U(0) = 1 and U(t) = C(t) U(t-1) with the commutator-free exponentials
// of Alvermann & Fehske, J. Comput. Phys. 230, 5930 (2011):
//   cf_order 2:  C = exp(-i h H(t - 1/2))
//   cf_order 4:  C = exp(-i h (a1 H1 + a2 H2)) exp(-i h (a2 H1 + a1 H2)),
//                H1,2 = H(t - 1/2 -+ sqrt(3)/6), a1,2 = (3 -+ 2 sqrt(3)) / 12,
// H(x) the order-k polynomial through H(n1-k..n1), n1 = max(t, k) (cntr_utilities_impl.hpp:
// 285-312).  The factors form a sequential chain of nt products of size1 x size1 blocks, so
// they are built on the host; the O(nt^2 + nt ntau) fill of G runs on the device.
namespace gtd {
typedef std::complex<double> hc;
inline void mm(int s, const hc *A, const hc *B, hc *C) {
    for (int i = 0; i < s; i++)
        for (int j = 0; j < s; j++) {
            hc acc = 0.0;
            for (int l = 0; l < s; l++) acc += A[i * s + l] * B[l * s + j];
            C[i * s + j] = acc;
        }
}
// out = exp(-i dt A) by scaling and squaring of the Taylor series (Moler & Van Loan, SIAM
// Rev. 45, 3 (2003), method 3): X = -i dt A / 2^q with ||X||_inf <= 1/4, 20 terms (the
// remainder is below 1e-25 of ||X||), then q squarings
inline void expmi(int s, const hc *A, double dt, hc *out) {
    const int es = s * s;
    std::vector<hc> X(es), T(es), E(es), W(es);
    double nrm = 0.0;
    for (int i = 0; i < s; i++) {
        double r = 0.0;
        for (int j = 0; j < s; j++) r += std::abs(A[i * s + j]);
        nrm = std::max(nrm, r * std::fabs(dt));
    }
    int q = 0;
    while (nrm > 0.25) {
        nrm *= 0.5;
        q++;
    }
    const double sc = dt / std::ldexp(1.0, q);
    for (int l = 0; l < es; l++) X[l] = hc(0.0, -sc) * A[l];
    for (int l = 0; l < es; l++) E[l] = T[l] = (l % (s + 1) == 0) ? 1.0 : 0.0;
    for (int p = 1; p <= 20; p++) {
        mm(s, T.data(), X.data(), W.data());
        for (int l = 0; l < es; l++) {
            T[l] = W[l] / double(p);
            E[l] += T[l];
        }
    }
    for (int i = 0; i < q; i++) {
        mm(s, E.data(), E.data(), W.data());
        E.swap(W);
    }
    std::copy(E.begin(), E.end(), out);
}
// the order-k polynomial through H(n1-k..n1) at x; H(t) at H + (t+1) es
inline void interp(const hc *H, int s, int n1, int k, double x, hc *out) {
    const WeightTable &T = host_weights();
    const int es = s * s;
    std::fill(out, out + es, hc(0.0));
    const double t0 = n1 - k;
    for (int l = 0; l <= k; l++) {
        double w = T.P[k][l], t1 = 1.0;
        for (int p = 1; p <= k; p++) {
            t1 *= x - t0;
            w += t1 * T.P[k][p * (k + 1) + l];
        }
        const hc *h = H + (int64_t)(n1 - k + l + 1) * es;
        for (int e = 0; e < es; e++) out[e] += w * h[e];
    }
}
// H(n+1) from H(n-k..n) (cntr_utilities_impl.hpp:187-231)
inline void extrap(hc *H, int s, int n, int k) {
    const WeightTable &T = host_weights();
    const int es = s * s;
    hc *dst = H + (int64_t)(n + 2) * es;
    if (k == 0) {
        std::copy(H + (int64_t)(n + 1) * es, H + (int64_t)(n + 2) * es, dst);
        return;
    }
    std::vector<hc> acc(es, hc(0.0));
    for (int l = 0; l <= k; l++) {
        double w = 0.0;
        for (int a = 0; a <= k; a++) w += (1 - 2 * (a % 2)) * T.P[k][a * (k + 1) + l];
        const hc *h = H + (int64_t)(n - l + 1) * es;
        for (int e = 0; e < es; e++) acc[e] += w * h[e];
    }
    std::copy(acc.begin(), acc.end(), dst);
}
// U(0..tlast) into U (tlast+1 blocks), before the chemical-potential phases
inline void propagators(hc *H, int s, double h, int k, int order, bool fixHam, int tlast, hc *U) {
    const int es = s * s;
    std::vector<hc> h1(es), h2(es), arg(es), e1(es), e2(es), tmp(es);
    const double a1 = (3.0 - 2.0 * std::sqrt(3.0)) / 12.0, a2 = (3.0 + 2.0 * std::sqrt(3.0)) / 12.0;
    for (int e = 0; e < es; e++) U[e] = (e % (s + 1) == 0) ? 1.0 : 0.0;
    for (int t = 1; t <= tlast; t++) {
        const int n1 = t <= k ? k : t;
        if (!fixHam && t > k) extrap(H, s, t - 1, k);
        hc *Ut = U + (int64_t)t * es;
        const hc *Up = U + (int64_t)(t - 1) * es;
        if (order == 2) {
            interp(H, s, n1, k, t - 0.5, h1.data());
            expmi(s, h1.data(), h, e1.data());
            mm(s, e1.data(), Up, Ut);
        } else {
            interp(H, s, n1, k, t - 0.5 - std::sqrt(3.0) / 6.0, h1.data());
            interp(H, s, n1, k, t - 0.5 + std::sqrt(3.0) / 6.0, h2.data());
            for (int e = 0; e < es; e++) arg[e] = a1 * h1[e] + a2 * h2[e];
            expmi(s, arg.data(), h, e1.data());
            for (int e = 0; e < es; e++) arg[e] = a2 * h1[e] + a1 * h2[e];
            expmi(s, arg.data(), h, e2.data());
            mm(s, e1.data(), e2.data(), tmp.data());
            mm(s, tmp.data(), Up, Ut);
        }
    }
}
}  // namespace gtd

// tstp < -1: every component; tstp = -1: mat; tstp >= 0: timestep tstp from U(0..tstp)
// (already phased).  scr: (tlast+3) ES complex entries plus S doubles.
template <int S>
int green_td_launch(const cntrx_herm *G, const cplx *Hmu, const cplx *U, int tlast, int tstp, double beta,
                    cplx *scr, hipStream_t st) {
    constexpr int ES = S * S;
    cplx *dH = scr, *V = dH + ES, *dU = V + ES;
    double *ev = (double *)(dU + (int64_t)(tlast + 1) * ES);
    HIPCHK(hipMemcpyAsync(dH, Hmu, ES * sizeof(cplx), hipMemcpyHostToDevice, st));
    if (tlast >= 0)
        HIPCHK(hipMemcpyAsync(dU, U, (size_t)(tlast + 1) * ES * sizeof(cplx), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));  // Hmu and U live in host vectors of the caller
    hipLaunchKernelGGL((k_eig<S>), dim3(1), dim3(64), 0, st, dH, V, ev);
    LAUNCHCHK();
    if (tstp >= 0) {
        const int64_t tot = (int64_t)(G->ntau + 1 + 2 * (tstp + 1)) * ES;
        hipLaunchKernelGGL((k_green_fill_tstp<S>), dim3(nblocks(tot, 256)), dim3(256), 0, st, HermView(*G),
                           tstp, V, ev, dU, beta);
    } else {
        int64_t tot = (int64_t)(G->ntau + 1) * ES;  // tstp = -1: the mat entries alone
        if (tstp < -1 && G->nt >= 0)
            tot += (int64_t)(G->nt + 1) * (G->ntau + 1) * ES + (int64_t)(G->nt + 1) * (G->nt + 2) / 2 * ES;
        hipLaunchKernelGGL((k_green_fill<S>), dim3(nblocks(tot, 256)), dim3(256), 0, st, HermView(*G), V, ev,
                           dU, beta, tstp == -1 ? 1 : 0);
    }
    LAUNCHCHK();
    return CNTRX_OK;
}

int extra_scratch(int64_t elems, cplx **out, hipStream_t st) {
    std::vector<Job> none;
    Stage sg;
    sg.jobs = &none;
    sg.extra_elems = elems;
    (void)st;
    void *base;
    RC(workspace(Carve::round(sizeof(Job)) + Carve::round(elems * sizeof(cplx)) + 1024, &base));
    Carve cv(base);
    cv.take<Job>(1);
    *out = cv.take<cplx>(elems);
    return CNTRX_OK;
}

// dyson_mat body (cntr_dyson_impl.hpp:1393-1418) with H(-1) given on the host (es
// complex, interleaved).  `stage`: pinned host memory of 2 es entries that stays
// untouched until the stream has passed the call, for the small uploads; NULL uploads
// from pageable memory and synchronises the stream.  No force_hermitian here.  active: the
// caller's size1 when G and Sigma are embedded copies (0 otherwise).
int dyson_mat_impl(const cntrx_herm *G, double mu, const double *Hm1, const cntrx_herm *Sigma,
                   double beta, int SolveOrder, int method, hipStream_t st, cplx *stage, int active = 0) {
    RC(check_herm(G, "G", -1, false));
    RC(check_herm(Sigma, "Sigma", -1, false));
    RC(same_shape(G, Sigma, "G", "Sigma"));
    if (method != CNTRX_MAT_FOURIER && method != CNTRX_MAT_FIXPOINT && method != CNTRX_MAT_CG)
        return fail(CNTRX_EINVAL, "dyson_mat: unknown method %d", method);
    if (SolveOrder < 0 || SolveOrder > 5) return fail(CNTRX_EINVAL, "SolveOrder %d out of range 0..5", SolveOrder);
    if (G->ntau % 2) return fail(CNTRX_EINVAL, "ntau must be even");
    if (G->ntau > 4096) return fail(CNTRX_EUNSUPPORTED, "Matsubara solve: ntau > 4096");
    const int S = G->size1, ES = S * S, ntau = G->ntau, sig = G->sig;
    const int64_t nm = (int64_t)(ntau + 1) * ES;
    if (method == CNTRX_MAT_FOURIER) {
        const int pcf = 10;
        int64_t nf = (int64_t)ntau * (2 * pcf + 1);
        cplx *scr;
        RC(extra_scratch(nm + nf * ES + 2 * ES, &scr, st));
        cplx *sdft = scr, *Z = sdft + nm, *hj = Z + nf * ES, *one = hj + ES;
        // hj = H(-1) - mu 1 ; coefficient of the first-order tail = identity
        std::vector<cplx> hloc(stage ? 0 : 2 * ES);
        cplx *hh = stage ? stage : hloc.data();
        for (int i = 0; i < ES; i++) {
            hh[i] = mk(Hm1[2 * i] - ((i / S == i % S) ? mu : 0.0), Hm1[2 * i + 1]);
            hh[ES + i] = mk((i / S == i % S) ? 1.0 : 0.0);
        }
        HIPCHK(hipMemcpyAsync(hj, hh, sizeof(cplx) * 2 * ES, hipMemcpyHostToDevice, st));
        if (!stage) HIPCHK(hipStreamSynchronize(st));
        const size_t rb = (size_t)2 * ntau * sizeof(cplx);
        hipLaunchKernelGGL(k_mat_dft, dim3(nblocks(nm, 4)), dim3(256), rb, st, sdft,
                           (const cplx *)Sigma->mat, ntau, ES, sig);
        LAUNCHCHK();
        DISPATCH_S(S, {
            hipLaunchKernelGGL((k_freq_dyson<S>), dim3(nblocks(nf, 256)), dim3(256), 0, st, Z,
                               (const cplx *)Sigma->mat, sdft, hj, ntau, sig, beta, pcf);
            LAUNCHCHK();
            hipLaunchKernelGGL((k_freq_sum_dyson<S>), dim3(nblocks(nm, 4)), dim3(256), rb, st,
                               (cplx *)G->mat, Z, one, ntau, sig, beta, pcf);
            LAUNCHCHK();
        });
    } else {
        // fixpoint (cntr_dyson_impl.hpp:451-468) or conjugate gradients (:522-540):
        // [1 + F] G = G0 with F = -G0*Sigma, and for the latter Fcc = -Sigma*G0
        const bool cg = method == CNTRX_MAT_CG;
        const int64_t solver = cg ? vie2_cg_scratch(ntau, ES) : 2 * nm + vie2_fourier_scratch(ntau, 5, ES);
        cplx *scr;
        RC(extra_scratch(3 * nm + solver + 4 * ES + (int64_t)S, &scr, st));
        cplx *G0m = scr, *F = G0m + nm, *Fcc = F + nm, *fix = Fcc + nm;
        // free Matsubara GF from H(-1)
        cntrx_herm g0 = {-1, ntau, S, sig, nullptr, nullptr, nullptr, (double *)G0m};
        cplx *gscr = fix + solver;
        DISPATCH_S(S, {
            RC(green_launch<S>(&g0, mu, Hm1, beta, 1.0, gscr, st, stage));
            RC(conv_mat_launch<S>(F, G0m, (const cplx *)Sigma->mat, ntau, SolveOrder, sig, beta, st));
            hipLaunchKernelGGL(k_axpby, dim3(nblocks(nm, 256)), dim3(256), 0, st, F, F, -1.0, 0.0, nm);
            LAUNCHCHK();
            if (cg) {
                RC(conv_mat_launch<S>(Fcc, (const cplx *)Sigma->mat, G0m, ntau, SolveOrder, sig, beta, st));
                hipLaunchKernelGGL(k_axpby, dim3(nblocks(nm, 256)), dim3(256), 0, st, Fcc, Fcc, -1.0, 0.0, nm);
                LAUNCHCHK();
                RC(vie2_cg_launch<S>((cplx *)G->mat, F, Fcc, G0m, ntau, sig, beta, SolveOrder, 40, 1e-12,
                                     active ? active : S, fix, st));
            } else {
                RC(vie2_fixpoint_launch<S>((cplx *)G->mat, F, G0m, ntau, sig, beta, SolveOrder, 6, fix, st));
            }
        });
    }
    return CNTRX_OK;
}

// vie2_mat body (cntr_vie2_impl.hpp:1230-1250).  active: the caller's size1 when G, F, Fcc
// and Q are embedded copies (0 otherwise), for the conjugate-gradient inner products.
int vie2_mat_impl(const cntrx_herm *G, const cntrx_herm *F, const cntrx_herm *Fcc, const cntrx_herm *Q,
                  double beta, int SolveOrder, int method, hipStream_t st, int active) {
    RC(check_herm(G, "G", -1, false));
    RC(check_herm(F, "F", -1, false));
    RC(check_herm(Fcc, "Fcc", -1, false));
    RC(check_herm(Q, "Q", -1, false));
    RC(same_shape(G, F, "G", "F"));
    RC(same_shape(G, Fcc, "G", "Fcc"));
    RC(same_shape(G, Q, "G", "Q"));
    if (method != CNTRX_MAT_FOURIER && method != CNTRX_MAT_FIXPOINT && method != CNTRX_MAT_CG)
        return fail(CNTRX_EINVAL, "vie2_mat: unknown method %d", method);
    if (SolveOrder < 0 || SolveOrder > 5) return fail(CNTRX_EINVAL, "SolveOrder %d out of range 0..5", SolveOrder);
    if (G->ntau % 2) return fail(CNTRX_EINVAL, "ntau must be even");
    if (G->ntau > 4096) return fail(CNTRX_EUNSUPPORTED, "Matsubara solve: ntau > 4096");
    const int S = G->size1, ES = S * S, ntau = G->ntau;
    if (method == CNTRX_MAT_CG) {
        // the Fourier method uses F alone; conjugate gradients take F and Fcc
        // (cntr_vie2_impl.hpp:1240-1243: at most 40 iterations, tolerance 1e-12)
        cplx *scr;
        RC(extra_scratch(vie2_cg_scratch(ntau, ES), &scr, st));
        DISPATCH_S(S, RC((vie2_cg_launch<S>((cplx *)G->mat, (const cplx *)F->mat, (const cplx *)Fcc->mat,
                                            (const cplx *)Q->mat, ntau, G->sig, beta, SolveOrder, 40, 1e-12,
                                            active ? active : S, scr, st))));
        return CNTRX_OK;
    }
    int pcf = (method == CNTRX_MAT_FOURIER) ? 20 : 5;
    int64_t need = vie2_fourier_scratch(ntau, pcf, ES) + 2 * (int64_t)(ntau + 1) * ES;
    cplx *scr;
    RC(extra_scratch(need, &scr, st));
    if (method == CNTRX_MAT_FOURIER) {
        DISPATCH_S(S, RC((vie2_fourier_launch<S>((cplx *)G->mat, (const cplx *)F->mat,
                                                 (const cplx *)Q->mat, ntau, G->sig, beta, pcf, scr, st))));
    } else {
        DISPATCH_S(S, RC((vie2_fixpoint_launch<S>((cplx *)G->mat, (const cplx *)F->mat,
                                                  (const cplx *)Q->mat, ntau, G->sig, beta,
                                                  SolveOrder, 6, scr, st))));
    }
    return CNTRX_OK;
}

#include "embed.inc"

}  // namespace

// ====================================================================================
// C ABI
// ====================================================================================
extern "C" {

const char *cntrx_last_error(void) { return g_err.c_str(); }

int cntrx_init(int device) {
    if (device >= 0) HIPCHK(hipSetDevice(device));
    return ensure_init();
}

int cntrx_wait_timeouts(long *count) {
    RC(ensure_init());
    if (!count) return fail(CNTRX_EINVAL, "NULL");
    int v = 0;
    HIPCHK(hipMemcpyFromSymbol(&v, HIP_SYMBOL(g_rmw_timeouts), sizeof(int), 0, hipMemcpyDeviceToHost));
    *count = v;
    return CNTRX_OK;
}

int cntrx_integrator_tables(int k, double *greg, double *omega, double *bd, double *P, double *D,
                            double *W, double *rcorr) {
    if (k < 0 || k > 5) return fail(CNTRX_EINVAL, "k out of range");
    const WeightTable &T = host_weights();
    int k1 = k + 1;
    if (greg) memcpy(greg, T.greg[k], sizeof(double) * 4 * k1 * k1);
    if (omega) memcpy(omega, T.omega[k], sizeof(double) * k1);
    if (bd) memcpy(bd, T.bd[k], sizeof(double) * (k1 + 1));
    if (P) memcpy(P, T.P[k], sizeof(double) * k1 * k1);
    if (D) memcpy(D, T.D[k], sizeof(double) * k1 * k1);
    if (W) memcpy(W, T.W[k], sizeof(double) * k1 * k1 * k1);
    if (rcorr && k > 1) memcpy(rcorr, T.rc[k], sizeof(double) * (k - 1) * k1 * k1);
    return CNTRX_OK;
}

int cntrx_dyson_timestep(int n, const cntrx_herm *G, double mu, const cntrx_func *H,
                         const cntrx_herm *Sigma, double beta, double h, int SolveOrder,
                         void *stream) {
    if (G && embed_size(G->size1))
        return embedded_dyson_timestep(1, &n, G, &mu, H, Sigma, &beta, &h, SolveOrder, (hipStream_t)stream);
    return dyson_timestep_impl(1, &n, G, &mu, H, Sigma, &beta, &h, SolveOrder, (hipStream_t)stream);
}

int cntrx_dyson_timestep_noalloc(int n, const cntrx_herm *G, double mu, const cntrx_func *H,
                                 const cntrx_herm *Sigma, double beta, double h, int SolveOrder,
                                 void *stream) {
    // arguments and the scratch size are checked before the device is touched
    const int k = SolveOrder;
    if (k < 1 || k > 5) return fail(CNTRX_EINVAL, "SolveOrder %d out of range 1..5", k);
    if (!G || !H || !Sigma) return fail(CNTRX_EINVAL, "NULL argument");
    RC(check_herm(G, "G", n));
    RC(check_herm(Sigma, "Sigma", n));
    RC(same_shape(G, Sigma, "G", "Sigma"));
    if (n <= k) return fail(CNTRX_EINVAL, "dyson_timestep needs n > SolveOrder (n=%d)", n);
    if (!H->data || H->nt < n || H->size1 != G->size1)
        return fail(CNTRX_EINVAL, "H: bad function descriptor");
    if (G->sig != Sigma->sig) return fail(CNTRX_EINVAL, "G and Sigma differ in sig");
    const int S = G->size1, ES = S * S, NP = 256 / ES, ntau = G->ntau;
    const int tvper = tv_tile_points(S, ntau), ntv = ntau / tvper + 1;
    const int last = std::max(n, k), nles = last / NP + 1;
    const int ntmax = std::max(G->nt, k);
    Job J = make_job(G, Sigma, Sigma, G, G, nullptr, H, n, k, mu, beta, h);
    // the size-1 lesser convolution of one solve runs as the one-column GEMM, as in
    // cntrx_dyson_timestep (lesconv_mm_ldb of a single job with n > k)
    const int mm_ldb = (S == 1) ? 16 : 0;
    size_t off = 0;
    auto take = [&](size_t bytes) {
        const size_t o = off;
        off = Carve::round(off + bytes);
        return o;
    };
    const size_t o_job = take(sizeof(Job)), o_tv = take(ntv * sizeof(int2)),
                 o_les = take(nles * sizeof(int2));
    const size_t o_ws = take(job_workspace_blocks(ntmax) * ES * sizeof(cplx));
    const size_t o_bm = take(mm_ldb ? lesconv_mm_elems(J, mm_ldb) * sizeof(cplx) : 0);
    if (off > NOALLOC_SCRATCH_BYTES)
        return fail(CNTRX_EINVAL,
                    "dyson_timestep_noalloc: nt=%d ntau=%d size1=%d needs %zu bytes of scratch, "
                    "more than its %zu",
                    G->nt, ntau, S, off, NOALLOC_SCRATCH_BYTES);
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    if (dev < 0 || dev >= 64) return fail(CNTRX_EINVAL, "dyson_timestep_noalloc: device %d", dev);
    const DevState &D = g_dev[dev];
    if (!D.ready || !D.scratch)
        return fail(CNTRX_EINVAL, "dyson_timestep_noalloc: call cntrx_init(%d) first", dev);
    char *base = D.scratch;
    cplx *ws = (cplx *)(base + o_ws);
    const int64_t nb = ntmax + 1;
    J.gles = ws;
    J.qqret = ws + nb * ES;
    J.qqles = ws + 2 * nb * ES;
    J.small = ws + 3 * nb * ES;
    Job *djob = (Job *)(base + o_job);
    int2 *dtv = (int2 *)(base + o_tv), *dles = (int2 *)(base + o_les);
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(k_single_job_setup, dim3(nblocks(std::max(ntv, nles), 256)), dim3(256), 0, st, J,
                       djob, dtv, ntv, tvper, dles, nles, NP);
    LAUNCHCHK();
    DISPATCH_S(S, RC((launch_timestep_raw<S, MODE_DYSON>(djob, dtv, ntv, dles, nles, ntau,
                                                          (cplx *)(base + o_bm), 1, last, st, mm_ldb))));
    return CNTRX_OK;
}

int cntrx_dyson_timestep_batched(int count, const int *n, const cntrx_herm *G, const double *mu,
                                 const cntrx_func *H, const cntrx_herm *Sigma, const double *beta,
                                 const double *h, int SolveOrder, void *stream) {
    if (count > 0 && n && mu && H && Sigma && beta && h && any_embedded(G, count))
        return embedded_dyson_timestep(count, n, G, mu, H, Sigma, beta, h, SolveOrder, (hipStream_t)stream);
    return dyson_timestep_impl(count, n, G, mu, H, Sigma, beta, h, SolveOrder, (hipStream_t)stream);
}

int cntrx_autotune(int enable) {
    g_autotune = enable != 0;
    return CNTRX_OK;
}

int cntrx_autotune_decisions(int cap, double *rows8, int *count) {
    if (!count || (cap > 0 && !rows8)) return fail(CNTRX_EINVAL, "autotune_decisions: NULL");
    const int n = std::min(cap, (int)g_tuned.size());
    for (int i = 0; i < n; i++) {
        const TuneRec &r = g_tuned[i];
        double *o = rows8 + 8 * i;
        o[0] = r.S;
        o[1] = r.bucket;
        o[2] = r.what;
        o[3] = r.ncand;
        o[4] = r.choice;
        o[5] = r.ms[0];
        o[6] = r.ms[1];
        o[7] = r.ms[2];
    }
    *count = (int)g_tuned.size();
    return CNTRX_OK;
}

int cntrx_dyson_start(const cntrx_herm *G, double mu, const cntrx_func *H, const cntrx_herm *Sigma,
                      double beta, double h, int SolveOrder, void *stream) {
    RC(ensure_init());
    int k = SolveOrder;
    if (k < 1 || k > 5) return fail(CNTRX_EINVAL, "SolveOrder %d out of range 1..5", k);
    if (G && embed_size(G->size1)) return embedded_dyson_start(G, mu, H, Sigma, beta, h, k, (hipStream_t)stream);
    RC(check_herm(G, "G", k));
    RC(check_herm(Sigma, "Sigma", k));
    RC(same_shape(G, Sigma, "G", "Sigma"));
    if (!H || !H->data || H->nt < k || H->size1 != G->size1) return fail(CNTRX_EINVAL, "bad H");
    return start_impl(MODE_DYSON, G, Sigma, Sigma, nullptr, H, mu, beta, h, k, (hipStream_t)stream);
}

int cntrx_vie2_start(const cntrx_herm *G, const cntrx_herm *F, const cntrx_herm *Fcc,
                     const cntrx_herm *Q, double beta, double h, int SolveOrder, void *stream) {
    RC(ensure_init());
    int k = SolveOrder;
    if (k < 0 || k > 5) return fail(CNTRX_EINVAL, "SolveOrder %d out of range", k);
    if (G && embed_size(G->size1)) return embedded_vie2(-2, G, F, Fcc, Q, beta, h, k, 0, (hipStream_t)stream);
    RC(check_herm(G, "G", k));
    RC(check_herm(F, "F", k));
    RC(check_herm(Fcc, "Fcc", k));
    RC(check_herm(Q, "Q", k));
    RC(same_shape(G, F, "G", "F"));
    RC(same_shape(G, Q, "G", "Q"));
    return start_impl(MODE_VIE2, G, F, Fcc, Q, nullptr, 0.0, beta, h, k, (hipStream_t)stream);
}

int cntrx_vie2_mat(const cntrx_herm *G, const cntrx_herm *F, const cntrx_herm *Fcc,
                   const cntrx_herm *Q, double beta, int SolveOrder, int method, void *stream) {
    RC(ensure_init());
    hipStream_t st = (hipStream_t)stream;
    if (G && embed_size(G->size1)) return embedded_vie2(-1, G, F, Fcc, Q, beta, 0.0, SolveOrder, method, st);
    return vie2_mat_impl(G, F, Fcc, Q, beta, SolveOrder, method, st, 0);
}

int cntrx_vie2_timestep(int n, const cntrx_herm *G, const cntrx_herm *F, const cntrx_herm *Fcc,
                        const cntrx_herm *Q, double beta, double h, int SolveOrder, int method,
                        void *stream) {
    RC(ensure_init());
    if (n == -1) return cntrx_vie2_mat(G, F, Fcc, Q, beta, SolveOrder, method, stream);
    if (n < 0) return fail(CNTRX_EINVAL, "n=%d", n);
    if (n <= SolveOrder) return cntrx_vie2_start(G, F, Fcc, Q, beta, h, n, stream);
    int k = SolveOrder;
    if (k < 1 || k > 5) return fail(CNTRX_EINVAL, "SolveOrder %d out of range 1..5", k);
    if (G && embed_size(G->size1)) return embedded_vie2(n, G, F, Fcc, Q, beta, h, k, method, (hipStream_t)stream);
    RC(check_herm(G, "G", n));
    RC(check_herm(F, "F", n));
    RC(check_herm(Fcc, "Fcc", n));
    RC(check_herm(Q, "Q", n));
    RC(same_shape(G, F, "G", "F"));
    RC(same_shape(G, Q, "G", "Q"));
    return vie2_timestep_impl(n, G, F, Fcc, Q, beta, h, k, (hipStream_t)stream);
}

int cntrx_dyson_mat(const cntrx_herm *G, double mu, const cntrx_func *H, const cntrx_herm *Sigma,
                    double beta, int SolveOrder, int method, int force_hermitian, void *stream) {
    RC(ensure_init());
    hipStream_t st = (hipStream_t)stream;
    if (G && embed_size(G->size1))
        return embedded_dyson_mat(G, mu, H, Sigma, beta, SolveOrder, method, force_hermitian, st);
    RC(check_herm(G, "G", -1, false));
    if (!H || !H->data || !G || H->size1 != G->size1) return fail(CNTRX_EINVAL, "bad H");
    // H(-1), the equilibrium Hamiltonian, from device memory
    std::vector<double> Hh(2 * G->size1 * G->size1);
    HIPCHK(hipMemcpyAsync(Hh.data(), H->data, sizeof(double) * Hh.size(), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    RC(dyson_mat_impl(G, mu, Hh.data(), Sigma, beta, SolveOrder, method, st, nullptr));
    if (force_hermitian) return cntrx_force_matsubara_hermitian(G, stream);
    return CNTRX_OK;
}

int cntrx_force_matsubara_hermitian(const cntrx_herm *G, void *stream) {
    RC(ensure_init());
    if (G && embed_size(G->size1)) return embedded_force_hermitian(G, (hipStream_t)stream);
    RC(check_herm(G, "G", -1, false));
    const int S = G->size1;
    DISPATCH_S(S, {
        hipLaunchKernelGGL((k_force_herm<S>), dim3(nblocks((int64_t)(G->ntau + 1) * S * S, 256)),
                           dim3(256), 0, (hipStream_t)stream, (cplx *)G->mat, G->ntau);
        LAUNCHCHK();
    });
    return CNTRX_OK;
}

int cntrx_convolution_timestep(int n, const cntrx_herm *C, const cntrx_herm *A,
                               const cntrx_herm *Acc, const cntrx_herm *B, const cntrx_herm *Bcc,
                               double beta, double h, int SolveOrder, void *stream) {
    RC(ensure_init());
    hipStream_t st = (hipStream_t)stream;
    int k = SolveOrder;
    if (k < 1 || k > 5) return fail(CNTRX_EINVAL, "SolveOrder %d out of range 1..5", k);
    if (n < -1) return fail(CNTRX_EINVAL, "convolution_timestep: n %d < -1", n);
    if (C && embed_size(C->size1)) return embedded_convolution(n, C, A, Acc, B, Bcc, beta, h, k, st);
    const cntrx_herm *all[5] = {C, A, Acc, B, Bcc};
    const char *nm5[5] = {"C", "A", "Acc", "B", "Bcc"};
    for (int i = 0; i < 5; i++) {
        RC(check_herm(all[i], nm5[i], n == -1 ? -1 : (i == 0 ? n : std::max(n, k)), n != -1));
        RC(same_shape(C, all[i], "C", nm5[i]));
    }
    for (int i = 1; i < 5; i++)
        if (all[i]->mat == C->mat) return fail(CNTRX_EINVAL, "C must not alias %s", nm5[i]);
    const int S = C->size1, NP = 256 / (S * S);
    if (n == -1) {
        DISPATCH_S(S, RC((conv_mat_launch<S>((cplx *)C->mat, (const cplx *)A->mat,
                                             (const cplx *)B->mat, C->ntau, k, A->sig, beta, st))));
        return CNTRX_OK;
    }
    std::vector<Job> jobs(1);
    jobs[0] = make_job(C, A, Acc, B, Bcc, nullptr, nullptr, n, k, 0.0, beta, h);
    Stage sg;
    sg.jobs = &jobs;
    sg.ntmax = {std::max(C->nt, k)};
    sg.lists.resize(3);
    tiles_over(sg.lists[0], 0, C->ntau, tv_tile_points(S, C->ntau));
    tiles_over(sg.lists[1], 0, n, NP);
    tiles_over(sg.lists[2], 0, n, NP);
    RC(stage_upload(sg, st));
    DISPATCH_S(S, RC(launch_conv<S>(sg, st)));
    return CNTRX_OK;
}

}  // extern "C"

namespace {
// the bubbles take size1 at run time; element indices must lie inside each matrix
int bubble_check(int tstp, const cntrx_herm *C, int c1, int c2, const cntrx_herm *A, const cntrx_herm *Acc,
                 int a1, int a2, const cntrx_herm *B, const cntrx_herm *Bcc, int b1, int b2) {
    if (tstp < -1) return fail(CNTRX_EINVAL, "bubble: tstp %d < -1", tstp);
    const cntrx_herm *all[5] = {C, A, Acc, B, Bcc};
    for (int i = 0; i < 5; i++) RC(check_herm(all[i], "bubble argument", tstp, tstp >= 0, true));
    for (int i = 1; i < 5; i++)
        if (all[i]->ntau != C->ntau) return fail(CNTRX_EINVAL, "bubble: arguments differ in ntau");
    if (A->size1 != Acc->size1 || B->size1 != Bcc->size1)
        return fail(CNTRX_EINVAL, "bubble: A/Acc or B/Bcc differ in size1");
    auto in = [](int i, const cntrx_herm *X) { return i >= 0 && i < X->size1; };
    if (!in(c1, C) || !in(c2, C) || !in(a1, A) || !in(a2, A) || !in(b1, B) || !in(b2, B))
        return fail(CNTRX_EINVAL, "bubble: element index out of range");
    return CNTRX_OK;
}
}  // namespace

extern "C" {

int cntrx_bubble1(int tstp, const cntrx_herm *C, int c1, int c2, const cntrx_herm *A,
                  const cntrx_herm *Acc, int a1, int a2, const cntrx_herm *B, const cntrx_herm *Bcc,
                  int b1, int b2, void *stream) {
    RC(ensure_init());
    RC(bubble_check(tstp, C, c1, c2, A, Acc, a1, a2, B, Bcc, b1, b2));
    hipLaunchKernelGGL(k_bubble, dim3(nblocks(tstp + 2 + C->ntau, 256)), dim3(256), 0,
                       (hipStream_t)stream, 1, tstp, HermView(*C), c1, c2, HermView(*A), HermView(*Acc),
                       a1, a2, HermView(*B), HermView(*Bcc), b1, b2);
    LAUNCHCHK();
    return CNTRX_OK;
}

int cntrx_bubble2(int tstp, const cntrx_herm *C, int c1, int c2, const cntrx_herm *A,
                  const cntrx_herm *Acc, int a1, int a2, const cntrx_herm *B, const cntrx_herm *Bcc,
                  int b1, int b2, void *stream) {
    RC(ensure_init());
    RC(bubble_check(tstp, C, c1, c2, A, Acc, a1, a2, B, Bcc, b1, b2));
    hipLaunchKernelGGL(k_bubble, dim3(nblocks(tstp + 2 + C->ntau, 256)), dim3(256), 0,
                       (hipStream_t)stream, 2, tstp, HermView(*C), c1, c2, HermView(*A), HermView(*Acc),
                       a1, a2, HermView(*B), HermView(*Bcc), b1, b2);
    LAUNCHCHK();
    return CNTRX_OK;
}

int cntrx_extrapolate_timestep(int n, const cntrx_herm *G, int SolveOrder, void *stream) {
    RC(ensure_init());
    RC(check_herm(G, "G", n + 1, true, true));
    if (n < SolveOrder || SolveOrder < 0 || SolveOrder > 5)
        return fail(CNTRX_EINVAL, "extrapolate_timestep: n out of range");
    int64_t tot = (int64_t)(G->ntau + 1 + 2 * (n + 2)) * G->size1 * G->size1;
    hipLaunchKernelGGL(k_extrap, dim3(nblocks(tot, 256)), dim3(256), 0, (hipStream_t)stream, n,
                       HermView(*G), SolveOrder);
    LAUNCHCHK();
    return CNTRX_OK;
}

int cntrx_distance_norm2(int tstp, const cntrx_herm *g1, const cntrx_herm *g2, double *out,
                         void *stream) {
    RC(ensure_init());
    hipStream_t st = (hipStream_t)stream;
    if (tstp < -1) return fail(CNTRX_EINVAL, "distance_norm2: tstp %d < -1", tstp);
    RC(check_herm(g1, "g1", tstp, tstp >= 0, true));
    RC(check_herm(g2, "g2", tstp, tstp >= 0, true));
    RC(same_shape(g1, g2, "g1", "g2"));
    if (!out) return fail(CNTRX_EINVAL, "out is NULL");
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    DevState &D = g_dev[dev];
    hipLaunchKernelGGL(k_axpby, dim3(1), dim3(256), 0, st, (cplx *)D.red, (const cplx *)D.red, 0.0, 0.0,
                       (int64_t)1);
    int tot = (tstp == -1) ? g1->ntau + 1 : 2 * (tstp + 1) + g1->ntau + 1;
    hipLaunchKernelGGL(k_dist, dim3(nblocks(tot, 256)), dim3(256), 0, st, tstp, HermView(*g1),
                       HermView(*g2), D.red);
    LAUNCHCHK();
    HIPCHK(hipMemcpyAsync(D.red_host, D.red, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    *out = D.red_host[0];
    return CNTRX_OK;
}

int cntrx_green_from_H(const cntrx_herm *G, double mu, const double *H0_host, double beta, double h,
                       void *stream) {
    RC(ensure_init());
    hipStream_t st = (hipStream_t)stream;
    if (!H0_host) return fail(CNTRX_EINVAL, "H0 is NULL");
    if (G && embed_size(G->size1)) return embedded_green_from_H(G, mu, H0_host, beta, h, st);
    RC(check_herm(G, "G", -1, G && G->nt >= 0));
    const int S = G->size1, ES = S * S;
    cplx *scr;
    RC(extra_scratch((int64_t)(std::max(G->nt, 0) + 4) * ES + S, &scr, st));
    DISPATCH_S(S, RC(green_launch<S>(G, mu, H0_host, beta, h, scr, st)));
    return CNTRX_OK;
}

int cntrx_green_from_H_td(const cntrx_herm *G, double mu, const cntrx_func *H, double beta, double h,
                          int SolveOrder, int cf_order, int fixHam, int tstp, void *stream) {
    hipStream_t st = (hipStream_t)stream;
    RC(check_herm(G, "G", -1, G && G->nt >= 0, true));
    if (!supported_size(G->size1) && !embed_size(G->size1))
        return fail(CNTRX_EUNSUPPORTED, "green_from_H: size1=%d not supported (1..8)", G->size1);
    if (!H || !H->data) return fail(CNTRX_EINVAL, "green_from_H: H is NULL");
    if (H->size1 != G->size1) return fail(CNTRX_EINVAL, "green_from_H: H and G sizes differ");
    if (cf_order != 2 && cf_order != 4) return fail(CNTRX_EINVAL, "green_from_H: cf_order must be 2 or 4");
    if (SolveOrder < 0 || SolveOrder > 5) return fail(CNTRX_EINVAL, "green_from_H: SolveOrder out of 0..5");
    if (tstp < -2 || tstp > G->nt) return fail(CNTRX_EINVAL, "green_from_H: tstp out of range");
    if (!(beta > 0.0) || !(h > 0.0)) return fail(CNTRX_EINVAL, "green_from_H: beta and h must be > 0");
    const int tlast = tstp == -2 ? G->nt : tstp;
    if (tlast >= 1 && H->nt < std::max(tlast, SolveOrder))
        return fail(CNTRX_EINVAL, "green_from_H: H needs timesteps up to %d", std::max(tlast, SolveOrder));
    RC(ensure_init());
    const int S = G->size1, ES = S * S;
    typedef std::complex<double> hc;
    hc *Hh = (hc *)H->data;
    std::vector<hc> Hmu(ES), U((size_t)(std::max(tlast, 0) + 1) * ES);
    for (int e = 0; e < ES; e++) Hmu[e] = -Hh[e];
    for (int i = 0; i < S; i++) Hmu[i * S + i] += mu;
    if (tlast >= 0) {
        gtd::propagators(Hh, S, h, SolveOrder, cf_order, fixHam != 0, tlast, U.data());
        if (tstp == -2) {
            for (int n = 0; n <= tlast; n++)
                for (int e = 0; e < ES; e++) U[(size_t)n * ES + e] *= std::polar(1.0, mu * h * n);
        } else {
            // the timestep form (cntr_equilibrium_impl.hpp:995-1000) multiplies U(tstp) by
            // e^{i mu h n} for every n = 0..tstp and leaves U(n < tstp) as they are
            for (int n = 0; n <= tlast; n++)
                for (int e = 0; e < ES; e++) U[(size_t)tlast * ES + e] *= std::polar(1.0, mu * h * n);
        }
    }
    cplx *scr;
    if (embed_size(S)) {
        // padding: Hmu = mu - (mu + 1) = -1 and U = 1 on the padding diagonal
        const int sp = embed_size(S), ESP = sp * sp;
        const size_t nb = (size_t)std::max(tlast, 0) + 1;
        std::vector<hc> Hp(ESP), Up(nb * ESP);
        for (int a = 0; a < sp; a++)
            for (int c = 0; c < sp; c++) {
                const bool in = a < S && c < S;
                Hp[a * sp + c] = in ? Hmu[a * S + c] : hc(a == c ? -1.0 : 0.0);
                for (size_t b = 0; b < nb; b++)
                    Up[b * ESP + a * sp + c] = in ? U[b * ES + a * S + c] : hc(a == c ? 1.0 : 0.0);
            }
        Embed E(st);
        int g;
        RC(E.herm(G, tlast, "G", &g));
        RC(E.setup());
        RC(extra_scratch((int64_t)(std::max(tlast, 0) + 4) * ESP + sp, &scr, st));
        DISPATCH_S(sp, RC(green_td_launch<S>(&E.H[g], (const cplx *)Hp.data(), (const cplx *)Up.data(), tlast,
                                             tstp, beta, scr, st)));
        if (tstp == -1) return E.back(g, 0, -1, true);
        if (tstp >= 0) return E.back(g, tstp, tstp, false);
        return E.back(g, 0, tlast, true);
    }
    RC(extra_scratch((int64_t)(std::max(tlast, 0) + 4) * ES + S, &scr, st));
    DISPATCH_S(S, RC(green_td_launch<S>(G, (const cplx *)Hmu.data(), (const cplx *)U.data(), tlast,
                                        tstp, beta, scr, st)));
    return CNTRX_OK;
}

int cntrx_timestep_lincomb(int tstp, const cntrx_herm *C, int count, const cntrx_herm *X,
                           const double *w, void *stream) {
    RC(ensure_init());
    hipStream_t st = (hipStream_t)stream;
    if (tstp < -1) return fail(CNTRX_EINVAL, "timestep_lincomb: tstp %d < -1", tstp);
    RC(check_herm(C, "C", tstp, tstp >= 0, true));
    if (count <= 0 || !X || !w) return fail(CNTRX_EINVAL, "bad lincomb arguments");
    std::vector<HermView> xv(count);
    for (int i = 0; i < count; i++) {
        RC(check_herm(&X[i], "X", tstp, tstp >= 0, true));
        RC(same_shape(C, &X[i], "C", "X"));
        xv[i] = HermView(X[i]);
    }
    size_t bytes = Carve::round(count * sizeof(HermView)) + Carve::round(count * sizeof(double));
    void *base;
    RC(workspace(bytes + 1024, &base));
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    Pinned &P = g_pin[dev];
    if (P.pending) {
        HIPCHK(hipEventSynchronize(P.ev));
        P.pending = false;
    }
    if (P.bytes < bytes) {
        if (P.buf) HIPCHK(hipHostFree(P.buf));
        P.bytes = std::max(bytes, (size_t)(1 << 16));
        HIPCHK(hipHostMalloc((void **)&P.buf, P.bytes));
        if (!P.ev) HIPCHK(hipEventCreateWithFlags(&P.ev, hipEventDisableTiming));
    }
    size_t woff = Carve::round(count * sizeof(HermView));
    memcpy(P.buf, xv.data(), count * sizeof(HermView));
    memcpy(P.buf + woff, w, count * sizeof(double));
    HIPCHK(hipMemcpyAsync(base, P.buf, bytes, hipMemcpyHostToDevice, st));
    HIPCHK(hipEventRecord(P.ev, st));
    P.pending = true;
    int64_t nent = (tstp == -1) ? C->ntau + 1 : (int64_t)2 * (tstp + 1) + C->ntau + 1;
    hipLaunchKernelGGL(k_lincomb, dim3(nblocks(nent * C->size1 * C->size1, 256)), dim3(256), 0, st, tstp,
                       HermView(*C), (const HermView *)base, (const double *)((char *)base + woff), count);
    LAUNCHCHK();
    return CNTRX_OK;
}

int cntrx_timestep_scale_by_function(int tstp, const cntrx_herm *C, const cntrx_func *F_host,
                                     double weight, int left, void *stream) {
    RC(ensure_init());
    hipStream_t st = (hipStream_t)stream;
    RC(check_herm(C, "C", tstp, tstp >= 0));
    if (C->size1 != 1) return fail(CNTRX_EUNSUPPORTED, "scale_by_function: size1 must be 1");
    if (!F_host || !F_host->data || F_host->nt < tstp) return fail(CNTRX_EINVAL, "bad function");
    cplx *scr;
    RC(extra_scratch(F_host->nt + 2, &scr, st));
    HIPCHK(hipMemcpyAsync(scr, F_host->data, sizeof(cplx) * (F_host->nt + 2), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    hipLaunchKernelGGL(k_scale_fn, dim3(nblocks(2 * (tstp + 1) + C->ntau + 1, 256)), dim3(256), 0, st, tstp,
                       HermView(*C), (const cplx *)scr, weight, left);
    LAUNCHCHK();
    return CNTRX_OK;
}

int cntrx_incr_convolution(int tstp, const unsigned char *mask_host, double alpha_re, double alpha_im,
                           const cntrx_herm *C, const cntrx_herm *A, const cntrx_herm *Acc,
                           const double *f0, const double *ft, const cntrx_herm *B, const cntrx_herm *Bcc,
                           double beta, double h, int SolveOrder, void *stream) {
    RC(ensure_init());
    hipStream_t st = (hipStream_t)stream;
    const int k = SolveOrder;
    if (k < 1 || k > 5) return fail(CNTRX_EINVAL, "SolveOrder %d out of range 1..5", k);
    if (tstp < -1) return fail(CNTRX_EINVAL, "incr_convolution: tstp %d < -1", tstp);
    if (C && embed_size(C->size1))
        return embedded_incr_convolution(tstp, mask_host, alpha_re, alpha_im, C, A, Acc, f0, ft, B, Bcc, beta, h, k,
                                         st);
    const cntrx_herm *all[5] = {C, A, Acc, B, Bcc};
    const char *nm5[5] = {"C", "A", "Acc", "B", "Bcc"};
    for (int i = 0; i < 5; i++) {
        RC(check_herm(all[i], nm5[i], tstp == -1 ? -1 : (i == 0 ? tstp : std::max(tstp, k)), tstp != -1));
        RC(same_shape(C, all[i], "C", nm5[i]));
    }
    for (int i = 1; i < 5; i++)
        if (all[i]->mat == C->mat) return fail(CNTRX_EINVAL, "C must not alias %s", nm5[i]);
    if (C->ntau < k) return fail(CNTRX_EINVAL, "incr_convolution: ntau %d < SolveOrder %d", C->ntau, k);
    const int S = C->size1, ES = S * S, ntau = C->ntau, n1 = std::max(tstp, k);
    // selected outputs: mat (tstp = -1), else ret, tv, les in one list each
    const int nsec = (tstp == -1) ? 1 : 3;
    const int len[3] = {tstp == -1 ? ntau + 1 : tstp + 1, ntau + 1, tstp + 1};
    Stage sg;
    std::vector<Job> nojobs;
    sg.jobs = &nojobs;
    sg.lists.resize(nsec);
    int off = 0;
    for (int s = 0; s < nsec; s++) {
        for (int i = 0; i < len[s]; i++)
            if (!mask_host || mask_host[off + i]) sg.lists[s].push_back(make_int2(s, i));
        off += len[s];
    }
    const bool les = nsec == 3 && !sg.lists[2].empty();
    const int64_t nj = n1 + 1, nm = ntau + 1;
    const int64_t e_ar = (tstp == -1) ? 0 : nj * ES, e_bmat = f0 ? nm * ES : 0;
    const int64_t e_les = les ? (2 * nj + nm) * ES : 0;
    sg.extra_elems = e_ar + e_bmat + e_les + 4;
    RC(stage_upload(sg, st));
    IncrOps o;
    o.C = HermView(*C);
    o.A = HermView(*A);
    o.Acc = HermView(*Acc);
    o.B = HermView(*B);
    o.Bcc = HermView(*Bcc);
    o.f0 = (const cplx *)f0;
    o.ft = (const cplx *)ft;
    o.ar = sg.extra;
    o.bmat = f0 ? sg.extra + e_ar : nullptr;
    o.badv = sg.extra + e_ar + e_bmat;
    o.bles = o.badv + nj * ES;
    o.bvt = o.bles + nj * ES;
    o.tstp = tstp;
    o.n1 = n1;
    o.k = k;
    o.ntau = ntau;
    o.h = h;
    o.beta = beta;
    o.alpha = mk(alpha_re, alpha_im);
    if (tstp == -1) o.ft = nullptr;
    const int NP = 256 / ES;
    DISPATCH_S(S, {
        const int64_t blocks = (tstp == -1 ? nm : 2 * nj + 2 * nm) * ES;
        if (f0 || tstp >= 0) {
            if (tstp == -1) {  // only bmat: skip the ar part of the grid
                IncrOps q = o;
                q.n1 = -1;
                hipLaunchKernelGGL((k_incr_stage<S>), dim3(nblocks(nm * ES, 256)), dim3(256), 0, st, q, 0);
            } else {
                hipLaunchKernelGGL((k_incr_stage<S>), dim3(nblocks(blocks, 256)), dim3(256), 0, st, o,
                                   les ? 1 : 0);
            }
            LAUNCHCHK();
        }
        if (tstp == -1) {
            const int c = (int)sg.lists[0].size();
            o.sel = sg.dlists[0];
            if (c) hipLaunchKernelGGL((k_incr_mat<S>), dim3(nblocks(c, NP)), dim3(256), 0, st, o, c);
            LAUNCHCHK();
            return CNTRX_OK;
        }
        const int cr = (int)sg.lists[0].size(), ct = (int)sg.lists[1].size(), cl = (int)sg.lists[2].size();
        o.sel = sg.dlists[0];
        if (cr) hipLaunchKernelGGL((k_incr_ret<S>), dim3(cr), dim3(256), 0, st, o);
        LAUNCHCHK();
        o.sel = sg.dlists[1];
        if (ct) hipLaunchKernelGGL((k_incr_tv<S>), dim3(nblocks(ct, NP)), dim3(256), 0, st, o, ct);
        LAUNCHCHK();
        o.sel = sg.dlists[2];
        if (cl) hipLaunchKernelGGL((k_incr_les<S>), dim3(cl), dim3(256), 0, st, o);
        LAUNCHCHK();
    });
    return CNTRX_OK;
}

}  // extern "C"

// ====================================================================================
// k-batched lattice driver (host side)
// ====================================================================================
struct cntrx_lattice {
    int nk, ntau, nt, k;
    double U, mu, beta, h;
    cplx *slab = nullptr;
    int64_t stride = 0;  // complex elements between consecutive G_k
    std::vector<cntrx_herm> G;
    cntrx_herm Gloc, Gold, Sig, Pol;
    cplx *Hdata = nullptr;
    std::vector<cntrx_func> H;
    Job *jobs = nullptr;
    int2 *tiles_tv = nullptr, *tiles_les = nullptr;
    int ntv = 0, nles = 0;
    // fused: the tv strips and the retarded rows of an iteration run as one launch
    // (k_tvret1): nretpad retarded-row workgroups, then npairs pairs of tv tiles
    bool fused = false;
    int2 *pairs_tv = nullptr;
    int npairs = 0, nretpad = 0;
    HermView *views = nullptr;  // G_k views (batched extrapolation)
    int *tstp_dev = nullptr;    // current timestep, read by the graph's kernels
    cplx *Bm = nullptr;         // K x ldb operand of the lesser-convolution GEMM
    int ldb = 0;
    LatCtl *ctl = nullptr;      // device-side iteration control
    LatCtl *ctl_host = nullptr; // pinned mirror, read once per batch of iterations
    int last_iters = 4;         // iterations of the previous timestep: the next batch size
    void *dev_block = nullptr;
    hipStream_t st = nullptr;   // private stream (graph capture needs a non-null stream)
    hipStream_t st2 = nullptr;  // second stream: the retarded row beside k_tv in x_iter
    hipEvent_t fork = nullptr, join = nullptr;
    // One gated corrector iteration is six parts: Sigma, the four Dyson kernels and the
    // G_loc/distance kernel.  x_iter replays all of it; with profiling the kernel being
    // timed is launched between two HIP events and the parts before and after it are
    // replayed as graphs (x_pre[q], x_post[q]).  prof = 1 times all four kernels.
    hipGraph_t g_iter = nullptr, g_pre[4] = {}, g_post[4] = {};
    hipGraphExec_t x_iter = nullptr, x_pre[4] = {}, x_post[4] = {};
    int prof = 0;  // 0 off, 1 all four kernels, 2 + q kernel q only, 6 sampled: every
                   // 8th iteration times one kernel, the four in turn
    long rot = 0;  // iterations seen by the sampled mode
    std::vector<hipEvent_t> ev;
    std::vector<int> slot_q;  // kernel timed in each enqueued slot
    struct Sample {
        int kernel, tstp;
        float ms;
    };
    std::vector<Sample> samples;  // every timed launch since the mode was set
    double ms[4] = {0, 0, 0, 0};
    long launches[4] = {0, 0, 0, 0};
    double tv_bytes = 0.0;
    // device-driven run (cntrx_lattice_run): chunk = prologue + B iterations + snapshot.
    // A sampled chunk times kernel q of its first iteration: prologue and parts 0..q,
    // kernel q+1 between two events, then parts q+2..5, B-1 iterations and the snapshot.
    // The kernels are launched one by one: a boundary between two graphs, or between a
    // graph and a kernel, left the GPU idle for 12-14 us, against 0-1 us between kernels.
    // With CNTRX_LATTICE_GRAPHS=1 in the environment at create, a chunk is one graph
    // instead: x_chunk[B], or x_head[q] + kernel + x_tail[q][B] (captured on first use).
    static constexpr int RUN_MAXB = 8, RUN_RING = 8, RUN_DEPTH = 2;
    bool run_graphs = false;
    bool snap_next = false;  // the next k_gloc_dist enqueued ends a chunk: it writes the snapshot
    LatSnap *ring = nullptr;      // pinned host ring read by the host
    LatSnap *ring_dev = nullptr;  // its device address, written by the chunk's last k_gloc_dist
    hipGraph_t g_chunk[RUN_MAXB + 1] = {}, g_head[4] = {}, g_tail[4][RUN_MAXB + 1] = {};
    hipGraphExec_t x_chunk[RUN_MAXB + 1] = {}, x_head[4] = {}, x_tail[4][RUN_MAXB + 1] = {};
    hipEvent_t run_end[RUN_RING] = {};      // after each chunk (no timing)
    hipEvent_t run_ev[2 * RUN_RING] = {};   // around the sampled kernel of each chunk
    long run_rot = 0;                       // chunks enqueued under the sampled mode
};

namespace {

cntrx_herm carve_herm(cplx *&p, int nt, int ntau, int sig) {
    int64_t ntri = (int64_t)(nt + 1) * (nt + 2) / 2, ntv = (int64_t)(nt + 1) * (ntau + 1);
    cntrx_herm d;
    d.nt = nt;
    d.ntau = ntau;
    d.size1 = 1;
    d.sig = sig;
    d.ret = (double *)p;
    p += ntri;
    d.les = (double *)p;
    p += ntri;
    d.tv = (double *)p;
    p += ntv;
    d.mat = (double *)p;
    p += ntau + 1;
    return d;
}

int nslice_blocks(cntrx_lattice *L) { return nblocks(2 * (L->nt + 1) + L->ntau + 1, 256); }

// G_loc(t) = (1/Nk) sum_k G_k(t) with t read from device memory
int lattice_gloc(cntrx_lattice *L, hipStream_t st, const int *gate = nullptr) {
    hipLaunchKernelGGL(k_gloc, dim3(nblocks(2 * (L->nt + 1) + L->ntau + 1, 64)), dim3(256), 0, st,
                       (const int *)L->tstp_dev, HermView(L->Gloc), HermView(L->G[0]), L->stride, L->nk,
                       1.0 / L->nk, gate);
    LAUNCHCHK();
    return CNTRX_OK;
}

// Parts of one gated corrector iteration: 0 Sigma from G_loc (formed by the kernel
// before it); 1 k_tv; 2 k_retrec; 3 lesser convolution (k_lesconv_mm); 4 k_lesrec;
// 5 G_old = G_loc, G_loc = mean over k, distance and iteration control, and Sigma of the
// next iteration when there is one (k_gloc_dist).
int lattice_part(cntrx_lattice *L, int part, hipStream_t st) {
    const int *tp = L->tstp_dev;
    switch (part) {
        case 0:
            // nothing to launch: Sigma and P of an iteration are formed at the end of the
            // kernel that wrote its G_loc (k_prologue, or the previous k_gloc_dist)
            break;
        case 1:
            if (L->fused) {
                const dim3 grid(L->nretpad + L->npairs);
                if (L->nt < 512)
                    hipLaunchKernelGGL(k_tvret1<512>, grid, dim3(64 * PL_NW), 0, st, L->jobs, L->nk, L->nretpad,
                                       L->pairs_tv);
                else
                    hipLaunchKernelGGL(k_tvret1<NQ1>, grid, dim3(64 * PL_NW), 0, st, L->jobs, L->nk, L->nretpad,
                                       L->pairs_tv);
            } else {
                launch_tv<1, MODE_DYSON>(L->jobs, L->tiles_tv, L->ntv, L->ntau, st);
            }
            break;
        case 2:
            if (!L->fused) launch_ret<1, MODE_DYSON, true>(L->jobs, L->nk, L->nt, st);
            break;
        case 3: launch_lesconv_mm<MODE_DYSON>(L->jobs, L->nk, L->nt, L->ntau, L->Bm, L->ldb, st); break;
        case 4: launch_les<1, MODE_DYSON, true>(L->jobs, L->nk, L->nt, st); break;
        default:
            hipLaunchKernelGGL(k_gloc_dist, dim3(nblocks(2 * (L->nt + 1) + L->ntau + 1, 16)), dim3(256),
                               0, st, tp, L->ctl, HermView(L->Gloc), HermView(L->Gold), HermView(L->G[0]),
                               L->stride, L->nk, 1.0 / L->nk, HermView(L->Sig), HermView(L->Pol), -L->U * L->U,
                               L->snap_next ? L->ring_dev : nullptr, (int)cntrx_lattice::RUN_RING);
    }
    LAUNCHCHK();
    return CNTRX_OK;
}

// Parts p0..p1 onto L->st while it is being captured.  The retarded row does not read
// the mixed component (cntr_dyson_impl.hpp:79-223), so when both parts 1 and 2 are
// included, k_retrec forks onto the second stream and runs beside k_tv; the lesser
// kernels join both.
int enqueue_parts(cntrx_lattice *L, int p0, int p1) {
    int rc = CNTRX_OK;
    const bool fork = !L->fused && p0 <= 1 && p1 >= 2;
    for (int p = p0; p <= p1 && rc == CNTRX_OK; p++) {
        if (fork && p == 1) {
            if (hipEventRecord(L->fork, L->st) != hipSuccess ||
                hipStreamWaitEvent(L->st2, L->fork, 0) != hipSuccess)
                rc = fail(CNTRX_EHIP, "graph fork failed");
            if (rc == CNTRX_OK) rc = lattice_part(L, 2, L->st2);
            if (rc == CNTRX_OK && hipEventRecord(L->join, L->st2) != hipSuccess)
                rc = fail(CNTRX_EHIP, "graph join failed");
            if (rc == CNTRX_OK) rc = lattice_part(L, 1, L->st);
            if (rc == CNTRX_OK && hipStreamWaitEvent(L->st, L->join, 0) != hipSuccess)
                rc = fail(CNTRX_EHIP, "graph join failed");
            p = 2;
            continue;
        }
        rc = lattice_part(L, p, L->st);
    }
    return rc;
}

// what body() enqueues on L->st, as one graph
template <class F>
int capture_fn(cntrx_lattice *L, F &&body, hipGraph_t *g, hipGraphExec_t *ex) {
    HIPCHK(hipStreamBeginCapture(L->st, hipStreamCaptureModeThreadLocal));
    const int rc = body();
    hipError_t ec = hipStreamEndCapture(L->st, g);
    if (rc != CNTRX_OK) return rc;
    if (ec != hipSuccess) return fail(CNTRX_EHIP, "graph capture failed: %s", hipGetErrorString(ec));
    HIPCHK(hipGraphInstantiate(ex, *g, nullptr, nullptr, 0));
    return CNTRX_OK;
}

// parts p0..p1 as one graph
int capture(cntrx_lattice *L, int p0, int p1, hipGraph_t *g, hipGraphExec_t *ex) {
    return capture_fn(L, [&] { return enqueue_parts(L, p0, p1); }, g, ex);
}

// start of a timestep (k_prologue): mode 0 begins timestep n (per-timestep path); mode 1
// begins the next timestep if the current one has stopped iterating (device-driven run)
template <int K>
void launch_prologue(cntrx_lattice *L, hipStream_t st, int mode, int n, double tol, int maxiter) {
    hipLaunchKernelGGL((k_prologue<K>), dim3(nblocks(2 * (L->nt + 1) + L->ntau + 1, 8)), dim3(256), 0, st,
                       mode, n, L->jobs, L->nk, L->tstp_dev, L->ctl, tol, maxiter, HermView(L->Gloc),
                       HermView(L->G[0]), L->stride, L->nk, 1.0 / L->nk, HermView(L->Sig), HermView(L->Pol),
                       -L->U * L->U);
}
int enqueue_prologue(cntrx_lattice *L, hipStream_t st, int mode, int n, double tol, int maxiter) {
    switch (L->k) {
        case 1: launch_prologue<1>(L, st, mode, n, tol, maxiter); break;
        case 2: launch_prologue<2>(L, st, mode, n, tol, maxiter); break;
        case 3: launch_prologue<3>(L, st, mode, n, tol, maxiter); break;
        case 4: launch_prologue<4>(L, st, mode, n, tol, maxiter); break;
        default: launch_prologue<5>(L, st, mode, n, tol, maxiter); break;
    }
    LAUNCHCHK();
    return CNTRX_OK;
}

// parts p0..5 of the current iteration, then `more` whole iterations; the chunk's last
// k_gloc_dist also writes the control block into the snapshot ring
int enqueue_to_chunk_end(cntrx_lattice *L, int p0, int more) {
    L->snap_next = (more == 0);
    int rc = enqueue_parts(L, p0, 5);
    for (int i = 0; i < more && rc == CNTRX_OK; i++) {
        L->snap_next = (i == more - 1);
        rc = enqueue_parts(L, 0, 5);
    }
    L->snap_next = false;
    return rc;
}

// the chunk graphs, captured on first use
int chunk_graph(cntrx_lattice *L, int B, hipGraphExec_t *out) {
    if (!L->x_chunk[B])
        RC(capture_fn(L, [&] {
            RC(enqueue_prologue(L, L->st, 1, 0, 0.0, 0));
            return enqueue_to_chunk_end(L, 0, B - 1);
        }, &L->g_chunk[B], &L->x_chunk[B]));
    *out = L->x_chunk[B];
    return CNTRX_OK;
}

int head_graph(cntrx_lattice *L, int q, hipGraphExec_t *out) {
    if (!L->x_head[q])
        RC(capture_fn(L, [&] {
            RC(enqueue_prologue(L, L->st, 1, 0, 0.0, 0));
            return enqueue_parts(L, 0, q);
        }, &L->g_head[q], &L->x_head[q]));
    *out = L->x_head[q];
    return CNTRX_OK;
}

int tail_graph(cntrx_lattice *L, int q, int B, hipGraphExec_t *out) {
    if (!L->x_tail[q][B])
        RC(capture_fn(L, [&] {
            return enqueue_to_chunk_end(L, q + 2, B - 1);
        }, &L->g_tail[q][B], &L->x_tail[q][B]));
    *out = L->x_tail[q][B];
    return CNTRX_OK;
}

// enqueue one gated iteration; with profiling, events ev[5 slot ..] bracket the timed
// kernels: all four (prof 1: ev[0..4] around parts 1..4) or kernel q (ev[0], ev[1])
int lattice_iteration(cntrx_lattice *L, hipStream_t st, int slot) {
    int q = -1;
    if (L->prof >= 2 && L->prof <= 5) q = L->prof - 2;
    if (L->prof == 6) {
        const long r = L->rot++;
        if (r % 8 == 0) q = (int)((r / 8) % 4);
    }
    if (L->prof != 1) L->slot_q[slot] = q;
    if (L->prof == 0 || (L->prof == 6 && q < 0)) {
        HIPCHK(hipGraphLaunch(L->x_iter, st));
        return CNTRX_OK;
    }
    hipEvent_t *e = &L->ev[5 * slot];
    if (L->prof == 1) {
        RC(lattice_part(L, 0, st));
        for (int p = 1; p <= 4; p++) {
            HIPCHK(hipEventRecord(e[p - 1], st));
            RC(lattice_part(L, p, st));
        }
        HIPCHK(hipEventRecord(e[4], st));
        RC(lattice_part(L, 5, st));
        return CNTRX_OK;
    }
    HIPCHK(hipGraphLaunch(L->x_pre[q], st));
    HIPCHK(hipEventRecord(e[0], st));
    RC(lattice_part(L, q + 1, st));
    HIPCHK(hipEventRecord(e[1], st));
    HIPCHK(hipGraphLaunch(L->x_post[q], st));
    return CNTRX_OK;
}

int lattice_sigma_host(cntrx_lattice *L, int t, hipStream_t st) {
    hipLaunchKernelGGL(k_set_n, dim3(1), dim3(256), 0, st, (Job *)nullptr, 0, t, L->tstp_dev);
    hipLaunchKernelGGL(k_sigma_local, dim3(nblocks(t + 2 + L->ntau, 256)), dim3(256), 0, st,
                       (const int *)L->tstp_dev, HermView(L->Sig), HermView(L->Pol), HermView(L->Gloc),
                       -L->U * L->U, 0.0, (const int *)nullptr);
    LAUNCHCHK();
    return CNTRX_OK;
}

}  // namespace

extern "C" {

int cntrx_copy(void *dst, const void *src, size_t bytes, int kind, void *stream) {
    RC(ensure_init());
    hipMemcpyKind kd = kind == 1 ? hipMemcpyHostToDevice
                                 : (kind == 2 ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice);
    if (kind < 1 || kind > 3) return fail(CNTRX_EINVAL, "copy kind %d", kind);
    HIPCHK(hipMemcpyAsync(dst, src, bytes, kd, (hipStream_t)stream));
    HIPCHK(hipStreamSynchronize((hipStream_t)stream));
    return CNTRX_OK;
}

int cntrx_device_alloc(size_t bytes, void **out) {
    RC(ensure_init());
    if (!out) return fail(CNTRX_EINVAL, "device_alloc: NULL");
    *out = nullptr;
    if (bytes == 0) return CNTRX_OK;
    if (hipMalloc(out, bytes) != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(CNTRX_ENOMEM, "device_alloc: %zu bytes", bytes);
    }
    return CNTRX_OK;
}

int cntrx_device_free(void *p) {
    if (!p) return CNTRX_OK;
    HIPCHK(hipFree(p));
    return CNTRX_OK;
}

int cntrx_lattice_create(int nk, const double *epsk, double U, double mu, double beta, int ntau,
                         int nt, double h, int SolveOrder, cntrx_lattice **out) {
    RC(ensure_init());
    if (!out || !epsk || nk < 1 || nt <= SolveOrder || ntau < 2 || SolveOrder < 1 || SolveOrder > 5)
        return fail(CNTRX_EINVAL, "lattice_create: bad arguments");
    cntrx_lattice *L = new cntrx_lattice();
    L->nk = nk;
    L->ntau = ntau;
    L->nt = nt;
    L->k = SolveOrder;
    L->U = U;
    L->mu = mu;
    L->beta = beta;
    L->h = h;
    HIPCHK(hipStreamCreateWithFlags(&L->st, hipStreamNonBlocking));
    HIPCHK(hipStreamCreateWithFlags(&L->st2, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&L->fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&L->join, hipEventDisableTiming));
    int64_t per = (int64_t)(nt + 1) * (nt + 2) + (int64_t)(nt + 1) * (ntau + 1) + ntau + 1;
    L->stride = per;
    int64_t tot = per * (nk + 4) + (int64_t)nk * (nt + 2);
    if (hipMalloc(&L->slab, tot * sizeof(cplx)) != hipSuccess) {
        delete L;
        (void)hipGetLastError();
        return fail(CNTRX_ENOMEM, "lattice: cannot allocate %lld bytes", (long long)(tot * sizeof(cplx)));
    }
    HIPCHK(hipMemset(L->slab, 0, tot * sizeof(cplx)));
    cplx *p = L->slab;
    for (int q = 0; q < nk; q++) L->G.push_back(carve_herm(p, nt, ntau, -1));
    L->Gloc = carve_herm(p, nt, ntau, -1);
    L->Gold = carve_herm(p, nt, ntau, -1);
    L->Sig = carve_herm(p, nt, ntau, -1);
    L->Pol = carve_herm(p, nt, ntau, 1);
    L->Hdata = p;
    std::vector<cplx> hh((size_t)nk * (nt + 2));
    for (int q = 0; q < nk; q++)
        for (int t = 0; t < nt + 2; t++) hh[(size_t)q * (nt + 2) + t] = mk(epsk[q]);
    HIPCHK(hipMemcpy(L->Hdata, hh.data(), hh.size() * sizeof(cplx), hipMemcpyHostToDevice));
    for (int q = 0; q < nk; q++) {
        cntrx_func f;
        f.nt = nt;
        f.size1 = 1;
        f.data = (double *)(L->Hdata + (size_t)q * (nt + 2));
        L->H.push_back(f);
    }
    std::vector<Job> jobs(nk);
    std::vector<int2> ttv, tles;
    for (int q = 0; q < nk; q++) {
        jobs[q] = make_job(&L->G[q], &L->Sig, &L->Sig, &L->G[q], &L->G[q], nullptr, &L->H[q],
                           SolveOrder + 1, SolveOrder, mu, beta, h);
        tiles_over(tles, q, nt, 256);
    }
    ttv = xcd_tiles(nk, ntau, tv_tile_points(1, ntau));
    // tv tile pairs of the fused launch, in the same XCD-aware order (two 64-point tiles each)
    const bool fused = ntau < NT1 && nt < NQ1;
    std::vector<int2> tpair;
    if (fused) tpair = xcd_tiles(nk, ntau, 2 * TV1P, true);
    int64_t scr_per = job_workspace_blocks(nt);
    size_t bytes = Carve::round(nk * sizeof(Job)) + Carve::round(ttv.size() * sizeof(int2)) +
                   Carve::round(std::max<size_t>(tpair.size(), 1) * sizeof(int2)) +
                   Carve::round(tles.size() * sizeof(int2)) + Carve::round(nk * sizeof(HermView)) +
                   4 * 256 + Carve::round(nk * scr_per * sizeof(cplx)) + 4096 +
                   Carve::round((size_t)(ntau + 2 + nt) * ((nk + 15) / 16 * 16) * sizeof(cplx));
    if (hipMalloc(&L->dev_block, bytes) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(L->slab);
        delete L;
        return fail(CNTRX_ENOMEM, "lattice: cannot allocate job block");
    }
    Carve cv(L->dev_block);
    L->jobs = cv.take<Job>(nk);
    L->tiles_tv = cv.take<int2>(ttv.size());
    L->tiles_les = cv.take<int2>(tles.size());
    L->pairs_tv = cv.take<int2>(std::max<size_t>(tpair.size(), 1));
    L->fused = fused;
    L->npairs = (int)tpair.size();
    L->nretpad = (nk + 7) / 8 * 8;
    L->views = cv.take<HermView>(nk);
    L->tstp_dev = cv.take<int>(1);
    L->ctl = cv.take<LatCtl>(1);
    HIPCHK(hipMemset(L->ctl, 0, sizeof(LatCtl)));  // k_prologue's ticket starts at 0
    L->ldb = (nk + 15) / 16 * 16;
    L->Bm = cv.take<cplx>((size_t)(ntau + 2 + nt) * L->ldb);
    cplx *scr = cv.take<cplx>((size_t)nk * scr_per);
    for (int q = 0; q < nk; q++) {
        cplx *sc = scr + (size_t)q * scr_per;
        jobs[q].gles = sc;
        jobs[q].qqret = sc + (nt + 1);
        jobs[q].qqles = sc + 2 * (int64_t)(nt + 1);
        jobs[q].small = sc + 3 * (int64_t)(nt + 1);
        jobs[q].gate = &L->ctl->active;
    }
    L->ntv = (int)ttv.size();
    L->nles = (int)tles.size();
    std::vector<HermView> views(nk);
    for (int q = 0; q < nk; q++) views[q] = HermView(L->G[q]);
    HIPCHK(hipHostMalloc((void **)&L->ctl_host, sizeof(LatCtl)));
    HIPCHK(hipHostMalloc((void **)&L->ring, cntrx_lattice::RUN_RING * sizeof(LatSnap),
                         hipHostMallocMapped | hipHostMallocCoherent));
    std::memset((void *)L->ring, 0, cntrx_lattice::RUN_RING * sizeof(LatSnap));
    HIPCHK(hipHostGetDevicePointer((void **)&L->ring_dev, L->ring, 0));
    // The run's events carry no system-scope fence: the host reads nothing through them but
    // the snapshot ring, which is coherent host memory that k_gloc_dist publishes with its
    // own system-scope fence.  A fenced marker between two chunks writes back and
    // invalidates the L2 of every XCD, and the next chunk's first kernels wait for it.
    for (int i = 0; i < cntrx_lattice::RUN_RING; i++) {
        HIPCHK(hipEventCreateWithFlags(&L->run_end[i], hipEventDisableTiming | hipEventDisableSystemFence));
        HIPCHK(hipEventCreateWithFlags(&L->run_ev[2 * i], hipEventDisableSystemFence));
        HIPCHK(hipEventCreateWithFlags(&L->run_ev[2 * i + 1], hipEventDisableSystemFence));
    }
    HIPCHK(hipMemcpy(L->jobs, jobs.data(), nk * sizeof(Job), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L->tiles_tv, ttv.data(), ttv.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L->tiles_les, tles.data(), tles.size() * sizeof(int2), hipMemcpyHostToDevice));
    if (fused) HIPCHK(hipMemcpy(L->pairs_tv, tpair.data(), tpair.size() * sizeof(int2), hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(L->views, views.data(), nk * sizeof(HermView), hipMemcpyHostToDevice));
    // one gated corrector iteration captured once and replayed for every iteration of
    // every step (the timestep and the gate live in device memory)
    RC(capture(L, 0, 5, &L->g_iter, &L->x_iter));
    for (int q = 0; q < 4; q++) {
        RC(capture(L, 0, q, &L->g_pre[q], &L->x_pre[q]));
        RC(capture(L, q + 2, 5, &L->g_post[q], &L->x_post[q]));
    }
    // the device-driven run's chunk graphs (the sampled mode's are captured when it is set)
    const char *rg = getenv("CNTRX_LATTICE_GRAPHS");
    L->run_graphs = rg && rg[0] == '1';
    for (int b = 1; b <= cntrx_lattice::RUN_MAXB && L->run_graphs; b++) {
        hipGraphExec_t x;
        RC(chunk_graph(L, b, &x));
    }
    *out = L;
    return CNTRX_OK;
}

int cntrx_lattice_profile(cntrx_lattice *L, int enable) {
    if (!L) return fail(CNTRX_EINVAL, "NULL");
    if (enable < 0 || enable > 6) return fail(CNTRX_EINVAL, "lattice_profile: mode %d", enable);
    L->prof = enable;
    L->rot = 0;
    L->run_rot = 0;
    if (enable == 6 && L->run_graphs) {
        // the sampled chunks' graphs, captured here rather than inside a timed run
        HIPCHK(hipStreamSynchronize(L->st));
        for (int q = 0; q < 4; q++) {
            hipGraphExec_t x;
            RC(head_graph(L, q, &x));
            for (int b = 1; b <= cntrx_lattice::RUN_MAXB; b++) RC(tail_graph(L, q, b, &x));
        }
    }
    L->samples.clear();
    for (int i = 0; i < 4; i++) {
        L->ms[i] = 0.0;
        L->launches[i] = 0;
    }
    L->tv_bytes = 0.0;
    return CNTRX_OK;
}

int cntrx_lattice_kernel_stats(cntrx_lattice *L, double *ms4, long *launches4, double *tv_bytes) {
    if (!L || !ms4 || !launches4 || !tv_bytes) return fail(CNTRX_EINVAL, "NULL");
    for (int i = 0; i < 4; i++) {
        ms4[i] = L->ms[i];
        launches4[i] = L->launches[i];
    }
    *tv_bytes = L->tv_bytes;
    return CNTRX_OK;
}

int cntrx_lattice_fused(cntrx_lattice *L, int *fused) {
    if (!L || !fused) return fail(CNTRX_EINVAL, "NULL");
    *fused = L->fused ? 1 : 0;
    return CNTRX_OK;
}

int cntrx_lattice_kernel_samples(cntrx_lattice *L, long cap, int *kernel, int *tstp, float *ms,
                                 long *count) {
    if (!L || !count || (cap > 0 && (!kernel || !tstp || !ms))) return fail(CNTRX_EINVAL, "NULL");
    const long n = (long)L->samples.size();
    for (long i = 0; i < n && i < cap; i++) {
        kernel[i] = L->samples[i].kernel;
        tstp[i] = L->samples[i].tstp;
        ms[i] = L->samples[i].ms;
    }
    *count = n;
    return CNTRX_OK;
}

int cntrx_lattice_destroy(cntrx_lattice *L) {
    if (!L) return CNTRX_OK;
    HIPCHK(hipDeviceSynchronize());
    RC(pin_quiesce());
    for (hipEvent_t e : L->ev) HIPCHK(hipEventDestroy(e));
    if (L->x_iter) HIPCHK(hipGraphExecDestroy(L->x_iter));
    if (L->g_iter) HIPCHK(hipGraphDestroy(L->g_iter));
    for (int q = 0; q < 4; q++) {
        if (L->x_pre[q]) HIPCHK(hipGraphExecDestroy(L->x_pre[q]));
        if (L->g_pre[q]) HIPCHK(hipGraphDestroy(L->g_pre[q]));
        if (L->x_post[q]) HIPCHK(hipGraphExecDestroy(L->x_post[q]));
        if (L->g_post[q]) HIPCHK(hipGraphDestroy(L->g_post[q]));
        if (L->x_head[q]) HIPCHK(hipGraphExecDestroy(L->x_head[q]));
        if (L->g_head[q]) HIPCHK(hipGraphDestroy(L->g_head[q]));
        for (int b = 0; b <= cntrx_lattice::RUN_MAXB; b++) {
            if (L->x_tail[q][b]) HIPCHK(hipGraphExecDestroy(L->x_tail[q][b]));
            if (L->g_tail[q][b]) HIPCHK(hipGraphDestroy(L->g_tail[q][b]));
        }
    }
    for (int b = 0; b <= cntrx_lattice::RUN_MAXB; b++) {
        if (L->x_chunk[b]) HIPCHK(hipGraphExecDestroy(L->x_chunk[b]));
        if (L->g_chunk[b]) HIPCHK(hipGraphDestroy(L->g_chunk[b]));
    }
    for (int i = 0; i < cntrx_lattice::RUN_RING; i++) {
        if (L->run_end[i]) HIPCHK(hipEventDestroy(L->run_end[i]));
        if (L->run_ev[2 * i]) HIPCHK(hipEventDestroy(L->run_ev[2 * i]));
        if (L->run_ev[2 * i + 1]) HIPCHK(hipEventDestroy(L->run_ev[2 * i + 1]));
    }
    if (L->ring) HIPCHK(hipHostFree(L->ring));
    if (L->st) HIPCHK(hipStreamDestroy(L->st));
    if (L->st2) HIPCHK(hipStreamDestroy(L->st2));
    if (L->fork) HIPCHK(hipEventDestroy(L->fork));
    if (L->join) HIPCHK(hipEventDestroy(L->join));
    if (L->ctl_host) HIPCHK(hipHostFree(L->ctl_host));
    if (L->dev_block) HIPCHK(hipFree(L->dev_block));
    if (L->slab) HIPCHK(hipFree(L->slab));
    delete L;
    return CNTRX_OK;
}

int cntrx_lattice_matrix(cntrx_lattice *L, int which, cntrx_herm *out) {
    if (!L || !out) return fail(CNTRX_EINVAL, "NULL");
    if (which >= 0 && which < L->nk) *out = L->G[which];
    else if (which == -1) *out = L->Gloc;
    else if (which == -2) *out = L->Sig;
    else return fail(CNTRX_EINVAL, "lattice_matrix: which=%d", which);
    return CNTRX_OK;
}

int cntrx_lattice_equilibrium(cntrx_lattice *L, void *stream) {
    (void)stream;
    for (int q = 0; q < L->nk; q++)
        RC(cntrx_dyson_mat(&L->G[q], L->mu, &L->H[q], &L->Sig, L->beta, L->k, CNTRX_MAT_FIXPOINT, 1,
                           L->st));
    HIPCHK(hipStreamSynchronize(L->st));
    return CNTRX_OK;
}

int cntrx_lattice_bootstrap(cntrx_lattice *L, double tol, int maxiter, int *iters, void *stream) {
    (void)stream;
    hipStream_t st = L->st;
    int it = 0;
    for (; it < maxiter; it++) {
        for (int t = 0; t <= L->k; t++) RC(lattice_sigma_host(L, t, st));
        for (int q = 0; q < L->nk; q++)
            RC(cntrx_dyson_start(&L->G[q], L->mu, &L->H[q], &L->Sig, L->beta, L->h, L->k, st));
        double err = 0.0;
        for (int t = 0; t <= L->k; t++) {
            hipLaunchKernelGGL(k_set_n, dim3(1), dim3(256), 0, st, (Job *)nullptr, 0, t, L->tstp_dev);
            hipLaunchKernelGGL(k_copy_slice, dim3(nslice_blocks(L)), dim3(256), 0, st,
                               (const int *)L->tstp_dev, HermView(L->Gold), HermView(L->Gloc));
            RC(lattice_gloc(L, st));
            double e;
            RC(cntrx_distance_norm2(t, &L->Gloc, &L->Gold, &e, st));
            err += e;
        }
        if (err < tol && it > 0) {
            it++;
            break;
        }
    }
    if (iters) *iters = it;
    return CNTRX_OK;
}

int cntrx_lattice_step(cntrx_lattice *L, int tstp, int batched, double tol, int maxiter, int *iters,
                       double *err_out, void *stream) {
    (void)stream;
    if (!L) return fail(CNTRX_EINVAL, "NULL");
    hipStream_t st = L->st;
    if (tstp <= L->k || tstp > L->nt) return fail(CNTRX_EINVAL, "lattice_step: tstp %d", tstp);
    if (maxiter < 1) return fail(CNTRX_EINVAL, "lattice_step: maxiter %d", maxiter);
    RC(enqueue_prologue(L, st, 0, tstp, tol, maxiter));
    int done = 0;
    if (batched) {
        // enqueue as many gated iterations as the previous timestep needed, read the
        // control block once, then continue one iteration at a time if necessary
        int batch = std::max(1, std::min(L->last_iters, maxiter));
        for (;;) {
            if ((int)L->slot_q.size() < batch) L->slot_q.resize(batch);
            if (L->prof && (int)L->ev.size() < 5 * batch) {
                size_t old = L->ev.size();
                L->ev.resize(5 * batch);
                for (size_t i = old; i < L->ev.size(); i++) HIPCHK(hipEventCreate(&L->ev[i]));
            }
            for (int i = 0; i < batch; i++) RC(lattice_iteration(L, st, i));
            HIPCHK(hipMemcpyAsync(L->ctl_host, L->ctl, sizeof(LatCtl), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            const int ran = L->ctl_host->iters - done;
            if (L->prof) {
                for (int i = 0; i < ran; i++) {
                    for (int q = 0; q < 4; q++) {
                        if (L->prof != 1 && q != L->slot_q[i]) continue;
                        const int e0 = (L->prof == 1) ? 5 * i + q : 5 * i;
                        float ms = 0.f;
                        HIPCHK(hipEventElapsedTime(&ms, L->ev[e0], L->ev[e0 + 1]));
                        L->ms[q] += ms;
                        L->launches[q]++;
                        L->samples.push_back({q, tstp, ms});
                        if (q == 0) {
                            double row = 16.0 * (L->ntau + 1);
                            L->tv_bytes += L->nk * (tstp + 2) * row + row + 16.0 * (tstp + 1);
                        }
                    }
                }
                // iterations enqueued beyond convergence never ran: undo their rotation
                if (L->prof == 6) L->rot -= batch - ran;
            }
            done = L->ctl_host->iters;
            if (!L->ctl_host->active) break;
            batch = 1;
        }
        L->last_iters = done;
    } else {
        // plain: Sigma, then cntrx_dyson_timestep once per k, then the same G_loc /
        // distance / control kernel; the host reads the control block every iteration
        for (;;) {
            RC(lattice_sigma_host(L, tstp, st));
            for (int q = 0; q < L->nk; q++)
                RC(cntrx_dyson_timestep(tstp, &L->G[q], L->mu, &L->H[q], &L->Sig, L->beta, L->h, L->k, st));
            RC(lattice_part(L, 5, st));
            HIPCHK(hipMemcpyAsync(L->ctl_host, L->ctl, sizeof(LatCtl), hipMemcpyDeviceToHost, st));
            HIPCHK(hipStreamSynchronize(st));
            done = L->ctl_host->iters;
            if (!L->ctl_host->active) break;
        }
    }
    if (iters) *iters = done;
    if (err_out) *err_out = L->ctl_host->last;
    return CNTRX_OK;
}

int cntrx_lattice_run(cntrx_lattice *L, int t0, int t1, double tol, int maxiter, int *iters,
                      double *errs, void *stream) {
    (void)stream;
    if (!L || !iters || !errs) return fail(CNTRX_EINVAL, "NULL");
    if (t0 <= L->k || t1 > L->nt || t1 < t0) return fail(CNTRX_EINVAL, "lattice_run: timesteps %d..%d", t0, t1);
    if (maxiter < 1) return fail(CNTRX_EINVAL, "lattice_run: maxiter %d", maxiter);
    if (L->prof != 0 && L->prof != 6)
        return fail(CNTRX_EINVAL, "lattice_run: profile mode %d (run supports 0 and 6)", L->prof);
    using LC = cntrx_lattice;
    hipStream_t st = L->st;
    for (int t = t0; t <= t1; t++) {
        iters[t - t0] = 0;
        errs[t - t0] = 0.0;
    }
    // every graph this run can use (normally captured already by create and profile),
    // captured while the stream is idle: the lattice's entry points return with it drained
    const int bmax = std::min(maxiter, (int)LC::RUN_MAXB);
    for (int b = 1; b <= bmax && L->run_graphs; b++) {
        hipGraphExec_t x;
        RC(chunk_graph(L, b, &x));
        for (int q = 0; q < 4 && L->prof == 6; q++) {
            RC(head_graph(L, q, &x));
            RC(tail_graph(L, q, b, &x));
        }
    }
    hipLaunchKernelGGL(k_run_init, dim3(1), dim3(64), 0, st, L->ctl, t0 - 1, t1, tol, maxiter);
    LAUNCHCHK();
    // a timestep takes at most ceil(maxiter / B) + 1 chunks; bound the loop by that
    const long max_chunks = (long)(t1 - t0 + 1) * (maxiter + 1) + LC::RUN_DEPTH + 2;
    LatSnap prev;
    prev.cur = t0 - 1;
    prev.active = 0;
    prev.iters = 0;
    prev.seq = 0;
    prev.last = 0.0;
    int B = std::max(1, std::min(std::min(L->last_iters, maxiter), (int)LC::RUN_MAXB));
    int sq[LC::RUN_RING];
    unsigned int enq = 0, fin = 0;
    bool over = false;
    while (true) {
        while (!over && enq - fin < (unsigned int)LC::RUN_DEPTH) {
            if ((long)enq >= max_chunks) return fail(CNTRX_EHIP, "lattice_run: no progress after %u chunks", enq);
            const int slot = (int)(enq % LC::RUN_RING);
            int q = -1;
            if (L->prof == 6 && (L->run_rot++ % 4) == 0) q = (int)((L->run_rot / 4) % 4);
            if (q < 0 && L->run_graphs) {
                hipGraphExec_t x;
                RC(chunk_graph(L, B, &x));
                HIPCHK(hipGraphLaunch(x, st));
            } else if (q < 0) {
                RC(enqueue_prologue(L, st, 1, 0, 0.0, 0));
                RC(enqueue_to_chunk_end(L, 0, B - 1));
            } else {
                hipGraphExec_t xh = nullptr, xt = nullptr;
                if (L->run_graphs) {
                    RC(head_graph(L, q, &xh));
                    RC(tail_graph(L, q, B, &xt));
                    HIPCHK(hipGraphLaunch(xh, st));
                } else {
                    RC(enqueue_prologue(L, st, 1, 0, 0.0, 0));
                    RC(enqueue_parts(L, 0, q));
                }
                HIPCHK(hipEventRecord(L->run_ev[2 * slot], st));
                RC(lattice_part(L, q + 1, st));
                HIPCHK(hipEventRecord(L->run_ev[2 * slot + 1], st));
                if (L->run_graphs) {
                    HIPCHK(hipGraphLaunch(xt, st));
                } else {
                    RC(enqueue_to_chunk_end(L, q + 2, B - 1));
                }
            }
            HIPCHK(hipEventRecord(L->run_end[slot], st));
            sq[slot] = q;
            enq++;
        }
        if (fin == enq) break;
        const int slot = (int)(fin % LC::RUN_RING);
        HIPCHK(hipEventSynchronize(L->run_end[slot]));
        LatSnap s;
        std::atomic_thread_fence(std::memory_order_acquire);
        std::memcpy(&s, (const void *)&L->ring[slot], sizeof(LatSnap));
        if (s.seq != fin) return fail(CNTRX_EHIP, "lattice_run: snapshot %u holds chunk %u", fin, s.seq);
        const bool began = s.cur != prev.cur;
        const int ran = began ? s.iters : s.iters - prev.iters;
        if (sq[slot] >= 0 && ran >= 1) {
            // the timed kernel belongs to the chunk's first iteration, which ran
            const int q = sq[slot];
            float ms = 0.f;
            HIPCHK(hipEventElapsedTime(&ms, L->run_ev[2 * slot], L->run_ev[2 * slot + 1]));
            L->ms[q] += ms;
            L->launches[q]++;
            L->samples.push_back({q, s.cur, ms});
            if (q == 0) {
                const double row = 16.0 * (L->ntau + 1);
                L->tv_bytes += L->nk * (s.cur + 2) * row + row + 16.0 * (s.cur + 1);
            }
        }
        if (s.active == 0 && s.cur >= t0) {
            iters[s.cur - t0] = s.iters;
            errs[s.cur - t0] = s.last;
            L->last_iters = s.iters;
            B = std::max(1, std::min(std::min(s.iters, maxiter), (int)LC::RUN_MAXB));
            if (s.cur == t1) over = true;
        }
        prev = s;
        fin++;
    }
    return CNTRX_OK;
}

}  // extern "C"

// ====================================================================================
// continuous batching of independent solves
// ====================================================================================
#include "server.inc"

// ====================================================================================
// text checkpoints (print_to_file / read_from_file)
// ====================================================================================
#include "textio.inc"
