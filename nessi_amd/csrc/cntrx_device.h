// Device-side building blocks of the cntrx HIP library (gfx950, complex FP64).
//
// Storage follows cntr::herm_matrix exactly (libcntr/cntr/cntr_herm_matrix_impl.hpp:
// 456-478): ret(i,j<=i) row-packed lower triangle, les(i<=j,j) column-packed upper
// triangle, tv(i,m) dense, mat(m); each element is a row-major NORB x NORB block of
// interleaved complex doubles.  Offsets are 64-bit.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cntrx.h"

namespace cntrx {

// complex double, layout of std::complex<double> (interleaved re, im)
struct __attribute__((aligned(16))) cplx {
    double x, y;
};

__host__ __device__ inline cplx mk(double r, double i = 0.0) {
    cplx z;
    z.x = r;
    z.y = i;
    return z;
}
__host__ __device__ inline cplx operator+(cplx a, cplx b) { return mk(a.x + b.x, a.y + b.y); }
__host__ __device__ inline cplx operator-(cplx a, cplx b) { return mk(a.x - b.x, a.y - b.y); }
__host__ __device__ inline cplx operator-(cplx a) { return mk(-a.x, -a.y); }
__host__ __device__ inline cplx operator*(cplx a, cplx b) {
    return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__host__ __device__ inline cplx operator*(double s, cplx a) { return mk(s * a.x, s * a.y); }
__host__ __device__ inline cplx operator*(cplx a, double s) { return mk(s * a.x, s * a.y); }
__host__ __device__ inline cplx &operator+=(cplx &a, cplx b) {
    a.x += b.x;
    a.y += b.y;
    return a;
}
__host__ __device__ inline cplx &operator-=(cplx &a, cplx b) {
    a.x -= b.x;
    a.y -= b.y;
    return a;
}
__host__ __device__ inline cplx conj(cplx a) { return mk(a.x, -a.y); }
__host__ __device__ inline double cabs2(cplx a) { return a.x * a.x + a.y * a.y; }
// a += b*c
__host__ __device__ inline void cfma(cplx &a, cplx b, cplx c) {
    a.x = fma(b.x, c.x, a.x);
    a.x = fma(-b.y, c.y, a.x);
    a.y = fma(b.x, c.y, a.y);
    a.y = fma(b.y, c.x, a.y);
}
// Smith's algorithm for a/b
__host__ __device__ inline cplx cdiv(cplx a, cplx b) {
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, d = b.x + b.y * r;
        return mk((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    double r = b.x / b.y, d = b.x * r + b.y;
    return mk((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__host__ __device__ inline cplx cexp_i(double phi) {  // e^{i phi}
    double s, c;
    sincos(phi, &s, &c);
    return mk(c, s);
}

// --------------------------------------------------------------------------------
// Integrator weights for orders k = 0..5 (libcntr/cntr/integration.hpp:128-514),
// generated on the host (weights.cpp) and kept in constant memory.
// --------------------------------------------------------------------------------
struct WeightTable {
    double greg[6][144];   // (2k+2)^2 start table, row n < 2k+2
    double omega[6][6];    // Gregory end weights
    double bd[6][7];       // BDF(k+1) weights
    double P[6][36];       // poly_interpolation (k+1)^2
    double D[6][36];       // poly_differentiation (k+1)^2
    double W[6][216];      // poly_integration (k+1)^3
    double rc[6][144];     // rcorr (k-1)(k+1)^2
};

// defined once: the library is built as a single translation unit (cntrx_lib.hip)
__constant__ WeightTable c_w;

__device__ inline double gw(int k, int n, int j) {
    int k1 = k + 1, k2 = 2 * k1;
    if (n >= k2) {
        if (j > k && j < n - k) return 1.0;
        if (j < k1) return c_w.omega[k][j];
        return c_w.omega[k][n - j];
    }
    return c_w.greg[k][n * k2 + j];
}
__device__ inline double gomega(int k, int j) { return c_w.omega[k][j]; }

// Gregory weights with the end weights held in registers: the interior value 1 and the
// omega selection cost a few selects instead of a (lane-divergent) constant-memory load;
// only the short start rows n < 2k+2 read the table.
struct GregRegs {
    double om[6];
    int k;
};
__device__ inline GregRegs greg_regs(int k) {
    GregRegs g;
    g.k = k;
#pragma unroll
    for (int p = 0; p < 6; p++) g.om[p] = (p <= k) ? c_w.omega[k][p] : 0.0;
    return g;
}
__device__ __forceinline__ double gw_fast(const GregRegs &g, int n, int j) {
    const int k = g.k;
    if (n >= 2 * k + 2) {
        if (j > k && j < n - k) return 1.0;
        const int i = (j <= k) ? j : n - j;
        double w = g.om[0];
#pragma unroll
        for (int p = 1; p < 6; p++) w = (i == p) ? g.om[p] : w;
        return w;
    }
    return c_w.greg[k][n * (2 * k + 2) + j];
}
// Gregory quadrature sum_{l=0}^{nn} w(nn, l) f(l) (libcntr/cntr/integration.hpp:382-390):
// the start table for nn < 2k+2; otherwise omega at both ends and an unweighted
// interior, which is the long loop and carries no weight arithmetic at all.
template <class F>
__device__ __forceinline__ cplx greg_sum(const GregRegs &G, int nn, F f) {
    const int k = G.k;
    cplx a = mk(0.0), b = mk(0.0);
    if (nn < 2 * k + 2) {
        const double *row = &c_w.greg[k][nn * (2 * k + 2)];
        for (int l = 0; l <= nn; l++) a += row[l] * f(l);
        return a;
    }
    int l = k + 1;
    for (; l + 1 < nn - k; l += 2) {
        a += f(l);
        b += f(l + 1);
    }
    if (l < nn - k) a += f(l);
#pragma unroll
    for (int p = 0; p < 6; p++)
        if (p <= k) b += G.om[p] * (f(p) + f(nn - p));
    return a + b;
}
__device__ inline double bdw(int k, int l) { return c_w.bd[k][l]; }
__device__ inline double pdiff(int k, int i, int l) { return c_w.D[k][i * (k + 1) + l]; }
__device__ inline double pint(int k, int i, int j, int l) {
    return c_w.W[k][(i * (k + 1) + j) * (k + 1) + l];
}
__device__ inline double pinterp(int k, int a, int l) { return c_w.P[k][a * (k + 1) + l]; }
__device__ inline double rcorr(int k, int m, int j, int l) {
    if (k < 2) return 0.0;
    return c_w.rc[k][((m - 1) * (k + 1) + j) * (k + 1) + l];
}

// --------------------------------------------------------------------------------
// Packed-storage views
// --------------------------------------------------------------------------------
// Pointer to device global memory.  Stored in address space 1 so that every access
// derived from it compiles to global_load/global_store rather than flat memory ops
// (a flat op also counts against the LDS wait counter, so an LDS wait would wait for
// outstanding global traffic); converts to a plain pointer for arithmetic.
template <class T>
struct GPtr {
    typedef __attribute__((address_space(1))) T gT;
    gT *p_;
    GPtr() = default;
    __host__ __device__ GPtr(T *p) : p_((gT *)p) {}
    __host__ __device__ operator T *() const { return (T *)p_; }
};

struct HermView {
    int nt, ntau, s, sig;
    GPtr<cplx> ret, les, tv, mat;
    __host__ __device__ HermView() {}
    __host__ __device__ HermView(const cntrx_herm &d)
        : nt(d.nt), ntau(d.ntau), s(d.size1), sig(d.sig), ret((cplx *)d.ret),
          les((cplx *)d.les), tv((cplx *)d.tv), mat((cplx *)d.mat) {}
    __host__ __device__ int64_t es() const { return (int64_t)s * s; }
    __device__ cplx *R(int i, int j) const {
        return (cplx *)ret + ((int64_t)i * (i + 1) / 2 + j) * es();
    }
    __device__ cplx *L(int i, int j) const {
        return (cplx *)les + ((int64_t)j * (j + 1) / 2 + i) * es();
    }
    __device__ cplx *TV(int i, int m) const {
        return (cplx *)tv + ((int64_t)i * (ntau + 1) + m) * es();
    }
    __device__ cplx *M(int m) const { return (cplx *)mat + (int64_t)m * es(); }
};

// Read element (a,b) of A^R(i,j) extended to i<j by A^R(i,j) = -[A^R(j,i)]^dagger
// (libcntr/cntr/cntr_herm_matrix_impl.hpp:569-934 symmetries).
__device__ inline cplx ret_any(const HermView &A, const HermView &Acc, int i, int j, int a,
                               int b) {
    if (i >= j) return A.R(i, j)[a * A.s + b];
    cplx v = Acc.R(j, i)[b * A.s + a];
    return mk(-v.x, v.y);
}
// A^<(i,j) for any i,j: les(i>j) = -[les(j,i)]^dagger
__device__ inline cplx les_any(const HermView &A, const HermView &Acc, int i, int j, int a,
                               int b) {
    if (i <= j) return A.L(i, j)[a * A.s + b];
    cplx v = Acc.L(j, i)[b * A.s + a];
    return mk(-v.x, v.y);
}

// --------------------------------------------------------------------------------
// Workgroup-cooperative dense solve in LDS: Gauss-Jordan with partial pivoting on
// an N x (N+R) augmented complex matrix (row-major, leading dimension ld).
// All threads of the workgroup must call it.
// --------------------------------------------------------------------------------
// Small dense solve by one wavefront in registers: A is N x (N+1) row-major in LDS
// (leading dimension N+1, right-hand side in the last column), N <= 6.  Lane j holds
// column j; Gauss-Jordan with partial pivoting as lds_gauss_jordan, with the pivot
// column and row values broadcast by v_readlane.  Every lane of the calling wave must
// call it; on return lane r < N of x holds solution row r.
__device__ __forceinline__ double rl_dbl(double v, int l) {
    long long b = __double_as_longlong(v);
    int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
    int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ cplx rl_cplx(cplx v, int l) { return mk(rl_dbl(v.x, l), rl_dbl(v.y, l)); }

__device__ __forceinline__ cplx wave_solve_small(const cplx *A, int N) {
    const int lane = threadIdx.x & 63, ld = N + 1;
    cplx col[6];
#pragma unroll
    for (int r = 0; r < 6; r++) col[r] = (r < N && lane <= N) ? A[r * ld + lane] : mk(0.0);
#pragma unroll
    for (int c = 0; c < 6; c++) {
        if (c >= N) break;
        // pivot: largest |A(r,c)|^2 for r >= c, first one on ties (column c is lane c)
        int p = c;
        double best = rl_dbl(cabs2(col[c]), c);
#pragma unroll
        for (int r = c + 1; r < 6; r++) {
            if (r < N) {
                const double v = rl_dbl(cabs2(col[r]), c);
                if (v > best) {
                    best = v;
                    p = r;
                }
            }
        }
        if (p != c) {
#pragma unroll
            for (int r = c + 1; r < 6; r++) {
                if (r == p) {
                    const cplx tmp = col[c];
                    col[c] = col[r];
                    col[r] = tmp;
                }
            }
        }
        const cplx d = rl_cplx(col[c], c);
        col[c] = cdiv(col[c], d);
#pragma unroll
        for (int r = 0; r < 6; r++) {
            if (r < N && r != c) {
                const cplx f = rl_cplx(col[r], c);
                if (lane == c) col[r] = mk(0.0);
                else col[r] -= f * col[c];
            }
        }
    }
    cplx x = mk(0.0);
#pragma unroll
    for (int r = 0; r < 6; r++) {
        const cplx v = rl_cplx(col[r], N);
        if (lane == r) x = v;
    }
    return x;
}

__device__ inline void lds_gauss_jordan(cplx *A, int N, int R, int ld, int *piv_sh) {
    const int tid = threadIdx.x, nth = blockDim.x;
    const int ncol = N + R;
    for (int c = 0; c < N; c++) {
        if (tid == 0) {
            int p = c;
            double best = cabs2(A[c * ld + c]);
            for (int r = c + 1; r < N; r++) {
                double v = cabs2(A[r * ld + c]);
                if (v > best) {
                    best = v;
                    p = r;
                }
            }
            *piv_sh = p;
        }
        __syncthreads();
        int p = *piv_sh;
        if (p != c)
            for (int j = tid; j < ncol; j += nth) {
                cplx t = A[c * ld + j];
                A[c * ld + j] = A[p * ld + j];
                A[p * ld + j] = t;
            }
        __syncthreads();
        cplx d = A[c * ld + c];
        __syncthreads();
        for (int j = tid; j < ncol; j += nth) A[c * ld + j] = cdiv(A[c * ld + j], d);
        __syncthreads();
        for (int idx = tid; idx < N * ncol; idx += nth) {
            int r = idx / ncol, j = idx % ncol;
            if (r == c || j == c) continue;
            cplx f = A[r * ld + c];
            A[r * ld + j] -= f * A[c * ld + j];
        }
        __syncthreads();
        for (int r = tid; r < N; r += nth)
            if (r != c) A[r * ld + c] = mk(0.0);
        __syncthreads();
    }
}

// Single-thread inverse of an s x s block (s <= 8) by Gauss-Jordan with partial
// pivoting; in/out may not alias.
template <int S>
__device__ inline void block_inverse(const cplx *in, cplx *out) {
    cplx a[S * S], b[S * S];
#pragma unroll
    for (int i = 0; i < S * S; i++) {
        a[i] = in[i];
        b[i] = mk((i / S) == (i % S) ? 1.0 : 0.0);
    }
    for (int c = 0; c < S; c++) {
        int p = c;
        double best = cabs2(a[c * S + c]);
        for (int r = c + 1; r < S; r++) {
            double v = cabs2(a[r * S + c]);
            if (v > best) {
                best = v;
                p = r;
            }
        }
        if (p != c)
            for (int j = 0; j < S; j++) {
                cplx t = a[c * S + j];
                a[c * S + j] = a[p * S + j];
                a[p * S + j] = t;
                t = b[c * S + j];
                b[c * S + j] = b[p * S + j];
                b[p * S + j] = t;
            }
        cplx d = a[c * S + c];
        for (int j = 0; j < S; j++) {
            a[c * S + j] = cdiv(a[c * S + j], d);
            b[c * S + j] = cdiv(b[c * S + j], d);
        }
        for (int r = 0; r < S; r++) {
            if (r == c) continue;
            cplx f = a[r * S + c];
            for (int j = 0; j < S; j++) {
                a[r * S + j] -= f * a[c * S + j];
                b[r * S + j] -= f * b[c * S + j];
            }
        }
    }
#pragma unroll
    for (int i = 0; i < S * S; i++) out[i] = b[i];
}
template <>
__device__ inline void block_inverse<1>(const cplx *in, cplx *out) {
    out[0] = cdiv(mk(1.0), in[0]);
}

}  // namespace cntrx
