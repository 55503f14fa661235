// CPU oracle for the Kadanoff-Baym time-stepping path -- TEST INFRASTRUCTURE ONLY.
//
// A serial C++ restatement of the reference libcntr algorithms (no Eigen), used as
// the checker for the HIP implementation in nessi_amd/.  The product library never
// links or calls this file.  Reference paths are relative to libcntr/cntr/.
//
// Parity status per function is recorded in the header comment of each block:
//   "pinned" = checked by a test against reference fixtures, closed forms or brute
//   force; the tests live in tests/test_oracle_*.py.
#include "cntr_oracle.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <complex>
#include <cstring>
#include <mutex>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

typedef std::complex<double> cplx;
const double kPi = 3.14159265358979323846;
const cplx kI(0.0, 1.0);

// ---------------------------------------------------------------------------------
// Integrator weights (integration.hpp:128-514, integration.cpp:15-1184).
// The reference reads hard-coded decimal tables; here every weight is generated from
// its defining property in 80-bit arithmetic and rounded once to double:
//   * poly_interpolation: inverse Vandermonde of nodes 0..k (integration.cpp:15-35)
//   * poly_differentiation / poly_integration: derivative / integral of the
//     interpolating polynomial (integration.cpp:37-86)
//   * bd_weights: backward differentiation of order k+1 = derivative at node k+1 of
//     the order-(k+1) interpolant, nodes reversed (integration.hpp:74-80)
//   * gregory omega: end corrections fixed by exactness for polynomials of degree
//     <= k against the Euler-Maclaurin end terms; the start table holds
//     poly_integration(0,n,.) for n<=k and the overlapping Gregory sums for
//     k < n < 2k+2 (integration.cpp:796-804 assembles the same table layout)
//   * rcorr(m,j,l) = int_0^m L_j(m-x) L_l(x) dx over the Lagrange basis
//     (integration.hpp:43-55)
// ---------------------------------------------------------------------------------
// Exact rational arithmetic on __int128 (all weights for k <= 6 have small
// numerators and denominators, so nothing overflows).
struct Rat {
    __int128 p = 0, q = 1;
    Rat() {}
    Rat(long long a) : p(a), q(1) {}
    Rat(__int128 a, __int128 b) : p(a), q(b) { norm(); }
    static __int128 g(__int128 a, __int128 b) {
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
        __int128 d = g(p, q);
        if (d > 1) {
            p /= d;
            q /= d;
        }
    }
    Rat operator+(const Rat &o) const { return Rat(p * o.q + o.p * q, q * o.q); }
    Rat operator-(const Rat &o) const { return Rat(p * o.q - o.p * q, q * o.q); }
    Rat operator*(const Rat &o) const { return Rat(p * o.p, q * o.q); }
    Rat operator/(const Rat &o) const { return Rat(p * o.q, q * o.p); }
    Rat &operator+=(const Rat &o) { return *this = *this + o; }
    Rat &operator-=(const Rat &o) { return *this = *this - o; }
    bool zero() const { return p == 0; }
    double d() const { return (double)((long double)p / (long double)q); }
};

static void rat_inverse(int n, std::vector<Rat> a, std::vector<Rat> &inv) {
    inv.assign(n * n, Rat(0));
    for (int i = 0; i < n; i++) inv[i * n + i] = Rat(1);
    for (int c = 0; c < n; c++) {
        int p = c;
        while (a[p * n + c].zero()) p++;
        if (p != c)
            for (int j = 0; j < n; j++) {
                std::swap(a[c * n + j], a[p * n + j]);
                std::swap(inv[c * n + j], inv[p * n + j]);
            }
        Rat d = a[c * n + c];
        for (int j = 0; j < n; j++) {
            a[c * n + j] = a[c * n + j] / d;
            inv[c * n + j] = inv[c * n + j] / d;
        }
        for (int r = 0; r < n; r++) {
            if (r == c || a[r * n + c].zero()) continue;
            Rat f = a[r * n + c];
            for (int j = 0; j < n; j++) {
                a[r * n + j] -= f * a[c * n + j];
                inv[r * n + j] -= f * inv[c * n + j];
            }
        }
    }
}

static Rat rpow(long long x, int e) {
    Rat r(1);
    for (int i = 0; i < e; i++) r = r * Rat(x);
    return r;
}

// coefficients c[alpha][l]: P(x) = sum_alpha x^alpha sum_l c[alpha][l] f_l, nodes 0..k
static std::vector<Rat> interp_coeffs(int k) {
    int k1 = k + 1;
    std::vector<Rat> V(k1 * k1), P;
    for (int i = 0; i <= k; i++)
        for (int n = 0; n <= k; n++) V[i * k1 + n] = rpow(i, n);
    rat_inverse(k1, V, P);
    return P;
}

struct Integ {
    int k = -1;
    std::vector<double> P, D, W, bd, greg, omega, rc;
    double poly_interpolation(int a, int l) const { return P[a * (k + 1) + l]; }
    double poly_differentiation(int i, int l) const { return D[i * (k + 1) + l]; }
    double poly_integration(int i, int j, int l) const {
        return W[(i * (k + 1) + j) * (k + 1) + l];
    }
    double bd_weights(int l) const { return bd[l]; }
    double gregory_omega(int j) const { return omega[j]; }
    // integration.hpp:382-390
    double gregory_weights(int n, int j) const {
        int k1 = k + 1, k2 = 2 * k1;
        if (n >= k2) {
            if (j > k && j < n - k) return 1.0;
            if (j < k1) return omega[j];
            return omega[n - j];
        }
        return greg[n * k2 + j];
    }
    // integration.hpp:463-468
    double rcorr(int m, int j, int l) const {
        int k1 = k + 1;
        if (k < 2) return 0.0;
        return rc[((m - 1) * k1 + j) * k1 + l];
    }

    void build(int kk) {
        k = kk;
        int k1 = k + 1, k2 = 2 * k1;
        std::vector<Rat> c = interp_coeffs(k);
        P.resize(k1 * k1);
        D.resize(k1 * k1);
        W.resize(k1 * k1 * k1);
        for (int i = 0; i < k1 * k1; i++) P[i] = c[i].d();
        for (int i = 0; i <= k; i++)
            for (int l = 0; l <= k; l++) {
                Rat s;  // d/dx sum_n x^n c[n][l] at x=i
                for (int n = 1; n <= k; n++) s += Rat(n) * rpow(i, n - 1) * c[n * k1 + l];
                D[i * k1 + l] = s.d();
            }
        std::vector<Rat> Wr(k1 * k1 * k1);
        for (int i = 0; i <= k; i++)
            for (int j = 0; j <= k; j++)
                for (int l = 0; l <= k; l++) {
                    Rat s;
                    for (int r = 0; r <= k; r++)
                        s += c[r * k1 + l] * (rpow(j, r + 1) - rpow(i, r + 1)) / Rat(r + 1);
                    Wr[(i * k1 + j) * k1 + l] = s;
                    W[(i * k1 + j) * k1 + l] = s.d();
                }
        // BDF of order k+1: derivative at node k+1 of interpolant on nodes 0..k+1
        {
            int q = k + 1, q1 = q + 1;
            std::vector<Rat> cq = interp_coeffs(q);
            bd.resize(q1);
            for (int l = 0; l <= q; l++) {
                Rat s;
                for (int n = 1; n <= q; n++) s += Rat(n) * rpow(q, n - 1) * cq[n * q1 + (q - l)];
                bd[l] = s.d();
            }
        }
        // Gregory end weights: sum_j (omega_j - 1) j^p = R_p,  p = 0..k, with
        // R_0 = -1/2, R_p = B_{p+1}/(p+1) for odd p, 0 for even p > 0.
        {
            const Rat bern[] = {Rat(1), Rat(-1, 2), Rat(1, 6), Rat(0), Rat(-1, 30), Rat(0),
                                Rat(1, 42), Rat(0), Rat(-1, 30), Rat(0), Rat(5, 66)};
            std::vector<Rat> V(k1 * k1), Vi;
            for (int p = 0; p <= k; p++)
                for (int j = 0; j <= k; j++) V[p * k1 + j] = rpow(j, p);
            rat_inverse(k1, V, Vi);
            std::vector<Rat> R(k1), om(k1);
            for (int p = 0; p <= k; p++) {
                if (p == 0) R[p] = Rat(-1, 2);
                else if (p % 2 == 1) R[p] = bern[p + 1] / Rat(p + 1);
            }
            omega.resize(k1);
            for (int j = 0; j <= k; j++) {
                Rat a;
                for (int p = 0; p <= k; p++) a += Vi[j * k1 + p] * R[p];
                om[j] = Rat(1) + a;
                omega[j] = om[j].d();
            }
            greg.assign(k2 * k2, 0.0);
            for (int n = 0; n <= k; n++)
                for (int i = 0; i <= k; i++) greg[n * k2 + i] = Wr[(0 * k1 + n) * k1 + i].d();
            for (int n = k1; n < k2; n++)
                for (int i = 0; i <= n; i++) {
                    Rat w(1);
                    if (i <= k) w += om[i] - Rat(1);
                    if (i >= n - k) w += om[n - i] - Rat(1);
                    greg[n * k2 + i] = w.d();
                }
        }
        // rcorr(m,j,l) = int_0^m L_j(m-x) L_l(x) dx, m = 1..k-1
        rc.assign(k > 1 ? (k - 1) * k1 * k1 : 0, 0.0);
        for (int m = 1; m < k; m++) {
            for (int j = 0; j <= k; j++) {
                std::vector<Rat> a(k1);  // coefficients of L_j(m - x) in powers of x
                for (int e = 0; e <= k; e++) {
                    long long binom = 1;
                    for (int q = 0; q <= e; q++) {
                        Rat term = c[e * k1 + j] * Rat(binom) * rpow(m, e - q);
                        if (q % 2) a[q] -= term;
                        else a[q] += term;
                        binom = binom * (e - q) / (q + 1);
                    }
                }
                for (int l = 0; l <= k; l++) {
                    std::vector<Rat> prod(2 * k1);
                    for (int p = 0; p <= k; p++)
                        for (int q = 0; q <= k; q++) prod[p + q] += a[p] * c[q * k1 + l];
                    Rat s;
                    for (int e = 0; e < 2 * k1; e++) s += prod[e] * rpow(m, e + 1) / Rat(e + 1);
                    rc[((m - 1) * k1 + j) * k1 + l] = s.d();
                }
            }
        }
    }
};

// integration.hpp:506-514 caches I<T>(k) lazily without synchronisation and maps
// k=0 to a default Integrator(0); here all orders 0..5 are built once under a lock.
static const Integ &I(int k) {
    static Integ tab[6];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int kk = 0; kk <= 5; kk++) tab[kk].build(kk);
    });
    return tab[k];
}

// ---------------------------------------------------------------------------------
// Storage views (cntr_herm_matrix_impl.hpp:456-478) and small dense kernels
// (cntr_elements.hpp:22-529).  Solves use LU with partial pivoting where the
// reference uses Eigen FullPivLU (linalg_eigen.cpp:91-104); for the diagonally
// dominant systems of the time stepping both agree to rounding.
// ---------------------------------------------------------------------------------
struct Herm {
    int nt, ntau, s, sig;
    cplx *ret_, *les_, *tv_, *mat_;
    explicit Herm(const orc_herm *d)
        : nt(d->nt), ntau(d->ntau), s(d->size1), sig(d->sig),
          ret_((cplx *)d->ret), les_((cplx *)d->les), tv_((cplx *)d->tv),
          mat_((cplx *)d->mat) {}
    Herm(int nt_, int ntau_, int s_, int sig_, std::vector<cplx> &store)
        : nt(nt_), ntau(ntau_), s(s_), sig(sig_) {
        long es = (long)s * s;
        long ntri = nt >= 0 ? (long)(nt + 1) * (nt + 2) / 2 : 0;
        long ntv = nt >= 0 ? (long)(nt + 1) * (ntau + 1) : 0;
        store.assign((2 * ntri + ntv + ntau + 1) * es, cplx(0));
        ret_ = store.data();
        les_ = ret_ + ntri * es;
        tv_ = les_ + ntri * es;
        mat_ = tv_ + ntv * es;
    }
    long es() const { return (long)s * s; }
    cplx *ret(int i, int j) const { return ret_ + ((long)i * (i + 1) / 2 + j) * es(); }
    cplx *les(int i, int j) const { return les_ + ((long)j * (j + 1) / 2 + i) * es(); }
    cplx *tv(int i, int m) const { return tv_ + ((long)i * (ntau + 1) + m) * es(); }
    cplx *mat(int m) const { return mat_ + (long)m * es(); }
};

struct Func {
    int nt, s;
    cplx *d;
    explicit Func(const orc_func *f) : nt(f->nt), s(f->size1), d((cplx *)f->data) {}
    cplx *ptr(int t) const { return d + (long)(t + 1) * s * s; }
};

typedef std::vector<cplx> Buf;

static inline void e_zero(int s, cplx *z) { std::fill(z, z + s * s, cplx(0)); }
static inline void e_set(int s, cplx *z, const cplx *a) { std::copy(a, a + s * s, z); }
static inline void e_diag(int s, cplx *z, cplx c) {
    e_zero(s, z);
    for (int i = 0; i < s; i++) z[i * s + i] = c;
}
// z = a^dagger   (element_conj, cntr_elements.hpp:38-47)
static inline void e_herm(int s, cplx *z, const cplx *a) {
    for (int l = 0; l < s; l++)
        for (int m = 0; m < s; m++) z[l * s + m] = std::conj(a[m * s + l]);
}
// z = -a^dagger  (element_minusconj, cntr_elements.hpp:56-66)
static inline void e_mherm(int s, cplx *z, const cplx *a) {
    for (int l = 0; l < s; l++)
        for (int m = 0; m < s; m++) {
            cplx t = a[m * s + l];
            z[l * s + m] = cplx(-t.real(), t.imag());
        }
}
static inline void e_smul(int s, cplx *z, cplx c) {
    for (int i = 0; i < s * s; i++) z[i] *= c;
}
static inline void e_incr(int s, cplx *z, const cplx *a) {
    for (int i = 0; i < s * s; i++) z[i] += a[i];
}
static inline void e_incr(int s, cplx *z, cplx c, const cplx *a) {
    for (int i = 0; i < s * s; i++) z[i] += c * a[i];
}
// z += a*b
static inline void e_incr_mm(int s, cplx *z, const cplx *a, const cplx *b) {
    if (s == 1) {
        z[0] += a[0] * b[0];
        return;
    }
    for (int i = 0; i < s; i++)
        for (int j = 0; j < s; j++) {
            cplx acc = 0;
            for (int l = 0; l < s; l++) acc += a[i * s + l] * b[l * s + j];
            z[i * s + j] += acc;
        }
}
// z += c*(a*b)
static inline void e_incr_mm(int s, cplx *z, cplx c, const cplx *a, const cplx *b) {
    if (s == 1) {
        z[0] += c * a[0] * b[0];
        return;
    }
    for (int i = 0; i < s; i++)
        for (int j = 0; j < s; j++) {
            cplx acc = 0;
            for (int l = 0; l < s; l++) acc += a[i * s + l] * b[l * s + j];
            z[i * s + j] += c * acc;
        }
}
static inline void e_mult(int s, cplx *z, const cplx *a, const cplx *b) {
    e_zero(s, z);
    e_incr_mm(s, z, a, b);
}

// Solve sum_m M(l,m) X(m) = Q(l) for nb blocks of size s (element_linsolve_right,
// cntr_elements.hpp:442-477).
static void solve_right(int s, int nb, cplx *X, const cplx *M, const cplx *Q) {
    int N = s * nb;
    long es = (long)s * s;
    if (N == 1) {
        X[0] = Q[0] / M[0];
        return;
    }
    Buf A((long)N * N), B((long)N * s);
    for (int l = 0; l < nb; l++)
        for (int m = 0; m < nb; m++)
            for (int p = 0; p < s; p++)
                for (int q = 0; q < s; q++)
                    A[(long)(l * s + p) * N + m * s + q] = M[(l * nb + m) * es + p * s + q];
    for (int l = 0; l < nb; l++)
        for (int p = 0; p < s; p++)
            for (int q = 0; q < s; q++) B[(long)(l * s + p) * s + q] = Q[l * es + p * s + q];
    std::vector<int> piv(N);
    for (int c = 0; c < N; c++) {
        int p = c;
        double best = std::abs(A[(long)c * N + c]);
        for (int r = c + 1; r < N; r++) {
            double v = std::abs(A[(long)r * N + c]);
            if (v > best) {
                best = v;
                p = r;
            }
        }
        if (p != c) {
            for (int j = 0; j < N; j++) std::swap(A[(long)c * N + j], A[(long)p * N + j]);
            for (int j = 0; j < s; j++) std::swap(B[(long)c * s + j], B[(long)p * s + j]);
        }
        cplx d = A[(long)c * N + c];
        for (int r = c + 1; r < N; r++) {
            cplx f = A[(long)r * N + c] / d;
            if (f == cplx(0)) continue;
            for (int j = c + 1; j < N; j++) A[(long)r * N + j] -= f * A[(long)c * N + j];
            for (int j = 0; j < s; j++) B[(long)r * s + j] -= f * B[(long)c * s + j];
        }
    }
    for (int c = N - 1; c >= 0; c--)
        for (int j = 0; j < s; j++) {
            cplx acc = B[(long)c * s + j];
            for (int t = c + 1; t < N; t++) acc -= A[(long)c * N + t] * B[(long)t * s + j];
            B[(long)c * s + j] = acc / A[(long)c * N + c];
        }
    for (int l = 0; l < nb; l++)
        for (int p = 0; p < s; p++)
            for (int q = 0; q < s; q++) X[l * es + p * s + q] = B[(long)(l * s + p) * s + q];
}

// Solve sum_m X(m) M(l,m) = Q(l) (element_linsolve_left, cntr_elements.hpp:484-491):
// conjugate-transpose every block, solve right, transpose back.
static void solve_left(int s, int nb, cplx *X, const cplx *M, const cplx *Q) {
    long es = (long)s * s;
    if (s * nb == 1) {
        X[0] = Q[0] / M[0];
        return;
    }
    Buf Mc(nb * nb * es), Qc(nb * es), Xc(nb * es);
    for (int b = 0; b < nb * nb; b++) e_herm(s, Mc.data() + b * es, M + b * es);
    for (int b = 0; b < nb; b++) e_herm(s, Qc.data() + b * es, Q + b * es);
    solve_right(s, nb, Xc.data(), Mc.data(), Qc.data());
    for (int b = 0; b < nb; b++) e_herm(s, X + b * es, Xc.data() + b * es);
}

static void e_inverse(int s, cplx *z, const cplx *a) {
    Buf one(s * s);
    e_diag(s, one.data(), 1.0);
    solve_right(s, 1, z, a, one.data());
}

// ---------------------------------------------------------------------------------
// Matsubara integrals (cntr_convolution_impl.hpp:85-193 and 409-515)
// ---------------------------------------------------------------------------------
static void matsubara_integral_1(int s, int m, int ntau, cplx *C, const cplx *A,
                                 const cplx *B, const Integ &In, int sig) {
    int k = In.k, k1 = k + 1, k2 = 2 * k1;
    long es = (long)s * s;
    Buf c1(es, 0), c2(es, 0);
    if (m >= k2 - 1) {
        for (int j = 0; j <= k; j++)
            e_incr_mm(s, c1.data(), In.gregory_omega(j), A + (m - j) * es, B + j * es);
        for (int j = k1; j < m - k; j++) e_incr_mm(s, c1.data(), A + (m - j) * es, B + j * es);
        for (int j = m - k; j <= m; j++)
            e_incr_mm(s, c1.data(), In.gregory_omega(m - j), A + (m - j) * es, B + j * es);
    } else if (m >= k) {
        for (int j = 0; j <= m; j++)
            e_incr_mm(s, c1.data(), In.gregory_weights(m, j), A + (m - j) * es, B + j * es);
    } else if (m > 0) {
        for (int l = 0; l <= k; l++)
            for (int j = 0; j <= k; j++)
                e_incr_mm(s, c1.data(), In.rcorr(m, l, j), A + l * es, B + j * es);
    }
    if (ntau - m >= k2 - 1) {
        int t = 0;
        for (int j = m; j <= m + k; j++, t++)
            e_incr_mm(s, c2.data(), In.gregory_omega(j - m), A + (ntau - t) * es, B + j * es);
        for (int j = m + k1; j < ntau - k; j++, t++)
            e_incr_mm(s, c2.data(), A + (ntau - t) * es, B + j * es);
        for (int j = ntau - k; j <= ntau; j++, t++)
            e_incr_mm(s, c2.data(), In.gregory_omega(ntau - j), A + (ntau - t) * es,
                      B + j * es);
    } else if (ntau - m >= k) {
        int t = 0;
        for (int j = m; j <= ntau; j++, t++)
            e_incr_mm(s, c2.data(), In.gregory_weights(ntau - m, ntau - j),
                      A + (ntau - t) * es, B + j * es);
    } else if (ntau - m > 0) {
        for (int l = 0; l <= k; l++)
            for (int j = 0; j <= k; j++)
                e_incr_mm(s, c2.data(), In.rcorr(ntau - m, l, j), A + (ntau - l) * es,
                          B + (ntau - j) * es);
    }
    for (long l = 0; l < es; l++) C[l] = c1[l] + cplx(sig, 0.0) * c2[l];
}

static void matsubara_integral_2(int s, int m, int ntau, cplx *C, const cplx *A,
                                 const cplx *B, const Integ &In, int sig) {
    int k = In.k, k1 = k + 1, k2 = 2 * k1;
    long es = (long)s * s;
    Buf c1(es, 0), c2(es, 0);
    if (m == 0) {
    } else if (m < k) {
        for (int j = 0; j <= k; j++)
            for (int l = 0; l <= k; l++)
                e_incr_mm(s, c1.data(), In.rcorr(m, l, j), A + j * es, B + (ntau - l) * es);
    } else if (m < k2 - 1) {
        for (int l = 0; l <= m; l++)
            e_incr_mm(s, c1.data(), In.gregory_weights(m, l), A + (m - l) * es,
                      B + (ntau - l) * es);
    } else {
        for (int l = 0; l <= k; l++)
            e_incr_mm(s, c1.data(), In.gregory_omega(l), A + (m - l) * es, B + (ntau - l) * es);
        for (int l = k1; l < m - k; l++) e_incr_mm(s, c1.data(), A + (m - l) * es, B + (ntau - l) * es);
        for (int l = m - k; l <= m; l++)
            e_incr_mm(s, c1.data(), In.gregory_omega(m - l), A + (m - l) * es,
                      B + (ntau - l) * es);
    }
    if (m == ntau) {
    } else if (m > ntau - k) {
        for (int l = 0; l <= k; l++)
            for (int j = 0; j <= k; j++)
                e_incr_mm(s, c2.data(), In.rcorr(ntau - m, l, j), A + (ntau - l) * es, B + j * es);
    } else if (m > ntau - k2 + 1) {
        for (int l = 0; l <= ntau - m; l++)
            e_incr_mm(s, c2.data(), In.gregory_weights(ntau - m, l), A + (m + l) * es, B + l * es);
    } else {
        int t = 0;
        for (int l = m; l <= m + k; l++, t++)
            e_incr_mm(s, c2.data(), In.gregory_omega(l - m), A + l * es, B + t * es);
        for (int l = m + k1; l < ntau - k; l++, t++) e_incr_mm(s, c2.data(), A + l * es, B + t * es);
        for (int l = ntau - k; l <= ntau; l++, t++)
            e_incr_mm(s, c2.data(), In.gregory_omega(ntau - l), A + l * es, B + t * es);
    }
    for (long l = 0; l < es; l++) C[l] = c2[l] + cplx(sig, 0.0) * c1[l];
}

// cntr_convolution_impl.hpp:545-571
static void convolution_matsubara(Herm &C, Herm &A, Herm &B, const Integ &In, double beta) {
    int ntau = A.ntau, s = C.s;
    for (int m = 0; m <= ntau; m++)
        matsubara_integral_1(s, m, ntau, C.mat(m), A.mat(0), B.mat(0), In, A.sig);
    double dtau = beta / ntau;
    long tot = (long)(ntau + 1) * C.es();
    for (long l = 0; l < tot; l++) C.mat_[l] *= dtau;
}

// ---------------------------------------------------------------------------------
// Real-time convolution pieces (cntr_convolution_impl.hpp:740-1546)
// ---------------------------------------------------------------------------------
// cntr_convolution_impl.hpp:740-855
static void convolution_timestep_ret(int n, Herm &C, Herm &A, Herm &Acc, Herm &B, Herm &Bcc,
                                     const Integ &In, double h) {
    int k = In.k, s = C.s;
    long es = C.es();
    Buf res((long)(n + 1) * es, 0), at(es), bt(es);
    if (n >= k) {
        for (int m = 0; m <= n; m++) {
            for (long l = 0; l < es; l++) at[l] = A.ret(n, m)[l] * h;
            cplx *bret = B.ret(m, 0);
            int j = 0;
            if (m < n - k) {
                for (; j < m - k; j++) e_incr_mm(s, res.data() + j * es, at.data(), bret + j * es);
            } else {
                double w = In.gregory_omega(n - m);
                for (; j < m - k; j++)
                    e_incr_mm(s, res.data() + j * es, w, at.data(), bret + j * es);
            }
            for (j = std::max(0, m - k); j <= m; j++)
                e_incr_mm(s, res.data() + j * es, In.gregory_weights(n - j, n - m), at.data(),
                          bret + j * es);
        }
        for (int m = n - k; m < n; m++) {
            for (long l = 0; l < es; l++) at[l] = A.ret(n, m)[l] * h;
            for (int j = m + 1; j <= n; j++) {
                e_herm(s, bt.data(), Bcc.ret(j, m));
                e_incr_mm(s, res.data() + j * es, -In.gregory_weights(n - j, n - m), at.data(),
                          bt.data());
            }
        }
    } else {
        for (int j = 0; j <= n; j++) {
            cplx *cr = res.data() + j * es;
            for (int m = 0; m <= k; m++) {
                double w = In.poly_integration(j, n, m) * h;
                if (m >= j) e_set(s, bt.data(), B.ret(m, j));
                else {
                    e_herm(s, bt.data(), Bcc.ret(j, m));
                    w = -w;
                }
                if (n >= m) e_set(s, at.data(), A.ret(n, m));
                else {
                    e_herm(s, at.data(), Acc.ret(m, n));
                    w = -w;
                }
                e_incr_mm(s, cr, w, at.data(), bt.data());
            }
        }
    }
    std::copy(res.begin(), res.end(), C.ret(n, 0));
}

// cntr_convolution_impl.hpp:900-974: ctv(m) = dtau*[A^tv(n) * B^M](m) + h sum_j w A^R B^tv
static void convolution_timestep_tv(int n, cplx *ctv, Herm &A, Herm &Acc, Herm &B,
                                    const Integ &In, double beta, double h) {
    int k = In.k, s = A.s, ntau = A.ntau, n1 = std::max(n, k);
    long es = A.es();
    double dtau = beta / ntau;
    Buf ct(es), at(es);
    for (int m = 0; m <= ntau; m++) {
        matsubara_integral_2(s, m, ntau, ct.data(), A.tv(n, 0), B.mat(0), In, B.sig);
        for (long l = 0; l < es; l++) ctv[m * es + l] = dtau * ct[l];
    }
    for (int j = 0; j <= n1; j++) {
        double w = In.gregory_weights(n, j);
        if (n < j) e_mherm(s, at.data(), Acc.ret(j, n));
        else e_set(s, at.data(), A.ret(n, j));
        e_smul(s, at.data(), h);
        cplx *btv = B.tv(j, 0);
        if (w != 1.0) {
            for (int m = 0; m <= ntau; m++)
                e_incr_mm(s, ctv + m * es, w, at.data(), btv + m * es);
        } else {
            for (int m = 0; m <= ntau; m++) e_incr_mm(s, ctv + m * es, at.data(), btv + m * es);
        }
    }
}

// cntr_convolution_impl.hpp:1076-1139 (j1=0, j2=max(n,k))
static void convolution_timestep_les_tvvt(int n, cplx *cles, Herm &A, Herm &Bcc,
                                          const Integ &In, double beta) {
    int k = In.k, k1 = k + 1, k2 = 2 * k1, s = A.s, ntau = A.ntau, sig = A.sig;
    int n1 = std::max(n, k);
    long es = A.es();
    double dtau = beta / ntau;
    Buf bt((long)(ntau + 1) * es);
    cplx idtau(0, -dtau);
    for (int m = 0; m <= ntau; m++) e_herm(s, bt.data() + m * es, Bcc.tv(n, ntau - m));
    for (long l = 0; l < (long)(ntau + 1) * es; l++) bt[l] *= idtau * (-(double)sig);
    for (int j = 0; j <= n1; j++) {
        cplx *atv = A.tv(j, 0), *c = cles + j * es;
        if (ntau < k2 - 1) {
            for (int m = 0; m <= ntau; m++)
                e_incr_mm(s, c, In.gregory_weights(ntau, m), atv + m * es, bt.data() + m * es);
        } else {
            for (int m = 0; m <= k; m++)
                e_incr_mm(s, c, In.gregory_omega(m), atv + m * es, bt.data() + m * es);
            for (int m = k1; m < ntau - k; m++) e_incr_mm(s, c, atv + m * es, bt.data() + m * es);
            for (int m = ntau - k; m <= ntau; m++)
                e_incr_mm(s, c, In.gregory_omega(ntau - m), atv + m * es, bt.data() + m * es);
        }
    }
}

// cntr_convolution_impl.hpp:1245-1305 (j1=0, j2=max(n,k))
static void convolution_timestep_les_lesadv(int n, cplx *cles, Herm &A, Herm &Acc, Herm &B,
                                            Herm &Bcc, const Integ &In, double h) {
    int k = In.k, s = A.s, n1 = std::max(n, k);
    long es = A.es();
    Buf badv((long)(n1 + 1) * es), al(es);
    for (int m = 0; m <= n1; m++) {
        double w = h * In.gregory_weights(n, m);
        if (m <= n) {
            e_herm(s, badv.data() + m * es, Bcc.ret(n, m));
            for (long l = 0; l < es; l++) badv[m * es + l] *= w;
        } else {
            for (long l = 0; l < es; l++) badv[m * es + l] = -w * B.ret(m, n)[l];
        }
    }
    for (int j = 0; j <= n1; j++)
        for (int m = 0; m < j; m++) {
            e_mherm(s, al.data(), Acc.les(m, j));
            e_incr_mm(s, cles + j * es, al.data(), badv.data() + m * es);
        }
    for (int m = 0; m <= n1; m++) {
        int jmax = std::min(n1, m);
        for (int j = 0; j <= jmax; j++) e_incr_mm(s, cles + j * es, A.les(j, m), badv.data() + m * es);
    }
}

// cntr_convolution_impl.hpp:1395-1483
static void convolution_timestep_les_retles(int n, cplx *cles, Herm &A, Herm &Acc, Herm &B,
                                            Herm &Bcc, const Integ &In, double h) {
    int k = In.k, k1 = k + 1, k2 = 2 * k1, s = A.s, n1 = std::max(n, k);
    long es = A.es();
    Buf bt((long)(n1 + 1) * es), at(es);
    for (int m = 0; m <= n1; m++) {
        if (m <= n) {
            for (long l = 0; l < es; l++) bt[m * es + l] = h * B.les(m, n)[l];
        } else {
            e_herm(s, bt.data() + m * es, Bcc.les(n, m));
            for (long l = 0; l < es; l++) bt[m * es + l] *= -h;
        }
    }
    for (int j = 0; j <= n; j++) {
        cplx *c = cles + j * es;
        cplx *ar = A.ret(j, 0);
        if (j >= k2 - 1) {
            for (int m = 0; m <= k; m++)
                e_incr_mm(s, c, In.gregory_omega(m), ar + m * es, bt.data() + m * es);
            for (int m = k1; m < j - k; m++) e_incr_mm(s, c, ar + m * es, bt.data() + m * es);
            for (int m = j - k; m <= j; m++)
                e_incr_mm(s, c, In.gregory_omega(j - m), ar + m * es, bt.data() + m * es);
        } else {
            for (int m = 0; m <= j; m++)
                e_incr_mm(s, c, In.gregory_weights(j, m), ar + m * es, bt.data() + m * es);
        }
        if (j < k) {
            for (int m = j + 1; m <= k; m++) {
                e_herm(s, at.data(), Acc.ret(m, j));
                e_incr_mm(s, c, -In.gregory_weights(j, m), at.data(), bt.data() + m * es);
            }
        }
    }
}

// cntr_convolution_impl.hpp:1581-1609
static void convolution_timestep(int n, Herm &C, Herm &A, Herm &Acc, Herm &B, Herm &Bcc,
                                 const Integ &In, double beta, double h) {
    if (n == -1) {
        convolution_matsubara(C, A, B, In, beta);
        return;
    }
    long es = C.es();
    int ntau = C.ntau, n1 = std::max(n, In.k);
    convolution_timestep_ret(n, C, A, Acc, B, Bcc, In, h);
    Buf ctv((long)(ntau + 1) * es);
    convolution_timestep_tv(n, ctv.data(), A, Acc, B, In, beta, h);
    std::copy(ctv.begin(), ctv.end(), C.tv(n, 0));
    Buf cles((long)(n1 + 1) * es, 0);
    convolution_timestep_les_tvvt(n, cles.data(), A, Bcc, In, beta);
    convolution_timestep_les_lesadv(n, cles.data(), A, Acc, B, Bcc, In, h);
    convolution_timestep_les_retles(n, cles.data(), A, Acc, B, Bcc, In, h);
    for (int m = 0; m <= n; m++) e_set(C.s, C.les(m, n), cles.data() + m * es);
}

// ---------------------------------------------------------------------------------
// Free Green's function for constant H (cntr_equilibrium_impl.hpp:44-100, 873-953).
// Hermitian eigen-decomposition by cyclic Jacobi rotations.
// ---------------------------------------------------------------------------------
static void herm_eig(int s, const cplx *A0, double *ev, cplx *V) {
    Buf A(A0, A0 + s * s);
    e_diag(s, V, 1.0);
    for (int sweep = 0; sweep < 100; sweep++) {
        double off = 0;
        for (int p = 0; p < s; p++)
            for (int q = p + 1; q < s; q++) off += std::norm(A[p * s + q]);
        if (off < 1e-300) break;
        for (int p = 0; p < s; p++)
            for (int q = p + 1; q < s; q++) {
                cplx apq = A[p * s + q];
                double mag = std::abs(apq);
                if (mag < 1e-300) continue;
                double app = A[p * s + p].real(), aqq = A[q * s + q].real();
                double theta = 0.5 * std::atan2(2 * mag, aqq - app);
                double c = std::cos(theta), sn = std::sin(theta);
                cplx ph = apq / mag;  // e^{i phi}
                // unitary J: columns p,q rotated: J_pp = c, J_qq = c, J_pq = s*ph, J_qp = -s*conj(ph)
                for (int i = 0; i < s; i++) {  // A <- A J
                    cplx aip = A[i * s + p], aiq = A[i * s + q];
                    A[i * s + p] = c * aip - sn * std::conj(ph) * aiq;
                    A[i * s + q] = sn * ph * aip + c * aiq;
                }
                for (int j = 0; j < s; j++) {  // A <- J^dagger A
                    cplx apj = A[p * s + j], aqj = A[q * s + j];
                    A[p * s + j] = c * apj - sn * ph * aqj;
                    A[q * s + j] = sn * std::conj(ph) * apj + c * aqj;
                }
                for (int i = 0; i < s; i++) {  // V <- V J
                    cplx vip = V[i * s + p], viq = V[i * s + q];
                    V[i * s + p] = c * vip - sn * std::conj(ph) * viq;
                    V[i * s + q] = sn * ph * vip + c * viq;
                }
            }
    }
    for (int i = 0; i < s; i++) ev[i] = A[i * s + i].real();
}

static double fermi(double beta, double w) {
    double arg = w * beta;
    if (std::fabs(arg) > 100.0) return arg > 0 ? 0.0 : 1.0;
    return 1.0 / (1.0 + std::exp(arg));
}
static double fermi_exp(double beta, double tau, double w) {
    if (w < 0) return std::exp(w * tau) * fermi(beta, w);
    return std::exp((tau - beta) * w) * fermi(beta, -w);
}
static double bose(double beta, double w) {
    double arg = w * beta;
    if (arg < 0) return -1.0 - bose(beta, -w);
    if (std::fabs(arg) > 100.0) return 0.0;
    if (arg < 1e-10) return 1.0 / arg;
    return 1.0 / (std::exp(arg) - 1.0);
}
static double bose_exp(double beta, double tau, double w) {
    if (w < 0) return std::exp(w * tau) * bose(beta, w);
    return -std::exp((tau - beta) * w) * bose(beta, -w);
}

// value = V diag(d) V^dagger
static void vdv(int s, cplx *out, const cplx *V, const double *d) {
    for (int i = 0; i < s; i++)
        for (int j = 0; j < s; j++) {
            cplx acc = 0;
            for (int l = 0; l < s; l++) acc += V[i * s + l] * d[l] * std::conj(V[j * s + l]);
            out[i * s + j] = acc;
        }
}
static void vdv(int s, cplx *out, const cplx *V, const cplx *d) {
    for (int i = 0; i < s; i++)
        for (int j = 0; j < s; j++) {
            cplx acc = 0;
            for (int l = 0; l < s; l++) acc += V[i * s + l] * d[l] * std::conj(V[j * s + l]);
            out[i * s + j] = acc;
        }
}

static void green_from_H_const(Herm &G, double mu, const cplx *H0, double beta, double h) {
    int s = G.s, ntau = G.ntau, nt = G.nt, sig = G.sig;
    long es = G.es();
    double dtau = beta / ntau;
    Buf Hmu(es), V(es), val(es);
    for (long l = 0; l < es; l++) Hmu[l] = -H0[l];
    for (int i = 0; i < s; i++) Hmu[i * s + i] += mu;
    std::vector<double> ev(s), evm(s), d(s);
    herm_eig(s, Hmu.data(), ev.data(), V.data());
    for (int i = 0; i < s; i++) evm[i] = -ev[i];
    for (int m = 0; m <= ntau; m++) {
        double tau = m * dtau;
        for (int i = 0; i < s; i++)
            d[i] = sig == -1 ? -fermi_exp(beta, tau, ev[i]) : bose_exp(beta, tau, ev[i]);
        vdv(s, G.mat(m), V.data(), d.data());
    }
    if (nt < 0) return;
    // U(t) = exp(i t (mu - H)) = V diag(e^{i t ev}) V^dagger
    Buf U((long)(nt + 1) * es);
    std::vector<cplx> ph(s);
    for (int n = 0; n <= nt; n++) {
        for (int i = 0; i < s; i++) ph[i] = std::exp(kI * (n * h * ev[i]));
        vdv(s, U.data() + n * es, V.data(), ph.data());
    }
    Buf tmp(es), tmp2(es), Ud(es);
    for (int m = 0; m <= ntau; m++) {
        double tau = m * dtau;
        for (int i = 0; i < s; i++)
            d[i] = sig == -1 ? fermi_exp(beta, tau, evm[i]) : bose_exp(beta, tau, evm[i]);
        vdv(s, val.data(), V.data(), d.data());
        for (int n = 0; n <= nt; n++) {
            e_mult(s, G.tv(n, m), U.data() + n * es, val.data());
            e_smul(s, G.tv(n, m), sig == -1 ? kI : -kI);
        }
    }
    for (int i = 0; i < s; i++) d[i] = sig == -1 ? fermi(beta, evm[i]) : -bose(beta, evm[i]);
    vdv(s, val.data(), V.data(), d.data());
    for (int m = 0; m <= nt; m++)
        for (int n = 0; n <= m; n++) {
            e_herm(s, Ud.data(), U.data() + n * es);
            e_mult(s, G.ret(m, n), U.data() + m * es, Ud.data());
            e_smul(s, G.ret(m, n), -kI);
            e_herm(s, Ud.data(), U.data() + m * es);
            e_mult(s, tmp.data(), U.data() + n * es, val.data());
            e_mult(s, G.les(n, m), tmp.data(), Ud.data());
            e_smul(s, G.les(n, m), kI);
        }
}

// ---------------------------------------------------------------------------------
// green_from_H with a time-dependent H (cntr_equilibrium_impl.hpp:754-870, 955-1040):
// U(-1) = U(0) = 1 and U(t) = C(t) U(t-1) with the commutator-free exponentials of
// Alvermann & Fehske, J. Comput. Phys. 230, 5930 (2011):
//   CF2:  C = exp(-i h H(t - 1/2))
//   CF4:2 C = exp(-i h (a1 H1 + a2 H2)) exp(-i h (a2 H1 + a1 H2)),
//         H1,2 = H(t - 1/2 -+ sqrt(3)/6), a1,2 = (3 -+ 2 sqrt(3)) / 12,
// with H(x) the order-k polynomial through H(n1-k..n1), n1 = max(t, k)
// (interpolation, cntr_utilities_impl.hpp:285-312).  Without fixHam, H(t) for t > k is
// first overwritten by its extrapolation from H(t-1-k..t-1) (cntr_utilities_impl.hpp:
// 187-231).  The reference exponentiates with Eigen's MatrixBase::exp; these Hermitian
// arguments are exponentiated through their eigendecomposition here.
// ---------------------------------------------------------------------------------
static void func_interp(const Func &H, int n1, int k, double x, cplx *out) {
    const int s = H.s;
    const Integ &In = I(k);
    e_zero(s, out);
    const double t0 = n1 - k;
    for (int l = 0; l <= k; l++) {
        double w = In.poly_interpolation(0, l), t1 = 1.0;
        for (int p = 1; p <= k; p++) {
            t1 *= x - t0;
            w += t1 * In.poly_interpolation(p, l);
        }
        e_incr(s, out, w, H.ptr(n1 - k + l));
    }
}
static void func_extrapolate(int n, Func &H, int k) {
    const int s = H.s;
    Buf acc(s * s);
    if (k == 0) {
        e_set(s, H.ptr(n + 1), H.ptr(n));
        return;
    }
    const Integ &In = I(k);
    std::vector<double> p1(k + 1, 0.0);
    for (int m = 0; m <= k; m++)
        for (int l = 0; l <= k; l++) p1[m] += (1 - 2 * (l % 2)) * In.poly_interpolation(l, m);
    e_zero(s, acc.data());
    for (int l = 0; l <= k; l++) e_incr(s, acc.data(), p1[l], H.ptr(n - l));
    e_set(s, H.ptr(n + 1), acc.data());
}
// out = exp(-i dt A) for Hermitian A
static void expm_herm(int s, const cplx *A, double dt, cplx *out) {
    Buf a(A, A + s * s), V(s * s);
    std::vector<double> ev(s);
    herm_eig(s, a.data(), ev.data(), V.data());
    std::vector<cplx> ph(s);
    for (int i = 0; i < s; i++) ph[i] = std::exp(-kI * (dt * ev[i]));
    vdv(s, out, V.data(), ph.data());
}
// U(t) for t = 0..tlast (U.data() + t es), before the chemical-potential phase
static void td_propagators(Func &H, double h, int k, int order, bool fixHam, int tlast, Buf &U) {
    const int s = H.s;
    const long es = (long)s * s;
    U.assign((long)(tlast + 1) * es, cplx(0));
    e_diag(s, U.data(), cplx(1));
    Buf h1(es), h2(es), arg(es), e1(es), e2(es), tmp(es);
    const double a1 = (3.0 - 2.0 * std::sqrt(3.0)) / 12.0, a2 = (3.0 + 2.0 * std::sqrt(3.0)) / 12.0;
    for (int t = 1; t <= tlast; t++) {
        const int n1 = t <= k ? k : t;
        if (!fixHam && t > k) func_extrapolate(t - 1, H, k);
        if (order == 2) {
            func_interp(H, n1, k, t - 0.5, h1.data());
            expm_herm(s, h1.data(), h, e1.data());
            e_mult(s, U.data() + t * es, e1.data(), U.data() + (t - 1) * es);
        } else {
            func_interp(H, n1, k, t - 0.5 - std::sqrt(3.0) / 6.0, h1.data());
            func_interp(H, n1, k, t - 0.5 + std::sqrt(3.0) / 6.0, h2.data());
            for (long l = 0; l < es; l++) arg[l] = a1 * h1[l] + a2 * h2[l];
            expm_herm(s, arg.data(), h, e1.data());
            for (long l = 0; l < es; l++) arg[l] = a2 * h1[l] + a1 * h2[l];
            expm_herm(s, arg.data(), h, e2.data());
            e_mult(s, tmp.data(), e1.data(), e2.data());
            e_mult(s, U.data() + t * es, tmp.data(), U.data() + (t - 1) * es);
        }
    }
}
// tstp < -1: every component (cntr_equilibrium_impl.hpp:793-870); tstp >= -1: the timestep
// form (:955-1040), which fills timestep tstp only and, as the reference does, multiplies
// U(tstp) by the phase e^{i mu h n} for every n = 0..tstp and leaves U(n < tstp) unphased
static void green_from_H_td(Herm &G, double mu, Func &H, double beta, double h, int k, int order,
                            bool fixHam, int tstp) {
    const int s = G.s, ntau = G.ntau, sig = G.sig;
    const long es = G.es();
    const double dtau = beta / ntau;
    Buf Hmu(es), V(es), val(es), tmp(es), Ud(es);
    for (long l = 0; l < es; l++) Hmu[l] = -H.ptr(-1)[l];
    for (int i = 0; i < s; i++) Hmu[i * s + i] += mu;
    std::vector<double> ev(s), evm(s), d(s);
    herm_eig(s, Hmu.data(), ev.data(), V.data());
    for (int i = 0; i < s; i++) evm[i] = -ev[i];
    const bool all = tstp < -1;
    if (all || tstp == -1) {
        for (int m = 0; m <= ntau; m++) {
            const double tau = m * dtau;
            for (int i = 0; i < s; i++)
                d[i] = sig == -1 ? -fermi_exp(beta, tau, ev[i]) : bose_exp(beta, tau, ev[i]);
            vdv(s, G.mat(m), V.data(), d.data());
        }
        if (!all || G.nt < 0) return;
    }
    const int tlast = all ? G.nt : tstp;
    Buf U;
    td_propagators(H, h, k, order, fixHam, tlast, U);
    if (all) {
        for (int n = 0; n <= tlast; n++) e_smul(s, U.data() + n * es, std::exp(kI * (mu * h * n)));
    } else {
        for (int n = 0; n <= tlast; n++) e_smul(s, U.data() + tlast * es, std::exp(kI * (mu * h * n)));
    }
    const int n0 = all ? 0 : tlast;
    for (int m = 0; m <= ntau; m++) {
        const double tau = m * dtau;
        for (int i = 0; i < s; i++)
            d[i] = sig == -1 ? fermi_exp(beta, tau, evm[i]) : bose_exp(beta, tau, evm[i]);
        vdv(s, val.data(), V.data(), d.data());
        for (int n = n0; n <= tlast; n++) {
            e_mult(s, G.tv(n, m), U.data() + n * es, val.data());
            e_smul(s, G.tv(n, m), sig == -1 ? kI : -kI);
        }
    }
    for (int i = 0; i < s; i++) d[i] = sig == -1 ? fermi(beta, evm[i]) : -bose(beta, evm[i]);
    vdv(s, val.data(), V.data(), d.data());
    for (int m = n0; m <= tlast; m++)
        for (int n = 0; n <= m; n++) {
            e_herm(s, Ud.data(), U.data() + n * es);
            e_mult(s, G.ret(m, n), U.data() + m * es, Ud.data());
            e_smul(s, G.ret(m, n), -kI);
            e_herm(s, Ud.data(), U.data() + m * es);
            e_mult(s, tmp.data(), U.data() + n * es, val.data());
            e_mult(s, G.les(n, m), tmp.data(), Ud.data());
            e_smul(s, G.les(n, m), kI);
        }
}

// ---------------------------------------------------------------------------------
// Matsubara Fourier tools (cntr_matsubara_impl.hpp:70-242, fourier.cpp:43-90).
// The cubic DFT end corrections follow Numerical Recipes 3rd ed. sec. 13.9 (dftcor).
// ---------------------------------------------------------------------------------
static void dftcorr_cubic(double th, double *corfac, cplx *endcor) {
    double ar[4], ai[4];
    if (std::fabs(th) < 5.0e-2) {
        double t = th, t2 = t * t, t4 = t2 * t2, t6 = t4 * t2;
        *corfac = 1.0 - (11.0 / 720.0) * t4 + (23.0 / 15120.0) * t6;
        ar[0] = (-2.0 / 3.0) + t2 / 45.0 + (103.0 / 15120.0) * t4 - (169.0 / 226800.0) * t6;
        ar[1] = (7.0 / 24.0) - (7.0 / 180.0) * t2 + (5.0 / 3456.0) * t4 - (7.0 / 259200.0) * t6;
        ar[2] = (-1.0 / 6.0) + t2 / 45.0 - (5.0 / 6048.0) * t4 + t6 / 64800.0;
        ar[3] = (1.0 / 24.0) - t2 / 180.0 + (5.0 / 24192.0) * t4 - t6 / 259200.0;
        ai[0] = t * (2.0 / 45.0 + (2.0 / 105.0) * t2 - (8.0 / 2835.0) * t4 + (86.0 / 467775.0) * t6);
        ai[1] = t * (7.0 / 72.0 - t2 / 168.0 + (11.0 / 72576.0) * t4 - (13.0 / 5987520.0) * t6);
        ai[2] = t * (-7.0 / 90.0 + t2 / 210.0 - (11.0 / 90720.0) * t4 + (13.0 / 7484400.0) * t6);
        ai[3] = t * (7.0 / 360.0 - t2 / 840.0 + (11.0 / 362880.0) * t4 - (13.0 / 29937600.0) * t6);
    } else {
        double cth = std::cos(th), sth = std::sin(th);
        double ctth = cth * cth - sth * sth, stth = 2.0 * sth * cth;
        double th2 = th * th, th4 = th2 * th2, tmth2 = 3.0 - th2, spth2 = 6.0 + th2;
        double sth4i = 1.0 / (6.0 * th4), tth4i = 2.0 * sth4i;
        *corfac = tth4i * spth2 * (3.0 - 4.0 * cth + ctth);
        ar[0] = sth4i * (-42.0 + 5.0 * th2 + spth2 * (8.0 * cth - ctth));
        ai[0] = sth4i * (th * (-12.0 + 6.0 * th2) + spth2 * stth);
        ar[1] = sth4i * (14.0 * tmth2 - 7.0 * spth2 * cth);
        ai[1] = sth4i * (30.0 * th - 5.0 * spth2 * sth);
        ar[2] = tth4i * (-4.0 * tmth2 + 2.0 * spth2 * cth);
        ai[2] = tth4i * (-12.0 * th + 2.0 * spth2 * sth);
        ar[3] = sth4i * (2.0 * tmth2 - spth2 * cth);
        ai[3] = sth4i * (6.0 * th - spth2 * sth);
    }
    for (int j = 0; j < 4; j++) endcor[j] = cplx(ar[j], ai[j]);
}

// mdft[m] = sum_r G(r) exp(i pi (2m+one) r / ntau), m = 0..ntau
static void matsubara_dft(cplx *mdft, Herm &G, int sig) {
    int ntau = G.ntau, s = G.s;
    long es = G.es();
    double one = sig == -1 ? 1.0 : 0.0;
    for (int m = 0; m <= ntau; m++) {
        cplx *z = mdft + m * es;
        e_zero(s, z);
        for (int r = 0; r <= ntau; r++) {
            double arg = ((2 * m + one) * r * kPi) / ntau;
            cplx ef(std::cos(arg), std::sin(arg));
            for (long l = 0; l < es; l++) z[l] += G.mat(r)[l] * ef;
        }
    }
}

static void matsubara_ft(cplx *res, int m, Herm &G, const cplx *mdft, int sig, double beta) {
    int ntau = G.ntau;
    long es = G.es();
    double one = sig == -1 ? 1.0 : 0.0, dtau = beta / ntau, corfac;
    cplx bc[4];
    dftcorr_cubic((2 * m + one) * kPi / ntau, &corfac, bc);
    int m1 = m;
    while (m1 < 0) m1 += ntau;
    while (m1 >= ntau) m1 -= ntau;
    Buf z1(es, 0);
    for (int r = 0; r <= 3; r++) {
        for (long l = 0; l < es; l++) z1[l] += bc[r] * G.mat(r)[l];
        cplx cc = (1.0 * sig) * std::conj(bc[r]);
        for (long l = 0; l < es; l++) z1[l] += cc * G.mat(ntau - r)[l];
    }
    for (long l = 0; l < es; l++) res[l] = dtau * (z1[l] + mdft[m1 * es + l] * corfac);
}

static void first_order_tail(cplx *xmat, const cplx *coeff, double beta, long es, int ntau,
                             int sig) {
    for (int r = 0; r <= ntau; r++) {
        for (long l = 0; l < es; l++) xmat[r * es + l] = coeff[l] * (-0.5);
        if (sig == 1) {
            double tau = r * (beta / ntau);
            for (long l = 0; l < es; l++) xmat[r * es + l] += coeff[l] * (tau / beta);
        }
    }
}

// cntr_dyson_impl.hpp:355-447
static void dyson_mat_fourier(Herm &G, Herm &Sigma, double mu, const cplx *H0, double beta) {
    int ntau = G.ntau, s = G.s, sig = G.sig, pcf = 10, m2 = ntau / 2;
    long es = G.es();
    Buf sdft((long)(ntau + 1) * es), siomn(es), z1(es), z2(es), zinv(es), one(es), hj(es);
    Buf gmat((long)(ntau + 1) * es);
    e_diag(s, one.data(), 1.0);
    for (long l = 0; l < es; l++) hj[l] = H0[l] - mu * one[l];
    matsubara_dft(sdft.data(), Sigma, sig);
    first_order_tail(gmat.data(), one.data(), beta, es, ntau, sig);
    for (int m = -m2; m <= m2 - 1; m++)
        for (int p = -pcf; p <= pcf; p++) {
            int mm = m + p * ntau;
            cplx iomn(0, kPi / beta * (2 * mm + (sig == -1 ? 1 : 0)));
            matsubara_ft(siomn.data(), mm, Sigma, sdft.data(), sig, beta);
            for (long l = 0; l < es; l++) z1[l] = siomn[l] + hj[l];
            if (sig == 1 && mm == 0) {
                e_inverse(s, zinv.data(), z1.data());
                for (long l = 0; l < es; l++) z2[l] = -zinv[l];
            } else {
                for (long l = 0; l < es; l++) z2[l] = iomn * one[l] - z1[l];
                e_inverse(s, zinv.data(), z2.data());
                e_smul(s, zinv.data(), 1.0 / iomn);
                e_mult(s, z2.data(), zinv.data(), z1.data());
            }
            e_smul(s, z2.data(), 1.0 / beta);
            for (int r = 0; r <= ntau; r++) {
                cplx ef = std::exp(-(r * (beta / ntau)) * iomn);
                for (long l = 0; l < es; l++) gmat[r * es + l] += z2[l] * ef;
            }
        }
    std::copy(gmat.begin(), gmat.end(), G.mat(0));
}

// cntr_vie2_impl.hpp:304-456
static void vie2_mat_fourier(Herm &G, Herm &F, Herm &Q, double beta, int pcf) {
    int ntau = G.ntau, s = G.s, sig = G.sig, m2 = ntau / 2;
    long es = G.es();
    double dtau = beta / ntau, msone = sig == -1 ? 1.0 : 0.0;
    Buf xmat((long)(ntau + 1) * es), fdft((long)(ntau + 1) * es), qdft((long)(ntau + 1) * es);
    Buf qmasy(es), z1(es), z2(es), z3(es), zinv(es), fio(es), qio(es), qio1(es), one(es);
    e_diag(s, one.data(), 1.0);
    for (long l = 0; l < es; l++) qmasy[l] = -Q.mat(0)[l] + (double)sig * Q.mat(ntau)[l];
    first_order_tail(xmat.data(), qmasy.data(), beta, es, ntau, sig);
    matsubara_dft(fdft.data(), F, sig);
    matsubara_dft(qdft.data(), Q, sig);
    for (int p = -pcf; p <= pcf; p++) {
        int mlow = 0, mhigh = 0;
        if (sig == 1) {
            mhigh = p < 0 ? 0 : 1;
            mlow = p > 0 ? -1 : 0;
        }
        for (int m = -m2 - mlow; m <= m2 - 1 + mhigh; m++) {
            int mm = m + p * ntau;
            double omn = (2 * mm + msone) * kPi / (ntau * dtau);
            matsubara_ft(fio.data(), mm, F, fdft.data(), sig, beta);
            matsubara_ft(qio.data(), mm, Q, qdft.data(), sig, beta);
            e_set(s, z3.data(), qmasy.data());
            if (sig == 1 && mm == 0) e_smul(s, z3.data(), cplx(0.0, 0.0));
            else e_smul(s, z3.data(), cplx(0.0, -1.0 / omn));
            for (long l = 0; l < es; l++) qio1[l] = qio[l] - z3[l];
            for (long l = 0; l < es; l++) z1[l] = fio[l] + one[l];
            e_mult(s, z2.data(), fio.data(), z3.data());
            for (long l = 0; l < es; l++) z2[l] = qio1[l] - z2[l];
            e_inverse(s, zinv.data(), z1.data());
            e_mult(s, z1.data(), zinv.data(), z2.data());
            e_smul(s, z1.data(), 1.0 / beta);
            for (int r = 0; r <= ntau; r++) {
                double arg = omn * r * dtau;
                cplx ef(std::cos(arg), -std::sin(arg));
                for (long l = 0; l < es; l++) xmat[r * es + l] += z1[l] * ef;
            }
        }
    }
    std::copy(xmat.begin(), xmat.end(), G.mat(0));
}

// cntr_vie2_impl.hpp:459-482
static void vie2_mat_fixpoint(Herm &G, Herm &F, Herm &Q, double beta, const Integ &In,
                              int nfix) {
    int ntau = G.ntau, s = G.s;
    long es = G.es();
    vie2_mat_fourier(G, F, Q, beta, 5);
    std::vector<cplx> sq, sd;
    Herm Qn(-1, ntau, s, Q.sig, sq), dG(-1, ntau, s, G.sig, sd);
    for (int it = 0; it < nfix; it++) {
        convolution_matsubara(Qn, F, G, In, beta);
        for (int r = 0; r <= ntau; r++) {
            e_incr(s, Qn.mat(r), G.mat(r));
            e_incr(s, Qn.mat(r), -1.0, Q.mat(r));
        }
        vie2_mat_fourier(dG, F, Qn, beta, 5);
        for (long l = 0; l < (long)(ntau + 1) * es; l++) G.mat_[l] -= dG.mat_[l];
    }
}

// cntr_vie2_impl.hpp:485-602 (vie2_mat_steep_dispatch): conjugate gradients on
// A G = Q, A X = X + (F*X + X*Fcc)/2 (Matsubara convolutions), started from the Fourier
// solution with pcf = 3, in the inner product
//   <X, Y> = sum_r w_r sum_ab Re (X_r^dagger Y_r)_ab
// (every element of the block product, :518-531), w_r = omega_r for r <= k,
// omega_(ntau-r) for r >= ntau-k, 1 between.  No iteration unless <R,R> > maxerr; stop
// once sqrt(<R,R>) < maxerr (:596).
static double cg_dot(const Herm &X, const Herm &Y, const Integ &In) {
    int ntau = X.ntau, s = X.s, k = In.k;
    long es = X.es();
    Buf xd(es), t(es);
    double sum = 0.0;
    for (int r = 0; r <= ntau; r++) {
        double w = r <= k ? In.gregory_omega(r) : (r >= ntau - k ? In.gregory_omega(ntau - r) : 1.0);
        e_herm(s, xd.data(), X.mat(r));
        e_mult(s, t.data(), xd.data(), Y.mat(r));
        for (long l = 0; l < es; l++) sum += w * t[l].real();
    }
    return sum;
}

static void vie2_mat_steep(Herm &G, Herm &F, Herm &Fcc, Herm &Q, double beta, const Integ &In,
                           int maxiter, double maxerr) {
    int ntau = G.ntau, s = G.s;
    long n = (long)(ntau + 1) * G.es();
    vie2_mat_fourier(G, F, Q, beta, 3);
    if (maxiter <= 0) return;
    std::vector<cplx> sr, sp, sap, s1, s2;
    Herm R(-1, ntau, s, Q.sig, sr), P(-1, ntau, s, Q.sig, sp), AP(-1, ntau, s, Q.sig, sap);
    Herm C1(-1, ntau, s, Q.sig, s1), C2(-1, ntau, s, Q.sig, s2);
    convolution_matsubara(C1, F, G, In, beta);
    convolution_matsubara(C2, G, Fcc, In, beta);
    // R = -((C1 + C2)/2 + G - Q)  (:504-513)
    for (long l = 0; l < n; l++) {
        cplx r = (C1.mat_[l] + C2.mat_[l]) * 0.5;
        r += G.mat_[l];
        r += -1.0 * Q.mat_[l];
        R.mat_[l] = -r;
        P.mat_[l] = R.mat_[l];
    }
    double rsold = cg_dot(R, R, In);
    if (!(rsold > maxerr)) return;
    for (int iter = 0; iter < maxiter; iter++) {
        convolution_matsubara(C1, F, P, In, beta);
        convolution_matsubara(C2, P, Fcc, In, beta);
        for (long l = 0; l < n; l++) {
            AP.mat_[l] = (C1.mat_[l] + C2.mat_[l]) * 0.5;
            AP.mat_[l] += P.mat_[l];
        }
        double alpha = rsold / cg_dot(P, AP, In);
        for (long l = 0; l < n; l++) {
            G.mat_[l] += alpha * P.mat_[l];
            R.mat_[l] += -alpha * AP.mat_[l];
        }
        double rsnew = cg_dot(R, R, In);
        for (long l = 0; l < n; l++) P.mat_[l] = R.mat_[l] + (rsnew / rsold) * P.mat_[l];
        rsold = rsnew;
        if (std::sqrt(rsold) < maxerr) break;
    }
}

static void force_matsubara_hermitian(Herm &G) {
    int s = G.s;
    long es = G.es();
    Buf t(es);
    for (int m = 0; m <= G.ntau; m++) {
        e_herm(s, t.data(), G.mat(m));
        for (long l = 0; l < es; l++) G.mat(m)[l] = 0.5 * (G.mat(m)[l] + t[l]);
    }
}

// ---------------------------------------------------------------------------------
// Dyson equation (cntr_dyson_impl.hpp:79-1022)
// ---------------------------------------------------------------------------------
// cntr_dyson_impl.hpp:79-223
static void dyson_timestep_ret(int n, Herm &G, double mu, Func &H, Herm &S, const Integ &In,
                               double h) {
    int k = In.k, k1 = k + 1, k2 = 2 * k1, s = G.s;
    long es = G.es();
    Buf gt(k * es), qq((long)(n + 1) * es, 0), mm(k * k * es, 0), one(es), st(es), hj(es), tmp(es);
    std::vector<cplx> diffw(k1 + 1);
    e_diag(s, one.data(), 1.0);
    std::fill(G.ret(n, 0), G.ret(n, 0) + (long)(n + 1) * es, cplx(0));
    e_diag(s, G.ret(n, n), -kI);
    for (int j = 1; j <= k; j++) {
        int p = j - 1;
        for (int l = 0; l <= k; l++) {
            cplx cw = kI / h * In.poly_differentiation(j, l);
            if (l == 0) e_incr(s, qq.data() + p * es, -cw, G.ret(n, n));
            else e_incr(s, mm.data() + (p * k + l - 1) * es, cw, one.data());
        }
        for (long l = 0; l < es; l++) tmp[l] = -H.ptr(n - j)[l] + mu * one[l];
        e_incr(s, mm.data() + (p * k + p) * es, tmp.data());
        for (int l = 0; l <= k; l++) {
            double w = h * In.gregory_weights(j, l);
            if (n - l >= n - j) e_set(s, st.data(), S.ret(n - l, n - j));
            else {
                e_herm(s, st.data(), S.ret(n - j, n - l));
                w = -w;
            }
            if (l == 0) e_incr_mm(s, qq.data() + p * es, w, G.ret(n, n), st.data());
            else e_incr(s, mm.data() + (p * k + l - 1) * es, -w, st.data());
        }
    }
    solve_left(s, k, gt.data(), mm.data(), qq.data());
    for (int j = 1; j <= k; j++) e_set(s, G.ret(n, n - j), gt.data() + (j - 1) * es);
    std::fill(qq.begin(), qq.end(), cplx(0));
    for (int m = n - k; m <= n; m++) {
        cplx *gr = G.ret(n, m);
        for (int j = 0; j <= n - k2; j++)
            e_incr_mm(s, qq.data() + j * es, In.gregory_weights(n - j, n - m), gr, S.ret(m, j));
        for (int j = std::max(0, n - k2 + 1); j < n - k; j++)
            e_incr_mm(s, qq.data() + j * es, In.gregory_weights(n - j, n - m), gr, S.ret(m, j));
    }
    for (int p = 0; p <= k1; p++) diffw[p] = In.bd_weights(p) * kI / h;
    double w0 = h * In.gregory_omega(0);
    for (int l = k + 1; l <= n; l++) {
        int j = n - l;
        e_herm(s, hj.data(), H.ptr(j));
        cplx *q = qq.data() + j * es;
        for (long i = 0; i < es; i++) {
            q[i] *= h;
            for (int p = 1; p <= k1; p++) q[i] += -diffw[p] * G.ret(n, n - l + p)[i];
        }
        e_set(s, st.data(), S.ret(j, j));
        for (long i = 0; i < es; i++)
            mm[i] = diffw[0] * one[i] + mu * one[i] - hj[i] - w0 * st[i];
        solve_left(s, 1, G.ret(n, j), mm.data(), q);
        cplx *gr = G.ret(n, j);
        for (int j2 = 0; j2 < j; j2++)
            e_incr_mm(s, qq.data() + j2 * es, In.gregory_weights(n - j2, n - j), gr, S.ret(j, j2));
    }
}

// cntr_dyson_impl.hpp:257-344
static void dyson_start_ret(Herm &G, double mu, Func &H, Herm &S, const Integ &In, double h) {
    int k = In.k, s = G.s;
    long es = G.es();
    Buf qq(k * es), mm(k * k * es), one(es), gt(k * es), st(es);
    e_diag(s, one.data(), 1.0);
    cplx minusi(0, -1);
    for (int n = 0; n <= k; n++) e_diag(s, G.ret(n, n), cplx(0, -1.0));
    for (int j = 0; j < k; j++) {
        int dim = k - j;
        std::fill(qq.begin(), qq.end(), cplx(0));
        std::fill(mm.begin(), mm.end(), cplx(0));
        for (int n = j + 1; n <= k; n++) {
            int p = n - (j + 1);
            for (int l = 0; l <= k; l++) {
                int q = l - (j + 1);
                if (l <= j) {
                    e_mherm(s, gt.data(), G.ret(j, l));
                    e_zero(s, st.data());
                    e_incr_mm(s, st.data(), S.ret(n, l), gt.data());
                    for (long i = 0; i < es; i++)
                        qq[p * es + i] += minusi / h * In.poly_differentiation(n, l) * gt[i] +
                                          h * In.poly_integration(j, n, l) * st[i];
                } else {
                    cplx *mpq = mm.data() + (p * dim + q) * es;
                    for (long i = 0; i < es; i++)
                        mpq[i] = -minusi / h * In.poly_differentiation(n, l) * one[i];
                    if (l == n)
                        for (long i = 0; i < es; i++) mpq[i] += mu * one[i] - H.ptr(l)[i];
                    double w = h * In.poly_integration(j, n, l);
                    if (n >= l) e_set(s, st.data(), S.ret(n, l));
                    else {
                        e_herm(s, st.data(), S.ret(l, n));
                        w = -w;
                    }
                    for (long i = 0; i < es; i++) mpq[i] -= w * st[i];
                }
            }
        }
        solve_right(s, dim, gt.data(), mm.data(), qq.data());
        for (int n = j + 1; n <= k; n++) e_set(s, G.ret(n, j), gt.data() + (n - j - 1) * es);
    }
}

// cntr_dyson_impl.hpp:650-711
static void dyson_timestep_tv(int n, Herm &G, double mu, const cplx *Hn, Herm &S,
                              const Integ &In, double beta, double h) {
    int k = In.k, s = G.s, ntau = G.ntau;
    long es = G.es();
    Buf mm(es), qq(es), one(es);
    e_diag(s, one.data(), 1.0);
    std::fill(G.tv(n, 0), G.tv(n, 0) + (long)(ntau + 1) * es, cplx(0));
    Buf ctv((long)(ntau + 1) * es);
    convolution_timestep_tv(n, ctv.data(), S, S, G, In, beta, h);
    std::copy(ctv.begin(), ctv.end(), G.tv(n, 0));
    cplx ih(0, 1 / h);
    for (int m = n - k - 1; m < n; m++) {
        cplx cw = ih * In.bd_weights(n - m);
        cplx *g = G.tv(m, 0), *g1 = G.tv(n, 0);
        for (long l = 0; l < (long)(ntau + 1) * es; l++) g1[l] -= cw * g[l];
    }
    double w = -h * In.gregory_weights(n, 0);
    for (long l = 0; l < es; l++)
        mm[l] = ih * In.bd_weights(0) * one[l] + w * S.ret(n, n)[l] + mu * one[l] - Hn[l];
    for (int j = 0; j <= ntau; j++) {
        e_set(s, qq.data(), G.tv(n, j));
        solve_right(s, 1, G.tv(n, j), mm.data(), qq.data());
    }
}

// cntr_dyson_impl.hpp:746-847
static void dyson_start_tv(Herm &G, double mu, Func &H, Herm &S, const Integ &In, double beta,
                           double h) {
    int k = In.k, s = G.s, ntau = G.ntau, sig = G.sig;
    long es = G.es();
    double dtau = beta / ntau;
    Buf qq(k * es), mm(k * k * es), st(es), gt(k * es), one(es), tmp(es);
    e_diag(s, one.data(), 1.0);
    for (int m = 0; m <= ntau; m++)
        for (long l = 0; l < es; l++) G.tv(0, m)[l] = ((double)sig) * kI * G.mat(ntau - m)[l];
    for (int n = 1; n <= k; n++)
        for (int m = 0; m <= ntau; m++) {
            matsubara_integral_2(s, m, ntau, gt.data(), S.tv(n, 0), G.mat(0), In, G.sig);
            for (long l = 0; l < es; l++) G.tv(n, m)[l] = dtau * gt[l];
        }
    for (int m = 0; m <= ntau; m++) {
        std::fill(mm.begin(), mm.end(), cplx(0));
        std::fill(qq.begin(), qq.end(), cplx(0));
        for (int n = 1; n <= k; n++) {
            int p = n - 1;
            for (int j = 0; j <= k; j++) {
                cplx cw = kI / h * In.poly_differentiation(n, j);
                if (j == 0) {
                    for (long l = 0; l < es; l++) qq[p * es + l] -= cw * G.tv(0, m)[l];
                } else {
                    for (long l = 0; l < es; l++) mm[(p * k + j - 1) * es + l] += cw * one[l];
                }
            }
            for (long l = 0; l < es; l++) tmp[l] = -H.ptr(n)[l] + mu * one[l];
            e_incr(s, mm.data() + (p * k + p) * es, tmp.data());
            for (int j = 0; j <= k; j++) {
                cplx cw = h * In.gregory_weights(n, j);
                if (j == 0) {
                    e_incr_mm(s, qq.data() + p * es, cw, S.ret(n, 0), G.tv(0, m));
                } else {
                    if (n >= j) e_set(s, st.data(), S.ret(n, j));
                    else e_mherm(s, st.data(), S.ret(j, n));
                    for (long l = 0; l < es; l++) mm[(p * k + j - 1) * es + l] += -cw * st[l];
                }
            }
            e_incr(s, qq.data() + p * es, G.tv(n, m));
        }
        solve_right(s, k, gt.data(), mm.data(), qq.data());
        for (int n = 1; n <= k; n++) e_set(s, G.tv(n, m), gt.data() + (n - 1) * es);
    }
}

// cntr_dyson_impl.hpp:892-1022
static void dyson_timestep_les(int n, Herm &G, double mu, Func &H, Herm &S, const Integ &In,
                               double beta, double h) {
    int k = In.k, k1 = k + 1, s = G.s, n1 = std::max(n, k);
    long es = G.es();
    Buf qq(k * es), mm(k * k * es), gt(es), st(es), one(es), gles((long)(n1 + 1) * es, 0);
    e_diag(s, one.data(), 1.0);
    convolution_timestep_les_tvvt(n, gles.data(), S, G, In, beta);
    convolution_timestep_les_lesadv(n, gles.data(), S, S, G, G, In, h);
    e_mherm(s, gles.data(), G.tv(n, 0));
    std::fill(mm.begin(), mm.end(), cplx(0));
    std::fill(qq.begin(), qq.end(), cplx(0));
    for (int j = 1; j <= k; j++) {
        int p = j - 1;
        e_incr(s, qq.data() + p * es, gles.data() + j * es);
        for (int m = 0; m <= k; m++) {
            cplx cw = kI / h * In.poly_differentiation(j, m);
            if (m == 0) {
                for (long l = 0; l < es; l++) qq[p * es + l] -= cw * gles[l];
            } else {
                for (long l = 0; l < es; l++) mm[(p * k + m - 1) * es + l] += cw * one[l];
            }
        }
        for (long l = 0; l < es; l++) gt[l] = -H.ptr(j)[l] + mu * one[l];
        e_incr(s, mm.data() + (p * k + p) * es, gt.data());
        for (int m = 0; m <= k; m++) {
            cplx cw = h * In.gregory_weights(j, m);
            if (m == 0) {
                e_incr_mm(s, qq.data() + p * es, cw, S.ret(j, 0), gles.data());
            } else {
                if (j >= m) e_set(s, st.data(), S.ret(j, m));
                else e_mherm(s, st.data(), S.ret(m, j));
                for (long l = 0; l < es; l++) mm[(p * k + m - 1) * es + l] += -cw * st[l];
            }
        }
    }
    solve_right(s, k, gles.data() + es, mm.data(), qq.data());
    for (int j = k + 1; j <= n; j++) {
        cplx *q = qq.data();
        e_set(s, q, gles.data() + j * es);
        for (int p = 1; p <= k1; p++) {
            cplx cw = -kI / h * In.bd_weights(p);
            for (long l = 0; l < es; l++) q[l] += cw * gles[(j - p) * es + l];
        }
        for (long l = 0; l < es; l++)
            mm[l] = kI / h * In.bd_weights(0) * one[l] - H.ptr(j)[l] + mu * one[l] -
                    h * S.ret(j, j)[l] * In.gregory_weights(j, j);
        for (int m = 0; m < j; m++)
            e_incr_mm(s, q, h * In.gregory_weights(j, m), S.ret(j, m), gles.data() + m * es);
        solve_right(s, 1, gles.data() + j * es, mm.data(), q);
    }
    for (int j = 0; j <= n; j++) e_set(s, G.les(j, n), gles.data() + j * es);
}

// cntr_dyson_impl.hpp:1618-1637
static void dyson_timestep(int n, Herm &G, double mu, Func &H, Herm &S, const Integ &In,
                           double beta, double h) {
    dyson_timestep_ret(n, G, mu, H, S, In, h);
    dyson_timestep_tv(n, G, mu, H.ptr(n), S, In, beta, h);
    dyson_timestep_les(n, G, mu, H, S, In, beta, h);
}

// cntr_dyson_impl.hpp:1502-1518
static void dyson_start(Herm &G, double mu, Func &H, Herm &S, const Integ &In, double beta,
                        double h) {
    dyson_start_ret(G, mu, H, S, In, h);
    dyson_start_tv(G, mu, H, S, In, beta, h);
    for (int n = 0; n <= In.k; n++) dyson_timestep_les(n, G, mu, H, S, In, beta, h);
}

// cntr_dyson_impl.hpp:450-468 + 1393-1418
static void dyson_mat(Herm &G, double mu, Func &H, Herm &S, const Integ &In, double beta,
                      int method, bool force_herm) {
    int ntau = G.ntau, s = G.s;
    if (method == 0) {
        dyson_mat_fourier(G, S, mu, H.ptr(-1), beta);
    } else if (method == 1) {
        // cntr_dyson_impl.hpp:522-540, 1403-1406: F = -G0*Sigma, Fcc = -Sigma*G0
        std::vector<cplx> s0, s1, s2;
        Herm G0(-1, ntau, s, G.sig, s0), G0xS(-1, ntau, s, G.sig, s1), SxG0(-1, ntau, s, G.sig, s2);
        green_from_H_const(G0, mu, H.ptr(-1), beta, 1.0);
        convolution_matsubara(G0xS, G0, S, In, beta);
        convolution_matsubara(SxG0, S, G0, In, beta);
        for (auto &z : s1) z = -z;
        for (auto &z : s2) z = -z;
        vie2_mat_steep(G, G0xS, SxG0, G0, beta, In, 40, 1.0e-12);
    } else {
        std::vector<cplx> s0, s1;
        Herm G0(-1, ntau, s, G.sig, s0), G0xS(-1, ntau, s, G.sig, s1);
        green_from_H_const(G0, mu, H.ptr(-1), beta, 1.0);
        convolution_matsubara(G0xS, G0, S, In, beta);
        for (auto &z : s1) z = -z;
        vie2_mat_fixpoint(G, G0xS, G0, beta, In, 6);
    }
    if (force_herm) force_matsubara_hermitian(G);
}

// ---------------------------------------------------------------------------------
// VIE2 [1+F]G = Q (cntr_vie2_impl.hpp:59-999, 1597-1651)
// ---------------------------------------------------------------------------------
// cntr_vie2_impl.hpp:59-181 (called with F and Fcc swapped, :1615)
static void vie2_timestep_ret(int n, Herm &G, Herm &F, Herm &Fcc, Herm &Q, const Integ &In,
                              double h) {
    int k = In.k, k1 = k + 1, k2 = 2 * k1, s = G.s;
    long es = G.es();
    Buf gt(k * es), qq((long)(n + 1) * es, 0), mm(k * k * es, 0), one(es), st(es);
    e_diag(s, one.data(), 1.0);
    std::fill(G.ret(n, 0), G.ret(n, 0) + (long)(n + 1) * es, cplx(0));
    e_set(s, G.ret(n, n), Q.ret(n, n));
    for (int j = 1; j <= k; j++) {
        int p = j - 1;
        e_incr(s, qq.data() + p * es, Q.ret(n, n - j));
        e_incr(s, mm.data() + (p + k * p) * es, one.data());
        for (int l = 0; l <= k; l++) {
            double w = -h * In.gregory_weights(j, l);
            if (n - l >= n - j) e_set(s, st.data(), F.ret(n - l, n - j));
            else {
                e_herm(s, st.data(), Fcc.ret(n - j, n - l));
                w = -w;
            }
            if (l == 0) e_incr_mm(s, qq.data() + p * es, w, G.ret(n, n), st.data());
            else e_incr(s, mm.data() + (p * k + l - 1) * es, -w, st.data());
        }
    }
    solve_left(s, k, gt.data(), mm.data(), qq.data());
    for (int j = 1; j <= k; j++) e_set(s, G.ret(n, n - j), gt.data() + (j - 1) * es);
    std::fill(qq.begin(), qq.end(), cplx(0));
    for (int m = n - k; m <= n; m++) {
        cplx *gr = G.ret(n, m);
        for (int j = 0; j <= n - k2; j++)
            e_incr_mm(s, qq.data() + j * es, -In.gregory_weights(n - j, n - m), gr, F.ret(m, j));
        for (int j = std::max(0, n - k2 + 1); j < n - k; j++)
            e_incr_mm(s, qq.data() + j * es, -In.gregory_weights(n - j, n - m), gr, F.ret(m, j));
    }
    double w0 = -h * In.gregory_omega(0);
    for (int l = k + 1; l <= n; l++) {
        int j = n - l;
        cplx *q = qq.data() + j * es;
        for (long i = 0; i < es; i++) q[i] *= h;
        for (long i = 0; i < es; i++) mm[i] = one[i] - w0 * F.ret(j, j)[i];
        e_incr(s, q, Q.ret(n, j));
        solve_left(s, 1, G.ret(n, j), mm.data(), q);
        cplx *gr = G.ret(n, j);
        for (int j2 = 0; j2 < j; j2++)
            e_incr_mm(s, qq.data() + j2 * es, -In.gregory_weights(n - j2, n - j), gr, F.ret(j, j2));
    }
}

// cntr_vie2_impl.hpp:214-293
static void vie2_start_ret(Herm &G, Herm &F, Herm &Fcc, Herm &Q, const Integ &In, double h) {
    int k = In.k, k1 = k + 1, s = G.s;
    long es = G.es();
    Buf qq(k1 * es), mm(k1 * k1 * es), one(es), gt(k1 * es), st(es);
    e_diag(s, one.data(), 1.0);
    for (int n = 0; n <= k; n++) e_set(s, G.ret(n, n), Q.ret(n, n));
    for (int j = 0; j < k; j++) {
        int dim = k - j;
        std::fill(qq.begin(), qq.end(), cplx(0));
        std::fill(mm.begin(), mm.end(), cplx(0));
        for (int n = j + 1; n <= k; n++) {
            int p = n - (j + 1);
            e_incr(s, qq.data() + p * es, Q.ret(n, j));
            for (int l = 0; l <= k; l++) {
                int q = l - (j + 1);
                if (l <= j) {
                    e_mherm(s, gt.data(), G.ret(j, l));
                    e_zero(s, st.data());
                    e_incr_mm(s, st.data(), F.ret(n, l), gt.data());
                    for (long i = 0; i < es; i++)
                        qq[p * es + i] += -h * In.poly_integration(j, n, l) * st[i];
                } else {
                    cplx *mpq = mm.data() + (p * dim + q) * es;
                    e_zero(s, mpq);
                    if (l == n) e_incr(s, mpq, one.data());
                    double w = -h * In.poly_integration(j, n, l);
                    if (n >= l) e_set(s, st.data(), F.ret(n, l));
                    else {
                        e_herm(s, st.data(), Fcc.ret(l, n));
                        w = -w;
                    }
                    for (long i = 0; i < es; i++) mpq[i] -= w * st[i];
                }
            }
        }
        solve_right(s, dim, gt.data(), mm.data(), qq.data());
        for (int n = j + 1; n <= k; n++) e_set(s, G.ret(n, j), gt.data() + (n - j - 1) * es);
    }
}

// cntr_vie2_impl.hpp:644-691
static void vie2_timestep_tv(int n, Herm &G, Herm &F, Herm &Fcc, Herm &Q, const Integ &In,
                             double beta, double h) {
    int s = G.s, ntau = G.ntau;
    long es = G.es();
    Buf mm(es), qq(es), one(es);
    e_diag(s, one.data(), 1.0);
    std::fill(G.tv(n, 0), G.tv(n, 0) + (long)(ntau + 1) * es, cplx(0));
    Buf ctv((long)(ntau + 1) * es);
    convolution_timestep_tv(n, ctv.data(), F, Fcc, G, In, beta, h);
    std::copy(ctv.begin(), ctv.end(), G.tv(n, 0));
    double w = h * In.gregory_weights(n, 0);
    for (long l = 0; l < es; l++) mm[l] = one[l] + w * F.ret(n, n)[l];
    for (int j = 0; j <= ntau; j++) {
        for (long l = 0; l < es; l++) qq[l] = -G.tv(n, j)[l] + Q.tv(n, j)[l];
        solve_right(s, 1, G.tv(n, j), mm.data(), qq.data());
    }
}

// cntr_vie2_impl.hpp:726-802
static void vie2_start_tv(Herm &G, Herm &F, Herm &Fcc, Herm &Q, const Integ &In, double beta,
                          double h) {
    int k = In.k, k1 = k + 1, s = G.s, ntau = G.ntau;
    long es = G.es();
    double dtau = beta / ntau;
    Buf qq(k1 * es), mm(k1 * k1 * es), st(es), gt(k1 * es), one(es);
    e_diag(s, one.data(), 1.0);
    for (int n = 0; n <= k; n++)
        for (int m = 0; m <= ntau; m++) {
            matsubara_integral_2(s, m, ntau, gt.data(), F.tv(n, 0), G.mat(0), In, G.sig);
            for (long l = 0; l < es; l++) G.tv(n, m)[l] = -dtau * gt[l];
        }
    for (int m = 0; m <= ntau; m++) {
        std::fill(mm.begin(), mm.end(), cplx(0));
        std::fill(qq.begin(), qq.end(), cplx(0));
        for (int n = 0; n <= k; n++) {
            e_incr(s, mm.data() + (n * k1 + n) * es, one.data());
            for (int j = 0; j <= k; j++) {
                cplx cw = h * In.gregory_weights(n, j);
                if (n >= j) e_set(s, st.data(), F.ret(n, j));
                else e_mherm(s, st.data(), Fcc.ret(j, n));
                for (long l = 0; l < es; l++) mm[(n * k1 + j) * es + l] += cw * st[l];
            }
            e_incr(s, qq.data() + n * es, G.tv(n, m));
            e_incr(s, qq.data() + n * es, Q.tv(n, m));
        }
        solve_right(s, k1, gt.data(), mm.data(), qq.data());
        for (int n = 0; n <= k; n++) e_set(s, G.tv(n, m), gt.data() + n * es);
    }
}

// cntr_vie2_impl.hpp:846-957
static void vie2_timestep_les(int n, Herm &G, Herm &F, Herm &Fcc, Herm &Q, const Integ &In,
                              double beta, double h) {
    int k = In.k, k1 = k + 1, s = G.s, n1 = std::max(n, k);
    long es = G.es();
    Buf qq(k1 * es), mm(k1 * k1 * es), st(es), qt(es), one(es), gles((long)(n1 + 1) * es, 0);
    e_diag(s, one.data(), 1.0);
    convolution_timestep_les_tvvt(n, gles.data(), F, G, In, beta);
    convolution_timestep_les_lesadv(n, gles.data(), F, Fcc, G, G, In, h);
    for (auto &z : gles) z = -z;
    std::fill(mm.begin(), mm.end(), cplx(0));
    std::fill(qq.begin(), qq.end(), cplx(0));
    for (int j = 0; j <= k; j++) {
        int p = j;
        if (j <= n) e_incr(s, qq.data() + p * es, Q.les(j, n));
        else {
            e_mherm(s, qt.data(), Q.les(n, j));
            e_incr(s, qq.data() + p * es, qt.data());
        }
        e_incr(s, qq.data() + p * es, gles.data() + j * es);
        e_incr(s, mm.data() + (p * k1 + p) * es, one.data());
        for (int m = 0; m <= k; m++) {
            cplx cw = -h * In.gregory_weights(j, m);
            if (j >= m) e_set(s, st.data(), F.ret(j, m));
            else e_mherm(s, st.data(), Fcc.ret(m, j));
            for (long l = 0; l < es; l++) mm[(p * k1 + m) * es + l] += -cw * st[l];
        }
    }
    solve_right(s, k1, gles.data(), mm.data(), qq.data());
    for (int j = k + 1; j <= n; j++) {
        cplx *q = qq.data();
        e_set(s, q, gles.data() + j * es);
        e_incr(s, q, Q.les(j, n));
        for (long l = 0; l < es; l++)
            mm[l] = one[l] + h * F.ret(j, j)[l] * In.gregory_weights(j, j);
        for (int m = 0; m < j; m++)
            e_incr_mm(s, q, -h * In.gregory_weights(j, m), F.ret(j, m), gles.data() + m * es);
        solve_right(s, 1, gles.data() + j * es, mm.data(), q);
    }
    for (int j = 0; j <= n; j++) e_set(s, G.les(j, n), gles.data() + j * es);
}

static void vie2_start(Herm &G, Herm &F, Herm &Fcc, Herm &Q, const Integ &In, double beta,
                       double h) {
    vie2_start_ret(G, F, Fcc, Q, In, h);
    vie2_start_tv(G, F, Fcc, Q, In, beta, h);
    for (int n = 0; n <= In.k; n++) vie2_timestep_les(n, G, F, Fcc, Q, In, beta, h);
}

// ---------------------------------------------------------------------------------
// Bubbles (cntr_bubble_impl.hpp:23-128)
// ---------------------------------------------------------------------------------
static void bubble(int which, int tstp, Herm &C, int c1, int c2, Herm &A, Herm &Acc, int a1,
                   int a2, Herm &B, Herm &Bcc, int b1, int b2) {
    int ntau = C.ntau, sa = A.s, sb = B.s, sc = C.s, sigb = B.sig;
    int a12 = a1 * sa + a2, a21 = a2 * sa + a1, b12 = b1 * sb + b2, b21 = b2 * sb + b1;
    int c12 = c1 * sc + c2;
    if (tstp == -1) {
        for (int m = 0; m <= ntau; m++) {
            if (which == 1)
                C.mat(m)[c12] = (-1.0 * sigb) * A.mat(m)[a12] * B.mat(ntau - m)[b21];
            else
                C.mat(m)[c12] = -A.mat(m)[a12] * B.mat(m)[b12];
        }
        return;
    }
    for (int m = 0; m <= tstp; m++) {
        cplx aret = A.ret(tstp, m)[a12], accles_m = Acc.les(m, tstp)[a21];
        cplx bret = B.ret(tstp, m)[which == 1 ? b21 : b12];
        if (which == 1) {
            cplx bles21_tt1 = -std::conj(Bcc.les(m, tstp)[b12]);
            cplx bgtr21_tt1 = bret + bles21_tt1;
            cplx bles21_t1t = B.les(m, tstp)[b21];
            cplx bgtr21_t1t = -std::conj(Bcc.ret(tstp, m)[b12]) + bles21_t1t;
            cplx ales12_tt1 = -std::conj(accles_m);
            cplx agtr12_tt1 = aret + ales12_tt1;
            cplx cgtr = kI * agtr12_tt1 * bles21_t1t;
            cplx cles = kI * ales12_tt1 * bgtr21_t1t;
            C.ret(tstp, m)[c12] = cgtr - cles;
            C.les(m, tstp)[c12] = kI * A.les(m, tstp)[a12] * bgtr21_tt1;
        } else {
            cplx bgtr12 = bret - std::conj(Bcc.les(m, tstp)[b21]);
            cplx agtr12 = aret - std::conj(accles_m);
            cplx cgtr = kI * agtr12 * bgtr12;
            cplx cles = kI * std::conj(accles_m) * std::conj(Bcc.les(m, tstp)[b21]);
            C.ret(tstp, m)[c12] = cgtr - cles;
            C.les(m, tstp)[c12] = kI * A.les(m, tstp)[a12] * B.les(m, tstp)[b12];
        }
    }
    for (int m = 0; m <= ntau; m++) {
        if (which == 1) {
            cplx bvt21 = (-1.0 * sigb) * std::conj(Bcc.tv(tstp, ntau - m)[b12]);
            C.tv(tstp, m)[c12] = kI * A.tv(tstp, m)[a12] * bvt21;
        } else {
            C.tv(tstp, m)[c12] = kI * A.tv(tstp, m)[a12] * B.tv(tstp, m)[b12];
        }
    }
}

// ---------------------------------------------------------------------------------
// Utilities (cntr_utilities_impl.hpp:21-160, 606-640)
// ---------------------------------------------------------------------------------
static void extrapolate_timestep(int n, Herm &G, const Integ &In) {
    int ntau = G.ntau, k = In.k, s = G.s, n1 = n + 1;
    long es = G.es();
    Buf gt(es), g1(es);
    if (k == 0) {
        for (int m = 0; m <= ntau; m++) e_set(s, G.tv(n1, m), G.tv(n, m));
        for (int j = 0; j <= n; j++) e_set(s, G.ret(n1, n1 - j), G.ret(n, n - j));
        e_set(s, G.ret(n1, 0), G.ret(n, 0));
        for (int j = 1; j <= n1; j++) e_set(s, G.les(j, n1), G.les(j - 1, n));
        e_set(s, G.les(0, n1), G.les(0, n));
        return;
    }
    std::vector<double> p1(k + 1, 0.0);
    for (int m = 0; m <= k; m++)
        for (int l = 0; l <= k; l++) p1[m] += (1 - 2 * (l % 2)) * In.poly_interpolation(l, m);
    for (int m = 0; m <= ntau; m++) {
        e_zero(s, gt.data());
        for (int l = 0; l <= k; l++) e_incr(s, gt.data(), p1[l], G.tv(n - l, m));
        e_set(s, G.tv(n1, m), gt.data());
    }
    for (int j = 0; j <= n - k; j++) {
        e_zero(s, gt.data());
        for (int l = 0; l <= k; l++) e_incr(s, gt.data(), p1[l], G.ret(n - l, n - l - j));
        e_set(s, G.ret(n1, n1 - j), gt.data());
    }
    for (int j = 0; j <= k; j++) {
        e_zero(s, gt.data());
        for (int l = 0; l <= k; l++) {
            if (n - l <= j) e_mherm(s, g1.data(), G.ret(j, n - l));
            else e_set(s, g1.data(), G.ret(n - l, j));
            e_incr(s, gt.data(), p1[l], g1.data());
        }
        e_set(s, G.ret(n1, j), gt.data());
    }
    for (int j = 0; j <= k; j++) {
        e_zero(s, gt.data());
        for (int l = 0; l <= k; l++) {
            if (n - l < j) e_mherm(s, g1.data(), G.les(n - l, j));
            else e_set(s, g1.data(), G.les(j, n - l));
            e_incr(s, gt.data(), p1[l], g1.data());
        }
        e_set(s, G.les(j, n1), gt.data());
    }
    for (int j = k + 1; j <= n1; j++) {
        e_zero(s, gt.data());
        for (int l = 0; l <= k; l++) e_incr(s, gt.data(), p1[l], G.les(j - l - 1, n - l));
        e_set(s, G.les(j, n1), gt.data());
    }
}

static double fro(int s, const cplx *a, const cplx *b) {
    double r = 0;
    for (int i = 0; i < s * s; i++) r += std::norm(a[i] - b[i]);
    return std::sqrt(r);
}

static double distance_norm2(int tstp, Herm &g1, Herm &g2) {
    int s = g1.s;
    double err = 0;
    if (tstp == -1) {
        for (int i = 0; i <= g1.ntau; i++) err += fro(s, g1.mat(i), g2.mat(i));
        return err;
    }
    for (int i = 0; i <= tstp; i++) err += fro(s, g1.ret(tstp, i), g2.ret(tstp, i));
    for (int i = 0; i <= g1.ntau; i++) err += fro(s, g1.tv(tstp, i), g2.tv(tstp, i));
    for (int i = 0; i <= tstp; i++) err += fro(s, g1.les(i, tstp), g2.les(i, tstp));
    return err;
}

// ---------------------------------------------------------------------------------
// incr_convolution (cntr_convolution_impl.hpp:4163-4805): C += alpha * A*f*B for the
// output indices selected by mask (length ntau+1 for tstp=-1, else
// (tstp+1) ret + (ntau+1) tv + (tstp+1) les).
// ---------------------------------------------------------------------------------
static void incr_convolution_mat(const unsigned char *mask, cplx alpha, Herm &C, Herm &A,
                                 const cplx *f0, Herm &B, const Integ &In, double beta) {
    int ntau = A.ntau, s = A.s;
    long es = A.es();
    cplx adtau = alpha * beta * (1.0 / ntau);
    Buf ct(es), bm;
    const cplx *bt = B.mat(0);
    if (f0) {
        bm.resize((long)(ntau + 1) * es);
        for (int m = 0; m <= ntau; m++) e_mult(s, bm.data() + m * es, f0, B.mat(m));
        bt = bm.data();
    }
    for (int m = 0; m <= ntau; m++)
        if (mask[m]) {
            matsubara_integral_1(s, m, ntau, ct.data(), A.mat(0), bt, In, A.sig);
            e_incr(s, C.mat(m), adtau, ct.data());
        }
}

static void incr_convolution_ret(int tstp, const unsigned char *mask, cplx alpha, Herm &C,
                                 Herm &A, Herm &Acc, const cplx *ft, Herm &B, Herm &Bcc,
                                 const Integ &In, double h) {
    int s = A.s, K = In.k, n1 = std::max(tstp, K);
    long es = A.es();
    cplx adt = alpha * h;
    Buf ar((long)(n1 + 1) * es), at(es), bt(es), c0(es);
    for (int j = 0; j <= n1; j++) {
        if (j <= tstp) e_set(s, at.data(), A.ret(tstp, j));
        else e_mherm(s, at.data(), Acc.ret(j, tstp));
        if (ft) e_mult(s, ar.data() + j * es, at.data(), ft + j * es);
        else e_set(s, ar.data() + j * es, at.data());
    }
    for (int n = 0; n <= tstp; n++) {
        if (!mask[n]) continue;
        int n2 = tstp - n;
        e_zero(s, c0.data());
        if (tstp < K) {
            for (int j = 0; j <= K; j++) {
                double w = In.poly_integration(n, tstp, j);
                if (j >= n) e_incr_mm(s, c0.data(), w, ar.data() + j * es, B.ret(j, n));
                else {
                    e_mherm(s, bt.data(), Bcc.ret(n, j));
                    e_incr_mm(s, c0.data(), w, ar.data() + j * es, bt.data());
                }
            }
        } else if (n2 < K) {
            for (int l = 0; l <= K; l++) {
                int j = tstp - l;
                double w = In.gregory_weights(n2, l);
                if (j >= n) e_incr_mm(s, c0.data(), w, ar.data() + j * es, B.ret(j, n));
                else {
                    e_mherm(s, bt.data(), Bcc.ret(n, j));
                    e_incr_mm(s, c0.data(), w, ar.data() + j * es, bt.data());
                }
            }
        } else if (n2 <= 2 * K + 2) {
            for (int l = 0; l <= n2; l++) {
                int j = n + l;
                e_incr_mm(s, c0.data(), In.gregory_weights(n2, l), ar.data() + j * es, B.ret(j, n));
            }
        } else {
            for (int j = n; j <= n + K; j++)
                e_incr_mm(s, c0.data(), In.gregory_omega(j - n), ar.data() + j * es, B.ret(j, n));
            for (int j = n + K + 1; j < tstp - K; j++)
                e_incr_mm(s, c0.data(), ar.data() + j * es, B.ret(j, n));
            for (int l = 0; l <= K; l++) {
                int j = tstp - l;
                e_incr_mm(s, c0.data(), In.gregory_omega(l), ar.data() + j * es, B.ret(j, n));
            }
        }
        e_incr(s, C.ret(tstp, n), adt, c0.data());
    }
}

// incr_convolution_tv (cntr_convolution_impl.hpp:4396-4492): C^tv(tstp, m) +=
// alpha dtau [A^tv(tstp,.) o f0 B^M](m) + alpha h sum_n w A^R(tstp,n) f(n) B^tv(n, m)
static void incr_convolution_tv(int tstp, const unsigned char *mask, cplx alpha, Herm &C,
                                Herm &A, Herm &Acc, const cplx *f0, const cplx *ft, Herm &B,
                                const Integ &In, double beta, double h) {
    int ntau = A.ntau, s = A.s, K = In.k, n1 = std::max(tstp, K);
    long es = A.es();
    cplx adt = alpha * h, adtau = alpha * beta * (1.0 / ntau);
    Buf ar((long)(n1 + 1) * es), at(es), c1(es), c2(es), bm;
    for (int j = 0; j <= n1; j++) {
        if (j <= tstp) e_set(s, at.data(), A.ret(tstp, j));
        else e_mherm(s, at.data(), Acc.ret(j, tstp));
        if (ft) e_mult(s, ar.data() + j * es, at.data(), ft + j * es);
        else e_set(s, ar.data() + j * es, at.data());
    }
    const cplx *bmat = B.mat(0);
    if (f0) {
        bm.resize((long)(ntau + 1) * es);
        for (int m = 0; m <= ntau; m++) e_mult(s, bm.data() + m * es, f0, B.mat(m));
        bmat = bm.data();
    }
    for (int m = 0; m <= ntau; m++) {
        if (!mask[m]) continue;
        matsubara_integral_2(s, m, ntau, c1.data(), A.tv(tstp, 0), bmat, In, B.sig);
        e_smul(s, c1.data(), adtau);
        e_zero(s, c2.data());
        if (tstp <= 2 * K + 2) {
            for (int n = 0; n <= n1; n++)
                e_incr_mm(s, c2.data(), In.gregory_weights(tstp, n), ar.data() + n * es, B.tv(n, m));
        } else {
            for (int n = 0; n <= K; n++)
                e_incr_mm(s, c2.data(), In.gregory_omega(n), ar.data() + n * es, B.tv(n, m));
            for (int n = K + 1; n < tstp - K; n++)
                e_incr_mm(s, c2.data(), ar.data() + n * es, B.tv(n, m));
            for (int n = tstp - K; n <= tstp; n++)
                e_incr_mm(s, c2.data(), In.gregory_omega(tstp - n), ar.data() + n * es, B.tv(n, m));
        }
        e_smul(s, c2.data(), adt);
        e_incr(s, C.tv(tstp, m), c1.data());
        e_incr(s, C.tv(tstp, m), c2.data());
    }
}

// incr_convolution_les (cntr_convolution_impl.hpp:4539-4748): C^<(n, tstp) +=
// alpha h [A^R f B^<] + alpha h [A^< f B^A] - i alpha dtau [A^tv f0 B^vt], n = 0..tstp
static void incr_convolution_les(int tstp, const unsigned char *mask, cplx alpha, Herm &C,
                                 Herm &A, Herm &Acc, const cplx *f0, const cplx *ft, Herm &B,
                                 Herm &Bcc, const Integ &In, double beta, double h) {
    int ntau = A.ntau, s = A.s, K = In.k, n1 = std::max(tstp, K);
    long es = A.es();
    cplx adt = alpha * h, adtau = alpha * beta * (1.0 / ntau);
    Buf badv((long)(n1 + 1) * es), bles((long)(n1 + 1) * es), bvt((long)(ntau + 1) * es);
    Buf bt(es), at(es), c1(es), c2(es), c3(es);
    // badv(j) = f(j) B^A(j, tstp), bles(j) = f(j) B^<(j, tstp), bvt(m) = f0 B^vt(m, tstp)
    for (int j = 0; j <= n1; j++) {
        cplx *ba = badv.data() + j * es, *bl = bles.data() + j * es;
        if (j <= tstp) {
            e_herm(s, ba, Bcc.ret(tstp, j));
            e_set(s, bl, B.les(j, tstp));
        } else {
            e_set(s, ba, B.ret(j, tstp));
            e_smul(s, ba, cplx(-1, 0));
            e_mherm(s, bl, Bcc.les(tstp, j));
        }
        if (ft) {
            e_mult(s, bt.data(), ft + j * es, ba);
            e_set(s, ba, bt.data());
            e_mult(s, bt.data(), ft + j * es, bl);
            e_set(s, bl, bt.data());
        }
    }
    for (int m = 0; m <= ntau; m++) {
        cplx *bv = bvt.data() + m * es;
        if (B.sig == -1) e_herm(s, bv, Bcc.tv(tstp, ntau - m));
        else e_mherm(s, bv, Bcc.tv(tstp, ntau - m));
        if (f0) {
            e_mult(s, bt.data(), f0, bv);
            e_set(s, bv, bt.data());
        }
    }
    for (int n = 0; n <= tstp; n++) {
        if (!mask[n]) continue;
        // int_0^n dj A^R(n, j) bles(j)
        e_zero(s, c1.data());
        if (n <= 2 * K + 2) {
            int nup = std::max(n, K);
            for (int j = 0; j <= nup; j++) {
                double w = In.gregory_weights(n, j);
                if (j <= n) e_incr_mm(s, c1.data(), w, A.ret(n, j), bles.data() + j * es);
                else {
                    e_mherm(s, at.data(), Acc.ret(j, n));
                    e_incr_mm(s, c1.data(), w, at.data(), bles.data() + j * es);
                }
            }
        } else {
            for (int j = 0; j <= K; j++)
                e_incr_mm(s, c1.data(), In.gregory_omega(j), A.ret(n, j), bles.data() + j * es);
            for (int j = K + 1; j < n - K; j++)
                e_incr_mm(s, c1.data(), A.ret(n, j), bles.data() + j * es);
            for (int j = n - K; j <= n; j++)
                e_incr_mm(s, c1.data(), In.gregory_omega(n - j), A.ret(n, j), bles.data() + j * es);
        }
        // int_0^beta dm A^tv(n, m) bvt(m)
        e_zero(s, c2.data());
        for (int m = 0; m <= K; m++)
            e_incr_mm(s, c2.data(), In.gregory_omega(m), A.tv(n, m), bvt.data() + m * es);
        for (int m = K + 1; m < ntau - K; m++) e_incr_mm(s, c2.data(), A.tv(n, m), bvt.data() + m * es);
        for (int m = ntau - K; m <= ntau; m++)
            e_incr_mm(s, c2.data(), In.gregory_omega(ntau - m), A.tv(n, m), bvt.data() + m * es);
        // int_0^tstp dj A^<(n, j) badv(j); A^<(n, j) = -A^<(j, n)^dagger for j < n
        e_zero(s, c3.data());
        // the reference takes j == n from Acc in the short-time branch, from A otherwise
        const int jc = (tstp <= 2 * K + 2) ? n + 1 : n;
        auto ales = [&](int j) -> const cplx * {
            if (j < jc) {
                e_mherm(s, at.data(), Acc.les(j, n));
                return at.data();
            }
            return A.les(n, j);
        };
        if (tstp <= 2 * K + 2) {
            for (int j = 0; j <= n1; j++) {
                const cplx *a = ales(j);
                e_incr_mm(s, c3.data(), In.gregory_weights(tstp, j), a, badv.data() + j * es);
            }
        } else {
            for (int j = 0; j <= K; j++) {
                const cplx *a = ales(j);
                e_incr_mm(s, c3.data(), In.gregory_omega(j), a, badv.data() + j * es);
            }
            for (int j = K + 1; j < tstp - K; j++) {
                const cplx *a = ales(j);
                e_incr_mm(s, c3.data(), a, badv.data() + j * es);
            }
            for (int j = tstp - K; j <= tstp; j++) {
                const cplx *a = ales(j);
                e_incr_mm(s, c3.data(), In.gregory_omega(tstp - j), a, badv.data() + j * es);
            }
        }
        e_incr(s, C.les(n, tstp), adt, c1.data());
        e_incr(s, C.les(n, tstp), adt, c3.data());
        e_incr(s, C.les(n, tstp), adtau * cplx(0.0, -1.0), c2.data());
    }
}

}  // namespace

// =================================================================================
// C ABI
// =================================================================================
extern "C" {

int orc_integrator_tables(int k, double *greg, double *omega, double *bd, double *P, double *D,
                          double *W, double *rcorr) {
    if (k < 0 || k > 5) return -1;
    const Integ &In = I(k);
    int k1 = k + 1;
    if (greg) std::copy(In.greg.begin(), In.greg.end(), greg);
    if (omega) std::copy(In.omega.begin(), In.omega.end(), omega);
    if (bd) std::copy(In.bd.begin(), In.bd.end(), bd);
    if (P) std::copy(In.P.begin(), In.P.end(), P);
    if (D) std::copy(In.D.begin(), In.D.end(), D);
    if (W) std::copy(In.W.begin(), In.W.end(), W);
    if (rcorr && k > 1) std::copy(In.rc.begin(), In.rc.end(), rcorr);
    (void)k1;
    return 0;
}

double orc_gregory_weight(int k, int n, int j) { return I(k).gregory_weights(n, j); }

int orc_convolution_matsubara(orc_herm *C, orc_herm *A, orc_herm *B, double beta, int k) {
    Herm c(C), a(A), b(B);
    convolution_matsubara(c, a, b, I(k), beta);
    return 0;
}

int orc_convolution_timestep(int n, orc_herm *C, orc_herm *A, orc_herm *Acc, orc_herm *B,
                             orc_herm *Bcc, double beta, double h, int k) {
    Herm c(C), a(A), acc(Acc), b(B), bcc(Bcc);
    convolution_timestep(n, c, a, acc, b, bcc, I(k), beta, h);
    return 0;
}

int orc_incr_convolution(int tstp, const unsigned char *mask, double are, double aim,
                         orc_herm *C, orc_herm *A, orc_herm *Acc, const double *f0,
                         const double *ft, orc_herm *B, orc_herm *Bcc, double beta, double h,
                         int k) {
    Herm c(C), a(A), acc(Acc), b(B), bcc(Bcc);
    cplx alpha(are, aim);
    const Integ &In = I(k);
    const int ntau = a.ntau;
    // NULL mask: every output, as the public incr_convolution (:4786-4805)
    std::vector<unsigned char> all;
    if (!mask) {
        all.assign(tstp == -1 ? ntau + 1 : 2 * (tstp + 1) + ntau + 1, 1);
        mask = all.data();
    }
    if (tstp == -1) {
        incr_convolution_mat(mask, alpha, c, a, (const cplx *)f0, b, In, beta);
        return 0;
    }
    if (tstp < 0) return -1;
    incr_convolution_ret(tstp, mask, alpha, c, a, acc, (const cplx *)ft, b, bcc, In, h);
    incr_convolution_tv(tstp, mask + tstp + 1, alpha, c, a, acc, (const cplx *)f0,
                        (const cplx *)ft, b, In, beta, h);
    incr_convolution_les(tstp, mask + tstp + 1 + ntau + 1, alpha, c, a, acc, (const cplx *)f0,
                         (const cplx *)ft, b, bcc, In, beta, h);
    return 0;
}

// One propagation request of the batching engine, restated serially: Sigma = V^2 g with
// g the free GF of the constant bath Hamiltonian (green_from_H,
// cntr_equilibrium_impl.hpp:873-953), then the reference's dyson(...) sequence
// (cntr_dyson_impl.hpp:1725-1734): dyson_mat (force_hermitian), dyson_start and
// dyson_timestep for n = k+1..nt_stop.  Sigma must have all components allocated.
int orc_dyson_solve(orc_herm *G, orc_herm *Sigma, double mu, orc_func *H, const double *Hbath,
                    double V, double beta, double h, int k, int method, int nt_stop) {
    if (method != 0 && method != 2) return -2;
    Herm g(G), s(Sigma);
    Func f(H);
    green_from_H_const(s, mu, (const cplx *)Hbath, beta, h);
    const long es = (long)s.s * s.s;
    const long ntri = (long)(s.nt + 1) * (s.nt + 2) / 2, ntv = (long)(s.nt + 1) * (s.ntau + 1);
    const double v2 = V * V;
    for (long i = 0; i < ntri * es; i++) {
        s.ret_[i] *= v2;
        s.les_[i] *= v2;
    }
    for (long i = 0; i < ntv * es; i++) s.tv_[i] *= v2;
    for (long i = 0; i < (long)(s.ntau + 1) * es; i++) s.mat_[i] *= v2;
    const Integ &In = I(k);
    dyson_mat(g, mu, f, s, In, beta, method, true);
    dyson_start(g, mu, f, s, In, beta, h);
    for (int n = k + 1; n <= nt_stop; n++) dyson_timestep(n, g, mu, f, s, In, beta, h);
    return 0;
}
int orc_dyson_mat(orc_herm *G, double mu, orc_func *H, orc_herm *Sigma, double beta, int k,
                  int method, int force_hermitian) {
    if (method < 0 || method > 2) return -2;
    Herm g(G), s(Sigma);
    Func f(H);
    dyson_mat(g, mu, f, s, I(k), beta, method, force_hermitian != 0);
    return 0;
}

int orc_dyson_start(orc_herm *G, double mu, orc_func *H, orc_herm *Sigma, double beta,
                    double h, int k) {
    Herm g(G), s(Sigma);
    Func f(H);
    dyson_start(g, mu, f, s, I(k), beta, h);
    return 0;
}

int orc_dyson_timestep(int n, orc_herm *G, double mu, orc_func *H, orc_herm *Sigma,
                       double beta, double h, int k) {
    if (n <= k) return -1;
    Herm g(G), s(Sigma);
    Func f(H);
    dyson_timestep(n, g, mu, f, s, I(k), beta, h);
    return 0;
}

int orc_vie2_mat(orc_herm *G, orc_herm *F, orc_herm *Fcc, orc_herm *Q, double beta, int k,
                 int method) {
    if (method < 0 || method > 2) return -2;
    Herm g(G), f(F), fcc(Fcc), q(Q);
    if (method == 0) vie2_mat_fourier(g, f, q, beta, 20);
    else if (method == 1) vie2_mat_steep(g, f, fcc, q, beta, I(k), 40, 1.0e-12);  // :1240-1243
    else vie2_mat_fixpoint(g, f, q, beta, I(k), 6);
    return 0;
}

int orc_vie2_start(orc_herm *G, orc_herm *F, orc_herm *Fcc, orc_herm *Q, double beta, double h,
                   int k) {
    Herm g(G), f(F), fcc(Fcc), q(Q);
    vie2_start(g, f, fcc, q, I(k), beta, h);
    return 0;
}

int orc_vie2_timestep(int n, orc_herm *G, orc_herm *F, orc_herm *Fcc, orc_herm *Q, double beta,
                      double h, int k, int method) {
    Herm g(G), f(F), fcc(Fcc), q(Q);
    if (n == -1) return orc_vie2_mat(G, F, Fcc, Q, beta, k, method);
    if (n <= k) {
        vie2_start(g, f, fcc, q, I(n), beta, h);
        return 0;
    }
    const Integ &In = I(k);
    vie2_timestep_ret(n, g, fcc, f, q, In, h);
    vie2_timestep_tv(n, g, f, fcc, q, In, beta, h);
    vie2_timestep_les(n, g, f, fcc, q, In, beta, h);
    return 0;
}

int orc_bubble1(int tstp, orc_herm *C, int c1, int c2, orc_herm *A, orc_herm *Acc, int a1,
                int a2, orc_herm *B, orc_herm *Bcc, int b1, int b2) {
    Herm c(C), a(A), acc(Acc), b(B), bcc(Bcc);
    bubble(1, tstp, c, c1, c2, a, acc, a1, a2, b, bcc, b1, b2);
    return 0;
}

int orc_bubble2(int tstp, orc_herm *C, int c1, int c2, orc_herm *A, orc_herm *Acc, int a1,
                int a2, orc_herm *B, orc_herm *Bcc, int b1, int b2) {
    Herm c(C), a(A), acc(Acc), b(B), bcc(Bcc);
    bubble(2, tstp, c, c1, c2, a, acc, a1, a2, b, bcc, b1, b2);
    return 0;
}

int orc_extrapolate_timestep(int n, orc_herm *G, int k) {
    Herm g(G);
    if (n + 1 > g.nt || n < k) return -1;
    extrapolate_timestep(n, g, I(k));
    return 0;
}

double orc_distance_norm2(int tstp, orc_herm *g1, orc_herm *g2) {
    Herm a(g1), b(g2);
    return distance_norm2(tstp, a, b);
}

int orc_green_from_H_const(orc_herm *G, double mu, const double *H0, double beta, double h) {
    Herm g(G);
    green_from_H_const(g, mu, (const cplx *)H0, beta, h);
    return 0;
}

int orc_green_from_H_td(orc_herm *G, double mu, orc_func *H, double beta, double h, int SolveOrder,
                        int cf_order, int fixHam, int tstp) {
    if (cf_order != 2 && cf_order != 4) return -1;
    Herm g(G);
    Func f(H);
    green_from_H_td(g, mu, f, beta, h, SolveOrder, cf_order, fixHam != 0, tstp);
    return 0;
}

int orc_force_matsubara_hermitian(orc_herm *G) {
    Herm g(G);
    force_matsubara_hermitian(g);
    return 0;
}

// Self-consistent second-order lattice run (the k-batched 1D chain workload):
// per timestep, extrapolate every G_k, then iterate
//   G_loc = (1/Nk) sum_k G_k;  P = -Bubble1(G_loc,G_loc);  P <- U P U;
//   Sigma = Bubble2(G_loc, P);  dyson_timestep for every k (OpenMP over k)
// until distance_norm2(G_loc_new, G_loc_old) < tol.  U(-1)=0 so the Matsubara and
// mixed self-energies vanish and the k-resolved equilibrium state is free.
// Mirrors examples/programs/hubbard_chain_2b.cpp:310-322 with G -> G_loc.
int orc_lattice_run(int nk, const double *epsk, double U, double beta, int ntau, int nt,
                    double h, int k, int nt_stop, double tol, int maxiter, int nthreads,
                    int setup_threads, double *out_density, int *out_iters, double *out_seconds,
                    orc_herm *gloc_out, int nk_out, const int *kidx, orc_herm *gk_out) {
    for (int i = 0; i < nk_out; i++)
        if (!kidx || !gk_out || kidx[i] < 0 || kidx[i] >= nk) return -1;
#ifdef _OPENMP
    omp_set_num_threads(setup_threads > 0 ? setup_threads : (nthreads > 0 ? nthreads : 1));
#endif
    const Integ &In = I(k);
    std::vector<std::vector<cplx>> store(nk);
    std::vector<Herm> G;
    G.reserve(nk);
    for (int q = 0; q < nk; q++) G.emplace_back(nt, ntau, 1, -1, store[q]);
    std::vector<cplx> sl, sloc, sp, sold;
    Herm Gloc(nt, ntau, 1, -1, sl), Sig(nt, ntau, 1, -1, sloc), Pol(nt, ntau, 1, 1, sp);
    Herm Gold(nt, ntau, 1, -1, sold);
    std::vector<std::vector<cplx>> hstore(nk);
    std::vector<orc_func> hf(nk);
    for (int q = 0; q < nk; q++) {
        hstore[q].assign(nt + 2, cplx(epsk[q], 0));
        hf[q].nt = nt;
        hf[q].size1 = 1;
        hf[q].data = (double *)hstore[q].data();
    }
    auto t0 = std::chrono::steady_clock::now();
    auto gloc_step = [&](int t) {
        long nr = t + 1;
        for (long i = 0; i < nr; i++) {
            cplx a = 0, b = 0;
            for (int q = 0; q < nk; q++) {
                a += G[q].ret(t, i)[0];
                b += G[q].les(i, t)[0];
            }
            Gloc.ret(t, i)[0] = a / (double)nk;
            Gloc.les(i, t)[0] = b / (double)nk;
        }
        for (int m = 0; m <= ntau; m++) {
            cplx a = 0;
            for (int q = 0; q < nk; q++) a += G[q].tv(t, m)[0];
            Gloc.tv(t, m)[0] = a / (double)nk;
        }
    };
    auto sigma_step = [&](int t) {
        bubble(1, t, Pol, 0, 0, Gloc, Gloc, 0, 0, Gloc, Gloc, 0, 0);
        for (int i = 0; i <= t; i++) {
            Pol.ret(t, i)[0] *= -U * U;
            Pol.les(i, t)[0] *= -U * U;
        }
        for (int m = 0; m <= ntau; m++) Pol.tv(t, m)[0] *= 0.0;  // U(-1) = 0
        bubble(2, t, Sig, 0, 0, Gloc, Gloc, 0, 0, Pol, Pol, 0, 0);
    };
    auto copy_step = [&](Herm &dst, Herm &src, int t) {
        for (int i = 0; i <= t; i++) {
            dst.ret(t, i)[0] = src.ret(t, i)[0];
            dst.les(i, t)[0] = src.les(i, t)[0];
        }
        for (int m = 0; m <= ntau; m++) dst.tv(t, m)[0] = src.tv(t, m)[0];
    };
    // equilibrium: Sigma^M = 0 -> free Matsubara function per k
#pragma omp parallel for schedule(static)
    for (int q = 0; q < nk; q++) {
        Func f(&hf[q]);
        dyson_mat(G[q], 0.0, f, Sig, In, beta, 2, true);
    }
    // bootstrap 0..k
    int total_iters = 0;
    for (int it = 0; it < maxiter; it++) {
        for (int t = 0; t <= k; t++) sigma_step(t);
#pragma omp parallel for schedule(static)
        for (int q = 0; q < nk; q++) {
            Func f(&hf[q]);
            dyson_start(G[q], 0.0, f, Sig, In, beta, h);
        }
        double err = 0;
        for (int t = 0; t <= k; t++) {
            copy_step(Gold, Gloc, t);
            gloc_step(t);
            err += distance_norm2(t, Gloc, Gold);
        }
        total_iters++;
        if (err < tol && it > 0) break;
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#endif
    auto ts = std::chrono::steady_clock::now();  // stepping only (setup excluded)
    for (int t = k + 1; t <= nt_stop; t++) {
#pragma omp parallel for schedule(static)
        for (int q = 0; q < nk; q++) extrapolate_timestep(t - 1, G[q], In);
        gloc_step(t);
        for (int it = 0; it < maxiter; it++) {
            sigma_step(t);
#pragma omp parallel for schedule(static)
            for (int q = 0; q < nk; q++) {
                Func f(&hf[q]);
                dyson_timestep(t, G[q], 0.0, f, Sig, In, beta, h);
            }
            copy_step(Gold, Gloc, t);
            gloc_step(t);
            total_iters++;
            if (distance_norm2(t, Gloc, Gold) < tol) break;
        }
    }
    auto t1 = std::chrono::steady_clock::now();
    if (out_seconds) *out_seconds = std::chrono::duration<double>(t1 - ts).count();
    (void)t0;
    if (out_iters) *out_iters = total_iters;
    if (out_density)
        for (int t = 0; t <= nt_stop; t++) out_density[t] = (kI * -1.0 * Gloc.les(t, t)[0]).real();
    const long ntri = (long)(nt + 1) * (nt + 2) / 2, ntv = (long)(nt + 1) * (ntau + 1);
    auto copy_out = [&](Herm &src, orc_herm *dst) {
        Herm o(dst);
        std::copy(src.ret_, src.ret_ + ntri, o.ret_);
        std::copy(src.les_, src.les_ + ntri, o.les_);
        std::copy(src.tv_, src.tv_ + ntv, o.tv_);
        std::copy(src.mat_, src.mat_ + ntau + 1, o.mat_);
    };
    if (gloc_out) copy_out(Gloc, gloc_out);
    for (int i = 0; i < nk_out; i++) copy_out(G[kidx[i]], &gk_out[i]);
    return 0;
}

}  // extern "C"
