/* C++ drop-in for the reference's cntr:: API on the Kadanoff-Baym path, over the C ABI
 * of libcntrx.so (cntrx.h).  Class and function names, argument order and defaults follow
 * libcntr: herm_matrix (cntr/cntr_herm_matrix_decl.hpp:44-301), function
 * (cntr/cntr_function_decl.hpp:26-107), dyson_* (cntr/cntr_dyson_decl.hpp:40-93), vie2_*
 * (cntr/cntr_vie2_decl.hpp:20-76), convolution_timestep (cntr/cntr_convolution_decl.hpp:
 * 31-66), Bubble1/Bubble2 (cntr/cntr_bubble_decl.hpp:62-301), extrapolate_timestep and
 * distance_norm2 (cntr/cntr_utilities_decl.hpp), green_from_H for constant H
 * (cntr/cntr_equilibrium_decl.hpp:73), and integration::I<T>(k)
 * (cntr/integration.hpp:506-514).  A caller includes this header in place of "cntr.hpp"
 * and links -lcntrx; no HIP header is needed.
 *
 * Storage.  Every herm_matrix and function keeps the reference's packed layout twice: in
 * host memory, which the accessors return, and in device memory, which the solvers use.
 *   - Host writes mark their timestep (ret row i, les column j, tv row i, mat = -1) as
 *     newer on the host.  The next solver call that uses the object uploads those
 *     timesteps first.  A pointer from retptr(i,j), lesptr(i,j), tvptr(i,m) or matptr(m)
 *     may be used to write within that timestep's strip; for other writes through raw
 *     pointers call mark_host_modified().
 *   - A solver call copies what it wrote (one timestep, or 0..SolveOrder for the start
 *     phase, or mat) back to the host before it returns.
 * Errors: the reference asserts or aborts; here a failing libcntrx call prints
 * cntrx_last_error() and aborts.
 * Not provided (outside the path): Eigen types (cdmatrix below is a small dense matrix
 * with the operations the callers use), HDF5/text I/O, MPI, the _omp variants,
 * time-dependent green_from_H, herm_matrix_timestep. */
#ifndef CNTR_AMD_HPP
#define CNTR_AMD_HPP

#include <complex>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <utility>
#include <vector>

#include "cntrx.h"

#ifndef MAX_SOLVE_ORDER
#define MAX_SOLVE_ORDER 5
#endif
#ifndef CNTR_MAT_FOURIER
#define CNTR_MAT_FOURIER CNTRX_MAT_FOURIER
#define CNTR_MAT_CG CNTRX_MAT_CG
#define CNTR_MAT_FIXPOINT CNTRX_MAT_FIXPOINT
#endif
#ifndef FERMION
#define FERMION -1
#define BOSON 1
#endif

/* Dense complex matrix, row-major: the subset of Eigen::MatrixXcd (the reference's
 * cdmatrix, cntr/eigen_typedef.h) that the callers of the path use.  data() is row-major,
 * where Eigen's default is column-major. */
class cdmatrix {
  public:
    typedef std::complex<double> Scalar;
    cdmatrix() : r_(0), c_(0) {}
    cdmatrix(int r, int c) : r_(r), c_(c), a_((size_t)r * c) {}
    int rows() const { return r_; }
    int cols() const { return c_; }
    void resize(int r, int c) {
        if ((size_t)r * c != a_.size()) a_.assign((size_t)r * c, Scalar(0.0));
        r_ = r;
        c_ = c;
    }
    Scalar &operator()(int i, int j) { return a_[(size_t)i * c_ + j]; }
    const Scalar &operator()(int i, int j) const { return a_[(size_t)i * c_ + j]; }
    Scalar *data() { return a_.data(); }
    const Scalar *data() const { return a_.data(); }
    static cdmatrix Zero(int r, int c) { return cdmatrix(r, c); }
    static cdmatrix Identity(int r, int c) {
        cdmatrix m(r, c);
        for (int i = 0; i < r && i < c; i++) m(i, i) = 1.0;
        return m;
    }
    cdmatrix &operator+=(const cdmatrix &b) {
        for (size_t i = 0; i < a_.size(); i++) a_[i] += b.a_[i];
        return *this;
    }
    cdmatrix &operator-=(const cdmatrix &b) {
        for (size_t i = 0; i < a_.size(); i++) a_[i] -= b.a_[i];
        return *this;
    }
    cdmatrix &operator*=(Scalar s) {
        for (size_t i = 0; i < a_.size(); i++) a_[i] *= s;
        return *this;
    }
    friend cdmatrix operator+(cdmatrix a, const cdmatrix &b) { return a += b; }
    friend cdmatrix operator-(cdmatrix a, const cdmatrix &b) { return a -= b; }
    friend cdmatrix operator*(cdmatrix a, Scalar s) { return a *= s; }
    friend cdmatrix operator*(Scalar s, cdmatrix a) { return a *= s; }

  private:
    int r_, c_;
    std::vector<Scalar> a_;
};

namespace integration {
/* The order k is all the library needs: it holds the weight tables itself. */
template <typename T>
class Integrator {
  public:
    explicit Integrator(int k = 0) : k_(k) {}
    int get_k() const { return k_; }
    int k() const { return k_; }

  private:
    int k_;
};
template <typename T>
Integrator<T> &I(int k) {
    static Integrator<T> tab[MAX_SOLVE_ORDER + 1] = {Integrator<T>(0), Integrator<T>(1), Integrator<T>(2),
                                                     Integrator<T>(3), Integrator<T>(4), Integrator<T>(5)};
    if (k < 0 || k > MAX_SOLVE_ORDER) {
        std::fprintf(stderr, "integration::I: order %d out of range\n", k);
        std::abort();
    }
    return tab[k];
}
}  // namespace integration

namespace cntr {

typedef std::complex<double> cdouble;

namespace amd_detail {
inline void check(int rc, const char *what) {
    if (rc != CNTRX_OK) {
        std::fprintf(stderr, "cntr::%s (libcntrx): %s\n", what, cntrx_last_error());
        std::abort();
    }
}
// one device allocation, freed with the object
class DevMem {
  public:
    DevMem() : p_(nullptr), n_(0) {}
    ~DevMem() { release(); }
    DevMem(const DevMem &) = delete;
    DevMem &operator=(const DevMem &) = delete;
    void *get(size_t bytes) {
        if (bytes != n_) {
            release();
            if (bytes) check(cntrx_device_alloc(bytes, &p_), "device_alloc");
            n_ = bytes;
        }
        return p_;
    }
    void release() {
        if (p_) cntrx_device_free(p_);
        p_ = nullptr;
        n_ = 0;
    }
    void swap(DevMem &o) {
        std::swap(p_, o.p_);
        std::swap(n_, o.n_);
    }

  private:
    void *p_;
    size_t n_;
};
inline void up(void *dst, const void *src, size_t bytes) {
    if (bytes) check(cntrx_copy(dst, src, bytes, 1, nullptr), "copy");
}
inline void down(void *dst, const void *src, size_t bytes) {
    if (bytes) check(cntrx_copy(dst, src, bytes, 2, nullptr), "copy");
}
}  // namespace amd_detail

template <typename T>
class function;
template <typename T>
class herm_matrix;

/* H(t), t = -1..nt (t = -1 the Matsubara value), size1 x size1 blocks. */
template <>
class function<double> {
  public:
    typedef std::complex<double> cplx;
    function() : nt_(-2), size1_(0), es_(0), dirty_(false) {}
    function(int nt, int size1 = 1)
        : nt_(nt), size1_(size1), es_(size1 * size1), data_((size_t)(nt + 2) * size1 * size1), dirty_(true) {}
    function(const function &o) : nt_(o.nt_), size1_(o.size1_), es_(o.es_), data_(o.data_), dirty_(true) {}
    function &operator=(const function &o) {
        nt_ = o.nt_;
        size1_ = o.size1_;
        es_ = o.es_;
        data_ = o.data_;
        dirty_ = true;
        return *this;
    }
    int nt() const { return nt_; }
    int size1() const { return size1_; }
    int size2() const { return size1_; }
    int element_size() const { return es_; }
    cplx *ptr(int t) {
        dirty_ = true;
        return &data_[(size_t)(t + 1) * es_];
    }
    const cplx *ptr(int t) const { return &data_[(size_t)(t + 1) * es_]; }
    void set_value(int t, cplx x) { ptr(t)[0] = x; }
    void set_value(int t, const cplx *M) { std::memcpy((void *)ptr(t), M, es_ * sizeof(cplx)); }
    void set_value(int t, const cdmatrix &M) { set_value(t, M.data()); }
    void set_constant(const cplx *M) {
        for (int t = -1; t <= nt_; t++) set_value(t, M);
    }
    void set_constant(const cdmatrix &M) { set_constant(M.data()); }
    void set_constant(cplx x) {
        for (int t = -1; t <= nt_; t++) set_value(t, x);
    }
    void get_value(int t, cplx &x) const { x = ptr(t)[0]; }
    void get_value(int t, cplx *M) const { std::memcpy((void *)M, ptr(t), es_ * sizeof(cplx)); }
    void get_value(int t, cdmatrix &M) const {
        M.resize(size1_, size1_);
        get_value(t, M.data());
    }
    void mark_host_modified() { dirty_ = true; }
    // text files of the reference's format (cntr/cntr_function_impl.hpp:719-824)
    void print_to_file(const char *file, int precision = 16) const {
        const cntrx_func &d = dev_sync();
        amd_detail::check(cntrx_function_print_to_file(&d, file, precision, nullptr), "function::print_to_file");
    }
    void read_from_file(const char *file) {
        int n, s1;
        amd_detail::check(cntrx_function_file_header(file, &n, &s1), "function::read_from_file");
        if (n > nt_ || s1 != size1_) *this = function(n, s1);
        const cntrx_func &d = dev_sync();
        amd_detail::check(cntrx_function_read_from_file(&d, file, nullptr), "function::read_from_file");
        amd_detail::down(data_.data(), d.data, data_.size() * sizeof(cplx));
    }
    void read_from_file(int nt1, const char *file) {
        const cntrx_func &d = dev_sync();
        amd_detail::check(cntrx_function_read_from_file_nt1(&d, nt1, file, nullptr), "function::read_from_file");
        amd_detail::down(data_.data(), d.data, data_.size() * sizeof(cplx));
    }
    // device descriptor, uploaded first if the host copy is newer
    const cntrx_func &dev_sync() const {
        if (dirty_) {
            void *p = dev_.get(data_.size() * sizeof(cplx));
            amd_detail::up(p, data_.data(), data_.size() * sizeof(cplx));
            d_.nt = nt_;
            d_.size1 = size1_;
            d_.data = (double *)p;
            dirty_ = false;
        }
        return d_;
    }

  private:
    int nt_, size1_, es_;
    std::vector<cplx> data_;
    mutable amd_detail::DevMem dev_;
    mutable cntrx_func d_;
    mutable bool dirty_;
};

/* Contour function G(t,t') in the reference's packed layout: ret(i,j) j <= i by rows,
 * les(i,j) i <= j by columns, tv(i,m), mat(m) (cntr/cntr_herm_matrix_impl.hpp:456-478). */
template <>
class herm_matrix<double> {
  public:
    typedef std::complex<double> cplx;
    typedef double scalar_type;
    herm_matrix() : nt_(-2), ntau_(0), size1_(0), sig_(-1), es_(0), all_dirty_(true) {}
    herm_matrix(herm_matrix &&g) noexcept : nt_(-2), ntau_(0), size1_(0), sig_(-1), es_(0), all_dirty_(true) {
        swap_in(g);
    }
    herm_matrix(int nt, int ntau, int size1 = 1, int sig = -1) { init(nt, ntau, size1, sig); }
    herm_matrix(int nt, int ntau, int size1, int size2, int sig) {
        if (size2 != size1) {
            std::fprintf(stderr, "cntr::herm_matrix: size1 != size2 is not provided\n");
            std::abort();
        }
        init(nt, ntau, size1, sig);
    }
    herm_matrix(const herm_matrix &g) : all_dirty_(true) { *this = g; }
    herm_matrix &operator=(const herm_matrix &g) {
        if (this == &g) return *this;
        nt_ = g.nt_;
        ntau_ = g.ntau_;
        size1_ = g.size1_;
        sig_ = g.sig_;
        es_ = g.es_;
        ret_ = g.ret_;
        les_ = g.les_;
        tv_ = g.tv_;
        mat_ = g.mat_;
        dirty_.assign(nt_ + 2, 1);
        all_dirty_ = true;
        dev_.release();  // the shape may have changed: a new device copy on next use
        block_ = nullptr;
        return *this;
    }
    herm_matrix &operator=(herm_matrix &&g) noexcept {
        swap_in(g);
        return *this;
    }

    int element_size() const { return es_; }
    int size1() const { return size1_; }
    int size2() const { return size1_; }
    int ntau() const { return ntau_; }
    int nt() const { return nt_; }
    int sig() const { return sig_; }
    void set_sig(int sig) { sig_ = sig; }

    // the reference's offsets (cntr/cntr_herm_matrix_impl.hpp:456-478); non-const access
    // marks the element's timestep as written on the host
    cplx *retptr(int i, int j) {
        touch(i);
        return &ret_[roff(i, j)];
    }
    cplx *lesptr(int i, int j) {
        touch(j);
        return &les_[loff(i, j)];
    }
    cplx *tvptr(int i, int m) {
        touch(i);
        return &tv_[toff(i, m)];
    }
    cplx *matptr(int m) {
        touch(-1);
        return &mat_[(size_t)m * es_];
    }
    const cplx *retptr(int i, int j) const { return &ret_[roff(i, j)]; }
    const cplx *lesptr(int i, int j) const { return &les_[loff(i, j)]; }
    const cplx *tvptr(int i, int m) const { return &tv_[toff(i, m)]; }
    const cplx *matptr(int m) const { return &mat_[(size_t)m * es_]; }

    // element access with the reference's symmetries (cntr/cntr_herm_matrix_impl.hpp:569-963)
    void get_ret(int i, int j, cplx *M) const {
        if (i >= j) read(retptr(i, j), M);
        else read_minus_conj(retptr(j, i), M);
    }
    void get_les(int i, int j, cplx *M) const {
        if (i <= j) read(lesptr(i, j), M);
        else read_minus_conj(lesptr(j, i), M);
    }
    void get_gtr(int i, int j, cplx *M) const {
        std::vector<cplx> M1(es_);
        get_ret(i, j, M);
        get_les(i, j, M1.data());
        for (int e = 0; e < es_; e++) M[e] += M1[e];
    }
    void get_tv(int i, int m, cplx *M) const { read(tvptr(i, m), M); }
    void get_vt(int m, int j, cplx *M) const {
        read_minus_conj(tvptr(j, ntau_ - m), M);
        if (sig_ == -1)
            for (int e = 0; e < es_; e++) M[e] = -M[e];
    }
    void get_mat(int m, cplx *M) const { read(matptr(m), M); }
    void get_matminus(int m, cplx *M) const {
        read(matptr(ntau_ - m), M);
        if (sig_ == -1)
            for (int e = 0; e < es_; e++) M[e] = -M[e];
    }
    void get_ret(int i, int j, cplx &x) const { get_ret(i, j, &x); }
    void get_les(int i, int j, cplx &x) const { get_les(i, j, &x); }
    void get_gtr(int i, int j, cplx &x) const { get_gtr(i, j, &x); }
    void get_tv(int i, int m, cplx &x) const { get_tv(i, m, &x); }
    void get_vt(int m, int j, cplx &x) const { get_vt(m, j, &x); }
    void get_mat(int m, cplx &x) const { get_mat(m, &x); }
    void get_matminus(int m, cplx &x) const { get_matminus(m, &x); }
    void get_ret(int i, int j, cdmatrix &M) const { get_ret(i, j, sized(M)); }
    void get_les(int i, int j, cdmatrix &M) const { get_les(i, j, sized(M)); }
    void get_gtr(int i, int j, cdmatrix &M) const { get_gtr(i, j, sized(M)); }
    void get_tv(int i, int m, cdmatrix &M) const { get_tv(i, m, sized(M)); }
    void get_vt(int m, int j, cdmatrix &M) const { get_vt(m, j, sized(M)); }
    void get_mat(int m, cdmatrix &M) const { get_mat(m, sized(M)); }
    void get_matminus(int m, cdmatrix &M) const { get_matminus(m, sized(M)); }

    // density_matrix (cntr/cntr_herm_matrix_impl.hpp:983-1028)
    cplx density_matrix(int tstp) const {
        cplx x;
        if (tstp == -1) {
            get_mat(ntau_, x);
            return -x;
        }
        get_les(tstp, tstp, x);
        return cplx(0.0, sig_) * x;
    }
    void density_matrix(int tstp, cdmatrix &M) const {
        cplx *m = sized(M);
        if (tstp == -1) {
            get_mat(ntau_, m);
            for (int e = 0; e < es_; e++) m[e] = -m[e];
        } else {
            get_les(tstp, tstp, m);
            for (int e = 0; e < es_; e++) m[e] *= cplx(0.0, sig_);
        }
    }

    void set_ret(int i, int j, cplx x) { retptr(i, j)[0] = x; }
    void set_les(int i, int j, cplx x) { lesptr(i, j)[0] = x; }
    void set_tv(int i, int m, cplx x) { tvptr(i, m)[0] = x; }
    void set_mat(int m, cplx x) { matptr(m)[0] = x; }
    void set_ret(int i, int j, const cdmatrix &M) { std::memcpy((void *)retptr(i, j), M.data(), es_ * sizeof(cplx)); }
    void set_les(int i, int j, const cdmatrix &M) { std::memcpy((void *)lesptr(i, j), M.data(), es_ * sizeof(cplx)); }
    void set_tv(int i, int m, const cdmatrix &M) { std::memcpy((void *)tvptr(i, m), M.data(), es_ * sizeof(cplx)); }
    void set_mat(int m, const cdmatrix &M) { std::memcpy((void *)matptr(m), M.data(), es_ * sizeof(cplx)); }

    // timestep algebra on the host copy (cntr/cntr_herm_matrix_impl.hpp:2064-2400, 3339-3400)
    void set_timestep_zero(int tstp) {
        for_strips(tstp, [](cplx *x, size_t n, const cplx *) {
            for (size_t i = 0; i < n; i++) x[i] = 0.0;
        });
    }
    void smul(int tstp, double weight) { smul(tstp, cplx(weight, 0.0)); }
    void smul(int tstp, cplx weight) {
        for_strips(tstp, [weight](cplx *x, size_t n, const cplx *) {
            for (size_t i = 0; i < n; i++) x[i] *= weight;
        });
    }
    void set_timestep(int tstp, const herm_matrix &g) {
        for_strips(tstp, [](cplx *x, size_t n, const cplx *y) { std::memcpy((void *)x, y, n * sizeof(cplx)); }, &g);
    }
    void incr_timestep(int tstp, const herm_matrix &g, cplx alpha = 1.0) {
        for_strips(tstp, [alpha](cplx *x, size_t n, const cplx *y) {
            for (size_t i = 0; i < n; i++) x[i] += alpha * y[i];
        }, &g);
    }
    // element (i1, i2) <- element (j1, j2) of g at timestep tstp
    // (cntr/cntr_herm_matrix_impl.hpp:2313-2364)
    void set_matrixelement(int tstp, int i1, int i2, const herm_matrix &g, int j1, int j2) {
        const int s = i1 * size1_ + i2, u = j1 * g.size1_ + j2;
        if (tstp == -1) {
            for (int m = 0; m <= ntau_; m++) matptr(m)[s] = g.matptr(m)[u];
            return;
        }
        for (int j = 0; j <= tstp; j++) retptr(tstp, j)[s] = g.retptr(tstp, j)[u];
        for (int m = 0; m <= ntau_; m++) tvptr(tstp, m)[s] = g.tvptr(tstp, m)[u];
        for (int i = 0; i <= tstp; i++) lesptr(i, tstp)[s] = g.lesptr(i, tstp)[u];
    }
    void set_matrixelement(int tstp, int i1, int i2, const herm_matrix &g) {
        set_matrixelement(tstp, i1, i2, g, i1, i2);
    }
    void mark_host_modified() {
        all_dirty_ = true;
    }
    // text files of the reference's format (cntr/cntr_herm_matrix_impl.hpp:1308-1566)
    void print_to_file(const char *file, int precision = 16) const {
        const cntrx_herm &d = dev_sync();
        amd_detail::check(cntrx_herm_print_to_file(&d, file, precision, nullptr), "herm_matrix::print_to_file");
    }
    void read_from_file(const char *file) {
        int n, m, s1, sg;
        amd_detail::check(cntrx_herm_file_header(file, &n, &m, &s1, &sg), "herm_matrix::read_from_file");
        if (n > nt_ || m != ntau_ || s1 != size1_) init(n, m, s1, sig_);
        const cntrx_herm &d = dev_sync();
        amd_detail::check(cntrx_herm_read_from_file(&d, file, &sg, nullptr), "herm_matrix::read_from_file");
        sig_ = sg;
        dev_pull_all();
    }
    void read_from_file(int nt1, const char *file) {
        const cntrx_herm &d = dev_sync();
        amd_detail::check(cntrx_herm_read_from_file_nt1(&d, nt1, file, nullptr), "herm_matrix::read_from_file");
        dev_pull_all();
    }

    // device side, for the free functions below
    const cntrx_herm &dev_sync() const {
        if (!block_) {
            block_ = (cplx *)dev_.get(total() * sizeof(cplx));
            d_.nt = nt_;
            d_.ntau = ntau_;
            d_.size1 = size1_;
            d_.sig = sig_;
            d_.ret = (double *)block_;
            d_.les = (double *)(block_ + ret_.size());
            d_.tv = (double *)(block_ + ret_.size() + les_.size());
            d_.mat = (double *)(block_ + ret_.size() + les_.size() + tv_.size());
            all_dirty_ = true;
        }
        d_.sig = sig_;
        if (all_dirty_) {
            amd_detail::up(d_.ret, ret_.data(), ret_.size() * sizeof(cplx));
            amd_detail::up(d_.les, les_.data(), les_.size() * sizeof(cplx));
            amd_detail::up(d_.tv, tv_.data(), tv_.size() * sizeof(cplx));
            amd_detail::up(d_.mat, mat_.data(), mat_.size() * sizeof(cplx));
            std::fill(dirty_.begin(), dirty_.end(), 0);
            all_dirty_ = false;
        } else {
            for (int t = -1; t <= nt_; t++)
                if (dirty_[t + 1]) {
                    move_timestep(t, true);
                    dirty_[t + 1] = 0;
                }
        }
        return d_;
    }
    // timestep t (-1: mat) device -> host, after a solver wrote it
    void dev_pull(int t) const { move_timestep(t, false); }
    void dev_pull_all() const {
        amd_detail::down((void *)ret_.data(), d_.ret, ret_.size() * sizeof(cplx));
        amd_detail::down((void *)les_.data(), d_.les, les_.size() * sizeof(cplx));
        amd_detail::down((void *)tv_.data(), d_.tv, tv_.size() * sizeof(cplx));
        amd_detail::down((void *)mat_.data(), d_.mat, mat_.size() * sizeof(cplx));
    }

  private:
    void init(int nt, int ntau, int size1, int sig) {
        nt_ = nt;
        ntau_ = ntau;
        size1_ = size1;
        sig_ = sig;
        es_ = size1 * size1;
        const size_t ntri = nt >= 0 ? (size_t)(nt + 1) * (nt + 2) / 2 : 0;
        const size_t ntv = nt >= 0 ? (size_t)(nt + 1) * (ntau + 1) : 0;
        ret_.assign(ntri * es_, cplx(0.0));
        les_.assign(ntri * es_, cplx(0.0));
        tv_.assign(ntv * es_, cplx(0.0));
        mat_.assign((size_t)(ntau + 1) * es_, cplx(0.0));
        dirty_.assign(nt + 2, 1);
        all_dirty_ = true;
        block_ = nullptr;
    }
    void swap_in(herm_matrix &g) {
        std::swap(nt_, g.nt_);
        std::swap(ntau_, g.ntau_);
        std::swap(size1_, g.size1_);
        std::swap(sig_, g.sig_);
        std::swap(es_, g.es_);
        ret_.swap(g.ret_);
        les_.swap(g.les_);
        tv_.swap(g.tv_);
        mat_.swap(g.mat_);
        dirty_.swap(g.dirty_);
        std::swap(all_dirty_, g.all_dirty_);
        dev_.swap(g.dev_);
        std::swap(block_, g.block_);
        std::swap(d_, g.d_);
    }
    size_t total() const { return ret_.size() + les_.size() + tv_.size() + mat_.size(); }
    size_t roff(int i, int j) const { return ((size_t)i * (i + 1) / 2 + j) * es_; }
    size_t loff(int i, int j) const { return ((size_t)j * (j + 1) / 2 + i) * es_; }
    size_t toff(int i, int m) const { return ((size_t)i * (ntau_ + 1) + m) * es_; }
    void touch(int t) {
        if (t + 1 >= 0 && t + 1 < (int)dirty_.size()) dirty_[t + 1] = 1;
        else all_dirty_ = true;
    }
    cplx *sized(cdmatrix &M) const {
        M.resize(size1_, size1_);
        return M.data();
    }
    void read(const cplx *x, cplx *M) const { std::memcpy((void *)M, x, es_ * sizeof(cplx)); }
    void read_minus_conj(const cplx *x, cplx *M) const {
        for (int r = 0; r < size1_; r++)
            for (int s = 0; s < size1_; s++) {
                const cplx w = x[s * size1_ + r];
                M[r * size1_ + s] = cplx(-w.real(), w.imag());
            }
    }
    // f(x, n, y) on the strips of timestep tstp: mat for -1, else ret row, tv row, les column
    template <class F>
    void for_strips(int tstp, F f, const herm_matrix *g = nullptr) {
        if (tstp == -1) {
            f(matptr(0), (size_t)(ntau_ + 1) * es_, g ? g->matptr(0) : nullptr);
            return;
        }
        f(retptr(tstp, 0), (size_t)(tstp + 1) * es_, g ? g->retptr(tstp, 0) : nullptr);
        f(tvptr(tstp, 0), (size_t)(ntau_ + 1) * es_, g ? g->tvptr(tstp, 0) : nullptr);
        f(lesptr(0, tstp), (size_t)(tstp + 1) * es_, g ? g->lesptr(0, tstp) : nullptr);
    }
    // one timestep's strips host <-> device
    void move_timestep(int t, bool to_device) const {
        const size_t B = sizeof(cplx);
        auto mv = [to_device](void *host, void *dev, size_t bytes) {
            if (to_device) amd_detail::up(dev, host, bytes);
            else amd_detail::down(host, dev, bytes);
        };
        if (t == -1) {
            mv((void *)mat_.data(), d_.mat, mat_.size() * B);
            return;
        }
        const size_t n = (size_t)(t + 1) * es_;
        mv((void *)&ret_[roff(t, 0)], (cplx *)d_.ret + roff(t, 0), n * B);
        mv((void *)&les_[loff(0, t)], (cplx *)d_.les + loff(0, t), n * B);
        mv((void *)&tv_[toff(t, 0)], (cplx *)d_.tv + toff(t, 0), (size_t)(ntau_ + 1) * es_ * B);
    }

    int nt_, ntau_, size1_, sig_, es_;
    std::vector<cplx> ret_, les_, tv_, mat_;
    mutable std::vector<unsigned char> dirty_;  // timestep t at t + 1, written on the host
    mutable bool all_dirty_;
    mutable amd_detail::DevMem dev_;
    mutable cplx *block_ = nullptr;
    mutable cntrx_herm d_{};
};

typedef herm_matrix<double> GREEN_AMD;

namespace amd_detail {
inline void pull_range(const herm_matrix<double> &G, int t0, int t1) {
    for (int t = t0; t <= t1; t++) G.dev_pull(t);
}
}  // namespace amd_detail

/* ---- Dyson equation (cntr/cntr_dyson_decl.hpp:78-93, Integrator forms :51-73) ---- */
inline void dyson_mat(herm_matrix<double> &G, double mu, function<double> &H, herm_matrix<double> &Sigma,
                      double beta, const int SolveOrder = MAX_SOLVE_ORDER, const int method = CNTR_MAT_FIXPOINT,
                      const bool force_hermitian = true) {
    const cntrx_herm &g = G.dev_sync(), &s = Sigma.dev_sync();
    const cntrx_func &h = H.dev_sync();
    amd_detail::check(cntrx_dyson_mat(&g, mu, &h, &s, beta, SolveOrder, method, force_hermitian ? 1 : 0, nullptr),
                      "dyson_mat");
    G.dev_pull(-1);
}
inline void dyson_start(herm_matrix<double> &G, double mu, function<double> &H, herm_matrix<double> &Sigma,
                        double beta, double h, const int SolveOrder = MAX_SOLVE_ORDER) {
    const cntrx_herm &g = G.dev_sync(), &s = Sigma.dev_sync();
    const cntrx_func &f = H.dev_sync();
    amd_detail::check(cntrx_dyson_start(&g, mu, &f, &s, beta, h, SolveOrder, nullptr), "dyson_start");
    amd_detail::pull_range(G, 0, SolveOrder < G.nt() ? SolveOrder : G.nt());
}
inline void dyson_timestep(int n, herm_matrix<double> &G, double mu, function<double> &H,
                           herm_matrix<double> &Sigma, double beta, double h,
                           const int SolveOrder = MAX_SOLVE_ORDER) {
    const cntrx_herm &g = G.dev_sync(), &s = Sigma.dev_sync();
    const cntrx_func &f = H.dev_sync();
    amd_detail::check(cntrx_dyson_timestep(n, &g, mu, &f, &s, beta, h, SolveOrder, nullptr), "dyson_timestep");
    G.dev_pull(n);
}
// the reference's dyson(...) wrapper (cntr/cntr_dyson_impl.hpp:1725-1734)
inline void dyson(herm_matrix<double> &G, double mu, function<double> &H, herm_matrix<double> &Sigma, double beta,
                  double h, const int SolveOrder = MAX_SOLVE_ORDER, const int method = CNTR_MAT_FIXPOINT,
                  const bool force_hermitian = true) {
    dyson_mat(G, mu, H, Sigma, beta, SolveOrder, method, force_hermitian);
    dyson_start(G, mu, H, Sigma, beta, h, SolveOrder);
    for (int n = SolveOrder + 1; n <= G.nt(); n++) dyson_timestep(n, G, mu, H, Sigma, beta, h, SolveOrder);
}
inline void dyson_mat(herm_matrix<double> &G, herm_matrix<double> &Sigma, double mu, function<double> &H,
                      integration::Integrator<double> &I, double beta, const int method = CNTR_MAT_FIXPOINT,
                      const bool force_hermitian = true) {
    dyson_mat(G, mu, H, Sigma, beta, I.get_k(), method, force_hermitian);
}
inline void dyson_start(herm_matrix<double> &G, double mu, function<double> &H, herm_matrix<double> &Sigma,
                        integration::Integrator<double> &I, double beta, double h) {
    dyson_start(G, mu, H, Sigma, beta, h, I.get_k());
}
inline void dyson_timestep(int n, herm_matrix<double> &G, double mu, function<double> &H,
                           herm_matrix<double> &Sigma, integration::Integrator<double> &I, double beta, double h) {
    dyson_timestep(n, G, mu, H, Sigma, beta, h, I.get_k());
}

/* ---- integral form [1 + F] G = Q (cntr/cntr_vie2_decl.hpp:20-76) ---- */
inline void vie2_mat(herm_matrix<double> &G, herm_matrix<double> &F, herm_matrix<double> &Fcc,
                     herm_matrix<double> &Q, double beta, const int SolveOrder = MAX_SOLVE_ORDER,
                     const int method = CNTR_MAT_FIXPOINT) {
    const cntrx_herm &g = G.dev_sync(), &f = F.dev_sync(), &fc = Fcc.dev_sync(), &q = Q.dev_sync();
    amd_detail::check(cntrx_vie2_mat(&g, &f, &fc, &q, beta, SolveOrder, method, nullptr), "vie2_mat");
    G.dev_pull(-1);
}
inline void vie2_start(herm_matrix<double> &G, herm_matrix<double> &F, herm_matrix<double> &Fcc,
                       herm_matrix<double> &Q, double beta, double h, const int SolveOrder = MAX_SOLVE_ORDER) {
    const cntrx_herm &g = G.dev_sync(), &f = F.dev_sync(), &fc = Fcc.dev_sync(), &q = Q.dev_sync();
    amd_detail::check(cntrx_vie2_start(&g, &f, &fc, &q, beta, h, SolveOrder, nullptr), "vie2_start");
    amd_detail::pull_range(G, 0, SolveOrder < G.nt() ? SolveOrder : G.nt());
}
inline void vie2_timestep(int n, herm_matrix<double> &G, herm_matrix<double> &F, herm_matrix<double> &Fcc,
                          herm_matrix<double> &Q, double beta, double h, const int SolveOrder = MAX_SOLVE_ORDER,
                          const int matsubara_method = CNTR_MAT_FIXPOINT) {
    const cntrx_herm &g = G.dev_sync(), &f = F.dev_sync(), &fc = Fcc.dev_sync(), &q = Q.dev_sync();
    amd_detail::check(cntrx_vie2_timestep(n, &g, &f, &fc, &q, beta, h, SolveOrder, matsubara_method, nullptr),
                      "vie2_timestep");
    if (n == -1) G.dev_pull(-1);
    else if (n <= SolveOrder) amd_detail::pull_range(G, 0, n);
    else G.dev_pull(n);
}
inline void vie2_start(herm_matrix<double> &G, herm_matrix<double> &F, herm_matrix<double> &Fcc,
                       herm_matrix<double> &Q, integration::Integrator<double> &I, double beta, double h) {
    vie2_start(G, F, Fcc, Q, beta, h, I.get_k());
}
inline void vie2_timestep(int n, herm_matrix<double> &G, herm_matrix<double> &F, herm_matrix<double> &Fcc,
                          herm_matrix<double> &Q, integration::Integrator<double> &I, double beta, double h,
                          const int matsubara_method = CNTR_MAT_FIXPOINT) {
    vie2_timestep(n, G, F, Fcc, Q, beta, h, I.get_k(), matsubara_method);
}

/* ---- convolution C = A * B (cntr/cntr_convolution_decl.hpp:31-66) ---- */
inline void convolution_timestep(int n, herm_matrix<double> &C, herm_matrix<double> &A, herm_matrix<double> &Acc,
                                 herm_matrix<double> &B, herm_matrix<double> &Bcc, double beta, double h,
                                 int SolveOrder = MAX_SOLVE_ORDER) {
    const cntrx_herm &c = C.dev_sync(), &a = A.dev_sync(), &ac = Acc.dev_sync(), &b = B.dev_sync(),
                     &bc = Bcc.dev_sync();
    amd_detail::check(cntrx_convolution_timestep(n, &c, &a, &ac, &b, &bc, beta, h, SolveOrder, nullptr),
                      "convolution_timestep");
    C.dev_pull(n);
}
inline void convolution_timestep(int n, herm_matrix<double> &C, herm_matrix<double> &A, herm_matrix<double> &B,
                                 double beta, double h, int SolveOrder = MAX_SOLVE_ORDER) {
    convolution_timestep(n, C, A, A, B, B, beta, h, SolveOrder);
}
inline void convolution_timestep(int n, herm_matrix<double> &C, herm_matrix<double> &A, herm_matrix<double> &B,
                                 integration::Integrator<double> &I, double beta, double h) {
    convolution_timestep(n, C, A, A, B, B, beta, h, I.get_k());
}
inline void convolution_timestep(int n, herm_matrix<double> &C, herm_matrix<double> &A, herm_matrix<double> &Acc,
                                 herm_matrix<double> &B, herm_matrix<double> &Bcc,
                                 integration::Integrator<double> &I, double beta, double h) {
    convolution_timestep(n, C, A, Acc, B, Bcc, beta, h, I.get_k());
}

/* ---- bubbles (cntr/cntr_bubble_decl.hpp:62-301) ---- */
inline void Bubble1(int tstp, herm_matrix<double> &C, int c1, int c2, herm_matrix<double> &A,
                    herm_matrix<double> &Acc, int a1, int a2, herm_matrix<double> &B, herm_matrix<double> &Bcc,
                    int b1, int b2) {
    const cntrx_herm &c = C.dev_sync(), &a = A.dev_sync(), &ac = Acc.dev_sync(), &b = B.dev_sync(),
                     &bc = Bcc.dev_sync();
    amd_detail::check(cntrx_bubble1(tstp, &c, c1, c2, &a, &ac, a1, a2, &b, &bc, b1, b2, nullptr), "Bubble1");
    C.dev_pull(tstp);
}
inline void Bubble1(int tstp, herm_matrix<double> &C, int c1, int c2, herm_matrix<double> &A, int a1, int a2,
                    herm_matrix<double> &B, int b1, int b2) {
    Bubble1(tstp, C, c1, c2, A, A, a1, a2, B, B, b1, b2);
}
inline void Bubble1(int tstp, herm_matrix<double> &C, herm_matrix<double> &A, herm_matrix<double> &Acc,
                    herm_matrix<double> &B, herm_matrix<double> &Bcc) {
    Bubble1(tstp, C, 0, 0, A, Acc, 0, 0, B, Bcc, 0, 0);
}
inline void Bubble1(int tstp, herm_matrix<double> &C, herm_matrix<double> &A, herm_matrix<double> &B) {
    Bubble1(tstp, C, 0, 0, A, A, 0, 0, B, B, 0, 0);
}
inline void Bubble2(int tstp, herm_matrix<double> &C, int c1, int c2, herm_matrix<double> &A,
                    herm_matrix<double> &Acc, int a1, int a2, herm_matrix<double> &B, herm_matrix<double> &Bcc,
                    int b1, int b2) {
    const cntrx_herm &c = C.dev_sync(), &a = A.dev_sync(), &ac = Acc.dev_sync(), &b = B.dev_sync(),
                     &bc = Bcc.dev_sync();
    amd_detail::check(cntrx_bubble2(tstp, &c, c1, c2, &a, &ac, a1, a2, &b, &bc, b1, b2, nullptr), "Bubble2");
    C.dev_pull(tstp);
}
inline void Bubble2(int tstp, herm_matrix<double> &C, int c1, int c2, herm_matrix<double> &A, int a1, int a2,
                    herm_matrix<double> &B, int b1, int b2) {
    Bubble2(tstp, C, c1, c2, A, A, a1, a2, B, B, b1, b2);
}
inline void Bubble2(int tstp, herm_matrix<double> &C, herm_matrix<double> &A, herm_matrix<double> &Acc,
                    herm_matrix<double> &B, herm_matrix<double> &Bcc) {
    Bubble2(tstp, C, 0, 0, A, Acc, 0, 0, B, Bcc, 0, 0);
}
inline void Bubble2(int tstp, herm_matrix<double> &C, herm_matrix<double> &A, herm_matrix<double> &B) {
    Bubble2(tstp, C, 0, 0, A, A, 0, 0, B, B, 0, 0);
}

/* ---- utilities (cntr/cntr_utilities_decl.hpp, cntr/cntr_equilibrium_decl.hpp:73) ---- */
inline void extrapolate_timestep(int n, herm_matrix<double> &G, int SolveOrder = MAX_SOLVE_ORDER) {
    const cntrx_herm &g = G.dev_sync();
    amd_detail::check(cntrx_extrapolate_timestep(n, &g, SolveOrder, nullptr), "extrapolate_timestep");
    G.dev_pull(n + 1);
}
inline void extrapolate_timestep(int n, herm_matrix<double> &G, integration::Integrator<double> &I) {
    extrapolate_timestep(n, G, I.get_k());
}
inline double distance_norm2(int tstp, herm_matrix<double> &g1, herm_matrix<double> &g2) {
    const cntrx_herm &a = g1.dev_sync(), &b = g2.dev_sync();
    double d = 0.0;
    amd_detail::check(cntrx_distance_norm2(tstp, &a, &b, &d, nullptr), "distance_norm2");
    return d;
}
inline void force_matsubara_hermitian(herm_matrix<double> &G) {
    const cntrx_herm &g = G.dev_sync();
    amd_detail::check(cntrx_force_matsubara_hermitian(&g, nullptr), "force_matsubara_hermitian");
    G.dev_pull(-1);
}
// constant H: eps is size1 x size1, row-major
inline void green_from_H(herm_matrix<double> &G, double mu, const cdmatrix &eps, double beta, double h) {
    const cntrx_herm &g = G.dev_sync();
    amd_detail::check(cntrx_green_from_H(&g, mu, (const double *)eps.data(), beta, h, nullptr), "green_from_H");
    G.dev_pull_all();
}

}  // namespace cntr

#endif
