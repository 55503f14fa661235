/* C++ drop-in for the reference's cntr:: API on the Kadanoff-Baym path, over the C ABI
 * of libcntrx.so (cntrx.h).  Class and function names, argument order and defaults follow
 * libcntr: herm_matrix (cntr/cntr_herm_matrix_decl.hpp:44-301), function
 * (cntr/cntr_function_decl.hpp:26-107), dyson_* (cntr/cntr_dyson_decl.hpp:40-93), vie2_*
 * (cntr/cntr_vie2_decl.hpp:20-76), convolution_timestep (cntr/cntr_convolution_decl.hpp:
 * 31-66), Bubble1/Bubble2 (cntr/cntr_bubble_decl.hpp:62-301), extrapolate_timestep and
 * distance_norm2 (cntr/cntr_utilities_decl.hpp), green_from_H for constant and
 * time-dependent H (cntr/cntr_equilibrium_decl.hpp:73-98), and integration::I<T>(k)
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
 * herm_matrix_timestep and herm_matrix_timestep_view hold or point at one timestep in host
 * memory; bubbles and distance_norm2 accept them in place of a herm_matrix.
 * Not provided (outside the path): Eigen types (cdmatrix below is a small dense matrix
 * with the operations the callers use), HDF5 I/O (the text format is provided), MPI, the
 * _omp variants. */
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
template <typename T>
class herm_matrix_timestep;
template <typename T>
class herm_matrix_timestep_view;

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
    // timestep objects (cntr/cntr_herm_matrix_impl.hpp:2098-2310): copy timestep tstp out
    // (resizing the timestep object), copy it in, add alpha times it, or set one element
    void get_timestep(int tstp, herm_matrix_timestep<double> &t) const;
    void set_timestep(int tstp, const herm_matrix_timestep<double> &t);
    void incr_timestep(int tstp, const herm_matrix_timestep<double> &t, cplx alpha = 1.0);
    void incr_timestep(int tstp, const herm_matrix_timestep<double> &t, double alpha) {
        incr_timestep(tstp, t, cplx(alpha, 0.0));
    }
    void set_matrixelement(int tstp, int i1, int i2, const herm_matrix_timestep<double> &g, int j1, int j2);
    void get_timestep_into(int tstp, herm_matrix_timestep<double> &t) const;
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
    friend class herm_matrix_timestep_view<double>;
    void incr_or_set(int tstp, const herm_matrix_timestep<double> &t, cplx alpha, bool incr);
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

/* One timestep of a contour function (cntr/cntr_herm_matrix_timestep_decl.hpp:33-251) in
 * the reference's owned layout (cntr/cntr_herm_matrix_timestep_impl.hpp:95-110):
 *   tstp >= 0:  [ret(tstp, 0..tstp) | tv(tstp, 0..ntau) | les(0..tstp, tstp)]
 *   tstp = -1:  [mat(0..ntau)]
 * The data live in host memory.  The bubbles and distance_norm2 run on the device: each
 * call uploads the strips and hands the library a cntrx_herm whose component pointers are
 * offset so that the packed-layout addresses of timestep tstp (ret row, tv row, les column)
 * fall on them; only timestep tstp is ever addressed through it. */
template <>
class herm_matrix_timestep<double> {
  public:
    typedef std::complex<double> cplx;
    typedef double scalar_type;
    herm_matrix_timestep() : tstp_(-2), ntau_(0), size1_(0), sig_(-1), es_(0) {}
    herm_matrix_timestep(int tstp, int ntau, int size1 = 1) { init(tstp, ntau, size1, -1); }
    herm_matrix_timestep(int tstp, int ntau, int size1, int sig) { init(tstp, ntau, size1, sig); }
    herm_matrix_timestep(int tstp, int ntau, int size1, int size2, int sig) {
        if (size2 != size1) {
            std::fprintf(stderr, "cntr::herm_matrix_timestep: size1 != size2 is not provided\n");
            std::abort();
        }
        init(tstp, ntau, size1, sig);
    }
    herm_matrix_timestep(const herm_matrix_timestep &g)
        : tstp_(g.tstp_), ntau_(g.ntau_), size1_(g.size1_), sig_(g.sig_), es_(g.es_), data_(g.data_) {}
    herm_matrix_timestep &operator=(const herm_matrix_timestep &g) {
        if (this != &g) {
            tstp_ = g.tstp_;
            ntau_ = g.ntau_;
            size1_ = g.size1_;
            sig_ = g.sig_;
            es_ = g.es_;
            data_ = g.data_;
        }
        return *this;
    }
    herm_matrix_timestep(herm_matrix_timestep &&g) noexcept : tstp_(-2), ntau_(0), size1_(0), sig_(-1), es_(0) {
        swap_in(g);
    }
    herm_matrix_timestep &operator=(herm_matrix_timestep &&g) noexcept {
        swap_in(g);
        return *this;
    }

    void resize(int tstp, int ntau, int size1) { init(tstp, ntau, size1, sig_); }
    void clear() { std::fill(data_.begin(), data_.end(), cplx(0.0)); }
    int size1() const { return size1_; }
    int size2() const { return size1_; }
    int tstp() const { return tstp_; }
    int ntau() const { return ntau_; }
    int sig() const { return sig_; }
    void set_sig(int sig) { sig_ = sig; }
    int element_size() const { return es_; }
    int total_size() const { return (int)(data_.size() / (es_ ? es_ : 1)); }

    cplx *retptr(int i) { return &data_[(size_t)i * es_]; }
    cplx *tvptr(int m) { return &data_[(size_t)(tstp_ + 1 + m) * es_]; }
    cplx *lesptr(int i) { return &data_[(size_t)(tstp_ + 1 + ntau_ + 1 + i) * es_]; }
    cplx *matptr(int m) { return &data_[(size_t)m * es_]; }
    const cplx *retptr(int i) const { return &data_[(size_t)i * es_]; }
    const cplx *tvptr(int m) const { return &data_[(size_t)(tstp_ + 1 + m) * es_]; }
    const cplx *lesptr(int i) const { return &data_[(size_t)(tstp_ + 1 + ntau_ + 1 + i) * es_]; }
    const cplx *matptr(int m) const { return &data_[(size_t)m * es_]; }

    // element access (cntr/cntr_herm_matrix_timestep_impl.hpp:518-1093): ret(i, j) needs
    // i or j = tstp, les(i, j) likewise, tv(i, m) needs i = tstp; the other triangle is
    // read through the Hermitian symmetry
    void get_ret(int i, int j, cplx *M) const {
        if (i == tstp_) read(retptr(j), M);
        else { want(j == tstp_, "get_ret"); read_minus_conj(retptr(i), M); }
    }
    void get_les(int i, int j, cplx *M) const {
        if (j == tstp_) read(lesptr(i), M);
        else { want(i == tstp_, "get_les"); read_minus_conj(lesptr(j), M); }
    }
    void get_tv(int i, int m, cplx *M) const {
        want(i == tstp_, "get_tv");
        read(tvptr(m), M);
    }
    void get_mat(int m, cplx *M) const {
        want(tstp_ == -1, "get_mat");
        read(matptr(m), M);
    }
    void get_matminus(int m, cplx *M) const {
        want(tstp_ == -1, "get_matminus");
        read(matptr(ntau_ - m), M);
        if (sig_ == -1)
            for (int e = 0; e < es_; e++) M[e] = -M[e];
    }
    void get_ret(int i, int j, cplx &x) const { get_ret(i, j, &x); }
    void get_les(int i, int j, cplx &x) const { get_les(i, j, &x); }
    void get_tv(int i, int m, cplx &x) const { get_tv(i, m, &x); }
    void get_mat(int m, cplx &x) const { get_mat(m, &x); }
    void get_matminus(int m, cplx &x) const { get_matminus(m, &x); }
    void get_ret(int i, int j, cdmatrix &M) const { get_ret(i, j, sized(M)); }
    void get_les(int i, int j, cdmatrix &M) const { get_les(i, j, sized(M)); }
    void get_tv(int i, int m, cdmatrix &M) const { get_tv(i, m, sized(M)); }
    void get_mat(int m, cdmatrix &M) const { get_mat(m, sized(M)); }
    void get_matminus(int m, cdmatrix &M) const { get_matminus(m, sized(M)); }
    void set_ret(int i, int j, const cdmatrix &M) { write_at(i, j, M.data(), 0); }
    void set_les(int i, int j, const cdmatrix &M) { write_at(i, j, M.data(), 1); }
    void set_tv(int i, int m, const cdmatrix &M) {
        want(i == tstp_, "set_tv");
        std::memcpy((void *)tvptr(m), M.data(), es_ * sizeof(cplx));
    }
    void set_mat(int m, const cdmatrix &M) {
        want(tstp_ == -1, "set_mat");
        std::memcpy((void *)matptr(m), M.data(), es_ * sizeof(cplx));
    }
    void set_ret(int i, int j, cplx x) { write_at(i, j, &x, 0); }
    void set_les(int i, int j, cplx x) { write_at(i, j, &x, 1); }
    void set_tv(int i, int m, cplx x) {
        want(i == tstp_, "set_tv");
        tvptr(m)[0] = x;
    }
    void set_mat(int m, cplx x) {
        want(tstp_ == -1, "set_mat");
        matptr(m)[0] = x;
    }
    // density_matrix (cntr/cntr_herm_matrix_timestep_impl.hpp:1116-1224)
    cplx density_matrix(int tstp) const {
        want(tstp == tstp_, "density_matrix");
        cplx x;
        if (tstp_ == -1) {
            read(matptr(ntau_), &x);
            return -x;
        }
        read(lesptr(tstp_), &x);
        return cplx(0.0, sig_) * x;
    }
    void density_matrix(int tstp, cplx &rho) const { rho = density_matrix(tstp); }
    void density_matrix(int tstp, cdmatrix &M) const {
        want(tstp == tstp_, "density_matrix");
        cplx *m = sized(M);
        if (tstp_ == -1) {
            read(matptr(ntau_), m);
            for (int e = 0; e < es_; e++) m[e] = -m[e];
        } else {
            read(lesptr(tstp_), m);
            for (int e = 0; e < es_; e++) m[e] *= cplx(0.0, sig_);
        }
    }

    // timestep algebra (cntr/cntr_herm_matrix_timestep_impl.hpp:413-470, 1380-1540)
    void set_timestep_zero(int tstp) {
        want(tstp == tstp_, "set_timestep_zero");
        clear();
    }
    void set_timestep(int tstp, const herm_matrix<double> &g) { g.get_timestep_into(tstp, *this); }
    void set_timestep(int tstp, const herm_matrix_timestep &g) {
        want(tstp == tstp_ && g.tstp_ == tstp_ && g.ntau_ == ntau_ && g.size1_ == size1_, "set_timestep");
        data_ = g.data_;
    }
    void incr_timestep(int tstp, const herm_matrix_timestep &g, double weight = 1.0) {
        want(tstp == tstp_ && g.tstp_ == tstp_ && g.ntau_ == ntau_ && g.size1_ == size1_, "incr_timestep");
        for (size_t i = 0; i < data_.size(); i++) data_[i] += weight * g.data_[i];
    }
    void incr_timestep(int tstp, const herm_matrix<double> &g, double weight = 1.0) {
        herm_matrix_timestep t(tstp, ntau_, size1_, sig_);
        g.get_timestep_into(tstp, t);
        incr_timestep(tstp, t, weight);
    }
    void incr(const herm_matrix_timestep &g, double weight = 1.0) { incr_timestep(tstp_, g, weight); }
    void incr(const herm_matrix<double> &g, double weight = 1.0) { incr_timestep(tstp_, g, weight); }
    void smul(int tstp, double weight) {
        want(tstp == tstp_, "smul");
        smul(weight);
    }
    void smul(double weight) {
        for (auto &x : data_) x *= weight;
    }
    // left_multiply: C(t,t') -> weight f(t) C(t,t'); right_multiply: C(t,t') f(t') weight
    // (cntr/cntr_herm_matrix_timestep_impl.hpp:1570-1823); f(-1) acts on the tau side
    void left_multiply(int tstp, const function<double> &f, double weight = 1.0) {
        want(tstp == tstp_, "left_multiply");
        left_multiply(f, weight);
    }
    void right_multiply(int tstp, const function<double> &f, double weight = 1.0) {
        want(tstp == tstp_, "right_multiply");
        right_multiply(f, weight);
    }
    void left_multiply(const function<double> &f, double weight = 1.0);
    void right_multiply(const function<double> &f, double weight = 1.0);
    // element (i1, i2) <- element (j1, j2) of g at timestep tstp
    // (cntr/cntr_herm_matrix_timestep_impl.hpp:1880-2020)
    void set_matrixelement(int tstp, int i1, int i2, const herm_matrix_timestep &g, int j1, int j2) {
        want(tstp == tstp_ && g.tstp_ == tstp_ && g.ntau_ == ntau_, "set_matrixelement");
        const int s = i1 * size1_ + i2, u = j1 * g.size1_ + j2, n = total_size();
        for (int e = 0; e < n; e++) data_[(size_t)e * es_ + s] = g.data_[(size_t)e * g.es_ + u];
    }
    void set_matrixelement(int i1, int i2, const herm_matrix_timestep &g, int j1, int j2) {
        set_matrixelement(tstp_, i1, i2, g, j1, j2);
    }
    void set_matrixelement(int tstp, int i1, int i2, const herm_matrix<double> &g, int j1, int j2) {
        herm_matrix_timestep t(tstp, ntau_, g.size1(), g.sig());
        g.get_timestep_into(tstp, t);
        set_matrixelement(tstp, i1, i2, t, j1, j2);
    }
    void set_matrixelement(int i1, int i2, const herm_matrix<double> &g, int j1, int j2) {
        set_matrixelement(tstp_, i1, i2, g, j1, j2);
    }

    // device side, for the free functions below: a descriptor valid at timestep `tstp`
    // only, and the strips back to the host after a call wrote them
    const cntrx_herm &dev_sync(int tstp) const {
        want(tstp == tstp_, "device call at another timestep");
        const size_t es = (size_t)es_;
        cplx *blk = (cplx *)dev_.get(data_.size() * sizeof(cplx));
        amd_detail::up(blk, data_.data(), data_.size() * sizeof(cplx));
        d_.nt = tstp_ < 0 ? -1 : tstp_;
        d_.ntau = ntau_;
        d_.size1 = size1_;
        d_.sig = sig_;
        if (tstp_ == -1) {
            d_.ret = d_.les = d_.tv = (double *)blk;
            d_.mat = (double *)blk;
        } else {
            const size_t tri = (size_t)tstp_ * (tstp_ + 1) / 2;
            d_.ret = (double *)(blk - tri * es);
            d_.tv = (double *)(blk + (size_t)(tstp_ + 1) * es - (size_t)tstp_ * (ntau_ + 1) * es);
            d_.les = (double *)(blk + (size_t)(tstp_ + 1 + ntau_ + 1) * es - tri * es);
            d_.mat = (double *)blk;
        }
        return d_;
    }
    void dev_pull(int) const {
        amd_detail::down((void *)data_.data(), dev_.get(data_.size() * sizeof(cplx)), data_.size() * sizeof(cplx));
    }

  private:
    friend class herm_matrix<double>;
    friend class herm_matrix_timestep_view<double>;
    void init(int tstp, int ntau, int size1, int sig) {
        tstp_ = tstp;
        ntau_ = ntau;
        size1_ = size1;
        sig_ = sig;
        es_ = size1 * size1;
        const size_t n = tstp >= 0 ? (size_t)2 * (tstp + 1) + ntau + 1 : (size_t)ntau + 1;
        data_.assign(n * es_, cplx(0.0));
    }
    void swap_in(herm_matrix_timestep &g) {
        std::swap(tstp_, g.tstp_);
        std::swap(ntau_, g.ntau_);
        std::swap(size1_, g.size1_);
        std::swap(sig_, g.sig_);
        std::swap(es_, g.es_);
        data_.swap(g.data_);
        dev_.swap(g.dev_);
    }
    void want(bool ok, const char *what) const {
        if (!ok) {
            std::fprintf(stderr, "cntr::herm_matrix_timestep::%s: index outside timestep %d\n", what, tstp_);
            std::abort();
        }
    }
    void write_at(int i, int j, const cplx *M, int les) {
        cplx *x;
        if (les == 0) {
            want(i == tstp_, "set_ret");
            x = retptr(j);
        } else {
            want(j == tstp_, "set_les");
            x = lesptr(i);
        }
        std::memcpy((void *)x, M, es_ * sizeof(cplx));
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
    // z = weight * a b for size1 x size1 blocks
    void block_mult(cplx *z, const cplx *a, const cplx *b, double weight) const {
        std::vector<cplx> t(es_);
        for (int r = 0; r < size1_; r++)
            for (int c = 0; c < size1_; c++) {
                cplx acc = 0.0;
                for (int l = 0; l < size1_; l++) acc += a[r * size1_ + l] * b[l * size1_ + c];
                t[r * size1_ + c] = weight * acc;
            }
        std::memcpy((void *)z, t.data(), es_ * sizeof(cplx));
    }

    int tstp_, ntau_, size1_, sig_, es_;
    std::vector<cplx> data_;
    mutable amd_detail::DevMem dev_;
    mutable cntrx_herm d_{};
};

inline void herm_matrix_timestep<double>::left_multiply(const function<double> &f, double weight) {
    if (tstp_ == -1) {
        for (int m = 0; m <= ntau_; m++) block_mult(matptr(m), f.ptr(-1), matptr(m), weight);
        return;
    }
    for (int j = 0; j <= tstp_; j++) block_mult(retptr(j), f.ptr(tstp_), retptr(j), weight);
    for (int m = 0; m <= ntau_; m++) block_mult(tvptr(m), f.ptr(tstp_), tvptr(m), weight);
    for (int i = 0; i <= tstp_; i++) block_mult(lesptr(i), f.ptr(i), lesptr(i), weight);
}
inline void herm_matrix_timestep<double>::right_multiply(const function<double> &f, double weight) {
    if (tstp_ == -1) {
        for (int m = 0; m <= ntau_; m++) block_mult(matptr(m), matptr(m), f.ptr(-1), weight);
        return;
    }
    for (int j = 0; j <= tstp_; j++) block_mult(retptr(j), retptr(j), f.ptr(j), weight);
    for (int m = 0; m <= ntau_; m++) block_mult(tvptr(m), tvptr(m), f.ptr(-1), weight);
    for (int i = 0; i <= tstp_; i++) block_mult(lesptr(i), lesptr(i), f.ptr(tstp_), weight);
}

/* A view of one timestep of a herm_matrix or a herm_matrix_timestep
 * (cntr/cntr_herm_matrix_timestep_view_decl.hpp:34-194): pointers to its ret row, tv row
 * and les column (or its mat strip for tstp = -1) in the host copy.  Writes through a view
 * of a herm_matrix mark that timestep for upload before the next solver call. */
template <>
class herm_matrix_timestep_view<double> {
  public:
    typedef std::complex<double> cplx;
    typedef double scalar_type;
    herm_matrix_timestep_view()
        : tstp_(-2), ntau_(0), size1_(0), sig_(-1), es_(0), ret_(nullptr), tv_(nullptr), les_(nullptr),
          mat_(nullptr), owner_(nullptr) {}
    herm_matrix_timestep_view(int tstp, herm_matrix<double> &g) { set_to_data(tstp, g); }
    herm_matrix_timestep_view(herm_matrix_timestep<double> &g) { set_to_data(g); }
    herm_matrix_timestep_view(int tstp, herm_matrix_timestep<double> &g) {
        g.want(tstp == g.tstp(), "view");
        set_to_data(g);
    }
    void set_to_data(int tstp, herm_matrix<double> &g) {
        tstp_ = tstp;
        ntau_ = g.ntau();
        size1_ = g.size1();
        sig_ = g.sig();
        es_ = g.element_size();
        owner_ = &g;
        if (tstp == -1) {
            mat_ = g.matptr(0);
            ret_ = tv_ = les_ = nullptr;
        } else {
            ret_ = g.retptr(tstp, 0);
            tv_ = g.tvptr(tstp, 0);
            les_ = g.lesptr(0, tstp);
            mat_ = nullptr;
        }
    }
    void set_to_data(herm_matrix_timestep<double> &g) {
        tstp_ = g.tstp();
        ntau_ = g.ntau();
        size1_ = g.size1();
        sig_ = g.sig();
        es_ = g.element_size();
        owner_ = nullptr;
        if (tstp_ == -1) {
            mat_ = g.matptr(0);
            ret_ = tv_ = les_ = nullptr;
        } else {
            ret_ = g.retptr(0);
            tv_ = g.tvptr(0);
            les_ = g.lesptr(0);
            mat_ = nullptr;
        }
    }
    int size1() const { return size1_; }
    int size2() const { return size1_; }
    int ntau() const { return ntau_; }
    int tstp() const { return tstp_; }
    int sig() const { return sig_; }
    cplx *retptr(int i) { return ret_ + (size_t)i * es_; }
    cplx *tvptr(int m) { return tv_ + (size_t)m * es_; }
    cplx *lesptr(int i) { return les_ + (size_t)i * es_; }
    cplx *matptr(int m) { return mat_ + (size_t)m * es_; }

    // operations on the viewed data (cntr/cntr_herm_matrix_timestep_view_impl.hpp)
    void set_timestep_zero(int tstp) {
        check_t(tstp);
        strips([](cplx *x, size_t n, const cplx *) {
            for (size_t i = 0; i < n; i++) x[i] = 0.0;
        });
    }
    void get_data(herm_matrix_timestep_view &g) {
        check_like(g);
        strips([](cplx *x, size_t n, const cplx *y) { std::memcpy((void *)x, y, n * sizeof(cplx)); }, &g);
    }
    void incr_timestep(herm_matrix_timestep_view &g, cplx alpha) {
        check_like(g);
        strips([alpha](cplx *x, size_t n, const cplx *y) {
            for (size_t i = 0; i < n; i++) x[i] += alpha * y[i];
        }, &g);
    }
    void incr_timestep(herm_matrix_timestep_view &g, double alpha = 1.0) { incr_timestep(g, cplx(alpha, 0.0)); }
    template <class GG>
    void incr_timestep(GG &g, cplx alpha) {
        herm_matrix_timestep_view v(tstp_, g);
        incr_timestep(v, alpha);
    }
    template <class GG>
    void incr_timestep(GG &g, double alpha = 1.0) {
        incr_timestep(g, cplx(alpha, 0.0));
    }
    template <class GG>
    void set_timestep(int tstp, GG &g) {
        check_t(tstp);
        herm_matrix_timestep_view v(tstp, g);
        get_data(v);
    }
    void smul(int tstp, double weight) {
        check_t(tstp);
        strips([weight](cplx *x, size_t n, const cplx *) {
            for (size_t i = 0; i < n; i++) x[i] *= weight;
        });
    }
    cplx density_matrix(int tstp) const {
        check_t(tstp);
        if (tstp_ == -1) return -mat_[(size_t)ntau_ * es_];
        return cplx(0.0, sig_) * les_[(size_t)tstp_ * es_];
    }
    void density_matrix(int tstp, cdmatrix &M) {
        check_t(tstp);
        M.resize(size1_, size1_);
        const cplx *x = tstp_ == -1 ? mat_ + (size_t)ntau_ * es_ : les_ + (size_t)tstp_ * es_;
        const cplx f = tstp_ == -1 ? cplx(-1.0, 0.0) : cplx(0.0, sig_);
        for (int e = 0; e < es_; e++) M.data()[e] = f * x[e];
    }
    void set_matrixelement(int i1, int i2, herm_matrix_timestep_view &g, int j1, int j2) {
        if (g.tstp_ != tstp_ || g.ntau_ != ntau_) bad("set_matrixelement");
        const int s = i1 * size1_ + i2, u = j1 * g.size1_ + j2;
        touch();
        if (tstp_ == -1) {
            for (int m = 0; m <= ntau_; m++) mat_[(size_t)m * es_ + s] = g.mat_[(size_t)m * g.es_ + u];
            return;
        }
        for (int j = 0; j <= tstp_; j++) ret_[(size_t)j * es_ + s] = g.ret_[(size_t)j * g.es_ + u];
        for (int m = 0; m <= ntau_; m++) tv_[(size_t)m * es_ + s] = g.tv_[(size_t)m * g.es_ + u];
        for (int i = 0; i <= tstp_; i++) les_[(size_t)i * es_ + s] = g.les_[(size_t)i * g.es_ + u];
    }
    template <class GG>
    void set_matrixelement(int i1, int i2, GG &g, int j1, int j2) {
        herm_matrix_timestep_view v(tstp_, g);
        set_matrixelement(i1, i2, v, j1, j2);
    }

  private:
    void bad(const char *what) const {
        std::fprintf(stderr, "cntr::herm_matrix_timestep_view::%s: shapes or timesteps differ\n", what);
        std::abort();
    }
    void check_t(int tstp) const {
        if (tstp != tstp_) bad("timestep");
    }
    void check_like(const herm_matrix_timestep_view &g) const {
        if (g.tstp_ != tstp_ || g.ntau_ != ntau_ || g.size1_ != size1_) bad("operand");
    }
    void touch();
    template <class F>
    void strips(F f, herm_matrix_timestep_view *g = nullptr) {
        touch();
        if (tstp_ == -1) {
            f(mat_, (size_t)(ntau_ + 1) * es_, g ? g->mat_ : nullptr);
            return;
        }
        f(ret_, (size_t)(tstp_ + 1) * es_, g ? g->ret_ : nullptr);
        f(tv_, (size_t)(ntau_ + 1) * es_, g ? g->tv_ : nullptr);
        f(les_, (size_t)(tstp_ + 1) * es_, g ? g->les_ : nullptr);
    }
    int tstp_, ntau_, size1_, sig_, es_;
    cplx *ret_, *tv_, *les_, *mat_;
    herm_matrix<double> *owner_;
};

// ---- herm_matrix <-> timestep objects ----
inline void herm_matrix<double>::get_timestep_into(int tstp, herm_matrix_timestep<double> &t) const {
    if (t.tstp_ != tstp || t.ntau_ != ntau_ || t.size1_ != size1_) t.init(tstp, ntau_, size1_, sig_);
    t.sig_ = sig_;
    if (tstp == -1) {
        std::memcpy((void *)t.matptr(0), matptr(0), (size_t)(ntau_ + 1) * es_ * sizeof(cplx));
        return;
    }
    std::memcpy((void *)t.retptr(0), retptr(tstp, 0), (size_t)(tstp + 1) * es_ * sizeof(cplx));
    std::memcpy((void *)t.tvptr(0), tvptr(tstp, 0), (size_t)(ntau_ + 1) * es_ * sizeof(cplx));
    std::memcpy((void *)t.lesptr(0), lesptr(0, tstp), (size_t)(tstp + 1) * es_ * sizeof(cplx));
}
inline void herm_matrix<double>::get_timestep(int tstp, herm_matrix_timestep<double> &t) const {
    get_timestep_into(tstp, t);
}
inline void herm_matrix<double>::set_timestep(int tstp, const herm_matrix_timestep<double> &t) {
    incr_or_set(tstp, t, cplx(1.0, 0.0), false);
}
inline void herm_matrix<double>::incr_timestep(int tstp, const herm_matrix_timestep<double> &t, cplx alpha) {
    incr_or_set(tstp, t, alpha, true);
}
inline void herm_matrix<double>::incr_or_set(int tstp, const herm_matrix_timestep<double> &t, cplx alpha,
                                              bool incr) {
    if (t.tstp_ != tstp || t.ntau_ != ntau_ || t.size1_ != size1_) {
        std::fprintf(stderr, "cntr::herm_matrix::set_timestep: timestep object of another shape\n");
        std::abort();
    }
    auto op = [alpha, incr](cplx *x, const cplx *y, size_t n) {
        for (size_t i = 0; i < n; i++) x[i] = incr ? x[i] + alpha * y[i] : y[i];
    };
    if (tstp == -1) {
        op(matptr(0), t.matptr(0), (size_t)(ntau_ + 1) * es_);
        return;
    }
    op(retptr(tstp, 0), t.retptr(0), (size_t)(tstp + 1) * es_);
    op(tvptr(tstp, 0), t.tvptr(0), (size_t)(ntau_ + 1) * es_);
    op(lesptr(0, tstp), t.lesptr(0), (size_t)(tstp + 1) * es_);
}
inline void herm_matrix<double>::set_matrixelement(int tstp, int i1, int i2, const herm_matrix_timestep<double> &g,
                                                    int j1, int j2) {
    herm_matrix_timestep<double> t(tstp, ntau_, size1_, sig_);
    get_timestep_into(tstp, t);
    t.set_matrixelement(tstp, i1, i2, g, j1, j2);
    set_timestep(tstp, t);
}
inline void herm_matrix_timestep_view<double>::touch() {
    if (owner_) {
        if (tstp_ == -1) owner_->matptr(0);
        else owner_->retptr(tstp_, 0);
    }
}

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

/* ---- bubbles (cntr/cntr_bubble_decl.hpp:62-301) ----
 * C, A and B may each be a herm_matrix or a herm_matrix_timestep of timestep tstp, as in
 * the reference's templates (GGC, GGA, GGB). */
namespace amd_detail {
inline const cntrx_herm &dev_at(const herm_matrix<double> &G, int) { return G.dev_sync(); }
inline const cntrx_herm &dev_at(const herm_matrix_timestep<double> &G, int tstp) { return G.dev_sync(tstp); }
}  // namespace amd_detail
template <class GGC, class GGA, class GGB>
inline void Bubble1(int tstp, GGC &C, int c1, int c2, GGA &A, GGA &Acc, int a1, int a2, GGB &B, GGB &Bcc,
                    int b1, int b2) {
    cntrx_herm c = amd_detail::dev_at(C, tstp), a = amd_detail::dev_at(A, tstp), ac = amd_detail::dev_at(Acc, tstp),
               b = amd_detail::dev_at(B, tstp), bc = amd_detail::dev_at(Bcc, tstp);
    amd_detail::check(cntrx_bubble1(tstp, &c, c1, c2, &a, &ac, a1, a2, &b, &bc, b1, b2, nullptr), "Bubble1");
    C.dev_pull(tstp);
}
template <class GGC, class GGA, class GGB>
inline void Bubble1(int tstp, GGC &C, int c1, int c2, GGA &A, int a1, int a2, GGB &B, int b1, int b2) {
    Bubble1(tstp, C, c1, c2, A, A, a1, a2, B, B, b1, b2);
}
template <class GGC, class GGA, class GGB>
inline void Bubble1(int tstp, GGC &C, GGA &A, GGA &Acc, GGB &B, GGB &Bcc) {
    Bubble1(tstp, C, 0, 0, A, Acc, 0, 0, B, Bcc, 0, 0);
}
template <class GGC, class GGA, class GGB>
inline void Bubble1(int tstp, GGC &C, GGA &A, GGB &B) {
    Bubble1(tstp, C, 0, 0, A, A, 0, 0, B, B, 0, 0);
}
template <class GGC, class GGA, class GGB>
inline void Bubble2(int tstp, GGC &C, int c1, int c2, GGA &A, GGA &Acc, int a1, int a2, GGB &B, GGB &Bcc,
                    int b1, int b2) {
    cntrx_herm c = amd_detail::dev_at(C, tstp), a = amd_detail::dev_at(A, tstp), ac = amd_detail::dev_at(Acc, tstp),
               b = amd_detail::dev_at(B, tstp), bc = amd_detail::dev_at(Bcc, tstp);
    amd_detail::check(cntrx_bubble2(tstp, &c, c1, c2, &a, &ac, a1, a2, &b, &bc, b1, b2, nullptr), "Bubble2");
    C.dev_pull(tstp);
}
template <class GGC, class GGA, class GGB>
inline void Bubble2(int tstp, GGC &C, int c1, int c2, GGA &A, int a1, int a2, GGB &B, int b1, int b2) {
    Bubble2(tstp, C, c1, c2, A, A, a1, a2, B, B, b1, b2);
}
template <class GGC, class GGA, class GGB>
inline void Bubble2(int tstp, GGC &C, GGA &A, GGA &Acc, GGB &B, GGB &Bcc) {
    Bubble2(tstp, C, 0, 0, A, Acc, 0, 0, B, Bcc, 0, 0);
}
template <class GGC, class GGA, class GGB>
inline void Bubble2(int tstp, GGC &C, GGA &A, GGB &B) {
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
// g1 and g2 may each be a herm_matrix or a herm_matrix_timestep of timestep tstp
template <class GG1, class GG2>
inline double distance_norm2(int tstp, GG1 &g1, GG2 &g2) {
    const cntrx_herm a = amd_detail::dev_at(g1, tstp), b = amd_detail::dev_at(g2, tstp);
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
// one timestep of it (cntr/cntr_equilibrium_decl.hpp:82-85)
inline void green_from_H(int tstp, herm_matrix_timestep<double> &G, double mu, const cdmatrix &eps, double beta,
                         double h) {
    herm_matrix<double> full(tstp < 0 ? -1 : tstp, G.ntau(), G.size1(), G.sig());
    green_from_H(full, mu, eps, beta, h);
    full.get_timestep(tstp, G);
}
// time-dependent H (cntr/cntr_equilibrium_decl.hpp:79-85, 93-98): commutator-free exponential
// propagators of order cf_order (2, 4) with H interpolated at order SolveOrder; without fixHam
// H(t > SolveOrder) is replaced by its extrapolation, in eps, as the reference does
namespace amd_detail {
inline void green_td(herm_matrix<double> &G, double mu, function<double> &eps, double beta, double h, bool fixHam,
                     int SolveOrder, int cf_order, int tstp) {
    const cntrx_herm &g = G.dev_sync();
    cntrx_func f;
    f.nt = eps.nt();
    f.size1 = eps.size1();
    f.data = (double *)eps.ptr(-1);  // host data; ptr() marks eps for upload, which fixHam = false needs
    check(cntrx_green_from_H_td(&g, mu, &f, beta, h, SolveOrder, cf_order, fixHam ? 1 : 0, tstp, nullptr),
          "green_from_H");
    if (tstp == CNTRX_ALL_TIMESTEPS) G.dev_pull_all();
    else G.dev_pull(tstp);
}
}  // namespace amd_detail
inline void green_from_H(herm_matrix<double> &G, double mu, function<double> &eps, double beta, double h,
                         int SolveOrder = MAX_SOLVE_ORDER, int cf_order = 4) {
    amd_detail::green_td(G, mu, eps, beta, h, true, SolveOrder, cf_order, CNTRX_ALL_TIMESTEPS);
}
inline void green_from_H(herm_matrix<double> &G, double mu, function<double> &eps, double beta, double h,
                         bool fixHam, int SolveOrder, int cf_order) {
    amd_detail::green_td(G, mu, eps, beta, h, fixHam, SolveOrder, cf_order, CNTRX_ALL_TIMESTEPS);
}
inline void green_from_H(int tstp, herm_matrix<double> &G, double mu, function<double> &eps, double beta, double h,
                         bool fixHam = false, int SolveOrder = MAX_SOLVE_ORDER, int cf_order = 4) {
    amd_detail::green_td(G, mu, eps, beta, h, fixHam, SolveOrder, cf_order, tstp);
}
inline void green_from_H(int tstp, herm_matrix_timestep<double> &G, double mu, function<double> &eps, double beta,
                         double h, bool fixHam = false, int SolveOrder = MAX_SOLVE_ORDER, int cf_order = 4) {
    herm_matrix<double> full(tstp < 0 ? -1 : tstp, G.ntau(), G.size1(), G.sig());
    amd_detail::green_td(full, mu, eps, beta, h, fixHam, SolveOrder, cf_order, tstp);
    full.get_timestep(tstp, G);
}

}  // namespace cntr

#endif
