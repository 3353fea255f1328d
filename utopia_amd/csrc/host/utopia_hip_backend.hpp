// utopia HIP backend: HipVector / HipCrsMatrix.
//
// The tensor types the solver templates run on, in place of PetscVector
// (backend/petsc/utopia_petsc_Vector.hpp:39-930) and PetscMatrix
// (backend/petsc/utopia_petsc_Matrix.hpp:103-1028). Every numeric method is one
// call into the C ABI of include/utopia_hip.h; no HIP type appears here.
// Errors from the C ABI are raised as utopia::HipError.
#pragma once

#include <cmath>
#include <cstdint>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "utopia_hip.h"

namespace utopia {

// Backend tag beside BLAS=1, PETSC=10, TRILINOS=10000, KOKKOS=100000, SERIAL_HOMEMADE=200000
// (utopia_core/utopia_Traits.hpp:12-21).
enum { HIP = 1000000 };

class HipError : public std::runtime_error {
public:
    HipError(int code, const std::string &what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

inline void hip_check(int status, const char *what) {
    if (status != UHIP_OK) throw HipError(status, std::string(what) + ": " + uhip_last_error());
}

// One device + one stream + reduction workspace; shared by everything made from it.
class HipContext {
public:
    explicit HipContext(int device = 0, void *stream = nullptr) {
        hip_check(uhip_ctx_create(device, stream, &ctx_), "uhip_ctx_create");
    }
    ~HipContext() { uhip_ctx_destroy(ctx_); }
    HipContext(const HipContext &) = delete;
    HipContext &operator=(const HipContext &) = delete;
    uhip_ctx get() const { return ctx_; }
    double *scalars() const { return uhip_ctx_scalars(ctx_); }
    void sync() const { hip_check(uhip_sync(ctx_), "uhip_sync"); }
    void read_scalars(const double *src, int n, double *dst) const {
        hip_check(uhip_read_scalars(ctx_, src, n, dst), "uhip_read_scalars");
    }
    void fetch_scalars(int slot, const double *src, int n) const {
        hip_check(uhip_fetch_scalars(ctx_, slot, src, n), "uhip_fetch_scalars");
    }
    void fetch_wait(int slot, int n, double *dst) const {
        hip_check(uhip_fetch_wait(ctx_, slot, n, dst), "uhip_fetch_wait");
    }

private:
    uhip_ctx ctx_ = nullptr;
};

using HipContextPtr = std::shared_ptr<HipContext>;

// utopia's operation tags (utopia_core/expressions/utopia_Operators.hpp:53-401) are only declared
// here: the surface takes them by reference, so a translation unit that includes utopia's core
// passes utopia's own tags, and one that does not defines them itself.
class Plus;
class Min;
class Max;
class Abs;

// One GPU driven by one process: the communicator of every HIP tensor
// (utopia_core/interfaces/utopia_Communicator.hpp:18-300, the SelfCommunicator case).
class HipCommunicator {
public:
    int rank() const { return 0; }
    int size() const { return 1; }
    void barrier() const {}
    template <typename T>
    T sum(const T &v) const { return v; }
    template <typename T>
    T min(const T &v) const { return v; }
    template <typename T>
    T max(const T &v) const { return v; }
    static const HipCommunicator &get() {
        static const HipCommunicator c;
        return c;
    }
};

// Vector layout (utopia_core/interfaces/utopia_Layout.hpp): local and global sizes agree on one GPU.
struct HipLayout {
    int64_t local_size_ = 0, size_ = 0;
    int64_t local_size() const { return local_size_; }
    int64_t size() const { return size_; }
    const HipCommunicator &comm() const { return HipCommunicator::get(); }
    bool same(const HipLayout &o) const { return local_size_ == o.local_size_ && size_ == o.size_; }
};

// Range of locally owned indices (utopia_core/utopia_Range.hpp)
struct HipRange {
    int64_t begin_ = 0, end_ = 0;
    int64_t begin() const { return begin_; }
    int64_t end() const { return end_; }
    int64_t extent() const { return end_ - begin_; }
    bool inside(int64_t i) const { return begin_ <= i && i < end_; }
};

template <typename T>
struct HipDtype;
template <>
struct HipDtype<double> {
    static constexpr int value = UHIP_F64;
};
template <>
struct HipDtype<float> {
    static constexpr int value = UHIP_F32;
};

// ---------------------------------------------------------------------------
// HipVectorT: contiguous device vector (FP64: HipVector, FP32: HipVector32).
// Deep copy on copy-construction (PetscVector semantics, utopia_petsc_Vector.hpp:441-458),
// move supported, wrap() for caller-owned memory (no free).
// ---------------------------------------------------------------------------
template <typename T>
class HipVectorT {
public:
    using Scalar = T;
    using SizeType = int64_t;
    static constexpr int dtype = HipDtype<T>::value;

    HipVectorT() = default;
    HipVectorT(HipContextPtr ctx, int64_t n) : ctx_(std::move(ctx)) { allocate(n); }
    HipVectorT(HipContextPtr ctx, int64_t n, T value) : ctx_(std::move(ctx)) {
        allocate(n);
        set(value);
    }
    ~HipVectorT() { release(); }

    HipVectorT(const HipVectorT &o) : ctx_(o.ctx_) {
        if (o.n_ > 0 || o.d_) {
            allocate(o.n_);
            copy_from(o);
        }
    }
    HipVectorT(HipVectorT &&o) noexcept : ctx_(std::move(o.ctx_)), d_(o.d_), n_(o.n_), own_(o.own_) {
        o.d_ = nullptr;
        o.n_ = 0;
    }
    HipVectorT &operator=(const HipVectorT &o) {
        if (this == &o) return *this;
        if (!ctx_) ctx_ = o.ctx_;
        if (n_ != o.n_ || !d_) {
            release();
            allocate(o.n_);
        }
        copy_from(o);
        return *this;
    }
    HipVectorT &operator=(HipVectorT &&o) noexcept {
        if (this == &o) return *this;
        release();
        ctx_ = std::move(o.ctx_);
        d_ = o.d_;
        n_ = o.n_;
        own_ = o.own_;
        o.d_ = nullptr;
        o.n_ = 0;
        return *this;
    }

    // PetscVector::wrap: caller-owned device memory, never freed here.
    void wrap(HipContextPtr ctx, T *ptr, int64_t n) {
        release();
        ctx_ = std::move(ctx);
        d_ = ptr;
        n_ = n;
        own_ = false;
    }

    void zeros(int64_t n) { values(n, T(0)); }
    void values(int64_t n, T v) {
        if (n != n_ || !d_) {
            release();
            allocate(n);
        }
        set(v);
    }
    void zeros_like(const HipVectorT &o) {
        if (!ctx_) ctx_ = o.ctx_;
        zeros(o.size());
    }

    int64_t size() const { return n_; }
    // layout / range / comm of PetscVector (utopia_petsc_Vector.hpp:150-152,289): one GPU owns every entry
    int64_t local_size() const { return n_; }
    HipLayout layout() const { return HipLayout{n_, n_}; }
    HipRange range() const { return HipRange{0, n_}; }
    const HipCommunicator &comm() const { return HipCommunicator::get(); }
    void zeros(const HipLayout &l) { zeros(l.local_size()); }
    void values(const HipLayout &l, T v) { values(l.local_size(), v); }
    void construct(const HipVectorT &o) { *this = o; }
    void assign(const HipVectorT &o) { *this = o; }
    void clear() { release(); }
    bool empty() const { return d_ == nullptr; }
    T *data() { return d_; }
    const T *data() const { return d_; }
    const HipContextPtr &context() const { return ctx_; }
    uhip_ctx ctx() const { return ctx_->get(); }
    bool is_alias(const HipVectorT &o) const { return d_ == o.d_; }

    // BLAS1 -----------------------------------------------------------------
    void set(T v);
    void copy_from(const HipVectorT &o);
    void axpy(T a, const HipVectorT &x);           // this += a x
    void axpby(T a, const HipVectorT &x, T b);     // this = a x + b this
    void scale(T a);
    void shift(T a);
    void abs();
    void e_mul(const HipVectorT &x);               // this = this .* x
    void e_div(const HipVectorT &x);               // this = this ./ x
    void e_min(const HipVectorT &x);               // this = min(this, x)  (VecPointwiseMin)
    void e_max(const HipVectorT &x);               // this = max(this, x)  (VecPointwiseMax)
    void reciprocal(T num);                        // VecReciprocal, then * num unless num == 1
    void replace(T from, T to);                    // entries equal to from become to
    double dot(const HipVectorT &y) const;
    double norm2() const;
    double norm_infty() const;
    double sum() const;
    double min() const;
    double max() const;
    bool has_nan_or_inf() const;
    // Reducible / Transformable with utopia's tags (utopia_petsc_Vector.hpp:338-397)
    double reduce(const Plus &) const { return sum(); }
    double reduce(const Min &) const { return min(); }
    double reduce(const Max &) const { return max(); }
    void transform(const Abs &) { abs(); }

    // Host mirror for utopia's local views (Read / Write / ReadAndWrite locks, l_get / l_set / l_add;
    // utopia_petsc_Vector.hpp:160-250, utopia_core/utopia_DeviceView.hpp:67-111): a lock copies the
    // vector to the host, the element accessors work on that copy, and a write unlock copies it back.
    // For element-wise set-up code written against those views; the solvers' hot loops run in kernels.
    void read_lock() const { mirror_down(); }
    void read_unlock() const { mirror_.clear(); }
    template <class Mode>
    void write_lock(Mode) { mirror_.assign((size_t)n_, T(0)); }
    template <class Mode>
    void write_unlock(Mode) { mirror_up(); }
    template <class Mode>
    void read_and_write_lock(Mode) { mirror_down(); }
    template <class Mode>
    void read_and_write_unlock(Mode) { mirror_up(); }
    T l_get(int64_t i) const { return mirror_.at((size_t)i); }
    void l_set(int64_t i, T v) { mirror_.at((size_t)i) = v; }
    void l_add(int64_t i, T v) { mirror_.at((size_t)i) += v; }
    T get(int64_t i) const { return l_get(i); }
    void set(int64_t i, T v) { l_set(i, v); }

    void upload(const T *host) { hip_check(uhip_memcpy_htod(ctx(), d_, host, sizeof(T) * n_), "upload"); }
    void download(T *host) const { hip_check(uhip_memcpy_dtoh(ctx(), host, d_, sizeof(T) * n_), "download"); }
    std::vector<T> to_host() const {
        std::vector<T> h((size_t)n_);
        download(h.data());
        return h;
    }

private:
    void mirror_down() const {
        mirror_.resize((size_t)n_);
        if (n_ > 0) download(mirror_.data());
    }
    void mirror_up() {
        if (n_ > 0 && mirror_.size() == (size_t)n_) upload(mirror_.data());
        mirror_.clear();
    }
    void allocate(int64_t n) {
        if (!ctx_) throw HipError(UHIP_ERR_INVALID, "HipVector: no context");
        void *p = nullptr;
        hip_check(uhip_malloc(ctx(), &p, sizeof(T) * (size_t)(n > 0 ? n : 1)), "HipVector alloc");
        d_ = (T *)p;
        n_ = n;
        own_ = true;
    }
    void release() {
        if (d_ && own_ && ctx_) uhip_free(ctx(), d_);
        d_ = nullptr;
        n_ = 0;
        own_ = true;
    }
    double reduce(int op) const {
        double *s = ctx_->scalars() + (UHIP_NUM_SCALARS - 1);
        int st = dtype == UHIP_F64 ? uhip_vec_reduce_f64(ctx(), n_, (const double *)d_, op, s)
                                   : uhip_vec_reduce_f32(ctx(), n_, (const float *)d_, op, s);
        hip_check(st, "reduce");
        double h = 0;
        ctx_->read_scalars(s, 1, &h);
        return h;
    }

    HipContextPtr ctx_;
    T *d_ = nullptr;
    int64_t n_ = 0;
    bool own_ = true;
    mutable std::vector<T> mirror_;  // host copy while a lock is held
};

using HipVector = HipVectorT<double>;
using HipVector32 = HipVectorT<float>;

#define UTOPIA_HIP_BLAS1(T, SUF)                                                                              \
    template <>                                                                                               \
    inline void HipVectorT<T>::set(T v) {                                                                     \
        hip_check(uhip_vec_set_##SUF(ctx(), n_, v, d_), "set");                                               \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::copy_from(const HipVectorT<T> &o) {                                            \
        hip_check(uhip_vec_copy_##SUF(ctx(), n_, o.d_, d_), "copy");                                          \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::axpy(T a, const HipVectorT<T> &x) {                                            \
        hip_check(uhip_vec_axpy_##SUF(ctx(), n_, a, x.d_, d_), "axpy");                                       \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::axpby(T a, const HipVectorT<T> &x, T b) {                                      \
        hip_check(uhip_vec_axpby_##SUF(ctx(), n_, a, x.d_, b, d_), "axpby");                                  \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::scale(T a) {                                                                   \
        hip_check(uhip_vec_scale_##SUF(ctx(), n_, a, d_), "scale");                                           \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::shift(T a) {                                                                   \
        hip_check(uhip_vec_shift_##SUF(ctx(), n_, a, d_), "shift");                                           \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::abs() {                                                                        \
        hip_check(uhip_vec_abs_##SUF(ctx(), n_, d_, d_), "abs");                                              \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::e_mul(const HipVectorT<T> &x) {                                                \
        hip_check(uhip_vec_emul_##SUF(ctx(), n_, d_, x.d_, d_), "e_mul");                                     \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::e_div(const HipVectorT<T> &x) {                                                \
        hip_check(uhip_vec_ediv_##SUF(ctx(), n_, d_, x.d_, d_), "e_div");                                     \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::e_min(const HipVectorT<T> &x) {                                                \
        hip_check(uhip_vec_emin_##SUF(ctx(), n_, d_, x.d_, d_), "e_min");                                     \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::e_max(const HipVectorT<T> &x) {                                                \
        hip_check(uhip_vec_emax_##SUF(ctx(), n_, d_, x.d_, d_), "e_max");                                     \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::reciprocal(T num) {                                                            \
        hip_check(uhip_vec_reciprocal_##SUF(ctx(), n_, num, d_, d_), "reciprocal");                           \
    }                                                                                                         \
    template <>                                                                                               \
    inline void HipVectorT<T>::replace(T from, T to) {                                                        \
        hip_check(uhip_vec_replace_##SUF(ctx(), n_, from, to, d_), "replace");                                \
    }                                                                                                         \
    template <>                                                                                               \
    inline double HipVectorT<T>::dot(const HipVectorT<T> &y) const {                                          \
        double *s = ctx_->scalars() + (UHIP_NUM_SCALARS - 1);                                                 \
        hip_check(uhip_vec_dot_##SUF(ctx(), n_, d_, y.d_, s), "dot");                                         \
        double h = 0;                                                                                         \
        ctx_->read_scalars(s, 1, &h);                                                                         \
        return h;                                                                                             \
    }                                                                                                         \
    template <>                                                                                               \
    inline double HipVectorT<T>::norm2() const {                                                              \
        double *s = ctx_->scalars() + (UHIP_NUM_SCALARS - 1);                                                 \
        hip_check(uhip_vec_norm2_##SUF(ctx(), n_, d_, s), "norm2");                                           \
        double h = 0;                                                                                         \
        ctx_->read_scalars(s, 1, &h);                                                                         \
        return h;                                                                                             \
    }

UTOPIA_HIP_BLAS1(double, f64)
UTOPIA_HIP_BLAS1(float, f32)
#undef UTOPIA_HIP_BLAS1

template <typename T>
double HipVectorT<T>::norm_infty() const {
    return reduce(3);
}
template <typename T>
double HipVectorT<T>::sum() const {
    return reduce(0);
}
template <typename T>
double HipVectorT<T>::min() const {
    return reduce(1);
}
template <typename T>
double HipVectorT<T>::max() const {
    return reduce(2);
}
template <typename T>
bool HipVectorT<T>::has_nan_or_inf() const {
    return reduce(4) > 0;
}

template <typename T>
inline double dot(const HipVectorT<T> &a, const HipVectorT<T> &b) {
    return a.dot(b);
}
template <typename T>
inline double norm2(const HipVectorT<T> &a) {
    return a.norm2();
}

// ---------------------------------------------------------------------------
// HipCrsMatrixT: CSR on the device (int32 indices) + row-block SpMV plan.
// ---------------------------------------------------------------------------
template <typename T>
class HipCrsMatrixT {
public:
    using Scalar = T;
    using SizeType = int64_t;
    using Vector = HipVectorT<T>;
    static constexpr int dtype = HipDtype<T>::value;

    HipCrsMatrixT() { std::memset(&m_, 0, sizeof(m_)); }
    ~HipCrsMatrixT() { release(); }
    HipCrsMatrixT(const HipCrsMatrixT &) = delete;
    HipCrsMatrixT &operator=(const HipCrsMatrixT &) = delete;
    HipCrsMatrixT(HipCrsMatrixT &&o) noexcept : ctx_(std::move(o.ctx_)), m_(o.m_), own_(o.own_) {
        std::memset(&o.m_, 0, sizeof(o.m_));
    }
    HipCrsMatrixT &operator=(HipCrsMatrixT &&o) noexcept {
        if (this != &o) {
            release();
            ctx_ = std::move(o.ctx_);
            m_ = o.m_;
            own_ = o.own_;
            std::memset(&o.m_, 0, sizeof(o.m_));
        }
        return *this;
    }

    // Copy host CSR arrays to the device (PetscMatrix::crs, utopia_petsc_Matrix.hpp:362).
    void from_host(HipContextPtr ctx, int32_t rows, int32_t cols, const int32_t *rowptr, const int32_t *col,
                   const T *val, int32_t block_nnz = 0) {
        release();
        ctx_ = std::move(ctx);
        own_ = true;
        m_.nrows = rows;
        m_.ncols = cols;
        m_.nnz = rowptr[rows];
        m_.dtype = dtype;
        void *p = nullptr;
        hip_check(uhip_malloc(c(), &p, sizeof(int32_t) * ((size_t)rows + 1)), "alloc rowptr");
        m_.rowptr = (int32_t *)p;
        hip_check(uhip_malloc(c(), &p, sizeof(int32_t) * (size_t)(m_.nnz > 0 ? m_.nnz : 1)), "alloc col");
        m_.col = (int32_t *)p;
        hip_check(uhip_malloc(c(), &p, sizeof(T) * (size_t)(m_.nnz > 0 ? m_.nnz : 1)), "alloc val");
        m_.val = p;
        hip_check(uhip_memcpy_htod(c(), m_.rowptr, rowptr, sizeof(int32_t) * ((size_t)rows + 1)), "h2d rowptr");
        hip_check(uhip_memcpy_htod(c(), m_.col, col, sizeof(int32_t) * (size_t)m_.nnz), "h2d col");
        hip_check(uhip_memcpy_htod(c(), m_.val, val, sizeof(T) * (size_t)m_.nnz), "h2d val");
        plan(block_nnz);
    }

    // PetscMatrix::wrap (utopia_petsc_Matrix.hpp:651-659): device arrays owned by the caller.
    void wrap(HipContextPtr ctx, int32_t rows, int32_t cols, int64_t nnz, int32_t *rowptr, int32_t *col, T *val,
              int32_t block_nnz = 0) {
        release();
        ctx_ = std::move(ctx);
        own_ = false;
        m_.nrows = rows;
        m_.ncols = cols;
        m_.nnz = nnz;
        m_.rowptr = rowptr;
        m_.col = col;
        m_.val = val;
        m_.dtype = dtype;
        plan(block_nnz);
    }

    // Take ownership of arrays produced by the C ABI (transpose / ptap).
    void adopt(HipContextPtr ctx, const uhip_csr &m, int32_t block_nnz = 0) {
        release();
        ctx_ = std::move(ctx);
        own_ = true;
        m_ = m;
        // the plan belongs to the source descriptor: start from none
        m_.blocks = nullptr;
        m_.slice_off = nullptr;
        m_.sell_col = nullptr;
        m_.sell_val = nullptr;
        m_.dia_off = nullptr;
        m_.dia_val = nullptr;
        m_.dia_mask = nullptr;
        m_.nblocks = m_.nslices = m_.nparts = m_.ndiag = m_.dia_sym = m_.spmm_kernel = m_.vi_nvals = 0;
        m_.vi_base = nullptr;
        m_.vi_table = nullptr;
        m_.tr = nullptr;
        m_.dia_nan = 0;
        m_.dia_s1 = m_.dia_s2 = 0;
        m_.vi_nsig = 0;
        m_.dia_w27 = 0;
        m_.dia_delta = nullptr;
        m_.format = UHIP_FMT_ROWBLOCK;
        plan(block_nnz);
    }

    void plan(int32_t block_nnz) { hip_check(uhip_csr_plan(c(), &m_, block_nnz), "uhip_csr_plan"); }

    int64_t rows() const { return m_.nrows; }
    int64_t cols() const { return m_.ncols; }
    int64_t nnz() const { return m_.nnz; }
    // sizes of PetscMatrix (utopia_petsc_Matrix.hpp): one GPU owns every row
    int64_t local_rows() const { return m_.nrows; }
    int64_t local_cols() const { return m_.ncols; }
    HipRange row_range() const { return HipRange{0, m_.nrows}; }
    const HipCommunicator &comm() const { return HipCommunicator::get(); }

    // PetscMatrix::read(f(i, j, a_ij)) (utopia_petsc_Matrix.hpp:482) over a host copy of the CSR
    template <class F>
    void read(F f) const {
        std::vector<int32_t> rp, col;
        std::vector<T> val;
        to_host(rp, col, val);
        for (int32_t i = 0; i < m_.nrows; ++i)
            for (int32_t e = rp[(size_t)i]; e < rp[(size_t)i + 1]; ++e) f((int64_t)i, (int64_t)col[(size_t)e], val[(size_t)e]);
    }
    // PetscMatrix::transform_values(f) (utopia_petsc_Matrix.hpp:476): a_ij = f(a_ij) on a host copy,
    // then the plan is rebuilt
    template <class F>
    void transform_values(F f) {
        std::vector<int32_t> rp, col;
        std::vector<T> val;
        to_host(rp, col, val);
        for (auto &v : val) v = f(v);
        if (!val.empty()) hip_check(uhip_memcpy_htod(c(), m_.val, val.data(), sizeof(T) * val.size()), "transform_values");
        replan();
    }
    // ||A||_inf = max_i sum_j |a_ij| (MatNorm NORM_INFINITY)
    double norm_infty() const {
        double best = 0;
        std::vector<int32_t> rp, col;
        std::vector<T> val;
        to_host(rp, col, val);
        for (int32_t i = 0; i < m_.nrows; ++i) {
            double s = 0;
            for (int32_t e = rp[(size_t)i]; e < rp[(size_t)i + 1]; ++e) s += std::fabs((double)val[(size_t)e]);
            best = s > best ? s : best;
        }
        return best;
    }
    void to_host(std::vector<int32_t> &rp, std::vector<int32_t> &col, std::vector<T> &val) const {
        rp.assign((size_t)m_.nrows + 1, 0);
        col.assign((size_t)m_.nnz, 0);
        val.assign((size_t)m_.nnz, T(0));
        hip_check(uhip_memcpy_dtoh(c(), rp.data(), m_.rowptr, sizeof(int32_t) * rp.size()), "d2h rowptr");
        if (m_.nnz) {
            hip_check(uhip_memcpy_dtoh(c(), col.data(), m_.col, sizeof(int32_t) * col.size()), "d2h col");
            hip_check(uhip_memcpy_dtoh(c(), val.data(), m_.val, sizeof(T) * val.size()), "d2h val");
        }
    }
    bool empty() const { return m_.rowptr == nullptr; }
    const uhip_csr &raw() const { return m_; }
    uhip_csr &raw() { return m_; }
    const HipContextPtr &context() const { return ctx_; }
    uhip_ctx c() const { return ctx_->get(); }

    // y = A x  (PetscMatrix::multiply, utopia_petsc_Matrix.hpp:570)
    void multiply(const Vector &x, Vector &y) const {
        if (y.size() != rows() || y.empty()) y = Vector(ctx_, rows());
        if (x.is_alias(y)) {  // PetscMatrix::multiply copies aliased input (utopia_petsc_Matrix.cpp:1328-1333)
            Vector tmp(x);
            hip_check(uhip_spmv(c(), &m_, tmp.data(), y.data()), "spmv");
            return;
        }
        hip_check(uhip_spmv(c(), &m_, x.data(), y.data()), "spmv");
    }
    void apply(const Vector &x, Vector &y) const { multiply(x, y); }

    // y = A^T x  (PetscMatrix::transpose_multiply -> MatMultTranspose, utopia_petsc_Matrix.cpp:1389);
    // the transpose and its plan are built on the first call and kept with A's plan
    void transpose_multiply(const Vector &x, Vector &y) {
        if (y.size() != cols() || y.empty()) y = Vector(ctx_, cols());
        hip_check(uhip_spmv_transpose(c(), &m_, x.data(), y.data()), "transpose_multiply");
    }

    // r = b - A x  (Eval<ResidualExpr>)
    void residual(const Vector &x, const Vector &b, Vector &r) const {
        hip_check(uhip_spmv_residual(c(), &m_, x.data(), b.data(), dtype, r.data()), "residual");
    }

    // d = diag(A)  (PetscMatrix::build_diag, utopia_petsc_Matrix.hpp:814)
    void build_diag(Vector &d) const {
        if (d.size() != rows() || d.empty()) d = Vector(ctx_, rows());
        hip_check(uhip_csr_diag(c(), &m_, d.data()), "diag");
    }

    // A_ii += alpha d_i (PetscMatrix::shift_diag -> MatDiagonalSet ADD_VALUES,
    // backend/petsc/utopia_petsc_Matrix.cpp:1783-1788); the plan is rebuilt
    void shift_diag(const Vector &d, T alpha = T(1)) {
        hip_check(uhip_csr_shift_diag(c(), &m_, d.data(), (double)alpha), "shift_diag");
        replan();
    }

    // B = A^T (PetscMatrix::transpose, utopia_petsc_Matrix.hpp:601)
    HipCrsMatrixT transpose() const {
        static_assert(std::is_same<T, double>::value, "transpose: FP64 setup only");
        uhip_csr B;
        std::memset(&B, 0, sizeof(B));
        hip_check(uhip_csr_transpose(c(), &m_, &B), "transpose");
        HipCrsMatrixT out;
        out.adopt(ctx_, B);
        return out;
    }

    // FP32 copy of the values, same pattern.
    HipCrsMatrixT<float> cast_f32() const;

    // Deep copy (same dtype and pattern), automatic plan.
    HipCrsMatrixT copy() const {
        uhip_csr m = m_;
        void *p = nullptr;
        const size_t vb = sizeof(T);
        hip_check(uhip_malloc(c(), &p, sizeof(int32_t) * ((size_t)m_.nrows + 1)), "alloc rowptr");
        m.rowptr = (int32_t *)p;
        hip_check(uhip_malloc(c(), &p, sizeof(int32_t) * (size_t)(m_.nnz > 0 ? m_.nnz : 1)), "alloc col");
        m.col = (int32_t *)p;
        hip_check(uhip_malloc(c(), &p, vb * (size_t)(m_.nnz > 0 ? m_.nnz : 1)), "alloc val");
        m.val = p;
        hip_check(uhip_memcpy_dtod(c(), m.rowptr, m_.rowptr, sizeof(int32_t) * ((size_t)m_.nrows + 1)), "copy rowptr");
        hip_check(uhip_memcpy_dtod(c(), m.col, m_.col, sizeof(int32_t) * (size_t)m_.nnz), "copy col");
        hip_check(uhip_memcpy_dtod(c(), m.val, m_.val, vb * (size_t)m_.nnz), "copy val");
        HipCrsMatrixT o;
        o.adopt(ctx_, m, 0);
        return o;
    }
    // Re-plan after the values changed in place (DIA / SELL-64 plans copy them).
    void replan() { plan(m_.format == UHIP_FMT_ROWBLOCK ? (m_.block_capk == 5 ? 1024 : (m_.block_capk == 20 ? 4096 : 2048)) : 0); }

private:
    void release() {
        if (!ctx_) return;
        if (own_) {
            uhip_csr_free(c(), &m_);
        } else {
            uhip_csr_plan_destroy(c(), &m_);
        }
        std::memset(&m_, 0, sizeof(m_));
    }

    HipContextPtr ctx_;
    uhip_csr m_;
    bool own_ = true;

    template <typename U>
    friend class HipCrsMatrixT;
};

using HipCrsMatrix = HipCrsMatrixT<double>;
using HipCrsMatrix32 = HipCrsMatrixT<float>;

template <>
inline HipCrsMatrixT<float> HipCrsMatrixT<double>::cast_f32() const {
    HipCrsMatrixT<float> o;
    uhip_csr m = m_;
    void *p = nullptr;
    hip_check(uhip_malloc(c(), &p, sizeof(int32_t) * ((size_t)m_.nrows + 1)), "alloc rowptr");
    m.rowptr = (int32_t *)p;
    hip_check(uhip_malloc(c(), &p, sizeof(int32_t) * (size_t)(m_.nnz > 0 ? m_.nnz : 1)), "alloc col");
    m.col = (int32_t *)p;
    hip_check(uhip_malloc(c(), &p, sizeof(float) * (size_t)(m_.nnz > 0 ? m_.nnz : 1)), "alloc val");
    m.val = p;
    m.dtype = UHIP_F32;
    hip_check(uhip_memcpy_dtod(c(), m.rowptr, m_.rowptr, sizeof(int32_t) * ((size_t)m_.nrows + 1)), "copy rowptr");
    hip_check(uhip_memcpy_dtod(c(), m.col, m_.col, sizeof(int32_t) * (size_t)m_.nnz), "copy col");
    hip_check(uhip_csr_cast_values_f64_f32(c(), m_.nnz, (const double *)m_.val, (float *)m.val), "cast vals");
    o.adopt(ctx_, m, m_.format != UHIP_FMT_ROWBLOCK ? 0 : (m_.block_capk == 5 ? 1024 : (m_.block_capk == 20 ? 4096 : 2048)));
    return o;
}

// Ac = P^T A P  (utopia::ptap -> MatPtAP, backend/petsc/eval/utopia_petsc_Eval_Multiply.cpp:12-61)
inline HipCrsMatrix ptap(const HipCrsMatrix &A, const HipCrsMatrix &P) {
    uhip_csr Ac;
    std::memset(&Ac, 0, sizeof(Ac));
    hip_check(uhip_csr_ptap(A.c(), &A.raw(), &P.raw(), &Ac), "ptap");
    HipCrsMatrix out;
    out.adopt(A.context(), Ac);
    return out;
}

}  // namespace utopia
