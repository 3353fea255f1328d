// Batched MG-preconditioned CG over many independent SPD systems with K
// right-hand sides each.
//
// The systems are one block-diagonal CSR matrix (segments of rows); every
// kernel covers all systems at once and every CG scalar (alpha, beta, norms)
// exists per (system, right-hand side). Per system and right-hand side the
// arithmetic is that of ConjugateGradient<..., HOMEMADE>
// (solvers/smoothers/utopia_ConjugateGradient_impl.hpp:195-294) preconditioned by one
// Multigrid V-cycle (solvers/multilevel/utopia_Multigrid.hpp:252-367) with
// Chebyshev3level smoothing (solvers/smoothers/utopia_Chebyshev3level.hpp:123-178)
// and a dense coarse solve; converged pairs take zero steps. One iteration of
// the whole batch is captured once as a hipGraph and replayed.
#pragma once

#include <chrono>
#include <cstdlib>
#include <map>
#include <memory>
#include <string>

#include "utopia_hip_solvers.hpp"

namespace utopia {
// the solvers of the HIP backend (the tensor types, HipVector and HipCrsMatrix, are in utopia)
namespace hip {

class DeviceInts {
public:
    DeviceInts() = default;
    DeviceInts(const HipContextPtr &ctx, const std::vector<int32_t> &h) { assign(ctx, h); }
    ~DeviceInts() { release(); }
    DeviceInts(const DeviceInts &) = delete;
    DeviceInts &operator=(const DeviceInts &) = delete;
    DeviceInts(DeviceInts &&o) noexcept : ctx_(std::move(o.ctx_)), d_(o.d_), n_(o.n_) { o.d_ = nullptr; }
    DeviceInts &operator=(DeviceInts &&o) noexcept {
        release();
        ctx_ = std::move(o.ctx_);
        d_ = o.d_;
        n_ = o.n_;
        o.d_ = nullptr;
        return *this;
    }
    // same size: overwrite in place (pointers captured in a graph stay valid)
    void fill(const std::vector<int32_t> &h) {
        if ((int64_t)h.size() != n_) throw HipError(UHIP_ERR_INVALID, "DeviceInts::fill size");
        if (n_) hip_check(uhip_memcpy_htod(ctx_->get(), d_, h.data(), sizeof(int32_t) * (size_t)n_), "h2d ints");
    }
    void assign(const HipContextPtr &ctx, const std::vector<int32_t> &h) {
        release();
        ctx_ = ctx;
        n_ = (int64_t)h.size();
        void *p = nullptr;
        hip_check(uhip_malloc(ctx->get(), &p, sizeof(int32_t) * (size_t)std::max<int64_t>(n_, 1)), "alloc ints");
        d_ = (int32_t *)p;
        if (n_) hip_check(uhip_memcpy_htod(ctx->get(), d_, h.data(), sizeof(int32_t) * (size_t)n_), "h2d ints");
    }
    int32_t *data() const { return d_; }
    int64_t size() const { return n_; }

private:
    void release() {
        if (d_ && ctx_) uhip_free(ctx_->get(), d_);
        d_ = nullptr;
    }
    HipContextPtr ctx_;
    int32_t *d_ = nullptr;
    int64_t n_ = 0;
};

// Segment layout of one level: row -> segment map and fixed reduction chunks.
struct SegLayout {
    std::vector<int32_t> seg;  // nseg + 1 row offsets
    DeviceInts rowseg, c_row0, c_row1, seg_c0;
    int32_t nchunks = 0;
    void build(const HipContextPtr &ctx, const std::vector<int32_t> &offsets, int32_t chunk_rows = 1024) {
        seg = offsets;
        const int32_t ns = (int32_t)seg.size() - 1;
        std::vector<int32_t> rs((size_t)seg.back()), r0, r1, c0((size_t)ns + 1, 0);
        for (int32_t s = 0; s < ns; ++s) {
            for (int32_t i = seg[(size_t)s]; i < seg[(size_t)s + 1]; ++i) rs[(size_t)i] = s;
            c0[(size_t)s] = (int32_t)r0.size();
            for (int32_t a = seg[(size_t)s]; a < seg[(size_t)s + 1]; a += chunk_rows) {
                r0.push_back(a);
                r1.push_back(std::min(a + chunk_rows, seg[(size_t)s + 1]));
            }
        }
        c0[(size_t)ns] = (int32_t)r0.size();
        nchunks = (int32_t)r0.size();
        rowseg.assign(ctx, rs);
        c_row0.assign(ctx, r0);
        c_row1.assign(ctx, r1);
        seg_c0.assign(ctx, c0);
    }
    int32_t nseg() const { return (int32_t)seg.size() - 1; }
    int32_t rows() const { return seg.back(); }
    // out[s*K + j] = (sqrt of) sum over segment s of x[., j] * y[., j]
    void dot(uhip_ctx c, int K, const void *x, const void *y, int dtype, double *out, bool take_sqrt) const {
        hip_check(uhip_segdot(c, K, nchunks, c_row0.data(), c_row1.data(), nseg(), seg_c0.data(), x, y, dtype, out,
                              take_sqrt ? 1 : 0),
                  "segdot");
    }
};

template <typename T>
struct BatchLevel {
    std::shared_ptr<const HipCrsMatrixT<T>> A, R, P;
    HipVectorT<T> c, r, rR, rk, pa, pb;  // N_l * K
    std::vector<HipVector> coef_pre, coef_post;  // per Chebyshev step: 3 * nseg (a_prev, beta, a_cur)
};

class BatchedMultigridCG {
public:
    BatchedMultigridCG(HipContextPtr ctx, int K, bool mixed) : ctx_(std::move(ctx)), K_(K), mixed_(mixed) {}
    ~BatchedMultigridCG() { drop_graph(); }

    void set_param(const std::string &k, double v) {
        params_[k] = v;
        // pre-size the reduction workspace (partial sums per dot slot) of this solver's context
        if (k == "workspace_reserve") hip_check(uhip_ctx_reserve(ctx_->get(), (int64_t)v), "workspace_reserve");
    }
    double param(const std::string &k, double def) const {
        auto it = params_.find(k);
        return it == params_.end() ? def : it->second;
    }

    // Block-diagonal operator with segment offsets; interpolations coarse->fine,
    // each with the segment offsets of its coarse side.
    void setup(const std::vector<int32_t> &seg, std::shared_ptr<HipCrsMatrix> A,
               const std::vector<std::shared_ptr<HipCrsMatrix>> &interps,
               const std::vector<std::vector<int32_t>> &coarse_segs) {
        const auto t0 = std::chrono::steady_clock::now();
        drop_graph();
        const int L = (int)interps.size() + 1;
        if ((int)coarse_segs.size() != L - 1) throw HipError(UHIP_ERR_INVALID, "batched: one segment list per level");
        // MultiLevelBase defaults (utopia_MultiLevelBase.hpp:36-37): 3 pre-, 3 post-smoothing steps
        pre_ = (int)param("pre_smoothing_steps", 3);
        post_ = (int)param("post_smoothing_steps", 3);
        if (pre_ < 1 || post_ < 1) throw HipError(UHIP_ERR_INVALID, "batched: smoothing steps >= 1");
        A64_.assign((size_t)L, nullptr);
        R64_.assign((size_t)L, nullptr);
        P64_.assign((size_t)L, nullptr);
        layout_.clear();
        layout_.resize((size_t)L);
        A64_[(size_t)L - 1] = A;
        layout_[(size_t)L - 1].build(ctx_, seg);
        for (int l = L - 1; l > 0; --l) {
            P64_[(size_t)l] = interps[(size_t)l - 1];
            R64_[(size_t)l] = std::make_shared<HipCrsMatrix>(interps[(size_t)l - 1]->transpose());
            A64_[(size_t)l - 1] = std::make_shared<HipCrsMatrix>(ptap(*A64_[(size_t)l], *interps[(size_t)l - 1]));
            layout_[(size_t)l - 1].build(ctx_, coarse_segs[(size_t)l - 1]);
        }
        // Chebyshev bounds per system (PowerMethod, utopia_PowerMethod.hpp:95-146), levels 1..L-1
        eig_max_.assign((size_t)L, {});
        eig_min_.assign((size_t)L, {});
        for (int l = 1; l < L; ++l) power_method(l);
        coarse_setup();
        build_levels();
        setup_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // X (N*K, row-major, initial guess and result), B (N*K), device pointers.
    int64_t solve(const double *B, double *X) {
        const auto t0 = std::chrono::steady_clock::now();
        const auto &top = layout_.back();
        const int32_t n = top.rows();
        const int32_t m = top.nseg() * K_;
        uhip_ctx c = ctx_->get();
        const int64_t nk = (int64_t)n * K_;
        if (x_.size() != nk) {
            drop_graph();
            for (HipVector *v : {&x_, &b_, &r_, &p_, &Ap_, &z_}) *v = HipVector(ctx_, nk, 0.0);
            for (HipVector *v : {&rz_, &pap_, &alpha_, &beta_, &rznew_, &rr_, &atol_}) *v = HipVector(ctx_, m, 0.0);
            flags_.assign(ctx_, std::vector<int32_t>((size_t)m, 0));
            conv_it_.assign(ctx_, std::vector<int32_t>((size_t)m, -1));
            counters_.assign(ctx_, std::vector<int32_t>(2, 0));
        }
        hip_check(uhip_memcpy_dtod(c, b_.data(), B, sizeof(double) * (size_t)nk), "copy B");
        hip_check(uhip_memcpy_dtod(c, x_.data(), X, sizeof(double) * (size_t)nk), "copy X");
        const double rtol = param("rtol", 1e-8);
        const int64_t max_it = (int64_t)param("max_it", 1000);
        // atol per (system, rhs) = rtol * |b|
        top.dot(c, K_, b_.data(), b_.data(), UHIP_F64, atol_.data(), true);
        atol_.scale(rtol);
        reset_state(m);
        // r = b - A x ; z = M r ; p = z ; rz = r.z
        hip_check(uhip_spmm(c, &A64_.back()->raw(), K_, x_.data(), UHIP_F64, r_.data(), UHIP_F64, 2, b_.data(),
                            UHIP_F64),
                  "r = b - Ax");
        precondition();
        p_.copy_from(z_);
        top.dot(c, K_, r_.data(), z_.data(), UHIP_F64, rz_.data(), false);
        const bool use_graph = param("use_graph", 1) != 0;
        // iterations per replay: the captured kernels are gated on the active-pair count
        // (uhip_ctx_set_gate), so the iterations of a replay past convergence do no work,
        // and converged pairs take zero steps in the ones before
        const int gi = std::max(1, (int)param("graph_iterations", 16));
        if (gi != graph_iters_) {
            drop_graph();
            graph_iters_ = gi;
        }
        int64_t it = 0;
        int32_t cnt[2] = {m, 0};  // active pairs, iterations done (device counters)
        graph_replays_ = 0;
        while (it < max_it && cnt[0] > 0) {
            if (use_graph && max_it - it >= graph_iters_) {
                // the graph holds the reduction workspace's address: capture again when an
                // operation on this context has reallocated it since
                if (!graph_exec_ || uhip_ctx_workspace_generation(c) != graph_gen_) capture();
                hip_check(uhip_graph_launch(c, graph_exec_), "graph launch");
                ++graph_replays_;
            } else {  // eager (or the last iterations before max_it)
                iteration();
            }
            hip_check(uhip_memcpy_dtoh(c, cnt, counters_.data(), sizeof(cnt)), "counters");
            it = cnt[1];
        }
        iterations_ = it;
        hip_check(uhip_memcpy_dtod(c, X, x_.data(), sizeof(double) * (size_t)nk), "copy X back");
        ctx_->sync();
        solve_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        return it;
    }

    void gradient(const double *lambda, const double *x, double *grad) const {
        hip_check(uhip_pattern_grad(ctx_->get(), &A64_.back()->raw(), K_, lambda, x, grad), "pattern grad");
    }

    std::vector<int32_t> conv_iterations() const { return download_ints(conv_it_); }
    std::vector<int32_t> flags() const { return download_ints(flags_); }
    const std::vector<double> &eig_max(int l) const { return eig_max_[(size_t)l]; }
    double setup_seconds() const { return setup_s_; }
    double solve_seconds() const { return solve_s_; }
    int64_t iterations() const { return iterations_; }
    int64_t graph_replays() const { return graph_replays_; }
    int64_t graph_captures() const { return graph_captures_; }
    int levels() const { return (int)A64_.size(); }
    const HipCrsMatrix &level_operator(int l) const { return *A64_[(size_t)l]; }
    void set_eig_max(int l, const std::vector<double> &v) {
        eig_max_[(size_t)l] = v;
        for (size_t s = 0; s < v.size(); ++s) eig_min_[(size_t)l][s] = param("scale_min_eig", 0.06) * v[s];
        drop_graph();
        build_levels();
    }

private:
    std::vector<int32_t> download_ints(const DeviceInts &d) const {
        std::vector<int32_t> h((size_t)d.size());
        if (d.size()) hip_check(uhip_memcpy_dtoh(ctx_->get(), h.data(), d.data(), sizeof(int32_t) * h.size()), "d2h");
        return h;
    }

    void reset_state(int32_t m) {
        flags_.fill(std::vector<int32_t>((size_t)m, 0));
        conv_it_.fill(std::vector<int32_t>((size_t)m, -1));
        counters_.fill(std::vector<int32_t>{m, 0});
    }

    // PowerMethod per system on level l (FP64 operator), rand() initial vector
    void power_method(int l) {
        const HipCrsMatrix &A = *A64_[(size_t)l];
        const SegLayout &lay = layout_[(size_t)l];
        const int32_t n = lay.rows(), ns = lay.nseg();
        const double tol = param("power_method_tol", 1e-2);
        const int64_t max_it = (int64_t)param("power_method_max_it", 30);
        const double smax = param("scale_max_eig", 1.2), smin = param("scale_min_eig", 0.06);
        uhip_ctx c = ctx_->get();
        std::vector<double> h((size_t)n);
        for (int32_t i = 0; i < n; ++i) h[(size_t)i] = ((double)std::rand() / (RAND_MAX)) + 1;
        HipVector v(ctx_, n), w(ctx_, n), nrm(ctx_, ns);
        v.upload(h.data());
        lay.dot(c, 1, v.data(), v.data(), UHIP_F64, nrm.data(), true);
        hip_check(uhip_seg_scale(c, n, lay.rowseg.data(), nrm.data(), nullptr, v.data(), v.data()), "normalise");
        std::vector<double> lam((size_t)ns, 0.0), old((size_t)ns, 0.0);
        std::vector<int32_t> done((size_t)ns, 0);
        DeviceInts done_d(ctx_, done);
        for (int64_t it = 1;; ++it) {
            hip_check(uhip_spmv(c, &A.raw(), v.data(), w.data()), "A v");
            lay.dot(c, 1, w.data(), w.data(), UHIP_F64, nrm.data(), true);
            hip_check(uhip_seg_scale(c, n, lay.rowseg.data(), nrm.data(), done_d.data(), w.data(), v.data()), "v");
            std::vector<double> nh((size_t)ns);
            nrm.download(nh.data());
            bool all = true;
            for (int32_t s = 0; s < ns; ++s) {
                if (done[(size_t)s]) continue;
                lam[(size_t)s] = nh[(size_t)s];
                if (std::abs(old[(size_t)s] - lam[(size_t)s]) < tol || it > max_it) done[(size_t)s] = 1;
                old[(size_t)s] = lam[(size_t)s];
                all = all && done[(size_t)s];
            }
            if (all) break;
            done_d.assign(ctx_, done);
        }
        eig_max_[(size_t)l].resize((size_t)ns);
        eig_min_[(size_t)l].resize((size_t)ns);
        for (int32_t s = 0; s < ns; ++s) {
            eig_max_[(size_t)l][(size_t)s] = smax * lam[(size_t)s];
            eig_min_[(size_t)l][(size_t)s] = smin * eig_max_[(size_t)l][(size_t)s];
        }
    }

    // Dense inverse of every diagonal block of the coarsest operator, kept as a
    // block-diagonal CSR (each block's rows are the dense inverse rows).
    void coarse_setup() {
        const HipCrsMatrix &A0 = *A64_[0];
        const SegLayout &lay = layout_[0];
        uhip_ctx c = ctx_->get();
        const int32_t ns = lay.nseg();
        std::vector<int32_t> rp(1, 0), col;
        int64_t tot = 0;
        for (int32_t s = 0; s < ns; ++s) {
            const int32_t a = lay.seg[(size_t)s], m = lay.seg[(size_t)s + 1] - a;
            for (int32_t i = 0; i < m; ++i) {
                for (int32_t j = 0; j < m; ++j) col.push_back(a + j);
                rp.push_back(rp.back() + m);
            }
            tot += (int64_t)m * m;
        }
        HipVector vals(ctx_, tot);
        // dense block of A0 per segment: copy rows of the CSR into a dense buffer on the host side of the ABI
        std::vector<int32_t> arp((size_t)A0.rows() + 1), acol((size_t)A0.nnz());
        std::vector<double> aval((size_t)A0.nnz());
        hip_check(uhip_memcpy_dtoh(c, arp.data(), A0.raw().rowptr, sizeof(int32_t) * arp.size()), "rowptr");
        hip_check(uhip_memcpy_dtoh(c, acol.data(), A0.raw().col, sizeof(int32_t) * acol.size()), "col");
        hip_check(uhip_memcpy_dtoh(c, aval.data(), A0.raw().val, sizeof(double) * aval.size()), "val");
        int64_t off = 0;
        for (int32_t s = 0; s < ns; ++s) {
            const int32_t a = lay.seg[(size_t)s], m = lay.seg[(size_t)s + 1] - a;
            std::vector<double> dense((size_t)m * m, 0.0);
            for (int32_t i = 0; i < m; ++i)
                for (int32_t e = arp[(size_t)(a + i)]; e < arp[(size_t)(a + i) + 1]; ++e)
                    dense[(size_t)i * m + (acol[(size_t)e] - a)] += aval[(size_t)e];
            HipVector Ld(ctx_, (int64_t)m * m);
            Ld.upload(dense.data());
            int info = 0;
            hip_check(uhip_dense_cholesky(c, m, Ld.data(), m, &info), "block cholesky");
            if (info) throw HipError(UHIP_ERR_NUMERIC, "batched: coarse block " + std::to_string(s) + " not SPD");
            hip_check(uhip_dense_cholesky_inverse(c, m, Ld.data(), m, vals.data() + off), "block inverse");
            off += (int64_t)m * m;
        }
        coarse_inv_ = std::make_shared<HipCrsMatrix>();
        std::vector<double> hv((size_t)tot);
        vals.download(hv.data());
        coarse_inv_->from_host(ctx_, A0.rows(), A0.rows(), rp.data(), col.data(), hv.data(), 2048);
    }

    template <typename T>
    void build_level_set(std::vector<BatchLevel<T>> &lv) {
        const int L = (int)A64_.size();
        lv.clear();
        lv.resize((size_t)L);
        // reps > 0: the multi-RHS kernel of every level operator and transfer is chosen by timing
        const int reps = (int)param("autotune", 0);
        auto tune = [&](const uhip_csr &m) {
            if (reps > 0)
                hip_check(uhip_csr_autotune_spmm(ctx_->get(), const_cast<uhip_csr *>(&m), K_, reps, nullptr, nullptr),
                          "autotune spmm");
        };
        if (reps > 0) {
            tune(A64_.back()->raw());  // the FP64 operator of the outer CG
            if (coarse_inv_) tune(coarse_inv_->raw());
        }
        for (int l = 0; l < L; ++l) {
            auto &d = lv[(size_t)l];
            d.A = cast<T>(A64_[(size_t)l]);
            tune(d.A->raw());
            if (l > 0) {
                d.R = cast<T>(R64_[(size_t)l]);
                d.P = cast<T>(P64_[(size_t)l]);
                tune(d.R->raw());
                tune(d.P->raw());
            }
            const int64_t nk = (int64_t)layout_[(size_t)l].rows() * K_;
            for (HipVectorT<T> *v : {&d.c, &d.r, &d.rR, &d.rk, &d.pa, &d.pb}) *v = HipVectorT<T>(ctx_, nk, T(0));
            if (l > 0) {
                d.coef_pre = coefficients(l, pre_);
                d.coef_post = coefficients(l, post_);
            }
        }
    }

    template <typename T>
    std::shared_ptr<const HipCrsMatrixT<T>> cast(const std::shared_ptr<HipCrsMatrix> &m);

    std::vector<HipVector> coefficients(int l, int nu) {
        const int32_t ns = layout_[(size_t)l].nseg();
        std::vector<std::vector<double>> a((size_t)ns), b((size_t)ns);
        for (int32_t s = 0; s < ns; ++s)
            chebyshev_coefficients(eig_max_[(size_t)l][(size_t)s], eig_min_[(size_t)l][(size_t)s], std::max(nu, 1),
                                   a[(size_t)s], b[(size_t)s]);
        std::vector<HipVector> out;
        const int steps = std::max(nu - 1, 1);
        for (int k = 1; k <= steps; ++k) {
            std::vector<double> h((size_t)3 * ns);
            for (int32_t s = 0; s < ns; ++s) {
                if (nu == 1) {
                    h[(size_t)3 * s] = a[(size_t)s][0];
                    h[(size_t)3 * s + 1] = h[(size_t)3 * s + 2] = 0.0;
                } else {
                    h[(size_t)3 * s] = a[(size_t)s][(size_t)k - 1];
                    h[(size_t)3 * s + 1] = b[(size_t)s][(size_t)k];
                    h[(size_t)3 * s + 2] = a[(size_t)s][(size_t)k];
                }
            }
            HipVector v(ctx_, 3 * ns);
            v.upload(h.data());
            out.push_back(std::move(v));
        }
        return out;
    }

    void build_levels() {
        if (mixed_) build_level_set(lv32_);
        else build_level_set(lv64_);
    }

    template <typename T>
    void smooth(std::vector<BatchLevel<T>> &lv, int l, const void *b, int b_dt, bool zero_guess, void *out,
                int out_dt, int nu, const std::vector<HipVector> &coef) {
        constexpr int dt = HipDtype<T>::value;
        auto &L = lv[(size_t)l];
        const auto &lay = layout_[(size_t)l];
        uhip_ctx c = ctx_->get();
        const void *pprev;
        int p_dt;
        if (zero_guess) {
            pprev = b;
            p_dt = b_dt;
        } else {
            hip_check(uhip_spmm(c, &L.A->raw(), K_, L.c.data(), dt, L.pa.data(), dt, 2, b, b_dt), "p0 = b - Ax");
            pprev = L.pa.data();
            p_dt = dt;
        }
        if (nu == 1) {
            hip_check(uhip_axpy_seg(c, lay.rows(), K_, lay.rowseg.data(), coef[0].data(), pprev, p_dt,
                                    zero_guess ? nullptr : L.c.data(), out, out_dt, dt),
                      "cheb nu=1");
            return;
        }
        T *pbuf[2] = {L.pb.data(), L.pa.data()};
        for (int k = 1; k < nu; ++k) {
            const bool last = (k == nu - 1);
            T *pnext = pbuf[(k - 1) & 1];
            if (pnext == pprev) pnext = pbuf[k & 1];
            hip_check(uhip_spmm_cheb_step(c, &L.A->raw(), K_, pprev, p_dt, k == 1 ? nullptr : L.rk.data(),
                                          (k == 1 && zero_guess) ? nullptr : L.c.data(),
                                          last ? out : (void *)L.c.data(), last ? out_dt : dt, L.rk.data(), pnext,
                                          lay.rowseg.data(), coef[(size_t)k - 1].data(), last ? 1 : 0),
                      "cheb step");
            pprev = pnext;
            p_dt = dt;
        }
    }

    template <typename T>
    void cycle(std::vector<BatchLevel<T>> &lv, int l, const void *r, int r_dt, void *out, int out_dt) {
        constexpr int dt = HipDtype<T>::value;
        uhip_ctx c = ctx_->get();
        if (l == 0) {
            hip_check(uhip_spmm(c, &coarse_inv_->raw(), K_, r, r_dt, out, out_dt, 0, nullptr, 0), "coarse solve");
            return;
        }
        auto &L = lv[(size_t)l];
        auto &Lc = lv[(size_t)l - 1];
        smooth(lv, l, r, r_dt, true, L.c.data(), dt, pre_, L.coef_pre);
        hip_check(uhip_spmm(c, &L.A->raw(), K_, L.c.data(), dt, L.rR.data(), dt, 2, r, r_dt), "residual");
        hip_check(uhip_spmm(c, &L.R->raw(), K_, L.rR.data(), dt, Lc.r.data(), dt, 0, nullptr, 0), "restrict");
        cycle(lv, l - 1, Lc.r.data(), dt, Lc.c.data(), dt);
        hip_check(uhip_spmm(c, &L.P->raw(), K_, Lc.c.data(), dt, L.c.data(), dt, 1, nullptr, 0), "interpolate");
        smooth(lv, l, r, r_dt, false, out, out_dt, post_, L.coef_post);
    }

    void precondition() {
        const int top = (int)A64_.size() - 1;
        if (mixed_) cycle(lv32_, top, r_.data(), UHIP_F64, z_.data(), UHIP_F64);
        else cycle(lv64_, top, r_.data(), UHIP_F64, z_.data(), UHIP_F64);
    }

    // one CG iteration of every (system, rhs) pair; no host synchronisation
    void iteration() {
        uhip_ctx c = ctx_->get();
        const auto &top = layout_.back();
        const int32_t n = top.rows(), m = top.nseg() * K_;
        double *state[7] = {rz_.data(), pap_.data(), alpha_.data(), beta_.data(), rznew_.data(), rr_.data(),
                            atol_.data()};
        hip_check(uhip_spmm(c, &A64_.back()->raw(), K_, p_.data(), UHIP_F64, Ap_.data(), UHIP_F64, 0, nullptr, 0), "Ap");
        top.dot(c, K_, p_.data(), Ap_.data(), UHIP_F64, pap_.data(), false);
        hip_check(uhip_bcg_step(c, 0, n, K_, m, top.rowseg.data(), x_.data(), p_.data(), r_.data(), Ap_.data(),
                                z_.data(), state, flags_.data(), conv_it_.data(), counters_.data() + 1,
                                counters_.data()),
                  "bcg update");
        top.dot(c, K_, r_.data(), r_.data(), UHIP_F64, rr_.data(), true);
        hip_check(uhip_bcg_step(c, 1, n, K_, m, top.rowseg.data(), x_.data(), p_.data(), r_.data(), Ap_.data(),
                                z_.data(), state, flags_.data(), conv_it_.data(), counters_.data() + 1,
                                counters_.data()),
                  "bcg check");
        precondition();
        top.dot(c, K_, z_.data(), r_.data(), UHIP_F64, rznew_.data(), false);
        hip_check(uhip_bcg_step(c, 2, n, K_, m, top.rowseg.data(), x_.data(), p_.data(), r_.data(), Ap_.data(),
                                z_.data(), state, flags_.data(), conv_it_.data(), counters_.data() + 1,
                                counters_.data()),
                  "bcg p update");
    }

    void capture() {
        drop_graph();
        // reserve the reduction workspace before capture (no allocation inside a graph)
        int64_t maxparts = 0;
        for (auto &lay : layout_) maxparts = std::max<int64_t>(maxparts, (int64_t)lay.nchunks * K_);
        hip_check(uhip_ctx_reserve(ctx_->get(), maxparts), "reserve");
        graph_gen_ = uhip_ctx_workspace_generation(ctx_->get());
        hip_check(uhip_graph_begin(ctx_->get()), "capture begin");
        hip_check(uhip_ctx_set_gate(ctx_->get(), counters_.data()), "gate");
        try {
            for (int i = 0; i < graph_iters_; ++i) iteration();
        } catch (...) {
            uhip_ctx_set_gate(ctx_->get(), nullptr);
            void *g = nullptr;
            uhip_graph_end(ctx_->get(), &g);
            uhip_graph_destroy(g);
            throw;
        }
        hip_check(uhip_ctx_set_gate(ctx_->get(), nullptr), "gate");
        hip_check(uhip_graph_end(ctx_->get(), &graph_exec_), "capture end");
        ++graph_captures_;
        if (uhip_ctx_workspace_generation(ctx_->get()) != graph_gen_) {
            drop_graph();
            throw HipError(UHIP_ERR_INVALID, "batched CG: reduction workspace reallocated during capture");
        }
    }
    void drop_graph() {
        if (graph_exec_) uhip_graph_destroy(graph_exec_);
        graph_exec_ = nullptr;
    }

    HipContextPtr ctx_;
    int K_;
    bool mixed_;
    int pre_ = 2, post_ = 2;
    std::map<std::string, double> params_;
    std::vector<std::shared_ptr<HipCrsMatrix>> A64_, R64_, P64_;
    std::vector<SegLayout> layout_;
    std::vector<std::vector<double>> eig_max_, eig_min_;
    std::shared_ptr<HipCrsMatrix> coarse_inv_;
    std::vector<BatchLevel<float>> lv32_;
    std::vector<BatchLevel<double>> lv64_;
    HipVector x_, b_, r_, p_, Ap_, z_;
    HipVector rz_, pap_, alpha_, beta_, rznew_, rr_, atol_;
    DeviceInts flags_, conv_it_, counters_;
    void *graph_exec_ = nullptr;
    int64_t graph_gen_ = -1;  // workspace generation the graph was captured at
    int graph_iters_ = 1;
    int64_t iterations_ = 0, graph_replays_ = 0, graph_captures_ = 0;
    double setup_s_ = 0, solve_s_ = 0;
};

template <>
inline std::shared_ptr<const HipCrsMatrixT<double>> BatchedMultigridCG::cast<double>(
    const std::shared_ptr<HipCrsMatrix> &m) {
    return m;
}
template <>
inline std::shared_ptr<const HipCrsMatrixT<float>> BatchedMultigridCG::cast<float>(
    const std::shared_ptr<HipCrsMatrix> &m) {
    return std::make_shared<HipCrsMatrix32>(m->cast_f32());
}

}  // namespace hip
}  // namespace utopia
