// AlgebraicMultigrid over the HIP backend:
//   Agglomerate         solvers/multilevel/agglomerate/utopia_Agglomerate.hpp:12-187
//   BlockAgglomerate    solvers/multilevel/agglomerate/utopia_BlockAgglomerate.hpp:12-163
//   AlgebraicMultigrid  solvers/multilevel/utopia_AlgebraicMultigrid.hpp:14-173
// The aggregation is a greedy pass whose every decision depends on the rows
// before it, so it runs on the host over the downloaded CSR arrays (setup
// only, O(nnz)); the coarse operators P^T A P are the device SpGEMMs and the
// cycles are the device Multigrid.
#pragma once

#include <algorithm>
#include <cmath>
#include <utility>
#include <vector>

#include "utopia_hip_multigrid.hpp"

namespace utopia {
// the solvers of the HIP backend (the tensor types, HipVector and HipCrsMatrix, are in utopia)
namespace hip {

class Agglomerate {
public:
    void bmax(int64_t v) { bmax_ = v; }
    void weight(double v) { weight_ = v; }
    int64_t bmax() const { return bmax_; }
    double weight() const { return weight_; }
    // AlgebraicMultigrid::ensure_agglomerator (utopia_AlgebraicMultigrid.hpp:87-103):
    // block_size 2, 3 or 4 selects BlockAgglomerate, anything else the scalar pass.
    // component < 0 means BlockSize - 1 (BlockAgglomerate's default).
    void block_size(int v) { block_size_ = v; }
    int block_size() const { return (block_size_ >= 2 && block_size_ <= 4) ? block_size_ : 1; }
    void component(int v) { component_ = v; }
    int component() const { return component_ < 0 ? block_size() - 1 : component_; }

    // BlockAgglomerate::create_transfer (:37-141) on the B x B block graph of a matrix
    // with n = B * nb rows: block (I, J) exists where any entry of rows I*B.. and columns
    // J*B.. is stored, its weight is the entry (I*B + c, J*B + c) (0 if absent). a_max[I]
    // is the largest |weight| off the diagonal block; block row I not yet in an aggregate
    // opens one and takes up to bmax blocks J not yet in one with weight * a_max[I] <
    // |w(I, J)|, in column order; if it takes none it joins the aggregate of its largest
    // |w| block (possibly its own, leaving it open); open blocks become singletons in
    // order. P(I*B + s, parent[I]*B + s) = 1 for every sub-row s.
    static void aggregate_blocks(int32_t n, const int32_t *rp, const int32_t *col, const double *val, int B, int c,
                                 int64_t bmax, double weight, std::vector<int32_t> &parent, int32_t &n_coarse) {
        if (B < 2 || n % B != 0 || c < 0 || c >= B)
            throw HipError(UHIP_ERR_INVALID, "BlockAgglomerate: rows must be a multiple of the block size, 0 <= component < block size");
        const int32_t nb = n / B;
        std::vector<std::vector<std::pair<int32_t, double>>> brow((size_t)nb);  // (J, weight), J ascending
        std::vector<int32_t> js;
        for (int32_t I = 0; I < nb; ++I) {
            js.clear();
            for (int32_t r = I * B; r < (I + 1) * B; ++r)
                for (int32_t e = rp[r]; e < rp[r + 1]; ++e) js.push_back(col[e] / B);
            std::sort(js.begin(), js.end());
            js.erase(std::unique(js.begin(), js.end()), js.end());
            auto &row = brow[(size_t)I];
            row.reserve(js.size());
            for (int32_t J : js) row.emplace_back(J, 0.0);
            const int32_t r = I * B + c;
            for (int32_t e = rp[r]; e < rp[r + 1]; ++e) {
                if (col[e] % B != c) continue;
                const int32_t J = col[e] / B;
                auto it = std::lower_bound(row.begin(), row.end(), std::make_pair(J, -HUGE_VAL));
                it->second = val[e];
            }
        }
        std::vector<double> a_max((size_t)nb, 0.0);
        for (int32_t I = 0; I < nb; ++I)
            for (const auto &b : brow[(size_t)I])
                if (b.first != I) a_max[(size_t)I] = std::max(a_max[(size_t)I], std::fabs(b.second));
        parent.assign((size_t)nb, -1);
        int32_t nc = 0;
        for (int32_t I = 0; I < nb; ++I) {
            if (parent[(size_t)I] != -1) continue;
            parent[(size_t)I] = nc;
            const double strong = weight * a_max[(size_t)I];
            int64_t taken = 0;
            for (const auto &b : brow[(size_t)I]) {
                if (strong < std::fabs(b.second) && parent[(size_t)b.first] == -1) {
                    parent[(size_t)b.first] = nc;
                    ++taken;
                }
                if (taken >= bmax) break;
            }
            if (taken > 0) {
                ++nc;
                continue;
            }
            parent[(size_t)I] = -1;
            double best = 0.0;
            int32_t arg = -1;
            for (const auto &b : brow[(size_t)I]) {
                const double a = std::fabs(b.second);
                if (a > best) {
                    best = a;
                    arg = b.first;
                }
            }
            if (arg != -1) parent[(size_t)I] = parent[(size_t)arg];
        }
        for (int32_t I = 0; I < nb; ++I)
            if (parent[(size_t)I] == -1) parent[(size_t)I] = nc++;
        n_coarse = nc;
    }

    // Agglomerate::create_transfer_matrix (:50-178): parent[i] = aggregate of row i
    // (-2: a row with a single stored value joins no aggregate). Row i not yet in an
    // aggregate opens one and takes up to bmax strongly coupled (weight * max_j|a_ij| <
    // |a_ij|) rows not yet in one, in column order; if it takes none it joins the
    // aggregate of its largest-magnitude entry (which may be its own diagonal, leaving
    // it open). Rows still open become singletons in row order. P(i, parent[i]) is 1,
    // or 0 where row i has no nonzero off-diagonal.
    static void aggregate(int32_t n, const int32_t *rp, const int32_t *col, const double *val, int64_t bmax,
                          double weight, std::vector<int32_t> &parent, std::vector<double> &a_max, int32_t &n_coarse) {
        parent.assign((size_t)n, -1);
        a_max.assign((size_t)n, 0.0);
        for (int32_t i = 0; i < n; ++i)
            if (rp[i + 1] - rp[i] == 1) parent[(size_t)i] = -2;
        for (int32_t i = 0; i < n; ++i)
            for (int32_t e = rp[i]; e < rp[i + 1]; ++e)
                if (col[e] != i) a_max[(size_t)i] = std::max(a_max[(size_t)i], std::fabs(val[e]));
        int32_t nc = 0;
        for (int32_t i = 0; i < n; ++i) {
            if (parent[(size_t)i] != -1) continue;
            parent[(size_t)i] = nc;
            const double strong = weight * a_max[(size_t)i];
            int64_t taken = 0;
            for (int32_t e = rp[i]; e < rp[i + 1]; ++e) {
                const int32_t j = col[e];
                if (strong < std::fabs(val[e]) && parent[(size_t)j] == -1) {
                    parent[(size_t)j] = nc;
                    ++taken;
                }
                if (taken >= bmax) break;
            }
            if (taken > 0) {
                ++nc;
                continue;
            }
            parent[(size_t)i] = -1;
            double best = 0.0;
            int32_t arg = -1;
            for (int32_t e = rp[i]; e < rp[i + 1]; ++e) {
                const double a = std::fabs(val[e]);
                if (a > best) {
                    best = a;
                    arg = col[e];
                }
            }
            if (arg != -1) parent[(size_t)i] = parent[(size_t)arg];
        }
        for (int32_t i = 0; i < n; ++i)
            if (parent[(size_t)i] == -1) parent[(size_t)i] = nc++;
        n_coarse = nc;
    }

    std::shared_ptr<HipCrsMatrix> create_transfer_matrix(const HipCrsMatrix &A) const {
        const uhip_csr &m = A.raw();
        if (m.dtype != UHIP_F64) throw HipError(UHIP_ERR_INVALID, "Agglomerate: FP64 operator expected");
        const int32_t n = m.nrows;
        std::vector<int32_t> rp((size_t)n + 1), col((size_t)std::max<int64_t>(m.nnz, 1));
        std::vector<double> val(col.size());
        hip_check(uhip_memcpy_dtoh(A.c(), rp.data(), m.rowptr, sizeof(int32_t) * ((size_t)n + 1)), "rowptr");
        hip_check(uhip_memcpy_dtoh(A.c(), col.data(), m.col, sizeof(int32_t) * (size_t)m.nnz), "col");
        hip_check(uhip_memcpy_dtoh(A.c(), val.data(), m.val, sizeof(double) * (size_t)m.nnz), "val");
        std::vector<int32_t> parent;
        std::vector<double> a_max;
        int32_t nc = 0;
        const int B = block_size();
        if (B > 1) {
            aggregate_blocks(n, rp.data(), col.data(), val.data(), B, component(), bmax_, weight_, parent, nc);
            std::vector<int32_t> prp((size_t)n + 1), pcol((size_t)n);
            std::vector<double> pval((size_t)n, 1.0);
            for (int32_t i = 0; i < n; ++i) {
                prp[(size_t)i] = i;
                pcol[(size_t)i] = parent[(size_t)(i / B)] * B + i % B;
            }
            prp[(size_t)n] = n;
            auto P = std::make_shared<HipCrsMatrix>();
            P->from_host(A.context(), n, nc * B, prp.data(), pcol.data(), pval.data());
            return P;
        }
        aggregate(n, rp.data(), col.data(), val.data(), bmax_, weight_, parent, a_max, nc);
        std::vector<int32_t> prp((size_t)n + 1, 0), pcol;
        std::vector<double> pval;
        pcol.reserve((size_t)n);
        pval.reserve((size_t)n);
        for (int32_t i = 0; i < n; ++i) {
            if (parent[(size_t)i] >= 0) {
                pcol.push_back(parent[(size_t)i]);
                pval.push_back(a_max[(size_t)i] != 0.0 ? 1.0 : 0.0);
            }
            prp[(size_t)i + 1] = (int32_t)pcol.size();
        }
        auto P = std::make_shared<HipCrsMatrix>();
        P->from_host(A.context(), n, nc, prp.data(), pcol.data(), pval.data());
        return P;
    }

private:
    int64_t bmax_ = 3;
    double weight_ = 1.0 / 3;
    int block_size_ = 1;
    int component_ = -1;
};

class AlgebraicMultigrid final : public IterativeSolver<HipCrsMatrix, HipVector> {
public:
    AlgebraicMultigrid(const std::shared_ptr<Multigrid> &algorithm, const std::shared_ptr<Agglomerate> &agglomerator)
        : algo_(algorithm), agg_(agglomerator) {}

    void n_levels(int v) { n_levels_ = v; }
    int n_levels() const { return n_levels_; }
    const std::shared_ptr<Multigrid> &algorithm() const { return algo_; }
    const std::shared_ptr<Agglomerate> &agglomerator() const { return agg_; }
    double setup_seconds() const { return setup_s_; }

    // AlgebraicMultigrid::max_it forwards to the Multigrid (:150-153)
    void max_it(int64_t v) {
        IterativeSolver::max_it(v);
        algo_->max_it(v);
    }

    // AlgebraicMultigrid::update + AlgebraicMultigridBuilder::build (:23-52, :128-137):
    // from the finest level down, P_l = agglomerate(A_{l+1}), A_l = P_l^T A_{l+1} P_l.
    void update(const std::shared_ptr<const HipCrsMatrix> &op) override {
        const auto t0 = std::chrono::steady_clock::now();
        IterativeSolver::update(op);
        const int L = n_levels_;
        if (L < 2) throw HipError(UHIP_ERR_INVALID, "AlgebraicMultigrid: n_levels >= 2");
        std::vector<std::shared_ptr<HipCrsMatrix>> transfers((size_t)L - 1);
        std::vector<std::shared_ptr<const HipCrsMatrix>> matrices((size_t)L);
        matrices[(size_t)L - 1] = op;
        std::shared_ptr<const HipCrsMatrix> last = op;
        for (int l = L - 2; l >= 0; --l) {
            auto P = agg_->create_transfer_matrix(*last);
            auto Ac = std::make_shared<HipCrsMatrix>(ptap(*last, *P));
            transfers[(size_t)l] = P;
            matrices[(size_t)l] = Ac;
            last = Ac;
        }
        algo_->set_transfer_operators(transfers);
        algo_->set_linear_operators(matrices);
        algo_->set_perform_galerkin_assembly(false);
        algo_->update(op);
        setup_s_ = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // AlgebraicMultigrid::apply (:118-123)
    bool apply(const HipVector &rhs, HipVector &x) override {
        const bool ok = algo_->apply(rhs, x);
        this->status_ = algo_->solution_status();
        return ok;
    }
    bool apply_zero_initial_guess(const HipVector &rhs, HipVector &z) override {
        return algo_->apply_zero_initial_guess(rhs, z);
    }
    bool apply_zero_initial_guess_dot(const HipVector &rhs, HipVector &z, double *dot_d) override {
        return algo_->apply_zero_initial_guess_dot(rhs, z, dot_d);
    }

private:
    std::shared_ptr<Multigrid> algo_;
    std::shared_ptr<Agglomerate> agg_;
    int n_levels_ = 2;
    double setup_s_ = 0.0;
};

}  // namespace hip
}  // namespace utopia
