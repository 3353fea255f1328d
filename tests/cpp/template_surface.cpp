// A solver template written against the tensor surface only (the way utopia's templated solvers
// reach a backend: multiply/apply, dot, norm2, axpy, scale, e_mul, reciprocal, layout, local views,
// reduce/transform with operation tags), instantiated on utopia::HipVector / utopia::HipCrsMatrix.
// It restates ConjugateGradient<M, V, HOMEMADE>::preconditioned_solve with a Jacobi preconditioner
// (solvers/smoothers/utopia_ConjugateGradient_impl.hpp:195-294) step by step, unfused, and checks the
// result against the backend's fused ConjugateGradient on the same system. Exit status 0 = pass.
#include <cmath>
#include <cstdio>
#include <vector>

#include "utopia_hip_solvers.hpp"

namespace utopia {
// the operation tags utopia's core defines (utopia_core/expressions/utopia_Operators.hpp)
class Plus {};
class Min {};
class Max {};
class Abs {};
}  // namespace utopia

template <class Matrix, class Vector>
struct TemplatedJacobiCG {
    double atol = 1e-10;
    long max_it = 5000, iterations = 0;

    bool solve(const Matrix &A, const Vector &b, Vector &x) {
        using Scalar = typename Vector::Scalar;
        Vector d, r(b), z, p, Ap, r_new(b), z_new;
        A.build_diag(d);
        d.reciprocal(Scalar(1));           // InvDiagPreconditioner, nonzero diagonal
        x.zeros(b.layout());
        A.multiply(x, Ap);
        r = b;
        r.axpy(Scalar(-1), Ap);            // r = b - A x
        z = r;
        z.e_mul(d);                        // z = D^-1 r
        p = z;
        double rz = dot(r, z);
        for (long it = 0; it < max_it; ++it) {
            A.apply(p, Ap);
            const double alpha = rz / dot(p, Ap);
            x.axpy(Scalar(alpha), p);
            r_new = r;
            r_new.axpy(Scalar(-alpha), Ap);
            if (norm2(r_new) <= atol) {
                iterations = it;
                return true;
            }
            z_new = r_new;
            z_new.e_mul(d);
            const double rz_new = dot(z_new, r_new);
            p.axpby(Scalar(1), z_new, Scalar(rz_new / rz));  // p = z_new + beta p
            rz = rz_new;
            r = r_new;
        }
        iterations = max_it;
        return false;
    }
};

static int fails = 0;
#define CHECK(c)                                                   \
    do {                                                           \
        if (!(c)) {                                                \
            std::printf("FAIL %s (line %d)\n", #c, __LINE__);      \
            ++fails;                                               \
        }                                                          \
    } while (0)

int main() {
    using namespace utopia;
    auto ctx = std::make_shared<HipContext>(0);
    const int nx = 48, ny = 40;
    const int64_t nnz = uhip_gen_nnz(0, nx, ny, 1);
    const int32_t n = nx * ny;
    std::vector<int32_t> rp(n + 1), col(nnz);
    std::vector<double> val(nnz), bh(n);
    uhip_gen_stencil(0, nx, ny, 1, 0, 0.0, rp.data(), col.data(), val.data());
    uhip_gen_uniform(1, n, -1, 1, bh.data());
    HipCrsMatrix A;
    A.from_host(ctx, n, n, rp.data(), col.data(), val.data());
    HipVector b(ctx, n);
    b.upload(bh.data());
    const double atol = 1e-10 * b.norm2();

    // the surface itself
    CHECK(b.layout().local_size() == n && b.layout().size() == n && b.comm().size() == 1);
    CHECK(b.range().begin() == 0 && b.range().end() == n && A.row_range().extent() == n);
    HipVector a(b);
    a.transform(Abs());
    CHECK(std::fabs(a.reduce(Min())) >= 0 && a.reduce(Max()) <= 1.0 && a.reduce(Plus()) > 0);
    {
        a.read_lock();
        double s = 0;
        for (int64_t i = a.range().begin(); i < a.range().end(); ++i) s += a.l_get(i);
        a.read_unlock();
        CHECK(std::fabs(s - a.reduce(Plus())) <= 1e-10 * s);
    }
    {
        HipVector w(ctx, n, 0.0);
        w.write_lock(0);
        for (int64_t i = 0; i < n; ++i) w.l_set(i, (double)i);
        w.write_unlock(0);
        CHECK(w.reduce(Max()) == n - 1);
        HipVector m(ctx, n, 100.0);
        m.e_min(w);
        CHECK(m.reduce(Max()) == 100.0);
        m.e_max(w);  // max(min(100, i), i) = i
        CHECK(m.reduce(Plus()) == (double)(n - 1) * n / 2);
    }
    double rowsum = 0;
    A.read([&](int64_t i, int64_t j, double v) { rowsum += (i == 0) ? std::fabs(v) : 0.0; });
    CHECK(rowsum == 6.0 && A.norm_infty() == 8.0);
    HipVector y1, y2;
    A.multiply(b, y1);
    A.transpose_multiply(b, y2);  // symmetric: A^T b == A b
    y1.axpy(-1.0, y2);
    CHECK(y1.norm_infty() == 0.0);

    // the templated solver on the HIP tensors against the backend's fused CG
    TemplatedJacobiCG<HipCrsMatrix, HipVector> tcg;
    tcg.atol = atol;
    HipVector xt(ctx, n, 0.0);
    CHECK(tcg.solve(A, b, xt));
    hip::ConjugateGradient<HipCrsMatrix, HipVector> cg;
    cg.set_preconditioner(std::make_shared<hip::InvDiagPreconditioner<HipCrsMatrix, HipVector>>());
    cg.atol(atol);
    cg.rtol(0);
    cg.stol(0);
    cg.max_it(5000);
    HipVector xf(ctx, n, 0.0);
    cg.solve(A, b, xf);
    const long itf = (long)cg.solution_status().iterates;
    HipVector res(b);
    A.multiply(xt, y1);
    res.axpy(-1.0, y1);
    HipVector diff(xt);
    diff.axpy(-1.0, xf);
    std::printf("templated CG: %ld iterations, fused CG: %ld, |b - A x| = %.3e (atol %.3e), |x_t - x_f|/|x_f| = %.3e\n",
                tcg.iterations, itf, res.norm2(), atol, diff.norm2() / xf.norm2());
    CHECK(std::labs(tcg.iterations - itf) <= 1);
    CHECK(res.norm2() <= 1.01 * atol);
    CHECK(diff.norm2() <= 1e-8 * xf.norm2());
    std::printf(fails ? "FAILED (%d)\n" : "OK\n", fails);
    return fails ? 1 : 0;
}
