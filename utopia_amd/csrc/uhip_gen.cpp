// Seeded problem generators: CSR stencil operators, trilinear prolongations and
// random vectors. Host code; every path (GPU solve, CPU oracle, benchmark)
// consumes exactly these arrays.
#include <cmath>
#include <cstdint>
#include <new>
#include <vector>

#include "../../include/utopia_hip.h"

namespace {

struct SplitMix64 {
    uint64_t s;
    explicit SplitMix64(uint64_t seed) : s(seed) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ull);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * 0x1.0p-53; }
    // Box-Muller, cosine branch only: two uniforms per normal.
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
    }
};

inline double harmonic(double a, double b) { return 2.0 * a * b / (a + b); }

}  // namespace

extern "C" {

int64_t uhip_gen_nnz(int kind, int32_t nx, int32_t ny, int32_t nz) {
    const int64_t X = nx, Y = ny, Z = nz;
    switch (kind) {
        case 0:
        case 4:
            return 5 * X * Y - 2 * X - 2 * Y;
        case 1:
        case 2:
        case 5:
            return 7 * X * Y * Z - 2 * (X * Y + Y * Z + X * Z);
        case 3:
            return 9 * X * Y - 6 * X - 6 * Y + 4;
        default:
            return -1;
    }
}

static int gen_stencil(int kind, int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double shift, int32_t *rowptr,
                       int32_t *col, double *val);

// Kind 5: the kind-2 operator P A P^T, P the seeded permutation of its full 64-row blocks (a
// tail block of fewer than 64 rows keeps its place). Rows keep their sorted column order.
static int gen_block_permuted(int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double shift, int32_t *rowptr,
                              int32_t *col, double *val) {
    const int64_t N = (int64_t)nx * ny * nz, nnz = uhip_gen_nnz(2, nx, ny, nz);
    std::vector<int32_t> rp((size_t)N + 1), c((size_t)nnz);
    std::vector<double> v((size_t)nnz);
    if (const int st = gen_stencil(2, nx, ny, nz, seed, shift, rp.data(), c.data(), v.data())) return st;
    const int64_t nb = N / 64, full = nb * 64;
    std::vector<int64_t> to(nb), from(nb);  // to[old block] = new block; from = inverse
    for (int64_t i = 0; i < nb; ++i) to[i] = i;
    SplitMix64 rng(seed ^ 0x5851F42D4C957F2Dull);
    for (int64_t i = nb - 1; i > 0; --i) {  // Fisher-Yates
        const int64_t j = (int64_t)(rng.next() % (uint64_t)(i + 1));
        const int64_t t = to[i];
        to[i] = to[j];
        to[j] = t;
    }
    for (int64_t i = 0; i < nb; ++i) from[to[i]] = i;
    int64_t e = 0;
    rowptr[0] = 0;
    for (int64_t q = 0; q < N; ++q) {
        const int64_t r = q < full ? from[q >> 6] * 64 + (q & 63) : q;
        const int64_t s0 = e;
        for (int64_t k = rp[r]; k < rp[r + 1]; ++k) {
            const int64_t cc = c[k] < full ? to[c[k] >> 6] * 64 + (c[k] & 63) : c[k];
            int64_t t = e++;
            for (; t > s0 && col[t - 1] > cc; --t) col[t] = col[t - 1], val[t] = val[t - 1];
            col[t] = (int32_t)cc;
            val[t] = v[k];
        }
        rowptr[q + 1] = (int32_t)e;
    }
    return UHIP_OK;
}

int uhip_gen_stencil(int kind, int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double shift, int32_t *rowptr,
                     int32_t *col, double *val) {
    try {
        return gen_stencil(kind, nx, ny, nz, seed, shift, rowptr, col, val);
    } catch (const std::bad_alloc &) {  // edge-coefficient arrays of a huge grid
        return UHIP_ERR_ALLOC;
    }
}

static int gen_stencil(int kind, int32_t nx, int32_t ny, int32_t nz, uint64_t seed, double shift, int32_t *rowptr,
                       int32_t *col, double *val) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !rowptr || !col || !val) return UHIP_ERR_INVALID;
    int64_t e = 0;
    rowptr[0] = 0;
    if (kind == 0 || kind == 3 || kind == 4) {
        if (nz != 1) return UHIP_ERR_INVALID;
        std::vector<double> hx, vy;  // kind 4 edge coefficients
        if (kind == 4) {
            SplitMix64 rng(seed);
            hx.resize((size_t)(nx + 1) * ny);
            vy.resize((size_t)nx * (ny + 1));
            for (auto &c : hx) c = std::exp(0.5 * rng.normal());
            for (auto &c : vy) c = std::exp(0.5 * rng.normal());
        }
        // h-edge (x,y), x in [0,nx]: between (x-1,y) and (x,y); v-edge (x,y), y in [0,ny]: between (x,y-1) and (x,y)
        auto H = [&](int x, int y) { return hx[(size_t)y * (nx + 1) + x]; };
        auto V = [&](int x, int y) { return vy[(size_t)y * nx + x]; };
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int64_t i = x + (int64_t)nx * y;
                if (kind == 3) {
                    for (int dy = -1; dy <= 1; ++dy)
                        for (int dx = -1; dx <= 1; ++dx) {
                            const int xx = x + dx, yy = y + dy;
                            if (xx < 0 || xx >= nx || yy < 0 || yy >= ny) continue;
                            col[e] = (int32_t)(xx + (int64_t)nx * yy);
                            val[e] = (dx == 0 && dy == 0) ? 8.0 + shift : -1.0;
                            ++e;
                        }
                } else {
                    const double cs = kind == 4 ? V(x, y) : 1.0, cn = kind == 4 ? V(x, y + 1) : 1.0;
                    const double cw = kind == 4 ? H(x, y) : 1.0, ce = kind == 4 ? H(x + 1, y) : 1.0;
                    if (y > 0) col[e] = (int32_t)(i - nx), val[e++] = -cs;
                    if (x > 0) col[e] = (int32_t)(i - 1), val[e++] = -cw;
                    col[e] = (int32_t)i, val[e++] = (kind == 4 ? (cs + cw + ce + cn) : 4.0) + shift;
                    if (x < nx - 1) col[e] = (int32_t)(i + 1), val[e++] = -ce;
                    if (y < ny - 1) col[e] = (int32_t)(i + nx), val[e++] = -cn;
                }
                rowptr[i + 1] = (int32_t)e;
            }
        return UHIP_OK;
    }
    if (kind == 5) return gen_block_permuted(nx, ny, nz, seed, shift, rowptr, col, val);
    if (kind != 1 && kind != 2) return UHIP_ERR_INVALID;
    const int64_t sx = 1, sy = nx, sz = (int64_t)nx * ny, N = sz * nz;
    std::vector<double> kap;
    if (kind == 2) {
        SplitMix64 rng(seed);
        kap.resize((size_t)N);
        for (auto &k : kap) k = std::exp(rng.normal());
    }
    for (int z = 0; z < nz; ++z)
        for (int y = 0; y < ny; ++y)
            for (int x = 0; x < nx; ++x) {
                const int64_t i = x + sy * y + sz * z;
                const bool has[6] = {z > 0, y > 0, x > 0, x < nx - 1, y < ny - 1, z < nz - 1};
                const int64_t off[6] = {-sz, -sy, -sx, sx, sy, sz};
                double c[6];
                double diag = 0.0;
                for (int f = 0; f < 6; ++f) {
                    if (kind == 1) {
                        c[f] = 1.0;
                    } else {
                        c[f] = has[f] ? harmonic(kap[(size_t)i], kap[(size_t)(i + off[f])]) : kap[(size_t)i];
                    }
                    diag += c[f];
                }
                for (int f = 0; f < 3; ++f)
                    if (has[f]) col[e] = (int32_t)(i + off[f]), val[e++] = -c[f];
                col[e] = (int32_t)i, val[e++] = (kind == 1 ? 6.0 : diag) + shift;
                for (int f = 3; f < 6; ++f)
                    if (has[f]) col[e] = (int32_t)(i + off[f]), val[e++] = -c[f];
                rowptr[i + 1] = (int32_t)e;
            }
    return UHIP_OK;
}

}  // extern "C"

namespace {

// 1D linear interpolation weights of fine point f onto a coarse grid of nc points.
int weights1d(int f, int nf, int nc, int *cidx, double *w) {
    if (nf == 1) {
        cidx[0] = 0;
        w[0] = 1.0;
        return 1;
    }
    if (f & 1) {
        cidx[0] = (f - 1) / 2;
        w[0] = 1.0;
        return 1;
    }
    int m = 0;
    if (f / 2 - 1 >= 0) cidx[m] = f / 2 - 1, w[m++] = 0.5;
    if (f / 2 < nc) cidx[m] = f / 2, w[m++] = 0.5;
    return m;
}

}  // namespace

extern "C" {

int64_t uhip_gen_prolongation_nnz(int32_t nx, int32_t ny, int32_t nz) {
    auto count1d = [](int nf) -> int64_t {
        if (nf == 1) return 1;
        const int nc = nf / 2;
        int64_t s = 0;
        int ci[2];
        double w[2];
        for (int f = 0; f < nf; ++f) s += weights1d(f, nf, nc, ci, w);
        return s;
    };
    return count1d(nx) * count1d(ny) * count1d(nz);
}

int uhip_gen_prolongation(int32_t nx, int32_t ny, int32_t nz, int32_t *rowptr, int32_t *col, double *val) {
    if (nx <= 0 || ny <= 0 || nz <= 0 || !rowptr || !col || !val) return UHIP_ERR_INVALID;
    const int cx = nx > 1 ? nx / 2 : 1, cy = ny > 1 ? ny / 2 : 1, cz = nz > 1 ? nz / 2 : 1;
    int64_t e = 0;
    rowptr[0] = 0;
    int ix[2], iy[2], iz[2];
    double wx[2], wy[2], wz[2];
    for (int z = 0; z < nz; ++z) {
        const int mz = weights1d(z, nz, cz, iz, wz);
        for (int y = 0; y < ny; ++y) {
            const int my = weights1d(y, ny, cy, iy, wy);
            for (int x = 0; x < nx; ++x) {
                const int mx = weights1d(x, nx, cx, ix, wx);
                for (int a = 0; a < mz; ++a)
                    for (int b = 0; b < my; ++b)
                        for (int c = 0; c < mx; ++c) {
                            col[e] = (int32_t)(ix[c] + (int64_t)cx * (iy[b] + (int64_t)cy * iz[a]));
                            val[e] = wz[a] * wy[b] * wx[c];
                            ++e;
                        }
                rowptr[x + (int64_t)nx * (y + (int64_t)ny * z) + 1] = (int32_t)e;
            }
        }
    }
    return UHIP_OK;
}

int uhip_gen_uniform(uint64_t seed, int64_t n, double lo, double hi, double *x) {
    if (n < 0 || (n > 0 && !x)) return UHIP_ERR_INVALID;
    SplitMix64 rng(seed);
    for (int64_t i = 0; i < n; ++i) x[i] = lo + (hi - lo) * rng.uniform();
    return UHIP_OK;
}

int uhip_gen_normal(uint64_t seed, int64_t n, double mu, double sigma, double *x) {
    if (n < 0 || (n > 0 && !x)) return UHIP_ERR_INVALID;
    SplitMix64 rng(seed);
    for (int64_t i = 0; i < n; ++i) x[i] = mu + sigma * rng.normal();
    return UHIP_OK;
}

}  // extern "C"
