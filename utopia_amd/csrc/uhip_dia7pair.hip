// Temporal blocking of a Chebyshev smoothing pair on the plane-walking 7-point operator: the
// two products of "last smoothing step, then the residual of its result" (pre-smoothing from
// a zero guess) or "residual, then the last smoothing step on it" (post-smoothing) in one
// pass over the matrix. Per-row arithmetic is that of EpiCheb / EpiResidual on
// k_spmv_dia7r, so both outputs equal the two separate launches bit for bit; only the dot of
// the post-smoothing pair is summed in another order.
//
// Layout. A workgroup of NT threads owns M = CORE * NT consecutive rows of a plane (a chunk) in
// each of P consecutive planes [zs, ze) and walks the planes. The first product runs two planes
// ahead of the second: in one step it computes plane f for the chunk and for S1 rows on either
// side (the y halo the second product needs) into a 4-plane ring in LDS, while the second
// product of plane f - 2 reads planes f - 3 .. f - 1 from the ring (the +-1 and +-S1
// neighbours from its own plane's slot, +-S2 from the slots beside it). The loads of plane f
// are issued before the second product of plane f - 2 is computed, so they are in flight while
// it runs, and one barrier per step suffices (the slot written is never one read in the same
// step). The first product also runs on the planes zs - 1 and ze next to the workgroup's own
// planes. A chunk row's seven values, loaded for the first product, are held in registers for
// its second product two planes later; the -S2 value and x at the planes below and at the row
// roll in registers as in k_spmv_dia7r, and x at r -+ 1 comes from the neighbouring lane by DPP
// where a wave's rows are consecutive (FP64 MG-PCG at 256^3: -3 ms per solve). Entries a row does not hold are NaN in the value arrays
// (dia_nan), so ring slots that are never written (planes outside the grid, halo rows outside
// the plane) meet only dropped products.
#include "uhip_spmv.hpp"


namespace uhip {

constexpr int kPairMaxS1 = 256;  // y halo rows on either side (one halo row per thread: 2 S1 <= NT)
constexpr int kPairPlanes = 32;  // planes per workgroup (measured: 8 and 16 are slower at 256^3)
// A grid with fewer planes takes fewer per workgroup, down to kPairMinPlanes, so that the pair
// keeps kPairMinGroups workgroups (one per CU) where it can.
constexpr int kPairMinGroups = 256;
constexpr int kPairMinPlanes = 8;

// threads per workgroup and chunk rows per thread: FP32 1024 x 2, FP64 512 x 2 (three planes of a
// row's values live in registers)
template <typename T>
struct PairShape {
    static constexpr int NT = sizeof(T) == 4 ? 1024 : 512;
    static constexpr int CORE = 2;
    // FP64 issues the next plane's loads before the barrier ending a step (they fly through it
    // and the next second product: measured faster); FP32 issues them at the start of the step
    // (measured faster there)
    static constexpr bool EARLY = sizeof(T) == 8;
};

// MODE 0 (pre-smoothing pair): c = a_prev b + a_cur (-(-b + a_prev A b) + beta b), the last
// Chebyshev step from a zero guess on x1 = b, stored to out1; then r_R = b - A c to out2.
// MODE 1 (post-smoothing pair): pa = b - A x1 (not stored); then the last Chebyshev step with
// pprev = pa, xin = x1: out2 = x1 + a_prev pa + a_cur (-(-pa + a_prev A pa) + beta pa), and with
// DOT part += out2 . dw.
template <typename T, int NT, int CORE, int MODE, bool DOT>
__global__ __launch_bounds__(NT) void k_dia7_pair(int32_t S1, int32_t S2, int32_t chunks, int32_t nz, int32_t P,
                                                  const T *__restrict__ val, int64_t ld, const T *__restrict__ x1,
                                                  const T *__restrict__ b, T *__restrict__ out1,
                                                  T *__restrict__ out2, const double *__restrict__ dw, T a_prev,
                                                  T beta, T a_cur, double *__restrict__ partials) {
    constexpr int M = CORE * NT;
    constexpr int NS = CORE + 1;  // row slots per thread: CORE chunk rows and one halo row
    __shared__ T ring[4][M + 2 * kPairMaxS1];
    const int blk = xcd_remap(blockIdx.x, gridDim.x);
    const int zc = blk / chunks, cc = blk - zc * chunks;
    const int zs = zc * P, ze = min(zs + P, nz);
    const int32_t o0 = cc * M, base = o0 - S1;  // ring index 0 <-> plane offset base
    const int t = threadIdx.x, lane = t & 63;
    const T nan = (T)__builtin_nan("");
    const T *v0 = val, *v1 = val + ld, *vS1 = val + 2 * ld, *vS2 = val + 3 * ld;
    const int32_t xlast = nz * S2 - 1;
    auto cl = [xlast](int32_t v) { return min(max(v, 0), xlast); };

    int32_t off[NS];
    bool on[NS];
#pragma unroll
    for (int k = 0; k < CORE; ++k) {
        off[k] = o0 + t + k * NT;
        on[k] = true;
    }
    off[CORE] = t < S1 ? base + t : o0 + M + (t - S1);
    on[CORE] = t < 2 * S1 && off[CORE] >= 0 && off[CORE] < S2;

    const int q0 = max(zs - 1, 0), q1 = min(ze, nz - 1);
    T prev_a2[NS], xprev[NS], xcur[NS];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
        prev_a2[j] = nan;
        xprev[j] = T(0);
        xcur[j] = T(0);
        if (on[j]) {
            const int32_t r = q0 * S2 + off[j];
            if (q0 >= 1) prev_a2[j] = vS2[r - S2];
            xprev[j] = x1[cl(r - S2)];
            xcur[j] = x1[r];
        }
    }
    double part = 0.0;
    struct Ld {
        T a0, m1, a1, aS1, mS1, aS2, xe, xe2, xmS1, xpS1, xn, b;
    };
    // x at r -+ 1 is the neighbouring lane's x_r (a DPP move; lanes 0 and 63 load theirs) where a
    // wave's 64 rows are consecutive and all active: always in the chunk slots, and in the halo
    // slot when S1 is a multiple of 64 (the left and right halo rows then never share a wave,
    // and a wave's halo rows are all inside the plane or all outside). Otherwise the halo slot
    // loads both neighbours (xe, xe2).
    const bool hdpp = (S1 & 63) == 0;
    auto dpp = [hdpp](int j) { return j < CORE || hdpp; };
    // the first product's loads of plane f
    auto issue = [&](int f, Ld(&L)[NS]) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (!on[j]) continue;
            const int32_t r = f * S2 + off[j];
            L[j].a0 = ld_stream(v0 + r);
            const typename PairOf<T>::t w =
                __builtin_nontemporal_load((const typename PairOf<T>::t *)(v1 + max(r - 1, 0)));
            L[j].m1 = r >= 1 ? w.x : nan;
            L[j].a1 = r >= 1 ? w.y : w.x;
            L[j].aS1 = vS1[r];
            const T ms = vS1[max(r - S1, 0)];
            L[j].mS1 = r >= S1 ? ms : nan;
            L[j].aS2 = ld_stream(vS2 + r);
            L[j].xe = T(0);
            L[j].xe2 = T(0);
            if (!dpp(j) || lane == 0 || lane == 63) L[j].xe = x1[cl(dpp(j) && lane == 63 ? r + 1 : r - 1)];
            if (!dpp(j)) L[j].xe2 = x1[cl(r + 1)];
            L[j].xmS1 = x1[cl(r - S1)];
            L[j].xpS1 = x1[cl(r + S1)];
            L[j].xn = x1[cl(r + S2)];
            if constexpr (MODE == 1) L[j].b = ld_stream(b + r);
        }
    };
    // the first product of plane f into its ring slot; the chunk rows' values and x1 to fv, fx
    auto first = [&](int f, T *slot, const Ld(&L)[NS], T(&fv)[CORE][7], T(&fx)[CORE]) {
#pragma unroll
        for (int j = 0; j < NS; ++j) {
            if (!on[j]) continue;
            const T vals[7] = {prev_a2[j], L[j].mS1, L[j].m1, L[j].a0, L[j].a1, L[j].aS1, L[j].aS2};
            const T fb = lane_from_below(xcur[j]), fa = lane_from_above(xcur[j]);
            const T xm1 = dpp(j) && lane != 0 ? fb : L[j].xe;
            const T xp1 = dpp(j) ? (lane == 63 ? L[j].xe : fa) : L[j].xe2;
            const T xs[7] = {xprev[j], L[j].xmS1, xm1, xcur[j], xp1, L[j].xpS1, L[j].xn};
            T acc = T(0);
#pragma unroll
            for (int d = 0; d < 7; ++d) {
                const T pr = vals[d] * xs[d];
                acc = (vals[d] == vals[d]) ? acc + pr : acc;
            }
            T v;
            if constexpr (MODE == 0) {
                const T pp = xcur[j];
                const T rk = -pp + a_prev * acc;
                const T pk = -rk + beta * pp;
                T xv = T(0);
                xv = xv + a_prev * pp;
                xv = xv + a_cur * pk;
                v = xv;
            } else {
                v = L[j].b - acc;
            }
            slot[off[j] - base] = v;
            if (j < CORE) {
                if constexpr (MODE == 0)
                    if (f >= zs && f < ze) st_stream(out1 + f * S2 + off[j], v);
#pragma unroll
                for (int d = 0; d < 7; ++d) fv[j][d] = vals[d];
                fx[j] = xcur[j];
            }
            prev_a2[j] = L[j].aS2;
            xprev[j] = xcur[j];
            xcur[j] = L[j].xn;
        }
    };
    // the second product of plane p from the ring slots of planes p - 1, p, p + 1
    auto second = [&](int p, const T *cm, const T *c0, const T *cp, const T(&hv)[CORE][7], const T(&hx)[CORE],
                      const double(&hd)[CORE]) {
#pragma unroll
        for (int kk = 0; kk < CORE; ++kk) {
            const int32_t r = p * S2 + off[kk];
            const int i = off[kk] - base;
            const T xs[7] = {cm[i], c0[i - S1], c0[i - 1], c0[i], c0[i + 1], c0[i + S1], cp[i]};
            T acc = T(0);
#pragma unroll
            for (int d = 0; d < 7; ++d) {
                const T pr = hv[kk][d] * xs[d];
                acc = (hv[kk][d] == hv[kk][d]) ? acc + pr : acc;
            }
            if constexpr (MODE == 0) {
                // r_R is read next by the restriction: in FP32 (67 MB at 256^3) a cached store
                // leaves it in the MALL for that read (measured: mixed MG-PCG -0.7 ms)
                if constexpr (sizeof(T) == 4) out2[r] = (T)(hx[kk] - acc);
                else st_stream(out2 + r, (T)(hx[kk] - acc));
            } else {
                const T pp = c0[i];
                const T rk = -pp + a_prev * acc;
                const T pk = -rk + beta * pp;
                T xv = hx[kk];
                xv = xv + a_prev * pp;
                xv = xv + a_cur * pk;
                st_stream(out2 + r, xv);
                if constexpr (DOT) part += (double)xv * hd[kk];
            }
        }
    };
    auto dw_of = [&](int p, double(&d)[CORE]) {
#pragma unroll
        for (int kk = 0; kk < CORE; ++kk) {
            d[kk] = 0.0;
            if constexpr (DOT)
                if (p >= zs && p < ze) d[kk] = ld_stream(dw + p * S2 + off[kk]);
        }
    };
    // the first product runs two planes ahead of the second; 4 ring slots, 1 barrier per plane
    T hv[CORE][7], hx[CORE], nv[CORE][7], nx[CORE];
    double hd[CORE];
#pragma unroll
    for (int k = 0; k < CORE; ++k) {
        hx[k] = nx[k] = T(0);
        hd[k] = 0.0;
#pragma unroll
        for (int d = 0; d < 7; ++d) hv[k][d] = nv[k][d] = nan;
    }
    const int kend = ze + 1 - q0;  // the step whose second product is plane ze - 1
    constexpr bool EARLY = PairShape<T>::EARLY;
    Ld L[NS];
    if (EARLY) issue(q0, L);
    for (int k = 0; k <= kend; ++k) {
        const int f = q0 + k, p = f - 2;
        if (!EARLY && f <= q1) issue(f, L);
        double fd[CORE];
        dw_of(p + 1, fd);
        if (p >= zs && p < ze) second(p, ring[(p + 3) & 3], ring[p & 3], ring[(p + 1) & 3], hv, hx, hd);
        T fv[CORE][7], fx[CORE];
#pragma unroll
        for (int kk = 0; kk < CORE; ++kk) {
            fx[kk] = T(0);
#pragma unroll
            for (int d = 0; d < 7; ++d) fv[kk][d] = nan;
        }
        if (f <= q1) first(f, ring[f & 3], L, fv, fx);
        if (EARLY && f + 1 <= q1) issue(f + 1, L);
#pragma unroll
        for (int kk = 0; kk < CORE; ++kk) {
            hx[kk] = nx[kk];
            nx[kk] = fx[kk];
            hd[kk] = fd[kk];
#pragma unroll
            for (int d = 0; d < 7; ++d) {
                hv[kk][d] = nv[kk][d];
                nv[kk][d] = fv[kk][d];
            }
        }
        __syncthreads();
    }
    if constexpr (DOT) {
        const double s = block_sum<NT>(part);
        if (threadIdx.x == 0) partials[blk] = s;
    }
}

// The pair applies to a plane-walking operator whose planes split into whole chunks of
// PairShape rows (2048 FP32, 1024 FP64) and whose y halo fits one row per thread.
static int pair_planes(const uhip_csr *A) {
    if (A->format != UHIP_FMT_DIA || A->dia_sym != 2 || A->dia_nan != 1) return 0;
    const int32_t S1 = A->dia_s1, S2 = A->dia_s2;
    if (S1 < 1 || S1 > kPairMaxS1 || S2 <= 0 || A->nrows % S2 != 0) return 0;
    const int32_t m = A->dtype == UHIP_F64 ? PairShape<double>::NT * PairShape<double>::CORE
                                           : PairShape<float>::NT * PairShape<float>::CORE;
    if (S2 % m != 0) return 0;
    const int64_t planes = (int64_t)(A->nrows / S2) * (S2 / m);  // chunk-planes to share out
    return (int)std::max<int64_t>(kPairMinPlanes, std::min<int64_t>(kPairPlanes, planes / kPairMinGroups));
}
static bool pair_ok(const uhip_csr *A) { return pair_planes(A) > 0; }
static int64_t pair_groups(const uhip_csr *A) {
    const int P = pair_planes(A);
    if (P == 0) return 0;
    const int32_t m = A->dtype == UHIP_F64 ? PairShape<double>::NT * PairShape<double>::CORE
                                           : PairShape<float>::NT * PairShape<float>::CORE;
    const int64_t nz = A->nrows / A->dia_s2;
    return (int64_t)(A->dia_s2 / m) * ((nz + P - 1) / P);
}

template <typename T, int MODE, bool DOT>
static int pair_launch(uhip_ctx ctx, const uhip_csr *A, const T *x1, const T *b, T *out1, T *out2,
                       const double *dw, double a_prev, double beta, double a_cur, double *dot_d) {
    constexpr int NT = PairShape<T>::NT, CORE = PairShape<T>::CORE;
    const int32_t S1 = A->dia_s1, S2 = A->dia_s2, nz = A->nrows / S2;
    const int32_t chunks = S2 / (CORE * NT);
    const int32_t P = pair_planes(A);
    const int64_t nblk = (int64_t)chunks * ((nz + P - 1) / P);
    if (nblk <= 0) return UHIP_OK;
    if (nblk > INT32_MAX) return set_error(UHIP_ERR_INVALID, "dia7 pair: grid too large");
    if (DOT) UHIP_TRY(ensure_partials(ctx, nblk));
    hipLaunchKernelGGL((k_dia7_pair<T, NT, CORE, MODE, DOT>), dim3((unsigned)nblk), dim3(NT), 0, ctx->stream, S1, S2,
                       chunks, nz, P, (const T *)A->dia_val, (int64_t)dia_ld(A->nrows), x1, b, out1, out2, dw,
                       (T)a_prev, (T)beta, (T)a_cur, DOT ? ctx->partials : nullptr);
    UHIP_CHECK_LAUNCH();
    if (DOT) UHIP_TRY(finish_sums(ctx, 1, nblk, ctx->max_partials, dot_d, 0u));
    return UHIP_OK;
}

template <typename T, int MODE>
static int pair_typed(uhip_ctx ctx, const uhip_csr *A, const void *x1, const void *b, void *out1, void *out2,
                      const double *dw, double a_prev, double beta, double a_cur, double *dot_d) {
    const T *xt = (const T *)x1, *bt = (const T *)b;
    T *o1 = (T *)out1, *o2 = (T *)out2;
    if (dw) return pair_launch<T, MODE, true>(ctx, A, xt, bt, o1, o2, dw, a_prev, beta, a_cur, dot_d);
    return pair_launch<T, MODE, false>(ctx, A, xt, bt, o1, o2, dw, a_prev, beta, a_cur, dot_d);
}

}  // namespace uhip

using namespace uhip;

extern "C" {

int uhip_cheb_pair_supported(const uhip_csr *A) { return A && pair_ok(A) ? 1 : 0; }

int64_t uhip_cheb_pair_workgroups(const uhip_csr *A) { return A ? pair_groups(A) : 0; }

int uhip_cheb_pre_residual(uhip_ctx ctx, const uhip_csr *A, const void *b, double a_prev, double beta,
                           double a_cur, void *c, void *rR) {
    UHIP_REQUIRE(ctx && A && b && c && rR, "uhip_cheb_pre_residual: null argument");
    UHIP_REQUIRE(c != b && rR != b && rR != c, "uhip_cheb_pre_residual: b, c and rR must not overlap");
    if (!pair_ok(A))
        return set_error(UHIP_ERR_UNSUPPORTED, "uhip_cheb_pre_residual: not a plane-walking 7-point operator "
                                               "with whole 1024-row (FP64) or 2048-row (FP32) chunks per plane");
    if (A->dtype == UHIP_F64)
        return pair_typed<double, 0>(ctx, A, b, nullptr, c, rR, nullptr, a_prev, beta, a_cur, nullptr);
    return pair_typed<float, 0>(ctx, A, b, nullptr, c, rR, nullptr, a_prev, beta, a_cur, nullptr);
}

int uhip_cheb_residual_post(uhip_ctx ctx, const uhip_csr *A, const void *x, const void *b, double a_prev,
                            double beta, double a_cur, void *out, const double *dot_with, double *dot_d) {
    UHIP_REQUIRE(ctx && A && x && b && out, "uhip_cheb_residual_post: null argument");
    UHIP_REQUIRE(out != x && out != b, "uhip_cheb_residual_post: out must not overlap x or b");
    UHIP_REQUIRE(!dot_with || dot_d, "uhip_cheb_residual_post: dot_with without dot_d");
    if (!pair_ok(A))
        return set_error(UHIP_ERR_UNSUPPORTED, "uhip_cheb_residual_post: not a plane-walking 7-point operator "
                                               "with whole 1024-row (FP64) or 2048-row (FP32) chunks per plane");
    if (A->dtype == UHIP_F64)
        return pair_typed<double, 1>(ctx, A, x, b, nullptr, out, dot_with, a_prev, beta, a_cur, dot_d);
    return pair_typed<float, 1>(ctx, A, x, b, nullptr, out, dot_with, a_prev, beta, a_cur, dot_d);
}

}  // extern "C"
