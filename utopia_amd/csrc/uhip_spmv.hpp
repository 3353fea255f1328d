// Row-block CSR SpMV skeleton shared by every sparse kernel of the library.
//
// A plan splits the rows into blocks whose nonzeros fit in LDS. One workgroup
// per block:
//   phase 1  the block's (val, col) stream is read fully coalesced (lane t
//            takes entries e0+t, e0+t+WG, ...), x[col] is gathered through the
//            L2/MALL, and the products val*x go to LDS;
//   phase 2  lane t owns row r0+t (+WG ...) and sums its products from LDS in
//            column order, then hands (row, sum) to an epilogue that fuses the
//            per-row vector work (residual, Chebyshev step, correction, dots).
// The sequential per-row sum makes each y_i bit-identical to a textbook CSR
// loop with the products rounded before they are added.
#pragma once

#include "uhip_common.hpp"

namespace uhip {

constexpr int kSpmvWG = 256;

template <typename T>
__device__ __forceinline__ T ld_stream(const T *p) {
    return __builtin_nontemporal_load(p);
}

// Streaming (non-temporal) store of a vector entry that this kernel writes once.
template <typename T>
__device__ __forceinline__ void st_stream(T *p, T v) {
    __builtin_nontemporal_store(v, p);
}

// Epilogues: operator()(row, sum, part) with part[kDots] per-thread partial sums
// that the kernel reduces per block into partials[d*stride + block].
template <typename TAcc, int CAPK, class Epi, typename TA, typename TX>
__global__ __launch_bounds__(kSpmvWG) void k_spmv_rowblock(const int32_t *__restrict__ blocks,
                                                           const int32_t *__restrict__ rowptr,
                                                           const int32_t *__restrict__ col,
                                                           const TA *__restrict__ val, const TX *__restrict__ x,
                                                           Epi epi, double *__restrict__ partials, int64_t stride) {
    constexpr int CAP = kSpmvWG * CAPK;
    constexpr int ND = Epi::kDots > 0 ? Epi::kDots : 1;
    __shared__ TAcc prod[CAP];
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int r0 = blocks[b], r1 = blocks[b + 1];
    const int e0 = rowptr[r0], e1 = rowptr[r1];
    const int tid = threadIdx.x;
    double part[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) part[d] = 0.0;

    if (e1 - e0 <= CAP) {
        int c[CAPK];
        TA v[CAPK];
#pragma unroll
        for (int k = 0; k < CAPK; ++k) {
            const int e = e0 + tid + k * kSpmvWG;
            if (e < e1) {
                c[k] = ld_stream(col + e);
                v[k] = ld_stream(val + e);
            }
        }
#pragma unroll
        for (int k = 0; k < CAPK; ++k) {
            const int e = e0 + tid + k * kSpmvWG;
            if (e < e1) prod[e - e0] = (TAcc)v[k] * (TAcc)x[c[k]];
        }
        __syncthreads();
        for (int r = r0 + tid; r < r1; r += kSpmvWG) {
            const int a = rowptr[r] - e0, z = rowptr[r + 1] - e0;
            TAcc s = TAcc(0);
            for (int j = a; j < z; ++j) s += prod[j];
            epi(r, s, part);
        }
    } else {
        // A single row longer than the LDS budget (the plan isolates such rows).
        TAcc s = TAcc(0);
        for (int e = e0 + tid; e < e1; e += kSpmvWG) s += (TAcc)val[e] * (TAcc)x[col[e]];
        s = block_sum<kSpmvWG>(s);
        if (tid == 0) epi(r0, s, part);
    }
    if constexpr (Epi::kDots > 0) {
#pragma unroll
        for (int d = 0; d < Epi::kDots; ++d) {
            const double t = block_sum<kSpmvWG>(part[d]);
            if (tid == 0) partials[d * stride + b] = t;
        }
    }
}

// SELL-64: one wavefront per slice of 64 consecutive rows, lane = row. Entry j
// of the slice's rows sits at slice_off[s] + 64*j + lane, so every val/col
// load of the wave is one contiguous 256/512-byte segment. A lane stops at its
// own row length (from rowptr), so padding is never read; lanes sum their row
// in column order.
constexpr int kSellWaves = 4;  // slices per workgroup

template <typename TAcc, class Epi, typename TA, typename TX>
__global__ __launch_bounds__(64 * kSellWaves) void k_spmv_sell(int32_t nrows, int32_t nslices,
                                                              const int32_t *__restrict__ rowptr,
                                                              const int32_t *__restrict__ slice_off,
                                                              const int32_t *__restrict__ col,
                                                              const TA *__restrict__ val, const TX *__restrict__ x,
                                                              Epi epi, double *__restrict__ partials, int64_t stride) {
    constexpr int ND = Epi::kDots > 0 ? Epi::kDots : 1;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63;
    const int s = b * kSellWaves + (threadIdx.x >> 6);
    double part[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) part[d] = 0.0;
    const int32_t row = s * 64 + lane;
    if (s < nslices && row < nrows) {
        const int32_t o0 = slice_off[s];
        const int32_t len = rowptr[row + 1] - rowptr[row];
        const int32_t *c = col + o0 + lane;
        const TA *v = val + o0 + lane;
        TAcc acc = TAcc(0);
        int j = 0;
        for (; j + 8 <= len; j += 8) {
            int ci[8];
            TA vi[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                ci[u] = ld_stream(c + (j + u) * 64);
                vi[u] = ld_stream(v + (j + u) * 64);
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) acc += (TAcc)vi[u] * (TAcc)x[ci[u]];
        }
        if (j < len) {
            int ci[8];
            TA vi[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j + u < len) {
                    ci[u] = ld_stream(c + (j + u) * 64);
                    vi[u] = ld_stream(v + (j + u) * 64);
                }
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (j + u < len) acc += (TAcc)vi[u] * (TAcc)x[ci[u]];
        }
        epi(row, acc, part);
    }
    if constexpr (Epi::kDots > 0) {
#pragma unroll
        for (int d = 0; d < Epi::kDots; ++d) {
            const double t = block_sum<64 * kSellWaves>(part[d]);
            if (threadIdx.x == 0) partials[d * stride + b] = t;
        }
    }
}

// DIA: diagonal d of row r holds A[r, r + off[d]] at val[d * ld + r]
// (ld = nrows rounded up to 64; no column indices), and bit d of mask[r] says
// whether the row stores that entry. A wavefront owns kDiaRows * 64
// consecutive rows; lane l takes rows base + 64 i + l, so every value, mask and
// x load instruction of the wave reads one contiguous segment. All loads of a
// group of diagonals are issued before the products (absent entries read x[0])
// and absent products are discarded by a select, so each row sum adds exactly
// the row's CSR entries in column order: bit-identical to the sequential CSR sum.
constexpr int kDiaWG = 256;
// rows per thread
template <typename TA>
struct DiaRows {
    static constexpr int R = 2;  // rows per thread
    static constexpr int G = 8;  // diagonals whose loads are in flight together
};
inline int32_t dia_rows_per_wg(int dtype) { return kDiaWG * (dtype == UHIP_F64 ? DiaRows<double>::R : DiaRows<float>::R); }

__host__ __device__ constexpr int32_t dia_ld(int32_t n) { return (n + 63) & ~63; }

// SYM: val holds the offsets >= 0 only (diagonals ndiag/2 .. ndiag-1); diagonal d < ndiag/2
// (offset -o) is read from the stored +o at row r - o. Those values are read twice (the
// mirror read o rows later hits in L2), so SYM loads them through the cache; the full
// layout reads every value once and streams it.
// NANSEL: the matrix holds no NaN and the entries a row does not hold are NaN in val (dia_nan):
// products are selected on the value, and the mask is not read. Taken for the full layout (a
// non-symmetric Galerkin level in FP64: measured faster); the symmetric layout keeps the mask,
// which measured faster than the select plus its mirror-before-row-0 test.
template <typename TAcc, class Epi, typename TA, typename TX, typename TM, bool SYM, bool NANSEL = false,
          int R = DiaRows<TA>::R, int G = DiaRows<TA>::G>
__global__ __launch_bounds__(kDiaWG) void k_spmv_dia(int32_t nrows, int32_t ndiag, const int32_t *__restrict__ off,
                                                     const TA *__restrict__ val, const TM *__restrict__ mask,
                                                     const TX *__restrict__ x, Epi epi, double *__restrict__ partials,
                                                     int64_t stride) {
    static_assert(!(SYM && NANSEL), "the symmetric layout selects on the mask (a mirror before row 0 reads row 0)");
    constexpr int ND = Epi::kDots > 0 ? Epi::kDots : 1;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63;
    const int32_t base = (b * (kDiaWG / 64) + (int32_t)(threadIdx.x >> 6)) * (64 * R);
    const int64_t ld = dia_ld(nrows);
    double part[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) part[d] = 0.0;
    if (base < nrows) {
        int32_t row[R];
        uint32_t m[R];
        TAcc acc[R];
#pragma unroll
        for (int i = 0; i < R; ++i) {
            row[i] = base + 64 * i + lane;
            if constexpr (NANSEL) m[i] = 0u;
            else m[i] = (uint32_t)mask[min(row[i], nrows - 1)];
            acc[i] = TAcc(0);
        }
        typename Epi::In ein[R];
#pragma unroll
        for (int i = 0; i < R; ++i) ein[i] = epi.load(min(row[i], nrows - 1));
#pragma unroll
        for (int i = 0; i < R; ++i) m[i] = row[i] < nrows ? m[i] : 0u;
        // x is read at the clamped column whether or not the entry exists (the mask
        // only selects in the sum), so no load waits for the mask
        const int32_t xlast = nrows - 1;
        const int32_t k = ndiag >> 1;
        for (int d = 0; d < ndiag; d += G) {
            TA vv[G][R];
            TX xv[G][R];
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (d + u < ndiag) {
                    const int32_t dd = d + u, o = off[dd];
                    if constexpr (SYM) {
                        const TA *vd = val + (int64_t)(dd < k ? k - dd : dd - k) * ld;
                        const int32_t sh = dd < k ? o : 0;
#pragma unroll
                        for (int i = 0; i < R; ++i) {
                            vv[u][i] = vd[min(max(row[i] + sh, 0), (int32_t)ld - 1)];
                            xv[u][i] = x[min(max(row[i] + o, 0), xlast)];
                        }
                    } else {
                        const TA *vd = val + (int64_t)dd * ld;
#pragma unroll
                        for (int i = 0; i < R; ++i) {
                            vv[u][i] = ld_stream(vd + min(row[i], (int32_t)ld - 1));
                            xv[u][i] = x[min(max(row[i] + o, 0), xlast)];
                        }
                    }
                }
            }
#pragma unroll
            for (int u = 0; u < G; ++u) {
                if (d + u < ndiag) {
#pragma unroll
                    for (int i = 0; i < R; ++i) {
                        const TAcc p = (TAcc)vv[u][i] * (TAcc)xv[u][i];
                        if constexpr (NANSEL) acc[i] = (vv[u][i] == vv[u][i]) ? acc[i] + p : acc[i];
                        else acc[i] = ((m[i] >> (d + u)) & 1u) ? acc[i] + p : acc[i];
                    }
                }
            }
        }
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (row[i] < nrows) epi.apply(row[i], acc[i], ein[i], part);
    }
    if constexpr (Epi::kDots > 0) {
#pragma unroll
        for (int d = 0; d < Epi::kDots; ++d) {
            const double t = block_sum<kDiaWG>(part[d]);
            if (threadIdx.x == 0) partials[d * stride + b] = t;
        }
    }
}

// planes per wave of the 7-point kernel (k_spmv_dia7r): FP64 walks 8 planes with the loads of
// the next 2 in flight; FP32 issues the loads of its 4 planes up front
template <typename TA>
struct Dia7Planes {
    static constexpr int P = sizeof(TA) == 8 ? 8 : 4;
    static constexpr int LA = sizeof(TA) == 8 ? 2 : 4;
};
inline int dia7_planes(int dtype) { return dtype == UHIP_F64 ? Dia7Planes<double>::P : Dia7Planes<float>::P; }

inline int32_t dia7z_workgroups(int32_t nrows, int32_t S2, int planes) {
    const int64_t slabs = ((int64_t)nrows + (int64_t)planes * S2 - 1) / ((int64_t)planes * S2);
    const int64_t waves = slabs * (S2 / 64);
    return (int32_t)((waves + kDiaWG / 64 - 1) / (kDiaWG / 64));
}

// A value from the neighbouring lane of the wavefront, by DPP: lane i gets lane i - 1's value
// (wave_shr:1) or lane i + 1's (wave_shl:1); lane 0, resp. 63, keeps its own.
__device__ __forceinline__ double lane_from_below(double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), 0x138, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), 0x138, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_from_above(double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(v), __double2loint(v), 0x130, 0xf, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(v), __double2hiint(v), 0x130, 0xf, 0xf, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float lane_from_below(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x138, 0xf, 0xf, false));
}
__device__ __forceinline__ float lane_from_above(float v) {
    return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(v), __float_as_int(v), 0x130, 0xf, 0xf, false));
}

// Epilogues that read the product's own input at row r (the x of A x) take it from the
// kernel's registers in the rolling 7-point kernel: they define load_nox (the other inputs)
// and apply_x (the row sum, those inputs and x_r).
template <class Epi, class = void>
struct EpiTakesX {
    static constexpr bool value = false;
};
template <class Epi>
struct EpiTakesX<Epi, decltype((void)&Epi::template apply_x<double>, void())> {
    static constexpr bool value = true;
};

// ±1 diagonal with its -1 mirror: the stored +1 value of rows r - 1 and r as one 2-vector load
template <typename T>
struct PairOf;
template <>
struct PairOf<double> {
    typedef double t __attribute__((ext_vector_type(2), aligned(8)));
};
template <>
struct PairOf<float> {
    typedef float t __attribute__((ext_vector_type(2), aligned(4)));
};

// Symmetric 7-point DIA (offsets -S2, -S1, -1, 0, +1, +S1, +S2, S2 a multiple of 64; stored:
// main, +1, +S1, +S2), streamed plane by plane: a wave owns 64 consecutive rows (a segment) of
// each of P consecutive planes, lane l row slab * P * S2 + z * S2 + seg * 64 + l, and walks the
// planes with the loads of the next LA planes in flight (LA = P: all issued up front). Per row:
//  - the -S2 entry is the +S2 value of the row one plane back (registers), x at +-S2 the
//    neighbouring planes' x (registers; x at +S2 is loaded with the plane and becomes x_r of
//    the next one);
//  - the -1 entry is the +1 value of row r - 1, loaded with the row's own +1 value as one
//    2-vector load (no cross-lane exchange, no extra load at the segment edge);
//  - the -S1 entry is the +S1 value of row r - S1, read through the cache (the wave S1/64
//    segments back read it);
//  - x at r -+ 1 is the neighbouring lane's x_r, moved across lanes by DPP (lanes 0 and 63 load
//    theirs): one load per row fewer (FP64 y = A x 147.6 -> 140.1 us at 256^3);
//  - entries the row does not hold are NaN in the value arrays of a NaN-free matrix (NANSEL, the
//    plan's dia_nan) and are dropped by a select on the value, or dropped by the row's mask bits.
// Every row sum adds the row's CSR entries in column order: bit-identical to the CSR sum.
template <typename TAcc, class Epi, typename TA, typename TX, int P, int LA, bool NANSEL>
__global__ __launch_bounds__(kDiaWG) void k_spmv_dia7r(int32_t nrows, const int32_t *__restrict__ off,
                                                       const TA *__restrict__ val, const uint8_t *__restrict__ mask,
                                                       const TX *__restrict__ x, Epi epi,
                                                       double *__restrict__ partials, int64_t stride) {
    constexpr int ND = Epi::kDots > 0 ? Epi::kDots : 1;
    constexpr bool TX_IN = EpiTakesX<Epi>::value;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63;
    const int64_t ld = dia_ld(nrows);
    const int32_t S1 = off[5], S2 = off[6];
    const int32_t segs = S2 >> 6;
    const int32_t wv = b * (kDiaWG / 64) + (int32_t)(threadIdx.x >> 6);
    const int32_t slab = wv / segs, seg = wv - slab * segs;
    const int64_t slab0 = (int64_t)slab * P * S2;
    double part[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) part[d] = 0.0;
    if (slab0 < nrows) {
        const TA *v0 = val, *v1 = val + ld, *vS1 = val + 2 * ld, *vS2 = val + 3 * ld;
        const TA nan = (TA)__builtin_nan("");
        const int32_t last = (int32_t)ld - 1, xlast = nrows - 1;
        auto cl = [xlast](int32_t v) { return min(max(v, 0), xlast); };
        const int32_t rb = (int32_t)slab0 + seg * 64 + lane;
        struct Plane {
            TA a0, a1, m1, aS1, aS2, mS1;
            uint32_t m;
            typename Epi::In ein;
            TX xe, xmS1, xpS1, xn;
        };
        // x and the mirrors are read at clamped rows whether or not the entry exists (the select
        // drops it), so no load waits for another; mirrors before row 0 are NaN
        auto load = [&](int z) {
            Plane q;
            const int32_t r = rb + z * S2;
            const int32_t rc = min(r, last);
            if constexpr (!NANSEL) q.m = r < nrows ? (uint32_t)mask[min(r, xlast)] : 0u;
            else q.m = 0u;
            q.a0 = ld_stream(v0 + rc);
            const typename PairOf<TA>::t w =
                __builtin_nontemporal_load((const typename PairOf<TA>::t *)(v1 + max(rc - 1, 0)));
            q.m1 = rc >= 1 ? w.x : nan;
            q.a1 = rc >= 1 ? w.y : w.x;
            q.aS1 = vS1[rc];
            const TA ms = vS1[max(rc - S1, 0)];
            q.mS1 = rc >= S1 ? ms : nan;
            q.aS2 = ld_stream(vS2 + rc);
            if constexpr (TX_IN) q.ein = epi.load_nox(min(r, xlast));
            else q.ein = epi.load(min(r, xlast));
            // x at r -+ 1 is the neighbouring lane's x_r (a DPP move); lanes 0 and 63 load theirs
            q.xe = TX(0);
            if (lane == 0 || lane == 63) q.xe = x[cl(lane == 0 ? r - 1 : r + 1)];
            q.xmS1 = x[cl(r - S1)];
            q.xpS1 = x[cl(r + S1)];
            q.xn = x[cl(r + S2)];
            return q;
        };
        const TA ps = ld_stream(vS2 + max(min(rb, last) - S2, 0));
        TA prev_aS2 = rb >= S2 ? ps : nan;
        TX prev_xc = x[cl(rb - S2)];
        TX cur_xc = x[cl(rb)];
        Plane q[LA + 1];
#pragma unroll
        for (int i = 0; i < LA; ++i) q[i] = load(i);
        // with every plane's loads issued up front (LA == P) the walk is unrolled
#pragma unroll(LA >= P ? P : 1)
        for (int z = 0; z < P; ++z) {
            if (z + LA < P) q[LA] = load(z + LA);
            const Plane &c = q[0];
            const TA vals[7] = {prev_aS2, c.mS1, c.m1, c.a0, c.a1, c.aS1, c.aS2};
            const TX fb = lane_from_below(cur_xc), fa = lane_from_above(cur_xc);
            const TX xs[7] = {prev_xc, c.xmS1, lane == 0 ? c.xe : fb, cur_xc, lane == 63 ? c.xe : fa, c.xpS1, c.xn};
            TAcc acc = TAcc(0);
#pragma unroll
            for (int d = 0; d < 7; ++d) {
                const TAcc pr = (TAcc)vals[d] * (TAcc)xs[d];
                if constexpr (NANSEL) acc = (vals[d] == vals[d]) ? acc + pr : acc;
                else acc = ((c.m >> d) & 1u) ? acc + pr : acc;
            }
            const int32_t r = rb + z * S2;
            if (r < nrows) {
                if constexpr (TX_IN) epi.apply_x(r, acc, c.ein, cur_xc, part);
                else epi.apply(r, acc, c.ein, part);
            }
            prev_aS2 = c.aS2;
            prev_xc = cur_xc;
            cur_xc = c.xn;
#pragma unroll
            for (int i = 0; i < LA; ++i) q[i] = q[i + 1];
        }
    }
    if constexpr (Epi::kDots > 0) {
#pragma unroll
        for (int d = 0; d < Epi::kDots; ++d) {
            const double t = block_sum<kDiaWG>(part[d]);
            if (threadIdx.x == 0) partials[d * stride + b] = t;
        }
    }
}

// planes per wave of the 27-point kernel (k_spmv_dia27r) and planes of loads in flight
// (measured on the 128^3 Galerkin level: 4 and 1 lead 8 / 1, 8 / 2 and 16 / 1 in FP64, tie in FP32)
template <typename TA>
struct Dia27Planes {
    static constexpr int P = 4;
    static constexpr int LA = 1;
};
inline int dia27_planes(int dtype) { return dtype == UHIP_F64 ? Dia27Planes<double>::P : Dia27Planes<float>::P; }
// fewest waves for which the 27-point walk beats the generic kernels: 1024 (one 4-wave
// workgroup per CU) - the 64^3 level of a 256^3 hierarchy; a 64 x 64 x 8 level has 128
constexpr int64_t kDia27MinWaves = 1024;
// value arrays up to this size are read through the caches by the 27-point kernel (half the
// MI355X's 256 MB MALL: the level's vectors share it). Measured at 256^3: the FP32 128^3 level
// (117 MB) gains 2.3 ms per MG-PCG solve; the FP64 one (453 MB in its full layout) lost 5.5 ms.
constexpr int64_t kDiaCachedBytes = int64_t(128) << 20;

// 27-point DIA (offsets dz S2 + dy S1 + dx for dz, dy, dx in {-1, 0, 1}, S2 a multiple of 64:
// a Galerkin operator of a 3D grid), streamed plane by plane as in k_spmv_dia7r: a wave owns 64
// consecutive rows of each of P consecutive planes and walks the planes, the loads of the next
// LA planes in flight. x at the nine (dy, dx) neighbours of the plane ahead is loaded once per
// plane and rolls through registers for the row's two other planes.
// SYM (the plan keeps the offsets >= 0, values symmetric bit for bit): the -o entry of row r is
// the stored +o value of row r - o:
//  - -1, and the -S2 +- 1 pair of the next plane, come with a row's own +1, +S2 - 1 and +S2 + 1
//    values as 2-vector loads (rows r - 1, r or r, r + 1);
//  - -S2 is the row's +S2 value one plane back (registers);
//  - -S1 +- 1, -S1 and the six -S2 +- S1 (+- 1) entries are read through the cache (the waves
//    owning those rows read them as their own);
// absent entries are dropped by the row's mask bits. The full layout streams all 27 values and
// selects on the NaN of absent entries (NANSEL) or on the mask. Every row sum adds the row's
// CSR entries in column order: bit-identical to the CSR sum.
// CACHED: the values are read through the caches instead of streamed (non-temporal), so that
// value arrays small enough to stay in the 256 MB last-level cache (MALL) are served from it
// by the next product on the same level (a V-cycle runs four per level).
// DELTA (SYM, values symmetric only up to rounding): lower entry d of row r is its mirror with
// delta[d][r] added to the bit pattern (int32 for FP64, int8 for FP32), exact.
template <typename TA>
struct Dia27Delta {
    typedef int32_t t;
    static __device__ __forceinline__ double apply(double v, int32_t d) {
        return __longlong_as_double(__double_as_longlong(v) + (long long)d);
    }
};
template <>
struct Dia27Delta<float> {
    typedef int8_t t;
    static __device__ __forceinline__ float apply(float v, int8_t d) { return __int_as_float(__float_as_int(v) + (int)d); }
};
template <typename TAcc, class Epi, typename TA, typename TX, int P, int LA, bool SYM, bool NANSEL, bool CACHED,
          bool DELTA = false>
__global__ __launch_bounds__(kDiaWG) void k_spmv_dia27r(int32_t nrows, int32_t S1, int32_t S2,
                                                        const TA *__restrict__ val, const uint32_t *__restrict__ mask,
                                                        const typename Dia27Delta<TA>::t *__restrict__ delta,
                                                        const TX *__restrict__ x, Epi epi,
                                                        double *__restrict__ partials, int64_t stride) {
    static_assert(!(SYM && NANSEL), "the symmetric layout selects on the mask");
    static_assert(!DELTA || SYM, "bit-pattern deltas belong to symmetric storage");
    constexpr int ND = Epi::kDots > 0 ? Epi::kDots : 1;
    constexpr bool TX_IN = EpiTakesX<Epi>::value;
    constexpr int NV = SYM ? 14 : 27;
    typedef typename PairOf<TA>::t Pair;
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63;
    const int64_t ld = dia_ld(nrows);
    const int32_t segs = S2 >> 6;
    const int32_t wv = b * (kDiaWG / 64) + (int32_t)(threadIdx.x >> 6);
    const int32_t slab = wv / segs, seg = wv - slab * segs;
    const int64_t slab0 = (int64_t)slab * P * S2;
    double part[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) part[d] = 0.0;
    if (slab0 < nrows) {
        const int32_t last = (int32_t)ld - 1, xlast = nrows - 1;
        auto cl = [xlast](int32_t v) { return min(max(v, 0), xlast); };
        auto V = [&](int k) { return val + (int64_t)k * ld; };
        auto lv = [](const TA *p) -> TA {
            if constexpr (CACHED) return *p;
            else return ld_stream(p);
        };
        auto lp = [](const TA *p) -> Pair {
            if constexpr (CACHED) return *(const Pair *)p;
            else return __builtin_nontemporal_load((const Pair *)p);
        };
        const int32_t rb = (int32_t)slab0 + seg * 64 + lane;
        struct Plane {
            TA v[NV];
            TA m1, nm1, np1, my[3], mz[6];  // SYM: mirrors
            typename Dia27Delta<TA>::t dl[DELTA ? 13 : 1];  // DELTA: the lower entries' bit-pattern differences
            uint32_t m;
            typename Epi::In ein;
            TX xn[9];
        };
        // values, mirrors and x are read at clamped rows whether or not the entry exists (the
        // select drops it), so no load waits for another
        auto load = [&](int z) {
            Plane q;
            const int32_t r = rb + z * S2;
            const int32_t rc = min(r, last);
            if constexpr (!NANSEL) q.m = r < nrows ? mask[min(r, xlast)] : 0u;
            else q.m = 0u;
            if constexpr (SYM) {
                // stored k: offsets 0, 1, S1 - 1, S1, S1 + 1, S2 - S1 - 1, S2 - S1, S2 - S1 + 1,
                // S2 - 1, S2, S2 + 1, S2 + S1 - 1, S2 + S1, S2 + S1 + 1
                q.v[0] = lv(V(0) + rc);
                const Pair w1 = lp(V(1) + max(rc - 1, 0));
                q.m1 = w1.x;
                q.v[1] = rc >= 1 ? w1.y : w1.x;
                q.v[2] = V(2)[rc];
                q.v[3] = V(3)[rc];
                q.v[4] = V(4)[rc];
                q.my[0] = V(4)[max(rc - S1 - 1, 0)];
                q.my[1] = V(3)[max(rc - S1, 0)];
                q.my[2] = V(2)[max(rc - S1 + 1, 0)];
                q.v[5] = V(5)[rc];
                q.v[6] = V(6)[rc];
                q.v[7] = V(7)[rc];
                const Pair w8 = lp(V(8) + min(rc, last - 1));
                q.v[8] = rc < last ? w8.x : w8.y;
                q.np1 = w8.y;
                q.v[9] = lv(V(9) + rc);
                const Pair w10 = lp(V(10) + max(rc - 1, 0));
                q.nm1 = w10.x;
                q.v[10] = rc >= 1 ? w10.y : w10.x;
                q.v[11] = V(11)[rc];
                q.v[12] = V(12)[rc];
                q.v[13] = V(13)[rc];
                q.mz[0] = V(13)[max(rc - S2 - S1 - 1, 0)];
                q.mz[1] = V(12)[max(rc - S2 - S1, 0)];
                q.mz[2] = V(11)[max(rc - S2 - S1 + 1, 0)];
                q.mz[3] = V(7)[max(rc - S2 + S1 - 1, 0)];
                q.mz[4] = V(6)[max(rc - S2 + S1, 0)];
                q.mz[5] = V(5)[max(rc - S2 + S1 + 1, 0)];
                if constexpr (DELTA)
#pragma unroll
                    for (int d = 0; d < 13; ++d) q.dl[d] = ld_stream(delta + (int64_t)d * ld + rc);
            } else {
#pragma unroll
                for (int d = 0; d < 27; ++d) q.v[d] = lv(V(d) + rc);
            }
            if constexpr (TX_IN) q.ein = epi.load_nox(min(r, xlast));
            else q.ein = epi.load(min(r, xlast));
#pragma unroll
            for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
                for (int dx = -1; dx <= 1; ++dx) q.xn[3 * (dy + 1) + dx + 1] = x[cl(r + S2 + dy * S1 + dx)];
            return q;
        };
        // the -S2 - 1, -S2 and -S2 + 1 values of the first plane: the previous plane's +S2 + 1
        // value of row r - S2 - 1, +S2 of row r - S2, +S2 - 1 of row r - S2 + 1
        TA cnm1 = TA(0), ca9 = TA(0), cnp1 = TA(0);
        if constexpr (SYM) {
            const int32_t rp = min(max(rb - S2, 0), last);
            cnm1 = V(10)[max(rp - 1, 0)];
            ca9 = V(9)[rp];
            cnp1 = V(8)[min(rp + 1, last)];
        }
        TX xp[9], xc[9];
#pragma unroll
        for (int dy = -1; dy <= 1; ++dy)
#pragma unroll
            for (int dx = -1; dx <= 1; ++dx) {
                xp[3 * (dy + 1) + dx + 1] = x[cl(rb - S2 + dy * S1 + dx)];
                xc[3 * (dy + 1) + dx + 1] = x[cl(rb + dy * S1 + dx)];
            }
        Plane q[LA + 1];
#pragma unroll
        for (int i = 0; i < LA; ++i) q[i] = load(i);
#pragma unroll(LA >= P ? P : 1)
        for (int z = 0; z < P; ++z) {
            if (z + LA < P) q[LA] = load(z + LA);
            const Plane &c = q[0];
            TA vals[27];
            if constexpr (SYM) {
                const TA lo[13] = {c.mz[0], c.mz[1], c.mz[2], cnm1,    ca9,     cnp1,   c.mz[3],
                                   c.mz[4], c.mz[5], c.my[0], c.my[1], c.my[2], c.m1};
#pragma unroll
                for (int d = 0; d < 13; ++d) {
                    if constexpr (DELTA) vals[d] = Dia27Delta<TA>::apply(lo[d], c.dl[d]);
                    else vals[d] = lo[d];
                }
#pragma unroll
                for (int d = 0; d < 14; ++d) vals[13 + d] = c.v[d];
            } else {
#pragma unroll
                for (int d = 0; d < 27; ++d) vals[d] = c.v[d];
            }
            TAcc acc = TAcc(0);
#pragma unroll
            for (int d = 0; d < 27; ++d) {
                const TX xd = d < 9 ? xp[d] : (d < 18 ? xc[d - 9] : c.xn[d - 18]);
                const TAcc pr = (TAcc)vals[d] * (TAcc)xd;
                if constexpr (NANSEL) acc = (vals[d] == vals[d]) ? acc + pr : acc;
                else acc = ((c.m >> d) & 1u) ? acc + pr : acc;
            }
            const int32_t r = rb + z * S2;
            if (r < nrows) {
                if constexpr (TX_IN) epi.apply_x(r, acc, c.ein, xc[4], part);
                else epi.apply(r, acc, c.ein, part);
            }
            if constexpr (SYM) {
                cnm1 = c.nm1;
                ca9 = c.v[9];
                cnp1 = c.np1;
            }
#pragma unroll
            for (int k = 0; k < 9; ++k) {
                xp[k] = xc[k];
                xc[k] = c.xn[k];
            }
#pragma unroll
            for (int i = 0; i < LA; ++i) q[i] = q[i + 1];
        }
    }
    if constexpr (Epi::kDots > 0) {
#pragma unroll
        for (int d = 0; d < Epi::kDots; ++d) {
            const double t = block_sum<kDiaWG>(part[d]);
            if (threadIdx.x == 0) partials[d * stride + b] = t;
        }
    }
}

// Rows per lane of the VI kernel: for at most 8 slots 4 (FP32) or 2 (FP64), else 1
// (a 27-slot restriction row is latency-bound: more rows per lane only lengthen it).
template <typename TA>
constexpr int vi_rows(int32_t ndiag) {
    return ndiag <= 8 ? (sizeof(TA) == 8 ? 2 : 4) : 1;
}
inline int vi_rows_per_lane(int dtype, int32_t ndiag) {
    return dtype == UHIP_F64 ? vi_rows<double>(ndiag) : vi_rows<float>(ndiag);
}

// VI (value-indexed, row-relative): slot d of row r holds the entry at column
// base[r] + off[d] (bit d of mask[r]) with value table[code], the code being byte
// d % 8 of the 64-bit word codes[r * NW + d / 8]; the table (at most 256 values)
// is staged in LDS. A row's codes arrive in NW wide loads, a slot that no row of
// the wavefront holds costs nothing (wave-uniform branch), and the masked,
// column-ordered sum makes each row sum the sequential CSR sum bit for bit.
// SIG (vi_nsig > 0): mask holds one uint8 signature id per row and codes the signatures (mask
// word, then the NW code words), staged in LDS.
template <typename TAcc, class Epi, typename TA, typename TX, typename TM, int R, bool SIG = false>
__global__ __launch_bounds__(kDiaWG) void k_spmv_vi(int32_t nrows, int32_t ndiag, const int32_t *__restrict__ off,
                                                    const uint64_t *__restrict__ codes, const TA *__restrict__ table,
                                                    int32_t nvals, const int32_t *__restrict__ vbase,
                                                    const TM *__restrict__ mask, const TX *__restrict__ x, Epi epi,
                                                    double *__restrict__ partials, int64_t stride, int32_t nsig) {
    constexpr int ND = Epi::kDots > 0 ? Epi::kDots : 1;
    constexpr int NW = (int)sizeof(TM);  // 64-bit code words per row: 8 slots each
    constexpr bool kFlat = R == 1 && NW >= 2;
    __shared__ TA tab[256];
    __shared__ int32_t so[32];
    __shared__ uint64_t stab[SIG ? 256 * (NW + 1) : 1];
    const uint8_t *sid = (const uint8_t *)mask;
    for (int i = threadIdx.x; i < nvals; i += kDiaWG) tab[i] = table[i];
    if constexpr (SIG)
        for (int i = threadIdx.x; i < nsig * (NW + 1); i += kDiaWG) stab[i] = codes[i];
    if (kFlat && threadIdx.x < 32) so[threadIdx.x] = (int32_t)threadIdx.x < ndiag ? off[threadIdx.x] : 0;
    __syncthreads();
    auto row_mask = [&](int32_t rc) -> uint32_t {
        if constexpr (SIG) return (uint32_t)stab[(int)sid[rc] * (NW + 1)];
        else return (uint32_t)mask[rc];
    };
    auto row_code = [&](int32_t rc, int w) -> uint64_t {
        if constexpr (SIG) return stab[(int)sid[rc] * (NW + 1) + 1 + w];
        else return ld_stream(codes + (int64_t)rc * NW + w);
    };
    const int b = xcd_remap(blockIdx.x, gridDim.x);
    const int lane = threadIdx.x & 63;
    if constexpr (kFlat) {
        // Wide rows (a restriction's 27 slots), one row per lane and no branch: all code
        // words, then all 8 * NW gathers (an absent slot reads x at the row's base column,
        // always in bounds), then the masked sum in slot order. Every load is in flight
        // before the first use.
        double part[ND];
#pragma unroll
        for (int d = 0; d < ND; ++d) part[d] = 0.0;
        const int32_t row = b * kDiaWG + (int32_t)threadIdx.x;
        const int32_t rc = min(row, nrows - 1);
        const uint32_t m = row < nrows ? row_mask(rc) : 0u;
        const int32_t rb = vbase[rc];
        const typename Epi::In ein = epi.load(rc);
        uint64_t c[NW];
#pragma unroll
        for (int w = 0; w < NW; ++w) c[w] = row_code(rc, w);
        TX xv[8 * NW];
#pragma unroll
        for (int d = 0; d < 8 * NW; ++d) xv[d] = d < ndiag ? x[((m >> d) & 1u) ? rb + so[d] : rb] : TX(0);
        TAcc acc = TAcc(0);
#pragma unroll
        for (int d = 0; d < 8 * NW; ++d) {
            const TAcc p = (TAcc)tab[(uint32_t)(c[d / 8] >> (8 * (d % 8))) & 0xffu] * (TAcc)xv[d];
            acc = ((m >> d) & 1u) ? acc + p : acc;
        }
        if (row < nrows) epi.apply(row, acc, ein, part);
        if constexpr (Epi::kDots > 0) {
#pragma unroll
            for (int d = 0; d < Epi::kDots; ++d) {
                const double t = block_sum<kDiaWG>(part[d]);
                if (threadIdx.x == 0) partials[d * stride + b] = t;
            }
        }
        return;
    }
    const int32_t base = (b * (kDiaWG / 64) + (int32_t)(threadIdx.x >> 6)) * (64 * R);
    double part[ND];
#pragma unroll
    for (int d = 0; d < ND; ++d) part[d] = 0.0;
    if (base < nrows) {
        int32_t row[R], rb[R];
        uint32_t m[R];
        TAcc acc[R];
        // Loads from a clamped row (no branch per row, so they are all in flight
        // together); rows past the end then get an empty mask.
        uint32_t wm = 0;
#pragma unroll
        for (int i = 0; i < R; ++i) {
            row[i] = base + 64 * i + lane;
            const int32_t rc = min(row[i], nrows - 1);
            m[i] = row_mask(rc);
            rb[i] = vbase[rc];
            acc[i] = TAcc(0);
        }
        typename Epi::In ein[R];
#pragma unroll
        for (int i = 0; i < R; ++i) ein[i] = epi.load(min(row[i], nrows - 1));
#pragma unroll
        for (int i = 0; i < R; ++i) {
            m[i] = row[i] < nrows ? m[i] : 0u;
            wm |= m[i];
        }
#pragma unroll
        for (int sh = 1; sh < 64; sh <<= 1) wm |= (uint32_t)__shfl_xor((int)wm, sh);
        wm = __builtin_amdgcn_readfirstlane(wm);  // slots held by some row of the wavefront
#pragma unroll
        for (int w = 0; w < NW; ++w) {
            if (8 * w >= ndiag) break;
            uint64_t c[R];
#pragma unroll
            for (int i = 0; i < R; ++i) c[i] = row_code(min(row[i], nrows - 1), w);
            TX xv[8][R];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int d = 8 * w + u;
                if (d < ndiag && ((wm >> d) & 1u)) {
                    const int32_t o = off[d];
#pragma unroll
                    for (int i = 0; i < R; ++i) xv[u][i] = ((m[i] >> d) & 1u) ? x[rb[i] + o] : TX(0);
                } else {
#pragma unroll
                    for (int i = 0; i < R; ++i) xv[u][i] = TX(0);
                }
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int d = 8 * w + u;
#pragma unroll
                for (int i = 0; i < R; ++i) {
                    const TAcc p = (TAcc)tab[(uint32_t)(c[i] >> (8 * u)) & 0xffu] * (TAcc)xv[u][i];
                    acc[i] = ((m[i] >> d) & 1u) ? acc[i] + p : acc[i];
                }
            }
        }
#pragma unroll
        for (int i = 0; i < R; ++i)
            if (row[i] < nrows) epi.apply(row[i], acc[i], ein[i], part);
    }
    if constexpr (Epi::kDots > 0) {
#pragma unroll
        for (int d = 0; d < Epi::kDots; ++d) {
            const double t = block_sum<kDiaWG>(part[d]);
            if (threadIdx.x == 0) partials[d * stride + b] = t;
        }
    }
}

// Dispatch over the plan's format and LDS budget (block_nnz = 4/5 * CAPK * WG).
// Matrix values have the accumulation type TAcc.
// With parts, the per-workgroup partial sums of a dot-carrying epilogue go to parts (cap
// entries per dot) instead of the context's workspace, which is then neither read nor grown.
template <typename TAcc, typename TX, class Epi>
int launch_rowblock(uhip_ctx ctx, const uhip_csr *A, const TX *x, const Epi &epi, double *parts_in = nullptr,
                    int64_t cap = 0) {
    using TA = TAcc;
    if (parts_in && Epi::kDots > 0) {
        const int64_t need = A->format == UHIP_FMT_ROWBLOCK ? A->nblocks : A->nparts;
        UHIP_REQUIRE(need <= cap, "spmv: the plan needs %lld partial sums, the buffer holds %lld", (long long)need,
                     (long long)cap);
    }
    // partial sums: the caller's buffer, else the context's (grown on demand)
    auto partials_for = [&](int64_t need) -> int {
        if (parts_in || Epi::kDots == 0) return UHIP_OK;
        return ensure_partials(ctx, need);
    };
    if (A && A->format == UHIP_FMT_VI) {
        UHIP_REQUIRE(A->dia_off && A->dia_val && A->dia_mask && A->vi_base && A->vi_table && A->ndiag >= 1 &&
                         A->ndiag <= 32 && A->vi_nvals >= 1 && A->vi_nvals <= 256,
                     "spmv: VI plan missing");
        if (A->nrows == 0) return UHIP_OK;
        UHIP_TRY(partials_for(A->nparts));
        double *parts = parts_in ? parts_in : ctx->partials;
        const int64_t pstride = parts_in ? cap : ctx->max_partials;
        const dim3 g(A->nparts), blk(kDiaWG);
#define UHIP_VI_LAUNCH2(TM, RR, SIG)                                                                            \
    hipLaunchKernelGGL((k_spmv_vi<TAcc, Epi, TA, TX, TM, RR, SIG>), g, blk, 0, stream_of(ctx), A->nrows,          \
                       A->ndiag, A->dia_off, (const uint64_t *)A->dia_val, (const TA *)A->vi_table, A->vi_nvals, \
                       A->vi_base, (const TM *)A->dia_mask, x, epi, parts, pstride, A->vi_nsig)
#define UHIP_VI_LAUNCH(TM, RR)                                                                                  \
    do {                                                                                                        \
        if (A->vi_nsig > 0) UHIP_VI_LAUNCH2(TM, RR, true);                                                      \
        else UHIP_VI_LAUNCH2(TM, RR, false);                                                                    \
    } while (0)
        if (A->ndiag <= 8) UHIP_VI_LAUNCH(uint8_t, vi_rows<TA>(8));
        else if (A->ndiag <= 16) UHIP_VI_LAUNCH(uint16_t, 1);
        else UHIP_VI_LAUNCH(uint32_t, 1);
#undef UHIP_VI_LAUNCH
#undef UHIP_VI_LAUNCH2
        UHIP_CHECK_LAUNCH();
        return UHIP_OK;
    }
    if (A && A->format == UHIP_FMT_DIA) {
        UHIP_REQUIRE(A->dia_off && A->dia_val && A->dia_mask && A->ndiag >= 1 && A->ndiag <= 32,
                     "spmv: DIA plan missing");
        UHIP_REQUIRE(!A->dia_delta || (A->dia_w27 && A->dia_sym == 1),
                     "spmv: DIA lower-triangle deltas need the 27-point kernel");
        if (A->nrows == 0) return UHIP_OK;
        UHIP_TRY(partials_for(A->nparts));
        double *parts = parts_in ? parts_in : ctx->partials;
        const int64_t pstride = parts_in ? cap : ctx->max_partials;
        const dim3 g(A->nparts), blk(kDiaWG);
        const TA *v = (const TA *)A->dia_val;
#define UHIP_DIA_LAUNCH2(TM, SYM, NANSEL)                                                                         \
    hipLaunchKernelGGL((k_spmv_dia<TAcc, Epi, TA, TX, TM, SYM, NANSEL>), g, blk, 0, stream_of(ctx), A->nrows,     \
                       A->ndiag, A->dia_off, v, (const TM *)A->dia_mask, x, epi, parts, pstride)
#define UHIP_DIA_LAUNCH_FULL(TM)                                                                                  \
    do {                                                                                                          \
        if (A->dia_nan) UHIP_DIA_LAUNCH2(TM, false, true);                                                        \
        else UHIP_DIA_LAUNCH2(TM, false, false);                                                                  \
    } while (0)
        // symmetric 7-point storage: the plane-walking kernel (its grid is nparts, the partial
        // sums the finish reads), selecting on NaN-encoded absent entries when the plan has them
        if (A->dia_w27) {
            constexpr int P = Dia27Planes<TA>::P, LA = Dia27Planes<TA>::LA;
#define UHIP_DIA27_LAUNCH3(SYM, NANSEL, CACHED, DELTA)                                                            \
    hipLaunchKernelGGL((k_spmv_dia27r<TAcc, Epi, TA, TX, P, LA, SYM, NANSEL, CACHED, DELTA>), g, blk, 0,            \
                       stream_of(ctx), A->nrows, A->dia_s1, A->dia_s2, v, (const uint32_t *)A->dia_mask,          \
                       (const typename Dia27Delta<TA>::t *)A->dia_delta, x, epi, parts, pstride)
#define UHIP_DIA27_LAUNCH2(SYM, NANSEL, CACHED) UHIP_DIA27_LAUNCH3(SYM, NANSEL, CACHED, false)
#define UHIP_DIA27_LAUNCH(SYM, NANSEL)                                                                            \
    do {                                                                                                          \
        if (cached) UHIP_DIA27_LAUNCH2(SYM, NANSEL, true);                                                        \
        else UHIP_DIA27_LAUNCH2(SYM, NANSEL, false);                                                              \
    } while (0)
            const bool cached = (int64_t)(A->dia_sym ? 14 : 27) * dia_ld(A->nrows) * (int64_t)sizeof(TA) <= kDiaCachedBytes;
            if (A->dia_delta) {
                if (cached) UHIP_DIA27_LAUNCH3(true, false, true, true);
                else UHIP_DIA27_LAUNCH3(true, false, false, true);
            } else if (A->dia_sym) UHIP_DIA27_LAUNCH(true, false);
            else if (A->dia_nan) UHIP_DIA27_LAUNCH(false, true);
            else UHIP_DIA27_LAUNCH(false, false);
#undef UHIP_DIA27_LAUNCH
#undef UHIP_DIA27_LAUNCH2
#undef UHIP_DIA27_LAUNCH3
        } else if (A->dia_sym == 2) {
            constexpr int P = Dia7Planes<TA>::P, LA = Dia7Planes<TA>::LA;
            if (A->dia_nan)
                hipLaunchKernelGGL((k_spmv_dia7r<TAcc, Epi, TA, TX, P, LA, true>), g, blk, 0, stream_of(ctx), A->nrows,
                                   A->dia_off, v, (const uint8_t *)A->dia_mask, x, epi, parts, pstride);
            else
                hipLaunchKernelGGL((k_spmv_dia7r<TAcc, Epi, TA, TX, P, LA, false>), g, blk, 0, stream_of(ctx), A->nrows,
                                   A->dia_off, v, (const uint8_t *)A->dia_mask, x, epi, parts, pstride);
        } else if (A->ndiag <= 8) {
            if (A->dia_sym) UHIP_DIA_LAUNCH2(uint8_t, true, false);
            else UHIP_DIA_LAUNCH_FULL(uint8_t);
        } else if (A->ndiag <= 16) {
            if (A->dia_sym) UHIP_DIA_LAUNCH2(uint16_t, true, false);
            else UHIP_DIA_LAUNCH_FULL(uint16_t);
        } else {
            if (A->dia_sym) UHIP_DIA_LAUNCH2(uint32_t, true, false);
            else UHIP_DIA_LAUNCH_FULL(uint32_t);
        }
#undef UHIP_DIA_LAUNCH_FULL
#undef UHIP_DIA_LAUNCH2
        UHIP_CHECK_LAUNCH();
        return UHIP_OK;
    }
    if (A && A->format == UHIP_FMT_SELL64) {
        UHIP_REQUIRE(A->slice_off && A->sell_col && A->sell_val, "spmv: SELL-64 plan missing");
        if (A->nslices == 0) return UHIP_OK;
        UHIP_TRY(partials_for(A->nparts));
        double *parts = parts_in ? parts_in : ctx->partials;
        const int64_t pstride = parts_in ? cap : ctx->max_partials;
        hipLaunchKernelGGL((k_spmv_sell<TAcc, Epi, TA, TX>), dim3(A->nparts), dim3(64 * kSellWaves), 0,
                           stream_of(ctx), A->nrows, A->nslices, A->rowptr, A->slice_off, A->sell_col,
                           (const TA *)A->sell_val, x, epi, parts, pstride);
        UHIP_CHECK_LAUNCH();
        return UHIP_OK;
    }
    UHIP_REQUIRE(A && A->blocks && A->nblocks >= 0, "spmv: matrix has no row-block plan");
    if (A->nblocks == 0) return UHIP_OK;
    UHIP_TRY(partials_for(A->nblocks));
    const int capk = A->block_capk;
    const TA *val = (const TA *)A->val;
    hipStream_t s = stream_of(ctx);
    dim3 g(A->nblocks), blk(kSpmvWG);
    double *parts_rb = parts_in ? parts_in : ctx->partials;
    const int64_t st = parts_in ? cap : ctx->max_partials;
    switch (capk) {
        case 5:
            hipLaunchKernelGGL((k_spmv_rowblock<TAcc, 5, Epi, TA, TX>), g, blk, 0, s, A->blocks, A->rowptr, A->col,
                               val, x, epi, parts_rb, st);
            break;
        case 10:
            hipLaunchKernelGGL((k_spmv_rowblock<TAcc, 10, Epi, TA, TX>), g, blk, 0, s, A->blocks, A->rowptr, A->col,
                               val, x, epi, parts_rb, st);
            break;
        case 20:
            hipLaunchKernelGGL((k_spmv_rowblock<TAcc, 20, Epi, TA, TX>), g, blk, 0, s, A->blocks, A->rowptr, A->col,
                               val, x, epi, parts_rb, st);
            break;
        default:
            return set_error(UHIP_ERR_INVALID, "spmv: unsupported plan capk=%d", capk);
    }
    UHIP_CHECK_LAUNCH();
    return UHIP_OK;
}

// ---- epilogues ------------------------------------------------------------
// Each epilogue splits into load(r) - the row's vector inputs, which the DIA, plane-
// blocked and VI kernels issue together with the matrix loads - and apply(r, s, in,
// part), which combines them with the row sum s and stores. operator() does both.

struct EpiNoIn {};

// y = s
template <typename TAcc, typename TY>
struct EpiStore {
    static constexpr int kDots = 0;
    using In = EpiNoIn;
    TY *y;
    __device__ In load(int) const { return In{}; }
    __device__ void apply(int r, TAcc s, const In &, double *) const { st_stream(y + r, (TY)s); }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
};

// y = alpha*s + beta*y   (beta == 0 never reads y)
template <typename TAcc, typename TY>
struct EpiAxpby {
    static constexpr int kDots = 0;
    struct In {
        TAcc y;
    };
    TY *y;
    TAcc alpha, beta;
    __device__ In load(int r) const { return In{beta == TAcc(0) ? TAcc(0) : (TAcc)y[r]}; }
    __device__ void apply(int r, TAcc s, const In &in, double *) const {
        st_stream(y + r, (beta == TAcc(0)) ? (TY)(alpha * s) : (TY)(alpha * s + beta * in.y));
    }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
};

// y = alpha*s + beta*z
template <typename TAcc, typename TY>
struct EpiAxpbz {
    static constexpr int kDots = 0;
    struct In {
        TAcc z;
    };
    const TY *z;
    TY *y;
    TAcc alpha, beta;
    __device__ In load(int r) const { return In{(TAcc)z[r]}; }
    __device__ void apply(int r, TAcc s, const In &in, double *) const {
        st_stream(y + r, (TY)(alpha * s + beta * in.z));
    }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
};

// r = b - s
template <typename TAcc, typename TB, typename TY>
struct EpiResidual {
    static constexpr int kDots = 0;
    struct In {
        TAcc b;
    };
    const TB *b;
    TY *y;
    __device__ In load(int r) const { return In{(TAcc)ld_stream(b + r)}; }
    __device__ void apply(int r, TAcc s, const In &in, double *) const { st_stream(y + r, (TY)(in.b - s)); }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
};

// y = s, part += x_r * s  (p.Ap of CG); x is the product's input
template <typename TAcc, typename TX, typename TY>
struct EpiStoreDot {
    static constexpr int kDots = 1;
    struct In {
        double x;
    };
    const TX *x;
    TY *y;
    __device__ In load(int r) const { return In{(double)x[r]}; }
    __device__ void apply(int r, TAcc s, const In &in, double *part) const {
        st_stream(y + r, (TY)s);
        part[0] += in.x * (double)s;
    }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
    __device__ In load_nox(int) const { return In{0.0}; }
    template <typename TXV>
    __device__ void apply_x(int r, TAcc s, const In &, TXV xr, double *part) const {
        apply(r, s, In{(double)xr}, part);
    }
};

// y = s, part[0] += x_r s, part[1] += z_r x_r   (A x with x.Ax and z.x in one pass)
template <typename TAcc, typename TY, typename TX>
struct EpiStoreDot2 {
    static constexpr int kDots = 2;
    struct In {
        double x, z;
    };
    const TX *x;
    const double *z;
    TY *y;
    __device__ In load(int r) const { return In{(double)x[r], ld_stream(z + r)}; }
    __device__ void apply(int r, TAcc s, const In &in, double *part) const {
        st_stream(y + r, (TY)s);
        part[0] += in.x * (double)s;
        part[1] += in.z * in.x;
    }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
    __device__ In load_nox(int r) const { return In{0.0, ld_stream(z + r)}; }
    template <typename TXV>
    __device__ void apply_x(int r, TAcc s, const In &in, TXV xr, double *part) const {
        apply(r, s, In{(double)xr, in.z}, part);
    }
};

// y = mask .* s  (restriction followed by MultiLevelMask::apply); mask may be null
template <typename TAcc, typename TY>
struct EpiMaskStore {
    static constexpr int kDots = 0;
    struct In {
        TAcc m;
    };
    const TY *mask;
    TY *y;
    __device__ In load(int r) const { return In{mask ? (TAcc)mask[r] : TAcc(1)}; }
    __device__ void apply(int r, TAcc s, const In &in, double *) const {
        st_stream(y + r, mask ? (TY)(in.m * s) : (TY)s);
    }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
};

// y = y + s  (coarse-grid correction c += I c_c)
template <typename TAcc, typename TY>
struct EpiAddTo {
    static constexpr int kDots = 0;
    struct In {
        TAcc y;
    };
    TY *y;
    __device__ In load(int r) const { return In{(TAcc)y[r]}; }
    __device__ void apply(int r, TAcc s, const In &in, double *) const { st_stream(y + r, (TY)(in.y + s)); }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
};

// One Jacobi sweep fused with its SpMV (Jacobi::sweep, solvers/smoothers/utopia_Jacobi.hpp:132-138):
// with s = (A x)_r, r = b_r - s, then out_r = x_r + d_r*r (ADD, x read from xin; null xin is
// x = 0); without ADD out_r = d_r*r (PointJacobi::sweep with A = LU,
// solvers/smoothers/utopia_PointJacobi.hpp:132-142). The residual, the product and the sum
// are rounded one by one as the reference's three vector operations round them.
template <typename TAcc, typename TB, typename TXO, bool ADD>
struct EpiJacobi {
    static constexpr int kDots = 0;
    struct In {
        TAcc b, d, x;
    };
    const TB *b;
    const TAcc *d;
    const TAcc *xin;
    TXO *out;
    __device__ In load(int r) const {
        return In{(TAcc)ld_stream(b + r), ld_stream(d + r), (ADD && xin) ? xin[r] : TAcc(0)};
    }
    // xin is the product's input: the rolling 7-point kernel hands x_r over
    __device__ In load_nox(int r) const { return In{(TAcc)ld_stream(b + r), ld_stream(d + r), TAcc(0)}; }
    template <typename TXV>
    __device__ void apply_x(int r, TAcc s, const In &in0, TXV xr, double *part) const {
        In in = in0;
        in.x = ADD ? (TAcc)xr : TAcc(0);
        apply(r, s, in, part);
    }
    __device__ void apply(int r, TAcc s, const In &in, double *) const {
        const TAcc res = in.b - s;
        const TAcc t = in.d * res;
        st_stream(out + r, (TXO)(ADD ? in.x + t : t));
    }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
};

// One step of Chebyshev3level::solve (utopia_Chebyshev3level.hpp:138-175) fused
// with the SpMV that feeds it. With s = (A p_{k-1})_r:
//   r_k = r_{k-1} + a_prev*s ; p_k = -r_k + beta*p_{k-1} ; x += a_prev*p_{k-1}
// and on the last step also x += a_cur*p_k. rprev == null means r_{k-1} = -p_{k-1}
// (first step), xin == null means x = 0 (zero initial guess).
// With DOT, part += xout_r * dot_with_r (the z.r of the CG step that follows a
// preconditioner application, computed while z is written).
template <typename TAcc, typename TP, typename TV, typename TXI, typename TXO, bool DOT = false>
struct EpiCheb {
    static constexpr int kDots = DOT ? 1 : 0;
    struct In {
        TAcc pp, rp, x;
        double dw;
    };
    const TP *pprev;
    const TV *rprev;
    const TXI *xin;
    TXO *xout;
    TV *rout;
    TV *pout;
    TAcc a_prev, beta, a_cur;
    int last;
    const double *dot_with;
    __device__ In load(int r) const {
        In in;
        in.pp = (TAcc)pprev[r];
        in.rp = rprev ? (TAcc)ld_stream(rprev + r) : -in.pp;
        in.x = xin ? (TAcc)ld_stream(xin + r) : TAcc(0);
        in.dw = DOT ? ld_stream(dot_with + r) : 0.0;
        return in;
    }
    // pprev is the product's input: the rolling 7-point kernel hands p_{k-1} at row r over
    __device__ In load_nox(int r) const {
        In in;
        in.pp = TAcc(0);
        in.rp = rprev ? (TAcc)ld_stream(rprev + r) : TAcc(0);
        in.x = xin ? (TAcc)ld_stream(xin + r) : TAcc(0);
        in.dw = DOT ? ld_stream(dot_with + r) : 0.0;
        return in;
    }
    template <typename TXV>
    __device__ void apply_x(int r, TAcc s, const In &in0, TXV xr, double *part) const {
        In in = in0;
        in.pp = (TAcc)xr;
        if (!rprev) in.rp = -in.pp;
        apply(r, s, in, part);
    }
    __device__ void apply(int r, TAcc s, const In &in, double *part) const {
        const TAcc rk = in.rp + a_prev * s;
        const TAcc pk = -rk + beta * in.pp;
        TAcc xv = in.x;
        xv = xv + a_prev * in.pp;
        if (last) {
            xv = xv + a_cur * pk;
        } else {
            st_stream(rout + r, (TV)rk);
            st_stream(pout + r, (TV)pk);
        }
        const TXO xo = (TXO)xv;
        st_stream(xout + r, xo);
        if constexpr (DOT) part[0] += (double)xo * in.dw;
    }
    __device__ void operator()(int r, TAcc s, double *part) const { apply(r, s, load(r), part); }
};

}  // namespace uhip
