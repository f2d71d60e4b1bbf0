// Sparse mode: the precomputed interpolation matrix P (M x prod K) of
// NufftBase(mode='sparse') (mrrt/nufft/_nufft.py:751-877), used as
// x = P xk in the forward (:1383-1384) and xk = P^H y in the adjoint (:1505).
//
// The device holds P twice:
//
// * rows (forward): the value of entry j of the row at sorted position i
//   sits at [j*M + i], j in the reference's tap order (axis 0 fastest,
//   :813-823), and the row's first grid index per axis at [d*M + i].  The
//   taps of axis d are (first + j_d) mod K_d (the kd of :791), so the grid
//   index of every entry follows from the row's first taps and only the
//   values are streamed.  Rows follow the bin order of the points, so
//   neighbouring threads gather neighbouring grid points, and the
//   entry-major layout makes every value load coalesced.
// * columns (adjoint): the entries sorted by grid index (a stable sort, so
//   within a column they keep the sorted-row order), cut into pieces of at
//   most kSpPiece entries.  kSpLanes lanes sum one piece.  A column of one
//   piece is stored directly (zero for an empty column); the pieces of a
//   longer column write partial sums that sp_combine_kernel adds in order.
//   Every grid point is written exactly once and no atomics are used, so the
//   adjoint is bit-reproducible and needs no cleared grid.
#pragma once

#include "device_common.hpp"
#include "kernels_grid.hpp"

namespace mi355 {

#ifndef MI355_SP_LANES
#define MI355_SP_LANES 16
#endif
#ifndef MI355_SP_PIECE
#define MI355_SP_PIECE 64
#endif
constexpr int kSpLanes = MI355_SP_LANES;  // lanes summing one column piece (power of two, <= 64)
constexpr int kSpPiece = MI355_SP_PIECE;  // entries per column piece

struct SpPiece {
    int32_t k;      // grid point (column of P)
    int32_t start;  // first entry in the column arrays
    int32_t len;    // entries (0 for an empty column)
    int32_t slot;   // partial-sum slot, or -1 when the piece is the whole column
};

struct SpSplit {
    int32_t k;      // grid point
    int32_t first;  // first partial-sum slot
    int32_t count;  // pieces
    int32_t pad;
};

__device__ __forceinline__ double2 cmul_d(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename T>
__device__ __forceinline__ void store_val(T* v, int64_t at, double re, double) {
    v[at] = (T)re;
}
template <typename T>
__device__ __forceinline__ void store_val(cx<T>* v, int64_t at, double re, double im) {
    v[at] = mkcx<T>((T)re, (T)im);
}

// Row form.  Value = the reference's float64 / complex128 product of the
// per-axis factors ((ud0 * ud1) * ud2, :815-823), conjugated for the complex
// kernel (:830-831), times exp(i omega . n_shift) when given (:833-838), then
// rounded to the matrix dtype (:854-858).  Index = kd0 + K0 (kd1 + K1 kd2).
// Also writes the (grid index, entry id) pairs of the column sort.
//   coef[d]  (J_d, M) float64 (real kernel) or complex128 (complex kernel)
//   kidx[d]  (J_d, M) int32 grid index mod K_d
template <typename T, bool CV>
__global__ __launch_bounds__(256) void sp_fill_rows_kernel(
    const void* __restrict__ coef0, const void* __restrict__ coef1, const void* __restrict__ coef2,
    const int32_t* __restrict__ kidx0, const int32_t* __restrict__ kidx1, const int32_t* __restrict__ kidx2,
    const double2* __restrict__ phase, const int32_t* __restrict__ perm, int64_t M, int ndim, int J0, int J1,
    int J2, int K0, int K1, tab<T, CV>* __restrict__ val, int32_t* __restrict__ rbase, uint32_t* __restrict__ key,
    uint32_t* __restrict__ pay) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int64_t m = perm[i];
    const int nnzr = J0 * J1 * J2;
    rbase[i] = kidx0[m];
    if (ndim > 1) rbase[M + i] = kidx1[m];
    if (ndim > 2) rbase[2 * M + i] = kidx2[m];
    int j = 0;
    for (int j2 = 0; j2 < J2; ++j2) {
        for (int j1 = 0; j1 < J1; ++j1) {
            for (int j0 = 0; j0 < J0; ++j0, ++j) {
                int64_t kk = kidx0[(int64_t)j0 * M + m];
                if (ndim > 1) {
                    int64_t k12 = kidx1[(int64_t)j1 * M + m];
                    if (ndim > 2) k12 += (int64_t)K1 * kidx2[(int64_t)j2 * M + m];
                    kk += (int64_t)K0 * k12;
                }
                double re, im = 0.0;
                if (CV) {
                    double2 u = ((const double2*)coef0)[(int64_t)j0 * M + m];
                    if (ndim > 1) u = cmul_d(u, ((const double2*)coef1)[(int64_t)j1 * M + m]);
                    if (ndim > 2) u = cmul_d(u, ((const double2*)coef2)[(int64_t)j2 * M + m]);
                    u.y = -u.y;
                    if (phase) u = cmul_d(u, phase[m]);
                    re = u.x;
                    im = u.y;
                } else {
                    re = ((const double*)coef0)[(int64_t)j0 * M + m];
                    if (ndim > 1) re = re * ((const double*)coef1)[(int64_t)j1 * M + m];
                    if (ndim > 2) re = re * ((const double*)coef2)[(int64_t)j2 * M + m];
                }
                store_val<T>(val, (int64_t)j * M + i, re, im);
                key[i * nnzr + j] = (uint32_t)kk;
                pay[i * nnzr + j] = (uint32_t)(i * nnzr + j);
            }
        }
    }
}

static __global__ __launch_bounds__(256) void sp_iota_kernel(int32_t* __restrict__ v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (int32_t)i;
}

// Column form from the sorted (grid index, entry id) pairs: the sample index
// and the value of each entry.
template <typename T, bool CV>
__global__ __launch_bounds__(256) void sp_fill_cols_kernel(const uint32_t* __restrict__ pay, int64_t nnz, int nnzr,
                                                           int64_t M, const int32_t* __restrict__ perm,
                                                           const tab<T, CV>* __restrict__ val,
                                                           tab<T, CV>* __restrict__ tval,
                                                           int32_t* __restrict__ trow) {
    const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nnz) return;
    const int64_t e = pay[s];
    const int64_t i = e / nnzr;
    const int64_t j = e - i * nnzr;
    trow[s] = perm[i];
    tval[s] = val[j * M + i];
}

// ptr[k] = first sorted entry with grid index >= k, for k = 0..nk.
static __global__ __launch_bounds__(256) void sp_col_ptr_kernel(const uint32_t* __restrict__ keys, int64_t nnz, int64_t nk,
                                                         int32_t* __restrict__ ptr) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nk) return;
    int64_t lo = 0, hi = nnz;
    while (lo < hi) {
        const int64_t mid = (lo + hi) >> 1;
        if ((int64_t)keys[mid] < k) lo = mid + 1;
        else hi = mid;
    }
    ptr[k] = (int32_t)lo;
}

// Per column: pieces, partial-sum slots (split columns only) and split flag.
// Entry nk of each count array is 0 so that the exclusive scans end in the
// totals.
static __global__ __launch_bounds__(256) void sp_piece_count_kernel(const int32_t* __restrict__ ptr, int64_t nk,
                                                             int32_t* __restrict__ npc, int32_t* __restrict__ nsl,
                                                             int32_t* __restrict__ nsp) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k > nk) return;
    if (k == nk) {
        npc[k] = nsl[k] = nsp[k] = 0;
        return;
    }
    const int32_t len = ptr[k + 1] - ptr[k];
    const bool split = len > kSpPiece;
    const int32_t n = split ? (len + kSpPiece - 1) / kSpPiece : 1;
    npc[k] = n;
    nsl[k] = split ? n : 0;
    nsp[k] = split ? 1 : 0;
}

static __global__ __launch_bounds__(256) void sp_piece_emit_kernel(const int32_t* __restrict__ ptr, int64_t nk,
                                                            const int32_t* __restrict__ opc,
                                                            const int32_t* __restrict__ osl,
                                                            const int32_t* __restrict__ osp,
                                                            SpPiece* __restrict__ pieces,
                                                            SpSplit* __restrict__ splits) {
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nk) return;
    const int32_t start = ptr[k];
    const int32_t len = ptr[k + 1] - start;
    if (len <= kSpPiece) {
        pieces[opc[k]] = SpPiece{(int32_t)k, start, len, -1};
        return;
    }
    const int32_t n = (len + kSpPiece - 1) / kSpPiece;
    for (int32_t q = 0; q < n; ++q) {
        const int32_t l = len - q * kSpPiece < kSpPiece ? len - q * kSpPiece : kSpPiece;
        pieces[opc[k] + q] = SpPiece{(int32_t)k, start + q * kSpPiece, l, osl[k] + q};
    }
    splits[osp[k]] = SpSplit{(int32_t)k, osl[k], n, 0};
}

template <typename T>
__device__ __forceinline__ void sp_mac(cx<T>& acc, T v, cx<T> g) {
    acc.x += v * g.x;
    acc.y += v * g.y;
}
template <typename T>
__device__ __forceinline__ void sp_mac(cx<T>& acc, cx<T> v, cx<T> g) {
    acc.x += v.x * g.x - v.y * g.y;
    acc.y += v.x * g.y + v.y * g.x;
}
// acc += conj(v) * g
template <typename T>
__device__ __forceinline__ void sp_mac_conj(cx<T>& acc, T v, cx<T> g) {
    acc.x += v * g.x;
    acc.y += v * g.y;
}
template <typename T>
__device__ __forceinline__ void sp_mac_conj(cx<T>& acc, cx<T> v, cx<T> g) {
    acc.x += v.x * g.x + v.y * g.y;
    acc.y += v.x * g.y - v.y * g.x;
}

template <typename T>
__device__ __forceinline__ cx<T> sp_cmul(cx<T> a, cx<T> b) {
    return mkcx<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// (b + j) mod K for 0 <= b < K, 0 <= j < J
__device__ __forceinline__ int sp_wrap(int b, int j, int K) {
    int k = b + j;
    k = k >= K ? k - K : k;
    if (k >= K) k = k % K;  // only when K < J
    return k;
}

// Forward, y = P xk: one thread per sorted row, up to 4 repetitions per pass
// over the row.  phase (per sample, original order): phase_after, or null.
template <typename T, bool CV>
__global__ __launch_bounds__(256) void sp_interp_kernel(const tab<T, CV>* __restrict__ val,
                                                        const int32_t* __restrict__ rbase,
                                                        const int32_t* __restrict__ perm, int64_t M, int ndim,
                                                        int J0, int J1, int J2, int K0, int K1, int K2,
                                                        const cx<T>* __restrict__ grid, int64_t nk, int32_t nreps,
                                                        const cx<T>* __restrict__ phase, cx<T>* __restrict__ y) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M) return;
    const int64_t m = perm[i];
    const int b0 = rbase[i];
    const int b1 = ndim > 1 ? rbase[M + i] : 0;
    const int b2 = ndim > 2 ? rbase[2 * M + i] : 0;
    for (int32_t r0 = 0; r0 < nreps; r0 += 4) {
        const int nr = nreps - r0 < 4 ? nreps - r0 : 4;
        cx<T> acc[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) acc[q] = mkcx<T>(0, 0);
        const cx<T>* g0 = grid + (int64_t)r0 * nk;
        int j = 0;
        for (int j2 = 0; j2 < J2; ++j2) {
            const int64_t c2 = ndim > 2 ? (int64_t)sp_wrap(b2, j2, K2) * K1 : 0;
            for (int j1 = 0; j1 < J1; ++j1) {
                const int64_t c12 = ((ndim > 1 ? sp_wrap(b1, j1, K1) : 0) + c2) * K0;
                for (int j0 = 0; j0 < J0; ++j0, ++j) {
                    const tab<T, CV> v = val[(int64_t)j * M + i];
                    const int64_t c = c12 + sp_wrap(b0, j0, K0);
#pragma unroll
                    for (int q = 0; q < 4; ++q)
                        if (q < nr) sp_mac<T>(acc[q], v, g0[(int64_t)q * nk + c]);
                }
            }
        }
        const cx<T> ph = phase ? phase[m] : mkcx<T>(1, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q)
            if (q < nr) y[(int64_t)(r0 + q) * M + m] = phase ? sp_cmul<T>(acc[q], ph) : acc[q];
    }
}

template <typename T>
__device__ __forceinline__ cx<T> sp_phase_before_conj(const GridGeom& g, int64_t k, T c0, T c1, T c2) {
    int k1, k2, k3;
    split_index(k, g, k1, k2, k3);
    const T c[3] = {c0, c1, c2};
    cx<T> p = phase_before_at<T>(c, g.ndim, k1, k2, k3);
    p.y = -p.y;
    return p;
}

// Adjoint, xk = P^H y: kSpLanes lanes per column piece.  phase (per sample):
// phase_after, applied as conj(phase) * y, or null; with_pb multiplies each
// finished column by conj(phase_before) (:1519-1520).
template <typename T, bool CV>
__global__ __launch_bounds__(256) void sp_spread_kernel(const SpPiece* __restrict__ pieces, int64_t npieces,
                                                        const tab<T, CV>* __restrict__ tval,
                                                        const int32_t* __restrict__ trow,
                                                        const cx<T>* __restrict__ y, int64_t M, int32_t nreps,
                                                        const cx<T>* __restrict__ phase, GridGeom g, T pc0, T pc1,
                                                        T pc2, int with_pb, cx<T>* __restrict__ grid,
                                                        cx<T>* __restrict__ part, int64_t nslots) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t pi = t / kSpLanes;
    const int lane = (int)(t & (kSpLanes - 1));
    // blockDim is a multiple of kSpLanes: a piece's lanes leave together
    if (pi >= npieces) return;
    const SpPiece pc = pieces[pi];
    cx<T> pb = mkcx<T>(1, 0);
    if (with_pb && pc.slot < 0) pb = sp_phase_before_conj<T>(g, pc.k, pc0, pc1, pc2);
    for (int32_t r = 0; r < nreps; ++r) {
        cx<T> acc = mkcx<T>(0, 0);
        for (int32_t e = lane; e < pc.len; e += kSpLanes) {
            const int64_t s = (int64_t)pc.start + e;
            const int64_t m = trow[s];
            cx<T> v = y[(int64_t)r * M + m];
            if (phase) {
                const cx<T> ph = phase[m];
                v = mkcx<T>(v.x * ph.x + v.y * ph.y, v.y * ph.x - v.x * ph.y);
            }
            sp_mac_conj<T>(acc, tval[s], v);
        }
#pragma unroll
        for (int off = kSpLanes / 2; off > 0; off >>= 1) {
            acc.x += __shfl_xor(acc.x, off, kSpLanes);
            acc.y += __shfl_xor(acc.y, off, kSpLanes);
        }
        if (lane == 0) {
            if (pc.slot < 0) grid[(int64_t)r * g.nk + pc.k] = with_pb ? sp_cmul<T>(acc, pb) : acc;
            else part[(int64_t)r * nslots + pc.slot] = acc;
        }
    }
}

// out[r*M + m] = y[r*M + m] * conj(phase[m]) (:1497), once per sample so
// the column pieces gather one value per entry
template <typename T>
__global__ __launch_bounds__(256) void sp_conj_phase_kernel(const cx<T>* __restrict__ y,
                                                            const cx<T>* __restrict__ phase, int64_t M,
                                                            cx<T>* __restrict__ out) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = blockIdx.y;
    if (m >= M) return;
    const cx<T> v = y[r * M + m];
    const cx<T> ph = phase[m];
    out[r * M + m] = mkcx<T>(v.x * ph.x + v.y * ph.y, v.y * ph.x - v.x * ph.y);
}

// Adds the partial sums of each split column in slot order.
template <typename T>
__global__ __launch_bounds__(256) void sp_combine_kernel(const SpSplit* __restrict__ splits, int64_t nsplit,
                                                         const cx<T>* __restrict__ part, int64_t nslots,
                                                         int32_t nreps, GridGeom g, T pc0, T pc1, T pc2,
                                                         int with_pb, cx<T>* __restrict__ grid) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsplit) return;
    const SpSplit sp = splits[t];
    const cx<T> pb = with_pb ? sp_phase_before_conj<T>(g, sp.k, pc0, pc1, pc2) : mkcx<T>(1, 0);
    for (int32_t r = 0; r < nreps; ++r) {
        cx<T> acc = mkcx<T>(0, 0);
        for (int32_t q = 0; q < sp.count; ++q) {
            const cx<T> v = part[(int64_t)r * nslots + sp.first + q];
            acc.x += v.x;
            acc.y += v.y;
        }
        grid[(int64_t)r * g.nk + sp.k] = with_pb ? sp_cmul<T>(acc, pb) : acc;
    }
}

}  // namespace mi355
