// Table1 interpolation kernels: forward gather (grid -> samples) and adjoint
// scatter (samples -> grid), 1-D/2-D/3-D, real or complex kernel, float or
// double.  One thread per sample; the J^d coefficients are computed once per
// sample (separably, J per axis) and reused for every repetition (coil) of a
// batch, so a batch of C coils costs one coefficient evaluation per sample.
//
// The tap nesting and the order of every sum follow the reference C kernels
// (mrrt/nufft/_extensions/c/nufft_table.template.c:538-616 for the 2-D real
// forward, :623-703 complex forward, :308-380 / :417-495 adjoints, :710-921
// and :924-1165 in 3-D, :43-305 in 1-D).
#pragma once

#include "device_common.hpp"

namespace mi355 {

template <typename T, bool CK>
struct InterpArgs {
    const cx<T>* grid;      // (nreps, prod K) forward input / adjoint output
    cx<T>* out;             // forward: (nreps, M) samples
    const cx<T>* in;        // adjoint: (nreps, M) samples
    cx<T>* grid_out;        // adjoint grid
    const tab<T, CK>* h[3]; // centred tables
    const T* tm;            // (ndim, M)
    const cx<T>* phase;     // per-sample phase or null
    const int* perm;        // processing order or null
    int64_t M;
    int64_t kstride;        // prod K
    AxisGeom ax[3];
    int L;
    int nreps;
    int conj_phase;         // adjoint: multiply by conj(phase)
};

template <typename T, int NDIM, bool CK, int JMAX>
__device__ __forceinline__ void load_axes(const InterpArgs<T, CK>& a, int64_t m,
                                          T (&cr)[3][JMAX], T (&ci)[3][JMAX],
                                          int (&idx)[3][JMAX]) {
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        const T t = a.tm[(int64_t)d * a.M + m];
        int k0;
        if constexpr (CK) {
            k0 = axis_coefs<T, JMAX>(t, a.L, a.ax[d], a.h[d], cr[d], ci[d]);
        } else {
            k0 = axis_coefs<T, JMAX>(t, a.L, a.ax[d], a.h[d], cr[d]);
        }
        axis_index<JMAX>(k0, a.ax[d].J, a.ax[d].K, idx[d]);
    }
}

// acc += c * v  (complex kernel: full complex product, real kernel: scaling)
template <typename T, bool CK>
__device__ __forceinline__ void madd(T& ar, T& ai, T cr, T ci, T vr, T vi) {
    if constexpr (CK) {
        ar += cr * vr - ci * vi;
        ai += cr * vi + ci * vr;
    } else {
        ar += cr * vr;
        ai += cr * vi;
    }
}

// acc += c * v with fused multiply-adds: one rounding per step instead of
// two, and half the instructions of the real-kernel sum.  The plan's tiled
// gather uses it; madd keeps the reference's roundings (the raw table1
// entry points).
template <typename T, bool CK>
__device__ __forceinline__ void madd_fused(T& ar, T& ai, T cr, T ci, T vr, T vi) {
    if constexpr (CK) {
        ar = fma(cr, vr, fma(-ci, vi, ar));
        ai = fma(cr, vi, fma(ci, vr, ai));
    } else {
        ar = fma(cr, vr, ar);
        ai = fma(cr, vi, ai);
    }
}

template <typename T, bool CK, bool FUSED>
__device__ __forceinline__ void madd_sel(T& ar, T& ai, T cr, T ci, T vr, T vi) {
    if constexpr (FUSED) madd_fused<T, CK>(ar, ai, cr, ci, vr, vi);
    else madd<T, CK>(ar, ai, cr, ci, vr, vi);
}

template <typename T, int NDIM, bool CK, int JMAX>
__global__ __launch_bounds__(256) void interp_fwd_kernel(InterpArgs<T, CK> a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.M) return;
    const int64_t m = a.perm ? (int64_t)a.perm[i] : i;
    T cr[3][JMAX], ci[3][JMAX];
    int idx[3][JMAX];
    load_axes<T, NDIM, CK, JMAX>(a, m, cr, ci, idx);
    const int K1 = a.ax[0].K;
    const int K2 = NDIM > 1 ? a.ax[1].K : 1;
    cx<T> ph = mkcx<T>((T)1, (T)0);
    if (a.phase) ph = a.phase[m];
    for (int r = 0; r < a.nreps; ++r) {
        const cx<T>* __restrict__ g = a.grid + (int64_t)r * a.kstride;
        T sr = 0, si = 0;
        if constexpr (NDIM == 1) {
#pragma unroll
            for (int j1 = 0; j1 < JMAX; ++j1) {
                if (j1 < a.ax[0].J) {
                    const cx<T> v = g[idx[0][j1]];
                    madd<T, CK>(sr, si, cr[0][j1], ci[0][j1], v.x, v.y);
                }
            }
        } else if constexpr (NDIM == 2) {
#pragma unroll
            for (int j2 = 0; j2 < JMAX; ++j2) {
                if (j2 < a.ax[1].J) {
                    const cx<T>* __restrict__ row = g + (int64_t)idx[1][j2] * K1;
                    T s1r = 0, s1i = 0;
#pragma unroll
                    for (int j1 = 0; j1 < JMAX; ++j1) {
                        if (j1 < a.ax[0].J) {
                            const cx<T> v = row[idx[0][j1]];
                            madd<T, CK>(s1r, s1i, cr[0][j1], ci[0][j1], v.x, v.y);
                        }
                    }
                    madd<T, CK>(sr, si, cr[1][j2], ci[1][j2], s1r, s1i);
                }
            }
        } else {
            for (int j3 = 0; j3 < JMAX; ++j3) {
                if (j3 < a.ax[2].J) {
                    T s2r = 0, s2i = 0;
#pragma unroll
                    for (int j2 = 0; j2 < JMAX; ++j2) {
                        if (j2 < a.ax[1].J) {
                            const cx<T>* __restrict__ row =
                                g + ((int64_t)idx[2][j3] * K2 + idx[1][j2]) * K1;
                            T s1r = 0, s1i = 0;
#pragma unroll
                            for (int j1 = 0; j1 < JMAX; ++j1) {
                                if (j1 < a.ax[0].J) {
                                    const cx<T> v = row[idx[0][j1]];
                                    madd<T, CK>(s1r, s1i, cr[0][j1], ci[0][j1], v.x, v.y);
                                }
                            }
                            madd<T, CK>(s2r, s2i, cr[1][j2], ci[1][j2], s1r, s1i);
                        }
                    }
                    madd<T, CK>(sr, si, cr[2][j3], ci[2][j3], s2r, s2i);
                }
            }
        }
        cx<T> o;
        if (a.phase) {
            o.x = sr * ph.x - si * ph.y;
            o.y = sr * ph.y + si * ph.x;
        } else {
            o.x = sr;
            o.y = si;
        }
        a.out[(int64_t)r * a.M + m] = o;
    }
}

// v = conj(c) * f (complex kernel) or c * f (real kernel)
template <typename T, bool CK>
__device__ __forceinline__ void cmul_conj(T& vr, T& vi, T cr, T ci, T fr, T fi) {
    if constexpr (CK) {
        vr = cr * fr + ci * fi;
        vi = cr * fi - ci * fr;
    } else {
        vr = cr * fr;
        vi = cr * fi;
    }
}

template <typename T>
__device__ __forceinline__ void atomic_add_cx(cx<T>* p, T vr, T vi) {
    T* q = reinterpret_cast<T*>(p);
    unsafeAtomicAdd(q, vr);
    unsafeAtomicAdd(q + 1, vi);
}

template <typename T, int NDIM, bool CK, int JMAX>
__global__ __launch_bounds__(256) void spread_gm_kernel(InterpArgs<T, CK> a) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.M) return;
    const int64_t m = a.perm ? (int64_t)a.perm[i] : i;
    T cr[3][JMAX], ci[3][JMAX];
    int idx[3][JMAX];
    load_axes<T, NDIM, CK, JMAX>(a, m, cr, ci, idx);
    const int K1 = a.ax[0].K;
    const int K2 = NDIM > 1 ? a.ax[1].K : 1;
    cx<T> ph = mkcx<T>((T)1, (T)0);
    if (a.phase) ph = a.phase[m];
    for (int r = 0; r < a.nreps; ++r) {
        cx<T> f = a.in[(int64_t)r * a.M + m];
        if (a.phase) {  // f * conj(phase)
            const T fr = f.x * ph.x + f.y * ph.y;
            const T fi = f.y * ph.x - f.x * ph.y;
            f.x = fr;
            f.y = fi;
        }
        cx<T>* __restrict__ g = a.grid_out + (int64_t)r * a.kstride;
        if constexpr (NDIM == 1) {
#pragma unroll
            for (int j1 = 0; j1 < JMAX; ++j1) {
                if (j1 < a.ax[0].J) {
                    T vr, vi;
                    cmul_conj<T, CK>(vr, vi, cr[0][j1], ci[0][j1], f.x, f.y);
                    atomic_add_cx<T>(g + idx[0][j1], vr, vi);
                }
            }
        } else if constexpr (NDIM == 2) {
#pragma unroll
            for (int j2 = 0; j2 < JMAX; ++j2) {
                if (j2 < a.ax[1].J) {
                    T v2r, v2i;
                    cmul_conj<T, CK>(v2r, v2i, cr[1][j2], ci[1][j2], f.x, f.y);
                    cx<T>* __restrict__ row = g + (int64_t)idx[1][j2] * K1;
#pragma unroll
                    for (int j1 = 0; j1 < JMAX; ++j1) {
                        if (j1 < a.ax[0].J) {
                            T vr, vi;
                            cmul_conj<T, CK>(vr, vi, cr[0][j1], ci[0][j1], v2r, v2i);
                            atomic_add_cx<T>(row + idx[0][j1], vr, vi);
                        }
                    }
                }
            }
        } else {
            for (int j3 = 0; j3 < JMAX; ++j3) {
                if (j3 < a.ax[2].J) {
                    T v3r, v3i;
                    cmul_conj<T, CK>(v3r, v3i, cr[2][j3], ci[2][j3], f.x, f.y);
#pragma unroll
                    for (int j2 = 0; j2 < JMAX; ++j2) {
                        if (j2 < a.ax[1].J) {
                            T v2r, v2i;
                            cmul_conj<T, CK>(v2r, v2i, cr[1][j2], ci[1][j2], v3r, v3i);
                            cx<T>* __restrict__ row =
                                g + ((int64_t)idx[2][j3] * K2 + idx[1][j2]) * K1;
#pragma unroll
                            for (int j1 = 0; j1 < JMAX; ++j1) {
                                if (j1 < a.ax[0].J) {
                                    T vr, vi;
                                    cmul_conj<T, CK>(vr, vi, cr[0][j1], ci[0][j1], v2r, v2i);
                                    atomic_add_cx<T>(row + idx[0][j1], vr, vi);
                                }
                            }
                        }
                    }
                }
            }
        }
    }
}

}  // namespace mi355
