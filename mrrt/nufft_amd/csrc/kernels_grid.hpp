// Elementwise kernels around the FFT: deapodization + zero-pad onto the
// oversampled grid, the phase_before multiply, crop + scale + deapodization,
// grid zeroing, and the per-sample set-up (tm and per-sample phase).
#pragma once

#include "device_common.hpp"

namespace mi355 {

struct GridGeom {
    int ndim;
    int K[3];
    int N[3];
    int64_t nk;      // prod K
    int64_t xs[4];   // image element strides: axes 0..2, repetition
};

__device__ __forceinline__ void split_index(int64_t i, const GridGeom& g, int& k1, int& k2, int& k3) {
    const uint32_t u = (uint32_t)i;  // prod K < 2^31 is checked on the host
    k1 = (int)(u % (uint32_t)g.K[0]);
    const uint32_t q = u / (uint32_t)g.K[0];
    k2 = (int)(q % (uint32_t)g.K[1]);
    k3 = (int)(q / (uint32_t)g.K[1]);
}

// sn[n] as the reference forms it: an outer product in float64 starting from
// 1.0 (_nufft.py:736-748), cast to the operator's real dtype (:668-669).
template <typename T>
__device__ __forceinline__ T deapod_factor(const double* const* sn, int ndim, int k1, int k2, int k3) {
    double s = 1.0 * sn[0][k1];
    if (ndim > 1) s = s * sn[1][k2];
    if (ndim > 2) s = s * sn[2][k3];
    return (T)s;
}

// grid[r, k] = x[r, k] * sn[k] for k < N on every axis, else 0
// (_nufft.py:1326 then the zero-padding of fftn(xk, Kd) at :1331).
template <typename T>
__global__ __launch_bounds__(256) void deapod_pad_kernel(const cx<T>* __restrict__ x, cx<T>* __restrict__ grid,
                                                         GridGeom g, const double* sn0, const double* sn1,
                                                         const double* sn2) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nk) return;
    const int r = blockIdx.y;
    int k1, k2, k3;
    split_index(i, g, k1, k2, k3);
    cx<T> o = mkcx<T>((T)0, (T)0);
    if (k1 < g.N[0] && k2 < g.N[1] && k3 < g.N[2]) {
        const double* sn[3] = {sn0, sn1, sn2};
        const T s = deapod_factor<T>(sn, g.ndim, k1, k2, k3);
        const cx<T> v = x[(int64_t)k1 * g.xs[0] + (int64_t)k2 * g.xs[1] + (int64_t)k3 * g.xs[2] +
                          (int64_t)r * g.xs[3]];
        o.x = v.x * s;
        o.y = v.y * s;
    }
    grid[(int64_t)r * g.nk + i] = o;
}

// phase_before (_nufft.py:703-715) evaluated in T exactly as the reference:
// phi = ((c0*k0) + (c1*k1)) + (c2*k2) with c_d = (T)((2 pi / K_d) * n_mid_d),
// every product and sum rounded to T; then exp(i phi).
template <typename T>
__device__ __forceinline__ T phase_before_arg(const T* c, int ndim, int k1, int k2, int k3) {
    T phi = c[0] * (T)k1;
    if (ndim > 1) phi = phi + c[1] * (T)k2;
    if (ndim > 2) phi = phi + c[2] * (T)k3;
    return phi;
}

template <typename T>
__device__ __forceinline__ cx<T> phase_before_at(const T* c, int ndim, int k1, int k2, int k3) {
    return cis<T>(phase_before_arg<T>(c, ndim, k1, k2, k3));
}

// In-place grid *= phase_before (forward, then / scale_ortho: _nufft.py:1371,
// :1374-1375) or grid *= conj(phase_before) (adjoint, :1519-1520).
template <typename T>
__global__ __launch_bounds__(256) void phase_grid_kernel(cx<T>* __restrict__ grid, GridGeom g, int nreps,
                                                         T c0, T c1, T c2, int conj_phase, int ortho,
                                                         T scale_ortho) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nk) return;
    int k1, k2, k3;
    split_index(i, g, k1, k2, k3);
    const T c[3] = {c0, c1, c2};
    cx<T> p = phase_before_at<T>(c, g.ndim, k1, k2, k3);
    if (conj_phase) p.y = -p.y;
    for (int r = 0; r < nreps; ++r) {
        cx<T> v = grid[(int64_t)r * g.nk + i];
        cx<T> o;
        o.x = v.x * p.x - v.y * p.y;
        o.y = v.x * p.y + v.y * p.x;
        if (ortho) {
            o.x = o.x / scale_ortho;
            o.y = o.y / scale_ortho;
        }
        grid[(int64_t)r * g.nk + i] = o;
    }
}

// x[r, n] = grid[r, n] * scale * conj(sn[n]) for n in Nd (_nufft.py:1561-1572;
// sn is real).  scale folds the reference's ifftn 1/prod(Kd) with its
// prod(Kd) * adjoint_scalefactor (or scale_ortho * adjoint_scalefactor).
template <typename T>
__global__ __launch_bounds__(256) void crop_deapod_kernel(const cx<T>* __restrict__ grid, cx<T>* __restrict__ x,
                                                          GridGeom g, int64_t nn, const double* sn0,
                                                          const double* sn1, const double* sn2, T scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nn) return;
    const int r = blockIdx.y;
    const uint32_t u = (uint32_t)i;
    const int n1 = (int)(u % (uint32_t)g.N[0]);
    const uint32_t q = u / (uint32_t)g.N[0];
    const int n2 = (int)(q % (uint32_t)g.N[1]);
    const int n3 = (int)(q / (uint32_t)g.N[1]);
    const double* sn[3] = {sn0, sn1, sn2};
    const T s = deapod_factor<T>(sn, g.ndim, n1, n2, n3);
    const cx<T> v = grid[(int64_t)r * g.nk + ((int64_t)n3 * g.K[1] + n2) * g.K[0] + n1];
    cx<T> o;
    o.x = (v.x * scale) * s;
    o.y = (v.y * scale) * s;
    x[(int64_t)n1 * g.xs[0] + (int64_t)n2 * g.xs[1] + (int64_t)n3 * g.xs[2] + (int64_t)r * g.xs[3]] = o;
}

// The crop of crop_deapod_kernel over the whole grid, writing zeros over
// every grid point after reading it (the grid is left all zero).
template <typename T>
__global__ __launch_bounds__(256) void crop_deapod_rezero_kernel(cx<T>* __restrict__ grid, cx<T>* __restrict__ x,
                                                                 GridGeom g, const double* sn0, const double* sn1,
                                                                 const double* sn2, T scale) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= g.nk) return;
    const int r = blockIdx.y;
    cx<T>* gp = grid + (int64_t)r * g.nk + i;
    int k1, k2, k3;
    split_index(i, g, k1, k2, k3);
    if (k1 < g.N[0] && k2 < g.N[1] && k3 < g.N[2]) {
        const double* sn[3] = {sn0, sn1, sn2};
        const T s = deapod_factor<T>(sn, g.ndim, k1, k2, k3);
        const cx<T> v = *gp;
        cx<T> o;
        o.x = (v.x * scale) * s;
        o.y = (v.y * scale) * s;
        x[(int64_t)k1 * g.xs[0] + (int64_t)k2 * g.xs[1] + (int64_t)k3 * g.xs[2] + (int64_t)r * g.xs[3]] = o;
    }
    *gp = mkcx<T>((T)0, (T)0);
}

template <typename T>
__global__ __launch_bounds__(256) void zero_kernel(cx<T>* __restrict__ p, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = mkcx<T>((T)0, (T)0);
}

// Per-sample set-up.
//   tm[d*M + m] = (T)omega / (T)(2 pi / K_d)  in T (_nufft.py:338-342 after
//   the cast of omega at :315).  For float the quotient is formed in double
//   and rounded once, which equals the correctly rounded float quotient.
//   phase_mode 1: phase_after = exp(i omega . shift) (_nufft.py:717-724) from
//   the uncast omega, dot in float32 (dot_single) or float64;
//   phase_mode 2: phase_shift = exp(i omega_T . n_shift) in float64
//   (_nufft.py:898-901).
template <typename T, typename W>
__global__ __launch_bounds__(256) void set_points_kernel(const W* __restrict__ omega, int64_t M, int64_t sm,
                                                         int64_t sd, int ndim, double g0, double g1, double g2,
                                                         T* __restrict__ tm, cx<T>* __restrict__ phase,
                                                         int phase_mode, double s0, double s1, double s2,
                                                         int dot_single) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M) return;
    const double gam[3] = {g0, g1, g2};
    const double sh[3] = {s0, s1, s2};
    W w[3] = {0, 0, 0};
    T wr[3] = {0, 0, 0};
    for (int d = 0; d < ndim; ++d) {
        w[d] = omega[m * sm + (int64_t)d * sd];
        wr[d] = (T)w[d];
        const T gt = (T)gam[d];
        tm[(int64_t)d * M + m] = (T)((double)wr[d] / (double)gt);
    }
    if (phase_mode == 1) {
        double acc;
        if (dot_single) {
            float a = (float)w[0] * (float)sh[0];
            for (int d = 1; d < ndim; ++d) a = a + (float)w[d] * (float)sh[d];
            float sf, cf;
            double sd_, cd_;
            sincos((double)a, &sd_, &cd_);
            sf = (float)sd_;
            cf = (float)cd_;
            phase[m] = mkcx<T>((T)cf, (T)sf);
            return;
        }
        acc = (double)w[0] * sh[0];
        for (int d = 1; d < ndim; ++d) acc = acc + (double)w[d] * sh[d];
        double s, c;
        sincos(acc, &s, &c);
        phase[m] = mkcx<T>((T)c, (T)s);
    } else if (phase_mode == 2) {
        double acc = (double)wr[0] * sh[0];
        for (int d = 1; d < ndim; ++d) acc = acc + (double)wr[d] * sh[d];
        double s, c;
        sincos(acc, &s, &c);
        phase[m] = mkcx<T>((T)c, (T)s);
    }
}

}  // namespace mi355
