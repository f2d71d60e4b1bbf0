// Multi-coil (SENSE) encoding around the NUFFT: coil images from one image
// and the coil sensitivities, the conjugate-weighted coil combination, and
// per-sample density weights.  The reference leaves these steps to
// MRI_Operator of the external mrrt.operators (recommended in
// mrrt/nufft/_nufft.py:3-5, not in the reference tree); here each step is one
// HBM-bound pass.
//
// Images are flat: pixel n of image i at [i*npix + n], in any pixel order
// shared by the images and the sensitivity maps.
#pragma once

#include "device_common.hpp"

namespace mi355 {

// out[(f*C + c)*npix + n] = smap[c*npix + n] * x[f*npix + n]
template <typename T>
__global__ __launch_bounds__(256) void coil_expand_kernel(const cx<T>* __restrict__ x,
                                                          const cx<T>* __restrict__ smap, int64_t npix,
                                                          int32_t ncoils, cx<T>* __restrict__ out) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t f = blockIdx.y;
    if (n >= npix) return;
    const cx<T> v = x[f * npix + n];
    for (int32_t c = 0; c < ncoils; ++c) {
        const cx<T> s = smap[(int64_t)c * npix + n];
        out[(f * ncoils + c) * npix + n] = mkcx<T>(s.x * v.x - s.y * v.y, s.x * v.y + s.y * v.x);
    }
}

// x[f*npix + n] = sum_c conj(smap[c*npix + n]) * img[(f*C + c)*npix + n]
template <typename T>
__global__ __launch_bounds__(256) void coil_combine_kernel(const cx<T>* __restrict__ img,
                                                           const cx<T>* __restrict__ smap, int64_t npix,
                                                           int32_t ncoils, cx<T>* __restrict__ x) {
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t f = blockIdx.y;
    if (n >= npix) return;
    T re = 0, im = 0;
    for (int32_t c = 0; c < ncoils; ++c) {
        const cx<T> s = smap[(int64_t)c * npix + n];
        const cx<T> v = img[(f * ncoils + c) * npix + n];
        re += s.x * v.x + s.y * v.y;
        im += s.x * v.y - s.y * v.x;
    }
    x[f * npix + n] = mkcx<T>(re, im);
}

// out[r*M + m] = w[m] * y[r*M + m]
template <typename T>
__global__ __launch_bounds__(256) void weight_samples_kernel(const cx<T>* __restrict__ y, const T* __restrict__ w,
                                                             int64_t M, cx<T>* __restrict__ out) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t r = blockIdx.y;
    if (m >= M) return;
    const cx<T> v = y[r * M + m];
    const T s = w[m];
    out[r * M + m] = mkcx<T>(s * v.x, s * v.y);
}

}  // namespace mi355
