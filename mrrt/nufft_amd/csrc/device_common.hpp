// Device-side helpers shared by the table1 NUFFT kernels (gfx950).
//
// Arithmetic that decides which table cell and which grid point a tap uses
// follows the reference C kernels operation for operation
// (mrrt/nufft/_extensions/c/nufft_table.template.c:577-601):
//   k0  = 1 + floor(t - J/2.)          evaluated in double
//   p   = (t - k) * L                  evaluated in T (float or double)
//   n   = floor(p),  alf = p - n       T
//   c   = (1 - alf) * h[n + nc] + alf * h[n + 1 + nc],   nc = floor(J*L/2)
// FP contraction is disabled in this file so that every operation above is
// rounded exactly as in the reference (no fused multiply-add).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#pragma clang fp contract(off)

namespace mi355 {

template <typename T> struct CxOf;
template <> struct CxOf<float> { using type = float2; };
template <> struct CxOf<double> { using type = double2; };
template <typename T> using cx = typename CxOf<T>::type;

// table element: real T for the real kernel, complex T for the complex kernel
template <typename T, bool CK> struct TabOf { using type = T; };
template <typename T> struct TabOf<T, true> { using type = cx<T>; };
template <typename T, bool CK> using tab = typename TabOf<T, CK>::type;

template <typename T>
__device__ __forceinline__ cx<T> mkcx(T re, T im) {
    cx<T> c;
    c.x = re;
    c.y = im;
    return c;
}

__device__ __forceinline__ float floor_t(float v) { return floorf(v); }
__device__ __forceinline__ double floor_t(double v) { return floor(v); }
__device__ __forceinline__ float fmin_t(float a, float b) { return fminf(a, b); }
__device__ __forceinline__ double fmin_t(double a, double b) { return fmin(a, b); }
__device__ __forceinline__ float fmax_t(float a, float b) { return fmaxf(a, b); }
__device__ __forceinline__ double fmax_t(double a, double b) { return fmax(a, b); }

// Per-axis table geometry.  The table pointer is shifted to its centre
// (nufft_table.template.c:58-63) and the cell index n is kept in
// [nlo, nhi] = [-nc, J*L - 1 - nc] so that h[n] and h[n+1] stay inside the
// J*L+1 entries even for the rounding edge case p == J*L/2 (where the
// clamped cell gives the same coefficient) and for non-finite t.
struct AxisGeom {
    int K;      // oversampled grid size
    int J;      // taps
    int nlo;    // -nc
    int nhi;    // J*L - 1 - nc
};

// First grid index (mod K) of the J taps of coordinate t.
template <typename T>
__device__ __forceinline__ long long first_tap(T t, int J) {
    double kd = floor((double)t - 0.5 * (double)J);
    kd = fmin(fmax(kd, -1.0e9), 1.0e9);  // NaN -> -1e9, keeps the cast defined
    // |k + j| < 2^31 for every tap j, so callers convert tap positions
    // through int: the same value as from 64 bits, in one instruction
    // instead of a software int64 -> float conversion
    return 1 + (long long)kd;
}

// Workgroups are dispatched to the 8 XCDs of an MI355X round-robin by
// block index, and each XCD has its own L2.  xcd_remap turns block b of n
// into a logical index such that each XCD gets one contiguous range of
// logical indices (x*q + min(x, r) + b/8 for x = b % 8, n = 8q + r), so
// neighbouring work items, which share cache lines, share an L2.
// It is a bijection of [0, n).
__device__ __forceinline__ unsigned xcd_remap(unsigned b, unsigned n) {
    constexpr unsigned kXcds = 8;
    const unsigned q = n / kXcds, r = n % kXcds, x = b % kXcds, j = b / kXcds;
    return x * q + (x < r ? x : r) + j;
}

// k mod K in [0, K).  Points with omega in [-pi, pi) give k in [-K, K),
// handled without the (software) 64-bit division.
__device__ __forceinline__ int pos_mod(long long k, int K) {
    if (k >= -(long long)K && k < (long long)K) return (int)(k < 0 ? k + K : k);
    long long r = k % (long long)K;
    return (int)(r < 0 ? r + K : r);
}

// Coefficients of the J taps along one axis, real kernel.  Returns the grid
// index of the first tap modulo K.
template <typename T, int JMAX, bool FIXED = false>
__device__ __forceinline__ int axis_coefs(T t, int L, const AxisGeom& g,
                                          const T* __restrict__ hc,
                                          T (&c)[JMAX]) {
    const int J = FIXED ? JMAX : g.J;
    const long long k = first_tap(t, J);
    const T tL = (T)L;
#pragma unroll
    for (int j = 0; j < JMAX; ++j) {
        if (FIXED || j < g.J) {
            const T p = (t - (T)(int)(k + j)) * tL;
            const T pc = fmin_t(fmax_t(p, (T)g.nlo), (T)g.nhi);
            const int n = (int)floor_t(pc);
            const T alf = p - (T)n;
            c[j] = ((T)1 - alf) * hc[n] + alf * hc[n + 1];
        } else {
            c[j] = (T)0;
        }
    }
    return pos_mod(k, g.K);
}

// Complex-kernel variant (phasing='complex'): coefficient real and imaginary
// parts are interpolated separately (nufft_table.template.c:670-672).
template <typename T, int JMAX, bool FIXED = false>
__device__ __forceinline__ int axis_coefs(T t, int L, const AxisGeom& g,
                                          const cx<T>* __restrict__ hc,
                                          T (&cr)[JMAX], T (&ci)[JMAX]) {
    const int J = FIXED ? JMAX : g.J;
    const long long k = first_tap(t, J);
    const T tL = (T)L;
#pragma unroll
    for (int j = 0; j < JMAX; ++j) {
        if (FIXED || j < g.J) {
            const T p = (t - (T)(int)(k + j)) * tL;
            const T pc = fmin_t(fmax_t(p, (T)g.nlo), (T)g.nhi);
            const int n = (int)floor_t(pc);
            const T alf = p - (T)n;
            const cx<T> h0 = hc[n];
            const cx<T> h1 = hc[n + 1];
            cr[j] = ((T)1 - alf) * h0.x + alf * h1.x;
            ci[j] = ((T)1 - alf) * h0.y + alf * h1.y;
        } else {
            cr[j] = (T)0;
            ci[j] = (T)0;
        }
    }
    return pos_mod(k, g.K);
}

// Coefficients and their t-derivatives.  The derivative is that of the
// piecewise-linear table interpolant itself, L * (h[n+1] - h[n]) on the
// cell used by the coefficient, so it is exact for the operator as
// evaluated (finite differences agree away from cell boundaries).
template <typename T, int JMAX, bool FIXED = false>
__device__ __forceinline__ int axis_coefs_d(T t, int L, const AxisGeom& g, const T* __restrict__ hc,
                                            T (&c)[JMAX], T (&dc)[JMAX]) {
    const int J = FIXED ? JMAX : g.J;
    const long long k = first_tap(t, J);
    const T tL = (T)L;
#pragma unroll
    for (int j = 0; j < JMAX; ++j) {
        if (FIXED || j < g.J) {
            const T p = (t - (T)(int)(k + j)) * tL;
            const T pc = fmin_t(fmax_t(p, (T)g.nlo), (T)g.nhi);
            const int n = (int)floor_t(pc);
            const T alf = p - (T)n;
            const T h0 = hc[n], h1 = hc[n + 1];
            c[j] = ((T)1 - alf) * h0 + alf * h1;
            dc[j] = tL * (h1 - h0);
        } else {
            c[j] = (T)0;
            dc[j] = (T)0;
        }
    }
    return pos_mod(k, g.K);
}

template <typename T, int JMAX, bool FIXED = false>
__device__ __forceinline__ int axis_coefs_d(T t, int L, const AxisGeom& g, const cx<T>* __restrict__ hc,
                                            T (&cr)[JMAX], T (&ci)[JMAX], T (&dcr)[JMAX], T (&dci)[JMAX]) {
    const int J = FIXED ? JMAX : g.J;
    const long long k = first_tap(t, J);
    const T tL = (T)L;
#pragma unroll
    for (int j = 0; j < JMAX; ++j) {
        if (FIXED || j < g.J) {
            const T p = (t - (T)(int)(k + j)) * tL;
            const T pc = fmin_t(fmax_t(p, (T)g.nlo), (T)g.nhi);
            const int n = (int)floor_t(pc);
            const T alf = p - (T)n;
            const cx<T> h0 = hc[n];
            const cx<T> h1 = hc[n + 1];
            cr[j] = ((T)1 - alf) * h0.x + alf * h1.x;
            ci[j] = ((T)1 - alf) * h0.y + alf * h1.y;
            dcr[j] = tL * (h1.x - h0.x);
            dci[j] = tL * (h1.y - h0.y);
        } else {
            cr[j] = ci[j] = dcr[j] = dci[j] = (T)0;
        }
    }
    return pos_mod(k, g.K);
}

// FIXED: every axis has exactly JMAX taps (compile-time trip counts, no
// per-tap guards); otherwise J <= JMAX per axis at run time.

// Grid indices of the J taps along one axis (periodic wrap).
template <int JMAX>
__device__ __forceinline__ void axis_index(int k0, int J, int K, int (&idx)[JMAX]) {
    int k = k0;
#pragma unroll
    for (int j = 0; j < JMAX; ++j) {
        idx[j] = k;
        if (j < J) {
            ++k;
            if (k == K) k = 0;
        }
    }
}

// (cos phi, sin phi) of a real phase, evaluated in double and rounded to T.
template <typename T>
__device__ __forceinline__ cx<T> cis(T phi) {
    double s, c;
    sincos((double)phi, &s, &c);
    return mkcx<T>((T)c, (T)s);
}

// Faster variant for the per-grid-point phases of the tiled kernels: the
// float path uses the device sincosf (within 2 ulp after full range
// reduction, like numpy's complex64 exp of the same float32 argument);
// the double path is the double sincos.
__device__ __forceinline__ float2 cis_tile(float phi) {
    float s, c;
    sincosf(phi, &s, &c);
    return make_float2(c, s);
}
__device__ __forceinline__ double2 cis_tile(double phi) {
    double s, c;
    sincos(phi, &s, &c);
    return make_double2(c, s);
}

}  // namespace mi355
