// Power-of-two FFTs of the oversampled grid, pruned to the zero-padded image
// and fused with the deapodization.
//
// The forward transform of the table1 NUFFT is fftn(x * sn, Kd): only the
// first N_d entries of every axis are non-zero before the FFT.  The adjoint
// keeps only the first N_d entries of every axis after the inverse FFT.
// Transforming axis 0 first (forward) and last (adjoint), every pass works on
// the lines whose higher axes index the image (n_e < N_e) and whose lower
// axes are already (forward) or still (adjoint) full length:
//   forward axis d: lines (k_0..k_{d-1} < K, n_{d+1}.. < N), read t < N_d
//                   (the rest is zero), write t < K_d;
//   adjoint axis d: the same lines, read t < K_d, write t < N_d.
// Axis 0 reads the image (times sn) in the forward and writes the image
// (times scale * sn) in the adjoint, so deapod_pad and crop_deapod vanish
// into the FFT passes, and rows and planes outside the image are never
// read or written.
//
// A workgroup transforms lpw lines in LDS, TL threads per line, with a
// Stockham auto-sort FFT (radix-8 passes, a final radix-2 or radix-4 pass);
// twiddles come from a table of exp(-2 pi i m / K) rounded once from double,
// staged in LDS as one table per pass, ordered by butterfly.
#pragma once

#include "device_common.hpp"
#include "kernels_grid.hpp"

namespace mi355 {

constexpr int kFftMinK = 16;
constexpr int kFftMaxK = 2048;
constexpr int kFftLds = 65536;
constexpr int kFftMaxThreads = 1024;

// TL threads transform one line: one radix-8 butterfly each per pass up to
// K = 512, more per thread beyond.  The lines per workgroup (LPW) are chosen
// at launch; the LDS holds LPW padded lines and the twiddle table.
template <typename T, int K>
struct FftCfg {
    static constexpr int CXS = 2 * (int)sizeof(T);
    // a line holds K points at pad(t) = t + t / 32 (strided butterfly writes
    // spread over the banks), plus one so neighbouring lines start on
    // different banks
    static constexpr int KP = K + K / 32 + 1;
    static constexpr int TL = K / 8 < 64 ? K / 8 : 64;
    // each thread loads NL = K / TL points of a line, LF of them in flight
    // at once: the loads of a batch are all issued before any is used
    static constexpr int NL = K / TL;
    static constexpr int LF = NL < 8 ? NL : 8;
    static constexpr size_t lds_bytes(int lpw) { return (size_t)(lpw * KP + K) * CXS; }
};

__device__ __forceinline__ int pad(int t) { return t + (t >> 5); }

// Small DFTs: X[k] = sum_n x[n] exp(SIGN * 2 pi i n k / R)
template <typename T, int SIGN>
__device__ __forceinline__ cx<T> mul_i(cx<T> a) {  // a * (SIGN i)
    return SIGN < 0 ? mkcx<T>(a.y, -a.x) : mkcx<T>(-a.y, a.x);
}

template <typename T>
__device__ __forceinline__ cx<T> cadd(cx<T> a, cx<T> b) { return mkcx<T>(a.x + b.x, a.y + b.y); }
template <typename T>
__device__ __forceinline__ cx<T> csub(cx<T> a, cx<T> b) { return mkcx<T>(a.x - b.x, a.y - b.y); }
template <typename T>
__device__ __forceinline__ cx<T> cmulc(cx<T> a, cx<T> w) {
    return mkcx<T>(a.x * w.x - a.y * w.y, a.x * w.y + a.y * w.x);
}

template <typename T, int SIGN>
__device__ __forceinline__ void dft2(cx<T>* v) {
    const cx<T> a = v[0], b = v[1];
    v[0] = cadd<T>(a, b);
    v[1] = csub<T>(a, b);
}

template <typename T, int SIGN>
__device__ __forceinline__ void dft4(cx<T>* v) {
    const cx<T> c0 = cadd<T>(v[0], v[2]), c1 = csub<T>(v[0], v[2]);
    const cx<T> d0 = cadd<T>(v[1], v[3]), d1 = mul_i<T, SIGN>(csub<T>(v[1], v[3]));
    v[0] = cadd<T>(c0, d0);
    v[1] = cadd<T>(c1, d1);
    v[2] = csub<T>(c0, d0);
    v[3] = csub<T>(c1, d1);
}

// radix 8 as two radix-4 DFTs of the even and odd halves after one radix-2
// step with the twiddles exp(SIGN 2 pi i n / 8), n = 0..3
template <typename T, int SIGN>
__device__ __forceinline__ void dft8(cx<T>* v) {
    const T h = (T)0.70710678118654752440;
    cx<T> a[4], b[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
        a[n] = cadd<T>(v[n], v[n + 4]);
        b[n] = csub<T>(v[n], v[n + 4]);
    }
    // b1 *= (h, SIGN h); b2 *= SIGN i; b3 *= (-h, SIGN h)
    b[1] = mkcx<T>(h * (b[1].x - (T)SIGN * b[1].y), h * (b[1].y + (T)SIGN * b[1].x));
    b[2] = mul_i<T, SIGN>(b[2]);
    b[3] = mkcx<T>(-h * (b[3].x + (T)SIGN * b[3].y), h * ((T)SIGN * b[3].x - b[3].y));
    dft4<T, SIGN>(a);
    dft4<T, SIGN>(b);
#pragma unroll
    for (int k = 0; k < 4; ++k) {
        v[2 * k] = a[k];
        v[2 * k + 1] = b[k];
    }
}

template <typename T, int R, int SIGN>
__device__ __forceinline__ void dft_r(cx<T>* v) {
    if constexpr (R == 8) dft8<T, SIGN>(v);
    else if constexpr (R == 4) dft4<T, SIGN>(v);
    else dft2<T, SIGN>(v);
}

// Radix of the pass of sub-transform size ns (8, the last one 2 or 4) and
// where its twiddle table starts (the passes with ns > 1, in order).
template <int K>
__host__ __device__ constexpr int tw_radix(int ns) { return K / ns >= 8 ? 8 : K / ns; }
template <int K>
__host__ __device__ constexpr int tw_offset(int ns) {
    int off = 0;
    for (int s = 8; s < ns; s *= tw_radix<K>(s)) off += (tw_radix<K>(s) - 1) * s;
    return off;
}

// One Stockham pass of radix R over a line of K points in LDS, sub-transform
// size NS before the pass: butterfly j reads j + r K/R, multiplies by
// w^(k r K/(NS R)) with k = j mod NS, and writes (j / NS) NS R + k + r NS.
// Every thread of the workgroup calls it (it holds the barriers).
template <typename T, int K, int TL, int R, int NS, int SIGN>
__device__ __forceinline__ void stockham_pass(cx<T>* line, int tl, const cx<T>* __restrict__ tw) {
    constexpr int NBF = K / R;                      // butterflies per line
    constexpr int NB = (NBF + TL - 1) / TL;         // per thread
    cx<T> v[NB][R];
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int j = tl + q * TL;
        if (j < NBF) {
#pragma unroll
            for (int r = 0; r < R; ++r) v[q][r] = line[pad(j + r * NBF)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < NB; ++q) {
        const int j = tl + q * TL;
        if (j < NBF) {
            const int k = j % NS;
            if constexpr (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    cx<T> w = tw[tw_offset<K>(NS) + (r - 1) * NS + k];
                    if (SIGN > 0) w.y = -w.y;
                    v[q][r] = cmulc<T>(v[q][r], w);
                }
            }
            dft_r<T, R, SIGN>(v[q]);
            const int o = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) line[pad(o + r * NS)] = v[q][r];
        }
    }
    __syncthreads();
}

template <typename T, int K, int TL, int NS, int SIGN>
__device__ __forceinline__ void stockham_from(cx<T>* line, int tl, const cx<T>* __restrict__ tw) {
    if constexpr (NS < K) {
        constexpr int rest = K / NS;
        constexpr int R = rest >= 8 ? 8 : rest;  // radix 8, the last pass 2 or 4
        stockham_pass<T, K, TL, R, NS, SIGN>(line, tl, tw);
        stockham_from<T, K, TL, NS * R, SIGN>(line, tl, tw);
    }
}

// in-place FFT of one LDS line by TL threads (thread tl of the line)
template <typename T, int K, int TL, int SIGN>
__device__ __forceinline__ void fft_line(cx<T>* line, int tl, const cx<T>* __restrict__ tw) {
    stockham_from<T, K, TL, 1, SIGN>(line, tl, tw);
}

struct FftRowGeom {
    int N[3];          // image size
    int K[3];          // grid size
    int64_t nk;        // prod K
    int64_t xs[4];     // image strides: axes 0..2, repetition
    int64_t nlines;    // N1 * N2 * nimg
    int lpw;           // lines per workgroup (blockDim = lpw * TL)
    const void* mult;  // adjoint: complex T grid (one image) multiplying the loaded rows, or null
    int across;        // image side read / written with lanes across neighbouring lines (image axis 1
                       // is the contiguous one, as for C-ordered images), else along each line
};

// The LDS twiddles, one table per pass of sub-transform size ns > 1 and
// radix R: entry tw_offset(ns) + (r - 1) ns + k = w^(k r K / (ns R)), so that
// the lanes of a pass (consecutive k) read consecutive entries.  The tables
// take fewer than K entries in all.
template <typename T, int K>
__device__ __forceinline__ cx<T>* stage_twiddles(unsigned char* smem, int lpw, const cx<T>* __restrict__ tw) {
    cx<T>* tws = reinterpret_cast<cx<T>*>(smem) + lpw * FftCfg<T, K>::KP;
    for (int i = threadIdx.x; i < K; i += blockDim.x) {
        int ns = 8, off = 0, R = 1;
        while (ns < K) {
            R = tw_radix<K>(ns);
            if (i - off < (R - 1) * ns) break;
            off += (R - 1) * ns;
            ns *= R;
        }
        if (ns >= K) continue;  // past the last table
        const int e = i - off, r = 1 + e / ns, k = e % ns;
        tws[i] = tw[k * r * (K / (ns * R))];
    }
    return tws;
}

__device__ __forceinline__ void row_of_line(int64_t line, const FftRowGeom& g, int& n2, int& n3, int64_t& r) {
    int64_t q = line;
    n2 = (int)(q % g.N[1]);
    q /= g.N[1];
    n3 = (int)(q % g.N[2]);
    r = q / g.N[2];
}

// Axis 0, forward: grid[r, n2, n3, :] = FFT_K0(pad(x[:, n2, n3, r] * sn)) for
// n2 < N1, n3 < N2 (sn as deapod_factor).
template <typename T, int K, int SIGN>
__global__ __launch_bounds__(kFftMaxThreads) void fft_rows_pad_kernel(const cx<T>* __restrict__ x,
                                                                      cx<T>* __restrict__ grid, FftRowGeom g,
                                                                      const double* sn0, const double* sn1,
                                                                      const double* sn2, int ndim,
                                                                      const cx<T>* __restrict__ tw) {
    using C = FftCfg<T, K>;
    constexpr int TL = C::TL;
    extern __shared__ __align__(16) unsigned char smem[];
    const int li = threadIdx.x / TL, tl = threadIdx.x % TL;
    const int64_t line = (int64_t)blockIdx.x * g.lpw + li;
    const bool live = line < g.nlines;
    int n2 = 0, n3 = 0;
    int64_t r = 0;
    if (live) row_of_line(line, g, n2, n3, r);
    cx<T>* ln = reinterpret_cast<cx<T>*>(smem) + li * C::KP;
    const double* sn[3] = {sn0, sn1, sn2};
    // image loads: along the line, or across neighbouring lines (consecutive
    // threads on consecutive lines, which are adjacent in memory when image
    // axis 1 is the contiguous one)
    const int lj = g.across ? (int)(threadIdx.x % (unsigned)g.lpw) : li;
    const int t0 = g.across ? (int)(threadIdx.x / (unsigned)g.lpw) : tl;
    int m2 = n2, m3 = n3;
    int64_t mr = r;
    const int64_t mline = (int64_t)blockIdx.x * g.lpw + lj;
    const bool mlive = mline < g.nlines;
    if (g.across && mlive) row_of_line(mline, g, m2, m3, mr);
    cx<T>* lm = reinterpret_cast<cx<T>*>(smem) + lj * C::KP;
    const cx<T>* xl = x + (int64_t)m2 * g.xs[1] + (int64_t)m3 * g.xs[2] + mr * g.xs[3];
#pragma unroll
    for (int kb = 0; kb < C::NL; kb += C::LF) {
        cx<T> v[C::LF];
#pragma unroll
        for (int u = 0; u < C::LF; ++u) {
            const int t = t0 + (kb + u) * TL;
            v[u] = (mlive && t < g.N[0]) ? xl[(int64_t)t * g.xs[0]] : mkcx<T>((T)0, (T)0);
        }
#pragma unroll
        for (int u = 0; u < C::LF; ++u) {
            const int t = t0 + (kb + u) * TL;
            cx<T> o = mkcx<T>((T)0, (T)0);
            if (mlive && t < g.N[0]) {
                const T s = sn0 ? deapod_factor<T>(sn, ndim, t, m2, m3) : (T)1;
                o.x = v[u].x * s;
                o.y = v[u].y * s;
            }
            lm[pad(t)] = o;
        }
    }
    cx<T>* tws = stage_twiddles<T, K>(smem, g.lpw, tw);  // after the data loads are in flight
    __syncthreads();
    fft_line<T, K, TL, SIGN>(ln, tl, tws);
    if (live) {
        cx<T>* out = grid + r * g.nk + ((int64_t)n3 * g.K[1] + n2) * K;
        for (int t = tl; t < K; t += TL) out[t] = ln[pad(t)];
    }
}

// Axis 0, adjoint: x[:N0, n2, n3, r] = IFFT_K0(grid[r, n2, n3, :])[:N0] *
// scale * sn for n2 < N1, n3 < N2 (the arithmetic of crop_deapod_kernel).
//   rezero: write zeros over the rows it read, so that a grid passed through
//   the adjoint with zero_tail set on every other pass ends all zero.
template <typename T, int K, int SIGN>
__global__ __launch_bounds__(kFftMaxThreads) void fft_rows_crop_kernel(cx<T>* __restrict__ grid,
                                                                       cx<T>* __restrict__ x, FftRowGeom g,
                                                                       const double* sn0, const double* sn1,
                                                                       const double* sn2, int ndim, T scale,
                                                                       const cx<T>* __restrict__ tw, int rezero) {
    using C = FftCfg<T, K>;
    constexpr int TL = C::TL;
    extern __shared__ __align__(16) unsigned char smem[];
    const int li = threadIdx.x / TL, tl = threadIdx.x % TL;
    const int64_t line = (int64_t)blockIdx.x * g.lpw + li;
    const bool live = line < g.nlines;
    int n2 = 0, n3 = 0;
    int64_t r = 0;
    if (live) row_of_line(line, g, n2, n3, r);
    cx<T>* ln = reinterpret_cast<cx<T>*>(smem) + li * C::KP;
    cx<T>* in = grid + r * g.nk + ((int64_t)n3 * g.K[1] + n2) * K;
    const cx<T>* mul = g.mult ? (const cx<T>*)g.mult + ((int64_t)n3 * g.K[1] + n2) * K : nullptr;
#pragma unroll
    for (int kb = 0; kb < C::NL; kb += C::LF) {
        cx<T> v[C::LF], w[C::LF];
#pragma unroll
        for (int u = 0; u < C::LF; ++u) {
            const int t = tl + (kb + u) * TL;
            v[u] = live ? in[t] : mkcx<T>((T)0, (T)0);
            if (mul && live) w[u] = mul[t];
        }
#pragma unroll
        for (int u = 0; u < C::LF; ++u) {
            const int t = tl + (kb + u) * TL;
            ln[pad(t)] = (mul && live) ? cmulc<T>(v[u], w[u]) : v[u];
        }
    }
    cx<T>* tws = stage_twiddles<T, K>(smem, g.lpw, tw);
    __syncthreads();
    if (rezero && live)  // each thread clears the points it loaded
        for (int t = tl; t < K; t += TL) in[t] = mkcx<T>((T)0, (T)0);
    fft_line<T, K, TL, SIGN>(ln, tl, tws);
    // image stores: along the line, or across neighbouring lines (see
    // fft_rows_pad_kernel), which needs the whole block's lines done
    const int lj = g.across ? (int)(threadIdx.x % (unsigned)g.lpw) : li;
    const int t0 = g.across ? (int)(threadIdx.x / (unsigned)g.lpw) : tl;
    int m2 = n2, m3 = n3;
    int64_t mr = r;
    const int64_t mline = (int64_t)blockIdx.x * g.lpw + lj;
    const bool mlive = mline < g.nlines;
    if (g.across) {
        __syncthreads();
        if (mlive) row_of_line(mline, g, m2, m3, mr);
    }
    const cx<T>* lm = reinterpret_cast<const cx<T>*>(smem) + lj * C::KP;
    if (mlive) {
        const double* sn[3] = {sn0, sn1, sn2};
        for (int t = t0; t < g.N[0]; t += TL) {
            const T s = sn0 ? deapod_factor<T>(sn, ndim, t, m2, m3) : (T)1;
            const cx<T> v = lm[pad(t)];
            cx<T> o;
            o.x = (v.x * scale) * s;
            o.y = (v.y * scale) * s;
            x[(int64_t)t * g.xs[0] + (int64_t)m2 * g.xs[1] + (int64_t)m3 * g.xs[2] + mr * g.xs[3]] = o;
        }
    }
}

struct FftColGeom {
    int64_t S;        // stride of the axis = number of lines side by side (prod of lower K)
    int n_in;         // points read per line (the rest is zero)
    int n_out;        // points written per line
    int A;            // active indices of the next higher axis (1 if none)
    int64_t st_a;     // grid stride of that axis
    int64_t nk;       // prod K (stride of an image)
    int64_t nouter;   // A * nimg
    int lpw;          // neighbouring lines per workgroup, divides S (blockDim = lpw * TL)
    int zero_tail;    // also write zeros at n_out <= t < K
    const void* mult; // complex T grid (one image) multiplying the loaded lines, or null
    // Last forward pass / first adjoint pass of an ndim >= 2 grid: pb = 1
    // multiplies what the pass writes by phase_before (then divides by
    // scale_ortho if ortho), pb = 2 multiplies what it loads by
    // conj(phase_before); the element's axis-0 (and axis-1) index comes from
    // the inner line index i0 = k0 + K0 * k1 (K0 = k0_mask + 1, a power of two).
    int pb;
    int ortho;
    int pb_ndim;
    int k0_mask, k0_shift;
    double pbc[3];      // phase_before coefficients, rounded to T in the kernel
    double scale_ortho;
};

// grid value times phase_before (then / scale_ortho): the arithmetic of the
// tiled gather's tile staging, so the staged values are unchanged
template <typename T>
__device__ __forceinline__ cx<T> fft_pb_apply(cx<T> w, const FftColGeom& g, int64_t i0, int t, bool conj) {
    const T c[3] = {(T)g.pbc[0], (T)g.pbc[1], (T)g.pbc[2]};
    int k0, k1, k2;
    if (g.pb_ndim == 2) {
        k0 = (int)i0;
        k1 = t;
        k2 = 0;
    } else {
        k0 = (int)(i0 & g.k0_mask);
        k1 = (int)(i0 >> g.k0_shift);
        k2 = t;
    }
    cx<T> p = cis_tile(phase_before_arg<T>(c, g.pb_ndim, k0, k1, k2));
    if (conj) p.y = -p.y;
    cx<T> o;
    o.x = w.x * p.x - w.y * p.y;
    o.y = w.x * p.y + w.y * p.x;
    if (!conj && g.ortho) {
        o.x = o.x / (T)g.scale_ortho;
        o.y = o.y / (T)g.scale_ortho;
    }
    return o;
}

// Axis d >= 1 in place: lpw neighbouring lines (inner index i0..i0+lpw-1) of
// one outer index per workgroup; each row of the block is one contiguous
// lpw-element segment.
template <typename T, int K, int SIGN>
__global__ __launch_bounds__(kFftMaxThreads) void fft_cols_kernel(cx<T>* __restrict__ grid, FftColGeom g,
                                                                  const cx<T>* __restrict__ tw) {
    using C = FftCfg<T, K>;
    constexpr int TL = C::TL;
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* lds = reinterpret_cast<cx<T>*>(smem);
    const int lpw = g.lpw;
    const int64_t nchunk = g.S / lpw;
    const int64_t bid = xcd_remap(blockIdx.x, gridDim.x);  // neighbouring chunks share 128-byte lines
    const int64_t chunk = bid % nchunk;
    const int64_t o = bid / nchunk;
    const int64_t ibase = (o % g.A) * g.st_a + chunk * lpw;  // within one image
    const int64_t base = ibase + (o / g.A) * g.nk;
    const int li = threadIdx.x % lpw, t0 = threadIdx.x / lpw;  // load / store mapping
    const cx<T>* mul = (const cx<T>*)g.mult;
    const int64_t i0 = chunk * lpw + li;  // inner index of this thread's line
#pragma unroll
    for (int kb = 0; kb < C::NL; kb += C::LF) {
        cx<T> v[C::LF], w[C::LF];
#pragma unroll
        for (int u = 0; u < C::LF; ++u) {
            const int t = t0 + (kb + u) * TL;
            v[u] = t < g.n_in ? grid[base + (int64_t)t * g.S + li] : mkcx<T>((T)0, (T)0);
            if (mul && t < g.n_in) w[u] = mul[ibase + (int64_t)t * g.S + li];
        }
#pragma unroll
        for (int u = 0; u < C::LF; ++u) {
            const int t = t0 + (kb + u) * TL;
            cx<T> x = v[u];
            if (t < g.n_in) {
                if (mul) x = cmulc<T>(x, w[u]);
                if (g.pb == 2) x = fft_pb_apply<T>(x, g, i0, t, true);
            }
            lds[li * C::KP + pad(t)] = x;
        }
    }
    cx<T>* tws = stage_twiddles<T, K>(smem, g.lpw, tw);
    __syncthreads();
    fft_line<T, K, TL, SIGN>(lds + (threadIdx.x / TL) * C::KP, threadIdx.x % TL, tws);
    for (int t = t0; t < g.n_out; t += TL) {
        cx<T> v = lds[li * C::KP + pad(t)];
        if (g.pb == 1) v = fft_pb_apply<T>(v, g, i0, t, false);
        grid[base + (int64_t)t * g.S + li] = v;
    }
    if (g.zero_tail)
        for (int t = g.n_out + t0; t < K; t += TL) grid[base + (int64_t)t * g.S + li] = mkcx<T>((T)0, (T)0);
}

// Circulant embedding of a point-spread function given on the lattice
// d_e in [-N_e, N_e) (element q_e = d_e + N_e, strides ps): T[d mod K] =
// psf[d] * inv_nk for d_e > -N_e on every axis.  T must be zero beforehand;
// K_e >= 2 N_e keeps the wrapped indices distinct.
template <typename T>
__global__ __launch_bounds__(256) void toeplitz_embed_kernel(const cx<T>* __restrict__ psf, cx<T>* __restrict__ t,
                                                             int ndim, int N0, int N1, int N2, int K0, int K1,
                                                             int K2, int64_t ps0, int64_t ps1, int64_t ps2,
                                                             int64_t count, T inv_nk) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const int L0 = 2 * N0, L1 = 2 * N1;
    const int q0 = (int)(i % L0);
    const int64_t r = i / L0;
    const int q1 = ndim > 1 ? (int)(r % L1) : 0;
    const int q2 = ndim > 2 ? (int)(r / L1) : 0;
    const int d0 = q0 - N0, d1 = ndim > 1 ? q1 - N1 : 0, d2 = ndim > 2 ? q2 - N2 : 0;
    if (d0 == -N0 || (ndim > 1 && d1 == -N1) || (ndim > 2 && d2 == -N2)) return;
    const int k0 = d0 < 0 ? d0 + K0 : d0;
    const int k1 = d1 < 0 ? d1 + K1 : d1;
    const int k2 = d2 < 0 ? d2 + K2 : d2;
    const cx<T> v = psf[(int64_t)q0 * ps0 + (int64_t)q1 * ps1 + (int64_t)q2 * ps2];
    t[((int64_t)k2 * K1 + k1) * K0 + k0] = mkcx<T>(v.x * inv_nk, v.y * inv_nk);
}

}  // namespace mi355
