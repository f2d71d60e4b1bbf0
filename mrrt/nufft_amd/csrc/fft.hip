// Dispatch of the fused pruned FFT passes (kernels_fft.hpp): one pass along
// axis 0 or along an axis d >= 1, for every supported power-of-two length.
#include "internal.hpp"

namespace mi355host {

FftRowGeom row_geom(const FftDev& f, const int64_t* xs, int32_t nimg) {
    FftRowGeom g;
    for (int d = 0; d < 3; ++d) {
        g.N[d] = (int)f.N[d];
        g.K[d] = (int)f.K[d];
    }
    g.nk = f.nk;
    for (int d = 0; d < 4; ++d) g.xs[d] = 0;
    for (int d = 0; d < f.ndim; ++d) g.xs[d] = xs[d];
    g.xs[3] = xs[f.ndim];
    g.nlines = (int64_t)g.N[1] * g.N[2] * nimg;
    g.mult = nullptr;
    // C-ordered images (axis 1 contiguous): neighbouring lines are adjacent
    // in memory, so the image side is read / written across lines
    g.across = (f.ndim >= 2 && g.xs[1] == 1 && g.xs[0] != 1) ? 1 : 0;
    return g;
}

// lines of axis d >= 1: LPW neighbouring lines of prod(K_0..K_{d-1}) per
// workgroup, for every active index of the next axis and every image
FftColGeom col_geom(const FftDev& f, int d, bool adjoint, int32_t nimg) {
    FftColGeom g;
    g.S = 1;
    for (int e = 0; e < d; ++e) g.S *= f.K[e];
    g.n_in = (int)(adjoint ? f.K[d] : f.N[d]);
    g.n_out = (int)(adjoint ? f.N[d] : f.K[d]);
    g.A = d + 1 < f.ndim ? (int)f.N[d + 1] : 1;
    g.st_a = g.S * f.K[d];
    g.nk = f.nk;
    g.nouter = (int64_t)g.A * nimg;
    g.mult = nullptr;
    g.pb = 0;
    g.ortho = 0;
    g.pb_ndim = f.ndim;
    g.k0_mask = (int)f.K[0] - 1;
    g.k0_shift = 0;
    while (((int64_t)1 << g.k0_shift) < f.K[0]) ++g.k0_shift;
    g.pbc[0] = g.pbc[1] = g.pbc[2] = 0.0;
    g.scale_ortho = 1.0;
    return g;
}

// Workgroups a pass aims for (measured best on 512^2 grids): lines per workgroup grow only
// while the pass keeps at least this many workgroups.
constexpr int64_t kFftTargetBlocks = 128;
// LDS cap of the row passes over C-ordered images (room for 16 lines of 512
// complex64 points)
constexpr size_t kFftLdsAcross = 81920;

template <typename T, int K>
int launch_fft_rows(const FftDev& f, bool adjoint, const void* x, void* xo, const int64_t* xs, void* grid,
                    int32_t nimg, T scale, FftPassOpts o) {
    using C = FftCfg<T, K>;
    FftRowGeom g = row_geom(f, xs, nimg);
    g.mult = o.mult;
    // more lines per workgroup across C-ordered images, for longer contiguous
    // runs of the image side (16 lines of complex64: one 128-byte line)
    const int max_threads = g.across ? 1024 : 256;
    const size_t lds_cap = g.across ? kFftLdsAcross : (size_t)kFftLds;
    int lpw = 1;
    while (2 * lpw * C::TL <= max_threads && C::lds_bytes(2 * lpw) <= lds_cap &&
           g.nlines / (2 * lpw) >= kFftTargetBlocks)
        lpw *= 2;
    g.lpw = lpw;
    const int64_t nb = (g.nlines + lpw - 1) / lpw;
    const size_t lds = C::lds_bytes(lpw);
    const double* s0 = f.sn ? f.sn[0] : nullptr;
    const double* s1 = f.sn && f.ndim > 1 ? f.sn[1] : s0;
    const double* s2 = f.sn && f.ndim > 2 ? f.sn[2] : s0;
    if (adjoint)
        hipLaunchKernelGGL((fft_rows_crop_kernel<T, K, 1>), dim3((unsigned)nb), dim3(lpw * C::TL), lds, f.stream,
                           (cx<T>*)grid, (cx<T>*)xo, g, s0, s1, s2, f.ndim, scale, (const cx<T>*)f.tw[0],
                           o.rezero);
    else
        hipLaunchKernelGGL((fft_rows_pad_kernel<T, K, -1>), dim3((unsigned)nb), dim3(lpw * C::TL), lds, f.stream,
                           (const cx<T>*)x, (cx<T>*)grid, g, s0, s1, s2, f.ndim, (const cx<T>*)f.tw[0]);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

template <typename T, int K>
int launch_fft_cols(const FftDev& f, bool adjoint, int d, void* grid, int32_t nimg, FftPassOpts o) {
    using C = FftCfg<T, K>;
    FftColGeom g = col_geom(f, d, adjoint, nimg);
    g.zero_tail = o.rezero;
    g.mult = o.mult;
    if (o.pb) {
        g.pb = o.pb;
        for (int e = 0; e < 3; ++e) g.pbc[e] = o.pbc ? o.pbc[e] : 0.0;
        g.ortho = o.ortho;
        g.scale_ortho = o.scale_ortho;
    }
    // up to 128 contiguous bytes per row, while the pass keeps its workgroups
    int lpw = 1;
    while (2 * lpw * C::CXS <= 128 && g.S % (2 * lpw) == 0 && 2 * lpw * C::TL <= kFftMaxThreads &&
           C::lds_bytes(2 * lpw) <= (size_t)kFftLds && (g.S / (2 * lpw)) * g.nouter >= kFftTargetBlocks)
        lpw *= 2;
    g.lpw = lpw;
    const int64_t nb = (g.S / lpw) * g.nouter;
    if (nb >= ((int64_t)1 << 32)) return fail(MI355NUFFT_EINVAL, "FFT launch too large");
    const size_t lds = C::lds_bytes(lpw);
    if (adjoint)
        hipLaunchKernelGGL((fft_cols_kernel<T, K, 1>), dim3((unsigned)nb), dim3(lpw * C::TL), lds, f.stream,
                           (cx<T>*)grid, g, (const cx<T>*)f.tw[d]);
    else
        hipLaunchKernelGGL((fft_cols_kernel<T, K, -1>), dim3((unsigned)nb), dim3(lpw * C::TL), lds, f.stream,
                           (cx<T>*)grid, g, (const cx<T>*)f.tw[d]);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

#define MI355_FFT_SWITCH(KV, CALL)                                        \
    switch (KV) {                                                         \
        case 16: return CALL(16);                                         \
        case 32: return CALL(32);                                         \
        case 64: return CALL(64);                                         \
        case 128: return CALL(128);                                       \
        case 256: return CALL(256);                                       \
        case 512: return CALL(512);                                       \
        case 1024: return CALL(1024);                                     \
        case 2048: return CALL(2048);                                     \
        default: return fail(MI355NUFFT_EINVAL, "no fused FFT for K=%d", (int)(KV)); \
    }

template <typename T>
int fft_rows(const FftDev& f, bool adjoint, const void* x, void* xo, const int64_t* xs, void* grid, int32_t nimg,
             T scale, FftPassOpts o) {
#define MI355_ROWS(KK) launch_fft_rows<T, KK>(f, adjoint, x, xo, xs, grid, nimg, scale, o)
    MI355_FFT_SWITCH((int)f.K[0], MI355_ROWS)
#undef MI355_ROWS
}

template <typename T>
int fft_cols(const FftDev& f, bool adjoint, int d, void* grid, int32_t nimg, FftPassOpts o) {
#define MI355_COLS(KK) launch_fft_cols<T, KK>(f, adjoint, d, grid, nimg, o)
    MI355_FFT_SWITCH((int)f.K[d], MI355_COLS)
#undef MI355_COLS
}

template int fft_rows<float>(const FftDev&, bool, const void*, void*, const int64_t*, void*, int32_t, float,
                             FftPassOpts);
template int fft_rows<double>(const FftDev&, bool, const void*, void*, const int64_t*, void*, int32_t, double,
                              FftPassOpts);
template int fft_cols<float>(const FftDev&, bool, int, void*, int32_t, FftPassOpts);
template int fft_cols<double>(const FftDev&, bool, int, void*, int32_t, FftPassOpts);

}  // namespace mi355host
