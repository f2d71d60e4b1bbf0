// C ABI of the MI355X table1 NUFFT (declarations and contracts in
// include/mi355nufft.h).  One plan owns the device copies of the tables and
// deapodization factors, the per-sample state (tm, phase), the rocFFT plans
// with their work buffer and an optional grid scratch buffer.  Everything is
// enqueued on the plan's stream.
#include <hipcub/hipcub.hpp>

#include <mutex>

#include "internal.hpp"
#include "kernels_coil.hpp"
#include "kernels_probe.hpp"

namespace mi355host {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// Kernel timing: while timing is on and the stream is not being captured,
// each tiled gather or scatter launch goes through hipExtLaunchKernel with a
// start/stop event pair that the dispatch itself timestamps (the kernel's own
// execution, not the gap before it).
constexpr int kTimingPairs = 1024;

int timing_fold(mi355nufft_plan* p, int w) {
    for (int i = 0; i < p->tev_used[w]; ++i) {
        float ms = 0.f;
        HIP_TRY(hipEventSynchronize(p->tev[w][2 * i + 1]));
        HIP_TRY(hipEventElapsedTime(&ms, p->tev[w][2 * i], p->tev[w][2 * i + 1]));
        p->tev_ms[w] += ms;
        ++p->tev_n[w];
    }
    p->tev_used[w] = 0;
    return MI355NUFFT_OK;
}

// events for the next launch of kind w, or nulls when not timing
int timing_events(mi355nufft_plan* p, int w, hipEvent_t* start, hipEvent_t* stop) {
    *start = *stop = nullptr;
    if (!p->timing) return MI355NUFFT_OK;
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    HIP_TRY(hipStreamIsCapturing(p->stream, &cs));
    if (cs != hipStreamCaptureStatusNone) return MI355NUFFT_OK;
    if (p->tev_used[w] == kTimingPairs) RET_IF(timing_fold(p, w));
    const size_t need = 2 * (size_t)(p->tev_used[w] + 1);
    while (p->tev[w].size() < need) {
        hipEvent_t e;
        HIP_TRY(hipEventCreate(&e));
        p->tev[w].push_back(e);
    }
    const int i = p->tev_used[w]++;
    *start = p->tev[w][2 * i];
    *stop = p->tev[w][2 * i + 1];
    return MI355NUFFT_OK;
}

}  // namespace mi355host


namespace {

int check_geometry(int ndim, const int64_t* Kd, const int32_t* Jd, int32_t L) {
    if (ndim < 1 || ndim > 3) return fail(MI355NUFFT_EINVAL, "ndim must be 1, 2 or 3 (got %d)", ndim);
    if (!Kd || !Jd) return fail(MI355NUFFT_EINVAL, "Kd and Jd are required");
    if (L < 1) return fail(MI355NUFFT_EINVAL, "L must be positive (got %d)", L);
    int64_t nk = 1;
    for (int d = 0; d < ndim; ++d) {
        if (Kd[d] < 1) return fail(MI355NUFFT_EINVAL, "Kd[%d] must be positive", d);
        if (Jd[d] < 1 || Jd[d] > 16)
            return fail(MI355NUFFT_EINVAL, "Jd[%d]=%d outside the supported range 1..16", d, Jd[d]);
        if ((int64_t)Jd[d] * L + 1 > (int64_t)1 << 30)
            return fail(MI355NUFFT_EINVAL, "table length J*L+1 too large on axis %d", d);
        nk *= Kd[d];
    }
    if (nk >= ((int64_t)1 << 31)) return fail(MI355NUFFT_EINVAL, "prod(Kd) must be below 2^31");
    return MI355NUFFT_OK;
}


template <typename T, int NDIM, bool CK, int JMAX>
void launch_interp(bool adjoint, const InterpArgs<T, CK>& a, hipStream_t s) {
    if (a.M == 0) return;
    if (adjoint)
        hipLaunchKernelGGL((spread_gm_kernel<T, NDIM, CK, JMAX>), dim3(nblocks(a.M)), dim3(kBlock), 0, s, a);
    else
        hipLaunchKernelGGL((interp_fwd_kernel<T, NDIM, CK, JMAX>), dim3(nblocks(a.M)), dim3(kBlock), 0, s, a);
}

template <typename T, bool CK>
int dispatch_interp(bool adjoint, int ndim, int jmax, const InterpArgs<T, CK>& a, hipStream_t s) {
#define MI355_CASE(ND, JM)                                          \
    if (ndim == ND && jmax == JM) {                                 \
        launch_interp<T, ND, CK, JM>(adjoint, a, s);                \
        LAUNCH_CHECK();                                             \
        return MI355NUFFT_OK;                                       \
    }
    MI355_CASE(1, 4) MI355_CASE(1, 8) MI355_CASE(1, 16)
    MI355_CASE(2, 4) MI355_CASE(2, 8) MI355_CASE(2, 16)
    MI355_CASE(3, 4) MI355_CASE(3, 8) MI355_CASE(3, 16)
#undef MI355_CASE
    return fail(MI355NUFFT_EINVAL, "unsupported ndim/J combination");
}

template <typename T>
int zero_grid(void* grid, int64_t n, hipStream_t s) {
    if (n == 0) return MI355NUFFT_OK;
    hipLaunchKernelGGL((zero_kernel<T>), dim3(nblocks(n)), dim3(kBlock), 0, s, (cx<T>*)grid, n);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

template <typename T>
int run_table1(bool adjoint, int ndim, const int64_t* Kd, const int32_t* Jd, int32_t L, int complex_kernel,
               const void* samples_in, void* samples_out, const void* grid_in, void* grid_out,
               const void* const* h, const void* tm, int64_t M, int32_t nreps, const void* phase,
               hipStream_t s) {
    const int jm = jmax_of(ndim, Jd);
    if (complex_kernel) {
        InterpArgs<T, true> a;
        fill_interp_args<T, true>(a, ndim, Kd, Jd, L, h, tm, M, nreps);
        a.grid = (const cx<T>*)grid_in;
        a.grid_out = (cx<T>*)grid_out;
        a.in = (const cx<T>*)samples_in;
        a.out = (cx<T>*)samples_out;
        a.phase = (const cx<T>*)phase;
        return dispatch_interp<T, true>(adjoint, ndim, jm, a, s);
    }
    InterpArgs<T, false> a;
    fill_interp_args<T, false>(a, ndim, Kd, Jd, L, h, tm, M, nreps);
    a.grid = (const cx<T>*)grid_in;
    a.grid_out = (cx<T>*)grid_out;
    a.in = (const cx<T>*)samples_in;
    a.out = (cx<T>*)samples_out;
    a.phase = (const cx<T>*)phase;
    return dispatch_interp<T, false>(adjoint, ndim, jm, a, s);
}

GridGeom grid_geom(const mi355nufft_plan* p, const int64_t* xs) {
    GridGeom g;
    g.ndim = p->ndim;
    for (int d = 0; d < 3; ++d) {
        g.K[d] = (int)p->K[d];
        g.N[d] = (int)p->N[d];
    }
    g.nk = p->nk;
    for (int d = 0; d < 4; ++d) g.xs[d] = 0;
    if (xs) {
        for (int d = 0; d < p->ndim; ++d) g.xs[d] = xs[d];
        g.xs[3] = xs[p->ndim];
    }
    return g;
}

int check_plan(const mi355nufft_plan* p) {
    if (!p) return fail(MI355NUFFT_EINVAL, "null plan");
    return MI355NUFFT_OK;
}

int check_nreps(int32_t nreps) {
    if (nreps < 1 || nreps > 65535) return fail(MI355NUFFT_EINVAL, "nreps must be in 1..65535 (got %d)", nreps);
    return MI355NUFFT_OK;
}


int get_agrid(mi355nufft_plan* p, int32_t nimg) {
    const size_t need = (size_t)nimg * (size_t)p->nk * p->cx_size();
    if (need > p->agrid_bytes) {
        p->agrid_zero = 0;
        if (p->d_agrid) {
            HIP_TRY(hipStreamSynchronize(p->stream));
            HIP_TRY(hipFree(p->d_agrid));
            p->d_agrid = nullptr;
            p->agrid_bytes = 0;
        }
        if (hipMalloc(&p->d_agrid, need) != hipSuccess)
            return fail(MI355NUFFT_ENOMEM, "adjoint grid of %zu bytes", need);
        p->agrid_bytes = need;
    }
    return MI355NUFFT_OK;
}

bool stream_capturing(const mi355nufft_plan* p) {
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(p->stream, &cs) != hipSuccess) return true;  // be conservative
    return cs != hipStreamCaptureStatusNone;
}

int get_grid(mi355nufft_plan* p, int32_t nreps, void** grid) {
    if (*grid) return MI355NUFFT_OK;
    const size_t need = (size_t)nreps * (size_t)p->nk * p->cx_size();
    if (need > p->grid_bytes) {
        if (p->d_grid) {
            HIP_TRY(hipStreamSynchronize(p->stream));
            HIP_TRY(hipFree(p->d_grid));
            p->d_grid = nullptr;
            p->grid_bytes = 0;
        }
        if (hipMalloc(&p->d_grid, need) != hipSuccess)
            return fail(MI355NUFFT_ENOMEM, "grid scratch of %zu bytes", need);
        p->grid_bytes = need;
    }
    *grid = p->d_grid;
    return MI355NUFFT_OK;
}

const char* fft_status_str(rocfft_status s) {
    switch (s) {
        case rocfft_status_success: return "success";
        case rocfft_status_failure: return "failure";
        case rocfft_status_invalid_arg_value: return "invalid arg value";
        case rocfft_status_invalid_dimensions: return "invalid dimensions";
        case rocfft_status_invalid_array_type: return "invalid array type";
        case rocfft_status_invalid_strides: return "invalid strides";
        case rocfft_status_invalid_distance: return "invalid distance";
        case rocfft_status_invalid_offset: return "invalid offset";
        case rocfft_status_invalid_work_buffer: return "invalid work buffer";
        default: return "unknown";
    }
}

#define FFT_TRY(expr)                                                                    \
    do {                                                                                 \
        rocfft_status s_ = (expr);                                                       \
        if (s_ != rocfft_status_success)                                                 \
            return fail(MI355NUFFT_EFFT, "%s: %s", #expr, fft_status_str(s_));           \
    } while (0)

int get_fft(mi355nufft_plan* p, int32_t nreps, FftPlans** out) {
    for (auto& f : p->ffts) {
        if (f.nreps == nreps) {
            *out = &f;
            return MI355NUFFT_OK;
        }
    }
    static bool setup_done = false;
    if (!setup_done) {
        FFT_TRY(rocfft_setup());
        setup_done = true;
    }
    size_t lengths[3];
    for (int d = 0; d < p->ndim; ++d) lengths[d] = (size_t)p->K[d];  // fastest first
    const rocfft_precision prec = p->precision ? rocfft_precision_double : rocfft_precision_single;
    FftPlans f;
    f.nreps = nreps;
    f.fwd = nullptr;
    f.bwd = nullptr;
    rocfft_plan_description desc = nullptr;
    FFT_TRY(rocfft_plan_description_create(&desc));
    size_t dist = (size_t)p->nk;
    rocfft_status st = rocfft_plan_description_set_data_layout(
        desc, rocfft_array_type_complex_interleaved, rocfft_array_type_complex_interleaved, nullptr, nullptr, 0,
        nullptr, dist, 0, nullptr, dist);
    if (st == rocfft_status_success)
        st = rocfft_plan_create(&f.fwd, rocfft_placement_inplace, rocfft_transform_type_complex_forward, prec,
                                (size_t)p->ndim, lengths, (size_t)nreps, desc);
    if (st == rocfft_status_success)
        st = rocfft_plan_create(&f.bwd, rocfft_placement_inplace, rocfft_transform_type_complex_inverse, prec,
                                (size_t)p->ndim, lengths, (size_t)nreps, desc);
    rocfft_plan_description_destroy(desc);
    if (st != rocfft_status_success) {
        if (f.fwd) rocfft_plan_destroy(f.fwd);
        if (f.bwd) rocfft_plan_destroy(f.bwd);
        return fail(MI355NUFFT_EFFT, "rocfft_plan_create: %s", fft_status_str(st));
    }
    size_t w1 = 0, w2 = 0;
    FFT_TRY(rocfft_plan_get_work_buffer_size(f.fwd, &w1));
    FFT_TRY(rocfft_plan_get_work_buffer_size(f.bwd, &w2));
    const size_t w = w1 > w2 ? w1 : w2;
    if (!p->fft_info) FFT_TRY(rocfft_execution_info_create(&p->fft_info));
    if (w > p->fft_work_bytes) {
        if (p->fft_work) {
            HIP_TRY(hipStreamSynchronize(p->stream));
            HIP_TRY(hipFree(p->fft_work));
            p->fft_work = nullptr;
        }
        if (hipMalloc(&p->fft_work, w) != hipSuccess) return fail(MI355NUFFT_ENOMEM, "fft work %zu bytes", w);
        p->fft_work_bytes = w;
    }
    p->ffts.push_back(f);
    *out = &p->ffts.back();
    return MI355NUFFT_OK;
}

template <typename T>
int deapod_pad_impl(mi355nufft_plan* p, const void* x, const int64_t* xs, void* grid, int32_t nreps) {
    GridGeom g = grid_geom(p, xs);
    hipLaunchKernelGGL((deapod_pad_kernel<T>), dim3(nblocks(p->nk), nreps), dim3(kBlock), 0, p->stream,
                       (const cx<T>*)x, (cx<T>*)grid, g, p->d_sn[0], p->ndim > 1 ? p->d_sn[1] : p->d_sn[0],
                       p->ndim > 2 ? p->d_sn[2] : p->d_sn[0]);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

// with_phase = 0 multiplies by exp(i 0) = 1 (used for the ortho scaling of the
// complex kernel, which has no phase_before)
template <typename T>
int phase_grid_impl(mi355nufft_plan* p, void* grid, int32_t nreps, int with_phase, int conj_phase, int ortho) {
    GridGeom g = grid_geom(p, nullptr);
    const double* c = p->phase_coef;
    const T c0 = with_phase ? (T)c[0] : (T)0, c1 = with_phase ? (T)c[1] : (T)0, c2 = with_phase ? (T)c[2] : (T)0;
    hipLaunchKernelGGL((phase_grid_kernel<T>), dim3(nblocks(p->nk)), dim3(kBlock), 0, p->stream, (cx<T>*)grid,
                       g, nreps, c0, c1, c2, conj_phase, ortho, (T)p->scale_ortho);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

template <typename T>
int crop_deapod_impl(mi355nufft_plan* p, const void* grid, void* x, const int64_t* xs, int32_t nreps,
                     int rezero = 0) {
    GridGeom g = grid_geom(p, xs);
    double scale = p->adjoint_scale;
    if (p->ortho) scale = p->scale_ortho * p->adjoint_scale / (double)p->nk;
    if (rezero) {
        hipLaunchKernelGGL((crop_deapod_rezero_kernel<T>), dim3(nblocks(p->nk), nreps), dim3(kBlock), 0, p->stream,
                           (cx<T>*)grid, (cx<T>*)x, g, p->d_sn[0], p->ndim > 1 ? p->d_sn[1] : p->d_sn[0],
                           p->ndim > 2 ? p->d_sn[2] : p->d_sn[0], (T)scale);
        LAUNCH_CHECK();
        return MI355NUFFT_OK;
    }
    hipLaunchKernelGGL((crop_deapod_kernel<T>), dim3(nblocks(p->nn), nreps), dim3(kBlock), 0, p->stream,
                       (const cx<T>*)grid, (cx<T>*)x, g, p->nn, p->d_sn[0],
                       p->ndim > 1 ? p->d_sn[1] : p->d_sn[0], p->ndim > 2 ? p->d_sn[2] : p->d_sn[0], (T)scale);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

int fft_impl(mi355nufft_plan* p, void* grid, int32_t nreps, int direction);

// Fused pruned FFT passes (kernels_fft.hpp) for power-of-two grids.
bool use_fused_fft(const mi355nufft_plan* p) { return p->fft_method == 0 && p->fft_fused_ok; }


// Every K_d a power of two the fused passes support at this precision: one
// padded line and the twiddle table must fit the LDS budget (complex128
// lines of 2048 points do not).
bool fused_fft_sizes_ok(int ndim, const int64_t* Kd, size_t cx_size) {
    for (int d = 0; d < ndim; ++d) {
        const int64_t k = Kd[d];
        if (k < kFftMinK || k > kFftMaxK || (k & (k - 1)) != 0 ||
            (size_t)(2 * k + k / 32 + 1) * cx_size > (size_t)kFftLds)  // FftCfg::lds_bytes(1)
            return false;
    }
    return true;
}

// twiddles exp(-2 pi i m / K_d), m < K_d, rounded once from double
int make_twiddles(int ndim, const int64_t* Kd, int precision, void** tw) {
    for (int d = 0; d < ndim; ++d) {
        const int64_t k = Kd[d];
        std::vector<double> w((size_t)(2 * k));
        for (int64_t m = 0; m < k; ++m) {
            const double a = 2.0 * M_PI * (double)m / (double)k;
            w[(size_t)(2 * m)] = std::cos(a);
            w[(size_t)(2 * m + 1)] = -std::sin(a);
        }
        const size_t bytes = (size_t)k * 2 * (precision ? sizeof(double) : sizeof(float));
        std::vector<float> wf;
        const void* src = w.data();
        if (!precision) {
            wf.assign(w.begin(), w.end());
            src = wf.data();
        }
        if (hipMalloc(&tw[d], bytes) != hipSuccess) return fail(MI355NUFFT_ENOMEM, "FFT twiddle table");
        if (hipMemcpy(tw[d], src, bytes, hipMemcpyHostToDevice) != hipSuccess)
            return fail(MI355NUFFT_EHIP, "FFT twiddle upload");
    }
    return MI355NUFFT_OK;
}

double adjoint_image_scale(const mi355nufft_plan* p) {
    return p->ortho ? p->scale_ortho * p->adjoint_scale / (double)p->nk : p->adjoint_scale;
}

// image -> oversampled grid: fftn(x * sn, Kd) (_nufft.py:1326-1331); the
// phase_before of :1371 is applied by the gather
// pb (fused passes, ndim >= 2 only): the last pass also applies
// phase_before (real kernel) and the ortho scale, as the gather's staging
// would.
template <typename T>
int forward_grid_impl(mi355nufft_plan* p, const void* x, const int64_t* xs, void* grid, int32_t nimg,
                      int pb = 0) {
    if (!use_fused_fft(p)) {
        RET_IF(deapod_pad_impl<T>(p, x, xs, grid, nimg));
        return fft_impl(p, grid, nimg, -1);
    }
    const FftDev f = fft_dev(p);
    RET_IF(fft_rows<T>(f, false, x, nullptr, xs, grid, nimg, (T)1));
    for (int d = 1; d < p->ndim; ++d) {
        FftPassOpts o;
        if (pb && d == p->ndim - 1) {
            o.pb = 1;
            o.pbc = p->phasing == 0 ? p->phase_coef : nullptr;
            o.ortho = p->ortho;
            o.scale_ortho = p->scale_ortho;
        }
        RET_IF(fft_cols<T>(f, false, d, grid, nimg, o));
    }
    return MI355NUFFT_OK;
}

// oversampled grid -> image: ifftn, crop to Nd, scale, * conj(sn)
// (_nufft.py:1524-1572); the conj(phase_before) of :1519 is applied by the
// scatter.  The grid is overwritten; with rezero it is left all zero.
// pb (fused passes, real kernel, ndim >= 2 only): the first pass applies
// conj(phase_before) instead of the scatter.
template <typename T>
int adjoint_image_impl(mi355nufft_plan* p, void* grid, void* x, const int64_t* xs, int32_t nimg, int rezero,
                       int pb = 0) {
    if (!use_fused_fft(p)) {
        RET_IF(fft_impl(p, grid, nimg, 1));
        return crop_deapod_impl<T>(p, grid, x, xs, nimg, rezero);
    }
    const FftDev f = fft_dev(p);
    FftPassOpts o;
    o.rezero = rezero;
    for (int d = p->ndim - 1; d >= 1; --d) {
        FftPassOpts od = o;
        if (pb && d == p->ndim - 1) {
            od.pb = 2;
            od.pbc = p->phase_coef;
        }
        RET_IF(fft_cols<T>(f, true, d, grid, nimg, od));
    }
    return fft_rows<T>(f, true, nullptr, x, xs, grid, nimg, (T)adjoint_image_scale(p), o);
}

// ---------------------------------------------------------------- sparse mode
void sparse_free(mi355nufft_plan* p) {
    SparseState& s = p->sp;
    void* bufs[] = {s.val, s.rbase, s.perm, s.tval, s.trow, s.pieces, s.splits, s.part, s.ybuf};
    bool any = false;
    for (void* b : bufs) any = any || b;
    if (any) (void)hipStreamSynchronize(p->stream);
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    s = SparseState();
}

// Device temporaries of one call, freed after the stream drains on every
// exit path.
struct DevTemps {
    hipStream_t s;
    std::vector<void*> v;
    explicit DevTemps(hipStream_t st) : s(st) {}
    void* get(size_t bytes) {
        void* q = nullptr;
        if (hipMalloc(&q, bytes ? bytes : 1) != hipSuccess) return nullptr;
        v.push_back(q);
        return q;
    }
    ~DevTemps() {
        if (v.empty()) return;
        (void)hipStreamSynchronize(s);
        for (void* q : v) (void)hipFree(q);
    }
};

template <typename T, bool CV>
int sparse_build(mi355nufft_plan* p, const void* const* coef, const int32_t* const* kidx, const void* phase) {
    using V = tab<T, CV>;
    SparseState& s = p->sp;
    const int64_t M = p->M, nk = p->nk;
    int nnzr = 1;
    for (int d = 0; d < p->ndim; ++d) nnzr *= p->J[d];
    const int64_t nnz = M * nnzr;
    const int64_t i32max = ((int64_t)1 << 31) - 1;
    if (nnz > i32max || nk + nnz / kSpPiece + 2 > i32max)
        return fail(MI355NUFFT_EINVAL, "sparse matrix of %lld entries on %lld grid points exceeds the 2^31 index range",
                    (long long)nnz, (long long)nk);
    if (p->nframes > 1) return fail(MI355NUFFT_EINVAL, "sparse mode takes one frame");
    s.cvals = CV;
    s.nnzr = nnzr;
    s.M = M;
    s.nnz = nnz;
    auto alloc = [&](void** q, size_t bytes) { return hipMalloc(q, bytes ? bytes : 1) == hipSuccess; };
    if (!alloc(&s.val, (size_t)nnz * sizeof(V)) || !alloc((void**)&s.rbase, (size_t)M * p->ndim * 4) ||
        !alloc((void**)&s.perm, (size_t)(M ? M : 1) * 4) || !alloc(&s.tval, (size_t)nnz * sizeof(V)) ||
        !alloc((void**)&s.trow, (size_t)nnz * 4)) {
        sparse_free(p);
        return fail(MI355NUFFT_ENOMEM, "sparse matrix of %lld entries", (long long)nnz);
    }
    DevTemps tmp(p->stream);
    uint32_t* key[2] = {(uint32_t*)tmp.get((size_t)nnz * 4), (uint32_t*)tmp.get((size_t)nnz * 4)};
    uint32_t* pay[2] = {(uint32_t*)tmp.get((size_t)nnz * 4), (uint32_t*)tmp.get((size_t)nnz * 4)};
    int32_t* ptr = (int32_t*)tmp.get((size_t)(nk + 1) * 4);
    int32_t* cnt = (int32_t*)tmp.get((size_t)(nk + 1) * 4 * 3);
    int32_t* off = (int32_t*)tmp.get((size_t)(nk + 1) * 4 * 3);
    if (!key[0] || !key[1] || !pay[0] || !pay[1] || !ptr || !cnt || !off) {
        sparse_free(p);
        return fail(MI355NUFFT_ENOMEM, "sparse matrix build temporaries");
    }
    // row order: the bin order of the tiled kernels when present
    if (M > 0) {
        if (p->binned) {
            HIP_TRY(hipMemcpyAsync(s.perm, p->d_bin_i + 2 * p->bin_m_cap, (size_t)M * 4, hipMemcpyDeviceToDevice,
                                   p->stream));
        } else {
            hipLaunchKernelGGL(sp_iota_kernel, dim3(nblocks(M)), dim3(kBlock), 0, p->stream, s.perm, M);
            LAUNCH_CHECK();
        }
        const int J1 = p->ndim > 1 ? p->J[1] : 1, J2 = p->ndim > 2 ? p->J[2] : 1;
        hipLaunchKernelGGL((sp_fill_rows_kernel<T, CV>), dim3(nblocks(M)), dim3(kBlock), 0, p->stream, coef[0],
                           p->ndim > 1 ? coef[1] : nullptr, p->ndim > 2 ? coef[2] : nullptr, kidx[0],
                           p->ndim > 1 ? kidx[1] : nullptr, p->ndim > 2 ? kidx[2] : nullptr, (const double2*)phase,
                           s.perm, M, p->ndim, p->J[0], J1, J2, (int)p->K[0], (int)(p->ndim > 1 ? p->K[1] : 1),
                           (V*)s.val, s.rbase, key[0], pay[0]);
        LAUNCH_CHECK();
        int bits = 1;
        while (((int64_t)1 << bits) < nk) ++bits;
        size_t tb = 0;
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key[0], key[1], pay[0], pay[1], (int)nnz, 0, bits,
                                                   p->stream));
        void* ctmp = tmp.get(tb);
        if (!ctmp) {
            sparse_free(p);
            return fail(MI355NUFFT_ENOMEM, "sparse matrix sort storage");
        }
        HIP_TRY(hipcub::DeviceRadixSort::SortPairs(ctmp, tb, key[0], key[1], pay[0], pay[1], (int)nnz, 0, bits,
                                                   p->stream));
        hipLaunchKernelGGL((sp_fill_cols_kernel<T, CV>), dim3(nblocks(nnz)), dim3(kBlock), 0, p->stream, pay[1],
                           nnz, nnzr, M, s.perm, (const V*)s.val, (V*)s.tval, s.trow);
        LAUNCH_CHECK();
    }
    hipLaunchKernelGGL(sp_col_ptr_kernel, dim3(nblocks(nk + 1)), dim3(kBlock), 0, p->stream, key[1], nnz, nk, ptr);
    LAUNCH_CHECK();
    int32_t* npc = cnt;
    int32_t* nsl = cnt + (nk + 1);
    int32_t* nsp = cnt + 2 * (nk + 1);
    hipLaunchKernelGGL(sp_piece_count_kernel, dim3(nblocks(nk + 1)), dim3(kBlock), 0, p->stream, ptr, nk, npc, nsl,
                       nsp);
    LAUNCH_CHECK();
    size_t sb = 0;
    HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, sb, npc, off, (int)(nk + 1), p->stream));
    void* stmp = tmp.get(sb);
    if (!stmp) {
        sparse_free(p);
        return fail(MI355NUFFT_ENOMEM, "sparse matrix scan storage");
    }
    for (int q = 0; q < 3; ++q)
        HIP_TRY(hipcub::DeviceScan::ExclusiveSum(stmp, sb, cnt + q * (nk + 1), off + q * (nk + 1), (int)(nk + 1),
                                                 p->stream));
    int32_t tot[3];
    for (int q = 0; q < 3; ++q)
        HIP_TRY(hipMemcpyAsync(&tot[q], off + q * (nk + 1) + nk, 4, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    s.npieces = tot[0];
    s.nslots = tot[1];
    s.nsplit = tot[2];
    if (!alloc((void**)&s.pieces, (size_t)s.npieces * sizeof(SpPiece)) ||
        !alloc((void**)&s.splits, (size_t)s.nsplit * sizeof(SpSplit))) {
        sparse_free(p);
        return fail(MI355NUFFT_ENOMEM, "sparse matrix column pieces");
    }
    hipLaunchKernelGGL(sp_piece_emit_kernel, dim3(nblocks(nk)), dim3(kBlock), 0, p->stream, ptr, nk, off,
                       off + (nk + 1), off + 2 * (nk + 1), s.pieces, s.splits);
    LAUNCH_CHECK();
    HIP_TRY(hipStreamSynchronize(p->stream));
    s.on = 1;
    return MI355NUFFT_OK;
}

// Forward with the sparse matrix.  raw: y = P xk exactly (grid_only,
// _nufft.py:1376-1388); else phase_before and the ortho scale are applied in
// place on the grid first (:1371-1375) and y is multiplied by phase_after
// (:1391).
template <typename T>
int sparse_interp(mi355nufft_plan* p, const void* grid, void* y, int32_t nreps, bool raw) {
    const SparseState& s = p->sp;
    // (fft_phased: the forward's last FFT pass applied phase_before / ortho)
    if (!raw && !p->fft_phased && (p->phasing == 0 || p->ortho))
        RET_IF(phase_grid_impl<T>(p, const_cast<void*>(grid), nreps, p->phasing == 0, 0, p->ortho));
    if (s.M == 0) return MI355NUFFT_OK;
    const cx<T>* ph = (!raw && p->phasing == 0 && p->has_phase) ? (const cx<T>*)p->d_phase : nullptr;
    const int J1 = p->ndim > 1 ? p->J[1] : 1, J2 = p->ndim > 2 ? p->J[2] : 1;
    const int K1 = (int)(p->ndim > 1 ? p->K[1] : 1), K2 = (int)(p->ndim > 2 ? p->K[2] : 1);
    if (s.cvals)
        hipLaunchKernelGGL((sp_interp_kernel<T, true>), dim3(nblocks(s.M)), dim3(kBlock), 0, p->stream,
                           (const cx<T>*)s.val, s.rbase, s.perm, s.M, p->ndim, p->J[0], J1, J2, (int)p->K[0], K1,
                           K2, (const cx<T>*)grid, p->nk, nreps, ph, (cx<T>*)y);
    else
        hipLaunchKernelGGL((sp_interp_kernel<T, false>), dim3(nblocks(s.M)), dim3(kBlock), 0, p->stream,
                           (const T*)s.val, s.rbase, s.perm, s.M, p->ndim, p->J[0], J1, J2, (int)p->K[0], K1,
                           K2, (const cx<T>*)grid, p->nk, nreps, ph, (cx<T>*)y);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

// Adjoint with the sparse matrix; writes every grid point.  raw: grid =
// P^H y exactly (grid_only, :1505-1509); else y is multiplied by
// conj(phase_after) (:1497) and the grid by conj(phase_before) (:1519-1520).
// grow a plan-owned sparse-mode buffer (stream-ordered free)
int sparse_grow(mi355nufft_plan* p, void** buf, size_t* cap, size_t need, const char* what) {
    if (need <= *cap) return MI355NUFFT_OK;
    if (*buf) {
        HIP_TRY(hipStreamSynchronize(p->stream));
        (void)hipFree(*buf);
        *buf = nullptr;
        *cap = 0;
    }
    if (hipMalloc(buf, need ? need : 1) != hipSuccess) return fail(MI355NUFFT_ENOMEM, "%s", what);
    *cap = need;
    return MI355NUFFT_OK;
}

template <typename T>
int sparse_spread(mi355nufft_plan* p, const void* y, void* grid, int32_t nreps, bool raw) {
    SparseState& s = p->sp;
    RET_IF(sparse_grow(p, &s.part, &s.part_bytes, (size_t)s.nslots * (size_t)nreps * p->cx_size(),
                       "sparse partial sums"));
    if (!raw && p->phasing == 0 && p->has_phase && s.M > 0) {
        // y * conj(phase_after) once per sample
        RET_IF(sparse_grow(p, &s.ybuf, &s.ybuf_bytes, (size_t)s.M * (size_t)nreps * p->cx_size(),
                           "sparse sample buffer"));
        hipLaunchKernelGGL((sp_conj_phase_kernel<T>), dim3(nblocks(s.M), nreps), dim3(kBlock), 0, p->stream,
                           (const cx<T>*)y, (const cx<T>*)p->d_phase, s.M, (cx<T>*)s.ybuf);
        LAUNCH_CHECK();
        y = s.ybuf;
    }
    const cx<T>* ph = nullptr;
    // (fft_phased: the adjoint's first FFT pass applies conj(phase_before))
    const int with_pb = (!raw && p->phasing == 0 && !p->fft_phased) ? 1 : 0;
    const GridGeom g = grid_geom(p, nullptr);
    const T c0 = (T)p->phase_coef[0], c1 = (T)p->phase_coef[1], c2 = (T)p->phase_coef[2];
    const unsigned nb = (unsigned)((s.npieces * kSpLanes + kBlock - 1) / kBlock);
    if (s.cvals)
        hipLaunchKernelGGL((sp_spread_kernel<T, true>), dim3(nb), dim3(kBlock), 0, p->stream, s.pieces, s.npieces,
                           (const cx<T>*)s.tval, s.trow, (const cx<T>*)y, s.M, nreps, ph, g, c0, c1, c2, with_pb,
                           (cx<T>*)grid, (cx<T>*)s.part, s.nslots);
    else
        hipLaunchKernelGGL((sp_spread_kernel<T, false>), dim3(nb), dim3(kBlock), 0, p->stream, s.pieces, s.npieces,
                           (const T*)s.tval, s.trow, (const cx<T>*)y, s.M, nreps, ph, g, c0, c1, c2, with_pb,
                           (cx<T>*)grid, (cx<T>*)s.part, s.nslots);
    LAUNCH_CHECK();
    if (s.nsplit > 0) {
        hipLaunchKernelGGL((sp_combine_kernel<T>), dim3(nblocks(s.nsplit)), dim3(kBlock), 0, p->stream, s.splits,
                           s.nsplit, (const cx<T>*)s.part, s.nslots, nreps, g, c0, c1, c2, with_pb, (cx<T>*)grid);
        LAUNCH_CHECK();
    }
    return MI355NUFFT_OK;
}

template <typename T>
int interp_impl(mi355nufft_plan* p, const void* grid, void* y, int32_t nreps) {
    if (p->sp.on) return sparse_interp<T>(p, grid, y, nreps, false);
    // tiled path: phase_before and the ortho scale are applied while staging
    if (use_tiled(p)) return run_tiled_any<T>(p, false, nullptr, y, grid, nullptr, nreps);
    if (p->nframes > 1) return fail(MI355NUFFT_ESTATE, "frames need the tiled kernels");
    // global-memory path: phase_before (and 1/sqrt(prod K) if ortho) is
    // applied in place on the grid before the gather
    if (p->phasing == 0 || p->ortho)
        RET_IF(phase_grid_impl<T>(p, const_cast<void*>(grid), nreps, p->phasing == 0, 0, p->ortho));
    int32_t J[3] = {p->J[0], p->J[1], p->J[2]};
    return run_table1<T>(false, p->ndim, p->K, J, p->L, p->phasing, nullptr, y, grid, nullptr,
                         (const void* const*)p->d_tab, p->d_tm, p->M, nreps,
                         p->has_phase ? p->d_phase : nullptr, p->stream);
}

template <typename T>
int spread_two_phase(mi355nufft_plan* p, const void* y, void* grid, int32_t nreps);
bool use_two_phase(const mi355nufft_plan* p, int32_t nreps);

template <typename T>
int spread_impl(mi355nufft_plan* p, const void* y, void* grid, int32_t nreps, bool grid_is_zero = false) {
    if (p->sp.on) return sparse_spread<T>(p, y, grid, nreps, false);  // writes every grid point
    if (use_tiled(p) && use_two_phase(p, nreps)) return spread_two_phase<T>(p, y, grid, nreps);
    if (!grid_is_zero) RET_IF(zero_grid<T>(grid, (int64_t)p->nframes * nreps * p->nk, p->stream));
    if (use_tiled(p)) return run_tiled_any<T>(p, true, y, nullptr, nullptr, grid, nreps);
    if (p->nframes > 1) return fail(MI355NUFFT_ESTATE, "frames need the tiled kernels");
    int32_t J[3] = {p->J[0], p->J[1], p->J[2]};
    RET_IF(run_table1<T>(true, p->ndim, p->K, J, p->L, p->phasing, y, nullptr, nullptr, grid,
                         (const void* const*)p->d_tab, p->d_tm, p->M, nreps,
                         p->has_phase ? p->d_phase : nullptr, p->stream));
    if (p->phasing == 0) RET_IF(phase_grid_impl<T>(p, grid, nreps, 1, 1, 0));
    return MI355NUFFT_OK;
}

int fft_impl(mi355nufft_plan* p, void* grid, int32_t nreps, int direction) {
    FftPlans* f = nullptr;
    RET_IF(get_fft(p, nreps, &f));
    FFT_TRY(rocfft_execution_info_set_stream(p->fft_info, p->stream));
    if (p->fft_work_bytes) FFT_TRY(rocfft_execution_info_set_work_buffer(p->fft_info, p->fft_work, p->fft_work_bytes));
    void* bufs[1] = {grid};
    FFT_TRY(rocfft_execute(direction < 0 ? f->fwd : f->bwd, bufs, nullptr, p->fft_info));
    return MI355NUFFT_OK;
}

}  // namespace

namespace mi355host {

// Default bin geometry (autotuned on golden-angle radial and kooshball
// trajectories): 2-D 16x16 with P = 384 points per subproblem; 3-D 16^3
// with P = 3072 when the scatter tile is fixed point and no sparse matrix is
// set, else 16x8x8 with P = 1024; 1-D 512 with P = 512.  Requested sizes
// override.
void default_bins(const mi355nufft_plan* p, int B[3], int& P) {
    for (int d = 0; d < 3; ++d) B[d] = 1;
    if (p->ndim == 1) {
        B[0] = 512;
        P = 512;
    } else if (p->ndim == 2) {
        // P = 384 rather than 256: the Kd 512^2 golden-angle grid's 853
        // occupied bins then split into 1,001 subproblems, within the 1,024
        // 256-thread workgroups that 256 CUs hold at once, instead of 1,182
        // (scatter 18.5 -> 16.4 us, pair 55.8 -> 53.9 us; uniform random
        // points of the same count 55.5 -> 53.7 us)
        B[0] = 16;
        B[1] = 16;
        P = 384;
    } else if (fxp_eligible(p) && !p->sparse_bins) {
        // the fixed-point scatter tile of 21^3 cells takes 74 KB of LDS:
        // bigger bins cut the halo and the flush (measured at Kd 256^3:
        // scatter 0.59 -> 0.45 ms, gather unchanged).  P 3072: on the
        // kooshball, with the packed 64-bit tile cells, pair 0.82 ms at P
        // 2048, 0.81 at 2560, 0.79 at 3072
        B[0] = 16;
        B[1] = 16;
        B[2] = 16;
        P = 3072;
    } else {
        B[0] = 16;
        B[1] = 8;
        B[2] = 8;
        P = 1024;
    }
    for (int d = 0; d < p->ndim; ++d)
        if (p->bin_req[d] > 0) B[d] = p->bin_req[d];
    if (p->P_req > 0) P = p->P_req;
}

void setup_bin_geom(mi355nufft_plan* p) {
    BinGeom& bg = p->bg;
    int B[3], P;
    default_bins(p, B, P);
    bg.nbins = 1;
    bg.tile = 1;
    for (int d = 0; d < 3; ++d) {
        if (d < p->ndim) {
            bg.B[d] = B[d] < (int)p->K[d] ? B[d] : (int)p->K[d];
            bg.nb[d] = (int)((p->K[d] + bg.B[d] - 1) / bg.B[d]);
            bg.T[d] = bg.B[d] + p->J[d] - 1;
        } else {
            bg.B[d] = 1;
            bg.nb[d] = 1;
            bg.T[d] = 1;
        }
        bg.nbins *= bg.nb[d];
        bg.tile *= bg.T[d];
        bg.invT[d] = 1.0f / (float)bg.T[d];
    }
    bg.nbins_frame = bg.nbins;
    bg.nbins *= p->nframes;
    bg.P = P;
    p->bg_threads = p->threads_req > 0 ? p->threads_req : 256;
    // two-phase scatter: every grid point must be covered by at most 8
    // (bin, offset) pairs per axis, the capacity of tile_reduce_kernel
    p->reduce_ok = 1;
    for (int d = 0; d < 3; ++d) {
        p->reduce_j[d] = 1;
        if (d >= p->ndim) continue;
        const int B = bg.B[d], nb = bg.nb[d], T = bg.T[d], K = (int)p->K[d];
        int jm = (T - 1 + B - 1) / B + 2;
        if (jm > nb) jm = nb;
        p->reduce_j[d] = jm;
        std::vector<int> hits((size_t)K, 0);
        for (int b = 0; b < nb; ++b)
            for (int e = 0; e < T; ++e) ++hits[(size_t)(((int64_t)b * B + e) % K)];
        int mx = 0;
        for (int v : hits) mx = v > mx ? v : mx;
        if (mx > 8) p->reduce_ok = 0;
    }
}

}  // namespace mi355host

namespace {

template <typename Tp>
int grow(mi355nufft_plan* p, Tp** buf, int64_t& cap, int64_t need, int64_t per) {
    if (need <= cap && *buf) return MI355NUFFT_OK;
    if (*buf) {
        HIP_TRY(hipStreamSynchronize(p->stream));
        (void)hipFree(*buf);
        *buf = nullptr;
        cap = 0;
    }
    if (hipMalloc((void**)buf, (size_t)(need * per + 1) * sizeof(Tp)) != hipSuccess)
        return fail(MI355NUFFT_ENOMEM, "binning storage (%lld entries)", (long long)(need * per));
    cap = need;
    return MI355NUFFT_OK;
}

template <typename T>
int bin_points_impl(mi355nufft_plan* p) {
    p->binned = 0;
    if (p->method == 1 || p->M == 0) return MI355NUFFT_OK;
    setup_bin_geom(p);
    if (p->method == 4 && !p->reduce_ok)
        return fail(MI355NUFFT_EINVAL, "bins too small for the two-phase scatter");
    // the scatter's tile (complex float64 or int32 pairs) is never smaller
    // than the gather's (complex T), and both kernels share the binning
    if (tile_bytes(p) > p->lds_limit) {
        if (p->method >= 2 || p->nframes > 1)
            return fail(MI355NUFFT_EINVAL, "tile of %d points (%zu bytes) exceeds the LDS budget of %zu bytes",
                        p->bg.tile, tile_bytes(p), p->lds_limit);
        return MI355NUFFT_OK;
    }
    const int64_t M = p->M;
    RET_IF(grow(p, &p->d_bin_i, p->bin_m_cap, M, 3));
    if ((int64_t)p->bg.nbins * kBanks >= ((int64_t)1 << 31))
        return fail(MI355NUFFT_EINVAL, "too many bins (%d) for the bin sort", p->bg.nbins);
    RET_IF(grow(p, &p->d_bin_b, p->bin_b_cap, (int64_t)p->bg.nbins, 3 + kBanks));
    RET_IF(grow(p, &p->d_sub, p->bin_s_cap, max_subproblems(p), 3));
    if (M > p->sorted_cap) {
        if (p->d_tms) {
            HIP_TRY(hipStreamSynchronize(p->stream));
            (void)hipFree(p->d_tms);
            (void)hipFree(p->d_phs);
            p->d_tms = p->d_phs = nullptr;
            p->sorted_cap = 0;
        }
        if (hipMalloc(&p->d_tms, (size_t)M * p->ndim * p->real_size()) != hipSuccess ||
            hipMalloc(&p->d_phs, (size_t)M * p->cx_size()) != hipSuccess)
            return fail(MI355NUFFT_ENOMEM, "sorted points storage for M=%lld", (long long)M);
        p->sorted_cap = M;
    }
    RET_IF(launch_bin_sort<T>(p, SortGate{nullptr, 0}));
    p->binned = 1;
    return MI355NUFFT_OK;
}

int bin_points(mi355nufft_plan* p) { return p->precision ? bin_points_impl<double>(p) : bin_points_impl<float>(p); }


// Two-phase scatter (binning method 4, the deterministic mode): the tiled
// scatter writes each subproblem's tile to a slot, tile_reduce_kernel sums
// the slots covering each grid point in a fixed order and writes the grid
// once.  Measured slower than the atomic flush on MI355X (slot traffic), so
// the atomic flush is the default.
constexpr size_t kTileSlotBudget = (size_t)8 << 30;

size_t two_phase_bytes(const mi355nufft_plan* p, int32_t nreps) {
    return (size_t)max_subproblems(p) * (size_t)nreps * (size_t)p->bg.tile * p->cx_size();
}

bool use_two_phase(const mi355nufft_plan* p, int32_t nreps) {
    return p->method == 4 && p->reduce_ok && two_phase_bytes(p, nreps) <= kTileSlotBudget;
}

template <typename T>
int spread_two_phase(mi355nufft_plan* p, const void* y, void* grid, int32_t nreps) {
    const size_t need = two_phase_bytes(p, nreps);
    if (need > p->tiles_bytes) {
        if (p->d_tiles) {
            HIP_TRY(hipStreamSynchronize(p->stream));
            (void)hipFree(p->d_tiles);
            p->d_tiles = nullptr;
            p->tiles_bytes = 0;
        }
        if (hipMalloc(&p->d_tiles, need) != hipSuccess)
            return fail(MI355NUFFT_ENOMEM, "scatter tile slots of %zu bytes", need);
        p->tiles_bytes = need;
    }
    p->tiles_target = p->d_tiles;
    const int rc = run_tiled_any<T>(p, true, y, nullptr, nullptr, grid, nreps);
    p->tiles_target = nullptr;
    RET_IF(rc);
    const int nimg = p->nframes * nreps;
    const T c0 = (T)p->phase_coef[0], c1 = (T)p->phase_coef[1], c2 = (T)p->phase_coef[2];
    const int use_pb = p->phasing == 0;
#define MI355_RED(ND)                                                                                          \
    hipLaunchKernelGGL((tile_reduce_kernel<T, ND>), dim3(nblocks(p->nk), nimg), dim3(kBlock), 0, p->stream,     \
                       (const cx<T>*)p->d_tiles, (cx<T>*)grid, p->bg, bin_state(p), (int)p->K[0], (int)p->K[1],  \
                       (int)p->K[2], p->nk, nreps, c0, c1, c2, use_pb, p->reduce_j[0], p->reduce_j[1],          \
                       p->reduce_j[2])
    if (p->ndim == 1) MI355_RED(1);
    else if (p->ndim == 2) MI355_RED(2);
    else MI355_RED(3);
#undef MI355_RED
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

}  // namespace

// Toeplitz normal operator: A^H W A of a table1 NUFFT applied as one
// circulant convolution on a power-of-two grid K_d >= 2 N_d:
//   y = crop_N(IFFT_K(T * FFT_K(pad_K(x)))),  T = FFT_K(circ(psf)) / prod(K).
struct mi355nufft_toeplitz {
    int ndim = 0;
    int precision = 0;
    int64_t N[3] = {1, 1, 1};
    int64_t K[3] = {1, 1, 1};
    int64_t nk = 1;
    hipStream_t stream = nullptr;
    void* d_tw[3] = {nullptr, nullptr, nullptr};
    void* d_T = nullptr;      // spectrum of the embedded kernel, nk complex
    int have_T = 0;
    void* d_grid = nullptr;   // scratch, nimg * nk complex
    size_t grid_bytes = 0;
    size_t cx_size() const { return precision ? 2 * sizeof(double) : 2 * sizeof(float); }
};

namespace {

FftDev toeplitz_dev(const mi355nufft_toeplitz* tp, bool full) {
    FftDev f;
    f.ndim = tp->ndim;
    for (int d = 0; d < 3; ++d) {
        f.N[d] = full ? tp->K[d] : tp->N[d];
        f.K[d] = tp->K[d];
    }
    f.nk = tp->nk;
    f.tw = tp->d_tw;
    f.sn = nullptr;
    f.stream = tp->stream;
    return f;
}

int toeplitz_scratch(mi355nufft_toeplitz* tp, int32_t nimg) {
    const size_t need = (size_t)nimg * (size_t)tp->nk * tp->cx_size();
    if (need > tp->grid_bytes) {
        if (tp->d_grid) {
            HIP_TRY(hipStreamSynchronize(tp->stream));
            HIP_TRY(hipFree(tp->d_grid));
            tp->d_grid = nullptr;
            tp->grid_bytes = 0;
        }
        if (hipMalloc(&tp->d_grid, need) != hipSuccess)
            return fail(MI355NUFFT_ENOMEM, "Toeplitz grid of %zu bytes", need);
        tp->grid_bytes = need;
    }
    return MI355NUFFT_OK;
}

template <typename T>
int toeplitz_set_psf_impl(mi355nufft_toeplitz* tp, const void* psf, const int64_t* ps) {
    // embed into the scratch grid, then FFT it into d_T
    RET_IF(toeplitz_scratch(tp, 1));
    RET_IF(zero_grid<T>(tp->d_grid, tp->nk, tp->stream));
    int64_t count = 1;
    for (int d = 0; d < tp->ndim; ++d) count *= 2 * tp->N[d];
    const int64_t p0 = ps[0], p1 = tp->ndim > 1 ? ps[1] : 0, p2 = tp->ndim > 2 ? ps[2] : 0;
    hipLaunchKernelGGL((toeplitz_embed_kernel<T>), dim3(nblocks(count)), dim3(kBlock), 0, tp->stream,
                       (const cx<T>*)psf, (cx<T>*)tp->d_grid, tp->ndim, (int)tp->N[0], (int)tp->N[1], (int)tp->N[2],
                       (int)tp->K[0], (int)tp->K[1], (int)tp->K[2], p0, p1, p2, count, (T)(1.0 / (double)tp->nk));
    LAUNCH_CHECK();
    // full forward FFT of the embedded kernel
    const FftDev f = toeplitz_dev(tp, true);
    int64_t xs[4] = {1, tp->K[0], tp->K[0] * tp->K[1], tp->nk};
    if (tp->ndim == 1) xs[1] = tp->nk;
    if (tp->ndim == 2) xs[2] = tp->nk;
    RET_IF(fft_rows<T>(f, false, tp->d_grid, nullptr, xs, tp->d_T, 1, (T)1));
    for (int d = 1; d < tp->ndim; ++d) RET_IF(fft_cols<T>(f, false, d, tp->d_T, 1));
    tp->have_T = 1;
    return MI355NUFFT_OK;
}

template <typename T>
int toeplitz_apply_impl(mi355nufft_toeplitz* tp, const void* x, const int64_t* xs, void* y, const int64_t* ys,
                        int32_t nimg) {
    RET_IF(toeplitz_scratch(tp, nimg));
    const FftDev f = toeplitz_dev(tp, false);
    RET_IF(fft_rows<T>(f, false, x, nullptr, xs, tp->d_grid, nimg, (T)1));
    for (int d = 1; d < tp->ndim; ++d) RET_IF(fft_cols<T>(f, false, d, tp->d_grid, nimg));
    FftPassOpts o;
    o.mult = tp->d_T;  // on the first inverse pass, which reads the whole spectrum
    for (int d = tp->ndim - 1; d >= 1; --d) {
        RET_IF(fft_cols<T>(f, true, d, tp->d_grid, nimg, o));
        o.mult = nullptr;
    }
    return fft_rows<T>(f, true, nullptr, y, ys, tp->d_grid, nimg, (T)1, o);
}

// Whether mi355nufft_forward / _adjoint move phase_before (and the forward's
// ortho scale) into the fused FFT passes: power-of-two fused grids of two or
// three dimensions, with the sparse matrix or the tiled kernels (atomic
// flush for the adjoint).
bool fft_phase_forward(const mi355nufft_plan* p) {
    return p->ndim >= 2 && use_fused_fft(p) && (p->sp.on || use_tiled(p)) && (p->phasing == 0 || p->ortho);
}
bool fft_phase_adjoint(const mi355nufft_plan* p, int32_t nreps) {
    return p->ndim >= 2 && p->phasing == 0 && use_fused_fft(p) &&
           (p->sp.on || (use_tiled(p) && !use_two_phase(p, nreps)));
}

// Sets fft_phased for its scope.
struct FftPhased {
    mi355nufft_plan* p;
    FftPhased(mi355nufft_plan* q, bool on = true) : p(q) { p->fft_phased = on ? 1 : 0; }
    ~FftPhased() { p->fft_phased = 0; }
};

int need_points(const mi355nufft_plan* p) {
    if (!p->d_tm && p->M > 0) return fail(MI355NUFFT_ESTATE, "points not set");
    if (p->M == 0 && !p->d_tm) return fail(MI355NUFFT_ESTATE, "mi355nufft_set_points was not called");
    return MI355NUFFT_OK;
}


// ------------------------------------------------------- raw table1 entry points
// mi355nufft_table1_forward / _adjoint take the reference's unsorted points
// (tm, table pointers, grid) on every call, like the C kernels they replace.
// They run the tiled kernels on a bin sort of tm that is cached: a check
// kernel compares tm with the copy the cached sort was made from and stamps
// a per-call generation number when anything differs; the six sort launches
// are gated on that stamp, so an unchanged tm costs one compare pass and
// six empty launches, and a changed one (even in place, at the same address)
// is re-sorted within the same call.  The host never reads anything back.
// Geometries whose scatter tile does not fit the LDS budget run the
// global-memory kernels, which need no workspace.

struct RawKey {
    const void* tm = nullptr;
    int64_t M = -1;
    int ndim = 0;
    int precision = -1;
    int64_t K[3] = {0, 0, 0};
    int J[3] = {0, 0, 0};
    bool same(const RawKey& o) const {
        if (tm != o.tm || M != o.M || ndim != o.ndim || precision != o.precision) return false;
        for (int d = 0; d < 3; ++d)
            if (K[d] != o.K[d] || J[d] != o.J[d]) return false;
        return true;
    }
};

// Byte offsets of one workspace: gate word, copy of tm, bin arrays
// (bin_of, rank, perm), per-bin arrays, subproblem table, sorted tm.
struct RawLayout {
    size_t gate, tm_copy, bin_i, bin_b, sub, tms, total;
};

RawLayout raw_layout(int64_t M, int ndim, size_t rs, int64_t nbins, int64_t max_sub) {
    auto up = [](size_t v) { return (v + 255) & ~(size_t)255; };
    RawLayout L;
    size_t o = 0;
    L.gate = o;
    o = up(o + 256);
    L.tm_copy = o;
    o = up(o + (size_t)ndim * (size_t)M * rs);
    L.bin_i = o;
    o = up(o + (size_t)3 * (size_t)M * sizeof(int));
    L.bin_b = o;
    o = up(o + (size_t)(kBanks + 3) * (size_t)nbins * sizeof(int));
    L.sub = o;
    o = up(o + (size_t)(3 * max_sub + 1) * sizeof(int));
    L.tms = o;
    o = up(o + (size_t)ndim * (size_t)M * rs);
    L.total = o;
    return L;
}

// One workspace with the state of its cached sort.
struct RawWork {
    int device = -1;
    hipStream_t stream = nullptr;
    unsigned char* base = nullptr;
    size_t bytes = 0;
    RawKey key;            // points and geometry of the cached sort
    size_t key_total = 0;  // layout it was made in
    unsigned gen = 0;      // last generation number used
};

constexpr int kMaxDevices = 64;
// library-owned workspaces, one per (device, stream) in use, at most this
// many; beyond that the raw entry points use the global-memory kernels
constexpr size_t kMaxRawWorkspaces = 16;

std::mutex g_raw_mu;
std::vector<RawWork> g_raw_work;
int g_lds_limit[kMaxDevices] = {0};

size_t device_lds_limit(int dev) {
    if (dev < 0 || dev >= kMaxDevices) return 65536;
    if (!g_lds_limit[dev]) {
        int lds = 0;
        if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || lds <= 0)
            lds = 65536;
        g_lds_limit[dev] = lds;
    }
    return (size_t)g_lds_limit[dev];
}

// A plan describing the raw call: the caller's tables and tm, the default
// binning, no phases and no ortho scaling (phase_before / phase_after are
// applied by the reference's Python around these kernels).  Its binning
// storage points into the workspace (set by raw_bind).
void raw_plan(mi355nufft_plan& q, int ndim, const int64_t* Kd, const int32_t* Jd, int32_t L, int precision,
              int complex_kernel, const void* const* h, const void* tm, int64_t M, hipStream_t s, size_t lds) {
    q.ndim = ndim;
    for (int d = 0; d < ndim; ++d) {
        q.K[d] = Kd[d];
        q.J[d] = Jd[d];
        q.nk *= Kd[d];
        q.d_tab[d] = const_cast<void*>(h[d]);
    }
    q.L = L;
    q.precision = precision;
    q.phasing = complex_kernel ? 1 : 0;
    q.fft_phased = 1;  // no phase_before in the staging or the flush
    q.ortho = 0;
    q.stream = s;
    q.M = M;
    q.d_tm = const_cast<void*>(tm);
    q.has_phase = 0;
    q.nframes = 1;
    q.lds_limit = lds;
    q.exact_taps = 1;
    setup_bin_geom(&q);
}

RawLayout raw_bind(mi355nufft_plan& q, unsigned char* base) {
    const RawLayout L = raw_layout(q.M, q.ndim, q.real_size(), q.bg.nbins, max_subproblems(&q));
    q.d_bin_i = reinterpret_cast<int*>(base + L.bin_i);
    q.bin_m_cap = q.M;
    q.d_bin_b = reinterpret_cast<int*>(base + L.bin_b);
    q.bin_b_cap = q.bg.nbins;
    q.d_sub = reinterpret_cast<int*>(base + L.sub);
    q.bin_s_cap = max_subproblems(&q);
    q.d_tms = base + L.tms;
    q.d_phs = nullptr;
    q.binned = 1;
    return L;
}

// Enqueue the cached sort of q's points in workspace w (bound by raw_bind)
// and the tiled gather / scatter.
template <typename T>
int raw_tiled(mi355nufft_plan& q, RawWork& w, const RawLayout& L, bool adjoint, const void* fm_in, void* fm_out,
              const void* ck_in, void* ck_out, int32_t nreps) {
    RawKey key;
    key.tm = q.d_tm;
    key.M = q.M;
    key.ndim = q.ndim;
    key.precision = q.precision;
    for (int d = 0; d < q.ndim; ++d) {
        key.K[d] = q.K[d];
        key.J[d] = q.J[d];
    }
    const int force = (!w.key.same(key) || w.key_total != L.total) ? 1 : 0;
    w.gen = w.gen + 1 ? w.gen + 1 : 1;
    unsigned* gate = reinterpret_cast<unsigned*>(w.base + L.gate);
    const int64_t n = (int64_t)q.ndim * q.M;
    hipLaunchKernelGGL((tm_check_kernel<T>), dim3(nblocks(n)), dim3(kBlock), 0, q.stream, (const T*)q.d_tm,
                       (T*)(w.base + L.tm_copy), n, gate, w.gen, force);
    LAUNCH_CHECK();
    // a failed launch below leaves the key unset, so the next call re-sorts
    w.key = RawKey();
    RET_IF(launch_bin_sort<T>(&q, force ? SortGate{nullptr, 0} : SortGate{gate, w.gen}));
    w.key = key;
    w.key_total = L.total;
    return q.phasing ? run_tiled<T, true>(&q, adjoint, fm_in, fm_out, ck_in, ck_out, nreps)
                     : run_tiled<T, false>(&q, adjoint, fm_in, fm_out, ck_in, ck_out, nreps);
}

// The workspace of (device, stream), grown to `need` bytes; null when the
// table of workspaces is full or the allocation fails (callers then use the
// global-memory kernels).
RawWork* raw_workspace(int dev, hipStream_t s, size_t need) {
    RawWork* w = nullptr;
    for (auto& e : g_raw_work)
        if (e.device == dev && e.stream == s) w = &e;
    if (!w) {
        if (g_raw_work.size() >= kMaxRawWorkspaces) return nullptr;
        g_raw_work.emplace_back();
        w = &g_raw_work.back();
        w->device = dev;
        w->stream = s;
    }
    if (need > w->bytes) {
        if (w->base) {
            if (hipStreamSynchronize(s) != hipSuccess) return nullptr;
            (void)hipFree(w->base);
            w->base = nullptr;
            w->bytes = 0;
        }
        const size_t sz = need + need / 4;
        if (hipMalloc((void**)&w->base, sz) != hipSuccess) {
            w->base = nullptr;
            return nullptr;
        }
        w->bytes = sz;
        w->key = RawKey();  // fresh storage holds no sort
    }
    return w;
}

template <typename T>
int raw_table1(bool adjoint, int ndim, const int64_t* Kd, const int32_t* Jd, int32_t L, int complex_kernel,
               const void* fm_in, void* fm_out, const void* ck_in, void* ck_out, const void* const* h,
               const void* tm, int64_t M, int32_t nreps, hipStream_t s) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    if (M > 0) {
        mi355nufft_plan q;
        raw_plan(q, ndim, Kd, Jd, L, sizeof(T) == 8 ? 1 : 0, complex_kernel, h, tm, M, s, device_lds_limit(dev));
        if (tile_bytes(&q) <= q.lds_limit && (int64_t)q.bg.nbins * kBanks < ((int64_t)1 << 31)) {
            std::lock_guard<std::mutex> lock(g_raw_mu);
            const RawLayout need = raw_layout(M, ndim, sizeof(T), q.bg.nbins, max_subproblems(&q));
            RawWork* w = raw_workspace(dev, s, need.total);
            if (w) {
                const RawLayout Lb = raw_bind(q, w->base);
                return raw_tiled<T>(q, *w, Lb, adjoint, fm_in, fm_out, ck_in, ck_out, nreps);
            }
        }
    }
    return run_table1<T>(adjoint, ndim, Kd, Jd, L, complex_kernel, fm_in, fm_out, ck_in, ck_out, h, tm, M, nreps,
                         nullptr, s);
}

// ------------------------------------------------- no-allocation adjoint
// Scratch of mi355nufft_table1_adjoint_noalloc: a file-scope device array
// (each GPU holds its own copy, at its own address, for the life of the
// process).  Host code never uses its name as an address: the address on
// each device comes from hipGetSymbolAddress, taken once per device with
// that device current, and only that pointer reaches kernels.
constexpr size_t kNoallocScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

std::mutex g_noalloc_mu;
void* g_scratch_ptr[kMaxDevices] = {nullptr};
RawWork g_noalloc_work[kMaxDevices];

int noalloc_scratch(int dev, void** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail(MI355NUFFT_EINVAL, "device %d out of range", dev);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(MI355NUFFT_EHIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
        g_noalloc_work[dev].device = dev;
        g_noalloc_work[dev].base = (unsigned char*)ptr;
        g_noalloc_work[dev].bytes = kNoallocScratchBytes;
    }
    *out = g_scratch_ptr[dev];
    return MI355NUFFT_OK;
}

template <typename T>
int raw_adjoint_noalloc(int ndim, const int64_t* Kd, const int32_t* Jd, int32_t L, int complex_kernel,
                        const void* fm, void* ck, const void* const* h, const void* tm, int64_t M, int32_t nreps,
                        hipStream_t s) {
    int dev = 0;
    HIP_TRY(hipGetDevice(&dev));
    std::lock_guard<std::mutex> lock(g_noalloc_mu);
    void* base = nullptr;
    RET_IF(noalloc_scratch(dev, &base));
    int64_t nk = 1;
    for (int d = 0; d < ndim; ++d) nk *= Kd[d];
    mi355nufft_plan q;
    bool tiled = false;
    if (M > 0) {
        raw_plan(q, ndim, Kd, Jd, L, sizeof(T) == 8 ? 1 : 0, complex_kernel, h, tm, M, s, device_lds_limit(dev));
        tiled = tile_bytes(&q) <= q.lds_limit && (int64_t)q.bg.nbins * kBanks < ((int64_t)1 << 31);
    }
    if (tiled) {
        const RawLayout need = raw_layout(M, ndim, sizeof(T), q.bg.nbins, max_subproblems(&q));
        if (need.total > kNoallocScratchBytes)
            return fail(MI355NUFFT_EINVAL,
                        "mi355nufft_table1_adjoint_noalloc: %lld points need %zu bytes of scratch, more than the "
                        "%zu bytes it holds (use mi355nufft_table1_adjoint)",
                        (long long)M, need.total, kNoallocScratchBytes);
    }
    RET_IF(zero_grid<T>(ck, nk * nreps, s));  // the output starts from zero (template.c:336-337)
    if (M == 0) return MI355NUFFT_OK;
    if (!tiled)  // the global-memory scatter needs no scratch
        return run_table1<T>(true, ndim, Kd, Jd, L, complex_kernel, fm, nullptr, nullptr, ck, h, tm, M, nreps,
                             nullptr, s);
    RawWork& w = g_noalloc_work[dev];
    w.stream = s;
    const RawLayout Lb = raw_bind(q, (unsigned char*)base);
    return raw_tiled<T>(q, w, Lb, true, fm, nullptr, nullptr, ck, nreps);
}

}  // namespace

extern "C" {

const char* mi355nufft_last_error(void) { return g_err.c_str(); }

int mi355nufft_abi_version(void) { return MI355NUFFT_ABI_VERSION; }

int mi355nufft_plan_create(mi355nufft_plan** plan, int ndim, const int64_t* Nd, const int64_t* Kd,
                           const int32_t* Jd, int32_t L, int precision, int phasing, int ortho,
                           double adjoint_scale, const void* const* tables, const double* const* sn_factors,
                           const double* phase_coef, void* stream) {
    g_err.clear();
    if (!plan) return fail(MI355NUFFT_EINVAL, "plan output pointer is null");
    *plan = nullptr;
    RET_IF(check_geometry(ndim, Kd, Jd, L));
    if (!Nd || !tables || !sn_factors) return fail(MI355NUFFT_EINVAL, "Nd, tables and sn_factors are required");
    if (precision != 0 && precision != 1) return fail(MI355NUFFT_EINVAL, "precision must be 0 or 1");
    if (phasing != 0 && phasing != 1) return fail(MI355NUFFT_EINVAL, "phasing must be 0 or 1");
    if (phasing == 0 && !phase_coef) return fail(MI355NUFFT_EINVAL, "phase_coef is required for phasing 0");
    for (int d = 0; d < ndim; ++d) {
        if (Nd[d] < 1 || Nd[d] > Kd[d]) return fail(MI355NUFFT_EINVAL, "need 1 <= Nd[%d] <= Kd[%d]", d, d);
        if (!tables[d] || !sn_factors[d]) return fail(MI355NUFFT_EINVAL, "missing table or sn for axis %d", d);
    }
    mi355nufft_plan* p = new mi355nufft_plan();
    p->ndim = ndim;
    p->L = L;
    p->precision = precision;
    p->phasing = phasing;
    p->ortho = ortho ? 1 : 0;
    p->adjoint_scale = adjoint_scale;
    p->stream = (hipStream_t)stream;
    for (int d = 0; d < ndim; ++d) {
        p->N[d] = Nd[d];
        p->K[d] = Kd[d];
        p->J[d] = Jd[d];
        p->nk *= Kd[d];
        p->nn *= Nd[d];
        p->phase_coef[d] = phase_coef ? phase_coef[d] : 0.0;
    }
    p->scale_ortho = p->ortho ? std::sqrt((double)p->nk) : 1.0;
    {
        int dev = 0, lds = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && lds > 0)
            p->lds_limit = (size_t)lds;  // 160 KiB per workgroup on gfx950
    }
    // prod_d max |h_d|: bounds every coefficient product (each coefficient
    // interpolates two neighbouring table entries), for the fixed-point scatter
    p->tab_cmax = 1.0;
    for (int d = 0; d < ndim; ++d) {
        const int64_t n = (int64_t)Jd[d] * L + 1;
        double mx = 0.0;
        for (int64_t i = 0; i < n; ++i) {
            double v;
            if (phasing) {
                const double re = precision ? ((const double*)tables[d])[2 * i] : ((const float*)tables[d])[2 * i];
                const double im =
                    precision ? ((const double*)tables[d])[2 * i + 1] : ((const float*)tables[d])[2 * i + 1];
                v = std::sqrt(re * re + im * im);
            } else {
                v = std::fabs(precision ? ((const double*)tables[d])[i] : ((const float*)tables[d])[i]);
            }
            if (!(v <= mx)) mx = v;  // NaN propagates
        }
        p->tab_cmax *= mx;
    }
    if (!(p->tab_cmax > 0.0 && p->tab_cmax < 1e30)) p->tab_cmax = 0.0;  // unusable: float64 scatter tile
    for (int d = 0; d < ndim; ++d) {
        const size_t tab_bytes = (size_t)(Jd[d] * L + 1) * (phasing ? p->cx_size() : p->real_size());
        if (hipMalloc(&p->d_tab[d], tab_bytes) != hipSuccess ||
            hipMalloc((void**)&p->d_sn[d], (size_t)Nd[d] * sizeof(double)) != hipSuccess) {
            mi355nufft_plan_destroy(p);
            return fail(MI355NUFFT_ENOMEM, "table / sn allocation");
        }
        if (hipMemcpy(p->d_tab[d], tables[d], tab_bytes, hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(p->d_sn[d], sn_factors[d], (size_t)Nd[d] * sizeof(double), hipMemcpyHostToDevice) !=
                hipSuccess) {
            mi355nufft_plan_destroy(p);
            return fail(MI355NUFFT_EHIP, "table / sn upload");
        }
    }
    p->fft_fused_ok = fused_fft_sizes_ok(ndim, Kd, p->cx_size()) ? 1 : 0;
    if (p->fft_fused_ok) {
        const int rc = make_twiddles(ndim, Kd, precision, p->d_tw);
        if (rc) {
            mi355nufft_plan_destroy(p);
            return rc;
        }
    }
    *plan = p;
    return MI355NUFFT_OK;
}

int mi355nufft_plan_destroy(mi355nufft_plan* p) {
    if (!p) return MI355NUFFT_OK;
    (void)hipStreamSynchronize(p->stream);
    for (auto& f : p->ffts) {
        if (f.fwd) rocfft_plan_destroy(f.fwd);
        if (f.bwd) rocfft_plan_destroy(f.bwd);
    }
    if (p->fft_info) rocfft_execution_info_destroy(p->fft_info);
    for (int d = 0; d < 3; ++d) {
        if (p->d_tab[d]) (void)hipFree(p->d_tab[d]);
        if (p->d_sn[d]) (void)hipFree(p->d_sn[d]);
        if (p->d_tw[d]) (void)hipFree(p->d_tw[d]);
    }
    sparse_free(p);
    if (p->d_tm) (void)hipFree(p->d_tm);
    if (p->d_phase) (void)hipFree(p->d_phase);
    if (p->d_grid) (void)hipFree(p->d_grid);
    if (p->d_agrid) (void)hipFree(p->d_agrid);
    if (p->fft_work) (void)hipFree(p->fft_work);
    if (p->d_bin_i) (void)hipFree(p->d_bin_i);
    if (p->d_bin_b) (void)hipFree(p->d_bin_b);
    if (p->d_sub) (void)hipFree(p->d_sub);
    if (p->d_frame_off) (void)hipFree(p->d_frame_off);
    if (p->d_tiles) (void)hipFree(p->d_tiles);
    if (p->d_flush_count) (void)hipFree(p->d_flush_count);
    if (p->d_tms) (void)hipFree(p->d_tms);
    if (p->d_phs) (void)hipFree(p->d_phs);
    for (auto& v : p->tev)
        for (hipEvent_t e : v) (void)hipEventDestroy(e);
    delete p;
    return MI355NUFFT_OK;
}

int mi355nufft_set_stream(mi355nufft_plan* p, void* stream) {
    g_err.clear();
    RET_IF(check_plan(p));
    p->stream = (hipStream_t)stream;
    return MI355NUFFT_OK;
}

int mi355nufft_set_points(mi355nufft_plan* p, const void* omega, int omega_double, int64_t M, int64_t stride_m,
                          int64_t stride_d, const double* shift, int phase_mode, int dot_single) {
    return mi355nufft_set_points_frames(p, omega, omega_double, M, stride_m, stride_d, shift, phase_mode, dot_single,
                                        nullptr, 1);
}

int mi355nufft_set_points_frames(mi355nufft_plan* p, const void* omega, int omega_double, int64_t M,
                                 int64_t stride_m, int64_t stride_d, const double* shift, int phase_mode,
                                 int dot_single, const int64_t* frame_offsets, int32_t nframes) {
    g_err.clear();
    RET_IF(check_plan(p));
    if (nframes < 1 || nframes > 65535) return fail(MI355NUFFT_EINVAL, "nframes must be in 1..65535");
    if (nframes > 1) {
        if (!frame_offsets) return fail(MI355NUFFT_EINVAL, "frame_offsets is required for nframes > 1");
        if (frame_offsets[0] != 0 || frame_offsets[nframes] != M)
            return fail(MI355NUFFT_EINVAL, "frame_offsets must start at 0 and end at M");
        for (int f = 0; f < nframes; ++f)
            if (frame_offsets[f + 1] < frame_offsets[f])
                return fail(MI355NUFFT_EINVAL, "frame_offsets must be non-decreasing");
        if (p->method == 1) return fail(MI355NUFFT_EINVAL, "frames need the tiled kernels (binning method 0 or 2)");
        if (nframes + 1 > p->frame_cap) {
            if (p->d_frame_off) {
                HIP_TRY(hipStreamSynchronize(p->stream));
                (void)hipFree(p->d_frame_off);
                p->d_frame_off = nullptr;
            }
            if (hipMalloc((void**)&p->d_frame_off, (size_t)(nframes + 1) * sizeof(int64_t)) != hipSuccess)
                return fail(MI355NUFFT_ENOMEM, "frame offsets");
            p->frame_cap = nframes + 1;
        }
        HIP_TRY(hipStreamSynchronize(p->stream));
        HIP_TRY(hipMemcpy(p->d_frame_off, frame_offsets, (size_t)(nframes + 1) * sizeof(int64_t),
                          hipMemcpyHostToDevice));
    }
    sparse_free(p);  // the sparse matrix belongs to the previous points
    p->nframes = nframes;
    if (M < 0 || M >= ((int64_t)1 << 31)) return fail(MI355NUFFT_EINVAL, "M must be in [0, 2^31)");
    if (M > 0 && !omega) return fail(MI355NUFFT_EINVAL, "omega is null");
    if (phase_mode < 0 || phase_mode > 2) return fail(MI355NUFFT_EINVAL, "phase_mode must be 0, 1 or 2");
    if (phase_mode && !shift) return fail(MI355NUFFT_EINVAL, "shift is required for phase_mode %d", phase_mode);
    if (phase_mode == 1 && p->phasing != 0) return fail(MI355NUFFT_EINVAL, "phase_mode 1 needs phasing 0");
    if (phase_mode == 2 && p->phasing != 1) return fail(MI355NUFFT_EINVAL, "phase_mode 2 needs phasing 1");
    if (M > p->M_cap || !p->d_tm) {
        if (p->d_tm) {
            HIP_TRY(hipStreamSynchronize(p->stream));
            (void)hipFree(p->d_tm);
            p->d_tm = nullptr;
        }
        if (p->d_phase) {
            (void)hipFree(p->d_phase);
            p->d_phase = nullptr;
        }
        const int64_t cap = M > 0 ? M : 1;
        if (hipMalloc(&p->d_tm, (size_t)cap * p->ndim * p->real_size()) != hipSuccess ||
            hipMalloc(&p->d_phase, (size_t)cap * p->cx_size()) != hipSuccess)
            return fail(MI355NUFFT_ENOMEM, "points storage for M=%lld", (long long)M);
        p->M_cap = cap;
    }
    p->M = M;
    p->has_phase = phase_mode != 0;
    p->binned = 0;
    for (int d = 0; d < 3; ++d) p->shift[d] = (phase_mode && d < p->ndim) ? shift[d] : 0.0;
    if (M == 0) return MI355NUFFT_OK;
    double gam[3] = {0, 0, 0};
    for (int d = 0; d < p->ndim; ++d) gam[d] = 2.0 * M_PI / (double)p->K[d];
    double sh[3] = {0, 0, 0};
    if (shift)
        for (int d = 0; d < p->ndim; ++d) sh[d] = shift[d];
#define SP_LAUNCH(T, W)                                                                                     \
    hipLaunchKernelGGL((set_points_kernel<T, W>), dim3(nblocks(M)), dim3(kBlock), 0, p->stream,             \
                       (const W*)omega, M, stride_m, stride_d, p->ndim, gam[0], gam[1], gam[2], (T*)p->d_tm, \
                       (cx<T>*)p->d_phase, phase_mode, sh[0], sh[1], sh[2], dot_single)
    if (p->precision == 0) {
        if (omega_double) SP_LAUNCH(float, double);
        else SP_LAUNCH(float, float);
    } else {
        if (omega_double) SP_LAUNCH(double, double);
        else SP_LAUNCH(double, float);
    }
#undef SP_LAUNCH
    LAUNCH_CHECK();
    return bin_points(p);
}

int mi355nufft_set_binning(mi355nufft_plan* p, const int32_t* bin_size, int32_t max_points, int method) {
    return mi355nufft_set_tiling(p, bin_size, max_points, method, 0);
}

int mi355nufft_set_tiling(mi355nufft_plan* p, const int32_t* bin_size, int32_t max_points, int method,
                          int32_t threads) {
    g_err.clear();
    RET_IF(check_plan(p));
    if (threads != 0 && threads != 64 && threads != 128 && threads != 256)
        return fail(MI355NUFFT_EINVAL, "threads must be 0, 64, 128 or 256");
    p->threads_req = threads;
    if (method < 0 || method > 4) return fail(MI355NUFFT_EINVAL, "method must be 0..4");
    if (max_points < 0) return fail(MI355NUFFT_EINVAL, "max_points must be >= 0");
    for (int d = 0; d < p->ndim; ++d) {
        const int b = bin_size ? bin_size[d] : 0;
        if (b < 0 || b > 4096) return fail(MI355NUFFT_EINVAL, "bin_size[%d] must be in 0..4096", d);
        p->bin_req[d] = b;
    }
    p->P_req = max_points;
    p->method = method;
    if (p->d_tm && p->M > 0) return bin_points(p);
    if (method >= 2) {
        setup_bin_geom(p);
        if (tile_bytes(p) > p->lds_limit)
            return fail(MI355NUFFT_EINVAL, "tile of %d points exceeds the LDS budget", p->bg.tile);
        if (method == 4 && !p->reduce_ok)
            return fail(MI355NUFFT_EINVAL, "bins too small for the two-phase scatter");
    }
    return MI355NUFFT_OK;
}

int mi355nufft_set_timing(mi355nufft_plan* p, int32_t enable) {
    if (!p) return fail(MI355NUFFT_EINVAL, "null plan");
    if ((enable & 2) && !p->d_flush_count) {
        const size_t bytes = kFlushSlots * sizeof(unsigned long long);
        if (hipMalloc(&p->d_flush_count, bytes) != hipSuccess) {
            p->d_flush_count = nullptr;
            return fail(MI355NUFFT_ENOMEM, "flush counters");
        }
        HIP_TRY(hipMemsetAsync(p->d_flush_count, 0, bytes, p->stream));
    }
    p->timing = enable ? (enable & 3) | 1 : 0;
    return MI355NUFFT_OK;
}

int mi355nufft_read_flush_atomics(mi355nufft_plan* p, uint64_t* lane_ops) {
    if (!p || !lane_ops) return fail(MI355NUFFT_EINVAL, "null argument");
    *lane_ops = 0;
    if (!p->d_flush_count) return MI355NUFFT_OK;
    std::vector<unsigned long long> v(kFlushSlots, 0);
    const size_t bytes = kFlushSlots * sizeof(unsigned long long);
    HIP_TRY(hipMemcpyAsync(v.data(), p->d_flush_count, bytes, hipMemcpyDeviceToHost, p->stream));
    HIP_TRY(hipStreamSynchronize(p->stream));
    HIP_TRY(hipMemsetAsync(p->d_flush_count, 0, bytes, p->stream));
    uint64_t total = 0;
    for (unsigned long long c : v) total += c;
    *lane_ops = total;
    return MI355NUFFT_OK;
}

int mi355nufft_read_timing(mi355nufft_plan* p, int32_t which, int64_t* launches, double* total_ms) {
    if (!p || !launches || !total_ms) return fail(MI355NUFFT_EINVAL, "null argument");
    if (which != 0 && which != 1) return fail(MI355NUFFT_EINVAL, "which must be 0 (gather) or 1 (scatter)");
    RET_IF(timing_fold(p, which));
    *launches = p->tev_n[which];
    *total_ms = p->tev_ms[which];
    p->tev_n[which] = 0;
    p->tev_ms[which] = 0.0;
    return MI355NUFFT_OK;
}

int mi355nufft_plan_info(const mi355nufft_plan* p, int64_t* info, int32_t n) {
    g_err.clear();
    RET_IF(check_plan(p));
    if (!info || n < 1) return fail(MI355NUFFT_EINVAL, "info buffer required");
    const int64_t v[19] = {p->ndim, p->M, p->bg.B[0], p->bg.B[1], p->bg.B[2], p->bg.P, p->bg.nbins, p->bg.tile,
                           use_tiled(p) ? 1 : 0, p->binned ? max_subproblems(p) : 0, p->method, (int64_t)p->lds_limit,
                           p->nframes, (use_tiled(p) && use_two_phase(p, 1)) ? 1 : 0, p->bg_threads,
                           use_fused_fft(p) ? 1 : 0, p->sp.on ? p->sp.nnz : -1, p->sp.on ? p->sp.npieces : 0,
                           use_fxp(p) ? 1 : 0};
    for (int i = 0; i < n && i < 19; ++i) info[i] = v[i];
    return MI355NUFFT_OK;
}

int64_t mi355nufft_num_points(const mi355nufft_plan* p) { return p ? p->M : 0; }

int mi355nufft_deapod_pad(mi355nufft_plan* p, const void* x, const int64_t* xs, void* grid, int32_t nreps) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    if (!x || !xs || !grid) return fail(MI355NUFFT_EINVAL, "x, xs and grid are required");
    return p->precision ? deapod_pad_impl<double>(p, x, xs, grid, nreps)
                        : deapod_pad_impl<float>(p, x, xs, grid, nreps);
}

int mi355nufft_fft(mi355nufft_plan* p, void* grid, int32_t nreps, int direction) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    if (!grid) return fail(MI355NUFFT_EINVAL, "grid is null");
    if (direction != -1 && direction != 1) return fail(MI355NUFFT_EINVAL, "direction must be -1 or +1");
    return fft_impl(p, grid, nreps, direction);
}

int mi355nufft_interp(mi355nufft_plan* p, const void* grid, void* y, int32_t nreps) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    RET_IF(need_points(p));
    if (!grid || (!y && p->M > 0)) return fail(MI355NUFFT_EINVAL, "grid and y are required");
    return p->precision ? interp_impl<double>(p, grid, y, nreps) : interp_impl<float>(p, grid, y, nreps);
}

int mi355nufft_spread(mi355nufft_plan* p, const void* y, void* grid, int32_t nreps) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    RET_IF(need_points(p));
    if (!grid || (!y && p->M > 0)) return fail(MI355NUFFT_EINVAL, "grid and y are required");
    return p->precision ? spread_impl<double>(p, y, grid, nreps) : spread_impl<float>(p, y, grid, nreps);
}

int mi355nufft_interp_grad(mi355nufft_plan* p, const void* grid, const void* w, void* grad, int32_t nreps) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    RET_IF(need_points(p));
    if (p->sp.on) return fail(MI355NUFFT_ESTATE, "the trajectory gradient needs table mode (no sparse matrix)");
    if (p->M == 0) return MI355NUFFT_OK;
    if (!grid || !w || !grad) return fail(MI355NUFFT_EINVAL, "grid, w and grad are required");
    if (!use_tiled(p))
        return fail(MI355NUFFT_ESTATE, "the trajectory gradient runs on the tiled kernels (binning method 0 or 2)");
    if (p->precision)
        return p->phasing ? run_grad<double, true>(p, grid, w, grad, nreps) : run_grad<double, false>(p, grid, w, grad, nreps);
    return p->phasing ? run_grad<float, true>(p, grid, w, grad, nreps) : run_grad<float, false>(p, grid, w, grad, nreps);
}

int mi355nufft_crop_deapod(mi355nufft_plan* p, const void* grid, void* x, const int64_t* xs, int32_t nreps) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    if (!x || !xs || !grid) return fail(MI355NUFFT_EINVAL, "x, xs and grid are required");
    return p->precision ? crop_deapod_impl<double>(p, grid, x, xs, nreps)
                        : crop_deapod_impl<float>(p, grid, x, xs, nreps);
}

int mi355nufft_forward_grid(mi355nufft_plan* p, const void* x, const int64_t* xs, void* grid, int32_t nimg) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nimg));
    if (!x || !xs || !grid) return fail(MI355NUFFT_EINVAL, "x, xs and grid are required");
    return p->precision ? forward_grid_impl<double>(p, x, xs, grid, nimg)
                        : forward_grid_impl<float>(p, x, xs, grid, nimg);
}

int mi355nufft_adjoint_image(mi355nufft_plan* p, void* grid, void* x, const int64_t* xs, int32_t nimg) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nimg));
    if (!x || !xs || !grid) return fail(MI355NUFFT_EINVAL, "grid, x and xs are required");
    return p->precision ? adjoint_image_impl<double>(p, grid, x, xs, nimg, 0)
                        : adjoint_image_impl<float>(p, grid, x, xs, nimg, 0);
}

// Sparse-matrix rows follow the bin order of the sort made before they are
// built: with a matrix set, the default 3-D complex64 bins are the 16x8x8
// ones (measured faster for the row product at Kd 256^3 than the 16^3 bins
// of the fixed-point scatter tile).  Re-bins the points when that changes
// the geometry.
static int set_sparse_bins(mi355nufft_plan* p, int on) {
    if (p->sparse_bins == on) return MI355NUFFT_OK;
    int B0[3], P0, B1[3], P1;
    default_bins(p, B0, P0);
    p->sparse_bins = on;
    default_bins(p, B1, P1);
    const bool same = P0 == P1 && B0[0] == B1[0] && B0[1] == B1[1] && B0[2] == B1[2];
    if (same || !(p->d_tm && p->M > 0)) return MI355NUFFT_OK;
    return bin_points(p);
}

int mi355nufft_set_sparse(mi355nufft_plan* p, const void* const* coef, const int32_t* const* kidx,
                          const void* phase) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(need_points(p));
    sparse_free(p);
    if (p->M > 0) {
        if (!coef || !kidx) return fail(MI355NUFFT_EINVAL, "coef and kidx are required");
        for (int d = 0; d < p->ndim; ++d)
            if (!coef[d] || !kidx[d]) return fail(MI355NUFFT_EINVAL, "coef[%d] and kidx[%d] are required", d, d);
    }
    if (phase && p->phasing == 0) return fail(MI355NUFFT_EINVAL, "phase is for the complex kernel (phasing 1)");
    RET_IF(set_sparse_bins(p, 1));
    if (p->precision)
        return p->phasing ? sparse_build<double, true>(p, coef, kidx, phase)
                          : sparse_build<double, false>(p, coef, kidx, phase);
    return p->phasing ? sparse_build<float, true>(p, coef, kidx, phase) : sparse_build<float, false>(p, coef, kidx, phase);
}

int mi355nufft_clear_sparse(mi355nufft_plan* p) {
    g_err.clear();
    RET_IF(check_plan(p));
    sparse_free(p);
    return set_sparse_bins(p, 0);
}

int mi355nufft_sparse_apply(mi355nufft_plan* p, int adjoint, const void* in, void* out, int32_t nreps) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    if (!p->sp.on) return fail(MI355NUFFT_ESTATE, "no sparse matrix (mi355nufft_set_sparse)");
    if (!out || (!in && (adjoint ? p->M > 0 : true))) return fail(MI355NUFFT_EINVAL, "in and out are required");
    if (adjoint)
        return p->precision ? sparse_spread<double>(p, in, out, nreps, true) : sparse_spread<float>(p, in, out, nreps, true);
    return p->precision ? sparse_interp<double>(p, in, out, nreps, true) : sparse_interp<float>(p, in, out, nreps, true);
}

int mi355nufft_coil_expand(int precision, const void* x, const void* smaps, int64_t npix, int32_t nimg,
                           int32_t ncoils, void* out, void* stream) {
    g_err.clear();
    RET_IF(check_nreps(nimg));
    if (precision != 0 && precision != 1) return fail(MI355NUFFT_EINVAL, "precision must be 0 or 1");
    if (ncoils < 1 || (int64_t)nimg * ncoils > 65535)
        return fail(MI355NUFFT_EINVAL, "ncoils must be >= 1 with nimg * ncoils <= 65535");
    if (npix < 1 || npix >= ((int64_t)1 << 40)) return fail(MI355NUFFT_EINVAL, "npix out of range");
    if (!x || !smaps || !out) return fail(MI355NUFFT_EINVAL, "x, smaps and out are required");
    const dim3 grid(nblocks(npix), nimg);
    if (precision)
        hipLaunchKernelGGL(coil_expand_kernel<double>, grid, dim3(kBlock), 0, (hipStream_t)stream,
                           (const cx<double>*)x, (const cx<double>*)smaps, npix, ncoils, (cx<double>*)out);
    else
        hipLaunchKernelGGL(coil_expand_kernel<float>, grid, dim3(kBlock), 0, (hipStream_t)stream,
                           (const cx<float>*)x, (const cx<float>*)smaps, npix, ncoils, (cx<float>*)out);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

int mi355nufft_coil_combine(int precision, const void* img, const void* smaps, int64_t npix, int32_t nimg,
                            int32_t ncoils, void* x, void* stream) {
    g_err.clear();
    RET_IF(check_nreps(nimg));
    if (precision != 0 && precision != 1) return fail(MI355NUFFT_EINVAL, "precision must be 0 or 1");
    if (ncoils < 1 || (int64_t)nimg * ncoils > 65535)
        return fail(MI355NUFFT_EINVAL, "ncoils must be >= 1 with nimg * ncoils <= 65535");
    if (npix < 1 || npix >= ((int64_t)1 << 40)) return fail(MI355NUFFT_EINVAL, "npix out of range");
    if (!img || !smaps || !x) return fail(MI355NUFFT_EINVAL, "img, smaps and x are required");
    const dim3 grid(nblocks(npix), nimg);
    if (precision)
        hipLaunchKernelGGL(coil_combine_kernel<double>, grid, dim3(kBlock), 0, (hipStream_t)stream,
                           (const cx<double>*)img, (const cx<double>*)smaps, npix, ncoils, (cx<double>*)x);
    else
        hipLaunchKernelGGL(coil_combine_kernel<float>, grid, dim3(kBlock), 0, (hipStream_t)stream,
                           (const cx<float>*)img, (const cx<float>*)smaps, npix, ncoils, (cx<float>*)x);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

int mi355nufft_weight_samples(int precision, const void* y, const void* w, int64_t M, int32_t nreps, void* out,
                              void* stream) {
    g_err.clear();
    RET_IF(check_nreps(nreps));
    if (precision != 0 && precision != 1) return fail(MI355NUFFT_EINVAL, "precision must be 0 or 1");
    if (M < 0 || M >= ((int64_t)1 << 40)) return fail(MI355NUFFT_EINVAL, "M out of range");
    if (M == 0) return MI355NUFFT_OK;
    if (!y || !w || !out) return fail(MI355NUFFT_EINVAL, "y, w and out are required");
    const dim3 grid(nblocks(M), nreps);
    if (precision)
        hipLaunchKernelGGL(weight_samples_kernel<double>, grid, dim3(kBlock), 0, (hipStream_t)stream,
                           (const cx<double>*)y, (const double*)w, M, (cx<double>*)out);
    else
        hipLaunchKernelGGL(weight_samples_kernel<float>, grid, dim3(kBlock), 0, (hipStream_t)stream,
                           (const cx<float>*)y, (const float*)w, M, (cx<float>*)out);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

int mi355nufft_set_scatter(mi355nufft_plan* p, int accumulate) {
    g_err.clear();
    RET_IF(check_plan(p));
    if (accumulate < 0 || accumulate > 2)
        return fail(MI355NUFFT_EINVAL, "accumulate must be 0 (auto), 1 (float64) or 2 (32-bit fixed point)");
    p->scatter_acc = accumulate;
    // the tile size changes, and with it whether the tiled kernels fit
    if (p->d_tm && p->M > 0) return bin_points(p);
    return MI355NUFFT_OK;
}

int mi355nufft_measure_atomics(int kind, double* ops_per_s, void* stream) {
    g_err.clear();
    if (!ops_per_s) return fail(MI355NUFFT_EINVAL, "ops_per_s is null");
    if (kind < 0 || kind > 3) return fail(MI355NUFFT_EINVAL, "kind must be 0, 1, 2 or 3");
    hipStream_t s = (hipStream_t)stream;
    int dev = 0, ncu = 0;
    HIP_TRY(hipGetDevice(&dev));
    HIP_TRY(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    const unsigned nblk = (unsigned)(ncu > 0 ? 8 * ncu : 2048);
    const int iters = 2048;
    const long long gn = 64ll << 20;  // 256 MiB of float32
    const int greps = 4;
    const size_t bytes = kind == 2 ? (size_t)gn * sizeof(float) : (size_t)nblk * 32768;
    void* buf = nullptr;
    if (hipMalloc(&buf, bytes) != hipSuccess) return fail(MI355NUFFT_ENOMEM, "probe buffer of %zu bytes", bytes);
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = MI355NUFFT_OK;
    float ms = 0.f;
    auto launch = [&]() {
        if (kind == 0)
            hipLaunchKernelGGL(lds_atomic_probe_kernel<double>, dim3(nblk), dim3(256), 0, s, (double*)buf, iters);
        else if (kind == 1)
            hipLaunchKernelGGL(lds_atomic_probe_kernel<unsigned>, dim3(nblk), dim3(256), 0, s, (unsigned*)buf, iters);
        else if (kind == 3)
            hipLaunchKernelGGL(lds_atomic_probe_kernel<unsigned long long>, dim3(nblk), dim3(256), 0, s,
                               (unsigned long long*)buf, iters);
        else
            hipLaunchKernelGGL(global_atomic_probe_kernel, dim3(nblk), dim3(256), 0, s, (float*)buf, gn, greps);
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = fail(MI355NUFFT_EHIP, "event creation");
    } else {
        if (kind == 2 && hipMemsetAsync(buf, 0, bytes, s) != hipSuccess) rc = fail(MI355NUFFT_EHIP, "memset");
        if (!rc) {
            launch();  // warm-up
            if (hipGetLastError() != hipSuccess || hipEventRecord(e0, s) != hipSuccess) {
                rc = fail(MI355NUFFT_EHIP, "probe launch");
            } else {
                launch();
                if (hipGetLastError() != hipSuccess || hipEventRecord(e1, s) != hipSuccess ||
                    hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess)
                    rc = fail(MI355NUFFT_EHIP, "probe timing");
            }
        }
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipStreamSynchronize(s);
    (void)hipFree(buf);
    if (rc) return rc;
    // lane operations: LDS 64 lanes x iters per wave, 4 waves per workgroup
    const double ops = kind == 2 ? (double)gn * greps : (double)nblk * 256.0 * iters;
    *ops_per_s = ops / (ms * 1e-3);
    return MI355NUFFT_OK;
}

int mi355nufft_set_fft(mi355nufft_plan* p, int method) {
    g_err.clear();
    RET_IF(check_plan(p));
    if (method != 0 && method != 1) return fail(MI355NUFFT_EINVAL, "FFT method must be 0 (auto) or 1 (rocFFT)");
    p->fft_method = method;
    return MI355NUFFT_OK;
}

int mi355nufft_forward(mi355nufft_plan* p, const void* x, const int64_t* xs, void* y, int32_t nreps, void* grid) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    RET_IF(need_points(p));
    if (!x || !xs || (!y && p->M > 0)) return fail(MI355NUFFT_EINVAL, "x, xs and y are required");
    const int32_t nimg = p->nframes * nreps;
    RET_IF(check_nreps(nimg));
    RET_IF(get_grid(p, nimg, &grid));
    if (fft_phase_forward(p)) {
        // the last FFT pass applies phase_before / the ortho scale; the
        // gather stages the grid as it is
        RET_IF(p->precision ? forward_grid_impl<double>(p, x, xs, grid, nimg, 1)
                            : forward_grid_impl<float>(p, x, xs, grid, nimg, 1));
        FftPhased on(p);
        return p->precision ? interp_impl<double>(p, grid, y, nreps) : interp_impl<float>(p, grid, y, nreps);
    }
    RET_IF(mi355nufft_forward_grid(p, x, xs, grid, nimg));
    return mi355nufft_interp(p, grid, y, nreps);
}

int mi355nufft_adjoint(mi355nufft_plan* p, const void* y, void* x, const int64_t* xs, int32_t nreps, void* grid) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    RET_IF(need_points(p));
    if (!x || !xs || (!y && p->M > 0)) return fail(MI355NUFFT_EINVAL, "y, x and xs are required");
    const int32_t nimg = p->nframes * nreps;
    RET_IF(check_nreps(nimg));
    // phase: the first inverse FFT pass applies conj(phase_before) instead of
    // the scatter's flush
    const int phase = fft_phase_adjoint(p, nreps) ? 2 : 0;
    if (grid) {
        if (!phase) {
            RET_IF(mi355nufft_spread(p, y, grid, nreps));
            return mi355nufft_adjoint_image(p, grid, x, xs, nimg);
        }
        {
            FftPhased on(p);
            RET_IF(p->precision ? spread_impl<double>(p, y, grid, nreps) : spread_impl<float>(p, y, grid, nreps));
        }
        return p->precision ? adjoint_image_impl<double>(p, grid, x, xs, nimg, 0, phase)
                            : adjoint_image_impl<float>(p, grid, x, xs, nimg, 0, phase);
    }
    // Plan-owned adjoint grid, left all zero by every adjoint that completes
    // (the passes write zeros over what they read), so the scatter needs no
    // clearing pass when the flag says so.  A captured adjoint ends the same
    // way on every replay: it skips the clearing pass when the grid is zero
    // at capture, and capture leaves the flag as it was.
    RET_IF(get_agrid(p, nimg));
    const bool capturing = stream_capturing(p);
    const bool was_zero = p->agrid_zero != 0;
    const int rezero = 1;
    if (!capturing) p->agrid_zero = 0;
    RET_IF(need_points(p));
    if (!y && p->M > 0) return fail(MI355NUFFT_EINVAL, "y is required");
    int rc;
    {
        FftPhased on(p, phase != 0);
        rc = p->precision ? spread_impl<double>(p, y, p->d_agrid, nreps, was_zero)
                          : spread_impl<float>(p, y, p->d_agrid, nreps, was_zero);
    }
    if (rc == MI355NUFFT_OK)
        rc = p->precision ? adjoint_image_impl<double>(p, p->d_agrid, x, xs, nimg, rezero, phase)
                          : adjoint_image_impl<float>(p, p->d_agrid, x, xs, nimg, rezero, phase);
    if (rc != MI355NUFFT_OK) {
        // clear what a partial call left, for adjoints captured earlier
        if (!capturing)
            (void)(p->precision ? zero_grid<double>(p->d_agrid, (int64_t)p->nframes * nreps * p->nk, p->stream)
                                : zero_grid<float>(p->d_agrid, (int64_t)p->nframes * nreps * p->nk, p->stream));
        return rc;
    }
    if (!capturing) p->agrid_zero = rezero;
    return MI355NUFFT_OK;
}

int mi355nufft_table1_forward(int ndim, const int64_t* Kd, const int32_t* Jd, int32_t L, int precision,
                              int complex_kernel, const void* ck, const void* const* h, const void* tm, int64_t M,
                              void* fm, int32_t nreps, void* stream) {
    g_err.clear();
    RET_IF(check_geometry(ndim, Kd, Jd, L));
    RET_IF(check_nreps(nreps));
    if (M < 0 || M >= ((int64_t)1 << 31)) return fail(MI355NUFFT_EINVAL, "M must be in [0, 2^31)");
    if (!h) return fail(MI355NUFFT_EINVAL, "h is null");
    for (int d = 0; d < ndim; ++d)
        if (!h[d]) return fail(MI355NUFFT_EINVAL, "h[%d] is null", d);
    if (M > 0 && (!ck || !tm || !fm)) return fail(MI355NUFFT_EINVAL, "ck, tm and fm are required");
    hipStream_t s = (hipStream_t)stream;
    return precision ? raw_table1<double>(false, ndim, Kd, Jd, L, complex_kernel, nullptr, fm, ck, nullptr, h, tm,
                                          M, nreps, s)
                     : raw_table1<float>(false, ndim, Kd, Jd, L, complex_kernel, nullptr, fm, ck, nullptr, h, tm,
                                         M, nreps, s);
}

int mi355nufft_table1_adjoint(int ndim, const int64_t* Kd, const int32_t* Jd, int32_t L, int precision,
                              int complex_kernel, const void* fm, const void* const* h, const void* tm, int64_t M,
                              void* ck, int32_t nreps, void* stream) {
    g_err.clear();
    RET_IF(check_geometry(ndim, Kd, Jd, L));
    RET_IF(check_nreps(nreps));
    if (M < 0 || M >= ((int64_t)1 << 31)) return fail(MI355NUFFT_EINVAL, "M must be in [0, 2^31)");
    if (!h || !ck) return fail(MI355NUFFT_EINVAL, "h and ck are required");
    for (int d = 0; d < ndim; ++d)
        if (!h[d]) return fail(MI355NUFFT_EINVAL, "h[%d] is null", d);
    if (M > 0 && (!fm || !tm)) return fail(MI355NUFFT_EINVAL, "fm and tm are required");
    hipStream_t s = (hipStream_t)stream;
    int64_t nk = 1;
    for (int d = 0; d < ndim; ++d) nk *= Kd[d];
    if (precision) {
        RET_IF(zero_grid<double>(ck, nk * nreps, s));
        return raw_table1<double>(true, ndim, Kd, Jd, L, complex_kernel, fm, nullptr, nullptr, ck, h, tm, M, nreps,
                                  s);
    }
    RET_IF(zero_grid<float>(ck, nk * nreps, s));
    return raw_table1<float>(true, ndim, Kd, Jd, L, complex_kernel, fm, nullptr, nullptr, ck, h, tm, M, nreps, s);
}

int mi355nufft_table1_adjoint_noalloc(int ndim, const int64_t* Kd, const int32_t* Jd, int32_t L, int precision,
                                      int complex_kernel, const void* fm, const void* const* h, const void* tm,
                                      int64_t M, void* ck, int32_t nreps, void* stream) {
    g_err.clear();
    RET_IF(check_geometry(ndim, Kd, Jd, L));
    RET_IF(check_nreps(nreps));
    if (precision != 0 && precision != 1) return fail(MI355NUFFT_EINVAL, "precision must be 0 or 1");
    if (M < 0 || M >= ((int64_t)1 << 31)) return fail(MI355NUFFT_EINVAL, "M must be in [0, 2^31)");
    if (!h || !ck) return fail(MI355NUFFT_EINVAL, "h and ck are required");
    for (int d = 0; d < ndim; ++d)
        if (!h[d]) return fail(MI355NUFFT_EINVAL, "h[%d] is null", d);
    if (M > 0 && (!fm || !tm)) return fail(MI355NUFFT_EINVAL, "fm and tm are required");
    hipStream_t s = (hipStream_t)stream;
    return precision ? raw_adjoint_noalloc<double>(ndim, Kd, Jd, L, complex_kernel, fm, ck, h, tm, M, nreps, s)
                     : raw_adjoint_noalloc<float>(ndim, Kd, Jd, L, complex_kernel, fm, ck, h, tm, M, nreps, s);
}

int mi355nufft_toeplitz_create(mi355nufft_toeplitz** out, int ndim, const int64_t* Nd, int precision,
                               void* stream) {
    g_err.clear();
    if (!out) return fail(MI355NUFFT_EINVAL, "output pointer is null");
    *out = nullptr;
    if (ndim < 1 || ndim > 3) return fail(MI355NUFFT_EINVAL, "ndim must be 1, 2 or 3 (got %d)", ndim);
    if (!Nd) return fail(MI355NUFFT_EINVAL, "Nd is required");
    if (precision != 0 && precision != 1) return fail(MI355NUFFT_EINVAL, "precision must be 0 or 1");
    mi355nufft_toeplitz* tp = new mi355nufft_toeplitz();
    tp->ndim = ndim;
    tp->precision = precision;
    tp->stream = (hipStream_t)stream;
    for (int d = 0; d < ndim; ++d) {
        if (Nd[d] < 1) {
            delete tp;
            return fail(MI355NUFFT_EINVAL, "Nd[%d] must be positive", d);
        }
        int64_t k = kFftMinK;
        while (k < 2 * Nd[d]) k *= 2;
        tp->N[d] = Nd[d];
        tp->K[d] = k;
        tp->nk *= k;
    }
    if (!fused_fft_sizes_ok(ndim, tp->K, tp->cx_size())) {
        delete tp;
        return fail(MI355NUFFT_EINVAL, "circulant grid beyond the fused FFT sizes (2 N_d <= %d, complex128 %d)",
                    kFftMaxK, kFftMaxK / 2);
    }
    if (make_twiddles(ndim, tp->K, precision, tp->d_tw) ||
        hipMalloc(&tp->d_T, (size_t)tp->nk * tp->cx_size()) != hipSuccess) {
        for (int d = 0; d < 3; ++d)
            if (tp->d_tw[d]) (void)hipFree(tp->d_tw[d]);
        delete tp;
        return fail(MI355NUFFT_ENOMEM, "Toeplitz plan storage");
    }
    *out = tp;
    return MI355NUFFT_OK;
}

int mi355nufft_toeplitz_destroy(mi355nufft_toeplitz* tp) {
    if (!tp) return MI355NUFFT_OK;
    (void)hipStreamSynchronize(tp->stream);
    for (int d = 0; d < 3; ++d)
        if (tp->d_tw[d]) (void)hipFree(tp->d_tw[d]);
    if (tp->d_T) (void)hipFree(tp->d_T);
    if (tp->d_grid) (void)hipFree(tp->d_grid);
    delete tp;
    return MI355NUFFT_OK;
}

int mi355nufft_toeplitz_set_stream(mi355nufft_toeplitz* tp, void* stream) {
    g_err.clear();
    if (!tp) return fail(MI355NUFFT_EINVAL, "null Toeplitz plan");
    tp->stream = (hipStream_t)stream;
    return MI355NUFFT_OK;
}

int mi355nufft_toeplitz_grid(const mi355nufft_toeplitz* tp, int64_t* Kd) {
    g_err.clear();
    if (!tp || !Kd) return fail(MI355NUFFT_EINVAL, "null argument");
    for (int d = 0; d < tp->ndim; ++d) Kd[d] = tp->K[d];
    return MI355NUFFT_OK;
}

int mi355nufft_toeplitz_set_psf(mi355nufft_toeplitz* tp, const void* psf, const int64_t* ps) {
    g_err.clear();
    if (!tp) return fail(MI355NUFFT_EINVAL, "null Toeplitz plan");
    if (!psf || !ps) return fail(MI355NUFFT_EINVAL, "psf and ps are required");
    return tp->precision ? toeplitz_set_psf_impl<double>(tp, psf, ps) : toeplitz_set_psf_impl<float>(tp, psf, ps);
}

int mi355nufft_toeplitz_apply(mi355nufft_toeplitz* tp, const void* x, const int64_t* xs, void* y,
                              const int64_t* ys, int32_t nimg) {
    g_err.clear();
    if (!tp) return fail(MI355NUFFT_EINVAL, "null Toeplitz plan");
    RET_IF(check_nreps(nimg));
    if (!x || !xs || !y || !ys) return fail(MI355NUFFT_EINVAL, "x, xs, y and ys are required");
    if (!tp->have_T) return fail(MI355NUFFT_ESTATE, "mi355nufft_toeplitz_set_psf was not called");
    return tp->precision ? toeplitz_apply_impl<double>(tp, x, xs, y, ys, nimg)
                         : toeplitz_apply_impl<float>(tp, x, xs, y, ys, nimg);
}

}  // extern "C"
