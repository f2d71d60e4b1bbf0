// C ABI of the MI355X table1 NUFFT (declarations and contracts in
// include/mi355nufft.h).  One plan owns the device copies of the tables and
// deapodization factors, the per-sample state (tm, phase), the rocFFT plans
// with their work buffer and an optional grid scratch buffer.  Everything is
// enqueued on the plan's stream.
#include <hip/hip_runtime.h>
#include <rocfft/rocfft.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../../include/mi355nufft.h"
#include "kernels_grid.hpp"
#include "kernels_interp.hpp"

using namespace mi355;

namespace {

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

#define HIP_TRY(expr)                                                                   \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess)                                                           \
            return fail(MI355NUFFT_EHIP, "%s: %s (%s:%d)", #expr, hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                            \
    } while (0)

#define LAUNCH_CHECK()                                                                     \
    do {                                                                                   \
        hipError_t e_ = hipGetLastError();                                                 \
        if (e_ != hipSuccess)                                                              \
            return fail(MI355NUFFT_EHIP, "kernel launch: %s (%s:%d)", hipGetErrorString(e_), \
                        __FILE__, __LINE__);                                               \
    } while (0)

constexpr int kBlock = 256;

inline unsigned nblocks(int64_t n) { return (unsigned)((n + kBlock - 1) / kBlock); }

struct FftPlans {
    int nreps;
    rocfft_plan fwd;
    rocfft_plan bwd;
};

}  // namespace

struct mi355nufft_plan {
    int ndim = 0;
    int64_t N[3] = {1, 1, 1};
    int64_t K[3] = {1, 1, 1};
    int J[3] = {1, 1, 1};
    int L = 0;
    int precision = 0;  // 0 float, 1 double
    int phasing = 0;    // 0 real kernel, 1 complex kernel
    int ortho = 0;
    double adjoint_scale = 1.0;
    double scale_ortho = 1.0;
    double phase_coef[3] = {0, 0, 0};
    int64_t nk = 1;  // prod K
    int64_t nn = 1;  // prod N
    hipStream_t stream = nullptr;

    void* d_tab[3] = {nullptr, nullptr, nullptr};  // full tables (J*L+1 entries)
    double* d_sn[3] = {nullptr, nullptr, nullptr};

    int64_t M = 0;
    int64_t M_cap = 0;
    void* d_tm = nullptr;     // (ndim, M) real
    void* d_phase = nullptr;  // (M) complex or null
    int has_phase = 0;

    void* d_grid = nullptr;  // plan-owned grid scratch
    size_t grid_bytes = 0;

    std::vector<FftPlans> ffts;
    rocfft_execution_info fft_info = nullptr;
    void* fft_work = nullptr;
    size_t fft_work_bytes = 0;

    size_t real_size() const { return precision ? sizeof(double) : sizeof(float); }
    size_t cx_size() const { return 2 * real_size(); }
};

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

template <typename T, bool CK>
void fill_interp_args(InterpArgs<T, CK>& a, int ndim, const int64_t* Kd, const int32_t* Jd, int L,
                      const void* const* h, const void* tm, int64_t M, int nreps) {
    std::memset(&a, 0, sizeof(a));
    a.M = M;
    a.L = L;
    a.nreps = nreps;
    int64_t nk = 1;
    for (int d = 0; d < 3; ++d) {
        if (d < ndim) {
            const int JL = Jd[d] * L;
            const int nc = JL / 2;
            a.ax[d].K = (int)Kd[d];
            a.ax[d].J = Jd[d];
            a.ax[d].nlo = -nc;
            a.ax[d].nhi = JL - 1 - nc;
            a.h[d] = reinterpret_cast<const tab<T, CK>*>(h[d]) + nc;
            nk *= Kd[d];
        } else {
            a.ax[d].K = 1;
            a.ax[d].J = 1;
            a.h[d] = nullptr;
        }
    }
    a.kstride = nk;
    a.tm = reinterpret_cast<const T*>(tm);
}

int jmax_of(int ndim, const int32_t* Jd) {
    int j = 0;
    for (int d = 0; d < ndim; ++d) j = Jd[d] > j ? Jd[d] : j;
    return j <= 4 ? 4 : (j <= 8 ? 8 : 16);
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

#define RET_IF(x)                 \
    do {                          \
        int rc_ = (x);            \
        if (rc_) return rc_;      \
    } while (0)

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
int crop_deapod_impl(mi355nufft_plan* p, const void* grid, void* x, const int64_t* xs, int32_t nreps) {
    GridGeom g = grid_geom(p, xs);
    double scale = p->adjoint_scale;
    if (p->ortho) scale = p->scale_ortho * p->adjoint_scale / (double)p->nk;
    hipLaunchKernelGGL((crop_deapod_kernel<T>), dim3(nblocks(p->nn), nreps), dim3(kBlock), 0, p->stream,
                       (const cx<T>*)grid, (cx<T>*)x, g, p->nn, p->d_sn[0],
                       p->ndim > 1 ? p->d_sn[1] : p->d_sn[0], p->ndim > 2 ? p->d_sn[2] : p->d_sn[0], (T)scale);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

template <typename T>
int interp_impl(mi355nufft_plan* p, const void* grid, void* y, int32_t nreps) {
    // phase_before (and 1/sqrt(prod K) if ortho) is applied in place on the
    // grid, which is the plan's FFT output, before the gather
    if (p->phasing == 0 || p->ortho)
        RET_IF(phase_grid_impl<T>(p, const_cast<void*>(grid), nreps, p->phasing == 0, 0, p->ortho));
    int32_t J[3] = {p->J[0], p->J[1], p->J[2]};
    return run_table1<T>(false, p->ndim, p->K, J, p->L, p->phasing, nullptr, y, grid, nullptr,
                         (const void* const*)p->d_tab, p->d_tm, p->M, nreps,
                         p->has_phase ? p->d_phase : nullptr, p->stream);
}

template <typename T>
int spread_impl(mi355nufft_plan* p, const void* y, void* grid, int32_t nreps) {
    RET_IF(zero_grid<T>(grid, (int64_t)nreps * p->nk, p->stream));
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

int need_points(const mi355nufft_plan* p) {
    if (!p->d_tm && p->M > 0) return fail(MI355NUFFT_ESTATE, "points not set");
    if (p->M == 0 && !p->d_tm) return fail(MI355NUFFT_ESTATE, "mi355nufft_set_points was not called");
    return MI355NUFFT_OK;
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
    }
    if (p->d_tm) (void)hipFree(p->d_tm);
    if (p->d_phase) (void)hipFree(p->d_phase);
    if (p->d_grid) (void)hipFree(p->d_grid);
    if (p->fft_work) (void)hipFree(p->fft_work);
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
    g_err.clear();
    RET_IF(check_plan(p));
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

int mi355nufft_crop_deapod(mi355nufft_plan* p, const void* grid, void* x, const int64_t* xs, int32_t nreps) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    if (!x || !xs || !grid) return fail(MI355NUFFT_EINVAL, "x, xs and grid are required");
    return p->precision ? crop_deapod_impl<double>(p, grid, x, xs, nreps)
                        : crop_deapod_impl<float>(p, grid, x, xs, nreps);
}

int mi355nufft_forward(mi355nufft_plan* p, const void* x, const int64_t* xs, void* y, int32_t nreps, void* grid) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    RET_IF(need_points(p));
    if (!x || !xs || (!y && p->M > 0)) return fail(MI355NUFFT_EINVAL, "x, xs and y are required");
    RET_IF(get_grid(p, nreps, &grid));
    RET_IF(mi355nufft_deapod_pad(p, x, xs, grid, nreps));
    RET_IF(mi355nufft_fft(p, grid, nreps, -1));
    return mi355nufft_interp(p, grid, y, nreps);
}

int mi355nufft_adjoint(mi355nufft_plan* p, const void* y, void* x, const int64_t* xs, int32_t nreps, void* grid) {
    g_err.clear();
    RET_IF(check_plan(p));
    RET_IF(check_nreps(nreps));
    RET_IF(need_points(p));
    if (!x || !xs || (!y && p->M > 0)) return fail(MI355NUFFT_EINVAL, "y, x and xs are required");
    RET_IF(get_grid(p, nreps, &grid));
    RET_IF(mi355nufft_spread(p, y, grid, nreps));
    RET_IF(mi355nufft_fft(p, grid, nreps, 1));
    return mi355nufft_crop_deapod(p, grid, x, xs, nreps);
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
    return precision ? run_table1<double>(false, ndim, Kd, Jd, L, complex_kernel, nullptr, fm, ck, nullptr, h, tm,
                                          M, nreps, nullptr, s)
                     : run_table1<float>(false, ndim, Kd, Jd, L, complex_kernel, nullptr, fm, ck, nullptr, h, tm,
                                         M, nreps, nullptr, s);
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
        return run_table1<double>(true, ndim, Kd, Jd, L, complex_kernel, fm, nullptr, nullptr, ck, h, tm, M, nreps,
                                  nullptr, s);
    }
    RET_IF(zero_grid<float>(ck, nk * nreps, s));
    return run_table1<float>(true, ndim, Kd, Jd, L, complex_kernel, fm, nullptr, nullptr, ck, h, tm, M, nreps,
                             nullptr, s);
}

}  // extern "C"
