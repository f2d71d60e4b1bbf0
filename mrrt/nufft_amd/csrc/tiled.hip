// Dispatch of the tiled gather, scatter and trajectory-gradient kernels
// (kernels_binned.hpp) for one value type and kernel kind, chosen at compile
// time: built once per (MI355_TILED_T, MI355_TILED_CK) in {float, double} x
// {real kernel 0, complex kernel 1}.
#include "internal.hpp"

#if !defined(MI355_TILED_T) || !defined(MI355_TILED_CK)
#error "build with -DMI355_TILED_T=float|double -DMI355_TILED_CK=0|1"
#endif

namespace mi355host {

template <typename T, bool CK, int NDIM, int JMAX, bool FIXED>
void launch_tiled(bool adjoint, const TiledArgs<T, CK>& a, unsigned nsub, unsigned threads, size_t lds,
                  hipStream_t s, hipEvent_t ev0, hipEvent_t ev1) {
    // coefficients cached across repetitions (and, with the fixed-point
    // tile, the sample read by the scale's sum): only when a subproblem fits
    // one point per thread (the cache costs registers even when unused:
    // 135 VGPRs against 110 for the 2-D J = 6 fixed-point scatter)
    const bool reps = (a.ia.nreps > 1 || a.cmax > 0.f) && a.bg.P <= (int)threads;
    // a.rc == 2 for the scatter: two repetitions' tiles per pass
    const bool pair = adjoint && a.rc == 2;
    auto spread = pair ? spread_tiled_kernel<T, NDIM, CK, JMAX, FIXED, false, false, 2>
                       : reps ? spread_tiled_kernel<T, NDIM, CK, JMAX, FIXED, false, true>
                              : spread_tiled_kernel<T, NDIM, CK, JMAX, FIXED, false, false>;
    if constexpr (sizeof(T) == 4)
        if (a.cmax > 0.f)
            spread = pair ? spread_tiled_kernel<T, NDIM, CK, JMAX, FIXED, true, false, 2>
                          : reps ? spread_tiled_kernel<T, NDIM, CK, JMAX, FIXED, true, true>
                                 : spread_tiled_kernel<T, NDIM, CK, JMAX, FIXED, true, false>;
    auto interp = a.rc > 1 ? interp_tiled_kernel<T, NDIM, CK, JMAX, FIXED, true>
                           : interp_tiled_kernel<T, NDIM, CK, JMAX, FIXED, false>;
    if constexpr (FIXED)  // fused tap sums for the J = 6 kernels unless the reference's roundings are asked for
        if (!a.exact)
            interp = a.rc > 1 ? interp_tiled_kernel<T, NDIM, CK, JMAX, FIXED, true, true>
                              : interp_tiled_kernel<T, NDIM, CK, JMAX, FIXED, false, true>;
    if (ev0) {
        if (adjoint)
            hipExtLaunchKernelGGL(spread, dim3(nsub), dim3(threads), (uint32_t)lds, s, ev0, ev1, 0u, a);
        else
            hipExtLaunchKernelGGL(interp, dim3(nsub), dim3(threads), (uint32_t)lds, s, ev0, ev1, 0u, a);
        return;
    }
    if (adjoint)
        hipLaunchKernelGGL(spread, dim3(nsub), dim3(threads), lds, s, a);
    else
        hipLaunchKernelGGL(interp, dim3(nsub), dim3(threads), lds, s, a);
}

template <typename T, bool CK>
int run_tiled(mi355nufft_plan* p, bool adjoint, const void* samples_in, void* samples_out, const void* grid_in,
              void* grid_out, int32_t nreps) {
    TiledArgs<T, CK> a;
    std::memset(&a, 0, sizeof(a));
    int32_t J[3] = {p->J[0], p->J[1], p->J[2]};
    fill_interp_args<T, CK>(a.ia, p->ndim, p->K, J, p->L, (const void* const*)p->d_tab, p->d_tm, p->M, nreps);
    a.ia.grid = (const cx<T>*)grid_in;
    a.ia.grid_out = (cx<T>*)grid_out;
    a.ia.in = (const cx<T>*)samples_in;
    a.ia.out = (cx<T>*)samples_out;
    a.ia.phase = p->has_phase ? (const cx<T>*)p->d_phase : nullptr;
    a.bg = p->bg;
    a.bs = bin_state(p);
    a.use_pb = p->phasing == 0 && !p->fft_phased;
    for (int d = 0; d < 3; ++d) a.pb[d] = (T)p->phase_coef[d];
    a.ortho = adjoint ? 0 : (p->ortho && !p->fft_phased);
    a.scale_ortho = (T)p->scale_ortho;
    a.tiles = adjoint ? (cx<T>*)p->tiles_target : nullptr;
    a.tms = (const T*)p->d_tms;
    a.phs = p->has_phase ? (const cx<T>*)p->d_phs : nullptr;
    a.cmax = (adjoint && use_fxp(p)) ? (float)p->tab_cmax : 0.f;  // > 0: fixed-point scatter tile
    a.exact = p->exact_taps;
    a.own_ok = 1;
    for (int d = 0; d < p->ndim; ++d)
        if (p->K[d] % p->bg.B[d] != 0) a.own_ok = 0;
    const unsigned nsub = tiled_grid(p);
    // the gather stages up to kMaxRepChunk repetitions' tiles at once
    // (coefficients are then computed once per point for all of them),
    // within the LDS budget; the scatter one or two at a time
    const size_t one = adjoint ? tile_bytes(p) : gather_tile_bytes(p);
    const size_t chunk_lds = p->lds_limit < kRepChunkLds ? p->lds_limit : kRepChunkLds;
    int rc = adjoint ? 1 : (int)(chunk_lds / one);
    rc = rc < kMaxRepChunk ? rc : kMaxRepChunk;
    rc = rc < nreps ? rc : nreps;
    a.rc = rc > 1 ? rc : 1;
    // the scatter takes two repetitions per pass when both tiles fit the
    // chunk budget and its points would not fit the coefficient cache
    if (adjoint && nreps >= 2 && kScatterPairs && 2 * one <= chunk_lds && p->bg.P > p->bg_threads) a.rc = 2;
    const size_t lds = one * (size_t)a.rc;
    const int jm = jmax_of(p->ndim, J);
    const int jf = fixed_j(p->ndim, J);
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    RET_IF(timing_events(p, adjoint ? 1 : 0, &ev0, &ev1));
    // counted with the timed launches when asked for (the count's final
    // atomics lengthen a short launch, so timing alone leaves it out)
    a.flush_count = (adjoint && ev0 && (p->timing & 2)) ? p->d_flush_count : nullptr;
    a.flush_slots = kFlushSlots;
#define MI355_FCASE(ND, JF)                                                    \
    if (p->ndim == ND && jf == JF) {                                           \
        launch_tiled<T, CK, ND, JF, true>(adjoint, a, nsub, p->bg_threads, lds, p->stream, ev0, ev1); \
        LAUNCH_CHECK();                                                        \
        return MI355NUFFT_OK;                                                  \
    }
    MI355_FCASE(1, 6) MI355_FCASE(2, 6) MI355_FCASE(3, 6)
#undef MI355_FCASE
#define MI355_TCASE(ND, JM)                                                  \
    if (p->ndim == ND && jm == JM) {                                         \
        launch_tiled<T, CK, ND, JM, false>(adjoint, a, nsub, p->bg_threads, lds, p->stream, ev0, ev1); \
        LAUNCH_CHECK();                                                      \
        return MI355NUFFT_OK;                                                \
    }
    MI355_TCASE(1, 4) MI355_TCASE(1, 8) MI355_TCASE(1, 16)
    MI355_TCASE(2, 4) MI355_TCASE(2, 8) MI355_TCASE(2, 16)
    MI355_TCASE(3, 4) MI355_TCASE(3, 8) MI355_TCASE(3, 16)
#undef MI355_TCASE
    return fail(MI355NUFFT_EINVAL, "unsupported ndim/J combination");
}

template <typename T, bool CK, int NDIM, int JMAX, bool FIXED>
void launch_grad(const TiledArgs<T, CK>& a, const GradArgs<T>& g, unsigned nsub, unsigned threads, size_t lds,
                 hipStream_t s) {
    if (a.rc > 1)
        hipLaunchKernelGGL((grad_tiled_kernel<T, NDIM, CK, JMAX, FIXED, true>), dim3(nsub), dim3(threads), lds, s, a,
                           g);
    else
        hipLaunchKernelGGL((grad_tiled_kernel<T, NDIM, CK, JMAX, FIXED, false>), dim3(nsub), dim3(threads), lds, s,
                           a, g);
}

template <typename T, bool CK>
int run_grad(mi355nufft_plan* p, const void* grid, const void* w, void* grad, int32_t nreps) {
    TiledArgs<T, CK> a;
    std::memset(&a, 0, sizeof(a));
    int32_t J[3] = {p->J[0], p->J[1], p->J[2]};
    fill_interp_args<T, CK>(a.ia, p->ndim, p->K, J, p->L, (const void* const*)p->d_tab, p->d_tm, p->M, nreps);
    a.ia.grid = (const cx<T>*)grid;
    a.ia.phase = p->has_phase ? (const cx<T>*)p->d_phase : nullptr;
    a.bg = p->bg;
    a.bs = bin_state(p);
    a.use_pb = p->phasing == 0;
    for (int d = 0; d < 3; ++d) a.pb[d] = (T)p->phase_coef[d];
    a.ortho = p->ortho;
    a.scale_ortho = (T)p->scale_ortho;
    a.tms = (const T*)p->d_tms;
    a.phs = p->has_phase ? (const cx<T>*)p->d_phs : nullptr;
    GradArgs<T> g;
    g.w = (const cx<T>*)w;
    g.grad = (T*)grad;
    for (int d = 0; d < 3; ++d) {
        g.dtdw[d] = (T)((double)p->K[d] / (2.0 * M_PI));
        g.shift[d] = (T)p->shift[d];
    }
    const unsigned nsub = tiled_grid(p);
    {  // repetitions staged per pass, as in the gather
        const size_t chunk_lds = p->lds_limit < kRepChunkLds ? p->lds_limit : kRepChunkLds;
        int rc = (int)(chunk_lds / gather_tile_bytes(p));
        rc = rc < kMaxRepChunk ? rc : kMaxRepChunk;
        rc = rc < nreps ? rc : nreps;
        a.rc = rc > 1 ? rc : 1;
    }
    const size_t lds = gather_tile_bytes(p) * (size_t)a.rc;
    const int jm = jmax_of(p->ndim, J);
    const int jf = fixed_j(p->ndim, J);
#define MI355_GFCASE(ND, JF)                                           \
    if (p->ndim == ND && jf == JF) {                                   \
        launch_grad<T, CK, ND, JF, true>(a, g, nsub, p->bg_threads, lds, p->stream);  \
        LAUNCH_CHECK();                                                \
        return MI355NUFFT_OK;                                          \
    }
    MI355_GFCASE(1, 6) MI355_GFCASE(2, 6) MI355_GFCASE(3, 6)
#undef MI355_GFCASE
#define MI355_GCASE(ND, JM)                                            \
    if (p->ndim == ND && jm == JM) {                                   \
        launch_grad<T, CK, ND, JM, false>(a, g, nsub, p->bg_threads, lds, p->stream); \
        LAUNCH_CHECK();                                                \
        return MI355NUFFT_OK;                                          \
    }
    MI355_GCASE(1, 4) MI355_GCASE(1, 8) MI355_GCASE(1, 16)
    MI355_GCASE(2, 4) MI355_GCASE(2, 8) MI355_GCASE(2, 16)
    MI355_GCASE(3, 4) MI355_GCASE(3, 8) MI355_GCASE(3, 16)
#undef MI355_GCASE
    return fail(MI355NUFFT_EINVAL, "unsupported ndim/J combination");
}

template int run_tiled<MI355_TILED_T, (bool)MI355_TILED_CK>(mi355nufft_plan*, bool, const void*, void*, const void*,
                                                             void*, int32_t);
template int run_grad<MI355_TILED_T, (bool)MI355_TILED_CK>(mi355nufft_plan*, const void*, const void*, void*, int32_t);

}  // namespace mi355host
