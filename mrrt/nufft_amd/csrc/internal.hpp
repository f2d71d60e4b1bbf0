// Internal declarations shared by the translation units of libmi355nufft.so:
// error reporting, the plan, the binning helpers, and the tiled-kernel and
// FFT-pass dispatchers (defined in tiled.hip and fft.hip, explicitly
// instantiated there, so the kernel templates compile in parallel).
#pragma once

#include <hip/hip_ext.h>
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
#include "kernels_binned.hpp"
#include "kernels_fft.hpp"
#include "kernels_sparse.hpp"

namespace mi355host {

using namespace mi355;

// Message of the last failing call on this thread; fail() sets it and
// returns code.
extern thread_local std::string g_err;
int fail(int code, const char* fmt, ...);

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

#define RET_IF(x)                 \
    do {                          \
        int rc_ = (x);            \
        if (rc_) return rc_;      \
    } while (0)

struct FftPlans {
    int nreps;
    rocfft_plan fwd;
    rocfft_plan bwd;
};

// Sparse-mode matrix (kernels_sparse.hpp): row form for the forward, column
// pieces for the adjoint.
struct SparseState {
    int on = 0;
    int cvals = 0;  // complex values (complex kernel)
    int nnzr = 0;   // entries per row, prod J
    int64_t M = 0, nnz = 0, npieces = 0, nsplit = 0, nslots = 0;
    void* val = nullptr;       // (nnzr, M) real or complex, sorted rows
    int32_t* rbase = nullptr;  // (ndim, M) first grid index per axis of each sorted row
    int32_t* perm = nullptr;   // sorted row -> sample
    void* tval = nullptr;      // (nnz) values in column order
    int32_t* trow = nullptr;   // (nnz) sample of each column entry
    SpPiece* pieces = nullptr;
    SpSplit* splits = nullptr;
    void* part = nullptr;      // partial sums of split columns (nslots x nreps)
    size_t part_bytes = 0;
    void* ybuf = nullptr;      // samples times conj(phase_after) (M x nreps)
    size_t ybuf_bytes = 0;
};

}  // namespace mi355host

using namespace mi355;
using namespace mi355host;

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
    void* d_tw[3] = {nullptr, nullptr, nullptr};   // FFT twiddles exp(-2 pi i m / K_d), complex T
    int fft_fused_ok = 0;  // every K_d a power of two in [kFftMinK, kFftMaxK]
    int fft_method = 0;    // 0 auto (fused pruned FFT when possible), 1 rocFFT
    // set while mi355nufft_forward / _adjoint run with phase_before (and the
    // forward's ortho scale) applied by the fused FFT passes instead of the
    // tiled gather's staging / scatter's flush
    int fft_phased = 0;
    double* d_sn[3] = {nullptr, nullptr, nullptr};

    int64_t M = 0;
    int64_t M_cap = 0;
    void* d_tm = nullptr;     // (ndim, M) real
    void* d_phase = nullptr;  // (M) complex or null
    int has_phase = 0;
    double shift[3] = {0, 0, 0};  // exponent coefficients of the per-sample phase

    void* d_grid = nullptr;  // plan-owned grid scratch (forward)
    size_t grid_bytes = 0;
    // plan-owned adjoint grid; the fused adjoint FFT passes write zeros over
    // everything they read, so after an adjoint through it the buffer is all
    // zero again and the next scatter needs no separate clear
    void* d_agrid = nullptr;
    size_t agrid_bytes = 0;
    int agrid_zero = 0;  // d_agrid holds zeros (host-side state, stream order)

    // binning (tiled kernels)
    int method = 0;           // 0 auto, 1 global-memory, 2 tiled, 3 tiled + atomic flush, 4 tiled + two-phase
    int reduce_ok = 0;        // two-phase scatter possible for the current geometry
    int reduce_j[3] = {1, 1, 1};  // candidate bins per axis of the reduce kernel
    void* d_tiles = nullptr;  // two-phase scatter tile slots
    size_t tiles_bytes = 0;
    void* tiles_target = nullptr;  // set while a two-phase scatter runs
    // scatter accumulation: 0 auto (32-bit fixed point for 2-D and 3-D
    // complex64 when the table bound is known, else float64), 1 float64,
    // 2 32-bit fixed point for complex64
    int scatter_acc = 0;
    // prod_d max |h_d| over the plan's tables (bound of every coefficient
    // product); 0 when unknown (raw entry points: caller tables)
    double tab_cmax = 0.0;
    // 1: the gather's tap sums keep the reference's roundings (raw entry
    // points); 0: fused multiply-adds for the J = 6 kernels
    int exact_taps = 0;
    // 1 while a sparse matrix is set: its rows follow the bin order, and
    // the default bins are then the float64-tile ones
    int sparse_bins = 0;
    int bin_req[3] = {0, 0, 0};  // requested bin size per axis (0: default)
    int P_req = 0;               // requested points per subproblem (0: default)
    int threads_req = 0;         // requested threads per tiled workgroup (0: default)
    BinGeom bg{};
    int bg_threads = 256;     // threads per tiled workgroup
    int binned = 0;           // bin state valid for the current points
    int* d_bin_i = nullptr;   // one allocation: bin_of, rank, perm (3*M_cap)
    int* d_bin_b = nullptr;   // counts, bin_start, bin_sub, bank-key counts ((3 + kBanks) * nbins_cap)
    int* d_sub = nullptr;     // sub_bin, sub_start, sub_end (3*sub_cap) + n_sub
    void* d_tms = nullptr;    // tm in sorted order (ndim * M_cap real)
    void* d_phs = nullptr;    // per-sample phase in sorted order (M_cap complex)
    int64_t sorted_cap = 0;
    int64_t bin_m_cap = 0, bin_b_cap = 0, bin_s_cap = 0;
    size_t lds_limit = 65536;
    int nframes = 1;              // frames with their own points (grids per frame: nreps)
    int64_t* d_frame_off = nullptr;
    int64_t frame_cap = 0;

    // launch timing of the tiled gather (0) and scatter (1) kernels; bit 1:
    // also count the scatter's flush atomics
    int timing = 0;
    std::vector<hipEvent_t> tev[2];  // start/stop pairs, reused
    int tev_used[2] = {0, 0};
    double tev_ms[2] = {0, 0};       // folded totals of earlier pairs
    int64_t tev_n[2] = {0, 0};
    // global atomic lane operations of the scatter's tile flushes while
    // timing is on (mi355nufft_read_flush_atomics), kFlushSlots partial
    // counts summed on the host
    unsigned long long* d_flush_count = nullptr;

    SparseState sp;

    std::vector<FftPlans> ffts;
    rocfft_execution_info fft_info = nullptr;
    void* fft_work = nullptr;
    size_t fft_work_bytes = 0;

    size_t real_size() const { return precision ? sizeof(double) : sizeof(float); }
    size_t cx_size() const { return 2 * real_size(); }
};

namespace mi355host {

// Partial counters of the flush atomic count: workgroup b adds to slot
// b mod kFlushSlots, so the counting adds no same-address contention.
constexpr int kFlushSlots = 1024;

// Whether the scatter tile accumulates complex64 in 32-bit fixed point (see
// spread_tiled_kernel): by default in 2-D and 3-D, or on request; it needs
// the plan's table bound.  Measured: Kd 256^3 scatter 26% faster (about
// 8e-7 from the float64 tile), Kd 512^2 6% faster (about 3e-7).  Else the
// tile accumulates in float64.
inline bool fxp_eligible(const mi355nufft_plan* p) {
    if (p->precision != 0 || !(p->tab_cmax > 0.0 && p->tab_cmax < 1e30) || p->scatter_acc == 1) return false;
    return p->scatter_acc == 2 || p->ndim >= 2;
}
inline bool use_fxp(const mi355nufft_plan* p) {
    // the int32 tile's bound holds up to 2^20 points per subproblem
    return fxp_eligible(p) && p->bg.P <= (1 << 20);
}
inline size_t tile_bytes(const mi355nufft_plan* p) {
    return (size_t)p->bg.tile * 2 * (use_fxp(p) ? sizeof(int) : sizeof(double));
}
inline size_t gather_tile_bytes(const mi355nufft_plan* p) { return (size_t)p->bg.tile * p->cx_size(); }

inline bool use_tiled(const mi355nufft_plan* p) { return p->method != 1 && p->binned; }

inline int64_t max_subproblems(const mi355nufft_plan* p) { return p->bg.nbins + (p->M + p->bg.P - 1) / p->bg.P; }

// Workgroups of a persistent tiled launch: one per subproblem slot, at most
// 2048 (8 per CU on 256 CUs); workgroups stride over the subproblem list.
inline unsigned tiled_grid(const mi355nufft_plan* p) {
    const int64_t n = max_subproblems(p);
    const int64_t cap = (int64_t)2048 * 256 / p->bg_threads;
    return (unsigned)(n < cap ? n : cap);
}

inline BinState bin_state(const mi355nufft_plan* p) {
    BinState bs;
    const int64_t mc = p->bin_m_cap, bc = p->bin_b_cap, sc = p->bin_s_cap;
    bs.bin_of = p->d_bin_i;
    bs.rank = p->d_bin_i + mc;
    bs.perm = p->d_bin_i + 2 * mc;
    bs.bkt = p->d_bin_b;  // first: its rows are read as int4
    bs.counts = p->d_bin_b + kBanks * bc;
    bs.bin_start = bs.counts + bc;
    bs.bin_sub = bs.counts + 2 * bc;
    bs.sub_bin = p->d_sub;
    bs.sub_start = p->d_sub + sc;
    bs.sub_end = p->d_sub + 2 * sc;
    bs.n_sub = p->d_sub + 3 * sc;
    bs.frame_off = p->nframes > 1 ? p->d_frame_off : nullptr;
    bs.nframes = p->nframes;
    return bs;
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

inline int jmax_of(int ndim, const int32_t* Jd) {
    int j = 0;
    for (int d = 0; d < ndim; ++d) j = Jd[d] > j ? Jd[d] : j;
    return j <= 4 ? 4 : (j <= 8 ? 8 : 16);
}

// Repetitions (coils) whose tiles one tiled launch stages together, and the
// LDS they may take (more per workgroup costs occupancy: at 24 frames x 32
// coils of 512^2 grids the gather took 1.18 ms with 64 KB, 0.81 ms with
// 40 KB and 1.22 ms with 16 KB per workgroup at 32x32 bins).
constexpr int kMaxRepChunk = 8;
// The scatter accumulates two repetitions per pass (two tiles, one barrier
// round and one coefficient computation per point for both) when they fit
// kRepChunkLds and a subproblem exceeds the one-point-per-thread coefficient
// cache: 24 frames x 32 coils at 64x32 bins, scatter 1.99 -> 1.79 ms.
constexpr bool kScatterPairs = true;
constexpr size_t kRepChunkLds = 40960;

// Taps compiled as fixed trip counts when every axis uses the same J among
// these; other widths run the generic kernels (J <= JMAX checked per tap).
inline int fixed_j(int ndim, const int32_t* Jd) {
    for (int d = 1; d < ndim; ++d)
        if (Jd[d] != Jd[0]) return 0;
    return Jd[0] == 6 ? 6 : 0;
}


// The bin sort of the plan's points (p->d_tm, p->M, p->bg) into its binning
// storage, in six launches: clear the (bin, bank key) counters, count, sum
// per bin, scan (subproblem table), scatter (perm), and copy tm and the
// per-sample phase into sorted order.  With gate.p set every launch does
// nothing unless the check kernel stamped gate.gen (cached sort).
template <typename T>
int launch_bin_sort(const mi355nufft_plan* p, SortGate gate) {
    const int64_t M = p->M;
    const int64_t max_sub = max_subproblems(p);
    const BinState bs = bin_state(p);
    InterpArgs<T, false> ia;
    int32_t J[3] = {p->J[0], p->J[1], p->J[2]};
    const void* dummy[3] = {p->d_tab[0], p->d_tab[0], p->d_tab[0]};
    fill_interp_args<T, false>(ia, p->ndim, p->K, J, p->L, dummy, p->d_tm, M, 1);
    hipLaunchKernelGGL(fill_int_kernel, dim3(nblocks((int64_t)p->bg.nbins * kBanks)), dim3(kBlock), 0, p->stream,
                       bs.bkt, (int64_t)p->bg.nbins * kBanks, 0, gate);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((bin_count_kernel<T>), dim3(nblocks(M)), dim3(kBlock), 0, p->stream, (const T*)p->d_tm, M,
                       p->ndim, ia.ax[0], ia.ax[1], ia.ax[2], p->bg, bs, gate);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(bin_sum_kernel, dim3(nblocks(p->bg.nbins)), dim3(kBlock), 0, p->stream, bs, p->bg.nbins,
                       gate);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(bin_scan_kernel, dim3(1), dim3(1024), 0, p->stream, bs, p->bg.nbins, p->bg.P, (int)max_sub,
                       gate);
    LAUNCH_CHECK();
    hipLaunchKernelGGL(bin_scatter_kernel, dim3(nblocks(M)), dim3(kBlock), 0, p->stream, M, bs, p->bg.P, gate);
    LAUNCH_CHECK();
    hipLaunchKernelGGL((sort_points_kernel<T>), dim3(nblocks(M)), dim3(kBlock), 0, p->stream, (const T*)p->d_tm,
                       p->has_phase ? (const cx<T>*)p->d_phase : nullptr, M, p->ndim, (const int*)bs.perm,
                       (T*)p->d_tms, (cx<T>*)p->d_phs, gate);
    LAUNCH_CHECK();
    return MI355NUFFT_OK;
}

// bin geometry of the plan's binning request (defaults per ndim; main TU)
void setup_bin_geom(mi355nufft_plan* p);

// events for the next launch of kind w (0 gather, 1 scatter) while kernel
// timing is on, or nulls
int timing_events(mi355nufft_plan* p, int w, hipEvent_t* start, hipEvent_t* stop);

// tiled gather (adjoint false) or scatter (adjoint true) of nreps
// repetitions on the plan's binned points (tiled.hip)
template <typename T, bool CK>
int run_tiled(mi355nufft_plan* p, bool adjoint, const void* samples_in, void* samples_out, const void* grid_in,
              void* grid_out, int32_t nreps);
// trajectory gradient of the tiled gather (tiled.hip)
template <typename T, bool CK>
int run_grad(mi355nufft_plan* p, const void* grid, const void* w, void* grad, int32_t nreps);

template <typename T>
int run_tiled_any(mi355nufft_plan* p, bool adjoint, const void* samples_in, void* samples_out, const void* grid_in,
                  void* grid_out, int32_t nreps) {
    return p->phasing ? run_tiled<T, true>(p, adjoint, samples_in, samples_out, grid_in, grid_out, nreps)
                      : run_tiled<T, false>(p, adjoint, samples_in, samples_out, grid_in, grid_out, nreps);
}


// What the fused FFT passes need of a plan: sizes, twiddles, deapodization
// factors (null: none) and the stream.
struct FftDev {
    int ndim;
    int64_t N[3], K[3], nk;
    void* const* tw;
    const double* const* sn;
    hipStream_t stream;
};

inline FftDev fft_dev(const mi355nufft_plan* p) {
    FftDev f;
    f.ndim = p->ndim;
    for (int d = 0; d < 3; ++d) {
        f.N[d] = p->N[d];
        f.K[d] = p->K[d];
    }
    f.nk = p->nk;
    f.tw = p->d_tw;
    f.sn = p->d_sn;
    f.stream = p->stream;
    return f;
}


// Options of one pass: rezero (adjoint: leave the grid it reads all zero),
// mult (adjoint: multiply the loaded spectrum by this grid-shaped array),
// pb (column pass of an ndim >= 2 power-of-two grid: 1 multiplies the
// output by phase_before and divides by scale_ortho if ortho, 2 multiplies
// the input by conj(phase_before); pbc null means a zero phase).
struct FftPassOpts {
    int rezero = 0;
    const void* mult = nullptr;
    int pb = 0;
    const double* pbc = nullptr;
    int ortho = 0;
    double scale_ortho = 1.0;
};

// one pass along axis 0 (rows) or axis d >= 1 (columns) of the fused
// pruned FFT (fft.hip)
template <typename T>
int fft_rows(const FftDev& f, bool adjoint, const void* x, void* xo, const int64_t* xs, void* grid, int32_t nimg,
             T scale, FftPassOpts o = FftPassOpts());
template <typename T>
int fft_cols(const FftDev& f, bool adjoint, int d, void* grid, int32_t nimg, FftPassOpts o = FftPassOpts());

}  // namespace mi355host
