/*
 * mi355nufft.h -- C ABI of the MI355X (gfx950) table1 NUFFT.
 *
 * Drop-in boundary for the table-mode ("table1") forward/adjoint NUFFT of
 * mrrt.nufft.  Reference paths are relative to the reference tree.
 *
 * Conventions shared by every entry point
 *   - return value: 0 on success, a positive MI355NUFFT_E* code otherwise;
 *     mi355nufft_last_error() gives a message for the calling thread.
 *   - every data pointer is a DEVICE pointer owned by the caller, except the
 *     set-up arrays of mi355nufft_plan_create (host pointers, copied).
 *   - complex data is interleaved (re, im), complex64 for precision 0 and
 *     complex128 for precision 1; real data is float32 / float64.
 *   - grids are Fortran ordered: index k1 + K1*k2 + K1*K2*k3; a batch of
 *     nreps grids (coils / repetitions) is stored rep-major with distance
 *     prod(Kd).  Samples are stored rep-major: y[r*M + m].
 *   - all work is enqueued on the plan's HIP stream (asynchronous).  A plan
 *     serves one stream at a time; distinct plans may run concurrently.
 *   - sizes are int64 (the reference's int32 index products overflow beyond
 *     2^31 grid elements, nufft_table.template.c:407).
 */
#ifndef MI355NUFFT_H
#define MI355NUFFT_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MI355NUFFT_ABI_VERSION 2

enum {
    MI355NUFFT_OK = 0,
    MI355NUFFT_EINVAL = 1,  /* invalid argument (sizes, null pointers, dtype) */
    MI355NUFFT_EHIP = 2,    /* HIP runtime / kernel launch error */
    MI355NUFFT_EFFT = 3,    /* rocFFT error */
    MI355NUFFT_ENOMEM = 4,  /* device allocation failed */
    MI355NUFFT_ESTATE = 5   /* call out of order (e.g. no points set) */
};

typedef struct mi355nufft_plan mi355nufft_plan;
typedef struct mi355nufft_toeplitz mi355nufft_toeplitz;

/* Message for the last failing call on this thread ("" if none). */
const char *mi355nufft_last_error(void);

/* MI355NUFFT_ABI_VERSION of the loaded library. */
int mi355nufft_abi_version(void);

/*
 * Plan creation.  Replaces the device-side state built by the table branch of
 * NufftBase.__init__ (mrrt/nufft/_nufft.py:216-359, tables :880-935 and
 * :1195-1243, sn :727-748, phase_before :703-715) and _init_gpu (:362-390).
 *   ndim            1, 2 or 3
 *   Nd, Kd, Jd      image size, oversampled grid size, kernel taps per axis
 *   L               table oversampling (table length J_d*L + 1)
 *   precision       0: complex64 / float32, 1: complex128 / float64
 *   phasing         0: real kernel + phase_before/phase_after ('real'),
 *                   1: complex kernel ('complex')
 *   ortho           nonzero: FFT normalised by 1/sqrt(prod(Kd)) each way
 *   adjoint_scale   adjoint_scalefactor (_nufft.py:1564-1568)
 *   tables          host: ndim arrays of J_d*L+1 table values, real T for
 *                   phasing 0 or interleaved complex T for phasing 1
 *   sn_factors      host: ndim float64 arrays of N_d deapodization factors;
 *                   sn[n] = (T)(sn_0[n0] * sn_1[n1] * sn_2[n2]) (double product)
 *   phase_coef      host: ndim float64 values (2*pi/K_d)*n_mid_d used by
 *                   phase_before; NULL for phasing 1
 *   stream          hipStream_t used for all work of this plan
 */
int mi355nufft_plan_create(mi355nufft_plan **plan, int ndim, const int64_t *Nd,
                           const int64_t *Kd, const int32_t *Jd, int32_t L,
                           int precision, int phasing, int ortho,
                           double adjoint_scale, const void *const *tables,
                           const double *const *sn_factors,
                           const double *phase_coef, void *stream);

/* Release every device resource of the plan (waits for its stream). */
int mi355nufft_plan_destroy(mi355nufft_plan *plan);

/* Change the stream used by subsequent calls on this plan. */
int mi355nufft_set_stream(mi355nufft_plan *plan, void *stream);

/*
 * Non-uniform points.  Replaces tm = omega / (2 pi / K_d) (_nufft.py:338-342,
 * evaluated after the cast of omega to the real dtype at :315), phase_after
 * (_nufft.py:717-724, from the uncast omega at :313) and phase_shift of the
 * complex kernel (_nufft.py:898-901), and sorts the points into bins of
 * the oversampled grid for the tiled kernels.
 *   omega        device (M, ndim) real array, element (m, d) at
 *                omega[m*stride_m + d*stride_d]
 *   omega_double 0: omega is float32, 1: float64
 *   shift        ndim float64 phase shifts: (n_shift - n_mid) rounded to the
 *                real dtype (phase_mode 1) or n_shift (phase_mode 2)
 *   phase_mode   0: no per-sample phase; 1: phase_after of the real kernel;
 *                2: phase_shift of the complex kernel
 *   dot_single   phase_mode 1 only: nonzero evaluates omega . shift in
 *                float32 (float32 omega and precision 0), else in float64
 */
int mi355nufft_set_points(mi355nufft_plan *plan, const void *omega,
                          int omega_double, int64_t M, int64_t stride_m,
                          int64_t stride_d, const double *shift, int phase_mode,
                          int dot_single);

/*
 * Points of several frames at once (no counterpart in the reference, which
 * loops over repetitions in Python, _nufft.py:1069-1084): frame f owns points
 * [frame_offsets[f], frame_offsets[f+1]) of the concatenated omega (M in
 * total, uneven per frame), and nframes*nreps grids / images in the transform
 * calls (index f*nreps + r), while samples stay (nreps, M).  One launch of
 * each tiled kernel then covers every frame and repetition.
 *   frame_offsets  host array of nframes+1 int64, 0 = f_0 <= ... <= f_n = M
 * Other arguments as mi355nufft_set_points (which is the nframes == 1 case).
 */
int mi355nufft_set_points_frames(mi355nufft_plan *plan, const void *omega,
                                 int omega_double, int64_t M, int64_t stride_m,
                                 int64_t stride_d, const double *shift,
                                 int phase_mode, int dot_single,
                                 const int64_t *frame_offsets, int32_t nframes);

/* Number of points set by the last mi355nufft_set_points (0 if none). */
int64_t mi355nufft_num_points(const mi355nufft_plan *plan);

/*
 * Bin geometry and kernel choice for the interpolation stages (no counterpart
 * in the reference, whose kernels run one thread per sample unsorted).
 *   bin_size    ndim bin sizes in grid points (NULL or 0 entries: defaults
 *               2-D 16x16; 3-D 16x16x16 for complex64 with the fixed-point
 *               scatter tile and no sparse matrix set, else 16x8x8; 1-D
 *               512; the scatter tile, 8 bytes
 *               per point in fixed point or 16 in float64 (see
 *               mi355nufft_set_scatter), must fit the LDS budget)
 *   max_points  points per subproblem (0: default 384 in 2-D; 3072 in 3-D
 *               with 16x16x16 bins, else 1024; 512 in 1-D)
 *   method      0 auto (tiled when the tile fits in LDS), 1 global-memory
 *               kernels on unsorted points, 2 tiled (error if it cannot),
 *               3 tiled with the scatter flushed by global atomics,
 *               4 tiled with the two-phase scatter (tile slots + one
 *               reduction per grid point in a fixed order: bit-reproducible
 *               adjoints; needs the slots to fit 8 GiB and bins not smaller
 *               than the kernel overlap allows)
 * Re-sorts the current points if any.
 */
int mi355nufft_set_binning(mi355nufft_plan *plan, const int32_t *bin_size,
                           int32_t max_points, int method);

/*
 * mi355nufft_set_binning plus the workgroup size of the tiled kernels:
 * threads = 64, 128 or 256 (0: default 256).  No counterpart in the
 * reference.
 */
int mi355nufft_set_tiling(mi355nufft_plan *plan, const int32_t *bin_size,
                          int32_t max_points, int method, int32_t threads);

/*
 * Kernel timing.  While enabled, every tiled gather (which = 0) or scatter
 * (which = 1) launch on a stream that is not being captured carries a HIP
 * event pair on the plan's stream, timestamped by the kernel's dispatch
 * (hipExtLaunchKernel), so the pair spans the kernel's execution only.  mi355nufft_read_timing waits for
 * the recorded launches, returns their number and summed duration in
 * milliseconds, and resets both.  enable: 0 off, 1 events, 3 events and the
 * scatter's flush atomic counts (mi355nufft_read_flush_atomics).
 * Instrumentation for benchmarks; no counterpart in the reference.
 */
int mi355nufft_set_timing(mi355nufft_plan *plan, int32_t enable);
int mi355nufft_read_timing(mi355nufft_plan *plan, int32_t which, int64_t *launches,
                           double *total_ms);

/*
 * Global atomic lane operations (float atomic adds, one per component of a
 * flushed tile cell) made by the tiled scatter launches timed since the last
 * call, then reset.  Counted in the kernel while timing is enabled with bit 1
 * set (mi355nufft_set_timing(plan, 3)); waits for the plan's stream.  The flush's atomic rate is this count over the
 * scatter's timed duration.  Instrumentation for benchmarks; no counterpart
 * in the reference, whose scatter makes two atomics per tap.
 */
int mi355nufft_read_flush_atomics(mi355nufft_plan *plan, uint64_t *lane_ops);

/*
 * Accumulation of the tiled scatter's LDS tile (no counterpart in the
 * reference, whose scatter accumulates in the operator's precision):
 *   accumulate  0 auto: 2-D and 3-D complex64 plans accumulate in 32-bit
 *               fixed point (per subproblem a power-of-two scale chosen so
 *               that no cell sum can overflow; relative rounding of order
 *               P * 2^-31 per cell, independent of the order of the atomics;
 *               about 8e-7 relative L2 from the float64 tile on a Kd 256^3
 *               kooshball grid, 3e-7 on a Kd 512^2 golden-angle grid), all
 *               others in float64.  1: float64.  2: 32-bit fixed point for
 *               every complex64 plan (complex128 stays float64).
 * Re-sorts the current points if any (the tile size decides whether the
 * tiled kernels fit the LDS).
 */
int mi355nufft_set_scatter(mi355nufft_plan *plan, int accumulate);

/*
 * Plan description: info[0..18] = ndim, M, bin size (3), points per
 * subproblem, number of bins, tile points, tiled kernels active (0/1),
 * subproblem slots, method, LDS budget in bytes, number of frames,
 * two-phase scatter active (0/1), threads per tiled workgroup, fused pruned
 * FFT active (0/1), sparse-matrix entries (-1 in table mode), sparse-matrix
 * column pieces, 32-bit fixed-point scatter tile (0/1).
 */
int mi355nufft_plan_info(const mi355nufft_plan *plan, int64_t *info, int32_t n);

/*
 * Forward transform, image -> non-uniform samples.  Replaces nufft_forward
 * (mrrt/nufft/_nufft.py:1275-1397) in table mode, including NufftBase.fft's
 * repetitions (:425-452).
 *   x          device complex image batch; element (n0, n1, n2, r) at
 *              x[n0*xs[0] + n1*xs[1] + n2*xs[2] + r*xs[3]] (elements;
 *              unused axes ignored)
 *   y          device complex (nreps, M), written
 *   grid       device scratch of nreps*prod(Kd) complex, or NULL to use a
 *              plan-owned buffer
 */
int mi355nufft_forward(mi355nufft_plan *plan, const void *x, const int64_t *xs,
                       void *y, int32_t nreps, void *grid);

/*
 * Adjoint transform, non-uniform samples -> image.  Replaces nufft_adj
 * (mrrt/nufft/_nufft.py:1459-1578) in table mode.
 *   y          device complex (nreps, M)
 *   x          device complex image batch with element strides xs, written
 *   grid       device scratch of nreps*prod(Kd) complex (overwritten), or
 *              NULL to use a plan-owned buffer.  The last FFT stage writes
 *              zeros over the plan-owned buffer as it reads it, so each
 *              adjoint leaves it all zero and the next one skips clearing
 *              it.  A captured adjoint skips the clearing pass when the
 *              buffer is zero at capture (after one eager adjoint of the
 *              same nreps), and each replay leaves it zero again.  A call
 *              with a larger nreps reallocates the buffer, which ends the
 *              validity of adjoints captured before it.
 */
int mi355nufft_adjoint(mi355nufft_plan *plan, const void *y, void *x,
                       const int64_t *xs, int32_t nreps, void *grid);

/* ---- stages (each replaces one step of nufft_forward / nufft_adj) ---- */

/* x * sn zero-padded onto the Kd grid: _nufft.py:1326 and the zero-pad of
 * fftn(xk, Kd) at :1331. */
int mi355nufft_deapod_pad(mi355nufft_plan *plan, const void *x,
                          const int64_t *xs, void *grid, int32_t nreps);

/* In-place C2C FFT of nreps Kd grids (rocFFT); direction -1 forward
 * (unnormalised, :1331), +1 backward (unnormalised; the reference's ifftn
 * 1/prod(Kd) is folded into crop_deapod, :1524 and :1564-1568). */
int mi355nufft_fft(mi355nufft_plan *plan, void *grid, int32_t nreps,
                   int direction);

/* Forward interpolation: grid * phase_before (/ sqrt(prod Kd) if ortho),
 * table1 gather, * phase_after / phase_shift: _nufft.py:1370-1391 with
 * _nufft_table_interp (:998-1097).  The tiled kernels leave the grid
 * unchanged; the global-memory kernels and sparse mode multiply it in place
 * by phase_before (and 1/sqrt(prod Kd)), as the reference does at :1371. */
int mi355nufft_interp(mi355nufft_plan *plan, const void *grid, void *y,
                      int32_t nreps);

/* Adjoint gridding: y * conj(phase_after), table1 scatter onto a zeroed
 * grid, * conj(phase_before): _nufft.py:1495-1520 with _nufft_table_adj
 * (:1101-1191).  In sparse mode the interpolation is P^H y (:1505). */
int mi355nufft_spread(mi355nufft_plan *plan, const void *y, void *grid,
                      int32_t nreps);

/*
 * Trajectory gradient of the forward interpolation (no counterpart in the
 * reference, which has no autograd).  With y = mi355nufft_interp(grid):
 *   grad[m*ndim + d] = sum_r Re(conj(w[r*M + m]) * d y[r*M + m] / d omega[m, d])
 * including the derivative of phase_after / phase_shift; the table
 * interpolant is differentiated exactly (slope L*(h[n+1]-h[n]) per cell).
 *   grid   the forward FFT output (as passed to mi355nufft_interp)
 *   w      device complex (nreps, M) upstream gradient
 *   grad   device real (M, ndim), written (float32 / float64 per precision)
 * Needs the tiled kernels (binning method 0 or 2).
 */
int mi355nufft_interp_grad(mi355nufft_plan *plan, const void *grid,
                           const void *w, void *grad, int32_t nreps);

/* Crop [:N] of the backward-FFT grid, scale, * conj(sn): _nufft.py:1561-1572. */
int mi355nufft_crop_deapod(mi355nufft_plan *plan, const void *grid, void *x,
                           const int64_t *xs, int32_t nreps);

/* ---- fused halves of the transforms ---- */

/*
 * Image -> oversampled grid, grid = fftn(x * sn, Kd) for nimg images (x and
 * xs as mi355nufft_forward, nimg = frames * reps): _nufft.py:1326-1331.
 * phase_before (:1371) is not applied here; the gather applies it.  Grids
 * whose every K_d is a power of two in 16..2048 (16..1024 for complex128)
 * run pruned FFT passes that
 * read only the image (deapodization fused into the first pass); other
 * grids run mi355nufft_deapod_pad + mi355nufft_fft.
 */
int mi355nufft_forward_grid(mi355nufft_plan *plan, const void *x,
                            const int64_t *xs, void *grid, int32_t nimg);

/*
 * Oversampled grid -> image, x = ifftn(grid)[:Nd] * scale * conj(sn):
 * _nufft.py:1524-1572 (conj(phase_before) of :1519 is applied by the
 * scatter).  The grid is overwritten.  Power-of-two grids as above run
 * pruned passes that write only the image rows (crop and deapodization
 * fused into the last pass); others mi355nufft_fft + mi355nufft_crop_deapod.
 */
int mi355nufft_adjoint_image(mi355nufft_plan *plan, void *grid, void *x,
                             const int64_t *xs, int32_t nimg);

/*
 * FFT used by mi355nufft_forward / adjoint / forward_grid / adjoint_image:
 * method 0 auto (pruned fused passes where supported, else rocFFT), 1 rocFFT.
 * No counterpart in the reference (which calls the array module's fftn).
 */
int mi355nufft_set_fft(mi355nufft_plan *plan, int method);

/* ---- sparse mode ---- */

/*
 * Sparse-matrix mode: builds the interpolation matrix P (M x prod Kd) of
 * NufftBase._init_sparsemat (mrrt/nufft/_nufft.py:751-877) on the device from
 * its per-axis factors, for the points of the last mi355nufft_set_points.
 * Until mi355nufft_clear_sparse or the next mi355nufft_set_points, the
 * plan's forward / interp use x = P xk (:1383-1384) and adjoint / spread use
 * xk = P^H y (:1505) instead of the table1 kernels; phase_before,
 * phase_after, ortho and the FFTs are unchanged.
 *   coef[d]  device (J_d, M) factors ud[d] of :783-802: float64 c (phasing
 *            0) or complex128 exp(i gam (N_d-1)/2 arg) * c (phasing 1)
 *   kidx[d]  device (J_d, M) int32 mod(koff + 1..J_d, K_d) (:788-791):
 *            kidx[d][j*M + m] must equal (kidx[d][m] + j) mod K_d
 *   phase    device complex128 (M) exp(i omega . n_shift) of :833-838
 *            (phasing 1 with a nonzero n_shift), or NULL
 * Entry (m, kk) = prod_d ud[d] in float64 / complex128 (conjugated for
 * phasing 1, times phase), rounded to the plan's precision (:854-858).
 * The matrix is held twice: rows for the forward (M * prod(Jd) values and
 * each row's first grid index per axis) and columns for the adjoint
 * (M * prod(Jd) values and sample indices); the adjoint writes every
 * grid point once, without atomics, so it is bit-reproducible.  Needs
 * M * prod(Jd) < 2^31 and one frame.
 */
int mi355nufft_set_sparse(mi355nufft_plan *plan, const void *const *coef,
                          const int32_t *const *kidx, const void *phase);

/* Frees the sparse matrix; the plan returns to the table1 kernels. */
int mi355nufft_clear_sparse(mi355nufft_plan *plan);

/*
 * The bare products of the sparse matrix, without any phase or scale:
 * adjoint 0: out (nreps, M) = P in (nreps, prod Kd)   (obj.p * xk, :1384)
 * adjoint 1: out (nreps, prod Kd) = P^H in (nreps, M) (obj.p.H * xk, :1505)
 */
int mi355nufft_sparse_apply(mi355nufft_plan *plan, int adjoint, const void *in,
                            void *out, int32_t nreps);

/*
 * Atomic throughput of this GPU, measured by timing one kernel launch
 * (measurement only; no counterpart in the reference).  The rates the
 * scatter is compared with:
 *   kind 0  LDS float64 atomic adds (ds_add_f64) at scattered addresses
 *   kind 1  LDS 32-bit integer atomic adds (ds_add_u32) at scattered
 *           addresses
 *   kind 2  global float32 atomic adds without return, coalesced, over
 *           256 MiB
 *   kind 3  LDS 64-bit integer atomic adds (ds_add_u64) at scattered
 *           addresses (the 3-D fixed-point tile's packed cells)
 * ops_per_s receives lane operations per second over the whole device.
 * Allocates and frees its own buffer; waits for the stream.
 */
int mi355nufft_measure_atomics(int kind, double *ops_per_s, void *stream);

/* ---- multi-coil encoding (no plan) ---- */

/*
 * The multi-coil (SENSE) steps a caller wraps around the transforms.  The
 * reference leaves them to MRI_Operator of the external mrrt.operators
 * (recommended in mrrt/nufft/_nufft.py:3-5; not part of the tree).  Images
 * are flat, pixel n of image i at [i*npix + n], with one pixel order shared
 * by images and maps.  precision 0: complex64 / float32, 1: complex128 /
 * float64.  Enqueued on stream.
 *
 * coil_expand:  out[(f*C + c)*npix + n] = smaps[c*npix + n] * x[f*npix + n]
 *               for nimg images f and C = ncoils maps
 * coil_combine: x[f*npix + n] = sum_c conj(smaps[c*npix + n]) * img[(f*C + c)*npix + n]
 * weight_samples: out[r*M + m] = w[m] * y[r*M + m] (real density weights)
 */
int mi355nufft_coil_expand(int precision, const void *x, const void *smaps,
                           int64_t npix, int32_t nimg, int32_t ncoils,
                           void *out, void *stream);
int mi355nufft_coil_combine(int precision, const void *img, const void *smaps,
                            int64_t npix, int32_t nimg, int32_t ncoils,
                            void *x, void *stream);
int mi355nufft_weight_samples(int precision, const void *y, const void *w,
                              int64_t M, int32_t nreps, void *out, void *stream);

/* ---- raw table1 kernels (no plan) ---- */

/*
 * Table1 gather on device buffers.  Replaces the 12 C kernels
 * {float,double}_interp{1,2,3}_table1_{real,complex}_forward
 * (mrrt/nufft/_extensions/c/nufft_table.template.h:34-278, bodies in
 * nufft_table.template.c:43-141, 538-703, 710-921), the Cython bridge
 * _interp{1,2,3}_table_forward (_extensions/_nufft_table.pyx:29-117,
 * 215-325, 444-571) and the CUDA kern_forward launch (_nufft.py:1057-1084).
 *   ck      device complex grids (nreps, prod(Kd))
 *   h       host array of ndim DEVICE table pointers (J_d*L+1 entries each;
 *           real T, or interleaved complex T if complex_kernel)
 *   tm      device real (ndim, M): tm[d*M + m]  (the reference's p_tm)
 *   fm      device complex (nreps, M), written
 * The points run on the tiled kernels through a bin sort of tm that the
 * library caches per (device, stream) in a workspace it allocates (grown on
 * demand, kept until the process ends).  Every call compares tm with the
 * copy the cached sort was made from, on the device, and re-sorts within the
 * same call when anything differs, also when tm was rewritten in place; the
 * host reads nothing back, so the call stays asynchronous.  Geometries whose
 * scatter tile does not fit the LDS, and calls beyond 16 (device, stream)
 * workspaces, run the global-memory kernels (one thread per sample).
 * Calls on one stream are ordered; calls on different streams use different
 * workspaces.
 */
int mi355nufft_table1_forward(int ndim, const int64_t *Kd, const int32_t *Jd,
                              int32_t L, int precision, int complex_kernel,
                              const void *ck, const void *const *h,
                              const void *tm, int64_t M, void *fm,
                              int32_t nreps, void *stream);

/*
 * Table1 scatter ck[k] += conj(coef) * fm[m] onto zeroed grids.  Replaces
 * {float,double}_interp{1,2,3}_table1_{real,complex}_adj (nufft_table.template.c:
 * 144-305, 308-531, 924-1202), _interp{1,2,3}_table_adj (_nufft_table.pyx:
 * 120-212, 328-441, 574-704) and the CUDA kern_adj launch (_nufft.py:1166-1189).
 * The output grids are zeroed first, as the C kernels do (template.c:336-337).
 * Points, workspace and fallback as mi355nufft_table1_forward.
 */
int mi355nufft_table1_adjoint(int ndim, const int64_t *Kd, const int32_t *Jd,
                              int32_t L, int precision, int complex_kernel,
                              const void *fm, const void *const *h,
                              const void *tm, int64_t M, void *ck,
                              int32_t nreps, void *stream);

/*
 * mi355nufft_table1_adjoint that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, result and zeroed output as
 * mi355nufft_table1_adjoint; the bin sort of tm lives in a static device
 * array of 64 MiB inside the library (one per GPU, held until the process
 * ends) instead of a workspace the library allocates.  It needs
 * ndim*2*sizeof(real) + 12 bytes per point and 140 bytes per bin of the
 * default binning (2-D complex64: about 2.3 M points); larger inputs are
 * refused with MI355NUFFT_EINVAL before anything is enqueued.  Every call
 * on a GPU shares that GPU's array, which keeps the last call's sort (reused
 * while tm is unchanged, re-made otherwise): calls that overlap on one GPU
 * (on different streams, or from threads without ordering) race.
 */
int mi355nufft_table1_adjoint_noalloc(int ndim, const int64_t *Kd,
                                      const int32_t *Jd, int32_t L,
                                      int precision, int complex_kernel,
                                      const void *fm, const void *const *h,
                                      const void *tm, int64_t M, void *ck,
                                      int32_t nreps, void *stream);

/* ---- Toeplitz normal operator ---- */

/*
 * A^H W A of a table1 NUFFT as one circulant convolution.  (A^H W A x)[n] =
 * sum_n' x[n'] psf[n - n'] with psf[d] = c * sum_m w_m exp(i omega_m . d)
 * (c the operator's scale: adjoint_scalefactor, / prod(Kd) with ortho).
 * The plan embeds psf in a power-of-two grid K_d >= 2 N_d (16 <= K_d <= 2048,
 * complex128 <= 1024) and applies
 *   y = crop_N(IFFT_K(FFT_K(circ(psf)) * FFT_K(pad_K(x))))
 * with the pruned FFT passes (the spectrum multiply fused into the first
 * inverse pass).  Builds on nufft_adj(..., return_psf=True, grid_only=True)
 * (mrrt/nufft/_nufft.py:1459-1518), where the reference stops at the
 * gridded PSF.
 */
int mi355nufft_toeplitz_create(mi355nufft_toeplitz **plan, int ndim, const int64_t *Nd,
                               int precision, void *stream);
int mi355nufft_toeplitz_destroy(mi355nufft_toeplitz *plan);
int mi355nufft_toeplitz_set_stream(mi355nufft_toeplitz *plan, void *stream);
/* circulant grid sizes K_d (ndim entries) */
int mi355nufft_toeplitz_grid(const mi355nufft_toeplitz *plan, int64_t *Kd);
/*
 * psf: device complex array on the lattice d_e in [-N_e, N_e), element
 * (d_0 + N_0, d_1 + N_1, ...) at psf[sum_e (d_e + N_e) * ps[e]] (elements);
 * the d_e = -N_e entries are not used.  This is what an adjoint NUFFT of w
 * with image size 2 N and n_shift = N produces, scaled by c.
 */
int mi355nufft_toeplitz_set_psf(mi355nufft_toeplitz *plan, const void *psf, const int64_t *ps);
/* y = (A^H W A) x for nimg images; x and y strides as mi355nufft_forward's x. */
int mi355nufft_toeplitz_apply(mi355nufft_toeplitz *plan, const void *x, const int64_t *xs,
                              void *y, const int64_t *ys, int32_t nimg);

#ifdef __cplusplus
}
#endif

#endif /* MI355NUFFT_H */
