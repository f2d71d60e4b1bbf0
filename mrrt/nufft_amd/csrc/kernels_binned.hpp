// Bin-sorted, LDS-tiled table1 kernels.
//
// Points are sorted by the bin of their first tap (the low corner of their
// J^d footprint) on a grid of B_0 x B_1 x B_2 bins.  A bin's footprints lie in
// a tile of (B_d + J_d - 1) points per axis that starts at the bin's corner,
// so all taps of a bin address the tile with local, wrap-free indices; the
// periodic wrap happens once per tile point when it is staged or flushed.
// Each bin is cut into subproblems of at most P points; one workgroup runs
// one subproblem:
//   gather:  stage the tile (grid * phase_before / scale_ortho) in LDS, then
//            every point sums its taps from LDS (the reference's sum order);
//   scatter: accumulate every point's taps into an LDS tile with LDS atomics,
//            then flush the non-zero tile points to the grid with one global
//            atomic per component, multiplied by conj(phase_before).
// The launch grid is an upper bound on the subproblem count (bins + M/P),
// so no host synchronisation is needed; surplus workgroups exit at once.
#pragma once

#include <type_traits>

#include "device_common.hpp"
#include "kernels_grid.hpp"
#include "kernels_interp.hpp"

namespace mi355 {

struct BinGeom {
    int B[3];       // bin size per axis
    int nb[3];      // bins per axis, ceil(K / B)
    int T[3];       // tile size per axis, B + J - 1
    int tile;       // T0 * T1 * T2
    int nbins;      // bins of all frames: nframes * nbins_frame
    int nbins_frame;  // nb0 * nb1 * nb2
    int P;          // max points per subproblem
    float invT[3];  // 1 / T[d], for the tile index split (tile_to_grid)
};

struct BinState {
    int* bin_of;     // (M) bin * kBanks + bank key of each point
    int* rank;       // (M) rank among the points of its (bin, bank key)
    int* counts;     // (nbins) points per bin
    int* bin_start;  // (nbins) first sorted position of each bin
    int* perm;       // (M) sorted position -> point index
    int* sub_bin;    // (max_sub) bin of each subproblem
    int* sub_start;  // (max_sub) first sorted position of each subproblem
    int* sub_end;    // (max_sub) one past its last sorted position
    int* n_sub;      // (1) number of subproblems
    int* bin_sub;    // (nbins) first subproblem of each bin
    int* bkt;        // (nbins * kBanks) points per (bin, bank key)
    const int64_t* frame_off;  // (nframes + 1) first point of each frame, or null
    int nframes;
};

// Points inside a bin are ordered so that consecutive points start their
// footprints on different LDS banks.  The bank key of a point is the tile
// offset of its first tap modulo kBanks (in 8-byte tile entries); the order
// is by (rank inside the key, key, rotated by the rank in bins with an even
// number of subproblems), so any kBanks consecutive points of one rank level
// have distinct keys and points in the same grid cell are at least one level
// apart.  The same order serves the gather (tile reads) and
// the scatter (tile atomics).
constexpr int kBanks = 32;

// Optional gate of the sort kernels: with p set, a kernel does nothing unless
// *p == gen.  A check kernel earlier in the stream stamps gen when the points
// changed, so a sort can be enqueued conditionally without the host reading
// anything back (the cached sort of the raw table1 entry points).
struct SortGate {
    const unsigned* p;
    unsigned gen;
};

__device__ __forceinline__ bool gate_closed(const SortGate& g) {
    return g.p != nullptr && *(const volatile unsigned*)g.p != g.gen;
}

template <typename T>
__device__ __forceinline__ int first_tap_mod(T t, int J, int K) {
    return pos_mod(first_tap(t, J), K);
}

template <typename T>
__global__ __launch_bounds__(256) void bin_count_kernel(const T* __restrict__ tm, int64_t M, int ndim,
                                                        AxisGeom a0, AxisGeom a1, AxisGeom a2, BinGeom bg,
                                                        BinState bs, SortGate gate) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M || gate_closed(gate)) return;
    const AxisGeom ax[3] = {a0, a1, a2};
    int bin = 0, mul = 1, base = 0, tmul = 1;
    for (int d = 0; d < ndim; ++d) {
        const int k0 = first_tap_mod(tm[(int64_t)d * M + m], ax[d].J, ax[d].K);
        const int b = k0 / bg.B[d];
        bin += b * mul;
        mul *= bg.nb[d];
        base += (k0 - b * bg.B[d]) * tmul;
        tmul *= bg.T[d];
    }
    if (bs.frame_off) {  // frame of point m: last f with frame_off[f] <= m
        int lo = 0, hi = bs.nframes - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (bs.frame_off[mid] <= m) lo = mid;
            else hi = mid - 1;
        }
        bin += lo * bg.nbins_frame;
    }
    const int slot = bin * kBanks + (base & (kBanks - 1));
    bs.bin_of[m] = slot;
    bs.rank[m] = atomicAdd(&bs.bkt[slot], 1);
}

// counts[b] = points of bin b (sum over its bank keys)
static __global__ __launch_bounds__(256) void bin_sum_kernel(BinState bs, int nbins, SortGate gate) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbins || gate_closed(gate)) return;
    const int4* q = reinterpret_cast<const int4*>(bs.bkt + (int64_t)b * kBanks);
    int c = 0;
#pragma unroll
    for (int k = 0; k < kBanks / 4; ++k) {
        const int4 v = q[k];
        c += v.x + v.y + v.z + v.w;
    }
    bs.counts[b] = c;
}

// One workgroup of 1024 threads: exclusive scans of the bin counts and of the
// per-bin subproblem counts, then the subproblem table.
static __global__ __launch_bounds__(1024) void bin_scan_kernel(BinState bs, int nbins, int P, int max_sub,
                                                               SortGate gate) {
    __shared__ int s_pts[1024];
    __shared__ int s_sub[1024];
    if (gate_closed(gate)) return;  // uniform over the workgroup, before any barrier
    const int tid = threadIdx.x;
    const int chunk = (nbins + 1023) / 1024;
    const int b0 = tid * chunk;
    const int b1 = min(nbins, b0 + chunk);
    int pts = 0, subs = 0;
    for (int b = b0; b < b1; ++b) {
        const int c = bs.counts[b];
        pts += c;
        subs += (c + P - 1) / P;
    }
    s_pts[tid] = pts;
    s_sub[tid] = subs;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {  // Hillis-Steele inclusive scan
        int vp = 0, vs = 0;
        if (tid >= off) {
            vp = s_pts[tid - off];
            vs = s_sub[tid - off];
        }
        __syncthreads();
        s_pts[tid] += vp;
        s_sub[tid] += vs;
        __syncthreads();
    }
    int run_p = s_pts[tid] - pts;
    int run_s = s_sub[tid] - subs;
    for (int b = b0; b < b1; ++b) {
        const int c = bs.counts[b];
        bs.bin_start[b] = run_p;
        bs.bin_sub[b] = run_s;
        const int ns = (c + P - 1) / P;
        const int q = ns ? c / ns : 0, rem = ns ? c - q * ns : 0;
        for (int i = 0; i < ns && run_s + i < max_sub; ++i) {
            bs.sub_bin[run_s + i] = b;
            bs.sub_start[run_s + i] = run_p + i * q + min(i, rem);
            bs.sub_end[run_s + i] = run_p + (i + 1) * q + min(i + 1, rem);
        }
        run_p += c;
        run_s += ns;
    }
    if (tid == 1023) bs.n_sub[0] = min(s_sub[1023], max_sub);
}

// Sorted position of point m.  Its place in the bin's (rank, key) order is
// pos = #{(k', r') : r' < r} + #{k' before key : count(k') > r}.  A bin of c
// points has ns = ceil(c / P) subproblems, dealt round-robin over that order:
// subproblem j holds positions j, j + ns, j + 2 ns, ... (q or q + 1 points,
// c = q ns + rem), so every subproblem samples all rank levels and the dense
// cells of a bin do not pile up in its last subproblem.  Within a rank level
// the keys run in order, rotated by the rank when ns is even: with full
// levels a stride of ns would otherwise meet only 32 / gcd(ns, 32) keys
// (a subproblem of a bin with ns = 16 sees two, and its lanes queue on two
// LDS banks); rotated, consecutive levels supply the other residues.
static __global__ __launch_bounds__(256) void bin_scatter_kernel(int64_t M, BinState bs, int P, SortGate gate) {
    const int64_t m = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= M || gate_closed(gate)) return;
    const int slot = bs.bin_of[m];
    const int bin = slot / kBanks;
    const int key = slot - bin * kBanks;
    const int r = bs.rank[m];
    const int c = bs.counts[bin];
    const int ns = (c + P - 1) / P;
    const int rot = (ns & 1) ? 0 : r;  // key k comes at place (k - rot) mod kBanks of level r
    const int kr = (key - rot) & (kBanks - 1);
    const int4* row = reinterpret_cast<const int4*>(bs.bkt + (int64_t)bin * kBanks);
    int pos = 0;
#pragma unroll
    for (int k4 = 0; k4 < kBanks / 4; ++k4) {
        const int4 v = row[k4];
        const int cn[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int k = 4 * k4 + u;
            pos += min(cn[u], r) + ((((k - rot) & (kBanks - 1)) < kr && cn[u] > r) ? 1 : 0);
        }
    }
    const int q = c / ns, rem = c - q * ns;
    const int j = pos % ns;
    bs.perm[bs.bin_start[bin] + j * q + min(j, rem) + pos / ns] = (int)m;
}

static __global__ __launch_bounds__(256) void fill_int_kernel(int* __restrict__ p, int64_t n, int v, SortGate gate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n && !gate_closed(gate)) p[i] = v;
}

// Cached-sort check: compares the caller's tm with the copy the current sort
// was made from (bit patterns), copies what differs and stamps gen into *gate
// if anything did; force copies everything without comparing.
template <typename T>
__global__ __launch_bounds__(256) void tm_check_kernel(const T* __restrict__ tm, T* __restrict__ tm_copy, int64_t n,
                                                       unsigned* __restrict__ gate, unsigned gen, int force) {
    using U = typename std::conditional<sizeof(T) == 8, unsigned long long, unsigned>::type;
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const U a = reinterpret_cast<const U*>(tm)[i];
    U* c = reinterpret_cast<U*>(tm_copy) + i;
    if (force) {
        *c = a;
    } else if (*c != a) {
        *c = a;
        *gate = gen;  // every writer stores the same value
    }
}

template <typename T, bool CK>
struct TiledArgs {
    InterpArgs<T, CK> ia;   // tables, tm, samples, grid, per-sample phase, geometry
    BinGeom bg;
    BinState bs;
    T pb[3];                // phase_before coefficients (real kernel)
    int use_pb;             // multiply by phase_before (gather) / conj (scatter)
    int ortho;              // gather: divide staged values by scale_ortho
    T scale_ortho;
    cx<T>* tiles;           // scatter, two-phase: (sub, rep) tile slots; null = atomic flush
    const T* tms;           // tm in sorted order: tms[d*M + i] = tm[d*M + perm[i]]
    const cx<T>* phs;       // per-sample phase in sorted order, or null
    int rc;                 // gather: repetitions staged together (LDS holds rc tiles)
    int own_ok;             // scatter: every K_d is a multiple of B_d (see spread_tiled_kernel)
    float cmax;             // scatter: > 0 accumulates in fixed point; bound of |coefficient product|
    int exact;              // gather: 1 keeps the reference's roundings (no fused multiply-add)
    unsigned long long* flush_count;  // scatter: non-null adds the flush's global atomic lane operations
    int flush_slots;                  // power of two: workgroup b adds to flush_count[b % flush_slots]
};

// Copy tm and the per-sample phase into sorted (bin) order, so the tiled
// kernels read them coalesced instead of through perm.
template <typename T>
__global__ __launch_bounds__(256) void sort_points_kernel(const T* __restrict__ tm, const cx<T>* __restrict__ phase,
                                                          int64_t M, int ndim, const int* __restrict__ perm,
                                                          T* __restrict__ tms, cx<T>* __restrict__ phs,
                                                          SortGate gate) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= M || gate_closed(gate)) return;
    const int64_t m = perm[i];
    for (int d = 0; d < ndim; ++d) tms[(int64_t)d * M + i] = tm[(int64_t)d * M + m];
    if (phase) phs[i] = phase[m];
}

struct Sub {
    int start, end;
    int frame;  // frame of the bin (0 without frames)
    int o[3];   // tile origin (grid coordinates)
    int solo;   // the only subproblem of its bin
};

// Subproblem sid; the workgroups of a launch stride over the subproblem list
// (persistent launch), so every workgroup exits after n_sub is exhausted.
// Block b takes subproblems b, b + gridDim, ...: consecutive subproblems go
// to different XCDs, which spreads a dense bin's subproblems over all eight
// (an XCD-contiguous mapping measured 10-25% slower for that reason).
__device__ __forceinline__ void load_sub(const BinGeom& bg, const BinState& bs, int ndim, int sid, Sub& s) {
    const int bin = bs.sub_bin[sid];
    s.start = bs.sub_start[sid];
    s.end = bs.sub_end[sid];
    s.solo = bs.counts[bin] <= bg.P;
    s.frame = bin / bg.nbins_frame;
    int q = bin - s.frame * bg.nbins_frame;
    for (int d = 0; d < 3; ++d) {
        const int bd = d < ndim ? q % bg.nb[d] : 0;
        if (d < ndim) q /= bg.nb[d];
        s.o[d] = bd * bg.B[d];
    }
}

// floor(e / t) for 0 <= e < 2^22 through the float reciprocal: (e + 0.5) / t
// lies at least 0.5 / t from an integer, farther than the 2^-23 relative
// rounding of the product, so the truncation is exact (tiles are far below
// 2^22 points).  Replaces an integer division, which the GPU does in
// software.
__device__ __forceinline__ int div_small(int e, float inv_t) { return (int)(((float)e + 0.5f) * inv_t); }

// (o + e) mod K for 0 <= o < K, e >= 0: one conditional subtraction covers
// tiles no wider than the grid plus one bin; wider tiles (tiny grids) take
// the remainder.
__device__ __forceinline__ int wrap_add(int o, int e, int K) {
    int g = o + e;
    g = g >= K ? g - K : g;
    if (g >= K) g = g % K;
    return g;
}

// grid index of local tile point e (local coordinates e0, e1, e2); axes at
// or above NDIM are 0
template <int NDIM = 3>
__device__ __forceinline__ int64_t tile_to_grid(int e, const BinGeom& bg, const Sub& s, const AxisGeom* ax,
                                                int& g0, int& g1, int& g2, int& e0, int& e1, int& e2) {
    const int q = div_small(e, bg.invT[0]);
    e0 = e - q * bg.T[0];
    if (NDIM > 2) {
        e2 = div_small(q, bg.invT[1]);
        e1 = q - e2 * bg.T[1];
    } else {
        e2 = 0;
        e1 = NDIM > 1 ? q : 0;
    }
    g0 = wrap_add(s.o[0], e0, ax[0].K);
    g1 = NDIM > 1 ? wrap_add(s.o[1], e1, ax[1].K) : 0;
    g2 = NDIM > 2 ? wrap_add(s.o[2], e2, ax[2].K) : 0;
    return ((int64_t)g2 * ax[1].K + g1) * ax[0].K + g0;
}

// grid index of local tile point e
template <int NDIM = 3>
__device__ __forceinline__ int64_t tile_to_grid(int e, const BinGeom& bg, const Sub& s, const AxisGeom* ax,
                                                int& g0, int& g1, int& g2) {
    int e0, e1, e2;
    return tile_to_grid<NDIM>(e, bg, s, ax, g0, g1, g2, e0, e1, e2);
}

// Stage the tile of subproblem s from grid g into LDS: grid * phase_before
// (real kernel) / scale_ortho (ortho), the order of _nufft.py:1370-1375.
template <typename T, int NDIM, bool CK>
__device__ __forceinline__ void stage_tile(const TiledArgs<T, CK>& a, const Sub& s, const AxisGeom* ax,
                                           const cx<T>* __restrict__ g, cx<T>* tile) {
    const BinGeom& bg = a.bg;
    constexpr int U = 4;  // grid loads in flight per thread before the phase math
    const T c[3] = {a.pb[0], a.pb[1], a.pb[2]};
    for (int e0 = threadIdx.x; e0 < bg.tile; e0 += U * blockDim.x) {
        cx<T> v[U];
        T phi[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * blockDim.x;
            if (e < bg.tile) {
                int g0, g1, g2;
                const int64_t gi = tile_to_grid<NDIM>(e, bg, s, ax, g0, g1, g2);
                v[u] = g[gi];
                phi[u] = phase_before_arg<T>(c, NDIM, g0, g1, g2);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * blockDim.x;
            if (e < bg.tile) {
                cx<T> w = v[u];
                if (a.use_pb) {
                    const cx<T> p = cis_tile(phi[u]);
                    const T vr = w.x * p.x - w.y * p.y;
                    const T vi = w.x * p.y + w.y * p.x;
                    w.x = vr;
                    w.y = vi;
                }
                if (a.ortho) {
                    w.x = w.x / a.scale_ortho;
                    w.y = w.y / a.scale_ortho;
                }
                tile[e] = w;
            }
        }
    }
}

// Coefficients and tile offsets of one point's J^d taps (sorted position i).
template <typename T, bool CK, int JMAX>
struct PointTaps {
    T cr[3][JMAX];
    T ci[3][JMAX];   // complex kernel only
    int l[3];        // first tap in tile coordinates
    int64_t m;       // point index in the caller's order
    cx<T> ph;        // per-sample phase (1 if none)
};

template <typename T, int NDIM, bool CK, int JMAX, bool FIXED>
__device__ __forceinline__ void point_taps(const TiledArgs<T, CK>& a, const AxisGeom* ax, const Sub& s, int i,
                                           PointTaps<T, CK, JMAX>& pt) {
    const InterpArgs<T, CK>& ia = a.ia;
    pt.m = a.bs.perm[i];
    pt.ph = a.phs ? a.phs[i] : mkcx<T>((T)1, (T)0);
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        const T t = a.tms[(int64_t)d * ia.M + i];
        int k0;
        if constexpr (CK) k0 = axis_coefs<T, JMAX, FIXED>(t, ia.L, ax[d], ia.h[d], pt.cr[d], pt.ci[d]);
        else k0 = axis_coefs<T, JMAX, FIXED>(t, ia.L, ax[d], ia.h[d], pt.cr[d]);
        pt.l[d] = k0 - s.o[d];
    }
}

// point_taps with the point's positions, index and phase already loaded
template <typename T, int NDIM, bool CK, int JMAX, bool FIXED>
__device__ __forceinline__ void point_taps_pre(const TiledArgs<T, CK>& a, const AxisGeom* ax, const Sub& s,
                                               const T (&t)[3], int m, cx<T> ph, PointTaps<T, CK, JMAX>& pt) {
    const InterpArgs<T, CK>& ia = a.ia;
    pt.m = m;
    pt.ph = ph;
#pragma unroll
    for (int d = 0; d < NDIM; ++d) {
        int k0;
        if constexpr (CK) k0 = axis_coefs<T, JMAX, FIXED>(t[d], ia.L, ax[d], ia.h[d], pt.cr[d], pt.ci[d]);
        else k0 = axis_coefs<T, JMAX, FIXED>(t[d], ia.L, ax[d], ia.h[d], pt.cr[d]);
        pt.l[d] = k0 - s.o[d];
    }
}

// sum over the J^d taps of the staged tile, in the reference's order
// (FUSED: with fused multiply-adds)
template <typename T, int NDIM, bool CK, int JMAX, bool FIXED, bool FUSED = false>
__device__ __forceinline__ cx<T> gather_taps(const cx<T>* tile, const PointTaps<T, CK, JMAX>& pt,
                                             const AxisGeom* ax, int T0, int T1, bool has_ph) {
    T sr = 0, si = 0;
    if constexpr (NDIM == 1) {
#pragma unroll
        for (int j1 = 0; j1 < JMAX; ++j1)
            if (FIXED || j1 < ax[0].J) {
                const cx<T> v = tile[pt.l[0] + j1];
                madd_sel<T, CK, FUSED>(sr, si, pt.cr[0][j1], pt.ci[0][j1], v.x, v.y);
            }
    } else if constexpr (NDIM == 2) {
#pragma unroll
        for (int j2 = 0; j2 < JMAX; ++j2)
            if (FIXED || j2 < ax[1].J) {
                const cx<T>* row = tile + (pt.l[1] + j2) * T0 + pt.l[0];
                T s1r = 0, s1i = 0;
#pragma unroll
                for (int j1 = 0; j1 < JMAX; ++j1)
                    if (FIXED || j1 < ax[0].J) {
                        const cx<T> v = row[j1];
                        madd_sel<T, CK, FUSED>(s1r, s1i, pt.cr[0][j1], pt.ci[0][j1], v.x, v.y);
                    }
                madd_sel<T, CK, FUSED>(sr, si, pt.cr[1][j2], pt.ci[1][j2], s1r, s1i);
            }
    } else if constexpr (FIXED) {
        // The 36 rows of the (j2, j3) plane in the same order, each loaded
        // R rows ahead of its use: left alone, the compiler keeps two LDS
        // reads in flight and waits on each pair (measured 5% faster at
        // Kd 256^3 with R = 4, at the same occupancy).
        constexpr int J = JMAX, R = 4, NR = JMAX * JMAX;
        cx<T> buf[R][J];
        auto rowp = [&](int r) { return tile + ((pt.l[2] + r / J) * T1 + (pt.l[1] + r % J)) * T0 + pt.l[0]; };
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const cx<T>* q = rowp(r);
#pragma unroll
            for (int j1 = 0; j1 < J; ++j1) buf[r][j1] = q[j1];
        }
        T s2r = 0, s2i = 0;
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            cx<T> cur[J];
#pragma unroll
            for (int j1 = 0; j1 < J; ++j1) cur[j1] = buf[r % R][j1];
            if (r + R < NR) {
                const cx<T>* q = rowp(r + R);
#pragma unroll
                for (int j1 = 0; j1 < J; ++j1) buf[r % R][j1] = q[j1];
            }
            T s1r = 0, s1i = 0;
#pragma unroll
            for (int j1 = 0; j1 < J; ++j1)
                madd_sel<T, CK, FUSED>(s1r, s1i, pt.cr[0][j1], pt.ci[0][j1], cur[j1].x, cur[j1].y);
            madd_sel<T, CK, FUSED>(s2r, s2i, pt.cr[1][r % J], pt.ci[1][r % J], s1r, s1i);
            if (r % J == J - 1) {
                madd_sel<T, CK, FUSED>(sr, si, pt.cr[2][r / J], pt.ci[2][r / J], s2r, s2i);
                s2r = 0;
                s2i = 0;
            }
        }
    } else {
#pragma unroll 1
        for (int j3 = 0; j3 < JMAX; ++j3)
            if (j3 < ax[2].J) {
                T s2r = 0, s2i = 0;
#pragma unroll
                for (int j2 = 0; j2 < JMAX; ++j2)
                    if (FIXED || j2 < ax[1].J) {
                        const cx<T>* row = tile + ((pt.l[2] + j3) * T1 + (pt.l[1] + j2)) * T0 + pt.l[0];
                        T s1r = 0, s1i = 0;
#pragma unroll
                        for (int j1 = 0; j1 < JMAX; ++j1)
                            if (FIXED || j1 < ax[0].J) {
                                const cx<T> v = row[j1];
                                madd_sel<T, CK, FUSED>(s1r, s1i, pt.cr[0][j1], pt.ci[0][j1], v.x, v.y);
                            }
                        madd_sel<T, CK, FUSED>(s2r, s2i, pt.cr[1][j2], pt.ci[1][j2], s1r, s1i);
                    }
                madd_sel<T, CK, FUSED>(sr, si, pt.cr[2][j3], pt.ci[2][j3], s2r, s2i);
            }
    }
    if (!has_ph) return mkcx<T>(sr, si);
    cx<T> o;
    o.x = sr * pt.ph.x - si * pt.ph.y;
    o.y = sr * pt.ph.y + si * pt.ph.x;
    return o;
}

// add conj(coefficient product) * f at every tap of the float64 LDS tile
template <typename T, int NDIM, bool CK, int JMAX, bool FIXED>
__device__ __forceinline__ void scatter_taps(double* tre, double* tim, const PointTaps<T, CK, JMAX>& pt, cx<T> f,
                                             const AxisGeom* ax, int T0, int T1) {
    if constexpr (NDIM == 1) {
#pragma unroll
        for (int j1 = 0; j1 < JMAX; ++j1)
            if (FIXED || j1 < ax[0].J) {
                T vr, vi;
                cmul_conj<T, CK>(vr, vi, pt.cr[0][j1], pt.ci[0][j1], f.x, f.y);
                atomicAdd(tre + pt.l[0] + j1, (double)vr);
                atomicAdd(tim + pt.l[0] + j1, (double)vi);
            }
    } else if constexpr (NDIM == 2) {
        // real kernel: the axis-0 coefficient times the row value in float64
        // (the product of two floats is exact there), one conversion per
        // coefficient and per row instead of two per tap
        double c0d[JMAX];
        if constexpr (!CK) {
#pragma unroll
            for (int j1 = 0; j1 < JMAX; ++j1) c0d[j1] = (double)pt.cr[0][j1];
        }
#pragma unroll
        for (int j2 = 0; j2 < JMAX; ++j2)
            if (FIXED || j2 < ax[1].J) {
                T v2r, v2i;
                cmul_conj<T, CK>(v2r, v2i, pt.cr[1][j2], pt.ci[1][j2], f.x, f.y);
                const int row = (pt.l[1] + j2) * T0 + pt.l[0];
                const double v2rd = (double)v2r, v2id = (double)v2i;
#pragma unroll
                for (int j1 = 0; j1 < JMAX; ++j1)
                    if (FIXED || j1 < ax[0].J) {
                        if constexpr (CK) {
                            T vr, vi;
                            cmul_conj<T, CK>(vr, vi, pt.cr[0][j1], pt.ci[0][j1], v2r, v2i);
                            atomicAdd(tre + row + j1, (double)vr);
                            atomicAdd(tim + row + j1, (double)vi);
                        } else {
                            atomicAdd(tre + row + j1, c0d[j1] * v2rd);
                            atomicAdd(tim + row + j1, c0d[j1] * v2id);
                        }
                    }
            }
    } else {
        double c0d[JMAX];
        if constexpr (!CK) {
#pragma unroll
            for (int j1 = 0; j1 < JMAX; ++j1) c0d[j1] = (double)pt.cr[0][j1];
        }
#pragma unroll FIXED ? JMAX : 1
        for (int j3 = 0; j3 < JMAX; ++j3)
            if (FIXED || j3 < ax[2].J) {
                T v3r, v3i;
                cmul_conj<T, CK>(v3r, v3i, pt.cr[2][j3], pt.ci[2][j3], f.x, f.y);
#pragma unroll
                for (int j2 = 0; j2 < JMAX; ++j2)
                    if (FIXED || j2 < ax[1].J) {
                        T v2r, v2i;
                        cmul_conj<T, CK>(v2r, v2i, pt.cr[1][j2], pt.ci[1][j2], v3r, v3i);
                        const int row = ((pt.l[2] + j3) * T1 + (pt.l[1] + j2)) * T0 + pt.l[0];
                        const double v2rd = (double)v2r, v2id = (double)v2i;
#pragma unroll
                        for (int j1 = 0; j1 < JMAX; ++j1)
                            if (FIXED || j1 < ax[0].J) {
                                if constexpr (CK) {
                                    T vr, vi;
                                    cmul_conj<T, CK>(vr, vi, pt.cr[0][j1], pt.ci[0][j1], v2r, v2i);
                                    atomicAdd(tre + row + j1, (double)vr);
                                    atomicAdd(tim + row + j1, (double)vi);
                                } else {
                                    atomicAdd(tre + row + j1, c0d[j1] * v2rd);
                                    atomicAdd(tim + row + j1, c0d[j1] * v2id);
                                }
                            }
                    }
            }
    }
}

// A tap's two fixed-point components as one signed 64-bit integer,
// im * 2^32 + re.  Sums of such values are exact in 64 bits and give back
// both component sums while each stays within int32 (the tile's scale
// guarantees it): the low 32 bits hold sum(re), and (v - sum(re)) / 2^32 is
// sum(im) (fx_unpack).
__device__ __forceinline__ unsigned long long fx_pack(float vr, float vi) {
    const long long v = (long long)__float2int_rn(vi) * 4294967296LL + (long long)__float2int_rn(vr);
    return (unsigned long long)v;
}
__device__ __forceinline__ void fx_unpack(long long v, int& re, int& im) {
    re = (int)(unsigned)(unsigned long long)v;
    im = (int)((v - (long long)re) / 4294967296LL);
}

// Fixed-point scatter taps (complex64): the tap values of a sample already
// scaled to fixed-point units, rounded to the nearest integer and added to
// the tile.  PACK: one tile of packed 64-bit cells (fx_pack, one ds_add_u64
// per tap); else a plane of int32 real parts (tre) and one of imaginary
// parts (tim), two ds_add_u32 per tap.  Same taps, same products and the
// same nesting as scatter_taps.
template <bool PACK>
__device__ __forceinline__ void fx_add(int* tre, int* tim, int i, float vr, float vi) {
    if constexpr (PACK) {
        atomicAdd(reinterpret_cast<unsigned long long*>(tre) + i, fx_pack(vr, vi));
    } else {
        atomicAdd(tre + i, __float2int_rn(vr));
        atomicAdd(tim + i, __float2int_rn(vi));
    }
}

template <int NDIM, bool CK, int JMAX, bool FIXED, bool PACK>
__device__ __forceinline__ void scatter_taps_fx(int* tre, int* tim, const PointTaps<float, CK, JMAX>& pt, float2 f,
                                                const AxisGeom* ax, int T0, int T1) {
    if constexpr (NDIM == 1) {
#pragma unroll
        for (int j1 = 0; j1 < JMAX; ++j1)
            if (FIXED || j1 < ax[0].J) {
                float vr, vi;
                cmul_conj<float, CK>(vr, vi, pt.cr[0][j1], pt.ci[0][j1], f.x, f.y);
                fx_add<PACK>(tre, tim, pt.l[0] + j1, vr, vi);
            }
    } else {
        auto rows = [&](float v3r, float v3i, int plane) {
#pragma unroll
            for (int j2 = 0; j2 < JMAX; ++j2)
                if (FIXED || j2 < ax[1].J) {
                    float v2r, v2i;
                    cmul_conj<float, CK>(v2r, v2i, pt.cr[1][j2], pt.ci[1][j2], v3r, v3i);
                    const int row = (plane + pt.l[1] + j2) * T0 + pt.l[0];
#pragma unroll
                    for (int j1 = 0; j1 < JMAX; ++j1)
                        if (FIXED || j1 < ax[0].J) {
                            float vr, vi;
                            cmul_conj<float, CK>(vr, vi, pt.cr[0][j1], pt.ci[0][j1], v2r, v2i);
                            fx_add<PACK>(tre, tim, row + j1, vr, vi);
                        }
                }
        };
        if constexpr (NDIM == 2) {
            rows(f.x, f.y, 0);
        } else {
#pragma unroll FIXED ? JMAX : 1
            for (int j3 = 0; j3 < JMAX; ++j3)
                if (FIXED || j3 < ax[2].J) {
                    float v3r, v3i;
                    cmul_conj<float, CK>(v3r, v3i, pt.cr[2][j3], pt.ci[2][j3], f.x, f.y);
                    rows(v3r, v3i, (pt.l[2] + j3) * T1);
                }
        }
    }
}

// Sum of v over the workgroup, the same value in every thread (wave64
// butterfly, then the per-wave partial sums in a fixed order).  red holds one
// float per wave; the leading barrier protects it from a previous use.
__device__ __forceinline__ float block_sum(float v, float* red) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
    __syncthreads();
    float t = 0.f;
    const int nw = blockDim.x >> 6;
    for (int w = 0; w < nw; ++w) t += red[w];
    return t;
}

// MULTI: repetitions (coils) are processed a.rc at a time: a.rc tiles are
// staged, and each point's coefficients are computed once for all of them.
// FUSED: the tap sums use fused multiply-adds.
template <typename T, int NDIM, bool CK, int JMAX, bool FIXED, bool MULTI, bool FUSED = false>
__global__ __launch_bounds__(256) void interp_tiled_kernel(TiledArgs<T, CK> a) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* tiles = reinterpret_cast<cx<T>*>(smem);
    const InterpArgs<T, CK>& ia = a.ia;
    const BinGeom& bg = a.bg;
    const AxisGeom ax[3] = {ia.ax[0], ia.ax[1], ia.ax[2]};
    const int T0 = bg.T[0], T1 = bg.T[1];
    const int tid = threadIdx.x;
    const int n_sub = a.bs.n_sub[0];
    for (int sid = blockIdx.x; sid < n_sub; sid += gridDim.x) {
        Sub s;
        load_sub(bg, a.bs, NDIM, sid, s);
        if constexpr (!MULTI) {
            for (int r = 0; r < ia.nreps; ++r) {
                stage_tile<T, NDIM, CK>(a, s, ax, ia.grid + ((int64_t)s.frame * ia.nreps + r) * ia.kstride, tiles);
                __syncthreads();
                if constexpr (NDIM >= 2) {
                    // several points per thread (3-D subproblems hold up to
                    // 3,072, 2-D ones 384): the next point's position, index
                    // and phase load while this one's taps are summed
                    // (measured 8% faster at Kd 256^3, where the tile's LDS
                    // bounds occupancy, not the extra registers; 2-D at Kd
                    // 512^2 11.4 -> 11.1 us, 99 VGPRs, 4 waves per SIMD,
                    // enough for its 1,001 subproblems)
                    int i = s.start + tid;
                    T tn[3] = {0, 0, 0};
                    int mn = 0;
                    cx<T> pn = mkcx<T>((T)1, (T)0);
                    if (i < s.end) {
#pragma unroll
                        for (int d = 0; d < 3; ++d) tn[d] = a.tms[(int64_t)d * ia.M + i];
                        mn = a.bs.perm[i];
                        if (a.phs) pn = a.phs[i];
                    }
                    for (; i < s.end; i += blockDim.x) {
                        const T tc[3] = {tn[0], tn[1], tn[2]};
                        const int mc = mn;
                        const cx<T> pc = pn;
                        const int inext = i + blockDim.x;
                        if (inext < s.end) {
#pragma unroll
                            for (int d = 0; d < 3; ++d) tn[d] = a.tms[(int64_t)d * ia.M + inext];
                            mn = a.bs.perm[inext];
                            if (a.phs) pn = a.phs[inext];
                        }
                        PointTaps<T, CK, JMAX> pt;
                        point_taps_pre<T, NDIM, CK, JMAX, FIXED>(a, ax, s, tc, mc, pc, pt);
                        ia.out[(int64_t)r * ia.M + pt.m] =
                            gather_taps<T, NDIM, CK, JMAX, FIXED, FUSED>(tiles, pt, ax, T0, T1, a.phs);
                    }
                } else {
                    for (int i = s.start + tid; i < s.end; i += blockDim.x) {
                        PointTaps<T, CK, JMAX> pt;
                        point_taps<T, NDIM, CK, JMAX, FIXED>(a, ax, s, i, pt);
                        ia.out[(int64_t)r * ia.M + pt.m] =
                            gather_taps<T, NDIM, CK, JMAX, FIXED, FUSED>(tiles, pt, ax, T0, T1, a.phs);
                    }
                }
                __syncthreads();
            }
            continue;
        }
        for (int r0 = 0; r0 < ia.nreps; r0 += a.rc) {
            const int rc = min(a.rc, ia.nreps - r0);
            for (int rr = 0; rr < rc; ++rr)
                stage_tile<T, NDIM, CK>(a, s, ax, ia.grid + ((int64_t)s.frame * ia.nreps + r0 + rr) * ia.kstride,
                                        tiles + (int64_t)rr * bg.tile);
            __syncthreads();
            for (int i = s.start + tid; i < s.end; i += blockDim.x) {
                PointTaps<T, CK, JMAX> pt;
                point_taps<T, NDIM, CK, JMAX, FIXED>(a, ax, s, i, pt);
                for (int rr = 0; rr < rc; ++rr)
                    ia.out[(int64_t)(r0 + rr) * ia.M + pt.m] =
                        gather_taps<T, NDIM, CK, JMAX, FIXED, FUSED>(tiles + (int64_t)rr * bg.tile, pt, ax, T0, T1, a.phs);
            }
            __syncthreads();
        }
    }
}

template <typename T, bool CK, int JMAX>
__device__ __forceinline__ cx<T> times_conj_phase(cx<T> f, const PointTaps<T, CK, JMAX>& pt, bool has_ph) {
    if (!has_ph) return f;
    return mkcx<T>(f.x * pt.ph.x + f.y * pt.ph.y, f.y * pt.ph.x - f.x * pt.ph.y);
}

template <typename T, bool CK, int JMAX>
__device__ __forceinline__ cx<T> sample_times_conj_phase(const InterpArgs<T, CK>& ia, int r,
                                                         const PointTaps<T, CK, JMAX>& pt, bool has_ph) {
    return times_conj_phase<T, CK, JMAX>(ia.in[(int64_t)r * ia.M + pt.m], pt, has_ph);
}

// RC repetitions per pass, each with its own tile (RC = 2 when both tiles
// fit the LDS budget and a subproblem has more points than threads, so
// that each point's coefficients serve two repetitions; more tiles per pass
// were measured no faster, the LDS atomics dominating, and cost occupancy).
//
// FXP (complex64 only): the tile accumulates in 32-bit fixed point; in 3-D
// a cell's two components are packed in one 64-bit word (fx_pack), so that
// each tap is one ds_add_u64.  On gfx950 an LDS float64 atomic add at scattered
// addresses costs about 21 CU-cycles per wave64 instruction, a 64-bit
// integer one about 11, a 32-bit integer one about 7 (a float32 one about
// 194), and the LDS atomics are most of the scatter's time.
// Per subproblem and repetition the taps are scaled by 2^sh, sh = 31 - e with
// 2^(e-1) <= 1.001 A cmax < 2^e, A = sum over its samples of |Re y| + |Im y|
// and cmax bounding every coefficient product.  Each tap component is at
// most |Re y| + |Im y| times cmax, so every cell sum stays below
// 2^31 / 1.001 plus half a unit per point for the rounding: within int32
// for up to 2^20 points per subproblem (the integer adds wrap modulo 2^32,
// so only the final sums must fit).  Each tap is rounded to a step of at most
// 2^-30 1.001 A cmax.  The scale is a power of two, so scaling and unscaling
// are exact, and the integer sums do not depend on the order of the atomics.
// Non-finite samples make the whole tile NaN.  Without FXP (and for
// complex128) the tile accumulates in float64.
//
// REPS (subproblems of at most one point per thread): each thread computes
// its point's coefficients once, before the first repetition, and keeps them
// in registers for every repetition.  With the fixed-point tile the sample
// that the scale's block sum reads is also kept for the scatter, so the
// coefficient loads run beside the sample reads and no load follows the
// sum's barriers.  Taken for several repetitions, and for one with the
// fixed-point tile.
template <typename T, int NDIM, bool CK, int JMAX, bool FIXED, bool FXP, bool REPS = false, int RC = 1>
__global__ __launch_bounds__(256) void spread_tiled_kernel(TiledArgs<T, CK> a) {
    static_assert(!FXP || sizeof(T) == 4, "fixed-point accumulation is for complex64");
    static_assert(RC == 1 || !REPS, "the coefficient cache is for one repetition per pass");
    extern __shared__ __align__(16) unsigned char smem[];
    __shared__ float red[4];  // block_sum, at most 256 threads
    using Acc = typename std::conditional<FXP, int, double>::type;
    // FXP: packed 64-bit cells in 3-D (C3 scatter 0.424 -> 0.411 ms at 2,560
    // points per subproblem, 0.442 -> 0.403 ms at 3,072); two int32 planes in
    // 1-D and 2-D, where packing measured 0.3 us slower at Kd 512^2
    constexpr bool PACK = FXP && NDIM == 3;
    const InterpArgs<T, CK>& ia = a.ia;
    const BinGeom& bg = a.bg;
    // RC tiles, each its real then its imaginary plane
    Acc* const tiles = reinterpret_cast<Acc*>(smem);
    const AxisGeom ax[3] = {ia.ax[0], ia.ax[1], ia.ax[2]};
    const int T0 = bg.T[0], T1 = bg.T[1];
    const int tid = threadIdx.x;
    const int n_sub = a.bs.n_sub[0];
    unsigned natom = 0;  // global atomic lane operations of this thread's flushes
    for (int sid = blockIdx.x; sid < n_sub; sid += gridDim.x) {
        Sub s;
        load_sub(bg, a.bs, NDIM, sid, s);
        // REPS: this thread's point (sorted position ic), coefficients cached
        const bool cached = REPS && s.end - s.start <= (int)blockDim.x;
        const int ic = s.start + tid;
        PointTaps<T, CK, JMAX> pc;
        if constexpr (REPS)
            if (cached && ic < s.end) point_taps<T, NDIM, CK, JMAX, FIXED>(a, ax, s, ic, pc);
        // KEEP (1-D and 2-D, fixed point, one repetition per pass, no cache):
        // with at most two points per thread, the indices and samples that
        // the scale's sum reads are kept in registers for the scatter
        constexpr bool KEEP = FXP && !REPS && RC == 1 && NDIM < 3;
        const bool kept = KEEP && s.end - s.start <= 2 * (int)blockDim.x;
        const int i1 = ic + blockDim.x;
        // RC repetitions per pass, each with its own tile: one barrier round
        // and one coefficient computation per point serve all of them
        for (int r0 = 0; r0 < ia.nreps; r0 += RC) {
            const int rc = min(RC, ia.nreps - r0);  // uniform over the workgroup
            for (int e = tid; e < 2 * rc * bg.tile; e += blockDim.x) tiles[e] = 0;
            int sh[RC];     // FXP: tap values of repetition r0 + k are scaled by 2^sh[k]
            bool bad[RC];   // FXP: a non-finite sample
#pragma unroll
            for (int k = 0; k < RC; ++k) {
                sh[k] = 0;
                bad[k] = false;
            }
            cx<T> yc = mkcx<T>((T)0, (T)0);  // REPS, fixed point: this thread's sample of repetition r0
            int km0 = 0, km1 = 0;            // KEEP: indices and samples of points ic and i1
            cx<T> ky0 = yc, ky1 = yc;
            if constexpr (FXP) {
                float part[RC];
#pragma unroll
                for (int k = 0; k < RC; ++k) part[k] = 0.f;
                if (cached) {
                    if (ic < s.end) {
                        yc = ia.in[(int64_t)r0 * ia.M + pc.m];
                        part[0] = fabsf(yc.x) + fabsf(yc.y);
                    }
                } else if (kept) {
                    if (ic < s.end) {
                        km0 = a.bs.perm[ic];
                        ky0 = ia.in[(int64_t)r0 * ia.M + km0];
                        part[0] = fabsf(ky0.x) + fabsf(ky0.y);
                    }
                    if (i1 < s.end) {
                        km1 = a.bs.perm[i1];
                        ky1 = ia.in[(int64_t)r0 * ia.M + km1];
                        part[0] += fabsf(ky1.x) + fabsf(ky1.y);
                    }
                } else {
                    for (int i = s.start + tid; i < s.end; i += blockDim.x) {
                        const int64_t m = a.bs.perm[i];
#pragma unroll
                        for (int k = 0; k < RC; ++k)
                            if (k < rc) {
                                const cx<T> y = ia.in[(int64_t)(r0 + k) * ia.M + m];
                                part[k] += fabsf(y.x) + fabsf(y.y);  // |f| = |y| per component bound (|phase| = 1)
                            }
                    }
                }
#pragma unroll
                for (int k = 0; k < RC; ++k)
                    if (k < rc) {
                        const float A = block_sum(part[k], red) * a.cmax * 1.001f;
                        bad[k] = !(A <= 3.0e38f);
                        int e = 0;
                        (void)frexpf(A, &e);
                        sh[k] = 31 - e;
                    }
            }
            __syncthreads();
            // sample f (times conj(phase)) of repetition r0 + k at point pt
            auto add_sample = [&](int k, const PointTaps<T, CK, JMAX>& pt, cx<T> f) {
                Acc* tre = tiles + 2 * k * bg.tile;
                Acc* tim = tre + bg.tile;
                if constexpr (FXP) {
                    if (bad[k]) return;
                    f.x = ldexpf(f.x, sh[k]);
                    f.y = ldexpf(f.y, sh[k]);
                    scatter_taps_fx<NDIM, CK, JMAX, FIXED, PACK>(tre, tim, pt, f, ax, T0, T1);
                } else {
                    scatter_taps<T, NDIM, CK, JMAX, FIXED>(tre, tim, pt, f, ax, T0, T1);
                }
            };
            auto add_point = [&](const PointTaps<T, CK, JMAX>& pt) {
#pragma unroll
                for (int k = 0; k < RC; ++k) {
                    if (k >= rc) break;
                    add_sample(k, pt, sample_times_conj_phase<T, CK, JMAX>(ia, r0 + k, pt, a.phs));
                }
            };
            if (cached) {
                if (ic < s.end) {
                    if constexpr (FXP) add_sample(0, pc, times_conj_phase<T, CK, JMAX>(yc, pc, a.phs));
                    else add_point(pc);
                }
            } else if (kept) {
#pragma unroll
                for (int u = 0; u < 2; ++u) {
                    const int i = u ? i1 : ic;
                    if (i < s.end) {
                        T t[3];
#pragma unroll
                        for (int d = 0; d < NDIM; ++d) t[d] = a.tms[(int64_t)d * ia.M + i];
                        const cx<T> ph = a.phs ? a.phs[i] : mkcx<T>((T)1, (T)0);
                        PointTaps<T, CK, JMAX> pt;
                        point_taps_pre<T, NDIM, CK, JMAX, FIXED>(a, ax, s, t, u ? km1 : km0, ph, pt);
                        add_sample(0, pt, times_conj_phase<T, CK, JMAX>(u ? ky1 : ky0, pt, a.phs));
                    }
                }
            } else {
                for (int i = s.start + tid; i < s.end; i += blockDim.x) {
                    PointTaps<T, CK, JMAX> pt;
                    point_taps<T, NDIM, CK, JMAX, FIXED>(a, ax, s, i, pt);
                    add_point(pt);
                }
            }
            __syncthreads();
#pragma unroll
            for (int k = 0; k < RC; ++k) {
                if (k >= rc) break;
                const int r = r0 + k;
                const Acc* tre = tiles + 2 * k * bg.tile;
                const Acc* tim = tre + bg.tile;
                const int shk = sh[k];
                const bool badk = bad[k];
                auto value = [&](int e, T& vr, T& vi) {  // false: an empty cell
                    if constexpr (FXP) {
                        int ir, ii;
                        if constexpr (PACK) {
                            fx_unpack(reinterpret_cast<const long long*>(tre)[e], ir, ii);
                        } else {
                            ir = tre[e];
                            ii = tim[e];
                        }
                        if (badk) {
                            vr = vi = __int_as_float(0x7fc00000);
                            return true;
                        }
                        if (ir == 0 && ii == 0) return false;
                        vr = ldexpf((float)ir, -shk);
                        vi = ldexpf((float)ii, -shk);
                    } else {
                        const double ar = tre[e], ai = tim[e];
                        if (ar == 0.0 && ai == 0.0) return false;
                        vr = (T)ar;
                        vi = (T)ai;
                    }
                    return true;
                };
                if (a.tiles) {  // two-phase: write the whole tile to its slot, no phase
                    cx<T>* __restrict__ slot = a.tiles + ((int64_t)sid * ia.nreps + r) * bg.tile;
                    for (int e = tid; e < bg.tile; e += blockDim.x) {
                        T vr = 0, vi = 0;
                        (void)value(e, vr, vi);
                        slot[e] = mkcx<T>(vr, vi);
                    }
                    continue;
                }
                cx<T>* __restrict__ g = ia.grid_out + ((int64_t)s.frame * ia.nreps + r) * ia.kstride;
                // A bin's own cells past the previous bin's halo (local e_d in
                // [J_d - 1, B_d) on every axis) lie in no other tile when every
                // K_d is a multiple of B_d; when the bin has a single
                // subproblem they are written with plain stores (the grid is
                // zero there), everything else with atomics.
                const bool own = a.own_ok && s.solo;
                for (int e = tid; e < bg.tile; e += blockDim.x) {
                    T vr, vi;
                    if (!value(e, vr, vi)) continue;
                    int g0, g1, g2, e0, e1, e2;
                    const int64_t gi = tile_to_grid<NDIM>(e, bg, s, ax, g0, g1, g2, e0, e1, e2);
                    if (a.use_pb) {  // * conj(phase_before)
                        const T c[3] = {a.pb[0], a.pb[1], a.pb[2]};
                        const cx<T> p = cis_tile(phase_before_arg<T>(c, NDIM, g0, g1, g2));
                        const T wr = vr * p.x + vi * p.y;
                        const T wi = vi * p.x - vr * p.y;
                        vr = wr;
                        vi = wi;
                    }
                    bool mine = own && e0 >= ax[0].J - 1 && e0 < bg.B[0];
                    if (NDIM > 1) mine = mine && e1 >= ax[1].J - 1 && e1 < bg.B[1];
                    if (NDIM > 2) mine = mine && e2 >= ax[2].J - 1 && e2 < bg.B[2];
                    if (mine) {
                        g[gi] = mkcx<T>(vr, vi);
                    } else {
                        T* q = reinterpret_cast<T*>(g + gi);
                        unsafeAtomicAdd(q, vr);
                        unsafeAtomicAdd(q + 1, vi);
                        natom += 2;
                    }
                }
            }
            __syncthreads();
        }
    }
    if (a.flush_count) {  // uniform: one 64-bit atomic per wave, into this workgroup's slot
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) natom += __shfl_xor(natom, o);
        if ((tid & 63) == 0) atomicAdd(a.flush_count + (blockIdx.x & (a.flush_slots - 1)), (unsigned long long)natom);
    }
}

// Trajectory gradient.  For y[r, m] = phase[m] * sum_k psi(t_m - k) G_r[k]
// (G_r the staged grid), t_d = omega_d * K_d / (2 pi), phase = exp(i omega . s):
//   grad[m, d] = sum_r Re(conj(w[r, m]) * dy[r, m] / d omega_d)
//   dy/d omega_d = phase * (K_d / 2 pi) * sum_k d_d psi(t_m - k) G_r[k] + i s_d y.
// Every point belongs to exactly one subproblem, so its thread owns grad[m, :]
// and accumulates the repetitions without atomics.
template <typename T>
struct GradArgs {
    const cx<T>* w;   // (nreps, M) upstream gradient
    T* grad;          // (M, ndim) output, row-major like omega
    T dtdw[3];        // K_d / (2 pi)
    T shift[3];       // phase exponent coefficients s_d (0 if no phase)
};

template <typename T, bool CK>
__device__ __forceinline__ void cmadd(T& ar, T& ai, T cr, T ci, T vr, T vi) {
    madd<T, CK>(ar, ai, cr, ci, vr, vi);
}

// MULTI: a.rc repetitions' tiles are staged per pass and each point's
// coefficients and derivatives are computed once for all of them.
template <typename T, int NDIM, bool CK, int JMAX, bool FIXED, bool MULTI>
__global__ __launch_bounds__(256) void grad_tiled_kernel(TiledArgs<T, CK> a, GradArgs<T> gr) {
    extern __shared__ __align__(16) unsigned char smem[];
    cx<T>* tiles = reinterpret_cast<cx<T>*>(smem);
    const InterpArgs<T, CK>& ia = a.ia;
    const BinGeom& bg = a.bg;
    const AxisGeom ax[3] = {ia.ax[0], ia.ax[1], ia.ax[2]};
    const int T0 = bg.T[0], T1 = bg.T[1];
    const int tid = threadIdx.x;
    const int n_sub = a.bs.n_sub[0];
    const int RC = MULTI ? a.rc : 1;
    for (int sid = blockIdx.x; sid < n_sub; sid += gridDim.x) {
    Sub s;
    load_sub(bg, a.bs, NDIM, sid, s);
    for (int r0 = 0; r0 < ia.nreps; r0 += RC) {
        const int rc = MULTI ? min(RC, ia.nreps - r0) : 1;
        for (int rr = 0; rr < rc; ++rr)
            stage_tile<T, NDIM, CK>(a, s, ax, ia.grid + ((int64_t)s.frame * ia.nreps + r0 + rr) * ia.kstride,
                                    tiles + (int64_t)rr * bg.tile);
        __syncthreads();
        for (int i = s.start + tid; i < s.end; i += blockDim.x) {
            const int64_t m = a.bs.perm[i];
            T cr[3][JMAX], ci[3][JMAX], dr[3][JMAX], di[3][JMAX];
            int l[3];
#pragma unroll
            for (int d = 0; d < NDIM; ++d) {
                const T t = a.tms[(int64_t)d * ia.M + i];
                int k0;
                if constexpr (CK) k0 = axis_coefs_d<T, JMAX, FIXED>(t, ia.L, ax[d], ia.h[d], cr[d], ci[d], dr[d], di[d]);
                else k0 = axis_coefs_d<T, JMAX, FIXED>(t, ia.L, ax[d], ia.h[d], cr[d], dr[d]);
                l[d] = k0 - s.o[d];
            }
            cx<T> ph = mkcx<T>((T)1, (T)0);
            if (a.phs) ph = a.phs[i];
            T acc[3] = {0, 0, 0};  // sum over this pass's repetitions
            for (int rr = 0; rr < rc; ++rr) {
                const cx<T>* tile = tiles + (int64_t)rr * bg.tile;
                // S[0] = value, S[1 + d] = derivative along axis d
                T Sr[4] = {0, 0, 0, 0}, Si[4] = {0, 0, 0, 0};
                if constexpr (NDIM == 1) {
#pragma unroll
                    for (int j1 = 0; j1 < JMAX; ++j1)
                        if (FIXED || j1 < ax[0].J) {
                            const cx<T> v = tile[l[0] + j1];
                            cmadd<T, CK>(Sr[0], Si[0], cr[0][j1], ci[0][j1], v.x, v.y);
                            cmadd<T, CK>(Sr[1], Si[1], dr[0][j1], di[0][j1], v.x, v.y);
                        }
                } else if constexpr (NDIM == 2) {
#pragma unroll
                    for (int j2 = 0; j2 < JMAX; ++j2)
                        if (FIXED || j2 < ax[1].J) {
                            const cx<T>* row = tile + (l[1] + j2) * T0 + l[0];
                            T ar = 0, ai = 0, br = 0, bi = 0;
#pragma unroll
                            for (int j1 = 0; j1 < JMAX; ++j1)
                                if (FIXED || j1 < ax[0].J) {
                                    const cx<T> v = row[j1];
                                    cmadd<T, CK>(ar, ai, cr[0][j1], ci[0][j1], v.x, v.y);
                                    cmadd<T, CK>(br, bi, dr[0][j1], di[0][j1], v.x, v.y);
                                }
                            cmadd<T, CK>(Sr[0], Si[0], cr[1][j2], ci[1][j2], ar, ai);
                            cmadd<T, CK>(Sr[1], Si[1], cr[1][j2], ci[1][j2], br, bi);
                            cmadd<T, CK>(Sr[2], Si[2], dr[1][j2], di[1][j2], ar, ai);
                        }
                } else {
#pragma unroll FIXED ? JMAX : 1
                    for (int j3 = 0; j3 < JMAX; ++j3)
                        if (FIXED || j3 < ax[2].J) {
                            T pr = 0, pi = 0, qr = 0, qi = 0, ur = 0, ui = 0;
#pragma unroll
                            for (int j2 = 0; j2 < JMAX; ++j2)
                                if (FIXED || j2 < ax[1].J) {
                                    const cx<T>* row = tile + ((l[2] + j3) * T1 + (l[1] + j2)) * T0 + l[0];
                                    T ar = 0, ai = 0, br = 0, bi = 0;
#pragma unroll
                                    for (int j1 = 0; j1 < JMAX; ++j1)
                                        if (FIXED || j1 < ax[0].J) {
                                            const cx<T> v = row[j1];
                                            cmadd<T, CK>(ar, ai, cr[0][j1], ci[0][j1], v.x, v.y);
                                            cmadd<T, CK>(br, bi, dr[0][j1], di[0][j1], v.x, v.y);
                                        }
                                    cmadd<T, CK>(pr, pi, cr[1][j2], ci[1][j2], ar, ai);
                                    cmadd<T, CK>(qr, qi, cr[1][j2], ci[1][j2], br, bi);
                                    cmadd<T, CK>(ur, ui, dr[1][j2], di[1][j2], ar, ai);
                                }
                            cmadd<T, CK>(Sr[0], Si[0], cr[2][j3], ci[2][j3], pr, pi);
                            cmadd<T, CK>(Sr[1], Si[1], cr[2][j3], ci[2][j3], qr, qi);
                            cmadd<T, CK>(Sr[2], Si[2], cr[2][j3], ci[2][j3], ur, ui);
                            cmadd<T, CK>(Sr[3], Si[3], dr[2][j3], di[2][j3], pr, pi);
                        }
                }
                // y = S0 * ph
                const T yr = Sr[0] * ph.x - Si[0] * ph.y;
                const T yi = Sr[0] * ph.y + Si[0] * ph.x;
                const cx<T> w = gr.w[(int64_t)(r0 + rr) * ia.M + m];
#pragma unroll
                for (int d = 0; d < NDIM; ++d) {
                    const T er = (Sr[1 + d] * ph.x - Si[1 + d] * ph.y) * gr.dtdw[d] - gr.shift[d] * yi;
                    const T ei = (Sr[1 + d] * ph.y + Si[1 + d] * ph.x) * gr.dtdw[d] + gr.shift[d] * yr;
                    const T v = w.x * er + w.y * ei;
                    acc[d] = (rr == 0) ? v : acc[d] + v;
                }
            }
#pragma unroll
            for (int d = 0; d < NDIM; ++d) {
                T* o = gr.grad + m * NDIM + d;
                *o = (r0 == 0) ? acc[d] : *o + acc[d];
            }
        }
        __syncthreads();
    }
    }
}

// Second phase of the two-phase scatter: every grid point sums, in a fixed
// order, the tile slots of the subproblems whose tiles cover it, multiplies by
// conj(phase_before) and writes the grid (every point is written, so the grid
// needs no clearing; the result does not depend on scheduling).  The bins
// covering grid index k along an axis are those whose tile [o, o + T) (mod K)
// contains k: local offset e = (k - o) mod K (+ K, 2K, ... for tiles wider than
// the grid) below T.
template <typename T, int NDIM>
__global__ __launch_bounds__(256) void tile_reduce_kernel(const cx<T>* __restrict__ tiles, cx<T>* __restrict__ grid,
                                                          BinGeom bg, BinState bs, int K0, int K1, int K2,
                                                          int64_t nk, int nreps, T pb0, T pb1, T pb2, int use_pb,
                                                          int j0, int j1, int j2) {
    const int jmax[3] = {j0, j1, j2};
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nk) return;
    const int img = blockIdx.y;            // frame * nreps + rep
    const int frame = img / nreps;
    const int r = img - frame * nreps;
    const int K[3] = {K0, K1, K2};
    int k[3];
    {
        const uint32_t u = (uint32_t)i;
        k[0] = (int)(u % (uint32_t)K0);
        const uint32_t q = u / (uint32_t)K0;
        k[1] = (int)(q % (uint32_t)K1);
        k[2] = (int)(q / (uint32_t)K1);
    }
    // (bin, local offset) pairs covering k along each axis: bins at or below
    // the one containing k (periodically, jmax of them, set on the host), and
    // every offset e, e + K, ... inside the tile
    int cb[3][8], ce[3][8], nc[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
        nc[d] = 0;
        if (d >= NDIM) {
            cb[d][0] = 0;
            ce[d][0] = 0;
            nc[d] = 1;
            continue;
        }
        const int B = bg.B[d], nb = bg.nb[d], Tt = bg.T[d];
        const int bk = k[d] / B;
        for (int j = 0; j < jmax[d]; ++j) {
            int b = bk - j;
            if (b < 0) b += nb;
            int e = k[d] - b * B;
            if (e < 0) e += K[d];
            for (; e < Tt && nc[d] < 8; e += K[d]) {
                cb[d][nc[d]] = b;
                ce[d][nc[d]] = e;
                ++nc[d];
            }
        }
    }
    double sr = 0.0, si = 0.0;
    for (int a2 = 0; a2 < nc[2]; ++a2)
        for (int a1 = 0; a1 < nc[1]; ++a1)
            for (int a0 = 0; a0 < nc[0]; ++a0) {
                const int bin = frame * bg.nbins_frame + (cb[2][a2] * bg.nb[1] + cb[1][a1]) * bg.nb[0] + cb[0][a0];
                const int cnt = bs.counts[bin];
                if (cnt == 0) continue;
                const int nsb = (cnt + bg.P - 1) / bg.P;
                const int s0 = bs.bin_sub[bin];
                const int e = (ce[2][a2] * bg.T[1] + ce[1][a1]) * bg.T[0] + ce[0][a0];
                for (int sub = s0; sub < s0 + nsb; ++sub) {
                    const cx<T> v = tiles[((int64_t)sub * nreps + r) * bg.tile + e];
                    sr += (double)v.x;
                    si += (double)v.y;
                }
            }
    T vr = (T)sr, vi = (T)si;
    if (use_pb) {  // * conj(phase_before)
        const T c[3] = {pb0, pb1, pb2};
        const cx<T> p = cis_tile(phase_before_arg<T>(c, NDIM, k[0], k[1], k[2]));
        const T wr = vr * p.x + vi * p.y;
        const T wi = vi * p.x - vr * p.y;
        vr = wr;
        vi = wi;
    }
    grid[(int64_t)img * nk + i] = mkcx<T>(vr, vi);
}

}  // namespace mi355
