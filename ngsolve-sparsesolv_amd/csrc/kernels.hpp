// kernels.hpp -- gfx950 kernels of the ABMC-IC(0) Krylov path.
//
// All vectors are in ABMC order.  Every kernel of the iteration reads the
// solver state first and returns at once when the solve has stopped, so a
// captured iteration can be replayed past convergence without effect.
//
// Reductions are deterministic: each workgroup writes one partial (wave
// butterfly + fixed-order sum over the waves); the partials are summed in
// index order by a control kernel or by the producing kernel's last
// workgroup (control.hpp).
#pragma once

#include <type_traits>

#include "control.hpp"
#include "device_math.hpp"
#include "solver_state.hpp"

namespace ssa {

constexpr int kBlock = 256;          // threads per workgroup (4 wavefronts)
static_assert(kBlock == kCtrlThreads, "fused control steps reduce with kBlock threads");
constexpr index_t kPadColumn = -(1 << 30);   // padding slot of a sweep layout
constexpr int kWavesPerBlock = kBlock / 64;

// ---------------------------------------------------------------------------
// Workgroup order across the 8 XCDs.  The hardware hands workgroup b of a
// launch to XCD b % 8; consecutive slices of the SpMV and of a colour gather
// neighbouring rows, so the kernels that gather work on the logical block
// xcd_block(b, G): each XCD takes one contiguous range of logical blocks and
// its L2 sees the neighbours together.  A bijection of [0, G); partials stay
// indexed by the logical block, so results do not depend on the mapping.
// Used by the one-vector kernels (measured: SpMV 27.2 -> 25.9 us at 64^3);
// the 16-wide block kernels ran 3-9% slower with it and keep the launch
// order.
// ---------------------------------------------------------------------------
__device__ inline unsigned xcd_block(unsigned b, unsigned G)
{
    const unsigned per = (G + 7u) / 8u;          // logical blocks of a full XCD
    const unsigned full = G - (per - 1u) * 8u;   // XCDs that take `per` (the rest per - 1)
    const unsigned x = b % 8u, k = b / 8u;
    return x < full ? x * per + k : full * per + (x - full) * (per - 1u) + k;
}

// ---------------------------------------------------------------------------
// Blocks of W right-hand sides: every vector holds n rows of W values,
// element (i, r) at v[i*W + r], so a lane's W values are one contiguous,
// 16-byte aligned run (W >= 4) moved as dwordx4 loads and stores.  W = 1 is
// the one-vector layout.
// ---------------------------------------------------------------------------
template <typename P, int W>
__device__ inline void load_row(const P* __restrict__ p, P (&d)[W])
{
    if constexpr (W == 1) {
        d[0] = p[0];
    } else {
        static_assert((W * sizeof(P)) % 16 == 0, "a row is whole dwordx4 words");
        constexpr int kWords = int(W * sizeof(P) / 16);
        uint4 w[kWords];
        const uint4* q = reinterpret_cast<const uint4*>(p);
#pragma unroll
        for (int k = 0; k < kWords; ++k) w[k] = q[k];
        __builtin_memcpy(d, w, sizeof(w));
    }
}

template <typename P, int W>
__device__ inline void store_row(P* p, const P (&d)[W])
{
    if constexpr (W == 1) {
        p[0] = d[0];
    } else {
        constexpr int kWords = int(W * sizeof(P) / 16);
        uint4 w[kWords];
        __builtin_memcpy(w, d, sizeof(w));
        uint4* q = reinterpret_cast<uint4*>(p);
#pragma unroll
        for (int k = 0; k < kWords; ++k) q[k] = w[k];
    }
}

// W values summed over the workgroup with block_sum's tree for each of them
// (wave butterfly, then the waves in order); valid in threads 0..W-1, thread
// r holding value r.  `scratch` holds kWavesPerBlock*W values.
template <typename T, int W>
__device__ inline T block_sum_multi(const T (&v)[W], T* scratch)
{
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    T w[W];
#pragma unroll
    for (int r = 0; r < W; ++r) w[r] = wave_sum(v[r]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int r = 0; r < W; ++r) scratch[wid * W + r] = w[r];
    }
    __syncthreads();
    T s = zero<T>();
    if (threadIdx.x < W) {
        for (int q = 0; q < kWavesPerBlock; ++q) s = add(s, scratch[q * W + threadIdx.x]);
    }
    return s;
}

// ---------------------------------------------------------------------------
// SpMV over SELL-64 (one wavefront per 64-row slice; lane l of slice s
// computes row rows[s*64 + l], -1: none), optional fused dot
// DOT: 0 none, 1 sum x_i y_i, 2 sum conj(x_i) y_i
//
// W right-hand sides (interleaved rows): each lane loads a matrix entry once
// and applies it to all W vectors; the partials of right-hand side r are
// [r*gridDim.x, (r+1)*gridDim.x).  Right-hand sides that have stopped are
// still multiplied (their results are not read).  The fused convergence
// check (`check`) is for one vector only.  With `stream_nt` (one vector) the
// matrix entries load with the non-temporal hint.  x may be stored narrower (TX:
// the preconditioner's precision); its values are widened exactly before
// they are multiplied, so the products are those of the widened vector.
// ---------------------------------------------------------------------------
template <typename T, int DOT, int W, typename TX = T>
__global__ __launch_bounds__(kBlock) void k_spmv(
    const index_t* __restrict__ slice_off, const T* __restrict__ val,
    const index_t* __restrict__ col, const index_t* __restrict__ rows, index_t nslices,
    const TX* __restrict__ x, T* __restrict__ y, T* __restrict__ partials,
    const SolverState* __restrict__ st, Ctrl check, Systems sy, int stream_nt)
{
    const unsigned lb = W == 1 ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x;
    const int sys = sy.spmv_sys ? sy.spmv_sys[lb] : 0;
    int live = 0;   // some right-hand side (or the batch system) runs
#pragma unroll
    for (int r = 0; r < W; ++r) live |= !st || !st[sys + r].status;
    __shared__ T red[kWavesPerBlock * W];
    const int lane = threadIdx.x & 63;
    const index_t s = index_t(lb) * kWavesPerBlock + (threadIdx.x >> 6);
    const index_t sc = s < nslices ? s : nslices;
    const index_t o0 = slice_off[sc], o1 = s < nslices ? slice_off[s + 1] : o0;
    const index_t row = s < nslices ? rows[size_t(s) * 64 + lane] : -1;
    if (!live) {
        // consuming the offset keeps its load ahead of the status test
        // (offsets are >= 0)
        if (o0 < 0) y[0] = zero<T>();
        return;
    }
    T d[W];
#pragma unroll
    for (int r = 0; r < W; ++r) d[r] = zero<T>();
    if (s < nslices) {
        T acc[W];
#pragma unroll
        for (int r = 0; r < W; ++r) acc[r] = zero<T>();
        if constexpr (W == 1) {
            // two copies of the loop, so the hinted and the plain loads are
            // never merged into one
            const auto rows_loop = [&](auto nt) {
#pragma unroll 4
                for (index_t k = o0; k < o1; ++k) {
                    const size_t at = size_t(k) * 64 + lane;
                    const T v = nt.value ? load_nt(val + at) : val[at];
                    const index_t c = nt.value ? load_nt(col + at) : col[at];
                    acc[0] = fma_(v, cvt<T>(x[c]), acc[0]);
                }
            };
            if (stream_nt) rows_loop(std::true_type{});
            else rows_loop(std::false_type{});
        } else {
            // one entry's W gathered values fill the registers: the next
            // entry's column and value load while this entry's gathers are
            // in flight, so an entry costs one memory round trip, not two
            // (K is uniform over the slice, so the guard is a scalar branch)
            index_t cn = 0;
            T vn = zero<T>();
            if (o0 < o1) {
                cn = col[size_t(o0) * 64 + lane];
                vn = val[size_t(o0) * 64 + lane];
            }
            for (index_t k = o0; k < o1; ++k) {
                const index_t c = cn;
                const T v = vn;
                if (k + 1 < o1) {
                    const size_t at = size_t(k + 1) * 64 + lane;
                    cn = col[at];
                    vn = val[at];
                }
                TX g[W];
                load_row<TX, W>(x + size_t(c) * W, g);
#pragma unroll
                for (int r = 0; r < W; ++r) acc[r] = fma_(v, cvt<T>(g[r]), acc[r]);
            }
        }
        if (row >= 0) {
            store_row<T, W>(y + size_t(row) * W, acc);
            if (DOT) {
                TX xr[W];
                load_row<TX, W>(x + size_t(row) * W, xr);
#pragma unroll
                for (int r = 0; r < W; ++r) d[r] = mul(conj_if(cvt<T>(xr[r]), DOT == 2), acc[r]);
            }
        }
    }
    if (DOT) {
        if constexpr (W == 1) {
            T t = block_sum(d[0], red);
            if (threadIdx.x == 0) partials[lb] = t;
        } else {
            T t = block_sum_multi<T, W>(d, red);
            if (threadIdx.x < W) partials[size_t(threadIdx.x) * gridDim.x + lb] = t;
        }
    }
    if (W == 1 && check.part) {
        __shared__ double redd[kWavesPerBlock];
        const index_t local = index_t(lb) - (sy.spmv_g0 ? sy.spmv_g0[sys] : 0);
        ctrl_check_wg0(ctrl_for(check, sy, sys, local), redd);
    }
}

// ---------------------------------------------------------------------------
// Triangular sweep, regular blocks.  One lane per row; the rows of an ABMC
// block sit in consecutive lanes in processing order.
//   FWD: out_i = (in_i - sum L_ik out_k) * d_i^-1     (L below the diagonal)
//   BWD: out_i = in_i - d_i^-1 * sum U_ik out_k       (U = L^T above it)
// Columns >= 0 are finished rows of earlier colours: every lane loads its
// entries and gathers those values CH at a time (independent loads, so one
// or two memory round trips per colour).  Columns < 0 name an earlier
// position of the same block; those coefficients are kept in registers and
// applied once that position is final, its value broadcast with a
// wavefront shuffle (BR steps).
// ---------------------------------------------------------------------------
// External part of the oversized blocks of this phase (walk slots t0..t1):
//   FWD: a = in_i - sum_ext L_ik out_k      BWD: a = -sum_ext U_ik out_k
// written to sa[t] (and in_i to sb[t] for BWD); the walk kernel finishes them.
// Right-hand side r of a block uses the slots [r*ws, (r+1)*ws).
struct WalkExt {
    size_t t0, t1;
    const index_t* row;
    const index_t* ext_rp;
    const index_t* ext_ci;
    const void* ext_v;
    void* sa;
    void* sb;
    size_t ws;
};

template <typename P, typename F, bool FWD, int W>
__device__ inline void walk_external(const WalkExt& w, size_t t, const P* __restrict__ in,
                                     const P* out)
{
    if (t >= w.t1) return;
    const index_t row = w.row[t];
    if (row < 0) return;
    const F* ev = static_cast<const F*>(w.ext_v);
    P xin[W], a[W];
    load_row<P, W>(in + size_t(row) * W, xin);
#pragma unroll
    for (int r = 0; r < W; ++r) a[r] = FWD ? xin[r] : zero<P>();
    for (index_t k = w.ext_rp[t]; k < w.ext_rp[t + 1]; ++k) {
        const P v = cvt<P>(ev[k]);
        P g[W];
        load_row<P, W>(out + size_t(w.ext_ci[k]) * W, g);
#pragma unroll
        for (int r = 0; r < W; ++r) a[r] = fms_(v, g[r], a[r]);
    }
#pragma unroll
    for (int r = 0; r < W; ++r) {
        static_cast<P*>(w.sa)[r * w.ws + t] = a[r];
        if (!FWD) static_cast<P*>(w.sb)[r * w.ws + t] = xin[r];
    }
}

template <typename P>
__device__ inline P shfl(P v, int src);
template <>
__device__ inline double shfl<double>(double v, int src) { return __shfl(v, src, 64); }
template <>
__device__ inline float shfl<float>(float v, int src) { return __shfl(v, src, 64); }
template <>
__device__ inline c128 shfl<c128>(c128 v, int src) { return c128{__shfl(v.x, src, 64), __shfl(v.y, src, 64)}; }
template <>
__device__ inline c64 shfl<c64>(c64 v, int src) { return c64{__shfl(v.x, src, 64), __shfl(v.y, src, 64)}; }

struct SweepArrays {
    const index_t* lane_row;
    const index_t* lane_info;   // position in the block | slice's entry slots << 8
    const void* val;
    const index_t* col;
    const void* invd;
    const int* slice_sys;   // batches: system of every slice (null: one system)
    int stream_nt;          // entries loaded with the non-temporal hint
};

// One wavefront, one slice, whose entries start at slot o0 (wavefront-
// uniform, from the colour's first slot and stride: no load).  The lane's
// row and info word and the first batch of entries -- as many as the
// colour's stride, the longest slice -- are requested before `stopped` is
// tested, so the chain of dependent memory round trips is entries ->
// gathers.  The info word gives the slice's own entry count K: only the
// first K entries are gathered.  W right-hand sides (interleaved rows):
// each entry is loaded once and applied to all of them; CH entries per
// batch.
template <typename P, typename F, int BR, bool FWD, int CH, int W, bool NT = false>
__device__ inline void sweep_slice(index_t s, index_t o0, index_t stride, const SweepArrays& A,
                                   const P* __restrict__ in, P* out, int stopped)
{
    const F* __restrict__ val = static_cast<const F*>(A.val);
    const index_t* __restrict__ col = A.col;
    const int lane = threadIdx.x & 63;
    const size_t li = size_t(s) * 64 + lane;
    const index_t row = A.lane_row[li];
    const index_t info = A.lane_info[li];
    // slots in [o0, o0 + stride) are this slice's, padding past its K
    index_t c[CH];
    F v[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        c[j] = kPadColumn;
        v[j] = zero<F>();
        if (j < stride) {
            const size_t at = (size_t(o0) + j) * 64 + lane;
            c[j] = NT ? load_nt(col + at) : col[at];
            v[j] = NT ? load_nt(val + at) : val[at];
        }
    }
    if (stopped) {
        // consuming the info word and the first entry on this path keeps
        // their loads ahead of the status test (the compiler would otherwise
        // sink them into the only branch that uses them); info is >= 0
        if (info < 0) out[0] = cvt<P>(v[0]);
        return;
    }
    const int pos = info & 255;
    // the same in every lane of the slice
    const index_t o1 = o0 + __builtin_amdgcn_readfirstlane(info >> 8);
    const bool live = row >= 0;
    const index_t rr = live ? row : 0;
    // the backward sweep needs in_i only at the end: with several
    // right-hand sides it is loaded there (registers), with one up front
    constexpr bool kLateIn = !FWD && W > 1;
    P xin[W], acc[W];
    if (kLateIn) {
#pragma unroll
        for (int r = 0; r < W; ++r) xin[r] = zero<P>();
    } else {
        load_row<P, W>(in + size_t(rr) * W, xin);
    }
#pragma unroll
    for (int r = 0; r < W; ++r) acc[r] = FWD ? xin[r] : zero<P>();
    const P dinv = cvt<P>(static_cast<const F*>(A.invd)[rr]);
    F intv[BR > 1 ? BR - 1 : 1];
#pragma unroll
    for (int j = 0; j < (BR > 1 ? BR - 1 : 1); ++j) intv[j] = zero<F>();
    for (index_t k0 = o0;;) {
        // gathers from row 0 stand in for non-external entries; their
        // values are discarded
        P g[CH][W];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            if (k0 + j < o1) {
                const index_t cj = c[j] >= 0 ? c[j] : 0;
                load_row<P, W>(out + size_t(cj) * W, g[j]);
            } else {
                c[j] = kPadColumn;
#pragma unroll
                for (int r = 0; r < W; ++r) g[j][r] = zero<P>();
            }
        }
        // When a slice can take several chunks (several right-hand sides,
        // or chunks of 8 complex values), the next chunk's entries load
        // while this chunk's gathers are in flight (one memory round trip
        // per chunk instead of two).  The guard is uniform over the slice.
        constexpr bool kAhead = W > 1 || CH < 16;
        const index_t k1 = k0 + CH;
        index_t cn[CH];
        F vn[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            cn[j] = kPadColumn;
            vn[j] = zero<F>();
        }
        if (kAhead && k1 < o1) {
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                if (k1 + j < o1) {
                    const size_t at = (size_t(k1) + j) * 64 + lane;
                    cn[j] = NT ? load_nt(col + at) : col[at];
                    vn[j] = NT ? load_nt(val + at) : val[at];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const P vj = cvt<P>(v[j]);
#pragma unroll
            for (int r = 0; r < W; ++r) {
                const P t = fms_(vj, g[j][r], acc[r]);
                acc[r] = select_(c[j] >= 0, t, acc[r]);
            }
            // -1 - c is an earlier position of this block; padding slots
            // carry kPadColumn and match no position
#pragma unroll
            for (int q = 0; q < BR - 1; ++q) intv[q] = select_(c[j] == -1 - q, v[j], intv[q]);
        }
        k0 = k1;
        if (k0 >= o1) break;
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            if (kAhead) {
                c[j] = cn[j];
                v[j] = vn[j];
            } else if (k0 + j < o1) {
                const size_t at = (size_t(k0) + j) * 64 + lane;
                c[j] = NT ? load_nt(col + at) : col[at];
                v[j] = NT ? load_nt(val + at) : val[at];
            }
        }
    }
    if (kLateIn) load_row<P, W>(in + size_t(rr) * W, xin);
    const int base = lane - pos;
    P res[W];
#pragma unroll
    for (int r = 0; r < W; ++r) res[r] = zero<P>();
#pragma unroll
    for (int j = 0; j < BR; ++j) {
        if (live && pos == j) {
#pragma unroll
            for (int r = 0; r < W; ++r) res[r] = FWD ? mul(acc[r], dinv) : fma_(acc[r], dinv, xin[r]);
            store_row<P, W>(out + size_t(row) * W, res);
        }
        if (j < BR - 1) {
            // position j of this lane's block (clamped for idle lanes and
            // short blocks at the end of a slice; those lanes do not use it)
            const int src = min(base + j, 63);
            const P cj = cvt<P>(intv[j]);
#pragma unroll
            for (int r = 0; r < W; ++r) {
                const P vj = shfl(res[r], src);
                if (pos > j) acc[r] = fms_(cj, vj, acc[r]);
            }
        }
    }
}

// A row's W values as L lanes hold them: lane q of the row takes the
// 16-byte words q, q + L, q + 2L, ... of the row (RL = W / L values), so
// one load instruction reads 16 L contiguous bytes of every row it touches.
template <typename P, int W, int L>
__device__ inline void load_part(const P* __restrict__ row, int q, P (&d)[W / L])
{
    constexpr int kWords = int(W * sizeof(P) / 16) / L;
    static_assert(kWords * L * 16 == int(W * sizeof(P)), "a row is L lanes of whole dwordx4 words");
    uint4 w[kWords];
    const uint4* p = reinterpret_cast<const uint4*>(row) + q;
#pragma unroll
    for (int k = 0; k < kWords; ++k) w[k] = p[k * L];
    __builtin_memcpy(d, w, sizeof(w));
}

template <typename P, int W, int L>
__device__ inline void store_part(P* row, int q, const P (&d)[W / L])
{
    constexpr int kWords = int(W * sizeof(P) / 16) / L;
    uint4 w[kWords];
    __builtin_memcpy(w, d, sizeof(w));
    uint4* p = reinterpret_cast<uint4*>(row) + q;
#pragma unroll
    for (int k = 0; k < kWords; ++k) p[k * L] = w[k];
}

// sweep_slice with L lanes a row: wavefront `part` of the slice takes its
// lanes part*64/L .. (part+1)*64/L - 1 (whole blocks: the layout keeps a
// block within kBlockAlign = 16 lanes, and 64 / L is a multiple of 16),
// lane rl * L + q of the wavefront holding row rl's words q, q + L, ...
// With W values of a row in one lane, a gather instruction touches a cache
// line per lane; here L lanes share one, so the CUs' address units move L
// times the bytes per instruction.  Each right-hand side sees the same
// operations in the same order as in sweep_slice (entries in slot order,
// then the block's earlier positions), so the results are bit for bit.
template <typename P, typename F, int BR, bool FWD, int CH, int W, int L>
__device__ inline void sweep_slice_split(index_t s, int part, index_t o0, index_t stride,
                                         const SweepArrays& A, const P* __restrict__ in, P* out,
                                         int stopped)
{
    constexpr int RL = W / L, RPW = 64 / L;
    const F* __restrict__ val = static_cast<const F*>(A.val);
    const index_t* __restrict__ col = A.col;
    const int lane = threadIdx.x & 63;
    const int rl = lane / L, q = lane % L;
    const int sl = part * RPW + rl;        // the row's lane in the slice layout
    const size_t li = size_t(s) * 64 + sl;
    const index_t row = A.lane_row[li];
    const index_t info = A.lane_info[li];
    index_t c[CH];
    F v[CH];
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        c[j] = kPadColumn;
        v[j] = zero<F>();
        if (j < stride) {
            const size_t at = (size_t(o0) + j) * 64 + sl;
            c[j] = col[at];
            v[j] = val[at];
        }
    }
    if (stopped) {
        if (info < 0) out[0] = cvt<P>(v[0]);   // keeps the loads above the test
        return;
    }
    const int pos = info & 255;
    const index_t o1 = o0 + __builtin_amdgcn_readfirstlane(info >> 8);
    const bool live = row >= 0;
    const index_t rr = live ? row : 0;
    constexpr bool kLateIn = !FWD;
    P xin[RL], acc[RL];
    if (kLateIn) {
#pragma unroll
        for (int r = 0; r < RL; ++r) xin[r] = zero<P>();
    } else {
        load_part<P, W, L>(in + size_t(rr) * W, q, xin);
    }
#pragma unroll
    for (int r = 0; r < RL; ++r) acc[r] = FWD ? xin[r] : zero<P>();
    const P dinv = cvt<P>(static_cast<const F*>(A.invd)[rr]);
    F intv[BR > 1 ? BR - 1 : 1];
#pragma unroll
    for (int j = 0; j < (BR > 1 ? BR - 1 : 1); ++j) intv[j] = zero<F>();
    for (index_t k0 = o0;;) {
        P g[CH][RL];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            if (k0 + j < o1) {
                const index_t cj = c[j] >= 0 ? c[j] : 0;
                load_part<P, W, L>(out + size_t(cj) * W, q, g[j]);
            } else {
                c[j] = kPadColumn;
#pragma unroll
                for (int r = 0; r < RL; ++r) g[j][r] = zero<P>();
            }
        }
        const index_t k1 = k0 + CH;
        index_t cn[CH];
        F vn[CH];
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            cn[j] = kPadColumn;
            vn[j] = zero<F>();
        }
        if (k1 < o1) {
#pragma unroll
            for (int j = 0; j < CH; ++j) {
                if (k1 + j < o1) {
                    const size_t at = (size_t(k1) + j) * 64 + sl;
                    cn[j] = col[at];
                    vn[j] = val[at];
                }
            }
        }
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            const P vj = cvt<P>(v[j]);
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const P t = fms_(vj, g[j][r], acc[r]);
                acc[r] = select_(c[j] >= 0, t, acc[r]);
            }
#pragma unroll
            for (int p = 0; p < BR - 1; ++p) intv[p] = select_(c[j] == -1 - p, v[j], intv[p]);
        }
        k0 = k1;
        if (k0 >= o1) break;
#pragma unroll
        for (int j = 0; j < CH; ++j) {
            c[j] = cn[j];
            v[j] = vn[j];
        }
    }
    if (kLateIn) load_part<P, W, L>(in + size_t(rr) * W, q, xin);
    const int base = rl - pos;
    P res[RL];
#pragma unroll
    for (int r = 0; r < RL; ++r) res[r] = zero<P>();
#pragma unroll
    for (int j = 0; j < BR; ++j) {
        if (live && pos == j) {
#pragma unroll
            for (int r = 0; r < RL; ++r) res[r] = FWD ? mul(acc[r], dinv) : fma_(acc[r], dinv, xin[r]);
            store_part<P, W, L>(out + size_t(row) * W, q, res);
        }
        if (j < BR - 1) {
            // the same right-hand sides (lane q) of position j's row;
            // clamped for idle lanes and short blocks at the wavefront's end
            const int src = min(base + j, RPW - 1) * L + q;
            const P cj = cvt<P>(intv[j]);
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                const P vj = shfl(res[r], src);
                if (pos > j) acc[r] = fms_(cj, vj, acc[r]);
            }
        }
    }
}

// One colour: slices s0..s1, four per workgroup; extra workgroups compute
// the external part of the colour's oversized blocks.  W right-hand sides
// per lane; the colour is skipped once all of them have stopped.
template <int W>
__device__ inline int all_stopped(const SolverState* __restrict__ st)
{
    int stopped = 1;
#pragma unroll
    for (int r = 0; r < W; ++r) stopped &= st[r].status != 0;
    return stopped;
}

// Slice s0 + k has its entries at slot slot0 + k * stride.  WPB slices
// (wavefronts) per workgroup; the external part takes WPB * 64 rows a
// workgroup.
template <typename P, typename F, int BR, bool FWD, int CH, int W, int WPB = kWavesPerBlock>
__global__ __launch_bounds__(WPB * 64) void k_sweep(index_t s0, index_t s1, index_t slot0,
                                                    index_t stride, SweepArrays A,
                                                    const P* __restrict__ in, P* out,
                                                    WalkExt walk,
                                                    const SolverState* __restrict__ st)
{
    const index_t nreg = (s1 - s0 + WPB - 1) / WPB;
    const index_t lb = index_t(W == 1 ? xcd_block(blockIdx.x, gridDim.x) : blockIdx.x);
    if (lb >= nreg) {
        // batches: computed for every system (the results of a stopped
        // system's sweep are never read)
        if (!A.slice_sys && all_stopped<W>(st)) return;
        walk_external<P, F, FWD, W>(walk, walk.t0 + size_t(lb - nreg) * (WPB * 64) + threadIdx.x, in, out);
        return;
    }
    const index_t wave = WPB == 1 ? 0 : __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    const index_t s = s0 + lb * WPB + wave;
    if (s >= s1) return;                       // whole wavefront: no shuffle below
    const int stopped = all_stopped<W>(st + (A.slice_sys ? A.slice_sys[s] : 0));
    if (A.stream_nt)
        sweep_slice<P, F, BR, FWD, CH, W, true>(s, slot0 + (s - s0) * stride, stride, A, in, out, stopped);
    else
        sweep_slice<P, F, BR, FWD, CH, W>(s, slot0 + (s - s0) * stride, stride, A, in, out, stopped);
}

// One colour of a block of W right-hand sides, L wavefronts a slice
// (sweep_slice_split), WPB wavefronts a workgroup; workgroups past the
// slices' compute the external part of the colour's oversized blocks
// (WPB * 64 rows a workgroup).
template <typename P, typename F, int BR, bool FWD, int CH, int W, int L, int WPB>
__global__ __launch_bounds__(WPB * 64) void k_sweep_split(index_t s0, index_t s1, index_t slot0,
                                                          index_t stride, SweepArrays A,
                                                          const P* __restrict__ in, P* out,
                                                          WalkExt walk,
                                                          const SolverState* __restrict__ st)
{
    const index_t nreg = ((s1 - s0) * L + WPB - 1) / WPB;
    if (index_t(blockIdx.x) >= nreg) {
        if (!A.slice_sys && all_stopped<W>(st)) return;
        walk_external<P, F, FWD, W>(walk, walk.t0 + size_t(index_t(blockIdx.x) - nreg) * (WPB * 64) + threadIdx.x,
                                    in, out);
        return;
    }
    const index_t wv = index_t(blockIdx.x) * WPB +
                       (WPB == 1 ? 0 : __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6)));
    const index_t s = s0 + wv / L;
    if (s >= s1) return;                       // whole wavefront: no shuffle below
    const int stopped = all_stopped<W>(st + (A.slice_sys ? A.slice_sys[s] : 0));
    sweep_slice_split<P, F, BR, FWD, CH, W, L>(s, int(wv % L), slot0 + (s - s0) * stride, stride, A,
                                               in, out, stopped);
}

// Several consecutive small colours in one workgroup of NT threads: the
// wavefronts share each colour's slices, a workgroup barrier separates the
// colours (global writes of a workgroup are visible to it after the
// barrier).  NT = 1024 for real values, 512 for complex values or 16-row
// blocks; with several right-hand sides per lane 512 (at most 256 VGPRs,
// which those kernels fit without spills), and 256 for complex<double>
// values or 16-row blocks (register budget).  A colour of a 16-vector
// block is a chain of dependent loads of ~15 us per pass of the
// workgroup's wavefronts: at 256 threads the eddy78 tail (9 slices) took
// three passes.
template <typename P, int BR, int W = 1>
constexpr int tail_threads()
{
    if (W > 1) return (sizeof(P) > 8 || BR >= 16) ? 256 : 512;
    return (sizeof(P) > 8 || BR >= 16) ? 512 : 1024;
}
constexpr int kTailMaxPhases = 16;
struct TailPhases {
    int count;
    index_t slices[kTailMaxPhases + 1];
    index_t slot[kTailMaxPhases];     // first entry slot of the phase's slices
    index_t stride[kTailMaxPhases];   // entry slots from one slice to the next
};

template <typename P, typename F, int BR, bool FWD, int CH, int W>
__global__ __launch_bounds__((tail_threads<P, BR, W>())) void k_sweep_tail(TailPhases T, SweepArrays A,
                                                                     const P* __restrict__ in, P* out,
                                                                     const SolverState* __restrict__ st)
{
    if (!A.slice_sys && all_stopped<W>(st)) return;
    const index_t w = __builtin_amdgcn_readfirstlane(int(threadIdx.x >> 6));
    constexpr index_t kWaves = tail_threads<P, BR, W>() / 64;
    for (int ph = 0; ph < T.count; ++ph) {
        const index_t sp = T.slices[ph], stride = T.stride[ph];
        for (index_t s = sp + w; s < T.slices[ph + 1]; s += kWaves)
            if (A.stream_nt)
                sweep_slice<P, F, BR, FWD, CH, W, true>(s, T.slot[ph] + (s - sp) * stride, stride, A, in, out,
                                                        A.slice_sys ? st[A.slice_sys[s]].status : 0);
            else
                sweep_slice<P, F, BR, FWD, CH, W>(s, T.slot[ph] + (s - sp) * stride, stride, A, in, out,
                                                  A.slice_sys ? st[A.slice_sys[s]].status : 0);
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Triangular sweep, oversized blocks: one workgroup per block, rows grouped
// by dependency level; a workgroup barrier separates levels.  Values of the
// block's own rows are kept in LDS when `use_lds` (indexed by position in
// level order), otherwise re-read from `out`.
// ---------------------------------------------------------------------------
template <typename P, typename F, bool FWD>
__global__ __launch_bounds__(1024) void k_sweep_big(
    index_t b0, const index_t* __restrict__ big_pos, const index_t* __restrict__ big_lvl_off,
    const index_t* __restrict__ big_lvl, const index_t* __restrict__ big_rows,
    const index_t* __restrict__ big_rp, const index_t* __restrict__ big_ci,
    const F* __restrict__ big_v, const F* __restrict__ invd, const P* __restrict__ in,
    P* out, int use_lds, index_t es, const SolverState* __restrict__ st)
{
    // right-hand side r = blockIdx.y of a block: row i at i*es + r
    in += blockIdx.y;
    out += blockIdx.y;
    if (st && st[blockIdx.y].status) return;
    extern __shared__ unsigned char lds_raw[];
    P* lds = reinterpret_cast<P*>(lds_raw);
    const index_t bb = b0 + index_t(blockIdx.x);
    const index_t p0 = big_pos[bb];
    const index_t* lv = big_lvl + big_lvl_off[bb];
    const index_t nlev = big_lvl_off[bb + 1] - big_lvl_off[bb] - 1;
    for (index_t l = 0; l < nlev; ++l) {
        for (index_t q = lv[l] + index_t(threadIdx.x); q < lv[l + 1]; q += index_t(blockDim.x)) {
            const index_t gp = p0 + q;
            const index_t row = big_rows[gp];
            P acc = FWD ? in[size_t(row) * es] : zero<P>();
            for (index_t k = big_rp[gp]; k < big_rp[gp + 1]; ++k) {
                const index_t c = big_ci[k];
                P xv;
                if (c >= 0) xv = out[size_t(c) * es];
                else if (use_lds) xv = lds[-1 - c];
                else xv = out[size_t(big_rows[p0 + (-1 - c)]) * es];
                acc = fms_(cvt<P>(big_v[k]), xv, acc);
            }
            const P dinv = cvt<P>(invd[row]);
            const P res = FWD ? mul(acc, dinv) : fma_(acc, dinv, in[size_t(row) * es]);
            out[size_t(row) * es] = res;
            if (use_lds) lds[q] = res;
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// Triangular sweep, oversized blocks, single-wavefront walk (launched with
// one wavefront per block, after the phase kernel wrote the external parts).
// Steps run in dependency order; the rows of one step are independent and
// read the values of earlier steps from LDS (indexed by level position).
// The operands of step s+D are loaded while step s computes (D register
// sets, rotated by the unrolled j loop), so the chain of steps pays LDS
// latency, not memory latency.  A block of right-hand sides: one wavefront
// per (block, right-hand side r = blockIdx.y); r's walk slots start at r*ws
// and its rows are out[i*es + r].
// ---------------------------------------------------------------------------
// a step's operands: the external part (and, backward, the input) from the
// phase kernel, and the slot's packed record of R 16-byte words (K values,
// D^-1, int32 row, position and K LDS positions; word w of lane l of step s
// at rec[(s*R + w)*64 + l])
template <typename F, int KMAX>
constexpr int walk_words() { return int(((KMAX + 1) * sizeof(F) + (2 + KMAX) * 4 + 15) / 16); }

template <typename P, typename F, int KMAX>
struct WalkBuf {
    P a, b;
    uint4 w[walk_words<F, KMAX>()];
};

template <typename P, typename F, bool FWD, int KMAX, int D>
__global__ __launch_bounds__(64) void k_sweep_walk(
    index_t b0, const index_t* __restrict__ step_off, const uint4* __restrict__ rec,
    index_t zero_pos, P* out, const P* __restrict__ sa, const P* __restrict__ sb, index_t es,
    size_t ws, const SolverState* __restrict__ st, const int* __restrict__ walk_sys)
{
    constexpr int R = walk_words<F, KMAX>();
    const index_t bb = b0 + index_t(blockIdx.x);
    const int r = int(blockIdx.y);
    out += r;
    sa += size_t(r) * ws;
    sb += size_t(r) * ws;
    if (st && st[(walk_sys ? walk_sys[bb] : 0) + r].status) return;
    extern __shared__ unsigned char lds_raw[];
    P* lds = reinterpret_cast<P*>(lds_raw);   // [0, zero_pos): values; zero_pos: 0; +1: idle lanes
    const index_t s0 = step_off[bb], s1 = step_off[bb + 1];   // s1 - s0 is a multiple of D
    const unsigned ul = threadIdx.x;
    const index_t trash = zero_pos + 1;
    if (ul == 0) lds[zero_pos] = zero<P>();
    __syncthreads();
    WalkBuf<P, F, KMAX> buf[D];
    // the step arrays extend D idle steps past the last block, so every
    // prefetch below is in bounds and none needs a guard
    auto load = [&](index_t s, WalkBuf<P, F, KMAX>& B) {
        const size_t t0 = size_t(s) * 64;
        B.a = (sa + t0)[ul];
        B.b = FWD ? zero<P>() : (sb + t0)[ul];
        const uint4* q = rec + size_t(s) * R * 64;
#pragma unroll
        for (unsigned w = 0; w < unsigned(R); ++w) B.w[w] = q[w * 64u + ul];
    };
#pragma unroll
    for (int j = 0; j < D; ++j) load(s0 + j, buf[j]);
    // One full wait here, once: the compiler orders the loads above as it
    // likes, and the wait counts at the loop head merge this entry with the
    // back edge -- without it, every group of D steps would wait for nearly
    // all of its prefetches (the slot the first step reads may be the last
    // one loaded here).
    __builtin_amdgcn_s_waitcnt(0);
    for (index_t s = s0; s < s1; s += D) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            WalkBuf<P, F, KMAX>& B = buf[j];
            // the record's fields (register moves)
            F v[KMAX], d;
            index_t ints[2 + KMAX];
            __builtin_memcpy(v, B.w, sizeof(v));
            __builtin_memcpy(&d, reinterpret_cast<const unsigned char*>(B.w) + KMAX * sizeof(F), sizeof(F));
            __builtin_memcpy(ints, reinterpret_cast<const unsigned char*>(B.w) + (KMAX + 1) * sizeof(F), sizeof(ints));
            const index_t row = ints[0], pos = ints[1];
            // the chain of the walk is LDS write -> next step's LDS reads:
            // issue this step's reads first, then refill the slot the
            // previous step consumed while they are in flight (step
            // s + j - 1 + D; in the first group that slot still holds its
            // own step and is reloaded unchanged; the last refills reach
            // D - 2 steps past the block, inside the padded step arrays)
            P g[KMAX];
#pragma unroll
            for (int k = 0; k < KMAX; ++k) g[k] = lds[ints[2 + k]];
            const P a = B.a, xb = B.b;
            load(s + j - 1 + D, buf[(j + D - 1) % D]);
            P acc = a;
#pragma unroll
            for (int k = 0; k < KMAX; ++k) acc = fms_(cvt<P>(v[k]), g[k], acc);
            const P dinv = cvt<P>(d);
            const P res = FWD ? mul(acc, dinv) : fma_(acc, dinv, xb);
            // Later steps read these LDS values from other lanes.  One
            // wavefront's LDS operations execute in order, and the compiler
            // keeps a possibly aliasing LDS read after this write.
            lds[pos >= 0 ? pos : trash] = res;
            if (pos >= 0) out[size_t(row) * es] = res;
            __builtin_amdgcn_wave_barrier();
        }
    }
}

// ---------------------------------------------------------------------------
// Vector kernels (grid-stride; one partial per workgroup)
// ---------------------------------------------------------------------------

// partial sums of conj?(a_i) b_i, accumulated in T (operands converted;
// each product rounded, then added, as the block kernel adds them)
// rows a thread of the vector kernels loads before it uses any: its rows of
// a grid-stride loop come kRowsAhead at a time (one memory round trip for
// up to 4 rows), and the first group is requested before the kernel's
// control step sums the producer's partials
constexpr int kRowsAhead = 4;

template <typename T, typename TA, typename TB>
__global__ __launch_bounds__(kBlock) void k_dot(index_t n, const TA* __restrict__ a,
                                               const TB* __restrict__ b, int conj,
                                               T* __restrict__ partials,
                                               const SolverState* __restrict__ st, Ctrl check,
                                               Systems sy)
{
    const Span sp = vec_span(sy, n);
    if (st && st[sp.sys].status) return;
    __shared__ T red[kWavesPerBlock];
    T d = zero<T>();
    const index_t step = sp.G * kBlock;
    for (index_t jb = sp.local * kBlock + threadIdx.x; jb < sp.n; jb += kRowsAhead * step) {
        // a thread's next kRowsAhead rows load together, then add in order
        TA av[kRowsAhead];
        TB bv[kRowsAhead];
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const index_t j = jb + u * step;
            av[u] = j < sp.n ? a[sp.row0 + j] : TA{};
            bv[u] = j < sp.n ? b[sp.row0 + j] : TB{};
        }
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u)
            if (jb + u * step < sp.n) d = add(d, mul(conj_if(cvt<T>(av[u]), conj != 0), cvt<T>(bv[u])));
    }
    T t = block_sum(d, red);
    if (threadIdx.x == 0) partials[sp.slot] = t;
    if (check.part) {
        __shared__ double redd[kWavesPerBlock];
        ctrl_check_wg0(ctrl_for(check, sy, sp.sys, sp.local), redd);
    }
}

// partial sums of |a_i|^2 and (optionally) |b_i|^2
template <typename T>
__global__ __launch_bounds__(kBlock) void k_norms(index_t n, const T* __restrict__ a,
                                                 const T* __restrict__ b,
                                                 double* __restrict__ pa,
                                                 double* __restrict__ pb, Systems sy)
{
    const Span sp = vec_span(sy, n);
    __shared__ double red[kWavesPerBlock];
    double sa = 0.0, sb = 0.0;
    for (index_t i = sp.local * kBlock + threadIdx.x; i < sp.n; i += sp.G * kBlock) {
        sa += norm2(a[sp.row0 + i]);
        if (b) sb += norm2(b[sp.row0 + i]);
    }
    double t = block_sum(sa, red);
    if (threadIdx.x == 0) pa[sp.slot] = t;
    if (b) {
        t = block_sum(sb, red);
        if (threadIdx.x == 0) pb[sp.slot] = t;
    }
}

// r = b - Ap
template <typename T>
__global__ __launch_bounds__(kBlock) void k_residual(index_t n, const T* __restrict__ b,
                                                    const T* __restrict__ Ap, T* __restrict__ r)
{
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += index_t(gridDim.x) * kBlock)
        r[i] = sub(b[i], Ap[i]);
}

// dst[i] = src[map[i]]; vector blockIdx.y of several at the given strides
// out = S in, S = diag(s) per row (diagonal scaling; W values per row,
// interleaved).  Each component is multiplied by s in the vector's precision
// (in and out may be the same vector).
__device__ inline double scale_by(double v, double s) { return v * s; }
__device__ inline float scale_by(float v, double s) { return v * float(s); }
__device__ inline c128 scale_by(c128 v, double s) { return c128{v.x * s, v.y * s}; }
__device__ inline c64 scale_by(c64 v, double s) { return c64{v.x * float(s), v.y * float(s)}; }

template <typename P>
__global__ __launch_bounds__(kBlock) void k_scale_rows(size_t count, int W, const double* __restrict__ s,
                                                      const P* in, P* out)
{
    for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < count; e += size_t(gridDim.x) * kBlock)
        out[e] = scale_by(in[e], s[e / size_t(W)]);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_gather(index_t n, const index_t* __restrict__ map,
                                                  const T* __restrict__ src, index_t src_stride,
                                                  T* __restrict__ dst, index_t dst_stride)
{
    src += size_t(blockIdx.y) * size_t(src_stride);
    dst += size_t(blockIdx.y) * size_t(dst_stride);
    for (index_t i = index_t(blockIdx.x) * kBlock + threadIdx.x; i < n;
         i += index_t(gridDim.x) * kBlock) {
        const index_t m = map[i];
        dst[i] = m >= 0 ? src[m] : zero<T>();    // -1: a padding row of a batch
    }
}

// p[0 .. words) = 0: clears device state with a kernel (no memset)
static __global__ __launch_bounds__(kBlock) void k_zero_words(unsigned* __restrict__ p, size_t words)
{
    for (size_t e = size_t(blockIdx.x) * kBlock + threadIdx.x; e < words; e += size_t(gridDim.x) * kBlock)
        p[e] = 0u;
}

// dst = src, converted to the destination's precision
template <typename TS, typename TD>
__global__ __launch_bounds__(kBlock) void k_copy(index_t n, const TS* __restrict__ src,
                                                TD* __restrict__ dst,
                                                const SolverState* __restrict__ st, Systems sy)
{
    const Span sp = vec_span(sy, n);
    if (st && st[sp.sys].status) return;
    for (index_t i = sp.local * kBlock + threadIdx.x; i < sp.n; i += sp.G * kBlock)
        dst[sp.row0 + i] = cvt<TD>(src[sp.row0 + i]);
}

// CG: x += alpha p; r -= alpha Ap; partial |r|^2; r also in the
// preconditioner's precision when r_low is given
template <typename T, typename P>
__global__ __launch_bounds__(kBlock) void k_cg_update(index_t n, T* __restrict__ x,
                                                     T* __restrict__ r, const T* __restrict__ p,
                                                     const T* __restrict__ Ap,
                                                     double* __restrict__ partials,
                                                     const SolverState* __restrict__ st,
                                                     Ctrl ctl, P* __restrict__ r_low, Systems sy)
{
    const Span sp = vec_span(sy, n);
    if (st[sp.sys].status) return;
    // the operands of this thread's first rows load before alpha is summed
    const index_t step = sp.G * kBlock, j0 = sp.local * kBlock + threadIdx.x;
    T xv[kRowsAhead], pv[kRowsAhead], av[kRowsAhead], rv[kRowsAhead];
    auto load = [&](index_t jb) {
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const index_t j = jb + u * step;
            if (j >= sp.n) break;
            const index_t i = sp.row0 + j;
            xv[u] = x[i];
            pv[u] = p[i];
            av[u] = Ap[i];
            rv[u] = r[i];
        }
    };
    load(j0);
    __shared__ double red[kWavesPerBlock];
    __shared__ T redt[kWavesPerBlock];
    T alpha;
    if (!ctrl_alpha_all(ctrl_for(ctl, sy, sp.sys, sp.local), redt, alpha)) return;
    double s = 0.0;
    for (index_t jb = j0; jb < sp.n;) {
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const index_t j = jb + u * step;
            if (j >= sp.n) break;
            const index_t i = sp.row0 + j;
            x[i] = fma_(alpha, pv[u], xv[u]);
            const T ri = fms_(alpha, av[u], rv[u]);
            r[i] = ri;
            if (r_low) r_low[i] = cvt<P>(ri);
            s += norm2(ri);
        }
        jb += kRowsAhead * step;
        if (jb < sp.n) load(jb);
    }
    double t = block_sum(s, red);
    if (threadIdx.x == 0) partials[sp.slot] = t;
}

// best_x <- x when the last convergence check improved the best residual
// (best null: not tracked).  This runs also after a stop: the check that
// stopped the solve may have improved the best residual.
template <typename T>
__device__ inline void copy_if_improved(const Span& sp, const T* __restrict__ x,
                                        T* __restrict__ best, const SolverState* __restrict__ st)
{
    if (!best || !st[sp.sys].improved) return;
    for (index_t j = sp.local * kBlock + threadIdx.x; j < sp.n; j += sp.G * kBlock)
        best[sp.row0 + j] = x[sp.row0 + j];
}

// CG: p = z + beta p (and the best-x copy of this iteration); z may be in
// the preconditioner's precision
template <typename T, typename TZ>
__global__ __launch_bounds__(kBlock) void k_xpby(index_t n, const TZ* __restrict__ z,
                                                T* __restrict__ p, const T* __restrict__ x,
                                                T* __restrict__ best,
                                                const SolverState* __restrict__ st, Ctrl ctl,
                                                Systems sy)
{
    const Span sp = vec_span(sy, n);
    copy_if_improved(sp, x, best, st);
    if (st[sp.sys].status) return;
    // the operands of this thread's first rows load before beta is summed
    const index_t step = sp.G * kBlock, j0 = sp.local * kBlock + threadIdx.x;
    T pv[kRowsAhead];
    TZ zv[kRowsAhead];
    auto load = [&](index_t jb) {
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const index_t j = jb + u * step;
            if (j >= sp.n) break;
            const index_t i = sp.row0 + j;
            pv[u] = p[i];
            zv[u] = z[i];
        }
    };
    load(j0);
    __shared__ T redt[kWavesPerBlock];
    T beta;
    if (!ctrl_beta_all(ctrl_for(ctl, sy, sp.sys, sp.local), redt, beta)) return;
    for (index_t jb = j0; jb < sp.n;) {
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const index_t j = jb + u * step;
            if (j >= sp.n) break;
            p[sp.row0 + j] = fma_(beta, pv[u], cvt<T>(zv[u]));
        }
        jb += kRowsAhead * step;
        if (jb < sp.n) load(jb);
    }
}

// COCR: x += alpha p; r -= alpha q; rt -= alpha qt; partial |r|^2 (qt and
// rt may be in the preconditioner's precision: rt is then computed wide and
// stored narrow).  best <- x (before its update) when the last check
// improved the best residual: x is read here anyway, so the copy costs one
// store.  A solve that stops at an improving check ends with x itself as its
// best iterate (the caller takes x, not best, when `improved` is set).
template <typename T, typename TQ, typename TR = T>
__global__ __launch_bounds__(kBlock) void k_cocr_update(index_t n, T* __restrict__ x,
                                                       T* __restrict__ r, TR* __restrict__ rt,
                                                       const T* __restrict__ p,
                                                       const T* __restrict__ q,
                                                       const TQ* __restrict__ qt,
                                                       double* __restrict__ partials,
                                                       const SolverState* __restrict__ st,
                                                       Ctrl ctl, Systems sy, T* __restrict__ best)
{
    const Span sp = vec_span(sy, n);
    if (st[sp.sys].status) return;
    // set by the last check; this kernel's control step does not change it
    const bool copy = best && st[sp.sys].improved;
    // the operands of this thread's first rows load before alpha is summed
    const index_t step = sp.G * kBlock, j0 = sp.local * kBlock + threadIdx.x;
    T xv[kRowsAhead], pv[kRowsAhead], rv[kRowsAhead], qv[kRowsAhead];
    TR rtv[kRowsAhead];
    TQ qtv[kRowsAhead];
    auto load = [&](index_t jb) {
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const index_t j = jb + u * step;
            if (j >= sp.n) break;
            const index_t i = sp.row0 + j;
            xv[u] = x[i];
            pv[u] = p[i];
            rv[u] = r[i];
            qv[u] = q[i];
            rtv[u] = rt[i];
            qtv[u] = qt[i];
        }
    };
    load(j0);
    __shared__ double red[kWavesPerBlock];
    __shared__ T redt[kWavesPerBlock];
    T alpha;
    if (!ctrl_alpha_all(ctrl_for(ctl, sy, sp.sys, sp.local), redt, alpha)) return;
    double s = 0.0;
    for (index_t jb = j0; jb < sp.n;) {
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const index_t j = jb + u * step;
            if (j >= sp.n) break;
            const index_t i = sp.row0 + j;
            if (copy) best[i] = xv[u];
            x[i] = fma_(alpha, pv[u], xv[u]);
            const T ri = fms_(alpha, qv[u], rv[u]);
            r[i] = ri;
            rt[i] = cvt<TR>(fms_(alpha, cvt<T>(qtv[u]), cvt<T>(rtv[u])));
            s += norm2(ri);
        }
        jb += kRowsAhead * step;
        if (jb < sp.n) load(jb);
    }
    double t = block_sum(s, red);
    if (threadIdx.x == 0) partials[sp.slot] = t;
}

// COCR: p = rt + beta p; q = t + beta q; q also in the preconditioner's
// precision when q_low is given; rt may be stored in it (TR).  (The best-x
// copy is the next k_cocr_update's.)
template <typename T, typename P, typename TR = T>
__global__ __launch_bounds__(kBlock) void k_cocr_pq(index_t n, const TR* __restrict__ rt,
                                                   const T* __restrict__ t, T* __restrict__ p,
                                                   T* __restrict__ q,
                                                   const SolverState* __restrict__ st, Ctrl ctl,
                                                   P* __restrict__ q_low, Systems sy)
{
    const Span sp = vec_span(sy, n);
    if (st[sp.sys].status) return;
    // the operands of this thread's first rows load before beta is summed
    const index_t step = sp.G * kBlock, j0 = sp.local * kBlock + threadIdx.x;
    T pv[kRowsAhead], qv[kRowsAhead], tv[kRowsAhead];
    TR rtv[kRowsAhead];
    auto load = [&](index_t jb) {
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const index_t j = jb + u * step;
            if (j >= sp.n) break;
            const index_t i = sp.row0 + j;
            pv[u] = p[i];
            rtv[u] = rt[i];
            qv[u] = q[i];
            tv[u] = t[i];
        }
    };
    load(j0);
    __shared__ T redt[kWavesPerBlock];
    T beta;
    if (!ctrl_beta_all(ctrl_for(ctl, sy, sp.sys, sp.local), redt, beta)) return;
    for (index_t jb = j0; jb < sp.n;) {
#pragma unroll
        for (int u = 0; u < kRowsAhead; ++u) {
            const index_t j = jb + u * step;
            if (j >= sp.n) break;
            const index_t i = sp.row0 + j;
            p[i] = fma_(beta, pv[u], cvt<T>(rtv[u]));
            const T qi = fma_(beta, qv[u], tv[u]);
            q[i] = qi;
            if (q_low) q_low[i] = cvt<P>(qi);
        }
        jb += kRowsAhead * step;
        if (jb < sp.n) load(jb);
    }
}

}  // namespace ssa
