// mpm_kernels.hpp — gfx950 kernels of one MPM step (plain FP64 form).
//
// Particles are binned every step into blocks of E^dim dual cells (E = 4 in 3D, 8 in 2D). The dual cell of
// a particle is the first node of its non-zero B-spline window per axis, so every particle of a block
// touches only the block's tile of (E + W - 1)^dim nodes (W = 3 non-zero nodes per axis for the quadratic
// spline, 4 for the cubic). All particle and grid kernels work one block per workgroup.
//
// Step pipeline (one HIP stream, no host synchronisation inside a batch of steps):
//   k_setup        1 wavefront: dt (updateDt), grid geometry (remeshFixedInit/Cont + arange), capacity check
//   k_zero_grid    clear the block histogram (batched scenes; k_setup does it for a single scene)
//   k_bin_count    block of every particle + histogram (wavefront-aggregated atomics)
//   k_bin_scan     exclusive prefix sum of the histogram (one workgroup)
//   k_bin_scatter  counting-sort permutation: sorted slot -> particle of the previous layout
//   k_p2g          per block: order the particles by (dual cell, previous slot); one thread per
//                  (cell, first-axis window offset) sums its cell's contributions in registers; the
//                  partial sums are added node by node in a fixed order into the block's tile, which is
//                  written to a staging area (P2G + the stress force of explicitEulerUpdate)
//   k_grid_update  per node: sum the tiles of the (at most 2^dim) blocks that cover it in block order,
//                  normalise, explicit update, boundary conditions, flip
//   k_g2p          gather from an LDS copy of the block's tile: pic, flip, B, velocity and position
//                  update (G2P + the velocity blend + positionUpdate)
//   MUSL only: k_musl_scatter  re-scatter of the new momentum, like k_p2g (MUSL)
//              k_musl_update   per node: sum the tiles, divide by m_i/m_p (MUSL tail)
//   k_defgrad      gather grad v from LDS copies of the tiles of the workgroup's blocks (USL: the updated
//                  grid velocity; MUSL: the re-scattered one), F update, return map, Kirchhoff stress for
//                  the next step (deformationUpdate + plasticity)
//   k_finish       1 wavefront: moveObjects, time/frame bookkeeping of simulate()
//                  (k_finish_setup: the same and the next step's k_setup, between the steps of a batch)
//
// No floating-point atomics are used: every sum has a fixed order given the particle state, so a step is
// deterministic and two runs with equal inputs give equal bits. (A block with more than kMaxBP particles
// takes a slower path that accumulates its tile with LDS atomics; its sums are then not ordered.)
//
// Particle state is double-buffered. Every step reads the previous layout through the sort permutation
// and writes all fields in sorted order into the other buffer; each particle carries its original index
// (id) for downloads. Every kernel returns at once when StepState::stop is set, so the host can enqueue
// many steps and let the device stop at an error, a frame end, the final time or a capacity overflow.
#pragma once
#include <hip/hip_runtime.h>
#include <utility>
#include "mpm_math.hpp"

namespace mpm {

// f(std::integral_constant<int, i>()) for i = 0 .. N-1, for loops whose index must be a compile-time constant.
template <class F, int... I>
__device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, I...>) { (f(std::integral_constant<int, I>()), ...); }
template <int N, class F>
__device__ __forceinline__ void static_for(F&& f) { static_for_impl(f, std::make_integer_sequence<int, N>()); }

// One layout of the particle state (structure of arrays). TF is the type of the transfer quantities
// (velocity, affine matrix, Kirchhoff stress): double in the plain form, float in the split form. Position,
// deformation gradient and the plastic state are double in both.
template <class TF> struct PartStateT {
    double* x[3];
    TF* v[3];
    TF* B[9];
    double* F[9];
    TF* tau[6];
    double* eps_dev; double* eps_vol; double* eps_prad; double* dgamma; double* visc; double* muI;
    uint32_t* id;    // index of the particle in the caller's order
};

// Counting-sort work arrays.
struct SortBufs {
    uint32_t* key;       // block of every particle of the source layout
    uint32_t* perm;      // sorted slot -> particle of the source layout
    uint32_t* count;     // particles per block
    uint32_t* offset;    // exclusive prefix sum of count
    uint32_t* cursor;    // scatter cursors
    uint16_t* cell_off;  // per block: first sorted slot (relative to offset) of each dual cell, 65 entries
    uint8_t* cellk;      // dual cell within its block of every entry of key (BG::NC: the window is outside)
    uint8_t* cells;      // the same per sorted slot (k_bin_scatter), so that ordering a block needs no positions
    // Periodic boundaries (ghosts = 1): particle p near x = 0 or x = Lx also has a periodic copy, binned
    // under key[n + p] = nblocks + its block (kNoGhost: none). count/offset/cursor then have 2 nblocks
    // entries, and the copies' slots follow the particles' (perm entries kGhostBit | p).
    uint32_t ghosts;
};
constexpr uint32_t kGhostBit = 0x80000000u;
constexpr uint32_t kNoGhost = 0xFFFFFFFFu;

template <class TF> struct GridPtrsT {
    TF* mass;        // m_i
    TF* v[3];        // node velocity after the update (MUSL: after the re-scatter)
    TF* flip[3];     // v_new - v_old (formed when flip_ratio >= -1 or conservation is checked)
    TF* fric;        // MIBF friction coefficient
    double* coord[3];
};

// Per-block node tiles written by the scatter kernels: value v of tile node q of block b is at
// base[v * stride + b * NT + q].
template <class TF> struct StagingT {
    TF* base;
    uint32_t stride;
};

constexpr int kBlock = 256;
constexpr int kMaxBP = 2048;   // particles of a block ordered in LDS; larger blocks take the LDS-atomic path
constexpr int kMaxNV = 8;      // values per tile node (mass, 3 momentum, 3 force, friction)

template <int D> struct BlockEdge { static constexpr int v = D == 3 ? 4 : 8; };

// Block geometry.
template <int D, int SPL> struct BG {
    static constexpr int E = BlockEdge<D>::v;              // dual cells per axis
    static constexpr int W = SPL == 2 ? 3 : 4;              // non-zero window nodes per axis
    static constexpr int T = E + W - 1;                     // tile nodes per axis
    static constexpr int NC = D == 3 ? E * E * E : E * E;   // dual cells per block (64)
    static constexpr int NT = D == 3 ? T * T * T : T * T;   // tile nodes
    static constexpr int NS = D == 3 ? W * W : W;           // window nodes with a fixed first-axis offset
};


// order-preserving int64 key of a double (for exact atomic min/max)
__device__ __forceinline__ long long ord_key(double d) { long long b = __double_as_longlong(d); return b >= 0 ? b : (b ^ 0x7FFFFFFFFFFFFFFFLL); }
__device__ __forceinline__ double key_val(long long k) { long long b = k >= 0 ? k : (k ^ 0x7FFFFFFFFFFFFFFFLL); return __longlong_as_double(b); }

__device__ __forceinline__ void atomic_add_f64(double* a, double v) { unsafeAtomicAdd(a, v); }

// wavefront reductions (64 lanes)
__device__ __forceinline__ unsigned long long wave_max_u64(unsigned long long v) {
    for (int o = 32; o > 0; o >>= 1) { unsigned long long t = __shfl_xor(v, o, 64); v = t > v ? t : v; }
    return v;
}
__device__ __forceinline__ long long wave_min_i64(long long v) {
    for (int o = 32; o > 0; o >>= 1) { long long t = __shfl_xor(v, o, 64); v = t < v ? t : v; }
    return v;
}
__device__ __forceinline__ long long wave_max_i64(long long v) {
    for (int o = 32; o > 0; o >>= 1) { long long t = __shfl_xor(v, o, 64); v = t > v ? t : v; }
    return v;
}
__device__ __forceinline__ int wave_min_i32(int v) {
    for (int o = 32; o > 0; o >>= 1) { int t = __shfl_xor(v, o, 64); v = t < v ? t : v; }
    return v;
}
__device__ __forceinline__ int wave_max_i32(int v) {
    for (int o = 32; o > 0; o >>= 1) { int t = __shfl_xor(v, o, 64); v = t > v ? t : v; }
    return v;
}
__device__ __forceinline__ double wave_sum_f64(double v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_sum_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ uint32_t wave_max_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) { const uint32_t t = __shfl_xor(v, o, 64); v = t > v ? t : v; }
    return v;
}
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}

// Wavefront-aggregated atomicAdd(ctr + key, 1): each run of consecutive active lanes with equal keys shares one
// atomic, and the runs' atomics go out as one instruction (one memory round trip per wavefront, whatever the number
// of runs: the particles arrive in the previous step's block order, so a wavefront holds one or two runs, and
// unsorted input costs no more round trips). Each lane gets its own slot: its run's old value plus its rank in the
// run. Lanes with active == false take no slot. Two runs of one key in a wavefront get disjoint slot ranges in an
// order the hardware picks; the slot order inside a block is fixed later (order_block).
__device__ __forceinline__ uint32_t wave_claim(uint32_t* ctr, uint32_t key, bool active) {
    const int lane = threadIdx.x & 63;
    const uint32_t kprev = __shfl_up(key, 1, 64);
    const unsigned long long act = __ballot(active);
    const bool prev_act = lane > 0 && ((act >> (lane - 1)) & 1ull);
    const bool start = active && !(prev_act && kprev == key);
    const unsigned long long starts = __ballot(start);
    const unsigned long long upto = (2ull << lane) - 1ull;  // lanes 0..lane (all 64 bits for lane 63)
    const unsigned long long below = starts & upto;
    const int leader = below ? 63 - __clzll((long long)below) : lane;
    uint32_t base = 0;
    if (start) {
        const unsigned long long after = (starts | ~act) & ~upto;  // the lanes above where the run cannot go on
        const int end = after ? __ffsll((long long)after) - 1 : 64;
        base = atomicAdd(ctr + key, (uint32_t)(end - lane));
    }
    base = __shfl(base, leader, 64);
    return active ? base + (uint32_t)(lane - leader) : 0u;
}

// Per-thread accumulation of the step reductions (max speed for updateDt, bounds for remesh) over the
// particles a thread handles.
struct BoundsAcc {
    unsigned long long v2 = 0ull;
    long long mn[3] = {0x7FFFFFFFFFFFFFFFLL, 0x7FFFFFFFFFFFFFFFLL, 0x7FFFFFFFFFFFFFFFLL};
    long long mx[3] = {(-0x7FFFFFFFFFFFFFFFLL - 1), (-0x7FFFFFFFFFFFFFFFLL - 1), (-0x7FFFFFFFFFFFFFFFLL - 1)};
    template <int D> __device__ void add(const double* x, const double* v) {
        double s = v[0] * v[0];
        for (int d = 1; d < D; ++d) s += v[d] * v[d];
        const unsigned long long b = (unsigned long long)__double_as_longlong(s);
        v2 = b > v2 ? b : v2;
        for (int d = 0; d < D; ++d) { const long long k = ord_key(x[d]); mn[d] = k < mn[d] ? k : mn[d]; mx[d] = k > mx[d] ? k : mx[d]; }
    }
};

// Workgroup reduction of BoundsAcc (all threads call), then one set of atomics per workgroup into its slot.
template <int D>
__device__ __forceinline__ void reduce_bounds(StepState* S, const BoundsAcc& ba) {
    unsigned long long v2 = wave_max_u64(ba.v2);
    long long mn[3], mx[3];
    for (int d = 0; d < 3; ++d) { mn[d] = ba.mn[d]; mx[d] = ba.mx[d]; }
    for (int d = 0; d < D; ++d) { mn[d] = wave_min_i64(mn[d]); mx[d] = wave_max_i64(mx[d]); }
    __shared__ long long s_red[1024 / 64][7];
    const int wid = threadIdx.x >> 6, nw = (blockDim.x + 63) >> 6;
    if ((threadIdx.x & 63) == 0) {
        s_red[wid][0] = (long long)v2;
        for (int d = 0; d < 3; ++d) { s_red[wid][1 + d] = mn[d]; s_red[wid][4 + d] = mx[d]; }
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < nw; ++w) {
            v2 = max(v2, (unsigned long long)s_red[w][0]);
            for (int d = 0; d < D; ++d) { mn[d] = min(mn[d], s_red[w][1 + d]); mx[d] = max(mx[d], s_red[w][4 + d]); }
        }
        RedSlot& r = S->red[blockIdx.x % kRedSlots];
        atomicMax(&r.v2, v2);
        for (int d = 0; d < D; ++d) { atomicMin(&r.mn[d], mn[d]); atomicMax(&r.mx[d], mx[d]); }
    }
}

// Folds the bound slots into H (the caller's LDS copy of the head of S) and resets them (one wavefront of 64
// lanes calls). Min/max folds are idempotent, so a step that is retried after a grid reallocation sees the same
// bounds.
template <int D>
__device__ __forceinline__ void fold_bound_slots(StepState* S, StepState& H) {
    unsigned long long v2 = 0ull;
    long long mn[3], mx[3];
    for (int d = 0; d < 3; ++d) { mn[d] = 0x7FFFFFFFFFFFFFFFLL; mx[d] = (-0x7FFFFFFFFFFFFFFFLL - 1); }
    for (int s = threadIdx.x; s < kRedSlots; s += 64) {
        RedSlot& r = S->red[s];
        v2 = max(v2, r.v2);
        for (int d = 0; d < D; ++d) { mn[d] = min(mn[d], r.mn[d]); mx[d] = max(mx[d], r.mx[d]); }
        r.v2 = 0ull;
        for (int d = 0; d < 3; ++d) { r.mn[d] = 0x7FFFFFFFFFFFFFFFLL; r.mx[d] = (-0x7FFFFFFFFFFFFFFFLL - 1); }
    }
    v2 = wave_max_u64(v2);
    for (int d = 0; d < D; ++d) { mn[d] = wave_min_i64(mn[d]); mx[d] = wave_max_i64(mx[d]); }
    if (threadIdx.x == 0) {  // into H, the workgroup's copy of the head of S
        H.vmax2_bits = max(H.vmax2_bits, v2);
        for (int d = 0; d < D; ++d) { H.xmin_key[d] = min(H.xmin_key[d], mn[d]); H.xmax_key[d] = max(H.xmax_key[d], mx[d]); }
    }
}

// Adds a wavefront's plastic count to its workgroup's slot.
__device__ __forceinline__ void add_plastic_count(StepState* S, uint32_t pcount) {
    if (pcount) atomicAdd(&S->red[blockIdx.x % kRedSlots].plastic, pcount);
}

// Stencil of one particle: first node of its non-zero window per axis and the window's weights, weight
// derivatives and node coordinates (3 nodes per axis for the quadratic spline, 4 for the cubic).
template <int D, int SPL>
struct Stencil {
    static constexpr int cnt = SPL == 2 ? 3 : 4;
    int base[3];
    double w[3][cnt], dw[3][cnt], xi[3][cnt];
};

// Base node of the 4-node window: floor((x - origin)/dx) - 1, clamped at 0 when clamp is set
// (p2g.cpp:21-24 clamps, g2p.cpp:29-32 does not). Returns false if the window leaves the grid.
template <int D>
__device__ __forceinline__ bool stencil_base(int* b, const double* xp, const double* org, const uint32_t* n, double h1, bool clamp) {
    bool ok = true;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        // a NaN or out-of-range coordinate becomes an index outside every grid (no int overflow in v + 3)
        const double f = floor((xp[d] - org[d]) * h1);
        int v = f < 2.0e9 ? (f > -2.0e9 ? (int)f - 1 : -0x3FFFFFFF) : 0x3FFFFFFF;
        if (clamp && v < 0) v = 0;
        if (v < 0 || v + 3 >= (int)n[d]) ok = false;
        b[d] = v;
    }
    return ok;
}

// Node coordinates read from the grid's coordinate arrays.
struct GlobalCoords {
    double* const* c;
    const uint32_t* n;
    __device__ __forceinline__ double operator()(int d, int i) const { return c[d][i]; }
    __device__ __forceinline__ int clamp4(int d, int b) const { return max(min(b, (int)n[d] - 4), 0); }
};

// Node coordinates around one block staged in LDS: entries lo_d .. lo_d + LC - 1 of axis d (separate
// members rather than arrays, so that the view stays in registers).
template <int LC> struct LdsCoords {
    const double* c0; const double* c1; const double* c2;
    int lo0, lo1, lo2;
    __device__ __forceinline__ double operator()(int d, int i) const {
        return d == 0 ? c0[i - lo0] : (d == 1 ? c1[i - lo1] : c2[i - lo2]);
    }
    __device__ __forceinline__ int clamp4(int d, int b) const {
        const int lo = d == 0 ? lo0 : (d == 1 ? lo1 : lo2);
        return max(min(b, lo + LC - 4), lo);
    }
};

// Node K of a window: weight and derivative (BSpline::node) at the coordinate xi.
template <int SPL, int K, class TF>
__device__ __forceinline__ void window_node(double x, double xi, double h1, TF& w, TF& dw) {
    double wd, dwd;
    BSpline<SPL>::template node<K>((x - xi) * h1, wd, dwd);
    w = (TF)wd; dw = (TF)dwd;
}

// Evaluates the non-zero part of the reference's 4-node window (weights are those of wip/grad_wip,
// tools.hpp:222-252): for the quadratic spline exactly one end node has zero weight, the base node when
// |u| >= 1.5 there, else the last one, and skipping an exactly-zero term leaves every sum of the reference
// unchanged. The four coordinates per axis are fetched unconditionally (from a clamped index) so that
// their loads issue together.
template <int D, int SPL, class CS>
__device__ __forceinline__ bool make_stencil(Stencil<D, SPL>& st, const double* xp, const double* org, const uint32_t* n,
                                             const CS& cs, double h1, bool clamp) {
    int b[3] = {0, 0, 0};
    const bool ok = stencil_base<D>(b, xp, org, n, h1, clamp);
    double c4[3][4];
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int bb = cs.clamp4(d, b[d]);
#pragma unroll
        for (int a = 0; a < 4; ++a) c4[d][a] = cs(d, bb + a);
    }
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const bool shift = SPL == 2 && !(fabs((xp[d] - c4[d][0]) * h1) < 1.5);
        st.base[d] = b[d] + (shift ? 1 : 0);
#pragma unroll
        for (int a = 0; a < Stencil<D, SPL>::cnt; ++a) st.xi[d][a] = (SPL == 2 && MPM_SEL(shift)) ? c4[d][a + 1] : c4[d][a];
        window_node<SPL, 0>(xp[d], st.xi[d][0], h1, st.w[d][0], st.dw[d][0]);
        window_node<SPL, 1>(xp[d], st.xi[d][1], h1, st.w[d][1], st.dw[d][1]);
        window_node<SPL, 2>(xp[d], st.xi[d][2], h1, st.w[d][2], st.dw[d][2]);
        if (SPL == 3) window_node<SPL, SPL == 3 ? 3 : 2>(xp[d], st.xi[d][SPL == 3 ? 3 : 2], h1, st.w[d][SPL == 3 ? 3 : 2], st.dw[d][SPL == 3 ? 3 : 2]);
    }
    return ok;
}

// First node of the non-zero window per axis, as make_stencil chooses it; false if the 4-node window
// leaves the grid.
template <int D, int SPL, class CS>
__device__ __forceinline__ bool window_base(int* wb, const double* xp, const double* org, const uint32_t* n, const CS& cs,
                                            double h1, bool clamp) {
    int b[3] = {0, 0, 0};
    const bool ok = stencil_base<D>(b, xp, org, n, h1, clamp);
    for (int d = 0; d < D; ++d) {
        int s = 0;
        if (SPL == 2) s = fabs((xp[d] - cs(d, cs.clamp4(d, b[d]))) * h1) < 1.5 ? 0 : 1;
        wb[d] = b[d] + s;
    }
    return ok;
}

__device__ __forceinline__ uint32_t node_index3(int i, int j, int k, const uint32_t* n) { return ((uint32_t)i * n[1] + (uint32_t)j) * n[2] + (uint32_t)k; }

template <int D>
__device__ __forceinline__ uint32_t node_index(int i, int j, int k, const uint32_t* n) {
    return D == 3 ? node_index3(i, j, k, n) : (uint32_t)i * n[1] + (uint32_t)j;
}
// The same index computed in 64 bits, for address arithmetic in groups of loads: written in 32 bits, it becomes
// 64-bit multiply-adds whose upper addend is an undefined register, and when that register is the destination of
// a load still in flight the compiler waits for every outstanding load before the multiply-add.
template <int D>
__device__ __forceinline__ uint64_t node_offset(int i, int j, int k, const uint32_t* n) {
    const uint64_t r = (uint64_t)(uint32_t)i * n[1] + (uint32_t)j;
    return D == 3 ? r * n[2] + (uint32_t)k : r;
}


// ----------------------------------------------------------------------------------------------
// k_prepare: Kirchhoff stress of the initial F and the reductions that the first step needs.
// ----------------------------------------------------------------------------------------------
template <int D, class TF>
__global__ __launch_bounds__(kBlock) void k_prepare(Params P, StepState* S, PartStateT<TF> ps, uint32_t n) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = p < n;
    double x[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    if (active) {
        M<D> F;
        for (int i = 0; i < D * D; ++i) F.m[i] = ps.F[i][p];
        double tau[6];
        if (P.elastic_model == MPM_ELASTIC_NEOHOOKEAN) neohookean_tau<D>(F, P.mu, P.lambda, tau);
        else {
            M<D> U, W; V<D> s; svd<D>(F, U, s, W);
            V<D> ls; for (int i = 0; i < D; ++i) ls[i] = log(fabs(s[i]));
            hencky_tau<D>(U, ls, P.mu, P.lambda, tau);
        }
        for (int i = 0; i < SymPack<D>::n; ++i) ps.tau[i][p] = (TF)tau[i];
        for (int d = 0; d < D; ++d) { x[d] = ps.x[d][p]; v[d] = ps.v[d][p]; }
        ps.id[p] = p;
    }
    BoundsAcc ba;
    if (active) ba.add<D>(x, v);
    reduce_bounds<D>(S, ba);
}

// ----------------------------------------------------------------------------------------------
// k_setup: updateDt (update_dt.cpp:5-66) and remeshFixedInit/Cont (remesh.cpp:38-311).
// One workgroup of 64 lanes. The head of StepState (every field before the plates) is copied to LDS by all
// lanes and written back at the end, so that the serial decisions read LDS instead of waiting on a chain of
// global loads. Lane d < dim remeshes axis d (the snap and the node count of arange, tools.hpp:291-296, whose
// accumulating sum is sequential per axis) and later builds that axis's node coordinates; lane 0 then takes
// dt, gravity and the capacity and finiteness decisions.
// ----------------------------------------------------------------------------------------------
constexpr int kHeadWords = (int)(offsetof(StepState, plates) / 4);
// Node coordinates of each axis kept in LDS by the counting pass of arange (the first kCoordStage of them),
// so that only a longer axis's tail is walked a second time when the coordinates are stored.
constexpr uint32_t kCoordStage = 2048;
template <int D, class TF, bool ZERO = false>
__device__ __forceinline__ void setup_body(const Params& P, StepState* S, const GridPtrsT<TF>& g, const SortBufs* sb = nullptr) {
    __shared__ __attribute__((aligned(16))) uint32_t head_w[kHeadWords];
    __shared__ double lo_s[3], hi_s[3], stop_s[3], mn_s[3], mx_s[3], mni_s[3], mxi_s[3], loi_s[3], hii_s[3];
    __shared__ uint32_t cnt_s[3], nb_s[3], nbm_s[3], nbl_s[3];
    __shared__ int go;
    __shared__ double s_coord[D][kCoordStage];
    StepState& H = *reinterpret_cast<StepState*>(head_w);  // only head fields are used through H
    const uint32_t* Sw = reinterpret_cast<const uint32_t*>(S);
    for (int w = threadIdx.x; w < kHeadWords; w += 64) head_w[w] = Sw[w];
    __syncthreads();
    fold_bound_slots<D>(S, H);
    __syncthreads();
    const double dx = P.dx;
    const int t = threadIdx.x;
    if (t < D) {  // remesh of axis t
        const int d = t;
        double mn = key_val(H.xmin_key[d]), mx = key_val(H.xmax_key[d]);
        mn_s[d] = mn; mx_s[d] = mx;  // the particle bounds as reduced (the finiteness check reads these)
        double mni, mxi, loi, hii, lo, hi;
        if (H.step == 0) {
            if (P.grid_reference_point[0] < 1e10) {
                if (P.grid_reference_point[d] < mn) mn = P.grid_reference_point[d];
                else if (P.grid_reference_point[d] > mx) mx = P.grid_reference_point[d];
            }
            mxi = mx; mni = mn;
            loi = mn - dx * 2.0; hii = mx + dx * 2.0;
            lo = loi; hi = hii;
        } else {
            mni = H.min_init[d]; mxi = H.max_init[d]; loi = H.low_init[d]; hii = H.high_init[d];
            if (mx < mxi) { unsigned rf = (unsigned)floor(fmax(0.0, (mxi - mx) / dx - 1e-8 * dx)); hi = hii - rf * dx; }
            else { unsigned ef = (unsigned)floor(fmax(0.0, (mx - mxi) / dx - 1e-8 * dx)); hi = hii + ef * dx; }
            if (mn > mni) { unsigned rf = (unsigned)floor(fmax(0.0, (mn - mni) / dx - 1e-8 * dx)); lo = loi + rf * dx; }
            else { unsigned ef = (unsigned)floor(fmax(0.0, (mni - mn) / dx - 1e-8 * dx)); lo = loi - ef * dx; }
        }
        // node count of arange(lo, hi + dx, dx); periodic runs keep the fixed grid of remeshFixed(4)
        // (remesh.cpp:14-36): x and z from -5 dx to L + 6 dx, y from -dx to Ly + 6 dx
        double stop = hi + dx;
        if (P.pbc) { lo = d == 1 ? -dx : -dx * 5.0; stop = P.L[d] + 6.0 * dx; hi = stop - dx; }
        uint32_t c = 0;
        for (double val = lo; val < stop; val += dx) {
            if (++c > H.cap_axis) break;
            if (c <= kCoordStage) s_coord[d][c - 1] = val;
        }
        cnt_s[d] = c;
        lo_s[d] = lo; hi_s[d] = hi; stop_s[d] = stop;
        mni_s[d] = mni; mxi_s[d] = mxi; loi_s[d] = loi; hii_s[d] = hii;
    }
    if (t < 3) {  // blocks along axis t and the multiplier of their quotient
        const uint32_t nb = t < D ? (cnt_s[t] + BlockEdge<D>::v - 1) / BlockEdge<D>::v : 1u;
        nb_s[t] = nb;
        block_div_magic(nb, nbm_s[t], nbl_s[t]);
    }
    __syncthreads();
    if (t == 0) {
        go = 0;
        if (!H.stop && H.frame >= P.end_frame) { H.stop = 1; H.finished = 1; }
        if (!H.stop) {
            double max_speed = sqrt(__longlong_as_double((long long)H.vmax2_bits));
            uint32_t warn = 0;
            if (max_speed >= P.wave_speed) warn |= MPM_WARN_SUPERSONIC;
            double dt;
            if (fabs(max_speed) > 1e-10) { double dt_cfl = P.cfl * P.dx / max_speed; dt = dt_cfl < P.dt_max ? dt_cfl : P.dt_max; }
            else dt = P.dt_max;
            double c1 = P.frame_dt * (double)(H.frame + 1) - H.time; dt = c1 < dt ? c1 : dt;
            double c2 = P.final_time - H.time; dt = c2 < dt ? c2 : dt;
            double grav[3] = {H.gravity[0], H.gravity[1], H.gravity[2]};
            if (P.gravity_special) {
                for (int d = 0; d < D; ++d) grav[d] = (H.time < P.gravity_time) ? P.gravity_final[d] * H.time / P.gravity_time : P.gravity_final[d];
            }
            // A NaN or infinite velocity or position anywhere shows up in these reductions (a NaN's order key
            // sorts above +inf or below -inf). The reference steps on with it (dt falls back to dt_max and the
            // remesh snap's fmax drops the NaN); here the run stops before the step.
            bool state_finite = isfinite(max_speed);
            for (int d = 0; d < D; ++d) state_finite = state_finite && isfinite(mn_s[d]) && isfinite(mx_s[d]);
            uint32_t cnt[3] = {1, 1, 1};
            bool fits = true;
            for (int d = 0; d < D; ++d) {
                cnt[d] = cnt_s[d];
                if (cnt[d] > H.cap_axis || cnt[d] < 1) fits = false;
            }
            unsigned long long nodes = (unsigned long long)cnt[0] * cnt[1] * cnt[2];
            if (nodes > H.cap_nodes) fits = false;
            const uint32_t nb[3] = {nb_s[0], nb_s[1], nb_s[2]};
            const unsigned long long nblocks = (unsigned long long)nb[0] * nb[1] * nb[2];
            if (nblocks > H.cap_blocks) fits = false;
            bool finite = state_finite;
            for (int d = 0; d < D; ++d)
                if (!(lo_s[d] == lo_s[d]) || !(hi_s[d] == hi_s[d]) || !isfinite(lo_s[d]) || !isfinite(hi_s[d]) || cnt[d] < 1) finite = false;
            if (!(dt == dt)) finite = false;
            if (!finite) {
                H.err |= MPM_ERR_NONFINITE; H.stop = 1;
            } else if (!fits) {
                H.need_realloc = 1; H.stop = 1;
                for (int d = 0; d < D; ++d) H.n[d] = cnt[d];
                H.nblocks = (uint32_t)nblocks;
            } else {
                H.dt = dt; H.warn |= warn;
                for (int d = 0; d < 3; ++d) H.gravity[d] = grav[d];
                for (int d = 0; d < D; ++d) {
                    H.min_init[d] = mni_s[d]; H.max_init[d] = mxi_s[d]; H.low_init[d] = loi_s[d]; H.high_init[d] = hii_s[d];
                    H.origin[d] = lo_s[d]; H.n[d] = cnt[d];
                }
                if (D == 2) { H.n[2] = 1; H.origin[2] = 0.0; }
                H.grid_nodes = (uint32_t)nodes;
                for (int d = 0; d < 3; ++d) { H.nb[d] = nb[d]; H.nb_m[d] = nbm_s[d]; H.nb_l[d] = nbl_s[d]; }
                H.nblocks = (uint32_t)nblocks;
                // reset the reductions that this step's particle kernels fill
                H.vmax2_bits = 0ull;
                for (int d = 0; d < 3; ++d) { H.xmin_key[d] = 0x7FFFFFFFFFFFFFFFLL; H.xmax_key[d] = (-0x7FFFFFFFFFFFFFFFLL - 1); }
                H.plastic_count = 0;
                H.cons_grid_mass = 0; for (int d = 0; d < 3; ++d) { H.cons_grid_mom[d] = 0; H.cons_part_mom[d] = 0; }
                go = 1;
            }
        }
    }
    __syncthreads();
    uint32_t* Swr = reinterpret_cast<uint32_t*>(S);
    for (int w = threadIdx.x; w < kHeadWords; w += 64) Swr[w] = head_w[w];
    if (go) {  // the accumulating sum of arange (tools.hpp:291-296) as the counting pass formed it
        for (int d = 0; d < D; ++d) {
            const uint32_t staged = min(cnt_s[d], kCoordStage);
            for (uint32_t i = (uint32_t)t; i < staged; i += 64) g.coord[d][i] = s_coord[d][i];
        }
        if (t < D && cnt_s[t] > kCoordStage) {  // the tail of a longer axis continues the same sum
            const int d = t;
            const double stop = stop_s[d];
            uint32_t i = kCoordStage;
            for (double val = s_coord[d][kCoordStage - 1] + dx; val < stop; val += dx) g.coord[d][i++] = val;
        }
    }
    if (ZERO && go) {  // the block histogram of this step (k_zero_grid's work, without a launch of its own)
        const uint32_t nb = H.nblocks << sb->ghosts, n4 = nb / 4;
        uint4* c4 = reinterpret_cast<uint4*>(sb->count);  // the histogram starts a 256-byte aligned buffer
        for (uint32_t i = (uint32_t)t; i < n4; i += 64) c4[i] = make_uint4(0u, 0u, 0u, 0u);
        for (uint32_t i = 4 * n4 + (uint32_t)t; i < nb; i += 64) sb->count[i] = 0;
    }
}
// The single-scene form also clears the block histogram.
template <int D, class TF>
__global__ __launch_bounds__(64) void k_setup(Params P, StepState* S, GridPtrsT<TF> g, SortBufs sb) { setup_body<D, TF, true>(P, S, g, &sb); }


// ----------------------------------------------------------------------------------------------
// k_zero_grid: clears the block histogram (the grid arrays are written whole by k_grid_update).
// ----------------------------------------------------------------------------------------------
__device__ __forceinline__ void zero_grid_body(const StepState* S, const SortBufs& sb, uint32_t bid, uint32_t nbid) {
    if (S->stop) return;
    const uint32_t nb = S->nblocks << sb.ghosts;
    for (uint32_t i = bid * blockDim.x + threadIdx.x; i < nb; i += nbid * blockDim.x) sb.count[i] = 0;
}
template <int = 0>  // a template, so that every translation unit may instantiate it
__global__ __launch_bounds__(kBlock) void k_zero_grid(const StepState* S, SortBufs sb) { zero_grid_body(S, sb, blockIdx.x, gridDim.x); }

// Block coordinates of linear block b: the quotients by the block counts as multiplies (block_div), about a
// tenth of the instructions of a division by a run-time divisor.
template <int D>
__device__ __forceinline__ void block_coords(uint32_t b, const StepState* S, int* bc) {
    if (D == 3) {
        const uint32_t r = block_div(b, S->nb_m[2], S->nb_l[2]), q = block_div(r, S->nb_m[1], S->nb_l[1]);
        bc[2] = (int)(b - r * S->nb[2]); bc[1] = (int)(r - q * S->nb[1]); bc[0] = (int)q;
    } else {
        const uint32_t q = block_div(b, S->nb_m[1], S->nb_l[1]);
        bc[1] = (int)(b - q * S->nb[1]); bc[0] = (int)q; bc[2] = 0;
    }
}

// Block of a particle from the first node of its non-zero window (P2G's clamped base). A particle whose
// window leaves the grid goes to block 0, where k_p2g reports it. `cell` receives the window base's dual
// cell within the block (BG::NC when the window leaves the grid). A window that fits the grid starts at
// most at node n - 3 < E nb, so its block needs no clamping and the cell is the base modulo the block edge;
// with these coordinates order_block's own window base (k_p2g) is the same.
template <int D, int SPL>
__device__ __forceinline__ uint32_t block_of_base(const int* wb, const StepState* S, uint32_t& cell) {
    constexpr uint32_t E = BlockEdge<D>::v;
    uint32_t c[3] = {0, 0, 0};
    bool in = true;
    for (int d = 0; d < D; ++d) {
        const uint32_t b = (uint32_t)wb[d] / E;
        in = in && b < S->nb[d];
        c[d] = b < S->nb[d] ? b : S->nb[d] - 1;
    }
    if (in) {
        const uint32_t l0 = (uint32_t)wb[0] % E, l1 = (uint32_t)wb[1] % E, l2 = D == 3 ? (uint32_t)wb[2] % E : 0u;
        cell = D == 3 ? (l0 * E + l1) * E + l2 : l0 * E + l1;
    }
    return D == 3 ? (c[0] * S->nb[1] + c[1]) * S->nb[2] + c[2] : c[0] * S->nb[1] + c[1];
}
template <int D, int SPL, class TF>
__device__ __forceinline__ uint32_t block_key(const double* xp, const StepState* S, const GridPtrsT<TF>& g, double h1, uint32_t& cell) {
    const double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
    const uint32_t nn[3] = {S->n[0], S->n[1], S->n[2]};
    int wb[3];
    cell = BG<D, SPL>::NC;
    if (!window_base<D, SPL>(wb, xp, org, nn, GlobalCoords{g.coord, nn}, h1, true)) return 0u;
    return block_of_base<D, SPL>(wb, S, cell);
}
template <int D, int SPL, class TF>
__device__ __forceinline__ uint32_t block_key(const double* xp, const StepState* S, const GridPtrsT<TF>& g, double h1) {
    uint32_t cell;
    return block_key<D, SPL, TF>(xp, S, g, h1, cell);
}

// Periodic copy of a particle at x0 (PBCAddParticles, pbc.cpp:37-83): within 4 dx of x = Lx (strictly
// inside) the copy sits at -(Lx - x0); otherwise within 4 dx of x = 0 it sits at Lx + x0.
__device__ __forceinline__ bool periodic_copy(double Lx, double band, double x0, double& gx) {
    double diff = Lx - x0;
    if (diff < band && diff > 0) { gx = -diff; return true; }
    diff = x0;
    if (diff < band && diff >= 0) { gx = Lx + diff; return true; }
    return false;
}

// k_bin_count: block of every particle and the histogram (wavefront-aggregated atomics); with periodic
// boundaries also the block of each particle's periodic copy (histogram entries nblocks + block). A thread takes
// PER particles (PER * blockDim apart, so that every load stays coalesced) and issues their position loads, then
// their coordinate loads, together: the keys of every lane are computed, inactive lanes' from in-bounds dummy
// positions, so that no branch separates the groups of loads.
constexpr int kBinPer = 4;  // particles per thread of the single-scene binning kernels
template <int D, int SPL, class TF, int PER = 1>
__device__ __forceinline__ void bin_count_body(const Params& P, const StepState* S, const PartStateT<TF>& src, const SortBufs& sb,
                                               const GridPtrsT<TF>& g, uint32_t n, uint32_t bid) {
    if (S->stop) return;
    const uint32_t p0 = bid * blockDim.x * PER + threadIdx.x;
    double xp[PER][3];
    bool active[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t p = p0 + (uint32_t)k * blockDim.x;
        active[k] = p < n;
#pragma unroll
        for (int d = 0; d < 3; ++d) xp[k][d] = d < D && active[k] ? src.x[d][p] : 0.0;
    }
    // block_key in three passes over the PER particles: window bases, the first window node's coordinate (an
    // in-bounds, clamped index for every lane), then the window base's block and cell
    const double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
    const uint32_t nn[3] = {S->n[0], S->n[1], S->n[2]};
    const GlobalCoords gc{g.coord, nn};
    int wb[PER][3];
    bool ok[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) ok[k] = stencil_base<D>(wb[k], xp[k], org, nn, P.one_over_dx, true);
    double c0[PER][3];
#pragma unroll
    for (int k = 0; k < PER; ++k)
#pragma unroll
        for (int d = 0; d < D; ++d) c0[k][d] = SPL == 2 ? gc(d, gc.clamp4(d, wb[k][d])) : 0.0;
    uint32_t key[PER], cell[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
#pragma unroll
        for (int d = 0; d < D; ++d)  // the shift of window_base
            if (SPL == 2) wb[k][d] += fabs((xp[k][d] - c0[k][d]) * P.one_over_dx) < 1.5 ? 0 : 1;
        cell[k] = BG<D, SPL>::NC;
        key[k] = ok[k] ? block_of_base<D, SPL>(wb[k], S, cell[k]) : 0u;
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t p = p0 + (uint32_t)k * blockDim.x;
        if (active[k]) { sb.key[p] = key[k]; sb.cellk[p] = (uint8_t)cell[k]; }
        (void)wave_claim(sb.count, active[k] ? key[k] : 0u, active[k]);
    }
    if (sb.ghosts) {
#pragma unroll
        for (int k = 0; k < PER; ++k) {
            const uint32_t p = p0 + (uint32_t)k * blockDim.x;
            double gx = 0.0;
            const bool has = active[k] && periodic_copy(P.L[0], 4.0 * P.dx, xp[k][0], gx);
            uint32_t gkey = 0, gcell = BG<D, SPL>::NC;
            if (has) {
                xp[k][0] = gx;
                gkey = S->nblocks + block_key<D, SPL, TF>(xp[k], S, g, P.one_over_dx, gcell);
            }
            if (active[k]) { sb.key[n + p] = has ? gkey : kNoGhost; sb.cellk[n + p] = (uint8_t)gcell; }
            (void)wave_claim(sb.count, gkey, has);
        }
    }
}
template <int D, int SPL, class TF, int PER>
__global__ __launch_bounds__(kBlock) void k_bin_count(Params P, const StepState* S, PartStateT<TF> src, SortBufs sb, GridPtrsT<TF> g, uint32_t n) {
    bin_count_body<D, SPL, TF, PER>(P, S, src, sb, g, n, blockIdx.x);
}

// k_bin_scan: exclusive prefix sum of the block histogram; one workgroup of 1024 threads. The counts pass through
// LDS in chunks of kScanPer * 1024: they are loaded and the offsets stored with consecutive threads on consecutive
// entries (whole cache lines per wavefront instruction), while the scan itself runs over runs of kScanPer
// consecutive entries per thread, read from LDS, with the run sums scanned within each wavefront by shuffles and
// across the 16 wavefronts by the first one.
constexpr int kScanPer = 16;
constexpr uint32_t kScanChunk = 1024u * kScanPer;
__device__ __forceinline__ void bin_scan_body(StepState* S, const SortBufs& sb) {
    const uint32_t stop = S->stop, nblocks = S->nblocks;  // read together; nblocks read once (the stores alias S)
    if (stop) return;
    __shared__ __attribute__((aligned(16))) uint32_t s_cnt[kScanChunk];
    __shared__ uint32_t wsum[16];
    __shared__ uint32_t s_ghost_off;
    const uint32_t nb = nblocks << sb.ghosts;
    const uint32_t tid = threadIdx.x;
    const int lane = tid & 63, wid = tid >> 6;
    uint32_t carry = 0;
    for (uint32_t c0 = 0; c0 < nb; c0 += kScanChunk) {
        const uint32_t m = min(kScanChunk, nb - c0);
        uint32_t v[kScanPer];
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) { const uint32_t e = (uint32_t)k * 1024u + tid; v[k] = e < m ? sb.count[c0 + e] : 0u; }
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) s_cnt[(uint32_t)k * 1024u + tid] = v[k];
        __syncthreads();
        uint4* run = reinterpret_cast<uint4*>(s_cnt + tid * kScanPer);
        uint32_t r[kScanPer];
#pragma unroll
        for (int k = 0; k < kScanPer / 4; ++k) { const uint4 q = run[k]; r[4 * k] = q.x; r[4 * k + 1] = q.y; r[4 * k + 2] = q.z; r[4 * k + 3] = q.w; }
        uint32_t sum = 0;
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) { const uint32_t c = r[k]; r[k] = sum; sum += c; }
        uint32_t incl = sum;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (lane >= o) incl += t;
        }
        if (lane == 63) wsum[wid] = incl;
        __syncthreads();
        if (wid == 0) {
            uint32_t w = lane < 16 ? wsum[lane] : 0u;
            for (int o = 1; o < 16; o <<= 1) {
                const uint32_t t = __shfl_up(w, o, 64);
                if (lane >= o) w += t;
            }
            if (lane < 16) wsum[lane] = w;  // inclusive over the wavefronts
        }
        __syncthreads();
        const uint32_t before = carry + (wid ? wsum[wid - 1] : 0u) + incl - sum;
#pragma unroll
        for (int k = 0; k < kScanPer / 4; ++k)
            run[k] = make_uint4(before + r[4 * k], before + r[4 * k + 1], before + r[4 * k + 2], before + r[4 * k + 3]);
        const uint32_t chunk_total = wsum[15];
        __syncthreads();
#pragma unroll
        for (int k = 0; k < kScanPer; ++k) {
            const uint32_t e = (uint32_t)k * 1024u + tid;
            if (e < m) {
                const uint32_t off = s_cnt[e];
                sb.offset[c0 + e] = off; sb.cursor[c0 + e] = off;
                if (sb.ghosts && c0 + e == nblocks) s_ghost_off = off;  // entries from here on are copies
            }
        }
        carry += chunk_total;
        __syncthreads();  // the next chunk overwrites s_cnt and wsum
    }
    if (sb.ghosts && tid == 0) S->n_ghosts = carry - s_ghost_off;
}
template <int = 0>  // a template, so that every translation unit may instantiate it
__global__ __launch_bounds__(1024) void k_bin_scan(StepState* S, SortBufs sb) { bin_scan_body(S, sb); }

// k_bin_scatter: counting-sort permutation (the order inside a block is fixed later by k_p2g). With
// periodic boundaries it runs over 2n entries: entry n + p is the periodic copy of p, if any.
// A thread takes PER entries, PER * blockDim apart, with their loads and their claims issued together.
template <int PER = 1>
__device__ __forceinline__ void bin_scatter_body(const StepState* S, const SortBufs& sb, uint32_t n, uint32_t bid) {
    if (S->stop) return;
    const uint32_t ne = n << sb.ghosts;
    const uint32_t e0 = bid * blockDim.x * PER + threadIdx.x;
    uint32_t key[PER], ck[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t e = e0 + (uint32_t)k * blockDim.x;
        key[k] = e < ne ? sb.key[e] : kNoGhost;
        ck[k] = e < ne ? sb.cellk[e] : 0u;
    }
    uint32_t pos[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) pos[k] = wave_claim(sb.cursor, key[k] != kNoGhost ? key[k] : 0u, key[k] != kNoGhost);
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        const uint32_t e = e0 + (uint32_t)k * blockDim.x;
        if (key[k] != kNoGhost) { sb.perm[pos[k]] = e < n ? e : (kGhostBit | (e - n)); sb.cells[pos[k]] = (uint8_t)ck[k]; }
    }
}
template <int PER>
__global__ __launch_bounds__(kBlock) void k_bin_scatter(const StepState* S, SortBufs sb, uint32_t n) { bin_scatter_body<PER>(S, sb, n, blockIdx.x); }

// k_binning: cell base index and block of every particle of a layout (diagnostic download).
template <int D, int SPL, class TF>
__global__ __launch_bounds__(kBlock) void k_binning(Params P, const StepState* S, PartStateT<TF> ps, GridPtrsT<TF> g, uint32_t n,
                                                     int32_t* base_out, uint32_t* key_out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double xp[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) xp[d] = ps.x[d][p];
    for (int d = 0; d < D; ++d) base_out[(size_t)p * D + d] = (int)floor((xp[d] - S->origin[d]) * P.one_over_dx) - 1;
    key_out[p] = block_key<D, SPL, TF>(xp, S, g, P.one_over_dx);
}

// ----------------------------------------------------------------------------------------------
// Block scatter (k_p2g, k_musl_scatter). One workgroup per block; lane l of each wavefront owns dual
// cell l (a block has 64 dual cells). The W first-axis window offsets ("slices") are split into NG
// groups of SPP slices, one wavefront per group (and per half of the second-axis offsets when the
// partial sums would not fit in registers). A lane adds the contributions of its cell's particles, in
// their sorted order, in registers; then the wavefronts add their partial sums into the block's LDS tile
// one after another, every lane to a different node at each instruction. The sums have a fixed order,
// and the tile goes to the staging area.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL, int NV, class TF, int REGS_ = 0> struct ScatterShape {
    using G = BG<D, SPL>;
    // 32-bit registers for the partial sums. The quadratic P2G scatter (NV > D) takes enough for one
    // wavefront to cover all or most of a particle's window slices, which saves the repeated stencil and
    // particle loads of the other wavefronts (C2 split: k_p2g 289 -> 234 us at 229 VGPRs); the MUSL
    // re-scatter (NV == D) and the cubic scatters were measured as fast or faster with 40.
    static constexpr int REGS = REGS_ > 0 ? REGS_ : ((NV > D && SPL == 2) ? 108 : 40);
    static constexpr int WORDS = NV * (int)(sizeof(TF) / 4);
    static constexpr int SPP0 = REGS / (G::NS * WORDS);
    static constexpr int SPP = SPP0 < 1 ? 1 : (SPP0 > G::W ? G::W : SPP0);   // slices per wavefront
    static constexpr int NG = (G::W + SPP - 1) / SPP;                  // slice groups
    static constexpr int SPLIT = (SPP0 < 1 && G::W % 2 == 0) ? 2 : 1;  // halves of the second-axis offsets
    static constexpr int OBN = G::W / SPLIT;                            // second-axis offsets per wavefront
    static constexpr int NSH = D == 3 ? OBN * G::W : OBN;              // window nodes per slice per wavefront
    static constexpr int WAVES = NG * SPLIT;
    static constexpr int TB = 64 * WAVES;                               // workgroup size
};

// Node coordinates around a block: axis d, nodes E*bc[d] - 1 .. E*bc[d] + E + 2 (every 4-node window of the
// block's particles), zero outside the grid.
template <int D, int SPL> struct BlockCoordLDS {
    static constexpr int LC = BG<D, SPL>::E + 4;
    double c[3][LC];
    // ONEPASS: one entry per thread over all axes (a workgroup has at least 64 threads), so that the D axes cost
    // one memory latency instead of D (the fused G2P + re-scatter kernel: 258 -> 245 us on C2). The P2G scatter
    // keeps the loop per axis: the one pass pushed its 3D quadratic split form past 256 registers (456 bytes of
    // spill).
    template <bool ONEPASS = false>
    __device__ __forceinline__ void load(double* const* coord, const int* bc, const uint32_t* n) {
        if constexpr (ONEPASS) {
            static_assert(D * LC <= 64, "one pass of a wavefront");
            const int t = threadIdx.x;
            if (t < D * LC) {
                const int d = t / LC, i = t - d * LC;
                const int gi = BG<D, SPL>::E * bc[d] - 1 + i;
                const double* ca = d == 0 ? coord[0] : (d == 1 ? coord[1] : coord[D == 3 ? 2 : 1]);  // selects, not a load
                c[d][i] = (gi >= 0 && gi < (int)n[d]) ? ca[gi] : 0.0;
            }
            return;
        }
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int e = BG<D, SPL>::E * bc[d] - 1, nd = (int)n[d];
            for (int i = threadIdx.x; i < LC; i += blockDim.x) {
                const int gi = e + i;
                c[d][i] = (gi >= 0 && gi < nd) ? coord[d][gi] : 0.0;
            }
        }
    }
    __device__ __forceinline__ LdsCoords<LC> view(const int* bc) const {
        LdsCoords<LC> v;
        v.c0 = c[0]; v.c1 = c[1]; v.c2 = c[2];
        v.lo0 = BG<D, SPL>::E * bc[0] - 1; v.lo1 = BG<D, SPL>::E * bc[1] - 1; v.lo2 = BG<D, SPL>::E * bc[2] - 1;
        return v;
    }
};

// The coordinate entries of BlockCoordLDS, entry t of every axis in thread t (a workgroup has at least 64
// threads), loaded into registers by issue() and written to LDS by put(): a block kernel issues the loads first
// and puts them after other work that waits for memory anyway, so the coordinates cost no latency of their own.
template <int D, int SPL> struct CoordPrefetch {
    static constexpr int LC = BlockCoordLDS<D, SPL>::LC;
    double v[3] = {0.0, 0.0, 0.0};
    __device__ __forceinline__ void issue(double* const* coord, const int* bc, const uint32_t* n) {
        static_assert(LC <= 64, "one pass of a wavefront per axis");
        const int t = threadIdx.x;
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int gi = BG<D, SPL>::E * bc[d] - 1 + t;
            v[d] = (t < LC && gi >= 0 && gi < (int)n[d]) ? coord[d][gi] : 0.0;
        }
    }
    __device__ __forceinline__ void put(BlockCoordLDS<D, SPL>& c) const {
        const int t = threadIdx.x;
#pragma unroll
        for (int d = 0; d < D; ++d) if (t < LC) c.c[d][t] = v[d];
    }
};

template <int D, int SPL, int NV, class TF> struct ScatterLDS {
    using G = BG<D, SPL>;
    BlockCoordLDS<D, SPL> coords;
    uint32_t p[kMaxBP];              // source slots of the block's particles, ordered by (cell, slot)
    uint32_t cstart[G::NC + 2];      // bucket c = cell c occupies p[cstart[c] .. cstart[c+1]); bucket NC: outside
    union {
        struct { uint32_t tmp[kMaxBP]; uint32_t ccount[G::NC + 1]; uint32_t ccur[G::NC + 1]; uint8_t cell[kMaxBP]; } sort;
        double tile[G::NT * NV];     // the block's node sums, [value][node] (binary64 in both forms, see add_partials)
    } u;
};

// Tile-local index of tile node (t0, t1, t2).
template <int D, int SPL>
__device__ __forceinline__ int tile_index(int t0, int t1, int t2) {
    using G = BG<D, SPL>;
    return D == 3 ? (t0 * G::T + t1) * G::T + t2 : t0 * G::T + t1;
}

// Component-major gather tiles in LDS (k_g2p, k_g2p_musl). For the 3D cubic spline the rows are padded to SY = 8
// words and the planes to SX words, SX = 4 modulo 8, so that the 32 lanes of a half wavefront, which gather the
// same window node of 32 consecutive dual cells (c0 in {0, 1} or {2, 3}, c1, c2 in 0..3), read 32 different banks
// of the 32: c1 SY + c2 covers four groups of four banks and SX shifts them into the other four. Measured on the
// inclined flow, k_g2p 602 -> 532 us. The quadratic tiles stay unpadded: the same padding removed 90 % of the fused
// quadratic G2P's LDS bank conflicts (C2) but added index arithmetic and ran about 6 % longer.
template <int D, int SPL> struct GatherTile {
    using G = BG<D, SPL>;
    static constexpr bool PAD = D == 3 && SPL == 3;
    static_assert(!PAD || G::T <= 8, "rows padded to 8 words");
    static constexpr int SY = PAD ? 8 : G::T;
    static constexpr int SX = PAD ? G::T * SY + 4 : G::T * G::T;  // padded: >= T SY, 4 modulo 8
    static constexpr int N = PAD ? G::T * SX : G::NT;             // words per component
    __device__ __forceinline__ static int index(int t0, int t1, int t2) {
        return PAD ? t0 * SX + t1 * SY + t2 : tile_index<D, SPL>(t0, t1, t2);
    }
};

// Orders a block's particles and writes the sorted permutation, the destination ids and the cell
// offsets. Within a dual cell the particles are ranked by source slot; the block's slots then hold rank 0
// of every non-empty cell in cell order, then rank 1 of every cell with at least two particles, and so
// on, so that the 64 lanes of a scatter wavefront (one per cell) read consecutive slots at every step and
// the next step's reads of this layout are coalesced. Particles whose window is outside the block (their
// window leaves the grid) go last and are reported. L.cstart keeps the cell-major offsets (cell counts).
// GHOST: the block's periodic copies (perm entries kGhostBit | p, at their copy positions); they are only
// ordered in LDS for the scatter, and nothing is written back.
constexpr int kOrderLoads = 8;
// Ascending sort of 8 values in registers: Batcher's odd-even merge network, 19 compare-exchanges (checked on
// every 0-1 input, which by the 0-1 principle covers all inputs).
__device__ __forceinline__ void sort8_u32(uint32_t (&r)[8]) {
    auto cx = [&](int i, int j) { const uint32_t a = r[i], b = r[j]; r[i] = min(a, b); r[j] = max(a, b); };
    cx(0, 1); cx(2, 3); cx(4, 5); cx(6, 7);
    cx(0, 2); cx(1, 3); cx(4, 6); cx(5, 7);
    cx(1, 2); cx(5, 6);
    cx(0, 4); cx(1, 5); cx(2, 6); cx(3, 7);
    cx(2, 4); cx(3, 5);
    cx(1, 2); cx(3, 4); cx(5, 6);
}
// dst[k] = src[k] for k < n, global memory to LDS, with kOrderLoads loads in flight per thread.
template <class T, class S>
__device__ __forceinline__ void copy_to_lds(T* dst, const S* src, uint32_t n) {
    for (uint32_t k0 = threadIdx.x; k0 < n; k0 += kOrderLoads * blockDim.x) {
        S v[kOrderLoads];
#pragma unroll
        for (int u = 0; u < kOrderLoads; ++u) {
            const uint32_t k = k0 + (uint32_t)u * blockDim.x;
            v[u] = k < n ? src[k] : S(0);
        }
#pragma unroll
        for (int u = 0; u < kOrderLoads; ++u) {
            const uint32_t k = k0 + (uint32_t)u * blockDim.x;
            if (k < n) dst[k] = (T)v[u];
        }
    }
}
template <int D, int SPL, int NV, class TF, bool GHOST = false, class LT = ScatterLDS<D, SPL, NV, TF>>
__device__ __forceinline__ uint32_t order_block(LT& L, const Params& P, const StepState* S, const PartStateT<TF>& src,
                                const PartStateT<TF>& dst, const SortBufs& sb, const GridPtrsT<TF>& g, uint32_t b, uint32_t start,
                                uint32_t cnt, const int* bc) {
    using G = BG<D, SPL>;
    const int tid = threadIdx.x;
    for (int c = tid; c <= G::NC; c += blockDim.x) L.u.sort.ccount[c] = 0;
    __syncthreads();
    uint32_t err = 0;
    // Each entry's dual cell comes from the binning (sb.cells, the window base that k_bin_count found with the
    // same coordinates). A thread takes kOrderLoads of the block's entries at a time and issues their slot and
    // cell loads together: one memory latency per group (a one-wavefront workgroup orders about 8 particles
    // per thread).
    for (uint32_t k0 = tid; k0 < cnt; k0 += kOrderLoads * blockDim.x) {
        uint32_t pk[kOrderLoads], ck[kOrderLoads];
#pragma unroll
        for (int u = 0; u < kOrderLoads; ++u) {
            const uint32_t k = k0 + (uint32_t)u * blockDim.x;
            pk[u] = k < cnt ? sb.perm[start + k] : 0u;
            ck[u] = k < cnt ? (uint32_t)sb.cells[start + k] : 0u;
        }
#pragma unroll
        for (int u = 0; u < kOrderLoads; ++u) {
            const uint32_t k = k0 + (uint32_t)u * blockDim.x;
            if (k >= cnt) break;
            const uint32_t cell = min(ck[u], (uint32_t)G::NC);
            if (cell == G::NC) err |= MPM_ERR_GRID_BOUNDS;
            L.u.sort.tmp[k] = pk[u];
            L.u.sort.cell[k] = (uint8_t)cell;
            atomicAdd(&L.u.sort.ccount[cell], 1u);
        }
    }
    __syncthreads();
    if (tid < 64) {  // exclusive scan of the 64 cell buckets by one wavefront; the outside bucket goes last
        const uint32_t v = L.u.sort.ccount[tid];
        uint32_t incl = v;
        for (int o = 1; o < 64; o <<= 1) {
            const uint32_t t = __shfl_up(incl, o, 64);
            if (tid >= o) incl += t;
        }
        L.cstart[tid + 1] = incl;
        L.u.sort.ccur[tid] = incl - v;
        if (tid == 0) L.cstart[0] = 0;
        if (tid == 63) { L.u.sort.ccur[64] = incl; L.cstart[65] = incl + L.u.sort.ccount[64]; }
    }
    __syncthreads();
    for (uint32_t k = tid; k < cnt; k += blockDim.x) {
        const uint32_t slot = atomicAdd(&L.u.sort.ccur[L.u.sort.cell[k]], 1u);
        L.p[slot] = L.u.sort.tmp[k];
    }
    __syncthreads();
    // Each bucket ordered by source slot (the entries are distinct), then laid out rank-major: slot order is
    // rank 0 of every non-empty cell in cell order, then rank 1, and so on. A one-wavefront workgroup whose
    // buckets hold at most 8 entries (and none outside the block) sorts each bucket in its lane's registers
    // with a 19-comparator network and writes the rank-major layout from them. Otherwise a one-wavefront
    // workgroup sorts by insertion, one lane per bucket, and a larger one ranks: one thread per slot counts the
    // entries of its bucket below its own, so that every wavefront takes part instead of waiting for the first
    // (measured: the 8-wavefront cubic scatter 3234 -> 2898 us on C3; the one-wavefront quadratic scatter of C2
    // was faster by insertion, 234 against 245 us).
    bool laid_out = false;
    if (blockDim.x <= 64) {
        const uint32_t s0 = L.cstart[tid], n_c = L.cstart[tid + 1] - s0;
        if (wave_max_u32(n_c) <= 8u && L.cstart[G::NC] == cnt) {
            uint32_t r[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) r[i] = (uint32_t)i < n_c ? L.p[s0 + i] : 0xFFFFFFFFu;
            sort8_u32(r);
            const unsigned long long below = (1ull << tid) - 1ull;
            uint32_t base = 0;
#pragma unroll
            for (int j = 0; j < 8; ++j) {  // every read of L.p above precedes these writes (one wavefront)
                const bool has = (uint32_t)j < n_c;
                const unsigned long long m = __ballot(has);
                if (m == 0ull) break;
                if (has) L.p[base + (uint32_t)__popcll(m & below)] = r[j];
                base += (uint32_t)__popcll(m);
            }
            laid_out = true;
        }
    }
    if (!laid_out) {
        if (blockDim.x <= 64) {
            for (int c = tid; c <= G::NC; c += blockDim.x) {
                const uint32_t s0 = L.cstart[c], s1 = L.cstart[c + 1];
                for (uint32_t i = s0 + 1; i < s1; ++i) {
                    const uint32_t key = L.p[i];
                    uint32_t j = i;
                    while (j > s0 && L.p[j - 1] > key) { L.p[j] = L.p[j - 1]; --j; }
                    L.p[j] = key;
                }
            }
            __syncthreads();
            for (uint32_t k = tid; k < cnt; k += blockDim.x) L.u.sort.tmp[k] = L.p[k];
        } else {
            for (uint32_t i = tid; i < cnt; i += blockDim.x) {
                int c = 0;  // the last bucket starting at or before slot i
#pragma unroll
                for (int step = 64; step > 0; step >>= 1)
                    if (c + step <= G::NC && L.cstart[c + step] <= i) c += step;
                const uint32_t s0 = L.cstart[c], s1 = L.cstart[c + 1], key = L.p[i];
                uint32_t r = 0;
                for (uint32_t j = s0; j < s1; ++j) r += L.p[j] < key ? 1u : 0u;
                L.u.sort.tmp[s0 + r] = key;
            }
        }
        __syncthreads();
        if (tid < 64) {  // rank-major layout: one wavefront, lane = cell
            const uint32_t c0 = L.cstart[tid], n_c = L.cstart[tid + 1] - c0;
            const unsigned long long below = (1ull << tid) - 1ull;
            uint32_t base = 0;
            for (uint32_t j = 0;; ++j) {
                const bool has = j < n_c;
                const unsigned long long m = __ballot(has);
                if (m == 0ull) break;
                if (has) L.p[base + (uint32_t)__popcll(m & below)] = L.u.sort.tmp[c0 + j];
                base += (uint32_t)__popcll(m);
            }
        }
        for (uint32_t k = L.cstart[G::NC] + tid; k < cnt; k += blockDim.x) L.p[k] = L.u.sort.tmp[k];  // outside bucket last
    }
    __syncthreads();
    if (GHOST) return err;
    // The sorted permutation. The destination ids (dst.id[slot] = src.id[perm[slot]]) are written by the
    // deformation kernel, which loads perm[slot] anyway; here their loads would add a memory latency per block.
    for (uint32_t k = tid; k < cnt; k += blockDim.x) sb.perm[start + k] = L.p[k];
    for (int c = tid; c <= G::NC; c += blockDim.x) sb.cell_off[(size_t)b * (G::NC + 1) + c] = (uint16_t)L.cstart[c];
    return err;
}

template <int D, int SPL, int NV, class TF, class LT = ScatterLDS<D, SPL, NV, TF>>
__device__ __forceinline__ void zero_tile(LT& L) {
    for (int q = threadIdx.x; q < BG<D, SPL>::NT * NV; q += blockDim.x) L.u.tile[q] = 0.0;
}

template <int D, int SPL, int NV, class TF, class LT = ScatterLDS<D, SPL, NV, TF>>
__device__ __forceinline__ void write_tile(const LT& L, const StagingT<TF>& stg, uint32_t b) {
    using G = BG<D, SPL>;
    for (int q = threadIdx.x; q < G::NT; q += blockDim.x)
        for (int v = 0; v < NV; ++v) stg.base[(size_t)v * stg.stride + (size_t)b * G::NT + q] = (TF)L.u.tile[v * G::NT + q];
}

// DPP quad permutation: lane 4k + i takes the value of lane 4k + Pi (every lane active).
template <int P0, int P1, int P2, int P3, class T>
__device__ __forceinline__ T quad_perm(T v) {
    constexpr int ctrl = P0 | (P1 << 2) | (P2 << 4) | (P3 << 6);
    if constexpr (sizeof(T) == 4) {
        return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), ctrl, 0xF, 0xF, false));
    } else {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_mov_dpp((int)(b & 0xFFFFFFFFll), ctrl, 0xF, 0xF, false);
        const int hi = __builtin_amdgcn_mov_dpp((int)(b >> 32), ctrl, 0xF, 0xF, false);
        return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
    }
}

// Adds one wavefront's partial sums acc (slices a0 .. a0+SPP-1, second-axis half `half`) of dual cell
// `cell` into the tile. At each instruction every lane adds to a different node, but a later instruction
// may add to a node that another lane updated earlier, so the adds are LDS atomics: a plain load/add/store
// would let the compiler move later loads above earlier stores (legal for one lane, a lost update across
// lanes). A wavefront's LDS instructions execute in program order, so each node still receives its terms
// in a fixed order. The tile is binary64 in the split form too: LDS binary32 atomic adds issue about ten
// times slower than binary64 ones on gfx950 (measured: ~80 ns against ~7 ns per wavefront instruction
// per CU for this access pattern). A read-add-write per node with a compiler barrier between the nodes (the
// order kept by the program order of the LDS instructions) ran slower than the atomics: the cubic 8-value
// P2G 1642 -> 1714 us, the quadratic split P2G and fused G2P about 5 % slower.
template <int D, int SPL, int NV, class TF, class SS = ScatterShape<D, SPL, NV, TF>, class LT = ScatterLDS<D, SPL, NV, TF>>
__device__ __forceinline__ void add_partials(LT& L, int cell, int a0, int half, const TF (&acc)[SS::SPP][SS::NSH][NV]) {
    using G = BG<D, SPL>;
    int c[3] = {0, 0, 0};
    if (D == 3) { c[2] = cell % G::E; c[1] = (cell / G::E) % G::E; c[0] = cell / (G::E * G::E); }
    else { c[1] = cell % G::E; c[0] = cell / G::E; }
#pragma unroll
    for (int s = 0; s < SS::SPP; ++s) {
        const int a = a0 + s;
        if (a >= G::W) break;
#pragma unroll
        for (int obl = 0; obl < SS::OBN; ++obl) {
            const int ob = half * SS::OBN + obl;
            if constexpr (D == 3) {
                // The W last-axis partial sums of the row of four cells (c0, c1, 0..3) go to nodes c2 .. c2 + W - 1;
                // they are summed across the row's lanes first (DPP quad permutations), so that lane c2 adds node c2
                // ("lo": terms oc <= c2 from lanes c2 - oc) and, for c2 <= W - 2, node c2 + 4 ("hi": terms oc > c2
                // from lanes c2 + 4 - oc): two LDS atomics per value instead of W. Each node's terms are summed in a
                // fixed order.
                static_assert(G::E == 4, "rows of four cells: one DPP quad");
                const int qlo = tile_index<D, SPL>(c[0] + a, c[1] + ob, c[2]), qhi = tile_index<D, SPL>(c[0] + a, c[1] + ob, c[2] + 4);
                const bool has_hi = c[2] <= G::W - 2;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const TF* r = &acc[s][obl * G::W][0];
                    TF lo = r[v], hi = TF(0);
                    static_for<G::W - 1>([&](auto K) {
                        constexpr int oc = decltype(K)::value + 1;
                        const TF x = r[oc * NV + v];
                        constexpr int d1 = (1 - oc) < 0 ? 0 : (1 - oc), d2 = (2 - oc) < 0 ? 0 : (2 - oc);
                        constexpr int u1 = (5 - oc) < 3 ? (5 - oc) : 3, u2 = (6 - oc) < 3 ? (6 - oc) : 3;
                        const TF dn = quad_perm<0, d1, d2, 3 - oc>(x);  // lane i <- i - oc (lanes i < oc unused)
                        const TF up = quad_perm<4 - oc, u1, u2, 3>(x);  // lane i <- i + 4 - oc (lanes i >= oc unused)
                        lo += c[2] >= oc ? dn : TF(0);
                        hi += c[2] < oc ? up : TF(0);
                    });
                    atomicAdd(&L.u.tile[v * G::NT + qlo], (double)lo);
                    if (has_hi) atomicAdd(&L.u.tile[v * G::NT + qhi], (double)hi);
                }
            } else {
                const int q = tile_index<D, SPL>(c[0] + a, c[1] + ob, 0);
#pragma unroll
                for (int v = 0; v < NV; ++v) atomicAdd(&L.u.tile[v * G::NT + q], (double)acc[s][obl][v]);
            }
        }
    }
}

// Ordered accumulation (see above) of the block's particles into its tile, written to staging.
// contrib.load(p, i, part) fetches source particle p (sorted slot i); contrib.add(part, a0, half, acc, base)
// adds it to acc for slices a0.. and second-axis half `half` and returns false if its window leaves the grid.
template <int D, int SPL, int NV, class TF, class Contrib, bool NEXT = false>
__device__ __forceinline__ void accumulate_cells(const ScatterLDS<D, SPL, NV, TF>& L, const Contrib& contrib, uint32_t start,
                                                 TF (&acc)[ScatterShape<D, SPL, NV, TF>::SPP][ScatterShape<D, SPL, NV, TF>::NSH][NV]) {
    using SS = ScatterShape<D, SPL, NV, TF>;
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int grp = wave % SS::NG, half = wave / SS::NG, a0 = grp * SS::SPP;
    {  // rank-major layout (see order_block): at rank j the cells with more than j particles are consecutive
        const uint32_t n_c = L.cstart[lane + 1] - L.cstart[lane];
        const uint32_t n_max = wave_max_u32(n_c);
        const unsigned long long below = (1ull << lane) - 1ull;
        uint32_t kb = 0;
        int base[3];
        // NEXT: load the next rank's particle while this one is added (kP2GNextRank)
        if constexpr (NEXT) {
            typename Contrib::Part qn;
            bool hasn = 0 < n_c;
            {
                const unsigned long long m = __ballot(hasn);
                const uint32_t k = kb + (uint32_t)__popcll(m & below);
                kb += (uint32_t)__popcll(m);
                if (hasn) contrib.load(L.p[k], start + k, qn);
            }
            for (uint32_t j = 0; j < n_max; ++j) {
                const typename Contrib::Part q = qn;
                const bool has = hasn;
                hasn = j + 1 < n_c;
                const unsigned long long m = __ballot(hasn);
                const uint32_t k = kb + (uint32_t)__popcll(m & below);
                kb += (uint32_t)__popcll(m);
                if (hasn) contrib.load(L.p[k], start + k, qn);
                if (has) (void)contrib.add(q, a0, half, acc, base);
            }
        } else {
            for (uint32_t j = 0; j < n_max; ++j) {
                const bool has = j < n_c;
                const unsigned long long m = __ballot(has);
                const uint32_t k = kb + (uint32_t)__popcll(m & below);
                kb += (uint32_t)__popcll(m);
                if (has) {
                    typename Contrib::Part q;
                    contrib.load(L.p[k], start + k, q);
                    (void)contrib.add(q, a0, half, acc, base);
                }
            }
        }
    }
}

// The wavefronts' partial sums into the zeroed tile, one wavefront after another, and the tile to staging.
// The wavefronts' partial sums added into the zeroed tile one wavefront after another (no write to staging).
template <int D, int SPL, int NV, class TF, class SS = ScatterShape<D, SPL, NV, TF>, class LT = ScatterLDS<D, SPL, NV, TF>>
__device__ __forceinline__ void ordered_tile(LT& L, const TF (&acc)[SS::SPP][SS::NSH][NV]) {
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int grp = wave % SS::NG, half = wave / SS::NG, a0 = grp * SS::SPP;
    zero_tile<D, SPL, NV, TF>(L);
    for (int w = 0; w < SS::WAVES; ++w) {
        __syncthreads();
        if (wave == w) add_partials<D, SPL, NV, TF, SS>(L, lane, a0, half, acc);
    }
    __syncthreads();
}

template <int D, int SPL, int NV, class TF>
__device__ __forceinline__ void finish_tile(ScatterLDS<D, SPL, NV, TF>& L, const StagingT<TF>& stg, uint32_t b,
                                            const TF (&acc)[ScatterShape<D, SPL, NV, TF>::SPP][ScatterShape<D, SPL, NV, TF>::NSH][NV]) {
    ordered_tile<D, SPL, NV, TF>(L, acc);
    write_tile<D, SPL, NV, TF>(L, stg, b);
}

// Clears a partial-sum array of any register shape (its extents are the array's own).
template <int D, int SPL, int NV, class TF, int SPP, int NSH>
__device__ __forceinline__ void zero_acc(TF (&acc)[SPP][NSH][NV]) {
#pragma unroll
    for (int s = 0; s < SPP; ++s)
#pragma unroll
        for (int o = 0; o < NSH; ++o)
#pragma unroll
            for (int v = 0; v < NV; ++v) acc[s][o][v] = TF(0);
}

template <int D, int SPL, int NV, class TF, class Contrib, bool NEXT = false>
__device__ __forceinline__ void scatter_block(ScatterLDS<D, SPL, NV, TF>& L, const Contrib& contrib, const StagingT<TF>& stg, uint32_t b,
                              uint32_t start) {
    using SS = ScatterShape<D, SPL, NV, TF>;
    TF acc[SS::SPP][SS::NSH][NV];
    zero_acc<D, SPL, NV, TF>(acc);
    accumulate_cells<D, SPL, NV, TF, Contrib, NEXT>(L, contrib, start, acc);
    finish_tile<D, SPL, NV, TF>(L, stg, b, acc);
}

// Tile of a block with more than kMaxBP particles: every particle adds its window with LDS atomics.
// Sorted slots start .. start+cnt-1, then gstart .. gstart+gcnt-1 (the block's periodic copies, if any).
template <int D, int SPL, int NV, class TF, class Contrib, class SS = ScatterShape<D, SPL, NV, TF>, class LT = ScatterLDS<D, SPL, NV, TF>>
__device__ __forceinline__ uint32_t scatter_block_atomic(LT& L, const Contrib& contrib, const StagingT<TF>& stg, const SortBufs& sb,
                                         uint32_t b, uint32_t start, uint32_t cnt, const int* bc, uint32_t gstart = 0, uint32_t gcnt = 0) {
    using G = BG<D, SPL>;
    const int tid = threadIdx.x;
    __syncthreads();
    zero_tile<D, SPL, NV, TF>(L);
    __syncthreads();
    uint32_t err = 0;
    for (uint32_t k = tid; k < cnt + gcnt; k += blockDim.x) {
        const uint32_t i = k < cnt ? start + k : gstart + (k - cnt), p = sb.perm[i];
        typename Contrib::Part q;
        contrib.load(p, i, q);
        bool ok = true;
        for (int grp = 0; grp < SS::NG && ok; ++grp) {
            for (int half = 0; half < SS::SPLIT && ok; ++half) {
                const int a0 = grp * SS::SPP;
                TF acc[SS::SPP][SS::NSH][NV];
#pragma unroll
                for (int s = 0; s < SS::SPP; ++s)
#pragma unroll
                    for (int o = 0; o < SS::NSH; ++o)
#pragma unroll
                        for (int v = 0; v < NV; ++v) acc[s][o][v] = TF(0);
                int base[3] = {0, 0, 0};
                ok = contrib.add(q, a0, half, acc, base);
                int l[3] = {0, 0, 0};
                for (int d = 0; d < D; ++d) { l[d] = base[d] - G::E * bc[d]; ok = ok && l[d] >= 0 && l[d] < G::E; }
                if (!ok) { err |= MPM_ERR_GRID_BOUNDS; break; }
#pragma unroll
                for (int s = 0; s < SS::SPP; ++s) {
                    const int a = a0 + s;
                    if (a >= G::W) break;
#pragma unroll
                    for (int obl = 0; obl < SS::OBN; ++obl) {
                        const int ob = half * SS::OBN + obl;
                        if (D == 3) {
#pragma unroll
                            for (int oc = 0; oc < G::W; ++oc) {
                                const int qn = tile_index<D, SPL>(l[0] + a, l[1] + ob, l[2] + oc);
#pragma unroll
                                for (int v = 0; v < NV; ++v) atomicAdd(&L.u.tile[v * G::NT + qn], (double)acc[s][obl * G::W + oc][v]);
                            }
                        } else {
                            const int qn = tile_index<D, SPL>(l[0] + a, l[1] + ob, 0);
#pragma unroll
                            for (int v = 0; v < NV; ++v) atomicAdd(&L.u.tile[v * G::NT + qn], (double)acc[s][obl][v]);
                        }
                    }
                }
            }
        }
    }
    __syncthreads();
    write_tile<D, SPL, NV, TF>(L, stg, b);
    return err;
}

// Multiply-add used by the scatter arithmetic (a fused multiply-add in both forms).
template <class TF> __device__ __forceinline__ TF madd(TF a, TF b, TF c) { return fma(a, b, c); }

// The window of one particle as a scatter wavefront needs it: the first non-zero node wb per axis
// (make_stencil's base, with the same shift test), and weight, derivative and node-minus-particle offset of
// the first-axis nodes a0 .. a0+SPP-1, the second-axis nodes of `half` and all third-axis nodes. Every
// value is the expression make_stencil evaluates on the same operands; they are rounded to TF.
// a0 and half are uniform across the wavefront, so the piece of a first- or second-axis node is picked
// by a scalar branch.
template <int D, int SPL, int SPP, int OBN, class TF>
struct SliceStencil {
    static constexpr int W = BG<D, SPL>::W;
    int wb[3];
    TF w0[SPP], dw0[SPP], pd0[SPP];
    TF w1[OBN], dw1[OBN], pd1[OBN];
    TF w2[W], dw2[W], pd2[W];
};

template <int SPL, class CS, class TF>
__device__ __forceinline__ void node_vals(const CS& cs, int d, int i0, int k, double x, double h1, TF& w, TF& dw, TF& pd) {
    const double xi = cs(d, i0 + k);
    if (k == 0) window_node<SPL, 0>(x, xi, h1, w, dw);
    else if (k == 1) window_node<SPL, 1>(x, xi, h1, w, dw);
    else if (SPL == 3 && k == 3) window_node<SPL, SPL == 3 ? 3 : 2>(x, xi, h1, w, dw);
    else window_node<SPL, 2>(x, xi, h1, w, dw);
    pd = (TF)(xi - x);
}

template <int D, int SPL, int SPP, int OBN, class TF, class CS>
__device__ __forceinline__ bool slice_stencil(SliceStencil<D, SPL, SPP, OBN, TF>& ss, const double* xp, const double* org,
                                              const uint32_t* n, const CS& cs, double h1, bool clamp, int a0, int half) {
    constexpr int W = BG<D, SPL>::W;
    int b[3] = {0, 0, 0};
    const bool ok = stencil_base<D>(b, xp, org, n, h1, clamp);
    int i0[3] = {0, 0, 0};
#pragma unroll
    for (int d = 0; d < D; ++d) {
        const int bb = cs.clamp4(d, b[d]);
        int sh = 0;
        if (SPL == 2) {  // make_stencil skips the base node when its weight is exactly zero, i.e. |t| >= 1.5
            const double t = (xp[d] - cs(d, bb)) * h1;
            sh = MPM_SEL(fabs(t) < 1.5) ? 0 : 1;
        }
        ss.wb[d] = b[d] + sh;
        i0[d] = bb + sh;
    }
#pragma unroll
    for (int s = 0; s < SPP; ++s) {
        const int a = min(a0 + s, W - 1);
        node_vals<SPL>(cs, 0, i0[0], a, xp[0], h1, ss.w0[s], ss.dw0[s], ss.pd0[s]);
    }
#pragma unroll
    for (int o = 0; o < OBN; ++o) node_vals<SPL>(cs, 1, i0[1], half * OBN + o, xp[1], h1, ss.w1[o], ss.dw1[o], ss.pd1[o]);
    if (D == 3) {
#pragma unroll
        for (int o = 0; o < W; ++o) node_vals<SPL>(cs, 2, i0[2], o, xp[2], h1, ss.w2[o], ss.dw2[o], ss.pd2[o]);
    }
    return ok;
}

// ----------------------------------------------------------------------------------------------
// Staged scatter records (3D split form, cubic spline). A block's particles are taken in chunks of kStageCH
// slots; for each (particle, axis) pair one thread loads what that axis needs and writes the particle's
// per-axis factors to LDS, and the cell-owning wavefronts accumulate from LDS instead of each loading the
// particle and evaluating its stencil again. The value groups factor by window row (a, b) and last-axis node c:
//   mass      w0 w1 w2
//   momentum  w0 w1 w2 (v + B~ (x_i - x_p)) = w2 X(a,b) + w0 w1 Z(c),  X = w0 w1 (U0(a) + P1(b)),
//             U0(a) = v + B~[:,0] (x_0a - x_p0), P1(b) = B~[:,1] (x_1b - x_p1), Z(c) = w2 B~[:,2] (x_2c - x_p2)
//   force     T grad w / h1 = w2 F(a,b) + dw2 H(a,b),  F = T[:,0] dw0 w1 + T[:,1] w0 dw1,  H = T[:,2] w0 w1
// (B~ = D^-1 B, T the stress scaled as the force takes it). Record of chunk slot s, structure of arrays (value
// v at rec[v * kStageCH + s]):
//   per first-axis node a:  w0, dw0, U0[3]      per second-axis node b: w1, dw1, P1[3]
//   per last-axis node c:   w2, dw2, Z[3]       T[r][d] (9), MINW[d] (3), muI (MIBF)
// MINW: the smallest window weight per axis; a product of the three below 1e-24 (a node near the reference's
// 1e-25 weight cut) sends the particle to the direct per-node form instead.
// ----------------------------------------------------------------------------------------------
// Slots per chunk: measured on the 7.95 M-particle inclined flow with two blocks per CU (the P2G bounded to 128
// VGPRs), the cubic MIBF P2G ran 2005, 2286, 1631, 1886 and 1697 us at 64, 96, 128, 160 and 192; 256 and more
// leave room for one block per CU (over 80 KB of LDS).
constexpr int kStageCH = 128;

template <int D, int SPL, bool MU> struct StagedRec {
    static constexpr int W = BG<D, SPL>::W;
    static constexpr int OT = D * W * 5, OMIN = OT + D * D, OMU = OMIN + D, N = OMU + (MU ? 1 : 0);
    // axis d, node k: weight, derivative, then the three components of U0 / P1 / Z
    __device__ __forceinline__ static constexpr int w(int d, int k) { return (d * W + k) * 5; }
    __device__ __forceinline__ static constexpr int dw(int d, int k) { return (d * W + k) * 5 + 1; }
    __device__ __forceinline__ static constexpr int u(int d, int k, int r) { return (d * W + k) * 5 + 2 + r; }
    __device__ __forceinline__ static constexpr int t(int r, int d) { return OT + r * D + d; }
};

// Values of one (particle, axis) pair of a staged chunk, read from the chunk's raw values in LDS.
template <class TF> struct StagePre { double x, mu; TF b[3], t[3], v[3]; };
// A staged chunk's particle values travel from global memory to LDS without registers (global_load_lds): the
// loads of the next chunk are issued before the current chunk's accumulation and land during it. Raw value f of
// chunk slot s is the 32-bit word raw[f * kStageCH + s]; a binary64 value takes two words (low, high).
__device__ __forceinline__ double raw_f64(const uint32_t* raw, int f, int s) {
    return __hiloint2double((int)raw[(f + 1) * kStageCH + s], (int)raw[f * kStageCH + s]);
}
__device__ __forceinline__ float raw_f32(const uint32_t* raw, int f, int s) { return __uint_as_float(raw[f * kStageCH + s]); }

// Wavefronts of a staged scatter: one first-axis slice each, and the second axis split in halves when a
// slice's window nodes would need more than 64 partial sums per lane.
template <int D, int SPL, int NV> struct StagedShape {
    using G = BG<D, SPL>;
    static constexpr int SPP = 1, NG = G::W;
    static constexpr int SPLIT = (G::W % 2 == 0 && G::W * G::W * NV > 64) ? 2 : 1;
    static constexpr int OBN = G::W / SPLIT, NSH = OBN * G::W, WAVES = NG * SPLIT, TB = 64 * WAVES;
};

// One axis of a particle's window as the clamped P2G / MUSL stencil takes it (stencil_base + the shift of
// make_stencil): weights, derivatives and node-minus-particle offsets of the non-zero nodes; returns the first
// node of the non-zero window.
template <int SPL, class CS, class TF>
__device__ __forceinline__ int axis_window(const CS& cs, int d, double x, double org, double h1, TF* w, TF* dw, TF* pd) {
    constexpr int W = SPL == 2 ? 3 : 4;
    const double f = floor((x - org) * h1);
    int v = f < 2.0e9 ? (f > -2.0e9 ? (int)f - 1 : -0x3FFFFFFF) : 0x3FFFFFFF;
    if (v < 0) v = 0;
    const int bb = cs.clamp4(d, v);
    int sh = 0;
    if (SPL == 2) sh = MPM_SEL(fabs((x - cs(d, bb)) * h1) < 1.5) ? 0 : 1;
#pragma unroll
    for (int k = 0; k < W; ++k) node_vals<SPL>(cs, d, bb + sh, k, x, h1, w[k], dw[k], pd[k]);
    return v + sh;
}

// P2G contributions (p2g.cpp:12-66 and the force of explicit_euler_update.cpp:15-168): value 0 sum of
// weights, 1..D momentum sum w (v + B (x_i - x_p) D^-1) (weight > 1e-25 only), then either the force sum
// tau grad w in separate values (SEP) or the force folded into the momentum as - dt V0 / m_p tau grad w,
// and last the MIBF friction sum w muI. The products are arranged separably: B (x_i - x_p) D^-1 =
// sum_k D^-1 B[:,k] (x_i - x_p)_k with one term per axis node, and grad w = h1 (dw0 w1 w2, w0 dw1 w2,
// w0 w1 dw2) with h1 folded into the stress.
template <int D, int SPL, bool SEP, bool MIBF, class TF>
struct P2GContrib {
    static constexpr int NV = 1 + D + (SEP ? D : 0) + (MIBF ? 1 : 0);
    using SS = ScatterShape<D, SPL, NV, TF>;
    struct Part { double x[3]; TF v[3], B[9], t6[6]; double muI; };
    PartStateT<TF> src;
    LdsCoords<BG<D, SPL>::E + 4> cs;
    double org[3];
    uint32_t nn[3];
    double h1, aD, kf;
    double Lx, band;  // periodic interval and copy band (periodic copies only)
    bool apic;
    // Source particle p, or (kGhostBit | p) its periodic copy: the same state at the copy's position.
    __device__ __forceinline__ void load(uint32_t e, uint32_t, Part& q) const {
        const uint32_t p = e & ~kGhostBit;
        for (int d = 0; d < D; ++d) { q.x[d] = src.x[d][p]; q.v[d] = src.v[d][p]; }
        for (int k = 0; k < D * D; ++k) q.B[k] = apic ? src.B[k][p] : TF(0);
        for (int k = 0; k < SymPack<D>::n; ++k) q.t6[k] = src.tau[k][p];
        q.muI = MIBF ? src.muI[p] : 0.0;
        if (e & kGhostBit) { double gx = q.x[0]; (void)periodic_copy(Lx, band, q.x[0], gx); q.x[0] = gx; }
    }
    __device__ __forceinline__ bool add(const Part& q, int a0, int half, TF (&acc)[SS::SPP][SS::NSH][NV], int* base) const {
        constexpr int W = BG<D, SPL>::W;
        SliceStencil<D, SPL, SS::SPP, SS::OBN, TF> ss;
        const bool ok = slice_stencil(ss, q.x, org, nn, cs, h1, true, a0, half);
        for (int d = 0; d < D; ++d) base[d] = ss.wb[d];
        // stress columns scaled by h1 (force) or by -dt V0 / m_p h1 (folded into the momentum)
        TF tau[3][3];
        if (D == 3) { tau[0][0] = q.t6[0]; tau[1][1] = q.t6[1]; tau[2][2] = q.t6[2]; tau[0][1] = tau[1][0] = q.t6[3]; tau[0][2] = tau[2][0] = q.t6[4]; tau[1][2] = tau[2][1] = q.t6[5]; }
        else { tau[0][0] = q.t6[0]; tau[1][1] = q.t6[1]; tau[0][1] = tau[1][0] = q.t6[2]; }
        const TF fs = SEP ? (TF)h1 : (TF)(-kf * h1);
        TF T[3][3];
        for (int r = 0; r < D; ++r) for (int k = 0; k < D; ++k) T[r][k] = tau[r][k] * fs;
        const TF aDt = (TF)aD;
        TF Bc[3][3];  // D^-1 B[:,k]
        for (int r = 0; r < D; ++r) for (int k = 0; k < D; ++k) Bc[r][k] = q.B[r + k * D] * aDt;
        const TF muI = (TF)q.muI;
        // Factored form, used when no node of this wavefront's part of the window can fall under the
        // reference's 1e-25 weight cut (smallest per-axis weights multiplied, with a factor 10 of margin for
        // rounding): with w = w0 w1 w2 and grad w = h1 (dw0 w1 w2, w0 dw1 w2, w0 w1 dw2), every value is
        // w1 w2 A0[a] + w0 w2 C1[b] + w0 w1 C2[c], where A0 carries the velocity, the first-axis affine term
        // and the first-axis force term, C1 and C2 the other axes' affine and force terms. In 3D a window row
        // (a, b) is summed as w2 X + w0 w1 C2[c] with X = w1 A0 + w0 C1: 7 operations per node (mass and the
        // folded momentum) instead of about 24.
        TF mw0 = TF(1), mw1 = ss.w1[0], mw2 = TF(1);
#pragma unroll
        for (int s = 0; s < SS::SPP; ++s) if (a0 + s < W) mw0 = s == 0 ? ss.w0[0] : fmin(mw0, ss.w0[s]);
#pragma unroll
        for (int o = 1; o < SS::OBN; ++o) mw1 = fmin(mw1, ss.w1[o]);
        if (D == 3) {
            mw2 = ss.w2[0];
#pragma unroll
            for (int o = 1; o < W; ++o) mw2 = fmin(mw2, ss.w2[o]);
        }
        // Split form only: the plain form keeps the reference's per-node products, so that its sums round as
        // the reference's do.
        constexpr bool kFactored = sizeof(TF) == 4;
        if (kFactored && MPM_SEL(mw0 * mw1 * mw2 > TF(1e-24))) {
            if constexpr (kFactored) add_factored(ss, q, a0, T, Bc, muI, acc);
            return ok;
        }
#pragma unroll
        for (int s = 0; s < SS::SPP; ++s) {
            if (a0 + s >= W) break;
            TF u0[3];
            for (int r = 0; r < D; ++r) u0[r] = madd(Bc[r][0], ss.pd0[s], q.v[r]);
#pragma unroll
            for (int o1 = 0; o1 < SS::OBN; ++o1) {
                const TF w01 = ss.w0[s] * ss.w1[o1], dw0w1 = ss.dw0[s] * ss.w1[o1], w0dw1 = ss.w0[s] * ss.dw1[o1];
                TF u01[3];
                for (int r = 0; r < D; ++r) u01[r] = madd(Bc[r][1], ss.pd1[o1], u0[r]);
                if (D == 3) {
#pragma unroll
                    for (int o2 = 0; o2 < W; ++o2) {
                        const TF wt = w01 * ss.w2[o2];
                        const TF g[3] = {dw0w1 * ss.w2[o2], w0dw1 * ss.w2[o2], w01 * ss.dw2[o2]};
                        const TF wm = MPM_SEL(wt > TF(1e-25)) ? wt : TF(0);
                        TF* a = acc[s][o1 * W + o2];
                        a[0] += wm;
                        for (int r = 0; r < D; ++r) {
                            const TF u = madd(Bc[r][2], ss.pd2[o2], u01[r]);
                            TF f = T[r][0] * g[0];
                            f = madd(T[r][1], g[1], f);
                            f = madd(T[r][2], g[2], f);
                            if (SEP) { a[1 + r] = madd(wm, u, a[1 + r]); a[1 + D + r] += f; }
                            else a[1 + r] = madd(wm, u, a[1 + r] + f);
                        }
                        if (MIBF) a[NV - 1] = madd(wm, muI, a[NV - 1]);
                    }
                } else {
                    const TF wt = w01;
                    const TF g[2] = {dw0w1, w0dw1};
                    const TF wm = MPM_SEL(wt > TF(1e-25)) ? wt : TF(0);
                    TF* a = acc[s][o1];
                    a[0] += wm;
                    for (int r = 0; r < D; ++r) {
                        TF f = T[r][0] * g[0];
                        f = madd(T[r][1], g[1], f);
                        if (SEP) { a[1 + r] = madd(wm, u01[r], a[1 + r]); a[1 + D + r] += f; }
                        else a[1 + r] = madd(wm, u01[r], a[1 + r] + f);
                    }
                    if (MIBF) a[NV - 1] = madd(wm, muI, a[NV - 1]);
                }
            }
        }
        return ok;
    }
    // The factored form of add (see there); every node of the wavefront's part of the window has w > 1e-25.
    template <class SSt>
    __device__ __forceinline__ void add_factored(const SSt& ss, const Part& q, int a0, const TF (&T)[3][3], const TF (&Bc)[3][3],
                                                 TF muI, TF (&acc)[SS::SPP][SS::NSH][NV]) const {
        constexpr int W = BG<D, SPL>::W;
        constexpr int NC = SEP ? 2 : 1;  // momentum and force separately (SEP), or summed
        // last-axis terms C2[c] (3D), per node of the axis
        TF C2[D == 3 ? W : 1][NC][3];
        if (D == 3) {
#pragma unroll
            for (int o = 0; o < W; ++o)
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    const TF m = ss.w2[o] * (Bc[r][2] * ss.pd2[o]), f = T[r][2] * ss.dw2[o];
                    if (SEP) { C2[o][0][r] = m; C2[o][1][r] = f; } else C2[o][0][r] = m + f;
                }
        }
#pragma unroll
        for (int s = 0; s < SS::SPP; ++s) {
            if (a0 + s >= W) break;
            TF A0[NC][3];
#pragma unroll
            for (int r = 0; r < D; ++r) {
                const TF m = ss.w0[s] * madd(Bc[r][0], ss.pd0[s], q.v[r]), f = T[r][0] * ss.dw0[s];
                if (SEP) { A0[0][r] = m; A0[1][r] = f; } else A0[0][r] = m + f;
            }
#pragma unroll
            for (int o1 = 0; o1 < SS::OBN; ++o1) {
                TF C1[NC][3];
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    const TF m = ss.w1[o1] * (Bc[r][1] * ss.pd1[o1]), f = T[r][1] * ss.dw1[o1];
                    if (SEP) { C1[0][r] = m; C1[1][r] = f; } else C1[0][r] = m + f;
                }
                const TF w01 = ss.w0[s] * ss.w1[o1];
                if (D == 3) {
                    // row (a, b): w1 w2 A0 + w0 w2 C1 = w2 X with X = w1 A0 + w0 C1, so that a node adds
                    // w2 X + w0 w1 C2[c] (two multiply-adds per value instead of three and two products).
                    // Rows are scheduled one after another: hoisting every row's X ahead of the sums would
                    // keep them all live and push the kernel past 256 registers (one wavefront per SIMD).
                    __builtin_amdgcn_sched_barrier(0);
                    TF X[NC][3];
#pragma unroll
                    for (int c = 0; c < NC; ++c)
#pragma unroll
                        for (int r = 0; r < D; ++r) X[c][r] = madd(ss.w1[o1], A0[c][r], ss.w0[s] * C1[c][r]);
                    const TF w01m = MIBF ? w01 * muI : TF(0);
#pragma unroll
                    for (int o2 = 0; o2 < W; ++o2) {
                        TF* a = acc[s][o1 * W + o2];
                        a[0] = madd(w01, ss.w2[o2], a[0]);
#pragma unroll
                        for (int c = 0; c < NC; ++c)
#pragma unroll
                            for (int r = 0; r < D; ++r)
                                a[1 + c * D + r] = madd(w01, C2[o2][c][r], madd(ss.w2[o2], X[c][r], a[1 + c * D + r]));
                        if (MIBF) a[NV - 1] = madd(w01m, ss.w2[o2], a[NV - 1]);
                    }
                } else {
                    TF* a = acc[s][o1];
                    a[0] += w01;
#pragma unroll
                    for (int c = 0; c < NC; ++c)
#pragma unroll
                        for (int r = 0; r < D; ++r) a[1 + c * D + r] = madd(ss.w1[o1], A0[c][r], madd(ss.w0[s], C1[c][r], a[1 + c * D + r]));
                    if (MIBF) a[NV - 1] = madd(w01, muI, a[NV - 1]);
                }
            }
        }
    }

    // ---- staged scatter (3D split form, cubic spline) ----
    using SRec = StagedRec<D, SPL, MIBF>;
    using Pre = StagePre<TF>;
    // Raw values of a source particle p: x (6 words), B (9), tau (6), v (3), muI (2, MIBF).
    static constexpr int NRAW = 24 + (MIBF ? 2 : 0);
    template <int f>
    __device__ __forceinline__ bool raw_used() const { return apic || f < 6 || f >= 15; }
    template <int f>
    __device__ __forceinline__ const void* raw_addr(uint32_t p, uint32_t) const {
        if constexpr (f < 6) return reinterpret_cast<const uint32_t*>(src.x[f / 2] + p) + (f & 1);
        else if constexpr (f < 15) return src.B[f - 6] + p;
        else if constexpr (f < 21) return src.tau[f - 15] + p;
        else if constexpr (f < 24) return src.v[f - 21] + p;
        else return reinterpret_cast<const uint32_t*>(src.muI + p) + (f - 24);
    }
    // Axis d of chunk slot s (position, column d of B and of tau, velocity and muI on axis 0).
    template <int d>
    __device__ __forceinline__ void load_axis(const uint32_t* raw, int s, Pre& q) const {
        q.x = raw_f64(raw, 2 * d, s);
#pragma unroll
        for (int r = 0; r < D; ++r) {
            q.b[r] = apic ? raw_f32(raw, 6 + r + d * D, s) : TF(0);
            q.t[r] = raw_f32(raw, 15 + (r == d ? r : (r + d == 1 ? 3 : (r + d == 2 ? 4 : 5))), s);  // tau[r][d], (xx, yy, zz, xy, xz, yz)
            q.v[r] = d == 0 ? raw_f32(raw, 21 + r, s) : TF(0);
        }
        q.mu = (MIBF && d == 0) ? raw_f64(raw, 24, s) : 0.0;
    }
    // Phase 1: the axis-d part of the record of chunk slot s.
    template <int d>
    __device__ __forceinline__ void stage_axis(const Pre& q, TF* rec, int s) const {
        constexpr int W = BG<D, SPL>::W;
        TF w[W], dw[W], pd[W];
        (void)axis_window<SPL>(cs, d, q.x, org[d], h1, w, dw, pd);
        const TF aDt = (TF)aD;
        const TF fs = SEP ? (TF)h1 : (TF)(-kf * h1);
        TF mn = w[0];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            rec[SRec::w(d, k) * kStageCH + s] = w[k];
            rec[SRec::dw(d, k) * kStageCH + s] = dw[k];
            mn = fmin(mn, w[k]);
#pragma unroll
            for (int r = 0; r < D; ++r) {
                const TF bp = (q.b[r] * aDt) * pd[k];
                rec[SRec::u(d, k, r) * kStageCH + s] = d == 0 ? q.v[r] + bp : (d == 1 ? bp : w[k] * bp);
            }
        }
#pragma unroll
        for (int r = 0; r < D; ++r) rec[SRec::t(r, d) * kStageCH + s] = q.t[r] * fs;
        rec[(SRec::OMIN + d) * kStageCH + s] = mn;
        if (MIBF && d == 0) rec[SRec::OMU * kStageCH + s] = (TF)q.mu;
    }
    // Phase 2: the particle of chunk slot s into the partial sums of first-axis node a and second-axis half
    // `half`; false (nothing added) when a window weight is near the 1e-25 cut.
    template <class SSt>
    __device__ __forceinline__ bool accumulate_rec(const TF* rec, int s, int a, int half, TF (&acc)[1][SSt::NSH][NV]) const {
        constexpr int W = BG<D, SPL>::W;
        const TF mw = rec[SRec::OMIN * kStageCH + s] * rec[(SRec::OMIN + 1) * kStageCH + s] * rec[(SRec::OMIN + 2) * kStageCH + s];
        if (!MPM_SEL(mw > TF(1e-24))) return false;
        const TF w0 = rec[SRec::w(0, a) * kStageCH + s], dw0 = rec[SRec::dw(0, a) * kStageCH + s];
        const TF muI = MIBF ? rec[SRec::OMU * kStageCH + s] : TF(0);
        TF U0[3], T[3][3];
#pragma unroll
        for (int r = 0; r < D; ++r) {
            U0[r] = rec[SRec::u(0, a, r) * kStageCH + s];
#pragma unroll
            for (int d = 0; d < D; ++d) T[r][d] = rec[SRec::t(r, d) * kStageCH + s];
        }
#pragma unroll
        for (int o1 = 0; o1 < SSt::OBN; ++o1) {
            // rows one after another: reading every row's record values ahead of the arithmetic would keep them
            // all live beside the partial sums
            __builtin_amdgcn_sched_barrier(0);
            const int b = half * SSt::OBN + o1;
            const TF w1 = rec[SRec::w(1, b) * kStageCH + s], dw1 = rec[SRec::dw(1, b) * kStageCH + s];
            const TF w01 = w0 * w1, dw0w1 = dw0 * w1, w0dw1 = w0 * dw1;
            TF X[3], F[3], H[3];
#pragma unroll
            for (int r = 0; r < D; ++r) {
                X[r] = w01 * (U0[r] + rec[SRec::u(1, b, r) * kStageCH + s]);
                F[r] = madd(T[r][1], w0dw1, T[r][0] * dw0w1);
                H[r] = T[r][2] * w01;
                if (!SEP) X[r] = X[r] + F[r];  // momentum and force summed: w2 (X + F) + w01 Z + dw2 H
            }
            const TF w01mu = w01 * muI;
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const TF w2 = rec[SRec::w(2, c) * kStageCH + s], dw2 = rec[SRec::dw(2, c) * kStageCH + s];
                TF* A = acc[0][o1 * W + c];
                A[0] = madd(w01, w2, A[0]);
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    const TF Z = rec[SRec::u(2, c, r) * kStageCH + s];
                    if (SEP) {
                        A[1 + r] = madd(w2, X[r], madd(w01, Z, A[1 + r]));
                        A[1 + D + r] = madd(w2, F[r], madd(dw2, H[r], A[1 + D + r]));
                    } else {
                        A[1 + r] = madd(w2, X[r], madd(w01, Z, madd(dw2, H[r], A[1 + r])));
                    }
                }
                if (MIBF) A[NV - 1] = madd(w01mu, w2, A[NV - 1]);
            }
        }
        return true;
    }
    // Direct form (add's per-node products and the 1e-25 weight cut) of one particle, for particles near the
    // cut and blocks too large to order.
    struct Win { TF w[3][4], dw[3][4], pd[3][4]; int wb[3]; TF v[3], Bc[3][3], T[3][3]; TF muI; };
    __device__ __forceinline__ void window_of(uint32_t p, uint32_t, Win& z) const {
#pragma unroll
        for (int d = 0; d < D; ++d) z.wb[d] = axis_window<SPL>(cs, d, src.x[d][p], org[d], h1, z.w[d], z.dw[d], z.pd[d]);
        const TF aDt = (TF)aD;
        const TF fs = SEP ? (TF)h1 : (TF)(-kf * h1);
        TF t6[6];
        for (int k = 0; k < SymPack<D>::n; ++k) t6[k] = src.tau[k][p];
        TF tau[3][3];
        tau[0][0] = t6[0]; tau[1][1] = t6[1]; tau[2][2] = t6[2]; tau[0][1] = tau[1][0] = t6[3]; tau[0][2] = tau[2][0] = t6[4]; tau[1][2] = tau[2][1] = t6[5];
        for (int r = 0; r < D; ++r) {
            z.v[r] = src.v[r][p];
            for (int k = 0; k < D; ++k) { z.Bc[r][k] = apic ? src.B[r + k * D][p] * aDt : TF(0); z.T[r][k] = tau[r][k] * fs; }
        }
        z.muI = MIBF ? (TF)src.muI[p] : TF(0);
    }
    __device__ __forceinline__ void node_direct(const Win& z, int a, int b, int c, TF* out) const {
        TF u01[3];
        for (int r = 0; r < D; ++r) u01[r] = madd(z.Bc[r][1], z.pd[1][b], madd(z.Bc[r][0], z.pd[0][a], z.v[r]));
        const TF w01 = z.w[0][a] * z.w[1][b], dw0w1 = z.dw[0][a] * z.w[1][b], w0dw1 = z.w[0][a] * z.dw[1][b];
        const TF wt = w01 * z.w[2][c];
        const TF g[3] = {dw0w1 * z.w[2][c], w0dw1 * z.w[2][c], w01 * z.dw[2][c]};
        const TF wm = wt > TF(1e-25) ? wt : TF(0);
        out[0] = wm;
        for (int r = 0; r < D; ++r) {
            const TF u = madd(z.Bc[r][2], z.pd[2][c], u01[r]);
            TF f = z.T[r][0] * g[0];
            f = madd(z.T[r][1], g[1], f);
            f = madd(z.T[r][2], g[2], f);
            if (SEP) { out[1 + r] = wm * u; out[1 + D + r] = f; }
            else out[1 + r] = madd(wm, u, f);
        }
        if (MIBF) out[NV - 1] = wm * z.muI;
    }
};

// The MUSL re-scatter terms of window slice a of one particle (musl.cpp:41-56), added to acc[b W + c][r]:
// the weight w0 w1[b] w2[c], dropped at or under the reference's 1e-25 cut, times
// v_r + Bc_r0 pd0 + Bc_r1 pd1[b] + Bc_r2 pd2[c] (Bc = B D^-1, pd the node offsets x_i - x_p). The fused G2P +
// re-scatter kernel and k_musl_scatter both take this function, so their tiles are equal bit for bit. (A
// row-factored form, w2[c] X[b] + w0 w1[b] Z[c], took 20 fewer instructions per node but more registers: the
// fused kernel spilled more and ran 278 -> 298 us on C2.)
template <int W, class TF>
__device__ __forceinline__ void musl_slice(TF w0, TF pd0, const TF (&w1)[W], const TF (&pd1)[W], const TF (&w2)[W], const TF (&pd2)[W],
                                           const TF (&v)[3], const TF (&Bc)[3][3], TF (&acc)[W * W][3]) {
    TF u0[3];
#pragma unroll
    for (int r = 0; r < 3; ++r) u0[r] = madd(Bc[r][0], pd0, v[r]);
#pragma unroll
    for (int b = 0; b < W; ++b) {
        const TF w01 = w0 * w1[b];
        TF u01[3];
#pragma unroll
        for (int r = 0; r < 3; ++r) u01[r] = madd(Bc[r][1], pd1[b], u0[r]);
#pragma unroll
        for (int c = 0; c < W; ++c) {
            const TF wt = w01 * w2[c];
            const TF wm = MPM_SEL(wt > TF(1e-25)) ? wt : TF(0);
#pragma unroll
            for (int r = 0; r < 3; ++r) acc[b * W + c][r] = madd(wm, madd(Bc[r][2], pd2[c], u01[r]), acc[b * W + c][r]);
        }
    }
}
// Slices through musl_slice: the 3D split form, where a wavefront's partial sums cover whole slices.
template <int D, class SS, class TF> constexpr bool kMuslSlices = D == 3 && sizeof(TF) == 4 && SS::SPLIT == 1 && SS::OBN == SS::G::W;

// MUSL re-scatter contributions (musl.cpp:28-85): sum w (v_new + B_new (x_i - x_p) D^-1) over
// weight > 1e-25, with the position of the step's start (source layout, slot p) and the new velocity
// and affine matrix (destination layout, sorted slot i), arranged separably as in P2GContrib.
template <int D, int SPL, class TF>
struct MuslContrib {
    static constexpr int NV = D;
    using SS = ScatterShape<D, SPL, NV, TF>;
    struct Part { double x[3]; TF v[3], B[9]; };
    PartStateT<TF> src, dst;
    LdsCoords<BG<D, SPL>::E + 4> cs;
    double org[3];
    uint32_t nn[3];
    double h1, aD;
    bool apic;
    __device__ __forceinline__ void load(uint32_t p, uint32_t i, Part& q) const {
        for (int d = 0; d < D; ++d) { q.x[d] = src.x[d][p]; q.v[d] = dst.v[d][i]; }
        for (int k = 0; k < D * D; ++k) q.B[k] = apic ? dst.B[k][i] : TF(0);
    }
    __device__ __forceinline__ bool add(const Part& q, int a0, int half, TF (&acc)[SS::SPP][SS::NSH][NV], int* base) const {
        constexpr int W = BG<D, SPL>::W;
        SliceStencil<D, SPL, SS::SPP, SS::OBN, TF> ss;
        const bool ok = slice_stencil(ss, q.x, org, nn, cs, h1, true, a0, half);
        for (int d = 0; d < D; ++d) base[d] = ss.wb[d];
        const TF aDt = (TF)aD;
        TF Bc[3][3];
        for (int r = 0; r < D; ++r) for (int k = 0; k < D; ++k) Bc[r][k] = q.B[r + k * D] * aDt;
        if constexpr (kMuslSlices<D, SS, TF>) {
#pragma unroll
            for (int s = 0; s < SS::SPP; ++s) {
                if (a0 + s >= W) break;
                musl_slice<W, TF>(ss.w0[s], ss.pd0[s], ss.w1, ss.pd1, ss.w2, ss.pd2, q.v, Bc, acc[s]);
            }
            return ok;
        }
#pragma unroll
        for (int s = 0; s < SS::SPP; ++s) {
            if (a0 + s >= W) break;
            TF u0[3];
            for (int r = 0; r < D; ++r) u0[r] = madd(Bc[r][0], ss.pd0[s], q.v[r]);
#pragma unroll
            for (int o1 = 0; o1 < SS::OBN; ++o1) {
                const TF w01 = ss.w0[s] * ss.w1[o1];
                TF u01[3];
                for (int r = 0; r < D; ++r) u01[r] = madd(Bc[r][1], ss.pd1[o1], u0[r]);
                if (D == 3) {
#pragma unroll
                    for (int o2 = 0; o2 < W; ++o2) {
                        const TF wt = w01 * ss.w2[o2];
                        const TF wm = MPM_SEL(wt > TF(1e-25)) ? wt : TF(0);
                        TF* a = acc[s][o1 * W + o2];
                        for (int r = 0; r < D; ++r) a[r] = madd(wm, madd(Bc[r][2], ss.pd2[o2], u01[r]), a[r]);
                    }
                } else {
                    const TF wm = MPM_SEL(w01 > TF(1e-25)) ? w01 : TF(0);
                    TF* a = acc[s][o1];
                    for (int r = 0; r < D; ++r) a[r] = madd(wm, u01[r], a[r]);
                }
            }
        }
        return ok;
    }

    // ---- staged scatter (3D split form, cubic spline): the momentum sum w2 X(a,b) + w0 w1 Z(c) ----
    using SRec = StagedRec<D, SPL, false>;
    using Pre = StagePre<TF>;
    // Raw values: x of source particle p (6 words), B (9) and v (3) of destination slot i.
    static constexpr int NRAW = 18;
    template <int f>
    __device__ __forceinline__ bool raw_used() const { return apic || f < 6 || f >= 15; }
    template <int f>
    __device__ __forceinline__ const void* raw_addr(uint32_t p, uint32_t i) const {
        if constexpr (f < 6) return reinterpret_cast<const uint32_t*>(src.x[f / 2] + p) + (f & 1);
        else if constexpr (f < 15) return dst.B[f - 6] + i;
        else return dst.v[f - 15] + i;
    }
    template <int d>
    __device__ __forceinline__ void load_axis(const uint32_t* raw, int s, Pre& q) const {
        q.x = raw_f64(raw, 2 * d, s);
#pragma unroll
        for (int r = 0; r < D; ++r) {
            q.b[r] = apic ? raw_f32(raw, 6 + r + d * D, s) : TF(0);
            q.v[r] = d == 0 ? raw_f32(raw, 15 + r, s) : TF(0);
            q.t[r] = TF(0);
        }
        q.mu = 0.0;
    }
    template <int d>
    __device__ __forceinline__ void stage_axis(const Pre& q, TF* rec, int s) const {
        constexpr int W = BG<D, SPL>::W;
        TF w[W], dw[W], pd[W];
        (void)axis_window<SPL>(cs, d, q.x, org[d], h1, w, dw, pd);
        const TF aDt = (TF)aD;
        TF mn = w[0];
#pragma unroll
        for (int k = 0; k < W; ++k) {
            rec[SRec::w(d, k) * kStageCH + s] = w[k];
            mn = fmin(mn, w[k]);
#pragma unroll
            for (int r = 0; r < D; ++r) {
                const TF bp = (q.b[r] * aDt) * pd[k];
                rec[SRec::u(d, k, r) * kStageCH + s] = d == 0 ? q.v[r] + bp : (d == 1 ? bp : w[k] * bp);
            }
        }
        rec[(SRec::OMIN + d) * kStageCH + s] = mn;
    }
    template <class SSt>
    __device__ __forceinline__ bool accumulate_rec(const TF* rec, int s, int a, int half, TF (&acc)[1][SSt::NSH][NV]) const {
        constexpr int W = BG<D, SPL>::W;
        const TF mw = rec[SRec::OMIN * kStageCH + s] * rec[(SRec::OMIN + 1) * kStageCH + s] * rec[(SRec::OMIN + 2) * kStageCH + s];
        if (!MPM_SEL(mw > TF(1e-24))) return false;
        const TF w0 = rec[SRec::w(0, a) * kStageCH + s];
        TF U0[3];
#pragma unroll
        for (int r = 0; r < D; ++r) U0[r] = rec[SRec::u(0, a, r) * kStageCH + s];
#pragma unroll
        for (int o1 = 0; o1 < SSt::OBN; ++o1) {
            const int b = half * SSt::OBN + o1;
            const TF w01 = w0 * rec[SRec::w(1, b) * kStageCH + s];
            TF X[3];
#pragma unroll
            for (int r = 0; r < D; ++r) X[r] = w01 * (U0[r] + rec[SRec::u(1, b, r) * kStageCH + s]);
#pragma unroll
            for (int c = 0; c < W; ++c) {
                const TF w2 = rec[SRec::w(2, c) * kStageCH + s];
                TF* A = acc[0][o1 * W + c];
#pragma unroll
                for (int r = 0; r < D; ++r) A[r] = madd(w2, X[r], madd(w01, rec[SRec::u(2, c, r) * kStageCH + s], A[r]));
            }
        }
        return true;
    }
    struct Win { TF w[3][4], dw[3][4], pd[3][4]; int wb[3]; TF v[3], Bc[3][3]; };
    __device__ __forceinline__ void window_of(uint32_t p, uint32_t i, Win& z) const {
#pragma unroll
        for (int d = 0; d < D; ++d) z.wb[d] = axis_window<SPL>(cs, d, src.x[d][p], org[d], h1, z.w[d], z.dw[d], z.pd[d]);
        const TF aDt = (TF)aD;
        for (int r = 0; r < D; ++r) {
            z.v[r] = dst.v[r][i];
            for (int k = 0; k < D; ++k) z.Bc[r][k] = apic ? dst.B[r + k * D][i] * aDt : TF(0);
        }
    }
    __device__ __forceinline__ void node_direct(const Win& z, int a, int b, int c, TF* out) const {
        const TF w01 = z.w[0][a] * z.w[1][b];
        const TF wt = w01 * z.w[2][c];
        const TF wm = wt > TF(1e-25) ? wt : TF(0);
        for (int r = 0; r < D; ++r) {
            const TF u01 = madd(z.Bc[r][1], z.pd[1][b], madd(z.Bc[r][0], z.pd[0][a], z.v[r]));
            out[r] = wm * madd(z.Bc[r][2], z.pd[2][c], u01);
        }
    }
};

// ----------------------------------------------------------------------------------------------
// Staged scatter driver (see the record layout above). The 3D split form's cubic P2G and MUSL re-scatter use
// it; other forms keep scatter_block. The cubic scatter's wavefronts each cover 8 of a particle's 64 window
// nodes, so in scatter_block all 8 loaded every particle and evaluated its stencil again (C3: 2.48 -> 1.95 ms
// staged). The quadratic scatter covers a particle's 27 nodes in one wavefront and has nothing to share: staged,
// it ran 340 instead of 233 us on the 4,096,000-particle collapse.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL, class TF, bool PBC> constexpr bool kStagedScatter = D == 3 && SPL == 3 && sizeof(TF) == 4 && !PBC;

// Workgroup size of a scatter kernel: the staged shape or the register shape of scatter_block.
template <int D, int SPL, int NV, class TF, bool PBC> struct ScatterLaunch {
    static constexpr int TB = kStagedScatter<D, SPL, TF, PBC> ? StagedShape<D, SPL, NV>::TB : ScatterShape<D, SPL, NV, TF>::TB;
    // Minimum wavefronts per SIMD that the compiler must allow (amdgpu_waves_per_eu). The staged P2G scatter
    // is bounded to 4 (at most 128 VGPRs): two 8-wavefront blocks per CU instead of one at 158 VGPRs, at the
    // price of 18 spilled registers (C3: 2.10 -> 1.70 ms). The 4-value payload (folded force, no MIBF) does not fit in
    // 128 and other scatters are not bounded (1). Only the single-scene kernel takes the bound: the batched one spills
    // more under it (420 bytes), and with it and the deformation bound the served stream of inclined flows and
    // collapses ran 1.74e9 -> 1.49e9 particle-updates/s (1.70e9 with both bounds on the single-scene kernels only).
    // The one-wavefront quadratic 3D scatter keeps at least 2 (at most 256 VGPRs): its 108 registers of partial
    // sums, the next rank's particle and the row factors otherwise reach 258 with three accumulation registers
    // kept in AGPRs, and one wavefront per SIMD.
    static constexpr int WPE = kStagedScatter<D, SPL, TF, PBC> && NV > D + 1 ? 4 : (D == 3 && SPL == 2 ? 2 : 1);
};

template <int D, int SPL, int NV, class TF, int NR> struct StagedLDS {
    using G = BG<D, SPL>;
    BlockCoordLDS<D, SPL> coords;
    uint32_t p[kMaxBP];              // source slots, ordered as order_block orders them
    uint32_t cstart[G::NC + 2];
    uint32_t flag[kMaxBP / 32];      // slots whose window weights are near the 1e-25 cut (direct form)
    union {
        struct { uint32_t tmp[kMaxBP]; uint32_t ccount[G::NC + 1]; uint32_t ccur[G::NC + 1]; uint8_t cell[kMaxBP]; } sort;
        double tile[G::NT * NV];
        TF rec[NR * kStageCH];       // the records of one chunk
    } u;
};

// Phases 1 and 2 over the block's ordered particles, chunk by chunk. A chunk is a run of whole ranks (the
// rank-major layout of order_block: rank j holds the j-th particle of every cell with more than j) of at most
// kStageCH slots; a rank has at most 64 slots, so every chunk holds at least one. All wavefronts see the same
// 64 cells, hence the same chunks and barriers.
// Chunk boundaries: the run of whole ranks from rank j (slot kb) holding at most kStageCH slots; a rank has at
// most 64 slots (the rank-major layout of order_block), so a chunk holds at least one. Returns the end rank.
__device__ __forceinline__ uint32_t stage_chunk_end(uint32_t j, uint32_t kb, uint32_t n_max, uint32_t n_c, uint32_t& kb1) {
    uint32_t j1 = j;
    kb1 = kb;
    while (j1 < n_max) {
        const uint32_t c = (uint32_t)__popcll(__ballot(j1 < n_c));
        if (kb1 + c - kb > (uint32_t)kStageCH) break;
        kb1 += c;
        ++j1;
    }
    return j1;
}

// Issues the raw-value loads of the chunk of slots kb .. kb + m - 1 straight into LDS: load (f, g) brings raw value
// f of the 64 slots of group g (G = kStageCH / 64 groups) and goes to wavefront (f * G + g) mod NW. Completion is
// awaited with wait_vmem and made visible to the other wavefronts by a barrier.
using lds_u32 = __attribute__((address_space(3))) uint32_t;
template <int NW, class Contrib>
__device__ __forceinline__ void stage_prefetch(const Contrib& cf, lds_u32* raw, const uint32_t* Lp, uint32_t start, uint32_t kb, uint32_t m) {
    constexpr int G = kStageCH / 64;
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    // Every lane loads: a slot past the chunk (m >= 1) repeats the chunk's last one, into a word nobody reads; no
    // per-lane condition stays live across the loads. Every group's slots are read from LDS before the first
    // load (an LDS read after a load to LDS waits for it).
    uint32_t p[G], i[G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
        const uint32_t s = min((uint32_t)(g * 64 + lane), m - 1u);
        p[g] = Lp[kb + s];
        i[g] = start + kb + s;
    }
    // one branch on the wavefront index, then that wavefront's loads with compile-time fields
    static_for<NW>([&](auto WV) {
        constexpr int w = decltype(WV)::value;
        if (wave == w) {
            static_for<Contrib::NRAW * G>([&](auto QV) {
                constexpr int q = decltype(QV)::value, f = q / G, g = q % G;
                if constexpr (q % NW == w) {
                    if (cf.template raw_used<f>())
                        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)cf.template raw_addr<f>(p[g], i[g]),
                                                         raw + f * kStageCH + g * 64, 4, 0, 0);
                }
            });
        }
    });
}
// Waits for this wavefront's outstanding vector memory operations (the LDS-direct loads among them).
__device__ __forceinline__ void wait_vmem() { __builtin_amdgcn_s_waitcnt(0x0F70); }  // vmcnt(0); expcnt, lgkmcnt not waited

// Phase 1 of the chunk of m slots: the records, from the raw values. Pair t = tid, tid + blockDim.x, ... (at
// most NP of them) is axis t / kStageCH of slot t % kStageCH, so that the axis is the same across a wavefront
// (kStageCH is a multiple of 64) and each wavefront runs one axis's code.
template <int D, class Contrib, int NP, class TF>
__device__ __forceinline__ void stage_write(const Contrib& cf, uint32_t m, const uint32_t* raw, TF* rec) {
    static_assert(kStageCH % 64 == 0, "a wavefront's pairs must share their axis");
#pragma unroll
    for (int it = 0; it < NP; ++it) {
        const uint32_t t = threadIdx.x + (uint32_t)it * blockDim.x;
        const uint32_t d = __builtin_amdgcn_readfirstlane(t / (uint32_t)kStageCH), s = t % (uint32_t)kStageCH;
        if (d >= (uint32_t)D) break;
        if (s >= m) continue;
        typename Contrib::Pre q;
        // the axis is a template argument: indexing the particle arrays by a run-time axis would place the
        // kernel's arrays of pointers in private memory
        if (d == 0) { cf.template load_axis<0>(raw, (int)s, q); cf.template stage_axis<0>(q, rec, (int)s); }
        else if (d == 1) { cf.template load_axis<1>(raw, (int)s, q); cf.template stage_axis<1>(q, rec, (int)s); }
        else { cf.template load_axis<2>(raw, (int)s, q); cf.template stage_axis<2>(q, rec, (int)s); }
    }
}

// Phases 1 and 2 over the block's ordered particles, chunk by chunk. All wavefronts see the same 64 cells, hence
// the same chunks and barriers. The raw values of chunk k + 1 load into LDS during chunk k's accumulation.
template <int D, int SPL, int NV, class TF, class Contrib, class LT>
__device__ __forceinline__ void staged_accumulate(LT& L, const Contrib& cf, uint32_t start,
                                                  TF (&acc)[1][StagedShape<D, SPL, NV>::NSH][NV]) {
    using SS = StagedShape<D, SPL, NV>;
    constexpr int NP = (kStageCH * D + SS::TB - 1) / SS::TB;  // pairs per thread and chunk
    __shared__ uint32_t raw[Contrib::NRAW * kStageCH];
    lds_u32* const lraw = (lds_u32*)raw;  // the same array as an LDS-space pointer (the load-to-LDS destination)
    const int lane = threadIdx.x & 63, wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const int a = wave % SS::NG, half = wave / SS::NG;
    const uint32_t n_c = L.cstart[lane + 1] - L.cstart[lane];
    const uint32_t n_max = wave_max_u32(n_c);
    const unsigned long long below = (1ull << lane) - 1ull;
    uint32_t j = 0, kb = 0, kb1 = 0;
    uint32_t j1 = stage_chunk_end(0, 0, n_max, n_c, kb1);
    stage_prefetch<SS::WAVES>(cf, lraw, L.p, start, 0u, kb1);
    wait_vmem();
    __syncthreads();  // the first chunk's raw values are in LDS
    while (j < n_max) {
        stage_write<D, Contrib, NP>(cf, kb1 - kb, raw, L.u.rec);
        __syncthreads();  // the chunk's records are complete and its raw values read
        uint32_t kb2 = kb1;
        const uint32_t j2 = stage_chunk_end(j1, kb1, n_max, n_c, kb2);
        if (j1 < n_max) stage_prefetch<SS::WAVES>(cf, lraw, L.p, start, kb1, kb2 - kb1);
        uint32_t k = kb;
        for (; j < j1; ++j) {
            const bool has = j < n_c;
            const unsigned long long mk = __ballot(has);
            const uint32_t slot = k + (uint32_t)__popcll(mk & below);
            k += (uint32_t)__popcll(mk);
            if (has && !cf.template accumulate_rec<SS>(L.u.rec, (int)(slot - kb), a, half, acc))
                atomicOr(&L.flag[slot >> 5], 1u << (slot & 31));
        }
        wait_vmem();
        __syncthreads();  // every wavefront is done with the records, and the next chunk's raw values are in LDS
        kb = kb1; kb1 = kb2; j1 = j2;
    }
}

// Adds particle `slot`'s direct form to the tile: thread t takes node t of its window.
template <int D, int SPL, int NV, class TF, class Contrib, class LT>
__device__ __forceinline__ void staged_direct(LT& L, const Contrib& cf, uint32_t p, uint32_t i, const int* bc, bool atomic) {
    using G = BG<D, SPL>;
    typename Contrib::Win z;
    cf.window_of(p, i, z);
    const int t = threadIdx.x;
    if (t >= G::W * G::W * G::W) return;
    const int a = t / (G::W * G::W), b = (t / G::W) % G::W, c = t % G::W;
    TF out[NV];
    cf.node_direct(z, a, b, c, out);
    const int q = tile_index<D, SPL>(z.wb[0] - G::E * bc[0] + a, z.wb[1] - G::E * bc[1] + b, z.wb[2] - G::E * bc[2] + c);
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        if (atomic) atomicAdd(&L.u.tile[v * G::NT + q], (double)out[v]);
        else L.u.tile[v * G::NT + q] += (double)out[v];
    }
}

// The block's tile by the staged scatter, to staging: ordered partial sums, then the direct form of the
// flagged particles one after another in slot order.
template <int D, int SPL, int NV, class TF, class Contrib, class LT>
__device__ __forceinline__ void staged_scatter_block(LT& L, const Contrib& cf, const StagingT<TF>& stg, uint32_t b, uint32_t start,
                                                     uint32_t cnt, const int* bc) {
    using SS = StagedShape<D, SPL, NV>;
    for (uint32_t w = threadIdx.x; w < (uint32_t)(kMaxBP / 32); w += blockDim.x) L.flag[w] = 0u;
    TF acc[1][SS::NSH][NV];
#pragma unroll
    for (int o = 0; o < SS::NSH; ++o)
#pragma unroll
        for (int v = 0; v < NV; ++v) acc[0][o][v] = TF(0);
    staged_accumulate<D, SPL, NV, TF>(L, cf, start, acc);
    __syncthreads();  // the records are read; the tile takes their place
    ordered_tile<D, SPL, NV, TF, SS>(L, acc);
    for (uint32_t w = 0; w < (cnt + 31) / 32; ++w) {
        uint32_t bits = L.flag[w];
        while (bits) {
            const uint32_t slot = w * 32 + (uint32_t)(__ffs(bits) - 1);
            bits &= bits - 1u;
            staged_direct<D, SPL, NV, TF>(L, cf, L.p[slot], start + slot, bc, false);
            __syncthreads();
        }
    }
    write_tile<D, SPL, NV, TF>(L, stg, b);
}

// A block with more than kMaxBP particles: every particle adds its direct form with LDS atomics (unordered).
template <int D, int SPL, int NV, class TF, class Contrib, class LT>
__device__ __forceinline__ uint32_t staged_block_atomic(LT& L, const Contrib& cf, const StagingT<TF>& stg, const SortBufs& sb, uint32_t b,
                                                        uint32_t start, uint32_t cnt, const int* bc) {
    using G = BG<D, SPL>;
    constexpr int NN = G::W * G::W * G::W;
    __syncthreads();
    zero_tile<D, SPL, NV, TF>(L);
    __syncthreads();
    uint32_t err = 0;
    for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
        typename Contrib::Win z;
        cf.window_of(sb.perm[start + k], start + k, z);
        int l[3];
        bool in = true;
        for (int d = 0; d < D; ++d) { l[d] = z.wb[d] - G::E * bc[d]; in = in && l[d] >= 0 && l[d] < G::E; }
        if (!in) { err |= MPM_ERR_GRID_BOUNDS; continue; }
        for (int t = 0; t < NN; ++t) {
            const int a = t / (G::W * G::W), bb = (t / G::W) % G::W, c = t % G::W;
            TF out[NV];
            cf.node_direct(z, a, bb, c, out);
            const int q = tile_index<D, SPL>(l[0] + a, l[1] + bb, l[2] + c);
            for (int v = 0; v < NV; ++v) atomicAdd(&L.u.tile[v * G::NT + q], (double)out[v]);
        }
    }
    __syncthreads();
    write_tile<D, SPL, NV, TF>(L, stg, b);
    return err;
}

// ----------------------------------------------------------------------------------------------
// k_p2g: P2G (p2g.cpp:12-66) fused with the stress force of explicitEulerUpdate
// (explicit_euler_update.cpp:15-168). The reference adds the force only where m_i > 0; nodes with
// m_i = 0 are never updated or read, so the force is scattered unconditionally.
// ----------------------------------------------------------------------------------------------
// PBC: the variant for periodic contexts, which also scatters the block's periodic copies.
// The staged form of the block's P2G (3D split form, no periodic copies).
template <int D, int SPL, bool SEP, bool MIBF, class TF>
__device__ __forceinline__ void p2g_staged_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                                const SortBufs& sb, const GridPtrsT<TF>& g, const StagingT<TF>& stg, uint32_t b) {
    using Cf = P2GContrib<D, SPL, SEP, MIBF, TF>;
    constexpr int NV = Cf::NV;
    __shared__ StagedLDS<D, SPL, NV, TF, Cf::SRec::N> L;
    if (S->stop) return;
    if (b >= S->nblocks) return;
    const uint32_t cnt = sb.count[b];
    if (cnt == 0) return;
    const uint32_t start = sb.offset[b];
    int bc[3];
    block_coords<D>(b, S, bc);
    Cf cf;
    cf.src = src;
    for (int d = 0; d < 3; ++d) { cf.org[d] = S->origin[d]; cf.nn[d] = S->n[d]; }
    L.coords.load(g.coord, bc, cf.nn);
    cf.cs = L.coords.view(bc);
    __syncthreads();
    cf.h1 = P.one_over_dx; cf.aD = P.apicDinverse; cf.apic = P.flip_ratio < 0;
    cf.kf = S->dt * P.particle_volume / P.particle_mass;
    cf.Lx = P.L[0]; cf.band = 4.0 * P.dx;
    uint32_t err = 0;
    if (cnt > kMaxBP) {
        err |= staged_block_atomic<D, SPL, NV, TF>(L, cf, stg, sb, b, start, cnt, bc);
    } else {
        err |= order_block<D, SPL, NV, TF>(L, P, S, src, dst, sb, g, b, start, cnt, bc);
        staged_scatter_block<D, SPL, NV, TF>(L, cf, stg, b, start, cnt, bc);
    }
    if (S->check_conservation) {  // sum of m_p v_p (simulation.cpp:292-317)
        double pm[3] = {0, 0, 0};
        for (uint32_t k = threadIdx.x; k < cnt; k += blockDim.x) {
            const uint32_t p = sb.perm[start + k];
            for (int d = 0; d < D; ++d) pm[d] += P.particle_mass * src.v[d][p];
        }
        for (int d = 0; d < D; ++d) pm[d] = wave_sum_f64(pm[d]);
        if ((threadIdx.x & 63) == 0) for (int d = 0; d < D; ++d) atomic_add_f64(&S->cons_part_mom[d], pm[d]);
    }
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0 && err) atomicOr(&S->err, err);
}

template <int D, int SPL, bool SEP, bool MIBF, class TF, bool PBC = false, bool NEXT = false>
__device__ __forceinline__ void p2g_plain_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                               const SortBufs& sb, const GridPtrsT<TF>& g, const StagingT<TF>& stg, uint32_t b);

// The single-scene quadratic split P2G with folded force and no MIBF (one wavefront per block at 229 VGPRs, 2 per
// SIMD) loads the next rank's particle while it adds this one: 255 VGPRs, still 2 per SIMD, C2 k_p2g 225 -> 215
// us. Elsewhere it would cost occupancy: the batched P2G and the separate-force and plain forms reach 256 VGPRs
// and 1 wavefront per SIMD, and the MUSL re-scatter dropped from 6 to 4 and ran 153 -> 171 us.
template <int D, int SPL, bool SEP, bool MIBF, class TF>
constexpr bool kP2GNextRank = D == 3 && SPL == 2 && sizeof(TF) == 4 && !SEP && !MIBF;

template <int D, int SPL, bool SEP, bool MIBF, class TF, bool PBC = false, bool NEXT = false>
__device__ __forceinline__ void p2g_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                         const SortBufs& sb, const GridPtrsT<TF>& g, const StagingT<TF>& stg, uint32_t b) {
    if constexpr (kStagedScatter<D, SPL, TF, PBC>) p2g_staged_body<D, SPL, SEP, MIBF, TF>(P, S, src, dst, sb, g, stg, b);
    else p2g_plain_body<D, SPL, SEP, MIBF, TF, PBC, NEXT>(P, S, src, dst, sb, g, stg, b);
}

template <int D, int SPL, bool SEP, bool MIBF, class TF, bool PBC, bool NEXT>
__device__ __forceinline__ void p2g_plain_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                               const SortBufs& sb, const GridPtrsT<TF>& g, const StagingT<TF>& stg, uint32_t b) {
    constexpr int NV = P2GContrib<D, SPL, SEP, MIBF, TF>::NV;
    using SS = ScatterShape<D, SPL, NV, TF>;
    __shared__ ScatterLDS<D, SPL, NV, TF> L;
    if (S->stop) return;
    const uint32_t nbk = S->nblocks;
    if (b >= nbk) return;
    const uint32_t cnt = sb.count[b];
    // periodic copies binned into this block (added to the tile after the block's particles; PBCAddParticles
    // appends them after the particles too, pbc.cpp:37-83)
    const uint32_t gcnt = PBC ? sb.count[nbk + b] : 0u;
    if (cnt == 0 && gcnt == 0) return;
    const uint32_t start = sb.offset[b];
    const uint32_t gstart = gcnt ? sb.offset[nbk + b] : 0u;
    int bc[3];
    block_coords<D>(b, S, bc);
    P2GContrib<D, SPL, SEP, MIBF, TF> cf;
    cf.src = src;
    for (int d = 0; d < 3; ++d) { cf.org[d] = S->origin[d]; cf.nn[d] = S->n[d]; }
    // the block's coordinates are loaded with the ordering's loads and put into LDS after it
    CoordPrefetch<D, SPL> cp;
    cp.issue(g.coord, bc, cf.nn);
    cf.cs = L.coords.view(bc);
    cf.h1 = P.one_over_dx; cf.aD = P.apicDinverse; cf.apic = P.flip_ratio < 0;
    cf.kf = S->dt * P.particle_volume / P.particle_mass;
    cf.Lx = P.L[0]; cf.band = 4.0 * P.dx;
    uint32_t err = 0;
    if (cnt > kMaxBP || gcnt > kMaxBP) {
        if (cnt && cnt <= kMaxBP) err |= order_block<D, SPL, NV, TF>(L, P, S, src, dst, sb, g, b, start, cnt, bc);
        cp.put(L.coords);
        err |= scatter_block_atomic<D, SPL, NV, TF>(L, cf, stg, sb, b, start, cnt, bc, gstart, gcnt);
    } else if (gcnt == 0) {
        err |= order_block<D, SPL, NV, TF>(L, P, S, src, dst, sb, g, b, start, cnt, bc);
        cp.put(L.coords);
        __syncthreads();
        scatter_block<D, SPL, NV, TF, P2GContrib<D, SPL, SEP, MIBF, TF>, NEXT>(L, cf, stg, b, start);
    } else {
        cp.put(L.coords);
        __syncthreads();
        TF acc[SS::SPP][SS::NSH][NV];
        zero_acc<D, SPL, NV, TF>(acc);
        if (cnt) {
            err |= order_block<D, SPL, NV, TF>(L, P, S, src, dst, sb, g, b, start, cnt, bc);
            accumulate_cells<D, SPL, NV, TF>(L, cf, start, acc);
            __syncthreads();
        }
        err |= order_block<D, SPL, NV, TF, true>(L, P, S, src, dst, sb, g, b, gstart, gcnt, bc);
        accumulate_cells<D, SPL, NV, TF>(L, cf, gstart, acc);
        finish_tile<D, SPL, NV, TF>(L, stg, b, acc);
    }
    if (S->check_conservation) {  // sum of m_p v_p (simulation.cpp:292-317), periodic copies included
        double pm[3] = {0, 0, 0};
        for (uint32_t k = threadIdx.x; k < cnt + gcnt; k += blockDim.x) {
            const uint32_t p = sb.perm[k < cnt ? start + k : gstart + (k - cnt)] & ~kGhostBit;
            for (int d = 0; d < D; ++d) pm[d] += P.particle_mass * src.v[d][p];
        }
        for (int d = 0; d < D; ++d) pm[d] = wave_sum_f64(pm[d]);
        if ((threadIdx.x & 63) == 0) for (int d = 0; d < D; ++d) atomic_add_f64(&S->cons_part_mom[d], pm[d]);
    }
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0 && err) atomicOr(&S->err, err);
}
template <int D, int SPL, bool SEP, bool MIBF, class TF, bool PBC>
__global__ __launch_bounds__((ScatterLaunch<D, SPL, P2GContrib<D, SPL, SEP, MIBF, TF>::NV, TF, PBC>::TB))
__attribute__((amdgpu_waves_per_eu(ScatterLaunch<D, SPL, P2GContrib<D, SPL, SEP, MIBF, TF>::NV, TF, PBC>::WPE)))
void k_p2g(Params P, StepState* S, PartStateT<TF> src, PartStateT<TF> dst, SortBufs sb, GridPtrsT<TF> g, StagingT<TF> stg) {
    p2g_body<D, SPL, SEP, MIBF, TF, PBC, kP2GNextRank<D, SPL, SEP, MIBF, TF>>(P, S, src, dst, sb, g, stg, blockIdx.x);
}

// k_musl_scatter: the MUSL re-scatter, in the particle order that k_p2g fixed for the step.
template <int D, int SPL, class TF>
__device__ __forceinline__ void musl_staged_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                                 const SortBufs& sb, const GridPtrsT<TF>& g, const StagingT<TF>& stg, uint32_t b) {
    using G = BG<D, SPL>;
    using Cf = MuslContrib<D, SPL, TF>;
    __shared__ StagedLDS<D, SPL, D, TF, Cf::SRec::N> L;
    if (S->stop) return;
    if (b >= S->nblocks) return;
    const uint32_t cnt = sb.count[b];
    if (cnt == 0) return;
    const uint32_t start = sb.offset[b];
    int bc[3];
    block_coords<D>(b, S, bc);
    Cf cf;
    cf.src = src; cf.dst = dst;
    for (int d = 0; d < 3; ++d) { cf.org[d] = S->origin[d]; cf.nn[d] = S->n[d]; }
    L.coords.load(g.coord, bc, cf.nn);
    cf.cs = L.coords.view(bc);
    cf.h1 = P.one_over_dx; cf.aD = P.apicDinverse; cf.apic = P.flip_ratio < 0;
    uint32_t err = 0;
    if (cnt > kMaxBP) {
        err |= staged_block_atomic<D, SPL, D, TF>(L, cf, stg, sb, b, start, cnt, bc);
    } else {
        copy_to_lds(L.p, sb.perm + start, cnt);
        copy_to_lds(L.cstart, sb.cell_off + (size_t)b * (G::NC + 1), (uint32_t)G::NC + 1);
        if (threadIdx.x == 0) L.cstart[G::NC + 1] = cnt;
        __syncthreads();
        staged_scatter_block<D, SPL, D, TF>(L, cf, stg, b, start, cnt, bc);
    }
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0 && err) atomicOr(&S->err, err);
}

template <int D, int SPL, class TF>
__device__ __forceinline__ void musl_plain_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                                const SortBufs& sb, const GridPtrsT<TF>& g, const StagingT<TF>& stg, uint32_t b);

template <int D, int SPL, class TF>
__device__ __forceinline__ void musl_scatter_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                                  const SortBufs& sb, const GridPtrsT<TF>& g, const StagingT<TF>& stg, uint32_t b) {
    if constexpr (kStagedScatter<D, SPL, TF, false>) musl_staged_body<D, SPL, TF>(P, S, src, dst, sb, g, stg, b);
    else musl_plain_body<D, SPL, TF>(P, S, src, dst, sb, g, stg, b);
}

template <int D, int SPL, class TF>
__device__ __forceinline__ void musl_plain_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                                const SortBufs& sb, const GridPtrsT<TF>& g, const StagingT<TF>& stg, uint32_t b) {
    using G = BG<D, SPL>;
    __shared__ ScatterLDS<D, SPL, D, TF> L;
    if (S->stop) return;
    if (b >= S->nblocks) return;
    const uint32_t cnt = sb.count[b];
    if (cnt == 0) return;
    const uint32_t start = sb.offset[b];
    int bc[3];
    block_coords<D>(b, S, bc);
    MuslContrib<D, SPL, TF> cf;
    cf.src = src; cf.dst = dst;
    for (int d = 0; d < 3; ++d) { cf.org[d] = S->origin[d]; cf.nn[d] = S->n[d]; }
    L.coords.load(g.coord, bc, cf.nn);
    cf.cs = L.coords.view(bc);
    cf.h1 = P.one_over_dx; cf.aD = P.apicDinverse; cf.apic = P.flip_ratio < 0;
    uint32_t err = 0;
    if (cnt > kMaxBP) {
        err |= scatter_block_atomic<D, SPL, D, TF>(L, cf, stg, sb, b, start, cnt, bc);
    } else {
        copy_to_lds(L.p, sb.perm + start, cnt);
        copy_to_lds(L.cstart, sb.cell_off + (size_t)b * (G::NC + 1), (uint32_t)G::NC + 1);
        if (threadIdx.x == 0) L.cstart[G::NC + 1] = cnt;
        __syncthreads();
        scatter_block<D, SPL, D, TF>(L, cf, stg, b, start);
    }
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0 && err) atomicOr(&S->err, err);
}
template <int D, int SPL, class TF>
__global__ __launch_bounds__((ScatterLaunch<D, SPL, D, TF, false>::TB))
void k_musl_scatter(Params P, StepState* S, PartStateT<TF> src, PartStateT<TF> dst, SortBufs sb, GridPtrsT<TF> g, StagingT<TF> stg) {
    musl_scatter_body<D, SPL, TF>(P, S, src, dst, sb, g, stg, blockIdx.x);
}

// Sum over the blocks that cover node gx (at most two per axis: a tile spans E + W - 1 < 2 E nodes), in increasing
// block order, of the staged tile values. The 2^D candidate blocks' counts and tile values are loaded together,
// one memory latency per node instead of two per covering block; a candidate outside the grid's blocks or
// without particles adds nothing (its staging holds stale values, read but not used).
// with_copies: the tiles of the P2G scatter, which also exist for blocks that hold only periodic copies
// (the MUSL re-scatter has none).
// BlockGrid: the step's blocks per axis and block count, read from StepState once per workgroup (inside a loop
// that stores to global memory the compiler would reload them for every node).
struct BlockGrid {
    uint32_t nb[3], nblocks;
    __device__ __forceinline__ explicit BlockGrid(const StepState* S) : nb{S->nb[0], S->nb[1], S->nb[2]}, nblocks(S->nblocks) {}
};
template <int D, int SPL, int NV, class TF>
__device__ __forceinline__ void sum_staging(const BlockGrid& bgrid, const SortBufs& sb, const StagingT<TF>& stg, const int* gx, double (&s)[NV],
                                            bool with_copies = false) {
    using G = BG<D, SPL>;
    static_assert(G::T <= 2 * G::E, "a node lies in the tiles of at most two blocks per axis");
    constexpr int NCB = 1 << D;
    const uint32_t nb[3] = {bgrid.nb[0], bgrid.nb[1], D == 3 ? bgrid.nb[2] : 1u};
    const uint32_t nblocks = bgrid.nblocks;
    int lo[3] = {0, 0, 0}, hi[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
        const int l = gx[d] - (G::T - 1);
        lo[d] = l <= 0 ? 0 : (l + G::E - 1) / G::E;
        hi[d] = min(gx[d] / G::E, (int)nb[d] - 1);
    }
    bool ok[NCB];
    uint32_t cnt[NCB];
    TF val[NCB][NV];
#pragma unroll
    for (int c = 0; c < NCB; ++c) {  // candidate c: block (lo0 + c_0, lo1 + c_1, lo2 + c_2), last axis fastest
        int b[3] = {0, 0, 0};
#pragma unroll
        for (int d = 0; d < D; ++d) b[d] = lo[d] + ((c >> (D - 1 - d)) & 1);
        bool in = true;
#pragma unroll
        for (int d = 0; d < D; ++d) in = in && b[d] <= hi[d];
        ok[c] = in;
        const uint32_t bl = D == 3 ? ((uint32_t)b[0] * nb[1] + (uint32_t)b[1]) * nb[2] + (uint32_t)b[2] : (uint32_t)b[0] * nb[1] + (uint32_t)b[1];
        cnt[c] = in ? sb.count[bl] + (with_copies ? sb.count[nblocks + bl] : 0u) : 0u;
        const int q = in ? tile_index<D, SPL>(gx[0] - G::E * b[0], gx[1] - G::E * b[1], D == 3 ? gx[2] - G::E * b[2] : 0) : 0;
        const size_t at = (size_t)(in ? bl : 0u) * G::NT + q;
#pragma unroll
        for (int v = 0; v < NV; ++v) val[c][v] = stg.base[(size_t)v * stg.stride + at];
    }
#pragma unroll
    for (int v = 0; v < NV; ++v) s[v] = 0.0;
#pragma unroll
    for (int c = 0; c < NCB; ++c)
        if (ok[c] && cnt[c] != 0)
#pragma unroll
            for (int v = 0; v < NV; ++v) s[v] += (double)val[c][v];
}

// ----------------------------------------------------------------------------------------------
// k_grid_update: node sums, normalisation (p2g.cpp:84-98), explicit update and boundary conditions
// (explicit_euler_update.cpp:183-232, boundary_collision.cpp:5-394), flip = v_new - v_old.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL, bool SEP, bool MIBF, class TF>
__device__ __forceinline__ void grid_update_body(const Params& P, StepState* S, const SortBufs& sb, const GridPtrsT<TF>& g,
                                                 const StagingT<TF>& stg, uint32_t bid, uint32_t nbid) {
    constexpr int NV = P2GContrib<D, SPL, SEP, MIBF, TF>::NV;
    if (S->stop) return;
    const uint32_t nn = S->grid_nodes;
    const uint32_t n[3] = {S->n[0], S->n[1], S->n[2]};
    const double dt = S->dt;
    const double dtV = dt * P.particle_volume;
    double dtg[3]; for (int d = 0; d < 3; ++d) dtg[d] = dt * S->gravity[d];
    const bool cons = S->check_conservation;
    // the plates' state of this step, read once per workgroup (plate_move updates it in k_finish)
    __shared__ DevPlate s_pl[kMaxPlates];
    for (int k = threadIdx.x; k < P.n_plates; k += blockDim.x) s_pl[k] = S->plates[k];
    __syncthreads();
    double cm = 0.0, cmv[3] = {0, 0, 0};
    const BlockGrid bgrid(S);
    for (uint32_t i = bid * blockDim.x + threadIdx.x; i < nn; i += nbid * blockDim.x) {
        int gx[3] = {0, 0, 0};
        if (D == 3) { gx[2] = (int)(i % n[2]); const uint32_t r = i / n[2]; gx[1] = (int)(r % n[1]); gx[0] = (int)(r / n[1]); }
        else { gx[1] = (int)(i % n[1]); gx[0] = (int)(i / n[1]); }
        V<D> X;  // loaded with the sums
        for (int d = 0; d < D; ++d) X[d] = g.coord[d][gx[d]];
        double s[NV];
        sum_staging<D, SPL, NV, TF>(bgrid, sb, stg, gx, s, sb.ghosts != 0);
        const double sw = s[0];
        V<D> v;
        if (sw > 0) { for (int d = 0; d < D; ++d) v[d] = s[1 + d] / sw; }
        else { for (int d = 0; d < D; ++d) v[d] = 0.0; }
        const double fr = MIBF ? s[NV - 1] / sw : 0.0;
        const double mi = sw * P.particle_mass;
        if (cons) { cm += mi; for (int d = 0; d < D; ++d) cmv[d] += v[d] * mi; }
        V<D> vn, flip;
        if (mi > 0) {
            for (int d = 0; d < D; ++d) vn[d] = SEP ? v[d] + (((-dtV) * s[1 + D + d]) / mi + dtg[d]) : v[d] + dtg[d];
            boundary_collision<D>(P, s_pl, dt, fr, X, vn);
            for (int d = 0; d < D; ++d) flip[d] = vn[d] - v[d];
        } else {
            vn = v;
            for (int d = 0; d < D; ++d) flip[d] = 0.0;
        }
        g.mass[i] = (TF)mi;
        for (int d = 0; d < D; ++d) g.v[d][i] = (TF)vn[d];
        if (SEP) for (int d = 0; d < D; ++d) g.flip[d][i] = (TF)flip[d];
        if (MIBF) g.fric[i] = (TF)fr;
    }
    if (cons) {
        cm = wave_sum_f64(cm);
        for (int d = 0; d < D; ++d) cmv[d] = wave_sum_f64(cmv[d]);
        if ((threadIdx.x & 63) == 0) {
            atomic_add_f64(&S->cons_grid_mass, cm);
            for (int d = 0; d < D; ++d) atomic_add_f64(&S->cons_grid_mom[d], cmv[d]);
        }
    }
}
template <int D, int SPL, bool SEP, bool MIBF, class TF>
__global__ __launch_bounds__(kBlock) void k_grid_update(Params P, StepState* S, SortBufs sb, GridPtrsT<TF> g, StagingT<TF> stg) {
    grid_update_body<D, SPL, SEP, MIBF, TF>(P, S, sb, g, stg, blockIdx.x, gridDim.x);
}

// k_musl_update: MUSL normalisation, musl.cpp:182-188: v_i = (sum) / (m_i / m_p), zero where m_i == 0.
template <int D, int SPL, class TF>
__device__ __forceinline__ void musl_update_body(const Params& P, const StepState* S, const SortBufs& sb, const GridPtrsT<TF>& g,
                                                 const StagingT<TF>& stg, uint32_t bid, uint32_t nbid) {
    if (S->stop) return;
    const uint32_t nn = S->grid_nodes;
    const uint32_t n[3] = {S->n[0], S->n[1], S->n[2]};
    const BlockGrid bgrid(S);
    for (uint32_t i = bid * blockDim.x + threadIdx.x; i < nn; i += nbid * blockDim.x) {
        int gx[3] = {0, 0, 0};
        if (D == 3) { gx[2] = (int)(i % n[2]); const uint32_t r = i / n[2]; gx[1] = (int)(r % n[1]); gx[0] = (int)(r / n[1]); }
        else { gx[1] = (int)(i % n[1]); gx[0] = (int)(i / n[1]); }
        const double mi = g.mass[i];
        double s[D];
        sum_staging<D, SPL, D, TF>(bgrid, sb, stg, gx, s);  // loaded with the mass
        if (mi > 0) {
            const double q = mi / P.particle_mass;
            for (int d = 0; d < D; ++d) g.v[d][i] = (TF)(s[d] / q);
        } else {
            for (int d = 0; d < D; ++d) g.v[d][i] = TF(0);
        }
    }
}
template <int D, int SPL, class TF>
__global__ __launch_bounds__(kBlock) void k_musl_update(Params P, const StepState* S, SortBufs sb, GridPtrsT<TF> g, StagingT<TF> stg) {
    musl_update_body<D, SPL, TF>(P, S, sb, g, stg, blockIdx.x, gridDim.x);
}

// Velocity update of positionUpdate / MUSL (position_update.cpp:14-23, musl.cpp:13-25).
__device__ __forceinline__ double blend_velocity(double fr, double v, double flip, double pic) {
    if (fr < -1) return pic;
    if (fr < 0) return (-fr) * (v + flip) + (1 - (-fr)) * pic;
    return fr * (v + flip) + (1 - fr) * pic;
}

// Material of a deformation kernel, fixed at compile time: a plastic model (mpm_plastic_model) on Hencky
// elasticity, or kMatNeoHookean (no return map).
constexpr int kMatNeoHookean = 16;

// The state the deformation update reads of source particle p: F and the plastic scalars.
template <int D> struct DefgradIn { M<D> F; PlasticState ps; };
template <int D, class TF>
__device__ __forceinline__ void load_defgrad_in(const PartStateT<TF>& src, uint32_t p, DefgradIn<D>& in) {
    for (int k = 0; k < D * D; ++k) in.F.m[k] = src.F[k][p];
    in.ps.eps_dev = src.eps_dev[p]; in.ps.eps_vol = src.eps_vol[p]; in.ps.eps_prad = src.eps_prad[p];
    in.ps.dgamma = 0.0; in.ps.visc = src.visc[p]; in.ps.muI = src.muI[p];
}

// F <- F + (dt*grad v) * F, return map, Kirchhoff stress for the next step (deformation_update.cpp:13-46,
// plasticity.cpp:11-665), on the state `in` of a source particle; writes slot i of the destination layout.
template <int D, class TF, int MAT>
__device__ __forceinline__ void deform_and_project(const Params& P, double dt, const DefgradIn<D>& in,
                                                   const PartStateT<TF>& dst, uint32_t i, const double (&gv)[3][3],
                                                   uint32_t& err, uint32_t& pcount) {
    M<D> F = in.F, A;
    for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) A(r, c) = dt * gv[r][c];
    M<D> AF = mul<D>(A, F);
    for (int k = 0; k < D * D; ++k) F.m[k] = F.m[k] + AF.m[k];
    PlasticState ps = in.ps;
    double tau[6];
    // The split form decomposes b = F F^T (left_spectral: half the rotations of the two-sided SVD, results
    // within about 1e-12). The plain form keeps the two-sided SVD, whose rotations round like the reference's
    // JacobiSVD: its yield decisions must match the reference's one for one, including those of particles
    // whose strain is rounding noise (a uniform flow's grad v), which sit on a yield surface through the
    // origin (MCC with beta = 0, cohesionless Drucker-Prager) and fall either side of it with the last bit.
    constexpr bool kLeftSpectral = sizeof(TF) == 4;
    M<D> U; V<D> lam;
    if (MAT == kMatNeoHookean) {
        neohookean_tau<D>(F, P.mu, P.lambda, tau);
    } else if (kLeftSpectral && left_spectral<D>(F, U, lam, P.svd_tol)) {
        // singular values sqrt(lam), all within a factor 100 of each other and above the 1e-4 clamp
        V<D> ls;
        for (int k = 0; k < D; ++k) ls[k] = 0.5 * log_d_split(lam[k]);
        if (MAT != MPM_PLASTIC_NONE) {
            V<D> h = ls;
            RMResult r = return_map<D, MAT, false>(P, dt, h, ps);
            err |= r.err;
            pcount += r.plastic_incr;
            if (r.recompose) {  // exp(h) / sqrt(lam) = exp(h - ln sqrt(lam)): one exponential, no square root
                V<D> e; for (int k = 0; k < D; ++k) e[k] = exp(h[k] - ls[k]);
                F = recompose_left<D>(U, e, F);
                ls = h;
            }
        }
        hencky_tau<D>(U, ls, P.mu, P.lambda, tau);
    } else {
        M<D> W; V<D> s;
        svd<D>(F, U, s, W, P.svd_tol);
        V<D> ls;
        if (MAT != MPM_PLASTIC_NONE) {
            V<D> h; for (int k = 0; k < D; ++k) h[k] = log_d(fmax(fabs(s[k]), 1e-4));
            const V<D> h0 = h;
            RMResult r = return_map<D, MAT, !kLeftSpectral>(P, dt, h, ps);
            err |= r.err;
            pcount += r.plastic_incr;
            if (r.recompose) {  // the new F = U exp(h) W^T has singular values exp(h): its Hencky strain is h
                V<D> e; for (int k = 0; k < D; ++k) e[k] = exp(h[k]);
                F = udvt<D>(U, e, W);
                for (int k = 0; k < D; ++k) ls[k] = h[k];
            } else {
                // ln|s| is the h computed above unless a singular value is below the clamp 1e-4
                bool small = false;
                for (int k = 0; k < D; ++k) small = small || !(fabs(s[k]) >= 1e-4);
                if (small) { for (int k = 0; k < D; ++k) ls[k] = log_d(fabs(s[k])); }
                else { for (int k = 0; k < D; ++k) ls[k] = h0[k]; }
            }
        } else {
            for (int k = 0; k < D; ++k) ls[k] = log_d(fabs(s[k]));
        }
        hencky_tau<D>(U, ls, P.mu, P.lambda, tau);
    }
    for (int k = 0; k < D * D; ++k) dst.F[k][i] = F.m[k];
    for (int k = 0; k < SymPack<D>::n; ++k) dst.tau[k][i] = (TF)tau[k];
    dst.eps_dev[i] = ps.eps_dev; dst.eps_vol[i] = ps.eps_vol; dst.eps_prad[i] = ps.eps_prad;
    dst.dgamma[i] = ps.dgamma; dst.visc[i] = ps.visc; dst.muI[i] = ps.muI;
}

// Gather of G2P (g2p.cpp:35-67) and of the velocity gradient (deformation_update.cpp:217-230) over the
// non-zero nodes of the window, in the reference's (i,j,k) order.

// ----------------------------------------------------------------------------------------------
// Gather kernels: one workgroup per block; the block's tile of node values is copied to LDS once and
// every particle of the block gathers from it.
// ----------------------------------------------------------------------------------------------

// Node values in an LDS tile, in one of two layouts: component-major ([component][node]; k_g2p) or
// node-major (NM: the D components of a node side by side, padded to TileStride<D>, 4 in 3D and 2 in 2D;
// k_defgrad), where a gather reads a node with one wide load. Measured: node-major took k_defgrad's plain
// form from 390 to 346 us on C2, but made k_g2p slower (253 -> 283 us), so each kernel keeps its own.
template <int D> struct TileStride { static constexpr int v = D == 3 ? 4 : 2; };

// Copies the D node arrays src[0..D) over the block's tile into LDS; nodes outside the grid read as zero
// (no particle window reaches them).
// A thread issues the loads of all its nodes (at most kTileLoads) before it stores any, so that the tile costs
// one memory latency rather than one per node.
constexpr int kTileLoads = 4;
template <int D, int SPL, class TF, bool NM = false>
__device__ __forceinline__ void load_tile(TF* s, TF* const* src, const int* bc, const uint32_t* n) {
    using G = BG<D, SPL>;
    constexpr int ST = TileStride<D>::v;
    for (int q0 = threadIdx.x; q0 < G::NT; q0 += kTileLoads * blockDim.x) {
        TF val[kTileLoads][3];
#pragma unroll
        for (int u = 0; u < kTileLoads; ++u) {
            const int q = q0 + u * (int)blockDim.x;
            int t[3] = {0, 0, 0};
            if (D == 3) { t[2] = q % G::T; const int r = q / G::T; t[1] = r % G::T; t[0] = r / G::T; }
            else { t[1] = q % G::T; t[0] = q / G::T; }
            bool in = q < G::NT;
            int gx[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) { gx[d] = G::E * bc[d] + t[d]; in = in && gx[d] < (int)n[d]; }
            const uint64_t node = in ? node_offset<D>(gx[0], gx[1], gx[2], n) : 0ull;
#pragma unroll
            for (int c = 0; c < D; ++c) val[u][c] = in ? src[c][node] : TF(0);
        }
#pragma unroll
        for (int u = 0; u < kTileLoads; ++u) {
            const int q = q0 + u * (int)blockDim.x;
            if (q >= G::NT) break;
            if constexpr (NM) {
                for (int c = 0; c < D; ++c) s[q * ST + c] = val[u][c];
                for (int c = D; c < ST; ++c) s[q * ST + c] = TF(0);
            } else {
                using GT = GatherTile<D, SPL>;
                int qg = q;
                if constexpr (GT::PAD) { const int t2 = q % G::T, r = q / G::T; qg = GT::index(r / G::T, r % G::T, t2); }
                for (int c = 0; c < D; ++c) s[c * GT::N + qg] = val[u][c];
            }
        }
    }
}

// Full window of a gathering particle (G2P does not clamp, g2p.cpp:29-32) and its tile-local base.
template <int D, int SPL, class CS>
__device__ __forceinline__ bool gather_stencil(Stencil<D, SPL>& st, int* tl, const double* xp, const double* org, const uint32_t* n,
                                               const CS& cs, double h1, const int* bc) {
    using G = BG<D, SPL>;
    const bool ok = make_stencil<D, SPL>(st, xp, org, n, cs, h1, false);
    bool in = true;
    tl[2] = 0;
    for (int d = 0; d < D; ++d) { tl[d] = st.base[d] - G::E * bc[d]; in = in && tl[d] >= 0 && tl[d] < G::E; }
    return ok && in;
}

// Node values for gathers: from the LDS tile of a block (tile-local window base t0..t2) or from the grid.
template <int D, int SPL, class TF, bool NM = false> struct TileFetch {
    const TF* v; const TF* f;
    int t0, t1, t2;
    __device__ __forceinline__ void get(int a, int b, int c, TF* vi) const {
        const int q = NM ? tile_index<D, SPL>(t0 + a, t1 + b, t2 + c) : GatherTile<D, SPL>::index(t0 + a, t1 + b, t2 + c);
        if constexpr (!NM) {
            for (int d = 0; d < D; ++d) vi[d] = v[d * GatherTile<D, SPL>::N + q];
        } else if constexpr (D == 3 && sizeof(TF) == 4) {
            const float4 w = *reinterpret_cast<const float4*>(v + q * 4);
            vi[0] = w.x; vi[1] = w.y; vi[2] = w.z;
        } else if constexpr (D == 2 && sizeof(TF) == 4) {
            const float2 w = *reinterpret_cast<const float2*>(v + q * 2);
            vi[0] = w.x; vi[1] = w.y;
        } else {
            for (int d = 0; d < D; ++d) vi[d] = v[q * TileStride<D>::v + d];
        }
    }
    __device__ __forceinline__ TF flip(int a, int b, int c, int d) const {
        if constexpr (NM) return f[tile_index<D, SPL>(t0 + a, t1 + b, t2 + c) * TileStride<D>::v + d];
        else return f[d * GatherTile<D, SPL>::N + GatherTile<D, SPL>::index(t0 + a, t1 + b, t2 + c)];
    }
};
template <int D, class TF> struct GridFetch {
    const GridPtrsT<TF>& g;
    const uint32_t* n;
    int b0, b1, b2;
    __device__ __forceinline__ void get(int a, int b, int c, TF* vi) const {
        const uint32_t node = node_index<D>(b0 + a, b1 + b, D == 3 ? b2 + c : 0, n);
        for (int d = 0; d < D; ++d) vi[d] = g.v[d][node];
    }
    __device__ __forceinline__ TF flip(int a, int b, int c, int d) const {
        return g.flip[d][node_index<D>(b0 + a, b1 + b, D == 3 ? b2 + c : 0, n)];
    }
};

// Gather of G2P (g2p.cpp:35-67) or of the velocity gradient (deformation_update.cpp:217-230) over the
// non-zero nodes of the window, factored by rows. The weight of node (a, b, c) is w0[a] w1[b] w2[c] and
// its gradient h1 (dw0[a] w1[b] w2[c], w0[a] dw1[b] w2[c], w0[a] w1[b] dw2[c]). Along each row (a, b) of
// the window the node values are summed first with the last axis's factors, S = sum_c w2[c] v and
// Sd = sum_c f[c] v (f = w2 (x_2c - x_p,2) for B, dw2 for grad v), and the row then adds
//   pic += w0 w1 S,  B[:,0] += w0 (x_0a - x_p,0) w1 S,  B[:,1] += w0 w1 (x_1b - x_p,1) S,  B[:,2] += w0 w1 Sd,
//   grad v[:,0] += h1 dw0 w1 S,  grad v[:,1] += h1 w0 dw1 S,  grad v[:,2] += h1 w0 w1 Sd.
// The sums are those of the reference's (i,j,k) loops, grouped by row; they agree with the loop order to
// rounding. About 9 multiply-adds per node instead of about 30.
// FLIP (compile time, so that kernels without flip carry no flip loads) adds the flip sum.
// Used by the split form: the plain form keeps the reference's node order (gather_nodes), so that its
// sums round as the reference's do and no particle at a yield threshold takes the other branch.
// AP: APIC known at compile time (0 or 1) or read from apic (-1).
template <int D, int SPL, bool PIC, bool GRAD, bool FLIP, class TF, class Fetch, int AP = -1>
__device__ __forceinline__ void gather_rows(const Stencil<D, SPL>& st, const double* xp, double h1, const Fetch& fetch, bool apic_rt,
                                            TF* pic, TF* fl, TF (&Bp)[3][3], TF (&gv)[3][3]) {
    const bool apic = AP < 0 ? apic_rt : AP == 1;
    static_assert(PIC != GRAD, "one gather per kind");
    constexpr int C = Stencil<D, SPL>::cnt;
    constexpr int C2 = D == 3 ? C : 1;
    TF w[3][C], f[3][C];  // weights; per axis the second factor: w (x_i - x_p) for B, h1 dw for grad v
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < C; ++k) {
            w[d][k] = d < D ? (TF)st.w[d][k] : TF(1);
            if (d >= D) f[d][k] = TF(0);
            else if (PIC) f[d][k] = w[d][k] * (TF)(st.xi[d][k] - xp[d]);
            else f[d][k] = (TF)(st.dw[d][k] * h1);
        }
    const bool second = GRAD || apic;
#pragma unroll
    for (int a = 0; a < C; ++a) {
#pragma unroll
        for (int b = 0; b < C; ++b) {
            TF S[3] = {0, 0, 0}, Sd[3] = {0, 0, 0}, Sf[3] = {0, 0, 0};
#pragma unroll
            for (int c = 0; c < C2; ++c) {
                TF vi[3];
                fetch.get(a, b, c, vi);
                if (D == 3) {
#pragma unroll
                    for (int r = 0; r < D; ++r) {
                        S[r] = madd(w[2][c], vi[r], S[r]);
                        if (second) Sd[r] = madd(f[2][c], vi[r], Sd[r]);
                        if (PIC && FLIP) Sf[r] = madd(w[2][c], fetch.flip(a, b, c, r), Sf[r]);
                    }
                } else {
#pragma unroll
                    for (int r = 0; r < D; ++r) {
                        S[r] = vi[r];
                        if (PIC && FLIP) Sf[r] = fetch.flip(a, b, c, r);
                    }
                }
            }
            const TF w01 = w[0][a] * w[1][b];
            // rows are scheduled one after another: hoisting every row's loads ahead of the arithmetic
            // would keep all window values live at once
            __builtin_amdgcn_sched_barrier(0);
            if (PIC) {
                const TF b0 = f[0][a] * w[1][b], b1 = w[0][a] * f[1][b];
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    pic[r] = madd(w01, S[r], pic[r]);
                    if (FLIP) fl[r] = madd(w01, Sf[r], fl[r]);
                    if (apic) {
                        Bp[r][0] = madd(b0, S[r], Bp[r][0]);
                        Bp[r][1] = madd(b1, S[r], Bp[r][1]);
                        if (D == 3) Bp[r][2] = madd(w01, Sd[r], Bp[r][2]);
                    }
                }
            } else {
                const TF g0 = f[0][a] * w[1][b], g1 = w[0][a] * f[1][b];
#pragma unroll
                for (int r = 0; r < D; ++r) {
                    gv[r][0] = madd(g0, S[r], gv[r][0]);
                    gv[r][1] = madd(g1, S[r], gv[r][1]);
                    if (D == 3) gv[r][2] = madd(w01, Sd[r], gv[r][2]);
                }
            }
        }
    }
}

// The same gather node by node in the reference's (i,j,k) loop order, with its products and sums
// (g2p.cpp:35-67, deformation_update.cpp:27-40): the plain form's.
// The flip sum is switched by a run-time flag here: with a compile-time one the compiler keeps every node's
// flip values live at once in binary64 (the kernel then runs at one wavefront per SIMD).
template <int D, int SPL, bool PIC, bool GRAD, class TF, class Fetch>
__device__ __forceinline__ void gather_nodes(const Stencil<D, SPL>& st, const double* xp, double h1, const Fetch& fetch, bool apic,
                                             bool useflip, TF* pic, TF* fl, TF (&Bp)[3][3], TF (&gv)[3][3]) {
    const int c2 = (D == 3) ? Stencil<D, SPL>::cnt : 1;
#pragma unroll
    for (int a = 0; a < Stencil<D, SPL>::cnt; ++a) {
#pragma unroll
        for (int b = 0; b < Stencil<D, SPL>::cnt; ++b) {
#pragma unroll
            for (int ic = 0; ic < c2; ++ic) {
                const int c = (D == 3) ? ic : 0;
                TF vi[3];
                fetch.get(a, b, c, vi);
                if (PIC) {
                    double weight = st.w[0][a] * st.w[1][b];
                    if (D == 3) weight = weight * st.w[2][c];
                    const TF w = (TF)weight;
                    const TF pd[3] = {(TF)(st.xi[0][a] - xp[0]), (TF)(st.xi[1][b] - xp[1]), D == 3 ? (TF)(st.xi[2][c] - xp[2]) : TF(0)};
                    for (int r = 0; r < D; ++r) {
                        pic[r] += vi[r] * w;
                        if (apic) for (int k = 0; k < D; ++k) Bp[r][k] += vi[r] * pd[k] * w;
                    }
                    if (useflip) for (int r = 0; r < D; ++r) fl[r] += fetch.flip(a, b, c, r) * w;
                }
                if (GRAD) {
                    double gr[3];
                    if (D == 3) {
                        gr[0] = st.dw[0][a] * st.w[1][b] * st.w[2][c] * h1;
                        gr[1] = st.dw[1][b] * st.w[0][a] * st.w[2][c] * h1;
                        gr[2] = st.dw[2][c] * st.w[0][a] * st.w[1][b] * h1;
                    } else {
                        gr[0] = st.dw[0][a] * st.w[1][b] * h1;
                        gr[1] = st.dw[1][b] * st.w[0][a] * h1;
                    }
                    for (int r = 0; r < D; ++r) for (int k = 0; k < D; ++k) gv[r][k] += vi[r] * (TF)gr[k];
                }
            }
        }
    }
}

// useflip must equal FLIP; the plain form reads the run-time value.
template <int D, int SPL, bool PIC, bool GRAD, bool FLIP, class TF, class Fetch>
__device__ __forceinline__ void gather(const Stencil<D, SPL>& st, const double* xp, double h1, const Fetch& fetch, bool apic,
                                       bool useflip, TF* pic, TF* fl, TF (&Bp)[3][3], TF (&gv)[3][3]) {
    if constexpr (sizeof(TF) == 4 && PIC) {
        // one copy of the gather per APIC setting behind a branch on the (wave-uniform) flag: with the flag read
        // inside, every affine sum of every row took a select
        if (apic) gather_rows<D, SPL, PIC, GRAD, FLIP, TF, Fetch, 1>(st, xp, h1, fetch, true, pic, fl, Bp, gv);
        else gather_rows<D, SPL, PIC, GRAD, FLIP, TF, Fetch, 0>(st, xp, h1, fetch, false, pic, fl, Bp, gv);
    } else if constexpr (sizeof(TF) == 4) {
        gather_rows<D, SPL, PIC, GRAD, FLIP, TF>(st, xp, h1, fetch, apic, pic, fl, Bp, gv);
    } else {
        gather_nodes<D, SPL, PIC, GRAD, TF>(st, xp, h1, fetch, apic, useflip, pic, fl, Bp, gv);
    }
}

// Block prologue shared by the gather kernels: returns false when the block has nothing to do.
struct BlockRange { uint32_t b, start, cnt; int bc[3]; };
template <int D>
__device__ __forceinline__ bool block_range(const StepState* S, const SortBufs& sb, uint32_t b, BlockRange& r) {
    if (S->stop) return false;
    r.b = b;
    if (r.b >= S->nblocks) return false;
    r.cnt = sb.count[r.b];
    if (r.cnt == 0) return false;
    r.start = sb.offset[r.b];
    block_coords<D>(r.b, S, r.bc);
    return true;
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2. For a kernel that walks the
// block-sorted particles with one thread each, this maps the workgroups that land on one XCD to one
// contiguous range of particles (an x-slab of the domain, with the same particle count on every XCD), so
// that each L2 holds the grid nodes of one slab instead of the whole grid.
__device__ __forceinline__ uint32_t xcd_slab(uint32_t w, uint32_t nwg) {
    constexpr uint32_t kXcd = 8;
    const uint32_t q = nwg / kXcd, rem = nwg % kXcd, x = w % kXcd, i = w / kXcd;
    return x < rem ? x * (q + 1) + i : rem * (q + 1) + (x - rem) * q + i;
}

// Kernels with one thread per sorted particle slot (k_defgrad) gather from LDS copies of the node
// tiles of the blocks that their workgroup's slots lie in. Slots are in block order, so these are the blocks
// from the first slot's to the last slot's; at 8 particles per cell a block holds about 440 particles and a
// workgroup of 256 slots spans one or two. SlotTiles::n tiles (at most 4, and at most about 32 KB of LDS, so
// that four workgroups fit on a CU) are copied; a particle whose block is beyond them gathers from the grid
// arrays, with the same values, weights and order.
template <int D, int SPL, class TF, int SETS> struct SlotTiles {
    static constexpr int bytes = SETS * TileStride<D>::v * BG<D, SPL>::NT * (int)sizeof(TF);
    static constexpr int n = (32768 / bytes) < 1 ? 1 : ((32768 / bytes) > 4 ? 4 : (32768 / bytes));
};
// ----------------------------------------------------------------------------------------------
// k_g2p: G2P (g2p.cpp:17-77), the particle velocity update (USL: position_update.cpp:14-23; MUSL: the
// same formula in musl.cpp:13-25) and the position update (position_update.cpp:12). The deformation
// update, which reads the particles' positions of the step's start, follows in k_defgrad.
// ----------------------------------------------------------------------------------------------
// Threads of a k_g2p workgroup (one workgroup per block, each thread a particle at a time): one wavefront
// leaves fewer threads idle in a block's last pass (a block holds about 440 particles at 8 per cell).
constexpr int kG2PThreads = 64;

// G2P, velocity and position update of source particle p into slot i of the destination layout, from the block's
// LDS tiles: false if its window leaves the tile. st, xp, vs and Bp return the stencil, the position of the step's
// start, the new velocity and the new affine matrix as stored (B zero without APIC).
// XLOADED: xp already holds the particle's position (the caller loaded it ahead).
template <int D, int SPL, bool FLIP, class TF, class CS, bool XLOADED = false>
__device__ __forceinline__ bool g2p_one(const Params& P, const PartStateT<TF>& src, const PartStateT<TF>& dst, uint32_t p, uint32_t i,
                                        const double* org, const uint32_t* nn, const CS& cs, const int* bc, double h1, double dt,
                                        bool apic, bool flip_rt, const TF* s_v, const TF* s_f, Stencil<D, SPL>& st,
                                        double (&xp)[3], TF (&vs)[3], TF (&Bp)[3][3], BoundsAcc& ba) {
    if (!XLOADED) for (int d = 0; d < 3; ++d) xp[d] = d < D ? src.x[d][p] : 0.0;
    int tl[3];
    if (!gather_stencil<D, SPL>(st, tl, xp, org, nn, cs, h1, bc)) return false;
    TF pic[3] = {0, 0, 0}, fl[3] = {0, 0, 0}, gv[3][3];
    for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) Bp[r][c] = TF(0);
    gather<D, SPL, true, false, FLIP, TF>(st, xp, h1, TileFetch<D, SPL, TF>{s_v, s_f, tl[0], tl[1], tl[2]}, apic, flip_rt, pic, fl, Bp, gv);
    double vnew[3] = {0, 0, 0}, xnew[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
        const double vo = FLIP ? (double)src.v[d][p] : 0.0;
        vnew[d] = (double)(TF)blend_velocity(P.flip_ratio, vo, (double)fl[d], (double)pic[d]);
        xnew[d] = xp[d] + dt * (double)pic[d];
    }
    if (P.pbc) {  // position_update.cpp:26-33
        if (xnew[0] > P.L[0]) xnew[0] = xnew[0] - P.L[0];
        else if (xnew[0] < 0) xnew[0] = P.L[0] + xnew[0];
    }
    if (P.change_particle_positions && xnew[0] > 0.8) { xnew[0] = -0.1; xnew[1] += 0.27; }
    for (int d = 0; d < D; ++d) { vs[d] = (TF)vnew[d]; dst.v[d][i] = vs[d]; dst.x[d][i] = xnew[d]; }
    if (!apic) for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) Bp[r][c] = TF(0);
    for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) dst.B[r + c * D][i] = Bp[r][c];
    ba.add<D>(xnew, vnew);
    return true;
}

template <int D, int SPL, bool FLIP, class TF>
__device__ __forceinline__ void g2p_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                              const SortBufs& sb, const GridPtrsT<TF>& g, uint32_t b) {
    using G = BG<D, SPL>;
    __shared__ TF s_v[3 * GatherTile<D, SPL>::N], s_f[3 * GatherTile<D, SPL>::N];
    __shared__ BlockCoordLDS<D, SPL> s_c;
    BlockRange br;
    if (!block_range<D>(S, sb, b, br)) return;
    const double h1 = P.one_over_dx, dt = S->dt;
    const double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
    const uint32_t nn[3] = {S->n[0], S->n[1], S->n[2]};
    const bool apic = P.flip_ratio < 0;
    constexpr bool useflip = FLIP;  // flip_ratio >= -1
    const bool flip_rt = P.flip_ratio >= -1;
    load_tile<D, SPL, TF>(s_v, g.v, br.bc, nn);
    if (useflip) load_tile<D, SPL, TF>(s_f, g.flip, br.bc, nn);
    s_c.template load<true>(g.coord, br.bc, nn);
    __syncthreads();
    uint32_t err = 0;
    BoundsAcc ba;
#pragma unroll 1
    for (uint32_t k = threadIdx.x; k < br.cnt; k += blockDim.x) {
        const uint32_t i = br.start + k;
        Stencil<D, SPL> st;
        double xp[3];
        TF vs[3], Bp[3][3];
        if (!g2p_one<D, SPL, FLIP, TF>(P, src, dst, sb.perm[i], i, org, nn, s_c.view(br.bc), br.bc, h1, dt, apic, flip_rt, s_v, s_f,
                                       st, xp, vs, Bp, ba))
            err |= MPM_ERR_GRID_BOUNDS;
    }
    reduce_bounds<D>(S, ba);
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0 && err) atomicOr(&S->err, err);
}
template <int D, int SPL, bool FLIP, class TF>
__global__ __launch_bounds__(kG2PThreads) void k_g2p(Params P, StepState* S, PartStateT<TF> src, PartStateT<TF> dst, SortBufs sb, GridPtrsT<TF> g) {
    g2p_body<D, SPL, FLIP, TF>(P, S, src, dst, sb, g, blockIdx.x);
}

// ----------------------------------------------------------------------------------------------
// k_g2p_musl: k_g2p and the MUSL re-scatter (musl.cpp:28-96) of a block in one pass, for the 3D quadratic split
// form. One wavefront per block, lane = dual cell as in the P2G scatter: at rank j a lane updates its cell's j-th
// particle (g2p_one, written to the next layout as k_g2p writes it) and adds that particle's re-scatter terms
// w_i (v + B D^-1 (x_i - x_p)) over its 27 window nodes to register partial sums, with the gather's weights and
// offsets rounded to binary32, which are the values the re-scatter's own stencil gives. The tile goes to staging
// as k_musl_scatter's does. This saves the re-scatter's pass over the particles: its loads of x, v and B and the
// three wavefronts that each evaluated every particle's stencil again.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL, class TF> constexpr bool kFusedMusl = D == 3 && SPL == 2 && sizeof(TF) == 4;
// All three first-axis slices in one wavefront: 27 nodes x 3 values of partial sums.
template <int D, int SPL, class TF> using FusedMuslShape = ScatterShape<D, SPL, D, TF, 108>;

// The re-scatter terms of one particle (MuslContrib::add on the stencil's weights): acc[a][b W + c][r] +=
// w_abc (v_r + (B D^-1)_r0 (x_a - x_p) + (B D^-1)_r1 (y_b - y_p) + (B D^-1)_r2 (z_c - z_p)), slice by slice with
// musl_slice, as k_musl_scatter sums them, with weights at or under the 1e-25 cut dropped (g2p.cpp:44,
// musl.cpp:48).
template <int D, int SPL, class TF, class SS>
__device__ __forceinline__ void musl_terms(const Stencil<D, SPL>& st, const double* xp, const TF (&v)[3], const TF (&B)[3][3], TF aD,
                                           TF (&acc)[SS::SPP][SS::NSH][D]) {
    constexpr int W = BG<D, SPL>::W;
    static_assert(D == 3 && SS::SPP == W && SS::NSH == W * W, "one wavefront covers the whole window");
    TF w[3][W], pd[3][W];
#pragma unroll
    for (int d = 0; d < 3; ++d)
#pragma unroll
        for (int k = 0; k < W; ++k) { w[d][k] = (TF)st.w[d][k]; pd[d][k] = (TF)(st.xi[d][k] - xp[d]); }
    TF Bc[3][3];
    for (int r = 0; r < 3; ++r) for (int k = 0; k < 3; ++k) Bc[r][k] = B[r][k] * aD;
#pragma unroll
    for (int a = 0; a < W; ++a) musl_slice<W, TF>(w[0][a], pd[0][a], w[1], pd[1], w[2], pd[2], v, Bc, acc[a]);
}

template <int D, int SPL, bool FLIP, class TF>
__device__ __forceinline__ void g2p_musl_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                              const SortBufs& sb, const GridPtrsT<TF>& g, const StagingT<TF>& stg, uint32_t b) {
    using G = BG<D, SPL>;
    using SS = FusedMuslShape<D, SPL, TF>;
    __shared__ TF s_v[3 * GatherTile<D, SPL>::N], s_f[FLIP ? 3 * GatherTile<D, SPL>::N : 1];
    __shared__ ScatterLDS<D, SPL, D, TF> L;
    BlockRange br;
    if (!block_range<D>(S, sb, b, br)) return;
    const double h1 = P.one_over_dx, dt = S->dt;
    const double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
    const uint32_t nn[3] = {S->n[0], S->n[1], S->n[2]};
    const bool apic = P.flip_ratio < 0;
    const bool flip_rt = P.flip_ratio >= -1;
    // The block's coordinates, the first kOrderLoads wavefront-widths of its slots and its cell offsets are
    // loaded into registers first, then the velocity tiles; the stores to LDS then wait for all of them at once
    // (one memory latency for the prologue instead of one per array).
    CoordPrefetch<D, SPL> cp;
    cp.issue(g.coord, br.bc, nn);
    const bool ordered = br.cnt <= (uint32_t)kMaxBP;  // the block's cell order, as the P2G scatter left it
    uint32_t pk[kOrderLoads], co[2];
#pragma unroll
    for (int u = 0; u < kOrderLoads; ++u) {
        const uint32_t k = threadIdx.x + (uint32_t)u * blockDim.x;
        pk[u] = ordered && k < br.cnt ? sb.perm[br.start + k] : 0u;
    }
#pragma unroll
    for (int u = 0; u < 2; ++u) {
        const uint32_t c = threadIdx.x + (uint32_t)u * blockDim.x;
        co[u] = ordered && c <= (uint32_t)G::NC ? (uint32_t)sb.cell_off[(size_t)br.b * (G::NC + 1) + c] : 0u;
    }
    load_tile<D, SPL, TF>(s_v, g.v, br.bc, nn);
    if (FLIP) load_tile<D, SPL, TF>(s_f, g.flip, br.bc, nn);
    cp.put(L.coords);
    if (ordered) {  // slot start + k holds source particle p[k]
#pragma unroll
        for (int u = 0; u < kOrderLoads; ++u) {
            const uint32_t k = threadIdx.x + (uint32_t)u * blockDim.x;
            if (k < br.cnt) L.p[k] = pk[u];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u) {
            const uint32_t c = threadIdx.x + (uint32_t)u * blockDim.x;
            if (c <= (uint32_t)G::NC) L.cstart[c] = co[u];
        }
        const uint32_t head = kOrderLoads * blockDim.x;
        if (br.cnt > head) copy_to_lds(L.p + head, sb.perm + br.start + head, br.cnt - head);
        if (threadIdx.x == 0) L.cstart[G::NC + 1] = br.cnt;
    }
    __syncthreads();
    const auto cs = L.coords.view(br.bc);
    uint32_t err = 0;
    BoundsAcc ba;
    if (ordered) {
        TF acc[SS::SPP][SS::NSH][D];
        zero_acc<D, SPL, D, TF>(acc);
        const TF aD = (TF)P.apicDinverse;
        const int lane = threadIdx.x & 63;
        const uint32_t n_c = L.cstart[lane + 1] - L.cstart[lane];
        const uint32_t n_max = wave_max_u32(n_c);
        const unsigned long long below = (1ull << lane) - 1ull;
        // the next rank's position is loaded while this rank's particle is updated and scattered
        uint32_t kb = 0, kn = 0;
        bool hasn = 0 < n_c;
        double xn[3] = {0, 0, 0};
        {
            const unsigned long long m = __ballot(hasn);
            kn = kb + (uint32_t)__popcll(m & below);
            kb += (uint32_t)__popcll(m);
            if (hasn) for (int d = 0; d < 3; ++d) xn[d] = src.x[d][L.p[kn]];
        }
#pragma unroll 1
        for (uint32_t j = 0; j < n_max; ++j) {
            const bool has = hasn;
            const uint32_t k = kn;
            double xp[3] = {xn[0], xn[1], xn[2]};
            hasn = j + 1 < n_c;
            const unsigned long long m = __ballot(hasn);
            kn = kb + (uint32_t)__popcll(m & below);
            kb += (uint32_t)__popcll(m);
            if (hasn) for (int d = 0; d < 3; ++d) xn[d] = src.x[d][L.p[kn]];
            if (has) {
                Stencil<D, SPL> st;
                TF vs[3], Bp[3][3];
                if (g2p_one<D, SPL, FLIP, TF, decltype(cs), true>(P, src, dst, L.p[k], br.start + k, org, nn, cs, br.bc, h1, dt, apic,
                                                                  flip_rt, s_v, s_f, st, xp, vs, Bp, ba))
                    musl_terms<D, SPL, TF, SS>(st, xp, vs, Bp, aD, acc);
                else
                    err |= MPM_ERR_GRID_BOUNDS;
            }
        }
        // particles whose window base lies outside the block's cells (the P2G scatter already raised the error
        // bit for them): updated, not scattered
        for (uint32_t k = L.cstart[G::NC] + threadIdx.x; k < br.cnt; k += blockDim.x) {
            Stencil<D, SPL> st;
            double xp[3];
            TF vs[3], Bp[3][3];
            (void)g2p_one<D, SPL, FLIP, TF>(P, src, dst, L.p[k], br.start + k, org, nn, cs, br.bc, h1, dt, apic, flip_rt, s_v, s_f,
                                            st, xp, vs, Bp, ba);
            err |= MPM_ERR_GRID_BOUNDS;
        }
        ordered_tile<D, SPL, D, TF, SS>(L, acc);
        write_tile<D, SPL, D, TF>(L, stg, br.b);
    } else {
        // a block too dense to order: the update in slot order, then the re-scatter with LDS atomics as
        // k_musl_scatter does for it
#pragma unroll 1
        for (uint32_t k = threadIdx.x; k < br.cnt; k += blockDim.x) {
            const uint32_t i = br.start + k;
            Stencil<D, SPL> st;
            double xp[3];
            TF vs[3], Bp[3][3];
            if (!g2p_one<D, SPL, FLIP, TF>(P, src, dst, sb.perm[i], i, org, nn, cs, br.bc, h1, dt, apic, flip_rt, s_v, s_f, st, xp, vs, Bp,
                                           ba))
                err |= MPM_ERR_GRID_BOUNDS;
        }
        __syncthreads();  // the block's new velocities and affine matrices are written
        MuslContrib<D, SPL, TF> cf;
        cf.src = src; cf.dst = dst;
        for (int d = 0; d < 3; ++d) { cf.org[d] = org[d]; cf.nn[d] = nn[d]; }
        cf.cs = cs;
        cf.h1 = h1; cf.aD = P.apicDinverse; cf.apic = apic;
        err |= scatter_block_atomic<D, SPL, D, TF>(L, cf, stg, sb, br.b, br.start, br.cnt, br.bc);
    }
    reduce_bounds<D>(S, ba);
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0 && err) atomicOr(&S->err, err);
}
template <int D, int SPL, bool FLIP, class TF>
__global__ __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2))) void k_g2p_musl(Params P, StepState* S, PartStateT<TF> src, PartStateT<TF> dst, SortBufs sb, GridPtrsT<TF> g,
                                                 StagingT<TF> stg) {
    g2p_musl_body<D, SPL, FLIP, TF>(P, S, src, dst, sb, g, stg, blockIdx.x);
}

// The node-major velocity tiles of blocks blo .. blo + nt - 1 (at most NTL; tile t at s_v + t ST NT) and their node
// coordinates, for a workgroup of at least NTL * D * LC threads. Every thread issues its coordinate load and the
// loads of all its tile nodes (kTileLoads at a time) before it stores any, so that the tiles cost one memory
// latency instead of several per tile. Blocks without particles inside the span are loaded too; nodes outside
// the grid are zero.
template <int D, int SPL, class TF, int NTL>
__device__ __forceinline__ void load_slot_tiles(TF* s_v, BlockCoordLDS<D, SPL>* s_c, uint32_t blo,
                                                uint32_t nt, const StepState* S, TF* const* src, double* const* coord,
                                                const uint32_t* n) {
    using G = BG<D, SPL>;
    constexpr int ST = TileStride<D>::v, LC = BlockCoordLDS<D, SPL>::LC;
    static_assert(NTL * D * LC <= kBlock, "one coordinate entry per thread of a kBlock workgroup");
    const uint32_t tid = threadIdx.x;
    // coordinates: entry (t, d, i) per thread
    const bool hc = tid < nt * (uint32_t)(D * LC);
    int ct = 0, cd = 0, ci = 0;
    double cv = 0.0;
    if (hc) {
        ct = (int)(tid / (D * LC));
        const int r = (int)tid - ct * D * LC;
        cd = r / LC; ci = r - cd * LC;
        int bc[3];
        block_coords<D>(blo + (uint32_t)ct, S, bc);
        const int gi = G::E * bc[cd] - 1 + ci;
        // the axis's array chosen by selects: indexing the pointer array with a lane's axis would load the
        // pointer from memory and wait for every load in flight before the coordinate's own
        const double* ca = cd == 0 ? coord[0] : (cd == 1 ? coord[1] : coord[D == 3 ? 2 : 1]);
        cv = (gi >= 0 && gi < (int)n[cd]) ? ca[gi] : 0.0;
    }
    const uint32_t total = nt * (uint32_t)G::NT;
    for (uint32_t k0 = tid; k0 < total; k0 += kTileLoads * blockDim.x) {
        TF val[kTileLoads][3];
#pragma unroll
        for (int u = 0; u < kTileLoads; ++u) {
            const uint32_t e = k0 + (uint32_t)u * blockDim.x;
            const uint32_t t = e / (uint32_t)G::NT;
            const int q = (int)(e - t * (uint32_t)G::NT);
            int bc[3] = {0, 0, 0};
            if (e < total) block_coords<D>(blo + t, S, bc);
            int tq[3] = {0, 0, 0};
            if (D == 3) { tq[2] = q % G::T; const int r = q / G::T; tq[1] = r % G::T; tq[0] = r / G::T; }
            else { tq[1] = q % G::T; tq[0] = q / G::T; }
            bool in = e < total;
            int gx[3] = {0, 0, 0};
            for (int d = 0; d < D; ++d) { gx[d] = G::E * bc[d] + tq[d]; in = in && gx[d] < (int)n[d]; }
            const uint64_t node = in ? node_offset<D>(gx[0], gx[1], gx[2], n) : 0ull;
#pragma unroll
            for (int c = 0; c < D; ++c) val[u][c] = in ? src[c][node] : TF(0);
        }
#pragma unroll
        for (int u = 0; u < kTileLoads; ++u) {
            const uint32_t e = k0 + (uint32_t)u * blockDim.x;
            if (e >= total) break;
            const uint32_t t = e / (uint32_t)G::NT;
            const int q = (int)(e - t * (uint32_t)G::NT);
            TF* row = s_v + (size_t)t * (ST * G::NT) + q * ST;
            for (int c = 0; c < D; ++c) row[c] = val[u][c];
            for (int c = D; c < ST; ++c) row[c] = TF(0);
        }
    }
    if (hc) s_c[ct].c[cd][ci] = cv;
}

// ----------------------------------------------------------------------------------------------
// k_defgrad: deformationUpdate (deformation_update.cpp:16-46) with the return map, on the velocity of the
// step's grid update (USL) or of the MUSL re-scatter. One thread per sorted particle slot (the return map is
// long and register-heavy, so the kernel keeps one particle per thread and no loop); grad v is gathered from
// the LDS tiles of the workgroup's blocks (SlotTiles; a gather from the grid arrays alone took about 40 % of
// the kernel).
// ----------------------------------------------------------------------------------------------
template <int D, int SPL, class TF, int MAT>
__device__ __forceinline__ void defgrad_body(const Params& P, StepState* S, const PartStateT<TF>& src, const PartStateT<TF>& dst,
                                             const SortBufs& sb, const GridPtrsT<TF>& g, uint32_t n, uint32_t bid) {
    using G = BG<D, SPL>;
    constexpr int NTL = SlotTiles<D, SPL, TF, 1>::n;
    __shared__ __attribute__((aligned(16))) TF s_v[NTL][TileStride<D>::v * G::NT];
    __shared__ BlockCoordLDS<D, SPL> s_c[NTL];
    __shared__ uint32_t s_lo, s_hi;
    if (S->stop) return;
    const uint32_t base = bid * blockDim.x;
    if (base >= n) return;
    const uint32_t i = base + threadIdx.x;
    const bool active = i < n;
    const uint32_t nn[3] = {S->n[0], S->n[1], S->n[2]};
    // The blocks of the slots are those of the first and the last slot's particle (slots are in block order).
    // The split form loads each particle's slot, block, position, F and plastic scalars first, so that these
    // loads are in flight while the tiles load (C2: 357 -> 345 us). In the plain form that order takes the
    // kernel from 118 to 131 VGPRs and 337 to 370 us, so there one thread reads the span and every particle
    // is loaded after the tiles.
    constexpr bool kEarlyState = sizeof(TF) == 4;
    uint32_t p = 0, key = 0, idp = 0;
    double xp[3] = {0, 0, 0};
    DefgradIn<D> in;
    const uint32_t last = min(n, base + blockDim.x) - 1;
    if constexpr (kEarlyState) {
        if (active) {
            p = sb.perm[i];
            key = sb.key[p];
            idp = src.id[p];
            for (int d = 0; d < D; ++d) xp[d] = src.x[d][p];
            load_defgrad_in<D, TF>(src, p, in);
            if (i == base) s_lo = key;
            if (i == last) s_hi = key;
        }
    } else if (threadIdx.x == 0) {
        s_lo = sb.key[sb.perm[base]];
        s_hi = sb.key[sb.perm[last]];
    }
    __syncthreads();
    const uint32_t blo = s_lo, nt = min((uint32_t)NTL, s_hi - s_lo + 1);
    load_slot_tiles<D, SPL, TF, NTL>(&s_v[0][0], s_c, blo, nt, S, g.v, g.coord, nn);
    __syncthreads();
    uint32_t err = 0, pcount = 0;
    if (active) {
        if constexpr (!kEarlyState) {
            p = sb.perm[i];
            for (int d = 0; d < D; ++d) xp[d] = src.x[d][p];
            key = sb.key[p];
            idp = src.id[p];
            load_defgrad_in<D, TF>(src, p, in);
        }
        const double h1 = P.one_over_dx, dt = S->dt;
        const double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
        const uint32_t t = key - blo;
        Stencil<D, SPL> st;
        TF gvt[3][3] = {{0}}, pic[3], fl[3], Bp[3][3];
        bool ok;
        if (t < nt) {
            int bc[3], tl[3];
            block_coords<D>(blo + t, S, bc);
            ok = gather_stencil<D, SPL>(st, tl, xp, org, nn, s_c[t].view(bc), h1, bc);
            if (ok) gather<D, SPL, false, true, false, TF>(st, xp, h1, TileFetch<D, SPL, TF, true>{s_v[t], s_v[t], tl[0], tl[1], tl[2]}, false, false, pic, fl, Bp, gvt);
        } else {
            ok = make_stencil<D, SPL>(st, xp, org, nn, GlobalCoords{g.coord, nn}, h1, false);
            if (ok) gather<D, SPL, false, true, false, TF>(st, xp, h1, GridFetch<D, TF>{g, nn, st.base[0], st.base[1], st.base[2]}, false, false, pic, fl, Bp, gvt);
        }
        dst.id[i] = idp;  // the particle's id moves with it to its sorted slot (for every slot, in bounds or not)
        if (!ok) {
            err |= MPM_ERR_GRID_BOUNDS;
        } else {
            double gv[3][3];
            for (int r = 0; r < 3; ++r) for (int c = 0; c < 3; ++c) gv[r][c] = (double)gvt[r][c];
            deform_and_project<D, TF, MAT>(P, dt, in, dst, i, gv, err, pcount);
        }
    }
    pcount = wave_sum_u32(pcount);
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0) {
        add_plastic_count(S, pcount);
        if (err) atomicOr(&S->err, err);
    }
}
// At least 3 wavefronts per SIMD (at most 168 VGPRs): the MCC-family return maps take up to 211 and would run 2
// (C3, mu(I)-MCC: 1080 -> 935 us with 156 bytes of spill; a bound of 4 spilled more and ran 1139 us). Kernels
// under 168 VGPRs are unchanged. The batched kb_defgrad is not bounded (see ScatterLaunch::WPE).
constexpr int kDefgradWaves = 3;
template <int D, int SPL, class TF, int MAT>
__global__ __launch_bounds__(kBlock) __attribute__((amdgpu_waves_per_eu(kDefgradWaves))) void k_defgrad(Params P, StepState* S, PartStateT<TF> src, PartStateT<TF> dst, SortBufs sb, GridPtrsT<TF> g, uint32_t n) {
    defgrad_body<D, SPL, TF, MAT>(P, S, src, dst, sb, g, n, xcd_slab(blockIdx.x, gridDim.x));
}

// ----------------------------------------------------------------------------------------------
// k_stress_measures: the per-particle output quantities of saveParticleData (save_data.cpp:9-38):
// tau = P(F) F^T (HenckyPiola without the strain clamp of the return map, or NeoHookeanPiola),
// pressure p = -tr(tau)/dim, deviatoric measure q = sqrt(c tau_dev : tau_dev) with c = 3/2 (use_mises_q)
// or 1/2, Je = det F. Written at the caller's particle index.
// ----------------------------------------------------------------------------------------------
template <int D, class TF>
__global__ __launch_bounds__(kBlock) void k_stress_measures(Params P, PartStateT<TF> ps, uint32_t n, double* pr, double* qd, double* je) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    M<D> F;
    for (int k = 0; k < D * D; ++k) F.m[k] = ps.F[k][i];
    double tau[6];
    if (P.elastic_model == MPM_ELASTIC_NEOHOOKEAN) {
        neohookean_tau<D>(F, P.mu, P.lambda, tau);
    } else {
        M<D> U, W; V<D> s;
        svd<D>(F, U, s, W, P.svd_tol);
        V<D> ls; for (int k = 0; k < D; ++k) ls[k] = log_d(fabs(s[k]));
        hencky_tau<D>(U, ls, P.mu, P.lambda, tau);
    }
    double tr = tau[0] + tau[1] + (D == 3 ? tau[2] : 0.0);
    const double p = -tr / D;
    double dd = 0.0;
    for (int k = 0; k < D; ++k) { const double t = tau[k] + p; dd += t * t; }
    for (int k = D; k < SymPack<D>::n; ++k) dd += 2.0 * tau[k] * tau[k];  // off-diagonal entries appear twice
    const double q = P.use_mises_q ? sqrt(1.5 * dd) : sqrt(0.5 * dd);
    const uint32_t id = ps.id[i];
    pr[id] = p; qd[id] = q; je[id] = det(F);
}

// ----------------------------------------------------------------------------------------------
// k_finish: moveObjects (simulation.cpp:277-281) and the loop bookkeeping of simulate()
// (simulation.cpp:176-200), conservation checks (simulation.cpp:292-333) when requested.
// ----------------------------------------------------------------------------------------------
template <int D>
__device__ __forceinline__ void finish_lane0(const Params& P, StepState* S, uint32_t n_particles, uint32_t pc);
template <int D>
__device__ __forceinline__ void finish_body(const Params& P, StepState* S, uint32_t n_particles) {
    if (S->stop) return;
    uint32_t pc = 0;
    for (int s = threadIdx.x; s < kRedSlots; s += 64) { pc += S->red[s].plastic; S->red[s].plastic = 0; }
    pc = wave_sum_u32(pc);
    if (threadIdx.x == 0) finish_lane0<D>(P, S, n_particles, pc);
}
template <int D>
__device__ __forceinline__ void finish_lane0(const Params& P, StepState* S, uint32_t n_particles, uint32_t pc) {
    S->plastic_count += pc;
    const double dt = S->dt;
    for (int k = 0; k < P.n_plates; ++k) plate_move<D>(S->plates[k], dt, P.frame_dt, S->time);
    if (S->check_conservation) {
        double pm = P.particle_mass * (n_particles + (P.pbc ? S->n_ghosts : 0u));  // Np includes the copies here
        if (fabs(S->cons_grid_mass - pm) > 1e-10 * pm) S->err |= MPM_ERR_CONSERVATION;
        double dn = 0, pn = 0;
        for (int d = 0; d < D; ++d) { double e = S->cons_grid_mom[d] - S->cons_part_mom[d]; dn += e * e; pn += S->cons_part_mom[d] * S->cons_part_mom[d]; }
        if (sqrt(dn) > 1e-10 * sqrt(pn)) S->err |= MPM_ERR_CONSERVATION;
    }
    S->steps_done++;
    S->xsel ^= 1u;
    if (S->err & MPM_ERR_EXIT_MASK) { S->stop = 1; return; }
    S->time += dt;
    S->step++;
    S->frame_ended = 0;
    if (P.frame_dt * (double)(S->frame + 1) - S->time < 1e-15) {
        S->frame++;
        S->frame_ended = 1;
        if (S->stop_at_frame_end) S->stop = 1;
    }
    if (fabs(P.final_time - S->time) < 1e-15 || P.final_time < S->time) { S->finished = 1; S->stop = 1; }
    if (S->steps_done >= S->step_limit) S->stop = 1;
}
template <int D>
__global__ __launch_bounds__(64) void k_finish(Params P, StepState* S, uint32_t n_particles) { finish_body<D>(P, S, n_particles); }
// The end of one step and the set-up of the next in one launch (the steps of a batch but its last). The
// workgroup is one wavefront on one CU, so the barrier's workgroup-scope fence makes lane 0's stores to
// StepState visible to the set-up's loads.
template <int D, class TF>
__global__ __launch_bounds__(64) void k_finish_setup(Params P, StepState* S, GridPtrsT<TF> g, SortBufs sb, uint32_t n_particles) {
    finish_body<D>(P, S, n_particles);
    __syncthreads();
    setup_body<D, TF, true>(P, S, g, &sb);
}

}  // namespace mpm
