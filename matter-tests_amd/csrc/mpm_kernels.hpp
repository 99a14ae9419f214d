// mpm_kernels.hpp — gfx950 kernels of one MPM step (plain FP64 form).
//
// Step pipeline (one HIP stream, no host synchronisation inside a batch of steps):
//   k_setup        1 workgroup: dt (updateDt), grid geometry (remeshFixedInit/Cont + arange), capacity check
//   k_zero_grid    clear the node accumulators of the dense grid (resizeGrid) and the block histogram
//   k_bin_count    block key of every particle (4x4x4-node blocks in 3D, 8x8 in 2D) + histogram
//   k_bin_scan     exclusive prefix sum of the histogram (one workgroup)
//   k_bin_scatter  counting-sort permutation: sorted position -> particle of the previous layout
//   k_p2g          scatter mass, APIC momentum, MIBF friction and stress force of 256 sorted particles
//                  into an LDS tile over their stencils' bounding box, then flush the tile with one
//                  global atomic per non-zero node value (P2G + force part of explicitEulerUpdate)
//   k_grid_update  normalise, explicit update, boundary conditions, flip       (P2G tail + grid loop of
//                  explicitEulerUpdate + boundaryCollision)
//   USL:  k_g2p_fused    gather, F update, return map, position/velocity update  (G2P + deformationUpdate
//                        + plasticity + positionUpdate), reductions for the next step
//   MUSL: k_g2p_musl     gather, velocity update, new position, LDS-tiled momentum re-scatter (G2P + MUSL)
//         k_musl_norm    divide by m_i/m_p                                         (MUSL tail)
//         k_defgrad      gather grad v, F update, return map                       (deformationUpdate)
//   k_finish       1 thread: moveObjects, time/frame bookkeeping of simulate()
//
// Particle state is double-buffered. Every step reads the previous layout through the sort permutation
// and writes all fields in sorted order into the other buffer, so particles stay spatially sorted at no
// extra traffic; each particle carries its original index (id) for downloads.
// Every kernel returns at once when StepState::stop is set, so the host can enqueue many steps and
// let the device stop at an error, a frame end, the final time or a capacity overflow.
#pragma once
#include <hip/hip_runtime.h>
#include "mpm_math.hpp"

namespace mpm {

// One layout of the particle state (structure of arrays).
struct PartState {
    double* x[3];
    double* v[3];
    double* B[9];
    double* F[9];
    double* tau[6];
    double* eps_dev; double* eps_vol; double* eps_prad; double* dgamma; double* visc; double* muI;
    uint32_t* id;    // index of the particle in the caller's order
};

// Counting-sort work arrays.
struct SortBufs {
    uint32_t* key;     // block key per particle of the source layout
    uint32_t* perm;    // sorted position -> particle of the source layout
    uint32_t* count;   // particles per block
    uint32_t* offset;  // exclusive prefix sum of count
    uint32_t* cursor;  // scatter cursors
};

struct GridPtrs {
    double* mass;    // sum of weights during P2G, m_i afterwards
    double* mv[3];   // momentum (sum of w (v + B d D^-1))
    double* f[3];    // sum of tau grad w
    double* fric;    // MIBF: sum of w muI, then divided by sum of w
    double* v[3];    // node velocity after the update (MUSL: after re-scatter)
    double* flip[3];
    double* mv2[3];  // MUSL re-scatter accumulator
    double* coord[3];
};

constexpr int kBlock = 256;
constexpr int kTileCap = 576;  // nodes of one LDS tile (7 x 7 x 11 covers a chunk straddling two 3D blocks)

template <int D> struct BlockEdge { static constexpr int v = D == 3 ? 4 : 8; };

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
__device__ __forceinline__ uint32_t wave_or_u32(uint32_t v) {
    for (int o = 32; o > 0; o >>= 1) v |= __shfl_xor(v, o, 64);
    return v;
}

// Wavefront-aggregated atomicAdd(ctr + key, 1): lanes with equal keys share one atomic; each lane gets
// its own slot (old value + rank among the lanes with its key). Lanes with active == false take no slot.
__device__ __forceinline__ uint32_t wave_claim(uint32_t* ctr, uint32_t key, bool active) {
    const int lane = threadIdx.x & 63;
    unsigned long long pending = __ballot(active);
    uint32_t slot = 0;
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const uint32_t lk = __shfl(key, leader, 64);
        const unsigned long long same = __ballot(active && key == lk) & pending;
        uint32_t base = 0;
        if (lane == leader) base = atomicAdd(ctr + lk, (uint32_t)__popcll(same));
        base = __shfl(base, leader, 64);
        if ((same >> lane) & 1ull) slot = base + (uint32_t)__popcll(same & ((1ull << lane) - 1ull));
        pending &= ~same;
    }
    return slot;
}

// Per-step reductions of the new particle state (max speed for updateDt, bounds for remesh).
// Inactive lanes pass neutral values.
template <int D>
__device__ __forceinline__ void reduce_step_bounds(StepState* S, bool active, const double* x, const double* v) {
    unsigned long long v2 = 0ull;
    long long mn[3], mx[3];
    for (int d = 0; d < 3; ++d) { mn[d] = 0x7FFFFFFFFFFFFFFFLL; mx[d] = (-0x7FFFFFFFFFFFFFFFLL - 1); }
    if (active) {
        double s = v[0] * v[0]; for (int d = 1; d < D; ++d) s += v[d] * v[d];
        v2 = (unsigned long long)__double_as_longlong(s);
        for (int d = 0; d < D; ++d) { mn[d] = ord_key(x[d]); mx[d] = ord_key(x[d]); }
    }
    v2 = wave_max_u64(v2);
    for (int d = 0; d < D; ++d) { mn[d] = wave_min_i64(mn[d]); mx[d] = wave_max_i64(mx[d]); }
    // workgroup combine through LDS, then one set of atomics per workgroup into its slot
    __shared__ long long s_red[kBlock / 64][7];
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

// Folds the bound slots into StepState and resets them (one wavefront of 64 lanes calls). Min/max folds
// are idempotent, so a step that is retried after a grid reallocation sees the same bounds.
template <int D>
__device__ __forceinline__ void fold_bound_slots(StepState* S) {
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
    if (threadIdx.x == 0) {
        S->vmax2_bits = max(S->vmax2_bits, v2);
        for (int d = 0; d < D; ++d) { S->xmin_key[d] = min(S->xmin_key[d], mn[d]); S->xmax_key[d] = max(S->xmax_key[d], mx[d]); }
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
        int v = (int)floor((xp[d] - org[d]) * h1) - 1;
        if (clamp && v < 0) v = 0;
        if (v < 0 || v + 3 >= (int)n[d]) ok = false;
        b[d] = v;
    }
    return ok;
}

// Evaluates the 4-node window of the reference (weights are those of wip/grad_wip, tools.hpp:222-252) and
// keeps its non-zero part: for the quadratic spline exactly one end node has zero weight (both at
// f = 0.5), and skipping an exactly-zero term leaves every sum of the reference unchanged.
template <int D, int SPL>
__device__ __forceinline__ bool make_stencil(Stencil<D, SPL>& st, const double* xp, const double* org, const uint32_t* n,
                                             double* const* coord, double h1, bool clamp) {
    int b[3] = {0, 0, 0};
    if (!stencil_base<D>(b, xp, org, n, h1, clamp)) return false;
#pragma unroll
    for (int d = 0; d < D; ++d) {
        double w4[4], dw4[4], xi4[4];
#pragma unroll
        for (int a = 0; a < 4; ++a) {
            const double c = coord[d][b[d] + a];
            const double t = (xp[d] - c) * h1;
            xi4[a] = c;
            w4[a] = BSpline<SPL>::N(t);
            dw4[a] = BSpline<SPL>::dN(t);
        }
        const bool shift = SPL == 2 && w4[0] == 0.0;
        st.base[d] = b[d] + (shift ? 1 : 0);
#pragma unroll
        for (int a = 0; a < Stencil<D, SPL>::cnt; ++a) {
            if (SPL == 2) {
                st.w[d][a] = shift ? w4[a + 1] : w4[a];
                st.dw[d][a] = shift ? dw4[a + 1] : dw4[a];
                st.xi[d][a] = shift ? xi4[a + 1] : xi4[a];
            } else {
                st.w[d][a] = w4[a]; st.dw[d][a] = dw4[a]; st.xi[d][a] = xi4[a];
            }
        }
    }
    return true;
}

__device__ __forceinline__ uint32_t node_index3(int i, int j, int k, const uint32_t* n) { return ((uint32_t)i * n[1] + (uint32_t)j) * n[2] + (uint32_t)k; }

template <int D>
__device__ __forceinline__ uint32_t node_index(int i, int j, int k, const uint32_t* n) {
    return D == 3 ? node_index3(i, j, k, n) : (uint32_t)i * n[1] + (uint32_t)j;
}

// LDS tile over the node bounding box of the 4-node windows of a workgroup's particles.
struct Tile {
    int lo[3], ext[3], vol;
    bool fits;
};

// Workgroup-wide bounding box of the windows base..base+W-1 of the participating lanes (all threads call).
template <int D, int W>
__device__ __forceinline__ Tile tile_bbox(bool part, const int* base, int* s_box) {
    int mn[3], mx[3];
    for (int d = 0; d < 3; ++d) { mn[d] = part && d < D ? base[d] : 0x7FFFFFFF; mx[d] = part && d < D ? base[d] + W - 1 : -0x7FFFFFFF; }
    if (threadIdx.x < 6) s_box[threadIdx.x] = threadIdx.x < 3 ? 0x7FFFFFFF : -0x7FFFFFFF;
    __syncthreads();
    for (int d = 0; d < D; ++d) { mn[d] = wave_min_i32(mn[d]); mx[d] = wave_max_i32(mx[d]); }
    if ((threadIdx.x & 63) == 0) for (int d = 0; d < D; ++d) { atomicMin(&s_box[d], mn[d]); atomicMax(&s_box[3 + d], mx[d]); }
    __syncthreads();
    Tile t;
    t.vol = 1;
    bool empty = false;
    for (int d = 0; d < 3; ++d) {
        if (d < D) {
            t.lo[d] = s_box[d];
            empty = empty || s_box[3 + d] < s_box[d];  // no participating lane
            t.ext[d] = empty ? 1 : s_box[3 + d] - s_box[d] + 1;
        } else { t.lo[d] = 0; t.ext[d] = 1; }
    }
    if (empty) { t.vol = 0; t.fits = true; for (int d = 0; d < 3; ++d) { t.lo[d] = 0; t.ext[d] = 1; } return t; }
    long long vol = 1;
    for (int d = 0; d < D; ++d) vol *= t.ext[d];
    t.fits = vol <= kTileCap;
    t.vol = t.fits ? (int)vol : 0;
    return t;
}

template <int D>
__device__ __forceinline__ int tile_local(const Tile& t, int i, int j, int k) {
    return D == 3 ? ((i - t.lo[0]) * t.ext[1] + (j - t.lo[1])) * t.ext[2] + (k - t.lo[2]) : (i - t.lo[0]) * t.ext[1] + (j - t.lo[1]);
}

// Adds the non-zero tile values to the grid (one global atomic each); all threads call.
template <int D, int NV>
__device__ __forceinline__ void tile_flush(const Tile& t, const double* tile, double* const* dst, const uint32_t* n) {
    for (int q = threadIdx.x; q < t.vol; q += blockDim.x) {
        int i, j, k;
        if (D == 3) { k = q % t.ext[2]; int r = q / t.ext[2]; j = r % t.ext[1]; i = r / t.ext[1]; }
        else { j = q % t.ext[1]; i = q / t.ext[1]; k = 0; }
        const uint32_t node = node_index<D>(t.lo[0] + i, t.lo[1] + j, t.lo[2] + k, n);
        for (int v = 0; v < NV; ++v) {
            const double a = tile[v * kTileCap + q];
            if (a != 0.0) atomic_add_f64(dst[v] + node, a);
        }
    }
}

// ----------------------------------------------------------------------------------------------
// k_prepare: Kirchhoff stress of the initial F and the reductions that the first step needs.
// ----------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_prepare(Params P, StepState* S, PartState ps, uint32_t n) {
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
        for (int i = 0; i < SymPack<D>::n; ++i) ps.tau[i][p] = tau[i];
        for (int d = 0; d < D; ++d) { x[d] = ps.x[d][p]; v[d] = ps.v[d][p]; }
        ps.id[p] = p;
    }
    reduce_step_bounds<D>(S, active, x, v);
}

// ----------------------------------------------------------------------------------------------
// k_setup: updateDt (update_dt.cpp:5-66) and remeshFixedInit/Cont (remesh.cpp:38-311).
// One workgroup of 64 lanes; lane d builds the node coordinates of axis d (arange, tools.hpp:291-296).
// ----------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(64) void k_setup(Params P, StepState* S, GridPtrs g) {
    __shared__ double lo_s[3], stop_s[3];
    __shared__ int go;
    fold_bound_slots<D>(S);
    __syncthreads();
    if (threadIdx.x == 0) {
        go = 0;
        if (!S->stop && S->frame >= P.end_frame) { S->stop = 1; S->finished = 1; }
        if (!S->stop) {
            double max_speed = sqrt(__longlong_as_double((long long)S->vmax2_bits));
            uint32_t warn = 0;
            if (max_speed >= P.wave_speed) warn |= MPM_WARN_SUPERSONIC;
            double dt;
            if (fabs(max_speed) > 1e-10) { double dt_cfl = P.cfl * P.dx / max_speed; dt = dt_cfl < P.dt_max ? dt_cfl : P.dt_max; }
            else dt = P.dt_max;
            double c1 = P.frame_dt * (double)(S->frame + 1) - S->time; dt = c1 < dt ? c1 : dt;
            double c2 = P.final_time - S->time; dt = c2 < dt ? c2 : dt;
            double grav[3] = {S->gravity[0], S->gravity[1], S->gravity[2]};
            if (P.gravity_special) {
                for (int d = 0; d < D; ++d) grav[d] = (S->time < P.gravity_time) ? P.gravity_final[d] * S->time / P.gravity_time : P.gravity_final[d];
            }
            double mn[3], mx[3], lo[3], hi[3];
            for (int d = 0; d < D; ++d) { mn[d] = key_val(S->xmin_key[d]); mx[d] = key_val(S->xmax_key[d]); }
            double mni[3], mxi[3], loi[3], hii[3];
            const double dx = P.dx;
            if (S->step == 0) {
                if (P.grid_reference_point[0] < 1e10) {
                    for (int d = 0; d < D; ++d) {
                        if (P.grid_reference_point[d] < mn[d]) mn[d] = P.grid_reference_point[d];
                        else if (P.grid_reference_point[d] > mx[d]) mx[d] = P.grid_reference_point[d];
                    }
                }
                for (int d = 0; d < D; ++d) {
                    mxi[d] = mx[d]; mni[d] = mn[d];
                    loi[d] = mn[d] - dx * 2.0; hii[d] = mx[d] + dx * 2.0;
                    lo[d] = loi[d]; hi[d] = hii[d];
                }
            } else {
                for (int d = 0; d < D; ++d) {
                    mni[d] = S->min_init[d]; mxi[d] = S->max_init[d]; loi[d] = S->low_init[d]; hii[d] = S->high_init[d];
                    if (mx[d] < mxi[d]) { unsigned rf = (unsigned)floor(fmax(0.0, (mxi[d] - mx[d]) / dx - 1e-8 * dx)); hi[d] = hii[d] - rf * dx; }
                    else { unsigned ef = (unsigned)floor(fmax(0.0, (mx[d] - mxi[d]) / dx - 1e-8 * dx)); hi[d] = hii[d] + ef * dx; }
                    if (mn[d] > mni[d]) { unsigned rf = (unsigned)floor(fmax(0.0, (mn[d] - mni[d]) / dx - 1e-8 * dx)); lo[d] = loi[d] + rf * dx; }
                    else { unsigned ef = (unsigned)floor(fmax(0.0, (mni[d] - mn[d]) / dx - 1e-8 * dx)); lo[d] = loi[d] - ef * dx; }
                }
            }
            // node counts of arange(lo, hi + dx, dx)
            uint32_t cnt[3] = {1, 1, 1};
            bool fits = true;
            for (int d = 0; d < D; ++d) {
                double stop = hi[d] + dx;
                uint32_t c = 0;
                for (double val = lo[d]; val < stop; val += dx) { if (++c > S->cap_axis) break; }
                cnt[d] = c;
                if (c > S->cap_axis || c < 1) fits = false;
                lo_s[d] = lo[d]; stop_s[d] = stop;
            }
            unsigned long long nodes = (unsigned long long)cnt[0] * cnt[1] * cnt[2];
            if (nodes > S->cap_nodes) fits = false;
            uint32_t nb[3] = {1, 1, 1};
            for (int d = 0; d < D; ++d) nb[d] = (cnt[d] + BlockEdge<D>::v - 1) / BlockEdge<D>::v;
            const unsigned long long nblocks = (unsigned long long)nb[0] * nb[1] * nb[2];
            if (nblocks > S->cap_blocks) fits = false;
            bool finite = true;
            for (int d = 0; d < D; ++d) if (!(lo[d] == lo[d]) || !(hi[d] == hi[d]) || !isfinite(lo[d]) || !isfinite(hi[d]) || cnt[d] < 1) finite = false;
            if (!(dt == dt)) finite = false;
            if (!finite) {
                S->err |= MPM_ERR_NONFINITE; S->stop = 1;
            } else if (!fits) {
                S->need_realloc = 1; S->stop = 1;
                for (int d = 0; d < D; ++d) S->n[d] = cnt[d];
                S->nblocks = (uint32_t)nblocks;
            } else {
                S->dt = dt; S->warn |= warn;
                for (int d = 0; d < 3; ++d) S->gravity[d] = grav[d];
                for (int d = 0; d < D; ++d) {
                    S->min_init[d] = mni[d]; S->max_init[d] = mxi[d]; S->low_init[d] = loi[d]; S->high_init[d] = hii[d];
                    S->origin[d] = lo[d]; S->n[d] = cnt[d];
                }
                if (D == 2) { S->n[2] = 1; S->origin[2] = 0.0; }
                S->grid_nodes = (uint32_t)nodes;
                for (int d = 0; d < 3; ++d) S->nb[d] = nb[d];
                S->nblocks = (uint32_t)nblocks;
                // reset the reductions that this step's particle kernels fill
                S->vmax2_bits = 0ull;
                for (int d = 0; d < 3; ++d) { S->xmin_key[d] = 0x7FFFFFFFFFFFFFFFLL; S->xmax_key[d] = (-0x7FFFFFFFFFFFFFFFLL - 1); }
                S->plastic_count = 0;
                S->cons_grid_mass = 0; for (int d = 0; d < 3; ++d) { S->cons_grid_mom[d] = 0; S->cons_part_mom[d] = 0; }
                go = 1;
            }
        }
    }
    __syncthreads();
    if (go && threadIdx.x < (unsigned)D) {
        int d = threadIdx.x;
        double dx = P.dx, stop = stop_s[d];
        uint32_t i = 0;
        for (double val = lo_s[d]; val < stop; val += dx) g.coord[d][i++] = val;
    }
}

// ----------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_zero_grid(Params P, const StepState* S, GridPtrs g, SortBufs sb) {
    if (S->stop) return;
    const uint32_t nn = S->grid_nodes;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        g.mass[i] = 0.0;
        for (int d = 0; d < D; ++d) { g.mv[d][i] = 0.0; g.f[d][i] = 0.0; }
        if (P.use_mibf) g.fric[i] = 0.0;
        if (P.use_musl) for (int d = 0; d < D; ++d) g.mv2[d][i] = 0.0;
    }
    const uint32_t nb = S->nblocks;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nb; i += gridDim.x * blockDim.x) sb.count[i] = 0;
}

// Block key of a particle: the P2G window base (clamped at 0) divided by the block edge, clamped into
// the block grid (particles whose window leaves the grid are reported by the scatter/gather kernels).
template <int D>
__device__ __forceinline__ uint32_t block_key(const double* xp, const StepState* S, double h1) {
    uint32_t c[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) {
        int v = (int)floor((xp[d] - S->origin[d]) * h1) - 1;
        if (v < 0) v = 0;
        uint32_t b = (uint32_t)v / BlockEdge<D>::v;
        c[d] = b < S->nb[d] ? b : S->nb[d] - 1;
    }
    return D == 3 ? (c[0] * S->nb[1] + c[1]) * S->nb[2] + c[2] : c[0] * S->nb[1] + c[1];
}

// k_bin_count: block key and histogram (wavefront-aggregated atomics).
template <int D>
__global__ __launch_bounds__(kBlock) void k_bin_count(Params P, const StepState* S, PartState src, SortBufs sb, uint32_t n) {
    if (S->stop) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = p < n;
    uint32_t key = 0;
    if (active) {
        double xp[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) xp[d] = src.x[d][p];
        key = block_key<D>(xp, S, P.one_over_dx);
        sb.key[p] = key;
    }
    (void)wave_claim(sb.count, key, active);
}

// k_bin_scan: exclusive prefix sum of the block histogram; one workgroup of 1024 threads, each owning
// a contiguous run of blocks.
__global__ __launch_bounds__(1024) void k_bin_scan(const StepState* S, SortBufs sb) {
    if (S->stop) return;
    __shared__ uint32_t part[1024];
    const uint32_t nb = S->nblocks;
    const uint32_t per = (nb + 1023) / 1024;
    const uint32_t b0 = threadIdx.x * per, b1 = min(nb, b0 + per);
    uint32_t s = 0;
    for (uint32_t b = b0; b < b1; ++b) s += sb.count[b];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int o = 1; o < 1024; o <<= 1) {  // inclusive Hillis-Steele scan of the run sums
        uint32_t t = threadIdx.x >= (unsigned)o ? part[threadIdx.x - o] : 0u;
        __syncthreads();
        part[threadIdx.x] += t;
        __syncthreads();
    }
    uint32_t run = threadIdx.x ? part[threadIdx.x - 1] : 0u;
    for (uint32_t b = b0; b < b1; ++b) { sb.offset[b] = run; sb.cursor[b] = run; run += sb.count[b]; }
}

// k_bin_scatter: counting-sort permutation; also carries the particle ids into the new layout.
template <int D>
__global__ __launch_bounds__(kBlock) void k_bin_scatter(const StepState* S, PartState src, PartState dst, SortBufs sb, uint32_t n) {
    if (S->stop) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = p < n;
    const uint32_t key = active ? sb.key[p] : 0u;
    const uint32_t pos = wave_claim(sb.cursor, key, active);
    if (active) { sb.perm[pos] = p; dst.id[pos] = src.id[p]; }
}

// k_binning: cell base index and block key of every particle of a layout (diagnostic download).
template <int D>
__global__ __launch_bounds__(kBlock) void k_binning(Params P, const StepState* S, PartState ps, uint32_t n, int32_t* base_out,
                                                     uint32_t* key_out) {
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= n) return;
    double xp[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) xp[d] = ps.x[d][p];
    for (int d = 0; d < D; ++d) base_out[(size_t)p * D + d] = (int)floor((xp[d] - S->origin[d]) * P.one_over_dx) - 1;
    key_out[p] = block_key<D>(xp, S, P.one_over_dx);
}

// ----------------------------------------------------------------------------------------------
// k_p2g: p2g.cpp:12-66 (mass, APIC momentum, MIBF friction; weight > 1e-25) fused with the stress
// part of explicit_euler_update.cpp:15-168 (tau grad w). The reference adds the force only where
// m_i > 0; nodes with m_i = 0 are never updated or read, so the force is scattered unconditionally.
// One workgroup = 256 consecutive sorted particles, accumulated in an LDS tile over their windows'
// bounding box and flushed with global atomics; a chunk whose box exceeds the tile uses global atomics.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL>
__global__ __launch_bounds__(kBlock) void k_p2g(Params P, StepState* S, PartState src, SortBufs sb, GridPtrs g, uint32_t n) {
    if (S->stop) return;
    __shared__ double tile[8 * kTileCap];
    __shared__ int s_box[6];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < n;
    const double h1 = P.one_over_dx;
    const double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
    const uint32_t nn[3] = {S->n[0], S->n[1], S->n[2]};
    double xp[3] = {0, 0, 0}, vp[3] = {0, 0, 0};
    Stencil<D, SPL> st;
    bool ok = false;
    uint32_t p = 0;
    if (active) {
        p = sb.perm[i];
        for (int d = 0; d < D; ++d) { xp[d] = src.x[d][p]; vp[d] = src.v[d][p]; }
        ok = make_stencil<D, SPL>(st, xp, org, nn, g.coord, h1, true);
        if (!ok) atomicOr(&S->err, MPM_ERR_GRID_BOUNDS);
    }
    const Tile t = tile_bbox<D, Stencil<D, SPL>::cnt>(ok, st.base, s_box);
    const int nv = P.use_mibf ? 8 : 7;
    if (t.fits) {
        for (int q = threadIdx.x; q < nv * kTileCap; q += blockDim.x) tile[q] = 0.0;
        __syncthreads();
    }
    if (ok) {
        const bool apic = P.flip_ratio < 0;
        double Bm[9];
        if (apic) for (int k = 0; k < D * D; ++k) Bm[k] = src.B[k][p];
        double t6[6]; for (int k = 0; k < SymPack<D>::n; ++k) t6[k] = src.tau[k][p];
        double tau[3][3];
        if (D == 3) { tau[0][0] = t6[0]; tau[1][1] = t6[1]; tau[2][2] = t6[2]; tau[0][1] = tau[1][0] = t6[3]; tau[0][2] = tau[2][0] = t6[4]; tau[1][2] = tau[2][1] = t6[5]; }
        else { tau[0][0] = t6[0]; tau[1][1] = t6[1]; tau[0][1] = tau[1][0] = t6[2]; }
        const double muI = P.use_mibf ? src.muI[p] : 0.0;
        const double aD = P.apicDinverse;
        const int c2 = (D == 3) ? Stencil<D, SPL>::cnt : 1;
#pragma unroll
        for (int ia = 0; ia < Stencil<D, SPL>::cnt; ++ia) {
            const int a = ia;
#pragma unroll
            for (int ib = 0; ib < Stencil<D, SPL>::cnt; ++ib) {
                const int b = ib;
#pragma unroll
                for (int ic = 0; ic < c2; ++ic) {
                    const int c = (D == 3) ? ic : 0;
                    double weight = st.w[0][a] * st.w[1][b];
                    if (D == 3) weight = weight * st.w[2][c];
                    const int gi = st.base[0] + a, gj = st.base[1] + b, gk = (D == 3) ? st.base[2] + c : 0;
                    double val[8];
                    bool massive = weight > 1e-25;
                    val[0] = weight;
                    if (massive) {
                        const double pd[3] = {st.xi[0][a] - xp[0], st.xi[1][b] - xp[1], D == 3 ? st.xi[2][c] - xp[2] : 0.0};
                        for (int r = 0; r < D; ++r) {
                            double m = vp[r] * weight;
                            if (apic) {
                                double bd = Bm[r] * pd[0];
                                for (int k = 1; k < D; ++k) bd += Bm[r + k * D] * pd[k];
                                m += bd * aD * weight;
                            }
                            val[1 + r] = m;
                        }
                        val[7] = muI * weight;
                    }
                    double gr[3];
                    if (D == 3) {
                        gr[0] = st.dw[0][a] * st.w[1][b] * st.w[2][c] * h1;
                        gr[1] = st.dw[1][b] * st.w[0][a] * st.w[2][c] * h1;
                        gr[2] = st.dw[2][c] * st.w[0][a] * st.w[1][b] * h1;
                    } else {
                        gr[0] = st.dw[0][a] * st.w[1][b] * h1;
                        gr[1] = st.dw[1][b] * st.w[0][a] * h1;
                    }
                    for (int r = 0; r < D; ++r) {
                        double fr = tau[r][0] * gr[0];
                        for (int k = 1; k < D; ++k) fr += tau[r][k] * gr[k];
                        val[4 + r] = fr;
                    }
                    if (t.fits) {
                        const int q = tile_local<D>(t, gi, gj, gk);
                        if (massive) {
                            atomicAdd(&tile[q], val[0]);
                            for (int r = 0; r < D; ++r) atomicAdd(&tile[(1 + r) * kTileCap + q], val[1 + r]);
                            if (P.use_mibf) atomicAdd(&tile[7 * kTileCap + q], val[7]);
                        }
                        for (int r = 0; r < D; ++r) atomicAdd(&tile[(4 + r) * kTileCap + q], val[4 + r]);
                    } else {
                        const uint32_t node = node_index<D>(gi, gj, gk, nn);
                        if (massive) {
                            atomic_add_f64(&g.mass[node], val[0]);
                            for (int r = 0; r < D; ++r) atomic_add_f64(&g.mv[r][node], val[1 + r]);
                            if (P.use_mibf) atomic_add_f64(&g.fric[node], val[7]);
                        }
                        for (int r = 0; r < D; ++r) atomic_add_f64(&g.f[r][node], val[4 + r]);
                    }
                }
            }
        }
        if (S->check_conservation)
            for (int d = 0; d < D; ++d) atomic_add_f64(&S->cons_part_mom[d], P.particle_mass * vp[d]);
    }
    if (t.fits) {
        __syncthreads();
        double* dst[8] = {g.mass, g.mv[0], g.mv[1], g.mv[2], g.f[0], g.f[1], g.f[2], g.fric};
        if (D == 2) { dst[3] = nullptr; dst[6] = nullptr; }
        // tile rows: 0 mass, 1..3 momentum, 4..6 force, 7 friction (unused rows stay zero)
        for (int q = threadIdx.x; q < t.vol; q += blockDim.x) {
            int ii, jj, kk;
            if (D == 3) { kk = q % t.ext[2]; int r = q / t.ext[2]; jj = r % t.ext[1]; ii = r / t.ext[1]; }
            else { jj = q % t.ext[1]; ii = q / t.ext[1]; kk = 0; }
            const uint32_t node = node_index<D>(t.lo[0] + ii, t.lo[1] + jj, t.lo[2] + kk, nn);
            for (int v = 0; v < 8; ++v) {
                if (!dst[v] || (v == 7 && !P.use_mibf)) continue;
                const double a = tile[v * kTileCap + q];
                if (a != 0.0) atomic_add_f64(dst[v] + node, a);
            }
        }
    }
}

// ----------------------------------------------------------------------------------------------
// k_grid_update: normalisation (p2g.cpp:84-98), explicit update and boundary conditions
// (explicit_euler_update.cpp:183-232, boundary_collision.cpp:5-394), flip = v_new - v_old.
// ----------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_grid_update(Params P, StepState* S, GridPtrs g) {
    if (S->stop) return;
    const uint32_t nn = S->grid_nodes;
    const uint32_t n[3] = {S->n[0], S->n[1], S->n[2]};
    const double dt = S->dt;
    const double dtV = dt * P.particle_volume;
    double dtg[3]; for (int d = 0; d < 3; ++d) dtg[d] = dt * S->gravity[d];
    const bool cons = S->check_conservation;
    double cm = 0.0, cmv[3] = {0, 0, 0};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        const double sw = g.mass[i];
        V<D> v;
        double fr = 0.0;
        if (sw > 0) { for (int d = 0; d < D; ++d) v[d] = g.mv[d][i] / sw; }
        else { for (int d = 0; d < D; ++d) v[d] = 0.0; }
        if (P.use_mibf) fr = g.fric[i] / sw;
        const double mi = sw * P.particle_mass;
        if (cons) { cm += mi; for (int d = 0; d < D; ++d) cmv[d] += v[d] * mi; }
        V<D> vn, flip;
        if (mi > 0) {
            for (int d = 0; d < D; ++d) vn[d] = v[d] + (((-dtV) * g.f[d][i]) / mi + dtg[d]);
            V<D> X;
            if (D == 3) { uint32_t ii = i / (n[1] * n[2]); uint32_t jj = (i / n[2]) % n[1]; uint32_t kk = i % n[2]; X[0] = g.coord[0][ii]; X[1] = g.coord[1][jj]; X[D - 1] = g.coord[2][kk]; }
            else { uint32_t ii = i / n[1]; uint32_t jj = i % n[1]; X[0] = g.coord[0][ii]; X[1] = g.coord[1][jj]; }
            boundary_collision<D>(P, S->plates, dt, fr, X, vn);
            for (int d = 0; d < D; ++d) flip[d] = vn[d] - v[d];
        } else {
            vn = v;
            for (int d = 0; d < D; ++d) flip[d] = 0.0;
        }
        g.mass[i] = mi;
        for (int d = 0; d < D; ++d) { g.v[d][i] = vn[d]; g.flip[d][i] = flip[d]; }
        if (P.use_mibf) g.fric[i] = fr;
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

// Velocity update of positionUpdate / MUSL (position_update.cpp:14-23, musl.cpp:13-25).
__device__ __forceinline__ double blend_velocity(double fr, double v, double flip, double pic) {
    if (fr < -1) return pic;
    if (fr < 0) return (-fr) * (v + flip) + (1 - (-fr)) * pic;
    return fr * (v + flip) + (1 - fr) * pic;
}

// F <- F + (dt*grad v) * F, return map, Kirchhoff stress for the next step (deformation_update.cpp:13-46,
// plasticity.cpp:11-665). Reads particle p of the source layout, writes slot i of the destination layout.
template <int D>
__device__ __forceinline__ void deform_and_project(const Params& P, double dt, const PartState& src, uint32_t p,
                                                   const PartState& dst, uint32_t i, const double (&gv)[3][3],
                                                   uint32_t& err, uint32_t& pcount) {
    M<D> F, A;
    for (int k = 0; k < D * D; ++k) F.m[k] = src.F[k][p];
    for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) A(r, c) = dt * gv[r][c];
    M<D> AF = mul<D>(A, F);
    for (int k = 0; k < D * D; ++k) F.m[k] = F.m[k] + AF.m[k];
    PlasticState ps;
    ps.eps_dev = src.eps_dev[p]; ps.eps_vol = src.eps_vol[p]; ps.eps_prad = src.eps_prad[p];
    ps.dgamma = 0.0; ps.visc = src.visc[p]; ps.muI = src.muI[p];
    double tau[6];
    if (P.elastic_model == MPM_ELASTIC_NEOHOOKEAN) {
        neohookean_tau<D>(F, P.mu, P.lambda, tau);
    } else {
        M<D> U, W; V<D> s;
        svd<D>(F, U, s, W);
        V<D> ls;
        if (P.plastic_model != MPM_PLASTIC_NONE) {
            V<D> h; for (int k = 0; k < D; ++k) h[k] = log(fmax(fabs(s[k]), 1e-4));
            RMResult r = return_map<D>(P, dt, h, ps);
            err |= r.err;
            pcount += r.plastic_incr;
            if (r.recompose) {
                V<D> e; for (int k = 0; k < D; ++k) e[k] = exp(h[k]);
                F = udvt<D>(U, e, W);
                for (int k = 0; k < D; ++k) ls[k] = log(e[k]);
            } else {
                for (int k = 0; k < D; ++k) ls[k] = log(fabs(s[k]));
            }
        } else {
            for (int k = 0; k < D; ++k) ls[k] = log(fabs(s[k]));
        }
        hencky_tau<D>(U, ls, P.mu, P.lambda, tau);
    }
    for (int k = 0; k < D * D; ++k) dst.F[k][i] = F.m[k];
    for (int k = 0; k < SymPack<D>::n; ++k) dst.tau[k][i] = tau[k];
    dst.eps_dev[i] = ps.eps_dev; dst.eps_vol[i] = ps.eps_vol; dst.eps_prad[i] = ps.eps_prad;
    dst.dgamma[i] = ps.dgamma; dst.visc[i] = ps.visc; dst.muI[i] = ps.muI;
}

// Gather of G2P (g2p.cpp:35-67) and of the velocity gradient (deformation_update.cpp:217-230) over the
// non-zero nodes of the window, in the reference's (i,j,k) order.
template <int D, int SPL, bool PIC, bool GRAD>
__device__ __forceinline__ void gather(const Stencil<D, SPL>& st, const double* xp, const GridPtrs& g, const uint32_t* n,
                                       double h1, bool apic, bool useflip, double* pic, double* fl, double (&Bp)[3][3],
                                       double (&gv)[3][3]) {
    const int c2 = (D == 3) ? Stencil<D, SPL>::cnt : 1;
#pragma unroll
    for (int ia = 0; ia < Stencil<D, SPL>::cnt; ++ia) {
        const int a = ia;
#pragma unroll
        for (int ib = 0; ib < Stencil<D, SPL>::cnt; ++ib) {
            const int b = ib;
#pragma unroll
            for (int ic = 0; ic < c2; ++ic) {
                const int c = (D == 3) ? ic : 0;
                const uint32_t node = node_index<D>(st.base[0] + a, st.base[1] + b, (D == 3) ? st.base[2] + c : 0, n);
                double vi[3];
                for (int d = 0; d < D; ++d) vi[d] = g.v[d][node];
                if (PIC) {
                    double weight = st.w[0][a] * st.w[1][b];
                    if (D == 3) weight = weight * st.w[2][c];
                    const double pd[3] = {st.xi[0][a] - xp[0], st.xi[1][b] - xp[1], D == 3 ? st.xi[2][c] - xp[2] : 0.0};
                    for (int r = 0; r < D; ++r) {
                        pic[r] += vi[r] * weight;
                        if (apic) for (int k = 0; k < D; ++k) Bp[r][k] += vi[r] * pd[k] * weight;
                    }
                    if (useflip) for (int r = 0; r < D; ++r) fl[r] += g.flip[r][node] * weight;
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
                    for (int r = 0; r < D; ++r) for (int k = 0; k < D; ++k) gv[r][k] += vi[r] * gr[k];
                }
            }
        }
    }
}

// ----------------------------------------------------------------------------------------------
// k_g2p_fused (USL): G2P (g2p.cpp:17-77), deformationUpdate + plasticity, positionUpdate.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL>
__global__ __launch_bounds__(kBlock) void k_g2p_fused(Params P, StepState* S, PartState src, PartState dst, SortBufs sb,
                                                      GridPtrs g, uint32_t n) {
    if (S->stop) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t err = 0, pcount = 0;
    double xnew[3] = {0, 0, 0}, vnew[3] = {0, 0, 0};
    bool live = false;
    if (i < n) {
        const uint32_t p = sb.perm[i];
        const double h1 = P.one_over_dx, dt = S->dt;
        const double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
        const uint32_t nn[3] = {S->n[0], S->n[1], S->n[2]};
        double xp[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) xp[d] = src.x[d][p];
        Stencil<D, SPL> st;
        if (!make_stencil<D, SPL>(st, xp, org, nn, g.coord, h1, false)) {
            err |= MPM_ERR_GRID_BOUNDS;
        } else {
            live = true;
            const bool apic = P.flip_ratio < 0, useflip = P.flip_ratio >= -1;
            double pic[3] = {0, 0, 0}, fl[3] = {0, 0, 0}, Bp[3][3] = {{0}}, gv[3][3] = {{0}};
            gather<D, SPL, true, true>(st, xp, g, nn, h1, apic, useflip, pic, fl, Bp, gv);
            for (int d = 0; d < D; ++d) {
                const double vo = useflip ? src.v[d][p] : 0.0;
                vnew[d] = blend_velocity(P.flip_ratio, vo, fl[d], pic[d]);
            }
            for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) dst.B[r + c * D][i] = apic ? Bp[r][c] : 0.0;
            deform_and_project<D>(P, dt, src, p, dst, i, gv, err, pcount);
            for (int d = 0; d < D; ++d) xnew[d] = xp[d] + dt * pic[d];
            if (P.change_particle_positions && xnew[0] > 0.8) { xnew[0] = -0.1; xnew[1] += 0.27; }
            for (int d = 0; d < D; ++d) { dst.x[d][i] = xnew[d]; dst.v[d][i] = vnew[d]; }
        }
    }
    reduce_step_bounds<D>(S, live, xnew, vnew);
    pcount = wave_sum_u32(pcount);
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0) {
        add_plastic_count(S, pcount);
        if (err) atomicOr(&S->err, err);
    }
}

// ----------------------------------------------------------------------------------------------
// k_g2p_musl: G2P (g2p.cpp:17-77), MUSL velocity update and re-scatter (musl.cpp:13-85), position
// update of positionUpdate (position_update.cpp:12). The re-scatter goes through an LDS tile like P2G.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL>
__global__ __launch_bounds__(kBlock) void k_g2p_musl(Params P, StepState* S, PartState src, PartState dst, SortBufs sb,
                                                     GridPtrs g, uint32_t n) {
    if (S->stop) return;
    __shared__ double tile[3 * kTileCap];
    __shared__ int s_box[6];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t err = 0;
    double xnew[3] = {0, 0, 0}, vnew[3] = {0, 0, 0}, xp[3] = {0, 0, 0};
    double Bp[3][3] = {{0}};
    bool live = false;
    const double h1 = P.one_over_dx, dt = S->dt;
    const double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
    const uint32_t nn[3] = {S->n[0], S->n[1], S->n[2]};
    const bool apic = P.flip_ratio < 0, useflip = P.flip_ratio >= -1;
    Stencil<D, SPL> st;
    if (i < n) {
        const uint32_t p = sb.perm[i];
        for (int d = 0; d < D; ++d) xp[d] = src.x[d][p];
        if (!make_stencil<D, SPL>(st, xp, org, nn, g.coord, h1, false)) {
            err |= MPM_ERR_GRID_BOUNDS;
        } else {
            live = true;
            double pic[3] = {0, 0, 0}, fl[3] = {0, 0, 0}, gv[3][3];
            gather<D, SPL, true, false>(st, xp, g, nn, h1, apic, useflip, pic, fl, Bp, gv);
            for (int d = 0; d < D; ++d) {
                const double vo = useflip ? src.v[d][p] : 0.0;
                vnew[d] = blend_velocity(P.flip_ratio, vo, fl[d], pic[d]);
                xnew[d] = xp[d] + dt * pic[d];
            }
            if (P.change_particle_positions && xnew[0] > 0.8) { xnew[0] = -0.1; xnew[1] += 0.27; }
            for (int d = 0; d < D; ++d) { dst.v[d][i] = vnew[d]; dst.x[d][i] = xnew[d]; }
            for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) dst.B[r + c * D][i] = apic ? Bp[r][c] : 0.0;
            // the re-scatter clamps the base at 0 as P2G does (musl.cpp:131-135); it differs from the
            // gather window only if the gather window already left the grid, which is reported above
        }
    }
    const Tile t = tile_bbox<D, Stencil<D, SPL>::cnt>(live, st.base, s_box);
    if (t.fits) {
        for (int q = threadIdx.x; q < 3 * kTileCap; q += blockDim.x) tile[q] = 0.0;
        __syncthreads();
    }
    if (live) {
        const double aD = P.apicDinverse;
        const int c2 = (D == 3) ? Stencil<D, SPL>::cnt : 1;
#pragma unroll
        for (int ia = 0; ia < Stencil<D, SPL>::cnt; ++ia) {
            const int a = ia;
#pragma unroll
            for (int ib = 0; ib < Stencil<D, SPL>::cnt; ++ib) {
                const int b = ib;
#pragma unroll
                for (int ic = 0; ic < c2; ++ic) {
                    const int c = (D == 3) ? ic : 0;
                    double weight = st.w[0][a] * st.w[1][b];
                    if (D == 3) weight = weight * st.w[2][c];
                    if (!(weight > 1e-25)) continue;
                    const int gi = st.base[0] + a, gj = st.base[1] + b, gk = (D == 3) ? st.base[2] + c : 0;
                    const double pd[3] = {st.xi[0][a] - xp[0], st.xi[1][b] - xp[1], D == 3 ? st.xi[2][c] - xp[2] : 0.0};
                    double m[3];
                    for (int r = 0; r < D; ++r) {
                        m[r] = vnew[r] * weight;
                        if (apic) {
                            double bd = Bp[r][0] * pd[0];
                            for (int k = 1; k < D; ++k) bd += Bp[r][k] * pd[k];
                            m[r] += bd * aD * weight;
                        }
                    }
                    if (t.fits) {
                        const int q = tile_local<D>(t, gi, gj, gk);
                        for (int r = 0; r < D; ++r) atomicAdd(&tile[r * kTileCap + q], m[r]);
                    } else {
                        const uint32_t node = node_index<D>(gi, gj, gk, nn);
                        for (int r = 0; r < D; ++r) atomic_add_f64(&g.mv2[r][node], m[r]);
                    }
                }
            }
        }
    }
    if (t.fits) {
        __syncthreads();
        tile_flush<D, D>(t, tile, g.mv2, nn);
    }
    reduce_step_bounds<D>(S, live, xnew, vnew);
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0 && err) atomicOr(&S->err, err);
}

// MUSL normalisation, musl.cpp:182-188: v_i = (sum) / (m_i / m_p), zero where m_i == 0.
template <int D>
__global__ __launch_bounds__(kBlock) void k_musl_norm(Params P, const StepState* S, GridPtrs g) {
    if (S->stop) return;
    const uint32_t nn = S->grid_nodes;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        const double mi = g.mass[i];
        if (mi > 0) { const double q = mi / P.particle_mass; for (int d = 0; d < D; ++d) g.v[d][i] = g.mv2[d][i] / q; }
        else for (int d = 0; d < D; ++d) g.v[d][i] = 0.0;
    }
}

// k_defgrad (MUSL): deformationUpdate (deformation_update.cpp:16-46) on the re-scattered velocity.
template <int D, int SPL>
__global__ __launch_bounds__(kBlock) void k_defgrad(Params P, StepState* S, PartState src, PartState dst, SortBufs sb,
                                                    GridPtrs g, uint32_t n) {
    if (S->stop) return;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t err = 0, pcount = 0;
    if (i < n) {
        const uint32_t p = sb.perm[i];
        const double h1 = P.one_over_dx, dt = S->dt;
        const double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
        const uint32_t nn[3] = {S->n[0], S->n[1], S->n[2]};
        double xp[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) xp[d] = src.x[d][p];
        Stencil<D, SPL> st;
        if (!make_stencil<D, SPL>(st, xp, org, nn, g.coord, h1, false)) {
            err |= MPM_ERR_GRID_BOUNDS;
        } else {
            double gv[3][3] = {{0}}, pic[3], fl[3], Bp[3][3];
            gather<D, SPL, false, true>(st, xp, g, nn, h1, false, false, pic, fl, Bp, gv);
            deform_and_project<D>(P, dt, src, p, dst, i, gv, err, pcount);
        }
    }
    pcount = wave_sum_u32(pcount);
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0) {
        add_plastic_count(S, pcount);
        if (err) atomicOr(&S->err, err);
    }
}

// ----------------------------------------------------------------------------------------------
// k_finish: moveObjects (simulation.cpp:277-281) and the loop bookkeeping of simulate()
// (simulation.cpp:176-200), conservation checks (simulation.cpp:292-333) when requested.
// ----------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(64) void k_finish(Params P, StepState* S, uint32_t n_particles) {
    if (S->stop) return;
    uint32_t pc = 0;
    for (int s = threadIdx.x; s < kRedSlots; s += 64) { pc += S->red[s].plastic; S->red[s].plastic = 0; }
    pc = wave_sum_u32(pc);
    if (threadIdx.x != 0) return;
    S->plastic_count += pc;
    const double dt = S->dt;
    for (int k = 0; k < P.n_plates; ++k) plate_move<D>(S->plates[k], dt, P.frame_dt, S->time);
    if (S->check_conservation) {
        double pm = P.particle_mass * n_particles;
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
}

}  // namespace mpm
