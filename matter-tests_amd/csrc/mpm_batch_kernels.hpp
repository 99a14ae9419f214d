// mpm_batch_kernels.hpp — batched launches: many independent scenes advanced by one launch per kernel.
//
// The scenes of a batch that share dimension, spline degree, transfer type and the compile-time switches
// of the step (separate force, MIBF, material, MUSL, periodic boundaries) form a class. Each kernel of the
// step is launched once per class over the concatenated workgroup ranges of its scenes: workgroup w belongs
// to the scene s with
// start[s] <= w < start[s + 1] (a binary search over the class's range table, uniform across the
// workgroup), and runs the same body as the single-scene kernel on that scene's constants, loop state and
// buffers with its local workgroup index. Every scene keeps its own StepState, so dt, grid geometry,
// stop conditions and step budget stay per scene, and a scene's results do not depend on the other scenes
// of the launch.
#pragma once
#include "mpm_kernels.hpp"

namespace mpm {

// One scene of a batched launch. Both particle layouts are listed; the current one is StepState::xsel.
template <class TF> struct SceneRef {
    const Params* P;
    StepState* S;
    PartStateT<TF> ps[2];
    SortBufs sb;
    GridPtrsT<TF> g;
    StagingT<TF> stg;
    uint32_t n;
    uint32_t cap_axis, cap_nodes, cap_blocks;
};

// Workgroup ranges of a class: particle kernels (one workgroup per 256 particles), block kernels (one per
// block of the scene's capacity), node kernels (grid-stride), histogram clear (grid-stride over blocks).
enum BatchRangeKind { BR_PART, BR_BLOCK, BR_GRID, BR_ZERO, BR_KINDS };

struct BatchTable {
    const void* refs;         // SceneRef<TF>[ns]
    const uint32_t* start;    // BR_KINDS rows of ns + 1 prefix sums
    const uint32_t* limit;    // step budget of each scene for the current chunk of steps
    uint32_t ns;
};

// The scene tables and constants are read-only while a launch runs. Reading them through the constant
// address space lets the compiler use scalar loads for them anywhere in a kernel (as it does for kernel
// arguments); through generic pointers it cannot prove that no store of the kernel changed them.
#define MPM_CONST __attribute__((address_space(4)))
template <class T> __device__ __forceinline__ const T& as_const_space(const T* p) {
    return *(const T*)(const MPM_CONST T*)p;
}

// Scene of workgroup w in a range row (start[0] = 0, start[ns] = workgroups of the launch).
__device__ __forceinline__ uint32_t find_scene(const uint32_t* start, uint32_t ns, uint32_t w) {
    uint32_t lo = 0, hi = ns;
    while (hi - lo > 1) {
        const uint32_t mid = (lo + hi) >> 1;
        if (start[mid] <= w) lo = mid; else hi = mid;
    }
    return lo;
}

template <class TF> struct BatchSlot {
    const SceneRef<TF>* r;
    uint32_t local, count, xs;
    __device__ __forceinline__ BatchSlot(const BatchTable& T, int kind) {
        const uint32_t* st = T.start + (size_t)kind * (T.ns + 1);
        const uint32_t s = find_scene(st, T.ns, blockIdx.x);
        r = &as_const_space(static_cast<const SceneRef<TF>*>(T.refs) + s);
        local = blockIdx.x - st[s];
        count = st[s + 1] - st[s];
        xs = __builtin_amdgcn_readfirstlane(r->S->xsel) & 1u;
    }
    __device__ __forceinline__ const Params& P() const { return as_const_space(r->P); }
    __device__ __forceinline__ const PartStateT<TF>& src() const { return r->ps[xs]; }
    __device__ __forceinline__ const PartStateT<TF>& dst() const { return r->ps[xs ^ 1u]; }
};

// Start of a chunk of steps (the batched k_control): one workgroup per scene.
template <class TF>
__global__ __launch_bounds__(64) void kb_control(BatchTable T, uint32_t stop_at_frame_end, uint32_t check_cons) {
    if (threadIdx.x != 0) return;
    const SceneRef<TF>& r = static_cast<const SceneRef<TF>*>(T.refs)[blockIdx.x];
    StepState* S = r.S;
    const uint32_t lim = T.limit[blockIdx.x];
    S->stop = lim == 0 ? 1u : 0u; S->need_realloc = 0; S->steps_done = 0; S->err = 0; S->warn = 0; S->frame_ended = 0;
    S->stop_at_frame_end = stop_at_frame_end; S->check_conservation = check_cons; S->step_limit = lim;
    S->cap_axis = r.cap_axis; S->cap_nodes = r.cap_nodes; S->cap_blocks = r.cap_blocks;
}

template <int D, class TF>
__global__ __launch_bounds__(64) void kb_setup(BatchTable T) {
    const SceneRef<TF>& r = as_const_space(static_cast<const SceneRef<TF>*>(T.refs) + blockIdx.x);
    setup_body<D, TF>(as_const_space(r.P), r.S, r.g);
}

template <class TF>
__global__ __launch_bounds__(kBlock) void kb_zero_grid(BatchTable T) {
    BatchSlot<TF> q(T, BR_ZERO);
    zero_grid_body(q.r->S, q.r->sb, q.local, q.count);
}

template <int D, int SPL, class TF>
__global__ __launch_bounds__(kBlock) void kb_bin_count(BatchTable T) {
    BatchSlot<TF> q(T, BR_PART);
    bin_count_body<D, SPL, TF>(q.P(), q.r->S, q.src(), q.r->sb, q.r->g, q.r->n, q.local);
}

template <class TF>
__global__ __launch_bounds__(1024) void kb_bin_scan(BatchTable T) {
    const SceneRef<TF>& r = static_cast<const SceneRef<TF>*>(T.refs)[blockIdx.x];
    bin_scan_body(r.S, r.sb);
}

template <class TF>
__global__ __launch_bounds__(kBlock) void kb_bin_scatter(BatchTable T) {
    BatchSlot<TF> q(T, BR_PART);
    bin_scatter_body(q.r->S, q.r->sb, q.r->n, q.local);
}

template <int D, int SPL, bool SEP, bool MIBF, class TF, bool PBC>
__global__ __launch_bounds__((ScatterLaunch<D, SPL, P2GContrib<D, SPL, SEP, MIBF, TF>::NV, TF, PBC>::TB)) void kb_p2g(BatchTable T) {
    BatchSlot<TF> q(T, BR_BLOCK);
    p2g_body<D, SPL, SEP, MIBF, TF, PBC>(q.P(), q.r->S, q.src(), q.dst(), q.r->sb, q.r->g, q.r->stg, q.local);
}

template <int D, int SPL, bool SEP, bool MIBF, class TF>
__global__ __launch_bounds__(kBlock) void kb_grid_update(BatchTable T) {
    BatchSlot<TF> q(T, BR_GRID);
    grid_update_body<D, SPL, SEP, MIBF, TF>(q.P(), q.r->S, q.r->sb, q.r->g, q.r->stg, q.local, q.count);
}

template <int D, int SPL, bool FLIP, class TF>
__global__ __launch_bounds__(kG2PThreads) void kb_g2p(BatchTable T) {
    BatchSlot<TF> q(T, BR_BLOCK);
    g2p_body<D, SPL, FLIP, TF>(q.P(), q.r->S, q.src(), q.dst(), q.r->sb, q.r->g, q.local);
}

template <int D, int SPL, class TF>
__global__ __launch_bounds__((ScatterLaunch<D, SPL, D, TF, false>::TB)) void kb_musl_scatter(BatchTable T) {
    BatchSlot<TF> q(T, BR_BLOCK);
    musl_scatter_body<D, SPL, TF>(q.P(), q.r->S, q.src(), q.dst(), q.r->sb, q.r->g, q.r->stg, q.local);
}

template <int D, int SPL, class TF>
__global__ __launch_bounds__(kBlock) void kb_musl_update(BatchTable T) {
    BatchSlot<TF> q(T, BR_GRID);
    musl_update_body<D, SPL, TF>(q.P(), q.r->S, q.r->sb, q.r->g, q.r->stg, q.local, q.count);
}

template <int D, int SPL, class TF, int MAT>
__global__ __launch_bounds__(kBlock) void kb_defgrad(BatchTable T) {
    BatchSlot<TF> q(T, BR_PART);
    defgrad_body<D, SPL, TF, MAT>(q.P(), q.r->S, q.src(), q.dst(), q.r->sb, q.r->g, q.r->n, q.local);
}

template <int D, class TF>
__global__ __launch_bounds__(64) void kb_finish(BatchTable T) {
    const SceneRef<TF>& r = as_const_space(static_cast<const SceneRef<TF>*>(T.refs) + blockIdx.x);
    finish_body<D>(as_const_space(r.P), r.S, r.n);
}

// Copies the loop-state head of every scene's StepState (everything before the reduction slots) into one
// contiguous buffer, so that the host reads the whole batch with one transfer.
template <int = 0>  // a template, so that every translation unit may instantiate it
__global__ __launch_bounds__(kBlock) void kb_gather_state(StepState* const* S, uint32_t head_words, uint32_t* out) {
    const uint32_t* src = reinterpret_cast<const uint32_t*>(S[blockIdx.x]);
    uint32_t* dst = out + (size_t)blockIdx.x * head_words;
    for (uint32_t i = threadIdx.x; i < head_words; i += blockDim.x) dst[i] = src[i];
}

}  // namespace mpm
