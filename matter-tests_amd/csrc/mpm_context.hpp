// mpm_context.hpp — the engine context (one per mpm_create) and the host helpers shared by the
// translation units of the engine: the typed views of its buffers and per-launch event timing.
#pragma once
#include "mpm_batch_kernels.hpp"
#include <vector>
#include <utility>
#include <mutex>

// Device buffers of the node grid and of the per-block arrays of the step, with their capacities (nodes,
// nodes per axis, blocks). A context owns one store; mpm_step_noalloc binds a store carved from the device's
// static scratch instead.
struct GridStore {
    char* gbuf = nullptr;        // node arrays: mass, friction, v[dim], flip[dim]
    double* cbuf = nullptr;      // node coordinates, cap_axis per axis
    uint32_t* bbuf = nullptr;    // block histogram, offsets, cursors
    uint16_t* cellbuf = nullptr; // per-block dual-cell offsets
    char* stgbuf = nullptr;      // per-block node tiles (staging of the scatter kernels)
    uint64_t cap_nodes = 0;
    uint32_t cap_axis = 0;
    uint32_t cap_blocks = 0;
};

struct mpm_context {
    mpm_config cfg;
    mpm::Params P;
    std::vector<mpm::DevPlate> plates;
    int device = 0;
    hipStream_t stream = nullptr;
    uint32_t n = 0;
    bool ready = false;
    bool split = false;         // MPM_PRECISION_SPLIT: transfer quantities in float
    // particles: two layouts (double buffer) and the counting-sort arrays
    char* pbuf = nullptr;
    size_t pbuf_bytes = 0;
    uint32_t* ibuf = nullptr;   // ids (2 layouts), keys, permutation
    size_t ibuf_n = 0;
    mpm::PartStateT<double> ps[2]{};  // plain form
    mpm::PartStateT<float> pf[2]{};   // split form
    mpm::SortBufs sb{};
    // grid and block arrays: the store the context allocated, and the one its views (sb.count/offset/cursor/
    // cell_off, stg, gp, ...) point into now (own, or the device scratch after mpm_step_noalloc)
    GridStore own{};
    bool grid_bound = false;    // the views point into a store
    bool on_scratch = false;    // ... and that store is the device scratch
    uint64_t grid_gen = 0;      // incremented whenever the views change (batch tables cache them)
    uint32_t cap_blocks = 0;    // blocks of the bound store (sort arrays, staging)
    uint32_t act_blocks = 0;    // blocks the block kernels launch and a step may use (<= cap_blocks)
    uint64_t cap_nodes = 0;     // nodes of the bound store
    uint32_t cap_axis = 0;      // nodes per axis of the bound store's coordinate arrays
    mpm::StagingT<double> stg{};
    mpm::StagingT<float> stgf{};
    bool flip_formed = false;   // the last enqueued step wrote grid flip
    mpm::GridPtrsT<double> gp{};
    mpm::GridPtrsT<float> gf{};
    // step state
    mpm::StepState* S = nullptr;   // device
    mpm::StepState* Sh = nullptr;  // pinned host mirror
    uint32_t est_n[3] = {1, 1, 1};
    // per-kernel HIP-event timing (mpm_set_profiling)
    bool profiling = false;
    uint32_t prof_mask = ~0u;  // kernels (bit KS_*) timed while profiling
    std::vector<hipEvent_t> ev_pool;
    std::vector<std::pair<int, size_t>> ev_pending;  // (kernel id, index of start event)
    double kt_ms[16] = {0};
    uint64_t kt_n[16] = {0};
    // membership of a batch (mpm_batch_add): device copy of P, remaining step budget, result of the last call
    mpm_batch* batch = nullptr;
    mpm::Params* dP = nullptr;
    uint32_t budget = 0;
    mpm_step_info last{};
    // device lattices of the level-set objects (mpm_add_sdf_object)
    std::vector<void*> sdf_bufs;
    // HIP graphs of step batches (mpm_step on small scenes): the launches of `batch` steps captured once and
    // replayed while everything they bake in (launch shapes, buffers, parameters, layout parity) is unchanged
    struct StepGraph {
        uint32_t batch, xpar, flags, act_blocks, n;
        uint64_t grid_gen;
        const void* pbuf;
        const void* ibuf;
        mpm::Params P;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> step_graphs;
    bool step_graphs_on = true;  // mpm_set_step_graphs
};

namespace mpm {

// The typed views of a context's buffers for transfer type TF.
template <class TF> PartStateT<TF>* layouts(mpm_context* c);
template <> inline PartStateT<double>* layouts<double>(mpm_context* c) { return c->ps; }
template <> inline PartStateT<float>* layouts<float>(mpm_context* c) { return c->pf; }
template <class TF> GridPtrsT<TF>& grid_of(mpm_context* c);
template <> inline GridPtrsT<double>& grid_of<double>(mpm_context* c) { return c->gp; }
template <> inline GridPtrsT<float>& grid_of<float>(mpm_context* c) { return c->gf; }
template <class TF> StagingT<TF>& staging_of(mpm_context* c);
template <> inline StagingT<double>& staging_of<double>(mpm_context* c) { return c->stg; }
template <> inline StagingT<float>& staging_of<float>(mpm_context* c) { return c->stgf; }

inline const char* const kKernelNames[] = {"k_setup", "k_zero_grid", "k_p2g", "k_grid_update", "k_g2p",
                                    "k_musl_scatter", "k_musl_update", "k_defgrad", "k_finish", "k_bin_count", "k_bin_scan",
                                    "k_bin_scatter", "k_g2p_musl"};
enum KernelId { KS_SETUP, KS_ZERO, KS_P2G, KS_GRID, KS_G2P, KS_MUSL_SCATTER, KS_MUSL_UPDATE, KS_DEFGRAD,
                KS_FINISH, KS_BIN_COUNT, KS_BIN_SCAN, KS_BIN_SCATTER, KS_G2P_MUSL, KS_COUNT };

// Records an event pair around one launch when profiling; returns the launch's slot.
struct LaunchTimer {
    mpm_context* c; int id; size_t slot = 0; bool on;
    LaunchTimer(mpm_context* c_, int id_) : c(c_), id(id_), on(c_->profiling && ((c_->prof_mask >> id_) & 1u)) {
        if (!on) return;
        slot = c->ev_pending.size() * 2;
        while (c->ev_pool.size() < slot + 2) { hipEvent_t e; if (hipEventCreate(&e) != hipSuccess) { on = false; return; } c->ev_pool.push_back(e); }
        (void)hipEventRecord(c->ev_pool[slot], c->stream);
    }
    ~LaunchTimer() {
        if (!on) return;
        (void)hipEventRecord(c->ev_pool[slot + 1], c->stream);
        c->ev_pending.push_back({id, slot});
    }
};

inline void collect_timings(mpm_context* c) {
    for (auto& pr : c->ev_pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, c->ev_pool[pr.second], c->ev_pool[pr.second + 1]) == hipSuccess) {
            c->kt_ms[pr.first] += ms; c->kt_n[pr.first] += 1;
        }
    }
    c->ev_pending.clear();
}

// One step of a context for a spline degree and transfer type (mpm_step.hpp, instantiated per
// (dimension, degree, type) in its own translation unit so that the engine compiles in parallel).
// head: launch the step's k_setup (else the previous step's k_finish_setup did it); tail: end with k_finish
// (else with k_finish_setup, which also sets up the next step).
template <int D, int SPL, class TF> void enqueue_step(mpm_context* c, uint32_t xsel, bool check_cons, bool head, bool tail);

// Material index of a run: the plastic model on Hencky elasticity, kMatNeoHookean, or a closed-form
// Perzyna variant (kMatVMViscLinear, kMatDPViscLinear) when the parameters select that branch.
inline int material_of(const Params& P) {
    if (P.elastic_model == MPM_ELASTIC_NEOHOOKEAN) return kMatNeoHookean;
    if (P.plastic_model == MPM_PLASTIC_VMVISC && P.perzyna_exp == 1 && P.xi == 0) return kMatVMViscLinear;
    if (P.plastic_model == MPM_PLASTIC_DPVISC && P.perzyna_exp == 1) return kMatDPViscLinear;
    return P.plastic_model;
}

// Compile-time switches shared by the scenes of one batched launch.
struct BatchClassKey {
    int D, SPL, split, sep, flip, mibf, musl, mat, pbc;
    bool operator==(const BatchClassKey& o) const {
        return D == o.D && SPL == o.SPL && split == o.split && sep == o.sep && flip == o.flip && mibf == o.mibf && musl == o.musl &&
               mat == o.mat && pbc == o.pbc;
    }
};

// One class of a batch: its scenes, the device table of the launches and their workgroup counts.
struct BatchClass {
    BatchClassKey key{};
    std::vector<mpm_context*> members;
    BatchTable T{};
    uint32_t wg[BR_KINDS] = {0, 0, 0, 0};
};

// One step of every scene of a class (one launch per kernel), instantiated like enqueue_step.
template <int D, int SPL, class TF> void enqueue_batch_step(const BatchClass& bc, hipStream_t s);

// A mutex that another host thread holds while it makes HIP calls (allocations, synchronous copies) that
// must not run during a stream capture of this batch. A chunk that needs a new graph while the mutex is
// taken is launched without a graph instead.
void batch_set_capture_lock(mpm_batch* b, std::mutex* m);

}  // namespace mpm
