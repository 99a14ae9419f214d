// mpm_engine.hip — context management and the C ABI of include/mpm.h for the gfx950 MPM engine.
//
// A context owns one HIP stream, the particle state (structure of arrays in HBM), a dense node grid sized
// with headroom for a batch of steps, and a device-resident StepState. mpm_step enqueues batches of
// steps without host synchronisation; the device stops itself (StepState::stop) at errors, frame ends,
// the final time or when the next grid would not fit, and the host re-allocates and resumes.
#include "mpm_context.hpp"
#include <vector>
#include <cstring>
#include <cmath>
#include <algorithm>
#include <new>
#include <type_traits>
#include <cstddef>
#include <cstdio>
#include <chrono>
#include <string>

using namespace mpm;

// ---- Scratch of mpm_step_noalloc -----------------------------------------------------------------------
// A file-scope device array that holds the grid, block arrays and staging tiles of a no-allocation step.
// Every GPU that loads the library holds its own copy until the process ends. Host code must not pass the
// array's name to a kernel or a copy: that is a host-side stand-in, not a device address. The device address
// comes from hipGetSymbolAddress, taken once per device with that device current (scratch_address).
// Nothing in it has to start at zero: every step clears its block histogram (k_setup) and writes its
// grid, cell offsets and staging tiles before it reads them.
constexpr size_t kScratchBytes = MPM_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) char scratch[kScratchBytes];

namespace {

constexpr uint32_t kBatch = 32;
// Arrays are padded so that consecutive ones start 4 KiB (floats) or 8 KiB (doubles) apart modulo 64 KiB:
// equal indices of different arrays then fall into different memory channels.
constexpr size_t kArrayPad = 1024;

#define HIP_TRY(expr) do { hipError_t e_ = (expr); if (e_ != hipSuccess) return MPM_ERR_HIP; } while (0)

int valid_enum(int v, int lo, int hi) { return v >= lo && v <= hi; }

void derive_params(const mpm_config& c, Params& P) {
    std::memset(&P, 0, sizeof(P));
    P.dim = c.dim; P.spline_degree = c.spline_degree; P.elastic_model = c.elastic_model; P.plastic_model = c.plastic_model;
    P.hardening_law = c.hardening_law; P.use_pradhana = c.use_pradhana; P.use_mises_q = c.use_mises_q; P.use_mibf = c.use_mibf;
    P.use_musl = c.use_musl; P.gravity_special = c.gravity_special; P.change_particle_positions = c.change_particle_positions;
    P.pbc = c.pbc ? 1 : 0;
    P.L[0] = c.Lx; P.L[1] = c.Ly; P.L[2] = c.Lz;
    P.end_frame = c.end_frame;
    // simulate() set-up, simulation.cpp:94-145
    P.dx = c.dx;
    P.apicDinverse = (c.spline_degree == 3 ? 3.0 : 4.0) / (c.dx * c.dx);
    P.lambda = c.nu * c.E / ((1.0 + c.nu) * (1.0 - 2.0 * c.nu));
    P.mu = c.E / (2.0 * (1.0 + c.nu));
    P.K = P.lambda + 2.0 * P.mu / c.dim;  // calculateBulkModulus, simulation.cpp:288-290
    P.wave_speed = std::sqrt(c.E / c.rho);
    P.dt_max = c.cfl_elastic * c.dx / P.wave_speed;
    P.frame_dt = 1.0 / c.fps;
    P.one_over_dx = 1.0 / c.dx;
    if (c.use_mises_q) { P.q_prefac = std::sqrt(3.0) / std::sqrt(2.0); P.d_prefac = std::sqrt(2.0) / std::sqrt(3.0); }
    else { P.q_prefac = 1.0 / std::sqrt(2.0); P.d_prefac = std::sqrt(2.0); }
    P.e_mu_prefac = 2 * P.q_prefac * P.mu;
    P.f_mu_prefac = 2 * P.q_prefac / P.d_prefac * P.mu;
    P.rma_prefac = 2 * P.q_prefac * P.q_prefac;
    P.fac_Q = c.I_ref / (c.grain_diameter * std::sqrt(c.rho_s));
    P.final_time = c.end_frame * P.frame_dt;
    P.particle_mass = c.particle_mass; P.particle_volume = c.particle_volume;
    P.cfl = c.cfl; P.flip_ratio = c.flip_ratio; P.gravity_time = c.gravity_time;
    for (int d = 0; d < 3; ++d) { P.gravity_final[d] = d < c.dim ? c.gravity[d] : 0.0; P.grid_reference_point[d] = c.grid_reference_point[d]; }
    P.E = c.E; P.nu = c.nu; P.stress_tolerance = c.stress_tolerance; P.xi = c.xi;
    P.q_max = c.q_max; P.q_min = c.q_min; P.p_min = c.p_min; P.M = c.M; P.q_cohesion = c.q_cohesion;
    P.perzyna_exp = c.perzyna_exp; P.perzyna_visc = c.perzyna_visc; P.beta = c.beta; P.p0 = c.p0;
    P.rho_s = c.rho_s; P.grain_diameter = c.grain_diameter; P.I_ref = c.I_ref; P.mu_1 = c.mu_1; P.mu_2 = c.mu_2;
    // The split form's transfers carry about 1e-7 relative error; rotations below 1e-12 of the largest
    // diagonal entry change its stresses far less than that, and skipping them saves sweeps.
    P.svd_tol = c.precision == MPM_PRECISION_SPLIT ? 1e-12 : kSvdTolEigen;
}

size_t tile_nodes(int D, int spl) {
    return D == 3 ? (spl == 2 ? (size_t)BG<3, 2>::NT : (size_t)BG<3, 3>::NT) : (spl == 2 ? (size_t)BG<2, 2>::NT : (size_t)BG<2, 3>::NT);
}

// Bytes of one particle layout: double x, F and six plastic scalars; TF v, B and the symmetric tau.
size_t layout_bytes(int D, bool split) {
    const size_t nd = D + D * D + 6, nt = D + D * D + (D == 3 ? 6 : 3);
    return nd * sizeof(double) + nt * (split ? sizeof(float) : sizeof(double));
}

int alloc_particles(mpm_context* c, uint32_t n) {
    const int D = c->cfg.dim;
    const size_t need = 2 * layout_bytes(D, c->split) * ((size_t)n + kArrayPad) + 64;
    // id x2, key, perm, then the byte arrays cellk and cells; with periodic boundaries key, perm, cellk and
    // cells also hold one periodic copy per particle
    const size_t ks = (size_t)(c->cfg.pbc ? 2 : 1) * n;
    const size_t ineed = (size_t)2 * n + 2 * ks + (2 * ks + 3) / 4;
    if (!(c->pbuf && c->pbuf_bytes >= need)) {
        if (c->pbuf) { (void)hipFree(c->pbuf); c->pbuf = nullptr; }
        if (hipMalloc(&c->pbuf, need) != hipSuccess) { c->pbuf = nullptr; c->pbuf_bytes = 0; return MPM_ERR_OUT_OF_MEMORY; }
        c->pbuf_bytes = need;
    }
    if (!(c->ibuf && c->ibuf_n >= ineed)) {
        if (c->ibuf) { (void)hipFree(c->ibuf); c->ibuf = nullptr; }
        if (hipMalloc(&c->ibuf, ineed * sizeof(uint32_t)) != hipSuccess) { c->ibuf = nullptr; c->ibuf_n = 0; return MPM_ERR_OUT_OF_MEMORY; }
        c->ibuf_n = ineed;
    }
    return 0;
}

template <class TF>
void layout_particles(mpm_context* c, uint32_t n) {
    const int D = c->cfg.dim;
    char* q = c->pbuf;
    const size_t np = (size_t)n + kArrayPad;
    auto taked = [&]() { double* r = (double*)q; q += sizeof(double) * np; return r; };
    auto taket = [&]() { TF* r = (TF*)q; q += sizeof(TF) * np; return r; };
    for (int s = 0; s < 2; ++s) {
        PartStateT<TF> ps{};
        for (int d = 0; d < D; ++d) ps.x[d] = taked();
        for (int i = 0; i < D * D; ++i) ps.F[i] = taked();
        ps.eps_dev = taked(); ps.eps_vol = taked(); ps.eps_prad = taked(); ps.dgamma = taked(); ps.visc = taked(); ps.muI = taked();
        for (int d = 0; d < D; ++d) ps.v[d] = taket();
        for (int i = 0; i < D * D; ++i) ps.B[i] = taket();
        for (int i = 0; i < (D == 3 ? 6 : 3); ++i) ps.tau[i] = taket();
        ps.id = c->ibuf + (size_t)s * n;
        layouts<TF>(c)[s] = ps;
    }
    const size_t ks = c->cfg.pbc ? (size_t)2 * n : (size_t)n;
    c->sb.key = c->ibuf + (size_t)2 * n;
    c->sb.perm = c->sb.key + ks;
    c->sb.cellk = (uint8_t*)(c->sb.perm + ks);
    c->sb.cells = c->sb.cellk + ks;
    c->sb.ghosts = c->cfg.pbc ? 1u : 0u;
}

// Nodes per axis by which the grid may grow over `batch` steps: the particle box grows by at most about
// cfl * dx per side per step (dt <= cfl dx / vmax), plus the remesh snap. A step that outgrows it stops
// (StepState::need_realloc) and the host grows the grid and resumes, so the margin only sets how often that
// happens.
uint32_t growth_nodes(const mpm_context* c, uint32_t batch) {
    const double cfl = std::min(1.0, std::max(0.25, c->cfg.cfl));
    return 2 * (uint32_t)std::ceil(cfl * batch) + 6;
}

// Grid that the next `batch` steps may need: nodes per axis (largest axis), nodes, and blocks of the grid
// grown by the margin.
struct GridNeed { uint32_t axis; uint64_t nodes, blocks; };
GridNeed grid_need(const mpm_context* c, uint32_t batch) {
    const int D = c->cfg.dim;
    const uint32_t margin = growth_nodes(c, batch);
    GridNeed nd{0, 1, 1};
    const uint32_t be = D == 3 ? BlockEdge<3>::v : BlockEdge<2>::v;
    for (int d = 0; d < D; ++d) {
        const uint32_t a = c->est_n[d] + margin;
        nd.axis = std::max(nd.axis, a);
        nd.nodes *= a;
        nd.blocks *= (a + be - 1) / be;
    }
    return nd;
}

// Bytes of the parts of a store: block arrays (count, offset, cursor; with periodic boundaries each also
// holds the periodic copies' blocks after the particles'), dual-cell offsets, staging tiles, node arrays
// (mass, friction, v, flip) and node coordinates.
struct StoreBytes { size_t bbuf, cellbuf, stgbuf, gbuf, cbuf; };
StoreBytes store_bytes(const mpm_context* c, uint64_t blocks, uint64_t nodes, uint32_t axis) {
    const int D = c->cfg.dim;
    const size_t tf = c->split ? sizeof(float) : sizeof(double);
    const uint64_t bs = c->cfg.pbc ? 2 * blocks : blocks;
    StoreBytes b;
    b.bbuf = (size_t)3 * bs * sizeof(uint32_t);
    b.cellbuf = (size_t)65 * blocks * sizeof(uint16_t);
    b.stgbuf = (size_t)kMaxNV * tile_nodes(D, c->cfg.spline_degree) * blocks * tf;
    b.gbuf = (size_t)(1 + 2 * D + 1) * nodes * tf;
    b.cbuf = (size_t)3 * axis * sizeof(double);
    return b;
}

// Points the context's views (sort arrays, staging, grid) into store s.
void bind_store(mpm_context* c, const GridStore& s) {
    const int D = c->cfg.dim;
    const size_t nt = tile_nodes(D, c->cfg.spline_degree);
    const uint64_t bs = c->cfg.pbc ? 2ull * s.cap_blocks : (uint64_t)s.cap_blocks;
    c->sb.count = s.bbuf; c->sb.offset = s.bbuf + bs; c->sb.cursor = s.bbuf + 2 * bs;
    c->sb.cell_off = s.cellbuf;
    c->stg.base = (double*)s.stgbuf; c->stg.stride = (uint32_t)(nt * s.cap_blocks);
    c->stgf.base = (float*)s.stgbuf; c->stgf.stride = (uint32_t)(nt * s.cap_blocks);
    auto carve = [&](auto& g) {
        using TF = std::remove_pointer_t<decltype(g.mass)>;
        TF* q = (TF*)s.gbuf;
        auto take = [&]() { TF* r = q; q += s.cap_nodes; return r; };
        g.mass = take();
        g.fric = take();
        for (int d = 0; d < D; ++d) g.v[d] = take();
        for (int d = 0; d < D; ++d) g.flip[d] = take();
        for (int d = 0; d < 3; ++d) g.coord[d] = s.cbuf + (size_t)d * s.cap_axis;
    };
    carve(c->gp);
    carve(c->gf);
    c->cap_blocks = s.cap_blocks; c->cap_nodes = s.cap_nodes; c->cap_axis = s.cap_axis;
    c->act_blocks = std::min(c->act_blocks, c->cap_blocks);
    c->grid_bound = true;
    c->grid_gen++;
}

void unbind_store(mpm_context* c) {
    c->grid_bound = false; c->on_scratch = false;
    c->cap_blocks = 0; c->act_blocks = 0; c->cap_nodes = 0; c->cap_axis = 0;
    c->grid_gen++;
}

// Active blocks: the block kernels launch one workgroup per active block, and a step whose grid has more
// blocks stops for a reallocation. They follow the batch's need with some hysteresis (grow to 5/4 of it,
// shrink below half), so that batches keep their launch shapes and graphs for a while.
void set_active_blocks(mpm_context* c, uint64_t need) {
    if (need > c->act_blocks || 2 * need < c->act_blocks)
        c->act_blocks = (uint32_t)std::min<uint64_t>(c->cap_blocks, need + need / 4);
}

// Grows the context's own store to the grid of the next `batch` steps (reallocating) and binds it.
int ensure_grid(mpm_context* c, uint32_t batch) {
    const GridNeed nd = grid_need(c, batch);
    GridStore& o = c->own;
    bool changed = !c->grid_bound || c->on_scratch;
    if (nd.blocks > o.cap_blocks) {
        if (o.bbuf) { (void)hipFree(o.bbuf); o.bbuf = nullptr; }
        if (o.cellbuf) { (void)hipFree(o.cellbuf); o.cellbuf = nullptr; }
        if (o.stgbuf) { (void)hipFree(o.stgbuf); o.stgbuf = nullptr; }
        o.cap_blocks = 0;
        unbind_store(c);
        const size_t nt = tile_nodes(c->cfg.dim, c->cfg.spline_degree);
        if (nd.blocks * nt > 0xFFFFFFF0ull) return MPM_ERR_OUT_OF_MEMORY;
        const uint64_t alloc = std::min<uint64_t>(0xFFFFFFF0ull / nt, nd.blocks + nd.blocks / 4);  // room for the active count to grow
        const StoreBytes sz = store_bytes(c, alloc, 0, 0);
        if (hipMalloc(&o.bbuf, sz.bbuf) != hipSuccess) { o.bbuf = nullptr; return MPM_ERR_OUT_OF_MEMORY; }
        if (hipMalloc(&o.cellbuf, sz.cellbuf) != hipSuccess) { o.cellbuf = nullptr; return MPM_ERR_OUT_OF_MEMORY; }
        if (hipMalloc(&o.stgbuf, sz.stgbuf) != hipSuccess) { o.stgbuf = nullptr; return MPM_ERR_OUT_OF_MEMORY; }
        o.cap_blocks = (uint32_t)alloc;
        changed = true;
    }
    if (!(nd.axis <= o.cap_axis && nd.nodes <= o.cap_nodes)) {
        if (nd.nodes > 0xFFFFFFF0ull) return MPM_ERR_OUT_OF_MEMORY;
        const uint32_t want_axis = std::max(nd.axis, o.cap_axis);
        const uint64_t want_nodes = std::max(nd.nodes, o.cap_nodes);
        if (o.gbuf) { (void)hipFree(o.gbuf); o.gbuf = nullptr; }
        if (o.cbuf) { (void)hipFree(o.cbuf); o.cbuf = nullptr; }
        o.cap_nodes = 0; o.cap_axis = 0;
        unbind_store(c);
        const StoreBytes sz = store_bytes(c, 0, want_nodes, want_axis);
        if (hipMalloc(&o.gbuf, sz.gbuf) != hipSuccess) { o.gbuf = nullptr; return MPM_ERR_OUT_OF_MEMORY; }
        if (hipMalloc(&o.cbuf, sz.cbuf) != hipSuccess) { o.cbuf = nullptr; return MPM_ERR_OUT_OF_MEMORY; }
        o.cap_axis = want_axis; o.cap_nodes = want_nodes;
        changed = true;
    }
    if (changed) { bind_store(c, o); c->on_scratch = false; }
    set_active_blocks(c, nd.blocks);
    return 0;
}

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
char* g_scratch_ptr[kMaxScratchDevices] = {};

// Device address of `scratch` on device dev (current), taken once per device. 0 on success.
int scratch_address(int dev, char** out) {
    if (dev < 0 || dev >= kMaxScratchDevices) return MPM_ERR_UNSUPPORTED;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MPM_ERR_HIP;
        g_scratch_ptr[dev] = (char*)p;
    }
    *out = g_scratch_ptr[dev];
    return 0;
}

// Carves a store for the grid of the next `batch` steps out of the scratch at `base` and binds it; no
// device memory is allocated. MPM_ERR_SCRATCH_LIMIT when that grid does not fit.
int bind_scratch(mpm_context* c, uint32_t batch, char* base) {
    const GridNeed nd = grid_need(c, batch);
    if (nd.nodes > 0xFFFFFFF0ull || nd.blocks * tile_nodes(c->cfg.dim, c->cfg.spline_degree) > 0xFFFFFFF0ull) return MPM_ERR_SCRATCH_LIMIT;
    const StoreBytes sz = store_bytes(c, nd.blocks, nd.nodes, nd.axis);
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    const size_t o_cell = align(sz.bbuf), o_stg = o_cell + align(sz.cellbuf), o_g = o_stg + align(sz.stgbuf),
                 o_c = o_g + align(sz.gbuf), end = o_c + align(sz.cbuf);
    if (end > kScratchBytes) return MPM_ERR_SCRATCH_LIMIT;
    GridStore s;
    s.bbuf = (uint32_t*)base; s.cellbuf = (uint16_t*)(base + o_cell); s.stgbuf = base + o_stg;
    s.gbuf = base + o_g; s.cbuf = (double*)(base + o_c);
    s.cap_blocks = (uint32_t)nd.blocks; s.cap_nodes = nd.nodes; s.cap_axis = nd.axis;
    bind_store(c, s);
    c->on_scratch = true;
    c->act_blocks = s.cap_blocks;
    return 0;
}

__global__ void k_control(StepState* S, uint32_t stop_at_frame_end, uint32_t check_cons, uint32_t cap_axis, uint32_t cap_nodes,
                          uint32_t cap_blocks) {
    if (threadIdx.x != 0) return;
    S->stop = 0; S->need_realloc = 0; S->steps_done = 0; S->err = 0; S->warn = 0; S->frame_ended = 0;
    S->stop_at_frame_end = stop_at_frame_end; S->check_conservation = check_cons; S->step_limit = 0xFFFFFFFFu;
    S->cap_axis = cap_axis; S->cap_nodes = cap_nodes; S->cap_blocks = cap_blocks;
}

template <class TF>
void enqueue_tf(mpm_context* c, uint32_t xsel, bool check_cons, bool head, bool tail) {
    const int D = c->cfg.dim, SPL = c->cfg.spline_degree;
    if (D == 3 && SPL == 2) enqueue_step<3, 2, TF>(c, xsel, check_cons, head, tail);
    else if (D == 3 && SPL == 3) enqueue_step<3, 3, TF>(c, xsel, check_cons, head, tail);
    else if (D == 2 && SPL == 2) enqueue_step<2, 2, TF>(c, xsel, check_cons, head, tail);
    else enqueue_step<2, 3, TF>(c, xsel, check_cons, head, tail);
}

// Step j of a batch of `batch` steps: the first launches its own k_setup, the others are set up by the
// previous step's last launch, and the last ends with k_finish alone.
void enqueue(mpm_context* c, uint32_t xsel, bool check_cons, uint32_t j, uint32_t batch) {
    const bool head = j == 0, tail = j + 1 == batch;
    if (c->split) enqueue_tf<float>(c, xsel, check_cons, head, tail);
    else enqueue_tf<double>(c, xsel, check_cons, head, tail);
}

int sync_state(mpm_context* c) {
    HIP_TRY(hipMemcpyAsync(c->Sh, c->S, sizeof(StepState), hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// Upload of the caller's particle arrays (interleaved dim / dim*dim per particle) into layout 0.
template <class TF>
int upload_particles(mpm_context* c, uint32_t n, const mpm_particle_arrays* in) {
    const int D = c->cfg.dim;
    layout_particles<TF>(c, n);
    std::vector<double> col(n);
    std::vector<TF> colt(n);
    auto up = [&](double* dst) { return hipMemcpy(dst, col.data(), sizeof(double) * n, hipMemcpyHostToDevice); };
    auto upt = [&](TF* dst) {
        for (uint32_t p = 0; p < n; ++p) colt[p] = (TF)col[p];
        return hipMemcpy(dst, colt.data(), sizeof(TF) * n, hipMemcpyHostToDevice);
    };
    const PartStateT<TF>& L0 = layouts<TF>(c)[0];
    for (int d = 0; d < D; ++d) {
        for (uint32_t p = 0; p < n; ++p) col[p] = in->x[(size_t)p * D + d];
        HIP_TRY(up(L0.x[d]));
        for (uint32_t p = 0; p < n; ++p) col[p] = in->v ? in->v[(size_t)p * D + d] : 0.0;
        HIP_TRY(upt(L0.v[d]));
    }
    for (int i = 0; i < D * D; ++i) {
        const bool diag = (i % (D + 1)) == 0;
        for (uint32_t p = 0; p < n; ++p) col[p] = in->F ? in->F[(size_t)p * D * D + i] : (diag ? 1.0 : 0.0);
        HIP_TRY(up(L0.F[i]));
        for (uint32_t p = 0; p < n; ++p) col[p] = in->Bmat ? in->Bmat[(size_t)p * D * D + i] : 0.0;
        HIP_TRY(upt(L0.B[i]));
    }
    const double* sc[6] = {in->eps_pl_dev, in->eps_pl_vol, in->eps_pl_vol_pradhana, in->delta_gamma, in->viscosity, in->muI};
    double* dst[6] = {L0.eps_dev, L0.eps_vol, L0.eps_prad, L0.dgamma, L0.visc, L0.muI};
    for (int k = 0; k < 6; ++k) {
        for (uint32_t p = 0; p < n; ++p) col[p] = sc[k] ? sc[k][p] : 0.0;
        if (k == 5 && c->cfg.use_mibf) {  // simulate() fills muI when use_mibf (simulation.cpp:138-145)
            const double fill = (c->cfg.plastic_model == MPM_PLASTIC_DPMUI || c->cfg.plastic_model == MPM_PLASTIC_MCCMUI) ? c->cfg.mu_1 : c->cfg.M;
            std::fill(col.begin(), col.end(), fill);
        }
        HIP_TRY(up(dst[k]));
    }
    return 0;
}

// Download of the current layout into the caller's order (particle ids).
template <class TF>
int download_particles(mpm_context* c, mpm_particle_arrays* out) {
    const int D = c->cfg.dim;
    const uint32_t n = c->n;
    const PartStateT<TF>& L = layouts<TF>(c)[c->Sh->xsel & 1];
    std::vector<uint32_t> id(n);
    HIP_TRY(hipMemcpy(id.data(), L.id, sizeof(uint32_t) * n, hipMemcpyDeviceToHost));
    for (uint32_t p = 0; p < n; ++p) if (id[p] >= n) return MPM_ERR_HIP;
    std::vector<double> col(n);
    std::vector<TF> colt(n);
    auto down = [&](const double* src) { return hipMemcpy(col.data(), src, sizeof(double) * n, hipMemcpyDeviceToHost); };
    auto downt = [&](const TF* src) {
        hipError_t e = hipMemcpy(colt.data(), src, sizeof(TF) * n, hipMemcpyDeviceToHost);
        for (uint32_t p = 0; p < n; ++p) col[p] = (double)colt[p];
        return e;
    };
    for (int d = 0; d < D; ++d) {
        if (out->x) { HIP_TRY(down(L.x[d])); for (uint32_t p = 0; p < n; ++p) out->x[(size_t)id[p] * D + d] = col[p]; }
        if (out->v) { HIP_TRY(downt(L.v[d])); for (uint32_t p = 0; p < n; ++p) out->v[(size_t)id[p] * D + d] = col[p]; }
    }
    for (int i = 0; i < D * D; ++i) {
        if (out->F) { HIP_TRY(down(L.F[i])); for (uint32_t p = 0; p < n; ++p) out->F[(size_t)id[p] * D * D + i] = col[p]; }
        if (out->Bmat) { HIP_TRY(downt(L.B[i])); for (uint32_t p = 0; p < n; ++p) out->Bmat[(size_t)id[p] * D * D + i] = col[p]; }
    }
    double* dsts[6] = {out->eps_pl_dev, out->eps_pl_vol, out->eps_pl_vol_pradhana, out->delta_gamma, out->viscosity, out->muI};
    const double* srcs[6] = {L.eps_dev, L.eps_vol, L.eps_prad, L.dgamma, L.visc, L.muI};
    for (int k = 0; k < 6; ++k) {
        if (!dsts[k]) continue;
        HIP_TRY(down(srcs[k]));
        for (uint32_t p = 0; p < n; ++p) dsts[k][id[p]] = col[p];
    }
    return 0;
}

// Download of grid node arrays (converted to double).
template <class TF>
int download_grid(mpm_context* c, uint32_t nn, double* mass, double* v, double* flip, double* friction) {
    const int D = c->cfg.dim;
    const GridPtrsT<TF>& g = grid_of<TF>(c);
    std::vector<TF> col(nn);
    auto down = [&](const TF* src) { return hipMemcpy(col.data(), src, sizeof(TF) * nn, hipMemcpyDeviceToHost); };
    if (mass) { HIP_TRY(down(g.mass)); for (uint32_t l = 0; l < nn; ++l) mass[l] = (double)col[l]; }
    if (friction && c->cfg.use_mibf) { HIP_TRY(down(g.fric)); for (uint32_t l = 0; l < nn; ++l) friction[l] = (double)col[l]; }
    for (int d = 0; d < D; ++d) {
        if (v) { HIP_TRY(down(g.v[d])); for (uint32_t l = 0; l < nn; ++l) v[(size_t)l * D + d] = (double)col[l]; }
        if (flip && !c->flip_formed) { for (uint32_t l = 0; l < nn; ++l) flip[(size_t)l * D + d] = 0.0; }
        else if (flip) { HIP_TRY(down(g.flip[d])); for (uint32_t l = 0; l < nn; ++l) flip[(size_t)l * D + d] = (double)col[l]; }
    }
    return 0;
}

template <class TF>
void launch_binning(mpm_context* c, uint32_t n, int32_t* dbase, uint32_t* dkey) {
    const PartStateT<TF>& L = layouts<TF>(c)[c->Sh->xsel & 1];
    const GridPtrsT<TF>& g = grid_of<TF>(c);
    const uint32_t pblocks = (n + kBlock - 1) / kBlock;
    const int D = c->cfg.dim, spl = c->cfg.spline_degree;
    if (D == 3 && spl == 2) k_binning<3, 2, TF><<<pblocks, kBlock, 0, c->stream>>>(c->P, c->S, L, g, n, dbase, dkey);
    else if (D == 3) k_binning<3, 3, TF><<<pblocks, kBlock, 0, c->stream>>>(c->P, c->S, L, g, n, dbase, dkey);
    else if (spl == 2) k_binning<2, 2, TF><<<pblocks, kBlock, 0, c->stream>>>(c->P, c->S, L, g, n, dbase, dkey);
    else k_binning<2, 3, TF><<<pblocks, kBlock, 0, c->stream>>>(c->P, c->S, L, g, n, dbase, dkey);
}


// ---------------------------------------------------------------------------------------------------------
// Batches (mpm_batch_*): member contexts grouped into launch classes, one device table per class.
// ---------------------------------------------------------------------------------------------------------

constexpr uint32_t kHeadBytes = (uint32_t)offsetof(StepState, red);  // loop state read back per chunk
static_assert(kHeadBytes % 4 == 0, "StepState head must be whole words");

BatchClassKey class_key(const mpm_context* c, bool check_cons) {
    BatchClassKey k{};
    k.D = c->cfg.dim; k.SPL = c->cfg.spline_degree; k.split = c->split ? 1 : 0;
    k.sep = (c->P.flip_ratio >= -1 || check_cons) ? 1 : 0; k.flip = c->P.flip_ratio >= -1 ? 1 : 0;
    k.mibf = c->P.use_mibf ? 1 : 0; k.musl = c->P.use_musl ? 1 : 0;
    k.mat = material_of(c->P);
    k.pbc = c->cfg.pbc ? 1 : 0;
    return k;
}

uint32_t grid_workgroups(const mpm_context* c) {
    return (uint32_t)std::min<uint64_t>((c->cap_nodes + kBlock - 1) / kBlock, 8192);
}

template <class TF>
SceneRef<TF> scene_ref(mpm_context* c) {
    SceneRef<TF> r{};
    r.P = c->dP; r.S = c->S;
    r.ps[0] = layouts<TF>(c)[0]; r.ps[1] = layouts<TF>(c)[1];
    r.sb = c->sb; r.g = grid_of<TF>(c); r.stg = staging_of<TF>(c);
    r.n = c->n; r.cap_axis = c->cap_axis; r.cap_nodes = (uint32_t)c->cap_nodes; r.cap_blocks = c->act_blocks;
    return r;
}

template <class TF>
void enqueue_batch_class(const BatchClass& bc, hipStream_t s) {
    const BatchClassKey& k = bc.key;
    if (k.D == 3 && k.SPL == 2) enqueue_batch_step<3, 2, TF>(bc, s);
    else if (k.D == 3) enqueue_batch_step<3, 3, TF>(bc, s);
    else if (k.SPL == 2) enqueue_batch_step<2, 2, TF>(bc, s);
    else enqueue_batch_step<2, 3, TF>(bc, s);
}

}  // namespace

struct mpm_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    mpm_batch_options opt{};
    std::vector<mpm_context*> scenes;     // members, in insertion order
    std::vector<BatchClass> classes;
    std::vector<mpm_context*> order;      // members in class order (rows of the limit and state arrays)
    bool dirty = true;                    // classes or device tables must be rebuilt
    bool check_cons = false;              // variant the classes were built for
    char* dtab = nullptr; size_t dtab_bytes = 0;     // refs, range starts, limits, StepState pointers, heads
    uint32_t* hlimit = nullptr; uint32_t* hhead = nullptr; size_t hcap = 0;  // pinned
    uint32_t* dlimit = nullptr; uint32_t* dhead = nullptr; StepState** dSp = nullptr;
    struct Graph { uint32_t chunk, flags; hipGraphExec_t exec; };
    std::vector<Graph> graphs;            // instantiated chunks of the current tables, by (steps, flags)
    std::mutex* capture_lock = nullptr;   // see batch_set_capture_lock
};

namespace {

constexpr size_t kMaxGraphs = 4;

void drop_graph(mpm_batch* b) {
    for (auto& g : b->graphs) (void)hipGraphExecDestroy(g.exec);
    b->graphs.clear();
}

// Groups the members into classes and uploads every class's scene table and range starts.
int rebuild_tables(mpm_batch* b, bool check_cons) {
    drop_graph(b);
    b->classes.clear();
    for (mpm_context* c : b->scenes) {
        const BatchClassKey k = class_key(c, check_cons);
        auto it = std::find_if(b->classes.begin(), b->classes.end(), [&](const BatchClass& x) { return x.key == k; });
        if (it == b->classes.end()) { b->classes.push_back(BatchClass{}); b->classes.back().key = k; it = b->classes.end() - 1; }
        it->members.push_back(c);
    }
    b->order.clear();
    for (auto& bc : b->classes) for (mpm_context* c : bc.members) b->order.push_back(c);
    const size_t ns = b->order.size();
    // layout of the device table: per class [refs][starts]; then limits[ns], S pointers[ns], heads[ns]
    std::vector<size_t> ref_off, start_off;
    size_t off = 0;
    auto align = [](size_t v) { return (v + 255) & ~(size_t)255; };
    for (auto& bc : b->classes) {
        const size_t rs = bc.key.split ? sizeof(SceneRef<float>) : sizeof(SceneRef<double>);
        ref_off.push_back(off); off = align(off + rs * bc.members.size());
        start_off.push_back(off); off = align(off + sizeof(uint32_t) * BR_KINDS * (bc.members.size() + 1));
    }
    const size_t lim_off = off; off = align(off + sizeof(uint32_t) * std::max<size_t>(ns, 1));
    const size_t sp_off = off; off = align(off + sizeof(StepState*) * std::max<size_t>(ns, 1));
    const size_t head_off = off; off = align(off + (size_t)kHeadBytes * std::max<size_t>(ns, 1));
    if (off > b->dtab_bytes) {
        if (b->dtab) { (void)hipFree(b->dtab); b->dtab = nullptr; b->dtab_bytes = 0; }
        if (hipMalloc(&b->dtab, off) != hipSuccess) { b->dtab = nullptr; return MPM_ERR_OUT_OF_MEMORY; }
        b->dtab_bytes = off;
    }
    if (ns > b->hcap) {
        if (b->hlimit) { (void)hipHostFree(b->hlimit); b->hlimit = nullptr; }
        if (b->hhead) { (void)hipHostFree(b->hhead); b->hhead = nullptr; }
        b->hcap = 0;
        if (hipHostMalloc(&b->hlimit, sizeof(uint32_t) * ns, hipHostMallocDefault) != hipSuccess) { b->hlimit = nullptr; return MPM_ERR_OUT_OF_MEMORY; }
        if (hipHostMalloc(&b->hhead, (size_t)kHeadBytes * ns, hipHostMallocDefault) != hipSuccess) { b->hhead = nullptr; return MPM_ERR_OUT_OF_MEMORY; }
        b->hcap = ns;
    }
    std::vector<char> host(off, 0);
    b->dlimit = (uint32_t*)(b->dtab + lim_off);
    b->dSp = (StepState**)(b->dtab + sp_off);
    b->dhead = (uint32_t*)(b->dtab + head_off);
    for (size_t k = 0; k < b->classes.size(); ++k) {
        BatchClass& bc = b->classes[k];
        const uint32_t m = (uint32_t)bc.members.size();
        uint32_t* st = (uint32_t*)(host.data() + start_off[k]);
        for (int r = 0; r < BR_KINDS; ++r) st[r * (m + 1)] = 0;
        for (uint32_t i = 0; i < m; ++i) {
            mpm_context* c = bc.members[i];
            if (bc.key.split) { SceneRef<float> r = scene_ref<float>(c); std::memcpy(host.data() + ref_off[k] + i * sizeof(r), &r, sizeof(r)); }
            else { SceneRef<double> r = scene_ref<double>(c); std::memcpy(host.data() + ref_off[k] + i * sizeof(r), &r, sizeof(r)); }
            // particle kernels cover the periodic copies' entries too (k_bin_scatter); the others return early there
            const uint32_t wg[BR_KINDS] = {(uint32_t)((((uint64_t)c->n << c->sb.ghosts) + kBlock - 1) / kBlock), c->act_blocks,
                                           grid_workgroups(c), (c->act_blocks + kBlock - 1) / kBlock};
            for (int r = 0; r < BR_KINDS; ++r) st[r * (m + 1) + i + 1] = st[r * (m + 1) + i] + wg[r];
        }
        for (int r = 0; r < BR_KINDS; ++r) bc.wg[r] = st[r * (m + 1) + m];
        bc.T.refs = b->dtab + ref_off[k];
        bc.T.start = (const uint32_t*)(b->dtab + start_off[k]);
        bc.T.ns = m;
        size_t first = 0;
        for (size_t j = 0; j < k; ++j) first += b->classes[j].members.size();
        bc.T.limit = b->dlimit + first;
    }
    StepState** sp = (StepState**)(host.data() + sp_off);
    for (size_t i = 0; i < ns; ++i) sp[i] = b->order[i]->S;
    for (mpm_context* c : b->scenes) {  // the scene's constants, as they are now
        HIP_TRY(hipMemcpy(c->dP, &c->P, sizeof(Params), hipMemcpyHostToDevice));
        c->flip_formed = class_key(c, check_cons).sep != 0;
    }
    HIP_TRY(hipMemcpy(b->dtab, host.data(), off, hipMemcpyHostToDevice));
    b->check_cons = check_cons;
    b->dirty = false;
    return 0;
}

// One chunk: control, `chunk` steps of every class, then the loop-state heads into one buffer.
void enqueue_chunk(mpm_batch* b, uint32_t chunk, uint32_t flags) {
    const uint32_t sfe = (flags & MPM_STEP_STOP_AT_FRAME_END) ? 1u : 0u, cons = (flags & MPM_STEP_CHECK_CONSERVATION) ? 1u : 0u;
    for (const BatchClass& bc : b->classes) {
        if (bc.key.split) kb_control<float><<<bc.T.ns, 64, 0, b->stream>>>(bc.T, sfe, cons);
        else kb_control<double><<<bc.T.ns, 64, 0, b->stream>>>(bc.T, sfe, cons);
    }
    for (uint32_t j = 0; j < chunk; ++j)
        for (const BatchClass& bc : b->classes) {
            if (bc.key.split) enqueue_batch_class<float>(bc, b->stream);
            else enqueue_batch_class<double>(bc, b->stream);
        }
    kb_gather_state<><<<(uint32_t)b->order.size(), kBlock, 0, b->stream>>>(b->dSp, kHeadBytes / 4, b->dhead);
}

int launch_chunk(mpm_batch* b, uint32_t chunk, uint32_t flags) {
    if (!b->opt.use_graph) {
        enqueue_chunk(b, chunk, flags);
        return hipGetLastError() == hipSuccess ? 0 : MPM_ERR_HIP;
    }
    auto it = std::find_if(b->graphs.begin(), b->graphs.end(), [&](const mpm_batch::Graph& g) { return g.chunk == chunk && g.flags == flags; });
    if (it == b->graphs.end()) {
        std::unique_lock<std::mutex> lk;
        if (b->capture_lock) {
            lk = std::unique_lock<std::mutex>(*b->capture_lock, std::try_to_lock);
            if (!lk.owns_lock()) {  // another thread is making HIP calls: launch this chunk directly
                enqueue_chunk(b, chunk, flags);
                return hipGetLastError() == hipSuccess ? 0 : MPM_ERR_HIP;
            }
        }
        HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
        enqueue_chunk(b, chunk, flags);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(b->stream, &graph);
        if (e != hipSuccess || !graph) { if (graph) (void)hipGraphDestroy(graph); return MPM_ERR_HIP; }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return MPM_ERR_HIP;
        if (b->graphs.size() >= kMaxGraphs) { (void)hipGraphExecDestroy(b->graphs.front().exec); b->graphs.erase(b->graphs.begin()); }
        b->graphs.push_back({chunk, flags, exec});
        it = b->graphs.end() - 1;
    }
    HIP_TRY(hipGraphLaunch(it->exec, b->stream));
    return 0;
}

}  // namespace

namespace {
// Binary little-endian PLY with one element "vertex" of float64 properties, rows interleaved.
int write_ply(const char* path, size_t rows, const std::vector<std::string>& names, const std::vector<const double*>& cols,
              const std::vector<int>& stride) {
    FILE* f = std::fopen(path, "wb");
    if (!f) return MPM_ERR_INVALID_ARGUMENT;
    std::string h = "ply\nformat binary_little_endian 1.0\nelement vertex " + std::to_string(rows) + "\n";
    for (const std::string& nm : names) h += "property double " + nm + "\n";
    h += "end_header\n";
    bool ok = std::fwrite(h.data(), 1, h.size(), f) == h.size();
    std::vector<double> row(names.size());
    std::vector<double> buf;
    buf.reserve(names.size() * 4096);
    for (size_t r = 0; r < rows && ok; ++r) {
        for (size_t k = 0; k < names.size(); ++k) buf.push_back(cols[k][r * (size_t)stride[k]]);
        if (buf.size() >= names.size() * 4096 || r + 1 == rows) {
            ok = std::fwrite(buf.data(), sizeof(double), buf.size(), f) == buf.size();
            buf.clear();
        }
    }
    if (std::fclose(f) != 0) ok = false;
    return ok ? 0 : MPM_ERR_INVALID_ARGUMENT;
}
}  // namespace

namespace mpm {
void batch_set_capture_lock(mpm_batch* b, std::mutex* m) { if (b) b->capture_lock = m; }
}  // namespace mpm

extern "C" {

void mpm_config_defaults(mpm_config* c) {
    if (!c) return;
    std::memset(c, 0, sizeof(*c));
    c->dim = 3; c->spline_degree = 2; c->precision = MPM_PRECISION_FP64; c->n_threads = 1;
    c->elastic_model = MPM_ELASTIC_HENCKY; c->plastic_model = MPM_PLASTIC_NONE; c->hardening_law = MPM_HARD_EXPO_IMPL;
    c->use_pradhana = 1; c->end_frame = 1;
    for (int i = 0; i < 3; ++i) c->grid_reference_point[i] = 2e10;
    c->fps = 1; c->cfl = 0.5; c->cfl_elastic = 0.5; c->flip_ratio = -0.95; c->rho = 1000;
    c->Lx = c->Ly = c->Lz = 1;
    c->E = 1e6; c->nu = 0.3; c->stress_tolerance = 1e-5;
    c->q_max = 100; c->q_min = 100; c->p_min = -1e20; c->M = 1;
    c->perzyna_exp = 1; c->p0 = 1000;
    c->rho_s = 2500; c->grain_diameter = 1e-3; c->I_ref = 0.279;
    c->mu_1 = std::tan(20.9 * M_PI / 180.0); c->mu_2 = std::tan(32.8 * M_PI / 180.0);
}

int mpm_create(const mpm_config* cfg, mpm_context** out) {
    if (!cfg || !out) return MPM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    const mpm_config& c = *cfg;
    if (c.dim != 2 && c.dim != 3) return MPM_ERR_INVALID_ARGUMENT;
    if (c.spline_degree != 2 && c.spline_degree != 3) return MPM_ERR_UNSUPPORTED;
    if (c.precision != MPM_PRECISION_FP64 && c.precision != MPM_PRECISION_SPLIT) return MPM_ERR_INVALID_ARGUMENT;
    if (c.pbc && !(c.Lx > 0 && c.Ly > 0 && (c.dim == 2 || c.Lz > 0))) return MPM_ERR_INVALID_ARGUMENT;
    if (!valid_enum(c.elastic_model, 0, 1) || !valid_enum(c.plastic_model, 0, 10) || !valid_enum(c.hardening_law, 0, 4))
        return MPM_ERR_INVALID_ARGUMENT;
    if (c.plastic_model == MPM_PLASTIC_SNOW) return MPM_ERR_UNSUPPORTED;
    if (c.elastic_model != MPM_ELASTIC_HENCKY && c.plastic_model != MPM_PLASTIC_NONE) return MPM_ERR_UNSUPPORTED;  // simulation.cpp:80-84
    if (!(c.dx > 0) || !(c.fps > 0) || !(c.rho > 0)) return MPM_ERR_INVALID_ARGUMENT;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MPM_ERR_HIP;
    mpm_context* ctx = new (std::nothrow) mpm_context();
    if (!ctx) return MPM_ERR_OUT_OF_MEMORY;
    ctx->cfg = c;
    ctx->split = c.precision == MPM_PRECISION_SPLIT;
    derive_params(c, ctx->P);
    (void)hipGetDevice(&ctx->device);
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) { delete ctx; return MPM_ERR_HIP; }
    if (hipMalloc(&ctx->S, sizeof(StepState)) != hipSuccess) { (void)hipStreamDestroy(ctx->stream); delete ctx; return MPM_ERR_OUT_OF_MEMORY; }
    if (hipHostMalloc(&ctx->Sh, sizeof(StepState), hipHostMallocDefault) != hipSuccess) { (void)hipFree(ctx->S); (void)hipStreamDestroy(ctx->stream); delete ctx; return MPM_ERR_OUT_OF_MEMORY; }
    std::memset(ctx->Sh, 0, sizeof(StepState));
    *out = ctx;
    return 0;
}

void mpm_destroy(mpm_context* c) {
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->pbuf) (void)hipFree(c->pbuf);
    if (c->ibuf) (void)hipFree(c->ibuf);
    if (c->own.bbuf) (void)hipFree(c->own.bbuf);
    if (c->own.cellbuf) (void)hipFree(c->own.cellbuf);
    if (c->own.stgbuf) (void)hipFree(c->own.stgbuf);
    if (c->own.gbuf) (void)hipFree(c->own.gbuf);
    if (c->own.cbuf) (void)hipFree(c->own.cbuf);
    if (c->batch) (void)mpm_batch_remove(c->batch, c);
    if (c->S) (void)hipFree(c->S);
    if (c->Sh) (void)hipHostFree(c->Sh);
    if (c->dP) (void)hipFree(c->dP);
    for (void* q : c->sdf_bufs) (void)hipFree(q);
    for (hipEvent_t ev : c->ev_pool) (void)hipEventDestroy(ev);
    for (auto& g : c->step_graphs) (void)hipGraphExecDestroy(g.exec);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int mpm_add_plate(mpm_context* c, const mpm_plate* p) {
    if (!c || !p || c->batch) return MPM_ERR_INVALID_ARGUMENT;
    if ((int)c->plates.size() >= kMaxPlates) return MPM_ERR_UNSUPPORTED;
    if (!valid_enum(p->plate_type, 0, 5) || !valid_enum(p->bc, 0, 2)) return MPM_ERR_INVALID_ARGUMENT;
    DevPlate d{};
    d.pos_object = p->pos_object; d.pos_lower = p->pos_lower; d.pos_upper = p->pos_upper;
    for (int i = 0; i < 3; ++i) { d.v_object[i] = (i < c->cfg.dim) ? p->v_object[i] : 0.0; d.v_object_original[i] = d.v_object[i]; }
    d.friction = p->friction; d.vmin_factor = p->vmin_factor; d.load_factor = p->load_factor;
    d.plate_type = p->plate_type; d.bc = p->bc;
    c->plates.push_back(d);
    c->P.n_plates = (int)c->plates.size();
    if (c->ready) {  // particles already resident: append the plate to the device state too
        const size_t k = c->plates.size() - 1;
        HIP_TRY(hipStreamSynchronize(c->stream));
        c->Sh->plates[k] = d;
        HIP_TRY(hipMemcpy(&c->S->plates[k], &d, sizeof(DevPlate), hipMemcpyHostToDevice));
    }
    return 0;
}

int mpm_add_object(mpm_context* c, const mpm_object* o) {
    if (!c || !o || c->batch) return MPM_ERR_INVALID_ARGUMENT;
    if (c->P.n_objects >= kMaxObjects) return MPM_ERR_UNSUPPORTED;
    if (!valid_enum(o->kind, MPM_OBJECT_BOX_ROTATED, MPM_OBJECT_CURVE) || !valid_enum(o->bc, 0, 2)) return MPM_ERR_INVALID_ARGUMENT;
    if (o->kind == MPM_OBJECT_SILO && c->cfg.dim != 3) return MPM_ERR_UNSUPPORTED;
    DevObject& d = c->P.objects[c->P.n_objects];
    std::memset(&d, 0, sizeof(d));
    d.kind = o->kind; d.bc = o->bc; d.friction = o->friction; d.param = o->param;
    for (int i = 0; i < 3; ++i) { d.center[i] = o->center[i]; d.half_extents[i] = o->half_extents[i]; }
    const int D = c->cfg.dim;
    // the rotation arrives as dim x dim column-major
    for (int i = 0; i < D * D; ++i) d.rot[i] = o->rotation[i];
    c->P.n_objects++;
    return 0;
}

int mpm_add_sdf_object(mpm_context* c, const mpm_sdf* sd) {
    if (!c || !sd || !sd->phi || c->batch) return MPM_ERR_INVALID_ARGUMENT;
    if (c->P.n_objects >= kMaxObjects) return MPM_ERR_UNSUPPORTED;
    const int D = c->cfg.dim;
    if (!valid_enum(sd->bc, 0, 2) || !(sd->spacing > 0) || !std::isfinite(sd->spacing)) return MPM_ERR_INVALID_ARGUMENT;
    uint64_t npts = 1;
    for (int d = 0; d < 3; ++d) {
        const uint32_t want_min = d < D ? 2u : 1u;
        if (sd->n[d] < want_min || (d >= D && sd->n[d] != 1)) return MPM_ERR_INVALID_ARGUMENT;
        npts *= sd->n[d];
    }
    if (npts > (1ull << 28)) return MPM_ERR_INVALID_ARGUMENT;
    // phi and its gradient per lattice point: central differences, one-sided at the border
    std::vector<float> field(4 * npts, 0.0f);
    const uint32_t n0 = sd->n[0], n1 = sd->n[1], n2 = sd->n[2];
    const double h = sd->spacing;
    auto at = [&](uint32_t i, uint32_t j, uint32_t k) { return (double)sd->phi[((uint64_t)i * n1 + j) * n2 + k]; };
    for (uint32_t i = 0; i < n0; ++i)
        for (uint32_t j = 0; j < n1; ++j)
            for (uint32_t k = 0; k < n2; ++k) {
                const uint64_t q = ((uint64_t)i * n1 + j) * n2 + k;
                const uint32_t idx[3] = {i, j, k};
                field[4 * q] = sd->phi[q];
                for (int d = 0; d < D; ++d) {
                    uint32_t lo[3] = {i, j, k}, hi[3] = {i, j, k};
                    if (idx[d] > 0) lo[d] = idx[d] - 1;
                    if (idx[d] + 1 < sd->n[d]) hi[d] = idx[d] + 1;
                    const double g = (at(hi[0], hi[1], hi[2]) - at(lo[0], lo[1], lo[2])) / ((double)(hi[d] - lo[d]) * h);
                    field[4 * q + 1 + d] = (float)g;
                }
            }
    void* dbuf = nullptr;
    if (hipSetDevice(c->device) != hipSuccess) return MPM_ERR_HIP;
    if (hipMalloc(&dbuf, field.size() * sizeof(float)) != hipSuccess) return MPM_ERR_OUT_OF_MEMORY;
    if (hipMemcpy(dbuf, field.data(), field.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) { (void)hipFree(dbuf); return MPM_ERR_HIP; }
    c->sdf_bufs.push_back(dbuf);
    DevObject& o = c->P.objects[c->P.n_objects];
    std::memset(&o, 0, sizeof(o));
    o.kind = MPM_OBJECT_SDF; o.bc = sd->bc; o.friction = sd->friction; o.param = h;
    for (int d = 0; d < 3; ++d) { o.center[d] = d < D ? sd->origin[d] : 0.0; o.sdf.n[d] = sd->n[d]; }
    o.sdf.field = (const float*)dbuf;
    c->P.n_objects++;
    return 0;
}

int mpm_set_particles(mpm_context* c, uint32_t n, const mpm_particle_arrays* in) {
    if (!c || !in || !in->x || n == 0) return MPM_ERR_INVALID_ARGUMENT;
    if (c->batch) return MPM_ERR_INVALID_ARGUMENT;
    const int D = c->cfg.dim;
    if (hipSetDevice(c->device) != hipSuccess) return MPM_ERR_HIP;
    HIP_TRY(hipStreamSynchronize(c->stream));
    int rc = alloc_particles(c, n);
    if (rc) return rc;
    c->n = n;
    rc = c->split ? upload_particles<float>(c, n, in) : upload_particles<double>(c, n, in);
    if (rc) return rc;
    // initial grid estimate from the particle bounds (periodic: the fixed grid of remeshFixed(4), remesh.cpp:14-36)
    for (int d = 0; d < D && c->cfg.pbc; ++d) {
        const double L = d == 0 ? c->cfg.Lx : (d == 1 ? c->cfg.Ly : c->cfg.Lz);
        const double span = (L + 6.0 * c->cfg.dx - (d == 1 ? -c->cfg.dx : -5.0 * c->cfg.dx)) / c->cfg.dx;
        if (!(span < 4.0e6)) return MPM_ERR_OUT_OF_MEMORY;
        c->est_n[d] = (uint32_t)span + 2;
    }
    for (int d = 0; d < D && !c->cfg.pbc; ++d) {
        double mn = in->x[d], mx = in->x[d];
        for (uint32_t p = 1; p < n; ++p) { double v = in->x[(size_t)p * D + d]; mn = std::min(mn, v); mx = std::max(mx, v); }
        if (c->cfg.grid_reference_point[0] < 1e10) { mn = std::min(mn, c->cfg.grid_reference_point[d]); mx = std::max(mx, c->cfg.grid_reference_point[d]); }
        double span = (mx - mn) / c->cfg.dx;
        if (!(span < 4.0e6)) return MPM_ERR_OUT_OF_MEMORY;  // non-finite or absurd extent
        c->est_n[d] = (uint32_t)span + 8;
    }
    for (int d = D; d < 3; ++d) c->est_n[d] = 1;
    // step state
    StepState& S = *c->Sh;
    std::memset(&S, 0, sizeof(S));
    for (int d = 0; d < 3; ++d) { S.gravity[d] = c->P.gravity_final[d]; S.xmin_key[d] = 0x7FFFFFFFFFFFFFFFLL; S.xmax_key[d] = (-0x7FFFFFFFFFFFFFFFLL - 1); }
    for (size_t k = 0; k < c->plates.size(); ++k) S.plates[k] = c->plates[k];
    for (int s = 0; s < kRedSlots; ++s)
        for (int d = 0; d < 3; ++d) { S.red[s].mn[d] = 0x7FFFFFFFFFFFFFFFLL; S.red[s].mx[d] = (-0x7FFFFFFFFFFFFFFFLL - 1); }
    S.n[0] = S.n[1] = S.n[2] = 1;
    HIP_TRY(hipMemcpy(c->S, &S, sizeof(StepState), hipMemcpyHostToDevice));
    const uint32_t pblocks = (n + kBlock - 1) / kBlock;
    if (c->split) {
        if (D == 3) k_prepare<3, float><<<pblocks, kBlock, 0, c->stream>>>(c->P, c->S, c->pf[0], n);
        else k_prepare<2, float><<<pblocks, kBlock, 0, c->stream>>>(c->P, c->S, c->pf[0], n);
    } else {
        if (D == 3) k_prepare<3, double><<<pblocks, kBlock, 0, c->stream>>>(c->P, c->S, c->ps[0], n);
        else k_prepare<2, double><<<pblocks, kBlock, 0, c->stream>>>(c->P, c->S, c->ps[0], n);
    }
    HIP_TRY(hipGetLastError());
    rc = sync_state(c);
    if (rc) return rc;
    c->ready = true;
    return 0;
}

}  // extern "C"

namespace {
// Single-context step graphs. A batch of a scene with at most kGraphMaxParticles particles and at least
// kGraphMinSteps steps is captured once as a HIP graph and replayed: such a step is bound by the host's launch
// rate (about a dozen launches per step), not by the device. Larger scenes keep ordinary launches, whose
// enqueue cost hides behind the device work, so that a graph's capture would only add time.
constexpr uint32_t kGraphMaxParticles = 262144, kGraphMinSteps = 16;
constexpr size_t kMaxStepGraphs = 4;

int launch_step_graph(mpm_context* c, uint32_t batch, uint32_t xsel, uint32_t flags) {
    const bool check_cons = (flags & MPM_STEP_CHECK_CONSERVATION) != 0;
    auto match = [&](const mpm_context::StepGraph& g) {
        return g.batch == batch && g.xpar == (xsel & 1u) && g.flags == flags && g.act_blocks == c->act_blocks && g.n == c->n &&
               g.grid_gen == c->grid_gen && g.pbuf == (const void*)c->pbuf && g.ibuf == (const void*)c->ibuf &&
               std::memcmp(&g.P, &c->P, sizeof(Params)) == 0;
    };
    auto it = std::find_if(c->step_graphs.begin(), c->step_graphs.end(), match);
    if (it == c->step_graphs.end()) {
        HIP_TRY(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        k_control<<<1, 64, 0, c->stream>>>(c->S, (flags & MPM_STEP_STOP_AT_FRAME_END) ? 1u : 0u, check_cons ? 1u : 0u, c->cap_axis,
                                           (uint32_t)c->cap_nodes, c->act_blocks);
        for (uint32_t j = 0; j < batch; ++j) enqueue(c, xsel + j, check_cons, j, batch);
        hipGraph_t graph = nullptr;
        const hipError_t e = hipStreamEndCapture(c->stream, &graph);
        if (e != hipSuccess || !graph) { if (graph) (void)hipGraphDestroy(graph); return MPM_ERR_HIP; }
        hipGraphExec_t exec = nullptr;
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        if (ei != hipSuccess) return MPM_ERR_HIP;
        if (c->step_graphs.size() >= kMaxStepGraphs) {
            (void)hipGraphExecDestroy(c->step_graphs.front().exec);
            c->step_graphs.erase(c->step_graphs.begin());
        }
        mpm_context::StepGraph g;
        g.batch = batch; g.xpar = xsel & 1u; g.flags = flags; g.act_blocks = c->act_blocks; g.n = c->n; g.grid_gen = c->grid_gen;
        g.pbuf = c->pbuf; g.ibuf = c->ibuf; std::memcpy(&g.P, &c->P, sizeof(Params)); g.exec = exec;
        c->step_graphs.push_back(g);
        it = c->step_graphs.end() - 1;
    } else {
        c->flip_formed = c->P.flip_ratio >= -1 || check_cons;  // as enqueue_step records it
    }
    HIP_TRY(hipGraphLaunch(it->exec, c->stream));
    return 0;
}

// The loop of mpm_step: batches of steps enqueued without host synchronisation, the grid grown between
// batches. scratch == nullptr: the context's own store (allocated as needed). Otherwise the grid lives in
// the device scratch at `scratch`, batches are kNoAllocBatch steps (a smaller growth margin), and nothing
// is allocated.
constexpr uint32_t kNoAllocBatch = 8;
int step_loop(mpm_context* c, uint32_t nsteps, uint32_t flags, mpm_step_info* info, char* scratch) {
    uint32_t remaining = nsteps, done = 0, status = 0, reallocs = 0;
    const int D = c->cfg.dim;
    while (remaining > 0) {
        uint32_t batch = std::min(remaining, scratch ? kNoAllocBatch : kBatch);
        int rc = scratch ? bind_scratch(c, batch, scratch) : ensure_grid(c, batch);
        if (rc) { status |= rc; break; }
        uint32_t xsel = c->Sh->xsel;
        if (!scratch && !c->profiling && c->step_graphs_on && c->n <= kGraphMaxParticles && batch >= kGraphMinSteps) {
            rc = launch_step_graph(c, batch, xsel, flags);
            if (rc) { status |= rc; break; }
        } else {
            k_control<<<1, 64, 0, c->stream>>>(c->S, (flags & MPM_STEP_STOP_AT_FRAME_END) ? 1u : 0u,
                                               (flags & MPM_STEP_CHECK_CONSERVATION) ? 1u : 0u, c->cap_axis, (uint32_t)c->cap_nodes,
                                               c->act_blocks);
            for (uint32_t j = 0; j < batch; ++j) enqueue(c, xsel + j, (flags & MPM_STEP_CHECK_CONSERVATION) != 0, j, batch);
        }
        if (hipGetLastError() != hipSuccess) { status |= MPM_ERR_HIP; break; }
        rc = sync_state(c);
        if (rc) { status |= rc; break; }
        if (c->profiling) collect_timings(c);
        const StepState& S = *c->Sh;
        done += S.steps_done;
        status |= S.err | S.warn;
        remaining -= std::min(remaining, S.steps_done);
        for (int d = 0; d < D; ++d) c->est_n[d] = S.n[d];
        if (S.need_realloc) {  // next ensure_grid grows the capacity from S.n
            if (++reallocs > 8) { status |= MPM_ERR_OUT_OF_MEMORY; break; }
            continue;
        }
        reallocs = 0;
        if (S.stop) break;
        if (S.steps_done == 0) break;
    }
    const StepState& S = *c->Sh;
    if (info) {
        info->dt = S.dt; info->time = S.time; info->step = S.step; info->frame = S.frame;
        info->plastic_count = S.plastic_count; info->steps_done = done; info->status = status;
        info->frame_ended = (int32_t)S.frame_ended; info->finished = (int32_t)S.finished;
    }
    return (int)(status & MPM_ERR_EXIT_MASK);
}
}  // namespace

extern "C" {

int mpm_step(mpm_context* c, uint32_t nsteps, uint32_t flags, mpm_step_info* info) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    if (info) std::memset(info, 0, sizeof(*info));
    if (!c->ready) return MPM_ERR_NOT_READY;
    if (c->batch) return MPM_ERR_INVALID_ARGUMENT;  // a batch member advances with its batch
    if (hipSetDevice(c->device) != hipSuccess) return MPM_ERR_HIP;
    return step_loop(c, nsteps, flags, info, nullptr);
}

int mpm_step_noalloc(mpm_context* c, uint32_t nsteps, uint32_t flags, mpm_step_info* info) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    if (info) std::memset(info, 0, sizeof(*info));
    if (!c->ready) return MPM_ERR_NOT_READY;
    if (c->batch) return MPM_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(c->device) != hipSuccess) return MPM_ERR_HIP;
    char* base = nullptr;
    const int rc = scratch_address(c->device, &base);
    if (rc) { if (info) info->status = (uint32_t)rc; return rc; }
    // per-kernel timing would create HIP events: it is off for the call
    const bool prof = c->profiling;
    c->profiling = false;
    const int r = step_loop(c, nsteps, flags, info, base);
    c->profiling = prof;
    return r;
}

uint32_t mpm_num_particles(const mpm_context* c) { return c ? c->n : 0; }

int mpm_get_particles(mpm_context* c, mpm_particle_arrays* out) {
    if (!c || !out) return MPM_ERR_INVALID_ARGUMENT;
    if (!c->ready) return MPM_ERR_NOT_READY;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return c->split ? download_particles<float>(c, out) : download_particles<double>(c, out);
}

int mpm_get_grid(mpm_context* c, mpm_grid_info* info, double* x, double* y, double* z, double* mass, double* v,
                 double* flip, double* friction) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    if (!c->ready) return MPM_ERR_NOT_READY;
    HIP_TRY(hipStreamSynchronize(c->stream));
    const StepState& S = *c->Sh;
    const int D = c->cfg.dim;
    const uint32_t nn = S.grid_nodes;
    if (info) {
        info->Nx = S.n[0]; info->Ny = S.n[1]; info->Nz = D == 3 ? S.n[2] : 1; info->grid_nodes = nn;
        info->xc = S.origin[0]; info->yc = S.origin[1]; info->zc = D == 3 ? S.origin[2] : 0.0;
    }
    if (nn == 0 || !c->grid_bound) return 0;
    double* axes[3] = {x, y, z};
    for (int d = 0; d < D; ++d) if (axes[d]) HIP_TRY(hipMemcpy(axes[d], c->gp.coord[d], sizeof(double) * S.n[d], hipMemcpyDeviceToHost));
    return c->split ? download_grid<float>(c, nn, mass, v, flip, friction) : download_grid<double>(c, nn, mass, v, flip, friction);
}

int mpm_get_plates(mpm_context* c, mpm_plate* out, uint32_t cap, uint32_t* count) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    if (count) *count = (uint32_t)c->plates.size();
    const bool live = c->ready;
    for (uint32_t i = 0; i < c->plates.size() && i < cap && out; ++i) {
        const DevPlate& p = live ? c->Sh->plates[i] : c->plates[i];
        out[i].pos_object = p.pos_object; out[i].plate_type = p.plate_type; out[i].bc = p.bc; out[i].friction = p.friction;
        out[i].pos_lower = p.pos_lower; out[i].pos_upper = p.pos_upper;
        for (int d = 0; d < 3; ++d) out[i].v_object[d] = p.v_object[d];
        out[i].vmin_factor = p.vmin_factor; out[i].load_factor = p.load_factor;
    }
    return 0;
}

int mpm_get_binning(mpm_context* c, int32_t* cell_base, uint32_t* block_key) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    if (!c->ready || c->Sh->grid_nodes == 0) return MPM_ERR_NOT_READY;
    const int D = c->cfg.dim;
    const uint32_t n = c->n;
    HIP_TRY(hipStreamSynchronize(c->stream));
    int32_t* dbase = nullptr; uint32_t* dkey = nullptr;
    HIP_TRY(hipMalloc(&dbase, sizeof(int32_t) * (size_t)n * D));
    if (hipMalloc(&dkey, sizeof(uint32_t) * n) != hipSuccess) { (void)hipFree(dbase); return MPM_ERR_OUT_OF_MEMORY; }
    if (c->split) launch_binning<float>(c, n, dbase, dkey);
    else launch_binning<double>(c, n, dbase, dkey);
    const uint32_t* ids = c->split ? c->pf[c->Sh->xsel & 1].id : c->ps[c->Sh->xsel & 1].id;
    std::vector<int32_t> b((size_t)n * D); std::vector<uint32_t> kk(n), id(n);
    int rc = 0;
    if (hipStreamSynchronize(c->stream) != hipSuccess ||
        hipMemcpy(b.data(), dbase, sizeof(int32_t) * b.size(), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(kk.data(), dkey, sizeof(uint32_t) * n, hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(id.data(), ids, sizeof(uint32_t) * n, hipMemcpyDeviceToHost) != hipSuccess) rc = MPM_ERR_HIP;
    (void)hipFree(dbase); (void)hipFree(dkey);
    if (rc) return rc;
    for (uint32_t p = 0; p < n; ++p) {
        if (id[p] >= n) return MPM_ERR_HIP;
        if (cell_base) for (int d = 0; d < D; ++d) cell_base[(size_t)id[p] * D + d] = b[(size_t)p * D + d];
        if (block_key) block_key[id[p]] = kk[p];
    }
    return 0;
}

int mpm_get_stress_measures(mpm_context* c, double* p, double* q, double* Je) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    if (!c->ready) return MPM_ERR_NOT_READY;
    if (hipSetDevice(c->device) != hipSuccess) return MPM_ERR_HIP;
    HIP_TRY(hipStreamSynchronize(c->stream));
    const uint32_t n = c->n;
    double* d = nullptr;
    HIP_TRY(hipMalloc(&d, sizeof(double) * 3 * (size_t)n));
    const uint32_t blocks = (n + kBlock - 1) / kBlock;
    const uint32_t xs = c->Sh->xsel & 1;
    if (c->cfg.dim == 3) {
        if (c->split) k_stress_measures<3, float><<<blocks, kBlock, 0, c->stream>>>(c->P, c->pf[xs], n, d, d + n, d + 2 * (size_t)n);
        else k_stress_measures<3, double><<<blocks, kBlock, 0, c->stream>>>(c->P, c->ps[xs], n, d, d + n, d + 2 * (size_t)n);
    } else {
        if (c->split) k_stress_measures<2, float><<<blocks, kBlock, 0, c->stream>>>(c->P, c->pf[xs], n, d, d + n, d + 2 * (size_t)n);
        else k_stress_measures<2, double><<<blocks, kBlock, 0, c->stream>>>(c->P, c->ps[xs], n, d, d + n, d + 2 * (size_t)n);
    }
    int rc = 0;
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess) rc = MPM_ERR_HIP;
    double* outs[3] = {p, q, Je};
    for (int k = 0; k < 3 && !rc; ++k)
        if (outs[k] && hipMemcpy(outs[k], d + (size_t)k * n, sizeof(double) * n, hipMemcpyDeviceToHost) != hipSuccess) rc = MPM_ERR_HIP;
    (void)hipFree(d);
    return rc;
}


int mpm_write_particles_ply(mpm_context* c, const char* path) {
    if (!c || !path) return MPM_ERR_INVALID_ARGUMENT;
    if (!c->ready) return MPM_ERR_NOT_READY;
    const int D = c->cfg.dim;
    const size_t n = c->n;
    std::vector<double> x(n * D), v(n * D), sc[6], pq[3];
    for (auto& a : sc) a.resize(n);
    for (auto& a : pq) a.resize(n);
    mpm_particle_arrays a{};
    a.x = x.data(); a.v = v.data(); a.eps_pl_dev = sc[0].data(); a.eps_pl_vol = sc[1].data(); a.delta_gamma = sc[3].data();
    a.viscosity = sc[4].data(); a.muI = sc[5].data();
    int rc = mpm_get_particles(c, &a);
    if (!rc) rc = mpm_get_stress_measures(c, pq[0].data(), pq[1].data(), pq[2].data());
    if (rc) return rc;
    std::vector<std::string> names;
    std::vector<const double*> cols;
    std::vector<int> stride;
    const char* ax[3] = {"x", "y", "z"};
    const char* av[3] = {"vx", "vy", "vz"};
    for (int d = 0; d < D; ++d) { names.push_back(ax[d]); cols.push_back(x.data() + d); stride.push_back(D); }
    for (int d = 0; d < D; ++d) { names.push_back(av[d]); cols.push_back(v.data() + d); stride.push_back(D); }
    const char* pn[3] = {"p", "q", "Je"};
    for (int k = 0; k < 3; ++k) { names.push_back(pn[k]); cols.push_back(pq[k].data()); stride.push_back(1); }
    const int pm = c->cfg.plastic_model;
    if (pm != MPM_PLASTIC_NONE) {
        names.push_back("eps_pl_vol"); cols.push_back(sc[1].data()); stride.push_back(1);
        names.push_back("eps_pl_dev"); cols.push_back(sc[0].data()); stride.push_back(1);
        names.push_back("delta_gamma"); cols.push_back(sc[3].data()); stride.push_back(1);
        const bool mui = pm == MPM_PLASTIC_DPMUI || pm == MPM_PLASTIC_MCCMUI;
        if (((pm == MPM_PLASTIC_DPVISC || pm == MPM_PLASTIC_MCCVISC) && c->cfg.use_mibf) || mui) {
            names.push_back("muI"); cols.push_back(sc[5].data()); stride.push_back(1);
        }
        if (mui) { names.push_back("viscosity"); cols.push_back(sc[4].data()); stride.push_back(1); }
    }
    return write_ply(path, n, names, cols, stride);
}

int mpm_write_grid_ply(mpm_context* c, const char* path) {
    if (!c || !path) return MPM_ERR_INVALID_ARGUMENT;
    mpm_grid_info gi{};
    int rc = mpm_get_grid(c, &gi, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr);
    if (rc) return rc;
    const int D = c->cfg.dim;
    const size_t nn = gi.grid_nodes;
    std::vector<double> ax[3] = {std::vector<double>(gi.Nx), std::vector<double>(gi.Ny), std::vector<double>(gi.Nz)};
    std::vector<double> mass(nn), v(nn * D), fric(nn);
    rc = mpm_get_grid(c, &gi, ax[0].data(), ax[1].data(), D == 3 ? ax[2].data() : nullptr, mass.data(), v.data(), nullptr,
                      c->cfg.use_mibf ? fric.data() : nullptr);
    if (rc) return rc;
    std::vector<double> X(nn * D);  // node coordinates in ind(i,j,k) order, as saveGridData builds them
    size_t q = 0;
    for (uint32_t i = 0; i < gi.Nx; ++i)
        for (uint32_t j = 0; j < gi.Ny; ++j)
            for (uint32_t k = 0; k < (D == 3 ? gi.Nz : 1u); ++k, ++q) {
                X[q * D] = ax[0][i]; X[q * D + 1] = ax[1][j];
                if (D == 3) X[q * D + 2] = ax[2][k];
            }
    std::vector<std::string> names;
    std::vector<const double*> cols;
    std::vector<int> stride;
    const char* an[3] = {"x", "y", "z"};
    const char* avn[3] = {"vx", "vy", "vz"};
    for (int d = 0; d < D; ++d) { names.push_back(an[d]); cols.push_back(X.data() + d); stride.push_back(D); }
    for (int d = 0; d < D; ++d) { names.push_back(avn[d]); cols.push_back(v.data() + d); stride.push_back(D); }
    names.push_back("mass"); cols.push_back(mass.data()); stride.push_back(1);
    if (c->cfg.use_mibf) { names.push_back("friction"); cols.push_back(fric.data()); stride.push_back(1); }
    return write_ply(path, nn, names, cols, stride);
}

int mpm_set_profiling(mpm_context* c, int enable) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->ev_pending.clear();
    for (int k = 0; k < KS_COUNT; ++k) { c->kt_ms[k] = 0; c->kt_n[k] = 0; }
    c->profiling = enable != 0;
    c->prof_mask = ~0u;
    return 0;
}

int mpm_set_step_graphs(mpm_context* c, int enable) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    HIP_TRY(hipStreamSynchronize(c->stream));
    c->step_graphs_on = enable != 0;
    return 0;
}

int mpm_set_profiling_kernel(mpm_context* c, const char* name) {
    if (!c || !name) return MPM_ERR_INVALID_ARGUMENT;
    int id = -1;
    for (int k = 0; k < KS_COUNT; ++k) if (std::strcmp(name, kKernelNames[k]) == 0) id = k;
    if (id < 0) return MPM_ERR_INVALID_ARGUMENT;
    const int rc = mpm_set_profiling(c, 1);
    if (rc) return rc;
    c->prof_mask = 1u << id;
    return 0;
}

int mpm_get_kernel_times(mpm_context* c, mpm_kernel_timing* out, uint32_t cap, uint32_t* count) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    uint32_t n = 0;
    for (int k = 0; k < KS_COUNT; ++k) {
        if (!c->kt_n[k]) continue;
        if (out && n < cap) {
            std::memset(out[n].name, 0, sizeof(out[n].name));
            std::strncpy(out[n].name, kKernelNames[k], sizeof(out[n].name) - 1);
            out[n].total_ms = c->kt_ms[k]; out[n].launches = c->kt_n[k];
        }
        ++n;
    }
    if (count) *count = n;
    return 0;
}

const char* mpm_status_string(uint32_t s) {
    if (s == 0) return "ok";
    static const char* names[] = {"invalid argument", "unsupported option", "HIP runtime error", "out of memory",
                                  "negative Perzyna delta_gamma", "Newton iteration limit", "Newton derivative too small",
                                  "MCC return map found no admissible step", "stencil outside the grid", "non-finite state",
                                  "conservation check failed", "particles not set",
                                  "grid of the no-allocation step exceeds the device scratch"};
    for (int b = 0; b < 13; ++b) if (s & (1u << b)) return names[b];
    if (s & MPM_WARN_SUPERSONIC) return "particle faster than the elastic wave speed";
    return "unknown status";
}

int mpm_abi_version(void) { return MPM_ABI_VERSION; }

int mpm_batch_create(const mpm_batch_options* opt, mpm_batch** out) {
    if (!out) return MPM_ERR_INVALID_ARGUMENT;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return MPM_ERR_HIP;
    mpm_batch* b = new (std::nothrow) mpm_batch();
    if (!b) return MPM_ERR_OUT_OF_MEMORY;
    b->opt.chunk_steps = 16; b->opt.use_graph = 1;
    if (opt) {
        if (opt->chunk_steps) b->opt.chunk_steps = std::min<uint32_t>(opt->chunk_steps, 256);
        b->opt.use_graph = opt->use_graph;
    }
    (void)hipGetDevice(&b->device);
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return MPM_ERR_HIP; }
    *out = b;
    return 0;
}

void mpm_batch_destroy(mpm_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (mpm_context* c : b->scenes) c->batch = nullptr;
    drop_graph(b);
    if (b->dtab) (void)hipFree(b->dtab);
    if (b->hlimit) (void)hipHostFree(b->hlimit);
    if (b->hhead) (void)hipHostFree(b->hhead);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

int mpm_batch_add(mpm_batch* b, mpm_context* c, uint32_t nsteps) {
    if (!b || !c || c->batch) return MPM_ERR_INVALID_ARGUMENT;
    if (!c->ready) return MPM_ERR_NOT_READY;
    if (c->device != b->device) return MPM_ERR_INVALID_ARGUMENT;
    if (hipSetDevice(b->device) != hipSuccess) return MPM_ERR_HIP;
    HIP_TRY(hipStreamSynchronize(c->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    if (!c->dP && hipMalloc(&c->dP, sizeof(Params)) != hipSuccess) { c->dP = nullptr; return MPM_ERR_OUT_OF_MEMORY; }
    c->batch = b; c->budget = nsteps;
    std::memset(&c->last, 0, sizeof(c->last));
    b->scenes.push_back(c);
    b->dirty = true;
    return 0;
}

int mpm_batch_remove(mpm_batch* b, mpm_context* c) {
    if (!b || !c || c->batch != b) return MPM_ERR_INVALID_ARGUMENT;
    (void)hipSetDevice(b->device);
    HIP_TRY(hipStreamSynchronize(b->stream));
    b->scenes.erase(std::remove(b->scenes.begin(), b->scenes.end(), c), b->scenes.end());
    c->batch = nullptr;
    b->dirty = true;
    return 0;
}

int mpm_batch_set_budget(mpm_batch* b, mpm_context* c, uint32_t nsteps) {
    if (!b || !c || c->batch != b) return MPM_ERR_INVALID_ARGUMENT;
    c->budget = nsteps;
    return 0;
}

int mpm_batch_step(mpm_batch* b, uint32_t max_steps, uint32_t flags, mpm_batch_info* info) {
    if (!b) return MPM_ERR_INVALID_ARGUMENT;
    if (info) std::memset(info, 0, sizeof(*info));
    if (hipSetDevice(b->device) != hipSuccess) return MPM_ERR_HIP;
    const bool cons = (flags & MPM_STEP_CHECK_CONSERVATION) != 0;
    uint32_t status = 0, chunks = 0, reallocs = 0;
    uint64_t particle_steps = 0;
    // steps each member is to take in this call
    std::vector<uint32_t> want(b->scenes.size()), done(b->scenes.size(), 0);
    for (size_t i = 0; i < b->scenes.size(); ++i) {
        mpm_context* c = b->scenes[i];
        want[i] = std::min(max_steps, c->budget);
        std::memset(&c->last, 0, sizeof(c->last));
    }
    auto index_of = [&](const mpm_context* c) { return (size_t)(std::find(b->scenes.begin(), b->scenes.end(), c) - b->scenes.begin()); };
    while (!b->scenes.empty()) {
        uint32_t chunk = 0;
        for (size_t i = 0; i < b->scenes.size(); ++i) chunk = std::max(chunk, want[i] - done[i]);
        chunk = std::min(chunk, b->opt.chunk_steps);
        if (chunk == 0) break;
        // grid capacity for the chunk; a reallocation changes the member's buffers and so the tables
        for (size_t i = 0; i < b->scenes.size(); ++i) {
            mpm_context* c = b->scenes[i];
            const uint32_t lim = std::min(chunk, want[i] - done[i]);
            if (!lim && c->act_blocks && c->cap_nodes) continue;  // idle member whose buffers exist
            // the views and launch shapes a class table caches: any rebinding of the store bumps grid_gen
            const uint64_t gen_before = c->grid_gen;
            const uint32_t act_before = c->act_blocks;
            const int rc = ensure_grid(c, std::max(lim, 1u));
            if (rc) { status |= rc; break; }
            if (gen_before != c->grid_gen || act_before != c->act_blocks) b->dirty = true;
        }
        if (status) break;
        if (b->dirty || b->check_cons != cons) {
            const int rc = rebuild_tables(b, cons);
            if (rc) { status |= rc; break; }
        }
        for (size_t r = 0; r < b->order.size(); ++r) {
            const size_t i = index_of(b->order[r]);
            b->hlimit[r] = std::min(chunk, want[i] - done[i]);
        }
        const size_t ns = b->order.size();
        if (hipMemcpyAsync(b->dlimit, b->hlimit, sizeof(uint32_t) * ns, hipMemcpyHostToDevice, b->stream) != hipSuccess) { status |= MPM_ERR_HIP; break; }
        int rc = launch_chunk(b, chunk, flags);
        if (rc) { status |= rc; break; }
        if (hipMemcpyAsync(b->hhead, b->dhead, (size_t)kHeadBytes * ns, hipMemcpyDeviceToHost, b->stream) != hipSuccess ||
            hipStreamSynchronize(b->stream) != hipSuccess) { status |= MPM_ERR_HIP; break; }
        ++chunks;
        bool realloc = false;
        for (size_t r = 0; r < ns; ++r) {
            mpm_context* c = b->order[r];
            const size_t i = index_of(c);
            std::memcpy((void*)c->Sh, (const char*)b->hhead + (size_t)kHeadBytes * r, kHeadBytes);
            if (b->hlimit[r] == 0) continue;  // idle in this chunk
            const StepState& S = *c->Sh;
            done[i] += S.steps_done;
            c->budget -= std::min(c->budget, S.steps_done);
            particle_steps += (uint64_t)c->n * S.steps_done;
            mpm_step_info& L = c->last;
            L.steps_done += S.steps_done; L.status |= S.err | S.warn;
            L.dt = S.dt; L.time = S.time; L.step = S.step; L.frame = S.frame; L.plastic_count = S.plastic_count;
            L.frame_ended = (int32_t)S.frame_ended; L.finished = (int32_t)S.finished;
            for (int d = 0; d < c->cfg.dim; ++d) c->est_n[d] = S.n[d];
            if (S.need_realloc) { realloc = true; continue; }
            const bool ended = S.finished || (S.err & MPM_ERR_EXIT_MASK);
            if (ended) c->budget = 0;
            if (ended || (S.stop && S.frame_ended && S.stop_at_frame_end) || (S.steps_done < S.step_limit && S.stop)) want[i] = done[i];
        }
        if (realloc) {
            if (++reallocs > 8) { status |= MPM_ERR_OUT_OF_MEMORY; break; }
        } else {
            reallocs = 0;
        }
    }
    if (info) {
        info->n_scenes = (uint32_t)b->scenes.size();
        for (mpm_context* c : b->scenes) info->n_active += c->budget > 0 ? 1u : 0u;
        info->n_classes = (uint32_t)b->classes.size();
        info->chunks = chunks; info->status = status; info->particle_steps = particle_steps;
    }
    return (int)(status & MPM_ERR_EXIT_MASK);
}

int mpm_batch_autotune(mpm_batch* b, uint32_t max_steps, uint32_t* chosen, double* rate) {
    if (!b) return MPM_ERR_INVALID_ARGUMENT;
    const uint32_t cand[4] = {4, 8, 16, 32};
    const uint32_t keep = b->opt.chunk_steps;
    uint32_t best = keep, used = 0;
    double best_rate = 0.0;
    for (uint32_t k : cand) {
        if (used + 3 * k > max_steps) break;
        b->opt.chunk_steps = k;
        mpm_batch_info bi{};
        int rc = mpm_batch_step(b, k, 0, &bi);  // warm-up chunk: captures this length's graph
        used += k;
        if (rc) { b->opt.chunk_steps = keep; return rc; }
        const auto t0 = std::chrono::steady_clock::now();
        rc = mpm_batch_step(b, 2 * k, 0, &bi);
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        used += 2 * k;
        if (rc) { b->opt.chunk_steps = keep; return rc; }
        if (bi.particle_steps == 0) break;  // budgets spent
        const double r = (double)bi.particle_steps / sec;
        if (r > best_rate) { best_rate = r; best = k; }
    }
    b->opt.chunk_steps = best;
    if (chosen) *chosen = best;
    if (rate) *rate = best_rate;
    return 0;
}

int mpm_scene_info(const mpm_context* c, mpm_step_info* info, uint32_t* budget_left) {
    if (!c) return MPM_ERR_INVALID_ARGUMENT;
    if (info) *info = c->last;
    if (budget_left) *budget_left = c->budget;
    return 0;
}

uint64_t mpm_device_bytes(const mpm_context* c) {
    if (!c) return 0;
    uint64_t bytes = c->pbuf_bytes + c->ibuf_n * sizeof(uint32_t);
    const StoreBytes sz = store_bytes(c, c->own.cap_blocks, c->own.cap_nodes, c->own.cap_axis);  // the scratch is not the context's
    bytes += sz.bbuf + sz.cellbuf + sz.stgbuf + sz.gbuf + sz.cbuf;
    bytes += sizeof(StepState) + (c->dP ? sizeof(Params) : 0);
    return bytes;
}

}  // extern "C"
