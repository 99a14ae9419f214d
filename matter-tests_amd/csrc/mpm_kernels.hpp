// mpm_kernels.hpp — gfx950 kernels of one MPM step (plain FP64 form).
//
// Step pipeline (one HIP stream, no host synchronisation inside a batch of steps):
//   k_setup        1 workgroup: dt (updateDt), grid geometry (remeshFixedInit/Cont + arange), capacity check
//   k_zero_grid    clear the node accumulators of the dense grid (resizeGrid)
//   k_p2g          scatter mass, APIC momentum, MIBF friction and stress force  (P2G + force part of
//                  explicitEulerUpdate)
//   k_grid_update  normalise, explicit update, boundary conditions, flip       (P2G tail + grid loop of
//                  explicitEulerUpdate + boundaryCollision)
//   USL:  k_g2p_fused    gather, F update, return map, position/velocity update  (G2P + deformationUpdate
//                        + plasticity + positionUpdate), reductions for the next step
//   MUSL: k_g2p_musl     gather, velocity update, position to x_next, re-scatter momentum (G2P + MUSL)
//         k_musl_norm    divide by m_i/m_p                                         (MUSL tail)
//         k_defgrad      gather grad v, F update, return map                       (deformationUpdate)
//   k_finish       1 thread: moveObjects, time/frame bookkeeping of simulate()
// Every kernel returns at once when StepState::stop is set, so the host can enqueue many steps and
// let the device stop at an error, a frame end, the final time or a capacity overflow.
#pragma once
#include <hip/hip_runtime.h>
#include "mpm_math.hpp"

namespace mpm {

struct PartPtrs {
    double* x[3];
    double* xn[3];   // MUSL: positions after the step
    double* v[3];
    double* B[9];
    double* F[9];
    double* tau[6];
    double* eps_dev; double* eps_vol; double* eps_prad; double* dgamma; double* visc; double* muI;
    uint32_t n;
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
    if ((threadIdx.x & 63) == 0) {
        atomicMax(&S->vmax2_bits, v2);
        for (int d = 0; d < D; ++d) { atomicMin(&S->xmin_key[d], mn[d]); atomicMax(&S->xmax_key[d], mx[d]); }
    }
}

// Stencil of one particle: base node, per-axis weights/gradient factors and node coordinates.
template <int D, int SPL>
struct Stencil {
    int base[3];
    int lo[3];
    double w[3][4], dw[3][4], xi[3][4];
    static constexpr int cnt = SPL == 2 ? 3 : 4;
};

// Builds the stencil; clamp=true reproduces the max(0, .) of p2g.cpp:21-24, clamp=false the plain
// floor of g2p.cpp:29-32. Returns false if the 4-node window leaves the grid.
template <int D, int SPL>
__device__ __forceinline__ bool make_stencil(Stencil<D, SPL>& st, const double* xp, const double* org, const uint32_t* n,
                                             double* const* coord, double h1, bool clamp) {
    bool ok = true;
    for (int d = 0; d < D; ++d) {
        int b = (int)floor((xp[d] - org[d]) * h1) - 1;
        if (clamp && b < 0) b = 0;
        if (b < 0 || b + 3 >= (int)n[d]) ok = false;
        st.base[d] = b;
    }
    if (!ok) return false;
    for (int d = 0; d < D; ++d) {
        for (int a = 0; a < 4; ++a) {
            double c = coord[d][st.base[d] + a];
            double t = (xp[d] - c) * h1;
            st.xi[d][a] = c;
            st.w[d][a] = BSpline<SPL>::N(t);
            st.dw[d][a] = BSpline<SPL>::dN(t);
        }
        // quadratic: exactly one of the end nodes has zero weight (both at f = 0.5)
        st.lo[d] = (SPL == 2 && st.w[d][0] == 0.0) ? 1 : 0;
    }
    return true;
}

// ----------------------------------------------------------------------------------------------
// k_prepare: Kirchhoff stress of the initial F and the reductions that the first step needs.
// ----------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_prepare(Params P, StepState* S, PartPtrs pp) {
    uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    bool active = p < pp.n;
    double x[3] = {0, 0, 0}, v[3] = {0, 0, 0};
    if (active) {
        M<D> F;
        for (int i = 0; i < D * D; ++i) F.m[i] = pp.F[i][p];
        double tau[6];
        if (P.elastic_model == MPM_ELASTIC_NEOHOOKEAN) neohookean_tau<D>(F, P.mu, P.lambda, tau);
        else {
            M<D> U, W; V<D> s; svd<D>(F, U, s, W);
            V<D> ls; for (int i = 0; i < D; ++i) ls[i] = log(fabs(s[i]));
            hencky_tau<D>(U, ls, P.mu, P.lambda, tau);
        }
        for (int i = 0; i < SymPack<D>::n; ++i) pp.tau[i][p] = tau[i];
        for (int d = 0; d < D; ++d) { x[d] = pp.x[d][p]; v[d] = pp.v[d][p]; }
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
            bool finite = true;
            for (int d = 0; d < D; ++d) if (!(lo[d] == lo[d]) || !(hi[d] == hi[d]) || !isfinite(lo[d]) || !isfinite(hi[d]) || cnt[d] < 1) finite = false;
            if (!(dt == dt)) finite = false;
            if (!finite) {
                S->err |= MPM_ERR_NONFINITE; S->stop = 1;
            } else if (!fits) {
                S->need_realloc = 1; S->stop = 1;
                for (int d = 0; d < D; ++d) S->n[d] = cnt[d];
            } else {
                S->dt = dt; S->warn |= warn;
                for (int d = 0; d < 3; ++d) S->gravity[d] = grav[d];
                for (int d = 0; d < D; ++d) {
                    S->min_init[d] = mni[d]; S->max_init[d] = mxi[d]; S->low_init[d] = loi[d]; S->high_init[d] = hii[d];
                    S->origin[d] = lo[d]; S->n[d] = cnt[d];
                }
                if (D == 2) { S->n[2] = 1; S->origin[2] = 0.0; }
                S->grid_nodes = (uint32_t)nodes;
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
__global__ __launch_bounds__(kBlock) void k_zero_grid(Params P, const StepState* S, GridPtrs g) {
    if (S->stop) return;
    const uint32_t nn = S->grid_nodes;
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nn; i += gridDim.x * blockDim.x) {
        g.mass[i] = 0.0;
        for (int d = 0; d < D; ++d) { g.mv[d][i] = 0.0; g.f[d][i] = 0.0; }
        if (P.use_mibf) g.fric[i] = 0.0;
        if (P.use_musl) for (int d = 0; d < D; ++d) g.mv2[d][i] = 0.0;
    }
}

__device__ __forceinline__ uint32_t node_index3(int i, int j, int k, const uint32_t* n) { return ((uint32_t)i * n[1] + (uint32_t)j) * n[2] + (uint32_t)k; }

// ----------------------------------------------------------------------------------------------
// k_p2g: p2g.cpp:12-66 (mass, APIC momentum, MIBF friction; weight > 1e-25) fused with the stress
// part of explicit_euler_update.cpp:15-168 (tau grad w). The reference adds the force only where
// m_i > 0; nodes with m_i = 0 are never updated or read, so the force is scattered unconditionally.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL>
__global__ __launch_bounds__(kBlock) void k_p2g(Params P, StepState* S, PartPtrs pp, GridPtrs g) {
    if (S->stop) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= pp.n) return;
    const double h1 = P.one_over_dx;
    double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
    uint32_t n[3] = {S->n[0], S->n[1], S->n[2]};
    double xp[3] = {0, 0, 0}, vp[3] = {0, 0, 0};
    for (int d = 0; d < D; ++d) { xp[d] = pp.x[d][p]; vp[d] = pp.v[d][p]; }
    Stencil<D, SPL> st;
    if (!make_stencil<D, SPL>(st, xp, org, n, g.coord, h1, true)) { atomicOr(&S->err, MPM_ERR_GRID_BOUNDS); return; }
    const bool apic = P.flip_ratio < 0;
    double Bm[9];
    if (apic) for (int i = 0; i < D * D; ++i) Bm[i] = pp.B[i][p];
    double t6[6]; for (int i = 0; i < SymPack<D>::n; ++i) t6[i] = pp.tau[i][p];
    double tau[3][3];
    if (D == 3) { tau[0][0] = t6[0]; tau[1][1] = t6[1]; tau[2][2] = t6[2]; tau[0][1] = tau[1][0] = t6[3]; tau[0][2] = tau[2][0] = t6[4]; tau[1][2] = tau[2][1] = t6[5]; }
    else { tau[0][0] = t6[0]; tau[1][1] = t6[1]; tau[0][1] = tau[1][0] = t6[2]; }
    const double muI = P.use_mibf ? pp.muI[p] : 0.0;
    const double aD = P.apicDinverse;
    const int c2 = (D == 3) ? Stencil<D, SPL>::cnt : 1;
    for (int ia = 0; ia < Stencil<D, SPL>::cnt; ++ia) {
        const int a = st.lo[0] + ia;
        for (int ib = 0; ib < Stencil<D, SPL>::cnt; ++ib) {
            const int b = st.lo[1] + ib;
            for (int ic = 0; ic < c2; ++ic) {
                const int c = (D == 3) ? st.lo[2] + ic : 0;
                double weight = st.w[0][a] * st.w[1][b];
                if (D == 3) weight = weight * st.w[2][c];
                uint32_t node = (D == 3) ? node_index3(st.base[0] + a, st.base[1] + b, st.base[2] + c, n)
                                         : (uint32_t)(st.base[0] + a) * n[1] + (uint32_t)(st.base[1] + b);
                if (weight > 1e-25) {
                    atomic_add_f64(&g.mass[node], weight);
                    double pd[3] = {st.xi[0][a] - xp[0], st.xi[1][b] - xp[1], D == 3 ? st.xi[2][c] - xp[2] : 0.0};
                    for (int r = 0; r < D; ++r) {
                        double m = vp[r] * weight;
                        if (apic) {
                            double bd = Bm[r] * pd[0];
                            for (int k = 1; k < D; ++k) bd += Bm[r + k * D] * pd[k];
                            m += bd * aD * weight;
                        }
                        atomic_add_f64(&g.mv[r][node], m);
                    }
                    if (P.use_mibf) atomic_add_f64(&g.fric[node], muI * weight);
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
                    atomic_add_f64(&g.f[r][node], fr);
                }
            }
        }
    }
    if (S->check_conservation) {
        for (int d = 0; d < D; ++d) atomic_add_f64(&S->cons_part_mom[d], P.particle_mass * vp[d]);
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

// F <- F + (dt*grad v) * F, return map, Kirchhoff stress for the next step; writes F, tau and the
// plastic scalars of particle p (deformation_update.cpp:13-46, plasticity.cpp:11-665).
template <int D>
__device__ __forceinline__ void deform_and_project(const Params& P, double dt, const PartPtrs& pp, uint32_t p,
                                                   const double (&gv)[3][3], uint32_t& err, uint32_t& pcount) {
    M<D> F, A;
    for (int i = 0; i < D * D; ++i) F.m[i] = pp.F[i][p];
    for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) A(r, c) = dt * gv[r][c];
    M<D> AF = mul<D>(A, F);
    for (int i = 0; i < D * D; ++i) F.m[i] = F.m[i] + AF.m[i];
    PlasticState ps;
    ps.eps_dev = pp.eps_dev[p]; ps.eps_vol = pp.eps_vol[p]; ps.eps_prad = pp.eps_prad[p];
    ps.dgamma = 0.0; ps.visc = pp.visc[p]; ps.muI = pp.muI[p];
    double tau[6];
    if (P.elastic_model == MPM_ELASTIC_NEOHOOKEAN) {
        neohookean_tau<D>(F, P.mu, P.lambda, tau);
    } else {
        M<D> U, W; V<D> s;
        svd<D>(F, U, s, W);
        V<D> ls;
        if (P.plastic_model != MPM_PLASTIC_NONE) {
            V<D> h; for (int i = 0; i < D; ++i) h[i] = log(fmax(fabs(s[i]), 1e-4));
            RMResult r = return_map<D>(P, dt, h, ps);
            err |= r.err;
            pcount += r.plastic_incr;
            if (r.recompose) {
                V<D> e; for (int i = 0; i < D; ++i) e[i] = exp(h[i]);
                F = udvt<D>(U, e, W);
                for (int i = 0; i < D; ++i) ls[i] = log(e[i]);
            } else {
                for (int i = 0; i < D; ++i) ls[i] = log(fabs(s[i]));
            }
        } else {
            for (int i = 0; i < D; ++i) ls[i] = log(fabs(s[i]));
        }
        hencky_tau<D>(U, ls, P.mu, P.lambda, tau);
    }
    for (int i = 0; i < D * D; ++i) pp.F[i][p] = F.m[i];
    for (int i = 0; i < SymPack<D>::n; ++i) pp.tau[i][p] = tau[i];
    pp.eps_dev[p] = ps.eps_dev; pp.eps_vol[p] = ps.eps_vol; pp.eps_prad[p] = ps.eps_prad;
    pp.dgamma[p] = ps.dgamma; pp.visc[p] = ps.visc; pp.muI[p] = ps.muI;
}

// ----------------------------------------------------------------------------------------------
// k_g2p_fused (USL): G2P (g2p.cpp:17-77), deformationUpdate + plasticity, positionUpdate.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL>
__global__ __launch_bounds__(kBlock) void k_g2p_fused(Params P, StepState* S, PartPtrs pp, GridPtrs g) {
    if (S->stop) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = p < pp.n;
    uint32_t err = 0, pcount = 0;
    double xnew[3] = {0, 0, 0}, vnew[3] = {0, 0, 0};
    bool live = false;
    if (active) {
        const double h1 = P.one_over_dx, dt = S->dt;
        double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
        uint32_t n[3] = {S->n[0], S->n[1], S->n[2]};
        double xp[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) xp[d] = pp.x[d][p];
        Stencil<D, SPL> st;
        if (!make_stencil<D, SPL>(st, xp, org, n, g.coord, h1, false)) {
            err |= MPM_ERR_GRID_BOUNDS;
        } else {
            live = true;
            const bool apic = P.flip_ratio < 0, useflip = P.flip_ratio >= -1;
            double pic[3] = {0, 0, 0}, fl[3] = {0, 0, 0}, Bp[3][3] = {{0}}, gv[3][3] = {{0}};
            const int c2 = (D == 3) ? Stencil<D, SPL>::cnt : 1;
            for (int ia = 0; ia < Stencil<D, SPL>::cnt; ++ia) {
                const int a = st.lo[0] + ia;
                for (int ib = 0; ib < Stencil<D, SPL>::cnt; ++ib) {
                    const int b = st.lo[1] + ib;
                    for (int ic = 0; ic < c2; ++ic) {
                        const int c = (D == 3) ? st.lo[2] + ic : 0;
                        double weight = st.w[0][a] * st.w[1][b];
                        if (D == 3) weight = weight * st.w[2][c];
                        uint32_t node = (D == 3) ? node_index3(st.base[0] + a, st.base[1] + b, st.base[2] + c, n)
                                                 : (uint32_t)(st.base[0] + a) * n[1] + (uint32_t)(st.base[1] + b);
                        double vi[3];
                        for (int d = 0; d < D; ++d) vi[d] = g.v[d][node];
                        double pd[3] = {st.xi[0][a] - xp[0], st.xi[1][b] - xp[1], D == 3 ? st.xi[2][c] - xp[2] : 0.0};
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
                            pic[r] += vi[r] * weight;
                            if (apic) for (int k = 0; k < D; ++k) Bp[r][k] += vi[r] * pd[k] * weight;
                            for (int k = 0; k < D; ++k) gv[r][k] += vi[r] * gr[k];
                        }
                        if (useflip) for (int r = 0; r < D; ++r) fl[r] += g.flip[r][node] * weight;
                    }
                }
            }
            // velocity (positionUpdate, USL) and B
            for (int d = 0; d < D; ++d) {
                double vo = useflip ? pp.v[d][p] : 0.0;
                vnew[d] = blend_velocity(P.flip_ratio, vo, fl[d], pic[d]);
            }
            if (apic) for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) pp.B[r + c * D][p] = Bp[r][c];
            deform_and_project<D>(P, dt, pp, p, gv, err, pcount);
            for (int d = 0; d < D; ++d) xnew[d] = xp[d] + dt * pic[d];
            if (P.change_particle_positions && xnew[0] > 0.8) { xnew[0] = -0.1; xnew[1] += 0.27; }
            for (int d = 0; d < D; ++d) { pp.x[d][p] = xnew[d]; pp.v[d][p] = vnew[d]; }
        }
    }
    reduce_step_bounds<D>(S, live, xnew, vnew);
    pcount = wave_sum_u32(pcount);
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0) {
        if (pcount) atomicAdd(&S->plastic_count, pcount);
        if (err) atomicOr(&S->err, err);
    }
}

// ----------------------------------------------------------------------------------------------
// k_g2p_musl: G2P (g2p.cpp:17-77), MUSL velocity update and re-scatter (musl.cpp:13-85), position
// update of positionUpdate (position_update.cpp:12) into x_next.
// ----------------------------------------------------------------------------------------------
template <int D, int SPL>
__global__ __launch_bounds__(kBlock) void k_g2p_musl(Params P, StepState* S, PartPtrs pp, GridPtrs g) {
    if (S->stop) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = p < pp.n;
    uint32_t err = 0;
    double xnew[3] = {0, 0, 0}, vnew[3] = {0, 0, 0};
    bool live = false;
    if (active) {
        const double h1 = P.one_over_dx, dt = S->dt;
        double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
        uint32_t n[3] = {S->n[0], S->n[1], S->n[2]};
        double xp[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) xp[d] = pp.x[d][p];
        Stencil<D, SPL> st;
        if (!make_stencil<D, SPL>(st, xp, org, n, g.coord, h1, false)) {
            err |= MPM_ERR_GRID_BOUNDS;
        } else {
            live = true;
            const bool apic = P.flip_ratio < 0, useflip = P.flip_ratio >= -1;
            double pic[3] = {0, 0, 0}, fl[3] = {0, 0, 0}, Bp[3][3] = {{0}};
            const int c2 = (D == 3) ? Stencil<D, SPL>::cnt : 1;
            for (int ia = 0; ia < Stencil<D, SPL>::cnt; ++ia) {
                const int a = st.lo[0] + ia;
                for (int ib = 0; ib < Stencil<D, SPL>::cnt; ++ib) {
                    const int b = st.lo[1] + ib;
                    for (int ic = 0; ic < c2; ++ic) {
                        const int c = (D == 3) ? st.lo[2] + ic : 0;
                        double weight = st.w[0][a] * st.w[1][b];
                        if (D == 3) weight = weight * st.w[2][c];
                        uint32_t node = (D == 3) ? node_index3(st.base[0] + a, st.base[1] + b, st.base[2] + c, n)
                                                 : (uint32_t)(st.base[0] + a) * n[1] + (uint32_t)(st.base[1] + b);
                        double pd[3] = {st.xi[0][a] - xp[0], st.xi[1][b] - xp[1], D == 3 ? st.xi[2][c] - xp[2] : 0.0};
                        for (int r = 0; r < D; ++r) {
                            double vi = g.v[r][node];
                            pic[r] += vi * weight;
                            if (apic) for (int k = 0; k < D; ++k) Bp[r][k] += vi * pd[k] * weight;
                        }
                        if (useflip) for (int r = 0; r < D; ++r) fl[r] += g.flip[r][node] * weight;
                    }
                }
            }
            for (int d = 0; d < D; ++d) {
                double vo = useflip ? pp.v[d][p] : 0.0;
                vnew[d] = blend_velocity(P.flip_ratio, vo, fl[d], pic[d]);
                xnew[d] = xp[d] + dt * pic[d];
            }
            if (P.change_particle_positions && xnew[0] > 0.8) { xnew[0] = -0.1; xnew[1] += 0.27; }
            for (int d = 0; d < D; ++d) { pp.v[d][p] = vnew[d]; pp.xn[d][p] = xnew[d]; }
            if (apic) for (int c = 0; c < D; ++c) for (int r = 0; r < D; ++r) pp.B[r + c * D][p] = Bp[r][c];
            // re-scatter with the updated velocity and B (musl.cpp:128-170); base clamped as in P2G
            Stencil<D, SPL> sc;
            if (!make_stencil<D, SPL>(sc, xp, org, n, g.coord, h1, true)) err |= MPM_ERR_GRID_BOUNDS;
            else {
                const double aD = P.apicDinverse;
                for (int ia = 0; ia < Stencil<D, SPL>::cnt; ++ia) {
                    const int a = sc.lo[0] + ia;
                    for (int ib = 0; ib < Stencil<D, SPL>::cnt; ++ib) {
                        const int b = sc.lo[1] + ib;
                        for (int ic = 0; ic < c2; ++ic) {
                            const int c = (D == 3) ? sc.lo[2] + ic : 0;
                            double weight = sc.w[0][a] * sc.w[1][b];
                            if (D == 3) weight = weight * sc.w[2][c];
                            if (!(weight > 1e-25)) continue;
                            uint32_t node = (D == 3) ? node_index3(sc.base[0] + a, sc.base[1] + b, sc.base[2] + c, n)
                                                     : (uint32_t)(sc.base[0] + a) * n[1] + (uint32_t)(sc.base[1] + b);
                            double pd[3] = {sc.xi[0][a] - xp[0], sc.xi[1][b] - xp[1], D == 3 ? sc.xi[2][c] - xp[2] : 0.0};
                            for (int r = 0; r < D; ++r) {
                                double m = vnew[r] * weight;
                                if (apic) {
                                    double bd = Bp[r][0] * pd[0];
                                    for (int k = 1; k < D; ++k) bd += Bp[r][k] * pd[k];
                                    m += bd * aD * weight;
                                }
                                atomic_add_f64(&g.mv2[r][node], m);
                            }
                        }
                    }
                }
            }
        }
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
__global__ __launch_bounds__(kBlock) void k_defgrad(Params P, StepState* S, PartPtrs pp, GridPtrs g) {
    if (S->stop) return;
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    uint32_t err = 0, pcount = 0;
    if (p < pp.n) {
        const double h1 = P.one_over_dx, dt = S->dt;
        double org[3] = {S->origin[0], S->origin[1], S->origin[2]};
        uint32_t n[3] = {S->n[0], S->n[1], S->n[2]};
        double xp[3] = {0, 0, 0};
        for (int d = 0; d < D; ++d) xp[d] = pp.x[d][p];
        Stencil<D, SPL> st;
        if (!make_stencil<D, SPL>(st, xp, org, n, g.coord, h1, false)) {
            err |= MPM_ERR_GRID_BOUNDS;
        } else {
            double gv[3][3] = {{0}};
            const int c2 = (D == 3) ? Stencil<D, SPL>::cnt : 1;
            for (int ia = 0; ia < Stencil<D, SPL>::cnt; ++ia) {
                const int a = st.lo[0] + ia;
                for (int ib = 0; ib < Stencil<D, SPL>::cnt; ++ib) {
                    const int b = st.lo[1] + ib;
                    for (int ic = 0; ic < c2; ++ic) {
                        const int c = (D == 3) ? st.lo[2] + ic : 0;
                        uint32_t node = (D == 3) ? node_index3(st.base[0] + a, st.base[1] + b, st.base[2] + c, n)
                                                 : (uint32_t)(st.base[0] + a) * n[1] + (uint32_t)(st.base[1] + b);
                        double gr[3];
                        if (D == 3) {
                            gr[0] = st.dw[0][a] * st.w[1][b] * st.w[2][c] * h1;
                            gr[1] = st.dw[1][b] * st.w[0][a] * st.w[2][c] * h1;
                            gr[2] = st.dw[2][c] * st.w[0][a] * st.w[1][b] * h1;
                        } else {
                            gr[0] = st.dw[0][a] * st.w[1][b] * h1;
                            gr[1] = st.dw[1][b] * st.w[0][a] * h1;
                        }
                        for (int r = 0; r < D; ++r) { double vi = g.v[r][node]; for (int k = 0; k < D; ++k) gv[r][k] += vi * gr[k]; }
                    }
                }
            }
            deform_and_project<D>(P, dt, pp, p, gv, err, pcount);
        }
    }
    pcount = wave_sum_u32(pcount);
    err = wave_or_u32(err);
    if ((threadIdx.x & 63) == 0) {
        if (pcount) atomicAdd(&S->plastic_count, pcount);
        if (err) atomicOr(&S->err, err);
    }
}

// ----------------------------------------------------------------------------------------------
// k_finish: moveObjects (simulation.cpp:277-281) and the loop bookkeeping of simulate()
// (simulation.cpp:176-200), conservation checks (simulation.cpp:292-333) when requested.
// ----------------------------------------------------------------------------------------------
template <int D>
__global__ void k_finish(Params P, StepState* S, uint32_t n_particles) {
    if (threadIdx.x != 0 || S->stop) return;
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
    if (P.use_musl) S->xsel ^= 1u;
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
