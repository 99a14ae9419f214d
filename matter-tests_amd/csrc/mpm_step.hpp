// mpm_step.hpp — enqueues one MPM step (simulation.cpp:219-273) of a context on its stream. Included by
// the per-(dimension, spline degree, transfer type) translation units mpm_step_*.hip.
#pragma once
#include "mpm_context.hpp"
#include <type_traits>

namespace mpm {

// P2G and the grid update for one choice of separate force (flip needed or conservation checked) and MIBF.
template <int D, int SPL, bool SEP, bool MIBF, class TF>
void enqueue_p2g(mpm_context* c, const PartStateT<TF>& src, const PartStateT<TF>& dst, uint32_t gblocks) {
    constexpr int NV = P2GContrib<D, SPL, SEP, MIBF, TF>::NV;
    constexpr int TBP = ScatterLaunch<D, SPL, NV, TF, true>::TB, TB = ScatterLaunch<D, SPL, NV, TF, false>::TB;
    hipStream_t s = c->stream;
    const GridPtrsT<TF>& g = grid_of<TF>(c);
    const StagingT<TF>& stg = staging_of<TF>(c);
    {
        LaunchTimer t(c, KS_P2G);
        if (c->sb.ghosts) k_p2g<D, SPL, SEP, MIBF, TF, true><<<c->act_blocks, TBP, 0, s>>>(c->P, c->S, src, dst, c->sb, g, stg);
        else k_p2g<D, SPL, SEP, MIBF, TF, false><<<c->act_blocks, TB, 0, s>>>(c->P, c->S, src, dst, c->sb, g, stg);
    }
    { LaunchTimer t(c, KS_GRID); k_grid_update<D, SPL, SEP, MIBF, TF><<<gblocks, kBlock, 0, s>>>(c->P, c->S, c->sb, g, stg); }
}

// Calls f(std::integral_constant<int, MAT>) for material index mat (material_of).
template <class Fn>
void with_material_id(int mat, Fn&& f) {
    if (mat == kMatNeoHookean) { f(std::integral_constant<int, kMatNeoHookean>()); return; }
    switch (mat) {
    case MPM_PLASTIC_VM: f(std::integral_constant<int, MPM_PLASTIC_VM>()); break;
    case MPM_PLASTIC_DP: f(std::integral_constant<int, MPM_PLASTIC_DP>()); break;
    case MPM_PLASTIC_DPSOFT: f(std::integral_constant<int, MPM_PLASTIC_DPSOFT>()); break;
    case MPM_PLASTIC_MCC: f(std::integral_constant<int, MPM_PLASTIC_MCC>()); break;
    case MPM_PLASTIC_VMVISC: f(std::integral_constant<int, MPM_PLASTIC_VMVISC>()); break;
    case MPM_PLASTIC_DPVISC: f(std::integral_constant<int, MPM_PLASTIC_DPVISC>()); break;
    case MPM_PLASTIC_MCCVISC: f(std::integral_constant<int, MPM_PLASTIC_MCCVISC>()); break;
    case MPM_PLASTIC_DPMUI: f(std::integral_constant<int, MPM_PLASTIC_DPMUI>()); break;
    case MPM_PLASTIC_MCCMUI: f(std::integral_constant<int, MPM_PLASTIC_MCCMUI>()); break;
    case kMatVMViscLinear: f(std::integral_constant<int, kMatVMViscLinear>()); break;
    case kMatDPViscLinear: f(std::integral_constant<int, kMatDPViscLinear>()); break;
    default: f(std::integral_constant<int, MPM_PLASTIC_NONE>()); break;
    }
}

template <class Fn>
void with_material(const Params& P, Fn&& f) { with_material_id(material_of(P), static_cast<Fn&&>(f)); }

template <int D, int SPL, class TF>
void enqueue_step(mpm_context* c, uint32_t xsel, bool check_cons, bool head, bool tail) {
    hipStream_t s = c->stream;
    const Params& P = c->P;
    const PartStateT<TF>& src = layouts<TF>(c)[xsel & 1];
    const PartStateT<TF>& dst = layouts<TF>(c)[(xsel + 1) & 1];
    const GridPtrsT<TF>& g = grid_of<TF>(c);
    const StagingT<TF>& stg = staging_of<TF>(c);
    const uint32_t n = c->n;
    const uint32_t pblocks = (n + kBlock - 1) / kBlock;
    const uint32_t gblocks = (uint32_t)std::min<uint64_t>((c->cap_nodes + kBlock - 1) / kBlock, 8192);
    if (head) { LaunchTimer t(c, KS_SETUP); k_setup<D, TF><<<1, 64, 0, s>>>(P, c->S, g, c->sb); }  // else the previous step's last launch
    constexpr uint32_t kBinSpan = kBlock * kBinPer;  // particles per workgroup of the binning kernels
    const uint32_t bblocks = (n + kBinSpan - 1) / kBinSpan;
    { LaunchTimer t(c, KS_BIN_COUNT); k_bin_count<D, SPL, TF, kBinPer><<<bblocks, kBlock, 0, s>>>(P, c->S, src, c->sb, g, n); }
    { LaunchTimer t(c, KS_BIN_SCAN); k_bin_scan<><<<1, 1024, 0, s>>>(c->S, c->sb); }
    const uint32_t sblocks = (uint32_t)((((uint64_t)n << c->sb.ghosts) + kBinSpan - 1) / kBinSpan);  // + periodic copies
    { LaunchTimer t(c, KS_BIN_SCATTER); k_bin_scatter<kBinPer><<<sblocks, kBlock, 0, s>>>(c->S, c->sb, n); }
    const bool sep = P.flip_ratio >= -1 || check_cons;
    c->flip_formed = sep;
    if (sep && P.use_mibf) enqueue_p2g<D, SPL, true, true, TF>(c, src, dst, gblocks);
    else if (sep) enqueue_p2g<D, SPL, true, false, TF>(c, src, dst, gblocks);
    else if (P.use_mibf) enqueue_p2g<D, SPL, false, true, TF>(c, src, dst, gblocks);
    else enqueue_p2g<D, SPL, false, false, TF>(c, src, dst, gblocks);
    bool fused = false;  // G2P and the MUSL re-scatter in one pass (k_g2p_musl)
    if constexpr (kFusedMusl<D, SPL, TF>) {
        if (P.use_musl) {
            LaunchTimer t(c, KS_G2P_MUSL);
            if (P.flip_ratio >= -1) k_g2p_musl<D, SPL, true, TF><<<c->act_blocks, 64, 0, s>>>(P, c->S, src, dst, c->sb, g, stg);
            else k_g2p_musl<D, SPL, false, TF><<<c->act_blocks, 64, 0, s>>>(P, c->S, src, dst, c->sb, g, stg);
            fused = true;
        }
    }
    if (!fused) {
        LaunchTimer t(c, KS_G2P);
        if (P.flip_ratio >= -1) k_g2p<D, SPL, true, TF><<<c->act_blocks, kG2PThreads, 0, s>>>(P, c->S, src, dst, c->sb, g);
        else k_g2p<D, SPL, false, TF><<<c->act_blocks, kG2PThreads, 0, s>>>(P, c->S, src, dst, c->sb, g);
    }
    if (P.use_musl) {
        constexpr int TBM = ScatterLaunch<D, SPL, D, TF, false>::TB;
        if (!fused) { LaunchTimer t(c, KS_MUSL_SCATTER); k_musl_scatter<D, SPL, TF><<<c->act_blocks, TBM, 0, s>>>(P, c->S, src, dst, c->sb, g, stg); }
        { LaunchTimer t(c, KS_MUSL_UPDATE); k_musl_update<D, SPL, TF><<<gblocks, kBlock, 0, s>>>(P, c->S, c->sb, g, stg); }
    }
    {
        LaunchTimer t(c, KS_DEFGRAD);
        with_material(P, [&](auto mat) { k_defgrad<D, SPL, TF, decltype(mat)::value><<<pblocks, kBlock, 0, s>>>(P, c->S, src, dst, c->sb, g, n); });
    }
    {
        LaunchTimer t(c, KS_FINISH);
        if (tail) k_finish<D><<<1, 64, 0, s>>>(P, c->S, n);
        else k_finish_setup<D, TF><<<1, 64, 0, s>>>(P, c->S, g, c->sb, n);  // and the next step's k_setup
    }
}

template <int D, int SPL, bool SEP, bool MIBF, class TF>
void enqueue_batch_p2g(const BatchClass& bc, hipStream_t s) {
    constexpr int NV = P2GContrib<D, SPL, SEP, MIBF, TF>::NV;
    constexpr int TBP = ScatterLaunch<D, SPL, NV, TF, true>::TB, TB = ScatterLaunch<D, SPL, NV, TF, false>::TB;
    if (bc.key.pbc) kb_p2g<D, SPL, SEP, MIBF, TF, true><<<bc.wg[BR_BLOCK], TBP, 0, s>>>(bc.T);
    else kb_p2g<D, SPL, SEP, MIBF, TF, false><<<bc.wg[BR_BLOCK], TB, 0, s>>>(bc.T);
    kb_grid_update<D, SPL, SEP, MIBF, TF><<<bc.wg[BR_GRID], kBlock, 0, s>>>(bc.T);
}

template <int D, int SPL, class TF>
void enqueue_batch_step(const BatchClass& bc, hipStream_t s) {
    const BatchTable& T = bc.T;
    const BatchClassKey& k = bc.key;
    kb_setup<D, TF><<<T.ns, 64, 0, s>>>(T);
    kb_zero_grid<TF><<<bc.wg[BR_ZERO], kBlock, 0, s>>>(T);
    kb_bin_count<D, SPL, TF><<<bc.wg[BR_PART], kBlock, 0, s>>>(T);
    kb_bin_scan<TF><<<T.ns, 1024, 0, s>>>(T);
    kb_bin_scatter<TF><<<bc.wg[BR_PART], kBlock, 0, s>>>(T);
    if (k.sep && k.mibf) enqueue_batch_p2g<D, SPL, true, true, TF>(bc, s);
    else if (k.sep) enqueue_batch_p2g<D, SPL, true, false, TF>(bc, s);
    else if (k.mibf) enqueue_batch_p2g<D, SPL, false, true, TF>(bc, s);
    else enqueue_batch_p2g<D, SPL, false, false, TF>(bc, s);
    if (k.flip) kb_g2p<D, SPL, true, TF><<<bc.wg[BR_BLOCK], kG2PThreads, 0, s>>>(T);
    else kb_g2p<D, SPL, false, TF><<<bc.wg[BR_BLOCK], kG2PThreads, 0, s>>>(T);
    if (k.musl) {
        constexpr int TBM = ScatterLaunch<D, SPL, D, TF, false>::TB;
        kb_musl_scatter<D, SPL, TF><<<bc.wg[BR_BLOCK], TBM, 0, s>>>(T);
        kb_musl_update<D, SPL, TF><<<bc.wg[BR_GRID], kBlock, 0, s>>>(T);
    }
    with_material_id(k.mat, [&](auto mat) { kb_defgrad<D, SPL, TF, decltype(mat)::value><<<bc.wg[BR_PART], kBlock, 0, s>>>(T); });
    kb_finish<D, TF><<<T.ns, 64, 0, s>>>(T);
}

}  // namespace mpm

#define MPM_INSTANTIATE_STEP(D, SPL, TF) template void mpm::enqueue_step<D, SPL, TF>(mpm_context*, uint32_t, bool, bool, bool); \
    template void mpm::enqueue_batch_step<D, SPL, TF>(const mpm::BatchClass&, hipStream_t);
