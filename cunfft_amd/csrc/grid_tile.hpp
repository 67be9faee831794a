// grid_tile.hpp -- plane-tiled gridding kernels for d = 2, 3 on gfx950 (templates;
// grid_tile_f32.hip and grid_tile_f64.hip instantiate them per precision).
//
// Forward gridding f_j = sum_l g_l phi(x_j - l/n) (src/cunfft_kernel.cu:514-1093)
// and adjoint gridding g_l = sum_j f_j phi(x_j - l/n) (src/cunfft_kernel.cu:
// 608-1197), one workgroup per chunk of the nodes binned to one grid tile
// (bin.hip).  The window of a node covers W = 2m+2 cells per axis; its cells
// in the padded tile held in LDS are visited one plane (leading-axis index) at
// a time, with the lanes of a wavefront spread over the W rows x window width
// of the plane:
//   * each lane owns one 16-byte unit of a row: two FP32 complex cells, or one
//     FP64 cell, so every LDS access is one ds_read_b128 / ds_write_b128;
//   * one instruction holds the cells of a single node, so the lanes of an
//     instruction never touch the same cell and the adjoint accumulates with
//     plain read-add-write;
//   * the lane -> (row, unit) assignment is a compile-time table that puts 16
//     lanes on 16 distinct LDS bank groups for the chosen row pitch (gfx950
//     serves 16 lanes of a 128-bit access per clock from 64 banks of 4 bytes).
// Nodes are taken in batches: a prologue turns each node's descriptor into its
// window origin and its three weight vectors in LDS (one thread per weight,
// hardware exp2 for FP32, exp for FP64), and the gridding loop reads them back
// with broadcast LDS reads while the next batch's descriptors are loaded.
//
// Adjoint: the accumulated padded tile of each chunk is written to a slot in
// a destination-blocked order, and k_tile_fold sums, for every grid tile, the
// slot blocks of the chunks whose padded tiles cover it: no atomics, no zero
// fill, and a result that does not depend on scheduling.  The 3D adjoint
// on 8^3 tiles (FP32 and FP64) accumulates its slots on the matrix cores
// instead (grid_mfma.hpp); the fold is shared.
#pragma once
#include <map>
#include <mutex>

#include "bin.hpp"

namespace cunfft {

constexpr int kLdsBudget = 80 * 1024;   // a padded tile up to this size leaves room for 2 per CU
constexpr int kLdsMax = 150 * 1024;

// ------------------------------------------------- compile-time layout --
template <int N> struct LaneTab {
    uint16_t v[N];
};

// 16-byte LDS bank group (0..15) of item (q, r, u): plane q * PS, row r, unit u.
constexpr int unit_group(int q, int r, int u, int PS, int PP16, int RP16)
{
    return (q * PS * PP16 + r * RP16 + u) & 15;
}

constexpr int max_group_load(int G, int R, int NU, int PS, int P1, int RP, int CB)
{
    int cnt[16] = {};
    const int RP16 = RP * CB / 16, PP16 = P1 * RP * CB / 16;
    for (int q = 0; q < G; ++q)
        for (int r = 0; r < R; ++r)
            for (int u = 0; u < NU; ++u) ++cnt[unit_group(q, r, u, PS, PP16, RP16)];
    int mx = 0;
    for (int k = 0; k < 16; ++k) mx = cnt[k] > mx ? cnt[k] : mx;
    return mx;
}

// Row pitch (cells): the smallest one, from the needed row length RL up, whose
// items spread over the NG 16-lane groups without bank conflicts, within the
// LDS budget; otherwise the least conflicted one.
constexpr int choose_pitch(int G, int R, int NU, int PS, int P0, int P1, int RL, int CB, int NG)
{
    const int step = 16 / CB;
    const bool room = (long)P0 * P1 * RL * CB <= kLdsBudget;
    int best = RL, best_cost = 1 << 30;
    for (int RP = RL; RP < RL + 16; RP += step) {
        if (room && (long)P0 * P1 * RP * CB > kLdsBudget) break;
        if (!room && RP != RL) break;
        const int c = max_group_load(G, R, NU, PS, P1, RP, CB);
        const int cost = (c + NG - 1) / NG;
        if (cost < best_cost) {
            best_cost = cost;
            best = RP;
        }
    }
    return best;
}

// Lane table: items listed group-major by bank group, most loaded group
// first, and dealt round-robin over the NG groups of 16 lanes.  A bank group
// with at most NG items then never appears twice in one 16-lane group.  Within
// a group, the item of bank group p sits in lane p: any shift of the node's
// origin then keeps the 16 lanes on 16 bank groups (what 128-bit reads need)
// and each half of 8 lanes on 8 distinct bank groups mod 8 (what 128-bit
// writes need, measured on gfx950).
// Entry: q << 10 | r << 5 | u, or 0xFFFF for an idle lane.
template <int NT>
constexpr LaneTab<NT> make_lanes(int G, int R, int NU, int PS, int P1, int RP, int CB, int NG)
{
    LaneTab<NT> t{};
    for (int i = 0; i < NT; ++i) t.v[i] = 0xFFFF;
    const int RP16 = RP * CB / 16, PP16 = P1 * RP * CB / 16;
    int cnt[16] = {};
    for (int q = 0; q < G; ++q)
        for (int r = 0; r < R; ++r)
            for (int u = 0; u < NU; ++u) ++cnt[unit_group(q, r, u, PS, PP16, RP16)];
    int order[16] = {};
    for (int k = 0; k < 16; ++k) order[k] = k;
    for (int a = 0; a < 16; ++a)
        for (int b = a + 1; b < 16; ++b)
            if (cnt[order[b]] > cnt[order[a]]) {
                const int tmp = order[a];
                order[a] = order[b];
                order[b] = tmp;
            }
    int k = 0;
    for (int oi = 0; oi < 16; ++oi)
        for (int q = 0; q < G; ++q)
            for (int r = 0; r < R; ++r)
                for (int u = 0; u < NU; ++u)
                    if (unit_group(q, r, u, PS, PP16, RP16) == order[oi]) {
                        // lane order[oi] of the next group that has it free; if a bank
                        // group has more items than there are groups, any free lane
                        const int p = order[oi];
                        int slot = -1;
                        for (int g = 0; g < NT / 16 && slot < 0; ++g) {
                            const int grp = (k + g) % (NT / 16);
                            if (t.v[grp * 16 + p] == 0xFFFF) slot = grp * 16 + p;
                        }
                        for (int l = 0; l < NT && slot < 0; ++l)
                            if (t.v[l] == 0xFFFF) slot = l;
                        t.v[slot] = (uint16_t)(q << 10 | r << 5 | u);
                        ++k;
                    }
    return t;
}

// Layout of the padded tile in LDS and of one node's window over the lanes.
// PS is the plane step between the planes one wavefront visits (the number of
// wavefronts sharing the adjoint tile, 1 for the forward).
template <typename T, int D, int W, int PS, int V> struct TileLayout {
    using Real = T;
    static constexpr int PSV = PS;
    static constexpr int CB = (int)sizeof(cplx_t<T>);
    static constexpr int CPL = 16 / CB;                      // cells per lane unit
    static constexpr int T0 = TileShapeV<D, V>::T[0];
    static constexpr int T1 = D == 3 ? TileShapeV<D, V>::T[1] : 1;
    static constexpr int TL = TileShapeV<D, V>::T[D - 1];
    static constexpr int P0 = T0 + W - 1;
    static constexpr int P1 = D == 3 ? T1 + W - 1 : 1;
    static constexpr int PL = TL + W - 1;
    static constexpr int R = D == 3 ? W : 1;                 // window rows per plane
    static constexpr int NU = CPL == 2 ? W / 2 + 1 : W;      // lane units per window row
    static constexpr int RL = CPL == 2 ? 2 * ((TL - 1) / 2 + NU) : PL;  // row cells used
    static constexpr int OWN = (W + PS - 1) / PS;            // planes per wavefront per node
    static constexpr int GM = 64 / (R * NU) > 0 ? 64 / (R * NU) : 1;
    static constexpr int G = GM < OWN ? GM : OWN;            // planes per lane group
    static constexpr int NGRP = (OWN + G - 1) / G;
    static constexpr int I = G * R * NU;
    static constexpr int NG = (I + 15) / 16;
    static constexpr int NI = (NG + 3) / 4;                  // LDS instructions per plane group
    static constexpr int RP = choose_pitch(G, R, NU, PS, P0, P1, RL, CB, NG);
    static constexpr int PP = P1 * RP;
    static constexpr int CELLS = P0 * PP;
    // adjoint slot cells, rounded up to whole 16-byte units so every slot starts aligned
    static constexpr int SC = (P0 * P1 * PL + CPL - 1) / CPL * CPL;
    static constexpr int Q0 = (P0 + T0 - 1) / T0, Q1 = (P1 + T1 - 1) / T1,
                         QL = (PL + TL - 1) / TL;            // slot blocks per axis
    static constexpr LaneTab<NI * 64> lanes = make_lanes<NI * 64>(G, R, NU, PS, P1, RP, CB, NG);
    static constexpr bool fits = (long)CELLS * CB <= kLdsMax && W <= 18 && NU <= 31 && G < 64;
    // a plane group of some node can reach past its window (adjoint dummy units needed)
    static constexpr bool NEED_DUMMY = (PS - 1) + PS * (NGRP * G - 1) >= W;
    // A lane's dummy unit: the spare 16 bytes at the end of padded row `lane`
    // when the pitch leaves them (no row access reaches past RL cells), else
    // one of 64 units appended to the tile.
    static constexpr bool SPARE = (RP - RL) * CB >= 16 && (RL * CB) % 16 == 0 && P0 * P1 >= 64;
    static constexpr int DUMMY_UNITS = NEED_DUMMY && !SPARE ? 64 : 0;
    __device__ static uint32_t dummy_byte(int lane)
    {
        return SPARE ? (uint32_t)(lane * RP + RL) * CB : (uint32_t)(CELLS / CPL + lane) * 16;
    }
    // planes loaded per batch: at most 12 outstanding 16-byte LDS reads
    static constexpr int PB = 12 / NI > 0 ? (12 / NI < NGRP ? 12 / NI : NGRP) : 1;

    // Slot block q_t along an axis with tile T_ and padded extent P_.
    static constexpr int ext(int T_, int P_, int q) { return P_ - q * T_ < T_ ? P_ - q * T_ : T_; }
    // Offset of block (q0, q1, qL) in the destination-blocked slot order.
    static constexpr int block_off(int q0, int q1, int qL)
    {
        return q0 * T0 * P1 * PL + ext(T0, P0, q0) * (q1 * T1 * PL + ext(T1, P1, q1) * qL * TL);
    }
};

// ----------------------------------------------------------- helpers ----

// The Kaiser-Bessel and sinc weights out of line: the Gaussian path keeps its
// registers (the window is a run-time choice of the plan).
template <typename T> __device__ __noinline__ T alt_phi_call(int win, T s, int m, T shape)
{
    return win == WIN_SINC ? sinc_phi<T>(s, m, shape) : kb_phi<T>(s, (T)m, shape);
}

template <typename T> __device__ __forceinline__ T wexp(T v);
template <> __device__ __forceinline__ float wexp<float>(float v)
{
    return __builtin_amdgcn_exp2f(v);
}
template <> __device__ __forceinline__ double wexp<double>(double v) { return exp(v); }

template <typename T> struct Vec16;
template <> struct Vec16<float> { using type = float4; };
template <> struct Vec16<double> { using type = double2; };
template <typename T> using vec16_t = typename Vec16<T>::type;

template <int CTRL, int ROWS> __device__ __forceinline__ float dpp(float v)
{
    return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWS, 0xf, false));
}
template <int CTRL, int ROWS> __device__ __forceinline__ double dpp(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(uint32_t)b, CTRL, ROWS, 0xf, false);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(b >> 32), CTRL, ROWS, 0xf, false);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Sum of the 16 lanes of each row of the wavefront, in the row's last lane.
template <typename T> __device__ __forceinline__ T row_sum15(T v)
{
    v += dpp<0x111, 0xf>(v);
    v += dpp<0x112, 0xf>(v);
    v += dpp<0x114, 0xf>(v);
    v += dpp<0x118, 0xf>(v);
    return v;
}

// Compiler-only barrier: keeps a wavefront's LDS accesses in program order
// (the LDS executes the instructions of one wavefront in order).
__device__ __forceinline__ void lds_order() { __asm__ volatile("" ::: "memory"); }

// s_waitcnt operand waiting for LDS/scalar-memory count 0 only (vmcnt and
// expcnt at their maxima, gfx9 encoding).
constexpr int kWaitLgkm0 = 0xC07F;

// Per-lane item of the lane table: plane q * PS, row r, unit u of a node's
// window, as a cell offset from the window origin in the padded tile.
struct LaneItem {
    int off;    // cells
    int q;      // plane within the plane group (0 when a group holds one plane)
    int r;      // window row
    int u2;     // first column tap of the unit, before the parity shift
    int ok;
};

template <typename L> __device__ __forceinline__ LaneItem lane_item(int i, int lane)
{
    constexpr LaneTab<L::NI * 64> tab = L::lanes;
    const int e = tab.v[i * 64 + lane];
    LaneItem it;
    it.ok = e != 0xFFFF;
    const int q = it.ok ? e >> 10 : 0, r = it.ok ? (e >> 5) & 31 : 0, u = it.ok ? e & 31 : 0;
    it.q = L::G == 1 ? 0 : q;   // one plane per lane group: the plane index is uniform
    it.r = r;
    it.u2 = L::CPL * u;
    it.off = q * L::PSV * L::PP + r * L::RP + L::CPL * u;
    return it;
}

// Window weights of one batch of nodes, written to LDS by the wavefront that
// grids the batch with one lane per weight, so the gridding loop evaluates no
// exponentials: its lanes fetch their plane, row and column weights with
// broadcast LDS reads issued together with the node's tile accesses.
// NWP > 0 (adjoint with register runs): the plane weights of each node are
// grouped by the plane residue a mod NWP of its NWP wavefronts, w0[n][r * OWNP + i]
// = weight of plane r + NWP i (zero past W), so a wavefront reads its planes'
// weights as whole 16-byte vectors.
template <typename T, int D, int W, int NB, int NWP = 0> struct alignas(16) WinBatch {
    static constexpr int V0 = (W + 3) & ~3;
    static constexpr int OWN = NWP > 0 ? (W + NWP - 1) / NWP : 0;
    static constexpr int VW = 16 / (int)sizeof(T);
    static constexpr int OWNP = (OWN + VW - 1) / VW * VW;
    static constexpr int W0N = NWP > 0 ? NWP * OWNP : V0;
    T w0[NB][W0N];        // plane weights
    // row weights (d = 3); NWP > 0: times the node's coefficient, (re, im) per tap
    T w1[NB][NWP > 0 ? 2 * W : W];
    T w2[NB][W + 4];      // column weight of tap l at l + 1; zero at 0 and past W
    T fx[NB][D];          // c_t - m per axis (prologue input)
};

// Nodes per wavefront batch: the largest power of two up to MAXNB that keeps
// the workgroups per CU the padded tile alone allows.  LDS is handed out in
// granules (a 54048-byte adjoint workgroup ran two per CU where a 52768-byte
// one ran three), so the budget is rounded down to kLdsGranule.
constexpr int kLdsGranule = 2048;   // conservative LDS allocation granule (bytes)

template <typename T, int D, int W, int NWP = 0, int MAXNB = 16>
constexpr int batch_nodes(int tile_bytes, int nw, int extra_per_node)
{
    const int total = 160 * 1024;
    const int per_cu = total / tile_bytes > 0 ? total / tile_bytes : 1;
    const int budget = total / per_cu / kLdsGranule * kLdsGranule - tile_bytes;
    return MAXNB >= 32 && nw * ((int)sizeof(WinBatch<T, D, W, 32, NWP>) + 32 * extra_per_node) <= budget
               ? 32
           : nw * ((int)sizeof(WinBatch<T, D, W, 16, NWP>) + 16 * extra_per_node) <= budget  ? 16
           : nw * ((int)sizeof(WinBatch<T, D, W, 8, NWP>) + 8 * extra_per_node) <= budget ? 8
           : nw * ((int)sizeof(WinBatch<T, D, W, 4, NWP>) + 4 * extra_per_node) <= budget ? 4
                                                                                          : 2;
}

// Window origin of a node, held by the lane that loaded it.
struct NodeRec {
    int nbase;   // cell of the window origin in the padded tile (plane shift excluded)
    int lo0;
    int par;     // FP32: parity of the column origin (16-byte unit alignment)
};

// Batch prologue of one wavefront: lane n < nb turns node n's descriptor into
// its record and axis offsets, then every lane evaluates window weights.
// Within one wavefront LDS accesses execute in order, so the compiler
// barriers suffice between the phases.
template <typename T, int D, int W, int NB, typename L, int NWP = 0>
__device__ __forceinline__ NodeRec batch_prologue(WinBatch<T, D, W, NB, NWP> &wb, int nb, int lane,
                                                  const NodeDesc<T> &nd, const WinConst<T> &wc)
{
    using WB = WinBatch<T, D, W, NB, NWP>;
    if constexpr (NWP > 0) {
        // zero the padding of the grouped plane weights (whole array, 16-byte stores)
        using VT = vec16_t<T>;
        constexpr int NV = NB * WB::W0N / WB::VW;
        VT z;
        if constexpr (sizeof(T) == 4) z = make_float4(0.f, 0.f, 0.f, 0.f);
        else z = make_double2(0.0, 0.0);
#pragma unroll
        for (int i = lane; i < NV; i += 64) reinterpret_cast<VT *>(&wb.w0[0][0])[i] = z;
    }
    NodeRec r{0, 0, 0};
    if (lane < nb) {
        const uint32_t loc = (uint32_t)nd.loc;
        r.lo0 = loc & 255;
        const int lo1 = D == 3 ? (loc >> 8) & 255 : 0;
        const int loL = (loc >> (8 * (D - 1))) & 255;
        r.par = L::CPL == 2 ? (loL & 1) : 0;
        r.nbase = r.lo0 * L::PP + lo1 * L::RP + loL - r.par;
#pragma unroll
        for (int t = 0; t < D; ++t) wb.fx[lane][t] = nd.f[t];
        wb.w2[lane][0] = (T)0;
        wb.w2[lane][W + 1] = (T)0;
        wb.w2[lane][W + 2] = (T)0;
        wb.w2[lane][W + 3] = (T)0;
    }
    lds_order();
    // axes at compile time (the per-axis constants stay scalar), lanes over
    // (node, tap), fixed trip count so all offsets are read before the exps
    constexpr int NPASS = (NB * W + 63) / 64;
#pragma unroll
    for (int t = 0; t < D; ++t) {
        T sv[NPASS];
        int nv[NPASS], lv[NPASS];
#pragma unroll
        for (int ps = 0; ps < NPASS; ++ps) {
            int it = lane + 64 * ps;
            it = it < nb * W ? it : 0;
            nv[ps] = it / W;
            lv[ps] = it - nv[ps] * W;
            sv[ps] = wb.fx[nv[ps]][t] + (T)(wc.m - lv[ps]);
        }
#pragma unroll
        for (int ps = 0; ps < NPASS; ++ps) {
            if (lane + 64 * ps >= nb * W) break;
            const T w = wc.win != WIN_GAUSSIAN
                            ? alt_phi_call<T>(wc.win, sv[ps], wc.m, wc.shp[t])
                            : wc.nrm[t] * wexp<T>(-wc.kb[t] * sv[ps] * sv[ps]);
            if (t == 0) {
                if constexpr (NWP > 0)
                    wb.w0[nv[ps]][(lv[ps] % NWP) * WB::OWNP + lv[ps] / NWP] = w;
                else
                    wb.w0[nv[ps]][lv[ps]] = w;
            }
            else if (t == D - 1) wb.w2[nv[ps]][lv[ps] + 1] = w;
            else wb.w1[nv[ps]][lv[ps]] = w;
        }
    }
    lds_order();
    return r;
}

// Batch prologue shared by the NW wavefronts of a workgroup that all walk
// the same nodes (the 3D adjoint): each wavefront turns its lane n < nb
// descriptor into the node's record (as batch_prologue), wavefront 0 writes
// the per-node inputs, and the D axes x ceil(nb W / 64) weight passes are
// dealt over the wavefronts, so a weight is evaluated once per workgroup, not
// once per wavefront.  Ends with a barrier; the caller must not rewrite wb
// before every wavefront has read the previous batch (a barrier).
template <typename T> __device__ __forceinline__ T lane_read(T v, int src);
template <> __device__ __forceinline__ float lane_read<float>(float v, int src)
{
    return __int_as_float(__builtin_amdgcn_ds_bpermute(src * 4, __float_as_int(v)));
}
template <> __device__ __forceinline__ double lane_read<double>(double v, int src)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_ds_bpermute(src * 4, (int)(uint32_t)b);
    const int hi = __builtin_amdgcn_ds_bpermute(src * 4, (int)(b >> 32));
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

template <typename T, int D, int W, int NB, typename L, int NWP, int NW>
__device__ __forceinline__ NodeRec batch_prologue_wg(WinBatch<T, D, W, NB, NWP> &wb, int nb, int lane,
                                                     int wave, const NodeDesc<T> &nd,
                                                     const cplx_t<T> &fcn, const WinConst<T> &wc)
{
    static_assert(D == 3 && NWP > 0, "row weights are stored times the coefficient");
    using WB = WinBatch<T, D, W, NB, NWP>;
    NodeRec r{0, 0, 0};
    if (lane < nb) {
        const uint32_t loc = (uint32_t)nd.loc;
        r.lo0 = loc & 255;
        const int lo1 = D == 3 ? (loc >> 8) & 255 : 0;
        const int loL = (loc >> (8 * (D - 1))) & 255;
        r.par = L::CPL == 2 ? (loL & 1) : 0;
        r.nbase = r.lo0 * L::PP + lo1 * L::RP + loL - r.par;
    }
    if (wave == 0) {
        if constexpr (NWP > 0) {
            using VT = vec16_t<T>;
            constexpr int NV = NB * WB::W0N / WB::VW;
            VT z;
            if constexpr (sizeof(T) == 4) z = make_float4(0.f, 0.f, 0.f, 0.f);
            else z = make_double2(0.0, 0.0);
#pragma unroll
            for (int i = lane; i < NV; i += 64) reinterpret_cast<VT *>(&wb.w0[0][0])[i] = z;
        }
        if (lane < nb) {
#pragma unroll
            for (int t = 0; t < D; ++t) wb.fx[lane][t] = nd.f[t];
            wb.w2[lane][0] = (T)0;
            wb.w2[lane][W + 1] = (T)0;
            wb.w2[lane][W + 2] = (T)0;
            wb.w2[lane][W + 3] = (T)0;
        }
    }
    __syncthreads();
    constexpr int NPASS = (NB * W + 63) / 64;
    for (int cmb = wave; cmb < D * NPASS; cmb += NW) {   // wavefront-uniform axis and pass
        const int t = cmb / NPASS, ps = cmb - t * NPASS;
        const int it = lane + 64 * ps;
        // the row pass needs node n's coefficient, held by lane n of every wavefront
        const int nn = it / W < NB ? it / W : NB - 1;
        const T fr = t == 1 ? lane_read<T>(fcn.x, nn) : (T)0;
        const T fi = t == 1 ? lane_read<T>(fcn.y, nn) : (T)0;
        if (it < nb * W) {
            const int n = it / W, l = it - n * W;
            const T sv = wb.fx[n][t] + (T)(wc.m - l);
            const T kbt = t == 0 ? wc.kb[0] : (t == 1 ? wc.kb[1] : wc.kb[2]);
            const T nrt = t == 0 ? wc.nrm[0] : (t == 1 ? wc.nrm[1] : wc.nrm[2]);
            const T shpt = t == 0 ? wc.shp[0] : (t == 1 ? wc.shp[1] : wc.shp[2]);
            const T w = wc.win != WIN_GAUSSIAN ? alt_phi_call<T>(wc.win, sv, wc.m, shpt)
                                               : nrt * wexp<T>(-kbt * sv * sv);
            if (t == 0) {
                if constexpr (NWP > 0) wb.w0[n][(l % NWP) * WB::OWNP + l / NWP] = w;
                else wb.w0[n][l] = w;
            } else if (t == D - 1) {
                wb.w2[n][l + 1] = w;
            } else {
                wb.w1[n][2 * l] = w * fr;
                wb.w1[n][2 * l + 1] = w * fi;
            }
        }
    }
    __syncthreads();
    return r;
}

__device__ __forceinline__ int rlane_i(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ __forceinline__ float rlane_t(float v, int k)
{
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), k));
}
__device__ __forceinline__ double rlane_t(double v, int k)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(uint32_t)b, k);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), k);
    return __longlong_as_double((long long)(((uint64_t)(uint32_t)hi << 32) | (uint32_t)lo));
}

// Kahan step: s + e keeps the running sum to about twice the working
// precision (e holds the negated low part lost by the last additions).
template <typename T> __device__ __forceinline__ void kahan_add(T &s, T &e, T x)
{
    const T y = x - e;
    const T t = s + y;
    e = (t - s) - y;
    s = t;
}

// Fold sum: plain, or Neumaier's compensated sum (c collects the low parts;
// the result is s + c).
template <bool COMP, typename T> __device__ __forceinline__ void fold_add(T &s, T &c, T x)
{
    if constexpr (COMP) {
        const T t = s + x;
        c += fabs(s) >= fabs(x) ? (s - t) + x : (x - t) + s;
        s = t;
    } else {
        s += x;
    }
}

// Stages the halo-padded grid tile `tile_id` into LDS (rows of the pitch,
// 16 bytes per load, every padded row wrapped onto the grid).  rowoff_mem is
// LDS scratch for the padded rows' grid offsets (used when they fit in
// rowoff_bytes, otherwise each offset is recomputed per load).  The caller
// puts a barrier after.
template <typename T, int D, typename L, int NT>
__device__ __forceinline__ void stage_tile(const cplx_t<T> *__restrict__ g, uint32_t tile_id,
                                           cplx_t<T> *tile, const Geom &ge, const BinGeom &bg,
                                           void *rowoff_mem, size_t rowoff_bytes)
{
    using VT = vec16_t<T>;
    constexpr int CPL = L::CPL;
    // lane units staged per row: the whole pitch, spare units included, so the
    // lanes of a wavefront write consecutive LDS units (no bank conflicts)
    constexpr int NR = L::RP / CPL;
    const int tid = threadIdx.x;
    uint32_t rem = tile_id, o[3] = {0, 0, 0};
    for (int t = D - 1; t >= 0; --t) {
        o[t] = (rem % (uint32_t)bg.nt[t]) << bg.logT[t];
        rem /= (uint32_t)bg.nt[t];
    }
    const uint32_t n0 = (uint32_t)ge.n[0], nL = (uint32_t)ge.n[D - 1];
    const uint32_t n1 = D == 3 ? (uint32_t)ge.n[1] : 1;
    const uint32_t s1 = D == 3 ? (uint32_t)ge.n[2] : 0;
    const uint32_t s0 = D == 3 ? (uint32_t)(ge.n[1] * ge.n[2]) : (uint32_t)ge.n[1];
    VT *tv = reinterpret_cast<VT *>(tile);
    const VT *gv = reinterpret_cast<const VT *>(g);
    // A thread keeps one unit column un of the rows; its RS = NT / NR
    // sibling threads per column take rows r0, r0 + RS, ... (NT mod NR
    // threads idle), so the column's wrapped grid offset is computed once.
    // 8 loads in flight per thread before the LDS stores (one at a time,
    // each would wait out the full memory latency).
    constexpr int ROWS = L::P0 * L::P1, RS = NT / NR, SB = 8;
    static_assert(RS >= 1, "a workgroup covers a padded row");
    // wrapped grid offset of a padded row; when they fit, all of them are
    // computed once per workgroup into rowoff_mem (the caller's weight-batch
    // memory, unused until the tile is staged)
    auto row_offset = [&](int row) {
        const int p0 = row / L::P1, p1 = row - p0 * L::P1;
        // o + p < 2n: the padded tile reaches at most W - 1 cells past the grid
        uint32_t off = wrap_hi32(o[0] + p0, n0) * s0;
        if (D == 3) off += wrap_hi32(o[1] + p1, n1) * s1;
        return off;
    };
    const bool TABLE = ROWS * sizeof(uint32_t) <= rowoff_bytes;
    uint32_t *rowoff = static_cast<uint32_t *>(rowoff_mem);
    if (TABLE) {
        for (int row = tid; row < ROWS; row += NT) rowoff[row] = row_offset(row);
        __syncthreads();
    }
    const int un = tid % NR, r0 = tid / NR;
    const uint32_t colc = wrap_hi32(o[D - 1] + CPL * un, nL);
    if (r0 < RS) {
        for (int rb = r0; rb < ROWS; rb += SB * RS) {
            VT v[SB];
            int at[SB];
#pragma unroll
            for (int u = 0; u < SB; ++u) {
                const int row = rb + u * RS < ROWS ? rb + u * RS : ROWS - 1;
                at[u] = (row * L::RP) / CPL + un;
                v[u] = gv[((TABLE ? rowoff[row] : row_offset(row)) + colc) / CPL];
            }
            // unconditional stores (a thread past the last row rewrites it
            // with its own value): a conditional store would pull its load
            // back into the branch, one round trip at a time
#pragma unroll
            for (int u = 0; u < SB; ++u) tv[at[u]] = v[u];
        }
    }
}

// ------------------------------------------------------------ gather ----
// NW wavefronts share the staged tile and take batches of the chunk's nodes
// round-robin.  Each lane accumulates its unit over the node's planes and
// weighs the sum with its row and column weights; the sixteen lanes of each
// row are summed with DPP, and the four row sums of each node are added by
// the node's lane after the batch.
template <typename T, int D, int W, int NW, int V>
__device__ __forceinline__ void tile_gather(const cplx_t<T> *__restrict__ g,
                                            const NodeDesc<T> *__restrict__ desc,
                                            const uint32_t *__restrict__ perm, const Chunk ch,
                                            cplx_t<T> *__restrict__ f, const Geom &ge,
                                            const BinGeom &bg, const WinConst<T> &wc)
{
    using L = TileLayout<T, D, W, 1, V>;
    using VT = vec16_t<T>;
    using C = cplx_t<T>;
    constexpr int CPL = L::CPL;
    constexpr int NT = NW * 64;
    constexpr int PB = L::PB;
    constexpr int NB = batch_nodes<T, D, W>(L::CELLS * L::CB, NW, 4 * (int)sizeof(C));
    __shared__ __align__(16) C tile[L::CELLS];
    __shared__ WinBatch<T, D, W, NB> wbs[NW];
    __shared__ C red[NW][NB][4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nbat = (ch.count + NB - 1) / NB;

    // this wavefront's first batch in flight while the tile is staged
    NodeDesc<T> nd{};
    uint32_t pm = 0;
    {
        const uint32_t s = (uint32_t)wave * NB + lane;
        if (lane < NB && s < ch.count) {
            nd = desc[ch.start + s];
            pm = perm[ch.start + s];
        }
    }

    stage_tile<T, D, L, NT>(g, ch.tile, tile, ge, bg, &wbs[0], sizeof(wbs));
    __syncthreads();

    LaneItem it[L::NI];
#pragma unroll
    for (int i = 0; i < L::NI; ++i) it[i] = lane_item<L>(i, lane);
    const char *tb = reinterpret_cast<const char *>(tile);
    WinBatch<T, D, W, NB> &wb = wbs[wave];

    for (uint32_t bi = wave; bi < nbat; bi += NW) {
        const uint32_t s0 = bi * NB;
        const int nb = ch.count - s0 < (uint32_t)NB ? (int)(ch.count - s0) : NB;
        const NodeRec rc = batch_prologue<T, D, W, NB, L>(wb, nb, lane, nd, wc);
        const uint32_t pm_out = pm;
        {
            const uint32_t s = s0 + NW * NB + lane;
            if (lane < NB && s < ch.count) {
                nd = desc[ch.start + s];
                pm = perm[ch.start + s];
            }
        }
        for (int k = 0; k < nb; ++k) {
            const int nbase = rlane_i(rc.nbase, k), par = rlane_i(rc.par, k);
            T a0x[L::NI], a0y[L::NI], a1x[L::NI], a1y[L::NI];
#pragma unroll
            for (int i = 0; i < L::NI; ++i) a0x[i] = a0y[i] = a1x[i] = a1y[i] = 0;
            T part_x = 0, part_y = 0;
#pragma unroll
            for (int i = 0; i < L::NI; ++i) {
                if (!it[i].ok) continue;
                const uint32_t lb = (uint32_t)(nbase + it[i].off - it[i].q * L::PP) * L::CB;
                const int cw = it[i].u2 - par + 1;
                T w1v = 0, w2a = 0, w2b = 0;
#pragma unroll
                for (int g0 = 0; g0 < L::NGRP; g0 += PB) {
                    VT v[PB];
#pragma unroll
                    for (int j = 0; j < PB; ++j) {
                        const int gq = g0 + j;
                        if (gq >= L::NGRP) break;
                        int a = gq * L::G + it[i].q;
                        a = a < W ? a : W - 1;   // planes past the window: weight 0 below
                        v[j] = *reinterpret_cast<const VT *>(tb + lb + (uint32_t)(a * L::PP * L::CB));
                    }
                    // the node's weights are read only after its tile reads are
                    // issued, and all of them before any use: one LDS round trip
                    lds_order();
                    if (g0 == 0) {
                        w1v = D == 3 ? wb.w1[k][it[i].r] : (T)1;
                        w2a = wb.w2[k][cw];
                        w2b = CPL == 2 ? wb.w2[k][cw + 1] : (T)0;
                    }
                    T w0r[PB];
#pragma unroll
                    for (int j = 0; j < PB; ++j) {
                        const int a = (g0 + j) * L::G + it[i].q;
                        w0r[j] = wb.w0[k][a < W ? a : W - 1];
                    }
                    lds_order();
#pragma unroll
                    for (int j = 0; j < PB; ++j) {
                        const int gq = g0 + j;
                        if (gq >= L::NGRP) break;
                        const int a = gq * L::G + it[i].q;
                        T w0a = w0r[j];
                        if constexpr (L::G > 1) w0a = a < W ? w0a : (T)0;
                        if constexpr (CPL == 2) {
                            a0x[i] += w0a * v[j].x;
                            a0y[i] += w0a * v[j].y;
                            a1x[i] += w0a * v[j].z;
                            a1y[i] += w0a * v[j].w;
                        } else {
                            a0x[i] += w0a * v[j].x;
                            a0y[i] += w0a * v[j].y;
                        }
                    }
                }
                const T wu = w1v * w2a;
                part_x += wu * a0x[i];
                part_y += wu * a0y[i];
                if constexpr (CPL == 2) {
                    const T wv = w1v * w2b;
                    part_x += wv * a1x[i];
                    part_y += wv * a1y[i];
                }
            }
            part_x = row_sum15<T>(part_x);
            part_y = row_sum15<T>(part_y);
            if ((lane & 15) == 15) red[wave][k][lane >> 4] = cmake<T>(part_x, part_y);
        }
        lds_order();
        if (lane < nb) {
            const C r0 = red[wave][lane][0], r1 = red[wave][lane][1];
            const C r2 = red[wave][lane][2], r3 = red[wave][lane][3];
            f[pm_out] = cmake<T>((r0.x + r1.x) + (r2.x + r3.x), (r0.y + r1.y) + (r2.y + r3.y));
        }
        lds_order();
    }
}

template <typename T, int D, int W, int NW, int V>
__global__ __launch_bounds__(NW * 64) void k_tile_gather(
    const cplx_t<T> *__restrict__ g, const NodeDesc<T> *__restrict__ desc,
    const uint32_t *__restrict__ perm, const Chunk *__restrict__ chunks,
    const uint32_t *__restrict__ n_chunks, cplx_t<T> *__restrict__ f, Geom ge, BinGeom bg,
    WinConst<T> wc)
{
    const int64_t c = xcd_chunk(blockIdx.x, n_chunks[0]);
    if (c < 0) return;
    tile_gather<T, D, W, NW, V>(g, desc, perm, chunks[c], f, ge, bg, wc);
}

// The gather of every chunk of every job of a batch in one grid.
template <typename T, int D, int W, int NW, int V>
__global__ __launch_bounds__(NW * 64) void k_tile_gather_batch(const GridJob<T> *__restrict__ jobs,
                                                               const BlockMap *__restrict__ map)
{
    const BlockMap bm = map[blockIdx.x];
    const GridJob<T> &J = jobs[bm.job];
    tile_gather<T, D, W, NW, V>(J.g, static_cast<const NodeDesc<T> *>(J.desc), J.perm,
                                J.chunks[bm.blk], J.f_out, J.ge, J.bg, J.wc);
}

// --------------------------------------------------------- x-gradient ----
// d f_j / d x_jt = sum_l g_l n_t phi'(s_t) prod_{t' != t} phi(s_t'): the
// gather's tile and lane mapping with the derivative weights beside the
// weights (the gradient of src/cunfft_kernel.cu:918-1061's sum with the
// window index set held fixed).  Per lane and unit two plane sums, one with
// the plane weights and one with their derivatives; the row and column
// weights (or their derivatives) then give the D partial derivatives, summed
// over the lanes like the gather's value.  Output gx[j d + t] =
// Re(conj(v_j) d f_j / d x_jt) for the upstream v.
template <typename T, int D, int W, int NB> struct alignas(16) DwBatch {
    static constexpr int V0 = (W + 3) & ~3;
    T d0[NB][V0];       // plane-weight derivatives
    T d1[NB][W];        // row-weight derivatives (d = 3)
    T d2[NB][W + 4];    // column-weight derivatives, tap l at l + 1, zero at 0 and past W
};

template <typename T> __device__ __noinline__ T alt_dphi_call(int win, T s, int m, T shape)
{
    return win == WIN_SINC ? sinc_dphi<T>(s, m, shape) : kb_dphi<T>(s, (T)m, shape);
}

// After batch_prologue: the weights' x-derivatives n_t phi'(s) of the batch
// (Gaussian: -2 n_t s / b_t phi(s) from the weight; the others out of line).
template <typename T, int D, int W, int NB>
__device__ __forceinline__ void dw_prologue(DwBatch<T, D, W, NB> &db, const WinBatch<T, D, W, NB> &wb,
                                            int nb, int lane, const WinConst<T> &wc)
{
    if (lane < nb) {
        db.d2[lane][0] = (T)0;
        db.d2[lane][W + 1] = (T)0;
        db.d2[lane][W + 2] = (T)0;
        db.d2[lane][W + 3] = (T)0;
    }
    constexpr int NPASS = (NB * W + 63) / 64;
#pragma unroll
    for (int t = 0; t < D; ++t) {
#pragma unroll
        for (int ps = 0; ps < NPASS; ++ps) {
            const int it = lane + 64 * ps;
            if (it >= nb * W) break;
            const int n = it / W, l = it - n * W;
            const T sv = wb.fx[n][t] + (T)(wc.m - l);
            const T w = t == 0 ? wb.w0[n][l] : (t == D - 1 ? wb.w2[n][l + 1] : wb.w1[n][l]);
            const T dw = wc.win != WIN_GAUSSIAN ? wc.nn[t] * alt_dphi_call<T>(wc.win, sv, wc.m, wc.shp[t])
                                                : wc.dk[t] * sv * w;
            if (t == 0) db.d0[n][l] = dw;
            else if (t == D - 1) db.d2[n][l + 1] = dw;
            else db.d1[n][l] = dw;
        }
    }
    lds_order();
}

// Workgroup shape of the x-gradient: the gather's wavefronts, node batches
// sized with the derivative weights and D partial sums per node.
template <typename T, int D, int W, int V> struct GradShape {
    using L = TileLayout<T, D, W, 1, V>;
    static constexpr int NW = (2 * (L::CELLS * L::CB + 1024) > 160 * 1024) ? 8 : 4;
    static constexpr int NB = batch_nodes<T, D, W>(
        L::CELLS * L::CB, NW, 4 * D * (int)sizeof(cplx_t<T>) + (int)sizeof(DwBatch<T, D, W, 1>));
    static constexpr long lds = (long)L::CELLS * L::CB +
                                NW * ((long)sizeof(WinBatch<T, D, W, NB>) + (long)sizeof(DwBatch<T, D, W, NB>) +
                                      (long)NB * 4 * D * (long)sizeof(cplx_t<T>));
    static constexpr bool fits = L::fits && lds <= kLdsMax;
};

template <typename T, int D, int W, int NW, int V>
__device__ __forceinline__ void tile_grad(const cplx_t<T> *__restrict__ g,
                                          const NodeDesc<T> *__restrict__ desc,
                                          const uint32_t *__restrict__ perm, const Chunk ch,
                                          const cplx_t<T> *__restrict__ v, T *__restrict__ gx,
                                          const Geom &ge, const BinGeom &bg, const WinConst<T> &wc)
{
    using L = TileLayout<T, D, W, 1, V>;
    using VT = vec16_t<T>;
    using C = cplx_t<T>;
    constexpr int CPL = L::CPL;
    constexpr int NT = NW * 64;
    constexpr int PB = L::PB;
    constexpr int NB = GradShape<T, D, W, V>::NB;
    __shared__ __align__(16) C tile[L::CELLS];
    __shared__ WinBatch<T, D, W, NB> wbs[NW];
    __shared__ DwBatch<T, D, W, NB> dbs[NW];
    __shared__ C red[NW][NB][4][D];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const uint32_t nbat = (ch.count + NB - 1) / NB;

    NodeDesc<T> nd{};
    uint32_t pm = 0;
    {
        const uint32_t s = (uint32_t)wave * NB + lane;
        if (lane < NB && s < ch.count) {
            nd = desc[ch.start + s];
            pm = perm[ch.start + s];
        }
    }
    stage_tile<T, D, L, NT>(g, ch.tile, tile, ge, bg, &wbs[0], sizeof(wbs));
    __syncthreads();

    LaneItem it[L::NI];
#pragma unroll
    for (int i = 0; i < L::NI; ++i) it[i] = lane_item<L>(i, lane);
    const char *tb = reinterpret_cast<const char *>(tile);
    WinBatch<T, D, W, NB> &wb = wbs[wave];
    DwBatch<T, D, W, NB> &db = dbs[wave];

    for (uint32_t bi = wave; bi < nbat; bi += NW) {
        const uint32_t s0 = bi * NB;
        const int nb = ch.count - s0 < (uint32_t)NB ? (int)(ch.count - s0) : NB;
        const NodeRec rc = batch_prologue<T, D, W, NB, L>(wb, nb, lane, nd, wc);
        dw_prologue<T, D, W, NB>(db, wb, nb, lane, wc);
        const uint32_t pm_out = pm;
        {
            const uint32_t s = s0 + NW * NB + lane;
            if (lane < NB && s < ch.count) {
                nd = desc[ch.start + s];
                pm = perm[ch.start + s];
            }
        }
        for (int k = 0; k < nb; ++k) {
            const int nbase = rlane_i(rc.nbase, k), par = rlane_i(rc.par, k);
            T sx[D], sy[D];   // partial derivatives along each axis, this lane's share
#pragma unroll
            for (int t = 0; t < D; ++t) sx[t] = sy[t] = 0;
#pragma unroll
            for (int i = 0; i < L::NI; ++i) {
                if (!it[i].ok) continue;
                const uint32_t lb = (uint32_t)(nbase + it[i].off - it[i].q * L::PP) * L::CB;
                const int cw = it[i].u2 - par + 1;
                // plane sums of the unit's cells: with the weights (a) and the
                // derivative weights (b); cells 0 and 1 of the unit
                T a0x = 0, a0y = 0, a1x = 0, a1y = 0, b0x = 0, b0y = 0, b1x = 0, b1y = 0;
                T w1v = 0, w2a = 0, w2b = 0, d1v = 0, d2a = 0, d2b = 0;
#pragma unroll
                for (int g0 = 0; g0 < L::NGRP; g0 += PB) {
                    VT vv[PB];
#pragma unroll
                    for (int j = 0; j < PB; ++j) {
                        const int gq = g0 + j;
                        if (gq >= L::NGRP) break;
                        int a = gq * L::G + it[i].q;
                        a = a < W ? a : W - 1;   // planes past the window: weights 0 below
                        vv[j] = *reinterpret_cast<const VT *>(tb + lb + (uint32_t)(a * L::PP * L::CB));
                    }
                    lds_order();
                    if (g0 == 0) {
                        w1v = D == 3 ? wb.w1[k][it[i].r] : (T)1;
                        d1v = D == 3 ? db.d1[k][it[i].r] : (T)0;
                        w2a = wb.w2[k][cw];
                        d2a = db.d2[k][cw];
                        w2b = CPL == 2 ? wb.w2[k][cw + 1] : (T)0;
                        d2b = CPL == 2 ? db.d2[k][cw + 1] : (T)0;
                    }
                    T w0r[PB], d0r[PB];
#pragma unroll
                    for (int j = 0; j < PB; ++j) {
                        const int a = (g0 + j) * L::G + it[i].q;
                        w0r[j] = wb.w0[k][a < W ? a : W - 1];
                        d0r[j] = db.d0[k][a < W ? a : W - 1];
                    }
                    lds_order();
#pragma unroll
                    for (int j = 0; j < PB; ++j) {
                        const int gq = g0 + j;
                        if (gq >= L::NGRP) break;
                        const int a = gq * L::G + it[i].q;
                        T w0a = w0r[j], d0a = d0r[j];
                        if constexpr (L::G > 1) {
                            w0a = a < W ? w0a : (T)0;
                            d0a = a < W ? d0a : (T)0;
                        }
                        if constexpr (CPL == 2) {
                            a0x += w0a * vv[j].x;
                            a0y += w0a * vv[j].y;
                            a1x += w0a * vv[j].z;
                            a1y += w0a * vv[j].w;
                            b0x += d0a * vv[j].x;
                            b0y += d0a * vv[j].y;
                            b1x += d0a * vv[j].z;
                            b1y += d0a * vv[j].w;
                        } else {
                            a0x += w0a * vv[j].x;
                            a0y += w0a * vv[j].y;
                            b0x += d0a * vv[j].x;
                            b0y += d0a * vv[j].y;
                        }
                    }
                }
                // axis 0: plane derivative, row and column weights
                const T wu = w1v * w2a, wv = w1v * w2b;
                sx[0] += wu * b0x + wv * b1x;
                sy[0] += wu * b0y + wv * b1y;
                if constexpr (D == 3) {   // axis 1: row derivative
                    const T ru = d1v * w2a, rv = d1v * w2b;
                    sx[1] += ru * a0x + rv * a1x;
                    sy[1] += ru * a0y + rv * a1y;
                }
                // last axis: column derivative
                const T cu = w1v * d2a, cv = w1v * d2b;
                sx[D - 1] += cu * a0x + cv * a1x;
                sy[D - 1] += cu * a0y + cv * a1y;
            }
#pragma unroll
            for (int t = 0; t < D; ++t) {
                const T rx = row_sum15<T>(sx[t]), ry = row_sum15<T>(sy[t]);
                if ((lane & 15) == 15) red[wave][k][lane >> 4][t] = cmake<T>(rx, ry);
            }
        }
        lds_order();
        if (lane < nb) {
            const C vj = v[pm_out];
#pragma unroll
            for (int t = 0; t < D; ++t) {
                const C r0 = red[wave][lane][0][t], r1 = red[wave][lane][1][t];
                const C r2 = red[wave][lane][2][t], r3 = red[wave][lane][3][t];
                const T Sx = (r0.x + r1.x) + (r2.x + r3.x), Sy = (r0.y + r1.y) + (r2.y + r3.y);
                gx[(int64_t)pm_out * D + t] = vj.x * Sx + vj.y * Sy;
            }
        }
        lds_order();
    }
}

template <typename T, int D, int W, int NW, int V>
__global__ __launch_bounds__(NW * 64) void k_tile_grad(
    const cplx_t<T> *__restrict__ g, const NodeDesc<T> *__restrict__ desc,
    const uint32_t *__restrict__ perm, const Chunk *__restrict__ chunks,
    const uint32_t *__restrict__ n_chunks, const cplx_t<T> *__restrict__ v, T *__restrict__ gx, Geom ge,
    BinGeom bg, WinConst<T> wc)
{
    const int64_t c = xcd_chunk(blockIdx.x, n_chunks[0]);
    if (c < 0) return;
    tile_grad<T, D, W, NW, V>(g, desc, perm, chunks[c], v, gx, ge, bg, wc);
}

// ------------------------------------------------------------ spread ----
// Slot stores carry the nontemporal hint: the fold reads each slot once, after
// the whole spread has written every other slot, so keeping them in the caches
// only evicts what the fold and the next passes read (measured on gfx950: the
// 3D FP32 fold 8% faster, the spread unchanged).
__device__ __forceinline__ void stream_store(float4 *p, float4 v)
{
    typedef float f4 __attribute__((ext_vector_type(4)));
    const f4 t = {v.x, v.y, v.z, v.w};
    __builtin_nontemporal_store(t, reinterpret_cast<f4 *>(p));
}
__device__ __forceinline__ void stream_store(float2 *p, float2 v)
{
    typedef float f2 __attribute__((ext_vector_type(2)));
    const f2 t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<f2 *>(p));
}
__device__ __forceinline__ void stream_store(double2 *p, double2 v)
{
    typedef double d2 __attribute__((ext_vector_type(2)));
    const d2 t = {v.x, v.y};
    __builtin_nontemporal_store(t, reinterpret_cast<d2 *>(p));
}

// Padded tile (LDS) -> the chunk's slot, destination-blocked: the block that
// lands on a neighbouring tile is contiguous.  COMP: the slot receives the
// sum minus its Kahan term (the second tile at EOFF).
template <typename L, bool COMP, int NT, typename VT>
__device__ __forceinline__ void tile_to_slot(const VT *tile, cplx_t<typename L::Real> *__restrict__ slot)
{
    using T = typename L::Real;
    using C = cplx_t<T>;
    constexpr int CPL = L::CPL;
    constexpr uint32_t EOFF = COMP ? (L::CELLS / CPL + L::DUMMY_UNITS) * 16 : 0;
    const int tid = threadIdx.x;
    const C *tc = reinterpret_cast<const C *>(tile);
#pragma unroll
    for (int q0 = 0; q0 < L::Q0; ++q0)
#pragma unroll
        for (int q1 = 0; q1 < L::Q1; ++q1)
#pragma unroll
            for (int qL = 0; qL < L::QL; ++qL) {
                const int E0 = L::ext(L::T0, L::P0, q0), E1 = L::ext(L::T1, L::P1, q1),
                          EL = L::ext(L::TL, L::PL, qL);
                const int n = E0 * E1 * EL;
                C *dst = slot + L::block_off(q0, q1, qL);
                if (CPL == 2 && EL % 2 == 0) {
                    // two cells per 16-byte move: block rows and padded rows start
                    // on even cells (even pitch, even tile extents, even slot size)
                    const int np = n / 2;
                    for (int c0 = 0; c0 < np; c0 += 8 * NT) {
                        VT v[8];
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            int c = c0 + tid + u * NT;
                            c = 2 * (c < np ? c : np - 1);
                            const int r0 = c / (E1 * EL), rr = c - r0 * (E1 * EL);
                            const int r1 = rr / EL, rL = rr - r1 * EL;
                            const int src = (q0 * L::T0 + r0) * L::PP + (q1 * L::T1 + r1) * L::RP +
                                            qL * L::TL + rL;
                            v[u] = tile[src / 2];
                            if constexpr (COMP) {
                                const VT e = tile[src / 2 + EOFF / 16];
                                v[u].x -= e.x;
                                v[u].y -= e.y;
                                v[u].z -= e.z;
                                v[u].w -= e.w;
                            }
                        }
#pragma unroll
                        for (int u = 0; u < 8; ++u) {
                            const int c = c0 + tid + u * NT;
                            if (c < np) stream_store(reinterpret_cast<VT *>(dst) + c, v[u]);
                        }
                    }
                    continue;
                }
                for (int c0 = 0; c0 < n; c0 += 8 * NT) {
                    C v[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        int c = c0 + tid + u * NT;
                        c = c < n ? c : n - 1;
                        const int r0 = c / (E1 * EL), rr = c - r0 * (E1 * EL);
                        const int r1 = rr / EL, rL = rr - r1 * EL;
                        const int src = (q0 * L::T0 + r0) * L::PP + (q1 * L::T1 + r1) * L::RP +
                                        qL * L::TL + rL;
                        v[u] = tc[src];
                        if constexpr (COMP) {
                            const C e = tc[src + EOFF / sizeof(C)];
                            v[u] = cmake<T>(v[u].x - e.x, v[u].y - e.y);
                        }
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int c = c0 + tid + u * NT;
                        if (c < n) stream_store(dst + c, v[u]);
                    }
                }
            }
}

// NW wavefronts share the tile; wavefront w owns the planes p0 = w (mod NW),
// so wavefronts never touch the same cell, and every wavefront walks all the
// chunk's nodes in order (each wavefront keeps its own weight batch).  The
// cells of one node are disjoint, so a node's reads are all issued before its
// adds and writes; consecutive nodes of one wavefront may overlap, so their
// read-add-write sequences stay in program order.  When a plane group can
// reach past the window, the lanes concerned read and write a private dummy
// unit (TileLayout::dummy_byte).  fs holds the coefficients in sorted node order.
template <typename T, int D, int W, int NW, int V, bool COMP>
__device__ __forceinline__ void tile_spread(const NodeDesc<T> *__restrict__ desc,
                                            const cplx_t<T> *__restrict__ fs, const Chunk ch,
                                            cplx_t<T> *__restrict__ slot, const WinConst<T> &wc)
{
    using L = TileLayout<T, D, W, NW, V>;
    using VT = vec16_t<T>;
    using C = cplx_t<T>;
    constexpr int CPL = L::CPL;
    constexpr int NT = NW * 64;
    constexpr int PB = L::PB;
    constexpr bool DUMMY = L::NEED_DUMMY;
    constexpr int TILE_UNITS = L::CELLS / CPL + L::DUMMY_UNITS;
    constexpr int NB = batch_nodes<T, D, W>(TILE_UNITS * 16 * (COMP ? 2 : 1), NW, 0);
    // COMP: a second tile of Kahan compensation terms follows the sums
    constexpr uint32_t EOFF = COMP ? TILE_UNITS * 16 : 0;
    __shared__ __align__(16) VT tile[COMP ? 2 * TILE_UNITS : TILE_UNITS];
    __shared__ WinBatch<T, D, W, NB> wbs[NW];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;

    NodeDesc<T> nd{};
    C fv = cmake<T>(0, 0);
    if (lane < NB && (uint32_t)lane < ch.count) {
        nd = desc[ch.start + lane];
        fv = fs[ch.start + lane];
    }
    {
        VT z;
        if constexpr (CPL == 2) z = make_float4(0.f, 0.f, 0.f, 0.f);
        else z = make_double2(0.0, 0.0);
        for (int idx = tid; idx < (COMP ? 2 : 1) * TILE_UNITS; idx += NT) tile[idx] = z;
    }
    __syncthreads();

    LaneItem it[L::NI];
#pragma unroll
    for (int i = 0; i < L::NI; ++i) it[i] = lane_item<L>(i, lane);
    char *tb = reinterpret_cast<char *>(tile);
    const uint32_t dummy = L::dummy_byte(lane);
    WinBatch<T, D, W, NB> &wb = wbs[wave];

    for (uint32_t s0 = 0; s0 < ch.count; s0 += NB) {
        // scalar, so the node loop exits on a scalar compare
        const int nb = __builtin_amdgcn_readfirstlane(ch.count - s0 < (uint32_t)NB ? (int)(ch.count - s0) : NB);
        const NodeRec rc = batch_prologue<T, D, W, NB, L>(wb, nb, lane, nd, wc);
        const C fc = fv;
        {
            const uint32_t s = s0 + NB + lane;
            if (lane < NB && s < ch.count) {
                nd = desc[ch.start + s];
                fv = fs[ch.start + s];
            }
        }
        // lane n's view of node n for this wavefront, computed once per batch:
        // first plane residue a0, tile byte base, column weight offset
        int na0, ntb, nw2;
        {
            na0 = NW == 1 ? 0 : ((wave - rc.lo0) & (NW - 1));
            ntb = (rc.nbase + na0 * L::PP) * L::CB;
            nw2 = lane * (W + 4) - rc.par + 1;
        }
        const T *w2base = &wb.w2[0][0];
        for (int k = 0; k < nb; ++k) {
            const int a0 = NW == 1 ? 0 : rlane_i(na0, k);
            const uint32_t tbase = (uint32_t)rlane_i(ntb, k);
            const T *w2k = w2base + rlane_i(nw2, k);
            const T fr = rlane_t(fc.x, k), fi = rlane_t(fc.y, k);
#pragma unroll
            for (int i = 0; i < L::NI; ++i) {
                if (!it[i].ok) continue;
                const uint32_t lb = tbase + (uint32_t)it[i].off * L::CB;
                const int cw = it[i].u2;
#pragma unroll
                for (int g0 = 0; g0 < L::NGRP; g0 += PB) {
                    VT v[PB], ev[PB];
                    uint32_t ad[PB];
#pragma unroll
                    for (int j = 0; j < PB; ++j) {
                        const int gq = g0 + j;
                        if (gq >= L::NGRP) break;
                        ad[j] = lb + (uint32_t)(NW * gq * L::G * L::PP * L::CB);
                        if constexpr (DUMMY) {
                            if ((NW - 1) + NW * (gq * L::G + L::G - 1) >= W) {
                                const int a = a0 + NW * (gq * L::G + it[i].q);
                                ad[j] = a < W ? ad[j] : dummy;
                            }
                        }
                        v[j] = *reinterpret_cast<const VT *>(tb + ad[j]);
                        if constexpr (COMP) ev[j] = *reinterpret_cast<const VT *>(tb + EOFF + ad[j]);
                    }
                    // the node's weights are read only after its tile reads are
                    // issued, and all of them before any use: one LDS round trip
                    lds_order();
                    const T w1v = D == 3 ? wb.w1[k][it[i].r] : (T)1;
                    const T w2a = w2k[cw];
                    const T w2b = CPL == 2 ? w2k[cw + 1] : (T)0;
                    T w0r[PB];
#pragma unroll
                    for (int j = 0; j < PB; ++j) {
                        const int a = a0 + NW * ((g0 + j) * L::G + it[i].q);
                        w0r[j] = wb.w0[k][a < W ? a : W - 1];
                    }
                    lds_order();
                    const T wu = w1v * w2a;
                    const T fux = fr * wu, fuy = fi * wu;
                    T fvx = 0, fvy = 0;
                    if constexpr (CPL == 2) {
                        const T wv = w1v * w2b;
                        fvx = fr * wv;
                        fvy = fi * wv;
                    }
#pragma unroll
                    for (int j = 0; j < PB; ++j) {
                        const int gq = g0 + j;
                        if (gq >= L::NGRP) break;
                        const T w0a = w0r[j];
                        if constexpr (COMP) {
                            // Kahan: e carries the low-order part lost by the sums
                            kahan_add(v[j].x, ev[j].x, w0a * fux);
                            kahan_add(v[j].y, ev[j].y, w0a * fuy);
                            if constexpr (CPL == 2) {
                                kahan_add(v[j].z, ev[j].z, w0a * fvx);
                                kahan_add(v[j].w, ev[j].w, w0a * fvy);
                            }
                            *reinterpret_cast<VT *>(tb + EOFF + ad[j]) = ev[j];
                        } else if constexpr (CPL == 2) {
                            v[j].x += w0a * fux;
                            v[j].y += w0a * fuy;
                            v[j].z += w0a * fvx;
                            v[j].w += w0a * fvy;
                        } else {
                            v[j].x += w0a * fux;
                            v[j].y += w0a * fuy;
                        }
                        *reinterpret_cast<VT *>(tb + ad[j]) = v[j];
                    }
                }
            }
            lds_order();
        }
    }
    __syncthreads();

    tile_to_slot<L, COMP, NT>(tile, slot);
}

// Work-queue state of the persistent held-unit spread (k_tile_spread_queue):
// workgroups take chunks from a counter, and each prefetches the first nodes
// of its next chunk (descriptors and coefficients) while the current tile
// goes to its slot.  FP32 reads the coefficients through the permutation,
// which saves the sorted copy of f; FP64, at two workgroups per CU, hides
// the dependent 16-byte loads less well and keeps the copy.  Thread 0 takes
// the next ticket when a chunk starts and publishes it in LDS only after the
// node loop, so the atomic's round trip hides behind the chunk's work.
template <typename T> struct SpreadQueue {
    uint32_t *counter;      // work counter, zero at launch
    const Chunk *chunks;
    const uint32_t *perm;   // node of each sorted position: the coefficients are read
    const cplx_t<T> *f;     // through it (FP32), or from the sorted copy fs (perm NULL)
    const cplx_t<T> *fs;
    uint32_t n;             // chunk count
    uint32_t *s_next;       // LDS: the next chunk, published by thread 0
    uint32_t ticket;        // thread 0: its next chunk
    NodeDesc<T> nd;         // lanes < NB: the chunk's first nodes, prefetched
    cplx_t<T> fv;
};

// Adjoint with the tile units held across nodes (d = 3, one plane per lane
// group).  The nodes of a tile come sorted by row, column unit and plane of
// their window start, so consecutive nodes often cover the same rows and
// column units of the padded tile, and a wavefront's first plane
// p = lo_0 + ((w - lo_0) mod NW) is then often the same as well: the units
// the wavefront updates for the two nodes are the same.  A lane keeps them in
// registers until a node needs other units, and only then writes them back
// and reads the new ones, so a sequence of r such nodes costs the LDS accesses
// of one.  The plane weights come grouped by residue (zero past the window),
// so a node whose wavefront planes run past its window adds zero there.  Order
// of the sums per cell: the nodes in sorted order, as in the direct update.
template <typename T, int D, int W, int NW, int V, bool Q = false>
__device__ __forceinline__ void tile_spread_runs(const NodeDesc<T> *__restrict__ desc,
                                                 const cplx_t<T> *__restrict__ fs, const Chunk ch,
                                                 cplx_t<T> *__restrict__ slot, const WinConst<T> &wc,
                                                 SpreadQueue<T> *q = nullptr)
{
    using L = TileLayout<T, D, W, NW, V>;
    static_assert(D == 3 && L::G == 1, "units held across nodes: one plane per lane group");
    using VT = vec16_t<T>;
    using C = cplx_t<T>;
    constexpr int CPL = L::CPL;
    constexpr int NT = NW * 64;
    constexpr int NI = L::NI;
    constexpr int OWN = L::OWN;
    constexpr bool DUMMY = L::NEED_DUMMY;
    constexpr int TILE_UNITS = L::CELLS / CPL + L::DUMMY_UNITS;
    constexpr int NB = batch_nodes<T, D, W, NW, 32>(TILE_UNITS * 16, 1, 0);
    using WB = WinBatch<T, D, W, NB, NW>;
    __shared__ __align__(16) VT tile[TILE_UNITS];
    __shared__ WB wb;   // one batch of weights for all wavefronts
    // wave-uniform values stay in scalar registers (they drive the branch below)
    const int tid = threadIdx.x, lane = tid & 63, wave = __builtin_amdgcn_readfirstlane(tid >> 6);

    NodeDesc<T> nd{};
    C fv = cmake<T>(0, 0);
    if constexpr (Q) {
        nd = q->nd;   // prefetched during the previous chunk's slot write
        fv = q->fv;
        if (tid == 0) q->ticket = atomicAdd(q->counter, 1u);
    } else if (lane < NB && (uint32_t)lane < ch.count) {
        nd = desc[ch.start + lane];
        fv = fs[ch.start + lane];
    }
    {
        VT z;
        if constexpr (CPL == 2) z = make_float4(0.f, 0.f, 0.f, 0.f);
        else z = make_double2(0.0, 0.0);
        for (int idx = tid; idx < TILE_UNITS; idx += NT) tile[idx] = z;
    }
    __syncthreads();

    bool ok[NI];
    int ur[NI], uc[NI];
    uint32_t offb[NI];
#pragma unroll
    for (int i = 0; i < NI; ++i) {
        const LaneItem it = lane_item<L>(i, lane);
        ok[i] = it.ok;
        ur[i] = it.r;
        uc[i] = it.u2 + 1;
        offb[i] = (uint32_t)it.off * L::CB;
    }
    char *tb = reinterpret_cast<char *>(tile);
    const uint32_t dummy = L::dummy_byte(lane);

    VT v[NI][OWN];          // the held units: plane p_cur + NW i of lane item ii
    uint32_t ad[NI][OWN];
    int cur_key = -1;
    static_assert(L::T0 + NW - 2 < 64, "a wavefront's first plane fits the key's low 6 bits");
    // the batch's node count in a scalar register: the node loop's exit test is
    // then a scalar compare (left alone, the clamp is a vector med3)
    for (uint32_t s0 = 0; s0 < ch.count; s0 += NB) {
        const int nb = __builtin_amdgcn_readfirstlane(ch.count - s0 < (uint32_t)NB ? (int)(ch.count - s0) : NB);
        if (s0 > 0) __syncthreads();   // every wavefront is done with the previous batch
        const NodeRec rc = batch_prologue_wg<T, D, W, NB, L, NW, NW>(wb, nb, lane, wave, nd, fv, wc);
        {
            const uint32_t s = s0 + NB + lane;
            if (lane < NB && s < ch.count) {
                nd = desc[ch.start + s];
                if constexpr (Q) fv = q->perm ? q->f[q->perm[ch.start + s]] : q->fs[ch.start + s];
                else fv = fs[ch.start + s];
            }
        }
        // lane n's view of node n for this wavefront: key (row/column base and
        // first plane p), tile byte base of plane p, and element offsets of its
        // plane weights (residue a0) and column weights (parity shift)
        int nkey, ntb, nw0, nw2;
        {
            const int a0 = NW == 1 ? 0 : ((wave - rc.lo0) & (NW - 1));
            const int p = rc.lo0 + a0;
            const int rcb = rc.nbase - rc.lo0 * L::PP;
            nkey = rcb * 64 + p;
            ntb = (p * L::PP + rcb) * L::CB;
            nw0 = lane * WB::W0N + a0 * WB::OWNP;
            nw2 = lane * (W + 4) - rc.par;
        }
        const T *w0base = &wb.w0[0][0];
        const T *w2base = &wb.w2[0][0];
        for (int k = 0; k < nb; ++k) {
            const int key = __builtin_amdgcn_readfirstlane(rlane_i(nkey, k));
            if (key != cur_key) {
                // write back the held units, then read the node's (in order per wavefront)
                if (cur_key >= 0) {
#pragma unroll
                    for (int ii = 0; ii < NI; ++ii) {
                        if (!ok[ii]) continue;
#pragma unroll
                        for (int i = 0; i < OWN; ++i) *reinterpret_cast<VT *>(tb + ad[ii][i]) = v[ii][i];
                    }
                }
                const int p = key & 63;
                const uint32_t tbase = (uint32_t)rlane_i(ntb, k);
#pragma unroll
                for (int ii = 0; ii < NI; ++ii) {
                    if (!ok[ii]) continue;
                    const uint32_t lb = tbase + offb[ii];
#pragma unroll
                    for (int i = 0; i < OWN; ++i) {
                        uint32_t a = lb + (uint32_t)(NW * i * L::PP * L::CB);
                        if constexpr (DUMMY) a = p + NW * i < L::P0 ? a : dummy;
                        ad[ii][i] = a;
                        v[ii][i] = *reinterpret_cast<const VT *>(tb + a);
                    }
                }
                cur_key = key;
            }
            // the node's weights: its planes p + NW i (grouped by residue a0), rows, columns
            T w0[OWN];
            {
                const VT *src = reinterpret_cast<const VT *>(w0base + rlane_i(nw0, k));
#pragma unroll
                for (int q = 0; q < WB::OWNP / WB::VW; ++q) {
                    const VT e = src[q];
                    if constexpr (sizeof(T) == 4) {
                        if (4 * q + 0 < OWN) w0[4 * q + 0] = e.x;
                        if (4 * q + 1 < OWN) w0[4 * q + 1] = e.y;
                        if (4 * q + 2 < OWN) w0[4 * q + 2] = e.z;
                        if (4 * q + 3 < OWN) w0[4 * q + 3] = e.w;
                    } else {
                        if (2 * q + 0 < OWN) w0[2 * q + 0] = e.x;
                        if (2 * q + 1 < OWN) w0[2 * q + 1] = e.y;
                    }
                }
            }
            const T *w2k = w2base + rlane_i(nw2, k);
#pragma unroll
            for (int ii = 0; ii < NI; ++ii) {
                // row weight times the coefficient, then the column weights
                const T w1r = wb.w1[k][2 * ur[ii]], w1i = wb.w1[k][2 * ur[ii] + 1];
                const T w2a = w2k[uc[ii]];
                const T fux = w1r * w2a, fuy = w1i * w2a;
                T fvx = 0, fvy = 0;
                if constexpr (CPL == 2) {
                    const T w2b = w2k[uc[ii] + 1];
                    fvx = w1r * w2b;
                    fvy = w1i * w2b;
                }
#pragma unroll
                for (int i = 0; i < OWN; ++i) {
                    v[ii][i].x += w0[i] * fux;
                    v[ii][i].y += w0[i] * fuy;
                    if constexpr (CPL == 2) {
                        v[ii][i].z += w0[i] * fvx;
                        v[ii][i].w += w0[i] * fvy;
                    }
                }
            }
        }
        lds_order();
    }
    if (cur_key >= 0) {
#pragma unroll
        for (int ii = 0; ii < NI; ++ii) {
            if (!ok[ii]) continue;
#pragma unroll
            for (int i = 0; i < OWN; ++i) *reinterpret_cast<VT *>(tb + ad[ii][i]) = v[ii][i];
        }
    }
    if constexpr (Q) {
        if (tid == 0) *q->s_next = q->ticket;
    }
    __syncthreads();
    if constexpr (Q) {
        // the next chunk's first nodes, in flight during the slot write
        const uint32_t cn = *q->s_next;
        if (cn < q->n) {
            const Chunk nc = q->chunks[cn];
            if (lane < NB && (uint32_t)lane < nc.count) {
                q->nd = desc[nc.start + lane];
                q->fv = q->perm ? q->f[q->perm[nc.start + lane]] : q->fs[nc.start + lane];
            }
        }
    }
    tile_to_slot<L, false, NT>(tile, slot);
}

template <typename T, int D, int W, int NW, int V, bool COMP>
__global__ __launch_bounds__(NW * 64) void k_tile_spread(
    const NodeDesc<T> *__restrict__ desc, const cplx_t<T> *__restrict__ fs,
    const Chunk *__restrict__ chunks, const uint32_t *__restrict__ n_chunks,
    cplx_t<T> *__restrict__ slots, WinConst<T> wc)
{
    using L = TileLayout<T, D, W, NW, V>;
    if (blockIdx.x >= n_chunks[0]) return;
    if constexpr (D == 3 && !COMP && L::G == 1)
        tile_spread_runs<T, D, W, NW, V>(desc, fs, chunks[blockIdx.x],
                                         slots + (int64_t)blockIdx.x * L::SC, wc);
    else
        tile_spread<T, D, W, NW, V, COMP>(desc, fs, chunks[blockIdx.x],
                                          slots + (int64_t)blockIdx.x * L::SC, wc);
}

// Persistent form of the 3D held-unit spread: a grid of resident workgroups
// takes chunks from the binning's work counter (n_chunks[1], cleared before
// the launch) until it runs past the chunk count, so a chunk's first node
// loads overlap the previous chunk's slot write and no chunk waits for a
// workgroup launch.  Every workgroup takes one ticket past the last chunk and
// exits.
template <typename T, int D, int W, int NW, int V>
__global__ __launch_bounds__(NW * 64) void k_tile_spread_queue(
    const NodeDesc<T> *__restrict__ desc, const uint32_t *__restrict__ perm,
    const cplx_t<T> *__restrict__ f, const cplx_t<T> *__restrict__ fs, const Chunk *__restrict__ chunks,
    uint32_t *__restrict__ n_chunks, cplx_t<T> *__restrict__ slots, WinConst<T> wc)
{
    using L = TileLayout<T, D, W, NW, V>;
    __shared__ uint32_t s_next;
    const int tid = threadIdx.x, lane = tid & 63;
    SpreadQueue<T> q;
    q.counter = n_chunks + 1;
    q.chunks = chunks;
    q.perm = perm;
    q.f = f;
    q.fs = fs;
    q.n = n_chunks[0];
    q.s_next = &s_next;
    q.ticket = 0;
    q.nd = NodeDesc<T>{};
    q.fv = cmake<T>(0, 0);
    if (tid == 0) s_next = atomicAdd(q.counter, 1u);
    __syncthreads();
    uint32_t c = s_next;
    if (c < q.n) {
        const Chunk ch = chunks[c];
        constexpr int TILE_UNITS = L::CELLS / L::CPL + L::DUMMY_UNITS;
        constexpr int NB = batch_nodes<T, D, W, NW, 32>(TILE_UNITS * 16, 1, 0);
        if (lane < NB && (uint32_t)lane < ch.count) {
            q.nd = desc[ch.start + lane];
            q.fv = perm ? f[perm[ch.start + lane]] : fs[ch.start + lane];
        }
    }
    while (c < q.n) {
        tile_spread_runs<T, D, W, NW, V, true>(desc, nullptr, chunks[c], slots + (int64_t)c * L::SC, wc,
                                               &q);
        c = s_next;        // published before the chunk's last barrier
        __syncthreads();   // the slot write's LDS reads are done before the next chunk clears the tile
    }
}

template <typename T, int D, int W, int NW, int V>
__global__ __launch_bounds__(NW * 64) void k_tile_spread_batch(const GridJob<T> *__restrict__ jobs,
                                                               const BlockMap *__restrict__ map)
{
    using L = TileLayout<T, D, W, NW, V>;
    const BlockMap bm = map[blockIdx.x];
    const GridJob<T> &J = jobs[bm.job];
    if constexpr (D == 3 && L::G == 1)
        tile_spread_runs<T, D, W, NW, V>(static_cast<const NodeDesc<T> *>(J.desc), J.fs,
                                         J.chunks[bm.blk], J.slots + (int64_t)bm.blk * L::SC, J.wc);
    else
        tile_spread<T, D, W, NW, V, false>(static_cast<const NodeDesc<T> *>(J.desc), J.fs,
                                           J.chunks[bm.blk], J.slots + (int64_t)bm.blk * L::SC, J.wc);
}

// Coefficients in sorted node order for the adjoint: fs[s] = f[perm[s]].
template <typename T>
__global__ __launch_bounds__(256) void k_sorted_coeffs(const cplx_t<T> *__restrict__ f,
                                                       const uint32_t *__restrict__ perm,
                                                       cplx_t<T> *__restrict__ fs, int64_t M)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s < M) fs[s] = f[perm[s]];
}

template <typename T>
__global__ __launch_bounds__(256) void k_sorted_coeffs_batch(const GridJob<T> *__restrict__ jobs,
                                                             const BlockMap *__restrict__ map)
{
    const BlockMap bm = map[blockIdx.x];
    const GridJob<T> &J = jobs[bm.job];
    const int64_t s = (int64_t)bm.blk * 256 + threadIdx.x;
    if (s < J.ge.M) J.fs[s] = J.f_in[J.perm[s]];
}

// -------------------------------------------------------------- fold ----
// g on one grid tile = sum over the chunks of the tiles whose padded region
// covers it of the matching slot block.  Per-cell form (FP64; the FP32 form
// by column segments follows): one workgroup per tile and slice of 512 core
// cells (blockIdx.y); each block a thread reads is contiguous over the
// workgroup.  The first chunk of the
// candidate tiles of one leading-axis offset is loaded at once (a dense tile's
// further chunks follow), and the sums run in a fixed order: the result does
// not depend on scheduling.
template <typename T, int D, int W, int PS, int V, bool COMP>
__device__ __forceinline__ void tile_fold(const cplx_t<T> *__restrict__ slots,
                                          const uint32_t *__restrict__ tile_count,
                                          const uint32_t *__restrict__ chunk_start,
                                          cplx_t<T> *__restrict__ g, const Geom &ge,
                                          const BinGeom &bg, uint32_t tile, int slice)
{
    using L = TileLayout<T, D, W, PS, V>;
    using C = cplx_t<T>;
    constexpr int CORE = L::T0 * L::T1 * L::TL;
    constexpr int CPT = CORE >= 512 ? 2 : (CORE + 255) / 256;   // cells per thread per slice
    constexpr int SLICES = CORE >= 512 ? CORE / 512 : 1;
    const int slice0 = SLICES > 1 ? slice * 256 * CPT : 0;       // first core cell of this slice
    constexpr int QR = L::Q1 * L::QL;            // candidates per leading-axis offset
    uint32_t tk[3] = {0, 0, 0};
    {
        uint32_t rem = tile;
        for (int t = D - 1; t >= 0; --t) {
            tk[t] = rem % (uint32_t)bg.nt[t];
            rem /= (uint32_t)bg.nt[t];
        }
    }
    int r0[CPT], r1[CPT], rL[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int c = slice0 + threadIdx.x + 256 * k;
        r0[k] = c / (L::T1 * L::TL);
        r1[k] = (c / L::TL) % L::T1;
        rL[k] = c % L::TL;
    }
    T sx[CPT], sy[CPT], cx[CPT], cy[CPT];   // sums and (COMP) Neumaier corrections
#pragma unroll
    for (int k = 0; k < CPT; ++k) cx[k] = cy[k] = 0;
#pragma unroll
    for (int k = 0; k < CPT; ++k) sx[k] = sy[k] = 0;
#pragma unroll
    for (int q0 = 0; q0 < L::Q0; ++q0) {
        uint32_t c0[QR], nch[QR];
        int li[QR][CPT];
        bool ok[QR][CPT];
        C v[QR][CPT];
#pragma unroll
        for (int qq = 0; qq < QR; ++qq) {
            const int q1 = qq / L::QL, qL = qq % L::QL;
            const int E0 = L::ext(L::T0, L::P0, q0), E1 = L::ext(L::T1, L::P1, q1),
                      EL = L::ext(L::TL, L::PL, qL);
            const int qv[3] = {q0, D == 3 ? q1 : qL, qL};
            uint32_t src = 0;
            for (int t = 0; t < D; ++t) {
                const uint32_t ntt = (uint32_t)bg.nt[t];
                src = src * ntt + (tk[t] + ntt * 4 - (uint32_t)qv[t]) % ntt;
            }
            const uint32_t cnt = tile_count[src];
            c0[qq] = chunk_start[src];
            nch[qq] = (cnt + kChunk - 1) / kChunk;
            const int boff = L::block_off(q0, q1, qL);
#pragma unroll
            for (int k = 0; k < CPT; ++k) {
                ok[qq][k] = cnt > 0 && slice0 + threadIdx.x + 256 * k < CORE && r0[k] < E0 &&
                            r1[k] < E1 && rL[k] < EL;
                li[qq][k] = boff + (r0[k] * E1 + r1[k]) * EL + rL[k];
            }
        }
#pragma unroll
        for (int qq = 0; qq < QR; ++qq)
#pragma unroll
            for (int k = 0; k < CPT; ++k) {
                // a candidate with no nodes reads chunk 0 and adds nothing
                const C *sp = slots + (int64_t)(ok[qq][k] ? c0[qq] : 0) * L::SC;
                v[qq][k] = sp[ok[qq][k] ? li[qq][k] : 0];
            }
#pragma unroll
        for (int qq = 0; qq < QR; ++qq)
#pragma unroll
            for (int k = 0; k < CPT; ++k)
                if (ok[qq][k]) {
                    fold_add<COMP>(sx[k], cx[k], v[qq][k].x);
                    fold_add<COMP>(sy[k], cy[k], v[qq][k].y);
                }
        // further chunks of dense candidate tiles, four loads in flight
#pragma unroll
        for (int qq = 0; qq < QR; ++qq)
            for (uint32_t c = 1; c < nch[qq]; c += 4) {
                C w[4][CPT];
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const bool in = c + u < nch[qq];
                    const C *sp = slots + (int64_t)(in ? c0[qq] + c + u : 0) * L::SC;
#pragma unroll
                    for (int k = 0; k < CPT; ++k) w[u][k] = sp[ok[qq][k] && in ? li[qq][k] : 0];
                }
#pragma unroll
                for (int u = 0; u < 4; ++u)
#pragma unroll
                    for (int k = 0; k < CPT; ++k)
                        if (ok[qq][k] && c + u < nch[qq]) {
                            fold_add<COMP>(sx[k], cx[k], w[u][k].x);
                            fold_add<COMP>(sy[k], cy[k], w[u][k].y);
                        }
            }
    }
    const uint32_t nL = (uint32_t)ge.n[D - 1];
#pragma unroll
    for (int k = 0; k < CPT; ++k)
        if (slice0 + threadIdx.x + 256 * k < CORE) {
            uint32_t cell = (tk[0] * L::T0 + r0[k]) * (D == 3 ? (uint32_t)(ge.n[1] * ge.n[2]) : nL);
            if (D == 3) cell += (tk[1] * L::T1 + r1[k]) * nL;
            cell += tk[D - 1] * L::TL + rL[k];
            g[cell] = cmake<T>(sx[k] + cx[k], sy[k] + cy[k]);
        }
}

// Two consecutive cells with one 16-byte load (FP32; the FP64 form is two loads).
__device__ __forceinline__ void load_pair(const float2 *p, float2 &a, float2 &b)
{
    const float4 t = *reinterpret_cast<const float4 *>(p);
    a = make_float2(t.x, t.y);
    b = make_float2(t.z, t.w);
}
__device__ __forceinline__ void load_pair(const double2 *p, double2 &a, double2 &b)
{
    a = p[0];
    b = p[1];
}
__device__ __forceinline__ void store_pair(float2 *p, float2 a, float2 b)
{
    *reinterpret_cast<float4 *>(p) = make_float4(a.x, a.y, b.x, b.y);
}
__device__ __forceinline__ void store_pair(double2 *p, double2 a, double2 b)
{
    p[0] = a;
    p[1] = b;
}

// 3D fold by column segments: a thread of a tile owns the SEGC = kFoldSegBytes /
// (cell bytes) core cells (r0, r1, seg SEGC .. seg SEGC + SEGC-1) and sums them
// over all candidate slot blocks.  In a slot block of extents (E0, E1, EL)
// those cells are consecutive, so a thread reads kFoldSegBytes contiguous
// bytes with 16-byte loads and consecutive threads read consecutive pieces of
// the block.  32-byte pieces (4 FP32 cells) measured 2% faster than 64-byte
// ones on the 3D 256^3 fold (more threads per tile, fewer registers each).
// A block that does not cover the thread's column (E_t = 1 on a halo block:
// only r_t = 0 is covered) is read at r_t = 0 and not added.  Candidates go
// by (q0, q1): the first chunks of the q_L blocks in flight together, then the
// further chunks of dense tiles.  The order of the sums is fixed, so the grid
// is reproducible for a given binning.
// The column-segment fold serves FP32 grids (3D 256^3: 0.51 -> 0.44 ms); for
// FP64 the per-cell fold (tile_fold), whose loads are already 16 bytes per lane and
// contiguous over the wavefront, measured faster (0.74 vs 0.78 ms).
template <typename T, int D> constexpr bool kFoldColumns = sizeof(T) == 4;

template <typename L> struct Fold3Shape {
    static constexpr int SEGC = kFoldSegBytes / L::CB;      // cells per thread
    static constexpr int S = L::TL / SEGC;                  // segments per column
    static constexpr int TPT = L::T0 * L::T1 * S;           // threads per tile
    static constexpr int TPW = TPT <= 256 ? 256 / TPT : 1;  // tiles per workgroup
    static constexpr int WPT = TPT > 256 ? TPT / 256 : 1;   // workgroups per tile
    static_assert(L::TL % SEGC == 0 && (256 % TPT == 0 || TPT % 256 == 0),
                  "column segments tile the core and whole workgroups");
};

template <typename T, int D, int W, int PS, int V, bool COMP>
__device__ __forceinline__ void tile_fold_cols(const cplx_t<T> *__restrict__ slots,
                                               const uint32_t *__restrict__ tile_count,
                                               const uint32_t *__restrict__ chunk_start,
                                               cplx_t<T> *__restrict__ g, const Geom &ge,
                                               const BinGeom &bg, uint32_t blk)
{
    using L = TileLayout<T, D, W, PS, V>;
    using F = Fold3Shape<L>;
    using C = cplx_t<T>;
    constexpr int TL = L::TL, SEGC = F::SEGC, S = F::S;
    constexpr int ELAST = L::PL - (L::QL - 1) * TL;          // extent of the last q_L block
    constexpr int NLAST = ELAST < SEGC ? ELAST : SEGC;       // cells a thread loads there
    static_assert(L::T0 % 2 == 0 && (D == 2 || L::T1 % 2 == 0) && TL % 2 == 0,
                  "even tile extents keep slot rows and grid rows 16-byte aligned");
    const int tid = threadIdx.x;
    // a tile's threads fill whole wavefronts (TPT >= 64), so the tile and every
    // candidate-table entry below are wave-uniform: scalar loads, all issued
    // before the first slot load
    static_assert(F::TPT >= 64, "a wavefront never spans two tiles");
    const uint32_t tile = __builtin_amdgcn_readfirstlane(
        F::TPT <= 256 ? blk * F::TPW + (uint32_t)(tid / F::TPT) : blk / F::WPT);
    if (tile >= (uint32_t)bg.ntiles) return;   // no barriers follow
    const int tt = F::TPT <= 256 ? tid % F::TPT : (int)(blk % F::WPT) * 256 + tid;
    const int col = tt / S, seg = tt % S;
    const int r0 = col / L::T1, r1 = col % L::T1;
    // tile coordinates on axes 0, 1 (d = 3; 0 and 0 for d = 2) and L
    uint32_t tk[3] = {0, 0, 0};
    {
        uint32_t rem = tile;
        tk[2] = rem % (uint32_t)bg.nt[D - 1];
        rem /= (uint32_t)bg.nt[D - 1];
        if (D == 3) {
            tk[1] = rem % (uint32_t)bg.nt[1];
            rem /= (uint32_t)bg.nt[1];
        }
        tk[0] = rem;
    }
    const uint32_t n0 = (uint32_t)bg.nt[0], n1 = D == 3 ? (uint32_t)bg.nt[1] : 1u,
                   n2 = (uint32_t)bg.nt[D - 1];
    constexpr int Q0 = L::Q0, Q1 = L::Q1, QL = L::QL;
    // first chunk and chunk count of the candidate tile q back along each axis
    uint32_t c0[Q0][Q1][QL], nch[Q0][Q1][QL];
#pragma unroll
    for (int q0 = 0; q0 < Q0; ++q0)
#pragma unroll
        for (int q1 = 0; q1 < Q1; ++q1)
#pragma unroll
            for (int qL = 0; qL < QL; ++qL) {
                const uint32_t src = ((tk[0] + 4 * n0 - q0) % n0 * n1 + (tk[1] + 4 * n1 - q1) % n1) * n2 +
                                     (tk[2] + 4 * n2 - qL) % n2;
                c0[q0][q1][qL] = chunk_start[src];
                nch[q0][q1][qL] = (tile_count[src] + kChunk - 1) / kChunk;
            }
    T sx[SEGC], sy[SEGC], cx[SEGC], cy[SEGC];
#pragma unroll
    for (int i = 0; i < SEGC; ++i) sx[i] = sy[i] = cx[i] = cy[i] = 0;
    // the thread's cells of block (q0, q1, qL) of slot sp; full blocks: SEGC
    // aligned cells, the last q_L block: NLAST cells (index clamped in the block)
    auto load_full = [&](const C *sp, int o, C *dst) {
#pragma unroll
        for (int i = 0; i < SEGC / 2; ++i) load_pair(sp + o + 2 * i, dst[2 * i], dst[2 * i + 1]);
    };
    auto load_last = [&](const C *sp, int o, C *dst) {
#pragma unroll
        for (int i = 0; i < NLAST; ++i) {
            const int rl = seg * SEGC + i;
            dst[i] = sp[o + (rl < ELAST ? rl : ELAST - 1)];
        }
    };
#pragma unroll
    for (int q0 = 0; q0 < Q0; ++q0)
#pragma unroll
        for (int q1 = 0; q1 < Q1; ++q1) {
            const int E0 = L::ext(L::T0, L::P0, q0), E1 = L::ext(L::T1, L::P1, q1);
            const bool in01 = r0 < E0 && r1 < E1;
            // a halo block pair that covers none of the wavefront's columns
            // (E0 = 1 past the first rows of the tile) is not loaded at all
            if (!__any(in01)) continue;
            const int rc = (r0 < E0 ? r0 : 0) * E1 + (r1 < E1 ? r1 : 0);
            // the first chunk of the QL candidates of this (q0, q1) in flight together
            C v[QL][SEGC];
#pragma unroll
            for (int qL = 0; qL < QL; ++qL) {
                // a candidate with no nodes reads chunk 0 and adds nothing
                const C *sp = slots + (int64_t)(nch[q0][q1][qL] > 0 ? c0[q0][q1][qL] : 0) * L::SC;
                const int EL = L::ext(TL, L::PL, qL);
                const int o = L::block_off(q0, q1, qL) + rc * EL;
                if (EL == TL) load_full(sp, o + seg * SEGC, v[qL]);
                else load_last(sp, o, v[qL]);
            }
#pragma unroll
            for (int qL = 0; qL < QL; ++qL) {
                const int EL = L::ext(TL, L::PL, qL);
                const bool ok = nch[q0][q1][qL] > 0 && in01;
#pragma unroll
                for (int i = 0; i < SEGC; ++i)
                    if ((EL == TL || (i < NLAST && seg * SEGC + i < EL)) && ok) {
                        fold_add<COMP>(sx[i], cx[i], v[qL][i].x);
                        fold_add<COMP>(sy[i], cy[i], v[qL][i].y);
                    }
            }
            // further chunks of dense candidate tiles
#pragma unroll
            for (int qL = 0; qL < QL; ++qL) {
                const int EL = L::ext(TL, L::PL, qL);
                const int o = L::block_off(q0, q1, qL) + rc * EL;
                for (uint32_t c = 1; c < nch[q0][q1][qL]; ++c) {
                    const C *sp = slots + (int64_t)(c0[q0][q1][qL] + c) * L::SC;
                    C w[SEGC];
                    if (EL == TL) load_full(sp, o + seg * SEGC, w);
                    else load_last(sp, o, w);
#pragma unroll
                    for (int i = 0; i < SEGC; ++i)
                        if ((EL == TL || (i < NLAST && seg * SEGC + i < EL)) && in01) {
                            fold_add<COMP>(sx[i], cx[i], w[i].x);
                            fold_add<COMP>(sy[i], cy[i], w[i].y);
                        }
                }
            }
        }
    const uint32_t g1 = D == 3 ? (uint32_t)ge.n[1] : 1u, gL = (uint32_t)ge.n[D - 1];
    C *dst = g + (size_t)((tk[0] * L::T0 + r0) * g1 + tk[1] * L::T1 + r1) * gL + tk[2] * TL + seg * SEGC;
#pragma unroll
    for (int i = 0; i < SEGC / 2; ++i)
        store_pair(dst + 2 * i, cmake<T>(sx[2 * i] + cx[2 * i], sy[2 * i] + cy[2 * i]),
                   cmake<T>(sx[2 * i + 1] + cx[2 * i + 1], sy[2 * i + 1] + cy[2 * i + 1]));
}

template <typename T, int D, int W, int PS, int V, bool COMP>
__global__ __launch_bounds__(256) void k_tile_fold(const cplx_t<T> *__restrict__ slots,
                                                   const uint32_t *__restrict__ tile_count,
                                                   const uint32_t *__restrict__ chunk_start,
                                                   cplx_t<T> *__restrict__ g, Geom ge, BinGeom bg)
{
    if constexpr (kFoldColumns<T, D>)
        tile_fold_cols<T, D, W, PS, V, COMP>(slots, tile_count, chunk_start, g, ge, bg, blockIdx.x);
    else
        tile_fold<T, D, W, PS, V, COMP>(slots, tile_count, chunk_start, g, ge, bg, blockIdx.x,
                                        blockIdx.y);
}

// Batched fold: block blk of a job is tile blk / slices, slice blk % slices.
template <typename T, int D, int W, int PS, int V>
__global__ __launch_bounds__(256) void k_tile_fold_batch(const GridJob<T> *__restrict__ jobs,
                                                         const BlockMap *__restrict__ map)
{
    using L = TileLayout<T, D, W, PS, V>;
    constexpr int CORE = L::T0 * L::T1 * L::TL;
    constexpr int SLICES = CORE >= 512 ? CORE / 512 : 1;
    const BlockMap bm = map[blockIdx.x];
    const GridJob<T> &J = jobs[bm.job];
    if constexpr (kFoldColumns<T, D>)
        tile_fold_cols<T, D, W, PS, V, false>(J.slots, J.tile_count, J.chunk_start, J.g, J.ge, J.bg, bm.blk);
    else
        tile_fold<T, D, W, PS, V, false>(J.slots, J.tile_count, J.chunk_start, J.g, J.ge, J.bg,
                                         bm.blk / SLICES, (int)(bm.blk % SLICES));
}

}  // namespace cunfft

#include "grid_mfma.hpp"

namespace cunfft {

// ------------------------------------------------------------ dispatch --
// Wavefronts per workgroup.  A tile small enough for two or more workgroups
// per CU gets 4 forward wavefronts (node batches); a tile that fills the CU's
// LDS alone gets 8, so the CU still holds enough wavefronts to cover the LDS
// latency.  The 3D FP32 gather also runs 8: three of its 52 KB workgroups
// fill a CU's LDS, and 24 wavefronts per CU instead of 12 hide more of each
// node's LDS round trip (8^3 tiles at 256^3: 0.549 -> 0.488 ms; 16 wavefronts
// are capped at the same 24 per CU by registers).  The 3D adjoint runs 4
// wavefronts (each owning every fourth plane): with the units held across
// nodes and the shared weight batches it is bound by instruction issue at 3
// workgroups per CU, and the extra wavefronts beat the two dummy planes they
// add (FP32 1.33 -> 1.30 ms, FP64 3.04 -> 2.81 ms at 256^3 against 2
// wavefronts).
template <typename T, int D, int W, int V> struct Waves {
    static constexpr int tile_bytes = TileLayout<T, D, W, 1, V>::CELLS * TileLayout<T, D, W, 1, V>::CB;
    static constexpr bool alone = 2 * (tile_bytes + 1024) > 160 * 1024;
    static constexpr int gather = alone || (D == 3 && sizeof(T) == 4) ? 8 : 4;
    // 2D: a node's window is W rows of W cells, which one wavefront covers
    // with most lanes busy; splitting its rows over wavefronts would leave
    // most lanes idle, so the adjoint runs one wavefront on a small tile
    // (enough of them fit a CU's LDS to hide the latency), else 2
    static constexpr int spread = D == 2 && tile_bytes <= 16384 ? 1 : (alone || D == 3 ? 4 : 2);
};

template <typename T> WinConst<T> win_const(const Geom &ge)
{
    WinConst<T> w{};
    const double sc = sizeof(T) == 4 ? 1.4426950408889634 : 1.0;   // log2(e) for exp2
    for (int t = 0; t < 3; ++t) {
        w.kb[t] = (T)(ge.invb[t] * sc);
        w.nrm[t] = (T)ge.nrm[t];
        w.shp[t] = (T)ge.shp[t];
        w.dk[t] = (T)(-2.0 * (double)ge.n[t] * ge.invb[t]);
        w.nn[t] = (T)ge.n[t];
    }
    w.m = ge.m;
    w.win = ge.win;
    return w;
}

template <typename T, int D, int W, int V> constexpr bool fits_all()
{
    return TileLayout<T, D, W, 1, V>::fits &&
           TileLayout<T, D, W, Waves<T, D, W, V>::spread, V>::fits;
}

// The compensated adjoint (FP32 only) holds a second tile of Kahan terms.
template <typename T, int D, int W, int V> constexpr bool fits_comp()
{
    using LS = TileLayout<T, D, W, Waves<T, D, W, V>::spread, V>;
    return sizeof(T) == 4 && fits_all<T, D, W, V>() &&
           2L * (LS::CELLS * LS::CB + LS::DUMMY_UNITS * 16) <= kLdsMax;
}

// Runs f.template run<D, W, V>() for the runtime window width, dimension and
// tile variant; false when there is no such kernel.
template <int D, int V, typename F> static bool with_w(int W, F &f)
{
    switch (W) {
    case 4: return f.template run<D, 4, V>();
    case 6: return f.template run<D, 6, V>();
    case 8: return f.template run<D, 8, V>();
    case 10: return f.template run<D, 10, V>();
    case 12: return f.template run<D, 12, V>();
    case 14: return f.template run<D, 14, V>();
    case 16: return f.template run<D, 16, V>();
    case 18: return f.template run<D, 18, V>();
    }
    return false;
}

template <typename F> static bool with_dwv(int d, int W, int v, F &f)
{
    if (v < 0 || v >= kTileVariants) return false;
    if (d == 2) {
        if (v == 0) return with_w<2, 0>(W, f);
        if (v == 1) return with_w<2, 1>(W, f);
        return with_w<2, 2>(W, f);
    }
    if (d == 3) {
        if (v == 0) return with_w<3, 0>(W, f);
        if (v == 1) return with_w<3, 1>(W, f);
        return with_w<3, 2>(W, f);
    }
    return false;
}

template <typename T> struct FitOp {
    bool comp = false;
    int64_t slot_cells = 0;
    template <int D, int W, int V> bool run()
    {
        if constexpr (fits_all<T, D, W, V>()) {
            slot_cells = TileLayout<T, D, W, Waves<T, D, W, V>::spread, V>::SC;
            return !comp || fits_comp<T, D, W, V>();
        } else {
            return false;
        }
    }
};

template <typename T> bool tile_kernels_fit(const Geom &ge, int variant, bool comp)
{
    FitOp<T> op;
    op.comp = comp;
    return with_dwv(ge.d, ge.W, variant, op);
}

template <typename T> int64_t tile_slot_cells(const Geom &ge, int variant)
{
    FitOp<T> op;
    return with_dwv(ge.d, ge.W, variant, op) ? op.slot_cells : 0;
}

template <typename T> struct GatherOp {
    const Geom &ge;
    const Binning &b;
    const cplx_t<T> *g;
    cplx_t<T> *f;
    hipStream_t s;
    template <int D, int W, int V> bool run()
    {
        if constexpr (fits_all<T, D, W, V>()) {
            constexpr int NWG = Waves<T, D, W, V>::gather;
            k_tile_gather<T, D, W, NWG, V><<<xcd_grid(b.max_chunks), NWG * 64, 0, s>>>(
                    g, static_cast<const NodeDesc<T> *>(b.desc), b.perm, b.chunks, b.n_chunks, f,
                    ge, bin_geom(b), win_const<T>(ge));
            return true;
        } else {
            return false;
        }
    }
};

template <typename T> struct GradOp {
    const Geom &ge;
    const Binning &b;
    const cplx_t<T> *g;
    const cplx_t<T> *v;
    T *gx;
    hipStream_t s;
    template <int D, int W, int V> bool run()
    {
        if constexpr (fits_all<T, D, W, V>() && GradShape<T, D, W, V>::fits) {
            constexpr int NWG = GradShape<T, D, W, V>::NW;
            k_tile_grad<T, D, W, NWG, V><<<xcd_grid(b.max_chunks), NWG * 64, 0, s>>>(
                g, static_cast<const NodeDesc<T> *>(b.desc), b.perm, b.chunks, b.n_chunks, v, gx, ge,
                bin_geom(b), win_const<T>(ge));
            return true;
        } else {
            return false;
        }
    }
};

// (static: each precision's translation unit has its own copy)
static __global__ void k_clear_u32(uint32_t *p)
{
    if (threadIdx.x == 0) p[0] = 0u;
}

// Workgroups of kernel k (threads per workgroup, no dynamic LDS) that the
// current device holds at once: its CUs times the occupancy per CU (asked
// once per device and kernel; kernels of one signature differ in occupancy,
// so the cache is keyed by the kernel itself).
template <typename K> static int64_t resident_blocks(K k, int threads)
{
    static std::mutex mu;
    static std::map<std::pair<int, const void *>, int64_t> known;
    int dev = 0;
    CUNFFT_HIP(hipGetDevice(&dev));
    const std::pair<int, const void *> key{dev, reinterpret_cast<const void *>(k)};
    {
        std::lock_guard<std::mutex> lk(mu);
        const auto it = known.find(key);
        if (it != known.end()) return it->second;
    }
    int cus = 0, per = 0;
    CUNFFT_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    CUNFFT_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k, threads, 0));
    const int64_t r = (int64_t)std::max(1, cus) * std::max(1, per);
    std::lock_guard<std::mutex> lk(mu);
    known[key] = r;
    return r;
}

template <typename T> struct SpreadOp {
    const Geom &ge;
    const Binning &b;
    const cplx_t<T> *f;
    cplx_t<T> *g;
    hipStream_t s;
    hipEvent_t mid;
    template <int D, int W, int V> bool run()
    {
        if constexpr (fits_all<T, D, W, V>()) {
            constexpr int NWS = Waves<T, D, W, V>::spread;
            cplx_t<T> *slots = static_cast<cplx_t<T> *>(b.slots);
            cplx_t<T> *fs = static_cast<cplx_t<T> *>(b.fs);
            using L = TileLayout<T, D, W, NWS, V>;
            // the FP32 3D VALU spreads read f through the permutation themselves
            constexpr bool MFMA = D == 3 && V == 0 && kMfmaSpread<T, W>;
            constexpr bool MFMA2 = sizeof(T) == 4 && D == 2 && V == 1 && kMfmaSpread2<W>;
            constexpr bool QUEUE = D == 3 && L::G == 1;
            constexpr bool PERM_F = QUEUE && !MFMA && sizeof(T) == 4;
            if (!PERM_F || b.compensated)
                k_sorted_coeffs<T><<<(unsigned)((ge.M + 255) / 256), 256, 0, s>>>(f, b.perm, fs, ge.M);
            constexpr int CORE = L::T0 * L::T1 * L::TL;
            constexpr int SLICES = CORE >= 512 ? CORE / 512 : 1;
            const dim3 fold_grid = kFoldColumns<T, D>
                                       ? dim3((unsigned)tile_fold_blocks(D, V, b.ntiles, sizeof(T)), 1)
                                       : dim3((unsigned)b.ntiles, SLICES);
            if constexpr (fits_comp<T, D, W, V>()) {
                if (b.compensated) {
                    k_tile_spread<T, D, W, NWS, V, true><<<(unsigned)b.max_chunks, NWS * 64, 0, s>>>(
                        static_cast<const NodeDesc<T> *>(b.desc), fs, b.chunks, b.n_chunks, slots,
                        win_const<T>(ge));
                    if (mid) CUNFFT_HIP(hipEventRecord(mid, s));
                    k_tile_fold<T, D, W, NWS, V, true><<<fold_grid, 256, 0, s>>>(
                        slots, b.tile_count, b.chunk_start, g, ge, bin_geom(b));
                    return true;
                }
            }
            if constexpr (MFMA) {
                // persistent: clear the work counter (a kernel, not a memset), then one
                // resident grid
                auto km = ge.win == WIN_KAISER_BESSEL ? k_tile_spread_mfma<T, W, WIN_KAISER_BESSEL>
                          : ge.win == WIN_SINC        ? k_tile_spread_mfma<T, W, WIN_SINC>
                                                      : k_tile_spread_mfma<T, W, WIN_GAUSSIAN>;
                constexpr int NT = MfmaSpread<T, W>::NW * 64;
                k_clear_u32<<<1, 64, 0, s>>>(b.n_chunks + 1);
                const unsigned grid = (unsigned)std::min<int64_t>(b.max_chunks, resident_blocks(km, NT));
                km<<<grid, NT, 0, s>>>(static_cast<const NodeDesc<T> *>(b.desc), fs, b.chunks, b.n_chunks, slots,
                                       win_const<T>(ge));
            } else if constexpr (MFMA2) {
                k_tile_spread_mfma2<W><<<(unsigned)b.max_chunks, 64, 0, s>>>(
                    reinterpret_cast<const NodeDesc<float> *>(b.desc), reinterpret_cast<const cplx_t<float> *>(fs),
                    b.chunks, b.n_chunks, reinterpret_cast<cplx_t<float> *>(slots), win_const<float>(ge));
            } else if constexpr (QUEUE) {
                // persistent held-unit spread: clear the work counter (a kernel, not a
                // memset, so the launch sequence can be captured in a graph), then one
                // resident grid
                auto kq = k_tile_spread_queue<T, D, W, NWS, V>;
                k_clear_u32<<<1, 64, 0, s>>>(b.n_chunks + 1);
                const unsigned grid = (unsigned)std::min<int64_t>(b.max_chunks, resident_blocks(kq, NWS * 64));
                kq<<<grid, NWS * 64, 0, s>>>(static_cast<const NodeDesc<T> *>(b.desc),
                                             PERM_F ? b.perm : nullptr, f, fs, b.chunks, b.n_chunks,
                                             slots, win_const<T>(ge));
            } else {
                k_tile_spread<T, D, W, NWS, V, false><<<(unsigned)b.max_chunks, NWS * 64, 0, s>>>(
                    static_cast<const NodeDesc<T> *>(b.desc), fs, b.chunks, b.n_chunks, slots,
                    win_const<T>(ge));
            }
            if (mid) CUNFFT_HIP(hipEventRecord(mid, s));
            k_tile_fold<T, D, W, NWS, V, false><<<fold_grid, 256, 0, s>>>(
                slots, b.tile_count, b.chunk_start, g, ge, bin_geom(b));
            return true;
        } else {
            return false;
        }
    }
};

template <typename T> struct BatchOp {
    BatchStage st;
    const GridJob<T> *jobs;
    const BlockMap *map;
    uint32_t nblk;
    hipStream_t s;
    template <int D, int W, int V> bool run()
    {
        if constexpr (fits_all<T, D, W, V>()) {
            constexpr int NWG = Waves<T, D, W, V>::gather;
            constexpr int NWS = Waves<T, D, W, V>::spread;
            switch (st) {
            case BS_GATHER:
                k_tile_gather_batch<T, D, W, NWG, V><<<nblk, NWG * 64, 0, s>>>(jobs, map);
                return true;
            case BS_COEFFS:
                k_sorted_coeffs_batch<T><<<nblk, 256, 0, s>>>(jobs, map);
                return true;
            case BS_SPREAD:
                if constexpr (sizeof(T) == 4 && D == 2 && V == 1 && kMfmaSpread2<W>)
                    k_tile_spread_mfma2_batch<W>
                        <<<nblk, 64, 0, s>>>(reinterpret_cast<const GridJob<float> *>(jobs), map);
                else
                    k_tile_spread_batch<T, D, W, NWS, V><<<nblk, NWS * 64, 0, s>>>(jobs, map);
                return true;
            case BS_FOLD:
                k_tile_fold_batch<T, D, W, NWS, V><<<nblk, 256, 0, s>>>(jobs, map);
                return true;
            default:
                return false;
            }
        } else {
            return false;
        }
    }
};

template <typename T>
void launch_tile_batch(BatchStage st, const GridJob<T> *jobs, const BlockMap *map, uint32_t nblk,
                       int d, int W, int variant, hipStream_t s)
{
    if (nblk == 0) return;
    BatchOp<T> op{st, jobs, map, nblk, s};
    if (!with_dwv(d, W, variant, op))
        throw Error(CUNFFT_ERR_UNSUPPORTED, "no batched plane-tiled kernel for this stage");
    CUNFFT_HIP(hipGetLastError());
}

template <typename T>
void launch_tile_gather(const Geom &ge, const Binning &b, const cplx_t<T> *g, cplx_t<T> *f,
                        hipStream_t s)
{
    GatherOp<T> op{ge, b, g, f, s};
    if (!with_dwv(ge.d, ge.W, b.tile_variant, op))
        throw Error(CUNFFT_ERR_UNSUPPORTED, "no tiled gather for this geometry");
    CUNFFT_HIP(hipGetLastError());
}

template <typename T>
bool launch_tile_grad(const Geom &ge, const Binning &b, const cplx_t<T> *g, const cplx_t<T> *v, T *gx,
                      hipStream_t s)
{
    GradOp<T> op{ge, b, g, v, gx, s};
    if (!with_dwv(ge.d, ge.W, b.tile_variant, op)) return false;
    CUNFFT_HIP(hipGetLastError());
    return true;
}

template <typename T>
void launch_tile_spread(const Geom &ge, const Binning &b, const cplx_t<T> *f, cplx_t<T> *g,
                        hipStream_t s, hipEvent_t mid)
{
    SpreadOp<T> op{ge, b, f, g, s, mid};
    if (!with_dwv(ge.d, ge.W, b.tile_variant, op))
        throw Error(CUNFFT_ERR_UNSUPPORTED, "no tiled spread for this geometry");
    CUNFFT_HIP(hipGetLastError());
}

}  // namespace cunfft
