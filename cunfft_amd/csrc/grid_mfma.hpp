// Adjoint tile accumulation on the matrix cores (FP32 and FP64, d = 3, 8^3
// tiles).
//
// Reference behaviour: the adjoint gridding of the reference's spreading
// kernels (src/cunfft_kernel.cu:1098-1197, the adjoint convolution over the
// 2m+2 window points per axis) sums f_j phi(x_j - cell) into every cell of
// each node's window.  Here the sum over the nodes of one chunk is a matrix
// product.  With the padded tile's cells as (plane a, row b, column c) and the
// window weights separable,
//
//   tile[a][b][c] = sum_j  wa_j[a] * (f_j wb_j[b] wc_j[c])
//                 = sum_j  A[a][j] * B[j][(b, c, re/im)],
//
// A (planes x nodes) holding the plane weights and B (nodes x columns) the
// products of coefficient, row and column weight, real and imaginary part in
// adjacent columns.  v_mfma_f32_16x16x4f32 (v_mfma_f64_16x16x4f64) multiplies
// 16 planes by 4 nodes by 16 columns, accumulating in the plan's precision
// (products in that precision, as the VALU form).  FP64 accumulators take
// twice the registers, so an FP64 workgroup has 8 wavefronts sharing a tile's
// rows where an FP32 one has 4.
//
// * Planes.  The padded tile has P = 8 + W - 1 planes (17 at W = 10) but one
//   instruction covers 16.  A node's window spans W planes from its start
//   lo_0 in [0, 8), so the nodes are split into two sets: lo_0 <= 16 - W
//   (window inside planes [0, 16)) and the rest (inside [P - 16, P)).  Each
//   set has its own accumulators whose rows are its 16 planes, and each node
//   goes through the matrix cores once.  The sets are summed at the end.
// * Columns.  Each wavefront owns a band of rows b.  Per row it holds two
//   blocks of 16 columns (columns c in [0, 8) and [8, 16), re/im), and the
//   columns c >= 16 of all rows (P - 16 of them) are packed into extra
//   blocks spread over the wavefronts.
// * Per batch of 32 nodes the weights go to LDS once, in padded-tile
//   coordinates (zero outside the window), one table row per node: plane
//   weights per set row, row weights times the coefficient (re and im), column
//   weights.  The nodes of a set are padded to a multiple of 4 with zero
//   rows.  Descriptors and coefficients (the sorted copy, so no load waits
//   on another) load one batch ahead.  Wavefront t < 3 evaluates axis t for
//   the batch's nodes (lanes n and n + 32: node n, each half of the row's
//   cells), wavefront 3 clears the padding rows.
// * The accumulators go to an LDS image of the chunk's slot (the fold's
//   destination-blocked order), which the workgroup then copies out with
//   16-byte stores: whole cache lines, where storing the accumulators
//   directly (4-byte lanes, 64-byte runs) took more than the matrix work.
//
// The sums per cell run over the nodes of the chunk in a fixed order (sorted
// order within a set, in the instruction's fixed reduction order), so the
// result does not depend on scheduling.
#pragma once

namespace cunfft {

typedef float mf4_t __attribute__((ext_vector_type(4)));
typedef double md4_t __attribute__((ext_vector_type(4)));

// One 16x16x4 product per precision.  Both take A[m = lane % 16][k = lane / 16]
// and B[k = lane / 16][n = lane % 16]; accumulator register i of lane l holds
// row 4 (l / 16) + i (FP32) or l / 16 + 4 i (FP64) of column l % 16 (both
// layouts checked on the GPU with an in-bounds probe).
template <typename R> struct Mfma;
template <> struct Mfma<float> {
    using acc = mf4_t;
    __device__ static acc op(float a, float b, const acc &c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __host__ __device__ static constexpr int row(int g, int i) { return 4 * g + i; }
};
template <> struct Mfma<double> {
    using acc = md4_t;
    __device__ static acc op(double a, double b, const acc &c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __host__ __device__ static constexpr int row(int g, int i) { return g + 4 * i; }
};

template <typename R, int W> struct MfmaSpread {
    static constexpr int T = 8;                         // tile extent per axis
    static constexpr int P = T + W - 1;                 // padded extent per axis
    static constexpr int OFF1 = P > 16 ? P - 16 : 0;    // first plane of the second set
    static constexpr bool SETS2 = OFF1 > 0;
    static constexpr int LO_MAX0 = 16 - W;              // set 0: window starts lo_0 <= LO_MAX0
    // wavefronts per workgroup: FP64 accumulators take twice the registers, so
    // twice the wavefronts share a tile's rows
    static constexpr int NW = sizeof(R) == 8 ? 8 : 4;
    static constexpr int RW = (P + NW - 1) / NW;        // rows per wavefront (at most)
    static constexpr int NLO = P > 16 ? P - 16 : 0;     // columns past the two 8-column halves
    static constexpr int NLOB = (P * NLO * 2 + 15) / 16;   // their 16-column blocks
    static constexpr int LB = (NLOB + NW - 1) / NW;     // of those, per wavefront (at most)
    static constexpr int NB = 32;                       // nodes per batch
    static constexpr int NPOS = NB + 8;                 // batch rows: each set padded to 4 nodes
    static constexpr int RS = 20;                       // row / column weight table stride
    // one table row per batch position: plane weights [0, 16), row weights times the
    // coefficient (re, im) from O_B1, column weights from O_B2; 80 floats, so the
    // four rows one step reads start 16 banks apart
    static constexpr int O_B1 = 16, O_B2 = O_B1 + 2 * RS, ROW = 80;
    static_assert(O_B2 + RS <= ROW, "a table row holds its four parts");
    // wavefront w: rows [row0(w), row0(w) + rows(w)) and extra(w) extra blocks
    static constexpr int row0(int w) { return w * P / NW; }
    static constexpr int rows(int w) { return (w + 1) * P / NW - w * P / NW; }
    static constexpr int extra(int w)
    {
        int k = 0;
        for (int j = 0; j < LB; ++j) k += w + NW * j < NLOB ? 1 : 0;
        return k;
    }
    // (FP32) every wavefront's first row weight is 16-byte aligned in its table row
    static constexpr bool rows_aligned()
    {
        bool ok = sizeof(R) == 4 && (O_B1 % 4 == 0) && (RS % 4 == 0) && (ROW % 4 == 0);
        for (int w = 0; w < NW; ++w) ok = ok && row0(w) % 4 == 0;
        return ok;
    }
    // some wavefront has nr rows and nl extra blocks (at most two such shapes)
    static constexpr bool shape(int nr, int nl)
    {
        bool any = false;
        for (int w = 0; w < NW; ++w) any = any || (rows(w) == nr && extra(w) == nl);
        return any;
    }
    static_assert(W >= 2 && 2 * W + T <= 32, "every window start fits one of the two plane sets");
    static_assert(P <= RS, "a padded row fits the weight table stride");
};

template <typename R, int W> constexpr bool kMfmaSpread = W >= 4 && W <= 12;
// threads and resident workgroups per CU the kernel is built for
template <typename R> constexpr int kMfmaThreads = sizeof(R) == 8 ? 512 : 256;
template <typename R> constexpr int kMfmaPerCu = sizeof(R) == 8 ? 1 : 3;

// Element offset (units of R) of part ri of cell (a, b, c) in a slot of the
// 8^3 padded tile (the destination-blocked order of TileLayout::block_off).
template <typename L> __device__ __forceinline__ int mfma_slot_off(int a, int b, int c, int ri)
{
    const int q1 = b >> 3, qL = c >> 3;
    const int E1 = L::ext(L::T1, L::P1, q1), EL = L::ext(L::TL, L::PL, qL);
    return 2 * (L::block_off(a >> 3, q1, qL) + ((a & 7) * E1 + (b & 7)) * EL + (c & 7)) + ri;
}

// Float offsets in a slot of the 8^3 padded tile (the destination-blocked
// order of TileLayout::block_off) of part ri of the cells (a, b, c) a lane
// group g stores: planes a = 4g + i (i < 4, a < 16) at base + i * step, and
// the planes a = 16 + i past them at base16 + i * step.
struct MfmaSlotCol {
    int base, base16, step;
};
template <typename L> __device__ __forceinline__ MfmaSlotCol mfma_slot_col(int g, int b, int c, int ri)
{
    const int q1 = b >> 3, qL = c >> 3;
    const int E1 = L::ext(L::T1, L::P1, q1), EL = L::ext(L::TL, L::PL, qL);
    const int in = (b & 7) * EL + (c & 7);   // row and column within the block
    MfmaSlotCol o;
    o.base = 2 * (L::block_off(g >> 1, q1, qL) + 4 * (g & 1) * E1 * EL + in) + ri;
    o.base16 = 2 * (L::block_off(2, q1, qL) + in) + ri;
    o.step = 2 * E1 * EL;
    return o;
}

// Persistent: a grid of resident workgroups takes chunks from the work
// counter n_chunks[1] (zero at launch) until it runs past the chunk count
// n_chunks[0].  A workgroup holds the tickets of its current and its next
// chunk and takes a third while the current one runs, so the next chunk's
// record is in registers when its first batch is loaded; it exits at its
// first ticket past the last chunk (its tickets rise, so every chunk below
// the count is run).  The next batch's nodes (or the next chunk's first
// batch) load while the current one runs through the matrix cores, and the
// operands of the next 4-node step load while the current step's
// instructions issue.
// WIN: the window (inline, no calls: the accumulators stay in registers
// through the weight evaluation).
template <typename R, int W, int WIN>
__global__ __launch_bounds__(kMfmaThreads<R>, kMfmaPerCu<R>) void k_tile_spread_mfma(
    const NodeDesc<R> *__restrict__ desc, const cplx_t<R> *__restrict__ fs, const Chunk *__restrict__ chunks,
    uint32_t *__restrict__ n_chunks, cplx_t<R> *__restrict__ slots, WinConst<R> wc)
{
    using S = MfmaSpread<R, W>;
    using M = Mfma<R>;
    using acc_t = typename M::acc;
    using V16 = vec16_t<R>;                 // 16 bytes of R
    constexpr int E16 = 16 / (int)sizeof(R);
    constexpr int NT = S::NW * 64;
    static_assert(NT == kMfmaThreads<R>, "one thread per lane of the workgroup's wavefronts");
    using L = TileLayout<R, 3, W, 4, 0>;    // slot layout (shared with the fold)
    static_assert(L::T0 == S::T && L::T1 == S::T && L::TL == S::T && L::P0 == S::P, "8^3 tiles");
    constexpr int LBN = S::LB > 0 ? S::LB : 1;
    __shared__ __align__(16) R tw[S::NPOS][S::ROW];
    __shared__ __align__(16) R stg[2 * L::SC];   // the slot image
    __shared__ uint32_t s_tk[2];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    // operand lanes: A[m = lane % 16][k = lane / 16], B[k = lane / 16][n = lane % 16],
    // accumulator register i: D[M::row(lane / 16, i)][lane % 16]
    const int g = lane >> 4, nn = lane & 15, ri = nn & 1, cc = nn >> 1;
    const int rb = wave * S::P / S::NW;
    const int nrows = (wave + 1) * S::P / S::NW - rb;
    // the lane's column of each extra block: row lb_b, column lb_c (>= 16)
    int lb_b[LBN], lb_c[LBN];
    bool lb_ok[LBN];
#pragma unroll
    for (int j = 0; j < LBN; ++j) {
        const int blk = wave + S::NW * j;
        const int e = (16 * blk + nn) >> 1;
        const int b = e / (S::NLO > 0 ? S::NLO : 1);
        lb_ok[j] = S::LB > 0 && blk < S::NLOB && b < S::P;
        lb_b[j] = lb_ok[j] ? b : 0;
        lb_c[j] = lb_ok[j] ? 16 + e % (S::NLO > 0 ? S::NLO : 1) : 0;
    }
    uint32_t *counter = n_chunks + 1;
    const uint32_t n = n_chunks[0];
    if (tid == 0) {
        s_tk[0] = atomicAdd(counter, 1u);
        s_tk[1] = atomicAdd(counter, 1u);
    }
    __syncthreads();
    uint32_t c = s_tk[0], cn = s_tk[1];
    // lanes n and n + 32: node n of the next batch to run (the two halves of a
    // wavefront share the weight evaluation of a node)
    static_assert(S::NB == 32, "a batch fills half a wavefront");
    const int nl = lane & (S::NB - 1), half = lane >> 5;
    NodeDesc<R> nd{};
    cplx_t<R> fv = cmake<R>((R)0, (R)0);
    Chunk ch{0u, 0u, 0u, 0u}, chn{0u, 0u, 0u, 0u};
    if (cn < n) chn = chunks[cn];
    if (c < n) {
        ch = chunks[c];
        if ((uint32_t)nl < ch.count) {
            nd = desc[ch.start + nl];
            fv = fs[ch.start + nl];
        }
    }

    // per-lane offsets in a table row
    const int oa = nn, ob2 = S::O_B2 + cc, ob1 = S::O_B1 + ri * S::RS;
    int ox1[LBN], ox2[LBN];
#pragma unroll
    for (int j = 0; j < LBN; ++j) {
        ox1[j] = S::O_B1 + ri * S::RS + lb_b[j];
        ox2[j] = S::O_B2 + lb_c[j];
    }
    // The weight prologue: PW wavefronts per axis (two where the workgroup has
    // six or more), wavefront 3 PW clears the padding rows.  Wavefront wave < 3 PW
    // evaluates axis wave / PW; its lanes write the quads x0 = 4 q + 8 PW k of a
    // row, q = PW half + wave % PW.
    constexpr int PW = S::NW >= 6 ? 2 : 1;
    // FP64 Gaussian: a quad's weights from two exps,
    // exp(-b (s0 - u)^2) = exp(-b s0^2) exp(2 b s0)^u exp(-b u^2), with the last
    // factor per axis (s0 is at most m + 4 in magnitude where a weight is used,
    // so neither factor leaves the FP64 range)
    constexpr bool GQUAD = WIN == WIN_GAUSSIAN && sizeof(R) == 8;
    R gq1 = (R)1, gq2 = (R)1, gq3 = (R)1;
    if constexpr (GQUAD) {
        const int ax = wave / PW < 3 ? wave / PW : 2;
        const R kb_ax = ax == 0 ? wc.kb[0] : (ax == 1 ? wc.kb[1] : wc.kb[2]);
        gq1 = wexp<R>(-kb_ax);
        gq2 = wexp<R>((R)-4 * kb_ax);
        gq3 = wexp<R>((R)-9 * kb_ax);
    }
    const R *rowg = &tw[0][0] + g * S::ROW;   // step ks reads row 4 ks + g
    const R *rowb = rowg + ob1 + rb;          // the wavefront's first row weight
    const acc_t z4 = {(R)0, (R)0, (R)0, (R)0};

    // The chunk loop of a wavefront with NR rows and NL extra blocks, both
    // fixed at compile time (one instance per wavefront shape, at most two per
    // window width): no branches between the loads and the instructions of a
    // 4-node step, and the accumulators of one instance only.
    auto body = [&](auto nr_c, auto nl_c) {
        constexpr int NR = decltype(nr_c)::value, NL = decltype(nl_c)::value;
        acc_t acc0[S::RW][2], acc1[S::RW][2], accL0[LBN], accL1[LBN];

        // The 4-node steps [0, ks0) into set 0 and [ks0, ks1) into set 1; the
        // next step's operands load while the current step's MFMAs issue (two
        // operand sets in turn, so no copies between steps).
        auto ksteps = [&](int ks0, int ks1) {
            struct Ops {
                R a, w2lo, w2hi, w1[NR], x1[NL > 0 ? NL : 1], x2[NL > 0 ? NL : 1];
            };
            // per-lane table pointers of a step's operands; step ks + 1 is 4 rows on
            constexpr int STEP = 4 * S::ROW;
            struct Ptrs {
                const R *a, *w2, *w1, *x1[NL > 0 ? NL : 1], *x2[NL > 0 ? NL : 1];
            };
            auto ptrs = [&](int ks) {
                Ptrs q;
                q.a = rowg + ks * STEP + oa;
                q.w2 = rowg + ks * STEP + ob2;
                q.w1 = rowb + ks * STEP;
#pragma unroll
                for (int j = 0; j < NL; ++j) {
                    q.x1[j] = rowg + ks * STEP + ox1[j];
                    q.x2[j] = rowg + ks * STEP + ox2[j];
                }
                return q;
            };
            auto advance = [&](Ptrs &q, int steps) {
                q.a += steps * STEP;
                q.w2 += steps * STEP;
                q.w1 += steps * STEP;
#pragma unroll
                for (int j = 0; j < NL; ++j) {
                    q.x1[j] += steps * STEP;
                    q.x2[j] += steps * STEP;
                }
            };
            // the operands of the step `d` steps past q's (d a constant: immediate offsets)
            auto load = [&](const Ptrs &q, int d) {
                Ops o;
                o.a = q.a[d * STEP];
                o.w2lo = q.w2[d * STEP];
                o.w2hi = q.w2[d * STEP + 8];
                const R *w1p = q.w1 + d * STEP;
                if constexpr (S::rows_aligned() && NR >= 4) {
                    // the first four row weights in one 16-byte read
                    const float4 q = *reinterpret_cast<const float4 *>(w1p);
                    o.w1[0] = q.x;
                    o.w1[1] = q.y;
                    o.w1[2] = q.z;
                    o.w1[3] = q.w;
#pragma unroll
                    for (int r = 4; r < NR; ++r) o.w1[r] = w1p[r];
                } else {
#pragma unroll
                    for (int r = 0; r < NR; ++r) o.w1[r] = w1p[r];
                }
#pragma unroll
                for (int j = 0; j < NL; ++j) {
                    o.x1[j] = q.x1[j][d * STEP];
                    o.x2[j] = q.x2[j][d * STEP];
                }
                return o;
            };
            auto mma = [&](const Ops &o, acc_t(&acc)[S::RW][2], acc_t(&accL)[LBN]) {
#pragma unroll
                for (int r = 0; r < NR; ++r) {
                    acc[r][0] = M::op(o.a, o.w1[r] * o.w2lo, acc[r][0]);
                    acc[r][1] = M::op(o.a, o.w1[r] * o.w2hi, acc[r][1]);
                }
#pragma unroll
                for (int j = 0; j < NL; ++j) {
                    const R bv = lb_ok[j] ? o.x1[j] * o.x2[j] : (R)0;
                    accL[j] = M::op(o.a, bv, accL[j]);
                }
            };
            auto run = [&](int k0, int k1, acc_t(&acc)[S::RW][2], acc_t(&accL)[LBN]) {
                if (k0 >= k1) return;
                // (the scheduling barriers keep each step's loads ahead of the
                // previous step's instructions, which the scheduler would sink
                // to their uses)
                // a set's last step reads the operands of one step past it: the two
                // padded sets fill at most NB + 4 rows (their sizes are multiples of 4
                // summing to at most NB + 6), so those four rows are inside the table
                static_assert(S::NB % 4 == 0 && S::NPOS >= S::NB + 8, "one step past the last stays in the table");
                Ptrs q = ptrs(k0);
                Ops o0 = load(q, 0);
                for (int ks = k0;; ks += 2) {
                    const Ops o1 = load(q, 1);
                    __builtin_amdgcn_sched_barrier(0);
                    mma(o0, acc, accL);
                    if (ks + 1 >= k1) break;
                    o0 = load(q, 2);
                    advance(q, 2);
                    __builtin_amdgcn_sched_barrier(0);
                    mma(o1, acc, accL);
                    if (ks + 2 >= k1) break;
                }
            };
            run(0, ks0, acc0, accL0);
            if constexpr (S::SETS2) run(ks0, ks1, acc1, accL1);
        };
        uint32_t ticket = 0;
        while (c < n) {
#pragma unroll
            for (int r = 0; r < S::RW; ++r) acc0[r][0] = acc0[r][1] = acc1[r][0] = acc1[r][1] = z4;
#pragma unroll
            for (int j = 0; j < LBN; ++j) accL0[j] = accL1[j] = z4;
            if (tid == 0) ticket = atomicAdd(counter, 1u);
            for (uint32_t s0 = 0; s0 < ch.count; s0 += S::NB) {
                const int nb = __builtin_amdgcn_readfirstlane(
                    ch.count - s0 < (uint32_t)S::NB ? (int)(ch.count - s0) : S::NB);
                if (s0 > 0) __syncthreads();   // every wavefront is done with the previous batch's tables
                const bool valid = nl < nb;
                // the batch's node, loaded one batch ahead: opaque from here on, so
                // no use of it is moved up to its load (which would wait on the
                // load there, before the previous batch's steps)
                uint32_t loc = (uint32_t)nd.loc;
                R nf0 = nd.f[0], nf1 = nd.f[1], nf2 = nd.f[2], fre = fv.x, fim = fv.y;
                asm volatile("" : "+v"(loc), "+v"(nf0), "+v"(nf1), "+v"(nf2), "+v"(fre), "+v"(fim));
                const int lo0 = loc & 255, lo1 = (loc >> 8) & 255, loL = (loc >> 16) & 255;
                const bool s1 = S::SETS2 && valid && lo0 > S::LO_MAX0;
                // set masks over the lower half; rank = the node's place in its set
                const uint64_t m1 = __ballot(s1 && half == 0), m0 = __ballot(valid && !s1 && half == 0);
                const int n0 = __popcll(m0), n1 = __popcll(m1);
                const int p0pad = (n0 + 3) & ~3, npos = p0pad + ((n1 + 3) & ~3);
                const uint64_t mine = s1 ? m1 : m0;
                const int rank = __popcll(mine & ((1ull << nl) - 1ull));
                const int pos = s1 ? p0pad + rank : rank;
                if (wave < 3 * PW) {
                    if (valid) {
                        const int t = PW == 1 ? wave : wave / PW;   // axis
                        const int tq = PW == 1 ? half : PW * half + (wave - t * PW);
                        const R ft = t == 0 ? nf0 : (t == 1 ? nf1 : nf2);
                        const R kbt = t == 0 ? wc.kb[0] : (t == 1 ? wc.kb[1] : wc.kb[2]);
                        const R nrt = t == 0 ? wc.nrm[0] : (t == 1 ? wc.nrm[1] : wc.nrm[2]);
                        const R shpt = t == 0 ? wc.shp[0] : (t == 1 ? wc.shp[1] : wc.shp[2]);
                        // cell x of the table row has tap l = x - lo (weight zero outside [0, W))
                        const int lo = t == 0 ? lo0 - (s1 ? S::OFF1 : 0) : (t == 1 ? lo1 : loL);
                        const int nx = t == 0 ? 16 : S::RS;
                        R *row = &tw[pos][0] + (t == 0 ? 0 : (t == 1 ? S::O_B1 : S::O_B2));
                        const R sc0 = t == 1 ? fre : (R)1, sc1 = fim;
#pragma unroll 1
                        for (int x0 = 4 * tq; x0 < nx; x0 += 8 * PW) {
                            R v[4];
                            if constexpr (GQUAD) {
                                const R s0 = ft + (R)(wc.m - (x0 - lo));
                                const R e0 = nrt * wexp<R>(-kbt * s0 * s0);
                                const R r = wexp<R>((R)2 * kbt * s0);
                                const R r2 = r * r;
                                v[0] = e0;
                                v[1] = e0 * r * gq1;
                                v[2] = e0 * r2 * gq2;
                                v[3] = e0 * (r2 * r) * gq3;
                            }
#pragma unroll
                            for (int u = 0; u < 4; ++u) {
                                const int l = x0 + u - lo;
                                const R sv = ft + (R)(wc.m - l);
                                R w;
                                if constexpr (GQUAD) w = v[u];
                                else if constexpr (WIN == WIN_KAISER_BESSEL) w = kb_phi<R>(sv, (R)wc.m, shpt);
                                else if constexpr (WIN == WIN_SINC) w = sinc_phi<R>(sv, wc.m, shpt);
                                else w = nrt * wexp<R>(-kbt * sv * sv);
                                v[u] = l >= 0 && l < W ? w : (R)0;
                            }
#pragma unroll
                            for (int u = 0; u < 4; ++u) row[x0 + u] = v[u] * sc0;
                            if (t == 1) {
#pragma unroll
                                for (int u = 0; u < 4; ++u) row[S::RS + x0 + u] = v[u] * sc1;
                            }
                        }
                    }
                } else if (wave == 3 * PW) {
                    // zero rows for the padding positions [n0, p0pad) and [p0pad + n1, npos)
                    constexpr int U = S::ROW / E16;   // 16-byte units per row
                    const int pad0 = p0pad - n0, npad = npos - n0 - n1;
                    for (int i = lane; i < npad * U; i += 64) {
                        const int pi = i / U, u = i - pi * U;
                        const int pp = pi < pad0 ? n0 + pi : p0pad + n1 + (pi - pad0);
                        reinterpret_cast<V16 *>(&tw[pp][0])[u] = V16{};
                    }
                }
                __syncthreads();
                // the next batch of this chunk loads during this one's steps
                if (s0 + S::NB < ch.count) {
                    const uint32_t sn = s0 + S::NB + (uint32_t)nl;
                    if (sn < ch.count) {
                        nd = desc[ch.start + sn];
                        fv = fs[ch.start + sn];
                    }
                }
                ksteps(p0pad >> 2, S::SETS2 ? npos >> 2 : p0pad >> 2);
            }
            // the sum of the two sets per plane into the slot image; register i of lane
            // group g holds plane M::row(g, i) of set 0 and M::row(g, i) + OFF1 of set 1
            auto put = [&](const acc_t &v0, const acc_t &v1, int b, int cl, bool ok) {
                R up[4];   // set 1 at plane M::row(g, i)
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    if (!S::SETS2) {
                        up[i] = (R)0;
                    } else if constexpr (sizeof(R) == 4) {
                        // plane 4g + i of set 1 is row 4g + i - OFF1: this lane
                        // group's register i - OFF1, or the previous group's
                        if (i >= S::OFF1) {
                            up[i] = v1[i - S::OFF1];
                        } else {
                            const R t = __shfl(v1[4 - S::OFF1 + i], lane - 16);
                            up[i] = g > 0 ? t : (R)0;
                        }
                    } else {
                        // plane g + 4i of set 1 is row g - OFF1 + 4i: register i of lane
                        // group g - OFF1, or register i - 1 of group g - OFF1 + 4
                        const R t1 = __shfl(v1[i], lane - 16 * S::OFF1);
                        const R t2 = i > 0 ? __shfl(v1[i > 0 ? i - 1 : 0], lane + 16 * (4 - S::OFF1)) : (R)0;
                        up[i] = g >= S::OFF1 ? t1 : t2;
                    }
                }
                if (!ok) return;
                if constexpr (sizeof(R) == 4) {
                    const MfmaSlotCol o = mfma_slot_col<L>(g, b, cl, ri);
#pragma unroll
                    for (int i = 0; i < 4; ++i)
                        if (4 * g + i < S::P) stg[o.base + i * o.step] = v0[i] + up[i];
                    if (S::SETS2 && g == 3) {
#pragma unroll
                        for (int i = 0; i < S::OFF1; ++i) stg[o.base16 + i * o.step] = v1[4 - S::OFF1 + i];
                    }
                } else {
#pragma unroll
                    for (int i = 0; i < 4; ++i) {
                        const int a = g + 4 * i;
                        if (a < S::P) stg[mfma_slot_off<L>(a, b, cl, ri)] = v0[i] + up[i];
                    }
                    // planes 16 + j (rows 16 - OFF1 + j of set 1: register 3 of group 4 - OFF1 + j)
                    if (S::SETS2 && g >= 4 - S::OFF1)
                        stg[mfma_slot_off<L>(16 + g - (4 - S::OFF1), b, cl, ri)] = v1[3];
                }
            };
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                put(acc0[r][0], acc1[r][0], rb + r, cc, cc < S::P);
                put(acc0[r][1], acc1[r][1], rb + r, 8 + cc, 8 + cc < S::P);
            }
#pragma unroll
            for (int j = 0; j < NL; ++j) put(accL0[j], accL1[j], lb_b[j], lb_c[j], lb_ok[j]);
            if (tid == 0) s_tk[0] = ticket;
            __syncthreads();   // the image is whole and the ticket visible
            const uint32_t cur_c = c;
            c = cn;
            ch = chn;
            cn = s_tk[0];
            // the next chunk's first batch loads during the copy-out, and the record
            // of the chunk after it during the next chunk's steps
            if (c < n && (uint32_t)nl < ch.count) {
                nd = desc[ch.start + nl];
                fv = fs[ch.start + nl];
            }
            if (cn < n) chn = chunks[cn];
            // the next chunk writes the image only after its first batch's barrier,
            // which every wavefront reaches after its copy-out reads
            V16 *dst = reinterpret_cast<V16 *>(slots + (int64_t)cur_c * L::SC);
            const V16 *src = reinterpret_cast<const V16 *>(stg);
            for (int i = tid; i < 2 * L::SC / E16; i += NT) stream_store(dst + i, src[i]);
        }
    };

    using std::integral_constant;
    constexpr int RLO = S::P / S::NW, LLO = S::NLOB / S::NW;
    const int nlb = S::extra(wave);
    if constexpr (S::shape(RLO, LLO))
        if (nrows == RLO && nlb == LLO) return body(integral_constant<int, RLO>{}, integral_constant<int, LLO>{});
    if constexpr (S::shape(RLO, LLO + 1))
        if (nrows == RLO && nlb == LLO + 1)
            return body(integral_constant<int, RLO>{}, integral_constant<int, LLO + 1>{});
    if constexpr (S::shape(RLO + 1, LLO))
        if (nrows == RLO + 1 && nlb == LLO)
            return body(integral_constant<int, RLO + 1>{}, integral_constant<int, LLO>{});
    if constexpr (S::shape(RLO + 1, LLO + 1))
        if (nrows == RLO + 1 && nlb == LLO + 1)
            return body(integral_constant<int, RLO + 1>{}, integral_constant<int, LLO + 1>{});
}

// ------------------------------------------------- 2D, 16^2 tiles ----
//
// The 2D adjoint on 16^2 tiles (FP32): tile[a][c] = sum_j wa_j[a] (f_j wc_j[c]),
// a rank-K product of A (rows x nodes) and B (nodes x columns, re/im in
// adjacent columns).  The padded tile has P = 16 + W - 1 <= 32 rows, so one
// v_mfma_f32_32x32x2f32 covers every row: A[m = lane % 32][k = lane / 32],
// B[k = lane / 32][n = lane % 32], and accumulator register i of lane l holds
// D[i % 4 + 8 (i / 4) + 4 (l / 32)][l % 32] (checked on the GPU with an
// in-bounds probe).  Two 32-column blocks hold cells [0, 16) and [16, P).
// One wavefront per chunk; per batch of 32 nodes the weights go to LDS in
// padded-tile coordinates (zero outside the window), lanes n and n + 32
// evaluating node n's row and column weights (the batch's nodes loaded during
// the previous batch's steps); then two MFMAs per two nodes.
// The accumulators go straight to the slot: for a fixed register the 32
// lanes of a half-wavefront store 32 consecutive floats of one row.
typedef float mf16_t __attribute__((ext_vector_type(16)));

template <int W> struct MfmaSpread2 {
    static constexpr int T = 16;            // tile extent per axis
    static constexpr int P = T + W - 1;     // padded extent per axis
    static constexpr int NB = 32;           // nodes per batch
    static constexpr int O_B = 32;          // table row: 32 row weights, then 64 columns (re, im)
    static constexpr int ROW = 100;         // floats per table row (the two nodes a step reads start 4 banks apart)
    static_assert(P <= 32 && O_B + 64 <= ROW, "a padded tile fits one 32-row product and a table row");
};
template <int W> constexpr bool kMfmaSpread2 = W >= 2 && 16 + W - 1 <= 32;

template <int W>
__device__ __forceinline__ float mfma2_weight(int win, float s, int m, float nrm, float kb, float shp)
{
    if (win == WIN_KAISER_BESSEL) return kb_phi<float>(s, (float)m, shp);
    if (win == WIN_SINC) return sinc_phi<float>(s, m, shp);
    return nrm * wexp<float>(-kb * s * s);
}

template <int W>
__device__ __forceinline__ void tile_spread_mfma2(const NodeDesc<float> *__restrict__ desc,
                                                  const cplx_t<float> *__restrict__ fs, const Chunk ch,
                                                  cplx_t<float> *__restrict__ slot, const WinConst<float> &wc)
{
    using S = MfmaSpread2<W>;
    using L = TileLayout<float, 2, W, 1, 1>;   // slot layout (shared with the fold)
    static_assert(L::T0 == S::T && L::TL == S::T && L::P0 == S::P && L::PL == S::P && L::P1 == 1,
                  "16^2 tiles");
    __shared__ __align__(16) float tw[S::NB][S::ROW];
    const int lane = threadIdx.x, nl = lane & 31, half = lane >> 5;
    mf16_t acc0, acc1;
#pragma unroll
    for (int i = 0; i < 16; ++i) acc0[i] = acc1[i] = 0.f;
    // lane (nl, half): node nl of the next batch (the coefficient on half 1 only)
    NodeDesc<float> nd{};
    cplx_t<float> fv = cmake<float>(0.f, 0.f);
    if ((uint32_t)nl < ch.count) {
        nd = desc[ch.start + nl];
        if (half == 1) fv = fs[ch.start + nl];
    }
    for (uint32_t s0 = 0; s0 < ch.count; s0 += S::NB) {
        const int nb = ch.count - s0 < (uint32_t)S::NB ? (int)(ch.count - s0) : S::NB;
        if (s0 > 0) __syncthreads();   // the previous batch's steps have read the table
        // lane (nl, half): half 0 writes node nl's row weights, half 1 its column weights
        // times the coefficient; rows past the batch stay zero (a step reads two nodes)
        float *row = &tw[nl][0] + (half == 0 ? 0 : S::O_B);
        const int nz = half == 0 ? S::O_B / 4 : 16;   // 16-byte units of the lane's part
        for (int u = 0; u < nz; ++u) reinterpret_cast<float4 *>(row)[u] = make_float4(0.f, 0.f, 0.f, 0.f);
        if (nl < nb) {
            const int t = half;   // axis
            // the node, loaded during the previous batch: opaque here, so its first use
            // is not moved up to the load
            uint32_t loc = nd.loc;
            float ft = t == 0 ? nd.f[0] : nd.f[1], fre = t == 1 ? fv.x : 1.f, fim = t == 1 ? fv.y : 0.f;
            asm volatile("" : "+v"(loc), "+v"(ft), "+v"(fre), "+v"(fim));
            const int lo = (int)((loc >> (8 * t)) & 255u);
            const float nrm = wc.nrm[t], kb = wc.kb[t], shp = wc.shp[t];
#pragma unroll
            for (int l = 0; l < W; ++l) {
                const float w = mfma2_weight<W>(wc.win, ft + (float)(wc.m - l), wc.m, nrm, kb, shp);
                if (t == 0) {
                    row[lo + l] = w;
                } else {
                    row[2 * (lo + l)] = w * fre;
                    row[2 * (lo + l) + 1] = w * fim;
                }
            }
        }
        __syncthreads();
        // the next batch's node loads during this batch's steps
        if (s0 + S::NB + nl < ch.count) {
            nd = desc[ch.start + s0 + S::NB + nl];
            if (half == 1) fv = fs[ch.start + s0 + S::NB + nl];
        }
        // two nodes per step: lane (nl, half) supplies node k + half's row nl and columns nl
        for (int k = 0; k < nb; k += 2) {
            const float *r = &tw[k + half][0];
            const float a = r[nl];
            acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, r[S::O_B + nl], acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a, r[S::O_B + 32 + nl], acc1, 0, 0, 0);
        }
    }
    // column nl of block 0 is cell nl / 2 (part nl % 2), of block 1 cell 16 + nl / 2
    float *out = reinterpret_cast<float *>(slot);
    const int c0 = nl >> 1, ri = nl & 1, c1 = 16 + c0;
    constexpr int EL1 = L::ext(L::TL, L::PL, 1);
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const int a = (i & 3) + 8 * (i >> 2) + 4 * half;
        if (a < S::P) {
            const int qa = a >> 4, ra = a & 15;
            out[2 * (L::block_off(qa, 0, 0) + ra * S::T + c0) + ri] = acc0[i];
            if (c1 < S::P) out[2 * (L::block_off(qa, 0, 1) + ra * EL1 + (c1 - 16)) + ri] = acc1[i];
        }
    }
}

// One wavefront per chunk of a plan's binning (blocks past the chunk count return).
template <int W>
__global__ __launch_bounds__(64) void k_tile_spread_mfma2(const NodeDesc<float> *__restrict__ desc,
                                                          const cplx_t<float> *__restrict__ fs,
                                                          const Chunk *__restrict__ chunks,
                                                          const uint32_t *__restrict__ n_chunks,
                                                          cplx_t<float> *__restrict__ slots, WinConst<float> wc)
{
    using L = TileLayout<float, 2, W, 1, 1>;
    if (blockIdx.x >= n_chunks[0]) return;
    tile_spread_mfma2<W>(desc, fs, chunks[blockIdx.x], slots + (int64_t)blockIdx.x * L::SC, wc);
}

// The batch form: workgroup b runs chunk map[b].blk of job map[b].job.
template <int W>
__global__ __launch_bounds__(64) void k_tile_spread_mfma2_batch(const GridJob<float> *__restrict__ jobs,
                                                                const BlockMap *__restrict__ map)
{
    using L = TileLayout<float, 2, W, 1, 1>;
    const BlockMap bm = map[blockIdx.x];
    const GridJob<float> &J = jobs[bm.job];
    tile_spread_mfma2<W>(static_cast<const NodeDesc<float> *>(J.desc), J.fs, J.chunks[bm.blk],
                         J.slots + (int64_t)bm.blk * L::SC, J.wc);
}

}  // namespace cunfft
