// fft.hip -- LDS-resident Stockham FFT pass for gfx950 with fused roll-off.
//
// Replaces the three reference steps around the FFT:
//   forward  kernelRollOf_{1,2,3}d + memset  (src/cunfft_kernel.cu:240-412,
//            src/cunfft_wrapper.cu:32-75)  then cufftExec(.., CUFFT_FORWARD)
//   adjoint  cufftExec(.., CUFFT_INVERSE) then kernelRollOf_adjoint_{1,2,3}d
//            (src/cunfft_kernel.cu:264-467, src/cunfft_wrapper.cu:77-121)
// A d-dimensional transform is d passes, one per axis.  Each workgroup loads
// C whole lines of the axis into LDS, runs radix-4 (and one radix-2) Stockham
// stages there -- then radix-3 stages when L = 2^a 3^b is not a power of two
// -- and writes the lines back: one HBM read and one write per pass.  The forward passes skip lines that are still all zero and read only
// the N_t nonzero entries of each line; the first forward pass reads fhat and
// applies exp(fac k^2) on the way in.  The adjoint passes write only the N_t
// entries that survive the crop; the last one writes fhat with exp(fac k^2).
// Twiddles come from a per-length table computed in extended precision on
// the host.
#include <algorithm>
#include <type_traits>

#include "internal.hpp"

namespace cunfft {

constexpr int kFftThreads = 256;
// Elements C * L per workgroup: up to 4096 (an 8192-element FP32 variant,
// for 16 columns of a 512-point column pass, measured slower and was
// dropped).  The kernels hold E / 1024 radix-4 butterflies per thread in
// registers.
template <typename T> constexpr int fft_max_elems() { return 4096; }
constexpr int kFftSmallElems = 4096;
// global loads in flight per thread from fhat (FP64: fewer, the registers run
// short); loads from the grid, with no FP64 roll-off factor to hold, take twice
// as many (FP64 3D 256^3 passes 3% faster, FP32 unchanged)
template <typename T> constexpr int fft_load_batch() { return sizeof(T) == 4 ? 8 : 4; }
constexpr int kMaxRowsPerGroup = 64;

template <typename T>
__device__ __forceinline__ cplx_t<T> cmul(const cplx_t<T> a, const cplx_t<T> b)
{
    return cmake<T>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

template <typename T>
__device__ __forceinline__ cplx_t<T> twiddle(const cplx_t<T> *__restrict__ tw, int q, int sign)
{
    cplx_t<T> w = tw[q];
    if (sign > 0) w.y = -w.y;
    return w;
}

// 4-point DFT with kernel exp(sign 2 pi i r q / 4).
template <typename T> __device__ __forceinline__ void dft4(cplx_t<T> (&v)[4], int sign)
{
    const cplx_t<T> a0 = cmake<T>(v[0].x + v[2].x, v[0].y + v[2].y);
    const cplx_t<T> a1 = cmake<T>(v[0].x - v[2].x, v[0].y - v[2].y);
    const cplx_t<T> a2 = cmake<T>(v[1].x + v[3].x, v[1].y + v[3].y);
    const cplx_t<T> a3 = cmake<T>(v[1].x - v[3].x, v[1].y - v[3].y);
    // omega * a3 with omega = sign * i
    const T sr = sign > 0 ? -a3.y : a3.y;
    const T si = sign > 0 ? a3.x : -a3.x;
    v[0] = cmake<T>(a0.x + a2.x, a0.y + a2.y);
    v[2] = cmake<T>(a0.x - a2.x, a0.y - a2.y);
    v[1] = cmake<T>(a1.x + sr, a1.y + si);
    v[3] = cmake<T>(a1.x - sr, a1.y - si);
}

struct LineInfo {
    int64_t gbase;  // grid offset of the line's axis start (without the column)
    int64_t fbase;  // fhat offset of the row (passes on the last axis)
    double fac;     // product of roll-off factors of the axes before t
};

// Map the line index pa (support of the axes before t, FFT order, last axis
// fastest) to its grid row, its fhat row and its roll-off factor.
template <typename T>
__device__ LineInfo line_info(const FftPass<T> &P, int64_t pa)
{
    LineInfo li;
    int64_t rem = pa;
    int64_t a = 0, gmul = 1, fq = 0, fmul = 1;
    double fac = 1.0;
    // axes tt = t-1 .. 0, unrolled so the per-axis fields are read by constant
    // index (no private-memory copy of the descriptor)
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int tt = 1 - k;
        if (tt >= P.t) continue;
        const int64_t Nn = P.N[tt], hN = Nn >> 1;
        const int64_t p = rem % Nn;
        rem /= Nn;
        const int64_t i = p < hN ? p : p + (P.n[tt] - Nn);
        const int64_t q = p < hN ? p + hN : p - hN;
        a += i * gmul;
        gmul *= P.n[tt];
        fq += q * fmul;
        fmul *= Nn;
        fac *= P.roc[P.rocoff[tt] + p];
    }
    li.gbase = a * P.Lf * P.So;
    li.fbase = fq * (P.t == 0 ? P.N[0] : (P.t == 1 ? P.N[1] : P.N[2]));
    li.fac = fac;
    return li;
}

template <typename T, int E>
__device__ __forceinline__ void stage4(cplx_t<T> *buf, const cplx_t<T> *__restrict__ tw, int logL,
                                       int C, int ldsL, int logNs, int sign, int tid)
{
    const int logQ = logL - 2;
    const int Q = 1 << logQ;
    const int Ns = 1 << logNs;
    const int nb = C << logQ;
    constexpr int kMaxB4 = E / 4 / kFftThreads;
    cplx_t<T> v[kMaxB4][4];
#pragma unroll
    for (int k = 0; k < kMaxB4; ++k) {
        const int b = tid + k * kFftThreads;
        if (b < nb) {
            const int c = b >> logQ, j = b & (Q - 1);
            const cplx_t<T> *base = buf + c * ldsL + j;
            v[k][0] = base[0];
            v[k][1] = base[Q];
            v[k][2] = base[2 * Q];
            v[k][3] = base[3 * Q];
            if (Ns > 1) {
                const int step = (j & (Ns - 1)) << (logL - 2 - logNs);
                v[k][1] = cmul<T>(v[k][1], twiddle<T>(tw, step, sign));
                v[k][2] = cmul<T>(v[k][2], twiddle<T>(tw, 2 * step, sign));
                v[k][3] = cmul<T>(v[k][3], twiddle<T>(tw, 3 * step, sign));
            }
            dft4<T>(v[k], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < kMaxB4; ++k) {
        const int b = tid + k * kFftThreads;
        if (b < nb) {
            const int c = b >> logQ, j = b & (Q - 1);
            const int out = ((j >> logNs) << (logNs + 2)) + (j & (Ns - 1));
            cplx_t<T> *ob = buf + c * ldsL + out;
            ob[0] = v[k][0];
            ob[Ns] = v[k][1];
            ob[2 * Ns] = v[k][2];
            ob[3 * Ns] = v[k][3];
        }
    }
    __syncthreads();
}

// Final radix-2 stage (Ns = L/2): butterfly j combines j and j + L/2 in place.
template <typename T, int E>
__device__ __forceinline__ void stage2_last(cplx_t<T> *buf, const cplx_t<T> *__restrict__ tw,
                                            int logL, int C, int ldsL, int sign, int tid)
{
    const int logH = logL - 1;
    const int H = 1 << logH;
    const int nb = C << logH;
    constexpr int kMaxB2 = E / 2 / kFftThreads;
#pragma unroll
    for (int k = 0; k < kMaxB2; ++k) {
        const int b = tid + k * kFftThreads;
        if (b < nb) {
            const int c = b >> logH, j = b & (H - 1);
            cplx_t<T> *base = buf + c * ldsL + j;
            const cplx_t<T> v0 = base[0];
            const cplx_t<T> v1 = cmul<T>(base[H], twiddle<T>(tw, j, sign));
            base[0] = cmake<T>(v0.x + v1.x, v0.y + v1.y);
            base[H] = cmake<T>(v0.x - v1.x, v0.y - v1.y);
        }
    }
    __syncthreads();
}

// 16-point DFT in registers with kernel exp(sign 2 pi i r q / 16), outputs in
// natural order: n = 4 n1 + n2, k = k1 + 4 k2 gives four 4-point DFTs over
// n1, the twiddles exp(sign 2 pi i n2 k1 / 16), and four 4-point DFTs over n2.
template <typename T> __device__ __forceinline__ void dft16(cplx_t<T> (&v)[16], int sign)
{
    const T c1 = (T)0.92387953251128675613;   // cos(pi/8)
    const T s1 = (T)0.38268343236508977173;   // sin(pi/8)
    const T h = (T)0.70710678118654752440;    // sqrt(1/2)
    const T sg = (T)sign;
    cplx_t<T> y[4][4];
#pragma unroll
    for (int n2 = 0; n2 < 4; ++n2) {
        cplx_t<T> a[4] = {v[n2], v[4 + n2], v[8 + n2], v[12 + n2]};
        dft4<T>(a, sign);
#pragma unroll
        for (int k1 = 0; k1 < 4; ++k1) y[n2][k1] = a[k1];
    }
    // z exp(sign i theta) for cos theta = c, sin theta = s
    auto rot = [&](cplx_t<T> z, T c, T s) {
        return cmake<T>(z.x * c - z.y * (sg * s), z.x * (sg * s) + z.y * c);
    };
    y[1][1] = rot(y[1][1], c1, s1);    // pi/8
    y[1][2] = rot(y[1][2], h, h);      // pi/4
    y[1][3] = rot(y[1][3], s1, c1);    // 3pi/8
    y[2][1] = rot(y[2][1], h, h);      // pi/4
    y[2][2] = cmake<T>(-sg * y[2][2].y, sg * y[2][2].x);   // pi/2
    y[2][3] = rot(y[2][3], -h, h);     // 3pi/4
    y[3][1] = rot(y[3][1], s1, c1);    // 3pi/8
    y[3][2] = rot(y[3][2], -h, h);     // 3pi/4
    y[3][3] = rot(y[3][3], -c1, -s1);  // 9pi/8
#pragma unroll
    for (int k1 = 0; k1 < 4; ++k1) {
        cplx_t<T> b[4] = {y[0][k1], y[1][k1], y[2][k1], y[3][k1]};
        dft4<T>(b, sign);
#pragma unroll
        for (int k2 = 0; k2 < 4; ++k2) v[k1 + 4 * k2] = b[k2];
    }
}

// One radix-16 Stockham stage (power-of-two L, Ns = 2^logNs): butterfly j of
// line c reads j + r Q (Q = L / 16), multiplies input r by
// exp(sign 2 pi i r (j mod Ns) / (16 Ns)), and writes its 16-point DFT to
// (j div Ns) 16 Ns + (j mod Ns) + r Ns.  A layout flag selects, per side, the
// lines at pitch ldsL (element i at i) or the padded lines at pitch ldsP
// (element i at i + i / 16): the first stage (Ns = 1) writes with stride 16,
// which the padding spreads over all LDS banks.
template <typename T, int E, bool PIN, bool POUT>
__device__ __forceinline__ void stage16(cplx_t<T> *buf, const cplx_t<T> *__restrict__ tw, int logL,
                                        int C, int ldsL, int ldsP, int logNs, int sign, int tid)
{
    const int logQ = logL - 4;
    const int Q = 1 << logQ;
    const int Ns = 1 << logNs;
    const int nb = C << logQ;
    const int pin = PIN ? ldsP : ldsL, pout = POUT ? ldsP : ldsL;
    constexpr int KB = (E / 16 + kFftThreads - 1) / kFftThreads;
    cplx_t<T> v[KB][16];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int b = tid + k * kFftThreads;
        if (b < nb) {
            const int c = b >> logQ, j = b & (Q - 1);
            const cplx_t<T> *base = buf + c * pin;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = j + r * Q;
                v[k][r] = base[PIN ? i + (i >> 4) : i];
            }
            if (Ns > 1) {
                const int step = (j & (Ns - 1)) << (logL - 4 - logNs);
#pragma unroll
                for (int r = 1; r < 16; ++r) v[k][r] = cmul<T>(v[k][r], twiddle<T>(tw, r * step, sign));
            }
            dft16<T>(v[k], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int b = tid + k * kFftThreads;
        if (b < nb) {
            const int c = b >> logQ, j = b & (Q - 1);
            const int out = ((j >> logNs) << (logNs + 4)) + (j & (Ns - 1));
            cplx_t<T> *ob = buf + c * pout;
#pragma unroll
            for (int r = 0; r < 16; ++r) {
                const int i = out + r * Ns;
                ob[POUT ? i + (i >> 4) : i] = v[k][r];
            }
        }
    }
    __syncthreads();
}

// Lines of at least 256 points run two radix-16 stages first, the first
// writing the padded layout; LDS elements per line of the widest layout.
__host__ __device__ constexpr bool fft_radix16(int logL) { return logL >= 8; }
__host__ __device__ constexpr int fft_lds_pitch(int L, int logL, int ldsL)
{
    return fft_radix16(logL) && L + (L >> 4) > ldsL ? L + (L >> 4) : ldsL;
}

// a = q d + r for 0 <= a < 2^22 with rd = 1/d in FP32 (exact after one
// correction step)
__device__ __forceinline__ void divmod_small(int a, int d, float rd, int &q, int &r)
{
    q = (int)((float)a * rd);
    r = a - q * d;
    if (r < 0) {
        --q;
        r += d;
    } else if (r >= d) {
        ++q;
        r -= d;
    }
}

// R-point DFT with kernel exp(sign 2 pi i r q / R), R = 2, 3, 4.
template <typename T, int R> __device__ __forceinline__ void dft_r(cplx_t<T> (&v)[R], int sign)
{
    if constexpr (R == 2) {
        const cplx_t<T> a = v[0], b = v[1];
        v[0] = cmake<T>(a.x + b.x, a.y + b.y);
        v[1] = cmake<T>(a.x - b.x, a.y - b.y);
    } else if constexpr (R == 3) {
        // y1,2 = a - (b + c)/2 +- sign i (sqrt(3)/2)(b - c)
        const T h = (T)0.86602540378443864676;   // sqrt(3) / 2
        const cplx_t<T> a = v[0], b = v[1], c = v[2];
        const T sx = b.x + c.x, sy = b.y + c.y;
        const T dx = (T)sign * h * (b.x - c.x), dy = (T)sign * h * (b.y - c.y);
        const T mx = a.x - (T)0.5 * sx, my = a.y - (T)0.5 * sy;
        v[0] = cmake<T>(a.x + sx, a.y + sy);
        v[1] = cmake<T>(mx - dy, my + dx);
        v[2] = cmake<T>(mx + dy, my - dx);
    } else {
        dft4<T>(v, sign);
    }
}

// One Stockham stage of radix R for any L (stride Q = L/R, Ns = product of
// the radices before it): butterfly j of line c reads j + rQ, multiplies
// input r by exp(sign 2 pi i r (j mod Ns) / (R Ns)), and writes its R-point
// DFT to (j div Ns) R Ns + (j mod Ns) + r Ns.
template <typename T, int E, int R>
__device__ __forceinline__ void stage_any(cplx_t<T> *buf, const cplx_t<T> *__restrict__ tw, int L,
                                          int C, int ldsL, int Ns, int sign, int tid)
{
    const int Q = L / R;
    const int nb = C * Q;
    const int tws = L / (R * Ns);    // twiddle-table step of exp(2 pi i / (R Ns))
    const float rQ = 1.0f / (float)Q, rN = 1.0f / (float)Ns;
    constexpr int KB = (E / R + kFftThreads - 1) / kFftThreads;
    cplx_t<T> v[KB][R];
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int b = tid + k * kFftThreads;
        if (b < nb) {
            int c, j, jq, jr;
            divmod_small(b, Q, rQ, c, j);
            divmod_small(j, Ns, rN, jq, jr);
            const cplx_t<T> *base = buf + c * ldsL + j;
#pragma unroll
            for (int r = 0; r < R; ++r) v[k][r] = base[r * Q];
#pragma unroll
            for (int r = 1; r < R; ++r) v[k][r] = cmul<T>(v[k][r], twiddle<T>(tw, r * jr * tws, sign));
            dft_r<T, R>(v[k], sign);
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < KB; ++k) {
        const int b = tid + k * kFftThreads;
        if (b < nb) {
            int c, j, jq, jr;
            divmod_small(b, Q, rQ, c, j);
            divmod_small(j, Ns, rN, jq, jr);
            cplx_t<T> *ob = buf + c * ldsL + jq * (R * Ns) + jr;
#pragma unroll
            for (int r = 0; r < R; ++r) ob[r * Ns] = v[k][r];
        }
    }
    __syncthreads();
}

template <typename T, int E, int SPLIT>
__device__ __forceinline__ void fft_pass(const FftPass<T> &P, int64_t blk)
{
    using C2 = cplx_t<T>;
    extern __shared__ __align__(16) unsigned char smem_raw[];
    C2 *buf = reinterpret_cast<C2 *>(smem_raw);
    __shared__ int64_t s_gbase[kMaxRowsPerGroup];   // input offset of the line (column 0)
    __shared__ int64_t s_obase[kMaxRowsPerGroup];   // SPLIT_B: output offset of the line
    __shared__ int64_t s_fbase[kMaxRowsPerGroup];
    __shared__ double s_fac[kMaxRowsPerGroup];
    __shared__ int s_valid[kMaxRowsPerGroup];
    __shared__ int s_k2[kMaxRowsPerGroup];          // SPLIT_B: output index k2 of the line
    __shared__ int s_i1[kMaxRowsPerGroup];          // SPLIT_A: index i1 of the column

    const int tid = threadIdx.x;
    const int L = P.L, logL = P.logL, C = P.C, logC = P.logC, ldsL = P.ldsL;
    const bool rows = (P.S == 1);
    int64_t pa0, s0;
    if (rows) {
        pa0 = blk << logC;
        s0 = 0;
    } else {
        const int64_t nch = P.S >> logC;
        pa0 = blk / nch;
        s0 = (blk - pa0 * nch) << logC;
    }
    if (tid < (rows ? C : 1)) {
        const int64_t pa = pa0 + tid;
        const bool valid = pa < P.n_pa;
        int64_t pb = valid ? pa : 0, k2 = 0;
        if (SPLIT == SPLIT_B) {
            k2 = pb % P.L2;
            pb /= P.L2;
        }
        const LineInfo li = line_info<T>(P, pb);
        s_gbase[tid] = li.gbase + (SPLIT == SPLIT_B ? k2 * P.L1 * P.So : 0);
        s_obase[tid] = li.gbase;
        s_fbase[tid] = li.fbase;
        s_fac[tid] = li.fac;
        s_valid[tid] = valid;
        s_k2[tid] = (int)k2;
    }
    if (SPLIT == SPLIT_A && tid < C) s_i1[tid] = (int)((s0 + tid) / P.So);
    // the twiddle table of the axis, staged in LDS after the lines
    const int ldsP = L + (L >> 4);   // padded line pitch of the radix-16 stages
    C2 *twl = buf + C * fft_lds_pitch(L, logL, ldsL);
    for (int q = tid; q < L; q += kFftThreads) twl[q] = P.tw[q];
    __syncthreads();

    const bool pow2 = logL >= 0;
    const float rL = 1.0f / (float)L;
    const int total = C * L;
    const int64_t Lf = P.Lf;
    const int Nt = (int)(P.t == 0 ? P.N[0] : (P.t == 1 ? P.N[1] : P.N[2]));
    const int hN = Nt >> 1;
    const double *roct =
        P.roc + (P.t == 0 ? P.rocoff[0] : (P.t == 1 ? P.rocoff[1] : P.rocoff[2]));
    // the full-axis index of element i of line / column c before the pass
    auto in_index = [&](int i, int c) -> int64_t {
        if constexpr (SPLIT == SPLIT_A) return (int64_t)s_i1[c] + P.L1 * (int64_t)i;
        return i;
    };

    // loads in batches of kFftLoadBatch per thread, all issued before the
    // first LDS store (a pass is bound by memory latency otherwise): every
    // lane loads, from a clamped address where it has nothing to load, and
    // the value is zeroed after
    auto stage_in = [&](auto from_fhat_tag) {
        constexpr bool FROM_FHAT = decltype(from_fhat_tag)::value;
        constexpr int kFftLoadBatch = fft_load_batch<T>() * (FROM_FHAT ? 1 : 2);
        const bool in_pruned = P.mode & FFT_IN_PRUNED;
        for (int e0 = tid; e0 < total; e0 += kFftLoadBatch * kFftThreads) {
            C2 v[kFftLoadBatch];
            double fr[kFftLoadBatch];   // roll-off factor of a load from fhat
            int at[kFftLoadBatch];
            bool ok[kFftLoadBatch];
#pragma unroll
            for (int u = 0; u < kFftLoadBatch; ++u) {
                const int e = e0 + u * kFftThreads;
                int c, i;
                if (rows) {
                    if (pow2) {
                        c = e >> logL;
                        i = e & (L - 1);
                    } else {
                        divmod_small(e, L, rL, c, i);
                    }
                } else {
                    c = e & (C - 1);
                    i = e >> logC;
                }
                const bool in = e < total;
                const int li = rows && in ? c : 0;
                at[u] = in ? c * ldsL + i : -1;
                const int64_t ifl = in_index(i, in ? c : 0);
                const bool insupp = (ifl < hN) || (ifl >= Lf - hN);
                if constexpr (FROM_FHAT) {
                    ok[u] = in && s_valid[li] && insupp;
                    const int p = (int)(ifl < hN ? ifl : ifl - (Lf - Nt));
                    const int q = p < hN ? p + hN : p - hN;
                    const int64_t fb = s_fbase[li];
                    const double fl = s_fac[li];
                    v[u] = P.fhat_in[ok[u] ? fb + q : 0];
                    fr[u] = fl * roct[ok[u] ? p : 0];
                } else {
                    ok[u] = in && s_valid[li] && (!in_pruned || insupp);
                    const int64_t gb = s_gbase[li];
                    v[u] = P.g[ok[u] ? gb + (int64_t)i * P.S + s0 + (rows ? 0 : c) : 0];
                }
            }
#pragma unroll
            for (int u = 0; u < kFftLoadBatch; ++u) {
                C2 w = ok[u] ? v[u] : cmake<T>(0, 0);
                if constexpr (FROM_FHAT)
                    w = cmake<T>((T)((double)w.x * fr[u]), (T)((double)w.y * fr[u]));
                if constexpr (SPLIT == SPLIT_B) {
                    // w_Lf^(sign i1 k2), the angle reduced mod Lf and taken in FP64
                    if (at[u] >= 0) {
                        const int c = at[u] / ldsL, i = at[u] - c * ldsL;
                        const int64_t q = ((int64_t)i * s_k2[rows ? c : 0]) % Lf;
                        double sn, cs;
                        sincospi((double)P.sign * 2.0 * (double)q / (double)Lf, &sn, &cs);
                        const double wr = (double)w.x, wi = (double)w.y;
                        w = cmake<T>((T)(wr * cs - wi * sn), (T)(wr * sn + wi * cs));
                    }
                }
                if (at[u] >= 0) buf[at[u]] = w;
            }
        }
    };
    if (P.mode & FFT_FROM_FHAT)
        stage_in(std::true_type{});
    else
        stage_in(std::false_type{});
    __syncthreads();

    if (pow2) {
        int logNs = 0;
        if (fft_radix16(logL)) {
            stage16<T, E, false, true>(buf, twl, logL, C, ldsL, ldsP, 0, P.sign, tid);
            stage16<T, E, true, false>(buf, twl, logL, C, ldsL, ldsP, 4, P.sign, tid);
            logNs = 8;
            while (logNs + 4 <= logL) {
                stage16<T, E, false, false>(buf, twl, logL, C, ldsL, ldsP, logNs, P.sign, tid);
                logNs += 4;
            }
        }
        while (logNs + 2 <= logL) {
            stage4<T, E>(buf, twl, logL, C, ldsL, logNs, P.sign, tid);
            logNs += 2;
        }
        if (logNs < logL) stage2_last<T, E>(buf, twl, logL, C, ldsL, P.sign, tid);
    } else {
        // L = 2^a 3^b: radix 4 while the power-of-two part allows, one radix 2
        // if a is odd, then b radix-3 stages
        int p2 = L;
        for (int k = 0; k < P.n3; ++k) p2 /= 3;
        int Ns = 1;
        while (Ns * 4 <= p2) {
            stage_any<T, E, 4>(buf, twl, L, C, ldsL, Ns, P.sign, tid);
            Ns *= 4;
        }
        if (Ns < p2) {
            stage_any<T, E, 2>(buf, twl, L, C, ldsL, Ns, P.sign, tid);
            Ns *= 2;
        }
        for (int k = 0; k < P.n3; ++k) {
            stage_any<T, E, 3>(buf, twl, L, C, ldsL, Ns, P.sign, tid);
            Ns *= 3;
        }
    }

    for (int e = tid; e < total; e += kFftThreads) {
        int c, i;
        if (rows) {
            if (pow2) {
                c = e >> logL;
                i = e & (L - 1);
            } else {
                divmod_small(e, L, rL, c, i);
            }
        } else {
            c = e & (C - 1);
            i = e >> logC;
        }
        const int li = rows ? c : 0;
        if (!s_valid[li]) continue;
        if constexpr (SPLIT == SPLIT_A) {
            // the first half of a split axis: back in place, nothing pruned
            P.g[s_gbase[li] + (int64_t)i * P.S + s0 + (rows ? 0 : c)] = buf[c * ldsL + i];
            continue;
        }
        // full output index: i, or k2 + L2 k1 for the second half of a split axis
        const int64_t kf = SPLIT == SPLIT_B ? s_k2[li] + P.L2 * (int64_t)i : (int64_t)i;
        const bool insupp = (kf < hN) || (kf >= Lf - hN);
        if (P.mode & FFT_TO_FHAT) {
            if (insupp) {
                const int p = (int)(kf < hN ? kf : kf - (Lf - Nt));
                const int q = p < hN ? p + hN : p - hN;
                const C2 v = buf[c * ldsL + i];
                const double f = s_fac[li] * roct[p];
                P.fhat_out[s_fbase[li] + q] = cmake<T>((T)((double)v.x * f), (T)((double)v.y * f));
            }
        } else if (!(P.mode & FFT_OUT_PRUNED) || insupp) {
            if constexpr (SPLIT == SPLIT_B)
                P.dst[s_obase[li] + kf * P.So + s0 + (rows ? 0 : c)] = buf[c * ldsL + i];
            else
                P.g[s_gbase[li] + (int64_t)i * P.S + s0 + (rows ? 0 : c)] = buf[c * ldsL + i];
        }
    }
}

template <typename T, int E, int SPLIT>
__global__ __launch_bounds__(kFftThreads) void k_fft_pass(FftPass<T> P)
{
    fft_pass<T, E, SPLIT>(P, blockIdx.x);
}

// Passes of many transforms in one grid (a batch's jobs; LDS sized for the
// largest).  Batches run unsplit axes only.
template <typename T, int E>
__global__ __launch_bounds__(kFftThreads) void k_fft_pass_batch(const FftPass<T> *__restrict__ passes,
                                                                const BlockMap *__restrict__ map)
{
    const BlockMap bm = map[blockIdx.x];
    fft_pass<T, E, SPLIT_NONE>(passes[bm.job], bm.blk);
}

template <typename T>
void launch_fft_pass(const FftPass<T> &p, int64_t blocks, size_t lds_bytes, hipStream_t s)
{
    if (blocks <= 0) return;
    if (blocks >= (int64_t(1) << 32) / kFftThreads)
        throw Error(CUNFFT_ERR_UNSUPPORTED, "FFT pass needs too many workgroups");
    static_assert(fft_max_elems<float>() == kFftSmallElems && fft_max_elems<double>() == kFftSmallElems,
                  "one LDS line size");
    if (p.split == SPLIT_A)
        k_fft_pass<T, kFftSmallElems, SPLIT_A><<<(unsigned)blocks, kFftThreads, lds_bytes, s>>>(p);
    else if (p.split == SPLIT_B)
        k_fft_pass<T, kFftSmallElems, SPLIT_B><<<(unsigned)blocks, kFftThreads, lds_bytes, s>>>(p);
    else
        k_fft_pass<T, kFftSmallElems, SPLIT_NONE><<<(unsigned)blocks, kFftThreads, lds_bytes, s>>>(p);
    CUNFFT_HIP(hipGetLastError());
}

template <typename T>
void launch_fft_batch(const FftPass<T> *passes, const BlockMap *map, uint32_t nblk,
                      size_t lds_bytes, int max_elems, hipStream_t s)
{
    if (nblk == 0) return;
    if (max_elems > kFftSmallElems)
        k_fft_pass_batch<T, fft_max_elems<T>()><<<nblk, kFftThreads, lds_bytes, s>>>(passes, map);
    else
        k_fft_pass_batch<T, kFftSmallElems><<<nblk, kFftThreads, lds_bytes, s>>>(passes, map);
    CUNFFT_HIP(hipGetLastError());
}

static int ilog2_i(int64_t v)
{
    int l = 0;
    while ((int64_t(1) << l) < v) ++l;
    return l;
}

static int64_t prev_pow2_i(int64_t v)
{
    int64_t p = 1;
    while (2 * p <= v) p <<= 1;
    return p;
}

static int64_t next_pow2_i(int64_t v)
{
    int64_t p = 1;
    while (p < v) p <<= 1;
    return p;
}

// Lines per workgroup: rows (S = 1) are packed up to 2048 elements per
// workgroup; strided passes take C adjacent columns (coalesced C-wide reads)
// up to 4096 elements.  Sets C, logC, ldsL and the launch's blocks and LDS
// from P.L, P.S and P.n_pa.
template <typename T> static void pass_shape(FftLaunch<T> &out)
{
    FftPass<T> &P = out.P;
    P.logL = (P.L & (P.L - 1)) == 0 ? ilog2_i(P.L) : -1;
    P.n3 = 0;
    for (int l = P.L; l % 3 == 0; l /= 3) ++P.n3;
    const int64_t target = 2048;
    int64_t C;
    if (P.S == 1) {
        C = P.L >= target ? 1 : prev_pow2_i(target / P.L);
        C = std::min<int64_t>(C, 64);
        C = std::min<int64_t>(C, next_pow2_i(P.n_pa));
        out.blocks = (P.n_pa + C - 1) / C;
        P.ldsL = P.L;
    } else {
        // adjacent columns: C divides S (the largest power of two in S bounds it)
        const int64_t cdef = sizeof(T) == 8 ? 8 : 16;
        C = std::min<int64_t>(cdef, P.S & -P.S);
        C = std::min<int64_t>(C, std::max<int64_t>(1, prev_pow2_i(fft_max_elems<T>() / P.L)));
        out.blocks = P.n_pa * (P.S / C);
        P.ldsL = C > 1 ? P.L + 1 : P.L;
    }
    P.C = (int)C;
    P.logC = ilog2_i(C);
    // lines (at the widest pitch the stages use) + twiddles
    out.lds = ((size_t)C * fft_lds_pitch(P.L, P.logL, P.ldsL) + P.L) * sizeof(cplx_t<T>);
}

template <typename T>
static FftPass<T> pass_common(const Geom &ge, const double *roc, const int64_t *rocoff, int t,
                              bool forward, const cplx_t<T> *fhat_in, cplx_t<T> *fhat_out,
                              cplx_t<T> *g)
{
    FftPass<T> P{};
    P.g = g;
    P.fhat_in = fhat_in;
    P.fhat_out = fhat_out;
    P.roc = roc;
    for (int q = 0; q < 3; ++q) {
        P.rocoff[q] = rocoff[q];
        P.n[q] = ge.n[q];
        P.N[q] = ge.N[q];
    }
    P.d = ge.d;
    P.t = t;
    P.Lf = ge.n[t];
    P.So = 1;
    for (int q = t + 1; q < ge.d; ++q) P.So *= ge.n[q];
    P.n_pa = 1;
    for (int q = 0; q < t; ++q) P.n_pa *= ge.N[q];
    P.sign = forward ? -1 : +1;
    P.split = SPLIT_NONE;
    P.dst = nullptr;
    return P;
}

template <typename T>
FftLaunch<T> fft_pass_setup(const Geom &ge, const cplx_t<T> *tw, const double *roc,
                            const int64_t *rocoff, int t, bool forward, const cplx_t<T> *fhat_in,
                            cplx_t<T> *fhat_out, cplx_t<T> *g)
{
    const int d = ge.d;
    FftLaunch<T> out{};
    FftPass<T> &P = out.P;
    P = pass_common<T>(ge, roc, rocoff, t, forward, fhat_in, fhat_out, g);
    P.tw = tw;
    P.L = (int)ge.n[t];
    P.S = P.So;
    P.L1 = 1;
    P.L2 = P.L;
    if (forward)
        P.mode = (t == d - 1) ? FFT_FROM_FHAT : FFT_IN_PRUNED;
    else
        P.mode = (t == d - 1) ? FFT_TO_FHAT : FFT_OUT_PRUNED;
    pass_shape<T>(out);
    return out;
}

template <typename T>
int fft_axis_setup(const Geom &ge, const cplx_t<T> *const *tw, const double *roc,
                   const int64_t *rocoff, int t, bool forward, const cplx_t<T> *fhat_in,
                   cplx_t<T> *fhat_out, cplx_t<T> **g, cplx_t<T> **g_other, FftLaunch<T> *out)
{
    const int d = ge.d;
    if (fft_line_ok(ge.n[t])) {
        out[0] = fft_pass_setup<T>(ge, tw[0], roc, rocoff, t, forward, fhat_in, fhat_out, *g);
        return 1;
    }
    int64_t L1, L2;
    if (!fft_split(ge.n[t], L1, L2)) throw Error(CUNFFT_ERR_UNSUPPORTED, "no FFT for this axis length");
    if (!*g_other) throw Error(CUNFFT_ERR_STATE, "a split FFT axis needs the second grid buffer");
    // A: the L2-point FFTs over i2 (stride L1 S), in place, with the input
    // pruning (and the roll-off from fhat) of the axis
    FftLaunch<T> &A = out[0];
    A.P = pass_common<T>(ge, roc, rocoff, t, forward, fhat_in, fhat_out, *g);
    A.P.split = SPLIT_A;
    A.P.tw = tw[1];
    A.P.L = (int)L2;
    A.P.S = L1 * A.P.So;
    A.P.L1 = L1;
    A.P.L2 = L2;
    A.P.mode = forward ? ((t == d - 1) ? FFT_FROM_FHAT : FFT_IN_PRUNED) : 0;
    pass_shape<T>(A);
    // B: per k2 the L1-point FFT over i1 with the w_Lf twiddles, to the other
    // grid in natural order (or to fhat with the crop and roll-off)
    FftLaunch<T> &B = out[1];
    B.P = pass_common<T>(ge, roc, rocoff, t, forward, fhat_in, fhat_out, *g);
    B.P.split = SPLIT_B;
    B.P.tw = tw[0];
    B.P.L = (int)L1;
    B.P.S = B.P.So;
    B.P.L1 = L1;
    B.P.L2 = L2;
    B.P.n_pa *= L2;
    B.P.mode = forward ? 0 : ((t == d - 1) ? FFT_TO_FHAT : FFT_OUT_PRUNED);
    B.P.dst = B.P.mode == FFT_TO_FHAT ? nullptr : *g_other;
    pass_shape<T>(B);
    if (B.P.mode != FFT_TO_FHAT) std::swap(*g, *g_other);
    return 2;
}

template void launch_fft_pass<float>(const FftPass<float> &, int64_t, size_t, hipStream_t);
template void launch_fft_pass<double>(const FftPass<double> &, int64_t, size_t, hipStream_t);
template void launch_fft_batch<float>(const FftPass<float> *, const BlockMap *, uint32_t, size_t,
                                      int, hipStream_t);
template void launch_fft_batch<double>(const FftPass<double> *, const BlockMap *, uint32_t, size_t,
                                       int, hipStream_t);
template FftLaunch<float> fft_pass_setup<float>(const Geom &, const float2 *, const double *,
                                                const int64_t *, int, bool, const float2 *,
                                                float2 *, float2 *);
template FftLaunch<double> fft_pass_setup<double>(const Geom &, const double2 *, const double *,
                                                  const int64_t *, int, bool, const double2 *,
                                                  double2 *, double2 *);
template int fft_axis_setup<float>(const Geom &, const float2 *const *, const double *,
                                   const int64_t *, int, bool, const float2 *, float2 *, float2 **,
                                   float2 **, FftLaunch<float> *);
template int fft_axis_setup<double>(const Geom &, const double2 *const *, const double *,
                                    const int64_t *, int, bool, const double2 *, double2 *,
                                    double2 **, double2 **, FftLaunch<double> *);

}  // namespace cunfft
