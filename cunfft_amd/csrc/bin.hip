// bin.hip -- node binning and the tiled gridding kernels for gfx950.
//
// The reference grids one node per thread straight from global memory: the
// forward through a texture copy of the grid (src/cunfft_kernel.cu:514-1093,
// 1408-1418), the adjoint with two CAS-loop atomics per tap into global memory
// (src/cunfft_kernel.cu:97-130, 608-1197).  Here nodes are first sorted on the
// device by the grid tile holding their window start (counting sort: keys,
// histogram, scan, placement), and each workgroup then owns one tile:
//   forward  the padded tile (T_t + 2m + 1 cells per axis) is staged in LDS
//            once, and every node of the tile reads its (2m+2)^d taps there;
//   adjoint  the node contributions accumulate in an LDS copy of the padded
//            tile with LDS float/double atomics, and the tile is added to the
//            global grid once, with coalesced global atomics, skipping cells
//            no node reached.
// The weights, the window start and the FP64 index/fraction split are the ones
// of the node-parallel kernels, so both paths compute the same taps.
#include <hip/amd_detail/amd_hip_unsafe_atomics.h>

#include "bin.hpp"

namespace cunfft {

constexpr int kTileThreads = 256;
constexpr int kChunk = 256;
constexpr int kMaxTileBytes = 120 * 1024;

template <int D> struct TileShape;
template <> struct TileShape<1> { static constexpr int T[3] = {256, 1, 1}; };
template <> struct TileShape<2> { static constexpr int T[3] = {32, 32, 1}; };
template <> struct TileShape<3> { static constexpr int T[3] = {8, 8, 8}; };

template <int D, int W> struct Pad {
    static constexpr int P0 = TileShape<D>::T[0] + W - 1;
    static constexpr int P1 = D >= 2 ? TileShape<D>::T[1] + W - 1 : 1;
    static constexpr int P2 = D >= 3 ? TileShape<D>::T[2] + W - 1 : 1;
    static constexpr int PT = P0 * P1 * P2;
};

template <typename T, int D, int W> constexpr bool tile_fits()
{
    return (size_t)Pad<D, W>::PT * 2 * sizeof(T) <= (size_t)kMaxTileBytes;
}

struct BinGeom {
    int d;
    int logT[3];
    int64_t nt[3];
    int64_t ntiles;
};

static BinGeom bin_geom(const Binning &b)
{
    BinGeom g{};
    g.d = b.d;
    for (int t = 0; t < 3; ++t) {
        int l = 0;
        while ((int64_t(1) << l) < b.T[t]) ++l;
        g.logT[t] = l;
        g.nt[t] = b.nt[t];
    }
    g.ntiles = b.ntiles;
    return g;
}

// ----------------------------------------------------------- binning ----
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_bin_keys(const double *__restrict__ x,
                                                  uint32_t *__restrict__ keys,
                                                  uint32_t *__restrict__ tile_count, Geom ge,
                                                  BinGeom bg)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= ge.M) return;
    uint32_t key = 0;
    for (int t = 0; t < bg.d; ++t) {
        int64_t u;
        double c;
        window_start(x[j * bg.d + t], ge.n[t], ge.m, u, c);
        const uint32_t tc = (uint32_t)((u & (ge.n[t] - 1)) >> bg.logT[t]);
        key = key * (uint32_t)bg.nt[t] + tc;
    }
    keys[j] = key;
    atomicAdd(&tile_count[key], 1u);
}

// One workgroup: exclusive scans of the tile counts and of the chunk counts.
__global__ __launch_bounds__(1024) void k_bin_scan(const uint32_t *__restrict__ cnt,
                                                   uint32_t *__restrict__ start,
                                                   uint32_t *__restrict__ cursor,
                                                   uint32_t *__restrict__ chunk_start,
                                                   uint32_t *__restrict__ n_chunks, int64_t ntiles)
{
    __shared__ uint32_t sa[1024], sb[1024];
    const int tid = threadIdx.x;
    const int64_t per = (ntiles + 1023) / 1024;
    const int64_t lo = tid * per, hi = lo + per < ntiles ? lo + per : ntiles;
    uint32_t a = 0, b = 0;
    for (int64_t i = lo; i < hi; ++i) {
        a += cnt[i];
        b += (cnt[i] + kChunk - 1) / kChunk;
    }
    sa[tid] = a;
    sb[tid] = b;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        uint32_t va = 0, vb = 0;
        if (tid >= off) {
            va = sa[tid - off];
            vb = sb[tid - off];
        }
        __syncthreads();
        sa[tid] += va;
        sb[tid] += vb;
        __syncthreads();
    }
    uint32_t ra = sa[tid] - a, rb = sb[tid] - b;
    for (int64_t i = lo; i < hi; ++i) {
        const uint32_t c = cnt[i];
        start[i] = ra;
        chunk_start[i] = rb;
        cursor[i] = 0;
        ra += c;
        rb += (c + kChunk - 1) / kChunk;
    }
    if (tid == 1023) n_chunks[0] = sb[1023];
}

__global__ __launch_bounds__(256) void k_bin_place(const double *__restrict__ x,
                                                   const uint32_t *__restrict__ keys,
                                                   const uint32_t *__restrict__ start,
                                                   uint32_t *__restrict__ cursor,
                                                   uint32_t *__restrict__ perm,
                                                   double *__restrict__ xs, int64_t M, int d)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= M) return;
    const uint32_t k = keys[j];
    const uint32_t pos = start[k] + atomicAdd(&cursor[k], 1u);
    perm[pos] = (uint32_t)j;
    for (int t = 0; t < d; ++t) xs[(int64_t)pos * d + t] = x[j * d + t];
}

__global__ __launch_bounds__(256) void k_bin_chunks(const uint32_t *__restrict__ cnt,
                                                    const uint32_t *__restrict__ start,
                                                    const uint32_t *__restrict__ chunk_start,
                                                    Chunk *__restrict__ chunks, int64_t ntiles)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= ntiles) return;
    const uint32_t c = cnt[i];
    uint32_t o = chunk_start[i];
    for (uint32_t s = 0; s < c; s += kChunk, ++o) {
        Chunk ch;
        ch.tile = (uint32_t)i;
        ch.start = start[i] + s;
        ch.count = c - s < (uint32_t)kChunk ? c - s : (uint32_t)kChunk;
        ch.pad = 0;
        chunks[o] = ch;
    }
}

bool Binning::plan(const Geom &ge, int elem_bytes, int kernel_variant)
{
    variant = kernel_variant;
    slot_elem_bytes = elem_bytes;
    d = ge.d;
    usable = 0;
    const int W = ge.W;
    if (W > 18 || ge.M >= (int64_t(1) << 31)) return false;
    static const int shapes[4][3] = {{1, 1, 1}, {256, 1, 1}, {32, 32, 1}, {8, 8, 8}};
    int64_t bytes = 2 * elem_bytes;
    ntiles = 1;
    for (int t = 0; t < 3; ++t) {
        T[t] = t < d ? shapes[d][t] : 1;
        if (t < d && ge.n[t] < T[t]) return false;
        nt[t] = t < d ? ge.n[t] / T[t] : 1;
        ntiles *= nt[t];
        if (t < d) bytes *= T[t] + W - 1;
    }
    if (bytes > kMaxTileBytes) return false;
    if (d == 3 && W > 14) return false;
    if (ntiles >= (int64_t(1) << 31)) return false;
    usable = 1;
    return true;
}

static void *dev_alloc(size_t bytes)
{
    void *p = nullptr;
    CUNFFT_HIP(hipMalloc(&p, bytes ? bytes : 16));
    return p;
}

void Binning::build(const Geom &ge, const double *x, hipStream_t s)
{
    if (!usable) return;
    if (!built || M != ge.M) {
        release();
        M = ge.M;
        max_chunks = ntiles + (M + kChunk - 1) / kChunk;
        keys = (uint32_t *)dev_alloc(sizeof(uint32_t) * M);
        perm = (uint32_t *)dev_alloc(sizeof(uint32_t) * M);
        tile_count = (uint32_t *)dev_alloc(sizeof(uint32_t) * ntiles);
        tile_start = (uint32_t *)dev_alloc(sizeof(uint32_t) * ntiles);
        cursor = (uint32_t *)dev_alloc(sizeof(uint32_t) * ntiles);
        chunk_start = (uint32_t *)dev_alloc(sizeof(uint32_t) * ntiles);
        n_chunks = (uint32_t *)dev_alloc(sizeof(uint32_t));
        chunks = (Chunk *)dev_alloc(sizeof(Chunk) * max_chunks);
        xs = (double *)dev_alloc(sizeof(double) * M * d);
        if (d >= 2 && variant == 0) {
            int64_t pt = 1;
            for (int t = 0; t < d; ++t) pt *= T[t] + ge.W - 1;
            slots = dev_alloc((size_t)max_chunks * pt * 2 * slot_elem_bytes);
        }
        M = ge.M;
    }
    const BinGeom bg = bin_geom(*this);
    k_fill_u32<<<(unsigned)((ntiles + 255) / 256), 256, 0, s>>>(tile_count, ntiles, 0u);
    k_bin_keys<<<(unsigned)((M + 255) / 256), 256, 0, s>>>(x, keys, tile_count, ge, bg);
    k_bin_scan<<<1, 1024, 0, s>>>(tile_count, tile_start, cursor, chunk_start, n_chunks, ntiles);
    k_bin_place<<<(unsigned)((M + 255) / 256), 256, 0, s>>>(x, keys, tile_start, cursor, perm, xs,
                                                            M, d);
    k_bin_chunks<<<(unsigned)((ntiles + 255) / 256), 256, 0, s>>>(tile_count, tile_start,
                                                                  chunk_start, chunks, ntiles);
    CUNFFT_HIP(hipGetLastError());
    built = 1;
}

void Binning::release()
{
    if (keys) (void)hipFree(keys);
    if (perm) (void)hipFree(perm);
    if (tile_count) (void)hipFree(tile_count);
    if (tile_start) (void)hipFree(tile_start);
    if (cursor) (void)hipFree(cursor);
    if (chunk_start) (void)hipFree(chunk_start);
    if (n_chunks) (void)hipFree(n_chunks);
    if (chunks) (void)hipFree(chunks);
    if (xs) (void)hipFree(xs);
    if (slots) (void)hipFree(slots);
    keys = perm = tile_count = tile_start = cursor = chunk_start = n_chunks = nullptr;
    chunks = nullptr;
    xs = nullptr;
    slots = nullptr;
    built = 0;
    M = 0;
}

// ------------------------------------------------------ tile helpers ----
template <int D>
__device__ __forceinline__ void tile_origin(uint32_t tile, const BinGeom &bg, int64_t (&o)[3])
{
    uint32_t rem = tile;
#pragma unroll
    for (int t = D - 1; t >= 0; --t) {
        const uint32_t tc = rem % (uint32_t)bg.nt[t];
        rem /= (uint32_t)bg.nt[t];
        o[t] = (int64_t)tc << bg.logT[t];
    }
#pragma unroll
    for (int t = D; t < 3; ++t) o[t] = 0;
}

// Global grid offset of padded-tile cell idx.
template <int D, int W>
__device__ __forceinline__ int64_t tile_cell(int idx, const int64_t (&o)[3], const Geom &ge)
{
    using PD = Pad<D, W>;
    const int p0 = idx / (PD::P1 * PD::P2);
    const int r = idx - p0 * (PD::P1 * PD::P2);
    const int p1 = r / PD::P2;
    const int p2 = r - p1 * PD::P2;
    const int64_t c0 = (o[0] + p0) & (ge.n[0] - 1);
    if constexpr (D == 1) return c0;
    const int64_t c1 = (o[1] + p1) & (ge.n[1] - 1);
    if constexpr (D == 2) return c0 * ge.n[1] + c1;
    const int64_t c2 = (o[2] + p2) & (ge.n[2] - 1);
    return (c0 * ge.n[1] + c1) * ge.n[2] + c2;
}

// Window weights of the node at sorted position s and its LDS base offset.
template <typename T, int D, int W>
__device__ __forceinline__ int node_local(const double *__restrict__ xs, uint32_t s,
                                          const int64_t (&o)[3], const Geom &ge, T (&w)[D][W])
{
    using PD = Pad<D, W>;
    constexpr int stride[3] = {PD::P1 * PD::P2, PD::P2, 1};
    int base = 0;
#pragma unroll
    for (int t = 0; t < D; ++t) {
        int64_t u;
        double c;
        window_start(xs[(int64_t)s * D + t], ge.n[t], ge.m, u, c);
        int lo = (int)((u & (ge.n[t] - 1)) - o[t]);
        lo = lo < 0 ? 0 : (lo >= TileShape<D>::T[t] ? TileShape<D>::T[t] - 1 : lo);
        window_weights<T, W>(c, ge.invb[t], ge.nrm[t], w[t]);
        base += lo * stride[t];
    }
    return base;
}

// --------------------------------------------------------- tiled gather --
template <typename T, int D, int W>
__global__ __launch_bounds__(kTileThreads) void k_gather_tiled(
    const cplx_t<T> *__restrict__ g, const double *__restrict__ xs, const uint32_t *__restrict__ perm,
    const Chunk *__restrict__ chunks, const uint32_t *__restrict__ n_chunks, cplx_t<T> *__restrict__ f,
    Geom ge, BinGeom bg)
{
    using PD = Pad<D, W>;
    __shared__ cplx_t<T> tile[PD::PT];
    if (blockIdx.x >= n_chunks[0]) return;
    const Chunk ch = chunks[blockIdx.x];
    int64_t o[3];
    tile_origin<D>(ch.tile, bg, o);
    for (int idx = threadIdx.x; idx < PD::PT; idx += kTileThreads)
        tile[idx] = g[tile_cell<D, W>(idx, o, ge)];
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ch.count; i += kTileThreads) {
        const uint32_t s = ch.start + i;
        T w[D][W];
        const int base = node_local<T, D, W>(xs, s, o, ge, w);
        const cplx_t<T> *tb = tile + base;
        T sr = 0, si = 0;
        if constexpr (D == 1) {
#pragma unroll
            for (int a = 0; a < W; ++a) {
                const cplx_t<T> v = tb[a];
                sr += v.x * w[0][a];
                si += v.y * w[0][a];
            }
        } else if constexpr (D == 2) {
#pragma unroll
            for (int a = 0; a < W; ++a) {
                T r1 = 0, i1 = 0;
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    const cplx_t<T> v = tb[a * PD::P1 + c];
                    r1 += v.x * w[1][c];
                    i1 += v.y * w[1][c];
                }
                sr += r1 * w[0][a];
                si += i1 * w[0][a];
            }
        } else {
#pragma unroll
            for (int a = 0; a < W; ++a) {
                T r1 = 0, i1 = 0;
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    T r2 = 0, i2 = 0;
#pragma unroll
                    for (int e = 0; e < W; ++e) {
                        const cplx_t<T> v = tb[(a * PD::P1 + c) * PD::P2 + e];
                        r2 += v.x * w[2][e];
                        i2 += v.y * w[2][e];
                    }
                    r1 += r2 * w[1][c];
                    i1 += i2 * w[1][c];
                }
                sr += r1 * w[0][a];
                si += i1 * w[0][a];
            }
        }
        f[perm[s]] = cmake<T>(sr, si);
    }
}

// -------------------------------------------------------- tiled scatter --
template <typename T>
__device__ __forceinline__ void lds_add(T *p, T v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T, typename A, int D, int W>
__global__ __launch_bounds__(kTileThreads) void k_scatter_tiled(
    const double *__restrict__ xs, const uint32_t *__restrict__ perm, const cplx_t<T> *__restrict__ f,
    const Chunk *__restrict__ chunks, const uint32_t *__restrict__ n_chunks, cplx_t<T> *__restrict__ g,
    Geom ge, BinGeom bg)
{
    using PD = Pad<D, W>;
    __shared__ A acc_re[PD::PT];
    __shared__ A acc_im[PD::PT];
    if (blockIdx.x >= n_chunks[0]) return;
    const Chunk ch = chunks[blockIdx.x];
    int64_t o[3];
    tile_origin<D>(ch.tile, bg, o);
    for (int idx = threadIdx.x; idx < PD::PT; idx += kTileThreads) {
        acc_re[idx] = 0;
        acc_im[idx] = 0;
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < ch.count; i += kTileThreads) {
        const uint32_t s = ch.start + i;
        T w[D][W];
        const int base = node_local<T, D, W>(xs, s, o, ge, w);
        const cplx_t<T> fj = f[perm[s]];
        A *re = acc_re + base;
        A *im = acc_im + base;
        if constexpr (D == 1) {
#pragma unroll
            for (int a = 0; a < W; ++a) {
                lds_add<A>(re + a, (A)(fj.x * w[0][a]));
                lds_add<A>(im + a, (A)(fj.y * w[0][a]));
            }
        } else if constexpr (D == 2) {
#pragma unroll
            for (int a = 0; a < W; ++a) {
                const T fr = fj.x * w[0][a], fi = fj.y * w[0][a];
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    lds_add<A>(re + a * PD::P1 + c, (A)(fr * w[1][c]));
                    lds_add<A>(im + a * PD::P1 + c, (A)(fi * w[1][c]));
                }
            }
        } else {
#pragma unroll
            for (int a = 0; a < W; ++a) {
                const T fr0 = fj.x * w[0][a], fi0 = fj.y * w[0][a];
#pragma unroll
                for (int c = 0; c < W; ++c) {
                    const T fr = fr0 * w[1][c], fi = fi0 * w[1][c];
#pragma unroll
                    for (int e = 0; e < W; ++e) {
                        const int off = (a * PD::P1 + c) * PD::P2 + e;
                        lds_add<A>(re + off, (A)(fr * w[2][e]));
                        lds_add<A>(im + off, (A)(fi * w[2][e]));
                    }
                }
            }
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < PD::PT; idx += kTileThreads) {
        const A vr = acc_re[idx], vi = acc_im[idx];
        if (vr != (A)0 || vi != (A)0) {
            cplx_t<T> *p = g + tile_cell<D, W>(idx, o, ge);
            unsafeAtomicAdd(&p->x, (T)vr);
            unsafeAtomicAdd(&p->y, (T)vi);
        }
    }
}

// ------------------------------------------------- row-lane tiled gather --
// W lanes per node, one lane per tap along the last axis, floor(64/W) nodes per
// wavefront.  Each LDS read of a node touches W consecutive cells, so a
// wavefront instruction hits a few runs of consecutive addresses instead of 64
// scattered cells.  Lane e of a node evaluates only the e-th window weight of
// every axis and the node's lanes share them through LDS, so each lane pays d
// exponentials per node instead of d*W.  The node's W lane partials are summed
// with a segmented reduction through LDS.
template <int W> struct Rows {
    static constexpr int NPW = 64 / W;
};

// Compiler-only barrier: orders LDS accesses of one wavefront in program
// order (the LDS executes a wavefront's instructions in order).
__device__ __forceinline__ void wave_order() { __asm__ volatile("" ::: "memory"); }

template <typename T, int D, int W>
__global__ __launch_bounds__(kTileThreads) void k_gather_rows(
    const cplx_t<T> *__restrict__ g, const double *__restrict__ xs, const uint32_t *__restrict__ perm,
    const Chunk *__restrict__ chunks, const uint32_t *__restrict__ n_chunks, cplx_t<T> *__restrict__ f,
    Geom ge, BinGeom bg)
{
    static_assert(D >= 2, "row kernels are for d >= 2");
    using PD = Pad<D, W>;
    constexpr int NPW = Rows<W>::NPW;
    constexpr int stride[3] = {PD::P1 * PD::P2, PD::P2, 1};
    __shared__ cplx_t<T> tile[PD::PT];
    __shared__ cplx_t<T> red[kTileThreads];
    __shared__ T wsh[kTileThreads / 64][NPW][D][W];
    if (blockIdx.x >= n_chunks[0]) return;
    const Chunk ch = chunks[blockIdx.x];
    int64_t o[3];
    tile_origin<D>(ch.tile, bg, o);
    for (int idx = threadIdx.x; idx < PD::PT; idx += kTileThreads)
        tile[idx] = g[tile_cell<D, W>(idx, o, ge)];
    __syncthreads();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int slot = lane / W, e = lane - slot * W;
    const int ngroups = ((int)ch.count + NPW - 1) / NPW;
    for (int gi = wave; gi < ngroups; gi += kTileThreads / 64) {
        const int i = gi * NPW + slot;
        const bool valid = slot < NPW && i < (int)ch.count;
        int base = e;
        if (valid) {
#pragma unroll
            for (int t = 0; t < D; ++t) {
                int64_t u;
                double c;
                window_start(xs[(int64_t)(ch.start + i) * D + t], ge.n[t], ge.m, u, c);
                int lo = (int)((u & (ge.n[t] - 1)) - o[t]);
                lo = lo < 0 ? 0 : (lo >= TileShape<D>::T[t] ? TileShape<D>::T[t] - 1 : lo);
                base += lo * stride[t];
                const T sx = (T)(c - (double)e);
                wsh[wave][slot][t][e] = win_exp<T>(-sx * sx * (T)ge.invb[t]) * (T)ge.nrm[t];
            }
        }
        wave_order();
        T sr = 0, si = 0;
        if (valid) {
            const T *ws = &wsh[wave][slot][0][0];
            T w0[W];
#pragma unroll
            for (int a = 0; a < W; ++a) w0[a] = ws[a];
            const T we = ws[(D - 1) * W + e];
            const cplx_t<T> *tb = tile + base;
            if constexpr (D == 2) {
#pragma unroll
                for (int a = 0; a < W; ++a) {
                    const cplx_t<T> v = tb[a * PD::P1];
                    sr += v.x * w0[a];
                    si += v.y * w0[a];
                }
            } else {
                T w1[W];
#pragma unroll
                for (int c = 0; c < W; ++c) w1[c] = ws[W + c];
#pragma unroll
                for (int a = 0; a < W; ++a) {
                    T r1 = 0, i1 = 0;
#pragma unroll
                    for (int c = 0; c < W; ++c) {
                        const cplx_t<T> v = tb[(a * PD::P1 + c) * PD::P2];
                        r1 += v.x * w1[c];
                        i1 += v.y * w1[c];
                    }
                    sr += r1 * w0[a];
                    si += i1 * w0[a];
                }
            }
            sr *= we;
            si *= we;
        }
        red[threadIdx.x] = cmake<T>(sr, si);
        wave_order();
        if (valid && e == 0) {
            const cplx_t<T> *rp = red + threadIdx.x;
            T tr = 0, ti = 0;
#pragma unroll
            for (int k = 0; k < W; ++k) {
                tr += rp[k].x;
                ti += rp[k].y;
            }
            f[perm[ch.start + i]] = cmake<T>(tr, ti);
        }
        wave_order();
    }
}

// ------------------------------------------------ layered row scatter ----
// Four wavefronts share one tile chunk.  W lanes per node run along the last
// axis and floor(64/W) nodes form one step; accumulation is a plain LDS
// read-add-write.  Conflicts are excluded by construction:
//  * wavefront w updates only planes p0 = w (mod 4) of the padded tile, so two
//    wavefronts never touch the same cell;
//  * a step holds nodes of one class (window start lo0 mod 4), so the planes a
//    wavefront updates are the same for every lane of the step;
//  * a step holds nodes with distinct row keys (window starts on the leading
//    axes), the only way two lanes of one instruction could hit one cell: the
//    k-th node of each key goes to layer k and a step takes one layer;
//  * lanes of a step whose window starts differ meet the same cell at
//    different (plane, row) iterations, so the iterations are kept in program
//    order (a compiler barrier; the LDS runs a wavefront's instructions in
//    order), and each iteration is a read-add-write of one row per lane.
// Steps of a wavefront execute in order in the LDS, so the tile sums do not
// depend on scheduling.  Node descriptors are built 64 at a time in parallel.
template <typename T, int D, int W>
__global__ __launch_bounds__(kTileThreads) void k_scatter_layers(
    const double *__restrict__ xs, const uint32_t *__restrict__ perm, const cplx_t<T> *__restrict__ f,
    const Chunk *__restrict__ chunks, const uint32_t *__restrict__ n_chunks,
    cplx_t<T> *__restrict__ slots, Geom ge, BinGeom bg)
{
    static_assert(D >= 2, "layered kernels are for d >= 2");
    using PD = Pad<D, W>;
    constexpr int NPW = Rows<W>::NPW;
    constexpr int NW = kTileThreads / 64;       // wavefronts, = plane classes
    constexpr int stride[3] = {PD::P1 * PD::P2, PD::P2, 1};
    constexpr int NKEY = TileShape<D>::T[0] * (D == 3 ? TileShape<D>::T[1] : 1);
    constexpr int NGRP = NW * 64;               // (class, layer) groups
    static_assert(NGRP == kTileThreads, "one scan element per thread");
    constexpr int MAXORD = 64 + 64 * (NPW - 1); // padded group list of one batch
    constexpr int NQ = (W + NW - 1) / NW;       // planes per wavefront per node
    __shared__ cplx_t<T> acc[PD::PT];
    __shared__ int dbase[64], dlo0[64];
    __shared__ cplx_t<T> dcoef[64];
    __shared__ T dw[64][D][W];
    __shared__ int key_cnt[NKEY];
    __shared__ int grp_cnt[NGRP], grp_fill[NGRP], grp_start[NGRP + 1];
    __shared__ int order[MAXORD];
    if (blockIdx.x >= n_chunks[0]) return;
    const Chunk ch = chunks[blockIdx.x];
    int64_t o[3];
    tile_origin<D>(ch.tile, bg, o);
    for (int idx = threadIdx.x; idx < PD::PT; idx += kTileThreads) acc[idx] = cmake<T>(0, 0);
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int slot = lane / W, e = lane - slot * W;
    for (uint32_t b0 = 0; b0 < ch.count; b0 += 64) {
        const int nb = (int)(ch.count - b0 < 64u ? ch.count - b0 : 64u);
        for (int k = tid; k < NKEY; k += kTileThreads) key_cnt[k] = 0;
        for (int k = tid; k < NGRP; k += kTileThreads) {
            grp_cnt[k] = 0;
            grp_fill[k] = 0;
        }
        __syncthreads();
        int my_grp = 0;
        if (tid < nb) {
            const uint32_t s = ch.start + b0 + tid;
            int base = 0, key = 0, lo0 = 0;
#pragma unroll
            for (int t = 0; t < D; ++t) {
                int64_t u;
                double c;
                window_start(xs[(int64_t)s * D + t], ge.n[t], ge.m, u, c);
                int lo = (int)((u & (ge.n[t] - 1)) - o[t]);
                lo = lo < 0 ? 0 : (lo >= TileShape<D>::T[t] ? TileShape<D>::T[t] - 1 : lo);
                base += lo * stride[t];
                if (t == 0) lo0 = lo;
                if (t < D - 1) key = key * TileShape<D>::T[t] + lo;
                T wt[W];
                window_weights<T, W>(c, ge.invb[t], ge.nrm[t], wt);
#pragma unroll
                for (int l = 0; l < W; ++l) dw[tid][t][l] = wt[l];
            }
            dbase[tid] = base;
            dlo0[tid] = lo0;
            dcoef[tid] = f[perm[s]];
            const int layer = atomicAdd(&key_cnt[key], 1);
            my_grp = (lo0 & (NW - 1)) * 64 + layer;
            atomicAdd(&grp_cnt[my_grp], 1);
        }
        __syncthreads();
        {  // exclusive scan of the padded group sizes (NGRP == kTileThreads)
            const int mine = (grp_cnt[tid] + NPW - 1) / NPW * NPW;
            grp_start[tid + 1] = mine;
            __syncthreads();
            for (int off = 1; off < NGRP; off <<= 1) {
                const int v = tid + 1 > off ? grp_start[tid + 1 - off] : 0;
                __syncthreads();
                grp_start[tid + 1] += v;
                __syncthreads();
            }
            if (tid == 0) grp_start[0] = 0;
            __syncthreads();
        }
        const int total = grp_start[NGRP];
        for (int k = tid; k < total; k += kTileThreads) order[k] = -1;
        __syncthreads();
        if (tid < nb) order[grp_start[my_grp] + atomicAdd(&grp_fill[my_grp], 1)] = tid;
        __syncthreads();
        for (int st = 0; st < total; st += NPW) {
            const int node = slot < NPW ? order[st + slot] : -1;
            if (node >= 0) {
                const T *wn = &dw[node][0][0];
                const cplx_t<T> cf = dcoef[node];
                const T we = wn[(D - 1) * W + e];
                const T fr = cf.x * we, fi = cf.y * we;
                const int a0 = (wave - dlo0[node]) & (NW - 1);  // first plane of this wavefront
                cplx_t<T> *ab = acc + dbase[node] + e + a0 * stride[0];
                if constexpr (D == 2) {
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int a = a0 + q * NW;
                        if (a < W) {
                            const T wa = wn[a];
                            cplx_t<T> *p = ab + q * NW * stride[0];
                            cplx_t<T> v = *p;
                            v.x += fr * wa;
                            v.y += fi * wa;
                            *p = v;
                        }
                        // Another lane of this step may touch this plane at the next q
                        // (window starts 4 planes apart): keep q+1's loads after q's stores.
                        wave_order();
                    }
                } else {
                    T w1[W];
#pragma unroll
                    for (int c = 0; c < W; ++c) w1[c] = wn[W + c];
#pragma unroll
                    for (int q = 0; q < NQ; ++q) {
                        const int a = a0 + q * NW;
                        if (a < W) {
                            const T wa = wn[a];
                            const T fra = fr * wa, fia = fi * wa;
                            cplx_t<T> *pa = ab + q * NW * stride[0];
#pragma unroll
                            for (int c = 0; c < W; ++c) {
                                cplx_t<T> v = pa[c * PD::P2];
                                v.x += fra * w1[c];
                                v.y += fia * w1[c];
                                pa[c * PD::P2] = v;
                                // lanes whose row starts differ meet this row at another c
                                wave_order();
                            }
                        }
                        // Another lane of this step may touch this plane at the next q
                        // (window starts 4 planes apart): keep q+1's loads after q's stores.
                        wave_order();
                    }
                }
            }
        }
        __syncthreads();
    }
    cplx_t<T> *slotp = slots + (int64_t)blockIdx.x * PD::PT;
    for (int idx = threadIdx.x; idx < PD::PT; idx += kTileThreads) slotp[idx] = acc[idx];
}

// Fold: every grid cell is the sum, over the chunks of the tiles whose padded
// region covers it, of that chunk's slot value.  One workgroup per tile core:
// the candidate tiles and their chunk lists are the same for the whole
// workgroup, each thread owns a few core cells, and a core row reads a
// contiguous run of every slot.  Each cell is written once, with a plain
// store, in a fixed order: no atomics, no zero fill, and the adjoint grid is
// bit-reproducible.
template <typename T, int D, int W>
__global__ __launch_bounds__(256) void k_fold(const cplx_t<T> *__restrict__ slots,
                                              const uint32_t *__restrict__ tile_count,
                                              const uint32_t *__restrict__ chunk_start,
                                              cplx_t<T> *__restrict__ g, Geom ge, BinGeom bg)
{
    using PD = Pad<D, W>;
    constexpr int TT[3] = {TileShape<D>::T[0], TileShape<D>::T[1], TileShape<D>::T[2]};
    constexpr int CORE = TT[0] * TT[1] * TT[2];
    constexpr int CPT = (CORE + 255) / 256;            // core cells per thread
    constexpr int P[3] = {PD::P0, PD::P1, PD::P2};
    constexpr int stride[3] = {PD::P1 * PD::P2, PD::P2, 1};
    constexpr int Q[3] = {(P[0] - 1) / TT[0] + 1, D >= 2 ? (P[1] - 1) / TT[1] + 1 : 1,
                          D >= 3 ? (P[2] - 1) / TT[2] + 1 : 1};
    const uint32_t tile = blockIdx.x;
    int64_t o[3];
    uint32_t tk[3];
    {
        uint32_t rem = tile;
        for (int t = 2; t >= 0; --t) {
            const uint32_t ntt = t < D ? (uint32_t)bg.nt[t] : 1u;
            tk[t] = rem % ntt;
            rem /= ntt;
            o[t] = (int64_t)tk[t] * TT[t];
        }
    }
    int r[CPT][3];
    bool live[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) {
        const int idx = threadIdx.x + k * 256;
        live[k] = idx < CORE;
        const int id = live[k] ? idx : 0;
        r[k][0] = id / (TT[1] * TT[2]);
        r[k][1] = (id / TT[2]) % TT[1];
        r[k][2] = id % TT[2];
    }
    T sr[CPT], si[CPT];
#pragma unroll
    for (int k = 0; k < CPT; ++k) sr[k] = si[k] = 0;
    for (int q0 = 0; q0 < Q[0]; ++q0)
        for (int q1 = 0; q1 < Q[1]; ++q1)
            for (int q2 = 0; q2 < Q[2]; ++q2) {
                const int q[3] = {q0, q1, q2};
                uint32_t ct = 0;
                for (int t = 0; t < D; ++t) {
                    const uint32_t ntt = (uint32_t)bg.nt[t];
                    ct = ct * ntt + (tk[t] + ntt * 4 - (uint32_t)q[t]) % ntt;
                }
                const uint32_t cnt = tile_count[ct];
                if (cnt == 0) continue;
                const uint32_t c0 = chunk_start[ct], nch = (cnt + kChunk - 1) / kChunk;
                int li[CPT];
                bool ok[CPT];
#pragma unroll
                for (int k = 0; k < CPT; ++k) {
                    ok[k] = live[k];
                    li[k] = 0;
#pragma unroll
                    for (int t = 0; t < D; ++t) {
                        const int p = r[k][t] + q[t] * TT[t];
                        ok[k] = ok[k] && p < P[t];
                        li[k] += p * stride[t];
                    }
                }
                for (uint32_t c = 0; c < nch; ++c) {
                    const cplx_t<T> *sp = slots + (int64_t)(c0 + c) * PD::PT;
#pragma unroll
                    for (int k = 0; k < CPT; ++k)
                        if (ok[k]) {
                            const cplx_t<T> v = sp[li[k]];
                            sr[k] += v.x;
                            si[k] += v.y;
                        }
                }
            }
#pragma unroll
    for (int k = 0; k < CPT; ++k)
        if (live[k]) {
            int64_t cell = 0;
            for (int t = 0; t < D; ++t) cell = cell * ge.n[t] + o[t] + r[k][t];
            g[cell] = cmake<T>(sr[k], si[k]);
        }
}

// ------------------------------------------------------------ dispatch --
template <typename T, int D, int W>
static void gather_tiled(const Geom &ge, const Binning &b, const cplx_t<T> *g, cplx_t<T> *f,
                         hipStream_t s)
{
    if constexpr (tile_fits<T, D, W>()) {
        if (D >= 2 && b.variant == 0)
            k_gather_rows<T, (D >= 2 ? D : 2), W><<<(unsigned)b.max_chunks, kTileThreads, 0, s>>>(
                g, b.xs, b.perm, b.chunks, b.n_chunks, f, ge, bin_geom(b));
        else
            k_gather_tiled<T, D, W><<<(unsigned)b.max_chunks, kTileThreads, 0, s>>>(
                g, b.xs, b.perm, b.chunks, b.n_chunks, f, ge, bin_geom(b));
    } else {
        throw Error(CUNFFT_ERR_UNSUPPORTED, "tile does not fit in LDS");
    }
}

template <typename T, int D, int W>
static void scatter_tiled(const Geom &ge, const Binning &b, const cplx_t<T> *f, cplx_t<T> *g,
                          hipStream_t s, hipEvent_t mid)
{
    if constexpr (tile_fits<T, D, W>()) {
        if (D >= 2 && b.variant == 0) {
            cplx_t<T> *slots = static_cast<cplx_t<T> *>(b.slots);
            k_scatter_layers<T, (D >= 2 ? D : 2), W><<<(unsigned)b.max_chunks, kTileThreads, 0, s>>>(
                b.xs, b.perm, f, b.chunks, b.n_chunks, slots, ge, bin_geom(b));
            if (mid) CUNFFT_HIP(hipEventRecord(mid, s));
            k_fold<T, D, W><<<(unsigned)b.ntiles, 256, 0, s>>>(slots, b.tile_count, b.chunk_start,
                                                                g, ge, bin_geom(b));
        } else
            k_scatter_tiled<T, double, D, W><<<(unsigned)b.max_chunks, kTileThreads, 0, s>>>(
                b.xs, b.perm, f, b.chunks, b.n_chunks, g, ge, bin_geom(b));
    } else {
        throw Error(CUNFFT_ERR_UNSUPPORTED, "tile does not fit in LDS");
    }
}

#define CUNFFT_TILED_W(FN, D, ...)                \
    switch (ge.W) {                               \
    case 4: FN<T, D, 4>(ge, b, a0, a1, s __VA_ARGS__); break; \
    case 6: FN<T, D, 6>(ge, b, a0, a1, s __VA_ARGS__); break; \
    case 8: FN<T, D, 8>(ge, b, a0, a1, s __VA_ARGS__); break; \
    case 10: FN<T, D, 10>(ge, b, a0, a1, s __VA_ARGS__); break; \
    case 12: FN<T, D, 12>(ge, b, a0, a1, s __VA_ARGS__); break; \
    case 14: FN<T, D, 14>(ge, b, a0, a1, s __VA_ARGS__); break; \
    case 16: if constexpr (D < 3) FN<T, D, 16>(ge, b, a0, a1, s __VA_ARGS__); break; \
    case 18: if constexpr (D < 3) FN<T, D, 18>(ge, b, a0, a1, s __VA_ARGS__); break; \
    default: throw Error(CUNFFT_ERR_UNSUPPORTED, "tiled gridding needs m <= 8"); \
    }

template <typename T>
void launch_gather_binned(const Geom &ge, const Binning &b, const cplx_t<T> *g, const double *x,
                          cplx_t<T> *f, hipStream_t s)
{
    if (!b.usable || !b.built) {
        launch_gather_plain<T>(ge, g, x, f, s);
        return;
    }
    const cplx_t<T> *a0 = g;
    cplx_t<T> *a1 = f;
    switch (ge.d) {
    case 1: CUNFFT_TILED_W(gather_tiled, 1) break;
    case 2: CUNFFT_TILED_W(gather_tiled, 2) break;
    default: CUNFFT_TILED_W(gather_tiled, 3) break;
    }
    CUNFFT_HIP(hipGetLastError());
}

template <typename T>
bool launch_scatter_binned(const Geom &ge, const Binning &b, const double *x, const cplx_t<T> *f,
                           cplx_t<T> *g, hipStream_t s, hipEvent_t mid)
{
    const int64_t nt = ge.n[0] * ge.n[1] * ge.n[2];
    if (!b.usable || !b.built) {
        launch_zero<T>(g, nt, s);
        launch_scatter_atomic<T>(ge, x, f, g, s);
        return false;
    }
    if (!(ge.d >= 2 && b.variant == 0)) launch_zero<T>(g, nt, s);
    const cplx_t<T> *a0 = f;
    cplx_t<T> *a1 = g;
    switch (ge.d) {
    case 1: CUNFFT_TILED_W(scatter_tiled, 1, , mid) break;
    case 2: CUNFFT_TILED_W(scatter_tiled, 2, , mid) break;
    default: CUNFFT_TILED_W(scatter_tiled, 3, , mid) break;
    }
    CUNFFT_HIP(hipGetLastError());
    return ge.d >= 2 && b.variant == 0 && mid != nullptr;
}

template void launch_gather_binned<float>(const Geom &, const Binning &, const float2 *,
                                          const double *, float2 *, hipStream_t);
template void launch_gather_binned<double>(const Geom &, const Binning &, const double2 *,
                                           const double *, double2 *, hipStream_t);
template bool launch_scatter_binned<float>(const Geom &, const Binning &, const double *,
                                           const float2 *, float2 *, hipStream_t, hipEvent_t);
template bool launch_scatter_binned<double>(const Geom &, const Binning &, const double *,
                                            const double2 *, double2 *, hipStream_t, hipEvent_t);

}  // namespace cunfft
