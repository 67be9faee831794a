// bin.hip -- node binning and the node-per-lane tiled gridding kernels.
//
// The reference grids one node per thread straight from global memory: the
// forward through a texture copy of the grid (src/cunfft_kernel.cu:514-1093,
// 1408-1418), the adjoint with two CAS-loop atomics per tap into global memory
// (src/cunfft_kernel.cu:97-130, 608-1197).  Here nodes are first sorted on the
// device by the grid tile holding their window start (a stable radix sort of
// composite keys, then tile counts, scans and chunks), and each workgroup then
// owns one tile.  For d >= 2 the plane-tiled kernels of grid_tile.hpp use the
// sorted order; this file holds the binning itself and the node-per-lane
// tiled kernels used for d = 1 and under CUNFFT_NODE_LANES:
//   forward  the padded tile (T_t + 2m + 1 cells per axis) is staged in LDS
//            once, and every node of the tile reads its (2m+2)^d taps there;
//   adjoint  the node contributions accumulate in an LDS copy of the padded
//            tile with LDS atomics, and the tile is added to the global grid
//            once, with coalesced global atomics, skipping cells no node
//            reached.
// The weights, the window start and the FP64 index/fraction split are the ones
// of the node-parallel kernels, so both paths compute the same taps.
#include <hip/amd_detail/amd_hip_unsafe_atomics.h>

#include <algorithm>

#include "bin.hpp"

namespace cunfft {

constexpr int kTileThreads = 256;
constexpr int kMaxTileBytes = 120 * 1024;

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

BinGeom bin_geom(const Binning &b)
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
    g.sub_shift = b.elem_bytes == 4 ? 1 : 0;
    return g;
}

// ----------------------------------------------------------- binning ----
__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *p, int64_t n, uint32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) p[i] = v;
}

// Composite sort key of every node (its tile times sub_range plus, for the
// plane-tiled kernels, its sub_key within the tile; sub_range 1 otherwise)
// and (plane-tiled kernels) its window descriptor by node index: the FP64
// index/fraction split of window_cell, lo_t the window start relative to the
// origin of the node's tile, in [0, T_t) by construction of the key.  The
// tile counts come from the sorted keys (k_bin_tile_first / _count), not from
// atomics here: a node cluster would serialize them on a few counters.
template <typename T>
__global__ __launch_bounds__(256) void k_bin_keys(const double *__restrict__ x,
                                                  uint32_t *__restrict__ keys,
                                                  NodeDesc<T> *__restrict__ desc_j, Geom ge,
                                                  BinGeom bg, uint32_t sub_range)
{
    const int64_t j = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (j >= ge.M) return;
    uint32_t tile = 0, loc = 0;
    uint32_t lo[3] = {0, 0, 0};
    NodeDesc<T> nd{};
    for (int t = 0; t < bg.d; ++t) {
        double c;
        const int64_t w = window_cell(x[j * bg.d + t], ge.n[t], ge.m, c);
        const uint32_t tc = (uint32_t)(w >> bg.logT[t]);
        lo[t] = (uint32_t)(w - ((int64_t)tc << bg.logT[t]));
        loc |= lo[t] << (8 * t);
        nd.f[t] = (T)(c - (double)ge.m);
        tile = tile * (uint32_t)bg.nt[t] + tc;
    }
    keys[j] = sub_range > 1 ? tile * sub_range + sub_key(bg.d, lo, bg.logT, bg.sub_shift) : tile;
    if (desc_j) {
        nd.loc = loc;
        desc_j[j] = nd;
    }
}

// First sorted position of every nonempty tile, from the sorted composite keys.
__global__ __launch_bounds__(256) void k_bin_tile_first(const uint32_t *__restrict__ skeys, int64_t M,
                                                        uint32_t sub_range, uint32_t *__restrict__ first)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= M) return;
    const uint32_t t = skeys[s] / sub_range;
    if (s == 0 || skeys[s - 1] / sub_range != t) first[t] = (uint32_t)s;
}

// Node count of every nonempty tile, written at its last sorted position
// (empty tiles keep the zero they were filled with).
__global__ __launch_bounds__(256) void k_bin_tile_count(const uint32_t *__restrict__ skeys, int64_t M,
                                                        uint32_t sub_range,
                                                        const uint32_t *__restrict__ first,
                                                        uint32_t *__restrict__ cnt)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= M) return;
    const uint32_t t = skeys[s] / sub_range;
    if (s == M - 1 || skeys[s + 1] / sub_range != t) cnt[t] = (uint32_t)s + 1u - first[t];
}

// Exclusive scans of the tile counts and of the chunk counts, in three
// launches: sums per block of kScanBlock tiles, one workgroup scanning those
// sums, then each block's own scan plus its offset.
constexpr int kScanBlock = 1024;

// Inclusive scan of (a, b) over a workgroup of kScanBlock threads: wavefront
// scans with shuffles, then the wavefront totals through LDS.
__device__ __forceinline__ void block_scan2(uint32_t &a, uint32_t &b, uint32_t *wa, uint32_t *wb)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 1; off < 64; off <<= 1) {
        const uint32_t va = __shfl_up(a, off, 64), vb = __shfl_up(b, off, 64);
        if (lane >= off) {
            a += va;
            b += vb;
        }
    }
    if (lane == 63) {
        wa[wave] = a;
        wb[wave] = b;
    }
    __syncthreads();
    if (wave == 0) {
        uint32_t ta = lane < kScanBlock / 64 ? wa[lane] : 0u, tb = lane < kScanBlock / 64 ? wb[lane] : 0u;
        for (int off = 1; off < kScanBlock / 64; off <<= 1) {
            const uint32_t va = __shfl_up(ta, off, 64), vb = __shfl_up(tb, off, 64);
            if (lane >= off) {
                ta += va;
                tb += vb;
            }
        }
        if (lane < kScanBlock / 64) {
            wa[lane] = ta;
            wb[lane] = tb;
        }
    }
    __syncthreads();
    if (wave > 0) {
        a += wa[wave - 1];
        b += wb[wave - 1];
    }
}

// (a, b) = (count, chunks of count) of element i of cnt, or of the digit
// counts when chunks is false (b then unused).
__global__ __launch_bounds__(kScanBlock) void k_bin_scan_part(const uint32_t *__restrict__ cnt,
                                                              uint32_t *__restrict__ part,
                                                              int64_t ntiles, int64_t nparts)
{
    __shared__ uint32_t wa[kScanBlock / 64], wb[kScanBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const uint32_t c = i < ntiles ? cnt[i] : 0u;
    uint32_t a = c, b = (c + kChunk - 1) / kChunk;
    block_scan2(a, b, wa, wb);
    if (threadIdx.x == kScanBlock - 1) {
        part[blockIdx.x] = a;
        part[nparts + blockIdx.x] = b;
    }
}

// One workgroup: exclusive scan of the block sums in place; the chunk total
// (n_chunks may be NULL).
__global__ __launch_bounds__(kScanBlock) void k_bin_scan_top(uint32_t *__restrict__ part,
                                                             int64_t nparts,
                                                             uint32_t *__restrict__ n_chunks)
{
    __shared__ uint32_t wa[kScanBlock / 64], wb[kScanBlock / 64];
    uint32_t ca = 0, cb = 0;
    for (int64_t base = 0; base < nparts; base += kScanBlock) {
        const int64_t i = base + threadIdx.x;
        const uint32_t a0 = i < nparts ? part[i] : 0u, b0 = i < nparts ? part[nparts + i] : 0u;
        uint32_t a = a0, b = b0;
        block_scan2(a, b, wa, wb);
        if (i < nparts) {
            part[i] = ca + a - a0;
            part[nparts + i] = cb + b - b0;
        }
        __syncthreads();   // every thread has its block totals before wa/wb are reused
        ca += wa[kScanBlock / 64 - 1];
        cb += wb[kScanBlock / 64 - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0 && n_chunks) n_chunks[0] = cb;
}

// start = exclusive scan of cnt; chunk_start (may be NULL) = exclusive scan
// of the chunk counts.  start may be cnt itself (in place).
__global__ __launch_bounds__(kScanBlock) void k_bin_scan_apply(const uint32_t *cnt,
                                                               const uint32_t *__restrict__ part,
                                                               uint32_t *start,
                                                               uint32_t *__restrict__ chunk_start,
                                                               int64_t ntiles, int64_t nparts)
{
    __shared__ uint32_t wa[kScanBlock / 64], wb[kScanBlock / 64];
    const int64_t i = (int64_t)blockIdx.x * kScanBlock + threadIdx.x;
    const uint32_t c = i < ntiles ? cnt[i] : 0u, cc = (c + kChunk - 1) / kChunk;
    uint32_t a = c, b = cc;
    block_scan2(a, b, wa, wb);
    if (i < ntiles) {
        start[i] = part[blockIdx.x] + a - c;
        if (chunk_start) chunk_start[i] = part[nparts + blockIdx.x] + b - cc;
    }
}

// ------------------------------------------------------- radix sort ----
// Stable LSD radix sort of (composite key, node index) pairs, 8-bit digits.
// Each pass: per-block digit counts (k_rs_hist), an exclusive scan of the
// counts in digit-major order (digit d of block b at d * nblk + b), then a
// stable scatter (k_rs_scatter) in which block b walks its items in order,
// 256 at a time, and ranks each among the earlier items of its digit: within
// a wavefront from ballots on the digit's bits, across the block's wavefronts
// and rounds from LDS counters.  No atomic decides a position, so the result
// is the same on every run.
constexpr int kRsThreads = 256;
constexpr int kRsRounds = 16;
constexpr int kRsItems = kRsThreads * kRsRounds;   // items per block
constexpr int kRsDigits = 256;

__global__ __launch_bounds__(kRsThreads) void k_rs_hist(const uint32_t *__restrict__ keys, int64_t M,
                                                        int shift, uint32_t *__restrict__ hist,
                                                        uint32_t nblk)
{
    __shared__ uint32_t h[kRsDigits];
    h[threadIdx.x] = 0;
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * kRsItems + threadIdx.x;
#pragma unroll 4
    for (int k = 0; k < kRsRounds; ++k) {
        const int64_t i = i0 + (int64_t)k * kRsThreads;
        if (i < M) atomicAdd(&h[(keys[i] >> shift) & 255u], 1u);
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * nblk + blockIdx.x] = h[threadIdx.x];
}

// vals_in NULL: the values are the item indices (the first pass over node
// order).  keys_out may be NULL when the sorted keys are not needed.
__global__ __launch_bounds__(kRsThreads) void k_rs_scatter(const uint32_t *__restrict__ keys_in,
                                                           const uint32_t *__restrict__ vals_in,
                                                           int64_t M, int shift,
                                                           const uint32_t *__restrict__ offs,
                                                           uint32_t nblk,
                                                           uint32_t *__restrict__ keys_out,
                                                           uint32_t *__restrict__ vals_out)
{
    __shared__ uint32_t base[kRsDigits];                        // next position per digit
    __shared__ uint32_t wcnt[kRsThreads / 64][kRsDigits];       // this digit pass's count per wavefront
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    base[tid] = offs[(size_t)tid * nblk + blockIdx.x];
#pragma unroll
    for (int w = 0; w < kRsThreads / 64; ++w) wcnt[w][tid] = 0;
    const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    __syncthreads();
    const int64_t i0 = (int64_t)blockIdx.x * kRsItems + tid;
    for (int k = 0; k < kRsRounds; ++k) {
        const int64_t i = i0 + (int64_t)k * kRsThreads;
        if (((int64_t)blockIdx.x * kRsItems + (int64_t)k * kRsThreads) >= M) break;   // uniform
        const bool valid = i < M;
        const uint32_t key = valid ? keys_in[i] : 0u;
        const uint32_t val = valid ? (vals_in ? vals_in[i] : (uint32_t)i) : 0u;
        const uint32_t dg = (key >> shift) & 255u;
        // lanes of this wavefront holding the same digit (valid lanes only)
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int bit = 0; bit < 8; ++bit) {
            const uint64_t ones = __ballot(valid && ((dg >> bit) & 1u));
            peers &= ((dg >> bit) & 1u) ? ones : ~ones;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & below);
        if (valid && rank == 0) wcnt[wave][dg] = (uint32_t)__popcll(peers);
        __syncthreads();
        if (valid) {
            uint32_t pos = base[dg] + rank;
            for (int w = 0; w < wave; ++w) pos += wcnt[w][dg];
            if (keys_out) keys_out[pos] = key;
            vals_out[pos] = val;
        }
        __syncthreads();
        uint32_t tot = 0;
#pragma unroll
        for (int w = 0; w < kRsThreads / 64; ++w) {
            tot += wcnt[w][tid];
            wcnt[w][tid] = 0;
        }
        base[tid] += tot;
        __syncthreads();
    }
}

// Sorted-order copies of the per-node data: descriptors (plane-tiled
// kernels) or coordinates (node-lane kernels).
template <typename T>
__global__ __launch_bounds__(256) void k_bin_gather_desc(const uint32_t *__restrict__ perm,
                                                         const NodeDesc<T> *__restrict__ desc_j,
                                                         NodeDesc<T> *__restrict__ desc, int64_t M)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s < M) desc[s] = desc_j[perm[s]];
}

__global__ __launch_bounds__(256) void k_bin_gather_x(const uint32_t *__restrict__ perm,
                                                      const double *__restrict__ x,
                                                      double *__restrict__ xs, int64_t M, int d)
{
    const int64_t s = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (s >= M) return;
    const int64_t j = perm[s];
    for (int t = 0; t < d; ++t) xs[s * d + t] = x[j * d + t];
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

bool Binning::plan(const Geom &ge, int elem_bytes_, int kernel_variant, int tile_variant_)
{
    variant = kernel_variant;
    tile_variant = tile_variant_;
    elem_bytes = elem_bytes_;
    d = ge.d;
    usable = 0;
    const int W = ge.W;
    if (W > 18 || ge.M >= (int64_t(1) << 31)) return false;
    int64_t shape[3];
    tile_shape(d, planes() ? tile_variant : 0, shape);
    int64_t bytes = 2 * elem_bytes;
    ntiles = 1;
    int64_t ncells = 1;
    for (int t = 0; t < 3; ++t) {
        T[t] = t < d ? shape[t] : 1;
        if (t < d && (ge.n[t] < T[t] || ge.n[t] % T[t] != 0)) return false;
        nt[t] = t < d ? ge.n[t] / T[t] : 1;
        ntiles *= nt[t];
        ncells *= t < d ? ge.n[t] : 1;
        if (t < d) bytes *= T[t] + W - 1;
    }
    if (ntiles >= (int64_t(1) << 31)) return false;
    if (planes()) {
        // plane-tiled kernels: 32-bit grid offsets and the LDS tile must fit
        if (ncells > (int64_t(1) << 32)) return false;
        const bool fits = elem_bytes == 4 ? tile_kernels_fit<float>(ge, tile_variant, compensated)
                                          : tile_kernels_fit<double>(ge, tile_variant, false);
        if (!fits) return false;
        slot_cells = elem_bytes == 4 ? tile_slot_cells<float>(ge, tile_variant)
                                     : tile_slot_cells<double>(ge, tile_variant);
    } else {
        if (bytes > kMaxTileBytes) return false;
        if (d == 3 && W > 14) return false;
    }
    // the composite sort key must fit 32 bits
    const uint64_t sr = planes() ? sub_range(bin_geom(*this)) : 1u;
    const uint64_t kmax = (uint64_t)ntiles * sr;
    if (kmax > (uint64_t(1) << 32)) return false;
    int bits = 1;
    while ((uint64_t(1) << bits) < kmax) ++bits;
    sort_passes = (bits + 7) / 8;
    usable = 1;
    return true;
}

static size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

// Offsets of the buffers in the binning's device block (carve == true: set
// the pointers from base).  Sizes:
//   keys, keys2, perm, perm2    M u32 each
//   rs_hist                     256 x radix blocks u32
//   scan_part                   2 x the blocks of the larger scan
//   tile_count/_start, chunk_start  ntiles u32 each; n_chunks 2 (count, work counter)
//   chunks                      max_chunks
//   plane-tiled: desc, desc2    M descriptors each; slots, fs (own scratch)
//   node-lane:   xs             M d doubles
static size_t bin_layout(Binning &b, int64_t M, char *base)
{
    const int64_t nblk = (M + kRsItems - 1) / kRsItems;
    const int64_t max_chunks = b.ntiles + (M + kChunk - 1) / kChunk;
    const int64_t scan_n = std::max<int64_t>(b.ntiles, kRsDigits * nblk);
    const size_t db = b.elem_bytes == 4 ? sizeof(NodeDesc<float>) : sizeof(NodeDesc<double>);
    size_t off = 0;
    auto take = [&](size_t bytes) {
        char *p = base ? base + off : nullptr;
        off += align256(bytes ? bytes : 16);
        return p;
    };
    uint32_t *keys = (uint32_t *)take(4 * (size_t)M), *keys2 = (uint32_t *)take(4 * (size_t)M);
    uint32_t *perm = (uint32_t *)take(4 * (size_t)M), *perm2 = (uint32_t *)take(4 * (size_t)M);
    uint32_t *rs_hist = (uint32_t *)take(4 * (size_t)kRsDigits * nblk);
    uint32_t *scan_part = (uint32_t *)take(4 * 2 * (size_t)((scan_n + kScanBlock - 1) / kScanBlock));
    uint32_t *tile_count = (uint32_t *)take(4 * (size_t)b.ntiles);
    uint32_t *tile_start = (uint32_t *)take(4 * (size_t)b.ntiles);
    uint32_t *chunk_start = (uint32_t *)take(4 * (size_t)b.ntiles);
    uint32_t *n_chunks = (uint32_t *)take(8);   // chunk count, then the spread's work counter
    Chunk *chunks = (Chunk *)take(sizeof(Chunk) * (size_t)max_chunks);
    void *desc = nullptr, *desc2 = nullptr, *slots = nullptr, *fs = nullptr;
    double *xs = nullptr;
    if (b.planes()) {
        desc = take((size_t)M * db);
        desc2 = take((size_t)M * db);
        if (b.own_scratch) {
            slots = take((size_t)max_chunks * b.slot_cells * 2 * b.elem_bytes);
            fs = take((size_t)M * 2 * b.elem_bytes);
        }
    } else {
        xs = (double *)take(sizeof(double) * (size_t)M * b.d);
    }
    if (base) {
        b.keys = keys;
        b.keys2 = keys2;
        b.perm = perm;
        b.perm2 = perm2;
        b.rs_hist = rs_hist;
        b.scan_part = scan_part;
        b.tile_count = tile_count;
        b.tile_start = tile_start;
        b.chunk_start = chunk_start;
        b.n_chunks = n_chunks;
        b.chunks = chunks;
        b.desc = desc;
        b.desc2 = desc2;
        b.slots = slots;
        b.fs = fs;
        b.xs = xs;
        b.max_chunks = max_chunks;
        b.M = M;
    }
    return off;
}

size_t Binning::bytes_for(int64_t M_) const { return bin_layout(const_cast<Binning &>(*this), M_, nullptr); }

void Binning::carve(char *base, int64_t M_) { bin_layout(*this, M_, base); }

bool Binning::bind(void *base, size_t bytes, int64_t M_)
{
    if (!usable || bytes < bytes_for(M_)) return false;
    release();
    block = base;
    own_block = false;
    carve(static_cast<char *>(base), M_);
    return true;
}

// Scan of n counts cnt in place into their exclusive prefix (start == cnt
// allowed) and, for tile counts, of their chunk counts into chunk_start.
static void scan_counts(const uint32_t *cnt, uint32_t *start, uint32_t *chunk_start,
                        uint32_t *n_chunks, uint32_t *part, int64_t n, hipStream_t s)
{
    const int64_t nparts = (n + kScanBlock - 1) / kScanBlock;
    k_bin_scan_part<<<(unsigned)nparts, kScanBlock, 0, s>>>(cnt, part, n, nparts);
    k_bin_scan_top<<<1, kScanBlock, 0, s>>>(part, nparts, n_chunks);
    k_bin_scan_apply<<<(unsigned)nparts, kScanBlock, 0, s>>>(cnt, part, start, chunk_start, n, nparts);
}

// The binning passes: keys (and, for the plane-tiled kernels, descriptors by
// node), the radix sort of the keys, the tile counts from the sorted keys,
// tile scan and chunk list, and the sorted copies of the per-node data.
template <typename T> void Binning::run_build(const Geom &ge, const BinGeom &bg, const double *x,
                                              hipStream_t s)
{
    const unsigned gm = (unsigned)((M + 255) / 256);
    NodeDesc<T> *dj = planes() ? static_cast<NodeDesc<T> *>(desc2) : nullptr;   // by node
    const uint32_t sr = planes() ? sub_range(bg) : 1u;
    k_fill_u32<<<(unsigned)((ntiles + 255) / 256), 256, 0, s>>>(tile_count, ntiles, 0u);
    k_bin_keys<T><<<gm, 256, 0, s>>>(x, keys, dj, ge, bg, sr);
    // radix passes: pass p reads (kin, vin) and writes (kout, vout); the
    // destinations alternate so that the last pass writes perm
    const uint32_t nblk = (uint32_t)((M + kRsItems - 1) / kRsItems);
    uint32_t *kbuf[2] = {keys, keys2};
    uint32_t *vbuf[2] = {perm, perm2};
    uint32_t *skeys = nullptr;   // the sorted keys (written by the last pass)
    for (int p = 0; p < sort_passes; ++p) {
        const int dst = (sort_passes - 1 - p) & 1;   // the last pass writes buffer 0 (perm)
        const uint32_t *kin = p == 0 ? keys : kbuf[dst ^ 1];
        const uint32_t *vin = p == 0 ? nullptr : vbuf[dst ^ 1];
        // the first pass reads keys (buffer 0); it must not also write keys;
        // the last one writes its keys into whichever key buffer it does not read
        uint32_t *kout = p + 1 < sort_passes ? (p == 0 ? keys2 : kbuf[dst]) : (kin == keys ? keys2 : keys);
        if (p + 1 == sort_passes) skeys = kout;
        uint32_t *vout = vbuf[dst];
        k_rs_hist<<<nblk, kRsThreads, 0, s>>>(kin, M, 8 * p, rs_hist, nblk);
        scan_counts(rs_hist, rs_hist, nullptr, nullptr, scan_part, (int64_t)kRsDigits * nblk, s);
        k_rs_scatter<<<nblk, kRsThreads, 0, s>>>(kin, vin, M, 8 * p, rs_hist, nblk, kout, vout);
        if (p == 0) {
            // later passes ping-pong between keys2 and keys
            kbuf[dst] = keys2;
            kbuf[dst ^ 1] = keys;
        }
    }
    // tile counts (tile_start holds each tile's first position until the scan
    // overwrites it), then the scans and the chunk list
    k_bin_tile_first<<<gm, 256, 0, s>>>(skeys, M, sr, tile_start);
    k_bin_tile_count<<<gm, 256, 0, s>>>(skeys, M, sr, tile_start, tile_count);
    scan_counts(tile_count, tile_start, chunk_start, n_chunks, scan_part, ntiles, s);
    k_bin_chunks<<<(unsigned)((ntiles + 255) / 256), 256, 0, s>>>(tile_count, tile_start,
                                                                  chunk_start, chunks, ntiles);
    if (planes())
        k_bin_gather_desc<T><<<gm, 256, 0, s>>>(perm, dj, static_cast<NodeDesc<T> *>(desc), M);
    else
        k_bin_gather_x<<<gm, 256, 0, s>>>(perm, x, xs, M, d);
}

void Binning::build(const Geom &ge, const double *x, hipStream_t s)
{
    if (!usable) return;
    if (!block || M != ge.M) {
        if (block && !own_block)
            throw Error(CUNFFT_ERR_STATE, "bound binning memory is sized for another node count");
        release();
        const size_t bytes = bytes_for(ge.M);
        void *p = nullptr;
        CUNFFT_HIP(hipMalloc(&p, bytes));
        block = p;
        own_block = true;
        carve(static_cast<char *>(p), ge.M);
    }
    const BinGeom bg = bin_geom(*this);
    if (elem_bytes == 4) run_build<float>(ge, bg, x, s);
    else run_build<double>(ge, bg, x, s);
    CUNFFT_HIP(hipGetLastError());
    built = 1;
}

void Binning::release()
{
    if (block && own_block) (void)hipFree(block);
    block = nullptr;
    own_block = false;
    keys = keys2 = perm = perm2 = rs_hist = scan_part = nullptr;
    tile_count = tile_start = chunk_start = n_chunks = nullptr;
    chunks = nullptr;
    xs = nullptr;
    desc = desc2 = slots = fs = nullptr;
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
    const int64_t c0 = wrap_hi(o[0] + p0, ge.n[0]);
    if constexpr (D == 1) return c0;
    const int64_t c1 = wrap_hi(o[1] + p1, ge.n[1]);
    if constexpr (D == 2) return c0 * ge.n[1] + c1;
    const int64_t c2 = wrap_hi(o[2] + p2, ge.n[2]);
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
        double c;
        int lo = (int)(window_cell(xs[(int64_t)s * D + t], ge.n[t], ge.m, c) - o[t]);
        lo = lo < 0 ? 0 : (lo >= TileShape<D>::T[t] ? TileShape<D>::T[t] - 1 : lo);
        window_weights<T, W>(c, ge, t, w[t]);
        base += lo * stride[t];
    }
    return base;
}

// --------------------------------------------------------- tiled gather --
template <typename T, int D, int W>
__device__ __forceinline__ void gather_tiled(const cplx_t<T> *__restrict__ g,
                                             const double *__restrict__ xs,
                                             const uint32_t *__restrict__ perm, const Chunk ch,
                                             cplx_t<T> *__restrict__ f, const Geom &ge,
                                             const BinGeom &bg)
{
    using PD = Pad<D, W>;
    __shared__ cplx_t<T> tile[PD::PT];
    int64_t o[3];
    tile_origin<D>(ch.tile, bg, o);
    // 8 loads in flight per thread before the LDS stores (unconditional: a
    // lane past the end rewrites the last cell with its own value)
    for (int i0 = threadIdx.x; i0 < PD::PT; i0 += 8 * kTileThreads) {
        cplx_t<T> v[8];
        int at[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            at[u] = i0 + u * kTileThreads < PD::PT ? i0 + u * kTileThreads : PD::PT - 1;
            v[u] = g[tile_cell<D, W>(at[u], o, ge)];
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) tile[at[u]] = v[u];
    }
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

template <typename T, int D, int W>
__global__ __launch_bounds__(kTileThreads) void k_gather_tiled(
    const cplx_t<T> *__restrict__ g, const double *__restrict__ xs, const uint32_t *__restrict__ perm,
    const Chunk *__restrict__ chunks, const uint32_t *__restrict__ n_chunks, cplx_t<T> *__restrict__ f,
    Geom ge, BinGeom bg)
{
    if (blockIdx.x >= n_chunks[0]) return;
    gather_tiled<T, D, W>(g, xs, perm, chunks[blockIdx.x], f, ge, bg);
}

template <typename T, int D, int W>
__global__ __launch_bounds__(kTileThreads) void k_gather_tiled_batch(
    const GridJob<T> *__restrict__ jobs, const BlockMap *__restrict__ map)
{
    const BlockMap bm = map[blockIdx.x];
    const GridJob<T> &J = jobs[bm.job];
    gather_tiled<T, D, W>(J.g, J.xs, J.perm, J.chunks[bm.blk], J.f_out, J.ge, J.bg);
}

// -------------------------------------------------------- tiled scatter --
template <typename T>
__device__ __forceinline__ void lds_add(T *p, T v)
{
    __hip_atomic_fetch_add(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}

template <typename T, typename A, int D, int W>
__device__ __forceinline__ void scatter_tiled(const double *__restrict__ xs,
                                              const uint32_t *__restrict__ perm,
                                              const cplx_t<T> *__restrict__ f, const Chunk ch,
                                              cplx_t<T> *__restrict__ g, const Geom &ge,
                                              const BinGeom &bg)
{
    using PD = Pad<D, W>;
    __shared__ A acc_re[PD::PT];
    __shared__ A acc_im[PD::PT];
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

template <typename T, typename A, int D, int W>
__global__ __launch_bounds__(kTileThreads) void k_scatter_tiled(
    const double *__restrict__ xs, const uint32_t *__restrict__ perm, const cplx_t<T> *__restrict__ f,
    const Chunk *__restrict__ chunks, const uint32_t *__restrict__ n_chunks, cplx_t<T> *__restrict__ g,
    Geom ge, BinGeom bg)
{
    if (blockIdx.x >= n_chunks[0]) return;
    scatter_tiled<T, A, D, W>(xs, perm, f, chunks[blockIdx.x], g, ge, bg);
}

template <typename T, typename A, int D, int W>
__global__ __launch_bounds__(kTileThreads) void k_scatter_tiled_batch(
    const GridJob<T> *__restrict__ jobs, const BlockMap *__restrict__ map)
{
    const BlockMap bm = map[blockIdx.x];
    const GridJob<T> &J = jobs[bm.job];
    scatter_tiled<T, A, D, W>(J.xs, J.perm, J.f_in, J.chunks[bm.blk], J.g, J.ge, J.bg);
}

// Grid <- 0 for the node-lane scatter of every job, 1024 cells per block.
template <typename T>
__global__ __launch_bounds__(256) void k_zero_batch(const GridJob<T> *__restrict__ jobs,
                                                    const BlockMap *__restrict__ map)
{
    const BlockMap bm = map[blockIdx.x];
    const GridJob<T> &J = jobs[bm.job];
    const int64_t c0 = (int64_t)bm.blk * 1024 + threadIdx.x;
#pragma unroll
    for (int k = 0; k < 4; ++k)
        if (c0 + 256 * k < J.n_total) J.g[c0 + 256 * k] = cmake<T>(0, 0);
}

// ------------------------------------------------------------ dispatch --
template <typename T, int D, int W>
static void launch_gather_lanes(const Geom &ge, const Binning &b, const cplx_t<T> *g, cplx_t<T> *f,
                                hipStream_t s)
{
    if constexpr (tile_fits<T, D, W>()) {
        k_gather_tiled<T, D, W><<<(unsigned)b.max_chunks, kTileThreads, 0, s>>>(
            g, b.xs, b.perm, b.chunks, b.n_chunks, f, ge, bin_geom(b));
    } else {
        throw Error(CUNFFT_ERR_UNSUPPORTED, "tile does not fit in LDS");
    }
}

template <typename T, int D, int W>
static void launch_scatter_lanes(const Geom &ge, const Binning &b, const cplx_t<T> *f,
                                 cplx_t<T> *g, hipStream_t s)
{
    if constexpr (tile_fits<T, D, W>()) {
        k_scatter_tiled<T, double, D, W><<<(unsigned)b.max_chunks, kTileThreads, 0, s>>>(
            b.xs, b.perm, f, b.chunks, b.n_chunks, g, ge, bin_geom(b));
    } else {
        throw Error(CUNFFT_ERR_UNSUPPORTED, "tile does not fit in LDS");
    }
}

#define CUNFFT_TILED_W(FN, D)                                             \
    switch (ge.W) {                                                       \
    case 4: FN<T, D, 4>(ge, b, a0, a1, s); break;                         \
    case 6: FN<T, D, 6>(ge, b, a0, a1, s); break;                         \
    case 8: FN<T, D, 8>(ge, b, a0, a1, s); break;                         \
    case 10: FN<T, D, 10>(ge, b, a0, a1, s); break;                       \
    case 12: FN<T, D, 12>(ge, b, a0, a1, s); break;                       \
    case 14: FN<T, D, 14>(ge, b, a0, a1, s); break;                       \
    case 16: if constexpr (D < 3) FN<T, D, 16>(ge, b, a0, a1, s); break;  \
    case 18: if constexpr (D < 3) FN<T, D, 18>(ge, b, a0, a1, s); break;  \
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
    if (b.planes()) {
        launch_tile_gather<T>(ge, b, g, f, s);
        return;
    }
    const cplx_t<T> *a0 = g;
    cplx_t<T> *a1 = f;
    switch (ge.d) {
    case 1: CUNFFT_TILED_W(launch_gather_lanes, 1) break;
    case 2: CUNFFT_TILED_W(launch_gather_lanes, 2) break;
    default: CUNFFT_TILED_W(launch_gather_lanes, 3) break;
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
    if (b.planes()) {
        launch_tile_spread<T>(ge, b, f, g, s, mid);
        return mid != nullptr;
    }
    launch_zero<T>(g, nt, s);
    const cplx_t<T> *a0 = f;
    cplx_t<T> *a1 = g;
    switch (ge.d) {
    case 1: CUNFFT_TILED_W(launch_scatter_lanes, 1) break;
    case 2: CUNFFT_TILED_W(launch_scatter_lanes, 2) break;
    default: CUNFFT_TILED_W(launch_scatter_lanes, 3) break;
    }
    CUNFFT_HIP(hipGetLastError());
    return false;
}

// ------------------------------------------------------------ batches --
int64_t tile_fold_blocks(int d, int variant, int64_t ntiles, int elem_bytes)
{
    int64_t T[3];
    tile_shape(d, variant, T);
    if (d >= 2 && elem_bytes == 4) {   // column-segment fold: kFoldSegBytes of a core column per thread
        const int64_t tpt = T[0] * (d == 3 ? T[1] : 1) * (T[d - 1] * 2 * elem_bytes / kFoldSegBytes);
        return tpt <= 256 ? (ntiles + 256 / tpt - 1) / (256 / tpt) : ntiles * (tpt / 256);
    }
    const int64_t core = T[0] * T[1] * T[2];
    return ntiles * (core >= 512 ? core / 512 : 1);
}

template <typename T, int D, int W>
static void launch_lanes_batch(BatchStage st, const GridJob<T> *jobs, const BlockMap *map,
                               uint32_t nblk, hipStream_t s)
{
    if constexpr (tile_fits<T, D, W>()) {
        if (st == BS_GATHER)
            k_gather_tiled_batch<T, D, W><<<nblk, kTileThreads, 0, s>>>(jobs, map);
        else
            k_scatter_tiled_batch<T, double, D, W><<<nblk, kTileThreads, 0, s>>>(jobs, map);
    } else {
        throw Error(CUNFFT_ERR_UNSUPPORTED, "tile does not fit in LDS");
    }
}

template <typename T>
void launch_batch_stage(BatchStage st, const GridJob<T> *jobs, const BlockMap *map, uint32_t nblk,
                        int d, int W, int config, hipStream_t s)
{
    if (nblk == 0) return;
    if (st == BS_ZERO) {
        k_zero_batch<T><<<nblk, 256, 0, s>>>(jobs, map);
    } else if ((st == BS_GATHER && (d == 1 || config == kTileVariants)) || st == BS_SCATTER_LANES) {
        const BatchStage a0 = st;
        const BlockMap *a1 = map;
#define CUNFFT_LANES_W(D)                                                            \
    switch (W) {                                                                     \
    case 4: launch_lanes_batch<T, D, 4>(a0, jobs, a1, nblk, s); break;               \
    case 6: launch_lanes_batch<T, D, 6>(a0, jobs, a1, nblk, s); break;               \
    case 8: launch_lanes_batch<T, D, 8>(a0, jobs, a1, nblk, s); break;               \
    case 10: launch_lanes_batch<T, D, 10>(a0, jobs, a1, nblk, s); break;             \
    case 12: launch_lanes_batch<T, D, 12>(a0, jobs, a1, nblk, s); break;             \
    case 14: launch_lanes_batch<T, D, 14>(a0, jobs, a1, nblk, s); break;             \
    case 16: if constexpr (D < 3) launch_lanes_batch<T, D, 16>(a0, jobs, a1, nblk, s); break; \
    case 18: if constexpr (D < 3) launch_lanes_batch<T, D, 18>(a0, jobs, a1, nblk, s); break; \
    default: throw Error(CUNFFT_ERR_UNSUPPORTED, "tiled gridding needs m <= 8");     \
    }
        switch (d) {
        case 1: CUNFFT_LANES_W(1) break;
        case 2: CUNFFT_LANES_W(2) break;
        default: CUNFFT_LANES_W(3) break;
        }
#undef CUNFFT_LANES_W
    } else {
        launch_tile_batch<T>(st, jobs, map, nblk, d, W, config, s);
        return;
    }
    CUNFFT_HIP(hipGetLastError());
}

template void launch_batch_stage<float>(BatchStage, const GridJob<float> *, const BlockMap *,
                                        uint32_t, int, int, int, hipStream_t);
template void launch_batch_stage<double>(BatchStage, const GridJob<double> *, const BlockMap *,
                                         uint32_t, int, int, int, hipStream_t);

template void launch_gather_binned<float>(const Geom &, const Binning &, const float2 *,
                                          const double *, float2 *, hipStream_t);
template void launch_gather_binned<double>(const Geom &, const Binning &, const double2 *,
                                           const double *, double2 *, hipStream_t);
template bool launch_scatter_binned<float>(const Geom &, const Binning &, const double *,
                                           const float2 *, float2 *, hipStream_t, hipEvent_t);
template bool launch_scatter_binned<double>(const Geom &, const Binning &, const double *,
                                            const double2 *, double2 *, hipStream_t, hipEvent_t);

}  // namespace cunfft
