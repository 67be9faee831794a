// bin.hpp -- node binning shared by the tiled gridding kernels.
#pragma once
#include "internal.hpp"

namespace cunfft {

constexpr int kChunk = 256;   // nodes per chunk (one workgroup's share of a tile)
constexpr int kFoldSegBytes = 32;   // bytes of a core column one fold thread sums (FP32, d >= 2)

// Grid tile per dimension: the nodes whose window starts in one tile are
// gridded by one workgroup against the tile plus a (2m+1)-cell halo in LDS.
// The plane-tiled kernels (d = 2, 3) come in kTileVariants tile shapes; the
// autotuner picks one per plan.  Variant 0 is the default.
constexpr int kTileVariants = 3;
template <int D, int V> struct TileShapeV;
template <> struct TileShapeV<1, 0> { static constexpr int T[3] = {256, 1, 1}; };
template <> struct TileShapeV<2, 0> { static constexpr int T[3] = {32, 32, 1}; };
template <> struct TileShapeV<2, 1> { static constexpr int T[3] = {16, 16, 1}; };
template <> struct TileShapeV<2, 2> { static constexpr int T[3] = {64, 64, 1}; };
template <> struct TileShapeV<3, 0> { static constexpr int T[3] = {8, 8, 8}; };
template <> struct TileShapeV<3, 1> { static constexpr int T[3] = {8, 8, 16}; };
template <> struct TileShapeV<3, 2> { static constexpr int T[3] = {16, 16, 8}; };
template <int D> using TileShape = TileShapeV<D, 0>;

// Tile extent per axis of variant v in d dimensions (host side).
inline void tile_shape(int d, int v, int64_t (&T)[3])
{
    static const int shapes[4][kTileVariants][3] = {
        {{1, 1, 1}, {1, 1, 1}, {1, 1, 1}},
        {{256, 1, 1}, {256, 1, 1}, {256, 1, 1}},
        {{32, 32, 1}, {16, 16, 1}, {64, 64, 1}},
        {{8, 8, 8}, {8, 8, 16}, {16, 16, 8}}};
    for (int t = 0; t < 3; ++t) T[t] = shapes[d][v][t];
}

// Work item of the tiled kernels: up to kChunk sorted nodes of one tile.
struct Chunk {
    uint32_t tile;
    uint32_t start;  // first sorted position
    uint32_t count;
    uint32_t pad;
};

// Workgroups of a grid are dispatched round-robin over the 8 XCDs of an
// MI355X (workgroup b runs on XCD b mod 8), each XCD with its own L2.  The
// tile-staging kernels deal the chunks (tiles in row-major order) to the
// XCDs in groups of kXcdGroup consecutive chunks, group g to XCD g mod 8:
// within a group, neighbouring tiles, whose halo-padded regions overlap, and
// the chunks of one dense tile are staged through the same L2, while the
// groups of a dense region still spread over all XCDs (contiguous eighths
// would put a node cluster on one or two XCDs).  Chunk of workgroup b, or -1
// when b maps past the n chunks; the grid must hold xcd_grid(max_chunks)
// workgroups for any n <= max_chunks (workgroup b can map up to
// 7 (kXcdGroup - 1) chunks below b).
constexpr uint32_t kXcds = 8;
constexpr uint32_t kXcdGroup = 128;
__host__ __device__ __forceinline__ int64_t xcd_chunk(uint32_t b, uint32_t n)
{
    const uint32_t x = b % kXcds, local = b / kXcds;
    const uint64_t g = (uint64_t)(local / kXcdGroup) * kXcds + x;
    const uint64_t c = g * kXcdGroup + local % kXcdGroup;
    return c < n ? (int64_t)c : -1;
}
inline unsigned xcd_grid(int64_t max_chunks)
{
    const int64_t need = max_chunks + (int64_t)(kXcds - 1) * (kXcdGroup - 1);
    return (unsigned)((need + kXcds - 1) / kXcds * kXcds);
}

// Per-node window parameters in sorted order, written at binning time so the
// gridding kernels do no FP64 index arithmetic:
//   f[t]  = c_t - m, with c_t = n_t x_t - u_t the offset of the node from its
//           first window cell u_t (window tap l has argument f[t] + m - l);
//   loc   = tile-local first window cell, lo_0 | lo_1 << 8 | lo_2 << 16.
// FP32 plans keep f in FP32 (|f| < 2, so the rounding is below 1.2e-7).
template <typename T> struct NodeDesc;
template <> struct alignas(16) NodeDesc<float> {
    float f[3];
    uint32_t loc;
};
template <> struct alignas(16) NodeDesc<double> {
    double f[3];
    uint64_t loc;
};

struct BinGeom;

// Nodes sorted on the device by the grid tile that holds the start of their
// window, u_t mod n_t, and -- for the plane-tiled kernels -- within a tile by
// their window start there: row and column unit first, plane last (sub_key),
// so nodes whose windows cover the same rows and column units follow each
// other with rising plane (the adjoint accumulates such a run in registers
// and writes the tile once per run).  The sort is a stable LSD radix sort of
// the composite key tile * sub_range + sub_key over the node indices, so
// nodes with equal keys keep their index order: the sorted order, and with
// it the order of every adjoint sum, depends on the nodes only, never on
// scheduling.  Built once per node set on the plan stream and reused by
// every forward and adjoint call until the nodes change.  Tiles with many
// nodes are split into chunks so one workgroup never owns more than kChunk
// nodes.
//
// All device buffers live in one block: allocated by build() (own memory),
// or handed in by bind() (a caller's region; nothing is allocated or freed).
struct Binning {
    int built = 0;
    int usable = 0;             // tiled kernels apply to this geometry
    int variant = 0;            // 0: plane-tiled kernels (d >= 2); 1: one node per lane
    int tile_variant = 0;       // tile shape of the plane-tiled kernels (TileShapeV)
    int d = 0;
    int64_t T[3] = {1, 1, 1};   // tile extent per axis
    int64_t nt[3] = {1, 1, 1};  // tiles per axis
    int64_t ntiles = 0;
    int64_t M = 0;
    int64_t max_chunks = 0;
    int64_t slot_cells = 0;     // padded-tile cells of one adjoint slot
    int sort_passes = 0;        // radix passes (8-bit digits) over the composite key
    uint32_t *keys = nullptr;       // M: composite key of each node (node order), sort ping-pong
    uint32_t *keys2 = nullptr;      // M: sort ping-pong
    uint32_t *perm = nullptr;       // M: node at each sorted position
    uint32_t *perm2 = nullptr;      // M: sort ping-pong
    uint32_t *rs_hist = nullptr;    // 256 x radix blocks: digit counts, then their offsets
    uint32_t *scan_part = nullptr;  // block sums of the scans
    uint32_t *tile_count = nullptr; // ntiles
    uint32_t *tile_start = nullptr; // ntiles
    uint32_t *chunk_start = nullptr;// ntiles
    uint32_t *n_chunks = nullptr;   // 2: the chunk count, then the persistent spread's work counter
    Chunk *chunks = nullptr;        // max_chunks
    double *xs = nullptr;           // M*d: nodes in sorted order (node-lane kernels)
    void *desc = nullptr;           // M: NodeDesc<T> in sorted order (plane-tiled kernels)
    void *desc2 = nullptr;          // M: NodeDesc<T> by node index (binning scratch)
    void *slots = nullptr;          // max_chunks padded tiles (adjoint partial sums)
    void *fs = nullptr;             // M: adjoint coefficients in sorted order
    void *block = nullptr;          // the one allocation holding the buffers above
    bool own_block = false;         // block came from build() (else from bind())
    int elem_bytes = 8;             // bytes of one real component (4: FP32 plan, 8: FP64)
    bool own_scratch = true;        // include slots and fs (a batch brings its own)
    bool compensated = false;       // FP32 Kahan tile sums and Neumaier fold (CUNFFT_COMPENSATED)

    // Decide tile shape for this geometry; false if the tiled path does not apply.
    bool plan(const Geom &ge, int elem_bytes, int kernel_variant, int tile_variant = 0);
    // Bytes of the device block for M nodes (plan() first).
    size_t bytes_for(int64_t M) const;
    // Use the caller's region [base, base + bytes) for M nodes instead of an
    // allocation; false (nothing changed) if it is too small.  base is 256-byte
    // aligned.  The next build() sorts into it.
    bool bind(void *base, size_t bytes, int64_t M);
    void build(const Geom &ge, const double *x, hipStream_t s);
    template <typename T> void run_build(const Geom &ge, const BinGeom &bg, const double *x, hipStream_t s);
    void release();
    bool planes() const { return d >= 2 && variant == 0; }

  private:
    void carve(char *base, int64_t M);
};

struct BinGeom {
    int d;
    int logT[3];
    int64_t nt[3];
    int64_t ntiles;
    int sub_shift;   // column cells per lane unit, log2 (FP32 1, FP64 0)
};

// Order of a node within its tile (window start lo_t relative to the tile
// origin, tile extents T_t): plane lo_0 fastest, then the column unit
// lo_L >> sub_shift, then the row lo_1 (d = 3).  Range: sub_range.
__host__ __device__ inline uint32_t sub_key(int d, const uint32_t (&lo)[3], const int (&logT)[3],
                                            int sub_shift)
{
    const uint32_t T0 = 1u << logT[0];
    const uint32_t cu = lo[d - 1] >> sub_shift;
    const uint32_t nu = (1u << logT[d - 1]) >> sub_shift;
    return d == 3 ? lo[0] + T0 * (cu + nu * lo[1]) : lo[0] + T0 * cu;
}
inline uint32_t sub_range(const BinGeom &g)
{
    uint32_t r = (1u << g.logT[0]) * ((1u << g.logT[g.d - 1]) >> g.sub_shift);
    return g.d == 3 ? r << g.logT[1] : r;
}
BinGeom bin_geom(const Binning &b);

// Window constants of the plane-tiled kernels.
template <typename T> struct WinConst {
    T kb[3];    // Gaussian: FP32 log2(e) / b_t (for exp2); FP64 1 / b_t (for exp)
    T nrm[3];   // Gaussian: (pi b_t)^-1/2
    T shp[3];   // Kaiser-Bessel shape b_t, or sinc a_t
    T dk[3];    // Gaussian: -2 n_t / b_t (the weight's x-derivative is dk s phi(s))
    T nn[3];    // n_t
    int m;
    int win;    // Window
};
template <typename T> WinConst<T> win_const(const Geom &ge);

// ------------------------------------------------------------ batches --
// One (problem, right-hand side) pair of a batched launch: the problem's grid
// geometry and node binning, shared by its right-hand sides, and the pair's
// own buffers.
template <typename T> struct GridJob {
    const void *desc;              // NodeDesc<T> by sorted position (plane-tiled kernels)
    const double *xs;              // nodes by sorted position (node-lane kernels)
    const uint32_t *perm;          // node at each sorted position
    const Chunk *chunks;
    const uint32_t *tile_count;
    const uint32_t *chunk_start;
    cplx_t<T> *g;                  // oversampled grid
    cplx_t<T> *fs;                 // adjoint coefficients by sorted position
    cplx_t<T> *slots;              // adjoint partial tiles, one per chunk
    const cplx_t<T> *f_in;         // adjoint input, M coefficients
    cplx_t<T> *f_out;              // forward output, M values
    int64_t n_total;               // grid cells
    Geom ge;
    BinGeom bg;
    WinConst<T> wc;
};

// Stages of a batched transform; each is one launch over all jobs.
enum BatchStage : int {
    BS_GATHER = 0,      // forward gridding (blocks: the binning's chunks)
    BS_COEFFS,          // adjoint coefficients into sorted order (blocks of 256 nodes)
    BS_SPREAD,          // plane-tiled adjoint accumulation into slots (chunks)
    BS_FOLD,            // slots -> grid (tiles x fold slices)
    BS_ZERO,            // grid <- 0 for the node-lane scatter (blocks of 1024 cells)
    BS_SCATTER_LANES,   // node-lane adjoint accumulation (chunks)
    BS_COUNT
};
// Blocks of job-independent size in the fold: slices of 512 core cells.
int64_t tile_fold_blocks(int d, int variant, int64_t ntiles, int elem_bytes);   // fold workgroups of a job
// One launch of stage st over the jobs of a batch whose problems share the
// dimension d, window width W, precision T and kernel configuration
// (plane-tiled variant 0..kTileVariants-1, or kTileVariants for node lanes).
template <typename T>
void launch_batch_stage(BatchStage st, const GridJob<T> *jobs, const BlockMap *map, uint32_t nblk,
                        int d, int W, int config, hipStream_t s);
template <typename T>
void launch_tile_batch(BatchStage st, const GridJob<T> *jobs, const BlockMap *map, uint32_t nblk,
                       int d, int W, int variant, hipStream_t s);

template <typename T>
void launch_gather_binned(const Geom &ge, const Binning &b, const cplx_t<T> *g, const double *x,
                          cplx_t<T> *f, hipStream_t s);
// mid (optional) is recorded on s after the tile accumulation and before the
// fold, so callers can time the two adjoint gridding kernels separately.
// Returns true if mid was recorded (the path has a fold).
template <typename T>
bool launch_scatter_binned(const Geom &ge, const Binning &b, const double *x, const cplx_t<T> *f,
                           cplx_t<T> *g, hipStream_t s, hipEvent_t mid = nullptr);

// Plane-tiled kernels for d = 2, 3 (grid_tile.hip).
template <typename T> bool tile_kernels_fit(const Geom &ge, int variant, bool comp);
template <typename T> int64_t tile_slot_cells(const Geom &ge, int variant);
template <typename T>
void launch_tile_gather(const Geom &ge, const Binning &b, const cplx_t<T> *g, cplx_t<T> *f,
                        hipStream_t s);
template <typename T>
void launch_tile_spread(const Geom &ge, const Binning &b, const cplx_t<T> *f, cplx_t<T> *g,
                        hipStream_t s, hipEvent_t mid);
// x-gradient gx[j d + t] = Re(conj(v_j) d f_j / d x_jt) over the forward grid g
// on a plane-tiled binning; false (nothing launched) when no kernel fits.
template <typename T>
bool launch_tile_grad(const Geom &ge, const Binning &b, const cplx_t<T> *g, const cplx_t<T> *v, T *gx,
                      hipStream_t s);

}  // namespace cunfft
