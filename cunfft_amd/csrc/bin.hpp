// bin.hpp -- node binning shared by the tiled gridding kernels.
#pragma once
#include "internal.hpp"

namespace cunfft {

// Work item of the tiled kernels: up to kChunk sorted nodes of one tile.
struct Chunk {
    uint32_t tile;
    uint32_t start;  // first sorted position
    uint32_t count;
    uint32_t pad;
};

// Nodes sorted (counting sort on the device) by the grid tile that holds the
// start of their window, u_t mod n_t.  Built once per node set on the plan
// stream and reused by every forward and adjoint call until the nodes change.
// Tiles with many nodes are split into chunks so one workgroup never owns more
// than kChunk nodes.
struct Binning {
    int built = 0;
    int usable = 0;             // tiled kernels apply to this geometry
    int variant = 0;            // 0: lanes along the window row (d >= 2); 1: one node per lane
    int d = 0;
    int64_t T[3] = {1, 1, 1};   // tile extent per axis
    int64_t nt[3] = {1, 1, 1};  // tiles per axis
    int64_t ntiles = 0;
    int64_t M = 0;
    int64_t max_chunks = 0;
    uint32_t *keys = nullptr;       // M: tile of each node
    uint32_t *perm = nullptr;       // M: node at each sorted position
    uint32_t *tile_count = nullptr; // ntiles
    uint32_t *tile_start = nullptr; // ntiles
    uint32_t *cursor = nullptr;     // ntiles
    uint32_t *chunk_start = nullptr;// ntiles
    uint32_t *n_chunks = nullptr;   // 1
    Chunk *chunks = nullptr;        // max_chunks
    double *xs = nullptr;           // M*d: nodes in sorted order
    void *slots = nullptr;          // max_chunks padded tiles (adjoint partial sums)
    int slot_elem_bytes = 8;        // bytes of one real component of a slot cell

    // Decide tile shape for this geometry; false if the tiled path does not apply.
    bool plan(const Geom &ge, int elem_bytes, int kernel_variant);
    void build(const Geom &ge, const double *x, hipStream_t s);
    void release();
};

template <typename T>
void launch_gather_binned(const Geom &ge, const Binning &b, const cplx_t<T> *g, const double *x,
                          cplx_t<T> *f, hipStream_t s);
// mid (optional) is recorded on s after the tile accumulation and before the
// fold, so callers can time the two adjoint gridding kernels separately.
// Returns true if mid was recorded (the path has a fold).
template <typename T>
bool launch_scatter_binned(const Geom &ge, const Binning &b, const double *x, const cplx_t<T> *f,
                           cplx_t<T> *g, hipStream_t s, hipEvent_t mid = nullptr);

}  // namespace cunfft
