// batch.hip -- many uneven NFFT problems and right-hand sides per graph launch.
//
// A batch holds P problems of one dimension d, cut-off m and precision, each
// with its own N, n and nodes, and R right-hand sides per problem.  Every
// stage of the forward transform (the d FFT passes, the gather) and of the
// adjoint (coefficient sort, tile accumulation, fold -- or grid clear and
// node-lane scatter -- then the d FFT passes) is ONE launch over all P*R
// (problem, right-hand side) jobs: workgroup b looks up its job and block in
// a block map, so small problems share the GPU instead of each filling a
// handful of CUs.  The stage launches are captured once into a hipGraph and
// replayed with one hipGraphLaunch per call.  Each problem is binned once
// per node set; the binning serves all its right-hand sides.
//
// The reference has no batch interface; a loop over cunfft_transform /
// cunfft_adjoint (src/cunfft.cu:196-224) of single plans is what this
// replaces.
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <new>
#include <string>
#include <vector>

#include "cunfft.h"
#include "internal.hpp"
#include "bin.hpp"

namespace cunfft {

void report_error(const char *what);

template <typename T> struct BatchImpl {
    using C = cplx_t<T>;
    struct Prob {
        Geom ge{};
        int64_t N_total = 0, n_total = 0, M = 0;
        double *x = nullptr;           // d*M node coordinates (device copy)
        double *roc = nullptr;
        int64_t rocoff[3] = {0, 0, 0};
        FftTables *ft = nullptr;
        Binning bins[2];               // forward [0], adjoint [1] (shares [0] if equal)
        bool dirty[2] = {true, true};
        uint32_t nch[2] = {0, 0};      // chunks of each binning (host copy)
        C *fhat_in = nullptr;          // R * N_total, forward input
        C *f_out = nullptr;            // R * M, forward output
        C *f_in = nullptr;             // R * M, adjoint input
        C *fhat_out = nullptr;         // R * N_total, adjoint output
        C *g = nullptr;                // R * n_total grids
        C *fs = nullptr;               // R * M sorted adjoint coefficients
        C *slots = nullptr;            // R * slot_cap adjoint partial tiles
        int64_t slot_cap = 0;          // cells of slots per right-hand side
    };
    struct Launch {
        int fft_t = -1;                // axis of an FFT stage, -1 for a gridding stage
        BatchStage st = BS_GATHER;
        int dir = 0;                   // job table (binning) of a gridding stage
        uint32_t nblk = 0;
        size_t map_off = 0;            // into maps
        size_t pass_off = 0;           // into passes (FFT stages)
        size_t lds = 0;
        int max_elems = 0;             // largest C * L of an FFT stage
    };

    int d = 0, m = 0, W = 0, R = 0;
    int config[2] = {0, 0};
    std::vector<Prob> probs;
    GridJob<T> *jobs[2] = {nullptr, nullptr};
    FftPass<T> *passes = nullptr;
    BlockMap *maps = nullptr;
    uint32_t *nch_host = nullptr;      // pinned, 2 per problem
    std::vector<Launch> stages[2];     // forward, adjoint
    bool tables_ok = false;
    hipStream_t cap = nullptr;         // capture stream
    hipGraphExec_t exec[4] = {nullptr, nullptr, nullptr, nullptr};  // by directions 1..3
    hipEvent_t ev[2] = {nullptr, nullptr};

    int slot(int dir) const { return config[1] == config[0] ? 0 : dir; }
    bool planes(int dir) const { return d >= 2 && config[dir] < kTileVariants; }

    ~BatchImpl() { release(); }

    void drop_graphs()
    {
        for (auto &e : exec)
            if (e) {
                (void)hipGraphExecDestroy(e);
                e = nullptr;
            }
    }

    void release()
    {
        if (cap) (void)hipStreamSynchronize(cap);
        (void)hipDeviceSynchronize();
        drop_graphs();
        for (Prob &p : probs) {
            for (void *q : {(void *)p.x, (void *)p.roc, (void *)p.fhat_in, (void *)p.f_out,
                            (void *)p.f_in, (void *)p.fhat_out, (void *)p.g, (void *)p.fs,
                            (void *)p.slots})
                if (q) (void)hipFree(q);
            free_fft_tables(p.ft);
            p.bins[0].release();
            p.bins[1].release();
        }
        probs.clear();
        for (auto &j : jobs)
            if (j) (void)hipFree(j);
        if (passes) (void)hipFree(passes);
        if (maps) (void)hipFree(maps);
        if (nch_host) (void)hipHostFree(nch_host);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
        if (cap) (void)hipStreamDestroy(cap);
        jobs[0] = jobs[1] = nullptr;
        passes = nullptr;
        maps = nullptr;
        nch_host = nullptr;
        cap = nullptr;
        ev[0] = ev[1] = nullptr;
    }

    static void *alloc(size_t bytes)
    {
        void *q = nullptr;
        CUNFFT_HIP(hipMalloc(&q, bytes ? bytes : 16));
        return q;
    }

    void create(int d_, const uint_t *N, const uint_t *n, const uint_t *M, int P, int m_, int R_)
    {
        d = d_;
        m = m_;
        W = 2 * m + 2;
        R = R_;
        probs.resize(P);
        CUNFFT_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        CUNFFT_HIP(hipEventCreate(&ev[0]));
        CUNFFT_HIP(hipEventCreate(&ev[1]));
        CUNFFT_HIP(hipHostMalloc((void **)&nch_host, sizeof(uint32_t) * 2 * P, hipHostMallocDefault));
        for (int i = 0; i < P; ++i) {
            Prob &p = probs[i];
            int64_t NN[3], nn[3];
            for (int t = 0; t < d; ++t) {
                NN[t] = N[(size_t)i * d + t];
                int64_t v = 1;
                while (v < NN[t]) v <<= 1;
                nn[t] = n ? n[(size_t)i * d + t] : 2 * v;   // src/cunfft.cu:296-299
                if (NN[t] < 2 || (NN[t] & 1))
                    throw Error(CUNFFT_ERR_INVALID_ARG, "N_t must be even and >= 2");
                // batches run every axis as one pass (no split axes)
                if (nn[t] < NN[t] || !fft_line_ok(nn[t]))
                    throw Error(CUNFFT_ERR_UNSUPPORTED,
                                "n_t must be 2^a 3^b (a >= 1) in [N_t, 4096]");
                if (nn[t] < W) throw Error(CUNFFT_ERR_INVALID_ARG, "n_t must be >= 2m+2");
            }
            if (M[i] < 1) throw Error(CUNFFT_ERR_INVALID_ARG, "M must be >= 1");
            p.ge = make_geom(d, NN, nn, M[i], m);
            p.M = M[i];
            p.N_total = p.ge.N[0] * p.ge.N[1] * p.ge.N[2];
            p.n_total = p.ge.n[0] * p.ge.n[1] * p.ge.n[2];
            p.roc = make_rolloff(p.ge, p.rocoff);
            p.ft = make_fft_tables(sizeof(T) == 8 ? 64 : 32, d, nn);
            p.x = (double *)alloc(sizeof(double) * d * p.M);
            CUNFFT_HIP(hipMemset(p.x, 0, sizeof(double) * d * p.M));
            p.fhat_in = (C *)alloc(sizeof(C) * R * p.N_total);
            p.fhat_out = (C *)alloc(sizeof(C) * R * p.N_total);
            p.f_out = (C *)alloc(sizeof(C) * R * p.M);
            p.f_in = (C *)alloc(sizeof(C) * R * p.M);
            p.g = (C *)alloc(sizeof(C) * R * p.n_total);
            CUNFFT_HIP(hipMemset(p.fhat_in, 0, sizeof(C) * R * p.N_total));
            CUNFFT_HIP(hipMemset(p.f_in, 0, sizeof(C) * R * p.M));
        }
        if (!set_configs(0, 0))
            throw Error(CUNFFT_ERR_UNSUPPORTED,
                        "a problem's grid is smaller than a gridding tile (n_t >= 256 in 1D, "
                        "32 in 2D, 8 in 3D) or its window is too wide for the tiled kernels");
    }

    // Plan the binnings of every problem for (forward, adjoint) configurations
    // cf, ca; false (nothing changed) if either does not apply to a problem.
    bool set_configs(int cf, int ca)
    {
        for (const Prob &p : probs)
            for (int c : {cf, ca}) {
                Binning probe;
                probe.own_scratch = false;
                const bool lanes = c == kTileVariants;
                if (!probe.plan(p.ge, (int)sizeof(T), lanes ? 1 : 0, lanes ? 0 : c)) return false;
            }
        CUNFFT_HIP(hipDeviceSynchronize());
        drop_graphs();
        tables_ok = false;
        config[0] = cf;
        config[1] = ca;
        for (Prob &p : probs) {
            for (int k = 0; k < 2; ++k) {
                p.bins[k].release();
                p.bins[k].own_scratch = false;
                p.dirty[k] = true;
            }
            for (int k = 0; k < (cf == ca ? 1 : 2); ++k) {
                const int c = config[k];
                const bool lanes = c == kTileVariants;
                p.bins[k].plan(p.ge, (int)sizeof(T), lanes ? 1 : 0, lanes ? 0 : c);
            }
        }
        return true;
    }

    void set_nodes(int i, const double *x, hipStream_t s)
    {
        Prob &p = probs[i];
        CUNFFT_HIP(hipMemcpyAsync(p.x, x, sizeof(double) * d * p.M, hipMemcpyDeviceToDevice, s));
        p.dirty[0] = p.dirty[1] = true;
        tables_ok = false;
    }

    // Bins the problems whose nodes changed (on s), reads the chunk counts
    // back, sizes the adjoint scratch and rebuilds the job tables, FFT pass
    // descriptors and block maps; the graphs are re-captured after.
    void prepare(hipStream_t s)
    {
        if (tables_ok) return;
        const int nb = config[0] == config[1] ? 1 : 2;
        bool rebinned = false;
        for (size_t i = 0; i < probs.size(); ++i) {
            Prob &p = probs[i];
            for (int k = 0; k < nb; ++k) {
                if (!p.dirty[k]) continue;
                p.bins[k].build(p.ge, p.x, s);
                CUNFFT_HIP(hipMemcpyAsync(nch_host + 2 * i + k, p.bins[k].n_chunks,
                                          sizeof(uint32_t), hipMemcpyDeviceToHost, s));
                rebinned = true;
            }
        }
        if (rebinned) CUNFFT_HIP(hipStreamSynchronize(s));
        for (size_t i = 0; i < probs.size(); ++i) {
            Prob &p = probs[i];
            for (int k = 0; k < nb; ++k)
                if (p.dirty[k]) {
                    p.nch[k] = nch_host[2 * i + k];
                    p.dirty[k] = false;
                }
            if (nb == 1) p.nch[1] = p.nch[0];
            // adjoint scratch of the plane-tiled kernels
            if (planes(1)) {
                const Binning &b = p.bins[slot(1)];
                const int64_t need = (int64_t)p.nch[1] * b.slot_cells;
                if (need > p.slot_cap || !p.fs) {
                    CUNFFT_HIP(hipDeviceSynchronize());
                    if (p.slots) (void)hipFree(p.slots);
                    p.slots = nullptr;
                    p.slot_cap = 0;
                    p.slots = (C *)alloc(sizeof(C) * (size_t)R * std::max<int64_t>(need, 1));
                    p.slot_cap = std::max<int64_t>(need, 1);
                    if (!p.fs) p.fs = (C *)alloc(sizeof(C) * R * p.M);
                }
            }
        }
        build_tables();
        drop_graphs();
        tables_ok = true;
    }

    void build_tables()
    {
        const int P = (int)probs.size();
        const int J = P * R;
        std::vector<GridJob<T>> jh[2];
        for (int dir = 0; dir < 2; ++dir) {
            jh[dir].resize(J);
            for (int i = 0; i < P; ++i) {
                const Prob &p = probs[i];
                const Binning &b = p.bins[slot(dir)];
                for (int r = 0; r < R; ++r) {
                    GridJob<T> &J1 = jh[dir][(size_t)i * R + r];
                    J1 = GridJob<T>{};
                    J1.desc = b.desc;
                    J1.xs = b.xs;
                    J1.perm = b.perm;
                    J1.chunks = b.chunks;
                    J1.tile_count = b.tile_count;
                    J1.chunk_start = b.chunk_start;
                    J1.g = p.g + (size_t)r * p.n_total;
                    J1.fs = p.fs ? p.fs + (size_t)r * p.M : nullptr;
                    J1.slots = p.slots ? p.slots + (size_t)r * p.slot_cap : nullptr;
                    J1.f_in = p.f_in + (size_t)r * p.M;
                    J1.f_out = p.f_out + (size_t)r * p.M;
                    J1.n_total = p.n_total;
                    J1.ge = p.ge;
                    J1.bg = bin_geom(b);
                    J1.wc = win_const<T>(p.ge);
                }
            }
        }
        std::vector<FftPass<T>> ph;
        std::vector<BlockMap> mh;
        for (int dir = 0; dir < 2; ++dir) stages[dir].clear();
        // a block map with `blocks(i)` blocks for each job of problem i
        auto add_map = [&](Launch &L, auto blocks) {
            L.map_off = mh.size();
            for (int i = 0; i < P; ++i) {
                const uint32_t nbk = (uint32_t)blocks(i);
                for (int r = 0; r < R; ++r)
                    for (uint32_t k = 0; k < nbk; ++k) mh.push_back(BlockMap{(uint32_t)(i * R + r), k});
            }
            L.nblk = (uint32_t)(mh.size() - L.map_off);
        };
        auto add_fft = [&](int dir, int t) {
            Launch L;
            L.fft_t = t;
            L.pass_off = ph.size();
            L.map_off = mh.size();
            for (int i = 0; i < P; ++i) {
                const Prob &p = probs[i];
                for (int r = 0; r < R; ++r) {
                    const FftLaunch<T> f = fft_pass_setup<T>(
                        p.ge, static_cast<const C *>(p.ft->tw[t]), p.roc, p.rocoff, t, dir == 0,
                        p.fhat_in + (size_t)r * p.N_total, p.fhat_out + (size_t)r * p.N_total,
                        p.g + (size_t)r * p.n_total);
                    const uint32_t job = (uint32_t)(ph.size() - L.pass_off);
                    ph.push_back(f.P);
                    for (int64_t k = 0; k < f.blocks; ++k) mh.push_back(BlockMap{job, (uint32_t)k});
                    L.lds = std::max(L.lds, f.lds);
                    L.max_elems = std::max(L.max_elems, f.P.C * f.P.L);
                }
            }
            L.nblk = (uint32_t)(mh.size() - L.map_off);
            stages[dir].push_back(L);
        };
        auto add_grid = [&](int dir, BatchStage st, auto blocks) {
            Launch L;
            L.st = st;
            L.dir = dir;
            add_map(L, blocks);
            stages[dir].push_back(L);
        };
        // forward: FFT passes d-1 .. 0, then the gather
        for (int k = 0; k < d; ++k) add_fft(0, d - 1 - k);
        add_grid(0, BS_GATHER, [&](int i) { return probs[i].nch[0]; });
        // adjoint: gridding, then FFT passes 0 .. d-1
        if (planes(1)) {
            add_grid(1, BS_COEFFS, [&](int i) { return (probs[i].M + 255) / 256; });
            add_grid(1, BS_SPREAD, [&](int i) { return probs[i].nch[1]; });
            add_grid(1, BS_FOLD, [&](int i) {
                return tile_fold_blocks(d, config[1], probs[i].bins[slot(1)].ntiles, (int)sizeof(T));
            });
        } else {
            add_grid(1, BS_ZERO, [&](int i) { return (probs[i].n_total + 1023) / 1024; });
            add_grid(1, BS_SCATTER_LANES, [&](int i) { return probs[i].nch[1]; });
        }
        for (int k = 0; k < d; ++k) add_fft(1, k);
        for (int dir = 0; dir < 2; ++dir)
            for (const Launch &L : stages[dir])
                if (L.nblk != 0 && mh.size() > 0xffffffffull)
                    throw Error(CUNFFT_ERR_UNSUPPORTED, "batch too large for one launch");
        // upload
        CUNFFT_HIP(hipDeviceSynchronize());
        for (int dir = 0; dir < 2; ++dir) {
            if (jobs[dir]) (void)hipFree(jobs[dir]);
            jobs[dir] = nullptr;
            jobs[dir] = (GridJob<T> *)alloc(sizeof(GridJob<T>) * J);
            CUNFFT_HIP(hipMemcpy(jobs[dir], jh[dir].data(), sizeof(GridJob<T>) * J,
                                 hipMemcpyHostToDevice));
        }
        if (passes) (void)hipFree(passes);
        if (maps) (void)hipFree(maps);
        passes = nullptr;
        maps = nullptr;
        passes = (FftPass<T> *)alloc(sizeof(FftPass<T>) * ph.size());
        maps = (BlockMap *)alloc(sizeof(BlockMap) * mh.size());
        CUNFFT_HIP(hipMemcpy(passes, ph.data(), sizeof(FftPass<T>) * ph.size(), hipMemcpyHostToDevice));
        CUNFFT_HIP(hipMemcpy(maps, mh.data(), sizeof(BlockMap) * mh.size(), hipMemcpyHostToDevice));
    }

    void launch_stages(int dir, hipStream_t s)
    {
        for (const Launch &L : stages[dir]) {
            if (L.fft_t >= 0)
                launch_fft_batch<T>(passes + L.pass_off, maps + L.map_off, L.nblk, L.lds,
                                    L.max_elems, s);
            else
                launch_batch_stage<T>(L.st, jobs[L.dir], maps + L.map_off, L.nblk, d, W,
                                      config[dir], s);
        }
    }

    hipGraphExec_t graph(int dirs)
    {
        if (exec[dirs]) return exec[dirs];
        hipGraph_t gr = nullptr;
        CUNFFT_HIP(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
        try {
            if (dirs & 1) launch_stages(0, cap);
            if (dirs & 2) launch_stages(1, cap);
        } catch (...) {
            (void)hipStreamEndCapture(cap, &gr);
            if (gr) (void)hipGraphDestroy(gr);
            throw;
        }
        CUNFFT_HIP(hipStreamEndCapture(cap, &gr));
        hipGraphExec_t e = nullptr;
        const hipError_t err = hipGraphInstantiate(&e, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        CUNFFT_HIP(err);
        exec[dirs] = e;
        return e;
    }

    void execute(int dirs, hipStream_t s)
    {
        prepare(s);
        CUNFFT_HIP(hipGraphLaunch(graph(dirs), s));
    }

    // The inputs the graphs read (fhat_in, f_in of every problem) filled with
    // nonzero data for timing, and put back after: the node-lane scatter skips
    // its global atomics for zero partial sums, so timing zeroed buffers (as
    // create() leaves them) would flatter it.  Returns the saved copies.
    std::vector<void *> fill_inputs(hipStream_t s)
    {
        std::vector<void *> saved;
        try {
            for (Prob &p : probs) {
                const size_t bh = sizeof(C) * R * p.N_total, bf = sizeof(C) * R * p.M;
                saved.push_back(alloc(bh));
                CUNFFT_HIP(hipMemcpyAsync(saved.back(), p.fhat_in, bh, hipMemcpyDeviceToDevice, s));
                saved.push_back(alloc(bf));
                CUNFFT_HIP(hipMemcpyAsync(saved.back(), p.f_in, bf, hipMemcpyDeviceToDevice, s));
                // words 0x3ff00000: 1.875f, or 1.0000002 as a double
                CUNFFT_HIP(hipMemsetD32Async((hipDeviceptr_t)p.fhat_in, 0x3ff00000, bh / 4, s));
                CUNFFT_HIP(hipMemsetD32Async((hipDeviceptr_t)p.f_in, 0x3ff00000, bf / 4, s));
            }
        } catch (...) {
            restore_inputs(saved, s);
            throw;
        }
        return saved;
    }

    void restore_inputs(std::vector<void *> &saved, hipStream_t s)
    {
        for (size_t i = 0; i < saved.size(); ++i) {
            Prob &p = probs[i / 2];
            const bool fh = i % 2 == 0;
            (void)hipMemcpyAsync(fh ? (void *)p.fhat_in : (void *)p.f_in, saved[i],
                                 sizeof(C) * R * (fh ? p.N_total : p.M), hipMemcpyDeviceToDevice, s);
        }
        (void)hipStreamSynchronize(s);
        for (void *q : saved) (void)hipFree(q);
        saved.clear();
    }

    // Times each configuration (both directions in one configuration, the
    // batch's graphs of one direction each) on nonzero inputs and keeps the
    // fastest per tuned direction.  ms_out: forward times, then adjoint times
    // (-1: not timed).  The batch's inputs are unchanged afterwards; its
    // outputs hold the last timed configuration's results.
    void autotune(int dirs, int reps, float *ms_out, hipStream_t s)
    {
        std::vector<void *> saved = fill_inputs(s);
        try {
            autotune_timed(dirs, reps, ms_out, s);
        } catch (...) {
            restore_inputs(saved, s);
            throw;
        }
        restore_inputs(saved, s);
    }

    void autotune_timed(int dirs, int reps, float *ms_out, hipStream_t s)
    {
        constexpr int NC = kTileVariants + 1;
        if (reps <= 0) reps = 5;
        float ms[2][NC];
        int best[2] = {config[0], config[1]};
        float best_ms[2] = {-1.0f, -1.0f};
        for (int c = 0; c < NC; ++c) {
            ms[0][c] = ms[1][c] = -1.0f;
            if (d == 1 && c != 0) continue;
            if (!set_configs(c, c)) continue;
            prepare(s);
            for (int dir = 0; dir < 2; ++dir) {
                if (!(dirs & (1 << dir))) continue;
                hipGraphExec_t e = graph(1 << dir);
                CUNFFT_HIP(hipGraphLaunch(e, s));
                float t_min = 0;
                for (int r = 0; r < reps; ++r) {
                    CUNFFT_HIP(hipEventRecord(ev[0], s));
                    CUNFFT_HIP(hipGraphLaunch(e, s));
                    CUNFFT_HIP(hipEventRecord(ev[1], s));
                    CUNFFT_HIP(hipEventSynchronize(ev[1]));
                    float t = 0;
                    CUNFFT_HIP(hipEventElapsedTime(&t, ev[0], ev[1]));
                    t_min = r == 0 ? t : std::min(t_min, t);
                }
                ms[dir][c] = t_min;
                if (best_ms[dir] < 0 || t_min < best_ms[dir]) {
                    best_ms[dir] = t_min;
                    best[dir] = c;
                }
            }
        }
        set_configs(best[0], best[1]);
        if (ms_out)
            for (int dir = 0; dir < 2; ++dir)
                for (int c = 0; c < NC; ++c) ms_out[dir * NC + c] = ms[dir][c];
    }
};

}  // namespace cunfft

using namespace cunfft;

struct cunfft_batch {
    int prec = 64;
    void *impl = nullptr;
};

namespace {

template <typename F> int batch_guard(F &&fn)
{
    try {
        fn();
        return CUNFFT_SUCCESS;
    } catch (const Error &e) {
        report_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        report_error("host allocation failed");
        return CUNFFT_ERR_ALLOC;
    } catch (const std::exception &e) {
        report_error(e.what());
        return CUNFFT_ERR_HIP;
    }
}

// Runs fn(BatchImpl<T>&) for the batch's precision.
template <typename F> int with_batch(cunfft_batch *b, F &&fn)
{
    if (!b || !b->impl) return CUNFFT_ERR_INVALID_ARG;
    return batch_guard([&] {
        if (b->prec == 64)
            fn(*static_cast<BatchImpl<double> *>(b->impl));
        else
            fn(*static_cast<BatchImpl<float> *>(b->impl));
    });
}

}  // namespace

extern "C" {

int cunfft_batch_create(cunfft_batch **out, int precision, int dim, int n_problems,
                        const uint_t *N, const uint_t *n, const uint_t *M, int m, int n_rhs)
{
    if (!out) return CUNFFT_ERR_INVALID_ARG;
    *out = nullptr;
    if ((precision != 32 && precision != 64) || dim < 1 || dim > 3 || n_problems < 1 || !N ||
        !M || m < 1 || m > CUNFFT_MAX_CUT_OFF || n_rhs < 1) {
        report_error("cunfft_batch_create: invalid argument");
        return CUNFFT_ERR_INVALID_ARG;
    }
    cunfft_batch *b = new (std::nothrow) cunfft_batch();
    if (!b) return CUNFFT_ERR_ALLOC;
    b->prec = precision;
    const int st = batch_guard([&] {
        if (precision == 64) {
            auto *im = new BatchImpl<double>();
            b->impl = im;
            im->create(dim, N, n, M, n_problems, m, n_rhs);
        } else {
            auto *im = new BatchImpl<float>();
            b->impl = im;
            im->create(dim, N, n, M, n_problems, m, n_rhs);
        }
    });
    if (st != CUNFFT_SUCCESS) {
        cunfft_batch_destroy(b);
        return st;
    }
    *out = b;
    return CUNFFT_SUCCESS;
}

void cunfft_batch_destroy(cunfft_batch *b)
{
    if (!b) return;
    if (b->impl) {
        if (b->prec == 64)
            delete static_cast<BatchImpl<double> *>(b->impl);
        else
            delete static_cast<BatchImpl<float> *>(b->impl);
    }
    delete b;
}

int cunfft_batch_set_nodes_dev(cunfft_batch *b, int problem, const double *x_dev, void *stream)
{
    return with_batch(b, [&](auto &im) {
        if (problem < 0 || problem >= (int)im.probs.size() || !x_dev)
            throw Error(CUNFFT_ERR_INVALID_ARG, "no such problem, or x is NULL");
        im.set_nodes(problem, x_dev, static_cast<hipStream_t>(stream));
    });
}

int cunfft_batch_buffers(cunfft_batch *b, int problem, void **fhat_in, void **f_out, void **f_in,
                         void **fhat_out)
{
    return with_batch(b, [&](auto &im) {
        if (problem < 0 || problem >= (int)im.probs.size())
            throw Error(CUNFFT_ERR_INVALID_ARG, "no such problem");
        const auto &p = im.probs[problem];
        if (fhat_in) *fhat_in = p.fhat_in;
        if (f_out) *f_out = p.f_out;
        if (f_in) *f_in = p.f_in;
        if (fhat_out) *fhat_out = p.fhat_out;
    });
}

int cunfft_batch_execute(cunfft_batch *b, int directions, void *stream)
{
    return with_batch(b, [&](auto &im) {
        if (directions < 1 || directions > 3)
            throw Error(CUNFFT_ERR_INVALID_ARG, "directions must be 1, 2 or 3");
        im.execute(directions, static_cast<hipStream_t>(stream));
    });
}

int cunfft_batch_autotune(cunfft_batch *b, int directions, int reps, float *ms_out, void *stream)
{
    return with_batch(b, [&](auto &im) {
        if (directions < 1 || directions > 3)
            throw Error(CUNFFT_ERR_INVALID_ARG, "directions must be 1, 2 or 3");
        im.autotune(directions, reps, ms_out, static_cast<hipStream_t>(stream));
        CUNFFT_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
    });
}

int cunfft_batch_get_config(const cunfft_batch *b, int direction)
{
    if (!b || !b->impl || (direction != 1 && direction != 2)) return -1;
    return b->prec == 64 ? static_cast<const BatchImpl<double> *>(b->impl)->config[direction - 1]
                         : static_cast<const BatchImpl<float> *>(b->impl)->config[direction - 1];
}

int cunfft_batch_set_config(cunfft_batch *b, int directions, int config)
{
    return with_batch(b, [&](auto &im) {
        if (directions < 1 || directions > 3)
            throw Error(CUNFFT_ERR_INVALID_ARG, "directions must be 1, 2 or 3");
        if (config < 0 || config > kTileVariants || (im.d == 1 && config != 0))
            throw Error(CUNFFT_ERR_INVALID_ARG, "no such gridding configuration");
        const int cf = (directions & 1) ? config : im.config[0];
        const int ca = (directions & 2) ? config : im.config[1];
        if (!im.set_configs(cf, ca))
            throw Error(CUNFFT_ERR_UNSUPPORTED,
                        "the tiled kernels of this configuration do not fit a problem's window");
    });
}

}  // extern "C"
