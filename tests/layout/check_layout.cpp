// check_layout.cpp -- host check of every plane-tiled LDS layout the library
// instantiates (grid_tile.hpp), compiled for the host only and run by
// tests/test_layout.py.  For each (precision, d, W, tile shape) that
// fits_all() admits, for the gather layout (one plane step) and the spread
// layout (the spread's wavefront count as plane step):
//   coverage  every window item (plane q < G, row r < R, unit u < NU) sits
//             in exactly one lane of the lane table, no lane twice;
//   banks     within each 16-lane group the items' 16-byte units are
//             distinct mod 16 (ds_read_b128 rule) and within each 8-lane
//             half distinct mod 8 (ds_write_b128 rule) wherever the layout's
//             own pitch search found a conflict-free pitch; elsewhere the
//             conflict degree is printed beside the search's estimate.  The
//             check is independent of the node origin, which shifts every
//             lane of an instruction alike;
//   bounds    for every node origin in the tile, every plane a wavefront
//             touches (gather: clamped planes; spread: residue planes, with
//             the dummy unit past the window; held-unit 3D spread: planes
//             below P0, the dummy unit beyond), every access lies in the
//             padded tile, rows stay within the row length RL <= pitch, and
//             dummy units lie in the spare row ends or past the cells;
//   slots     the destination-blocked slot blocks partition the padded tile
//             exactly once, in order, within SC cells.
// Prints one line per layout and "LAYOUT_OK" when every check held.
#include <cstdio>
#include <cstdlib>
#include <set>
#include <vector>

#include "grid_tile.hpp"

using namespace cunfft;

static int g_fail = 0;
#define CHECK(cond, ...)                                         \
    do {                                                         \
        if (!(cond)) {                                           \
            ++g_fail;                                            \
            if (g_fail < 50) {                                   \
                printf("FAIL %s:%d: ", __FILE__, __LINE__);     \
                printf(__VA_ARGS__);                             \
                printf("\n");                                    \
            }                                                    \
        }                                                        \
    } while (0)

struct Item {
    int q, r, u, lane, inst;
};

template <typename L> std::vector<Item> items()
{
    std::vector<Item> out;
    for (int i = 0; i < L::NI; ++i)
        for (int lane = 0; lane < 64; ++lane) {
            const int e = L::lanes.v[i * 64 + lane];
            if (e == 0xFFFF) continue;
            out.push_back({e >> 10, (e >> 5) & 31, e & 31, lane, i});
        }
    return out;
}

// cells from the window origin of an item, as lane_item() forms it
template <typename L> int item_off(const Item &it) { return it.q * L::PSV * L::PP + it.r * L::RP + L::CPL * it.u; }

template <typename L> int check_layout(const char *name, int mode)   // mode 0 gather, 1 spread, 2 held-unit spread
{
    const int CB = L::CB, CPL = L::CPL;
    const std::vector<Item> its = items<L>();
    // coverage
    std::set<int> seen;
    for (const Item &it : its) {
        CHECK(it.q < L::G && it.r < L::R && it.u < L::NU, "%s item out of range", name);
        CHECK(seen.insert((it.q * 64 + it.r) * 64 + it.u).second, "%s item twice", name);
    }
    CHECK((int)seen.size() == L::G * L::R * L::NU, "%s items %zu of %d", name, seen.size(),
          L::G * L::R * L::NU);
    // banks (16-byte units; the pitch keeps rows on whole units)
    CHECK((L::RP * CB) % 16 == 0 && (L::PP * CB) % 16 == 0, "%s pitch not on 16-byte units", name);
    int worst_read = 0, worst_write = 0;
    for (int i = 0; i < L::NI; ++i)
        for (int g = 0; g < 4; ++g) {
            int cnt16[16] = {}, cnt8[2][8] = {};
            for (const Item &it : its) {
                if (it.inst != i || it.lane / 16 != g) continue;
                const int unit = item_off<L>(it) * CB / 16;
                const int c16 = ++cnt16[unit & 15];
                const int c8 = ++cnt8[(it.lane & 15) / 8][unit & 7];
                worst_read = c16 > worst_read ? c16 : worst_read;
                worst_write = c8 > worst_write ? c8 : worst_write;
            }
        }
    const int est = (max_group_load(L::G, L::R, L::NU, L::PSV, L::P1, L::RP, CB) + L::NG - 1) / L::NG;
    // where the pitch search found a conflict-free pitch the lanes must be
    // conflict-free; elsewhere the degree is reported (a cost, not a fault)
    if (est <= 1) CHECK(worst_read == 1 && worst_write == 1, "%s conflicts %d/%d", name, worst_read, worst_write);
    // bounds over all node origins
    const int T0 = L::T0, T1 = L::T1, TL = L::TL;
    const long cells = (long)L::CELLS;
    const long tile_bytes = cells * CB + (long)L::DUMMY_UNITS * 16;
    const int NW = L::PSV;
    for (int lo0 = 0; lo0 < T0; ++lo0)
        for (int lo1 = 0; lo1 < T1; ++lo1)
            for (int loL = 0; loL < TL; ++loL) {
                const int par = CPL == 2 ? (loL & 1) : 0;
                const long nbase = (long)lo0 * L::PP + (long)lo1 * L::RP + loL - par;
                for (const Item &it : its) {
                    const int col = loL - par + CPL * it.u;   // first cell of the unit in the row
                    CHECK(col >= 0 && col + CPL <= L::RL && L::RL <= L::RP, "%s column %d", name, col);
                    if (L::R > 1) CHECK(lo1 + it.r < L::P1, "%s row %d", name, lo1 + it.r);
                    for (int w = 0; w < (mode == 0 ? 1 : NW); ++w) {
                        const int a0 = mode == 0 ? 0 : ((w - lo0) % NW + NW) % NW;
                        const int ngrp = mode == 2 ? L::OWN : L::NGRP;
                        for (int gq = 0; gq < ngrp; ++gq) {
                            int plane;   // plane index relative to lo0
                            bool dummy = false;
                            const int Wt = (int)L::P0 - T0 + 1;   // taps per axis, W
                            if (mode == 0) {
                                const int a = gq * L::G + it.q;
                                plane = a < Wt ? a : Wt - 1;   // planes past the window: weight 0
                            } else if (mode == 1) {
                                const int a = a0 + NW * (gq * L::G + it.q);
                                dummy = a >= Wt;
                                plane = a;
                            } else {
                                const int a = a0 + NW * gq;
                                dummy = lo0 + a >= L::P0;
                                plane = a;
                            }
                            if (dummy) {
                                CHECK(L::NEED_DUMMY, "%s dummy access without dummy units", name);
                                continue;
                            }
                            // the kernels add the plane (which includes the item's plane in
                            // its group) to the item's offset without that plane
                            const long cell = nbase + (long)plane * L::PP + item_off<L>(it) - (long)it.q * NW * L::PP;
                            const long byte = cell * CB;
                            CHECK(lo0 + plane < L::P0, "%s plane %d (lo0 %d)", name, lo0 + plane, lo0);
                            CHECK(byte >= 0 && byte + 16 <= cells * CB, "%s byte %ld of %ld", name, byte, cells * CB);
                        }
                    }
                }
            }
    // dummy units
    if (L::NEED_DUMMY) {
        for (int lane = 0; lane < 64; ++lane) {
            const long b = L::SPARE ? (long)(lane * L::RP + L::RL) * CB : (long)(L::CELLS / CPL + lane) * 16;
            CHECK(b >= 0 && b + 16 <= tile_bytes, "%s dummy byte %ld", name, b);
            if (L::SPARE) CHECK((L::RP - L::RL) * CB >= 16 && lane < L::P0 * L::P1, "%s spare", name);
        }
    }
    // slot blocks partition the padded tile in order
    long off = 0;
    std::vector<char> hit((size_t)L::P0 * L::P1 * L::PL, 0);
    for (int q0 = 0; q0 < L::Q0; ++q0)
        for (int q1 = 0; q1 < L::Q1; ++q1)
            for (int qL = 0; qL < L::QL; ++qL) {
                const int E0 = L::ext(L::T0, L::P0, q0), E1 = L::ext(L::T1, L::P1, q1),
                          EL = L::ext(L::TL, L::PL, qL);
                CHECK(L::block_off(q0, q1, qL) == off, "%s block (%d,%d,%d) at %d, expected %ld", name, q0,
                      q1, qL, L::block_off(q0, q1, qL), off);
                CHECK(E0 > 0 && E1 > 0 && EL > 0, "%s empty block", name);
                for (int a = 0; a < E0; ++a)
                    for (int b = 0; b < E1; ++b)
                        for (int c = 0; c < EL; ++c) {
                            const size_t cell = ((size_t)(q0 * L::T0 + a) * L::P1 + q1 * L::T1 + b) * L::PL +
                                                qL * L::TL + c;
                            CHECK(cell < hit.size() && !hit[cell], "%s slot cell twice", name);
                            if (cell < hit.size()) hit[cell] = 1;
                        }
                off += (long)E0 * E1 * EL;
            }
    CHECK(off == (long)L::P0 * L::P1 * L::PL && off <= L::SC, "%s slots %ld of %d", name, off, L::SC);
    printf("%-34s G %d NI %d pitch %3d  conflicts read %d write %d (estimate %d)%s\n", name, L::G, L::NI,
           L::RP, worst_read, worst_write, est, L::NEED_DUMMY ? (L::SPARE ? "  dummy: spare" : "  dummy: appended") : "");
    return 0;
}

template <typename T, int D, int W, int V> void check_one()
{
    if constexpr (fits_all<T, D, W, V>()) {
        using LG = TileLayout<T, D, W, 1, V>;
        constexpr int NWS = Waves<T, D, W, V>::spread;
        using LS = TileLayout<T, D, W, NWS, V>;
        char name[96];
        snprintf(name, sizeof name, "%s %dD W=%2d V=%d gather", sizeof(T) == 4 ? "fp32" : "fp64", D, W, V);
        check_layout<LG>(name, 0);
        snprintf(name, sizeof name, "%s %dD W=%2d V=%d spread/%d", sizeof(T) == 4 ? "fp32" : "fp64", D, W, V, NWS);
        check_layout<LS>(name, 1);
        if constexpr (D == 3 && LS::G == 1) {
            snprintf(name, sizeof name, "%s %dD W=%2d V=%d held/%d", sizeof(T) == 4 ? "fp32" : "fp64", D, W, V, NWS);
            check_layout<LS>(name, 2);
        }
    }
}

template <typename T, int D, int V> void check_ws()
{
    check_one<T, D, 4, V>();
    check_one<T, D, 6, V>();
    check_one<T, D, 8, V>();
    check_one<T, D, 10, V>();
    check_one<T, D, 12, V>();
    check_one<T, D, 14, V>();
    check_one<T, D, 16, V>();
    check_one<T, D, 18, V>();
}

int main()
{
    check_ws<float, 2, 0>();
    check_ws<float, 2, 1>();
    check_ws<float, 2, 2>();
    check_ws<float, 3, 0>();
    check_ws<float, 3, 1>();
    check_ws<float, 3, 2>();
    check_ws<double, 2, 0>();
    check_ws<double, 2, 1>();
    check_ws<double, 2, 2>();
    check_ws<double, 3, 0>();
    check_ws<double, 3, 1>();
    check_ws<double, 3, 2>();
    if (g_fail) {
        printf("%d layout checks failed\n", g_fail);
        return 1;
    }
    printf("LAYOUT_OK\n");
    return 0;
}
