// CPU check of the launch plan (g-vom_amd/csrc/gvom_launch_plan.h) and of the launchers'
// host-side refusals, built by tests/test_launch_plan.py with hipcc (host code only) against
// libgvom_hip.so.  Every launcher call below is refused before it touches the device, so the
// program needs no GPU.  Prints one line per failed check and "ok" at the end.
#include <stdio.h>

#include "gvom_kernels.h"

namespace P = gvom_plan;

static int fails = 0;
#define CHECK(cond, ...)                       \
    do {                                       \
        if (!(cond)) {                         \
            ++fails;                           \
            if (fails < 20) {                  \
                printf("FAIL %s: ", #cond);    \
                printf(__VA_ARGS__);           \
                printf("\n");                  \
            }                                  \
        }                                      \
    } while (0)

int main() {
    const int xys[] = {4, 5, 7, 8, 13, 16, 37, 64, 100, 255, 256, 333, 512, 1000, 1024, 2048, 4096, 8192};
    const int zs[] = {4, 5, 13, 16, 32, 48, 64, 100, 128, 255, 256};
    long long cases = 0;
    for (int xy : xys)
        for (int z : zs)
            for (int B = 1; B <= P::MAX_BUFFER; ++B) {
                if (!P::params_in_domain(xy, z, B)) continue;
                ++cases;
                const int64_t V = (int64_t)xy * xy * z, X = (int64_t)xy * xy;
                const int smax = B < P::MAX_BATCH ? B : P::MAX_BATCH;
                for (int S = 1; S <= smax; ++S) {
                    // the direct cast's keys fit for every batch the API admits
                    CHECK(P::raycast_keys_fit(V, xy, z, S, false), "xy %d z %d B %d S %d", xy, z, B, S);
                    // and its writes stay inside the slot tables the API allocates
                    CHECK(P::slot_table_footprint(V, B, -1) <= P::alloc_slot_table_words(V, B), "xy %d z %d B %d", xy, z, B);
                    if (P::tiled_cast_fits(xy, z, B)) {
                        CHECK(P::raycast_keys_fit(V, xy, z, S, true), "tiled xy %d z %d B %d S %d", xy, z, B, S);
                        CHECK((int64_t)S * P::tile_words(xy, z) <= P::alloc_tile_scratch_words(xy, z, B), "tiled scratch");
                    }
                }
                // a tile of 4 x 4 columns covers the map, and its words are 64-byte aligned
                CHECK(P::tile_words(xy, z) >= V && P::tile_words(xy, z) % 16 == 0, "tile words xy %d z %d", xy, z);
                // the column scan touches the counts and chunk totals the API allocates
                CHECK(P::col_scan_count_words(X) <= P::alloc_col_count_words(X), "col counts xy %d", xy);
                CHECK(P::col_scan_chunk_words(X) + 1 <= P::alloc_chunk_tot_words(X), "chunk totals xy %d", xy);
                CHECK(P::col_chunks(X) * P::COLSCAN >= X && (P::col_chunks(X) - 1) * P::COLSCAN < X, "col chunks xy %d", xy);
            }
    // the domain: V (B + 1) < 2^31 with the sides' limits
    CHECK(P::params_in_domain(256, 64, 30), "headline shape with 30 slots");
    CHECK(!P::params_in_domain(8192, 256, 1), "V = 2^34");
    CHECK(!P::params_in_domain(3, 64, 1) && !P::params_in_domain(256, 3, 1) && !P::params_in_domain(256, 64, 31), "limits");
    // 1024 x 1024 x 48 with 30 slots: in the domain; its 60 physical slot buffers span 3.0e9
    // words, so keys formed from the physical buffer index would pass 2^31 -- the batch-relative
    // keys stay below it for every batch
    {
        const int64_t V = 1024LL * 1024 * 48;
        CHECK(P::params_in_domain(1024, 48, 30), "1024 x 1024 x 48, B 30");
        CHECK(P::alloc_slot_table_words(V, 30) >= P::KEY_END, "physical buffer span exceeds 2^31");
        for (int S = 1; S <= 30; ++S) CHECK(P::raycast_keys_fit(V, 1024, 48, S, false), "S %d", S);
    }
    // segment counts and radix histograms for key counts around the tiles
    const long long ns[] = {0, 1, 2047, 2048, 2049, 4095, 131072, 1000000, 2147483000LL};
    for (long long n : ns) {
        const int64_t nb = n > 0 ? (n + P::SEG_TILE - 1) / P::SEG_TILE : 1;
        for (int S = 1; S <= P::MAX_BATCH; ++S)
            CHECK(P::segment_block_footprint(n, S) <= P::segment_block_words(n), "segments n %lld S %d", n, S);
        CHECK(P::segment_block_words(n) >= (nb + 1) * (2 + 3 * P::MAX_BATCH), "segments n %lld", n);
        const int64_t tiles = n > 0 ? (n + P::RS_TILE - 1) / P::RS_TILE : 1;
        CHECK(P::radix_hist_words(n) >= 256 * tiles + 256, "radix n %lld", n);
        // four sets: the first pass's tile histogram, then tile counts + digit totals of up
        // to three lookback passes (32-bit keys)
        CHECK(P::radix_hist_alloc_words(n) == P::RADIX_HIST_SETS * P::radix_hist_words(n) && P::RADIX_HIST_SETS >= 4,
              "radix sets n %lld", n);
    }
    // the launchers refuse undersized buffers before any device work (null pointers: nothing is
    // dereferenced on a refused launch)
    {
        Geo g{};
        g.xy_res = 0.4;
        g.z_res = 0.2;
        g.XY = 256;
        g.Z = 64;
        g.V = 256LL * 256 * 64;
        g.X = 256LL * 256;
        g.B = 8;
        g.VB = 24;
        const int64_t n = 131072;
        const int64_t blk = P::segment_block_footprint(n, 3);
        const int64_t hw = P::radix_hist_words(n);
        CHECK(launch_segments(GVOM_F32, nullptr, nullptr, nullptr, n, nullptr, 3, g, nullptr, nullptr, hw, nullptr,
                              blk - 1, nullptr, nullptr, nullptr, nullptr) == hipErrorInvalidValue,
              "segments one word short");
        CHECK(launch_segments(GVOM_F32, nullptr, nullptr, nullptr, n, nullptr, 33, g, nullptr, nullptr, hw, nullptr, blk,
                              nullptr, nullptr, nullptr, nullptr) == hipErrorInvalidValue,
              "segments with 33 scans");
        CHECK(launch_segments(GVOM_F32, nullptr, nullptr, nullptr, n, nullptr, 3, g, nullptr, nullptr, hw - 1, nullptr,
                              blk, nullptr, nullptr, nullptr, nullptr) == hipErrorInvalidValue,
              "segments with a histogram set one word short");
        uint32_t *kout = nullptr, *vout = nullptr;
        CHECK(radix_sort(nullptr, nullptr, nullptr, nullptr, n, 32, 27, nullptr, P::radix_hist_words(n) - 1, nullptr,
                         &kout, &vout, nullptr) == hipErrorInvalidValue, "radix histogram one word short");
        CHECK(radix_sort(nullptr, nullptr, nullptr, nullptr, n, 32, 24, nullptr, P::radix_hist_words(n), nullptr, &kout,
                         &vout, nullptr) == hipErrorInvalidValue, "radix sort with more than 32 scans");
        const int64_t tw = P::tile_words(g.XY, g.Z);
        CHECK(launch_raycast(GVOM_F32, nullptr, nullptr, nullptr, 2, n, g, nullptr, nullptr, 2 * tw - 1, 256, 1, 0, 1,
                             -1, nullptr, 1, 0, 0, nullptr) == hipErrorInvalidValue, "tiled cast into a short scratch");
        CHECK(launch_raycast(GVOM_F32, nullptr, nullptr, nullptr, 2, n, g, nullptr, nullptr,
                             P::alloc_slot_table_words(g.V, g.B) - 1, 256, 1, 0, 0, -1, nullptr, 1, 0, 0, nullptr) ==
                  hipErrorInvalidValue, "direct cast into short slot tables");
        CHECK(launch_raycast(GVOM_F32, nullptr, nullptr, nullptr, 1, n, g, nullptr, nullptr, g.V - 1, 256, 1, 0, 0, 0,
                             nullptr, 1, 0, 0, nullptr) == hipErrorInvalidValue, "autotune cast into a short table");
        CHECK(launch_tile_convert(g, nullptr, 3, -1, nullptr, 3 * tw - 1, nullptr, P::alloc_slot_table_words(g.V, g.B),
                                  nullptr) == hipErrorInvalidValue, "conversion from a short scratch");
        CHECK(launch_tile_convert(g, nullptr, 3, -1, nullptr, 3 * tw, nullptr, P::alloc_slot_table_words(g.V, g.B) - 1,
                                  nullptr) == hipErrorInvalidValue, "conversion into short slot tables");
        CHECK(launch_scan_columns(nullptr, g.X - 1, g.X, nullptr, P::col_chunks(g.X), nullptr, nullptr) ==
                  hipErrorInvalidValue, "column scan over short counts");
        CHECK(launch_scan_columns(nullptr, g.X, g.X, nullptr, P::col_chunks(g.X) - 1, nullptr, nullptr) ==
                  hipErrorInvalidValue, "column scan with short chunk totals");
        // 2048 x 2048 x 64 with 6 slots is in the domain; 8 scans form keys up to 8 V - 1 =
        // 2^31 - 1, a batch of 9 would pass 2^31 and is refused by the launcher itself (the
        // API admits at most buffer_size scans per batch)
        Geo h = g;
        h.XY = 2048;
        h.Z = 64;
        h.V = 2048LL * 2048 * 64;
        h.X = 2048LL * 2048;
        h.B = 6;
        CHECK(P::params_in_domain(h.XY, h.Z, h.B), "2048 x 2048 x 64, B 6");
        CHECK(P::raycast_keys_fit(h.V, h.XY, h.Z, 8, false) && !P::raycast_keys_fit(h.V, h.XY, h.Z, 9, false),
              "2048 x 2048 x 64 direct keys");
        CHECK(launch_raycast(GVOM_F32, nullptr, nullptr, nullptr, 9, n, h, nullptr, nullptr,
                             P::alloc_slot_table_words(h.V, 9), 256, 1, 0, 0, -1, nullptr, 1, 0, 0, nullptr) ==
                  hipErrorInvalidValue, "keys past 2^31 refused");
    }
    printf("%lld shapes\n", cases);
    if (fails) {
        printf("%d failed\n", fails);
        return 1;
    }
    printf("ok\n");
    return 0;
}
