// Launch geometry of the kernels that work in fixed tiles: their tile sizes, the words each
// launch touches in the buffers it writes, and the sizes the C-ABI allocates for those
// buffers.  The launchers refuse a launch whose footprint exceeds the buffer they were given,
// so a tile or grid change that outgrows a buffer becomes an error on the host instead of an
// out-of-bounds access on the device.  Plain C++ with no HIP dependency: the CPU tests
// compile it with g++ and sweep the whole parameter domain.
#pragma once

#include <stdint.h>

namespace gvom_plan {

constexpr int RS_TILE = 2048;     // keys per radix-sort tile (1024 threads x 2)
constexpr int SEG_TILE = 2048;    // keys per segment-kernel block (256 threads x 8)
constexpr int COLSCAN = 4096;     // columns per column-scan block (1024 threads x 4)
constexpr int MAX_BATCH = 32;     // scans per process batch
constexpr int MAX_BUFFER = 30;    // ring slots
constexpr int64_t KEY_END = (int64_t)1 << 31;   // ray keys: the sign bit marks a dead lane

constexpr int64_t blocks(int64_t n, int64_t per) {
    return (n + per - 1) / per > 0 ? (n + per - 1) / per : 1;
}

// ---- radix sort and segments: RADIX_HIST_SETS sets of radix_hist_words(n) words.  Set 0:
// the first pass's per-tile histogram.  Set p >= 1: the tiles' published digit counts of
// pass p (at most three later passes: keys have <= 32 bits), the pass's digit totals, then
// 64 words (set 1: valid keys per scan, counted by the first pass) and one word per
// segment block (set 1: the segment kernel's published head counts).
constexpr int RADIX_HIST_SETS = 4;
constexpr int64_t radix_totals_at(int64_t n) { return 256 * blocks(n, RS_TILE); }
constexpr int64_t radix_scan_counts_at(int64_t n) { return radix_totals_at(n) + 256; }
constexpr int64_t segment_flags_at(int64_t n) { return radix_scan_counts_at(n) + 64; }
constexpr int64_t radix_hist_words(int64_t n) { return segment_flags_at(n) + blocks(n, SEG_TILE); }
constexpr int64_t radix_hist_alloc_words(int64_t n) { return RADIX_HIST_SETS * radix_hist_words(n); }

// ---- segment kernels: per block of SEG_TILE keys, W = 2 + 3 S counters (heads, valid keys,
// then per scan of the batch in-map heads, halo heads, valid keys), plus one row of totals;
// the C-ABI sizes the buffer for the largest batch
constexpr int64_t segment_block_footprint(int64_t n, int S) { return (blocks(n, SEG_TILE) + 1) * (2 + 3 * (int64_t)S); }
constexpr int64_t segment_block_words(int64_t n) { return segment_block_footprint(n, MAX_BATCH); }

// ---- column scan over X per-column counts: block b covers columns [b COLSCAN, (b + 1) COLSCAN)
// with 16-byte loads and stores where all four columns exist and single words at the tail, so
// it touches exactly the X counts; it writes one chunk total per block.
inline int64_t col_chunks(int64_t X) { return (X + COLSCAN - 1) / COLSCAN; }
inline int64_t col_scan_count_words(int64_t X) { return X; }
inline int64_t col_scan_chunk_words(int64_t X) { return col_chunks(X); }
// what the C-ABI allocates: counts (16 spare words) and chunk totals (+1 for the grand total)
inline int64_t alloc_col_count_words(int64_t X) { return X + 16; }
inline int64_t alloc_chunk_tot_words(int64_t X) { return col_chunks(X) + 1; }

// ---- tiled ray-cast scratch: 4 x 4 columns x 1 z per 16 words, one table per batch scan
inline int64_t tile_words(int32_t XY, int32_t Z) {
    const int64_t t = (XY + 3) / 4;
    return t * t * Z * 16;
}
inline int tile_tables(int B) { return B < MAX_BATCH ? B : MAX_BATCH; }
inline int64_t alloc_tile_scratch_words(int32_t XY, int32_t Z, int B) { return tile_words(XY, Z) * tile_tables(B); }
// the tiled cast's keys: scan s of a batch counts at s * tile_words + tiled index
inline bool tiled_cast_fits(int32_t XY, int32_t Z, int B) {
    return alloc_tile_scratch_words(XY, Z, B) <= KEY_END;
}

// ---- direct ray cast into the slot tables: scan s of a batch forms keys s * V + voxel, and a
// run head adds the offset of its scan's slot buffer (64-bit) when it updates the table
inline int64_t raycast_key_end(int64_t V, int32_t XY, int32_t Z, int S, bool tiled) {
    return (int64_t)S * (tiled ? tile_words(XY, Z) : V);
}
inline bool raycast_keys_fit(int64_t V, int32_t XY, int32_t Z, int S, bool tiled) {
    return raycast_key_end(V, XY, Z, S, tiled) <= KEY_END;
}

// ---- ring-patch ray order: patches of 64 / pr azimuths x pr rings of `stride` points
inline bool ray_patch_valid(int64_t stride, int pr) {
    if (pr < 1 || pr > 64 || (pr & (pr - 1)) != 0) return false;
    const int pa = 64 / pr;
    return stride >= pa && stride % pa == 0 && stride < ((int64_t)1 << 40);
}

// ---- slot tables: two physical buffers per ring slot
inline int64_t alloc_slot_table_words(int64_t V, int B) { return 2 * (int64_t)B * V; }
// the buffers a launch with force_slot (autotuning: every scan into buffer force_slot) or into
// the ring (new_buf of the slots the batch fills, all < 2 B) may write
inline int64_t slot_table_footprint(int64_t V, int B, int force_slot) {
    return force_slot >= 0 ? ((int64_t)force_slot + 1) * V : 2 * (int64_t)B * V;
}

// ---- the domain gvom_create accepts (gvom.py's constructor has no limits; these keep every
// index of the path in 32 bits)
inline bool params_in_domain(int32_t XY, int32_t Z, int B) {
    if (XY < 4 || XY > 8192 || Z < 4 || Z > 256 || B < 1 || B > MAX_BUFFER) return false;
    const int64_t V = (int64_t)XY * XY * Z;
    return V * (int64_t)(B + 1) < KEY_END;
}

}  // namespace gvom_plan
