// Host-side launchers of the gfx950 kernels (implemented next to the kernels).
#pragma once

#include "gvom_internal.h"
#include "gvom_launch_plan.h"

// ---- fills --------------------------------------------------------------------------
hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s);

// ---- scan stage (gvom_scan.hip) ------------------------------------------------------
// Transform, validity, own-voxel sort key per point (reference gvom.py:1122-1163).
// With hist (hwords words, radix_hist_entries): the sort's first-pass counts are accumulated
// into hist, which must be zero.
hipError_t launch_prepare(int dtype, const void *in, void *world, uint32_t *keys, uint32_t *vals,
                          uint8_t *valid, const ScanDesc *sd, int S, int64_t NT, const Geo &g,
                          DevState *st, uint8_t *rsec, uint32_t *hist, int64_t hwords, hipStream_t s);
// Tiled ray-cast scratch (4 x 4 x 1 voxels per 64-byte segment): words per scan, and the
// conversion into the slot tables (which also zeroes the scratch).
int64_t tile_scratch_words(const Geo &g);
hipError_t launch_tile_convert(const Geo &g, const DevState *st, int S, int force_slot, int32_t *scratch,
                               int64_t scratch_words, int32_t *tables, int64_t table_words, hipStream_t s);
// Groups valid rays by azimuth sector for the XCD-aware ray schedule.
hipError_t launch_ray_order(const uint8_t *rsec, int64_t NT, DevState *st, uint32_t *order, hipStream_t s);
// Exact-rule ray march counting passes into the slot tables (gvom.py:1174-1231), or into
// the tiled scratch; table_words = words of `slot_tables` (refused when the launch would
// write past them or its keys would not fit in 31 bits).
hipError_t launch_raycast(int dtype, const void *world, const uint8_t *valid, const ScanDesc *sd,
                          int S, int64_t NT, const Geo &g, const DevState *st, int32_t *slot_tables,
                          int64_t table_words, int block, int rays_per_thread, int use_box, int tiled,
                          int force_slot, const uint32_t *order, int spread, int64_t ring_stride, int patch_rings,
                          hipStream_t s);
// Stable LSD radix sort of (key, val) on the low `bits` (<= 32) key bits, 8-bit passes, one
// kernel per pass.  hist holds RADIX_HIST_SETS x hwords words: the first pass's per-tile
// counts, filled by launch_prepare (the last pass clears them for the next sort), and the
// later passes' published tile counts and digit totals, zeroed by launch_prepare.  A tile
// that gives up waiting sets st->wait_fault.  On return *kout/*vout point at whichever
// buffer pair holds the result.
// scan_shift: keys >> scan_shift is the scan of the batch (the first pass counts each scan's
// valid keys into set 1).
hipError_t radix_sort(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, int64_t n, int bits, int scan_shift,
                      uint32_t *hist, int64_t hwords, DevState *st, uint32_t **kout, uint32_t **vout,
                      hipStream_t s);
int64_t radix_hist_entries(int64_t n);
// words of the segment kernels' per-block counts (launch_segments' blk) for n sorted keys
int64_t segment_block_words(int64_t n);
// Segments of equal keys in the sorted array, per-scan bounds and slot cell counts, and the
// per-chunk moment partials, in one launch; hist / hwords as given to radix_sort (set 1
// holds the scans' valid-key counts and the blocks' published head counts).
hipError_t launch_segments(int dtype, const void *world, const uint32_t *skeys, const uint32_t *svals, int64_t n,
                           const ScanDesc *sd, int S, const Geo &g, DevState *st, uint32_t *hist, int64_t hwords,
                           uint32_t *blk, int64_t blk_words, int32_t *seg_start, uint32_t *seg_key, Moment *part,
                           hipStream_t s);
// Per-segment moment sums of the chunk partials + slot table occupancy marks +
// hit/total/min-height records.  All scan-stage writes to slot tables and records go to the
// buffers new_buf(flip, slot) of the slots the batch fills.
hipError_t launch_mom_seg(const int32_t *seg_start, const uint32_t *seg_key, int64_t seg_cap, const Geo &g,
                          DevState *st, int32_t *slot_tables, SlotRec *slot_recs, int64_t ncap, const Moment *part,
                          Moment *mom, hipStream_t s);
// Neighbourhood mean/covariance per occupied voxel (gvom.py:1250-1387).
hipError_t launch_metrics(const uint32_t *seg_key, const Moment *mom, int64_t seg_cap, const Geo &g,
                          DevState *st, const int32_t *slot_tables, SlotRec *slot_recs, int64_t ncap, int S,
                          hipStream_t s);
// One-wave commit of the S scans: slot origins, latest ego, buffer bits, cell counts, ring
// indices and the next combine's source shifts.
hipError_t launch_commit_process(DevState *st, const ScanDesc *sd, const Geo &g, int S, hipStream_t s);


// ---- combine stage (gvom_combine.hip) ----------------------------------------------------
// Per column it also writes the occupied-voxel count, the occupancy words (bit z of word
// z / 64), the inferred height and, for columns without occupancy, the height.
// zero_page: zero_page_words() zeroed int32 the row kernel reads instead of off-map words
// (xy_size % 4 == 0; null selects the general kernel).
int64_t zero_page_words();
hipError_t launch_comb_tables(const Geo &g, DevState *st, const int32_t *slot_tables, const int32_t *zero_page,
                              int32_t *comb_tables, int32_t *col_count, uint64_t *occ_bits, double *height,
                              double *inferred, hipStream_t s);
int64_t col_chunks(int64_t X);
hipError_t launch_scan_columns(int32_t *col_count, int64_t count_words, int64_t X, int32_t *chunk_tot,
                               int64_t chunk_words, DevState *st, hipStream_t s);
// Cell ids + compact list, then the record merge + eigenvalues + heights of occupied columns.
hipError_t launch_comb_records(const Geo &g, DevState *st, const int32_t *slot_tables,
                               const SlotRec *slot_recs, int64_t ncap, int32_t *comb_tables,
                               CombRec *comb_recs, int64_t ccap, const int32_t *col_base,
                               const int32_t *chunk_tot, const uint64_t *occ_bits, uint2 *occ_list, double *height,
                               hipStream_t s);
hipError_t launch_maps2d(const Geo &g, const DevState *st, const double *height, const double *inferred,
                         const int32_t *comb_tables, const CombRec *comb_recs, int64_t ccap, double *slope_x,
                         double *slope_y, double *rough, double *delta, int32_t *pos, int32_t *neg, int32_t *vis,
                         uint8_t *host_maps, hipStream_t s);


// ---- export (gvom_combine.hip) -----------------------------------------------------------
// Internal z-contiguous table -> reference layout and encoding.
hipError_t launch_export_table(const Geo &g, const int32_t *table, int32_t *out_ref, hipStream_t s);
hipError_t launch_debug_voxels(const Geo &g, const DevState *st, const int32_t *comb_tables,
                               const CombRec *comb_recs, int64_t ccap, int parity, float *out,
                               hipStream_t s);
hipError_t launch_debug_height(const Geo &g, const DevState *st, int parity, const double *height,
                               const double *rough, const double *sx, const double *sy,
                               const double *delta, float *out7, float *out3, hipStream_t s);


// ---- image kernels the reference defines but never calls (gvom_image.hip) -------------------
hipError_t launch_expand_binary(const int32_t *in, int32_t *out, int XY, double r, hipStream_t s);
hipError_t launch_lowpass_binary(const int32_t *in, int32_t *out, int XY, int filter_size, double fraction,
                                 hipStream_t s);
hipError_t launch_convolve(const double *in, const double *kern, double *out, int XY, int kernel_size,
                           hipStream_t s);
