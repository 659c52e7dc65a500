// Internal definitions shared by the gfx950 kernels and the C-ABI layer.
//
// HBM layout (z-contiguous): voxel (x, y, z) of an xy*xy*z map lives at
//   lin = (x * xy + y) * z_size + z,
// so a column c = x * xy + y is z_size consecutive int32 words and c is also the
// index of that column in every [x, y] C-order 2D map.  The reference stores
// v = x + y*xy + z*xy*xy (gvom.py:1167); the layout is internal and converted back
// on state export.
//
// Table encoding (one int32 per voxel, slot tables and combined tables alike):
//   t >= 0          : no hit; t ray passes (t == 0 is "unknown")
//   OCC | id        : occupied, compact record id (sign bit set)
// which is the reference's encoding (>= 0 id, -1 unknown, -t-1 free) with the
// sign flipped, chosen so that ray casting can count passes straight into the
// table and occupancy never needs a dense rewrite.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/gvom_abi.h"

#define GVOM_MAX_BUFFER 30
#define GVOM_MAX_BATCH 32
#define GVOM_OCC 0x80000000u
#define GVOM_NOKEY 0xFFFFFFFFu

// Per-scan description, in device memory (the batch kernels index it by scan).
struct ScanDesc {
    double T[12];      // rows 0..2 of the 4x4 sensor-to-world matrix (exact copy of the input values)
    double ego[3];
    int32_t t_mode;    // GVOM_NO_TRANSFORM, GVOM_F32 or GVOM_F64
    int32_t pad;
    int64_t offset;    // first point of the scan in the flat point array
    int64_t count;
};

// Ring and combine state, in device memory, so that a captured launch sequence
// replays correctly (slot choice, origins and ping-pong parity are read on device).
struct DevState {
    int32_t next_slot;     // gvom.py buffer_index
    int32_t last_slot;     // gvom.py last_buffer_index
    int32_t filled_mask;   // slots holding a scan
    int32_t parity;        // combined buffer holding the latest combined map
    int32_t have_last;     // a previous combined map exists
    int32_t nseg;          // segments of the current process batch
    int32_t nvalid;        // sorted keys != NOKEY in the current process batch
    int32_t comb_total;    // cell count of the combine in flight
    int32_t comb_cells[2];
    int32_t slot_cells[GVOM_MAX_BUFFER + 2];
    // Every ring slot has two physical table/record buffers, 2 slot and 2 slot + 1; bit slot of
    // `flip` selects the one holding the slot's scan.  A scan stage writes the other one and
    // its commit flips the bit, so a combine may still read the ring while the next scan
    // stage runs.
    int32_t flip;
    int32_t scan_cells[GVOM_MAX_BATCH];      // in-map cells of each scan of the batch in flight
    int32_t first_seg[GVOM_MAX_BATCH];
    int32_t inb_end[GVOM_MAX_BATCH];
    int32_t scan_pos[GVOM_MAX_BATCH];        // first sorted position of each scan's valid keys
    int32_t scan_valid[GVOM_MAX_BATCH];      // valid keys per scan
    int32_t chunk_off[GVOM_MAX_BATCH + 1];   // moment chunks before each scan (scan-relative chunking)
    // physical buffer each scan of the batch in flight writes, fixed before its commit: the
    // neighbourhood metrics read it while the commit may already rewrite flip and next_slot
    int32_t stage_buf[GVOM_MAX_BATCH];
    double slot_origin[GVOM_MAX_BUFFER][3];
    double comb_origin[2][3];
    double latest_ego[3];
    int32_t sec_count[8];    // valid rays per azimuth sector (XCD-aware ray order)
    int32_t sec_cursor[8];
    // per slot: shift of the newest slot's origin relative to the slot's (x, y, z voxels) and
    // whether the slot holds a scan; written with the ring indices at the end of every scan
    // stage for the combine that follows
    alignas(16) int32_t comb_shift[GVOM_MAX_BUFFER][4];
    // set when a block gave up waiting for counts an earlier block publishes (PUB_READY
    // protocol; a defect): the state getters then report an error
    int32_t wait_fault;
};

// Physical buffer of ring slot `slot`: the one holding its scan, and the one the scan stage in
// flight writes.
__host__ __device__ __forceinline__ int cur_buf(int32_t flip, int slot) { return 2 * slot + ((flip >> slot) & 1); }

// Counts a block publishes for the later blocks of the same launch (radix-sort passes,
// segments, cell ids): count | PUB_READY in a word that is zero before the launch.  A block
// publishes before it waits and waits only for lower blocks, which were dispatched before it
// or are dispatched as other work retires, so every wait ends; the poll bound (about 30 ms
// with s_sleep 1) only guards against a defect: the waiter then goes on without the count
// and flags the fault.
constexpr uint32_t PUB_READY = 0x80000000u;
constexpr int PUB_SPIN_LIMIT = 1 << 20;

// The count published at p; fault set (and 0 returned) when the bound is reached.
__device__ __forceinline__ uint32_t pub_wait(const uint32_t *p, int &fault) {
    uint32_t v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    for (int spins = 0; !(v & PUB_READY) && spins < PUB_SPIN_LIMIT; ++spins) {
        __builtin_amdgcn_s_sleep(1);
        v = __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    if (v & PUB_READY) return v & ~PUB_READY;
    fault = 1;
    return 0;
}
__host__ __device__ __forceinline__ int new_buf(int32_t flip, int slot) {
    return 2 * slot + 1 - ((flip >> slot) & 1);
}

// One occupied voxel of a scan slot (reference: hit_count, total_count, metrics f64[10],
// min_height, gvom.py:198-208, 1079-1085).
struct SlotRec {
    int32_t hit;
    int32_t total;
    float min_h;
    int32_t pad;
    double m[10];   // mean x,y,z; cov xx,xy,xz,yy,yz,zz; neighbourhood point count
};

// One occupied voxel of a combined map (reference: combined_* arrays, gvom.py:262-277,
// and voxels_eigenvalues, 315).
struct CombRec {
    int32_t hit;
    int32_t total;
    float min_h;
    float m[10];
    float ev[3];
};

// Per-segment point moments about the segment's own voxel corner.
struct Moment {
    double S[3];
    double Q[6];    // xx, xy, xz, yy, yz, zz
    int32_t n;
    int32_t pad;
};

// Geometry constants passed by value to every kernel.
struct Geo {
    double xy_res, z_res;
    double min_d2;
    int32_t XY, Z;
    int64_t V;      // XY*XY*Z
    int64_t X;      // XY*XY
    int32_t B;
    int32_t ex, ez; // eigen neighbourhood radii
    int32_t XYp, Zp; // padded (halo) dims
    int32_t VB;     // voxel key bits: in-bounds keys < 2^(VB-1), halo keys have bit VB-1 set
    double pos_thr, neg_thr, slope_thr, robot_h, robot_r, g2l;
};

#define GVOM_CHECK_LAUNCH() (hipGetLastError())
