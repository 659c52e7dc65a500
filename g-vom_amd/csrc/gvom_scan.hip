// Scan stage of the scan->map path on gfx950: one Process_pointcloud call (or a batch
// of them) turns points into a ring slot (reference scripts/gvom.py:105-220).
//
// Launch sequence for a batch of S scans over NT points (all stream-ordered, no host
// synchronisation):
//   fill      slot tables of the S target slots := 0
//   prepare   transform, validity, own-voxel sort key per point
//   raycast   one thread per ray, exact stepping rule, run-aggregated atomics
//   sort      stable LSD radix sort of (key, point index)
//   segments  runs of equal keys -> segment starts, per-scan bounds
//   moments   per segment: n, sum r, sum r r^T, min local z; marks occupancy in the
//             slot table and writes the hit/total/min-height record
//   metrics   per occupied voxel: mean and population covariance over the
//             neighbourhood, assembled from neighbouring segments' moments
//   commit    ring bookkeeping
//
// Numerics follow the reference's Numba-CUDA typing; this file is compiled with
// -ffp-contract=off so that no multiply-add is fused.
#include <hip/hip_runtime.h>

#include "gvom_kernels.h"

namespace {

__device__ __forceinline__ float py_maxf(float a, float b) { return (b > a) ? b : a; }
__device__ __forceinline__ float py_minf(float a, float b) { return (b < a) ? b : a; }

// Scan index of global point i (offsets ascending, S <= 32).
__device__ __forceinline__ int scan_of(const ScanDesc *__restrict__ sd, int S, int64_t i) {
    int lo = 0, hi = S - 1;
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (sd[mid].offset <= i) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// gvom.py:138-141: origin = floor(ego / res - size / 2), f64.
__device__ __forceinline__ void scan_origin(const ScanDesc &d, const Geo &g, double o[3]) {
    o[0] = floor((d.ego[0] / g.xy_res) - (double)g.XY / 2.0);
    o[1] = floor((d.ego[1] / g.xy_res) - (double)g.XY / 2.0);
    o[2] = floor((d.ego[2] / g.z_res) - (double)g.Z / 2.0);
}

__device__ __forceinline__ int64_t lin_of(int64_t x, int64_t y, int64_t z, const Geo &g) {
    return (x * g.XY + y) * g.Z + z;
}

// __transform_pointcloud (gvom.py:1122-1138): arithmetic in the promoted type of the
// point and matrix dtypes, stored back in the point dtype.
template <typename P>
__device__ __forceinline__ void transform_point(const ScanDesc &d, P &x, P &y, P &z) {
    if (d.t_mode == GVOM_NO_TRANSFORM) return;
    if (sizeof(P) == 4 && d.t_mode == GVOM_F32) {
        float t[12];
#pragma unroll
        for (int k = 0; k < 12; ++k) t[k] = (float)d.T[k];
        float fx = (float)x, fy = (float)y, fz = (float)z;
        float r0 = fx * t[0] + fy * t[1] + fz * t[2] + t[3];
        float r1 = fx * t[4] + fy * t[5] + fz * t[6] + t[7];
        float r2 = fx * t[8] + fy * t[9] + fz * t[10] + t[11];
        x = (P)r0; y = (P)r1; z = (P)r2;
    } else {
        double dx = (double)x, dy = (double)y, dz = (double)z;
        const double *t = d.T;
        double r0 = dx * t[0] + dy * t[1] + dz * t[2] + t[3];
        double r1 = dx * t[4] + dy * t[5] + dz * t[6] + t[7];
        double r2 = dx * t[8] + dy * t[9] + dz * t[10] + t[11];
        x = (P)r0; y = (P)r1; z = (P)r2;
    }
}

// Minimum-distance filter in the point dtype (gvom.py:1145-1149), plus the domain rule:
// non-finite points and points with |p / res| >= 2^60 are dropped.
template <typename P>
__device__ __forceinline__ bool point_valid(P x, P y, P z, const Geo &g) {
    double dx = (double)x, dy = (double)y, dz = (double)z;
    if (!isfinite(dx) || !isfinite(dy) || !isfinite(dz)) return false;
    const double lim = 1152921504606846976.0;  // 2^60
    if (fabs(dx / g.xy_res) >= lim || fabs(dy / g.xy_res) >= lim || fabs(dz / g.z_res) >= lim) return false;
    P d2 = x * x + y * y + z * z;
    return !((double)d2 < g.min_d2);
}

// ------------------------------------------------------------------------------------
// Azimuth sector (8 x 45 deg) of a ray's xy direction; used only to schedule rays.
__device__ __forceinline__ int ray_sector(double dx, double dy) {
    const double a = atan2(dy, dx) + 3.141592653589793;   // [0, 2 pi]
    int s = (int)(a * (8.0 / 6.283185307179586));
    return s < 0 ? 0 : (s > 7 ? 7 : s);
}

// Radix sort tiles: RS_TILE keys per block of every pass.
constexpr int RS_TILE = gvom_plan::RS_TILE;

// With `hist` (the radix sort's RADIX_HIST_SETS x hwords words; the first set, the first
// pass's per-tile histogram, is zero on entry: the sort's last pass clears it for the next
// scan stage), the per-tile digit counts of the first pass are accumulated here, so the sort
// starts with its first scatter, and the later passes' sets (published tile counts and digit
// totals) are zeroed.
template <typename P>
__global__ __launch_bounds__(256) void k_prepare(const P *__restrict__ in, P *__restrict__ world,
                                                 uint32_t *__restrict__ keys, uint32_t *__restrict__ vals,
                                                 uint8_t *__restrict__ valid,
                                                 const ScanDesc *__restrict__ sd, int S, int64_t NT,
                                                 Geo g, DevState *__restrict__ st, uint8_t *__restrict__ rsec,
                                                 uint32_t *__restrict__ hist, int64_t hwords) {
    __shared__ int32_t scount[8];
    __shared__ uint32_t lh[256];
    if (rsec && threadIdx.x < 8) scount[threadIdx.x] = 0;
    if (hist) lh[threadIdx.x] = 0;
    if (rsec || hist) __syncthreads();
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < NT) {
    int s = scan_of(sd, S, i);
    const ScanDesc &d = sd[s];
    P x = in[3 * i], y = in[3 * i + 1], z = in[3 * i + 2];
    transform_point<P>(d, x, y, z);
    world[3 * i] = x;
    world[3 * i + 1] = y;
    world[3 * i + 2] = z;
    bool ok = point_valid<P>(x, y, z, g);
    valid[i] = ok ? 1 : 0;
    uint32_t key = GVOM_NOKEY;
    if (ok) {
        double o[3];
        scan_origin(d, g, o);
        double fx = floor(((double)x / g.xy_res) - o[0]);
        double fy = floor(((double)y / g.xy_res) - o[1]);
        double fz = floor(((double)z / g.z_res) - o[2]);
        uint32_t sk = (uint32_t)s << g.VB;
        if (fx >= 0 && fx < g.XY && fy >= 0 && fy < g.XY && fz >= 0 && fz < g.Z) {
            key = sk | (uint32_t)lin_of((int64_t)fx, (int64_t)fy, (int64_t)fz, g);
        } else if (fx >= -g.ex && fx < g.XY + g.ex && fy >= -g.ex && fy < g.XY + g.ex &&
                   fz >= -g.ez && fz < g.Z + g.ez) {
            // own voxel in the halo: contributes to in-map neighbours (gvom.py:1266-1291)
            int64_t kx = (int64_t)fx + g.ex, ky = (int64_t)fy + g.ex, kz = (int64_t)fz + g.ez;
            key = sk | (1u << (g.VB - 1)) | (uint32_t)((kx * g.XYp + ky) * g.Zp + kz);
        }
    }
    keys[i] = key;
    vals[i] = (uint32_t)i;
    if (hist) atomicAdd(&lh[key & 0xFFu], 1u);
    if (rsec) {
        int sec = 255;
        if (ok) {
            sec = ray_sector((double)x - d.ego[0], (double)y - d.ego[1]);
            atomicAdd(&scount[sec], 1);
        }
        rsec[i] = (uint8_t)sec;
    }
    }
    if (hist) {
        // the words of sets 1 .. RADIX_HIST_SETS - 1 that a sort of NT keys uses
        const int64_t rw = gvom_plan::radix_hist_words(NT);
        const int64_t nz = (int64_t)(gvom_plan::RADIX_HIST_SETS - 1) * rw;
        for (int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; k < nz; k += (int64_t)gridDim.x * blockDim.x)
            hist[(1 + k / rw) * hwords + k % rw] = 0;
    }
    if (rsec || hist) __syncthreads();
    if (rsec && threadIdx.x < 8 && scount[threadIdx.x]) atomicAdd(&st->sec_count[threadIdx.x], scount[threadIdx.x]);
    // the block's 256 points lie in one sort tile (256 divides RS_TILE)
    if (hist && lh[threadIdx.x])
        atomicAdd(&hist[((int64_t)blockIdx.x * 256 / RS_TILE) * 256 + threadIdx.x], lh[threadIdx.x]);
}

// Groups valid rays by azimuth sector: each block appends its rays of sector s, in point
// order, at a position reserved with one atomic per (block, sector).
__global__ __launch_bounds__(256) void k_ray_order(const uint8_t *__restrict__ rsec, int64_t NT,
                                                   DevState *__restrict__ st, uint32_t *__restrict__ order) {
    __shared__ int32_t cnt[8], base[8];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (t < 8) cnt[t] = 0;
    __syncthreads();
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + t;
    const int sec = i < NT ? rsec[i] : 255;
    // rank within the block among rays of the same sector, in point order
    int rank = 0;
    for (int k = 0; k < 8; ++k) {
        const uint64_t m = __ballot(sec == k);
        if (sec == k) rank = __popcll(m & ((1ull << lane) - 1ull));
        __shared__ int32_t wc[4][8];
        if (lane == 0) wc[w][k] = __popcll(m);
        __syncthreads();
        if (sec == k)
            for (int ww = 0; ww < w; ++ww) rank += wc[ww][k];
        if (t == 0) {
            int c = 0;
            for (int ww = 0; ww < 4; ++ww) c += wc[ww][k];
            cnt[k] = c;
        }
        __syncthreads();
    }
    if (t < 8) {
        int32_t start = 0;
        for (int k = 0; k < t; ++k) start += st->sec_count[k];
        base[t] = cnt[t] ? start + atomicAdd(&st->sec_cursor[t], cnt[t]) : 0;
    }
    __syncthreads();
    if (sec < 8) order[base[sec] + rank] = (uint32_t)i;
}

// ------------------------------------------------------------------------------------
// Ray casting (__point_2_map, gvom.py:1174-1231).  State pt/end/slope in f32, length in
// f64, dominant axis = argmax |slope| with ties to y over x and z over both, each step
// adds +-1 to the dominant coordinate and slope/|slope_dom| to the others, the ray is
// abandoned at its first out-of-map step.
//
// The loop is kept convergent across the wavefront (dead lanes carry a unique
// sentinel key) so that at every step, lanes whose rays sit in the same voxel can be
// merged: consecutive points of a spinning-lidar scan have nearly equal directions,
// so near the sensor long runs of adjacent lanes hit the same voxel.  The head of
// each run of equal keys issues one atomic add of the run length.
//
// Termination: the dominant coordinate moves by exactly 1 voxel per step while
// |pt| < 2^23 (ego is restricted to |ego/res| < 2^22), and the ego starts inside the
// map, so a ray leaves the map within max(xy, z) + 1 steps; maxit bounds the loop at
// that number plus one.
// Near-sensor privatisation: every ray starts at the ego, so the voxels around it are
// the most contended counters.  With USE_BOX, a block keeps a BOX_XY x BOX_XY x BOX_Z
// box of counters centred on the ego voxel of its first ray's scan in LDS; steps of
// that scan inside the box add there, and the block flushes the non-zero counters
// with one global atomic each at its end.
// One ray's stepping state (gvom.py:1174-1213 typing: f32 position / slope, f64 length).
struct Ray {
    float p0, p1, p2, inc0, inc1, inc2;
    double length, rl1, dlen;
    int32_t o0, o1, o2;   // the scan's origin (integral, |o| < 2^23)
    bool alive;
};

// Ray from the scan's ego to world point i (gvom.py:1174-1208).
template <typename P>
__device__ __forceinline__ void ray_init(Ray &r, const ScanDesc &d, const Geo &g, const P *__restrict__ world,
                                         int64_t i) {
    double o[3];
    scan_origin(d, g, o);
    r.o0 = (int32_t)o[0]; r.o1 = (int32_t)o[1]; r.o2 = (int32_t)o[2];
    float pt0 = (float)(d.ego[0] / g.xy_res);
    float pt1 = (float)(d.ego[1] / g.xy_res);
    float pt2 = (float)(d.ego[2] / g.z_res);
    float e0 = (float)((double)world[3 * i] / g.xy_res);
    float e1 = (float)((double)world[3 * i + 1] / g.xy_res);
    float e2 = (float)((double)world[3 * i + 2] / g.z_res);
    float s0 = e0 - pt0, s1 = e1 - pt1, s2 = e2 - pt2;
    float rl = sqrtf(s0 * s0 + s1 * s1 + s2 * s2);
    s0 = s0 / rl;
    s1 = s1 / rl;
    s2 = s2 / rl;
    float a0 = fabsf(s0), a1 = fabsf(s1), a2 = fabsf(s2);
    float smax = py_maxf(a0, py_maxf(a1, a2));
    int si = 0;
    if (smax == a1) si = 1;
    if (smax == a2) si = 2;
    float sd_ = si == 0 ? s0 : (si == 1 ? s1 : s2);
    float ad = fabsf(sd_);
    float dir = sd_ / ad;
    // per-axis increments: +-1 on the dominant axis, slope/|slope_dom| elsewhere
    r.inc0 = si == 0 ? dir : s0 / ad;
    r.inc1 = si == 1 ? dir : s1 / ad;
    r.inc2 = si == 2 ? dir : s2 / ad;
    r.p0 = pt0; r.p1 = pt1; r.p2 = pt2;
    r.length = 0.0;
    r.rl1 = (double)rl - 1.0;
    r.dlen = fabs(1.0 / (double)sd_);
    r.alive = r.length < r.rl1;
}

// (int32_t)floorf(x) in one instruction: v_cvt_flr_i32_f32 converts rounding toward -inf with
// the same saturation and NaN rules as v_cvt_i32_f32, which the compiler otherwise emits after
// a separate v_floor_f32.
__device__ __forceinline__ int32_t floor_to_i32(float x) {
    int32_t r;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(r) : "v"(x));
    return r;
}

// One step of an alive ray (gvom.py:1209-1231): returns false when the voxel it enters is
// off the map (the ray is then dead), else its coordinates; advances length.
__device__ __forceinline__ bool ray_advance(Ray &r, const Geo &g, int32_t &ix, int32_t &iy, int32_t &iz) {
    r.p0 += r.inc0;
    r.p1 += r.inc1;
    r.p2 += r.inc2;
    // gvom.py:1215-1223 takes floor(pt - origin) in f64.  pt is f32 and origin an integer,
    // both below 2^23 in magnitude, so the f64 difference is exact and equals
    // floor(pt) - origin: one f32 floor and an integer subtraction per axis.
    const int32_t fx = (int32_t)floorf(r.p0) - r.o0;
    const int32_t fy = (int32_t)floorf(r.p1) - r.o1;
    const int32_t fz = (int32_t)floorf(r.p2) - r.o2;
    if ((uint32_t)fx >= (uint32_t)g.XY || (uint32_t)fy >= (uint32_t)g.XY || (uint32_t)fz >= (uint32_t)g.Z) {
        r.alive = false;
        return false;
    }
    r.length += r.dlen;
    r.alive = r.length < r.rl1;
    ix = fx;
    iy = fy;
    iz = fz;
    return true;
}

// ((ix * XY + iy) * Z + iz) as ix * (XY * Z) + iy * Z + iz: on the map ix, iy < XY <= 2^13 and
// XY * Z <= 2^21, so both products are 24-bit multiplies (full-rate v_mad_u32_u24 instead of a
// 64-bit multiply-add and a 32-bit multiply); a lane off the map gets a masked product, and its
// key is discarded.
__device__ __forceinline__ uint32_t lin_index(int32_t ix, int32_t iy, int32_t iz, const Geo &g) {
    return __umul24((uint32_t)ix, (uint32_t)g.XY * (uint32_t)g.Z) + __umul24((uint32_t)iy, (uint32_t)g.Z) +
           (uint32_t)iz;
}

// ray_advance for every lane at once, without branches (so a step costs no exec-mask
// changes): the key of the voxel the step enters, base + index(voxel), or the dead-lane
// sentinel 0xFFFFFFFF - lane when the ray was dead or leaves the map.  A dead lane's
// position keeps moving; nothing reads it.
template <bool TILED>
__device__ __forceinline__ uint32_t ray_step_key(Ray &r, const Geo &g, uint32_t base, int32_t ty, int lane);

// Branch-free run merge: as run_length, with the result selected instead of branched on.
__device__ __forceinline__ int run_length_sel(uint32_t key, int lane) {
    const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)key, 0x138, 0xF, 0xF, false);
    const bool is_voxel = key < 0x80000000u;
    const bool cont = (lane > 0) & (prev == key) & is_voxel;
    const uint64_t same = __ballot(cont);
    const uint64_t rest = (lane == 63) ? 0ull : (same >> (lane + 1));
    const int len = 1 + __builtin_ctzll(~rest);
    return (is_voxel & !cont) ? len : 0;
}

// Run merge across a wavefront: lanes whose keys are equal to the previous lane's
// continue that lane's run (the previous lane's key comes from a DPP wave shift right by
// one; lane 0 receives a dead-lane sentinel).  Returns the run length at the head lane of
// each run of voxel keys (< 2^31), 0 elsewhere.
__device__ __forceinline__ int run_length(uint32_t key, int lane) {
    const uint32_t prev = (uint32_t)__builtin_amdgcn_update_dpp((int)0xFFFFFFFF, (int)key, 0x138, 0xF, 0xF, false);
    const bool is_voxel = key < 0x80000000u;
    const bool cont = (lane > 0) && (prev == key) && is_voxel;
    const uint64_t same = __ballot(cont);
    if (!is_voxel || cont) return 0;
    const uint64_t rest = (lane == 63) ? 0ull : (same >> (lane + 1));
    return 1 + __builtin_ctzll(~rest);
}

// Ray-cast scratch layout: 4 x 4 columns x 1 z per 64-byte segment (tile t = (x/4 * ty + y/4)
// * z_size + z, word 16 t + 4 (x%4) + y%4).  A wavefront's run heads at one step lie on an
// arc of nearby voxels; device-scope atomics of one instruction that fall in the same 64-byte
// segment travel as one request, and the segment holds more of the arc's voxels in this
// layout than in the z-contiguous table (measured on the headline scan: 0.76M instead of
// 2.1M segments for 2.16M runs).  k_tile_convert moves the counts into the slot tables.
struct TileGeo {
    int32_t tx, ty;    // tiles along x, y
    int64_t vt;        // words per scan
};
__host__ __device__ __forceinline__ TileGeo tile_geo(const Geo &g) {
    TileGeo t;
    t.tx = (g.XY + 3) / 4;
    t.ty = (g.XY + 3) / 4;
    t.vt = (int64_t)t.tx * t.ty * g.Z * 16;
    return t;
}
__device__ __forceinline__ uint32_t tiled_index(int32_t ix, int32_t iy, int32_t iz, const Geo &g, int32_t ty) {
    // as lin_index: 24-bit products (ty * Z <= 2^19)
    return ((__umul24((uint32_t)(ix >> 2), (uint32_t)ty * (uint32_t)g.Z) + __umul24((uint32_t)(iy >> 2), (uint32_t)g.Z) +
             (uint32_t)iz) << 4) +
           ((uint32_t)(ix & 3) << 2) + (uint32_t)(iy & 3);
}

template <bool TILED>
__device__ __forceinline__ uint32_t ray_step_key(Ray &r, const Geo &g, uint32_t base, int32_t ty, int lane) {
    const bool was = r.alive;
    r.p0 += r.inc0;
    r.p1 += r.inc1;
    r.p2 += r.inc2;
    // floor(pt - origin) in f64 (gvom.py:1215-1223) = floor(pt) - origin, as in ray_advance
    const int32_t fx = floor_to_i32(r.p0) - r.o0;
    const int32_t fy = floor_to_i32(r.p1) - r.o1;
    const int32_t fz = floor_to_i32(r.p2) - r.o2;
    const bool in = ((uint32_t)fx < (uint32_t)g.XY) & ((uint32_t)fy < (uint32_t)g.XY) & ((uint32_t)fz < (uint32_t)g.Z);
    const bool step = was & in;
    r.length += r.dlen;   // read only while the ray is alive
    r.alive = step & (r.length < r.rl1);
    const uint32_t k = base + (TILED ? tiled_index(fx, fy, fz, g, ty) : lin_index(fx, fy, fz, g));
    return step ? k : 0xFFFFFFFFu - (uint32_t)lane;
}

// Ring-patch ray order.  For an organised scan (rings of `stride` consecutive points), a
// wavefront traces a patch of 64 / pr adjacent azimuths x pr adjacent rings instead of 64
// azimuths of one ring: at one step its rays sit over nearly the same column at different
// heights, so their run heads fall on few z-contiguous 64-byte segments of the slot table
// (measured offline on the 128 x 1024 headline scan, 2 x 32 patches: 0.30M segment requests
// against 0.76M for the tiled scratch and 2.1M for one ring per wavefront).  Within a patch
// the lanes run ring-fastest: adjacent lanes are adjacent rings at one azimuth, whose rays
// share voxels near the sensor, so the run merge of adjacent lanes catches them instead of
// several lanes of one instruction updating the same word.  Bands
// of pr complete rings are cut into patches; rays after the last complete band (and rays of a
// scan whose size is not a multiple of stride) follow in linear order.  A bijection on
// [0, NT): the counts do not depend on it.
__device__ __forceinline__ int64_t patch_ray(int64_t gw, int lane, int64_t NT, int64_t stride, int pr) {
    const int pa = 64 / pr;
    const int64_t per_band = stride / pa;              // patches across one band of pr rings
    const int64_t bands = NT / (stride * pr);          // complete bands
    const int64_t full = bands * per_band;             // wavefronts in complete bands
    if (gw < full) {
        const int64_t band = gw / per_band, col = gw - band * per_band;
        const int r = lane % pr, a = lane / pr;
        return (band * pr + r) * stride + col * pa + a;
    }
    return bands * stride * pr + ((gw - full) << 6) + lane;
}

constexpr int BOX_XY = 32, BOX_Z = 16;
constexpr int BOX_N = BOX_XY * BOX_XY * BOX_Z;   // 64 KiB of int32 counters

template <typename P, bool USE_BOX, bool TILED>
__global__ __launch_bounds__(256) void k_raycast(const P *__restrict__ world,
                                                 const uint8_t *__restrict__ valid,
                                                 const ScanDesc *__restrict__ sd, int S, int64_t NT,
                                                 Geo g, const DevState *__restrict__ st,
                                                 int32_t *__restrict__ tables, int maxit, int rpt, int force_slot,
                                                 const uint32_t *__restrict__ order, int spread,
                                                 int64_t ring_stride, int patch_rings) {
    extern __shared__ int32_t box[];
    // With `order` (XCD-aware schedule) block b traces chunk b / 8 of azimuth sector b % 8;
    // workgroups are dispatched to the 8 XCDs round-robin, so each XCD traces one sector.
    int64_t o_beg = 0, o_len = NT;
    int64_t chunk = blockIdx.x;
    if (order) {
        const int sec = blockIdx.x & 7;
        chunk = blockIdx.x >> 3;
        for (int k = 0; k < sec; ++k) o_beg += st->sec_count[k];
        o_len = st->sec_count[sec];
    }
#define RAY_INDEX(j) (order ? ((j) < o_len ? (int64_t)order[o_beg + (j)] : NT) : (j))
    const int lane = threadIdx.x & 63;
    const int next = st->next_slot, flip = st->flip;
    const TileGeo tg = tile_geo(g);
    // the table the scan stage writes for the batch's s-th scan (force_slot >= 0, autotuning:
    // table force_slot for every scan)
#define RAY_SLOT(s) (force_slot >= 0 ? force_slot : new_buf(flip, (next + (s)) % g.B))
    // the block's box: scan of its first ray, centred on that scan's ego voxel
    int bs = 0;
    int64_t bx0 = 0, by0 = 0, bz0 = 0;
    int64_t bbase = 0;   // the box's slot buffer, in words
    (void)bbase;
    if (USE_BOX) {
        int64_t i0 = RAY_INDEX(chunk * rpt * blockDim.x);
        if (i0 >= NT) i0 = NT > 0 ? NT - 1 : 0;
        bs = scan_of(sd, S, i0);
        double o[3];
        scan_origin(sd[bs], g, o);
        bx0 = (int64_t)floor((double)(float)(sd[bs].ego[0] / g.xy_res) - o[0]) - BOX_XY / 2;
        by0 = (int64_t)floor((double)(float)(sd[bs].ego[1] / g.xy_res) - o[1]) - BOX_XY / 2;
        bz0 = (int64_t)floor((double)(float)(sd[bs].ego[2] / g.z_res) - o[2]) - BOX_Z / 2;
        bbase = (int64_t)RAY_SLOT(bs) * g.V;
        for (int k = threadIdx.x; k < BOX_N; k += blockDim.x) box[k] = 0;
        __syncthreads();
    }
    for (int rep = 0; rep < rpt; ++rep) {
        // spread: the block's wavefronts come from across the scan (wavefront w of pass rep of
        // block b traces rays of global wavefront (rep * waves per block + w) * blocks + b), so
        // the long-ray rings do not end up together on a few CUs; lanes stay adjacent rays
        const int64_t wpb = blockDim.x >> 6;
        const int64_t gw = (spread && !order) ? ((int64_t)rep * wpb + (threadIdx.x >> 6)) * gridDim.x + chunk
                                              : (chunk * rpt + rep) * wpb + (threadIdx.x >> 6);
        const int64_t j = (patch_rings > 0 && !order) ? patch_ray(gw, lane, NT, ring_stride, patch_rings)
                                                      : (gw << 6) + lane;
        const int64_t i = RAY_INDEX(j);
        bool inbox_scan = false;
        Ray r;
        r.alive = false;
        uint32_t base = 0;
        int64_t toff = 0;
        if (i < NT && valid[i]) {
            const int s = scan_of(sd, S, i);
            inbox_scan = USE_BOX && s == bs;
            // keys: scan s of the batch at s * (words per table), so adjacent lanes of
            // different scans never merge; TILED counts into the s-th scratch table, else a
            // run head reaches the scan's slot buffer through the 64-bit offset toff
            if (TILED) {
                base = (uint32_t)(s * tg.vt);
            } else {
                base = (uint32_t)(s * g.V);
                toff = (int64_t)RAY_SLOT(s) * g.V - (int64_t)s * g.V;
            }
            ray_init<P>(r, sd[s], g, world, i);
        }
        if (!USE_BOX) {
            // two steps per round: both keys are formed before either run merge, so the two
            // steps' arithmetic overlaps; a ray takes at most max(xy, z) + 1 steps, so the
            // round that passes maxit has no live lane left
            for (int it = 0; it < maxit; it += 2) {
                if (__ballot(r.alive) == 0ull) break;
                const uint32_t k1 = ray_step_key<TILED>(r, g, base, tg.ty, lane);
                const uint32_t k2 = ray_step_key<TILED>(r, g, base, tg.ty, lane);
                const int l1 = run_length_sel(k1, lane);
                const int l2 = run_length_sel(k2, lane);
                if (l1 > 0) __hip_atomic_fetch_add(tables + toff + k1, l1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if (l2 > 0) __hip_atomic_fetch_add(tables + toff + k2, l2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            continue;
        }
        for (int it = 0; it < maxit; ++it) {
            if (__ballot(r.alive) == 0ull) break;
            uint32_t key = 0xFFFFFFFFu - (uint32_t)lane;   // unique per dead lane
            int lbox = -1;
            if (r.alive) {
                int32_t ix, iy, iz;
                const bool in = ray_advance(r, g, ix, iy, iz);
                if (in && TILED) {
                    key = base + tiled_index(ix, iy, iz, g, tg.ty);
                } else if (in) {
                    key = base + lin_index(ix, iy, iz, g);
                    if (USE_BOX && inbox_scan) {
                        const int32_t lx = ix - (int32_t)bx0, ly = iy - (int32_t)by0, lz = iz - (int32_t)bz0;
                        if (lx >= 0 && lx < BOX_XY && ly >= 0 && ly < BOX_XY && lz >= 0 && lz < BOX_Z)
                            lbox = (int)((lx * BOX_XY + ly) * BOX_Z + lz);
                    }
                }
            }
            const int len = run_length(key, lane);
            if (len > 0) {
                if (USE_BOX && lbox >= 0) atomicAdd(&box[lbox], len);
                else __hip_atomic_fetch_add(tables + toff + key, len, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
        }
    }
    if (USE_BOX) {
        __syncthreads();
        for (int k = threadIdx.x; k < BOX_N; k += blockDim.x) {
            const int32_t v = box[k];
            if (v == 0) continue;
            const int64_t lz = k % BOX_Z, lc = k / BOX_Z, ly = lc % BOX_XY, lx = lc / BOX_XY;
            const int64_t w = bbase + (int64_t)lin_of(bx0 + lx, by0 + ly, bz0 + lz, g);
            __hip_atomic_fetch_add(tables + w, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
#undef RAY_SLOT
#undef RAY_INDEX
}

// ------------------------------------------------------------------------------------
// Radix sort: 8-bit digits, tiles of RS_TILE keys (defined with k_prepare).
// Points per moment chunk (one wavefront, see mom_chunk).
constexpr int MCH = 64;

// Tiles of RS_TILE keys, RS_THREADS threads (2 keys each): more wavefronts per tile keep
// the short passes of a ~10^5-key sort from running on a quarter of the chip.
constexpr int RS_THREADS = 1024;
constexpr int RS_WAVES = RS_THREADS / 64;
constexpr int RS_ROWS = RS_TILE / RS_THREADS;   // 64-key rows per wavefront

// The lookback passes publish their tiles' digit counts (PUB_READY protocol); a tile that
// gives up waiting uses what it has, so its positions stay in [0, n).

// Ranks the tile's keys by the digit at `shift`, stably: rows in key order, waves in order.
// wcnt[w][d] ends as the count of digit d in wavefront w; lo = rank within the wavefront.
// With nh, the next digit of every key is counted into nh (the next pass's totals).
__device__ __forceinline__ void rs_rank_tile(const uint32_t *__restrict__ kin, const uint32_t *__restrict__ vin,
                                             int64_t n, int shift, int64_t base, int lane, int w,
                                             uint32_t (*wcnt)[256], uint32_t *nh, uint32_t (&kk)[RS_ROWS],
                                             uint32_t (&vv)[RS_ROWS], uint32_t (&lo)[RS_ROWS]) {
    const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
    for (int r = 0; r < RS_ROWS; ++r) {
        int64_t i = base + r * 64 + lane;
        bool in = i < n;
        uint32_t key = in ? kin[i] : 0u;
        kk[r] = key;
        vv[r] = in ? vin[i] : 0u;
        uint32_t d = (key >> shift) & 0xFFu;
        if (nh && in) atomicAdd(&nh[(key >> (shift + 8)) & 0xFFu], 1u);
        uint64_t m = __ballot(in);
#pragma unroll
        for (int b = 0; b < 8; ++b) {
            bool bit = (d >> b) & 1u;
            uint64_t bb = __ballot(bit);
            m &= bit ? bb : ~bb;
        }
        uint32_t rank = __popcll(m & lt), cnt = __popcll(m);
        uint32_t prior = wcnt[w][d];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        if (in && rank == cnt - 1) wcnt[w][d] = prior + cnt;
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        lo[r] = prior + rank;
    }
}

// Turns wcnt[.][d] into exclusive wave offsets (thread d < 256) and returns the tile's count
// of digit d.
__device__ __forceinline__ uint32_t rs_wave_offsets(uint32_t (*wcnt)[256], int d) {
    uint32_t a = 0;
#pragma unroll
    for (int ww = 0; ww < RS_WAVES; ++ww) {
        uint32_t c = wcnt[ww][d];
        wcnt[ww][d] = a;
        a += c;
    }
    return a;
}

// Exclusive prefix of the digits' totals inside the wavefront (thread d < 256 holds digit
// d's); the wavefront's sum goes to wtot[w].  After a barrier, rs_digit_add adds the sums of
// the wavefronts before.
__device__ __forceinline__ uint32_t rs_digit_scan(uint32_t total, int lane, int w, uint32_t *wtot) {
    uint32_t v = total;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    if (lane == 63) wtot[w] = v;
    return v - total;
}

__device__ __forceinline__ uint32_t rs_digit_add(const uint32_t *wtot, int w) {
    uint32_t wo = 0;
    for (int ww = 0; ww < w; ++ww) wo += wtot[ww];
    return wo;
}

__device__ __forceinline__ void rs_scatter_tile(uint32_t *__restrict__ kout, uint32_t *__restrict__ vout, int64_t n,
                                                int shift, int64_t base, int lane, int w, uint32_t (*wcnt)[256],
                                                const uint32_t *gbase, const uint32_t (&kk)[RS_ROWS],
                                                const uint32_t (&vv)[RS_ROWS], const uint32_t (&lo)[RS_ROWS]) {
#pragma unroll
    for (int r = 0; r < RS_ROWS; ++r) {
        int64_t i = base + r * 64 + lane;
        if (i < n) {
            uint32_t d = (kk[r] >> shift) & 0xFFu;
            uint32_t pos = gbase[d] + wcnt[w][d] + lo[r];
            kout[pos] = kk[r];
            vout[pos] = vv[r];
        }
    }
}

// First pass.  hist: its per-tile digit counts (block-major, from k_prepare).  tot_next
// (zero on entry, or null in a one-pass sort): the next digit's totals, accumulated here.
// scan_cnt (zero on entry): valid keys (!= GVOM_NOKEY) per scan, the scan being key >> VB.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(const uint32_t *__restrict__ kin,
                                                           const uint32_t *__restrict__ vin,
                                                           uint32_t *__restrict__ kout, uint32_t *__restrict__ vout,
                                                           int64_t n, int shift, const uint32_t *__restrict__ hist,
                                                           int nblk, uint32_t *__restrict__ tot_next, int VB,
                                                           uint32_t *__restrict__ scan_cnt) {
    __shared__ uint32_t wcnt[RS_WAVES][256];
    __shared__ uint32_t ptot[4][256], pbef[4][256];
    __shared__ uint32_t gbase[256], nh[256];
    __shared__ uint32_t wtot[4];
    __shared__ uint32_t scn[GVOM_MAX_BATCH];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int k = threadIdx.x; k < RS_WAVES * 256; k += RS_THREADS) (&wcnt[0][0])[k] = 0;
    if (threadIdx.x < 256) nh[threadIdx.x] = 0;
    if (threadIdx.x < GVOM_MAX_BATCH) scn[threadIdx.x] = 0;
    {
        // global offset of digit d for this block: keys of smaller digits in all blocks plus
        // keys of digit d in earlier blocks.  Four threads per digit each add every fourth
        // block's count (hist is block-major, so each round of loads coalesces).
        const int d = threadIdx.x & 255, q = threadIdx.x >> 8;
        const int me = (int)blockIdx.x;
        uint32_t before = 0, total = 0;
#pragma unroll 8
        for (int b = q; b < nblk; b += 4) {
            const uint32_t c = hist[(int64_t)b * 256 + d];
            total += c;
            before += (b < me) ? c : 0u;
        }
        ptot[q][d] = total;
        pbef[q][d] = before;
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        const int d = threadIdx.x;
        ptot[0][d] = rs_digit_scan(ptot[0][d] + ptot[1][d] + ptot[2][d] + ptot[3][d], lane, w, wtot) +
                     pbef[0][d] + pbef[1][d] + pbef[2][d] + pbef[3][d];
    }
    __syncthreads();
    if (threadIdx.x < 256) gbase[threadIdx.x] = rs_digit_add(wtot, w) + ptot[0][threadIdx.x];
    const int64_t base = (int64_t)blockIdx.x * RS_TILE + w * (RS_ROWS * 64);
    uint32_t kk[RS_ROWS], vv[RS_ROWS], lo[RS_ROWS];
    rs_rank_tile(kin, vin, n, shift, base, lane, w, wcnt, tot_next ? nh : nullptr, kk, vv, lo);
    // valid keys per scan: one LDS add per scan present in a row (rows are runs of consecutive
    // points, so usually of one scan)
#pragma unroll
    for (int r = 0; r < RS_ROWS; ++r) {
        const bool valid = base + r * 64 + lane < n && kk[r] != GVOM_NOKEY;
        const uint32_t sc = valid ? kk[r] >> VB : 0u;
        uint64_t rem = __ballot(valid);
        while (rem) {
            const int leader = __builtin_ctzll(rem);
            const uint32_t s0 = __shfl(sc, leader);
            const uint64_t m = __ballot(valid && sc == s0);
            if (lane == leader) atomicAdd(&scn[s0], (uint32_t)__popcll(m));
            rem &= ~m;
        }
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        rs_wave_offsets(wcnt, threadIdx.x);
        if (tot_next && nh[threadIdx.x]) atomicAdd(&tot_next[threadIdx.x], nh[threadIdx.x]);
    }
    if (threadIdx.x < GVOM_MAX_BATCH && scn[threadIdx.x]) atomicAdd(&scan_cnt[threadIdx.x], scn[threadIdx.x]);
    __syncthreads();
    rs_scatter_tile(kout, vout, n, shift, base, lane, w, wcnt, gbase, kk, vv, lo);
}

// Later passes, without a histogram kernel: each tile ranks its keys, publishes its digit
// counts (count | PUB_READY in status[tile][d], zero on entry), then adds the counts of the
// tiles before it, waiting for those not yet published; `tot` (this digit's totals,
// accumulated by the previous pass) gives the digits' starts.  A tile waits only on lower
// tiles, which were dispatched before it or are dispatched as other work retires, so every
// wait ends (and is bounded regardless).  tot_next as in k_rs_scatter; hclear (last pass) is
// the first pass's histogram, cleared for the next sort.
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter_lb(const uint32_t *__restrict__ kin,
                                                              const uint32_t *__restrict__ vin,
                                                              uint32_t *__restrict__ kout,
                                                              uint32_t *__restrict__ vout, int64_t n, int shift,
                                                              uint32_t *__restrict__ status,
                                                              const uint32_t *__restrict__ tot,
                                                              uint32_t *__restrict__ tot_next, int nblk,
                                                              uint32_t *__restrict__ hclear, int64_t hwords,
                                                              DevState *__restrict__ st) {
    __shared__ uint32_t wcnt[RS_WAVES][256];
    __shared__ uint32_t pbef[4][256];
    __shared__ uint32_t gbase[256], nh[256];
    __shared__ uint32_t wtot[4];
    __shared__ int fault;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int me = (int)blockIdx.x;
    for (int k = threadIdx.x; k < RS_WAVES * 256; k += RS_THREADS) (&wcnt[0][0])[k] = 0;
    if (threadIdx.x < 256) nh[threadIdx.x] = 0;
    if (threadIdx.x == 0) fault = 0;
    __syncthreads();
    const int64_t base = (int64_t)me * RS_TILE + w * (RS_ROWS * 64);
    uint32_t kk[RS_ROWS], vv[RS_ROWS], lo[RS_ROWS];
    rs_rank_tile(kin, vin, n, shift, base, lane, w, wcnt, tot_next ? nh : nullptr, kk, vv, lo);
    __syncthreads();
    if (threadIdx.x < 256) {
        const uint32_t c = rs_wave_offsets(wcnt, threadIdx.x);
        __hip_atomic_store(&status[(int64_t)me * 256 + threadIdx.x], c | PUB_READY, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        if (tot_next && nh[threadIdx.x]) atomicAdd(&tot_next[threadIdx.x], nh[threadIdx.x]);
    }
    {
        // counts of digit d in the tiles before this one; four threads per digit, eight loads
        // in flight per thread before any is waited on
        const int d = threadIdx.x & 255, q = threadIdx.x >> 8;
        uint32_t before = 0;
        for (int b0 = q; b0 < me; b0 += 32) {
            uint32_t v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int b = b0 + 4 * u;
                v[u] = b < me ? __hip_atomic_load(&status[(int64_t)b * 256 + d], __ATOMIC_RELAXED,
                                                  __HIP_MEMORY_SCOPE_AGENT)
                              : PUB_READY;
            }
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int b = b0 + 4 * u;
                int spins = 0;
                while (!(v[u] & PUB_READY) && spins < PUB_SPIN_LIMIT) {
                    __builtin_amdgcn_s_sleep(1);
                    v[u] = __hip_atomic_load(&status[(int64_t)b * 256 + d], __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
                    ++spins;
                }
                if (v[u] & PUB_READY) before += v[u] & ~PUB_READY;
                else fault = 1;
            }
        }
        pbef[q][d] = before;
    }
    __syncthreads();
    if (threadIdx.x < 256) {
        const int d = threadIdx.x;
        pbef[0][d] += rs_digit_scan(tot[d], lane, w, wtot) + pbef[1][d] + pbef[2][d] + pbef[3][d];
    }
    __syncthreads();
    if (threadIdx.x < 256) gbase[threadIdx.x] = rs_digit_add(wtot, w) + pbef[0][threadIdx.x];
    __syncthreads();
    rs_scatter_tile(kout, vout, n, shift, base, lane, w, wcnt, gbase, kk, vv, lo);
    if (hclear)
        for (int64_t k = (int64_t)me * RS_THREADS + threadIdx.x; k < hwords; k += (int64_t)gridDim.x * RS_THREADS)
            hclear[k] = 0;
    if (threadIdx.x == 0 && fault) st->wait_fault = 1;
}

// ------------------------------------------------------------------------------------
// Segments of the sorted key array.

// Segment kernels: blocks of 256 threads x 8 sorted keys, independent of the sort's tile.
constexpr int SEG_TILE = gvom_plan::SEG_TILE;   // 256 threads x 8 keys
static_assert(SEG_TILE == 256 * 8, "segment kernels: 256 threads x 8 keys per block");

// Totals of the per-block counts -> nseg, nvalid, end sentinel of the last segment,
// per-scan [first segment, end of in-map segments) and slot cell counts, by one block
// (blockDim >= 64).  Segments are ordered by scan, and within a scan in-map keys precede
// halo keys.
__device__ void seg_bounds_body(const uint32_t *__restrict__ blk, int nblk, int S, const Geo &g,
                                int32_t *__restrict__ seg_start, DevState *__restrict__ st) {
    __shared__ uint32_t tot[2 + 3 * GVOM_MAX_BATCH];
    const int t = threadIdx.x;
    const int W = 2 + 3 * S;
    for (int col = t; col < W; col += blockDim.x) {
        uint32_t a = 0;
#pragma unroll 8
        for (int b = 0; b < nblk; ++b) a += blk[(int64_t)b * W + col];
        tot[col] = a;
    }
    __syncthreads();
    if (t != 0) return;
    st->nseg = (int32_t)tot[0];
    st->nvalid = (int32_t)tot[1];
    seg_start[tot[0]] = (int32_t)tot[1];
    int32_t f = 0, pos = 0, ch = 0;
    for (int s = 0; s < S; ++s) {
        const int32_t in = (int32_t)tot[2 + 3 * s], halo = (int32_t)tot[3 + 3 * s];
        const int32_t nv = S > 1 ? (int32_t)tot[4 + 3 * s] : (int32_t)tot[1];
        st->first_seg[s] = f;
        st->inb_end[s] = f + in;
        st->scan_cells[s] = in;
        st->stage_buf[s] = new_buf(st->flip, (st->next_slot + s) % g.B);
        st->scan_pos[s] = pos;
        st->scan_valid[s] = nv;
        st->chunk_off[s] = ch;
        f += in + halo;
        pos += nv;
        ch += (nv + MCH - 1) / MCH;
    }
    st->chunk_off[S] = ch;
}

__global__ __launch_bounds__(64) void k_seg_bounds(const uint32_t *__restrict__ blk, int nblk, int S, Geo g,
                                                   int32_t *__restrict__ seg_start, DevState *__restrict__ st) {
    seg_bounds_body(blk, nblk, S, g, seg_start, st);
}

// ------------------------------------------------------------------------------------
// Per-segment moments about the segment's own voxel corner, in two deterministic steps:
//   k_segments: one wavefront per chunk of MCH sorted points (a lane per point) reduces
//               every run of equal keys inside the chunk and writes its partial Moment at
//               the run's last position;
//   k_mom_seg : one thread per segment adds the partials of the chunks it spans, in
//               chunk order.
// A segment of L points costs ceil(L / MCH) + 1 partial loads instead of L serial point
// loads, so the long segments of voxels next to the sensor no longer set the tail.
// k_mom_seg is also where the slot record is born: hit = points in the voxel
// (gvom.py:1170), total = ray passes + hits (1171, 1229), min_h = min local z over the
// voxel's points (1390-1421), and the table entry becomes OCC | id.

__device__ __forceinline__ void key_voxel(uint32_t key, const Geo &g, int &s, bool &halo, int64_t &wx, int64_t &wy,
                                          int64_t &wz) {
    s = (int)(key >> g.VB);
    const uint32_t vk = key & ((1u << g.VB) - 1u);
    halo = (vk >> (g.VB - 1)) & 1u;
    if (!halo) {
        const uint32_t l = vk, Z = (uint32_t)g.Z, XY = (uint32_t)g.XY;
        const uint32_t c = l / Z;
        wz = (int64_t)(l - c * Z);
        wx = (int64_t)(c / XY);
        wy = (int64_t)(c - (c / XY) * XY);
    } else {
        const uint32_t l = vk & ((1u << (g.VB - 1)) - 1u), Zp = (uint32_t)g.Zp, XYp = (uint32_t)g.XYp;
        const uint32_t c = l / Zp;
        wz = (int64_t)(l - c * Zp) - g.ez;
        wx = (int64_t)(c / XYp) - g.ex;
        wy = (int64_t)(c - (c / XYp) * XYp) - g.ex;
    }
}

struct MomAcc {
    double S0, S1, S2, Qxx, Qxy, Qxz, Qyy, Qyz, Qzz;
    int32_t n;
    float mh;
    __device__ void zero() {
        S0 = S1 = S2 = Qxx = Qxy = Qxz = Qyy = Qyz = Qzz = 0.0;
        n = 0;
        mh = 1.0f;
    }
};

__device__ __forceinline__ void store_moment(Moment *m, const MomAcc &a) {
    m->S[0] = a.S0; m->S[1] = a.S1; m->S[2] = a.S2;
    m->Q[0] = a.Qxx; m->Q[1] = a.Qxy; m->Q[2] = a.Qxz; m->Q[3] = a.Qyy; m->Q[4] = a.Qyz; m->Q[5] = a.Qzz;
    m->n = a.n;
    m->pad = __float_as_int(a.mh);
}

// Chunk partials of one wavefront: chunk t of the scan-relative chunking (each scan's
// valid keys, from its first sorted position, in chunks of MCH), a lane per point.  A
// segmented inclusive scan (fixed shuffle tree) leaves each run's sum in its last lane,
// which stores it at that point's sorted position.  scan_cnt: valid keys per scan (the
// sorted keys are ordered by scan), so chunks are cut relative to each scan's first sorted
// position and a scan's sums do not depend on the other scans of its batch.
template <typename P>
__device__ __forceinline__ void mom_chunk(int32_t t, const P *__restrict__ world, const uint32_t *__restrict__ skeys,
                                          const uint32_t *__restrict__ svals, const ScanDesc *__restrict__ sd, int S,
                                          const Geo &g, const uint32_t *__restrict__ scan_cnt,
                                          Moment *__restrict__ part) {
    const int lane = threadIdx.x & 63;
    // lane s < S: scan s's valid keys, first sorted position and first chunk
    const int32_t cnt_s = lane < S ? (int32_t)scan_cnt[lane] : 0;
    const int32_t nch_s = (cnt_s + MCH - 1) / MCH;
    int32_t pos_s = cnt_s, off_s = nch_s;
    for (int o = 1; o < 64; o <<= 1) {
        const int32_t up = __shfl_up(pos_s, o), uc = __shfl_up(off_s, o);
        if (lane >= o) {
            pos_s += up;
            off_s += uc;
        }
    }
    pos_s -= cnt_s;
    off_s -= nch_s;
    const int32_t total = __shfl(off_s + nch_s, S - 1);
    if (t >= total) return;   // uniform across the wavefront
    // the chunk's scan: the last one whose first chunk is <= t
    const uint64_t le = __ballot(lane < S && off_s <= t);
    const int sc = 63 - __builtin_clzll(le);
    const int64_t sbeg = __shfl(pos_s, sc), send = sbeg + __shfl(cnt_s, sc);
    const int64_t j = sbeg + (int64_t)(t - __shfl(off_s, sc)) * MCH + lane;
    const bool in = j < send;
    const uint32_t key = in ? skeys[j] : GVOM_NOKEY;
    const uint32_t prev = __shfl_up(key, 1);
    int f = (lane == 0 || key != prev) ? 1 : 0;   // run head
    double v[9];
    int32_t cnt = 0;
    float mh = 1.0f;
    for (int k = 0; k < 9; ++k) v[k] = 0.0;
    if (in) {
        int s;
        bool halo;
        int64_t wx, wy, wz;
        key_voxel(key, g, s, halo, wx, wy, wz);
        double o[3];
        scan_origin(sd[s], g, o);
        const int64_t i = svals[j];
        const double r0 = (((double)world[3 * i] / g.xy_res) - o[0]) - (double)wx;
        const double r1 = (((double)world[3 * i + 1] / g.xy_res) - o[1]) - (double)wy;
        const double r2 = (((double)world[3 * i + 2] / g.z_res) - o[2]) - (double)wz;
        v[0] = r0; v[1] = r1; v[2] = r2;
        v[3] = r0 * r0; v[4] = r0 * r1; v[5] = r0 * r2;
        v[6] = r1 * r1; v[7] = r1 * r2; v[8] = r2 * r2;
        cnt = 1;
        mh = py_minf(1.0f, (float)r2);
    }
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
        double pv[9];
#pragma unroll
        for (int k = 0; k < 9; ++k) pv[k] = __shfl_up(v[k], off);
        const int32_t pc = __shfl_up(cnt, off);
        const float pm = __shfl_up(mh, off);
        const int pf = __shfl_up(f, off);
        if (lane >= off) {
            if (!f) {
#pragma unroll
                for (int k = 0; k < 9; ++k) v[k] = pv[k] + v[k];
                cnt += pc;
                mh = py_minf(pm, mh);
            }
            f |= pf;
        }
    }
    // last lane of its run inside the chunk: the next lane starts a run (padding lanes carry
    // GVOM_NOKEY, so the last valid lane always ends one)
    const uint32_t next = __shfl_down(key, 1);
    const bool last = lane == 63 || next != key;
    if (in && last) {
        MomAcc a;
        a.S0 = v[0]; a.S1 = v[1]; a.S2 = v[2];
        a.Qxx = v[3]; a.Qxy = v[4]; a.Qxz = v[5]; a.Qyy = v[6]; a.Qyz = v[7]; a.Qzz = v[8];
        a.n = cnt;
        a.mh = mh;
        store_moment(part + j, a);
    }
}

// Segments of the sorted keys and the moment chunk partials, in one launch of 256-thread
// blocks.  Every wavefront writes the partials of one chunk (mom_chunk).  Blocks b < nbs
// also own sorted keys [b SEG_TILE, (b + 1) SEG_TILE): they count heads, valid keys and per
// scan s in-map heads, halo heads and valid keys (W = 2 + 3 S words into blk[b]), publish
// the head count (flags[b] = heads | PUB_READY, zero on entry), and after their chunk add the
// head counts of the blocks before (waiting for those not yet published; a block waits only
// on lower ones, as in the radix sort) to write their segment starts and keys in sorted
// order.  The last of them then waits for every block and derives the totals.
template <typename P>
__global__ __launch_bounds__(256) void k_segments(const uint32_t *__restrict__ k, const uint32_t *__restrict__ svals,
                                                  int64_t n, int nbs, const P *__restrict__ world,
                                                  const ScanDesc *__restrict__ sd, int S, Geo g,
                                                  DevState *__restrict__ st, const uint32_t *__restrict__ scan_cnt,
                                                  uint32_t *__restrict__ flags, uint32_t *__restrict__ blk,
                                                  int32_t *__restrict__ seg_start, uint32_t *__restrict__ seg_key,
                                                  Moment *__restrict__ part) {
    __shared__ uint32_t c[2 + 3 * GVOM_MAX_BATCH];
    __shared__ uint32_t wsum[4], wpre[4];
    __shared__ int fault;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    const int b = (int)blockIdx.x;
    const bool segb = b < nbs;   // uniform
    const int W = 2 + 3 * S;
    const int64_t base = (int64_t)b * SEG_TILE + t * 8;
    uint32_t hflags = 0, cnt = 0;
    if (segb) {
        if (t < W) c[t] = 0;
        if (t == 0) fault = 0;
        __syncthreads();
        uint32_t valid = 0;
        for (int jj = 0; jj < 8; ++jj) {
            const int64_t i = base + jj;
            if (i >= n) break;
            const uint32_t key = k[i];
            if (key == GVOM_NOKEY) continue;
            ++valid;
            const int sc = (int)(key >> g.VB);
            if (S > 1) atomicAdd(&c[4 + 3 * sc], 1u);
            if (i == 0 || k[i - 1] != key) {
                hflags |= 1u << jj;
                ++cnt;
                atomicAdd(&c[2 + 3 * sc + (int)((key >> (g.VB - 1)) & 1u)], 1u);
            }
        }
        uint32_t hs = cnt, vs = valid;
        for (int off = 32; off > 0; off >>= 1) {
            hs += __shfl_down(hs, off);
            vs += __shfl_down(vs, off);
        }
        if (lane == 0) {
            atomicAdd(&c[0], hs);
            atomicAdd(&c[1], vs);
        }
        __syncthreads();
        if (t == 0) {
            for (int q = 0; q < W; ++q) blk[(int64_t)b * W + q] = c[q];
            __hip_atomic_store(&flags[b], c[0] | PUB_READY, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
        }
    }
    mom_chunk<P>((int32_t)(b * 4 + w), world, k, svals, sd, S, g, scan_cnt, part);
    if (!segb) return;
    // heads in the blocks before this one
    uint32_t pre = 0;
    for (int bb = t; bb < b; bb += 256) pre += pub_wait(&flags[bb], fault);
    for (int off = 32; off > 0; off >>= 1) pre += __shfl_down(pre, off);
    if (lane == 0) wpre[w] = pre;
    uint32_t v = cnt;
    for (int off = 1; off < 64; off <<= 1) {
        uint32_t u = __shfl_up(v, off);
        if (lane >= off) v += u;
    }
    if (lane == 63) wsum[w] = v;
    __syncthreads();
    uint32_t r = wpre[0] + wpre[1] + wpre[2] + wpre[3];
    for (int ww = 0; ww < w; ++ww) r += wsum[ww];
    r += v - cnt;
    for (int jj = 0; jj < 8; ++jj) {
        if (hflags & (1u << jj)) {
            const int64_t i = base + jj;
            seg_start[r] = (int32_t)i;
            seg_key[r] = k[i];
            ++r;
        }
    }
    if (b == nbs - 1) {
        // every block's counts were published before the head counts this block waited for
        // (release / acquire): the totals
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        seg_bounds_body(blk, nbs, S, g, seg_start, st);
    }
    if (t == 0 && fault) st->wait_fault = 1;
}

__global__ __launch_bounds__(256) void k_mom_seg(const int32_t *__restrict__ seg_start,
                                                 const uint32_t *__restrict__ seg_key,
                                                 const Moment *__restrict__ part, Geo g,
                                                 const DevState *__restrict__ st,
                                                 int32_t *__restrict__ tables, SlotRec *__restrict__ recs,
                                                 int64_t ncap, Moment *__restrict__ mom) {
    const int64_t gi = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (gi >= st->nseg) return;
    const int32_t a = seg_start[gi], b = seg_start[gi + 1];
    const int64_t sbeg = st->scan_pos[(int)(seg_key[gi] >> g.VB)];
    const int64_t ra = a - sbeg, rb = b - sbeg;   // scan-relative positions
    Moment m;
    for (int k = 0; k < 3; ++k) m.S[k] = 0.0;
    for (int k = 0; k < 6; ++k) m.Q[k] = 0.0;
    int32_t n = 0;
    float mh = 1.0f;
    for (int64_t c = ra / MCH; c <= (rb - 1) / MCH; ++c) {
        const int64_t pos = sbeg + ((c + 1) * MCH < rb ? (c + 1) * MCH : rb) - 1;
        const Moment &p = part[pos];
        for (int k = 0; k < 3; ++k) m.S[k] += p.S[k];
        for (int k = 0; k < 6; ++k) m.Q[k] += p.Q[k];
        n += p.n;
        mh = py_minf(mh, __int_as_float(p.pad));
    }
    m.n = n;
    m.pad = 0;
    mom[gi] = m;
    const uint32_t key = seg_key[gi];
    const uint32_t vk = key & ((1u << g.VB) - 1u);
    if (!((vk >> (g.VB - 1)) & 1u)) {
        const int s = (int)(key >> g.VB);
        const int buf = new_buf(st->flip, (st->next_slot + s) % g.B);
        const int32_t id = (int32_t)(gi - st->first_seg[s]);
        int32_t *tv = tables + (int64_t)buf * g.V + vk;
        const int32_t passes = *tv;
        *tv = (int32_t)(GVOM_OCC | (uint32_t)id);
        SlotRec *r = recs + (int64_t)buf * ncap + id;
        r->hit = n;
        r->total = (int32_t)((uint32_t)passes + (uint32_t)n);
        r->min_h = mh;
        r->pad = 0;
    }
}

// Neighbourhood statistics of one occupied voxel u (gvom.py:1250-1387): every valid
// point p whose own voxel w lies within (+-xy_eigen_dist, +-z_eigen_dist) of u adds
// local = p/res - origin - u.  With r = p/res - origin - w and d = w - u, segment w
// contributes sum(local) = S_w + n_w d and, about the mean m with a = d - m,
// sum((local - m)(local - m)^T) = Q_w + S_w a^T + a S_w^T + n_w a a^T.
__device__ __forceinline__ int64_t neighbour_segment(int64_t x, int64_t y, int64_t z, int s, const Geo &g,
                                                     const int32_t *__restrict__ table, int32_t fs,
                                                     int32_t inb_end, int32_t nseg,
                                                     const uint32_t *__restrict__ seg_key) {
    if (x >= 0 && x < g.XY && y >= 0 && y < g.XY && z >= 0 && z < g.Z) {
        int32_t tv = table[(x * g.XY + y) * g.Z + z];
        if (tv >= 0) return -1;
        return fs + (int64_t)(tv & 0x7FFFFFFF);
    }
    // halo voxel: binary search among this scan's halo segments
    int64_t kx = x + g.ex, ky = y + g.ex, kz = z + g.ez;
    uint32_t key = ((uint32_t)s << g.VB) | (1u << (g.VB - 1)) | (uint32_t)((kx * g.XYp + ky) * g.Zp + kz);
    int32_t lo = inb_end, hi = nseg;
    while (lo < hi) {
        int32_t mid = (lo + hi) >> 1;
        if (seg_key[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return (lo < nseg && seg_key[lo] == key) ? lo : -1;
}

// 32 lanes (half a wavefront) per occupied voxel, one lane per neighbour; sums over
// neighbours are shuffle-tree reductions in a fixed lane order.
__device__ __forceinline__ double red32(double v) {
#pragma unroll
    for (int off = 16; off > 0; off >>= 1) v += __shfl_xor(v, off, 32);
    return v;
}

__global__ __launch_bounds__(256) void k_metrics(const uint32_t *__restrict__ seg_key,
                                                 const Moment *__restrict__ mom, Geo g,
                                                 DevState *__restrict__ st,
                                                 const int32_t *__restrict__ tables,
                                                 SlotRec *__restrict__ recs, int64_t ncap, int S) {
    const int l32 = threadIdx.x & 31;
    const int32_t nseg = st->nseg;
    const int64_t groups = ((int64_t)gridDim.x * blockDim.x) >> 5;
    const int nx = 2 * g.ex + 1, nz = 2 * g.ez + 1, nn = nx * nx * nz;
    // grid-stride over segments; each 32-lane group handles one occupied voxel at a time
    for (int64_t gi = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 5; gi < nseg; gi += groups) {
        const uint32_t key = seg_key[gi];
        const uint32_t vk = key & ((1u << g.VB) - 1u);
        if ((vk >> (g.VB - 1)) & 1u) continue;   // halo segment: no record (uniform in the group)
        const int s = (int)(key >> g.VB);
        const int buf = st->stage_buf[s];   // the commit may already run beside this kernel
        const int32_t fs = st->first_seg[s], ie = st->inb_end[s];
        const int32_t *table = tables + (int64_t)buf * g.V;
        const uint32_t Z = (uint32_t)g.Z, XY = (uint32_t)g.XY;
        const uint32_t uc = vk / Z;
        const int64_t uz = vk - uc * Z, ux = uc / XY, uy = uc - (uc / XY) * XY;

        // first neighbour of this lane (nn <= 32 for the default eigen distances): its
        // segment is looked up once and reused by the second pass
        int64_t w0 = -1;
        int dx0 = 0, dy0 = 0, dz0 = 0;
        if (l32 < nn) {
            dz0 = l32 % nz - g.ez;
            dy0 = (l32 / nz) % nx - g.ex;
            dx0 = l32 / (nz * nx) - g.ex;
            w0 = neighbour_segment(ux + dx0, uy + dy0, uz + dz0, s, g, table, fs, ie, nseg, seg_key);
        }
        double n = 0.0, m0 = 0.0, m1 = 0.0, m2 = 0.0;
        for (int k = l32; k < nn; k += 32) {
            int dx = dx0, dy = dy0, dz = dz0;
            int64_t w = w0;
            if (k != l32) {
                dz = k % nz - g.ez;
                dy = (k / nz) % nx - g.ex;
                dx = k / (nz * nx) - g.ex;
                w = neighbour_segment(ux + dx, uy + dy, uz + dz, s, g, table, fs, ie, nseg, seg_key);
            }
            if (w < 0) continue;
            const Moment &M = mom[w];
            double nw = (double)M.n;
            m0 += M.S[0] + nw * (double)dx;
            m1 += M.S[1] + nw * (double)dy;
            m2 += M.S[2] + nw * (double)dz;
            n += nw;
        }
        n = red32(n);
        m0 = red32(m0) / n;
        m1 = red32(m1) / n;
        m2 = red32(m2) / n;
        double cxx = 0, cxy = 0, cxz = 0, cyy = 0, cyz = 0, czz = 0;
        for (int k = l32; k < nn; k += 32) {
            int dx = dx0, dy = dy0, dz = dz0;
            int64_t w = w0;
            if (k != l32) {
                dz = k % nz - g.ez;
                dy = (k / nz) % nx - g.ex;
                dx = k / (nz * nx) - g.ex;
                w = neighbour_segment(ux + dx, uy + dy, uz + dz, s, g, table, fs, ie, nseg, seg_key);
            }
            if (w < 0) continue;
            const Moment &M = mom[w];
            double nw = (double)M.n;
            double a0 = (double)dx - m0, a1 = (double)dy - m1, a2 = (double)dz - m2;
            cxx += M.Q[0] + 2.0 * M.S[0] * a0 + nw * a0 * a0;
            cxy += M.Q[1] + M.S[0] * a1 + a0 * M.S[1] + nw * a0 * a1;
            cxz += M.Q[2] + M.S[0] * a2 + a0 * M.S[2] + nw * a0 * a2;
            cyy += M.Q[3] + 2.0 * M.S[1] * a1 + nw * a1 * a1;
            cyz += M.Q[4] + M.S[1] * a2 + a1 * M.S[2] + nw * a1 * a2;
            czz += M.Q[5] + 2.0 * M.S[2] * a2 + nw * a2 * a2;
        }
        cxx = red32(cxx); cxy = red32(cxy); cxz = red32(cxz);
        cyy = red32(cyy); cyz = red32(cyz); czz = red32(czz);
        if (l32 == 0) {
            SlotRec *r = recs + (int64_t)buf * ncap + (gi - fs);
            r->m[0] = m0;
            r->m[1] = m1;
            r->m[2] = m2;
            bool pos = n > 0;   // gvom.py:1383-1385
            r->m[3] = pos ? cxx / n : 0.0;
            r->m[4] = pos ? cxy / n : 0.0;
            r->m[5] = pos ? cxz / n : 0.0;
            r->m[6] = pos ? cyy / n : 0.0;
            r->m[7] = pos ? cyz / n : 0.0;
            r->m[8] = pos ? czz / n : 0.0;
            r->m[9] = n;
        }
    }
}

// Ring bookkeeping after the scan stage of S scans (gvom.py:211-216): slot origins for
// combine (gvom.py:208), the ego the height map uses (gvom.py:110-112), the written slots'
// buffer bits and cell counts, the ring indices and the next combine's source shifts.
// Everything the combine reads about the ring is written here, after the scan stage, so that
// the stage's kernels may run while a combine still reads the previous state.  One wavefront:
// lane s < S computes the origin of the batch's s-th slot, lane j < B reads slot j's origin
// and writes its combine shift, so the loads are issued together, not one after another.
__global__ __launch_bounds__(64) void k_commit_process(DevState *st, const ScanDesc *sd, Geo g, int S) {
    if (blockIdx.x != 0) return;
    const int lane = threadIdx.x;
    const int B = g.B;
    const int next = st->next_slot, mask0 = st->filled_mask, flip0 = st->flip;
    // lane s < S: the new origin of slot (next + s) % B and the scan's cell count
    double no[3] = {0.0, 0.0, 0.0};
    int32_t cells = 0;
    if (lane < S) {
        scan_origin(sd[lane], g, no);
        cells = st->scan_cells[lane];
    }
    // lane j < B: slot j's origin after this commit (new if the batch writes slot j)
    double so[3] = {0.0, 0.0, 0.0};
    const int wj = lane < B ? (lane - next + B) % B : 64;   // batch index that writes slot `lane`
    if (lane < B && wj >= S) {
        so[0] = st->slot_origin[lane][0];
        so[1] = st->slot_origin[lane][1];
        so[2] = st->slot_origin[lane][2];
    }
    const int src = wj < S ? wj : 0;
    const double n0 = __shfl(no[0], src), n1 = __shfl(no[1], src), n2 = __shfl(no[2], src);
    if (lane < B && wj < S) {
        so[0] = n0;
        so[1] = n1;
        so[2] = n2;
    }
    int mask = mask0, flip = flip0;
    for (int k = 0; k < S; ++k) {
        const int slot = (next + k) % B;
        mask |= 1 << slot;
        flip ^= 1 << slot;   // the buffer the stage wrote now holds the slot
    }
    const int last = (next + S - 1) % B;
    // combined origin = newest slot's (gvom.py:229); each slot is read at (combined voxel +
    // origin difference) (gvom.py:1016-1022)
    const double c0 = __shfl(so[0], last), c1 = __shfl(so[1], last), c2 = __shfl(so[2], last);
    if (lane < S) {
        const int slot = (next + lane) % B;
        st->slot_origin[slot][0] = no[0];
        st->slot_origin[slot][1] = no[1];
        st->slot_origin[slot][2] = no[2];
        st->slot_cells[slot] = cells;
    }
    if (lane < B) {
        st->comb_shift[lane][0] = (int32_t)(c0 - so[0]);
        st->comb_shift[lane][1] = (int32_t)(c1 - so[1]);
        st->comb_shift[lane][2] = (int32_t)(c2 - so[2]);
        st->comb_shift[lane][3] = (mask >> lane) & 1;
    }
    if (lane == 0) {
        st->latest_ego[0] = sd[S - 1].ego[0];
        st->latest_ego[1] = sd[S - 1].ego[1];
        st->latest_ego[2] = sd[S - 1].ego[2];
        st->filled_mask = mask;
        st->flip = flip;
        st->last_slot = last;
        st->next_slot = (next + S) % B;
    }
}


// Zero the slot tables of the S slots this batch writes (slot choice read on device).
// Each thread clears 4 consecutive words of one table.
__global__ __launch_bounds__(256) void k_clear_slots(int32_t *__restrict__ tables, const DevState *__restrict__ st,
                                                     int B, int32_t V) {
    // grid (ceil(V / 4096), S): block (b, s) clears words [4096 b, 4096 (b + 1)) of the table
    // of the batch's s-th slot; a wavefront stores 1 KiB contiguous per round
    const int s = blockIdx.y;
    const int buf = __builtin_amdgcn_readfirstlane(new_buf(st->flip, (st->next_slot + s) % B));
    int32_t *t = tables + (int64_t)buf * V;
    const int32_t base = (int32_t)blockIdx.x * 4096;
    if ((V & 3) == 0) {
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int32_t j = base + (k * 256 + (int32_t)threadIdx.x) * 4;
            if (j < V) *reinterpret_cast<int4 *>(t + j) = make_int4(0, 0, 0, 0);
        }
    } else {
        for (int k = 0; k < 16; ++k) {
            const int32_t j = base + k * 256 + (int32_t)threadIdx.x;
            if (j < V) t[j] = 0;
        }
    }
}

// One word per thread, for buffers that are not 16-byte aligned (fields inside DevState).
__global__ __launch_bounds__(256) void k_fill_i32_words(int32_t *__restrict__ p, int32_t v, int64_t n) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(256) void k_fill_i32(int32_t *__restrict__ p, int32_t v, int64_t n) {
    int64_t i = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) * 4;
    if (i + 3 < n) {
        *reinterpret_cast<int4 *>(p + i) = make_int4(v, v, v, v);
    } else {
        for (; i < n; ++i) p[i] = v;
    }
}

inline unsigned nblocks(int64_t n, int per) { return (unsigned)((n + per - 1) / per > 0 ? (n + per - 1) / per : 1); }

}  // namespace

// ---- launchers ----------------------------------------------------------------------
hipError_t launch_fill_i32(int32_t *p, int32_t v, int64_t n, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    if (((uintptr_t)p % 16) == 0)
        k_fill_i32<<<nblocks((n + 3) / 4, 256), 256, 0, s>>>(p, v, n);
    else
        k_fill_i32_words<<<nblocks(n, 256), 256, 0, s>>>(p, v, n);
    return hipGetLastError();
}

hipError_t launch_prepare(int dtype, const void *in, void *world, uint32_t *keys, uint32_t *vals,
                          uint8_t *valid, const ScanDesc *sd, int S, int64_t NT, const Geo &g,
                          DevState *st, uint8_t *rsec, uint32_t *hist, int64_t hwords, hipStream_t s) {
    unsigned nb = nblocks(NT, 256);
    if (dtype == GVOM_F32)
        k_prepare<float><<<nb, 256, 0, s>>>((const float *)in, (float *)world, keys, vals, valid, sd, S, NT, g, st,
                                             rsec, hist, hwords);
    else
        k_prepare<double><<<nb, 256, 0, s>>>((const double *)in, (double *)world, keys, vals, valid, sd, S, NT, g, st,
                                              rsec, hist, hwords);
    return hipGetLastError();
}

hipError_t launch_ray_order(const uint8_t *rsec, int64_t NT, DevState *st, uint32_t *order, hipStream_t s) {
    if (NT <= 0) return hipSuccess;
    k_ray_order<<<nblocks(NT, 256), 256, 0, s>>>(rsec, NT, st, order);
    return hipGetLastError();
}

hipError_t launch_raycast(int dtype, const void *world, const uint8_t *valid, const ScanDesc *sd,
                          int S, int64_t NT, const Geo &g, const DevState *st, int32_t *tables,
                          int64_t table_words, int block, int rpt, int use_box, int tiled, int force_slot,
                          const uint32_t *order, int spread, int64_t ring_stride, int patch_rings, hipStream_t s) {
    if (NT <= 0) return hipSuccess;
    if (block != 64 && block != 128 && block != 256) return hipErrorInvalidValue;
    if (tiled && use_box) return hipErrorInvalidValue;
    if (S < 1 || S > GVOM_MAX_BATCH || !gvom_plan::raycast_keys_fit(g.V, g.XY, g.Z, S, tiled != 0))
        return hipErrorInvalidValue;
    const int64_t need = tiled ? (int64_t)S * gvom_plan::tile_words(g.XY, g.Z)
                               : gvom_plan::slot_table_footprint(g.V, g.B, force_slot);
    if (need > table_words) return hipErrorInvalidValue;   // the launch would write past `tables`
    if (patch_rings != 0 && !gvom_plan::ray_patch_valid(ring_stride, patch_rings)) return hipErrorInvalidValue;
    if (rpt < 1) rpt = 1;
    int maxit = (g.XY > g.Z ? g.XY : g.Z) + 2;
    unsigned nb = nblocks(NT, block * rpt);
    if (order) nb *= 8;   // 8 sectors, each sized for every ray in the worst case
    const size_t lds = use_box ? BOX_N * sizeof(int32_t) : 0;
#define RAY_LAUNCH(P, BOX, TL)                                                                              \
    k_raycast<P, BOX, TL><<<nb, block, BOX ? lds : 0, s>>>((const P *)world, valid, sd, S, NT, g, st, tables, \
                                                           maxit, rpt, force_slot, order, spread, ring_stride, \
                                                           patch_rings)
    if (dtype == GVOM_F32) {
        if (tiled) RAY_LAUNCH(float, false, true);
        else if (use_box) RAY_LAUNCH(float, true, false);
        else RAY_LAUNCH(float, false, false);
    } else {
        if (tiled) RAY_LAUNCH(double, false, true);
        else if (use_box) RAY_LAUNCH(double, true, false);
        else RAY_LAUNCH(double, false, false);
    }
#undef RAY_LAUNCH
    return hipGetLastError();
}

// Tiled scratch -> slot tables, and the scratch back to zero: one block per 4 x 4 tile
// column of one scan of the batch.  The tile column's z_size * 16 words are contiguous in
// the scratch (16-byte loads and zero stores); each of its 4 x-rows is 4 * z_size
// contiguous words of the slot table and is written by one wavefront.
__global__ __launch_bounds__(256) void k_tile_convert(Geo g, const DevState *__restrict__ st, int force_slot,
                                                      int32_t *__restrict__ scratch, int32_t *__restrict__ tables) {
    // [z][17]: the tile's 16 columns at one z, padded by one word so that a wavefront reading
    // 4 z-consecutive words of 16 columns spreads over the banks
    __shared__ int32_t sh[256 * 17];   // z_size <= 256
    const TileGeo tg = tile_geo(g);
    const int tcol = blockIdx.x, s = blockIdx.y;
    const int txi = tcol / tg.ty, tyi = tcol - (tcol / tg.ty) * tg.ty;
    const int Z = g.Z, XY = g.XY;
    // the tile column is 64-byte aligned (vt and z_size * 16 are multiples of 16 words)
    int4 *src = reinterpret_cast<int4 *>(scratch + (int64_t)s * tg.vt + (int64_t)tcol * Z * 16);
    for (int k = threadIdx.x; k < Z * 4; k += 256) {
        const int4 v = src[k];
        // words no ray reached are zero already: skipping their store saves HBM writes
        if (v.x | v.y | v.z | v.w) src[k] = make_int4(0, 0, 0, 0);
        int32_t *d = sh + (k >> 2) * 17 + (k & 3) * 4;
        d[0] = v.x;
        d[1] = v.y;
        d[2] = v.z;
        d[3] = v.w;
    }
    __syncthreads();
    const int buf = force_slot >= 0 ? force_slot : new_buf(st->flip, (st->next_slot + s) % g.B);
    int32_t *tab = tables + (int64_t)buf * g.V;
    // wavefront a writes x-row a: columns (x, y0 .. y0 + ny - 1), ny * Z contiguous words
    const int a = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int x = txi * 4 + a;
    const int y0 = tyi * 4, ny = (XY - y0) < 4 ? (XY - y0) : 4;
    if (x >= XY) return;
    int32_t *row = tab + ((int64_t)x * XY + y0) * Z;
    const int n = ny * Z;
    if ((Z & 3) == 0) {
        // 16-byte stores: words k .. k + 3 are z .. z + 3 of column b
        for (int k = lane * 4; k < n; k += 256) {
            const int b = k / Z, z = k - (k / Z) * Z;
            const int32_t *q = sh + z * 17 + a * 4 + b;
            *reinterpret_cast<int4 *>(row + k) = make_int4(q[0], q[17], q[34], q[51]);
        }
    } else {
        for (int k = lane; k < n; k += 64) {
            const int b = k / Z, z = k - (k / Z) * Z;
            row[k] = sh[z * 17 + a * 4 + b];
        }
    }
}

hipError_t launch_tile_convert(const Geo &g, const DevState *st, int S, int force_slot, int32_t *scratch,
                               int64_t scratch_words, int32_t *tables, int64_t table_words, hipStream_t s) {
    const TileGeo tg = tile_geo(g);
    if (S < 1 || (int64_t)S * tg.vt > scratch_words ||
        gvom_plan::slot_table_footprint(g.V, g.B, force_slot) > table_words || g.Z > 256)
        return hipErrorInvalidValue;
    k_tile_convert<<<dim3((unsigned)(tg.tx * tg.ty), (unsigned)S), 256, 0, s>>>(g, st, force_slot, scratch, tables);
    return hipGetLastError();
}

int64_t tile_scratch_words(const Geo &g) { return gvom_plan::tile_words(g.XY, g.Z); }

int64_t radix_hist_entries(int64_t n) { return gvom_plan::radix_hist_words(n); }

int64_t segment_block_words(int64_t n) { return gvom_plan::segment_block_words(n); }

hipError_t radix_sort(uint32_t *k0, uint32_t *v0, uint32_t *k1, uint32_t *v1, int64_t n, int bits, int scan_shift,
                      uint32_t *hist, int64_t hwords, DevState *st, uint32_t **kout, uint32_t **vout, hipStream_t s) {
    *kout = k0;
    *vout = v0;
    if (n <= 0) return hipSuccess;
    if (bits > 32 || scan_shift < 1 || scan_shift > 31 || (bits - scan_shift) > 5) return hipErrorInvalidValue;
    if (gvom_plan::radix_hist_words(n) > hwords) return hipErrorInvalidValue;   // hist too small for n keys
    const int nblk = (int)nblocks(n, RS_TILE);
    const int passes = (bits + 7) / 8;
    // set p >= 1 of hist: pass p's published tile counts (256 per tile), then its digit totals
    auto status = [&](int p) { return hist + (int64_t)p * hwords; };
    auto totals = [&](int p) { return hist + (int64_t)p * hwords + gvom_plan::radix_totals_at(n); };
    uint32_t *ka = k0, *va = v0, *kb = k1, *vb = v1;
    for (int p = 0; p < passes; ++p) {
        uint32_t *tn = p + 1 < passes ? totals(p + 1) : nullptr;
        if (p == 0) {
            k_rs_scatter<<<nblk, RS_THREADS, 0, s>>>(ka, va, kb, vb, n, 0, hist, nblk, tn, scan_shift,
                                                      hist + hwords + gvom_plan::radix_scan_counts_at(n));
            if (passes == 1) {
                hipError_t e = launch_fill_i32((int32_t *)hist, 0, hwords, s);
                if (e != hipSuccess) return e;
            }
        } else {
            k_rs_scatter_lb<<<nblk, RS_THREADS, 0, s>>>(ka, va, kb, vb, n, 8 * p, status(p), totals(p), tn, nblk,
                                                         p + 1 == passes ? hist : nullptr, hwords, st);
        }
        uint32_t *tk = ka; ka = kb; kb = tk;
        uint32_t *tv = va; va = vb; vb = tv;
    }
    *kout = ka;
    *vout = va;
    return hipGetLastError();
}

hipError_t launch_segments(int dtype, const void *world, const uint32_t *skeys, const uint32_t *svals, int64_t n,
                           const ScanDesc *sd, int S, const Geo &g, DevState *st, uint32_t *hist, int64_t hwords,
                           uint32_t *blk, int64_t blk_words, int32_t *seg_start, uint32_t *seg_key, Moment *part,
                           hipStream_t s) {
    if (S < 1 || S > GVOM_MAX_BATCH || gvom_plan::segment_block_footprint(n, S) > blk_words)
        return hipErrorInvalidValue;   // blk too small for n keys of S scans
    if (n <= 0) {
        k_seg_bounds<<<1, 64, 0, s>>>(blk, 0, S, g, seg_start, st);
        return hipGetLastError();
    }
    if (gvom_plan::radix_hist_words(n) > hwords) return hipErrorInvalidValue;
    const int nbs = (int)nblocks(n, SEG_TILE);
    // one wavefront per chunk: at most ceil(n / MCH) + S scan-relative chunks, 4 per block
    const int64_t nchunk = (n + MCH - 1) / MCH + S;
    const int64_t nb = std::max<int64_t>(nbs, (nchunk + 3) / 4);
    const uint32_t *scan_cnt = hist + hwords + gvom_plan::radix_scan_counts_at(n);
    uint32_t *flags = hist + hwords + gvom_plan::segment_flags_at(n);
    if (dtype == GVOM_F32)
        k_segments<float><<<(unsigned)nb, 256, 0, s>>>(skeys, svals, n, nbs, (const float *)world, sd, S, g, st,
                                                       scan_cnt, flags, blk, seg_start, seg_key, part);
    else
        k_segments<double><<<(unsigned)nb, 256, 0, s>>>(skeys, svals, n, nbs, (const double *)world, sd, S, g, st,
                                                        scan_cnt, flags, blk, seg_start, seg_key, part);
    return hipGetLastError();
}

hipError_t launch_mom_seg(const int32_t *seg_start, const uint32_t *seg_key, int64_t seg_cap, const Geo &g,
                          DevState *st, int32_t *slot_tables, SlotRec *slot_recs, int64_t ncap, const Moment *part,
                          Moment *mom, hipStream_t s) {
    k_mom_seg<<<nblocks(seg_cap, 256), 256, 0, s>>>(seg_start, seg_key, part, g, st, slot_tables, slot_recs, ncap,
                                                    mom);
    return hipGetLastError();
}

hipError_t launch_metrics(const uint32_t *seg_key, const Moment *mom, int64_t seg_cap, const Geo &g,
                          DevState *st, const int32_t *slot_tables, SlotRec *slot_recs, int64_t ncap, int S,
                          hipStream_t s) {
    // grid-stride: enough 32-lane groups to fill the chip, never more than segments can need
    int64_t nb = nblocks(seg_cap * 32, 256);
    if (nb > 2048) nb = 2048;
    if (nb < 1) nb = 1;
    k_metrics<<<(unsigned)nb, 256, 0, s>>>(seg_key, mom, g, st, slot_tables, slot_recs, ncap, S);
    return hipGetLastError();
}

hipError_t launch_commit_process(DevState *st, const ScanDesc *sd, const Geo &g, int S, hipStream_t s) {
    k_commit_process<<<1, 64, 0, s>>>(st, sd, g, S);
    return hipGetLastError();
}

hipError_t launch_clear_slots(int32_t *slot_tables, const DevState *st, int S, const Geo &g, hipStream_t s) {
    k_clear_slots<<<dim3((unsigned)((g.V + 4095) / 4096), (unsigned)S), 256, 0, s>>>(slot_tables, st, g.B,
                                                                                    (int32_t)g.V);
    return hipGetLastError();
}

