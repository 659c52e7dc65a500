// Combine stage of the scan->map path on gfx950 (reference Gvom.combine_maps,
// scripts/gvom.py:222-393).
//
// Launch sequence (stream-ordered, no host synchronisation):
//   tables    one wavefront per z-contiguous column: for each voxel, union of the
//             filled slots' occupancy (slot order) and sum of their free passes, then
//             the previous combined map with its <= 10-pass carry-over rule
//             (gvom.py:1010-1063).  Writes OCC | source-mask or the pass count, and
//             the column's occupied count.
//   scan      exclusive scan of the column counts -> first cell id of each column,
//             so ids follow the z-contiguous voxel order (a legal permutation of the
//             reference's atomic order).
//   records   one wavefront per column: assign ids, merge the sources' records in
//             slot order then the previous map (gvom.py:889-980), eigenvalues
//             (1424-1487), and the column passes: height (561-576), inferred height
//             (580-590), positive-obstacle density window (530-555).
//   maps2d    16x16 column tiles with a 15-cell halo in LDS: slope and roughness
//             (718-805), guessed height delta (593-714), positive/negative obstacle
//             and visibility maps (445-555).
//   commit    ping-pong parity of the combined buffers, flipped by the column scan's
//             first thread (every later kernel of the combine reads the new parity), and
//             the new map's cell count, written by the record merge.
#include <hip/hip_runtime.h>

#include "gvom_kernels.h"


namespace {

constexpr uint32_t MASK31 = 0x7FFFFFFFu;
constexpr int LAST_BIT = 30;   // source-mask bit of the previous combined map

__device__ __forceinline__ float py_maxf(float a, float b) { return (b > a) ? b : a; }
__device__ __forceinline__ float py_minf(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ double py_max(double a, double b) { return (b > a) ? b : a; }
__device__ __forceinline__ double py_min(double a, double b) { return (b < a) ? b : a; }

struct SrcCol {
    const int32_t *col;   // source column base (nullptr when the shifted column is off-map)
    int32_t dz;
};

__device__ __forceinline__ SrcCol source_column(const int32_t *table, const double co[3], const double so[3],
                                                int64_t x, int64_t y, const Geo &g) {
    SrcCol r;
    int64_t dx = (int64_t)(co[0] - so[0]), dy = (int64_t)(co[1] - so[1]);
    r.dz = (int32_t)(co[2] - so[2]);
    int64_t ux = x + dx, uy = y + dy;
    if (ux < 0 || ux >= g.XY || uy < 0 || uy >= g.XY) r.col = nullptr;
    else r.col = table + (ux * g.XY + uy) * g.Z;
    return r;
}

// ------------------------------------------------------------------------------------
// One wavefront handles CPW consecutive columns (lane = z).  The shifted source columns
// of all sources of a group are computed from uniform data first; then every lane
// issues its CPW x group loads before it uses the first value, so a wavefront keeps
// CPW * (group + 1) independent loads in flight.
constexpr int CPW = 4;
constexpr int SRC_GROUP = 8;

__device__ __forceinline__ void col_xy(int64_t c, const Geo &g, int64_t &x, int64_t &y) {
    x = c / g.XY;
    y = c - x * g.XY;
}

// Wave-uniform int32 in a scalar register.
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }

__global__ __launch_bounds__(256) void k_comb_tables(Geo g, const DevState *__restrict__ st,
                                                     const int32_t *__restrict__ slot_tables,
                                                     int32_t *__restrict__ comb_tables,
                                                     int32_t *__restrict__ col_count,
                                                     uint64_t *__restrict__ occ_bits,
                                                     double *__restrict__ height,
                                                     double *__restrict__ inferred) {
    // Everything but the z index is uniform across the wavefront: column coordinates,
    // each source's shift and column base live in scalar registers, so a source word costs
    // one vector compare and one load (SGPR base + lane offset).  All offsets fit in 32 bits
    // (V * (B + 1) < 2^31).
    const int lane = threadIdx.x & 63;
    const int32_t XY = g.XY, Z = g.Z, X = (int32_t)g.X, V = (int32_t)g.V;
    const int32_t c0 = uni((int32_t)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * CPW);
    if (c0 >= X) return;
    const int parity = uni(st->parity), cur = 1 - parity;
    const int filled = uni(st->filled_mask), have_last = uni(st->have_last);
    const int last_slot = uni(st->last_slot), flip = uni(st->flip);
    const double *co = st->slot_origin[last_slot];   // combined origin = newest slot's (gvom.py:229)
    int32_t cx[CPW], cy[CPW];
    {
        const int32_t x0 = uni(c0 / XY), y0 = uni(c0 - (c0 / XY) * XY);
        int32_t x = x0, y = y0;
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
            cx[j] = x;
            cy[j] = y;
            if (++y == XY) { y = 0; ++x; }
        }
    }
    const int32_t *ltab = comb_tables + (int64_t)parity * V;
    uint32_t count[CPW] = {0, 0, 0, 0};
    int first_free[CPW] = {-1, -1, -1, -1};
    const int NW = (Z + 63) >> 6;   // occupancy words per column
    for (int z0 = 0; z0 < Z; z0 += 64) {
        const int z = z0 + lane;
        const bool zin = z < Z;
        // previous combined map
        int32_t vl[CPW];
        bool okl[CPW];
        {
            const int32_t dx = have_last ? uni((int32_t)(co[0] - st->comb_origin[parity][0])) : 0;
            const int32_t dy = have_last ? uni((int32_t)(co[1] - st->comb_origin[parity][1])) : 0;
            const int32_t dz = have_last ? uni((int32_t)(co[2] - st->comb_origin[parity][2])) : 0;
            const int zl = z + dz;
            const bool zok = have_last && zin && zl >= 0 && zl < Z;
#pragma unroll
            for (int j = 0; j < CPW; ++j) {
                const int32_t ux = cx[j] + dx, uy = cy[j] + dy;
                const bool cok = c0 + j < X && ux >= 0 && ux < XY && uy >= 0 && uy < XY;   // uniform
                const int32_t *col = ltab + uni(cok ? (ux * XY + uy) * Z : 0);
                okl[j] = cok && zok;
                vl[j] = okl[j] ? col[zl] : 0;
            }
        }
        uint32_t mask[CPW] = {0, 0, 0, 0}, P[CPW] = {0, 0, 0, 0};
        for (int s0 = 0; s0 < g.B; s0 += SRC_GROUP) {
            int32_t v[SRC_GROUP][CPW];
#pragma unroll
            for (int k = 0; k < SRC_GROUP; ++k) {
                const int s = s0 + k;
                const bool use = s < g.B && ((filled >> s) & 1);
                const double *so = st->slot_origin[use ? s : 0];
                const int32_t dx = uni((int32_t)(co[0] - so[0])), dy = uni((int32_t)(co[1] - so[1]));
                const int32_t dz = uni((int32_t)(co[2] - so[2]));
                const int zs = z + dz;
                const bool zok = use && zin && zs >= 0 && zs < Z;
                const int32_t *tab = slot_tables + (int64_t)cur_buf(flip, use ? s : 0) * V;
#pragma unroll
                for (int j = 0; j < CPW; ++j) {
                    const int32_t ux = cx[j] + dx, uy = cy[j] + dy;
                    const bool cok = c0 + j < X && ux >= 0 && ux < XY && uy >= 0 && uy < XY;   // uniform
                    const int32_t *col = tab + uni(cok ? (ux * XY + uy) * Z : 0);
                    // unused sources and off-map cells read as "no data" (0)
                    v[k][j] = (cok && zok) ? col[zs] : 0;
                }
            }
#pragma unroll
            for (int k = 0; k < SRC_GROUP; ++k)
#pragma unroll
                for (int j = 0; j < CPW; ++j) {
                    if (v[k][j] < 0) mask[j] |= 1u << (s0 + k);
                    else P[j] += (uint32_t)v[k][j];
                }
        }
        int32_t *out = comb_tables + (int64_t)cur * V + c0 * Z;
#pragma unroll
        for (int j = 0; j < CPW; ++j) {
            uint32_t m = mask[j], p = P[j];
            if (okl[j]) {
                if (m) {
                    if (vl[j] < 0) m |= 1u << LAST_BIT;
                } else if (vl[j] < 0) {
                    if (p <= 10u) m = 1u << LAST_BIT;   // carry-over (gvom.py:1058)
                } else {
                    p += (uint32_t)vl[j];
                }
            }
            const int32_t o = m ? (int32_t)(GVOM_OCC | m) : (int32_t)p;
            if (zin && c0 + j < X) out[j * Z + z] = o;
            const uint64_t ob = __ballot(m != 0);
            count[j] += __popcll(ob);
            if (lane == 0 && c0 + j < X) occ_bits[(int64_t)(c0 + j) * NW + (z0 >> 6)] = ob;
            // lowest free voxel with at least one pass (gvom.py:580-590)
            const uint64_t fb = __ballot(zin && m == 0 && p >= 1u);
            if (first_free[j] < 0 && fb) first_free[j] = z0 + __builtin_ctzll(fb);
        }
    }
    if (lane < CPW && c0 + lane < X) {
        uint32_t cnt = count[0];
        int ff = first_free[0];
#pragma unroll
        for (int j = 1; j < CPW; ++j)
            if (lane == j) {
                cnt = count[j];
                ff = first_free[j];
            }
        const int32_t c = c0 + lane;
        col_count[c] = (int32_t)cnt;
        inferred[c] = ff >= 0 ? ((double)ff + co[2]) * g.z_res : -1000.0;
        if (cnt == 0) {
            // no occupied voxel: the robot-radius default around the latest ego, measured to
            // the cell corner (gvom.py:561-570), else unknown (-1000); occupied columns get
            // their height from the record merge
            const int32_t x = c / XY, y = c - (c / XY) * XY;
            const double *ego = st->latest_ego;
            double h = -1000.0;
            const double xp = ((co[0] + (double)x) * g.xy_res) - ego[0];
            const double yp = ((co[1] + (double)y) * g.xy_res) - ego[1];
            if (xp * xp + yp * yp <= g.robot_r * g.robot_r) h = ego[2] - g.g2l;
            height[c] = h;
        }
    }
}

// Same combine for xy_size % 8 == 0, with less scalar work per column.  The RCPW columns of
// a group lie in one x row; each source's shift (precomputed by the scan stage's commit) gives
// one uniform column pointer, and a word that is off the map (column or z) is read from a
// zero page instead of being masked after the load, so no per-word predicate outlives its
// address.  A source's shift and column base are scalar work done once per group, so a wider
// group spends less of it per column: on 256 x 256 x 64 with 8 slots, 4 columns per wavefront
// took 29.4 us, 8 columns with 4 sources loaded together 28.6 us (8 sources together: 129
// VGPRs, 3 wavefronts per SIMD; 16 columns spill ~190 SGPRs).  Each wavefront handles GPW
// groups.  (The general kernel above issues ~1200 scalar instructions per wavefront for 4
// columns, which made it scalar-issue bound.)
constexpr int RCPW = 8;
constexpr int RSRC_GROUP = 4;
constexpr int RCPW_MAX = 8;
constexpr int GPW = 1;   // measured with 4 columns: 1 group per wavefront 34.4 us, 2 groups 35.3 us, 4 groups 36.8 us
constexpr int ZERO_PAGE_WORDS = RCPW_MAX * 256 + 64;   // covers lane + j * z_size for z_size <= 256

// combine rule of one voxel: union of the sources' occupancy and sum of their passes, then
// the previous map (gvom.py:1010-1063)
__device__ __forceinline__ int32_t comb_word(uint32_t m, uint32_t p, bool okl, int32_t vl, bool &occ, bool &freev) {
    if (okl) {
        if (m) {
            if (vl < 0) m |= 1u << LAST_BIT;
        } else if (vl < 0) {
            if (p <= 10u) m = 1u << LAST_BIT;   // carry-over (gvom.py:1058)
        } else {
            p += (uint32_t)vl;
        }
    }
    occ = m != 0;
    freev = m == 0 && p >= 1u;
    return m ? (int32_t)(GVOM_OCC | m) : (int32_t)p;
}

__global__ __launch_bounds__(256) void k_comb_rows(Geo g, const DevState *__restrict__ st,
                                                   const int32_t *__restrict__ slot_tables,
                                                   const int32_t *__restrict__ zero_page,
                                                   int32_t *__restrict__ comb_tables,
                                                   int32_t *__restrict__ col_count,
                                                   uint64_t *__restrict__ occ_bits,
                                                   double *__restrict__ height,
                                                   double *__restrict__ inferred) {
    const int lane = threadIdx.x & 63;
    const int32_t XY = g.XY, Z = g.Z, X = (int32_t)g.X, V = (int32_t)g.V;
    const int32_t cw = uni((int32_t)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * (RCPW * GPW));
    if (cw >= X) return;
    const int parity = uni(st->parity), cur = 1 - parity;
    const int have_last = uni(st->have_last);
    const int last_slot = uni(st->last_slot), flip = uni(st->flip);
    const double *co = st->slot_origin[last_slot];   // combined origin = newest slot's (gvom.py:229)
    const int32_t ldx = have_last ? uni((int32_t)(co[0] - st->comb_origin[parity][0])) : 0;
    const int32_t ldy = have_last ? uni((int32_t)(co[1] - st->comb_origin[parity][1])) : 0;
    const int32_t ldz = have_last ? uni((int32_t)(co[2] - st->comb_origin[parity][2])) : 0;
    const int32_t *ltab = comb_tables + (int64_t)parity * V;
    const int32_t *zp = zero_page + lane;
    const int NW = (Z + 63) >> 6;
    uint32_t count[GPW][RCPW], ffree[GPW][RCPW];
#pragma unroll
    for (int q = 0; q < GPW; ++q)
#pragma unroll
        for (int j = 0; j < RCPW; ++j) {
            count[q][j] = 0;
            ffree[q][j] = 0xFFFFFFFFu;
        }
    for (int z0 = 0; z0 < Z; z0 += 64) {
        const int z = z0 + lane;
        const bool zin = z < Z;
        uint32_t mask[GPW][RCPW], P[GPW][RCPW];
#pragma unroll
        for (int q = 0; q < GPW; ++q)
#pragma unroll
            for (int j = 0; j < RCPW; ++j) mask[q][j] = P[q][j] = 0;
        for (int s0 = 0; s0 < g.B; s0 += RSRC_GROUP) {
#pragma unroll
            for (int q = 0; q < GPW; ++q) {
                const int32_t c0 = cw + q * RCPW;
                const int32_t x0 = c0 / XY, y0 = c0 - (c0 / XY) * XY;
                int32_t v[RSRC_GROUP][RCPW];
#pragma unroll
                for (int k = 0; k < RSRC_GROUP; ++k) {
                    // shifts precomputed by the scan stage's commit (one 16-byte scalar load)
                    const int s = s0 + k < g.B ? s0 + k : 0;
                    const int4 sh = *reinterpret_cast<const int4 *>(st->comb_shift[s]);
                    const bool use = (s0 + k < g.B) & (sh.w != 0) & (c0 < X);
                    const int32_t ux = x0 + sh.x, uy0 = y0 + sh.y;
                    const bool rowok = use & ((uint32_t)ux < (uint32_t)XY);   // uniform
                    const int zs = z + sh.z;
                    const bool zok = zin & ((uint32_t)zs < (uint32_t)Z);
                    // column j of the source: cp + j * Z (wrapping arithmetic; dereferenced
                    // only where the column is on the map)
                    const int32_t *cp = slot_tables + (int64_t)cur_buf(flip, s) * V +
                                        (int32_t)(((uint32_t)ux * (uint32_t)XY + (uint32_t)uy0) * (uint32_t)Z);
#pragma unroll
                    for (int j = 0; j < RCPW; ++j) {
                        const bool cok = rowok & ((uint32_t)(uy0 + j) < (uint32_t)XY);   // uniform
                        const int32_t *a = (cok & zok) ? cp + j * Z + zs : zp + j * Z;
                        v[k][j] = *a;
                    }
                }
#pragma unroll
                for (int k = 0; k < RSRC_GROUP; ++k)
#pragma unroll
                    for (int j = 0; j < RCPW; ++j) {
                        if (v[k][j] < 0) mask[q][j] |= 1u << (s0 + k);
                        else P[q][j] += (uint32_t)v[k][j];
                    }
            }
        }
        const int zl = z + ldz;
        const bool zokl = (have_last != 0) & zin & ((uint32_t)zl < (uint32_t)Z);
#pragma unroll
        for (int q = 0; q < GPW; ++q) {
            const int32_t c0 = cw + q * RCPW;
            if (c0 >= X) break;
            const int32_t x0 = c0 / XY, y0 = c0 - (c0 / XY) * XY;
            const int32_t ux = x0 + ldx, uy0 = y0 + ldy;
            const bool rowok = (uint32_t)ux < (uint32_t)XY;
            const int32_t *cp = ltab + (int32_t)(((uint32_t)ux * (uint32_t)XY + (uint32_t)uy0) * (uint32_t)Z);
            int32_t vl[RCPW];
            bool okl[RCPW];
#pragma unroll
            for (int j = 0; j < RCPW; ++j) {
                okl[j] = rowok & zokl & ((uint32_t)(uy0 + j) < (uint32_t)XY);
                vl[j] = *(okl[j] ? cp + j * Z + zl : zp + j * Z);
            }
            int32_t *out = comb_tables + (int64_t)cur * V + (int64_t)c0 * Z;
#pragma unroll
            for (int j = 0; j < RCPW; ++j) {
                bool occ, freev;
                const int32_t o = comb_word(mask[q][j], P[q][j], okl[j], vl[j], occ, freev);
                if (zin) out[j * Z + z] = o;
                const uint64_t ob = __ballot(occ);
                count[q][j] += __popcll(ob);
                if (lane == 0) occ_bits[(int64_t)(c0 + j) * NW + (z0 >> 6)] = ob;
                // lowest free voxel with at least one pass (gvom.py:580-590)
                const uint64_t fb = __ballot(zin && freev);
                if (ffree[q][j] == 0xFFFFFFFFu && fb) ffree[q][j] = (uint32_t)(z0 + __builtin_ctzll(fb));
            }
        }
    }
    if (lane < GPW * RCPW) {
        const int32_t c = cw + lane;
        if (c < X) {
            uint32_t cnt = 0, ff = 0xFFFFFFFFu;
#pragma unroll
            for (int q = 0; q < GPW; ++q)
#pragma unroll
                for (int j = 0; j < RCPW; ++j)
                    if (lane == q * RCPW + j) {
                        cnt = count[q][j];
                        ff = ffree[q][j];
                    }
            col_count[c] = (int32_t)cnt;
            inferred[c] = ff != 0xFFFFFFFFu ? ((double)ff + co[2]) * g.z_res : -1000.0;
            if (cnt == 0) {
                // no occupied voxel: the robot-radius default around the latest ego, measured to
                // the cell corner (gvom.py:561-570), else unknown (-1000); occupied columns get
                // their height from the record merge
                const int32_t x = c / XY, y = c - (c / XY) * XY;
                const double *ego = st->latest_ego;
                double h = -1000.0;
                const double xp = ((co[0] + (double)x) * g.xy_res) - ego[0];
                const double yp = ((co[1] + (double)y) * g.xy_res) - ego[1];
                if (xp * xp + yp * yp <= g.robot_r * g.robot_r) h = ego[2] - g.g2l;
                height[c] = h;
            }
        }
    }
}

// ------------------------------------------------------------------------------------
// Column id scan: blocks of 1024 threads scan COLSCAN consecutive column counts in place
// (exclusive, int4 loads) and write their total; k_comb_records adds the totals of the
// earlier chunks.
constexpr int COLSCAN = gvom_plan::COLSCAN;   // 1024 threads x 4 columns
static_assert(COLSCAN == 1024 * 4, "column scan: 1024 threads x 4 columns per block");

__global__ __launch_bounds__(1024) void k_col_scan(int32_t *__restrict__ cnt, int64_t X,
                                                   int32_t *__restrict__ chunk_tot, DevState *__restrict__ st) {
    __shared__ int32_t wsum[16];
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    if (blockIdx.x == 0 && t == 0) {
        // origin of the combined map being built = newest slot's origin (gvom.py:229), and
        // the ping-pong flip (gvom.py:276-316): no other thread of this kernel reads the
        // parity, the combine's earlier kernel (the table pass) has finished, and its later
        // kernels read the new value
        const int cur = 1 - st->parity;
        const double *co = st->slot_origin[st->last_slot];
        st->comb_origin[cur][0] = co[0];
        st->comb_origin[cur][1] = co[1];
        st->comb_origin[cur][2] = co[2];
        st->parity = cur;
        st->have_last = 1;
    }
    const int64_t i = (int64_t)blockIdx.x * COLSCAN + 4 * t;
    // four counts per thread; past the last column they read as 0 and are not written
    int4 v = make_int4(0, 0, 0, 0);
    if (i + 3 < X) {
        v = *reinterpret_cast<const int4 *>(cnt + i);
    } else if (i < X) {
        v.x = cnt[i];
        if (i + 1 < X) v.y = cnt[i + 1];
        if (i + 2 < X) v.z = cnt[i + 2];
    }
    const int32_t sum = v.x + v.y + v.z + v.w;
    int32_t incl = sum;
    for (int off = 1; off < 64; off <<= 1) {
        int32_t u = __shfl_up(incl, off);
        if (lane >= off) incl += u;
    }
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int32_t pre = incl - sum;
    for (int ww = 0; ww < w; ++ww) pre += wsum[ww];
    int4 o;
    o.x = pre;
    o.y = pre + v.x;
    o.z = o.y + v.y;
    o.w = o.z + v.z;
    if (i + 3 < X) {
        *reinterpret_cast<int4 *>(cnt + i) = o;
    } else if (i < X) {
        cnt[i] = o.x;
        if (i + 1 < X) cnt[i + 1] = o.y;
        if (i + 2 < X) cnt[i + 2] = o.z;
    }
    if (t == 1023) chunk_tot[blockIdx.x] = pre + sum;
}

// ------------------------------------------------------------------------------------
// Record merge, __combine_metrics (gvom.py:889-980) with Numba's typing: the product of
// two combined (f32) entries is f32, everything touching an f64 operand is f64.
template <typename OT>
__device__ __forceinline__ void merge_record(float cm[10], int32_t &hit, int32_t &tot, float &mh, const OT om[10],
                                             int32_t ohit, int32_t otot, float omh) {
    using MT = decltype((float)0 * (OT)0);
    MT mx = (cm[0] * cm[9] + om[0] * om[9]) / (cm[9] + om[9]);
    MT my = (cm[1] * cm[9] + om[1] * om[9]) / (cm[9] + om[9]);
    MT mz = (cm[2] * cm[9] + om[2] * om[9]) / (cm[9] + om[9]);
    float c3 = (float)((cm[9] * cm[3] + om[9] * om[3] + cm[9] * (cm[0] - mx) * (cm[0] - mx) +
                        om[9] * (om[0] - mx) * (om[0] - mx)) / (cm[9] + om[9]));
    float c4 = (float)((cm[9] * cm[4] + om[9] * om[4] + cm[9] * (cm[0] - mx) * (cm[1] - my) +
                        om[9] * (om[0] - mx) * (om[1] - my)) / (cm[9] + om[9]));
    float c5 = (float)((cm[9] * cm[5] + om[9] * om[5] + cm[9] * (cm[0] - mx) * (cm[2] - mz) +
                        om[9] * (om[0] - mx) * (om[2] - mz)) / (cm[9] + om[9]));
    float c6 = (float)((cm[9] * cm[6] + om[9] * om[6] + cm[9] * (cm[1] - my) * (cm[1] - my) +
                        om[9] * (om[1] - my) * (om[1] - my)) / (cm[9] + om[9]));
    float c7 = (float)((cm[9] * cm[7] + om[9] * om[7] + cm[9] * (cm[1] - my) * (cm[2] - mz) +
                        om[9] * (om[1] - my) * (om[2] - mz)) / (cm[9] + om[9]));
    float c8 = (float)((cm[9] * cm[8] + om[9] * om[8] + cm[9] * (cm[2] - mz) * (cm[2] - mz) +
                        om[9] * (om[2] - mz) * (om[2] - mz)) / (cm[9] + om[9]));
    cm[3] = c3; cm[4] = c4; cm[5] = c5; cm[6] = c6; cm[7] = c7; cm[8] = c8;
    cm[0] = (float)mx; cm[1] = (float)my; cm[2] = (float)mz;
    cm[9] = (float)(cm[9] + om[9]);
    hit = (int32_t)((uint32_t)hit + (uint32_t)ohit);
    tot = (int32_t)((uint32_t)tot + (uint32_t)otot);
    mh = py_minf(mh, omh);
}

// __calculate_eigenvalues (gvom.py:1424-1487): f32 loads, f64 arithmetic, f32 stores,
// descending order; the middle one is 3q minus the two stored extremes.
__device__ __forceinline__ void eigen3(const float m[10], float ev[3]) {
    const double PI = 3.141592653589793;
    float xx = m[3], xy = m[4], xz = m[5], yy = m[6], yz = m[7], zz = m[8];
    float p1 = xy * xy + xz * xz + yz * yz;
    double q = (double)(xx + yy + zz) / 3.0;
    if (p1 == 0) {
        ev[0] = py_maxf(xx, py_maxf(yy, zz));
        ev[2] = py_minf(xx, py_minf(yy, zz));
        ev[1] = (float)(3.0 * q - (double)ev[0] - (double)ev[2]);
    } else {
        double p2 = ((double)xx - q) * ((double)xx - q) + ((double)yy - q) * ((double)yy - q) +
                    ((double)zz - q) * ((double)zz - q) + 2.0 * (double)p1;
        double p = sqrt(p2 / 6.0);
        double B0 = ((double)xx - q) / p, B1 = (double)xy / p, B2 = (double)xz / p;
        double B3 = ((double)yy - q) / p, B4 = (double)yz / p, B5 = ((double)zz - q) / p;
        double r = B0 * (B3 * B5 - B4 * B4) - B1 * (B1 * B5 - B4 * B2) + B2 * (B1 * B4 - B3 * B2);
        r = r / 2.0;
        double phi;
        if (r <= -1) phi = PI / 3.0;
        else if (r >= 1) phi = 0.0;
        else phi = acos(r) / 3.0;
        ev[0] = (float)(q + 2.0 * p * cos(phi));
        ev[2] = (float)(q + 2.0 * p * cos(phi + (2.0 * PI / 3.0)));
        ev[1] = (float)(3.0 * q - (double)ev[0] - (double)ev[2]);
    }
}

// Cell ids: one wavefront per column assigns ids in z order (first id of the column =
// earlier chunks' totals + in-chunk exclusive prefix), rewrites the table entry to
// OCC | id and appends (voxel, source mask) to the compact list.  It also writes the
// inferred height (lowest free voxel, gvom.py:580-590) and, for columns without an
// occupied voxel, the height-map default (robot-radius cell corner, gvom.py:561-570).
// Bit 31 of the list's voxel word marks the lowest occupied voxel of a column.
__global__ __launch_bounds__(256) void k_comb_ids(Geo g, const DevState *__restrict__ st,
                                                  int32_t *__restrict__ comb_tables,
                                                  const int32_t *__restrict__ col_base,
                                                  const int32_t *__restrict__ chunk_tot,
                                                  const uint64_t *__restrict__ occ_bits,
                                                  uint2 *__restrict__ occ_list) {
    // One lane per column, 64 columns per wavefront.  Only the occupied voxels are touched:
    // their z positions come from the occupancy words k_comb_tables wrote.
    const int lane = threadIdx.x & 63;
    const int32_t X = (int32_t)g.X, Z = g.Z;
    const int32_t cb = uni((int32_t)(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64);
    if (cb >= X) return;
    const int cur = st->parity;   // flipped by the column scan
    const int NW = (Z + 63) >> 6;
    int32_t *ctab = comb_tables + (int64_t)cur * g.V;
    // earlier chunks' totals (a wavefront's 64 columns lie in one chunk: 64 divides COLSCAN)
    uint32_t chunk_pre = 0;
    {
        const int32_t q = cb / COLSCAN;
        for (int32_t k = lane; k < q; k += 64) chunk_pre += (uint32_t)chunk_tot[k];
        for (int off = 32; off > 0; off >>= 1) chunk_pre += __shfl_xor(chunk_pre, off);
    }
    const int32_t c = cb + lane;
    if (c >= X) return;
    uint32_t id = chunk_pre + (uint32_t)col_base[c];
    bool lowest = true;
    for (int w = 0; w < NW; ++w) {
        uint64_t ob = occ_bits[(int64_t)c * NW + w];
        while (ob) {
            const int z = w * 64 + __builtin_ctzll(ob);
            ob &= ob - 1;
            const int32_t lin = c * Z + z;
            const int32_t v = ctab[lin];
            ctab[lin] = (int32_t)(GVOM_OCC | id);
            occ_list[id] = make_uint2((uint32_t)lin | (lowest ? 0x80000000u : 0u), (uint32_t)v & MASK31);
            lowest = false;
            ++id;
        }
    }
}

// Record merge per occupied voxel of the compact list (grid-stride).  Sources are taken
// in slot order, then the previous combined map (gvom.py:280-298).  Per group of
// MERGE_GROUP slots, all source-table loads are issued, then all record loads, then the
// merges run in order, so a thread has a group's gathers in flight together.  The
// thread holding a column's lowest occupied voxel writes its height
// (min_h + z + oz) * z_res (gvom.py:572-576).
constexpr int MERGE_GROUP = 4;

__global__ __launch_bounds__(256) void k_comb_merge(Geo g, DevState *__restrict__ st,
                                                    const int32_t *__restrict__ slot_tables,
                                                    const SlotRec *__restrict__ slot_recs, int64_t ncap,
                                                    const int32_t *__restrict__ comb_tables,
                                                    CombRec *__restrict__ comb_recs, int64_t ccap,
                                                    const uint2 *__restrict__ occ_list,
                                                    const int32_t *__restrict__ chunk_tot, int64_t nchunk,
                                                    double *__restrict__ height) {
    const int cur = st->parity, parity = 1 - cur, flip = st->flip;   // flipped by the column scan
    const double *co = st->comb_origin[cur];
    const int lane = threadIdx.x & 63;
    uint32_t C = 0;
    for (int64_t k = lane; k < nchunk; k += 64) C += (uint32_t)chunk_tot[k];
    for (int off = 32; off > 0; off >>= 1) C += __shfl_xor(C, off);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        // the new map's cell count (the commit of gvom.py:276-316)
        st->comb_cells[cur] = (int32_t)C;
        st->comb_total = (int32_t)C;
    }
    if ((int64_t)C > ccap) C = (uint32_t)ccap;   // the compact list and records hold ccap cells
    CombRec *crec = comb_recs + (int64_t)cur * ccap;
    const CombRec *lrec = comb_recs + (int64_t)parity * ccap;
    const int32_t *ltab = comb_tables + (int64_t)parity * g.V;
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    for (int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; id < (int64_t)C; id += stride) {
        const uint2 e = occ_list[id];
        const uint32_t lin = e.x & MASK31, mask = e.y;
        const bool lowest = (e.x >> 31) != 0;
        const int64_t z = lin % g.Z, c = lin / g.Z, x = c / g.XY, y = c % g.XY;
        float cm[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        int32_t hit = 0, tot = 0;
        float mh = 1.0f;
        for (int s0 = 0; s0 < g.B; s0 += MERGE_GROUP) {
            uint32_t sid[MERGE_GROUP];
            bool use[MERGE_GROUP];
#pragma unroll
            for (int k = 0; k < MERGE_GROUP; ++k) {
                const int s = s0 + k;
                use[k] = s < g.B && ((mask >> s) & 1u);
                const int32_t *addr = slot_tables;   // any valid word when the source is unused
                if (use[k]) {
                    SrcCol sc = source_column(slot_tables + (int64_t)cur_buf(flip, s) * g.V, co, st->slot_origin[s],
                                              x, y, g);
                    addr = sc.col + z + sc.dz;
                }
                sid[k] = (uint32_t)*addr & MASK31;
            }
            SlotRec r[MERGE_GROUP];
#pragma unroll
            for (int k = 0; k < MERGE_GROUP; ++k)
                r[k] = slot_recs[(int64_t)cur_buf(flip, use[k] ? s0 + k : 0) * ncap + (use[k] ? sid[k] : 0)];
#pragma unroll
            for (int k = 0; k < MERGE_GROUP; ++k)
                if (use[k]) merge_record<double>(cm, hit, tot, mh, r[k].m, r[k].hit, r[k].total, r[k].min_h);
        }
        if (mask & (1u << LAST_BIT)) {
            SrcCol sc = source_column(ltab, co, st->comb_origin[parity], x, y, g);
            const uint32_t lid = (uint32_t)sc.col[z + sc.dz] & MASK31;
            const CombRec &r = lrec[lid];
            merge_record<float>(cm, hit, tot, mh, r.m, r.hit, r.total, r.min_h);
        }
        CombRec out;
        out.hit = hit;
        out.total = tot;
        out.min_h = mh;
#pragma unroll
        for (int j = 0; j < 10; ++j) out.m[j] = cm[j];
        eigen3(cm, out.ev);
        crec[id] = out;
        if (lowest) height[c] = (((double)mh + (double)z) + co[2]) * g.z_res;
    }
}

// ------------------------------------------------------------------------------------
constexpr int TILE = 16;
constexpr int HALO = 15;
constexpr int LT = TILE + 2 * HALO;   // 46 <= 64: one 64-bit mask per tile row / column
constexpr int TT = TILE * TILE;       // cells per tile
constexpr int MT = 3 * TT;            // threads per block: three per cell
static_assert(MT % 64 == 0 && MT <= 1024 && LT <= 64, "whole wavefronts, one 64-bit mask per tile row");

// First set bit of m within [lo, hi) (bit positions), or -1.
__device__ __forceinline__ int first_in(uint64_t m, int lo, int hi) {
    if (hi <= lo) return -1;
    uint64_t r = (m >> lo);
    if (hi - lo < 64) r &= (1ull << (hi - lo)) - 1ull;
    return r ? lo + __builtin_ctzll(r) : -1;
}

// One block of MT = 3 TILE^2 threads per TILE x TILE tile: the first TILE^2 threads run the
// slope fit of their column, the second walk the same columns' positive-obstacle windows
// (dependent table and record loads), the third run the guessed-height ring search, so the
// three latency chains overlap; the first group then writes the maps.  (With the slope fit
// and the ring search on the same threads the kernel took 21.0 us on 256 x 256 cells, event
// timed; skipping either shortened it by 3.7 and 6.6 us.)
__global__ __launch_bounds__(MT) void k_maps2d(Geo g, const DevState *__restrict__ st,
                                                const double *__restrict__ height,
                                                const double *__restrict__ inferred,
                                                const int32_t *__restrict__ comb_tables,
                                                const CombRec *__restrict__ comb_recs, int64_t ccap,
                                                double *__restrict__ slope_x, double *__restrict__ slope_y,
                                                double *__restrict__ rough, double *__restrict__ delta,
                                                int32_t *__restrict__ pos, int32_t *__restrict__ neg,
                                                int32_t *__restrict__ vis, uint8_t *__restrict__ host_maps) {
    __shared__ double H[LT][LT];
    __shared__ uint64_t rowm[LT];   // bit ly set when H[lx][ly] is observed (rows: fixed x)
    __shared__ uint64_t colm[LT];   // bit lx set when H[lx][ly] is observed (columns: fixed y)
    __shared__ int32_t winpv[TT];   // window density result per column of the tile
    __shared__ double dhw[TT];      // guessed-height difference per column of the tile
    const int XY = g.XY;
    const int bx = blockIdx.y * TILE, by = blockIdx.x * TILE;   // tile origin (x rows, y fastest)
    for (int k = threadIdx.x; k < LT * LT; k += MT) {
        int lx = k / LT, ly = k % LT;
        int gx = bx - HALO + lx, gy = by - HALO + ly;
        H[lx][ly] = (gx >= 0 && gx < XY && gy >= 0 && gy < XY) ? height[(int64_t)gx * XY + gy] : -1000.0;
    }
    __syncthreads();
    {
        // observed masks by ballot: one row and one column per wavefront and round
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        for (int r = wv; r < LT; r += MT / 64) {
            bool ro = false, co = false;
            if (lane < LT) {
                ro = H[r][lane] > -1000;
                co = H[lane][r] > -1000;
            }
            const uint64_t mr = __ballot(ro), mc = __ballot(co);
            if (lane == 0) {
                rowm[r] = mr;
                colm[r] = mc;
            }
        }
    }
    __syncthreads();
    const int cell = threadIdx.x % TT, half = threadIdx.x / TT;
    const int tx = cell / TILE, ty = cell % TILE;
    const int x0 = bx + tx, y0 = by + ty;
    const bool act = x0 < XY && y0 < XY;
    const int64_t c0 = (int64_t)x0 * XY + y0;
#define HAT(gx, gy) H[(gx) - bx + HALO][(gy) - by + HALO]
    if (half == 1) {
        // __make_positive_obstacle_map's density (gvom.py:530-555): occupied voxels with
        // hit > 10 in [floor((h + pos)/zr - oz) + 1, floor((h + robot_h)/zr - oz)]; the slope
        // test that overrides it is applied by the first half
        int32_t pvw = 0;
        if (act) {
            const double h0 = HAT(x0, y0);
            const int cur = st->parity;   // flipped by the column scan
            const double oz = st->comb_origin[cur][2];
            const double flo = floor(((h0 + g.pos_thr) / g.z_res) - oz);
            const double fhi = floor(((h0 + g.robot_h) / g.z_res) - oz);
            if (flo + 1 >= 0 && flo + 1 < g.Z && fhi >= 0 && fhi < g.Z) {
                const int lo = (int)flo + 1, hi = (int)fhi;
                const int32_t *col = comb_tables + (int64_t)cur * g.V + c0 * g.Z;
                const CombRec *recs = comb_recs + (int64_t)cur * ccap;
                double dens = 0.0, nn = 0.0;
                // eight voxels per round: all table words, then all records, are in flight
                // together; the sums still run in z order
                for (int z0 = lo; z0 <= hi; z0 += 8) {
                    int32_t t[8], rh[8], rt[8];
#pragma unroll
                    for (int k = 0; k < 8; ++k) t[k] = (z0 + k <= hi) ? col[z0 + k] : 0;
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        rh[k] = -1;
                        rt[k] = 0;
                        if (t[k] < 0) {
                            const CombRec &r = recs[(uint32_t)t[k] & MASK31];
                            rh[k] = r.hit;
                            rt[k] = r.total;
                        }
                    }
#pragma unroll
                    for (int k = 0; k < 8; ++k)
                        if (rh[k] > 10) {
                            nn += (double)rt[k];
                            dens += (double)rh[k];
                        }
                }
                if (nn > 0.0) dens /= nn;
                pvw = (int32_t)(dens * 100.0);
            }
        }
        winpv[cell] = pvw;
        __syncthreads();
        return;
    }
    if (half == 2) {
        // __guess_height (gvom.py:593-714), its loop condition and the y_nh gating as written.
        // Each one-sided edge scan returns the first observed cell in scan order, i.e. the
        // lowest set bit of the row / column mask restricted to the scanned range (cells off
        // the map are never observed, matching the reference's bounds `continue`).
        double dh_out = 0.0;
        const double h0 = HAT(x0, y0);
        const double inf0 = act ? inferred[c0] : -1000.0;
        if (!(h0 > -1000) && !(inf0 == -1000.0)) {
            const int lx0 = x0 - bx + HALO, ly0 = y0 - by + HALO;
            bool xpd = false, xnd = false, ypd = false, ynd = false;
            double xph = -1000, xnh = -1000, yph = -1000, ynh = -1000;
            int i = 0;
            while (i < 15 && !(xnd && xnd && ypd && ynd)) {
                i += 1;
                if (!xpd) {
                    if (x0 + i < XY) {   // row x0+i, dy in [-i, i)
                        int f = first_in(rowm[lx0 + i], ly0 - i, ly0 + i);
                        if (f >= 0) { xph = H[lx0 + i][f]; xpd = true; }
                    } else xpd = true;
                }
                if (!xnd) {
                    if (x0 - i >= 0) {   // row x0-i, dy in (-i, i]
                        int f = first_in(rowm[lx0 - i], ly0 - i + 1, ly0 + i + 1);
                        if (f >= 0) { xnh = H[lx0 - i][f]; xnd = true; }
                    } else xnd = true;
                }
                if (!ypd) {
                    if (y0 + i < XY) {   // column y0+i, dx in (-i, i]
                        int f = first_in(colm[ly0 + i], lx0 - i + 1, lx0 + i + 1);
                        if (f >= 0) { yph = H[f][ly0 + i]; ypd = true; }
                    } else ypd = true;
                }
                if (!ynd) {
                    if (y0 - i >= 0) {   // column y0-i, dx in [-i, i)
                        int f = first_in(colm[ly0 - i], lx0 - i, lx0 + i);
                        if (f >= 0) { ynh = H[f][ly0 - i]; ynd = true; }
                    } else ynd = true;
                }
            }
            double mn = 1000.0, mxh = inf0;
            if (xph > -1000) { mn = py_min(xph, mn); mxh = py_max(xph, mxh); }
            if (xnh > -1000) { mn = py_min(xnh, mn); mxh = py_max(xnh, mxh); }
            if (yph > -1000) { mn = py_min(yph, mn); mxh = py_max(yph, mxh); }
            if (xnh > -1000) { mn = py_min(ynh, mn); mxh = py_max(ynh, mxh); }
            double dh = mxh - mn;
            if (dh > 0) dh_out = dh;
        }
        dhw[cell] = dh_out;
        if (act) delta[c0] = dh_out;
        __syncthreads();
        return;
    }
    // first group: lanes outside the map (partial tiles) compute on in-tile LDS only and
    // write nothing, so every lane reaches the one barrier below

    // __calculate_slope (gvom.py:718-805)
    double sx = 0.0, sy = 0.0, rg = -1.0;
    {
        const int xlo = x0 - 1 < 0 ? 0 : x0 - 1, xhi = x0 + 2 > XY ? XY : x0 + 2;
        const int ylo = y0 - 1 < 0 ? 0 : y0 - 1, yhi = y0 + 2 > XY ? XY : y0 + 2;
        int ngood = 0;
        for (int xx = xlo; xx < xhi; ++xx)
            for (int yy = ylo; yy < yhi; ++yy)
                if (HAT(xx, yy) > -1000) ++ngood;
        if (ngood >= 3) {
            // pts in window order (x outer, y inner) over observed cells, as gvom.py:736-749
            double mx = 0.0, my = 0.0, mz = 0.0;
            int i = 0;
            for (int xx = xlo; xx < xhi; ++xx)
                for (int yy = ylo; yy < yhi; ++yy) {
                    double hh = HAT(xx, yy);
                    if (hh > -1000) {
                        mx += (double)xx * g.xy_res;
                        my += (double)yy * g.xy_res;
                        mz += hh;
                        ++i;
                    }
                }
            mx /= (double)i;
            my /= (double)i;
            mz /= (double)i;
            double cxx = 0.0, cxy = 0.0, cxz = 0.0, cyy = 0.0, cyz = 0.0;
            for (int xx = xlo; xx < xhi; ++xx)
                for (int yy = ylo; yy < yhi; ++yy) {
                    double hh = HAT(xx, yy);
                    if (hh > -1000) {
                        double px = (double)xx * g.xy_res, py = (double)yy * g.xy_res;
                        cxx += (px - mx) * (px - mx);
                        cxy += (px - mx) * (py - my);
                        cxz += (px - mx) * (hh - mz);
                        cyy += (py - my) * (py - my);
                        cyz += (py - my) * (hh - mz);
                    }
                }
            double det = cxx * cyy - cxy * cxy;
            if (det != 0.0) {
                double a0 = (cyy * cxz - cxy * cyz) / det;
                double a1 = (cxx * cyz - cxy * cxz) / det;
                double m = sqrt(a0 * a0 + a1 * a1 + 1.0);
                a0 /= m;
                a1 /= m;
                double err = 0.0;
                for (int xx = xlo; xx < xhi; ++xx)
                    for (int yy = ylo; yy < yhi; ++yy) {
                        double hh = HAT(xx, yy);
                        if (hh > -1000) {
                            double px = (double)xx * g.xy_res, py = (double)yy * g.xy_res;
                            double e = (hh - mz) - (a0 * (px - mx) + a1 * (py - my));
                            err += e * e;
                        }
                    }
                err /= (double)ngood;
                if (err > 0) err = log(err);
                rg = err;
                sx = atan2(a0, 1.0 / m);
                sy = atan2(a1, 1.0 / m);
            }
        }
    }

#undef HAT
    if (act) {
        slope_x[c0] = sx;
        slope_y[c0] = sy;
        rough[c0] = rg;
    }
    // __make_positive_obstacle_map (gvom.py:516-555): slope test, else the window density
    __syncthreads();
    const double dh_out = dhw[cell];
    const double h0 = H[tx + HALO][ty + HALO];
    const int32_t pv = (sqrt(sx * sx + sy * sy) >= g.slope_thr) ? 100 : winpv[cell];
    // __make_negative_obstacle_map (506-512); __make_visability_map (445-452)
    if (act) {
        const int32_t nv = (dh_out > g.neg_thr) ? 100 : 0, vv = (h0 > -1000) ? 1 : 0;
        pos[c0] = pv;
        neg[c0] = nv;
        vis[c0] = vv;
        if (host_maps) {
            // the same four maps, [pos | neg | vis | rough], into pinned host memory
            const int64_t X = g.X;
            int32_t *hm = reinterpret_cast<int32_t *>(host_maps);
            hm[c0] = pv;
            hm[X + c0] = nv;
            hm[2 * X + c0] = vv;
            reinterpret_cast<double *>(hm + 3 * X)[c0] = rg;
        }
    }
}

// ------------------------------------------------------------------------------------
// Export: internal table -> reference layout (x fastest) and encoding.
__global__ __launch_bounds__(256) void k_export_table(Geo g, const int32_t *__restrict__ table,
                                                      int32_t *__restrict__ out) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.V) return;
    int64_t x = v % g.XY, y = (v / g.XY) % g.XY, z = v / g.X;
    int32_t t = table[(x * g.XY + y) * g.Z + z];
    out[v] = t < 0 ? (int32_t)((uint32_t)t & MASK31) : (int32_t)(-1 - (int64_t)t);
}

// __make_voxel_pointcloud (gvom.py:482-503), rows in cell-id order.
__global__ __launch_bounds__(256) void k_debug_voxels(Geo g, const DevState *__restrict__ st, int par,
                                                      const int32_t *__restrict__ table,
                                                      const CombRec *__restrict__ recs,
                                                      float *__restrict__ out) {
    int64_t v = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= g.V) return;
    int32_t t = table[v];
    if (t >= 0) return;
    uint32_t id = (uint32_t)t & MASK31;
    int64_t z = v % g.Z, c = v / g.Z, y = c % g.XY, x = c / g.XY;
    const double *o = st->comb_origin[par];
    const CombRec &r = recs[id];
    float *row = out + (int64_t)id * 8;
    row[0] = (float)(((double)x + o[0]) * g.xy_res);
    row[1] = (float)(((double)y + o[1]) * g.xy_res);
    row[2] = (float)(((double)z + o[2]) * g.z_res);
    row[3] = (float)((double)r.hit / (double)r.total);
    row[4] = (float)r.hit;
    row[5] = r.ev[0] - r.ev[1];
    row[6] = r.ev[1] - r.ev[2];
    row[7] = r.ev[2];
}

// __make_height_map_pointcloud (gvom.py:456-468) and __make_infered_height_map_pointcloud
// (471-479, fed with the guessed height delta, 439-440); row index x + y*xy.
__global__ __launch_bounds__(256) void k_debug_height(Geo g, const DevState *__restrict__ st, int par,
                                                      const double *__restrict__ height,
                                                      const double *__restrict__ rough,
                                                      const double *__restrict__ sx,
                                                      const double *__restrict__ sy,
                                                      const double *__restrict__ delta,
                                                      float *__restrict__ out7, float *__restrict__ out3) {
    int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= g.X) return;
    int64_t x = c / g.XY, y = c % g.XY;
    int64_t row = x + y * g.XY;
    const double *o = st->comb_origin[par];
    float fx = (float)(((double)x + o[0]) * g.xy_res);
    float fy = (float)(((double)y + o[1]) * g.xy_res);
    if (out7) {
        float *r = out7 + row * 7;
        r[0] = fx;
        r[1] = fy;
        r[2] = (float)(height[c] - g.z_res);
        r[3] = (float)rough[c];
        r[4] = (float)sx[c];
        r[5] = (float)sy[c];
        r[6] = (float)sqrt(sx[c] * sx[c] + sy[c] * sy[c]);
    }
    if (out3) {
        float *r = out3 + row * 3;
        r[0] = fx;
        r[1] = fy;
        r[2] = (float)(delta[c] - g.z_res);
    }
}

inline unsigned nblocks(int64_t n, int per) { return (unsigned)((n + per - 1) / per > 0 ? (n + per - 1) / per : 1); }

}  // namespace

int64_t zero_page_words() { return ZERO_PAGE_WORDS; }

hipError_t launch_comb_tables(const Geo &g, DevState *st, const int32_t *slot_tables, const int32_t *zero_page,
                              int32_t *comb_tables, int32_t *col_count, uint64_t *occ_bits, double *height,
                              double *inferred, hipStream_t s) {
    static_assert(RCPW <= RCPW_MAX, "the zero page covers RCPW_MAX columns");
    if (g.XY % RCPW == 0 && zero_page)
        k_comb_rows<<<nblocks((g.X + RCPW * GPW - 1) / (RCPW * GPW), 4), 256, 0, s>>>(g, st, slot_tables, zero_page,
                                                                                 comb_tables, col_count, occ_bits,
                                                                                 height, inferred);
    else
        k_comb_tables<<<nblocks((g.X + CPW - 1) / CPW, 4), 256, 0, s>>>(g, st, slot_tables, comb_tables, col_count,
                                                                        occ_bits, height, inferred);
    return hipGetLastError();
}

int64_t col_chunks(int64_t X) { return gvom_plan::col_chunks(X); }

hipError_t launch_scan_columns(int32_t *col_count, int64_t count_words, int64_t X, int32_t *chunk_tot,
                               int64_t chunk_words, DevState *st, hipStream_t s) {
    if (X < 1 || gvom_plan::col_scan_count_words(X) > count_words || gvom_plan::col_scan_chunk_words(X) > chunk_words)
        return hipErrorInvalidValue;
    k_col_scan<<<(unsigned)col_chunks(X), 1024, 0, s>>>(col_count, X, chunk_tot, st);
    return hipGetLastError();
}

hipError_t launch_comb_records(const Geo &g, DevState *st, const int32_t *slot_tables,
                               const SlotRec *slot_recs, int64_t ncap, int32_t *comb_tables,
                               CombRec *comb_recs, int64_t ccap, const int32_t *col_base,
                               const int32_t *chunk_tot, const uint64_t *occ_bits, uint2 *occ_list, double *height,
                               hipStream_t s) {
    k_comb_ids<<<nblocks((g.X + 63) / 64, 4), 256, 0, s>>>(g, st, comb_tables, col_base, chunk_tot, occ_bits,
                                                                 occ_list);
    k_comb_merge<<<2048, 256, 0, s>>>(g, st, slot_tables, slot_recs, ncap, comb_tables, comb_recs, ccap, occ_list,
                                      chunk_tot, col_chunks(g.X), height);
    return hipGetLastError();
}

hipError_t launch_maps2d(const Geo &g, const DevState *st, const double *height, const double *inferred,
                         const int32_t *comb_tables, const CombRec *comb_recs, int64_t ccap, double *slope_x,
                         double *slope_y, double *rough, double *delta, int32_t *pos, int32_t *neg, int32_t *vis,
                         uint8_t *host_maps, hipStream_t s) {
    dim3 grid((unsigned)((g.XY + TILE - 1) / TILE), (unsigned)((g.XY + TILE - 1) / TILE));
    k_maps2d<<<grid, MT, 0, s>>>(g, st, height, inferred, comb_tables, comb_recs, ccap, slope_x, slope_y, rough,
                                  delta, pos, neg, vis, host_maps);
    return hipGetLastError();
}


hipError_t launch_export_table(const Geo &g, const int32_t *table, int32_t *out_ref, hipStream_t s) {
    k_export_table<<<nblocks(g.V, 256), 256, 0, s>>>(g, table, out_ref);
    return hipGetLastError();
}

hipError_t launch_debug_voxels(const Geo &g, const DevState *st, const int32_t *comb_tables,
                               const CombRec *comb_recs, int64_t ccap, int parity, float *out, hipStream_t s) {
    k_debug_voxels<<<nblocks(g.V, 256), 256, 0, s>>>(g, st, parity, comb_tables + (int64_t)parity * g.V,
                                                     comb_recs + (int64_t)parity * ccap, out);
    return hipGetLastError();
}

hipError_t launch_debug_height(const Geo &g, const DevState *st, int parity, const double *height,
                               const double *rough, const double *sx, const double *sy, const double *delta,
                               float *out7, float *out3, hipStream_t s) {
    k_debug_height<<<nblocks(g.X, 256), 256, 0, s>>>(g, st, parity, height, rough, sx, sy, delta, out7, out3);
    return hipGetLastError();
}
