// C-ABI of libgvom_hip.so (declared in include/gvom_abi.h): handle lifetime, device
// memory, the ring-buffer bookkeeping of the reference class Gvom
// (scripts/gvom.py:21-101, 198-216, 225-308) and the launch sequences of the scan and
// combine stages.
#include <dlfcn.h>
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "gvom_kernels.h"

hipError_t launch_clear_slots(int32_t *slot_tables, const DevState *st, int S, const Geo &g, hipStream_t s);
int64_t diff_partials(int64_t n_points, int64_t X, int64_t ncap);
int64_t diff_rec_bytes();
hipError_t launch_diff_forward(const Geo &g, const DevState *st, int slot, const int32_t *table,
                               const SlotRec *recs, int64_t ncap, const int32_t *seg_start, const uint32_t *svals,
                               const double *world, int64_t n, void *dr, double *colh, int32_t *colpt,
                               int32_t *pmark, double *part, double *sums, double wh, hipStream_t s);
hipError_t launch_grid_payloads(const Geo &g, const int32_t *pos, const int32_t *neg, const double *rough,
                                const int32_t *vis, double density, double min_r, double max_r, int8_t *out,
                                hipStream_t s);
hipError_t launch_cloud_gather(const uint8_t *data, int64_t n, int64_t width, int64_t point_step, int64_t row_step,
                               const int32_t off[3], const int32_t type[3], int big, double *xyz, hipStream_t s);
hipError_t launch_probe_atomics(int32_t *table, uint32_t V, int blocks, int per_thread, int pattern, hipStream_t s);
hipError_t launch_diff_backward(const Geo &g, const DevState *st, int slot, const ScanDesc *sd,
                                const double *world, const uint8_t *valid, const int32_t *table, const void *dr,
                                const int32_t *pmark, const double *colh, const double *sums, double wh, double dL,
                                const double *dL_dev, int64_t n, double *gsensor, double *gworld, double *part,
                                double *grad6, hipStream_t s);

namespace {
thread_local std::string g_create_error;

int64_t bits_for(int64_t v) {  // smallest b with 2^b >= v
    int64_t b = 0;
    while (((int64_t)1 << b) < v) ++b;
    return b;
}
}  // namespace

// ---- device-static scratch of gvom_process_pointcloud_device_noalloc ----------------------
// The scan stage's point work buffers for scans of up to NOALLOC_POINTS points, carved from
// one file-scope device array.  Every GPU the library runs on holds its own copy until the
// process ends; all handles and calls on that GPU share it (calls that overlap on one GPU
// race).  Host code never uses the array's name as an address: each device's address comes
// from hipGetSymbolAddress, once per device, at handle creation.
constexpr int64_t NOALLOC_POINTS = 131072;

struct ScratchLayout {
    int64_t world, valid, k0, v0, k1, v1, hist, hwords, blk, blk_words, seg_key, seg_start, mom, part, rsec,
        rorder, total;
};

constexpr int64_t align256(int64_t b) { return (b + 255) / 256 * 256; }

constexpr ScratchLayout scratch_layout(int64_t n) {
    ScratchLayout L{};
    int64_t at = 0;
    L.world = at;     at = align256(at + n * 3 * (int64_t)sizeof(double));
    L.valid = at;     at = align256(at + n);
    L.k0 = at;        at = align256(at + n * 4);
    L.v0 = at;        at = align256(at + n * 4);
    L.k1 = at;        at = align256(at + n * 4);
    L.v1 = at;        at = align256(at + n * 4);
    L.hwords = gvom_plan::radix_hist_words(n);
    L.hist = at;      at = align256(at + gvom_plan::RADIX_HIST_SETS * L.hwords * 4);
    L.blk_words = gvom_plan::segment_block_words(n);
    L.blk = at;       at = align256(at + L.blk_words * 4);
    L.seg_key = at;   at = align256(at + n * 4);
    L.seg_start = at; at = align256(at + (n + 1) * 4);
    L.mom = at;       at = align256(at + n * (int64_t)sizeof(Moment));
    L.part = at;      at = align256(at + n * (int64_t)sizeof(Moment));
    L.rsec = at;      at = align256(at + n);
    L.rorder = at;    at = align256(at + n * 4);
    L.total = at;
    return L;
}

constexpr ScratchLayout NOALLOC_LAYOUT = scratch_layout(NOALLOC_POINTS);
static_assert(NOALLOC_LAYOUT.total < ((int64_t)32 << 20), "allocation-free scratch stays below 32 MiB");

static __device__ __align__(256) uint8_t scratch[NOALLOC_LAYOUT.total];

constexpr int MAX_DEVICES = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_dev[MAX_DEVICES] = {};

// The scratch array's address on the current device (looked up once per device).
static hipError_t device_scratch(int device, void **out) {
    if (device < 0 || device >= MAX_DEVICES) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_dev[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess) return e;
        g_scratch_dev[device] = p;
    }
    *out = g_scratch_dev[device];
    return hipSuccess;
}

struct gvom_handle {
    gvom_params p;
    Geo g;
    int device = 0;
    hipStream_t stream = nullptr;
    // second stream for the sort / segment / chunk-partial chain, which does not depend on
    // the ray cast; forked after k_prepare and joined before the segment sums
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // caller-stream hand-off of the stream-ordered entry points: ev_in on the caller's stream,
    // waited for by the handle's stream; ev_out the other way
    hipEvent_t ev_in = nullptr, ev_out = nullptr;
    std::mutex mu;
    std::string err;

    DevState *st = nullptr;
    int32_t *slot_tables = nullptr;
    SlotRec *slot_recs = nullptr;
    int64_t ncap = 0;
    int32_t *comb_tables = nullptr;
    CombRec *comb_recs = nullptr;
    int64_t ccap = 0;
    double *height = nullptr, *inferred = nullptr, *slope_x = nullptr, *slope_y = nullptr, *rough = nullptr,
           *delta = nullptr;
    int32_t *pos = nullptr, *neg = nullptr, *vis = nullptr, *col_count = nullptr;
    uint8_t *maps_block = nullptr;   // pos, neg, vis (int32) and rough (f64): 20 X bytes
    // Pinned host staging of the host-array entry points (hipHostMalloc, grown on demand):
    // two point buffers used in turn (ev_pin[k] marks the end of buffer k's last transfer)
    // and one for the four maps.
    void *pin_pts[2] = {nullptr, nullptr};
    size_t pin_pts_bytes[2] = {0, 0};
    hipEvent_t ev_pin[2] = {nullptr, nullptr};
    int pin_next = 0;
    void *pin_maps = nullptr, *pin_maps_dev = nullptr;   // host address and its device mapping
    uint64_t *occ_bits = nullptr;   // combined map: per column, bit z of word z / 64 = occupied
    int32_t *zero_page = nullptr;   // zeros the combine reads in place of off-map words
    uint2 *occ_list = nullptr;
    int32_t *export_buf = nullptr;
    int32_t *chunk_tot = nullptr;

    int64_t ptcap = 0;
    void *pts_in = nullptr;
    void *world = nullptr;
    uint8_t *valid = nullptr;
    uint32_t *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr, *hist = nullptr, *blk = nullptr,
             *seg_key = nullptr;
    int32_t *seg_start = nullptr;
    int64_t hwords = 0;   // words of each of the sort's two histograms
    Moment *mom = nullptr, *part = nullptr;
    ScanDesc *sd = nullptr;

    // host mirror of the ring state (gvom.py buffer_index / last_buffer_index / slots)
    int next_slot = 0, last_slot = 0, filled_mask = 0, parity = 0, have_last = 0;
    double slot_origin[GVOM_MAX_BUFFER][3] = {};
    double latest_ego[3] = {0, 0, 0};

    // Stage timing: one set of 10 events per process call (0..4, 9) and combine call (5..8);
    // gvom_stage_times averages every stage over all sets recorded since timing was enabled.
    // differentiable form: buffers (allocated on first use) and the forward's context
    void *diff_rec = nullptr;
    double *diff_colh = nullptr, *diff_part = nullptr, *diff_sums = nullptr;
    int32_t *diff_colpt = nullptr, *diff_pmark = nullptr;
    int64_t diff_rec_cap = 0, diff_pt_cap = 0, diff_part_cap = 0;
    const uint32_t *sorted_vals = nullptr;
    int64_t diff_n = -1;
    int diff_slot = 0;
    double diff_wh = 0.0;
    // the scan descriptor the last forward's kernels read: h->sd after an eager forward,
    // h->dg_sd after a replay of the captured forward (the backward reads its transform and ego)
    const ScanDesc *diff_sd = nullptr;

    // hipGraph capture: ops recorded while capturing (process = S scans, combine = -1)
    bool capturing = false;
    std::vector<int> cap_ops;
    std::vector<ScanDesc> cap_sd;   // descriptors of the captured scans, in capture order
    ScanDesc *graph_sd = nullptr;   // device pool the captured kernels read descriptors from
    int64_t graph_sd_cap = 0;
    int mirror_next = 0, mirror_last = 0, mirror_filled = 0, mirror_parity = 0, mirror_have = 0;
    // Bumped whenever a device buffer a captured sequence may reference is reallocated; a graph
    // captured under an older generation is refused instead of replayed against freed memory.
    int64_t buf_gen = 0;
    // Internal captures read their one descriptor from here instead of the graph pool.
    ScanDesc *cap_sd_dev = nullptr;
    // gvom_diff_forward's cached launch sequence (geometry-only, one-slot handles): captured on
    // the second call with the same points, count and weight, then replayed with a new
    // descriptor.  dk_*: the previous call's key.
    hipGraph_t dg_graph = nullptr;
    hipGraphExec_t dg_exec = nullptr;
    ScanDesc *dg_sd = nullptr;
    const void *dg_pts = nullptr, *dk_pts = nullptr;
    int64_t dg_n = -1, dk_n = -1, dg_gen = -1;
    double dg_wh = 0.0, dk_wh = 0.0;
    // A scan call that repeats the previous one's point count, dtype and source replays a
    // captured launch sequence: one launch instead of a dozen and no gaps between the kernels.
    // Host-array calls (gvom_process_pointcloud) need no transfer: the kernels read the points
    // and the scan's descriptor from the pinned staging buffer itself.  Device-pointer calls
    // without overlap (gvom_process_pointcloud_device) read the points from the caller's buffer
    // and the descriptor from the staging buffer.  One sequence per staging buffer.  A
    // sequence is replayed only for the source it was captured with (hp_src: the caller's
    // device buffer, or null for host arrays) and under the buffer generation, the staging
    // buffers (pin_gen, bumped when one is reallocated) and the launch settings (cfg_gen,
    // bumped by every setting that changes what a call launches) it was captured under.
    hipGraph_t hp_graph[2] = {nullptr, nullptr};
    hipGraphExec_t hp_exec[2] = {nullptr, nullptr};
    int64_t hp_n = -1, hp_gen = -1, hp_cfg = -1, hp_pin_gen = -1;
    int hp_dtype = -1;
    const void *hp_src = nullptr;
    int64_t pin_gen = 0;
    int64_t hk_n = -1;   // the previous eager scan call's point count, dtype and source
    int hk_dtype = -1;
    const void *hk_src = nullptr;
    int64_t cfg_gen = 0;

    bool timing = false;
    std::vector<hipEvent_t> ev_pool;
    std::vector<int> set_used;   // bit k set when event k of that set was recorded
    int cur_set = -1;
    int ray_block = 128, ray_rpt = 1, ray_box = 0, ray_xcd = 0, ray_spread = 1;
    int64_t ray_stride = 0;   // gvom_set_ray_rings: points per ring (0: linear ray order)
    int ray_patch = 0;        // rings per patch
    uint8_t *rsec = nullptr;
    uint32_t *rorder = nullptr;
    int8_t *grids = nullptr;   // OccupancyGrid payloads staged for host copies (5 * X bytes)
    uint8_t *raw = nullptr;    // PointCloud2 bytes staged from the host
    int64_t rawcap = 0;
    // tiled ray-cast scratch: min(B, 32) zeroed tables in the 4 x 4 x 1 segment layout
    int ray_tiled = 1;
    int32_t *tscratch = nullptr;
    int tscratch_scans = 0;
    int64_t tscratch_words = 0;
    // words of the buffers whose launches the host checks against their footprint
    int64_t slot_table_words = 0, col_count_words = 0, chunk_tot_words = 0;
    // Overlap of a device-pointer scan stage with the preceding combine (gvom_set_overlap):
    // the stage's first part (transform, keys, ray cast into the tiled scratch, sort chain)
    // runs on `front`, which waits only for ev_back; ev_back is recorded on the main stream
    // after every call that uses the scan stage's work buffers, but not after a combine.
    int overlap = 0;
    hipStream_t front = nullptr;
    hipEvent_t ev_front = nullptr, ev_back = nullptr;
    // An overlapped stage's commit waits only up to its segment sums (ev_front), which write
    // what the combine's table pass reads; its neighbourhood metrics (slot records' moments)
    // may still run on `front`, ending at ev_metrics.  metrics_pending: the main stream has
    // not waited for them yet; the combine waits just before its record merge, every other
    // call that touches the device at its start (join_metrics).
    hipEvent_t ev_metrics = nullptr;
    bool metrics_pending = false;
    // An overlapped stage inside a capture forks `front` from the main stream and `side` from
    // `front`.  hipStreamEndCapture has crashed on that nested fork under an older HIP runtime
    // than the one this library is built against (torch bundles one, and when torch is imported
    // first this library binds to it: both have the SONAME libamdhip64.so.7).  Under an older
    // runtime a captured sequence therefore runs its scan stages on the main stream, as with
    // overlap off; eager calls still overlap.
    int runtime_version = 0;
    bool nested_capture = true;
    int diff_geometry = 0;   // gvom_set_diff_geometry_only
    uint8_t *noalloc_scratch = nullptr;   // this device's copy of `scratch`
};

#define CKH(h, expr)                                                                 \
    do {                                                                             \
        hipError_t e_ = (expr);                                                      \
        if (e_ != hipSuccess) {                                                      \
            (h)->err = std::string(#expr) + ": " + hipGetErrorString(e_);            \
            return GVOM_EHIP;                                                        \
        }                                                                            \
    } while (0)

// Orders the main stream after the neighbourhood metrics an overlapped scan stage left
// running on `front`.
static hipError_t join_metrics(gvom_handle *h) {
    if (!h->metrics_pending) return hipSuccess;
    h->metrics_pending = false;
    return hipStreamWaitEvent(h->stream, h->ev_metrics, 0);
}

static void free_all(gvom_handle *h) {
    void *ptrs[] = {h->st, h->slot_tables, h->slot_recs, h->comb_tables, h->comb_recs, h->height, h->inferred,
                    h->slope_x, h->slope_y, h->maps_block, h->delta, h->occ_list,
                    h->col_count, h->chunk_tot, h->export_buf, h->pts_in, h->world, h->valid, h->k0, h->v0, h->k1, h->v1,
                    h->hist, h->blk, h->seg_key, h->seg_start, h->mom, h->part, h->sd, h->diff_rec,
                    h->diff_colh, h->diff_part, h->diff_sums, h->diff_colpt, h->diff_pmark, h->graph_sd, h->rsec, h->rorder, h->grids, h->raw, h->tscratch,
                    h->occ_bits, h->zero_page};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    if (h->dg_exec) (void)hipGraphExecDestroy(h->dg_exec);
    if (h->dg_graph) (void)hipGraphDestroy(h->dg_graph);
    if (h->dg_sd) (void)hipFree(h->dg_sd);
    for (int k = 0; k < 2; ++k) {
        if (h->hp_exec[k]) (void)hipGraphExecDestroy(h->hp_exec[k]);
        if (h->hp_graph[k]) (void)hipGraphDestroy(h->hp_graph[k]);
    }
    for (auto &e : h->ev_pool)
        if (e) (void)hipEventDestroy(e);
    if (h->ev_fork) (void)hipEventDestroy(h->ev_fork);
    if (h->ev_join) (void)hipEventDestroy(h->ev_join);
    if (h->ev_in) (void)hipEventDestroy(h->ev_in);
    for (int k = 0; k < 2; ++k) {
        if (h->ev_pin[k]) (void)hipEventDestroy(h->ev_pin[k]);
        if (h->pin_pts[k]) (void)hipHostFree(h->pin_pts[k]);
    }
    if (h->pin_maps) (void)hipHostFree(h->pin_maps);
    if (h->ev_out) (void)hipEventDestroy(h->ev_out);
    if (h->ev_front) (void)hipEventDestroy(h->ev_front);
    if (h->ev_back) (void)hipEventDestroy(h->ev_back);
    if (h->ev_metrics) (void)hipEventDestroy(h->ev_metrics);
    if (h->side) (void)hipStreamDestroy(h->side);
    if (h->front) (void)hipStreamDestroy(h->front);
    if (h->stream) (void)hipStreamDestroy(h->stream);
}

static int fail_create(gvom_handle *h, const std::string &msg, int code) {
    g_create_error = msg;
    if (h) {
        free_all(h);
        delete h;
    }
    return code;
}

template <typename T>
static hipError_t dalloc(T **p, int64_t count) {
    if (count <= 0) count = 1;
    return hipMalloc((void **)p, (size_t)count * sizeof(T));
}

extern "C" int gvom_create(const gvom_params *params, gvom_handle **out) {
    if (!params || !out) return fail_create(nullptr, "null argument", GVOM_EINVAL);
    *out = nullptr;
    const gvom_params &p = *params;
    if (!(p.xy_resolution > 0 && std::isfinite(p.xy_resolution) && p.z_resolution > 0 &&
          std::isfinite(p.z_resolution)))
        return fail_create(nullptr, "resolutions must be positive and finite", GVOM_EINVAL);
    if (p.xy_size < 4 || p.xy_size > 8192 || p.z_size < 4 || p.z_size > 256)
        return fail_create(nullptr, "xy_size must be in [4, 8192] and z_size in [4, 256]", GVOM_EINVAL);
    if (p.buffer_size < 1 || p.buffer_size > GVOM_MAX_BUFFER)
        return fail_create(nullptr, "buffer_size must be in [1, 30]", GVOM_EINVAL);
    if (p.xy_eigen_dist < 0 || p.xy_eigen_dist > 8 || p.z_eigen_dist < 0 || p.z_eigen_dist > 8)
        return fail_create(nullptr, "eigen distances must be in [0, 8]", GVOM_EINVAL);
    const int64_t XY = p.xy_size, Z = p.z_size, V = XY * XY * Z;
    if (!gvom_plan::params_in_domain(p.xy_size, p.z_size, p.buffer_size))
        return fail_create(nullptr, "map too large: voxels * (buffer_size + 1) must be < 2^31", GVOM_EINVAL);

    gvom_handle *h = new gvom_handle();
    h->p = p;
    Geo &g = h->g;
    g.xy_res = p.xy_resolution;
    g.z_res = p.z_resolution;
    g.min_d2 = p.min_distance * p.min_distance;
    g.XY = (int32_t)XY;
    g.Z = (int32_t)Z;
    g.V = V;
    g.X = XY * XY;
    g.B = p.buffer_size;
    g.ex = p.xy_eigen_dist;
    g.ez = p.z_eigen_dist;
    g.XYp = (int32_t)(XY + 2 * g.ex);
    g.Zp = (int32_t)(Z + 2 * g.ez);
    int64_t vpad = (int64_t)g.XYp * g.XYp * g.Zp;
    int64_t vmax = vpad > V ? vpad : V;
    g.VB = (int32_t)(bits_for(vmax + 1) + 1);
    if (g.VB > 27) return fail_create(h, "map too large for 32-bit voxel sort keys", GVOM_EINVAL);
    g.pos_thr = p.positive_obstacle_threshold;
    g.neg_thr = p.negative_obstacle_threshold;
    g.slope_thr = p.slope_obstacle_threshold;
    g.robot_h = p.robot_height;
    g.robot_r = p.robot_radius;
    g.g2l = p.ground_to_lidar_height;

    hipError_t e;
#define CKC(expr)                                                                              \
    do {                                                                                       \
        e = (expr);                                                                            \
        if (e != hipSuccess) return fail_create(h, std::string(#expr) + ": " + hipGetErrorString(e), GVOM_EHIP); \
    } while (0)
    CKC(hipGetDevice(&h->device));
    {
        void *sp = nullptr;
        CKC(device_scratch(h->device, &sp));   // the handle's device is current here
        h->noalloc_scratch = (uint8_t *)sp;
    }
    CKC(hipRuntimeGetVersion(&h->runtime_version));
    h->nested_capture = h->runtime_version / 100000 >= HIP_VERSION / 100000;   // major * 100 + minor
    CKC(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    CKC(hipStreamCreateWithFlags(&h->side, hipStreamNonBlocking));
    CKC(hipEventCreateWithFlags(&h->ev_fork, hipEventDisableTiming));
    CKC(hipEventCreateWithFlags(&h->ev_join, hipEventDisableTiming));
    CKC(hipEventCreateWithFlags(&h->ev_in, hipEventDisableTiming));
    CKC(hipEventCreateWithFlags(&h->ev_pin[0], hipEventDisableTiming));
    CKC(hipEventCreateWithFlags(&h->ev_pin[1], hipEventDisableTiming));
    CKC(hipEventCreateWithFlags(&h->ev_out, hipEventDisableTiming));
    CKC(hipStreamCreateWithFlags(&h->front, hipStreamNonBlocking));
    CKC(hipEventCreateWithFlags(&h->ev_front, hipEventDisableTiming));
    CKC(hipEventCreateWithFlags(&h->ev_back, hipEventDisableTiming));
    CKC(hipEventCreateWithFlags(&h->ev_metrics, hipEventDisableTiming));
    CKC(hipEventRecord(h->ev_back, h->stream));
    CKC(hipEventRecord(h->ev_pin[0], h->stream));
    CKC(hipEventRecord(h->ev_pin[1], h->stream));
    CKC(dalloc(&h->st, 1));
    CKC(hipMemset(h->st, 0, sizeof(DevState)));
    h->slot_table_words = gvom_plan::alloc_slot_table_words(V, g.B);   // two buffers per ring slot
    CKC(dalloc(&h->slot_tables, h->slot_table_words));
    h->ccap = V;
    CKC(dalloc(&h->comb_tables, 2 * V));
    CKC(dalloc(&h->comb_recs, 2 * h->ccap));
    const int64_t X = g.X;
    CKC(dalloc(&h->height, X));
    CKC(dalloc(&h->inferred, X));
    CKC(dalloc(&h->slope_x, X));
    CKC(dalloc(&h->slope_y, X));
    // the four output maps in one block [pos | neg | vis | rough], so a host combine copies
    // them with one transfer
    CKC(dalloc(&h->maps_block, 20 * X));
    h->pos = (int32_t *)h->maps_block;
    h->neg = h->pos + X;
    h->vis = h->neg + X;
    h->rough = (double *)(h->vis + X);
    CKC(dalloc(&h->delta, X));
    CKC(dalloc(&h->grids, 5 * X));
    CKC(dalloc(&h->occ_list, V));
    h->col_count_words = gvom_plan::alloc_col_count_words(X);
    CKC(dalloc(&h->col_count, h->col_count_words));
    CKC(dalloc(&h->occ_bits, X * ((g.Z + 63) / 64)));
    CKC(dalloc(&h->zero_page, zero_page_words()));
    CKC(hipMemset(h->zero_page, 0, sizeof(int32_t) * zero_page_words()));
    h->chunk_tot_words = gvom_plan::alloc_chunk_tot_words(X);
    CKC(dalloc(&h->chunk_tot, h->chunk_tot_words));
    CKC(dalloc(&h->sd, GVOM_MAX_BATCH));
    CKC(hipDeviceSynchronize());   // the fills above ran on the null stream
#undef CKC
    *out = h;
    return GVOM_OK;
}

extern "C" void gvom_destroy(gvom_handle *h) {
    if (!h) return;
    {
        std::lock_guard<std::mutex> lk(h->mu);
        if (h->stream) (void)hipStreamSynchronize(h->stream);
        if (h->front) (void)hipStreamSynchronize(h->front);
        if (h->side) (void)hipStreamSynchronize(h->side);
        free_all(h);
    }
    delete h;
}

extern "C" const char *gvom_last_error(const gvom_handle *h) {
    return h ? h->err.c_str() : g_create_error.c_str();
}

extern "C" void *gvom_stream(gvom_handle *h) {
    if (!h) return nullptr;
    // work the caller orders after the stream also follows pending metrics
    std::lock_guard<std::mutex> lk(h->mu);
    if (join_metrics(h) != hipSuccess) h->err = "hipStreamWaitEvent on the metrics event failed";
    return (void *)h->stream;
}

// Grow per-batch work buffers (points) and per-slot record capacity (cells per scan).
static int ensure_capacity(gvom_handle *h, int64_t max_scan, int64_t batch_pts) {
    const Geo &g = h->g;
    if (h->capturing) {
        int64_t need_ncap = max_scan < g.V ? max_scan : g.V;
        if (need_ncap > h->ncap || batch_pts > h->ptcap || h->ptcap < 1) {
            h->err = "buffers must be reserved (gvom_reserve) before graph capture";
            return GVOM_ESTATE;
        }
        return GVOM_OK;
    }
    int64_t need_ncap = max_scan < g.V ? max_scan : g.V;
    if (need_ncap < 1) need_ncap = 1;
    if (need_ncap > h->ncap) {
        int64_t nc = h->ncap * 2 > need_ncap ? h->ncap * 2 : need_ncap;
        if (nc > g.V) nc = g.V;
        SlotRec *nr = nullptr;
        CKH(h, dalloc(&nr, nc * 2 * g.B));   // two buffers per ring slot
        ++h->buf_gen;
        if (h->slot_recs) {
            for (int s = 0; s < 2 * g.B; ++s)
                CKH(h, hipMemcpyAsync(nr + (int64_t)s * nc, h->slot_recs + (int64_t)s * h->ncap,
                                      (size_t)h->ncap * sizeof(SlotRec), hipMemcpyDeviceToDevice, h->stream));
            CKH(h, hipStreamSynchronize(h->stream));
            CKH(h, hipFree(h->slot_recs));
        }
        h->slot_recs = nr;
        h->ncap = nc;
    }
    // Always hold work buffers, also for an empty first scan: the segment pass writes its
    // end sentinel even when there are no points.
    if (batch_pts < 1) batch_pts = 1;
    if (batch_pts > h->ptcap) {
        int64_t pc = h->ptcap * 2 > batch_pts ? h->ptcap * 2 : batch_pts;
        if (pc < 4096) pc = 4096;
        if (pc >= ((int64_t)1 << 31) - 1) {
            h->err = "too many points in one batch";
            return GVOM_EINVAL;
        }
        CKH(h, hipStreamSynchronize(h->stream));
        // Bumped before anything is freed: a captured sequence is refused from here on, also
        // when a reallocation below fails and leaves the buffers missing.
        ++h->buf_gen;
        h->ptcap = 0;
        void **ptrs[] = {(void **)&h->pts_in, (void **)&h->world, (void **)&h->valid, (void **)&h->k0,
                         (void **)&h->v0, (void **)&h->k1, (void **)&h->v1, (void **)&h->hist, (void **)&h->blk,
                         (void **)&h->seg_key, (void **)&h->seg_start, (void **)&h->mom, (void **)&h->part,
                         (void **)&h->rsec, (void **)&h->rorder};
        for (void **q : ptrs) {
            void *old = *q;
            *q = nullptr;
            if (old) CKH(h, hipFree(old));
        }
        CKH(h, hipMalloc(&h->pts_in, (size_t)pc * 3 * sizeof(double)));
        CKH(h, hipMalloc(&h->world, (size_t)pc * 3 * sizeof(double)));
        CKH(h, dalloc(&h->valid, pc));
        CKH(h, dalloc(&h->k0, pc));
        CKH(h, dalloc(&h->v0, pc));
        CKH(h, dalloc(&h->k1, pc));
        CKH(h, dalloc(&h->v1, pc));
        // the sort's histogram sets; the first must start at zero (see radix_sort)
        h->hwords = radix_hist_entries(pc);
        CKH(h, dalloc(&h->hist, gvom_plan::RADIX_HIST_SETS * h->hwords));
        CKH(h, hipMemsetAsync(h->hist, 0, sizeof(uint32_t) * gvom_plan::RADIX_HIST_SETS * h->hwords, h->stream));
        CKH(h, hipStreamSynchronize(h->stream));
        CKH(h, dalloc(&h->blk, segment_block_words(pc)));
        CKH(h, dalloc(&h->seg_key, pc));
        CKH(h, dalloc(&h->seg_start, pc + 1));
        CKH(h, dalloc(&h->mom, pc));
        CKH(h, dalloc(&h->part, pc));
        CKH(h, dalloc(&h->rsec, pc));
        CKH(h, dalloc(&h->rorder, pc));
        h->ptcap = pc;
    }
    return GVOM_OK;
}

static int ensure_tile_scratch(gvom_handle *h);
static bool tiled_fits(const Geo &g);

extern "C" int gvom_reserve(gvom_handle *h, int64_t max_points, int64_t max_batch_points) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (max_points < 0 || max_batch_points < 0) {
        h->err = "negative size";
        return GVOM_EINVAL;
    }
    if (max_batch_points < max_points) max_batch_points = max_points;
    const int rc = ensure_capacity(h, max_points, max_batch_points);
    if (rc || !h->ray_tiled || !tiled_fits(h->g) || h->capturing) return rc;
    return ensure_tile_scratch(h);   // the tiled cast's scratch, so later calls need not allocate
}

static void host_origin(const Geo &g, const double ego[3], double o[3]) {
    o[0] = std::floor((ego[0] / g.xy_res) - (double)g.XY / 2.0);
    o[1] = std::floor((ego[1] / g.xy_res) - (double)g.XY / 2.0);
    o[2] = std::floor((ego[2] / g.z_res) - (double)g.Z / 2.0);
}

// Timing events: process records 0..4 and 9 (end of the ray-cast kernel) of a new set,
// combine records 5..8 of the current set (or of a new one when that set already has a combine).
static constexpr int EV_PER_SET = 10;
static constexpr int MAX_SETS = 8192;

static void stage_event(gvom_handle *h, int k, hipStream_t on = nullptr) {
    if (!h->timing) return;
    if (k == 0 || (k == 5 && (h->cur_set < 0 || (h->set_used[h->cur_set] >> 5 & 1)))) {
        if ((int)h->set_used.size() >= MAX_SETS) return;
        h->set_used.push_back(0);
        h->cur_set = (int)h->set_used.size() - 1;
        while ((int)h->ev_pool.size() < EV_PER_SET * (h->cur_set + 1)) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) return;
            h->ev_pool.push_back(e);
        }
    }
    if (h->cur_set < 0) return;
    if (hipEventRecord(h->ev_pool[EV_PER_SET * h->cur_set + k], on ? on : h->stream) == hipSuccess)
        h->set_used[h->cur_set] |= 1 << k;
}

// Tiled ray-cast scratch: one zeroed table per scan a batch can hold (min(B, 32)); the
// conversion kernel leaves it zeroed after every use.
// Everything enqueued on the main stream so far may use the scan stage's work buffers: the
// next overlapped scan stage waits for it.
static hipError_t mark_back(gvom_handle *h) { return hipEventRecord(h->ev_back, h->stream); }


// Maps whose tiled scratch keys would not fit in 31 bits use the direct cast instead.
static bool tiled_fits(const Geo &g) { return gvom_plan::tiled_cast_fits(g.XY, g.Z, g.B); }

static int ensure_tile_scratch(gvom_handle *h) {
    if (h->tscratch) return GVOM_OK;
    const Geo &g = h->g;
    const int n = gvom_plan::tile_tables(g.B);
    const int64_t words = tile_scratch_words(g);
    if (!tiled_fits(g)) {
        h->err = "map too large for the tiled ray cast";
        return GVOM_EINVAL;
    }
    if (h->capturing) {
        h->err = "the tiled ray cast's scratch must exist before graph capture (process one scan first)";
        return GVOM_ESTATE;
    }
    CKH(h, dalloc(&h->tscratch, words * n));
    ++h->buf_gen;
    CKH(h, hipMemsetAsync(h->tscratch, 0, (size_t)words * n * sizeof(int32_t), h->stream));
    CKH(h, mark_back(h));
    h->tscratch_scans = n;
    h->tscratch_words = words * n;
    return GVOM_OK;
}

// The scan stage's per-point work buffers: the handle's own (grown on demand), or the
// device-static scratch of gvom_process_pointcloud_device_noalloc.
struct WorkBufs {
    void *pts_in = nullptr;
    void *world = nullptr;
    uint8_t *valid = nullptr;
    uint32_t *k0 = nullptr, *v0 = nullptr, *k1 = nullptr, *v1 = nullptr;
    uint32_t *hist = nullptr;
    int64_t hwords = 0;
    uint32_t *blk = nullptr;
    int64_t blk_words = 0;
    uint32_t *seg_key = nullptr;
    int32_t *seg_start = nullptr;
    Moment *mom = nullptr, *part = nullptr;
    uint8_t *rsec = nullptr;
    uint32_t *rorder = nullptr;
    int64_t cap = 0;          // points
    bool clear_hist = false;  // zero the first histogram before the transform kernel counts into it
};

static WorkBufs scratch_bufs(uint8_t *base) {
    const ScratchLayout &L = NOALLOC_LAYOUT;
    WorkBufs w;
    w.world = base + L.world;
    w.valid = base + L.valid;
    w.k0 = (uint32_t *)(base + L.k0);
    w.v0 = (uint32_t *)(base + L.v0);
    w.k1 = (uint32_t *)(base + L.k1);
    w.v1 = (uint32_t *)(base + L.v1);
    w.hist = (uint32_t *)(base + L.hist);
    w.hwords = L.hwords;
    w.blk = (uint32_t *)(base + L.blk);
    w.blk_words = L.blk_words;
    w.seg_key = (uint32_t *)(base + L.seg_key);
    w.seg_start = (int32_t *)(base + L.seg_start);
    w.mom = (Moment *)(base + L.mom);
    w.part = (Moment *)(base + L.part);
    w.rsec = base + L.rsec;
    w.rorder = (uint32_t *)(base + L.rorder);
    w.cap = NOALLOC_POINTS;
    w.clear_hist = true;   // another call may have left the shared histogram dirty
    return w;
}

static WorkBufs handle_bufs(gvom_handle *h) {
    WorkBufs w;
    w.pts_in = h->pts_in;
    w.world = h->world;
    w.valid = h->valid;
    w.k0 = h->k0;
    w.v0 = h->v0;
    w.k1 = h->k1;
    w.v1 = h->v1;
    w.hist = h->hist;
    w.hwords = h->hwords;
    w.blk = h->blk;
    w.blk_words = segment_block_words(h->ptcap);
    w.seg_key = h->seg_key;
    w.seg_start = h->seg_start;
    w.mom = h->mom;
    w.part = h->part;
    w.rsec = h->rsec;
    w.rorder = h->rorder;
    w.cap = h->ptcap;
    return w;
}

// The scan stage for S scans (gvom.py:105-220 for each).
// cast = false (gvom_diff_forward in geometry-only mode): no ray cast; the slot tables are
// cleared instead, so the slot holds the occupied voxels and records without pass counts.
// scratch: the scan stage runs on these work buffers and allocates nothing (null: the
// handle's buffers, grown as needed).
static int process_impl(gvom_handle *h, const void *xyz, const int64_t *offsets, int S, int dtype,
                        const double *ego, const void *T, int t_dtype, bool on_device, bool may_overlap,
                        bool cast = true, const WorkBufs *scratch = nullptr) {
    const Geo &g = h->g;
    if (S < 1 || S > GVOM_MAX_BATCH || S > g.B) {
        h->err = "number of scans must be in [1, min(32, buffer_size)]";
        return GVOM_EINVAL;
    }
    if (dtype != GVOM_F32 && dtype != GVOM_F64) {
        h->err = "point dtype must be GVOM_F32 or GVOM_F64";
        return GVOM_EINVAL;
    }
    if (T && t_dtype != GVOM_F32 && t_dtype != GVOM_F64 && t_dtype != GVOM_NO_TRANSFORM) {
        h->err = "transform dtype must be GVOM_F32, GVOM_F64 or GVOM_NO_TRANSFORM";
        return GVOM_EINVAL;
    }
    if (offsets[0] != 0) {
        h->err = "offsets[0] must be 0";
        return GVOM_EINVAL;
    }
    int64_t max_scan = 0;
    for (int s = 0; s < S; ++s) {
        int64_t n = offsets[s + 1] - offsets[s];
        if (n < 0) {
            h->err = "offsets must be non-decreasing";
            return GVOM_EINVAL;
        }
        if (n > max_scan) max_scan = n;
        for (int k = 0; k < 3; ++k) {
            const double res = k < 2 ? g.xy_res : g.z_res;
            if (!std::isfinite(ego[3 * s + k]) || std::fabs(ego[3 * s + k] / res) >= 4194304.0) {
                h->err = "ego position must be finite with |ego / resolution| < 2^22";
                return GVOM_EINVAL;
            }
        }
    }
    const int64_t NT = offsets[S];
    if (NT > 0 && !xyz) {
        h->err = "null point array";
        return GVOM_EINVAL;
    }
    int sbits = 0;
    while ((1 << sbits) < S) ++sbits;
    if (g.VB + sbits > 32) {
        h->err = "batch too large for 32-bit sort keys on this map size";
        return GVOM_EINVAL;
    }
    // A rejected call leaves a pending gvom_diff_backward valid unless its buffers moved.
    const int64_t gen0 = h->buf_gen;
    if (scratch) {
        const int64_t need_ncap = max_scan < g.V ? (max_scan > 0 ? max_scan : 1) : g.V;
        if (NT > scratch->cap) {
            h->err = "too many points for the allocation-free entry point (at most " +
                     std::to_string(scratch->cap) + ")";
            return GVOM_EINVAL;
        }
        if (need_ncap > h->ncap || !h->slot_recs) {
            h->err = "slot records too small for this scan: reserve them first (gvom_reserve)";
            return GVOM_ESTATE;
        }
    } else {
        int rc = ensure_capacity(h, max_scan, NT);
        if (h->buf_gen != gen0) h->diff_n = -1;
        if (rc) return rc;
    }
    const WorkBufs W = scratch ? *scratch : handle_bufs(h);
    if (!h->slot_recs || !W.world || !W.valid || !W.k0 || !W.v0 || !W.k1 || !W.v1 || !W.hist || !W.blk ||
        !W.seg_key || !W.seg_start || !W.mom || !W.part || !W.rsec || !W.rorder || W.cap < NT || W.cap < 1 ||
        (!on_device && !W.pts_in)) {
        h->err = "internal: work buffers not allocated";
        return GVOM_ESTATE;
    }

    ScanDesc sdh[GVOM_MAX_BATCH];
    std::memset(sdh, 0, sizeof(sdh));
    for (int s = 0; s < S; ++s) {
        ScanDesc &d = sdh[s];
        d.t_mode = (T && t_dtype != GVOM_NO_TRANSFORM) ? t_dtype : GVOM_NO_TRANSFORM;
        if (d.t_mode == GVOM_F32) {
            const float *t = (const float *)T + 16 * s;
            for (int k = 0; k < 12; ++k) d.T[k] = (double)t[k];
        } else if (d.t_mode == GVOM_F64) {
            const double *t = (const double *)T + 16 * s;
            for (int k = 0; k < 12; ++k) d.T[k] = t[k];
        }
        d.ego[0] = ego[3 * s];
        d.ego[1] = ego[3 * s + 1];
        d.ego[2] = ego[3 * s + 2];
        d.offset = offsets[s];
        d.count = offsets[s + 1] - offsets[s];
    }
    // without allocation, a handle whose tiled scratch does not exist yet casts directly
    // (same counts)
    const bool tiled = cast && h->ray_tiled != 0 && tiled_fits(g) && (!scratch || h->tscratch);
    if (tiled) {
        int rc_t = ensure_tile_scratch(h);
        if (h->buf_gen != gen0) h->diff_n = -1;
        if (rc_t) return rc_t;
    }
    hipStream_t st = h->stream;
    // The stage writes only work buffers, the tiled scratch and the slots' second buffers
    // (new_buf), never what a combine reads; with overlap on it runs on `front` and may proceed
    // while a combine enqueued before this call still runs on st.  Only its commit (ring
    // state, buffer bits) goes on st, after that combine.
    const bool ovl = h->overlap && on_device && may_overlap && (h->nested_capture || !h->capturing);
    hipStream_t fs = ovl ? h->front : st;
    // a stage on the main stream reuses the work buffers pending metrics still read; an
    // overlapped one runs on `front` after them
    if (!ovl) CKH(h, join_metrics(h));
    if (ovl) CKH(h, hipStreamWaitEvent(fs, h->ev_back, 0));
    const size_t esz = dtype == GVOM_F32 ? 4 : 8;
    const void *pts = xyz;
    ScanDesc *sd = h->sd;
    if (h->capturing) {
        if (!on_device) {
            h->err = "only device-pointer process calls can be captured";
            return GVOM_ESTATE;
        }
        if (h->cap_sd_dev) {
            sd = h->cap_sd_dev;   // an internal capture (one scan), its descriptor written before each replay
        } else {
            // descriptors come from the graph pool, written by gvom_graph_launch before each replay
            const int64_t at = (int64_t)h->cap_sd.size();
            if (at + S > h->graph_sd_cap) {
                h->err = "too many scans in one captured sequence";
                return GVOM_ECAP;
            }
            sd = h->graph_sd + at;
            for (int s = 0; s < S; ++s) h->cap_sd.push_back(sdh[s]);
            h->cap_ops.push_back(S);
        }
    } else {
        if (!on_device && NT > 0) {
            // host points: copied into a pinned staging buffer, then one asynchronous transfer;
            // the two buffers alternate, so this call waits only for the transfer two calls back
            const size_t bytes = (size_t)NT * 3 * esz;
            const int k = h->pin_next;
            CKH(h, hipEventSynchronize(h->ev_pin[k]));
            if (h->pin_pts_bytes[k] < bytes) {
                if (h->pin_pts[k]) CKH(h, hipHostFree(h->pin_pts[k]));
                h->pin_pts[k] = nullptr;
                h->pin_pts_bytes[k] = 0;
                ++h->pin_gen;
                CKH(h, hipHostMalloc(&h->pin_pts[k], bytes, hipHostMallocDefault));
                h->pin_pts_bytes[k] = bytes;
            }
            std::memcpy(h->pin_pts[k], xyz, bytes);
            CKH(h, hipMemcpyAsync(W.pts_in, h->pin_pts[k], bytes, hipMemcpyHostToDevice, fs));
            CKH(h, hipEventRecord(h->ev_pin[k], fs));
            h->pin_next = 1 - k;
            pts = W.pts_in;
        }
        CKH(h, hipMemcpyAsync(h->sd, sdh, sizeof(ScanDesc) * S, hipMemcpyHostToDevice, fs));
    }
    h->diff_n = -1;   // every check passed: the work buffers a pending gvom_diff_backward reads are reused below
    stage_event(h, 0, fs);
    const bool xcd = cast && h->ray_xcd != 0;
    if (xcd) CKH(h, launch_fill_i32(&h->st->sec_count[0], 0, 16, fs));   // sec_count and sec_cursor
    if (W.clear_hist) CKH(h, launch_fill_i32((int32_t *)W.hist, 0, W.hwords, fs));
    CKH(h, launch_prepare(dtype, pts, W.world, W.k0, W.v0, W.valid, sd, S, NT, g, h->st,
                          xcd ? W.rsec : nullptr, W.hist, W.hwords, fs));
    if (xcd) CKH(h, launch_ray_order(W.rsec, NT, h->st, W.rorder, fs));
    // fork: keys, segments and chunk partials on the side stream, concurrent with the cast
    CKH(h, hipEventRecord(h->ev_fork, fs));
    CKH(h, hipStreamWaitEvent(h->side, h->ev_fork, 0));
    uint32_t *sk = W.k0, *sv = W.v0;
    CKH(h, radix_sort(W.k0, W.v0, W.k1, W.v1, NT, g.VB + sbits, g.VB, W.hist, W.hwords, h->st, &sk, &sv, h->side));
    h->sorted_vals = sv;
    CKH(h, launch_segments(dtype, W.world, sk, sv, NT, sd, S, g, h->st, W.hist, W.hwords, W.blk, W.blk_words,
                           W.seg_start, W.seg_key, W.part, h->side));
    CKH(h, hipEventRecord(h->ev_join, h->side));
    // the tiled path's conversion writes every word of the slot tables; the direct cast counts
    // into cleared tables, cleared here, beside the sort chain (nothing before the cast reads them)
    if (!tiled) CKH(h, launch_clear_slots(h->slot_tables, h->st, S, g, fs));
    stage_event(h, 1, fs);
    if (cast)
        CKH(h, launch_raycast(dtype, W.world, W.valid, sd, S, NT, g, h->st, tiled ? h->tscratch : h->slot_tables,
                              tiled ? h->tscratch_words : h->slot_table_words,
                              h->ray_block, h->ray_rpt, tiled ? 0 : h->ray_box, tiled ? 1 : 0, -1,
                              xcd ? W.rorder : nullptr, h->ray_spread, h->ray_stride, h->ray_patch, fs));
    stage_event(h, 9, fs);
    if (tiled)
        CKH(h, launch_tile_convert(g, h->st, S, -1, h->tscratch, h->tscratch_words, h->slot_tables,
                                   h->slot_table_words, fs));
    stage_event(h, 2, fs);
    // join: the segment sums read the pass counts and the partials
    CKH(h, hipStreamWaitEvent(fs, h->ev_join, 0));
    stage_event(h, 3, fs);
    CKH(h, launch_mom_seg(W.seg_start, W.seg_key, NT, g, h->st, h->slot_tables, h->slot_recs, h->ncap, W.part,
                          W.mom, fs));
    if (ovl) {
        // back on the main stream, after any combine enqueued before this call: the commit
        // needs the tables' occupied words and the records' counts, which the segment sums
        // wrote; the metrics run on beside the commit and the next combine's table pass
        CKH(h, hipEventRecord(h->ev_front, fs));
        CKH(h, hipStreamWaitEvent(st, h->ev_front, 0));
        h->metrics_pending = false;   // `front` ran any earlier stage's metrics before ev_front
    }
    CKH(h, launch_metrics(W.seg_key, W.mom, NT, g, h->st, h->slot_tables, h->slot_recs, h->ncap, S, fs));
    if (ovl) {
        CKH(h, hipEventRecord(h->ev_metrics, fs));
        h->metrics_pending = true;
    }
    CKH(h, launch_commit_process(h->st, sd, g, S, st));
    stage_event(h, 4);
    CKH(h, mark_back(h));

    for (int s = 0; s < S; ++s) {
        int slot = (h->next_slot + s) % g.B;
        h->filled_mask |= 1 << slot;
        host_origin(g, &ego[3 * s], h->slot_origin[slot]);
    }
    h->last_slot = (h->next_slot + S - 1) % g.B;
    h->next_slot = (h->next_slot + S) % g.B;
    h->latest_ego[0] = ego[3 * (S - 1)];
    h->latest_ego[1] = ego[3 * (S - 1) + 1];
    h->latest_ego[2] = ego[3 * (S - 1) + 2];
    return GVOM_OK;
}

// ---- replayed host-array calls ---------------------------------------------------------------
static void drop_graph(hipGraphExec_t &exec, hipGraph_t &graph) {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    exec = nullptr;
    graph = nullptr;
}

// Where a replayed scan's descriptor lies in the pinned staging buffer: after the points of a
// host-array call, at the start for a device-pointer call (src set).
static size_t host_desc_offset(int64_t n, int dtype, const void *src) {
    if (src) return 0;
    const size_t bytes = (size_t)n * 3 * (dtype == GVOM_F32 ? 4 : 8);
    return (bytes + 255) & ~(size_t)255;
}

// Captures the scan stage of one scan of n points twice, into h->hp_exec[k] for k = 0, 1:
// sequence k reads the points from the device buffer src (a device-pointer call) or from
// pinned staging buffer k (src null: a host-array call), and the descriptor from staging
// buffer k at host_desc_offset(n, dtype, src).  Nothing runs and the host mirror is left as it
// was.  The previous call had the same count, dtype and source, so every work buffer already
// holds n points; a capture that reallocated a buffer anyway, or failed, keeps no sequence
// and the call runs the eager path.
static void capture_host_process(gvom_handle *h, int64_t n, int dtype, const double ego[3], const void *T,
                                 int t_dtype, const void *src) {
    for (int k = 0; k < 2; ++k) drop_graph(h->hp_exec[k], h->hp_graph[k]);
    h->hp_n = -1;
    const size_t off = host_desc_offset(n, dtype, src), need = off + sizeof(ScanDesc);
    if (hipStreamSynchronize(h->stream) != hipSuccess) return;   // nothing in flight reads the staging buffers
    void *dev[2] = {nullptr, nullptr};
    for (int k = 0; k < 2; ++k) {
        if (h->pin_pts_bytes[k] < need) {
            if (h->pin_pts[k]) (void)hipHostFree(h->pin_pts[k]);
            h->pin_pts[k] = nullptr;
            h->pin_pts_bytes[k] = 0;
            ++h->pin_gen;
            if (hipHostMalloc(&h->pin_pts[k], need, hipHostMallocDefault) != hipSuccess) {
                h->pin_pts[k] = nullptr;
                (void)hipGetLastError();
                return;
            }
            h->pin_pts_bytes[k] = need;
        }
        if (hipHostGetDevicePointer(&dev[k], h->pin_pts[k], 0) != hipSuccess || !dev[k]) {
            (void)hipGetLastError();
            return;
        }
    }
    const int m_next = h->next_slot, m_last = h->last_slot, m_filled = h->filled_mask;
    double m_origin[GVOM_MAX_BUFFER][3], m_ego[3];
    std::memcpy(m_origin, h->slot_origin, sizeof(m_origin));
    std::memcpy(m_ego, h->latest_ego, sizeof(m_ego));
    const std::string m_err = h->err;
    const int64_t gen = h->buf_gen;
    bool ok = true;
    for (int k = 0; k < 2 && ok; ++k) {
        if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
            (void)hipGetLastError();
            ok = false;
            break;
        }
        h->capturing = true;
        h->cap_sd_dev = reinterpret_cast<ScanDesc *>((uint8_t *)dev[k] + off);
        int64_t offs[2] = {0, n};
        const int rc = process_impl(h, src ? src : dev[k], offs, 1, dtype, ego, T, t_dtype, true, false);
        hipGraph_t graph = nullptr;
        const hipError_t ee = hipStreamEndCapture(h->stream, &graph);
        h->capturing = false;
        h->cap_sd_dev = nullptr;
        (void)mark_back(h);   // ev_back was recorded inside the capture
        h->next_slot = m_next;
        h->last_slot = m_last;
        h->filled_mask = m_filled;
        std::memcpy(h->slot_origin, m_origin, sizeof(m_origin));
        std::memcpy(h->latest_ego, m_ego, sizeof(m_ego));
        h->err = m_err;
        hipGraphExec_t exec = nullptr;
        if (rc || ee != hipSuccess || !graph || h->buf_gen != gen ||
            hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
            if (graph) (void)hipGraphDestroy(graph);
            (void)hipGetLastError();
            ok = false;
            break;
        }
        h->hp_graph[k] = graph;
        h->hp_exec[k] = exec;
    }
    if (!ok) {
        for (int k = 0; k < 2; ++k) drop_graph(h->hp_exec[k], h->hp_graph[k]);
        return;
    }
    h->hp_n = n;
    h->hp_dtype = dtype;
    h->hp_src = src;
    h->hp_gen = gen;
    h->hp_cfg = h->cfg_gen;
    h->hp_pin_gen = h->pin_gen;
}

constexpr int NOT_REPLAYED = 100;

// A scan call through a captured scan stage: the descriptor (and a host-array call's points)
// are written into one of the two pinned staging buffers (used in turn, so the call waits only
// for the sequence two calls back, the last reader of that buffer) and the sequence bound to
// that buffer is launched.  on_device: xyz is a device buffer the sequence reads in place.
// Returns NOT_REPLAYED when the call is not taken; the eager path then runs it and reports
// any error in its arguments.
static int process_host_replay(gvom_handle *h, const void *xyz, int64_t n, int dtype, const double ego[3],
                               const void *T, int t_dtype, bool on_device) {
    const Geo &g = h->g;
    if (h->capturing || h->timing || n <= 0 || !xyz || (dtype != GVOM_F32 && dtype != GVOM_F64))
        return NOT_REPLAYED;
    if (T && t_dtype != GVOM_F32 && t_dtype != GVOM_F64 && t_dtype != GVOM_NO_TRANSFORM) return NOT_REPLAYED;
    for (int k = 0; k < 3; ++k) {
        const double res = k < 2 ? g.xy_res : g.z_res;
        if (!std::isfinite(ego[k]) || std::fabs(ego[k] / res) >= 4194304.0) return NOT_REPLAYED;
    }
    // the sequence runs on the main stream and reuses the work buffers pending metrics read
    CKH(h, join_metrics(h));
    const void *src = on_device ? xyz : nullptr;
    auto valid = [&]() {
        return h->hp_exec[0] && h->hp_exec[1] && h->hp_n == n && h->hp_dtype == dtype && h->hp_src == src &&
               h->hp_gen == h->buf_gen && h->hp_cfg == h->cfg_gen && h->hp_pin_gen == h->pin_gen;
    };
    if (!valid()) {
        // captured on the second call in a row with this count, dtype and source
        if (n != h->hk_n || dtype != h->hk_dtype || src != h->hk_src) return NOT_REPLAYED;
        capture_host_process(h, n, dtype, ego, T, t_dtype, src);
        if (!valid()) return NOT_REPLAYED;
    }
    ScanDesc d;
    std::memset(&d, 0, sizeof(d));
    d.t_mode = (T && t_dtype != GVOM_NO_TRANSFORM) ? t_dtype : GVOM_NO_TRANSFORM;
    if (d.t_mode == GVOM_F32) {
        for (int k = 0; k < 12; ++k) d.T[k] = (double)((const float *)T)[k];
    } else if (d.t_mode == GVOM_F64) {
        for (int k = 0; k < 12; ++k) d.T[k] = ((const double *)T)[k];
    }
    for (int k = 0; k < 3; ++k) d.ego[k] = ego[k];
    d.offset = 0;
    d.count = n;
    const size_t pbytes = (size_t)n * 3 * (dtype == GVOM_F32 ? 4 : 8);
    const size_t off = host_desc_offset(n, dtype, src);
    const int k = h->pin_next;
    CKH(h, hipEventSynchronize(h->ev_pin[k]));
    if (!on_device) std::memcpy(h->pin_pts[k], xyz, pbytes);
    std::memcpy((uint8_t *)h->pin_pts[k] + off, &d, sizeof(d));
    h->diff_n = -1;   // the sequence reuses the work buffers a pending gvom_diff_backward reads
    CKH(h, hipGraphLaunch(h->hp_exec[k], h->stream));
    CKH(h, hipEventRecord(h->ev_pin[k], h->stream));
    h->pin_next = 1 - k;
    CKH(h, mark_back(h));
    // host mirror of the ring, as the eager call leaves it
    const int slot = h->next_slot;
    h->filled_mask |= 1 << slot;
    host_origin(g, ego, h->slot_origin[slot]);
    h->last_slot = slot;
    h->next_slot = (slot + 1) % g.B;
    for (int j = 0; j < 3; ++j) h->latest_ego[j] = ego[j];
    return GVOM_OK;
}

extern "C" int gvom_process_pointcloud(gvom_handle *h, const void *xyz, int64_t n, int32_t dtype,
                                       const double ego[3], const void *T, int32_t t_dtype) {
    if (!h || !ego || n < 0) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    const int rr = process_host_replay(h, xyz, n, dtype, ego, T, t_dtype, false);
    if (rr != NOT_REPLAYED) return rr;
    int64_t off[2] = {0, n};
    const int rc = process_impl(h, xyz, off, 1, dtype, ego, T, t_dtype, false, false);
    if (rc == GVOM_OK) {
        h->hk_n = n;
        h->hk_dtype = dtype;
        h->hk_src = nullptr;
    }
    return rc;
}

extern "C" int gvom_process_pointcloud_device(gvom_handle *h, const void *xyz, int64_t n, int32_t dtype,
                                              const double ego[3], const void *T, int32_t t_dtype) {
    if (!h || !ego || n < 0) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    // with overlap on, the stage forks from the previous commit and runs beside the combine
    // enqueued before it, which a sequence launched on the main stream could not
    const bool replayable = !h->overlap;
    if (replayable) {
        const int rr = process_host_replay(h, xyz, n, dtype, ego, T, t_dtype, true);
        if (rr != NOT_REPLAYED) return rr;
    }
    int64_t off[2] = {0, n};
    const int rc = process_impl(h, xyz, off, 1, dtype, ego, T, t_dtype, true, true);
    if (rc == GVOM_OK) {
        h->hk_n = replayable ? n : -1;
        h->hk_dtype = dtype;
        h->hk_src = xyz;
    }
    return rc;
}

extern "C" int gvom_process_pointcloud_device_noalloc(gvom_handle *h, const void *xyz, int64_t n, int32_t dtype,
                                                      const double ego[3], const void *T, int32_t t_dtype) {
    if (!h || !ego || n < 0) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    if (h->timing) {
        h->err = "stage timing creates events on demand: disable it for the allocation-free entry point";
        return GVOM_ESTATE;
    }
    if (!h->noalloc_scratch) {
        h->err = "internal: no device scratch";
        return GVOM_ESTATE;
    }
    if (n > NOALLOC_POINTS) {
        h->err = "too many points for the allocation-free entry point (at most " + std::to_string(NOALLOC_POINTS) + ")";
        return GVOM_EINVAL;
    }
    const WorkBufs w = scratch_bufs(h->noalloc_scratch);
    int64_t off[2] = {0, n};
    return process_impl(h, xyz, off, 1, dtype, ego, T, t_dtype, true, true, true, &w);
}

extern "C" int gvom_process_batch(gvom_handle *h, const void *xyz, const int64_t *offsets, int32_t n_scans,
                                  int32_t dtype, const double *ego, const void *T, int32_t t_dtype,
                                  int32_t on_device) {
    if (!h || !offsets || !ego) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    return process_impl(h, xyz, offsets, n_scans, dtype, ego, T, t_dtype, on_device != 0, on_device != 0);
}

// PointCloud2 input (gvom_ros.py:108): gather x, y, z into float64 points, then the
// scan stage as for a float64 array.
extern "C" int gvom_process_pointcloud2(gvom_handle *h, const void *data, const gvom_cloud_layout *L,
                                        const double ego[3], const void *T, int32_t t_dtype, int32_t on_device) {
    if (!h || !L || !ego) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    static const int fsize[9] = {0, 1, 1, 2, 2, 4, 4, 4, 8};
    if (L->height < 0 || L->width < 0 || L->point_step <= 0 || L->row_step < 0) {
        h->err = "PointCloud2 layout: negative size or non-positive point_step";
        return GVOM_EINVAL;
    }
    const int64_t n = L->height * L->width;
    if (L->width > 0 && L->row_step < L->width * L->point_step) {
        h->err = "PointCloud2 layout: row_step < width * point_step";
        return GVOM_EINVAL;
    }
    for (int k = 0; k < 3; ++k) {
        if (L->datatype[k] < 1 || L->datatype[k] > 8) {
            h->err = "PointCloud2 layout: x, y, z datatypes must be PointField codes 1..8";
            return GVOM_EINVAL;
        }
        if (L->offset[k] < 0 || L->offset[k] + fsize[L->datatype[k]] > L->point_step) {
            h->err = "PointCloud2 layout: field outside the point";
            return GVOM_EINVAL;
        }
    }
    if (n > 0 && !data) {
        h->err = "null PointCloud2 data";
        return GVOM_EINVAL;
    }
    const int64_t bytes = n > 0 ? (L->height - 1) * L->row_step + L->width * L->point_step : 0;
    int rc = ensure_capacity(h, n, n);
    if (rc) return rc;
    hipStream_t st = h->stream;
    const uint8_t *src = (const uint8_t *)data;
    if (!on_device && n > 0) {
        if (h->capturing) {
            h->err = "only device-pointer process calls can be captured";
            return GVOM_ESTATE;
        }
        if (bytes > h->rawcap) {
            CKH(h, hipStreamSynchronize(st));
            if (h->raw) CKH(h, hipFree(h->raw));
            h->raw = nullptr;
            h->rawcap = 0;
            CKH(h, dalloc(&h->raw, bytes));
            h->rawcap = bytes;
        }
        CKH(h, hipMemcpyAsync(h->raw, data, (size_t)bytes, hipMemcpyHostToDevice, st));
        src = h->raw;
    }
    double *xyz = (double *)h->pts_in;
    CKH(h, launch_cloud_gather(src, n, L->width, L->point_step, L->row_step, L->offset, L->datatype,
                               L->is_bigendian ? 1 : 0, xyz, st));
    int64_t off[2] = {0, n};
    return process_impl(h, xyz, off, 1, GVOM_F64, ego, T, t_dtype, true, false);
}

// The combine stage (gvom.py:222-393).
static int combine_impl(gvom_handle *h, uint8_t *host_maps = nullptr) {
    const Geo &g = h->g;
    if (!(h->filled_mask & (1 << h->last_slot))) return GVOM_NO_DATA;
    hipStream_t st = h->stream;
    if (h->capturing) h->cap_ops.push_back(-1);
    const int cur = 1 - h->parity;
    stage_event(h, 5);
    CKH(h, launch_comb_tables(g, h->st, h->slot_tables, h->zero_page, h->comb_tables, h->col_count, h->occ_bits,
                              h->height, h->inferred, st));
    stage_event(h, 6);
    CKH(h, launch_scan_columns(h->col_count, h->col_count_words, g.X, h->chunk_tot, h->chunk_tot_words, h->st, st));
    CKH(h, join_metrics(h));   // the record merge reads the newest stage's metrics
    CKH(h, launch_comb_records(g, h->st, h->slot_tables, h->slot_recs, h->ncap, h->comb_tables, h->comb_recs,
                               h->ccap, h->col_count, h->chunk_tot, h->occ_bits, h->occ_list, h->height, st));
    stage_event(h, 7);
    CKH(h, launch_maps2d(g, h->st, h->height, h->inferred, h->comb_tables, h->comb_recs, h->ccap, h->slope_x,
                         h->slope_y, h->rough, h->delta, h->pos, h->neg, h->vis, host_maps, st));
    stage_event(h, 8);
    h->parity = cur;
    h->have_last = 1;
    return GVOM_OK;
}

static void origin_metres(const gvom_handle *h, double out[3]) {
    const double *o = h->slot_origin[h->last_slot];
    out[0] = o[0] * h->g.xy_res;
    out[1] = o[1] * h->g.xy_res;
    out[2] = o[2] * h->g.z_res;
}

extern "C" int gvom_combine_maps(gvom_handle *h, double origin_m[3], int32_t *pos, int32_t *neg, double *rough,
                                 int32_t *vis) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    if (h->capturing) {
        h->err = "combine_maps copies to the host; capture gvom_combine_maps_device instead";
        return GVOM_ESTATE;
    }
    const size_t X = (size_t)h->g.X;
    hipStream_t st = h->stream;
    // the 2D-map pass also writes the four maps into pinned host memory (no separate
    // transfer), from where they are copied into the caller's arrays
    if (!h->pin_maps) {
        CKH(h, hipHostMalloc(&h->pin_maps, 20 * X, hipHostMallocDefault));
        CKH(h, hipHostGetDevicePointer(&h->pin_maps_dev, h->pin_maps, 0));
    }
    int rc = combine_impl(h, (uint8_t *)h->pin_maps_dev);
    if (rc) return rc;
    CKH(h, hipStreamSynchronize(st));
    const uint8_t *m = (const uint8_t *)h->pin_maps;
    if (pos) std::memcpy(pos, m, X * 4);
    if (neg) std::memcpy(neg, m + X * 4, X * 4);
    if (vis) std::memcpy(vis, m + X * 8, X * 4);
    if (rough) std::memcpy(rough, m + X * 12, X * 8);
    if (origin_m) origin_metres(h, origin_m);
    return GVOM_OK;
}

extern "C" int gvom_combine_maps_device(gvom_handle *h, double origin_m[3]) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    int rc = combine_impl(h);
    if (rc) return rc;
    if (origin_m) origin_metres(h, origin_m);
    return GVOM_OK;
}

extern "C" int gvom_occupancy_grids(gvom_handle *h, double density_threshold, double min_roughness,
                                    double max_roughness, int8_t *out, int32_t on_device, double origin_m[3]) {
    if (!h || !out) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    bool has = h->have_last != 0;
    if (h->capturing) {
        if (!on_device) {
            h->err = "occupancy grids to host memory cannot be captured; pass a device buffer";
            return GVOM_ESTATE;
        }
        for (int op : h->cap_ops) has = has || op == -1;
    }
    if (!has) return GVOM_NO_DATA;
    const size_t X = (size_t)h->g.X;
    int8_t *dst = on_device ? out : h->grids;
    CKH(h, launch_grid_payloads(h->g, h->pos, h->neg, h->rough, h->vis, density_threshold, min_roughness,
                                max_roughness, dst, h->stream));
    if (!on_device) {
        CKH(h, hipMemcpyAsync(out, h->grids, 5 * X, hipMemcpyDeviceToHost, h->stream));
        CKH(h, hipStreamSynchronize(h->stream));
    }
    if (origin_m) origin_metres(h, origin_m);
    return GVOM_OK;
}

// ---- image kernels the reference defines but never calls (gvom.py:807-885) ------------------
extern "C" int gvom_expand_binary(gvom_handle *h, const int32_t *in, int32_t *out, int32_t xy_size, double r) {
    if (!h || !in || !out || xy_size < 1) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    CKH(h, launch_expand_binary(in, out, xy_size, r, h->stream));
    return GVOM_OK;
}

extern "C" int gvom_lowpass_binary(gvom_handle *h, const int32_t *in, int32_t *out, int32_t xy_size,
                                   int32_t filter_size, double filter_fraction) {
    if (!h || !in || !out || xy_size < 1 || filter_size < 0) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    CKH(h, launch_lowpass_binary(in, out, xy_size, filter_size, filter_fraction, h->stream));
    return GVOM_OK;
}

extern "C" int gvom_convolve(gvom_handle *h, const double *in, const double *kernel, double *out, int32_t xy_size,
                             int32_t kernel_size) {
    if (!h || !in || !kernel || !out || xy_size < 1 || kernel_size < 1 || kernel_size % 2 == 0) {
        if (h) h->err = "convolve needs device buffers and an odd kernel_size";
        return GVOM_EINVAL;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    CKH(h, launch_convolve(in, kernel, out, xy_size, kernel_size, h->stream));
    return GVOM_OK;
}

extern "C" int gvom_probe_atomics(gvom_handle *h, int32_t pattern, int64_t n_atomics, float *ms) {
    if (!h || !ms || pattern < 0 || pattern > 5 || n_atomics <= 0) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (h->capturing) {
        h->err = "cannot probe while capturing";
        return GVOM_ESTATE;
    }
    const int blocks = 2048;
    int64_t per = n_atomics / ((int64_t)blocks * 256);
    if (per < 1) per = 1;
    if (per > (1 << 20)) per = 1 << 20;
    const uint32_t V = (uint32_t)h->g.V;
    int32_t *table = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = dalloc(&table, (int64_t)V);
    if (e == hipSuccess) e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e == hipSuccess) e = hipMemsetAsync(table, 0, (size_t)V * 4, h->stream);
    if (e == hipSuccess) e = launch_probe_atomics(table, V, blocks, (int)per, pattern, h->stream);   // warm-up
    if (e == hipSuccess) e = hipEventRecord(e0, h->stream);
    if (e == hipSuccess) e = launch_probe_atomics(table, V, blocks, (int)per, pattern, h->stream);
    if (e == hipSuccess) e = hipEventRecord(e1, h->stream);
    if (e == hipSuccess) e = hipEventSynchronize(e1);
    float t = 0.f;
    if (e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
    if (table) (void)hipFree(table);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    CKH(h, e);
    ms[0] = t;
    ms[1] = (float)((double)per * blocks * 256);
    return GVOM_OK;
}

extern "C" int gvom_device_maps(gvom_handle *h, int32_t **pos, int32_t **neg, double **rough, int32_t **vis) {
    if (!h) return GVOM_EINVAL;
    if (pos) *pos = h->pos;
    if (neg) *neg = h->neg;
    if (rough) *rough = h->rough;
    if (vis) *vis = h->vis;
    return GVOM_OK;
}

extern "C" int gvom_sync(gvom_handle *h) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    CKH(h, hipStreamSynchronize(h->stream));
    return GVOM_OK;
}

static int read_state(gvom_handle *h, DevState *out) {
    CKH(h, hipMemcpyAsync(out, h->st, sizeof(DevState), hipMemcpyDeviceToHost, h->stream));
    CKH(h, hipStreamSynchronize(h->stream));
    if (out->wait_fault) {
        h->err = "internal: a block gave up waiting for the counts of an earlier block";
        return GVOM_ESTATE;
    }
    return GVOM_OK;
}

static int export_table(gvom_handle *h, const int32_t *table, int32_t *out) {
    if (!h->export_buf) CKH(h, dalloc(&h->export_buf, h->g.V));
    CKH(h, launch_export_table(h->g, table, h->export_buf, h->stream));
    CKH(h, hipMemcpyAsync(out, h->export_buf, (size_t)h->g.V * 4, hipMemcpyDeviceToHost, h->stream));
    CKH(h, hipStreamSynchronize(h->stream));
    return GVOM_OK;
}

extern "C" int gvom_get_slot(gvom_handle *h, int32_t slot, int32_t *index_map, int64_t cap, int32_t *hit,
                             int32_t *total, float *min_h, double *metrics, double origin[3], int64_t *cells) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (slot < 0 || slot >= h->g.B) {
        h->err = "slot out of range";
        return GVOM_EINVAL;
    }
    if (!(h->filled_mask & (1 << slot))) return GVOM_NO_DATA;
    DevState ds;
    int rc = read_state(h, &ds);
    if (rc) return rc;
    const int64_t C = ds.slot_cells[slot];
    if (cells) *cells = C;
    if (origin)
        for (int k = 0; k < 3; ++k) origin[k] = ds.slot_origin[slot][k];
    if (index_map) {
        rc = export_table(h, h->slot_tables + (int64_t)cur_buf(ds.flip, slot) * h->g.V, index_map);
        if (rc) return rc;
    }
    if (hit || total || min_h || metrics) {
        if (cap < C) {
            h->err = "capacity smaller than the cell count";
            return GVOM_ECAP;
        }
        std::vector<SlotRec> recs((size_t)(C > 0 ? C : 1));
        if (C > 0) {
            CKH(h, hipMemcpy(recs.data(), h->slot_recs + (int64_t)cur_buf(ds.flip, slot) * h->ncap,
                             (size_t)C * sizeof(SlotRec), hipMemcpyDeviceToHost));
        }
        for (int64_t i = 0; i < C; ++i) {
            if (hit) hit[i] = recs[i].hit;
            if (total) total[i] = recs[i].total;
            if (min_h) min_h[i] = recs[i].min_h;
            if (metrics)
                for (int j = 0; j < 10; ++j) metrics[10 * i + j] = recs[i].m[j];
        }
    }
    return GVOM_OK;
}

extern "C" int gvom_get_combined(gvom_handle *h, int32_t *index_map, int64_t cap, int32_t *hit, int32_t *total,
                                 float *min_h, float *metrics, float *eigen, double origin[3], int64_t *cells) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (!h->have_last) return GVOM_NO_DATA;
    DevState ds;
    int rc = read_state(h, &ds);
    if (rc) return rc;
    const int par = h->parity;
    const int64_t C = ds.comb_cells[par];
    if (cells) *cells = C;
    if (origin)
        for (int k = 0; k < 3; ++k) origin[k] = ds.comb_origin[par][k];
    if (index_map) {
        rc = export_table(h, h->comb_tables + (int64_t)par * h->g.V, index_map);
        if (rc) return rc;
    }
    if (hit || total || min_h || metrics || eigen) {
        if (cap < C) {
            h->err = "capacity smaller than the cell count";
            return GVOM_ECAP;
        }
        std::vector<CombRec> recs((size_t)(C > 0 ? C : 1));
        if (C > 0)
            CKH(h, hipMemcpy(recs.data(), h->comb_recs + (int64_t)par * h->ccap, (size_t)C * sizeof(CombRec),
                             hipMemcpyDeviceToHost));
        for (int64_t i = 0; i < C; ++i) {
            if (hit) hit[i] = recs[i].hit;
            if (total) total[i] = recs[i].total;
            if (min_h) min_h[i] = recs[i].min_h;
            if (metrics)
                for (int j = 0; j < 10; ++j) metrics[10 * i + j] = recs[i].m[j];
            if (eigen)
                for (int j = 0; j < 3; ++j) eigen[3 * i + j] = recs[i].ev[j];
        }
    }
    return GVOM_OK;
}

extern "C" int gvom_get_maps(gvom_handle *h, double *height, double *inferred, double *slope_x, double *slope_y,
                             double *rough, double *delta, int32_t *pos, int32_t *neg, int32_t *vis) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (!h->have_last) return GVOM_NO_DATA;
    const size_t X = (size_t)h->g.X;
    hipStream_t st = h->stream;
    struct {
        void *dst;
        const void *src;
        size_t bytes;
    } cp[] = {{height, h->height, X * 8}, {inferred, h->inferred, X * 8}, {slope_x, h->slope_x, X * 8},
              {slope_y, h->slope_y, X * 8}, {rough, h->rough, X * 8},     {delta, h->delta, X * 8},
              {pos, h->pos, X * 4},         {neg, h->neg, X * 4},         {vis, h->vis, X * 4}};
    for (auto &c : cp)
        if (c.dst) CKH(h, hipMemcpyAsync(c.dst, c.src, c.bytes, hipMemcpyDeviceToHost, st));
    CKH(h, hipStreamSynchronize(st));
    return GVOM_OK;
}

extern "C" int gvom_ring_state(const gvom_handle *h, int32_t *buffer_index, int32_t *last_buffer_index,
                               int32_t *filled_mask, int32_t *have_combined) {
    if (!h) return GVOM_EINVAL;
    if (buffer_index) *buffer_index = h->next_slot;
    if (last_buffer_index) *last_buffer_index = h->last_slot;
    if (filled_mask) *filled_mask = h->filled_mask;
    if (have_combined) *have_combined = h->have_last;
    return GVOM_OK;
}

extern "C" int gvom_debug_voxel_map(gvom_handle *h, float *out, int64_t cap, int64_t *rows) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (!h->have_last) return GVOM_NO_DATA;
    DevState ds;
    int rc = read_state(h, &ds);
    if (rc) return rc;
    const int par = h->parity;
    const int64_t C = ds.comb_cells[par];
    if (rows) *rows = C;
    if (!out) return GVOM_OK;
    if (cap < C) {
        h->err = "capacity smaller than the cell count";
        return GVOM_ECAP;
    }
    if (C == 0) return GVOM_OK;
    float *dbuf = nullptr;
    CKH(h, dalloc(&dbuf, C * 8));
    hipError_t e = launch_debug_voxels(h->g, h->st, h->comb_tables, h->comb_recs, h->ccap, par, dbuf, h->stream);
    if (e == hipSuccess) e = hipMemcpyAsync(out, dbuf, (size_t)C * 8 * sizeof(float), hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dbuf);
    CKH(h, e);
    return GVOM_OK;
}

static int debug_height_common(gvom_handle *h, float *out7, float *out3) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (!h->have_last) return GVOM_NO_DATA;
    if (!out7 && !out3) return GVOM_OK;
    const int64_t X = h->g.X;
    float *dbuf = nullptr;
    const int w = out7 ? 7 : 3;
    CKH(h, dalloc(&dbuf, X * w));
    hipError_t e = launch_debug_height(h->g, h->st, h->parity, h->height, h->rough, h->slope_x, h->slope_y, h->delta,
                                       out7 ? dbuf : nullptr, out3 ? dbuf : nullptr, h->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(out7 ? (void *)out7 : (void *)out3, dbuf, (size_t)X * w * sizeof(float),
                           hipMemcpyDeviceToHost, h->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    (void)hipFree(dbuf);
    CKH(h, e);
    return GVOM_OK;
}

extern "C" int gvom_debug_height_map(gvom_handle *h, float *out) { return debug_height_common(h, out, nullptr); }

extern "C" int gvom_debug_inferred_height_map(gvom_handle *h, float *out) {
    return debug_height_common(h, nullptr, out);
}

extern "C" int gvom_set_timing(gvom_handle *h, int32_t enabled) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    h->timing = enabled != 0;
    h->set_used.clear();
    h->cur_set = -1;
    return GVOM_OK;
}

extern "C" int gvom_stage_times(gvom_handle *h, float *ms, int32_t n) {
    if (!h || !ms) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    CKH(h, hipStreamSynchronize(h->stream));
    static const int pairs[9][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 4}, {5, 6}, {6, 7}, {7, 8}, {1, 9}, {9, 2}};
    for (int k = 0; k < n; ++k) {
        ms[k] = 0.f;
        if (k >= 9) continue;
        double sum = 0.0;
        int cnt = 0;
        for (size_t q = 0; q < h->set_used.size(); ++q) {
            int u = h->set_used[q];
            if (!((u >> pairs[k][0]) & 1) || !((u >> pairs[k][1]) & 1)) continue;
            float t = 0.f;
            if (hipEventElapsedTime(&t, h->ev_pool[EV_PER_SET * q + pairs[k][0]],
                                    h->ev_pool[EV_PER_SET * q + pairs[k][1]]) == hipSuccess) {
                sum += t;
                ++cnt;
            }
        }
        ms[k] = cnt ? (float)(sum / cnt) : 0.f;
    }
    return GVOM_OK;
}

// ---- differentiable form -------------------------------------------------------------
// The loss kernels' buffers for n points (a reallocation invalidates captured sequences).
static int diff_buffers(gvom_handle *h, int64_t n) {
    const int64_t X = h->g.X;
    if (!h->diff_colh) {
        CKH(h, dalloc(&h->diff_colh, X));
        CKH(h, dalloc(&h->diff_colpt, X));
        CKH(h, dalloc(&h->diff_sums, 16));
        ++h->buf_gen;
    }
    if (h->diff_rec_cap < h->ncap) {
        if (h->diff_rec) CKH(h, hipFree(h->diff_rec));
        CKH(h, hipMalloc(&h->diff_rec, (size_t)h->ncap * (size_t)diff_rec_bytes()));
        h->diff_rec_cap = h->ncap;
        ++h->buf_gen;
    }
    const int64_t need_pt = n > 0 ? n : 1;
    if (h->diff_pt_cap < need_pt) {
        if (h->diff_pmark) CKH(h, hipFree(h->diff_pmark));
        CKH(h, dalloc(&h->diff_pmark, need_pt));
        h->diff_pt_cap = need_pt;
        ++h->buf_gen;
    }
    const int64_t np = diff_partials(need_pt, X, h->ncap);
    if (h->diff_part_cap < np) {
        if (h->diff_part) CKH(h, hipFree(h->diff_part));
        CKH(h, dalloc(&h->diff_part, np));
        h->diff_part_cap = np;
        ++h->buf_gen;
    }
    return GVOM_OK;
}

static hipError_t enqueue_diff_forward(gvom_handle *h, int64_t n, double w_height) {
    return launch_diff_forward(h->g, h->st, h->last_slot, h->slot_tables, h->slot_recs, h->ncap, h->seg_start,
                               h->sorted_vals, (const double *)h->world, n, h->diff_rec, h->diff_colh, h->diff_colpt,
                               h->diff_pmark, h->diff_part, h->diff_sums, w_height, h->stream);
}

// Captures the scan stage (no ray cast) and the loss kernels for these points into
// h->dg_exec, reading the scan descriptor from h->dg_sd.  Nothing runs; the host mirror is
// left as it was.  On any failure no graph is kept and the caller runs the eager path.
static void capture_diff_graph(gvom_handle *h, const void *points, int64_t n, const double T[16],
                               const double ego[3], double w_height) {
    if (h->dg_exec) (void)hipGraphExecDestroy(h->dg_exec);
    if (h->dg_graph) (void)hipGraphDestroy(h->dg_graph);
    h->dg_exec = nullptr;
    h->dg_graph = nullptr;
    if (!h->dg_sd && dalloc(&h->dg_sd, 1) != hipSuccess) return;
    // the previous call had the same point count, so every buffer already holds it
    if (!h->diff_colh || h->diff_pt_cap < n || h->diff_rec_cap < h->ncap ||
        h->diff_part_cap < diff_partials(n, h->g.X, h->ncap) || h->ptcap < n)
        return;
    if (hipStreamSynchronize(h->stream) != hipSuccess) return;
    const int m_next = h->next_slot, m_last = h->last_slot, m_filled = h->filled_mask;
    double m_origin[3], m_ego[3];
    for (int k = 0; k < 3; ++k) {
        m_origin[k] = h->slot_origin[0][k];
        m_ego[k] = h->latest_ego[k];
    }
    const std::string m_err = h->err;
    if (hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return;
    h->capturing = true;
    h->cap_sd_dev = h->dg_sd;
    int64_t off[2] = {0, n};
    const int rc = process_impl(h, points, off, 1, GVOM_F64, ego, T, GVOM_F64, true, false, false);
    const hipError_t e = rc ? hipErrorInvalidValue : enqueue_diff_forward(h, n, w_height);
    hipGraph_t graph = nullptr;
    const hipError_t ee = hipStreamEndCapture(h->stream, &graph);
    h->capturing = false;
    h->cap_sd_dev = nullptr;
    (void)mark_back(h);   // ev_back was recorded inside the capture
    h->next_slot = m_next;
    h->last_slot = m_last;
    h->filled_mask = m_filled;
    for (int k = 0; k < 3; ++k) {
        h->slot_origin[0][k] = m_origin[k];
        h->latest_ego[k] = m_ego[k];
    }
    h->err = m_err;
    hipGraphExec_t exec = nullptr;
    if (rc || e != hipSuccess || ee != hipSuccess || !graph ||
        hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        (void)hipGetLastError();
        return;
    }
    h->dg_graph = graph;
    h->dg_exec = exec;
    h->dg_pts = points;
    h->dg_n = n;
    h->dg_wh = w_height;
    h->dg_gen = h->buf_gen;
}

// The handle's stream waits for everything enqueued on the caller's stream so far.
static int join_caller(gvom_handle *h, hipStream_t caller) {
    CKH(h, hipEventRecord(h->ev_in, caller));
    CKH(h, hipStreamWaitEvent(h->stream, h->ev_in, 0));
    return GVOM_OK;
}

// The caller's stream waits for everything enqueued on the handle's stream so far.
static int release_caller(gvom_handle *h, hipStream_t caller) {
    CKH(h, hipEventRecord(h->ev_out, h->stream));
    CKH(h, hipStreamWaitEvent(caller, h->ev_out, 0));
    return GVOM_OK;
}

// out (host) or out_dev (device): {L, L_rough, w_height * Var}.  With `async`, the work is
// ordered after the caller's stream and the caller's stream after it, with no host sync.
static int diff_forward_impl(gvom_handle *h, const void *points, int64_t n, const double T[16], const double ego[3],
                             double w_height, double *out, double *out_dev, bool async, hipStream_t caller) {
    if (h->capturing) {
        h->err = "gvom_diff_forward synchronises and cannot be captured";
        return GVOM_ESTATE;
    }
    const Geo &g = h->g;
    // A repeated call (same device points, count and weight) on a geometry-only one-slot
    // handle replays a captured launch sequence with the new transform and ego: the slot the
    // loss reads is then always slot 0, and no ray cast (no launch configuration) is involved.
    const bool graphable = h->diff_geometry && g.B == 1 && !h->timing && n > 0 && points;
    const bool repeat = points == h->dk_pts && n == h->dk_n && w_height == h->dk_wh;
    auto cached = [&]() {
        return h->dg_exec && h->dg_pts == points && h->dg_n == n && h->dg_wh == w_height && h->dg_gen == h->buf_gen;
    };
    if (graphable && repeat && !cached()) capture_diff_graph(h, points, n, T, ego, w_height);
    const ScanDesc *used_sd = h->sd;
    if (async) {
        const int rc = join_caller(h, caller);
        if (rc) return rc;
    }
    if (graphable && repeat && cached()) {
        for (int k = 0; k < 3; ++k) {
            const double res = k < 2 ? g.xy_res : g.z_res;
            if (!std::isfinite(ego[k]) || std::fabs(ego[k] / res) >= 4194304.0) {
                h->err = "ego position must be finite with |ego / resolution| < 2^22";
                return GVOM_EINVAL;
            }
        }
        ScanDesc d;
        std::memset(&d, 0, sizeof(d));
        d.t_mode = GVOM_F64;
        for (int k = 0; k < 12; ++k) d.T[k] = T[k];
        for (int k = 0; k < 3; ++k) d.ego[k] = ego[k];
        d.offset = 0;
        d.count = n;
        h->diff_n = -1;
        CKH(h, hipMemcpyAsync(h->dg_sd, &d, sizeof(d), hipMemcpyHostToDevice, h->stream));
        CKH(h, hipGraphLaunch(h->dg_exec, h->stream));
        CKH(h, mark_back(h));
        used_sd = h->dg_sd;
        // the ring bookkeeping of one process call on a one-slot ring
        h->filled_mask |= 1;
        host_origin(g, ego, h->slot_origin[0]);
        h->last_slot = 0;
        h->next_slot = 0;
        for (int k = 0; k < 3; ++k) h->latest_ego[k] = ego[k];
    } else {
        int64_t off[2] = {0, n};
        int rc = process_impl(h, points, off, 1, GVOM_F64, ego, T, GVOM_F64, true, false, !h->diff_geometry);
        if (rc) return rc;
        rc = diff_buffers(h, n);
        if (rc) return rc;
        CKH(h, enqueue_diff_forward(h, n, w_height));
    }
    const int slot = h->last_slot;
    if (async) {
        if (out_dev)
            CKH(h, hipMemcpyAsync(out_dev, h->diff_sums + 4, 3 * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        const int rc = release_caller(h, caller);
        if (rc) return rc;
    } else {
        double sums[8];
        CKH(h, hipMemcpyAsync(sums, h->diff_sums, sizeof(sums), hipMemcpyDeviceToHost, h->stream));
        CKH(h, hipStreamSynchronize(h->stream));
        if (out) {
            out[0] = sums[4];
            out[1] = sums[5];
            out[2] = sums[6];
        }
    }
    h->diff_n = n;
    h->diff_slot = slot;
    h->diff_wh = w_height;
    h->diff_sd = used_sd;
    // the repeat key changes only when the call succeeded
    h->dk_pts = points;
    h->dk_n = n;
    h->dk_wh = w_height;
    return GVOM_OK;
}

extern "C" int gvom_diff_forward(gvom_handle *h, const void *points, int64_t n, const double T[16],
                                 const double ego[3], double w_height, double out[3]) {
    if (!h || !T || !ego || n < 0) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    return diff_forward_impl(h, points, n, T, ego, w_height, out, nullptr, false, nullptr);
}

extern "C" int gvom_diff_forward_async(gvom_handle *h, const void *points, int64_t n, const double T[16],
                                       const double ego[3], double w_height, double *out_dev, void *stream) {
    if (!h || !T || !ego || n < 0) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    return diff_forward_impl(h, points, n, T, ego, w_height, nullptr, out_dev, true, (hipStream_t)stream);
}

extern "C" int gvom_set_diff_geometry_only(gvom_handle *h, int32_t enabled) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    h->diff_geometry = enabled != 0;
    return GVOM_OK;
}

// dL_dev (device scalar) replaces dL when not null; grad_xi (host) or grad_xi_dev (device).
static int diff_backward_impl(gvom_handle *h, double dL, const double *dL_dev, double *grad_points,
                              double *grad_world, double *grad_xi, double *grad_xi_dev, bool async,
                              hipStream_t caller) {
    if (h->diff_n < 0 || !h->diff_sd) {
        h->err = "gvom_diff_backward needs a preceding gvom_diff_forward";
        return GVOM_ESTATE;
    }
    if (async) {
        const int rc = join_caller(h, caller);
        if (rc) return rc;
    }
    double *g6 = h->diff_sums + 8;
    CKH(h, launch_diff_backward(h->g, h->st, h->diff_slot, h->diff_sd, (const double *)h->world, h->valid, h->slot_tables,
                                h->diff_rec, h->diff_pmark, h->diff_colh, h->diff_sums, h->diff_wh, dL, dL_dev,
                                h->diff_n, grad_points, grad_world, h->diff_part, g6, h->stream));
    if (async) {
        if (grad_xi_dev)
            CKH(h, hipMemcpyAsync(grad_xi_dev, g6, 6 * sizeof(double), hipMemcpyDeviceToDevice, h->stream));
        return release_caller(h, caller);
    }
    double gx[6];
    CKH(h, hipMemcpyAsync(gx, g6, sizeof(gx), hipMemcpyDeviceToHost, h->stream));
    CKH(h, hipStreamSynchronize(h->stream));
    if (grad_xi)
        for (int k = 0; k < 6; ++k) grad_xi[k] = gx[k];
    return GVOM_OK;
}

extern "C" int gvom_diff_backward(gvom_handle *h, double dL, double *grad_points, double *grad_world,
                                  double grad_xi[6]) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    return diff_backward_impl(h, dL, nullptr, grad_points, grad_world, grad_xi, nullptr, false, nullptr);
}

extern "C" int gvom_diff_backward_async(gvom_handle *h, const double *dL_dev, double *grad_points,
                                        double *grad_world, double *grad_xi_dev, void *stream) {
    if (!h || !dL_dev) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    return diff_backward_impl(h, 0.0, dL_dev, grad_points, grad_world, nullptr, grad_xi_dev, true,
                              (hipStream_t)stream);
}

// ---- hipGraph capture of a launch sequence --------------------------------------------
struct gvom_graph {
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> ops;
    std::vector<ScanDesc> sd;
    int64_t gen = -1;   // the handle's buffer generation at capture
};

extern "C" int gvom_graph_begin(gvom_handle *h, int32_t max_scans) {
    if (!h || max_scans < 1) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (h->capturing) {
        h->err = "a capture is already in progress";
        return GVOM_ESTATE;
    }
    if (h->timing) {
        h->err = "disable timing before graph capture";
        return GVOM_ESTATE;
    }
    if (h->ray_tiled && tiled_fits(h->g)) {
        const int rc_t = ensure_tile_scratch(h);
        if (rc_t) return rc_t;
    }
    if (h->graph_sd_cap < max_scans) {
        if (h->graph_sd) CKH(h, hipFree(h->graph_sd));
        CKH(h, dalloc(&h->graph_sd, max_scans));
        h->graph_sd_cap = max_scans;
        ++h->buf_gen;
    }
    CKH(h, hipStreamSynchronize(h->stream));
    h->cap_ops.clear();
    h->cap_sd.clear();
    h->mirror_next = h->next_slot;
    h->mirror_last = h->last_slot;
    h->mirror_filled = h->filled_mask;
    h->mirror_parity = h->parity;
    h->mirror_have = h->have_last;
    CKH(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    h->capturing = true;
    // the first overlapped scan stage of the sequence forks from here
    CKH(h, mark_back(h));
    return GVOM_OK;
}

extern "C" int gvom_graph_end(gvom_handle *h, gvom_graph **out) {
    if (!h || !out) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    *out = nullptr;
    if (!h->capturing) {
        h->err = "no capture in progress";
        return GVOM_ESTATE;
    }
    h->capturing = false;
    // nothing ran: restore the host mirror; gvom_graph_launch replays the bookkeeping
    h->next_slot = h->mirror_next;
    h->last_slot = h->mirror_last;
    h->filled_mask = h->mirror_filled;
    h->parity = h->mirror_parity;
    h->have_last = h->mirror_have;
    gvom_graph *gr = new gvom_graph();
    hipError_t e = hipStreamEndCapture(h->stream, &gr->graph);
    // ev_back was last recorded inside the capture: record it again outside
    const hipError_t em = mark_back(h);
    if (e == hipSuccess) e = em;
    if (e == hipSuccess) e = hipGraphInstantiate(&gr->exec, gr->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        h->err = std::string("graph capture failed: ") + hipGetErrorString(e);
        if (gr->graph) (void)hipGraphDestroy(gr->graph);
        delete gr;
        return GVOM_EHIP;
    }
    gr->ops = h->cap_ops;
    gr->gen = h->buf_gen;
    gr->sd = h->cap_sd;
    *out = gr;
    return GVOM_OK;
}

extern "C" int gvom_graph_scans(const gvom_graph *gr) {
    return gr ? (int)gr->sd.size() : 0;
}

extern "C" int gvom_graph_launch(gvom_handle *h, gvom_graph *gr, const double *ego, const void *T, int32_t t_dtype) {
    if (!h || !gr || (!ego && !gr->sd.empty())) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (h->capturing) {
        h->err = "cannot launch while capturing";
        return GVOM_ESTATE;
    }
    const int n = (int)gr->sd.size();
    if (gr->gen != h->buf_gen) {
        h->err = "device buffers were reallocated after this sequence was captured (a larger scan, batch or "
                 "graph); capture it again";
        return GVOM_ESTATE;
    }
    if (n > h->graph_sd_cap) {
        h->err = "graph descriptor pool smaller than the graph";
        return GVOM_ESTATE;
    }
    std::vector<ScanDesc> sd = gr->sd;
    for (int k = 0; k < n; ++k) {
        for (int j = 0; j < 3; ++j) {
            const double res = j < 2 ? h->g.xy_res : h->g.z_res;
            if (!std::isfinite(ego[3 * k + j]) || std::fabs(ego[3 * k + j] / res) >= 4194304.0) {
                h->err = "ego position must be finite with |ego / resolution| < 2^22";
                return GVOM_EINVAL;
            }
            sd[k].ego[j] = ego[3 * k + j];
        }
        if (T && t_dtype == GVOM_F32) {
            for (int j = 0; j < 12; ++j) sd[k].T[j] = (double)((const float *)T)[16 * k + j];
            sd[k].t_mode = GVOM_F32;
        } else if (T && t_dtype == GVOM_F64) {
            for (int j = 0; j < 12; ++j) sd[k].T[j] = ((const double *)T)[16 * k + j];
            sd[k].t_mode = GVOM_F64;
        } else {
            sd[k].t_mode = GVOM_NO_TRANSFORM;
        }
    }
    if (n > 0) CKH(h, hipMemcpyAsync(h->graph_sd, sd.data(), sizeof(ScanDesc) * n, hipMemcpyHostToDevice, h->stream));
    CKH(h, hipGraphLaunch(gr->exec, h->stream));
    CKH(h, mark_back(h));
    // host mirror of the ring, as the captured calls would have left it
    int k = 0;
    for (int op : gr->ops) {
        if (op > 0) {
            for (int s = 0; s < op; ++s, ++k) {
                const int slot = (h->next_slot + s) % h->g.B;
                h->filled_mask |= 1 << slot;
                host_origin(h->g, sd[k].ego, h->slot_origin[slot]);
            }
            h->last_slot = (h->next_slot + op - 1) % h->g.B;
            h->next_slot = (h->next_slot + op) % h->g.B;
            for (int j = 0; j < 3; ++j) h->latest_ego[j] = sd[k - 1].ego[j];
        } else {
            h->parity = 1 - h->parity;
            h->have_last = 1;
        }
    }
    return GVOM_OK;
}

extern "C" void gvom_graph_destroy(gvom_graph *gr) {
    if (!gr) return;
    if (gr->exec) (void)hipGraphExecDestroy(gr->exec);
    if (gr->graph) (void)hipGraphDestroy(gr->graph);
    delete gr;
}

// ---- autotuner -----------------------------------------------------------------------------
// Times the ray-cast kernel on one real scan for every (block, rays per thread, LDS box)
// configuration, into a scratch table so the ring is not touched (the device state is
// saved and restored around it), and keeps the fastest.
extern "C" int gvom_autotune(gvom_handle *h, const void *points, int64_t n, int32_t dtype, const double ego[3],
                             const void *T, int32_t t_dtype, int32_t chosen[3], float *best_ms) {
    if (!h || !ego || n <= 0 || (dtype != GVOM_F32 && dtype != GVOM_F64)) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (h->capturing) {
        h->err = "cannot autotune while capturing";
        return GVOM_ESTATE;
    }
    const Geo &g = h->g;
    int rc = ensure_capacity(h, n, n);
    if (rc) return rc;
    h->diff_n = -1;
    ScanDesc d;
    std::memset(&d, 0, sizeof(d));
    d.t_mode = (T && t_dtype != GVOM_NO_TRANSFORM) ? t_dtype : GVOM_NO_TRANSFORM;
    for (int k = 0; k < 12 && d.t_mode != GVOM_NO_TRANSFORM; ++k)
        d.T[k] = d.t_mode == GVOM_F32 ? (double)((const float *)T)[k] : ((const double *)T)[k];
    for (int k = 0; k < 3; ++k) d.ego[k] = ego[k];
    d.offset = 0;
    d.count = n;
    hipStream_t st = h->stream;
    DevState *saved = nullptr;
    int32_t *scratch = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int out = GVOM_OK;
    int bb = h->ray_block, br = h->ray_rpt, bx = h->ray_box, bo = h->ray_xcd, bt = h->ray_tiled;
    float best = 1e30f;
    do {
        const bool can_tile = tiled_fits(g);
        if (can_tile && ensure_tile_scratch(h) != GVOM_OK) {
            out = GVOM_EHIP;
            break;
        }
        if (dalloc(&saved, 1) != hipSuccess || dalloc(&scratch, g.V) != hipSuccess ||
            hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
            h->err = "autotune: allocation failed";
            out = GVOM_EHIP;
            break;
        }
        if (hipMemcpyAsync(saved, h->st, sizeof(DevState), hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipMemcpyAsync(h->sd, &d, sizeof(ScanDesc), hipMemcpyHostToDevice, st) != hipSuccess ||
            launch_fill_i32(&h->st->sec_count[0], 0, 16, st) != hipSuccess ||
            launch_prepare(dtype, points, h->world, h->k0, h->v0, h->valid, h->sd, 1, n, g, h->st, h->rsec, nullptr, 0,
                           st) !=
                hipSuccess ||
            launch_ray_order(h->rsec, n, h->st, h->rorder, st) != hipSuccess) {
            h->err = "autotune: setup failed";
            out = GVOM_EHIP;
            break;
        }
        // each timed region holds what the configuration needs per scan: the table clear and
        // the cast into the table, or the cast into the tiled scratch and its conversion
        const int blocks[3] = {64, 128, 256}, rpts[3] = {1, 2, 4};
        for (int ib = 0; ib < 3 && out == GVOM_OK; ++ib)
            for (int ir = 0; ir < 3 && out == GVOM_OK; ++ir)
                for (int icfg = 0; icfg < 6 && out == GVOM_OK; ++icfg) {
                    // icfg: 0-3 table cast (bit 0 box, bit 1 sector order); 4-5 tiled (+ order)
                    const int itile = icfg >= 4, ibox = itile ? 0 : (icfg & 1), iord = itile ? (icfg & 1) : (icfg >> 1);
                    if (itile && !can_tile) continue;
                    float t_min = 1e30f;
                    for (int rep = 0; rep < 3; ++rep) {
                        float t = 0.f;
                        bool ok = hipEventRecord(e0, st) == hipSuccess;
                        if (ok && !itile) ok = launch_fill_i32(scratch, 0, g.V, st) == hipSuccess;
                        ok = ok && launch_raycast(dtype, h->world, h->valid, h->sd, 1, n, g, h->st,
                                                  itile ? h->tscratch : scratch, itile ? h->tscratch_words : g.V,
                                                  blocks[ib], rpts[ir], ibox, itile, 0,
                                                  iord ? h->rorder : nullptr, h->ray_spread, h->ray_stride,
                                                  h->ray_patch, st) == hipSuccess;
                        if (ok && itile)
                            ok = launch_tile_convert(g, h->st, 1, 0, h->tscratch, h->tscratch_words, scratch, g.V,
                                                     st) == hipSuccess;
                        ok = ok && hipEventRecord(e1, st) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                             hipEventElapsedTime(&t, e0, e1) == hipSuccess;
                        if (!ok) {
                            h->err = "autotune: timing run failed";
                            out = GVOM_EHIP;
                            break;
                        }
                        if (t < t_min) t_min = t;
                    }
                    if (out == GVOM_OK && t_min < best) {
                        best = t_min;
                        bb = blocks[ib];
                        br = rpts[ir];
                        bx = ibox;
                        bo = iord;
                        bt = itile;
                    }
                }
        if (hipMemcpyAsync(h->st, saved, sizeof(DevState), hipMemcpyDeviceToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess) {
            h->err = "autotune: restoring state failed";
            out = GVOM_EHIP;
        }
    } while (false);
    if (saved) (void)hipFree(saved);
    if (scratch) (void)hipFree(scratch);
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (out == GVOM_OK) {
        h->ray_block = bb;
        h->ray_rpt = br;
        h->ray_box = bx;
        h->ray_xcd = bo;
        h->ray_tiled = bt;
        ++h->cfg_gen;
        if (chosen) {
            chosen[0] = bb;
            chosen[1] = br;
            chosen[2] = bx | (bo << 1) | (bt << 2) | (h->ray_spread << 3);
        }
        if (best_ms) *best_ms = best;
    }
    return out;
}

extern "C" const char *gvom_hip_runtime_path(void) {
    Dl_info info;
    if (dladdr(reinterpret_cast<void *>(&hipRuntimeGetVersion), &info) && info.dli_fname) return info.dli_fname;
    return "";
}

extern "C" int gvom_hip_versions(int32_t *built, int32_t *running) {
    if (!built || !running) return GVOM_EINVAL;
    *built = HIP_VERSION;
    int v = 0;
    const hipError_t e = hipRuntimeGetVersion(&v);
    *running = v;
    return e == hipSuccess ? GVOM_OK : GVOM_EHIP;
}

extern "C" int gvom_set_overlap(gvom_handle *h, int32_t enabled) {
    if (!h) return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    CKH(h, join_metrics(h));
    if (h->capturing) {
        h->err = "cannot change overlap while capturing";
        return GVOM_ESTATE;
    }
    h->overlap = enabled ? 1 : 0;
    ++h->cfg_gen;
    return GVOM_OK;
}

extern "C" int gvom_set_ray_rings(gvom_handle *h, int64_t points_per_ring, int32_t patch_rings) {
    if (!h) return GVOM_EINVAL;
    if (points_per_ring != 0 && !gvom_plan::ray_patch_valid(points_per_ring, patch_rings)) {
        std::lock_guard<std::mutex> lk(h->mu);
        h->err = "points_per_ring must be a multiple of 64 / patch_rings (patch_rings a power of two <= 64), or 0";
        return GVOM_EINVAL;
    }
    std::lock_guard<std::mutex> lk(h->mu);
    h->ray_stride = points_per_ring;
    h->ray_patch = points_per_ring ? patch_rings : 0;
    ++h->cfg_gen;
    return GVOM_OK;
}

extern "C" int gvom_set_raycast_config(gvom_handle *h, int32_t block, int32_t rays_per_thread, int32_t flags) {
    if (!h || (block != 64 && block != 128 && block != 256) || rays_per_thread < 1 || rays_per_thread > 16 ||
        (flags & 1 && flags & 4))
        return GVOM_EINVAL;
    std::lock_guard<std::mutex> lk(h->mu);
    h->ray_block = block;
    h->ray_rpt = rays_per_thread;
    h->ray_box = flags & 1;
    h->ray_xcd = (flags >> 1) & 1;
    h->ray_tiled = (flags >> 2) & 1;
    h->ray_spread = (flags >> 3) & 1;
    ++h->cfg_gen;
    return GVOM_OK;
}
