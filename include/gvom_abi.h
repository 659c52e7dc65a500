/*
 * gvom_abi.h -- C-ABI of libgvom_hip.so, the MI355X (gfx950) implementation of
 * G-VOM's scan->map path.
 *
 * The reference exposes this path as one Python class, Gvom (scripts/gvom.py:8),
 * with no native boundary.  Each entry point below names the reference method or
 * behaviour it replaces.  Calling conventions:
 *   - plain pointers and sizes only; the library owns all device memory, the caller
 *     owns every host buffer it passes;
 *   - entry points return 0 on success, GVOM_NO_DATA (1) where the reference returns
 *     None, and a negative GVOM_E* code on error; gvom_last_error() gives the text;
 *   - every call on one handle is serialised by a per-handle mutex, and all device
 *     work of a handle runs in order on one HIP stream;
 *   - "_device" variants take device pointers and do not synchronise the stream.
 *
 * Map layout of all 2D outputs: C-order [x, y] arrays of xy_size * xy_size, as the
 * reference's combine_maps returns them (gvom.py:393).
 */
#ifndef GVOM_ABI_H
#define GVOM_ABI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GVOM_OK 0
#define GVOM_NO_DATA 1
#define GVOM_EINVAL (-1)
#define GVOM_EHIP (-2)
#define GVOM_ECAP (-3)
#define GVOM_ESTATE (-4)

#define GVOM_F32 0
#define GVOM_F64 1
#define GVOM_NO_TRANSFORM (-1)

/* The 14 constructor parameters of Gvom.__init__ (gvom.py:21-22), in order.
 * slope_obstacle_threshold is the reference's `slope_obsacle_threshold`. */
typedef struct gvom_params {
    double xy_resolution;
    double z_resolution;
    int32_t xy_size;
    int32_t z_size;
    int32_t buffer_size;
    int32_t reserved0;
    double min_distance;
    double positive_obstacle_threshold;
    double negative_obstacle_threshold;
    double slope_obstacle_threshold;
    double robot_height;
    double robot_radius;
    double ground_to_lidar_height;
    int32_t xy_eigen_dist;
    int32_t z_eigen_dist;
} gvom_params;

typedef struct gvom_handle gvom_handle;
typedef struct gvom_graph gvom_graph;

/* Replaces Gvom.__init__ (gvom.py:21-101).  Allocates the ring of buffer_size slots,
 * two combined maps and the 2D maps on the current HIP device. */
int gvom_create(const gvom_params *params, gvom_handle **out);

/* Releases everything gvom_create allocated (no reference counterpart: the
 * reference relies on Numba's garbage collection of device arrays). */
void gvom_destroy(gvom_handle *h);

/* Text of the last error on this handle (or of the last failed gvom_create when
 * h is NULL).  Replaces the Numba/CUDA exceptions the reference raises. */
const char *gvom_last_error(const gvom_handle *h);

/* Pre-sizes the per-scan work buffers and slot records for scans of up to
 * max_points points and batches of up to max_batch_points points, and creates the tiled
 * ray cast's scratch.  Optional: the process calls grow the buffers on demand (reference:
 * per-call cuda.device_array, gvom.py:124-174). */
int gvom_reserve(gvom_handle *h, int64_t max_points, int64_t max_batch_points);

/* Replaces Gvom.Process_pointcloud(pointcloud, ego_position, transform)
 * (gvom.py:105-220).  xyz: host array of n x 3 values of type dtype (GVOM_F32 or
 * GVOM_F64), C-contiguous.  ego: 3 doubles (copied, not aliased).  T: host 4x4
 * row-major matrix of type t_dtype, or NULL / t_dtype == GVOM_NO_TRANSFORM for no
 * transform.  The caller's arrays are not modified (the reference transforms its own
 * device copy, gvom.py:115, 1136-1138). */
int gvom_process_pointcloud(gvom_handle *h, const void *xyz, int64_t n, int32_t dtype,
                            const double ego[3], const void *T, int32_t t_dtype);

/* As gvom_process_pointcloud, with xyz a device pointer (read only).  ego and T are
 * host pointers.  Does not synchronise. */
int gvom_process_pointcloud_device(gvom_handle *h, const void *xyz_device, int64_t n,
                                   int32_t dtype, const double ego[3], const void *T,
                                   int32_t t_dtype);

/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under graph capture, tight serving loops, harnesses that forbid hipMalloc): as
 * gvom_process_pointcloud_device, with the same results, but the scan stage's point work
 * buffers come from a device-static scratch array (about 28 MiB, held by every GPU the
 * library runs on until the process ends) and the call makes no allocation.  Limits:
 * n <= 131072 points (larger n is rejected with GVOM_EINVAL); the slot records must hold
 * the scan's cells already (gvom_reserve(h, n, 0) once, or an earlier process call of at
 * least n points, else GVOM_ESTATE); stage timing must be off.  The scratch is shared by
 * every handle and call on one GPU and keeps the last call's contents: calls that overlap
 * on one GPU (several handles on several host threads or streams) race. */
int gvom_process_pointcloud_device_noalloc(gvom_handle *h, const void *xyz_device, int64_t n,
                                           int32_t dtype, const double ego[3], const void *T,
                                           int32_t t_dtype);

/* Several scans in one launch sequence, with the same result as n_scans
 * consecutive Process_pointcloud calls (gvom.py:105-220).  xyz: flat array of
 * offsets[n_scans] points (host, or device when on_device != 0); offsets: n_scans+1
 * host int64 prefix offsets; ego: n_scans x 3 host doubles; T: n_scans x 16 host
 * values of t_dtype or NULL.  n_scans must not exceed buffer_size. */
int gvom_process_batch(gvom_handle *h, const void *xyz, const int64_t *offsets, int32_t n_scans,
                       int32_t dtype, const double *ego, const void *T, int32_t t_dtype,
                       int32_t on_device);

/* Replaces Gvom.combine_maps() (gvom.py:222-393).  Writes the map origin in metres
 * and the four xy_size x xy_size maps to host memory and synchronises.  Returns
 * GVOM_NO_DATA when no scan has been processed (gvom.py:225-227 prints "ERROR: No data
 * in buffer" and returns None; the library leaves printing to its caller).  Any output pointer may be NULL to skip that copy. */
int gvom_combine_maps(gvom_handle *h, double origin_m[3], int32_t *positive_obstacle,
                      int32_t *negative_obstacle, double *roughness, int32_t *visibility);

/* combine_maps without host copies or synchronisation: the maps stay on the device
 * (see gvom_device_maps).  origin_m is written from the host-side mirror and may be
 * NULL. */
int gvom_combine_maps_device(gvom_handle *h, double origin_m[3]);

/* Device pointers of the latest combined 2D maps (valid until the next combine or
 * gvom_destroy): positive, negative, visibility (int32) and roughness (f64). */
int gvom_device_maps(gvom_handle *h, int32_t **positive_obstacle, int32_t **negative_obstacle,
                     double **roughness, int32_t **visibility);

/* Memory layout of a sensor_msgs/PointCloud2 message's data (the fields the node's
 * input conversion reads).  Point (r, c) starts at byte r * row_step + c * point_step;
 * offset / datatype are the x, y, z PointFields (datatype: PointField codes 1 INT8 ..
 * 8 FLOAT64). */
typedef struct {
    int64_t height, width;
    int64_t point_step, row_step;
    int32_t offset[3];
    int32_t datatype[3];
    int32_t is_bigendian;
    int32_t reserved;
} gvom_cloud_layout;

/* Replaces the node's input path (gvom_ros.py:108-109): ros_numpy's
 * pointcloud2_to_xyz_array (x, y, z as float64, rows with a non-finite coordinate
 * removed) followed by Process_pointcloud.  The x, y, z fields are gathered into float64
 * on the device; a point with a non-finite coordinate is dropped by the scan stage,
 * which leaves the slot identical to processing the compacted float64 array.  data is
 * the message's byte buffer (host memory, or device memory with on_device != 0: then
 * no host copy is made and the call may be captured).  ego and T as for
 * gvom_process_pointcloud.  (ros_numpy assumes row_step == width * point_step; padded
 * rows are addressed by row_step here.) */
int gvom_process_pointcloud2(gvom_handle *h, const void *data, const gvom_cloud_layout *layout,
                             const double ego[3], const void *T, int32_t t_dtype, int32_t on_device);

/* Replaces the payload formatting of the reference's ROS node (gvom_ros.py:137-164):
 * the int8 OccupancyGrid data it publishes from the latest combined maps, made on the
 * device.  out receives 5 payloads of xy_size^2 bytes, in the order GVOM_GRID_HARD,
 * _SOFT, _CERTAINTY, _NEGATIVE, _ROUGHNESS, each in the message's order (x fastest:
 * byte x + y*xy_size is map cell [x, y]).  Values follow the node's numpy expressions,
 * including its float64 -> int8 cast as on x86-64 (truncate to int32, keep the low
 * byte; so the default roughness range lands in [-128, 127] by wrapping).  The node
 * publishes the ground-certainty payload twice (ground_certainty_map and
 * all_ground_certainty_map).  on_device != 0: out is a device pointer, the call is
 * queued on the handle's stream and may be captured; otherwise out is host memory and
 * the call synchronises.  origin_m (may be NULL) receives the grid origin in metres
 * (info.origin.position x, y; the node sets z = 0).  Returns GVOM_NO_DATA before the
 * first combine. */
#define GVOM_GRID_HARD 0
#define GVOM_GRID_SOFT 1
#define GVOM_GRID_CERTAINTY 2
#define GVOM_GRID_NEGATIVE 3
#define GVOM_GRID_ROUGHNESS 4
int gvom_occupancy_grids(gvom_handle *h, double density_threshold, double min_roughness, double max_roughness,
                         int8_t *out, int32_t on_device, double origin_m[3]);

/* The reference's 2D image kernels that its pipeline never calls, on xy_size x xy_size
 * [x, y] row-major device buffers, enqueued on the handle's stream (no synchronisation):
 *   gvom_expand_binary  replaces Gvom.__expand_binary  (gvom.py:807-832): disk dilation of a
 *                       0/1 map, radius r (row offsets |i| <= floor(r), |j| <= floor(sqrt(r^2 - i^2)));
 *   gvom_lowpass_binary replaces Gvom.__lowpass_binary (gvom.py:834-862): 1 where the input is
 *                       non-zero and its (2 filter_size + 1)^2 in-map window mean >= filter_fraction;
 *   gvom_convolve       replaces Gvom.__convolve       (gvom.py:864-885) with its arithmetic as
 *                       written (every term reads input[x + r, y + r]), kernel kernel_size^2 f64
 *                       row-major, r = (kernel_size - 1) / 2; kernel_size must be odd (the
 *                       reference's own kernel never compiles: it passes a float r to range()). */
int gvom_expand_binary(gvom_handle *h, const int32_t *in, int32_t *out, int32_t xy_size, double r);
int gvom_lowpass_binary(gvom_handle *h, const int32_t *in, int32_t *out, int32_t xy_size, int32_t filter_size,
                        double filter_fraction);
int gvom_convolve(gvom_handle *h, const double *in, const double *kernel, double *out, int32_t xy_size,
                  int32_t kernel_size);

/* Atomic-throughput probe (no reference counterpart; a measurement aid): times one launch
 * of non-returning int32 atomic adds into a scratch table of V words on the handle's
 * stream.  pattern 0 scattered (device scope), 1 64 hot words, 2 scattered within the
 * region of the issuing XCD, 3 scattered with workgroup scope (timing only), 4 each
 * wavefront's 64 lanes on 64 consecutive words, 5 each wavefront on 8 lines of 8 lanes.  ms[0] =
 * kernel time in ms, ms[1] = atomics issued.  Synchronises; allocates and frees the table. */
int gvom_probe_atomics(gvom_handle *h, int32_t pattern, int64_t n_atomics, float ms[2]);

/* Waits for all work queued on the handle's stream. */
int gvom_sync(gvom_handle *h);

/* Replaces Gvom.make_debug_voxel_map (gvom.py:395-409): rows of 8 floats per
 * combined cell, in cell-id order.  cap = rows available in out; *rows receives the
 * cell count.  Returns GVOM_NO_DATA when no combine has run. */
int gvom_debug_voxel_map(gvom_handle *h, float *out, int64_t cap, int64_t *rows);

/* Replaces Gvom.make_debug_height_map (gvom.py:411-425): xy_size^2 rows of 7 floats. */
int gvom_debug_height_map(gvom_handle *h, float *out);

/* Replaces Gvom.make_debug_inferred_height_map (gvom.py:427-442): xy_size^2 rows of 3. */
int gvom_debug_inferred_height_map(gvom_handle *h, float *out);

/* State export for parity tests and checkpointing (the reference keeps this state
 * in its ring lists, gvom.py:50-56, and combined_* members, 58-92).  index_map is
 * written in the reference's layout (v = x + y*xy + z*xy*xy) and encoding (>= 0 cell
 * id, -1 unknown, -t-1 free with t passes); compact arrays have cap entries.  Any
 * output may be NULL.  Slot arrays: hit, total (int32), min_h (f32), metrics (f64
 * [C][10]).  Returns GVOM_NO_DATA for an empty slot. */
int gvom_get_slot(gvom_handle *h, int32_t slot, int32_t *index_map, int64_t cap, int32_t *hit,
                  int32_t *total, float *min_h, double *metrics, double origin[3], int64_t *cells);

/* The latest combined map: metrics f32[C][10], eigenvalues f32[C][3]
 * (gvom.py:276, 315). */
int gvom_get_combined(gvom_handle *h, int32_t *index_map, int64_t cap, int32_t *hit,
                      int32_t *total, float *min_h, float *metrics, float *eigen, double origin[3],
                      int64_t *cells);

/* The latest 2D maps, including the intermediate ones the reference keeps as members
 * (height_map, inferred_height_map, x/y_slope_map, roughness_map, guessed_height_delta,
 * gvom.py:327-359).  Each output may be NULL. */
int gvom_get_maps(gvom_handle *h, double *height, double *inferred, double *slope_x,
                  double *slope_y, double *roughness, double *height_delta, int32_t *positive,
                  int32_t *negative, int32_t *visibility);

/* Host-side ring state (gvom.py:211-216): write index, newest slot, filled-slot mask,
 * and whether a combined map exists. */
int gvom_ring_state(const gvom_handle *h, int32_t *buffer_index, int32_t *last_buffer_index,
                    int32_t *filled_mask, int32_t *have_combined);

/* Per-stage device time (HIP events on the handle's stream), in milliseconds, averaged
 * over every process / combine call made since timing was last enabled
 * (gvom_set_timing resets the record).  Stage order: 0 clear+prepare, 1 ray cast,
 * 2 the part of the sort / segments / chunk-partial chain (run on a second stream,
 * concurrently with the ray cast) that is not hidden under it, 3 segment sums+metrics,
 * 4 combine tables, 5 column scan+records+columns, 6 2D maps; then the two parts of stage 1:
 * 7 the ray-cast kernel alone, 8 the tiled scratch's conversion into the slot table (0 when the
 * direct-table cast runs).  n = entries available in ms (at most 9 are filled).  (The reference has commented-out
 * wall-clock prints at these points, gvom.py:107-217, 224, 321-322, 390.) */
int gvom_set_timing(gvom_handle *h, int32_t enabled);
int gvom_stage_times(gvom_handle *h, float *ms, int32_t n);

/* The handle's HIP stream (as void*), for callers that order their own work
 * against it. */
void *gvom_stream(gvom_handle *h);

/* Scan/combine overlap (no reference counterpart; off by default).  Every ring slot has two
 * physical buffers and a scan stage writes the one the ring does not hold, so when enabled a
 * device-pointer scan stage (gvom_process_pointcloud_device, gvom_process_batch with
 * on_device) runs on a second stream as soon as the previous scan stage has finished,
 * concurrently with a gvom_combine_maps_device enqueued in between; only its commit (ring
 * state, buffer selection) and everything else stay ordered on the handle's stream.  The
 * commit waits for the stage up to its segment sums; the stage's neighbourhood metrics (the
 * slot records' moments) may still run beside it and beside the next combine, whose record
 * merge waits for them, as does every other call on the handle when it starts (and
 * gvom_stream, so that work ordered after the stream it returns follows them).  The
 * caller's points must therefore be complete when the process call is made (not merely
 * enqueued on the handle's stream after the preceding combine).  Also applies inside a
 * hipGraph capture, unless the HIP runtime in use is older than the one the library was built
 * against (see gvom_hip_versions).  Fails while capturing. */
int gvom_set_overlap(gvom_handle *h, int32_t enabled);

/* The HIP version this library was built against and the version of the HIP runtime it is
 * bound to in this process (HIP_VERSION encoding: major * 10^7 + minor * 10^5 + patch).
 * Both runtimes carry the SONAME libamdhip64.so.7, so a process that loaded another copy
 * first (torch bundles one) binds the library to that copy.  When the running runtime is
 * older, a captured sequence does not overlap its scan stages with the combines (eager calls
 * still do).  No counterpart in the reference (diagnostics). */
int gvom_hip_versions(int32_t *built, int32_t *running);

/* File of the HIP runtime this library is bound to in this process, so that a caller's own
 * HIP calls (device buffers, copies) go to the same runtime.  No counterpart in the
 * reference. */
const char *gvom_hip_runtime_path(void);

/* hipGraph capture of a launch sequence (no reference counterpart).
 * Between gvom_graph_begin and gvom_graph_end, gvom_process_pointcloud_device,
 * gvom_process_batch(on_device=1) and gvom_combine_maps_device calls are recorded
 * instead of run.  Point buffers and point counts are fixed at capture; the ego
 * positions and transforms of all captured scans (capture order) are passed to every
 * gvom_graph_launch.  Buffers must be sized beforehand (gvom_reserve); calls that copy
 * from or to the host are refused while capturing.  max_scans bounds the scans one
 * sequence may hold. */
int gvom_graph_begin(gvom_handle *h, int32_t max_scans);
int gvom_graph_end(gvom_handle *h, gvom_graph **out);
/* Number of scans recorded in a captured sequence. */
int gvom_graph_scans(const gvom_graph *g);
/* Replays the sequence on the handle's stream with new ego positions (scans x 3 doubles)
 * and transforms (scans x 16 values of t_dtype, or NULL).  Does not synchronise.  Refused
 * (GVOM_ESTATE) when the handle reallocated device buffers after the capture (a larger scan,
 * batch or graph): capture the sequence again. */
int gvom_graph_launch(gvom_handle *h, gvom_graph *g, const double *ego, const void *T, int32_t t_dtype);
void gvom_graph_destroy(gvom_graph *g);

/* Autotuner (no reference counterpart): times the ray cast on one real scan (device
 * points, as gvom_process_pointcloud_device), without touching the ring, for block sizes
 * {64, 128, 256}, {1, 2, 4} rays per thread and the variants
 *   - cast into the slot table (after its clear), LDS near-sensor box off/on;
 *   - cast into a tiled scratch table (4 x 4 x 1 voxels per 64-byte segment, so the run
 *     heads of one wavefront step share fewer atomic requests) and its conversion;
 *   - each with the XCD-aware (azimuth-sector) ray order off/on,
 * and keeps the fastest.  chosen = {block, rays per thread, flags (bit 0 box, bit 1 sector
 * order, bit 2 tiled, bit 3 spread)}; best_ms = its time.  The spread bit (wavefronts of a
 * block taken from across the scan) is not searched: it is kept as set, since it slows the
 * cast alone but speeds up a cast that overlaps a combine.  Synchronises. */
int gvom_autotune(gvom_handle *h, const void *points, int64_t n, int32_t dtype, const double ego[3],
                  const void *T, int32_t t_dtype, int32_t chosen[3], float *best_ms);
/* Sets the ray-cast configuration directly (flags as for gvom_autotune; the box and the
 * tiled scratch exclude each other; bit 3 has no effect with bit 1).  A new handle starts at
 * {128, 1, tiled + spread}; the tiled scratch (min(buffer_size, 32) tables) is allocated by
 * the first process call or gvom_graph_begin. */
int gvom_set_raycast_config(gvom_handle *h, int32_t block, int32_t rays_per_thread, int32_t flags);

/* Ray order hint (no reference counterpart): the scans are organised in rings of
 * points_per_ring consecutive points (a spinning lidar's PointCloud2 rows, e.g. 1024 for a
 * 1024-azimuth sensor).  Each wavefront of the ray cast then traces a patch of
 * 64 / patch_rings adjacent azimuths x patch_rings adjacent rings, whose steps fall on few
 * z-contiguous 64-byte segments of the slot table, instead of 64 azimuths of one ring.  The
 * counts are the same for any order; the hint only changes the cast's speed, and a wrong
 * stride only makes it slower.  points_per_ring must be a multiple of 64 / patch_rings
 * (patch_rings a power of two up to 64); 0 restores the linear order (the default). */
int gvom_set_ray_rings(gvom_handle *h, int64_t points_per_ring, int32_t patch_rings);

/* Differentiable form (no reference counterpart).  Processes one scan like
 * gvom_process_pointcloud into the next ring slot (without the ray cast when
 * gvom_set_diff_geometry_only is on; points: device array of n x 3
 * float64 in the sensor frame; T: host 4x4 float64 sensor-to-world; ego: host 3
 * doubles) and evaluates
 *   L = sum over occupied voxels of the smallest eigenvalue of the voxel's
 *       neighbourhood covariance (the per-scan metrics of gvom.py:1250-1387)
 *     + w_height * variance over columns of the height of their lowest point
 *       (the height map's lowest voxel, gvom.py:572-576),
 * with voxel assignment and counts as the forward found them.  out = {L, L_rough,
 * w_height * Var}.  Synchronises. */
int gvom_diff_forward(gvom_handle *h, const void *points, int64_t n, const double T[16], const double ego[3],
                      double w_height, double out[3]);

/* With enabled != 0, gvom_diff_forward skips the ray cast: the ring slot it writes holds the
 * scan's occupied voxels and records but no ray-pass counts (free space reads as unknown,
 * record totals equal hit counts).  The loss and its gradient use neither, and are the same.
 * For handles that only evaluate the loss; a handle whose ring also feeds gvom_combine_maps
 * keeps the default (off).  On such a handle with buffer_size 1, a gvom_diff_forward repeated
 * with the same device points, count and w_height captures its launch sequence once and then
 * replays it with the new transform and ego (same results, fewer launches).  No counterpart
 * in the reference. */
int gvom_set_diff_geometry_only(gvom_handle *h, int32_t enabled);

/* Gradient of dL * L from the last gvom_diff_forward: grad_points (device, n x 3 f64,
 * sensor frame) and grad_world (device, n x 3 f64, world frame) may be NULL;
 * grad_xi (host, 6 doubles) receives dL/d(rho, phi) for T' = exp(xi^) T.  Synchronises. */
int gvom_diff_backward(gvom_handle *h, double dL, double *grad_points, double *grad_world, double grad_xi[6]);

/* Stream-ordered forms of gvom_diff_forward / gvom_diff_backward for an autograd boundary
 * (no reference counterpart; the reference hands arrays to the device at gvom.py:115).
 * `stream` is the caller's HIP stream (NULL = the null stream): the handle's work waits for
 * everything enqueued on it so far, and it waits for the handle's work, with no host
 * synchronisation.  out_dev: device, 3 doubles {L, L_rough, w_height * Var} (may be NULL);
 * dL_dev: device scalar read by the backward kernel; grad_xi_dev: device, 6 doubles (may be
 * NULL).  Results equal the synchronous forms' bit for bit. */
int gvom_diff_forward_async(gvom_handle *h, const void *points, int64_t n, const double T[16],
                            const double ego[3], double w_height, double *out_dev, void *stream);
int gvom_diff_backward_async(gvom_handle *h, const double *dL_dev, double *grad_points, double *grad_world,
                             double *grad_xi_dev, void *stream);

#ifdef __cplusplus
}
#endif

#endif /* GVOM_ABI_H */
