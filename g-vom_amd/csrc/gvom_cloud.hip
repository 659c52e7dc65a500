// PointCloud2 input: x, y, z fields of a sensor_msgs/PointCloud2 byte buffer gathered on
// the device into the (N, 3) float64 points the scan stage reads.
//
// The reference's ROS node converts each message with ros_numpy's
// pointcloud2_to_xyz_array (gvom_ros.py:108): a structured view of the buffer, the
// x/y/z fields copied into float64, rows with a non-finite coordinate removed.  Here
// every point is converted in place (no compaction): a point with a non-finite
// coordinate stays non-finite after the transform and the scan stage drops it, and the
// surviving points keep their relative order, so the scan's tables and statistics are
// the ones of the compacted array.
#include <hip/hip_runtime.h>

#include "gvom_internal.h"

namespace {

// sensor_msgs/PointField datatype codes
enum : int { PF_INT8 = 1, PF_UINT8, PF_INT16, PF_UINT16, PF_INT32, PF_UINT32, PF_FLOAT32, PF_FLOAT64 };

__device__ __forceinline__ uint64_t load_bytes(const uint8_t *p, int n, bool big) {
    uint64_t v = 0;
    for (int k = 0; k < n; ++k) {
        const uint64_t b = p[k];
        v |= big ? (b << (8 * (n - 1 - k))) : (b << (8 * k));
    }
    return v;
}

__device__ __forceinline__ double field_value(const uint8_t *p, int type, bool big) {
    switch (type) {
        case PF_INT8: return (double)(int8_t)p[0];
        case PF_UINT8: return (double)p[0];
        case PF_INT16: return (double)(int16_t)(uint16_t)load_bytes(p, 2, big);
        case PF_UINT16: return (double)(uint16_t)load_bytes(p, 2, big);
        case PF_INT32: return (double)(int32_t)(uint32_t)load_bytes(p, 4, big);
        case PF_UINT32: return (double)(uint32_t)load_bytes(p, 4, big);
        case PF_FLOAT32: return (double)__uint_as_float((uint32_t)load_bytes(p, 4, big));
        default: return __longlong_as_double((long long)load_bytes(p, 8, big));
    }
}

__global__ __launch_bounds__(256) void k_cloud_gather(const uint8_t *__restrict__ data, int64_t n, int64_t width,
                                                      int64_t point_step, int64_t row_step, int3 off, int3 type,
                                                      int big, double *__restrict__ xyz) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t r = i / width, c = i - r * width;
    const uint8_t *p = data + r * row_step + c * point_step;
    xyz[3 * i] = field_value(p + off.x, type.x, big != 0);
    xyz[3 * i + 1] = field_value(p + off.y, type.y, big != 0);
    xyz[3 * i + 2] = field_value(p + off.z, type.z, big != 0);
}

}  // namespace

hipError_t launch_cloud_gather(const uint8_t *data, int64_t n, int64_t width, int64_t point_step, int64_t row_step,
                               const int32_t off[3], const int32_t type[3], int big, double *xyz, hipStream_t s) {
    if (n <= 0) return hipSuccess;
    const unsigned nb = (unsigned)((n + 255) / 256);
    k_cloud_gather<<<nb, 256, 0, s>>>(data, n, width, point_step, row_step, make_int3(off[0], off[1], off[2]),
                                      make_int3(type[0], type[1], type[2]), big, xyz);
    return hipGetLastError();
}
