// OccupancyGrid payloads: the int8 maps the reference's ROS node publishes after every
// combine (gvom_ros.py:137-164), made on the device from the combined 2D maps.
//
// Five payloads, each xy_size^2 bytes in the message's order (np.reshape(..., order='F'):
// x fastest, so byte x + y*xy holds map cell [x, y]):
//   0 hard obstacles   max(100 * (positive > density_threshold), negative)   (gvom_ros.py:142)
//   1 soft obstacles   100 * (positive <= density_threshold) * (positive > 0) (146)
//   2 ground certainty visibility * 100                                       (151)
//   3 negative         negative                                               (157)
//   4 roughness        ((clip(r, min_r, max_r) + min_r) / (max_r - min_r)) * 100 (162)
// The 2D maps are [x, y] row-major on the device, so each workgroup transposes one
// 32 x 32 tile through LDS: reads run along y, payload writes along x, both coalesced.
#include <hip/hip_runtime.h>

#include "gvom_internal.h"

namespace {

constexpr int GT = 32;   // tile edge (cells)

// numpy's float64 -> int8 cast as built for x86-64 (the platform the ROS node runs on):
// truncation to int32 (cvttsd2si, whose out-of-range and NaN result is INT32_MIN), then
// the low byte.  E.g. -150.0 -> 106, 300.0 -> 44, NaN -> 0.
__device__ __forceinline__ int8_t f64_to_int8(double v) {
    int32_t t = INT32_MIN;
    if (v > -2147483649.0 && v < 2147483648.0) t = (int32_t)v;   // C conversion truncates
    return (int8_t)(uint8_t)((uint32_t)t & 0xFFu);
}

// numpy.minimum / numpy.maximum: a NaN in either operand is the result.
__device__ __forceinline__ double np_min(double a, double b) { return (a != a) ? a : ((b != b) ? b : (a < b ? a : b)); }
__device__ __forceinline__ double np_max(double a, double b) { return (a != a) ? a : ((b != b) ? b : (a > b ? a : b)); }

__global__ __launch_bounds__(256) void k_grid_payloads(const int32_t *__restrict__ pos, const int32_t *__restrict__ neg,
                                                       const double *__restrict__ rough,
                                                       const int32_t *__restrict__ vis, int XY, double density,
                                                       double min_r, double max_r, double span, int packed,
                                                       int8_t *__restrict__ out) {
    __shared__ int8_t tile[5][GT][GT + 4];   // [payload][y][x]
    const int t = threadIdx.x;
    const int x0 = blockIdx.x * GT, y0 = blockIdx.y * GT;
    const int64_t X = (int64_t)XY * XY;
    // load: thread -> (x = t / 8, y = 4 * (t % 8) .. +3), consecutive threads read along y
    {
        const int lx = t >> 3, ly = (t & 7) * 4;
        const int x = x0 + lx;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            const int y = y0 + ly + k;
            int8_t b0 = 0, b1 = 0, b2 = 0, b3 = 0, b4 = 0;
            if (x < XY && y < XY) {
                const int64_t c = (int64_t)x * XY + y;
                const int32_t p = pos[c], n = neg[c], v = vis[c];
                const double r = rough[c];
                const double pd = (double)p;
                const int64_t hard_pos = (pd > density) ? 100 : 0;
                b0 = (int8_t)(uint8_t)(uint64_t)(hard_pos > (int64_t)n ? hard_pos : (int64_t)n);
                b1 = (int8_t)((pd <= density && p > 0) ? 100 : 0);
                b2 = (int8_t)(uint8_t)((uint32_t)v * 100u);
                b3 = (int8_t)(uint8_t)(uint32_t)n;
                b4 = f64_to_int8(((np_max(np_min(r, max_r), min_r) + min_r) / span) * 100.0);
            }
            tile[0][ly + k][lx] = b0;
            tile[1][ly + k][lx] = b1;
            tile[2][ly + k][lx] = b2;
            tile[3][ly + k][lx] = b3;
            tile[4][ly + k][lx] = b4;
        }
    }
    __syncthreads();
    // store: thread -> (y = t / 8, x = 4 * (t % 8) .. +3), consecutive threads write along x
    const int ly = t >> 3, lx = (t & 7) * 4;
    const int y = y0 + ly, x = x0 + lx;
    if (y >= XY) return;
    for (int pl = 0; pl < 5; ++pl) {
        int8_t *o = out + pl * X + (int64_t)y * XY;
        if (packed) {   // XY % 4 == 0 and out word aligned: x .. x+3 all in range or all out
            if (x < XY) {
                const uint32_t w = (uint32_t)(uint8_t)tile[pl][ly][lx] | ((uint32_t)(uint8_t)tile[pl][ly][lx + 1] << 8) |
                                   ((uint32_t)(uint8_t)tile[pl][ly][lx + 2] << 16) |
                                   ((uint32_t)(uint8_t)tile[pl][ly][lx + 3] << 24);
                *reinterpret_cast<uint32_t *>(o + x) = w;
            }
        } else {
            for (int k = 0; k < 4; ++k)
                if (x + k < XY) o[x + k] = tile[pl][ly][lx + k];
        }
    }
}

}  // namespace

hipError_t launch_grid_payloads(const Geo &g, const int32_t *pos, const int32_t *neg, const double *rough,
                                const int32_t *vis, double density, double min_r, double max_r, int8_t *out,
                                hipStream_t s) {
    const int XY = (int)g.XY;
    dim3 grid((XY + GT - 1) / GT, (XY + GT - 1) / GT);
    // (max_r - min_r) as the node evaluates it: Python arithmetic on the parameters
    const int packed = (XY % 4 == 0) && ((uintptr_t)out % 4 == 0);
    k_grid_payloads<<<grid, 256, 0, s>>>(pos, neg, rough, vis, XY, density, min_r, max_r, max_r - min_r, packed,
                                         out);
    return hipGetLastError();
}
