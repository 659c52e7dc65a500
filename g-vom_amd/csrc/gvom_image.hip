// The reference's three 2D image kernels that its pipeline never calls (gvom.py:807-885):
// disk dilation and low-pass filtering of 0/1 maps, and a "convolution".  Each output cell is
// one thread; maps are [x, y] row-major (y fastest), as every 2D map here.
#include <hip/hip_runtime.h>

#include "gvom_internal.h"

namespace {

constexpr int IT = 16;   // 16 x 16 threads per block

// __expand_binary (gvom.py:807-832): 1 when any cell with |i| <= floor(r) and
// |j| <= floor(sqrt(r*r - i*i)) around (x, y), inside the map, holds exactly 1; else 0.
__global__ __launch_bounds__(IT * IT) void k_expand_binary(const int32_t *__restrict__ in, int32_t *__restrict__ out,
                                                           int XY, double r) {
    const int x = blockIdx.y * IT + threadIdx.y, y = blockIdx.x * IT + threadIdx.x;
    if (x >= XY || y >= XY) return;
    const int ri = (int)floor(r);
    for (int i = -ri; i <= ri; ++i) {
        const int dy = (int)floor(sqrt(r * r - (double)(i * i)));
        if (x + i < 0 || x + i >= XY) continue;
        for (int j = -dy; j <= dy; ++j) {
            if (y + j < 0 || y + j >= XY) continue;
            if (in[(int64_t)(x + i) * XY + (y + j)] == 1) {
                out[(int64_t)x * XY + y] = 1;
                return;
            }
        }
    }
    out[(int64_t)x * XY + y] = 0;
}

// __lowpass_binary (gvom.py:834-862): 0 where the input is 0; elsewhere 1 when the mean of
// the input over the in-map cells of the (2 f + 1)^2 window (f64 sum and count, window
// order x outer, y inner) reaches `fraction`, else 0.
__global__ __launch_bounds__(IT * IT) void k_lowpass_binary(const int32_t *__restrict__ in, int32_t *__restrict__ out,
                                                            int XY, int f, double fraction) {
    const int x = blockIdx.y * IT + threadIdx.y, y = blockIdx.x * IT + threadIdx.x;
    if (x >= XY || y >= XY) return;
    const int64_t c = (int64_t)x * XY + y;
    if (in[c] == 0) {
        out[c] = 0;
        return;
    }
    double val = 0.0, cnt = 0.0;
    for (int i = -f; i <= f; ++i) {
        if (x + i < 0 || x + i >= XY) continue;
        for (int j = -f; j <= f; ++j) {
            if (y + j < 0 || y + j >= XY) continue;
            val += (double)in[(int64_t)(x + i) * XY + (y + j)] * 1.0;
            cnt += 1.0;
        }
    }
    out[c] = (val / cnt) >= fraction ? 1 : 0;
}

// __convolve (gvom.py:864-885) as written: every term reads the input at (x + r, y + r), not
// at (x + i, y + j), so a cell inside the map by r is input[x + r, y + r] times each kernel
// entry, summed in kernel order (f64); other cells are 0.  r = (kernel_size - 1) / 2 for an
// odd kernel_size (the reference computes it with true division, a float that Numba's
// range() rejects, so its kernel never compiles).
__global__ __launch_bounds__(IT * IT) void k_convolve(const double *__restrict__ in, const double *__restrict__ kern,
                                                      double *__restrict__ out, int XY, int ks) {
    const int x = blockIdx.y * IT + threadIdx.y, y = blockIdx.x * IT + threadIdx.x;
    if (x >= XY || y >= XY) return;
    const int r = (ks - 1) / 2;
    double acc = 0.0;
    if (x + r >= 0 && x + r < XY && y + r >= 0 && y + r < XY) {
        const double v = in[(int64_t)(x + r) * XY + (y + r)];
        for (int x2 = 0; x2 <= 2 * r; ++x2)
            for (int y2 = 0; y2 <= 2 * r; ++y2) acc += v * kern[x2 * ks + y2];
    }
    out[(int64_t)x * XY + y] = acc;
}

dim3 image_grid(int XY) { return dim3((unsigned)((XY + IT - 1) / IT), (unsigned)((XY + IT - 1) / IT)); }

}  // namespace

hipError_t launch_expand_binary(const int32_t *in, int32_t *out, int XY, double r, hipStream_t s) {
    k_expand_binary<<<image_grid(XY), dim3(IT, IT), 0, s>>>(in, out, XY, r);
    return hipGetLastError();
}

hipError_t launch_lowpass_binary(const int32_t *in, int32_t *out, int XY, int filter_size, double fraction,
                                 hipStream_t s) {
    k_lowpass_binary<<<image_grid(XY), dim3(IT, IT), 0, s>>>(in, out, XY, filter_size, fraction);
    return hipGetLastError();
}

hipError_t launch_convolve(const double *in, const double *kern, double *out, int XY, int kernel_size,
                           hipStream_t s) {
    k_convolve<<<image_grid(XY), dim3(IT, IT), 0, s>>>(in, kern, out, XY, kernel_size);
    return hipGetLastError();
}
