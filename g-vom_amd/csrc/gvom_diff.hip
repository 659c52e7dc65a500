// Differentiable form of the scan stage (no reference counterpart): a scalar map loss
// over one scan and its gradient with respect to the points and to a 6-DoF se(3)
// perturbation of the sensor-to-world transform.
//
//   L = sum over occupied voxels u of lambda_min(C_u)  +  w_h * Var_c(h_c)
//
// C_u is the voxel's neighbourhood population covariance in voxel units (the per-scan
// metrics of gvom.py:1250-1387), h_c the world z of the lowest point (least local z,
// first in sorted order on ties) in the lowest occupied voxel of column c, so that
// h_c = (min_h + z + oz) * z_res of gvom.py:572-576 in exact arithmetic, and the variance
// runs over columns holding an occupied voxel.  Voxel assignment, counts and the argmin
// points are those of the forward pass and are held fixed.
//
// Gradients (q = p / res component-wise, l = q - u, m_u the voxel mean, v_u the unit
// eigenvector of lambda_min):
//   d lambda_min(C_u) / d l_i = (2 / n_u) v_u (v_u . (l_i - m_u))
//   d Var / d h_c             = 2 (h_c - mean h) / K
//   T' = exp(xi^) T, p' = exp(xi^) p:  dL/d rho = sum g_i,  dL/d phi = sum p_i x g_i
// Every reduction is two-stage (per-block partials, then one block in block order), so
// results are deterministic.
#include <hip/hip_runtime.h>

#include "gvom_kernels.h"

namespace {

constexpr int RB = 256;   // threads per reduction block

struct DiffRec {
    double v[3];
    double m[3];
    double coef;   // 2 / n
};

__device__ __forceinline__ double block_sum(double x, double *sh) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) x += __shfl_xor(x, off);
    if (lane == 0) sh[w] = x;
    __syncthreads();
    double t = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); ++k) t += sh[k];
    __syncthreads();
    return t;   // valid in thread 0
}

// Smallest eigenvalue and its unit eigenvector of a symmetric 3x3 (f64).  Eigenvalue by
// the trigonometric closed form; eigenvector = the largest cross product of two rows of
// (C - lambda I).  A zero matrix (single point) gives lambda = 0 and v = e_z, whose
// gradient term vanishes because l - m = 0 there.
__device__ void min_eig(double xx, double xy, double xz, double yy, double yz, double zz, double &lam,
                        double v[3]) {
    const double PI = 3.141592653589793;
    const double p1 = xy * xy + xz * xz + yz * yz;
    const double q = (xx + yy + zz) / 3.0;
    if (p1 == 0.0) {
        lam = fmin(xx, fmin(yy, zz));
        v[0] = lam == xx ? 1.0 : 0.0;
        v[1] = (lam == yy && lam != xx) ? 1.0 : 0.0;
        v[2] = (v[0] == 0.0 && v[1] == 0.0) ? 1.0 : 0.0;
        return;
    }
    const double p2 = (xx - q) * (xx - q) + (yy - q) * (yy - q) + (zz - q) * (zz - q) + 2.0 * p1;
    const double p = sqrt(p2 / 6.0);
    const double b0 = (xx - q) / p, b1 = xy / p, b2 = xz / p, b3 = (yy - q) / p, b4 = yz / p, b5 = (zz - q) / p;
    double r = (b0 * (b3 * b5 - b4 * b4) - b1 * (b1 * b5 - b4 * b2) + b2 * (b1 * b4 - b3 * b2)) / 2.0;
    r = fmin(1.0, fmax(-1.0, r));
    const double phi = acos(r) / 3.0;
    lam = q + 2.0 * p * cos(phi + 2.0 * PI / 3.0);
    const double r0[3] = {xx - lam, xy, xz}, r1[3] = {xy, yy - lam, yz}, r2[3] = {xz, yz, zz - lam};
    const double c01[3] = {r0[1] * r1[2] - r0[2] * r1[1], r0[2] * r1[0] - r0[0] * r1[2], r0[0] * r1[1] - r0[1] * r1[0]};
    const double c02[3] = {r0[1] * r2[2] - r0[2] * r2[1], r0[2] * r2[0] - r0[0] * r2[2], r0[0] * r2[1] - r0[1] * r2[0]};
    const double c12[3] = {r1[1] * r2[2] - r1[2] * r2[1], r1[2] * r2[0] - r1[0] * r2[2], r1[0] * r2[1] - r1[1] * r2[0]};
    const double n01 = c01[0] * c01[0] + c01[1] * c01[1] + c01[2] * c01[2];
    const double n02 = c02[0] * c02[0] + c02[1] * c02[1] + c02[2] * c02[2];
    const double n12 = c12[0] * c12[0] + c12[1] * c12[1] + c12[2] * c12[2];
    const double *c = c01;
    double nn = n01;
    if (n02 > nn) { c = c02; nn = n02; }
    if (n12 > nn) { c = c12; nn = n12; }
    // largest row of A - lambda I
    const double q0 = r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2];
    const double q1 = r1[0] * r1[0] + r1[1] * r1[1] + r1[2] * r1[2];
    const double q2 = r2[0] * r2[0] + r2[1] * r2[1] + r2[2] * r2[2];
    const double *rm = r0;
    double qm = q0;
    if (q1 > qm) { rm = r1; qm = q1; }
    if (q2 > qm) { rm = r2; qm = q2; }
    if (nn > 1e-20 * qm * qm) {
        // simple root: A - lambda I has rank 2 and the cross product of two rows spans its
        // null space
        const double inv = 1.0 / sqrt(nn);
        v[0] = c[0] * inv; v[1] = c[1] * inv; v[2] = c[2] * inv;
    } else if (qm > 0.0) {
        // double root (lambda_min = lambda_mid, e.g. collinear neighbourhood points): the
        // rows span one direction u and every unit vector orthogonal to u is an eigenvector
        // of lambda_min.  With collinear points every (p - mean) is parallel to u, so any
        // such choice gives lambda_min the zero gradient it has along its smooth branch.
        const double iu = 1.0 / sqrt(qm);
        const double u0 = rm[0] * iu, u1 = rm[1] * iu, u2 = rm[2] * iu;
        // cross u with the axis it is least aligned with
        const double a0 = fabs(u0), a1 = fabs(u1), a2 = fabs(u2);
        double w0, w1, w2;
        if (a0 <= a1 && a0 <= a2) { w0 = 0.0; w1 = u2; w2 = -u1; }          // u x e0
        else if (a1 <= a2) { w0 = -u2; w1 = 0.0; w2 = u0; }                 // u x e1
        else { w0 = u1; w1 = -u0; w2 = 0.0; }                               // u x e2
        const double iw = 1.0 / sqrt(w0 * w0 + w1 * w1 + w2 * w2);
        v[0] = w0 * iw; v[1] = w1 * iw; v[2] = w2 * iw;
    } else {
        // A = lambda I: every direction is an eigenvector
        v[0] = 0.0; v[1] = 0.0; v[2] = 1.0;
    }
}

// Per occupied voxel of the slot: lambda_min, eigenvector, mean, 2/n; block partial of
// sum lambda_min.
__global__ __launch_bounds__(RB) void k_diff_eigen(const DevState *__restrict__ st, int slot, const SlotRec *__restrict__ recs,
                                                   int64_t ncap, DiffRec *__restrict__ dr, double *__restrict__ part) {
    __shared__ double sh[RB / 64];
    const int64_t id = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t C = st->slot_cells[slot];
    recs += (int64_t)cur_buf(st->flip, slot) * ncap;
    double lam = 0.0;
    if (id < C) {
        const SlotRec &r = recs[id];
        double v[3];
        min_eig(r.m[3], r.m[4], r.m[5], r.m[6], r.m[7], r.m[8], lam, v);
        DiffRec d;
        d.v[0] = v[0]; d.v[1] = v[1]; d.v[2] = v[2];
        d.m[0] = r.m[0]; d.m[1] = r.m[1]; d.m[2] = r.m[2];
        d.coef = 2.0 / r.m[9];
        dr[id] = d;
    }
    const double t = block_sum(lam, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Per column, one wavefront (lane = z within a 64-voxel chunk): the lowest occupied voxel
// by ballot, then its points' argmin of world z over the lanes (ties: first in sorted
// order), h_c; per-block partials of (K, sum h) over the block's CW columns in order.
constexpr int CW = RB / 64;   // columns per block

__global__ __launch_bounds__(RB) void k_diff_columns(Geo g, const DevState *__restrict__ st, int slot,
                                                     const int32_t *__restrict__ table,
                                                     const int32_t *__restrict__ seg_start,
                                                     const uint32_t *__restrict__ svals,
                                                     const double *__restrict__ world, double *__restrict__ colh,
                                                     int32_t *__restrict__ colpt, int32_t *__restrict__ pmark,
                                                     double *__restrict__ partK, double *__restrict__ partH) {
    __shared__ double shK[CW], shH[CW];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int64_t c = (int64_t)blockIdx.x * CW + w;
    double k = 0.0, hsum = 0.0;
    if (c < g.X) {
        const int32_t *col = table + (int64_t)cur_buf(st->flip, slot) * g.V + c * g.Z;
        int zfound = -1;
        int32_t tfound = 0;
        for (int z0 = 0; z0 < g.Z && zfound < 0; z0 += 64) {
            const int z = z0 + lane;
            const int32_t t = z < g.Z ? col[z] : 0;
            const uint64_t m = __ballot(t < 0);
            if (m) {
                const int l = __builtin_ctzll(m);
                zfound = z0 + l;
                tfound = __shfl(t, l);
            }
        }
        int32_t pt = -1;
        double h = 0.0;
        if (zfound >= 0) {
            const int64_t seg = st->first_seg[0] + (int64_t)(tfound & 0x7FFFFFFF);
            const int32_t a = seg_start[seg], b = seg_start[seg + 1];
            double best = 0.0;
            int32_t bj = -1;
            for (int32_t j = a + lane; j < b; j += 64) {
                const double wz = world[3 * (int64_t)svals[j] + 2];
                if (bj < 0 || wz < best) {
                    best = wz;
                    bj = j;
                }
            }
            for (int off = 32; off > 0; off >>= 1) {
                const double ob = __shfl_xor(best, off);
                const int32_t oj = __shfl_xor(bj, off);
                if (oj >= 0 && (bj < 0 || ob < best || (ob == best && oj < bj))) {
                    best = ob;
                    bj = oj;
                }
            }
            if (bj >= 0) {
                pt = (int32_t)svals[bj];
                h = best;
            }
        }
        if (lane == 0) {
            colpt[c] = pt;
            colh[c] = h;
            if (pt >= 0) {
                pmark[pt] = (int32_t)c;
                k = 1.0;
                hsum = h;
            }
        }
    }
    if (lane == 0) {
        shK[w] = k;
        shH[w] = hsum;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double tk = 0.0, th = 0.0;
        for (int q = 0; q < CW; ++q) {
            tk += shK[q];
            th += shH[q];
        }
        partK[blockIdx.x] = tk;
        partH[blockIdx.x] = th;
    }
}

// Ordered sum of n partials in one block: each thread sums a contiguous run, then the block
// combines in thread order (valid in thread 0).
__device__ __forceinline__ double ordered_sum(const double *__restrict__ part, int64_t n, double *sh) {
    const int64_t per = (n + blockDim.x - 1) / blockDim.x;
    double x = 0.0;
    for (int64_t i = threadIdx.x * per; i < (threadIdx.x + 1) * per && i < n; ++i) x += part[i];
    return block_sum(x, sh);
}

// Up to 6 independent ordered sums, one block each.
struct SumJobs {
    const double *part[6];
    int64_t n[6];
    double *out[6];
};

__global__ __launch_bounds__(RB) void k_diff_sums(SumJobs jobs) {
    __shared__ double sh[RB / 64];
    const int k = blockIdx.x;
    const double t = ordered_sum(jobs.part[k], jobs.n[k], sh);
    if (threadIdx.x == 0) *jobs.out[k] = t;
}

// Per column: (h_c - mean)^2 partials.
__global__ __launch_bounds__(RB) void k_diff_colvar(Geo g, const double *__restrict__ colh,
                                                    const int32_t *__restrict__ colpt, const double *__restrict__ sums,
                                                    double *__restrict__ part) {
    __shared__ double sh[RB / 64];
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const double K = sums[1], mean = K > 0 ? sums[2] / K : 0.0;
    double x = 0.0;
    if (c < g.X && colpt[c] >= 0) {
        const double d = colh[c] - mean;
        x = d * d;
    }
    const double t = block_sum(x, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// Sum of the variance partials, then the final scalars:
// sums = [sum lambda, K, sum h, sum (h - mean)^2, L, L_rough, L_height]
__global__ __launch_bounds__(RB) void k_diff_loss(const double *__restrict__ part, int64_t n,
                                                  double *__restrict__ sums, double wh) {
    __shared__ double sh[RB / 64];
    const double t = ordered_sum(part, n, sh);
    if (threadIdx.x != 0) return;
    sums[3] = t;
    const double K = sums[1];
    const double var = K > 0 ? sums[3] / K : 0.0;
    sums[5] = sums[0];
    sums[6] = wh * var;
    sums[4] = sums[5] + sums[6];
}

// Per point: dL/dp (world) from every occupied voxel whose neighbourhood holds the point,
// plus the height term for argmin points; the sensor-frame gradient R^T g; block partials
// of sum g and sum p x g.
__global__ __launch_bounds__(RB) void k_diff_backward(Geo g, const DevState *__restrict__ st, int slot,
                                                      const ScanDesc *__restrict__ sd, const double *__restrict__ world,
                                                      const uint8_t *__restrict__ valid,
                                                      const int32_t *__restrict__ table,
                                                      const DiffRec *__restrict__ dr, const int32_t *__restrict__ pmark,
                                                      const double *__restrict__ colh, const double *__restrict__ sums,
                                                      double wh, double dL_arg, const double *__restrict__ dL_dev,
                                                      int64_t n, double *__restrict__ gsensor,
                                                      double *__restrict__ gworld, double *__restrict__ part6) {
    __shared__ double sh6[6][RB / 64];
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double gx = 0.0, gy = 0.0, gz = 0.0, px = 0.0, py = 0.0, pz = 0.0;
    if (i < n && valid[i]) {
        px = world[3 * i];
        py = world[3 * i + 1];
        pz = world[3 * i + 2];
        const ScanDesc &d = sd[0];
        const double o0 = floor((d.ego[0] / g.xy_res) - (double)g.XY / 2.0);
        const double o1 = floor((d.ego[1] / g.xy_res) - (double)g.XY / 2.0);
        const double o2 = floor((d.ego[2] / g.z_res) - (double)g.Z / 2.0);
        const double q0 = px / g.xy_res - o0, q1 = py / g.xy_res - o1, q2 = pz / g.z_res - o2;
        const double f0 = floor(q0), f1 = floor(q1), f2 = floor(q2);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;   // gradient in q (voxel units)
        if (f0 >= -g.ex - 1 && f0 <= g.XY + g.ex && f1 >= -g.ex - 1 && f1 <= g.XY + g.ex && f2 >= -g.ez - 1 &&
            f2 <= g.Z + g.ez) {
            const int64_t b0 = (int64_t)f0, b1 = (int64_t)f1, b2 = (int64_t)f2;
            const int32_t *tab = table + (int64_t)cur_buf(st->flip, slot) * g.V;
            if (g.ex == 1 && g.ez == 1) {
                // the 27 table words first, then the records of three neighbours at a time, so
                // the gathers overlap; terms are added in the loop order of the general case
                int32_t t[27];
#pragma unroll
                for (int k = 0; k < 27; ++k) {
                    const int64_t ux = b0 + k / 9 - 1, uy = b1 + (k / 3) % 3 - 1, uz = b2 + k % 3 - 1;
                    const bool in = ux >= 0 && ux < g.XY && uy >= 0 && uy < g.XY && uz >= 0 && uz < g.Z;
                    t[k] = in ? tab[(ux * g.XY + uy) * g.Z + uz] : 0;
                }
#pragma unroll
                for (int k0 = 0; k0 < 27; k0 += 3) {
                    DiffRec r[3];
#pragma unroll
                    for (int k = 0; k < 3; ++k) r[k] = dr[t[k0 + k] < 0 ? (t[k0 + k] & 0x7FFFFFFF) : 0];
#pragma unroll
                    for (int k = 0; k < 3; ++k) {
                        if (t[k0 + k] >= 0) continue;
                        const int kk = k0 + k;
                        const double l0 = q0 - (double)(b0 + kk / 9 - 1) - r[k].m[0],
                                     l1 = q1 - (double)(b1 + (kk / 3) % 3 - 1) - r[k].m[1],
                                     l2 = q2 - (double)(b2 + kk % 3 - 1) - r[k].m[2];
                        const double s = r[k].coef * (r[k].v[0] * l0 + r[k].v[1] * l1 + r[k].v[2] * l2);
                        a0 += s * r[k].v[0];
                        a1 += s * r[k].v[1];
                        a2 += s * r[k].v[2];
                    }
                }
            } else
            for (int64_t ux = b0 - g.ex; ux <= b0 + g.ex; ++ux) {
                if (ux < 0 || ux >= g.XY) continue;
                for (int64_t uy = b1 - g.ex; uy <= b1 + g.ex; ++uy) {
                    if (uy < 0 || uy >= g.XY) continue;
                    for (int64_t uz = b2 - g.ez; uz <= b2 + g.ez; ++uz) {
                        if (uz < 0 || uz >= g.Z) continue;
                        const int32_t t = tab[(ux * g.XY + uy) * g.Z + uz];
                        if (t >= 0) continue;
                        const DiffRec &r = dr[t & 0x7FFFFFFF];
                        const double l0 = q0 - (double)ux - r.m[0], l1 = q1 - (double)uy - r.m[1],
                                     l2 = q2 - (double)uz - r.m[2];
                        const double s = r.coef * (r.v[0] * l0 + r.v[1] * l1 + r.v[2] * l2);
                        a0 += s * r.v[0];
                        a1 += s * r.v[1];
                        a2 += s * r.v[2];
                    }
                }
            }
        }
        gx = a0 / g.xy_res;
        gy = a1 / g.xy_res;
        gz = a2 / g.z_res;
        const int32_t c = pmark[i];
        if (c >= 0) {
            const double K = sums[1], mean = sums[2] / K;
            gz += wh * 2.0 * (colh[c] - mean) / K;
        }
        const double dL = dL_dev ? *dL_dev : dL_arg;
        gx *= dL;
        gy *= dL;
        gz *= dL;
        if (gworld) {
            gworld[3 * i] = gx;
            gworld[3 * i + 1] = gy;
            gworld[3 * i + 2] = gz;
        }
        if (gsensor) {
            const double *T = d.T;   // rows of [R | t]
            gsensor[3 * i] = T[0] * gx + T[4] * gy + T[8] * gz;
            gsensor[3 * i + 1] = T[1] * gx + T[5] * gy + T[9] * gz;
            gsensor[3 * i + 2] = T[2] * gx + T[6] * gy + T[10] * gz;
        }
    } else if (i < n) {
        if (gworld) gworld[3 * i] = gworld[3 * i + 1] = gworld[3 * i + 2] = 0.0;
        if (gsensor) gsensor[3 * i] = gsensor[3 * i + 1] = gsensor[3 * i + 2] = 0.0;
    }
    // the six block partials in one pass (same order of additions as block_sum)
    double vals[6] = {gx, gy, gz, py * gz - pz * gy, pz * gx - px * gz, px * gy - py * gx};
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 6; ++k)
        for (int off = 32; off > 0; off >>= 1) vals[k] += __shfl_xor(vals[k], off);
    if (lane == 0)
#pragma unroll
        for (int k = 0; k < 6; ++k) sh6[k][w] = vals[k];
    __syncthreads();
    if (threadIdx.x < 6) {
        double t = 0.0;
        for (int q = 0; q < (int)(blockDim.x >> 6); ++q) t += sh6[threadIdx.x][q];
        part6[(int64_t)threadIdx.x * gridDim.x + blockIdx.x] = t;
    }
}

inline unsigned nb(int64_t n, int per) { return (unsigned)((n + per - 1) / per > 0 ? (n + per - 1) / per : 1); }

}  // namespace

int64_t diff_partials(int64_t n_points, int64_t X, int64_t ncap) {
    int64_t a = nb(n_points, RB), b = nb(X, CW), c = nb(ncap, RB);
    int64_t m = a > b ? a : b;
    m = m > c ? m : c;
    return 6 * m + 16;
}

int64_t diff_rec_bytes() { return (int64_t)sizeof(DiffRec); }

// sums: 8 doubles on the device; part: diff_partials() doubles.
hipError_t launch_diff_forward(const Geo &g, const DevState *st, int slot, const int32_t *table,
                               const SlotRec *recs, int64_t ncap, const int32_t *seg_start, const uint32_t *svals,
                               const double *world, int64_t n, void *dr, double *colh, int32_t *colpt,
                               int32_t *pmark, double *part, double *sums, double wh, hipStream_t s) {
    // partials: [0, m) eigenvalues, then [m, 2m) and [2m, 3m) the column counts and heights;
    // the variance partials reuse [0, m) once the first three sums are taken
    const unsigned ne = nb(ncap, RB), ncw = nb(g.X, CW), nc = nb(g.X, RB);
    const int64_t m = (diff_partials(n > 0 ? n : 1, g.X, ncap) - 16) / 6;
    hipError_t e = launch_fill_i32(pmark, -1, n > 0 ? n : 1, s);
    if (e != hipSuccess) return e;
    k_diff_eigen<<<ne, RB, 0, s>>>(st, slot, recs, ncap, (DiffRec *)dr, part);
    k_diff_columns<<<ncw, RB, 0, s>>>(g, st, slot, table, seg_start, svals, world, colh, colpt, pmark, part + m,
                                      part + 2 * m);
    SumJobs j3{};
    j3.part[0] = part;
    j3.n[0] = ne;
    j3.out[0] = sums + 0;
    j3.part[1] = part + m;
    j3.n[1] = ncw;
    j3.out[1] = sums + 1;
    j3.part[2] = part + 2 * m;
    j3.n[2] = ncw;
    j3.out[2] = sums + 2;
    k_diff_sums<<<3, RB, 0, s>>>(j3);
    k_diff_colvar<<<nc, RB, 0, s>>>(g, colh, colpt, sums, part);
    k_diff_loss<<<1, RB, 0, s>>>(part, nc, sums, wh);
    return hipGetLastError();
}

// grad6: 6 doubles on the device (d/d rho, d/d phi).
hipError_t launch_diff_backward(const Geo &g, const DevState *st, int slot, const ScanDesc *sd,
                                const double *world, const uint8_t *valid, const int32_t *table, const void *dr,
                                const int32_t *pmark, const double *colh, const double *sums, double wh, double dL,
                                const double *dL_dev, int64_t n, double *gsensor, double *gworld, double *part,
                                double *grad6, hipStream_t s) {
    const unsigned np = nb(n, RB);
    k_diff_backward<<<np, RB, 0, s>>>(g, st, slot, sd, world, valid, table, (const DiffRec *)dr, pmark, colh, sums,
                                      wh, dL, dL_dev, n, gsensor, gworld, part);
    SumJobs j6{};
    for (int k = 0; k < 6; ++k) {
        j6.part[k] = part + (int64_t)k * np;
        j6.n[k] = np;
        j6.out[k] = grad6 + k;
    }
    k_diff_sums<<<6, RB, 0, s>>>(j6);
    return hipGetLastError();
}
