// gocean_iewpf.hip -- the coarse-grid kernels of the two-stage implicit
// equal-weights particle filter, batched over the members of an ensemble.
//
// The reference (IEWPFOcean.iewpf_2stage, gpu_ocean/SWESimulators/
// IEWPFOcean.py:195-329, kernels in gpu_kernels/iewpf_kernels.cu and
// ocean_noise.cu) launches each of these once per particle.  Here member m's
// coarse random field is n = coarse_nx * coarse_ny floats at offsets[m] of a
// flat buffer (the batched model-error layout, rows of coarse_nx), and one
// launch covers every member:
//
//   k_iewpf_dots            |xi|^2, |nu|^2, xi . nu per member (float64)
//   k_iewpf_perpendicular   nu <- |nu| / |nu_perp| (nu - (xi.nu / |xi|^2) xi)
//   k_iewpf_axpby           xi <- a_m xi + b nu
//   k_iewpf_local_svd       for each drifter in order, the 7 x 7 coarse
//                           window around it <- svd (49 x 49) x window
//   k_iewpf_kalman_patch    one drifter's Kalman-gain patch e_x P_x + e_y P_y
//                           of every member written into its zeroed field
//   k_iewpf_kalman_window   the Kalman-gain term of every drifter added to
//                           every member's fine-grid state, from windowed
//                           basis fields (interpolated model error)

#include <algorithm>

#include "gocean_internal.hpp"

namespace gocean {
namespace {

constexpr int kDotNT = 256;
constexpr int kElemNT = 256;
constexpr int kSvdNT = 64;

__device__ __forceinline__ double wave_sum_d(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

__global__ __launch_bounds__(kDotNT) void k_iewpf_dots(const float* xi, const float* nu, const int64_t* offsets,
                                                      int n, double* out) {
    __shared__ double red[kDotNT / 64][3];
    const int m = blockIdx.x;
    const float* x = xi + offsets[m];
    const float* y = nu + offsets[m];
    double s[3] = {0.0, 0.0, 0.0};
    for (int i = threadIdx.x; i < n; i += kDotNT) {
        const double a = x[i], b = y[i];
        s[0] += a * a;
        s[1] += b * b;
        s[2] += a * b;
    }
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < 3; ++k) s[k] = wave_sum_d(s[k]);
    if (lane == 0)
        for (int k = 0; k < 3; ++k) red[w][k] = s[k];
    __syncthreads();
    if (threadIdx.x < 3) {
        double t = 0.0;
        for (int i = 0; i < kDotNT / 64; ++i) t += red[i][threadIdx.x];
        out[3 * m + threadIdx.x] = t;
    }
}

// makePerpendicular (ocean_noise.cu:206-243): remove xi's component from nu
// and rescale to nu's original norm.
__global__ __launch_bounds__(kElemNT) void k_iewpf_perpendicular(const float* xi, float* nu, const int64_t* offsets,
                                                                int n, const double* dots) {
    const int m = blockIdx.y;
    const int i = blockIdx.x * kElemNT + threadIdx.x;
    if (i >= n) return;
    const double xx = dots[3 * m], nn = dots[3 * m + 1], xn = dots[3 * m + 2];
    const double par = xn / xx;
    const double perp = nn - par * xn;
    const double scale = sqrt(nn / perp);
    const long long o = offsets[m] + i;
    nu[o] = (float)(scale * ((double)nu[o] - par * (double)xi[o]));
}

// blas_xaxpby (iewpf_kernels.cu:47-66) with a per-member a
__global__ __launch_bounds__(kElemNT) void k_iewpf_axpby(float* xi, const float* nu, const int64_t* offsets, int n,
                                                        const float* a, float b) {
    const int m = blockIdx.y;
    const int i = blockIdx.x * kElemNT + threadIdx.x;
    if (i >= n) return;
    const long long o = offsets[m] + i;
    xi[o] = a[m] * xi[o] + b * nu[o];
}

// localSVDOnGlobalXi (iewpf_kernels.cu:213-258) for every drifter of one
// member in order (windows of nearby drifters overlap, so the order matters).
__global__ __launch_bounds__(kSvdNT) void k_iewpf_local_svd(float* xi, const int64_t* offsets, int cnx, int cny,
                                                           const float* svd, const int32_t* cells, int n_drifters) {
    __shared__ float s_svd[49][50];
    __shared__ float s_win[49];
    const int m = blockIdx.x;
    const int t = threadIdx.x;
    float* x = xi + offsets[m];
    for (int k = t; k < 49 * 49; k += kSvdNT) s_svd[k / 49][k % 49] = svd[k];
    int gx = 0, gy = 0;
    for (int d = 0; d < n_drifters; ++d) {
        const int cx = cells[2 * d], cy = cells[2 * d + 1];
        if (cx < 0 || cx >= cnx || cy < 0 || cy >= cny) continue;   // uniform over the workgroup
        if (t < 49) {
            gx = cx - 3 + t % 7;
            gy = cy - 3 + t / 7;
            gx += gx < 0 ? cnx : 0;
            gx -= gx >= cnx ? cnx : 0;
            gy += gy < 0 ? cny : 0;
            gy -= gy >= cny ? cny : 0;
        }
        __syncthreads();                                // previous drifter's window written
        if (t < 49) s_win[t] = x[(long long)gy * cnx + gx];
        __syncthreads();
        if (t < 49) {
            float acc = 0.0f;
            for (int k = 0; k < 49; ++k) acc += s_svd[t][k] * s_win[k];
            x[(long long)gy * cnx + gx] = acc;
        }
    }
}

// setBufferToZero + halfTheKalmanGain (iewpf_kernels.cu:34-45, 115-176):
// member m's whole field is zero except the 7 x 7 window centred on coarse
// cell (cx, cy) (periodic), which holds e_x P_x + e_y P_y with e = e[m][d].
__global__ __launch_bounds__(kElemNT) void k_iewpf_kalman_patch(float* xi, const int64_t* offsets, int cnx, int cny,
                                                               const float* patch, const float* e, int n_drifters,
                                                               int d, int cx, int cy) {
    const int m = blockIdx.y;
    const int i = blockIdx.x * kElemNT + threadIdx.x;
    if (i >= cnx * cny) return;
    const int y = i / cnx, x = i - y * cnx;
    int rx = x - cx + 3, ry = y - cy + 3;
    rx += rx < 0 ? cnx : 0;
    rx -= rx >= cnx ? cnx : 0;
    ry += ry < 0 ? cny : 0;
    ry -= ry >= cny ? cny : 0;
    float v = 0.0f;
    if (rx < 7 && ry < 7) {
        const float ex = e[2 * ((long long)m * n_drifters + d)], ey = e[2 * ((long long)m * n_drifters + d) + 1];
        v = ex * patch[ry * 7 + rx] + ey * patch[49 + ry * 7 + rx];
    }
    xi[offsets[m] + i] = v;
}

// addKalmanGain with interpolated model error (IEWPFOcean.py:548-595): the
// term of drifter d is sum_k coef[d][k][m] basis[d][k] with five fields per
// drifter, non-zero on a window of win_h x win_w interior cells whose first
// cell is (origins[2d], origins[2d+1]) (row, column; periodic).  A gather:
// one workgroup per (256-cell segment of an interior row, field, 16
// members), thread = cell, so each member's row is one contiguous run.  tiles[2t], tiles[2t+1] are the
// segment's row and index; list[list_off[t] .. list_off[t+1]) the drifters
// whose window meets it, ascending.  Each cell adds its sum over them once,
// so the result does not depend on the launch order.
constexpr int kWinSeg = 256;
constexpr int kWinNT = 256;
constexpr int kWinPG = 16;    // consecutive members per workgroup

__global__ __launch_bounds__(kWinNT) void k_iewpf_kalman_window(float* state, int n_members, int nx, int ny,
                                                                long long member_stride, const float* basis,
                                                                int n_drifters, int wh, int ww, const int* origins,
                                                                const float* coef, int coef_stride,
                                                                const int* tiles, const int* list_off,
                                                                const int* list) {
    const int t = blockIdx.x, f = blockIdx.y;
    const int row = tiles[2 * t], seg = tiles[2 * t + 1];
    if (row < 0 || row >= ny || seg < 0) return;
    const int c = seg * kWinSeg + (int)threadIdx.x;
    const bool in = c < nx;
    const int pitch = nx + 4;
    const long long plane = (long long)(ny + 4) * pitch;
    const long long win = (long long)wh * ww;
    float* x = state + f * plane + (long long)(row + 2) * pitch + (c + 2);
    const int q0 = list_off[t], q1 = list_off[t + 1];
    const int mw = blockIdx.z * kWinPG;     // the workgroup's first member
    // the state is read first, so that its latency overlaps the sums
    float acc[kWinPG], xv[kWinPG];
#pragma unroll
    for (int g = 0; g < kWinPG; ++g) {
        const int m = mw + g;
        acc[g] = 0.0f;
        xv[g] = (in && m < n_members) ? x[m * member_stride] : 0.0f;
    }
    for (int q = q0; q < q1; ++q) {
        const int d = list[q];
        if ((unsigned)d >= (unsigned)n_drifters) continue;
        int i = row - origins[2 * d], j = c - origins[2 * d + 1];
        i += i < 0 ? ny : 0;
        j += j < 0 ? nx : 0;
        float b[5] = {0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
        if (in && (unsigned)i < (unsigned)wh && (unsigned)j < (unsigned)ww) {
            const float* bp = basis + ((long long)d * 15 + f) * win + (long long)i * ww + j;
#pragma unroll
            for (int k = 0; k < 5; ++k) b[k] = bp[k * 3 * win];
        }
        // the workgroup's kWinPG consecutive members: one contiguous run per k
        const float* e = coef + (long long)d * 5 * coef_stride + mw;
#pragma unroll
        for (int g = 0; g < kWinPG; ++g)
            acc[g] += (((e[g] * b[0] + e[coef_stride + g] * b[1]) + e[2 * coef_stride + g] * b[2])
                       + e[3 * coef_stride + g] * b[3]) + e[4 * coef_stride + g] * b[4];
    }
    if (in) {
#pragma unroll
        for (int g = 0; g < kWinPG; ++g) {
            const int m = mw + g;
            if (m < n_members) x[m * member_stride] = xv[g] + acc[g];
        }
    }
}

int check_members(const int64_t* offsets, int n_members, int n) {
    if (!offsets || n_members < 1 || n < 1) return set_error(GOCEAN_EINVAL, "offsets NULL or %d members of %d", n_members, n);
    return 0;
}

}  // namespace
}  // namespace gocean

using namespace gocean;

extern "C" {

int gocean_iewpf_dots(const float* xi, const float* nu, const int64_t* offsets, int32_t n_members, int32_t n,
                      double* out, void* stream) {
    clear_error();
    if (int e = check_members(offsets, n_members, n)) return e;
    if (!xi || !nu || !out) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    hipLaunchKernelGGL(k_iewpf_dots, dim3(n_members), dim3(kDotNT), 0, (hipStream_t)stream, xi, nu, offsets, n, out);
    return check_launch("k_iewpf_dots");
}

int gocean_iewpf_make_perpendicular(const float* xi, float* nu, const int64_t* offsets, int32_t n_members,
                                    int32_t n, const double* dots, void* stream) {
    clear_error();
    if (int e = check_members(offsets, n_members, n)) return e;
    if (!xi || !nu || !dots) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    const dim3 grid((n + kElemNT - 1) / kElemNT, n_members);
    hipLaunchKernelGGL(k_iewpf_perpendicular, grid, dim3(kElemNT), 0, (hipStream_t)stream, xi, nu, offsets, n, dots);
    return check_launch("k_iewpf_perpendicular");
}

int gocean_iewpf_axpby(float* xi, const float* nu, const int64_t* offsets, int32_t n_members, int32_t n,
                       const float* a, float b, void* stream) {
    clear_error();
    if (int e = check_members(offsets, n_members, n)) return e;
    if (!xi || !nu || !a) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    const dim3 grid((n + kElemNT - 1) / kElemNT, n_members);
    hipLaunchKernelGGL(k_iewpf_axpby, grid, dim3(kElemNT), 0, (hipStream_t)stream, xi, nu, offsets, n, a, b);
    return check_launch("k_iewpf_axpby");
}

int gocean_iewpf_local_svd(float* xi, const int64_t* offsets, int32_t n_members, int32_t coarse_nx,
                           int32_t coarse_ny, const float* svd, const int32_t* cells, int32_t n_drifters,
                           void* stream) {
    clear_error();
    if (int e = check_members(offsets, n_members, coarse_nx * coarse_ny)) return e;
    if (coarse_nx < 7 || coarse_ny < 7) return set_error(GOCEAN_EINVAL, "coarse grid %d x %d smaller than 7 x 7", coarse_nx, coarse_ny);
    if (n_drifters < 0) return set_error(GOCEAN_EINVAL, "drifters %d", n_drifters);
    if (!xi || !svd || (n_drifters > 0 && !cells)) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    if (n_drifters == 0) return 0;
    hipLaunchKernelGGL(k_iewpf_local_svd, dim3(n_members), dim3(kSvdNT), 0, (hipStream_t)stream, xi, offsets,
                       coarse_nx, coarse_ny, svd, cells, n_drifters);
    return check_launch("k_iewpf_local_svd");
}

int gocean_iewpf_kalman_patch(float* xi, const int64_t* offsets, int32_t n_members, int32_t coarse_nx,
                              int32_t coarse_ny, const float* patch, const float* e, int32_t n_drifters,
                              int32_t drifter, int32_t cx, int32_t cy, void* stream) {
    clear_error();
    if (int e_ = check_members(offsets, n_members, coarse_nx * coarse_ny)) return e_;
    if (coarse_nx < 7 || coarse_ny < 7) return set_error(GOCEAN_EINVAL, "coarse grid %d x %d smaller than 7 x 7", coarse_nx, coarse_ny);
    if (drifter < 0 || drifter >= n_drifters) return set_error(GOCEAN_EINVAL, "drifter %d of %d", drifter, n_drifters);
    if (cx < 0 || cx >= coarse_nx || cy < 0 || cy >= coarse_ny)
        return set_error(GOCEAN_EINVAL, "coarse cell (%d, %d) outside %d x %d", cx, cy, coarse_nx, coarse_ny);
    if (!xi || !patch || !e) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    const dim3 grid((coarse_nx * coarse_ny + kElemNT - 1) / kElemNT, n_members);
    hipLaunchKernelGGL(k_iewpf_kalman_patch, grid, dim3(kElemNT), 0, (hipStream_t)stream, xi, offsets, coarse_nx,
                       coarse_ny, patch, e, n_drifters, drifter, cx, cy);
    return check_launch("k_iewpf_kalman_patch");
}

int gocean_iewpf_kalman_window(float* state, int32_t n_members, int32_t nx, int32_t ny, int64_t member_stride,
                               const float* basis, int32_t n_drifters, int32_t win_h, int32_t win_w,
                               const int32_t* origins, const float* coef, int32_t coef_stride,
                               const int32_t* tiles, int32_t n_tiles, const int32_t* list_off, const int32_t* list,
                               void* stream) {
    clear_error();
    if (n_members < 1 || nx < 1 || ny < 1) return set_error(GOCEAN_EINVAL, "%d members of %d x %d", n_members, nx, ny);
    if (member_stride < 3LL * (nx + 4) * (ny + 4))
        return set_error(GOCEAN_EINVAL, "member stride %lld below three planes", (long long)member_stride);
    if (n_drifters < 0 || n_tiles < 0) return set_error(GOCEAN_EINVAL, "drifters %d, tiles %d", n_drifters, n_tiles);
    if (win_h < 1 || win_h > ny || win_w < 1 || win_w > nx)
        return set_error(GOCEAN_EINVAL, "window %d x %d outside 1..%d x 1..%d", win_h, win_w, ny, nx);
    if (n_tiles > 0 && n_drifters == 0) return set_error(GOCEAN_EINVAL, "tiles without drifters");
    if (coef_stride < (n_members + kWinPG - 1) / kWinPG * kWinPG)
        return set_error(GOCEAN_EINVAL, "coefficient stride %d below %d members rounded up to %d", coef_stride,
                         n_members, kWinPG);
    if (n_tiles == 0) return 0;
    if (!state || !basis || !origins || !coef || !tiles || !list_off || !list)
        return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    hipLaunchKernelGGL(k_iewpf_kalman_window, dim3(n_tiles, 3, (n_members + kWinPG - 1) / kWinPG),
                       dim3(kWinNT), 0, (hipStream_t)stream, state,
                       n_members, nx, ny, (long long)member_stride, basis, n_drifters, win_h, win_w, origins, coef,
                       coef_stride, tiles, list_off, list);
    return check_launch("k_iewpf_kalman_window");
}

}  // extern "C"
