// gocean_letkf.hip -- localised ETKF analysis on the device ensemble state.
//
// The reference (ETKFOcean.LETKF, gpu_ocean/SWESimulators/ETKFOcean.py:
// 583-744) downloads every member, loops over the observations of each group
// in Python, solves one N x N eigenproblem per observation with scipy, blends
// the local analyses into the forecast with Gaspari-Cohn weights and uploads
// every member again.  Here the state never leaves HBM:
//
//   k_letkf_coef   one workgroup per observation of a group: observes hu, hv
//                  of every member at the buoy cell (observe_particles_from_
//                  X_f, ETKFOcean.py:777-793) and reduces the local analysis
//                  to a rank-2 form (below); float64.
//   k_letkf_apply  one workgroup per (row of 64 cells the group's weights
//                  reach, field), its four waves splitting the members:
//                  blends the local analyses of every observation covering
//                  the cell into the forecast, for all members, in place
//                  (ETKFOcean.py:700-740).
//
// Rank-2 form.  One buoy observes two values, so with Y = M HX' / sqrt(N-1)
// (M^T M = R^-1, Y is 2 x N) the reference's A / (N-1) = ff I + Y^T Y has
// eigenvalues ff + s_j^2 on the two right singular vectors V_j of Y and ff
// on their complement.  Its inverse square root is a I + V diag(c) V^T with
// a = ff^-1/2, c_j = (ff + s_j^2)^-1/2 - a, and the gain applied to D is
// X' V z with z_j = s_j / (ff + s_j^2) * u_j . (M D) / sqrt(N-1).  So the
// local analysis of member e at a cell is
//
//   mean + a (x_e - mean) + sum_j p_j B_ej,   p_j = sum_k (x_k - mean) V_kj,
//   B_ej = c_j V_ej + z_j,
//
// the same numbers as the reference's solve + eigh on N x N matrices, from
// two dot products per cell and field.  ff = 1 for eta; hu, hv use the
// inflated ff when inflation is on (ETKFOcean.py:667-675, 706-720).

#include <algorithm>
#include <cmath>

#include "gocean_internal.hpp"

namespace gocean {
namespace {

constexpr int kCoefNT = 256;
constexpr int kCoefWaves = kCoefNT / 64;
constexpr int kMaxMembers = 1024;
constexpr int kTileW = 64;
constexpr int kTileH = 1;
constexpr int kChunks = 4;        // waves per apply workgroup, one member range each
constexpr int kRegMembers = 32;   // members a wave keeps in registers
constexpr int kMaxCover = 4;
constexpr int kCoefPerMember = 6;   // V0, V1, B_eta0, B_eta1, B_uv0, B_uv1

struct CoefArgs {
    const float* state;
    long long plane, member_stride;
    int pitch, nx, ny, n_members, n_obs;
    const int32_t* group;
    const int32_t* cells;
    const double* y;
    double r00, r01, r11;   // R^-1
    double m00, m01, m11;   // M upper triangular, M^T M = R^-1
    double inflation;       // 0: adaptive
    float* coef;
    gocean_letkf_coef_t* oc;
};

struct ApplyArgs {
    float* state;
    long long plane, member_stride;
    int pitch, nx, ny, n_members, n_obs, local_n;
    long long n_tile_obs;
    const double* w_loc;
    const int32_t* boxes;
    const float* coef;
    const gocean_letkf_coef_t* oc;
    const gocean_letkf_tile_t* tiles;
    const int32_t* tile_obs;
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Sums of K values over the workgroup, in a fixed order.
template <int K>
__device__ __forceinline__ void block_sums(double (&v)[K], double (*red)[4]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = wave_sum(v[k]);
    __syncthreads();
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) red[w][k] = v[k];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int i = 0; i < kCoefWaves; ++i) s += red[i][k];
        v[k] = s;
    }
}

__global__ __launch_bounds__(kCoefNT) void k_letkf_coef(CoefArgs A) {
    __shared__ float s_hu[kMaxMembers];
    __shared__ float s_hv[kMaxMembers];
    __shared__ double red[kCoefWaves][4];
    const int o = A.group[blockIdx.x];
    if (o < 0 || o >= A.n_obs) return;                 // uniform over the workgroup
    const int ix = A.cells[2 * o], iy = A.cells[2 * o + 1];
    if (ix < 0 || ix >= A.nx || iy < 0 || iy >= A.ny) return;
    const int N = A.n_members;
    const long long cell = (long long)(iy + 2) * A.pitch + ix + 2;

    // H X_f and its member mean (ETKFOcean.py:763-775), float64
    double s[2] = {0.0, 0.0};
    for (int k = threadIdx.x; k < N; k += kCoefNT) {
        const float u = A.state[k * A.member_stride + A.plane + cell];
        const float v = A.state[k * A.member_stride + 2 * A.plane + cell];
        s_hu[k] = u;
        s_hv[k] = v;
        s[0] += u;
        s[1] += v;
    }
    block_sums<2>(s, red);
    const double mu = s[0] / N, mv = s[1] / N;
    const double D0 = A.y[2 * o] - mu, D1 = A.y[2 * o + 1] - mv;
    double ff = 1.0 / (A.inflation * A.inflation);
    if (A.inflation == 0.0) {                           // Saetrom and Omre, ETKFOcean.py:667-671
        const double q = A.r00 * D0 * D0 + 2.0 * A.r01 * D0 * D1 + A.r11 * D1 * D1;
        ff = 1.0 / (1.0 + q / (N - 2));
    }
    const double isq = 1.0 / sqrt((double)(N - 1));

    // Gram matrix Y Y^T (2 x 2)
    double g[3] = {0.0, 0.0, 0.0};
    for (int k = threadIdx.x; k < N; k += kCoefNT) {
        const double p = (double)s_hu[k] - mu, q = (double)s_hv[k] - mv;
        const double y0 = (A.m00 * p + A.m01 * q) * isq, y1 = A.m11 * q * isq;
        g[0] += y0 * y0;
        g[1] += y0 * y1;
        g[2] += y1 * y1;
    }
    block_sums<3>(g, red);
    // symmetric 2 x 2 eigenproblem by one Jacobi rotation
    const double th = 0.5 * atan2(2.0 * g[1], g[0] - g[2]);
    const double cs = cos(th), sn = sin(th);
    const double l0 = fmax(g[0] * cs * cs + 2.0 * g[1] * sn * cs + g[2] * sn * sn, 0.0);
    const double l1 = fmax(g[0] * sn * sn - 2.0 * g[1] * sn * cs + g[2] * cs * cs, 0.0);
    const double lmax = fmax(l0, l1);
    const bool k0 = lmax > 0.0 && l0 > 1e-30 * lmax, k1 = lmax > 0.0 && l1 > 1e-30 * lmax;
    const double s0 = sqrt(l0), s1 = sqrt(l1);
    const double md0 = A.m00 * D0 + A.m01 * D1, md1 = A.m11 * D1;
    const double ud0 = cs * md0 + sn * md1, ud1 = -sn * md0 + cs * md1;   // u_j . (M D)
    double a[2], c0[2], c1[2], z0[2], z1[2];
    const double ffs[2] = {1.0, ff};
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        a[i] = 1.0 / sqrt(ffs[i]);
        c0[i] = k0 ? 1.0 / sqrt(ffs[i] + l0) - a[i] : 0.0;
        c1[i] = k1 ? 1.0 / sqrt(ffs[i] + l1) - a[i] : 0.0;
        z0[i] = k0 ? s0 / (ffs[i] + l0) * ud0 * isq : 0.0;
        z1[i] = k1 ? s1 / (ffs[i] + l1) * ud1 * isq : 0.0;
    }
    const double i0 = k0 ? 1.0 / s0 : 0.0, i1 = k1 ? 1.0 / s1 : 0.0;
    double sv[2] = {0.0, 0.0};
    float* out = A.coef + (long long)o * N * kCoefPerMember;
    for (int k = threadIdx.x; k < N; k += kCoefNT) {
        const double p = (double)s_hu[k] - mu, q = (double)s_hv[k] - mv;
        const double y0 = (A.m00 * p + A.m01 * q) * isq, y1 = A.m11 * q * isq;
        const float v0 = (float)((cs * y0 + sn * y1) * i0);
        const float v1 = (float)((-sn * y0 + cs * y1) * i1);
        sv[0] += v0;
        sv[1] += v1;
        float* c = out + (long long)k * kCoefPerMember;
        c[0] = v0;
        c[1] = v1;
        c[2] = (float)(c0[0] * v0 + z0[0]);
        c[3] = (float)(c1[0] * v1 + z1[0]);
        c[4] = (float)(c0[1] * v0 + z0[1]);
        c[5] = (float)(c1[1] * v1 + z1[1]);
    }
    block_sums<2>(sv, red);
    if (threadIdx.x == 0) {
        gocean_letkf_coef_t r;
        r.a_eta = (float)a[0];
        r.a_uv = (float)a[1];
        r.sv0 = sv[0];
        r.sv1 = sv[1];
        A.oc[o] = r;
    }
}

// One workgroup per (tile row of kTileW cells, field): wave w of the
// workgroup takes members [w P, (w+1) P), P = ceil(N / kChunks), of every
// cell of the row.  Three phases, the member sums of the first two combined
// across the waves through LDS in a fixed order: the member mean (float64);
// p_j = sum_k (x_k - mean) V_kj (float32: the anomalies are small numbers
// without cancellation); the blend.  With REG the wave's members stay in
// registers across the phases, so the state is read once and written once.
template <bool REG>
__global__ __launch_bounds__(kTileW* kChunks) void k_letkf_apply(ApplyArgs A) {
    __shared__ double red_s[kChunks][kTileW];
    __shared__ float red_p[kChunks][2 * kMaxCover][kTileW];
    const gocean_letkf_tile_t t = A.tiles[blockIdx.x];
    const int f = blockIdx.y;
    const int lane = (int)(threadIdx.x % kTileW), chunk = (int)(threadIdx.x / kTileW);
    const int x = t.x0 + lane, y = t.y0;
    const bool live = x >= 0 && x < A.nx && y >= 0 && y < A.ny && t.begin >= 0 && t.count >= 0 &&
                      (long long)t.begin + t.count <= A.n_tile_obs;
    const int L = A.local_n;

    // the group's observations whose weight reaches this cell, in group order
    int cov[kMaxCover];
    double wl[kMaxCover];
#pragma unroll
    for (int c = 0; c < kMaxCover; ++c) {
        cov[c] = 0;
        wl[c] = 0.0;
    }
    int nc = 0;
    double wsum = 0.0;
    for (int j = 0; live && j < t.count; ++j) {
        const int o = A.tile_obs[t.begin + j];
        if (o < 0 || o >= A.n_obs) continue;
        int rx = x - A.boxes[2 * o], ry = y - A.boxes[2 * o + 1];
        rx += rx < 0 ? A.nx : 0;
        ry += ry < 0 ? A.ny : 0;
        if (rx < 0 || rx >= L || ry < 0 || ry >= L) continue;
        const double w = A.w_loc[ry * L + rx];
        if (w == 0.0 || nc >= kMaxCover) continue;
#pragma unroll
        for (int c = 0; c < kMaxCover; ++c)
            if (c == nc) {
                cov[c] = o;
                wl[c] = w;
            }
        ++nc;
        wsum += w;
    }

    const float2* cf[kMaxCover];      // per member: (V0, V1), (B_eta0, B_eta1), (B_uv0, B_uv1)
#pragma unroll
    for (int c = 0; c < kMaxCover; ++c)
        cf[c] = reinterpret_cast<const float2*>(A.coef + (long long)cov[c] * A.n_members * kCoefPerMember);
    const int N = A.n_members;
    const int per = (N + kChunks - 1) / kChunks;
    const int k0 = chunk * per, k1 = min(N, k0 + per);
    const long long ms = A.member_stride;
    float* q = A.state + f * A.plane + (long long)(y + 2) * A.pitch + x + 2;
    constexpr int kPairs = kCoefPerMember / 2;

    // phase 1: member mean
    double s = 0.0;
    float xs[REG ? kRegMembers : 1];
    if (nc > 0) {
        if constexpr (REG) {
#pragma unroll
            for (int j = 0; j < kRegMembers; ++j)
                if (k0 + j < k1) xs[j] = q[(k0 + j) * ms];
#pragma unroll
            for (int j = 0; j < kRegMembers; ++j)
                if (k0 + j < k1) s += (double)xs[j];
        } else {
            for (int k = k0; k < k1; ++k) s += (double)q[k * ms];
        }
    }
    red_s[chunk][lane] = s;
    __syncthreads();
    s = 0.0;
#pragma unroll
    for (int h = 0; h < kChunks; ++h) s += red_s[h][lane];
    const float mean = (float)(s / N);

    // phase 2: p_j = sum_k (x_k - mean) V_kj
    float p0[kMaxCover], p1[kMaxCover];
#pragma unroll
    for (int c = 0; c < kMaxCover; ++c) p0[c] = p1[c] = 0.0f;
    auto project = [&](int k, float v) {
        const float d = v - mean;
#pragma unroll
        for (int c = 0; c < kMaxCover; ++c)
            if (c < nc) {
                const float2 V = cf[c][k * kPairs];
                p0[c] += d * V.x;
                p1[c] += d * V.y;
            }
    };
    if (nc > 0) {
        if constexpr (REG) {
#pragma unroll
            for (int j = 0; j < kRegMembers; ++j)
                if (k0 + j < k1) project(k0 + j, xs[j]);
        } else {
            for (int k = k0; k < k1; ++k) project(k, q[k * ms]);
        }
    }
#pragma unroll
    for (int c = 0; c < kMaxCover; ++c) {
        red_p[chunk][2 * c][lane] = p0[c];
        red_p[chunk][2 * c + 1][lane] = p1[c];
    }
    __syncthreads();
    if (nc == 0) return;                                // outside every box: forecast kept
#pragma unroll
    for (int c = 0; c < kMaxCover; ++c) p0[c] = p1[c] = 0.0f;
#pragma unroll
    for (int h = 0; h < kChunks; ++h) {
#pragma unroll
        for (int c = 0; c < kMaxCover; ++c) {
            p0[c] += red_p[h][2 * c][lane];
            p1[c] += red_p[h][2 * c + 1][lane];
        }
    }

    // phase 3: x_e <- W_f x_e + sum_o w_o (mean + a (x_e - mean) + p . B_e)
    const float wf = (float)(1.0 - wsum / fmax(wsum, 1.0));   // ETKFOcean.py:465-468
    float a[kMaxCover], w[kMaxCover];
#pragma unroll
    for (int c = 0; c < kMaxCover; ++c) {
        const gocean_letkf_coef_t r = A.oc[cov[c]];
        a[c] = f == 0 ? r.a_eta : r.a_uv;
        w[c] = (float)wl[c];
    }
    const int b = f == 0 ? 1 : 2;
    auto blend = [&](int k, float v) {
        const float d = v - mean;
        float acc = wf * v;
#pragma unroll
        for (int c = 0; c < kMaxCover; ++c)
            if (c < nc) {
                const float2 B = cf[c][k * kPairs + b];
                acc += w[c] * (mean + a[c] * d + p0[c] * B.x + p1[c] * B.y);
            }
        q[k * ms] = acc;
    };
    if constexpr (REG) {
#pragma unroll
        for (int j = 0; j < kRegMembers; ++j)
            if (k0 + j < k1) blend(k0 + j, xs[j]);
    } else {
        for (int k = k0; k < k1; ++k) blend(k, q[k * ms]);
    }
}

int check_layout(const gocean_letkf_layout_t* L) {
    if (!L) return set_error(GOCEAN_EINVAL, "layout is NULL");
    if (L->n_members < 2 || L->n_members > kMaxMembers)
        return set_error(GOCEAN_EINVAL, "members %d outside [2, %d]", L->n_members, kMaxMembers);
    if (L->nx < 1 || L->ny < 1) return set_error(GOCEAN_EINVAL, "domain %d x %d", L->nx, L->ny);
    if (L->pitch < L->nx + 4) return set_error(GOCEAN_EINVAL, "pitch %d < nx + 4", L->pitch);
    if (L->plane < (long long)(L->ny + 4) * L->pitch) return set_error(GOCEAN_EINVAL, "plane smaller than (ny+4) x pitch");
    if (L->member_stride < 3 * L->plane) return set_error(GOCEAN_EINVAL, "member stride smaller than three planes");
    return 0;
}

}  // namespace
}  // namespace gocean

using namespace gocean;

extern "C" {

int gocean_letkf_tile_shape(int32_t* tile_w, int32_t* tile_h, int32_t* max_cover, int32_t* max_members) {
    clear_error();
    if (tile_w) *tile_w = kTileW;
    if (tile_h) *tile_h = kTileH;
    if (max_cover) *max_cover = kMaxCover;
    if (max_members) *max_members = kMaxMembers;
    return 0;
}

int gocean_letkf_coefficients(const gocean_letkf_layout_t* lay, const float* state, const int32_t* group,
                              int32_t n_group, const int32_t* cells, const double* y, int32_t n_obs,
                              const double* obs_cov, double inflation, float* coef, gocean_letkf_coef_t* oc,
                              void* stream) {
    clear_error();
    if (int e = check_layout(lay)) return e;
    if (n_group < 0 || n_obs < 1) return set_error(GOCEAN_EINVAL, "group size %d, observations %d", n_group, n_obs);
    if (!state || !group || !cells || !y || !obs_cov || !coef || !oc) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    if (!(inflation >= 0.0) || !std::isfinite(inflation)) return set_error(GOCEAN_EINVAL, "inflation %g", inflation);
    if (inflation == 0.0 && lay->n_members < 3) return set_error(GOCEAN_EINVAL, "adaptive inflation needs 3 members");
    const double R00 = obs_cov[0], R01 = obs_cov[1], R10 = obs_cov[2], R11 = obs_cov[3];
    const double det = R00 * R11 - R01 * R10;
    if (!(R00 > 0.0) || !(det > 0.0) || R01 != R10)
        return set_error(GOCEAN_EINVAL, "observation covariance is not symmetric positive definite");
    CoefArgs A{};
    A.state = state;
    A.plane = lay->plane;
    A.member_stride = lay->member_stride;
    A.pitch = lay->pitch;
    A.nx = lay->nx;
    A.ny = lay->ny;
    A.n_members = lay->n_members;
    A.n_obs = n_obs;
    A.group = group;
    A.cells = cells;
    A.y = y;
    A.r00 = R11 / det;
    A.r01 = -R01 / det;
    A.r11 = R00 / det;
    A.m00 = std::sqrt(A.r00);
    A.m01 = A.r01 / A.m00;
    A.m11 = std::sqrt(A.r11 - A.m01 * A.m01);
    A.inflation = inflation;
    A.coef = coef;
    A.oc = oc;
    if (n_group == 0) return 0;
    hipLaunchKernelGGL(k_letkf_coef, dim3(n_group), dim3(kCoefNT), 0, (hipStream_t)stream, A);
    return check_launch("k_letkf_coef");
}

int gocean_letkf_apply(const gocean_letkf_layout_t* lay, float* state, const double* w_loc, int32_t local_n,
                       const int32_t* boxes, int32_t n_obs, const float* coef, const gocean_letkf_coef_t* oc,
                       const gocean_letkf_tile_t* tiles, int32_t n_tiles, const int32_t* tile_obs,
                       int32_t n_tile_obs, void* stream) {
    clear_error();
    if (int e = check_layout(lay)) return e;
    if (local_n < 1 || local_n % 2 == 0 || local_n > std::min(lay->nx, lay->ny))
        return set_error(GOCEAN_EINVAL, "local box side %d must be odd and fit the %d x %d domain", local_n, lay->nx,
                         lay->ny);
    if (n_obs < 1 || n_tiles < 0 || n_tile_obs < 0)
        return set_error(GOCEAN_EINVAL, "observations %d, tiles %d, tile entries %d", n_obs, n_tiles, n_tile_obs);
    if (!state || !w_loc || !boxes || !coef || !oc || (n_tiles > 0 && (!tiles || !tile_obs)))
        return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    ApplyArgs A{};
    A.state = state;
    A.plane = lay->plane;
    A.member_stride = lay->member_stride;
    A.pitch = lay->pitch;
    A.nx = lay->nx;
    A.ny = lay->ny;
    A.n_members = lay->n_members;
    A.n_obs = n_obs;
    A.local_n = local_n;
    A.n_tile_obs = n_tile_obs;
    A.w_loc = w_loc;
    A.boxes = boxes;
    A.coef = coef;
    A.oc = oc;
    A.tiles = tiles;
    A.tile_obs = tile_obs;
    if (n_tiles == 0) return 0;
    const dim3 grid(n_tiles, 3);
    if ((lay->n_members + kChunks - 1) / kChunks <= kRegMembers)
        hipLaunchKernelGGL(k_letkf_apply<true>, grid, dim3(kTileW * kChunks), 0, (hipStream_t)stream, A);
    else
        hipLaunchKernelGGL(k_letkf_apply<false>, grid, dim3(kTileW * kChunks), 0, (hipStream_t)stream, A);
    return check_launch("k_letkf_apply");
}

}  // extern "C"
