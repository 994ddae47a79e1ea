// gocean_ens.hip -- batched ensemble of CDKLM16 domains of any sizes.
//
// Replaces the reference's per-member Python loop (OceanModelEnsemble.modelStep,
// gpu_ocean/SWESimulators/OceanModelEnsemble.py:105-117;
// EnsembleFromFiles.stepToObservation, EnsembleFromFiles.py:302-325) with one
// launch per RK substep that covers every member's tiles.  Member state lives
// in caller-owned flat buffers located by per-member offsets; the library
// keeps a device table of the members and a tile -> member map, plus an
// optional captured HIP graph of n fixed-dt steps.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "../../include/gocean.h"
#include "adjoint_march.hpp"
#include "adjoint_tile.hpp"
#include "cdklm16_march.hpp"
#include "cdklm16_tile.hpp"
#include "gocean_internal.hpp"

namespace gocean {

// one member as the device sees it
struct EnsMember {
    int nx, ny, pitch, hi_pitch;
    long long state_off, plane, hi_off, hm_off, field_off;
    int tiles_x, tile_begin;
    float dx, dy, g, r, theta, desing, flux_eps, cutoff, land;
    float f, north_x, north_y;
    float wind_x, wind_y;          // wind stress (wind_lerp of the member's samples)
    int bc_n, bc_e, bc_s, bc_w;
    float vdx, udy;                // as SubstepArgs (with r_dx, r_dy, r_two_g below)
    // reverse mode (set by gocean_ensemble_adjoint_setup): the member's block
    // of the work buffer and its tiles in the recording-forward (32x16) and
    // adjoint (kAdjBW x kAdjBH) launches
    long long work_off;
    int ftiles_x, ftile_begin, atiles_x, atile_begin;
    int astrips_x, astrip_begin;   // the member's strips in the column-marching reverse launch
    float r_dx, r_dy, r_two_g, Dx, Dy, Ddx, Ddy, g_dx, g_dy;   // as AdjTileArgs / SubstepArgs
};

struct EnsLaunch {
    const EnsMember* members;
    const int* tile_member;
    const float* state_in;
    const float* state_qn;
    float* state_out;
    const float* Hi;
    const float* fields;
    const float* dt;           // one per member
    int rk_order, rk_step;
    int n_tiles;
    int march_bh;              // rows per strip (column-marching configurations)
};

// 512-thread tiles: 8 waves per SIMD, as k_substep
// the member's substep arguments (its constants, offsets and device dt)
template <int MODE>
__device__ __forceinline__ void member_args(const EnsLaunch& L, const EnsMember& M, int m, SubstepArgs& A) {
    const long long so = M.state_off, pl = M.plane;
    A.eta_in = L.state_in + so; A.hu_in = L.state_in + so + pl; A.hv_in = L.state_in + so + 2 * pl;
    A.eta_qn = L.state_qn + so; A.hu_qn = L.state_qn + so + pl; A.hv_qn = L.state_qn + so + 2 * pl;
    A.eta_out = L.state_out + so; A.hu_out = L.state_out + so + pl; A.hv_out = L.state_out + so + 2 * pl;
    A.Hi = L.Hi + M.hi_off;
    if (MODE == kModeField) {
        A.f_field = L.fields + M.field_off;
        A.nx_field = A.f_field + pl;
        A.ny_field = A.f_field + 2 * pl;
    } else {
        A.f_field = nullptr; A.nx_field = nullptr; A.ny_field = nullptr;
    }
    A.dt_dev = L.dt + m;
    A.nx = M.nx; A.ny = M.ny; A.pitch = M.pitch; A.hi_pitch = M.hi_pitch;
    A.dx = M.dx; A.dy = M.dy; A.g = M.g; A.r = M.r; A.theta = M.theta;
    A.desing_eps = M.desing; A.flux_eps = M.flux_eps; A.cutoff = M.cutoff; A.land = M.land;
    A.f = M.f; A.north_x = M.north_x; A.north_y = M.north_y;
    A.wind_x = M.wind_x; A.wind_y = M.wind_y;
    A.wind_grid = nullptr; A.wind_wx = 0; A.wind_wy = 0; A.wind_t = 0.0f;   // members: uniform wind
    A.dt = 0.0f;
    A.set_flags(M.bc_n, M.bc_e, M.bc_s, M.bc_w, L.rk_order, L.rk_step);
    A.vdx = M.vdx; A.udy = M.udy; A.r_two_g = M.r_two_g; A.r_dx = M.r_dx; A.r_dy = M.r_dy;
    A.tiles_x = M.tiles_x; A.tiles_y = 0;
}

// column-marching substep over every member's strips: wavefront -> strip
// through the tile map (one entry per strip), as k_march
template <bool FAST>
__global__ __launch_bounds__(256) void k_ens_march(EnsLaunch L) {
    const int wave = xcd_band(blockIdx.x, gridDim.x) * 4 + march::wave_in_group();
    if (wave >= L.n_tiles) return;
    const int m = L.tile_member[wave];
    if (L.dt[m] == 0.0f) return;                       // member done (see k_ens_substep)
    const EnsMember& M = L.members[m];
    SubstepArgs A;
    member_args<kModeZero>(L, M, m, A);
    const int lt = wave - M.tile_begin;
    const int sy = lt / M.tiles_x, sx = lt - sy * M.tiles_x;
    march::march_strip<FAST, false>(A, sx, sy, L.march_bh);   // members: uniform wind
}

template <int BW, int BH, int NT, int MODE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT >= 512 ? 8 : 1))) void k_ens_substep(EnsLaunch L) {
    __shared__ TileSmem<BW, BH, MODE> S;
    const int t = xcd_band(blockIdx.x, gridDim.x);
    const int m = L.tile_member[t];
    // dt 0 never occurs as a step length (the step loop only runs while time
    // is left): it marks a member that has reached its end time and is left
    // as it is, so members with different dt can share the launches of a
    // time-targeted step
    if (L.dt[m] == 0.0f) return;                       // uniform over the workgroup
    const EnsMember& M = L.members[m];
    SubstepArgs A;
    member_args<MODE>(L, M, m, A);
    const int lt = t - M.tile_begin;
    const int ty = lt / M.tiles_x, tx = lt - ty * M.tiles_x;
    if (tile_is_interior<BW, BH>(A, tx, ty))
        cdklm_tile<BW, BH, NT, MODE, false, false>(A, tx, ty, S);
    else
        cdklm_tile<BW, BH, NT, MODE>(A, tx, ty, S);
}

// per-member CFL minimum: one workgroup per member; dt[m] = courant * min
// (rounded to float), or the raw minimum when courant <= 0
__global__ __launch_bounds__(256) void k_ens_dt(const EnsMember* members, const float* state,
                                                const float* Hm, float courant, float* dt_out) {
    __shared__ float red[4];
    const EnsMember& M = members[blockIdx.x];
    const float* eta = state + M.state_off;
    const float* hu = eta + M.plane;
    const float* hv = eta + 2 * M.plane;
    const float* H = Hm + M.hm_off;
    const float qdx = 0.25f * M.dx, qdy = 0.25f * M.dy;
    float m = 1.0e5f;
    // rows of the interior, columns across the workgroup: no per-cell
    // index division
    for (int j = 2; j < M.ny + 2; ++j)
    for (int i = threadIdx.x + 2; i < M.nx + 2; i += blockDim.x) {
        const long long o = (long long)j * M.pitch + i;
        const float Hc = H[o];
        const float h = eta[o] + Hc;
        if (h <= 0.0f || fabsf(Hc - M.land) < kDryEps) continue;
        const float rh = rcp_refined(h);
        const float u = div_r(hu[o], h, rh), v = div_r(hv[o], h, rh);
        const float c = sqrt_rn(M.g * h);
        float d = qdx / fabsf(u + c);
        d = fminf(d, qdx / fabsf(u - c));
        d = fminf(d, qdy / fabsf(v + c));
        d = fminf(d, qdy / fabsf(v - c));
        m = fminf(m, d);
    }
    for (int o = 32; o > 0; o >>= 1) m = fminf(m, __shfl_xor(m, o, 64));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        m = fminf(fminf(red[0], red[1]), fminf(red[2], red[3]));
        dt_out[blockIdx.x] = courant > 0.0f ? courant * m : m;
    }
}

__global__ void k_ens_bathy(const EnsMember* members, float* Hi_all, float* Hm_all) {
    // one workgroup per member: halo (NS then EW) then cell depths
    const EnsMember& M = members[blockIdx.x];
    float* Hi = Hi_all + M.hi_off;
    float* Hm = Hm_all + M.hm_off;
    const int nx = M.nx, ny = M.ny, hp = M.hi_pitch;
    const bool pns = M.bc_n == kPeriodic, pew = M.bc_e == kPeriodic;
    for (int i = threadIdx.x; i <= nx + 4; i += blockDim.x) {
        if (pns) {
            Hi[i] = Hi[(long long)ny * hp + i];
            Hi[hp + i] = Hi[(long long)(ny + 1) * hp + i];
            for (int j = ny + 2; j <= ny + 4; ++j) Hi[(long long)j * hp + i] = Hi[(long long)(j - ny) * hp + i];
        } else {
            for (int j = 0; j < 2; ++j) Hi[(long long)j * hp + i] = Hi[(long long)(4 - j) * hp + i];
            for (int j = 0; j < 2; ++j) Hi[(long long)(ny + 4 - j) * hp + i] = Hi[(long long)(ny + j) * hp + i];
        }
    }
    __syncthreads();
    for (int j = threadIdx.x; j <= ny + 4; j += blockDim.x) {
        float* row = Hi + (long long)j * hp;
        if (pew) {
            row[0] = row[nx];
            row[1] = row[nx + 1];
            for (int i = nx + 2; i <= nx + 4; ++i) row[i] = row[i - nx];
        } else {
            for (int i = 0; i < 2; ++i) row[i] = row[4 - i];
            for (int i = 0; i < 2; ++i) row[nx + 4 - i] = row[nx + i];
        }
    }
    __syncthreads();
    const int n = (nx + 4) * (ny + 4);
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int j = k / (nx + 4), i = k % (nx + 4);
        auto flag = [&](float v) { return (fabsf(v - M.land) <= kDryEps) ? kDryFlag : v; };
        const float a = flag(Hi[(long long)j * hp + i]);
        const float b = flag(Hi[(long long)(j + 1) * hp + i]);
        const float c = flag(Hi[(long long)j * hp + i + 1]);
        const float d = flag(Hi[(long long)(j + 1) * hp + i + 1]);
        const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag)
                        - (float)(c == kDryFlag) - (float)(d == kDryFlag);
        Hm[(long long)j * M.pitch + i] = (wet == 0.0f) ? M.land : (((a + b) + c) + d) / wet;
    }
}

__global__ void k_ens_periodic(const EnsMember* members, float* state) {
    // one workgroup per member: NS ghost rows, then EW ghost columns
    const EnsMember& M = members[blockIdx.x];
    const int nx = M.nx, ny = M.ny, p = M.pitch;
    for (int f = 0; f < 3; ++f) {
        float* a = state + M.state_off + f * M.plane;
        if (M.bc_n == kPeriodic) {
            for (int i = threadIdx.x; i < nx + 4; i += blockDim.x) {
                a[i] = a[(long long)ny * p + i];
                a[p + i] = a[(long long)(ny + 1) * p + i];
                a[(long long)(ny + 2) * p + i] = a[2LL * p + i];
                a[(long long)(ny + 3) * p + i] = a[3LL * p + i];
            }
        }
        __syncthreads();
        if (M.bc_e == kPeriodic) {
            for (int j = threadIdx.x; j < ny + 4; j += blockDim.x) {
                float* row = a + (long long)j * p;
                row[0] = row[nx];
                row[1] = row[nx + 1];
                row[nx + 2] = row[2];
                row[nx + 3] = row[3];
            }
        }
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// reverse mode of the batched ensemble: the same per-tile device code as the
// single-domain adjoint, with each workgroup's member from a tile map.
// A state (3 planes), record (5 planes) or adjoint lives either in a flat
// buffer laid out like the ensemble state (member m at state_off) or in the
// member's block of the work buffer (plane index idx from work_off).
// ---------------------------------------------------------------------------
struct EnsRef {
    const float* base;
    int in_work;
    int idx;
};

__device__ __forceinline__ long long ens_off(const EnsMember& M, const EnsRef& r) {
    return r.in_work ? M.work_off + (long long)r.idx * M.plane : M.state_off;
}

constexpr int kEnsTapeBW = 32, kEnsTapeBH = 16, kEnsTapeNT = 512;

struct EnsTapeLaunch {
    const EnsMember* members;
    const int* tile_member;
    EnsRef in, qn, out, tape;
    const float* Hi;
    const float* dt;
    int rk_step;
};

// 8 waves per SIMD, as k_ens_substep
__global__ __launch_bounds__(kEnsTapeNT) __attribute__((amdgpu_waves_per_eu(8))) void k_ens_substep_tape(EnsTapeLaunch L) {
    __shared__ TileSmem<kEnsTapeBW, kEnsTapeBH, kModeZero, true> S;
    const int t = xcd_band(blockIdx.x, gridDim.x);
    const int m = L.tile_member[t];
    const EnsMember& M = L.members[m];
    SubstepArgs A;
    const long long pl = M.plane;
    const float* in = L.in.base + ens_off(M, L.in);
    const float* qn = L.qn.base + ens_off(M, L.qn);
    float* out = const_cast<float*>(L.out.base) + ens_off(M, L.out);
    float* tp = const_cast<float*>(L.tape.base) + ens_off(M, L.tape);
    A.eta_in = in; A.hu_in = in + pl; A.hv_in = in + 2 * pl;
    A.eta_qn = qn; A.hu_qn = qn + pl; A.hv_qn = qn + 2 * pl;
    A.eta_out = out; A.hu_out = out + pl; A.hv_out = out + 2 * pl;
    A.tape_u = tp; A.tape_v = tp + pl; A.tape_t = tp + 2 * pl; A.tape_s = tp + 3 * pl;
    A.tape_code = reinterpret_cast<uint32_t*>(tp + 4 * pl);
    A.Hi = L.Hi + M.hi_off;
    A.f_field = nullptr; A.nx_field = nullptr; A.ny_field = nullptr;
    A.dt_dev = L.dt + m;
    A.nx = M.nx; A.ny = M.ny; A.pitch = M.pitch; A.hi_pitch = M.hi_pitch;
    A.dx = M.dx; A.dy = M.dy; A.g = M.g; A.r = M.r; A.theta = M.theta;
    A.desing_eps = M.desing; A.flux_eps = M.flux_eps; A.cutoff = M.cutoff; A.land = M.land;
    A.f = M.f; A.north_x = M.north_x; A.north_y = M.north_y;
    A.wind_x = M.wind_x; A.wind_y = M.wind_y;
    A.wind_grid = nullptr; A.wind_wx = 0; A.wind_wy = 0; A.wind_t = 0.0f;   // members: uniform wind
    A.dt = 0.0f;
    A.set_flags(M.bc_n, M.bc_e, M.bc_s, M.bc_w, 2, L.rk_step);
    A.vdx = M.vdx; A.udy = M.udy; A.r_two_g = M.r_two_g; A.r_dx = M.r_dx; A.r_dy = M.r_dy;
    A.tiles_x = M.ftiles_x; A.tiles_y = 0;
    const int lt = t - M.ftile_begin;
    const int ty = lt / M.ftiles_x, tx = lt - ty * M.ftiles_x;
    if (tile_is_interior<kEnsTapeBW, kEnsTapeBH>(A, tx, ty))
        cdklm_tile<kEnsTapeBW, kEnsTapeBH, kEnsTapeNT, kModeZero, true, false>(A, tx, ty, S);
    else
        cdklm_tile<kEnsTapeBW, kEnsTapeBH, kEnsTapeNT, kModeZero, true>(A, tx, ty, S);
}

// column-marching recording substep over every member's strips (the strip
// map of the ensemble's marching configuration)
template <bool FAST>
__global__ __launch_bounds__(256) void k_ens_march_tape(EnsTapeLaunch L, int bh, int n_strips) {
    const int wave = xcd_band(blockIdx.x, gridDim.x) * 4 + march::wave_in_group();
    if (wave >= n_strips) return;
    const int m = L.tile_member[wave];
    const EnsMember& M = L.members[m];
    SubstepArgs A;
    const long long pl = M.plane;
    const float* in = L.in.base + ens_off(M, L.in);
    const float* qn = L.qn.base + ens_off(M, L.qn);
    float* out = const_cast<float*>(L.out.base) + ens_off(M, L.out);
    float* tp = const_cast<float*>(L.tape.base) + ens_off(M, L.tape);
    A.eta_in = in; A.hu_in = in + pl; A.hv_in = in + 2 * pl;
    A.eta_qn = qn; A.hu_qn = qn + pl; A.hv_qn = qn + 2 * pl;
    A.eta_out = out; A.hu_out = out + pl; A.hv_out = out + 2 * pl;
    A.tape_u = tp; A.tape_v = tp + pl; A.tape_t = tp + 2 * pl; A.tape_s = tp + 3 * pl;
    A.tape_code = reinterpret_cast<uint32_t*>(tp + 4 * pl);
    A.Hi = L.Hi + M.hi_off;
    A.f_field = nullptr; A.nx_field = nullptr; A.ny_field = nullptr;
    A.dt_dev = L.dt + m;
    A.nx = M.nx; A.ny = M.ny; A.pitch = M.pitch; A.hi_pitch = M.hi_pitch;
    A.dx = M.dx; A.dy = M.dy; A.g = M.g; A.r = M.r; A.theta = M.theta;
    A.desing_eps = M.desing; A.flux_eps = M.flux_eps; A.cutoff = M.cutoff; A.land = M.land;
    A.f = M.f; A.north_x = M.north_x; A.north_y = M.north_y;
    A.wind_x = M.wind_x; A.wind_y = M.wind_y;
    A.wind_grid = nullptr; A.wind_wx = 0; A.wind_wy = 0; A.wind_t = 0.0f;   // members: uniform wind
    A.dt = 0.0f;
    A.set_flags(M.bc_n, M.bc_e, M.bc_s, M.bc_w, 2, L.rk_step);
    A.vdx = M.vdx; A.udy = M.udy; A.r_two_g = M.r_two_g; A.r_dx = M.r_dx; A.r_dy = M.r_dy;
    A.tiles_x = M.tiles_x; A.tiles_y = 0;
    const int lt = wave - M.tile_begin;
    const int sy = lt / M.tiles_x, sx = lt - sy * M.tiles_x;
    march::march_strip<FAST, false, true>(A, sx, sy, bh);
}

struct EnsAdjLaunch {
    const EnsMember* members;
    const int* tile_member;
    EnsRef eta, tape, w, w_clamp, x, x_clamp, o;   // x.base NULL: no x term
    const float* Hi;
    const float* dt;
    float alpha, gamma;
    int w_fold, x_fold;                            // as AdjTileArgs
};

// prefetching tile body at 8 waves per SIMD (four 40 KB workgroups per CU)
__global__ __launch_bounds__(kAdjNT) __attribute__((amdgpu_waves_per_eu(8))) void k_ens_adj_tile(EnsAdjLaunch L) {
    __shared__ AdjSmem<kAdjBW, kAdjBH> S;
    const int t = xcd_band(blockIdx.x, gridDim.x);
    const int m = L.tile_member[t];
    const EnsMember& M = L.members[m];
    const long long pl = M.plane;
    AdjTileArgs A;
    A.eta = L.eta.base + ens_off(M, L.eta);
    const float* tp = L.tape.base + ens_off(M, L.tape);
    A.tU = tp; A.tV = tp + pl; A.tT = tp + 2 * pl; A.tS = tp + 3 * pl;
    A.tC = reinterpret_cast<const uint32_t*>(tp + 4 * pl);
    const float* w = L.w.base + ens_off(M, L.w);
    A.w_eta = w; A.w_hu = w + pl; A.w_hv = w + 2 * pl;
    A.w_clamp = L.w_clamp.base + ens_off(M, L.w_clamp);
    if (L.x.base) {
        const float* x = L.x.base + ens_off(M, L.x);
        A.x_eta = x; A.x_hu = x + pl; A.x_hv = x + 2 * pl;
        A.x_clamp = L.x_clamp.base + ens_off(M, L.x_clamp);
    } else {
        A.x_eta = nullptr; A.x_hu = nullptr; A.x_hv = nullptr; A.x_clamp = nullptr;
    }
    float* o = const_cast<float*>(L.o.base) + ens_off(M, L.o);
    A.o_eta = o; A.o_hu = o + pl; A.o_hv = o + 2 * pl;
    A.Hi = L.Hi + M.hi_off;
    A.dt_dev = L.dt + m; A.dt = 0.0f;
    A.alpha = L.alpha; A.gamma = L.gamma;
    A.w_fold = L.w_fold; A.x_fold = L.x_fold;
    A.nx = M.nx; A.ny = M.ny; A.pitch = M.pitch; A.hi_pitch = M.hi_pitch;
    A.dx = M.dx; A.dy = M.dy; A.g = M.g; A.theta = M.theta; A.desing = M.desing;
    A.flux_eps = M.flux_eps; A.cutoff = M.cutoff; A.land = M.land; A.f = M.f;
    A.bc_n = M.bc_n; A.bc_e = M.bc_e; A.bc_s = M.bc_s; A.bc_w = M.bc_w;
    A.tiles_x = M.atiles_x; A.tiles_y = 0;
    A.r_dx = M.r_dx; A.r_dy = M.r_dy; A.r_two_g = M.r_two_g; A.Dx = M.Dx; A.Dy = M.Dy;
    A.Ddx = M.Ddx; A.Ddy = M.Ddy; A.g_dx = M.g_dx; A.g_dy = M.g_dy;
    const int lt = t - M.atile_begin;
    const int ty = lt / M.atiles_x, tx = lt - ty * M.atiles_x;
    adj_tile_body<kAdjBW, kAdjBH, kAdjNT, true>(A, S, tx, ty, threadIdx.x);
}

// Column-marching reverse substep over every member's strips (wavefront ->
// strip through the strip map, as k_adj_march for one domain).  Each member's
// arrays are addressed from their first plane with 32-bit plane strides; the
// host takes this kernel only when every member's five record planes fit.
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_ens_adj_march(EnsAdjLaunch L,
                                                                                               int bh, int n_strips) {
    const int wave = xcd_band(blockIdx.x, gridDim.x) * 4 + march::wave_in_group();
    if (wave >= n_strips) return;
    const int m = L.tile_member[wave];
    const EnsMember& M = L.members[m];
    const uint32_t pb = (uint32_t)(M.plane * 4);       // plane stride in bytes
    adjm::MArgs A;
    A.eta = L.eta.base + ens_off(M, L.eta);
    A.tape = reinterpret_cast<const char*>(L.tape.base + ens_off(M, L.tape));
    A.w = reinterpret_cast<const char*>(L.w.base + ens_off(M, L.w));
    A.w_clamp = L.w_clamp.base + ens_off(M, L.w_clamp);
    if (L.x.base) {
        A.x = reinterpret_cast<const char*>(L.x.base + ens_off(M, L.x));
        A.x_clamp = L.x_clamp.base + ens_off(M, L.x_clamp);
    } else {
        A.x = nullptr;
        A.x_clamp = nullptr;
    }
    A.o = reinterpret_cast<char*>(const_cast<float*>(L.o.base) + ens_off(M, L.o));
    A.Hi = L.Hi + M.hi_off;
    A.dt_dev = L.dt + m;
    A.tape_stride = pb; A.w_stride = pb; A.x_stride = pb; A.o_stride = pb;
    A.dt = 0.0f; A.alpha = L.alpha; A.gamma = L.gamma;
    A.w_fold = L.w_fold; A.x_fold = L.x_fold;
    A.nx = M.nx; A.ny = M.ny; A.pitch = M.pitch; A.hi_pitch = M.hi_pitch;
    A.g = M.g; A.theta = M.theta; A.desing = M.desing; A.cutoff = M.cutoff; A.land = M.land; A.f = M.f;
    A.bc_n = M.bc_n; A.bc_e = M.bc_e; A.bc_s = M.bc_s; A.bc_w = M.bc_w;
    A.r_dx = M.r_dx; A.r_dy = M.r_dy; A.r_two_g = M.r_two_g; A.Dx = M.Dx; A.Dy = M.Dy;
    A.Ddx = M.Ddx; A.Ddy = M.Ddy; A.g_dx = M.g_dx; A.g_dy = M.g_dy;
    const int lt = wave - M.astrip_begin;
    const int sy = lt / M.astrips_x, sx = lt - sy * M.astrips_x;
    const int r0 = sy * bh;
    const int r1 = min(r0 + bh, M.ny + 4) - 1;
    adjm::adj_march_strip(A, sx, r0, r1);
}

// member blockIdx.y's three state planes from src to dst
__global__ __launch_bounds__(256) void k_ens_copy3(const EnsMember* members, EnsRef src, EnsRef dst) {
    const EnsMember& M = members[blockIdx.y];
    const float* a = src.base + ens_off(M, src);
    float* b = const_cast<float*>(dst.base) + ens_off(M, dst);
    const long long n = 3 * M.plane;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        b[i] = a[i];
}

__global__ __launch_bounds__(256) void k_ens_adj_fold(const EnsMember* members, EnsRef o) {
    const EnsMember& M = members[blockIdx.z];
    FoldArgs F;
    float* b = const_cast<float*>(o.base) + ens_off(M, o);
    F.e = b; F.u = b + M.plane; F.v = b + 2 * M.plane;
    F.nx = M.nx; F.ny = M.ny; F.pitch = M.pitch;
    F.bc_n = M.bc_n; F.bc_e = M.bc_e; F.bc_s = M.bc_s; F.bc_w = M.bc_w;
    fold_item(F, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x);
}

}  // namespace gocean

using namespace gocean;

struct gocean_ensemble {
    int n = 0;
    int rk_order = 2;
    int cfg = 1;
    int mode = kModeZero;
    int bw = 32, bh = 16, nt = 256;
    int n_tiles = 0;
    int device = 0;
    bool fast_arith = false;   // fast_arith_exact holds for every member
    EnsMember* d_members = nullptr;
    int* d_tile_member = nullptr;
    std::vector<EnsMember> members;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
    // reverse mode
    int adj_kmax = 0;
    int n_ftiles = 0, n_atiles = 0, max_side = 0;
    long long max_plane = 0;
    int* d_ftile_member = nullptr;
    int* d_atile_member = nullptr;
    // column-marching reverse launch: strip map, rows per strip, and whether
    // every member's planes fit its 32-bit offsets
    int* d_astrip_member = nullptr;
    int n_astrips = 0, adj_bh = 0;
    long long adj_cells = 0;
    bool adj_march_fits = false;
};

namespace gocean {

template <int BW, int BH, int NT, int MODE>
static void ens_launch_t(const EnsLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL((k_ens_substep<BW, BH, NT, MODE>), dim3(L.n_tiles), dim3(NT), 0, s, L);
}

static void ens_march(EnsLaunch L, int bh, bool fast_arith, hipStream_t s) {
    L.march_bh = bh;
    if (fast_arith)
        hipLaunchKernelGGL((k_ens_march<true>), dim3((L.n_tiles + 3) / 4), dim3(256), 0, s, L);
    else
        hipLaunchKernelGGL((k_ens_march<false>), dim3((L.n_tiles + 3) / 4), dim3(256), 0, s, L);
}

// rows per strip of the ensemble's column-marching launch (configuration
// with bh == 0): whole rounds of resident waves over all members' strips
static int ens_march_rows(const std::vector<EnsMember>& members, bool fast_arith) {
    std::vector<int> strips, ny;
    for (const EnsMember& M : members) {
        strips.push_back((M.nx + march::kOut - 1) / march::kOut);
        ny.push_back(M.ny);
    }
    const int resident = resident_waves(fast_arith ? (const void*)k_ens_march<true> : (const void*)k_ens_march<false>);
    return march_rows(strips.data(), ny.data(), (int)members.size(), resident);
}

static int ens_launch(const gocean_ensemble* E, EnsLaunch L, hipStream_t s) {
    L.members = E->d_members;
    L.tile_member = E->d_tile_member;
    L.n_tiles = E->n_tiles;
    if (E->mode == kModeField) {
        ens_launch_t<32, 8, 256, kModeField>(L, s);
    } else if (E->mode == kModeConst) {
        ens_launch_t<32, 8, 256, kModeConst>(L, s);
    } else {
        switch (E->cfg) {
            case 0: ens_launch_t<32, 8, 256, kModeZero>(L, s); break;
            case 1: ens_launch_t<32, 16, 256, kModeZero>(L, s); break;
            case 2: ens_launch_t<64, 8, 256, kModeZero>(L, s); break;
            case 3: ens_launch_t<16, 16, 256, kModeZero>(L, s); break;
            case 4: ens_launch_t<64, 4, 256, kModeZero>(L, s); break;
            case 5: ens_launch_t<32, 16, 512, kModeZero>(L, s); break;
            case 6: ens_launch_t<32, 14, 512, kModeZero>(L, s); break;
            case 7:
            case 8:
            case 9: ens_march(L, E->bh, E->fast_arith, s); break;
        }
    }
    return check_launch("ensemble substep launch");
}

static int ens_rk_steps(gocean_ensemble* E, int n_steps, float* s0, float* s1, float* s2,
                        const float* Hi, const float* fields, const float* dt, int rk_order,
                        hipStream_t s) {
    EnsLaunch L;
    std::memset(&L, 0, sizeof(L));
    L.Hi = Hi; L.fields = fields; L.dt = dt; L.rk_order = rk_order;
    for (int n = 0; n < n_steps; ++n) {
        int e = 0;
        if (rk_order == 2) {
            L.rk_step = 0; L.state_in = s0; L.state_qn = s1; L.state_out = s1;
            e = ens_launch(E, L, s);
            if (!e) { L.rk_step = 1; L.state_in = s1; L.state_qn = s0; L.state_out = s0; e = ens_launch(E, L, s); }
        } else if (rk_order == 1) {
            float* a = (n % 2 == 0) ? s0 : s1;
            float* b = (n % 2 == 0) ? s1 : s0;
            L.rk_step = 0; L.state_in = a; L.state_qn = b; L.state_out = b;
            e = ens_launch(E, L, s);
        } else {
            L.rk_step = 0; L.state_in = s0; L.state_qn = s1; L.state_out = s1;
            e = ens_launch(E, L, s);
            if (!e) { L.rk_step = 1; L.state_in = s1; L.state_qn = s0; L.state_out = s2; e = ens_launch(E, L, s); }
            if (!e) { L.rk_step = 2; L.state_in = s2; L.state_qn = s0; L.state_out = s0; e = ens_launch(E, L, s); }
        }
        if (e) return e;
    }
    return 0;
}

}  // namespace gocean

extern "C" {

int gocean_ensemble_create(const gocean_member_t* members, int n, int tile_config,
                           gocean_ensemble_t** out) {
    clear_error();
    if (!out) return set_error(GOCEAN_EINVAL, "out is NULL");
    *out = nullptr;
    if (n <= 0 || !members) return set_error(GOCEAN_EINVAL, "need at least one member");
    if (tile_config >= kNumTileConfigs)
        return set_error(GOCEAN_EINVAL, "tile configuration %d out of range", tile_config);
    gocean_ensemble* E = new (std::nothrow) gocean_ensemble();
    if (!E) return set_error(GOCEAN_ENOMEM, "host allocation failed");
    E->n = n;
    E->rk_order = members[0].domain.rk_order;
    long long mean_cells = 0;
    for (int k = 0; k < n; ++k) mean_cells += (long long)members[k].domain.nx * members[k].domain.ny;
    mean_cells /= n;
    E->cfg = tile_config < 0 ? default_tile_config(mean_cells) : tile_config;
    bool any_field = false, any_const = false;
    for (int k = 0; k < n; ++k) {
        const gocean_domain_t& d = members[k].domain;
        if (int e = validate_domain(&d)) {
            delete E;
            return set_error(e, "member %d: %s", k, last_error());
        }
        if (members[k].field_offset >= 0) any_field = true;
        else if (!(d.north_x == 0.0f && d.north_y == 1.0f)) any_const = true;
        if (d.rk_order != members[0].domain.rk_order) {
            delete E;
            return set_error(GOCEAN_EINVAL, "all members must use the same rk_order");
        }
    }
    if (any_field) {
        for (int k = 0; k < n; ++k)
            if (members[k].field_offset < 0) {
                delete E;
                return set_error(GOCEAN_EINVAL, "per-cell fields must be given for all members or none");
            }
    }
    E->mode = any_field ? kModeField : (any_const ? kModeConst : kModeZero);
    bool fits = true;
    for (int k = 0; k < n; ++k)
        fits = fits && march_fits(members[k].domain.ny, members[k].domain.pitch, members[k].domain.hi_pitch);
    if (E->cfg >= kFirstMarchConfig && !fits) {
        if (tile_config >= 0) {
            delete E;
            return set_error(GOCEAN_EINVAL, "tile configuration %d (column marching) needs member planes below 4 GiB",
                             tile_config);
        }
        E->cfg = kDefaultTileConfig;
    }
    E->fast_arith = true;
    for (int k = 0; k < n; ++k) {
        const gocean_domain_t& d = members[k].domain;
        E->fast_arith = E->fast_arith && fast_arith_exact(d.g, d.dx, d.dy, d.depth_cutoff);
    }
    const TileConfig& tc = (E->mode == kModeZero) ? kTileConfigs[E->cfg] : kTileConfigs[0];
    E->bw = tc.bw; E->bh = tc.bh; E->nt = tc.nt;
    if (E->bh == 0) {
        std::vector<EnsMember> shape(n);
        for (int k = 0; k < n; ++k) { shape[k].nx = members[k].domain.nx; shape[k].ny = members[k].domain.ny; }
        E->bh = ens_march_rows(shape, E->fast_arith);
    }
    std::vector<int> tile_member;
    E->members.resize(n);
    for (int k = 0; k < n; ++k) {
        const gocean_member_t& m = members[k];
        const gocean_domain_t& d = m.domain;
        EnsMember& M = E->members[k];
        std::memset(&M, 0, sizeof(M));
        M.nx = d.nx; M.ny = d.ny; M.pitch = d.pitch; M.hi_pitch = d.hi_pitch;
        M.state_off = m.state_offset; M.plane = (long long)(d.ny + 4) * d.pitch;
        M.hi_off = m.hi_offset; M.hm_off = m.hm_offset; M.field_off = m.field_offset;
        M.tiles_x = (d.nx + E->bw - 1) / E->bw;
        const int tiles_y = (d.ny + E->bh - 1) / E->bh;
        M.tile_begin = (int)tile_member.size();
        for (int t = 0; t < M.tiles_x * tiles_y; ++t) tile_member.push_back(k);
        M.dx = d.dx; M.dy = d.dy; M.g = d.g; M.r = d.r; M.theta = d.theta;
        M.desing = d.desing_eps; M.flux_eps = d.flux_slope_eps; M.cutoff = d.depth_cutoff;
        M.land = d.land_value; M.f = d.f; M.north_x = d.north_x; M.north_y = d.north_y;
        if (m.wind.grid) {
            delete E;
            return set_error(GOCEAN_EINVAL, "member %d: ensembles take uniform wind only (wind grid given)", k);
        }
        M.wind_x = wind_lerp(m.wind.t, m.wind.x0, m.wind.x1);
        M.wind_y = wind_lerp(m.wind.t, m.wind.y0, m.wind.y1);
        M.bc_n = d.bc_north; M.bc_e = d.bc_east; M.bc_s = d.bc_south; M.bc_w = d.bc_west;
        {
            SubstepArgs D;
            std::memset(&D, 0, sizeof(D));
            D.dx = d.dx; D.dy = d.dy; D.g = d.g;
            substep_derive(D);
            M.vdx = D.vdx; M.udy = D.udy; M.r_two_g = D.r_two_g; M.r_dx = D.r_dx; M.r_dy = D.r_dy;
        }
        if (m.state_offset < 0 || m.hi_offset < 0 || m.hm_offset < 0) {
            delete E;
            return set_error(GOCEAN_EINVAL, "member %d: negative buffer offset", k);
        }
    }
    E->n_tiles = (int)tile_member.size();
    (void)hipGetDevice(&E->device);
    if (hipMalloc(&E->d_members, sizeof(EnsMember) * n) != hipSuccess ||
        hipMalloc(&E->d_tile_member, sizeof(int) * E->n_tiles) != hipSuccess) {
        (void)hipGetLastError();
        (void)hipFree(E->d_members);
        (void)hipFree(E->d_tile_member);
        delete E;
        return set_error(GOCEAN_ENOMEM, "device allocation of the member table failed");
    }
    (void)hipMemcpy(E->d_members, E->members.data(), sizeof(EnsMember) * n, hipMemcpyHostToDevice);
    (void)hipMemcpy(E->d_tile_member, tile_member.data(), sizeof(int) * E->n_tiles, hipMemcpyHostToDevice);
    if (int e = check_launch("ensemble table upload")) {
        (void)hipFree(E->d_members);
        (void)hipFree(E->d_tile_member);
        delete E;
        return e;
    }
    *out = E;
    return 0;
}

int gocean_ensemble_set_tile_config(gocean_ensemble_t* E, int tile_config) {
    clear_error();
    if (!E) return set_error(GOCEAN_EINVAL, "ensemble is NULL");
    if (tile_config < 0) {
        long long mean_cells = 0;
        for (const EnsMember& M : E->members) mean_cells += (long long)M.nx * M.ny;
        tile_config = default_tile_config(mean_cells / E->n);
    }
    if (tile_config >= kNumTileConfigs)
        return set_error(GOCEAN_EINVAL, "tile configuration %d out of range", tile_config);
    // work already queued may still read the old tables
    if (hipDeviceSynchronize() != hipSuccess) return check_launch("ensemble tile configuration");
    E->cfg = tile_config;
    const TileConfig& tc = (E->mode == kModeZero) ? kTileConfigs[E->cfg] : kTileConfigs[0];
    E->bw = tc.bw; E->bh = tc.bh; E->nt = tc.nt;
    if (E->bh == 0) E->bh = ens_march_rows(E->members, E->fast_arith);
    std::vector<int> tile_member;
    for (int k = 0; k < E->n; ++k) {
        EnsMember& M = E->members[k];
        M.tiles_x = (M.nx + E->bw - 1) / E->bw;
        const int tiles_y = (M.ny + E->bh - 1) / E->bh;
        M.tile_begin = (int)tile_member.size();
        for (int t = 0; t < M.tiles_x * tiles_y; ++t) tile_member.push_back(k);
    }
    int* d_map = nullptr;
    if (hipMalloc(&d_map, sizeof(int) * tile_member.size()) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GOCEAN_ENOMEM, "device allocation of the tile map failed");
    }
    (void)hipMemcpy(d_map, tile_member.data(), sizeof(int) * tile_member.size(), hipMemcpyHostToDevice);
    (void)hipMemcpy(E->d_members, E->members.data(), sizeof(EnsMember) * E->n, hipMemcpyHostToDevice);
    (void)hipFree(E->d_tile_member);
    E->d_tile_member = d_map;
    E->n_tiles = (int)tile_member.size();
    // a captured graph holds the old launch shape: capture again
    if (E->graph_exec) (void)hipGraphExecDestroy(E->graph_exec);
    if (E->graph) (void)hipGraphDestroy(E->graph);
    E->graph_exec = nullptr;
    E->graph = nullptr;
    return check_launch("ensemble tile configuration");
}

int gocean_ensemble_destroy(gocean_ensemble_t* E) {
    clear_error();
    if (!E) return 0;
    if (E->graph_exec) (void)hipGraphExecDestroy(E->graph_exec);
    if (E->graph) (void)hipGraphDestroy(E->graph);
    if (E->capture_stream) (void)hipStreamDestroy(E->capture_stream);
    (void)hipFree(E->d_members);
    (void)hipFree(E->d_tile_member);
    (void)hipFree(E->d_ftile_member);
    (void)hipFree(E->d_atile_member);
    (void)hipFree(E->d_astrip_member);
    delete E;
    return 0;
}

int gocean_ensemble_info(const gocean_ensemble_t* E, int* n_members, int* n_tiles, int* mode) {
    clear_error();
    if (!E) return set_error(GOCEAN_EINVAL, "ensemble is NULL");
    if (n_members) *n_members = E->n;
    if (n_tiles) *n_tiles = E->n_tiles;
    if (mode) *mode = E->mode;
    return 0;
}

int gocean_ensemble_bathymetry_init(gocean_ensemble_t* E, float* Hi, float* Hm, void* stream) {
    clear_error();
    if (!E || !Hi || !Hm) return set_error(GOCEAN_EINVAL, "NULL argument");
    hipLaunchKernelGGL(k_ens_bathy, dim3(E->n), dim3(256), 0, (hipStream_t)stream, E->d_members, Hi, Hm);
    return check_launch("ensemble bathymetry launch");
}

int gocean_ensemble_periodic_bc(gocean_ensemble_t* E, float* state, void* stream) {
    clear_error();
    if (!E || !state) return set_error(GOCEAN_EINVAL, "NULL argument");
    hipLaunchKernelGGL(k_ens_periodic, dim3(E->n), dim3(256), 0, (hipStream_t)stream, E->d_members, state);
    return check_launch("ensemble periodic launch");
}

int gocean_ensemble_max_dt(gocean_ensemble_t* E, const float* state, const float* Hm, float courant,
                           float* dt_out, void* stream) {
    clear_error();
    if (!E || !state || !Hm || !dt_out) return set_error(GOCEAN_EINVAL, "NULL argument");
    hipLaunchKernelGGL(k_ens_dt, dim3(E->n), dim3(256), 0, (hipStream_t)stream, E->d_members, state, Hm,
                       courant, dt_out);
    return check_launch("ensemble max dt launch");
}

int gocean_ensemble_substep(gocean_ensemble_t* E, int rk_step, const float* state_in,
                            const float* state_qn, float* state_out, const float* Hi,
                            const float* fields, const float* dt, void* stream) {
    clear_error();
    if (!E || !state_in || !state_out || !Hi || !dt) return set_error(GOCEAN_EINVAL, "NULL argument");
    if (state_in == state_out) return set_error(GOCEAN_EINVAL, "input and output state must not alias");
    if (E->mode == kModeField && !fields) return set_error(GOCEAN_EINVAL, "members need the field buffer");
    if (rk_step < 0 || rk_step >= E->rk_order)
        return set_error(GOCEAN_EINVAL, "rk_step %d invalid for rk_order %d", rk_step, E->rk_order);
    EnsLaunch L;
    std::memset(&L, 0, sizeof(L));
    L.state_in = state_in; L.state_qn = state_qn ? state_qn : state_out; L.state_out = state_out;
    L.Hi = Hi; L.fields = fields; L.dt = dt;
    L.rk_order = E->rk_order; L.rk_step = rk_step;
    return ens_launch(E, L, (hipStream_t)stream);
}

int gocean_ensemble_rk_steps(gocean_ensemble_t* E, int n_steps, float* state0,
                             float* state1, float* state2, const float* Hi, const float* fields,
                             const float* dt, void* stream) {
    clear_error();
    const int rk_order = E ? E->rk_order : 0;
    if (!E || !state0 || !state1 || !Hi || !dt) return set_error(GOCEAN_EINVAL, "NULL argument");
    if (rk_order < 1 || rk_order > 3) return set_error(GOCEAN_EINVAL, "rk_order %d invalid", rk_order);
    if (rk_order == 3 && !state2) return set_error(GOCEAN_EINVAL, "rk_order 3 needs state2");
    if (rk_order == 1 && n_steps % 2) return set_error(GOCEAN_EINVAL, "rk_order 1 needs an even n_steps");
    if (E->mode == kModeField && !fields) return set_error(GOCEAN_EINVAL, "members need the field buffer");
    return ens_rk_steps(E, n_steps, state0, state1, state2, Hi, fields, dt, rk_order, (hipStream_t)stream);
}

int gocean_ensemble_capture(gocean_ensemble_t* E, int n_steps, float* state0,
                            float* state1, float* state2, const float* Hi, const float* fields,
                            const float* dt) {
    clear_error();
    const int rk_order = E ? E->rk_order : 0;
    if (!E || !state0 || !state1 || !Hi || !dt) return set_error(GOCEAN_EINVAL, "NULL argument");
    if (rk_order < 1 || rk_order > 3 || (rk_order == 3 && !state2) || (rk_order == 1 && n_steps % 2))
        return set_error(GOCEAN_EINVAL, "invalid rk_order / buffers / step count");
    if (E->graph_exec) { (void)hipGraphExecDestroy(E->graph_exec); E->graph_exec = nullptr; }
    if (E->graph) { (void)hipGraphDestroy(E->graph); E->graph = nullptr; }
    if (!E->capture_stream && hipStreamCreateWithFlags(&E->capture_stream, hipStreamNonBlocking) != hipSuccess)
        return check_launch("capture stream");
    if (hipStreamBeginCapture(E->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return check_launch("begin capture");
    const int e = ens_rk_steps(E, n_steps, state0, state1, state2, Hi, fields, dt, rk_order,
                               E->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(E->capture_stream, &g);
    if (e) { if (g) (void)hipGraphDestroy(g); return e; }
    if (ee != hipSuccess) return set_error((int)ee, "end capture: %s", hipGetErrorString(ee));
    E->graph = g;
    const hipError_t ei = hipGraphInstantiate(&E->graph_exec, g, nullptr, nullptr, 0);
    if (ei != hipSuccess) return set_error((int)ei, "graph instantiate: %s", hipGetErrorString(ei));
    return 0;
}

int gocean_ensemble_replay(gocean_ensemble_t* E, void* stream) {
    clear_error();
    if (!E) return set_error(GOCEAN_EINVAL, "ensemble is NULL");
    if (!E->graph_exec) return set_error(GOCEAN_ESTATE, "no captured graph: call gocean_ensemble_capture first");
    const hipError_t e = hipGraphLaunch(E->graph_exec, (hipStream_t)stream);
    if (e != hipSuccess) return set_error((int)e, "graph launch: %s", hipGetErrorString(e));
    return 0;
}


// ---------------------------------------------------------------------------
// reverse mode
// ---------------------------------------------------------------------------
long long gocean_ensemble_adjoint_setup(gocean_ensemble_t* E, int max_steps) {
    clear_error();
    if (!E) return set_error(GOCEAN_EINVAL, "ensemble is NULL");
    if (max_steps < 1) return set_error(GOCEAN_EINVAL, "max_steps must be >= 1");
    if (E->mode != kModeZero)
        return set_error(GOCEAN_EINVAL, "the adjoint supports constant f and zero grid rotation only");
    if (E->rk_order != 2) return set_error(GOCEAN_EINVAL, "the adjoint supports rk_order 2 only");
    std::vector<int> ftiles, atiles;
    long long off = 0;
    int max_side = 0;
    long long max_plane = 0;
    for (int k = 0; k < E->n; ++k) {
        EnsMember& M = E->members[k];
        if (M.r != 0.0f) return set_error(GOCEAN_EINVAL, "member %d: the adjoint supports r == 0 only", k);
        M.work_off = off;
        AdjTileArgs D;
        std::memset(&D, 0, sizeof(D));
        D.dx = M.dx; D.dy = M.dy; D.g = M.g; D.f = M.f;
        adj_derive(D);
        M.r_dx = D.r_dx; M.r_dy = D.r_dy; M.r_two_g = D.r_two_g; M.Dx = D.Dx; M.Dy = D.Dy;
        M.Ddx = D.Ddx; M.Ddy = D.Ddy; M.g_dx = D.g_dx; M.g_dy = D.g_dy;
        off += (16LL * max_steps + 9LL) * M.plane;
        M.ftiles_x = (M.nx + kEnsTapeBW - 1) / kEnsTapeBW;
        M.ftile_begin = (int)ftiles.size();
        for (int t = 0; t < M.ftiles_x * ((M.ny + kEnsTapeBH - 1) / kEnsTapeBH); ++t) ftiles.push_back(k);
        M.atiles_x = (M.nx + kAdjBW - 1) / kAdjBW;
        M.atile_begin = (int)atiles.size();
        for (int t = 0; t < M.atiles_x * ((M.ny + kAdjBH - 1) / kAdjBH); ++t) atiles.push_back(k);
        max_side = std::max(max_side, std::max(M.nx, M.ny));
        max_plane = std::max(max_plane, M.plane);
    }
    // column-marching reverse launch: one row count for all members' strips,
    // whole rounds of the kernel's resident waves
    std::vector<int> snx(E->n), sny(E->n), astrips;
    bool fits = true;
    long long cells = 0;
    for (int k = 0; k < E->n; ++k) {
        const EnsMember& M = E->members[k];
        snx[k] = (M.nx + adjm::kOut - 1) / adjm::kOut;
        sny[k] = M.ny + 4;
        fits = fits && 5LL * M.plane * 4 <= 0xFFFFFFFFLL && march_fits(M.ny, M.pitch, M.hi_pitch);
        cells += (long long)M.nx * M.ny;
    }
    static int resident = 0;
    if (!resident) resident = resident_waves((const void*)k_ens_adj_march);
    const int abh = march_rows(snx.data(), sny.data(), E->n, resident);
    for (int k = 0; k < E->n; ++k) {
        EnsMember& M = E->members[k];
        M.astrips_x = snx[k];
        M.astrip_begin = (int)astrips.size();
        for (int t = 0; t < snx[k] * ((sny[k] + abh - 1) / abh); ++t) astrips.push_back(k);
    }
    (void)hipFree(E->d_ftile_member);
    (void)hipFree(E->d_atile_member);
    (void)hipFree(E->d_astrip_member);
    E->d_ftile_member = nullptr;
    E->d_atile_member = nullptr;
    E->d_astrip_member = nullptr;
    if (hipMalloc(&E->d_ftile_member, sizeof(int) * ftiles.size()) != hipSuccess ||
        hipMalloc(&E->d_atile_member, sizeof(int) * atiles.size()) != hipSuccess ||
        hipMalloc(&E->d_astrip_member, sizeof(int) * astrips.size()) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GOCEAN_ENOMEM, "device allocation of the adjoint tile maps failed");
    }
    (void)hipMemcpy(E->d_ftile_member, ftiles.data(), sizeof(int) * ftiles.size(), hipMemcpyHostToDevice);
    (void)hipMemcpy(E->d_atile_member, atiles.data(), sizeof(int) * atiles.size(), hipMemcpyHostToDevice);
    (void)hipMemcpy(E->d_astrip_member, astrips.data(), sizeof(int) * astrips.size(), hipMemcpyHostToDevice);
    E->n_astrips = (int)astrips.size();
    E->adj_bh = abh;
    E->adj_cells = cells;
    E->adj_march_fits = fits;
    (void)hipMemcpy(E->d_members, E->members.data(), sizeof(EnsMember) * E->n, hipMemcpyHostToDevice);
    if (int e = check_launch("ensemble adjoint setup")) return e;
    E->adj_kmax = max_steps;
    E->n_ftiles = (int)ftiles.size();
    E->n_atiles = (int)atiles.size();
    E->max_side = max_side;
    E->max_plane = max_plane;
    return off;
}

}  // extern "C"

// Work of a reverse-mode segment, per member block (setup's max_steps K):
// Q_1..Q_K (3 planes each), Q*_0..Q*_{K-1} (3 each), branch records of the
// 2K substeps (5 each), lambda scratch (9).
struct EnsSegment {
    gocean_ensemble_t* E;
    const float* dt;
    const float* state0;
    const float* Hi;
    float* work;
    hipStream_t s;
    int K, k;
    EnsRef W(int idx) const { EnsRef r; r.base = work; r.in_work = 1; r.idx = idx; return r; }
    static EnsRef F(const float* b) { EnsRef r; r.base = b; r.in_work = 0; r.idx = 0; return r; }
    EnsRef Q(int i) const { return i == 0 ? F(state0) : W(3 * (i - 1)); }   // Q_i
    EnsRef Qs(int i) const { return W(3 * K + 3 * i); }                      // Q*_i
    EnsRef Tp(int i, int sub) const { return W(6 * K + 10 * i + 5 * sub); } // branch record

    int fwd(int rk_step, EnsRef in, EnsRef qn, EnsRef out, EnsRef tape) const {
        EnsTapeLaunch L;
        L.members = E->d_members; L.tile_member = E->d_ftile_member;
        L.in = in; L.qn = qn; L.out = out; L.tape = tape; L.Hi = Hi; L.dt = dt; L.rk_step = rk_step;
        if (E->cfg >= kFirstMarchConfig && E->mode == kModeZero) {
            // the ensemble's marching configuration records through its strips
            L.tile_member = E->d_tile_member;
            const dim3 grid((E->n_tiles + 3) / 4), block(256);
            if (E->fast_arith) hipLaunchKernelGGL((k_ens_march_tape<true>), grid, block, 0, s, L, E->bh, E->n_tiles);
            else hipLaunchKernelGGL((k_ens_march_tape<false>), grid, block, 0, s, L, E->bh, E->n_tiles);
        } else {
            hipLaunchKernelGGL(k_ens_substep_tape, dim3(E->n_ftiles), dim3(kEnsTapeNT), 0, s, L);
        }
        return check_launch("ensemble recording substep launch");
    }
    // w_fold / x_fold: w / x is an unfolded output of another substep
    // adjoint; fold_out: fold this one's output (the segment's result only)
    int adj(EnsRef eta, EnsRef tape, EnsRef w, EnsRef wc, EnsRef x, EnsRef xc, float alpha, float gamma,
            EnsRef o, int w_fold, int x_fold, bool fold_out) const {
        EnsAdjLaunch L;
        L.members = E->d_members; L.tile_member = E->d_atile_member;
        L.eta = eta; L.tape = tape; L.w = w; L.w_clamp = wc; L.x = x; L.x_clamp = xc; L.o = o;
        L.Hi = Hi; L.dt = dt; L.alpha = alpha; L.gamma = gamma; L.w_fold = w_fold; L.x_fold = x_fold;
        if (adj_use_march(E->adj_cells) && E->adj_march_fits) {
            L.tile_member = E->d_astrip_member;
            hipLaunchKernelGGL(k_ens_adj_march, dim3((E->n_astrips + 3) / 4), dim3(256), 0, s, L, E->adj_bh,
                               E->n_astrips);
        } else {
            hipLaunchKernelGGL(k_ens_adj_tile, dim3(E->n_atiles), dim3(kAdjNT), 0, s, L);
        }
        if (fold_out) {
            const int gx = (E->max_side + 16 + 255) / 256;
            hipLaunchKernelGGL(k_ens_adj_fold, dim3(gx, 9, E->n), dim3(256), 0, s, E->d_members, o);
        }
        return check_launch("ensemble adjoint launch");
    }
    int copy3(EnsRef src, EnsRef dst) const {
        const int gx = (int)std::min<long long>((3 * E->max_plane + 255) / 256, 4096);
        hipLaunchKernelGGL(k_ens_copy3, dim3(gx, E->n), dim3(256), 0, s, E->d_members, src, dst);
        return check_launch("ensemble state copy launch");
    }
    int record(float* state_out) const {
        // the final slot starts as the start state, so the cells a step does
        // not write (wall ghosts, pitch padding) carry over into state_out
        if (state_out)
            if (int e = copy3(F(state0), W(3 * (k - 1)))) return e;
        for (int i = 0; i < k; ++i) {
            if (int e = fwd(0, Q(i), Qs(i), Qs(i), Tp(i, 0))) return e;
            if (int e = fwd(1, Qs(i), Q(i), W(3 * i), Tp(i, 1))) return e;
        }
        if (state_out) return copy3(W(3 * (k - 1)), F(state_out));
        return 0;
    }
    int reverse(const float* lam_in, float* lam_out) const {
        EnsRef none; none.base = nullptr; none.in_work = 0; none.idx = 0;
        const EnsRef lam_s = W(16 * K), lam_a = W(16 * K + 3), lam_b = W(16 * K + 6);
        EnsRef le = F(lam_in);
        for (int i = k - 1; i >= 0; --i) {
            const EnsRef next = W(3 * i);          // Q_{i+1}
            EnsRef o;
            if (i == 0) o = F(lam_out);
            else o = (le.in_work && le.idx == lam_a.idx) ? lam_b : lam_a;
            const int l_unfolded = i < k - 1;
            if (int e = adj(Qs(i), Tp(i, 1), le, next, none, none, 0.5f, 0.0f, lam_s, l_unfolded, 0, false))
                return e;
            if (int e = adj(Q(i), Tp(i, 0), lam_s, Qs(i), le, next, 1.0f, 0.5f, o, 1, l_unfolded, i == 0))
                return e;
            le = o;
        }
        return 0;
    }
};

static int ens_segment(gocean_ensemble_t* E, int n_steps, const float* dt, const float* state0,
                       const float* Hi, float* work, long long work_floats, void* stream, EnsSegment& S) {
    if (!E || !dt || !state0 || !Hi || !work) return set_error(GOCEAN_EINVAL, "NULL argument");
    if (E->adj_kmax < 1) return set_error(GOCEAN_ESTATE, "call gocean_ensemble_adjoint_setup first");
    if (n_steps < 1 || n_steps > E->adj_kmax)
        return set_error(GOCEAN_EINVAL, "n_steps %d not in [1, %d]", n_steps, E->adj_kmax);
    const EnsMember& last = E->members[E->n - 1];
    if (work_floats < last.work_off + (16LL * E->adj_kmax + 9LL) * last.plane)
        return set_error(GOCEAN_EINVAL, "work buffer too small");
    S.E = E; S.dt = dt; S.state0 = state0; S.Hi = Hi; S.work = work; S.s = (hipStream_t)stream;
    S.K = E->adj_kmax; S.k = n_steps;
    return 0;
}

extern "C" {

int gocean_ensemble_adjoint_record(gocean_ensemble_t* E, int n_steps, const float* dt, const float* state0,
                                   const float* Hi, float* work, long long work_floats, float* state_out,
                                   void* stream) {
    clear_error();
    EnsSegment S;
    if (int e = ens_segment(E, n_steps, dt, state0, Hi, work, work_floats, stream, S)) return e;
    if (state_out == state0) return set_error(GOCEAN_EINVAL, "state_out must not alias state0");
    return S.record(state_out);
}

int gocean_ensemble_adjoint_reverse(gocean_ensemble_t* E, int n_steps, const float* dt, const float* state0,
                                    const float* lam_in, float* lam_out, const float* Hi, float* work,
                                    long long work_floats, void* stream) {
    clear_error();
    if (!lam_in || !lam_out) return set_error(GOCEAN_EINVAL, "NULL argument");
    EnsSegment S;
    if (int e = ens_segment(E, n_steps, dt, state0, Hi, work, work_floats, stream, S)) return e;
    return S.reverse(lam_in, lam_out);
}

int gocean_ensemble_adjoint_segment(gocean_ensemble_t* E, int n_steps, const float* dt,
                                    const float* state0, const float* lam_in, float* lam_out,
                                    const float* Hi, float* work, long long work_floats, void* stream) {
    clear_error();
    if (!lam_in || !lam_out) return set_error(GOCEAN_EINVAL, "NULL argument");
    EnsSegment S;
    if (int e = ens_segment(E, n_steps, dt, state0, Hi, work, work_floats, stream, S)) return e;
    if (int e = S.record(nullptr)) return e;
    return S.reverse(lam_in, lam_out);
}

}  // extern "C"
