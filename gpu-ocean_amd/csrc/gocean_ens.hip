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

#include <cstring>
#include <vector>

#include "../../include/gocean.h"
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
    float wind_t, wx0, wx1, wy0, wy1;
    int bc_n, bc_e, bc_s, bc_w;
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
};

template <int BW, int BH, int NT, int MODE>
__global__ __launch_bounds__(NT) void k_ens_substep(EnsLaunch L) {
    __shared__ TileSmem<BW, BH, MODE> S;
    const int t = xcd_band(blockIdx.x, gridDim.x);
    const int m = L.tile_member[t];
    const EnsMember& M = L.members[m];
    SubstepArgs A;
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
    A.wind_t = M.wind_t; A.wx0 = M.wx0; A.wx1 = M.wx1; A.wy0 = M.wy0; A.wy1 = M.wy1;
    A.dt = 0.0f;
    A.bc_n = M.bc_n; A.bc_e = M.bc_e; A.bc_s = M.bc_s; A.bc_w = M.bc_w;
    A.rk_order = L.rk_order; A.rk_step = L.rk_step;
    A.tiles_x = M.tiles_x; A.tiles_y = 0;
    const int lt = t - M.tile_begin;
    const int ty = lt / M.tiles_x, tx = lt - ty * M.tiles_x;
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
    const int n = M.nx * M.ny;
    for (int k = threadIdx.x; k < n; k += blockDim.x) {
        const int j = k / M.nx + 2, i = k % M.nx + 2;
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
    EnsMember* d_members = nullptr;
    int* d_tile_member = nullptr;
    std::vector<EnsMember> members;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    hipStream_t capture_stream = nullptr;
};

namespace gocean {

template <int BW, int BH, int NT, int MODE>
static void ens_launch_t(const EnsLaunch& L, hipStream_t s) {
    hipLaunchKernelGGL((k_ens_substep<BW, BH, NT, MODE>), dim3(L.n_tiles), dim3(NT), 0, s, L);
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
    E->cfg = tile_config < 0 ? kDefaultTileConfig : tile_config;
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
    const TileConfig& tc = (E->mode == kModeZero) ? kTileConfigs[E->cfg] : kTileConfigs[0];
    E->bw = tc.bw; E->bh = tc.bh; E->nt = tc.nt;
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
        M.wind_t = m.wind.t; M.wx0 = m.wind.x0; M.wx1 = m.wind.x1; M.wy0 = m.wind.y0; M.wy1 = m.wind.y1;
        M.bc_n = d.bc_north; M.bc_e = d.bc_east; M.bc_s = d.bc_south; M.bc_w = d.bc_west;
        if (m.state_offset < 0 || m.hi_offset < 0 || m.hm_offset < 0) {
            delete E;
            return set_error(GOCEAN_EINVAL, "member %d: negative buffer offset", k);
        }
    }
    E->n_tiles = (int)tile_member.size();
    hipGetDevice(&E->device);
    if (hipMalloc(&E->d_members, sizeof(EnsMember) * n) != hipSuccess ||
        hipMalloc(&E->d_tile_member, sizeof(int) * E->n_tiles) != hipSuccess) {
        (void)hipGetLastError();
        hipFree(E->d_members);
        hipFree(E->d_tile_member);
        delete E;
        return set_error(GOCEAN_ENOMEM, "device allocation of the member table failed");
    }
    hipMemcpy(E->d_members, E->members.data(), sizeof(EnsMember) * n, hipMemcpyHostToDevice);
    hipMemcpy(E->d_tile_member, tile_member.data(), sizeof(int) * E->n_tiles, hipMemcpyHostToDevice);
    if (int e = check_launch("ensemble table upload")) {
        hipFree(E->d_members);
        hipFree(E->d_tile_member);
        delete E;
        return e;
    }
    *out = E;
    return 0;
}

int gocean_ensemble_destroy(gocean_ensemble_t* E) {
    clear_error();
    if (!E) return 0;
    if (E->graph_exec) hipGraphExecDestroy(E->graph_exec);
    if (E->graph) hipGraphDestroy(E->graph);
    if (E->capture_stream) hipStreamDestroy(E->capture_stream);
    hipFree(E->d_members);
    hipFree(E->d_tile_member);
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
    if (E->graph_exec) { hipGraphExecDestroy(E->graph_exec); E->graph_exec = nullptr; }
    if (E->graph) { hipGraphDestroy(E->graph); E->graph = nullptr; }
    if (!E->capture_stream && hipStreamCreateWithFlags(&E->capture_stream, hipStreamNonBlocking) != hipSuccess)
        return check_launch("capture stream");
    if (hipStreamBeginCapture(E->capture_stream, hipStreamCaptureModeThreadLocal) != hipSuccess)
        return check_launch("begin capture");
    const int e = ens_rk_steps(E, n_steps, state0, state1, state2, Hi, fields, dt, rk_order,
                               E->capture_stream);
    hipGraph_t g = nullptr;
    const hipError_t ee = hipStreamEndCapture(E->capture_stream, &g);
    if (e) { if (g) hipGraphDestroy(g); return e; }
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

}  // extern "C"
