// gocean_drift.hip -- passive drifters and the observation operator.
//
// Drifters are advected by the bilinearly interpolated water velocity after
// each model step (CDKLM16.drifterStep, gpu_ocean/SWESimulators/
// CDKLM16.py:600-610 -> passiveDrifterKernel, gpu_kernels/driftKernels.cu:
// 108-210); data assimilation observes hu, hv at drifter positions
// (observeUnderlyingFlow, gpu_kernels/observationKernels.cu:29-82), per
// ensemble member in the reference (BaseOceanStateEnsemble.py:105-136).
//
// One thread per drifter.  Positions are (num + 1) x 2 floats: the drifters,
// then the observation position, which passiveDrifterKernel also moves.
// The batched forms cover every member of an ensemble with one launch: each
// block of kDriftNT drifters belongs to one member (block -> member map).

#include <algorithm>
#include <vector>

#include "gocean_internal.hpp"

namespace gocean {
namespace {

constexpr int kDriftNT = 64;

struct DriftMember {
    int nx, ny, pitch, periodic_ns, periodic_ew, x_zero, y_zero, num;
    float dx, dy, sensitivity, wind_factor, wind_t, wx0, wx1, wy0, wy1;
    long long eta_off, hu_off, hv_off, hm_off, pos_off, obs_off;
    int block_begin;
};

struct DriftLaunch {
    const DriftMember* members;   // device table, or NULL for `one`
    const int* map;
    DriftMember one;
    const float* eta;             // bases; member offsets eta_off, hu_off, hv_off
    const float* hu;
    const float* hv;
    const float* Hm;
    float* pos;
    const float* obs_pos;         // observe: positions to sample (NULL: pos)
    float* obs;
    const float* dt;              // per member (device) or NULL: dt_host
    float dt_host;
};

template <bool BATCH>
__device__ __forceinline__ const DriftMember& dmember(const DriftLaunch& L) {
    if constexpr (BATCH) return L.members[L.map[blockIdx.x]];
    else return L.one;
}

// h = Hm + eta, u = hu / h at a ghosted cell (driftKernels.cu:54-88)
__device__ __forceinline__ float velocity(const float* q, const float* eta, const float* Hm, int pitch,
                                          int j, int i) {
    const long long o = (long long)j * pitch + i;
    const float h = Hm[o] + eta[o];
    return q[o] / h;
}

// One drift step of one drifter, operation by operation as
// passiveDrifterKernel: cell index from ceil(x/dx), fraction from
// floor(x/dx) (the velocity is sampled one cell east/north of the drifter
// away from faces), uniform wind lerped in time, periodic wrap.
template <bool BATCH>
__global__ __launch_bounds__(kDriftNT) void k_drift(DriftLaunch L) {
    const DriftMember& M = dmember<BATCH>(L);
    const int k = (blockIdx.x - M.block_begin) * kDriftNT + threadIdx.x;
    if (k >= M.num + 1) return;
    const float dt = L.dt ? L.dt[BATCH ? L.map[blockIdx.x] : 0] : L.dt_host;
    const float* eta = L.eta + M.eta_off;
    const float* hu = L.hu + M.hu_off;
    const float* hv = L.hv + M.hv_off;
    const float* Hm = L.Hm + M.hm_off;
    float* p = L.pos + M.pos_off + 2ll * k;
    float x = p[0], y = p[1];
    const float qx = x / M.dx, qy = y / M.dy;
    const int cx = (int)(ceilf(qx) + (float)M.x_zero);
    const int cy = (int)(ceilf(qy) + (float)M.y_zero);
    const float fx = qx - floorf(qx);
    const float fy = qy - floorf(qy);
    // the reference reads whatever cells the position maps to; positions
    // more than a ghost cell outside the domain (or NaN) are clamped to the
    // array here so that every read stays inside it
    const int x0 = min(max(fx < 0.5f ? cx - 1 : cx, 0), M.nx + 2);
    const float xf = fx < 0.5f ? fx + 0.5f : fx - 0.5f;
    const int y0 = min(max(fy < 0.5f ? cy - 1 : cy, 0), M.ny + 2);
    const float yf = fy < 0.5f ? fy + 0.5f : fy - 0.5f;
    const int x1 = x0 + 1, y1 = y0 + 1;
    const float u00 = velocity(hu, eta, Hm, M.pitch, y0, x0), u10 = velocity(hu, eta, Hm, M.pitch, y0, x1);
    const float u01 = velocity(hu, eta, Hm, M.pitch, y1, x0), u11 = velocity(hu, eta, Hm, M.pitch, y1, x1);
    const float v00 = velocity(hv, eta, Hm, M.pitch, y0, x0), v10 = velocity(hv, eta, Hm, M.pitch, y0, x1);
    const float v01 = velocity(hv, eta, Hm, M.pitch, y1, x0), v11 = velocity(hv, eta, Hm, M.pitch, y1, x1);
    const float uy0 = (1.0f - xf) * u00 + xf * u10;
    const float uy1 = (1.0f - xf) * u01 + xf * u11;
    const float vy0 = (1.0f - xf) * v00 + xf * v10;
    const float vy1 = (1.0f - xf) * v01 + xf * v11;
    float u = (1.0f - yf) * uy0 + yf * uy1;
    float v = (1.0f - yf) * vy0 + yf * vy1;
    if (M.wind_factor != 0.0f) {
        u = u + (M.wind_t * M.wx1 + (1.0f - M.wind_t) * M.wx0) * M.wind_factor;
        v = v + (M.wind_t * M.wy1 + (1.0f - M.wind_t) * M.wy0) * M.wind_factor;
    }
    x = x + (M.sensitivity * u) * dt;
    y = y + (M.sensitivity * v) * dt;
    const float Lx = (float)M.nx * M.dx, Ly = (float)M.ny * M.dy;
    if (M.periodic_ew && x < 0.0f) x = x + Lx;
    if (M.periodic_ew && x > Lx) x = x - Lx;
    if (M.periodic_ns && y < 0.0f) y = y + Ly;
    if (M.periodic_ns && y > Ly) y = y - Ly;
    p[0] = x;
    p[1] = y;
}

// enforceBoundaryConditions (driftKernels.cu:214-250)
__global__ __launch_bounds__(kDriftNT) void k_drift_bc(float Lx, float Ly, int pns, int pew, int n,
                                                       float* pos) {
    const int k = blockIdx.x * kDriftNT + threadIdx.x;
    if (k >= n + 1) return;
    float x = pos[2 * k], y = pos[2 * k + 1];
    if (pew && x < 0.0f) x = x + Lx;
    if (pew && x > Lx) x = x - Lx;
    if (pns && y < 0.0f) y = y + Ly;
    if (pns && y > Ly) y = y - Ly;
    pos[2 * k] = x;
    pos[2 * k + 1] = y;
}

// observeUnderlyingFlow: obs[k] = (hu, hv) of the cell holding position k
template <bool BATCH>
__global__ __launch_bounds__(kDriftNT) void k_observe(DriftLaunch L) {
    const DriftMember& M = dmember<BATCH>(L);
    const int k = (blockIdx.x - M.block_begin) * kDriftNT + threadIdx.x;
    if (k >= M.num) return;
    const float* p = (L.obs_pos ? L.obs_pos : L.pos + M.pos_off) + 2ll * k;
    const int i = min(max((int)(floorf(p[0] / M.dx) + (float)M.x_zero), 0), M.nx + 3);
    const int j = min(max((int)(floorf(p[1] / M.dy) + (float)M.y_zero), 0), M.ny + 3);
    const long long o = (long long)j * M.pitch + i;
    float* out = L.obs + M.obs_off + 2ll * k;
    out[0] = L.hu[M.hu_off + o];
    out[1] = L.hv[M.hv_off + o];
}

int fill_drift(const gocean_domain_t* d, int num, float sensitivity, const gocean_wind_t* wind,
               float wind_factor, int x_zero, int y_zero, DriftMember& M) {
    if (int e = validate_domain(d)) return e;
    if (num < 0) return set_error(GOCEAN_EINVAL, "number of drifters %d < 0", num);
    if (wind_factor != 0.0f && !wind)
        return set_error(GOCEAN_EINVAL, "a wind drift factor needs the wind");
    if (wind_factor != 0.0f && wind->grid)
        return set_error(GOCEAN_EINVAL, "drifters take uniform wind only (wind grid given)");
    if (x_zero < 0 || y_zero < 0 || x_zero > 2 || y_zero > 2)
        return set_error(GOCEAN_EINVAL, "reference cells (%d, %d) outside the two ghost cells", x_zero, y_zero);
    M = DriftMember{};
    M.nx = d->nx;
    M.ny = d->ny;
    M.pitch = d->pitch;
    M.periodic_ns = d->bc_north == GOCEAN_BC_PERIODIC;
    M.periodic_ew = d->bc_east == GOCEAN_BC_PERIODIC;
    M.x_zero = x_zero;
    M.y_zero = y_zero;
    M.num = num;
    M.dx = d->dx;
    M.dy = d->dy;
    M.sensitivity = sensitivity;
    M.wind_factor = wind_factor;
    if (wind) {
        M.wind_t = wind->t;
        M.wx0 = wind->x0;
        M.wx1 = wind->x1;
        M.wy0 = wind->y0;
        M.wy1 = wind->y1;
    }
    return 0;
}

}  // namespace
}  // namespace gocean

using namespace gocean;

int gocean_drift(const gocean_domain_t* d, int num_drifters, float* positions, const float* eta,
                 const float* hu, const float* hv, const float* Hm, float dt, float sensitivity,
                 const gocean_wind_t* wind, float wind_drift_factor, int x_zero, int y_zero, void* stream) {
    clear_error();
    DriftLaunch L{};
    if (int e = fill_drift(d, num_drifters, sensitivity, wind, wind_drift_factor, x_zero, y_zero, L.one))
        return e;
    if (!positions || !eta || !hu || !hv || !Hm) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    L.eta = eta;
    L.hu = hu;
    L.hv = hv;
    L.Hm = Hm;
    L.pos = positions;
    L.dt_host = dt;
    const int blocks = (num_drifters + 1 + kDriftNT - 1) / kDriftNT;
    hipLaunchKernelGGL(k_drift<false>, dim3(blocks), dim3(kDriftNT), 0, (hipStream_t)stream, L);
    return check_launch("k_drift");
}

int gocean_drifters_enforce_bc(float domain_size_x, float domain_size_y, int periodic_ns,
                               int periodic_ew, int num_drifters, float* positions, void* stream) {
    clear_error();
    if (!positions || num_drifters < 0) return set_error(GOCEAN_EINVAL, "positions NULL or count < 0");
    const int blocks = (num_drifters + 1 + kDriftNT - 1) / kDriftNT;
    hipLaunchKernelGGL(k_drift_bc, dim3(blocks), dim3(kDriftNT), 0, (hipStream_t)stream, domain_size_x,
                       domain_size_y, periodic_ns, periodic_ew, num_drifters, positions);
    return check_launch("k_drift_bc");
}

int gocean_observe_flow(const gocean_domain_t* d, int num_drifters, const float* positions,
                        const float* hu, const float* hv, float* obs, int x_zero, int y_zero,
                        void* stream) {
    clear_error();
    DriftLaunch L{};
    if (int e = fill_drift(d, num_drifters, 1.0f, nullptr, 0.0f, x_zero, y_zero, L.one)) return e;
    if (!positions || !hu || !hv || !obs) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    if (num_drifters == 0) return 0;
    L.hu = hu;
    L.hv = hv;
    L.obs_pos = positions;
    L.obs = obs;
    const int blocks = (num_drifters + kDriftNT - 1) / kDriftNT;
    hipLaunchKernelGGL(k_observe<false>, dim3(blocks), dim3(kDriftNT), 0, (hipStream_t)stream, L);
    return check_launch("k_observe");
}

// ---------------------------------------------------------------------------
// batched
// ---------------------------------------------------------------------------
struct gocean_drift_batch {
    int n = 0, n_blocks = 0, n_obs_blocks = 0;
    std::vector<DriftMember> host;
    DriftMember* d_members = nullptr;
    int* d_map = nullptr;
    int* d_obs_map = nullptr;
    DriftMember* d_obs_members = nullptr;
};

int gocean_drift_batch_create(const gocean_drift_member_t* members, int n, gocean_drift_batch_t** out) {
    clear_error();
    if (!members || n < 1 || !out) return set_error(GOCEAN_EINVAL, "need members, n >= 1 and out");
    auto* B = new gocean_drift_batch;
    B->n = n;
    B->host.resize(n);
    std::vector<int> map, omap;
    std::vector<DriftMember> obs_members(n);
    for (int m = 0; m < n; ++m) {
        const gocean_drift_member_t& G = members[m];
        DriftMember& M = B->host[m];
        gocean_wind_t w = G.wind;
        if (int e = fill_drift(&G.domain, G.num_drifters, G.sensitivity, &w, G.wind_drift_factor, 2, 2, M)) {
            delete B;
            return set_error(e, "member %d: %s", m, last_error());
        }
        const long long plane = (long long)(G.domain.ny + 4) * G.domain.pitch;
        M.eta_off = G.state_offset;
        M.hu_off = G.state_offset + plane;
        M.hv_off = G.state_offset + 2 * plane;
        M.hm_off = G.hm_offset;
        M.pos_off = G.positions_offset;
        M.obs_off = G.observations_offset;
        M.block_begin = (int)map.size();
        map.insert(map.end(), (G.num_drifters + 1 + kDriftNT - 1) / kDriftNT, m);
        obs_members[m] = M;
        obs_members[m].block_begin = (int)omap.size();
        omap.insert(omap.end(), (G.num_drifters + kDriftNT - 1) / kDriftNT, m);
    }
    B->n_blocks = (int)map.size();
    B->n_obs_blocks = (int)omap.size();
    bool ok = hipMalloc(&B->d_members, sizeof(DriftMember) * n) == hipSuccess &&
              hipMalloc(&B->d_obs_members, sizeof(DriftMember) * n) == hipSuccess &&
              hipMalloc(&B->d_map, sizeof(int) * std::max<size_t>(1, map.size())) == hipSuccess &&
              hipMalloc(&B->d_obs_map, sizeof(int) * std::max<size_t>(1, omap.size())) == hipSuccess;
    ok = ok && hipMemcpy(B->d_members, B->host.data(), sizeof(DriftMember) * n, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->d_obs_members, obs_members.data(), sizeof(DriftMember) * n, hipMemcpyHostToDevice) == hipSuccess &&
         (map.empty() || hipMemcpy(B->d_map, map.data(), sizeof(int) * map.size(), hipMemcpyHostToDevice) == hipSuccess) &&
         (omap.empty() || hipMemcpy(B->d_obs_map, omap.data(), sizeof(int) * omap.size(), hipMemcpyHostToDevice) == hipSuccess);
    if (!ok) {
        gocean_drift_batch_destroy(B);
        (void)hipGetLastError();
        return set_error(GOCEAN_ENOMEM, "device allocation of the drifter member table failed");
    }
    *out = B;
    return 0;
}

int gocean_drift_batch_destroy(gocean_drift_batch_t* B) {
    if (!B) return 0;
    (void)hipFree(B->d_members);
    (void)hipFree(B->d_obs_members);
    (void)hipFree(B->d_map);
    (void)hipFree(B->d_obs_map);
    delete B;
    return 0;
}

int gocean_drift_batch_step(gocean_drift_batch_t* B, const float* state, const float* Hm, float* positions,
                            const float* dt, void* stream) {
    clear_error();
    if (!B || !state || !Hm || !positions || !dt) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    if (B->n_blocks == 0) return 0;
    DriftLaunch L{};
    L.members = B->d_members;
    L.map = B->d_map;
    L.eta = L.hu = L.hv = state;
    L.Hm = Hm;
    L.pos = positions;
    L.dt = dt;
    hipLaunchKernelGGL(k_drift<true>, dim3(B->n_blocks), dim3(kDriftNT), 0, (hipStream_t)stream, L);
    return check_launch("k_drift (batch)");
}

int gocean_observe_flow_batch(gocean_drift_batch_t* B, const float* state, const float* positions,
                              int shared, float* observations, void* stream) {
    clear_error();
    if (!B || !state || !positions || !observations) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    if (B->n_obs_blocks == 0) return 0;
    DriftLaunch L{};
    L.members = B->d_obs_members;
    L.map = B->d_obs_map;
    L.eta = L.hu = L.hv = state;
    L.pos = (float*)positions;
    L.obs_pos = shared ? positions : nullptr;
    L.obs = observations;
    hipLaunchKernelGGL(k_observe<true>, dim3(B->n_obs_blocks), dim3(kDriftNT), 0, (hipStream_t)stream, L);
    return check_launch("k_observe (batch)");
}
