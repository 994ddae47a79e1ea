// gocean_fwd.hip -- forward CDKLM16 kernels and their C-ABI launchers.
//
// Kernels: bathymetry halo + cell depth, periodic ghost exchange, the CFL
// time-step reduction and the tiled RK substep (cdklm16_tile.hpp).  See
// include/gocean.h for the contract of each entry point.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <unordered_map>

#include "../../include/gocean.h"
#include "cdklm16_march.hpp"
#include "cdklm16_tile.hpp"
#include "gocean_internal.hpp"

namespace gocean {

// ---------------------------------------------------------------------------
// error reporting shared by every translation unit
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;

int set_error(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code == 0 ? -1 : code;
}

void clear_error() { g_last_error.clear(); }

const char* last_error() { return g_last_error.c_str(); }

int check_launch(const char* what) {
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error((int)e, "%s: %s", what, hipGetErrorString(e));
    return 0;
}

int validate_domain(const gocean_domain_t* d) {
    if (!d) return set_error(GOCEAN_EINVAL, "domain descriptor is NULL");
    if (d->nx < 2 || d->ny < 2)
        return set_error(GOCEAN_EINVAL, "domain too small: nx=%d ny=%d (need >= 2)", d->nx, d->ny);
    if (d->pitch < d->nx + 4)
        return set_error(GOCEAN_EINVAL, "pitch %d < nx+4 = %d", d->pitch, d->nx + 4);
    if (d->hi_pitch < d->nx + 5)
        return set_error(GOCEAN_EINVAL, "hi_pitch %d < nx+5 = %d", d->hi_pitch, d->nx + 5);
    const int bcs[4] = {d->bc_north, d->bc_east, d->bc_south, d->bc_west};
    for (int k = 0; k < 4; ++k)
        if (bcs[k] != kWall && bcs[k] != kPeriodic)
            return set_error(GOCEAN_EINVAL,
                             "boundary condition %d not supported (1 wall, 2 periodic)", bcs[k]);
    if ((d->bc_north == kPeriodic) != (d->bc_south == kPeriodic))
        return set_error(GOCEAN_EINVAL, "periodic north/south must be paired");
    if ((d->bc_east == kPeriodic) != (d->bc_west == kPeriodic))
        return set_error(GOCEAN_EINVAL, "periodic east/west must be paired");
    if (d->rk_order < 1 || d->rk_order > 3)
        return set_error(GOCEAN_EINVAL, "rk_order %d not in {1,2,3}", d->rk_order);
    if (d->rk_order == 3 && d->r != 0.0f)
        return set_error(GOCEAN_EINVAL, "rk_order 3 requires r == 0");
    if (!(d->dx > 0.0f) || !(d->dy > 0.0f) || !(d->g > 0.0f))
        return set_error(GOCEAN_EINVAL, "dx, dy and g must be positive");
    return 0;
}

// ---------------------------------------------------------------------------
// division by a launch constant with one residual correction
// ---------------------------------------------------------------------------
// For q = RN(n r), r = RN(1/d), one correction RN(q + RN(n - d q) r) is the
// correctly rounded n/d for most d but is not guaranteed for all.  Scaling n
// by a power of two scales q, the residual and the quotient exactly, so the
// check over the 2^23 significands of n in [1, 2) decides it for every n of
// normal magnitude (and -n by symmetry).  Done once per divisor value on the
// host, with the hardware FMA when the CPU has one.
namespace {

__attribute__((target("fma"))) bool one_correction_scan_fma(float d) {
    const float r = 1.0f / d;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        uint32_t bits = 0x3f800000u | m;
        float n;
        std::memcpy(&n, &bits, 4);
        const float q = n * r;
        const float e = __builtin_fmaf(-d, q, n);
        const float q1 = __builtin_fmaf(e, r, q);
        if (q1 != n / d) return false;
    }
    return true;
}

bool one_correction_scan_libm(float d) {
    const float r = 1.0f / d;
    for (uint32_t m = 0; m < (1u << 23); ++m) {
        uint32_t bits = 0x3f800000u | m;
        float n;
        std::memcpy(&n, &bits, 4);
        const float q = n * r;
        const float e = std::fmaf(-d, q, n);
        const float q1 = std::fmaf(e, r, q);
        if (q1 != n / d) return false;
    }
    return true;
}

std::mutex g_div_mu;
std::unordered_map<uint32_t, bool> g_div_ok;

}  // namespace

bool one_correction_exact(float d) {
    if (!(d > 0.0f) || !std::isnormal(d) || !std::isnormal(1.0f / d)) return false;
    uint32_t key;
    std::memcpy(&key, &d, 4);
    {
        std::lock_guard<std::mutex> lock(g_div_mu);
        auto it = g_div_ok.find(key);
        if (it != g_div_ok.end()) return it->second;
    }
    const bool ok = __builtin_cpu_supports("fma") ? one_correction_scan_fma(d) : one_correction_scan_libm(d);
    std::lock_guard<std::mutex> lock(g_div_mu);
    g_div_ok[key] = ok;
    return ok;
}

bool fast_arith_exact(float g, float dx, float dy, float cutoff) {
    // sqrt_newton equals sqrt_rn for arguments >= 2^-60; the wave speeds take
    // it only for depths above the cutoff, whose g h rounds to >= RN(g cutoff)
    const float lo = g * cutoff;
    if (!(g > 0.0f) || !std::isfinite(g) || !(lo >= 0x1p-60f)) return false;
    return one_correction_exact(2.0f * g) && one_correction_exact(dx) && one_correction_exact(dy);
}

void fill_args(const gocean_domain_t* d, SubstepArgs& A) {
    std::memset(&A, 0, sizeof(A));
    A.nx = d->nx; A.ny = d->ny; A.pitch = d->pitch; A.hi_pitch = d->hi_pitch;
    A.dx = d->dx; A.dy = d->dy; A.g = d->g; A.r = d->r; A.theta = d->theta;
    A.desing_eps = d->desing_eps; A.flux_eps = d->flux_slope_eps;
    A.cutoff = d->depth_cutoff; A.land = d->land_value;
    A.f = d->f; A.north_x = d->north_x; A.north_y = d->north_y;
    A.set_flags(d->bc_north, d->bc_east, d->bc_south, d->bc_west, d->rk_order, 0);
    substep_derive(A);
    A.set_fast_arith(fast_arith_exact(A.g, A.dx, A.dy, A.cutoff));
}

// ---------------------------------------------------------------------------
// tile configurations
// ---------------------------------------------------------------------------
const TileConfig kTileConfigs[] = {
    {32, 8, 256, "32x8/256"},
    {32, 16, 256, "32x16/256"},
    {64, 8, 256, "64x8/256"},
    {16, 16, 256, "16x16/256"},
    {64, 4, 256, "64x4/256"},
    {32, 16, 512, "32x16/512"},
    // 14 rows: the slope and face phases fit two passes of 512 threads
    // (16 rows need a third pass for 96 / 48 threads)
    {32, 14, 512, "32x14/512"},
    // column-marching strips (zero rotation): 60 output columns per
    // wavefront, BH rows; four independent wavefronts per workgroup
    {60, 8, 256, "march60x8/256"},
    {60, 16, 256, "march60x16/256"},
    {60, 0, 256, "march60xauto/256"},    // rows per strip chosen to fill whole rounds
};
const int kNumTileConfigs = sizeof(kTileConfigs) / sizeof(kTileConfigs[0]);
const int kDefaultTileConfig = 1;
const int kMarchAutoConfig = 9;

// 512-thread tiles: 8 waves per SIMD (four 38 KB workgroups per CU) at the
// price of a few scalar-register spills into vector-register lanes; measured
// faster than the 7 waves the registers otherwise allow
template <int BW, int BH, int NT, int MODE>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(NT >= 512 ? 8 : 1))) void k_substep(SubstepArgs A) {
    __shared__ TileSmem<BW, BH, MODE> S;
    const int t = xcd_band(blockIdx.x, gridDim.x);
    const int ty = t / A.tiles_x, tx = t - ty * A.tiles_x;
    if (tile_is_interior<BW, BH>(A, tx, ty))
        cdklm_tile<BW, BH, NT, MODE, false, false>(A, tx, ty, S);
    else
        cdklm_tile<BW, BH, NT, MODE>(A, tx, ty, S);
}

// column-marching substep: wavefront w of the grid takes strip (sx, sy),
// sx fastest, workgroups banded over the XCDs
template <bool FAST, bool WGRID>
__global__ __launch_bounds__(256) void k_march(SubstepArgs A) {
    const int wave = xcd_band(blockIdx.x, gridDim.x) * 4 + march::wave_in_group();
    if (wave >= A.tiles_x * A.tiles_y) return;
    const int sy = wave / A.tiles_x, sx = wave - sy * A.tiles_x;
    march::march_strip<FAST, WGRID>(A, sx, sy, A.march_bh);
}

// waves of a kernel that the device holds at once: 4 SIMDs per CU, each
// holding as many waves as its 512 vector registers per lane allow
int resident_waves(const void* kernel) {
    int regs = 128;
    hipFuncAttributes fa;
    if (hipFuncGetAttributes(&fa, kernel) == hipSuccess && fa.numRegs > 0) regs = fa.numRegs;
    else (void)hipGetLastError();
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
        (void)hipGetLastError();
        cus = 256;
    }
    const int alloc = (regs + 7) / 8 * 8;
    const int per_simd = std::max(1, std::min(8, 512 / alloc));
    return cus * 4 * per_simd;
}

// Rows per strip for the column-marching kernel over domains of the given
// strip counts and heights: each wave's time grows with its rows plus about
// three rows of prologue, and waves run in rounds of `resident`; pick the
// row count (8..256) with the least rounds * (rows + 3).
int march_rows(const int* strips, const int* ny, int n, int resident) {
    int best = 8;
    double best_cost = 1e300;
    for (int bh = 8; bh <= 256; ++bh) {
        long long waves = 0;
        for (int k = 0; k < n; ++k) waves += (long long)strips[k] * ((ny[k] + bh - 1) / bh);
        const long long rounds = (waves + resident - 1) / resident;
        const double cost = (double)rounds * (bh + 3);
        if (cost < best_cost - 1e-9) { best_cost = cost; best = bh; }
    }
    return best;
}

// the marching kernel for the launch's arithmetic and wind variants
static const void* march_kernel(bool fast, bool wgrid) {
    if (fast) return wgrid ? (const void*)k_march<true, true> : (const void*)k_march<true, false>;
    return wgrid ? (const void*)k_march<false, true> : (const void*)k_march<false, false>;
}

// sqrt_newton against sqrt_rn over the float bit patterns lo .. hi
__global__ __launch_bounds__(256) void k_sqrt_check(uint32_t lo, uint32_t hi, unsigned long long* count) {
    const uint64_t n = (uint64_t)hi - lo + 1;
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    unsigned long long bad = 0;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) {
        const float x = __uint_as_float(lo + (uint32_t)i);
        bad += __float_as_uint(sqrt_newton(x)) != __float_as_uint(sqrt_rn(x));
    }
    if (bad) atomicAdd(count, bad);
}

// bh > 0: fixed rows per strip; 0: balanced by march_rows
static void launch_march(SubstepArgs A, int bh, hipStream_t s) {
    A.tiles_x = (A.nx + march::kOut - 1) / march::kOut;
    const bool fast = A.fast_arith(), wgrid = A.wind_grid != nullptr;
    if (bh <= 0) {
        static int resident[4] = {0, 0, 0, 0};
        const int k = (fast ? 2 : 0) + (wgrid ? 1 : 0);
        if (!resident[k]) resident[k] = resident_waves(march_kernel(fast, wgrid));
        bh = march_rows(&A.tiles_x, &A.ny, 1, resident[k]);
    }
    A.march_bh = bh;
    A.tiles_y = (A.ny + bh - 1) / bh;
    const int waves = A.tiles_x * A.tiles_y;
    const dim3 grid((waves + 3) / 4), block(256);
    if (fast) {
        if (wgrid) hipLaunchKernelGGL((k_march<true, true>), grid, block, 0, s, A);
        else hipLaunchKernelGGL((k_march<true, false>), grid, block, 0, s, A);
    } else {
        if (wgrid) hipLaunchKernelGGL((k_march<false, true>), grid, block, 0, s, A);
        else hipLaunchKernelGGL((k_march<false, false>), grid, block, 0, s, A);
    }
}

template <int BW, int BH, int NT, int MODE>
static void launch_substep_t(SubstepArgs A, hipStream_t s) {
    A.tiles_x = (A.nx + BW - 1) / BW;
    A.tiles_y = (A.ny + BH - 1) / BH;
    const int nb = A.tiles_x * A.tiles_y;
    hipLaunchKernelGGL((k_substep<BW, BH, NT, MODE>), dim3(nb), dim3(NT), 0, s, A);
}

// recording variant (zero rotation only): also writes the branch record
template <int BW, int BH, int NT>
__global__ __launch_bounds__(NT) void k_substep_tape(SubstepArgs A) {
    __shared__ TileSmem<BW, BH, kModeZero, true> S;
    const int t = xcd_band(blockIdx.x, gridDim.x);
    const int ty = t / A.tiles_x, tx = t - ty * A.tiles_x;
    if (tile_is_interior<BW, BH>(A, tx, ty))
        cdklm_tile<BW, BH, NT, kModeZero, true, false>(A, tx, ty, S);
    else
        cdklm_tile<BW, BH, NT, kModeZero, true>(A, tx, ty, S);
}

template <int BW, int BH, int NT>
static void launch_tape_t(SubstepArgs A, hipStream_t s) {
    A.tiles_x = (A.nx + BW - 1) / BW;
    A.tiles_y = (A.ny + BH - 1) / BH;
    hipLaunchKernelGGL((k_substep_tape<BW, BH, NT>), dim3(A.tiles_x * A.tiles_y), dim3(NT), 0, s, A);
}

// column-marching recording variant (uniform wind only: the adjoint's scope)
template <bool FAST>
__global__ __launch_bounds__(256) void k_march_tape(SubstepArgs A) {
    const int wave = xcd_band(blockIdx.x, gridDim.x) * 4 + march::wave_in_group();
    if (wave >= A.tiles_x * A.tiles_y) return;
    const int sy = wave / A.tiles_x, sx = wave - sy * A.tiles_x;
    march::march_strip<FAST, false, true>(A, sx, sy, A.march_bh);
}

// bh > 0: fixed rows per strip; 0: balanced by march_rows
static void launch_march_tape(SubstepArgs A, int bh, hipStream_t s) {
    A.tiles_x = (A.nx + march::kOut - 1) / march::kOut;
    const bool fast = A.fast_arith();
    if (bh <= 0) {
        static int resident[2] = {0, 0};
        const void* k = fast ? (const void*)k_march_tape<true> : (const void*)k_march_tape<false>;
        if (!resident[fast]) resident[fast] = resident_waves(k);
        bh = march_rows(&A.tiles_x, &A.ny, 1, resident[fast]);
    }
    A.march_bh = bh;
    A.tiles_y = (A.ny + bh - 1) / bh;
    const int waves = A.tiles_x * A.tiles_y;
    if (fast) hipLaunchKernelGGL((k_march_tape<true>), dim3((waves + 3) / 4), dim3(256), 0, s, A);
    else hipLaunchKernelGGL((k_march_tape<false>), dim3((waves + 3) / 4), dim3(256), 0, s, A);
}

// tile configuration of the recording forward (process-wide; the record is
// laid out per cell, so every configuration writes the same record); -1:
// the column-marching kernel from a million cells on, 32x16/512 below
static int g_record_cfg = -1;

int launch_substep_tape(SubstepArgs A, hipStream_t s) {
    int cfg = g_record_cfg;
    if (cfg < 0) cfg = default_tile_config((long long)A.nx * A.ny) >= kFirstMarchConfig ? kMarchAutoConfig : 5;
    if (cfg >= kFirstMarchConfig && !march_fits(A.ny, A.pitch, A.hi_pitch)) cfg = 5;
    switch (cfg) {
        case 0: launch_tape_t<32, 8, 256>(A, s); break;
        case 1: launch_tape_t<32, 16, 256>(A, s); break;
        case 2: launch_tape_t<64, 8, 256>(A, s); break;
        case 3: launch_tape_t<16, 16, 256>(A, s); break;
        case 4: launch_tape_t<64, 4, 256>(A, s); break;
        case 6: launch_tape_t<32, 14, 512>(A, s); break;
        case 7: launch_march_tape(A, 8, s); break;
        case 8: launch_march_tape(A, 16, s); break;
        case 9: launch_march_tape(A, 0, s); break;
        default: launch_tape_t<32, 16, 512>(A, s); break;
    }
    return check_launch("recording substep launch");
}

int substep_mode(const SubstepArgs& A, bool field) {
    if (field) return kModeField;
    return (A.north_x == 0.0f && A.north_y == 1.0f) ? kModeZero : kModeConst;
}

// default configuration: the column-marching kernel with balanced rows
// per strip from a million cells on (where it is faster), else the 2-D tile
int default_tile_config(long long cells) { return cells >= (1LL << 20) ? kMarchAutoConfig : kDefaultTileConfig; }

// the column-marching kernel addresses a plane with 32-bit byte offsets
bool march_fits(int ny, int pitch, int hi_pitch) {
    return (uint64_t)(ny + 4) * (uint64_t)pitch * 4u <= 0xFFFFFFFFull &&
           (uint64_t)(ny + 5) * (uint64_t)hi_pitch * 4u <= 0xFFFFFFFFull;
}

int launch_substep(const SubstepArgs& A, int cfg, bool field, hipStream_t s) {
    const bool fits = march_fits(A.ny, A.pitch, A.hi_pitch);
    if (cfg < 0) {
        cfg = default_tile_config((long long)A.nx * A.ny);
        if (cfg >= kFirstMarchConfig && !fits) cfg = kDefaultTileConfig;
    }
    if (cfg >= kNumTileConfigs)
        return set_error(GOCEAN_EINVAL, "tile configuration %d out of range [0,%d)", cfg, kNumTileConfigs);
    if (cfg >= kFirstMarchConfig && !fits)
        return set_error(GOCEAN_EINVAL, "tile configuration %d (column marching) needs planes below 4 GiB", cfg);
    const int mode = substep_mode(A, field);
    if (mode == kModeField) {
        // per-cell Coriolis / north fields: one general configuration
        launch_substep_t<32, 8, 256, kModeField>(A, s);
    } else if (mode == kModeConst) {
        launch_substep_t<32, 8, 256, kModeConst>(A, s);
    } else {
        switch (cfg) {
            case 0: launch_substep_t<32, 8, 256, kModeZero>(A, s); break;
            case 1: launch_substep_t<32, 16, 256, kModeZero>(A, s); break;
            case 2: launch_substep_t<64, 8, 256, kModeZero>(A, s); break;
            case 3: launch_substep_t<16, 16, 256, kModeZero>(A, s); break;
            case 4: launch_substep_t<64, 4, 256, kModeZero>(A, s); break;
            case 5: launch_substep_t<32, 16, 512, kModeZero>(A, s); break;
            case 6: launch_substep_t<32, 14, 512, kModeZero>(A, s); break;
            case 7: launch_march(A, 8, s); break;
            case 8: launch_march(A, 16, s); break;
            case 9: launch_march(A, 0, s); break;
        }
    }
    return check_launch("cdklm substep launch");
}

// ---------------------------------------------------------------------------
// bathymetry: Hi halo (boundary_kernels.cu:121-246) and Hm (initBm :38-107)
// ---------------------------------------------------------------------------
__global__ void k_hi_halo_ns(float* Hi, int nx, int ny, int hp, int periodic) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;   // column
    if (i > nx + 4) return;
    if (periodic) {
        Hi[0 * hp + i] = Hi[(size_t)ny * hp + i];
        Hi[1 * hp + i] = Hi[(size_t)(ny + 1) * hp + i];
        for (int j = ny + 2; j <= ny + 4; ++j) Hi[(size_t)j * hp + i] = Hi[(size_t)(j - ny) * hp + i];
    } else {
        for (int j = 0; j < 2; ++j) Hi[(size_t)j * hp + i] = Hi[(size_t)(4 - j) * hp + i];
        for (int j = 0; j < 2; ++j) Hi[(size_t)(ny + 4 - j) * hp + i] = Hi[(size_t)(ny + j) * hp + i];
    }
}

__global__ void k_hi_halo_ew(float* Hi, int nx, int ny, int hp, int periodic) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;   // row
    if (j > ny + 4) return;
    float* row = Hi + (size_t)j * hp;
    if (periodic) {
        row[0] = row[nx];
        row[1] = row[nx + 1];
        for (int i = nx + 2; i <= nx + 4; ++i) row[i] = row[i - nx];
    } else {
        for (int i = 0; i < 2; ++i) row[i] = row[4 - i];
        for (int i = 0; i < 2; ++i) row[nx + 4 - i] = row[nx + i];
    }
}

__global__ void k_init_bm(const float* Hi, float* Hm, int nx, int ny, int hp, int pitch, float land) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    if (i >= nx + 4 || j >= ny + 4) return;
    auto flag = [&](float v) { return (fabsf(v - land) <= kDryEps) ? kDryFlag : v; };
    const float a = flag(Hi[(size_t)j * hp + i]);
    const float b = flag(Hi[(size_t)(j + 1) * hp + i]);
    const float c = flag(Hi[(size_t)j * hp + i + 1]);
    const float d = flag(Hi[(size_t)(j + 1) * hp + i + 1]);
    const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag)
                    - (float)(c == kDryFlag) - (float)(d == kDryFlag);
    Hm[(size_t)j * pitch + i] = (wet == 0.0f) ? land : (((a + b) + c) + d) / wet;
}

// ---------------------------------------------------------------------------
// periodic ghost exchange (boundary_kernels.cu:35-110), NS then EW
// ---------------------------------------------------------------------------
__global__ void k_periodic_ns(float* e, float* u, float* v, int nx, int ny, int pitch) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;        // 0..3 -> ghost rows 0,1,ny+2,ny+3
    if (i > nx + 3) return;
    const int j = g < 2 ? g : ny + g;
    const int src = g < 2 ? j + ny : j - ny;
    const size_t d = (size_t)j * pitch + i, s = (size_t)src * pitch + i;
    e[d] = e[s]; u[d] = u[s]; v[d] = v[s];
}

__global__ void k_periodic_ew(float* e, float* u, float* v, int nx, int ny, int pitch) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int g = blockIdx.y;
    if (j > ny + 3) return;
    const int i = g < 2 ? g : nx + g;
    const int src = g < 2 ? i + nx : i - nx;
    const size_t d = (size_t)j * pitch + i, s = (size_t)j * pitch + src;
    e[d] = e[s]; u[d] = u[s]; v[d] = v[s];
}

int periodic_bc(const gocean_domain_t* d, float* e, float* u, float* v, hipStream_t s) {
    if (d->bc_north == kPeriodic) {
        hipLaunchKernelGGL(k_periodic_ns, dim3((d->nx + 4 + 255) / 256, 4), dim3(256), 0, s,
                           e, u, v, d->nx, d->ny, d->pitch);
    }
    if (d->bc_east == kPeriodic) {
        hipLaunchKernelGGL(k_periodic_ew, dim3((d->ny + 4 + 255) / 256, 4), dim3(256), 0, s,
                           e, u, v, d->nx, d->ny, d->pitch);
    }
    return check_launch("periodic boundary launch");
}

// ---------------------------------------------------------------------------
// CFL time step (max_dt.cu:42-194).  The minimum is exact, so any reduction
// order gives the reference's value; partial minima go through a buffer the
// caller provides (or the static array of gocean_max_dt_static), so the two
// launches are graph-capturable and need no memset: every launch writes each
// partial it reads.
// ---------------------------------------------------------------------------
constexpr int kDtBlocks = kMaxDtPartials;
constexpr int kDtThreads = 256;

__device__ __forceinline__ float wave_min(float v) {
    for (int o = 32; o > 0; o >>= 1) v = fminf(v, __shfl_xor(v, o, 64));
    return v;
}

__device__ __forceinline__ float block_min(float v, float* red) {
    v = wave_min(v);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) red[w] = v;
    __syncthreads();
    if (w == 0) {
        v = (lane < (int)(blockDim.x >> 6)) ? red[lane] : 1.0e5f;
        v = wave_min(v);
    }
    return v;
}

__global__ __launch_bounds__(kDtThreads) void k_dt_partial(const float* eta, const float* hu,
                                                           const float* hv, const float* Hm,
                                                           int nx, int ny, int pitch, float dx, float dy,
                                                           float g, float land, float* partial) {
    __shared__ float red[kDtThreads / 64];
    float m = 1.0e5f;   // FLT_MAX of max_dt.cu:29
    const float qdx = 0.25f * dx, qdy = 0.25f * dy;
    // one interior row per workgroup at a time, columns across its threads
    for (int j = blockIdx.x + 2; j < ny + 2; j += gridDim.x)
    for (int i = threadIdx.x + 2; i < nx + 2; i += blockDim.x) {
        const size_t o = (size_t)j * pitch + i;
        const float H = Hm[o];
        const float h = eta[o] + H;
        if (h <= 0.0f || fabsf(H - land) < kDryEps) continue;
        const float u = hu[o] / h, v = hv[o] / h;
        const float c = sqrtf(g * h);
        float d = qdx / fabsf(u + c);
        d = fminf(d, qdx / fabsf(u - c));
        d = fminf(d, qdy / fabsf(v + c));
        d = fminf(d, qdy / fabsf(v - c));
        m = fminf(m, d);
    }
    m = block_min(m, red);
    if (threadIdx.x == 0) partial[blockIdx.x] = m;
}

__global__ __launch_bounds__(kDtThreads) void k_dt_final(const float* partial, int n, float courant,
                                                         float* out) {
    __shared__ float red[kDtThreads / 64];
    float m = 1.0e5f;
    for (int k = threadIdx.x; k < n; k += blockDim.x) m = fminf(m, partial[k]);
    m = block_min(m, red);
    if (threadIdx.x == 0) {
        out[0] = m;
        out[1] = courant > 0.0f ? courant * m : m;
    }
}

int max_dt_partials(const gocean_domain_t* d) { return d->ny < kDtBlocks ? d->ny : kDtBlocks; }

int max_dt_into(const gocean_domain_t* d, const float* eta, const float* hu, const float* hv,
                const float* Hm, float courant, float* partial, float* out, hipStream_t s) {
    const int nb = max_dt_partials(d);
    hipLaunchKernelGGL(k_dt_partial, dim3(nb), dim3(kDtThreads), 0, s, eta, hu, hv, Hm, d->nx,
                       d->ny, d->pitch, d->dx, d->dy, d->g, d->land_value, partial);
    hipLaunchKernelGGL(k_dt_final, dim3(1), dim3(kDtThreads), 0, s, partial, nb, courant, out);
    return check_launch("max dt launch");
}

int max_dt(const gocean_domain_t* d, const float* eta, const float* hu, const float* hv,
           const float* Hm, float courant, float* out, hipStream_t s) {
    return max_dt_into(d, eta, hu, hv, Hm, courant, out + 2, out, s);
}

}  // namespace gocean

// ===========================================================================
// C ABI
// ===========================================================================
using namespace gocean;

extern "C" {

const char* gocean_version(void) { return "gocean-mi355x 0.1 (gfx950)"; }

const char* gocean_last_error(void) { return last_error(); }

int gocean_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

int gocean_adjoint_set_record_tile_config(int cfg) {
    clear_error();
    if (cfg < 0) cfg = -1;
    if (cfg >= kNumTileConfigs)
        return set_error(GOCEAN_EINVAL, "tile configuration %d out of range [0,%d)", cfg, kNumTileConfigs);
    g_record_cfg = cfg;
    return 0;
}

int gocean_adjoint_record_tile_config(void) { return g_record_cfg; }

int gocean_default_tile_config(int nx, int ny) { return default_tile_config((long long)nx * ny); }

int gocean_sqrt_newton_check(uint32_t lo_bits, uint32_t hi_bits, unsigned long long* mismatches) {
    clear_error();
    if (!mismatches) return set_error(GOCEAN_EINVAL, "mismatches is null");
    if (lo_bits > hi_bits || hi_bits > 0x7F7FFFFFu)
        return set_error(GOCEAN_EINVAL, "range [%08x, %08x] is not within the positive finite floats", lo_bits,
                         hi_bits);
    unsigned long long* d = nullptr;
    if (hipMalloc(&d, sizeof(unsigned long long)) != hipSuccess) {
        (void)hipGetLastError();
        return set_error(GOCEAN_ENOMEM, "hipMalloc of the mismatch counter failed");
    }
    hipError_t e = hipMemset(d, 0, sizeof(unsigned long long));
    int rc = (e == hipSuccess) ? GOCEAN_OK : set_error((int)e, "hipMemset: %s", hipGetErrorString(e));
    if (rc == GOCEAN_OK) {
        hipLaunchKernelGGL(k_sqrt_check, dim3(4096), dim3(256), 0, 0, lo_bits, hi_bits, d);
        rc = check_launch("sqrt check launch");
    }
    if (rc == GOCEAN_OK) {
        e = hipMemcpy(mismatches, d, sizeof(unsigned long long), hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = set_error((int)e, "hipMemcpy: %s", hipGetErrorString(e));
    }
    (void)hipFree(d);
    return rc;
}

int gocean_tile_configs(const char** names, int max) {
    for (int k = 0; k < kNumTileConfigs && k < max; ++k) names[k] = kTileConfigs[k].name;
    return kNumTileConfigs;
}

int gocean_bathymetry_init(const gocean_domain_t* d, float* Hi, float* Hm, void* stream) {
    clear_error();
    if (int e = validate_domain(d)) return e;
    if (!Hi || !Hm) return set_error(GOCEAN_EINVAL, "Hi/Hm is NULL");
    hipStream_t s = (hipStream_t)stream;
    const int nsp = d->bc_north == kPeriodic && d->bc_south == kPeriodic;
    const int ewp = d->bc_east == kPeriodic && d->bc_west == kPeriodic;
    hipLaunchKernelGGL(k_hi_halo_ns, dim3((d->nx + 5 + 255) / 256), dim3(256), 0, s, Hi, d->nx,
                       d->ny, d->hi_pitch, nsp);
    hipLaunchKernelGGL(k_hi_halo_ew, dim3((d->ny + 5 + 255) / 256), dim3(256), 0, s, Hi, d->nx,
                       d->ny, d->hi_pitch, ewp);
    hipLaunchKernelGGL(k_init_bm, dim3((d->nx + 4 + 15) / 16, (d->ny + 4 + 15) / 16), dim3(16, 16),
                       0, s, Hi, Hm, d->nx, d->ny, d->hi_pitch, d->pitch, d->land_value);
    return check_launch("bathymetry init launch");
}

int gocean_periodic_bc(const gocean_domain_t* d, float* eta, float* hu, float* hv, void* stream) {
    clear_error();
    if (int e = validate_domain(d)) return e;
    if (!eta || !hu || !hv) return set_error(GOCEAN_EINVAL, "state pointer is NULL");
    return periodic_bc(d, eta, hu, hv, (hipStream_t)stream);
}

int gocean_max_dt(const gocean_domain_t* d, const float* eta, const float* hu, const float* hv,
                  const float* Hm, float courant, float* out_dev, void* stream) {
    clear_error();
    if (int e = validate_domain(d)) return e;
    if (!eta || !hu || !hv || !Hm || !out_dev) return set_error(GOCEAN_EINVAL, "NULL pointer");
    return max_dt(d, eta, hu, hv, Hm, courant, out_dev, (hipStream_t)stream);
}

int gocean_cdklm_substep(const gocean_domain_t* d, float dt, const float* dt_dev, int rk_step,
                         const float* eta_in, const float* hu_in, const float* hv_in,
                         const float* eta_qn, const float* hu_qn, const float* hv_qn,
                         float* eta_out, float* hu_out, float* hv_out, const float* Hi,
                         const gocean_fields_t* fields, const gocean_wind_t* wind,
                         int tile_config, void* stream) {
    clear_error();
    if (int e = validate_domain(d)) return e;
    if (!eta_in || !hu_in || !hv_in || !eta_out || !hu_out || !hv_out || !Hi)
        return set_error(GOCEAN_EINVAL, "state/bathymetry pointer is NULL");
    if (rk_step < 0 || rk_step >= d->rk_order)
        return set_error(GOCEAN_EINVAL, "rk_step %d invalid for rk_order %d", rk_step, d->rk_order);
    if (eta_in == eta_out || hu_in == hu_out || hv_in == hv_out)
        return set_error(GOCEAN_EINVAL, "input and output state must not alias");
    SubstepArgs A;
    fill_args(d, A);
    A.eta_in = eta_in; A.hu_in = hu_in; A.hv_in = hv_in;
    A.eta_out = eta_out; A.hu_out = hu_out; A.hv_out = hv_out;
    A.eta_qn = eta_qn ? eta_qn : eta_out;
    A.hu_qn = hu_qn ? hu_qn : hu_out;
    A.hv_qn = hv_qn ? hv_qn : hv_out;
    A.Hi = Hi;
    A.dt = dt;
    A.dt_dev = dt_dev;
    A.set_rk_step(rk_step);
    bool field = false;
    if (fields && (fields->f || fields->north_x || fields->north_y)) {
        if (!fields->f || !fields->north_x || !fields->north_y)
            return set_error(GOCEAN_EINVAL, "per-cell fields must give f, north_x and north_y together");
        A.f_field = fields->f; A.nx_field = fields->north_x; A.ny_field = fields->north_y;
        field = true;
    }
    if (wind) {
        A.wind_x = wind_lerp(wind->t, wind->x0, wind->x1);
        A.wind_y = wind_lerp(wind->t, wind->y0, wind->y1);
        if (wind->grid) {
            if (wind->grid_nx < 1 || wind->grid_ny < 1)
                return set_error(GOCEAN_EINVAL, "wind grid %d x %d: both sides must be >= 1",
                                 wind->grid_nx, wind->grid_ny);
            A.wind_grid = wind->grid;
            A.wind_wx = wind->grid_nx;
            A.wind_wy = wind->grid_ny;
            A.wind_t = wind->t;
        }
    }
    return launch_substep(A, tile_config, field, (hipStream_t)stream);
}

int gocean_rk_steps(const gocean_domain_t* d, int n_steps, float dt, const float* dt_dev,
                    float* eta0, float* hu0, float* hv0, float* eta1, float* hu1, float* hv1,
                    float* eta2, float* hu2, float* hv2, const float* Hi,
                    const gocean_fields_t* fields, const gocean_wind_t* wind, int tile_config,
                    void* stream) {
    clear_error();
    if (int e = validate_domain(d)) return e;
    if (n_steps < 0) return set_error(GOCEAN_EINVAL, "n_steps < 0");
    if (d->rk_order == 3 && (!eta2 || !hu2 || !hv2))
        return set_error(GOCEAN_EINVAL, "rk_order 3 needs the third buffer set");
    if (d->rk_order == 1 && n_steps % 2 != 0)
        return set_error(GOCEAN_EINVAL,
                         "rk_order 1 swaps buffers every step; call with an even n_steps");
    for (int n = 0; n < n_steps; ++n) {
        int e = 0;
        if (d->rk_order == 2) {
            e = gocean_cdklm_substep(d, dt, dt_dev, 0, eta0, hu0, hv0, nullptr, nullptr, nullptr,
                                     eta1, hu1, hv1, Hi, fields, wind, tile_config, stream);
            if (!e)
                e = gocean_cdklm_substep(d, dt, dt_dev, 1, eta1, hu1, hv1, nullptr, nullptr, nullptr,
                                         eta0, hu0, hv0, Hi, fields, wind, tile_config, stream);
        } else if (d->rk_order == 1) {
            // Euler: h0 -> h1, then the caller's buffers swap; two steps
            // bring the state back into h0.
            if (n % 2 == 0)
                e = gocean_cdklm_substep(d, dt, dt_dev, 0, eta0, hu0, hv0, nullptr, nullptr, nullptr,
                                         eta1, hu1, hv1, Hi, fields, wind, tile_config, stream);
            else
                e = gocean_cdklm_substep(d, dt, dt_dev, 0, eta1, hu1, hv1, nullptr, nullptr, nullptr,
                                         eta0, hu0, hv0, Hi, fields, wind, tile_config, stream);
        } else {
            // SSP-RK3; the middle stage goes to the third buffer instead of
            // overwriting its own input as the reference does
            e = gocean_cdklm_substep(d, dt, dt_dev, 0, eta0, hu0, hv0, nullptr, nullptr, nullptr,
                                     eta1, hu1, hv1, Hi, fields, wind, tile_config, stream);
            if (!e)
                e = gocean_cdklm_substep(d, dt, dt_dev, 1, eta1, hu1, hv1, eta0, hu0, hv0, eta2, hu2,
                                         hv2, Hi, fields, wind, tile_config, stream);
            if (!e)
                e = gocean_cdklm_substep(d, dt, dt_dev, 2, eta2, hu2, hv2, nullptr, nullptr, nullptr,
                                         eta0, hu0, hv0, Hi, fields, wind, tile_config, stream);
        }
        if (e) return e;
    }
    return 0;
}

}  // extern "C"
