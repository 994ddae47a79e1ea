// gocean_noise.hip -- model error for the CDKLM16 state (OceanStateNoise).
//
// The reference perturbs the ocean state with a SOAR-correlated random eta
// field and the geostrophically balanced (hu, hv) that goes with it
// (gpu_ocean/SWESimulators/OceanStateNoise.py:399-496,
//  gpu_kernels/ocean_noise.cu).  Four kernels cover it:
//
//   k_noise_rng      random field: LCG + Box-Muller (ocean_noise.cu:80-195)
//                    or Philox4x32-10 (in place of pycuda's XORWOW stream)
//   k_noise_soar     coarse eta perturbation = SOAR covariance (cutoff 2)
//                    applied to the random field (ocean_noise.cu:248-370)
//   k_noise_balance  add d_eta and the balanced (d_hu, d_hv) to the state,
//                    directly (factor 1, ocean_noise.cu:379-505) or after
//                    bicubic interpolation from the coarse grid (odd factor
//                    > 1, ocean_noise.cu:522-772)
//   k_noise_fused1   SOAR and balance in one launch for factor 1
//   k_noise_norms / k_noise_perp   sum xi^2, nu^2, xi.nu (reductions.cu) and
//                    the perpendicular pair (ocean_noise.cu:206-243)
//
// Every kernel runs over a table of members, so one launch perturbs every
// member of an ensemble (the single-domain entry points use a table of one
// held in the kernel arguments).  All are memory-bound and light: per fine
// cell the balance reads the coarse field, Hi and the state and writes the
// state back; the SOAR and generator kernels touch the (coarse) random field.

#include <algorithm>
#include <cmath>
#include <vector>

#include "gocean_internal.hpp"

namespace gocean {
namespace {

constexpr int kCut = 2;                    // config.soar_cutoff
constexpr int kRngNT = 256;
constexpr int kSoarBW = 32, kSoarBH = 8;   // SOAR tile over the coarse buffer
constexpr int kBalBW = 64, kBalBH = 4;     // balance tile over fine interior cells
constexpr int kFuBW = 64, kFuBH = 16;      // fused SOAR + balance tile (factor 1)
constexpr int kFuNT = 256;                 // its threads: four rows of the tile each
constexpr int kNormNT = 1024;

struct NoiseMember {
    int nx, ny, pitch, hi_pitch;
    long long eta_off, hu_off, hv_off, hi_off;
    long long f_off, nxf_off, nyf_off;          // -1: constant f / north
    float dx, dy, g, f, beta, north_x, north_y, land;
    int factor, cnx, cny, periodic_ns, periodic_ew;
    float cdx, cdy, q0;
    int rand_nx, rand_ny, seed_nx, seed_ny;
    long long seed_off, rand_off, coarse_off;
    float w[(2 * kCut + 1) * (2 * kCut + 1)];   // (1 + d/L) exp(-d/L), row-major (dy, dx)
    int rng_begin, soar_begin, soar_tiles_x, bal_begin, bal_tiles_x, fu_begin, fu_tiles_x;
    int id;
    // set on the host from the fields above (fill_member): g / f when the
    // Coriolis parameter is one constant (beta 0, no f field), else 0; 2 dx,
    // 2 dy and their reciprocals rounded to float (RN(1/d)), for correctly
    // rounded divisions without the generic sequence
    float gc_const, two_dx, two_dy, r_two_dx, r_two_dy;
};

struct NoiseLaunch {
    const NoiseMember* members;   // device table, or NULL for `one`
    const int* map;               // block -> member (NULL with `one`)
    NoiseMember one;
    float* eta;
    float* hu;
    float* hv;
    const float* Hi;
    const float* ff;
    const float* fnx;
    const float* fny;
    unsigned long long* seed;
    float* random;
    float* coarse;
    // Philox counter advanced by the first kernel after the generator (the
    // batched perturbation folds the advance into its SOAR launch), or NULL
    unsigned long long* bump;
};

// The member a block works on.  Single domain: the table of one in the
// kernel arguments; batch: the device table through the block -> member map.
// Kept as two instantiations so that every field read is a scalar load from
// a known address space.
template <bool BATCH>
__device__ __forceinline__ const NoiseMember& member_of(const NoiseLaunch& L) {
    if constexpr (BATCH) return L.members[L.map[blockIdx.x]];
    else return L.one;
}

// a mod n for a within a few periods of [0, n): conditional wraps instead of
// an integer division
__device__ __forceinline__ int pmod(int a, int n) {
    while (a < 0) a += n;
    while (a >= n) a -= n;
    return a;
}

// ---------------------------------------------------------------------------
// generators
// ---------------------------------------------------------------------------
constexpr unsigned long long kLcgA = 1103515245ull, kLcgC = 12345ull, kLcgM = 2147483647ull;

// (s a + c) mod (2^31 - 1) with uint64 wrap-around, as ocean_noise.cu:85.
// Seeds below 2^31 (all the generator produces) keep s a + c < 2^62, which
// folds by 2^31 = 1 (mod 2^31 - 1) without a 64-bit division.
__device__ __forceinline__ unsigned long long lcg_next(unsigned long long s) {
    const unsigned long long x = s * kLcgA + kLcgC;
    if (s >> 31) return x % kLcgM;
    unsigned long long y = (x & kLcgM) + (x >> 31);
    y = (y & kLcgM) + (y >> 31);
    return y >= kLcgM ? y - kLcgM : y;
}

// float(seed / 2^31): the exact quotient rounded once to float
__device__ __forceinline__ float lcg_unit(unsigned long long s) {
    return (float)s * 0x1p-31f;
}

__device__ __forceinline__ void box_muller(float u1, float u2, float& n1, float& n2) {
    const float r = sqrtf(-2.0f * logf(u1));
    n1 = r * cospif(2.0f * u2);
    n2 = r * sinpif(2.0f * u2);
}

// Philox4x32-10 (Salmon, Moraes, Dror & Shaw, SC'11)
__device__ __forceinline__ uint4 philox4x32_10(uint4 c, unsigned k0, unsigned k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        const unsigned hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
        const unsigned hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
        c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
    }
    return c;
}

// (k + 1/2) 2^-23 for the top 23 bits k: exact in float, strictly in (0, 1)
__device__ __forceinline__ float philox_unit(unsigned x) {
    return (float(x >> 9) + 0.5f) * 0x1p-23f;
}

// kind: GOCEAN_RNG_*.  LCG: one thread per seed, two numbers each, the
// column rule of ocean_noise.cu:186-192.  Philox: one thread per four
// consecutive numbers of the member's row-major field.
template <bool BATCH>
__global__ __launch_bounds__(kRngNT) void k_noise_rng(NoiseLaunch L, int kind,
                                                      const unsigned long long* counter,
                                                      unsigned long long key) {
    const NoiseMember& M = member_of<BATCH>(L);
    const int t = (blockIdx.x - M.rng_begin) * kRngNT + threadIdx.x;
    float* rnd = L.random + M.rand_off;
    if (kind <= GOCEAN_RNG_LCG_UNIFORM) {
        if (t >= M.seed_nx * M.seed_ny) return;
        const int y = t / M.seed_nx, x = t - y * M.seed_nx;
        unsigned long long* sp = L.seed + M.seed_off + t;
        unsigned long long s = *sp;
        s = lcg_next(s);
        const float u1 = lcg_unit(s);
        s = lcg_next(s);
        const float u2 = lcg_unit(s);
        *sp = s;
        float n1 = u1, n2 = u2;
        if (kind == GOCEAN_RNG_LCG_NORMAL) box_muller(u1, u2, n1, n2);
        float* row = rnd + (long long)y * M.rand_nx;
        if (2 * x + 1 < M.rand_nx) {
            row[2 * x] = n1;
            row[2 * x + 1] = n2;
        } else if (2 * x == M.rand_nx) {
            row[2 * x] = n1;
        }
        return;
    }
    const long long n = (long long)M.rand_nx * M.rand_ny;
    const long long e = 4ll * t;
    if (e >= n) return;
    const unsigned long long c = *counter;
    const uint4 r = philox4x32_10(make_uint4((unsigned)t, (unsigned)c, (unsigned)(c >> 32), (unsigned)M.id),
                                  (unsigned)key, (unsigned)(key >> 32));
    float v[4] = {philox_unit(r.x), philox_unit(r.y), philox_unit(r.z), philox_unit(r.w)};
    if (kind == GOCEAN_RNG_PHILOX_NORMAL) {
        box_muller(v[0], v[1], v[0], v[1]);
        box_muller(v[2], v[3], v[2], v[3]);
    }
    for (int k = 0; k < 4; ++k)
        if (e + k < n) rnd[e + k] = v[k];
}

__global__ void k_counter_bump(unsigned long long* counter) { *counter += 1ull; }

// the counter advance of a launch that runs after the generator read it; at
// the end of the kernel, so that the atomic orders none of its loads
__device__ __forceinline__ void bump_counter(const NoiseLaunch& L) {
    if (L.bump && blockIdx.x == 0 && threadIdx.x == 0) atomicAdd(L.bump, 1ull);
}

// ---------------------------------------------------------------------------
// SOAR on the coarse grid: coarse(a) = q0 scale sum_{|o| <= 2} w(o) xi(a+o).
// Coarse index a reads random index a+2 (+o) when the random field carries
// 2+cutoff ghost cells, and (a-2+o) mod n when periodic (ocean_noise.cu:
// 305-327); the sum runs dy-outer, dx-inner as the reference's loop.
// ---------------------------------------------------------------------------
template <bool BATCH>
__global__ __launch_bounds__(kSoarBW * kSoarBH) void k_noise_soar(NoiseLaunch L, float q0_scale,
                                                                  float scale, int additive) {
    constexpr int W = kSoarBW + 2 * kCut, H = kSoarBH + 2 * kCut, NW = (2 * kCut + 1) * (2 * kCut + 1);
    __shared__ float xi[H][W + 1];
    const NoiseMember& M = member_of<BATCH>(L);
    float w[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) w[k] = M.w[k];
    const int lt = blockIdx.x - M.soar_begin;
    const int by = (lt / M.soar_tiles_x) * kSoarBH;
    const int bx = (lt % M.soar_tiles_x) * kSoarBW;
    const float* rnd = L.random + M.rand_off;
    for (int k = threadIdx.x; k < H * W; k += kSoarBW * kSoarBH) {
        const int j = k / W, i = k - j * W;
        const int gy = M.periodic_ns ? pmod(by + j - 2 * kCut, M.cny) : min(by + j, M.rand_ny - 1);
        const int gx = M.periodic_ew ? pmod(bx + i - 2 * kCut, M.cnx) : min(bx + i, M.rand_nx - 1);
        xi[j][i] = rnd[(long long)gy * M.rand_nx + gx];
    }
    __syncthreads();
    const int tx = threadIdx.x % kSoarBW, ty = threadIdx.x / kSoarBW;
    const int ax = bx + tx, ay = by + ty;
    if (ax >= M.cnx + 4 || ay >= M.cny + 4) return;
    float acc = 0.0f;
#pragma unroll
    for (int oy = 0; oy <= 2 * kCut; ++oy)
#pragma unroll
        for (int ox = 0; ox <= 2 * kCut; ++ox)
            acc = acc + w[oy * (2 * kCut + 1) + ox] * xi[ty + oy][tx + ox];
    const float v = ((M.q0 * q0_scale) * acc) * scale;
    float* c = L.coarse + M.coarse_off + (long long)ay * (M.cnx + 4) + ax;
    *c = additive ? *c + v : v;
    bump_counter(L);   // thread 0 of block 0 (coarse cell (0, 0)) gets here
}

// ---------------------------------------------------------------------------
// balance
// ---------------------------------------------------------------------------
// Coarse cell and position in it of fine ghosted index t, in the float
// arithmetic of ocean_noise.cu:614-637 (ghost cells 2 on both grids).
__device__ __forceinline__ void coarse_cell(int t, int off, float d, float cd, int cn, int& ci,
                                            float& rel) {
    const float x = ((float)(t - 2) + 0.5f + (float)off) * d;
    ci = (int)floorf((x / cd) + 2.0f - 0.5f);
    if (ci > cn + 1) ci -= 1;
    const float cx = ((float)(ci - 2) + 0.5f) * cd;
    rel = (x - cx) / cd;
}

// Cubic Hermite weights: (1, x, x^2, x^3) times the reference's bicubic
// matrix (gpu_kernels/common.cu:723-735): values at 0 and 1, slopes at 0 and 1.
__device__ __forceinline__ float4 hermite(float x) {
    const float x2 = x * x, x3 = x2 * x;
    return make_float4(1.0f - 3.0f * x2 + 2.0f * x3, 3.0f * x2 - 2.0f * x3, x - 2.0f * x2 + x3,
                       x3 - x2);
}

// The reference's bicubic surface (ocean_noise.cu:652-700: corner values,
// central-difference slopes and cross slopes, B F B^T) is the tensor product
// of the 1-D Hermite cubic with central-difference slopes, so along each
// axis the four coarse values c(-1), c(0), c(1), c(2) enter with the weights
// (-h10/2, h00 - h11/2, h01 + h10/2, h11/2) of the Hermite basis h.
__device__ __forceinline__ float4 cr_weights(float r) {
    const float4 h = hermite(r);
    return make_float4(-0.5f * h.z, h.x - 0.5f * h.w, h.y + 0.5f * h.z, 0.5f * h.w);
}

// Coarse window per block, large enough for factors >= 3 (the window spans
// floor((BW+1)/factor) + 5 columns); larger windows fall back to reads from
// the coarse buffer.
constexpr int kWinW = kBalBW / 3 + 8, kWinH = kBalBH / 3 + 8;
// threads [0, BW+2) set up the columns, [kRowT, kRowT+BH+2) the rows
constexpr int kRowT = ((kBalBW + 2 + 63) / 64) * 64;
static_assert(kRowT + kBalBH + 2 <= kBalBW * kBalBH, "column/row set-up threads");

// Balance of one interior cell given d_eta and its central differences:
// h_mid, Coriolis, the geostrophic (d_hu, d_hv) and the land rules of
// geostrophicBalance (interp false, ocean_noise.cu:443-504) or
// bicubicInterpolation (interp true, :720-770).
__device__ __forceinline__ void balance_cell(const NoiseLaunch& L, const NoiseMember& M, int ti, int tj,
                                             float d, float ex, float ey, bool interp) {
    const float* Hi = L.Hi + M.hi_off;
    const long long h0 = (long long)tj * M.hi_pitch + ti;
    const float H00 = Hi[h0], H01 = Hi[h0 + 1], H10 = Hi[h0 + M.hi_pitch], H11 = Hi[h0 + M.hi_pitch + 1];
    const bool dry = H00 == M.land || H01 == M.land || H10 == M.land || H11 == M.land;
    const float Hmid = 0.25f * (H00 + H01 + H10 + H11);

    const long long s = (long long)tj * M.pitch + ti;
    float nxv = M.north_x, nyv = M.north_y;
    if (M.nxf_off >= 0) {
        nxv = L.fnx[M.nxf_off + s];
        nyv = L.fny[M.nyf_off + s];
    }
    float cor = M.f + M.beta * (((float)ti + 0.5f) * M.dx * nxv + ((float)tj + 0.5f) * M.dy * nyv);
    if (M.f == 0.0f) cor = L.ff[M.f_off + s];
    const float gc = M.g / cor;

    float* eta = L.eta + M.eta_off + s;
    float* hu = L.hu + M.hu_off + s;
    float* hv = L.hv + M.hv_off + s;
    const float e0 = *eta;
    if (interp) {
        const float hm = Hmid + d;
        float dhu = -gc * hm * ey;
        float dhv = gc * hm * ex;
        if (dry) { d = 0.0f; dhu = 0.0f; dhv = 0.0f; }
        *eta = e0 + d;
        *hu = *hu + dhu;
        *hv = *hv + dhv;
    } else {
        const float hm = d + Hmid + e0;
        const float dhu = -gc * hm * ey;
        const float dhv = gc * hm * ex;
        if (dry) {
            *eta = 0.0f;
            *hu = 0.0f;
            *hv = 0.0f;
        } else {
            *eta = e0 + d;
            *hu = *hu + dhu;
            *hv = *hv + dhv;
        }
    }
}


// One thread per fine interior cell (ghosted ti = 2..nx+1, tj = 2..ny+1).
// Factor 1: d_eta is the coarse buffer itself, h_mid includes the current
// eta and a cell with a land corner gets eta = hu = hv = 0
// (ocean_noise.cu:443-504).  Factor > 1: d_eta on the tile and one ring is
// interpolated into LDS first, from a coarse window staged in LDS; h_mid =
// H_mid + d_eta and a land cell gets no perturbation (:607-770).
template <bool BATCH>
__global__ __launch_bounds__(kBalBW * kBalBH) void k_noise_balance(NoiseLaunch L, int offset_i,
                                                                   int offset_j) {
    __shared__ float de[kBalBH + 2][kBalBW + 3];
    __shared__ float win[kWinH][kWinW + 1];
    __shared__ float4 wxs[kBalBW + 2], wys[kBalBH + 2];
    __shared__ int cxs[kBalBW + 2], cys[kBalBH + 2];
    const NoiseMember& M = member_of<BATCH>(L);
    const int lt = blockIdx.x - M.bal_begin;
    const int by = 2 + (lt / M.bal_tiles_x) * kBalBH;
    const int bx = 2 + (lt % M.bal_tiles_x) * kBalBW;
    const float* C = L.coarse + M.coarse_off;
    const int cp = M.cnx + 4;
    const bool interp = M.factor > 1;
    if (interp) {   // uniform per block
        // coarse cells of the first and last fine column / row of the tile's ring
        int c0, c1, r0, r1;
        float tmp;
        coarse_cell(bx - 1, offset_i, M.dx, M.cdx, M.cnx, c0, tmp);
        coarse_cell(min(bx + kBalBW, M.nx + 2), offset_i, M.dx, M.cdx, M.cnx, c1, tmp);
        coarse_cell(by - 1, offset_j, M.dy, M.cdy, M.cny, r0, tmp);
        coarse_cell(min(by + kBalBH, M.ny + 2), offset_j, M.dy, M.cdy, M.cny, r1, tmp);
        const int wx0 = c0 - 1, wy0 = r0 - 1;
        const int ww = c1 - c0 + 4, wh = r1 - r0 + 4;
        const bool staged = ww <= kWinW && wh <= kWinH;
        const int t = threadIdx.x;
        if (t < kBalBW + 2) {
            int ci;
            float rx;
            coarse_cell(bx - 1 + t, offset_i, M.dx, M.cdx, M.cnx, ci, rx);
            cxs[t] = ci;
            wxs[t] = cr_weights(rx);
        } else if (t >= kRowT && t < kRowT + kBalBH + 2) {
            int cj;
            float ry;
            coarse_cell(by - 1 + (t - kRowT), offset_j, M.dy, M.cdy, M.cny, cj, ry);
            cys[t - kRowT] = cj;
            wys[t - kRowT] = cr_weights(ry);
        }
        if (staged)
            for (int k = t; k < wh * ww; k += kBalBW * kBalBH) {
                const int j = k / ww, i = k - j * ww;
                const int gy = min(max(wy0 + j, 0), M.cny + 3);
                const int gx = min(max(wx0 + i, 0), M.cnx + 3);
                win[j][i] = C[(long long)gy * cp + gx];
            }
        __syncthreads();
        for (int k = t; k < (kBalBH + 2) * (kBalBW + 2); k += kBalBW * kBalBH) {
            const int j = k / (kBalBW + 2), i = k - j * (kBalBW + 2);
            if (bx + i - 1 > M.nx + 2 || by + j - 1 > M.ny + 2) continue;   // beyond this domain's ring
            const float4 wx = wxs[i], wy = wys[j];
            const int ci = cxs[i], cj = cys[j];
            float v = 0.0f;
            if (staged) {
                const int lx = ci - 1 - wx0, ly = cj - 1 - wy0;
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const float* row = &win[ly + b][lx];
                    const float rv = wx.x * row[0] + wx.y * row[1] + wx.z * row[2] + wx.w * row[3];
                    v = v + (b == 0 ? wy.x : b == 1 ? wy.y : b == 2 ? wy.z : wy.w) * rv;
                }
            } else {
#pragma unroll
                for (int b = 0; b < 4; ++b) {
                    const long long ro = (long long)min(max(cj - 1 + b, 0), M.cny + 3) * cp;
                    float rv = 0.0f;
#pragma unroll
                    for (int a = 0; a < 4; ++a) {
                        const float wa = a == 0 ? wx.x : a == 1 ? wx.y : a == 2 ? wx.z : wx.w;
                        rv = rv + wa * C[ro + min(max(ci - 1 + a, 0), M.cnx + 3)];
                    }
                    v = v + (b == 0 ? wy.x : b == 1 ? wy.y : b == 2 ? wy.z : wy.w) * rv;
                }
            }
            de[j][i] = v;
        }
        __syncthreads();
    }
    const int tx = threadIdx.x % kBalBW, ty = threadIdx.x / kBalBW;
    const int ti = bx + tx, tj = by + ty;
    if (ti >= M.nx + 2 || tj >= M.ny + 2) return;

    float d, ex, ey;
    if (interp) {
        d = de[ty + 1][tx + 1];
        ex = (de[ty + 1][tx + 2] - de[ty + 1][tx]) / (2.0f * M.dx);
        ey = (de[ty + 2][tx + 1] - de[ty][tx + 1]) / (2.0f * M.dy);
    } else {
        const long long c0 = (long long)tj * cp + ti;
        d = C[c0];
        ex = (C[c0 + 1] - C[c0 - 1]) / (2.0f * M.dx);
        ey = (C[c0 + cp] - C[c0 - cp]) / (2.0f * M.dy);
    }
    balance_cell(L, M, ti, tj, d, ex, ey, interp);
}

// SOAR and geostrophic balance in one launch for interpolation factor 1
// (the SOAR kernel followed by geostrophicBalance, ocean_noise.cu:264-505):
// the block computes d_eta for its 64x16 tile and a 2-cell ring in LDS from
// the random field (2-cell SOAR halo on top), writes the coarse-buffer cells
// it owns (the tile's interior, plus the ghost rings next to it on the
// domain's edges: each coarse cell once), and balances its interior cells.
// Each thread balances four cells of a column of the tile; their depths and
// state are loaded before the first barrier, so that those loads overlap the
// random-field and SOAR stages.  The divisions by 2 dx, 2 dy and (constant
// Coriolis) f are correctly rounded with reciprocals formed on the host, as
// the generic division rounds them.
template <bool BATCH>
__global__ __launch_bounds__(kFuNT) void k_noise_fused1(NoiseLaunch L, float q0_scale, float scale) {
    constexpr int DW = kFuBW + 4, DH = kFuBH + 4;             // d_eta: tile + 2-cell ring
    constexpr int RW = DW + 2 * kCut, RH = DH + 2 * kCut;       // random values: + SOAR halo
    constexpr int NW = (2 * kCut + 1) * (2 * kCut + 1);
    constexpr int RPT = kFuBH * kFuBW / kFuNT;                  // cells per thread
    __shared__ float xi[RH][RW + 1];
    __shared__ float de[DH][DW + 1];
    const NoiseMember& M = member_of<BATCH>(L);
    const int lt = blockIdx.x - M.fu_begin;
    const int by = 2 + (lt / M.fu_tiles_x) * kFuBH;
    const int bx = 2 + (lt % M.fu_tiles_x) * kFuBW;
    const int ay0 = by - 2, ax0 = bx - 2;     // coarse (= fine, ghosted) index of de[0][0]
    const int tx = threadIdx.x % kFuBW, ty0 = threadIdx.x / kFuBW;

    // the balance inputs of this thread's cells (rows ty0, ty0 + 4, ...)
    const float* Hi = L.Hi + M.hi_off;
    const int ti = bx + tx;
    float H00[RPT], H01[RPT], H10[RPT], H11[RPT], e0[RPT], u0[RPT], v0[RPT];
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int tj = by + ty0 + r * (kFuNT / kFuBW);
        H00[r] = H01[r] = H10[r] = H11[r] = e0[r] = u0[r] = v0[r] = 0.0f;
        if (ti < M.nx + 2 && tj < M.ny + 2) {
            const long long h0 = (long long)tj * M.hi_pitch + ti;
            H00[r] = Hi[h0]; H01[r] = Hi[h0 + 1]; H10[r] = Hi[h0 + M.hi_pitch]; H11[r] = Hi[h0 + M.hi_pitch + 1];
            const long long st = (long long)tj * M.pitch + ti;
            e0[r] = L.eta[M.eta_off + st]; u0[r] = L.hu[M.hu_off + st]; v0[r] = L.hv[M.hv_off + st];
        }
    }

    const float* rnd = L.random + M.rand_off;
    for (int k = threadIdx.x; k < RH * RW; k += kFuNT) {
        const int j = k / RW, i = k - j * RW;
        const int gy = M.periodic_ns ? pmod(ay0 + j - 2 * kCut, M.cny) : min(ay0 + j, M.rand_ny - 1);
        const int gx = M.periodic_ew ? pmod(ax0 + i - 2 * kCut, M.cnx) : min(ax0 + i, M.rand_nx - 1);
        xi[j][i] = rnd[(long long)gy * M.rand_nx + gx];
    }
    __syncthreads();
    float w[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) w[k] = M.w[k];
    const float q = M.q0 * q0_scale;
    const int cp = M.cnx + 4;
    float* C = L.coarse + M.coarse_off;
    for (int k = threadIdx.x; k < DH * DW; k += kFuNT) {
        const int j = k / DW, i = k - j * DW;
        const int ay = ay0 + j, ax = ax0 + i;
        if (ay > M.cny + 3 || ax > M.cnx + 3) continue;
        float acc = 0.0f;
#pragma unroll
        for (int oy = 0; oy <= 2 * kCut; ++oy)
#pragma unroll
            for (int ox = 0; ox <= 2 * kCut; ++ox)
                acc = acc + w[oy * (2 * kCut + 1) + ox] * xi[j + oy][i + ox];
        const float v = (q * acc) * scale;
        de[j][i] = v;
        // owner: the tile whose interior holds the cell clamped to the interior
        const int oy = min(max(ay, 2), M.ny + 1), ox = min(max(ax, 2), M.nx + 1);
        if (oy >= by && oy < by + kFuBH && ox >= bx && ox < bx + kFuBW) C[(long long)ay * cp + ax] = v;
    }
    __syncthreads();
    if (ti >= M.nx + 2) return;
#pragma unroll
    for (int r = 0; r < RPT; ++r) {
        const int ty = ty0 + r * (kFuNT / kFuBW), tj = by + ty;
        if (tj >= M.ny + 2) break;
        const float d = de[ty + 2][tx + 2];
        const float ex = div_r(de[ty + 2][tx + 3] - de[ty + 2][tx + 1], M.two_dx, M.r_two_dx);
        const float ey = div_r(de[ty + 3][tx + 2] - de[ty + 1][tx + 2], M.two_dy, M.r_two_dy);
        // balance (geostrophicBalance, ocean_noise.cu:443-504)
        const bool dry = H00[r] == M.land || H01[r] == M.land || H10[r] == M.land || H11[r] == M.land;
        const float Hmid = 0.25f * (H00[r] + H01[r] + H10[r] + H11[r]);
        float gc = M.gc_const;
        if (gc == 0.0f) {
            const long long st = (long long)tj * M.pitch + ti;
            float nxv = M.north_x, nyv = M.north_y;
            if (M.nxf_off >= 0) {
                nxv = L.fnx[M.nxf_off + st];
                nyv = L.fny[M.nyf_off + st];
            }
            float cor = M.f + M.beta * (((float)ti + 0.5f) * M.dx * nxv + ((float)tj + 0.5f) * M.dy * nyv);
            if (M.f == 0.0f) cor = L.ff[M.f_off + st];
            gc = M.g / cor;
        }
        const float hm = d + Hmid + e0[r];
        const float dhu = -gc * hm * ey;
        const float dhv = gc * hm * ex;
        const long long st = (long long)tj * M.pitch + ti;
        if (dry) {
            L.eta[M.eta_off + st] = 0.0f;
            L.hu[M.hu_off + st] = 0.0f;
            L.hv[M.hv_off + st] = 0.0f;
        } else {
            L.eta[M.eta_off + st] = e0[r] + d;
            L.hu[M.hu_off + st] = u0[r] + dhu;
            L.hv[M.hv_off + st] = v0[r] + dhv;
        }
    }
    bump_counter(L);   // thread 0 of block 0 (column 2) gets here
}

// ---------------------------------------------------------------------------
// norms and the perpendicular pair
// ---------------------------------------------------------------------------
// out[0..2] = sum xi^2, sum nu^2, sum xi nu (nu may be NULL: only the
// first), one block.  Compensated per-thread sums, fixed-order tree.
__global__ __launch_bounds__(kNormNT) void k_noise_norms(NoiseLaunch L, const float* xi_base,
                                                         const float* nu_base, float* out) {
    __shared__ float red[3][kNormNT];
    const NoiseMember& M = L.one;
    const long long n = (long long)M.rand_nx * M.rand_ny;
    const float* xi = xi_base + M.rand_off;
    const float* nu = nu_base ? nu_base + M.rand_off : nullptr;
    float s[3] = {0.f, 0.f, 0.f}, c[3] = {0.f, 0.f, 0.f};
    for (long long k = threadIdx.x; k < n; k += kNormNT) {
        const float a = xi[k];
        const float b = nu ? nu[k] : 0.0f;
        const float v[3] = {a * a, b * b, a * b};
#pragma unroll
        for (int q = 0; q < 3; ++q) {   // Kahan
            const float y = v[q] - c[q];
            const float t = s[q] + y;
            c[q] = (t - s[q]) - y;
            s[q] = t;
        }
    }
#pragma unroll
    for (int q = 0; q < 3; ++q) red[q][threadIdx.x] = s[q];
    __syncthreads();
    for (int w = kNormNT / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w)
#pragma unroll
            for (int q = 0; q < 3; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = red[0][0];
        if (nu) {
            out[1] = red[1][0];
            out[2] = red[2][0];
        }
    }
}

// nu <- sqrt(|nu|^2 / |nu_perp|^2) (nu - (xi.nu / |xi|^2) xi), one thread per
// element of the member's field (ocean_noise.cu:206-243).
__global__ __launch_bounds__(kRngNT) void k_noise_perp(NoiseLaunch L, const float* xi_base,
                                                       float* nu_base, const float* norms) {
    const NoiseMember& M = L.one;
    const long long k = (long long)blockIdx.x * kRngNT + threadIdx.x;
    if (k >= (long long)M.rand_nx * M.rand_ny) return;
    const float* r = norms;
    const float par = r[2] / r[0];
    const float perp = r[1] - par * r[2];
    float* nu = nu_base + M.rand_off + k;
    *nu = sqrtf(r[1] / perp) * (*nu - par * xi_base[M.rand_off + k]);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
struct Sizes {
    int cnx, cny, rnx, rny, snx, sny, gx, gy;
};

int noise_sizes(const gocean_noise_t* nz, Sizes& S) {
    if (!nz) return set_error(GOCEAN_EINVAL, "noise descriptor is NULL");
    if (nz->nx < 1 || nz->ny < 1)
        return set_error(GOCEAN_EINVAL, "noise domain too small: nx=%d ny=%d", nz->nx, nz->ny);
    const int f = nz->interpolation_factor;
    if (f < 1 || f % 2 == 0)
        return set_error(GOCEAN_EINVAL, "interpolation_factor must be a positive odd integer (got %d)", f);
    if (nz->nx % f || nz->ny % f)
        return set_error(GOCEAN_EINVAL, "nx=%d and ny=%d must be divisible by the interpolation factor %d",
                         nz->nx, nz->ny, f);
    S.cnx = nz->nx / f;
    S.cny = nz->ny / f;
    S.gx = nz->periodic_ew ? 0 : 2 + kCut;
    S.gy = nz->periodic_ns ? 0 : 2 + kCut;
    S.rnx = S.cnx + 2 * S.gx;
    S.rny = S.cny + 2 * S.gy;
    S.snx = (S.rnx + 1) / 2;
    S.sny = S.rny;
    return 0;
}

int validate_noise(const gocean_noise_t* nz, bool need_state, Sizes& S) {
    if (int e = noise_sizes(nz, S)) return e;
    if (need_state) {
        if (nz->pitch < nz->nx + 4) return set_error(GOCEAN_EINVAL, "pitch %d < nx+4 = %d", nz->pitch, nz->nx + 4);
        if (nz->hi_pitch < nz->nx + 5)
            return set_error(GOCEAN_EINVAL, "hi_pitch %d < nx+5 = %d", nz->hi_pitch, nz->nx + 5);
        if (!(nz->dx > 0.f) || !(nz->dy > 0.f)) return set_error(GOCEAN_EINVAL, "dx, dy must be positive");
    }
    if (!(nz->soar_L > 0.f)) return set_error(GOCEAN_EINVAL, "soar_L must be positive (got %g)", nz->soar_L);
    return 0;
}

void fill_member(const gocean_noise_t* nz, const Sizes& S, NoiseMember& M) {
    M = NoiseMember{};
    M.nx = nz->nx;
    M.ny = nz->ny;
    M.pitch = nz->pitch;
    M.hi_pitch = nz->hi_pitch;
    M.f_off = M.nxf_off = M.nyf_off = -1;
    M.dx = nz->dx;
    M.dy = nz->dy;
    M.g = nz->g;
    M.f = nz->f;
    M.beta = nz->beta;
    M.north_x = nz->north_x;
    M.north_y = nz->north_y;
    M.land = nz->land_value;
    M.factor = nz->interpolation_factor;
    M.cnx = S.cnx;
    M.cny = S.cny;
    M.periodic_ns = nz->periodic_ns != 0;
    M.periodic_ew = nz->periodic_ew != 0;
    // coarse_dx = float32(dx * factor) (OceanStateNoise.py:105-106)
    M.cdx = (float)((double)nz->dx * M.factor);
    M.cdy = (float)((double)nz->dy * M.factor);
    M.q0 = nz->soar_q0;
    M.rand_nx = S.rnx;
    M.rand_ny = S.rny;
    M.seed_nx = S.snx;
    M.seed_ny = S.sny;
    // constant Coriolis parameter: beta 0 makes f + beta*(...) exactly f
    M.gc_const = (nz->beta == 0.0f && nz->f != 0.0f) ? nz->g / nz->f : 0.0f;
    M.two_dx = 2.0f * nz->dx;
    M.two_dy = 2.0f * nz->dy;
    M.r_two_dx = 1.0f / M.two_dx;
    M.r_two_dy = 1.0f / M.two_dy;
    const double L = nz->soar_L;
    for (int oy = -kCut; oy <= kCut; ++oy)
        for (int ox = -kCut; ox <= kCut; ++ox) {
            const double d = std::sqrt((double)M.cdx * M.cdx * ox * ox + (double)M.cdy * M.cdy * oy * oy);
            M.w[(oy + kCut) * (2 * kCut + 1) + ox + kCut] = (float)((1.0 + d / L) * std::exp(-d / L));
        }
}

int rng_blocks(const NoiseMember& M, int kind) {
    if (kind <= GOCEAN_RNG_LCG_UNIFORM) return (M.seed_nx * M.seed_ny + kRngNT - 1) / kRngNT;
    const long long n4 = ((long long)M.rand_nx * M.rand_ny + 3) / 4;
    return (int)((n4 + kRngNT - 1) / kRngNT);
}
int perp_blocks(const NoiseMember& M) {
    return (int)(((long long)M.rand_nx * M.rand_ny + kRngNT - 1) / kRngNT);
}
void soar_tiles(const NoiseMember& M, int& tx, int& ty) {
    tx = (M.cnx + 4 + kSoarBW - 1) / kSoarBW;
    ty = (M.cny + 4 + kSoarBH - 1) / kSoarBH;
}
void bal_tiles(const NoiseMember& M, int& tx, int& ty) {
    tx = (M.nx + kBalBW - 1) / kBalBW;
    ty = (M.ny + kBalBH - 1) / kBalBH;
}
void fu_tiles(const NoiseMember& M, int& tx, int& ty) {
    tx = (M.nx + kFuBW - 1) / kFuBW;
    ty = (M.ny + kFuBH - 1) / kFuBH;
}

int check_kind(int kind) {
    if (kind < GOCEAN_RNG_LCG_NORMAL || kind > GOCEAN_RNG_PHILOX_UNIFORM)
        return set_error(GOCEAN_EINVAL, "unknown generator kind %d", kind);
    return 0;
}

int check_coriolis(const gocean_noise_t* nz, bool have_f_field) {
    if (nz->f == 0.0f && !have_f_field)
        return set_error(GOCEAN_EINVAL,
                         "f == 0 needs a Coriolis field (geostrophic balance divides by f)");
    return 0;
}

}  // namespace
}  // namespace gocean

using namespace gocean;

// ---------------------------------------------------------------------------
// single domain
// ---------------------------------------------------------------------------
int gocean_noise_sizes(const gocean_noise_t* nz, int32_t sizes[8]) {
    clear_error();
    Sizes S;
    if (int e = noise_sizes(nz, S)) return e;
    if (!sizes) return set_error(GOCEAN_EINVAL, "sizes is NULL");
    const int v[8] = {S.cnx, S.cny, S.rnx, S.rny, S.snx, S.sny, S.gx, S.gy};
    for (int k = 0; k < 8; ++k) sizes[k] = v[k];
    return 0;
}

int gocean_noise_generate(const gocean_noise_t* nz, int kind, uint64_t* seed, float* random,
                          uint64_t key, uint64_t* counter, void* stream) {
    clear_error();
    Sizes S;
    if (int e = noise_sizes(nz, S)) return e;
    if (int e = check_kind(kind)) return e;
    if (!random) return set_error(GOCEAN_EINVAL, "random is NULL");
    const bool lcg = kind <= GOCEAN_RNG_LCG_UNIFORM;
    if (lcg && !seed) return set_error(GOCEAN_EINVAL, "LCG generator needs the seed buffer");
    if (!lcg && !counter) return set_error(GOCEAN_EINVAL, "Philox generator needs the device counter");
    NoiseLaunch L{};
    fill_member(nz, S, L.one);
    L.seed = (unsigned long long*)seed;
    L.random = random;
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_noise_rng<false>, dim3(rng_blocks(L.one, kind)), dim3(kRngNT), 0, s, L, kind,
                       (const unsigned long long*)counter, (unsigned long long)key);
    if (!lcg) hipLaunchKernelGGL(k_counter_bump, dim3(1), dim3(1), 0, s, (unsigned long long*)counter);
    return check_launch("k_noise_rng");
}

int gocean_noise_soar(const gocean_noise_t* nz, float q0, float scale, const float* random,
                      float* coarse, int additive, void* stream) {
    clear_error();
    Sizes S;
    if (int e = validate_noise(nz, false, S)) return e;
    if (!random || !coarse) return set_error(GOCEAN_EINVAL, "random/coarse is NULL");
    NoiseLaunch L{};
    fill_member(nz, S, L.one);
    L.one.q0 = q0;
    int tx, ty;
    soar_tiles(L.one, tx, ty);
    L.one.soar_tiles_x = tx;
    L.random = (float*)random;
    L.coarse = coarse;
    hipLaunchKernelGGL(k_noise_soar<false>, dim3(tx * ty), dim3(kSoarBW * kSoarBH), 0, (hipStream_t)stream, L,
                       1.0f, scale, additive);
    return check_launch("k_noise_soar");
}

int gocean_noise_balance(const gocean_noise_t* nz, const float* coarse, float* eta, float* hu,
                         float* hv, const float* Hi, const gocean_fields_t* fields, int offset_i,
                         int offset_j, void* stream) {
    clear_error();
    Sizes S;
    if (int e = validate_noise(nz, true, S)) return e;
    if (!coarse || !eta || !hu || !hv || !Hi) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    const bool have_f = fields && fields->f;
    const bool have_n = fields && (fields->north_x || fields->north_y);
    if (have_n && !(fields->north_x && fields->north_y))
        return set_error(GOCEAN_EINVAL, "north_x and north_y fields must be given together");
    if (int e = check_coriolis(nz, have_f)) return e;
    const int half = nz->interpolation_factor / 2;
    if (offset_i < -half || offset_i > half || offset_j < -half || offset_j > half)
        return set_error(GOCEAN_EINVAL, "alignment offsets (%d, %d) outside [-%d, %d]", offset_i,
                         offset_j, half, half);
    NoiseLaunch L{};
    fill_member(nz, S, L.one);
    int tx, ty;
    bal_tiles(L.one, tx, ty);
    L.one.bal_tiles_x = tx;
    L.eta = eta;
    L.hu = hu;
    L.hv = hv;
    L.Hi = Hi;
    L.coarse = (float*)coarse;
    if (have_f) {
        L.ff = fields->f;
        L.one.f_off = 0;
    }
    if (have_n) {
        L.fnx = fields->north_x;
        L.fny = fields->north_y;
        L.one.nxf_off = L.one.nyf_off = 0;
    }
    hipLaunchKernelGGL(k_noise_balance<false>, dim3(tx * ty), dim3(kBalBW * kBalBH), 0, (hipStream_t)stream,
                       L, offset_i, offset_j);
    return check_launch("k_noise_balance");
}

int gocean_noise_soar_balance(const gocean_noise_t* nz, float q0, float scale, const float* random,
                              float* coarse, float* eta, float* hu, float* hv, const float* Hi,
                              const gocean_fields_t* fields, void* stream) {
    clear_error();
    Sizes S;
    if (int e = validate_noise(nz, true, S)) return e;
    if (nz->interpolation_factor != 1)
        return set_error(GOCEAN_EINVAL, "the fused SOAR + balance needs interpolation factor 1 (got %d)",
                         nz->interpolation_factor);
    if (!random || !coarse || !eta || !hu || !hv || !Hi) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    const bool have_f = fields && fields->f;
    const bool have_n = fields && (fields->north_x || fields->north_y);
    if (have_n && !(fields->north_x && fields->north_y))
        return set_error(GOCEAN_EINVAL, "north_x and north_y fields must be given together");
    if (int e = check_coriolis(nz, have_f)) return e;
    NoiseLaunch L{};
    fill_member(nz, S, L.one);
    L.one.q0 = q0;
    int tx, ty;
    fu_tiles(L.one, tx, ty);
    L.one.fu_tiles_x = tx;
    L.eta = eta;
    L.hu = hu;
    L.hv = hv;
    L.Hi = Hi;
    L.random = (float*)random;
    L.coarse = coarse;
    if (have_f) {
        L.ff = fields->f;
        L.one.f_off = 0;
    }
    if (have_n) {
        L.fnx = fields->north_x;
        L.fny = fields->north_y;
        L.one.nxf_off = L.one.nyf_off = 0;
    }
    hipLaunchKernelGGL(k_noise_fused1<false>, dim3(tx * ty), dim3(kFuNT), 0, (hipStream_t)stream,
                       L, 1.0f, scale);
    return check_launch("k_noise_fused1");
}

int gocean_noise_norms(const gocean_noise_t* nz, const float* xi, const float* nu, float* out,
                       void* stream) {
    clear_error();
    Sizes S;
    if (int e = noise_sizes(nz, S)) return e;
    if (!xi || !out) return set_error(GOCEAN_EINVAL, "xi/out is NULL");
    NoiseLaunch L{};
    fill_member(nz, S, L.one);
    hipLaunchKernelGGL(k_noise_norms, dim3(1), dim3(kNormNT), 0, (hipStream_t)stream, L, xi, nu, out);
    return check_launch("k_noise_norms");
}

int gocean_noise_make_perpendicular(const gocean_noise_t* nz, const float* xi, float* nu,
                                    const float* norms, void* stream) {
    clear_error();
    Sizes S;
    if (int e = noise_sizes(nz, S)) return e;
    if (!xi || !nu || !norms) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    NoiseLaunch L{};
    fill_member(nz, S, L.one);
    hipLaunchKernelGGL(k_noise_perp, dim3(perp_blocks(L.one)), dim3(kRngNT), 0, (hipStream_t)stream, L,
                       xi, nu, norms);
    return check_launch("k_noise_perp");
}

// ---------------------------------------------------------------------------
// batched members
// ---------------------------------------------------------------------------
struct gocean_noise_batch {
    int n = 0, kind = 0;
    std::vector<NoiseMember> host;
    NoiseMember* d_members = nullptr;
    int* d_rng_map = nullptr;
    int* d_soar_map = nullptr;
    int* d_bal_map = nullptr;
    int* d_fu_map = nullptr;
    int n_rng = 0, n_soar = 0, n_bal = 0, n_fu = 0;
    bool any_field = false;
    bool all_f1 = true;     // every member at interpolation factor 1: fused SOAR + balance
    long long totals[3] = {0, 0, 0};
};

int gocean_noise_batch_create(const gocean_noise_member_t* members, int n, int kind,
                              gocean_noise_batch_t** out) {
    clear_error();
    if (!members || n < 1 || !out) return set_error(GOCEAN_EINVAL, "need members, n >= 1 and out");
    if (int e = check_kind(kind)) return e;
    auto* B = new gocean_noise_batch;
    B->n = n;
    B->kind = kind;
    B->host.resize(n);
    std::vector<int> rng_map, soar_map, bal_map, fu_map;
    long long seed_off = 0, rand_off = 0, coarse_off = 0;
    for (int m = 0; m < n; ++m) {
        const gocean_noise_member_t& G = members[m];
        Sizes S;
        int e = validate_noise(&G.noise, true, S);
        if (!e) e = check_coriolis(&G.noise, G.field_offset >= 0);
        if (e) {
            delete B;
            return set_error(e, "member %d: %s", m, last_error());
        }
        NoiseMember& M = B->host[m];
        fill_member(&G.noise, S, M);
        M.id = m;
        const long long plane = (long long)(G.noise.ny + 4) * G.noise.pitch;
        M.eta_off = G.state_offset;
        M.hu_off = G.state_offset + plane;
        M.hv_off = G.state_offset + 2 * plane;
        M.hi_off = G.hi_offset;
        if (G.field_offset >= 0) {
            M.f_off = G.field_offset;
            M.nxf_off = G.field_offset + plane;
            M.nyf_off = G.field_offset + 2 * plane;
            B->any_field = true;
        }
        // 64-element alignment of every member's block
        M.seed_off = seed_off;
        M.rand_off = rand_off;
        M.coarse_off = coarse_off;
        seed_off += ((long long)M.seed_nx * M.seed_ny + 63) / 64 * 64;
        rand_off += ((long long)M.rand_nx * M.rand_ny + 63) / 64 * 64;
        coarse_off += ((long long)(M.cnx + 4) * (M.cny + 4) + 63) / 64 * 64;
        M.rng_begin = (int)rng_map.size();
        rng_map.insert(rng_map.end(), rng_blocks(M, kind), m);
        if (M.factor != 1) B->all_f1 = false;
        int tx, ty;
        soar_tiles(M, tx, ty);
        M.soar_begin = (int)soar_map.size();
        M.soar_tiles_x = tx;
        soar_map.insert(soar_map.end(), tx * ty, m);
        bal_tiles(M, tx, ty);
        M.bal_begin = (int)bal_map.size();
        M.bal_tiles_x = tx;
        bal_map.insert(bal_map.end(), tx * ty, m);
        fu_tiles(M, tx, ty);
        M.fu_begin = (int)fu_map.size();
        M.fu_tiles_x = tx;
        fu_map.insert(fu_map.end(), tx * ty, m);
    }
    B->totals[0] = seed_off;
    B->totals[1] = rand_off;
    B->totals[2] = coarse_off;
    B->n_rng = (int)rng_map.size();
    B->n_soar = (int)soar_map.size();
    B->n_bal = (int)bal_map.size();
    B->n_fu = (int)fu_map.size();
    bool ok = hipMalloc(&B->d_members, sizeof(NoiseMember) * n) == hipSuccess &&
              hipMalloc(&B->d_rng_map, sizeof(int) * rng_map.size()) == hipSuccess &&
              hipMalloc(&B->d_soar_map, sizeof(int) * soar_map.size()) == hipSuccess &&
              hipMalloc(&B->d_bal_map, sizeof(int) * bal_map.size()) == hipSuccess &&
              hipMalloc(&B->d_fu_map, sizeof(int) * fu_map.size()) == hipSuccess;
    ok = ok &&
         hipMemcpy(B->d_members, B->host.data(), sizeof(NoiseMember) * n, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->d_rng_map, rng_map.data(), sizeof(int) * rng_map.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->d_soar_map, soar_map.data(), sizeof(int) * soar_map.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->d_bal_map, bal_map.data(), sizeof(int) * bal_map.size(), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(B->d_fu_map, fu_map.data(), sizeof(int) * fu_map.size(), hipMemcpyHostToDevice) == hipSuccess;
    if (!ok) {
        gocean_noise_batch_destroy(B);
        (void)hipGetLastError();
        return set_error(GOCEAN_ENOMEM, "device allocation of the noise member table failed");
    }
    *out = B;
    return 0;
}

int gocean_noise_batch_destroy(gocean_noise_batch_t* B) {
    if (!B) return 0;
    (void)hipFree(B->d_members);
    (void)hipFree(B->d_rng_map);
    (void)hipFree(B->d_soar_map);
    (void)hipFree(B->d_bal_map);
    (void)hipFree(B->d_fu_map);
    delete B;
    return 0;
}

int gocean_noise_batch_layout(const gocean_noise_batch_t* B, int64_t* offsets, int64_t totals[3]) {
    clear_error();
    if (!B) return set_error(GOCEAN_EINVAL, "batch is NULL");
    if (offsets)
        for (int m = 0; m < B->n; ++m) {
            offsets[3 * m] = B->host[m].seed_off;
            offsets[3 * m + 1] = B->host[m].rand_off;
            offsets[3 * m + 2] = B->host[m].coarse_off;
        }
    if (totals)
        for (int k = 0; k < 3; ++k) totals[k] = B->totals[k];
    return 0;
}

static NoiseLaunch batch_launch(const gocean_noise_batch_t* B) {
    NoiseLaunch L{};
    L.members = B->d_members;
    return L;
}

// bump: advance the Philox counter in its own launch (else the caller's next
// launch does)
static int batch_generate(gocean_noise_batch_t* B, uint64_t* seed, float* random, uint64_t key,
                          uint64_t* counter, void* stream, bool bump) {
    if (!B || !random) return set_error(GOCEAN_EINVAL, "batch/random is NULL");
    const bool lcg = B->kind <= GOCEAN_RNG_LCG_UNIFORM;
    if (lcg && !seed) return set_error(GOCEAN_EINVAL, "LCG generator needs the seed buffer");
    if (!lcg && !counter) return set_error(GOCEAN_EINVAL, "Philox generator needs the device counter");
    NoiseLaunch L = batch_launch(B);
    L.map = B->d_rng_map;
    L.seed = (unsigned long long*)seed;
    L.random = random;
    const hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(k_noise_rng<true>, dim3(B->n_rng), dim3(kRngNT), 0, s, L, B->kind,
                       (const unsigned long long*)counter, (unsigned long long)key);
    if (!lcg && bump) hipLaunchKernelGGL(k_counter_bump, dim3(1), dim3(1), 0, s, (unsigned long long*)counter);
    return check_launch("k_noise_rng (batch)");
}

int gocean_noise_batch_generate(gocean_noise_batch_t* B, uint64_t* seed, float* random, uint64_t key,
                                uint64_t* counter, void* stream) {
    clear_error();
    return batch_generate(B, seed, random, key, counter, stream, true);
}

int gocean_noise_batch_perturb(gocean_noise_batch_t* B, float* state, const float* Hi,
                               const float* fields, float q0_scale, float perturbation_scale,
                               int update_random, uint64_t* seed, float* random, float* coarse,
                               uint64_t key, uint64_t* counter, void* stream) {
    clear_error();
    if (!B || !state || !Hi || !random || !coarse) return set_error(GOCEAN_EINVAL, "a buffer is NULL");
    if (B->any_field && !fields) return set_error(GOCEAN_EINVAL, "members use fields but fields is NULL");
    const bool philox = B->kind > GOCEAN_RNG_LCG_UNIFORM;
    if (update_random)
        if (int e = batch_generate(B, seed, random, key, counter, stream, false)) return e;
    const hipStream_t s = (hipStream_t)stream;
    NoiseLaunch L = batch_launch(B);
    if (update_random && philox) L.bump = (unsigned long long*)counter;   // advanced by the next launch
    L.random = random;
    L.coarse = coarse;
    L.eta = L.hu = L.hv = state;
    L.Hi = Hi;
    L.ff = L.fnx = L.fny = fields;
    if (B->all_f1) {
        L.map = B->d_fu_map;
        hipLaunchKernelGGL(k_noise_fused1<true>, dim3(B->n_fu), dim3(kFuNT), 0, s, L, q0_scale,
                           perturbation_scale);
        return check_launch("model error (batch, fused)");
    }
    L.map = B->d_soar_map;
    hipLaunchKernelGGL(k_noise_soar<true>, dim3(B->n_soar), dim3(kSoarBW * kSoarBH), 0, s, L, q0_scale,
                       perturbation_scale, 0);
    L.map = B->d_bal_map;
    hipLaunchKernelGGL(k_noise_balance<true>, dim3(B->n_bal), dim3(kBalBW * kBalBH), 0, s, L, 0, 0);
    return check_launch("model error (batch)");
}
