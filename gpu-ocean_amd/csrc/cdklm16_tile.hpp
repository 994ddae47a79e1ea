// cdklm16_tile.hpp -- one CDKLM16 Runge-Kutta substep on one output tile.
//
// Device code for gfx950.  The arithmetic restates
// gpu_ocean/SWESimulators/gpu_kernels/CDKLM16_kernel.cu:80-1006 and
// common.cu:532-551,822-824 with the reference's operand order and rounding
// points (the reference builds with --fmad=false; this file must be built
// with -ffp-contract=off), and every division and square root is correctly
// rounded, so that a float32 result equals the reference's on the same inputs.
//
// The structure is this library's own:
//   * one workgroup owns a BW x BH tile of output cells; the state tile with a
//     2-cell halo is loaded once, with wall (mirror) and periodic (wrap)
//     boundaries applied by index remapping at load time, so no ghost-cell
//     kernels run between substeps;
//   * (eta, u, v) conversion, x/y slopes (with the positivity adjustment) and
//     every face flux are computed exactly once per tile and kept in LDS; the
//     reference evaluates each face from both sides (4 flux evaluations per
//     cell) and keeps only one slope set live at a time;
//   * the cell depth Hm is rebuilt from the corner depths in LDS, so the
//     substep reads 3 state fields + Hi and writes 3 fields;
//   * work items of every phase are spread over all NT threads of the group.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gocean {

constexpr float kDryFlag = 1.0e-30f;   // CDKLM16_kernel.cu:37
constexpr float kDryEps = 1.0e-3f;     // CDKLM16_kernel.cu:38
constexpr int kWall = 1;
constexpr int kPeriodic = 2;

// Everything one substep of one domain needs.  Plain launches pass it as a
// kernel argument; the ensemble kernel builds it per workgroup from a table.
struct SubstepArgs {
    const float* eta_in; const float* hu_in; const float* hv_in;
    const float* eta_qn; const float* hu_qn; const float* hv_qn;
    float* eta_out; float* hu_out; float* hv_out;
    const float* Hi;
    const float* f_field; const float* nx_field; const float* ny_field;
    const float* dt_dev;
    int nx, ny, pitch, hi_pitch;
    float dx, dy, g, r, theta, desing_eps, flux_eps, cutoff, land;
    float f, north_x, north_y;
    // wind stress at the substep's time, interpolated on the host
    // (wind_lerp: the kernel's former expression, so the same bits)
    float wind_x, wind_y;
    // gridded wind (nullptr: the uniform wind above): X0, X1, Y0, Y1 planes
    // of wind_wy x wind_wx, time weight wind_t.  Initialised here so that
    // every construction of the struct (host or device) starts uniform.
    const float* wind_grid = nullptr;
    int wind_wx = 0, wind_wy = 0;
    float wind_t = 0.0f;
    float dt;
    // bits 0-7: north, east, south, west boundary codes (2 bits each); bits
    // 8-9: rk_order; bits 10-11: rk_step.  One word instead of six keeps the
    // kernel within the scalar registers of 8 waves per SIMD.
    uint32_t flags;
    int tiles_x, tiles_y;
    int march_bh = 8;          // rows per strip of the column-marching kernel
    __host__ __device__ int bc_n() const { return (int)(flags & 3u); }
    __host__ __device__ int bc_e() const { return (int)((flags >> 2) & 3u); }
    __host__ __device__ int bc_s() const { return (int)((flags >> 4) & 3u); }
    __host__ __device__ int bc_w() const { return (int)((flags >> 6) & 3u); }
    __host__ __device__ int rk_order() const { return (int)((flags >> 8) & 3u); }
    __host__ __device__ int rk_step() const { return (int)((flags >> 10) & 3u); }
    // bit 12: divisions by dx, dy and 2g may take one residual correction
    __host__ __device__ bool fast_arith() const { return (flags >> 12) & 1u; }
    __host__ __device__ void set_fast_arith(bool on) { flags = (flags & ~(1u << 12)) | ((on ? 1u : 0u) << 12); }
    __host__ __device__ void set_flags(int n, int e, int s_, int w, int order, int step) {
        flags = (uint32_t)(n & 3) | ((uint32_t)(e & 3) << 2) | ((uint32_t)(s_ & 3) << 4) |
                ((uint32_t)(w & 3) << 6) | ((uint32_t)(order & 3) << 8) | ((uint32_t)(step & 3) << 10);
    }
    __host__ __device__ void set_rk_step(int step) {
        flags = (flags & ~(3u << 10)) | ((uint32_t)(step & 3) << 10);
    }
    // derived by substep_derive on the host, once per launch instead of in
    // every thread: dx/(2g), dy/(2g), 1/(2g), 1/dx, 1/dy (IEEE divisions on
    // either side, so the same bits)
    float vdx, udy, r_two_g, r_dx, r_dy;
    // branch record of the substep (written only by the recording variant):
    // per padded cell u, v, the x / y face depth terms and a code word
    float* tape_u; float* tape_v; float* tape_t; float* tape_s;
    uint32_t* tape_code;
};

// wind stress at interpolation weight t between two samples
__host__ __device__ inline float wind_lerp(float t, float a0, float a1) { return t * a1 + (1.0f - t) * a0; }

// One axis of the reference's wind texture lookup (common.cu:616-650 from
// CDKLM16_kernel.cu:827-828): normalised coordinate s = (ti + 0.5) / n of
// the ghost-inclusive cell index ti, texel coordinate s * w - 0.5 (texel
// centres at (k + 0.5) / w), clamped neighbours and the linear weight.
struct WindAxis {
    int i0, i1;
    float a;
};

__device__ __forceinline__ WindAxis wind_axis(int ti, int n, int w) {
    const float s = ((float)ti + 0.5f) / (float)n;
    const float x = s * (float)w - 0.5f;
    const float x0 = floorf(x);
    WindAxis o;
    o.a = x - x0;
    const int k = (int)x0;
    o.i0 = min(max(k, 0), w - 1);
    o.i1 = min(max(k + 1, 0), w - 1);
    return o;
}

// bilinear value of one wind plane (row length w) at (ax, ay)
__device__ __forceinline__ float wind_bilinear(const float* F, int w, const WindAxis& ax, const WindAxis& ay) {
    const float a = ax.a, b = ay.a;
    const float* r0 = F + (size_t)ay.i0 * w;
    const float* r1 = F + (size_t)ay.i1 * w;
    return ((((1.0f - a) * (1.0f - b)) * r0[ax.i0] + (a * (1.0f - b)) * r0[ax.i1])
            + ((1.0f - a) * b) * r1[ax.i0]) + (a * b) * r1[ax.i1];
}

// gridded wind stress (X, Y) of one cell, interpolated in time
__device__ __forceinline__ void wind_grid_at(const float* grid, int wx, int wy, float wt,
                                             const WindAxis& ax, const WindAxis& ay, float& X, float& Y) {
    const size_t pl = (size_t)wx * wy;
    const float x0 = wind_bilinear(grid, wx, ax, ay), x1 = wind_bilinear(grid + pl, wx, ax, ay);
    const float y0 = wind_bilinear(grid + 2 * pl, wx, ax, ay), y1 = wind_bilinear(grid + 3 * pl, wx, ax, ay);
    X = wind_lerp(wt, x0, x1);
    Y = wind_lerp(wt, y0, y1);
}

inline void substep_derive(SubstepArgs& A) {
    const float two_g = 2.0f * A.g;
    A.vdx = A.dx / two_g;
    A.udy = A.dy / two_g;
    A.r_two_g = 1.0f / two_g;
    A.r_dx = 1.0f / A.dx;
    A.r_dy = 1.0f / A.dy;
}

// ---------------------------------------------------------------------------
// scalar helpers
// ---------------------------------------------------------------------------

// minmod of (b, c, f): 0.25*s(b)*(s(b)+s(c))*(s(c)+s(f))*min(|b|,|c|,|f|)
// (common.cu:544-551).  The sign product is exactly +1, -1 or 0, so the value
// is min(|b|,|c|,|f|) with b's sign when all three sign bits agree, else zero.
__host__ __device__ __forceinline__ float minmod_raw(float b, float c, float f) {
    const float m = fminf(fminf(fabsf(b), fabsf(c)), fabsf(f));
    const uint32_t ub = __builtin_bit_cast(uint32_t, b), uc = __builtin_bit_cast(uint32_t, c);
    const uint32_t uf = __builtin_bit_cast(uint32_t, f);
    const bool agree = (((ub ^ uc) | (uc ^ uf)) & 0x80000000u) == 0u;
    return agree ? copysignf(m, b) : 0.0f;
}

// common.cu:532-542
__host__ __device__ __forceinline__ float minmod_slope(float l, float c, float r, float theta) {
    return minmod_raw((c - l) * theta, (r - l) * 0.5f, (r - c) * theta);
}

// which candidate minmod_raw(b, c, f) returns: 0, 1, 2, or 3 when the sign
// bits disagree (result zero).  Ties pick the first; the value is the same.
__host__ __device__ __forceinline__ uint32_t minmod_sel(float b, float c, float f) {
    const uint32_t ub = __builtin_bit_cast(uint32_t, b), uc = __builtin_bit_cast(uint32_t, c);
    const uint32_t uf = __builtin_bit_cast(uint32_t, f);
    if ((((ub ^ uc) | (uc ^ uf)) & 0x80000000u) != 0u) return 3u;
    const float ab = fabsf(b), ac = fabsf(c), af = fabsf(f);
    if (ab <= ac && ab <= af) return 0u;
    if (ac <= af) return 1u;
    return 2u;
}

// Branch record of one cell (tape code word):
//   bits  0- 7: x slope centre: u slope sel, v slope sel, K sel, adjustment
//   bits  8-15: y slope centre: u slope sel, v slope sel, L sel, adjustment
//   bits 16-20: flux of the cell's east face, bits 21-25: of its north face
// sel as minmod_sel; adjustment 0 none, 1 low side, 2 high side (overrides 1);
// flux: a+ source (0 zero, 1 minus side, 2 plus side), a- source (same), and
// bit 4 when |a+ - a-| < flux_eps (zero flux).
constexpr int kCodeX = 0, kCodeY = 8, kCodeFx = 16, kCodeFy = 21;

// common.cu:822-824.  The reference's "0.5*eps" is a double literal, so that
// one addition is evaluated in double there and rounded to float.  The
// float32 sum below gives the same bits.  With c = 0.5*eps (exact in float)
// of exponent E and 0 <= a = h*h/(2 eps) < c (the branch needs |h| < eps):
// if a >= 2^(E-28), every bit of a and c is a multiple of 2^(E-51) and the
// sum is below 2^(E+2), so the double sum is exact and one rounding to float
// remains; if a < 2^(E-28), a is below half an ulp of c and both forms give
// c.  (Checked exhaustively over every float32 h in [2^-17, 0.1) for
// eps = 0.1 by the test suite.)
__host__ __device__ __forceinline__ float desingularize(float h, float hu, float eps) {
    const float a = (h * h) / (2.0f * eps);
    const float b = a + 0.5f * eps;
    return hu / fmaxf(fminf(b, eps), fabsf(h));
}

// Correctly rounded float division and square root without the range
// scaling of the generic library sequences.  The scaled forms only differ for
// operands near the overflow / denormal boundaries, which this scheme never
// produces (depths >= 1e-5, speeds and fluxes far inside the normal range).
//
// rcp_refined: hardware reciprocal plus one Newton step, the same reciprocal
// the generic division sequence builds.
__device__ __forceinline__ float rcp_refined(float d) {
    const float r = __builtin_amdgcn_rcpf(d);
    const float e = fmaf(-d, r, 1.0f);
    return fmaf(e, r, r);
}

// n/d from a reciprocal r of d that is RN(1/d) or rcp_refined(d): two
// residual corrections (Markstein), identical to the core of the generic
// sequence, give the correctly rounded quotient.
__device__ __forceinline__ float div_r(float n, float d, float r) {
    float q = n * r;
    float e = fmaf(-d, q, n);
    q = fmaf(e, r, q);
    e = fmaf(-d, q, n);
    return fmaf(e, r, q);
}

// n/d for a launch-constant divisor d (dx, dy, 2g) with r = RN(1/d) formed on
// the host: one residual correction.  The host enables it (SubstepArgs
// fast_arith()) only for divisors for which it has checked, over every float
// significand of n, that this gives the correctly rounded quotient (see
// fast_arith_exact in gocean_fwd.hip); otherwise div_r's two corrections are
// used.
__device__ __forceinline__ float div_c1(float n, float d, float r) {
    const float q = n * r;
    const float e = fmaf(-d, q, n);
    return fmaf(e, r, q);
}

template <bool FAST>
__device__ __forceinline__ float div_k(float n, float d, float r) {
    if constexpr (FAST) return div_c1(n, d, r);
    else return div_r(n, d, r);
}

// correctly rounded sqrt for x == 0 or x >= 2^-96: hardware estimate, then
// pick among the two neighbours by the sign of the exact residual.
__device__ __forceinline__ float sqrt_rn(float x) {
    float s = __builtin_amdgcn_sqrtf(x);
    const float sm = __uint_as_float(__float_as_uint(s) - 1u);
    const float sp = __uint_as_float(__float_as_uint(s) + 1u);
    const float vm = fmaf(-sm, s, x);
    const float vp = fmaf(-sp, s, x);
    s = (vm <= 0.0f) ? sm : s;
    s = (vp > 0.0f) ? sp : s;
    return s;
}

// sqrt(x) from the hardware reciprocal square root y and one Newton step:
// g = RN(x y), r = x - g^2 (one fma), result RN(g + r * y/2).  Checked
// exhaustively on the device against sqrt_rn (gocean_sqrt_newton_check): the
// same bits for every float in [2^-60, FLT_MAX]; near the bottom of the
// normal range they differ.  Five instructions against sqrt_rn's nine.  The
// kernels take it (FAST) only where the argument is g h with h above the
// depth cutoff, and the host enables FAST only when RN(g cutoff) >= 2^-60
// (fast_arith_exact in gocean_fwd.hip).
__device__ __forceinline__ float sqrt_newton(float x) {
    const float y = __builtin_amdgcn_rsqf(x);
    const float g = x * y;
    const float h = 0.5f * y;
    const float r = fmaf(-g, g, x);
    return fmaf(r, h, g);
}

template <bool FAST>
__device__ __forceinline__ float sqrt_k(float x) {
    if constexpr (FAST) return sqrt_newton(x);
    else return sqrt_rn(x);
}

struct Consts {
    float g, half_g, two_g, theta, theta_g;
    float dx, dy, vdx, udy;        // vdx = DX/(2g), udy = DY/(2g)
    float desing, flux_eps, cutoff, land;
    float r_two_g, r_dx, r_dy;     // RN(1/(2g)), RN(1/dx), RN(1/dy)
};

// Central-upwind flux of primitive face states Qm=(hm,um,vm), Qp=(hp,up,vp)
// (CDKLM16_F_func + CDKLM16_flux, CDKLM16_kernel.cu:80-140).
template <bool REC = false>
__device__ __forceinline__ void cu_flux(float hm, float um, float vm,
                                        float hp, float up, float vp,
                                        const Consts& K,
                                        float& Fx, float& Fy, float& Fz,
                                        uint32_t* code = nullptr) {
    float Fpx = 0.0f, Fpy = 0.0f, Fpz = 0.0f, u_p = 0.0f, c_p = 0.0f;
    if (hp > K.cutoff) {
        const float q = hp * up;
        Fpx = q;
        Fpy = q * up + (K.half_g * hp) * hp;
        Fpz = q * vp;
        u_p = up;
        c_p = sqrt_rn(K.g * hp);
    }
    float Fmx = 0.0f, Fmy = 0.0f, Fmz = 0.0f, u_m = 0.0f, c_m = 0.0f;
    if (hm > K.cutoff) {
        const float q = hm * um;
        Fmx = q;
        Fmy = q * um + (K.half_g * hm) * hm;
        Fmz = q * vm;
        u_m = um;
        c_m = sqrt_rn(K.g * hm);
    }
    const float am = fminf(fminf(u_m - c_m, u_p - c_p), 0.0f);
    const float ap = fmaxf(fmaxf(u_m + c_m, u_p + c_p), 0.0f);
    const float den = ap - am;
    if constexpr (REC) {
        const float b1 = u_m + c_m, b2 = u_p + c_p, a1 = u_m - c_m, a2 = u_p - c_p;
        const uint32_t cp = (fmaxf(b1, b2) > 0.0f) ? ((b1 >= b2) ? 1u : 2u) : 0u;
        const uint32_t cm = (fminf(a1, a2) < 0.0f) ? ((a1 <= a2) ? 1u : 2u) : 0u;
        *code = cp | (cm << 2) | ((fabsf(den) < K.flux_eps) ? 16u : 0u);
    }
    if (fabsf(den) < K.flux_eps) {
        Fx = 0.0f; Fy = 0.0f; Fz = 0.0f;
        return;
    }
    const float apam = ap * am;
    const float rden = rcp_refined(den);
    Fx = div_r((ap * Fmx - am * Fpx) + apam * (hp - hm), den, rden);
    Fy = div_r((ap * Fmy - am * Fpy) + apam * (Fpx - Fmx), den, rden);
    Fz = (um + up > 0.0f) ? Fmz : Fpz;
}

// ---------------------------------------------------------------------------
// LDS layout of one tile
// ---------------------------------------------------------------------------
// MODE selects how the Coriolis parameter and the north vector enter:
//   kModeZero  constant f, north = (0, 1) (zero grid rotation): with this
//              north vector the reference's face term (K + dx*f*v_north)/2g
//              equals the cell's source-term expression (K + dx*f*v)/2g, so it
//              is computed once per cell (T, S below) instead of per face;
//   kModeConst constant f and any constant north vector;
//   kModeField per-cell f and north (the reference's textures).  The
//              reference evaluates each face with the north vector of the
//              thread's own cell, so the two neighbours of a face may disagree;
//              both evaluations are kept.
enum { kModeZero = 0, kModeConst = 1, kModeField = 2 };

template <int BW, int BH, int MODE, bool TAPE = false>
struct TileSmem {
    static constexpr bool FIELD = MODE == kModeField;
    // branch codes of the recording variant (see kCodeX ...), 8 bits each
    // (slope codes) or 5 (flux codes): bytes keep the 32x16 tile under 40 KB
    uint8_t CodeX[TAPE ? BH : 1][TAPE ? BW + 2 : 1];
    uint8_t CodeY[TAPE ? BH + 2 : 1][TAPE ? BW : 1];
    uint8_t CodeF[TAPE ? BH : 1][TAPE ? BW + 1 : 1];
    uint8_t CodeG[TAPE ? BH + 1 : 1][TAPE ? BW : 1];
    float Hi[BH + 5][BW + 5];          // corner depths, land flagged
    float E[BH + 4][BW + 4];           // eta (dry / land adjusted)
    float U[BH + 4][BW + 4];           // u
    float V[BH + 4][BW + 4];           // v
    // x slopes of u, v and K (kModeZero: T = (K + dx f v_flip)/2g instead of K)
    float Ux[BH][BW + 2], Vx[BH][BW + 2], Kx[BH][BW + 2];
    // y slopes of u, v and L (kModeZero: S = (L - dy f u_flip)/2g instead of L)
    float Uy[BH + 2][BW], Vy[BH + 2][BW], Ly[BH + 2][BW];
    float F[3][BH][BW + 1];            // x faces (west cell's north vector)
    float G[3][BH + 1][BW];            // y faces (south cell's east vector)
    float Fc[FIELD ? BH + 4 : 1][FIELD ? BW + 4 : 1];
    float Nx[FIELD ? BH + 4 : 1][FIELD ? BW + 4 : 1];
    float Ny[FIELD ? BH + 4 : 1][FIELD ? BW + 4 : 1];
    float F2[FIELD ? 3 : 1][FIELD ? BH : 1][FIELD ? BW + 1 : 1];
    float G2[FIELD ? 3 : 1][FIELD ? BH + 1 : 1][FIELD ? BW : 1];
};

// Map a workgroup index to a tile index so that each of the 8 XCDs (which
// receive workgroups round-robin) works on one contiguous band of tiles and
// shares halo rows through its own L2.
__device__ __forceinline__ int xcd_band(int b, int nb) {
    const int q = nb >> 3, rem = nb & 7;
    const int xcd = b & 7, idx = b >> 3;
    return (xcd < rem) ? xcd * (q + 1) + idx : rem * (q + 1) + (xcd - rem) * q + idx;
}

__device__ __forceinline__ Consts make_consts(const SubstepArgs& A) {
    Consts K;
    K.g = A.g;
    K.half_g = 0.5f * A.g;
    K.two_g = 2.0f * A.g;
    K.theta = A.theta;
    K.theta_g = A.theta * A.g;
    K.dx = A.dx;
    K.dy = A.dy;
    K.vdx = A.vdx;
    K.udy = A.udy;
    K.desing = A.desing_eps;
    K.flux_eps = A.flux_eps;
    K.cutoff = A.cutoff;
    K.land = A.land;
    K.r_two_g = A.r_two_g;
    K.r_dx = A.r_dx;
    K.r_dy = A.r_dy;
    return K;
}

// ---------------------------------------------------------------------------
// the tile
// ---------------------------------------------------------------------------
// A tile whose every access (padded rows by-1 .. by+BH+4, columns bx-1 ..
// bx+BW+4) lies at least one cell inside the interior rows 2 .. ny+1 and
// columns 2 .. nx+1: no boundary remap, wall flip, clamp or periodic ghost
// write applies to it.
template <int BW, int BH>
__device__ __forceinline__ bool tile_is_interior(const SubstepArgs& A, int tile_x, int tile_y) {
    const int bx = tile_x * BW, by = tile_y * BH;
    return bx >= 4 && by >= 4 && bx + BW + 4 <= A.nx && by + BH + 4 <= A.ny;
}

// EDGE = false is the body for tiles with tile_is_interior: the boundary
// flags are constant false and the clamps are dropped, which changes no
// result there and removes their instructions.
template <int BW, int BH, int NT, int MODE, bool TAPE = false, bool EDGE = true>
__device__ __forceinline__ void cdklm_tile(const SubstepArgs& A, int tile_x, int tile_y,
                                           TileSmem<BW, BH, MODE, TAPE>& S) {
    static_assert(!TAPE || MODE == kModeZero, "the branch record is defined for zero rotation");
    constexpr bool FIELD = MODE == kModeField;
    constexpr bool ZERO = MODE == kModeZero;
    const int tid = threadIdx.x;
    const int nx = A.nx, ny = A.ny, pitch = A.pitch;
    const int bx = tile_x * BW;     // padded column of the tile's first halo column
    const int by = tile_y * BH;     // padded row of the tile's first halo row
    const Consts K = make_consts(A);

    const bool wN = EDGE && A.bc_n() == kWall, wS = EDGE && A.bc_s() == kWall;
    const bool wE = EDGE && A.bc_e() == kWall, wW = EDGE && A.bc_w() == kWall;
    const bool pNS = EDGE && A.bc_n() == kPeriodic, pEW = EDGE && A.bc_e() == kPeriodic;

    // ---- phase A: corner depths (CDKLM16_kernel.cu:509-525) ----
    for (int it = tid; it < (BH + 5) * (BW + 5); it += NT) {
        const int r = it / (BW + 5), c = it - r * (BW + 5);
        const int gr = EDGE ? min(by + r, ny + 4) : by + r, gc = EDGE ? min(bx + c, nx + 4) : bx + c;
        float h = A.Hi[(size_t)gr * A.hi_pitch + gc];
        if (fabsf(h - K.land) < kDryEps) h = kDryFlag;
        S.Hi[r][c] = h;
    }
    __syncthreads();

    // ---- phase B: load state with boundary remap, convert to (eta,u,v)
    //      (:490-505, handleWallBC :373-424, :551-592) ----
    // Only the cross-shaped part of the halo is ever used; the four 2x2
    // corner blocks are skipped.
    constexpr int kMid = BH * (BW + 4);
    constexpr int kConv = kMid + 4 * BW;
    for (int it = tid; it < kConv; it += NT) {
        int rr, cc;
        if (it < kMid) {
            rr = 2 + it / (BW + 4);
            cc = it - (rr - 2) * (BW + 4);
        } else {
            const int k = it - kMid;
            const int band = k / BW;
            rr = band < 2 ? band : BH + band;     // 0,1,BH+2,BH+3
            cc = 2 + (k - band * BW);
        }
        const int R = EDGE ? min(by + rr, ny + 3) : by + rr, C = EDGE ? min(bx + cc, nx + 3) : bx + cc;
        int sr = R, sc = C;
        bool neg_u = false, neg_v = false;
        if (R < 2) {
            if (wS) { sr = 3 - R; neg_v = true; } else if (pNS) { sr = R + ny; }
        } else if (R > ny + 1) {
            if (wN) { sr = 2 * ny + 3 - R; neg_v = true; } else if (pNS) { sr = R - ny; }
        }
        if (C < 2) {
            if (wW) { sc = 3 - C; neg_u = true; } else if (pEW) { sc = C + nx; }
        } else if (C > nx + 1) {
            if (wE) { sc = 2 * nx + 3 - C; neg_u = true; } else if (pEW) { sc = C - nx; }
        }
        const size_t src = (size_t)sr * pitch + sc;
        const float eta = A.eta_in[src];
        float hu = A.hu_in[src];
        float hv = A.hv_in[src];
        if (neg_u) hu = -hu;
        if (neg_v) hv = -hv;
        // cell depth as initBm computes it (initBm_kernel.cu:84-105)
        const float a = S.Hi[rr][cc], b = S.Hi[rr + 1][cc];
        const float c = S.Hi[rr][cc + 1], d = S.Hi[rr + 1][cc + 1];
        const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag)
                        - (float)(c == kDryFlag) - (float)(d == kDryFlag);
        const float sum = ((a + b) + c) + d;
        const float Hm = (wet == 4.0f) ? sum * 0.25f : (wet == 0.0f ? K.land : sum / wet);
        const float h = eta + Hm;
        float E, U, V;
        if (fabsf(Hm - K.land) <= kDryEps) {
            E = kDryFlag; U = 0.0f; V = 0.0f;
        } else if (h < K.desing) {
            if (h <= K.cutoff) {
                E = -Hm + K.cutoff; U = 0.0f; V = 0.0f;
            } else {
                E = eta;
                U = desingularize(h, hu, K.desing);
                V = desingularize(h, hv, K.desing);
            }
        } else {
            E = eta;
            const float rh = rcp_refined(h);
            U = div_r(hu, h, rh);
            V = div_r(hv, h, rh);
        }
        S.E[rr][cc] = E;
        S.U[rr][cc] = U;
        S.V[rr][cc] = V;
        if constexpr (FIELD) {
            const size_t o = (size_t)R * pitch + C;   // textures are read unremapped
            S.Fc[rr][cc] = A.f_field[o];
            S.Nx[rr][cc] = A.nx_field[o];
            S.Ny[rr][cc] = A.ny_field[o];
        }
    }
    __syncthreads();

    auto fcell = [&](int rr, int cc) -> float {
        if constexpr (FIELD) return S.Fc[rr][cc];
        return A.f;
    };
    auto nxcell = [&](int rr, int cc) -> float {
        if constexpr (FIELD) return S.Nx[rr][cc];
        return A.north_x;
    };
    auto nycell = [&](int rr, int cc) -> float {
        if constexpr (FIELD) return S.Ny[rr][cc];
        return A.north_y;
    };
    // geostrophic sign flip of the tangential velocity in wall ghost cells
    auto ghost_col_flip = [&](int C) -> bool {
        return (wW && C < 2) || (wE && C > nx + 1);
    };
    auto ghost_row_flip = [&](int R) -> bool {
        return (wS && R < 2) || (wN && R > ny + 1);
    };

    // ---- phase C: slopes and K/L reconstruction (:607-666, :148-191,
    //      :711-770, :198-241) ----
    constexpr int kSx = BH * (BW + 2);
    constexpr int kSy = (BH + 2) * BW;
    for (int it = tid; it < kSx + kSy; it += NT) {
        if (it < kSx) {
            const int j = it / (BW + 2), i = it - j * (BW + 2);
            const int rr = j + 2, cc = i + 1;
            const int C = bx + cc;
            const float el = S.E[rr][cc - 1], ec = S.E[rr][cc], er = S.E[rr][cc + 1];
            const float ul = S.U[rr][cc - 1], uc = S.U[rr][cc], ur = S.U[rr][cc + 1];
            float vl = S.V[rr][cc - 1], vc = S.V[rr][cc], vr = S.V[rr][cc + 1];
            S.Ux[j][i] = minmod_slope(ul, uc, ur, K.theta);
            S.Vx[j][i] = minmod_slope(vl, vc, vr, K.theta);
            const float vc_raw = vc;
            const bool flip_c = ghost_col_flip(C);
            if (ghost_col_flip(C - 1)) vl = -vl;
            if (flip_c) vc = -vc;
            if (ghost_col_flip(C + 1)) vr = -vr;
            const float fc = fcell(rr, cc);
            float fv_l, fv_c, fv_r;
            if constexpr (ZERO) {
                // (0*u + 1*v)*f == v*f up to the sign of a zero
                fv_l = vl * fc; fv_c = vc * fc; fv_r = vr * fc;
            } else {
                const float n_x = nxcell(rr, cc), n_y = nycell(rr, cc);
                fv_l = (n_x * ul + n_y * vl) * fcell(rr, cc - 1);
                fv_c = (n_x * uc + n_y * vc) * fc;
                fv_r = (n_x * ur + n_y * vr) * fcell(rr, cc + 1);
            }
            const float bw_ = K.theta_g * ((ec - el) - K.vdx * (fv_c + fv_l));
            const float ce_ = K.half_g * ((er - el) - K.vdx * ((fv_r + 2.0f * fv_c) + fv_l));
            const float fw_ = K.theta_g * ((er - ec) - K.vdx * (fv_c + fv_r));
            float Kx = minmod_raw(bw_, ce_, fw_);
            // positivity adjustment; the reference flips v only in west ghost
            // column 1 here (its east test, bx+k > NX+2, never hits a column
            // an interior cell reads)
            const bool flip_a = wW && C < 2;
            const float dfc = K.dx * fc;
            const float dxfv = dfc * (flip_a ? -vc_raw : vc_raw);
            const float H_west = 0.5f * (S.Hi[rr][cc] + S.Hi[rr + 1][cc]);
            const float H_east = 0.5f * (S.Hi[rr][cc + 1] + S.Hi[rr + 1][cc + 1]);
            const float t = div_r(Kx + dxfv, K.two_g, K.r_two_g);
            const float h_west = (ec + H_west) - t;
            const float h_east = (ec + H_east) + t;
            bool adj = false;
            if constexpr (TAPE) {
                S.CodeX[j][i] = (uint8_t)(minmod_sel((uc - ul) * K.theta, (ur - ul) * 0.5f, (ur - uc) * K.theta)
                                 | (minmod_sel((vc_raw - S.V[rr][cc - 1]) * K.theta,
                                               (S.V[rr][cc + 1] - S.V[rr][cc - 1]) * 0.5f,
                                               (S.V[rr][cc + 1] - vc_raw) * K.theta) << 2)
                                 | (minmod_sel(bw_, ce_, fw_) << 4)
                                 | ((!(h_east > 0.0f)) ? (2u << 6) : (!(h_west > 0.0f)) ? (1u << 6) : 0u));
            }
            if (!(h_west > 0.0f)) { Kx = (-dxfv) + K.two_g * (ec + H_west); adj = true; }
            if (!(h_east > 0.0f)) { Kx = (-dxfv) - K.two_g * (ec + H_east); adj = true; }
            if constexpr (ZERO) {
                // T = (K + dx f v)/2g with the face's wall flip; equals t
                // unless the adjustment fired or the flips differ
                float T = t;
                if (adj || flip_c != flip_a)
                    T = div_r(Kx + dfc * (flip_c ? -vc_raw : vc_raw), K.two_g, K.r_two_g);
                S.Kx[j][i] = T;
            } else {
                S.Kx[j][i] = Kx;
            }
        } else {
            const int k = it - kSx;
            const int j = k / BW, i = k - j * BW;
            const int rr = j + 1, cc = i + 2;
            const int R = by + rr;
            const float el = S.E[rr - 1][cc], ec = S.E[rr][cc], eu = S.E[rr + 1][cc];
            float ul = S.U[rr - 1][cc], uc = S.U[rr][cc], uu = S.U[rr + 1][cc];
            const float vl = S.V[rr - 1][cc], vc = S.V[rr][cc], vu = S.V[rr + 1][cc];
            S.Uy[j][i] = minmod_slope(ul, uc, uu, K.theta);
            S.Vy[j][i] = minmod_slope(vl, vc, vu, K.theta);
            const float uc_raw = uc;
            const bool flip_c = ghost_row_flip(R);
            if (ghost_row_flip(R - 1)) ul = -ul;
            if (flip_c) uc = -uc;
            if (ghost_row_flip(R + 1)) uu = -uu;
            const float fc = fcell(rr, cc);
            float fu_l, fu_c, fu_u;
            if constexpr (ZERO) {
                // east = (1, -0): (1*u + (-0)*v)*f == u*f up to a zero's sign
                fu_l = ul * fc; fu_c = uc * fc; fu_u = uu * fc;
            } else {
                const float n_x = nxcell(rr, cc), n_y = nycell(rr, cc);
                const float e_x = n_y, e_y = -n_x;
                fu_l = (e_x * ul + e_y * vl) * fcell(rr - 1, cc);
                fu_c = (e_x * uc + e_y * vc) * fc;
                fu_u = (e_x * uu + e_y * vu) * fcell(rr + 1, cc);
            }
            const float bw_ = K.theta_g * ((ec - el) + K.udy * (fu_c + fu_l));
            const float ce_ = K.half_g * ((eu - el) + K.udy * ((fu_u + 2.0f * fu_c) + fu_l));
            const float fw_ = K.theta_g * ((eu - ec) + K.udy * (fu_c + fu_u));
            float Ly = minmod_raw(bw_, ce_, fw_);
            const bool flip_a = wS && R < 2;
            const float dfc = K.dy * fc;
            const float dyfu = dfc * (flip_a ? -uc_raw : uc_raw);
            const float H_south = 0.5f * (S.Hi[rr][cc] + S.Hi[rr][cc + 1]);
            const float H_north = 0.5f * (S.Hi[rr + 1][cc] + S.Hi[rr + 1][cc + 1]);
            const float t = div_r(Ly - dyfu, K.two_g, K.r_two_g);
            const float h_south = (ec + H_south) - t;
            const float h_north = (ec + H_north) + t;
            bool adj = false;
            if constexpr (TAPE) {
                S.CodeY[j][i] = (uint8_t)(minmod_sel((uc_raw - S.U[rr - 1][cc]) * K.theta,
                                            (S.U[rr + 1][cc] - S.U[rr - 1][cc]) * 0.5f,
                                            (S.U[rr + 1][cc] - uc_raw) * K.theta)
                                 | (minmod_sel((vc - vl) * K.theta, (vu - vl) * 0.5f, (vu - vc) * K.theta) << 2)
                                 | (minmod_sel(bw_, ce_, fw_) << 4)
                                 | ((!(h_north > 0.0f)) ? (2u << 6) : (!(h_south > 0.0f)) ? (1u << 6) : 0u));
            }
            if (!(h_south > 0.0f)) { Ly = dyfu + K.two_g * (ec + H_south); adj = true; }
            if (!(h_north > 0.0f)) { Ly = dyfu - K.two_g * (ec + H_north); adj = true; }
            if constexpr (ZERO) {
                float Sv = t;
                if (adj || flip_c != flip_a)
                    Sv = div_r(Ly - dfc * (flip_c ? -uc_raw : uc_raw), K.two_g, K.r_two_g);
                S.Ly[j][i] = Sv;
            } else {
                S.Ly[j][i] = Ly;
            }
        }
    }
    __syncthreads();

    // ---- phase D: face fluxes, each face once (:246-370) ----
    constexpr int kFx = BH * (BW + 1);
    constexpr int kFy = (BH + 1) * BW;
    for (int it = tid; it < kFx + kFy; it += NT) {
        if (it < kFx) {
            const int j = it / (BW + 1), i = it - j * (BW + 1);
            const int rr = j + 2, cm = i + 1, cp = i + 2;   // tile columns of m and p
            const int Cp = bx + cp;
            const float eta_p = S.E[rr][cp], eta_m = S.E[rr][cm];
            float Fx = 0.0f, Fy = 0.0f, Fz = 0.0f, Fx2 = 0.0f, Fy2 = 0.0f, Fz2 = 0.0f;
            uint32_t fcode = 0u;
            if (!(eta_p == kDryFlag || eta_m == kDryFlag)) {
                const float up = S.U[rr][cp], um = S.U[rr][cm];
                float vp = S.V[rr][cp], vm = S.V[rr][cm];
                const float Rpx = up - 0.5f * S.Ux[j][i + 1];
                const float Rpy = vp - 0.5f * S.Vx[j][i + 1];
                const float Rmx = um + 0.5f * S.Ux[j][i];
                const float Rmy = vm + 0.5f * S.Vx[j][i];
                const float H_face = 0.5f * (S.Hi[rr][cp] + S.Hi[rr + 1][cp]);
                if constexpr (ZERO) {
                    const float hp = fmaxf(0.0f, (eta_p + H_face) - S.Kx[j][i + 1]);
                    const float hm = fmaxf(0.0f, (eta_m + H_face) + S.Kx[j][i]);
                    cu_flux<TAPE>(hm, Rmx, Rmy, hp, Rpx, Rpy, K, Fx, Fy, Fz, &fcode);
                } else {
                    const float Kp = S.Kx[j][i + 1], Km = S.Kx[j][i];
                    if (wW && Cp == 2) vm = -vm;
                    if (wE && Cp == nx + 2) vp = -vp;
                    const float dfp = K.dx * fcell(rr, cp), dfm = K.dx * fcell(rr, cm);
                    {
                        const float n_x = nxcell(rr, cm), n_y = nycell(rr, cm);
                        const float vpn = up * n_x + vp * n_y;
                        const float vmn = um * n_x + vm * n_y;
                        const float hp = fmaxf(0.0f, (eta_p + H_face) - div_r(Kp + dfp * vpn, K.two_g, K.r_two_g));
                        const float hm = fmaxf(0.0f, (eta_m + H_face) + div_r(Km + dfm * vmn, K.two_g, K.r_two_g));
                        cu_flux(hm, Rmx, Rmy, hp, Rpx, Rpy, K, Fx, Fy, Fz);
                    }
                    if constexpr (FIELD) {
                        const float n_x = nxcell(rr, cp), n_y = nycell(rr, cp);
                        const float vpn = up * n_x + vp * n_y;
                        const float vmn = um * n_x + vm * n_y;
                        const float hp = fmaxf(0.0f, (eta_p + H_face) - div_r(Kp + dfp * vpn, K.two_g, K.r_two_g));
                        const float hm = fmaxf(0.0f, (eta_m + H_face) + div_r(Km + dfm * vmn, K.two_g, K.r_two_g));
                        cu_flux(hm, Rmx, Rmy, hp, Rpx, Rpy, K, Fx2, Fy2, Fz2);
                    }
                }
            }
            S.F[0][j][i] = Fx; S.F[1][j][i] = Fy; S.F[2][j][i] = Fz;
            if constexpr (TAPE) S.CodeF[j][i] = (uint8_t)fcode;
            if constexpr (FIELD) { S.F2[0][j][i] = Fx2; S.F2[1][j][i] = Fy2; S.F2[2][j][i] = Fz2; }
        } else {
            const int k = it - kFx;
            const int j = k / BW, i = k - j * BW;
            const int rm = j + 1, rp = j + 2, cc = i + 2;
            const int Rp = by + rp;
            const float eta_p = S.E[rp][cc], eta_m = S.E[rm][cc];
            float Gx = 0.0f, Gy = 0.0f, Gz = 0.0f, Gx2 = 0.0f, Gy2 = 0.0f, Gz2 = 0.0f;
            uint32_t gcode = 0u;
            if (!(eta_p == kDryFlag || eta_m == kDryFlag)) {
                float up = S.U[rp][cc], um = S.U[rm][cc];
                const float vp = S.V[rp][cc], vm = S.V[rm][cc];
                const float Rpx = up - 0.5f * S.Uy[j + 1][i];
                const float Rpy = vp - 0.5f * S.Vy[j + 1][i];
                const float Rmx = um + 0.5f * S.Uy[j][i];
                const float Rmy = vm + 0.5f * S.Vy[j][i];
                const float H_face = 0.5f * (S.Hi[rp][cc] + S.Hi[rp][cc + 1]);
                if constexpr (ZERO) {
                    const float hp = fmaxf(0.0f, (eta_p + H_face) - S.Ly[j + 1][i]);
                    const float hm = fmaxf(0.0f, (eta_m + H_face) + S.Ly[j][i]);
                    float a, b, c;
                    cu_flux<TAPE>(hm, Rmy, Rmx, hp, Rpy, Rpx, K, a, b, c, &gcode);
                    Gx = a; Gy = c; Gz = b;
                } else {
                    const float Lp = S.Ly[j + 1][i], Lm = S.Ly[j][i];
                    if (wS && Rp == 2) um = -um;
                    if (wN && Rp == ny + 2) up = -up;
                    const float dfp = K.dy * fcell(rp, cc), dfm = K.dy * fcell(rm, cc);
                    {
                        const float e_x = nycell(rm, cc), e_y = -nxcell(rm, cc);
                        const float upe = up * e_x + vp * e_y;
                        const float ume = um * e_x + vm * e_y;
                        const float hp = fmaxf(0.0f, (eta_p + H_face) - div_r(Lp - dfp * upe, K.two_g, K.r_two_g));
                        const float hm = fmaxf(0.0f, (eta_m + H_face) + div_r(Lm - dfm * ume, K.two_g, K.r_two_g));
                        float a, b, c;
                        cu_flux(hm, Rmy, Rmx, hp, Rpy, Rpx, K, a, b, c);
                        Gx = a; Gy = c; Gz = b;
                    }
                    if constexpr (FIELD) {
                        const float e_x = nycell(rp, cc), e_y = -nxcell(rp, cc);
                        const float upe = up * e_x + vp * e_y;
                        const float ume = um * e_x + vm * e_y;
                        const float hp = fmaxf(0.0f, (eta_p + H_face) - div_r(Lp - dfp * upe, K.two_g, K.r_two_g));
                        const float hm = fmaxf(0.0f, (eta_m + H_face) + div_r(Lm - dfm * ume, K.two_g, K.r_two_g));
                        float a, b, c;
                        cu_flux(hm, Rmy, Rmx, hp, Rpy, Rpx, K, a, b, c);
                        Gx2 = a; Gy2 = c; Gz2 = b;
                    }
                }
            }
            S.G[0][j][i] = Gx; S.G[1][j][i] = Gy; S.G[2][j][i] = Gz;
            if constexpr (TAPE) S.CodeG[j][i] = (uint8_t)gcode;
            if constexpr (FIELD) { S.G2[0][j][i] = Gx2; S.G2[1][j][i] = Gy2; S.G2[2][j][i] = Gz2; }
        }
    }
    __syncthreads();

    // ---- branch record (recording variant only) ----
    // Each padded cell is written by one tile: interior cells by the tile
    // that owns them; the ghost columns / rows the reverse pass reads (u, v
    // of the two ghost cells, the depth term and code of the first one) by
    // the boundary tile next to them.
    if constexpr (TAPE) {
        const bool own_w = tile_x == 0, own_s = tile_y == 0;
        const bool own_e = nx + 1 >= bx + 2 && nx + 1 < bx + BW + 2;
        const bool own_n = ny + 1 >= by + 2 && ny + 1 < by + BH + 2;
        for (int it = tid; it < kConv; it += NT) {
            int rr, cc;
            if (it < kMid) {
                rr = 2 + it / (BW + 4);
                cc = it - (rr - 2) * (BW + 4);
            } else {
                const int k = it - kMid;
                const int band = k / BW;
                rr = band < 2 ? band : BH + band;
                cc = 2 + (k - band * BW);
            }
            const int R = by + rr, C = bx + cc;
            if (R > ny + 3 || C > nx + 3) continue;
            const bool ri = R >= 2 && R <= ny + 1, ci = C >= 2 && C <= nx + 1;
            bool own;
            if (ri && ci) own = rr >= 2 && rr < BH + 2 && cc >= 2 && cc < BW + 2;
            else if (ri) own = rr >= 2 && rr < BH + 2 && (C < 2 ? own_w : own_e);
            else if (ci) own = cc >= 2 && cc < BW + 2 && (R < 2 ? own_s : own_n);
            else own = false;
            if (!own) continue;
            const size_t o = (size_t)R * pitch + C;
            A.tape_u[o] = S.U[rr][cc];
            A.tape_v[o] = S.V[rr][cc];
            uint32_t code = 0u;
            // x slope centre (interior rows, tile columns 1..BW+2)
            if (ri && cc >= 1 && cc <= BW + 2 && C >= 1 && C <= nx + 2) {
                A.tape_t[o] = S.Kx[rr - 2][cc - 1];
                code |= (uint32_t)S.CodeX[rr - 2][cc - 1] << kCodeX;
                if (C <= nx + 1 && cc <= BW + 1) code |= (uint32_t)S.CodeF[rr - 2][cc - 1] << kCodeFx;
            }
            // y slope centre (interior columns, tile rows 1..BH+2)
            if (ci && rr >= 1 && rr <= BH + 2 && R >= 1 && R <= ny + 2) {
                A.tape_s[o] = S.Ly[rr - 1][cc - 2];
                code |= (uint32_t)S.CodeY[rr - 1][cc - 2] << kCodeY;
                if (R <= ny + 1 && rr <= BH + 1) code |= (uint32_t)S.CodeG[rr - 1][cc - 2] << kCodeFy;
            }
            A.tape_code[o] = code;
        }
    }

    // ---- phase E: sources, RK update, store (:802-1004) ----
    const float dt = A.dt_dev ? *A.dt_dev : A.dt;
    const bool friction = A.r > 0.0f;
    for (int it = tid; it < BW * BH; it += NT) {
        const int j = it / BW, i = it - j * BW;
        const int rr = j + 2, cc = i + 2;
        const int R = by + rr, C = bx + cc;
        if (R > ny + 1 || C > nx + 1) continue;
        const float Hm = 0.25f * (((S.Hi[rr][cc] + S.Hi[rr + 1][cc]) + S.Hi[rr][cc + 1]) + S.Hi[rr + 1][cc + 1]);
        const float ec = S.E[rr][cc], uc = S.U[rr][cc], vc = S.V[rr][cc];
        const float hreg = ec + Hm;
        float hu = 0.0f, hv = 0.0f;
        if (hreg > K.cutoff) {
            hu = uc * hreg;
            hv = vc * hreg;
        }
        // west face is F[i] (evaluated by this cell as "p"), east face F[i+1]
        float fex, fey, fez, fwx, fwy, fwz, gnx, gny, gnz, gsx, gsy, gsz;
        fex = S.F[0][j][i + 1]; fey = S.F[1][j][i + 1]; fez = S.F[2][j][i + 1];
        gnx = S.G[0][j + 1][i]; gny = S.G[1][j + 1][i]; gnz = S.G[2][j + 1][i];
        if constexpr (FIELD) {
            fwx = S.F2[0][j][i]; fwy = S.F2[1][j][i]; fwz = S.F2[2][j][i];
            gsx = S.G2[0][j][i]; gsy = S.G2[1][j][i]; gsz = S.G2[2][j][i];
        } else {
            fwx = S.F[0][j][i]; fwy = S.F[1][j][i]; fwz = S.F[2][j][i];
            gsx = S.G[0][j][i]; gsy = S.G[1][j][i]; gsz = S.G[2][j][i];
        }
        const float fd0 = div_r(fex - fwx, K.dx, K.r_dx) + div_r(gnx - gsx, K.dy, K.r_dy);
        const float fd1 = div_r(fey - fwy, K.dx, K.r_dx) + div_r(gny - gsy, K.dy, K.r_dy);
        const float fd2 = div_r(fez - fwz, K.dx, K.r_dx) + div_r(gnz - gsz, K.dy, K.r_dy);

        const float fc = fcell(rr, cc);
        float tx, ty;
        if constexpr (ZERO) {
            tx = S.Kx[j][i + 1];
            ty = S.Ly[j + 1][i];
        } else {
            tx = div_r(S.Kx[j][i + 1] + (K.dx * fc) * vc, K.two_g, K.r_two_g);
            ty = div_r(S.Ly[j + 1][i] - (K.dy * fc) * uc, K.two_g, K.r_two_g);
        }
        const float eta_west = ec - tx, eta_east = ec + tx;
        const float eta_south = ec - ty, eta_north = ec + ty;

        float st1 = 0.0f, st2 = 0.0f;
        if (hreg >= K.cutoff && ec != kDryFlag) {
            float X = A.wind_x, Y = A.wind_y;
            // the recording variant serves the adjoint, which takes uniform wind
            if (!TAPE && A.wind_grid)
                wind_grid_at(A.wind_grid, A.wind_wx, A.wind_wy, A.wind_t, wind_axis(C, nx, A.wind_wx),
                             wind_axis(R, ny, A.wind_wy), X, Y);
            const float RHxp = 0.5f * (S.Hi[rr][cc + 1] + S.Hi[rr + 1][cc + 1]);
            const float RHxm = 0.5f * (S.Hi[rr][cc] + S.Hi[rr + 1][cc]);
            const float RHyp = 0.5f * (S.Hi[rr + 1][cc] + S.Hi[rr + 1][cc + 1]);
            const float RHym = 0.5f * (S.Hi[rr][cc] + S.Hi[rr][cc + 1]);
            const float H_x = RHxp - RHxm;
            const float H_y = RHyp - RHym;
            const float eta_sn = 0.5f * (eta_north + eta_south);
            const float eta_we = 0.5f * (eta_west + eta_east);
            const float b1 = (K.g * (eta_we + Hm)) * H_x;
            const float b2 = (K.g * (eta_sn + Hm)) * H_y;
            float hu_cor, hv_cor;
            if constexpr (ZERO) {
                // north (0,1): the rotations reduce to f*hv and (-f)*hu
                hu_cor = fc * hv;
                hv_cor = (-fc) * hu;
            } else {
                const float n_x = nxcell(rr, cc), n_y = nycell(rr, cc);
                const float e_x = n_y, e_y = -n_x;
                const float hu_east = hu * e_x + hv * e_y;
                const float hv_north = hu * n_x + hv * n_y;
                const float hu_east_cor = fc * hv_north;
                const float hv_north_cor = (-fc) * hu_east;
                const float up_x = -n_x, up_y = n_y;
                const float right_x = up_y, right_y = -up_x;
                hu_cor = right_x * hu_east_cor + right_y * hv_north_cor;
                hv_cor = up_x * hu_east_cor + up_y * hv_north_cor;
            }
            st1 = (X + hu_cor) + div_r(b1, K.dx, K.r_dx);
            st2 = (Y + hv_cor) + div_r(b2, K.dy, K.r_dy);
        }
        const float L1 = -fd0;
        const float L2 = -fd1 + st1;
        const float L3 = -fd2 + st2;

        float Cf = 0.0f;
        if (friction) {   // quadratic bottom friction (:884-900)
            const float h = hreg;
            if (h < K.desing) {
                const float u = desingularize(h, hu, K.desing);
                const float v = desingularize(h, hv, K.desing);
                Cf = (dt * A.r) * desingularize(h, sqrtf(u * u + v * v), K.desing);
            } else {
                const float u = hu / h, v = hv / h;
                Cf = ((dt * A.r) * sqrtf(u * u + v * v)) / h;
            }
        }

        const size_t o = (size_t)R * pitch + C;
        float ue, uu, uv;
        if (A.rk_order() < 3) {
            if (A.rk_step() == 0) {
                ue = ec + dt * L1;
                uu = hu + dt * L2;
                uv = hv + dt * L3;
                if (friction) {   // x / (1 + 0) == x, so r == 0 skips it
                    uu = uu / (1.0f + Cf);
                    uv = uv / (1.0f + Cf);
                }
            } else {
                const float ea = A.eta_qn[o], ha = A.hu_qn[o], va = A.hv_qn[o];
                ue = 0.5f * (ea + (ec + dt * L1));
                uu = 0.5f * (ha + (hu + dt * L2));
                uv = 0.5f * (va + (hv + dt * L3));
                if (friction) {
                    uu = uu / (1.0f + 0.5f * Cf);
                    uv = uv / (1.0f + 0.5f * Cf);
                }
            }
        } else {
            if (A.rk_step() == 0) {
                ue = ec + dt * L1;
                uu = hu + dt * L2;
                uv = hv + dt * L3;
            } else if (A.rk_step() == 1) {
                const float ea = A.eta_qn[o], ha = A.hu_qn[o], va = A.hv_qn[o];
                ue = 0.75f * ea + 0.25f * (ec + dt * L1);
                uu = 0.75f * ha + 0.25f * (hu + dt * L2);
                uv = 0.75f * va + 0.25f * (hv + dt * L3);
            } else {
                const float ea = A.eta_qn[o], ha = A.hu_qn[o], va = A.hv_qn[o];
                ue = (ea + 2.0f * (ec + dt * L1)) / 3.0f;
                uu = (ha + 2.0f * (hu + dt * L2)) / 3.0f;
                uv = (va + 2.0f * (hv + dt * L3)) / 3.0f;
            }
        }
        if (ue + Hm <= K.cutoff) {
            ue = -Hm + K.cutoff;
            uu = 0.0f;
            uv = 0.0f;
        }
        ue = fmaxf(-Hm + K.cutoff, ue);
        A.eta_out[o] = ue;
        A.hu_out[o] = uu;
        A.hv_out[o] = uv;
        // keep the periodic ghost cells of the output consistent, as the
        // reference's boundary kernels would after this substep; only cells
        // within two rows / columns of a periodic side have ghost images
        if ((pNS && (R <= 3 || R >= ny)) || (pEW && (C <= 3 || C >= nx))) {
            // ghost rows / columns this cell is the periodic image of:
            // ghost index g in {0,1} copies g+n, g in {n+2,n+3} copies g-n
            int rows[3] = {R, -1, -1}, cols[3] = {C, -1, -1};
            int nr = 1, nc = 1;
            if (pNS) {
                if (R + ny <= ny + 3) rows[nr++] = R + ny;
                if (R - ny >= 0) rows[nr++] = R - ny;
            }
            if (pEW) {
                if (C + nx <= nx + 3) cols[nc++] = C + nx;
                if (C - nx >= 0) cols[nc++] = C - nx;
            }
            for (int a = 0; a < nr; ++a) {
                for (int b = 0; b < nc; ++b) {
                    if (a == 0 && b == 0) continue;
                    const size_t q = (size_t)rows[a] * pitch + cols[b];
                    A.eta_out[q] = ue; A.hu_out[q] = uu; A.hv_out[q] = uv;
                }
            }
        }
    }
}

}  // namespace gocean
