// adjoint_tile.hpp -- device code of the CDKLM16 substep adjoint (one output
// tile per workgroup) and of the ghost-cell fold; shared by the single-domain
// entry points (gocean_adj.hip) and the batched ensemble (gocean_ens.hip).
//
// The forward substep is  out = P(Q) = base(Q) + dt*L(Q)  (rk_step 0) and
// out = 0.5*(Q^n + P(Q*))  (rk_step 1), then the dry clamp, where L is the
// flux divergence plus sources of
// gpu_ocean/SWESimulators/gpu_kernels/CDKLM16_kernel.cu:429-1004 and
// base(Q) = (E, u*h_reg, v*h_reg) the converted state the update starts from
// (:597-602, :905-920).  A tile computes  o = alpha * J_P(Q)^T w + gamma * x
// for its output cells, with every branch fixed to the forward's choice as
// stored in the branch record the recording forward variant writes:
//   * cell data of the tile and a 3-cell halo go to LDS;
//   * every face is differentiated once (flux VJP), the slope / depth-term
//     VJPs are applied once per slope centre, and each output cell gathers
//     its contributions in a fixed order (deterministic result);
//   * ghost cells get their own adjoint entries, which the fold then adds to
//     the interior cells they are images of (wall mirror with the normal
//     momentum negated, periodic wrap).
// Branch-deciding quantities are read from the record or recomputed with the
// forward's exact operations; values that only scale derivatives use fast
// reciprocal / square root.
#pragma once

#include <hip/hip_runtime.h>

#include "cdklm16_tile.hpp"

namespace gocean {

struct AdjTileArgs {
    const float* eta;                     // substep input eta (images read for ghosts)
    const float* tU; const float* tV; const float* tT; const float* tS;
    const uint32_t* tC;                   // branch record of the substep
    const float* w_eta; const float* w_hu; const float* w_hv; const float* w_clamp;
    const float* x_eta; const float* x_hu; const float* x_hv; const float* x_clamp;
    float* o_eta; float* o_hu; float* o_hv;
    const float* Hi;
    const float* dt_dev;
    float dt, alpha, gamma;
    // w / x hold the ghost-cell entries of an adjoint output whose fold has
    // not run: the fold's sums are formed while reading (same order, so the
    // same bits as folding first)
    int w_fold, x_fold;
    int nx, ny, pitch, hi_pitch;
    float dx, dy, g, theta, desing, flux_eps, cutoff, land, f;
    int bc_n, bc_e, bc_s, bc_w;
    int tiles_x, tiles_y;
    // derived on the host (derivative values only, never branch decisions):
    // 1/dx, 1/dy, 1/(2g), dx/(2g), dy/(2g), dx*f, dy*f, g/dx, g/dy
    float r_dx, r_dy, r_two_g, Dx, Dy, Ddx, Ddy, g_dx, g_dy;
};

// fill the derived constants of AdjTileArgs from its base fields
__host__ __device__ inline void adj_derive(AdjTileArgs& A) {
    A.r_dx = 1.0f / A.dx;
    A.r_dy = 1.0f / A.dy;
    A.r_two_g = 1.0f / (2.0f * A.g);
    A.Dx = A.dx / (2.0f * A.g);
    A.Dy = A.dy / (2.0f * A.g);
    A.Ddx = A.dx * A.f;
    A.Ddy = A.dy * A.f;
    A.g_dx = A.g / A.dx;
    A.g_dy = A.g / A.dy;
}

__host__ __device__ inline float adj_rcp(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_rcpf(x);
#else
    return 1.0f / x;
#endif
}

__host__ __device__ inline float adj_sqrt(float x) {
#ifdef __HIP_DEVICE_COMPILE__
    return __builtin_amdgcn_sqrtf(x);
#else
    return sqrtf(x);
#endif
}

// value of the minmod candidate `sel` (the forward's minmod_slope result,
// bit for bit) and its derivative pattern
__host__ __device__ inline float slope_val(uint32_t sel, float l, float c, float r, float theta) {
    // all three candidates and a select: no divergent branches in a wave
    // whose cells chose different candidates
    const float a = (c - l) * theta, b = (r - l) * 0.5f, d = (r - c) * theta;
    return sel == 0u ? a : (sel == 1u ? b : (sel == 2u ? d : 0.0f));
}

// d(candidate sel)/d(l, c, r): (c-l)*theta, (r-l)/2, (r-c)*theta, or 0
__host__ __device__ inline void slope_coef(uint32_t sel, float theta, float& kl, float& kc, float& kr) {
    kl = sel == 0u ? -theta : (sel == 1u ? -0.5f : 0.0f);
    kc = sel == 0u ? theta : (sel == 2u ? -theta : 0.0f);
    kr = sel == 1u ? 0.5f : (sel == 2u ? theta : 0.0f);
}

// VJP of the face depth term T = (K - sig*Dd*s_c*t_c)/2g of one slope centre
// (sig = -1 for x faces, +1 for y faces; t = tangential velocity, v for x and
// u for y; s_* the wall ghost sign flips, s_a the adjustment's own flip) with
// K the minmod candidate `ks` of
//   theta g ((E_c - E_l) + sig D f (s_c t_c + s_l t_l)),
//   g/2 ((E_r - E_l) + sig D f ((s_r t_r + 2 s_c t_c) + s_l t_l)),
//   theta g ((E_r - E_c) + sig D f (s_c t_c + s_r t_r)),
// or the positivity adjustment sig Dd s_a t_c +- 2g (E_c + H).  Adds
// d/d(E, t) of cells l, c, r times gT.
// NOADJ: the caller knows adj == 0 (the same operations without the branch)
template <bool NOADJ = false>
__host__ __device__ inline void depth_term_vjp(uint32_t ks, uint32_t adj, float sig, float theta, float g,
                                               float f, float Dd, float D, float r_two_g, float sl, float sc,
                                               float sr, float sa, float gT, float* gE, float* gt) {
    const float two_g = 2.0f * g;
    const float gK = gT * r_two_g;
    float el, ec, er, tl, tc, tr;
    if (!NOADJ && adj != 0u) {
        el = 0.0f; er = 0.0f; ec = adj == 1u ? two_g : -two_g;
        tl = 0.0f; tr = 0.0f; tc = sig * Dd * sa;
    } else {
        const float q = ks == 1u ? 0.5f * g : (ks == 3u ? 0.0f : theta * g);
        // E pattern of the candidate: (c - l), (r - l), (r - c)
        el = (ks == 0u || ks == 1u) ? -q : 0.0f;
        ec = ks == 0u ? q : (ks == 2u ? -q : 0.0f);
        er = (ks == 1u || ks == 2u) ? q : 0.0f;
        const float m = q * sig * D * f;
        // f*t pattern: (l, c) for 0, (l, 2c, r) for 1, (c, r) for 2
        tl = (ks == 0u || ks == 1u) ? m * sl : 0.0f;
        tc = ks == 1u ? 2.0f * m * sc : (ks == 3u ? 0.0f : m * sc);
        tr = (ks == 1u || ks == 2u) ? m * sr : 0.0f;
    }
    tc += -sig * Dd * sc;                       // T's own tangential term
    gE[0] += el * gK; gE[1] += ec * gK; gE[2] += er * gK;
    gt[0] += tl * gK; gt[1] += tc * gK; gt[2] += tr * gK;
}

// VJP of the central-upwind flux in (depth, normal velocity, tangential
// velocity) form: mass = F0, normal momentum = F1, tangential momentum = F2
// (CDKLM16_flux, CDKLM16_kernel.cu:101-140), branch choices from `code`.
// WET: the caller knows the flux is not zeroed (code bit 4 clear) and both
// depths exceed the cutoff (the same operations without the branches)
template <bool WET = false>
__host__ __device__ inline void flux_vjp(float hm, float nm, float tm, float hp, float np_, float tp,
                                         uint32_t code, float g, float cutoff,
                                         float l0, float l1, float l2,
                                         float& ghm, float& gnm, float& gtm,
                                         float& ghp, float& gnp, float& gtp) {
    ghm = gnm = gtm = ghp = gnp = gtp = 0.0f;
    if (!WET && (code & 16u)) return;             // |a+ - a-| < flux_eps: zero flux
    const bool wp = WET || hp > cutoff, wm = WET || hm > cutoff;
    float Fpx = 0.0f, Fpy = 0.0f, c_p = 0.0f, u_p = 0.0f;
    if (wp) { Fpx = hp * np_; Fpy = hp * np_ * np_ + 0.5f * g * hp * hp; u_p = np_; c_p = adj_sqrt(g * hp); }
    float Fmx = 0.0f, Fmy = 0.0f, c_m = 0.0f, u_m = 0.0f;
    if (wm) { Fmx = hm * nm; Fmy = hm * nm * nm + 0.5f * g * hm * hm; u_m = nm; c_m = adj_sqrt(g * hm); }
    const uint32_t cap = code & 3u, cam = (code >> 2) & 3u;
    const float ap = cap == 1u ? u_m + c_m : (cap == 2u ? u_p + c_p : 0.0f);
    const float am = cam == 1u ? u_m - c_m : (cam == 2u ? u_p - c_p : 0.0f);
    const float rden = adj_rcp(ap - am);
    const float apam = ap * am;
    const float F0 = ((ap * Fmx - am * Fpx) + apam * (hp - hm)) * rden;
    const float F1 = ((ap * Fmy - am * Fpy) + apam * (Fpx - Fmx)) * rden;
    const float g0 = l0 * rden, g1 = l1 * rden;
    const float gden = -(l0 * F0 + l1 * F1) * rden;
    const float gap = g0 * (Fmx + am * (hp - hm)) + g1 * (Fmy + am * (Fpx - Fmx)) + gden;
    const float gam = g0 * (-Fpx + ap * (hp - hm)) + g1 * (-Fpy + ap * (Fpx - Fmx)) - gden;
    const float gFmx = g0 * ap - g1 * apam;
    const float gFpx = -g0 * am + g1 * apam;
    const float gFmy = g1 * ap;
    const float gFpy = -g1 * am;
    ghp += g0 * apam;
    ghm -= g0 * apam;
    // tangential momentum upwinded on the sign of nm + np (recomputed exactly)
    const float gFmz = (nm + np_ > 0.0f) ? l2 : 0.0f;
    const float gFpz = (nm + np_ > 0.0f) ? 0.0f : l2;
    float gum = 0.0f, gcm = 0.0f, gup = 0.0f, gcp = 0.0f;
    if (cap == 1u) { gum += gap; gcm += gap; } else if (cap == 2u) { gup += gap; gcp += gap; }
    if (cam == 1u) { gum += gam; gcm -= gam; } else if (cam == 2u) { gup += gam; gcp -= gam; }
    if (wp) {
        ghp += gFpx * np_ + gFpy * (np_ * np_ + g * hp) + gFpz * np_ * tp + gcp * g * 0.5f * adj_rcp(c_p);
        gnp += gFpx * hp + gFpy * 2.0f * hp * np_ + gFpz * hp * tp + gup;
        gtp += gFpz * hp * np_;
    }
    if (wm) {
        ghm += gFmx * nm + gFmy * (nm * nm + g * hm) + gFmz * nm * tm + gcm * g * 0.5f * adj_rcp(c_m);
        gnm += gFmx * hm + gFmy * 2.0f * hm * nm + gFmz * hm * tm + gum;
        gtm += gFmz * hm * nm;
    }
}

// ---------------------------------------------------------------------------
// LDS of one adjoint tile.  Cell (R, C) sits at [R - by + 1][C - bx + 1]
// (a 3-cell halo around the BW x BH block of interior cells).
// ---------------------------------------------------------------------------
template <int BW, int BH>
struct AdjSmem {
    static constexpr int W = BW + 6, H = BH + 6;
    static constexpr int NF = (BH * W > H * BW) ? BH * W : H * BW;
    float Hi[H + 1][W + 1];
    float E[H][W], U[H][W], V[H][W];
    uint32_t code[H][W];                 // record word; bits 28-29: conversion branch
    // lambda at interior cells (clamp-masked) and the gathered d/d(E, u, v)
    // of output cells, only read inside the outer ring: cell (lr, lc) at
    // [lr - 1][lc - 1]
    float W0[H - 2][W - 2], W1[H - 2][W - 2], W2[H - 2][W - 2];
    float A0[H - 2][W - 2], A1[H - 2][W - 2], A2[H - 2][W - 2];
    float FD[6][NF];                     // face -> (E, u, v) of its minus and plus cell
    // face -> (T, u slope, v slope) of minus / plus in rows 0-5 until the
    // slope-centre stage has read them, then slope centre -> (E, u, v) of
    // cells -1, 0, +1 in rows 0-8 (written after a barrier)
    float FC[9][NF];
};

enum { kBrLand = 0, kBrDry = 1, kBrDesing = 2, kBrWet = 3 };

template <int BW, int BH>
struct AdjTile {
    using Smem = AdjSmem<BW, BH>;
    static constexpr int W = Smem::W, H = Smem::H;
    const AdjTileArgs& A;
    Smem& S;
    int bx, by;                  // padded index of the block's first interior cell is (by+2, bx+2)
    int C0, C1, R0, R1;          // output cell range (ghost cells included)
    int fx0, fx1, fy0, fy1;      // x faces (minus cell column) / y faces (minus cell row) computed
    float dt;

    __host__ __device__ AdjTile(const AdjTileArgs& a, Smem& s, int tx, int ty) : A(a), S(s) {
        bx = tx * BW;
        by = ty * BH;
        const bool own_e = A.nx + 1 >= bx + 2 && A.nx + 1 < bx + BW + 2;
        const bool own_n = A.ny + 1 >= by + 2 && A.ny + 1 < by + BH + 2;
        C0 = tx == 0 ? 0 : bx + 2;
        C1 = own_e ? A.nx + 3 : bx + BW + 1;
        R0 = ty == 0 ? 0 : by + 2;
        R1 = own_n ? A.ny + 3 : by + BH + 1;
        fx0 = C0 - 2 > 1 ? C0 - 2 : 1;
        fx1 = C1 + 1 < A.nx + 1 ? C1 + 1 : A.nx + 1;
        fy0 = R0 - 2 > 1 ? R0 - 2 : 1;
        fy1 = R1 + 1 < A.ny + 1 ? R1 + 1 : A.ny + 1;
        dt = A.dt_dev ? *A.dt_dev : A.dt;
    }

    __host__ __device__ int Rof(int lr) const { return by - 1 + lr; }
    __host__ __device__ int Cof(int lc) const { return bx - 1 + lc; }
    __host__ __device__ bool ghost_r(int R) const { return R < 2 || R > A.ny + 1; }
    __host__ __device__ bool ghost_c(int C) const { return C < 2 || C > A.nx + 1; }
    __host__ __device__ bool is_output(int R, int C) const {
        return R >= R0 && R <= R1 && C >= C0 && C <= C1 && !(ghost_r(R) && ghost_c(C));
    }
    __host__ __device__ float hflag(float v) const {
        return (fabsf(v - A.land) < kDryEps) ? kDryFlag : v;
    }
    // tangential-velocity sign of a wall ghost column / row in the K / L terms
    __host__ __device__ float col_flip(int C) const {
        return ((A.bc_w == kWall && C < 2) || (A.bc_e == kWall && C > A.nx + 1)) ? -1.0f : 1.0f;
    }
    __host__ __device__ float row_flip(int R) const {
        return ((A.bc_s == kWall && R < 2) || (A.bc_n == kWall && R > A.ny + 1)) ? -1.0f : 1.0f;
    }
    // the forward clamped the cell dry: its output eta (v) sits at the cutoff
    __host__ __device__ bool clamped_v(bool has, float v, int lr, int lc) const {
        if (!has) return false;
        const float Hreg = 0.25f * (((S.Hi[lr][lc] + S.Hi[lr + 1][lc]) + S.Hi[lr][lc + 1]) + S.Hi[lr + 1][lc + 1]);
        return v <= -Hreg + A.cutoff;
    }

    // a[R, C] of an interior cell plus the ghost entries whose image it is,
    // summed as fold_cell sums them (field 1 = hu, 2 = hv: the normal
    // momentum of a wall image is negated)
    __host__ __device__ float fold_read(const float* a, int field, int R, int C) const {
        const int nx = A.nx, ny = A.ny;
        const size_t pitch = (size_t)A.pitch;
        float s = a[R * pitch + C];
        if (R <= 3) {
            if (A.bc_s == kWall) { const float g = a[(3 - R) * pitch + C]; s += field == 2 ? -g : g; }
            else s += a[(R + ny) * pitch + C];
        }
        if (R >= ny) {
            if (A.bc_n == kWall) { const float g = a[(2 * ny + 3 - R) * pitch + C]; s += field == 2 ? -g : g; }
            else s += a[(R - ny) * pitch + C];
        }
        if (C <= 3) {
            if (A.bc_w == kWall) { const float g = a[R * pitch + 3 - C]; s += field == 1 ? -g : g; }
            else s += a[R * pitch + C + nx];
        }
        if (C >= nx) {
            if (A.bc_e == kWall) { const float g = a[R * pitch + 2 * nx + 3 - C]; s += field == 1 ? -g : g; }
            else s += a[R * pitch + C - nx];
        }
        return s;
    }

    // cell depth from the four corner depths in LDS, land corners left out
    // (the forward's operations, so h = eta + Hm is the forward's h)
    __host__ __device__ float cell_hm(int lr, int lc) const {
        const float a = S.Hi[lr][lc], b = S.Hi[lr + 1][lc], c = S.Hi[lr][lc + 1], d = S.Hi[lr + 1][lc + 1];
        const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag) - (float)(c == kDryFlag)
                          - (float)(d == kDryFlag);
        const float sum = ((a + b) + c) + d;
        return (wet == 4.0f) ? sum * 0.25f : (wet == 0.0f ? A.land : sum / wet);
    }

    // ---- stage 0: corner depths, land flagged
    __host__ __device__ void stage_hi(int it) {
        const int lr = it / (W + 1), lc = it - lr * (W + 1);
        const int R = Rof(lr), C = Cof(lc);
        float v = 0.0f;
        if (R >= 0 && R <= A.ny + 4 && C >= 0 && C <= A.nx + 4) v = hflag(A.Hi[(size_t)R * A.hi_pitch + C]);
        S.Hi[lr][lc] = v;
    }
    static constexpr int kHiItems = (H + 1) * (W + 1);

    // ---- stage 1: cell data
    // The stage's global loads are issued by load_cell before stage 0, so
    // their latency overlaps the corner-depth stage; stage_cells then needs
    // only LDS and these registers.
    struct CellIn {
        float eta, U, V, wc, w0, w1, w2;
        uint32_t tc;
    };
    __host__ __device__ bool cell_used(int lr, int lc, int R, int C) const {
        const bool in_band = (lr >= 3 && lr < BH + 3) || (lc >= 3 && lc < BW + 3);
        return in_band && R >= 0 && R <= A.ny + 3 && C >= 0 && C <= A.nx + 3 && !(ghost_r(R) && ghost_c(C));
    }
    // interior cells with a ghost image (the ones fold_read adds to)
    __host__ __device__ bool near_boundary(int R, int C) const {
        return R <= 3 || R >= A.ny || C <= 3 || C >= A.nx;
    }
    __host__ __device__ void load_cell(int it, CellIn& in) const {
        const int lr = it / W, lc = it - lr * W;
        const int R = Rof(lr), C = Cof(lc);
        const int nx = A.nx, ny = A.ny;
        in.eta = in.U = in.V = in.wc = in.w0 = in.w1 = in.w2 = 0.0f;
        in.tc = 0u;
        if (!cell_used(lr, lc, R, C)) return;
        // eta of the image cell (wall mirror / periodic wrap)
        int sr = R, sc = C;
        if (R < 2) sr = (A.bc_s == kWall) ? 3 - R : R + ny;
        else if (R > ny + 1) sr = (A.bc_n == kWall) ? 2 * ny + 3 - R : R - ny;
        if (C < 2) sc = (A.bc_w == kWall) ? 3 - C : C + nx;
        else if (C > nx + 1) sc = (A.bc_e == kWall) ? 2 * nx + 3 - C : C - nx;
        in.eta = A.eta[(size_t)sr * A.pitch + sc];
        const size_t o = (size_t)R * A.pitch + C;
        in.U = A.tU[o];
        in.V = A.tV[o];
        const bool ri = !ghost_r(R), ci = !ghost_c(C);
        if ((ri && C >= 1 && C <= nx + 2) || (ci && R >= 1 && R <= ny + 2)) in.tc = A.tC[o];
        if (ri && ci) {
            if (A.w_clamp) in.wc = A.w_clamp[o];
            in.w0 = A.w_eta[o]; in.w1 = A.w_hu[o]; in.w2 = A.w_hv[o];
        }
    }
    __host__ __device__ void stage_cells(int it, const CellIn& in) {
        const int lr = it / W, lc = it - lr * W;
        const int R = Rof(lr), C = Cof(lc);
        const int nx = A.nx, ny = A.ny;
        float E = 0.0f, U = 0.0f, V = 0.0f;
        float w0 = 0.0f, w1 = 0.0f, w2 = 0.0f;
        uint32_t code = 0u;
        if (cell_used(lr, lc, R, C)) {
            const float eta = in.eta;
            const float Hm = cell_hm(lr, lc);
            const float h = eta + Hm;
            uint32_t br;
            if (fabsf(Hm - A.land) <= kDryEps) { E = kDryFlag; br = kBrLand; }
            else if (h < A.desing) {
                if (h <= A.cutoff) { E = -Hm + A.cutoff; br = kBrDry; }
                else { E = eta; br = kBrDesing; }
            } else { E = eta; br = kBrWet; }
            U = in.U;
            V = in.V;
            const bool ri = !ghost_r(R), ci = !ghost_c(C);
            code = br << 28;
            if (ri && C >= 1 && C <= nx + 2) code |= in.tc & 0x001F00FFu;
            if (ci && R >= 1 && R <= ny + 2) code |= in.tc & 0x03E0FF00u;
            if (ri && ci) {
                if (!clamped_v(A.w_clamp != nullptr, in.wc, lr, lc)) {
                    if (A.w_fold && near_boundary(R, C)) {
                        w0 = fold_read(A.w_eta, 0, R, C); w1 = fold_read(A.w_hu, 1, R, C);
                        w2 = fold_read(A.w_hv, 2, R, C);
                    } else { w0 = in.w0; w1 = in.w1; w2 = in.w2; }
                }
            }
        }
        S.E[lr][lc] = E; S.U[lr][lc] = U; S.V[lr][lc] = V; S.code[lr][lc] = code;
        if (lr >= 1 && lr + 1 < H && lc >= 1 && lc + 1 < W) {
            S.W0[lr - 1][lc - 1] = w0; S.W1[lr - 1][lc - 1] = w1; S.W2[lr - 1][lc - 1] = w2;
            S.A0[lr - 1][lc - 1] = 0.0f; S.A1[lr - 1][lc - 1] = 0.0f; S.A2[lr - 1][lc - 1] = 0.0f;
        }
    }
    static constexpr int kCellItems = H * W;

    // ---- x pass --------------------------------------------------------
    // faces: row j (lr = j + 3), minus cell lc = fi; index j * W + fi
    static constexpr int kXItems = BH * W;
    __host__ __device__ bool x_row_ok(int j) const { return by + 2 + j <= A.ny + 1; }

    __host__ __device__ bool xface_used(int j, int fi) const {
        const int cm = Cof(fi);
        return x_row_ok(j) && cm >= fx0 && cm <= fx1 && fi >= 1 && fi + 2 < W;
    }
    // the face's depth terms from the record (rows 2..ny+1, columns
    // 1..nx+2), loaded before stage 0
    __host__ __device__ void load_xface(int it, float& tm, float& tp) const {
        const int j = it / W, fi = it - j * W;
        tm = tp = 0.0f;
        if (!xface_used(j, fi)) return;
        const size_t om = (size_t)Rof(j + 3) * A.pitch + Cof(fi);
        tm = A.tT[om];
        tp = A.tT[om + 1];
    }
    __host__ __device__ void stage_xface(int it, float tm, float tp) {
        const int j = it / W, fi = it - j * W;
        const int lr = j + 3;
        float out[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (xface_used(j, fi)) {
            const int m = fi, p = fi + 1;
            const float Em = S.E[lr][m], Ep = S.E[lr][p];
            if (!(Em == kDryFlag || Ep == kDryFlag)) {
                const uint32_t km = S.code[lr][m], kp = S.code[lr][p];
                const float rdx = dt * A.r_dx;
                const float l0 = (S.W0[lr - 1][p - 1] - S.W0[lr - 1][m - 1]) * rdx;
                const float l1 = (S.W1[lr - 1][p - 1] - S.W1[lr - 1][m - 1]) * rdx;
                const float l2 = (S.W2[lr - 1][p - 1] - S.W2[lr - 1][m - 1]) * rdx;
                const float th = A.theta;
                const float Sum = slope_val(km & 3u, S.U[lr][m - 1], S.U[lr][m], S.U[lr][m + 1], th);
                const float Svm = slope_val((km >> 2) & 3u, S.V[lr][m - 1], S.V[lr][m], S.V[lr][m + 1], th);
                const float Sup = slope_val(kp & 3u, S.U[lr][p - 1], S.U[lr][p], S.U[lr][p + 1], th);
                const float Svp = slope_val((kp >> 2) & 3u, S.V[lr][p - 1], S.V[lr][p], S.V[lr][p + 1], th);
                const float Rpx = S.U[lr][p] - 0.5f * Sup, Rpy = S.V[lr][p] - 0.5f * Svp;
                const float Rmx = S.U[lr][m] + 0.5f * Sum, Rmy = S.V[lr][m] + 0.5f * Svm;
                const float Hf = 0.5f * (S.Hi[lr][p] + S.Hi[lr + 1][p]);
                const float hp_raw = (Ep + Hf) - tp;
                const float hm_raw = (Em + Hf) + tm;
                const float hp = fmaxf(0.0f, hp_raw), hm = fmaxf(0.0f, hm_raw);
                float ghm, gnm, gtm, ghp, gnp, gtp;
                flux_vjp(hm, Rmx, Rmy, hp, Rpx, Rpy, (km >> kCodeFx) & 31u, A.g, A.cutoff, l0, l1, l2,
                         ghm, gnm, gtm, ghp, gnp, gtp);
                const float ghm_r = hm_raw > 0.0f ? ghm : 0.0f;
                const float ghp_r = hp_raw > 0.0f ? ghp : 0.0f;
                out[0] = ghm_r; out[1] = gnm; out[2] = gtm;          // direct: E, u, v of minus cell
                out[3] = ghp_r; out[4] = gnp; out[5] = gtp;          // and of plus cell
                out[6] = ghm_r; out[7] = 0.5f * gnm; out[8] = 0.5f * gtm;     // T, u slope, v slope (minus)
                out[9] = -ghp_r; out[10] = -0.5f * gnp; out[11] = -0.5f * gtp; // (plus)
            }
        }
        const int k = j * W + fi;
        for (int q = 0; q < 6; ++q) { S.FD[q][k] = out[q]; S.FC[q][k] = out[6 + q]; }
    }

    // slope centre c (lc = fi, row lr): gather from its two faces, apply the
    // u / v slope VJPs and the depth-term VJP (K with minmod / adjustment)
    __host__ __device__ void stage_xcentre(int it, float (&out)[9]) const {
        const int j = it / W, fi = it - j * W;
        const int lr = j + 3, c = Cof(fi);
        for (int q = 0; q < 9; ++q) out[q] = 0.0f;
        if (x_row_ok(j) && c >= fx0 && c <= fx1 + 1 && fi >= 1 && fi + 1 < W) {
            const int kl = j * W + fi - 1, kr = j * W + fi;   // face to the left (c is plus), right (c is minus)
            const float gT = S.FC[3][kl] + S.FC[0][kr];
            const float gSu = S.FC[4][kl] + S.FC[1][kr];
            const float gSv = S.FC[5][kl] + S.FC[2][kr];
            const uint32_t code = S.code[lr][fi];
            const float th = A.theta;
            float cl, cc, cr;
            slope_coef(code & 3u, th, cl, cc, cr);
            out[1] = cl * gSu; out[4] = cc * gSu; out[7] = cr * gSu;
            slope_coef((code >> 2) & 3u, th, cl, cc, cr);
            out[2] = cl * gSv; out[5] = cc * gSv; out[8] = cr * gSv;
            // T = (K + dx f s_c v_c) / 2g: x direction, sigma = -1, t = v
            const int C = c;
            const float sa = (A.bc_w == kWall && C < 2) ? -1.0f : 1.0f;
            float gE[3] = {0.0f, 0.0f, 0.0f}, gt[3] = {0.0f, 0.0f, 0.0f};
            depth_term_vjp((code >> 4) & 3u, (code >> 6) & 3u, -1.0f, th, A.g, A.f, A.Ddx, A.Dx, A.r_two_g,
                           col_flip(C - 1), col_flip(C), col_flip(C + 1), sa, gT, gE, gt);
            out[0] += gE[0]; out[3] += gE[1]; out[6] += gE[2];
            out[2] += gt[0]; out[5] += gt[1]; out[8] += gt[2];
        }
    }
    // centre item it's results, stored once every centre has read FC
    __host__ __device__ void store_centre(int it, const float (&out)[9]) {
        for (int q = 0; q < 9; ++q) S.FC[q][it] = out[q];
    }

    __host__ __device__ void stage_xgather(int it) {
        const int j = it / W, fi = it - j * W;
        const int lr = j + 3;
        if (!x_row_ok(j) || fi < 1 || fi + 1 >= W) return;
        if (!is_output(Rof(lr), Cof(fi))) return;
        const int k = j * W + fi;
        // faces: left (cell is plus), right (cell is minus); centres -1, 0, +1
        S.A0[lr - 1][fi - 1] += ((S.FD[3][k - 1] + S.FD[0][k]) + (S.FC[6][k - 1] + S.FC[3][k])) + S.FC[0][k + 1];
        S.A1[lr - 1][fi - 1] += ((S.FD[4][k - 1] + S.FD[1][k]) + (S.FC[7][k - 1] + S.FC[4][k])) + S.FC[1][k + 1];
        S.A2[lr - 1][fi - 1] += ((S.FD[5][k - 1] + S.FD[2][k]) + (S.FC[8][k - 1] + S.FC[5][k])) + S.FC[2][k + 1];
    }

    // ---- y pass --------------------------------------------------------
    // faces: column i (lc = i + 3), minus cell lr = fj; index fj * BW + i
    static constexpr int kYItems = H * BW;
    __host__ __device__ bool y_col_ok(int i) const { return bx + 2 + i <= A.nx + 1; }

    __host__ __device__ bool yface_used(int fj, int i) const {
        const int rm = Rof(fj);
        return y_col_ok(i) && rm >= fy0 && rm <= fy1 && fj >= 1 && fj + 2 < H;
    }
    // the face's depth terms from the record (columns 2..nx+1, rows
    // 1..ny+2), loaded before stage 0
    __host__ __device__ void load_yface(int it, float& tm, float& tp) const {
        const int fj = it / BW, i = it - fj * BW;
        tm = tp = 0.0f;
        if (!yface_used(fj, i)) return;
        const size_t om = (size_t)Rof(fj) * A.pitch + Cof(i + 3);
        tm = A.tS[om];
        tp = A.tS[om + A.pitch];
    }
    __host__ __device__ void stage_yface(int it, float tm, float tp) {
        const int fj = it / BW, i = it - fj * BW;
        const int lc = i + 3;
        float out[12] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
        if (yface_used(fj, i)) {
            const int m = fj, p = fj + 1;
            const float Em = S.E[m][lc], Ep = S.E[p][lc];
            if (!(Em == kDryFlag || Ep == kDryFlag)) {
                const uint32_t km = S.code[m][lc], kp = S.code[p][lc];
                const float rdy = dt * A.r_dy;
                // flux order (mass, normal = v momentum, tangential = u momentum)
                const float l0 = (S.W0[p - 1][lc - 1] - S.W0[m - 1][lc - 1]) * rdy;
                const float l1 = (S.W2[p - 1][lc - 1] - S.W2[m - 1][lc - 1]) * rdy;
                const float l2 = (S.W1[p - 1][lc - 1] - S.W1[m - 1][lc - 1]) * rdy;
                const float th = A.theta;
                const float Sum = slope_val((km >> 8) & 3u, S.U[m - 1][lc], S.U[m][lc], S.U[m + 1][lc], th);
                const float Svm = slope_val((km >> 10) & 3u, S.V[m - 1][lc], S.V[m][lc], S.V[m + 1][lc], th);
                const float Sup = slope_val((kp >> 8) & 3u, S.U[p - 1][lc], S.U[p][lc], S.U[p + 1][lc], th);
                const float Svp = slope_val((kp >> 10) & 3u, S.V[p - 1][lc], S.V[p][lc], S.V[p + 1][lc], th);
                const float Rpx = S.U[p][lc] - 0.5f * Sup, Rpy = S.V[p][lc] - 0.5f * Svp;
                const float Rmx = S.U[m][lc] + 0.5f * Sum, Rmy = S.V[m][lc] + 0.5f * Svm;
                const float Hf = 0.5f * (S.Hi[p][lc] + S.Hi[p][lc + 1]);
                const float hp_raw = (Ep + Hf) - tp;
                const float hm_raw = (Em + Hf) + tm;
                const float hp = fmaxf(0.0f, hp_raw), hm = fmaxf(0.0f, hm_raw);
                float ghm, gnm, gtm, ghp, gnp, gtp;
                flux_vjp(hm, Rmy, Rmx, hp, Rpy, Rpx, (km >> kCodeFy) & 31u, A.g, A.cutoff, l0, l1, l2,
                         ghm, gnm, gtm, ghp, gnp, gtp);
                const float ghm_r = hm_raw > 0.0f ? ghm : 0.0f;
                const float ghp_r = hp_raw > 0.0f ? ghp : 0.0f;
                // back to (E, u, v): normal = v, tangential = u
                out[0] = ghm_r; out[1] = gtm; out[2] = gnm;
                out[3] = ghp_r; out[4] = gtp; out[5] = gnp;
                out[6] = ghm_r; out[7] = 0.5f * gtm; out[8] = 0.5f * gnm;
                out[9] = -ghp_r; out[10] = -0.5f * gtp; out[11] = -0.5f * gnp;
            }
        }
        const int k = fj * BW + i;
        for (int q = 0; q < 6; ++q) { S.FD[q][k] = out[q]; S.FC[q][k] = out[6 + q]; }
    }

    __host__ __device__ void stage_ycentre(int it, float (&out)[9]) const {
        const int fj = it / BW, i = it - fj * BW;
        const int lc = i + 3, c = Rof(fj);
        for (int q = 0; q < 9; ++q) out[q] = 0.0f;
        if (y_col_ok(i) && c >= fy0 && c <= fy1 + 1 && fj >= 1 && fj + 1 < H) {
            const int kl = (fj - 1) * BW + i, kr = fj * BW + i;
            const float gT = S.FC[3][kl] + S.FC[0][kr];
            const float gSu = S.FC[4][kl] + S.FC[1][kr];
            const float gSv = S.FC[5][kl] + S.FC[2][kr];
            const uint32_t code = S.code[fj][lc] >> 8;
            const float th = A.theta;
            float cl, cc, cr;
            slope_coef(code & 3u, th, cl, cc, cr);
            out[1] = cl * gSu; out[4] = cc * gSu; out[7] = cr * gSu;
            slope_coef((code >> 2) & 3u, th, cl, cc, cr);
            out[2] = cl * gSv; out[5] = cc * gSv; out[8] = cr * gSv;
            // S = (L - dy f s_c u_c) / 2g: y direction, sigma = +1, t = u
            const int R = c;
            const float sa = (A.bc_s == kWall && R < 2) ? -1.0f : 1.0f;
            float gE[3] = {0.0f, 0.0f, 0.0f}, gt[3] = {0.0f, 0.0f, 0.0f};
            depth_term_vjp((code >> 4) & 3u, (code >> 6) & 3u, 1.0f, th, A.g, A.f, A.Ddy, A.Dy, A.r_two_g,
                           row_flip(R - 1), row_flip(R), row_flip(R + 1), sa, gT, gE, gt);
            out[0] += gE[0]; out[3] += gE[1]; out[6] += gE[2];
            out[1] += gt[0]; out[4] += gt[1]; out[7] += gt[2];
        }
    }

    __host__ __device__ void stage_ygather(int it) {
        const int fj = it / BW, i = it - fj * BW;
        const int lc = i + 3;
        if (!y_col_ok(i) || fj < 1 || fj + 1 >= H) return;
        if (!is_output(Rof(fj), Cof(lc))) return;
        const int k = fj * BW + i;
        S.A0[fj - 1][lc - 1] += ((S.FD[3][k - BW] + S.FD[0][k]) + (S.FC[6][k - BW] + S.FC[3][k])) + S.FC[0][k + BW];
        S.A1[fj - 1][lc - 1] += ((S.FD[4][k - BW] + S.FD[1][k]) + (S.FC[7][k - BW] + S.FC[4][k])) + S.FC[1][k + BW];
        S.A2[fj - 1][lc - 1] += ((S.FD[5][k - BW] + S.FD[2][k]) + (S.FC[8][k - BW] + S.FC[5][k])) + S.FC[2][k + BW];
    }

    // ---- output: base / source VJP (interior cells), conversion VJP, store
    // The x term and its clamp value are loaded before stage 0 (load_out).
    struct OutIn {
        float xe, xu, xv, xc;
    };
    __host__ __device__ bool out_used(int lr, int lc, int R, int C) const {
        return lr >= 1 && lc >= 1 && lr + 1 < H && lc + 1 < W && is_output(R, C);
    }
    __host__ __device__ void load_out(int it, OutIn& in) const {
        const int lr = it / W, lc = it - lr * W;
        const int R = Rof(lr), C = Cof(lc);
        in.xe = in.xu = in.xv = in.xc = 0.0f;
        if (!A.x_eta || !out_used(lr, lc, R, C) || ghost_r(R) || ghost_c(C)) return;
        const size_t o = (size_t)R * A.pitch + C;
        if (A.x_clamp) in.xc = A.x_clamp[o];
        in.xe = A.x_eta[o]; in.xu = A.x_hu[o]; in.xv = A.x_hv[o];
    }
    __host__ __device__ void stage_out(int it, const OutIn& in) {
        const int lr = it / W, lc = it - lr * W;
        const int R = Rof(lr), C = Cof(lc);
        if (!out_used(lr, lc, R, C)) return;
        float gE = S.A0[lr - 1][lc - 1], gU = S.A1[lr - 1][lc - 1], gV = S.A2[lr - 1][lc - 1];
        const size_t o = (size_t)R * A.pitch + C;
        const bool interior = !ghost_r(R) && !ghost_c(C);
        const float E = S.E[lr][lc], U = S.U[lr][lc], V = S.V[lr][lc];
        float xe = 0.0f, xu = 0.0f, xv = 0.0f;
        if (interior) {
            const float w0 = S.W0[lr - 1][lc - 1], w1 = S.W1[lr - 1][lc - 1], w2 = S.W2[lr - 1][lc - 1];
            const float a = S.Hi[lr][lc], b = S.Hi[lr + 1][lc], c = S.Hi[lr][lc + 1], d = S.Hi[lr + 1][lc + 1];
            const float Hreg = 0.25f * (((a + b) + c) + d);
            const float hreg = E + Hreg;
            gE += w0;
            float ghu = w1, ghv = w2;
            if (hreg >= A.cutoff && E != kDryFlag) {
                const float H_x = 0.5f * (c + d) - 0.5f * (a + b);
                const float H_y = 0.5f * (b + d) - 0.5f * (a + c);
                ghv += dt * w1 * A.f;
                gE += dt * w1 * A.g_dx * H_x;
                ghu += -dt * w2 * A.f;
                gE += dt * w2 * A.g_dy * H_y;
            }
            if (hreg > A.cutoff) {
                gU += ghu * hreg; gE += ghu * U;
                gV += ghv * hreg; gE += ghv * V;
            }
            if (A.x_eta && !clamped_v(A.x_clamp != nullptr, in.xc, lr, lc)) {
                if (A.x_fold && near_boundary(R, C)) {
                    xe = fold_read(A.x_eta, 0, R, C); xu = fold_read(A.x_hu, 1, R, C);
                    xv = fold_read(A.x_hv, 2, R, C);
                } else { xe = in.xe; xu = in.xu; xv = in.xv; }
            }
        }
        // conversion (eta, hu, hv) -> (E, u, v)
        const uint32_t br = S.code[lr][lc] >> 28;
        // wet / desingularised cells have E = eta, so h = E + Hm as in the forward
        const float h = (br == kBrWet || br == kBrDesing) ? E + cell_hm(lr, lc) : 0.0f;
        float g_eta = 0.0f, g_hu = 0.0f, g_hv = 0.0f;
        if (br == kBrWet) {
            const float rh = adj_rcp(h);
            g_eta = gE - (gU * U + gV * V) * rh;
            g_hu = gU * rh;
            g_hv = gV * rh;
        } else if (br == kBrDesing) {
            const float eps = A.desing;
            const float r_eps = adj_rcp(eps);
            const float aq = h * h * (0.5f * r_eps) + 0.5f * eps;
            const float mn = fminf(aq, eps);
            const float den = fmaxf(mn, fabsf(h));
            float dden = 0.0f;
            if (fabsf(h) >= mn) dden = (h >= 0.0f) ? 1.0f : -1.0f;
            else if (aq <= eps) dden = h * r_eps;
            const float rden = adj_rcp(den);
            g_eta = gE - (gU * U + gV * V) * dden * rden;
            g_hu = gU * rden;
            g_hv = gV * rden;
        }
        A.o_eta[o] = A.alpha * g_eta + A.gamma * xe;
        A.o_hu[o] = A.alpha * g_hu + A.gamma * xu;
        A.o_hv[o] = A.alpha * g_hv + A.gamma * xv;
    }
};

#ifdef __HIP_DEVICE_COMPILE__
#define ADJ_SYNC() __syncthreads()
#else
#define ADJ_SYNC()
#endif

// every stage of one tile, one item per thread and stage (NT threads cover
// the largest stage).  PREFETCH issues every stage's global loads first,
// before the corner depths, so that no stage after the first barrier waits
// on memory; it costs scalar registers (the batched kernel, which runs many
// rounds of tiles, gains from it; one 512^2 domain, a single round, does not)
template <int BW, int BH, int NT, bool PREFETCH>
__host__ __device__ inline void adj_tile_body(const AdjTileArgs& A, AdjSmem<BW, BH>& S, int tx, int ty,
                                             int tid) {
    using Tl = AdjTile<BW, BH>;
    static_assert(NT >= Tl::kCellItems && NT >= Tl::kXItems && NT >= Tl::kYItems,
                  "one item per thread in every stage");
    Tl T(A, S, tx, ty);
    typename Tl::CellIn cin;
    typename Tl::OutIn oin;
    float xtm = 0.0f, xtp = 0.0f, ytm = 0.0f, ytp = 0.0f;
    const bool cell = tid < Tl::kCellItems, xi = tid < Tl::kXItems, yi = tid < Tl::kYItems;
    if (PREFETCH) {
        if (cell) { T.load_cell(tid, cin); T.load_out(tid, oin); }
        if (xi) T.load_xface(tid, xtm, xtp);
        if (yi) T.load_yface(tid, ytm, ytp);
    }
    for (int it = tid; it < Tl::kHiItems; it += NT) T.stage_hi(it);
    ADJ_SYNC();
    if (cell) {
        if (!PREFETCH) T.load_cell(tid, cin);
        T.stage_cells(tid, cin);
    }
    ADJ_SYNC();
    if (xi) {
        if (!PREFETCH) T.load_xface(tid, xtm, xtp);
        T.stage_xface(tid, xtm, xtp);
    }
    ADJ_SYNC();
    {
        // slope centres overwrite the face rows they read: results are held
        // in registers until every centre has read them
        float out[9];
        if (xi) T.stage_xcentre(tid, out);
        ADJ_SYNC();
        if (xi) T.store_centre(tid, out);
    }
    ADJ_SYNC();
    if (xi) T.stage_xgather(tid);
    ADJ_SYNC();
    if (yi) {
        if (!PREFETCH) T.load_yface(tid, ytm, ytp);
        T.stage_yface(tid, ytm, ytp);
    }
    ADJ_SYNC();
    {
        float out[9];
        if (yi) T.stage_ycentre(tid, out);
        ADJ_SYNC();
        if (yi) T.store_centre(tid, out);
    }
    ADJ_SYNC();
    if (yi) T.stage_ygather(tid);
    ADJ_SYNC();
    if (cell) {
        if (!PREFETCH) T.load_out(tid, oin);
        T.stage_out(tid, oin);
    }
}

constexpr int kAdjBW = 16, kAdjBH = 16, kAdjNT = 512;

// Ghost cells back to the interior cells they are images of.  One thread per
// interior cell within two cells of the boundary; ghost entries are read by
// exactly one thread and cleared by it; the 16 corner entries are cleared by
// the last block row.
struct FoldArgs {
    float* e; float* u; float* v;
    int nx, ny, pitch;
    int bc_n, bc_e, bc_s, bc_w;
};

__host__ __device__ inline void fold_cell(const FoldArgs& F, int R, int C) {
    const int nx = F.nx, ny = F.ny;
    const size_t o = (size_t)R * F.pitch + C;
    float se = F.e[o], su = F.u[o], sv = F.v[o];
    auto take = [&](int gr, int gc, float fu, float fv) {
        const size_t q = (size_t)gr * F.pitch + gc;
        se += F.e[q]; su += fu * F.u[q]; sv += fv * F.v[q];
        F.e[q] = 0.0f; F.u[q] = 0.0f; F.v[q] = 0.0f;
    };
    if (R <= 3) {          // south side ghosts whose image is R
        if (F.bc_s == kWall) take(3 - R, C, 1.0f, -1.0f);
        else take(R + ny, C, 1.0f, 1.0f);
    }
    if (R >= ny) {         // north side
        if (F.bc_n == kWall) take(2 * ny + 3 - R, C, 1.0f, -1.0f);
        else take(R - ny, C, 1.0f, 1.0f);
    }
    if (C <= 3) {
        if (F.bc_w == kWall) take(R, 3 - C, -1.0f, 1.0f);
        else take(R, C + nx, 1.0f, 1.0f);
    }
    if (C >= nx) {
        if (F.bc_e == kWall) take(R, 2 * nx + 3 - C, -1.0f, 1.0f);
        else take(R, C - nx, 1.0f, 1.0f);
    }
    F.e[o] = se; F.u[o] = su; F.v[o] = sv;
}

// blockIdx.y 0..3: rows 2, 3, ny, ny+1 (all interior columns); 4..7: columns
// 2, 3, nx, nx+1 over rows 4..ny-1; 8: corners
__host__ __device__ inline void fold_item(const FoldArgs& F, int y, int x) {
    const int nx = F.nx, ny = F.ny;
    if (y < 4) {
        const int R = y < 2 ? 2 + y : ny + (y - 2);
        if ((y >= 2 && R <= 3) || R > ny + 1) return;
        const int C = 2 + x;
        if (C > nx + 1) return;
        fold_cell(F, R, C);
    } else if (y < 8) {
        const int k = y - 4;
        const int C = k < 2 ? 2 + k : nx + (k - 2);
        if ((k >= 2 && C <= 3) || C > nx + 1) return;
        const int R = 4 + x;
        if (R > ny - 1) return;
        fold_cell(F, R, C);
    } else if (x < 16) {
        const int R = (x & 3) < 2 ? (x & 3) : ny + (x & 3);
        const int C = (x >> 2) < 2 ? (x >> 2) : nx + (x >> 2);
        const size_t q = (size_t)R * F.pitch + C;
        F.e[q] = 0.0f; F.u[q] = 0.0f; F.v[q] = 0.0f;
    }
}

}  // namespace gocean
