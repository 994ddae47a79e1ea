// gocean_adj.hip -- hand-written adjoint (vector-Jacobian product) of the
// CDKLM16 substep, for the fixed-dt SSP-RK2 step.
//
// The forward substep is  out = P(Q) = base(Q) + dt*L(Q)  (rk_step 0) and
// out = 0.5*(Q^n + P(Q*))  (rk_step 1), where L is the flux divergence plus
// sources of gpu_ocean/SWESimulators/gpu_kernels/CDKLM16_kernel.cu:429-1004
// and base(Q) = (E, u*h_reg, v*h_reg) the converted state the update starts
// from (:597-602, :905-920).  gocean_adjoint_substep returns
//     alpha * J_P(Q)^T w  +  gamma * extra
// with every branch (minmod candidate, wave-speed min/max, upwinding,
// positivity adjustment, wet/dry/desingularised cells) fixed to the forward
// choice.  Scope: zero grid rotation, constant f, walls/periodic, no
// friction.  Cells the forward update clamped to the dry state are recognised
// from the forward outputs (optional arguments) and pass no adjoint.
//
// Structure (all kernels one thread per face or cell, deterministic gathers,
// no atomics):
//   k_adj_xface / k_adj_yface  recompute the four cells around a face
//       (conversion, slopes, adjustment, face depth term), evaluate the face
//       flux, and apply its VJP with lambda_F = dt*(w_p - w_m)/dx back to the
//       converted (E, u, v) of those four cells -> 12 scratch planes each;
//   k_adj_cell   gathers the face contributions of a cell, adds the base and
//       source-term VJP (output cells), applies the conversion VJP and writes
//       the cell's adjoint, ghost cells included;
//   k_adj_fold_ns / k_adj_fold_ew  map ghost-cell adjoints back through the
//       wall mirror / periodic wrap into interior cells and clear the ghosts.
#include <hip/hip_runtime.h>

#include <cstring>

#include "../../include/gocean.h"
#include "cdklm16_tile.hpp"
#include "gocean_internal.hpp"

namespace gocean {

struct AdjArgs {
    const float* eta; const float* hu; const float* hv;     // substep input Q
    const float* w_eta; const float* w_hu; const float* w_hv;
    const float* x_eta; const float* x_hu; const float* x_hv;
    const float* w_clamp;   // forward output (eta) whose dry clamp masks w
    const float* x_clamp;   // forward output (eta) whose dry clamp masks x
    float* o_eta; float* o_hu; float* o_hv;
    const float* Hi;
    float* xc;       // 12 planes: x-face contributions
    float* yc;       // 12 planes: y-face contributions
    const float* dt_dev;
    float dt, alpha, gamma;
    int nx, ny, pitch, hi_pitch;
    long long plane;
    float dx, dy, g, theta, desing, flux_eps, cutoff, land, f;
    int bc_n, bc_e, bc_s, bc_w;
};

struct Cell {
    float E, U, V;
};

// state at padded (R, C) read the way the forward tile reads it (wall mirror
// with the normal momentum negated, periodic wrap), converted to (E, u, v)
__host__ __device__ inline Cell adj_cell(const AdjArgs& A, int R, int C, float* h_out = nullptr,
                                         int* br = nullptr, float* Hm_out = nullptr) {
    const int nx = A.nx, ny = A.ny;
    int sr = R, sc = C;
    bool nu = false, nv = false;
    const bool wN = A.bc_n == kWall, wS = A.bc_s == kWall, wE = A.bc_e == kWall, wW = A.bc_w == kWall;
    const bool pNS = A.bc_n == kPeriodic, pEW = A.bc_e == kPeriodic;
    if (R < 2) { if (wS) { sr = 3 - R; nv = true; } else if (pNS) sr = R + ny; }
    else if (R > ny + 1) { if (wN) { sr = 2 * ny + 3 - R; nv = true; } else if (pNS) sr = R - ny; }
    if (C < 2) { if (wW) { sc = 3 - C; nu = true; } else if (pEW) sc = C + nx; }
    else if (C > nx + 1) { if (wE) { sc = 2 * nx + 3 - C; nu = true; } else if (pEW) sc = C - nx; }
    const size_t o = (size_t)sr * A.pitch + sc;
    const float eta = A.eta[o];
    float hu = A.hu[o], hv = A.hv[o];
    if (nu) hu = -hu;
    if (nv) hv = -hv;
    auto flag = [&](float v) { return (fabsf(v - A.land) < kDryEps) ? kDryFlag : v; };
    const float a = flag(A.Hi[(size_t)R * A.hi_pitch + C]);
    const float b = flag(A.Hi[(size_t)(R + 1) * A.hi_pitch + C]);
    const float c = flag(A.Hi[(size_t)R * A.hi_pitch + C + 1]);
    const float d = flag(A.Hi[(size_t)(R + 1) * A.hi_pitch + C + 1]);
    const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag) - (float)(c == kDryFlag) - (float)(d == kDryFlag);
    const float sum = ((a + b) + c) + d;
    const float Hm = (wet == 0.0f) ? A.land : sum / wet;
    const float h = eta + Hm;
    Cell r;
    int branch;
    if (fabsf(Hm - A.land) <= kDryEps) { r.E = kDryFlag; r.U = 0.0f; r.V = 0.0f; branch = 0; }
    else if (h < A.desing) {
        if (h <= A.cutoff) { r.E = -Hm + A.cutoff; r.U = 0.0f; r.V = 0.0f; branch = 1; }
        else { r.E = eta; r.U = desingularize(h, hu, A.desing); r.V = desingularize(h, hv, A.desing); branch = 2; }
    } else { r.E = eta; r.U = hu / h; r.V = hv / h; branch = 3; }
    if (h_out) *h_out = h;
    if (br) *br = branch;
    if (Hm_out) *Hm_out = Hm;
    return r;
}

__host__ __device__ inline float hi_flag(const AdjArgs& A, int r, int c) {
    const float v = A.Hi[(size_t)r * A.hi_pitch + c];
    return (fabsf(v - A.land) < kDryEps) ? kDryFlag : v;
}

// true where the forward update clamped cell (R, C) to the dry state
// (eta = -Hm + cutoff, hu = hv = 0): that output is constant, so the adjoint
// arriving at it is dropped (CDKLM16_kernel.cu:986-1001)
__host__ __device__ inline bool dry_clamped(const AdjArgs& A, const float* out_eta, int R, int C) {
    if (!out_eta) return false;
    const float Hreg = 0.25f * (((hi_flag(A, R, C) + hi_flag(A, R + 1, C)) + hi_flag(A, R, C + 1))
                                + hi_flag(A, R + 1, C + 1));
    return out_eta[(size_t)R * A.pitch + C] <= -Hreg + A.cutoff;
}

// index of the minmod candidate that is the result (0,1,2) or -1 for zero
__host__ __device__ inline int minmod_sel(float b, float c, float f) {
    const uint32_t ub = __builtin_bit_cast(uint32_t, b), uc = __builtin_bit_cast(uint32_t, c);
    const uint32_t uf = __builtin_bit_cast(uint32_t, f);
    if ((((ub ^ uc) | (uc ^ uf)) & 0x80000000u) != 0u) return -1;
    const float ab = fabsf(b), ac = fabsf(c), af = fabsf(f);
    if (ab <= ac && ab <= af) return 0;
    if (ac <= af) return 1;
    return 2;
}

// VJP of the central-upwind flux (CDKLM16_flux) w.r.t. (hm,um,vm,hp,up,vp)
__host__ __device__ inline void flux_vjp(float hm, float um, float vm, float hp, float up, float vp,
                         float g, float cutoff, float feps, float gx, float gy, float gz,
                         float& ghm, float& gum, float& gvm, float& ghp, float& gup, float& gvp) {
    ghm = gum = gvm = ghp = gup = gvp = 0.0f;
    const bool wp = hp > cutoff, wm = hm > cutoff;
    float Fpx = 0, Fpy = 0, u_p = 0, c_p = 0;
    if (wp) { Fpx = hp * up; Fpy = hp * up * up + 0.5f * g * hp * hp; u_p = up; c_p = sqrtf(g * hp); }
    float Fmx = 0, Fmy = 0, u_m = 0, c_m = 0;
    if (wm) { Fmx = hm * um; Fmy = hm * um * um + 0.5f * g * hm * hm; u_m = um; c_m = sqrtf(g * hm); }
    const float a1 = u_m - c_m, a2 = u_p - c_p;
    const float b1 = u_m + c_m, b2 = u_p + c_p;
    const float am = fminf(fminf(a1, a2), 0.0f);
    const float ap = fmaxf(fmaxf(b1, b2), 0.0f);
    const float den = ap - am;
    if (fabsf(den) < feps) return;
    const float apam = ap * am;
    const float Fx = ((ap * Fmx - am * Fpx) + apam * (hp - hm)) / den;
    const float Fy = ((ap * Fmy - am * Fpy) + apam * (Fpx - Fmx)) / den;
    const float gNx = gx / den, gNy = gy / den;
    const float gden = -(gx * Fx + gy * Fy) / den;
    const float gap = gNx * (Fmx + am * (hp - hm)) + gNy * (Fmy + am * (Fpx - Fmx)) + gden;
    const float gam = gNx * (-Fpx + ap * (hp - hm)) + gNy * (-Fpy + ap * (Fpx - Fmx)) - gden;
    float gFmx = gNx * ap - gNy * apam;
    float gFpx = -gNx * am + gNy * apam;
    const float gFmy = gNy * ap;
    const float gFpy = -gNy * am;
    ghp += gNx * apam;
    ghm -= gNx * apam;
    float gFmz = 0.0f, gFpz = 0.0f;
    if (um + up > 0.0f) gFmz = gz; else gFpz = gz;
    float gu_m = 0, gc_m = 0, gu_p = 0, gc_p = 0;
    if (ap > 0.0f) {                       // ap = max(b1, b2, 0)
        if (b1 >= b2) { gu_m += gap; gc_m += gap; } else { gu_p += gap; gc_p += gap; }
    }
    if (am < 0.0f) {                       // am = min(a1, a2, 0)
        if (a1 <= a2) { gu_m += gam; gc_m -= gam; } else { gu_p += gam; gc_p -= gam; }
    }
    if (wp) {
        ghp += gFpx * up + gFpy * (up * up + g * hp) + gFpz * up * vp + gc_p * g / (2.0f * c_p);
        gup += gFpx * hp + gFpy * 2.0f * hp * up + gFpz * hp * vp + gu_p;
        gvp += gFpz * hp * up;
    }
    if (wm) {
        ghm += gFmx * um + gFmy * (um * um + g * hm) + gFmz * um * vm + gc_m * g / (2.0f * c_m);
        gum += gFmx * hm + gFmy * 2.0f * hm * um + gFmz * hm * vm + gu_m;
        gvm += gFmz * hm * um;
    }
}

// the minmod slope of a scalar and its VJP onto (l, c, r)
__host__ __device__ inline void slope_vjp(float l, float c, float r, float theta, float gs,
                                          float& gl, float& gc, float& gr) {
    const int s = minmod_sel((c - l) * theta, (r - l) * 0.5f, (r - c) * theta);
    if (s == 0) { gc += theta * gs; gl -= theta * gs; }
    else if (s == 1) { gr += 0.5f * gs; gl -= 0.5f * gs; }
    else if (s == 2) { gr += theta * gs; gc -= theta * gs; }
}

// One slope centre k (0..3 among the four cells) along a line.  dir 0: x
// (tangential v, K term -Vd*f*v), dir 1: y (tangential u, L term +Ud*f*u).
// Computes the primal slopes (Sn = normal-velocity slope, St = tangential,
// Tt = face depth term T or S) and, when given adjoints, accumulates the VJP.
struct LineCells {
    float E[4], U[4], V[4];
    int line_idx[4];        // global column (x) or row (y) of each cell
};

template <int DIR>
__host__ __device__ inline void slope_term(const AdjArgs& A, const LineCells& L, int k, int other,
                           float& Su, float& Sv, float& Tt, bool adjoint,
                           float gT, float gSu, float gSv, float* gE, float* gU, float* gV) {
    // other = the fixed row (x) or column (y) index of the line
    const float g = A.g, theta = A.theta, f = A.f;
    const float two_g = 2.0f * g;
    const int i0 = k - 1, i1 = k, i2 = k + 1;
    const float El = L.E[i0], Ec = L.E[i1], Er = L.E[i2];
    const float Ul = L.U[i0], Uc = L.U[i1], Ur = L.U[i2];
    const float Vl = L.V[i0], Vc = L.V[i1], Vr = L.V[i2];
    const int Cc = L.line_idx[i1];
    auto flip = [&](int idx) -> float {
        if (DIR == 0) return ((A.bc_w == kWall && idx < 2) || (A.bc_e == kWall && idx > A.nx + 1)) ? -1.0f : 1.0f;
        return ((A.bc_s == kWall && idx < 2) || (A.bc_n == kWall && idx > A.ny + 1)) ? -1.0f : 1.0f;
    };
    const float sl = flip(L.line_idx[i0]), sc = flip(Cc), sr = flip(L.line_idx[i2]);
    Su = minmod_slope(Ul, Uc, Ur, theta);
    Sv = minmod_slope(Vl, Vc, Vr, theta);
    // tangential velocity in the geostrophic term
    const float tl = (DIR == 0 ? Vl : Ul) * sl, tc = (DIR == 0 ? Vc : Uc) * sc, tr = (DIR == 0 ? Vr : Ur) * sr;
    const float fl = tl * f, fc = tc * f, fr = tr * f;
    const float D = (DIR == 0 ? A.dx : A.dy) / two_g;
    const float sg = (DIR == 0) ? -1.0f : 1.0f;     // x: minus, y: plus
    const float cb = theta * g * ((Ec - El) + sg * D * (fc + fl));
    const float cc = 0.5f * g * ((Er - El) + sg * D * ((fr + 2.0f * fc) + fl));
    const float cf = theta * g * ((Er - Ec) + sg * D * (fc + fr));
    float K = minmod_raw(cb, cc, cf);
    const int sel = minmod_sel(cb, cc, cf);
    // positivity adjustment (x: west/east faces, y: south/north faces)
    float Hlo, Hhi;
    if (DIR == 0) {
        const int R = other;
        Hlo = 0.5f * (hi_flag(A, R, Cc) + hi_flag(A, R + 1, Cc));
        Hhi = 0.5f * (hi_flag(A, R, Cc + 1) + hi_flag(A, R + 1, Cc + 1));
    } else {
        const int C = other;
        Hlo = 0.5f * (hi_flag(A, Cc, C) + hi_flag(A, Cc, C + 1));
        Hhi = 0.5f * (hi_flag(A, Cc + 1, C) + hi_flag(A, Cc + 1, C + 1));
    }
    const bool flip_a = (DIR == 0) ? (A.bc_w == kWall && Cc < 2) : (A.bc_s == kWall && Cc < 2);
    const float traw = (DIR == 0 ? Vc : Uc);
    const float t_adj = flip_a ? -traw : traw;
    const float Dd = (DIR == 0 ? A.dx : A.dy) * f;
    // x: t = (K + dx f v)/2g; y: t = (L - dy f u)/2g
    const float tq = (K - sg * Dd * t_adj) / two_g;
    const float hlo = (Ec + Hlo) - tq, hhi = (Ec + Hhi) + tq;
    int adj = 0;
    if (!(hlo > 0.0f)) { K = sg * Dd * t_adj + two_g * (Ec + Hlo); adj = 1; }
    if (!(hhi > 0.0f)) { K = sg * Dd * t_adj - two_g * (Ec + Hhi); adj = 2; }
    Tt = (K - sg * Dd * tc) / two_g;
    if (!adjoint) return;
    // ---- VJP ----
    const float gK = gT / two_g;
    // Tt's tangential term
    const float gtc_T = gT * (-sg * Dd) / two_g;
    float gfl = 0, gfc = 0, gfr = 0, gtadj = 0;
    if (adj == 2) { gE[i1] += -two_g * gK; gtadj += sg * Dd * gK; }
    else if (adj == 1) { gE[i1] += two_g * gK; gtadj += sg * Dd * gK; }
    else if (sel == 0) {
        const float q = theta * g * gK;
        gE[i1] += q; gE[i0] -= q; gfc += q * sg * D; gfl += q * sg * D;
    } else if (sel == 1) {
        const float q = 0.5f * g * gK;
        gE[i2] += q; gE[i0] -= q; gfr += q * sg * D; gfc += 2.0f * q * sg * D; gfl += q * sg * D;
    } else if (sel == 2) {
        const float q = theta * g * gK;
        gE[i2] += q; gE[i1] -= q; gfc += q * sg * D; gfr += q * sg * D;
    }
    // f*t terms back to the tangential velocity of each cell
    float* gT_ = (DIR == 0) ? gV : gU;
    gT_[i0] += f * sl * gfl;
    gT_[i1] += f * sc * gfc + sc * gtc_T + (flip_a ? -gtadj : gtadj);
    gT_[i2] += f * sr * gfr;
    // velocity slopes
    slope_vjp(Ul, Uc, Ur, theta, gSu, gU[i0], gU[i1], gU[i2]);
    slope_vjp(Vl, Vc, Vr, theta, gSv, gV[i0], gV[i1], gV[i2]);
}

template <int DIR>
__host__ __device__ inline void face_adjoint(const AdjArgs& A, int line, int m) {
    // DIR 0: x-face in row `line` between columns m and m+1
    // DIR 1: y-face in column `line` between rows m and m+1
    const float dt = A.dt_dev ? *A.dt_dev : A.dt;
    LineCells L;
    for (int k = 0; k < 4; ++k) {
        const int idx = m - 1 + k;
        L.line_idx[k] = idx;
        const Cell c = (DIR == 0) ? adj_cell(A, line, idx) : adj_cell(A, idx, line);
        L.E[k] = c.E; L.U[k] = c.U; L.V[k] = c.V;
    }
    float gE[4] = {0, 0, 0, 0}, gU[4] = {0, 0, 0, 0}, gV[4] = {0, 0, 0, 0};
    const bool dry = (L.E[1] == kDryFlag || L.E[2] == kDryFlag);
    if (!dry) {
        // output weights of the two adjacent cells (0 for ghost cells)
        const int n_line = (DIR == 0) ? A.nx : A.ny;
        float wm[3] = {0, 0, 0}, wp[3] = {0, 0, 0};
        auto wload = [&](int idx, float* w) {
            if (idx < 2 || idx > n_line + 1) return;
            const int R = (DIR == 0) ? line : idx, C = (DIR == 0) ? idx : line;
            if (dry_clamped(A, A.w_clamp, R, C)) return;
            const size_t o = (size_t)R * A.pitch + C;
            w[0] = A.w_eta[o]; w[1] = A.w_hu[o]; w[2] = A.w_hv[o];
        };
        wload(m, wm);
        wload(m + 1, wp);
        const float D = (DIR == 0) ? A.dx : A.dy;
        const float gF0 = dt * (wp[0] - wm[0]) / D;
        const float gF1 = dt * (wp[1] - wm[1]) / D;
        const float gF2 = dt * (wp[2] - wm[2]) / D;
        // primal slopes and depth terms at m (k=1) and p (k=2)
        float Sum, Svm, Tm, Sup, Svp, Tp;
        slope_term<DIR>(A, L, 1, line, Sum, Svm, Tm, false, 0, 0, 0, gE, gU, gV);
        slope_term<DIR>(A, L, 2, line, Sup, Svp, Tp, false, 0, 0, 0, gE, gU, gV);
        float Hf;
        if (DIR == 0) Hf = 0.5f * (hi_flag(A, line, m + 1) + hi_flag(A, line + 1, m + 1));
        else Hf = 0.5f * (hi_flag(A, m + 1, line) + hi_flag(A, m + 1, line + 1));
        const float Rpx = L.U[2] - 0.5f * Sup, Rpy = L.V[2] - 0.5f * Svp;
        const float Rmx = L.U[1] + 0.5f * Sum, Rmy = L.V[1] + 0.5f * Svm;
        const float hp_raw = (L.E[2] + Hf) - Tp;
        const float hm_raw = (L.E[1] + Hf) + Tm;
        const float hp = fmaxf(0.0f, hp_raw), hm = fmaxf(0.0f, hm_raw);
        float ghm, gn_m, gt_m, ghp, gn_p, gt_p;
        if (DIR == 0) {
            flux_vjp(hm, Rmx, Rmy, hp, Rpx, Rpy, A.g, A.cutoff, A.flux_eps, gF0, gF1, gF2,
                     ghm, gn_m, gt_m, ghp, gn_p, gt_p);
        } else {
            // flux(hm, Rmy, Rmx, hp, Rpy, Rpx) = (a, b, c); G = (a, c, b)
            flux_vjp(hm, Rmy, Rmx, hp, Rpy, Rpx, A.g, A.cutoff, A.flux_eps, gF0, gF2, gF1,
                     ghm, gn_m, gt_m, ghp, gn_p, gt_p);
        }
        // map back to (Rx, Ry) adjoints
        float gRmx, gRmy, gRpx, gRpy;
        if (DIR == 0) { gRmx = gn_m; gRmy = gt_m; gRpx = gn_p; gRpy = gt_p; }
        else { gRmy = gn_m; gRmx = gt_m; gRpy = gn_p; gRpx = gt_p; }
        float gTp = 0, gTm = 0;
        if (hp_raw > 0.0f) { gE[2] += ghp; gTp = -ghp; }
        if (hm_raw > 0.0f) { gE[1] += ghm; gTm = ghm; }
        gU[2] += gRpx; gV[2] += gRpy;
        gU[1] += gRmx; gV[1] += gRmy;
        const float gSup = -0.5f * gRpx, gSvp = -0.5f * gRpy;
        const float gSum = 0.5f * gRmx, gSvm = 0.5f * gRmy;
        float d0, d1, d2;
        slope_term<DIR>(A, L, 1, line, d0, d1, d2, true, gTm, gSum, gSvm, gE, gU, gV);
        slope_term<DIR>(A, L, 2, line, d0, d1, d2, true, gTp, gSup, gSvp, gE, gU, gV);
    }
    float* out = (DIR == 0) ? A.xc : A.yc;
    const size_t o = (DIR == 0) ? (size_t)line * A.pitch + m : (size_t)m * A.pitch + line;
    for (int k = 0; k < 4; ++k) {
        out[(size_t)(3 * k + 0) * A.plane + o] = gE[k];
        out[(size_t)(3 * k + 1) * A.plane + o] = gU[k];
        out[(size_t)(3 * k + 2) * A.plane + o] = gV[k];
    }
}

__global__ __launch_bounds__(256) void k_adj_xface(AdjArgs A) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x + 1;     // 1..nx+1
    const int r = blockIdx.y + 2;                                  // 2..ny+1
    if (m > A.nx + 1) return;
    face_adjoint<0>(A, r, m);
}

__global__ __launch_bounds__(256) void k_adj_yface(AdjArgs A) {
    const int c = blockIdx.x * blockDim.x + threadIdx.x + 2;     // 2..nx+1
    const int m = blockIdx.y + 1;                                  // 1..ny+1
    if (c > A.nx + 1) return;
    face_adjoint<1>(A, c, m);
}

__host__ __device__ inline void cell_adjoint(const AdjArgs& A, int R, int C) {
    const int nx = A.nx, ny = A.ny;
    const size_t o = (size_t)R * A.pitch + C;
    const bool rint = R >= 2 && R <= ny + 1, cint = C >= 2 && C <= nx + 1;
    float gE = 0, gU = 0, gV = 0;
    if (rint) {   // x faces in this row with left column C-2..C+1
        for (int k = 0; k < 4; ++k) {
            const int cm = C + 1 - k;          // this cell is cell k of the face's window
            if (cm < 1 || cm > nx + 1) continue;
            const size_t q = (size_t)R * A.pitch + cm;
            gE += A.xc[(size_t)(3 * k + 0) * A.plane + q];
            gU += A.xc[(size_t)(3 * k + 1) * A.plane + q];
            gV += A.xc[(size_t)(3 * k + 2) * A.plane + q];
        }
    }
    if (cint) {   // y faces in this column with lower row R-2..R+1
        for (int k = 0; k < 4; ++k) {
            const int rm = R + 1 - k;
            if (rm < 1 || rm > ny + 1) continue;
            const size_t q = (size_t)rm * A.pitch + C;
            gE += A.yc[(size_t)(3 * k + 0) * A.plane + q];
            gU += A.yc[(size_t)(3 * k + 1) * A.plane + q];
            gV += A.yc[(size_t)(3 * k + 2) * A.plane + q];
        }
    }
    if (!(rint || cint)) {
        A.o_eta[o] = 0.0f; A.o_hu[o] = 0.0f; A.o_hv[o] = 0.0f;
        return;
    }
    float h, Hm;
    int br;
    const Cell c = adj_cell(A, R, C, &h, &br, &Hm);
    float xe = 0, xu = 0, xv = 0;
    if (rint && cint) {
        // base (E, u*h_reg, v*h_reg) and source terms of this output cell
        const float dt = A.dt_dev ? *A.dt_dev : A.dt;
        float w0 = A.w_eta[o], w1 = A.w_hu[o], w2 = A.w_hv[o];
        if (dry_clamped(A, A.w_clamp, R, C)) w0 = w1 = w2 = 0.0f;
        const float a = hi_flag(A, R, C), b = hi_flag(A, R + 1, C);
        const float cc = hi_flag(A, R, C + 1), d = hi_flag(A, R + 1, C + 1);
        const float Hreg = 0.25f * (((a + b) + cc) + d);
        const float hreg = c.E + Hreg;
        gE += w0;
        float ghuc = w1, ghvc = w2;
        if (hreg >= A.cutoff && c.E != kDryFlag) {
            const float H_x = 0.5f * (cc + d) - 0.5f * (a + b);
            const float H_y = 0.5f * (b + d) - 0.5f * (a + cc);
            ghvc += dt * w1 * A.f;
            gE += dt * w1 * A.g * H_x / A.dx;
            ghuc += -dt * w2 * A.f;
            gE += dt * w2 * A.g * H_y / A.dy;
        }
        if (hreg > A.cutoff) {
            gU += ghuc * hreg; gE += ghuc * c.U;
            gV += ghvc * hreg; gE += ghvc * c.V;
        }
        if (A.x_eta && !dry_clamped(A, A.x_clamp, R, C)) { xe = A.x_eta[o]; xu = A.x_hu[o]; xv = A.x_hv[o]; }
    }
    // conversion VJP
    float g_eta = 0, g_hu = 0, g_hv = 0;
    if (br == 3) {
        g_eta = gE - (gU * c.U + gV * c.V) / h;
        g_hu = gU / h;
        g_hv = gV / h;
    } else if (br == 2) {
        const float eps = A.desing;
        const float aq = h * h / (2.0f * eps) + 0.5f * eps;
        const float mn = fminf(aq, eps);
        const float den = fmaxf(mn, fabsf(h));
        float dden = 0.0f;
        if (fabsf(h) >= mn) dden = (h >= 0.0f) ? 1.0f : -1.0f;
        else if (aq <= eps) dden = h / eps;
        g_eta = gE - (gU * c.U + gV * c.V) * dden / den;
        g_hu = gU / den;
        g_hv = gV / den;
    }
    A.o_eta[o] = A.alpha * g_eta + A.gamma * xe;
    A.o_hu[o] = A.alpha * g_hu + A.gamma * xu;
    A.o_hv[o] = A.alpha * g_hv + A.gamma * xv;
}

__global__ __launch_bounds__(256) void k_adj_cell(AdjArgs A) {
    const int C = blockIdx.x * blockDim.x + threadIdx.x;
    const int R = blockIdx.y;
    if (C > A.nx + 3 || R > A.ny + 3) return;
    cell_adjoint(A, R, C);
}

// ghost adjoints back to the interior cells they are images of
__host__ __device__ inline void fold_ns(const AdjArgs& A, int C) {
    const int ny = A.ny;
    auto add = [&](int Rg, int Ri, float sv) {
        const size_t g = (size_t)Rg * A.pitch + C, i = (size_t)Ri * A.pitch + C;
        A.o_eta[i] += A.o_eta[g]; A.o_hu[i] += A.o_hu[g]; A.o_hv[i] += sv * A.o_hv[g];
        A.o_eta[g] = 0.0f; A.o_hu[g] = 0.0f; A.o_hv[g] = 0.0f;
    };
    if (A.bc_s == kWall) { add(1, 2, -1.0f); add(0, 3, -1.0f); }
    else if (A.bc_s == kPeriodic) { add(0, ny, 1.0f); add(1, ny + 1, 1.0f); }
    if (A.bc_n == kWall) { add(ny + 2, ny + 1, -1.0f); add(ny + 3, ny, -1.0f); }
    else if (A.bc_n == kPeriodic) { add(ny + 2, 2, 1.0f); add(ny + 3, 3, 1.0f); }
}

__host__ __device__ inline void fold_ew(const AdjArgs& A, int R) {
    const int nx = A.nx;
    auto add = [&](int Cg, int Ci, float su) {
        const size_t g = (size_t)R * A.pitch + Cg, i = (size_t)R * A.pitch + Ci;
        A.o_eta[i] += A.o_eta[g]; A.o_hu[i] += su * A.o_hu[g]; A.o_hv[i] += A.o_hv[g];
        A.o_eta[g] = 0.0f; A.o_hu[g] = 0.0f; A.o_hv[g] = 0.0f;
    };
    if (A.bc_w == kWall) { add(1, 2, -1.0f); add(0, 3, -1.0f); }
    else if (A.bc_w == kPeriodic) { add(0, nx, 1.0f); add(1, nx + 1, 1.0f); }
    if (A.bc_e == kWall) { add(nx + 2, nx + 1, -1.0f); add(nx + 3, nx, -1.0f); }
    else if (A.bc_e == kPeriodic) { add(nx + 2, 2, 1.0f); add(nx + 3, 3, 1.0f); }
}

__global__ void k_adj_fold_ns(AdjArgs A) {
    const int C = blockIdx.x * blockDim.x + threadIdx.x + 2;
    if (C <= A.nx + 1) fold_ns(A, C);
}

__global__ void k_adj_fold_ew(AdjArgs A) {
    const int R = blockIdx.x * blockDim.x + threadIdx.x + 2;
    if (R <= A.ny + 1) fold_ew(A, R);
}

static int adjoint_substep(const AdjArgs& A, hipStream_t s) {
    hipLaunchKernelGGL(k_adj_xface, dim3((A.nx + 1 + 255) / 256, A.ny), dim3(256), 0, s, A);
    hipLaunchKernelGGL(k_adj_yface, dim3((A.nx + 255) / 256, A.ny + 1), dim3(256), 0, s, A);
    hipLaunchKernelGGL(k_adj_cell, dim3((A.nx + 4 + 255) / 256, A.ny + 4), dim3(256), 0, s, A);
    hipLaunchKernelGGL(k_adj_fold_ns, dim3((A.nx + 255) / 256), dim3(256), 0, s, A);
    hipLaunchKernelGGL(k_adj_fold_ew, dim3((A.ny + 255) / 256), dim3(256), 0, s, A);
    return check_launch("adjoint substep launch");
}

static int adj_validate(const gocean_domain_t* d) {
    if (int e = validate_domain(d)) return e;
    if (!(d->north_x == 0.0f && d->north_y == 1.0f))
        return set_error(GOCEAN_EINVAL, "the adjoint supports zero grid rotation only (north = (0,1))");
    if (d->r != 0.0f) return set_error(GOCEAN_EINVAL, "the adjoint supports r == 0 only");
    if (d->rk_order != 2) return set_error(GOCEAN_EINVAL, "the adjoint supports rk_order 2 only");
    return 0;
}

static void adj_fill(const gocean_domain_t* d, AdjArgs& A) {
    std::memset(&A, 0, sizeof(A));
    A.nx = d->nx; A.ny = d->ny; A.pitch = d->pitch; A.hi_pitch = d->hi_pitch;
    A.plane = (long long)(d->ny + 4) * d->pitch;
    A.dx = d->dx; A.dy = d->dy; A.g = d->g; A.theta = d->theta; A.desing = d->desing_eps;
    A.flux_eps = d->flux_slope_eps; A.cutoff = d->depth_cutoff; A.land = d->land_value; A.f = d->f;
    A.bc_n = d->bc_north; A.bc_e = d->bc_east; A.bc_s = d->bc_south; A.bc_w = d->bc_west;
}

}  // namespace gocean

using namespace gocean;

extern "C" {

long long gocean_adjoint_scratch_floats(const gocean_domain_t* d) {
    if (!d) return -1;
    return 27LL * (long long)(d->ny + 4) * d->pitch;
}

int gocean_adjoint_substep(const gocean_domain_t* d, float dt, const float* dt_dev,
                           const float* eta, const float* hu, const float* hv,
                           const float* w_eta, const float* w_hu, const float* w_hv, const float* w_clamp,
                           const float* x_eta, const float* x_hu, const float* x_hv, const float* x_clamp,
                           float alpha, float gamma, float* o_eta, float* o_hu, float* o_hv,
                           const float* Hi, float* scratch, void* stream) {
    clear_error();
    if (int e = adj_validate(d)) return e;
    if (!eta || !hu || !hv || !w_eta || !w_hu || !w_hv || !o_eta || !o_hu || !o_hv || !Hi || !scratch)
        return set_error(GOCEAN_EINVAL, "NULL argument");
    if ((x_eta || x_hu || x_hv) && !(x_eta && x_hu && x_hv))
        return set_error(GOCEAN_EINVAL, "extra term needs all three fields");
    AdjArgs A;
    adj_fill(d, A);
    A.eta = eta; A.hu = hu; A.hv = hv;
    A.w_eta = w_eta; A.w_hu = w_hu; A.w_hv = w_hv;
    A.x_eta = x_eta; A.x_hu = x_hu; A.x_hv = x_hv;
    A.w_clamp = w_clamp; A.x_clamp = x_clamp;
    A.o_eta = o_eta; A.o_hu = o_hu; A.o_hv = o_hv;
    A.Hi = Hi; A.xc = scratch; A.yc = scratch + 12 * A.plane;
    A.dt = dt; A.dt_dev = dt_dev; A.alpha = alpha; A.gamma = gamma;
    return adjoint_substep(A, (hipStream_t)stream);
}

int gocean_adjoint_rk2_step(const gocean_domain_t* d, float dt, const float* dt_dev,
                            const float* eta_n, const float* hu_n, const float* hv_n,
                            const float* eta_s, const float* hu_s, const float* hv_s,
                            const float* eta_next, const float* l_eta, const float* l_hu, const float* l_hv,
                            float* o_eta, float* o_hu, float* o_hv,
                            const float* Hi, float* scratch, void* stream) {
    clear_error();
    if (int e = adj_validate(d)) return e;
    if (!eta_n || !hu_n || !hv_n || !eta_s || !hu_s || !hv_s || !l_eta || !l_hu || !l_hv ||
        !o_eta || !o_hu || !o_hv || !Hi || !scratch)
        return set_error(GOCEAN_EINVAL, "NULL argument");
    if (o_eta == l_eta || o_hu == l_hu || o_hv == l_hv)
        return set_error(GOCEAN_EINVAL, "output adjoint must not alias the input adjoint");
    const long long plane = (long long)(d->ny + 4) * d->pitch;
    float* ls = scratch + 24 * plane;          // lambda* (3 planes)
    // lambda* = 0.5 * J_P(Q*)^T lambda^{n+1}
    int e = gocean_adjoint_substep(d, dt, dt_dev, eta_s, hu_s, hv_s, l_eta, l_hu, l_hv, eta_next,
                                   nullptr, nullptr, nullptr, nullptr, 0.5f, 0.0f,
                                   ls, ls + plane, ls + 2 * plane, Hi, scratch, stream);
    if (e) return e;
    // lambda^n = J_P(Q^n)^T lambda* + 0.5 * lambda^{n+1}
    return gocean_adjoint_substep(d, dt, dt_dev, eta_n, hu_n, hv_n, ls, ls + plane, ls + 2 * plane, eta_s,
                                  l_eta, l_hu, l_hv, eta_next, 1.0f, 0.5f, o_eta, o_hu, o_hv, Hi, scratch,
                                  stream);
}

}  // extern "C"
