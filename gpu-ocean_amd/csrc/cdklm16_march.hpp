// cdklm16_march.hpp -- one CDKLM16 Runge-Kutta substep, column-marching form.
//
// Device code for gfx950, constant Coriolis parameter and zero grid rotation
// (north = (0, 1)), the configuration of every benchmark workload.  The
// arithmetic is the one of cdklm16_tile.hpp (the reference's operand order
// and rounding points of gpu_ocean/SWESimulators/gpu_kernels/
// CDKLM16_kernel.cu:80-1006, correctly rounded division and square root), so
// the two kernels give the same bits; what differs is how work is laid out:
//
//   * one wavefront owns a strip of 64 padded columns (lane = column) and
//     walks up BH output rows; lanes 2..61 write output, lanes 0, 1, 62, 63
//     are the strip's x halo;
//   * every per-row quantity is formed once per lane and kept in registers
//     while it is needed: the converted (eta, u, v) of the last three rows,
//     the y slopes of the last two, the corner depths of the last four, and
//     the flux through the lane's north face, which is the next row's south
//     face -- so the y direction needs no data exchange at all;
//   * the x neighbours (left and right lane) come through ds_bpermute, which
//     moves a register across lanes without LDS storage or barriers;
//   * boundary handling is per lane (the column's wall mirror / periodic
//     wrap, fixed for the whole strip) or per row (uniform over the wave), so
//     interior and boundary strips run the same code with no per-cell index
//     arithmetic; loads address one uniform row pointer plus the lane's
//     32-bit column offset;
//   * no LDS, no __syncthreads: each wave is independent and the occupancy
//     is set by vector registers alone.
//
// Work per output cell: the tile kernel converts (BH+4)(BW+4)-4*4 cells for
// BW*BH outputs and forms every x and y quantity on its halo; here every
// quantity is formed 64/60 times per output plus (BH+4)/BH for the y halo
// of the conversion only.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "cdklm16_tile.hpp"

namespace gocean {
namespace march {

constexpr int kLanes = 64;
constexpr int kOut = 60;          // output columns per strip (lanes 2..61)

// index of the wavefront within its workgroup, as a scalar: everything
// derived from it (strip, rows, row addresses) stays in scalar registers
__device__ __forceinline__ int wave_in_group() {
    return __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
}

// value of `v` in the lane whose byte address (lane * 4) is `addr`
__device__ __forceinline__ float lane_get(float v, int addr) {
    return __int_as_float(__builtin_amdgcn_ds_bpermute(addr, __float_as_int(v)));
}

// element at byte offset `off` of a row that starts at the wave-uniform
// address `row`: the load / store then takes the row from scalar registers
// and only the lane's 32-bit column offset from a vector register
__device__ __forceinline__ float ld(const float* row, uint32_t off) {
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(row) + off);
}
__device__ __forceinline__ void st(float* row, uint32_t off, float v) {
    *reinterpret_cast<float*>(reinterpret_cast<char*>(row) + off) = v;
}

// corner depth with land flagged (CDKLM16_kernel.cu:509-525)
__device__ __forceinline__ float flag_land(float h, const Consts& K) {
    return (fabsf(h - K.land) < kDryEps) ? kDryFlag : h;
}

// converted state of one cell: (eta, u, v) with dry / land handling, plus
// the cell depth as the update uses it (0.25 * the corner sum)
struct Cell {
    float E, U, V, Hm4;
};

// (:546-602, initBm :84-105).  a = Hi[R][C], b = Hi[R+1][C], c = Hi[R][C+1],
// d = Hi[R+1][C+1] (land flagged); ab = a + b is shared with H_west.
__device__ __forceinline__ Cell convert(float eta, float hu, float hv, float ab, float a, float b,
                                        float c, float d, const Consts& K) {
    Cell o;
    const float sum = (ab + c) + d;
    o.Hm4 = sum * 0.25f;
    float Hm = o.Hm4;
    const bool any_dry = (a == kDryFlag) | (b == kDryFlag) | (c == kDryFlag) | (d == kDryFlag);
    if (any_dry) {
        const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag)
                        - (float)(c == kDryFlag) - (float)(d == kDryFlag);
        Hm = (wet == 0.0f) ? K.land : sum / wet;
    }
    const float h = eta + Hm;
    if (fabsf(Hm - K.land) <= kDryEps) {
        o.E = kDryFlag; o.U = 0.0f; o.V = 0.0f;
    } else if (h < K.desing) {
        if (h <= K.cutoff) {
            o.E = -Hm + K.cutoff; o.U = 0.0f; o.V = 0.0f;
        } else {
            o.E = eta;
            o.U = desingularize(h, hu, K.desing);
            o.V = desingularize(h, hv, K.desing);
        }
    } else {
        o.E = eta;
        const float rh = rcp_refined(h);
        o.U = div_r(hu, h, rh);
        o.V = div_r(hv, h, rh);
    }
    return o;
}

// x or y slope centre of one cell (x: :607-666 with adjustSlopes_x :148-191;
// y: :711-770 with adjustSlopes_y :198-241), zero-rotation form: returns the
// minmod slopes of the two velocities and T = (K + d f v)/2g for the face
// reconstruction.  l, c, r: the cell and its two neighbours along the
// direction; `a*` the velocity along the direction's minmod only, `b*` the
// velocity entering the geostrophic term (v for x, u for y); Hlo / Hhi the
// face-averaged depths on the two sides; SGN +1 for y, -1 for x (the sign of
// the geostrophic term); fl* the wall flips of the three cells; fa the
// adjustment's own flip.
template <int SGN>
__device__ __forceinline__ void slope_centre(float el, float ec, float er,
                                             float al, float ac, float ar,
                                             float bl, float bc, float br,
                                             float Hlo, float Hhi,
                                             bool fll, bool flc, bool flr, bool fa,
                                             float dd, float vd, const Consts& K, float fc,
                                             float& Sa, float& Sb, float& T) {
    Sa = minmod_slope(al, ac, ar, K.theta);
    Sb = minmod_slope(bl, bc, br, K.theta);
    const float b_raw = bc;
    if (fll) bl = -bl;
    if (flc) bc = -bc;
    if (flr) br = -br;
    const float fl = bl * fc, fcc = bc * fc, fr = br * fc;
    float bw, ce, fw;
    if constexpr (SGN < 0) {
        bw = K.theta_g * ((ec - el) - vd * (fcc + fl));
        ce = K.half_g * ((er - el) - vd * ((fr + 2.0f * fcc) + fl));
        fw = K.theta_g * ((er - ec) - vd * (fcc + fr));
    } else {
        bw = K.theta_g * ((ec - el) + vd * (fcc + fl));
        ce = K.half_g * ((er - el) + vd * ((fr + 2.0f * fcc) + fl));
        fw = K.theta_g * ((er - ec) + vd * (fcc + fr));
    }
    float Kv = minmod_raw(bw, ce, fw);
    const float dfc = dd * fc;
    const float dfb = dfc * (fa ? -b_raw : b_raw);
    // x: t = (K + d f v)/2g; y: t = (L - d f u)/2g
    const float t = (SGN < 0) ? div_r(Kv + dfb, K.two_g, K.r_two_g) : div_r(Kv - dfb, K.two_g, K.r_two_g);
    const float hlo = (ec + Hlo) - t;
    const float hhi = (ec + Hhi) + t;
    bool adj = false;
    if constexpr (SGN < 0) {
        if (!(hlo > 0.0f)) { Kv = (-dfb) + K.two_g * (ec + Hlo); adj = true; }
        if (!(hhi > 0.0f)) { Kv = (-dfb) - K.two_g * (ec + Hhi); adj = true; }
    } else {
        if (!(hlo > 0.0f)) { Kv = dfb + K.two_g * (ec + Hlo); adj = true; }
        if (!(hhi > 0.0f)) { Kv = dfb - K.two_g * (ec + Hhi); adj = true; }
    }
    T = t;
    if (adj || flc != fa) {
        const float db = dfc * (flc ? -b_raw : b_raw);
        T = (SGN < 0) ? div_r(Kv + db, K.two_g, K.r_two_g) : div_r(Kv - db, K.two_g, K.r_two_g);
    }
}

// one face between cell m (low side) and p (high side) (:246-370):
// reconstructed states, face depths, central-upwind flux.  For x faces
// (a, b) = (u, v) and the result is (F.x, F.y, F.z); for y faces
// (a, b) = (v, u) and the flux comes back as (G.x, G.z, G.y) -- the caller
// swaps, as the reference's computeGFaceFlux does.
__device__ __forceinline__ void face(float em, float am, float bm, float Sam, float Sbm, float Tm,
                                     float ep, float ap, float bp, float Sap, float Sbp, float Tp,
                                     float Hf, const Consts& K, float& F0, float& F1, float& F2) {
    F0 = 0.0f; F1 = 0.0f; F2 = 0.0f;
    if (!(ep == kDryFlag || em == kDryFlag)) {
        const float Rpa = ap - 0.5f * Sap;
        const float Rpb = bp - 0.5f * Sbp;
        const float Rma = am + 0.5f * Sam;
        const float Rmb = bm + 0.5f * Sbm;
        const float hp = fmaxf(0.0f, (ep + Hf) - Tp);
        const float hm = fmaxf(0.0f, (em + Hf) + Tm);
        cu_flux(hm, Rma, Rmb, hp, Rpa, Rpb, K, F0, F1, F2);
    }
}

// One strip: padded columns sx*60 .. sx*60+63 (lanes), output rows
// 2 + sy*BH .. min(2 + (sy+1)*BH, ny+2) - 1.
template <int BH>
__device__ __forceinline__ void march_strip(const SubstepArgs& A, int sx, int sy) {
    const int lane = threadIdx.x & (kLanes - 1);
    const int nx = A.nx, ny = A.ny, pitch = A.pitch, hp = A.hi_pitch;
    const Consts K = make_consts(A);
    const int C = sx * kOut + lane;                  // padded column of the lane
    const int Cl = min(C, nx + 3);                   // load column (clamped)
    const int j0 = 2 + sy * BH;
    const int j1 = min(j0 + BH, ny + 2);
    const bool wN = A.bc_n() == kWall, wS = A.bc_s() == kWall;
    const bool wE = A.bc_e() == kWall, wW = A.bc_w() == kWall;
    const bool pNS = A.bc_n() == kPeriodic, pEW = A.bc_e() == kPeriodic;

    // the lane's source column (wall mirror with u negated, or periodic wrap)
    int sc = Cl;
    bool neg_u = false;
    if (Cl < 2) {
        if (wW) { sc = 3 - Cl; neg_u = true; } else if (pEW) { sc = Cl + nx; }
    } else if (Cl > nx + 1) {
        if (wE) { sc = 2 * nx + 3 - Cl; neg_u = true; } else if (pEW) { sc = Cl - nx; }
    }
    auto col_flip = [&](int c) { return (wW && c < 2) || (wE && c > nx + 1); };
    const bool xfl = col_flip(C - 1), xfc = col_flip(C), xfr = col_flip(C + 1);
    const bool xfa = wW && C < 2;
    auto row_flip = [&](int r) { return (wS && r < 2) || (wN && r > ny + 1); };
    const int hc = min(C, nx + 4), hn = min(C + 1, nx + 4);   // corner-depth columns
    // byte offsets of the lane's columns within a row
    const uint32_t osc = (uint32_t)sc * 4u, ohc = (uint32_t)hc * 4u, ohn = (uint32_t)hn * 4u;
    const uint32_t ocl = (uint32_t)Cl * 4u;
    const int addr_r = ((lane + 1) & (kLanes - 1)) << 2;
    const int addr_l = ((lane - 1) & (kLanes - 1)) << 2;

    const float fc = A.f;
    const float dt = A.dt_dev ? *A.dt_dev : A.dt;

    // state row R of the input (remapped) and the corner-depth row R
    auto load_state = [&](int R, float& e, float& u, float& v) {
        int sr = R;
        bool neg_v = false;
        if (R < 2) {
            if (wS) { sr = 3 - R; neg_v = true; } else if (pNS) { sr = R + ny; }
        } else if (R > ny + 1) {
            if (wN) { sr = 2 * ny + 3 - R; neg_v = true; } else if (pNS) { sr = R - ny; }
        }
        const size_t ro = (size_t)sr * pitch;
        e = ld(A.eta_in + ro, osc);
        u = ld(A.hu_in + ro, osc);
        v = ld(A.hv_in + ro, osc);
        if (neg_u) u = -u;
        if (neg_v) v = -v;
    };
    auto load_hi = [&](int R, float& a, float& b) {
        const size_t ro = (size_t)R * hp;
        a = flag_land(ld(A.Hi + ro, ohc), K);
        b = flag_land(ld(A.Hi + ro, ohn), K);
    };

    // ---- prologue: conversion of rows j0-2 .. j0+1, y slopes of rows
    //      j0-1 and j0, the flux through the face (j0-1 | j0) ----
    float hcA, hnA, hcB, hnB, hcC, hnC, hcD, hnD;      // corner rows j-2 .. j+1 of the window
    load_hi(j0 - 2, hcA, hnA);
    load_hi(j0 - 1, hcB, hnB);
    load_hi(j0, hcC, hnC);
    load_hi(j0 + 1, hcD, hnD);
    Cell c0, c1, c2, c3;                             // rows j0-2, j0-1, j0, j0+1
    {
        float e, u, v, hcE, hnE;
        load_state(j0 - 2, e, u, v);
        c0 = convert(e, u, v, hcA + hcB, hcA, hcB, hnA, hnB, K);
        load_state(j0 - 1, e, u, v);
        c1 = convert(e, u, v, hcB + hcC, hcB, hcC, hnB, hnC, K);
        load_state(j0, e, u, v);
        c2 = convert(e, u, v, hcC + hcD, hcC, hcD, hnC, hnD, K);
        load_hi(j0 + 2, hcE, hnE);
        load_state(j0 + 1, e, u, v);
        c3 = convert(e, u, v, hcD + hcE, hcD, hcE, hnD, hnE, K);
        // slide the corner window to rows j0-1 .. j0+2
        hcA = hcB; hnA = hnB; hcB = hcC; hnB = hnC; hcC = hcD; hnC = hnD; hcD = hcE; hnD = hnE;
    }
    // corner rows now: A = j0-1, B = j0, C = j0+1, D = j0+2
    // face-averaged depths of the south faces: Hs(R) = 0.5 (Hi[R][C] + Hi[R][C+1])
    float Hs_m1 = 0.5f * (hcA + hnA);    // row j0-1's south face
    float Hs_0 = 0.5f * (hcB + hnB);     // row j0's south face (= row j0-1's north)
    float Hs_1 = 0.5f * (hcC + hnC);     // row j0+1's south face
    // y slopes of row j0-1 (cells c0, c1, c2) and row j0 (c1, c2, c3)
    float Uy0, Vy0, S0, Uy1, Vy1, S1;
    slope_centre<+1>(c0.E, c1.E, c2.E, c0.V, c1.V, c2.V, c0.U, c1.U, c2.U, Hs_m1, Hs_0,
                     row_flip(j0 - 2), row_flip(j0 - 1), row_flip(j0), wS && j0 - 1 < 2,
                     K.dy, K.udy, K, fc, Vy0, Uy0, S0);
    slope_centre<+1>(c1.E, c2.E, c3.E, c1.V, c2.V, c3.V, c1.U, c2.U, c3.U, Hs_0, Hs_1,
                     row_flip(j0 - 1), row_flip(j0), row_flip(j0 + 1), wS && j0 < 2,
                     K.dy, K.udy, K, fc, Vy1, Uy1, S1);
    // flux through (j0-1 | j0): m = row j0-1 (c1), p = row j0 (c2)
    float gsx, gsy, gsz;
    {
        float a, b, c;
        face(c1.E, c1.V, c1.U, Vy0, Uy0, S0, c2.E, c2.V, c2.U, Vy1, Uy1, S1, Hs_0, K, a, b, c);
        gsx = a; gsy = c; gsz = b;
    }
    // window: (c2, c3) = rows j, j+1; (Uy1, Vy1, S1) = y slopes of row j;
    // corner rows B = j, C = j+1, D = j+2; Hs_0 = south face of row j,
    // Hs_1 = its north face
    Cell cj = c2, cj1 = c3;
    float Uyj = Uy1, Vyj = Vy1, Sj = S1;
    float hcj = hcB, hnj = hnB, hcj1 = hcC, hnj1 = hnC, hcj2 = hcD, hnj2 = hnD;
    float Hsj = Hs_0, Hnj = Hs_1;
    const bool friction = A.r > 0.0f;
    // gridded wind: the lane's column weights, fixed for the strip
    const bool wgrid = A.wind_grid != nullptr;
    WindAxis wax;
    wax.i0 = wax.i1 = 0; wax.a = 0.0f;
    if (wgrid) wax = wind_axis(C, nx, A.wind_wx);
    const int rk_order = A.rk_order(), rk_step = A.rk_step();
    const bool store_lane = lane >= 2 && lane < 2 + kOut && C <= nx + 1;

    for (int j = j0; j < j1; ++j) {
        // loads for row j+2 (conversion) and Q^n of row j (second stages)
        float e2, u2, v2, hcj3, hnj3;
        load_state(j + 2, e2, u2, v2);
        load_hi(j + 3 <= ny + 4 ? j + 3 : ny + 4, hcj3, hnj3);
        const size_t rj = (size_t)j * pitch;
        float ea = 0.0f, ha = 0.0f, va = 0.0f;
        if (rk_step > 0) { ea = ld(A.eta_qn + rj, ocl); ha = ld(A.hu_qn + rj, ocl); va = ld(A.hv_qn + rj, ocl); }

        // ---- x slope centre of row j (neighbours by lane exchange) ----
        const float el = lane_get(cj.E, addr_l), er = lane_get(cj.E, addr_r);
        const float ul = lane_get(cj.U, addr_l), ur = lane_get(cj.U, addr_r);
        const float vl = lane_get(cj.V, addr_l), vr = lane_get(cj.V, addr_r);
        const float abj = hcj + hcj1;
        const float Hw = 0.5f * abj;                     // west face depth of the cell
        const float He = 0.5f * (hnj + hnj1);            // east face depth
        float Ux, Vx, T;
        slope_centre<-1>(el, cj.E, er, ul, cj.U, ur, vl, cj.V, vr, Hw, He, xfl, xfc, xfr, xfa,
                         K.dx, K.vdx, K, fc, Ux, Vx, T);
        // ---- east face (C | C+1) ----
        const float Uxr = lane_get(Ux, addr_r), Vxr = lane_get(Vx, addr_r), Tr = lane_get(T, addr_r);
        float fex, fey, fez;
        face(cj.E, cj.U, cj.V, Ux, Vx, T, er, ur, vr, Uxr, Vxr, Tr, He, K, fex, fey, fez);

        // ---- conversion of row j+2, y slopes of row j+1, north face ----
        const Cell cj2 = convert(e2, u2, v2, hcj2 + hcj3, hcj2, hcj3, hnj2, hnj3, K);
        const float Hnj1 = 0.5f * (hcj2 + hnj2);         // north face of row j+1
        float Uyn, Vyn, Sn;
        slope_centre<+1>(cj.E, cj1.E, cj2.E, cj.V, cj1.V, cj2.V, cj.U, cj1.U, cj2.U, Hnj, Hnj1,
                         row_flip(j), row_flip(j + 1), row_flip(j + 2), wS && j + 1 < 2,
                         K.dy, K.udy, K, fc, Vyn, Uyn, Sn);
        float gnx, gny, gnz;
        {
            float a, b, c;
            face(cj.E, cj.V, cj.U, Vyj, Uyj, Sj, cj1.E, cj1.V, cj1.U, Vyn, Uyn, Sn, Hnj, K, a, b, c);
            gnx = a; gny = c; gnz = b;
        }
        // ---- west face from the left lane ----
        const float fwx = lane_get(fex, addr_l), fwy = lane_get(fey, addr_l), fwz = lane_get(fez, addr_l);

        // ---- sources, RK update, store (:802-1004) ----
        if (store_lane) {
            const float Hm = cj.Hm4;
            const float ec = cj.E, uc = cj.U, vc = cj.V;
            const float hreg = ec + Hm;
            float hu = 0.0f, hv = 0.0f;
            if (hreg > K.cutoff) {
                hu = uc * hreg;
                hv = vc * hreg;
            }
            const float fd0 = div_r(fex - fwx, K.dx, K.r_dx) + div_r(gnx - gsx, K.dy, K.r_dy);
            const float fd1 = div_r(fey - fwy, K.dx, K.r_dx) + div_r(gny - gsy, K.dy, K.r_dy);
            const float fd2 = div_r(fez - fwz, K.dx, K.r_dx) + div_r(gnz - gsz, K.dy, K.r_dy);
            const float eta_west = ec - T, eta_east = ec + T;
            const float eta_south = ec - Sj, eta_north = ec + Sj;
            float st1 = 0.0f, st2 = 0.0f;
            if (hreg >= K.cutoff && ec != kDryFlag) {
                float X = A.wind_x, Y = A.wind_y;
                if (wgrid)
                    wind_grid_at(A.wind_grid, A.wind_wx, A.wind_wy, A.wind_t, wax, wind_axis(j, ny, A.wind_wy),
                                 X, Y);
                const float H_x = He - Hw;
                const float H_y = Hnj - Hsj;
                const float eta_sn = 0.5f * (eta_north + eta_south);
                const float eta_we = 0.5f * (eta_west + eta_east);
                const float b1 = (K.g * (eta_we + Hm)) * H_x;
                const float b2 = (K.g * (eta_sn + Hm)) * H_y;
                const float hu_cor = fc * hv;
                const float hv_cor = (-fc) * hu;
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
                    const float uu = desingularize(h, hu, K.desing);
                    const float vv = desingularize(h, hv, K.desing);
                    Cf = (dt * A.r) * desingularize(h, sqrtf(uu * uu + vv * vv), K.desing);
                } else {
                    const float uu = hu / h, vv = hv / h;
                    Cf = ((dt * A.r) * sqrtf(uu * uu + vv * vv)) / h;
                }
            }
            float ue, uu, uv;
            if (rk_order < 3) {
                if (rk_step == 0) {
                    ue = ec + dt * L1;
                    uu = hu + dt * L2;
                    uv = hv + dt * L3;
                    if (friction) {
                        uu = uu / (1.0f + Cf);
                        uv = uv / (1.0f + Cf);
                    }
                } else {
                    ue = 0.5f * (ea + (ec + dt * L1));
                    uu = 0.5f * (ha + (hu + dt * L2));
                    uv = 0.5f * (va + (hv + dt * L3));
                    if (friction) {
                        uu = uu / (1.0f + 0.5f * Cf);
                        uv = uv / (1.0f + 0.5f * Cf);
                    }
                }
            } else {
                if (rk_step == 0) {
                    ue = ec + dt * L1;
                    uu = hu + dt * L2;
                    uv = hv + dt * L3;
                } else if (rk_step == 1) {
                    ue = 0.75f * ea + 0.25f * (ec + dt * L1);
                    uu = 0.75f * ha + 0.25f * (hu + dt * L2);
                    uv = 0.75f * va + 0.25f * (hv + dt * L3);
                } else {
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
            st(A.eta_out + rj, ocl, ue);
            st(A.hu_out + rj, ocl, uu);
            st(A.hv_out + rj, ocl, uv);
            // periodic ghost images of the output (cells within two rows /
            // columns of a periodic side)
            if ((pNS && (j <= 3 || j >= ny)) || (pEW && (C <= 3 || C >= nx))) {
                int rows[3] = {j, -1, -1}, cols[3] = {C, -1, -1};
                int nr = 1, nc = 1;
                if (pNS) {
                    if (j + ny <= ny + 3) rows[nr++] = j + ny;
                    if (j - ny >= 0) rows[nr++] = j - ny;
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

        // ---- slide the window one row north ----
        gsx = gnx; gsy = gny; gsz = gnz;
        cj = cj1; cj1 = cj2;
        Uyj = Uyn; Vyj = Vyn; Sj = Sn;
        hcj = hcj1; hnj = hnj1; hcj1 = hcj2; hnj1 = hnj2; hcj2 = hcj3; hnj2 = hnj3;
        Hsj = Hnj; Hnj = Hnj1;
    }
}

}  // namespace march
}  // namespace gocean
