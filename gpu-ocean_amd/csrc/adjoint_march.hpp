// adjoint_march.hpp -- the substep adjoint of adjoint_tile.hpp in
// column-marching form.
//
// Same inputs as the tile adjoint (the substep's eta, its branch record,
// lambda and the extra term), same arithmetic and the same order of every
// sum, so the two give the same bits; what differs is the layout of the
// work:
//
//   * one wavefront owns a strip of 64 padded columns (lane L is column
//     58*sx + L - 1) and walks up the rows of its row strip; a cell's
//     gradient depends on cells up to three columns away (face -> slope
//     centre -> cell, each reaching one or two cells), so lanes 3..60 are
//     complete and written, plus the ghost columns at the domain's edges
//     (which depend on nothing beyond the domain);
//   * x neighbours come through ds_bpermute; the y direction is a pipeline
//     in registers: iteration k loads row k+2, differentiates the y face
//     (k | k+1) and the y slope centre k, and completes row k-1, whose x
//     faces and centres it differentiates from the row data it holds;
//   * loads and stores are a scalar plane address plus the row's byte offset
//     and the lane's 32-bit column offset; the three fields of lambda, of
//     the extra term and of the output, and the five planes of the record,
//     are each one base pointer and a plane stride;
//   * no LDS, no barriers, no per-item index arithmetic.
//
// The per-face and per-centre vector-Jacobian products are the functions of
// adjoint_tile.hpp (flux_vjp, slope_val, slope_coef, depth_term_vjp).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "adjoint_tile.hpp"
#include "cdklm16_march.hpp"

namespace gocean {
namespace adjm {

constexpr int kLanes = 64;
constexpr int kOut = 58;          // columns advanced per strip (lanes 3..60)

using march::lane_get;

// Launch arguments.  Planes of one array are `*_stride` bytes apart; the
// record holds u, v, T, S and the code word in that order.
struct MArgs {
    const float* eta;
    const char* tape;
    const char* w;
    const float* w_clamp;
    const char* x;                 // NULL: no extra term
    const float* x_clamp;
    char* o;
    const float* Hi;
    const float* dt_dev;
    uint32_t tape_stride, w_stride, x_stride, o_stride;   // bytes; every offset stays below 4 GiB
    float dt, alpha, gamma;
    int w_fold, x_fold;
    int nx, ny, pitch, hi_pitch;
    float g, theta, desing, cutoff, land, f;
    int bc_n, bc_e, bc_s, bc_w;
    float r_dx, r_dy, r_two_g, Dx, Dy, Ddx, Ddy, g_dx, g_dy;
};

// base + a 32-bit byte offset (plane + row + column): one scalar base per
// array and the offset in a vector register, the global_load form with a
// scalar base
__device__ __forceinline__ float ldb(const void* base, uint32_t plane_off, uint32_t row_off, uint32_t lane_off) {
    const uint32_t off = lane_off + (row_off + plane_off);
    return *reinterpret_cast<const float*>(reinterpret_cast<const char*>(base) + off);
}
__device__ __forceinline__ void stb(void* base, uint32_t plane_off, uint32_t row_off, uint32_t lane_off, float v) {
    const uint32_t off = lane_off + (row_off + plane_off);
    *reinterpret_cast<float*>(reinterpret_cast<char*>(base) + off) = v;
}

// one padded row as a lane holds it
struct Row {
    float E, U, V;
    uint32_t code;        // record bits of the cell (masked as stage_cells) | branch << 28
    float T, S;           // x / y depth terms of the record (0 where none)
    float W0, W1, W2;     // lambda of an interior cell, clamp-masked and folded
    float hc, hn;         // corner depths Hi[R][C], Hi[R][C+1], land flagged
};

struct V3 {
    float a, b, c;
};

__device__ __forceinline__ V3 v3z() { return V3{0.0f, 0.0f, 0.0f}; }
__device__ __forceinline__ V3 v3get(const V3& v, int addr) {
    return V3{lane_get(v.a, addr), lane_get(v.b, addr), lane_get(v.c, addr)};
}

// vector-Jacobian product of one face: direct terms of the minus / plus
// cell (dm, dp: d/d(E, u, v)) and the terms of its two slope centres (cm, cp:
// d/d(T, u slope, v slope)), as out[0..11] of stage_xface / stage_yface
struct FaceOut {
    V3 dm, dp, cm, cp;
};

// centre out[0..8]: d/d(E, u, v) of cells -1 (lo), 0 (mid), +1 (hi)
struct CentreOut {
    V3 lo, mid, hi;
};

// the loaded words of one row, before the conversion branch and the clamp
// mask (fetched one iteration ahead of their use)
struct Raw {
    float eta, U, V, T, S, W0, W1, W2, wc;
    uint32_t tc;
};

// pipeline state carried into iteration k (see adj_march_strip): row k-1's
// y sum is ((face (k-2|k-1) plus + face (k-1|k) minus) + (centre k-2 hi +
// centre k-1 mid)) + centre k lo, and the two inner sums are formed in
// iteration k-1
struct Carry {
    V3 p1, p2;                              // row k-1's inner sums
    V3 dp, cp, c6;                          // face (k-1|k): plus direct, plus centre terms; centre k-1 hi
    float Suy, Svy;                         // y slopes of row k
    float phc, phn;                         // corners of row k+2 (flagged)
    Raw raw;                                // row k+2 as loaded
    float qhc, qhn;                         // corners of row k+3 as loaded
};

struct Strip {
    const MArgs& A;
    int lane, C;
    bool col_ok, ci;
    uint32_t oc, osc, ohc, ohn;             // byte offsets: column, image column, corner columns
    uint32_t rowb, hrowb;                   // row bytes of the state / Hi planes
    float dt;
    int addr_l, addr_r;

    __device__ Strip(const MArgs& a, int sx) : A(a) {
        lane = threadIdx.x & (kLanes - 1);
        C = sx * kOut + lane - 1;
        col_ok = C >= 0 && C <= A.nx + 3;
        ci = C >= 2 && C <= A.nx + 1;
        int sc = C;
        if (C < 2) sc = (A.bc_w == kWall) ? 3 - C : C + A.nx;
        else if (C > A.nx + 1) sc = (A.bc_e == kWall) ? 2 * A.nx + 3 - C : C - A.nx;
        const int cl = min(max(C, 0), A.nx + 3);          // loads of out-of-range lanes stay in the row
        oc = (uint32_t)cl * 4u;
        osc = (uint32_t)min(max(sc, 0), A.nx + 3) * 4u;
        ohc = (uint32_t)min(max(C, 0), A.nx + 4) * 4u;
        ohn = (uint32_t)min(max(C + 1, 0), A.nx + 4) * 4u;
        rowb = (uint32_t)A.pitch * 4u;
        hrowb = (uint32_t)A.hi_pitch * 4u;
        dt = A.dt_dev ? *A.dt_dev : A.dt;
        addr_l = ((lane - 1) & (kLanes - 1)) << 2;
        addr_r = ((lane + 1) & (kLanes - 1)) << 2;
    }

    __device__ bool ghost_r(int R) const { return R < 2 || R > A.ny + 1; }
    __device__ float hflag(float v) const { return (fabsf(v - A.land) < kDryEps) ? kDryFlag : v; }
    // corner depths Hi[R][C], Hi[R][C+1] with land flagged (0 outside, as
    // stage_hi leaves them)
    __device__ void corners(int R, float& hc, float& hn) const {
        float a, b;
        corners_raw(R, a, b);
        corners_flag(R, a, b, hc, hn);
    }
    __device__ void corners_raw(int R, float& a, float& b) const {
        a = 0.0f; b = 0.0f;
        if (R < 0 || R > A.ny + 4) return;
        const uint32_t ro = (uint32_t)R * hrowb;
        a = ldb(A.Hi, 0, ro, ohc);
        b = ldb(A.Hi, 0, ro, ohn);
    }
    __device__ void corners_flag(int R, float a, float b, float& hc, float& hn) const {
        hc = 0.0f; hn = 0.0f;
        if (R < 0 || R > A.ny + 4) return;
        if (C >= 0 && C <= A.nx + 4) hc = hflag(a);
        if (C + 1 >= 0 && C + 1 <= A.nx + 4) hn = hflag(b);
    }
    __device__ float col_flip(int c) const {
        return ((A.bc_w == kWall && c < 2) || (A.bc_e == kWall && c > A.nx + 1)) ? -1.0f : 1.0f;
    }
    __device__ float row_flip(int r) const {
        return ((A.bc_s == kWall && r < 2) || (A.bc_n == kWall && r > A.ny + 1)) ? -1.0f : 1.0f;
    }
    __device__ bool near_boundary(int R) const { return R <= 3 || R >= A.ny || C <= 3 || C >= A.nx; }

    // field `field` of a 3-plane array at (R, C) plus the ghost entries whose
    // image it is (AdjTile::fold_read)
    __device__ float fold_read(const char* base, uint32_t stride, int field, int R) const {
        const int nx = A.nx, ny = A.ny;
        const uint32_t po = (uint32_t)field * stride;
        const uint32_t rr = (uint32_t)R * rowb;
        float s = ldb(base, po, rr, oc);
        if (R <= 3) {
            if (A.bc_s == kWall) { const float g = ldb(base, po, (uint32_t)(3 - R) * rowb, oc); s += field == 2 ? -g : g; }
            else s += ldb(base, po, (uint32_t)(R + ny) * rowb, oc);
        }
        if (R >= ny) {
            if (A.bc_n == kWall) {
                const float g = ldb(base, po, (uint32_t)(2 * ny + 3 - R) * rowb, oc);
                s += field == 2 ? -g : g;
            } else s += ldb(base, po, (uint32_t)(R - ny) * rowb, oc);
        }
        if (C <= 3) {
            if (A.bc_w == kWall) { const float g = ldb(base, po, rr, (uint32_t)(3 - C) * 4u); s += field == 1 ? -g : g; }
            else s += ldb(base, po, rr, (uint32_t)(C + nx) * 4u);
        }
        if (C >= nx) {
            if (A.bc_e == kWall) {
                const float g = ldb(base, po, rr, (uint32_t)(2 * nx + 3 - C) * 4u);
                s += field == 1 ? -g : g;
            } else s += ldb(base, po, rr, (uint32_t)(C - nx) * 4u);
        }
        return s;
    }

    // the loads of row R of the cell data (stage_cells)
    __device__ void fetch(int R, Raw& q) const {
        q.eta = 0.0f; q.U = 0.0f; q.V = 0.0f; q.T = 0.0f; q.S = 0.0f;
        q.W0 = 0.0f; q.W1 = 0.0f; q.W2 = 0.0f; q.wc = 0.0f; q.tc = 0u;
        const int nx = A.nx, ny = A.ny;
        if (R < 0 || R > ny + 3) return;                 // uniform
        const bool ri = !ghost_r(R);
        if (!col_ok || (!ri && !ci)) return;            // outside, or a ghost corner (never read)
        int sr = R;
        if (R < 2) sr = (A.bc_s == kWall) ? 3 - R : R + ny;
        else if (R > ny + 1) sr = (A.bc_n == kWall) ? 2 * ny + 3 - R : R - ny;
        const uint32_t ro = (uint32_t)R * rowb;
        q.eta = ldb(A.eta, 0, (uint32_t)sr * rowb, osc);
        const uint32_t ts = A.tape_stride;
        q.U = ldb(A.tape, 0, ro, oc);
        q.V = ldb(A.tape, ts, ro, oc);
        const bool xrec = ri && C >= 1 && C <= nx + 2, yrec = ci && R >= 1 && R <= ny + 2;
        if (xrec || yrec) q.tc = __float_as_uint(ldb(A.tape, 4 * ts, ro, oc));
        if (xrec) q.T = ldb(A.tape, 2 * ts, ro, oc);
        if (yrec) q.S = ldb(A.tape, 3 * ts, ro, oc);
        if (ri && ci) {
            if (A.w_clamp) q.wc = ldb(A.w_clamp, 0, ro, oc);
            const uint32_t ws = A.w_stride;
            q.W0 = ldb(A.w, 0, ro, oc); q.W1 = ldb(A.w, ws, ro, oc); q.W2 = ldb(A.w, 2 * ws, ro, oc);
        }
    }

    // row R from its loaded words: conversion branch, code word, clamp mask
    // and fold of lambda; a, c = corners of row R (its own hc, hn), b, d =
    // corners of row R+1
    __device__ void finish(int R, const Raw& q, float a, float c, float b, float d, Row& o) const {
        o.E = 0.0f; o.U = 0.0f; o.V = 0.0f; o.code = 0u; o.T = 0.0f; o.S = 0.0f;
        o.W0 = 0.0f; o.W1 = 0.0f; o.W2 = 0.0f;
        o.hc = a; o.hn = c;
        const int nx = A.nx, ny = A.ny;
        if (R < 0 || R > ny + 3) return;                 // uniform
        const bool ri = !ghost_r(R);
        if (!col_ok || (!ri && !ci)) return;
        const float eta = q.eta;
        o.U = q.U;
        o.V = q.V;
        const bool xrec = ri && C >= 1 && C <= nx + 2, yrec = ci && R >= 1 && R <= ny + 2;
        o.T = q.T;
        o.S = q.S;
        // conversion branch (cell_hm: the forward's h = eta + Hm)
        const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag) - (float)(c == kDryFlag)
                          - (float)(d == kDryFlag);
        const float sum = ((a + b) + c) + d;
        const float Hm = (wet == 4.0f) ? sum * 0.25f : (wet == 0.0f ? A.land : sum / wet);
        const float h = eta + Hm;
        uint32_t br;
        if (fabsf(Hm - A.land) <= kDryEps) { o.E = kDryFlag; br = kBrLand; }
        else if (h < A.desing) {
            if (h <= A.cutoff) { o.E = -Hm + A.cutoff; br = kBrDry; }
            else { o.E = eta; br = kBrDesing; }
        } else { o.E = eta; br = kBrWet; }
        uint32_t code = br << 28;
        if (xrec) code |= q.tc & 0x001F00FFu;
        if (yrec) code |= q.tc & 0x03E0FF00u;
        o.code = code;
        if (ri && ci) {
            bool clamped = false;
            if (A.w_clamp) {
                const float Hreg = 0.25f * (((a + b) + c) + d);
                clamped = q.wc <= -Hreg + A.cutoff;
            }
            if (!clamped) {
                const uint32_t ws = A.w_stride;
                if (A.w_fold && near_boundary(R)) {
                    o.W0 = fold_read(A.w, ws, 0, R); o.W1 = fold_read(A.w, ws, 1, R); o.W2 = fold_read(A.w, ws, 2, R);
                } else {
                    o.W0 = q.W0; o.W1 = q.W1; o.W2 = q.W2;
                }
            }
        }
    }

    __device__ void load(int R, float a, float c, float b, float d, Row& o) const {
        Raw q;
        fetch(R, q);
        finish(R, q, a, c, b, d, o);
    }

    // ---- x direction of row R (interior rows): the gathered d/d(E, u, v)
    //      of the lane's cell (stage_xgather's sum); hnR1 = Hi[R+1][C+1] ----
    __device__ V3 x_row(const Row& r, float hnR1) const {
        const float th = A.theta;
        // own slopes (the face reads its two cells' slopes)
        const float Ul = lane_get(r.U, addr_l), Ur = lane_get(r.U, addr_r);
        const float Vl = lane_get(r.V, addr_l), Vr = lane_get(r.V, addr_r);
        const float Su = slope_val(r.code & 3u, Ul, r.U, Ur, th);
        const float Sv = slope_val((r.code >> 2) & 3u, Vl, r.V, Vr, th);
        // the east face (C | C+1): plus cell from the right lane
        const float Ep = lane_get(r.E, addr_r);
        const float Sup = lane_get(Su, addr_r), Svp = lane_get(Sv, addr_r);
        const float tp = lane_get(r.T, addr_r);
        const float W0p = lane_get(r.W0, addr_r), W1p = lane_get(r.W1, addr_r), W2p = lane_get(r.W2, addr_r);
        FaceOut F;
        F.dm = v3z(); F.dp = v3z(); F.cm = v3z(); F.cp = v3z();
        {
            const bool valid = C >= 1 && C <= A.nx + 1 && !(r.E == kDryFlag || Ep == kDryFlag);
            const float rdx = dt * A.r_dx;
            const float l0 = (W0p - r.W0) * rdx;
            const float l1 = (W1p - r.W1) * rdx;
            const float l2 = (W2p - r.W2) * rdx;
            const float Rpx = Ur - 0.5f * Sup, Rpy = Vr - 0.5f * Svp;
            const float Rmx = r.U + 0.5f * Su, Rmy = r.V + 0.5f * Sv;
            const float Hf = 0.5f * (r.hn + hnR1);
            const float hp_raw = (Ep + Hf) - tp;
            const float hm_raw = (r.E + Hf) + r.T;
            const float hp = fmaxf(0.0f, hp_raw), hm = fmaxf(0.0f, hm_raw);
            const uint32_t fcode = (r.code >> kCodeFx) & 31u;
            float ghm, gnm, gtm, ghp, gnp, gtp;
            bool done = false;
            if (__all(valid && hp > A.cutoff && hm > A.cutoff && !(fcode & 16u))) {
                flux_vjp<true>(hm, Rmx, Rmy, hp, Rpx, Rpy, fcode, A.g, A.cutoff, l0, l1, l2, ghm, gnm, gtm, ghp,
                               gnp, gtp);
                done = true;
            } else if (valid) {
                flux_vjp(hm, Rmx, Rmy, hp, Rpx, Rpy, fcode, A.g, A.cutoff, l0, l1, l2, ghm, gnm, gtm, ghp, gnp,
                         gtp);
                done = true;
            }
            if (done) {
                const float ghm_r = hm_raw > 0.0f ? ghm : 0.0f;
                const float ghp_r = hp_raw > 0.0f ? ghp : 0.0f;
                F.dm = V3{ghm_r, gnm, gtm};
                F.dp = V3{ghp_r, gnp, gtp};
                F.cm = V3{ghm_r, 0.5f * gnm, 0.5f * gtm};
                F.cp = V3{-ghp_r, -0.5f * gnp, -0.5f * gtp};
            }
        }
        // slope centre C: plus side of the west face, minus side of the east one
        const V3 wcp = v3get(F.cp, addr_l);
        CentreOut Q;
        Q.lo = v3z(); Q.mid = v3z(); Q.hi = v3z();
        if (C >= 1 && C <= A.nx + 2) {
            const float gT = wcp.a + F.cm.a;
            const float gSu = wcp.b + F.cm.b;
            const float gSv = wcp.c + F.cm.c;
            float out[9];
            for (int q = 0; q < 9; ++q) out[q] = 0.0f;
            float cl, cc, cr;
            slope_coef(r.code & 3u, th, cl, cc, cr);
            out[1] = cl * gSu; out[4] = cc * gSu; out[7] = cr * gSu;
            slope_coef((r.code >> 2) & 3u, th, cl, cc, cr);
            out[2] = cl * gSv; out[5] = cc * gSv; out[8] = cr * gSv;
            const float sa = (A.bc_w == kWall && C < 2) ? -1.0f : 1.0f;
            float gE[3] = {0.0f, 0.0f, 0.0f}, gt[3] = {0.0f, 0.0f, 0.0f};
            const uint32_t ks = (r.code >> 4) & 3u, adj = (r.code >> 6) & 3u;
            if (__all(adj == 0u))
                depth_term_vjp<true>(ks, adj, -1.0f, th, A.g, A.f, A.Ddx, A.Dx, A.r_two_g, col_flip(C - 1),
                                     col_flip(C), col_flip(C + 1), sa, gT, gE, gt);
            else
                depth_term_vjp(ks, adj, -1.0f, th, A.g, A.f, A.Ddx, A.Dx, A.r_two_g, col_flip(C - 1), col_flip(C),
                               col_flip(C + 1), sa, gT, gE, gt);
            out[0] += gE[0]; out[3] += gE[1]; out[6] += gE[2];
            out[2] += gt[0]; out[5] += gt[1]; out[8] += gt[2];
            Q.lo = V3{out[0], out[1], out[2]};
            Q.mid = V3{out[3], out[4], out[5]};
            Q.hi = V3{out[6], out[7], out[8]};
        }
        // gather: ((west face's plus direct + east face's minus direct) +
        // (centre C-1's +1 + centre C's own)) + centre C+1's -1
        const V3 wdp = v3get(F.dp, addr_l);
        const V3 lhi = v3get(Q.hi, addr_l);
        const V3 rlo = v3get(Q.lo, addr_r);
        return V3{((wdp.a + F.dm.a) + (lhi.a + Q.mid.a)) + rlo.a,
                  ((wdp.b + F.dm.b) + (lhi.b + Q.mid.b)) + rlo.b,
                  ((wdp.c + F.dm.c) + (lhi.c + Q.mid.c)) + rlo.c};
    }

    // ---- y face (k | k+1) of the lane's column (interior columns): rm =
    //      row k, rp = row k+1 with the y slopes of both ----
    __device__ FaceOut y_face(int k, const Row& rm, const Row& rp, float Sum, float Svm, float Sup,
                              float Svp) const {
        FaceOut F;
        F.dm = v3z(); F.dp = v3z(); F.cm = v3z(); F.cp = v3z();
        if (!(k >= 1 && k <= A.ny + 1)) return F;                       // uniform
        const bool valid = ci && !(rm.E == kDryFlag || rp.E == kDryFlag);
        const float rdy = dt * A.r_dy;
        // flux order (mass, normal = v momentum, tangential = u momentum)
        const float l0 = (rp.W0 - rm.W0) * rdy;
        const float l1 = (rp.W2 - rm.W2) * rdy;
        const float l2 = (rp.W1 - rm.W1) * rdy;
        const float Rpx = rp.U - 0.5f * Sup, Rpy = rp.V - 0.5f * Svp;
        const float Rmx = rm.U + 0.5f * Sum, Rmy = rm.V + 0.5f * Svm;
        const float Hf = 0.5f * (rp.hc + rp.hn);
        const float hp_raw = (rp.E + Hf) - rp.S;
        const float hm_raw = (rm.E + Hf) + rm.S;
        const float hp = fmaxf(0.0f, hp_raw), hm = fmaxf(0.0f, hm_raw);
        const uint32_t fcode = (rm.code >> kCodeFy) & 31u;
        float ghm, gnm, gtm, ghp, gnp, gtp;
        if (__all(valid && hp > A.cutoff && hm > A.cutoff && !(fcode & 16u))) {
            flux_vjp<true>(hm, Rmy, Rmx, hp, Rpy, Rpx, fcode, A.g, A.cutoff, l0, l1, l2, ghm, gnm, gtm, ghp, gnp,
                           gtp);
        } else {
            if (!valid) return F;
            flux_vjp(hm, Rmy, Rmx, hp, Rpy, Rpx, fcode, A.g, A.cutoff, l0, l1, l2, ghm, gnm, gtm, ghp, gnp, gtp);
        }
        const float ghm_r = hm_raw > 0.0f ? ghm : 0.0f;
        const float ghp_r = hp_raw > 0.0f ? ghp : 0.0f;
        // back to (E, u, v): normal = v, tangential = u
        F.dm = V3{ghm_r, gtm, gnm};
        F.dp = V3{ghp_r, gtp, gnp};
        F.cm = V3{ghm_r, 0.5f * gtm, 0.5f * gnm};
        F.cp = V3{-ghp_r, -0.5f * gtp, -0.5f * gnp};
        return F;
    }

    // ---- y slope centre k: pw = plus-side terms of face (k-1 | k), fm =
    //      minus-side terms of face (k | k+1) ----
    __device__ CentreOut y_centre(int k, uint32_t code_k, const V3& pw, const V3& fm) const {
        CentreOut Q;
        Q.lo = v3z(); Q.mid = v3z(); Q.hi = v3z();
        if (!(ci && k >= 1 && k <= A.ny + 2)) return Q;
        const float gT = pw.a + fm.a;
        const float gSu = pw.b + fm.b;
        const float gSv = pw.c + fm.c;
        const uint32_t code = code_k >> 8;
        const float th = A.theta;
        float out[9];
        for (int q = 0; q < 9; ++q) out[q] = 0.0f;
        float cl, cc, cr;
        slope_coef(code & 3u, th, cl, cc, cr);
        out[1] = cl * gSu; out[4] = cc * gSu; out[7] = cr * gSu;
        slope_coef((code >> 2) & 3u, th, cl, cc, cr);
        out[2] = cl * gSv; out[5] = cc * gSv; out[8] = cr * gSv;
        const float sa = (A.bc_s == kWall && k < 2) ? -1.0f : 1.0f;
        float gE[3] = {0.0f, 0.0f, 0.0f}, gt[3] = {0.0f, 0.0f, 0.0f};
        const uint32_t ks = (code >> 4) & 3u, adj = (code >> 6) & 3u;
        if (__all(adj == 0u))
            depth_term_vjp<true>(ks, adj, 1.0f, th, A.g, A.f, A.Ddy, A.Dy, A.r_two_g, row_flip(k - 1), row_flip(k),
                                 row_flip(k + 1), sa, gT, gE, gt);
        else
            depth_term_vjp(ks, adj, 1.0f, th, A.g, A.f, A.Ddy, A.Dy, A.r_two_g, row_flip(k - 1), row_flip(k),
                           row_flip(k + 1), sa, gT, gE, gt);
        out[0] += gE[0]; out[3] += gE[1]; out[6] += gE[2];
        out[1] += gt[0]; out[4] += gt[1]; out[7] += gt[2];
        Q.lo = V3{out[0], out[1], out[2]};
        Q.mid = V3{out[3], out[4], out[5]};
        Q.hi = V3{out[6], out[7], out[8]};
        return Q;
    }

    // ---- output cell (R, C) (stage_out): gathered (gE, gU, gV), corners
    //      a = Hi[R][C], c = Hi[R][C+1] (the row's own), b = Hi[R+1][C],
    //      d = Hi[R+1][C+1] ----
    __device__ void output(int R, const Row& r, float gE, float gU, float gV, float b, float d) const {
        const float a = r.hc, c = r.hn;
        const uint32_t ro = (uint32_t)R * rowb;
        const bool interior = !ghost_r(R) && ci;
        const float E = r.E, U = r.U, V = r.V;
        float xe = 0.0f, xu = 0.0f, xv = 0.0f;
        if (interior) {
            const float w0 = r.W0, w1 = r.W1, w2 = r.W2;
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
            if (A.x) {
                bool clamped = false;
                if (A.x_clamp) clamped = ldb(A.x_clamp, 0, ro, oc) <= -Hreg + A.cutoff;
                if (!clamped) {
                    const uint32_t xs = A.x_stride;
                    if (A.x_fold && near_boundary(R)) {
                        xe = fold_read(A.x, xs, 0, R); xu = fold_read(A.x, xs, 1, R); xv = fold_read(A.x, xs, 2, R);
                    } else {
                        xe = ldb(A.x, 0, ro, oc); xu = ldb(A.x, xs, ro, oc); xv = ldb(A.x, 2 * xs, ro, oc);
                    }
                }
            }
        }
        // conversion (eta, hu, hv) -> (E, u, v)
        const uint32_t br = r.code >> 28;
        float h = 0.0f;
        if (br == kBrWet || br == kBrDesing) {
            const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag) - (float)(c == kDryFlag)
                              - (float)(d == kDryFlag);
            const float sum = ((a + b) + c) + d;
            h = E + ((wet == 4.0f) ? sum * 0.25f : (wet == 0.0f ? A.land : sum / wet));
        }
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
        const uint32_t os = A.o_stride;
        stb(A.o, 0, ro, oc, A.alpha * g_eta + A.gamma * xe);
        stb(A.o, os, ro, oc, A.alpha * g_hu + A.gamma * xu);
        stb(A.o, 2 * os, ro, oc, A.alpha * g_hv + A.gamma * xv);
    }

    // One pipeline iteration k: ra = row k-1, rb = row k, rc = row k+1 are
    // held, rd receives row k+2.  Returns false when the strip is done.
    __device__ bool step(int k, int r0, int r1, bool out_lane, Row& ra, Row& rb, Row& rc, Row& rd,
                         Carry& Z) const {
        if (k > r1 + 1) return false;
        // row k+2 (loaded in the previous iteration) with the corners of rows
        // k+2 and k+3, then the loads of row k+3 and of the corners of row k+4
        // for the next iteration
        float qc, qn;
        corners_flag(k + 3, Z.qhc, Z.qhn, qc, qn);
        finish(k + 2, Z.raw, Z.phc, Z.phn, qc, qn, rd);
        Z.phc = qc; Z.phn = qn;
        fetch(k + 3, Z.raw);
        corners_raw(k + 4, Z.qhc, Z.qhn);
        // the x direction of row k-1 needs nothing of row k+2: it runs while
        // those loads are in flight
        const int R = k - 1;
        const bool ri = R >= 2 && R <= A.ny + 1;
        V3 xs = v3z();
        if (R >= r0 && ri) xs = x_row(ra, rb.hn);       // every lane takes part (lane exchange)
        float Suy2, Svy2;
        Suy2 = slope_val((rc.code >> 8) & 3u, rb.U, rc.U, rd.U, A.theta);
        Svy2 = slope_val((rc.code >> 10) & 3u, rb.V, rc.V, rd.V, A.theta);
        // y face (k | k+1) and y centre k (plus side of (k-1 | k) carried)
        const FaceOut F = y_face(k, rb, rc, Z.Suy, Z.Svy, Suy2, Svy2);
        const CentreOut Q = y_centre(k, rb.code, Z.cp, F.cm);
        // row k-1's y sum, and row k's inner sums for the next iteration
        const V3 ys{(Z.p1.a + Z.p2.a) + Q.lo.a, (Z.p1.b + Z.p2.b) + Q.lo.b, (Z.p1.c + Z.p2.c) + Q.lo.c};
        Z.p1 = V3{Z.dp.a + F.dm.a, Z.dp.b + F.dm.b, Z.dp.c + F.dm.c};
        Z.p2 = V3{Z.c6.a + Q.mid.a, Z.c6.b + Q.mid.b, Z.c6.c + Q.mid.c};
        Z.dp = F.dp; Z.cp = F.cp; Z.c6 = Q.hi;
        Z.Suy = Suy2; Z.Svy = Svy2;
        // row k-1 is complete
        if (R >= r0) {
            if (out_lane && R <= A.ny + 3 && (ri || ci)) {
                float gE = 0.0f, gU = 0.0f, gV = 0.0f;
                if (ri) { gE += xs.a; gU += xs.b; gV += xs.c; }
                if (ci) { gE += ys.a; gU += ys.b; gV += ys.c; }
                output(R, ra, gE, gU, gV, rb.hc, rb.hn);
            }
        }
        return true;
    }
};

// One strip: lanes = padded columns 58*sx - 1 .. 58*sx + 62, output rows
// r0 .. r1 (padded, ghost rows included).
__device__ __forceinline__ void adj_march_strip(const MArgs& A, int sx, int r0, int r1) {
    const Strip P(A, sx);
    const int lane = P.lane, C = P.C, nx = A.nx;
    // lanes 3..60 have every cell they depend on in the strip; so do the
    // west ghost columns in lanes 1, 2 of the first strip (nothing lies west
    // of them) and the east ghost columns up to lane 62 (they depend on no
    // slope centre east of column nx+2; lane 63's right neighbour wraps)
    const bool out_lane = P.col_ok && ((lane >= 3 && lane <= 60) || (sx == 0 && lane >= 1 && lane <= 2) ||
                                       (C >= nx + 2 && lane <= 62));
    // first iteration k = r0 - 2: rows k-1 .. k+1 and the y slopes of row k
    const int k0 = r0 - 2;
    float h0c, h0n, h1c, h1n, h2c, h2n, h3c, h3n;
    P.corners(k0 - 1, h0c, h0n);
    P.corners(k0, h1c, h1n);
    P.corners(k0 + 1, h2c, h2n);
    P.corners(k0 + 2, h3c, h3n);
    Row s0, s1, s2, s3;
    P.load(k0 - 1, h0c, h0n, h1c, h1n, s0);
    P.load(k0, h1c, h1n, h2c, h2n, s1);
    P.load(k0 + 1, h2c, h2n, h3c, h3n, s2);
    Carry Z;
    Z.Suy = slope_val((s1.code >> 8) & 3u, s0.U, s1.U, s2.U, A.theta);
    Z.Svy = slope_val((s1.code >> 10) & 3u, s0.V, s1.V, s2.V, A.theta);
    Z.phc = h3c; Z.phn = h3n;
    P.fetch(k0 + 2, Z.raw);
    P.corners_raw(k0 + 3, Z.qhc, Z.qhn);
    Z.p1 = v3z(); Z.p2 = v3z(); Z.dp = v3z(); Z.cp = v3z(); Z.c6 = v3z();
    for (int k = k0; P.step(k, r0, r1, out_lane, s0, s1, s2, s3, Z); ++k) {
        s0 = s1; s1 = s2; s2 = s3;
    }
}

}  // namespace adjm
}  // namespace gocean
