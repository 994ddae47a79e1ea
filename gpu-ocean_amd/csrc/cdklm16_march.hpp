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
//     walks up bh output rows; lanes 2..61 write output, lanes 0, 1, 62, 63
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
// quantity is formed 64/60 times per output, plus a prologue of four
// conversions, two y slopes and a face per strip of bh rows.  The host picks
// bh so that the strips fill whole rounds of resident waves (march_rows).
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

// Each plane is addressed as a wave-uniform row address (scalar registers:
// the plane's base plus the row's byte offset) plus the lane's column byte
// offset (a vector register fixed for the strip), so a row's loads and
// stores take no vector address arithmetic.
using rsrc_t = const float*;

__device__ __forceinline__ rsrc_t plane(const float* p) { return p; }
// The row address, held in scalar registers and opaque to the optimizer:
// without this, loop-invariant code motion regroups (plane + row) + lane
// into (plane + lane) + row, which costs a 64-bit vector add per access
// and two vector registers per plane.  The empty asm emits no instruction.
using gchar = __attribute__((address_space(1))) char;     // global memory
__device__ __forceinline__ gchar* srow(rsrc_t r, uint32_t row_off) {
    gchar* row = (gchar*)(const_cast<char*>(reinterpret_cast<const char*>(r))) + row_off;
    asm("" : "+s"(row));
    return row;
}
// a wave-uniform flag as an opaque 32-bit scalar: held across the row loop
// in one scalar register and tested with a scalar compare, where a bool
// would be kept as a 64-bit lane mask (and spilled under register pressure)
__device__ __forceinline__ int sflag(bool b) {
    int v = __builtin_amdgcn_readfirstlane(b ? 1 : 0);   // float compares give a vector result
    asm("" : "+s"(v));
    return v;
}

// the lane's 32-bit offset, zero-extended at each access (scalar base plus
// 32-bit vector offset is the load's addressing form)
__device__ __forceinline__ uint32_t voff(uint32_t lane_off) {
    asm("" : "+v"(lane_off));
    return lane_off;
}
__device__ __forceinline__ float ld(rsrc_t r, uint32_t lane_off, uint32_t row_off) {
    return *(const __attribute__((address_space(1))) float*)(srow(r, row_off) + voff(lane_off));
}
__device__ __forceinline__ void st(rsrc_t r, uint32_t lane_off, uint32_t row_off, float v) {
    *(__attribute__((address_space(1))) float*)(srow(r, row_off) + voff(lane_off)) = v;
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
    const float h = eta + Hm;
    // every lane wet and deep: the plain branch alone
    if (__all(!any_dry && !(fabsf(Hm - K.land) <= kDryEps) && !(h < K.desing))) {
        o.E = eta;
        const float rh = rcp_refined(h);
        o.U = div_r(hu, h, rh);
        o.V = div_r(hv, h, rh);
        return o;
    }
    if (any_dry) {
        const float wet = 4.0f - (float)(a == kDryFlag) - (float)(b == kDryFlag)
                        - (float)(c == kDryFlag) - (float)(d == kDryFlag);
        Hm = (wet == 0.0f) ? K.land : sum / wet;
    }
    const float hh = eta + Hm;
    if (fabsf(Hm - K.land) <= kDryEps) {
        o.E = kDryFlag; o.U = 0.0f; o.V = 0.0f;
    } else if (hh < K.desing) {
        if (hh <= K.cutoff) {
            o.E = -Hm + K.cutoff; o.U = 0.0f; o.V = 0.0f;
        } else {
            o.E = eta;
            o.U = desingularize(hh, hu, K.desing);
            o.V = desingularize(hh, hv, K.desing);
        }
    } else {
        o.E = eta;
        const float rh = rcp_refined(hh);
        o.U = div_r(hu, hh, rh);
        o.V = div_r(hv, hh, rh);
    }
    return o;
}

// The backward candidates of a cell's three minmod slopes equal, bit for
// bit, the forward candidates of its low-side neighbour: (c - l) * theta of
// the cell is (r - c) * theta of the neighbour, and the geostrophic candidate
// theta g ((e_c - e_l) -+ d/(2g) (f_c + f_l)) is the neighbour's forward one
// up to the order of the commutative sum f_c + f_l.  So each slope centre
// takes them from its neighbour (the previous row in y, the left lane in x)
// and hands its own forward ones on.
struct Carry {
    float a;     // (a_c - a_l) * theta
    float b;     // (b_c - b_l) * theta
    float K;     // geostrophic backward candidate
    float fb;    // f * b_l (with b_l's wall flip)
};

// x or y slope centre of one cell (x: :607-666 with adjustSlopes_x :148-191;
// y: :711-770 with adjustSlopes_y :198-241), zero-rotation form.  l, c, r:
// the cell and its two neighbours along the direction; `a` the velocity
// along the direction (minmod only), `b` the velocity in the geostrophic
// term (v for x, u for y); Hlo / Hhi the face-averaged depths on the two
// sides; SGN -1 for x, +1 for y (the sign of the geostrophic term); flc /
// flr the wall flips of the cell and its high neighbour; fa the
// adjustment's own flip.  `in`: the backward candidates (from the low
// neighbour); `out`: this cell's forward candidates.  Returns the minmod
// slopes Sa, Sb and T = (K + d f v)/2g (x) or (L - d f u)/2g (y).
template <int SGN>
__device__ __forceinline__ void slope_fwd(float ec, float er, float ac, float ar, float bc, float br,
                                          bool flc, bool flr, float vd, const Consts& K, float fc,
                                          Carry& out, float& fr) {
    out.a = (ar - ac) * K.theta;
    out.b = (br - bc) * K.theta;
    out.fb = (flc ? -bc : bc) * fc;
    fr = (flr ? -br : br) * fc;
    if constexpr (SGN < 0) out.K = K.theta_g * ((er - ec) - vd * (out.fb + fr));
    else out.K = K.theta_g * ((er - ec) + vd * (out.fb + fr));
}

// REC: also the slope centre's branch record (tape code byte): the minmod
// selections of the u and v slopes and of K / L, and which positivity
// adjustment fired (x: a = u; y: a = v)
template <int SGN, bool FAST, bool REC = false>
__device__ __forceinline__ void slope_rest(float el, float ec, float er, float al, float ar, float bl, float bc,
                                           float br, float Hlo, float Hhi, bool flc, bool fa, float dd, float vd,
                                           const Consts& K, float fc, const Carry& in, const Carry& out, float fr,
                                           float& Sa, float& Sb, float& T, float& dlo, float& dhi,
                                           uint32_t* code8 = nullptr) {
    Sa = minmod_raw(in.a, (ar - al) * 0.5f, out.a);
    Sb = minmod_raw(in.b, (br - bl) * 0.5f, out.b);
    const float fl = in.fb, fcc = out.fb;
    float ce;
    if constexpr (SGN < 0) ce = K.half_g * ((er - el) - vd * ((fr + 2.0f * fcc) + fl));
    else ce = K.half_g * ((er - el) + vd * ((fr + 2.0f * fcc) + fl));
    float Kv = minmod_raw(in.K, ce, out.K);
    const float dfc = dd * fc;
    const float dfb = dfc * (fa ? -bc : bc);
    const float t = (SGN < 0) ? div_k<FAST>(Kv + dfb, K.two_g, K.r_two_g) : div_k<FAST>(Kv - dfb, K.two_g, K.r_two_g);
    const float hlo = (ec + Hlo) - t;
    const float hhi = (ec + Hhi) + t;
    if constexpr (REC) {
        const uint32_t sa = minmod_sel(in.a, (ar - al) * 0.5f, out.a);
        const uint32_t sb = minmod_sel(in.b, (br - bl) * 0.5f, out.b);
        const uint32_t sk = minmod_sel(in.K, ce, out.K);
        const uint32_t ad = !(hhi > 0.0f) ? 2u : (!(hlo > 0.0f) ? 1u : 0u);
        *code8 = (SGN < 0 ? (sa | (sb << 2)) : (sb | (sa << 2))) | (sk << 4) | (ad << 6);
    }
    bool adj = false;
    if constexpr (SGN < 0) {
        if (!(hlo > 0.0f)) { Kv = (-dfb) + K.two_g * (ec + Hlo); adj = true; }
        if (!(hhi > 0.0f)) { Kv = (-dfb) - K.two_g * (ec + Hhi); adj = true; }
    } else {
        if (!(hlo > 0.0f)) { Kv = dfb + K.two_g * (ec + Hlo); adj = true; }
        if (!(hhi > 0.0f)) { Kv = dfb - K.two_g * (ec + Hhi); adj = true; }
    }
    // T and the depths of the cell's two faces as the face fluxes form them,
    // (e + H_face) -/+ T: with T == t they are hlo / hhi above
    T = t;
    dlo = hlo;
    dhi = hhi;
    if (adj || flc != fa) {
        const float db = dfc * (flc ? -bc : bc);
        T = (SGN < 0) ? div_k<FAST>(Kv + db, K.two_g, K.r_two_g) : div_k<FAST>(Kv - db, K.two_g, K.r_two_g);
        dlo = (ec + Hlo) - T;
        dhi = (ec + Hhi) + T;
    }
}

// the backward candidates of a cell computed directly (the first y slope of
// a strip has no previous row)
template <int SGN>
__device__ __forceinline__ Carry slope_bwd(float el, float ec, float al, float ac, float bl, float bc,
                                           bool fll, bool flc, float vd, const Consts& K, float fc) {
    Carry o;
    o.a = (ac - al) * K.theta;
    o.b = (bc - bl) * K.theta;
    o.fb = (fll ? -bl : bl) * fc;
    const float fcc = (flc ? -bc : bc) * fc;
    if constexpr (SGN < 0) o.K = K.theta_g * ((ec - el) - vd * (fcc + o.fb));
    else o.K = K.theta_g * ((ec - el) + vd * (fcc + o.fb));
    return o;
}

// one face between cell m (low side) and p (high side) (:246-370):
// reconstructed states and central-upwind flux.  dm / dp: the face depths
// (e + H_face) + T of m and (e + H_face) - T of p before the clamp at zero,
// as the slope centres formed them.  For x faces (a, b) = (u, v) and the
// result is (F.x, F.y, F.z); for y faces (a, b) = (v, u) and the flux comes
// back as (G.x, G.z, G.y) -- the caller swaps, as the reference's
// computeGFaceFlux does.  When every lane of the wave has a wet face
// (neither side dry-flagged, both depths above the cutoff, wave-speed spread
// at least flux_eps), the same operations run without the selects of the
// dry and zero-flux cases (and, with FAST, the wave speeds take the Newton
// square root: both depths are above the cutoff there).
// REC: also the face's branch record (which states give a+ and a-, and
// the zero-flux bit), as cu_flux<true>.
template <bool FAST, bool REC = false>
__device__ __forceinline__ void face(float em, float am, float bm, float hSam, float hSbm, float dm,
                                     float ep, float ap, float bp, float hSap, float hSbp, float dp,
                                     const Consts& K, float& F0, float& F1, float& F2, uint32_t* code = nullptr) {
    const float hp = fmaxf(0.0f, dp);
    const float hm = fmaxf(0.0f, dm);
    if (__all(!(ep == kDryFlag || em == kDryFlag) && hp > K.cutoff && hm > K.cutoff)) {
        const float up = ap - hSap, vp = bp - hSbp;
        const float um = am + hSam, vm = bm + hSbm;
        const float qp = hp * up;
        const float Fpx = qp, Fpy = qp * up + (K.half_g * hp) * hp, Fpz = qp * vp;
        const float c_p = sqrt_k<FAST>(K.g * hp);
        const float qm = hm * um;
        const float Fmx = qm, Fmy = qm * um + (K.half_g * hm) * hm, Fmz = qm * vm;
        const float c_m = sqrt_k<FAST>(K.g * hm);
        const float amin = fminf(fminf(um - c_m, up - c_p), 0.0f);
        const float amax = fmaxf(fmaxf(um + c_m, up + c_p), 0.0f);
        const float den = amax - amin;
        if constexpr (REC) {
            const float b1 = um + c_m, b2 = up + c_p, a1 = um - c_m, a2 = up - c_p;
            const uint32_t cp = (fmaxf(b1, b2) > 0.0f) ? ((b1 >= b2) ? 1u : 2u) : 0u;
            const uint32_t cm = (fminf(a1, a2) < 0.0f) ? ((a1 <= a2) ? 1u : 2u) : 0u;
            *code = cp | (cm << 2) | ((fabsf(den) < K.flux_eps) ? 16u : 0u);
        }
        const float apam = amax * amin;
        const float rden = rcp_refined(den);
        F0 = div_r((amax * Fmx - amin * Fpx) + apam * (hp - hm), den, rden);
        F1 = div_r((amax * Fmy - amin * Fpy) + apam * (Fpx - Fmx), den, rden);
        F2 = (um + up > 0.0f) ? Fmz : Fpz;
        if (!__all(!(fabsf(den) < K.flux_eps))) {
            const bool z = fabsf(den) < K.flux_eps;
            F0 = z ? 0.0f : F0; F1 = z ? 0.0f : F1; F2 = z ? 0.0f : F2;
        }
        return;
    }
    F0 = 0.0f; F1 = 0.0f; F2 = 0.0f;
    if constexpr (REC) *code = 0u;
    if (!(ep == kDryFlag || em == kDryFlag))
        cu_flux<REC>(hm, am + hSam, bm + hSbm, hp, ap - hSap, bp - hSbp, K, F0, F1, F2, code);
}

// One strip: padded columns sx*60 .. sx*60+63 (lanes), output rows
// 2 + sy*bh .. min(2 + (sy+1)*bh, ny+2) - 1.  WGRID: the wind stress comes
// from A.wind_grid (else the uniform A.wind_x, A.wind_y); a separate
// instantiation, because the per-lane interpolation state and lookups would
// otherwise hold registers in every launch.  REC: also write the branch
// record (A.tape_*) of the reverse pass, the same words the tile kernel's
// recording variant writes.
template <bool FAST, bool WGRID, bool REC = false>
__device__ __forceinline__ void march_strip(const SubstepArgs& A, int sx, int sy, int bh) {
    const int lane = threadIdx.x & (kLanes - 1);
    const int nx = A.nx, ny = A.ny;
    const uint32_t rowb = (uint32_t)A.pitch * 4u, hrowb = (uint32_t)A.hi_pitch * 4u;   // row bytes
    const Consts K = make_consts(A);
    const int C = sx * kOut + lane;                  // padded column of the lane
    const int Cl = min(C, nx + 3);                   // load column (clamped)
    const int j0 = 2 + sy * bh;
    const int j1 = min(j0 + bh, ny + 2);
    const int wN = sflag(A.bc_n() == kWall), wS = sflag(A.bc_s() == kWall);
    const bool wE = A.bc_e() == kWall, wW = A.bc_w() == kWall;
    const int pNS = sflag(A.bc_n() == kPeriodic), pEW = sflag(A.bc_e() == kPeriodic);

    // the lane's source column (wall mirror with u negated, or periodic wrap)
    int sc = Cl;
    bool neg_u = false;
    if (Cl < 2) {
        if (wW) { sc = 3 - Cl; neg_u = true; } else if (pEW) { sc = Cl + nx; }
    } else if (Cl > nx + 1) {
        if (wE) { sc = 2 * nx + 3 - Cl; neg_u = true; } else if (pEW) { sc = Cl - nx; }
    }
    auto col_flip = [&](int c) { return (wW && c < 2) || (wE && c > nx + 1); };
    const bool xfc = col_flip(C), xfr = col_flip(C + 1);
    const bool xfa = wW && C < 2;
    auto row_flip = [&](int r) { return (wS && r < 2) || (wN && r > ny + 1); };
    const int hc = min(C, nx + 4), hn = min(C + 1, nx + 4);   // corner-depth columns
    // byte offsets of the lane's columns within a row
    const uint32_t osc = (uint32_t)sc * 4u, ohc = (uint32_t)hc * 4u, ohn = (uint32_t)hn * 4u;
    const uint32_t ocl = (uint32_t)Cl * 4u;
    const int addr_r = ((lane + 1) & (kLanes - 1)) << 2;
    const int addr_l = ((lane - 1) & (kLanes - 1)) << 2;
    const rsrc_t r_ein = plane(A.eta_in), r_uin = plane(A.hu_in), r_vin = plane(A.hv_in);
    const rsrc_t r_hi = plane(A.Hi);
    // branch record: each cell's words are written by one strip -- interior
    // columns by lanes 2..61, the ghost columns by the first and last strip
    // (ghost corners are never written, as in the tile kernel)
    const bool col_in = C >= 2 && C <= nx + 1;
    const bool rec_lane = REC && (col_in ? (lane >= 2 && lane <= 61)
                                         : (C <= 1 ? sx == 0 : (C <= nx + 3 && sx == A.tiles_x - 1)));
    auto rec_store = [&](int R, float u, float v) {
        const uint32_t ro = (uint32_t)R * rowb;
        st(plane(A.tape_u), ocl, ro, u);
        st(plane(A.tape_v), ocl, ro, v);
    };
    auto rec_code = [&](int R, uint32_t code) {
        st(plane(reinterpret_cast<const float*>(A.tape_code)), ocl, (uint32_t)R * rowb, __uint_as_float(code));
    };

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
        const uint32_t ro = (uint32_t)sr * rowb;
        e = ld(r_ein, osc, ro);
        u = ld(r_uin, osc, ro);
        v = ld(r_vin, osc, ro);
        if (neg_u) u = -u;
        if (neg_v) v = -v;
    };
    auto load_hi = [&](int R, float& a, float& b) {
        const uint32_t ro = (uint32_t)R * hrowb;
        a = flag_land(ld(r_hi, ohc, ro), K);
        b = flag_land(ld(r_hi, ohn, ro), K);
    };

    // ---- prologue: conversion of rows j0-2 .. j0+1, y slopes of rows
    //      j0-1 and j0, the flux through the face (j0-1 | j0) ----
    float hcA, hnA, hcB, hnB, hcC, hnC, hcD, hnD;      // corner rows j0-2 .. j0+1
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
    const float Hs_m1 = 0.5f * (hcA + hnA);    // row j0-1's south face
    const float Hs_0 = 0.5f * (hcB + hnB);     // row j0's south face (= row j0-1's north)
    const float Hs_1 = 0.5f * (hcC + hnC);     // row j0+1's south face
    // y slopes of row j0-1 (cells c0, c1, c2) and row j0 (c1, c2, c3); in y
    // (a, b) = (v, u).  Faces use half slopes, kept as such.
    float hUy1, hVy1, S1, dhi1;
    Carry ycar;
    float gsx, gsy, gsz;
    uint32_t yc1 = 0u;                               // record byte of row j0's y slope centre
    {
        float Uy0, Vy0, S0, dlo0, dhi0, Uy1, Vy1, dlo1;
        uint32_t yc0 = 0u, gc0 = 0u;
        const Carry bwd = slope_bwd<+1>(c0.E, c1.E, c0.V, c1.V, c0.U, c1.U, row_flip(j0 - 2), row_flip(j0 - 1),
                                        K.udy, K, fc);
        Carry fwd;
        float fr;
        slope_fwd<+1>(c1.E, c2.E, c1.V, c2.V, c1.U, c2.U, row_flip(j0 - 1), row_flip(j0), K.udy, K, fc, fwd, fr);
        slope_rest<+1, FAST, REC>(c0.E, c1.E, c2.E, c0.V, c2.V, c0.U, c1.U, c2.U, Hs_m1, Hs_0, row_flip(j0 - 1),
                                  wS && j0 - 1 < 2, K.dy, K.udy, K, fc, bwd, fwd, fr, Vy0, Uy0, S0, dlo0, dhi0, &yc0);
        Carry fwd1;
        slope_fwd<+1>(c2.E, c3.E, c2.V, c3.V, c2.U, c3.U, row_flip(j0), row_flip(j0 + 1), K.udy, K, fc, fwd1, fr);
        slope_rest<+1, FAST, REC>(c1.E, c2.E, c3.E, c1.V, c3.V, c1.U, c2.U, c3.U, Hs_0, Hs_1, row_flip(j0),
                                  wS && j0 < 2, K.dy, K.udy, K, fc, fwd, fwd1, fr, Vy1, Uy1, S1, dlo1, dhi1, &yc1);
        ycar = fwd1;
        hUy1 = 0.5f * Uy1;
        hVy1 = 0.5f * Vy1;
        // flux through (j0-1 | j0): m = row j0-1 (c1), p = row j0 (c2)
        float a, b, c;
        face<FAST, REC>(c1.E, c1.V, c1.U, 0.5f * Vy0, 0.5f * Uy0, dhi0, c2.E, c2.V, c2.U, hVy1, hUy1, dlo1, K, a, b,
                        c, &gc0);
        gsx = a; gsy = c; gsz = b;
        // the first strip of a column records ghost rows 0 and 1 (interior
        // columns; row 1's y slope centre and its north face (1 | 2))
        if (REC && sy == 0 && rec_lane && col_in) {
            rec_store(0, c0.U, c0.V);
            rec_code(0, 0u);
            rec_store(1, c1.U, c1.V);
            st(plane(A.tape_s), ocl, rowb, S0);
            rec_code(1, (yc0 << kCodeY) | (gc0 << kCodeFy));
        }
    }
    // window: (cj, cj1) = rows j, j+1; (hUyj, hVyj, Sj, dhij) = half y slopes,
    // depth term and north-face depth of row j; ycar = forward candidates of
    // row j; corner rows j, j+1, j+2; Hsj = south face of row j, Hnj = its
    // north face
    Cell cj = c2, cj1 = c3;
    float hUyj = hUy1, hVyj = hVy1, Sj = S1, dhij = dhi1;
    float hcj = hcB, hnj = hnB, hcj1 = hcC, hnj1 = hnC, hcj2 = hcD, hnj2 = hnD;
    float Hsj = Hs_0, Hnj = Hs_1;
    const int friction = sflag(A.r > 0.0f);
    // gridded wind: the lane's column weights, fixed for the strip
    constexpr bool wgrid = WGRID;
    WindAxis wax;
    wax.i0 = wax.i1 = 0; wax.a = 0.0f;
    if (wgrid) wax = wind_axis(C, nx, A.wind_wx);
    const int rk_order = A.rk_order(), rk_step = A.rk_step();
    const bool store_lane = lane >= 2 && lane < 2 + kOut && C <= nx + 1;
    const rsrc_t r_eqn = plane(A.eta_qn), r_uqn = plane(A.hu_qn), r_vqn = plane(A.hv_qn);
    const rsrc_t r_eo = plane(A.eta_out), r_uo = plane(A.hu_out), r_vo = plane(A.hv_out);

    uint32_t ycj = yc1;                               // record byte of row j's y slope centre
    for (int j = j0; j < j1; ++j) {
        // loads for row j+2 (conversion) and Q^n of row j (second stages)
        float e2, u2, v2, hcj3, hnj3;
        load_state(j + 2, e2, u2, v2);
        load_hi(j + 3, hcj3, hnj3);
        const uint32_t rj = (uint32_t)j * rowb;
        float ea = 0.0f, ha = 0.0f, va = 0.0f;
        if (rk_step > 0) { ea = ld(r_eqn, ocl, rj); ha = ld(r_uqn, ocl, rj); va = ld(r_vqn, ocl, rj); }

        // ---- x slope centre of row j (neighbours by lane exchange) ----
        const float el = lane_get(cj.E, addr_l), er = lane_get(cj.E, addr_r);
        const float ul = lane_get(cj.U, addr_l), ur = lane_get(cj.U, addr_r);
        const float vl = lane_get(cj.V, addr_l), vr = lane_get(cj.V, addr_r);
        const float Hw = 0.5f * (hcj + hcj1);            // west face depth of the cell
        const float He = 0.5f * (hnj + hnj1);            // east face depth
        float hUx, hVx, T, dlox, dhix;
        uint32_t xc = 0u, fce = 0u, ycn = 0u, gcn = 0u;   // record bytes of row j (x, east face, north face), row j+1 (y)
        {
            Carry fwd, bwd;
            float fr, Ux, Vx;
            slope_fwd<-1>(cj.E, er, cj.U, ur, cj.V, vr, xfc, xfr, K.vdx, K, fc, fwd, fr);
            bwd.a = lane_get(fwd.a, addr_l);
            bwd.b = lane_get(fwd.b, addr_l);
            bwd.K = lane_get(fwd.K, addr_l);
            bwd.fb = lane_get(fwd.fb, addr_l);
            slope_rest<-1, FAST, REC>(el, cj.E, er, ul, ur, vl, cj.V, vr, Hw, He, xfc, xfa, K.dx, K.vdx, K, fc, bwd,
                                      fwd, fr, Ux, Vx, T, dlox, dhix, &xc);
            hUx = 0.5f * Ux;
            hVx = 0.5f * Vx;
        }
        // ---- east face (C | C+1) ----
        const float hUxr = lane_get(hUx, addr_r), hVxr = lane_get(hVx, addr_r), dpr = lane_get(dlox, addr_r);
        float fex, fey, fez;
        face<FAST, REC>(cj.E, cj.U, cj.V, hUx, hVx, dhix, er, ur, vr, hUxr, hVxr, dpr, K, fex, fey, fez, &fce);

        // ---- conversion of row j+2, y slopes of row j+1, north face ----
        const Cell cj2 = convert(e2, u2, v2, hcj2 + hcj3, hcj2, hcj3, hnj2, hnj3, K);
        const float Hnj1 = 0.5f * (hcj2 + hnj2);         // north face of row j+1
        float hUyn, hVyn, Sn, dlon, dhin;
        {
            Carry fwd;
            float fr, Uyn, Vyn;
            slope_fwd<+1>(cj1.E, cj2.E, cj1.V, cj2.V, cj1.U, cj2.U, row_flip(j + 1), row_flip(j + 2), K.udy, K,
                          fc, fwd, fr);
            slope_rest<+1, FAST, REC>(cj.E, cj1.E, cj2.E, cj.V, cj2.V, cj.U, cj1.U, cj2.U, Hnj, Hnj1, row_flip(j + 1),
                                      wS && j + 1 < 2, K.dy, K.udy, K, fc, ycar, fwd, fr, Vyn, Uyn, Sn, dlon, dhin,
                                      &ycn);
            ycar = fwd;
            hUyn = 0.5f * Uyn;
            hVyn = 0.5f * Vyn;
        }
        float gnx, gny, gnz;
        {
            float a, b, c;
            face<FAST, REC>(cj.E, cj.V, cj.U, hVyj, hUyj, dhij, cj1.E, cj1.V, cj1.U, hVyn, hUyn, dlon, K, a, b, c,
                            &gcn);
            gnx = a; gny = c; gnz = b;
        }
        // ---- branch record of row j (an interior row) ----
        if (REC && rec_lane) {
            rec_store(j, cj.U, cj.V);
            uint32_t code = 0u;
            if (C >= 1 && C <= nx + 2) {
                st(plane(A.tape_t), ocl, rj, T);
                code |= xc << kCodeX;
                if (C <= nx + 1) code |= fce << kCodeFx;
            }
            if (col_in) {
                st(plane(A.tape_s), ocl, rj, Sj);
                code |= (ycj << kCodeY) | (gcn << kCodeFy);
            }
            rec_code(j, code);
        }
        // ---- west face from the left lane ----
        const float fwx = lane_get(fex, addr_l), fwy = lane_get(fey, addr_l), fwz = lane_get(fez, addr_l);

        // ---- sources, RK update, store (:802-1004) ----
        if (store_lane) {
            const float Hm = cj.Hm4;
            const float ec = cj.E, uc = cj.U, vc = cj.V;
            const float hreg = ec + Hm;
            const float fd0 = div_k<FAST>(fex - fwx, K.dx, K.r_dx) + div_k<FAST>(gnx - gsx, K.dy, K.r_dy);
            const float fd1 = div_k<FAST>(fey - fwy, K.dx, K.r_dx) + div_k<FAST>(gny - gsy, K.dy, K.r_dy);
            const float fd2 = div_k<FAST>(fez - fwz, K.dx, K.r_dx) + div_k<FAST>(gnz - gsz, K.dy, K.r_dy);
            const float eta_west = ec - T, eta_east = ec + T;
            const float eta_south = ec - Sj, eta_north = ec + Sj;
            // wind, Coriolis and bottom slope of a wet cell
            auto sources = [&](float hu, float hv, float& st1, float& st2) {
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
                st1 = (X + hu_cor) + div_k<FAST>(b1, K.dx, K.r_dx);
                st2 = (Y + hv_cor) + div_k<FAST>(b2, K.dy, K.r_dy);
            };
            float hu, hv, st1, st2;
            if (__all(hreg > K.cutoff && ec != kDryFlag)) {
                // every writing lane wet: no selects
                hu = uc * hreg;
                hv = vc * hreg;
                sources(hu, hv, st1, st2);
            } else {
                hu = 0.0f; hv = 0.0f; st1 = 0.0f; st2 = 0.0f;
                if (hreg > K.cutoff) {
                    hu = uc * hreg;
                    hv = vc * hreg;
                }
                if (hreg >= K.cutoff && ec != kDryFlag) sources(hu, hv, st1, st2);
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
            st(r_eo, ocl, rj, ue);
            st(r_uo, ocl, rj, uu);
            st(r_vo, ocl, rj, uv);
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
                        const uint32_t ro = (uint32_t)rows[a] * rowb, co = (uint32_t)cols[b] * 4u;
                        st(r_eo, co, ro, ue); st(r_uo, co, ro, uu); st(r_vo, co, ro, uv);
                    }
                }
            }
        }

        // ---- slide the window one row north ----
        ycj = ycn;
        gsx = gnx; gsy = gny; gsz = gnz;
        cj = cj1; cj1 = cj2;
        hUyj = hUyn; hVyj = hVyn; Sj = Sn; dhij = dhin;
        hcj = hcj1; hnj = hnj1; hcj1 = hcj2; hnj1 = hnj2; hcj2 = hcj3; hnj2 = hnj3;
        Hsj = Hnj; Hnj = Hnj1;
    }
    // the last strip of a column records ghost rows ny+2 (its y slope
    // centre) and ny+3 (interior columns)
    if (REC && j1 == ny + 2 && rec_lane && col_in) {
        rec_store(ny + 2, cj.U, cj.V);
        st(plane(A.tape_s), ocl, (uint32_t)(ny + 2) * rowb, Sj);
        rec_code(ny + 2, ycj << kCodeY);
        rec_store(ny + 3, cj1.U, cj1.V);
        rec_code(ny + 3, 0u);
    }
}

}  // namespace march
}  // namespace gocean
