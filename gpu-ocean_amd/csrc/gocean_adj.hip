// gocean_adj.hip -- reverse mode of the fixed-dt CDKLM16 SSP-RK2 step for one
// domain: the C ABI entry points and their launches.  The device code of a
// substep adjoint is in adjoint_tile.hpp.  Per substep adjoint:
//   * the forward substep re-run in its recording variant writes the branch
//     record (u, v, the x / y face depth terms, and a code word with every
//     limiter / adjustment / wave-speed choice) for each cell;
//   * k_adj_tile: one workgroup per output tile (adjoint_tile.hpp);
//   * k_adj_fold: ghost-cell adjoints back to their interior images, after
//     the last substep of a call only: a substep adjoint that reads another
//     one's output folds the ghost entries in while reading.
#include <hip/hip_runtime.h>

#include <cstring>
#include <vector>

#include "../../include/gocean.h"
#include "adjoint_march.hpp"
#include "adjoint_tile.hpp"
#include "gocean_internal.hpp"

namespace gocean {

__global__ __launch_bounds__(kAdjNT) void k_adj_tile(AdjTileArgs A) {
    __shared__ AdjSmem<kAdjBW, kAdjBH> S;
    const int t = xcd_band(blockIdx.x, gridDim.x);
    const int ty = t / A.tiles_x, tx = t - ty * A.tiles_x;
    adj_tile_body<kAdjBW, kAdjBH, kAdjNT, false>(A, S, tx, ty, threadIdx.x);
}

// column-marching form (adjoint_march.hpp): one wavefront per strip of 58
// output columns and bh rows (padded rows 0 .. ny+3, ghost rows included)
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3))) void k_adj_march(adjm::MArgs A, int strips_x, int bh, int n_strips) {
    const int wave = xcd_band(blockIdx.x, gridDim.x) * 4 + march::wave_in_group();
    if (wave >= n_strips) return;
    const int sy = wave / strips_x, sx = wave - sy * strips_x;
    const int r0 = sy * bh;
    const int r1 = min(r0 + bh, A.ny + 4) - 1;
    adjm::adj_march_strip(A, sx, r0, r1);
}

// reverse kernel: 0 tile, 1 column-marching, -1 by size (marching from
// 2 M cells on: at 1024^2 the two take the same time, at 2048^2 marching
// is 25 % faster)
int g_adj_kernel = -1;

bool adj_use_march(long long cells) { return g_adj_kernel == 1 || (g_adj_kernel < 0 && cells >= (2LL << 20)); }

// rows per strip of k_adj_march for domains of the given widths and heights
int adj_march_rows(const int* nx, const int* ny, int n) {
    static int resident = 0;
    if (!resident) resident = resident_waves((const void*)k_adj_march);
    std::vector<int> strips(n), rows(n);
    for (int k = 0; k < n; ++k) {
        strips[k] = (nx[k] + adjm::kOut - 1) / adjm::kOut;
        rows[k] = ny[k] + 4;
    }
    return march_rows(strips.data(), rows.data(), n, resident);
}

__global__ __launch_bounds__(256) void k_adj_fold(FoldArgs F) {
    fold_item(F, blockIdx.y, blockIdx.x * blockDim.x + threadIdx.x);
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int adj_validate(const gocean_domain_t* d) {
    if (int e = validate_domain(d)) return e;
    if (!(d->north_x == 0.0f && d->north_y == 1.0f))
        return set_error(GOCEAN_EINVAL, "the adjoint supports zero grid rotation only (north = (0,1))");
    if (d->r != 0.0f) return set_error(GOCEAN_EINVAL, "the adjoint supports r == 0 only");
    if (d->rk_order != 2) return set_error(GOCEAN_EINVAL, "the adjoint supports rk_order 2 only");
    return 0;
}

static void adj_fill(const gocean_domain_t* d, AdjTileArgs& A) {
    std::memset(&A, 0, sizeof(A));
    A.nx = d->nx; A.ny = d->ny; A.pitch = d->pitch; A.hi_pitch = d->hi_pitch;
    A.dx = d->dx; A.dy = d->dy; A.g = d->g; A.theta = d->theta; A.desing = d->desing_eps;
    A.flux_eps = d->flux_slope_eps; A.cutoff = d->depth_cutoff; A.land = d->land_value; A.f = d->f;
    A.bc_n = d->bc_north; A.bc_e = d->bc_east; A.bc_s = d->bc_south; A.bc_w = d->bc_west;
    A.tiles_x = (d->nx + kAdjBW - 1) / kAdjBW;
    A.tiles_y = (d->ny + kAdjBH - 1) / kAdjBH;
    adj_derive(A);
}

struct Tape {
    float* u; float* v; float* t; float* s; uint32_t* code;
};

static Tape tape_at(float* base, long long plane) {
    Tape T;
    T.u = base; T.v = base + plane; T.t = base + 2 * plane; T.s = base + 3 * plane;
    T.code = reinterpret_cast<uint32_t*>(base + 4 * plane);
    return T;
}

// forward substep that also writes the branch record
static int fwd_tape(const gocean_domain_t* d, const gocean_wind_t* wind, float dt, const float* dt_dev,
                    int rk_step,
                    const float* e_in, const float* u_in, const float* v_in,
                    const float* e_qn, const float* u_qn, const float* v_qn,
                    float* e_out, float* u_out, float* v_out, const float* Hi, const Tape& T,
                    hipStream_t s) {
    SubstepArgs A;
    fill_args(d, A);
    A.dt = dt; A.dt_dev = dt_dev;
    A.set_rk_step(rk_step);
    if (wind && wind->grid)
        return set_error(GOCEAN_EINVAL, "the adjoint takes uniform wind only (wind grid given)");
    if (wind) {
        A.wind_x = wind_lerp(wind->t, wind->x0, wind->x1);
        A.wind_y = wind_lerp(wind->t, wind->y0, wind->y1);
    }
    A.eta_in = e_in; A.hu_in = u_in; A.hv_in = v_in;
    A.eta_qn = e_qn ? e_qn : e_out; A.hu_qn = u_qn ? u_qn : u_out; A.hv_qn = v_qn ? v_qn : v_out;
    A.eta_out = e_out; A.hu_out = u_out; A.hv_out = v_out;
    A.Hi = Hi;
    A.tape_u = T.u; A.tape_v = T.v; A.tape_t = T.t; A.tape_s = T.s; A.tape_code = T.code;
    return launch_substep_tape(A, s);
}

// the three fields of an array as one base and a byte stride, when they are
// equally spaced (the reverse entry points' buffers always are)
static bool triple(const float* a, const float* b, const float* c, const char*& base, long long& stride) {
    base = reinterpret_cast<const char*>(a);
    if (!a) { stride = 0; return true; }
    stride = (long long)(reinterpret_cast<const char*>(b) - base);
    return reinterpret_cast<const char*>(c) - reinterpret_cast<const char*>(b) == stride;
}

// the marching kernel's arguments, false when an array's planes are not
// equally spaced in increasing order or its last plane ends beyond 4 GiB of
// its base (then the tile kernel runs)
bool adj_march_args(const AdjTileArgs& A, adjm::MArgs& M) {
    const long long plane = (long long)(A.ny + 4) * A.pitch * 4;     // bytes
    auto fits = [&](long long st, int planes) {
        return st >= 0 && st * (planes - 1) + plane <= 0xFFFFFFFFLL;
    };
    const char* p;
    long long st;
    if (!triple(A.w_eta, A.w_hu, A.w_hv, p, st) || !fits(st, 3)) return false;
    M.w = p; M.w_stride = (uint32_t)st;
    if (!triple(A.x_eta, A.x_hu, A.x_hv, p, st) || !fits(st, 3)) return false;
    M.x = p; M.x_stride = (uint32_t)st;
    if (!triple(A.o_eta, A.o_hu, A.o_hv, p, st) || !fits(st, 3)) return false;
    M.o = const_cast<char*>(p); M.o_stride = (uint32_t)st;
    const char* t = reinterpret_cast<const char*>(A.tU);
    const long long ts = (long long)(reinterpret_cast<const char*>(A.tV) - t);
    if (reinterpret_cast<const char*>(A.tT) != t + 2 * ts || reinterpret_cast<const char*>(A.tS) != t + 3 * ts ||
        reinterpret_cast<const char*>(A.tC) != t + 4 * ts || !fits(ts, 5))
        return false;
    M.tape = t; M.tape_stride = (uint32_t)ts;
    M.eta = A.eta; M.w_clamp = A.w_clamp; M.x_clamp = A.x_clamp; M.Hi = A.Hi; M.dt_dev = A.dt_dev;
    M.dt = A.dt; M.alpha = A.alpha; M.gamma = A.gamma; M.w_fold = A.w_fold; M.x_fold = A.x_fold;
    M.nx = A.nx; M.ny = A.ny; M.pitch = A.pitch; M.hi_pitch = A.hi_pitch;
    M.g = A.g; M.theta = A.theta; M.desing = A.desing; M.cutoff = A.cutoff; M.land = A.land; M.f = A.f;
    M.bc_n = A.bc_n; M.bc_e = A.bc_e; M.bc_s = A.bc_s; M.bc_w = A.bc_w;
    M.r_dx = A.r_dx; M.r_dy = A.r_dy; M.r_two_g = A.r_two_g; M.Dx = A.Dx; M.Dy = A.Dy;
    M.Ddx = A.Ddx; M.Ddy = A.Ddy; M.g_dx = A.g_dx; M.g_dy = A.g_dy;
    return true;
}

static int adj_launch(const AdjTileArgs& A, bool fold_out, hipStream_t s) {
    adjm::MArgs M;
    const bool march = adj_use_march((long long)A.nx * A.ny);
    if (march && adj_march_args(A, M)) {
        const int sx = (A.nx + adjm::kOut - 1) / adjm::kOut;
        const int bh = adj_march_rows(&A.nx, &A.ny, 1);
        const int n = sx * ((A.ny + 4 + bh - 1) / bh);
        hipLaunchKernelGGL(k_adj_march, dim3((n + 3) / 4), dim3(256), 0, s, M, sx, bh, n);
    } else {
        const int nb = A.tiles_x * A.tiles_y;
        hipLaunchKernelGGL(k_adj_tile, dim3(nb), dim3(kAdjNT), 0, s, A);
    }
    if (!fold_out) return check_launch("adjoint substep launch");
    FoldArgs F;
    F.e = A.o_eta; F.u = A.o_hu; F.v = A.o_hv;
    F.nx = A.nx; F.ny = A.ny; F.pitch = A.pitch;
    F.bc_n = A.bc_n; F.bc_e = A.bc_e; F.bc_s = A.bc_s; F.bc_w = A.bc_w;
    const int n = (A.nx > A.ny ? A.nx : A.ny) + 16;
    hipLaunchKernelGGL(k_adj_fold, dim3((n + 255) / 256, 9), dim3(256), 0, s, F);
    return check_launch("adjoint substep launch");
}

// adjoint of one substep at the state whose eta is e_in, with its record T;
// w_fold / x_fold: w / x is an unfolded output of another substep adjoint;
// fold_out: fold the ghost entries of this one's output (else left for the
// reader)
static int adj_substep(const gocean_domain_t* d, float dt, const float* dt_dev, const float* e_in,
                       const Tape& T, const float* w_e, const float* w_u, const float* w_v,
                       const float* w_clamp, const float* x_e, const float* x_u, const float* x_v,
                       const float* x_clamp, float alpha, float gamma, float* o_e, float* o_u, float* o_v,
                       const float* Hi, int w_fold, int x_fold, bool fold_out, hipStream_t s) {
    AdjTileArgs A;
    adj_fill(d, A);
    A.eta = e_in;
    A.tU = T.u; A.tV = T.v; A.tT = T.t; A.tS = T.s; A.tC = T.code;
    A.w_eta = w_e; A.w_hu = w_u; A.w_hv = w_v; A.w_clamp = w_clamp;
    A.x_eta = x_e; A.x_hu = x_u; A.x_hv = x_v; A.x_clamp = x_clamp;
    A.o_eta = o_e; A.o_hu = o_u; A.o_hv = o_v;
    A.Hi = Hi; A.dt = dt; A.dt_dev = dt_dev; A.alpha = alpha; A.gamma = gamma;
    A.w_fold = w_fold; A.x_fold = x_fold;
    return adj_launch(A, fold_out, s);
}

static long long plane_of(const gocean_domain_t* d) { return (long long)(d->ny + 4) * d->pitch; }

}  // namespace gocean

using namespace gocean;

extern "C" {

int gocean_adjoint_set_reverse_kernel(int kind) {
    clear_error();
    if (kind < 0) kind = -1;
    if (kind > 1) return set_error(GOCEAN_EINVAL, "reverse kernel %d: 0 (tile), 1 (column-marching), -1 (by size)", kind);
    g_adj_kernel = kind;
    return 0;
}

int gocean_adjoint_reverse_kernel(void) { return g_adj_kernel; }

long long gocean_adjoint_scratch_floats(const gocean_domain_t* d) {
    if (!d) return -1;
    return 19LL * plane_of(d);
}

int gocean_adjoint_substep(const gocean_domain_t* d, const gocean_wind_t* wind, float dt,
                           const float* dt_dev, const float* eta, const float* hu, const float* hv,
                           const float* w_eta, const float* w_hu, const float* w_hv,
                           const float* w_clamp, const float* x_eta, const float* x_hu,
                           const float* x_hv, const float* x_clamp, float alpha, float gamma,
                           float* o_eta, float* o_hu, float* o_hv, const float* Hi,
                           float* scratch, void* stream) {
    clear_error();
    if (int e = adj_validate(d)) return e;
    if (!eta || !hu || !hv || !w_eta || !w_hu || !w_hv || !o_eta || !o_hu || !o_hv || !Hi || !scratch)
        return set_error(GOCEAN_EINVAL, "NULL argument");
    if ((x_eta || x_hu || x_hv) && !(x_eta && x_hu && x_hv))
        return set_error(GOCEAN_EINVAL, "extra term needs all three fields");
    hipStream_t s = (hipStream_t)stream;
    const long long pl = plane_of(d);
    float* fo = scratch;                           // forward output P(Q): 3 planes
    const Tape T = tape_at(scratch + 3 * pl, pl);   // record: 5 planes
    if (int e = fwd_tape(d, wind, dt, dt_dev, 0, eta, hu, hv, nullptr, nullptr, nullptr, fo, fo + pl,
                         fo + 2 * pl, Hi, T, s))
        return e;
    return adj_substep(d, dt, dt_dev, eta, T, w_eta, w_hu, w_hv, w_clamp ? w_clamp : fo, x_eta, x_hu,
                       x_hv, x_clamp, alpha, gamma, o_eta, o_hu, o_hv, Hi, 0, 0, true, s);
}

int gocean_adjoint_rk2_step(const gocean_domain_t* d, const gocean_wind_t* wind, float dt,
                            const float* dt_dev, const float* eta_n, const float* hu_n, const float* hv_n,
                            const float* eta_s, const float* hu_s, const float* hv_s,
                            const float* eta_next, const float* l_eta, const float* l_hu,
                            const float* l_hv, float* o_eta, float* o_hu, float* o_hv,
                            const float* Hi, float* scratch, void* stream) {
    clear_error();
    if (int e = adj_validate(d)) return e;
    if (!eta_n || !hu_n || !hv_n || !eta_s || !hu_s || !hv_s || !l_eta || !l_hu || !l_hv ||
        !o_eta || !o_hu || !o_hv || !Hi || !scratch)
        return set_error(GOCEAN_EINVAL, "NULL argument");
    if (o_eta == l_eta || o_hu == l_hu || o_hv == l_hv)
        return set_error(GOCEAN_EINVAL, "output adjoint must not alias the input adjoint");
    hipStream_t s = (hipStream_t)stream;
    const long long pl = plane_of(d);
    float* qs = scratch;                     // Q* recomputed (clamp mask of lambda*)
    float* qn1 = scratch + 3 * pl;           // Q^{n+1} (clamp mask of lambda^{n+1})
    float* ls = scratch + 6 * pl;            // lambda*
    const Tape T0 = tape_at(scratch + 9 * pl, pl);
    const Tape T1 = tape_at(scratch + 14 * pl, pl);
    if (int e = fwd_tape(d, wind, dt, dt_dev, 0, eta_n, hu_n, hv_n, nullptr, nullptr, nullptr, qs, qs + pl,
                         qs + 2 * pl, Hi, T0, s))
        return e;
    if (int e = fwd_tape(d, wind, dt, dt_dev, 1, eta_s, hu_s, hv_s, eta_n, hu_n, hv_n, qn1, qn1 + pl,
                         qn1 + 2 * pl, Hi, T1, s))
        return e;
    const float* clamp_next = eta_next ? eta_next : qn1;
    // lambda* = 0.5 * J_P(Q*)^T lambda^{n+1}
    if (int e = adj_substep(d, dt, dt_dev, eta_s, T1, l_eta, l_hu, l_hv, clamp_next, nullptr, nullptr,
                            nullptr, nullptr, 0.5f, 0.0f, ls, ls + pl, ls + 2 * pl, Hi, 0, 0, false, s))
        return e;
    // lambda^n = J_P(Q^n)^T lambda* + 0.5 * lambda^{n+1}
    return adj_substep(d, dt, dt_dev, eta_n, T0, ls, ls + pl, ls + 2 * pl, qs, l_eta, l_hu, l_hv,
                       clamp_next, 1.0f, 0.5f, o_eta, o_hu, o_hv, Hi, 1, 0, true, s);
}

long long gocean_adjoint_segment_work_floats(const gocean_domain_t* d, int n_steps) {
    if (!d || n_steps < 1) return -1;
    return (16LL * n_steps + 9LL) * plane_of(d);
}

// Work layout of a segment of k steps, plane = (ny + 4) * pitch floats:
// Q_1..Q_k (3 planes each), Q*_0..Q*_{k-1} (3 each), branch records of the
// 2k substeps (5 each), lambda scratch (9).
static int seg_check(const gocean_domain_t* d, int n_steps, const void* a, const void* b, const void* c,
                     const void* Hi, const float* work, long long work_floats) {
    if (int e = adj_validate(d)) return e;
    if (n_steps < 1) return set_error(GOCEAN_EINVAL, "n_steps must be >= 1");
    if (!a || !b || !c || !Hi || !work) return set_error(GOCEAN_EINVAL, "NULL argument");
    if (work_floats < gocean_adjoint_segment_work_floats(d, n_steps))
        return set_error(GOCEAN_EINVAL, "work buffer too small: %lld floats, need %lld", work_floats,
                         gocean_adjoint_segment_work_floats(d, n_steps));
    return 0;
}

int gocean_adjoint_record(const gocean_domain_t* d, const gocean_wind_t* wind, int n_steps, float dt,
                          const float* dt_dev, const float* eta0, const float* hu0, const float* hv0,
                          const float* Hi, float* work, long long work_floats, void* stream) {
    clear_error();
    if (int e = seg_check(d, n_steps, eta0, hu0, hv0, Hi, work, work_floats)) return e;
    hipStream_t s = (hipStream_t)stream;
    const long long pl = plane_of(d);
    const int k = n_steps;
    float* Qb = work;
    float* Qs = work + 3LL * k * pl;
    float* Tb = work + 6LL * k * pl;
    auto Qh = [&](int i, int f) -> const float* {
        if (i == 0) return f == 0 ? eta0 : (f == 1 ? hu0 : hv0);
        return Qb + (3LL * (i - 1) + f) * pl;
    };
    // Q*_i = P(Q_i), Q_{i+1} = (Q_i + P(Q*_i)) / 2, with the records
    for (int i = 0; i < k; ++i) {
        float* qs = Qs + 3LL * i * pl;
        float* qn = Qb + 3LL * i * pl;       // Q_{i+1}
        if (int e = fwd_tape(d, wind, dt, dt_dev, 0, Qh(i, 0), Qh(i, 1), Qh(i, 2), nullptr, nullptr, nullptr, qs,
                             qs + pl, qs + 2 * pl, Hi, tape_at(Tb + 10LL * i * pl, pl), s))
            return e;
        if (int e = fwd_tape(d, wind, dt, dt_dev, 1, qs, qs + pl, qs + 2 * pl, Qh(i, 0), Qh(i, 1), Qh(i, 2), qn,
                             qn + pl, qn + 2 * pl, Hi, tape_at(Tb + (10LL * i + 5) * pl, pl), s))
            return e;
    }
    return 0;
}

int gocean_adjoint_reverse(const gocean_domain_t* d, int n_steps, float dt, const float* dt_dev,
                           const float* eta0, const float* l_eta, const float* l_hu, const float* l_hv,
                           float* o_eta, float* o_hu, float* o_hv, const float* Hi, float* work,
                           long long work_floats, void* stream) {
    clear_error();
    if (int e = seg_check(d, n_steps, eta0, l_eta, o_eta, Hi, work, work_floats)) return e;
    if (!l_hu || !l_hv || !o_hu || !o_hv) return set_error(GOCEAN_EINVAL, "NULL argument");
    hipStream_t s = (hipStream_t)stream;
    const long long pl = plane_of(d);
    const int k = n_steps;
    float* Qb = work;
    float* Qs = work + 3LL * k * pl;
    float* Tb = work + 6LL * k * pl;
    float* Lb = work + 16LL * k * pl;
    auto Q = [&](int i) -> const float* { return i == 0 ? eta0 : Qb + 3LL * (i - 1) * pl; };
    float* lam_s = Lb;
    float* lam_a = Lb + 3 * pl;
    float* lam_b = Lb + 6 * pl;
    const float* le = l_eta; const float* lu = l_hu; const float* lv = l_hv;
    for (int i = k - 1; i >= 0; --i) {
        const float* qs = Qs + 3LL * i * pl;
        const float* clamp_next = Q(i + 1);
        float* oe; float* ou; float* ov;
        if (i == 0) { oe = o_eta; ou = o_hu; ov = o_hv; }
        else {
            float* dst = (le == lam_a) ? lam_b : lam_a;
            oe = dst; ou = dst + pl; ov = dst + 2 * pl;
        }
        // lambda^{i+1} is the caller's (folded) for i = k-1, else this
        // loop's unfolded output; only the segment's result is folded
        const int l_unfolded = i < k - 1;
        if (int e = adj_substep(d, dt, dt_dev, qs, tape_at(Tb + (10LL * i + 5) * pl, pl), le, lu, lv,
                                clamp_next, nullptr, nullptr, nullptr, nullptr, 0.5f, 0.0f, lam_s,
                                lam_s + pl, lam_s + 2 * pl, Hi, l_unfolded, 0, false, s))
            return e;
        if (int e = adj_substep(d, dt, dt_dev, Q(i), tape_at(Tb + 10LL * i * pl, pl), lam_s, lam_s + pl,
                                lam_s + 2 * pl, qs, le, lu, lv, clamp_next, 1.0f, 0.5f, oe, ou, ov, Hi,
                                1, l_unfolded, i == 0, s))
            return e;
        le = oe; lu = ou; lv = ov;
    }
    return 0;
}

int gocean_adjoint_segment(const gocean_domain_t* d, const gocean_wind_t* wind, int n_steps, float dt,
                           const float* dt_dev, const float* eta0, const float* hu0, const float* hv0,
                           const float* l_eta, const float* l_hu, const float* l_hv,
                           float* o_eta, float* o_hu, float* o_hv, const float* Hi,
                           float* work, long long work_floats, void* stream) {
    if (int e = gocean_adjoint_record(d, wind, n_steps, dt, dt_dev, eta0, hu0, hv0, Hi, work, work_floats,
                                      stream))
        return e;
    return gocean_adjoint_reverse(d, n_steps, dt, dt_dev, eta0, l_eta, l_hu, l_hv, o_eta, o_hu, o_hv, Hi, work,
                                  work_floats, stream);
}

}  // extern "C"
