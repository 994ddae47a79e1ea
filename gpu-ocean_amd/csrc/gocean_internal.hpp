// gocean_internal.hpp -- declarations shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdarg>

#include "../../include/gocean.h"
#include "cdklm16_tile.hpp"

namespace gocean {

struct TileConfig {
    int bw, bh, nt;
    const char* name;
};

extern const TileConfig kTileConfigs[];
extern const int kNumTileConfigs;
extern const int kDefaultTileConfig;
extern const int kMarchAutoConfig;
// configurations kFirstMarchConfig .. kNumTileConfigs-1 are column marching
constexpr int kFirstMarchConfig = 7;
// configuration used when the caller passes -1
int default_tile_config(long long cells);

int set_error(int code, const char* fmt, ...);
void clear_error();
const char* last_error();
int check_launch(const char* what);
int validate_domain(const gocean_domain_t* d);
void fill_args(const gocean_domain_t* d, SubstepArgs& A);
// true when one residual correction divides every float by d correctly
// rounded (checked once per value of d, on the host)
bool one_correction_exact(float d);
// the substep's FAST arithmetic is exact for these constants: one-correction
// division by dx, dy and 2g, and the Newton square root of g h for h above
// the depth cutoff
bool fast_arith_exact(float g, float dx, float dy, float cutoff);
// the column-marching kernels' 32-bit plane offsets cover a domain of this
// height and pitches
bool march_fits(int ny, int pitch, int hi_pitch);
// column-marching launch shape: waves the device holds at once for a kernel,
// and the rows per strip that fill whole rounds of them
int resident_waves(const void* kernel);
int march_rows(const int* strips, const int* ny, int n, int resident);
int launch_substep(const SubstepArgs& A, int cfg, bool field, hipStream_t s);
// zero-rotation substep that also writes the branch record (A.tape_*)
int launch_substep_tape(SubstepArgs A, hipStream_t s);
int periodic_bc(const gocean_domain_t* d, float* e, float* u, float* v, hipStream_t s);
// CFL reduction: out[0..1] as gocean_max_dt, partial minima (one per
// reduction workgroup, max_dt_partials(d) <= kMaxDtPartials of them) in
// `partial`; max_dt keeps them in out[2..]
constexpr int kMaxDtPartials = 1024;
int max_dt_partials(const gocean_domain_t* d);
int max_dt_into(const gocean_domain_t* d, const float* eta, const float* hu, const float* hv,
                const float* Hm, float courant, float* partial, float* out, hipStream_t s);
int max_dt(const gocean_domain_t* d, const float* eta, const float* hu, const float* hv,
           const float* Hm, float courant, float* out, hipStream_t s);
// substep-adjoint kernel of the reverse entry points (0 tile, 1 marching)
// and the marching kernel's rows per strip for domains of these sizes
extern int g_adj_kernel;
int adj_march_rows(const int* nx, const int* ny, int n);
// the reverse kernel for a launch over this many cells: column marching
// (true) or tile, by g_adj_kernel
bool adj_use_march(long long cells);

}  // namespace gocean
