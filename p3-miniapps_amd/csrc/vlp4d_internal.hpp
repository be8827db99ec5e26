// vlp4d_internal.hpp -- shared declarations of the HIP library (not part of the C-ABI).
//
// Device data layout (all x fastest, like the reference's layout_left views):
//   F    : interior f, N0*N1*N2*N3                         (positions 0..N-1 per dim)
//   CXY  : (N0+4)*(N1+4)*N2*N3, x,y on -2..N+1, v interior  (spline along x and y)
//   TV   : (N0+4)*(N1+4)*(N2+4)*N3, also holds the x-pass output (N0+4)*N1*N2*N3
//   C4   : positions -3..N+2 along y, vx, vy (the coefficients fill -2..N+1,
//          the outer layer stays 0) and -2..N+1 along x, in rows of N0+4:
//          position (p0, p1, p2, p3) at (p0+3) + s1 (p1+3) + s2 (p2+3) +
//          s3 (p3+3) with the pitches of c4_pitch.  The x layer -3 / N0+2
//          has no storage of its own (it overlaps the neighbouring row):
//          feet stay within [x_min - dx, x_max], so a stencil reaches N0+2
//          only from a foot on x_max, where that point's weight is 0, and
//          -3 only through rounding at x_min - dx, with a weight ~1e-16.  Rows without x padding make the coefficients of
//          one (vx, vy) contiguous, so the vy spline pass writes whole
//          64-byte blocks (planes padded to 16 elements, base aligned by
//          the library).
// The reference's (N+6)^4 halo arrays of f are never materialised: the
// periodic copies and alpha-sums of its halo exchange are computed inside the
// spline kernels from the interior line they extend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "../../include/vlp4d.h"

namespace vlp {

// Pitches of the C4 storage (elements): row (y), plane (vx), volume (vy).
struct C4Pitch {
  long s1, s2, s3;
};
__host__ __device__ inline C4Pitch c4_pitch(long N0, long N1, long N2) {
  // + 1: the last row's N0+2 slot stays inside the plane's padding
  const long s1 = N0 + 4, s2 = (s1 * (N1 + 6) + 1 + 15) / 16 * 16;
  return C4Pitch{s1, s2, s2 * (N2 + 6)};
}
// Elements of the C4 storage including a zeroed tail.
inline long c4_elems(long N0, long N1, long N2, long N3) { return c4_pitch(N0, N1, N2).s3 * (N3 + 6) + 64; }

// Host helpers of vlp4d_api.cpp: record an error for vlp_last_error(NULL)
// (returns code); fill dx from the box and check the configuration the way
// vlp_create does (VLP_OK or an error code, already recorded).
int set_error(int code, const char *msg);
int prepare_config(vlp_config *cfg);

// Everything a kernel needs from the configuration (one rank).
struct Geom {
  int n[4];        // points per dimension
  double minPhy[4];
  double maxPhy[4];
  double dx[4];
  double inv_dx[4];  // 1./dx, VMO/Advection.hpp:319,430
  int idcase;
};

// Addressing of a strided-line spline pass.  Thread index t -> (i0, i1, o)
// with i0 fastest; the line of N interior values starts at
// in + i0*is0 + i1*is1 + o*ios with stride ils; the N+4 coefficients
// (positions -2..N+1) go to out + i0*os0 + i1*os1 + o*oos with stride ols.
struct LineMap {
  int N;
  long X0, X1, A;
  long is0, is1, ios, ils;
  long os0, os1, oos, ols;
};

// Launchers.  T is the storage type (double or float); arith32 selects float
// interpolation weights and accumulation for float storage (else double).
template <typename T> void launch_init_case(const Geom &g, T *f, hipStream_t s);
template <typename T> void launch_spline_x(const Geom &g, const T *f, T *tx, hipStream_t s);
// Kernel choice of a strided spline pass: SPLINE_AUTO the preference order
// (128-position LDS tiles where N is a multiple of 32, else 64-position
// tiles, else the line kernels), SPLINE_TILE64 64-position tiles (or the line
// kernels), SPLINE_LINES the line kernels, SPLINE_TILE256 256-position tiles
// where they fit in LDS (else the AUTO choice).  The VLP4D_SPLINE_* environment
// variables override it.
enum { SPLINE_AUTO = 0, SPLINE_TILE64 = 1, SPLINE_LINES = 2, SPLINE_TILE256 = 3 };
template <typename T>
void launch_spline_line(const T *in, T *out, const LineMap &m, hipStream_t s, int mode = SPLINE_AUTO);
// The strided spline passes as MFMA matrix products (spline_mfma.hip):
// OUT = M IN per outer index, M the exact (N+4) x N line operator, row-major
// with Kp = spline_mfma_kp(N) columns, spline_mfma_matrix_elems(N) values
// (rows padded to a multiple of 16), built by build_spline_mfma_matrix from
// the double line kernel (work: N*N + spline_mfma_matrix_elems(N) doubles).
bool spline_mfma_eligible(const LineMap &m, size_t tsz);
template <typename T> void launch_spline_mfma(const T *in, T *out, const LineMap &m, const T *Mm, int Kp, hipStream_t s);
int spline_mfma_kp(int N);
size_t spline_mfma_matrix_elems(int N);
template <typename T> void build_spline_mfma_matrix(int N, T *Mm, double *work, hipStream_t s);
// Tile width (64 or 128 positions) of the LDS-tiled kernel launch_spline_line
// picks for this line family and storage size, 0 when it takes a line kernel.
int spline_tile_width(const LineMap &m, size_t tsz, int mode = SPLINE_AUTO);
template <typename T>
void launch_advect2d_rho(const Geom &g, const T *cxy, double dt_half, bool arith32, double *rho_part, int *err,
                        hipStream_t s);
// Returns true when rho and the L2 norm of fout are still to be computed
// (launch_moments); false when the kernel produced them.
template <typename T>
bool launch_advect4d(const Geom &g, const T *c4, const double *ex, const double *ey, const double *exy, double dt,
                     bool arith32, T *fout, double *rho_part, double *l2_part, int *err, hipStream_t s,
                     int ys_tvx = 0, int ys_nb = 0);
// Fused spline x/y + 2-D half-step advection + density partials (xy_fused.hip):
// f -> CXY and rho_part[grp][vy][y][x] for grp < G.  Returns false without
// launching when the geometry does not fit; xy_fused_groups gives G (0: the
// separate x, y and advect2d kernels run instead).
template <typename T, typename A>
bool launch_xy_adv2d(const Geom &g, const T *f, T *cxy, double dt_half, double *rho_part, int *err, hipStream_t s);
int xy_fused_groups(const Geom &g, size_t tsz);

// y-streaming 4-D advection (f only; rho and L2 by launch_moments afterwards).
// Returns false without launching when the geometry does not fit it.
template <typename T, typename A>
// ys_tvx / ys_nb (0: the preference order) select the vx tile and LDS plane
// buffers; the environment overrides both.
bool launch_advect4d_ystream(const Geom &g, const T *c4, const double *exy, double dt, T *fout, int *err,
                             hipStream_t s, int ys_tvx = 0, int ys_nb = 0);
// Whether the streaming kernel takes (tvx, nb) for this geometry and storage
// size (the autotuner's candidate filter).
bool ystream_eligible(const Geom &g, size_t tsz, int tvx, int nb);
// Whether (tvx, nb) is the preference order's choice for this geometry (the
// autotuner's incumbent).
bool ystream_preferred(const Geom &g, size_t tsz, int tvx, int nb);
// vx tile of the y-streaming kernel for a geometry (0: not used).
int ystream_tvx(const Geom &g, bool fp32_storage, bool arith32);
template <typename T>
void launch_moments(const Geom &g, const T *f, double *rho_part, double *l2_part, hipStream_t s);

// Number of l2 partial sums written by launch_advect4d / launch_moments.
long l2_partials(const Geom &g);

// Kernel variants chosen for a geometry and precision mode (for tests and
// reports): bit 0: x-pass uses the multi-lane scan kernel, bit 1: strided
// passes use the multi-lane scan kernel for all three dims, bit 2: 4-D
// advection uses the vx-blocked kernel, bit 3: 4-cell blocks (else 2), bit 4:
// E rows staged in LDS, bit 5: XCD-aware workgroup order, bit 6: float
// advection arithmetic, bit 7: the y-streaming 4-D advection replaces the
// vx-blocked one, bits 8-11: its vx tile, bit 12: the fused xy spline + 2-D
// advection kernel replaces the x, y and advect2d kernels, bits 13 / 14: the
// vx / vy spline pass runs 128-position tiles (four lanes per 128-point line),
// bit 15: the vx and vy passes run as MFMA matrix products (VLP4D_SPLINE_MFMA=1).
int kernel_variants(const Geom &g, bool fp32_storage, bool arith32);

// Adjoint of one step (adjoint.hip), FP64.
// Lagrange split-step vlp4d (lagrange.hip), FP64, arrays without halo.
void launch_lag_init(const Geom &g, double *f, hipStream_t s);
void launch_lag_x(const Geom &g, const double *in, double *out, double dt, hipStream_t s);
void launch_lag_y(const Geom &g, const double *in, double *out, double dt, double *rho_part, hipStream_t s);
void launch_lag_v(const Geom &g, int d, const double *in, double *out, const double *e, double dt, hipStream_t s);
// Fused x/2 + y/2 (one pass over f) when the (x, y) plane fits in LDS:
// lag_xy_chunks gives the vx runs per vy, 0 when not eligible;
// launch_lag_xy returns false for nchunk 0.
// The stencil tables of the x/2, y/2 passes (lag_tables_bytes) are built once
// per context for the half step dt.
int lag_xy_chunks(const Geom &g);
size_t lag_tables_bytes(const Geom &g);
void launch_lag_tables(const Geom &g, void *tab, double dt, hipStream_t s);
bool launch_lag_xy(const Geom &g, const double *in, double *out, const void *tab, int nchunk, hipStream_t s);

// Work buffers of the reproducible adjoint scatter: scl (4 doubles: the
// fixed-point scale S, 1/S, and the max|g| bits), axy (2 per cell), part
// (2 N0 N1 N3).
struct AdjWork {
  double *scl, *axy, *part;
};
// VLP4D_ADJ_BLOCK unset or nonzero: the reproducible blocked forms
bool adj_blocked();
// With wk: lc4 receives 64-bit fixed-point sums (convert with
// launch_c4_block(..., wk->scl)) and lE is written (not added to); without:
// floating-point atomics into both.
void launch_adj_advect4d(const Geom &g, const double *c4, const double *ex, const double *ey, double dt,
                         const double *gin, double *lc4, double *lE, bool field_term, hipStream_t s,
                         const AdjWork *wk = nullptr);
// out[o][n][i] (+)= sum_m Mt[n][m] in[o][m][i]; with M (= Mt^T, Min x Nout
// row-major) given, by LDS tiles of 64 columns
void launch_axis_contract(const double *Mt, int Min, int Nout, const double *in, double *out, long O, long I,
                          bool accumulate, hipStream_t s, const double *M = nullptr);
void launch_c4_block(const Geom &g, const double *c4s, double *blk, hipStream_t s, const double *scl = nullptr);
void launch_poisson_T(const double *lE, const double *kx, const double *ky, double *lrho, int N0, int N1,
                      hipStream_t s);
void launch_adj_advect2d(const double *lrho, const Geom &g, double dth, double *lcxy, hipStream_t s);
// Field-energy adjoint helpers: lE = scale * exy (n doubles); g(x, y, v) +=
// scale * lrho(x, y) over ncell cells; out(x, y) = sum_v w(v) in(x, y, v).
void launch_scale_pairs(const double *exy, double scale, long n, double *lE, hipStream_t s);
void launch_add_density_adj(const double *lrho, double scale, long nxy, long ncell, double *g, hipStream_t s);
void launch_vmoment(const double *in, const double *w, long nxy, long nv, double *out, hipStream_t s);

// Field solve (FP64): rho from the per-vy partial sums, Poisson by DFT with
// the reference's FFT conventions, field diagnostics; optional L2 norm.
struct FieldBufs {
  double *rho;     // N0*N1 : rho, then phi (Poisson cases)
  double *ex, *ey; // N0*N1
  double *exy;     // N0*N1 (Ex, Ey) pairs, the layout the streaming advection stages
  double *spec1;   // (N0/2+1)*N1 complex: r2c along x
  double *spec2;   // 3 * (N0/2+1)*N1 complex: inverse along y (phi, ex, ey)
  double *twx, *twy;  // exp(-2 pi i m/N) tables, complex
  double *diag_part;  // 4*N1
  double *rho_tmp;    // RHO_SLICES*N0*N1: slice sums of many density partials
  double *diag;       // 5*(nbiter+1): nrj, nrjx, nrjy, mass, l2
  int nbiter;
};
// rho_outer: number of N0*N1 density-partial planes in rho_part (N3, or
// G*N3 after the fused xy kernel).  Beyond 64 planes they are first summed in
// RHO_SLICES slices by a wide kernel.
constexpr int RHO_SLICES = 16;
// iter_dev (optional): device-side iteration counter for graph replays; the
// diagnostics then go to index *iter_dev + 1, which is stored back.
void launch_field(const Geom &g, const FieldBufs &b, const double *rho_part, long rho_outer, const double *l2_part,
                  long n_l2, int iter, bool record_diag, hipStream_t s, int *iter_dev = nullptr);

// Field solve of a caller-held f (field_solve.hip): field_solve_geom checks
// the arguments, fills g and the scratch size in doubles; field_solve_enqueue
// launches the moments, twiddle and field kernels on s with that scratch.
int field_solve_geom(const vlp_config *cfg, const void *f, int f_bits, double *phi, double *ex, double *ey,
                     double *diag, Geom &g, size_t &scratch_doubles);
int field_solve_enqueue(const Geom &g, const void *f, int f_bits, double *phi, double *ex, double *ey,
                        double *diag, double *scratch, hipStream_t s);

}  // namespace vlp
