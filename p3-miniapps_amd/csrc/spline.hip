// spline.hip -- spline prefilter kernels (gfx950).
//
// Reference: Spline::computeCoeff (miniapps/vlp4d_mpi/openmp/Spline.hpp:114-199)
// on lines extended by the one-rank halo exchange with the spline boundary
// condition (VMO/Communication.hpp:713-898, VMO/Communication.cpp:541-570):
//   g[-3]      = sum_{j=1..8} alpha^j f[N-2-j]   (boundary_condition, min side)
//   g[-2..N+1] = f[i mod N]                      (periodic copy)
//   g[N+2]     = sum_{j=1..8} alpha^j f[1+j]     (boundary_condition, max side)
//   causal      cp[-2] = g[-3]+g[-2],  cp[n] = g[n] + alpha cp[n-1]
//   anticausal  c[N+1] = sqrt3 (g[N+2] + g[N+1] + sum_{k=1}^{N+3} alpha^k g[N+1-k])
//               c[n]   = beta cp[n] + alpha c[n+1],  n = N..-2
// The halo never exists in memory: each kernel reads the N interior values of
// a line and writes its N+4 coefficients.  Arithmetic is in double for both
// storage precisions.
//
// Kernels:
//   k_spline_x_scan   lines along the contiguous x axis; a group of P lanes
//                     splits each line into chunks of C and runs both
//                     recursions as a blocked linear-recurrence scan.
//   k_spline_line_scan lines along a strided axis with the same chunked scan;
//                     groups of lanes walk the contiguous axis, so every load
//                     and store of a wave is a few contiguous segments.
//   k_spline_x / k_spline_line  single-lane LDS fallbacks for sizes without a
//                     matching instantiation.
#include "spline_scan.hpp"

#include <cstdlib>

namespace vlp {

#pragma clang fp contract(off)
SplineConsts spline_consts() {
  SplineConsts c;
  c.sqrt3 = std::sqrt(3.0);
  c.alpha = c.sqrt3 - 2;
  c.beta = c.sqrt3 * (1 - c.alpha * c.alpha);
  return c;
}
#pragma clang fp contract(on)

// ---------------------------------------------------------------------------
// Generic single-lane recursion on a strided buffer (LDS fallbacks).
// buf[(i+2)*st] holds f[i] (i = 0..N-1) on entry and c[i] (i = -2..N+1) on exit.
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ void spline_recursion(T *buf, int st, int N, const SplineConsts sc) {
  const double a = sc.alpha, b = sc.beta, s3 = sc.sqrt3;
#define FV(i) ((double)buf[((i) + 2) * st])
  double gm3 = 0., gp2 = 0., ap = a;
  for (int j = 1; j <= 8; j++) {  // MMAX = 8, VMO/Config.hpp:8
    gm3 += FV(N - 2 - j) * ap;
    gp2 += FV(1 + j) * ap;
    ap *= a;
  }
  const double f0 = FV(0), f1 = FV(1);
  double fnend = gp2 + f1;  // VMO/Spline.hpp:144-149
  double ak = a;
  fnend += f0 * ak;  // g[N] = f[0]
  ak *= a;
  for (int k = 2; k <= N + 1; k++) {  // g[N-1] .. g[0]
    fnend += FV(N + 1 - k) * ak;
    ak *= a;
  }
  fnend += FV(N - 1) * ak;  // g[-1]
  ak *= a;
  fnend += FV(N - 2) * ak;  // g[-2]
  double cp = gm3 + FV(N - 2);  // VMO/Spline.hpp:132-135
  buf[0] = (T)cp;
  cp = FV(N - 1) + a * cp;
  buf[st] = (T)cp;
  for (int p = 0; p < N; p++) {
    cp = FV(p) + a * cp;
    buf[(p + 2) * st] = (T)cp;
  }
  cp = f0 + a * cp;
  buf[(N + 2) * st] = (T)cp;
  double c = s3 * fnend;  // VMO/Spline.hpp:151-155
  buf[(N + 3) * st] = (T)c;
  for (int p = N; p >= -2; p--) {
    c = b * (double)buf[(p + 2) * st] + a * c;
    buf[(p + 2) * st] = (T)c;
  }
#undef FV
}

__device__ __forceinline__ void line_coords(long t, const LineMap &m, long &in_off, long &out_off) {
  const long r = t / m.X0;
  const long i0 = t - r * m.X0;
  const long o = r / m.X1;
  const long i1 = r - o * m.X1;
  in_off = i0 * m.is0 + i1 * m.is1 + o * m.ios;
  out_off = i0 * m.os0 + i1 * m.os1 + o * m.oos;
}

// x-pass: workgroup of 256 lanes = 256/P lines, staged through LDS with
// coalesced loads and stores; the recursion runs as scan_recursion.
template <typename T, int C>
__global__ __launch_bounds__(256) void k_spline_x_scan(const T *__restrict__ f, T *__restrict__ tx, int P, int logP,
                                                       long nlines, SplineConsts sc, double aC, double aN3) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int N = P * C, M = N + 4;
  const int R = M | 1;
  const int lpw = 256 >> logP;
  const long l0 = (long)blockIdx.x * lpw;
  const int nl = (int)((nlines - l0) < lpw ? (nlines - l0) : lpw);
  const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
  for (int l = wave; l < nl; l += 4) {
    const T *src = f + (l0 + l) * N;
    for (int x = lane; x < N; x += 64) lds[l * R + x + 2] = src[x];
  }
  __syncthreads();
  const int li = threadIdx.x >> logP;
  const int p = threadIdx.x & (P - 1);
  const bool act = li < nl;
  T *row = lds + (act ? li : 0) * R;
  double v[C];
#pragma unroll
  for (int i = 0; i < C; i++) v[i] = (double)row[2 + p * C + i];
  double cm2, cm1, cN, cN1;
  scan_recursion<C>(v, p, P, sc, aC, aN3, cm2, cm1, cN, cN1);
  // each lane rewrites only the slots it read (plus the empty edge slots)
  if (act) {
#pragma unroll
    for (int i = 0; i < C; i++) row[2 + p * C + i] = (T)v[i];
    if (p == 0) {
      row[0] = (T)cm2;
      row[1] = (T)cm1;
    }
    if (p == P - 1) {
      row[N + 2] = (T)cN;
      row[N + 3] = (T)cN1;
    }
  }
  __syncthreads();
  for (int l = wave; l < nl; l += 4) {
    T *dst = tx + (l0 + l) * M;
    for (int x = lane; x < M; x += 64) dst[x] = lds[l * R + x];
  }
}

// x-pass fallback: 64 lanes, 64 lines in LDS, one lane per line.
template <typename T, int LINES>
__global__ __launch_bounds__(64) void k_spline_x(const T *__restrict__ f, T *__restrict__ tx, int N, long nlines,
                                                 SplineConsts sc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const int P = (N + 4) | 1;  // odd row stride: conflict-free column access
  const long l0 = (long)blockIdx.x * LINES;
  const int nl = (int)((nlines - l0) < LINES ? (nlines - l0) : LINES);
  const T *src = f + l0 * N;
  for (int l = 0; l < nl; l++)
    for (int x = threadIdx.x; x < N; x += 64) lds[l * P + x + 2] = src[(long)l * N + x];
  __syncthreads();
  if ((int)threadIdx.x < nl) spline_recursion(lds + threadIdx.x * P, 1, N, sc);
  __syncthreads();
  const int M = N + 4;
  T *dst = tx + l0 * M;
  for (int l = 0; l < nl; l++)
    for (int x = threadIdx.x; x < M; x += 64) dst[(long)l * M + x] = lds[l * P + x];
}

// (P, C) with N = P*C, P a power of two <= 16, C in [10, 16]; C = 0 if none.
static void x_split(int N, int &P, int &logP, int &C) {
  for (logP = 4; logP >= 0; logP--) {
    P = 1 << logP;
    if (N % P) continue;
    C = N / P;
    if (C >= 10 && C <= 16) return;
  }
  P = 1;
  logP = 0;
  C = 0;
}

template <typename T, int C>
static void launch_x_scan(const T *f, T *tx, int P, int logP, long nlines, hipStream_t s) {
  const int N = P * C;
  const int lpw = 256 >> logP;
  const size_t lds = (size_t)lpw * ((N + 4) | 1) * sizeof(T);
  const SplineConsts sc = spline_consts();
  double aC = 1.0, aN3 = 1.0;
  for (int i = 0; i < C; i++) aC *= sc.alpha;
  for (int i = 0; i < N + 3; i++) aN3 *= sc.alpha;
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)k_spline_x_scan<T, C>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    attr = true;
  }
  const long nblk = (nlines + lpw - 1) / lpw;
  hipLaunchKernelGGL((k_spline_x_scan<T, C>), dim3((unsigned)nblk), dim3(256), lds, s, f, tx, P, logP, nlines, sc, aC,
                     aN3);
}

template <typename T> void launch_spline_x(const Geom &g, const T *f, T *tx, hipStream_t s) {
  const int N = g.n[0];
  const long nlines = (long)g.n[1] * g.n[2] * g.n[3];
  int P, logP, C;
  x_split(N, P, logP, C);
  switch (C) {
    case 10: return launch_x_scan<T, 10>(f, tx, P, logP, nlines, s);
    case 11: return launch_x_scan<T, 11>(f, tx, P, logP, nlines, s);
    case 12: return launch_x_scan<T, 12>(f, tx, P, logP, nlines, s);
    case 13: return launch_x_scan<T, 13>(f, tx, P, logP, nlines, s);
    case 14: return launch_x_scan<T, 14>(f, tx, P, logP, nlines, s);
    case 15: return launch_x_scan<T, 15>(f, tx, P, logP, nlines, s);
    case 16: return launch_x_scan<T, 16>(f, tx, P, logP, nlines, s);
    default: break;
  }
  constexpr int LINES = 64;
  const size_t lds = (size_t)LINES * ((N + 4) | 1) * sizeof(T);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)k_spline_x<T, LINES>, hipFuncAttributeMaxDynamicSharedMemorySize,
                              160 * 1024);
    attr = true;
  }
  const long nblk = (nlines + LINES - 1) / LINES;
  hipLaunchKernelGGL((k_spline_x<T, LINES>), dim3((unsigned)nblk), dim3(64), lds, s, f, tx, N, nlines,
                     spline_consts());
}

// ---------------------------------------------------------------------------
// Strided passes.
// ---------------------------------------------------------------------------

// Strided lines: groups of P lanes per line (lane p owns chunk p), groups
// walk the contiguous inner index, so a wave's access for chunk element i is
// 64/P consecutive elements of each of P chunk positions.
template <typename T, int C>
__global__ __launch_bounds__(256) void k_spline_line_scan(const T *__restrict__ in, T *__restrict__ out, LineMap m,
                                                          int P, int logP, SplineConsts sc, double aC, double aN3) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long grp = t >> logP;
  const int p = (int)(t & (P - 1));
  if (grp >= m.X0 * m.X1 * m.A) return;  // whole groups leave together
  long ioff, ooff;
  line_coords(grp, m, ioff, ooff);
  const long ils = m.ils, ols = m.ols;
  const T *src = in + ioff + (long)p * C * ils;
  double v[C];
#pragma unroll
  for (int i = 0; i < C; i++) v[i] = (double)src[i * ils];
  double cm2, cm1, cN, cN1;
  scan_recursion<C>(v, p, P, sc, aC, aN3, cm2, cm1, cN, cN1);
  T *dst = out + ooff;
  T *dc = dst + ((long)p * C + 2) * ols;
#pragma unroll
  for (int i = 0; i < C; i++) dc[i * ols] = (T)v[i];
  const int N = P * C;
  if (p == 0) {
    dst[0] = (T)cm2;
    dst[ols] = (T)cm1;
  }
  if (p == P - 1) {
    dst[(long)(N + 2) * ols] = (T)cN;
    dst[(long)(N + 3) * ols] = (T)cN1;
  }
}

// Strided lines through an LDS tile: the workgroup takes XT consecutive
// contiguous-index positions i0 (a tile column each) of one outer index and
// the whole line along the strided axis.  The N input rows of the tile (XT
// values each, contiguous in memory) arrive by LDS-DMA, 16 bytes per lane;
// the chunked-scan recursion runs down the tile columns (P lanes per column)
// in place, leaving the N+4 coefficient rows, which go out as contiguous
// segments.  Used when the map's i0 index is contiguous (m.is0 == m.os0 ==
// 1); i1 and the outer index o select the tile's line family.  When the input
// is contiguous across i1 as well (is1 == X0) the tiles run over the combined
// index i0 + X0 i1 and each tile column finds its output position once.
typedef __attribute__((address_space(3))) void lds_void_t;
// 16-byte chunks: aligned in LDS, element-aligned in global memory
typedef float f4a __attribute__((ext_vector_type(4)));
typedef double d2a __attribute__((ext_vector_type(2)));
typedef float f4e __attribute__((ext_vector_type(4), aligned(4)));
typedef double d2e __attribute__((ext_vector_type(2), aligned(8)));

template <typename T, int C, int XT>
__global__ __launch_bounds__(XT == 256 ? 1024 : 512) void k_spline_tile(const T *__restrict__ in, T *__restrict__ out, LineMap m, int P,
                                                     SplineConsts sc, double aC, double aN3) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *tile = reinterpret_cast<T *>(smem_raw);  // [N+4][XT]
  constexpr int VPL = 16 / (int)sizeof(T), CPR = XT / VPL;  // 16-byte chunks per row
  static_assert(CPR <= 64, "a tile row must fit one 64-lane DMA instruction");
  const int N = P * C;
  const bool comb = m.is1 == m.X0;
  const long XC = comb ? m.X0 * m.X1 : m.X0;  // contiguous input extent
  const long ntile = (XC + XT - 1) / XT;
  // consecutive tiles, whose rows share memory pages, on one XCD: blocks go
  // to the 8 XCDs round-robin, so block b takes logical tile (b mod 8) G/8 +
  // b/8 over the first 8 floor(G/8) blocks (the rest in order); each XCD's
  // address-translation cache then serves a run of neighbouring tiles
  const unsigned G8 = gridDim.x & ~7u, bk = blockIdx.x;
  const long lb = bk < G8 ? (long)(bk & 7u) * (G8 >> 3) + (bk >> 3) : (long)bk;
  const long tl = lb % ntile, ob = lb / ntile;
  const long i1 = comb ? 0 : ob % m.X1, o = comb ? ob : ob / m.X1;
  const long i0 = tl * XT;
  const int nx = (int)((XC - i0) < XT ? (XC - i0) : XT);
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nw = blockDim.x >> 6;
  const T *src = in + i1 * m.is1 + o * m.ios + i0;
  // rows k = 0..N-1 into tile rows 2..N+1: 64 lanes cover 64/CPR rows per DMA
  {
    const int rpi = 64 / CPR;  // rows per instruction
    const int r = lane / CPR, c = lane % CPR;
    for (int k0 = wave * rpi; k0 < N; k0 += nw * rpi) {
      const int k = k0 + r;
      if (k < N && c * VPL < nx)
        __builtin_amdgcn_global_load_lds((const void *)(src + (long)k * m.ils + c * VPL),
                                         (lds_void_t *)(tile + (long)(k0 + 2) * XT), 16, 0, 0);
    }
  }
  __syncthreads();
  // recursion down column col = tid / P, lane p = tid % P of the column group
  {
    const int col = tid / P, p = tid & (P - 1);
    const bool act = col < nx;
    T *cp = tile + (act ? col : 0);
    double v[C];
#pragma unroll
    for (int i = 0; i < C; i++) v[i] = (double)cp[(2 + p * C + i) * XT];
    double cm2, cm1, cN, cN1;
    scan_recursion<C>(v, p, P, sc, aC, aN3, cm2, cm1, cN, cN1);
    __syncthreads();  // all reads of the input rows done before rows are rewritten
    if (act) {
#pragma unroll
      for (int i = 0; i < C; i++) cp[(2 + p * C + i) * XT] = (T)v[i];
      if (p == 0) {
        cp[0] = (T)cm2;
        cp[XT] = (T)cm1;
      }
      if (p == P - 1) {
        cp[(N + 2) * XT] = (T)cN;
        cp[(N + 3) * XT] = (T)cN1;
      }
    }
  }
  __syncthreads();
  // N+4 coefficient rows out.  When the contiguous extent X0 splits into
  // whole 16-byte chunks, no chunk straddles two runs of X0, so a chunk stays
  // contiguous in the output: a lane keeps one chunk column of the tile and
  // writes it down the rows, one 16-byte LDS read and one multi-dword store
  // per row (output rows are only element-aligned, which global stores
  // allow).  Else element by element.
  T *dst = out + i1 * m.os1 + o * m.oos;
  if (m.X0 % VPL == 0) {
    const int c = tid % CPR, rpp = (int)blockDim.x / CPR;
    if (c * VPL < nx) {
      const long J = i0 + (long)c * VPL;
      const long ox = comb ? (J % m.X0) + (J / m.X0) * m.os1 : J;
      for (int k = tid / CPR; k < N + 4; k += rpp) {
        const T *s = tile + k * XT + c * VPL;
        T *d = dst + (long)k * m.ols + ox;
        if constexpr (VPL == 4) {
          *(f4e *)d = *(const f4a *)s;
        } else {
          *(d2e *)d = *(const d2a *)s;
        }
      }
    }
  } else {
    for (int x = lane; x < nx; x += 64) {
      const long J = i0 + x;
      const long ox = comb ? (J % m.X0) + (J / m.X0) * m.os1 : J;
      for (int k = wave; k < N + 4; k += nw) dst[(long)k * m.ols + ox] = tile[k * XT + x];
    }
  }
}

// LDS fallback: one lane per line, the line in an LDS column of the lane.
template <typename T>
__global__ __launch_bounds__(64) void k_spline_line(const T *__restrict__ in, T *__restrict__ out, LineMap m,
                                                    SplineConsts sc) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T *lds = reinterpret_cast<T *>(smem_raw);
  const long t = (long)blockIdx.x * 64 + threadIdx.x;
  if (t >= m.X0 * m.X1 * m.A) return;
  long ioff, ooff;
  line_coords(t, m, ioff, ooff);
  const T *src = in + ioff;
  T *buf = lds + threadIdx.x;
  for (int k = 0; k < m.N; k++) buf[(k + 2) * 64] = src[k * m.ils];
  spline_recursion(buf, 64, m.N, sc);
  T *dst = out + ooff;
  for (int k = 0; k < m.N + 4; k++) dst[k * m.ols] = buf[k * 64];
}

// (P, C) with N = P*C, P a power of two <= 16 and C in [10, cmax], smallest P
// first; C = 0 when there is none.
static void line_split(int N, int cmax, int &P, int &logP, int &C) {
  for (logP = 0; logP <= 4; logP++) {
    P = 1 << logP;
    if (N % P) continue;
    C = N / P;
    if (C >= 10 && C <= cmax) return;
  }
  P = 1;
  logP = 0;
  C = 0;
}

template <typename T, int C>
static void launch_line_scan(const T *in, T *out, const LineMap &m, int P, int logP, hipStream_t s) {
  const SplineConsts sc = spline_consts();
  double aC = 1.0, aN3 = 1.0;
  for (int i = 0; i < C; i++) aC *= sc.alpha;
  for (int i = 0; i < P * C + 3; i++) aN3 *= sc.alpha;
  const long nthr = m.X0 * m.X1 * m.A * P;
  hipLaunchKernelGGL((k_spline_line_scan<T, C>), dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, s, in, out, m, P,
                     logP, sc, aC, aN3);
}

#define VLP_LINE_CASE(CC) \
  case CC: return launch_line_scan<T, CC>(in, out, m, P, logP, s);

template <typename T, int C, int XT>
static void launch_tile(const T *in, T *out, const LineMap &m, int P, hipStream_t s) {
  const SplineConsts sc = spline_consts();
  double aC = 1.0, aN3 = 1.0;
  for (int i = 0; i < C; i++) aC *= sc.alpha;
  for (int i = 0; i < P * C + 3; i++) aN3 *= sc.alpha;
  const bool comb = m.is1 == m.X0;
  const long XC = comb ? m.X0 * m.X1 : m.X0;
  const long ntile = (XC + XT - 1) / XT;
  const size_t lds = (size_t)(P * C + 4) * XT * sizeof(T);
  const void *fn = (const void *)k_spline_tile<T, C, XT>;
  if (lds > 64 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  hipLaunchKernelGGL((k_spline_tile<T, C, XT>), dim3((unsigned)(ntile * (comb ? 1 : m.X1) * m.A)), dim3(XT * P), lds,
                     s, in, out, m, P, sc, aC, aN3);
}

// Tile kernel when the contiguous index is i0 alone and the line splits into
// P * C with C in [10, 16] and 64-position tiles (P * 64 <= 512 lanes), or --
// wide tiles, N a multiple of 32 -- C = 32 and 128-position tiles (P * 128 <=
// 512), the tile in LDS.  Wide tiles move twice the bytes per page a tile
// touches: each tile row sits on its own memory page when the line stride is
// large, and the address translations per byte halve.
static int tile_chunk(const LineMap &m, size_t tsz, int &P, int &XT, int mode = SPLINE_AUTO) {
  if (mode == SPLINE_LINES) return 0;
  if (const char *e = getenv("VLP4D_SPLINE_TILE"))
    if (atoi(e) == 0) return 0;
  if (m.is0 != 1 || m.os0 != 1) return 0;
  // VLP4D_SPLINE_WIDE=0: 64-position tiles only; =2: 256-position tiles
  // where they fit (float storage at N = 128: 135 KB of LDS, 1024 lanes)
  int wide = mode == SPLINE_TILE64 ? 0 : mode == SPLINE_TILE256 ? 2 : 1;
  if (const char *e = getenv("VLP4D_SPLINE_WIDE")) wide = atoi(e);
  // (a tile row is moved by one DMA instruction of 64 lanes x 16 bytes, so a
  // 256-position row needs 4-byte elements)
  if (wide == 2 && tsz * 256 <= 1024 && m.N % 32 == 0 && (m.N / 32) * 256 <= 1024 &&
      ((m.N / 32) & (m.N / 32 - 1)) == 0 &&
      (size_t)(m.N + 4) * 256 * tsz <= 160 * 1024 && (m.ils * tsz) % 16 == 0 && (m.is1 * tsz) % 16 == 0 &&
      (m.ios * tsz) % 16 == 0 && ((m.is1 == m.X0 ? m.X0 * m.X1 : m.X0) * (long)tsz) % 16 == 0) {
    P = m.N / 32;
    XT = 256;
    return 32;
  }
  if (wide && m.N % 32 == 0 && (m.N / 32) * 128 <= 512 && ((m.N / 32) & (m.N / 32 - 1)) == 0 &&
      (size_t)(m.N + 4) * 128 * tsz <= 160 * 1024 && (m.ils * tsz) % 16 == 0 && (m.is1 * tsz) % 16 == 0 &&
      (m.ios * tsz) % 16 == 0 && ((m.is1 == m.X0 ? m.X0 * m.X1 : m.X0) * (long)tsz) % 16 == 0) {
    P = m.N / 32;
    XT = 128;
    return 32;
  }
  XT = 64;
  int logP, C;
  x_split(m.N, P, logP, C);
  if (!C || P * 64 > 512) return 0;
  if ((size_t)(m.N + 4) * 64 * tsz > 160 * 1024) return 0;
  // input rows start 16-byte aligned (LDS-DMA of whole 16-byte chunks; the
  // buffers themselves are 256-byte aligned)
  if ((m.ils * tsz) % 16 || (m.is1 * tsz) % 16 || (m.ios * tsz) % 16) return 0;
  const long XC = m.is1 == m.X0 ? m.X0 * m.X1 : m.X0;
  if ((XC * (long)tsz) % 16) return 0;  // the last tile's chunks end inside the line family
  return C;
}

template <typename T> void launch_spline_line(const T *in, T *out, const LineMap &m, hipStream_t s, int mode) {
  int P, logP, C;
  {
    int Pt, XT;
    switch (tile_chunk(m, sizeof(T), Pt, XT, mode)) {
      case 10: return launch_tile<T, 10, 64>(in, out, m, Pt, s);
      case 11: return launch_tile<T, 11, 64>(in, out, m, Pt, s);
      case 12: return launch_tile<T, 12, 64>(in, out, m, Pt, s);
      case 13: return launch_tile<T, 13, 64>(in, out, m, Pt, s);
      case 14: return launch_tile<T, 14, 64>(in, out, m, Pt, s);
      case 15: return launch_tile<T, 15, 64>(in, out, m, Pt, s);
      case 16: return launch_tile<T, 16, 64>(in, out, m, Pt, s);
      case 32:
        if constexpr (sizeof(T) * 256 <= 1024)  // float only (tile_chunk checks the same)
          if (XT == 256) return launch_tile<T, 32, 256>(in, out, m, Pt, s);
        return launch_tile<T, 32, 128>(in, out, m, Pt, s);
      default: break;
    }
  }
  line_split(m.N, 32, P, logP, C);
  switch (C) {
    VLP_LINE_CASE(10) VLP_LINE_CASE(11) VLP_LINE_CASE(12) VLP_LINE_CASE(13) VLP_LINE_CASE(14) VLP_LINE_CASE(15)
    VLP_LINE_CASE(16) VLP_LINE_CASE(17) VLP_LINE_CASE(18) VLP_LINE_CASE(19) VLP_LINE_CASE(20) VLP_LINE_CASE(21)
    VLP_LINE_CASE(22) VLP_LINE_CASE(23) VLP_LINE_CASE(24) VLP_LINE_CASE(25) VLP_LINE_CASE(26) VLP_LINE_CASE(27)
    VLP_LINE_CASE(28) VLP_LINE_CASE(29) VLP_LINE_CASE(30) VLP_LINE_CASE(31) VLP_LINE_CASE(32)
    default: break;
  }
  const long nthr = m.X0 * m.X1 * m.A;
  const size_t lds = (size_t)64 * (m.N + 4) * sizeof(T);
  static bool attr = false;
  if (!attr) {
    (void)hipFuncSetAttribute((const void *)k_spline_line<T>, hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024);
    attr = true;
  }
  hipLaunchKernelGGL(k_spline_line<T>, dim3((unsigned)((nthr + 63) / 64)), dim3(64), lds, s, in, out, m,
                     spline_consts());
}
#undef VLP_LINE_CASE

int spline_tile_width(const LineMap &m, size_t tsz, int mode) {
  int P, XT;
  return tile_chunk(m, tsz, P, XT, mode) ? XT : 0;
}

int spline_variant_bits(const Geom &g) {
  int P, logP, C;
  x_split(g.n[0], P, logP, C);
  int bits = C ? 1 : 0;
  bool all = true;
  for (int d = 1; d < 4; d++) {
    line_split(g.n[d], 32, P, logP, C);
    all = all && C != 0;
  }
  return bits | (all ? 2 : 0);
}

template void launch_spline_x<double>(const Geom &, const double *, double *, hipStream_t);
template void launch_spline_x<float>(const Geom &, const float *, float *, hipStream_t);
template void launch_spline_line<double>(const double *, double *, const LineMap &, hipStream_t, int);
template void launch_spline_line<float>(const float *, float *, const LineMap &, hipStream_t, int);

}  // namespace vlp
