// spline_mfma.hip -- the strided spline passes (vx, vy) as matrix products on
// the MFMA units (gfx950).
//
// Every line of N values maps to its N+4 coefficients (positions -2..N+1) by
// one fixed linear operator, M = R E: E the periodic copy plus the MMAX-term
// alpha-sum extension of the halo exchange (VMO/Communication.hpp:756-898),
// R the causal/anticausal prefilter (VMO/Spline.hpp:114-199).  The line
// solves of a pass are therefore one matrix product per outer index,
//   OUT[m][c] = sum_i M[m][i] IN[i][c],   m < N+4, i < N, c the contiguous index,
// which this kernel runs with v_mfma_f64_16x16x4f64 (double storage) or
// v_mfma_f32_16x16x4f32 (float storage).  M is the exact reference operator
// (built once per context by the line kernel on unit vectors), not the
// circulant inverse of (1, 4, 1)/6, which differs from it at the |alpha|^9
// level near the line ends.
//
// Workgroup: Mr/16 waves (Mr = N+4 rounded up to 16), wave w owning output
// rows [16 w, 16 w + 16); A fragments (its 16 rows of M) stay in registers
// across the workgroup's column tiles.  A tile of 64 contiguous columns of
// the N input rows is staged in LDS; each wave then issues Kp/4 x 4 MFMAs
// (four 16-column blocks) and writes its 16 x 64 block.
// Lane layout of the 16x16x4 MFMAs: lane l supplies A[l % 16][l / 16] and
// B[l / 16][l % 16]; it receives, in register r, D[4 (l / 16) + r][l % 16]
// (f32) or D[l / 16 + 4 r][l % 16] (f64), as measured on gfx950.
#include "device_common.hpp"

namespace vlp {

namespace {

typedef double mf_v4d __attribute__((ext_vector_type(4)));
typedef float mf_v4f __attribute__((ext_vector_type(4)));

template <typename T> struct MfmaOps;
template <> struct MfmaOps<double> {
  typedef mf_v4d acc_t;
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
};
template <> struct MfmaOps<float> {
  typedef mf_v4f acc_t;
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
  static __device__ __forceinline__ acc_t mma(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
};

constexpr int MF_COLS = 64;  // columns per LDS tile

// Address of contiguous index c of outer index r (the LineMap's two forms:
// "comb" when i0 and i1 together are contiguous, as for the vy pass).
struct MfAddr {
  long XC;          // contiguous columns per outer index
  long X0;          // length of i0 (output split of c for comb)
  bool comb;
  __device__ __forceinline__ void bases(const LineMap &m, long r, long &in0, long &out0) const {
    if (comb) {
      in0 = r * m.ios;
      out0 = r * m.oos;
    } else {
      const long i1 = r % m.X1, o = r / m.X1;
      in0 = i1 * m.is1 + o * m.ios;
      out0 = i1 * m.os1 + o * m.oos;
    }
  }
  __device__ __forceinline__ long out_col(const LineMap &m, long c) const {
    return comb ? (c % X0) + (c / X0) * m.os1 : c;
  }
};

}  // namespace

// KS = Kp / 4 (the k-steps; A fragments in registers), or 0: A read from
// global memory at each k-step (any N).
template <typename T, int KS>
__global__ __launch_bounds__(1024) void k_spline_mfma(const T *__restrict__ in, T *__restrict__ out, LineMap m,
                                                      const T *__restrict__ Mm, int Kp, MfAddr ad, int tiles_per_wg,
                                                      long ntile) {
  extern __shared__ __align__(16) unsigned char mf_smem[];
  T *B = (T *)mf_smem;  // [N rows][MF_COLS]
  using Ops = MfmaOps<T>;
  using acc_t = typename Ops::acc_t;
  const int N = m.N, Mrows = N + 4;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int nthr = blockDim.x;
  const int ks = KS ? KS : Kp / 4;
  const long r = blockIdx.y;  // outer index
  long in0, out0;
  ad.bases(m, r, in0, out0);
  const int arow = 16 * w + (lane & 15), akc = lane >> 4;
  T areg[KS ? KS : 1];
  if constexpr (KS > 0) {
#pragma unroll
    for (int s = 0; s < KS; s++) areg[s] = Mm[(long)arow * Kp + 4 * s + akc];
  }
  constexpr int VPC = 16 / (int)sizeof(T);  // values per 16-byte chunk
  const int chunks_per_row = MF_COLS / VPC;
  const long t0 = (long)blockIdx.x * tiles_per_wg;
  for (long t = t0; t < t0 + tiles_per_wg && t < ntile; t++) {
    const long c0 = t * MF_COLS;
    __syncthreads();  // the previous tile's B reads are done
    for (int q = tid; q < N * chunks_per_row; q += nthr) {
      const int i = q / chunks_per_row, cc = (q - i * chunks_per_row) * VPC;
      const long c = c0 + cc;
      typedef T vec_t __attribute__((ext_vector_type(VPC)));
      vec_t v = {};
      if (c < ad.XC) v = *(const vec_t *)(in + in0 + (long)i * m.ils + c);
      *(vec_t *)(B + i * MF_COLS + cc) = v;
    }
    __syncthreads();
    acc_t acc[4];
#pragma unroll
    for (int nb = 0; nb < 4; nb++) acc[nb] = acc_t{};
    for (int s = 0; s < ks; s++) {
      T a;
      if constexpr (KS > 0) a = areg[s];
      else a = Mm[(long)arow * Kp + 4 * s + akc];
      const int brow = 4 * s + akc;
#pragma unroll
      for (int nb = 0; nb < 4; nb++) {
        const T b = brow < N ? B[brow * MF_COLS + 16 * nb + (lane & 15)] : (T)0;
        acc[nb] = Ops::mma(a, b, acc[nb]);
      }
    }
#pragma unroll
    for (int nb = 0; nb < 4; nb++) {
      const long c = c0 + 16 * nb + (lane & 15);
      if (c >= ad.XC) continue;
      const long oc = out0 + ad.out_col(m, c);
#pragma unroll
      for (int rr = 0; rr < 4; rr++) {
        const int mrow = 16 * w + Ops::row(lane, rr);
        if (mrow < Mrows) out[oc + (long)mrow * m.ols] = acc[nb][rr];
      }
    }
  }
}

// Whether the MFMA pass takes this line family: unit-stride contiguous index
// (whole 16-byte chunks), N + 4 <= 256 rows (at most 16 waves).
bool spline_mfma_eligible(const LineMap &m, size_t tsz) {
  if (m.is0 != 1 || m.os0 != 1 || m.N + 4 > 256) return false;
  const bool comb = m.is1 == m.X0;
  const long XC = comb ? m.X0 * m.X1 : m.X0;
  return ((XC * (long)tsz) % 16 == 0) && ((m.ils * (long)tsz) % 16 == 0) && ((m.is1 * (long)tsz) % 16 == 0) &&
         ((m.ios * (long)tsz) % 16 == 0);
}

template <typename T> void launch_spline_mfma(const T *in, T *out, const LineMap &m, const T *Mm, int Kp, hipStream_t s) {
  MfAddr ad;
  ad.comb = m.is1 == m.X0;
  ad.XC = ad.comb ? m.X0 * m.X1 : m.X0;
  ad.X0 = m.X0;
  const long R = ad.comb ? m.A : m.X1 * m.A;
  const int Mr = (m.N + 4 + 15) / 16 * 16;
  const long ntile = (ad.XC + MF_COLS - 1) / MF_COLS;
  const int tpw = 4;
  const size_t lds = (size_t)m.N * MF_COLS * sizeof(T);
  dim3 grid((unsigned)((ntile + tpw - 1) / tpw), (unsigned)R);
  if (Kp == 128) {
    const void *fn = (const void *)k_spline_mfma<T, 32>;
    if (lds > 64 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_spline_mfma<T, 32>), grid, dim3(4 * Mr), lds, s, in, out, m, Mm, Kp, ad, tpw, ntile);
  } else {
    const void *fn = (const void *)k_spline_mfma<T, 0>;
    if (lds > 64 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL((k_spline_mfma<T, 0>), grid, dim3(4 * Mr), lds, s, in, out, m, Mm, Kp, ad, tpw, ntile);
  }
}

// The operator M of lines of length N, row-major (N+4) x Kp padded to Mr
// rows (Kp = N rounded up to 4; padding zero): the double line kernel on the
// N unit vectors, then a conversion for float.  Mm must hold Mr * Kp values.
__global__ void k_mfma_unit(double *id, int N) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < N * N) id[t] = (t % N == t / N) ? 1.0 : 0.0;
}
template <typename T> __global__ void k_mfma_cvt(const double *src, T *dst, long n) {
  const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) dst[t] = (T)src[t];
}

int spline_mfma_kp(int N) { return (N + 3) / 4 * 4; }
size_t spline_mfma_matrix_elems(int N) { return (size_t)((N + 4 + 15) / 16 * 16) * spline_mfma_kp(N); }

template <typename T> void build_spline_mfma_matrix(int N, T *Mm, double *work, hipStream_t s) {
  // work: N*N (identity) + matrix_elems doubles
  const int Kp = spline_mfma_kp(N);
  const size_t ne = spline_mfma_matrix_elems(N);
  double *id = work, *Md = work + (size_t)N * N;
  hipLaunchKernelGGL(k_mfma_unit, dim3((unsigned)((N * N + 255) / 256)), dim3(256), 0, s, id, N);
  (void)hipMemsetAsync(Md, 0, ne * sizeof(double), s);
  // line j = e_j (stride 1 inside the line, N apart); coefficient m of line j
  // lands at M[m][j] = Md[m * Kp + j]
  LineMap q{};
  q.N = N;
  q.X0 = N; q.X1 = 1; q.A = 1;
  q.is0 = N; q.is1 = 0; q.ios = 0; q.ils = 1;
  q.os0 = 1; q.os1 = 0; q.oos = 0; q.ols = Kp;
  launch_spline_line<double>(id, Md, q, s);
  hipLaunchKernelGGL(k_mfma_cvt<T>, dim3((unsigned)((ne + 255) / 256)), dim3(256), 0, s, Md, Mm, (long)ne);
}

template void launch_spline_mfma<double>(const double *, double *, const LineMap &, const double *, int, hipStream_t);
template void launch_spline_mfma<float>(const float *, float *, const LineMap &, const float *, int, hipStream_t);
template void build_spline_mfma_matrix<double>(int, double *, double *, hipStream_t);
template void build_spline_mfma_matrix<float>(int, float *, double *, hipStream_t);

}  // namespace vlp
