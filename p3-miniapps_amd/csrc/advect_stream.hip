// advect_stream.hip -- y-streaming 4-D advection (gfx950).
//
// Same arithmetic as advect_4D of the reference (computeFeet
// VMO/Advection.hpp:166-218, interp_4D :221-296, the loop of :409-498); the
// work is organised so that every C4 coefficient a workgroup needs crosses
// from L2 into the CU once, through LDS.
//
// Workgroup = (all x) x (TVX consecutive vx) at one vy.  It walks y: at step q
// it holds one y-plane of C4 in LDS -- positions x -2..N0+1, vx tile-2 ..
// tile+TVX+1, vy ivy-2..ivy+2 (a slab) -- with the next two in flight.
// Lane (x, vx) keeps the four cells (x, y, vx, vy) whose y-stencil contains
// the current plane "open": their B-spline weights along x, y, vx, vy and an
// accumulator.  For vy > 0 every regular foot has floor(y*) = y-1 (for vy <= 0,
// y), so cell y reads planes y+oy-1 .. y+oy+2 with oy fixed per workgroup, is
// opened when its first plane arrives and written after its fourth.  Each 4-wide
// x row read from LDS then feeds all four open cells: 16 rows x 4 values per
// step and per lane, 340 FMAs per cell, no re-reads from L2.
//
// A lane reads one set of rows for its four cells when they agree on their x,
// vx and vy stencil offsets (each in {-1, 0} relative to the cell), else one
// set per distinct offset pattern.  A foot outside this pattern (displacement
// beyond one cell, or y offset different from the workgroup's) is interpolated
// by the direct 4^4 gather from C4 in global memory at the step it is opened
// and takes no streamed contribution.
#include "advect_common.hpp"

#include <cstdlib>

namespace vlp {

namespace {

// One open cell: B-spline weights along x, y, vx, vy ([dim][k]), running
// sum, and the packed stencil offsets (ox+1) | (ovx+1) << 1 | (ovy+1) << 2,
// or -1 when the cell takes no streamed contribution (with the same key as
// a bit, so that the four cells' keys compare by two bit operations).  (Keeping the four y
// weights rather than re-evaluating the one a plane needs from the
// fractional y position measured 1 % faster with double arithmetic, at 246
// registers.)
template <typename A> struct OpenCell {
  A w[4][4];
  A acc;
  // 1 << key (key = (ox + 1) | (ovx + 1) << 1 | (ovy + 1) << 2, the stencil
  // offsets of a streamed foot), or 0 without a streamed contribution
  int kbit;
};

typedef __attribute__((address_space(3))) void lds_void;


// Per-workgroup constants and per-lane state of the streaming loop.  The slab
// is kept in the storage type T (moved by LDS-DMA, 16 bytes per lane) and
// converted to the arithmetic type A as it is read.  Its rows are contiguous
// in LDS, so the slab is moved as one run of 16-byte chunks: DMA instruction j
// fills LDS bytes [1024 j, 1024 j + 1024) of the buffer, lane l taking chunk
// 64 j + l wherever it sits in C4; instructions are dealt to the waves in
// turn, and each lane keeps the C4 byte offset (from the plane's first row)
// of its chunk in every instruction of its wave.  E rows sit in LDS as
// (Ex, Ey) pairs: [buffer][row k1 = 0..2][EPR], moved 64 points per instruction.
//
// computeFeet's 3x3 Lagrange interpolation of E at the half-step position
// (x - dt/2 vx, y - dt/2 vy) separates: the x part (base column ip0 and
// weights cx) depends on the lane's (x, vx) only, the y part (row ip1 and
// weights cy) on (y, vy), which is uniform in the workgroup at a step.  A lane
// therefore keeps G[k1] = sum_k0 cx[k0] E(ip0+k0, ip1+k1) for the three rows
// and, as ip1 advances by one per step, computes only the new row.
template <typename T, typename A, int TVX, int NB, int RXC> struct YStream {
  static constexpr int VPL = 16 / (int)sizeof(T);  // slab values per lane load
  // double arithmetic reads the foot constants from LDS (see open())
  static constexpr bool KLDS = sizeof(A) == sizeof(double);
  static constexpr int MS = 8;  // slab DMA instructions per wave per plane, at most
  const T *__restrict__ c4;
  const double *__restrict__ exy;  // (Ex, Ey) pairs, [N1][N0]
  T *__restrict__ fout;
  T *slab;         // NB buffers of slab_bytes
  double *eb;      // [NB][3][EPR][2]
  double *kconst;  // LDS: domain constants of the foot computation (KC_* layout)
  // LDS: per y-step values uniform in the workgroup, [q][QT]: the y of cell
  // q, computeFeet's E row ip1 (as a double) and its three y Lagrange weights
  double *qtab;
  static constexpr int QT = 5;
  T *fbase;  // f^{n+1} at (x, y = 0, vx, vy) of this lane
  static constexpr int KC_MIN = 0, KC_INV = 4, KC_LMIN = 8, KC_LMAX = 12, KC_VHI = 16, KC_N = 18;
  int slab_bytes;  // whole 1 KB blocks
  int EPR, cpr;    // E points per LDS row (whole 64-point chunks), chunks per row
  unsigned voff[MS];  // C4 byte offset of this lane's chunk in the wave's k-th slab DMA
  int nslot;       // slab DMA instructions of this wave (wave-uniform)
  Geom g;
  FeetConsts fc;
  double dt;
  int N0, N1, N2, N3, RX, nsteps, oyw;
  int ix, tv, ivx, ivy;
  bool active;
  bool wave_has_cells;  // some lane of this wave owns cells (wave-uniform)
  double x, vx, vy;
  long S1, S2, S3;
  int F0;
  int wave, nwave, lane;
  int force;  // bit 0: every lane reads rows per cell; bit 1: every foot by direct gather
  int cbase;  // C4 element of (x -2, y -3, vx tile-2, vy ivy-2)
  double cx[3];          // x Lagrange weights of the half-step foot
  int jx[3];             // wrapped E columns ip0 .. ip0+2
  double gx[3], gy[3];   // x-interpolated E of rows ip1 .. ip1+2
  int ip1p;              // ip1 of the previous step (wave-uniform)
  OpenCell<A> cell[4];
  int bad;

  // E row ip1 of the cell y (computeFeet: round of the half-step y) - 1;
  // uniform in the workgroup.
  __device__ __forceinline__ int erow0(int y) const {
    return __builtin_amdgcn_readfirstlane((int)qtab[QT * y + 1]);
  }

  // Start moving step q's slab and E rows into LDS buffer b by LDS-DMA.
  // Returns the number of DMA instructions this wave issued (wave-uniform).
  __device__ __forceinline__ int fetch(int q, int b) {
    const int p = q + oyw - 1;  // y-position of the plane
    const char *src = (const char *)c4 + ((long)cbase + (long)F0 * (p + 3)) * (long)sizeof(T);
    // the slot bookkeeping is recomputed each step from these (a few scalar
    // operations) rather than kept live across the loop, where it would be
    // spilled from the scalar registers
    int ns = nslot, lb = b * slab_bytes + 1024 * wave, ls = 1024 * nwave;
    asm volatile("" : "+s"(ns), "+s"(lb), "+s"(ls));
    char *sb = (char *)slab + lb;
    int n = 0;
#pragma unroll
    for (int k = 0; k < MS; k++)
      if (k < ns) {
        // a fresh 32-bit offset per load keeps the address in the form
        // (scalar base + 32-bit lane offset) the load encodes directly,
        // instead of a 64-bit vector add per load
        unsigned o = voff[k];
        asm volatile("" : "+v"(o));
        __builtin_amdgcn_global_load_lds((const void *)(src + o), (lds_void *)(sb + k * ls), 16, 0, 0);
        n++;
      }
    if (q < N1) {
      const int ip1 = erow0(q);
      char *el = (char *)(eb + b * 6 * EPR);
      for (int e = wave; e < 3 * cpr; e += nwave, n++) {
        const int k1 = (e >= cpr) + (e >= 2 * cpr), c = e - k1 * cpr;
        int row = ip1 + k1;
        // periodic wrap (row is wave-uniform; a scalar loop, no division
        // constants kept live across the y loop)
        while (row < 0) row += N1;
        while (row >= N1) row -= N1;
        // points beyond N0 (last chunk) repeat the row's last point into the padding
        const int pt = min(64 * c + lane, N0 - 1);
        __builtin_amdgcn_global_load_lds((const void *)(exy + 2 * ((long)row * N0 + pt)),
                                         (lds_void *)(el + 16 * (k1 * EPR + 64 * c)), 16, 0, 0);
      }
    }
    return n;
  }

  // Per step, uniform in the workgroup: the y part of computeFeet for cell
  // y = q, and the update of the lane's x-interpolated E rows (buffer b).
  __device__ __forceinline__ void field_rows(int q, int b, double cy[3], double &yq) {
    const double *t = qtab + QT * q;
    yq = t[0];
    const int ip1 = __builtin_amdgcn_readfirstlane((int)t[1]);
    cy[0] = t[2];
    cy[1] = t[3];
    cy[2] = t[4];
    const double2 *er = (const double2 *)(eb + b * 6 * EPR);
    if (ip1 == ip1p + 1) {
      gx[0] = gx[1];
      gy[0] = gy[1];
      gx[1] = gx[2];
      gy[1] = gy[2];
      const double2 e0 = er[2 * EPR + jx[0]], e1 = er[2 * EPR + jx[1]], e2 = er[2 * EPR + jx[2]];
      gx[2] = cx[0] * e0.x + cx[1] * e1.x + cx[2] * e2.x;
      gy[2] = cx[0] * e0.y + cx[1] * e1.y + cx[2] * e2.y;
    } else {
#pragma unroll
      for (int k1 = 0; k1 < 3; k1++) {
        const double2 e0 = er[k1 * EPR + jx[0]], e1 = er[k1 * EPR + jx[1]], e2 = er[k1 * EPR + jx[2]];
        gx[k1] = cx[0] * e0.x + cx[1] * e1.x + cx[2] * e2.x;
        gy[k1] = cx[0] * e0.y + cx[1] * e1.y + cx[2] * e2.y;
      }
    }
    ip1p = ip1;
  }

  // Open cell y = q in slot R: the rest of computeFeet (VMO/Advection.hpp:
  // 166-218), the bound check, B-spline weights and stencil offsets.  A cell
  // that takes no streamed contribution gets no key bit and zero x-weights,
  // so that a common pass adds nothing to it.
  template <int R> __device__ __forceinline__ void open(int q, double yq, const double cy[3]) {
    OpenCell<A> &c = cell[R];
    c.acc = 0;
    c.kbit = 0;
    if (q >= N1) {
      // no cell: zero x-weights make its share of a common pass zero
#pragma unroll
      for (int k = 0; k < 4; k++) c.w[0][k] = 0;
      return;
    }
    // with double arithmetic the domain constants come from LDS (ds_read,
    // wave-uniform address) at each open rather than living in scalar
    // registers for the whole loop, where they would be spilled; float
    // arithmetic (vector registers held to three waves per SIMD) keeps them
    // in scalar registers
    double minp[4], invd[4], lmin[4], lmax[4], vhi[2];
    if constexpr (KLDS) {
      const double *kc = kconst;
      asm volatile("" : "+s"(kc));
#pragma unroll
      for (int j = 0; j < 4; j++) {
        minp[j] = kc[KC_MIN + j];
        invd[j] = kc[KC_INV + j];
        lmin[j] = kc[KC_LMIN + j];
        lmax[j] = kc[KC_LMAX + j];
      }
      vhi[0] = kc[KC_VHI];
      vhi[1] = kc[KC_VHI + 1];
    } else {
#pragma unroll
      for (int j = 0; j < 4; j++) {
        minp[j] = g.minPhy[j];
        invd[j] = g.inv_dx[j];
        lmin[j] = fc.locmin[j];
        lmax[j] = fc.locmax[j];
      }
      vhi[0] = g.minPhy[2] + fc.rxwidth[2];
      vhi[1] = g.minPhy[3] + fc.rxwidth[3];
    }
    const double ftmp1 = cy[0] * gx[0] + cy[1] * gx[1] + cy[2] * gx[2];
    const double ftmp2 = cy[0] * gy[0] + cy[1] * gy[1] + cy[2] * gy[2];
    const double xt2 = vx - 0.5 * dt * ftmp1;
    const double xt3 = vy - 0.5 * dt * ftmp2;
    double xs[4];
    xs[0] = x - dt * xt2;
    xs[1] = yq - dt * xt3;
    // the velocity clamp max(min(v*, vmax), vmin) as v_min / v_max: they
    // differ from the comparisons only for a NaN v*, which makes x* NaN
    // too, so the bound check below still fails
    xs[2] = fmax(fmin(vx - dt * ftmp1, vhi[0]), minp[2]);
    xs[3] = fmax(fmin(vy - dt * ftmp2, vhi[1]), minp[3]);
    // all eight comparisons, no short cut (NaN fails them)
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 4; j++) ok = ok & (xs[j] >= lmin[j]) & (xs[j] <= lmax[j]);
    int ip[4];
#pragma unroll
    for (int j = 0; j < 4; j++) {
      const double fj = invd[j] * (xs[j] - minp[j]);
      const double fl = floor(fj);
      ip[j] = ok ? (int)fl : 0;
      // fj - floor(fj) is fj - ip exactly wherever the foot is valid
      bspline_eta<A>((A)(fj - fl), c.w[j]);
    }
    const int ox = ip[0] - ix, oy = ip[1] - q, ovx = ip[2] - ivx, ovy = ip[3] - ivy;
    const bool fast = ok & ((unsigned)(ox + 1) < 2u) & (oy == oyw) & ((unsigned)(ovx + 1) < 2u) &
                      ((unsigned)(ovy + 1) < 2u) & !(force & 2);
    c.kbit = fast ? 1 << ((ox + 1) | ((ovx + 1) << 1) | ((ovy + 1) << 2)) : 0;
    // out of domain: the reference stops with "Time step is too large"
    bad |= !ok;
    if (ok & !fast) {
      // a foot outside the streamed pattern: direct gather.  A copy keeps the
      // slot's weights out of the callee's reach (the per-lane state then
      // stays in registers).
      A e[4][4];
#pragma unroll
      for (int j = 0; j < 4; j++)
#pragma unroll
        for (int k = 0; k < 4; k++) e[j][k] = c.w[j][k];
      // the C4 strides are rebuilt here, on this rare path, rather than
      // kept in scalar registers across the loop
      int n0 = N0, n1 = N1, n2 = N2;
      asm volatile("" : "+s"(n0), "+s"(n1), "+s"(n2));
      const C4Pitch cp = c4_pitch(n0, n1, n2);
      c.acc = gather4d<T, A>(c4, cp.s1, cp.s2, cp.s3, ip, e);
#pragma unroll
      for (int k = 0; k < 4; k++) c.w[0][k] = 0;  // no share of a common pass
    }
  }

  // LDS element of the first value of row (kvx = 0, kvy = 0) for a key.
  __device__ __forceinline__ int row_base(int key) const {
    const int ox = (key & 1) - 1, ovx = ((key >> 1) & 1) - 1, ovy = ((key >> 2) & 1) - 1;
    return (tv + 1 + ovx) * RX + (1 + ovy) * (TVX + 4) * RX + ix + 1 + ox;
  }

  // The plane's contribution t[c] to each of the four open cells from the 16
  // rows at LDS element sp (one stencil key): rows (kvx, kvy) in order, the
  // next row's LDS reads issued before the current row's FMAs (one row of
  // look-ahead hides the LDS latency).  With the slab row pitch known at
  // compile time (RXC = N0 + 4), rows kvx = 0, 1 and 2, 3 of a vy row are
  // read from one base address each, the second row through the immediate
  // offsets of the paired LDS reads (an empty asm keeps the compiler from
  // rebuilding a base per row): 8 address additions per pass instead of 16.
  __device__ __forceinline__ void pass_rows(const T *sp, A t[4]) const {
    const int VS = (TVX + 4) * (RXC ? RXC : RX);  // vy row stride
    typedef const __attribute__((address_space(3))) T lds_t;
    lds_t *pb = (lds_t *)sp;
    if constexpr (RXC != 0) asm volatile("" : "+v"(pb));
    A u[4];
#pragma unroll
    for (int c = 0; c < 4; c++) t[c] = u[c] = 0;
    T r[2][4];
#pragma unroll
    for (int k = 0; k < 4; k++) r[0][k] = RXC ? pb[k] : sp[k];
#pragma unroll
    for (int i = 0; i < 16; i++) {
      const int kvx = i & 3, kvy = i >> 2;
      // row i complete (only LDS reads are waited on; the slab DMA stays in
      // flight), then row i+1's reads go out ahead of row i's FMAs
      __builtin_amdgcn_s_waitcnt(0xC07F);  // lgkmcnt(0)
      if (i + 1 < 16) {
        const int n = i + 1;
        if constexpr (RXC != 0) {
          if ((n & 1) == 0) {  // first row of a pair: a new base
            pb = (lds_t *)(sp + (n >> 2) * VS + (n & 2) * RXC);
            asm volatile("" : "+v"(pb));
          }
          const int o = (n & 1) * RXC;
#pragma unroll
          for (int k = 0; k < 4; k++) r[n & 1][k] = pb[o + k];
        } else {
          const T *rn = sp + (n >> 2) * VS + (n & 3) * RX;
#pragma unroll
          for (int k = 0; k < 4; k++) r[n & 1][k] = rn[k];
        }
      }
      __builtin_amdgcn_sched_barrier(0);
      const A r0 = (A)r[i & 1][0], r1 = (A)r[i & 1][1], r2 = (A)r[i & 1][2], r3 = (A)r[i & 1][3];
#pragma unroll
      for (int c = 0; c < 4; c++) {
        const A *w0 = cell[c].w[0];
        const A d = w0[0] * r0 + w0[1] * r1 + w0[2] * r2 + w0[3] * r3;
        u[c] += cell[c].w[2][kvx] * d;
      }
      if (kvx == 3) {
#pragma unroll
        for (int c = 0; c < 4; c++) {
          t[c] += cell[c].w[3][kvy] * u[c];
          u[c] = 0;
        }
      }
      __builtin_amdgcn_sched_barrier(0);
    }
  }

  // Add plane q (LDS buffer b) to the four open cells; the cell opened at
  // step q - j sits in slot (R - j) & 3 and takes y-weight j.  When every
  // lane's streamed cells agree on their stencil key (the usual case), one
  // pass over the 16 rows of that key serves all four cells; cells without a
  // streamed contribution have zero x-weights, so their share is zero.
  // Otherwise one pass reads the rows of one key per lane (the first key
  // among its cells still to do) and adds to the cells with that key; the
  // wave repeats until no lane has cells left.  force & 1 takes the second
  // form with one cell per pass.
  template <int R> __device__ __forceinline__ void accumulate(int b) {
    const T *s = (const T *)((const char *)slab + b * slab_bytes);
    // the distinct keys of the open cells as bits: one key (or none) when
    // at most one bit is set
    const int m = cell[0].kbit | cell[1].kbit | cell[2].kbit | cell[3].kbit;
    const bool same = ((m & (m - 1)) == 0) & !(force & 1);
    A t[4];
    if (__builtin_amdgcn_ballot_w64(!same) == 0) {
      const int k0 = m ? __builtin_ctz(m) : 0;
      pass_rows(s + row_base(k0), t);
      close_plane<R, 0>(t, 15u);
      close_plane<R, 1>(t, 15u);
      close_plane<R, 2>(t, 15u);
      close_plane<R, 3>(t, 15u);
      return;
    }
    unsigned todo = 0;
#pragma unroll
    for (int c = 0; c < 4; c++) todo |= (cell[c].kbit ? 1u : 0u) << c;
    while (__builtin_amdgcn_ballot_w64(todo != 0)) {
      int kp = 0;
#pragma unroll
      for (int c = 3; c >= 0; c--)
        if ((todo >> c) & 1) kp = cell[c].kbit;
      unsigned sel = 0;
#pragma unroll
      for (int c = 0; c < 4; c++)
        if (((todo >> c) & 1) && cell[c].kbit == kp && !((force & 1) && sel)) sel |= 1u << c;
      todo &= ~sel;
      pass_rows(s + row_base(kp ? __builtin_ctz(kp) : 0), t);
      close_plane<R, 0>(t, sel);
      close_plane<R, 1>(t, sel);
      close_plane<R, 2>(t, sel);
      close_plane<R, 3>(t, sel);
    }
  }

  // The cell opened J steps ago takes the plane's contribution with its y
  // weight J, when sel has its bit.
  template <int R, int J> __device__ __forceinline__ void close_plane(const A t[4], unsigned sel) {
    constexpr int c = (R - J) & 3;
    OpenCell<A> &oc = cell[c];
    const A add = oc.w[1][J] * t[c];
    if (sel == 15u) oc.acc += add;
    else oc.acc += ((sel >> c) & 1) ? add : (A)0;
  }

  // Write cell y = q - 3 (slot (R + 1) & 3).
  template <int R> __device__ __forceinline__ void close(int q) {
    const int y = q - 3;
    if (y < 0 || y >= N1) return;
    const A v = cell[(R + 1) & 3].acc;
    fbase[(long)N0 * y] = (T)v;
  }

  // Step q: plane q sits in LDS buffer b = q mod NB; vector memory operations
  // of a wave retire in issue order, so a vmcnt wait on all but the newest
  // operations completes every older load.
  //  NB = 3: the loads of plane q+2 are issued last; the wait before the
  //   barrier leaves exactly those in flight, so plane q+1 is complete in LDS
  //   for every wave after the barrier.
  //  NB = 2: the loads of plane q+1 are issued first and overlap the step;
  //   the wait leaves only this step's store of f in flight.
  template <int R> __device__ __forceinline__ void step(int q, int b) {
    if (NB == 2 && q + 1 < nsteps) fetch(q + 1, b ^ 1);
    if (active) {
      double cy[3] = {0., 0., 0.}, yq = 0.;
      if (q < N1) field_rows(q, b, cy, yq);  // no cell opens beyond the last y
      open<R>(q, yq, cy);
      accumulate<R>(b);
      close<R>(q);
    }
    int n = 0;
    if (NB == 3) {
      if (q + 2 < nsteps) n = fetch(q + 2, b == 0 ? 2 : b - 1);
    } else {
      n = (wave_has_cells && q >= 3 && q - 3 < N1) ? 1 : 0;
    }
    wait_vm(n);
    __builtin_amdgcn_s_barrier();
  }
};

// Workgroup -> (vx tile, vy).  Workgroups go to the 8 XCDs (each with its
// own L2) round-robin: block b runs on XCD b mod 8, as the (b / 8)-th of
// that XCD.  The workgroups that stage the same C4 rows are those of the 5
// neighbouring vy and the neighbouring vx tiles; with remap, XCD x takes the
// vx tiles [x ntile/8, (x+1) ntile/8) for every vy, vy slower and tile
// faster, so that the workgroups resident on one XCD at a time cover a
// compact block of (tile, vy) and share their rows through its L2, while
// every XCD gets its share of each vy (the vy = 0 row, whose feet mostly take
// the direct gather, is the slowest).  Without remap (or when the tiles do
// not split over the XCDs) block b is (b mod ntile, b / ntile).
__device__ __forceinline__ void ys_block_tile(int b, int grid, int ntile, int N3, int remap, int &tile, int &ivy) {
  constexpr int NX = 8;
  if (remap && ntile % NX == 0 && grid == ntile * N3) {
    const int x = b % NX, k = b / NX;
    const int tx = ntile / NX;  // tiles per XCD
    ivy = k / tx;
    tile = x * tx + k % tx;
  } else {
    tile = b % ntile;
    ivy = b / ntile;
  }
}

}  // namespace

template <typename T, typename A, int TVX, int NB, int RXC>
__global__ __launch_bounds__(512) __attribute__((amdgpu_waves_per_eu(sizeof(A) == 4 ? 3 : 1)))
void k_advect4d_ys(const T *__restrict__ c4, const double *__restrict__ exy, Geom g,
                                                     FeetConsts fc, double dt, T *__restrict__ fout,
                                                     int *__restrict__ err, int force, int remap) {
  extern __shared__ __align__(16) unsigned char ys_smem[];
  YStream<T, A, TVX, NB, RXC> s;
  s.force = force;
  s.c4 = c4;
  s.exy = exy;
  s.fout = fout;
  s.g = g;
  s.fc = fc;
  s.dt = dt;
  s.N0 = g.n[0];
  s.N1 = g.n[1];
  s.N2 = g.n[2];
  s.N3 = g.n[3];
  s.RX = s.N0 + 4;
  const int per_row = s.RX / s.VPL;                 // 16-byte chunks per slab row
  const int nchunk = 5 * (TVX + 4) * per_row;
  const int ninstr = (nchunk + 63) >> 6;
  s.slab_bytes = ninstr * 1024;
  s.cpr = (s.N0 + 63) >> 6;
  s.EPR = 64 * s.cpr;
  s.slab = (T *)ys_smem;
  s.eb = (double *)(ys_smem + NB * (size_t)s.slab_bytes);
  s.kconst = s.eb + NB * 6 * s.EPR;
  s.qtab = s.kconst + s.KC_N;
  if (threadIdx.x == 0) {
    double *k = s.kconst;
#pragma unroll
    for (int j = 0; j < 4; j++) {
      k[s.KC_MIN + j] = g.minPhy[j];
      k[s.KC_INV + j] = g.inv_dx[j];
      k[s.KC_LMIN + j] = fc.locmin[j];
      k[s.KC_LMAX + j] = fc.locmax[j];
    }
    k[s.KC_VHI] = g.minPhy[2] + fc.rxwidth[2];
    k[s.KC_VHI + 1] = g.minPhy[3] + fc.rxwidth[3];
  }
  s.nsteps = s.N1 + 3;
  const int ntile = s.N2 / TVX;
  int tile, ivy;
  ys_block_tile(blockIdx.x, gridDim.x, ntile, s.N3, remap, tile, ivy);
  s.ivy = ivy;
  {  // the y part of computeFeet for every cell y of this vy (VMO/Advection.hpp:174-196)
    const double vyw = g.minPhy[3] + ivy * g.dx[3];
    for (int q = threadIdx.x; q < s.N1; q += blockDim.x) {
      const double yq = g.minPhy[1] + q * g.dx[1];
      const double xt1 = half_y(g, yq, vyw, dt);
      const int ip1 = (int)(round(xt1) - 1);
      const double py = xt1 - ip1;
      double *t = s.qtab + s.QT * q;
      t[0] = yq;
      t[1] = (double)ip1;
      t[2] = .5 * (py - 1.) * (py - 2.);
      t[3] = -1. * (py) * (py - 2.);
      t[4] = .5 * (py) * (py - 1.);
    }
  }
  __syncthreads();  // both tables are written before the first fetch reads them
  const int tid = threadIdx.x;
  s.active = tid < s.N0 * TVX;
  s.ix = s.active ? tid % s.N0 : 0;
  s.tv = s.active ? tid / s.N0 : 0;
  s.ivx = tile * TVX + s.tv;
  s.x = g.minPhy[0] + s.ix * g.dx[0];
  s.vx = g.minPhy[2] + s.ivx * g.dx[2];
  s.vy = g.minPhy[3] + s.ivy * g.dx[3];
  s.oyw = s.vy > 0. ? -1 : 0;
  s.fbase = fout + s.ix + (long)s.N0 * s.N1 * (s.ivx + (long)s.N2 * s.ivy);
  {
    const C4Pitch cp = c4_pitch(s.N0, s.N1, s.N2);
    s.F0 = (int)cp.s1;
    s.S1 = cp.s1;
    s.S2 = cp.s2;
    s.S3 = cp.s3;
  }
  s.wave = __builtin_amdgcn_readfirstlane(tid >> 6);
  s.lane = tid & 63;
  s.nwave = (int)(blockDim.x >> 6);
  s.nslot = s.wave < ninstr ? (ninstr - 1 - s.wave) / s.nwave + 1 : 0;
  {
#pragma unroll
    for (int k = 0; k < s.MS; k++) {
      // chunks past the slab repeat its last chunk into the buffer's padding
      const int c = min(64 * (s.wave + k * s.nwave) + s.lane, nchunk - 1);
      const int row = c / per_row, col = c - row * per_row;
      const int kvx = row % (TVX + 4), kvy = row / (TVX + 4);
      s.voff[k] = (unsigned)((s.S2 * kvx + s.S3 * kvy + (long)s.VPL * col) * (long)sizeof(T));
    }
  }
  {  // x part of computeFeet's E interpolation (VMO/Advection.hpp:174-196)
    double xt0 = s.x - 0.5 * dt * s.vx;
    xt0 = g.inv_dx[0] * (xt0 - g.minPhy[0]);
    const int ip0 = (int)(round(xt0) - 1);
    const double px = xt0 - ip0;
    s.cx[0] = .5 * (px - 1.) * (px - 2.);
    s.cx[1] = -1. * (px) * (px - 2.);
    s.cx[2] = .5 * (px) * (px - 1.);
#pragma unroll
    for (int k0 = 0; k0 < 3; k0++) s.jx[k0] = ((ip0 + k0) % s.N0 + s.N0) % s.N0;
  }
  s.ip1p = -(1 << 30);
  // position (x -2, y -3, vx tile-2, vy ivy-2) sits at storage (1, 0, tile+1, ivy+1)
  s.cbase = (int)(1 + s.S2 * (tile * TVX + 1) + s.S3 * (s.ivy + 1));
#pragma unroll
  for (int c = 0; c < 4; c++) {
    s.cell[c].kbit = 0;
    s.cell[c].acc = 0;
#pragma unroll
    for (int j = 0; j < 4; j++)
#pragma unroll
      for (int k = 0; k < 4; k++) s.cell[c].w[j][k] = 0;
  }
  s.bad = 0;
  s.wave_has_cells = s.wave * 64 < s.N0 * TVX;
  if (NB == 3) {  // planes 0 and 1 in flight; wait for plane 0
    s.fetch(0, 0);
    wait_vm(s.fetch(1, 1));
  } else {
    s.fetch(0, 0);
    wait_vm(0);
  }
  __builtin_amdgcn_s_barrier();
  int b = 0;  // q mod NB
  for (int q0 = 0; q0 < s.nsteps; q0 += 4) {
    s.template step<0>(q0, b);
    b = b == NB - 1 ? 0 : b + 1;
    if (q0 + 1 < s.nsteps) s.template step<1>(q0 + 1, b);
    b = b == NB - 1 ? 0 : b + 1;
    if (q0 + 2 < s.nsteps) s.template step<2>(q0 + 2, b);
    b = b == NB - 1 ? 0 : b + 1;
    if (q0 + 3 < s.nsteps) s.template step<3>(q0 + 3, b);
    b = b == NB - 1 ? 0 : b + 1;
  }
  if (s.bad) atomicOr(err, 2);
}

namespace {

struct YsPlan {
  int tvx = 0;  // 0: not eligible
  int nb = 3;   // LDS plane buffers
  int block = 0;
  size_t lds = 0;
};

// Slab DMA instructions per plane and E points per LDS row of the streaming
// kernel (the same arithmetic as its set-up).
long ys_slab_instr(long N0, long tvx, size_t tsz) { return (5 * (tvx + 4) * ((N0 + 4) * (long)tsz / 16) + 63) / 64; }
long ys_epr(long N0) { return (N0 + 63) / 64 * 64; }

// want_tvx / want_nb: a chosen vx tile and buffer count (0: free); the
// environment (VLP4D_YS_TVX, VLP4D_YS_NBUF) overrides them.
YsPlan ys_plan(const Geom &g, size_t tsz, int want_tvx = 0, int want_nb = 0) {
  YsPlan p;
  const long N0 = g.n[0], N1 = g.n[1], N2 = g.n[2], N3 = g.n[3];
  const long c4n = c4_elems(N0, N1, N2, N3);
  if (N0 % 2 || c4n >= (1L << 31)) return p;
  int force = want_tvx, force_nb = want_nb;
  if (const char *e = getenv("VLP4D_YS_TVX")) force = atoi(e);
  if (const char *e = getenv("VLP4D_YS_NBUF")) force_nb = atoi(e);
  // Preference order measured at 128^4: two or more small workgroups per CU
  // (vx tile 2, two buffers), whose barriers are independent, run best for
  // both slab types (double: 8.0 ms against 8.7 at tile 4; float storage 7.5
  // against 8.1; float arithmetic, three waves per SIMD, 5.2 against 6.2).
  static const int pref_f[][2] = {{2, 2}, {4, 2}, {8, 2}, {4, 3}, {2, 3}, {8, 3}};
  static const int pref_d[][2] = {{2, 2}, {4, 2}, {8, 2}, {4, 3}, {8, 3}, {2, 3}};
  for (const auto &c : (tsz == sizeof(float) ? pref_f : pref_d)) {
    const int tvx = c[0], nb = c[1];
    if (force && tvx != force) continue;
    if (force_nb && nb != force_nb) continue;
    if (N2 % tvx) continue;
    const long lanes = N0 * tvx;
    if (lanes > 512) continue;
    const int block = (int)((lanes + 63) / 64 * 64);
    if ((N0 + 4) % (16 / (long)tsz) || N0 % 16) continue;  // whole 16-byte slab chunks per row
    const long ninstr = ys_slab_instr(N0, tvx, tsz);
    if ((ninstr + block / 64 - 1) / (block / 64) > 8) continue;  // DMA slots per wave
    const size_t lds = nb * (size_t)ninstr * 1024 + nb * 3 * (size_t)ys_epr(N0) * 2 * sizeof(double) +
                       (18 + 5 * (size_t)N1) * sizeof(double);  // domain constants, per-step table
    if (lds > 160 * 1024) continue;
    p.tvx = tvx;
    p.nb = nb;
    p.block = block;
    p.lds = lds;
    return p;
  }
  return p;
}

template <typename T, typename A, int TVX, int NB, int RXC>
void launch_ys_t(const Geom &g, const YsPlan &p, const T *c4, const double *exy, double dt, T *fout, int *err,
                 hipStream_t s) {
  const void *fn = (const void *)k_advect4d_ys<T, A, TVX, NB, RXC>;
  if (p.lds > 64 * 1024) (void)hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)p.lds);
  const unsigned grid = (unsigned)((g.n[2] / TVX) * g.n[3]);
  int force = 0;  // test switch: VLP4D_YS_FORCE=1 rows per cell, 2 direct gather, 3 both
  if (const char *e = getenv("VLP4D_YS_FORCE")) force = atoi(e) & 3;
  int remap = 1;  // VLP4D_YS_REMAP=0: block b -> (b mod ntile, b / ntile)
  if (const char *e = getenv("VLP4D_YS_REMAP")) remap = atoi(e) != 0;
  hipLaunchKernelGGL((k_advect4d_ys<T, A, TVX, NB, RXC>), dim3(grid), dim3(p.block), p.lds, s, c4, exy, g,
                     feet_consts(g), dt, fout, err, force, remap);
}

template <typename T, typename A, int TVX>
void launch_ys_nb(const Geom &g, const YsPlan &p, const T *c4, const double *exy, double dt, T *fout, int *err,
                  hipStream_t s) {
  // N0 = 128 with the vx tile of 2 (the tile chosen at 128^4) takes the
  // kernel with the slab row pitch 132 built in
  if constexpr (TVX == 2) {
    if (g.n[0] == 128) {
      if (p.nb == 2) launch_ys_t<T, A, TVX, 2, 132>(g, p, c4, exy, dt, fout, err, s);
      else launch_ys_t<T, A, TVX, 3, 132>(g, p, c4, exy, dt, fout, err, s);
      return;
    }
  }
  if (p.nb == 2) launch_ys_t<T, A, TVX, 2, 0>(g, p, c4, exy, dt, fout, err, s);
  else launch_ys_t<T, A, TVX, 3, 0>(g, p, c4, exy, dt, fout, err, s);
}

}  // namespace

// Returns false (nothing launched) when the geometry does not fit the
// streaming kernel.
template <typename T, typename A>
bool launch_advect4d_ystream(const Geom &g, const T *c4, const double *exy, double dt, T *fout, int *err,
                             hipStream_t s, int ys_tvx, int ys_nb) {
  if (const char *e = getenv("VLP4D_YSTREAM"))
    if (atoi(e) == 0) return false;
  YsPlan p = ys_plan(g, sizeof(T), ys_tvx, ys_nb);
  if (!p.tvx && (ys_tvx || ys_nb)) p = ys_plan(g, sizeof(T));  // a choice this geometry cannot take
  switch (p.tvx) {
    case 8: launch_ys_nb<T, A, 8>(g, p, c4, exy, dt, fout, err, s); return true;
    case 4: launch_ys_nb<T, A, 4>(g, p, c4, exy, dt, fout, err, s); return true;
    case 2: launch_ys_nb<T, A, 2>(g, p, c4, exy, dt, fout, err, s); return true;
    default: return false;
  }
}

bool ystream_eligible(const Geom &g, size_t tsz, int tvx, int nb) {
  const YsPlan p = ys_plan(g, tsz, tvx, nb);
  return p.tvx == tvx && p.nb == nb;
}

bool ystream_preferred(const Geom &g, size_t tsz, int tvx, int nb) {
  const YsPlan p = ys_plan(g, tsz);
  return p.tvx == tvx && p.nb == nb;
}

int ystream_tvx(const Geom &g, bool fp32_storage, bool arith32) {
  if (const char *e = getenv("VLP4D_YSTREAM"))
    if (atoi(e) == 0) return 0;
  (void)arith32;
  return ys_plan(g, fp32_storage ? sizeof(float) : sizeof(double)).tvx;
}

template bool launch_advect4d_ystream<double, double>(const Geom &, const double *, const double *, double, double *,
                                                      int *, hipStream_t, int, int);
template bool launch_advect4d_ystream<float, double>(const Geom &, const float *, const double *, double, float *,
                                                     int *, hipStream_t, int, int);
template bool launch_advect4d_ystream<float, float>(const Geom &, const float *, const double *, double, float *, int *,
                                                    hipStream_t, int, int);

}  // namespace vlp
