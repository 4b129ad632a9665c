// sdv_ke.hip -- sinc-DVR kinetic energy apply (mult_ke): FP64 or FP32 matrix
// cores (k_ke_xy, k_ke_z), and an FP64 VALU form (k_ke_axis).
//
// Reference: mult_ke -> mult_general(option=1) -> mult_general_withbcast ->
// mult_allpar -> mult_allone -> mult_all0 (SD/coreproject.f90:2104-2116,
// 1321-1390, 1634-1666, 2050-2097):
//   out = sum_d  invJ_d * ( K_d contracted along axis d of (invJ_d * in) )
//       + scalediag * in                                   (ECS, scalingflag=1)
//   out = sum_d  K_d contracted along axis d of in         (scalingflag=0)
// with K_d(p,q) = ketot(q-p) the real Toeplitz sinc stencil
// (SD/init_sinc_new.f90:441-464).  ECS enters through per-axis complex
// Jacobians; scalediag(x,y,z) = s_1(x) + s_2(y) + s_3(z) (SD/init_sinc_new.f90:517-521).
//
// Matrix cores (n <= 96): two launches, the x and y contractions of each
// (orbital, z) plane from one staged copy of the plane (k_ke_xy), then the z
// contraction added from registers (k_ke_z): `in` read twice and `out`
// written, read and written again, 5 n^3 values per orbital against 8 for one
// launch per axis.  VALU form (n > 96 or precision 2): one launch per axis, a
// workgroup stages a tile of lines of the axis in LDS (Jacobian pre-scale fused
// into the load), contracts them with the stencil held in LDS, applies the
// post-scale and accumulates into `out`.
#include "sdv_internal.h"

namespace sdv {

constexpr int kKeThreads = 256;
// flags of the axis kernels: multiply each axis's result by its coefficient
// (mult_general option 2, mycoefs); axis 0 adds into `out` instead of
// overwriting it (the first derivative accumulates onto the kinetic apply).
constexpr int kKeCoef = 1, kKeAccumulate = 2;

// maskflag (scalingflag 0): divide_mask / mult_mask of SD/coreproject.f90:1273-1319
// divide by, then multiply with, the edge mask of the x, y and z coordinate of
// grid point g, in that order.
__device__ __forceinline__ z64 mask_divide(z64 v, const double* m, size_t g, int n) {
  const double mx = m[g % n], my = m[(g / n) % n], mz = m[g / ((size_t)n * n)];
  v = mk(v.re / mx, v.im / mx);
  v = mk(v.re / my, v.im / my);
  return mk(v.re / mz, v.im / mz);
}
__device__ __forceinline__ z64 mask_multiply(z64 v, const double* m, size_t g, int n) {
  return scale(scale(scale(v, m[g % n]), m[(g / n) % n]), m[g / ((size_t)n * n)]);
}
// the same, for a point whose coordinates are known
__device__ __forceinline__ z64 mask_divide_xyz(z64 v, const double* m, int x, int y, int z) {
  const double mx = m[x], my = m[y], mz = m[z];
  v = mk(v.re / mx, v.im / mx);
  v = mk(v.re / my, v.im / my);
  return mk(v.re / mz, v.im / mz);
}
__device__ __forceinline__ z64 mask_multiply_xyz(z64 v, const double* m, int x, int y, int z) {
  return scale(scale(scale(v, m[x]), m[y]), m[z]);
}

// Axis-d pass.  Element (a, p, b) sits at a + s*p + s*n*b with s = n^d.
// A tile is `la` consecutive inner indices a (d > 0) or `la` consecutive
// lines b (d == 0), for one orbital (blockIdx.z).
template <int AXIS>
__global__ void __launch_bounds__(kKeThreads) k_ke_axis(int n, const double* __restrict__ ke,
                                                         const z64* __restrict__ invj, const z64* __restrict__ sdiag,
                                                         const double* __restrict__ mask,
                                                         const z64* __restrict__ in, z64* __restrict__ out, int la,
                                                         z64 coef, int flags) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  double* st = reinterpret_cast<double*>(smem_raw);      // 2n-1 stencil
  z64* buf = reinterpret_cast<z64*>(st + ((2 * n - 1 + 1) & ~1));  // la * n
  const int tid = threadIdx.x;
  const size_t n3 = (size_t)n * n * n;
  const z64* src = in + (size_t)blockIdx.z * n3;
  z64* dst = out + (size_t)blockIdx.z * n3;
  const size_t s = AXIS == 0 ? 1 : (AXIS == 1 ? (size_t)n : (size_t)n * n);
  const size_t nouter = n3 / (s * n);
  // tile coordinates
  size_t a0, b0;
  int nl;
  if (AXIS == 0) {
    a0 = 0;
    b0 = (size_t)blockIdx.x * la;
    nl = (int)min((size_t)la, nouter - b0);
  } else {
    const size_t tiles_a = (s + la - 1) / la;
    a0 = (size_t)(blockIdx.x % tiles_a) * la;
    b0 = blockIdx.x / tiles_a;
    nl = (int)min((size_t)la, s - a0);
  }
  for (int i = tid; i < 2 * n - 1; i += kKeThreads) st[i] = ke[i];
  const z64* ij = invj ? invj + (size_t)AXIS * n : nullptr;
  for (int e = tid; e < nl * n; e += kKeThreads) {
    int l, p;
    size_t g;
    if (AXIS == 0) {
      l = e / n;
      p = e - l * n;
      g = (size_t)p + (size_t)n * (b0 + l);
    } else {
      p = e / nl;
      l = e - p * nl;
      g = (a0 + l) + s * p + s * n * b0;
    }
    z64 v = src[g];
    if (ij) v = v * ij[p];  // temp = in * invjacobian(:,d)   :1344-1346
    if (mask) v = mask_divide(v, mask, g, n);
    buf[l * n + p] = v;
  }
  __syncthreads();
  for (int e = tid; e < nl * n; e += kKeThreads) {
    int l, q;
    size_t g;
    if (AXIS == 0) {
      l = e / n;
      q = e - l * n;
      g = (size_t)q + (size_t)n * (b0 + l);
    } else {
      q = e / nl;
      l = e - q * nl;
      g = (a0 + l) + s * q + s * n * b0;
    }
    // mult_all0: out(.,q,.) = sum_p temp(.,p,.) * ketot(p,q), ketot(p,q) = kevect(q-p)
    double re = 0.0, im = 0.0;
    const z64* row = buf + l * n;
    for (int p = 0; p < n; ++p) {
      const double k = st[q - p + n - 1];
      re = fma(row[p].re, k, re);
      im = fma(row[p].im, k, im);
    }
    z64 r = mk(re, im);
    if (ij) r = r * ij[q];  // temp2 * invjacobian(:,d)   :1366-1368
    if (mask) r = mask_multiply(r, mask, g, n);
    if (flags & kKeCoef) r = r * coef;  // * mycoefs(jj)   :1368,1374
    if (AXIS == 0 && !(flags & kKeAccumulate)) {
      z64 o = mk(0.0, 0.0);
      dst[g] = o + r;
    } else if (AXIS != 2) {
      dst[g] = dst[g] + r;
    } else {
      z64 o = dst[g] + r;
      if (sdiag) {  // out + in * scalediag   :1382-1388
        const int x = (int)(g % n), y = (int)((g / n) % n), z = (int)(g / ((size_t)n * n));
        const z64 sd = ((mk(0.0, 0.0) + sdiag[x]) + sdiag[n + y]) + sdiag[2 * n + z];
        o = o + src[g] * sd;
      }
      dst[g] = o;
    }
  }
}

// ------------------------------------------------------------ MFMA form
// Each axis contraction is a matrix product on the matrix cores,
//   out(q, line) = sum_p  K(q, p) temp(p, line),   K(q, p) = ketot(q - p),
// with the real stencil as one operand and the real and imaginary parts of
// the lines as the other.  T = double: v_mfma_f64_16x16x4f64; T = float (mixed
// form): v_mfma_f32_16x16x4f32 on operands rounded once from FP64, scales and
// accumulation into `out` in FP64.
template <typename T>
struct Mma;
template <>
struct Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 run(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // accumulator element v of lane l holds row 4v + l/16 (gfx950, checked)
  static __device__ __forceinline__ int row(int l, int v) { return 4 * v + (l >> 4); }
};
template <>
struct Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 run(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  // accumulator element v of lane l holds row 4 (l/16) + v (gfx950, checked)
  static __device__ __forceinline__ int row(int l, int v) { return 4 * (l >> 4) + v; }
};

constexpr int kKeMaxQT = 6;                  // q tiles of 16: n <= 96

// ------------------------------------------------------- fused x/y pass
// The x and y contractions of one (orbital, z) plane from one staged copy of
// the plane (the z contraction follows in k_ke_z):
//   rx(x, y) = post_x(x) sum_p K(x, p) pre_x(p) in(p, y)
//   ry(x, y) = post_y(y) sum_p K(y, p) pre_y(p) in(x, p)
//   out(x, y) = ((out or 0) + rx) + ry          (out: the z pass's result)
// with K(q, p) = ketot(q - p), pre/post the ECS 1/J of the axis, the edge
// mask divided out before and multiplied back after, and the axis
// coefficient (SD/coreproject.f90:1321-1390).  Both terms are 16 x 16 output
// tiles D[y][x] on the matrix cores: for rx the plane is the A operand (rows
// y, k = p) and the stencil the B operand (k = p, columns x); for ry the
// stencil is A (rows y) and the plane B (columns x), so the two land in the
// same accumulator layout and every tile is written once, straight from the
// accumulators (16 consecutive x per row).  One persistent workgroup per CU
// (the FP64 plane takes 105 KB of LDS at n = 81) walks the (orbital, z)
// planes and loads the next plane into registers while the current plane's
// MFMAs run.  ECS in FP32: the plane is staged twice, pre-scaled by 1/J of x
// and of y in FP64 and rounded once (as the per-axis kernels do); in FP64 the
// pre-scale is applied to the operand as it is read.
template <typename T>
__device__ __forceinline__ cplx<T> cvt(z64 v) {
  return mk((T)v.re, (T)v.im);
}

constexpr int kKeXYWaves = 12;
constexpr int kKeXYThreads = 64 * kKeXYWaves;
constexpr int kKeXOn = 4, kKeYOn = 8;
// scaling of the x/y pass, a template parameter so the plain form carries no
// ECS or mask code (ECS and the edge mask exclude each other)
constexpr int kKePlain = 0, kKeEcs = 1, kKeMask = 2;

template <typename T>
__host__ __device__ constexpr bool ke_xy_two_planes(bool ecs) {
  return ecs && sizeof(T) == 4;
}
// LDS: the stencil as a table of K(q - p) for every q < 16 NT and p < NP
// (zero beyond the stencil), padded to 16 bytes; the x and y 1/J tables (2 NP
// z64, zero beyond n; ECS) or the mask (n doubles, padded); the plane(s): NP
// rows of pitch NP | 1, zero in the rows and columns n .. NP-1, so the
// contraction's padding needs no test in the MFMA loop.
__host__ __device__ constexpr int ke_xy_np(int n) { return (n + 3) & ~3; }
__host__ __device__ constexpr int ke_xy_sk(int n) { return 16 * ((n + 15) / 16) + ke_xy_np(n) - 1; }
template <typename T>
__host__ __device__ constexpr size_t ke_xy_tab_bytes(int n, int mode) {
  return (((size_t)ke_xy_sk(n) * sizeof(T) + 15) & ~(size_t)15) +
         (mode == kKeEcs ? 2 * (size_t)ke_xy_np(n) * sizeof(z64)
                         : mode == kKeMask ? (((size_t)n * 8 + 15) & ~(size_t)15) : 0);
}
template <typename T>
static size_t ke_xy_lds_bytes(int n, int mode) {
  return ke_xy_tab_bytes<T>(n, mode) + (ke_xy_two_planes<T>(mode == kKeEcs) ? 2 : 1) * (size_t)ke_xy_np(n) *
                                           (ke_xy_np(n) | 1) * sizeof(cplx<T>);
}

// BOTH: the x and y axes are both on (every kinetic apply; the velocity gauge
// may switch one off)
template <typename T, int LD, int MODE, bool BOTH>
__global__ void __launch_bounds__(kKeXYThreads) k_ke_xy(int n, int nplane, const double* __restrict__ ke,
                                                         const z64* __restrict__ invj,
                                                         const double* __restrict__ mask,
                                                         const z64* __restrict__ in, z64* __restrict__ out, z64 cx,
                                                         z64 cy, int flags) {
  typedef typename Mma<T>::v4 v4;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int NP = ke_xy_np(n), pitch = NP | 1, ne = n * n, SK = ke_xy_sk(n), OFF = NP - 1;
  // MFMA tiles per axis; with n = 16 k + 1 (81 = 5 x 16 + 1) the last row and
  // column (2n - 1 points) are dot products on the VALU instead of a sixth tile
  // row and column that would hold one point in sixteen
  const bool edge = n % 16 == 1;
  const int NT = edge ? n / 16 : (n + 15) / 16;
  const size_t n2 = (size_t)n * n, n3 = n2 * n;
  constexpr bool ECS = MODE == kKeEcs, MASK = MODE == kKeMask, two = ke_xy_two_planes<T>(ECS);
  T* sk = reinterpret_cast<T*>(smem_raw);  // sk[q - p + OFF] = K(q, p)
  unsigned char* tab = smem_raw + (((size_t)SK * sizeof(T) + 15) & ~(size_t)15);
  z64* ijs = reinterpret_cast<z64*>(tab);       // ECS: 1/J of x (ijs[p]) and of y (ijs[NP + p])
  double* ms = reinterpret_cast<double*>(tab);  // mask
  cplx<T>* plx = reinterpret_cast<cplx<T>*>(smem_raw + ke_xy_tab_bytes<T>(n, MODE));
  cplx<T>* ply = two ? plx + (size_t)NP * pitch : plx;
  const z64* ijx = ijs;
  const z64* ijy = ijs + NP;
  const bool on_x = BOTH || (flags & kKeXOn), on_y = BOTH || (flags & kKeYOn);
  const int tid = threadIdx.x, w = tid >> 6, l = tid & 63, i16 = l & 15, k4 = l >> 4;
  for (int i = tid; i < SK; i += kKeXYThreads) {
    const int d = i - OFF;
    sk[i] = (d > -n && d < n) ? (T)ke[d + n - 1] : (T)0;
  }
  if (ECS)
    for (int i = tid; i < 2 * NP; i += kKeXYThreads) {
      const int p = i < NP ? i : i - NP;
      ijs[i] = p < n ? invj[(i < NP ? 0 : n) + p] : mk(0.0, 0.0);
    }
  if (MASK)
    for (int i = tid; i < n; i += kKeXYThreads) ms[i] = mask[i];
  for (int i = tid; i < (two ? 2 : 1) * NP * pitch; i += kKeXYThreads) plx[i] = mk((T)0, (T)0);

  // Tasks: the planes in full rounds of gridDim.x, then the R planes left
  // over each split S ways (S = gridDim.x / R, at most one per tile), sub-task
  // s taking the tiles and edge points congruent to s mod S: every workgroup
  // gets the same number of tasks instead of R workgroups running a last
  // round alone (81^3 with 16 orbitals: 1296 planes = 5 x 256 + 16).
  const int G = gridDim.x, R = nplane % G, NT2 = NT * NT;
  const int S = R ? max(1, min(G / R, NT2)) : 1, nfull = nplane - R, ntask = nfull + R * S;
  auto plane_of = [&](int t) { return t < nfull ? t : nfull + (t - nfull) / S; };
  // ECS (FP64: operands scaled as they are read) keeps no next plane in
  // registers: it would spill there
  constexpr bool PF = !ECS;
  z64 pf[PF ? LD : 1];  // the next plane, in flight during the current plane's MFMAs
  auto prefetch = [&](int task) {
    const int pl = plane_of(task);
    const z64* src = in + (size_t)(pl / n) * n3 + (size_t)(pl % n) * n2;
#pragma unroll
    for (int k = 0; k < (PF ? LD : 0); ++k) {
      const int e = tid + k * kKeXYThreads;
      if (e < ne) pf[k] = src[e];
    }
  };
  // out(x, y) of plane z of orbital o from the two contractions
  // (prev: out(x, y) as read before the contractions, or 0)
  auto emit = [&](int o, int z, int x, int y, z64 rx, z64 ry, z64 prev) {
    const size_t g = (size_t)o * n3 + (size_t)z * n2 + (size_t)y * n + x;
    z64 r = prev;
    if (on_x) {
      if (ECS) rx = rx * ijx[x];                          // temp2 * invjacobian   :1366-1368
      if (MASK) rx = mask_multiply_xyz(rx, ms, x, y, z);  // mult_mask   :1373-1374
      if (flags & kKeCoef) rx = rx * cx;                  // * mycoefs   :1368,1374
      r = r + rx;
    }
    if (on_y) {
      if (ECS) ry = ry * ijy[y];
      if (MASK) ry = mask_multiply_xyz(ry, ms, x, y, z);
      if (flags & kKeCoef) ry = ry * cy;
      r = r + ry;
    }
    out[g] = r;
  };
  if (PF && (int)blockIdx.x < ntask) prefetch(blockIdx.x);
  for (int task = blockIdx.x; task < ntask; task += G) {
    const int pl = plane_of(task), sub = task < nfull ? 0 : (task - nfull) % S, ss = task < nfull ? 1 : S;
    const int o = pl / n, z = pl - o * n;
    const z64* src = in + (size_t)o * n3 + (size_t)z * n2;
    __syncthreads();  // every wave is done with the previous plane (and the tables are stored)
    // element e = tid + k * threads sits at (y, x); stepped without divisions
    int y = tid / n, x = tid - (tid / n) * n;
    const int dy = kKeXYThreads / n, dx = kKeXYThreads - dy * n;
#pragma unroll
    for (int k = 0; k < LD; ++k) {
      const int e = tid + k * kKeXYThreads;
      if (e < ne) {
        z64 v = PF ? pf[PF ? k : 0] : src[e];
        if (MASK) v = mask_divide_xyz(v, ms, x, y, z);  // divide_mask   :1355-1356
        if (two) {  // temp = in * invjacobian(:,d), rounded once   :1344-1346
          plx[y * pitch + x] = cvt<T>(v * ijx[x]);
          ply[y * pitch + x] = cvt<T>(v * ijy[y]);
        } else {
          plx[y * pitch + x] = cvt<T>(v);
        }
      }
      x += dx;
      y += dy;
      if (x >= n) {
        x -= n;
        ++y;
      }
    }
    __syncthreads();
    if (PF && task + G < ntask) prefetch(task + G);
    for (int tile = sub + ss * w; tile < NT2; tile += ss * kKeXYWaves) {
      const int yt = tile / NT, xt = tile - yt * NT;
      // operand addresses at p0 = 0 (rows and columns beyond the plane read
      // a valid row: their outputs are discarded)
      const cplx<T>* ax = plx + min(16 * yt + i16, NP - 1) * pitch + k4;  // rx: A = plane(row y, col p), + p0
      const cplx<T>* by = ply + k4 * pitch + min(16 * xt + i16, NP - 1);  // ry: B = plane(row p, col x), + p0 pitch
      const T* bx = sk + 16 * xt + i16 - k4 + OFF;                         // rx: B = K(x, p), - p0
      const T* ay = sk + 16 * yt + i16 - k4 + OFF;                         // ry: A = K(y, p), - p0
      // the tile's values of out, read now and added after the contractions
      const int xo = 16 * xt + i16;
      z64 prev[4];
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int y = 16 * yt + Mma<T>::row(l, v);
        prev[v] = mk(0.0, 0.0);
        if ((flags & kKeAccumulate) && y < n && xo < n) prev[v] = out[(size_t)o * n3 + (size_t)z * n2 + (size_t)y * n + xo];
      }
      v4 xr = {0, 0, 0, 0}, xi = {0, 0, 0, 0}, yr = {0, 0, 0, 0}, yi = {0, 0, 0, 0};
      // two operand sets, each read a half step ahead of its MFMAs (no
      // register copies, so no wait is forced inside the half step)
      auto step = [&](const cplx<T>& da, const cplx<T>& db, T sa, T sb, int p) {
        if (on_x) {
          const cplx<T> a = (ECS && !two) ? cvt<T>(mk((double)da.re, (double)da.im) * ijx[p + k4]) : da;
          xr = Mma<T>::run(a.re, sa, xr);
          xi = Mma<T>::run(a.im, sa, xi);
        }
        if (on_y) {
          const cplx<T> b = (ECS && !two) ? cvt<T>(mk((double)db.re, (double)db.im) * ijy[p + k4]) : db;
          yr = Mma<T>::run(sb, b.re, yr);
          yi = Mma<T>::run(sb, b.im, yi);
        }
      };
      cplx<T> da0 = ax[0], db0 = by[0], da1 = da0, db1 = db0;
      T sa0 = bx[0], sb0 = ay[0], sa1 = sa0, sb1 = sb0;
      for (int p0 = 0; p0 < NP; p0 += 8) {
        const bool h1 = p0 + 4 < NP, h2 = p0 + 8 < NP;
        if (h1) {
          if (on_x) {
            da1 = ax[p0 + 4];
            sa1 = bx[-(p0 + 4)];
          }
          if (on_y) {
            db1 = by[(p0 + 4) * pitch];
            sb1 = ay[-(p0 + 4)];
          }
        }
        step(da0, db0, sa0, sb0, p0);
        if (h2) {
          if (on_x) {
            da0 = ax[p0 + 8];
            sa0 = bx[-(p0 + 8)];
          }
          if (on_y) {
            db0 = by[(p0 + 8) * pitch];
            sb0 = ay[-(p0 + 8)];
          }
        }
        if (h1) step(da1, db1, sa1, sb1, p0 + 4);
      }
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int y = 16 * yt + Mma<T>::row(l, v);
        if (y < n && xo < n) emit(o, z, xo, y, mk((double)xr[v], (double)xi[v]), mk((double)yr[v], (double)yi[v]), prev[v]);
      }
    }
    // the last row and column on the VALU, by the wavefronts that hold one
    // tile fewer (lane e: point (n-1, e) for e < n, else (e-n, n-1))
    if (edge) {
      const int nts = (NT2 - sub + ss - 1) / ss;  // tiles of this (sub-)task
      const int w0 = nts < kKeXYWaves ? nts : nts % kKeXYWaves, e = sub + ss * ((w - w0) * 64 + l);
      if (w >= w0 && e < 2 * n - 1) {
        const int x = e < n ? n - 1 : e - n, y = e < n ? e : n - 1;
        double xre = 0.0, xim = 0.0, yre = 0.0, yim = 0.0;
        for (int p = 0; p < n; ++p) {
          cplx<T> a = plx[y * pitch + p];  // rx: K(x, p) pre_x(p) in(p, y)
          cplx<T> b = ply[p * pitch + x];  // ry: K(y, p) pre_y(p) in(x, p)
          if (ECS && !two) {
            a = cvt<T>(mk((double)a.re, (double)a.im) * ijx[p]);
            b = cvt<T>(mk((double)b.re, (double)b.im) * ijy[p]);
          }
          const T kx = sk[x - p + OFF], ky = sk[y - p + OFF];
          xre = fma((double)a.re, (double)kx, xre);
          xim = fma((double)a.im, (double)kx, xim);
          yre = fma((double)b.re, (double)ky, yre);
          yim = fma((double)b.im, (double)ky, yim);
        }
        const z64 prev = (flags & kKeAccumulate) ? out[(size_t)o * n3 + (size_t)z * n2 + (size_t)y * n + x] : mk(0.0, 0.0);
        emit(o, z, x, y, mk(xre, xim), mk(yre, yim), prev);
      }
    }
  }
}
// ---------------------------------------------------------------- z pass
// The z contraction, the first of the two passes (SD/coreproject.f90:1321-1390):
//   out(x, y, q) = (out or 0) + post_z(q) sum_p K(q, p) pre_z(p) in(x, y, p)  [+ scalediag in]
// from registers: one wavefront per 16 consecutive (x, y) columns of one
// orbital.  Lane l holds the B operand of every 4-step s of the contraction,
// column 16 cb + l % 16 at z = 4 s + l / 16 (for one s the wavefront reads
// four runs of 16 consecutive columns, 256 bytes each in FP64); the stencil
// is the A operand, read from a Toeplitz table in LDS; each 16-row q tile is
// added into `out` straight from its accumulators, 16 consecutive columns per
// row.  The data never passes through LDS and the wavefronts share no
// barrier, so a CU's wavefronts overlap one another's loads and MFMAs.  It
// runs first because it is the memory-bound pass: writing `out` without
// reading it leaves the read-modify-write of `out` to the MFMA-bound x/y
// pass, which reads it ahead of each tile (81^3 FP64, 16 orbitals: z pass
// 108 -> 90 us, whole apply 0.27 -> 0.25 ms).
constexpr int kKeZWaves = 4;

// three wavefronts per SIMD (at most 168 registers): the overlap of one
// wavefront's loads with another's MFMAs is what this pass runs on
template <typename T, int NS, int MODE>
__global__ void __launch_bounds__(64 * kKeZWaves) __attribute__((amdgpu_waves_per_eu(3))) k_ke_z(int n, int ntask, const double* __restrict__ ke,
                                                          const z64* __restrict__ invj,
                                                          const z64* __restrict__ sdiag,
                                                          const double* __restrict__ mask,
                                                          const z64* __restrict__ in, z64* __restrict__ out, z64 cz,
                                                          int flags) {
  typedef typename Mma<T>::v4 v4;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int NP = ke_xy_np(n), SK = ke_xy_sk(n), OFF = NP - 1, QT = (n + 15) / 16, ns = NP / 4;
  constexpr bool ECS = MODE == kKeEcs, MASK = MODE == kKeMask;
  T* sk = reinterpret_cast<T*>(smem_raw);  // sk[q - p + OFF] = K(q, p), zero beyond the stencil
  const int tid = threadIdx.x, l = tid & 63, i16 = l & 15, k4 = l >> 4;
  for (int i = tid; i < SK; i += 64 * kKeZWaves) {
    const int d = i - OFF;
    sk[i] = (d > -n && d < n) ? (T)ke[d + n - 1] : (T)0;
  }
  __syncthreads();
  const int task = blockIdx.x * kKeZWaves + (tid >> 6);
  if (task >= ntask) return;  // after the kernel's only barrier
  const int n2 = n * n, cbs = (n2 + 15) / 16;
  const int o = task / cbs, cb = task - o * cbs;
  const int col = 16 * cb + i16;  // (x, y) = (col % n, col / n)
  const bool colok = col < n2;
  const int xc = colok ? col % n : 0, yc = colok ? col / n : 0;
  const size_t base = (size_t)o * n2 * n + (colok ? col : 0);
  const z64* ijz = ECS ? invj + 2 * n : nullptr;
  cplx<T> t[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int p = 4 * s + k4;
    t[s] = mk((T)0, (T)0);
    if (s < ns && p < n && colok) {
      z64 v = in[base + (size_t)p * n2];
      if (ECS) v = v * ijz[p];                              // temp = in * invjacobian(:,3)   :1344-1346
      if (MASK) v = mask_divide_xyz(v, mask, xc, yc, p);    // divide_mask   :1355-1356
      t[s] = cvt<T>(v);
    }
  }
  // out(x, y, q) from the contraction r
  auto put = [&](int q, z64 r) {
    const size_t g = base + (size_t)q * n2;
    if (ECS) r = r * ijz[q];                              // temp2 * invjacobian   :1366-1368
    if (MASK) r = mask_multiply_xyz(r, mask, xc, yc, q);  // mult_mask   :1373-1374
    if (flags & kKeCoef) r = r * cz;                      // * mycoefs   :1368,1374
    z64 acc = (flags & kKeAccumulate) ? out[g] + r : mk(0.0, 0.0) + r;
    if (sdiag) {  // out + in * scalediag   :1382-1388
      const z64 sd = ((mk(0.0, 0.0) + sdiag[xc]) + sdiag[n + yc]) + sdiag[2 * n + q];
      acc = acc + in[g] * sd;
    }
    out[g] = acc;
  };
  // with n = 16 k + 1 the last row q = n - 1 is a dot product on the VALU
  // (each lane's share of p, summed over the column's four lanes) instead of
  // a seventeenth-row tile
  const bool edge = n % 16 == 1;
  const int QTM = edge ? n / 16 : QT;
  const T* arow = sk + i16 - k4 + OFF;  // A(q, p) = sk[16 qt + i16 - (4 s + k4) + OFF]
  for (int qt = 0; qt < QTM; ++qt) {
    v4 ar = {0, 0, 0, 0}, ai = {0, 0, 0, 0};
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (s < ns) {
        const T a = arow[16 * qt - 4 * s];
        ar = Mma<T>::run(a, t[s].re, ar);
        ai = Mma<T>::run(a, t[s].im, ai);
      }
    }
#pragma unroll
    for (int v = 0; v < 4; ++v) {
      const int q = 16 * qt + Mma<T>::row(l, v);
      if (q < n && colok) put(q, mk((double)ar[v], (double)ai[v]));
    }
  }
  if (edge) {
    double re = 0.0, im = 0.0;
    const T* krow = sk + (n - 1) - k4 + OFF;  // K(n - 1, 4 s + k4)
#pragma unroll
    for (int s = 0; s < NS; ++s) {
      if (s < ns) {
        const double k = (double)krow[-4 * s];
        re = fma((double)t[s].re, k, re);
        im = fma((double)t[s].im, k, im);
      }
    }
    re += __shfl_xor(re, 16);
    im += __shfl_xor(im, 16);
    re += __shfl_xor(re, 32);
    im += __shfl_xor(im, 32);
    if (k4 == 0 && colok) put(n - 1, mk(re, im));
  }
}

template <typename T, int NS, int MODE>
static int ke_z_launch1(int n, int howmany, hipStream_t st, const KeArgsHost& a) {
  const int ntask = howmany * ((n * n + 15) / 16);
  const size_t sh = (((size_t)ke_xy_sk(n) * sizeof(T) + 15) & ~(size_t)15);
  hipLaunchKernelGGL((k_ke_z<T, NS, MODE>), dim3((ntask + kKeZWaves - 1) / kKeZWaves), dim3(64 * kKeZWaves), sh, st,
                     n, ntask, a.ke, (const z64*)a.invj, (const z64*)a.sdiag, a.mask, (const z64*)a.in, (z64*)a.out,
                     a.coef[2], (a.use_coef ? kKeCoef : 0) | (a.accumulate ? kKeAccumulate : 0));
  return SDV_OK;
}
template <typename T>
static int ke_z_launch(int n, hipStream_t st, const KeArgsHost& a) {
  const int mode = a.invj ? kKeEcs : a.mask ? kKeMask : kKePlain;
  // register classes of the contraction's B operands: 4-steps <= 12, 21 (n <= 84), 24
  const int ns = ke_xy_np(n) / 4;
  if (mode == kKeEcs)
    return ns <= 12 ? ke_z_launch1<T, 12, kKeEcs>(n, a.howmany, st, a)
         : ns <= 21 ? ke_z_launch1<T, 21, kKeEcs>(n, a.howmany, st, a)
                    : ke_z_launch1<T, 24, kKeEcs>(n, a.howmany, st, a);
  if (mode == kKeMask)
    return ns <= 12 ? ke_z_launch1<T, 12, kKeMask>(n, a.howmany, st, a)
         : ns <= 21 ? ke_z_launch1<T, 21, kKeMask>(n, a.howmany, st, a)
                    : ke_z_launch1<T, 24, kKeMask>(n, a.howmany, st, a);
  return ns <= 12 ? ke_z_launch1<T, 12, kKePlain>(n, a.howmany, st, a)
       : ns <= 21 ? ke_z_launch1<T, 21, kKePlain>(n, a.howmany, st, a)
                  : ke_z_launch1<T, 24, kKePlain>(n, a.howmany, st, a);
}

// compute units of the current device (the persistent x/y pass takes one
// workgroup per CU), cached per device
static int cu_count() { return device_cu_count(); }

// An axis with a zero coefficient is skipped, as the reference skips it
// (SD/coreproject.f90:1339); only when the axes accumulate into `out`.
static bool axis_on(const KeArgsHost& a, int d) {
  return !(a.use_coef && a.accumulate && a.coef[d].re == 0.0 && a.coef[d].im == 0.0);
}

template <typename T, int LD, int MODE, bool BOTH>
static int ke_xy_launch2(int grid, size_t shxy, hipStream_t st, int n, int ntask, const KeArgsHost& a, const z64* ij,
                         int fl) {
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_ke_xy<T, LD, MODE, BOTH>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shxy));
  hipLaunchKernelGGL((k_ke_xy<T, LD, MODE, BOTH>), dim3(grid), dim3(kKeXYThreads), shxy, st, n, ntask, a.ke, ij,
                     a.mask, (const z64*)a.in, (z64*)a.out, a.coef[0], a.coef[1], fl);
  return SDV_OK;
}
template <typename T, int LD, int MODE>
static int ke_xy_launch1(int grid, size_t shxy, hipStream_t st, int n, int ntask, const KeArgsHost& a, const z64* ij,
                         int fl) {
  return (fl & kKeXOn) && (fl & kKeYOn) ? ke_xy_launch2<T, LD, MODE, true>(grid, shxy, st, n, ntask, a, ij, fl)
                                        : ke_xy_launch2<T, LD, MODE, false>(grid, shxy, st, n, ntask, a, ij, fl);
}
template <typename T, int LD>
static int ke_xy_launch(int mode, int grid, size_t shxy, hipStream_t st, int n, int ntask, const KeArgsHost& a,
                        const z64* ij, int fl) {
  if (mode == kKeEcs) return ke_xy_launch1<T, LD, kKeEcs>(grid, shxy, st, n, ntask, a, ij, fl);
  if (mode == kKeMask) return ke_xy_launch1<T, LD, kKeMask>(grid, shxy, st, n, ntask, a, ij, fl);
  return ke_xy_launch1<T, LD, kKePlain>(grid, shxy, st, n, ntask, a, ij, fl);
}

template <typename T>
static int ke_launch_mma(const KeArgsHost& a, hipStream_t st) {
  const int n = a.n;
  SDV_REQUIRE(n <= 16 * kKeMaxQT, SDV_ERR_UNSUPPORTED, "grid edge too large for the MFMA kinetic tile");
  const z64* ij = (const z64*)a.invj;
  const int flags = (a.use_coef ? kKeCoef : 0) | (a.accumulate ? kKeAccumulate : 0);
  const bool on0 = axis_on(a, 0), on1 = axis_on(a, 1), on2 = axis_on(a, 2);
  if (on2) {  // z first: out = (out or 0) + z term (+ scalediag in)
    const int rc = ke_z_launch<T>(n, st, a);
    if (rc != SDV_OK) return rc;
  }
  if (on0 || on1) {  // x and y together: k_ke_xy, adding into the z pass's out
    const int ld = (n * n + kKeXYThreads - 1) / kKeXYThreads;
    const int mode = ij ? kKeEcs : a.mask ? kKeMask : kKePlain;
    const size_t shxy = ke_xy_lds_bytes<T>(n, mode);
    SDV_REQUIRE(ld <= 12 && shxy <= 160 * 1024, SDV_ERR_UNSUPPORTED, "grid edge too large for the x/y kinetic pass");
    const int ntask = a.howmany * n;
    const int grid = std::min(ntask, cu_count());
    const int fl = flags | (on0 ? kKeXOn : 0) | (on1 ? kKeYOn : 0) | (on2 ? kKeAccumulate : 0);
    const int rc = ld <= 9 ? ke_xy_launch<T, 9>(mode, grid, shxy, st, n, ntask, a, ij, fl)
                           : ke_xy_launch<T, 12>(mode, grid, shxy, st, n, ntask, a, ij, fl);
    if (rc != SDV_OK) return rc;
  }
  SDV_HIP_CHECK(hipGetLastError());
  return SDV_OK;
}

// precision 0: FP64 matrix cores; 1: FP32 matrix cores (mixed form); the
// FP64 VALU/LDS form (k_ke_axis) for precision 2 and for grid edges beyond
// the matrix-core tile (n > 96).
int ke_launch(const KeArgsHost& a, int precision, hipStream_t st, Prof* prof) {
  if (a.howmany <= 0) return SDV_OK;
  const int n = a.n;
  hipEvent_t e0 = prof ? prof->begin(st) : nullptr;
  if ((precision == 0 || precision == 1) && n <= 16 * kKeMaxQT) {
    const int rc = precision == 0 ? ke_launch_mma<double>(a, st) : ke_launch_mma<float>(a, st);
    if (rc != SDV_OK) return rc;
    if (prof) prof->end(SDV_STAGE_KE, e0, st);
    return SDV_OK;
  }
  const size_t n3 = (size_t)n * n * n;
  const int la0 = 32;
  const size_t sh0 = ((size_t)((2 * n - 1 + 1) & ~1)) * sizeof(double) + (size_t)la0 * n * sizeof(z64);
  SDV_REQUIRE(sh0 <= 64 * 1024, SDV_ERR_UNSUPPORTED, "grid edge too large for the KE tile");
  const z64* ij = (const z64*)a.invj;
  const z64* sd = (const z64*)a.sdiag;
  const int flags = (a.use_coef ? kKeCoef : 0) | (a.accumulate ? kKeAccumulate : 0);
  if (axis_on(a, 0)) {
    const unsigned nb = (unsigned)((n3 / n + la0 - 1) / la0);
    hipLaunchKernelGGL((k_ke_axis<0>), dim3(nb, 1, a.howmany), dim3(kKeThreads), sh0, st, n, a.ke, ij, sd,
                       a.mask, (const z64*)a.in, (z64*)a.out, la0, a.coef[0], flags);
  }
  if (axis_on(a, 1)) {
    const size_t s = n;
    const unsigned nb = (unsigned)(((s + la0 - 1) / la0) * (n3 / (s * n)));
    hipLaunchKernelGGL((k_ke_axis<1>), dim3(nb, 1, a.howmany), dim3(kKeThreads), sh0, st, n, a.ke, ij, sd,
                       a.mask, (const z64*)a.in, (z64*)a.out, la0, a.coef[1], flags);
  }
  if (axis_on(a, 2)) {
    const size_t s = (size_t)n * n;
    const unsigned nb = (unsigned)(((s + la0 - 1) / la0) * (n3 / (s * n)));
    hipLaunchKernelGGL((k_ke_axis<2>), dim3(nb, 1, a.howmany), dim3(kKeThreads), sh0, st, n, a.ke, ij, sd,
                       a.mask, (const z64*)a.in, (z64*)a.out, la0, a.coef[2], flags);
  }
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_KE, e0, st);
  return SDV_OK;
}

}  // namespace sdv
