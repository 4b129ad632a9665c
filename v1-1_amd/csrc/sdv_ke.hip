// sdv_ke.hip -- sinc-DVR kinetic energy apply (mult_ke): FP64 or FP32 matrix
// cores (k_ke_mma), and an FP64 VALU form (k_ke_axis).
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
// One launch per axis: a workgroup stages a tile of lines of the axis in LDS
// (Jacobian pre-scale fused into the load), contracts them with the stencil
// held in LDS, applies the post-scale and accumulates into `out`.
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
// The per-axis contraction as a matrix product on the matrix cores:
//   out(q, line) = sum_p  K(q, p) temp(p, line),   K(q, p) = ketot(q - p)
// with the real stencil as the A operand (16 q x 4 p per instruction) and the
// real and imaginary parts of a 16-line tile as two B operands.  A workgroup
// (4 wavefronts) stages 32 lines x n positions (pre-scaled, zero-padded to a
// multiple of 4) in LDS; wave w takes lines 16 (w & 1) .. +16 and every other
// 16-row q tile; results go back through LDS so the post-scale and the
// accumulation into `out` are coalesced.  T = double: v_mfma_f64_16x16x4f64;
// T = float (mixed form): v_mfma_f32_16x16x4f32 on operands rounded once from
// FP64, scales and accumulation into `out` in FP64.
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

constexpr int kKeLines = 32;                 // lines per workgroup tile
constexpr int kKeMaxQT = 6;                  // q tiles of 16: n <= 96

template <typename T, int AXIS>
__global__ void __launch_bounds__(kKeThreads) k_ke_mma(int n, const double* __restrict__ ke,
                                                        const z64* __restrict__ invj, const z64* __restrict__ sdiag,
                                                        const double* __restrict__ mask,
                                                        const z64* __restrict__ in, z64* __restrict__ out, z64 coef,
                                                        int flags) {
  typedef typename Mma<T>::v4 v4;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int NP = (n + 3) & ~3, QT = (n + 15) / 16, pitch = NP + 1;
  T* st = reinterpret_cast<T*>(smem_raw);                                  // 2n-1 stencil
  cplx<T>* buf = reinterpret_cast<cplx<T>*>(smem_raw + (((2 * n - 1) * sizeof(T) + 15) & ~(size_t)15));
  const int tid = threadIdx.x;
  const size_t n3 = (size_t)n * n * n;
  const z64* src = in + (size_t)blockIdx.z * n3;
  z64* dst = out + (size_t)blockIdx.z * n3;
  const size_t s = AXIS == 0 ? 1 : (AXIS == 1 ? (size_t)n : (size_t)n * n);
  const size_t nouter = n3 / (s * n);
  size_t a0, b0;
  int nl;
  if (AXIS == 0) {
    a0 = 0;
    b0 = (size_t)blockIdx.x * kKeLines;
    nl = (int)min((size_t)kKeLines, nouter - b0);
  } else {
    const size_t tiles_a = (s + kKeLines - 1) / kKeLines;
    a0 = (size_t)(blockIdx.x % tiles_a) * kKeLines;
    b0 = blockIdx.x / tiles_a;
    nl = (int)min((size_t)kKeLines, s - a0);
  }
  for (int i = tid; i < 2 * n - 1; i += kKeThreads) st[i] = (T)ke[i];
  const z64* ij = invj ? invj + (size_t)AXIS * n : nullptr;
  // stage: buf[line][p] = temp(p, line); lines >= nl and p >= n are zero
  for (int e = tid; e < kKeLines * NP; e += kKeThreads) {
    int l, p;
    if (AXIS == 0) {
      l = e / NP;
      p = e - l * NP;
    } else {
      p = e / kKeLines;
      l = e - p * kKeLines;
    }
    cplx<T> v = mk((T)0, (T)0);
    if (l < nl && p < n) {
      const size_t g = AXIS == 0 ? (size_t)p + (size_t)n * (b0 + l) : (a0 + l) + s * p + s * n * b0;
      z64 x = src[g];
      if (ij) x = x * ij[p];  // temp = in * invjacobian(:,d)   SD/coreproject.f90:1344-1346
      if (mask) x = mask_divide(x, mask, g, n);  // divide_mask   :1355-1356
      v = mk((T)x.re, (T)x.im);
    }
    buf[l * pitch + p] = v;
  }
  __syncthreads();
  const int w = tid >> 6, l = tid & 63;
  const int lane_line = 16 * (w & 1) + (l & 15);  // B column / output line of this lane
  const int kk = l >> 4;                          // A column / B row within the 4-step
  v4 acc_re[(kKeMaxQT + 1) / 2], acc_im[(kKeMaxQT + 1) / 2];
#pragma unroll
  for (int t = 0; t < (kKeMaxQT + 1) / 2; ++t) {
    acc_re[t] = v4{0, 0, 0, 0};
    acc_im[t] = v4{0, 0, 0, 0};
  }
  const cplx<T>* brow = buf + lane_line * pitch + kk;
  for (int p0 = 0; p0 < NP; p0 += 4) {
    const cplx<T> b = brow[p0];
#pragma unroll
    for (int t = 0; t < (kKeMaxQT + 1) / 2; ++t) {
      const int qt = 2 * t + (w >> 1);
      if (qt < QT) {
        const int q = 16 * qt + (l & 15);
        const int d = min(max(q - (p0 + kk) + n - 1, 0), 2 * n - 2);  // clamped: B is zero where unused
        const T a = st[d];
        acc_re[t] = Mma<T>::run(a, b.re, acc_re[t]);
        acc_im[t] = Mma<T>::run(a, b.im, acc_im[t]);
      }
    }
  }
  __syncthreads();  // every wave has read buf
#pragma unroll
  for (int t = 0; t < (kKeMaxQT + 1) / 2; ++t) {
    const int qt = 2 * t + (w >> 1);
    if (qt < QT) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int q = 16 * qt + Mma<T>::row(l, v);
        if (q < NP) buf[lane_line * pitch + q] = mk(acc_re[t][v], acc_im[t][v]);
      }
    }
  }
  __syncthreads();
  for (int e = tid; e < nl * n; e += kKeThreads) {
    int ln, q;
    size_t g;
    if (AXIS == 0) {
      ln = e / n;
      q = e - ln * n;
      g = (size_t)q + (size_t)n * (b0 + ln);
    } else {
      q = e / nl;
      ln = e - q * nl;
      g = (a0 + ln) + s * q + s * n * b0;
    }
    const cplx<T> rv = buf[ln * pitch + q];
    z64 r = mk((double)rv.re, (double)rv.im);
    if (ij) r = r * ij[q];  // temp2 * invjacobian(:,d)   :1366-1368
    if (mask) r = mask_multiply(r, mask, g, n);  // mult_mask   :1373-1374
    if (flags & kKeCoef) r = r * coef;  // * mycoefs(jj)   :1368,1374
    if (AXIS == 0 && !(flags & kKeAccumulate)) {
      dst[g] = mk(0.0, 0.0) + r;
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

// An axis with a zero coefficient is skipped, as the reference skips it
// (SD/coreproject.f90:1339); only when the axes accumulate into `out`.
static bool axis_on(const KeArgsHost& a, int d) {
  return !(a.use_coef && a.accumulate && a.coef[d].re == 0.0 && a.coef[d].im == 0.0);
}

template <typename T>
static int ke_launch_mma(const KeArgsHost& a, hipStream_t st) {
  const int n = a.n;
  const size_t n3 = (size_t)n * n * n;
  const int NP = (n + 3) & ~3;
  const size_t sh = (((2 * n - 1) * sizeof(T) + 15) & ~(size_t)15) + (size_t)kKeLines * (NP + 1) * sizeof(cplx<T>);
  SDV_REQUIRE(n <= 16 * kKeMaxQT, SDV_ERR_UNSUPPORTED, "grid edge too large for the MFMA kinetic tile");
  const z64* ij = (const z64*)a.invj;
  const z64* sd = (const z64*)a.sdiag;
  const int flags = (a.use_coef ? kKeCoef : 0) | (a.accumulate ? kKeAccumulate : 0);
  if (axis_on(a, 0)) {
    const unsigned nb0 = (unsigned)((n3 / n + kKeLines - 1) / kKeLines);
    hipLaunchKernelGGL((k_ke_mma<T, 0>), dim3(nb0, 1, a.howmany), dim3(kKeThreads), sh, st, n, a.ke, ij, sd,
                       a.mask, (const z64*)a.in, (z64*)a.out, a.coef[0], flags);
  }
  if (axis_on(a, 1)) {
    const unsigned nb1 = (unsigned)(((size_t)(n + kKeLines - 1) / kKeLines) * n);
    hipLaunchKernelGGL((k_ke_mma<T, 1>), dim3(nb1, 1, a.howmany), dim3(kKeThreads), sh, st, n, a.ke, ij, sd,
                       a.mask, (const z64*)a.in, (z64*)a.out, a.coef[1], flags);
  }
  if (axis_on(a, 2)) {
    const unsigned nb2 = (unsigned)(((size_t)n * n + kKeLines - 1) / kKeLines);
    hipLaunchKernelGGL((k_ke_mma<T, 2>), dim3(nb2, 1, a.howmany), dim3(kKeThreads), sh, st, n, a.ke, ij, sd,
                       a.mask, (const z64*)a.in, (z64*)a.out, a.coef[2], flags);
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
