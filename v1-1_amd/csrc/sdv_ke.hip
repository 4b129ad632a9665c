// sdv_ke.hip -- sinc-DVR kinetic energy apply (mult_ke), FP64.
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

// Axis-d pass.  Element (a, p, b) sits at a + s*p + s*n*b with s = n^d.
// A tile is `la` consecutive inner indices a (d > 0) or `la` consecutive
// lines b (d == 0), for one orbital (blockIdx.z).
template <int AXIS>
__global__ void __launch_bounds__(kKeThreads) k_ke_axis(int n, const double* __restrict__ ke,
                                                         const z64* __restrict__ invj, const z64* __restrict__ sdiag,
                                                         const z64* __restrict__ in, z64* __restrict__ out, int la) {
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
    if (AXIS == 0) {
      z64 o = mk(0.0, 0.0);
      dst[g] = o + r;
    } else if (AXIS == 1) {
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

int ke_launch(const KeArgsHost& a, int precision, hipStream_t st, Prof* prof) {
  (void)precision;
  if (a.howmany <= 0) return SDV_OK;
  const int n = a.n;
  const size_t n3 = (size_t)n * n * n;
  const int la0 = 32;
  const size_t sh0 = ((size_t)((2 * n - 1 + 1) & ~1)) * sizeof(double) + (size_t)la0 * n * sizeof(z64);
  SDV_REQUIRE(sh0 <= 64 * 1024, SDV_ERR_UNSUPPORTED, "grid edge too large for the KE tile");
  hipEvent_t e0 = prof ? prof->begin(st) : nullptr;
  const z64* ij = (const z64*)a.invj;
  const z64* sd = (const z64*)a.sdiag;
  {
    const unsigned nb = (unsigned)((n3 / n + la0 - 1) / la0);
    hipLaunchKernelGGL((k_ke_axis<0>), dim3(nb, 1, a.howmany), dim3(kKeThreads), sh0, st, n, a.ke, ij, sd,
                       (const z64*)a.in, (z64*)a.out, la0);
  }
  {
    const size_t s = n;
    const unsigned nb = (unsigned)(((s + la0 - 1) / la0) * (n3 / (s * n)));
    hipLaunchKernelGGL((k_ke_axis<1>), dim3(nb, 1, a.howmany), dim3(kKeThreads), sh0, st, n, a.ke, ij, sd,
                       (const z64*)a.in, (z64*)a.out, la0);
  }
  {
    const size_t s = (size_t)n * n;
    const unsigned nb = (unsigned)(((s + la0 - 1) / la0) * (n3 / (s * n)));
    hipLaunchKernelGGL((k_ke_axis<2>), dim3(nb, 1, a.howmany), dim3(kKeThreads), sh0, st, n, a.ke, ij, sd,
                       (const z64*)a.in, (z64*)a.out, la0);
  }
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_KE, e0, st);
  return SDV_OK;
}

}  // namespace sdv
