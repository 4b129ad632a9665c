// sdv_onebody.hip -- one-electron matrix elements over the grid, FP64.
//
//   sparseops_matel (S/matel.f90:241-293):  opmatel(a, b) = sum_x f(in1(x, a)) * (T in2)(x, b)
//   pot_matel       (S/matel.f90:113-169):  potmatel(a, b) = sum_x f(in1(x, a)) * pot(x) * in2(x, b)
// with f = conjugation for hermnorm (MYGEMM with CNORMCHAR 'C') and the
// identity for cnorm ('T'), T the kinetic apply (mult_ke).  The grid sum is
// split over workgroups (x chunks); partial N_a x N_b blocks are summed in a
// fixed order by a second kernel, so results are deterministic.
#include "sdv_internal.h"

namespace sdv {

constexpr int kDotChunk = 8192;  // grid points per workgroup
constexpr int kDotTile = 64;     // grid points staged per step

__global__ void __launch_bounds__(256) k_dot_partial(int n3, int na, int nb, const z64* __restrict__ A,
                                                      const z64* __restrict__ B, const double* __restrict__ w,
                                                      int conj_a, z64* __restrict__ part) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  z64* As = reinterpret_cast<z64*>(smem_raw);  // [na][kDotTile]
  z64* Bs = As + (size_t)na * kDotTile;         // [nb][kDotTile]
  const int x0 = blockIdx.x * kDotChunk;
  const int x1 = min(n3, x0 + kDotChunk);
  const int npair = na * nb;
  // each thread owns up to 4 (a, b) entries
  z64 acc[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = mk(0.0, 0.0);
  for (int t0 = x0; t0 < x1; t0 += kDotTile) {
    const int nt = min(kDotTile, x1 - t0);
    __syncthreads();
    for (int e = threadIdx.x; e < na * kDotTile; e += 256) {
      const int a = e / kDotTile, i = e - a * kDotTile;
      z64 v = mk(0.0, 0.0);
      if (i < nt) {
        v = A[(size_t)a * n3 + t0 + i];
        if (conj_a) v.im = -v.im;
        if (w) v = scale(v, w[t0 + i]);
      }
      As[e] = v;
    }
    for (int e = threadIdx.x; e < nb * kDotTile; e += 256) {
      const int b = e / kDotTile, i = e - b * kDotTile;
      Bs[e] = i < nt ? B[(size_t)b * n3 + t0 + i] : mk(0.0, 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int pr = threadIdx.x + 256 * j;
      if (pr < npair) {
        const int a = pr % na, b = pr / na;
        const z64* ar = As + (size_t)a * kDotTile;
        const z64* br = Bs + (size_t)b * kDotTile;
        z64 s = acc[j];
        for (int i = 0; i < nt; ++i) {
          s.re = fma(ar[i].re, br[i].re, fma(-ar[i].im, br[i].im, s.re));
          s.im = fma(ar[i].re, br[i].im, fma(ar[i].im, br[i].re, s.im));
        }
        acc[j] = s;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int pr = threadIdx.x + 256 * j;
    if (pr < npair) part[(size_t)blockIdx.x * npair + pr] = acc[j];
  }
}

// out(a + ld * b) = sum over chunks, in chunk order
__global__ void k_dot_reduce(int nchunk, int na, int nb, int ld, const z64* __restrict__ part, z64* __restrict__ out) {
  const int pr = blockIdx.x * blockDim.x + threadIdx.x;
  if (pr >= na * nb) return;
  z64 s = mk(0.0, 0.0);
  for (int c = 0; c < nchunk; ++c) s = s + part[(size_t)c * na * nb + pr];
  const int a = pr % na, b = pr / na;
  out[(size_t)a + (size_t)ld * b] = s;
}

size_t dot_scratch_bytes(int n3, int na, int nb) {
  const size_t nchunk = ((size_t)n3 + kDotChunk - 1) / kDotChunk;
  return nchunk * (size_t)na * nb * sizeof(z64);
}

int dot_launch(int n3, int na, int nb, const void* A, const void* B, const double* w, int conj_a, void* scratch,
               void* out, int ld, hipStream_t st) {
  SDV_REQUIRE(na >= 1 && nb >= 1 && na * nb <= 1024, SDV_ERR_UNSUPPORTED, "at most 1024 matrix elements per call");
  const int nchunk = (n3 + kDotChunk - 1) / kDotChunk;
  const size_t sh = (size_t)(na + nb) * kDotTile * sizeof(z64);
  SDV_REQUIRE(sh <= 160 * 1024, SDV_ERR_UNSUPPORTED, "too many orbitals for the staged dot");
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_dot_partial, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  hipLaunchKernelGGL(k_dot_partial, dim3(nchunk), dim3(256), sh, st, n3, na, nb, (const z64*)A, (const z64*)B, w,
                     conj_a, (z64*)scratch);
  hipLaunchKernelGGL(k_dot_reduce, dim3((na * nb + 255) / 256), dim3(256), 0, st, nchunk, na, nb, ld,
                     (const z64*)scratch, (z64*)out);
  SDV_HIP_CHECK(hipGetLastError());
  return SDV_OK;
}

}  // namespace sdv

namespace sdv {

// ---------------------------------------------------------- orbital operator
// out(x, j) = alpha * sum_i A(x, i) C(i + ldc j) + beta * out(x, j) for j in
// [0, nout), i in [0, nin); C is a small complex matrix (column-major, ldc).
// One thread per grid point and 16 output columns; C's tile in LDS.
constexpr int kOgCols = 16;

__global__ void __launch_bounds__(256) k_orb_gemm(int n3, int nin, int nout, const z64* __restrict__ A,
                                                   const z64* __restrict__ C, int ldc, z64 alpha, z64 beta,
                                                   z64* __restrict__ out) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  z64* cs = reinterpret_cast<z64*>(smem_raw);  // [nin][kOgCols]
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int j0 = blockIdx.y * kOgCols;
  const int nj = min(kOgCols, nout - j0);
  for (int e = threadIdx.x; e < nin * kOgCols; e += 256) {
    const int i = e / kOgCols, j = e - i * kOgCols;
    cs[e] = j < nj ? C[(size_t)i + (size_t)ldc * (j0 + j)] : mk(0.0, 0.0);
  }
  __syncthreads();
  if (x >= n3) return;
  z64 acc[kOgCols];
#pragma unroll
  for (int j = 0; j < kOgCols; ++j) acc[j] = mk(0.0, 0.0);
  for (int i = 0; i < nin; ++i) {
    const z64 a = A[(size_t)x + (size_t)n3 * i];
#pragma unroll
    for (int j = 0; j < kOgCols; ++j) {
      const z64 c = cs[i * kOgCols + j];
      acc[j].re = fma(a.re, c.re, fma(-a.im, c.im, acc[j].re));
      acc[j].im = fma(a.re, c.im, fma(a.im, c.re, acc[j].im));
    }
  }
#pragma unroll
  for (int j = 0; j < kOgCols; ++j) {
    if (j < nj) {
      z64* o = out + (size_t)x + (size_t)n3 * (j0 + j);
      z64 r = alpha * acc[j];
      if (beta.re != 0.0 || beta.im != 0.0) r = r + beta * (*o);
      *o = r;
    }
  }
}

int orb_gemm_launch(int n3, int nin, int nout, const void* A, const void* C, int ldc, z64 alpha, z64 beta, void* out,
                    hipStream_t st) {
  if (nout <= 0) return SDV_OK;
  SDV_REQUIRE(nin >= 1 && nin <= 512, SDV_ERR_UNSUPPORTED, "orbital count outside 1..512");
  const size_t sh = (size_t)nin * kOgCols * sizeof(z64);
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_orb_gemm, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  hipLaunchKernelGGL(k_orb_gemm, dim3((n3 + 255) / 256, (nout + kOgCols - 1) / kOgCols), dim3(256), sh, st, n3, nin,
                     nout, (const z64*)A, (const z64*)C, ldc, alpha, beta, (z64*)out);
  SDV_HIP_CHECK(hipGetLastError());
  return SDV_OK;
}

// s(x, j) = ((ke + pot w) + pulse) + rp, actreduced00's spfmult sum
// (S/derivs.f90:691-764); pulse = length-gauge lenmultiply
// (S/proputils.f90:157-182): ((w dz(z)) tz + (w dy(y)) ty) + (w dx(x)) tx
// with the per-axis dipoles dip (the grid coordinate, complex-scaled under ECS);
// an axis whose field value is zero is skipped (bit 0 x, 1 y, 2 z of axes).
__global__ void k_orb_sum(size_t total, int n, const z64* __restrict__ ke, const double* __restrict__ pot,
                          const z64* __restrict__ w, const z64* __restrict__ rp, const z64* __restrict__ dip, z64 tx,
                          z64 ty, z64 tz, int axes, z64* __restrict__ s) {
  const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const size_t n3 = (size_t)n * n * n, g = e % n3;
  z64 v = ke[e];
  if (pot) v = v + scale(w[e], pot[g]);
  if (axes) {
    const z64 we = w[e];
    z64 p = mk(0.0, 0.0);
    if (axes & 4) p = p + (we * dip[g / ((size_t)n * n)]) * tz;
    if (axes & 2) p = p + (we * dip[(g / n) % n]) * ty;
    if (axes & 1) p = p + (we * dip[g % n]) * tx;
    v = v + p;
  }
  s[e] = v + rp[e];
}

int orb_sum_launch(int n, int ncols, const void* ke, const double* pot, const void* w, const void* rp,
                   const void* dip, const double* tdpot, void* s, hipStream_t st) {
  const size_t total = (size_t)n * n * n * ncols;
  if (!total) return SDV_OK;
  const z64 zero = mk(0.0, 0.0);
  const z64 tx = tdpot ? mk(tdpot[0], tdpot[1]) : zero, ty = tdpot ? mk(tdpot[2], tdpot[3]) : zero,
            tz = tdpot ? mk(tdpot[4], tdpot[5]) : zero;
  auto on = [](z64 t) { return (t.re != 0.0 || t.im != 0.0) ? 1 : 0; };
  const int axes = on(tx) | (on(ty) << 1) | (on(tz) << 2);
  hipLaunchKernelGGL(k_orb_sum, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, total, n, (const z64*)ke,
                     pot, (const z64*)w, (const z64*)rp, (const z64*)dip, tx, ty, tz, axes, (z64*)s);
  SDV_HIP_CHECK(hipGetLastError());
  return SDV_OK;
}

}  // namespace sdv
