// sdv_meanfield.hip -- mean-field potential build and apply, FP64.
//
//   get_reducedpot0 (S/mean.f90:556-558):
//     reducedpot(x, c) = sum_k twoereduced(x, k) * twoden(k, c),  k, c in [0, N^2)
//   mult_reducedpot (SD/coreproject.f90:1200-1235):
//     out(x, i) = sum_k reducedpot(x, k, i) * in(x, k)
#include "sdv_internal.h"

namespace sdv {

constexpr int kRpCols = 16;   // output columns per workgroup
constexpr int kRpKc = 64;     // twoden rows staged per step

// Each thread owns one grid point x and kRpCols output columns; the
// twoden tile is staged in LDS and read as broadcasts.
__global__ void __launch_bounds__(256) k_reducedpot(int n3, int NN, const z64* __restrict__ tr,
                                                     const z64* __restrict__ g, z64* __restrict__ out) {
  __shared__ z64 gs[kRpKc][kRpCols];
  const int x = blockIdx.x * 256 + threadIdx.x;
  const int c0 = blockIdx.y * kRpCols;
  z64 acc[kRpCols];
#pragma unroll
  for (int j = 0; j < kRpCols; ++j) acc[j] = mk(0.0, 0.0);
  for (int k0 = 0; k0 < NN; k0 += kRpKc) {
    __syncthreads();
    for (int e = threadIdx.x; e < kRpKc * kRpCols; e += 256) {
      const int kk = e / kRpCols, j = e - kk * kRpCols;
      const int k = k0 + kk, c = c0 + j;
      gs[kk][j] = (k < NN && c < NN) ? g[(size_t)k + (size_t)NN * c] : mk(0.0, 0.0);
    }
    __syncthreads();
    if (x < n3) {
      const int kmax = min(kRpKc, NN - k0);
      for (int kk = 0; kk < kmax; ++kk) {
        const z64 t = tr[(size_t)x + (size_t)n3 * (k0 + kk)];
#pragma unroll
        for (int j = 0; j < kRpCols; ++j) {
          const z64 gv = gs[kk][j];
          acc[j].re = fma(t.re, gv.re, fma(-t.im, gv.im, acc[j].re));
          acc[j].im = fma(t.re, gv.im, fma(t.im, gv.re, acc[j].im));
        }
      }
    }
  }
  if (x < n3) {
#pragma unroll
    for (int j = 0; j < kRpCols; ++j)
      if (c0 + j < NN) out[(size_t)x + (size_t)n3 * (c0 + j)] = acc[j];
  }
}

__global__ void k_mult_reducedpot(int n3, int N, const z64* __restrict__ in, const z64* __restrict__ rp,
                                  z64* __restrict__ out) {
  const int x = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;  // output orbital (relative to firstspf)
  if (x >= n3) return;
  z64 acc = mk(0.0, 0.0);
  for (int k = 0; k < N; ++k) {
    const z64 r = rp[(size_t)x + (size_t)n3 * (k + (size_t)N * i)];
    const z64 v = in[(size_t)x + (size_t)n3 * k];
    acc = acc + r * v;
  }
  out[(size_t)x + (size_t)n3 * i] = acc;
}

int reducedpot_launch(int n3, int N, const void* tr, const void* twoden, void* out, hipStream_t st) {
  const int NN = N * N;
  hipLaunchKernelGGL(k_reducedpot, dim3((n3 + 255) / 256, (NN + kRpCols - 1) / kRpCols), dim3(256), 0, st, n3, NN,
                     (const z64*)tr, (const z64*)twoden, (z64*)out);
  SDV_HIP_CHECK(hipGetLastError());
  return SDV_OK;
}

int mult_reducedpot_launch(int n3, int N, int nout, const void* inspfs, const void* rp, void* out, hipStream_t st) {
  if (nout <= 0) return SDV_OK;
  hipLaunchKernelGGL(k_mult_reducedpot, dim3((n3 + 255) / 256, nout), dim3(256), 0, st, n3, N,
                     (const z64*)inspfs, (const z64*)rp, (z64*)out);
  SDV_HIP_CHECK(hipGetLastError());
  return SDV_OK;
}

}  // namespace sdv
