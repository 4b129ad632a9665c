// sdv_meanfield.hip -- mean-field potential build and apply, FP64.
//
//   get_reducedpot0 (S/mean.f90:556-558):
//     reducedpot(x, c) = sum_k twoereduced(x, k) * twoden(k, c),  k, c in [0, N^2)
//     (a complex GEMM, on the FP64 matrix cores)
//   mult_reducedpot (SD/coreproject.f90:1200-1235):
//     out(x, i) = sum_k reducedpot(x, k, i) * in(x, k)
#include "sdv_internal.h"

namespace sdv {

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

// ---- the same product on the FP64 matrix cores
// reducedpot (M x K) . twoden (K x K), M = n^3, K = N^2, as four real
// products per complex one:  Re += Re A Re B - Im A Im B,  Im += Re A Im B +
// Im A Re B, on v_mfma_f64_16x16x4f64 (A operand: lane l holds row l % 16,
// k l / 16; B operand: k l / 16, column l % 16; accumulator element v of lane
// l is row 4 v + l / 16, column l % 16, as checked on gfx950).  A workgroup
// of eight waves computes 64 rows x 256 columns (wave: 32 rows x 64
// columns, 2 x 4 tiles of 16 x 16, 16 accumulators); the operands come
// straight from global memory: A's 16 rows of one k are 256 contiguous bytes,
// and twoden is first transposed (k-major, 1 MB at 16 orbitals, L2 resident)
// so that B's 16 columns of one k are too.  8 n^3 N^4 real flops
// (0.28 TFLOP at 81^3 with 16 orbitals), compute-bound on the matrix cores.
typedef double f64x4 __attribute__((ext_vector_type(4)));
constexpr int kRmRows = 64, kRmCols = 256, kRmThreads = 512;

__global__ void k_transpose_z64(int K, const z64* __restrict__ g, z64* __restrict__ gt) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;  // gt index: k * K + c
  if (i >= (long long)K * K) return;
  const int k = (int)(i / K), c = (int)(i - (long long)k * K);
  gt[i] = g[(size_t)k + (size_t)K * c];
}

#ifndef SDV_RP_3M
#define SDV_RP_3M 1
#endif
__global__ void __launch_bounds__(kRmThreads) k_reducedpot_mfma(int M, int K, const z64* __restrict__ A,
                                                               const z64* __restrict__ Bt, z64* __restrict__ C) {
  const int w = threadIdx.x >> 6, l = threadIdx.x & 63;
  const int lr = l & 15, lk = l >> 4;
  const long long x0 = (long long)blockIdx.x * kRmRows + 32 * (w & 1);
  const int c0 = blockIdx.y * kRmCols + 64 * (w >> 1);
  // Three real products per complex one (Gauss): P1 = Re A Re B, P2 = Im A Im B,
  // P3 = (Re A + Im A)(Re B + Im B); Re C = P1 - P2, Im C = P3 - P1 - P2.
  // With SDV_RP_3M=0, four: Re C = Re A Re B - Im A Im B, Im C = Re A Im B +
  // Im A Re B.
  constexpr int NP = SDV_RP_3M ? 3 : 2;
  f64x4 acc[NP][2][4];
#pragma unroll
  for (int p = 0; p < NP; ++p)
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) acc[p][rt][ct] = f64x4{0, 0, 0, 0};
#pragma unroll 2
  for (int k0 = 0; k0 < K; k0 += 4) {
    const int k = k0 + lk;
    z64 a[2], b[4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
      const long long x = x0 + 16 * rt + lr;
      a[rt] = (x < M && k < K) ? A[x + (long long)M * k] : mk(0.0, 0.0);
    }
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      const int c = c0 + 16 * ct + lr;
      b[ct] = (c < K && k < K) ? Bt[(long long)k * K + c] : mk(0.0, 0.0);
    }
    if (SDV_RP_3M) {
      double as[2], bs[4];
#pragma unroll
      for (int rt = 0; rt < 2; ++rt) as[rt] = a[rt].re + a[rt].im;
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) bs[ct] = b[ct].re + b[ct].im;
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
          acc[0][rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt].re, b[ct].re, acc[0][rt][ct], 0, 0, 0);
          acc[1][rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt].im, b[ct].im, acc[1][rt][ct], 0, 0, 0);
          acc[NP - 1][rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(as[rt], bs[ct], acc[NP - 1][rt][ct], 0, 0, 0);
        }
    } else {
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
          acc[0][rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt].re, b[ct].re, acc[0][rt][ct], 0, 0, 0);
          acc[0][rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(-a[rt].im, b[ct].im, acc[0][rt][ct], 0, 0, 0);
          acc[1][rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt].re, b[ct].im, acc[1][rt][ct], 0, 0, 0);
          acc[1][rt][ct] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[rt].im, b[ct].re, acc[1][rt][ct], 0, 0, 0);
        }
    }
  }
#pragma unroll
  for (int rt = 0; rt < 2; ++rt)
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      const int c = c0 + 16 * ct + lr;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const long long x = x0 + 16 * rt + 4 * v + lk;
        const double p1 = acc[0][rt][ct][v], p2 = acc[1][rt][ct][v];
        const z64 r = SDV_RP_3M ? mk(p1 - p2, acc[NP - 1][rt][ct][v] - p1 - p2) : mk(p1, p2);
        if (x < M && c < K) C[x + (long long)M * c] = r;
      }
    }
}

size_t reducedpot_scratch_bytes(int N) { return (size_t)N * N * N * N * sizeof(z64); }

int reducedpot_launch(int n3, int N, const void* tr, const void* twoden, void* out, void* scratch, hipStream_t st) {
  const int NN = N * N;
  const long long nt = (long long)NN * NN;
  hipLaunchKernelGGL(k_transpose_z64, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, NN,
                     (const z64*)twoden, (z64*)scratch);
  SDV_HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_reducedpot_mfma, dim3((unsigned)((n3 + kRmRows - 1) / kRmRows), (NN + kRmCols - 1) / kRmCols),
                     dim3(kRmThreads), 0, st, n3, NN, (const z64*)tr, (const z64*)scratch, (z64*)out);
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
