// sdv_matel.hip -- two-electron matrix elements from pair potentials, FP64.
//
// Reference: call_twoe_matelxxx00, SD/coreproject.f90:1083-1100:
//   for (1a,1b): myden = inspfs10(:,1a) * CONJUGATE(inspfs20(:,1b))
//                twoematel(1a,1b,:,low:high) = ZGEMV('T', twoereduced, myden)
// i.e. twoematel(1a,1b,2a,2b) = sum_x rho_{1a1b}(x) V_{2a2b}(x), no conjugation
// of V.  Here the N^2 GEMVs become one real split-K GEMM over the grid:
//   G[2r+s][2c+t] = sum_x part_s(rho_r(x)) part_t(V_c(x)),  s,t in {re, im}
// followed by an assembly kernel that forms every complex element from the
// four real products, including the mirrored/conjugated pairs.  Reductions
// over the split-K partials run in a fixed order (results do not depend on
// scheduling).
#include "sdv_internal.h"

namespace sdv {

constexpr int kTM = 64, kTN = 64, kKC = 16;

__global__ void k_form_rho(const z64* __restrict__ phi1, const z64* __restrict__ phi2, const int2* __restrict__ rows,
                           int n3, int conj_b, z64* __restrict__ D) {
  const int r = blockIdx.y;
  const int2 ab = rows[r];
  const z64* pa = phi1 + (size_t)ab.x * n3;
  const z64* pb = phi2 + (size_t)ab.y * n3;
  z64* d = D + (size_t)r * n3;
  for (int x = blockIdx.x * blockDim.x + threadIdx.x; x < n3; x += gridDim.x * blockDim.x) {
    z64 w = pb[x];
    if (conj_b) w.im = -w.im;
    d[x] = pa[x] * w;
  }
}

__global__ void __launch_bounds__(256) k_gemm_tn(const z64* __restrict__ D, const z64* __restrict__ tr,
                                                  const int* __restrict__ cols, int nrows, int ncols, int n3,
                                                  int kchunk, double* __restrict__ part) {
  __shared__ double As[kKC][kTM + 2];
  __shared__ double Bs[kKC][kTN + 2];
  const int tid = threadIdx.x;
  const int tx = tid % 16, ty = tid / 16;
  const int R2 = 2 * nrows, C2 = 2 * ncols;
  const int r0 = blockIdx.y * (kTM / 2), c0 = blockIdx.x * (kTN / 2);
  const int x0 = blockIdx.z * kchunk;
  const int x1 = min(n3, x0 + kchunk);
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int xb = x0; xb < x1; xb += kKC) {
#pragma unroll
    for (int q = 0; q < 2; ++q) {
      const int e = tid + 256 * q;
      const int xx = e % kKC, rr = e / kKC;
      const int x = xb + xx;
      z64 va = mk(0.0, 0.0), vb = mk(0.0, 0.0);
      if (x < x1) {
        if (r0 + rr < nrows) va = D[(size_t)(r0 + rr) * n3 + x];
        if (c0 + rr < ncols) vb = tr[(size_t)cols[c0 + rr] * n3 + x];
      }
      As[xx][2 * rr] = va.re;
      As[xx][2 * rr + 1] = va.im;
      Bs[xx][2 * rr] = vb.re;
      Bs[xx][2 * rr + 1] = vb.im;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kKC; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) a[i] = As[kk][ty * 4 + i];
#pragma unroll
      for (int j = 0; j < 4; ++j) b[j] = Bs[kk][tx * 4 + j];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  double* p = part + (size_t)blockIdx.z * R2 * C2;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const int row = 2 * r0 + ty * 4 + i;
    if (row >= R2) continue;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int col = 2 * c0 + tx * 4 + j;
      if (col < C2) p[(size_t)row * C2 + col] = acc[i][j];
    }
  }
}

__global__ void k_reduce_parts(const double* __restrict__ part, int nsplit, long len, double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= len) return;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * len + i];
  out[i] = s;
}

static int ksplit_for(int tiles) {
  int ks = 2048 / (tiles > 0 ? tiles : 1);
  return ks < 1 ? 1 : (ks > 256 ? 256 : ks);
}

size_t matel_scratch_bytes(int n3, int nrows, int ncols) {
  const int tiles = ((2 * nrows + kTM - 1) / kTM) * ((2 * ncols + kTN - 1) / kTN);
  const int ks = ksplit_for(tiles);
  return (size_t)nrows * n3 * sizeof(z64) + (size_t)ks * (2 * nrows) * (2 * ncols) * sizeof(double) + 256;
}

int matel_launch(const MatelArgsHost& a, hipStream_t st, Prof* prof) {
  if (a.nrows <= 0 || a.ncols <= 0) return SDV_OK;
  const int tilesR = (2 * a.nrows + kTM - 1) / kTM, tilesC = (2 * a.ncols + kTN - 1) / kTN;
  const int ks = ksplit_for(tilesR * tilesC);
  int kchunk = (a.n3 + ks - 1) / ks;
  kchunk = ((kchunk + kKC - 1) / kKC) * kKC;
  const int ksplit = (a.n3 + kchunk - 1) / kchunk;
  z64* D = (z64*)a.scratch;
  double* part = (double*)((char*)a.scratch + (size_t)a.nrows * a.n3 * sizeof(z64));
  hipEvent_t e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL(k_form_rho, dim3(256, a.nrows), dim3(256), 0, st, (const z64*)a.phi1, (const z64*)a.phi2,
                     a.rows_dev, a.n3, a.conj_b, D);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_MATEL_RHO, e0, st);
  e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL(k_gemm_tn, dim3(tilesC, tilesR, ksplit), dim3(256), 0, st, D, (const z64*)a.tr, a.cols_dev,
                     a.nrows, a.ncols, a.n3, kchunk, part);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_MATEL_GEMM, e0, st);
  e0 = prof ? prof->begin(st) : nullptr;
  const long len = (long)(2 * a.nrows) * (2 * a.ncols);
  hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, part, ksplit, len, a.gout);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_MATEL_REDUCE, e0, st);
  return SDV_OK;
}

}  // namespace sdv
