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
//
// The GEMM runs on FP64 MFMA (v_mfma_f64_16x16x4f64); pair densities are
// formed from the orbitals while the A tile is staged, never stored.  When the
// row and column pair sets coincide (identical bra/ket orbitals, full
// lowspf..highspf range) G is symmetric: sum rho_P V_Q = sum rho_Q V_P for the
// real even Coulomb kernel, which makes the real product matrix symmetric, so
// only the upper-triangle tiles are computed.
#include "sdv_internal.h"

namespace sdv {

#ifndef SDV_MATEL_KKC
#define SDV_MATEL_KKC 16
#endif
constexpr int kTM = 64, kTN = 64, kKC = SDV_MATEL_KKC;
constexpr int kEPT = kKC * (kTM / 2) / 256;  // staged (grid point, pair) elements per thread and stage

typedef double d4 __attribute__((ext_vector_type(4)));

// 64 x 64 real output tile per workgroup (4 waves, 32 x 32 each = 2 x 2
// MFMA 16x16 tiles), kKC grid points per LDS stage.  MFMA f64 16x16x4
// operands: lane l holds A[l%16][l/16] and B[l/16][l%16]; the accumulator
// holds D[4*v + l/16][l%16] in element v (checked on gfx950 with a probe).
__global__ void __launch_bounds__(256) k_gemm_mfma(const z64* __restrict__ phi1, const z64* __restrict__ phi2,
                                                    const int2* __restrict__ rows, int conj_b,
                                                    const z64* __restrict__ tr, const int* __restrict__ cols,
                                                    int nrows, int ncols, int n3, int kchunk, int sym, int tiles_c,
                                                    double* __restrict__ part) {
  __shared__ double As[kKC][kTM + 2];
  __shared__ double Bs[kKC][kTN + 2];
  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wr = wave >> 1, wc = wave & 1;
  int by, bx;
  if (sym) {  // upper-triangle tile t -> (by, bx), bx >= by
    int t = blockIdx.x;
    by = 0;
    while (t >= tiles_c - by) {
      t -= tiles_c - by;
      ++by;
    }
    bx = by + t;
  } else {
    bx = blockIdx.x % tiles_c;
    by = blockIdx.x / tiles_c;
  }
  const int R2 = 2 * nrows, C2 = 2 * ncols;
  const int r0 = by * (kTM / 2), c0 = bx * (kTN / 2);
  const int x0 = blockIdx.z * kchunk;
  const int x1 = min(n3, x0 + kchunk);
  d4 acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  // each thread stages two (grid point, pair) elements of A and of B per stage
  int2 rab[kEPT];
  const z64* bcol[kEPT];
#pragma unroll
  for (int q = 0; q < kEPT; ++q) {
    const int rr = (tid + 256 * q) / kKC;
    rab[q] = (r0 + rr < nrows) ? rows[r0 + rr] : int2{-1, -1};
    bcol[q] = (c0 + rr < ncols) ? tr + (size_t)cols[c0 + rr] * n3 : nullptr;
  }
  // The next stage's orbitals and potentials are loaded into registers while
  // the current stage's MFMAs run (register prefetch, one LDS buffer).
  z64 fa[kEPT], fw[kEPT], fb[kEPT];
  auto fetch = [&](int xb) {
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
      const int e = tid + 256 * q;
      const int x = xb + e % kKC;
      fa[q] = fw[q] = fb[q] = mk(0.0, 0.0);
      if (x < x1) {
        if (rab[q].x >= 0) {
          fw[q] = phi2[(size_t)rab[q].y * n3 + x];
          fa[q] = phi1[(size_t)rab[q].x * n3 + x];
        }
        if (bcol[q]) fb[q] = bcol[q][x];
      }
    }
  };
  fetch(x0);
  for (int xb = x0; xb < x1; xb += kKC) {
#pragma unroll
    for (int q = 0; q < kEPT; ++q) {
      const int e = tid + 256 * q;
      const int xx = e % kKC, rr = e / kKC;
      z64 w = fw[q];  // rho = phi1(:,a) * CONJUGATE(phi2(:,b))   SD/coreproject.f90:1094
      if (conj_b) w.im = -w.im;
      const z64 va = fa[q] * w, vb = fb[q];
      As[xx][2 * rr] = va.re;
      As[xx][2 * rr + 1] = va.im;
      Bs[xx][2 * rr] = vb.re;
      Bs[xx][2 * rr + 1] = vb.im;
    }
    __syncthreads();
    if (xb + kKC < x1) fetch(xb + kKC);
#pragma unroll
    for (int kk = 0; kk < kKC; kk += 4) {
      double a[2], b[2];
#pragma unroll
      for (int i = 0; i < 2; ++i) a[i] = As[kk + (lane >> 4)][wr * 32 + i * 16 + (lane & 15)];
#pragma unroll
      for (int j = 0; j < 2; ++j) b[j] = Bs[kk + (lane >> 4)][wc * 32 + j * 16 + (lane & 15)];
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  double* p = part + (size_t)blockIdx.z * R2 * C2;
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = 2 * r0 + wr * 32 + i * 16 + 4 * v + (lane >> 4);
        const int col = 2 * c0 + wc * 32 + j * 16 + (lane & 15);
        if (row < R2 && col < C2) p[(size_t)row * C2 + col] = acc[i][j][v];
      }
}

__global__ void k_reduce_parts(const double* __restrict__ part, int nsplit, int R2, int C2, int sym,
                               double* __restrict__ out) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long len = (long)R2 * C2;
  if (i >= len) return;
  const int row = (int)(i / C2), col = (int)(i % C2);
  // with the symmetric tiling only tiles with col-tile >= row-tile exist
  const long src = (sym && col / kTN < row / kTM) ? (long)col * C2 + row : i;
  double s = 0.0;
  for (int k = 0; k < nsplit; ++k) s += part[(size_t)k * len + src];
  out[i] = s;
}

static int ksplit_for(int tiles) {
  int ks = 2048 / (tiles > 0 ? tiles : 1);
  return ks < 1 ? 1 : (ks > 256 ? 256 : ks);
}

// Tiling of one reduction: output tiles (upper triangle only when symmetric)
// and the split of the grid-point sum.  Shared by the scratch sizing and the
// launch so the two can never disagree.
struct MatelTiling {
  int tilesR, tilesC, ntiles, kchunk, ksplit;
  bool sym;
};
static MatelTiling matel_tiling(int n3, int nrows, int ncols, bool symmetric) {
  MatelTiling t;
  t.tilesR = (2 * nrows + kTM - 1) / kTM;
  t.tilesC = (2 * ncols + kTN - 1) / kTN;
  t.sym = symmetric && nrows == ncols;
  t.ntiles = t.sym ? t.tilesC * (t.tilesC + 1) / 2 : t.tilesR * t.tilesC;
  const int ks = ksplit_for(t.ntiles);
  t.kchunk = (n3 + ks - 1) / ks;
  t.kchunk = ((t.kchunk + kKC - 1) / kKC) * kKC;
  t.ksplit = (n3 + t.kchunk - 1) / t.kchunk;
  return t;
}

size_t matel_scratch_bytes(int n3, int nrows, int ncols, int symmetric) {
  const MatelTiling t = matel_tiling(n3, nrows, ncols, symmetric != 0);
  return (size_t)t.ksplit * (2 * nrows) * (2 * ncols) * sizeof(double) + 256;
}

int matel_launch(const MatelArgsHost& a, hipStream_t st, Prof* prof) {
  if (a.nrows <= 0 || a.ncols <= 0) return SDV_OK;
  const MatelTiling t = matel_tiling(a.n3, a.nrows, a.ncols, a.symmetric != 0);
  const bool sym = t.sym;
  const int tilesC = t.tilesC, ntiles = t.ntiles, kchunk = t.kchunk, ksplit = t.ksplit;
  SDV_REQUIRE(a.scratch_bytes >= matel_scratch_bytes(a.n3, a.nrows, a.ncols, a.symmetric), SDV_ERR_ARG,
              "matrix-element scratch buffer too small");
  double* part = (double*)a.scratch;
  hipEvent_t e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL(k_gemm_mfma, dim3(ntiles, 1, ksplit), dim3(256), 0, st, (const z64*)a.phi1, (const z64*)a.phi2,
                     a.rows_dev, a.conj_b, (const z64*)a.tr, a.cols_dev, a.nrows, a.ncols, a.n3, kchunk,
                     sym ? 1 : 0, tilesC, part);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_MATEL_GEMM, e0, st);
  e0 = prof ? prof->begin(st) : nullptr;
  const long len = (long)(2 * a.nrows) * (2 * a.ncols);
  hipLaunchKernelGGL(k_reduce_parts, dim3((unsigned)((len + 255) / 256)), dim3(256), 0, st, part, ksplit,
                     2 * a.nrows, 2 * a.ncols, sym ? 1 : 0, a.gout);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_MATEL_REDUCE, e0, st);
  return SDV_OK;
}

}  // namespace sdv
