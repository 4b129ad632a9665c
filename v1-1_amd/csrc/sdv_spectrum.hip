// sdv_spectrum.hip -- device construction of the Toeplitz 1/r Coulomb kernel
// and of its circulant spectrum, in FP64.
//
// Restates GETINVERSE / RVECTOR / TOOTHEIGMULT (SD/tooth.f90:12-231):
//   R(i,j,k)  = D^3 / |(iD, jD, kD)|,  R(0) = 0                    (:135-146)
//   TR        = -sum_p [T(i-p) R(p,j,k) + T(j-p) R(i,p,k) + T(k-p) R(i,j,p)]
//               on the small box |i|,|j|,|k| <= nsmall, p in [-nbig, nbig]
//               (:158-166, with the intended slice -NBIG-J:NBIG-J at :165)
//   TR(0)    += 2 pi                                                 (:44)
//   T_small   = 1-D sinc Laplacian on 2 nsmall + 1 points            (:48-53)
//   Q         = W^(x3) [ (W^T)^(x3) TR / (E_i + E_j + E_k) ]         (:63-79)
//   TINV      = R + Q on the small box, R elsewhere                  (:83-89)
// then threed_two = TINV / D^3 (SD/sincproject.f90:155-156).
//
// The eigenpairs of T_small come from a parallel (round-robin) Jacobi run in
// one workgroup (the reference calls LAPACK DSYEV, :64); Q is the unique
// solution of the linear system, so any accurate eigen-decomposition gives it.
//
// The circulant holds K at offsets |o| <= n-1 on each axis (real, even), so
// its DFT is real and even:  Khat(k) = sum_{i,j,l} w_i w_j w_l K(i,j,l)
// cos(2pi k_x i/M) cos(2pi k_y j/M) cos(2pi k_z l/M) with w_0 = 1, w_i = 2.
// It is evaluated as three separable cosine contractions on the octant and
// stored for k in [0, m]^3 with the 1/M^3 of SD/toeplitz_subs.f90:578.
#include <algorithm>
#include <cmath>

#include "sdv_internal.h"

namespace sdv {

__device__ __forceinline__ double rvec(int i, int j, int k, double delta) {
  if (i == 0 && j == 0 && k == 0) return 0.0;
  const double pi_ = i * delta, pj = j * delta, pk = k * delta;
  return (delta * delta * delta) / sqrt(pi_ * pi_ + pj * pj + pk * pk);
}

__device__ __forceinline__ double alltt(int m, double delta) {
  if (m == 0) return (kPiD * kPiD) / (6.0 * delta * delta);
  const double sgn = (m & 1) ? -1.0 : 1.0;
  return sgn * (1.0 / (delta * delta)) / ((double)m * (double)m);
}

__global__ void k_tr(double* TR, int nsmall, int nbig, double delta) {
  const int NS = 2 * nsmall + 1;
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)NS * NS * NS) return;
  const int i = (int)(idx % NS) - nsmall;
  const int j = (int)((idx / NS) % NS) - nsmall;
  const int k = (int)(idx / ((long)NS * NS)) - nsmall;
  double s = 0.0;
  for (int p = -nbig; p <= nbig; ++p)
    s += alltt(p - i, delta) * rvec(p, j, k, delta) + alltt(p - j, delta) * rvec(i, p, k, delta) +
         alltt(p - k, delta) * rvec(i, j, p, delta);
  double v = -s;
  if (i == 0 && j == 0 && k == 0) v += 2.0 * kPiD;
  TR[idx] = v;
}

// One-workgroup parallel Jacobi on the NS x NS sinc Laplacian (column-major),
// in the round-robin ordering: NS (+1 when odd) indices play a round-robin
// tournament, each of the NS - 1 (or NS) rounds pairs every index once, and
// the rotations of one round act on disjoint row/column pairs, so their
// angles are computed together from the matrix at the start of the round and
// applied together (columns, then rows).  A sweep is one full tournament.
// Outputs eigenvectors W (columns) and eigenvalues E.
constexpr int kJacThreads = 1024;
__global__ void __launch_bounds__(kJacThreads) k_jacobi(int nsmall, double delta, double* Wout, double* Eout) {
  const int NS = 2 * nsmall + 1;
  const int NP = NS + (NS & 1);  // players (one idle index when NS is odd)
  const int H = NP / 2;          // pairs per round
  extern __shared__ double sh[];
  double* A = sh;             // NS*NS
  double* W = sh + NS * NS;   // NS*NS
  double* red = W + NS * NS;  // kJacThreads
  double* cs = red + kJacThreads;               // 2 per pair
  int* pq = reinterpret_cast<int*>(cs + 2 * H);  // 2 per pair
  const int tid = threadIdx.x;
  const double mycon = (kPiD / delta) * (kPiD / delta) / 6.0;
  for (int e = tid; e < NS * NS; e += kJacThreads) {
    const int i = e % NS, j = e / NS;
    double t;
    if (i == j)
      t = mycon;
    else {
      const int d = i - j;
      t = ((d & 1) ? -1.0 : 1.0) / (delta * delta * (double)d * (double)d);
    }
    A[e] = t;
    W[e] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  auto reduce = [&](double v) -> double {
    red[tid] = v;
    __syncthreads();
    for (int s = kJacThreads / 2; s > 0; s >>= 1) {
      if (tid < s) red[tid] += red[tid + s];
      __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
  };
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, tot = 0.0;
    for (int e = tid; e < NS * NS; e += kJacThreads) {
      const double v = A[e] * A[e];
      tot += v;
      if (e % NS != e / NS) off += v;
    }
    // converged when the off-diagonal norm is 1e-14 of the whole: rounding
    // keeps it near 1e-15.5 after that, so further sweeps only churn
    const double offs = reduce(off), tots = reduce(tot);
    if (offs <= 1e-28 * tots) break;
    for (int r = 0; r < NP - 1; ++r) {
      // circle method: index NP-1 fixed, the others rotate; round r pairs
      // (NP-1, r) and ((r + k) mod (NP-1), (r - k) mod (NP-1)), k = 1 .. H-1
      if (tid < H) {
        int p, q;
        if (tid == 0) {
          p = r;
          q = NP - 1;
        } else {
          p = (r + tid) % (NP - 1);
          q = (r - tid + (NP - 1)) % (NP - 1);
        }
        if (p > q) {
          const int t = p;
          p = q;
          q = t;
        }
        double c = 1.0, s = 0.0;
        if (q < NS) {
          const double apq = A[p + NS * q];
          if (fabs(apq) >= 1e-300) {
            const double tau = (A[q + NS * q] - A[p + NS * p]) / (2.0 * apq);
            const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        cs[2 * tid] = c;
        cs[2 * tid + 1] = s;
        pq[2 * tid] = p;
        pq[2 * tid + 1] = q;
      }
      __syncthreads();
      for (int e = tid; e < H * NS; e += kJacThreads) {  // columns p, q of A and W
        const int h = e / NS, k = e - h * NS;
        const double c = cs[2 * h], s = cs[2 * h + 1];
        if (s == 0.0) continue;
        const int p = pq[2 * h], q = pq[2 * h + 1];
        const double akp = A[k + NS * p], akq = A[k + NS * q];
        A[k + NS * p] = c * akp - s * akq;
        A[k + NS * q] = s * akp + c * akq;
        const double wkp = W[k + NS * p], wkq = W[k + NS * q];
        W[k + NS * p] = c * wkp - s * wkq;
        W[k + NS * q] = s * wkp + c * wkq;
      }
      __syncthreads();
      for (int e = tid; e < H * NS; e += kJacThreads) {  // rows p, q of A
        const int h = e / NS, k = e - h * NS;
        const double c = cs[2 * h], s = cs[2 * h + 1];
        if (s == 0.0) continue;
        const int p = pq[2 * h], q = pq[2 * h + 1];
        const double apk = A[p + NS * k], aqk = A[q + NS * k];
        A[p + NS * k] = c * apk - s * aqk;
        A[q + NS * k] = s * apk + c * aqk;
      }
      __syncthreads();
    }
  }
  for (int e = tid; e < NS * NS; e += kJacThreads) Wout[e] = W[e];
  for (int i = tid; i < NS; i += kJacThreads) Eout[i] = A[i + NS * i];
}

// Y(.., a, ..) = sum_l Wm(a, l) X(.., l, ..) along one axis of an NS^3 cube;
// trans = 1 uses Wm^T.
__global__ void k_axis_mat(const double* X, double* Y, const double* Wm, int NS, int axis, int trans) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long n3 = (long)NS * NS * NS;
  if (idx >= n3) return;
  int c[3] = {(int)(idx % NS), (int)((idx / NS) % NS), (int)(idx / ((long)NS * NS))};
  const long stride = axis == 0 ? 1 : (axis == 1 ? NS : (long)NS * NS);
  const int a = c[axis];
  const long base = idx - (long)a * stride;
  double s = 0.0;
  for (int l = 0; l < NS; ++l) {
    const double w = trans ? Wm[l + NS * a] : Wm[a + NS * l];
    s += w * X[base + (long)l * stride];
  }
  Y[idx] = s;
}

__global__ void k_eig_divide(double* X, const double* E, int NS) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)NS * NS * NS) return;
  const int i = idx % NS, j = (idx / NS) % NS, k = idx / ((long)NS * NS);
  X[idx] = X[idx] / (E[i] + E[j] + E[k]);
}

// K on the first octant (i,j,k in [0,n)), divided by D^3.
__global__ void k_kernel_octant(double* koct, const double* Q, int n, int nsmall, double delta) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n * n) return;
  const int i = idx % n, j = (idx / n) % n, k = idx / ((long)n * n);
  double v = rvec(i, j, k, delta);
  if (i <= nsmall && j <= nsmall && k <= nsmall) {
    const int NS = 2 * nsmall + 1;
    v += Q[(i + nsmall) + (long)NS * ((j + nsmall) + (long)NS * (k + nsmall))];
  }
  koct[idx] = v / (delta * delta * delta);
}

// out(a, k, b) = sum_i C(k, i) in(a, i, b); in dims (A, I, B), out (A, K, B)
__global__ void k_cos_axis(const double* in, double* out, int A, int I, int B, int K, int M) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)A * K * B) return;
  const int a = idx % A;
  const int k = (idx / A) % K;
  const long b = idx / ((long)A * K);
  double s = 0.0;
  for (int i = 0; i < I; ++i) {
    const int r = (int)(((long)k * i) % M);
    const double c = cospi(2.0 * (double)r / (double)M) * (i == 0 ? 1.0 : 2.0) / (double)M;
    s += c * in[a + (long)A * (i + (long)I * b)];
  }
  out[idx] = s;
}

// S[kx + m1 (ky + m1 kz)] -> khat[kky + m1 (kkz + m1 kkx)], plus FP32 copy
__global__ void k_khat_permute(const double* S, double* kh64, float* kh32, int m1) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)m1 * m1 * m1) return;
  const int kx = idx % m1, ky = (idx / m1) % m1, kz = idx / ((long)m1 * m1);
  const long o = kz + (long)m1 * (ky + (long)m1 * kx);  // [kkx][kky][kkz]: the z frequencies of one column adjacent
  kh64[o] = S[idx];
  kh32[o] = (float)S[idx];
}

__global__ void k_twiddles(z64* tw64, c32* tw32, int M) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= M) return;
  double s, c;
  sincospi(2.0 * (double)k / (double)M, &s, &c);
  tw64[k] = mk(c, -s);
  tw32[k] = mk((float)c, (float)-s);
}

// Nuclear potential of get_3dpoisson (SD/sincproject.f90:192-250): a centre
// at half-spacing offsets cs sees grid point i at offset o = i + (cs-n+1)/2;
// inside the cube |o| <= n-1 the value is the kernel threed_two (from the
// octant, the kernel being even in each axis), outside it is
// 1/(spacing*sqrt(o.o)) with the sqrt in default REAL (realkind4: 4-byte,
// builds without -r8).  pot = sum over centres of -Z_c * value.
__global__ void k_nuclear_pot(int n, double spacing, const double* koct, int ncent, const int* cs, const double* Z,
                              int realkind4, double* pot) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= (long)n * n * n) return;
  const int i = idx % n, j = (idx / n) % n, k = idx / ((long)n * n);
  double acc = 0.0;
  for (int c = 0; c < ncent; ++c) {
    const int ox = i + (-n + cs[3 * c] + 1) / 2;
    const int oy = j + (-n + cs[3 * c + 1] + 1) / 2;
    const int oz = k + (-n + cs[3 * c + 2] + 1) / 2;
    double v;
    if (abs(ox) <= n - 1 && abs(oy) <= n - 1 && abs(oz) <= n - 1) {
      v = koct[abs(ox) + (long)n * (abs(oy) + (long)n * abs(oz))];
    } else {
      const int s2 = ox * ox + oy * oy + oz * oz;
      const double rval = realkind4 ? spacing * (double)sqrtf((float)s2) : spacing * sqrt((double)s2);
      v = 1.0 / rval;
    }
    acc = acc - Z[c] * v;
  }
  pot[idx] = acc;
}

int nuclear_pot_launch(const Spectrum& s, double spacing, int ncent, const int* cs_dev, const double* z_dev,
                       int realkind4, double* pot_dev, hipStream_t st) {
  const int n = s.g.n;
  const long n3 = (long)n * n * n;
  hipLaunchKernelGGL(k_nuclear_pot, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, n, spacing, s.koct, ncent,
                     cs_dev, z_dev, realkind4, pot_dev);
  SDV_HIP_CHECK(hipGetLastError());
  return SDV_OK;
}

static inline unsigned nblk(long n, int t) { return (unsigned)((n + t - 1) / t); }

__global__ void k_real_part(const z64* __restrict__ in, double* __restrict__ out, long n) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = in[i].re;
}

// The FP32 form's charge carrier (see Spectrum::gsub): g(x) = exp(-|x|^2 /
// (2 w^2)) normalised to sum 1 over the grid, centred on the grid, w = max(0.7
// bohr, 1.5 spacing); K*g through the FP64 pipeline.
static int build_charge_carrier(Spectrum* s, double spacing, hipStream_t st) {
  const int n = s->g.n;
  const long n3 = (long)n * n * n;
  const double w = std::max(0.7, 1.5 * spacing);
  std::vector<double> g1(n), g(n3);
  for (int i = 0; i < n; ++i) {
    const double x = (i - 0.5 * (n - 1)) * spacing;  // grid points of SD/SineDVR.f90:16-20, centred
    g1[i] = std::exp(-x * x / (2.0 * w * w));
  }
  double tot = 0.0;
  for (long k = 0; k < n; ++k)
    for (long j = 0; j < n; ++j)
      for (long i = 0; i < n; ++i) tot += g[i + n * (j + n * k)] = g1[i] * g1[j] * g1[k];
  std::vector<z64> gz(n3);
  for (long i = 0; i < n3; ++i) {
    g[i] /= tot;
    gz[i] = mk(g[i], 0.0);
  }
  struct Tmp {
    void *den = nullptr, *out = nullptr, *work = nullptr, *item = nullptr;
    ~Tmp() {
      for (void* p : {den, out, work, item}) (void)hipFree(p);
    }
  } t;
  SDV_HIP_CHECK(hipMalloc(&s->gsub, n3 * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&s->vsub, n3 * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&t.den, n3 * sizeof(z64)));
  SDV_HIP_CHECK(hipMalloc(&t.out, n3 * sizeof(z64)));
  SDV_HIP_CHECK(hipMalloc(&t.work, conv_work_bytes(s->g, 1, 0)));
  SDV_HIP_CHECK(hipMalloc(&t.item, sizeof(PairItem)));
  const PairItem it{0, 0, 0, -1};
  SDV_HIP_CHECK(hipMemcpyAsync(s->gsub, g.data(), n3 * sizeof(double), hipMemcpyHostToDevice, st));
  SDV_HIP_CHECK(hipMemcpyAsync(t.den, gz.data(), n3 * sizeof(z64), hipMemcpyHostToDevice, st));
  SDV_HIP_CHECK(hipMemcpyAsync(t.item, &it, sizeof(PairItem), hipMemcpyHostToDevice, st));
  ConvArgsHost a;
  a.n = n;
  a.npairs = 1;
  a.phi1 = a.phi2 = t.den;
  a.density_mode = 1;
  a.conj_b = 0;
  a.mirror_conj = 0;
  a.pairs_dev = (const PairItem*)t.item;
  a.work = t.work;
  a.out = t.out;
  {
    const int rc = conv_launch(*s, a, 0, st, nullptr);
    if (rc != SDV_OK) return rc;
  }
  hipLaunchKernelGGL(k_real_part, dim3((unsigned)((n3 + 255) / 256)), dim3(256), 0, st, (const z64*)t.out, s->vsub,
                     n3);
  SDV_HIP_CHECK(hipGetLastError());
  SDV_HIP_CHECK(hipStreamSynchronize(st));  // the host vectors and temporaries end here
  return SDV_OK;
}

int build_spectrum(int n, double spacing, int nbig, int nsmall, Spectrum* s, hipStream_t st, int min_m) {
  SDV_REQUIRE(nbig > nsmall, SDV_ERR_ARG, "toothnbig must exceed toothnsmall (SD/tooth.f90:114-117)");
  SDV_REQUIRE(nbig >= n - 1, SDV_ERR_ARG, "toothnbig must be >= numpoints-1 (SD/sincproject.f90:151-153)");
  SDV_REQUIRE(choose_geom(n, &s->g, min_m), SDV_ERR_UNSUPPORTED, "grid edge beyond the supported circulant sizes");
  const Geom& g = s->g;
  const int NS = 2 * nsmall + 1;
  const long ns3 = (long)NS * NS * NS;
  const long n3 = (long)n * n * n;
  const int m1 = g.m + 1;
  // temporaries of the construction, released on every exit path
  struct Tmp {
    double *TR = nullptr, *T2 = nullptr, *W = nullptr, *E = nullptr, *S1 = nullptr, *S2 = nullptr, *S3 = nullptr;
    ~Tmp() {
      for (double* p : {TR, T2, W, E, S1, S2, S3}) (void)hipFree(p);
    }
  } t;
  double *&TR = t.TR, *&T2 = t.T2, *&W = t.W, *&E = t.E, *&S1 = t.S1, *&S2 = t.S2, *&S3 = t.S3;
  SDV_HIP_CHECK(hipMalloc(&TR, ns3 * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&T2, ns3 * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&W, (size_t)NS * NS * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&E, (size_t)NS * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&s->koct, n3 * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&S1, (size_t)m1 * n * n * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&S2, (size_t)m1 * m1 * n * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&S3, (size_t)m1 * m1 * m1 * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&s->khat64, (size_t)m1 * m1 * m1 * sizeof(double)));
  SDV_HIP_CHECK(hipMalloc(&s->khat32, (size_t)m1 * m1 * m1 * sizeof(float)));
  SDV_HIP_CHECK(hipMalloc(&s->tw64, (size_t)g.M * sizeof(z64)));
  SDV_HIP_CHECK(hipMalloc(&s->tw32, (size_t)g.M * sizeof(c32)));

  hipLaunchKernelGGL(k_tr, dim3(nblk(ns3, 256)), dim3(256), 0, st, TR, nsmall, nbig, spacing);
  const size_t hj = (size_t)(NS + (NS & 1)) / 2;
  const size_t shj = (2 * (size_t)NS * NS + kJacThreads + 2 * hj) * sizeof(double) + 2 * hj * sizeof(int);
  SDV_REQUIRE(shj <= 160 * 1024, SDV_ERR_UNSUPPORTED, "toothnsmall too large for the LDS eigensolver");
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shj));
  hipLaunchKernelGGL(k_jacobi, dim3(1), dim3(kJacThreads), shj, st, nsmall, spacing, W, E);
  // QQ = (W^T)^(x3) TR ; QQ /= E_i+E_j+E_k ; Q = W^(x3) QQ   (TOOTHEIGMULT twice)
  hipLaunchKernelGGL(k_axis_mat, dim3(nblk(ns3, 256)), dim3(256), 0, st, TR, T2, W, NS, 2, 1);
  hipLaunchKernelGGL(k_axis_mat, dim3(nblk(ns3, 256)), dim3(256), 0, st, T2, TR, W, NS, 1, 1);
  hipLaunchKernelGGL(k_axis_mat, dim3(nblk(ns3, 256)), dim3(256), 0, st, TR, T2, W, NS, 0, 1);
  hipLaunchKernelGGL(k_eig_divide, dim3(nblk(ns3, 256)), dim3(256), 0, st, T2, E, NS);
  hipLaunchKernelGGL(k_axis_mat, dim3(nblk(ns3, 256)), dim3(256), 0, st, T2, TR, W, NS, 2, 0);
  hipLaunchKernelGGL(k_axis_mat, dim3(nblk(ns3, 256)), dim3(256), 0, st, TR, T2, W, NS, 1, 0);
  hipLaunchKernelGGL(k_axis_mat, dim3(nblk(ns3, 256)), dim3(256), 0, st, T2, TR, W, NS, 0, 0);
  hipLaunchKernelGGL(k_kernel_octant, dim3(nblk(n3, 256)), dim3(256), 0, st, s->koct, TR, n, nsmall, spacing);
  // separable cosine transform of the octant
  hipLaunchKernelGGL(k_cos_axis, dim3(nblk((long)m1 * n * n, 256)), dim3(256), 0, st, s->koct, S1, 1, n, n * n, m1,
                     g.M);
  hipLaunchKernelGGL(k_cos_axis, dim3(nblk((long)m1 * m1 * n, 256)), dim3(256), 0, st, S1, S2, m1, n, n, m1, g.M);
  hipLaunchKernelGGL(k_cos_axis, dim3(nblk((long)m1 * m1 * m1, 256)), dim3(256), 0, st, S2, S3, m1 * m1, n, 1, m1,
                     g.M);
  hipLaunchKernelGGL(k_khat_permute, dim3(nblk((long)m1 * m1 * m1, 256)), dim3(256), 0, st, S3, s->khat64,
                     s->khat32, m1);
  hipLaunchKernelGGL(k_twiddles, dim3(nblk(g.M, 256)), dim3(256), 0, st, (z64*)s->tw64, (c32*)s->tw32, g.M);
  SDV_HIP_CHECK(hipGetLastError());
  SDV_HIP_CHECK(hipStreamSynchronize(st));
  return build_charge_carrier(s, spacing, st);
}

void free_spectrum(Spectrum* s) {
  hipFree(s->tw64);
  hipFree(s->tw32);
  hipFree(s->khat64);
  hipFree(s->khat32);
  hipFree(s->koct);
  hipFree(s->gsub);
  hipFree(s->vsub);
  *s = Spectrum();
}

}  // namespace sdv
