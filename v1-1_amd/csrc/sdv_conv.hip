// sdv_conv.hip -- circulant-embedded Toeplitz 1/r convolution of orbital-pair
// densities on gfx950 (the op_tinv / circ3d_sub path of the reference,
// SD/coreproject.f90:366-537 and SD/toeplitz_subs.f90:561-589).
//
// The reference embeds each n^3 density into a (2n)^3 circulant and runs
// 3 x (1-D FFT + transpose) forward, a pointwise multiply by the kernel
// spectrum, and 3 x (1-D FFT + transpose) back (SD/toeplitz_subs.f90:264-294).
// Here the circulant edge is M = 2m with m = P*Q >= n (any M >= 2n-1 gives
// the same linear convolution), and the transform is pruned and fused:
//
//   k_xfwd : per pencil of x-lines: rho = phi_a * CONJ(phi_b) formed in
//            registers (never stored), zero-padded FFT along x split into the
//            even/odd frequency channels (two m-point FFTs), written as
//            planes X[pair][kx][z][y].                      (R n^3, W 2m n^2)
//   k_yz   : one (pair, kx) plane per workgroup, resident in LDS: y-forward,
//            z-forward, multiply by the real even spectrum, z-inverse,
//            y-inverse -- all pruned to the n x n data region.  (R+W 2m n^2)
//   k_xinv : x-inverse of both channels, keeps x < n, writes V and its
//            mirror column (conj for hermnorm) into twoereduced. (R 2m n^2, W n^3)
//
// Each m-point line FFT is a two-step Cooley-Tukey (P-point DFT in registers,
// twiddle, LDS exchange, Q-point DFT in registers); frequency k1 + P k2 is
// kept at position k1 Q + k2 (the x planes, and the y positions of the FP32
// y/z plane) or at k1 + P k2 (the y positions of the FP64 y/z plane), and
// the inverse consumes that order directly.
#include <algorithm>
#include <cstring>
#include <type_traits>

#include "sdv_dft.h"
#include "sdv_internal.h"

namespace sdv {

template <typename T>
__device__ __forceinline__ cplx<T> cvt(z64 v) {
  return mk((T)v.re, (T)v.im);
}

// Walks e = t, t + step, ... over a row-major (rows x width) array without a
// per-element integer division: (row, col) advance incrementally.
struct RowCol {
  int row, col, drow, dcol, width;
  __device__ __forceinline__ RowCol(int t, int step, int w) : width(w) {
    row = t / w;
    col = t - row * w;
    drow = step / w;
    dcol = step - drow * w;
  }
  __device__ __forceinline__ void next() {
    col += dcol;
    row += drow;
    if (col >= width) {
      col -= width;
      ++row;
    }
  }
};

#ifndef SDV_X_THREADS
#define SDV_X_THREADS 512
#endif
constexpr int kXThreads = SDV_X_THREADS;

template <typename T>
struct KArgs {
  int n, n2, n3, npairs;
  int ld;              // plane stride of X (plane_stride(n) >= n2)
  const z64* phi1;
  const z64* phi2;
  int density_mode, conj_b, mirror_conj;
  const PairItem* pairs;
  cplx<T>* X;
  z64* out;
  const cplx<T>* tw;
  const T* khat;
  int lds_tw2, lds_y;  // runtime LDS offsets (elements) of the k_yz twiddle table and plane
  const double* gsub;  // FP32 form, bra = ket: charge carrier g and K*g of the spectrum (else nullptr)
  const double* vsub;
};

// Everything one work item (one pair's convolution) needs, resolved from the
// single-set arguments or from a multi-set batch table.
template <typename T>
struct ItemView {
  int n, n2, n3;
  int ld;              // plane stride of X (plane_stride(n) >= n2)
  const z64* pa;       // phi1 column a (orbital mode) or density column (density mode)
  const z64* pb;       // phi2 column b
  int density_mode, conj_b, mirror_conj;
  int pack;            // two real diagonal densities |phi_a|^2 + i |phi_b|^2 (outputs Re to out, Im to outm)
  cplx<T>* X;          // this pair's 2m planes of n^2
  z64* out;            // output column
  z64* outm;           // mirror output column or nullptr
  const cplx<T>* tw;
  const T* khat;
  int lds_tw2, lds_y;
  const double* gsub;  // diagonal pair of the FP32 form: rho - g is convolved, K*g added to the output
  const double* vsub;
};

template <typename T>
__device__ __forceinline__ ItemView<T> make_view(const KArgs<T>& a, int pidx, int R) {
  const PairItem it = a.pairs[pidx];
  ItemView<T> v;
  v.n = a.n;
  v.n2 = a.n2;
  v.n3 = a.n3;
  v.ld = a.ld;
  v.pa = a.phi1 + (size_t)it.a * a.n3;
  v.pb = a.phi2 + (size_t)it.b * a.n3;
  v.density_mode = a.density_mode;
  v.conj_b = a.conj_b;
  v.mirror_conj = a.mirror_conj;
  v.X = a.X + (size_t)pidx * 2 * R * a.ld;
  v.out = a.out + (size_t)it.col * a.n3;
  v.outm = it.mcol >= 0 ? a.out + (size_t)it.mcol * a.n3 : nullptr;
  v.tw = a.tw;
  v.khat = a.khat;
  v.lds_tw2 = a.lds_tw2;
  v.lds_y = a.lds_y;
  v.pack = it.pack;
  const bool sub = a.gsub && !a.density_mode && (it.a == it.b || it.pack);
  v.gsub = sub ? a.gsub : nullptr;
  v.vsub = sub ? a.vsub : nullptr;
  return v;
}

// Workgroups of one dispatch go to the 8 XCDs round robin in the order of
// their linear index, and each XCD has its own L2.  xcd_span renumbers the
// linear index L of a grid of `total` workgroups so that the workgroups one
// XCD receives form one contiguous range of the logical order (XCD x takes
// the x-th of 8 near-equal ranges).  It is a bijection of [0, total) whatever
// the hardware order, so only cache reuse depends on it.
constexpr unsigned kXcds = 8;
__device__ __forceinline__ unsigned xcd_span(unsigned L, unsigned total) {
  const unsigned x = L % kXcds, j = L / kXcds, q = total / kXcds, r = total % kXcds;
  return x < r ? x * (q + 1) + j : r * (q + 1) + (x - r) * q + j;
}

template <int P, int Q, int XT_ = kXThreads>
struct XCfg {
  static constexpr int MX = P > Q ? P : Q;
  static constexpr int XT = XT_;        // workgroup size of the x kernels
  static constexpr int LA = XT / MX;    // x-lines per workgroup (56 at m = 81: 0.9 KB runs per plane row)
};

// y/z plane kernel: every y-row and z-column FFT of the plane is owned by one
// lane group of G = max(P, Q) lanes inside a single wavefront (GPW groups
// per wave), so the exchanges inside a line FFT need no workgroup barrier.
// A group could take LPG lines one after the other (smaller workgroups, two
// per CU for FP32's 54 KB of LDS at m = 81); measured at 81^3 that is 1.4x
// slower in FP32 than one line per group at 12 waves, and 1.2x slower in FP64
// (6 waves, no register spill, against 32 bytes per lane of spill at 12 with
// the odd twiddles in registers), so
// LPG = 1.  Static wave priorities that differ between the three waves of a
// SIMD (to stagger their LDS and VALU phases) change nothing measurable.
// LDS after the plane.  FP32: the kernel stages its x-frequency plane's slice
// of the spectrum, (m+1)^2 reals, when the workgroup's LDS still fits, so the
// z pass reads it from LDS in the middle of its column FFTs.  FP64: the rows
// of the parked cy = 0 output that fit in the rest of the 160 KB (42 of 81 at
// m = 81, six whole wavefronts of rows) stay in LDS instead of HBM; measured
// faster there than staging the spectrum (which was neutral in FP64).
#ifndef SDV_YZ_KHAT_LDS
#define SDV_YZ_KHAT_LDS 1
#endif
#ifndef SDV_YZ_TW_REGS
#define SDV_YZ_TW_REGS 0
#endif
#ifndef SDV_YZ_TWB
#define SDV_YZ_TWB 3
#endif

// Plane pitch in elements.  FP64 (16-byte elements): the smallest p >= R with
// p = 9 mod 16, which together with the exchange slot maps of body_yz makes
// every LDS read of the plane free of bank conflicts (64 banks of 4 bytes, a
// b128 access served 16 lanes at a time: at p = 81 half of them are two-way
// conflicted, 8.2 against 4.9 cycles per wave instruction, while writes cost
// the same either way).  FP32: R | 1.
constexpr int yz_pitch(int R, int es) { return es == 16 ? R + ((9 - R) & 15) : (R | 1); }
constexpr size_t yz_base_bytes(int R, int es) {
  return (size_t)(3 * R + (R & 1) + R * yz_pitch(R, es)) * es;  // tables (2R + R + pad) and plane, bytes
}
constexpr bool yz_khat_in_lds(int R, int es) {
  return es == 8 && SDV_YZ_KHAT_LDS && yz_base_bytes(R, es) + (size_t)(R + 1) * (R + 1) * (es / 2) <= 160 * 1024;
}
// rows z < yz_park_rows park their cy = 0 output in LDS (FP64)
constexpr int yz_park_rows(int R, int es) {
  return es != 16 ? 0 : (int)((160 * 1024 - yz_base_bytes(R, es)) / ((size_t)R * es) < (size_t)R
                                  ? (160 * 1024 - yz_base_bytes(R, es)) / ((size_t)R * es) : (size_t)R);
}
constexpr size_t yz_lds_bytes(int R, int es) {
  return yz_base_bytes(R, es) + (yz_khat_in_lds(R, es) ? (size_t)(R + 1) * (R + 1) * (es / 2) : 0) +
         (size_t)yz_park_rows(R, es) * R * es;
}

template <typename T, int P, int Q>
struct YZCfg;
// Wavefronts per SIMD the y/z kernel is compiled for: one workgroup's
// (NT / 256), except FP32 at m = 81, compiled for 6 -- two workgroups of 12
// waves per CU (the plane, its tables and the spectrum slice take 81 KB of
// LDS, so two fit the 160 KB).  That holds it to 80 VGPRs with 52 bytes of
// spill, and measured at 81^3 the second workgroup wins: 1.40-1.43 ->
// 1.33-1.35 ms per 128-convolution launch.  Smaller planes already fit
// several workgroups by LDS, and the same cap only adds spills there (the
// 64-set batch 216k -> 192k pair potentials/s, 41^3 275k -> 230k).
#ifndef SDV_YZ_WPE32
#define SDV_YZ_WPE32 6
#endif
template <typename T, int P, int Q>
constexpr int yz_wpe() {
  return sizeof(T) == 4 && P * Q == 81 ? SDV_YZ_WPE32 : (YZCfg<T, P, Q>::NT + 255) / 256;
}

template <typename T, int P, int Q>
struct YZCfg {
  static constexpr int R = P * Q;
  static constexpr int G = P > Q ? P : Q;
  static constexpr int GPW = 64 / G;                                   // lane groups per wavefront
  static constexpr int LPG = 1;                                        // lines per group
  static constexpr int NG = (R + LPG - 1) / LPG;                       // groups needed
  static constexpr int NT = 64 * ((NG + GPW - 1) / GPW);
  static constexpr int NGB = NT / 64 * GPW;                            // groups in the workgroup
};

// ---------------------------------------------------------------- x forward
template <typename T, int P, int Q, int XT = kXThreads>
__device__ __forceinline__ void xfwd_step2(const ItemView<T>& a, const cplx<T>* buf, int l0, int nl, int c) {
  constexpr int R = P * Q, LA = XCfg<P, Q, XT>::LA, pitch = R | 1;
  const int tid = threadIdx.x;
  const int line = tid % LA, k1 = tid / LA;
  if (k1 < P && line < nl) {
    cplx<T> u[Q];
#pragma unroll
    for (int j2 = 0; j2 < Q; ++j2) u[j2] = buf[line * pitch + k1 * Q + j2];
    DFT<T, Q, false>::run(u);
    cplx<T>* dst = a.X + ((size_t)c * R + (size_t)k1 * Q) * a.ld + l0 + line;
#pragma unroll
    for (int k2 = 0; k2 < Q; ++k2) dst[(size_t)k2 * a.ld] = u[k2];
  }
}

// x forward of lines [chunk*LA, chunk*LA + LA) of one pair (XT threads).
template <typename T, int P, int Q, int XT = kXThreads>
__device__ __forceinline__ void body_xfwd(const ItemView<T>& a, int chunk, unsigned char* smem_raw) {
  constexpr int R = P * Q, LA = XCfg<P, Q, XT>::LA, pitch = R | 1, M = 2 * R;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + M;
  const int tid = threadIdx.x;
  const int n = a.n;
  const int l0 = chunk * LA;
  const int nl = min(LA, a.n2 - l0);
  for (int i = tid; i < M; i += XCfg<P, Q, XT>::XT) tw[i] = a.tw[i];
  const z64* pa = a.pa + (size_t)l0 * n;
  const z64* pb = a.pb + (size_t)l0 * n;
  // The loads of this thread are issued in NB batches, each batch before
  // the first of its values is consumed (a loop that waits for each load in
  // turn pays the memory latency ITER times: 0.85 -> 0.73 ms for the FP64
  // kernel at 81^3, 136 pairs).  FP32 takes two batches: the FP64 orbital
  // values of one batch hold half the registers, so four workgroups fit a
  // CU instead of two (0.49 -> 0.46 ms); FP64 is held to two workgroups by
  // its LDS and takes one.
  constexpr int ITER = (LA * R + XCfg<P, Q, XT>::XT - 1) / XCfg<P, Q, XT>::XT;  // n <= R
  constexpr int NB = sizeof(T) == 4 ? 2 : 1;
  constexpr int IB = (ITER + NB - 1) / NB;
  const int ne = nl * n;
  const double* gsub = a.gsub ? a.gsub + (size_t)l0 * n : nullptr;
  RowCol rc(tid, XCfg<P, Q, XT>::XT, n);
#pragma unroll
  for (int bch = 0; bch < NB; ++bch) {
    z64 ua[IB], ub[IB];
    double gs[IB];
#pragma unroll
    for (int k = 0; k < IB; ++k) {
      const int i = bch * IB + k;
      const int e = tid + i * XCfg<P, Q, XT>::XT;
      gs[k] = 0.0;
      if (i < ITER && e < ne) {
        ua[k] = pa[e];
        if (!a.density_mode) ub[k] = pb[e];
        if (gsub) gs[k] = gsub[e];
      }
    }
#pragma unroll
    for (int k = 0; k < IB; ++k) {
      const int i = bch * IB + k;
      if (i >= ITER) break;
      if (tid + i * XCfg<P, Q, XT>::XT < ne) {
        z64 v = ua[k];
        if (a.pack) {  // |phi_a|^2 - g + i (|phi_b|^2 - g)
          const z64 w = ub[k];
          const z64 da = v * mk(v.re, -v.im), db = w * mk(w.re, -w.im);
          v = mk(da.re - gs[k], db.re - gs[k]);
        } else {
          if (!a.density_mode) {  // SD/coreproject.f90:1057-1065: inspfs10(:,a) * CONJUGATE(inspfs20(:,b))
            z64 w = ub[k];
            if (a.conj_b) w.im = -w.im;
            v = v * w;
          }
          v.re -= gs[k];  // neutral remainder of a diagonal pair density (FP32 form)
        }
        buf[rc.row * pitch + rc.col] = cvt<T>(v);
      }
      rc.next();
    }
  }
  __syncthreads();
  const int line = tid % LA, j2 = tid / LA;
  const bool act = (j2 < Q) && (line < nl);
  cplx<T> v[P], b[P];
  if (act) {
#pragma unroll
    for (int j1 = 0; j1 < P; ++j1) {
      const int x = j1 * Q + j2;
      cplx<T> u = (x < n) ? buf[line * pitch + x] : mk((T)0, (T)0);
      v[j1] = u;
      b[j1] = u * tw[x];  // odd channel: x * w_M^x
    }
    DFT<T, P, false>::run(v);
    DFT<T, P, false>::run(b);
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) {
      const cplx<T> w = tw[(2 * j2 * k1) % M];
      v[k1] = v[k1] * w;
      b[k1] = b[k1] * w;
      buf[line * pitch + k1 * Q + j2] = v[k1];
    }
  }
  __syncthreads();
  xfwd_step2<T, P, Q, XT>(a, buf, l0, nl, 0);
  __syncthreads();
  if (act) {
#pragma unroll
    for (int k1 = 0; k1 < P; ++k1) buf[line * pitch + k1 * Q + j2] = b[k1];
  }
  __syncthreads();
  xfwd_step2<T, P, Q, XT>(a, buf, l0, nl, 1);
}

// Workgroup sizes of the x kernels per precision, measured at 81^3: the FP32
// x forward at 288 threads (32 lines, 0.47 -> 0.39 ms), the FP64 x inverse at
// 360 (40 lines, 0.84 -> 0.81 ms), the other two at 512 (the FP32 x inverse
// runs the same at 288).  The batched kernels share one line-chunk list
// between their x forward and x inverse: 288 in FP32, 512 in FP64.  The
// arithmetic of a line does not depend on the size.
template <typename T>
struct XThreads {
  static constexpr int fwd = sizeof(T) == 4 ? 288 : kXThreads;
  static constexpr int inv = sizeof(T) == 8 ? 360 : kXThreads;
  static constexpr int batch = sizeof(T) == 4 ? 288 : kXThreads;  // both batched x kernels
};

template <typename T, int P, int Q>
__global__ void __launch_bounds__(XThreads<T>::fwd) k_xfwd(KArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // pairs run fastest in the logical order, so the workgroups resident at
  // one time work on the same line chunk of many pairs, and each XCD takes
  // one range of line chunks, so each orbital chunk is fetched into one
  // XCD's L2 and served to all its pairs there (pair-fastest over the raw
  // index spreads every chunk's pairs over all 8 XCDs and fetches the
  // orbitals 8 times: 1.08 GB per launch at 81^3 with 16 orbitals;
  // 0.60 -> 0.57 ms FP64, FP32 unchanged)
  const unsigned L = xcd_span(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  body_xfwd<T, P, Q, XThreads<T>::fwd>(make_view(a, L % gridDim.x, P * Q), L / gridDim.x, smem_raw);
}

// ------------------------------------------------------------- y/z planes
// One (pair, kx) plane per workgroup.  LDS (elements of cplx<T>):
//   tw[M]  = w_M^k ;  T2[a*P + b] = w_m^(a b) (a < Q, b < P)
//   Y      = m x pitch, the plane: [ky-position][z] in FP64 (a column of one
//            ky contiguous: the column FFTs, which hold the most registers,
//            address it from one base with immediate offsets), [z][ky] in FP32
// Per y channel cy (even, odd):
//   rows    : y-forward of row z in place (input y < n from HBM), position
//             k1 + P*k2 (FP64) or k1*Q + k2 (FP32) holds frequency k1 + P*k2;
//   columns : for the column of one ky, both z channels one after the other
//             by the same lane group (its input kept in registers), each a
//             forward FFT, the multiply by the real even spectrum and the
//             inverse, accumulated; the column's m slots are the exchange
//             scratch of its own FFTs;
//   rows    : y-inverse of row z, this channel's output part.
// Five workgroup barriers per plane.  The cy = 0 output part stays in
// registers (FP32, and FP64 on full grids) or is parked (FP64 with n < m,
// whose registers are needed for the column FFTs; the LDS after the plane is
// reserved for the parked rows on full grids too, where it goes unused):
// rows z < yz_park_rows in LDS after the plane, each lane in
// slots of its own, the other rows in the plane's own HBM slots, which the
// same lanes have just read for the cy = 1 input; the parked values are loaded
// before the last y-inverse, so their latency overlaps it.  The pitch is
// yz_pitch (FP32: other row pitches than R | 1 measured no faster, even ones
// 18 % slower).
// The LDS offsets are runtime values, so the compiler cannot prove the
// twiddle tables distinct from the plane and does not pin hoisted twiddles
// in VGPRs across phases.

// Ordering of LDS accesses between lanes of one wavefront (a wave's LDS
// operations execute in order; this keeps the compiler from reordering).
__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// V is KArgs<T> (single set: the fields stay in the kernel-argument segment,
// where the compiler can reload them instead of holding them in registers) or
// ItemView<T>; base points at plane number `plane` (of 2m) of this pair.
// With GUARD (blocks larger than NT) lanes at or beyond NT only take part in
// the barriers.
// FULL: the grid edge equals the circulant half-edge (n = m = R, e.g. 81^3),
// so every row, column and position guard against n is known true at compile
// time.
template <typename T, int P, int Q, bool GUARD, typename V, bool FULL = false>
__device__ __forceinline__ void body_yz(const V& a, cplx<T>* base, int plane, unsigned char* smem_raw) {
  constexpr int R = P * Q, M = 2 * R, pitch = yz_pitch(R, (int)sizeof(cplx<T>));
  typedef YZCfg<T, P, Q> C;
  constexpr int G = C::G, GPW = C::GPW, NT = C::NT;
  // the cy = 0 output part stays in registers in FP32 and in FP64 on full
  // grids (150 VGPRs at m = 81, no spill; 2.58 -> 2.45 ms at 81^3 against
  // parking); FP64 with n < m parks it (keeping it there spills)
  constexpr bool KEEP = sizeof(T) == 4 || FULL;
  // element (ky position p, row z) at Y[p * SP + z * SR]: FP64 keeps a
  // column contiguous (its column FFTs are the register-bound phase), FP32
  // keeps a row contiguous
  constexpr int SP = sizeof(T) == 8 ? pitch : 1, SR = sizeof(T) == 8 ? 1 : pitch;
  // Slots of the forward exchanges and of the y-forward output.  FP64 puts the
  // value of input index j2 and output index k1 of a line FFT at slot
  // j2 P + k1 (written with the writer lane's j2 in the high digit, read with
  // the reader lane's k1 in the low one) and frequency k1 + P k2 at position
  // k1 + P k2, so that with its pitch every read has consecutive slots across
  // a lane group; FP32 uses slot k1 Q + j2 and position k1 Q + k2.
  constexpr bool XR = sizeof(T) == 8;
  auto xslot = [](int j2, int k1) { return XR ? j2 * P + k1 : k1 * Q + j2; };
  auto fpos = [](int k1, int k2) { return XR ? k1 + P * k2 : k1 * Q + k2; };
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* T2 = tw + a.lds_tw2;
  cplx<T>* Y = tw + a.lds_y;
  const int t = (!GUARD || threadIdx.x < NT) ? (int)threadIdx.x : (1 << 29);
  const int lane = t & 63, gi = lane / G, li = lane - gi * G;
  const int grp = gi < GPW ? (t >> 6) * GPW + gi : (1 << 28);  // the one line (row, then column) of this group
  const int n = FULL ? R : a.n;
  constexpr bool KH_LDS = yz_khat_in_lds(R, (int)sizeof(cplx<T>));
  T* khs = reinterpret_cast<T*>(Y + R * pitch);
  constexpr int PR = yz_park_rows(R, (int)sizeof(cplx<T>));
  cplx<T>* park = Y + R * pitch;

  // this plane's x frequency and its spectrum slice
  const int cx = plane / R, posx = plane - cx * R;
  const int kx = 2 * (posx / Q + P * (posx % Q)) + cx;
  const T* khg = a.khat + (size_t)min(kx, M - kx) * (R + 1) * (R + 1);
  auto stage_kh = [&]() {  // FP32: every load in flight before the first store
    constexpr int KI = ((R + 1) * (R + 1) + NT - 1) / NT;
    T kv[KI];
#pragma unroll
    for (int i = 0; i < KI; ++i) kv[i] = t + i * NT < (R + 1) * (R + 1) ? khg[t + i * NT] : (T)0;
#pragma unroll
    for (int i = 0; i < KI; ++i)
      if (t + i * NT < (R + 1) * (R + 1)) khs[t + i * NT] = kv[i];
  };

  // Odd-channel twiddle w_M^(j1 Q + li) = w_M^li * w_2P^j1: the correctly
  // rounded table entry from LDS.  SDV_YZ_TW_REGS=1 forms it in FP64 as the
  // lane's w_M^li (one LDS read) times the compile-time w_2P^j1, saving P - 1
  // 16-byte LDS reads per use at one complex multiply each; its registers
  // raise the kernel's spill from 20 to 32 bytes per lane and it runs 3.5 %
  // slower at 81^3, so both precisions read the table.
  constexpr bool TW_REG = sizeof(T) == 8 && SDV_YZ_TW_REGS;
  auto odd_base = [&]() -> cplx<T> { return TW_REG ? tw[li] : mk((T)0, (T)0); };
  auto odd_tw = [&](const cplx<T>& wb, int j1) -> cplx<T> {
    if (TW_REG) return j1 == 0 ? wb : wb * mk((T)kTw<2 * P>.c[j1], (T)-kTw<2 * P>.s[j1]);
    return tw[j1 * Q + li];
  };
  // FP32: the lane's step-1 twiddles read into registers before the DFT they
  // follow (the compiler cannot tell the table from the plane, so a table
  // read left between the plane stores waits for every store before it;
  // 1.54 -> 1.51 ms at 81^3).  FP64 reads them in the store loop: its
  // registers hold the column input and accumulator, and hoisting spills.
  // FP32 compiled for two workgroups per CU (m = 81, 80 VGPRs) reads them in
  // the store loop too, as FP64 does (spill 52 -> 20 bytes with the column
  // twiddles below also left in LDS; 1.34-1.38 -> 1.29-1.32 ms).
  constexpr bool TWO_WG = sizeof(T) == 4 && yz_wpe<T, P, Q>() > 3;
  constexpr bool HOIST = sizeof(T) == 4 && !TWO_WG;
  auto t2_load = [&](cplx<T>* t2v, bool inv) {
#pragma unroll
    for (int k = 0; k < P; ++k) t2v[k] = inv ? T2[k * P + li] : T2[li * P + k];
  };
  // FP64 stores a step's K outputs v[k] * tab(k) with the table reads issued
  // TWB at a time ahead of the stores they scale.  A read placed between two
  // plane stores can only be issued after them (the table and plane offsets
  // are runtime values, so the compiler must assume they alias) and its
  // result waits for every LDS operation before it: one read per store costs
  // a full LDS round trip per element.  Batches keep TWB twiddles live
  // (hoisting all nine spills at 166 VGPRs).
  // Every step-1 twiddle table has w^0 = 1 at index 0 (T2[l P] and T2[l]),
  // so output 0 is stored unscaled: no read, no multiply.
  constexpr int TWB = SDV_YZ_TWB;
  auto mul_store = [&](auto kk, const cplx<T>* v, auto tab, auto st) {
    constexpr int K = decltype(kk)::value;
#pragma unroll
    for (int k0 = 0; k0 < K; k0 += TWB) {
      cplx<T> w[TWB];
#pragma unroll
      for (int b = 0; b < TWB; ++b)
        if (k0 + b < K && k0 + b > 0) w[b] = tab(k0 + b);
#pragma unroll
      for (int b = 0; b < TWB; ++b)
        if (k0 + b < K) st(k0 + b, k0 + b == 0 ? v[0] : v[k0 + b] * w[b]);
    }
  };
  // acc + v * conj(w): in FP64 each component one expression (two fused
  // multiply-adds, 4 VALU operations instead of 6); FP32 rounds the product
  // first, as the mixed form's error figures were measured with (fused, the
  // RMS error of the 41^3 c-norm diagonal pair potentials is unchanged but
  // their largest single-element error went from 8.1e-8 to 1.5e-7)
  auto macc_conj = [](const cplx<T>& acc, const cplx<T>& v, const cplx<T>& w) -> cplx<T> {
    if constexpr (sizeof(T) == 8)
      return mk(acc.re + v.re * w.re + v.im * w.im, acc.im + v.im * w.re - v.re * w.im);
    else
      return acc + v * conj(w);
  };


  // row input x[j1] = plane(z = g, y = j1 Q + li) of the plane at b
  auto load_row = [&](const cplx<T>* b, int g, cplx<T>* x) {
    if (g < n && li < Q) {
      const cplx<T>* grow = b + (size_t)g * n;
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) {
        const int y = j1 * Q + li;
        x[j1] = y < n ? grow[y] : mk((T)0, (T)0);
      }
    }
  };
  // y-forward of row g for channel cy (row z = g: position p at yrow[p * SP])
  auto yfwd = [&](int cy, int g, const cplx<T>* x) {
    const bool rowok = g < n;
    cplx<T>* yrow = Y + g * SR;
    if (rowok && li < Q) {
      cplx<T> v[P];
      const cplx<T> wb = odd_base();
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) {
        const int y = j1 * Q + li;
        v[j1] = (cy && y < n) ? x[j1] * odd_tw(wb, j1) : x[j1];
      }
      cplx<T> t2v[HOIST ? P : 1];
      if constexpr (HOIST) t2_load(t2v, false);
      DFT<T, P, false>::run(v);
      if constexpr (HOIST) {
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) yrow[xslot(li, k1) * SP] = k1 == 0 ? v[0] : v[k1] * t2v[HOIST ? k1 : 0];
      } else {
        mul_store(std::integral_constant<int, P>{}, v, [&](int k1) { return T2[li * P + k1]; },
                  [&](int k1, const cplx<T>& x) { yrow[xslot(li, k1) * SP] = x; });
      }
    }
    wave_sync();
    if (rowok && li < P) {
      cplx<T> u[Q];
#pragma unroll
      for (int j2 = 0; j2 < Q; ++j2) u[j2] = yrow[xslot(j2, li) * SP];
      DFT<T, Q, false>::run(u);
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) yrow[fpos(li, k2) * SP] = u[k2];
    }
  };
  // both z channels of column c (ky position), spectrum of y channel cy
  auto zpass = [&](int cy, int c) {
    if (c >= R) return;
    const T* kh = KH_LDS ? khs : khg;
    cplx<T>* ycol = Y + c * SP;  // column c: row z at ycol[z * SR]
    const int ky = XR ? 2 * c + cy : 2 * (c / Q + P * (c - (c / Q) * Q)) + cy;
    const int kcol = (R + 1) * min(ky, M - ky);  // this column's z frequencies, adjacent; 32-bit offset from a uniform base
    cplx<T> u[P], acc[P];
    if (li < Q) {
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) {
        const int z = j1 * Q + li;
        u[j1] = z < n ? ycol[z * SR] : mk((T)0, (T)0);
      }
    }
    // FP32 with P = Q: the nine T2 twiddles of this lane serve the forward and
    // the inverse of both channels (T2[a P + b] is symmetric in a, b), so they
    // are read from LDS once per column (not with TWO_WG)
    constexpr bool T2C = sizeof(T) == 4 && P == Q && !TWO_WG;
    constexpr bool H2 = HOIST && !T2C;
    cplx<T> t2c[T2C ? P : 1];
    if (T2C && li < Q) {
#pragma unroll
      for (int k = 0; k < (T2C ? P : 1); ++k) t2c[k] = T2[li * P + k];
    }
    auto t2f = [&](int k1) -> cplx<T> { return T2C ? t2c[T2C ? k1 : 0] : T2[li * P + k1]; };
    auto t2i = [&](int j2) -> cplx<T> { return T2C ? t2c[T2C ? j2 : 0] : T2[j2 * P + li]; };
    // one z channel: forward, spectral multiply, inverse, accumulated
    auto zchan = [&](auto czv) {
      const int cz = czv;
      if (li < Q) {
        cplx<T> v[P];
        const cplx<T> wb = odd_base();
#pragma unroll
        for (int j1 = 0; j1 < P; ++j1) v[j1] = cz ? u[j1] * odd_tw(wb, j1) : u[j1];
        cplx<T> t2v[H2 ? P : 1];
        if constexpr (H2) t2_load(t2v, false);
        DFT<T, P, false>::run(v);
        if constexpr (HOIST) {
#pragma unroll
          for (int k1 = 0; k1 < P; ++k1)
            ycol[xslot(li, k1) * SR] = k1 == 0 ? v[0] : v[k1] * (H2 ? t2v[H2 ? k1 : 0] : t2f(k1));
        } else {
          mul_store(std::integral_constant<int, P>{}, v, t2f,
                    [&](int k1, const cplx<T>& x) { ycol[xslot(li, k1) * SR] = x; });
        }
      }
      wave_sync();
      if (li < P) {  // step 2, spectral multiply, inverse step 1
        cplx<T> w[Q];
#pragma unroll
        for (int j2 = 0; j2 < Q; ++j2) w[j2] = ycol[xslot(j2, li) * SR];
        cplx<T> t2v[H2 ? Q : 1];
        if constexpr (H2) t2_load(t2v, true);
        DFT<T, Q, false>::run(w);
#pragma unroll
        for (int k2 = 0; k2 < Q; ++k2) {
          const int kz = 2 * (li + P * k2) + cz;
          w[k2] = scale(w[k2], kh[kcol + min(kz, M - kz)]);
        }
        DFT<T, Q, true>::run(w);
        if constexpr (HOIST) {
#pragma unroll
          for (int j2 = 0; j2 < Q; ++j2)
            ycol[(li * Q + j2) * SR] = j2 == 0 ? w[0] : w[j2] * conj(H2 ? t2v[H2 ? j2 : 0] : t2i(j2));
        } else {
          mul_store(std::integral_constant<int, Q>{}, w, [&](int j2) { return conj(t2i(j2)); },
                    [&](int j2, const cplx<T>& x) { ycol[(li * Q + j2) * SR] = x; });
        }
      }
      wave_sync();
      if (li < Q) {  // inverse step 2; the odd channel carries w_M^-z
        cplx<T> v[P];
#pragma unroll
        for (int k1 = 0; k1 < P; ++k1) v[k1] = ycol[(k1 * Q + li) * SR];
        DFT<T, P, true>::run(v);
        const cplx<T> wb = odd_base();
#pragma unroll
        for (int j1 = 0; j1 < P; ++j1) acc[j1] = cz ? macc_conj(acc[j1], v[j1], odd_tw(wb, j1)) : v[j1];
      }
    };
    // both channels unrolled (cz known at compile time): with a runtime
    // channel loop FP64 branched around every odd-channel twiddle and held
    // 166 VGPRs; unrolled it holds 114-148 (2.68 -> 2.58 ms at 81^3)
    zchan(std::integral_constant<int, 0>{});
    zchan(std::integral_constant<int, 1>{});
    if (li < Q) {
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) {
        const int z = j1 * Q + li;
        if (z < n) ycol[z * SR] = acc[j1];
      }
    }
  };
  // y-inverse of row g for channel cy -> o[j1] at y = j1 Q + li; the odd
  // channel's output is added to add[j1] (the even channel's part) in the
  // same fused multiply-adds that apply its w_M^-y
  auto yinv = [&](int cy, int g, cplx<T>* o, const cplx<T>* add) {
    const bool rowok = g < n;
    cplx<T>* yrow = Y + g * SR;
    if (rowok && li < P) {
      cplx<T> u[Q], t2v[HOIST ? Q : 1];
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) u[k2] = yrow[fpos(li, k2) * SP];
      if constexpr (HOIST) t2_load(t2v, true);
      DFT<T, Q, true>::run(u);
      if constexpr (HOIST) {
#pragma unroll
        for (int j2 = 0; j2 < Q; ++j2) yrow[(li * Q + j2) * SP] = j2 == 0 ? u[0] : u[j2] * conj(t2v[HOIST ? j2 : 0]);
      } else {
        mul_store(std::integral_constant<int, Q>{}, u, [&](int j2) { return conj(T2[j2 * P + li]); },
                  [&](int j2, const cplx<T>& x) { yrow[(li * Q + j2) * SP] = x; });
      }
    }
    wave_sync();
    if (rowok && li < Q) {
      cplx<T> v[P];
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) v[k1] = yrow[(k1 * Q + li) * SP];
      DFT<T, P, true>::run(v);
      const cplx<T> wb = odd_base();
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) o[j1] = cy ? macc_conj(add[j1], v[j1], odd_tw(wb, j1)) : v[j1];
    }
  };

  for (int i = t; i < M; i += NT) tw[i] = a.tw[i];
  for (int i = t; i < R; i += NT) {
    const int aa = i / P, bb = i - aa * P;
    T2[i] = a.tw[(2 * aa * bb) % M];
  }
  if (KH_LDS) stage_kh();
  cplx<T> x[P];
  load_row(base, grp, x);  // in flight during the table stores and the barrier
  __syncthreads();         // tables (and the spectrum slice)
  yfwd(0, grp, x);
  __syncthreads();
  zpass(0, grp);
  __syncthreads();
  const int g = grp;
  cplx<T> o[P], keep[P];
  yinv(0, g, o, o);
  load_row(base, g, x);  // the cy = 1 input: the same rows again, mostly from L2
  if (KEEP) {
#pragma unroll
    for (int j1 = 0; j1 < P; ++j1) keep[j1] = o[j1];  // cy = 0 output part, in registers
    yfwd(1, g, x);
  } else {
    yfwd(1, g, x);  // consumes x before its HBM slots take the parked cy = 0 output
    if (g < PR) {
      if (li < Q) {
#pragma unroll
        for (int j1 = 0; j1 < P; ++j1) park[(g * Q + li) * P + j1] = o[j1];
      }
    } else if (g < n && li < Q) {
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) {
        const int y = j1 * Q + li;
        if (y < n) base[(size_t)g * n + y] = o[j1];
      }
    }
  }
  __syncthreads();
  zpass(1, grp);
  __syncthreads();
  if (!KEEP) {
    if (g < PR) {
      if (li < Q) {
#pragma unroll
        for (int j1 = 0; j1 < P; ++j1) keep[j1] = park[(g * Q + li) * P + j1];
      }
    } else if (g < n && li < Q) {  // the parked cy = 0 part, in flight during the y-inverse
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) {
        const int y = j1 * Q + li;
        keep[j1] = y < n ? base[(size_t)g * n + y] : mk((T)0, (T)0);
      }
    }
  }
  yinv(1, g, o, keep);  // o = keep + the odd channel's part
  if (g < n && li < Q) {
#pragma unroll
    for (int j1 = 0; j1 < P; ++j1) {
      const int y = j1 * Q + li;
      if (y < n) base[(size_t)g * n + y] = o[j1];
    }
  }
}

// Logical order plane-major, pairs fastest, one range per XCD: an XCD works
// on a few x-frequency planes and keeps their spectrum slices in its L2 (FP32
// 81^3: 1.60 -> 1.57 ms; FP64 unchanged).  One plane per workgroup: a
// persistent form (the CU's workgroup walking a strided share of its XCD's
// range, the plane body in a function of its own) measured 2.83 -> 3.20 ms in
// FP64 at 81^3, the hardware's dynamic workgroup dispatch balancing better.
template <typename T, int P, int Q, bool FULL>
__global__ void __launch_bounds__((YZCfg<T, P, Q>::NT)) __attribute__((amdgpu_waves_per_eu(yz_wpe<T, P, Q>()))) k_yz(KArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const unsigned L = xcd_span(blockIdx.x + blockIdx.y * gridDim.x, gridDim.x * gridDim.y);
  const unsigned plane = L / gridDim.y, pair = L % gridDim.y;
  body_yz<T, P, Q, false, KArgs<T>, FULL>(a, a.X + ((size_t)pair * 2 * (P * Q) + plane) * a.ld, (int)plane,
                                         smem_raw);
}

// ---------------------------------------------------------------- x inverse
// x inverse of lines [chunk*LA, chunk*LA + LA) of one pair and the output
// scatter (XT threads).
template <typename T, int P, int Q, int XT = kXThreads>
__device__ __forceinline__ void body_xinv(const ItemView<T>& a, int chunk, unsigned char* smem_raw) {
  constexpr int R = P * Q, LA = XCfg<P, Q, XT>::LA, pitch = R | 1, M = 2 * R;
  cplx<T>* tw = reinterpret_cast<cplx<T>*>(smem_raw);
  cplx<T>* buf = tw + M;
  const int tid = threadIdx.x;
  const int n = a.n;
  const int l0 = chunk * LA;
  const int nl = min(LA, a.n2 - l0);
  for (int i = tid; i < M; i += XCfg<P, Q, XT>::XT) tw[i] = a.tw[i];
  const int line = tid % LA, jj = tid / LA;
  cplx<T> acc[P];
#pragma unroll
  for (int j = 0; j < P; ++j) acc[j] = mk((T)0, (T)0);
  for (int c = 0; c < 2; ++c) {
    const cplx<T>* src = a.X + (size_t)c * R * a.ld + l0;
    // every load of the tile is issued before the first LDS store, as in the
    // x forward (one load in flight per thread leaves a CU far short of the
    // bytes in flight HBM latency needs: 0.89 -> 0.84 ms FP64, 0.69 -> 0.63 ms
    // FP32 at 81^3, 136 pairs).  Row and column come from the element index:
    // stepping a row/column pair through the unrolled loop kept the loaded
    // values in scratch memory instead of registers.
    constexpr int ITER = (LA * R + XCfg<P, Q, XT>::XT - 1) / XCfg<P, Q, XT>::XT;
    const int ne = R * nl;
    cplx<T> ld[ITER];
#pragma unroll
    for (int i = 0; i < ITER; ++i) {
      const int e = tid + i * XCfg<P, Q, XT>::XT;
      const int row = e / nl;
      ld[i] = e < ne ? src[(size_t)row * a.ld + (e - row * nl)] : mk((T)0, (T)0);
    }
    {
      RowCol rc(tid, XCfg<P, Q, XT>::XT, nl);
#pragma unroll
      for (int i = 0; i < ITER; ++i, rc.next())
        if (tid + i * XCfg<P, Q, XT>::XT < ne) buf[rc.col * pitch + rc.row] = ld[i];
    }
    __syncthreads();
    if (jj < P && line < nl) {  // inverse step 1: task (line, k1)
      cplx<T> u[Q];
#pragma unroll
      for (int k2 = 0; k2 < Q; ++k2) u[k2] = buf[line * pitch + jj * Q + k2];
      DFT<T, Q, true>::run(u);
#pragma unroll
      for (int j2 = 0; j2 < Q; ++j2) buf[line * pitch + jj * Q + j2] = u[j2] * conj(tw[(2 * j2 * jj) % M]);
    }
    __syncthreads();
    if (jj < Q && line < nl) {  // inverse step 2: task (line, j2)
      cplx<T> v[P];
#pragma unroll
      for (int k1 = 0; k1 < P; ++k1) v[k1] = buf[line * pitch + k1 * Q + jj];
      DFT<T, P, true>::run(v);
#pragma unroll
      for (int j1 = 0; j1 < P; ++j1) acc[j1] = acc[j1] + (c ? v[j1] * conj(tw[j1 * Q + jj]) : v[j1]);
    }
    __syncthreads();
  }
  if (jj < Q && line < nl) {
#pragma unroll
    for (int j1 = 0; j1 < P; ++j1) {
      const int x = j1 * Q + jj;
      if (x < n) buf[line * pitch + x] = acc[j1];
    }
  }
  __syncthreads();
  // twoereduced(:, col) and the mirror column (SD/coreproject.f90:521,528)
  z64* o = a.out + (size_t)l0 * n;
  z64* om = a.outm ? a.outm + (size_t)l0 * n : nullptr;
  const double* vsub = a.vsub ? a.vsub + (size_t)l0 * n : nullptr;
  RowCol rc(tid, XCfg<P, Q, XT>::XT, n);
  for (int e = tid; e < nl * n; e += XCfg<P, Q, XT>::XT, rc.next()) {
    const cplx<T> v = buf[rc.row * pitch + rc.col];
    z64 d = mk((double)v.re, (double)v.im);
    if (vsub) {  // K*g of the charge carrier, in FP64
      d.re += vsub[e];
      if (a.pack) d.im += vsub[e];
    }
    if (a.pack) {  // two diagonal pair potentials, real
      o[e] = mk(d.re, 0.0);
      om[e] = mk(d.im, 0.0);
      continue;
    }
    o[e] = d;
    if (om) {
      if (a.mirror_conj) d.im = -d.im;
      om[e] = d;
    }
  }
}

template <typename T, int P, int Q>
__global__ void __launch_bounds__(XThreads<T>::inv) k_xinv(KArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  body_xinv<T, P, Q, XThreads<T>::inv>(make_view(a, blockIdx.y, P * Q), blockIdx.x, smem_raw);
}

// ---------------------------------------------------------------- dispatch
template <typename T, int P, int Q>
static int launch_pq(const Spectrum& s, const ConvArgsHost& h, hipStream_t st, Prof* prof) {
  constexpr int R = P * Q, pitch = R | 1, M = 2 * R;
  constexpr int TF = XThreads<T>::fwd, TI = XThreads<T>::inv;
  constexpr int LAF = XCfg<P, Q, TF>::LA, LAI = XCfg<P, Q, TI>::LA;
  KArgs<T> a;
  a.n = h.n;
  a.n2 = h.n * h.n;
  a.n3 = h.n * h.n * h.n;
  a.ld = plane_stride(h.n);
  a.npairs = h.npairs;
  a.phi1 = (const z64*)h.phi1;
  a.phi2 = (const z64*)h.phi2;
  a.density_mode = h.density_mode;
  a.conj_b = h.conj_b;
  a.mirror_conj = h.mirror_conj;
  a.pairs = h.pairs_dev;
  a.X = (cplx<T>*)h.work;
  a.out = (z64*)h.out;
  a.tw = (const cplx<T>*)(sizeof(T) == 8 ? s.tw64 : s.tw32);
  a.khat = (const T*)(sizeof(T) == 8 ? (const void*)s.khat64 : (const void*)s.khat32);
  const bool carrier = sizeof(T) == 4 && !h.density_mode && h.phi1 == h.phi2 && s.gsub && s.vsub;
  a.gsub = carrier ? s.gsub : nullptr;
  a.vsub = carrier ? s.vsub : nullptr;
  const size_t shXF = (size_t)(M + LAF * pitch) * sizeof(cplx<T>);
  const size_t shXI = (size_t)(M + LAI * pitch) * sizeof(cplx<T>);
  a.lds_tw2 = M;
  a.lds_y = M + R + ((M + R) & 1);  // keep the workspaces 16-byte aligned for FP32
  const size_t shB = yz_lds_bytes(R, (int)sizeof(cplx<T>));
  SDV_REQUIRE(shB <= 160 * 1024 && shB >= (size_t)(a.lds_y + R * yz_pitch(R, (int)sizeof(cplx<T>))) * sizeof(cplx<T>), SDV_ERR_UNSUPPORTED,
              "y/z plane does not fit in LDS");
  const bool full = h.n == R;
  const void* kyz = full ? (const void*)k_yz<T, P, Q, true> : (const void*)k_yz<T, P, Q, false>;
  SDV_HIP_CHECK(hipFuncSetAttribute(kyz, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shB));
  SDV_REQUIRE(shXF <= 160 * 1024 && shXI <= 160 * 1024, SDV_ERR_UNSUPPORTED, "x lines do not fit in LDS");
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_xfwd<T, P, Q>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shXF));
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_xinv<T, P, Q>, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shXI));
  dim3 gXF(h.npairs, (a.n2 + LAF - 1) / LAF);  // pairs fastest (renumbered by XCD inside)
  dim3 gXI((a.n2 + LAI - 1) / LAI, h.npairs);
  dim3 gB(2 * R, h.npairs);
  hipEvent_t e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL((k_xfwd<T, P, Q>), gXF, dim3(TF), shXF, st, a);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_XFWD, e0, st);
  e0 = prof ? prof->begin(st) : nullptr;
  if (full)
    hipLaunchKernelGGL((k_yz<T, P, Q, true>), gB, dim3(YZCfg<T, P, Q>::NT), shB, st, a);
  else
    hipLaunchKernelGGL((k_yz<T, P, Q, false>), gB, dim3(YZCfg<T, P, Q>::NT), shB, st, a);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_YZ, e0, st);
  e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL((k_xinv<T, P, Q>), gXI, dim3(TI), shXI, st, a);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_XINV, e0, st);
  return SDV_OK;
}

// ---------------------------------------------------------- multi-set batch
// Persistent kernels over a batch of orbital sets of uneven grid sizes: a
// static work list (one entry per line chunk or plane of every pair of every
// set, sorted by cost so the round-robin stays balanced) is walked by a grid
// sized to the machine; each entry dispatches on its set's circulant size.
template <typename T>
struct BSetDev {
  const z64* phi;
  z64* out;
  const cplx<T>* tw;
  const T* khat;
  const double* gsub;  // FP32 form: charge carrier (bra = ket always holds in a batch)
  const double* vsub;
  int n, m, conj_b, mirror_conj;
};

template <typename T>
__device__ __forceinline__ ItemView<T> batch_view(const BSetDev<T>& s, const BItemDev& it, cplx<T>* X) {
  ItemView<T> v;
  v.n = s.n;
  v.n2 = s.n * s.n;
  v.n3 = s.n * s.n * s.n;
  v.ld = plane_stride(s.n);
  v.pa = s.phi + (size_t)it.a * v.n3;
  v.pb = s.phi + (size_t)it.b * v.n3;
  v.density_mode = 0;
  v.conj_b = s.conj_b;
  v.mirror_conj = s.mirror_conj;
  v.X = X + it.xoff;
  v.out = s.out + (size_t)it.col * v.n3;
  v.outm = it.mcol >= 0 ? s.out + (size_t)it.mcol * v.n3 : nullptr;
  v.tw = s.tw;
  v.khat = s.khat;
  const int R = s.m, M = 2 * R;
  v.lds_tw2 = M;
  v.lds_y = M + R + ((M + R) & 1);
  v.pack = it.pack;
  const bool sub = s.gsub && (it.a == it.b || it.pack);
  v.gsub = sub ? s.gsub : nullptr;
  v.vsub = sub ? s.vsub : nullptr;
  return v;
}

#define SDV_SIZE_SWITCH(m_, CALL) \
  switch (m_) {                   \
    case 9: CALL(3, 3); break;    \
    case 12: CALL(4, 3); break;   \
    case 16: CALL(4, 4); break;   \
    case 21: CALL(7, 3); break;   \
    case 25: CALL(5, 5); break;   \
    case 27: CALL(9, 3); break;   \
    case 32: CALL(8, 4); break;   \
    case 36: CALL(6, 6); break;   \
    case 45: CALL(9, 5); break;   \
    case 48: CALL(8, 6); break;   \
    case 54: CALL(9, 6); break;   \
    case 63: CALL(9, 7); break;   \
    case 64: CALL(8, 8); break;   \
    case 72: CALL(9, 8); break;   \
    case 81: CALL(9, 9); break;   \
    default: break;               \
  }

// Separately allocated per-size entry points: the multi-set kernels call
// these rather than inlining every size, so each body keeps its own register
// allocation and the kernel needs only the largest.
template <typename T, int P, int Q>
__device__ __noinline__ void call_xfwd(const ItemView<T>* v, int sub, unsigned char* smem) {
  body_xfwd<T, P, Q, XThreads<T>::batch>(*v, sub, smem);
}
template <typename T, int P, int Q>
__device__ __noinline__ void call_yz(const ItemView<T>* v, int sub, unsigned char* smem) {
  if (v->n == P * Q)  // the full-grid form, as the single-set and grouped paths choose it
    body_yz<T, P, Q, true, ItemView<T>, true>(*v, v->X + (size_t)sub * v->ld, sub, smem);
  else
    body_yz<T, P, Q, true, ItemView<T>, false>(*v, v->X + (size_t)sub * v->ld, sub, smem);
}
template <typename T, int P, int Q>
__device__ __noinline__ void call_xinv(const ItemView<T>* v, int sub, unsigned char* smem) {
  body_xinv<T, P, Q, XThreads<T>::batch>(*v, sub, smem);
}

template <typename T>
__global__ void __launch_bounds__(XThreads<T>::batch) k_xfwd_multi(const BSetDev<T>* __restrict__ sets,
                                                     const BItemDev* __restrict__ items,
                                                     const BWork* __restrict__ work, int nwork, cplx<T>* X) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const BWork wk = work[w];
    const BItemDev it = items[wk.item];
    const BSetDev<T> st = sets[it.set];
    const ItemView<T> v = batch_view(st, it, X);
#define SDV_CALL_XFWD(P_, Q_) call_xfwd<T, P_, Q_>(&v, wk.sub, smem_raw)
    SDV_SIZE_SWITCH(st.m, SDV_CALL_XFWD)
#undef SDV_CALL_XFWD
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(768) k_yz_multi(const BSetDev<T>* __restrict__ sets,
                                                   const BItemDev* __restrict__ items,
                                                   const BWork* __restrict__ work, int nwork, cplx<T>* X) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const BWork wk = work[w];
    const BItemDev it = items[wk.item];
    const BSetDev<T> st = sets[it.set];
    const ItemView<T> v = batch_view(st, it, X);
#define SDV_CALL_YZ(P_, Q_) call_yz<T, P_, Q_>(&v, wk.sub, smem_raw)
    SDV_SIZE_SWITCH(st.m, SDV_CALL_YZ)
#undef SDV_CALL_YZ
    __syncthreads();
  }
}

template <typename T>
__global__ void __launch_bounds__(XThreads<T>::batch) k_xinv_multi(const BSetDev<T>* __restrict__ sets,
                                                     const BItemDev* __restrict__ items,
                                                     const BWork* __restrict__ work, int nwork, cplx<T>* X) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const BWork wk = work[w];
    const BItemDev it = items[wk.item];
    const BSetDev<T> st = sets[it.set];
    const ItemView<T> v = batch_view(st, it, X);
#define SDV_CALL_XINV(P_, Q_) call_xinv<T, P_, Q_>(&v, wk.sub, smem_raw)
    SDV_SIZE_SWITCH(st.m, SDV_CALL_XINV)
#undef SDV_CALL_XINV
    __syncthreads();
  }
}

// SDV_SIZE_SWITCH is itself the switch statement; unsupported m gives 0.
static int lines_per_block(int m, int precision) {
#define SDV_LA(P_, Q_) \
  return precision ? XCfg<P_, Q_, XThreads<float>::batch>::LA : XCfg<P_, Q_, XThreads<double>::batch>::LA
  SDV_SIZE_SWITCH(m, SDV_LA)
#undef SDV_LA
  return 0;
}
static int yz_threads(int m, int precision) {
#define SDV_NT(P_, Q_) return precision ? YZCfg<float, P_, Q_>::NT : YZCfg<double, P_, Q_>::NT
  SDV_SIZE_SWITCH(m, SDV_NT)
#undef SDV_NT
  return 0;
}

// Grouped schedule: one work entry per workgroup, block shape and LDS of
// this circulant size (as the single-set kernels).
template <typename T, int P, int Q>
__global__ void __launch_bounds__(XThreads<T>::batch) k_xfwd_grp(const BSetDev<T>* __restrict__ sets,
                                                   const BItemDev* __restrict__ items,
                                                   const BWork* __restrict__ work, cplx<T>* X) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // the work list is chunk-major within a size; each XCD takes one range of
  // it, as in the single-set x forward (64-set batch: 2-3 % per step)
  const BWork wk = work[xcd_span(blockIdx.x, gridDim.x)];
  const BItemDev it = items[wk.item];
  body_xfwd<T, P, Q, XThreads<T>::batch>(batch_view(sets[it.set], it, X), wk.sub, smem_raw);
}

template <typename T, int P, int Q, bool FULL>
__global__ void __launch_bounds__((YZCfg<T, P, Q>::NT)) __attribute__((amdgpu_waves_per_eu(yz_wpe<T, P, Q>()))) k_yz_grp(const BSetDev<T>* __restrict__ sets,
                                                               const BItemDev* __restrict__ items,
                                                               const BWork* __restrict__ work, cplx<T>* X) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  // the group's list is plane-major with pairs fastest; each XCD takes one
  // range of it, as in the single-set y/z kernel
  const BWork wk = work[xcd_span(blockIdx.x, gridDim.x)];
  const BItemDev it = items[wk.item];
  const ItemView<T> v = batch_view(sets[it.set], it, X);
  body_yz<T, P, Q, false, ItemView<T>, FULL>(v, v.X + (size_t)wk.sub * v.ld, wk.sub, smem_raw);
}

template <typename T, int P, int Q>
__global__ void __launch_bounds__(XThreads<T>::batch) k_xinv_grp(const BSetDev<T>* __restrict__ sets,
                                                   const BItemDev* __restrict__ items,
                                                   const BWork* __restrict__ work, cplx<T>* X) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const BWork wk = work[blockIdx.x];
  const BItemDev it = items[wk.item];
  body_xinv<T, P, Q, XThreads<T>::batch>(batch_view(sets[it.set], it, X), wk.sub, smem_raw);
}

// LDS of the x kernels and of the y/z kernel for half-edge m (elements of
// cplx<T>).
static size_t lds_x_elems(int m, int precision) {
  return (size_t)2 * m + (size_t)lines_per_block(m, precision) * (m | 1);
}
static size_t lds_yz_elems(int m, int elem_bytes) {
  return (yz_lds_bytes(m, elem_bytes) + elem_bytes - 1) / elem_bytes;
}

BatchPlan batch_plan(const std::vector<BatchSetHost>& sets, int precision, int mode, int grid_mult, int ncu) {
  BatchPlan bp;
  bp.precision = precision;
  bp.mode = mode;
  const size_t es = precision == 0 ? sizeof(z64) : sizeof(c32);
  long long xoff = 0;
  size_t shx = 0, shy = 0;
  int nty = 64;
  // key = 2 m + (n == m): the grouped schedule launches one group per key, so
  // sets whose grid edge equals the half-edge run the full-grid y/z kernel as
  // they do in the single-set path (same per-pair arithmetic).
  struct Cost {
    long long c;
    int key;
    BWork w;
  };
  std::vector<Cost> wx, wy;
  for (int si = 0; si < (int)sets.size(); ++si) {
    const BatchSetHost& s = sets[si];
    const int n = s.spec->g.n, m = s.spec->g.m;
    const int key = 2 * m + (n == m ? 1 : 0);
    const int LA = lines_per_block(m, precision);
    if (LA <= 0 || yz_threads(m, precision) <= 0 || m >= 128) {  // every context's m comes from choose_geom
      bp.items.clear();
      bp.grid_x = bp.grid_y = 0;
      return bp;
    }
    shx = std::max(shx, lds_x_elems(m, precision) * es);
    shy = std::max(shy, lds_yz_elems(m, (int)es) * es);
    nty = std::max(nty, yz_threads(m, precision));
    for (const PairItem& p : *s.items) {
      BItemDev it;
      it.set = si;
      it.a = p.a;
      it.b = p.b;
      it.col = p.col;
      it.mcol = p.mcol;
      it.pack = p.pack;
      it.xoff = xoff;
      const int ii = (int)bp.items.size();
      bp.items.push_back(it);
      xoff += (long long)2 * m * plane_stride(n);
      const int nch = (n * n + LA - 1) / LA;
      for (int c = 0; c < nch; ++c) wx.push_back(Cost{(long long)LA * m, key, BWork{ii, c}});
      for (int pl = 0; pl < 2 * m; ++pl) wy.push_back(Cost{(long long)m * m * m, key, BWork{ii, pl}});
    }
  }
  // FUSED: largest cost first so the round-robin of the persistent grid stays
  // balanced; GROUPED: by size (largest first), each size contiguous.
  // planes: GROUPED plane-major within a size (pairs fastest), so the
  // workgroups of one XCD share the spectrum slices of a few planes
  auto order = [mode](const Cost& a, const Cost& b) {
    if (mode == BATCH_GROUPED && a.key != b.key) return a.key > b.key;
    if (a.c != b.c) return a.c > b.c;
    if (mode == BATCH_GROUPED && a.w.sub != b.w.sub) return a.w.sub < b.w.sub;
    if (a.w.item != b.w.item) return a.w.item < b.w.item;
    return a.w.sub < b.w.sub;
  };
  // line chunks: chunk-major within a size, so the workgroups resident at
  // one time read the same orbital chunks (L2 reuse across pairs)
  auto order_x = [mode](const Cost& a, const Cost& b) {
    if (mode == BATCH_GROUPED && a.key != b.key) return a.key > b.key;
    if (a.c != b.c) return a.c > b.c;
    if (a.w.sub != b.w.sub) return a.w.sub < b.w.sub;
    return a.w.item < b.w.item;
  };
  std::stable_sort(wx.begin(), wx.end(), order_x);
  std::stable_sort(wy.begin(), wy.end(), order);
  for (auto& c : wx) bp.wx.push_back(c.w);
  for (auto& c : wy) bp.wy.push_back(c.w);
  if (mode == BATCH_GROUPED) {
    for (size_t i = 0; i < wy.size();) {
      BatchGroup g;
      const int key = wy[i].key;
      g.m = key / 2;
      g.full = (key & 1) != 0;
      g.wy0 = (int)i;
      while (i < wy.size() && wy[i].key == key) ++i;
      g.nwy = (int)i - g.wy0;
      g.shx = lds_x_elems(g.m, precision) * es;
      g.shy = lds_yz_elems(g.m, (int)es) * es;
      bp.groups.push_back(g);
    }
    size_t gi = 0;
    for (size_t i = 0; i < wx.size(); ++gi) {
      const int key = wx[i].key;
      const int b0 = (int)i;
      while (i < wx.size() && wx[i].key == key) ++i;
      if (gi >= bp.groups.size() || 2 * bp.groups[gi].m + (bp.groups[gi].full ? 1 : 0) != key) {  // same keys, same order
        bp.items.clear();
        bp.grid_x = bp.grid_y = 0;
        return bp;
      }
      bp.groups[gi].wx0 = b0;
      bp.groups[gi].nwx = (int)i - b0;
    }
  }
  bp.x_elems = xoff;
  bp.x_bytes = (size_t)xoff * es;
  bp.shx = shx;
  bp.shy = shy;
  bp.nty = nty;
  bp.grid_x = std::max(1, std::min((int)bp.wx.size(), grid_mult * ncu * 4));
  bp.grid_y = std::max(1, std::min((int)bp.wy.size(), grid_mult * ncu));
  return bp;
}

template <typename T>
static int batch_launch_fused(const BatchPlan& bp, const BatchDev& d, hipStream_t st, Prof* prof) {
  SDV_REQUIRE(bp.shy <= 160 * 1024, SDV_ERR_UNSUPPORTED, "batch plane does not fit in LDS");
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_yz_multi<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)bp.shy));
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_xfwd_multi<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)bp.shx));
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_xinv_multi<T>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)bp.shx));
  const BSetDev<T>* sets = (const BSetDev<T>*)d.sets;
  hipEvent_t e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL(k_xfwd_multi<T>, dim3(bp.grid_x), dim3(XThreads<T>::batch), bp.shx, st, sets, d.items, d.wx, (int)bp.wx.size(),
                     (cplx<T>*)d.X);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_XFWD, e0, st);
  e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL(k_yz_multi<T>, dim3(bp.grid_y), dim3(bp.nty), bp.shy, st, sets, d.items, d.wy,
                     (int)bp.wy.size(), (cplx<T>*)d.X);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_YZ, e0, st);
  e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL(k_xinv_multi<T>, dim3(bp.grid_x), dim3(XThreads<T>::batch), bp.shx, st, sets, d.items, d.wx, (int)bp.wx.size(),
                     (cplx<T>*)d.X);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_XINV, e0, st);
  return SDV_OK;
}

template <typename T, int P, int Q>
static int launch_group(const BatchGroup& g, const BatchDev& d, hipStream_t st, Prof* prof) {
  SDV_REQUIRE(g.shy <= 160 * 1024, SDV_ERR_UNSUPPORTED, "batch plane does not fit in LDS");
  const void* kyz = g.full ? (const void*)k_yz_grp<T, P, Q, true> : (const void*)k_yz_grp<T, P, Q, false>;
  SDV_HIP_CHECK(hipFuncSetAttribute(kyz, hipFuncAttributeMaxDynamicSharedMemorySize, (int)g.shy));
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_xfwd_grp<T, P, Q>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)g.shx));
  SDV_HIP_CHECK(hipFuncSetAttribute((const void*)k_xinv_grp<T, P, Q>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)g.shx));
  const BSetDev<T>* sets = (const BSetDev<T>*)d.sets;
  cplx<T>* X = (cplx<T>*)d.X;
  hipEvent_t e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL((k_xfwd_grp<T, P, Q>), dim3(g.nwx), dim3(XThreads<T>::batch), g.shx, st, sets, d.items, d.wx + g.wx0, X);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_XFWD, e0, st);
  e0 = prof ? prof->begin(st) : nullptr;
  if (g.full)
    hipLaunchKernelGGL((k_yz_grp<T, P, Q, true>), dim3(g.nwy), dim3(YZCfg<T, P, Q>::NT), g.shy, st, sets, d.items,
                       d.wy + g.wy0, X);
  else
    hipLaunchKernelGGL((k_yz_grp<T, P, Q, false>), dim3(g.nwy), dim3(YZCfg<T, P, Q>::NT), g.shy, st, sets, d.items,
                       d.wy + g.wy0, X);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_YZ, e0, st);
  e0 = prof ? prof->begin(st) : nullptr;
  hipLaunchKernelGGL((k_xinv_grp<T, P, Q>), dim3(g.nwx), dim3(XThreads<T>::batch), g.shx, st, sets, d.items, d.wx + g.wx0, X);
  SDV_HIP_CHECK(hipGetLastError());
  if (prof) prof->end(SDV_STAGE_XINV, e0, st);
  return SDV_OK;
}

// Sizes run on forked streams (round robin), joined back into `st` at the
// end; under stream capture the forks become parallel graph branches.  With
// per-stage profiling on, everything stays on `st`.
template <typename T>
static int batch_launch_grouped(const BatchPlan& bp, const BatchDev& d, hipStream_t st, Prof* prof) {
  const bool fork = d.nside > 0 && bp.groups.size() > 1 && !(prof && prof->on);
  const int ns = fork ? std::min(d.nside, (int)bp.groups.size()) : 0;
  if (fork) {
    SDV_HIP_CHECK(hipEventRecord(d.ev[0], st));
    for (int i = 0; i < ns; ++i) SDV_HIP_CHECK(hipStreamWaitEvent(d.side[i], d.ev[0], 0));
  }
  for (size_t gi = 0; gi < bp.groups.size(); ++gi) {
    const BatchGroup& g = bp.groups[gi];
    hipStream_t s = fork ? d.side[gi % ns] : st;
    int rc = SDV_ERR_UNSUPPORTED;
#define SDV_CALL_GRP(P_, Q_) rc = launch_group<T, P_, Q_>(g, d, s, fork ? nullptr : prof)
    SDV_SIZE_SWITCH(g.m, SDV_CALL_GRP)
#undef SDV_CALL_GRP
    if (rc != SDV_OK) return rc;
  }
  if (fork) {
    for (int i = 0; i < ns; ++i) {
      SDV_HIP_CHECK(hipEventRecord(d.ev[1 + i], d.side[i]));
      SDV_HIP_CHECK(hipStreamWaitEvent(st, d.ev[1 + i], 0));
    }
  }
  return SDV_OK;
}

int batch_launch(const BatchPlan& bp, const BatchDev& d, hipStream_t st, Prof* prof) {
  if (bp.mode == BATCH_GROUPED)
    return bp.precision == 0 ? batch_launch_grouped<double>(bp, d, st, prof)
                             : batch_launch_grouped<float>(bp, d, st, prof);
  return bp.precision == 0 ? batch_launch_fused<double>(bp, d, st, prof) : batch_launch_fused<float>(bp, d, st, prof);
}

size_t batch_set_bytes(int precision) { return precision == 0 ? sizeof(BSetDev<double>) : sizeof(BSetDev<float>); }

void batch_fill_set(int precision, const BatchSetHost& s, void* dst) {
  const int cn = s.cnorm;
  if (precision == 0) {
    BSetDev<double> b;
    b.phi = (const z64*)s.phi;
    b.out = (z64*)s.out;
    b.tw = (const z64*)s.spec->tw64;
    b.khat = s.spec->khat64;
    b.gsub = b.vsub = nullptr;
    b.n = s.spec->g.n;
    b.m = s.spec->g.m;
    b.conj_b = cn ? 0 : 1;
    b.mirror_conj = cn ? 0 : 1;
    std::memcpy(dst, &b, sizeof b);
  } else {
    BSetDev<float> b;
    b.phi = (const z64*)s.phi;
    b.out = (z64*)s.out;
    b.tw = (const c32*)s.spec->tw32;
    b.khat = s.spec->khat32;
    b.gsub = s.spec->gsub;
    b.vsub = s.spec->vsub;
    b.n = s.spec->g.n;
    b.m = s.spec->g.m;
    b.conj_b = cn ? 0 : 1;
    b.mirror_conj = cn ? 0 : 1;
    std::memcpy(dst, &b, sizeof b);
  }
}
size_t batch_item_bytes() { return sizeof(BItemDev); }
size_t batch_work_bytes() { return sizeof(BWork); }

// Supported half-edges m = P*Q (smallest m >= n is chosen).
struct PQ {
  int m, P, Q;
};
static const PQ kSizes[] = {{9, 3, 3},  {12, 4, 3}, {16, 4, 4}, {21, 7, 3}, {25, 5, 5},
                            {27, 9, 3}, {32, 8, 4}, {36, 6, 6}, {45, 9, 5}, {48, 8, 6},
                            {54, 9, 6}, {63, 9, 7}, {64, 8, 8}, {72, 9, 8}, {81, 9, 9}};

bool choose_geom(int n, Geom* g, int min_m) {
  for (const PQ& p : kSizes)
    if (p.m >= n && p.m >= min_m) {
      g->n = n;
      g->m = p.m;
      g->M = 2 * p.m;
      g->P = p.P;
      g->Q = p.Q;
      return true;
    }
  return false;
}

int circulant_sizes(int* out, int cap) {
  const int k = (int)(sizeof(kSizes) / sizeof(kSizes[0]));
  for (int i = 0; i < k && i < cap; ++i) out[i] = kSizes[i].m;
  return k;
}

size_t conv_work_bytes(const Geom& g, int npairs, int precision) {
  const size_t es = precision == 0 ? sizeof(z64) : sizeof(c32);
  return (size_t)npairs * 2 * g.m * (size_t)plane_stride(g.n) * es;
}

template <typename T>
static int conv_dispatch(const Spectrum& s, const ConvArgsHost& h, hipStream_t st, Prof* prof) {
  switch (s.g.m) {
#define SDV_CASE(m_, P_, Q_) \
  case m_:                   \
    return launch_pq<T, P_, Q_>(s, h, st, prof);
    SDV_CASE(9, 3, 3)
    SDV_CASE(12, 4, 3)
    SDV_CASE(16, 4, 4)
    SDV_CASE(21, 7, 3)
    SDV_CASE(25, 5, 5)
    SDV_CASE(27, 9, 3)
    SDV_CASE(32, 8, 4)
    SDV_CASE(36, 6, 6)
    SDV_CASE(45, 9, 5)
    SDV_CASE(48, 8, 6)
    SDV_CASE(54, 9, 6)
    SDV_CASE(63, 9, 7)
    SDV_CASE(64, 8, 8)
    SDV_CASE(72, 9, 8)
    SDV_CASE(81, 9, 9)
#undef SDV_CASE
    default:
      set_error("unsupported circulant half-edge");
      return SDV_ERR_UNSUPPORTED;
  }
}

int conv_launch(const Spectrum& s, const ConvArgsHost& h, int precision, hipStream_t st, Prof* prof) {
  if (h.npairs <= 0) return SDV_OK;
  return precision == 0 ? conv_dispatch<double>(s, h, st, prof) : conv_dispatch<float>(s, h, st, prof);
}

}  // namespace sdv
