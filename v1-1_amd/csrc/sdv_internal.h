// sdv_internal.h -- internal (C++) interfaces between the library's
// translation units.  Nothing here crosses the C-ABI.
#pragma once
#include <vector>
#include <utility>

#include "sdv_common.h"

namespace sdv {

// Per-kernel HIP-event timing on the launch stream (enabled by the caller
// through sdv_profile).  Stage ids are the SDV_STAGE_* values of sincdvr.h.
struct Prof {
  bool on = false;
  struct Rec {
    int stage;
    hipEvent_t a, b;
  };
  std::vector<Rec> recs;
  hipEvent_t begin(hipStream_t st) {
    hipEvent_t e = nullptr;
    if (on && hipEventCreate(&e) == hipSuccess) (void)hipEventRecord(e, st);
    return e;
  }
  void end(int stage, hipEvent_t a, hipStream_t st) {
    if (!on || !a) return;
    hipEvent_t b = nullptr;
    if (hipEventCreate(&b) != hipSuccess) return;
    (void)hipEventRecord(b, st);
    recs.push_back(Rec{stage, a, b});
  }
};

// One convolution work item: output column `col` of the (n^3, ncols) result
// gets K * rho where rho = phi1[:,a] * CONJUGATE(phi2[:,b]) (orbital mode) or
// rho = den[:,a] (density mode).  mcol >= 0 also receives the mirror pair
// (b,a): conj(V) for hermnorm, V for cnorm.
// pack = 1 (hermnorm, bra = ket): two diagonal pairs (a,a) and (b,b), whose
// densities |phi|^2 are real, in one convolution of rho_aa + i rho_bb; column
// col receives Re V, column mcol Im V.
struct PairItem {
  int a, b, col, mcol;
  int pack = 0;
};

// Geometry of one circulant: grid edge n, half edge m = P*Q >= n, M = 2m.
struct Geom {
  int n = 0, m = 0, M = 0, P = 0, Q = 0;
};
// Smallest supported half-edge m >= max(n, min_m); false if none.
bool choose_geom(int n, Geom* g, int min_m = 0);
int circulant_sizes(int* out, int cap);

// Device-resident tables for one geometry.
struct Spectrum {
  Geom g;
  void* tw64 = nullptr;     // M x z64   exp(-2 pi i k / M)
  void* tw32 = nullptr;     // M x c32
  double* khat64 = nullptr; // (m+1)^3, [kkx][kky][kkz], includes 1/M^3
  float* khat32 = nullptr;
  double* koct = nullptr;   // n^3 octant of the real-space kernel K(i,j,k), i,j,k >= 0
  // Charge carrier of the FP32 form: a unit-charge Gaussian g on the grid
  // (n^3, real) and its potential K*g computed once in FP64.  The FP32
  // pipeline convolves rho_aa - g and adds K*g back in FP64, so the large
  // low-frequency part of a diagonal pair density never passes through FP32
  // butterflies (the linear convolution is unchanged).
  double* gsub = nullptr;
  double* vsub = nullptr;
};

struct ConvArgsHost {
  int n;
  int npairs;
  const void* phi1;   // z64 (n^3, N) or density (n^3, ncols)
  const void* phi2;
  int density_mode;   // 1: phi1 columns are densities
  int conj_b;         // 1: conjugate the second factor (hermnorm)
  int mirror_conj;    // 1: mirror columns receive conj(V)
  const PairItem* pairs_dev;
  void* work;         // intermediate, npairs * 2m planes of plane_stride(n) complex<T>
  void* out;          // z64 (n^3, ncols)
};

// Elements between consecutive x-frequency planes of the convolution work
// buffer: the n^2 lines of a plane rounded up to 16, so every plane starts on
// a 256-byte boundary and the x kernels' runs of consecutive lines (56 FP64
// lines = 7 cache lines at m = 81) do not straddle an extra cache line.
__host__ __device__ inline int plane_stride(int n) { return (n * n + 15) & ~15; }

// Compute units of the current device (cached per device; 256 if unknown).
inline int device_cu_count() {
  static int cache[64] = {0};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
  if (!cache[dev]) {
    int c = 0;
    cache[dev] = hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && c > 0 ? c : 256;
  }
  return cache[dev];
}

// Convolution pipeline: x-forward (pair density fused), y/z plane kernel with
// the spectral multiply, x-inverse with output scatter.  precision 0 = FP64,
// 1 = FP32 butterflies.
int conv_launch(const Spectrum& s, const ConvArgsHost& a, int precision, hipStream_t st, Prof* prof);
size_t conv_work_bytes(const Geom& g, int npairs, int precision);

// Multi-set batch (persistent kernels over uneven grid sizes).
struct BatchSetHost {
  const Spectrum* spec;
  const std::vector<PairItem>* items;
  const void* phi;  // z64 (n^3, N): bra and ket orbitals (same array)
  void* out;        // z64 twoereduced (n^3, N, N)
  int cnorm;
};
struct BItemDev {
  int set, a, b, col, mcol, pack;
  long long xoff;  // element offset of this pair's planes in the batch buffer
};
struct BWork {
  int item, sub;  // line chunk (x kernels) or kx plane (y/z kernel)
};
// Work of one circulant size in the grouped schedule: ranges of the
// size-sorted work lists and the LDS its kernels need.
struct BatchGroup {
  int m = 0;
  int wx0 = 0, nwx = 0, wy0 = 0, nwy = 0;
  size_t shx = 0, shy = 0;
  bool full = false;  // every set of the group has grid edge n = m (the y/z kernel's n guards drop)
};
// Schedules: FUSED = one persistent launch per stage over every set (block
// shape and LDS of the largest size); GROUPED = one launch per stage per
// circulant size, sizes on forked streams (branches of the captured graph).
enum BatchMode { BATCH_FUSED = 0, BATCH_GROUPED = 1 };
struct BatchPlan {
  int precision = 0;
  int mode = BATCH_GROUPED;
  std::vector<BItemDev> items;
  std::vector<BWork> wx, wy;
  std::vector<BatchGroup> groups;
  long long x_elems = 0;
  size_t x_bytes = 0, shx = 0, shy = 0;
  int nty = 64, grid_x = 1, grid_y = 1;
};
struct BatchDev {
  const void* sets;
  const BItemDev* items;
  const BWork* wx;
  const BWork* wy;
  void* X;
  int nside = 0;              // forked streams available to the grouped schedule
  hipStream_t* side = nullptr;
  hipEvent_t* ev = nullptr;   // 1 + nside events
};
BatchPlan batch_plan(const std::vector<BatchSetHost>& sets, int precision, int mode, int grid_mult, int ncu);
int batch_launch(const BatchPlan& bp, const BatchDev& d, hipStream_t st, Prof* prof);
size_t batch_set_bytes(int precision);
void batch_fill_set(int precision, const BatchSetHost& s, void* dst);
size_t batch_item_bytes();
size_t batch_work_bytes();

// Kernel construction (tooth + placement + spectrum), FP64.
int build_spectrum(int n, double spacing, int nbig, int nsmall, Spectrum* s, hipStream_t st, int min_m = 0);
void free_spectrum(Spectrum* s);
int nuclear_pot_launch(const Spectrum& s, double spacing, int ncent, const int* cs_dev, const double* z_dev,
                       int realkind4, double* pot_dev, hipStream_t st);

// twoematel reduction: G = sum_x D_row(x) V_col(x)
struct MatelArgsHost {
  int n3, N;
  int nrows;                 // number of row pairs
  const int2* rows_dev;      // (1a, 1b) 0-based, rho = phi1[1a] * CONJ(phi2[1b])
  int ncols;
  const int* cols_dev;       // twoereduced column index of each computed V
  const void* phi1;
  const void* phi2;
  int conj_b;
  const void* tr;            // z64 twoereduced base
  int symmetric;             // rows and cols are the same pair list
  double* gout;              // (2*nrows) x (2*ncols) real products, row-major
  void* scratch;             // matel_scratch_bytes()
  size_t scratch_bytes;
};
size_t matel_scratch_bytes(int n3, int nrows, int ncols, int symmetric);
int matel_launch(const MatelArgsHost& a, hipStream_t st, Prof* prof);

// Kinetic apply
struct KeArgsHost {
  int n, howmany;
  const double* ke;          // 2n-1 stencil (device)
  const void* invj;          // 3 x n z64 per-axis 1/J (device) or null
  const void* sdiag;         // 3 x n z64 per-axis scalediag terms (device) or null
  const double* mask = nullptr;  // n edge-mask values (device), or null: maskflag without ECS
  const void* in;            // z64 (n^3, howmany)
  void* out;
  // mult_general option 2 (first derivative, velocity gauge): axis d's result
  // times coef[d] (an axis with coef 0 is skipped), added into `out`.
  z64 coef[3] = {mk(1.0, 0.0), mk(1.0, 0.0), mk(1.0, 0.0)};
  int use_coef = 0;
  int accumulate = 0;        // 1: every axis adds into `out` (0: axis 0 overwrites it)
};
int ke_launch(const KeArgsHost& a, int precision, hipStream_t st, Prof* prof);

// One-electron matrix elements: out(a + ld b) = sum_x f(A(x, a)) w(x) B(x, b)
// (f = conj if conj_a; w = 1 if null); scratch of dot_scratch_bytes.
size_t dot_scratch_bytes(int n3, int na, int nb);
int dot_launch(int n3, int na, int nb, const void* A, const void* B, const double* w, int conj_a, void* scratch,
               void* out, int ld, hipStream_t st);

// Orbital operator pieces: out = alpha A C + beta out (A: n3 x nin device
// columns, C: nin x nout device, column-major ldc); s = ke + pot w + pulse
// (tdpot: 3 complex field values x, y, z on the host, or null) + rp.
int orb_gemm_launch(int n3, int nin, int nout, const void* A, const void* C, int ldc, z64 alpha, z64 beta, void* out,
                    hipStream_t st);
int orb_sum_launch(int n, int ncols, const void* ke, const double* pot, const void* w, const void* rp,
                   const void* dip, const double* tdpot, void* s, hipStream_t st);

// Mean field
size_t reducedpot_scratch_bytes(int N);  // twoden transposed, N^4 complex
int reducedpot_launch(int n3, int N, const void* tr, const void* twoden, void* out, void* scratch, hipStream_t st);
int mult_reducedpot_launch(int n3, int N, int nout, const void* inspfs, const void* rp, void* out,
                           hipStream_t st);

}  // namespace sdv
