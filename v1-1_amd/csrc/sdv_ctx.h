// sdv_ctx.h -- the context behind the C-ABI handle (sdv_ctx) and the host
// helpers the entry points share (index plans, growable buffers, shared
// kernel spectra).  Internal: nothing here crosses the C-ABI.
#pragma once
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "sdv_internal.h"

namespace sdv {

// Index lists of one call shape, uploaded once and kept on the device: the
// convolution items, and the row/column maps of the matrix-element assembly.
struct IdxPlan {
  std::vector<PairItem> items;
  std::vector<int2> rows, rmap, cmap;
  std::vector<int> cols;
  void* dev = nullptr;
  size_t off_items = 0, off_rows = 0, off_cols = 0, off_rmap = 0, off_cmap = 0;
};

}  // namespace sdv

struct sdv_ctx {
  sdv_config cfg;
  sdv::Prof prof;
  std::map<std::tuple<int, int, int, int>, sdv::IdxPlan> plans;  // (kind, low, high, sym)
  int n = 0, n3 = 0, N = 0;
  int precision = 0;
  int pair_batch = 0;                      // pairs per pipeline launch; 0: as many as the work budget holds
  size_t work_budget = (size_t)8 << 30;    // bytes of convolution work buffer for pair_batch = 0
  int auto_m = 0;      // half-edge choose_geom picks for n (the default circulant)
  sdv::Spectrum* spec = nullptr;  // shared with every context of the same grid (spectrum cache)
  double* ke = nullptr;
  double* fd = nullptr;        // first-derivative stencil, negated (velocity gauge)
  sdv::z64* invj = nullptr;
  sdv::z64* invsqj = nullptr;  // sqrt(1/J) per axis (ECS; velocity gauge)
  sdv::z64* sdiag = nullptr;
  sdv::z64* dip = nullptr;  // dipole (coordinate, complex-scaled with ECS) per grid index, same on the three axes
  double* mask = nullptr;  // edge mask (n values, same on the three axes) when maskflag and not scalingflag
  hipStream_t stream = nullptr;
  // growable device buffers
  struct Buf {
    void* p = nullptr;
    size_t cap = 0;
  };
  Buf work, scratch, gout, h_in1, h_in2, h_out1, h_out2, ob_ke, ob_scr, h_pot;
  Buf ar_w, ar_rp, ar_s, ar_s2, ar_dots, h_small, h_out3;  // orbital operator (actreduced00)
  Buf rp_bt;  // get_reducedpot: twoden transposed
};

namespace sdv {

inline int ensure(sdv_ctx::Buf& b, size_t bytes) {
  if (b.cap >= bytes) return SDV_OK;
  if (b.p) hipFree(b.p);
  b.p = nullptr;
  b.cap = 0;
  if (hipMalloc(&b.p, bytes) != hipSuccess) {
    b.p = nullptr;
    set_error("device allocation of " + std::to_string(bytes) + " bytes failed");
    return SDV_ERR_NOMEM;
  }
  b.cap = bytes;
  return SDV_OK;
}

#define SDV_TRY(expr)          \
  do {                         \
    int _rc = (expr);          \
    if (_rc != SDV_OK) return _rc; \
  } while (0)

// Pair bookkeeping of call_twoe_matelxxx00 for (lowspf, highspf, symmetric),
// built and uploaded once per shape.
int twoe_plan(sdv_ctx* c, int low, int high, bool sym, IdxPlan** out);

// Kernel spectra are shared by every context on the same grid (edge,
// spacing, tooth parameters, circulant half-edge): acquired at sdv_init /
// sdv_set_circulant, released at sdv_free; built once, kept resident.
int spectrum_acquire(int n, double spacing, int nbig, int nsmall, int min_m, hipStream_t st, Spectrum** out);
void spectrum_retain(Spectrum* s);   // one more reference on a cached spectrum (batches hold one per set)
void spectrum_release(Spectrum* s);

}  // namespace sdv
