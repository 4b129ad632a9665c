// sdv_abi.hip -- C-ABI entry points and host-side orchestration.
//
// Mirrors the SINCDVR project interface of the reference (op_tinv,
// call_twoe_matel00, mult_ke, mult_reducedpot, get_reducedpot0's GEMM) on
// device-resident data.  No computation of the operators happens on the host:
// the host builds index lists (pair bookkeeping) and launches kernels.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include <mutex>

#include "sdv_ctx.h"

namespace sdv {

static thread_local std::string g_err;  // per host thread, like errno
void set_error(const std::string& msg) { g_err = msg; }
const char* last_error() { return g_err.c_str(); }

// Per-axis tables: KE stencil (SD/SineDVR.f90:22-33), ECS 1/J and the
// per-axis scalediag terms (SD/init_sinc_new.f90:5-166, 493-521).
__device__ double fscaled_d(double x) {
  const double s = sin(kPiD * x);
  return (-1.0) / 12.0 / kPiD * s * s * s - 1.0 / 2.0 / kPiD * s + x / 2.0;
}
__device__ double jscaled_d(double x) {
  const double c = cos(kPiD * x);
  return 1.0 / 4.0 * c * c * c - 3.0 / 4.0 * c + 0.5;
}
__device__ double djscaled_d(double x) {
  const double s = sin(kPiD * x);
  return 0.75 * kPiD * s * s * s;
}
__device__ double ddjscaled_d(double x) {
  const double s = sin(kPiD * x), c = cos(kPiD * x);
  return 9.0 / 4.0 * kPiD * kPiD * s * s * c;
}

// dip(i) = dipole (coordinate) of grid index i on one axis: the grid point,
// or the complex-scaled coordinate F(x) with ECS (get_dipoles,
// SD/init_sinc_new.f90:527,619-645; F of :5-30).
// Principal square root (branch cut on the negative real axis, as Fortran's
// complex SQRT), by the half-angle formulas written to avoid cancellation.
__device__ z64 csqrt_principal(z64 z) {
  const double r = hypot(z.re, z.im);
  if (r == 0.0) return mk(0.0, 0.0);
  if (z.re >= 0.0) {
    const double t = sqrt(0.5 * (r + z.re));
    return mk(t, z.im / (2.0 * t));
  }
  const double t = sqrt(0.5 * (r - z.re));
  return mk(fabs(z.im) / (2.0 * t), copysign(t, z.im));
}

__global__ void k_init_tables(int n, double spacing, int ecs, double theta, double stretch, double R0, double S,
                              double* ke, double* fd, z64* invj, z64* invsqj, z64* sdiag, z64* dip) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < 2 * n - 1) {
    const int i = t - (n - 1);
    double k;
    if (i == 0)
      k = kPiD * kPiD / 3.0;
    else
      k = 2.0 / ((double)i * (double)i) * ((i & 1) ? -1.0 : 1.0);
    ke[t] = k / 2.0 / (spacing * spacing);
    // first derivative fdtot(i) = (-1)^i / i / spacing, 0 at i = 0 (SD/SineDVR.f90:25-33),
    // stored negated: mult_all0 option 2 contracts with alpha = -1 (SD/coreproject.f90:2089-2092)
    const double f = i == 0 ? 0.0 : 1.0 / (double)i * ((i & 1) ? -1.0 : 1.0);
    fd[t] = -(f / spacing);
  }
  if (!ecs && t < n) {
    const double p1 = 1 * spacing, pn = n * spacing;
    dip[t] = mk((t + 1) * spacing - (p1 + pn) / 2.0, 0.0);
  }
  if (ecs && t < n) {
    const double p1 = 1 * spacing, pn = n * spacing;
    const double x = (t + 1) * spacing - (p1 + pn) / 2.0;
    const z64 fac = mk(cos(theta) * stretch, sin(theta) * stretch);
    const z64 fm1 = mk(fac.re - 1.0, fac.im);
    z64 j, dj, ddj;
    const z64 zero = mk(0.0, 0.0);
    z64 f;
    if (x <= -(R0 + S)) {
      f = mk(x, 0.0) + fm1 * mk(x + R0 + S / 2.0, 0.0);
      j = fac;
      dj = zero;
      ddj = zero;
    } else if (x >= R0 + S) {
      f = mk(x, 0.0) + fm1 * mk(x - R0 - S / 2.0, 0.0);
      j = fac;
      dj = zero;
      ddj = zero;
    } else if (fabs(x) <= R0) {
      f = mk(x, 0.0);
      j = mk(1.0, 0.0);
      dj = zero;
      ddj = zero;
    } else if (x >= R0) {
      const double xx = (x - R0) / S;
      f = mk(x, 0.0) + scale(fm1, S * fscaled_d(xx));
      j = mk(1.0, 0.0) + scale(fm1, jscaled_d(xx));
      dj = scale(fm1, djscaled_d(xx) / S);
      ddj = scale(fm1, ddjscaled_d(xx) / (S * S));
    } else {
      const double xx = (-x - R0) / S;
      f = mk(x, 0.0) - scale(fm1, S * fscaled_d(xx));
      j = mk(1.0, 0.0) + scale(fm1, jscaled_d(xx));
      dj = scale(mk(-fm1.re, -fm1.im), djscaled_d(xx) / S);
      ddj = scale(fm1, ddjscaled_d(xx) / (S * S));
    }
    dip[t] = f;
    const double d = j.re * j.re + j.im * j.im;
    const z64 a = mk(j.re / d, -j.im / d);  // 1/J
    const z64 a2 = a * a, a3 = a2 * a, a4 = a2 * a2;
    const z64 t1 = scale(a4 * (dj * dj), 3.0 / 8.0);
    const z64 t2 = scale(a3 * ddj, 1.0 / 4.0);
    const z64 sa = csqrt_principal(a);  // invsqrtjacobian = sqrt(invjacobian)   SD/init_sinc_new.f90:504
    for (int d3 = 0; d3 < 3; ++d3) {
      invj[d3 * n + t] = a;
      invsqj[d3 * n + t] = sa;
      sdiag[d3 * n + t] = t1 - t2;
    }
  }
}

// twoematel(1a,1b,2a,2b) from the four real products of (row r, col c).
__global__ void k_assemble(int N, int nnn, const int2* __restrict__ rmap, const int2* __restrict__ cmap,
                           const double* __restrict__ g, int C2, z64* __restrict__ out) {
  const long idx = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long tot = (long)N * N * N * nnn;
  if (idx >= tot) return;
  const int rr = (int)(idx % ((long)N * N));     // (1a, 1b)
  const int cc = (int)(idx / ((long)N * N));     // (2a, 2b-low)
  const int2 rm = rmap[rr], cm = cmap[cc];
  const double* gr = g + (size_t)(2 * rm.x) * C2 + 2 * cm.x;
  const double RR = gr[0], RI = gr[1], IR = gr[C2], II = gr[C2 + 1];
  z64 v;
  if (!rm.y && !cm.y)
    v = mk(RR - II, RI + IR);          // sum rho V
  else if (rm.y && !cm.y)
    v = mk(RR + II, RI - IR);          // sum conj(rho) V
  else if (!rm.y && cm.y)
    v = mk(RR + II, IR - RI);          // sum rho conj(V)
  else
    v = mk(RR - II, -(RI + IR));       // sum conj(rho) conj(V)
  out[idx] = v;
}

}  // namespace sdv

using namespace sdv;

// ---- shared kernel spectra
namespace {
using SpecKey = std::tuple<int, int, long long, int, int, int>;  // device, n, spacing bits, nbig, nsmall, m
struct SpecEntry {
  sdv::Spectrum* s;
  int refs;
};
std::mutex g_spec_mu;
std::map<SpecKey, SpecEntry> g_specs;
}  // namespace

int sdv::spectrum_acquire(int n, double spacing, int nbig, int nsmall, int min_m, hipStream_t st, Spectrum** out) {
  Geom g;
  SDV_REQUIRE(choose_geom(n, &g, min_m), SDV_ERR_UNSUPPORTED, "grid edge beyond the supported circulant sizes");
  int dev = 0;
  SDV_HIP_CHECK(hipGetDevice(&dev));
  long long bits = 0;
  std::memcpy(&bits, &spacing, sizeof bits);
  const SpecKey key(dev, n, bits, nbig, nsmall, g.m);
  std::lock_guard<std::mutex> lk(g_spec_mu);
  auto it = g_specs.find(key);
  if (it != g_specs.end()) {
    it->second.refs++;
    *out = it->second.s;
    return SDV_OK;
  }
  Spectrum* s = new Spectrum();
  int rc = build_spectrum(n, spacing, nbig, nsmall, s, st, min_m);
  if (rc == SDV_OK && hipStreamSynchronize(st) != hipSuccess) {  // other streams and contexts use it next
    set_error("spectrum construction failed");
    rc = SDV_ERR_HIP;
  }
  if (rc != SDV_OK) {
    free_spectrum(s);
    delete s;
    return rc;
  }
  g_specs.emplace(key, SpecEntry{s, 1});
  *out = s;
  return SDV_OK;
}

void sdv::spectrum_retain(Spectrum* s) {
  std::lock_guard<std::mutex> lk(g_spec_mu);
  for (auto& kv : g_specs)
    if (kv.second.s == s) {
      kv.second.refs++;
      return;
    }
}

void sdv::spectrum_release(Spectrum* s) {
  if (!s) return;
  std::lock_guard<std::mutex> lk(g_spec_mu);
  for (auto it = g_specs.begin(); it != g_specs.end(); ++it)
    if (it->second.s == s) {
      if (--it->second.refs == 0) {
        free_spectrum(s);
        delete s;
        g_specs.erase(it);
      }
      return;
    }
}

static int scratch_ptr(void** out);

extern "C" {

void sdv_config_default(sdv_config* c) {
  std::memset(c, 0, sizeof(*c));
  c->numpoints = 15;
  c->spacing = 0.25;
  c->toothnbig = 240;
  c->toothnsmall = 40;
  c->numspf = 1;
  c->scalingstretch = 1.0;
  c->scalingdistance = 10000.0;
  c->smoothness = 5.0;
}

const char* sdv_last_error(void) { return last_error(); }

int sdv_init(const sdv_config* cfg, sdv_ctx** out) {
  SDV_REQUIRE(cfg && out, SDV_ERR_ARG, "null argument");
  SDV_REQUIRE(cfg->numpoints >= 2, SDV_ERR_ARG, "numpoints must be >= 2");
  SDV_REQUIRE(cfg->spacing > 0, SDV_ERR_ARG, "spacing must be positive");
  SDV_REQUIRE(cfg->numspf >= 1, SDV_ERR_ARG, "numspf must be >= 1");
  SDV_REQUIRE(cfg->precision == 0 || cfg->precision == 1, SDV_ERR_ARG, "precision must be 0 or 1");
  SDV_REQUIRE(!cfg->maskflag || (cfg->masknumpoints >= 0 && cfg->masknumpoints <= cfg->numpoints), SDV_ERR_ARG,
              "masknumpoints not allowed (SD/sincproject.f90:55-60)");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
    set_error("no HIP device visible");
    return SDV_ERR_NODEVICE;
  }
  int dev = 0;
  SDV_HIP_CHECK(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  SDV_HIP_CHECK(hipGetDeviceProperties(&prop, dev));
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) {
    set_error(std::string("library is built for gfx950, device is ") + prop.gcnArchName);
    return SDV_ERR_NODEVICE;
  }
  {
    void* sp = nullptr;  // per-device address of the allocation-free entry point's scratch
    SDV_TRY(scratch_ptr(&sp));
  }
  sdv_ctx* c = new sdv_ctx();
  c->cfg = *cfg;
  c->n = cfg->numpoints;
  c->n3 = c->n * c->n * c->n;
  c->N = cfg->numspf;
  c->precision = cfg->precision;
  if (const char* pb = std::getenv("SDV_PAIR_BATCH")) c->pair_batch = std::max(1, std::atoi(pb));
  int rc = SDV_OK;
  if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
    set_error("hipStreamCreate failed");
    delete c;
    return SDV_ERR_HIP;
  }
  rc = spectrum_acquire(c->n, cfg->spacing, cfg->toothnbig, cfg->toothnsmall, 0, c->stream, &c->spec);
  if (rc == SDV_OK) c->auto_m = c->spec->g.m;
  if (rc == SDV_OK) {
    const int n = c->n;
    if (hipMalloc(&c->ke, (2 * n) * sizeof(double)) != hipSuccess) rc = SDV_ERR_NOMEM;
    if (rc == SDV_OK && hipMalloc(&c->fd, (2 * n) * sizeof(double)) != hipSuccess) rc = SDV_ERR_NOMEM;
    if (rc == SDV_OK && cfg->scalingflag) {
      if (hipMalloc(&c->invj, 3 * n * sizeof(z64)) != hipSuccess) rc = SDV_ERR_NOMEM;
      if (rc == SDV_OK && hipMalloc(&c->invsqj, 3 * n * sizeof(z64)) != hipSuccess) rc = SDV_ERR_NOMEM;
      if (rc == SDV_OK && hipMalloc(&c->sdiag, 3 * n * sizeof(z64)) != hipSuccess) rc = SDV_ERR_NOMEM;
    }
    if (rc == SDV_OK && hipMalloc(&c->dip, n * sizeof(z64)) != hipSuccess) rc = SDV_ERR_NOMEM;
    if (rc == SDV_OK && cfg->maskflag && !cfg->scalingflag) {  // maskfunction, SD/init_sinc_new.f90:580-599
      std::vector<double> m(n, 1.0);
      const int k = cfg->masknumpoints;
      for (int ii = 1; ii <= k; ++ii) {
        const double rsum = (1.0 + std::cos(kPiD * ii / (k + 1))) / 2.0;
        m[(k + 1 - ii) - 1] *= rsum;
        m[(n - k + ii) - 1] *= rsum;
      }
      if (hipMalloc(&c->mask, n * sizeof(double)) != hipSuccess ||
          hipMemcpy(c->mask, m.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        rc = SDV_ERR_NOMEM;
    }
    if (rc == SDV_OK) {
      hipLaunchKernelGGL(k_init_tables, dim3((2 * n + 255) / 256), dim3(256), 0, c->stream, n, cfg->spacing,
                         cfg->scalingflag ? 1 : 0, cfg->scalingtheta, cfg->scalingstretch, cfg->scalingdistance,
                         cfg->smoothness, c->ke, c->fd, c->invj, c->invsqj, c->sdiag, c->dip);
      if (hipStreamSynchronize(c->stream) != hipSuccess) rc = SDV_ERR_HIP;
    }
    if (rc == SDV_ERR_NOMEM) set_error("device allocation failed in sdv_init");
  }
  if (rc != SDV_OK) {
    sdv_free(c);
    return rc;
  }
  *out = c;
  return SDV_OK;
}

int sdv_free(sdv_ctx* c) {
  if (!c) return SDV_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  spectrum_release(c->spec);
  hipFree(c->ke);
  hipFree(c->fd);
  hipFree(c->invj);
  hipFree(c->invsqj);
  hipFree(c->sdiag);
  hipFree(c->mask);
  hipFree(c->dip);
  for (sdv_ctx::Buf* b : {&c->work, &c->scratch, &c->gout, &c->h_in1, &c->h_in2, &c->h_out1, &c->h_out2, &c->ob_ke,
                          &c->ob_scr, &c->h_pot, &c->ar_w, &c->ar_rp, &c->ar_s, &c->ar_s2, &c->ar_dots,
                          &c->h_small, &c->h_out3, &c->rp_bt})
    hipFree(b->p);
  for (auto& kv : c->plans) hipFree(kv.second.dev);
  for (auto& r : c->prof.recs) {
    hipEventDestroy(r.a);
    hipEventDestroy(r.b);
  }
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return SDV_OK;
}

int sdv_set_precision(sdv_ctx* c, int p) {
  SDV_REQUIRE(c, SDV_ERR_ARG, "null context");
  SDV_REQUIRE(p == 0 || p == 1, SDV_ERR_ARG, "precision must be 0 or 1");
  c->precision = p;
  return SDV_OK;
}

int sdv_circulant_edge(sdv_ctx* c) { return c && c->spec ? c->spec->g.M : SDV_ERR_ARG; }

int sdv_circulant_sizes(int* half_edges, int cap) {
  SDV_REQUIRE(cap >= 0 && (half_edges || cap == 0), SDV_ERR_ARG, "bad argument");
  return circulant_sizes(half_edges, cap);
}

int sdv_set_circulant(sdv_ctx* c, int half_edge) {
  SDV_REQUIRE(c, SDV_ERR_ARG, "null context");
  if (half_edge != 0) {
    Geom g;
    SDV_REQUIRE(half_edge >= c->n && choose_geom(c->n, &g, half_edge) && g.m == half_edge, SDV_ERR_ARG,
                "half edge must be a supported size >= numpoints");
  }
  if (half_edge == c->spec->g.m || (half_edge == 0 && c->spec->g.m == c->auto_m)) return SDV_OK;
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  SDV_HIP_CHECK(hipDeviceSynchronize());
  Spectrum* ns = nullptr;
  SDV_TRY(spectrum_acquire(c->n, c->cfg.spacing, c->cfg.toothnbig, c->cfg.toothnsmall, half_edge, c->stream, &ns));
  spectrum_release(c->spec);
  c->spec = ns;
  return SDV_OK;
}

int sdv_set_pair_batch(sdv_ctx* c, int pairs) {
  SDV_REQUIRE(c && pairs >= 0, SDV_ERR_ARG, "pair batch must be >= 0 (0: fill the work budget)");
  c->pair_batch = pairs;
  return SDV_OK;
}

int sdv_nuclear_potential(sdv_ctx* c, int numcenters, const int* centershift, const double* nuccharges,
                          int realkind4, double* pot) {
  SDV_REQUIRE(c && pot && numcenters >= 0, SDV_ERR_ARG, "bad argument");
  SDV_REQUIRE(numcenters == 0 || (centershift && nuccharges), SDV_ERR_ARG, "null centre arrays");
  const int n = c->n;
  for (int i = 0; i < 3 * numcenters; ++i) {  // SD/sincproject.f90:196-204
    const int pp = -n + centershift[i] + 1, qq = n + centershift[i] - 1;
    SDV_REQUIRE(pp % 2 == 0 && qq % 2 == 0, SDV_ERR_ARG,
                "centershift parity: with numpoints odd centershift must be even and vice versa");
  }
  int* dcs = nullptr;
  double *dz = nullptr, *dpot = nullptr;
  const size_t n3 = (size_t)c->n3;
  int rc = SDV_OK;
  if (hipMalloc(&dcs, 3 * (numcenters + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc(&dz, (numcenters + 1) * sizeof(double)) != hipSuccess ||
      hipMalloc(&dpot, n3 * sizeof(double)) != hipSuccess) {
    set_error("device allocation failed in sdv_nuclear_potential");
    rc = SDV_ERR_NOMEM;
  }
  if (rc == SDV_OK && numcenters > 0) {
    if (hipMemcpy(dcs, centershift, 3 * numcenters * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dz, nuccharges, numcenters * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      rc = SDV_ERR_HIP;
  }
  if (rc == SDV_OK) rc = nuclear_pot_launch(*c->spec, c->cfg.spacing, numcenters, dcs, dz, realkind4, dpot, c->stream);
  if (rc == SDV_OK && (hipMemcpyAsync(pot, dpot, n3 * sizeof(double), hipMemcpyDefault, c->stream) != hipSuccess ||
                       hipStreamSynchronize(c->stream) != hipSuccess))
    rc = SDV_ERR_HIP;
  hipFree(dcs);
  hipFree(dz);
  hipFree(dpot);
  if (rc == SDV_ERR_HIP) set_error("HIP error in sdv_nuclear_potential");
  return rc;
}

int sdv_get_kernel_octant(sdv_ctx* c, double* koct) {
  SDV_REQUIRE(c && koct, SDV_ERR_ARG, "null argument");
  SDV_HIP_CHECK(hipMemcpy(koct, c->spec->koct, (size_t)c->n3 * sizeof(double), hipMemcpyDeviceToHost));
  return SDV_OK;
}

int sdv_get_ke_stencil(sdv_ctx* c, double* ketot) {
  SDV_REQUIRE(c && ketot, SDV_ERR_ARG, "null argument");
  SDV_HIP_CHECK(hipMemcpy(ketot, c->ke, (size_t)(2 * c->n - 1) * sizeof(double), hipMemcpyDeviceToHost));
  return SDV_OK;
}

}  // extern "C"

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Upload a plan's index lists once; later calls with the same shape reuse them.
static int upload_plan(IdxPlan& pl) {
  pl.off_items = 0;
  pl.off_rows = align256(pl.items.size() * sizeof(PairItem));
  pl.off_cols = align256(pl.off_rows + pl.rows.size() * sizeof(int2));
  pl.off_rmap = align256(pl.off_cols + pl.cols.size() * sizeof(int));
  pl.off_cmap = align256(pl.off_rmap + pl.rmap.size() * sizeof(int2));
  const size_t bytes = pl.off_cmap + pl.cmap.size() * sizeof(int2) + 16;
  if (hipMalloc(&pl.dev, bytes) != hipSuccess) {
    pl.dev = nullptr;
    set_error("device allocation of the index plan failed");
    return SDV_ERR_NOMEM;
  }
  char* d = (char*)pl.dev;
  auto put = [&](size_t off, const void* src, size_t nb) -> hipError_t {
    return nb ? hipMemcpy(d + off, src, nb, hipMemcpyHostToDevice) : hipSuccess;
  };
  SDV_HIP_CHECK(put(pl.off_items, pl.items.data(), pl.items.size() * sizeof(PairItem)));
  SDV_HIP_CHECK(put(pl.off_rows, pl.rows.data(), pl.rows.size() * sizeof(int2)));
  SDV_HIP_CHECK(put(pl.off_cols, pl.cols.data(), pl.cols.size() * sizeof(int)));
  SDV_HIP_CHECK(put(pl.off_rmap, pl.rmap.data(), pl.rmap.size() * sizeof(int2)));
  SDV_HIP_CHECK(put(pl.off_cmap, pl.cmap.data(), pl.cmap.size() * sizeof(int2)));
  return SDV_OK;
}

// Pair bookkeeping of call_twoe_matelxxx00 (SD/coreproject.f90:1039-1100):
// twoereduced(:, 2a, 2b), 2a in [1,N], 2b in [low,high], column
// (2a-1) + N (2b-low); twoematel(1a,1b,2a,2b).  With identical bra and ket
// orbitals each unordered pair is convolved once; its mirror column gets
// conj(V) (hermnorm) or V (cnorm), and rows (1a,1b) with 1a > 1b reuse the
// density of (1b,1a) the same way.
int sdv::twoe_plan(sdv_ctx* c, int low, int high, bool sym, IdxPlan** out) {
  auto key = std::make_tuple(1, low, high, sym ? 1 : 0);
  auto it = c->plans.find(key);
  if (it != c->plans.end()) {
    *out = &it->second;
    return SDV_OK;
  }
  const int N = c->N;
  const int nnn = high - low + 1;
  const bool herm = c->cfg.cnormflag == 0;
  IdxPlan pl;
  auto col = [&](int a2, int b2) { return (a2 - 1) + N * (b2 - low); };
  pl.cmap.assign((size_t)N * nnn, int2{-1, 0});
  for (int b2 = low; b2 <= high; ++b2)
    for (int a2 = 1; a2 <= N; ++a2) {
      const bool mirror_needed = sym && a2 != b2 && a2 >= low && a2 <= high;
      if (mirror_needed && a2 > b2) continue;  // produced as the mirror of (b2, a2)
      const int ci = (int)pl.items.size();
      pl.items.push_back(PairItem{a2 - 1, b2 - 1, col(a2, b2), mirror_needed ? col(b2, a2) : -1});
      pl.cmap[col(a2, b2)] = int2{ci, 0};
      if (mirror_needed) pl.cmap[col(b2, a2)] = int2{ci, herm ? 1 : 0};
    }
  pl.rmap.resize((size_t)N * N);
  if (sym) {
    std::vector<int> ridx((size_t)N * N, -1);
    for (int b1 = 0; b1 < N; ++b1)
      for (int a1 = 0; a1 <= b1; ++a1) {
        ridx[a1 + N * b1] = (int)pl.rows.size();
        pl.rows.push_back(int2{a1, b1});
      }
    for (int b1 = 0; b1 < N; ++b1)
      for (int a1 = 0; a1 < N; ++a1)
        pl.rmap[a1 + N * b1] = a1 <= b1 ? int2{ridx[a1 + N * b1], 0} : int2{ridx[b1 + N * a1], herm ? 1 : 0};
  } else {
    for (int b1 = 0; b1 < N; ++b1)
      for (int a1 = 0; a1 < N; ++a1) {
        pl.rmap[a1 + N * b1] = int2{(int)pl.rows.size(), 0};
        pl.rows.push_back(int2{a1, b1});
      }
  }
  pl.cols.resize(pl.items.size());
  for (size_t i = 0; i < pl.items.size(); ++i) pl.cols[i] = pl.items[i].col;
  if (sym && herm) {
    // The densities of diagonal pairs are real (|phi_a|^2): two of them share
    // one complex convolution.  The reduction's column list above stays per
    // pair; only the convolution's work list changes.
    std::vector<PairItem> conv;
    std::vector<size_t> diag;
    for (size_t i = 0; i < pl.items.size(); ++i) {
      if (pl.items[i].a == pl.items[i].b)
        diag.push_back(i);
      else
        conv.push_back(pl.items[i]);
    }
    for (size_t k = 0; k + 1 < diag.size(); k += 2) {
      const PairItem& p = pl.items[diag[k]];
      const PairItem& q = pl.items[diag[k + 1]];
      conv.push_back(PairItem{p.a, q.a, p.col, q.col, 1});
    }
    if (diag.size() % 2) conv.push_back(pl.items[diag.back()]);
    pl.items.swap(conv);
  }
  SDV_TRY(upload_plan(pl));
  auto ins = c->plans.emplace(key, std::move(pl));
  *out = &ins.first->second;
  return SDV_OK;
}

static int density_plan(sdv_ctx* c, int allsize, IdxPlan** out) {
  auto key = std::make_tuple(0, allsize, 0, 0);
  auto it = c->plans.find(key);
  if (it != c->plans.end()) {
    *out = &it->second;
    return SDV_OK;
  }
  IdxPlan pl;
  pl.items.resize(allsize);
  for (int i = 0; i < allsize; ++i) pl.items[i] = PairItem{i, 0, i, -1};
  SDV_TRY(upload_plan(pl));
  auto ins = c->plans.emplace(key, std::move(pl));
  *out = &ins.first->second;
  return SDV_OK;
}

// Run the convolution over a plan's items, in batches sized to the work buffer.
static int run_pairs(sdv_ctx* c, const IdxPlan& pl, const void* phi1, const void* phi2, int density_mode, void* out,
                     hipStream_t st) {
  const int cn = c->cfg.cnormflag;
  const int np = (int)pl.items.size();
  if (np == 0) return SDV_OK;
  int batch = c->pair_batch;
  if (batch <= 0) {  // every pair of the call in one launch when the budget allows (288 GB of HBM)
    const size_t per = conv_work_bytes(c->spec->g, 1, c->precision);
    batch = (int)std::max<size_t>(1, std::min<size_t>((size_t)np, c->work_budget / per));
  }
  batch = std::min(np, batch);
  SDV_TRY(ensure(c->work, conv_work_bytes(c->spec->g, batch, c->precision)));
  for (int p0 = 0; p0 < np; p0 += batch) {
    ConvArgsHost a;
    a.n = c->n;
    a.npairs = std::min(batch, np - p0);
    a.phi1 = phi1;
    a.phi2 = phi2;
    a.density_mode = density_mode;
    a.conj_b = (cn || density_mode) ? 0 : 1;
    a.mirror_conj = cn ? 0 : 1;
    a.pairs_dev = (const PairItem*)((char*)pl.dev + pl.off_items) + p0;
    a.work = c->work.p;
    a.out = out;
    SDV_TRY(conv_launch(*c->spec, a, c->precision, st, &c->prof));
  }
  return SDV_OK;
}

extern "C" {

int sdv_op_tinv_dev(sdv_ctx* c, const void* den, void* red, int allsize, int circsize, void* stream) {
  SDV_REQUIRE(c && den && red, SDV_ERR_ARG, "null argument");
  SDV_REQUIRE(allsize >= 0 && circsize > 0 && allsize % circsize == 0, SDV_ERR_ARG,
              "SIZE ERROR OP_TINV: allsize must be a multiple of circsize");
  hipStream_t st = (hipStream_t)stream;
  const int dbg = c->cfg.debugflag;
  if (dbg == 4040 || dbg == 4242 || dbg == 4343) {  // SD/coreproject.f90:260-263
    SDV_HIP_CHECK(hipMemsetAsync(red, 0, (size_t)c->n3 * allsize * sizeof(z64), st));
    return SDV_OK;
  }
  if (allsize == 0) return SDV_OK;
  IdxPlan* pl = nullptr;
  SDV_TRY(density_plan(c, allsize, &pl));
  return run_pairs(c, *pl, den, den, 1, red, st);
}

}  // extern "C"

// ---- allocation-free op_tinv
// The device memory of sdv_op_tinv_noalloc_dev: the pair items of one chunk
// at the front, the chunk's convolution intermediates behind them.  One copy
// per device (the array below), held until the process ends.  Host code never
// passes the array's name to a kernel: the name is a host-side stand-in, and
// each device's address comes from hipGetSymbolAddress (scratch_ptr).
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
constexpr int kNoAllocItems = 4096;  // PairItem slots at the front (80 KB)
constexpr size_t kNoAllocItemBytes = ((size_t)kNoAllocItems * sizeof(PairItem) + 255) & ~(size_t)255;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
}  // namespace

// The current device's address of `scratch`, looked up once per device (at
// sdv_init, else at the entry point's first call on that device).
static int scratch_ptr(void** out) {
  int dev = 0;
  SDV_HIP_CHECK(hipGetDevice(&dev));
  SDV_REQUIRE(dev >= 0 && dev < kMaxDevices, SDV_ERR_UNSUPPORTED, "device ordinal beyond the scratch table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      set_error(std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
      return SDV_ERR_HIP;
    }
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return SDV_OK;
}

__global__ void k_density_items(PairItem* items, int p0, int np) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < np) items[j] = PairItem{p0 + j, 0, p0 + j, -1, 0};
}

__global__ void k_zero_z64(z64* p, long long count) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < count;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = mk(0.0, 0.0);
}

extern "C" {

int sdv_op_tinv_noalloc_dev(sdv_ctx* c, const void* den, void* red, int allsize, int circsize, void* stream) {
  SDV_REQUIRE(c && den && red, SDV_ERR_ARG, "null argument");
  SDV_REQUIRE(allsize >= 0 && circsize > 0 && allsize % circsize == 0, SDV_ERR_ARG,
              "SIZE ERROR OP_TINV: allsize must be a multiple of circsize");
  hipStream_t st = (hipStream_t)stream;
  const int dbg = c->cfg.debugflag;
  if (dbg == 4040 || dbg == 4242 || dbg == 4343) {  // SD/coreproject.f90:260-263
    const long long cnt = (long long)c->n3 * allsize;
    if (cnt > 0) {
      hipLaunchKernelGGL(k_zero_z64, dim3((unsigned)std::min<long long>((cnt + 255) / 256, 4096)), dim3(256), 0, st,
                         (z64*)red, cnt);
      SDV_HIP_CHECK(hipGetLastError());
    }
    return SDV_OK;
  }
  if (allsize == 0) return SDV_OK;
  void* base = nullptr;
  SDV_TRY(scratch_ptr(&base));
  const size_t per = conv_work_bytes(c->spec->g, 1, c->precision);
  SDV_REQUIRE(per <= kNoAllocScratchBytes - kNoAllocItemBytes, SDV_ERR_UNSUPPORTED,
              "one convolution's work buffer exceeds the 128 MiB allocation-free scratch");
  const int chunk = (int)std::min<size_t>((size_t)kNoAllocItems, (kNoAllocScratchBytes - kNoAllocItemBytes) / per);
  PairItem* items = (PairItem*)base;
  void* work = (char*)base + kNoAllocItemBytes;
  for (int p0 = 0; p0 < allsize; p0 += chunk) {
    const int np = std::min(chunk, allsize - p0);
    // Items are rewritten per chunk; the stream orders this after the previous
    // chunk's kernels, which read them.
    hipLaunchKernelGGL(k_density_items, dim3((np + 255) / 256), dim3(256), 0, st, items, p0, np);
    SDV_HIP_CHECK(hipGetLastError());
    ConvArgsHost a;
    a.n = c->n;
    a.npairs = np;
    a.phi1 = den;
    a.phi2 = den;
    a.density_mode = 1;
    a.conj_b = 0;
    a.mirror_conj = c->cfg.cnormflag ? 0 : 1;
    a.pairs_dev = items;
    a.work = work;
    a.out = red;
    SDV_TRY(conv_launch(*c->spec, a, c->precision, st, nullptr));
  }
  return SDV_OK;
}

static int check_twoe_args(sdv_ctx* c, int low, int high) {
  SDV_REQUIRE(low >= 1 && high <= c->N, SDV_ERR_ARG, "programmer fail: lowspf/highspf out of range");
  return SDV_OK;
}

int sdv_twoereduced_dev(sdv_ctx* c, int low, int high, const void* phi1, const void* phi2, void* twoereduced,
                        void* stream) {
  SDV_REQUIRE(c && phi1 && phi2 && twoereduced, SDV_ERR_ARG, "null argument");
  if (high < low) return SDV_OK;
  SDV_TRY(check_twoe_args(c, low, high));
  hipStream_t st = (hipStream_t)stream;
  if (c->cfg.notwoflag) {
    SDV_HIP_CHECK(hipMemsetAsync(twoereduced, 0, (size_t)c->n3 * c->N * (high - low + 1) * sizeof(z64), st));
    return SDV_OK;
  }
  IdxPlan* pl = nullptr;
  SDV_TRY(twoe_plan(c, low, high, phi1 == phi2, &pl));
  return run_pairs(c, *pl, phi1, phi2, 0, twoereduced, st);
}

int sdv_twoe_convolutions(sdv_ctx* c, int low, int high, int same) {
  SDV_REQUIRE(c, SDV_ERR_ARG, "null context");
  if (high < low) return 0;
  SDV_TRY(check_twoe_args(c, low, high));
  IdxPlan* pl = nullptr;
  SDV_TRY(twoe_plan(c, low, high, same != 0, &pl));
  return (int)pl->items.size();
}

int sdv_call_twoe_matel00_dev(sdv_ctx* c, int low, int high, const void* phi1, const void* phi2, void* twoematel,
                              void* twoereduced, void* stream) {
  SDV_REQUIRE(c && phi1 && phi2 && twoematel && twoereduced, SDV_ERR_ARG, "null argument");
  const int N = c->N, n3 = c->n3;
  const int nnn = high - low + 1;
  if (nnn <= 0) return SDV_OK;  // SD/coreproject.f90:969-972
  SDV_TRY(check_twoe_args(c, low, high));
  hipStream_t st = (hipStream_t)stream;
  if (c->cfg.notwoflag) {  // :1019-1023
    SDV_HIP_CHECK(hipMemsetAsync(twoereduced, 0, (size_t)n3 * N * nnn * sizeof(z64), st));
    SDV_HIP_CHECK(hipMemsetAsync(twoematel, 0, (size_t)N * N * N * nnn * sizeof(z64), st));
    return SDV_OK;
  }
  const bool sym = (phi1 == phi2);
  IdxPlan* pl = nullptr;
  SDV_TRY(twoe_plan(c, low, high, sym, &pl));
  SDV_TRY(run_pairs(c, *pl, phi1, phi2, 0, twoereduced, st));
  const int nrows = (int)pl->rows.size(), ncols = (int)pl->cols.size();
  const int msym = (sym && low == 1 && high == N) ? 1 : 0;
  const size_t mbytes = matel_scratch_bytes(n3, nrows, ncols, msym);
  SDV_TRY(ensure(c->scratch, mbytes));
  SDV_TRY(ensure(c->gout, (size_t)(2 * nrows) * (2 * ncols) * sizeof(double)));
  const char* ib = (const char*)pl->dev;
  MatelArgsHost m;
  m.n3 = n3;
  m.N = N;
  m.nrows = nrows;
  m.rows_dev = (const int2*)(ib + pl->off_rows);
  m.ncols = ncols;
  m.cols_dev = (const int*)(ib + pl->off_cols);
  m.phi1 = phi1;
  m.phi2 = phi2;
  m.conj_b = c->cfg.cnormflag ? 0 : 1;
  m.tr = twoereduced;
  m.symmetric = msym;
  m.gout = (double*)c->gout.p;
  m.scratch = c->scratch.p;
  m.scratch_bytes = c->scratch.cap;
  SDV_TRY(matel_launch(m, st, &c->prof));
  const long tot = (long)N * N * N * nnn;
  hipEvent_t e0 = c->prof.begin(st);
  hipLaunchKernelGGL(k_assemble, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, N, nnn,
                     (const int2*)(ib + pl->off_rmap), (const int2*)(ib + pl->off_cmap), (const double*)c->gout.p,
                     2 * ncols, (z64*)twoematel);
  SDV_HIP_CHECK(hipGetLastError());
  c->prof.end(SDV_STAGE_MATEL_ASSEMBLE, e0, st);
  return SDV_OK;
}

// ---- multi-set batch
}  // extern "C"

static int prof_reset(Prof* p, int enable) {
  SDV_HIP_CHECK(hipDeviceSynchronize());
  for (auto& r : p->recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  p->recs.clear();
  p->on = enable != 0;
  return SDV_OK;
}

static int prof_read(const Prof& p, double* ms, int* count) {
  for (int i = 0; i < SDV_NSTAGES; ++i) {
    ms[i] = 0.0;
    count[i] = 0;
  }
  for (const auto& r : p.recs) {
    SDV_HIP_CHECK(hipEventSynchronize(r.b));
    float t = 0.f;
    SDV_HIP_CHECK(hipEventElapsedTime(&t, r.a, r.b));
    if (r.stage >= 0 && r.stage < SDV_NSTAGES) {
      ms[r.stage] += t;
      count[r.stage] += 1;
    }
  }
  return SDV_OK;
}

struct sdv_batch {
  // The batch holds its own reference on every set's kernel spectrum and its
  // own copy of every set's pair list, so a context may change its circulant
  // (sdv_set_circulant) or be freed while the batch lives: the batch keeps
  // computing with the spectrum it was built with.
  std::vector<Spectrum*> specs;
  std::vector<std::vector<PairItem>> items;  // reserved up front: addresses stay put
  std::vector<BatchSetHost> sets;
  long long pairs = 0;  // pair potentials delivered per step (unordered pairs of every set)
  BatchPlan plan;
  int precision = 0;
  int mode = BATCH_GROUPED;
  int grid_mult = 1;
  int nstreams = 4;
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> ev;
  void* dsets = nullptr;
  void* ditems = nullptr;
  void* dwx = nullptr;
  void* dwy = nullptr;
  void* X = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t graph_stream = nullptr;
  Prof prof;
};

static void batch_release(sdv_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  (void)hipFree(b->dsets);
  (void)hipFree(b->ditems);
  (void)hipFree(b->dwx);
  (void)hipFree(b->dwy);
  (void)hipFree(b->X);
  b->dsets = b->ditems = b->dwx = b->dwy = b->X = nullptr;
  for (hipStream_t st : b->side) (void)hipStreamDestroy(st);
  for (hipEvent_t e : b->ev) (void)hipEventDestroy(e);
  b->side.clear();
  b->ev.clear();
}

static int batch_build(sdv_batch* b) {
  batch_release(b);
  int dev = 0, ncu = 256;
  SDV_HIP_CHECK(hipGetDevice(&dev));
  SDV_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  b->plan = batch_plan(b->sets, b->precision, b->mode, b->grid_mult, ncu);
  SDV_REQUIRE(b->plan.grid_x > 0 && b->plan.grid_y > 0, SDV_ERR_UNSUPPORTED, "batch holds an unsupported grid");
  const size_t sb = batch_set_bytes(b->precision);
  std::vector<char> hs(sb * b->sets.size());
  for (size_t i = 0; i < b->sets.size(); ++i) batch_fill_set(b->precision, b->sets[i], hs.data() + i * sb);
  const BatchPlan& p = b->plan;
  if (hipMalloc(&b->dsets, hs.size() + 16) != hipSuccess ||
      hipMalloc(&b->ditems, p.items.size() * batch_item_bytes() + 16) != hipSuccess ||
      hipMalloc(&b->dwx, p.wx.size() * batch_work_bytes() + 16) != hipSuccess ||
      hipMalloc(&b->dwy, p.wy.size() * batch_work_bytes() + 16) != hipSuccess ||
      hipMalloc(&b->X, p.x_bytes + 256) != hipSuccess) {
    set_error("device allocation failed for the batch");
    batch_release(b);
    return SDV_ERR_NOMEM;
  }
  SDV_HIP_CHECK(hipMemcpy(b->dsets, hs.data(), hs.size(), hipMemcpyHostToDevice));
  SDV_HIP_CHECK(hipMemcpy(b->ditems, p.items.data(), p.items.size() * batch_item_bytes(), hipMemcpyHostToDevice));
  SDV_HIP_CHECK(hipMemcpy(b->dwx, p.wx.data(), p.wx.size() * batch_work_bytes(), hipMemcpyHostToDevice));
  SDV_HIP_CHECK(hipMemcpy(b->dwy, p.wy.data(), p.wy.size() * batch_work_bytes(), hipMemcpyHostToDevice));
  if (b->mode == BATCH_GROUPED && p.groups.size() > 1) {
    const int ns = std::min(b->nstreams, (int)p.groups.size());
    for (int i = 0; i < ns; ++i) {
      hipStream_t st = nullptr;
      SDV_HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      b->side.push_back(st);
    }
    for (int i = 0; i < ns + 1 && ns > 0; ++i) {
      hipEvent_t e = nullptr;
      SDV_HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      b->ev.push_back(e);
    }
  }
  return SDV_OK;
}

static BatchDev batch_dev(const sdv_batch* b) {
  BatchDev d;
  d.sets = b->dsets;
  d.items = (const BItemDev*)b->ditems;
  d.wx = (const BWork*)b->dwx;
  d.wy = (const BWork*)b->dwy;
  d.X = b->X;
  d.nside = (int)b->side.size();
  d.side = d.nside ? const_cast<hipStream_t*>(b->side.data()) : nullptr;
  d.ev = d.nside ? const_cast<hipEvent_t*>(b->ev.data()) : nullptr;
  return d;
}

extern "C" {

int sdv_batch_create(int nsets, const sdv_batch_set* sets, int precision, sdv_batch** out) {
  SDV_REQUIRE(out && sets && nsets > 0, SDV_ERR_ARG, "bad batch arguments");
  SDV_REQUIRE(precision == 0 || precision == 1, SDV_ERR_ARG, "precision must be 0 or 1");
  sdv_batch* b = new sdv_batch();
  b->precision = precision;
  b->items.reserve(nsets);
  for (int i = 0; i < nsets; ++i) {
    sdv_ctx* c = sets[i].ctx;
    if (!c || !sets[i].inspfs || !sets[i].twoereduced || sets[i].numspf != c->N) {
      set_error("batch set " + std::to_string(i) + ": null pointer or numspf differs from its context");
      sdv_batch_free(b);
      return SDV_ERR_ARG;
    }
    IdxPlan* pl = nullptr;
    int rc = twoe_plan(c, 1, c->N, true, &pl);
    if (rc != SDV_OK) {
      sdv_batch_free(b);
      return rc;
    }
    spectrum_retain(c->spec);
    b->specs.push_back(c->spec);
    b->items.push_back(pl->items);
    BatchSetHost h;
    h.spec = c->spec;
    h.items = &b->items.back();
    b->pairs += (long long)pl->cols.size();
    h.phi = sets[i].inspfs;
    h.out = sets[i].twoereduced;
    h.cnorm = c->cfg.cnormflag;
    b->sets.push_back(h);
  }
  int rc = batch_build(b);
  if (rc != SDV_OK) {
    sdv_batch_free(b);
    return rc;
  }
  *out = b;
  return SDV_OK;
}

int sdv_batch_set_schedule(sdv_batch* b, int mode, int grid_mult, int nstreams) {
  SDV_REQUIRE(b, SDV_ERR_ARG, "null batch");
  SDV_REQUIRE(mode == SDV_BATCH_FUSED || mode == SDV_BATCH_GROUPED, SDV_ERR_ARG, "unknown batch schedule");
  SDV_REQUIRE(grid_mult >= 1 && grid_mult <= 64, SDV_ERR_ARG, "grid multiplier must be in [1, 64]");
  SDV_REQUIRE(nstreams >= 0 && nstreams <= 8, SDV_ERR_ARG, "stream count must be in [0, 8]");
  SDV_HIP_CHECK(hipDeviceSynchronize());
  b->mode = mode;
  b->grid_mult = grid_mult;
  b->nstreams = nstreams;
  return batch_build(b);
}

int sdv_batch_run(sdv_batch* b, void* stream) {
  SDV_REQUIRE(b, SDV_ERR_ARG, "null batch");
  hipStream_t st = (hipStream_t)stream;
  if (b->exec && st == b->graph_stream) {
    SDV_HIP_CHECK(hipGraphLaunch(b->exec, st));
    return SDV_OK;
  }
  return batch_launch(b->plan, batch_dev(b), st, &b->prof);
}

int sdv_batch_capture(sdv_batch* b, void* stream) {
  SDV_REQUIRE(b && stream, SDV_ERR_ARG, "capture needs a batch and a non-default stream");
  hipStream_t st = (hipStream_t)stream;
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
  SDV_HIP_CHECK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  const int rc = batch_launch(b->plan, batch_dev(b), st, nullptr);
  hipGraph_t g = nullptr;
  const hipError_t e = hipStreamEndCapture(st, &g);
  if (rc != SDV_OK) return rc;
  SDV_HIP_CHECK(e);
  b->graph = g;
  SDV_HIP_CHECK(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
  b->graph_stream = st;
  return SDV_OK;
}

int sdv_batch_stats(sdv_batch* b, long long* npairs, long long* work_bytes) {
  SDV_REQUIRE(b, SDV_ERR_ARG, "null batch");
  if (npairs) *npairs = b->pairs;
  if (work_bytes) *work_bytes = (long long)b->plan.x_bytes;
  return SDV_OK;
}

int sdv_batch_free(sdv_batch* b) {
  if (!b) return SDV_OK;
  (void)hipDeviceSynchronize();
  batch_release(b);
  for (auto& r : b->prof.recs) {
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  for (Spectrum* sp : b->specs) spectrum_release(sp);
  delete b;
  return SDV_OK;
}

int sdv_profile(sdv_ctx* c, int enable) {
  SDV_REQUIRE(c, SDV_ERR_ARG, "null context");
  return prof_reset(&c->prof, enable);
}

int sdv_profile_read(sdv_ctx* c, double* ms, int* count) {
  SDV_REQUIRE(c && ms && count, SDV_ERR_ARG, "null argument");
  return prof_read(c->prof, ms, count);
}

int sdv_batch_profile(sdv_batch* b, int enable) {
  SDV_REQUIRE(b, SDV_ERR_ARG, "null batch");
  return prof_reset(&b->prof, enable);
}

int sdv_batch_profile_read(sdv_batch* b, double* ms, int* count) {
  SDV_REQUIRE(b && ms && count, SDV_ERR_ARG, "null argument");
  return prof_read(b->prof, ms, count);
}

int sdv_mult_ke_dev(sdv_ctx* c, const void* in, void* out, int howmany, void* stream) {
  SDV_REQUIRE(c && in && out, SDV_ERR_ARG, "null argument");
  SDV_REQUIRE(howmany >= 0, SDV_ERR_ARG, "howmany must be >= 0");
  KeArgsHost a;
  a.n = c->n;
  a.howmany = howmany;
  a.ke = c->ke;
  a.invj = c->invj;
  a.sdiag = c->sdiag;
  a.mask = c->mask;
  a.in = in;
  a.out = out;
  return ke_launch(a, c->precision, (hipStream_t)stream, &c->prof);
}

int sdv_sparseops_matel_dev(sdv_ctx* c, int low, int high, const void* inspfs1, const void* inspfs2, void* opmatel,
                            void* stream) {
  SDV_REQUIRE(c && inspfs1 && inspfs2 && opmatel, SDV_ERR_ARG, "null argument");
  SDV_REQUIRE(low >= 1 && high <= c->N, SDV_ERR_ARG, "spf range outside 1..numspf");
  if (high < low) return SDV_OK;
  const int num = high - low + 1;
  const size_t col = (size_t)c->n3 * sizeof(z64);
  hipStream_t st = (hipStream_t)stream;
  SDV_TRY(ensure(c->ob_ke, col * num));
  SDV_TRY(ensure(c->ob_scr, dot_scratch_bytes(c->n3, c->N, num)));
  KeArgsHost a;
  a.n = c->n;
  a.howmany = num;
  a.ke = c->ke;
  a.invj = c->invj;
  a.sdiag = c->sdiag;
  a.mask = c->mask;
  a.in = (const char*)inspfs2 + col * (low - 1);
  a.out = c->ob_ke.p;
  SDV_TRY(ke_launch(a, c->precision, st, &c->prof));
  return dot_launch(c->n3, c->N, num, inspfs1, c->ob_ke.p, nullptr, c->cfg.cnormflag ? 0 : 1, c->ob_scr.p, opmatel,
                    c->N, st);
}

int sdv_pot_matel_dev(sdv_ctx* c, int low, int high, const void* inspfs1, const void* inspfs2, const double* pot,
                      void* potmatel, void* stream) {
  SDV_REQUIRE(c && inspfs1 && inspfs2 && pot && potmatel, SDV_ERR_ARG, "null argument");
  SDV_REQUIRE(low >= 1 && high <= c->N, SDV_ERR_ARG, "spf range outside 1..numspf");
  if (high < low) return SDV_OK;
  const int num = high - low + 1;
  const size_t col = (size_t)c->n3 * sizeof(z64);
  SDV_TRY(ensure(c->ob_scr, dot_scratch_bytes(c->n3, c->N, num)));
  return dot_launch(c->n3, c->N, num, inspfs1, (const char*)inspfs2 + col * (low - 1), pot,
                    c->cfg.cnormflag ? 0 : 1, c->ob_scr.p, potmatel, c->N, (hipStream_t)stream);
}

int sdv_actreduced00_dev(sdv_ctx* c, const void* inspfs, const void* projspfs, const void* denmat,
                         const void* invdenmat, const void* reducedpot, const double* pot, const double* tdpot,
                         int dentimeflag, int projflag, double timefac_re, double timefac_im, void* outspfs,
                         void* stream) {
  SDV_REQUIRE(c && inspfs && denmat && reducedpot && outspfs, SDV_ERR_ARG, "null argument");
  SDV_REQUIRE(!dentimeflag || invdenmat, SDV_ERR_ARG, "dentimeflag needs invdenmat");
  SDV_REQUIRE(!projflag || projspfs, SDV_ERR_ARG, "projflag needs projspfs");
  hipStream_t st = (hipStream_t)stream;
  const int N = c->N, n3 = c->n3;
  const size_t cols = (size_t)n3 * N * sizeof(z64);
  SDV_TRY(ensure(c->ar_w, cols));
  SDV_TRY(ensure(c->ob_ke, cols));
  SDV_TRY(ensure(c->ar_rp, cols));
  SDV_TRY(ensure(c->ar_s, cols));
  SDV_TRY(ensure(c->ar_s2, cols));
  SDV_TRY(ensure(c->ar_dots, (size_t)N * N * sizeof(z64)));
  SDV_TRY(ensure(c->ob_scr, dot_scratch_bytes(n3, N, N)));
  const z64 one = mk(1.0, 0.0), zero = mk(0.0, 0.0);
  // spfinvrsq = spfinvr = inspfs . denmat   (S/derivs.f90:670-688, bondpoints(1) = 1)
  SDV_TRY(orb_gemm_launch(n3, N, N, inspfs, denmat, N, one, zero, c->ar_w.p, st));
  KeArgsHost a;
  a.n = c->n;
  a.howmany = N;
  a.ke = c->ke;
  a.invj = c->invj;
  a.sdiag = c->sdiag;
  a.mask = c->mask;
  a.in = c->ar_w.p;
  a.out = c->ob_ke.p;
  SDV_TRY(ke_launch(a, c->precision, st, &c->prof));                                       // :691-693
  const bool vel = tdpot && c->cfg.velflag;
  if (vel) {  // velmultiply (:747-750): sum_d (-i A_d) W_d (-fdtot) W_d spfinvr, added onto the kinetic apply
    KeArgsHost v = a;
    v.ke = c->fd;
    v.invj = c->invsqj;  // invsqrtjacobian under ECS (SD/coreproject.f90:1325-1331,1356-1362)
    v.sdiag = nullptr;
    for (int d = 0; d < 3; ++d) v.coef[d] = mk(tdpot[2 * d + 1], -tdpot[2 * d]);  // (0,-1) * tdpot_d   :2128-2130
    v.use_coef = 1;
    v.accumulate = 1;
    SDV_TRY(ke_launch(v, c->precision, st, nullptr));
  }
  SDV_TRY(mult_reducedpot_launch(n3, N, N, inspfs, reducedpot, c->ar_rp.p, st));           // :762-764
  SDV_TRY(orb_sum_launch(c->n, N, c->ob_ke.p, pot, c->ar_w.p, c->ar_rp.p, c->dip, vel ? nullptr : tdpot,
                         c->ar_s.p, st));  // + mult_pot :698, length-gauge pulse :742-746
  void* cur = c->ar_s.p;
  if (dentimeflag) {  // spfmult . invdenmat * timefac   :780-787
    SDV_TRY(orb_gemm_launch(n3, N, N, cur, invdenmat, N, mk(timefac_re, timefac_im), zero, c->ar_s2.p, st));
    cur = c->ar_s2.p;
  }
  SDV_HIP_CHECK(hipMemcpyAsync(outspfs, cur, cols, hipMemcpyDeviceToDevice, st));
  if (projflag) {  // out = spfmult - projspfs . <projspfs | spfmult>   project00, S/proputils.f90:10-60
    SDV_TRY(dot_launch(n3, N, N, projspfs, cur, nullptr, c->cfg.cnormflag ? 0 : 1, c->ob_scr.p, c->ar_dots.p, N, st));
    SDV_TRY(orb_gemm_launch(n3, N, N, projspfs, c->ar_dots.p, N, mk(-1.0, 0.0), one, outspfs, st));
  }
  return SDV_OK;
}

int sdv_mult_reducedpot_dev(sdv_ctx* c, int first, int last, const void* inspfs, void* outspfs,
                            const void* reducedpot, void* stream) {
  SDV_REQUIRE(c && inspfs && outspfs && reducedpot, SDV_ERR_ARG, "null argument");
  if (last < first) return SDV_OK;
  return mult_reducedpot_launch(c->n3, c->N, last - first + 1, inspfs, reducedpot, outspfs, (hipStream_t)stream);
}

int sdv_get_reducedpot_dev(sdv_ctx* c, const void* tr, const void* twoden, void* rp, void* stream) {
  SDV_REQUIRE(c && tr && twoden && rp, SDV_ERR_ARG, "null argument");
  SDV_TRY(ensure(c->rp_bt, reducedpot_scratch_bytes(c->N)));
  return reducedpot_launch(c->n3, c->N, tr, twoden, rp, c->rp_bt.p, (hipStream_t)stream);
}

// ---- host-pointer (Fortran drop-in) variants: stage, run, copy back, sync.
static int stage_in(sdv_ctx* c, sdv_ctx::Buf& b, const void* h, size_t bytes, const void** d) {
  SDV_TRY(ensure(b, bytes));
  SDV_HIP_CHECK(hipMemcpyAsync(b.p, h, bytes, hipMemcpyDefault, c->stream));
  *d = b.p;
  return SDV_OK;
}

int sdv_op_tinv(sdv_ctx* c, const double* den, double* red, int allsize, int circsize) {
  SDV_REQUIRE(c && den && red, SDV_ERR_ARG, "null argument");
  const size_t bytes = (size_t)c->n3 * (allsize > 0 ? allsize : 0) * sizeof(z64);
  const void* dd;
  SDV_TRY(stage_in(c, c->h_in1, den, bytes, &dd));
  SDV_TRY(ensure(c->h_out1, bytes));
  SDV_TRY(sdv_op_tinv_dev(c, dd, c->h_out1.p, allsize, circsize, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(red, c->h_out1.p, bytes, hipMemcpyDefault, c->stream));
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SDV_OK;
}

int sdv_call_twoe_matel00(sdv_ctx* c, int low, int high, const double* phi1, const double* phi2, double* tm,
                          double* tr) {
  SDV_REQUIRE(c && phi1 && phi2 && tm && tr, SDV_ERR_ARG, "null argument");
  const int N = c->N, nnn = high - low + 1;
  if (nnn <= 0) return SDV_OK;
  const size_t pb = (size_t)c->n3 * N * sizeof(z64);
  const void *d1, *d2;
  SDV_TRY(stage_in(c, c->h_in1, phi1, pb, &d1));
  if (phi2 == phi1)
    d2 = d1;
  else
    SDV_TRY(stage_in(c, c->h_in2, phi2, pb, &d2));
  const size_t trb = (size_t)c->n3 * N * nnn * sizeof(z64), tmb = (size_t)N * N * N * nnn * sizeof(z64);
  SDV_TRY(ensure(c->h_out1, trb));
  SDV_TRY(ensure(c->h_out2, tmb));
  SDV_TRY(sdv_call_twoe_matel00_dev(c, low, high, d1, d2, c->h_out2.p, c->h_out1.p, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(tr, c->h_out1.p, trb, hipMemcpyDefault, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(tm, c->h_out2.p, tmb, hipMemcpyDefault, c->stream));
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SDV_OK;
}

int sdv_twoereduced(sdv_ctx* c, int low, int high, const double* phi1, const double* phi2, double* tr) {
  SDV_REQUIRE(c && phi1 && phi2 && tr, SDV_ERR_ARG, "null argument");
  const int N = c->N, nnn = high - low + 1;
  if (nnn <= 0) return SDV_OK;
  const size_t pb = (size_t)c->n3 * N * sizeof(z64);
  const void *d1, *d2;
  SDV_TRY(stage_in(c, c->h_in1, phi1, pb, &d1));
  if (phi2 == phi1)
    d2 = d1;
  else
    SDV_TRY(stage_in(c, c->h_in2, phi2, pb, &d2));
  const size_t trb = (size_t)c->n3 * N * nnn * sizeof(z64);
  SDV_TRY(ensure(c->h_out1, trb));
  SDV_TRY(sdv_twoereduced_dev(c, low, high, d1, d2, c->h_out1.p, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(tr, c->h_out1.p, trb, hipMemcpyDefault, c->stream));
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SDV_OK;
}

int sdv_mult_ke(sdv_ctx* c, const double* in, double* out, int howmany) {
  SDV_REQUIRE(c && in && out, SDV_ERR_ARG, "null argument");
  const size_t bytes = (size_t)c->n3 * (howmany > 0 ? howmany : 0) * sizeof(z64);
  const void* di;
  SDV_TRY(stage_in(c, c->h_in1, in, bytes, &di));
  SDV_TRY(ensure(c->h_out1, bytes));
  SDV_TRY(sdv_mult_ke_dev(c, di, c->h_out1.p, howmany, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(out, c->h_out1.p, bytes, hipMemcpyDefault, c->stream));
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SDV_OK;
}

int sdv_sparseops_matel(sdv_ctx* c, int low, int high, const double* inspfs1, const double* inspfs2,
                        double* opmatel) {
  SDV_REQUIRE(c && inspfs1 && inspfs2 && opmatel, SDV_ERR_ARG, "null argument");
  if (high < low) return SDV_OK;
  const size_t pb = (size_t)c->n3 * c->N * sizeof(z64);
  const void *d1 = nullptr, *d2 = nullptr;
  SDV_TRY(stage_in(c, c->h_in1, inspfs1, pb, &d1));
  SDV_TRY(stage_in(c, c->h_in2, inspfs2, pb, &d2));
  const size_t ob = (size_t)c->N * (high - low + 1) * sizeof(z64);
  SDV_TRY(ensure(c->h_out2, ob));
  SDV_TRY(sdv_sparseops_matel_dev(c, low, high, d1, d2, c->h_out2.p, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(opmatel, c->h_out2.p, ob, hipMemcpyDefault, c->stream));
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SDV_OK;
}

int sdv_pot_matel(sdv_ctx* c, int low, int high, const double* inspfs1, const double* inspfs2, const double* pot,
                  double* potmatel) {
  SDV_REQUIRE(c && inspfs1 && inspfs2 && pot && potmatel, SDV_ERR_ARG, "null argument");
  if (high < low) return SDV_OK;
  const size_t pb = (size_t)c->n3 * c->N * sizeof(z64);
  const void *d1 = nullptr, *d2 = nullptr, *dp = nullptr;
  SDV_TRY(stage_in(c, c->h_in1, inspfs1, pb, &d1));
  SDV_TRY(stage_in(c, c->h_in2, inspfs2, pb, &d2));
  SDV_TRY(stage_in(c, c->h_pot, pot, (size_t)c->n3 * sizeof(double), &dp));
  const size_t ob = (size_t)c->N * (high - low + 1) * sizeof(z64);
  SDV_TRY(ensure(c->h_out2, ob));
  SDV_TRY(sdv_pot_matel_dev(c, low, high, d1, d2, (const double*)dp, c->h_out2.p, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(potmatel, c->h_out2.p, ob, hipMemcpyDefault, c->stream));
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SDV_OK;
}

int sdv_actreduced00(sdv_ctx* c, const double* inspfs, const double* projspfs, const double* denmat,
                     const double* invdenmat, const double* reducedpot, const double* pot, const double* tdpot,
                     int dentimeflag, int projflag, double timefac_re, double timefac_im, double* outspfs) {
  SDV_REQUIRE(c && inspfs && denmat && reducedpot && outspfs, SDV_ERR_ARG, "null argument");
  const size_t pb = (size_t)c->n3 * c->N * sizeof(z64), nn = (size_t)c->N * c->N * sizeof(z64);
  const void *din = nullptr, *dproj = nullptr, *drp = nullptr, *dpot = nullptr;
  SDV_TRY(stage_in(c, c->h_in1, inspfs, pb, &din));
  if (projflag && projspfs) SDV_TRY(stage_in(c, c->h_in2, projspfs, pb, &dproj));
  SDV_TRY(stage_in(c, c->h_out2, reducedpot, pb * c->N, &drp));
  if (pot) SDV_TRY(stage_in(c, c->h_pot, pot, (size_t)c->n3 * sizeof(double), &dpot));
  SDV_TRY(ensure(c->h_small, 2 * nn));
  char* dm = (char*)c->h_small.p;
  SDV_HIP_CHECK(hipMemcpyAsync(dm, denmat, nn, hipMemcpyHostToDevice, c->stream));
  if (dentimeflag && invdenmat) SDV_HIP_CHECK(hipMemcpyAsync(dm + nn, invdenmat, nn, hipMemcpyHostToDevice, c->stream));
  SDV_TRY(ensure(c->h_out3, pb));
  SDV_TRY(sdv_actreduced00_dev(c, din, dproj, dm, (dentimeflag && invdenmat) ? dm + nn : nullptr, drp,
                               (const double*)dpot, tdpot, dentimeflag, projflag, timefac_re, timefac_im,
                               c->h_out3.p, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(outspfs, c->h_out3.p, pb, hipMemcpyDeviceToHost, c->stream));
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SDV_OK;
}

int sdv_mult_reducedpot(sdv_ctx* c, int first, int last, const double* inspfs, double* outspfs,
                        const double* reducedpot) {
  SDV_REQUIRE(c && inspfs && outspfs && reducedpot, SDV_ERR_ARG, "null argument");
  if (last < first) return SDV_OK;
  const int nout = last - first + 1;
  const void *di, *dr;
  SDV_TRY(stage_in(c, c->h_in1, inspfs, (size_t)c->n3 * c->N * sizeof(z64), &di));
  SDV_TRY(stage_in(c, c->h_in2, reducedpot, (size_t)c->n3 * c->N * nout * sizeof(z64), &dr));
  SDV_TRY(ensure(c->h_out1, (size_t)c->n3 * nout * sizeof(z64)));
  SDV_TRY(sdv_mult_reducedpot_dev(c, first, last, di, c->h_out1.p, dr, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(outspfs, c->h_out1.p, (size_t)c->n3 * nout * sizeof(z64), hipMemcpyDefault,
                               c->stream));
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SDV_OK;
}

int sdv_get_reducedpot(sdv_ctx* c, const double* tr, const double* twoden, double* rp) {
  SDV_REQUIRE(c && tr && twoden && rp, SDV_ERR_ARG, "null argument");
  const size_t NN = (size_t)c->N * c->N;
  const void *dt, *dg;
  SDV_TRY(stage_in(c, c->h_in1, tr, (size_t)c->n3 * NN * sizeof(z64), &dt));
  SDV_TRY(stage_in(c, c->h_in2, twoden, NN * NN * sizeof(z64), &dg));
  SDV_TRY(ensure(c->h_out1, (size_t)c->n3 * NN * sizeof(z64)));
  SDV_TRY(sdv_get_reducedpot_dev(c, dt, dg, c->h_out1.p, c->stream));
  SDV_HIP_CHECK(hipMemcpyAsync(rp, c->h_out1.p, (size_t)c->n3 * NN * sizeof(z64), hipMemcpyDefault, c->stream));
  SDV_HIP_CHECK(hipStreamSynchronize(c->stream));
  return SDV_OK;
}

}  // extern "C"
