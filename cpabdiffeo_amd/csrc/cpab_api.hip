// C ABI of libcpab_hip.so: context, argument checks, launch configuration and
// kernel dispatch.  Declarations and their reference counterparts are in
// include/cpab_hip.h.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cpab_hip.h"
#include "cpab_kernels.h"
#include "cpab_mcmc.h"
#include "cpab_adjoint.h"
#include "cpab_batch.h"

#pragma clang fp contract(off)

using namespace cpab;

struct cpab_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int max_lds = 65536;
  int n_cu = 256;
  bool prof = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev;  // pending timed launches
  std::vector<int> ev_n;                               // kernel launches each pair brackets
  std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_free;
  int64_t prof_n = 0;
  double prof_ms = 0.0;
  double* scratch = nullptr;  // reduction partials
  size_t scratch_bytes = 0;
  double* ws = nullptr;       // adjoint checkpoints and partials
  size_t ws_bytes = 0;
  WorkItem* items = nullptr;  // work items of cpab_calc_T_items
  size_t items_cap = 0;
  unsigned int* queue = nullptr;
  std::vector<WorkItem> h_items;
  unsigned char* static_scratch = nullptr;  // this device's copy of `scratch` (cpab_calc_T_items_noalloc)
};

// Scratch of cpab_calc_T_items_noalloc, for callers that cannot allocate:
// the launch's queue counter (the first 256 bytes) and up to
// kStaticItems work items.  One copy per GPU, held until the process ends;
// every call on a GPU shares it.  Host code never passes this name to a
// kernel: it is a host-side stand-in; the device address comes from
// hipGetSymbolAddress (cpab_create, once per device).
constexpr int kStaticItems = 4096;
constexpr size_t kStaticQueueBytes = 256;
constexpr size_t kStaticScratchBytes = kStaticQueueBytes + (size_t)kStaticItems * sizeof(WorkItem);
static __device__ __align__(256) unsigned char scratch[kStaticScratchBytes];
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static unsigned char* g_scratch_ptr[kMaxDevices] = {};

namespace {

// The substep length of calc_T: dt / nStepsODE (any positive count).
double dtdiv(double dt, int nS) { return dt / (double)(nS > 0 ? nS : 1); }

int fail(cpab_ctx* ctx, int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  if (ctx) ctx->err = buf;
  return code;
}

#define HIP_OK(ctx, call)                                                                     \
  do {                                                                                        \
    hipError_t e_ = (call);                                                                   \
    if (e_ != hipSuccess) return fail(ctx, CPAB_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)); \
  } while (0)

bool is_pow2(double x) {
  if (!(x > 0.0) || !std::isfinite(x)) return false;
  int e;
  const double m = std::frexp(x, &e);
  if (m != 0.5) return false;
  const double r = 1.0 / x;
  return std::isfinite(r) && r * x == 1.0;
}

// Validate the geometry and fill the lookup constants the way the reference
// kernels compute them inline (cpab/gpu/calc_2Dto2D.cu:55-56,104,124;
// calc_3Dto3D.cu:57-59).
int make_geom(cpab_ctx* ctx, int dim, int tess, int nC0, int nC1, int nC2, double ix, double iy,
              double iz, Geom* g, bool* pow2) {
  if (dim < 1 || dim > 3) return fail(ctx, CPAB_ERR_UNSUPPORTED, "dim must be 1, 2 or 3 (got %d)", dim);
  if (tess != TESS_I && tess != TESS_II) return fail(ctx, CPAB_ERR_ARG, "tess must be 1 (I) or 2 (II), got %d", tess);
  if (tess == TESS_I && dim == 1) return fail(ctx, CPAB_ERR_UNSUPPORTED, "tess I needs dim 2 or 3");
  const int n[3] = {nC0, nC1, nC2};
  const double inc[3] = {ix, iy, iz};
  for (int d = 0; d < dim; d++) {
    if (n[d] < 1) return fail(ctx, CPAB_ERR_ARG, "nC%d must be >= 1 (got %d)", d, n[d]);
    if (!(inc[d] > 0.0) || !std::isfinite(inc[d])) return fail(ctx, CPAB_ERR_ARG, "inc %d must be positive and finite", d);
  }
  std::memset(g, 0, sizeof *g);
  g->nC0 = nC0;
  g->nC1 = nC1;
  g->nC2 = nC2;
  g->inc_x = ix;
  g->inc_y = iy;
  g->inc_z = iz;
  bool p2 = true;
  for (int d = 0; d < dim; d++) p2 = p2 && is_pow2(inc[d]);
  *pow2 = p2;
  g->rinc_x = 1.0 / ix;
  g->rinc_y = 1.0 / iy;
  g->rinc_z = 1.0 / iz;
  g->top_x = nC0 * ix - 0.0000000001;
  g->top_y = nC1 * iy - 0.0000000001;
  g->top_z = nC2 * ix - 0.0000000001;  // the reference clamps z with inc_x (calc_3Dto3D.cu:59)
  g->wall_x = nC0 * ix;
  g->wall_y = nC1 * iy;
  // The type I clamps must leave a non-empty range, or the box index of a
  // clamped point would be negative.
  if (tess == TESS_I && !(g->top_x > 0.0 && g->top_y > 0.0 && (dim < 3 || g->top_z > 0.0)))
    return fail(ctx, CPAB_ERR_ARG, "cells too small: nC*inc must exceed 1e-10");
  return CPAB_OK;
}

int check_nC(cpab_ctx* ctx, int dim, int tess, int nC, int nC0, int nC1, int nC2) {
  int64_t boxes = nC0;
  if (dim >= 2) boxes *= nC1;
  if (dim >= 3) boxes *= nC2;
  const int64_t want = boxes * (tess == TESS_II ? 1 : (dim == 2 ? 4 : 5));
  // the lookups form cell ids with 24-bit multiplies
  if (want >= (1 << 24)) return fail(ctx, CPAB_ERR_UNSUPPORTED, "tessellations of 2^24 cells or more are not supported");
  if (nC != want)
    return fail(ctx, CPAB_ERR_ARG, "nC=%d does not match the tessellation (%lld cells)", nC, (long long)want);
  return CPAB_OK;
}

bool overlap(const void* a, size_t na, const void* b, size_t nb) {
  const char* x = static_cast<const char*>(a);
  const char* y = static_cast<const char*>(b);
  return x < y + nb && y < x + na;
}

// Integration launch shape: 64-lane waves, each owning a contiguous range of
// points; enough waves for several per SIMD.
struct Shape {
  int block;
  int ptsPerWave;
  int64_t grid;
  size_t lds;
  bool use_lds;
  bool fast;  // enough waves per SIMD to be throughput-bound: the FAST 2D lookup pays
};

Shape integrate_shape(const cpab_ctx* ctx, int64_t nPts, int nC, int ne, int ntables) {
  Shape s;
  const size_t tbytes = (size_t)ntables * nC * ne * sizeof(double);
  s.use_lds = tbytes <= (size_t)ctx->max_lds;
  s.lds = s.use_lds ? tbytes : 0;
  s.block = (s.lds > 32768) ? 1024 : 256;
  // waves wanted: 8 per SIMD (4 SIMDs per CU); each lane takes ceil of the rest
  const int64_t want_waves = (int64_t)ctx->n_cu * 4 * 8;
  int64_t per_lane = (nPts + want_waves * 64 - 1) / (want_waves * 64);
  if (per_lane < 1) per_lane = 1;
  s.ptsPerWave = (int)(per_lane * 64);
  const int64_t waves = (nPts + s.ptsPerWave - 1) / s.ptsPerWave;
  const int wpb = s.block / 64;
  s.grid = (waves + wpb - 1) / wpb;
  s.fast = waves >= (int64_t)ctx->n_cu * 4 * 4;  // four waves per SIMD or more
  return s;
}

template <typename K>
void allow_lds(K kernel, size_t bytes) {
  if (bytes > 65536) (void)hipFuncSetAttribute((const void*)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

int prof_begin(cpab_ctx* ctx, hipEvent_t* a, hipEvent_t* b) {
  if (!ctx->prof) return CPAB_OK;
  if (ctx->ev_free.empty()) {
    hipEvent_t x, y;
    HIP_OK(ctx, hipEventCreate(&x));
    HIP_OK(ctx, hipEventCreate(&y));
    ctx->ev_free.push_back({x, y});
  }
  *a = ctx->ev_free.back().first;
  *b = ctx->ev_free.back().second;
  ctx->ev_free.pop_back();
  HIP_OK(ctx, hipEventRecord(*a, ctx->stream));
  return CPAB_OK;
}

int prof_end(cpab_ctx* ctx, hipEvent_t a, hipEvent_t b, int launches = 1) {
  if (!ctx->prof) return CPAB_OK;
  HIP_OK(ctx, hipEventRecord(b, ctx->stream));
  ctx->ev.push_back({a, b});
  ctx->ev_n.push_back(launches);
  return CPAB_OK;
}

template <int DIM, int TESS, bool POW2, bool LDS>
int launch_integrate_t(cpab_ctx* ctx, int mode, const Shape& s, const double* pos0, double* out,
                       const double* Trels, const double* As, double h, int nT, int nS, int64_t nPts,
                       const Geom& g, int nC, int32_t* nfb, const PrepSrc& ps) {
  if constexpr (DIM == 2 && TESS == TESS_I && POW2) {
    if (mode == 0 && s.fast) {
      auto k = k_integrate<DIM, TESS, POW2, LDS, 0, true>;
      allow_lds(k, s.lds);
      hipLaunchKernelGGL(k, dim3((unsigned)s.grid), dim3(s.block), s.lds, ctx->stream, pos0, out, Trels, As, h, nT,
                         nS, nPts, g, nC, s.ptsPerWave, nfb, ps);
      HIP_OK(ctx, hipGetLastError());
      return CPAB_OK;
    }
  }
  if (mode == 0) {
    auto k = k_integrate<DIM, TESS, POW2, LDS, 0>;
    allow_lds(k, s.lds);
    hipLaunchKernelGGL(k, dim3((unsigned)s.grid), dim3(s.block), s.lds, ctx->stream, pos0, out, Trels, As, h, nT, nS,
                       nPts, g, nC, s.ptsPerWave, nfb, ps);
  } else {
    auto k = k_integrate<DIM, TESS, POW2, LDS, 1>;
    allow_lds(k, s.lds);
    hipLaunchKernelGGL(k, dim3((unsigned)s.grid), dim3(s.block), s.lds, ctx->stream, pos0, out, Trels, As, h, nT, nS,
                       nPts, g, nC, s.ptsPerWave, nfb, ps);
  }
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

template <int DIM, int TESS>
int launch_integrate_d(cpab_ctx* ctx, int mode, bool pow2, const Shape& s, const double* pos0, double* out,
                       const double* Trels, const double* As, double h, int nT, int nS, int64_t nPts,
                       const Geom& g, int nC, int32_t* nfb, const PrepSrc& ps) {
  if (pow2) {
    if (s.use_lds) return launch_integrate_t<DIM, TESS, true, true>(ctx, mode, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps);
    return launch_integrate_t<DIM, TESS, true, false>(ctx, mode, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps);
  }
  if (s.use_lds) return launch_integrate_t<DIM, TESS, false, true>(ctx, mode, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps);
  return launch_integrate_t<DIM, TESS, false, false>(ctx, mode, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps);
}

// ps.raw != nullptr (the prep fused into the launch) only with s.use_lds.
int launch_integrate(cpab_ctx* ctx, int mode, int dim, int tess, bool pow2, const Shape& s, const double* pos0,
                     double* out, const double* Trels, const double* As, double h, int nT, int nS, int64_t nPts,
                     const Geom& g, int nC, int32_t* nfb, const PrepSrc& ps) {
  if (dim == 1) return launch_integrate_d<1, TESS_II>(ctx, mode, pow2, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps);
  if (dim == 2) {
    if (tess == TESS_I) return launch_integrate_d<2, TESS_I>(ctx, mode, pow2, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps);
    return launch_integrate_d<2, TESS_II>(ctx, mode, pow2, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps);
  }
  if (tess == TESS_I) return launch_integrate_d<3, TESS_I>(ctx, mode, pow2, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps);
  return launch_integrate_d<3, TESS_II>(ctx, mode, pow2, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps);
}

// Generic 3-way dispatch for the simple per-point kernels.
#define DISPATCH_GEOM(dim, tess, pow2, ...)                                          \
  do {                                                                                 \
    if ((dim) == 1) {                                                                  \
      if (pow2) { constexpr int D_ = 1, T_ = TESS_II; constexpr bool P_ = true; __VA_ARGS__; } \
      else { constexpr int D_ = 1, T_ = TESS_II; constexpr bool P_ = false; __VA_ARGS__; }    \
    } else if ((dim) == 2) {                                                           \
      if ((tess) == TESS_I) {                                                          \
        if (pow2) { constexpr int D_ = 2, T_ = TESS_I; constexpr bool P_ = true; __VA_ARGS__; } \
        else { constexpr int D_ = 2, T_ = TESS_I; constexpr bool P_ = false; __VA_ARGS__; }   \
      } else {                                                                         \
        if (pow2) { constexpr int D_ = 2, T_ = TESS_II; constexpr bool P_ = true; __VA_ARGS__; } \
        else { constexpr int D_ = 2, T_ = TESS_II; constexpr bool P_ = false; __VA_ARGS__; }  \
      }                                                                                \
    } else {                                                                           \
      if ((tess) == TESS_I) {                                                          \
        if (pow2) { constexpr int D_ = 3, T_ = TESS_I; constexpr bool P_ = true; __VA_ARGS__; } \
        else { constexpr int D_ = 3, T_ = TESS_I; constexpr bool P_ = false; __VA_ARGS__; }   \
      } else {                                                                         \
        if (pow2) { constexpr int D_ = 3, T_ = TESS_II; constexpr bool P_ = true; __VA_ARGS__; } \
        else { constexpr int D_ = 3, T_ = TESS_II; constexpr bool P_ = false; __VA_ARGS__; }  \
      }                                                                                \
    }                                                                                  \
  } while (0)

int ensure_scratch(cpab_ctx* ctx, size_t bytes) {
  if (ctx->scratch_bytes >= bytes) return CPAB_OK;
  if (ctx->scratch) {
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OK(ctx, hipFree(ctx->scratch));
    ctx->scratch = nullptr;
    ctx->scratch_bytes = 0;
  }
  HIP_OK(ctx, hipMalloc(&ctx->scratch, bytes));
  ctx->scratch_bytes = bytes;
  return CPAB_OK;
}

}  // namespace

// ====================================================================== API
extern "C" {

const char* cpab_version(void) { return "cpab_hip 0.1 gfx950 (fp64 integration, no fp contraction)"; }

int cpab_create(int device, cpab_ctx** out) {
  if (!out) return CPAB_ERR_ARG;
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n <= 0) return CPAB_ERR_HIP;
  if (device < 0 || device >= n) return CPAB_ERR_ARG;
  if (hipSetDevice(device) != hipSuccess) return CPAB_ERR_HIP;
  cpab_ctx* ctx = new (std::nothrow) cpab_ctx();
  if (!ctx) return CPAB_ERR_NOMEM;
  ctx->device = device;
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && v > 0)
    ctx->max_lds = v;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && v > 0) ctx->n_cu = v;
  if (device < kMaxDevices) {  // the device address of this GPU's `scratch`, looked up once per device
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
      void* p = nullptr;
      if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        delete ctx;
        return CPAB_ERR_HIP;
      }
      g_scratch_ptr[device] = static_cast<unsigned char*>(p);
    }
    ctx->static_scratch = g_scratch_ptr[device];
  }
  *out = ctx;
  return CPAB_OK;
}

int cpab_destroy(cpab_ctx* ctx) {
  if (!ctx) return CPAB_OK;
  (void)hipSetDevice(ctx->device);
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  else (void)hipDeviceSynchronize();
  for (auto& p : ctx->ev) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
  for (auto& p : ctx->ev_free) { (void)hipEventDestroy(p.first); (void)hipEventDestroy(p.second); }
  if (ctx->scratch) (void)hipFree(ctx->scratch);
  if (ctx->ws) (void)hipFree(ctx->ws);
  if (ctx->items) (void)hipFree(ctx->items);
  if (ctx->queue) (void)hipFree(ctx->queue);
  delete ctx;
  return CPAB_OK;
}

const char* cpab_last_error(const cpab_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int cpab_set_stream(cpab_ctx* ctx, void* stream) {
  if (!ctx) return CPAB_ERR_ARG;
  ctx->stream = static_cast<hipStream_t>(stream);
  return CPAB_OK;
}

int cpab_synchronize(cpab_ctx* ctx) {
  if (!ctx) return CPAB_ERR_ARG;
  HIP_OK(ctx, hipSetDevice(ctx->device));
  HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
  return CPAB_OK;
}

int cpab_prep_trels(cpab_ctx* ctx, int dim, int nC, const double* As, int sign, double dt, double* sA_out,
                    double* T_out) {
  if (!ctx) return CPAB_ERR_ARG;
  if (dim < 1 || dim > 3) return fail(ctx, CPAB_ERR_UNSUPPORTED, "dim must be 1, 2 or 3");
  if (nC < 1) return fail(ctx, CPAB_ERR_ARG, "nC must be >= 1");
  if (sign != 1 && sign != -1) return fail(ctx, CPAB_ERR_ARG, "sign must be +1 or -1");
  if (!As || !T_out) return fail(ctx, CPAB_ERR_ARG, "null table");
  const size_t nb = (size_t)nC * dim * (dim + 1) * sizeof(double);
  if (overlap(As, nb, T_out, nb) || (sA_out && overlap(sA_out, nb, T_out, nb)))
    return fail(ctx, CPAB_ERR_ALIAS, "Trels_out overlaps an input");
  const dim3 grid((nC + 255) / 256, 1), block(256);
  if (dim == 1) hipLaunchKernelGGL(k_prep<1>, grid, block, 0, ctx->stream, As, nC, sign, dt, sA_out, T_out, 12, (int64_t)0);
  else if (dim == 2) hipLaunchKernelGGL(k_prep<2>, grid, block, 0, ctx->stream, As, nC, sign, dt, sA_out, T_out, 12, (int64_t)0);
  else hipLaunchKernelGGL(k_prep<3>, grid, block, 0, ctx->stream, As, nC, sign, dt, sA_out, T_out, 12, (int64_t)0);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

int cpab_expm(cpab_ctx* ctx, int dim, int nC, const double* As, double* Ts, int p) {
  if (!ctx) return CPAB_ERR_ARG;
  if (dim != 2 && dim != 3) return fail(ctx, CPAB_ERR_UNSUPPORTED, "expm: dim must be 2 or 3");
  if (nC < 1 || p < 1) return fail(ctx, CPAB_ERR_ARG, "expm: nC and p must be >= 1");
  const dim3 grid((nC + 511) / 512), block(512);
  if (dim == 2) hipLaunchKernelGGL(k_expm<2>, grid, block, 0, ctx->stream, As, Ts, nC, p);
  else hipLaunchKernelGGL(k_expm<3>, grid, block, 0, ctx->stream, As, Ts, nC, p);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

static int integrate_common(cpab_ctx* ctx, int mode, int dim, int tess, int nC, const double* pos0, double* out,
                            const double* Trels, const double* As, double dt, int nT, int nS, int64_t nPts,
                            int nC0, int nC1, int nC2, double ix, double iy, double iz, int32_t* nfb,
                            const double* raw = nullptr, int sign = 1) {
  if (!ctx) return CPAB_ERR_ARG;
  Geom g;
  bool pow2;
  int r = make_geom(ctx, dim, tess, nC0, nC1, nC2, ix, iy, iz, &g, &pow2);
  if (r) return r;
  if ((r = check_nC(ctx, dim, tess, nC, nC0, nC1, nC2))) return r;
  if (nPts < 0) return fail(ctx, CPAB_ERR_ARG, "nPts must be >= 0");
  if (nT < 0 || nS < 0) return fail(ctx, CPAB_ERR_ARG, "nTimeSteps and nStepsODE must be >= 0");
  if (nPts == 0) return CPAB_OK;
  if (!pos0 || !out || !Trels || !As) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const size_t pb = (size_t)nPts * dim * sizeof(double);
  const size_t ob = mode == 0 ? pb : pb * (size_t)(nT > 0 ? nT : 1);
  if (overlap(pos0, pb, out, ob)) return fail(ctx, CPAB_ERR_ALIAS, "pts and out cannot point to the same memory");
  const int ne = dim * (dim + 1);
  const Shape s = integrate_shape(ctx, nPts, nC, ne, 2);
  if (s.grid > 0x7fffffffLL) return fail(ctx, CPAB_ERR_ARG, "too many points for one launch");
  // h = dt/double(nStepsOdeSolver) (calc_2Dto2D.cu:383)
  const double h = dt / (double)nS;
  // the prep fused into the launch when the tables fit LDS; else its own launch first
  PrepSrc ps{nullptr, nullptr, nullptr, 0.0, 1};
  if (raw) {
    if (s.use_lds) {
      ps = PrepSrc{raw, const_cast<double*>(As), const_cast<double*>(Trels), dt, sign};
    } else if ((r = cpab_prep_trels(ctx, dim, nC, raw, sign, dt, const_cast<double*>(As), const_cast<double*>(Trels)))) {
      return r;
    }
  }
  hipEvent_t a = nullptr, b = nullptr;
  if ((r = prof_begin(ctx, &a, &b))) return r;
  if ((r = launch_integrate(ctx, mode, dim, tess, pow2, s, pos0, out, Trels, As, h, nT, nS, nPts, g, nC, nfb, ps))) return r;
  return prof_end(ctx, a, b);
}

int cpab_calc_T(cpab_ctx* ctx, int dim, int tess, int nC, const double* pos0, double* pos, const double* Trels,
                const double* As, double dt, int nT, int nS, int64_t nPts, int nC0, int nC1, int nC2, double ix,
                double iy, double iz, int32_t* n_fallback) {
  return integrate_common(ctx, 0, dim, tess, nC, pos0, pos, Trels, As, dt, nT, nS, nPts, nC0, nC1, nC2, ix, iy, iz,
                          n_fallback);
}

int cpab_calc_T_prep(cpab_ctx* ctx, int dim, int tess, int nC, const double* pos0, double* pos, const double* As,
                     int sign, double dt, int nT, int nS, int64_t nPts, int nC0, int nC1, int nC2, double ix,
                     double iy, double iz, double* signedAs_out, double* Trels_out, int32_t* n_fallback) {
  if (!ctx) return CPAB_ERR_ARG;
  if (sign != 1 && sign != -1) return fail(ctx, CPAB_ERR_ARG, "sign must be +1 or -1");
  if (!As || !signedAs_out || !Trels_out) return fail(ctx, CPAB_ERR_ARG, "null table");
  if (nC < 1) return fail(ctx, CPAB_ERR_ARG, "nC must be >= 1");
  const size_t nb = (size_t)nC * dim * (dim + 1) * sizeof(double);
  if (overlap(As, nb, Trels_out, nb) || overlap(As, nb, signedAs_out, nb) || overlap(signedAs_out, nb, Trels_out, nb))
    return fail(ctx, CPAB_ERR_ALIAS, "the table outputs overlap each other or As");
  if (nPts == 0) return cpab_prep_trels(ctx, dim, nC, As, sign, dt, signedAs_out, Trels_out);  // the tables still
  return integrate_common(ctx, 0, dim, tess, nC, pos0, pos, Trels_out, signedAs_out, dt, nT, nS, nPts, nC0, nC1, nC2,
                          ix, iy, iz, n_fallback, As, sign);
}

int cpab_calc_trajectory(cpab_ctx* ctx, int dim, int tess, int nC, const double* pos0, double* traj,
                         const double* Trels, const double* As, double dt, int nT, int nS, int64_t nPts, int nC0,
                         int nC1, int nC2, double ix, double iy, double iz) {
  return integrate_common(ctx, 1, dim, tess, nC, pos0, traj, Trels, As, dt, nT, nS, nPts, nC0, nC1, nC2, ix, iy, iz,
                          nullptr);
}

int cpab_calc_T_simple(cpab_ctx* ctx, int dim, int tess, int nC, const double* pos0, double* pos, const double* As,
                       double dt, int nT, int nS, int64_t nPts, int nC0, int nC1, int nC2, double ix, double iy,
                       double iz) {
  if (!ctx) return CPAB_ERR_ARG;
  Geom g;
  bool pow2;
  int r = make_geom(ctx, dim, tess, nC0, nC1, nC2, ix, iy, iz, &g, &pow2);
  if (r) return r;
  if ((r = check_nC(ctx, dim, tess, nC, nC0, nC1, nC2))) return r;
  if (nPts < 0 || nT < 0 || nS < 0) return fail(ctx, CPAB_ERR_ARG, "negative size");
  if (nPts == 0) return CPAB_OK;
  if (!pos0 || !pos || !As) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const size_t pb = (size_t)nPts * dim * sizeof(double);
  if (overlap(pos0, pb, pos, pb)) return fail(ctx, CPAB_ERR_ALIAS, "pts and out cannot point to the same memory");
  const int64_t steps64 = (int64_t)nS * nT;
  if (steps64 > 0x7fffffffLL) return fail(ctx, CPAB_ERR_ARG, "nStepsODE*nTimeSteps overflows int");
  const double h = dt / (double)nS;
  const size_t tb = (size_t)nC * dim * (dim + 1) * sizeof(double);
  const bool lds = tb <= (size_t)ctx->max_lds;
  const int block = 256;
  const dim3 grid((unsigned)((nPts + block - 1) / block));
  hipEvent_t a = nullptr, b = nullptr;
  if ((r = prof_begin(ctx, &a, &b))) return r;
  DISPATCH_GEOM(dim, tess, pow2, {
    if (lds) {
      auto k = k_simple<D_, T_, P_, true>;
      allow_lds(k, tb);
      hipLaunchKernelGGL(k, grid, dim3(block), tb, ctx->stream, pos0, pos, As, h, (int)steps64, nPts, g, nC);
    } else {
      hipLaunchKernelGGL((k_simple<D_, T_, P_, false>), grid, dim3(block), 0, ctx->stream, pos0, pos, As, h,
                         (int)steps64, nPts, g, nC);
    }
  });
  HIP_OK(ctx, hipGetLastError());
  return prof_end(ctx, a, b);
}

int cpab_calc_grad_theta(cpab_ctx* ctx, int dim, int tess, int nC, const double* pos0, double* pos,
                         const double* As, const double* Bs, double* grad, int d, double dt, int nT, int nS,
                         int64_t nPts, int nC0, int nC1, int nC2, double ix, double iy, double iz) {
  if (!ctx) return CPAB_ERR_ARG;
  Geom g;
  bool pow2;
  int r = make_geom(ctx, dim, tess, nC0, nC1, nC2, ix, iy, iz, &g, &pow2);
  if (r) return r;
  if ((r = check_nC(ctx, dim, tess, nC, nC0, nC1, nC2))) return r;
  if (nPts < 0 || nT < 0 || nS < 0 || d < 1) return fail(ctx, CPAB_ERR_ARG, "bad sizes");
  if (nPts == 0) return CPAB_OK;
  if (!pos0 || !pos || !As || !Bs || !grad) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const size_t pb = (size_t)nPts * dim * sizeof(double);
  if (overlap(pos0, pb, pos, pb)) return fail(ctx, CPAB_ERR_ALIAS, "pts and out cannot point to the same memory");
  const int64_t steps64 = (int64_t)nS * nT;
  if (steps64 > 0x7fffffffLL) return fail(ctx, CPAB_ERR_ARG, "nStepsODE*nTimeSteps overflows int");
  const int ne = dim * (dim + 1);
  const double h = dt / (double)nS;
  // columns per thread: 8 when As + 8 basis tables fit in LDS, else 2, else
  // 2 with the tables read from global memory
  const size_t lds8 = (size_t)9 * nC * ne * sizeof(double), lds2 = (size_t)3 * nC * ne * sizeof(double);
  const int jc = lds8 <= (size_t)ctx->max_lds ? 8 : 2;
  const bool in_lds = jc == 8 || lds2 <= (size_t)ctx->max_lds;
  const size_t lds = !in_lds ? 0 : (jc == 8 ? lds8 : lds2);
  const dim3 grid((unsigned)((nPts + 255) / 256), (unsigned)((d + jc - 1) / jc));
  DISPATCH_GEOM(dim, tess, pow2, {
    if (jc == 8) {
      auto k = k_grad_theta<D_, T_, P_, 8, true>;
      allow_lds(k, lds);
      hipLaunchKernelGGL(k, grid, dim3(256), lds, ctx->stream, pos0, pos, As, Bs, grad, d, h, (int)steps64, nPts,
                         g, nC);
    } else if (in_lds) {
      auto k = k_grad_theta<D_, T_, P_, 2, true>;
      allow_lds(k, lds);
      hipLaunchKernelGGL(k, grid, dim3(256), lds, ctx->stream, pos0, pos, As, Bs, grad, d, h, (int)steps64, nPts,
                         g, nC);
    } else {
      hipLaunchKernelGGL((k_grad_theta<D_, T_, P_, 2, false>), grid, dim3(256), 0, ctx->stream, pos0, pos, As, Bs,
                         grad, d, h, (int)steps64, nPts, g, nC);
    }
  });
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

int cpab_calc_v(cpab_ctx* ctx, int dim, int tess, int nC, const double* pts, double* vel, const double* As,
                int64_t nPts, int nC0, int nC1, int nC2, double ix, double iy, double iz) {
  if (!ctx) return CPAB_ERR_ARG;
  Geom g;
  bool pow2;
  int r = make_geom(ctx, dim, tess, nC0, nC1, nC2, ix, iy, iz, &g, &pow2);
  if (r) return r;
  if ((r = check_nC(ctx, dim, tess, nC, nC0, nC1, nC2))) return r;
  if (nPts < 0) return fail(ctx, CPAB_ERR_ARG, "nPts must be >= 0");
  if (nPts == 0) return CPAB_OK;
  if (!pts || !vel || !As) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const size_t pb = (size_t)nPts * dim * sizeof(double);
  if (overlap(pts, pb, vel, pb)) return fail(ctx, CPAB_ERR_ALIAS, "pts and v cannot point to the same memory");
  const size_t tb = (size_t)nC * dim * (dim + 1) * sizeof(double);
  const bool lds = tb <= (size_t)ctx->max_lds;
  const int block = 256;
  const dim3 grid((unsigned)((nPts + block - 1) / block));
  DISPATCH_GEOM(dim, tess, pow2, {
    if (lds) {
      auto k = k_velocity<D_, T_, P_, true>;
      allow_lds(k, tb);
      hipLaunchKernelGGL(k, grid, dim3(block), tb, ctx->stream, pts, vel, As, nPts, g, nC);
    } else {
      hipLaunchKernelGGL((k_velocity<D_, T_, P_, false>), grid, dim3(block), 0, ctx->stream, pts, vel, As, nPts, g,
                         nC);
    }
  });
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

int cpab_calc_cell_idx(cpab_ctx* ctx, int dim, int tess, const double* pts, int32_t* cell, int64_t nPts, int nC0,
                       int nC1, int nC2, double ix, double iy, double iz) {
  if (!ctx) return CPAB_ERR_ARG;
  Geom g;
  bool pow2;
  int r = make_geom(ctx, dim, tess, nC0, nC1, nC2, ix, iy, iz, &g, &pow2);
  if (r) return r;
  if (nPts < 0) return fail(ctx, CPAB_ERR_ARG, "nPts must be >= 0");
  if (nPts == 0) return CPAB_OK;
  if (!pts || !cell) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const int block = 256;
  const dim3 grid((unsigned)((nPts + block - 1) / block));
  DISPATCH_GEOM(dim, tess, pow2, {
    hipLaunchKernelGGL((k_cell_idx<D_, T_, P_>), grid, dim3(block), 0, ctx->stream, pts, cell, nPts, g);
  });
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

int cpab_remap2d(cpab_ctx* ctx, const double* pts, const float* img, float* out, int H, int W) {
  if (!ctx) return CPAB_ERR_ARG;
  if (H < 1 || W < 1) return fail(ctx, CPAB_ERR_ARG, "image must be non-empty");
  if (!pts || !img || !out) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const int64_t n = (int64_t)H * W;
  if (overlap(img, n * 4, out, n * 4)) return fail(ctx, CPAB_ERR_ALIAS, "img and out cannot share memory");
  hipLaunchKernelGGL(k_remap2d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, pts, img, out, H, W);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

int cpab_remap3d(cpab_ctx* ctx, const double* pts, const float* vol, float* out, int D, int H, int W) {
  if (!ctx) return CPAB_ERR_ARG;
  if (D < 1 || H < 1 || W < 1) return fail(ctx, CPAB_ERR_ARG, "volume must be non-empty");
  if (!pts || !vol || !out) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const int64_t n = (int64_t)D * H * W;
  if (overlap(vol, n * 4, out, n * 4)) return fail(ctx, CPAB_ERR_ALIAS, "vol and out cannot share memory");
  hipLaunchKernelGGL(k_remap3d, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, pts, vol, out, D, H,
                     W);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

int cpab_ssd_ll(cpab_ctx* ctx, const float* x, const float* y, int64_t n, double sigma, double* ll) {
  if (!ctx) return CPAB_ERR_ARG;
  if (n < 1) return fail(ctx, CPAB_ERR_ARG, "n must be >= 1");
  if (!(sigma > 0.0)) return fail(ctx, CPAB_ERR_ARG, "sigma must be positive");
  if (!x || !y || !ll) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  int nb = (int)((n + 255) / 256);
  if (nb > 1024) nb = 1024;
  int r = ensure_scratch(ctx, (size_t)nb * sizeof(double));
  if (r) return r;
  hipLaunchKernelGGL(k_ssd_partial, dim3(nb), dim3(256), 0, ctx->stream, x, y, n, ctx->scratch);
  hipLaunchKernelGGL(k_reduce_ll, dim3(1), dim3(64), 0, ctx->stream, ctx->scratch, nb, sigma, ll);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

// ------------------------------------------------------------- adjoint
}  // extern "C"

namespace {

int ensure_ws(cpab_ctx* ctx, size_t bytes) {
  if (ctx->ws_bytes >= bytes) return CPAB_OK;
  if (ctx->ws) {
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    HIP_OK(ctx, hipFree(ctx->ws));
    ctx->ws = nullptr;
    ctx->ws_bytes = 0;
  }
  hipError_t e = hipMalloc(&ctx->ws, bytes);
  if (e != hipSuccess) return fail(ctx, CPAB_ERR_NOMEM, "hipMalloc(%zu) for the adjoint workspace: %s", bytes,
                                   hipGetErrorString(e));
  ctx->ws_bytes = bytes;
  return CPAB_OK;
}

// Shared by the adjoint entries: checks, workspace, launch, reductions.
// Trels non-null: the calc_T map (Trel proposals + RK2 redos), else
// calc_T_simple's RK2.
int adjoint2d(cpab_ctx* ctx, int mode, int tess, int nC, int nInst, int64_t nPts, int H, int W, const double* pos0,
              const double* lam, const float* src, const float* dst, const double* Trels, const double* As, double dt,
              int nT, int nS, int nC0, int nC1, double ix, double iy, double* pos_out, double* loss, double* gradA,
              uint32_t* sig, double* nredo) {
  Geom g;
  bool pow2;
  int r = make_geom(ctx, 2, tess, nC0, nC1, 0, ix, iy, 1.0, &g, &pow2);
  if (r) return r;
  if ((r = check_nC(ctx, 2, tess, nC, nC0, nC1, 0))) return r;
  // (the tables' LDS bound below keeps nC far lower; the kernel packs cells in 16 bits)
  if (nC > 65535) return fail(ctx, CPAB_ERR_UNSUPPORTED, "the adjoint takes at most 65535 cells");
  if (nInst < 1 || nInst > 65535) return fail(ctx, CPAB_ERR_ARG, "nInst must be in [1, 65535]");
  if (nPts < 0 || nT < 0 || nS < 0) return fail(ctx, CPAB_ERR_ARG, "bad sizes");
  if (nS > ADJ_NSMAX) return fail(ctx, CPAB_ERR_UNSUPPORTED, "nStepsODE > %d is not supported by the adjoint", ADJ_NSMAX);
  if (!As || !gradA) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const bool fwd = Trels != nullptr;
  const int NT = nC * 6, NG = fwd ? 2 : 1;
  // block of 4, 2 or 1 waves: the per-wave gradient tables must fit in LDS
  int waves = 4;
  while (waves > 1 && adj_lds_doubles(nC, fwd, waves) * sizeof(double) > (size_t)ctx->max_lds) waves >>= 1;
  const size_t lds = adj_lds_doubles(nC, fwd, waves) * sizeof(double);
  if (lds > (size_t)ctx->max_lds) return fail(ctx, CPAB_ERR_UNSUPPORTED, "tables exceed LDS");
  if (nPts == 0 || nT == 0) {  // the identity map: zero gradient (and loss of the unwarped images in MODE 1)
    if (mode == 1) return fail(ctx, CPAB_ERR_UNSUPPORTED, "empty integration in the SSD adjoint");
    HIP_OK(ctx, hipMemsetAsync(gradA, 0, sizeof(double) * NT * nInst, ctx->stream));
    if (pos_out && nPts > 0)
      HIP_OK(ctx, hipMemcpyAsync(pos_out, pos0, sizeof(double) * 2 * nPts * nInst, hipMemcpyDeviceToDevice,
                                 ctx->stream));
    return CPAB_OK;
  }
  const int block = 64 * waves;
  const int64_t nbx = (nPts + block - 1) / block;
  if (nbx > 0x7fffffffLL) return fail(ctx, CPAB_ERR_ARG, "too many points");
  const size_t ck = (size_t)nInst * nT * nPts * 2, pa = (size_t)nInst * nbx * (NG * NT + 2);
  const size_t gt = fwd ? (size_t)nInst * NT : 0;
  const size_t rb = fwd ? (size_t)nInst * ((nT + 63) / 64) * nPts : 0;  // 64-bit words
  if ((r = ensure_ws(ctx, (ck + pa + gt + rb) * sizeof(double)))) return r;
  AdjDesc D{};
  D.nInst = nInst;
  D.nC = nC;
  D.nPts = nPts;
  D.H = H;
  D.W = W;
  D.As = As;
  D.Trels = Trels;
  D.pos0 = pos0;
  D.lam = lam;
  D.pos_out = pos_out;
  D.src = src;
  D.dst = dst;
  D.sig = sig;
  D.ckpt = ctx->ws;
  D.partial = ctx->ws + ck;
  D.redo_bits = fwd ? reinterpret_cast<uint64_t*>(ctx->ws + ck + pa + gt) : nullptr;
  const bool ns10 = nS <= 10;  // the substep starts of a redone step in 10 registers pairs, else 16
  D.h = dt / (double)nS;
  D.nT = nT;
  D.nS = nS;
  D.tile8 = (mode == 1 && H % 8 == 0 && W % 8 == 0) ? 1 : 0;
  D.g = g;
  // instances' first blocks (the domain corner's long chain) first, while
  // the blocks per instance fit the grid's y dimension
  D.inst_major = nbx <= 65535 ? 1 : 0;
  const dim3 grid = D.inst_major ? dim3((unsigned)nInst, (unsigned)nbx) : dim3((unsigned)nbx, (unsigned)nInst);
  hipEvent_t a = nullptr, b = nullptr;
  if ((r = prof_begin(ctx, &a, &b))) return r;
#define ADJ_LAUNCH4(T_, P_, M_, F_)                                        \
  do {                                                                     \
    if (ns10) {                                                            \
      auto k = k_adjoint2d<T_, P_, M_, F_, 10>;                            \
      allow_lds(k, lds);                                                   \
      hipLaunchKernelGGL(k, grid, dim3(block), lds, ctx->stream, D);       \
    } else {                                                               \
      auto k = k_adjoint2d<T_, P_, M_, F_, ADJ_NSMAX>;                     \
      allow_lds(k, lds);                                                   \
      hipLaunchKernelGGL(k, grid, dim3(block), lds, ctx->stream, D);       \
    }                                                                      \
  } while (0)
#define ADJ_LAUNCH3(T_, P_, M_)                                            \
  do {                                                                     \
    if (fwd) ADJ_LAUNCH4(T_, P_, M_, true);                                \
    else ADJ_LAUNCH4(T_, P_, M_, false);                                   \
  } while (0)
#define ADJ_LAUNCH(T_, P_)                   \
  do {                                       \
    if (mode == 0) ADJ_LAUNCH3(T_, P_, 0);   \
    else ADJ_LAUNCH3(T_, P_, 1);             \
  } while (0)
  if (tess == TESS_I) {
    if (pow2) ADJ_LAUNCH(TESS_I, true);
    else ADJ_LAUNCH(TESS_I, false);
  } else {
    if (pow2) ADJ_LAUNCH(TESS_II, true);
    else ADJ_LAUNCH(TESS_II, false);
  }
#undef ADJ_LAUNCH
#undef ADJ_LAUNCH3
#undef ADJ_LAUNCH4
  HIP_OK(ctx, hipGetLastError());
  if ((r = prof_end(ctx, a, b))) return r;
  double* gradT = fwd ? ctx->ws + ck + pa : nullptr;
  hipLaunchKernelGGL(k_adj_reduce, dim3((unsigned)((NG * NT + 2 + 63) / 64), (unsigned)nInst), dim3(64), 0, ctx->stream,
                     D.partial, (int)nbx, NT, NG, gradA, gradT, loss, nredo);
  if (fwd) {
    const int64_t n = (int64_t)nInst * nC;
    hipLaunchKernelGGL(k_adj_expm, dim3((unsigned)((n + 63) / 64)), dim3(64), 0, ctx->stream, As, gradT, gradA, nC,
                       nInst, dt);
  }
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

}  // namespace

extern "C" {

int cpab_calc_T_simple_vjp(cpab_ctx* ctx, int tess, int nC, int nInst, const double* pos0, const double* lam,
                           const double* As, double dt, int nTimeSteps, int nStepsODE, int64_t nPts, int nC0, int nC1,
                           double inc_x, double inc_y, double* pos_out, double* gradA) {
  if (!ctx) return CPAB_ERR_ARG;
  if (nPts > 0 && (!pos0 || !lam)) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  return adjoint2d(ctx, 0, tess, nC, nInst, nPts, 0, 0, pos0, lam, nullptr, nullptr, nullptr, As, dt, nTimeSteps,
                   nStepsODE, nC0, nC1, inc_x, inc_y, pos_out, nullptr, gradA, nullptr, nullptr);
}

int cpab_ssd_grad2d(cpab_ctx* ctx, int tess, int nC, int nInst, const double* As, const float* src, const float* dst,
                    int H, int W, double dt, int nTimeSteps, int nStepsODE, int nC0, int nC1, double inc_x,
                    double inc_y, double* pos_out, double* loss, double* gradA) {
  if (!ctx) return CPAB_ERR_ARG;
  if (H < 1 || W < 1) return fail(ctx, CPAB_ERR_ARG, "images must be non-empty");
  if (!src || !dst || !loss) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  return adjoint2d(ctx, 1, tess, nC, nInst, (int64_t)H * W, H, W, nullptr, nullptr, src, dst, nullptr, As, dt,
                   nTimeSteps, nStepsODE, nC0, nC1, inc_x, inc_y, pos_out, loss, gradA, nullptr, nullptr);
}

int cpab_calc_T_vjp(cpab_ctx* ctx, int tess, int nC, int nInst, const double* pos0, const double* lam,
                    const double* Trels, const double* As, double dt, int nTimeSteps, int nStepsODE, int64_t nPts,
                    int nC0, int nC1, double inc_x, double inc_y, double* pos_out, double* gradA) {
  if (!ctx) return CPAB_ERR_ARG;
  if (!Trels) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  if (nPts > 0 && (!pos0 || !lam)) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  return adjoint2d(ctx, 0, tess, nC, nInst, nPts, 0, 0, pos0, lam, nullptr, nullptr, Trels, As, dt, nTimeSteps,
                   nStepsODE, nC0, nC1, inc_x, inc_y, pos_out, nullptr, gradA, nullptr, nullptr);
}

int cpab_ssd_grad2d_calc_T(cpab_ctx* ctx, int tess, int nC, int nInst, const double* Trels, const double* As,
                           const float* src, const float* dst, int H, int W, double dt, int nTimeSteps, int nStepsODE,
                           int nC0, int nC1, double inc_x, double inc_y, double* pos_out, double* loss,
                           double* gradA, uint32_t* sig, double* n_redo) {
  if (!ctx) return CPAB_ERR_ARG;
  if (H < 1 || W < 1) return fail(ctx, CPAB_ERR_ARG, "images must be non-empty");
  if (!src || !dst || !loss || !Trels) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  return adjoint2d(ctx, 1, tess, nC, nInst, (int64_t)H * W, H, W, nullptr, nullptr, src, dst, Trels, As, dt,
                   nTimeSteps, nStepsODE, nC0, nC1, inc_x, inc_y, pos_out, loss, gradA, sig, n_redo);
}

// ------------------------------------------------------------- batch
}  // extern "C"

struct cpab_batch {
  cpab_ctx* ctx = nullptr;
  int dim = 2, tess = TESS_I, nInst = 0, maxNC = 0;
  bool pow2 = true;
  int64_t nCells = 0, nPts = 0, tableDoubles = 0;
  std::vector<BatchInst> h;
  BatchInst* d_inst = nullptr;
  unsigned int* queue = nullptr;
  BatchTile* d_tiles = nullptr;
  int nTiles = 0, tile = 0;
  // captured launch (reset + persistent kernel) and the arguments it was captured with
  hipStream_t cap = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* key[4] = {nullptr, nullptr, nullptr, nullptr};
  double key_h = 0.0;
  int key_i[5] = {0, 0, 0, 0, 0};
};

namespace {

void batch_drop_graph(cpab_batch* b) {
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  b->exec = nullptr;
  b->graph = nullptr;
}

int batch_tiles(cpab_batch* b, int tile) {
  if (b->d_tiles && b->tile == tile) return CPAB_OK;
  cpab_ctx* ctx = b->ctx;
  std::vector<BatchTile> t;
  for (int i = 0; i < b->nInst; i++) {
    const BatchInst& I = b->h[i];
    for (int64_t s = 0; s < I.nPts; s += tile) {
      BatchTile x{};
      x.inst = i;
      x.begin = I.pts + s;
      x.end = I.pts + (s + tile < I.nPts ? s + tile : I.nPts);
      t.push_back(x);
    }
  }
  if (t.size() > 0x7fffffff) return fail(ctx, CPAB_ERR_ARG, "too many tiles");
  HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
  if (b->d_tiles) HIP_OK(ctx, hipFree(b->d_tiles));
  b->d_tiles = nullptr;
  b->nTiles = (int)t.size();
  b->tile = tile;
  batch_drop_graph(b);
  if (t.empty()) return CPAB_OK;
  HIP_OK(ctx, hipMalloc(&b->d_tiles, sizeof(BatchTile) * t.size()));
  HIP_OK(ctx, hipMemcpy(b->d_tiles, t.data(), sizeof(BatchTile) * t.size(), hipMemcpyHostToDevice));
  return CPAB_OK;
}

template <int DIM, int TESS, bool POW2>
int batch_launch(cpab_batch* b, hipStream_t s, const cpab_batch_sched& q, const double* tables, const double* pos0,
                 double* out, double h, int nT, int nS, int32_t* nfb) {
  constexpr int NE = DIM * (DIM + 1);
  cpab_ctx* ctx = b->ctx;
  // both tables of the largest tessellation in LDS (measured on MI355X: the
  // Trels alone in LDS with the As read from global memory, which doubles the
  // occupancy of a 16 x 16 x 4-triangle batch, ran its best schedule 4 %
  // slower, substeps being a quarter of its trips)
  size_t lds = (size_t)2 * b->maxNC * NE * sizeof(double);
  // tiles of up to kBucketTileMax points are taken in an order by position
  // built in LDS after the tables (measured on MI355X, 1 024 2D instances of
  // uniform random points, tiles of 16 384: 54.4 -> 43.9 ms)
  int sort_off = 0;
  const size_t lds_sorted = lds + sizeof(unsigned int) * kBucketMax + sizeof(uint16_t) * (size_t)q.tile;
  if (q.tile <= kBucketTileMax && lds_sorted <= (size_t)ctx->max_lds) {
    sort_off = 2 * b->maxNC * NE;
    lds = lds_sorted;
  }
  auto k = k_integrate_batch<DIM, TESS, POW2>;
  allow_lds(k, lds);
  int per_cu = 0;
  HIP_OK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, q.block, lds));
  if (per_cu < 1) return fail(ctx, CPAB_ERR_UNSUPPORTED, "the batch kernel does not fit a CU with block %d", q.block);
  int grid = ctx->n_cu * per_cu;
  const int64_t chunks = ((int64_t)b->nTiles + q.chunk - 1) / q.chunk;
  if (grid > chunks) grid = (int)chunks;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, b->queue);
  hipLaunchKernelGGL(k, dim3(grid), dim3(q.block), lds, s, b->d_inst, b->d_tiles, b->nTiles, q.chunk, b->queue,
                     tables, pos0, out, h, nT, nS, nfb, sort_off);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

int batch_dispatch(cpab_batch* b, hipStream_t s, const cpab_batch_sched& q, const double* tables,
                   const double* pos0, double* out, double h, int nT, int nS, int32_t* nfb) {
#define BATCH_CASE(D_, T_, P_) \
  if (b->dim == D_ && b->tess == T_ && b->pow2 == P_) return batch_launch<D_, T_, P_>(b, s, q, tables, pos0, out, h, nT, nS, nfb);
  BATCH_CASE(1, TESS_II, true)
  BATCH_CASE(1, TESS_II, false)
  BATCH_CASE(2, TESS_I, true)
  BATCH_CASE(2, TESS_I, false)
  BATCH_CASE(2, TESS_II, true)
  BATCH_CASE(2, TESS_II, false)
  BATCH_CASE(3, TESS_I, true)
  BATCH_CASE(3, TESS_I, false)
  BATCH_CASE(3, TESS_II, true)
  BATCH_CASE(3, TESS_II, false)
#undef BATCH_CASE
  return fail(b->ctx, CPAB_ERR_UNSUPPORTED, "no batch kernel for dim %d tess %d", b->dim, b->tess);
}

}  // namespace

extern "C" {

int cpab_batch_create(cpab_ctx* ctx, int dim, int tess, int nInst, const cpab_batch_inst* insts, cpab_batch** out) {
  if (!ctx || !insts || !out) return CPAB_ERR_ARG;
  *out = nullptr;
  if (nInst < 1) return fail(ctx, CPAB_ERR_ARG, "nInst must be >= 1");
  const int NE = dim * (dim + 1);
  std::vector<BatchInst> h(nInst);
  int64_t cells = 0, pts = 0, tab = 0;
  int maxNC = 0;
  bool pow2_all = true, pow2_any = false;
  for (int i = 0; i < nInst; i++) {
    const cpab_batch_inst& q = insts[i];
    bool pow2;
    BatchInst& I = h[i];
    std::memset(&I, 0, sizeof I);
    int r = make_geom(ctx, dim, tess, q.nC0, q.nC1, q.nC2, q.inc_x, q.inc_y, q.inc_z, &I.g, &pow2);
    if (r) return r;
    int64_t boxes = q.nC0;
    if (dim >= 2) boxes *= q.nC1;
    if (dim >= 3) boxes *= q.nC2;
    const int64_t nC = boxes * (tess == TESS_II ? 1 : (dim == 2 ? 4 : 5));
    if ((r = check_nC(ctx, dim, tess, (int)nC, q.nC0, q.nC1, q.nC2))) return r;
    if (q.nPts < 0) return fail(ctx, CPAB_ERR_ARG, "instance %d: nPts < 0", i);
    pow2_all = pow2_all && pow2;
    pow2_any = pow2_any || pow2;
    I.nC = (int)nC;
    I.tab = tab;
    I.pts = pts;
    I.nPts = q.nPts;
    I.cell = cells;
    tab += 2 * nC * NE;
    pts += q.nPts;
    cells += nC;
    if (nC > maxNC) maxNC = (int)nC;
  }
  if (pow2_any && !pow2_all)
    return fail(ctx, CPAB_ERR_UNSUPPORTED,
                "a batch mixes power-of-two and other increments: make one batch of each");
  if ((size_t)2 * maxNC * NE * sizeof(double) > (size_t)ctx->max_lds)
    return fail(ctx, CPAB_ERR_UNSUPPORTED, "the largest tessellation's tables exceed LDS");
  cpab_batch* b = new (std::nothrow) cpab_batch();
  if (!b) return CPAB_ERR_NOMEM;
  b->ctx = ctx;
  b->dim = dim;
  b->tess = tess;
  b->nInst = nInst;
  b->maxNC = maxNC;
  b->pow2 = pow2_all;
  b->nCells = cells;
  b->nPts = pts;
  b->tableDoubles = tab;
  b->h = h;
  hipError_t e = hipMalloc(&b->d_inst, sizeof(BatchInst) * nInst);
  if (e == hipSuccess) e = hipMalloc(&b->queue, 256);
  if (e == hipSuccess) e = hipMemcpy(b->d_inst, h.data(), sizeof(BatchInst) * nInst, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    if (b->d_inst) (void)hipFree(b->d_inst);
    if (b->queue) (void)hipFree(b->queue);
    delete b;
    return fail(ctx, CPAB_ERR_NOMEM, "batch: %s", hipGetErrorString(e));
  }
  *out = b;
  return CPAB_OK;
}

int cpab_batch_sizes(cpab_batch* b, int64_t* nCells, int64_t* nPts, int64_t* tableDoubles, int* maxNC) {
  if (!b) return CPAB_ERR_ARG;
  if (nCells) *nCells = b->nCells;
  if (nPts) *nPts = b->nPts;
  if (tableDoubles) *tableDoubles = b->tableDoubles;
  if (maxNC) *maxNC = b->maxNC;
  return CPAB_OK;
}

int cpab_batch_prep(cpab_batch* b, const double* As, int sign, double dt, double* tables) {
  if (!b) return CPAB_ERR_ARG;
  cpab_ctx* ctx = b->ctx;
  if (sign != 1 && sign != -1) return fail(ctx, CPAB_ERR_ARG, "sign must be 1 or -1");
  if (!As || !tables) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const int64_t nb = (b->nCells + 255) / 256;
  if (b->dim == 1)
    hipLaunchKernelGGL(k_batch_prep<1>, dim3((unsigned)nb), dim3(256), 0, ctx->stream, b->d_inst, b->nInst, As,
                       b->nCells, sign, dt, tables, 12);
  else if (b->dim == 2)
    hipLaunchKernelGGL(k_batch_prep<2>, dim3((unsigned)nb), dim3(256), 0, ctx->stream, b->d_inst, b->nInst, As,
                       b->nCells, sign, dt, tables, 12);
  else
    hipLaunchKernelGGL(k_batch_prep<3>, dim3((unsigned)nb), dim3(256), 0, ctx->stream, b->d_inst, b->nInst, As,
                       b->nCells, sign, dt, tables, 12);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

int cpab_batch_calc_T(cpab_batch* b, const double* tables, const double* pos0, double* out, double dt, int nT,
                      int nS, const cpab_batch_sched* sched, int use_graph, int32_t* nfb) {
  if (!b) return CPAB_ERR_ARG;
  cpab_ctx* ctx = b->ctx;
  cpab_batch_sched q = sched ? *sched : cpab_batch_sched{16384, 1024, 4};
  if (q.tile < 64 || q.block < 64 || q.block > 1024 || (q.block & 63) || q.chunk < 1)
    return fail(ctx, CPAB_ERR_ARG, "schedule: tile >= 64, block a multiple of 64 up to 1024, chunk >= 1");
  if (nT < 0 || nS < 0) return fail(ctx, CPAB_ERR_ARG, "negative step counts");
  if (b->nPts == 0) return CPAB_OK;
  if (!tables || !pos0 || !out) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const size_t pb = (size_t)b->nPts * b->dim * sizeof(double);
  if (overlap(pos0, pb, out, pb)) return fail(ctx, CPAB_ERR_ALIAS, "pts and out cannot point to the same memory");
  int r = batch_tiles(b, q.tile);
  if (r) return r;
  const double h = dt / (double)(nS > 0 ? nS : 1);
  hipEvent_t a = nullptr, e = nullptr;
  if ((r = prof_begin(ctx, &a, &e))) return r;
  if (!use_graph) {
    if ((r = batch_dispatch(b, ctx->stream, q, tables, pos0, out, h, nT, nS, nfb))) return r;
  } else {
    const void* key[4] = {tables, pos0, out, nfb};
    const int key_i[5] = {nT, nS, q.tile, q.block, q.chunk};
    bool same = b->exec && b->key_h == h;
    for (int i = 0; i < 4 && same; i++) same = b->key[i] == key[i];
    for (int i = 0; i < 5 && same; i++) same = b->key_i[i] == key_i[i];
    if (!same) {
      batch_drop_graph(b);
      if (!b->cap) HIP_OK(ctx, hipStreamCreate(&b->cap));
      HIP_OK(ctx, hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal));
      r = batch_dispatch(b, b->cap, q, tables, pos0, out, h, nT, nS, nfb);
      hipGraph_t gph = nullptr;
      hipError_t ee = hipStreamEndCapture(b->cap, &gph);
      if (r) return r;
      if (ee != hipSuccess) return fail(ctx, CPAB_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(ee));
      b->graph = gph;
      HIP_OK(ctx, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
      for (int i = 0; i < 4; i++) b->key[i] = key[i];
      for (int i = 0; i < 5; i++) b->key_i[i] = key_i[i];
      b->key_h = h;
    }
    HIP_OK(ctx, hipGraphLaunch(b->exec, ctx->stream));
  }
  return prof_end(ctx, a, e);
}

int cpab_batch_destroy(cpab_batch* b) {
  if (!b) return CPAB_OK;
  (void)hipStreamSynchronize(b->ctx->stream);
  batch_drop_graph(b);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  if (b->d_tiles) (void)hipFree(b->d_tiles);
  if (b->d_inst) (void)hipFree(b->d_inst);
  if (b->queue) (void)hipFree(b->queue);
  delete b;
  return CPAB_OK;
}

// ------------------------------------------------------------- work items
}  // extern "C"

namespace {

template <int DIM, int TESS, bool POW2>
int items_launch(cpab_ctx* ctx, const WorkItem* items, unsigned int* queue, int nItems, int maxNC, int64_t maxLen,
                 const cpab_batch_sched& q, double h, int nT, int nS) {
  constexpr int NE = DIM * (DIM + 1);
  size_t lds = (size_t)2 * maxNC * NE * sizeof(double);
  // 1D items of up to kBucketTileMax points are taken in an order by position
  // built in LDS after the tables (tile_bucket_order).  Measured on MI355X in
  // the served mix: 1D requests of random points p50 0.49 -> 0.44 ms; the 2D
  // and 3D requests there are grids whose items are whole rows, already in
  // an order by position, and the sort's passes made them 1-3 % slower.
  int sort_off = 0;
  const size_t lds_sorted = lds + sizeof(unsigned int) * kBucketMax + sizeof(uint16_t) * (size_t)maxLen;
  if (DIM == 1 && maxLen <= kBucketTileMax && lds_sorted <= (size_t)ctx->max_lds) {
    sort_off = 2 * maxNC * NE;
    lds = lds_sorted;
  }
  auto k = k_integrate_items<DIM, TESS, POW2>;
  allow_lds(k, lds);
  int per_cu = 0;
  HIP_OK(ctx, hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k, q.block, lds));
  if (per_cu < 1) return fail(ctx, CPAB_ERR_UNSUPPORTED, "the item kernel does not fit a CU with block %d", q.block);
  int grid = ctx->n_cu * per_cu;
  const int64_t chunks = ((int64_t)nItems + q.chunk - 1) / q.chunk;
  if (grid > chunks) grid = (int)chunks;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, ctx->stream, queue);
  hipLaunchKernelGGL(k, dim3(grid), dim3(q.block), lds, ctx->stream, items, nItems, q.chunk, queue, h, nT, nS,
                     sort_off);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

}  // namespace

extern "C" {

int cpab_theta2avees(cpab_ctx* ctx, int D, int d, int m, const double* B, const double* theta, double* Av) {
  if (!ctx) return CPAB_ERR_ARG;
  if (D < 1 || d < 1 || m < 0) return fail(ctx, CPAB_ERR_ARG, "bad sizes");
  if (m == 0) return CPAB_OK;
  if (!B || !theta || !Av) return fail(ctx, CPAB_ERR_ARG, "null pointer argument");
  const int64_t n = (int64_t)m * D;
  hipLaunchKernelGGL(k_theta2avees, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx->stream, B, theta, Av, D, d,
                     m);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

}  // extern "C"

namespace {

// Checks of a work list and the device form of item i (no allocation).
int item_convert(cpab_ctx* ctx, int dim, int tess, const cpab_work_item& q2, int i, WorkItem* W, bool* pow2) {
  std::memset(W, 0, sizeof *W);
  int r = make_geom(ctx, dim, tess, q2.nC0, q2.nC1, q2.nC2, q2.inc_x, q2.inc_y, q2.inc_z, &W->g, pow2);
  if (r) return r;
  int64_t boxes = q2.nC0;
  if (dim >= 2) boxes *= q2.nC1;
  if (dim >= 3) boxes *= q2.nC2;
  const int64_t nC = boxes * (tess == TESS_II ? 1 : (dim == 2 ? 4 : 5));
  if ((r = check_nC(ctx, dim, tess, (int)nC, q2.nC0, q2.nC1, q2.nC2))) return r;
  if (!q2.trels || !q2.As || !q2.pos0 || !q2.out || q2.begin < 0 || q2.end < q2.begin)
    return fail(ctx, CPAB_ERR_ARG, "item %d: null pointer or bad range", i);
  W->nC = (int)nC;
  W->trels = q2.trels;
  W->as = q2.As;
  W->pos0 = q2.pos0;
  W->out = q2.out;
  W->begin = q2.begin;
  W->end = q2.end;
  return CPAB_OK;
}

// Argument checks shared by the item entries; the schedule in *q, the
// largest tessellation and the increment kind of the list.
int items_check(cpab_ctx* ctx, int dim, int tess, int nItems, const cpab_work_item* items, int nTimeSteps,
                int nStepsODE, const cpab_batch_sched* sched, cpab_batch_sched* q, int* maxNC, bool* pow2_all,
                int64_t* maxLen) {
  if (nItems < 0 || (nItems > 0 && !items)) return fail(ctx, CPAB_ERR_ARG, "bad item list");
  if (nTimeSteps < 0 || nStepsODE < 0) return fail(ctx, CPAB_ERR_ARG, "negative step counts");
  *q = sched ? *sched : cpab_batch_sched{16384, 1024, 1};
  if (q->block < 64 || q->block > 1024 || (q->block & 63) || q->chunk < 1)
    return fail(ctx, CPAB_ERR_ARG, "schedule: block a multiple of 64 up to 1024, chunk >= 1");
  bool all = true, any = false;
  int mx = 0;
  int64_t len = 0;
  for (int i = 0; i < nItems; i++) {
    WorkItem W;
    bool pow2;
    int r = item_convert(ctx, dim, tess, items[i], i, &W, &pow2);
    if (r) return r;
    all = all && pow2;
    any = any || pow2;
    if (W.nC > mx) mx = W.nC;
    if (W.end - W.begin > len) len = W.end - W.begin;
  }
  if (any && !all)
    return fail(ctx, CPAB_ERR_UNSUPPORTED, "items mix power-of-two and other increments: launch each kind apart");
  const int NE = dim * (dim + 1);
  if ((size_t)2 * mx * NE * sizeof(double) > (size_t)ctx->max_lds)
    return fail(ctx, CPAB_ERR_UNSUPPORTED, "the largest tessellation's tables exceed LDS");
  *maxNC = mx;
  *pow2_all = all;
  *maxLen = len;
  return CPAB_OK;
}

int items_dispatch(cpab_ctx* ctx, const WorkItem* d_items, unsigned int* queue, int dim, int tess, bool pow2_all,
                   int nItems, int maxNC, int64_t maxLen, const cpab_batch_sched& q, double dt, int nTimeSteps,
                   int nStepsODE) {
  const double h = dtdiv(dt, nStepsODE);
  int r = CPAB_OK;
#define ITEMS_CASE(D_, T_, P_)                                                                                    \
  if (dim == D_ && tess == T_ && pow2_all == P_)                                                                 \
    r = items_launch<D_, T_, P_>(ctx, d_items, queue, nItems, maxNC, maxLen, q, h, nTimeSteps, nStepsODE);       \
  else
  ITEMS_CASE(1, TESS_II, true)
  ITEMS_CASE(1, TESS_II, false)
  ITEMS_CASE(2, TESS_I, true)
  ITEMS_CASE(2, TESS_I, false)
  ITEMS_CASE(2, TESS_II, true)
  ITEMS_CASE(2, TESS_II, false)
  ITEMS_CASE(3, TESS_I, true)
  ITEMS_CASE(3, TESS_I, false)
  ITEMS_CASE(3, TESS_II, true)
  ITEMS_CASE(3, TESS_II, false)
  r = fail(ctx, CPAB_ERR_UNSUPPORTED, "no item kernel for dim %d tess %d", dim, tess);
#undef ITEMS_CASE
  return r;
}

}  // namespace

extern "C" {

int cpab_calc_T_items(cpab_ctx* ctx, int dim, int tess, int nItems, const cpab_work_item* items, double dt,
                      int nTimeSteps, int nStepsODE, const cpab_batch_sched* sched) {
  if (!ctx) return CPAB_ERR_ARG;
  cpab_batch_sched q;
  int maxNC = 0;
  bool pow2_all = true;
  int64_t maxLen = 0;
  int r = items_check(ctx, dim, tess, nItems, items, nTimeSteps, nStepsODE, sched, &q, &maxNC, &pow2_all, &maxLen);
  if (r) return r;
  if (nItems == 0) return CPAB_OK;
  std::vector<WorkItem>& hv = ctx->h_items;
  hv.resize(nItems);
  for (int i = 0; i < nItems; i++) {
    bool pow2;
    if ((r = item_convert(ctx, dim, tess, items[i], i, &hv[i], &pow2))) return r;
  }
  if (!ctx->queue) HIP_OK(ctx, hipMalloc(&ctx->queue, 256));
  if (ctx->items_cap < (size_t)nItems) {
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->items) HIP_OK(ctx, hipFree(ctx->items));
    ctx->items = nullptr;
    size_t cap = 1024;
    while (cap < (size_t)nItems) cap *= 2;
    HIP_OK(ctx, hipMalloc(&ctx->items, cap * sizeof(WorkItem)));
    ctx->items_cap = cap;
  }
  // pageable source: the copy is staged before the call returns, so the host
  // vector can be refilled by the next call; stream order protects the device
  // buffer from a launch still reading it
  HIP_OK(ctx, hipMemcpyAsync(ctx->items, hv.data(), sizeof(WorkItem) * nItems, hipMemcpyHostToDevice, ctx->stream));
  hipEvent_t a = nullptr, e = nullptr;
  if ((r = prof_begin(ctx, &a, &e))) return r;
  r = items_dispatch(ctx, ctx->items, ctx->queue, dim, tess, pow2_all, nItems, maxNC, maxLen, q, dt, nTimeSteps,
                     nStepsODE);
  if (r) return r;
  return prof_end(ctx, a, e);
}

int cpab_calc_T_items_noalloc(cpab_ctx* ctx, int dim, int tess, int nItems, const cpab_work_item* items, double dt,
                              int nTimeSteps, int nStepsODE, const cpab_batch_sched* sched) {
  if (!ctx) return CPAB_ERR_ARG;
  if (nItems > kStaticItems)
    return fail(ctx, CPAB_ERR_ARG, "%d work items: the allocation-free entry takes at most %d per call", nItems,
                kStaticItems);
  if (!ctx->static_scratch)
    return fail(ctx, CPAB_ERR_UNSUPPORTED, "no static scratch on device %d (devices 0..%d have one)", ctx->device,
                kMaxDevices - 1);
  cpab_batch_sched q;
  int maxNC = 0;
  bool pow2_all = true;
  int64_t maxLen = 0;
  int r = items_check(ctx, dim, tess, nItems, items, nTimeSteps, nStepsODE, sched, &q, &maxNC, &pow2_all, &maxLen);
  if (r) return r;
  if (nItems == 0) return CPAB_OK;
  unsigned int* queue = reinterpret_cast<unsigned int*>(ctx->static_scratch);
  WorkItem* d_items = reinterpret_cast<WorkItem*>(ctx->static_scratch + kStaticQueueBytes);
  // the device form of the list, converted in stack-sized chunks; each
  // pageable copy is staged before it returns, so the chunk can be refilled
  constexpr int kChunk = 64;
  WorkItem buf[kChunk];
  for (int i0 = 0; i0 < nItems; i0 += kChunk) {
    const int n = nItems - i0 < kChunk ? nItems - i0 : kChunk;
    for (int i = 0; i < n; i++) {
      bool pow2;
      if ((r = item_convert(ctx, dim, tess, items[i0 + i], i0 + i, &buf[i], &pow2))) return r;
    }
    HIP_OK(ctx, hipMemcpyAsync(d_items + i0, buf, sizeof(WorkItem) * n, hipMemcpyHostToDevice, ctx->stream));
  }
  return items_dispatch(ctx, d_items, queue, dim, tess, pow2_all, nItems, maxNC, maxLen, q, dt, nTimeSteps,
                        nStepsODE);
}

// ------------------------------------------------------------- MCMC
}  // extern "C"

// Launches per captured graph: graph launches cost several microseconds each.
constexpr int kGraphSteps = 16;
// Speculation depth chosen when the descriptor leaves it to the library:
// enough slots that one launch integrates about 2^20 points.
constexpr int64_t kSpecPoints = 1 << 20;
constexpr int kMaxSpec = 64;  // slots per chain and launch (a 64-bit mask tracks them)
constexpr int kAbortCheckpoints = 3;  // steps at which a one-pixel-per-lane wave polls for early abort

// Pixel order of a chain's integration (work index -> pixel): the domain's
// top row and left column first, then the interior, in 8 x 8 tiles when W
// and H are multiples of 8 (tile by tile, rows within a tile, the tiles'
// pixels of that row and column left out), else row by row.  The row and
// column lie on the domain boundary, where the zero boundary velocity keeps
// their points within rounding of the wall, so their lookups take the
// general path nearly every trip; in waves of their own they no longer send
// the interior lanes of their tiles down it (one chain: +2 % measured on
// MI355X).  The first pixel is the domain corner (0, 0).
static std::vector<uint32_t> reg2d_pixel_order(int H, int W, int ptsPerWave) {
  std::vector<uint32_t> o;
  o.reserve((size_t)H * W);
  // the top row, then the left column: the pixel (0, 0) first
  for (int x = 0; x < W; x++) o.push_back((uint32_t)x);
  for (int y = 1; y < H; y++) o.push_back((uint32_t)y * W);
  // The interior in 8 x 8 tiles from (1, 1); what is left of it at the right
  // and the bottom (under 8 wide) in strips: the right strip row by row, the
  // bottom strip column by column (a wave's pixels then lie in a ~7 x 9 patch).
  const int nTx = (W - 1) / 8, nTy = (H - 1) / 8, xs = 1 + 8 * nTx, ys = 1 + 8 * nTy;
  for (int y = 1; y < ys; y++)
    for (int x = xs; x < W; x++) o.push_back((uint32_t)y * W + x);
  for (int x = 1; x < W; x++)
    for (int y = ys; y < H; y++) o.push_back((uint32_t)y * W + x);
  // The full tiles last, as many as possible starting on a wave's first
  // pixel (the ones that do not fit go first), so each wave's pixels are
  // whole tiles; that needs H W to be a multiple of 64.
  const int64_t nFull = (int64_t)nTx * nTy, nPts = (int64_t)H * W;
  int64_t nHead = 0;
  if (nPts % 64 == 0 && nFull > 0)
    while (nHead < nFull && (nPts - 64 * (nFull - nHead)) % ptsPerWave != 0) nHead++;
  for (int64_t i = 0; i < nFull; i++) {
    const int64_t tile = i < nHead ? nFull - nHead + i : i - nHead;  // the head's tiles are the last ones
    const int tx = (int)(tile % nTx), ty = (int)(tile / nTx);
    for (int v = 0; v < 8; v++)
      for (int u = 0; u < 8; u++) o.push_back((uint32_t)(1 + ty * 8 + v) * W + (uint32_t)(1 + tx * 8 + u));
  }
  return o;
}

struct cpab_reg2d {
  cpab_ctx* ctx = nullptr;
  McmcDesc D{};
  int K_cap = 1;  // slots the scratch holds (cpab_reg2d_set_spec_depth may go up to it)
  int block = MCMC_BLOCK;  // threads per block of the step kernel
  bool pow2 = true;
  double* scratch = nullptr;  // theta_prop | ll_prop | tables | partials | slot_res | lz | Bt | done_slot | stop | done
  hipStream_t cap = nullptr;  // capture stream, created once
  // one captured kGraphSteps-launch graph per speculation depth: the launches
  // carry the descriptor by value and only D.K differs between depths
  std::map<int, std::pair<hipGraph_t, hipGraphExec_t>> graphs;
  int64_t n_replayed = 0, n_direct = 0;  // step launches by graph replay / directly
  std::vector<int64_t> host_step;
};

namespace {

size_t reg2d_lds(const McmcDesc& D) {
  const int a = 2 * D.nC * 6, b = mcmc_finish_doubles(D.K);
  return (size_t)(a > b ? a : b) * sizeof(double);
}

// Tables of the first nk armed proposals of every chain, the L z vectors of
// the chain's current step counter and the next dispatch order.
void reg2d_launch_tables(cpab_reg2d* r, hipStream_t s, int nk) {
  const McmcDesc& D = r->D;
  const size_t sh = (size_t)mcmc_tables_doubles(D.nC, D.d, D.nBlocks) * sizeof(double);
  const int rows = 6 * D.nC;
  const unsigned threads = rows >= 1024 ? 1024u : (unsigned)((rows + 63) / 64 * 64);
  hipLaunchKernelGGL(k_mcmc_tables, dim3((unsigned)(nk + D.K + (D.order ? 1 : 0)), D.nChains), dim3(threads), sh, s,
                     D, nk);
}

// One step launch and the tables of the proposals it arms.
int reg2d_launch_step(cpab_reg2d* r, hipStream_t s, int init) {
  McmcDesc& D = r->D;
  const size_t sh = reg2d_lds(D);
  // several chains one step per launch: the chain on grid axis x (dispatched
  // fastest), so every chain's first pixel block (the domain corner's, the
  // longest) starts first (MI355X, 256 chains of 256x256: +14 %; with
  // speculation, 16 chains x 4 slots measured 7 % slower, so not there)
  const bool swap = D.nChains > 1 && D.K == 1 && D.nBlocks <= 65535;
  const dim3 grid = swap ? dim3((unsigned)D.nChains, (unsigned)(D.nBlocks * D.K))
                         : dim3((unsigned)(D.nBlocks * D.K), (unsigned)D.nChains);
  // events bracket the two kernels when profiling (not inside a capture)
  const bool timed = r->ctx->prof && s == r->ctx->stream;
  hipEvent_t a = nullptr, b = nullptr;
  if (timed) {
    int rc = prof_begin(r->ctx, &a, &b);
    if (rc) return rc;
  }
  // the fast lookup pays once the launch holds four waves per SIMD or more
  // (a lone one-step chain is latency-bound: one wave per SIMD)
  const int64_t waves = (int64_t)grid.x * grid.y * (r->block / 64);
  const bool fast = waves >= (int64_t)r->ctx->n_cu * 4 * 4;
#define STEP_LAUNCH_S(T_, P_, S_)                                                                               \
  do {                                                                                                          \
    if (r->block == 64) hipLaunchKernelGGL((k_mcmc_step<T_, P_, 64, true, S_>), grid, dim3(64), sh, s, D, init); \
    else if (fast) hipLaunchKernelGGL((k_mcmc_step<T_, P_, 256, true, S_>), grid, dim3(256), sh, s, D, init);    \
    else hipLaunchKernelGGL((k_mcmc_step<T_, P_, 256, false, S_>), grid, dim3(256), sh, s, D, init);             \
  } while (0)
#define STEP_LAUNCH(T_, P_)              \
  do {                                   \
    if (swap) STEP_LAUNCH_S(T_, P_, true); \
    else STEP_LAUNCH_S(T_, P_, false);     \
  } while (0)
  if (D.tess == TESS_I) {
    if (r->pow2) STEP_LAUNCH(TESS_I, true);
    else STEP_LAUNCH(TESS_I, false);
  } else {
    if (r->pow2) STEP_LAUNCH(TESS_II, true);
    else STEP_LAUNCH(TESS_II, false);
  }
#undef STEP_LAUNCH
#undef STEP_LAUNCH_S
  reg2d_launch_tables(r, s, D.K);
  if (timed) {
    int rc = prof_end(r->ctx, a, b);
    if (rc) return rc;
  }
  HIP_OK(r->ctx, hipGetLastError());
  return CPAB_OK;
}

// n launches of the step kernel: replays of the captured kGraphSteps-launch
// graph, then the remainder directly.
int reg2d_launches(cpab_reg2d* r, int64_t n, int use_graph) {
  cpab_ctx* ctx = r->ctx;
  if (!use_graph) {
    for (int64_t i = 0; i < n; i++) {
      int rc = reg2d_launch_step(r, ctx->stream, 0);
      if (rc) return rc;
      r->n_direct++;
    }
    return CPAB_OK;
  }
  const int64_t full = n / kGraphSteps;
  hipGraphExec_t exec = nullptr;
  auto it = r->graphs.find(r->D.K);
  if (it != r->graphs.end()) exec = it->second.second;
  if (!exec && full > 0) {
    // capture kGraphSteps launches on a private stream; replays go to the context's stream
    if (!r->cap) HIP_OK(ctx, hipStreamCreate(&r->cap));
    HIP_OK(ctx, hipStreamBeginCapture(r->cap, hipStreamCaptureModeThreadLocal));
    int rc = CPAB_OK;
    for (int i = 0; i < kGraphSteps && rc == CPAB_OK; i++) rc = reg2d_launch_step(r, r->cap, 0);
    hipGraph_t gph = nullptr;
    hipError_t e = hipStreamEndCapture(r->cap, &gph);
    if (rc) {
      if (gph) (void)hipGraphDestroy(gph);
      return rc;
    }
    if (e != hipSuccess) return fail(ctx, CPAB_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(&exec, gph, nullptr, nullptr, 0);
    if (e != hipSuccess) {
      (void)hipGraphDestroy(gph);
      return fail(ctx, CPAB_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    r->graphs[r->D.K] = std::make_pair(gph, exec);
  }
  for (int64_t i = 0; i < full; i++) {
    hipEvent_t a = nullptr, b = nullptr;
    int rc = prof_begin(ctx, &a, &b);
    if (rc) return rc;
    HIP_OK(ctx, hipGraphLaunch(exec, ctx->stream));
    if ((rc = prof_end(ctx, a, b, kGraphSteps))) return rc;
    r->n_replayed += kGraphSteps;
  }
  for (int64_t i = full * kGraphSteps; i < n; i++) {
    int rc = reg2d_launch_step(r, ctx->stream, 0);
    if (rc) return rc;
    r->n_direct++;
  }
  return CPAB_OK;
}

}  // namespace

extern "C" {

int cpab_reg2d_create(cpab_ctx* ctx, const cpab_reg2d_desc* in, cpab_reg2d** out) {
  if (!ctx || !in || !out) return CPAB_ERR_ARG;
  *out = nullptr;
  const cpab_reg2d_desc& q = *in;
  Geom g;
  bool pow2;
  int rc = make_geom(ctx, 2, q.tess, q.nC0, q.nC1, 0, q.inc_x, q.inc_y, 1.0, &g, &pow2);
  if (rc) return rc;
  if ((rc = check_nC(ctx, 2, q.tess, q.nC, q.nC0, q.nC1, 0))) return rc;
  if (q.nChains < 1 || q.nChains > 65535) return fail(ctx, CPAB_ERR_ARG, "nChains must be in [1, 65535]");
  if (q.d < 1 || q.d > 8192) return fail(ctx, CPAB_ERR_ARG, "d must be in [1, 8192]");
  if (q.H < 1 || q.W < 1) return fail(ctx, CPAB_ERR_ARG, "images must be non-empty");
  if (!(q.sigma > 0.0)) return fail(ctx, CPAB_ERR_ARG, "sigma must be positive");
  if (q.nTimeSteps < 0 || q.nStepsODE < 0) return fail(ctx, CPAB_ERR_ARG, "negative step counts");
  if (!q.B || !q.Lprop || !q.src || !q.dst || !q.theta || !q.ll || !q.n_accepted || !q.step)
    return fail(ctx, CPAB_ERR_ARG, "null pointer in the descriptor");
  if (q.spec_depth < 0 || q.spec_depth > kMaxSpec)
    return fail(ctx, CPAB_ERR_ARG, "spec_depth must be in [0, %d]", kMaxSpec);
  const int64_t nPts = (int64_t)q.H * q.W;
  int K = q.spec_depth;
  if (K == 0) {
    const int64_t k = kSpecPoints / (nPts * (int64_t)q.nChains);
    K = k < 1 ? 1 : (k > 16 ? 16 : (int)k);
  }
  // the re-arming workspace (normals and proposals of every slot) must fit in LDS
  while (K > 1 && (size_t)mcmc_arm_doubles(K, q.d) * sizeof(double) > (size_t)ctx->max_lds) K--;
  // room for four times the library's own choice, for later tuning of the
  // depth (with early abort, deep speculation costs little when it fails)
  int K_cap = K;
  {
    const int64_t k4 = 4 * kSpecPoints / (nPts * (int64_t)q.nChains);
    const int auto4 = k4 < 1 ? 1 : (k4 > kMaxSpec ? kMaxSpec : (int)k4);
    if (auto4 > K_cap) K_cap = auto4;
    while (K_cap > K && (size_t)mcmc_arm_doubles(K_cap, q.d) * sizeof(double) > (size_t)ctx->max_lds) K_cap--;
  }
  if ((size_t)(2 * q.nC * 6) * sizeof(double) > (size_t)ctx->max_lds ||
      (size_t)mcmc_tables_doubles(q.nC, q.d, 1024) * sizeof(double) > (size_t)ctx->max_lds ||
      (size_t)mcmc_arm_doubles(K, q.d) * sizeof(double) > (size_t)ctx->max_lds)
    return fail(ctx, CPAB_ERR_UNSUPPORTED, "per-cell tables or proposal workspace exceed LDS");
  cpab_reg2d* r = new (std::nothrow) cpab_reg2d();
  if (!r) return CPAB_ERR_NOMEM;
  r->ctx = ctx;
  r->pow2 = pow2;
  McmcDesc& D = r->D;
  D.nChains = q.nChains;
  D.d = q.d;
  D.nC = q.nC;
  D.nC0 = q.nC0;
  D.nC1 = q.nC1;
  D.tess = q.tess;
  D.inc_x = q.inc_x;
  D.inc_y = q.inc_y;
  D.H = q.H;
  D.W = q.W;
  D.sigma = q.sigma;
  D.dt = q.dt;
  D.h = q.dt / (double)q.nStepsODE;
  D.nT = q.nTimeSteps;
  D.nS = q.nStepsODE;
  D.use_ave_ll = q.use_ave_ll;
  D.inverted = q.inverted_acceptance;
  D.trace_len = q.trace ? q.trace_len : 0;
  D.seed_lo = (uint32_t)q.seed;
  D.seed_hi = (uint32_t)(q.seed >> 32);
  D.L = q.Lprop;
  D.src = q.src;
  D.dst = q.dst;
  D.theta = q.theta;
  D.ll = q.ll;
  D.n_acc = q.n_accepted;
  D.step = q.step;
  D.trace = q.trace;
  D.nfb = reinterpret_cast<unsigned long long*>(q.n_fallback);
  D.stats = q.stats;
  D.K = K;
  D.g = g;

  D.poll = kAbortCheckpoints;
  // One pixel per lane while a launch holds few waves per SIMD (latency
  // hiding needs the waves); with many (batched chains) 2 or 4 pixels per
  // lane with refill keep lanes busy past their slowest pixel (measured on
  // MI355X: 1 chain x 20 slots best at 64, 256 chains at 256, +13 %).
  if (q.pts_per_wave != 0) {
    if (q.pts_per_wave < 64 || q.pts_per_wave > 4096 || (q.pts_per_wave % 64) != 0) {
      delete r;
      return fail(ctx, CPAB_ERR_ARG, "pts_per_wave must be a multiple of 64 in [64, 4096] or 0");
    }
    D.ptsPerWave = q.pts_per_wave;
  } else {
    const int64_t waves64 = (int64_t)q.nChains * K * ((nPts + 63) / 64);
    const int64_t full = (int64_t)ctx->n_cu * 4 * 8;
    D.ptsPerWave = waves64 >= 32 * full ? 256 : (waves64 >= 8 * full ? 128 : 64);
    // many short-armed chains: one-wave blocks release CUs sooner
    if (K == 1 && waves64 >= 8 * full) r->block = 64;
  }
  const int64_t waves = (nPts + D.ptsPerWave - 1) / D.ptsPerWave;
  D.nBlocks = (int)((waves + r->block / 64 - 1) / (r->block / 64));
  if ((int64_t)D.nBlocks * K > 0x7fffffff) {
    delete r;
    return fail(ctx, CPAB_ERR_ARG, "image too large for the launch grid");
  }
  r->K_cap = K_cap;
  // scratch, zeroed once: theta_prop | ll_prop | tables | partial | partial_nfb | partial_px | slot_res |
  // lz | Bt | lp | lp_prop | done_slot | stop | done | order | ready, abort_after (in doubles; done_slot
  // starts on a 128-byte boundary)
  const size_t nck = (size_t)q.nChains * K_cap;
  const size_t nB = (size_t)q.nC * 6 * q.d;
  const size_t o_ds = (nck * q.d + q.nChains + nck * 2 * q.nC * 6 + 3 * nck * D.nBlocks + 3 * nck + 2 * nck * q.d +
                       nB + q.nChains + nck + 15) / 16 * 16;
  const size_t n_ds = nck * DONE_STRIDE / 2;
  const size_t n_ord = ((size_t)q.nChains * D.nBlocks + 1) / 2;  // ints, in doubles
  const size_t o_ctl = o_ds + n_ds + 2 * q.nChains + n_ord;
  const size_t o_pm = o_ctl + q.nChains + (q.nChains + 1) / 2;  // after ready (64-bit) and abort_after (int) per chain
  const size_t n = o_pm + (nPts + 1) / 2;                       // the pixel order (uint32)
  hipError_t e = hipMalloc(&r->scratch, n * sizeof(double));
  // zeroed on the context's stream, so the kernels after it see zeros
  if (e == hipSuccess) e = hipMemsetAsync(r->scratch, 0, n * sizeof(double), ctx->stream);
  if (e != hipSuccess) {
    if (r->scratch) (void)hipFree(r->scratch);
    delete r;
    return fail(ctx, CPAB_ERR_NOMEM, "hipMalloc: %s", hipGetErrorString(e));
  }
  D.theta_prop = r->scratch;
  D.ll_prop = D.theta_prop + nck * q.d;
  D.tables = D.ll_prop + q.nChains;
  D.partial = D.tables + nck * 2 * q.nC * 6;
  D.partial_nfb = D.partial + nck * D.nBlocks;
  D.partial_px = D.partial_nfb + nck * D.nBlocks;
  D.slot_res = D.partial_px + nck * D.nBlocks;
  D.lz = D.slot_res + 3 * nck;
  double* Bt = D.lz + 2 * nck * q.d;
  D.Bt = Bt;
  D.lp = Bt + nB;
  D.lp_prop = D.lp + q.nChains;
  D.wlp = q.wlp;
  D.J = q.prior_prec;
  D.integrate_corner = q.integrate_corner != 0;
  D.done_slot = reinterpret_cast<unsigned int*>(r->scratch + o_ds);
  D.stop = reinterpret_cast<int64_t*>(r->scratch + o_ds + n_ds);
  D.done = reinterpret_cast<unsigned int*>(D.stop + q.nChains);
  // blocks in dispatch order by their work when lanes take several pixels
  D.order = D.ptsPerWave > 64 ? reinterpret_cast<int*>(r->scratch + o_ds + n_ds + 2 * q.nChains) : nullptr;
  D.ready = reinterpret_cast<unsigned long long*>(r->scratch + o_ctl);
  D.abort_after = reinterpret_cast<int*>(D.ready + q.nChains);
  uint32_t* pixmap = reinterpret_cast<uint32_t*>(r->scratch + o_pm);
  D.pixmap = pixmap;
  r->host_step.resize(q.nChains);
  {  // no slot is aborted before the first launch; the pixel order
    const std::vector<int> none(q.nChains, kNoAbort);
    const std::vector<uint32_t> order = reg2d_pixel_order(q.H, q.W, D.ptsPerWave);
    e = hipMemcpyAsync(D.abort_after, none.data(), sizeof(int) * q.nChains, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(pixmap, order.data(), sizeof(uint32_t) * order.size(), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) {
      (void)hipFree(r->scratch);
      delete r;
      return fail(ctx, CPAB_ERR_HIP, "hipMemcpyAsync: %s", hipGetErrorString(e));
    }
  }
  // B is read transposed by the tables kernel (a copy: later changes of B are not seen)
  hipLaunchKernelGGL(k_transpose, dim3((unsigned)((nB + 255) / 256)), dim3(256), 0, ctx->stream, q.B, Bt,
                     q.nC * 6, q.d);
  e = hipGetLastError();
  if (e != hipSuccess) {
    (void)hipFree(r->scratch);
    delete r;
    return fail(ctx, CPAB_ERR_HIP, "k_transpose: %s", hipGetErrorString(e));
  }
  *out = r;
  return CPAB_OK;
}

int cpab_reg2d_init(cpab_reg2d* r) {
  if (!r) return CPAB_ERR_ARG;
  McmcDesc& D = r->D;
  const size_t sh = (size_t)mcmc_arm_doubles(D.K, D.d) * sizeof(double);
  hipLaunchKernelGGL(k_mcmc_prep, dim3(D.nChains), dim3(256), sh, r->ctx->stream, D, 0);
  reg2d_launch_tables(r, r->ctx->stream, 1);
  HIP_OK(r->ctx, hipGetLastError());
  return reg2d_launch_step(r, r->ctx->stream, 1);
}

int cpab_reg2d_run(cpab_reg2d* r, int nSteps, int use_graph) {
  if (!r) return CPAB_ERR_ARG;
  cpab_ctx* ctx = r->ctx;
  McmcDesc& D = r->D;
  if (nSteps < 0) return fail(ctx, CPAB_ERR_ARG, "nSteps must be >= 0");
  if (nSteps == 0) return CPAB_OK;
  hipLaunchKernelGGL(k_mcmc_set_stop, dim3((D.nChains + 255) / 256), dim3(256), 0, ctx->stream, D,
                     (int64_t)nSteps);
  HIP_OK(ctx, hipGetLastError());
  if (D.K == 1) return reg2d_launches(r, nSteps, use_graph);  // one step per launch
  // A launch decides 1..K steps per chain; launch the fewest that could end
  // the run, read the counters back, repeat.  Launches past a chain's stop
  // leave it unchanged.
  std::vector<int64_t>& hs = r->host_step;
  HIP_OK(ctx, hipMemcpyAsync(hs.data(), D.step, sizeof(int64_t) * D.nChains, hipMemcpyDeviceToHost, ctx->stream));
  HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
  std::vector<int64_t> stop(hs.begin(), hs.end());
  for (auto& v : stop) v += nSteps;
  int64_t left = nSteps;
  while (left > 0) {
    int rc = reg2d_launches(r, (left + D.K - 1) / D.K, use_graph);
    if (rc) return rc;
    HIP_OK(ctx, hipMemcpyAsync(hs.data(), D.step, sizeof(int64_t) * D.nChains, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
    left = 0;
    for (int c = 0; c < D.nChains; c++) left = stop[c] - hs[c] > left ? stop[c] - hs[c] : left;
  }
  return CPAB_OK;
}

int cpab_reg2d_last_proposal(cpab_reg2d* r, double* theta_prop, double* ll_prop) {
  if (!r) return CPAB_ERR_ARG;
  cpab_ctx* ctx = r->ctx;
  const McmcDesc& D = r->D;
  if (theta_prop)  // slot 0 of every chain
    HIP_OK(ctx, hipMemcpy2DAsync(theta_prop, sizeof(double) * D.d, D.theta_prop, sizeof(double) * D.K * D.d,
                                 sizeof(double) * D.d, D.nChains, hipMemcpyDeviceToDevice, ctx->stream));
  if (ll_prop)
    HIP_OK(ctx, hipMemcpyAsync(ll_prop, D.ll_prop, sizeof(double) * D.nChains, hipMemcpyDeviceToDevice,
                               ctx->stream));
  return CPAB_OK;
}

int cpab_reg2d_spec_depth(cpab_reg2d* r, int* K) {
  if (!r || !K) return CPAB_ERR_ARG;
  *K = r->D.K;
  return CPAB_OK;
}

int cpab_reg2d_set_spec_depth(cpab_reg2d* r, int K) {
  if (!r) return CPAB_ERR_ARG;
  cpab_ctx* ctx = r->ctx;
  McmcDesc& D = r->D;
  if (K < 1 || K > r->K_cap) return fail(ctx, CPAB_ERR_ARG, "spec_depth must be in [1, %d] for this chain set", r->K_cap);
  if (K == D.K) return CPAB_OK;
  if ((int64_t)D.nBlocks * K > 0x7fffffff) return fail(ctx, CPAB_ERR_ARG, "image too large for the launch grid");
  // launches in flight read the scratch the re-arming below rewrites; graphs
  // captured at other depths stay (the descriptor differs only in K)
  HIP_OK(ctx, hipStreamSynchronize(ctx->stream));
  D.K = K;
  // re-arm the proposals of the current step at the new depth: the same
  // numbers the chain would use, so the chain is unchanged
  const size_t sh = (size_t)mcmc_arm_doubles(D.K, D.d) * sizeof(double);
  hipLaunchKernelGGL(k_mcmc_prep, dim3(D.nChains), dim3(256), sh, ctx->stream, D, 1);
  reg2d_launch_tables(r, ctx->stream, D.K);
  HIP_OK(ctx, hipGetLastError());
  return CPAB_OK;
}

int cpab_reg2d_launch_counts(cpab_reg2d* r, int64_t* replayed, int64_t* direct) {
  if (!r) return CPAB_ERR_ARG;
  if (replayed) *replayed = r->n_replayed;
  if (direct) *direct = r->n_direct;
  return CPAB_OK;
}

int cpab_reg2d_destroy(cpab_reg2d* r) {
  if (!r) return CPAB_OK;
  (void)hipStreamSynchronize(r->ctx->stream);
  for (auto& g : r->graphs) {
    (void)hipGraphExecDestroy(g.second.second);
    (void)hipGraphDestroy(g.second.first);
  }
  if (r->cap) (void)hipStreamDestroy(r->cap);
  if (r->scratch) (void)hipFree(r->scratch);
  delete r;
  return CPAB_OK;
}

int cpab_prof_enable(cpab_ctx* ctx, int on) {
  if (!ctx) return CPAB_ERR_ARG;
  ctx->prof = on != 0;
  return CPAB_OK;
}

int cpab_prof_read(cpab_ctx* ctx, int64_t* n, double* ms) {
  if (!ctx) return CPAB_ERR_ARG;
  double total = ctx->prof_ms;
  int64_t cnt = ctx->prof_n;
  for (size_t i = 0; i < ctx->ev.size(); i++) {
    auto& p = ctx->ev[i];
    HIP_OK(ctx, hipEventSynchronize(p.second));
    float t = 0.f;
    HIP_OK(ctx, hipEventElapsedTime(&t, p.first, p.second));
    total += t;
    cnt += ctx->ev_n[i];
    ctx->ev_free.push_back(p);
  }
  ctx->ev.clear();
  ctx->ev_n.clear();
  ctx->prof_ms = 0.0;
  ctx->prof_n = 0;
  if (n) *n = cnt;
  if (ms) *ms = total;
  return CPAB_OK;
}

}  // extern "C"
