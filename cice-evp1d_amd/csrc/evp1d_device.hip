// evp1d_device.hip - gfx950 kernels and the C ABI of libevp1d.so.
//
// Device layout (one handle = one 1D point list, see include/evp1d.h):
//   vel   double2[navel]      (uvel, vvel) interleaved: one 16-byte load per
//                             neighbour gather in stress
//   sig   double[12][ldsig]   stressp_1..4, stressm_1..4, stress12_1..4, SoA
//   s15, s27, s36, s48 double2[navel]
//                             str1/str5, str2/str7, str3/str6, str4/str8: the
//                             pairs stepu gathers from one neighbour each
//                             (own, nw, sse, sw), so every gather is 16 bytes
//   skip  uint8[na]           bit0 = skipT, bit1 = skipU
//   idx   int32[6][na]        ee, ne, se, nw, sw, sse, 0-based
//   everything else           double[na] SoA, streamed once per kernel
//
// Kernels: one thread per 1D point; stress and stepu are separate phases
// because stepu at a point gathers the str of three neighbours written by
// stress in the same subcycle, and the next stress gathers the velocities of
// three neighbours written by stepu.  The subcycle loop stays on the device as
// a captured hipGraph or as one cooperative kernel with grid barriers.
#include <hip/hip_runtime.h>
#include <hip/hip_cooperative_groups.h>

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "../../include/evp1d.h"
#include "evp1d_point.hpp"

#ifndef EVP1D_FMA
#define EVP1D_FMA 0
#endif

namespace cg = cooperative_groups;
using evp1d::Scalars;

// ---------------------------------------------------------------------------
// error reporting
// ---------------------------------------------------------------------------
static thread_local int g_err_code = 0;
static thread_local char g_err_msg[512] = "";

static int fail(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err_msg, sizeof(g_err_msg), fmt, ap);
  va_end(ap);
  g_err_code = code;
  return code;
}

#define HIP_TRY(expr)                                                          \
  do {                                                                         \
    hipError_t e_ = (expr);                                                    \
    if (e_ != hipSuccess)                                                      \
      return fail(EVP1D_EHIP, "%s failed: %s (%s:%d)", #expr,                  \
                  hipGetErrorString(e_), __FILE__, __LINE__);                  \
  } while (0)

// ---------------------------------------------------------------------------
// device view of a handle (passed by value to every kernel)
// ---------------------------------------------------------------------------
struct DevView {
  int na, navel;
  long ldsig;
  const int32_t *ee, *ne, *se, *nw, *sw, *sse;
  const uint8_t *skip;
  double2 *vel;
  double *sig;
  double2 *s15, *s27, *s36, *s48;
  const double *dxT, *dyT, *strength, *hte, *htn, *htem1, *htnm1;
  const double *cw, *aiu, *uocn, *vocn, *waterx, *watery, *forcex, *forcey;
  const double *umassdti, *fm, *uarear, *uvel_init, *vvel_init, *tbu;
  double *strintx, *strinty, *cb;
  int nhalo;
  const int32_t *halo_dst, *halo_src;
  Scalars s;
};

__device__ __forceinline__ void stress_at(const DevView &d, int iw) {
  if (d.skip[iw] & 1) return;                       // bench_v2b.F90:127
  const double2 cc = d.vel[iw];
  const double2 ve = d.vel[d.ee[iw]];
  const double2 vn = d.vel[d.ne[iw]];
  const double2 vs = d.vel[d.se[iw]];
  double sig[12];
#pragma unroll
  for (int c = 0; c < 12; ++c) sig[c] = d.sig[c * d.ldsig + iw];
  evp1d::StressOut o;
  evp1d::stress_point(d.s, cc.x, cc.y, ve.x, ve.y, vs.x, vs.y, vn.x, vn.y,
                      d.dxT[iw], d.dyT[iw], d.strength[iw], d.hte[iw], d.htn[iw],
                      d.htem1[iw], d.htnm1[iw], sig, o);
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    d.sig[c * d.ldsig + iw] = o.sp[c];
    d.sig[(4 + c) * d.ldsig + iw] = o.sm[c];
    d.sig[(8 + c) * d.ldsig + iw] = o.s12[c];
  }
  d.s15[iw] = make_double2(o.str[0], o.str[4]);
  d.s27[iw] = make_double2(o.str[1], o.str[6]);
  d.s36[iw] = make_double2(o.str[2], o.str[5]);
  d.s48[iw] = make_double2(o.str[3], o.str[7]);
}

__device__ __forceinline__ void stepu_at(const DevView &d, int iw, bool write_diag) {
  if (d.skip[iw] & 2) return;                       // bench_v2b.F90:507
  const double2 v = d.vel[iw];
  const double2 a = d.s15[iw];
  const double2 b = d.s27[d.nw[iw]];
  const double2 c = d.s36[d.sse[iw]];
  const double2 e = d.s48[d.sw[iw]];
  evp1d::StepuOut o;
  evp1d::stepu_point(d.s, v.x, v.y, d.cw[iw], d.aiu[iw], d.uocn[iw], d.vocn[iw],
                     d.waterx[iw], d.watery[iw], d.forcex[iw], d.forcey[iw],
                     d.umassdti[iw], d.fm[iw], d.uarear[iw], d.uvel_init[iw],
                     d.vvel_init[iw], d.tbu[iw],
                     a.x, b.x, c.x, e.x, a.y, c.y, b.y, e.y, o);
  d.vel[iw] = make_double2(o.uvel, o.vvel);
  // strintx/strinty/Cb are overwritten every subcycle and read by nothing
  // inside the loop; only the values of the final subcycle are observable.
  if (write_diag) {
    d.strintx[iw] = o.strintx;
    d.strinty[iw] = o.strinty;
    d.cb[iw] = o.Cb;
  }
}

__device__ __forceinline__ void halo_at(const DevView &d, int k) {
  const int dst = d.halo_dst[k];
  const int src = d.halo_src[k];
  d.vel[dst] = (src >= 0) ? d.vel[src] : make_double2(0.0, 0.0);
}

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_stress(DevView d) {
  const int iw = blockIdx.x * kBlock + threadIdx.x;
  if (iw < d.na) stress_at(d, iw);
}

__global__ __launch_bounds__(kBlock) void k_stepu(DevView d, int write_diag) {
  const int iw = blockIdx.x * kBlock + threadIdx.x;
  if (iw < d.na) stepu_at(d, iw, write_diag != 0);
}

__global__ __launch_bounds__(kBlock) void k_halo(DevView d) {
  const int k = blockIdx.x * kBlock + threadIdx.x;
  if (k < d.nhalo) halo_at(d, k);
}

// Whole subcycle loop in one cooperative launch.  Every wave runs the same
// number of grid barriers (2 or 3 per subcycle) and leaves the loop together.
__global__ __launch_bounds__(kBlock) void k_persistent(DevView d, int nsub) {
  cg::grid_group grid = cg::this_grid();
  const int stride = gridDim.x * kBlock;
  const int t0 = blockIdx.x * kBlock + threadIdx.x;
  for (int k = 0; k < nsub; ++k) {
    for (int iw = t0; iw < d.na; iw += stride) stress_at(d, iw);
    grid.sync();
    const bool last = (k == nsub - 1);
    for (int iw = t0; iw < d.na; iw += stride) stepu_at(d, iw, last);
    grid.sync();
    if (d.nhalo > 0) {
      for (int g = t0; g < d.nhalo; g += stride) halo_at(d, g);
      grid.sync();
    }
  }
}

__global__ void k_put_comp(double2 *dst, const double *src, int n, int comp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    if (comp == 0) dst[i].x = src[i]; else dst[i].y = src[i];
  }
}

__global__ void k_get_comp(double *dst, const double2 *src, int n, int comp) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (comp == 0) ? src[i].x : src[i].y;
}

// ---------------------------------------------------------------------------
// handle
// ---------------------------------------------------------------------------
struct evp1d_handle {
  int na = 0, navel = 0;
  long ldsig = 0;
  int device = 0;
  evp1d_params p{};
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  std::vector<void *> allocs;
  DevView v{};
  double *scratch = nullptr;      // navel doubles, staging for component copies
  int32_t *idx_dev[6] = {};
  uint8_t *skip_dev = nullptr;
  std::vector<uint8_t> skip_host;
  bool idx_set[6] = {};
  int32_t *halo_dev = nullptr;    // 2*cap: dst then src
  int halo_cap = 0;
  int mode = EVP1D_MODE_AUTO;
  int coop_blocks = 0;
  hipGraph_t graph = nullptr;
  hipGraphExec_t graph_exec = nullptr;
  int graph_nsub = -1;
  bool graph_dirty = true;
};

static Scalars scalars_from(const evp1d_params &p) {
  Scalars s;
  s.capping = p.capping; s.e_factor = p.e_factor; s.epp2i = p.epp2i;
  s.deltaminEVP = p.deltaminEVP; s.arlx1i = p.arlx1i; s.denom1 = p.denom1;
  s.brlx = p.brlx; s.revp = p.revp; s.Ktens = p.Ktens; s.u0 = p.u0;
  s.cosw = p.cosw; s.sinw = p.sinw; s.rhow = p.rhow;
  return s;
}

static int dev_alloc(evp1d_handle *h, void **ptr, size_t bytes) {
  if (bytes == 0) bytes = 256;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess)
    return fail(EVP1D_ENOMEM, "hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
  h->allocs.push_back(*ptr);
  HIP_TRY(hipMemsetAsync(*ptr, 0, bytes, h->stream));
  return EVP1D_OK;
}

static void drop_graph(evp1d_handle *h) {
  if (h->graph_exec) (void)hipGraphExecDestroy(h->graph_exec);
  if (h->graph) (void)hipGraphDestroy(h->graph);
  h->graph_exec = nullptr;
  h->graph = nullptr;
  h->graph_nsub = -1;
}

static void release(evp1d_handle *h) {
  if (!h) return;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  drop_graph(h);
  for (void *p : h->allocs) (void)hipFree(p);
  if (h->ev0) (void)hipEventDestroy(h->ev0);
  if (h->ev1) (void)hipEventDestroy(h->ev1);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
}

// field table ---------------------------------------------------------------
enum FieldKind { FK_PLAIN, FK_VEL, FK_STR, FK_SIG };
struct FieldRef {
  FieldKind kind;
  double *plain;   // FK_PLAIN
  double2 *pair;   // FK_VEL / FK_STR
  int comp;        // component of pair
  int len;
};

static bool field_ref(evp1d_handle *h, int f, FieldRef &r) {
  DevView &v = h->v;
  r = FieldRef{FK_PLAIN, nullptr, nullptr, 0, h->na};
  auto plain = [&](const double *p) { r.kind = FK_PLAIN; r.plain = const_cast<double *>(p); };
  auto pair = [&](double2 *p, int c, FieldKind k) { r.kind = k; r.pair = p; r.comp = c; r.len = h->navel; };
  if (f >= EVP1D_STRESSP_1 && f <= EVP1D_STRESS12_4) {
    r.kind = FK_SIG;
    r.plain = v.sig + (long)(f - EVP1D_STRESSP_1) * v.ldsig;
    return true;
  }
  switch (f) {
    case EVP1D_STRENGTH: plain(v.strength); return true;
    case EVP1D_UVEL: pair(v.vel, 0, FK_VEL); return true;
    case EVP1D_VVEL: pair(v.vel, 1, FK_VEL); return true;
    case EVP1D_DXT: plain(v.dxT); return true;
    case EVP1D_DYT: plain(v.dyT); return true;
    case EVP1D_UAREAR: plain(v.uarear); return true;
    case EVP1D_CDN_OCN: plain(v.cw); return true;
    case EVP1D_AIU: plain(v.aiu); return true;
    case EVP1D_UOCN: plain(v.uocn); return true;
    case EVP1D_VOCN: plain(v.vocn); return true;
    case EVP1D_WATERX: plain(v.waterx); return true;
    case EVP1D_WATERY: plain(v.watery); return true;
    case EVP1D_FORCEX: plain(v.forcex); return true;
    case EVP1D_FORCEY: plain(v.forcey); return true;
    case EVP1D_UMASSDTI: plain(v.umassdti); return true;
    case EVP1D_FM: plain(v.fm); return true;
    case EVP1D_STRINTX: plain(v.strintx); return true;
    case EVP1D_STRINTY: plain(v.strinty); return true;
    case EVP1D_TBU: plain(v.tbu); return true;
    case EVP1D_HTE: plain(v.hte); return true;
    case EVP1D_HTN: plain(v.htn); return true;
    case EVP1D_HTEM1: plain(v.htem1); return true;
    case EVP1D_HTNM1: plain(v.htnm1); return true;
    case EVP1D_UVEL_INIT: plain(v.uvel_init); return true;
    case EVP1D_VVEL_INIT: plain(v.vvel_init); return true;
    case EVP1D_CB: plain(v.cb); return true;
    case EVP1D_STR1: pair(v.s15, 0, FK_STR); return true;
    case EVP1D_STR5: pair(v.s15, 1, FK_STR); return true;
    case EVP1D_STR2: pair(v.s27, 0, FK_STR); return true;
    case EVP1D_STR7: pair(v.s27, 1, FK_STR); return true;
    case EVP1D_STR3: pair(v.s36, 0, FK_STR); return true;
    case EVP1D_STR6: pair(v.s36, 1, FK_STR); return true;
    case EVP1D_STR4: pair(v.s48, 0, FK_STR); return true;
    case EVP1D_STR8: pair(v.s48, 1, FK_STR); return true;
    default: return false;
  }
}

static bool ready(const evp1d_handle *h) {
  for (int k = 0; k < 6; ++k)
    if (!h->idx_set[k]) return false;
  return true;
}

// ---------------------------------------------------------------------------
// launches
// ---------------------------------------------------------------------------
static inline int blocks_for(int n) { return (n + kBlock - 1) / kBlock; }

static int launch_stress(evp1d_handle *h) {
  if (h->na == 0) return EVP1D_OK;
  hipLaunchKernelGGL(k_stress, dim3(blocks_for(h->na)), dim3(kBlock), 0, h->stream, h->v);
  HIP_TRY(hipGetLastError());
  return EVP1D_OK;
}

static int launch_stepu(evp1d_handle *h, bool write_diag) {
  if (h->na == 0) return EVP1D_OK;
  hipLaunchKernelGGL(k_stepu, dim3(blocks_for(h->na)), dim3(kBlock), 0, h->stream, h->v,
                     write_diag ? 1 : 0);
  HIP_TRY(hipGetLastError());
  return EVP1D_OK;
}

static int launch_halo(evp1d_handle *h) {
  if (h->v.nhalo == 0) return EVP1D_OK;
  hipLaunchKernelGGL(k_halo, dim3(blocks_for(h->v.nhalo)), dim3(kBlock), 0, h->stream, h->v);
  HIP_TRY(hipGetLastError());
  return EVP1D_OK;
}

static int issue_subcycles(evp1d_handle *h, int nsub) {
  for (int k = 0; k < nsub; ++k) {
    int rc = launch_stress(h);
    if (rc) return rc;
    rc = launch_stepu(h, k == nsub - 1);
    if (rc) return rc;
    rc = launch_halo(h);
    if (rc) return rc;
  }
  return EVP1D_OK;
}

static int run_graph(evp1d_handle *h, int nsub) {
  if (h->graph_dirty || h->graph_nsub != nsub || !h->graph_exec) {
    drop_graph(h);
    HIP_TRY(hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    int rc = issue_subcycles(h, nsub);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(h->stream, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess)
      return fail(EVP1D_EHIP, "hipStreamEndCapture failed: %s", hipGetErrorString(e));
    h->graph = g;
    HIP_TRY(hipGraphInstantiate(&h->graph_exec, h->graph, nullptr, nullptr, 0));
    h->graph_nsub = nsub;
    h->graph_dirty = false;
  }
  HIP_TRY(hipGraphLaunch(h->graph_exec, h->stream));
  return EVP1D_OK;
}

static int run_persistent(evp1d_handle *h, int nsub) {
  if (h->coop_blocks <= 0) {
    int per_cu = 0, cus = 0;
    HIP_TRY(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_persistent, kBlock, 0));
    HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    if (per_cu <= 0 || cus <= 0) return fail(EVP1D_EHIP, "no co-resident blocks for k_persistent");
    h->coop_blocks = per_cu * cus;
  }
  const int needed = std::max(blocks_for(h->na), 1);
  int grid = std::min(h->coop_blocks, needed);
  DevView v = h->v;
  void *args[] = {&v, &nsub};
  HIP_TRY(hipLaunchCooperativeKernel((const void *)k_persistent, dim3(grid), dim3(kBlock), args, 0,
                                     h->stream));
  return EVP1D_OK;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" {

// error channel for the host-only translation units of the library
int evp1d__set_error(int code, const char *msg) { return fail(code, "%s", msg); }

int evp1d_set_evp_parameters(int32_t ndte, double dt, evp1d_params *p) {
  (void)dt;   // unused, as in src/ice.F90:119 (dtei is commented out)
  if (!p) return fail(EVP1D_EINVAL, "evp1d_set_evp_parameters: p is NULL");
  if (ndte < 1) return fail(EVP1D_EINVAL, "evp1d_set_evp_parameters: ndte=%d < 1", ndte);
  // src/ice.F90:80-107 (parameters) and 109-130 (derived values)
  const double c1 = 1.0, c2 = 2.0;
  const double elasticDamp = 0.36, e_plasticpot = c2, e_yieldcurve = c2;
  std::memset(p, 0, sizeof(*p));
  p->ndte = ndte;
  p->epp2i = c1 / (e_plasticpot * e_plasticpot);
  p->capping = c1;
  p->e_factor = (e_yieldcurve * e_yieldcurve) / (e_plasticpot * e_plasticpot * e_plasticpot * e_plasticpot);
  p->arlx = c2 * elasticDamp * (double)ndte;
  p->arlx1i = c1 / p->arlx;
  p->brlx = (double)ndte;
  p->denom1 = c1 / (c1 + p->arlx1i);
  p->deltaminEVP = 10e-9;
  p->revp = 0.0;
  p->Ktens = 0.0;
  p->u0 = 5e-5;
  p->cosw = 1.0;
  p->sinw = 0.0;
  p->rhow = 1026.0;
  return EVP1D_OK;
}

int evp1d_create(int32_t na, int32_t navel, const evp1d_params *p, evp1d_handle **out) {
  if (!out) return fail(EVP1D_EINVAL, "evp1d_create: out is NULL");
  *out = nullptr;
  if (!p) return fail(EVP1D_EINVAL, "evp1d_create: params is NULL");
  if (na < 0 || navel < na)
    return fail(EVP1D_EINVAL, "evp1d_create: need 0 <= na <= navel (na=%d navel=%d)", na, navel);
  evp1d_handle *h = new (std::nothrow) evp1d_handle();
  if (!h) return fail(EVP1D_ENOMEM, "evp1d_create: host allocation failed");
  h->na = na;
  h->navel = navel;
  h->p = *p;
  h->ldsig = ((long)na + 63) / 64 * 64;
  int rc = EVP1D_OK;
  auto chk = [&](hipError_t e, const char *what) {
    if (e != hipSuccess && rc == EVP1D_OK) rc = fail(EVP1D_EHIP, "%s: %s", what, hipGetErrorString(e));
  };
  chk(hipGetDevice(&h->device), "hipGetDevice");
  chk(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking), "hipStreamCreate");
  chk(hipEventCreate(&h->ev0), "hipEventCreate");
  chk(hipEventCreate(&h->ev1), "hipEventCreate");
  if (rc) { release(h); return rc; }

  DevView &v = h->v;
  v.na = na;
  v.navel = navel;
  v.ldsig = h->ldsig;
  const size_t nA = (size_t)na * sizeof(double), nV = (size_t)navel * sizeof(double2);
  auto A = [&](void **ptr, size_t bytes) { if (!rc) rc = dev_alloc(h, ptr, bytes); };
  A((void **)&v.vel, nV);
  A((void **)&v.s15, nV); A((void **)&v.s27, nV); A((void **)&v.s36, nV); A((void **)&v.s48, nV);
  A((void **)&v.sig, (size_t)12 * h->ldsig * sizeof(double));
  const double **cplain[] = {&v.dxT, &v.dyT, &v.strength, &v.hte, &v.htn, &v.htem1, &v.htnm1,
                             &v.cw, &v.aiu, &v.uocn, &v.vocn, &v.waterx, &v.watery, &v.forcex,
                             &v.forcey, &v.umassdti, &v.fm, &v.uarear, &v.uvel_init, &v.vvel_init,
                             &v.tbu};
  for (const double **pp : cplain) A((void **)pp, nA);
  A((void **)&v.strintx, nA); A((void **)&v.strinty, nA); A((void **)&v.cb, nA);
  for (int k = 0; k < 6; ++k) A((void **)&h->idx_dev[k], (size_t)na * sizeof(int32_t));
  A((void **)&h->skip_dev, (size_t)na);
  A((void **)&h->scratch, (size_t)navel * sizeof(double));
  if (rc) { release(h); return rc; }
  v.ee = h->idx_dev[0]; v.ne = h->idx_dev[1]; v.se = h->idx_dev[2];
  v.nw = h->idx_dev[3]; v.sw = h->idx_dev[4]; v.sse = h->idx_dev[5];
  v.skip = h->skip_dev;
  v.nhalo = 0;
  v.s = scalars_from(*p);
  h->skip_host.assign((size_t)na, 0);
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    release(h);
    return fail(EVP1D_EHIP, "evp1d_create: initial clear failed");
  }
  *out = h;
  return EVP1D_OK;
}

int evp1d_destroy(evp1d_handle *h) {
  release(h);
  return EVP1D_OK;
}

int evp1d_set_params(evp1d_handle *h, const evp1d_params *p) {
  if (!h || !p) return fail(EVP1D_EINVAL, "evp1d_set_params: NULL argument");
  h->p = *p;
  h->v.s = scalars_from(*p);
  h->graph_dirty = true;
  return EVP1D_OK;
}

int evp1d_get_params(const evp1d_handle *h, evp1d_params *p) {
  if (!h || !p) return fail(EVP1D_EINVAL, "evp1d_get_params: NULL argument");
  *p = h->p;
  return EVP1D_OK;
}

int evp1d_get_sizes(const evp1d_handle *h, int32_t *na, int32_t *navel) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_sizes: NULL handle");
  if (na) *na = h->na;
  if (navel) *navel = h->navel;
  return EVP1D_OK;
}

int evp1d_set_field(evp1d_handle *h, int32_t field, const double *src, int32_t n) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_field: NULL handle");
  FieldRef r;
  if (!field_ref(h, field, r)) return fail(EVP1D_EINVAL, "evp1d_set_field: unknown field %d", field);
  if (n != r.len) return fail(EVP1D_EINVAL, "evp1d_set_field: field %d has length %d, got %d", field, r.len, n);
  if (n == 0) return EVP1D_OK;
  if (!src) return fail(EVP1D_EINVAL, "evp1d_set_field: src is NULL");
  if (r.kind == FK_PLAIN || r.kind == FK_SIG) {
    HIP_TRY(hipMemcpyAsync(r.plain, src, (size_t)n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  } else {
    HIP_TRY(hipMemcpyAsync(h->scratch, src, (size_t)n * sizeof(double), hipMemcpyHostToDevice, h->stream));
    hipLaunchKernelGGL(k_put_comp, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, r.pair,
                       (const double *)h->scratch, n, r.comp);
    HIP_TRY(hipGetLastError());
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

int evp1d_get_field(evp1d_handle *h, int32_t field, double *dst, int32_t n) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_get_field: NULL handle");
  FieldRef r;
  if (!field_ref(h, field, r)) return fail(EVP1D_EINVAL, "evp1d_get_field: unknown field %d", field);
  if (n != r.len) return fail(EVP1D_EINVAL, "evp1d_get_field: field %d has length %d, got %d", field, r.len, n);
  if (n == 0) return EVP1D_OK;
  if (!dst) return fail(EVP1D_EINVAL, "evp1d_get_field: dst is NULL");
  if (r.kind == FK_PLAIN || r.kind == FK_SIG) {
    HIP_TRY(hipMemcpyAsync(dst, r.plain, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  } else {
    hipLaunchKernelGGL(k_get_comp, dim3(blocks_for(n)), dim3(kBlock), 0, h->stream, h->scratch,
                       (const double2 *)r.pair, n, r.comp);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(dst, h->scratch, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  }
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

int evp1d_set_index(evp1d_handle *h, int32_t which, const int32_t *idx, int32_t n) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_index: NULL handle");
  if (which < EVP1D_EE || which > EVP1D_SSE) return fail(EVP1D_EINVAL, "evp1d_set_index: unknown list %d", which);
  if (n != h->na) return fail(EVP1D_EINVAL, "evp1d_set_index: length %d != na %d", n, h->na);
  if (n > 0 && !idx) return fail(EVP1D_EINVAL, "evp1d_set_index: idx is NULL");
  std::vector<int32_t> z((size_t)n);
  for (int i = 0; i < n; ++i) {
    const int32_t q = idx[i];
    if (q < 1 || q > h->navel)
      return fail(EVP1D_EINDEX, "evp1d_set_index: list %d entry %d = %d outside 1..%d", which, i + 1, q,
                  h->navel);
    z[(size_t)i] = q - 1;
  }
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(h->idx_dev[which], z.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice,
                           h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  h->idx_set[which] = true;
  return EVP1D_OK;
}

int evp1d_set_skip(evp1d_handle *h, int32_t which, const int32_t *flags, int32_t n) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_skip: NULL handle");
  if (which != EVP1D_SKIPT && which != EVP1D_SKIPU) return fail(EVP1D_EINVAL, "evp1d_set_skip: unknown flag %d", which);
  if (n != h->na) return fail(EVP1D_EINVAL, "evp1d_set_skip: length %d != na %d", n, h->na);
  if (n > 0 && !flags) return fail(EVP1D_EINVAL, "evp1d_set_skip: flags is NULL");
  const uint8_t bit = (uint8_t)(1u << which);
  for (int i = 0; i < n; ++i) {
    uint8_t &b = h->skip_host[(size_t)i];
    b = (uint8_t)(flags[i] ? (b | bit) : (b & ~bit));
  }
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(h->skip_dev, h->skip_host.data(), (size_t)n, hipMemcpyHostToDevice, h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  return EVP1D_OK;
}

int evp1d_set_halo(evp1d_handle *h, int32_t n, const int32_t *dst, const int32_t *src) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_halo: NULL handle");
  if (n < 0) return fail(EVP1D_EINVAL, "evp1d_set_halo: n=%d < 0", n);
  if (n > 0 && (!dst || !src)) return fail(EVP1D_EINVAL, "evp1d_set_halo: NULL list");
  std::vector<int32_t> buf((size_t)2 * n);
  std::vector<char> seen((size_t)(h->navel - h->na), 0);
  for (int k = 0; k < n; ++k) {
    if (dst[k] <= h->na || dst[k] > h->navel)
      return fail(EVP1D_EINDEX, "evp1d_set_halo: dst[%d]=%d not a ghost slot (%d..%d)", k + 1, dst[k], h->na + 1,
                  h->navel);
    if (src[k] < 0 || src[k] > h->na)
      return fail(EVP1D_EINDEX, "evp1d_set_halo: src[%d]=%d outside 0..%d", k + 1, src[k], h->na);
    char &s = seen[(size_t)(dst[k] - h->na - 1)];
    if (s) return fail(EVP1D_EINDEX, "evp1d_set_halo: ghost slot %d listed twice", dst[k]);
    s = 1;
    buf[(size_t)k] = dst[k] - 1;
    buf[(size_t)n + k] = src[k] - 1;
  }
  if (n > h->halo_cap) {
    int rc = dev_alloc(h, (void **)&h->halo_dev, (size_t)2 * n * sizeof(int32_t));
    if (rc) return rc;
    h->halo_cap = n;
  }
  if (n > 0) {
    HIP_TRY(hipMemcpyAsync(h->halo_dev, buf.data(), (size_t)2 * n * sizeof(int32_t), hipMemcpyHostToDevice,
                           h->stream));
    HIP_TRY(hipStreamSynchronize(h->stream));
  }
  h->v.nhalo = n;
  h->v.halo_dst = h->halo_dev;
  h->v.halo_src = h->halo_dev ? h->halo_dev + n : nullptr;
  h->graph_dirty = true;
  return EVP1D_OK;
}

int evp1d_stress(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_stress: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_stress: neighbour lists not set");
  return launch_stress(h);
}

int evp1d_stepu(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_stepu: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_stepu: neighbour lists not set");
  return launch_stepu(h, true);
}

int evp1d_halo(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_halo: NULL handle");
  return launch_halo(h);
}

int evp1d_set_mode(evp1d_handle *h, int32_t mode) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_set_mode: NULL handle");
  if (mode < EVP1D_MODE_LAUNCH || mode > EVP1D_MODE_AUTO) return fail(EVP1D_EINVAL, "evp1d_set_mode: bad mode %d", mode);
  h->mode = mode;
  return EVP1D_OK;
}

int evp1d_get_mode(const evp1d_handle *h, int32_t *mode) {
  if (!h || !mode) return fail(EVP1D_EINVAL, "evp1d_get_mode: NULL argument");
  *mode = h->mode;
  return EVP1D_OK;
}

int evp1d_run(evp1d_handle *h, int32_t nsub) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_run: NULL handle");
  if (nsub < 0) return fail(EVP1D_EINVAL, "evp1d_run: nsub=%d < 0", nsub);
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_run: neighbour lists not set");
  if (nsub == 0 || h->na == 0) return EVP1D_OK;
  switch (h->mode) {
    case EVP1D_MODE_LAUNCH: return issue_subcycles(h, nsub);
    case EVP1D_MODE_PERSISTENT: return run_persistent(h, nsub);
    case EVP1D_MODE_GRAPH:
    case EVP1D_MODE_AUTO:
    default: return run_graph(h, nsub);
  }
}

int evp1d_run_individual(evp1d_handle *h, int32_t which, int32_t nsub) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_run_individual: NULL handle");
  if (!ready(h)) return fail(EVP1D_ESTATE, "evp1d_run_individual: neighbour lists not set");
  if (which != 0 && which != 1) return fail(EVP1D_EINVAL, "evp1d_run_individual: which=%d", which);
  for (int k = 0; k < nsub; ++k) {
    const int rc = which == 0 ? launch_stress(h) : launch_stepu(h, k == nsub - 1);
    if (rc) return rc;
  }
  return EVP1D_OK;
}

int evp1d_synchronize(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_synchronize: NULL handle");
  HIP_TRY(hipStreamSynchronize(h->stream));
  return EVP1D_OK;
}

int evp1d_timer_start(evp1d_handle *h) {
  if (!h) return fail(EVP1D_EINVAL, "evp1d_timer_start: NULL handle");
  HIP_TRY(hipEventRecord(h->ev0, h->stream));
  return EVP1D_OK;
}

int evp1d_timer_stop(evp1d_handle *h, float *ms) {
  if (!h || !ms) return fail(EVP1D_EINVAL, "evp1d_timer_stop: NULL argument");
  HIP_TRY(hipEventRecord(h->ev1, h->stream));
  HIP_TRY(hipEventSynchronize(h->ev1));
  HIP_TRY(hipEventElapsedTime(ms, h->ev0, h->ev1));
  return EVP1D_OK;
}

int evp1d_last_error(char *buf, int32_t n) {
  if (buf && n > 0) {
    std::strncpy(buf, g_err_msg, (size_t)n - 1);
    buf[n - 1] = '\0';
  }
  return g_err_code;
}

int evp1d_build_info(int32_t *abi_version, int32_t *fma) {
  if (abi_version) *abi_version = EVP1D_ABI_VERSION;
  if (fma) *fma = EVP1D_FMA;
  return EVP1D_OK;
}

}  // extern "C"
