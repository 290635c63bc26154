// txg_api.hip -- host side of the C ABI declared in include/taxila_lbm.h.
//
// A handle owns one lattice domain on the current HIP device: two population
// buffers (post-collision state, ping-pong), the stencil density, per-node
// link masks and mineral tags, and a stream.  The reference procedures each
// entry point replaces are cited in the header.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "taxila_lbm.h"
#include "txg_dispatch.h"
#include "txg_kernels.h"

using namespace txg;

namespace {

thread_local std::string g_create_error;

// reduction kernels (FP64, deterministic: fixed per-block order, then a
// single-block pass in index order)
__global__ void k_partial_sum(const double* __restrict__ a, const uint32_t* __restrict__ mask, int n,
                              double* __restrict__ part) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256)
    if (!(mask[i] & kSelfSolid)) s += a[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_final_sum(const double* __restrict__ part, int nb, double* __restrict__ out) {
  __shared__ double sh[256];
  double s = 0.0;
  for (int i = threadIdx.x; i < nb; i += 256) s += part[i];
  sh[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = sh[0];
}

// momentum of the common-velocity definition: j + F/2
__global__ void k_j_half_f(double* __restrict__ out, const double* __restrict__ j, const double* __restrict__ F, int n) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i < n) out[i] = j[i] + .5 * F[i];
}

// max |(old - new)/new| over node-components with new != 0; old := new
__global__ void k_delta(double* __restrict__ old, const double* __restrict__ cur, int n, int first,
                        double* __restrict__ part) {
  __shared__ double sh[256];
  double mx = 0.0;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) {
    double r = cur[i];
    double o = first ? 0.0 : old[i];
    double v = o + -1.0 * r;
    double q = (r != 0.0) ? v / r : 0.0;
    mx = fmax(mx, fabs(q));
    old[i] = r;
  }
  sh[threadIdx.x] = mx;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) sh[threadIdx.x] = fmax(sh[threadIdx.x], sh[threadIdx.x + w]);
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

template <class R>
__global__ void k_fill_rho(R* __restrict__ rho, const double* __restrict__ src, const uint32_t* __restrict__ mask,
                           int N, int Np, int C) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  bool solid = (mask[i] & kSelfSolid) != 0;
  for (int m = 0; m < C; m++) {
    R v = solid ? (R)NAN : (R)(src ? src[(size_t)i * C + m] : 0.0);
    rho[(size_t)m * Np + i] = v;
  }
}

// AoS FP64 fi(1:C,0:b,node) -> SoA storage
template <class S, class R>
__global__ void k_load_fi(S* __restrict__ f, const double* __restrict__ src, int N, int Np, int C, int Q,
                          const R* __restrict__ shift) {
  int i = blockIdx.x * 256 + threadIdx.x;
  if (i >= N) return;
  for (int n = 0; n < Q; n++)
    for (int m = 0; m < C; m++) {
      R v = (R)src[((size_t)i * Q + n) * C + m];
      st<S, R>(f + (size_t)(m * Q + n) * Np, i, v, shift[m * Q + n]);
    }
}

inline int cdiv(long a, int b) { return (int)((a + b - 1) / b); }

}  // namespace

struct txg_handle {
  txg_desc d;
  int D = 3, Q = 19, C = 1, N = 0, Np = 0;
  int prec = TXG_FP64;
  int device = 0;
  hipStream_t stream = nullptr;
  KernelSet ks{};
  size_t ssize = 8;  // storage bytes per population
  void* f[2] = {nullptr, nullptr};
  int cur = 0;
  void* rho = nullptr;  // R[C*Np]
  uint32_t* mask = nullptr;
  uint8_t* tag = nullptr;
  void* u_ic = nullptr;  // R[N*D*C]
  void* shift = nullptr; // R[C*Q] storage shifts for k_load_fi
  // FP64 export arrays
  double* ex = nullptr;
  size_t ex_len = 0;
  double* red = nullptr;  // reduction scratch
  double* rho_old = nullptr;
  int have_old = 0;
  Params<double> p64{};
  Params<float> p32{};
  std::vector<uint32_t> hmask;
  std::vector<uint8_t> htag;
  std::vector<int32_t> bb_node, bb_dir, fs_node, fs_dir, fs_min;
  // event timing of the step passes (txg_profile_enable)
  int prof = 0;
  int ev_next = 0;
  std::vector<hipEvent_t> ev_pool;
  std::vector<std::pair<int, int>> ev_pairs[2];  // indices into ev_pool per pass
  int walls_set = 0, state_set = 0, fi_set = 0, initialized = 0;
  int local_state = 0;  // f[cur] holds post-stream populations (after init)
  int rho_valid = 0;    // rho[] holds the stencil density of the current state
  int kernel = 0;       // txg_tune: 0 auto, 1 two-pass, 2 one-pass
  int tile = 0;         // index into kTileX/kTileY
  int zseg = 0;         // z-planes per one-pass workgroup (0: default)
  std::string err;

  int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    err = buf;
    return code;
  }
  int hip(hipError_t e, const char* what) {
    if (e == hipSuccess) return TXG_OK;
    return fail(TXG_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
  }
  const void* params() const { return prec == TXG_FP64 ? (const void*)&p64 : (const void*)&p32; }
};

namespace {

template <class S, class R>
Bufs<S, R> make_bufs(txg_handle* h, int in) {
  Bufs<S, R> b;
  b.fin = (const S*)h->f[in];
  b.fout = (S*)h->f[1 - in];
  b.rho = (R*)h->rho;
  b.mask = h->mask;
  b.tag = h->tag;
  return b;
}

// all kernels take a Bufs by pointer; this holds it for the call
struct AnyBufs {
  Bufs<double, double> b64;
  Bufs<float, float> b32;
  Bufs<__half, float> b16;
  const void* ptr = nullptr;
};

AnyBufs bufs(txg_handle* h, int in) {
  AnyBufs a;
  if (h->prec == TXG_FP64) {
    a.b64 = make_bufs<double, double>(h, in);
    a.ptr = &a.b64;
  } else if (h->prec == TXG_FP32) {
    a.b32 = make_bufs<float, float>(h, in);
    a.ptr = &a.b32;
  } else {
    a.b16 = make_bufs<__half, float>(h, in);
    a.ptr = &a.b16;
  }
  return a;
}

template <class R>
void fill_params(const txg_desc& d, int N, int Np, int flags, int dirichlet, Params<R>& p) {
  memset(&p, 0, sizeof(p));
  p.nx = d.nx;
  p.ny = d.ny;
  p.nz = d.ndims == 3 ? d.nz : 1;
  p.nxy = d.nx * d.ny;
  p.N = N;
  p.Np = Np;
  for (int i = 0; i < 3; i++) p.periodic[i] = d.periodic[i] ? 1 : 0;
  if (d.ndims == 2) p.periodic[2] = 0;
  p.flags = flags;
  p.nminerals = d.nminerals;
  p.dirichlet = dirichlet;
  const int C = d.ncomponents;
  for (int m = 0; m < C; m++) {
    const double cs2 = 1.0 / 3.;                 // lbm_relaxation.F90:79
    const double dk = 1. - 2. / (3. * d.mm[m]);  // lbm_component.F90:158
    const bool srt = d.relaxation_mode == TXG_SRT;
    const double tau = srt ? d.tau[m] : 1.0;
    const double s_c = srt ? 1.0 / d.tau[m] : d.s_c[m];  // lbm_relaxation.F90:133
    p.tau[m] = (R)tau;
    p.mmot[m] = (R)(d.mm[m] * s_c);
    p.mm[m] = (R)d.mm[m];
    p.dk[m] = (R)dk;
    p.cs2[m] = (R)cs2;
    p.a15[m] = (R)(1.5 * (1. - dk));
    p.two_cs2[m] = (R)(2. * cs2);
    p.two_cs2_sq[m] = (R)(2. * cs2 * cs2);
    p.feq0a[m] = (R)((1. + dk * 5.) / 6.);
    p.omega[m] = (R)(1.0 / tau);
    p.ic2[m] = (R)(1.0 / cs2);
    p.ic45[m] = (R)(1.0 / (2. * cs2 * cs2));
    p.ic15[m] = (R)(1.0 / (2. * cs2));
    const double rates[7] = {d.s_c[m], d.s_e[m], d.s_e2[m], d.s_q[m], d.s_nu[m], d.s_pi[m], d.s_m[m]};
    for (int k = 0; k < 7; k++) p.rate[m][k] = (R)rates[k];
    for (int n = 0; n < C; n++) p.g[m][n] = (R)d.g[m][n];
    for (int k = 0; k < d.nminerals; k++) p.gw[k][m] = (R)d.gw[k][m];
    for (int f = 0; f < 6; f++) p.bcrho[f][m] = (R)d.bc_density[f][m];
    p.shift0[m] = (R)0;
  }
  for (int i = 0; i < 3; i++) p.gvt[i] = (R)d.gvt[i];
  double ffw[16];
  ff_weights(d.ndims, d.isotropy_order, ffw);
  for (int k = 0; k < 11; k++) {
    p.fw[k] = (R)ffw[k];
    p.fw2[k] = (R)(2. * ffw[k]);
    p.fw3[k] = (R)(3. * ffw[k]);
    p.fwt4[k] = (R)(ffw[k] * 4.);
    p.fwt9[k] = (R)(ffw[k] * 9.);
  }
  p.c0 = (R)6.0;
  p.ffeps = (R)(double)1.e-12f;
  p.kax = (R)(d.ndims == 3 ? (double)(1.f / 6.f) : (double)(1.f / 3.f));
  p.kdg = (R)(double)(1.f / 12.f);
}

const int kC3[19][3] = {{0, 0, 0},  {1, 0, 0},  {0, 1, 0},  {-1, 0, 0}, {0, -1, 0}, {0, 0, 1},  {0, 0, -1},
                        {1, 1, 0},  {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}, {1, 0, 1},  {-1, 0, 1}, {-1, 0, -1},
                        {1, 0, -1}, {0, 1, 1},  {0, -1, 1}, {0, -1, -1}, {0, 1, -1}};
const int kC2[9][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {-1, 0, 0}, {0, -1, 0},
                       {1, 1, 0}, {-1, 1, 0}, {-1, -1, 0}, {1, -1, 0}};

int ensure_export(txg_handle* h) {
  size_t need = (size_t)h->N * ((size_t)h->C * h->Q + h->C + h->D + 2 * (size_t)h->C * h->D);
  if (h->ex && h->ex_len >= need) return TXG_OK;
  if (h->ex) (void)hipFree(h->ex);
  h->ex = nullptr;
  if (hipMalloc(&h->ex, need * sizeof(double)) != hipSuccess)
    return h->fail(TXG_ERR_NOMEM, "export buffer allocation failed (%zu doubles)", need);
  h->ex_len = need;
  return TXG_OK;
}

bool use_fused(const txg_handle* h) {
  if (h->kernel == 1) return false;
  // the one-pass kernel addresses each component's populations with 32-bit
  // byte offsets
  if ((uint64_t)h->Q * h->Np * h->ssize >= (1ull << 32)) return false;
  return h->ks.fused[h->tile] != nullptr;
}

FusedGeom fused_geom(const txg_handle* h) {
  FusedGeom g;
  const int tx = kTileX[h->tile], ty = kTileY[h->tile];
  const int nz = h->D == 3 ? h->d.nz : 1;
  g.ntx = (h->d.nx + tx - 1) / tx;
  g.nty = (h->d.ny + ty - 1) / ty;
  g.ntiles = g.ntx * g.nty;
  int zs = h->zseg;
  if (zs <= 0) {
    // aim for about 4 workgroups per CU on 256 CUs, segments of >= 8 planes
    const int want = 1024;
    int nseg = std::max(1, want / std::max(1, g.ntiles));
    zs = std::max(8, (nz + nseg - 1) / nseg);
  }
  g.zseg = std::min(zs, nz);
  g.nseg = (nz + g.zseg - 1) / g.zseg;
  g.nblocks = g.ntiles * g.nseg;
  return g;
}

// make rho[] the stencil density of the current state (the one-pass step
// keeps it in LDS only)
int ensure_rho(txg_handle* h) {
  if (h->local_state || h->rho_valid) return TXG_OK;
  AnyBufs b = bufs(h, h->cur);
  int rc = h->hip(h->ks.density(h->params(), b.ptr, 1, 1, h->N, h->stream), "density kernel");
  if (rc) return rc;
  h->rho_valid = 1;
  return TXG_OK;
}

ExportPtrs export_ptrs(txg_handle* h) {
  ExportPtrs e;
  size_t N = h->N;
  e.f = h->ex;
  e.rho = e.f + N * h->C * h->Q;
  e.up = e.rho + N * h->C;
  e.F = e.up + N * h->D;
  e.j = e.F + N * h->C * h->D;
  return e;
}

// Compute the current reference state (f, rho, u', F, j) into the export
// arrays (FP64, SoA).
int run_export(txg_handle* h) {
  if (!h->initialized) return h->fail(TXG_ERR_STATE, "txg_init has not been called");
  int rc = ensure_export(h);
  if (rc) return rc;
  rc = ensure_rho(h);
  if (rc) return rc;
  ExportPtrs e = export_ptrs(h);
  AnyBufs b = bufs(h, h->cur);
  int pulled = h->local_state ? 0 : 1;
  rc = h->hip(h->ks.collide(h->params(), b.ptr, pulled, pulled, &e, h->N, h->stream), "export kernel");
  if (rc) return rc;
  return h->hip(hipStreamSynchronize(h->stream), "export sync");
}

}  // namespace

extern "C" {

int txg_desc_init(txg_desc* d) {
  if (!d) return TXG_ERR_ARG;
  memset(d, 0, sizeof(*d));
  d->ndims = 3;
  d->ncomponents = 1;
  d->nx = d->ny = d->nz = 1;
  d->isotropy_order = 4;
  d->relaxation_mode = TXG_SRT;
  d->precision = TXG_FP64;
  d->nminerals = 0;
  for (int m = 0; m < TXG_MAX_COMPONENTS; m++) {
    d->mm[m] = 1.0;
    d->tau[m] = 1.0;
    d->s_c[m] = d->s_e[m] = d->s_e2[m] = d->s_q[m] = d->s_nu[m] = d->s_pi[m] = d->s_m[m] = 1.0;
  }
  return TXG_OK;
}

int txg_create(const txg_desc* d, txg_handle** out) {
  g_create_error.clear();
  auto bad = [](int code, const char* msg) {
    g_create_error = msg;
    return code;
  };
  if (!d || !out) return bad(TXG_ERR_ARG, "null argument");
  *out = nullptr;
  if (d->ndims != 2 && d->ndims != 3) return bad(TXG_ERR_ARG, "ndims must be 2 or 3");
  if (d->ncomponents < 1 || d->ncomponents > TXG_MAX_COMPONENTS) return bad(TXG_ERR_ARG, "ncomponents out of range");
  if (d->ncomponents > 2) return bad(TXG_ERR_UNSUPPORTED, "the HIP step is built for 1 or 2 components");
  if (d->nx < 3 || d->ny < 3 || (d->ndims == 3 && d->nz < 3)) return bad(TXG_ERR_ARG, "domain must be at least 3 per side");
  const int iso = d->isotropy_order;
  if (!(iso == 0 || iso == 4 || iso == 8 || (iso == 10 && d->ndims == 2)))
    return bad(TXG_ERR_ARG, "isotropy_order must be 0, 4, 8 (or 10 for D2Q9)");
  if (d->relaxation_mode != TXG_SRT && d->relaxation_mode != TXG_MRT) return bad(TXG_ERR_ARG, "bad relaxation_mode");
  if (d->precision < TXG_FP64 || d->precision > TXG_FP32_FP16) return bad(TXG_ERR_ARG, "bad precision");
  if (d->nminerals < 0 || d->nminerals > TXG_MAX_MINERALS) return bad(TXG_ERR_ARG, "nminerals out of range");
  long long N = (long long)d->nx * d->ny * (d->ndims == 3 ? d->nz : 1);
  const int Q = d->ndims == 3 ? 19 : 9;
  if (N * Q * d->ncomponents >= (1LL << 31) - 4096)
    return bad(TXG_ERR_UNSUPPORTED, "domain too large for 32-bit population indexing");
  for (int m = 0; m < d->ncomponents; m++) {
    if (!(d->mm[m] > 0)) return bad(TXG_ERR_ARG, "mm must be positive");
    if (d->relaxation_mode == TXG_SRT && !(d->tau[m] > 0.5)) return bad(TXG_ERR_ARG, "tau must exceed 1/2");
  }
  int dirichlet = 0;
  for (int f = 0; f < 6; f++) {
    const int dim = f / 2;
    if (dim >= d->ndims) continue;
    if (d->bc_flags[f] == TXG_BC_DENSITY) {
      if (d->periodic[dim]) return bad(TXG_ERR_ARG, "density BC on a periodic face");
      dirichlet |= 1 << f;
    } else if (d->bc_flags[f] != TXG_BC_NULL) {
      return bad(TXG_ERR_UNSUPPORTED, "only periodic, wall and density faces are implemented");
    }
  }

  txg_handle* h = new txg_handle();
  h->d = *d;
  h->D = d->ndims;
  h->Q = Q;
  h->C = d->ncomponents;
  h->N = (int)N;
  h->Np = (int)((N + 255) / 256 * 256);
  h->prec = d->precision;
  h->ssize = h->prec == TXG_FP64 ? 8 : (h->prec == TXG_FP32 ? 4 : 2);
  bool ok = h->prec == TXG_FP64 ? kernels_fp64(h->D, h->C, iso, &h->ks)
                                : kernels_fp32(h->D, h->C, iso, h->prec == TXG_FP32_FP16, &h->ks);
  if (!ok) {
    delete h;
    return bad(TXG_ERR_UNSUPPORTED, "no kernel instantiation for this lattice/components/isotropy");
  }
  // forcing switches (lbm_flow.F90:209-215, lbm_walls.F90:117-125, eps 1.e-15 default real)
  const double eps15 = (double)1.e-15f;
  int flags = 0;
  for (int m = 0; m < h->C; m++)
    for (int n = 0; n < h->C; n++)
      if (std::fabs(d->g[m][n]) > eps15) flags |= kFlagFF;
  for (int k = 0; k < d->nminerals; k++)
    for (int m = 0; m < h->C; m++)
      if (std::fabs(d->gw[k][m]) > eps15) flags |= kFlagFS;
  if (d->body_forces) flags |= kFlagBody;
  if (d->relaxation_mode == TXG_MRT) flags |= kFlagMRT;
  if (iso == 0) flags &= ~kFlagFF;
  fill_params<double>(*d, h->N, h->Np, flags, dirichlet, h->p64);
  fill_params<float>(*d, h->N, h->Np, flags, dirichlet, h->p32);
  if (h->prec == TXG_FP32_FP16)
    for (int m = 0; m < h->C; m++) h->p32.shift0[m] = (float)d->rho0[m];

  if (hipGetDevice(&h->device) != hipSuccess || hipStreamCreate(&h->stream) != hipSuccess) {
    delete h;
    return bad(TXG_ERR_HIP, "no HIP device / stream creation failed");
  }
  const size_t R = h->prec == TXG_FP64 ? 8 : 4;
  const size_t fbytes = (size_t)h->C * Q * h->Np * h->ssize;
  bool alloc_ok = hipMalloc(&h->f[0], fbytes) == hipSuccess && hipMalloc(&h->f[1], fbytes) == hipSuccess &&
                  hipMalloc(&h->rho, (size_t)h->C * h->Np * R) == hipSuccess &&
                  hipMalloc((void**)&h->mask, (size_t)h->Np * 4) == hipSuccess &&
                  hipMalloc((void**)&h->tag, (size_t)h->Np) == hipSuccess &&
                  hipMalloc(&h->shift, (size_t)h->C * Q * R) == hipSuccess &&
                  hipMalloc((void**)&h->red, 4096 * sizeof(double)) == hipSuccess;
  if (!alloc_ok) {
    txg_destroy(h);
    return bad(TXG_ERR_NOMEM, "device allocation failed");
  }
  (void)hipMemsetAsync(h->f[0], 0, fbytes, h->stream);
  (void)hipMemsetAsync(h->f[1], 0, fbytes, h->stream);
  (void)hipMemsetAsync(h->mask, 0, (size_t)h->Np * 4, h->stream);
  (void)hipMemsetAsync(h->tag, 0, (size_t)h->Np, h->stream);
  {
    // storage shifts w_n * rho0_m (zero unless shifted FP16)
    std::vector<double> wq(Q);
    for (int n = 0; n < Q; n++) wq[n] = h->D == 3 ? Lat<3>::w[n] : Lat<2>::w[n];
    if (h->prec == TXG_FP64) {
      std::vector<double> s((size_t)h->C * Q, 0.0);
      (void)hipMemcpy(h->shift, s.data(), s.size() * 8, hipMemcpyHostToDevice);
    } else {
      std::vector<float> s((size_t)h->C * Q, 0.0f);
      for (int m = 0; m < h->C; m++)
        for (int n = 0; n < Q; n++) s[m * Q + n] = (float)wq[n] * h->p32.shift0[m];
      (void)hipMemcpy(h->shift, s.data(), s.size() * 4, hipMemcpyHostToDevice);
    }
  }
  if (hipStreamSynchronize(h->stream) != hipSuccess) {
    txg_destroy(h);
    return bad(TXG_ERR_HIP, "initialisation failed");
  }
  *out = h;
  return TXG_OK;
}

int txg_set_walls(txg_handle* h, const double* walls) {
  if (!h || !walls) return TXG_ERR_ARG;
  const txg_desc& d = h->d;
  const int nx = d.nx, ny = d.ny, nz = h->D == 3 ? d.nz : 1, N = h->N, Q = h->Q, D = h->D;
  const int (*cv)[3] = D == 3 ? kC3 : kC2;
  const int reach = d.isotropy_order == 8 ? 2 : (d.isotropy_order == 10 ? 3 : 1);
  for (int i = 0; i < N; i++) {
    double t = walls[i];
    if (!(t >= 0.0) || std::isinf(t)) return h->fail(TXG_ERR_ARG, "wall tag %g at node %d is not >= 0", t, i);
    if (t == 900.0 || t == 901.0 || t == 902.0)
      return h->fail(TXG_ERR_UNSUPPORTED, "specular wall tags 900-902 are not implemented on the GPU");
  }
  // tag value of the cell at (x,y,z), 999 outside a non-periodic face
  // (WallsSetGhostNodes + WallsCommunicate, lbm_walls.F90:185-244)
  auto cell = [&](int x, int y, int z, long* idx) -> double {
    int c[3] = {x, y, z}, n3[3] = {nx, ny, nz};
    for (int k = 0; k < D; k++) {
      if (c[k] < 0 || c[k] >= n3[k]) {
        if (!d.periodic[k]) return 999.0;
        c[k] = ((c[k] % n3[k]) + n3[k]) % n3[k];
      }
    }
    long i = c[0] + (long)nx * (c[1] + (long)ny * c[2]);
    if (idx) *idx = i;
    return walls[i];
  };
  h->hmask.assign(h->Np, 0u);
  h->htag.assign(h->Np, 0);
  h->bb_node.clear();
  h->bb_dir.clear();
  h->fs_node.clear();
  h->fs_dir.clear();
  h->fs_min.clear();
  for (int i = 0; i < N; i++) {
    double t = walls[i];
    if (t != 0.0) {
      int mineral = (int)t;
      h->htag[i] = (t > 0.0 && t < 998.0 && mineral >= 1 && mineral <= d.nminerals) ? (uint8_t)mineral : 0;
    }
  }
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        long node = x + (long)nx * (y + (long)ny * z);
        uint32_t mk = 0;
        if (walls[node] != 0.0) {
          h->hmask[node] = kSelfSolid;
          continue;
        }
        for (int n = 1; n < Q; n++) {
          // bounce-back link: the pull source x - c_n is a wall
          double s = cell(x - cv[n][0], y - cv[n][1], z - cv[n][2], nullptr);
          if (s > 0.0) {
            mk |= 1u << n;
            h->bb_node.push_back((int32_t)node);
            h->bb_dir.push_back(n);
          }
          // wettability link: the neighbour x + c_n is a mineral wall
          double tn = cell(x + cv[n][0], y + cv[n][1], z + cv[n][2], nullptr);
          int mineral = (int)tn;
          if (tn > 0.0 && tn < 998.0 && mineral >= 1 && mineral <= d.nminerals) {
            mk |= kFSNear;
            h->fs_node.push_back((int32_t)node);
            h->fs_dir.push_back(n);
            h->fs_min.push_back(mineral);
          }
        }
        bool clean = true;
        for (int dz = (D == 3 ? -reach : 0); dz <= (D == 3 ? reach : 0) && clean; dz++)
          for (int dy = -reach; dy <= reach && clean; dy++)
            for (int dx = -reach; dx <= reach && clean; dx++)
              if (cell(x + dx, y + dy, z + dz, nullptr) != 0.0) clean = false;
        if (clean) mk |= kFFClean;
        const int pos[3] = {x, y, z}, hi[3] = {nx - 1, ny - 1, nz - 1};
        for (int f = 0; f < 2 * D; f++) {
          if (!(h->p64.dirichlet & (1 << f))) continue;
          int dim = f / 2;
          if (pos[dim] == ((f & 1) ? hi[dim] : 0)) mk |= 1u << (kFaceShift + f);
        }
        h->hmask[node] = mk;
      }
  int rc = h->hip(hipMemcpyAsync(h->mask, h->hmask.data(), (size_t)h->Np * 4, hipMemcpyHostToDevice, h->stream),
                  "mask upload");
  if (rc) return rc;
  rc = h->hip(hipMemcpyAsync(h->tag, h->htag.data(), (size_t)h->Np, hipMemcpyHostToDevice, h->stream), "tag upload");
  if (rc) return rc;
  rc = h->hip(hipStreamSynchronize(h->stream), "walls sync");
  if (rc) return rc;
  h->walls_set = 1;
  h->initialized = 0;
  return TXG_OK;
}

int txg_set_state(txg_handle* h, const double* rho, const double* u) {
  if (!h || !rho) return TXG_ERR_ARG;
  if (!h->walls_set) return h->fail(TXG_ERR_STATE, "txg_set_walls must come first");
  const size_t R = h->prec == TXG_FP64 ? 8 : 4;
  double* tmp = nullptr;
  size_t nr = (size_t)h->N * h->C, nu = (size_t)h->N * h->C * h->D;
  if (hipMalloc(&tmp, std::max(nr, nu) * 8) != hipSuccess) return h->fail(TXG_ERR_NOMEM, "staging allocation failed");
  int rc = h->hip(hipMemcpy(tmp, rho, nr * 8, hipMemcpyHostToDevice), "rho upload");
  if (!rc) {
    if (R == 8)
      hipLaunchKernelGGL(k_fill_rho<double>, dim3(cdiv(h->N, 256)), dim3(256), 0, h->stream, (double*)h->rho, tmp,
                         h->mask, h->N, h->Np, h->C);
    else
      hipLaunchKernelGGL(k_fill_rho<float>, dim3(cdiv(h->N, 256)), dim3(256), 0, h->stream, (float*)h->rho, tmp,
                         h->mask, h->N, h->Np, h->C);
    rc = h->hip(hipStreamSynchronize(h->stream), "rho fill");
  }
  if (!rc) {
    if (!h->u_ic && hipMalloc(&h->u_ic, nu * R) != hipSuccess) rc = h->fail(TXG_ERR_NOMEM, "u allocation failed");
  }
  if (!rc) {
    std::vector<double> uz(nu, 0.0);
    const double* us = u ? u : uz.data();
    if (R == 8) {
      rc = h->hip(hipMemcpy(h->u_ic, us, nu * 8, hipMemcpyHostToDevice), "u upload");
    } else {
      std::vector<float> uf(nu);
      for (size_t i = 0; i < nu; i++) uf[i] = (float)us[i];
      rc = h->hip(hipMemcpy(h->u_ic, uf.data(), nu * 4, hipMemcpyHostToDevice), "u upload");
    }
  }
  (void)hipFree(tmp);
  if (rc) return rc;
  h->state_set = 1;
  h->fi_set = 0;
  h->initialized = 0;
  return TXG_OK;
}

int txg_set_fi(txg_handle* h, const double* fi) {
  if (!h || !fi) return TXG_ERR_ARG;
  if (!h->walls_set) return h->fail(TXG_ERR_STATE, "txg_set_walls must come first");
  size_t n = (size_t)h->N * h->C * h->Q;
  double* tmp = nullptr;
  if (hipMalloc(&tmp, n * 8) != hipSuccess) return h->fail(TXG_ERR_NOMEM, "staging allocation failed");
  int rc = h->hip(hipMemcpy(tmp, fi, n * 8, hipMemcpyHostToDevice), "fi upload");
  if (!rc) {
    dim3 g(cdiv(h->N, 256));
    if (h->prec == TXG_FP64)
      hipLaunchKernelGGL((k_load_fi<double, double>), g, dim3(256), 0, h->stream, (double*)h->f[0], tmp, h->N, h->Np,
                         h->C, h->Q, (const double*)h->shift);
    else if (h->prec == TXG_FP32)
      hipLaunchKernelGGL((k_load_fi<float, float>), g, dim3(256), 0, h->stream, (float*)h->f[0], tmp, h->N, h->Np,
                         h->C, h->Q, (const float*)h->shift);
    else
      hipLaunchKernelGGL((k_load_fi<__half, float>), g, dim3(256), 0, h->stream, (__half*)h->f[0], tmp, h->N, h->Np,
                         h->C, h->Q, (const float*)h->shift);
    rc = h->hip(hipStreamSynchronize(h->stream), "fi load");
  }
  (void)hipFree(tmp);
  if (rc) return rc;
  // solid nodes keep NaN in the stencil density
  if (h->prec == TXG_FP64)
    hipLaunchKernelGGL(k_fill_rho<double>, dim3(cdiv(h->N, 256)), dim3(256), 0, h->stream, (double*)h->rho,
                       (const double*)nullptr, h->mask, h->N, h->Np, h->C);
  else
    hipLaunchKernelGGL(k_fill_rho<float>, dim3(cdiv(h->N, 256)), dim3(256), 0, h->stream, (float*)h->rho,
                       (const double*)nullptr, h->mask, h->N, h->Np, h->C);
  h->cur = 0;
  h->fi_set = 1;
  h->state_set = 1;
  h->initialized = 0;
  return h->hip(hipStreamSynchronize(h->stream), "rho reset");
}

int txg_init(txg_handle* h) {
  if (!h) return TXG_ERR_ARG;
  if (!h->walls_set || !h->state_set) return h->fail(TXG_ERR_STATE, "set walls and state before txg_init");
  int rc;
  if (!h->fi_set) {
    // FlowFiInit: f_0 into buffer 0 (the kernel writes b.fout = f[1 - in])
    AnyBufs b = bufs(h, 1);
    rc = h->hip(h->ks.fiinit(h->params(), b.ptr, h->u_ic, h->N, h->stream), "fiinit kernel");
    if (rc) return rc;
    h->cur = 0;
  }
  // FlowUpdateMoments: rho = sum f (no boundary conditions at start-up)
  AnyBufs b = bufs(h, h->cur);
  rc = h->hip(h->ks.density(h->params(), b.ptr, 0, 0, h->N, h->stream), "density kernel");
  if (rc) return rc;
  h->local_state = 1;
  h->rho_valid = 1;
  h->initialized = 1;
  h->have_old = 0;
  return h->hip(hipStreamSynchronize(h->stream), "init sync");
}

namespace {
// record an event pair around one launch of pass `which` when profiling
struct ProfScope {
  txg_handle* h;
  int which;
  int a = -1;
  ProfScope(txg_handle* hh, int w) : h(hh), which(w) {
    if (!h->prof) return;
    a = take();
    (void)hipEventRecord(h->ev_pool[a], h->stream);
  }
  ~ProfScope() {
    if (a < 0) return;
    int b = take();
    (void)hipEventRecord(h->ev_pool[b], h->stream);
    h->ev_pairs[which].push_back({a, b});
  }
  int take() {
    if (h->ev_next >= (int)h->ev_pool.size()) {
      hipEvent_t e;
      (void)hipEventCreate(&e);
      h->ev_pool.push_back(e);
    }
    return h->ev_next++;
  }
};
}  // namespace

int txg_step(txg_handle* h, int32_t nsteps) {
  if (!h || nsteps < 0) return TXG_ERR_ARG;
  if (!h->initialized) return h->fail(TXG_ERR_STATE, "txg_init has not been called");
  const bool fused = use_fused(h);
  const FusedGeom g = fused ? fused_geom(h) : FusedGeom{};
  for (int s = 0; s < nsteps; s++) {
    int rc;
    if (fused && !h->local_state) {
      // one pass: pull, density (LDS), forces, collide
      AnyBufs b = bufs(h, h->cur);
      {
        ProfScope ps(h, 0);
        rc = h->hip(h->ks.fused[h->tile](h->params(), b.ptr, &g, h->stream), "one-pass kernel");
      }
      if (rc) return rc;
      h->cur = 1 - h->cur;
      h->rho_valid = 0;
      continue;
    }
    // two passes: collide (from the local initial state, or pulled), then
    // the density of the new state for the next collide
    AnyBufs b = bufs(h, h->cur);
    int pulled = h->local_state ? 0 : 1;
    {
      ProfScope ps(h, 0);
      rc = h->hip(h->ks.collide(h->params(), b.ptr, pulled, pulled, nullptr, h->N, h->stream), "collide kernel");
    }
    if (rc) return rc;
    h->cur = 1 - h->cur;
    h->local_state = 0;
    h->rho_valid = 0;
    if (fused) continue;
    AnyBufs b2 = bufs(h, h->cur);
    {
      ProfScope ps(h, 1);
      rc = h->hip(h->ks.density(h->params(), b2.ptr, 1, 1, h->N, h->stream), "density kernel");
    }
    if (rc) return rc;
    h->rho_valid = 1;
  }
  return TXG_OK;
}

int txg_tune(txg_handle* h, int32_t kernel, int32_t tile_x, int32_t tile_y, int32_t zseg) {
  if (!h || kernel < 0 || kernel > 2 || zseg < 0) return TXG_ERR_ARG;
  int tile = h->tile;
  if (tile_x != 0 || tile_y != 0) {
    tile = -1;
    for (int i = 0; i < kNumTiles; i++)
      if (kTileX[i] == tile_x && kTileY[i] == tile_y) tile = i;
    if (tile < 0) return h->fail(TXG_ERR_UNSUPPORTED, "tile %dx%d is not instantiated", tile_x, tile_y);
  }
  if (kernel == 2 && !h->ks.fused[tile])
    return h->fail(TXG_ERR_UNSUPPORTED, "no one-pass kernel for this lattice");
  if (kernel == 2 && (uint64_t)h->Q * h->Np * h->ssize >= (1ull << 32))
    return h->fail(TXG_ERR_UNSUPPORTED, "one-pass kernel needs < 4 GiB of populations per component");
  int rc = h->hip(hipStreamSynchronize(h->stream), "tune sync");
  if (rc) return rc;
  h->kernel = kernel;
  h->tile = tile;
  h->zseg = zseg;
  return TXG_OK;
}

int txg_profile_enable(txg_handle* h, int32_t on) {
  if (!h) return TXG_ERR_ARG;
  int rc = h->hip(hipStreamSynchronize(h->stream), "profile sync");
  if (rc) return rc;
  h->ev_pairs[0].clear();
  h->ev_pairs[1].clear();
  h->ev_next = 0;
  h->prof = on ? 1 : 0;
  return TXG_OK;
}

int txg_profile_read(txg_handle* h, double* ms, int32_t* n) {
  if (!h || !ms || !n) return TXG_ERR_ARG;
  int rc = h->hip(hipStreamSynchronize(h->stream), "profile sync");
  if (rc) return rc;
  for (int w = 0; w < 2; w++) {
    double tot = 0.0;
    for (auto& pr : h->ev_pairs[w]) {
      float t = 0.f;
      rc = h->hip(hipEventElapsedTime(&t, h->ev_pool[pr.first], h->ev_pool[pr.second]), "event time");
      if (rc) return rc;
      tot += t;
    }
    ms[w] = tot;
    n[w] = (int32_t)h->ev_pairs[w].size();
  }
  return TXG_OK;
}

int txg_sync(txg_handle* h) {
  if (!h) return TXG_ERR_ARG;
  return h->hip(hipStreamSynchronize(h->stream), "stream sync");
}

int txg_get_fi(txg_handle* h, double* fi) {
  if (!h || !fi) return TXG_ERR_ARG;
  int rc = run_export(h);
  if (rc) return rc;
  const size_t N = h->N, C = h->C, Q = h->Q;
  std::vector<double> tmp(N * C * Q);
  rc = h->hip(hipMemcpy(tmp.data(), export_ptrs(h).f, tmp.size() * 8, hipMemcpyDeviceToHost), "fi download");
  if (rc) return rc;
  for (size_t m = 0; m < C; m++)
    for (size_t n = 0; n < Q; n++) {
      const double* src = tmp.data() + (m * Q + n) * N;
      for (size_t i = 0; i < N; i++) fi[(i * Q + n) * C + m] = h->hmask[i] & kSelfSolid ? 0.0 : src[i];
    }
  return TXG_OK;
}

int txg_get_moments(txg_handle* h, double* rho, double* flux, double* forces) {
  if (!h) return TXG_ERR_ARG;
  int rc = run_export(h);
  if (rc) return rc;
  const size_t N = h->N, C = h->C, D = h->D;
  ExportPtrs e = export_ptrs(h);
  std::vector<double> r(N * C), up(N * D), F(N * C * D);
  rc = h->hip(hipMemcpy(r.data(), e.rho, r.size() * 8, hipMemcpyDeviceToHost), "rho download");
  if (!rc) rc = h->hip(hipMemcpy(up.data(), e.up, up.size() * 8, hipMemcpyDeviceToHost), "u download");
  if (!rc) rc = h->hip(hipMemcpy(F.data(), e.F, F.size() * 8, hipMemcpyDeviceToHost), "F download");
  if (rc) return rc;
  for (size_t i = 0; i < N; i++) {
    const bool solid = h->hmask[i] & kSelfSolid;
    for (size_t m = 0; m < C; m++) {
      if (rho) rho[i * C + m] = solid ? 0.0 : r[m * N + i];
      for (size_t d = 0; d < D; d++) {
        if (flux) flux[(i * D + d) * C + m] = solid ? 0.0 : up[d * N + i];
        if (forces) forces[(i * D + d) * C + m] = solid ? 0.0 : F[(m * D + d) * N + i];
      }
    }
  }
  return TXG_OK;
}

int txg_diagnostics(txg_handle* h, double* rhot, double* prs, double* velt) {
  if (!h) return TXG_ERR_ARG;
  int rc = run_export(h);
  if (rc) return rc;
  const size_t N = h->N, C = h->C, D = h->D;
  ExportPtrs e = export_ptrs(h);
  std::vector<double> r(N * C), F(N * C * D), j(N * C * D);
  rc = h->hip(hipMemcpy(r.data(), e.rho, r.size() * 8, hipMemcpyDeviceToHost), "rho download");
  if (!rc) rc = h->hip(hipMemcpy(F.data(), e.F, F.size() * 8, hipMemcpyDeviceToHost), "F download");
  if (!rc) rc = h->hip(hipMemcpy(j.data(), e.j, j.size() * 8, hipMemcpyDeviceToHost), "j download");
  if (rc) return rc;
  const txg_desc& d = h->d;
  // FlowUpdateDiagnosticsD3/D2 (lbm_flow.F90:654-758), psi = rho
  for (size_t i = 0; i < N; i++) {
    if (h->hmask[i] & kSelfSolid) {
      if (rhot) rhot[i] = 0.0;
      if (prs) prs[i] = 0.0;
      if (velt)
        for (size_t dd = 0; dd < D; dd++) velt[i * D + dd] = 0.0;
      continue;
    }
    double rt = 0.0;
    for (size_t m = 0; m < C; m++) rt = rt + r[m * N + i] * d.mm[m];
    double p = rt / 3.;
    if (C > 1)
      for (size_t m = 0; m < C; m++) {
        double s = 0.0;
        for (size_t mp = 0; mp < C; mp++) s = s + d.g[m][mp] * r[mp * N + i];
        p = p + 6.0 / 2. * r[m * N + i] * s;
      }
    if (rhot) rhot[i] = rt;
    if (prs) prs[i] = p;
    for (size_t dd = 0; dd < D; dd++) {
      double s = 0.0;
      for (size_t m = 0; m < C; m++) {
        double u = j[(m * D + dd) * N + i] + .5 * F[(m * D + dd) * N + i];
        s = s + u * d.mm[m];
      }
      if (velt) velt[i * D + dd] = s / rt;
    }
  }
  return TXG_OK;
}

int txg_reduce(txg_handle* h, double* out) {
  if (!h || !out) return TXG_ERR_ARG;
  int rc = run_export(h);
  if (rc) return rc;
  ExportPtrs e = export_ptrs(h);
  const int nb = 1024;
  const int C = h->C, D = h->D;
  double* part = h->red;
  double* res = h->red + nb;
  // out[m] = sum rho_m ; out[C + m*D + d] = sum (j + F/2)_{m,d}
  for (int m = 0; m < C; m++) {
    hipLaunchKernelGGL(k_partial_sum, dim3(nb), dim3(256), 0, h->stream, e.rho + (size_t)m * h->N, h->mask, h->N, part);
    hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, h->stream, part, nb, res + m);
  }
  double* tmp = e.f;  // the population export is not needed here
  for (int m = 0; m < C; m++)
    for (int d = 0; d < D; d++) {
      const size_t off = (size_t)(m * D + d) * h->N;
      hipLaunchKernelGGL(k_j_half_f, dim3(cdiv(h->N, 256)), dim3(256), 0, h->stream, tmp, e.j + off, e.F + off, h->N);
      hipLaunchKernelGGL(k_partial_sum, dim3(nb), dim3(256), 0, h->stream, tmp, h->mask, h->N, part);
      hipLaunchKernelGGL(k_final_sum, dim3(1), dim3(256), 0, h->stream, part, nb, res + C + m * D + d);
    }
  rc = h->hip(hipStreamSynchronize(h->stream), "reduce sync");
  if (rc) return rc;
  return h->hip(hipMemcpy(out, res, (size_t)(C + C * D) * 8, hipMemcpyDeviceToHost), "reduce download");
}

int txg_delta_norm(txg_handle* h, double* norm) {
  if (!h || !norm) return TXG_ERR_ARG;
  int rc = run_export(h);
  if (rc) return rc;
  const size_t n = (size_t)h->N * h->C;
  if (!h->rho_old && hipMalloc(&h->rho_old, n * 8) != hipSuccess) return h->fail(TXG_ERR_NOMEM, "rho_old allocation");
  // export rho is [m*N + node]; the reference Vec is node-major, but max is
  // order independent
  ExportPtrs e = export_ptrs(h);
  const int nb = 1024;
  hipLaunchKernelGGL(k_delta, dim3(nb), dim3(256), 0, h->stream, h->rho_old, e.rho, (int)n, h->have_old ? 0 : 1,
                     h->red);
  std::vector<double> part(nb);
  rc = h->hip(hipMemcpy(part.data(), h->red, nb * 8, hipMemcpyDeviceToHost), "delta download");
  if (rc) return rc;
  double mx = 0.0;
  for (double v : part) mx = std::max(mx, v);
  *norm = mx;
  h->have_old = 1;
  return TXG_OK;
}

int txg_last_error(const txg_handle* h, char* buf, int32_t len) {
  const std::string& s = h ? h->err : g_create_error;
  if (buf && len > 0) {
    size_t k = std::min((size_t)(len - 1), s.size());
    memcpy(buf, s.data(), k);
    buf[k] = 0;
  }
  return (int)s.size();
}

int txg_destroy(txg_handle* h) {
  if (!h) return TXG_OK;
  if (h->stream) (void)hipStreamSynchronize(h->stream);
  (void)hipFree(h->f[0]);
  (void)hipFree(h->f[1]);
  (void)hipFree(h->rho);
  (void)hipFree(h->mask);
  (void)hipFree(h->tag);
  (void)hipFree(h->u_ic);
  (void)hipFree(h->shift);
  (void)hipFree(h->ex);
  (void)hipFree(h->red);
  (void)hipFree(h->rho_old);
  for (hipEvent_t e : h->ev_pool) (void)hipEventDestroy(e);
  if (h->stream) (void)hipStreamDestroy(h->stream);
  delete h;
  return TXG_OK;
}

}  // extern "C"
