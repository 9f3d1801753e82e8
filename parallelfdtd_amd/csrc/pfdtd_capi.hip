// pfdtd_capi.hip -- C ABI of the SRL_FORWARD core (declarations and reference citations in
// include/pfdtd.h). Host code: context, device memory, coefficient tables, the step loop.
#include "../../include/pfdtd.h"
#include "pfdtd_internal.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

using namespace pfdtd;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_last_error = buf;
  return code;
}

#define HIP_TRY(expr)                                                                        \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      return fail(PFDTD_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),        \
                  __FILE__, __LINE__);                                                       \
  } while (0)

constexpr int kCoefNum = 20;   // MATERIAL_COEF_NUM, src/base/MaterialHandler.h:29
constexpr int kProgressMod = 100;  // PROGRESS_MOD, src/kernels/kernels3d.h:36

size_t elem_size(int precision) {
  switch (precision) {
  case PFDTD_F32:
    return 4;
  case PFDTD_F64:
    return 8;
  case PFDTD_F16:
  case PFDTD_BF16:
  case PFDTD_F16_DIFF:
    return 2;
  default:
    return 0;
  }
}

// Storage format of the current field P (the difference form keeps P in binary16).
int p_storage(int precision) { return precision == PFDTD_F16_DIFF ? PFDTD_F16 : precision; }

template <typename T>
void dfree(T *&p) {
  if (p) (void)hipFree((void *)p);
  p = nullptr;
}

}  // namespace

struct pfdtd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;

  // domain
  int64_t X = 0, Y = 0, Z = 0, N = 0;
  uint8_t *d_pos = nullptr, *d_mat = nullptr;
  uint16_t *d_pm = nullptr;  // pos | mat << 8, built for the two-level update (pm_ok: current)
  uint32_t *d_air_rows = nullptr;  // its air-row bits (built with it; null: not available)
  uint32_t *d_air_rows_p = nullptr;  // the same for the two-column sweep's tiles (X even)
  uint32_t *d_air_rows3 = nullptr;   // the same for the three-level sweep's tiles
  bool air3_ok = false;              // d_air_rows3 built for the current domain (or no room)
  bool pm_ok = false;
  int64_t dom_cap = 0;  // voxels the position/material buffers hold
  bool has_domain = false;
  DomainScan scan{};
  DomainScan *d_scan = nullptr;

  // materials / parameters
  bool has_materials = false;
  std::vector<float> table;  // [n_mat][20]
  int n_mat = 0;
  unsigned octave = 0;
  double lambda = 0.0;
  int flags = 0;
  CoefF *d_coef_f = nullptr;
  CoefD *d_coef_d = nullptr;
  float lam2_f = 0.f;
  double lam2_d = 0.0;
  CoefF air_f{};          // K = 6 row entry shared by all air voxels
  CoefD air_d{};
  int air_uniform = 0;    // 1 when every material gives the same K = 6 entry (both tables)
  int update_type = 0;    // enum UpdateType: 0 SRL_FORWARD, 1 SHARED, 2 SRL (centred)
  int scheme = 0;         // boundary formulation: 0 forward (types 0, 1), 2 centred (type 2)
  unsigned degenerate = 0;  // inside voxels without a centred stencil (label classification)
  int ry = 0;             // y rows per thread of the update kernel (0: automatic, plan_single)
  int zc = 0;             // slices per z-chunk of the update kernel (with ry)
  int levels = 0;         // time levels per sweep: 1, 2, or 0 = by precision
  int fzc = 0;            // slices per z-chunk of the two-level update (0: automatic)
  int fshape = PFDTD_SWEEP_AUTO;  // block shape of the two-level update

  // sources / receivers
  int n_src = 0, n_samples = 0;
  std::vector<int32_t> src_xyz, src_kind;
  int64_t *d_src_idx = nullptr;
  int src_distinct = 1;  // every in-grid source on its own element: injection may run in parallel
  int32_t *d_src_xyz = nullptr;  // grid coordinates, (-2^30,-2^30,-2^30) outside the grid
  int32_t *d_src_kind = nullptr;
  float *d_sig_f = nullptr;
  double *d_sig_d = nullptr;
  int n_rec = 0;
  std::vector<int32_t> rec_xyz;
  // batch: fields hold 2^lgB values per voxel, slot innermost, or (grouped) 2^lgB whole fields
  // one after another; per-source/receiver slot
  int lgB = 0;
  int grouped = 0;
  std::vector<int32_t> src_slot, rec_slot;
  int32_t *d_src_slot = nullptr, *d_rec_slot = nullptr;
  int64_t *d_rec_idx = nullptr;
  int32_t *d_rec_xyz = nullptr;

  // fields
  int precision = -1;
  int scale_exp = 0;      // scale of the allocated fields
  int scale_req = 0;      // requested F16/BF16 storage scale (pfdtd_prepare_fields)
  void *d_P = nullptr, *d_Pp = nullptr;    // current and past field
  void *d_P2 = nullptr, *d_Pp2 = nullptr;  // spare pair written by the two-level update
  size_t field_bytes = 0;

  // responses
  double *d_resp = nullptr;
  int resp_steps = 0;
  size_t resp_cap = 0;

  // step-by-step state
  int last_direction = 1;

  // options / timing
  int poll_interval = kProgressMod;
  bool kernel_timing = false;
  double upd_ms = 0.0;
  int64_t upd_launches = 0;
  std::vector<hipEvent_t> events;
  hipEvent_t ev_begin = nullptr, ev_end = nullptr;
};

namespace pfdtd {

void record_error(const char *msg) { g_last_error = msg; }

// kernels3d.cu:513 (literal `mat*20*+params[3]`) vs kernels3d.cu:639 (intended)
int64_t material_slot(unsigned m, unsigned octave, bool literal) {
  if (literal) {
    const unsigned base = m * 20u;
    return (int64_t)(unsigned)((float)base * (float)octave);
  }
  return (int64_t)m * kCoefNum + (int64_t)octave;
}

// The per-(K, m) coefficients, evaluated in the operation order of the reference update body
// (kernels3d.cu:513-526) so that a[K,m] multiplies exactly like the per-node expression.
void build_coef_tables(const float *table20, int n_mat, unsigned octave, double lambda,
                       bool literal, CoefF *cf, CoefD *cd) {
#pragma clang fp contract(off)
  const float lam_f = (float)lambda;                 // params[0]
  const float lam2_f = (float)(lambda * lambda);     // params[1]
  const double lam_d = lambda;
  const double lam2_d = lambda * lambda;
  for (int i = 0; i < kCoefEntries; ++i) {
    cf[i] = CoefF{0.f, 0.f, 0.f, 0.f};
    cd[i] = CoefD{0.0, 0.0, 0.0, 0.0};
  }
  for (int K = 0; K <= 6; ++K) {
    for (int m = 0; m < 256; ++m) {
      const int64_t slot = material_slot((unsigned)m, octave, literal);
      const float bm = slot < (int64_t)n_mat * kCoefNum ? table20[(size_t)slot] : 0.f;
      const float position = (float)K;
      const float beta = (float)(0.5 * (double)(bm * (6.f - position) * lam_f));
      CoefF &f = cf[(size_t)K * 256 + m];
      f.inv = 1.f / (1.f + beta);
      f.omb = 1.f - beta;
      f.tk = 2.f - position * lam2_f;
      const double bmd = (double)bm;
      const double positiond = (double)K;
      const double betad = 0.5 * (bmd * (6.f - positiond) * lam_d);
      CoefD &d = cd[(size_t)K * 256 + m];
      d.inv = 1.f / (1.f + betad);
      d.omb = 1.f - betad;
      d.tk = 2.f - positiond * lam2_d;
    }
  }
}

// Centred-boundary coefficients, in the operation order of fdtd3dStdKowalczykMaterials
// (kernels3d.cu:626-664): beta = table[m*20+octave] * lambda * D (D = number of DIR bits),
// entries {1/(1+beta), beta-1, 2-6 lambda^2}.
void build_centred_tables(const float *table20, int n_mat, unsigned octave, double lambda,
                          CoefF *cf, CoefD *cd) {
#pragma clang fp contract(off)
  const float lam_f = (float)lambda;
  const float lam2_f = (float)(lambda * lambda);
  const double lam_d = lambda;
  const double lam2_d = lambda * lambda;
  const float c_f = 2 - 6 * lam2_f;
  const double c_d = 2 - 6 * lam2_d;
  for (int i = 0; i < kCoefEntries; ++i) {
    cf[i] = CoefF{0.f, 0.f, 0.f, 0.f};
    cd[i] = CoefD{0.0, 0.0, 0.0, 0.0};
  }
  for (int D = 0; D <= 3; ++D) {
    for (int m = 0; m < 256; ++m) {
      const int64_t slot = material_slot((unsigned)m, octave, false);
      const float bm = slot < (int64_t)n_mat * kCoefNum ? table20[(size_t)slot] : 0.f;
      const float beta = bm * lam_f * (float)D;
      CoefF &f = cf[(size_t)D * 256 + m];
      f.inv = 1.f / (1.f + beta);
      f.omb = beta - 1.f;
      f.tk = c_f;
      const double betad = (double)bm * lam_d * (double)D;
      CoefD &d = cd[(size_t)D * 256 + m];
      d.inv = 1.f / (1.f + betad);
      d.omb = betad - 1.f;
      d.tk = c_d;
    }
  }
}

}  // namespace pfdtd

// Staging planes of pfdtd_capture_slice_noalloc: FP32 values (PFDTD_STATIC_CAPTURE_MAX * 4
// bytes) followed by the position bytes. Each GPU holds its own copy at its own address; host
// code reaches it only through the address hipGetSymbolAddress returns on that device.
static __device__ __align__(256) unsigned char scratch[(size_t)PFDTD_STATIC_CAPTURE_MAX * 5];

namespace {

int check_ctx(pfdtd_ctx *ctx) {
  if (!ctx) return fail(PFDTD_E_ARG, "null context");
  hipError_t e = hipSetDevice(ctx->device);
  if (e != hipSuccess)
    return fail(PFDTD_E_HIP, "hipSetDevice(%d): %s", ctx->device, hipGetErrorString(e));
  return PFDTD_OK;
}

std::mutex g_scratch_mu;
std::vector<void *> g_scratch_ptr;  // device address of `scratch`, per device (null: unresolved)

// Resolves `scratch` on `device`, which must be current; once per device (pfdtd_ctx_create).
int resolve_scratch(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize(device + 1, nullptr);
  if (g_scratch_ptr[device]) return PFDTD_OK;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p)
    return fail(PFDTD_E_HIP, "static capture buffer on device %d: %s", device,
                hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return PFDTD_OK;
}

void *scratch_on(int device) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return device < (int)g_scratch_ptr.size() ? g_scratch_ptr[device] : nullptr;
}

int64_t lin_idx(const pfdtd_ctx *c, const int32_t *q) {
  if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= c->X || q[1] >= c->Y || q[2] >= c->Z)
    return -1;
  return ((int64_t)q[2] * c->Y + q[1]) * c->X + q[0];
}

// Slots one two-level launch advances (grouped batch), 1 otherwise.
int batch_slots(const pfdtd_ctx *c) { return c->grouped ? 1 << c->lgB : 1; }

// Element index of a source/receiver in the (possibly batched) field, -1 outside the grid.
int64_t elem_idx(const pfdtd_ctx *c, const int32_t *q, int slot) {
  const int64_t v = lin_idx(c, q);
  if (v < 0) return -1;
  return c->grouped ? (int64_t)slot * c->N + v : (v << c->lgB) + slot;
}

// 1 when no two in-grid sources share a field element (injection order then cannot matter).
int distinct_indices(const std::vector<int64_t> &idx) {
  std::vector<int64_t> v;
  v.reserve(idx.size());
  for (int64_t i : idx)
    if (i >= 0) v.push_back(i);
  std::sort(v.begin(), v.end());
  return std::adjacent_find(v.begin(), v.end()) == v.end() ? 1 : 0;
}

constexpr int32_t kOutside = -(1 << 30);

// Device copy of grid coordinates with entries outside the grid replaced by kOutside.
int upload_xyz(const pfdtd_ctx *c, const std::vector<int32_t> &xyz, int32_t *dst) {
  const size_t n = xyz.size() / 3;
  if (!n) return PFDTD_OK;
  std::vector<int32_t> v(xyz);
  for (size_t i = 0; i < n; ++i)
    if (!c->has_domain || lin_idx(c, &v[3 * i]) < 0) v[3 * i] = v[3 * i + 1] = v[3 * i + 2] = kOutside;
  hipError_t e = hipMemcpy(dst, v.data(), sizeof(int32_t) * 3 * n, hipMemcpyHostToDevice);
  if (e != hipSuccess) return fail(PFDTD_E_HIP, "coordinate upload: %s", hipGetErrorString(e));
  return PFDTD_OK;
}

void build_coefficients(const pfdtd_ctx *c, std::vector<CoefF> &cf, std::vector<CoefD> &cd) {
  cf.assign(kCoefEntries, CoefF{0.f, 0.f, 0.f, 0.f});
  cd.assign(kCoefEntries, CoefD{0.0, 0.0, 0.0, 0.0});
  if (c->scheme == 2)
    build_centred_tables(c->table.data(), c->n_mat, c->octave, c->lambda, cf.data(), cd.data());
  else
    build_coef_tables(c->table.data(), c->n_mat, c->octave, c->lambda,
                      (c->flags & PFDTD_MATSLOT_LITERAL) != 0, cf.data(), cd.data());
}

int scan_domain(pfdtd_ctx *c) {
  HIP_TRY(hipMemsetAsync(c->d_scan, 0, sizeof(DomainScan), c->stream));
  HIP_TRY(launch_scan_domain(c->d_pos, c->d_mat, (int)c->X, (int)c->Y, (int)c->Z, c->scheme,
                             c->d_scan, c->stream));
  HIP_TRY(hipMemcpyAsync(&c->scan, c->d_scan, sizeof(DomainScan), hipMemcpyDeviceToHost,
                         c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return PFDTD_OK;
}

// Time levels per sweep: the explicit setting, else three for FP32 and FP64 (unbatched or in
// the grouped batch layout) and two otherwise. Same buffers, alternating settings: on the
// 2048x1536x768 room three levels take FP32 from 4.37 to 3.85 ms per step and FP64 from 8.61
// to 7.4-7.7, on 1024x768x384 FP32 from 0.71 to 0.55 and FP64 from 1.19 to 1.06; 16-bit
// storage stays on the two-column two-level sweep (2.67 ms against 3.46 for three levels of one
// column). Two levels were faster than single steps for every precision (FP32 7.35 vs 7.74 ms
// per step on the large room).
int levels_for(const pfdtd_ctx *c, int precision) {
  if (c->levels) return c->levels;
  if ((precision == PFDTD_F32 || precision == PFDTD_F64) && (c->lgB == 0 || c->grouped))
    return 3;
  return 2;
}

void free_fields(pfdtd_ctx *c) {
  dfree(c->d_P);
  dfree(c->d_Pp);
  dfree(c->d_P2);
  dfree(c->d_Pp2);
  c->field_bytes = 0;
  c->precision = -1;
}

int alloc_domain(pfdtd_ctx *c, int64_t X, int64_t Y, int64_t Z) {
  if (X < 3 || Y < 3 || Z < 3)
    return fail(PFDTD_E_ARG, "grid %lldx%lldx%lld: every dimension must be >= 3",
                (long long)X, (long long)Y, (long long)Z);
  if (X > (1 << 30) || Y > 65535 || Z > 65535)
    return fail(PFDTD_E_ARG, "grid %lldx%lldx%lld exceeds X<=2^30, Y<=65535, Z<=65535",
                (long long)X, (long long)Y, (long long)Z);
  const int64_t N = X * Y * Z;
  if (c->d_pos && c->dom_cap != N) {
    dfree(c->d_pos);
    dfree(c->d_mat);
    dfree(c->d_pm);
    c->dom_cap = 0;
  }
  c->pm_ok = false;
  c->air3_ok = false;
  dfree(c->d_air_rows);  // sized by the grid's shape, not only its voxel count
  dfree(c->d_air_rows_p);
  dfree(c->d_air_rows3);
  // fields sized for the previous grid are dropped
  free_fields(c);
  if (!c->d_pos) {
    hipError_t e1 = hipMalloc((void **)&c->d_pos, (size_t)N);
    hipError_t e2 = e1 == hipSuccess ? hipMalloc((void **)&c->d_mat, (size_t)N) : e1;
    if (e1 != hipSuccess || e2 != hipSuccess) {
      dfree(c->d_pos);
      dfree(c->d_mat);
      (void)hipGetLastError();
      return fail(PFDTD_E_NOMEM, "cannot allocate %lld-voxel domain bytes", (long long)N);
    }
    c->dom_cap = N;
  }
  c->X = X;
  c->Y = Y;
  c->Z = Z;
  c->N = N;
  c->has_domain = false;
  return PFDTD_OK;
}

int finish_domain(pfdtd_ctx *c) {
  int rc = scan_domain(c);
  if (rc) return rc;
  c->has_domain = true;
  // source/receiver linear indices depend on the grid
  if (c->n_src) {
    std::vector<int64_t> idx(c->n_src);
    for (int s = 0; s < c->n_src; ++s) idx[s] = elem_idx(c, &c->src_xyz[3 * s], c->src_slot[s]);
    c->src_distinct = distinct_indices(idx);
    HIP_TRY(hipMemcpy(c->d_src_idx, idx.data(), sizeof(int64_t) * c->n_src,
                      hipMemcpyHostToDevice));
    rc = upload_xyz(c, c->src_xyz, c->d_src_xyz);
    if (rc) return rc;
  }
  if (c->n_rec) {
    std::vector<int64_t> idx(c->n_rec);
    for (int r = 0; r < c->n_rec; ++r) idx[r] = elem_idx(c, &c->rec_xyz[3 * r], c->rec_slot[r]);
    HIP_TRY(hipMemcpy(c->d_rec_idx, idx.data(), sizeof(int64_t) * c->n_rec,
                      hipMemcpyHostToDevice));
    rc = upload_xyz(c, c->rec_xyz, c->d_rec_xyz);
    if (rc) return rc;
  }
  return PFDTD_OK;
}

// Checks everything the update relies on before the first launch of a run.
int check_runnable(pfdtd_ctx *c) {
  if (!c->has_domain) return fail(PFDTD_E_STATE, "no domain set");
  if (!c->has_materials) return fail(PFDTD_E_STATE, "no materials set");
  if (c->scan.shell_inside)
    return fail(PFDTD_E_DOMAIN,
                "%u voxels of the outer layer are inside; the update requires an outside shell",
                c->scan.shell_inside);
  if (c->scan.bad_k)
    return fail(PFDTD_E_DOMAIN, "%u inside voxels have K > 6 in their position byte",
                c->scan.bad_k);
  if (c->scheme == 2 && c->degenerate)
    return fail(PFDTD_E_DOMAIN,
                "%u inside voxels have no air neighbour on some axis (no centred stencil)",
                c->degenerate);
  if (c->scheme == 2 && c->lgB)
    return fail(PFDTD_E_STATE, "source batching runs the forward update types only");
  const bool literal = c->scheme == 0 && (c->flags & PFDTD_MATSLOT_LITERAL) != 0;
  for (int m = 0; m < 256; ++m) {
    if (!((c->scan.mat_present[m >> 5] >> (m & 31)) & 1u)) continue;
    if (material_slot((unsigned)m, c->octave, literal) >= (int64_t)c->n_mat * kCoefNum)
      return fail(PFDTD_E_DOMAIN,
                  "material byte %d addresses coefficient slot %lld outside the %d-material "
                  "table",
                  m, (long long)material_slot((unsigned)m, c->octave, literal), c->n_mat);
  }
  return PFDTD_OK;
}

int ensure_fields(pfdtd_ctx *c, int precision, int scale_exp) {
  const size_t es = elem_size(precision);
  if (!es) return fail(PFDTD_E_ARG, "unknown precision %d", precision);
  if ((precision == PFDTD_F32 || precision == PFDTD_F64) && scale_exp != 0)
    return fail(PFDTD_E_ARG, "scale_exp applies to F16/BF16 storage only");
  if (scale_exp < -60 || scale_exp > 60) return fail(PFDTD_E_ARG, "scale_exp out of range");
  const size_t bytes = es * ((size_t)c->N << c->lgB);
  const bool want4 = (c->lgB == 0 || c->grouped) && levels_for(c, precision) >= 2;
  if (c->d_P && c->field_bytes == bytes && (!want4 || c->d_P2)) {
    c->precision = precision;
    c->scale_exp = scale_exp;
    return PFDTD_OK;
  }
  free_fields(c);
  hipError_t e = hipMalloc(&c->d_P, bytes);
  if (e == hipSuccess) e = hipMalloc(&c->d_Pp, bytes);
  if (e == hipSuccess && want4) e = hipMalloc(&c->d_P2, bytes);
  if (e == hipSuccess && want4) e = hipMalloc(&c->d_Pp2, bytes);
  if (e != hipSuccess) {
    free_fields(c);
    (void)hipGetLastError();
    return fail(PFDTD_E_NOMEM, "cannot allocate %d pressure fields of %zu bytes", want4 ? 4 : 2,
                bytes);
  }
  c->field_bytes = bytes;
  c->precision = precision;
  c->scale_exp = scale_exp;
  return PFDTD_OK;
}

// Zero current and past fields. Of the spare pair only the first and last z-slices are cleared:
// the two-level update writes every voxel of slices 1..Z-2 of both spare buffers before they
// are read (x/y edge voxels as 0), and never writes slices 0 and Z-1, which the next updates
// read as z-neighbours once the pairs are swapped.
int zero_fields(pfdtd_ctx *c) {
  HIP_TRY(hipMemsetAsync(c->d_P, 0, c->field_bytes, c->stream));
  HIP_TRY(hipMemsetAsync(c->d_Pp, 0, c->field_bytes, c->stream));
  // a grouped batch holds one whole field per slot, each with its own first and last slice
  const int slots = batch_slots(c);
  const size_t per_slot = c->field_bytes / (size_t)slots;
  const size_t slice = per_slot / (size_t)c->Z;
  for (void *q : {c->d_P2, c->d_Pp2}) {
    if (!q) continue;
    for (int sl = 0; sl < slots; ++sl) {
      char *base = (char *)q + (size_t)sl * per_slot;
      HIP_TRY(hipMemsetAsync(base, 0, slice, c->stream));
      HIP_TRY(hipMemsetAsync(base + (size_t)(c->Z - 1) * slice, 0, slice, c->stream));
    }
  }
  c->last_direction = 1;
  return PFDTD_OK;
}

int ensure_responses(pfdtd_ctx *c, int n_steps) {
  const size_t need = (size_t)c->n_rec * (size_t)n_steps;
  if (need > c->resp_cap) {
    dfree(c->d_resp);
    if (hipMalloc((void **)&c->d_resp, need * sizeof(double)) != hipSuccess) {
      (void)hipGetLastError();
      c->resp_cap = 0;
      return fail(PFDTD_E_NOMEM, "cannot allocate responses");
    }
    c->resp_cap = need;
  }
  c->resp_steps = n_steps;
  if (need) HIP_TRY(hipMemsetAsync(c->d_resp, 0, need * sizeof(double), c->stream));
  return PFDTD_OK;
}

InjectArgs inject_args(const pfdtd_ctx *c, int gather_step, int inject_step) {
  InjectArgs a;
  a.rec_idx = c->d_rec_idx;
  a.n_rec = c->n_rec;
  a.resp = c->d_resp;
  a.resp_steps = c->resp_steps;
  a.gather_step = (c->n_rec > 0) ? gather_step : -1;
  a.src_idx = c->d_src_idx;
  a.src_kind = c->d_src_kind;
  a.n_src = c->n_src;
  a.sig_f = c->d_sig_f;
  a.sig_d = c->d_sig_d;
  a.n_samples = c->n_samples;
  a.inject_step = (c->n_src > 0) ? inject_step : -1;
  a.soft_overwrite = (c->flags & PFDTD_SOFT_OVERWRITE) ? 1 : 0;
  a.scale_exp = c->scale_exp;
  a.distinct = c->src_distinct;
  return a;
}

// Material columns the update kernels stage in LDS: the highest material byte present in the
// domain + 1 (every present byte has been checked against the table, check_runnable).
int coef_cols(const pfdtd_ctx *c) {
  for (int m = 255; m > 0; --m)
    if ((c->scan.mat_present[m >> 5] >> (m & 31)) & 1u) return m + 1;
  return 1;
}

hipError_t launch_update_any(pfdtd_ctx *c, const UpdateGeom &g);

// Grouped batch: one single-level update of every slot (each slot's fields are one domain).
hipError_t launch_update_grouped(pfdtd_ctx *c, const UpdateGeom &g) {
  const size_t slot_bytes = c->field_bytes >> c->lgB;
  void *P = c->d_P, *Pp = c->d_Pp;
  const int lg = c->lgB;
  hipError_t e = hipSuccess;
  c->lgB = 0;
  for (int sl = 0; sl < (1 << lg) && e == hipSuccess; ++sl) {
    c->d_P = (char *)P + (size_t)sl * slot_bytes;
    c->d_Pp = (char *)Pp + (size_t)sl * slot_bytes;
    e = launch_update_any(c, g);
  }
  c->lgB = lg;
  c->d_P = P;
  c->d_Pp = Pp;
  return e;
}

hipError_t launch_update_any(pfdtd_ctx *c, const UpdateGeom &g) {
  if (c->grouped && c->lgB > 0) return launch_update_grouped(c, g);
  const int ncol = coef_cols(c);
  switch (c->precision) {
  case PFDTD_F32:
    return launch_update_f32(g, UpdateArgs<float, CoefF, float>{c->d_pos, c->d_mat,
                             (const float *)c->d_P, (float *)c->d_Pp, c->d_coef_f, ncol,
                             c->air_f, c->air_uniform, c->lam2_f, c->scheme}, c->stream);
  case PFDTD_F64:
    return launch_update_f64(g, UpdateArgs<double, CoefD, double>{c->d_pos, c->d_mat,
                             (const double *)c->d_P, (double *)c->d_Pp, c->d_coef_d, ncol,
                             c->air_d, c->air_uniform, c->lam2_d, c->scheme}, c->stream);
  case PFDTD_F16:
    return launch_update_f16(g, UpdateArgs<h16, CoefF, float>{c->d_pos, c->d_mat,
                             (const h16 *)c->d_P, (h16 *)c->d_Pp, c->d_coef_f, ncol,
                             c->air_f, c->air_uniform, c->lam2_f, c->scheme}, c->stream);
  case PFDTD_BF16:
    return launch_update_bf16(g, UpdateArgs<b16, CoefF, float>{c->d_pos, c->d_mat,
                              (const b16 *)c->d_P, (b16 *)c->d_Pp, c->d_coef_f, ncol,
                              c->air_f, c->air_uniform, c->lam2_f, c->scheme}, c->stream);
  default:
    return hipErrorInvalidValue;
  }
}

// The two-level update needs the spare pair and no source on the never-updated slices
// z = 0 and z = Z-1 (their values would have to be carried between the four buffers).
bool fused_eligible(const pfdtd_ctx *c) {
  if ((c->lgB > 0 && !c->grouped) || levels_for(c, c->precision) < 2 || !c->d_P2 || !c->d_Pp2)
    return false;
  // the two-level kernel addresses each z-slice through a buffer descriptor (32-bit range)
  const uint64_t esz = c->precision == PFDTD_F64 ? 8 : (c->precision == PFDTD_F32 ? 4 : 2);
  if ((uint64_t)c->X * (uint64_t)c->Y * esz >= (1ull << 32)) return false;
  for (int s = 0; s < c->n_src; ++s) {
    const int32_t *q = &c->src_xyz[3 * s];
    if (lin_idx(c, q) >= 0 && (q[2] == 0 || q[2] == c->Z - 1)) return false;
  }
  return true;
}

template <typename S, typename CT, typename C>
Fused2Args<S, CT, C> fused_args(const pfdtd_ctx *c, const CT *coef, CT air, C lam2, int step,
                                bool pair) {
  Fused2Args<S, CT, C> a;
  a.pm = c->d_pm;
  a.air_rows = pair ? c->d_air_rows_p : c->d_air_rows;
  a.A = (const S *)c->d_P;
  a.B = (const S *)c->d_Pp;
  a.Cn = (S *)c->d_Pp2;
  a.Dn = (S *)c->d_P2;
  a.coef = coef;
  a.ncol = coef_cols(c);
  a.air = air;
  a.air_uniform = c->air_uniform;
  a.lam2 = lam2;
  a.src_xyz = c->d_src_xyz;
  a.src_kind = c->d_src_kind;
  a.n_src = c->n_src;
  a.sig_f = c->d_sig_f;
  a.sig_d = c->d_sig_d;
  a.n_samples = c->n_samples;
  a.inject_step = (c->n_src > 0 && step + 1 < c->resp_steps) ? step + 1 : -1;
  a.levels = 2;
  a.diff = c->precision == PFDTD_F16_DIFF ? 1 : 0;
  if (c->grouped && c->lgB > 0) {
    a.n_slots = 1 << c->lgB;
    a.slot_elems = c->N;
    a.src_slot = c->d_src_slot;
    a.rec_slot = c->d_rec_slot;
  }
  a.soft_overwrite = (c->flags & PFDTD_SOFT_OVERWRITE) ? 1 : 0;
  a.scale_exp = c->scale_exp;
  a.rec_xyz = c->d_rec_xyz;
  a.n_rec = c->n_rec;
  a.resp = c->d_resp;
  a.resp_steps = c->resp_steps;
  a.gather_step = step;
  return a;
}

// Steps `step` and `step+1` in one sweep: P^{step+1} lands in d_Pp2, P^{step+2} in d_P2. The
// difference form stores P in d_P2 and D in d_Pp2 after `levels` steps (2, or 1 for its last).
hipError_t launch_update2_any(pfdtd_ctx *c, const FusedGeom &g, int step, int levels = 2) {
  if (c->precision == PFDTD_F16_DIFF) {
    auto a = fused_args<h16>(c, c->d_coef_f, c->air_f, c->lam2_f, step, g.pair);
    a.levels = levels;
    return launch_update2_f16(g, a, c->stream);
  }
  switch (c->precision) {
  case PFDTD_F32:
    return launch_update2_f32(
        g, fused_args<float>(c, c->d_coef_f, c->air_f, c->lam2_f, step, g.pair), c->stream);
  case PFDTD_F64:
    return launch_update2_f64(
        g, fused_args<double>(c, c->d_coef_d, c->air_d, c->lam2_d, step, g.pair), c->stream);
  case PFDTD_F16:
    return launch_update2_f16(
        g, fused_args<h16>(c, c->d_coef_f, c->air_f, c->lam2_f, step, g.pair), c->stream);
  case PFDTD_BF16:
    return launch_update2_bf16(
        g, fused_args<b16>(c, c->d_coef_f, c->air_f, c->lam2_f, step, g.pair), c->stream);
  default:
    return hipErrorInvalidValue;
  }
}

// Steps `step`, `step+1` and `step+2` in one sweep (forward scheme): P^{step+2} lands in d_Pp2
// and P^{step+3} in d_P2; the receivers of `step` and `step+1` are gathered and the
// sources of `step+1` and `step+2` injected inside.
template <typename S, typename CT, typename C>
Fused2Args<S, CT, C> fused3_args(const pfdtd_ctx *c, const CT *coef, CT air, C lam2, int step) {
  auto a = fused_args<S>(c, coef, air, lam2, step, false);
  a.air_rows = c->d_air_rows3;
  a.levels = 3;
  a.gather_step2 = c->n_rec > 0 ? step + 1 : -1;
  a.inject_step2 = (c->n_src > 0 && step + 2 < c->resp_steps) ? step + 2 : -1;
  return a;
}
hipError_t launch_update3(pfdtd_ctx *c, const FusedGeom &g, int step) {
  switch (c->precision) {
  case PFDTD_F32:
    return launch_update3_f32(
        g, fused3_args<float>(c, c->d_coef_f, c->air_f, c->lam2_f, step), c->stream);
  case PFDTD_F64:
    return launch_update3_f64(
        g, fused3_args<double>(c, c->d_coef_d, c->air_d, c->lam2_d, step), c->stream);
  case PFDTD_F16:
    return launch_update3_f16(
        g, fused3_args<h16>(c, c->d_coef_f, c->air_f, c->lam2_f, step), c->stream);
  case PFDTD_BF16:
    return launch_update3_bf16(
        g, fused3_args<b16>(c, c->d_coef_f, c->air_f, c->lam2_f, step), c->stream);
  default:
    return hipErrorInvalidValue;
  }
}

// Single-step update tiling: the explicit setting, else by precision for unbatched runs and the
// grouped layout's single steps (FP32 4 rows per thread x 4-slice chunks, FP64 4 x 8, 16-bit
// storage 2 x 8) and 2 x 8 for the interleaved batched update. Per single step on the
// 2048x1536x768 room: FP32 6.97 ms (4 x 4) against 7.19 (2 x 8), FP64 13.38 (4 x 8) against
// 13.80, FP16 4.72 (2 x 8) against 5.17 (4 x 4), BF16 4.80 against 6.03; interleaved batches of
// 8-64 slots on 512x384x256 run 13 % slower per source with 4 x 4 than with 2 x 8.
UpdateGeom plan_single(const pfdtd_ctx *c) {
  int ry = c->ry, zc = c->zc;
  const int lgB = c->grouped ? 0 : c->lgB;
  if (ry == 0) {
    ry = 2;
    zc = 8;
    if (lgB == 0 && c->precision == PFDTD_F32) {
      ry = 4;
      zc = 4;
    } else if (lgB == 0 && c->precision == PFDTD_F64) {
      ry = 4;
    }
  }
  return plan_update_batch((int)c->X, (int)c->Y, (int)c->Z, ry, zc, lgB);
}

int ensure_events(pfdtd_ctx *c, size_t n) {
  while (c->events.size() < n) {
    hipEvent_t e;
    HIP_TRY(hipEventCreate(&e));
    c->events.push_back(e);
  }
  return PFDTD_OK;
}

}  // namespace

extern "C" {

const char *pfdtd_last_error(void) { return g_last_error.c_str(); }

int pfdtd_version(void) { return 1; }

int pfdtd_device_count(int *count) {
  if (!count) return fail(PFDTD_E_ARG, "null count");
  HIP_TRY(hipGetDeviceCount(count));
  return PFDTD_OK;
}

int pfdtd_device_mem(int device, uint64_t *free_bytes, uint64_t *total_bytes) {
  if (!free_bytes || !total_bytes) return fail(PFDTD_E_ARG, "null output");
  HIP_TRY(hipSetDevice(device));
  size_t f = 0, t = 0;
  HIP_TRY(hipMemGetInfo(&f, &t));
  *free_bytes = f;
  *total_bytes = t;
  return PFDTD_OK;
}

int pfdtd_ctx_create(int device, pfdtd_ctx **out) {
  if (!out) return fail(PFDTD_E_ARG, "null output pointer");
  *out = nullptr;
  int n = 0;
  HIP_TRY(hipGetDeviceCount(&n));
  if (device < 0 || device >= n)
    return fail(PFDTD_E_ARG, "device %d not present (%d devices)", device, n);
  HIP_TRY(hipSetDevice(device));
  int rs = resolve_scratch(device);
  if (rs) return rs;
  pfdtd_ctx *c = new pfdtd_ctx();
  c->device = device;
  hipError_t e = hipStreamCreate(&c->stream);
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_scan, sizeof(DomainScan));
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_coef_f, sizeof(CoefF) * kCoefEntries);
  if (e == hipSuccess) e = hipMalloc((void **)&c->d_coef_d, sizeof(CoefD) * kCoefEntries);
  if (e == hipSuccess) e = hipEventCreate(&c->ev_begin);
  if (e == hipSuccess) e = hipEventCreate(&c->ev_end);
  if (e != hipSuccess) {
    pfdtd_ctx_destroy(c);
    return fail(PFDTD_E_HIP, "context setup on device %d: %s", device, hipGetErrorString(e));
  }
  *out = c;
  return PFDTD_OK;
}

int pfdtd_ctx_destroy(pfdtd_ctx *c) {
  if (!c) return PFDTD_OK;
  (void)hipSetDevice(c->device);
  if (c->stream) (void)hipStreamSynchronize(c->stream);
  dfree(c->d_pos);
  dfree(c->d_mat);
  dfree(c->d_pm);
  dfree(c->d_air_rows);
  dfree(c->d_air_rows_p);
  dfree(c->d_air_rows3);
  dfree(c->d_scan);
  dfree(c->d_coef_f);
  dfree(c->d_coef_d);
  dfree(c->d_src_idx);
  dfree(c->d_src_xyz);
  dfree(c->d_src_kind);
  dfree(c->d_src_slot);
  dfree(c->d_rec_slot);
  dfree(c->d_sig_f);
  dfree(c->d_sig_d);
  dfree(c->d_rec_idx);
  dfree(c->d_rec_xyz);
  free_fields(c);
  dfree(c->d_resp);
  for (hipEvent_t e : c->events) (void)hipEventDestroy(e);
  if (c->ev_begin) (void)hipEventDestroy(c->ev_begin);
  if (c->ev_end) (void)hipEventDestroy(c->ev_end);
  if (c->stream) (void)hipStreamDestroy(c->stream);
  delete c;
  return PFDTD_OK;
}

int pfdtd_set_domain(pfdtd_ctx *c, const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y,
                     int64_t Z) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!pos || !mat) return fail(PFDTD_E_ARG, "null position or material array");
  rc = alloc_domain(c, X, Y, Z);
  if (rc) return rc;
  c->degenerate = 0;
  HIP_TRY(hipMemcpy(c->d_pos, pos, (size_t)c->N, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_mat, mat, (size_t)c->N, hipMemcpyHostToDevice));
  return finish_domain(c);
}

// Classifies a device-resident label grid into the context's domain bytes for the current
// boundary family (alloc_domain done by the caller).
static int classify_device_labels(pfdtd_ctx *c, const uint8_t *d_lab) {
  hipError_t e = hipSuccess;
  unsigned *d_bad = nullptr;
  c->degenerate = 0;
  if (c->scheme == 2) {
    e = hipMalloc((void **)&d_bad, sizeof(unsigned));
    if (e == hipSuccess) e = hipMemsetAsync(d_bad, 0, sizeof(unsigned), c->stream);
    if (e == hipSuccess)
      e = launch_classify_labels_centred(d_lab, c->d_pos, c->d_mat, (int)c->X, (int)c->Y,
                                         (int)c->Z, d_bad, c->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(&c->degenerate, d_bad, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream);
  } else {
    e = launch_classify_labels(d_lab, c->d_pos, c->d_mat, (int)c->X, (int)c->Y, (int)c->Z,
                               c->stream);
  }
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (d_bad) (void)hipFree(d_bad);
  if (e != hipSuccess) return fail(PFDTD_E_HIP, "label classification: %s", hipGetErrorString(e));
  return finish_domain(c);
}

int pfdtd_set_domain_labels(pfdtd_ctx *c, const uint8_t *labels, int64_t X, int64_t Y,
                            int64_t Z) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!labels) return fail(PFDTD_E_ARG, "null label array");
  rc = alloc_domain(c, X, Y, Z);
  if (rc) return rc;
  uint8_t *d_lab = nullptr;
  if (hipMalloc((void **)&d_lab, (size_t)c->N) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PFDTD_E_NOMEM, "cannot allocate the label grid");
  }
  hipError_t e = hipMemcpy(d_lab, labels, (size_t)c->N, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(d_lab);
    return fail(PFDTD_E_HIP, "label upload: %s", hipGetErrorString(e));
  }
  rc = classify_device_labels(c, d_lab);
  (void)hipFree(d_lab);
  return rc;
}

int pfdtd_voxel_grid(const float *vertices, int64_t n_vertices, float dx, const int32_t *pad,
                     int64_t *dims) {
  if (!vertices || !dims || n_vertices <= 0) return fail(PFDTD_E_ARG, "null or empty vertex array");
  if (pfdtd::voxel_prepare(vertices, n_vertices, dx, pad, dims, nullptr))
    return fail(PFDTD_E_ARG, "voxel grid: dx must be positive and finite, vertices finite, "
                             "and the grid at most 32768 nodes per axis");
  return PFDTD_OK;
}

int pfdtd_voxelize(pfdtd_ctx *c, const float *vertices, int64_t n_vertices,
                   const uint32_t *indices, int64_t n_triangles, const uint8_t *tri_material,
                   float dx, const int32_t *pad, int64_t *dims, uint8_t *labels) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!vertices || n_vertices <= 0) return fail(PFDTD_E_ARG, "null or empty vertex array");
  if (!indices || n_triangles <= 0) return fail(PFDTD_E_ARG, "null or empty triangle array");
  for (int64_t i = 0; i < 3 * n_triangles; ++i)
    if ((int64_t)indices[i] >= n_vertices)
      return fail(PFDTD_E_ARG, "triangle %lld names vertex %u of %lld", (long long)(i / 3),
                  indices[i], (long long)n_vertices);
  if (tri_material)
    for (int64_t t = 0; t < n_triangles; ++t)
      if (tri_material[t] > 254)
        return fail(PFDTD_E_ARG, "triangle %lld: material %u above 254", (long long)t,
                    tri_material[t]);
  int64_t d[3];
  std::vector<int32_t> q;
  if (pfdtd::voxel_prepare(vertices, n_vertices, dx, pad, d, &q))
    return fail(PFDTD_E_ARG, "voxel grid: dx must be positive and finite, vertices finite, "
                             "and the grid at most 32768 nodes per axis");
  rc = alloc_domain(c, d[0], d[1], d[2]);
  if (rc) return rc;
  uint8_t *d_lab = nullptr;
  if (hipMalloc((void **)&d_lab, (size_t)c->N) != hipSuccess) {
    (void)hipGetLastError();
    return fail(PFDTD_E_NOMEM, "cannot allocate the label grid");
  }
  hipError_t e = pfdtd::voxelize_labels(q, indices, n_triangles, tri_material, d, d_lab, c->stream);
  if (e == hipSuccess && labels) e = hipMemcpy(labels, d_lab, (size_t)c->N, hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    (void)hipFree(d_lab);
    (void)hipGetLastError();
    return fail(e == hipErrorOutOfMemory ? PFDTD_E_NOMEM : PFDTD_E_HIP, "voxelization: %s",
                hipGetErrorString(e));
  }
  rc = classify_device_labels(c, d_lab);
  (void)hipFree(d_lab);
  if (rc == PFDTD_OK && dims) {
    dims[0] = d[0];
    dims[1] = d[1];
    dims[2] = d[2];
  }
  return rc;
}

int pfdtd_set_domain_codes(pfdtd_ctx *c, const uint8_t *codes, const uint8_t *mat, int64_t X,
                           int64_t Y, int64_t Z) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!codes || !mat) return fail(PFDTD_E_ARG, "null code or material array");
  rc = alloc_domain(c, X, Y, Z);
  if (rc) return rc;
  HIP_TRY(hipMemcpy(c->d_pos, codes, (size_t)c->N, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_mat, mat, (size_t)c->N, hipMemcpyHostToDevice));
  c->degenerate = 0;
  if (c->scheme == 2)
    HIP_TRY(launch_codes_to_centred(c->d_pos, c->d_mat, c->N, c->stream));
  else
    HIP_TRY(launch_codes_to_bilbao(c->d_pos, c->d_mat, c->N, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return finish_domain(c);
}

int pfdtd_get_domain(pfdtd_ctx *c, uint8_t *pos, uint8_t *mat) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->has_domain) return fail(PFDTD_E_STATE, "no domain set");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (pos) HIP_TRY(hipMemcpy(pos, c->d_pos, (size_t)c->N, hipMemcpyDeviceToHost));
  if (mat) HIP_TRY(hipMemcpy(mat, c->d_mat, (size_t)c->N, hipMemcpyDeviceToHost));
  return PFDTD_OK;
}

int pfdtd_get_dims(pfdtd_ctx *c, int64_t *X, int64_t *Y, int64_t *Z) {
  if (!c || !X || !Y || !Z) return fail(PFDTD_E_ARG, "null argument");
  *X = c->X;
  *Y = c->Y;
  *Z = c->Z;
  return PFDTD_OK;
}

int pfdtd_get_counts(pfdtd_ctx *c, uint64_t *air, uint64_t *boundary) {
  if (!c || !air || !boundary) return fail(PFDTD_E_ARG, "null argument");
  if (!c->has_domain) return fail(PFDTD_E_STATE, "no domain set");
  *air = c->scan.air;
  *boundary = c->scan.boundary;
  return PFDTD_OK;
}

// Builds and uploads the coefficient tables of the current scheme and the air entry (K = 6 for
// the forward scheme, D = 0 for the centred one, where it holds for every material byte).
static int upload_coefficients(pfdtd_ctx *c) {
  std::vector<CoefF> cf;
  std::vector<CoefD> cd;
  build_coefficients(c, cf, cd);
  const size_t air_row = c->scheme == 2 ? 0 : 6 * 256;
  c->air_f = cf[air_row];
  c->air_d = cd[air_row];
  c->air_uniform = 1;
  for (int m = 0; m < 256; ++m) {
    if (std::memcmp(&cf[air_row + m], &c->air_f, sizeof(CoefF)) != 0 ||
        std::memcmp(&cd[air_row + m], &c->air_d, sizeof(CoefD)) != 0)
      c->air_uniform = 0;
  }
  HIP_TRY(hipMemcpy(c->d_coef_f, cf.data(), sizeof(CoefF) * cf.size(), hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_coef_d, cd.data(), sizeof(CoefD) * cd.size(), hipMemcpyHostToDevice));
  return PFDTD_OK;
}

int pfdtd_set_materials(pfdtd_ctx *c, const float *table, int n_mat, int n_coef,
                        unsigned octave, double lambda, int flags) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!table) return fail(PFDTD_E_ARG, "null material table");
  if (n_mat < 1 || n_mat > 256) return fail(PFDTD_E_ARG, "n_mat %d outside 1..256", n_mat);
  if (n_coef < 1 || n_coef > kCoefNum)
    return fail(PFDTD_E_ARG, "n_coef %d outside 1..%d", n_coef, kCoefNum);
  if (octave >= (unsigned)kCoefNum) return fail(PFDTD_E_ARG, "octave %u >= 20", octave);
  if (!(lambda > 0.0) || !std::isfinite(lambda)) return fail(PFDTD_E_ARG, "bad lambda");
  if (flags & ~(PFDTD_MATSLOT_LITERAL | PFDTD_SOFT_OVERWRITE))
    return fail(PFDTD_E_ARG, "unknown flags 0x%x", flags);
  c->table.assign((size_t)n_mat * kCoefNum, 0.f);
  for (int m = 0; m < n_mat; ++m)
    for (int j = 0; j < n_coef; ++j) c->table[(size_t)m * kCoefNum + j] = table[(size_t)m * n_coef + j];
  c->n_mat = n_mat;
  c->octave = octave;
  c->lambda = lambda;
  c->flags = flags;
  c->lam2_f = (float)(lambda * lambda);
  c->lam2_d = lambda * lambda;
  c->has_materials = true;
  return upload_coefficients(c);
}

int pfdtd_transparent_signals(pfdtd_ctx *c, int n_src, int n_samples, const float *grid_ir,
                              int n_ir, const float *in_f32, float *out_f32,
                              const double *in_f64, double *out_f64) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n_src < 0 || n_samples < 0 || n_ir < 0) return fail(PFDTD_E_ARG, "negative count");
  if (n_ir > 0 && !grid_ir) return fail(PFDTD_E_ARG, "null grid impulse response");
  if ((in_f32 == nullptr) != (out_f32 == nullptr) || (in_f64 == nullptr) != (out_f64 == nullptr))
    return fail(PFDTD_E_ARG, "each precision needs both an input and an output array");
  const size_t n = (size_t)n_src * (size_t)n_samples;
  if (!n) return PFDTD_OK;
  float *d_h = nullptr;
  void *d_in = nullptr, *d_out = nullptr;
  hipError_t e = hipMalloc((void **)&d_h, sizeof(float) * (size_t)(n_ir > 0 ? n_ir : 1));
  if (e == hipSuccess && n_ir > 0)
    e = hipMemcpy(d_h, grid_ir, sizeof(float) * (size_t)n_ir, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMalloc(&d_in, sizeof(double) * n);
  if (e == hipSuccess) e = hipMalloc(&d_out, sizeof(double) * n);
  for (int pass = 0; pass < 2 && e == hipSuccess; ++pass) {
    const bool f64 = pass == 1;
    const void *in = f64 ? (const void *)in_f64 : (const void *)in_f32;
    void *out = f64 ? (void *)out_f64 : (void *)out_f32;
    if (!in) continue;
    const size_t bytes = (f64 ? sizeof(double) : sizeof(float)) * n;
    e = hipMemcpy(d_in, in, bytes, hipMemcpyHostToDevice);
    if (e == hipSuccess)
      e = launch_transparent(f64, d_in, d_h, n_ir, n_src, n_samples, d_out, c->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_out, bytes, hipMemcpyDeviceToHost);
  }
  (void)hipFree(d_h);
  if (d_in) (void)hipFree(d_in);
  if (d_out) (void)hipFree(d_out);
  if (e != hipSuccess) return fail(PFDTD_E_HIP, "transparent signals: %s", hipGetErrorString(e));
  return PFDTD_OK;
}

int pfdtd_set_update_type(pfdtd_ctx *c, int type) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (type < 0 || type > 2) return fail(PFDTD_E_ARG, "update type %d not in {0,1,2}", type);
  const int scheme = type == 2 ? 2 : 0;
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->update_type = type;
  if (scheme != c->scheme) {
    // position bytes encode the boundary formulation: a classified domain no longer applies
    c->scheme = scheme;
    c->has_domain = false;
    if (c->has_materials) return upload_coefficients(c);
  }
  return PFDTD_OK;
}

int pfdtd_set_sources(pfdtd_ctx *c, int n_src, const int32_t *xyz, const int32_t *kind,
                      int n_samples, const float *signal_f32, const double *signal_f64) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n_src < 0 || n_samples < 0) return fail(PFDTD_E_ARG, "negative count");
  if (n_src > 0 && (!xyz || !kind)) return fail(PFDTD_E_ARG, "null source arrays");
  if (n_src > 0 && n_samples > 0 && !signal_f32 && !signal_f64)
    return fail(PFDTD_E_ARG, "no source signal");
  for (int s = 0; s < n_src; ++s)
    if (kind[s] < PFDTD_SRC_HARD || kind[s] > PFDTD_SRC_TRANSPARENT)
      return fail(PFDTD_E_ARG, "source %d: unknown kind %d", s, kind[s]);
  HIP_TRY(hipStreamSynchronize(c->stream));
  dfree(c->d_src_idx);
  dfree(c->d_src_xyz);
  dfree(c->d_src_kind);
  dfree(c->d_src_slot);
  dfree(c->d_sig_f);
  dfree(c->d_sig_d);
  c->n_src = 0;
  c->n_samples = 0;
  c->src_xyz.assign(xyz, xyz + 3 * (size_t)n_src);
  c->src_kind.assign(kind, kind + n_src);
  c->src_slot.assign((size_t)n_src, 0);
  if (n_src == 0) return PFDTD_OK;
  const size_t ns = (size_t)n_src * (size_t)n_samples;
  std::vector<float> sf(ns);
  std::vector<double> sd(ns);
  for (size_t i = 0; i < ns; ++i) {
    sf[i] = signal_f32 ? signal_f32[i] : (float)signal_f64[i];
    sd[i] = signal_f64 ? signal_f64[i] : (double)signal_f32[i];
  }
  std::vector<int64_t> idx(n_src, -1);
  if (c->has_domain)
    for (int s = 0; s < n_src; ++s) idx[s] = elem_idx(c, &c->src_xyz[3 * s], 0);
  HIP_TRY(hipMalloc((void **)&c->d_src_idx, sizeof(int64_t) * n_src));
  HIP_TRY(hipMalloc((void **)&c->d_src_kind, sizeof(int32_t) * n_src));
  HIP_TRY(hipMalloc((void **)&c->d_src_slot, sizeof(int32_t) * n_src));
  HIP_TRY(hipMemcpy(c->d_src_slot, c->src_slot.data(), sizeof(int32_t) * n_src, hipMemcpyHostToDevice));
  HIP_TRY(hipMalloc((void **)&c->d_src_xyz, sizeof(int32_t) * 3 * n_src));
  {
    int rc2 = upload_xyz(c, c->src_xyz, c->d_src_xyz);
    if (rc2) return rc2;
  }
  HIP_TRY(hipMalloc((void **)&c->d_sig_f, sizeof(float) * (ns ? ns : 1)));
  HIP_TRY(hipMalloc((void **)&c->d_sig_d, sizeof(double) * (ns ? ns : 1)));
  c->src_distinct = distinct_indices(idx);
  HIP_TRY(hipMemcpy(c->d_src_idx, idx.data(), sizeof(int64_t) * n_src, hipMemcpyHostToDevice));
  HIP_TRY(hipMemcpy(c->d_src_kind, kind, sizeof(int32_t) * n_src, hipMemcpyHostToDevice));
  if (ns) {
    HIP_TRY(hipMemcpy(c->d_sig_f, sf.data(), sizeof(float) * ns, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_sig_d, sd.data(), sizeof(double) * ns, hipMemcpyHostToDevice));
  }
  c->n_src = n_src;
  c->n_samples = n_samples;
  return PFDTD_OK;
}

int pfdtd_set_receivers(pfdtd_ctx *c, int n_rec, const int32_t *xyz) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n_rec < 0) return fail(PFDTD_E_ARG, "negative receiver count");
  if (n_rec > 0 && !xyz) return fail(PFDTD_E_ARG, "null receiver array");
  HIP_TRY(hipStreamSynchronize(c->stream));
  dfree(c->d_rec_idx);
  dfree(c->d_rec_xyz);
  dfree(c->d_rec_slot);
  c->n_rec = 0;
  c->rec_xyz.assign(xyz, xyz + 3 * (size_t)n_rec);
  c->rec_slot.assign((size_t)n_rec, 0);
  c->resp_steps = 0;
  if (n_rec == 0) return PFDTD_OK;
  std::vector<int64_t> idx(n_rec, -1);
  if (c->has_domain)
    for (int r = 0; r < n_rec; ++r) idx[r] = elem_idx(c, &c->rec_xyz[3 * r], 0);
  HIP_TRY(hipMalloc((void **)&c->d_rec_idx, sizeof(int64_t) * n_rec));
  HIP_TRY(hipMalloc((void **)&c->d_rec_xyz, sizeof(int32_t) * 3 * n_rec));
  HIP_TRY(hipMalloc((void **)&c->d_rec_slot, sizeof(int32_t) * n_rec));
  HIP_TRY(hipMemcpy(c->d_rec_slot, c->rec_slot.data(), sizeof(int32_t) * n_rec, hipMemcpyHostToDevice));
  {
    int rc2 = upload_xyz(c, c->rec_xyz, c->d_rec_xyz);
    if (rc2) return rc2;
  }
  HIP_TRY(hipMemcpy(c->d_rec_idx, idx.data(), sizeof(int64_t) * n_rec, hipMemcpyHostToDevice));
  c->n_rec = n_rec;
  return PFDTD_OK;
}

int pfdtd_prepare_fields(pfdtd_ctx *c, int precision, int scale_exp) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->has_domain) return fail(PFDTD_E_STATE, "no domain set");
  rc = ensure_fields(c, precision, scale_exp);
  if (rc) return rc;
  if (precision == PFDTD_F16 || precision == PFDTD_BF16 || precision == PFDTD_F16_DIFF)
    c->scale_req = scale_exp;
  return zero_fields(c);
}

int pfdtd_reset_fields(pfdtd_ctx *c) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->d_P) return fail(PFDTD_E_STATE, "fields not prepared");
  return zero_fields(c);
}

// The two-level sweep reads position and material as one 16-bit plane (2 B per voxel), built
// once per domain with its air-row bits; without room for the plane *fused is cleared and the
// caller takes single steps.
static int ensure_pm(pfdtd_ctx *c, bool *fused) {
  if (!*fused || c->pm_ok) return PFDTD_OK;
  if (!c->d_pm && hipMalloc((void **)&c->d_pm, sizeof(uint16_t) * (size_t)c->N) != hipSuccess) {
    (void)hipGetLastError();
    c->d_pm = nullptr;
    *fused = false;
    return PFDTD_OK;
  }
  HIP_TRY(launch_pack_pm(c->d_pos, c->d_mat, c->d_pm, c->N, c->stream));
  // the air-row bits only save loads: without room for them the sweep reads every row
  for (int pair = 0; pair < 2; ++pair) {
    uint32_t *&rows = pair ? c->d_air_rows_p : c->d_air_rows;
    if (pair && c->X % 2 != 0) continue;  // the two-column sweep needs X even
    const size_t words = air_rows_words((int)c->X, (int)c->Y, (int)c->Z, pair != 0);
    if (!rows && hipMalloc((void **)&rows, sizeof(uint32_t) * words) != hipSuccess) {
      (void)hipGetLastError();
      rows = nullptr;
    }
    if (rows && launch_air_rows(c->d_pm, rows, (int)c->X, (int)c->Y, (int)c->Z, c->scheme,
                                pair != 0, c->stream) != hipSuccess) {
      (void)hipGetLastError();
      dfree(rows);
    }
  }
  c->pm_ok = true;
  return PFDTD_OK;
}

// The air-row bits of the three-level sweep's tiles, built at its first run on a domain (left
// null without room for them: the sweep then reads every row's packed value).
static int ensure_air_rows3(pfdtd_ctx *c) {
  if (c->air3_ok) return PFDTD_OK;
  const size_t words = air_rows_words((int)c->X, (int)c->Y, (int)c->Z, 2);
  if (!c->d_air_rows3 &&
      hipMalloc((void **)&c->d_air_rows3, sizeof(uint32_t) * words) != hipSuccess) {
    (void)hipGetLastError();
    c->d_air_rows3 = nullptr;
  }
  if (c->d_air_rows3 && launch_air_rows(c->d_pm, c->d_air_rows3, (int)c->X, (int)c->Y,
                                        (int)c->Z, c->scheme, 2, c->stream) != hipSuccess) {
    (void)hipGetLastError();
    dfree(c->d_air_rows3);
  }
  c->air3_ok = true;
  return PFDTD_OK;
}

int pfdtd_run(pfdtd_ctx *c, int precision, int n_steps, pfdtd_interrupt_cb interrupt,
              pfdtd_progress_cb progress, void *user, int *steps_done, double *ms_elapsed) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n_steps < 0) return fail(PFDTD_E_ARG, "negative step count");
  rc = check_runnable(c);
  if (rc) return rc;
  const int scale = (precision == PFDTD_F16 || precision == PFDTD_BF16 ||
                     precision == PFDTD_F16_DIFF) ? c->scale_req : 0;
  rc = ensure_fields(c, precision, scale);
  if (rc) return rc;
  const bool diff = precision == PFDTD_F16_DIFF;
  if (diff && (!fused_eligible(c) || c->scheme != 0 || c->X % 2 != 0 || c->lgB > 0 ||
               c->X * c->Y * 2 >= ((int64_t)1 << 31)))
    return fail(PFDTD_E_STATE, "the difference form runs through the two-column sweep only: forward "
                "update, X even, X*Y below 2^30, no batch, no source on the slices z = 0 or Z-1");
  rc = zero_fields(c);
  if (rc) return rc;
  rc = ensure_responses(c, n_steps);
  if (rc) return rc;
  if (steps_done) *steps_done = 0;
  c->upd_ms = 0.0;
  c->upd_launches = 0;
  const UpdateGeom g = plan_single(c);
  const int bundle = c->poll_interval > 0 ? c->poll_interval : kProgressMod;
  if (c->kernel_timing) {
    rc = ensure_events(c, 2 * (size_t)bundle);
    if (rc) return rc;
  }

  HIP_TRY(hipEventRecord(c->ev_begin, c->stream));
  // step 0's source samples go into the (zero) current field before the first update
  if (n_steps > 0) {
    InjectArgs ia = inject_args(c, -1, 0);
    if (diff) ia.dfield = c->d_Pp;
    HIP_TRY(launch_inject_gather(p_storage(c->precision), c->d_P, ia, c->stream));
  }

  bool fused = fused_eligible(c);
  rc = ensure_pm(c, &fused);
  if (rc) return rc;
  const FusedGeom fg = plan_fused((int)c->X, (int)c->Y, (int)c->Z, c->fzc, c->precision, false,
                                  diff ? PFDTD_SWEEP_PAIR : c->fshape, c->scheme, batch_slots(c));
  if (diff && (!fused || !fg.pair))
    return fail(PFDTD_E_NOMEM, "no room for the packed plane the two-column sweep reads");
  // three levels per sweep with the forward scheme when asked for (not the difference form);
  // the centred scheme takes the two-level sweep
  const bool three = fused && !diff && c->scheme == 0 && levels_for(c, c->precision) == 3;
  FusedGeom fg3;
  if (three) {
    fg3 = plan_fused3((int)c->X, (int)c->Y, (int)c->Z, c->fzc, c->precision, c->fshape,
                      batch_slots(c));
    rc = ensure_air_rows3(c);
    if (rc) return rc;
  }
  int done = 0;
  auto wall0 = std::chrono::steady_clock::now();
  while (done < n_steps) {
    if (interrupt && interrupt(user)) break;  // kernels3d.cu:86-89
    const int nb = std::min(bundle, n_steps - done);
    int nev = 0;
    for (int k = 0; k < nb;) {
      const int step = done + k;
      // four steps left take two two-level sweeps rather than three levels and a single step
      if (three && k + 3 <= nb && nb - k != 4) {
        if (c->kernel_timing) HIP_TRY(hipEventRecord(c->events[2 * nev], c->stream));
        HIP_TRY(launch_update3(c, fg3, step));
        if (c->kernel_timing) HIP_TRY(hipEventRecord(c->events[2 * nev + 1], c->stream));
        ++nev;
        // P^{step+3} (d_P2): gather step+2, inject step+3
        const int next = step + 3 < n_steps ? step + 3 : -1;
        HIP_TRY(launch_inject_gather(p_storage(c->precision), c->d_P2,
                                     inject_args(c, step + 2, next), c->stream));
        std::swap(c->d_P, c->d_P2);
        std::swap(c->d_Pp, c->d_Pp2);
        k += 3;
        c->upd_launches += 3;
      } else if (fused && k + 2 <= nb) {
        if (c->kernel_timing) HIP_TRY(hipEventRecord(c->events[2 * nev], c->stream));
        HIP_TRY(launch_update2_any(c, fg, step));
        if (c->kernel_timing) HIP_TRY(hipEventRecord(c->events[2 * nev + 1], c->stream));
        ++nev;
        // P^{step+2} (d_P2): gather step+1, inject step+2
        const int next = step + 2 < n_steps ? step + 2 : -1;
        InjectArgs ia = inject_args(c, step + 1, next);
        if (diff) ia.dfield = c->d_Pp2;
        HIP_TRY(launch_inject_gather(p_storage(c->precision), c->d_P2, ia, c->stream));
        std::swap(c->d_P, c->d_P2);
        std::swap(c->d_Pp, c->d_Pp2);
        k += 2;
        c->upd_launches += 2;
      } else if (diff) {
        // the difference form's odd last step: one level of the sweep (gather of `step` and the
        // injection of step+1 happen inside it)
        if (c->kernel_timing) HIP_TRY(hipEventRecord(c->events[2 * nev], c->stream));
        HIP_TRY(launch_update2_any(c, fg, step, 1));
        if (c->kernel_timing) HIP_TRY(hipEventRecord(c->events[2 * nev + 1], c->stream));
        ++nev;
        std::swap(c->d_P, c->d_P2);
        std::swap(c->d_Pp, c->d_Pp2);
        k += 1;
        c->upd_launches += 1;
      } else {
        if (c->kernel_timing) HIP_TRY(hipEventRecord(c->events[2 * nev], c->stream));
        HIP_TRY(launch_update_any(c, g));
        if (c->kernel_timing) HIP_TRY(hipEventRecord(c->events[2 * nev + 1], c->stream));
        ++nev;
        // after the update the new field lives in d_Pp: gather this step, inject the next
        const int next = step + 1 < n_steps ? step + 1 : -1;
        HIP_TRY(launch_inject_gather(c->precision, c->d_Pp, inject_args(c, step, next), c->stream));
        std::swap(c->d_P, c->d_Pp);  // flipPressurePointers, cudaMesh.h:766-779
        k += 1;
        c->upd_launches += 1;
      }
    }
    if (interrupt || progress || c->kernel_timing) {
      HIP_TRY(hipStreamSynchronize(c->stream));
      if (c->kernel_timing) {
        for (int k = 0; k < nev; ++k) {
          float ms = 0.f;
          HIP_TRY(hipEventElapsedTime(&ms, c->events[2 * k], c->events[2 * k + 1]));
          c->upd_ms += ms;
        }
      }
    }
    const int first = done;
    done += nb;
    if (progress) {
      // PROGRESS_MOD steps: report at every multiple of 100 passed in this bundle
      const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - wall0).count();
      for (int s = ((first + kProgressMod - 1) / kProgressMod) * kProgressMod; s < done; s += kProgressMod)
        progress(s, n_steps, (float)(sec / done), user);
    }
  }
  HIP_TRY(hipEventRecord(c->ev_end, c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (ms_elapsed) {
    float ms = 0.f;
    HIP_TRY(hipEventElapsedTime(&ms, c->ev_begin, c->ev_end));
    *ms_elapsed = ms;
  }
  if (steps_done) *steps_done = done;
  return PFDTD_OK;
}

int pfdtd_begin_steps(pfdtd_ctx *c, int precision, int n_steps) {
  if (precision == PFDTD_F16_DIFF)
    return fail(PFDTD_E_ARG, "the difference form (PFDTD_F16_DIFF) runs through pfdtd_run only");
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n_steps < 0) return fail(PFDTD_E_ARG, "negative step count");
  rc = check_runnable(c);
  if (rc) return rc;
  const int scale = (precision == PFDTD_F16 || precision == PFDTD_BF16) ? c->scale_req : 0;
  rc = ensure_fields(c, precision, scale);
  if (rc) return rc;
  rc = zero_fields(c);
  if (rc) return rc;
  return ensure_responses(c, n_steps);
}

int pfdtd_step(pfdtd_ctx *c, int step, int direction) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->d_P || c->precision < 0) return fail(PFDTD_E_STATE, "fields not prepared");
  if (direction != 1 && direction != -1) return fail(PFDTD_E_ARG, "direction must be +1 or -1");
  rc = check_runnable(c);
  if (rc) return rc;
  // launchFDTD3dStep: inject this step's samples, update, flip unless the direction changed,
  // then read receivers into response[step] (kernels3d.cu:390-475).
  if (step >= 0) HIP_TRY(launch_inject_gather(c->precision, c->d_P, inject_args(c, -1, step), c->stream));
  const UpdateGeom g = plan_single(c);
  HIP_TRY(launch_update_any(c, g));
  if (c->last_direction == direction) std::swap(c->d_P, c->d_Pp);
  c->last_direction = direction;
  if (step >= 0 && step < c->resp_steps)
    HIP_TRY(launch_inject_gather(c->precision, c->d_P, inject_args(c, step, -1), c->stream));
  HIP_TRY(hipStreamSynchronize(c->stream));
  return PFDTD_OK;
}

int pfdtd_steps(pfdtd_ctx *c, int first_step, int n_steps) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->d_P || c->precision < 0) return fail(PFDTD_E_STATE, "fields not prepared");
  if (first_step < 0 || n_steps < 0 || (int64_t)first_step + n_steps > c->resp_steps)
    return fail(PFDTD_E_ARG, "steps %d..%d outside the %d steps of pfdtd_begin_steps", first_step,
                first_step + n_steps - 1, c->resp_steps);
  rc = check_runnable(c);
  if (rc) return rc;
  if (n_steps == 0) return PFDTD_OK;
  // after a backward step the next forward one keeps the pointers (the flip rule of
  // pfdtd_step): take it on its own
  if (c->last_direction != 1) {
    rc = pfdtd_step(c, first_step, 1);
    if (rc) return rc;
    ++first_step;
    --n_steps;
    if (n_steps == 0) return PFDTD_OK;
  }
  const UpdateGeom g = plan_single(c);
  bool fused = fused_eligible(c);
  rc = ensure_pm(c, &fused);
  if (rc) return rc;
  const FusedGeom fg = plan_fused((int)c->X, (int)c->Y, (int)c->Z, c->fzc, c->precision, false,
                                  c->fshape, c->scheme, batch_slots(c));
  const bool three = fused && c->scheme == 0 && levels_for(c, c->precision) == 3;
  FusedGeom fg3;
  if (three) {
    fg3 = plan_fused3((int)c->X, (int)c->Y, (int)c->Z, c->fzc, c->precision, c->fshape,
                      batch_slots(c));
    rc = ensure_air_rows3(c);
    if (rc) return rc;
  }
  // the sequence of pfdtd_step calls: inject(s), update, gather(s) for s = first .. first+n-1,
  // with the injection of each following step applied right after the previous gather
  HIP_TRY(launch_inject_gather(c->precision, c->d_P, inject_args(c, -1, first_step), c->stream));
  for (int k = 0; k < n_steps;) {
    const int step = first_step + k;
    if (three && k + 3 <= n_steps && n_steps - k != 4) {
      HIP_TRY(launch_update3(c, fg3, step));
      const int next = k + 3 < n_steps ? step + 3 : -1;
      HIP_TRY(launch_inject_gather(c->precision, c->d_P2, inject_args(c, step + 2, next), c->stream));
      std::swap(c->d_P, c->d_P2);
      std::swap(c->d_Pp, c->d_Pp2);
      k += 3;
    } else if (fused && k + 2 <= n_steps) {
      HIP_TRY(launch_update2_any(c, fg, step));
      const int next = k + 2 < n_steps ? step + 2 : -1;
      HIP_TRY(launch_inject_gather(c->precision, c->d_P2, inject_args(c, step + 1, next), c->stream));
      std::swap(c->d_P, c->d_P2);
      std::swap(c->d_Pp, c->d_Pp2);
      k += 2;
    } else {
      HIP_TRY(launch_update_any(c, g));
      const int next = k + 1 < n_steps ? step + 1 : -1;
      HIP_TRY(launch_inject_gather(c->precision, c->d_Pp, inject_args(c, step, next), c->stream));
      std::swap(c->d_P, c->d_Pp);
      k += 1;
    }
  }
  c->last_direction = 1;
  HIP_TRY(hipStreamSynchronize(c->stream));
  return PFDTD_OK;
}

int pfdtd_get_response_shape(pfdtd_ctx *c, int *n_rec, int *n_steps) {
  if (!c || !n_rec || !n_steps) return fail(PFDTD_E_ARG, "null argument");
  *n_rec = c->n_rec;
  *n_steps = c->resp_steps;
  return PFDTD_OK;
}

int pfdtd_get_responses_f64(pfdtd_ctx *c, double *out, int64_t capacity) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const int64_t need = (int64_t)c->n_rec * c->resp_steps;
  if (!out || capacity < need) return fail(PFDTD_E_ARG, "response buffer too small (%lld < %lld)", (long long)capacity, (long long)need);
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (need) HIP_TRY(hipMemcpy(out, c->d_resp, sizeof(double) * need, hipMemcpyDeviceToHost));
  return PFDTD_OK;
}

int pfdtd_get_responses_f32(pfdtd_ctx *c, float *out, int64_t capacity) {
  int rc = check_ctx(c);
  if (rc) return rc;
  const int64_t need = (int64_t)c->n_rec * c->resp_steps;
  if (!out || capacity < need) return fail(PFDTD_E_ARG, "response buffer too small (%lld < %lld)", (long long)capacity, (long long)need);
  std::vector<double> tmp((size_t)need);
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (need) HIP_TRY(hipMemcpy(tmp.data(), c->d_resp, sizeof(double) * need, hipMemcpyDeviceToHost));
  for (int64_t i = 0; i < need; ++i) out[i] = (float)tmp[(size_t)i];
  return PFDTD_OK;
}

namespace {

// Checks of pfdtd_capture_slice(_noalloc); *n = elements of the plane.
int capture_plane(pfdtd_ctx *c, int64_t slice, int orientation, const float *p_out,
                  const uint8_t *pos_out, int64_t *n) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->lgB) return fail(PFDTD_E_STATE, "capture of batched fields is not supported (use pfdtd_get_fields_raw)");
  if (!c->d_P || c->precision < 0) return fail(PFDTD_E_STATE, "fields not prepared");
  if (!p_out || !pos_out) return fail(PFDTD_E_ARG, "null output");
  int64_t lim, U, V;
  switch (orientation) {
  case 0: lim = c->Z; U = c->X; V = c->Y; break;
  case 1: lim = c->Y; U = c->X; V = c->Z; break;
  case 2: lim = c->X; U = c->Y; V = c->Z; break;
  default: return fail(PFDTD_E_ARG, "orientation %d not in {0,1,2}", orientation);
  }
  if (slice < 0 || slice >= lim)
    return fail(PFDTD_E_ARG, "slice %lld out of bounds %lld, no capture made", (long long)slice, (long long)lim);
  *n = U * V;
  return PFDTD_OK;
}

}  // namespace

int pfdtd_capture_slice(pfdtd_ctx *c, int64_t slice, int orientation, float *p_out,
                        uint8_t *pos_out) {
  int64_t n = 0;
  int rc = capture_plane(c, slice, orientation, p_out, pos_out, &n);
  if (rc) return rc;
  float *d_p = nullptr;
  uint8_t *d_q = nullptr;
  HIP_TRY(hipMalloc((void **)&d_p, sizeof(float) * n));
  hipError_t e = hipMalloc((void **)&d_q, (size_t)n);
  if (e == hipSuccess)
    e = launch_capture_slice(p_storage(c->precision), c->d_P, c->d_pos, (int)c->X, (int)c->Y, (int)c->Z,
                             slice, orientation, c->scale_exp, d_p, d_q, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(p_out, d_p, sizeof(float) * n, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(pos_out, d_q, (size_t)n, hipMemcpyDeviceToHost);
  (void)hipFree(d_p);
  if (d_q) (void)hipFree(d_q);
  if (e != hipSuccess) return fail(PFDTD_E_HIP, "capture_slice: %s", hipGetErrorString(e));
  return PFDTD_OK;
}

int pfdtd_capture_slice_noalloc(pfdtd_ctx *c, int64_t slice, int orientation, float *p_out,
                                uint8_t *pos_out) {
  int64_t n = 0;
  int rc = capture_plane(c, slice, orientation, p_out, pos_out, &n);
  if (rc) return rc;
  if (n > PFDTD_STATIC_CAPTURE_MAX)
    return fail(PFDTD_E_ARG, "plane of %lld elements exceeds the static capture buffer (%lld)",
                (long long)n, (long long)PFDTD_STATIC_CAPTURE_MAX);
  unsigned char *base = (unsigned char *)scratch_on(c->device);
  if (!base) return fail(PFDTD_E_STATE, "static capture buffer not resolved on device %d", c->device);
  float *d_p = (float *)base;
  uint8_t *d_q = base + (size_t)PFDTD_STATIC_CAPTURE_MAX * sizeof(float);
  hipError_t e = launch_capture_slice(p_storage(c->precision), c->d_P, c->d_pos, (int)c->X, (int)c->Y,
                                      (int)c->Z, slice, orientation, c->scale_exp, d_p, d_q,
                                      c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(p_out, d_p, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(pos_out, d_q, (size_t)n, hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e != hipSuccess) return fail(PFDTD_E_HIP, "capture_slice_noalloc: %s", hipGetErrorString(e));
  return PFDTD_OK;
}

int pfdtd_capture_field(pfdtd_ctx *c, float *out, int64_t capacity) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (c->lgB) return fail(PFDTD_E_STATE, "capture of batched fields is not supported (use pfdtd_get_fields_raw)");
  if (!c->d_P || c->precision < 0) return fail(PFDTD_E_STATE, "fields not prepared");
  if (!out || capacity < c->N) return fail(PFDTD_E_ARG, "output buffer too small");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (c->precision == PFDTD_F32) {
    HIP_TRY(hipMemcpy(out, c->d_P, sizeof(float) * c->N, hipMemcpyDeviceToHost));
    return PFDTD_OK;
  }
  float *d_tmp = nullptr;
  HIP_TRY(hipMalloc((void **)&d_tmp, sizeof(float) * c->N));
  hipError_t e = launch_field_to_f32(p_storage(c->precision), c->d_P, c->N, c->scale_exp, d_tmp, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess) e = hipMemcpy(out, d_tmp, sizeof(float) * c->N, hipMemcpyDeviceToHost);
  (void)hipFree(d_tmp);
  if (e != hipSuccess) return fail(PFDTD_E_HIP, "capture_field: %s", hipGetErrorString(e));
  return PFDTD_OK;
}

int pfdtd_get_fields_raw(pfdtd_ctx *c, void *p_out, void *p_past_out, int64_t capacity_bytes) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (!c->d_P) return fail(PFDTD_E_STATE, "fields not prepared");
  if (capacity_bytes < (int64_t)c->field_bytes) return fail(PFDTD_E_ARG, "output buffer too small");
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (p_out) HIP_TRY(hipMemcpy(p_out, c->d_P, c->field_bytes, hipMemcpyDeviceToHost));
  if (p_past_out) HIP_TRY(hipMemcpy(p_past_out, c->d_Pp, c->field_bytes, hipMemcpyDeviceToHost));
  return PFDTD_OK;
}

int pfdtd_set_kernel_timing(pfdtd_ctx *c, int enabled) {
  if (!c) return fail(PFDTD_E_ARG, "null context");
  c->kernel_timing = enabled != 0;
  return PFDTD_OK;
}

int pfdtd_get_kernel_timing(pfdtd_ctx *c, double *update_ms_total, int64_t *update_launches) {
  if (!c || !update_ms_total || !update_launches) return fail(PFDTD_E_ARG, "null argument");
  *update_ms_total = c->upd_ms;
  *update_launches = c->upd_launches;
  return PFDTD_OK;
}

int pfdtd_set_update_tiling(pfdtd_ctx *c, int rows, int zchunk) {
  if (!c) return fail(PFDTD_E_ARG, "null context");
  if (rows == 0 && zchunk == 0) {  // back to the automatic tiling
    c->ry = 0;
    c->zc = 0;
    return PFDTD_OK;
  }
  if (rows != 1 && rows != 2 && rows != 4 && rows != 8)
    return fail(PFDTD_E_ARG, "rows per thread %d not in {1,2,4,8}", rows);
  if (zchunk < 1 || zchunk > 65535) return fail(PFDTD_E_ARG, "z-chunk %d outside 1..65535", zchunk);
  c->ry = rows;
  c->zc = zchunk;
  return PFDTD_OK;
}

int pfdtd_set_time_levels(pfdtd_ctx *c, int levels, int zchunk) {
  if (!c) return fail(PFDTD_E_ARG, "null context");
  if (levels < 0 || levels > 3)
    return fail(PFDTD_E_ARG, "time levels %d not in {0,1,2,3}", levels);
  if (zchunk < 0 || zchunk > 65535) return fail(PFDTD_E_ARG, "z-chunk %d outside 0..65535", zchunk);
  if ((levels == 1) != (c->levels == 1)) {
    HIP_TRY(hipSetDevice(c->device));
    HIP_TRY(hipStreamSynchronize(c->stream));
    free_fields(c);  // reallocated with the right number of buffers (2 or 4) at the next run
  }
  c->levels = levels;
  c->fzc = zchunk;
  return PFDTD_OK;
}

int pfdtd_set_sweep_shape(pfdtd_ctx *c, int shape) {
  if (!c) return fail(PFDTD_E_ARG, "null context");
  if (shape != PFDTD_SWEEP_AUTO && shape != PFDTD_SWEEP_COLUMN && shape != PFDTD_SWEEP_WIDE &&
      shape != PFDTD_SWEEP_PAIR)
    return fail(PFDTD_E_ARG, "sweep shape %d not AUTO (0), COLUMN (1), WIDE (2) or PAIR (3)", shape);
  c->fshape = shape;
  return PFDTD_OK;
}

// Re-derives the device element indices of sources and receivers (grid, batch or slots changed).
static int refresh_indices(pfdtd_ctx *c) {
  if (c->n_src) {
    std::vector<int64_t> idx(c->n_src);
    for (int s = 0; s < c->n_src; ++s) idx[s] = elem_idx(c, &c->src_xyz[3 * s], c->src_slot[s]);
    c->src_distinct = distinct_indices(idx);
    HIP_TRY(hipMemcpy(c->d_src_idx, idx.data(), sizeof(int64_t) * c->n_src, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_src_slot, c->src_slot.data(), sizeof(int32_t) * c->n_src,
                      hipMemcpyHostToDevice));
  }
  if (c->n_rec) {
    std::vector<int64_t> idx(c->n_rec);
    for (int r = 0; r < c->n_rec; ++r) idx[r] = elem_idx(c, &c->rec_xyz[3 * r], c->rec_slot[r]);
    HIP_TRY(hipMemcpy(c->d_rec_idx, idx.data(), sizeof(int64_t) * c->n_rec, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_rec_slot, c->rec_slot.data(), sizeof(int32_t) * c->n_rec,
                      hipMemcpyHostToDevice));
  }
  return PFDTD_OK;
}

int pfdtd_set_batch(pfdtd_ctx *c, int slots) {
  int rc = check_ctx(c);
  if (rc) return rc;
  int lg = 0;
  while ((1 << lg) < slots && lg < 9) ++lg;
  if (slots < 1 || slots > 256 || (1 << lg) != slots)
    return fail(PFDTD_E_ARG, "batch width %d is not a power of two in 1..256", slots);
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (lg != c->lgB) free_fields(c);
  c->lgB = lg;
  for (int32_t &s : c->src_slot) s = 0;
  for (int32_t &r : c->rec_slot) r = 0;
  return refresh_indices(c);
}

int pfdtd_set_batch_layout(pfdtd_ctx *c, int layout) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (layout != PFDTD_BATCH_INTERLEAVED && layout != PFDTD_BATCH_GROUPED)
    return fail(PFDTD_E_ARG, "batch layout %d not INTERLEAVED (0) or GROUPED (1)", layout);
  HIP_TRY(hipStreamSynchronize(c->stream));
  if (layout != c->grouped) free_fields(c);
  c->grouped = layout;
  return refresh_indices(c);
}

int pfdtd_set_source_slots(pfdtd_ctx *c, int n_src, const int32_t *slot) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n_src != c->n_src || (n_src && !slot)) return fail(PFDTD_E_ARG, "expected %d source slots", c->n_src);
  for (int s = 0; s < n_src; ++s)
    if (slot[s] < 0 || slot[s] >= (1 << c->lgB)) return fail(PFDTD_E_ARG, "source %d: slot %d outside the batch", s, slot[s]);
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->src_slot.assign(slot, slot + n_src);
  return refresh_indices(c);
}

int pfdtd_set_receiver_slots(pfdtd_ctx *c, int n_rec, const int32_t *slot) {
  int rc = check_ctx(c);
  if (rc) return rc;
  if (n_rec != c->n_rec || (n_rec && !slot)) return fail(PFDTD_E_ARG, "expected %d receiver slots", c->n_rec);
  for (int r = 0; r < n_rec; ++r)
    if (slot[r] < 0 || slot[r] >= (1 << c->lgB)) return fail(PFDTD_E_ARG, "receiver %d: slot %d outside the batch", r, slot[r]);
  HIP_TRY(hipStreamSynchronize(c->stream));
  c->rec_slot.assign(slot, slot + n_rec);
  return refresh_indices(c);
}

int pfdtd_set_poll_interval(pfdtd_ctx *c, int steps) {
  if (!c) return fail(PFDTD_E_ARG, "null context");
  if (steps < 1 || steps > 100000) return fail(PFDTD_E_ARG, "poll interval %d outside 1..100000", steps);
  c->poll_interval = steps;
  return PFDTD_OK;
}

int pfdtd_hbm_probe(int device, int kind, int64_t bytes_per_array, int iters,
                    double *ms_per_launch, double *bytes_moved_per_launch) {
  if (kind < PFDTD_PROBE_READ || kind > PFDTD_PROBE_UPDATE)
    return fail(PFDTD_E_ARG, "probe kind %d not in 0..4", kind);
  if (iters < 1 || !ms_per_launch || !bytes_moved_per_launch)
    return fail(PFDTD_E_ARG, "iterations < 1 or null output");
  const int64_t unit = (int64_t)sizeof(float4) * kProbeBlockElems;
  const int64_t blocks = bytes_per_array / unit;
  if (blocks < 1 || blocks > 0x7FFFFFFF)
    return fail(PFDTD_E_ARG, "array of %lld bytes outside 16 KiB .. 32 TiB", (long long)bytes_per_array);
  const size_t bytes = (size_t)(blocks * unit);
  const int arrays = kind == PFDTD_PROBE_TRIAD ? 3 : (kind <= PFDTD_PROBE_WRITE ? 1 : 2);
  const int moved = kind <= PFDTD_PROBE_WRITE ? 1 : (kind == PFDTD_PROBE_COPY ? 2 : 3);
  HIP_TRY(hipSetDevice(device));
  float4 *buf[3] = {nullptr, nullptr, nullptr};
  float *sink = nullptr;
  hipStream_t st = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  float ms = 0.f;
  hipError_t e = hipStreamCreate(&st);
  for (int i = 0; i < arrays && e == hipSuccess; ++i) {
    e = hipMalloc((void **)&buf[i], bytes);
    if (e == hipSuccess) e = hipMemsetAsync(buf[i], 0, bytes, st);
  }
  if (e == hipSuccess) e = hipMalloc((void **)&sink, 256 * sizeof(float));
  if (e == hipSuccess) e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e == hipSuccess) e = launch_hbm_probe(kind, buf[0], buf[1], buf[2], sink, blocks, st);
  if (e == hipSuccess) e = hipEventRecord(e0, st);
  for (int it = 0; it < iters && e == hipSuccess; ++it)
    e = launch_hbm_probe(kind, buf[0], buf[1], buf[2], sink, blocks, st);
  if (e == hipSuccess) e = hipEventRecord(e1, st);
  if (e == hipSuccess) e = hipEventSynchronize(e1);
  if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
  if (st) (void)hipStreamSynchronize(st);
  for (float4 *q : buf) dfree(q);
  dfree(sink);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  if (st) (void)hipStreamDestroy(st);
  if (e != hipSuccess) return fail(PFDTD_E_HIP, "hbm probe: %s", hipGetErrorString(e));
  *ms_per_launch = (double)ms / iters;
  *bytes_moved_per_launch = (double)bytes * moved;
  return PFDTD_OK;
}

}  // extern "C"
