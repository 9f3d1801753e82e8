// pfdtd_pool.hip -- serving pool: many independent impulse-response jobs (rooms of different
// sizes, FP32 / FP16 / BF16 storage) resident in one HBM arena and advanced together.
//
// Device side. Every resident job has a descriptor (PoolJobDev: its buffers, dimensions,
// coefficients, sources and receivers) and a state (step, current buffer, past buffer) in device
// memory, indexed by the job's slot. Jobs are advanced in rounds of two steps. A job whose
// sources avoid the never-updated first and last z slices owns four field buffers and is
// advanced by the two-level sweep (srl_update2_kernel in pool mode, pfdtd_fused.hip): ONE launch
// per storage type enumerates the sweep tiles of every such job (a block finds its job by a
// binary search over the launch table's tile prefix) and does two steps per job, or one for a
// job with a single step left; pool_inject2_kernel (one block per job) then gathers and injects
// the second step and rotates the buffer roles. Other jobs take two single steps per round:
// pool_update_kernel (flattened single-step tiles) and pool_inject_kernel, which gathers,
// injects the next samples and swaps current and past. Because the kernels read the state from
// device memory, a bundle of rounds is replay-invariant: it is captured once into a hipGraph per
// membership of the pool and replayed until a job joins or leaves. A job whose last step is
// done is skipped by the kernels until the host drops it from the launch.
//
// Host side. Jobs are admitted (pfdtd_pool_submit) only when their buffers fit in the arena's
// free space and a slot is free; the arena is one allocation made at pool creation, so
// admission and release never call hipMalloc / hipFree (which would synchronise the device and
// stall the jobs in flight). A job's set-up (label upload, classification, zeroing, first
// injection) runs on a separate upload stream while the run stream keeps stepping the resident
// jobs. Bundles are enqueued at most `depth` ahead; completion is observed through events.
#include "../../include/pfdtd.h"
#include "pfdtd_device.h"

#include <algorithm>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <deque>
#include <map>
#include <string>
#include <vector>

namespace pfdtd {

constexpr int kPoolTW = 64;  // tile width: rooms of any width waste fewer than 64 lanes per row

template <typename S, int RY>
__global__ __launch_bounds__(kTX) void pool_update_kernel(const PoolJobDev *__restrict__ jobs,
                                                          const PoolState *__restrict__ state,
                                                          PoolLaunch L, int zc) {
  const int bid = (int)blockIdx.x;
  int lo = 0, hi = L.n - 1;
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (L.tile_end[mid] > bid)
      hi = mid;
    else
      lo = mid + 1;
  }
  const int slot = L.slot[lo];
  const PoolJobDev &J = jobs[slot];
  const PoolState st = state[slot];
  if (st.step >= J.n_steps) return;
  const int t = bid - (lo ? L.tile_end[lo - 1] : 0);
  const int bx = t % J.gx;
  const int r = t / J.gx;
  const int by = r % J.gy;
  const int bz = r / J.gy;
  const S *P = (const S *)J.F[st.cur];
  S *Pp = (S *)J.F[st.past];
  // every thread of the block serves the same job, so the whole block reaches the barrier
  extern __shared__ __align__(16) unsigned char s_dyn[];
  CoefF *s_coef = reinterpret_cast<CoefF *>(s_dyn);
  stage_coef(s_coef, J.coef, (unsigned)J.ncol);
  srl_tile<S, RY, kPoolTW>(J.pos, J.mat, P, Pp, s_coef, (unsigned)J.ncol, J.air, J.air_uniform,
                           J.lam2, J.X, J.Y, J.Z, zc, bx, by, bz);
}

// Gather/inject arguments of job J at its current state.
inline __device__ InjectArgs pool_inject_args(const PoolJobDev &J, int gather_step,
                                              int inject_step) {
  InjectArgs a;
  a.rec_idx = J.rec_idx;
  a.n_rec = J.n_rec;
  a.resp = J.resp;
  a.resp_steps = J.n_steps;
  a.gather_step = J.n_rec > 0 ? gather_step : -1;
  a.src_idx = J.src_idx;
  a.src_kind = J.src_kind;
  a.n_src = J.n_src;
  a.sig_f = J.sig;
  a.sig_d = nullptr;
  a.n_samples = J.n_samples;
  a.inject_step = (J.n_src > 0 && inject_step < J.n_steps) ? inject_step : -1;
  a.soft_overwrite = J.soft_overwrite;
  a.scale_exp = J.scale_exp;
  a.distinct = J.distinct;
  return a;
}

// After a single step: gather step n on P^{n+1} (written over the past buffer), inject step
// n+1 into it, then P^{n+1} becomes the current field.
template <typename S>
__global__ __launch_bounds__(256) void pool_inject_kernel(const PoolJobDev *__restrict__ jobs,
                                                          PoolState *__restrict__ state,
                                                          PoolLaunch L) {
  const int slot = L.slot[blockIdx.x];
  const PoolJobDev &J = jobs[slot];
  const PoolState st = state[slot];
  if (st.step >= J.n_steps) return;
  inject_gather_block<S>((S *)J.F[st.past], pool_inject_args(J, st.step, st.step + 1));
  __syncthreads();
  if (threadIdx.x == 0) state[slot] = PoolState{st.step + 1, st.past, st.cur};
}

// After a two-level sweep (pool_update2 in pfdtd_fused.hip, which gathered step n and
// injected step n+1 itself): with two steps done, gather step n+1 on P^{n+2} and inject step
// n+2 into it; P^{n+2} and P^{n+1} (the spare buffers o2, o1) become current and past. With
// one step done (the job's last), P^{n+1} in o1 becomes current.
template <typename S>
__global__ __launch_bounds__(256) void pool_inject2_kernel(const PoolJobDev *__restrict__ jobs,
                                                           PoolState *__restrict__ state,
                                                           PoolLaunch L) {
  const int slot = L.slot[blockIdx.x];
  const PoolJobDev &J = jobs[slot];
  const PoolState st = state[slot];
  const int remaining = J.n_steps - st.step;
  if (remaining <= 0) return;
  int o1, o2;
  pool_spare(st.cur, st.past, o1, o2);
  if (remaining >= 2) {
    inject_gather_block<S>((S *)J.F[o2], pool_inject_args(J, st.step + 1, st.step + 2));
    __syncthreads();
    if (threadIdx.x == 0) state[slot] = PoolState{st.step + 2, o2, o1};
  } else if (threadIdx.x == 0) {
    state[slot] = PoolState{st.step + 1, o1, st.cur};
  }
}

template <typename S>
hipError_t launch_pool_update_t(int ry, int tiles, const PoolJobDev *jobs, const PoolState *st,
                                const PoolLaunch &L, int zc, hipStream_t stream) {
  const size_t lds = sizeof(CoefF) * kStageRows * (size_t)L.ncol_max;
  switch (ry) {
  case 1:
    hipLaunchKernelGGL((pool_update_kernel<S, 1>), dim3(tiles), dim3(kTX), lds, stream, jobs, st, L, zc);
    break;
  case 2:
    hipLaunchKernelGGL((pool_update_kernel<S, 2>), dim3(tiles), dim3(kTX), lds, stream, jobs, st, L, zc);
    break;
  case 4:
    hipLaunchKernelGGL((pool_update_kernel<S, 4>), dim3(tiles), dim3(kTX), lds, stream, jobs, st, L, zc);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pool_update(int precision, int ry, int tiles, const PoolJobDev *jobs,
                              const PoolState *st, const PoolLaunch &L, int zc,
                              hipStream_t stream) {
  switch (precision) {
  case PFDTD_F32: return launch_pool_update_t<float>(ry, tiles, jobs, st, L, zc, stream);
  case PFDTD_F16: return launch_pool_update_t<h16>(ry, tiles, jobs, st, L, zc, stream);
  case PFDTD_BF16: return launch_pool_update_t<b16>(ry, tiles, jobs, st, L, zc, stream);
  default: return hipErrorInvalidValue;
  }
}

hipError_t launch_pool_inject2(int precision, const PoolJobDev *jobs, PoolState *st,
                               const PoolLaunch &L, hipStream_t stream) {
  const dim3 grid(L.n), block(256);
  switch (precision) {
  case PFDTD_F32: hipLaunchKernelGGL(pool_inject2_kernel<float>, grid, block, 0, stream, jobs, st, L); break;
  case PFDTD_F16: hipLaunchKernelGGL(pool_inject2_kernel<h16>, grid, block, 0, stream, jobs, st, L); break;
  case PFDTD_BF16: hipLaunchKernelGGL(pool_inject2_kernel<b16>, grid, block, 0, stream, jobs, st, L); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_pool_inject(int precision, const PoolJobDev *jobs, PoolState *st,
                              const PoolLaunch &L, hipStream_t stream) {
  const dim3 grid(L.n), block(256);
  switch (precision) {
  case PFDTD_F32: hipLaunchKernelGGL(pool_inject_kernel<float>, grid, block, 0, stream, jobs, st, L); break;
  case PFDTD_F16: hipLaunchKernelGGL(pool_inject_kernel<h16>, grid, block, 0, stream, jobs, st, L); break;
  case PFDTD_BF16: hipLaunchKernelGGL(pool_inject_kernel<b16>, grid, block, 0, stream, jobs, st, L); break;
  default: return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

}  // namespace pfdtd

using namespace pfdtd;

namespace {

int pfail(int code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  record_error(buf);
  return code;
}

#define POOL_TRY(expr)                                                                       \
  do {                                                                                       \
    hipError_t _e = (expr);                                                                  \
    if (_e != hipSuccess)                                                                    \
      return pfail(PFDTD_E_HIP, "%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e),       \
                   __FILE__, __LINE__);                                                      \
  } while (0)

constexpr size_t kAlign = 256;
constexpr int kCoefNumP = 20;

size_t align_up(size_t v) { return (v + kAlign - 1) & ~(kAlign - 1); }

size_t storage_bytes(int precision) {
  switch (precision) {
  case PFDTD_F32: return 4;
  case PFDTD_F16:
  case PFDTD_BF16: return 2;
  default: return 0;
  }
}

// z-chunk of the two-level sweep in the pool: the C5 capacity probe (200 rooms of 64..512 per
// side at once) sustains 479 Gvox/s with 16 slices, 480 with 8 and 460 with 32
constexpr int kPoolFusedZc = 16;

// Byte offsets of one job's buffers inside its arena block (a two-level job adds two field
// buffers, the packed byte plane, its air-row bits and the source/receiver coordinates).
struct JobLayout {
  size_t F[4], pos, mat, pm, air_rows, labels, coef, src_idx, src_xyz, src_kind, sig, rec_idx,
      rec_xyz, resp, total;
};

JobLayout layout_job(int X, int Y, int Z, size_t es, int n_src, int n_samples, int n_rec,
                     int n_steps, bool fused) {
  const int64_t N = (int64_t)X * Y * Z;
  JobLayout L{};
  size_t o = 0;
  auto take = [&](size_t bytes) {
    const size_t at = o;
    o += align_up(bytes ? bytes : 1);
    return at;
  };
  for (int b = 0; b < 4; ++b) L.F[b] = (b < 2 || fused) ? take(es * (size_t)N) : 0;
  L.pos = take((size_t)N);
  L.mat = take((size_t)N);
  L.pm = fused ? take(sizeof(uint16_t) * (size_t)N) : 0;
  L.air_rows = fused ? take(sizeof(uint32_t) * air_rows_words(X, Y, Z)) : 0;
  L.labels = take((size_t)N);
  L.coef = take(sizeof(CoefF) * kCoefEntries);
  L.src_idx = take(sizeof(int64_t) * (size_t)n_src);
  L.src_xyz = fused ? take(sizeof(int32_t) * 3 * (size_t)n_src) : 0;
  L.src_kind = take(sizeof(int32_t) * (size_t)n_src);
  L.sig = take(sizeof(float) * (size_t)n_src * (size_t)n_samples);
  L.rec_idx = take(sizeof(int64_t) * (size_t)n_rec);
  L.rec_xyz = fused ? take(sizeof(int32_t) * 3 * (size_t)n_rec) : 0;
  L.resp = take(sizeof(double) * (size_t)n_rec * (size_t)n_steps);
  L.total = o;
  return L;
}

enum JobPhase { kActive = 0, kFinishing = 1, kFinished = 2 };

struct PoolJobHost {
  int64_t id = 0;
  int slot = -1;
  int precision = 0;
  int n_steps = 0, n_rec = 0;
  int64_t voxels = 0;
  int tiles = 0;  // tiles of the job's update kernel (two-level or single-step)
  bool fused = false;
  int ncol = 1;
  size_t off = 0, bytes = 0;
  int64_t steps_enqueued = 0;
  uint64_t final_seq = 0;
  JobPhase phase = kActive;
  bool reported = false;
  double *d_resp = nullptr;
};

}  // namespace

struct pfdtd_pool {
  int device = 0;
  hipStream_t run = nullptr, upload = nullptr;
  char *arena = nullptr;
  size_t arena_bytes = 0;
  std::map<size_t, size_t> free_blocks;  // offset -> size
  PoolJobDev *d_jobs = nullptr;
  PoolState *d_state = nullptr;
  std::vector<int64_t> slot_job;  // job id per slot, -1 free
  std::map<int64_t, PoolJobHost> jobs;
  int64_t next_id = 1;

  // options
  int bundle = 32;
  int use_graph = 1;
  int ry = 2;
  int zc = 8;
  int max_jobs = kPoolMaxJobs;
  int depth = 2;

  // launch membership
  bool dirty = true;
  std::vector<int64_t> members;  // job ids in the current launch
  int prec_list[3] = {PFDTD_F32, PFDTD_F16, PFDTD_BF16};
  PoolLaunch launch[3][2];  // [storage type][0 single-step jobs, 1 two-level jobs]
  int tiles[3][2] = {{0, 0}, {0, 0}, {0, 0}};
  hipGraphExec_t exec = nullptr;
  uint64_t exec_last_seq = 0;
  std::vector<std::pair<uint64_t, hipGraphExec_t>> retired;

  // bundles in flight
  uint64_t seq = 0;            // bundles enqueued
  uint64_t completed_seq = 0;  // every bundle <= this has completed
  std::deque<std::pair<uint64_t, hipEvent_t>> inflight;
  std::vector<hipEvent_t> spare_events;
  std::deque<int64_t> done_queue;

  // statistics
  int64_t graph_builds = 0;
  double voxel_steps = 0.0;
};

namespace {

int check_pool(pfdtd_pool *p) {
  if (!p) return pfail(PFDTD_E_ARG, "null pool");
  hipError_t e = hipSetDevice(p->device);
  if (e != hipSuccess) return pfail(PFDTD_E_HIP, "hipSetDevice(%d): %s", p->device, hipGetErrorString(e));
  return PFDTD_OK;
}

bool arena_alloc(pfdtd_pool *p, size_t bytes, size_t *off) {
  for (auto it = p->free_blocks.begin(); it != p->free_blocks.end(); ++it) {
    if (it->second >= bytes) {
      *off = it->first;
      const size_t rest = it->second - bytes;
      const size_t at = it->first + bytes;
      p->free_blocks.erase(it);
      if (rest) p->free_blocks[at] = rest;
      return true;
    }
  }
  return false;
}

void arena_free(pfdtd_pool *p, size_t off, size_t bytes) {
  auto it = p->free_blocks.emplace(off, bytes).first;
  auto next = std::next(it);
  if (next != p->free_blocks.end() && it->first + it->second == next->first) {
    it->second += next->second;
    p->free_blocks.erase(next);
  }
  if (it != p->free_blocks.begin()) {
    auto prev = std::prev(it);
    if (prev->first + prev->second == it->first) {
      prev->second += it->second;
      p->free_blocks.erase(it);
    }
  }
}

size_t arena_free_total(const pfdtd_pool *p, size_t *largest) {
  size_t tot = 0, big = 0;
  for (const auto &kv : p->free_blocks) {
    tot += kv.second;
    big = std::max(big, kv.second);
  }
  if (largest) *largest = big;
  return tot;
}

// Marks finished every job whose last bundle is covered by `completed_seq`; retires graphs.
void note_completed(pfdtd_pool *p) {
  for (auto &kv : p->jobs) {
    PoolJobHost &j = kv.second;
    if (j.phase == kFinishing && j.final_seq <= p->completed_seq) {
      j.phase = kFinished;
      p->done_queue.push_back(j.id);
    }
  }
  for (size_t i = 0; i < p->retired.size();) {
    if (p->retired[i].first <= p->completed_seq) {
      (void)hipGraphExecDestroy(p->retired[i].second);
      p->retired[i] = p->retired.back();
      p->retired.pop_back();
    } else {
      ++i;
    }
  }
}

// Non-blocking: pops every completed bundle event (`block` waits for the oldest first).
int reap(pfdtd_pool *p, bool block) {
  while (!p->inflight.empty()) {
    auto &front = p->inflight.front();
    hipError_t e = block ? hipEventSynchronize(front.second) : hipEventQuery(front.second);
    if (e == hipErrorNotReady) break;
    if (e != hipSuccess) return pfail(PFDTD_E_HIP, "bundle event: %s", hipGetErrorString(e));
    p->completed_seq = front.first;
    p->spare_events.push_back(front.second);
    p->inflight.pop_front();
    block = false;
  }
  note_completed(p);
  return PFDTD_OK;
}

// A bundle is bundle/2 rounds; each round advances two-level jobs by one sweep (two steps, or
// the last one) and single-step jobs by two steps.
int enqueue_bundle_kernels(pfdtd_pool *p) {
  for (int k = 0; k < p->bundle; k += 2) {
    for (int q = 0; q < 3; ++q)
      if (p->launch[q][1].n) {
        POOL_TRY(launch_pool_update2(p->prec_list[q], p->tiles[q][1], p->d_jobs, p->d_state,
                                     p->launch[q][1], kPoolFusedZc, p->run));
        POOL_TRY(launch_pool_inject2(p->prec_list[q], p->d_jobs, p->d_state, p->launch[q][1],
                                     p->run));
      }
    for (int half = 0; half < 2; ++half)
      for (int q = 0; q < 3; ++q)
        if (p->launch[q][0].n) {
          POOL_TRY(launch_pool_update(p->prec_list[q], p->ry, p->tiles[q][0], p->d_jobs,
                                      p->d_state, p->launch[q][0], p->zc, p->run));
          POOL_TRY(launch_pool_inject(p->prec_list[q], p->d_jobs, p->d_state, p->launch[q][0],
                                      p->run));
        }
  }
  return PFDTD_OK;
}

int rebuild_membership(pfdtd_pool *p) {
  p->members.clear();
  for (int q = 0; q < 3; ++q)
    for (int f = 0; f < 2; ++f) {
      p->launch[q][f].n = 0;
      p->launch[q][f].ncol_max = 1;
      p->tiles[q][f] = 0;
    }
  for (auto &kv : p->jobs) {
    PoolJobHost &j = kv.second;
    if (j.phase != kActive) continue;
    const int q = j.precision == PFDTD_F32 ? 0 : (j.precision == PFDTD_F16 ? 1 : 2);
    const int f = j.fused ? 1 : 0;
    PoolLaunch &L = p->launch[q][f];
    L.slot[L.n] = j.slot;
    L.ncol_max = std::max(L.ncol_max, j.ncol);
    p->tiles[q][f] += j.tiles;
    L.tile_end[L.n] = p->tiles[q][f];
    ++L.n;
    p->members.push_back(j.id);
  }
  if (p->exec) {
    p->retired.emplace_back(p->exec_last_seq, p->exec);
    p->exec = nullptr;
  }
  if (p->use_graph && !p->members.empty()) {
    hipGraph_t graph = nullptr;
    POOL_TRY(hipStreamBeginCapture(p->run, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_bundle_kernels(p);
    hipError_t e = hipStreamEndCapture(p->run, &graph);
    if (rc) {
      if (graph) (void)hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return pfail(PFDTD_E_HIP, "graph capture: %s", hipGetErrorString(e));
    e = hipGraphInstantiateWithFlags(&p->exec, graph, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return pfail(PFDTD_E_HIP, "graph instantiate: %s", hipGetErrorString(e));
    ++p->graph_builds;
  }
  p->dirty = false;
  return PFDTD_OK;
}

// Host-side preconditions of a job: the update needs an outside outer layer and every solid
// label must name a material of the table.
int check_job(const pfdtd_job *j) {
  if (!j) return pfail(PFDTD_E_ARG, "null job");
  if (!j->labels) return pfail(PFDTD_E_ARG, "null labels");
  if (j->X < 3 || j->Y < 3 || j->Z < 3) return pfail(PFDTD_E_ARG, "grid %dx%dx%d below 3x3x3", j->X, j->Y, j->Z);
  if ((int64_t)j->X * j->Y * j->Z > ((int64_t)1 << 34)) return pfail(PFDTD_E_ARG, "grid too large");
  if (!storage_bytes(j->precision))
    return pfail(PFDTD_E_ARG, "pool precision %d: F32, F16 or BF16 only", j->precision);
  if (j->precision == PFDTD_F32 && j->scale_exp != 0) return pfail(PFDTD_E_ARG, "FP32 fields take no scale");
  if (j->scale_exp < -64 || j->scale_exp > 64) return pfail(PFDTD_E_ARG, "scale exponent out of range");
  if (j->n_steps < 1) return pfail(PFDTD_E_ARG, "n_steps must be positive");
  if (!j->materials || j->n_mat < 1 || j->n_mat > 255 || j->n_coef < 1 || j->n_coef > kCoefNumP)
    return pfail(PFDTD_E_ARG, "bad material table");
  if (j->octave >= (uint32_t)kCoefNumP) return pfail(PFDTD_E_ARG, "octave %u >= 20", j->octave);
  if (!(j->lambda > 0.0)) return pfail(PFDTD_E_ARG, "bad lambda");
  if (j->flags & ~(PFDTD_MATSLOT_LITERAL | PFDTD_SOFT_OVERWRITE)) return pfail(PFDTD_E_ARG, "unknown flags");
  if (j->n_src < 0 || j->n_rec < 0 || j->n_samples < 0) return pfail(PFDTD_E_ARG, "negative count");
  if (j->n_src > 0 && (!j->src_xyz || !j->src_kind || (j->n_samples > 0 && !j->signals)))
    return pfail(PFDTD_E_ARG, "null source arrays");
  if (j->n_rec > 0 && !j->rec_xyz) return pfail(PFDTD_E_ARG, "null receiver array");
  for (int s = 0; s < j->n_src; ++s)
    if (j->src_kind[s] < PFDTD_SRC_HARD || j->src_kind[s] > PFDTD_SRC_TRANSPARENT)
      return pfail(PFDTD_E_ARG, "source %d: unknown kind %d", s, j->src_kind[s]);
  const int64_t X = j->X, Y = j->Y, Z = j->Z, XY = X * Y;
  const uint8_t *lab = j->labels;
  for (int64_t z = 0; z < Z; ++z)
    for (int64_t y = 0; y < Y; ++y) {
      const uint8_t *row = lab + z * XY + y * X;
      if (z == 0 || z == Z - 1 || y == 0 || y == Y - 1) {
        for (int64_t x = 0; x < X; ++x)
          if (!row[x]) return pfail(PFDTD_E_DOMAIN, "inside voxel on the outer layer at (%lld,%lld,%lld)", (long long)x, (long long)y, (long long)z);
      } else if (!row[0] || !row[X - 1]) {
        return pfail(PFDTD_E_DOMAIN, "inside voxel on the outer layer in row (y=%lld, z=%lld)", (long long)y, (long long)z);
      }
    }
  uint8_t mx = 0;
  const int64_t N = XY * Z;
  for (int64_t i = 0; i < N; ++i) mx = lab[i] > mx ? lab[i] : mx;
  if (mx > j->n_mat) return pfail(PFDTD_E_DOMAIN, "label %u names material %u of %d", mx, mx - 1u, j->n_mat);
  return PFDTD_OK;
}

// Whether a job is advanced by the two-level sweep: not when a source sits on the first or last
// z slice (never updated, but read by its neighbours: the single-step order is needed there)
// or when one z slice exceeds the sweep's 32-bit buffer range.
bool job_fused(const pfdtd_job *j) {
  const size_t es = storage_bytes(j->precision);
  if ((uint64_t)j->X * (uint64_t)j->Y * es >= (1ull << 32)) return false;
  for (int s = 0; s < j->n_src; ++s) {
    const int32_t *q = &j->src_xyz[3 * s];
    const bool in = q[0] >= 0 && q[1] >= 0 && q[2] >= 0 && q[0] < j->X && q[1] < j->Y && q[2] < j->Z;
    if (in && (q[2] == 0 || q[2] == j->Z - 1)) return false;
  }
  return true;
}

int64_t grid_index(const int32_t *q, int X, int Y, int Z) {
  if (q[0] < 0 || q[1] < 0 || q[2] < 0 || q[0] >= X || q[1] >= Y || q[2] >= Z) return -1;
  return ((int64_t)q[2] * Y + q[1]) * X + q[0];
}

}  // namespace

extern "C" {

int pfdtd_pool_create(int device, int64_t arena_bytes, pfdtd_pool **out) {
  if (!out) return pfail(PFDTD_E_ARG, "null output");
  *out = nullptr;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || device < 0 || device >= n)
    return pfail(PFDTD_E_ARG, "device %d not present", device);
  POOL_TRY(hipSetDevice(device));
  size_t fr = 0, tot = 0;
  POOL_TRY(hipMemGetInfo(&fr, &tot));
  size_t want = arena_bytes > 0 ? (size_t)arena_bytes : (fr / 16) * 14;  // default: 87.5 % of free
  want = align_up(want);
  if (want > fr) return pfail(PFDTD_E_NOMEM, "arena of %zu bytes exceeds free HBM (%zu)", want, fr);
  pfdtd_pool *p = new pfdtd_pool();
  p->device = device;
  p->slot_job.assign(kPoolMaxJobs, -1);
  hipError_t e = hipMalloc((void **)&p->arena, want);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_jobs, sizeof(PoolJobDev) * kPoolMaxJobs);
  if (e == hipSuccess) e = hipMalloc((void **)&p->d_state, sizeof(PoolState) * kPoolMaxJobs);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->run, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipStreamCreateWithFlags(&p->upload, hipStreamNonBlocking);
  if (e != hipSuccess) {
    pfdtd_pool_destroy(p);
    return pfail(PFDTD_E_NOMEM, "pool allocation: %s", hipGetErrorString(e));
  }
  p->arena_bytes = want;
  p->free_blocks[0] = want;
  *out = p;
  return PFDTD_OK;
}

int pfdtd_pool_destroy(pfdtd_pool *p) {
  if (!p) return PFDTD_OK;
  (void)hipSetDevice(p->device);
  if (p->run) (void)hipStreamSynchronize(p->run);
  if (p->upload) (void)hipStreamSynchronize(p->upload);
  if (p->exec) (void)hipGraphExecDestroy(p->exec);
  for (auto &r : p->retired) (void)hipGraphExecDestroy(r.second);
  for (auto &f : p->inflight) (void)hipEventDestroy(f.second);
  for (auto ev : p->spare_events) (void)hipEventDestroy(ev);
  if (p->arena) (void)hipFree(p->arena);
  if (p->d_jobs) (void)hipFree(p->d_jobs);
  if (p->d_state) (void)hipFree(p->d_state);
  if (p->run) (void)hipStreamDestroy(p->run);
  if (p->upload) (void)hipStreamDestroy(p->upload);
  delete p;
  return PFDTD_OK;
}

int pfdtd_pool_set_options(pfdtd_pool *p, int bundle_steps, int use_graph, int rows, int zchunk,
                           int max_jobs) {
  int rc = check_pool(p);
  if (rc) return rc;
  if (bundle_steps < 1 || bundle_steps > 4096) return pfail(PFDTD_E_ARG, "bundle_steps outside 1..4096");
  if (rows != 1 && rows != 2 && rows != 4) return pfail(PFDTD_E_ARG, "rows per thread %d not in {1,2,4}", rows);
  if (zchunk < 1) return pfail(PFDTD_E_ARG, "zchunk must be positive");
  if (max_jobs < 1 || max_jobs > kPoolMaxJobs) return pfail(PFDTD_E_ARG, "max_jobs outside 1..%d", kPoolMaxJobs);
  // tile geometry of resident jobs depends on rows / zchunk: change only on an empty pool
  if ((rows != p->ry || zchunk != p->zc) && !p->jobs.empty())
    return pfail(PFDTD_E_STATE, "tiling can change only while the pool is empty");
  p->bundle = bundle_steps + (bundle_steps & 1);  // whole rounds of two steps
  p->use_graph = use_graph ? 1 : 0;
  p->ry = rows;
  p->zc = zchunk;
  p->max_jobs = max_jobs;
  p->dirty = true;
  return PFDTD_OK;
}

int pfdtd_pool_job_bytes(const pfdtd_job *j, int64_t *bytes) {
  if (!j || !bytes) return pfail(PFDTD_E_ARG, "null argument");
  const size_t es = storage_bytes(j->precision);
  if (!es || j->X < 1 || j->Y < 1 || j->Z < 1 || j->n_steps < 0 || j->n_src < 0 || j->n_rec < 0)
    return pfail(PFDTD_E_ARG, "bad job dimensions or precision");
  *bytes = (int64_t)layout_job(j->X, j->Y, j->Z, es, j->n_src, j->n_samples, j->n_rec, j->n_steps,
                              j->n_src == 0 || j->src_xyz ? job_fused(j) : false)
               .total;
  return PFDTD_OK;
}

int pfdtd_pool_mem(pfdtd_pool *p, int64_t *arena_bytes, int64_t *free_bytes, int64_t *largest_free) {
  if (!p) return pfail(PFDTD_E_ARG, "null pool");
  size_t big = 0;
  const size_t fr = arena_free_total(p, &big);
  if (arena_bytes) *arena_bytes = (int64_t)p->arena_bytes;
  if (free_bytes) *free_bytes = (int64_t)fr;
  if (largest_free) *largest_free = (int64_t)big;
  return PFDTD_OK;
}

int pfdtd_pool_submit(pfdtd_pool *p, const pfdtd_job *j, int64_t *job_id) {
  int rc = check_pool(p);
  if (rc) return rc;
  if (!job_id) return pfail(PFDTD_E_ARG, "null job id output");
  rc = check_job(j);
  if (rc) return rc;
  const int resident = (int)p->jobs.size();
  if (resident >= p->max_jobs) return pfail(PFDTD_E_NOMEM, "all %d job slots in use", p->max_jobs);
  int slot = -1;
  for (int s = 0; s < kPoolMaxJobs; ++s)
    if (p->slot_job[s] < 0) {
      slot = s;
      break;
    }
  if (slot < 0) return pfail(PFDTD_E_NOMEM, "no free job slot");
  const int X = j->X, Y = j->Y, Z = j->Z;
  const int64_t N = (int64_t)X * Y * Z;
  const size_t es = storage_bytes(j->precision);
  const bool fused = job_fused(j);
  const JobLayout L = layout_job(X, Y, Z, es, j->n_src, j->n_samples, j->n_rec, j->n_steps, fused);
  size_t off = 0;
  if (!arena_alloc(p, L.total, &off)) {
    size_t big = 0;
    arena_free_total(p, &big);
    return pfail(PFDTD_E_NOMEM, "job needs %zu bytes, largest free block %zu", L.total, big);
  }
  char *base = p->arena + off;

  // host-side tables
  std::vector<float> table((size_t)j->n_mat * kCoefNumP, 0.f);
  for (int m = 0; m < j->n_mat; ++m)
    for (int c = 0; c < j->n_coef; ++c) table[(size_t)m * kCoefNumP + c] = j->materials[(size_t)m * j->n_coef + c];
  std::vector<CoefF> cf(kCoefEntries);
  std::vector<CoefD> cd(kCoefEntries);
  build_coef_tables(table.data(), j->n_mat, j->octave, j->lambda,
                    (j->flags & PFDTD_MATSLOT_LITERAL) != 0, cf.data(), cd.data());
  int air_uniform = 1;
  for (int m = 0; m < 256; ++m)
    if (std::memcmp(&cf[6 * 256 + m], &cf[6 * 256], sizeof(CoefF)) != 0) air_uniform = 0;
  std::vector<int64_t> sidx(j->n_src), ridx(j->n_rec);
  for (int s = 0; s < j->n_src; ++s) sidx[s] = grid_index(&j->src_xyz[3 * s], X, Y, Z);
  for (int r = 0; r < j->n_rec; ++r) ridx[r] = grid_index(&j->rec_xyz[3 * r], X, Y, Z);

  PoolJobDev d{};
  d.pos = (const uint8_t *)(base + L.pos);
  d.mat = (const uint8_t *)(base + L.mat);
  d.pm = fused ? (const uint16_t *)(base + L.pm) : nullptr;
  d.air_rows = fused ? (const uint32_t *)(base + L.air_rows) : nullptr;
  for (int b = 0; b < 4; ++b) d.F[b] = (b < 2 || fused) ? base + L.F[b] : nullptr;
  d.coef = (const CoefF *)(base + L.coef);
  d.ncol = j->n_mat;  // labels were checked to name materials 0..n_mat-1 only
  d.air = cf[6 * 256];
  d.air_uniform = air_uniform;
  d.lam2 = (float)(j->lambda * j->lambda);
  d.X = X;
  d.Y = Y;
  d.Z = Z;
  const int rows_per_tile = p->ry * (kTX / kPoolTW);
  d.gx = (X + kPoolTW - 1) / kPoolTW;
  d.gy = (Y + rows_per_tile - 1) / rows_per_tile;
  const int gz = (Z - 2 + p->zc - 1) / p->zc;
  const FusedGeom fg = plan_fused(X, Y, Z, kPoolFusedZc, j->precision == PFDTD_F32 ? 0 : 2, true);
  d.fx = (int)fg.grid.x;
  d.fy = (int)fg.grid.y;
  d.fused = fused ? 1 : 0;
  const int64_t tiles = fused ? (int64_t)fg.grid.x * fg.grid.y * fg.grid.z
                              : (int64_t)d.gx * d.gy * gz;
  if (tiles > (int64_t)1 << 30) {
    arena_free(p, off, L.total);
    return pfail(PFDTD_E_ARG, "job grid too large for one launch");
  }
  d.rec_idx = (const int64_t *)(base + L.rec_idx);
  d.rec_xyz = fused ? (const int32_t *)(base + L.rec_xyz) : nullptr;
  d.resp = (double *)(base + L.resp);
  d.src_idx = (const int64_t *)(base + L.src_idx);
  d.src_xyz = fused ? (const int32_t *)(base + L.src_xyz) : nullptr;
  d.src_kind = (const int32_t *)(base + L.src_kind);
  d.sig = (const float *)(base + L.sig);
  d.n_rec = j->n_rec;
  d.n_src = j->n_src;
  d.n_samples = j->n_samples;
  d.n_steps = j->n_steps;
  d.soft_overwrite = (j->flags & PFDTD_SOFT_OVERWRITE) ? 1 : 0;
  d.scale_exp = j->scale_exp;
  {
    std::vector<int64_t> v;
    for (int64_t i : sidx)
      if (i >= 0) v.push_back(i);
    std::sort(v.begin(), v.end());
    d.distinct = std::adjacent_find(v.begin(), v.end()) == v.end() ? 1 : 0;
  }
  const PoolState st0{0, 0, 1};

  // set-up on the upload stream, concurrent with the resident jobs' steps
  hipStream_t u = p->upload;
  hipError_t e = hipMemcpyAsync(base + L.labels, j->labels, (size_t)N, hipMemcpyHostToDevice, u);
  if (e == hipSuccess)
    e = launch_classify_labels((const uint8_t *)(base + L.labels), (uint8_t *)(base + L.pos),
                               (uint8_t *)(base + L.mat), X, Y, Z, u);
  for (int b = 0; b < 4 && e == hipSuccess; ++b)
    if (d.F[b]) e = hipMemsetAsync(d.F[b], 0, es * (size_t)N, u);
  if (e == hipSuccess && fused)
    e = launch_pack_pm(d.pos, d.mat, (uint16_t *)(base + L.pm), N, u);
  if (e == hipSuccess && fused)
    e = launch_air_rows(d.pm, (uint32_t *)(base + L.air_rows), X, Y, Z, 0, false, u);
  if (e == hipSuccess && fused && j->n_src)
    e = hipMemcpyAsync(base + L.src_xyz, j->src_xyz, sizeof(int32_t) * 3 * j->n_src, hipMemcpyHostToDevice, u);
  if (e == hipSuccess && fused && j->n_rec)
    e = hipMemcpyAsync(base + L.rec_xyz, j->rec_xyz, sizeof(int32_t) * 3 * j->n_rec, hipMemcpyHostToDevice, u);
  if (e == hipSuccess && j->n_rec)
    e = hipMemsetAsync(base + L.resp, 0, sizeof(double) * (size_t)j->n_rec * j->n_steps, u);
  if (e == hipSuccess) e = hipMemcpyAsync(base + L.coef, cf.data(), sizeof(CoefF) * kCoefEntries, hipMemcpyHostToDevice, u);
  if (e == hipSuccess && j->n_src) {
    e = hipMemcpyAsync(base + L.src_idx, sidx.data(), sizeof(int64_t) * j->n_src, hipMemcpyHostToDevice, u);
    if (e == hipSuccess)
      e = hipMemcpyAsync(base + L.src_kind, j->src_kind, sizeof(int32_t) * j->n_src, hipMemcpyHostToDevice, u);
    if (e == hipSuccess && j->n_samples)
      e = hipMemcpyAsync(base + L.sig, j->signals, sizeof(float) * (size_t)j->n_src * j->n_samples,
                         hipMemcpyHostToDevice, u);
  }
  if (e == hipSuccess && j->n_rec)
    e = hipMemcpyAsync(base + L.rec_idx, ridx.data(), sizeof(int64_t) * j->n_rec, hipMemcpyHostToDevice, u);
  if (e == hipSuccess && j->n_src) {
    // step 0's samples go into the (zero) current field before the first update
    InjectArgs a{};
    a.rec_idx = d.rec_idx;
    a.n_rec = 0;
    a.resp = d.resp;
    a.resp_steps = j->n_steps;
    a.gather_step = -1;
    a.src_idx = d.src_idx;
    a.src_kind = d.src_kind;
    a.n_src = j->n_src;
    a.sig_f = d.sig;
    a.sig_d = nullptr;
    a.n_samples = j->n_samples;
    a.inject_step = 0;
    a.soft_overwrite = d.soft_overwrite;
    a.scale_exp = j->scale_exp;
    a.distinct = d.distinct;
    e = launch_inject_gather(j->precision, d.F[0], a, u);
  }
  if (e == hipSuccess) e = hipMemcpyAsync(p->d_jobs + slot, &d, sizeof(d), hipMemcpyHostToDevice, u);
  if (e == hipSuccess) e = hipMemcpyAsync(p->d_state + slot, &st0, sizeof(st0), hipMemcpyHostToDevice, u);
  // the host arrays above are read by the copies: wait for this job's set-up only
  const hipError_t e2 = hipStreamSynchronize(u);
  if (e == hipSuccess) e = e2;
  if (e != hipSuccess) {
    arena_free(p, off, L.total);
    return pfail(PFDTD_E_HIP, "job set-up: %s", hipGetErrorString(e));
  }

  PoolJobHost h;
  h.id = p->next_id++;
  h.slot = slot;
  h.precision = j->precision;
  h.fused = fused;
  h.ncol = j->n_mat;
  h.n_steps = j->n_steps;
  h.n_rec = j->n_rec;
  h.voxels = N;
  h.tiles = (int)tiles;
  h.off = off;
  h.bytes = L.total;
  h.d_resp = d.resp;
  p->jobs[h.id] = h;
  p->slot_job[slot] = h.id;
  p->dirty = true;
  *job_id = h.id;
  return PFDTD_OK;
}

int pfdtd_pool_advance(pfdtd_pool *p, int *n_active) {
  int rc = check_pool(p);
  if (rc) return rc;
  rc = reap(p, false);
  if (rc) return rc;
  if (p->dirty) {
    rc = rebuild_membership(p);
    if (rc) return rc;
  }
  if (n_active) *n_active = (int)p->members.size();
  if (p->members.empty()) return PFDTD_OK;
  while ((int)p->inflight.size() >= p->depth) {
    rc = reap(p, true);
    if (rc) return rc;
  }
  if (p->use_graph) {
    POOL_TRY(hipGraphLaunch(p->exec, p->run));
  } else {
    rc = enqueue_bundle_kernels(p);
    if (rc) return rc;
  }
  hipEvent_t ev = nullptr;
  if (!p->spare_events.empty()) {
    ev = p->spare_events.back();
    p->spare_events.pop_back();
  } else {
    POOL_TRY(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  }
  POOL_TRY(hipEventRecord(ev, p->run));
  const uint64_t s = ++p->seq;
  p->inflight.emplace_back(s, ev);
  p->exec_last_seq = s;
  for (int64_t id : p->members) {
    PoolJobHost &j = p->jobs[id];
    const int64_t before = j.steps_enqueued;
    j.steps_enqueued = std::min<int64_t>(j.n_steps, before + p->bundle);
    p->voxel_steps += (double)j.voxels * (double)(j.steps_enqueued - before);
    if (j.steps_enqueued >= j.n_steps) {
      j.phase = kFinishing;
      j.final_seq = s;
      p->dirty = true;
    }
  }
  return PFDTD_OK;
}

int pfdtd_pool_poll(pfdtd_pool *p, int64_t *ids, int capacity, int *n_done) {
  int rc = check_pool(p);
  if (rc) return rc;
  if (!n_done || (capacity > 0 && !ids)) return pfail(PFDTD_E_ARG, "null output");
  rc = reap(p, false);
  if (rc) return rc;
  int n = 0;
  while (n < capacity && !p->done_queue.empty()) {
    ids[n++] = p->done_queue.front();
    p->done_queue.pop_front();
  }
  *n_done = n;
  return PFDTD_OK;
}

int pfdtd_pool_drain(pfdtd_pool *p) {
  int rc = check_pool(p);
  if (rc) return rc;
  while (!p->inflight.empty()) {
    rc = reap(p, true);
    if (rc) return rc;
  }
  return PFDTD_OK;
}

int pfdtd_pool_job_info(pfdtd_pool *p, int64_t id, int *phase, int *n_rec, int *n_steps) {
  if (!p) return pfail(PFDTD_E_ARG, "null pool");
  auto it = p->jobs.find(id);
  if (it == p->jobs.end()) return pfail(PFDTD_E_ARG, "no job %lld", (long long)id);
  if (phase) *phase = (int)it->second.phase;
  if (n_rec) *n_rec = it->second.n_rec;
  if (n_steps) *n_steps = it->second.n_steps;
  return PFDTD_OK;
}

int pfdtd_pool_get_responses(pfdtd_pool *p, int64_t id, float *out_f32, double *out_f64,
                             int64_t capacity) {
  int rc = check_pool(p);
  if (rc) return rc;
  auto it = p->jobs.find(id);
  if (it == p->jobs.end()) return pfail(PFDTD_E_ARG, "no job %lld", (long long)id);
  const PoolJobHost &j = it->second;
  if (j.phase != kFinished) return pfail(PFDTD_E_STATE, "job %lld has not finished", (long long)id);
  const size_t n = (size_t)j.n_rec * j.n_steps;
  if (capacity < (int64_t)n || (!out_f32 && !out_f64)) return pfail(PFDTD_E_ARG, "output buffer too small");
  std::vector<double> tmp(n);
  if (n) {
    POOL_TRY(hipMemcpyAsync(tmp.data(), j.d_resp, sizeof(double) * n, hipMemcpyDeviceToHost, p->upload));
    POOL_TRY(hipStreamSynchronize(p->upload));
  }
  for (size_t i = 0; i < n; ++i) {
    if (out_f64) out_f64[i] = tmp[i];
    if (out_f32) out_f32[i] = (float)tmp[i];
  }
  return PFDTD_OK;
}

int pfdtd_pool_release(pfdtd_pool *p, int64_t id) {
  int rc = check_pool(p);
  if (rc) return rc;
  auto it = p->jobs.find(id);
  if (it == p->jobs.end()) return pfail(PFDTD_E_ARG, "no job %lld", (long long)id);
  if (it->second.phase != kFinished) return pfail(PFDTD_E_STATE, "job %lld has not finished", (long long)id);
  arena_free(p, it->second.off, it->second.bytes);
  p->slot_job[it->second.slot] = -1;
  p->jobs.erase(it);
  return PFDTD_OK;
}

int pfdtd_pool_stats(pfdtd_pool *p, int64_t *bundles, int64_t *graph_builds, double *voxel_steps,
                     int *resident) {
  if (!p) return pfail(PFDTD_E_ARG, "null pool");
  if (bundles) *bundles = (int64_t)p->seq;
  if (graph_builds) *graph_builds = p->graph_builds;
  if (voxel_steps) *voxel_steps = p->voxel_steps;
  if (resident) *resident = (int)p->jobs.size();
  return PFDTD_OK;
}

}  // extern "C"
