// pfdtd_fused.hip -- two SRL_FORWARD time steps per sweep (temporal blocking).
//
// Reads P^n (A), P^{n-1} (B) and the packed position|material plane once and writes P^{n+1} (C)
// and P^{n+2} (D): 18 bytes of HBM traffic per two FP32 node-updates instead of 28. The four
// field buffers are distinct, so blocks never read what another block writes and need no
// synchronisation between blocks.
//
// Tiling: each wavefront owns a 60-column x TY-row x-y tile (TY = 6 for FP32, 4 for FP64, 8 for
// 16-bit storage) and one z-chunk. Its 64 lanes cover the 60 owned columns plus two halo
// columns on each side. P^{n+1} is computed on the owned tile plus a one-voxel ring (lanes
// 1..62, TY+2 rows), from P^n on a two-voxel ring (TY+4 rows); P^{n+2} is then computed on the
// owned tile. x-neighbours come from cross-lane DPP shifts, y- and z-neighbours from registers:
// the wave marches z with P^n of three slices and P^{n+1} of three slices in registers, the
// second time level trailing the first by one slice. Every field and byte access of a slice is a
// buffer instruction (descriptor of the slice in scalar registers, loop-invariant 32-bit lane
// offsets), and all loads of a slice are issued before any is consumed. A block's wave tiles
// (4 x 4 for FP32 on wide rooms, a column of 4 otherwise) step z together (one barrier per
// slice) so that their overlapping halo columns and rows meet in cache, and the boundary
// coefficients are gathered from the block's LDS copy of the table.
// Rooms are mostly air: an air-row plane (one bit per slice, x tile and row, 1/480 of the
// packed plane's bytes) marks the tile rows whose voxels are all plain air, and a wave whose
// P^{n+1} rows of a slice are all marked takes the air value instead of loading the packed
// plane -- the value it would have loaded, so the results do not change.
//
// The same kernel serves two callers: a context's domain (POOL = false) and the two-level jobs
// of one serving-pool launch (POOL = true: the block finds its job in the launch table and takes
// its arguments from the job descriptor; a job with one step left runs only the first level).
// SCHEME selects the boundary formulation of a context's domain: forward (Bilbao, update types
// SRL_FORWARD and SHARED) or centred (Kowalczyk, update type SRL).
//
// Source injection of step n+1 and the receiver gather of step n happen inside the kernel on
// P^{n+1}, in the reference's order (gather before the next injection; sources in index order),
// so every node's arithmetic -- and storage rounding -- equals two single steps bit for bit.
#include "../../include/pfdtd.h"
#include "pfdtd_device.h"

#include <algorithm>
#include <type_traits>

namespace pfdtd {

constexpr int kFO = 60;       // owned columns per wave tile
constexpr int kPO = 124;      // owned columns per wave tile of the two-column sweep
constexpr int kFListCap = 32; // per-wave list capacity for sources / receivers (LDS)
constexpr int kFO3 = 58;      // owned columns per wave tile of the three-level sweep
constexpr int kF3TY = 4;      // owned rows per wave tile of the three-level sweep
constexpr int kThreeBlocks = kSquareBlocks;  // block shape of the three-level sweep

__device__ __forceinline__ bool in_box(const int32_t *q, int x0, int x1, int y0, int y1, int z0,
                                       int z1) {
  return q[0] >= x0 && q[0] < x1 && q[1] >= y0 && q[1] < y1 && q[2] >= z0 && q[2] < z1;
}

// Order-preserving wave-level compaction of the entries of `xyz` inside a box (and, with a slot
// array, of batch slot `slot`) into `list`. Returns the count, or -1 when it exceeds the
// capacity (the caller then scans all entries).
__device__ __forceinline__ int wave_collect(const int32_t *xyz, int n, int x0, int x1, int y0,
                                            int y1, int z0, int z1, int *list, int lane,
                                            const int32_t *slotv = nullptr, int slot = 0) {
  int cnt = 0;
  for (int base = 0; base < n; base += 64) {
    const int s = base + lane;
    const bool hit = s < n && in_box(xyz + 3 * s, x0, x1, y0, y1, z0, z1) &&
                     (!slotv || slotv[s] == slot);
    const uint64_t m = __ballot(hit);
    const int pre = __popcll(m & ((1ull << lane) - 1ull));
    if (hit && cnt >= 0 && cnt + pre < kFListCap) list[cnt + pre] = s;
    if (cnt >= 0) cnt += __popcll(m);
    if (cnt > kFListCap) cnt = -1;
  }
  return cnt;
}

// x-neighbours across the wavefront with DPP shifts: wave_shl:1 gives lane i the value of
// lane i+1, wave_shr:1 the value of lane i-1; lanes without a source read 0 (bound_ctrl; they
// are halo lanes whose results are never used). With no `old` operand to keep, the compiler
// folds each shift into the add that consumes it (v_add_f32_dpp).
__device__ __forceinline__ float lane_next(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x130, 0xF, 0xF, true));
}
__device__ __forceinline__ float lane_prev(float v) {
  return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x138, 0xF, 0xF, true));
}
__device__ __forceinline__ double lane_next(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x130, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x130, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane_prev(double v) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x138, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x138, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}

// Buffer access to one z-slice: the descriptor (base = the slice, range = its bytes) lives in
// scalar registers and every lane passes a loop-invariant 32-bit byte offset, so no per-slice
// 64-bit addresses are formed in vector registers.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t slice_rsrc(const void *p, uint32_t nbytes) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(p), (short)0, (int)nbytes,
                                           0x00020000);
}
template <typename T>
__device__ __forceinline__ T bld(__amdgpu_buffer_rsrc_t r, uint32_t e) {
  if constexpr (sizeof(T) == 1) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b8(r, (int)e, 0, 0));
  } else if constexpr (sizeof(T) == 2) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b16(r, (int)(e * 2u), 0, 0));
  } else if constexpr (sizeof(T) == 4) {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(e * 4u), 0, 0));
  } else {
    return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(e * 8u), 0, 0));
  }
}
template <typename T>
__device__ __forceinline__ void bst(__amdgpu_buffer_rsrc_t r, uint32_t e, T v) {
  if constexpr (sizeof(T) == 2) {
    __builtin_amdgcn_raw_buffer_store_b16(__builtin_bit_cast(unsigned short, v), r, (int)(e * 2u), 0, 0);
  } else if constexpr (sizeof(T) == 4) {
    __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned int, v), r, (int)(e * 4u), 0, 0);
  } else {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)(e * 8u), 0, 0);
  }
}

// Packed position|material<<8 of rows 2i and 2i+1 in one register.
template <int NB>
__device__ __forceinline__ unsigned pm_of(const uint32_t (&b)[NB], int j) {
  return (b[j >> 1] >> ((j & 1) * 16)) & 0xFFFFu;
}

// Owned rows per wave tile TY (even: the byte planes of two rows share a register). Measured
// on the 2048x1536x768 room: FP32 6, the most that keeps the kernel within 128 VGPRs (four waves
// per SIMD; 3 % faster than 4 rows at five); 2-byte fields 8 (120 VGPRs, four waves per SIMD;
// 11 % faster than 10 rows at three waves, 10 % faster than 6 rows at five); FP64 4.
template <typename S>
struct FusedTile {
  static constexpr int TY = 8;
};
template <>
struct FusedTile<float> {
  static constexpr int TY = 6;
};
template <>
struct FusedTile<double> {
  static constexpr int TY = 4;
};
// The centred update keeps six separate neighbours per node: 2-byte fields take 6-row tiles
// there (110-113 VGPRs, four waves per SIMD; 8 rows need 236).
template <typename S, int SCHEME>
constexpr int fused_ty() {
  return (SCHEME == 2 && sizeof(S) == 2) ? 6 : FusedTile<S>::TY;
}
static int fused_ty_host(int precision, int scheme) {
  if (precision == 0) return fused_ty<float, 0>();
  if (precision == 1) return fused_ty<double, 0>();
  return scheme == 2 ? fused_ty<h16, 2>() : fused_ty<h16, 0>();
}
// A block holds WX x WY wave tiles (x-fastest) that step z together, so that the halo columns
// and rows they share meet in cache: a column of 1 x 4 tiles; wide blocks of 4 x 4 tiles (FP32,
// forward scheme), 1024 threads covering 240 columns x 24 rows, one block per CU; or square
// blocks of 2 x 2 tiles (FP64 with the forward scheme, FP32 with the centred one). On the
// 2048x1536x768 room the wide FP32 block is 18 % faster than the column (4 x 2: 4 %, 2 x 4: 8 %
// slower than 4 x 4); on the 512x384x256 room, with 3 of its 12 x tiles past the grid and few
// blocks per CU, it is 7 % slower. Square blocks against the column on the large room, one
// buffer set, five allocation offsets: FP64 forward 0.8-5.6 % faster, FP32 centred 0.8-6 %
// faster (2 x 1 blocks: FP64 1.2-2.4 %); FP64 centred within 2 % either way. The 16-bit sweep is
// not bound by HBM traffic: every wider block (2 x 2, 2 x 4, 4 x 2, 4 x 4) was 1-24 % slower
// than the column, 6 % with the centred scheme. Serving-pool launches (small rooms) use the
// column.
template <int SHAPE>
struct FusedShape {
  static constexpr int WX = SHAPE == kWideBlocks ? 4 : SHAPE == kSquareBlocks ? 2 : 1;
  static constexpr int WY = SHAPE == kSquareBlocks ? 2 : 4;
};
template <int SHAPE>
constexpr int fused_waves() {
  return FusedShape<SHAPE>::WX * FusedShape<SHAPE>::WY;
}

// Kernel arguments: one context's domain, or the two-level jobs of one serving-pool launch.
template <typename S, typename CT, typename C>
struct Fused2Ctx {
  Fused2Args<S, CT, C> a;
  int X, Y, Z, zc;
};
struct Fused2Pool {
  const PoolJobDev *jobs;
  const PoolState *state;
  PoolLaunch L;
  int zc;
};

// Four waves per SIMD (at most 128 VGPRs) for the forward update with FP32 and 16-bit storage;
// FP64 and the centred update (six separate neighbours per node) need more registers.
template <typename S, int SCHEME>
constexpr int fused_min_waves() {
  return (sizeof(S) == 8 || SCHEME == 2) ? 1 : 4;
}

// SCHEME 0: forward (Bilbao) boundaries, update types SRL_FORWARD and SHARED; SCHEME 2: the
// centred (Kowalczyk) boundaries of update type SRL, through the same `centred_node` as the
// single-step kernel (the reference's operation order).
template <typename S, typename CT, typename C, bool POOL, int SHAPE, int SCHEME>
__global__ __launch_bounds__((64 * fused_waves<SHAPE>()), (fused_min_waves<S, SCHEME>())) void srl_update2_kernel(
    typename std::conditional<POOL, Fused2Pool, Fused2Ctx<S, CT, C>>::type p) {
#pragma clang fp contract(off)
  constexpr int kFTY = fused_ty<S, SCHEME>();
  constexpr int WX = FusedShape<SHAPE>::WX, WY = FusedShape<SHAPE>::WY, kFWaves = WX * WY;
  static_assert(kFTY % 2 == 0, "byte planes of two rows share one register");
  constexpr int NA = kFTY + 4;   // P^n rows per wave tile (2-voxel y halo)
  constexpr int NQ = kFTY + 2;   // P^{n+1} rows (1-voxel y halo)
  constexpr int NB = NQ / 2;     // packed byte registers
  __shared__ int s_src[kFWaves][kFListCap];
  __shared__ int s_rec[kFWaves][kFListCap];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  // XCD-aware tile order: consecutive block ids are dealt round-robin to the 8 XCDs (each with
  // its own L2); remapping them so that every XCD walks a contiguous run of x-fastest tiles keeps
  // the halo columns and rows that neighbouring tiles share in one L2 while both are resident.
  Fused2Args<S, CT, C> a;
  int X, Y, Z, zc, bxi, byi, bzi;
  bool two = true;  // second time level (a pool job with one step left takes only the first)
  int slot = 0;
  if constexpr (!POOL) {
    a = p.a;
    X = p.X;
    Y = p.Y;
    Z = p.Z;
    zc = p.zc;
    const unsigned nx = gridDim.x, ny = gridDim.y, total = nx * ny * gridDim.z;
    const unsigned lin = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
    const unsigned full = total & ~7u;
    unsigned tile = lin < full ? (lin & 7u) * (full >> 3) + (lin >> 3) : lin;
    // grouped batch: the slots of one spatial tile are consecutive, so one XCD's L2 serves
    // the tile's packed plane and air-row bits to all of them
    slot = (int)(tile % (unsigned)a.n_slots);
    tile /= (unsigned)a.n_slots;
    bxi = (int)(tile % nx);
    byi = (int)((tile / nx) % ny);
    bzi = (int)(tile / (nx * ny));
    if (slot) {
      const int64_t off = (int64_t)slot * a.slot_elems;
      a.A += off;
      a.B += off;
      a.Cn += off;
      a.Dn += off;
    }
  } else {
    // the flattened tiles of every job of the launch; the block finds its job by a binary
    // search over the launch table's cumulative tile counts
    const unsigned total = gridDim.x, lin = blockIdx.x, full = total & ~7u;
    const int tile = (int)(lin < full ? (lin & 7u) * (full >> 3) + (lin >> 3) : lin);
    int lo = 0, hi = p.L.n - 1;
    while (lo < hi) {
      const int mid = (lo + hi) >> 1;
      if (p.L.tile_end[mid] > tile)
        hi = mid;
      else
        lo = mid + 1;
    }
    const PoolJobDev &J = p.jobs[p.L.slot[lo]];
    const PoolState st = p.state[p.L.slot[lo]];
    const int remaining = J.n_steps - st.step;
    if (remaining <= 0) return;  // the whole block: no barrier has been reached
    two = remaining >= 2;
    int o1, o2;
    pool_spare(st.cur, st.past, o1, o2);
    a.pm = J.pm;
    a.air_rows = J.air_rows;
    a.A = (const S *)J.F[st.cur];
    a.B = (const S *)J.F[st.past];
    a.Cn = (S *)J.F[o1];
    a.Dn = (S *)J.F[o2];
    a.coef = J.coef;
    a.ncol = J.ncol;
    a.air = J.air;
    a.air_uniform = J.air_uniform;
    a.lam2 = J.lam2;
    a.src_xyz = J.src_xyz;
    a.src_kind = J.src_kind;
    a.n_src = J.n_src;
    a.sig_f = J.sig;
    a.sig_d = nullptr;
    a.n_samples = J.n_samples;
    a.inject_step = (J.n_src > 0 && st.step + 1 < J.n_steps) ? st.step + 1 : -1;
    a.soft_overwrite = J.soft_overwrite;
    a.scale_exp = J.scale_exp;
    a.rec_xyz = J.rec_xyz;
    a.n_rec = J.n_rec;
    a.resp = J.resp;
    a.resp_steps = J.n_steps;
    a.gather_step = st.step;
    X = J.X;
    Y = J.Y;
    Z = J.Z;
    zc = p.zc;
    const int t = tile - (lo ? p.L.tile_end[lo - 1] : 0);
    bxi = t % J.fx;
    byi = (t / J.fx) % J.fy;
    bzi = t / (J.fx * J.fy);
  }
  const int xt = bxi * WX + wave % WX;  // this wave's x tile
  const int xb = xt * kFO;  // first owned column
  const int x = xb - 2 + lane;      // this lane's column
  const int yt = (byi * WY + wave / WX) * kFTY;  // first owned row
  const int z0 = 1 + bzi * zc;
  const int z1 = min(z0 + zc, Z - 1);  // owned slices [z0, z1)

  // The (K, material) coefficient rows the domain can address, staged in LDS: gathers from
  // LDS neither go to L2 nor share the vector-memory counter with the field loads.
  extern __shared__ __align__(16) unsigned char s_dyn[];
  typedef typename CoefOf<C>::T CE;
  CE *s_coef = reinterpret_cast<CE *>(s_dyn);
  const unsigned ncol = (unsigned)a.ncol;
  for (unsigned i = threadIdx.x; i < kStageRows * ncol; i += 64 * kFWaves)
    s_coef[i] = a.coef[(i / ncol) * 256u + i % ncol];

  // Sources that land on this wave's P^{n+1} region (owned tile + 1-voxel ring, slices
  // z0-1 .. z1) and receivers on its owned tile.
  const int nsrc = wave_collect(a.src_xyz, a.n_src, xb - 1, xb + kFO + 1, yt - 1, yt + kFTY + 1,
                                z0 - 1, z1 + 1, s_src[wave], lane, a.src_slot, slot);
  const int nrec = wave_collect(a.rec_xyz, a.n_rec, xb, xb + kFO, yt, yt + kFTY, z0, z1,
                                s_rec[wave], lane, a.rec_slot, slot);
  __syncthreads();
  // Waves whose tile lies past the grid stay in the z loop (every row masked) so that the
  // block's waves can step z together.

  const size_t XY = (size_t)X * (size_t)Y;
  const uint32_t sbytes = (uint32_t)(XY * sizeof(S));  // host guarantees < 2^32
  const int xc = min(max(x, 0), X - 1);
  // Rows r = 0..NA-1 of the P^n block are y = yt - 2 + r (clamped to the grid for loads).
  // Node (x, y_r) is updated when its column and its row are off the outer layer, and stored
  // when the lane owns its column and its row lies in the grid. The row parts are wave-uniform
  // bit masks (scalar registers), the column parts one lane condition each, so the per-row
  // conditions cost no 64-bit lane masks kept across the loop.
  uint32_t ro[NA];
  uint32_t row_upd = 0;  // bit r: row y = yt - 2 + r is off the outer layer
#pragma unroll
  for (int r = 0; r < NA; ++r) {
    const int y = yt - 2 + r;
    ro[r] = (uint32_t)(min(max(y, 0), Y - 1) * X + xc);
    if (y > 0 && y < Y - 1) row_upd |= 1u << r;
  }
  const bool col_upd = x > 0 && x < X - 1;
  const bool owned_lane = lane >= 2 && lane < 2 + kFO && x < X;
  uint32_t row_st = 0;  // bit o: owned row o (y = yt + o) lies in the grid
#pragma unroll
  for (int o = 0; o < kFTY; ++o)
    if (yt + o < Y) row_st |= 1u << o;
  const double inv_scale = ldexp(1.0, -a.scale_exp);
  const C scale = (C)ldexp(1.0, a.scale_exp);

  const typename CoefOf<C>::T air = a.air;
  const C lam2 = a.lam2;
  // TABLE: gather the (K, material) entry. Otherwise every position byte of the stage is 0x86
  // (air, inside: the factor pb >> 7 is 1 and 1 * inv == inv) and every K = 6 entry is `air`.
  auto node = [&](auto table, unsigned pm, C pc, C s, C past) -> C {
    if constexpr (decltype(table)::value) {
      const unsigned pb = pm & 0xFFu;
      const typename CoefOf<C>::T e = s_coef[min(pb & 0x7Fu, 6u) * ncol + (pm >> 8)];
      const C aa = (C)(pb >> 7) * e.inv;
      return aa * ((e.tk * pc + lam2 * s) - e.omb * past);
    } else {
      return air.inv * ((air.tk * pc + lam2 * s) - air.omb * past);
    }
  };
  const bool air_uniform = a.air_uniform != 0;
  // wave-uniform: does any of the NQ packed rows hold a voxel other than air (0x86)?
  auto any_nonair = [&](const uint32_t (&b)[NB]) -> bool {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < NB; ++i) m |= ((b[i] & 0xFFu) ^ 0x86u) | (((b[i] >> 16) & 0xFFu) ^ 0x86u);
    return !air_uniform || __ballot(m != 0u) != 0;
  };
  // the same for the centred scheme's air byte 0x80; its air entry holds for every material
  // byte, so the material does not matter
  auto any_noncair = [&](const uint32_t (&b)[NB]) -> bool {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < NB; ++i) m |= ((b[i] & 0xFFu) ^ 0x80u) | (((b[i] >> 16) & 0xFFu) ^ 0x80u);
    return __ballot(m != 0u) != 0;
  };
  const S *__restrict__ gA = a.A;
  const S *__restrict__ gB = a.B;
  S *__restrict__ gC = a.Cn;
  S *__restrict__ gD = a.Dn;
  const uint16_t *__restrict__ gpm = a.pm;
  // This wave's P^{n+1} rows yt-1 .. yt+kFTY are bits yt .. yt+NQ-1 of its tile's air-row
  // words (bit b of word w: row 32 w + b - 1). Waves past the grid count as air: every one of
  // their rows is masked.
  // The words are read through the constant address space (scalar loads, counted apart from
  // the slice's vector loads), one slice ahead of their use.
  typedef const __attribute__((address_space(4))) uint32_t *ConstU32;
  const uint32_t *__restrict__ gar = a.air_rows;
  const int nyw = Y / 32 + 2;
  const size_t ar_slice = (size_t)((X + kFO - 1) / kFO) * (size_t)nyw;
  const bool past_grid = yt >= Y || xb >= X;
  const uint32_t *ar_tile = gar && !past_grid ? gar + (size_t)xt * nyw + (yt >> 5) : nullptr;
  auto air_window = [&](int k) -> uint64_t {
    if (!gar || k < 1 || k > Z - 2) return 0;
    if (past_grid) return ~0ull;
    const ConstU32 w = (ConstU32)(ar_tile + (size_t)k * ar_slice);
    return (((uint64_t)w[1] << 32) | w[0]) >> (yt & 31);
  };
  constexpr uint64_t kAllRows = (1ull << NQ) - 1ull;
  // the packed value of plain air (position 0x86 forward, 0x80 centred; material 0), twice
  constexpr uint32_t kAirPair = SCHEME == 2 ? 0x00800080u : 0x00860086u;
  uint64_t ar_win = air_window(z0 - 1);  // the window of the slice the next body call updates

  // One z-iteration: first level at slice k (rows 1..NQ), second level at slice k-1 (rows
  // 2..NQ-1). am/ac/ap: P^n at k-1, k, k+1 (NA rows); q0/q1/q2: P^{n+1} at k-2, k-1, k (rows
  // 1..NQ at index row-1); bp/bc: packed bytes at k-1 and k (rows 1..NQ).
  auto body = [&](C (&am)[NA], C (&ac)[NA], C (&ap)[NA], C (&q0)[NQ], C (&q1)[NQ], C (&q2)[NQ],
                  uint32_t (&bp)[NB], uint32_t (&bc)[NB], int k) {
    // fresh copies each slice: the per-row conditions derived from them are not hoisted out of
    // the loop as lane masks
    uint32_t ru = __builtin_amdgcn_readfirstlane(row_upd);
    uint32_t rs = __builtin_amdgcn_readfirstlane(row_st);
    asm volatile("" : "+s"(ru), "+s"(rs));
    // the block's waves load overlapping halo rows and columns: keeping them on the same slice lets
    // the cache serve the overlap once
    __builtin_amdgcn_s_barrier();
    const bool upd1 = k >= 1 && k <= Z - 2;
    if (k + 1 <= Z - 1) {
      const auto rA = slice_rsrc(gA + (size_t)(k + 1) * XY, sbytes);
#pragma unroll
      for (int r = 0; r < NA; ++r) ap[r] = Fmt<S>::val(bld<S>(rA, ro[r]));
    } else {
#pragma unroll
      for (int r = 0; r < NA; ++r) ap[r] = (C)0;
    }
    C past[NQ];
    const bool air1 = upd1 && (ar_win & kAllRows) == kAllRows;  // all P^{n+1} rows plain air
    ar_win = air_window(k + 1);
    if (upd1) {
      const auto rB = slice_rsrc(gB + (size_t)k * XY, sbytes);
#pragma unroll
      for (int j = 0; j < NQ; ++j) past[j] = Fmt<S>::val(bld<S>(rB, ro[j + 1]));
      if (air1) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < NB; ++i) bc[i] = kAirPair;
      } else {
        const auto rPM = slice_rsrc(gpm + (size_t)k * XY, (uint32_t)(XY * 2));
        uint32_t pmv[NQ];
#pragma unroll
        for (int j = 0; j < NQ; ++j) pmv[j] = bld<uint16_t>(rPM, ro[j + 1]);
        // every load of the slice is in flight before any result is combined
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < NB; ++i) bc[i] = pmv[2 * i] | (pmv[2 * i + 1] << 16);
      }
    } else {
#pragma unroll
      for (int j = 0; j < NQ; ++j) past[j] = (C)0;
#pragma unroll
      for (int i = 0; i < NB; ++i) bc[i] = 0;
    }

    // ---- first time level: P^{n+1}[k]
    auto level1 = [&](auto table) {
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        const int r = j + 1;
        C s = ap[r] + am[r];                 // z+1, z-1
        s = s + ac[r + 1];                   // y+1
        s = s + ac[r - 1];                   // y-1
        s = s + lane_next(ac[r]);            // x+1
        s = s + lane_prev(ac[r]);            // x-1
        const C v = Fmt<S>::round(node(table, pm_of(bc, j), ac[r], s, past[j]));
        q2[j] = (upd1 && ((ru >> r) & 1u) && col_upd) ? v : (C)0;
      }
    };
    if constexpr (SCHEME == 2) {
      // slices whose rows are all air (the air-row bits, or the packed bytes) skip the byte
      // decode and the table gather of every node
      auto level1c = [&](auto allair) {
#pragma unroll
        for (int j = 0; j < NQ; ++j) {
          const int r = j + 1;
          const unsigned pmv = pm_of(bc, j);
          C w;
          if constexpr (decltype(allair)::value)
            w = centred_air<C>(ac[r], ap[r], am[r], ac[r + 1], ac[r - 1], lane_next(ac[r]),
                               lane_prev(ac[r]), past[j], air, lam2);
          else
            w = centred_node<C>(pmv & 0xFFu, pmv >> 8, ac[r], ap[r], am[r], ac[r + 1],
                                ac[r - 1], lane_next(ac[r]), lane_prev(ac[r]), past[j], s_coef,
                                ncol, air, lam2);
          const C v = Fmt<S>::round(w);
          q2[j] = (upd1 && ((ru >> r) & 1u) && col_upd) ? v : (C)0;
        }
      };
      if (air1 || !any_noncair(bc))
        level1c(std::true_type());
      else
        level1c(std::false_type());
    } else if (air1 ? !air_uniform : any_nonair(bc)) {
      level1(std::true_type());
    } else {
      level1(std::false_type());
    }

    // ---- receivers (before injection) and sources on P^{n+1}[k]
    if (nrec != 0) {
      const int cnt = nrec < 0 ? a.n_rec : nrec;
      for (int t = 0; t < cnt; ++t) {
        const int id = nrec < 0 ? t : s_rec[wave][t];
        const int32_t *q = a.rec_xyz + 3 * id;
        if (q[2] != k || q[0] != x) continue;
        if (nrec < 0 && (!in_box(q, xb, xb + kFO, yt, yt + kFTY, z0, z1) ||
                         (a.rec_slot && a.rec_slot[id] != slot)))
          continue;
#pragma unroll
        for (int j = 0; j < NQ; ++j)
          if (q[1] == yt - 1 + j)
            a.resp[(int64_t)id * a.resp_steps + a.gather_step] = (double)q2[j] * inv_scale;
      }
    }
    if (nsrc != 0 && a.inject_step >= 0) {
      const int cnt = nsrc < 0 ? a.n_src : nsrc;
      for (int t = 0; t < cnt; ++t) {
        const int id = nsrc < 0 ? t : s_src[wave][t];
        const int32_t *q = a.src_xyz + 3 * id;
        if (q[2] != k || q[0] != x) continue;
        if (nsrc < 0 && a.src_slot && a.src_slot[id] != slot) continue;
        C smp = (C)0;
        if (a.inject_step < a.n_samples) {
          const int64_t off = (int64_t)id * a.n_samples + a.inject_step;
          if (sizeof(C) == sizeof(double))
            smp = (C)a.sig_d[off];
          else
            smp = (C)a.sig_f[off];
        }
        smp = smp * scale;
        const bool hard = a.src_kind[id] == 0 || a.soft_overwrite;
#pragma unroll
        for (int j = 0; j < NQ; ++j)
          if (q[1] == yt - 1 + j) q2[j] = hard ? Fmt<S>::round(smp) : Fmt<S>::round(q2[j] + smp);
      }
    }

    // ---- store the owned part of P^{n+1}[k]
    if (k >= z0 && k < z1) {
      const auto rC = slice_rsrc(gC + (size_t)k * XY, sbytes);
#pragma unroll
      for (int o = 0; o < kFTY; ++o)
        if (((rs >> o) & 1u) && owned_lane) bst<S>(rC, ro[o + 2], Fmt<S>::raw(q2[o + 1]));
    }

    // ---- second time level: P^{n+2}[k-1] on the owned rows
    if (two && k - 1 >= z0) {
      const auto rD = slice_rsrc(gD + (size_t)(k - 1) * XY, sbytes);
      auto level2 = [&](auto table) {
#pragma unroll
        for (int o = 0; o < kFTY; ++o) {
          const int j = o + 1;
          C s = q2[j] + q0[j];               // z+1, z-1
          s = s + q1[j + 1];                 // y+1
          s = s + q1[j - 1];                 // y-1
          s = s + lane_next(q1[j]);          // x+1
          s = s + lane_prev(q1[j]);          // x-1
          C v = node(table, pm_of(bp, j), q1[j], s, am[o + 2]);
          v = (((ru >> (o + 2)) & 1u) && col_upd) ? v : (C)0;
          if (((rs >> o) & 1u) && owned_lane) bst<S>(rD, ro[o + 2], Fmt<S>::raw(v));
        }
      };
      if constexpr (SCHEME == 2) {
        auto level2c = [&](auto allair) {
#pragma unroll
          for (int o = 0; o < kFTY; ++o) {
            const int j = o + 1;
            const unsigned pmv = pm_of(bp, j);
            C v;
            if constexpr (decltype(allair)::value)
              v = centred_air<C>(q1[j], q2[j], q0[j], q1[j + 1], q1[j - 1], lane_next(q1[j]),
                                 lane_prev(q1[j]), am[o + 2], air, lam2);
            else
              v = centred_node<C>(pmv & 0xFFu, pmv >> 8, q1[j], q2[j], q0[j], q1[j + 1],
                                  q1[j - 1], lane_next(q1[j]), lane_prev(q1[j]), am[o + 2],
                                  s_coef, ncol, air, lam2);
            v = (((ru >> (o + 2)) & 1u) && col_upd) ? v : (C)0;
            if (((rs >> o) & 1u) && owned_lane) bst<S>(rD, ro[o + 2], Fmt<S>::raw(v));
          }
        };
        if (any_noncair(bp))
          level2c(std::false_type());
        else
          level2c(std::true_type());
      } else if (any_nonair(bp)) {
        level2(std::true_type());
      } else {
        level2(std::false_type());
      }
    }
  };

  C A0[NA], A1[NA], A2[NA];
  C Q0[NQ], Q1[NQ], Q2[NQ];
  uint32_t B0[NB], B1[NB], B2[NB];
  const int k0 = z0 - 1;
  const int k1 = z1;
  {
    const auto rK = slice_rsrc(gA + (size_t)k0 * XY, sbytes);
#pragma unroll
    for (int r = 0; r < NA; ++r) A1[r] = Fmt<S>::val(bld<S>(rK, ro[r]));
    if (k0 >= 1) {
      const auto rKm = slice_rsrc(gA + (size_t)(k0 - 1) * XY, sbytes);
#pragma unroll
      for (int r = 0; r < NA; ++r) A0[r] = Fmt<S>::val(bld<S>(rKm, ro[r]));
    } else {
#pragma unroll
      for (int r = 0; r < NA; ++r) A0[r] = (C)0;
    }
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      Q0[j] = (C)0;
      Q1[j] = (C)0;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) B0[i] = 0;
  }
  // Unrolled by three so the three-slice queues rotate by renaming, not by copies.
  for (int k = k0;;) {
    body(A0, A1, A2, Q0, Q1, Q2, B0, B1, k);
    if (++k > k1) break;
    body(A1, A2, A0, Q1, Q2, Q0, B1, B2, k);
    if (++k > k1) break;
    body(A2, A0, A1, Q2, Q0, Q1, B2, B0, k);
    if (++k > k1) break;
  }
}

// ---- Three time levels per launch (forward scheme) -------------------------------------------
// P^{n+1}, P^{n+2} and P^{n+3} in one z-sweep: z-iteration k takes the first level at slice k,
// the second at k-1 and the third at k-2, from three-slice register queues of P^n, P^{n+1} and
// P^{n+2}. Only P^{n+2} (Cn) and P^{n+3} (Dn) -- the next launch's past and current fields --
// are stored, so a launch moves the two-level sweep's bytes for three steps. A wave tile owns
// 58 columns (lanes 3..60: each level consumes one halo column on either side) and 4 rows (one
// halo row per level). The sources of steps n+1 and n+2 are injected, and the receivers of
// steps n and n+1 gathered, inside the sweep in the reference's order; step n+2's gather and
// step n+3's injection follow on P^{n+3} (launch_inject_gather), as after the two-level sweep.
// Node arithmetic, sum order and storage rounding are the two-level sweep's, so the results are
// three single steps bit for bit.
template <typename S, typename CT, typename C, int SHAPE, int BG>
__global__ __launch_bounds__((64 * fused_waves<SHAPE>()), (fused_min_waves<S, 0>())) void srl_update3_kernel(
    Fused2Ctx<S, CT, C> p) {
#pragma clang fp contract(off)
  constexpr int TY = kF3TY;
  constexpr int WX = FusedShape<SHAPE>::WX, WY = FusedShape<SHAPE>::WY, kW = WX * WY;
  constexpr int NA = TY + 6;  // P^n rows yt-3 .. yt+TY+2
  constexpr int N1 = TY + 4;  // P^{n+1} rows yt-2 .. yt+TY+1
  constexpr int N2 = TY + 2;  // P^{n+2} rows yt-1 .. yt+TY
  constexpr int NB = N1 / 2;  // packed byte registers of the N1 rows
  __shared__ int s_src[kW][kFListCap];
  __shared__ int s_rec[kW][kFListCap];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  Fused2Args<S, CT, C> a = p.a;
  const int X = p.X, Y = p.Y, Z = p.Z, zc = p.zc;
  // XCD-aware tile order and grouped-batch slots, as in srl_update2_kernel
  const unsigned nx = gridDim.x, ny = gridDim.y, total = nx * ny * gridDim.z;
  const unsigned lin = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const unsigned full = total & ~7u;
  unsigned tile = lin < full ? (lin & 7u) * (full >> 3) + (lin >> 3) : lin;
  const int slot = (int)(tile % (unsigned)a.n_slots);
  tile /= (unsigned)a.n_slots;
  if (slot) {
    const int64_t off = (int64_t)slot * a.slot_elems;
    a.A += off;
    a.B += off;
    a.Cn += off;
    a.Dn += off;
  }
  // x-fastest over bands of BG block rows, the rows of a band interleaved: a block's
  // y-neighbours in the band are resident with it, while concurrent blocks still stream along x
  // (DRAM pages). On the 2048x1536x768 room, FP32: reads 3.97e10 -> 3.39e10 B per launch with
  // bands of two, 3.90 -> 3.85 ms per step (three rows 3.87, four 3.88 at 3.11e10 B, y-fastest
  // order 4.94 at 3.27e10 B); FP64: 7.66 / 7.34 / 7.26 ms for bands of one / two / three.
  const unsigned tl = tile % (nx * ny);
  const unsigned band = tl / ((unsigned)BG * nx), in_band = tl % ((unsigned)BG * nx);
  const unsigned band_rows = min((unsigned)BG, ny - band * (unsigned)BG);
  const int byi = (int)(band * (unsigned)BG + in_band % band_rows);
  const int bxi = (int)(in_band / band_rows);
  const int bzi = (int)(tile / (nx * ny));
  const int xt = bxi * WX + wave % WX;            // this wave's x tile
  const int xb = xt * kFO3;                       // first owned column
  const int x = xb - 3 + lane;                    // this lane's column
  const int yt = (byi * WY + wave / WX) * TY;     // first owned row
  const int z0 = 1 + bzi * zc;
  const int z1 = min(z0 + zc, Z - 1);             // owned slices [z0, z1)

  extern __shared__ __align__(16) unsigned char s_dyn[];
  typedef typename CoefOf<C>::T CE;
  CE *s_coef = reinterpret_cast<CE *>(s_dyn);
  const unsigned ncol = (unsigned)a.ncol;
  for (unsigned i = threadIdx.x; i < kStageRows * ncol; i += 64 * kW)
    s_coef[i] = a.coef[(i / ncol) * 256u + i % ncol];
  // sources on this wave's first-level region (owned tile + 2-voxel ring, slices z0-2 .. z1+1);
  // receivers on its owned tile
  const int nsrc = wave_collect(a.src_xyz, a.n_src, xb - 2, xb + kFO3 + 2, yt - 2, yt + TY + 2,
                                z0 - 2, z1 + 2, s_src[wave], lane, a.src_slot, slot);
  const int nrec = wave_collect(a.rec_xyz, a.n_rec, xb, xb + kFO3, yt, yt + TY, z0, z1,
                                s_rec[wave], lane, a.rec_slot, slot);
  __syncthreads();

  const size_t XY = (size_t)X * (size_t)Y;
  const uint32_t sbytes = (uint32_t)(XY * sizeof(S));  // host guarantees < 2^32
  const int xc = min(max(x, 0), X - 1);
  // rows r = 0..NA-1 of the P^n block are y = yt - 3 + r (clamped for loads); row bits as in
  // srl_update2_kernel
  uint32_t ro[NA];
  uint32_t row_upd = 0;
#pragma unroll
  for (int r = 0; r < NA; ++r) {
    const int y = yt - 3 + r;
    ro[r] = (uint32_t)(min(max(y, 0), Y - 1) * X + xc);
    if (y > 0 && y < Y - 1) row_upd |= 1u << r;
  }
  const bool col_upd = x > 0 && x < X - 1;
  const bool owned_lane = lane >= 3 && lane < 3 + kFO3 && x < X;
  uint32_t row_st = 0;
#pragma unroll
  for (int o = 0; o < TY; ++o)
    if (yt + o < Y) row_st |= 1u << o;
  const double inv_scale = ldexp(1.0, -a.scale_exp);
  const C scale = (C)ldexp(1.0, a.scale_exp);
  const CE air = a.air;
  const C lam2 = a.lam2;
  auto node = [&](auto table, unsigned pm, C pc, C s, C past) -> C {
    if constexpr (decltype(table)::value) {
      const unsigned pb = pm & 0xFFu;
      const CE e = s_coef[min(pb & 0x7Fu, 6u) * ncol + (pm >> 8)];
      const C aa = (C)(pb >> 7) * e.inv;
      return aa * ((e.tk * pc + lam2 * s) - e.omb * past);
    } else {
      return air.inv * ((air.tk * pc + lam2 * s) - air.omb * past);
    }
  };
  const bool air_uniform = a.air_uniform != 0;
  auto any_nonair = [&](const uint32_t (&b)[NB]) -> bool {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < NB; ++i) m |= ((b[i] & 0xFFu) ^ 0x86u) | (((b[i] >> 16) & 0xFFu) ^ 0x86u);
    return !air_uniform || __ballot(m != 0u) != 0;
  };
  const S *__restrict__ gA = a.A;
  const S *__restrict__ gB = a.B;
  S *__restrict__ gC = a.Cn;
  S *__restrict__ gD = a.Dn;
  const uint16_t *__restrict__ gpm = a.pm;
  // air-row bits of the 58-column tiles (bit b of word w: row 32 w + b - 2), so this wave's
  // first-level rows yt-2 .. yt+TY+1 start at bit yt
  typedef const __attribute__((address_space(4))) uint32_t *ConstU32;
  const uint32_t *__restrict__ gar = a.air_rows;
  const int nyw = Y / 32 + 2;
  const size_t ar_slice = (size_t)((X + kFO3 - 1) / kFO3) * (size_t)nyw;
  const bool past_grid = yt >= Y || xb >= X;
  const uint32_t *ar_tile = gar && !past_grid ? gar + (size_t)xt * nyw + (yt >> 5) : nullptr;
  auto air_window = [&](int k) -> uint64_t {
    if (!gar || k < 1 || k > Z - 2) return 0;
    if (past_grid) return ~0ull;
    const ConstU32 w = (ConstU32)(ar_tile + (size_t)k * ar_slice);
    return (((uint64_t)w[1] << 32) | w[0]) >> (yt & 31);
  };
  constexpr uint64_t kAllRows = (1ull << N1) - 1ull;
  uint64_t ar_win = air_window(z0 - 2);

  // receivers of one level on the values `q` (rows y0 .. y0+n-1) at slice k, then the sources
  // of step `inj` on them
  auto gather_inject = [&](auto &q, int y0, int k, int gstep, int inj) {
    constexpr int n = (int)(sizeof(q) / sizeof(q[0]));
    if (nrec != 0 && gstep >= 0) {
      const int cnt = nrec < 0 ? a.n_rec : nrec;
      for (int t = 0; t < cnt; ++t) {
        const int id = nrec < 0 ? t : s_rec[wave][t];
        const int32_t *qq = a.rec_xyz + 3 * id;
        if (qq[2] != k || qq[0] != x) continue;
        if (nrec < 0 && (!in_box(qq, xb, xb + kFO3, yt, yt + TY, z0, z1) ||
                         (a.rec_slot && a.rec_slot[id] != slot)))
          continue;
#pragma unroll
        for (int j = 0; j < n; ++j)
          if (qq[1] == y0 + j) a.resp[(int64_t)id * a.resp_steps + gstep] = (double)q[j] * inv_scale;
      }
    }
    if (nsrc != 0 && inj >= 0) {
      const int cnt = nsrc < 0 ? a.n_src : nsrc;
      for (int t = 0; t < cnt; ++t) {
        const int id = nsrc < 0 ? t : s_src[wave][t];
        const int32_t *qq = a.src_xyz + 3 * id;
        if (qq[2] != k || qq[0] != x) continue;
        if (nsrc < 0 && a.src_slot && a.src_slot[id] != slot) continue;
        C smp = (C)0;
        if (inj < a.n_samples) {
          const int64_t off = (int64_t)id * a.n_samples + inj;
          if (sizeof(C) == sizeof(double))
            smp = (C)a.sig_d[off];
          else
            smp = (C)a.sig_f[off];
        }
        smp = smp * scale;
        const bool hard = a.src_kind[id] == 0 || a.soft_overwrite;
#pragma unroll
        for (int j = 0; j < n; ++j)
          if (qq[1] == y0 + j) q[j] = hard ? Fmt<S>::round(smp) : Fmt<S>::round(q[j] + smp);
      }
    }
  };

  // am/ac/ap: P^n at k-1, k, k+1; qa/qb/qc: P^{n+1} at k-2, k-1, k; ra/rb/rc: P^{n+2} at
  // k-3, k-2, k-1; bpp/bp/bc: packed bytes at k-2, k-1, k
  auto body = [&](C (&am)[NA], C (&ac)[NA], C (&ap)[NA], C (&qa)[N1], C (&qb)[N1], C (&qc)[N1],
                  C (&ra)[N2], C (&rb)[N2], C (&rc)[N2], uint32_t (&bpp)[NB], uint32_t (&bp)[NB],
                  uint32_t (&bc)[NB], int k) {
    uint32_t ru = __builtin_amdgcn_readfirstlane(row_upd);
    uint32_t rs = __builtin_amdgcn_readfirstlane(row_st);
    asm volatile("" : "+s"(ru), "+s"(rs));
    __builtin_amdgcn_s_barrier();
    const bool upd1 = k >= 1 && k <= Z - 2;
    if (k + 1 >= 0 && k + 1 <= Z - 1) {
      const auto rA = slice_rsrc(gA + (size_t)(k + 1) * XY, sbytes);
#pragma unroll
      for (int r = 0; r < NA; ++r) ap[r] = Fmt<S>::val(bld<S>(rA, ro[r]));
    } else {
#pragma unroll
      for (int r = 0; r < NA; ++r) ap[r] = (C)0;
    }
    C past[N1];
    const bool air1 = upd1 && (ar_win & kAllRows) == kAllRows;
    ar_win = air_window(k + 1);
    if (upd1) {
      const auto rB = slice_rsrc(gB + (size_t)k * XY, sbytes);
#pragma unroll
      for (int j = 0; j < N1; ++j) past[j] = Fmt<S>::val(bld<S>(rB, ro[j + 1]));
      if (air1) {
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < NB; ++i) bc[i] = 0x00860086u;
      } else {
        const auto rPM = slice_rsrc(gpm + (size_t)k * XY, (uint32_t)(XY * 2));
        uint32_t pmv[N1];
#pragma unroll
        for (int j = 0; j < N1; ++j) pmv[j] = bld<uint16_t>(rPM, ro[j + 1]);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < NB; ++i) bc[i] = pmv[2 * i] | (pmv[2 * i + 1] << 16);
      }
    } else {
#pragma unroll
      for (int j = 0; j < N1; ++j) past[j] = (C)0;
#pragma unroll
      for (int i = 0; i < NB; ++i) bc[i] = 0;
    }

    // ---- first level: P^{n+1}[k], rows yt-2 .. yt+TY+1
    auto level1 = [&](auto table) {
#pragma unroll
      for (int j = 0; j < N1; ++j) {
        const int r = j + 1;
        C s = ap[r] + am[r];                 // z+1, z-1
        s = s + ac[r + 1];                   // y+1
        s = s + ac[r - 1];                   // y-1
        s = s + lane_next(ac[r]);            // x+1
        s = s + lane_prev(ac[r]);            // x-1
        const C v = Fmt<S>::round(node(table, pm_of(bc, j), ac[r], s, past[j]));
        qc[j] = (upd1 && ((ru >> r) & 1u) && col_upd) ? v : (C)0;
      }
    };
    if (air1 ? !air_uniform : any_nonair(bc))
      level1(std::true_type());
    else
      level1(std::false_type());
    gather_inject(qc, yt - 2, k, a.gather_step, a.inject_step);

    // ---- second level: P^{n+2}[k-1], rows yt-1 .. yt+TY
    const int k2 = k - 1;
    if (k2 >= z0 - 1) {
      const bool upd2 = k2 >= 1 && k2 <= Z - 2;
      auto level2 = [&](auto table) {
#pragma unroll
        for (int i = 0; i < N2; ++i) {
          const int j = i + 1;
          C s = qc[j] + qa[j];               // z+1, z-1
          s = s + qb[j + 1];                 // y+1
          s = s + qb[j - 1];                 // y-1
          s = s + lane_next(qb[j]);          // x+1
          s = s + lane_prev(qb[j]);          // x-1
          const C v = Fmt<S>::round(node(table, pm_of(bp, j), qb[j], s, am[i + 2]));
          rc[i] = (upd2 && ((ru >> (i + 2)) & 1u) && col_upd) ? v : (C)0;
        }
      };
      if (any_nonair(bp))
        level2(std::true_type());
      else
        level2(std::false_type());
      gather_inject(rc, yt - 1, k2, a.gather_step2, a.inject_step2);
      if (k2 >= z0 && k2 < z1) {
        const auto rC = slice_rsrc(gC + (size_t)k2 * XY, sbytes);
#pragma unroll
        for (int o = 0; o < TY; ++o)
          if (((rs >> o) & 1u) && owned_lane) bst<S>(rC, ro[o + 3], Fmt<S>::raw(rc[o + 1]));
      }
    }

    // ---- third level: P^{n+3}[k-2] on the owned rows
    const int k3 = k - 2;
    if (k3 >= z0) {
      const auto rD = slice_rsrc(gD + (size_t)k3 * XY, sbytes);
      auto level3 = [&](auto table) {
#pragma unroll
        for (int o = 0; o < TY; ++o) {
          const int i = o + 1;
          C s = rc[i] + ra[i];               // z+1, z-1
          s = s + rb[i + 1];                 // y+1
          s = s + rb[i - 1];                 // y-1
          s = s + lane_next(rb[i]);          // x+1
          s = s + lane_prev(rb[i]);          // x-1
          C v = node(table, pm_of(bpp, o + 2), rb[i], s, qa[o + 2]);
          v = (((ru >> (o + 3)) & 1u) && col_upd) ? v : (C)0;
          if (((rs >> o) & 1u) && owned_lane) bst<S>(rD, ro[o + 3], Fmt<S>::raw(v));
        }
      };
      if (any_nonair(bpp))
        level3(std::true_type());
      else
        level3(std::false_type());
    }
  };

  C A0[NA], A1[NA], A2[NA];
  C Q0[N1], Q1[N1], Q2[N1];
  C R0[N2], R1[N2], R2[N2];
  uint32_t B0[NB], B1[NB], B2[NB];
  const int k0 = z0 - 2;
  const int k1 = z1 + 1;
  if (k0 >= 0) {
    const auto rK = slice_rsrc(gA + (size_t)k0 * XY, sbytes);
#pragma unroll
    for (int r = 0; r < NA; ++r) A1[r] = Fmt<S>::val(bld<S>(rK, ro[r]));
  } else {
#pragma unroll
    for (int r = 0; r < NA; ++r) A1[r] = (C)0;
  }
  if (k0 >= 1) {
    const auto rKm = slice_rsrc(gA + (size_t)(k0 - 1) * XY, sbytes);
#pragma unroll
    for (int r = 0; r < NA; ++r) A0[r] = Fmt<S>::val(bld<S>(rKm, ro[r]));
  } else {
#pragma unroll
    for (int r = 0; r < NA; ++r) A0[r] = (C)0;
  }
#pragma unroll
  for (int j = 0; j < N1; ++j) {
    Q0[j] = (C)0;
    Q1[j] = (C)0;
  }
#pragma unroll
  for (int i = 0; i < N2; ++i) {
    R0[i] = (C)0;
    R1[i] = (C)0;
    R2[i] = (C)0;
  }
#pragma unroll
  for (int i = 0; i < NB; ++i) {
    B0[i] = 0;
    B1[i] = 0;
  }
  // unrolled by three so the three-slice queues rotate by renaming
  for (int k = k0;;) {
    body(A0, A1, A2, Q0, Q1, Q2, R0, R1, R2, B0, B1, B2, k);
    if (++k > k1) break;
    body(A1, A2, A0, Q1, Q2, Q0, R1, R2, R0, B1, B2, B0, k);
    if (++k > k1) break;
    body(A2, A0, A1, Q2, Q0, Q1, R2, R0, R1, B2, B0, B1, k);
    if (++k > k1) break;
  }
}

// ---- Two x-columns per lane (16-bit storage) ---------------------------------------------------
// With one column per lane the 16-bit sweep issues one 2-byte load per voxel row: it is bound by
// load issue and latency, not by bytes. Here lane l holds the adjacent columns xb-2+2l and
// xb-1+2l and moves them with one 32-bit access, half the vector-memory instructions per byte:
// a wave covers 128 columns, lanes 1..62 own 124 of them and lanes 0 and 63 hold the two halo
// columns on either side. X must be even, so that every pair starts on a 4-byte boundary (rows
// and slices of even length). The x-neighbours of a column are the other half of its own pair
// and one half of the neighbouring lane's pair (a DPP shift); everything else -- the order of
// the sums, the coefficient gathers, the rounding, injection and gather -- is the one-column
// sweep's, so the results are the same bits.
// Two adjacent columns of one row, as FP32 values (a plain aggregate: HIP's vector types are
// unions, which keep the row queues from being promoted to registers).
struct Pair2 {
  float x, y;
};

// One column of a pair chosen per lane. The register barrier keeps the compiler from turning the
// select of two loads into one load through a selected address, which would put the row queue
// in scratch memory.
__device__ __forceinline__ float pick(const Pair2 &p, bool hi) {
  float lo = p.x, h = p.y;
  asm("" : "+v"(lo), "+v"(h));
  return hi ? h : lo;
}

template <typename S>
__device__ __forceinline__ float lo_half(uint32_t u) {
  return Fmt<S>::val(S{(uint16_t)(u & 0xFFFFu)});
}
template <typename S>
__device__ __forceinline__ float hi_half(uint32_t u) {
  return Fmt<S>::val(S{(uint16_t)(u >> 16)});
}
template <typename S>
__device__ __forceinline__ uint32_t pack_halves(float a, float b) {
  return (uint32_t)Fmt<S>::bits(a) | ((uint32_t)Fmt<S>::bits(b) << 16);
}
__device__ __forceinline__ uint32_t bld_pair(__amdgpu_buffer_rsrc_t r, uint32_t e) {
  return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)(e * 2u), 0, 0);
}
__device__ __forceinline__ void bst_pair(__amdgpu_buffer_rsrc_t r, uint32_t e, uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)(e * 2u), 0, 0);
}
// The same at a byte offset (the difference-form sweep keeps byte offsets, one per tile row).
__device__ __forceinline__ uint32_t bld_pair_at(__amdgpu_buffer_rsrc_t r, uint32_t b) {
  return (uint32_t)__builtin_amdgcn_raw_buffer_load_b32(r, (int)b, 0, 0);
}
__device__ __forceinline__ void bst_pair_at(__amdgpu_buffer_rsrc_t r, uint32_t b, uint32_t v) {
  __builtin_amdgcn_raw_buffer_store_b32(v, r, (int)b, 0, 0);
}

// A lane that must not store sets bit 31 of its vector offset. Slices of the two-column sweep
// are under 2^31 bytes (the host checks it), so that offset is past every slice's range and the
// buffer range check drops the write: masked stores need no lane-masked branch.
constexpr uint32_t kDropShift = 0x80000000u;
// pack_halves for values that are already exact in the storage format (rounded results): the
// conversion is exact, so it needs no register barrier (and no copy of a value still in use).
template <typename S>
__device__ __forceinline__ uint16_t exact_bits(float v) {
  if constexpr (std::is_same<S, h16>::value)
    return __builtin_bit_cast(uint16_t, (_Float16)v);
  else
    return (uint16_t)(__float_as_uint(v) >> 16);
}
template <typename S>
__device__ __forceinline__ uint32_t pack_exact(float a, float b) {
  return (uint32_t)exact_bits<S>(a) | ((uint32_t)exact_bits<S>(b) << 16);
}

// Rows per wave tile of the two-column sweep: 7 (167 VGPRs, three waves per SIMD). Each level
// recomputes one halo row on either side and loads two more P^n rows, so taller tiles do less
// work per owned row for fewer resident waves. Per step on the 2048x1536x768 room, FP16, blocks
// of 2 x 2 waves, one buffer set: 4 rows (119 VGPRs, four waves) 2.642 ms, 5 rows 2.798, 6 rows
// 2.622, 7 rows 2.581, 8 rows at two waves 3.388 (at three they spill 72 B per lane); with 7
// rows, in another session, 16 / 18 / 20 / 22 / 24-slice z-chunks 2.467 / 2.457 / 2.466 / 2.446
// / 2.462 (22 is the automatic choice; 4-row tiles: 8 / 12 / 32 slices 9.4 / 2.4 / 1.3 % slower
// than 16, 20 slices 0.4-0.7 % faster).
constexpr int kPairTY = 7;
// Rows per wave tile of the difference-form sweep (160 VGPRs, three waves per SIMD; 5 and 6 rows
// spill there, and 6 rows at two waves per SIMD take 2.956 against 2.861 ms per step).
constexpr int kDiffTY = 4;
// Block shapes (WX x-tiles by WY y-tiles of waves, x fastest): 2 x 2 for the plain two-column
// sweep, 1 x 4 for the difference form. Per step on the 2048x1536x768 room, one buffer set,
// shapes alternating every 20 launches: plain, 7-row tiles, 2x2 2.574 ms, 2x1 2.632, 4x1 2.747,
// 1x2 2.857, 1x4 2.868, 1x1 2.878 (4-row tiles, three allocation offsets: 2x2 2.66, 2x1 2.69,
// 1x4 2.73, 1x2 2.74, 1x1 2.81, 4x1 2.88, 2x4 2.95); difference form 1x4 and 2x2 2.86 (within
// 0.3 %), 1x2 2.89, 1x1 2.93. A wave tile past the grid in x (17 tiles of 124 columns on that
// room) computes on clamped columns and stores nothing; returning early from it, with the
// block's per-slice barriers kept, measured slower.
constexpr int kPairWX = 2;
constexpr int kPairWY = 2;
constexpr int kDiffWX = 1;
constexpr int kDiffWY = 4;

// The two-column sweep of plain 16-bit storage (P^n in A, P^{n-1} in B, P^{n+1} to Cn, P^{n+2}
// to Dn). Per-node selects and lane-masked stores, row offsets in registers, no prefetch: on the
// 2048x1536x768 room this form is 5 % faster than the branch-free prefetching form of the
// difference-form kernel below (which issues every buffer access with a scalar row offset, two
// issue cycles each, to fit its registers).
template <typename S, int WX, int WY>
__global__ __launch_bounds__(64 * WX * WY, 3 / ((WX * WY + 3) / 4)) void srl_update2_pair_kernel(
    Fused2Ctx<S, CoefF, float> p) {
#pragma clang fp contract(off)
  typedef float C;
  constexpr int TY = kPairTY;
  constexpr int NA = TY + 4;  // P^n rows (2-voxel y halo)
  constexpr int NQ = TY + 2;  // P^{n+1} rows (1-voxel y halo)
  __shared__ int s_src[WX * WY][kFListCap];
  __shared__ int s_rec[WX * WY][kFListCap];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  Fused2Args<S, CoefF, float> a = p.a;
  const int X = p.X, Y = p.Y, Z = p.Z, zc = p.zc;
  // XCD-aware tile order, as in srl_update2_kernel (grouped-batch slots innermost)
  const unsigned nx = gridDim.x, ny = gridDim.y, total = nx * ny * gridDim.z;
  const unsigned lin = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const unsigned full = total & ~7u;
  unsigned tile = lin < full ? (lin & 7u) * (full >> 3) + (lin >> 3) : lin;
  const int slot = (int)(tile % (unsigned)a.n_slots);
  tile /= (unsigned)a.n_slots;
  if (slot) {
    const int64_t off = (int64_t)slot * a.slot_elems;
    a.A += off;
    a.B += off;
    a.Cn += off;
    a.Dn += off;
  }
  // a block holds WX x WY wave tiles, x fastest; a tile past the grid computes and stores
  // nothing (its columns clamp to X - 2 and no lane owns a column)
  const int xt = (int)(tile % nx) * WX + wave % WX;
  const int byi = (int)((tile / nx) % ny);
  const int bzi = (int)(tile / (nx * ny));
  const int xb = xt * kPO;               // first owned column
  const int x0 = xb - 2 + 2 * lane;      // this lane's columns: x0 and x0 + 1
  const int yt = (byi * WY + wave / WX) * TY;  // first owned row
  const int z0 = 1 + bzi * zc;
  const int z1 = min(z0 + zc, Z - 1);

  extern __shared__ __align__(16) unsigned char s_dyn[];
  CoefF *s_coef = reinterpret_cast<CoefF *>(s_dyn);
  const unsigned ncol = (unsigned)a.ncol;
  for (unsigned i = threadIdx.x; i < kStageRows * ncol; i += 64 * WX * WY)
    s_coef[i] = a.coef[(i / ncol) * 256u + i % ncol];
  const int nsrc = wave_collect(a.src_xyz, a.n_src, xb - 1, xb + kPO + 1, yt - 1, yt + TY + 1,
                                z0 - 1, z1 + 1, s_src[wave], lane, a.src_slot, slot);
  const int nrec = wave_collect(a.rec_xyz, a.n_rec, xb, xb + kPO, yt, yt + TY, z0, z1,
                                s_rec[wave], lane, a.rec_slot, slot);
  __syncthreads();

  const size_t XY = (size_t)X * (size_t)Y;
  const uint32_t sbytes = (uint32_t)(XY * sizeof(S));  // host guarantees < 2^32
  const int xc = min(max(x0, 0), X - 2);  // an even column of the grid
  uint32_t ro[NA];
  uint32_t row_upd = 0;
#pragma unroll
  for (int r = 0; r < NA; ++r) {
    const int y = yt - 2 + r;
    ro[r] = (uint32_t)(min(max(y, 0), Y - 1) * X + xc);
    if (y > 0 && y < Y - 1) row_upd |= 1u << r;
  }
  const bool cu0 = x0 > 0 && x0 < X - 1;
  const bool cu1 = x0 + 1 > 0 && x0 + 1 < X - 1;
  const bool owned_lane = lane >= 1 && lane <= 62 && x0 < X;
  uint32_t row_st = 0;
#pragma unroll
  for (int o = 0; o < TY; ++o)
    if (yt + o < Y) row_st |= 1u << o;
  const double inv_scale = ldexp(1.0, -a.scale_exp);
  const C scale = (C)ldexp(1.0, a.scale_exp);
  const CoefF air = a.air;
  const C lam2 = a.lam2;
  // Plain air with the entry (inv, omb, tk) = (1, 1, 0) -- every air voxel when lambda^2 rounds
  // to 1/3 in FP32, since 2 - 6 lambda^2 is then exactly 0 -- takes
  //   fma(0, p, lambda^2 s) - p_past,
  // the same bits as inv*((tk*p + lambda^2 s) - omb*p_past): the products by 1 are exact, and the
  // product 0*p is exact, so the fused add rounds exactly as the separate one (signed zeros
  // included). Otherwise every slice takes the table.
  const bool air_fast = a.air_uniform != 0 && air.inv == 1.f && air.omb == 1.f && air.tk == 0.f;
  auto node = [&](auto table, unsigned pm, C pc, C s, C past) -> C {
    if constexpr (decltype(table)::value) {
      const unsigned pb = pm & 0xFFu;
      const CoefF e = s_coef[min(pb & 0x7Fu, 6u) * ncol + (pm >> 8)];
      const C aa = (C)(pb >> 7) * e.inv;
      return aa * ((e.tk * pc + lam2 * s) - e.omb * past);
    } else {
      return __builtin_fmaf(0.f, pc, lam2 * s) - past;
    }
  };
  auto any_nonair = [&](const uint32_t (&b)[NQ]) -> bool {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < NQ; ++i) m |= ((b[i] & 0xFFu) ^ 0x86u) | (((b[i] >> 16) & 0xFFu) ^ 0x86u);
    return __ballot(m != 0u) != 0;
  };
  const S *__restrict__ gA = a.A;
  const S *__restrict__ gB = a.B;
  S *__restrict__ gC = a.Cn;
  S *__restrict__ gD = a.Dn;
  const uint16_t *__restrict__ gpm = a.pm;
  typedef const __attribute__((address_space(4))) uint32_t *ConstU32;
  const uint32_t *__restrict__ gar = a.air_rows;
  const int nyw = Y / 32 + 2;
  const size_t ar_slice = (size_t)((X + kPO - 1) / kPO) * (size_t)nyw;
  const bool past_grid = yt >= Y || xb >= X;
  const uint32_t *ar_tile = gar && !past_grid ? gar + (size_t)xt * nyw + (yt >> 5) : nullptr;
  auto air_window = [&](int k) -> uint64_t {
    if (!gar || k < 1 || k > Z - 2) return 0;
    if (past_grid) return ~0ull;
    const ConstU32 w = (ConstU32)(ar_tile + (size_t)k * ar_slice);
    return (((uint64_t)w[1] << 32) | w[0]) >> (yt & 31);
  };
  constexpr uint64_t kAllRows = (1ull << NQ) - 1ull;
  uint64_t ar_win = air_window(z0 - 1);

  // P^n rows as (column x0, column x0+1) pairs of FP32 values, P^{n+1} likewise, packed
  // position|material of both columns in one register per row.
  // The first level at slice k reads A[k+1], B[k] and the packed plane at k: fetch() issues
  // those loads into raw 32-bit registers (pairs of 16-bit values). Issuing slice k+1's loads
  // one body ahead (4-row tiles: 20 more VGPRs, three waves per SIMD instead of four) was 2 %
  // slower.
  auto fetch = [&](int k, uint32_t (&ua)[NA], uint32_t (&ub)[NQ], uint32_t (&um)[NQ],
                   bool &uair, uint64_t win) {
    const bool upd = k >= 1 && k <= Z - 2;
    if (k + 1 <= Z - 1) {
      const auto rA = slice_rsrc(gA + (size_t)(k + 1) * XY, sbytes);
#pragma unroll
      for (int r = 0; r < NA; ++r) ua[r] = bld_pair(rA, ro[r]);
    } else {
#pragma unroll
      for (int r = 0; r < NA; ++r) ua[r] = 0;
    }
    uair = upd && (win & kAllRows) == kAllRows;
    if (upd) {
      const auto rB = slice_rsrc(gB + (size_t)k * XY, sbytes);
#pragma unroll
      for (int j = 0; j < NQ; ++j) ub[j] = bld_pair(rB, ro[j + 1]);
      if (uair) {
#pragma unroll
        for (int j = 0; j < NQ; ++j) um[j] = 0x00860086u;
      } else {
        const auto rPM = slice_rsrc(gpm + (size_t)k * XY, (uint32_t)(XY * 2));
#pragma unroll
        for (int j = 0; j < NQ; ++j) um[j] = bld_pair(rPM, ro[j + 1]);
      }
    } else {
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        ub[j] = 0;
        um[j] = 0;
      }
    }
  };
  auto body = [&](Pair2 (&am)[NA], Pair2 (&ac)[NA], Pair2 (&ap)[NA], Pair2 (&q0)[NQ],
                  Pair2 (&q1)[NQ], Pair2 (&q2)[NQ], uint32_t (&bp)[NQ], uint32_t (&bc)[NQ],
                  int k) {
    uint32_t ru = __builtin_amdgcn_readfirstlane(row_upd);
    uint32_t rs = __builtin_amdgcn_readfirstlane(row_st);
    asm volatile("" : "+s"(ru), "+s"(rs));
    __builtin_amdgcn_s_barrier();
    const bool upd1 = k >= 1 && k <= Z - 2;
    uint32_t ua[NA], pu[NQ];
    bool air1;
    fetch(k, ua, pu, bc, air1, ar_win);  // ar_win: the air-row window of slice k
    ar_win = air_window(k + 1);
    // every load of the slice is in flight before any result is combined
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < NA; ++r) ap[r] = Pair2{lo_half<S>(ua[r]), hi_half<S>(ua[r])};

    // ---- first time level: P^{n+1}[k]
    auto level1 = [&](auto table) {
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        const int r = j + 1;
        C s0 = ap[r].x + am[r].x;            // z+1, z-1
        s0 = s0 + ac[r + 1].x;               // y+1
        s0 = s0 + ac[r - 1].x;               // y-1
        s0 = s0 + ac[r].y;                   // x+1: the pair's other column
        s0 = s0 + lane_prev(ac[r].y);        // x-1: the previous lane's upper column
        C s1 = ap[r].y + am[r].y;
        s1 = s1 + ac[r + 1].y;
        s1 = s1 + ac[r - 1].y;
        s1 = s1 + lane_next(ac[r].x);        // x+1: the next lane's lower column
        s1 = s1 + ac[r].x;                   // x-1
        const C pp0 = lo_half<S>(pu[j]);
        const C pp1 = hi_half<S>(pu[j]);
        const C w0 = node(table, bc[j] & 0xFFFFu, ac[r].x, s0, pp0);
        const C w1 = node(table, bc[j] >> 16, ac[r].y, s1, pp1);
        const bool row = upd1 && ((ru >> r) & 1u);
        q2[j].x = (row && cu0) ? Fmt<S>::round(w0) : (C)0;
        q2[j].y = (row && cu1) ? Fmt<S>::round(w1) : (C)0;
      }
    };
    if (!air_fast || (!air1 && any_nonair(bc))) {
      level1(std::true_type());
    } else {
      level1(std::false_type());
    }

    // ---- receivers (before injection) and sources on P^{n+1}[k]
    if (nrec != 0) {
      const int cnt = nrec < 0 ? a.n_rec : nrec;
      for (int t = 0; t < cnt; ++t) {
        const int id = nrec < 0 ? t : s_rec[wave][t];
        const int32_t *q = a.rec_xyz + 3 * id;
        if (q[2] != k || (q[0] & ~1) != x0) continue;
        if (nrec < 0 && (!in_box(q, xb, xb + kPO, yt, yt + TY, z0, z1) ||
                         (a.rec_slot && a.rec_slot[id] != slot)))
          continue;
        const bool hi = q[0] & 1;
#pragma unroll
        for (int j = 0; j < NQ; ++j)
          if (q[1] == yt - 1 + j)
            a.resp[(int64_t)id * a.resp_steps + a.gather_step] =
                (double)pick(q2[j], hi) * inv_scale;
      }
    }
    if (nsrc != 0 && a.inject_step >= 0) {
      const int cnt = nsrc < 0 ? a.n_src : nsrc;
      for (int t = 0; t < cnt; ++t) {
        const int id = nsrc < 0 ? t : s_src[wave][t];
        const int32_t *q = a.src_xyz + 3 * id;
        if (q[2] != k || (q[0] & ~1) != x0) continue;
        if (nsrc < 0 && a.src_slot && a.src_slot[id] != slot) continue;
        C smp = (C)0;
        if (a.inject_step < a.n_samples) smp = a.sig_f[(int64_t)id * a.n_samples + a.inject_step];
        smp = smp * scale;
        const bool hard = a.src_kind[id] == 0 || a.soft_overwrite;
        const bool hi = q[0] & 1;
        // value selects only: a reference chosen at run time would put the queue in memory
#pragma unroll
        for (int j = 0; j < NQ; ++j)
          if (q[1] == yt - 1 + j) {
            const C old = pick(q2[j], hi);
            const C v = hard ? Fmt<S>::round(smp) : Fmt<S>::round(old + smp);
            q2[j].x = hi ? q2[j].x : v;
            q2[j].y = hi ? v : q2[j].y;
          }
      }
    }

    // ---- store the owned part of P^{n+1}[k]
    if (k >= z0 && k < z1) {
      const auto rC = slice_rsrc(gC + (size_t)k * XY, sbytes);
#pragma unroll
      for (int o = 0; o < TY; ++o)
        if (((rs >> o) & 1u) && owned_lane)
          bst_pair(rC, ro[o + 2], pack_halves<S>(q2[o + 1].x, q2[o + 1].y));
    }

    // ---- second time level: P^{n+2}[k-1] on the owned rows
    if (k - 1 >= z0) {
      const auto rD = slice_rsrc(gD + (size_t)(k - 1) * XY, sbytes);
      auto level2 = [&](auto table) {
#pragma unroll
        for (int o = 0; o < TY; ++o) {
          const int j = o + 1;
          C s0 = q2[j].x + q0[j].x;          // z+1, z-1
          s0 = s0 + q1[j + 1].x;             // y+1
          s0 = s0 + q1[j - 1].x;             // y-1
          s0 = s0 + q1[j].y;                 // x+1
          s0 = s0 + lane_prev(q1[j].y);      // x-1
          C s1 = q2[j].y + q0[j].y;
          s1 = s1 + q1[j + 1].y;
          s1 = s1 + q1[j - 1].y;
          s1 = s1 + lane_next(q1[j].x);
          s1 = s1 + q1[j].x;
          C v0 = node(table, bp[j] & 0xFFFFu, q1[j].x, s0, am[o + 2].x);
          C v1 = node(table, bp[j] >> 16, q1[j].y, s1, am[o + 2].y);
          const bool row = (ru >> (o + 2)) & 1u;
          v0 = (row && cu0) ? v0 : (C)0;
          v1 = (row && cu1) ? v1 : (C)0;
          if (((rs >> o) & 1u) && owned_lane) bst_pair(rD, ro[o + 2], pack_halves<S>(v0, v1));
        }
      };
      if (!air_fast || any_nonair(bp)) {
        level2(std::true_type());
      } else {
        level2(std::false_type());
      }
    }
  };

  Pair2 A0[NA], A1[NA], A2[NA];
  Pair2 Q0[NQ], Q1[NQ], Q2[NQ];
  uint32_t B0[NQ], B1[NQ], B2[NQ];
  const int k0 = z0 - 1;
  const int k1 = z1;
  {
    const auto rK = slice_rsrc(gA + (size_t)k0 * XY, sbytes);
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      const uint32_t u = bld_pair(rK, ro[r]);
      A1[r] = Pair2{lo_half<S>(u), hi_half<S>(u)};
    }
    if (k0 >= 1) {
      const auto rKm = slice_rsrc(gA + (size_t)(k0 - 1) * XY, sbytes);
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        const uint32_t u = bld_pair(rKm, ro[r]);
        A0[r] = Pair2{lo_half<S>(u), hi_half<S>(u)};
      }
    } else {
#pragma unroll
      for (int r = 0; r < NA; ++r) A0[r] = Pair2{0.f, 0.f};
    }
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      Q0[j] = Pair2{0.f, 0.f};
      Q1[j] = Pair2{0.f, 0.f};
      B0[j] = 0;
    }
  }
  for (int k = k0;;) {
    body(A0, A1, A2, Q0, Q1, Q2, B0, B1, k);
    if (++k > k1) break;
    body(A1, A2, A0, Q1, Q2, Q0, B1, B2, k);
    if (++k > k1) break;
    body(A2, A0, A1, Q2, Q0, Q1, B2, B0, k);
    if (++k > k1) break;
  }
}

// The difference form (PFDTD_F16_DIFF). A holds P^n and B holds D^n = P^n - P^{n-1}; the
// past value of a node is P^n - D^n (FP32), and each level stores fl16(ret) and fl16(ret - P),
// ret the FP32 result: P^{n+2} into Dn and D^{n+2} into Cn (both at the second level), or with
// one level (a.levels == 1) P^{n+1} into Dn and D^{n+1} into Cn. A source sample keeps the past
// value: D = fl16(P_new - (P_old - D_old)). The level-1 differences of the owned rows stay in
// registers for the second level (dcur, dprev), so this variant runs at three waves per SIMD.
// It is written branch-free: every lane evaluates every node and edge waves zero the nodes
// outside the updated range afterwards, masked stores drop through the buffer range check, the
// loads of the next slice are issued as soon as the first level has consumed its own, and the
// row offsets are vector registers (with scalar row offsets every buffer access takes two issue
// cycles; 1.4 % slower). At three waves per SIMD this is 9 % faster than the selects and
// lane-masked stores of the plain kernel above.
template <int WX, int WY>
__global__ __launch_bounds__(64 * WX * WY, 3 / ((WX * WY + 3) / 4)) void srl_update2_pair_diff_kernel(
    Fused2Ctx<h16, CoefF, float> p) {
#pragma clang fp contract(off)
  typedef h16 S;
  typedef float C;
  constexpr int TY = kDiffTY;
  constexpr int NA = TY + 4;  // P^n rows (2-voxel y halo)
  constexpr int NQ = TY + 2;  // P^{n+1} rows (1-voxel y halo)
  __shared__ int s_src[WX * WY][kFListCap];
  __shared__ int s_rec[WX * WY][kFListCap];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
  Fused2Args<S, CoefF, float> a = p.a;
  const int X = p.X, Y = p.Y, Z = p.Z, zc = p.zc;
  // XCD-aware tile order, as in srl_update2_kernel (grouped-batch slots innermost)
  const unsigned nx = gridDim.x, ny = gridDim.y, total = nx * ny * gridDim.z;
  const unsigned lin = blockIdx.x + nx * (blockIdx.y + ny * blockIdx.z);
  const unsigned full = total & ~7u;
  unsigned tile = lin < full ? (lin & 7u) * (full >> 3) + (lin >> 3) : lin;
  const int slot = (int)(tile % (unsigned)a.n_slots);
  tile /= (unsigned)a.n_slots;
  if (slot) {
    const int64_t off = (int64_t)slot * a.slot_elems;
    a.A += off;
    a.B += off;
    a.Cn += off;
    a.Dn += off;
  }
  // a block holds WX x WY wave tiles, x fastest; a tile past the grid computes and stores
  // nothing (its columns clamp to X - 2 and no lane owns a column)
  const int xt = (int)(tile % nx) * WX + wave % WX;
  const int byi = (int)((tile / nx) % ny);
  const int bzi = (int)(tile / (nx * ny));
  const int xb = xt * kPO;               // first owned column
  const int x0 = xb - 2 + 2 * lane;      // this lane's columns: x0 and x0 + 1
  const int yt = (byi * WY + wave / WX) * TY;  // first owned row
  const int z0 = 1 + bzi * zc;
  const int z1 = min(z0 + zc, Z - 1);
  const bool two = a.levels >= 2;  // the second time level (the odd last step takes one)

  extern __shared__ __align__(16) unsigned char s_dyn[];
  CoefF *s_coef = reinterpret_cast<CoefF *>(s_dyn);
  const unsigned ncol = (unsigned)a.ncol;
  for (unsigned i = threadIdx.x; i < kStageRows * ncol; i += 64 * WX * WY)
    s_coef[i] = a.coef[(i / ncol) * 256u + i % ncol];
  const int nsrc = wave_collect(a.src_xyz, a.n_src, xb - 1, xb + kPO + 1, yt - 1, yt + TY + 1,
                                z0 - 1, z1 + 1, s_src[wave], lane, a.src_slot, slot);
  const int nrec = wave_collect(a.rec_xyz, a.n_rec, xb, xb + kPO, yt, yt + TY, z0, z1,
                                s_rec[wave], lane, a.rec_slot, slot);
  __syncthreads();

  const size_t XY = (size_t)X * (size_t)Y;
  const uint32_t sbytes = (uint32_t)(XY * sizeof(S));  // host guarantees < 2^31
  const int xc = min(max(x0, 0), X - 2);  // an even column of the grid
  const uint32_t vcol = (uint32_t)xc * 2u;  // byte offset of the lane's pair within a row
  // byte offset of tile row r (rows outside the grid read the nearest grid row, never stored)
  auto srow = [&](int r) -> uint32_t {
    return (uint32_t)(min(max(yt - 2 + r, 0), Y - 1) * X) * 2u;
  };
  uint32_t row_upd = 0;
#pragma unroll
  for (int r = 0; r < NA; ++r) {
    const int y = yt - 2 + r;
    if (y > 0 && y < Y - 1) row_upd |= 1u << r;
  }
  // whether this lane's two columns are updated (not 0 or X-1): evaluated where needed, on the
  // edge waves only, from a copy of x0 the compiler cannot hoist, so that no register holds it
  // across the sweep
  auto col_upd = [&]() {
    int xl = x0;
    asm volatile("" : "+v"(xl));
    return std::pair<bool, bool>(xl > 0 && xl < X - 1, xl + 1 > 0 && xl + 1 < X - 1);
  };
  const bool owned_lane = lane >= 1 && lane <= 62 && x0 < X;
  const uint32_t vst = owned_lane ? 0u : kDropShift;  // halo lanes, lanes past the grid
  uint32_t rob[NA];  // byte offset of the lane's pair in tile row r
#pragma unroll
  for (int r = 0; r < NA; ++r) rob[r] = srow(r) + vcol;
  const bool rows_in = yt + TY <= Y;  // every owned row lies in the grid
  // every column of the wave (xb-2 .. xb+125) lies in 1..X-2: no lane needs a column select
  const bool x_inner = xb - 2 >= 1 && xb + 2 * 64 - 3 <= X - 2;
  const double inv_scale = ldexp(1.0, -a.scale_exp);
  const C scale = (C)ldexp(1.0, a.scale_exp);
  const CoefF air = a.air;
  const C lam2 = a.lam2;
  // Plain air with the entry (inv, omb, tk) = (1, 1, 0) -- every air voxel when lambda^2 rounds
  // to 1/3 in FP32, since 2 - 6 lambda^2 is then exactly 0 -- takes
  //   fma(0, p, lambda^2 s) - p_past,
  // the same bits as inv*((tk*p + lambda^2 s) - omb*p_past): the products by 1 are exact, and the
  // product 0*p is exact, so the fused add rounds exactly as the separate one (signed zeros
  // included). Otherwise every slice takes the table.
  const bool air_fast = a.air_uniform != 0 && air.inv == 1.f && air.omb == 1.f && air.tk == 0.f;
  auto node = [&](auto table, unsigned pm, C pc, C s, C past) -> C {
    if constexpr (decltype(table)::value) {
      const unsigned pb = pm & 0xFFu;
      const CoefF e = s_coef[min(pb & 0x7Fu, 6u) * ncol + (pm >> 8)];
      const C aa = (C)(pb >> 7) * e.inv;
      return aa * ((e.tk * pc + lam2 * s) - e.omb * past);
    } else {
      return __builtin_fmaf(0.f, pc, lam2 * s) - past;
    }
  };
  auto any_nonair = [&](const uint32_t (&b)[NQ]) -> bool {
    uint32_t m = 0;
#pragma unroll
    for (int i = 0; i < NQ; ++i) m |= ((b[i] & 0xFFu) ^ 0x86u) | (((b[i] >> 16) & 0xFFu) ^ 0x86u);
    return __ballot(m != 0u) != 0;
  };
  const S *__restrict__ gA = a.A;
  const S *__restrict__ gB = a.B;
  S *__restrict__ gC = a.Cn;
  S *__restrict__ gD = a.Dn;
  const uint16_t *__restrict__ gpm = a.pm;
  typedef const __attribute__((address_space(4))) uint32_t *ConstU32;
  const uint32_t *__restrict__ gar = a.air_rows;
  const int nyw = Y / 32 + 2;
  const size_t ar_slice = (size_t)((X + kPO - 1) / kPO) * (size_t)nyw;
  const bool past_grid = yt >= Y || xb >= X;
  const uint32_t *ar_tile = gar && !past_grid ? gar + (size_t)xt * nyw + (yt >> 5) : nullptr;
  auto air_window = [&](int k) -> uint64_t {
    if (!gar || k < 1 || k > Z - 2) return 0;
    if (past_grid) return ~0ull;
    const ConstU32 w = (ConstU32)(ar_tile + (size_t)k * ar_slice);
    return (((uint64_t)w[1] << 32) | w[0]) >> (yt & 31);
  };
  constexpr uint64_t kAllRows = (1ull << NQ) - 1ull;
  Pair2 dprev[TY];  // D^{n+1} of the owned rows at the slice the second level updates
#pragma unroll
  for (int o = 0; o < TY; ++o) dprev[o] = Pair2{0.f, 0.f};

  // P^n rows as (column x0, column x0+1) pairs of FP32 values, P^{n+1} likewise, packed
  // position|material of both columns in one register per row. The first level at slice k
  // reads P^n[k+1] (fetch_a), D^n[k] (fetch_b) and the packed plane at k (fetch_pm: constant
  // air values when the air-row bits say every row is plain air), as raw 32-bit registers.
  auto fetch_a = [&](int k, uint32_t (&ua)[NA]) {
    if (k + 1 <= Z - 1) {
      const auto rA = slice_rsrc(gA + (size_t)(k + 1) * XY, sbytes);
#pragma unroll
      for (int r = 0; r < NA; ++r) ua[r] = bld_pair_at(rA, rob[r]);
    } else {
#pragma unroll
      for (int r = 0; r < NA; ++r) ua[r] = 0;
    }
  };
  auto fetch_b = [&](int k, uint32_t (&ub)[NQ]) {
    const bool upd = k >= 1 && k <= Z - 2;
    if (upd) {
      const auto rB = slice_rsrc(gB + (size_t)k * XY, sbytes);
#pragma unroll
      for (int j = 0; j < NQ; ++j) ub[j] = bld_pair_at(rB, rob[j + 1]);
    } else {
#pragma unroll
      for (int j = 0; j < NQ; ++j) ub[j] = 0;
    }
  };
  auto fetch_pm = [&](int k, uint32_t (&um)[NQ], bool &uair, uint64_t win) {
    const bool upd = k >= 1 && k <= Z - 2;
    uair = upd && (win & kAllRows) == kAllRows;
    if (uair) {
#pragma unroll
      for (int j = 0; j < NQ; ++j) um[j] = 0x00860086u;
    } else if (upd) {
      const auto rPM = slice_rsrc(gpm + (size_t)k * XY, (uint32_t)(XY * 2));
#pragma unroll
      for (int j = 0; j < NQ; ++j) um[j] = bld_pair_at(rPM, rob[j + 1]);
    } else {
#pragma unroll
      for (int j = 0; j < NQ; ++j) um[j] = 0;
    }
  };
  // Software pipeline across slices: the loads the first level of slice k+1 reads (P^n[k+2],
  // D^n[k+1], the packed plane at k+1) are issued as soon as the first level of slice k has
  // consumed its own, into the same registers (ua, pu: raw 32-bit pairs; the packed plane into
  // the third array of the rotation), so they are in flight while this wave runs its second
  // level and stores. air_k: whether slice k's rows are all plain air.
  uint32_t ua[NA], pu[NQ];
  bool air_k = false;
  auto body = [&](Pair2 (&am)[NA], Pair2 (&ac)[NA], Pair2 (&ap)[NA], Pair2 (&q0)[NQ],
                  Pair2 (&q1)[NQ], Pair2 (&q2)[NQ], uint32_t (&bp)[NQ], uint32_t (&bc)[NQ],
                  uint32_t (&bn)[NQ], int k) {
    uint32_t ru = __builtin_amdgcn_readfirstlane(row_upd);
    asm volatile("" : "+s"(ru));
    __builtin_amdgcn_s_barrier();
    const bool upd1 = k >= 1 && k <= Z - 2;
    const bool air1 = air_k;
    const uint64_t win_next = air_window(k + 1);  // for the loads of slice k+1
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < NA; ++r) ap[r] = Pair2{lo_half<S>(ua[r]), hi_half<S>(ua[r])};

    // ---- first time level: P^{n+1}[k]
    Pair2 dcur[TY];  // D^{n+1}[k] of the owned rows
    auto level1 = [&](auto table) {
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        const int r = j + 1;
        C s0 = ap[r].x + am[r].x;            // z+1, z-1
        s0 = s0 + ac[r + 1].x;               // y+1
        s0 = s0 + ac[r - 1].x;               // y-1
        s0 = s0 + ac[r].y;                   // x+1: the pair's other column
        s0 = s0 + lane_prev(ac[r].y);        // x-1: the previous lane's upper column
        C s1 = ap[r].y + am[r].y;
        s1 = s1 + ac[r + 1].y;
        s1 = s1 + ac[r - 1].y;
        s1 = s1 + lane_next(ac[r].x);        // x+1: the next lane's lower column
        s1 = s1 + ac[r].x;                   // x-1
        const C pp0 = ac[r].x - lo_half<S>(pu[j]);  // the past value P^n - D^n
        const C pp1 = ac[r].y - hi_half<S>(pu[j]);
        // every lane evaluates every node (all operands come from in-range loads or read 0);
        // the nodes outside the updated range are zeroed afterwards
        const C w0 = node(table, bc[j] & 0xFFFFu, ac[r].x, s0, pp0);
        const C w1 = node(table, bc[j] >> 16, ac[r].y, s1, pp1);
        q2[j].x = Fmt<S>::round(w0);
        q2[j].y = Fmt<S>::round(w1);
        if (j >= 1 && j <= TY) {
          dcur[j - 1].x = Fmt<S>::round(w0 - ac[r].x);
          dcur[j - 1].y = Fmt<S>::round(w1 - ac[r].y);
        }
      }
    };
    if (!air_fast || (!air1 && any_nonair(bc))) {
      level1(std::true_type());
    } else {
      level1(std::false_type());
    }
    // Zero the nodes outside the updated range: the rows and slices the grid does not update,
    // and the columns 0 and X-1. Only waves at the grid's edges take this branch. Selects inside
    // the node loop (`row && cu ? round(w) : 0`) made every node a lane-masked branch of its
    // own -- exec-mask saves and restores per node, and DPP shifts that no longer fold into
    // their adds -- or, written with bitwise masks, cost five instructions per select.
    if (!(upd1 && x_inner && ((ru >> 1) & kAllRows) == kAllRows)) {
      const auto [cu0, cu1] = col_upd();
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        const bool row = upd1 && ((ru >> (j + 1)) & 1u);
        const bool m0 = row && cu0, m1 = row && cu1;
        q2[j].x = m0 ? q2[j].x : (C)0;
        q2[j].y = m1 ? q2[j].y : (C)0;
        if (j >= 1 && j <= TY) {
          dcur[j - 1].x = m0 ? dcur[j - 1].x : (C)0;
          dcur[j - 1].y = m1 ? dcur[j - 1].y : (C)0;
        }
      }
    }
    // the first level is done with ua / pu: issue slice k+1's loads into them
    __builtin_amdgcn_sched_barrier(0);
    if (k + 1 <= z1) {
      fetch_a(k + 1, ua);
      fetch_b(k + 1, pu);
      fetch_pm(k + 1, bn, air_k, win_next);
    }
    __builtin_amdgcn_sched_barrier(0);

    // ---- receivers (before injection) and sources on P^{n+1}[k]
    if (nrec != 0) {
      const int cnt = nrec < 0 ? a.n_rec : nrec;
      for (int t = 0; t < cnt; ++t) {
        const int id = nrec < 0 ? t : s_rec[wave][t];
        const int32_t *q = a.rec_xyz + 3 * id;
        if (q[2] != k || (q[0] & ~1) != x0) continue;
        if (nrec < 0 && (!in_box(q, xb, xb + kPO, yt, yt + TY, z0, z1) ||
                         (a.rec_slot && a.rec_slot[id] != slot)))
          continue;
        const bool hi = q[0] & 1;
#pragma unroll
        for (int j = 0; j < NQ; ++j)
          if (q[1] == yt - 1 + j)
            a.resp[(int64_t)id * a.resp_steps + a.gather_step] =
                (double)pick(q2[j], hi) * inv_scale;
      }
    }
    if (nsrc != 0 && a.inject_step >= 0) {
      const int cnt = nsrc < 0 ? a.n_src : nsrc;
      for (int t = 0; t < cnt; ++t) {
        const int id = nsrc < 0 ? t : s_src[wave][t];
        const int32_t *q = a.src_xyz + 3 * id;
        if (q[2] != k || (q[0] & ~1) != x0) continue;
        if (nsrc < 0 && a.src_slot && a.src_slot[id] != slot) continue;
        C smp = (C)0;
        if (a.inject_step < a.n_samples) smp = a.sig_f[(int64_t)id * a.n_samples + a.inject_step];
        smp = smp * scale;
        const bool hard = a.src_kind[id] == 0 || a.soft_overwrite;
        const bool hi = q[0] & 1;
        // value selects only: a reference chosen at run time would put the queue in memory
#pragma unroll
        for (int j = 0; j < NQ; ++j)
          if (q[1] == yt - 1 + j) {
            const C old = pick(q2[j], hi);
            const C v = hard ? Fmt<S>::round(smp) : Fmt<S>::round(old + smp);
            q2[j].x = hi ? q2[j].x : v;
            q2[j].y = hi ? v : q2[j].y;
            if (j >= 1 && j <= TY) {  // owned rows: keep the past value
              const C past = old - pick(dcur[j - 1], hi);
              const C d = Fmt<S>::round(v - past);
              dcur[j - 1].x = hi ? dcur[j - 1].x : d;
              dcur[j - 1].y = hi ? d : dcur[j - 1].y;
            }
          }
      }
    }

    // ---- one level only (the odd last step): P^{n+1}[k] into Dn and D^{n+1}[k] into Cn; with
    // two levels the second level stores
    if (k >= z0 && k < z1 && !two) {
      const auto rC = slice_rsrc(gD + (size_t)k * XY, sbytes);
      const auto rDd = slice_rsrc(gC + (size_t)k * XY, sbytes);
      auto store_rows = [&](int nrows) {
#pragma unroll
        for (int o = 0; o < TY; ++o) {
          if (o >= nrows) break;
          bst_pair_at(rC, rob[o + 2] | vst, pack_exact<S>(q2[o + 1].x, q2[o + 1].y));
          bst_pair_at(rDd, rob[o + 2] | vst, pack_exact<S>(dcur[o].x, dcur[o].y));
        }
      };
      if (rows_in)
        store_rows(TY);
      else  // the last row of wave tiles: rows past the grid skip their stores
        store_rows(Y - yt);
    }

    // ---- second time level: P^{n+2}[k-1] on the owned rows
    if (two && k - 1 >= z0) {
      const auto rD = slice_rsrc(gD + (size_t)(k - 1) * XY, sbytes);
      const auto rCd = slice_rsrc(gC + (size_t)(k - 1) * XY, sbytes);
      constexpr uint32_t kOwnedRows = (1u << TY) - 1u;
      const bool inner2 = x_inner && ((ru >> 2) & kOwnedRows) == kOwnedRows;
      auto level2 = [&](auto table) {
#pragma unroll
        for (int o = 0; o < TY; ++o) {
          const int j = o + 1;
          C s0 = q2[j].x + q0[j].x;          // z+1, z-1
          s0 = s0 + q1[j + 1].x;             // y+1
          s0 = s0 + q1[j - 1].x;             // y-1
          s0 = s0 + q1[j].y;                 // x+1
          s0 = s0 + lane_prev(q1[j].y);      // x-1
          C s1 = q2[j].y + q0[j].y;
          s1 = s1 + q1[j + 1].y;
          s1 = s1 + q1[j - 1].y;
          s1 = s1 + lane_next(q1[j].x);
          s1 = s1 + q1[j].x;
          const C pp0 = q1[j].x - dprev[o].x;  // the past value P^{n+1} - D^{n+1}
          const C pp1 = q1[j].y - dprev[o].y;
          C v0 = node(table, bp[j] & 0xFFFFu, q1[j].x, s0, pp0);
          C v1 = node(table, bp[j] >> 16, q1[j].y, s1, pp1);
          C e0 = Fmt<S>::round(v0 - q1[j].x);
          C e1 = Fmt<S>::round(v1 - q1[j].y);
          uint32_t vs = vst;
          if (!inner2) {  // edge waves: zero the nodes outside the updated range
            const auto [cu0, cu1] = col_upd();
            const bool row = (ru >> (o + 2)) & 1u;
            const bool m0 = row && cu0, m1 = row && cu1;
            v0 = m0 ? v0 : (C)0;
            v1 = m1 ? v1 : (C)0;
            e0 = m0 ? e0 : (C)0;
            e1 = m1 ? e1 : (C)0;
            if (yt + o >= Y) vs |= kDropShift;  // a row past the grid: drop its stores
          }
          bst_pair_at(rCd, rob[o + 2] | vs, pack_exact<S>(e0, e1));
          bst_pair_at(rD, rob[o + 2] | vs, pack_halves<S>(v0, v1));
        }
      };
      if (!air_fast || any_nonair(bp)) {
        level2(std::true_type());
      } else {
        level2(std::false_type());
      }
    }
#pragma unroll
    for (int o = 0; o < TY; ++o) dprev[o] = dcur[o];
  };

  Pair2 A0[NA], A1[NA], A2[NA];
  Pair2 Q0[NQ], Q1[NQ], Q2[NQ];
  uint32_t B0[NQ], B1[NQ], B2[NQ];
  const int k0 = z0 - 1;
  const int k1 = z1;
  {
    fetch_a(k0, ua);  // the first body's slice
    fetch_b(k0, pu);
    fetch_pm(k0, B1, air_k, air_window(k0));
    const auto rK = slice_rsrc(gA + (size_t)k0 * XY, sbytes);
#pragma unroll
    for (int r = 0; r < NA; ++r) {
      const uint32_t u = bld_pair_at(rK, rob[r]);
      A1[r] = Pair2{lo_half<S>(u), hi_half<S>(u)};
    }
    if (k0 >= 1) {
      const auto rKm = slice_rsrc(gA + (size_t)(k0 - 1) * XY, sbytes);
#pragma unroll
      for (int r = 0; r < NA; ++r) {
        const uint32_t u = bld_pair_at(rKm, rob[r]);
        A0[r] = Pair2{lo_half<S>(u), hi_half<S>(u)};
      }
    } else {
#pragma unroll
      for (int r = 0; r < NA; ++r) A0[r] = Pair2{0.f, 0.f};
    }
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
      Q0[j] = Pair2{0.f, 0.f};
      Q1[j] = Pair2{0.f, 0.f};
      B0[j] = 0;
    }
  }
  for (int k = k0;;) {
    body(A0, A1, A2, Q0, Q1, Q2, B0, B1, B2, k);
    if (++k > k1) break;
    body(A1, A2, A0, Q1, Q2, Q0, B1, B2, B0, k);
    if (++k > k1) break;
    body(A2, A0, A1, Q2, Q0, Q1, B2, B0, B1, k);
    if (++k > k1) break;
  }
}

// Air-row plane: bit b of word (z * ntx + t) * nyw + w is set when every voxel of row
// y = 32 w + b - 1 of slice z in columns [60 t - 1, 60 t + 61) -- the columns whose first-level
// results wave tile t uses -- is plain air (packed value `air`: position 0x86 for the forward
// scheme, 0x80 for the centred one; material 0); rows outside the grid are set. One wave per
// word, lanes over the columns; blocks stride over words and slices.
// The same for the tiles of the two-column sweep (tile width kPO, window [124 t - 1, 124 t + 125)),
// and (h = 2) for those of the three-level sweep: rows y = 32 w + b - 2, columns
// [58 t - 2, 58 t + 60).
__global__ __launch_bounds__(64) void air_rows_kernel(const uint16_t *__restrict__ pm,
                                                      uint32_t *__restrict__ out, int X, int Y,
                                                      int Z, int ntx, int nyw, unsigned air,
                                                      int tw, int h) {
  const int t = blockIdx.x;
  const int lane = threadIdx.x;
  for (int z = blockIdx.z; z < Z; z += gridDim.z) {
    const uint16_t *slice = pm + (size_t)z * (size_t)X * (size_t)Y;
    for (int w = blockIdx.y; w < nyw; w += gridDim.y) {
      uint32_t word = 0;
      for (int b = 0; b < 32; ++b) {
        const int y = 32 * w + b - h;
        bool other = false;
        for (int c = lane; c < tw + 2 * h; c += 64) {
          const int x = t * tw - h + c;
          other = other || (x >= 0 && x < X && y >= 0 && y < Y && slice[(size_t)y * X + x] != air);
        }
        if (__ballot(other) == 0) word |= 1u << b;
      }
      if (lane == 0) out[((size_t)z * ntx + t) * nyw + w] = word;
    }
  }
}

static int air_rows_tile(int kind) { return kind == 2 ? kFO3 : kind == 1 ? kPO : kFO; }

size_t air_rows_words(int X, int Y, int Z, int kind) {
  const int tw = air_rows_tile(kind);
  return (size_t)((X + tw - 1) / tw) * (size_t)(Y / 32 + 2) * (size_t)Z;
}

hipError_t launch_air_rows(const uint16_t *pm, uint32_t *rows, int X, int Y, int Z, int scheme,
                           int kind, hipStream_t stream) {
  const int tw = air_rows_tile(kind);
  const int ntx = (X + tw - 1) / tw, nyw = Y / 32 + 2;
  const dim3 grid((unsigned)ntx, (unsigned)std::min(nyw, 4096), (unsigned)std::min(Z, 4096));
  const unsigned air = scheme == 2 ? 0x0080u : 0x0086u;
  hipLaunchKernelGGL(air_rows_kernel, grid, dim3(64), 0, stream, pm, rows, X, Y, Z, ntx, nyw, air,
                     tw, kind == 2 ? 2 : 1);
  return hipGetLastError();
}

FusedGeom plan_fused(int X, int Y, int Z, int zc, int precision, bool pool, int shape,
                     int scheme, int slots) {
  const int kFTY = fused_ty_host(precision, scheme);
  FusedGeom g;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  const int zi = Z - 2;
  const int ntx = (X + kFO - 1) / kFO;
  // automatic on rooms at least 16 x tiles of 60 wide (more than 900 columns), where at most
  // 3 of 16 tiles fall past the grid: wide blocks for FP32 with the forward scheme, square blocks for FP64
  // with the forward scheme and for FP32 with the centred one
  g.scheme = scheme;
  g.blocks = kColumnBlocks;
  if (!pool && precision == 0 && scheme == 0 &&
      (shape == PFDTD_SWEEP_WIDE || (shape == PFDTD_SWEEP_AUTO && ntx >= 16)))
    g.blocks = kWideBlocks;
  else if (!pool && shape == PFDTD_SWEEP_AUTO && ntx >= 16 &&
           ((precision == 1 && scheme == 0) || (precision == 0 && scheme == 2)))
    g.blocks = kSquareBlocks;
  // 16-bit storage, forward scheme, X even, slices under 2^31 bytes: two columns per lane
  g.pair = !pool && precision >= 2 && scheme == 0 && X % 2 == 0 &&
           (int64_t)X * Y * 2 < (int64_t)kDropShift &&
           (shape == PFDTD_SWEEP_AUTO || shape == PFDTD_SWEEP_PAIR);
  if (g.pair) {
    const int wx = precision == PFDTD_F16_DIFF ? kDiffWX : kPairWX;
    const int wy = precision == PFDTD_F16_DIFF ? kDiffWY : kPairWY;
    const int rows = (precision == PFDTD_F16_DIFF ? kDiffTY : kPairTY) * wy;
    const int pbx = ((X + kPO - 1) / kPO + wx - 1) / wx;
    const int64_t pxy = (int64_t)pbx * ((Y + rows - 1) / rows) * slots;
    if (zc < 1) {
      // 22 slices (the difference form 16), then 16 and 8, while the sweep has fewer than 4096
      // blocks; on 2048x1536x768 the plain sweep with 7-row tiles takes 2.554-2.555 ms per step
      // with 22-slice chunks against 2.576-2.578 with 20 (BF16 2.754 against 2.794-2.801; same
      // buffers, two processes)
      zc = precision == PFDTD_F16_DIFF ? 16 : 22;
      while (zc > 8 && pxy * ((zi + zc - 1) / zc) < 4096) zc = zc == 22 ? 16 : zc / 2;
    }
    g.zc = zc > zi ? zi : zc;
    g.grid = dim3(pbx, (Y + rows - 1) / rows, ((zi + g.zc - 1) / g.zc) * slots);
    g.block = dim3(64 * wx * wy, 1, 1);
    return g;
  }
  const int WX = g.blocks == kWideBlocks     ? FusedShape<kWideBlocks>::WX
                 : g.blocks == kSquareBlocks ? FusedShape<kSquareBlocks>::WX
                                             : FusedShape<kColumnBlocks>::WX;
  const int WY = g.blocks == kWideBlocks     ? FusedShape<kWideBlocks>::WY
                 : g.blocks == kSquareBlocks ? FusedShape<kSquareBlocks>::WY
                                             : FusedShape<kColumnBlocks>::WY;
  const int64_t bxy =
      (int64_t)((ntx + WX - 1) / WX) * ((Y + kFTY * WY - 1) / (kFTY * WY)) * slots;
  if (zc < 1) {
    // automatic: the longest of 32 (16 for FP64 and 2-byte fields), 16 and 8 slices that still
    // gives 4096 blocks (16 per CU), so small rooms keep every CU busy to the end of the sweep;
    // each chunk re-reads one slice of P^n on either side, so longer chunks move fewer bytes. On
    // the 2048x1536x768 room 16-slice chunks are 1 % faster than 32 with FP64 and 2-byte fields
    // (and 2 % slower with FP32).
    zc = precision == PFDTD_F32 ? 32 : 16;
    while (zc > 8 && bxy * ((zi + zc - 1) / zc) < 4096) zc /= 2;
  }
  g.zc = zc > zi ? zi : zc;
  g.grid = dim3((ntx + WX - 1) / WX, (Y + kFTY * WY - 1) / (kFTY * WY),
                ((zi + g.zc - 1) / g.zc) * slots);
  g.block = dim3(64 * WX * WY, 1, 1);
  return g;
}

template <typename S, typename CT, typename C>
static hipError_t launch_fused(const FusedGeom &g, const Fused2Args<S, CT, C> &a,
                               hipStream_t stream) {
  const size_t lds = sizeof(CT) * kStageRows * (size_t)a.ncol;
  const Fused2Ctx<S, CT, C> p{a, g.X, g.Y, g.Z, g.zc};
  if constexpr (sizeof(S) == 2) {
    if (g.pair) {
      if constexpr (std::is_same<S, h16>::value) {
        if (a.diff) {
          hipLaunchKernelGGL((srl_update2_pair_diff_kernel<kDiffWX, kDiffWY>), g.grid, g.block, lds, stream, p);
          return hipGetLastError();
        }
      }
      hipLaunchKernelGGL((srl_update2_pair_kernel<S, kPairWX, kPairWY>), g.grid, g.block, lds, stream, p);
      return hipGetLastError();
    }
  }
  if (g.scheme == 2) {
    if constexpr (std::is_same<S, float>::value) {
      if (g.blocks == kSquareBlocks) {
        hipLaunchKernelGGL((srl_update2_kernel<S, CT, C, false, kSquareBlocks, 2>), g.grid,
                           g.block, lds, stream, p);
        return hipGetLastError();
      }
    }
    hipLaunchKernelGGL((srl_update2_kernel<S, CT, C, false, kColumnBlocks, 2>), g.grid, g.block,
                       lds, stream, p);
    return hipGetLastError();
  }
  if constexpr (std::is_same<S, float>::value) {
    if (g.blocks == kWideBlocks) {
      hipLaunchKernelGGL((srl_update2_kernel<S, CT, C, false, kWideBlocks, 0>), g.grid, g.block,
                         lds, stream, p);
      return hipGetLastError();
    }
  }
  if constexpr (std::is_same<S, double>::value) {
    if (g.blocks == kSquareBlocks) {
      hipLaunchKernelGGL((srl_update2_kernel<S, CT, C, false, kSquareBlocks, 0>), g.grid,
                         g.block, lds, stream, p);
      return hipGetLastError();
    }
  }
  hipLaunchKernelGGL((srl_update2_kernel<S, CT, C, false, kColumnBlocks, 0>), g.grid, g.block,
                     lds, stream, p);
  return hipGetLastError();
}

hipError_t launch_pool_update2(int precision, int tiles, const PoolJobDev *jobs,
                               const PoolState *st, const PoolLaunch &L, int zc,
                               hipStream_t stream) {
  const Fused2Pool p{jobs, st, L, zc};
  const size_t lds = sizeof(CoefF) * kStageRows * (size_t)L.ncol_max;
  const dim3 grid((unsigned)tiles);
  switch (precision) {
  case PFDTD_F32:
    hipLaunchKernelGGL((srl_update2_kernel<float, CoefF, float, true, kColumnBlocks, 0>), grid,
                       dim3(64 * fused_waves<kColumnBlocks>()), lds, stream, p);
    break;
  case PFDTD_F16:
    hipLaunchKernelGGL((srl_update2_kernel<h16, CoefF, float, true, kColumnBlocks, 0>), grid,
                       dim3(64 * fused_waves<kColumnBlocks>()), lds, stream, p);
    break;
  case PFDTD_BF16:
    hipLaunchKernelGGL((srl_update2_kernel<b16, CoefF, float, true, kColumnBlocks, 0>), grid,
                       dim3(64 * fused_waves<kColumnBlocks>()), lds, stream, p);
    break;
  default:
    return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

hipError_t launch_update2_f32(const FusedGeom &g, const Fused2Args<float, CoefF, float> &a,
                              hipStream_t stream) {
  return launch_fused(g, a, stream);
}

FusedGeom plan_fused3(int X, int Y, int Z, int zc, int precision, int shape, int slots) {
  FusedGeom g;
  g.X = X;
  g.Y = Y;
  g.Z = Z;
  g.scheme = 0;
  g.pair = false;
  g.three = true;
  g.blocks = shape == PFDTD_SWEEP_COLUMN ? kColumnBlocks : kThreeBlocks;
  const int zi = Z - 2;
  const int ntx = (X + kFO3 - 1) / kFO3;
  const int WX = g.blocks == kColumnBlocks ? FusedShape<kColumnBlocks>::WX
                                            : FusedShape<kSquareBlocks>::WX;
  const int WY = g.blocks == kColumnBlocks ? FusedShape<kColumnBlocks>::WY
                                            : FusedShape<kSquareBlocks>::WY;
  const int64_t bxy =
      (int64_t)((ntx + WX - 1) / WX) * ((Y + kF3TY * WY - 1) / (kF3TY * WY)) * slots;
  if (zc < 1) {
    // automatic: the longest of 32, 16 and 8 slices (FP64: 96, 48, 24, 12) that gives 2048
    // blocks (each chunk also computes four slices of its neighbours, so shorter chunks cost
    // more than in the two-level sweep: on 512x384x256 16 slices are 7 % faster than 8, on
    // 130x98x66 8 slices 40 % faster than 16). On 2048x1536x768 FP32 takes 3.85 ms per step
    // with 32 slices (24 / 40 / 48: 3.90 / 3.94 / 3.90); FP64, at two waves per SIMD, 7.37-7.71
    // with 96 (32 / 48 / 64: 7.83-8.09 / 7.45-7.87 / 7.39-7.79).
    zc = precision == PFDTD_F64 ? 96 : 32;
    while (zc > 8 && bxy * ((zi + zc - 1) / zc) < 2048) zc /= 2;
  }
  g.zc = zc > zi ? zi : zc;
  g.grid = dim3((ntx + WX - 1) / WX, (Y + kF3TY * WY - 1) / (kF3TY * WY),
                ((zi + g.zc - 1) / g.zc) * slots);
  g.block = dim3(64 * WX * WY, 1, 1);
  return g;
}

template <typename S, typename CT, typename C>
static hipError_t launch_fused3(const FusedGeom &g, const Fused2Args<S, CT, C> &a,
                                hipStream_t stream) {
  const size_t lds = sizeof(CT) * kStageRows * (size_t)a.ncol;
  const Fused2Ctx<S, CT, C> p{a, g.X, g.Y, g.Z, g.zc};
  // square blocks (kThreeBlocks), or the column of 1 x 4 tiles with PFDTD_SWEEP_COLUMN
  // bands of three block rows for FP64, two otherwise
  constexpr int BG = sizeof(S) == 8 ? 3 : 2;
  if (g.blocks == kColumnBlocks)
    hipLaunchKernelGGL((srl_update3_kernel<S, CT, C, kColumnBlocks, BG>), g.grid, g.block, lds,
                       stream, p);
  else
    hipLaunchKernelGGL((srl_update3_kernel<S, CT, C, kSquareBlocks, BG>), g.grid, g.block, lds,
                       stream, p);
  return hipGetLastError();
}
hipError_t launch_update3_f32(const FusedGeom &g, const Fused2Args<float, CoefF, float> &a,
                              hipStream_t stream) {
  return launch_fused3(g, a, stream);
}
hipError_t launch_update3_f64(const FusedGeom &g, const Fused2Args<double, CoefD, double> &a,
                              hipStream_t stream) {
  return launch_fused3(g, a, stream);
}
hipError_t launch_update3_f16(const FusedGeom &g, const Fused2Args<h16, CoefF, float> &a,
                              hipStream_t stream) {
  return launch_fused3(g, a, stream);
}
hipError_t launch_update3_bf16(const FusedGeom &g, const Fused2Args<b16, CoefF, float> &a,
                               hipStream_t stream) {
  return launch_fused3(g, a, stream);
}
hipError_t launch_update2_f64(const FusedGeom &g, const Fused2Args<double, CoefD, double> &a,
                              hipStream_t stream) {
  return launch_fused(g, a, stream);
}
hipError_t launch_update2_f16(const FusedGeom &g, const Fused2Args<h16, CoefF, float> &a,
                              hipStream_t stream) {
  return launch_fused(g, a, stream);
}
hipError_t launch_update2_bf16(const FusedGeom &g, const Fused2Args<b16, CoefF, float> &a,
                               hipStream_t stream) {
  return launch_fused(g, a, stream);
}

}  // namespace pfdtd
