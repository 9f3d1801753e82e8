// pfdtd_internal.h -- declarations shared by the HIP kernel and C-ABI translation units.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

namespace pfdtd {

// Per-(K, material) update coefficients of the SRL_FORWARD rule
// (src/kernels/kernels3d.cu:508-528):
//   inv = 1/(1+beta_e), omb = 1-beta_e, tk = 2-K*lambda^2,
//   beta_e = 0.5*(beta[m]*(6-K)*lambda).
// Indexed [K*256 + material byte] for K in 0..6; row 7 is all zero and catches any K the
// domain check should already have rejected.
struct CoefF {
  float inv, omb, tk, pad;
};
struct CoefD {
  double inv, omb, tk, pad;
};
constexpr int kCoefRows = 8;
constexpr int kCoefEntries = kCoefRows * 256;
// Rows the update kernels stage in LDS: K = 0..6 (forward) or D = 0..3 (centred). Inside voxels
// never name K > 6 (the domain scan rejects them) and outside voxels are multiplied by 0, so
// the row index is clamped to kStageRows - 1; FP64 tables of 256 materials stay at 56 KB.
constexpr int kStageRows = 7;

// Storage tags for the half-width field formats.
struct h16 {
  uint16_t bits;
};
struct b16 {
  uint16_t bits;
};

// Result of the domain scan (validate_domain).
struct DomainScan {
  unsigned long long air;       // pos == 0x86
  unsigned long long boundary;  // pos != 0 && pos != 0x86
  unsigned int shell_inside;    // voxels on the outer layer with bit 7 set
  unsigned int bad_k;           // inside voxels with K > 6
  unsigned int mat_present[8];  // 256-bit set of material bytes that occur
};

// Update kernel launch geometry.
struct UpdateGeom {
  int X, Y, Z;
  int ry;       // y rows per thread (1, 2 or 4)
  int zc;       // interior slices per block
  int lgB;      // log2 of the batch width (0: unbatched)
  dim3 grid, block;
};
UpdateGeom plan_update(int X, int Y, int Z, int ry, int zc);
// Batched update: fields hold 2^lgB values per voxel, slot innermost.
UpdateGeom plan_update_batch(int X, int Y, int Z, int ry, int zc, int lgB);

// Arguments of one update launch: fields in storage type S, coefficients of compute type C.
// `air` holds the coefficients shared by every air voxel (position byte 0x86); air_uniform
// is 0 when the table's K = 6 row differs between materials (non-finite admittances), in
// which case every voxel reads the table.
template <typename S, typename CT, typename C>
struct UpdateArgs {
  const uint8_t *pos;
  const uint8_t *mat;
  const S *P;
  S *Pp;
  const CT *coef;
  int ncol;  // material columns staged in LDS: highest material byte in the domain + 1
  CT air;
  int air_uniform;
  C lam2;
  int scheme;  // 0 forward (Bilbao) boundaries, 2 centred (Kowalczyk) boundaries
};

// ---- serving pool (pfdtd_pool.hip; the two-level pool launch lives in pfdtd_fused.hip)
constexpr int kPoolMaxJobs = 64;

// One resident job. Field buffers F[0..3]: jobs advanced by single steps use F[0] and F[1];
// jobs advanced by the two-level sweep (fused != 0) own all four.
struct PoolJobDev {
  const uint8_t *pos;
  const uint8_t *mat;
  const uint16_t *pm;  // pos | mat << 8 (two-level jobs)
  const uint32_t *air_rows;  // air-row bits of the two-level wave tiles (two-level jobs)
  void *F[4];
  const CoefF *coef;
  int ncol;  // material columns of the table the job's domain addresses (its n_mat)
  CoefF air;
  int air_uniform;
  float lam2;
  int X, Y, Z;
  int gx, gy;      // single-step tiles along x and y (z tiles follow from the tile count)
  int fx, fy;      // two-level tiles along x and y
  int fused;       // advanced by the two-level sweep
  const int64_t *rec_idx;
  const int32_t *rec_xyz;
  double *resp;  // [n_rec][n_steps]
  const int64_t *src_idx;
  const int32_t *src_xyz;
  const int32_t *src_kind;
  const float *sig;  // [n_src][n_samples]
  int n_rec, n_src, n_samples, n_steps;
  int soft_overwrite, scale_exp;
  int distinct;  // sources on distinct voxels: injection may run in parallel
};

// Per-job progress in device memory: steps done and the buffers holding P^n (cur) and
// P^{n-1} (past).
struct PoolState {
  int step;
  int cur, past;
};

// The two buffers of a four-buffer job that hold neither P^n nor P^{n-1}, in index order.
__host__ __device__ inline void pool_spare(int cur, int past, int &o1, int &o2) {
  o1 = -1;
  o2 = -1;
  for (int i = 0; i < 4; ++i)
    if (i != cur && i != past) {
      if (o1 < 0)
        o1 = i;
      else
        o2 = i;
    }
}

// Launch table, passed by value: the jobs of one launch and their cumulative tile counts.
struct PoolLaunch {
  int n;
  int ncol_max;  // largest material column count of the launch's jobs (LDS bytes per block)
  int slot[kPoolMaxJobs];
  int tile_end[kPoolMaxJobs];
};

// Two-level sweep of every job of a launch (two steps, or one for a job with one step left),
// then per job: gather/inject of the second step and the buffer rotation.
hipError_t launch_pool_update2(int precision, int tiles, const PoolJobDev *jobs,
                               const PoolState *st, const PoolLaunch &L, int zc,
                               hipStream_t stream);
hipError_t launch_pool_inject2(int precision, const PoolJobDev *jobs, PoolState *st,
                               const PoolLaunch &L, hipStream_t stream);

// Launchers (all asynchronous on `stream`; return the launch status).
hipError_t launch_update_f32(const UpdateGeom &g, const UpdateArgs<float, CoefF, float> &a,
                             hipStream_t stream);
hipError_t launch_update_f64(const UpdateGeom &g, const UpdateArgs<double, CoefD, double> &a,
                             hipStream_t stream);
hipError_t launch_update_f16(const UpdateGeom &g, const UpdateArgs<h16, CoefF, float> &a,
                             hipStream_t stream);
hipError_t launch_update_bf16(const UpdateGeom &g, const UpdateArgs<b16, CoefF, float> &a,
                              hipStream_t stream);

// Two-time-level update (temporal blocking): reads P^n (A), P^{n-1} (B), writes P^{n+1} (Cn)
// with the sources of step n+1 injected, and P^{n+2} (Dn); gathers receivers of step n from
// P^{n+1} before the injection. Coordinates of sources/receivers outside the grid are
// (-2^30, -2^30, -2^30).
template <typename S, typename CT, typename C>
struct Fused2Args {
  const uint16_t *pm;  // position byte | material byte << 8 per voxel
  const uint32_t *air_rows;  // air-row bits of the wave tiles (launch_air_rows), or null
  const S *A;
  const S *B;
  S *Cn;
  S *Dn;
  const CT *coef;
  int ncol;  // material columns staged in LDS: highest material byte in the domain + 1
  CT air;
  int air_uniform;
  C lam2;
  const int32_t *src_xyz;
  const int32_t *src_kind;
  int n_src;
  const float *sig_f;
  const double *sig_d;
  int n_samples;
  int inject_step;
  int soft_overwrite;
  int scale_exp;
  const int32_t *rec_xyz;
  int n_rec;
  double *resp;
  int resp_steps;
  int gather_step;
  int levels = 2;  // time levels of this launch (1 only for the difference form's last step)
  // three-level sweep: the receivers of step gather_step2 are gathered on P^{n+2} before the
  // sources of step inject_step2 are injected into it (-1: none)
  int gather_step2 = -1;
  int inject_step2 = -1;
  int diff = 0;    // 1: the difference form (B holds D = P^n - P^{n-1}; 16-bit storage only)
  // Grouped batch (slot-outermost fields): the launch advances n_slots source configurations,
  // slot s's fields at element offset s * slot_elems; sources and receivers carry their slot.
  int n_slots = 1;
  int64_t slot_elems = 0;
  const int32_t *src_slot = nullptr;
  const int32_t *rec_slot = nullptr;
};
// Block shapes of the one-column two-level sweep (wave tiles in x by y)
constexpr int kColumnBlocks = 0, kWideBlocks = 1, kSquareBlocks = 2;
struct FusedGeom {
  int X, Y, Z;
  int zc;
  int blocks;  // kColumnBlocks (1 x 4 wave tiles), kWideBlocks (4 x 4) or kSquareBlocks (2 x 2)
  bool pair;   // 16-bit storage: two columns per lane (124-column wave tiles, X even)
  bool three = false;  // three time levels per launch (58-column wave tiles, forward scheme)
  int scheme;  // 0 forward boundaries, 2 centred
  dim3 grid, block;
};
// Grid of the two-level sweep; `pool`: serving-pool launches (1 x 4 blocks); `shape`: one of
// PFDTD_SWEEP_AUTO / _COLUMN / _WIDE; `scheme`: 0 forward, 2 centred boundaries.
FusedGeom plan_fused(int X, int Y, int Z, int zc, int precision, bool pool = false,
                     int shape = 0, int scheme = 0, int slots = 1);
// Grid of the three-level sweep (forward scheme): 58-column, 4-row wave tiles in blocks of
// 2 x 2 tiles, or (`shape` PFDTD_SWEEP_COLUMN) 1 x 4.
FusedGeom plan_fused3(int X, int Y, int Z, int zc, int precision, int shape, int slots = 1);
// One bit per (z-slice, two-level x tile, row): set when the tile's row holds only plain air,
// so the sweep takes the packed plane's value without loading it. Words for an X x Y x Z grid,
// and the kernel that builds them from the packed plane.
// `kind`: 0 the one-column two-level sweep's tiles (60 columns), 1 the two-column sweep's (124),
// 2 the three-level sweep's (58 columns, a 2-voxel window ring).
size_t air_rows_words(int X, int Y, int Z, int kind = 0);
hipError_t launch_air_rows(const uint16_t *pm, uint32_t *rows, int X, int Y, int Z, int scheme,
                           int kind, hipStream_t stream);
hipError_t launch_update2_f32(const FusedGeom &g, const Fused2Args<float, CoefF, float> &a,
                              hipStream_t stream);
hipError_t launch_update2_f64(const FusedGeom &g, const Fused2Args<double, CoefD, double> &a,
                              hipStream_t stream);
hipError_t launch_update2_f16(const FusedGeom &g, const Fused2Args<h16, CoefF, float> &a,
                              hipStream_t stream);
hipError_t launch_update2_bf16(const FusedGeom &g, const Fused2Args<b16, CoefF, float> &a,
                               hipStream_t stream);
// Three time levels (the same arguments; P^{n+2} to Cn, P^{n+3} to Dn; forward scheme).
hipError_t launch_update3_f32(const FusedGeom &g, const Fused2Args<float, CoefF, float> &a,
                              hipStream_t stream);
hipError_t launch_update3_f64(const FusedGeom &g, const Fused2Args<double, CoefD, double> &a,
                              hipStream_t stream);
hipError_t launch_update3_f16(const FusedGeom &g, const Fused2Args<h16, CoefF, float> &a,
                              hipStream_t stream);
hipError_t launch_update3_bf16(const FusedGeom &g, const Fused2Args<b16, CoefF, float> &a,
                               hipStream_t stream);

// Receiver gather of time index `gather_step` (skipped when < 0) followed by injection of
// source samples of time index `inject_step` (skipped when < 0) into the same field.
struct InjectArgs {
  const int64_t *rec_idx;  // -1: outside the grid
  int n_rec;
  double *resp;            // [n_rec][resp_steps]
  int resp_steps;
  int gather_step;
  const int64_t *src_idx;  // -1: outside the grid
  const int32_t *src_kind;
  int n_src;
  const float *sig_f;      // [n_src][n_samples]
  const double *sig_d;
  int n_samples;
  int inject_step;
  int soft_overwrite;
  int scale_exp;
  int distinct;  // 1: no two sources share an element, so they may be applied in parallel
  void *dfield = nullptr;  // the difference form's D field (injection keeps P - D), or null
};
hipError_t launch_inject_gather(int precision, void *field, const InjectArgs &a,
                                hipStream_t stream);

// Material slot of material m at `octave` (intended m*20+octave, or the reference's literal
// m*20*octave) and the per-(K, m) coefficient tables (kCoefEntries entries each) built from a
// [n_mat][20] admittance table.
int64_t material_slot(unsigned m, unsigned octave, bool literal);
// Sets the message pfdtd_last_error() returns on this thread.
void record_error(const char *msg);
void build_coef_tables(const float *table20, int n_mat, unsigned octave, double lambda,
                       bool literal, CoefF *cf, CoefD *cd);

hipError_t launch_classify_labels(const uint8_t *labels, uint8_t *pos, uint8_t *mat, int X,
                                  int Y, int Z, hipStream_t stream);
hipError_t launch_codes_to_bilbao(uint8_t *pos, uint8_t *mat, int64_t n, hipStream_t stream);
// pm[i] = pos[i] | mat[i] << 8 (the two-level update's packed byte plane).
hipError_t launch_pack_pm(const uint8_t *pos, const uint8_t *mat, uint16_t *pm, int64_t n,
                          hipStream_t stream);
hipError_t launch_scan_domain(const uint8_t *pos, const uint8_t *mat, int X, int Y, int Z,
                              int scheme, DomainScan *d_out, hipStream_t stream);
// Centred-scheme (update type SRL) classification of labels / voxelizer codes.
hipError_t launch_classify_labels_centred(const uint8_t *labels, uint8_t *pos, uint8_t *mat,
                                          int X, int Y, int Z, unsigned int *degenerate,
                                          hipStream_t stream);
hipError_t launch_codes_to_centred(uint8_t *pos, uint8_t *mat, int64_t n, hipStream_t stream);
// Transparent-source filter of n_src signals of n_samples each (FP32 or FP64).
hipError_t launch_transparent(bool f64, const void *in, const float *h, int n_ir, int n_src,
                              int n_samples, void *out, hipStream_t stream);
// Centred-scheme coefficient tables: entry (D, m), D = number of DIR bits (0..3).
void build_centred_tables(const float *table20, int n_mat, unsigned octave, double lambda,
                          CoefF *cf, CoefD *cd);
// HBM streaming probe (pfdtd_probe.hip): one launch streams `blocks` * kProbeBlockElems float4
// elements of each array the kind uses (a, then b, c; sink takes nothing in practice).
constexpr int64_t kProbeBlockElems = 1024;
hipError_t launch_hbm_probe(int kind, float4 *a, float4 *b, float4 *c, float *sink,
                            int64_t blocks, hipStream_t st);
// Gathers a plane of the field (unscaled FP32) and of the position bytes.
hipError_t launch_capture_slice(int precision, const void *field, const uint8_t *pos, int X,
                                int Y, int Z, int64_t slice, int orientation, int scale_exp,
                                float *out_p, uint8_t *out_pos, hipStream_t stream);
// Triangle-mesh voxelization (pfdtd_voxel.hip): fixed-point vertices and grid dims of a mesh
// (0, or -1 on a bad dx / non-finite vertex / grid above 32768 per axis), then the label grid
// of the mesh written to d_labels (dims[0]*dims[1]*dims[2] bytes). Triangle indices must
// already be checked against the vertex count.
int voxel_prepare(const float *v, int64_t nv, float dx, const int32_t *pad, int64_t dims[3],
                  std::vector<int32_t> *q);
hipError_t voxelize_labels(const std::vector<int32_t> &q, const uint32_t *tri, int64_t n_tri,
                           const uint8_t *tmat, const int64_t dims[3], uint8_t *d_labels,
                           hipStream_t s);
// Whole field to unscaled FP32.
hipError_t launch_field_to_f32(int precision, const void *field, int64_t n, int scale_exp,
                               float *out, hipStream_t stream);

}  // namespace pfdtd
