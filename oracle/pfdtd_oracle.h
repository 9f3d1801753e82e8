/*
 * pfdtd_oracle.h -- CPU restatement of ParallelFDTD's SRL_FORWARD time-stepping core.
 *
 * TEST INFRASTRUCTURE ONLY. This is the parity oracle: a plain-C restatement of the
 * reference algorithm used by tests/, __graft_entry__.smoke() and the cpu_baseline leg
 * of bench.py as the *checker*. The product library (parallelfdtd_amd/libpfdtd.so)
 * never links, loads or calls it.
 *
 * Every function cites the reference file:line it follows (paths relative to the
 * reference repository root). Floating-point expressions keep the reference's operand
 * order and are compiled with -ffp-contract=off so that each product and sum is a
 * separately rounded IEEE operation.
 *
 * Pinning (what checks each function independently of this file's own formulas) is
 * listed per function below and enforced by tests/test_oracle_*.py.
 */
#ifndef PFDTD_ORACLE_H
#define PFDTD_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Source kinds: enum SrcType, src/base/SrcRec.h:33 */
#define ORC_SRC_HARD 0
#define ORC_SRC_SOFT 1
#define ORC_SRC_TRANSPARENT 2
/* Input kinds: enum InputType, src/base/SrcRec.h:34 */
#define ORC_IN_IMPULSE 0
#define ORC_IN_GAUSSIAN 1
#define ORC_IN_SINE 2
#define ORC_IN_DATA 3

/* Field storage for the FP32-arithmetic run (precision study, no reference counterpart). */
#define ORC_STORE_F32 0
#define ORC_STORE_F16 1
#define ORC_STORE_BF16 2
/* Difference form (no reference counterpart): the stored state is P^n and D^n = P^n - P^{n-1},
 * both binary16 (scaled). Each update forms the past value as P^n - D^n in FP32, evaluates the
 * reference expression, and stores P^{n+1} = fl16(ret) and D^{n+1} = fl16(ret - P^n), ret the
 * FP32 result. A source sample changes P^n and keeps the past value: D^n = fl16(P^n_new - past).
 * P_out / Pp_out then hold P and D. */
#define ORC_STORE_F16_DIFF 3
/* Sum-difference form (no reference counterpart): the stored state is S^n = P^n + P^{n-1} and
 * D^n = P^n - P^{n-1}, both binary16 (scaled). Each update reconstructs P^n = (S^n + D^n)/2 and the
 * past value (S^n - D^n)/2 in FP32, evaluates the reference expression, and stores
 * S^{n+1} = fl16(ret + P^n), D^{n+1} = fl16(ret - P^n). A source sample changes P^n and keeps the
 * past value. Receivers read the reconstructed P. P_out / Pp_out then hold S and D. */
#define ORC_STORE_F16_SUMDIFF 4

/* Bit 2 of the `literal` argument of orc_run_*: update type SRL (centred Kowalczyk boundary,
 * enum UpdateType, src/base/SimulationParameters.h:39) instead of SRL_FORWARD. */
#define ORC_SCHEME_KOWALCZYK 4

int orc_version(void);

/* Voxelizer code k (0..27) -> Bilbao position byte, material byte zeroed for k==0.
 * Follows toBilbao, src/kernels/cudaMesh.cu:328-361.
 * Pinned by: the code->K groups documented in toKowalczyk, cudaMesh.cu:372-479 (1-8 three
 * air directions, 9-20 four, 21-26 five, 27 six), checked exhaustively in tests. */
uint8_t orc_to_bilbao(uint8_t k, uint8_t *mat_inout);

/* Classify a label grid (label 0 = inside/air, label m+1 = solid voxel of material m).
 * pos = 0x80|K for inside voxels (K = number of inside face neighbours, out-of-grid counts
 * as solid), 0 for solid voxels. mat = material of the first solid face neighbour in the
 * order -x,+x,-y,+y,-z,+z for inside voxels with K<6, else 0.
 * counts[0] = air (pos==0x86), counts[1] = boundary (pos not 0 and not 0x86), following
 * calcBoundaries, src/kernels/cudaMesh.cu:500-514; counts[2+K] = histogram of K (inside only).
 * Pinned by: closed-form box-room counts, brute-force neighbour counting in numpy. */
void orc_classify_labels(const uint8_t *label, int64_t X, int64_t Y, int64_t Z,
                         uint8_t *pos, uint8_t *mat, uint64_t counts[9]);

/* Voxelizer code k (0..27) -> centred-scheme (update type SRL) position byte: toKowalczyk,
 * src/kernels/cudaMesh.cu:363-480, restated from each code's list of air directions.
 * Pinned by: the exhaustive byte table of toKowalczyk transcribed as data in the tests, and
 * consistency with orc_classify_labels_kow on every 3x3x3 neighbourhood. */
uint8_t orc_to_kowalczyk(uint8_t k, uint8_t *mat_inout);

/* Label grid -> centred-scheme position/material bytes; returns the number of inside voxels
 * with no air neighbour on some axis (no centred stencil). counts: air (0x80), boundary,
 * degenerate. Pinned by: orc_to_kowalczyk on codes derived from each voxel's neighbourhood. */
uint64_t orc_classify_labels_kow(const uint8_t *label, int64_t X, int64_t Y, int64_t Z,
                                 uint8_t *pos, uint8_t *mat, uint64_t counts[3]);

/* Parameter vector [lambda, lambda^2, 1/3, octave]:
 * SimulationParameters::getParameterPtr(+Double), src/base/SimulationParameters.cpp:379-396.
 * Pinned by: tests/SimulationParametersTest.cpp:80-97. */
void orc_param_vector_f(double lambda, unsigned octave, float out[4]);
void orc_param_vector_d(double lambda, unsigned octave, double out[4]);

/* Reflectance <-> admittance, (1-R)/(1+R): src/global_includes.h:30-38.
 * Pinned by: involution property and R=0 -> 1, R=1 -> 0. */
float orc_r2y_f(float r);

/* dx = c / (fs * lambda): SimulationParameters::getDx, SimulationParameters.cpp:362-364
 * (double arithmetic, float result). Pinned by SimulationParametersTest.cpp:16-22. */
float orc_dx(float c, unsigned fs, double lambda);

/* Element index of a position in metres: Position::getElementIdx, src/base/SrcRec.cpp:26-34,
 * with nv::ROUND = floor(x+0.5f) (src/math/geomMath.h:126-128) and the +1 padding offset of
 * SimulationParameters::getSource/ReceiverElementCoordinates, SimulationParameters.cpp:200-224.
 * Pinned by: hand-computed indices in tests. */
void orc_element_idx(float px, float py, float pz, unsigned fs, float c, double lambda,
                     int add_padding, int32_t out[3]);

/* Regular source sample: SimulationParameters::getRegularSourceSample(+Double),
 * SimulationParameters.cpp:262-298 and :312-348. DATA reads past the end return 0
 * (getInputDataSample, :161-174). Pinned by: SimulationParametersTest.cpp:126-185. */
float orc_source_sample_f(int input_type, unsigned step, unsigned fs, const float *data,
                          unsigned n_data);
double orc_source_sample_d(int input_type, unsigned step, unsigned fs, const double *data,
                           unsigned n_data);

/* Transparent source sample s(n) - sum_{i<n} h(n-i) s(i):
 * getTransparentSourceSample(+Double), SimulationParameters.cpp:300-310, :350-360.
 * Pinned by: SimulationParametersTest.cpp:196-207 (s_T(3)=1/3 with the grid IR h[2]=-1/3). */
float orc_transparent_sample_f(int input_type, unsigned step, unsigned fs, const float *data,
                               unsigned n_data, const float *grid_ir, unsigned n_ir);
double orc_transparent_sample_d(int input_type, unsigned step, unsigned fs, const double *data,
                                unsigned n_data, const float *grid_ir, unsigned n_ir);

/* Index into the [material][20] coefficient table used by the update for material byte m.
 * intended: m*20 + octave (the centred kernel's form, kernels3d.cu:639).
 * literal:  (unsigned)((float)(m*20) * (float)octave), i.e. the forward kernel's
 *           `mat*20*+params[3]`, kernels3d.cu:513.
 * Pinned by: literal==intended at (octave 1, m) vs (octave 0, m); slot 0 at octave 0. */
int64_t orc_material_slot(unsigned m, unsigned octave, int literal);

/* Full step loop. Follows launchFDTD3d / launchFDTD3dDouble, src/kernels/kernels3d.cu:83-191
 * (inject all sources into P, update, flip, gather receivers), with the per-node update of
 * fdtd3dStdMaterials, kernels3d.cu:485-529, applied to every node of slices z in [1, Z-2]
 * (launch coverage kernels3d.cu:112-123). Nodes with x or y on the grid edge are written 0;
 * the reference writes switchBit*(...) there and switchBit is 0 under the precondition that
 * the outer voxel layer is outside (checked: returns -2 otherwise).
 *
 * mat_table: n_mat*20 coefficients laid out [material][coef] (MaterialHandler.cpp:100-164).
 * src_xyz / rec_xyz: grid indices (x,y,z) per source/receiver (already padded).
 * src_type: ORC_SRC_*; signals: [n_src][n_steps] samples of the source signal (already
 * transparent-filtered where applicable). Sources/receivers outside the grid are skipped
 * (receivers then stay 0, as kernels3d.cu:168).
 * literal: bit0 = Q1 material slot (see orc_material_slot), bit1 = Q2 soft source overwrites
 * (cudaMesh.h:362-366) instead of adding, bit2 (ORC_SCHEME_KOWALCZYK) = centred update
 * fdtd3dStdKowalczykMaterials (kernels3d.cu:609-665; its material slot is always m*20+octave).
 * responses: [n_rec][n_steps], response[r][n] = P^{n+1}[rec] (post-flip read, kernels3d.cu:164-173).
 * P_out / Pp_out (optional): the final current and past fields after the last flip.
 * storage (f32 entry only): ORC_STORE_* with field values stored scaled by 2^scale_exp,
 * rounded to nearest even in the storage format after every write.
 * Returns 0 on success, -1 bad argument, -2 precondition violated.
 *
 * Pinned by: the free-field closed form h[n] = E[U_n(c)] (tiny grids vs lattice-walk sums),
 * K=6 air reduction P_next = S/3 - P_past, energy conservation at beta=0, decay at beta>0,
 * mirror symmetry, and brute-force golden fixtures from an independent numpy evaluation. */
int orc_run_f(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z,
              const float *mat_table, int n_mat, unsigned octave, double lambda,
              const int32_t *src_xyz, const int32_t *src_type, const float *signals, int n_src,
              const int32_t *rec_xyz, int n_rec, int n_steps, int literal, int storage,
              int scale_exp, int threads, float *responses, float *P_out, float *Pp_out);

int orc_run_d(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z,
              const double *mat_table, int n_mat, unsigned octave, double lambda,
              const int32_t *src_xyz, const int32_t *src_type, const double *signals, int n_src,
              const int32_t *rec_xyz, int n_rec, int n_steps, int literal, int threads,
              double *responses, double *P_out, double *Pp_out);

/* Slice gather of a field and the position bytes: captureSliceFast / captureSliceKernel,
 * src/kernels/visualizationUtils.cu:130-254, :500-527. orientation 0 = xy at z=slice
 * (out[y*X+x]), 1 = xz at y=slice (out[z*X+x]), 2 = yz at x=slice (out[z*Y+y]); in xz and yz
 * the row z = Z-1 is 0 (the kernel's limit dim * (partition size - 1), one partition).
 * Returns -1 when the slice is out of range (the reference logs and skips).
 * Pinned by numpy slicing of random fields in every orientation and the out-of-range rule
 * (tests/test_oracle_capture.py). */
int orc_capture_slice_f(const float *P, const uint8_t *pos, int64_t X, int64_t Y, int64_t Z,
                        int64_t slice, int orientation, float *out_p, uint8_t *out_pos);

/* Single-step mode: launchFDTD3dStep (kernels3d.cu:376-482) called n_calls times as
 * App::executeStep calls it (App.cpp:407-436), call c with step index steps[c] (0 <= steps[c] <
 * n_steps) and direction dirs[c] (+1 or -1). Each call injects the source samples of step
 * steps[c] (signals [n_src][n_steps]) into the current field, runs the update (current, past ->
 * past buffer), flips the field pointers only when dirs[c] equals the previous call's
 * direction (initially +1: the function-static past_step_directon, kernels3d.cu:386, :462-465),
 * then writes response[r][steps[c]] from the current field. Responses no call writes are 0.
 * Other arguments and the return code as orc_run_f.
 * Pinned by (tests/test_oracle_steps.py): all-forward calls equal orc_run_f bit for bit; on a
 * direction change the current field stays p^K and the past buffer holds p^{K+1}, both bit for
 * bit; with rigid walls (FP64 arithmetic through FP32 fields aside) reversed calls retrace the
 * forward fields p^{K-1}, p^{K-2}, ... to rounding. */
int orc_step_calls_f(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z,
                     const float *mat_table, int n_mat, unsigned octave, double lambda,
                     const int32_t *src_xyz, const int32_t *src_type, const float *signals,
                     int n_src, const int32_t *rec_xyz, int n_rec, int n_steps, int literal,
                     int storage, int scale_exp, int threads, const int32_t *steps,
                     const int32_t *dirs, int n_calls, float *responses, float *P_out,
                     float *Pp_out);

/* Triangle-mesh solid voxelization into a label grid (0 = air inside the mesh, m+1 = solid
 * voxel of material m), the input of orc_classify_labels / orc_classify_labels_kow.
 *
 * The reference delegates this to the external hakarlss/Voxelizer library (call site
 * voxelizeGeometry, src/kernels/voxelizationUtils.cu:58-131; App::initializeMesh,
 * src/App.cpp:178-187; build README.md:24), which is not vendored and has no pinned version
 * in the reference, so there is no reference source to follow. README.md:100 states its
 * method: intersection tests decide inside/outside, so the mesh must be watertight. This
 * function states the definition the HIP voxelizer implements:
 *  - frame: vertices minus their per-axis minimum (float subtraction, as
 *    GeometryHandler::initialize, src/base/GeometryHandler.cpp:64-73), then fixed point
 *    q = llrint(v * 4096 / dx) (double arithmetic);
 *  - node (i,j,k) sits at ((i-1)dx, (j-1)dx, (k-1)dx): the +1 padding offset of
 *    Position::getElementIdx (src/base/SrcRec.cpp:26-34); grid = floor(extent/dx) + 3 nodes
 *    per axis, rounded up to a multiple of pad[a] (the (32,4,1) of App.cpp:193 / padWithZeros,
 *    src/kernels/cudaMesh.cu:253-326, added at the high end);
 *  - inside (air) = odd number of crossings of the x-ray through the node centre, counting a
 *    triangle at column (y,z) when the column point lies inside its yz projection under a
 *    top-left rule (exact integer edge functions), the crossing at
 *    x* = xA + (wB (xB-xA) + wC (xC-xA)) / W (double, no contraction) and flipping nodes
 *    with x >= x*;
 *  - solid nodes take material 1 + tmat[t] of the triangle t crossing the link to their first
 *    air face neighbour (order -x,+x,-y,+y,-z,+z; links rasterised per axis in the same way;
 *    the lowest triangle index when several cross one link), 1 when no air neighbour or no
 *    recorded crossing.
 * Returns 0, -1 (bad dx / non-finite vertex / grid above 32768 per axis), -2 (triangle index
 * out of range), -3 (material above 254).
 * Pinned by (tests/test_oracle_voxel.py): closed-form box rooms (air counts and per-face
 * materials, walls on and between node planes), nested solids, brute-force half-space
 * point-in-convex-polyhedron tests for rotated boxes and tetrahedra, the enclosed volume of
 * the reference's own meshes (python/Data/box.json, larun_hytti.json), and invariance under
 * triangle order, vertex rotation, winding and translation. */
int orc_voxel_grid(const float *vertices, int64_t n_vertices, float dx, const int32_t pad[3],
                   int64_t dims[3]);
int orc_voxelize(const float *vertices, int64_t n_vertices, const uint32_t *tri,
                 int64_t n_tri, const uint8_t *tmat, float dx, const int32_t pad[3],
                 int64_t dims[3], uint8_t *labels);

/* Storage rounding helpers (round to nearest even), exposed for tests. */
uint16_t orc_f32_to_f16(float f);
float orc_f16_to_f32(uint16_t h);
uint16_t orc_f32_to_bf16(float f);
float orc_bf16_to_f32(uint16_t h);

#ifdef __cplusplus
}
#endif
#endif
