/*
 * pfdtd.h -- C ABI of the MI355X (gfx950) SRL_FORWARD time-stepping core.
 *
 * This is the boundary between host code (the FDTD::App facade, Python through ctypes or
 * pybind11, a MEX gateway) and the hand-written HIP kernels in libpfdtd.so. Plain C types
 * only: no HIP or torch types cross it. Every entry point returns an int status
 * (PFDTD_OK == 0, negative on error) and leaves a message for pfdtd_last_error().
 *
 * Each declaration names the reference interface it replaces (paths relative to the
 * juuli/ParallelFDTD repository root).
 *
 * Grid layout everywhere: idx = (z*Y + y)*X + x, x fastest (src/kernels/cudaMesh.h:247-249).
 * Precondition (checked): every voxel of the outer layer is outside (position byte bit 7
 * clear); the reference only updates slices z in [1, Z-2] and wraps x/y edges into
 * neighbouring rows (src/kernels/kernels3d.cu:112-123, :516-521), which is only well
 * defined under this precondition.
 */
#ifndef PFDTD_H
#define PFDTD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ------------------------------------------------------------------ */
#define PFDTD_OK 0
#define PFDTD_E_ARG (-1)    /* bad argument (null pointer, size, index, enum value) */
#define PFDTD_E_HIP (-2)    /* a HIP runtime call failed */
#define PFDTD_E_NOMEM (-3)  /* device or host allocation failed / request exceeds free HBM */
#define PFDTD_E_STATE (-4)  /* call order: e.g. run before a domain and materials are set */
#define PFDTD_E_DOMAIN (-5) /* domain violates a precondition (shell, K range, material slot) */

/* ---- field precision --------------------------------------------------------------- */
#define PFDTD_F32 0  /* FP32 fields and arithmetic: launchFDTD3d */
#define PFDTD_F64 1  /* FP64 fields and arithmetic: launchFDTD3dDouble */
#define PFDTD_F16 2  /* FP32 arithmetic, binary16 field storage, power-of-two scaled */
#define PFDTD_BF16 3 /* FP32 arithmetic, bfloat16 field storage, power-of-two scaled */
/* FP32 arithmetic, binary16 storage of P^n and D^n = P^n - P^{n-1} (difference form, power-of-two
 * scaled; no reference counterpart): a storage rounding of P no longer changes the velocity
 * P^n - P^{n-1}, which cuts the in-band error growth of the scheme's marginal low-frequency
 * modes. pfdtd_run only, through the two-column sweep: forward update, X even, no batch, no
 * source on the slices z = 0 or Z-1. pfdtd_get_fields_raw returns P and D. */
#define PFDTD_F16_DIFF 4

/* ---- source kinds: enum SrcType, src/base/SrcRec.h:33 ----------------------------- */
#define PFDTD_SRC_HARD 0
#define PFDTD_SRC_SOFT 1
#define PFDTD_SRC_TRANSPARENT 2 /* injected like SOFT; the signal is pre-filtered by the host */

/* ---- material-slot and soft-source conventions (reference quirks) ------------------ */
#define PFDTD_MATSLOT_INTENDED 0 /* beta = table[mat*20 + octave] (kernels3d.cu:639) */
#define PFDTD_MATSLOT_LITERAL 1  /* beta = table[mat*20*octave]   (kernels3d.cu:513) */
#define PFDTD_SOFT_ADD 0         /* soft source adds (CudaMeshTest.cpp:311-314) */
#define PFDTD_SOFT_OVERWRITE 2   /* soft source overwrites, as addSample does (cudaMesh.h:362-366) */

/* ---- block shape of the two-level sweep (pfdtd_set_sweep_shape) -------------------- */
#define PFDTD_SWEEP_AUTO 0   /* two levels, on rooms of at least 901 columns (16 tiles of 60):
                                FP32 wide, FP64 and centred FP32 blocks of 2 x 2 wave tiles;
                                else column; 16-bit storage: pair when X is even, else column;
                                three levels: blocks of 2 x 2 wave tiles */
#define PFDTD_SWEEP_COLUMN 1 /* blocks of 1 x 4 wave tiles (60 columns x 4 tiles of rows; the
                                three-level sweep: 58 columns) */
#define PFDTD_SWEEP_WIDE 2   /* FP32: blocks of 4 x 4 wave tiles (240 columns, 1024 threads) */
#define PFDTD_SWEEP_PAIR 3   /* 16-bit storage, X even: two columns per lane (124-column wave
                                tiles in blocks of 2 x 2, the difference form 1 x 4; one
                                32-bit access per pair); otherwise as AUTO */

typedef struct pfdtd_ctx pfdtd_ctx;

/* InterruptCallback bool(*)(void), src/App.h:33 -- nonzero stops the run. */
typedef int (*pfdtd_interrupt_cb)(void *user);
/* ProgressCallback void(*)(int,int,float), src/App.h:34 -- called every 100 steps. */
typedef void (*pfdtd_progress_cb)(int step, int n_steps, float seconds_per_step, void *user);

/* Last error message of the calling thread (cudasafe's message, src/kernels/cudaUtils.h:47-51). */
const char *pfdtd_last_error(void);
int pfdtd_version(void);

/* Device queries: App::queryDevices, src/App.cpp:91-111. */
int pfdtd_device_count(int *count);
int pfdtd_device_mem(int device, uint64_t *free_bytes, uint64_t *total_bytes);

/* Context on one device: App::initializeDevices, src/App.cpp:125-131 (without resetting the
 * device: other contexts may share it). All device memory belongs to the context. */
int pfdtd_ctx_create(int device, pfdtd_ctx **out);
/* App::close / CudaMesh::destroyPartitions, src/App.cpp:438-443, src/kernels/cudaMesh.h:160-182. */
int pfdtd_ctx_destroy(pfdtd_ctx *ctx);

/* ---- domain ------------------------------------------------------------------------- */
/* Pre-classified position and material bytes: CudaMesh::setupMesh(+Double),
 * src/kernels/cudaMesh.cu:26-153 (without block padding). Position byte = 0x80|K for inside
 * voxels (K = air face-neighbour count, 0..6), 0 outside. Copies the host arrays. */
int pfdtd_set_domain(pfdtd_ctx *ctx, const uint8_t *pos, const uint8_t *mat, int64_t X,
                     int64_t Y, int64_t Z);
/* Label grid classified on the GPU: label 0 = inside, label m+1 = solid voxel of material m.
 * Replaces voxelizeGeometry + toBilbaoScheme (src/kernels/voxelizationUtils.cu:55-136,
 * src/kernels/cudaMesh.cu:204-251) for voxel-built rooms: K = inside face neighbours
 * (out-of-grid counts as solid), material byte = material of the first solid face neighbour
 * in the order -x,+x,-y,+y,-z,+z for inside voxels with K<6, else 0. */
int pfdtd_set_domain_labels(pfdtd_ctx *ctx, const uint8_t *labels, int64_t X, int64_t Y,
                            int64_t Z);
/* Voxelizer node codes (0..27) mapped on the GPU by toBilbao, src/kernels/cudaMesh.cu:328-361
 * (toBilbaoKernel :492-498); material bytes as given (zeroed for code 0). */
int pfdtd_set_domain_codes(pfdtd_ctx *ctx, const uint8_t *codes, const uint8_t *mat, int64_t X,
                           int64_t Y, int64_t Z);
/* Triangle-mesh solid voxelization on the GPU, then classification for the current update
 * type: replaces voxelizeGeometry (src/kernels/voxelizationUtils.cu:58-131, the external
 * hakarlss/Voxelizer) + padWithZeros + toBilbaoScheme/toKowalczykScheme as App::initializeMesh
 * runs them (src/App.cpp:178-211, src/kernels/cudaMesh.cu:26-153). vertices [n_vertices][3]
 * (metres), indices [n_triangles][3], tri_material [n_triangles] (unique material index per
 * triangle, MaterialHandler::getMaterialIdxPtr; null = all 0; at most 254), dx = voxel edge.
 * The mesh must be closed (README.md:100): a node is air when the ray along +x through its
 * centre has crossed the surface an odd number of times. Frame and grid: vertices minus their
 * per-axis minimum; node (i,j,k) at ((i-1)dx, (j-1)dx, (k-1)dx) (the +1 padding offset of
 * Position::getElementIdx, src/base/SrcRec.cpp:26-34); floor(extent/dx)+3 nodes per axis,
 * rounded up to a multiple of pad[a] (null = 1; App uses 32,4,1 as App.cpp:193). Solid nodes
 * next to air carry the material of the triangle crossing the link to their first air
 * neighbour (-x,+x,-y,+y,-z,+z). dims (nullable) receives X,Y,Z; labels (nullable, X*Y*Z
 * bytes, x fastest) the label grid (0 air, m+1 solid of material m). */
int pfdtd_voxelize(pfdtd_ctx *ctx, const float *vertices, int64_t n_vertices,
                   const uint32_t *indices, int64_t n_triangles, const uint8_t *tri_material,
                   float dx, const int32_t *pad, int64_t *dims, uint8_t *labels);
/* Grid dims pfdtd_voxelize would produce (host only, no context). */
int pfdtd_voxel_grid(const float *vertices, int64_t n_vertices, float dx, const int32_t *pad,
                     int64_t *dims);
/* Device -> host copy of the classified bytes (bit-exactness checks). */
int pfdtd_get_domain(pfdtd_ctx *ctx, uint8_t *pos, uint8_t *mat);
int pfdtd_get_dims(pfdtd_ctx *ctx, int64_t *X, int64_t *Y, int64_t *Z);
/* CudaMesh::getNumberOfAirElements / getNumberOfBoundaryElements (src/kernels/cudaMesh.h:242-243),
 * counted like calcBoundaries, src/kernels/cudaMesh.cu:500-514. */
int pfdtd_get_counts(pfdtd_ctx *ctx, uint64_t *air, uint64_t *boundary);

/* ---- materials and parameters ------------------------------------------------------ */
/* Material table [n_mat][n_coef] of admittances (MaterialHandler::getMaterialCoefficientPtr,
 * src/base/MaterialHandler.cpp:100-164; coefficients beyond n_coef up to 20 are zero) and the
 * parameter vector [lambda, lambda^2, 1/3, octave] (SimulationParameters::getParameterPtr,
 * src/base/SimulationParameters.cpp:379-396). flags: PFDTD_MATSLOT_* | PFDTD_SOFT_*.
 * The per-(K, material) update coefficients are built from these on the host in the
 * reference's own operation order (FP32 and FP64 tables). */
int pfdtd_set_materials(pfdtd_ctx *ctx, const float *table, int n_mat, int n_coef,
                        unsigned octave, double lambda, int flags);

/* ---- sources and receivers ----------------------------------------------------------- */
/* Sources at grid indices xyz[3*n] (already +1 padded as
 * SimulationParameters::getSourceElementCoordinates, SimulationParameters.cpp:200-211),
 * kinds PFDTD_SRC_*, and their signals [n_src][n_samples] as the FP32 and FP64 runs see them
 * (getSourceSample / getSourceSampleDouble, SimulationParameters.cpp:139-159). Either signal
 * array may be NULL, in which case it is converted from the other. Steps past n_samples
 * inject 0. Sources outside the grid are skipped (setSample, cudaMesh.h:321-338). */
int pfdtd_set_sources(pfdtd_ctx *ctx, int n_src, const int32_t *xyz, const int32_t *kind,
                      int n_samples, const float *signal_f32, const double *signal_f64);
/* Receivers at grid indices xyz[3*n] (getReceiverElementCoordinates,
 * SimulationParameters.cpp:213-224). Receivers outside the grid record 0 (kernels3d.cu:168). */
int pfdtd_set_receivers(pfdtd_ctx *ctx, int n_rec, const int32_t *xyz);

/* ---- transparent sources -------------------------------------------------------------- */
/* out[s][n] = in[s][n] - sum_{i<n} h[n-i] in[s][i] for n_src signals of n_samples samples, on
 * the device: getTransparentSourceSample(+Double), SimulationParameters.cpp:300-310, :350-360
 * (the same left-to-right sum per sample, FP32 or FP64; h = the FP32 grid impulse response,
 * readGridIr, 0 past n_ir). Either precision pair may be null. The reference evaluates this
 * O(n^2) sum on the host inside its step loop. */
int pfdtd_transparent_signals(pfdtd_ctx *ctx, int n_src, int n_samples, const float *grid_ir,
                              int n_ir, const float *in_f32, float *out_f32,
                              const double *in_f64, double *out_f64);

/* ---- update type -------------------------------------------------------------------- */
/* enum UpdateType (src/base/SimulationParameters.h:39): 0 SRL_FORWARD (forward boundaries,
 * fdtd3dStdMaterials, kernels3d.cu:485-529; the default), 1 SHARED (the same update: the
 * reference's sliced kernel fdtd3dSliced, kernels3d.cu:531-607, implements this formula; its
 * unmasked position byte and LDS race are not reproduced), 2 SRL (centred boundaries,
 * fdtd3dStdKowalczykMaterials, kernels3d.cu:609-665). The domain's position bytes encode the
 * boundary formulation (toBilbao vs toKowalczyk, cudaMesh.cu:66-74): set the type before the
 * domain; switching between the forward and centred families discards a classified domain.
 * Labels are classified for the centred scheme as toKowalczyk encodes voxelizer codes (inside
 * voxels with no air neighbour on some axis have no centred stencil: PFDTD_E_DOMAIN at run).
 * Type 2 runs the single-level, unbatched update. */
int pfdtd_set_update_type(pfdtd_ctx *ctx, int type);

/* ---- source batching ---------------------------------------------------------------- */
/* Advance `slots` independent source configurations on the same domain in one launch per
 * step (power of two, 1..256; 1 = unbatched, the default). Fields then hold `slots` values per
 * voxel, slot innermost, and every source and receiver belongs to one slot (default 0; set
 * them after pfdtd_set_sources / pfdtd_set_receivers). Each slot evolves exactly as an
 * unbatched run with only its own sources would. No reference counterpart: the reference runs
 * one source configuration per simulation (App::runSimulation, src/App.cpp:308-348). Batched
 * runs in the interleaved layout use the single-level update; slice/field capture is
 * unavailable (raw fields are). */
int pfdtd_set_batch(pfdtd_ctx *ctx, int slots);
/* Field layout of a batch: PFDTD_BATCH_INTERLEAVED (default; slot innermost, element
 * voxel * slots + slot, advanced by the single-level batched update that reads each voxel's
 * position and material bytes once for all slots) or PFDTD_BATCH_GROUPED (slot outermost,
 * element slot * N + voxel: each slot's fields have the unbatched layout, and one launch of the
 * two-level sweep advances every slot, the slots of one tile in consecutive blocks so that the
 * tile's packed byte plane and air-row bits are read once into one XCD's L2). Results per slot
 * are the same in both. No reference counterpart. */
#define PFDTD_BATCH_INTERLEAVED 0
#define PFDTD_BATCH_GROUPED 1
int pfdtd_set_batch_layout(pfdtd_ctx *ctx, int layout);
int pfdtd_set_source_slots(pfdtd_ctx *ctx, int n_src, const int32_t *slot);
int pfdtd_set_receiver_slots(pfdtd_ctx *ctx, int n_rec, const int32_t *slot);

/* ---- running ------------------------------------------------------------------------ */
/* Allocate (if needed) and zero the two pressure fields in the given precision.
 * scale_exp: fields are stored multiplied by 2^scale_exp (F16/BF16 only; 0 otherwise).
 * CudaMesh::makePartition's field allocation (cudaMesh.h:700-746) + resetPressures
 * (cudaMesh.h:782-791, App::resetPressureMesh src/App.cpp:397-401). */
int pfdtd_prepare_fields(pfdtd_ctx *ctx, int precision, int scale_exp);
int pfdtd_reset_fields(pfdtd_ctx *ctx);

/* Whole run from step 0: launchFDTD3d / launchFDTD3dDouble, src/kernels/kernels3d.cu:31-374.
 * Fields are zeroed first. Responses stay on the device until pfdtd_get_responses_*.
 * interrupt/progress may be NULL; they are polled between step bundles (every 100 steps for
 * progress). steps_done receives the number of completed steps (< n_steps when interrupted;
 * later samples are 0, kernels3d.cu:86-89). ms_elapsed (optional) receives device time of the
 * whole step loop measured with HIP events on the context's stream. */
int pfdtd_run(pfdtd_ctx *ctx, int precision, int n_steps, pfdtd_interrupt_cb interrupt,
              pfdtd_progress_cb progress, void *user, int *steps_done, double *ms_elapsed);

/* One step at time index `step` (source samples of that step, receivers written to response
 * sample `step`): launchFDTD3dStep, src/kernels/kernels3d.cu:376-482, used by App::executeStep
 * (src/App.cpp:407-436). direction is +1 or -1; when it differs from the previous call's the
 * field pointers are not flipped (time reversal, kernels3d.cu:462-465). Requires
 * pfdtd_prepare_fields and a response length (n_steps) set by pfdtd_begin_steps. */
int pfdtd_begin_steps(pfdtd_ctx *ctx, int precision, int n_steps);
int pfdtd_step(pfdtd_ctx *ctx, int step, int direction);
/* Steps first_step .. first_step + n_steps - 1 forward in one call: the same fields and
 * responses as that many pfdtd_step(ctx, s, +1) calls, advanced in triples or pairs by the
 * three- or two-level sweep where pfdtd_run would use them (one synchronisation at the end). first_step + n_steps may not
 * exceed the response length set by pfdtd_begin_steps. No reference counterpart: the
 * reference's capture loop (App::runCapture, src/App.cpp:350-379) launches one step at a time;
 * FDTD::App::runCapture calls this for the steps between two captures. */
int pfdtd_steps(pfdtd_ctx *ctx, int first_step, int n_steps);

/* Responses [n_rec][n_steps] of the last run or step sequence (kernels3d.cu:184-191). */
int pfdtd_get_responses_f32(pfdtd_ctx *ctx, float *out, int64_t capacity);
int pfdtd_get_responses_f64(pfdtd_ctx *ctx, double *out, int64_t capacity);
int pfdtd_get_response_shape(pfdtd_ctx *ctx, int *n_rec, int *n_steps);

/* Slice of the current pressure field and of the position bytes: captureSliceFast,
 * src/kernels/visualizationUtils.cu:130-254 (captureSliceKernel :500-527). orientation 0 = xy
 * (out[y*X+x] at z=slice), 1 = xz (out[z*X+x] at y=slice), 2 = yz (out[z*Y+y] at x=slice).
 * Values are unscaled FP32. PFDTD_E_ARG when the slice is out of range (the reference logs and
 * skips the capture). */
int pfdtd_capture_slice(pfdtd_ctx *ctx, int64_t slice, int orientation, float *p_out,
                        uint8_t *pos_out);
/* Largest plane, in elements, that pfdtd_capture_slice_noalloc accepts (every plane of a
 * 2048 x 1536 x 768 room fits). */
#define PFDTD_STATIC_CAPTURE_MAX 4194304
/* pfdtd_capture_slice without any device allocation. No counterpart in the reference: it is
 * for callers that cannot allocate device memory (tight capture or serving loops, harnesses
 * that forbid hipMalloc). The plane is staged in a static device buffer of
 * PFDTD_STATIC_CAPTURE_MAX * 5 bytes (20 MiB) that every GPU the library runs on holds until
 * the process ends; a plane larger than PFDTD_STATIC_CAPTURE_MAX elements is refused with
 * PFDTD_E_ARG. Every call on a GPU shares that GPU's buffer, so calls that overlap on one GPU
 * (from several host threads or contexts) race. The buffer keeps the last call's contents;
 * nothing in it needs clearing, since each call writes the whole plane before copying it out.
 * Same arguments, results and errors as pfdtd_capture_slice. */
int pfdtd_capture_slice_noalloc(pfdtd_ctx *ctx, int64_t slice, int orientation, float *p_out,
                                uint8_t *pos_out);
/* Whole current field as unscaled FP32: captureMesh, visualizationUtils.cu:111-128. */
int pfdtd_capture_field(pfdtd_ctx *ctx, float *out, int64_t capacity);
/* Whole current and past fields in storage precision (raw bytes), for checks. */
int pfdtd_get_fields_raw(pfdtd_ctx *ctx, void *p_out, void *p_past_out, int64_t capacity_bytes);

/* HBM streaming probe: the read, write and read+write ceilings the update kernels are measured
 * against. kind: PFDTD_PROBE_READ (sum of a), _WRITE (a = s), _COPY (b = a), _TRIAD
 * (a = b + s*c), _UPDATE (a = s*a + b in place: the leapfrog update's access pattern). Each
 * array has bytes_per_array bytes (rounded down to 16 KiB); one warm-up launch, then `iters`
 * timed launches (HIP events). Outputs: average launch time and the bytes one launch reads and
 * writes (arrays read + arrays written). The arrays are allocated and freed inside the call.
 * No reference counterpart: a measurement aid of this build. */
#define PFDTD_PROBE_READ 0
#define PFDTD_PROBE_WRITE 1
#define PFDTD_PROBE_COPY 2
#define PFDTD_PROBE_TRIAD 3
#define PFDTD_PROBE_UPDATE 4
int pfdtd_hbm_probe(int device, int kind, int64_t bytes_per_array, int iters,
                    double *ms_per_launch, double *bytes_moved_per_launch);

/* Timing of the update kernels inside pfdtd_run: when enabled, every update launch is bracketed
 * by HIP events on the context's stream. update_ms_total is their summed device time and
 * update_launches the number of time steps they advanced (a two-level launch counts 2). */
int pfdtd_set_kernel_timing(pfdtd_ctx *ctx, int enabled);
/* Steps between host polls of the interrupt/progress callbacks inside pfdtd_run (default 100,
 * PROGRESS_MOD in src/kernels/kernels3d.h:36). The reference polls the interrupt every step
 * (kernels3d.cu:86); polling here costs one stream synchronisation per interval. */
int pfdtd_set_poll_interval(pfdtd_ctx *ctx, int steps);
/* Time levels advanced per sweep of the field (1, 2 or 3; 0 = default: 3 for FP32 and FP64
 * (unbatched or grouped batches), else 2 -- the faster settings on MI355X) and slices per z-chunk of the sweep
 * (0 = automatic: the longest of 32 (16-bit storage: 22 for the two-column sweep, 16 for the
 * difference form and one column), 16 and 8 slices that gives the two-level sweep 4096 blocks; the three-level sweep 2048 blocks, from 32 slices (FP64: 96)
 * down). With 2, pfdtd_run advances pairs of
 * steps in one kernel that reads P^n, P^{n-1} and the byte planes once and writes P^{n+1} and
 * P^{n+2}; with 3 (forward update; the centred one and the difference form take 2), triples of
 * steps in one kernel that writes only P^{n+2} and P^{n+3}. Sources are injected and receivers
 * gathered inside the kernel; four field buffers. A remainder of one step takes a single-step
 * update (so does every step while a source sits on slice z = 0 or z = Z-1), four steps left
 * take two pairs. Non-grouped batches advance one level per sweep. Results are bit-identical
 * to single steps. No reference counterpart: the reference's loop (kernels3d.cu:83-191)
 * launches one update per step. */
int pfdtd_set_time_levels(pfdtd_ctx *ctx, int levels, int zchunk);
/* Block shape of the two-level sweep: PFDTD_SWEEP_AUTO (default), _COLUMN, _WIDE (FP32 only)
 * or _PAIR (16-bit storage on rooms of even X; elsewhere it acts as AUTO). The wave tiles of a block step z together so their
 * shared halo columns and rows meet in cache: wide blocks share more of them but leave up to
 * three tiles past the grid's last column, so they pay off on wide rooms. Results do not
 * depend on it. No reference counterpart (a tuning knob of the two-level sweep). */
int pfdtd_set_sweep_shape(pfdtd_ctx *ctx, int shape);
/* Update-kernel tiling: y rows per thread (1, 2, 4 or 8) and slices per z-chunk; (0, 0) = the
 * default, automatic: 4 rows and 4-slice chunks for FP32, 4 and 8 for FP64, 2 and 8 for 16-bit
 * storage and for interleaved batches. Results do not depend on it (every voxel's arithmetic
 * is the same); it only changes the streaming pattern. Replaces the fixed 32x4 block of App::initializeMesh
 * (src/App.cpp:193) and the one-block-per-slice grid of launchFDTD3d (kernels3d.cu:42-45). */
int pfdtd_set_update_tiling(pfdtd_ctx *ctx, int rows, int zchunk);
int pfdtd_get_kernel_timing(pfdtd_ctx *ctx, double *update_ms_total, int64_t *update_launches);


/* ---- serving pool ------------------------------------------------------------------- */
/* A pool holds many independent impulse-response jobs (rooms of different sizes, each with
 * its own materials, sources, receivers, step count and FP32 / FP16 / BF16 storage) resident in
 * one HBM arena, and advances all of them together: one step of every resident job of one
 * storage type is one launch of a multi-job update kernel plus one launch of a multi-job
 * gather/inject kernel; `bundle_steps` steps are captured into a hipGraph that is replayed
 * until a job joins or leaves. Jobs are admitted while their buffers fit in the arena's free
 * space and a slot is free (at most 64 resident), so a job submitted while others run starts
 * at the next bundle and overlaps them on the device. Every job's arithmetic is that of a
 * single-context run of the same inputs (bit-identical responses). A pool is driven by one
 * host thread at a time (calls on one pool must not overlap); pools are independent.
 * No reference counterpart: the reference runs one simulation per App (App::runSimulation,
 * src/App.cpp:308-348); a job's inputs and semantics are those of that call with the voxel
 * entry point, sources as App::addSource (src/App.h) with DATA signals, and receivers as
 * App::addReceiver. */
typedef struct pfdtd_pool pfdtd_pool;

typedef struct pfdtd_job {
  const uint8_t *labels;   /* [Z][Y][X]: 0 inside, m+1 solid of material m; outer layer solid */
  int32_t X, Y, Z;
  const float *materials;  /* [n_mat][n_coef] admittances */
  int32_t n_mat, n_coef;
  uint32_t octave;
  double lambda;
  int32_t flags;           /* PFDTD_MATSLOT_LITERAL | PFDTD_SOFT_OVERWRITE */
  int32_t n_src;
  const int32_t *src_xyz;  /* [n_src][3] grid indices */
  const int32_t *src_kind; /* PFDTD_SRC_* */
  int32_t n_samples;
  const float *signals;    /* [n_src][n_samples]; samples past the end are 0 */
  int32_t n_rec;
  const int32_t *rec_xyz;  /* [n_rec][3] grid indices */
  int32_t n_steps;
  int32_t precision;       /* PFDTD_F32, PFDTD_F16 or PFDTD_BF16 */
  int32_t scale_exp;       /* F16/BF16 storage scale 2^scale_exp (0 for F32) */
} pfdtd_job;

/* arena_bytes <= 0: 7/8 of the device's free memory. */
int pfdtd_pool_create(int device, int64_t arena_bytes, pfdtd_pool **out);
int pfdtd_pool_destroy(pfdtd_pool *pool);
/* Steps per bundle (default 32; rounded up to an even count, as jobs advance in rounds of two
 * steps), hipGraph replay (1, default) or direct launches (0), single-step update tiling
 * (rows 1/2/4, z-chunk; only while empty) and the resident-job limit (1..64). */
int pfdtd_pool_set_options(pfdtd_pool *pool, int bundle_steps, int use_graph, int rows,
                           int zchunk, int max_jobs);
/* Arena bytes a job occupies while resident. */
int pfdtd_pool_job_bytes(const pfdtd_job *job, int64_t *bytes);
int pfdtd_pool_mem(pfdtd_pool *pool, int64_t *arena_bytes, int64_t *free_bytes,
                   int64_t *largest_free);
/* Admits a job (its inputs are copied before the call returns). PFDTD_E_NOMEM: it does not
 * fit now (arena space or slots); submit again after jobs are released. PFDTD_E_DOMAIN: the
 * outer voxel layer is not solid or a label names a material outside the table. */
int pfdtd_pool_submit(pfdtd_pool *pool, const pfdtd_job *job, int64_t *job_id);
/* Enqueues one bundle for every resident unfinished job (returns without waiting, unless two
 * bundles are already in flight). *n_active = jobs in the bundle (0: nothing enqueued). */
int pfdtd_pool_advance(pfdtd_pool *pool, int *n_active);
/* Non-blocking: ids of jobs whose last step has completed since the previous poll. */
int pfdtd_pool_poll(pfdtd_pool *pool, int64_t *ids, int capacity, int *n_done);
/* Waits for every enqueued bundle. */
int pfdtd_pool_drain(pfdtd_pool *pool);
/* phase: 0 running, 1 last bundle enqueued, 2 finished. */
int pfdtd_pool_job_info(pfdtd_pool *pool, int64_t job_id, int *phase, int *n_rec, int *n_steps);
/* Responses of a finished job, [n_rec][n_steps] (layout of pfdtd_get_responses_f32). */
int pfdtd_pool_get_responses(pfdtd_pool *pool, int64_t job_id, float *out_f32, double *out_f64,
                             int64_t capacity);
/* Frees a finished job's arena space and slot. */
int pfdtd_pool_release(pfdtd_pool *pool, int64_t job_id);
int pfdtd_pool_stats(pfdtd_pool *pool, int64_t *bundles, int64_t *graph_builds,
                     double *voxel_steps, int *resident);

#ifdef __cplusplus
}
#endif
#endif
