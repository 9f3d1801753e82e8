/*
 * pfdtd_oracle.c -- CPU restatement of ParallelFDTD's SRL_FORWARD time-stepping core.
 *
 * TEST INFRASTRUCTURE ONLY (the parity oracle). Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load this library, always as the checker or the timed CPU
 * baseline, never as a product path. See pfdtd_oracle.h for the per-function pins.
 *
 * Build: see oracle/Makefile (gcc -O2 -fopenmp -ffp-contract=off).
 */
#include "pfdtd_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_PI 3.14159265358979323846 /* src/math/geomMath.h:68 */
#define ORC_COEF_NUM 20               /* MATERIAL_COEF_NUM, src/base/MaterialHandler.h:29 */
#define ORC_POS_MASK 0x7F             /* FORWARD_POSITION_MASK, src/kernels/cudaMesh.h:36 */
#define ORC_INSIDE_SHIFT 7            /* INSIDE_SWITCH, src/kernels/cudaMesh.h:35 */

int orc_version(void) { return 1; }

/* ------------------------------------------------------------------------------------ */
/* Classification                                                                        */
/* ------------------------------------------------------------------------------------ */

/* toBilbao, src/kernels/cudaMesh.cu:328-361 */
uint8_t orc_to_bilbao(uint8_t k, uint8_t *mat_inout) {
  if (k == 0) {
    if (mat_inout) *mat_inout = 0;
    return 0;
  }
  if (k <= 8) return (uint8_t)(3 | 0x80);
  if (k <= 20) return (uint8_t)(4 | 0x80);
  if (k <= 26) return (uint8_t)(5 | 0x80);
  if (k == 27) return (uint8_t)(6 | 0x80);
  return k; /* codes above 27 fall through every branch and are left unchanged */
}

void orc_classify_labels(const uint8_t *label, int64_t X, int64_t Y, int64_t Z, uint8_t *pos,
                         uint8_t *mat, uint64_t counts[9]) {
  const int64_t XY = X * Y;
  uint64_t c[9];
  memset(c, 0, sizeof(c));
  for (int64_t z = 0; z < Z; ++z) {
    for (int64_t y = 0; y < Y; ++y) {
      for (int64_t x = 0; x < X; ++x) {
        const int64_t i = z * XY + y * X + x;
        if (label[i] != 0) {
          pos[i] = 0;
          mat[i] = 0;
          continue;
        }
        /* neighbour order -x,+x,-y,+y,-z,+z; out-of-grid neighbours are solid material 0 */
        const int inb[6] = {x > 0, x < X - 1, y > 0, y < Y - 1, z > 0, z < Z - 1};
        const int64_t off[6] = {-1, 1, -X, X, -XY, XY};
        int K = 0, found = 0;
        unsigned m = 0;
        for (int d = 0; d < 6; ++d) {
          const unsigned lab = inb[d] ? label[i + off[d]] : 1u;
          if (lab == 0) {
            ++K;
          } else if (!found) {
            found = 1;
            m = inb[d] ? lab - 1u : 0u;
          }
        }
        pos[i] = (uint8_t)(0x80 | K);
        mat[i] = (uint8_t)(K < 6 ? m : 0);
        c[2 + K]++;
        if (pos[i] == 0x86)
          c[0]++;
        else
          c[1]++;
      }
    }
  }
  if (counts) memcpy(counts, c, sizeof(c));
}

/* toKowalczyk, src/kernels/cudaMesh.cu:363-480: voxelizer code -> centred-scheme position
 * byte (0x80 inside | DIR_X 0x01 | DIR_Y 0x02 | DIR_Z 0x04 | SIGN_X 0x10 | SIGN_Y 0x20 |
 * SIGN_Z 0x40, cudaMesh.h:35-44). Restated from the code comments' air-direction lists
 * ("Left/Right" = -x/+x, "In/Out" = -y/+y, "Down/Up" = -z/+z): an axis with exactly one air
 * neighbour gets its DIR bit, and its SIGN bit selects the present neighbour (+x, +y, -z). */
uint8_t orc_to_kowalczyk(uint8_t k, uint8_t *mat_inout) {
  /* air-neighbour sets per code: bit0 -x, bit1 +x, bit2 -y, bit3 +y, bit4 -z, bit5 +z */
  static const uint8_t present[28] = {
      0,
      0x01 | 0x04 | 0x10, 0x02 | 0x04 | 0x10, 0x01 | 0x08 | 0x10, 0x02 | 0x08 | 0x10, /* 1-4 down */
      0x01 | 0x04 | 0x20, 0x02 | 0x04 | 0x20, 0x01 | 0x08 | 0x20, 0x02 | 0x08 | 0x20, /* 5-8 up */
      0x10 | 0x03 | 0x04, 0x10 | 0x03 | 0x08, 0x10 | 0x01 | 0x0C, 0x10 | 0x02 | 0x0C, /* 9-12 */
      0x20 | 0x03 | 0x04, 0x20 | 0x03 | 0x08, 0x20 | 0x01 | 0x0C, 0x20 | 0x02 | 0x0C, /* 13-16 */
      0x30 | 0x01 | 0x04, 0x30 | 0x02 | 0x04, 0x30 | 0x01 | 0x08, 0x30 | 0x02 | 0x08, /* 17-20 */
      0x03 | 0x0C | 0x10, 0x03 | 0x08 | 0x30, 0x03 | 0x04 | 0x30, 0x02 | 0x0C | 0x30, /* 21-24 */
      0x01 | 0x0C | 0x30, 0x03 | 0x0C | 0x20, 0x3F};                                 /* 25-27 */
  if (k == 0) {
    if (mat_inout) *mat_inout = 0;
    return 0;
  }
  if (k > 27) return k;
  const uint8_t a = present[k];
  uint8_t pb = 0x80;
  if ((a & 0x03) == 0x01) pb |= 0x01;        /* only -x present: mirror -x */
  if ((a & 0x03) == 0x02) pb |= 0x01 | 0x10; /* only +x present */
  if ((a & 0x0C) == 0x04) pb |= 0x02;
  if ((a & 0x0C) == 0x08) pb |= 0x02 | 0x20;
  if ((a & 0x30) == 0x10) pb |= 0x04 | 0x40; /* only -z present: SIGN_Z ("down") */
  if ((a & 0x30) == 0x20) pb |= 0x04;
  return pb;
}

/* Centred-scheme classification of a label grid (same conventions as orc_classify_labels):
 * inside voxels get 0x80 | DIR/SIGN bits as orc_to_kowalczyk would give for their air-neighbour
 * set; an inside voxel with no air neighbour on some axis has no centred stencil (the voxelizer
 * never emits one) and is counted in counts[2]. counts[0] = air (0x80), counts[1] = boundary,
 * following calcBoundaries with air value 0x80 (cudaMesh.cu:176-186). Returns counts[2]. */
uint64_t orc_classify_labels_kow(const uint8_t *label, int64_t X, int64_t Y, int64_t Z,
                                 uint8_t *pos, uint8_t *mat, uint64_t counts[3]) {
  const int64_t XY = X * Y;
  uint64_t c[3] = {0, 0, 0};
  for (int64_t z = 0; z < Z; ++z)
    for (int64_t y = 0; y < Y; ++y)
      for (int64_t x = 0; x < X; ++x) {
        const int64_t i = z * XY + y * X + x;
        if (label[i] != 0) {
          pos[i] = 0;
          mat[i] = 0;
          continue;
        }
        const int inb[6] = {x > 0, x < X - 1, y > 0, y < Y - 1, z > 0, z < Z - 1};
        const int64_t off[6] = {-1, 1, -X, X, -XY, XY};
        unsigned air = 0, m = 0;
        int found = 0;
        for (int d = 0; d < 6; ++d) {
          const unsigned lab = inb[d] ? label[i + off[d]] : 1u;
          if (lab == 0)
            air |= 1u << d;
          else if (!found) {
            found = 1;
            m = inb[d] ? lab - 1u : 0u;
          }
        }
        uint8_t pb = 0x80;
        int bad = 0;
        for (int a = 0; a < 3; ++a) {
          const unsigned two = (air >> (2 * a)) & 3u;
          if (two == 0) bad = 1;
          if (two == 1) pb |= (uint8_t)((1u << a) | (a == 2 ? 0x40u : 0u));
          if (two == 2) pb |= (uint8_t)((1u << a) | (a == 2 ? 0u : (0x10u << a)));
        }
        pos[i] = pb;
        mat[i] = (uint8_t)(air == 0x3F ? 0 : m);
        if (bad) c[2]++;
        if (pb == 0x80)
          c[0]++;
        else
          c[1]++;
      }
  if (counts) memcpy(counts, c, sizeof(c));
  return c[2];
}

/* ------------------------------------------------------------------------------------ */
/* Parameters, materials, coordinates, source signals                                   */
/* ------------------------------------------------------------------------------------ */

/* SimulationParameters::getParameterPtr, SimulationParameters.cpp:379-387 */
void orc_param_vector_f(double lambda, unsigned octave, float out[4]) {
  out[0] = (float)lambda;
  out[1] = (float)(lambda * lambda);
  out[2] = 1.f / 3.f;
  out[3] = (float)octave;
}

/* SimulationParameters::getParameterPtrDouble, SimulationParameters.cpp:389-396 */
void orc_param_vector_d(double lambda, unsigned octave, double out[4]) {
  out[0] = lambda;
  out[1] = lambda * lambda;
  out[2] = (double)1 / (double)3;
  out[3] = (double)octave;
}

/* reflection2Admitance, src/global_includes.h:30-33 */
float orc_r2y_f(float r) { return (1.f - r) / (1.f + r); }

/* SimulationParameters::getDx, SimulationParameters.cpp:362-364 */
float orc_dx(float c, unsigned fs, double lambda) {
  return (float)((double)c / ((double)fs * lambda));
}

/* Position::getElementIdx, SrcRec.cpp:26-34; nv::ROUND, geomMath.h:126-128;
 * padding +1, SimulationParameters.cpp:200-224 */
void orc_element_idx(float px, float py, float pz, unsigned fs, float c, double lambda,
                     int add_padding, int32_t out[3]) {
  const float lam = (float)lambda;
  const float dx = c / ((float)fs * lam);
  out[0] = (int32_t)floorf(px / dx + 0.5f);
  out[1] = (int32_t)floorf(py / dx + 0.5f);
  out[2] = (int32_t)floorf(pz / dx + 0.5f);
  if (add_padding) {
    out[0] += 1;
    out[1] += 1;
    out[2] += 1;
  }
}

/* getRegularSourceSample, SimulationParameters.cpp:262-298 */
float orc_source_sample_f(int input_type, unsigned step, unsigned fs, const float *data,
                          unsigned n_data) {
  float sample = 0.f;
  switch (input_type) {
  case ORC_IN_IMPULSE:
    sample += (step == 1) ? 1.f : 0.f;
    break;
  case ORC_IN_GAUSSIAN: {
    const float t0 = 40;
    const float width = 4;
    const float e = ((float)step - t0) / width;
    sample += expf(-0.5f * (e * e));
    break;
  }
  case ORC_IN_DATA:
    sample += (data && step < n_data) ? data[step] : 0.f;
    break;
  case ORC_IN_SINE: {
    const float freq = 120;
    const float t = (float)step / (float)fs;
    sample += sinf(2.f * (float)ORC_PI * freq * t);
    break;
  }
  default:
    break;
  }
  return sample;
}

/* getRegularSourceSampleDouble, SimulationParameters.cpp:312-348 */
double orc_source_sample_d(int input_type, unsigned step, unsigned fs, const double *data,
                           unsigned n_data) {
  double sample = 0.0;
  switch (input_type) {
  case ORC_IN_IMPULSE:
    sample += (step == 1) ? 1.0 : 0.0;
    break;
  case ORC_IN_GAUSSIAN: {
    const double t0 = 40;
    const double width = 4;
    const double e = ((float)((double)step - t0) / width);
    sample += exp(-0.5f * (e * e));
    break;
  }
  case ORC_IN_DATA:
    sample += (data && step < n_data) ? data[step] : 0.0;
    break;
  case ORC_IN_SINE: {
    const double freq = 120;
    const double t = (double)step / (double)fs;
    sample += sin(2.f * (double)ORC_PI * freq * t);
    break;
  }
  default:
    break;
  }
  return sample;
}

/* getGridIrDataSample, SimulationParameters.cpp:191-198: 0 past the end */
static float grid_ir_at(const float *h, unsigned n, unsigned i) { return (h && i < n) ? h[i] : 0.f; }

/* getTransparentSourceSample, SimulationParameters.cpp:300-310 */
float orc_transparent_sample_f(int input_type, unsigned step, unsigned fs, const float *data,
                               unsigned n_data, const float *grid_ir, unsigned n_ir) {
  float sample = 0.f;
  for (unsigned i = 0; i < step; i++)
    sample += grid_ir_at(grid_ir, n_ir, step - i) *
              orc_source_sample_f(input_type, i, fs, data, n_data);
  return orc_source_sample_f(input_type, step, fs, data, n_data) - sample;
}

/* getTransparentSourceSampleDouble, SimulationParameters.cpp:350-360 */
double orc_transparent_sample_d(int input_type, unsigned step, unsigned fs, const double *data,
                                unsigned n_data, const float *grid_ir, unsigned n_ir) {
  double sample = 0.0;
  for (unsigned i = 0; i < step; i++)
    sample += grid_ir_at(grid_ir, n_ir, step - i) *
              orc_source_sample_d(input_type, i, fs, data, n_data);
  return orc_source_sample_d(input_type, step, fs, data, n_data) - sample;
}

/* kernels3d.cu:513 (literal, `mat*20*+params[3]`) vs kernels3d.cu:639 (intended) */
int64_t orc_material_slot(unsigned m, unsigned octave, int literal) {
  if (literal) {
    const unsigned base = m * 20u;
    return (int64_t)(unsigned)((float)base * (float)octave);
  }
  return (int64_t)m * ORC_COEF_NUM + (int64_t)octave;
}

/* ------------------------------------------------------------------------------------ */
/* Storage rounding (round to nearest even)                                             */
/* ------------------------------------------------------------------------------------ */

uint16_t orc_f32_to_f16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  const uint32_t sign = (x >> 16) & 0x8000u;
  const uint32_t ax = x & 0x7FFFFFFFu;
  if (ax >= 0x7F800000u) return (uint16_t)(sign | 0x7C00u | (ax > 0x7F800000u ? 0x200u : 0u));
  if (ax >= 0x477FF000u) return (uint16_t)(sign | 0x7C00u); /* >= 65520 rounds to inf */
  if (ax < 0x38800000u) {                                    /* below 2^-14: subnormal */
    if (ax < 0x33000000u) return (uint16_t)sign;             /* below 2^-25 */
    const uint32_t e = ax >> 23;
    const uint32_t m = (ax & 0x7FFFFFu) | 0x800000u;
    const uint32_t shift = 126u - e; /* value in units of 2^-24 = m >> shift */
    uint32_t q = m >> shift;
    const uint32_t rem = m & ((1u << shift) - 1u);
    const uint32_t half = 1u << (shift - 1u);
    if (rem > half || (rem == half && (q & 1u))) q++;
    return (uint16_t)(sign | q);
  }
  const uint32_t e = (ax >> 23) - 127u + 15u;
  const uint32_t m = ax & 0x7FFFFFu;
  uint32_t q = (e << 10) | (m >> 13);
  const uint32_t rem = m & 0x1FFFu;
  if (rem > 0x1000u || (rem == 0x1000u && (q & 1u))) q++;
  return (uint16_t)(sign | q);
}

float orc_f16_to_f32(uint16_t h) {
  const uint32_t sign = ((uint32_t)h & 0x8000u) << 16;
  const uint32_t e = ((uint32_t)h >> 10) & 0x1Fu;
  const uint32_t m = (uint32_t)h & 0x3FFu;
  uint32_t x;
  if (e == 0) {
    if (m == 0) {
      x = sign;
    } else {
      const float v = (float)m * 5.9604644775390625e-08f; /* m * 2^-24, exact */
      uint32_t vb;
      memcpy(&vb, &v, 4);
      x = sign | vb;
    }
  } else if (e == 31) {
    x = sign | 0x7F800000u | (m << 13);
  } else {
    x = sign | ((e - 15u + 127u) << 23) | (m << 13);
  }
  float f;
  memcpy(&f, &x, 4);
  return f;
}

uint16_t orc_f32_to_bf16(float f) {
  uint32_t x;
  memcpy(&x, &f, 4);
  if ((x & 0x7FFFFFFFu) > 0x7F800000u) return (uint16_t)((x >> 16) | 0x40u);
  const uint32_t lsb = (x >> 16) & 1u;
  x += 0x7FFFu + lsb;
  return (uint16_t)(x >> 16);
}

float orc_bf16_to_f32(uint16_t h) {
  const uint32_t x = (uint32_t)h << 16;
  float f;
  memcpy(&f, &x, 4);
  return f;
}

static inline float store_round(float v, int storage) {
  if (storage == ORC_STORE_F16) return orc_f16_to_f32(orc_f32_to_f16(v));
  if (storage == ORC_STORE_BF16) return orc_bf16_to_f32(orc_f32_to_bf16(v));
  return v;
}

/* ------------------------------------------------------------------------------------ */
/* Step loop                                                                             */
/* ------------------------------------------------------------------------------------ */

static int shell_is_outside(const uint8_t *pos, int64_t X, int64_t Y, int64_t Z) {
  const int64_t XY = X * Y;
  for (int64_t z = 0; z < Z; ++z)
    for (int64_t y = 0; y < Y; ++y) {
      const int edge_zy = (z == 0 || z == Z - 1 || y == 0 || y == Y - 1);
      for (int64_t x = 0; x < X; ++x) {
        if (!edge_zy && x != 0 && x != X - 1) continue;
        if (pos[z * XY + y * X + x] >> ORC_INSIDE_SHIFT) return 0;
      }
    }
  return 1;
}

static int in_grid(const int32_t *p, int64_t X, int64_t Y, int64_t Z) {
  return p[0] >= 0 && p[1] >= 0 && p[2] >= 0 && p[0] < X && p[1] < Y && p[2] < Z;
}

static int64_t lin(const int32_t *p, int64_t X, int64_t Y) {
  return ((int64_t)p[2] * Y + p[1]) * X + p[0];
}

static int check_common(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z,
                        int n_mat, unsigned octave, int literal, int n_src, int n_rec,
                        int n_steps) {
  if (!pos || !mat || X < 3 || Y < 3 || Z < 3 || n_mat < 1 || n_mat > 256 || n_src < 0 ||
      n_rec < 0 || n_steps < 0)
    return -1;
  if (octave >= ORC_COEF_NUM) return -1;
  /* every material byte present must address a slot inside the table */
  int seen[256];
  memset(seen, 0, sizeof(seen));
  const int64_t N = X * Y * Z;
  for (int64_t i = 0; i < N; ++i) seen[mat[i]] = 1;
  for (int m = 0; m < 256; ++m)
    if (seen[m] && orc_material_slot((unsigned)m, octave, (literal & ORC_SCHEME_KOWALCZYK) ? 0 : (literal & 1)) >= (int64_t)n_mat * ORC_COEF_NUM)
      return -1;
  if (!shell_is_outside(pos, X, Y, Z)) return -2;
  return 0;
}

/* One fdtd3dStdMaterials<float> sweep (kernels3d.cu:485-529) over slices z in [1, Z-2]. */
static void sweep_f(const uint8_t *pos, const uint8_t *mat, const float *P, float *Pp,
                    const float *beta_of_mat, float lam, float lam2, int64_t X, int64_t Y,
                    int64_t Z, int storage, int threads) {
  const int64_t XY = X * Y;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int64_t z = 1; z <= Z - 2; ++z) {
    for (int64_t y = 0; y < Y; ++y) {
      const int64_t row = z * XY + y * X;
      if (y == 0 || y == Y - 1) {
        for (int64_t x = 0; x < X; ++x) Pp[row + x] = 0.f;
        continue;
      }
      Pp[row] = 0.f;
      Pp[row + X - 1] = 0.f;
      for (int64_t x = 1; x <= X - 2; ++x) {
        const int64_t i = row + x;
        const uint8_t pb = pos[i];
        const float position = (float)(pb & ORC_POS_MASK);
        const float switch_bit = (float)(pb >> ORC_INSIDE_SHIFT);
        const float beta = (float)(0.5 * (double)(beta_of_mat[mat[i]] * (6.f - position) * lam));
        const float p_zp = P[i + XY];
        const float p_zm = P[i - XY];
        const float p_yp = P[i + X];
        const float p_ym = P[i - X];
        const float p_xp = P[i + 1];
        const float p_xm = P[i - 1];
        const float p = P[i];
        const float p_past = Pp[i];
        const float S = p_zp + p_zm + p_yp + p_ym + p_xp + p_xm;
        const float ret = switch_bit * (1.f / (1.f + beta)) *
                          ((2.f - position * lam2) * p + lam2 * S - (1.f - beta) * p_past);
        Pp[i] = store_round(ret, storage);
      }
    }
  }
}

/* The same sweep in the difference form (ORC_STORE_F16_DIFF): A = P^n and B = D^n in, C = P^{n+1}
 * out, B overwritten with D^{n+1}. Each node reads only its own B, so B is updated in place. */
static void sweep_diff_f(const uint8_t *pos, const uint8_t *mat, const float *A, float *B,
                         float *C, const float *beta_of_mat, float lam, float lam2, int64_t X,
                         int64_t Y, int64_t Z, int threads) {
  const int64_t XY = X * Y;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int64_t z = 1; z <= Z - 2; ++z) {
    for (int64_t y = 0; y < Y; ++y) {
      const int64_t row = z * XY + y * X;
      if (y == 0 || y == Y - 1) {
        for (int64_t x = 0; x < X; ++x) C[row + x] = B[row + x] = 0.f;
        continue;
      }
      C[row] = B[row] = 0.f;
      C[row + X - 1] = B[row + X - 1] = 0.f;
      for (int64_t x = 1; x <= X - 2; ++x) {
        const int64_t i = row + x;
        const uint8_t pb = pos[i];
        const float position = (float)(pb & ORC_POS_MASK);
        const float switch_bit = (float)(pb >> ORC_INSIDE_SHIFT);
        const float beta = (float)(0.5 * (double)(beta_of_mat[mat[i]] * (6.f - position) * lam));
        const float p = A[i];
        const float p_past = p - B[i];
        const float S = A[i + XY] + A[i - XY] + A[i + X] + A[i - X] + A[i + 1] + A[i - 1];
        const float ret = switch_bit * (1.f / (1.f + beta)) *
                          ((2.f - position * lam2) * p + lam2 * S - (1.f - beta) * p_past);
        C[i] = store_round(ret, ORC_STORE_F16);
        B[i] = store_round(ret - p, ORC_STORE_F16);
      }
    }
  }
}

/* The sweep in the sum-difference form (ORC_STORE_F16_SUMDIFF): Pc = P^n and Pp = P^{n-1}
 * reconstructed from (S, D) before the sweep; S and D are overwritten with the new state. */
static void sweep_sumdiff_f(const uint8_t *pos, const uint8_t *mat, const float *Pc,
                            const float *Pp, float *S, float *D, const float *beta_of_mat,
                            float lam, float lam2, int64_t X, int64_t Y, int64_t Z,
                            int threads) {
  const int64_t XY = X * Y;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int64_t z = 1; z <= Z - 2; ++z) {
    for (int64_t y = 0; y < Y; ++y) {
      const int64_t row = z * XY + y * X;
      if (y == 0 || y == Y - 1) {
        for (int64_t x = 0; x < X; ++x) S[row + x] = D[row + x] = 0.f;
        continue;
      }
      S[row] = D[row] = 0.f;
      S[row + X - 1] = D[row + X - 1] = 0.f;
      for (int64_t x = 1; x <= X - 2; ++x) {
        const int64_t i = row + x;
        const uint8_t pb = pos[i];
        const float position = (float)(pb & ORC_POS_MASK);
        const float switch_bit = (float)(pb >> ORC_INSIDE_SHIFT);
        const float beta = (float)(0.5 * (double)(beta_of_mat[mat[i]] * (6.f - position) * lam));
        const float p = Pc[i];
        const float sum = Pc[i + XY] + Pc[i - XY] + Pc[i + X] + Pc[i - X] + Pc[i + 1] + Pc[i - 1];
        const float ret = switch_bit * (1.f / (1.f + beta)) *
                          ((2.f - position * lam2) * p + lam2 * sum - (1.f - beta) * Pp[i]);
        S[i] = store_round(ret + p, ORC_STORE_F16);
        D[i] = store_round(ret - p, ORC_STORE_F16);
      }
    }
  }
}

/* fdtd3dStdMaterials<double> sweep (kernels3d.cu:485-529). */
static void sweep_d(const uint8_t *pos, const uint8_t *mat, const double *P, double *Pp,
                    const double *beta_of_mat, double lam, double lam2, int64_t X, int64_t Y,
                    int64_t Z, int threads) {
  const int64_t XY = X * Y;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int64_t z = 1; z <= Z - 2; ++z) {
    for (int64_t y = 0; y < Y; ++y) {
      const int64_t row = z * XY + y * X;
      if (y == 0 || y == Y - 1) {
        for (int64_t x = 0; x < X; ++x) Pp[row + x] = 0.0;
        continue;
      }
      Pp[row] = 0.0;
      Pp[row + X - 1] = 0.0;
      for (int64_t x = 1; x <= X - 2; ++x) {
        const int64_t i = row + x;
        const uint8_t pb = pos[i];
        const double position = (double)(pb & ORC_POS_MASK);
        const double switch_bit = (double)(pb >> ORC_INSIDE_SHIFT);
        const double beta = 0.5 * (beta_of_mat[mat[i]] * (6.f - position) * lam);
        const double p_zp = P[i + XY];
        const double p_zm = P[i - XY];
        const double p_yp = P[i + X];
        const double p_ym = P[i - X];
        const double p_xp = P[i + 1];
        const double p_xm = P[i - 1];
        const double p = P[i];
        const double p_past = Pp[i];
        const double S = p_zp + p_zm + p_yp + p_ym + p_xp + p_xm;
        const double ret = switch_bit * (1.f / (1.f + beta)) *
                           ((2.f - position * lam2) * p + lam2 * S - (1.f - beta) * p_past);
        Pp[i] = ret;
      }
    }
  }
}

/* One fdtd3dStdKowalczykMaterials<float> sweep (kernels3d.cu:609-665, update type SRL, the
 * centred boundary): the missing neighbour of an axis with a DIR bit is replaced by the present
 * one that its SIGN bit selects (p_x[0] = x-1, p_x[1] = x+1, p_y likewise, p_z[0] = z+1,
 * p_z[1] = z-1); beta = table[mat*20+octave] * lambda * (number of DIR bits). */
static void sweep_kow_f(const uint8_t *pos, const uint8_t *mat, const float *P, float *Pp,
                        const float *beta_of_mat, float lam, float lam2, int64_t X, int64_t Y,
                        int64_t Z, int storage, int threads) {
  const int64_t XY = X * Y;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int64_t z = 1; z <= Z - 2; ++z) {
    for (int64_t y = 0; y < Y; ++y) {
      const int64_t row = z * XY + y * X;
      if (y == 0 || y == Y - 1) {
        for (int64_t x = 0; x < X; ++x) Pp[row + x] = 0.f;
        continue;
      }
      Pp[row] = 0.f;
      Pp[row + X - 1] = 0.f;
      for (int64_t x = 1; x <= X - 2; ++x) {
        const int64_t i = row + x;
        const uint8_t pb = pos[i];
        const float switch_bit = (float)(pb >> ORC_INSIDE_SHIFT);
        const float dir_x = (float)(pb & 0x01);
        const float dir_y = (float)((pb & 0x02) >> 1);
        const float dir_z = (float)((pb & 0x04) >> 2);
        const float beta = beta_of_mat[mat[i]] * lam * (dir_x + dir_y + dir_z);
        float p_z[2], p_y[2], p_x[2];
        p_z[1] = P[i - XY];
        p_z[0] = P[i + XY];
        p_y[0] = P[i - X];
        p_y[1] = P[i + X];
        p_x[0] = P[i - 1];
        p_x[1] = P[i + 1];
        const float p = P[i];
        const float p_past = Pp[i];
        const int sx = (pb & 0x10) >> 4, sy = (pb & 0x20) >> 5, sz = (pb & 0x40) >> 6;
        const float s_b = p_x[sx] * dir_x + p_y[sy] * dir_y + p_z[sz] * dir_z;
        const float S = (p_x[0] + p_x[1] + p_y[0] + p_y[1] + p_z[0] + p_z[1] + s_b) * lam2;
        const float ret =
            (S - (2 - 6 * lam2) * p + (beta - 1.f) * p_past) * switch_bit * (1.f / (1.f + beta));
        Pp[i] = store_round(ret, storage);
      }
    }
  }
}

/* fdtd3dStdKowalczykMaterials<double> sweep (kernels3d.cu:609-665). */
static void sweep_kow_d(const uint8_t *pos, const uint8_t *mat, const double *P, double *Pp,
                        const double *beta_of_mat, double lam, double lam2, int64_t X, int64_t Y,
                        int64_t Z, int threads) {
  const int64_t XY = X * Y;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
  for (int64_t z = 1; z <= Z - 2; ++z) {
    for (int64_t y = 0; y < Y; ++y) {
      const int64_t row = z * XY + y * X;
      if (y == 0 || y == Y - 1) {
        for (int64_t x = 0; x < X; ++x) Pp[row + x] = 0.0;
        continue;
      }
      Pp[row] = 0.0;
      Pp[row + X - 1] = 0.0;
      for (int64_t x = 1; x <= X - 2; ++x) {
        const int64_t i = row + x;
        const uint8_t pb = pos[i];
        const double switch_bit = (double)(pb >> ORC_INSIDE_SHIFT);
        const double dir_x = (double)(pb & 0x01);
        const double dir_y = (double)((pb & 0x02) >> 1);
        const double dir_z = (double)((pb & 0x04) >> 2);
        const double beta = beta_of_mat[mat[i]] * lam * (dir_x + dir_y + dir_z);
        double p_z[2], p_y[2], p_x[2];
        p_z[1] = P[i - XY];
        p_z[0] = P[i + XY];
        p_y[0] = P[i - X];
        p_y[1] = P[i + X];
        p_x[0] = P[i - 1];
        p_x[1] = P[i + 1];
        const double p = P[i];
        const double p_past = Pp[i];
        const int sx = (pb & 0x10) >> 4, sy = (pb & 0x20) >> 5, sz = (pb & 0x40) >> 6;
        const double s_b = p_x[sx] * dir_x + p_y[sy] * dir_y + p_z[sz] * dir_z;
        const double S = (p_x[0] + p_x[1] + p_y[0] + p_y[1] + p_z[0] + p_z[1] + s_b) * lam2;
        Pp[i] = (S - (2 - 6 * lam2) * p + (beta - 1.f) * p_past) * switch_bit * (1.f / (1.f + beta));
      }
    }
  }
}

/* Step loop of the difference form: the source loop, the sweep and the receiver loop of orc_run_f
 * (kernels3d.cu:83-191) on the state (P, D). */
static int run_diff_f(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z,
                      const float *beta_of_mat, const float params[4], const int32_t *src_xyz,
                      const int32_t *src_type, const float *signals, int n_src,
                      const int32_t *rec_xyz, int n_rec, int n_steps, int literal, int scale_exp,
                      int threads, float *responses, float *P_out, float *D_out) {
  const int64_t N = X * Y * Z;
  float *A = (float *)calloc((size_t)N, sizeof(float));
  float *B = (float *)calloc((size_t)N, sizeof(float));
  float *Cn = (float *)calloc((size_t)N, sizeof(float));
  if (!A || !B || !Cn) {
    free(A);
    free(B);
    free(Cn);
    return -1;
  }
  const float scale = ldexpf(1.f, scale_exp);
  const float inv_scale = ldexpf(1.f, -scale_exp);
  if (responses) memset(responses, 0, sizeof(float) * (size_t)n_rec * (size_t)n_steps);
  for (int n = 0; n < n_steps; ++n) {
    for (int s = 0; s < n_src; ++s) {
      const int32_t *q = src_xyz + 3 * s;
      if (!in_grid(q, X, Y, Z)) continue;
      const int64_t i = lin(q, X, Y);
      const float smp = signals[(size_t)s * (size_t)n_steps + (size_t)n] * scale;
      const float past = A[i] - B[i];
      if (src_type[s] == ORC_SRC_HARD || (literal & 2))
        A[i] = store_round(smp, ORC_STORE_F16);
      else
        A[i] = store_round(A[i] + smp, ORC_STORE_F16);
      B[i] = store_round(A[i] - past, ORC_STORE_F16);
    }
    sweep_diff_f(pos, mat, A, B, Cn, beta_of_mat, params[0], params[1], X, Y, Z, threads);
    float *t = A;
    A = Cn;
    Cn = t;
    for (int r = 0; r < n_rec; ++r) {
      const int32_t *q = rec_xyz + 3 * r;
      if (!in_grid(q, X, Y, Z) || !responses) continue;
      responses[(size_t)r * (size_t)n_steps + (size_t)n] = A[lin(q, X, Y)] * inv_scale;
    }
  }
  if (P_out) memcpy(P_out, A, sizeof(float) * (size_t)N);
  if (D_out) memcpy(D_out, B, sizeof(float) * (size_t)N);
  free(A);
  free(B);
  free(Cn);
  return 0;
}

/* Step loop of the sum-difference form: the source loop, the sweep and the receiver loop of
 * orc_run_f (kernels3d.cu:83-191) on the state (S, D). */
static int run_sumdiff_f(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z,
                         const float *beta_of_mat, const float params[4], const int32_t *src_xyz,
                         const int32_t *src_type, const float *signals, int n_src,
                         const int32_t *rec_xyz, int n_rec, int n_steps, int literal,
                         int scale_exp, int threads, float *responses, float *S_out,
                         float *D_out) {
  const int64_t N = X * Y * Z;
  float *S = (float *)calloc((size_t)N, sizeof(float));
  float *D = (float *)calloc((size_t)N, sizeof(float));
  float *Pc = (float *)calloc((size_t)N, sizeof(float));
  float *Pp = (float *)calloc((size_t)N, sizeof(float));
  if (!S || !D || !Pc || !Pp) {
    free(S);
    free(D);
    free(Pc);
    free(Pp);
    return -1;
  }
  const float scale = ldexpf(1.f, scale_exp);
  const float inv_scale = ldexpf(1.f, -scale_exp);
  if (responses) memset(responses, 0, sizeof(float) * (size_t)n_rec * (size_t)n_steps);
  for (int n = 0; n < n_steps; ++n) {
    for (int s = 0; s < n_src; ++s) {
      const int32_t *q = src_xyz + 3 * s;
      if (!in_grid(q, X, Y, Z)) continue;
      const int64_t i = lin(q, X, Y);
      const float smp = signals[(size_t)s * (size_t)n_steps + (size_t)n] * scale;
      const float p = (S[i] + D[i]) * 0.5f;
      const float past = (S[i] - D[i]) * 0.5f;
      const float pn = (src_type[s] == ORC_SRC_HARD || (literal & 2)) ? smp : p + smp;
      S[i] = store_round(pn + past, ORC_STORE_F16);
      D[i] = store_round(pn - past, ORC_STORE_F16);
    }
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) schedule(static)
#endif
    for (int64_t i = 0; i < N; ++i) {
      Pc[i] = (S[i] + D[i]) * 0.5f;
      Pp[i] = (S[i] - D[i]) * 0.5f;
    }
    sweep_sumdiff_f(pos, mat, Pc, Pp, S, D, beta_of_mat, params[0], params[1], X, Y, Z, threads);
    for (int r = 0; r < n_rec; ++r) {
      const int32_t *q = rec_xyz + 3 * r;
      if (!in_grid(q, X, Y, Z) || !responses) continue;
      const int64_t i = lin(q, X, Y);
      responses[(size_t)r * (size_t)n_steps + (size_t)n] = (S[i] + D[i]) * 0.5f * inv_scale;
    }
  }
  if (S_out) memcpy(S_out, S, sizeof(float) * (size_t)N);
  if (D_out) memcpy(D_out, D, sizeof(float) * (size_t)N);
  free(S);
  free(D);
  free(Pc);
  free(Pp);
  return 0;
}

int orc_run_f(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z,
              const float *mat_table, int n_mat, unsigned octave, double lambda,
              const int32_t *src_xyz, const int32_t *src_type, const float *signals, int n_src,
              const int32_t *rec_xyz, int n_rec, int n_steps, int literal, int storage,
              int scale_exp, int threads, float *responses, float *P_out, float *Pp_out) {
  int rc = check_common(pos, mat, X, Y, Z, n_mat, octave, literal, n_src, n_rec, n_steps);
  if (rc) return rc;
  if (!mat_table || storage < 0 || storage > 4 || scale_exp < -100 || scale_exp > 100) return -1;
  if (storage >= ORC_STORE_F16_DIFF && (literal & ORC_SCHEME_KOWALCZYK)) return -1;
  if (threads < 1) threads = 1;
  float params[4];
  orc_param_vector_f(lambda, octave, params);
  float beta_of_mat[256];
  for (int m = 0; m < 256; ++m) {
    const int64_t slot = orc_material_slot((unsigned)m, octave, (literal & ORC_SCHEME_KOWALCZYK) ? 0 : (literal & 1));
    beta_of_mat[m] = slot < (int64_t)n_mat * ORC_COEF_NUM ? mat_table[slot] : 0.f;
  }
  if (storage == ORC_STORE_F16_SUMDIFF)
    return run_sumdiff_f(pos, mat, X, Y, Z, beta_of_mat, params, src_xyz, src_type, signals,
                         n_src, rec_xyz, n_rec, n_steps, literal, scale_exp, threads, responses,
                         P_out, Pp_out);
  if (storage == ORC_STORE_F16_DIFF)
    return run_diff_f(pos, mat, X, Y, Z, beta_of_mat, params, src_xyz, src_type, signals, n_src,
                      rec_xyz, n_rec, n_steps, literal, scale_exp, threads, responses, P_out,
                      Pp_out);
  const int64_t N = X * Y * Z;
  float *A = (float *)calloc((size_t)N, sizeof(float));
  float *B = (float *)calloc((size_t)N, sizeof(float));
  if (!A || !B) {
    free(A);
    free(B);
    return -1;
  }
  float *P = A, *Pp = B;
  const float scale = ldexpf(1.f, scale_exp);
  const float inv_scale = ldexpf(1.f, -scale_exp);
  if (responses) memset(responses, 0, sizeof(float) * (size_t)n_rec * (size_t)n_steps);
  for (int n = 0; n < n_steps; ++n) {
    /* source loop, kernels3d.cu:93-104 */
    for (int s = 0; s < n_src; ++s) {
      const int32_t *q = src_xyz + 3 * s;
      if (!in_grid(q, X, Y, Z)) continue;
      const int64_t i = lin(q, X, Y);
      const float smp = signals[(size_t)s * (size_t)n_steps + (size_t)n] * scale;
      if (src_type[s] == ORC_SRC_HARD || (literal & 2))
        P[i] = store_round(smp, storage);
      else
        P[i] = store_round(P[i] + smp, storage);
    }
    if (literal & ORC_SCHEME_KOWALCZYK)
      sweep_kow_f(pos, mat, P, Pp, beta_of_mat, params[0], params[1], X, Y, Z, storage, threads);
    else
      sweep_f(pos, mat, P, Pp, beta_of_mat, params[0], params[1], X, Y, Z, storage, threads);
    float *t = P; /* flipPressurePointers, cudaMesh.h:766-779 */
    P = Pp;
    Pp = t;
    /* receiver loop, kernels3d.cu:164-173 */
    for (int r = 0; r < n_rec; ++r) {
      const int32_t *q = rec_xyz + 3 * r;
      if (!in_grid(q, X, Y, Z) || !responses) continue;
      responses[(size_t)r * (size_t)n_steps + (size_t)n] = P[lin(q, X, Y)] * inv_scale;
    }
  }
  if (P_out) memcpy(P_out, P, sizeof(float) * (size_t)N);
  if (Pp_out) memcpy(Pp_out, Pp, sizeof(float) * (size_t)N);
  free(A);
  free(B);
  return 0;
}

int orc_run_d(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z,
              const double *mat_table, int n_mat, unsigned octave, double lambda,
              const int32_t *src_xyz, const int32_t *src_type, const double *signals, int n_src,
              const int32_t *rec_xyz, int n_rec, int n_steps, int literal, int threads,
              double *responses, double *P_out, double *Pp_out) {
  int rc = check_common(pos, mat, X, Y, Z, n_mat, octave, literal, n_src, n_rec, n_steps);
  if (rc) return rc;
  if (!mat_table) return -1;
  if (threads < 1) threads = 1;
  double params[4];
  orc_param_vector_d(lambda, octave, params);
  double beta_of_mat[256];
  for (int m = 0; m < 256; ++m) {
    const int64_t slot = orc_material_slot((unsigned)m, octave, (literal & ORC_SCHEME_KOWALCZYK) ? 0 : (literal & 1));
    beta_of_mat[m] = slot < (int64_t)n_mat * ORC_COEF_NUM ? mat_table[slot] : 0.0;
  }
  const int64_t N = X * Y * Z;
  double *A = (double *)calloc((size_t)N, sizeof(double));
  double *B = (double *)calloc((size_t)N, sizeof(double));
  if (!A || !B) {
    free(A);
    free(B);
    return -1;
  }
  double *P = A, *Pp = B;
  if (responses) memset(responses, 0, sizeof(double) * (size_t)n_rec * (size_t)n_steps);
  for (int n = 0; n < n_steps; ++n) {
    for (int s = 0; s < n_src; ++s) {
      const int32_t *q = src_xyz + 3 * s;
      if (!in_grid(q, X, Y, Z)) continue;
      const int64_t i = lin(q, X, Y);
      const double smp = signals[(size_t)s * (size_t)n_steps + (size_t)n];
      if (src_type[s] == ORC_SRC_HARD || (literal & 2))
        P[i] = smp;
      else
        P[i] = P[i] + smp;
    }
    if (literal & ORC_SCHEME_KOWALCZYK)
      sweep_kow_d(pos, mat, P, Pp, beta_of_mat, params[0], params[1], X, Y, Z, threads);
    else
      sweep_d(pos, mat, P, Pp, beta_of_mat, params[0], params[1], X, Y, Z, threads);
    double *t = P;
    P = Pp;
    Pp = t;
    for (int r = 0; r < n_rec; ++r) {
      const int32_t *q = rec_xyz + 3 * r;
      if (!in_grid(q, X, Y, Z) || !responses) continue;
      responses[(size_t)r * (size_t)n_steps + (size_t)n] = P[lin(q, X, Y)];
    }
  }
  if (P_out) memcpy(P_out, P, sizeof(double) * (size_t)N);
  if (Pp_out) memcpy(Pp_out, Pp, sizeof(double) * (size_t)N);
  free(A);
  free(B);
  return 0;
}

/* captureSliceFast / captureSliceKernel, visualizationUtils.cu:130-254, :500-527.
 * xy (orientation 0) is a plain copy of slice z (getSlice / getPositionSlice, cudaMesh.h:601-646).
 * xz and yz are gathered by captureSliceKernel into calloc'd planes with the limit
 * dim * (partition size - 1) (visualizationUtils.cu:178, :190, :522): with the whole grid in one
 * partition of Z slices, the rows z = 0..Z-2 are written and row z = Z-1 stays 0. */
int orc_capture_slice_f(const float *P, const uint8_t *pos, int64_t X, int64_t Y, int64_t Z,
                        int64_t slice, int orientation, float *out_p, uint8_t *out_pos) {
  const int64_t XY = X * Y;
  if (slice < 0) return -1;
  if (orientation == 0) {
    if (slice >= Z) return -1;
    for (int64_t y = 0; y < Y; ++y)
      for (int64_t x = 0; x < X; ++x) {
        out_p[y * X + x] = P[slice * XY + y * X + x];
        out_pos[y * X + x] = pos[slice * XY + y * X + x];
      }
    return 0;
  }
  if (orientation == 1) {
    if (slice >= Y) return -1;
    for (int64_t z = 0; z < Z; ++z)
      for (int64_t x = 0; x < X; ++x) {
        const int keep = z < Z - 1;
        out_p[z * X + x] = keep ? P[z * XY + slice * X + x] : 0.f;
        out_pos[z * X + x] = keep ? pos[z * XY + slice * X + x] : 0;
      }
    return 0;
  }
  if (orientation == 2) {
    if (slice >= X) return -1;
    for (int64_t z = 0; z < Z; ++z)
      for (int64_t y = 0; y < Y; ++y) {
        const int keep = z < Z - 1;
        out_p[z * Y + y] = keep ? P[z * XY + y * X + slice] : 0.f;
        out_pos[z * Y + y] = keep ? pos[z * XY + y * X + slice] : 0;
      }
    return 0;
  }
  return -1;
}

/* launchFDTD3dStep called once per entry of steps[] / dirs[] (kernels3d.cu:376-482), as
 * App::executeStep drives it (App.cpp:407-436). */
int orc_step_calls_f(const uint8_t *pos, const uint8_t *mat, int64_t X, int64_t Y, int64_t Z,
                     const float *mat_table, int n_mat, unsigned octave, double lambda,
                     const int32_t *src_xyz, const int32_t *src_type, const float *signals,
                     int n_src, const int32_t *rec_xyz, int n_rec, int n_steps, int literal,
                     int storage, int scale_exp, int threads, const int32_t *steps,
                     const int32_t *dirs, int n_calls, float *responses, float *P_out,
                     float *Pp_out) {
  int rc = check_common(pos, mat, X, Y, Z, n_mat, octave, literal, n_src, n_rec, n_steps);
  if (rc) return rc;
  if (!mat_table || storage < 0 || storage > 2 || scale_exp < -100 || scale_exp > 100) return -1;
  if (n_calls < 0 || (n_calls > 0 && (!steps || !dirs))) return -1;
  for (int c = 0; c < n_calls; ++c)
    if (steps[c] < 0 || steps[c] >= n_steps || (dirs[c] != 1 && dirs[c] != -1)) return -1;
  if (threads < 1) threads = 1;
  float params[4];
  orc_param_vector_f(lambda, octave, params);
  float beta_of_mat[256];
  for (int m = 0; m < 256; ++m) {
    const int64_t slot = orc_material_slot((unsigned)m, octave, (literal & ORC_SCHEME_KOWALCZYK) ? 0 : (literal & 1));
    beta_of_mat[m] = slot < (int64_t)n_mat * ORC_COEF_NUM ? mat_table[slot] : 0.f;
  }
  const int64_t N = X * Y * Z;
  float *A = (float *)calloc((size_t)N, sizeof(float));
  float *B = (float *)calloc((size_t)N, sizeof(float));
  if (!A || !B) {
    free(A);
    free(B);
    return -1;
  }
  float *P = A, *Pp = B;
  const float scale = ldexpf(1.f, scale_exp);
  const float inv_scale = ldexpf(1.f, -scale_exp);
  if (responses) memset(responses, 0, sizeof(float) * (size_t)n_rec * (size_t)n_steps);
  int past_direction = 1; /* the function-static past_step_directon, kernels3d.cu:386 */
  for (int c = 0; c < n_calls; ++c) {
    const int n = steps[c];
    /* source loop: samples of step `n` into the current field, kernels3d.cu:400-409 */
    for (int s = 0; s < n_src; ++s) {
      const int32_t *q = src_xyz + 3 * s;
      if (!in_grid(q, X, Y, Z)) continue;
      const int64_t i = lin(q, X, Y);
      const float smp = signals[(size_t)s * (size_t)n_steps + (size_t)n] * scale;
      if (src_type[s] == ORC_SRC_HARD || (literal & 2))
        P[i] = store_round(smp, storage);
      else
        P[i] = store_round(P[i] + smp, storage);
    }
    if (literal & ORC_SCHEME_KOWALCZYK)
      sweep_kow_f(pos, mat, P, Pp, beta_of_mat, params[0], params[1], X, Y, Z, storage, threads);
    else
      sweep_f(pos, mat, P, Pp, beta_of_mat, params[0], params[1], X, Y, Z, storage, threads);
    /* flip only while the direction is unchanged, kernels3d.cu:462-465 */
    if (past_direction == dirs[c]) {
      float *t = P;
      P = Pp;
      Pp = t;
    }
    past_direction = dirs[c];
    /* receiver loop: sample `n` from the current field, kernels3d.cu:470-475 */
    for (int r = 0; r < n_rec; ++r) {
      const int32_t *q = rec_xyz + 3 * r;
      if (!in_grid(q, X, Y, Z) || !responses) continue;
      responses[(size_t)r * (size_t)n_steps + (size_t)n] = P[lin(q, X, Y)] * inv_scale;
    }
  }
  if (P_out) memcpy(P_out, P, sizeof(float) * (size_t)N);
  if (Pp_out) memcpy(Pp_out, Pp, sizeof(float) * (size_t)N);
  free(A);
  free(B);
  return 0;
}

/* ------------------------------------------------------------------------------------ */
/* Triangle-mesh solid voxelization (definition in pfdtd_oracle.h; no reference source)    */
/* ------------------------------------------------------------------------------------ */

#define VOX_SUB 4096
#define VOX_MAX_DIM 32768
#define VOX_NONE 0xFFFFFFFFu

/* Fixed-point vertices (q may be NULL) and per-axis maximum. */
static int vox_fixed(const float *v, int64_t nv, float dx, int32_t *q, int64_t ext[3]) {
  if (!(dx > 0.f) || !isfinite(dx) || nv <= 0 || !v) return -1;
  float mn[3];
  for (int a = 0; a < 3; ++a) mn[a] = v[a];
  for (int64_t i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      const float x = v[3 * i + a];
      if (!isfinite(x)) return -1;
      if (x < mn[a]) mn[a] = x;
    }
  ext[0] = ext[1] = ext[2] = 0;
  const double lim = (double)(VOX_MAX_DIM - 3) * VOX_SUB;
  for (int64_t i = 0; i < nv; ++i)
    for (int a = 0; a < 3; ++a) {
      const float r = v[3 * i + a] - mn[a];
      const double s = ((double)r * (double)VOX_SUB) / (double)dx;
      if (!(s <= lim)) return -1;
      const long long k = llrint(s);
      if (q) q[3 * i + a] = (int32_t)k;
      if (k > ext[a]) ext[a] = k;
    }
  return 0;
}

int orc_voxel_grid(const float *vertices, int64_t n_vertices, float dx, const int32_t pad[3],
                   int64_t dims[3]) {
  int64_t ext[3];
  if (vox_fixed(vertices, n_vertices, dx, NULL, ext)) return -1;
  for (int a = 0; a < 3; ++a) {
    const int64_t p = (pad && pad[a] > 0) ? pad[a] : 1;
    int64_t n = ext[a] / VOX_SUB + 3;
    n = (n + p - 1) / p * p;
    if (n > VOX_MAX_DIM) return -1;
    dims[a] = n;
  }
  return 0;
}

static int64_t vox_edge(int64_t ux, int64_t uy, int64_t vx, int64_t vy, int64_t px, int64_t py) {
  return (vx - ux) * (py - uy) - (vy - uy) * (px - ux);
}

/* a point on an edge belongs to the triangle on the edge's left (CCW) side iff the edge
 * direction (ex, ey) points down, or left-to-right when horizontal: the minimum-side edges
 * of a CCW triangle, so an axis-aligned square [0,L)^2 owns the columns of [0,L)^2 */
static int vox_incl(int64_t w, int64_t ex, int64_t ey) {
  return w > 0 || (w == 0 && (ey < 0 || (ey == 0 && ex > 0)));
}

static void vox_axis(const int32_t *q, const uint32_t *tri, int64_t nt, const int64_t d[3],
                     int axis, uint8_t *flip, uint32_t *link) {
  const int au = (axis + 1) % 3, av = (axis + 2) % 3;
  for (int64_t t = 0; t < nt; ++t) {
    const int32_t *A = q + 3 * (int64_t)tri[3 * t], *B = q + 3 * (int64_t)tri[3 * t + 1],
                  *C = q + 3 * (int64_t)tri[3 * t + 2];
    const int64_t ax = A[au], ay = A[av], bx = B[au], by = B[av], cx = C[au], cy = C[av];
    const int64_t area = vox_edge(ax, ay, bx, by, cx, cy);
    if (area == 0) continue;
    const int64_t s = area > 0 ? 1 : -1;
    int64_t umin = ax, umax = ax, vmin = ay, vmax = ay;
    if (bx < umin) umin = bx;
    if (cx < umin) umin = cx;
    if (bx > umax) umax = bx;
    if (cx > umax) umax = cx;
    if (by < vmin) vmin = by;
    if (cy < vmin) vmin = cy;
    if (by > vmax) vmax = by;
    if (cy > vmax) vmax = cy;
    int64_t u0 = (umin + VOX_SUB - 1) / VOX_SUB + 1, u1 = umax / VOX_SUB + 1;
    int64_t v0 = (vmin + VOX_SUB - 1) / VOX_SUB + 1, v1 = vmax / VOX_SUB + 1;
    if (u1 > d[au] - 1) u1 = d[au] - 1;
    if (v1 > d[av] - 1) v1 = d[av] - 1;
    for (int64_t nv_ = v0; nv_ <= v1; ++nv_)
      for (int64_t nu = u0; nu <= u1; ++nu) {
        const int64_t px = (nu - 1) * VOX_SUB, py = (nv_ - 1) * VOX_SUB;
        const int64_t wa = s * vox_edge(bx, by, cx, cy, px, py);
        const int64_t wb = s * vox_edge(cx, cy, ax, ay, px, py);
        const int64_t wc = s * vox_edge(ax, ay, bx, by, px, py);
        if (!vox_incl(wa, s * (cx - bx), s * (cy - by)) ||
            !vox_incl(wb, s * (ax - cx), s * (ay - cy)) ||
            !vox_incl(wc, s * (bx - ax), s * (by - ay)))
          continue;
        const double W = (double)(s * area);
        const double num = (double)wb * (double)(B[axis] - A[axis]) +
                           (double)wc * (double)(C[axis] - A[axis]);
        const double h = (double)A[axis] + num / W;
        int64_t n0 = (int64_t)ceil(h / (double)VOX_SUB) + 1;
        if (n0 < 1) n0 = 1;
        if (n0 >= d[axis]) continue;
        int64_t c[3];
        c[axis] = n0;
        c[au] = nu;
        c[av] = nv_;
        const int64_t idx = (c[2] * d[1] + c[1]) * d[0] + c[0];
        if (flip) flip[idx] ^= 1;
        if ((uint32_t)t < link[idx]) link[idx] = (uint32_t)t;
      }
  }
}

int orc_voxelize(const float *vertices, int64_t n_vertices, const uint32_t *tri,
                 int64_t n_tri, const uint8_t *tmat, float dx, const int32_t pad[3],
                 int64_t dims[3], uint8_t *labels) {
  if (orc_voxel_grid(vertices, n_vertices, dx, pad, dims)) return -1;
  for (int64_t i = 0; i < 3 * n_tri; ++i)
    if ((int64_t)tri[i] >= n_vertices) return -2;
  if (tmat)
    for (int64_t t = 0; t < n_tri; ++t)
      if (tmat[t] > 254) return -3;
  int64_t ext[3];
  int32_t *q = (int32_t *)malloc(sizeof(int32_t) * 3 * (size_t)n_vertices);
  vox_fixed(vertices, n_vertices, dx, q, ext);
  const int64_t X = dims[0], Y = dims[1], Z = dims[2], N = X * Y * Z;
  uint8_t *air = (uint8_t *)calloc((size_t)N, 1);
  uint32_t *link = (uint32_t *)malloc(sizeof(uint32_t) * 3 * (size_t)N);
  memset(link, 0xFF, sizeof(uint32_t) * 3 * (size_t)N);
  for (int axis = 0; axis < 3; ++axis)
    vox_axis(q, tri, n_tri, dims, axis, axis == 0 ? air : NULL, link + axis * N);
  for (int64_t r = 0; r < Y * Z; ++r) { /* prefix parity along each x row */
    uint8_t p = 0;
    for (int64_t i = 0; i < X; ++i) {
      p ^= air[r * X + i];
      air[r * X + i] = p;
    }
  }
  for (int64_t z = 0; z < Z; ++z)
    for (int64_t y = 0; y < Y; ++y)
      for (int64_t x = 0; x < X; ++x) {
        const int64_t idx = (z * Y + y) * X + x;
        if (air[idx]) {
          labels[idx] = 0;
          continue;
        }
        /* neighbour offsets and the node holding the link: -x,+x,-y,+y,-z,+z */
        const int ok[6] = {x > 0, x + 1 < X, y > 0, y + 1 < Y, z > 0, z + 1 < Z};
        const int64_t nb[6] = {idx - 1, idx + 1, idx - X, idx + X, idx - X * Y, idx + X * Y};
        const int64_t lk[6] = {idx, idx + 1, N + idx, N + idx + X, 2 * N + idx,
                               2 * N + idx + X * Y};
        uint8_t lab = 1;
        for (int k = 0; k < 6; ++k) {
          if (!ok[k] || !air[nb[k]]) continue;
          const uint32_t t = link[lk[k]];
          if (t != VOX_NONE) lab = (uint8_t)(1 + (tmat ? tmat[t] : 0));
          break;
        }
        labels[idx] = lab;
      }
  free(air);
  free(link);
  free(q);
  return 0;
}
