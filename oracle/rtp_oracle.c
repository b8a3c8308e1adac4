/*
 * CPU oracle — TEST INFRASTRUCTURE ONLY (see rtp_oracle.h for the pinned builtins).
 * Build: gcc -O2 -ffp-contract=off -fno-fast-math -fopenmp (oracle/Makefile).
 *
 * Every function restates one reference kernel or host routine; the citation above each
 * names the file:line it follows.  Loops over particles stand for the OpenCL NDRange.
 */
#include "rtp_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_FLOAT_EPS 0.00000001f /* define.cl:6 */
#define ORC_PI_F 3.1415927f       /* utils/Math.hpp:69 */

int orc_num_threads(void)
{
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  if (n > 0)
    omp_set_num_threads(n);
#else
  (void)n;
#endif
}

/* ------------------------------------------------------------------------------------
 * Host-side constants
 * ---------------------------------------------------------------------------------- */

/* utils/Utils.cpp:24-29: std::fixed << setprecision(10) << val, then the OpenCL compiler
 * parses the literal (correctly rounded, like strtof). */
float orc_float_roundtrip(float v)
{
  char buf[128];
  snprintf(buf, sizeof(buf), "%.10f", (double)v);
  return strtof(buf, NULL);
}

/* physics/ocl/Fluids.cpp:104-119 and physics/ocl/Boids.cpp:104-113. */
void orc_make_consts(const float box[3], const int32_t res[3], uint32_t maxPartsInCell,
                     float gravity, orc_consts* out)
{
  memset(out, 0, sizeof(*out));
  const float effectRadius = box[0] / (float)res[0];                         /* Fluids.cpp:104 */
  out->effectRadius = orc_float_roundtrip(effectRadius);                      /* :107 */
  out->effectRadiusSq =
    orc_float_roundtrip(1.0f * box[0] * box[0] / (float)(res[0] * res[0])); /* Boids.cpp:104 */
  for (int a = 0; a < 3; ++a) {
    out->absWall[a] = orc_float_roundtrip(box[a] / 2.0f);                     /* :108-110 */
    out->res[a] = res[a];                                                     /* :111-113 */
  }
  out->cellSize = orc_float_roundtrip(box[0] / (float)res[0]);                /* :114 */
  out->numCells = (uint32_t)res[0] * (uint32_t)res[1] * (uint32_t)res[2];     /* Model.hpp:89 */
  out->maxPartsInCell = maxPartsInCell;                                       /* :116 */
  out->poly6Coeff =
    orc_float_roundtrip(315.0f / (64.0f * ORC_PI_F * powf(effectRadius, 9.f))); /* :117 */
  out->spikyCoeff = orc_float_roundtrip(15.0f / (ORC_PI_F * powf(effectRadius, 6.f))); /* :118 */
  out->maxVel = orc_float_roundtrip(30.0f);                                   /* :119 */
  out->gravity = gravity;                                                     /* define.cl:8 */
}

/* ------------------------------------------------------------------------------------
 * Pinned OpenCL builtins
 * ---------------------------------------------------------------------------------- */

static inline float dot4(orc_f4 a, orc_f4 b)
{
  return a.x * b.x + a.y * b.y + a.z * b.z + a.w * b.w;
}

static inline float len4(orc_f4 v) { return sqrtf(dot4(v, v)); }

static inline orc_f4 sub4(orc_f4 a, orc_f4 b)
{
  orc_f4 r = { a.x - b.x, a.y - b.y, a.z - b.z, a.w - b.w };
  return r;
}

static inline orc_f4 add4(orc_f4 a, orc_f4 b)
{
  orc_f4 r = { a.x + b.x, a.y + b.y, a.z + b.z, a.w + b.w };
  return r;
}

static inline orc_f4 mul4s(orc_f4 a, float s)
{
  orc_f4 r = { a.x * s, a.y * s, a.z * s, a.w * s };
  return r;
}

static inline orc_f4 muls4(float s, orc_f4 a)
{
  orc_f4 r = { s * a.x, s * a.y, s * a.z, s * a.w };
  return r;
}

static inline orc_f4 div4s(orc_f4 a, float s)
{
  orc_f4 r = { a.x / s, a.y / s, a.z / s, a.w / s };
  return r;
}

static inline orc_f4 normalize4(orc_f4 v)
{
  if (v.x == 0.0f && v.y == 0.0f && v.z == 0.0f && v.w == 0.0f)
    return v;
  return div4s(v, sqrtf(dot4(v, v)));
}

/* OpenCL cross(float4, float4): xyz cross product, w = 0. */
static inline orc_f4 cross4(orc_f4 a, orc_f4 b)
{
  orc_f4 r = { a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x, 0.0f };
  return r;
}

static inline float clampf(float x, float lo, float hi) { return fminf(fmaxf(x, lo), hi); }

static inline float stepf(float edge, float x) { return x < edge ? 0.0f : 1.0f; }

float orc_pown(float x, uint32_t n)
{
  if (n == 0)
    return 1.0f;
  float r = x;
  for (uint32_t k = 1; k < n; ++k)
    r = r * x;
  return r;
}

/* ------------------------------------------------------------------------------------
 * sph.cl
 * ---------------------------------------------------------------------------------- */

/* sph.cl:10-14 */
static inline float poly6(orc_f4 vec, float h, const orc_consts* c)
{
  const float r = len4(vec);
  return (1.0f - stepf(h, r)) * c->poly6Coeff * orc_pown(h * h - r * r, 3);
}

/* sph.cl:16-19 */
static inline float poly6L(float r, float h, const orc_consts* c)
{
  return (1.0f - stepf(h, r)) * c->poly6Coeff * orc_pown(h * h - r * r, 3);
}

/* sph.cl:26-34 */
static inline orc_f4 grad_spiky(orc_f4 vec, float h, const orc_consts* c)
{
  const float r = len4(vec);
  if (r <= ORC_FLOAT_EPS) {
    orc_f4 z = { 0.0f, 0.0f, 0.0f, 0.0f };
    return z;
  }
  const float t = 1.0f - stepf(h, r);
  const float p2 = orc_pown(h - r, 2);
  orc_f4 g = mul4s(vec, t);
  g = mul4s(g, c->spikyCoeff);
  g = mul4s(g, -3.0f);
  g = mul4s(g, p2);
  return div4s(g, r);
}

/* fluids.cl:51-57 */
static inline float art_pressure(orc_f4 vec, const orc_fluid_params* p, const orc_consts* c)
{
  if (p->isArtPressureEnabled == 0)
    return 0.0f;
  const float h = c->effectRadius;
  const float ratio = poly6(vec, h, c) / poly6L(p->artPressureRadius * h, h, c);
  return -p->artPressureCoeff * orc_pown(ratio, p->artPressureExp);
}

float orc_poly6(const orc_f4* vec, const orc_consts* c) { return poly6(*vec, c->effectRadius, c); }

void orc_grad_spiky(const orc_f4* vec, const orc_consts* c, orc_f4* out)
{
  *out = grad_spiky(*vec, c->effectRadius, c);
}

float orc_art_pressure(const orc_f4* vec, const orc_fluid_params* p, const orc_consts* c)
{
  return art_pressure(*vec, p, c);
}

/* ------------------------------------------------------------------------------------
 * grid.cl
 * ---------------------------------------------------------------------------------- */

/* grid.cl:14-23 getCell3DIndexFromPos */
static inline void cell3d(orc_f4 pos, const orc_consts* c, uint32_t out[3])
{
  const float p[3] = { pos.x, pos.y, pos.z };
  for (int a = 0; a < 3; ++a) {
    const float w = c->absWall[a];
    const float shifted = clampf(p[a], -w, w) + w;
    out[a] = (uint32_t)floorf(shifted / c->cellSize);
  }
}

void orc_cell3d(const orc_f4* pos, const orc_consts* c, uint32_t out[3]) { cell3d(*pos, c, out); }

/* grid.cl:28-38 getCell1DIndexFromPos */
static inline uint32_t cell1d(orc_f4 pos, const orc_consts* c)
{
  uint32_t c3[3];
  cell3d(pos, c, c3);
  const uint32_t ry = (uint32_t)c->res[1], rz = (uint32_t)c->res[2];
  return c3[0] * rz * ry + c3[1] * rz + c3[2];
}

/* grid.cl:65-71 */
void orc_reset_cell_ids(uint32_t* key, uint32_t maxN, const orc_consts* c)
{
  for (uint32_t i = 0; i < maxN; ++i)
    key[i] = c->numCells * 2 + i;
}

/* grid.cl:76-86 */
void orc_fill_cell_ids(const orc_f4* pos, uint32_t* key, uint32_t n, const orc_consts* c)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    key[i] = cell1d(pos[i], c);
}

/* grid.cl:91-96 */
void orc_reset_start_end(orc_u2* se, const orc_consts* c)
{
  for (uint32_t i = 0; i < c->numCells; ++i) {
    se[i].x = 1;
    se[i].y = 0;
  }
}

/* grid.cl:101-117 (work-item 0 never writes: quirk kept) */
void orc_fill_start(const uint32_t* key, orc_u2* se, uint32_t n, const orc_consts* c)
{
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t cur = key[i];
    if (i > 0 && cur < c->numCells) {
      if (cur != key[i - 1])
        se[cur].x = i;
    }
  }
}

/* grid.cl:122-138.  `ID != get_global_size(0)` is always true, so key[ID+1] is read for the
 * last work-item; keyLen (= maxN) bounds that read, and a read past the end counts as a
 * different key. */
void orc_fill_end(const uint32_t* key, orc_u2* se, uint32_t n, uint32_t keyLen, const orc_consts* c)
{
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t cur = key[i];
    if (cur < c->numCells) {
      const int differs = (i + 1 >= keyLen) ? 1 : (cur != key[i + 1]);
      if (differs)
        se[cur].y = i;
    }
  }
}

/* grid.cl:143-152 */
void orc_adjust_end(orc_u2* se, const orc_consts* c)
{
  for (uint32_t i = 0; i < c->numCells; ++i) {
    const orc_u2 s = se[i];
    if (s.y > s.x) {
      const uint32_t span = s.y - s.x;
      const uint32_t capped = span < c->maxPartsInCell ? span : c->maxPartsInCell;
      se[i].y = s.x + capped;
    }
  }
}

/* ------------------------------------------------------------------------------------
 * radixSort.cl + RadixSort.cpp:122-189.  256 radix, 8 bits, 4 passes over all maxN keys.
 * The reference splits each pass over 512 work-items (128 groups x 4 items) that each own
 * a contiguous chunk of maxN/512 keys; its histogram layout hist[digit*512 + item] makes the
 * exclusive scan give every (digit, chunk) its base, so the scatter is a stable counting
 * sort per pass.  The serial loop below visits the chunks in the same order.
 * ---------------------------------------------------------------------------------- */
int orc_radix_sort(uint32_t* key, uint32_t* idx, uint32_t maxN)
{
  if (maxN % 512u != 0)
    return -1;
  uint32_t* keyTmp = (uint32_t*)malloc(sizeof(uint32_t) * (maxN ? maxN : 1));
  uint32_t* idxTmp = (uint32_t*)malloc(sizeof(uint32_t) * (maxN ? maxN : 1));
  if (!keyTmp || !idxTmp) {
    free(keyTmp);
    free(idxTmp);
    return -2;
  }
  for (uint32_t i = 0; i < maxN; ++i) /* resetIndex radixSort.cl:171-174 */
    idx[i] = i;
  uint32_t *kin = key, *iin = idx, *kout = keyTmp, *iout = idxTmp;
  for (int pass = 0; pass < 4; ++pass) {
    uint32_t hist[256];
    memset(hist, 0, sizeof(hist));
    for (uint32_t i = 0; i < maxN; ++i) /* histogram radixSort.cl:19-54 */
      ++hist[(kin[i] >> (pass * 8)) & 255u];
    uint32_t run = 0; /* scan + merge radixSort.cl:59-130 */
    for (int d = 0; d < 256; ++d) {
      const uint32_t h = hist[d];
      hist[d] = run;
      run += h;
    }
    for (uint32_t i = 0; i < maxN; ++i) { /* reorder radixSort.cl:132-169 */
      const uint32_t k = kin[i];
      const uint32_t pos = hist[(k >> (pass * 8)) & 255u]++;
      kout[pos] = k;
      iout[pos] = iin[i];
    }
    uint32_t* t = kin; /* RadixSort.cpp:159-160 swapBuffers */
    kin = kout;
    kout = t;
    t = iin;
    iin = iout;
    iout = t;
  }
  /* four swaps: the sorted data is back in key/idx */
  free(keyTmp);
  free(idxTmp);
  return 0;
}

/* RadixSort.cpp:163-175 (copyBuffer into a temp) + permutateFloat4 radixSort.cl:179-188 */
void orc_permute_f4(const uint32_t* idx, orc_f4* buf, orc_f4* tmp, uint32_t maxN)
{
  memcpy(tmp, buf, sizeof(orc_f4) * maxN);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)maxN; ++i)
    buf[i] = tmp[idx[i]];
}

/* ------------------------------------------------------------------------------------
 * fluids.cl
 * ---------------------------------------------------------------------------------- */

/* The 27 cells visited by every PBF neighbour kernel, in the reference's loop order
 * (fluids.cl:101-113: x outer, z inner, periodic wrap; the range test after the modulo is
 * dead code and is omitted). */
static inline void fld_neighbour_cells(orc_f4 pos, const orc_consts* c, int32_t cells[27])
{
  uint32_t c3[3];
  cell3d(pos, c, c3);
  const int32_t rx = c->res[0], ry = c->res[1], rz = c->res[2];
  int k = 0;
  for (int iX = -1; iX <= 1; ++iX)
    for (int iY = -1; iY <= 1; ++iY)
      for (int iZ = -1; iZ <= 1; ++iZ) {
        const int32_t nx = ((int32_t)c3[0] + iX + rx) % rx;
        const int32_t ny = ((int32_t)c3[1] + iY + ry) % ry;
        const int32_t nz = ((int32_t)c3[2] + iZ + rz) % rz;
        cells[k++] = (nx * ry + ny) * rz + nz;
      }
}

#define ORC_FLD_FOR_NEIGHBOURS(c, pos, e)                                                 \
  int32_t cells_[27];                                                                    \
  fld_neighbour_cells(pos, c, cells_);                                                   \
  for (int k_ = 0; k_ < 27; ++k_)                                                        \
    for (uint32_t e = se[cells_[k_]].x; e <= se[cells_[k_]].y; ++e)

/* fluids.cl:62-74 */
void orc_fld_predict(const orc_f4* pos, const orc_f4* vel, const orc_fluid_params* p,
                     orc_f4* pred, uint32_t n, const orc_consts* c)
{
  const orc_f4 g = { 0.0f, -c->gravity, 0.0f, 0.0f }; /* define.cl:9 GRAVITY_ACC */
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 newVel = add4(vel[i], mul4s(g, p->timeStep));
    pred[i] = add4(pos[i], mul4s(newVel, p->timeStep));
  }
}

/* fluids.cl:435-439 */
void orc_fld_boundary(orc_f4* pred, uint32_t n, const orc_consts* c)
{
  const float lox = -c->absWall[0] + 0.01f, loy = -c->absWall[1] + 0.01f,
              loz = -c->absWall[2] + 0.01f;
  const float hix = c->absWall[0] - 0.1f, hiy = c->absWall[1] - 0.1f, hiz = c->absWall[2] - 0.1f;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    orc_f4 q = pred[i];
    q.x = clampf(q.x, lox, hix);
    q.y = clampf(q.y, loy, hiy);
    q.z = clampf(q.z, loz, hiz);
    q.w = clampf(q.w, 0.0f, 0.0f);
    pred[i] = q;
  }
}

/* fluids.cl:80-126 */
void orc_fld_density(const orc_f4* pred, const orc_u2* se, const orc_fluid_params* p,
                     float* density, uint32_t n, const orc_consts* c)
{
  (void)p;
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    float fluidDensity = 0.0f;
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      fluidDensity += poly6(sub4(pos, pred[e]), h, c);
    }
    density[i] = fluidDensity;
  }
}

/* fluids.cl:131-193 */
void orc_fld_lambda(const orc_f4* pred, const float* density, const orc_u2* se,
                    const orc_fluid_params* p, float* lambda, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const float densityC = density[i] / p->restDensity - 1.0f;
    orc_f4 sumGradCi = { 0.0f, 0.0f, 0.0f, 0.0f };
    float sumSqGradC = 0.0f;
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      const orc_f4 vec = sub4(pos, pred[e]);
      const orc_f4 grad = grad_spiky(vec, h, c);
      sumGradCi = add4(sumGradCi, grad);
      sumSqGradC += dot4(grad, grad);
    }
    sumSqGradC += dot4(sumGradCi, sumGradCi);
    sumSqGradC /= p->restDensity * p->restDensity;
    lambda[i] = -densityC / (sumSqGradC + p->relaxCFM);
  }
}

/* fluids.cl:198-247 */
void orc_fld_correction(const float* lambda, const orc_u2* se, const orc_f4* pred,
                        const orc_fluid_params* p, orc_f4* corrOut, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const float lambdaI = lambda[i];
    orc_f4 corr = { 0.0f, 0.0f, 0.0f, 0.0f };
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      const orc_f4 vec = sub4(pos, pred[e]);
      const float s = lambdaI + lambda[e] + art_pressure(vec, p, c);
      corr = add4(corr, muls4(s, grad_spiky(vec, h, c)));
    }
    corrOut[i] = div4s(corr, p->restDensity);
  }
}

/* fluids.cl:252-258 */
void orc_fld_correct_pos(const orc_f4* corr, orc_f4* pred, uint32_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    pred[i] = add4(pred[i], corr[i]);
}

/* fluids.cl:263-273 */
void orc_fld_update_vel(const orc_f4* pred, const orc_f4* pos, const orc_fluid_params* p,
                        orc_f4* vel, uint32_t n, const orc_consts* c)
{
  const float denom = p->timeStep + ORC_FLOAT_EPS;
  const float mv = c->maxVel;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 d = div4s(sub4(pred[i], pos[i]), denom);
    orc_f4 v = { clampf(d.x, -mv, mv), clampf(d.y, -mv, mv), clampf(d.z, -mv, mv),
                 clampf(d.w, -mv, mv) };
    vel[i] = v;
  }
}

/* fluids.cl:278-324 */
void orc_fld_vorticity(const orc_f4* pred, const orc_u2* se, const orc_f4* vel,
                       const orc_fluid_params* p, orc_f4* vortOut, uint32_t n, const orc_consts* c)
{
  (void)p;
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const orc_f4 velocity = vel[i];
    orc_f4 vort = { 0.0f, 0.0f, 0.0f, 0.0f };
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      vort = add4(vort, cross4(sub4(vel[e], velocity), grad_spiky(sub4(pos, pred[e]), h, c)));
    }
    vortOut[i] = vort;
  }
}

/* fluids.cl:329-377.  Each work-item writes only its own vel and reads no other vel. */
void orc_fld_confinement(const orc_f4* pred, const orc_u2* se, const orc_f4* vortIn,
                         const orc_fluid_params* p, orc_f4* vel, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const orc_f4 vorticity = vortIn[i];
    orc_f4 nv = { 0.0f, 0.0f, 0.0f, 0.0f };
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      nv = add4(nv, muls4(len4(vortIn[e]), grad_spiky(sub4(pos, pred[e]), h, c)));
    }
    const orc_f4 conf = mul4s(muls4(p->vorticityConfCoeff, cross4(normalize4(nv), vorticity)),
                              p->timeStep);
    vel[i] = add4(vel[i], conf);
  }
}

/* fluids.cl:383-430 */
void orc_fld_xsph(const orc_f4* pred, const orc_u2* se, const orc_f4* velIn,
                  const orc_fluid_params* p, orc_f4* velOut, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const orc_f4 velocity = velIn[i];
    orc_f4 visc = { 0.0f, 0.0f, 0.0f, 0.0f };
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      visc = add4(visc, mul4s(sub4(velIn[e], velocity), poly6(sub4(pos, pred[e]), h, c)));
    }
    velOut[i] = add4(velocity, muls4(p->xsphViscosityCoeff, visc));
  }
}

/* fluids.cl:444-450 */
void orc_fld_update_pos(const orc_f4* pred, orc_f4* pos, uint32_t n)
{
  memcpy(pos, pred, sizeof(orc_f4) * n);
}

/* Inputs of the learned warm-start (defined by this project, no reference counterpart).
 * The sums follow fld_computeDensity / fld_computeConstraintFactor (fluids.cl:80-193). */
void orc_fld_features(const orc_f4* pred, const float* density, const orc_u2* se,
                      const float* lambda, const orc_f4* vel, const orc_fluid_params* p,
                      float* feat, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
  const float lo[3] = { -c->absWall[0] + 0.01f, -c->absWall[1] + 0.01f, -c->absWall[2] + 0.01f };
  const float hi[3] = { c->absWall[0] - 0.1f, c->absWall[1] - 0.1f, c->absWall[2] - 0.1f };
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    orc_f4 sumGradCi = { 0.0f, 0.0f, 0.0f, 0.0f };
    float sumSqGradC = 0.0f, count = 0.0f;
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      const orc_f4 vec = sub4(pos, pred[e]);
      if (len4(vec) < h)
        count += 1.0f;
      const orc_f4 grad = grad_spiky(vec, h, c);
      sumGradCi = add4(sumGradCi, grad);
      sumSqGradC += dot4(grad, grad);
    }
    const float selfSq = dot4(sumGradCi, sumGradCi);
    sumSqGradC += selfSq;
    sumSqGradC /= p->restDensity * p->restDensity;
    const float w[6] = { pos.x - lo[0], hi[0] - pos.x, pos.y - lo[1], hi[1] - pos.y,
                         pos.z - lo[2], hi[2] - pos.z };
    float m0 = 3.0e38f, m1 = 3.0e38f;
    for (int k = 0; k < 6; ++k) {
      const float d = fmaxf(w[k], 0.0f);
      if (d < m0) {
        m1 = m0;
        m0 = d;
      } else if (d < m1) {
        m1 = d;
      }
    }
    const orc_f4 v = vel[i];
    float* f = feat + (size_t)i * 8u;
    f[0] = density[i] / p->restDensity - 1.0f;
    f[1] = lambda[i];
    f[2] = count;
    f[3] = sumSqGradC;
    f[4] = selfSq / (p->restDensity * p->restDensity);
    f[5] = sqrtf(v.x * v.x + v.y * v.y + v.z * v.z);
    f[6] = m0;
    f[7] = m1;
  }
}

/* ------------------------------------------------------------------------------------
 * boids.cl
 * ---------------------------------------------------------------------------------- */

/* boids.cl:46-132 */
void orc_bd_rules3d(const orc_f4* position, const orc_f4* velocity, const orc_u2* se,
                    const orc_boids_params* prm, orc_f4* acc, uint32_t n, const orc_consts* c)
{
  const int32_t rx = c->res[0], ry = c->res[1], rz = c->res[2];
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = position[i];
    uint32_t c3[3];
    cell3d(pos, c, c3);
    int count = 0;
    orc_f4 newAcc = { 0.0f, 0.0f, 0.0f, 0.0f };
    orc_f4 avgPos = { 0.0f, 0.0f, 0.0f, 0.0f };
    orc_f4 avgVel = { 0.0f, 0.0f, 0.0f, 0.0f };
    orc_f4 repulse = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int iX = -1; iX <= 1; ++iX)
      for (int iY = -1; iY <= 1; ++iY)
        for (int iZ = -1; iZ <= 1; ++iZ) {
          const int32_t nx = (int32_t)c3[0] + iX, ny = (int32_t)c3[1] + iY,
                        nz = (int32_t)c3[2] + iZ;
          if (nx < 0 || ny < 0 || nz < 0 || nx >= rx || ny >= ry || nz >= rz)
            continue;
          const int32_t cn = (nx * ry + ny) * rz + nz;
          for (uint32_t e = se[cn].x; e <= se[cn].y; ++e) {
            const orc_f4 posN = position[e];
            const orc_f4 vec = sub4(pos, posN);
            const float sq = dot4(vec, vec);
            if (sq < c->effectRadiusSq && sq > ORC_FLOAT_EPS) {
              avgPos = add4(avgPos, posN);
              avgVel = add4(avgVel, normalize4(velocity[e]));
              repulse = add4(repulse, div4s(vec, sq));
              ++count;
            }
          }
        }
    if (count != 0) {
      avgPos = div4s(avgPos, (float)count);
      avgPos = sub4(avgPos, pos);
      avgPos = mul4s(normalize4(avgPos), prm->velocityScale);
      avgVel = mul4s(normalize4(avgVel), prm->velocityScale);
      repulse = mul4s(normalize4(repulse), prm->velocityScale);
      newAcc = add4(add4(mul4s(avgVel, prm->alignmentScale), mul4s(repulse, prm->separationScale)),
                    mul4s(avgPos, prm->cohesionScale));
    }
    acc[i] = newAcc;
  }
}

/* boids.cl:137-221 */
void orc_bd_rules2d(const orc_f4* position, const orc_f4* velocity, const orc_u2* se,
                    const orc_boids_params* prm, orc_f4* acc, uint32_t n, const orc_consts* c)
{
  const int32_t rx = c->res[0], ry = c->res[1];
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = position[i];
    uint32_t c3[3];
    cell3d(pos, c, c3);
    int count = 0;
    orc_f4 newAcc = { 0.0f, 0.0f, 0.0f, 0.0f };
    orc_f4 avgPos = { 0.0f, 0.0f, 0.0f, 0.0f };
    orc_f4 avgVel = { 0.0f, 0.0f, 0.0f, 0.0f };
    orc_f4 repulse = { 0.0f, 0.0f, 0.0f, 0.0f };
    for (int iY = -1; iY <= 1; ++iY)
      for (int iZ = -1; iZ <= 1; ++iZ) {
        const int32_t nx = (int32_t)c3[0], ny = (int32_t)c3[1] + iY, nz = (int32_t)c3[2] + iZ;
        if (nx < 0 || ny < 0 || nz < 0 || nx >= rx || ny >= rx || nz >= rx)
          continue;
        const uint32_t cn = (uint32_t)((rx / 2 * rx + ny) * ry + nz);
        if (cn >= c->numCells)
          continue;
        for (uint32_t e = se[cn].x; e <= se[cn].y; ++e) {
          const orc_f4 posN = position[e];
          const orc_f4 vec = sub4(pos, posN);
          const float sq = dot4(vec, vec);
          if (sq < c->effectRadiusSq && sq > ORC_FLOAT_EPS) {
            avgPos = add4(avgPos, posN);
            avgVel = add4(avgVel, normalize4(velocity[e]));
            repulse = add4(repulse, div4s(vec, sq));
            ++count;
          }
        }
      }
    if (count != 0) {
      avgPos = div4s(avgPos, (float)count);
      avgPos = sub4(avgPos, pos);
      avgPos = mul4s(normalize4(avgPos), prm->velocityScale);
      avgVel = mul4s(normalize4(avgVel), prm->velocityScale);
      repulse = mul4s(normalize4(repulse), prm->velocityScale);
      newAcc = add4(add4(mul4s(avgVel, prm->alignmentScale), mul4s(repulse, prm->separationScale)),
                    mul4s(avgPos, prm->cohesionScale));
    }
    acc[i] = newAcc;
  }
}

/* boids.cl:226-241 (MAX_STEERING 0.5f, boids.cl:10) */
void orc_bd_add_target(const orc_f4* pos, const float target[3], float radius, int32_t sign,
                       orc_f4* acc, uint32_t n)
{
  const orc_f4 t = { target[0], target[1], target[2], 0.0f };
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 vec = sub4(t, pos[i]);
    const float dist = len4(vec);
    if (dist < radius)
      acc[i] = add4(acc[i], mul4s(muls4((float)sign, vec), clampf(1.3f / dist, 0.0f, 1.4f * 0.5f)));
  }
}

/* boids.cl:289-315 */
void orc_bd_update_pos_periodic(const orc_f4* vel, float dt, orc_f4* pos, uint32_t n,
                                const orc_consts* c)
{
  const float wx = c->absWall[0], wy = c->absWall[1], wz = c->absWall[2];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 np = add4(pos[i], mul4s(vel[i], dt));
    orc_f4 cp = { clampf(np.x, -wx, wx), clampf(np.y, -wy, wy), clampf(np.z, -wz, wz),
                  clampf(np.w, 0.0f, 0.0f) };
    if (!(cp.x == np.x))
      cp.x *= -1.0f;
    if (!(cp.y == np.y))
      cp.y *= -1.0f;
    if (!(cp.z == np.z))
      cp.z *= -1.0f;
    pos[i] = cp;
  }
}

/* boids.cl:246-259 */
void orc_bd_update_vel(const orc_f4* acc, float dt, float maxVel, orc_f4* vel, uint32_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 newVel = add4(vel[i], mul4s(acc[i], dt));
    const float norm = clampf(len4(newVel), 0.2f * maxVel, maxVel);
    vel[i] = mul4s(normalize4(newVel), norm);
  }
}

/* boids.cl:264-284 */
void orc_bd_update_pos_wall(orc_f4* vel, float dt, orc_f4* pos, uint32_t n, const orc_consts* c)
{
  const float wx = c->absWall[0], wy = c->absWall[1], wz = c->absWall[2];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 np = add4(pos[i], mul4s(vel[i], dt));
    orc_f4 cp = { clampf(np.x, -wx, wx), clampf(np.y, -wy, wy), clampf(np.z, -wz, wz),
                  clampf(np.w, 0.0f, 0.0f) };
    pos[i] = cp;
    if (!(cp.x == np.x && cp.y == np.y && cp.z == np.z))
      vel[i] = mul4s(vel[i], -0.5f);
  }
}

/* ------------------------------------------------------------------------------------
 * Host step drivers
 * ---------------------------------------------------------------------------------- */

/* physics/ocl/Fluids.cpp:409-457, the physics of one frame.  col (p_col), when given,
 * rides along in the cell sort's permutation like the other payloads (:417). */
static int fluids_physics(orc_fluids* s, orc_f4* col)
{
  const orc_consts* c = &s->c;
  const uint32_t n = s->n;
  orc_fld_predict(s->pos, s->vel, &s->p, s->pred, n, c);             /* :412 */
  orc_fill_cell_ids(s->pred, s->key, n, c);                            /* :415 */
  const int rc = orc_radix_sort(s->key, s->idx, s->maxN);              /* :417 */
  if (rc)
    return rc;
  orc_permute_f4(s->idx, s->pos, s->tmp, s->maxN);                     /* p_pos */
  if (col)
    orc_permute_f4(s->idx, col, s->tmp, s->maxN);                      /* p_col */
  orc_permute_f4(s->idx, s->vel, s->tmp, s->maxN);                     /* p_vel */
  orc_permute_f4(s->idx, s->pred, s->tmp, s->maxN);                    /* p_predPos */
  orc_reset_start_end(s->se, c);                                       /* :419 */
  orc_fill_start(s->key, s->se, n, c);                                 /* :420 */
  orc_fill_end(s->key, s->se, n, s->maxN, c);                          /* :421 */
  orc_adjust_end(s->se, c);                                            /* :423-424 */
  for (uint32_t it = 0; it < s->iters; ++it) {                         /* :427-439 */
    orc_fld_boundary(s->pred, n, c);
    orc_fld_density(s->pred, s->se, &s->p, s->density, n, c);
    orc_fld_lambda(s->pred, s->density, s->se, &s->p, s->lambda, n, c);
    orc_fld_correction(s->lambda, s->se, s->pred, &s->p, s->corr, n, c);
    orc_fld_correct_pos(s->corr, s->pred, n);
  }
  orc_fld_update_vel(s->pred, s->pos, &s->p, s->vel, n, c);           /* :442 */
  if (s->p.isVorticityConfEnabled) {                                   /* :444-454 */
    orc_fld_vorticity(s->pred, s->se, s->vel, &s->p, s->vort, n, c);
    orc_fld_confinement(s->pred, s->se, s->vort, &s->p, s->vel, n, c);
    memcpy(s->velIn, s->vel, sizeof(orc_f4) * s->maxN);
    orc_fld_xsph(s->pred, s->se, s->velIn, &s->p, s->vel, n, c);
  }
  orc_fld_update_pos(s->pred, s->pos, n);                              /* :457 */
  return 0;
}

/* physics/ocl/Fluids.cpp:400-471 (pause, colour, grid detector and camera sort omitted). */
int orc_fluids_update(orc_fluids* s) { return fluids_physics(s, NULL); }

/* physics/ocl/Boids.cpp:332-372, the step of one frame.  col (p_col), when given, rides
 * along in the cell sort's permutation (:337). */
static int boids_physics(orc_boids* s, orc_f4* col)
{
  const orc_consts* c = &s->c;
  const uint32_t n = s->n;
  orc_fill_cell_ids(s->pos, s->key, n, c);                             /* :335 */
  const int rc = orc_radix_sort(s->key, s->idx, s->maxN);              /* :337 */
  if (rc)
    return rc;
  orc_permute_f4(s->idx, s->pos, s->tmp, s->maxN);
  if (col)
    orc_permute_f4(s->idx, col, s->tmp, s->maxN);
  orc_permute_f4(s->idx, s->vel, s->tmp, s->maxN);
  orc_permute_f4(s->idx, s->acc, s->tmp, s->maxN);
  orc_reset_start_end(s->se, c);                                       /* :339-344 */
  orc_fill_start(s->key, s->se, n, c);
  orc_fill_end(s->key, s->se, n, s->maxN, c);
  orc_adjust_end(s->se, c);
  if (s->dim == 2)                                                     /* :346-349 */
    orc_bd_rules2d(s->pos, s->vel, s->se, &s->p, s->acc, n, c);
  else
    orc_bd_rules3d(s->pos, s->vel, s->se, &s->p, s->acc, n, c);
  if (s->targetOn)                                                     /* :351-358 */
    orc_bd_add_target(s->pos, s->target, s->targetRadius, s->targetSign, s->acc, n);
  orc_bd_update_vel(s->acc, s->timeStep, s->p.velocityScale, s->vel, n); /* :360-361 */
  if (s->cyclic)                                                       /* :363-372 */
    orc_bd_update_pos_periodic(s->vel, s->timeStep, s->pos, n, c);
  else
    orc_bd_update_pos_wall(s->vel, s->timeStep, s->pos, n, c);
  return 0;
}

/* physics/ocl/Boids.cpp:323-384 without the rendering steps. */
int orc_boids_update(orc_boids* s) { return boids_physics(s, NULL); }

/* The learned substep's fallback set (no reference counterpart; the rule the library applies
 * after its learned iterations).  A particle whose |density/rho0 - 1| exceeds tau is a seed;
 * every particle a seed's 27-cell walk (fluids.cl:101-123) meets at length(p_seed - p_j) < h
 * joins the set, and sorted element 0 (met by no walk) when its own walk meets a seed there.
 * Every other particle within h of the set is its halo.  act: n flags (2 = seed, 1 = other
 * member, 3 = halo, 0 = outside). */
void orc_fallback_set(const orc_f4* pred, const orc_u2* se, const float* density, float rho0,
                      float tau, uint8_t* act, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
  memset(act, 0, n);
  for (uint32_t i = 0; i < n; ++i) {
    if (!(fabsf(density[i] / rho0 - 1.0f) > tau))
      continue;
    act[i] = 2;
    const orc_f4 pos = pred[i];
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      if (len4(sub4(pos, pred[e])) < h && act[e] == 0)
        act[e] = 1;
    }
  }
  /* Sorted element 0 is in no cell list (grid.cl:95-108), so no seed's walk meets it: it
   * joins when a seed lies within h on its own walk. */
  if (n > 0 && act[0] == 0) {
    const orc_f4 pos = pred[0];
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      if (len4(sub4(pos, pred[e])) < h && fabsf(density[e] / rho0 - 1.0f) > tau)
        act[0] = 1;
    }
  }
  /* The halo: every other particle within h of a set member (element 0 by its own walk). */
  for (uint32_t i = 0; i < n; ++i) {
    if (act[i] != 1 && act[i] != 2)
      continue;
    const orc_f4 pos = pred[i];
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      if (len4(sub4(pos, pred[e])) < h && act[e] == 0)
        act[e] = 3;
    }
  }
  if (n > 0 && act[0] == 0) {
    const orc_f4 pos = pred[0];
    ORC_FLD_FOR_NEIGHBOURS(c, pos, e)
    {
      if (len4(sub4(pos, pred[e])) < h && (act[e] == 1 || act[e] == 2))
        act[0] = 3;
    }
  }
}

/* ------------------------------------------------------------------------------------
 * Rendering side of a frame
 * ---------------------------------------------------------------------------------- */

#define ORC_FAR_DIST 1000000.0f /* define.cl:10 */

/* utils.cl:35-38 */
void orc_reset_camera_dist(uint32_t* dist, uint32_t maxN)
{
  for (uint32_t i = 0; i < maxN; ++i)
    dist[i] = (uint32_t)ORC_FAR_DIST;
}

/* utils.cl:43-52 (cameraPos is a float3 buffer: its first element's xyz) */
void orc_fill_camera_dist(const orc_f4* pos, const float cam[3], uint32_t* dist, uint32_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const float dx = pos[i].x - cam[0], dy = pos[i].y - cam[1], dz = pos[i].z - cam[2];
    const float len = sqrtf((dx * dx + dy * dy) + dz * dz);
    dist[i] = (uint32_t)fmaxf(ORC_FAR_DIST - len * 100.0f, 0.0f);
  }
}

/* fluids.cl:458-478 */
void orc_fld_fill_color(const float* density, const orc_fluid_params* p, orc_f4* col, uint32_t n)
{
  const orc_f4 blue = { 0.0f, 0.1f, 1.0f, 0.5f };
  const orc_f4 lightBlue = { 0.7f, 0.7f, 1.0f, 0.5f };
  const orc_f4 darkBlue = { 0.0f, 0.0f, 0.8f, 0.5f };
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const float constraint = 1.0f - density[i] / p->restDensity;
    orc_f4 color = blue;
    if (constraint > 0.0f)
      color = add4(color, div4s(muls4(constraint, sub4(lightBlue, blue)), 0.35f));
    else if (constraint < 0.0f)
      color = add4(color, div4s(muls4(constraint, sub4(blue, darkBlue)), 0.35f));
    col[i] = color;
  }
}

/* boids.cl:38-41 */
void orc_bd_fill_color(orc_f4* col, uint32_t count)
{
  const orc_f4 v = { 1.0f, 0.02f, 0.02f, 0.5f };
  for (uint32_t i = 0; i < count; ++i)
    col[i] = v;
}

/* grid.cl:43-46 */
void orc_reset_detector(float* det, const orc_consts* c)
{
  memset(det, 0, sizeof(float) * 8 * (size_t)c->numCells);
}

/* grid.cl:51-60: the scalar 1.0f widens to all eight components of the float8 */
void orc_fill_detector(const orc_f4* pos, float* det, uint32_t n, const orc_consts* c)
{
  for (uint32_t i = 0; i < n; ++i) {
    const uint32_t cell = cell1d(pos[i], c);
    if (cell < c->numCells)
      for (int k = 0; k < 8; ++k)
        det[(size_t)cell * 8 + k] = 1.0f;
  }
}

/* Fluids.cpp:466-468 / Boids.cpp:379-381: the camera distances of the active particles and
 * the second RadixSort::sort, over maxN keys, carrying four float4 payloads. */
static int camera_sort(const orc_render* r, uint32_t* idx, uint32_t n, uint32_t maxN,
                       orc_f4* tmp, orc_f4* a, orc_f4* b, orc_f4* c, orc_f4* d)
{
  orc_fill_camera_dist(a, r->cameraPos, r->cameraDist, n);
  const int rc = orc_radix_sort(r->cameraDist, idx, maxN);
  if (rc)
    return rc;
  orc_permute_f4(idx, a, tmp, maxN);
  orc_permute_f4(idx, b, tmp, maxN);
  orc_permute_f4(idx, c, tmp, maxN);
  orc_permute_f4(idx, d, tmp, maxN);
  return 0;
}

/* physics/ocl/Fluids.cpp:400-471 */
int orc_fluids_frame(orc_fluids* s, orc_render* r, int paused)
{
  if (!paused) {
    const int rc = fluids_physics(s, r->col);
    if (rc)
      return rc;
    orc_reset_detector(r->detector, &s->c);                          /* :460 */
    orc_fill_detector(s->pos, r->detector, s->n, &s->c);             /* :461 */
    orc_fld_fill_color(s->density, &s->p, r->col, s->n);             /* :462 */
  }
  return camera_sort(r, s->idx, s->n, s->maxN, s->tmp, s->pos, r->col, s->vel, s->pred);
}

/* physics/ocl/Boids.cpp:323-384 */
int orc_boids_frame(orc_boids* s, orc_render* r, int paused)
{
  if (!paused) {
    const int rc = boids_physics(s, r->col);
    if (rc)
      return rc;
    orc_reset_detector(r->detector, &s->c);                          /* :375 */
    orc_fill_detector(s->pos, r->detector, s->n, &s->c);             /* :376 */
  }
  return camera_sort(r, s->idx, s->n, s->maxN, s->tmp, s->pos, r->col, s->vel, s->acc);
}

/* ------------------------------------------------------------------------------------
 * clouds.cl + Clouds.cpp (the Clouds PBF variant)
 * ---------------------------------------------------------------------------------- */

/* The exponential of clouds.cl (exp of a float in externalHeatSource and the wind of
 * cld_updatePosition; exp of a double in saturationVaporDensity, whose literals are double)
 * is vendor-defined in OpenCL.  Pinned here by a published method, Cody-Waite argument
 * reduction (x = k ln2 + r, ln2 split in a 31-bit head and a tail, |r| <= ln2/2) followed by
 * the degree-13 Taylor polynomial of e^r in Horner form with fused multiply-adds, and 2^k
 * applied as two exact power-of-two factors; the float version rounds the double result.
 * Both are within 1 ulp of the correctly rounded result (checked against libm). */
static const double ORC_EXP_C[14] = {
  0x1.0000000000000p+0,  0x1.0000000000000p+0,  0x1.0000000000000p-1,  0x1.5555555555555p-3,
  0x1.5555555555555p-5,  0x1.1111111111111p-7,  0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
  0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
  0x1.1eed8eff8d898p-29, 0x1.6124613a86d09p-33
};

static inline double pow2i(int k) /* 2^k for -1022 <= k <= 1023 */
{
  const uint64_t b = (uint64_t)(k + 1023) << 52;
  double d;
  memcpy(&d, &b, sizeof(d));
  return d;
}

double orc_exp64(double x)
{
  if (x != x)
    return x + x;
  if (x > 709.782712893384)
    return INFINITY;
  if (x < -745.1332191019412)
    return 0.0;
  const double k = rint(x * 0x1.71547652b82fep+0);
  double r = fma(-k, 0x1.62e42fec00000p-1, x);
  r = fma(-k, 0x1.d1cf79abc9e3bp-32, r);
  double p = ORC_EXP_C[13];
  for (int i = 12; i >= 0; --i)
    p = fma(p, r, ORC_EXP_C[i]);
  const int ki = (int)k;
  const int k1 = ki / 2;
  const int k2 = ki - k1;
  return (p * pow2i(k1)) * pow2i(k2);
}

float orc_exp32(float x) { return (float)orc_exp64((double)x); }

/* clouds.cl:67-70 */
float orc_cld_ext_heat(float altitude, const orc_consts* c)
{
  return clampf(orc_exp32(-(altitude + c->absWall[1]) / 3.0f), 0.0f, 1.0f);
}

/* clouds.cl:75-86 */
float orc_cld_env_temp(float altitude, const orc_consts* c)
{
  const float coeff = 3.5f;
  const float tempGround = 293.0f;
  return -coeff * (altitude + c->absWall[1]) + tempGround;
}

/* clouds.cl:91-96: the literals are double, so the expression is evaluated in double and
 * the result rounded to the float return type. */
float orc_cld_svd(float temperature)
{
  const double t = (double)temperature;
  return (float)(217 * orc_exp64(19.5 - 4303.4 / (t - 29.5)) / t);
}

/* clouds.cl:116-121 */
void orc_cld_init_temp(const orc_f4* pos, float* temp, uint32_t count, const orc_consts* c)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)count; ++i)
    temp[i] = orc_cld_env_temp(pos[i].y, c);
}

/* clouds.cl:127-133 */
void orc_cld_init_vapor(const orc_cloud_params* q, const float* temp, float* vapor, uint32_t count)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)count; ++i)
    vapor[i] = q->initVaporDensityCoeff * orc_cld_svd(temp[i]);
}

/* clouds.cl:159-167 */
void orc_cld_heat_ground(const float* tempIn, const orc_f4* pos, const orc_cloud_params* q,
                         float* temp, uint32_t n, const orc_consts* c)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    temp[i] = fminf(tempIn[i] + orc_cld_ext_heat(pos[i].y, c) * q->groundHeatCoeff * q->timeStep,
                    313.0f);
}

/* clouds.cl:173-183 */
void orc_cld_buoyancy(const float* temp, const orc_f4* pos, const float* cloudDens,
                      const orc_cloud_params* q, float* buoyancy, uint32_t n, const orc_consts* c)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const float envTemp = orc_cld_env_temp(pos[i].y, c);
    buoyancy[i] = q->buoyancyCoeff * (temp[i] - envTemp) / envTemp -
                  q->gravCoeff * c->gravity * cloudDens[i];
  }
}

/* clouds.cl:189-198 */
void orc_cld_adiabatic(const float* temp, const orc_f4* vel, const orc_cloud_params* q,
                       float* tempOut, uint32_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    tempOut[i] = fmaxf(temp[i] - q->adiabaticLapseRate * vel[i].y * q->timeStep, 223.0f);
}

/* clouds.cl:207-220 */
void orc_cld_generate(const float* temp, const float* vapor, const float* cloudDens,
                      const orc_cloud_params* q, float* gen, uint32_t n)
{
  (void)cloudDens;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    gen[i] = q->phaseTransitionRate * (vapor[i] - orc_cld_svd(temp[i]));
}

/* clouds.cl:226-237 */
void orc_cld_phase(const float* vaporIn, const float* cloudIn, const float* gen,
                   const orc_cloud_params* q, float* vapor, float* cloudDens, uint32_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    cloudDens[i] = fmaxf(cloudIn[i] + gen[i] * q->timeStep, 0.0f);
    vapor[i] = fmaxf(vaporIn[i] - gen[i] * q->timeStep, 0.0f);
  }
}

/* clouds.cl:243-251 */
void orc_cld_latent(const float* tempIn, const float* gen, const orc_cloud_params* q, float* temp,
                    uint32_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    temp[i] = tempIn[i] + fmaxf(q->latentHeatCoeff * gen[i] * q->timeStep, 0.0f);
}

/* clouds.cl:257-273 */
void orc_cld_predict(const orc_f4* pos, const orc_f4* vel, const float* buoyancy,
                     const orc_cloud_params* q, orc_f4* pred, orc_f4* totCorr, uint32_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 f = { 0.0f, buoyancy[i], 0.0f, 0.0f };
    const orc_f4 predVel = add4(vel[i], mul4s(f, q->timeStep));
    totCorr[i] = mul4s(predVel, q->timeStep);
    pred[i] = add4(pos[i], mul4s(predVel, q->timeStep));
  }
}

/* clouds.cl:279-300: periodic in x and z (x - 2 clamp(x)), wall in y */
void orc_cld_boundary(orc_f4* pred, uint32_t n, const orc_consts* c)
{
  const float wx = c->absWall[0], wy = c->absWall[1], wz = c->absWall[2];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 p = pred[i];
    const orc_f4 cl = { clampf(p.x, -wx, wx), clampf(p.y, -wy, wy), clampf(p.z, -wz, wz),
                        clampf(p.w, 0.0f, 0.0f) };
    if (fabsf(p.x) > wx)
      pred[i].x = p.x - 2 * cl.x;
    if (fabsf(p.y) > wy)
      pred[i].y = cl.y;
    if (fabsf(p.z) > wz)
      pred[i].z = p.z - 2 * cl.z;
  }
}

/* The cells of clouds.cl:322-353 in the loop's order (x outer, z inner): x and z wrap, and
 * a wrapped neighbour's position is shifted by absWall * (+-2); cells past the y walls are
 * skipped.  Returns the number of cells. */
static inline int cld_neighbour_cells(orc_f4 pos, const orc_consts* c, int32_t cells[27],
                                      orc_f4 shift[27])
{
  uint32_t u3[3];
  cell3d(pos, c, u3);
  const int32_t cx = (int32_t)u3[0], cy = (int32_t)u3[1], cz = (int32_t)u3[2];
  const int32_t rx = c->res[0], ry = c->res[1], rz = c->res[2];
  const orc_f4 absWall = { c->absWall[0], c->absWall[1], c->absWall[2], 0.0f };
  int k = 0;
  for (int iX = -1; iX <= 1; ++iX)
    for (int iY = -1; iY <= 1; ++iY)
      for (int iZ = -1; iZ <= 1; ++iZ) {
        orc_f4 sign = { 0.0f, 0.0f, 0.0f, 0.0f };
        const int32_t nx = (cx + iX + rx) % rx, ny = (cy + iY + ry) % ry, nz = (cz + iZ + rz) % rz;
        if (cx + iX >= rx)
          sign.x = 2.0f;
        else if (cx + iX < 0)
          sign.x = -2.0f;
        if (cy + iY >= ry || cy + iY < 0)
          continue;
        if (cz + iZ >= rz)
          sign.z = 2.0f;
        else if (cz + iZ < 0)
          sign.z = -2.0f;
        cells[k] = (nx * ry + ny) * rz + nz;
        const orc_f4 s = { absWall.x * sign.x, absWall.y * sign.y, absWall.z * sign.z,
                           absWall.w * sign.w };
        shift[k] = s;
        ++k;
      }
  return k;
}

/* vec = pos - p_e - absWall * sign, over every candidate e of the cells above */
#define ORC_CLD_FOR_NEIGHBOURS(c, pos, src, e, vec)                                        \
  int32_t cells_[27];                                                                    \
  orc_f4 shift_[27];                                                                     \
  const int ncells_ = cld_neighbour_cells(pos, c, cells_, shift_);                       \
  for (int k_ = 0; k_ < ncells_; ++k_)                                                   \
    for (uint32_t e = se[cells_[k_]].x; e <= se[cells_[k_]].y; ++e)                      \
      for (orc_f4 vec = sub4(sub4(pos, src[e]), shift_[k_]), *once_ = &vec; once_;       \
           once_ = NULL)

/* clouds.cl:305-359 */
void orc_cld_density(const orc_f4* pred, const orc_u2* se, const orc_fluid_params* p,
                     float* density, uint32_t n, const orc_consts* c)
{
  (void)p;
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    float fluidDensity = 0.0f;
    ORC_CLD_FOR_NEIGHBOURS(c, pos, pred, e, vec)
    {
      fluidDensity += poly6(vec, h, c);
    }
    density[i] = fluidDensity;
  }
}

/* clouds.cl:365-435 */
void orc_cld_lambda(const orc_f4* pred, const float* density, const orc_u2* se,
                    const orc_fluid_params* p, float* lambda, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const float densityC = density[i] / p->restDensity - 1.0f;
    orc_f4 sumGradCi = { 0.0f, 0.0f, 0.0f, 0.0f };
    float sumSqGradC = 0.0f;
    ORC_CLD_FOR_NEIGHBOURS(c, pos, pred, e, vec)
    {
      const orc_f4 grad = grad_spiky(vec, h, c);
      sumGradCi = add4(sumGradCi, grad);
      sumSqGradC += dot4(grad, grad);
    }
    sumSqGradC += dot4(sumGradCi, sumGradCi);
    sumSqGradC /= p->restDensity * p->restDensity;
    lambda[i] = -densityC / (sumSqGradC + p->relaxCFM);
  }
}

/* clouds.cl:441-502 */
void orc_cld_correction(const float* lambda, const orc_u2* se, const orc_f4* pred,
                        const orc_fluid_params* p, orc_f4* corrOut, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const float lambdaI = lambda[i];
    orc_f4 corr = { 0.0f, 0.0f, 0.0f, 0.0f };
    ORC_CLD_FOR_NEIGHBOURS(c, pos, pred, e, vec)
    {
      const float s = lambdaI + lambda[e] + art_pressure(vec, p, c);
      corr = add4(corr, muls4(s, grad_spiky(vec, h, c)));
    }
    corrOut[i] = div4s(corr, p->restDensity);
  }
}

/* clouds.cl:508-569 */
void orc_cld_laplacian_temp(const orc_f4* posP, const float* tempP, const orc_u2* se,
                            const orc_cloud_params* q, float* lap, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = posP[i];
    const float temp = tempP[i];
    float laplacian = 0.0f;
    ORC_CLD_FOR_NEIGHBOURS(c, pos, posP, e, vec)
    {
      laplacian += (temp - tempP[e]) * dot4(vec, grad_spiky(vec, h, c)) / (dot4(vec, vec) + ORC_FLOAT_EPS);
    }
    lap[i] = laplacian / q->restDensity;
  }
}

/* clouds.cl:575-648 */
void orc_cld_lambda_temp(const orc_f4* posP, const float* lap, const orc_u2* se,
                         const orc_cloud_params* q, float* lambdaT, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = posP[i];
    const float laplacian = lap[i];
    float sumGradCi = 0.0f, sumSqGradC = 0.0f;
    ORC_CLD_FOR_NEIGHBOURS(c, pos, posP, e, vec)
    {
      const orc_f4 grad = grad_spiky(vec, h, c);
      const float derivativeTemp = dot4(vec, grad) / (dot4(vec, vec) * q->restDensity + ORC_FLOAT_EPS);
      sumGradCi += derivativeTemp;
      sumSqGradC += derivativeTemp * derivativeTemp;
    }
    sumSqGradC += sumGradCi * sumGradCi;
    lambdaT[i] = -laplacian / (sumSqGradC + q->relaxCFM);
  }
}

/* clouds.cl:654-721 */
void orc_cld_correction_temp(const float* lambdaT, const orc_u2* se, const orc_f4* posP,
                             const orc_cloud_params* q, float* corrT, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = posP[i];
    const float lambdaI = lambdaT[i];
    float corr = 0.0f;
    ORC_CLD_FOR_NEIGHBOURS(c, pos, posP, e, vec)
    {
      const orc_f4 grad = grad_spiky(vec, h, c);
      const float derivativeTemp = dot4(vec, grad) / (dot4(vec, vec) * q->restDensity + ORC_FLOAT_EPS);
      corr += (lambdaI + lambdaT[e]) * derivativeTemp;
    }
    corrT[i] = corr;
  }
}

/* clouds.cl:727-784 */
void orc_cld_vorticity(const orc_f4* pred, const orc_u2* se, const orc_f4* vel,
                       const orc_fluid_params* p, orc_f4* vort, uint32_t n, const orc_consts* c)
{
  (void)p;
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const orc_f4 velocity = vel[i];
    orc_f4 w = { 0.0f, 0.0f, 0.0f, 0.0f };
    ORC_CLD_FOR_NEIGHBOURS(c, pos, pred, e, vec)
    {
      w = add4(w, cross4(sub4(vel[e], velocity), grad_spiky(vec, h, c)));
    }
    vort[i] = w;
  }
}

/* clouds.cl:790-850 */
void orc_cld_confinement(const orc_f4* pred, const orc_u2* se, const orc_f4* vort,
                         const orc_fluid_params* p, orc_f4* vel, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const orc_f4 vorticity = vort[i];
    orc_f4 nn = { 0.0f, 0.0f, 0.0f, 0.0f };
    ORC_CLD_FOR_NEIGHBOURS(c, pos, pred, e, vec)
    {
      nn = add4(nn, muls4(len4(vort[e]), grad_spiky(vec, h, c)));
    }
    const orc_f4 f = mul4s(muls4(p->vorticityConfCoeff, cross4(normalize4(nn), vorticity)), p->timeStep);
    vel[i] = add4(vel[i], f);
  }
}

/* clouds.cl:856-914 */
void orc_cld_xsph(const orc_f4* pred, const orc_u2* se, const orc_f4* velIn,
                  const orc_fluid_params* p, orc_f4* velOut, uint32_t n, const orc_consts* c)
{
  const float h = c->effectRadius;
#pragma omp parallel for schedule(dynamic, 256)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 pos = pred[i];
    const orc_f4 velocity = velIn[i];
    orc_f4 viscosity = { 0.0f, 0.0f, 0.0f, 0.0f };
    ORC_CLD_FOR_NEIGHBOURS(c, pos, pred, e, vec)
    {
      viscosity = add4(viscosity, mul4s(sub4(velIn[e], velocity), poly6(vec, h, c)));
    }
    velOut[i] = add4(velocity, muls4(p->xsphViscosityCoeff, viscosity));
  }
}

/* clouds.cl:931-935 */
void orc_cld_correct_temp(const float* corrT, float* temp, uint32_t n)
{
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i)
    temp[i] += 0.3f * corrT[i];
}

/* clouds.cl:942-952 */
void orc_cld_update_pos(const orc_f4* pred, const orc_cloud_params* q, orc_f4* pos, uint32_t n,
                        const orc_consts* c)
{
  const float wy = c->absWall[1];
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    pos[i] = pred[i];
    pos[i].x += (1 - orc_exp32(-(pred[i].y + wy) * 0.2f)) * q->windCoeff * q->timeStep *
                (float)(q->dim - 2);
    pos[i].z += (1 - orc_exp32(-(pred[i].y + wy) * 0.3f)) * 0.7f * q->windCoeff * q->timeStep;
  }
}

/* clouds.cl:958-966 */
void orc_cld_update_vel(const orc_f4* totCorr, const orc_fluid_params* p, orc_f4* vel, uint32_t n,
                        const orc_consts* c)
{
  const float m = c->maxVel;
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)n; ++i) {
    const orc_f4 v = div4s(totCorr[i], p->timeStep + ORC_FLOAT_EPS);
    const orc_f4 r = { clampf(v.x, -m, m), clampf(v.y, -m, m), clampf(v.z, -m, m), clampf(v.w, -m, m) };
    vel[i] = r;
  }
}

/* RadixSort.cpp:163-189 permutateFloat for a float payload */
void orc_permute_f1(const uint32_t* idx, float* buf, float* tmp, uint32_t maxN)
{
  memcpy(tmp, buf, sizeof(float) * maxN);
#pragma omp parallel for schedule(static)
  for (int64_t i = 0; i < (int64_t)maxN; ++i)
    buf[i] = tmp[idx[i]];
}

/* physics/ocl/Clouds.cpp:503-610 without the rendering steps (grid detector, colour, camera
 * sort). */
int orc_clouds_update(orc_clouds* s)
{
  const orc_consts* c = &s->c;
  const uint32_t n = s->n, maxN = s->maxN;
  memcpy(s->tempIn, s->temp, sizeof(float) * maxN);                      /* :516 */
  orc_cld_heat_ground(s->tempIn, s->pos, &s->q, s->temp, n, c);          /* :517 */
  orc_cld_buoyancy(s->temp, s->pos, s->cloud, &s->q, s->buoyancy, n, c); /* :519 */
  orc_cld_adiabatic(s->temp, s->vel, &s->q, s->tempIn, n);               /* :521 */
  orc_cld_generate(s->tempIn, s->vapor, s->cloud, &s->q, s->gen, n);     /* :523 */
  memcpy(s->vaporIn, s->vapor, sizeof(float) * maxN);                    /* :525-526 */
  memcpy(s->cloudIn, s->cloud, sizeof(float) * maxN);
  orc_cld_phase(s->vaporIn, s->cloudIn, s->gen, &s->q, s->vapor, s->cloud, n); /* :527 */
  orc_cld_latent(s->tempIn, s->gen, &s->q, s->temp, n);                  /* :529 */
  orc_cld_predict(s->pos, s->vel, s->buoyancy, &s->q, s->pred, s->totCorr, n); /* :534 */
  orc_cld_boundary(s->pred, n, c);                                       /* :537-538 */
  orc_fill_cell_ids(s->pred, s->key, n, c);                              /* :541 */
  const int rc = orc_radix_sort(s->key, s->idx, maxN);                   /* :543 */
  if (rc)
    return rc;
  orc_permute_f4(s->idx, s->pos, s->tmp, maxN);
  orc_permute_f4(s->idx, s->vel, s->tmp, maxN);
  orc_permute_f4(s->idx, s->pred, s->tmp, maxN);
  orc_permute_f4(s->idx, s->totCorr, s->tmp, maxN);
  orc_permute_f1(s->idx, s->temp, s->ftmp, maxN);
  orc_permute_f1(s->idx, s->buoyancy, s->ftmp, maxN);
  orc_permute_f1(s->idx, s->vapor, s->ftmp, maxN);
  orc_permute_f1(s->idx, s->cloud, s->ftmp, maxN);
  orc_permute_f1(s->idx, s->partID, s->ftmp, maxN);
  orc_reset_start_end(s->se, c);                                         /* :545-547 */
  orc_fill_start(s->key, s->se, n, c);
  orc_fill_end(s->key, s->se, n, maxN, c);
  orc_adjust_end(s->se, c);                                              /* :549-550 */
  if (s->q.isTempSmoothingEnabled) {                                     /* :554-567 */
    orc_cld_laplacian_temp(s->pos, s->temp, s->se, &s->q, s->lap, n, c);
    orc_cld_lambda_temp(s->pos, s->lap, s->se, &s->q, s->lambdaT, n, c);
    orc_cld_correction_temp(s->lambdaT, s->se, s->pos, &s->q, s->corrT, n, c);
    orc_cld_correct_temp(s->corrT, s->temp, n);
  }
  for (uint32_t it = 0; it < s->iters; ++it) {                           /* :570-587 */
    orc_cld_density(s->pred, s->se, &s->p, s->density, n, c);
    orc_cld_lambda(s->pred, s->density, s->se, &s->p, s->lambda, n, c);
    orc_cld_correction(s->lambda, s->se, s->pred, &s->p, s->corr, n, c);
    orc_fld_correct_pos(s->corr, s->pred, n);
    orc_fld_correct_pos(s->corr, s->totCorr, n);
    orc_cld_boundary(s->pred, n, c);
  }
  orc_cld_update_vel(s->totCorr, &s->p, s->vel, n, c);                   /* :590 */
  if (s->p.isVorticityConfEnabled) {                                     /* :592-602 */
    orc_cld_vorticity(s->pred, s->se, s->vel, &s->p, s->vort, n, c);
    orc_cld_confinement(s->pred, s->se, s->vort, &s->p, s->vel, n, c);
    memcpy(s->velIn, s->vel, sizeof(orc_f4) * maxN);
    orc_cld_xsph(s->pred, s->se, s->velIn, &s->p, s->vel, n, c);
  }
  orc_cld_update_pos(s->pred, &s->q, s->pos, n, c);                      /* :605 */
  return 0;
}
