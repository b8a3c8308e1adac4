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
 * joins the set.  act: n flags (2 = seed, 1 = other member, 0 = outside). */
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
