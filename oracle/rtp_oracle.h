/*
 * CPU oracle for the RealTimeParticles physics hot path — TEST INFRASTRUCTURE ONLY.
 *
 * This is a plain-C restatement of the reference's OpenCL kernels and host step orders
 * (physics/ocl/kernels/{define,sph,grid,radixSort,fluids,boids,clouds}.cl,
 *  physics/ocl/Fluids.cpp:97-126,400-471, physics/ocl/Boids.cpp:96-120,323-384,
 *  physics/ocl/Clouds.cpp:126-145,503-610, physics/utils/RadixSort.cpp:122-189).  Only the test suite, the smoke check and the
 * benchmark's CPU-baseline leg may load it; the HIP library never links or calls it.
 *
 * Each kernel is one C function that loops over the reference's work-items.  The
 * arithmetic follows the OpenCL expressions as written (left-to-right, no contraction:
 * build with -ffp-contract=off).  The OpenCL builtins whose precision the reference
 * leaves to the vendor are pinned as follows (all within the OpenCL 1.2 error bounds):
 *   dot(a,b)        = ((a.x*b.x + a.y*b.y) + a.z*b.z) + a.w*b.w
 *   length/fast_length(v)       = sqrtf(dot(v,v))           (IEEE sqrt)
 *   normalize/fast_normalize(v) = v / sqrtf(dot(v,v)), and 0 for the zero vector
 *   pow(x, n) for integral n    = ((x*x)*x)*...  (n-1 left-to-right products; 1 for n==0)
 *   clamp(x,lo,hi)  = fminf(fmaxf(x,lo),hi);  step(e,x) = x < e ? 0 : 1
 *   division        = IEEE division (no -cl-fast-relaxed-math reciprocal)
 *   denormals       = IEEE (not flushed)
 *   exp             = orc_exp64 / orc_exp32 below (Cody-Waite + Taylor; the Clouds model)
 */
#ifndef RTP_ORACLE_H
#define RTP_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct { float x, y, z, w; } orc_f4;
typedef struct { uint32_t x, y; } orc_u2;

/* physics/ocl/kernels/define.cl:13-25 — FluidParams, 44 bytes, passed by value. */
typedef struct {
  float restDensity;
  float relaxCFM;
  float timeStep;
  uint32_t dim;
  uint32_t isArtPressureEnabled;
  float artPressureRadius;
  float artPressureCoeff;
  uint32_t artPressureExp;
  uint32_t isVorticityConfEnabled;
  float vorticityConfCoeff;
  float xsphViscosityCoeff;
} orc_fluid_params;

/* physics/ocl/kernels/boids.cl:12-17 — BoidsRuleParams, 16 bytes. */
typedef struct {
  float velocityScale;
  float alignmentScale;
  float separationScale;
  float cohesionScale;
} orc_boids_params;

/* The compile-time "-D" constants of the reference programs (Fluids.cpp:104-119,
 * Boids.cpp:104-113) after their FloatToStr round trip (Utils.cpp:24-29). */
typedef struct {
  float effectRadius;     /* EFFECT_RADIUS          (fluids) */
  float effectRadiusSq;   /* EFFECT_RADIUS_SQUARED  (boids)  */
  float absWall[3];       /* ABS_WALL_X/Y/Z */
  int32_t res[3];         /* GRID_RES_X/Y/Z */
  float cellSize;         /* GRID_CELL_SIZE_XYZ */
  uint32_t numCells;      /* GRID_NUM_CELLS */
  uint32_t maxPartsInCell;/* NUM_MAX_PARTS_IN_CELL */
  float poly6Coeff;       /* POLY6_COEFF */
  float spikyCoeff;       /* SPIKY_COEFF */
  float maxVel;           /* MAX_VEL */
  float gravity;          /* ABS_GRAVITY_ACC_Y (define.cl:8 literal, 9.81f in the reference) */
} orc_consts;

/* FloatToStr(v, 10) followed by the OpenCL compiler's parse of the literal. */
float orc_float_roundtrip(float v);
/* Fluids.cpp:104-119 / Boids.cpp:104-113 with box given as floats (integers in the reference). */
void orc_make_consts(const float box[3], const int32_t res[3], uint32_t maxPartsInCell,
                     float gravity, orc_consts* out);

/* ---- pinned builtins exposed for the pinning tests ---- */
float orc_pown(float x, uint32_t n);
float orc_poly6(const orc_f4* vec, const orc_consts* c);
void orc_grad_spiky(const orc_f4* vec, const orc_consts* c, orc_f4* out);
float orc_art_pressure(const orc_f4* vec, const orc_fluid_params* p, const orc_consts* c);

/* ---- grid.cl ---- */
void orc_cell3d(const orc_f4* pos, const orc_consts* c, uint32_t out[3]);
void orc_reset_cell_ids(uint32_t* key, uint32_t maxN, const orc_consts* c);
void orc_fill_cell_ids(const orc_f4* pos, uint32_t* key, uint32_t n, const orc_consts* c);
void orc_reset_start_end(orc_u2* se, const orc_consts* c);
void orc_fill_start(const uint32_t* key, orc_u2* se, uint32_t n, const orc_consts* c);
void orc_fill_end(const uint32_t* key, orc_u2* se, uint32_t n, uint32_t keyLen, const orc_consts* c);
void orc_adjust_end(orc_u2* se, const orc_consts* c);

/* ---- radixSort.cl + RadixSort.cpp ---- */
/* Returns 0, or -1 when maxN is not a multiple of 512 (RadixSort.cpp:34-35 precondition). */
int orc_radix_sort(uint32_t* key, uint32_t* idx, uint32_t maxN);
void orc_permute_f4(const uint32_t* idx, orc_f4* buf, orc_f4* tmp, uint32_t maxN);

/* ---- fluids.cl (one function per kernel, work-items 0..n-1) ---- */
void orc_fld_predict(const orc_f4* pos, const orc_f4* vel, const orc_fluid_params* p,
                     orc_f4* pred, uint32_t n, const orc_consts* c);
void orc_fld_boundary(orc_f4* pred, uint32_t n, const orc_consts* c);
void orc_fld_density(const orc_f4* pred, const orc_u2* se, const orc_fluid_params* p,
                     float* density, uint32_t n, const orc_consts* c);
void orc_fld_lambda(const orc_f4* pred, const float* density, const orc_u2* se,
                    const orc_fluid_params* p, float* lambda, uint32_t n, const orc_consts* c);
void orc_fld_correction(const float* lambda, const orc_u2* se, const orc_f4* pred,
                        const orc_fluid_params* p, orc_f4* corr, uint32_t n, const orc_consts* c);
void orc_fld_correct_pos(const orc_f4* corr, orc_f4* pred, uint32_t n);
void orc_fld_update_vel(const orc_f4* pred, const orc_f4* pos, const orc_fluid_params* p,
                        orc_f4* vel, uint32_t n, const orc_consts* c);
void orc_fld_vorticity(const orc_f4* pred, const orc_u2* se, const orc_f4* vel,
                       const orc_fluid_params* p, orc_f4* vort, uint32_t n, const orc_consts* c);
void orc_fld_confinement(const orc_f4* pred, const orc_u2* se, const orc_f4* vort,
                         const orc_fluid_params* p, orc_f4* vel, uint32_t n, const orc_consts* c);
void orc_fld_xsph(const orc_f4* pred, const orc_u2* se, const orc_f4* velIn,
                  const orc_fluid_params* p, orc_f4* velOut, uint32_t n, const orc_consts* c);
void orc_fld_update_pos(const orc_f4* pred, orc_f4* pos, uint32_t n);

/* ---- learned warm-start inputs (no reference counterpart) ----
 * For each particle, from the density pass of the first Jacobi iteration on pred:
 * C, lambda[i] (that pass's own), neighbour count (r < h incl. self), sum |grad C|^2
 * (lambda's denominator without relaxCFM), the particle's own |grad_i C|^2, |vel[i]|, and
 * the two smallest distances to the walls of the boundary clamp box (negative distances
 * count as 0).  feat is n x 8. */
void orc_fld_features(const orc_f4* pred, const float* density, const orc_u2* se,
                      const float* lambda, const orc_f4* vel, const orc_fluid_params* p,
                      float* feat, uint32_t n, const orc_consts* c);
/* The learned substep's fallback set: seeds are the particles with |density/rho0 - 1| > tau;
 * every particle within h of a seed joins them (met by the seed's 27-cell walk, or, for
 * sorted element 0, by its own).  The halo: every other particle within h of the set.
 * act: n flags, 2 = seed, 1 = other member, 3 = halo, 0 = outside. */
void orc_fallback_set(const orc_f4* pred, const orc_u2* se, const float* density, float rho0,
                      float tau, uint8_t* act, uint32_t n, const orc_consts* c);

/* ---- boids.cl ---- */
void orc_bd_rules3d(const orc_f4* pos, const orc_f4* vel, const orc_u2* se,
                    const orc_boids_params* p, orc_f4* acc, uint32_t n, const orc_consts* c);
void orc_bd_update_vel(const orc_f4* acc, float dt, float maxVel, orc_f4* vel, uint32_t n);
void orc_bd_update_pos_wall(orc_f4* vel, float dt, orc_f4* pos, uint32_t n, const orc_consts* c);
/* boids.cl:137-221: the 9 cells of the YZ plane, at neighbour index
 * (RX/2 * RX + y) * RY + z with every coordinate range-checked against RX, as written (the
 * two agree with the grid's own indexing on cubic grids).  An index past the cell table,
 * which only non-cubic grids produce and where the reference reads out of bounds, is
 * skipped. */
void orc_bd_rules2d(const orc_f4* pos, const orc_f4* vel, const orc_u2* se,
                    const orc_boids_params* p, orc_f4* acc, uint32_t n, const orc_consts* c);
/* boids.cl:226-241: acc += sign * (target - pos) * clamp(1.3 / dist, 0, 1.4 * 0.5) within
 * the radius. */
void orc_bd_add_target(const orc_f4* pos, const float target[3], float radius, int32_t sign,
                       orc_f4* acc, uint32_t n);
/* boids.cl:289-315: a coordinate clamped at a wall re-enters at the opposite wall. */
void orc_bd_update_pos_periodic(const orc_f4* vel, float dt, orc_f4* pos, uint32_t n,
                                const orc_consts* c);

/* ---- host step drivers (caller-owned arrays, all maxN long except se = numCells) ---- */
typedef struct {
  orc_consts c;
  orc_fluid_params p;
  uint32_t maxN, n, iters;
  orc_f4 *pos, *vel, *pred, *corr, *vort, *velIn, *tmp;
  float *density, *lambda;
  uint32_t *key, *idx;
  orc_u2* se;
} orc_fluids;

typedef struct {
  orc_consts c;
  orc_boids_params p;
  float timeStep;
  uint32_t maxN, n;
  orc_f4 *pos, *vel, *acc, *tmp;
  uint32_t *key, *idx;
  orc_u2* se;
  /* variants of Boids::update (Boids.cpp:346-372) */
  int32_t dim;          /* 2: bd_applyBoidsRulesWithGrid2D, else the 3-D rules */
  int32_t cyclic;       /* Boundary::CyclicWall: bd_updatePosAndApplyPeriodicBC */
  int32_t targetOn;     /* isTargetActivated(): bd_addTargetRule */
  float target[3];      /* the target position of this frame (Target::pos) */
  float targetRadius;   /* TargetParams.targetRadiusEffect */
  int32_t targetSign;   /* TargetParams.targetSignEffect */
} orc_boids;

/* Fluids::update (Fluids.cpp:400-471) without GL, colour or the camera-distance sort. */
int orc_fluids_update(orc_fluids* s);
/* Boids::update (Boids.cpp:323-384) without GL: the rules of s->dim, the target rule when
 * s->targetOn, the cyclic or bouncing wall. */
int orc_boids_update(orc_boids* s);

/* ---- rendering side of a frame (utils.cl:35-77, grid.cl:43-60, fluids.cl:458-478,
 *      boids.cl:38-41) ---- */
/* The buffers a frame's rendering steps touch: u_cameraPos, p_col (maxN), p_cameraDist
 * (maxN) and c_partDetector (numCells x float8). */
typedef struct {
  float cameraPos[3];
  orc_f4* col;
  uint32_t* cameraDist;
  float* detector;
} orc_render;

/* utils.cl:35-38 over maxN entries. */
void orc_reset_camera_dist(uint32_t* dist, uint32_t maxN);
/* utils.cl:43-52: dist = (uint)max(FAR_DIST - length(pos.xyz - cam) * 100, 0) with
 * length(v) = sqrtf((v.x*v.x + v.y*v.y) + v.z*v.z) and max = fmaxf. */
void orc_fill_camera_dist(const orc_f4* pos, const float cam[3], uint32_t* dist, uint32_t n);
/* fluids.cl:458-478 fld_fillFluidColor. */
void orc_fld_fill_color(const float* density, const orc_fluid_params* p, orc_f4* col, uint32_t n);
/* boids.cl:38-41 bd_fillBoidsColor. */
void orc_bd_fill_color(orc_f4* col, uint32_t count);
/* grid.cl:43-46 and :51-60 (every component of a hit cell's float8 is 1). */
void orc_reset_detector(float* det, const orc_consts* c);
void orc_fill_detector(const orc_f4* pos, float* det, uint32_t n, const orc_consts* c);

/* One whole Fluids::update (Fluids.cpp:400-471): unless paused, the physics with p_col
 * permuted by the cell sort, then the grid detector and the colour; always the camera
 * distances and their sort over maxN, permuting p_pos, p_col, p_vel and p_predPos. */
int orc_fluids_frame(orc_fluids* s, orc_render* r, int paused);
/* One whole Boids::update (Boids.cpp:323-384) with the bouncing wall: unless paused, the
 * step with p_col permuted by the cell sort and the grid detector; always the camera
 * sort, permuting p_pos, p_col, p_vel and p_acc. */
int orc_boids_frame(orc_boids* s, orc_render* r, int paused);

/* ---- clouds.cl + Clouds.cpp (the Clouds PBF variant) ---- */
/* clouds.cl:18-32 CloudParams (52 bytes). */
typedef struct {
  uint32_t dim;
  float timeStep;
  float restDensity;
  float groundHeatCoeff;
  float buoyancyCoeff;
  float gravCoeff;
  float adiabaticLapseRate;
  float phaseTransitionRate;
  float latentHeatCoeff;
  uint32_t isTempSmoothingEnabled;
  float relaxCFM;
  float initVaporDensityCoeff;
  float windCoeff;
} orc_cloud_params;

/* exp pinned by Cody-Waite reduction + degree-13 Taylor polynomial (double; the float form
 * rounds the double result). */
double orc_exp64(double x);
float orc_exp32(float x);
float orc_cld_ext_heat(float altitude, const orc_consts* c);  /* clouds.cl:67-70 */
float orc_cld_env_temp(float altitude, const orc_consts* c);  /* clouds.cl:75-86 */
float orc_cld_svd(float temperature);                          /* clouds.cl:91-96 */
void orc_cld_init_temp(const orc_f4* pos, float* temp, uint32_t count, const orc_consts* c);
void orc_cld_init_vapor(const orc_cloud_params* q, const float* temp, float* vapor, uint32_t count);
void orc_cld_heat_ground(const float* tempIn, const orc_f4* pos, const orc_cloud_params* q,
                         float* temp, uint32_t n, const orc_consts* c);
void orc_cld_buoyancy(const float* temp, const orc_f4* pos, const float* cloudDens,
                      const orc_cloud_params* q, float* buoyancy, uint32_t n, const orc_consts* c);
void orc_cld_adiabatic(const float* temp, const orc_f4* vel, const orc_cloud_params* q,
                       float* tempOut, uint32_t n);
void orc_cld_generate(const float* temp, const float* vapor, const float* cloudDens,
                      const orc_cloud_params* q, float* gen, uint32_t n);
void orc_cld_phase(const float* vaporIn, const float* cloudIn, const float* gen,
                   const orc_cloud_params* q, float* vapor, float* cloudDens, uint32_t n);
void orc_cld_latent(const float* tempIn, const float* gen, const orc_cloud_params* q, float* temp,
                    uint32_t n);
void orc_cld_predict(const orc_f4* pos, const orc_f4* vel, const float* buoyancy,
                     const orc_cloud_params* q, orc_f4* pred, orc_f4* totCorr, uint32_t n);
void orc_cld_boundary(orc_f4* pred, uint32_t n, const orc_consts* c);
void orc_cld_density(const orc_f4* pred, const orc_u2* se, const orc_fluid_params* p,
                     float* density, uint32_t n, const orc_consts* c);
void orc_cld_lambda(const orc_f4* pred, const float* density, const orc_u2* se,
                    const orc_fluid_params* p, float* lambda, uint32_t n, const orc_consts* c);
void orc_cld_correction(const float* lambda, const orc_u2* se, const orc_f4* pred,
                        const orc_fluid_params* p, orc_f4* corr, uint32_t n, const orc_consts* c);
void orc_cld_laplacian_temp(const orc_f4* posP, const float* tempP, const orc_u2* se,
                            const orc_cloud_params* q, float* lap, uint32_t n, const orc_consts* c);
void orc_cld_lambda_temp(const orc_f4* posP, const float* lap, const orc_u2* se,
                         const orc_cloud_params* q, float* lambdaT, uint32_t n, const orc_consts* c);
void orc_cld_correction_temp(const float* lambdaT, const orc_u2* se, const orc_f4* posP,
                             const orc_cloud_params* q, float* corrT, uint32_t n, const orc_consts* c);
void orc_cld_vorticity(const orc_f4* pred, const orc_u2* se, const orc_f4* vel,
                       const orc_fluid_params* p, orc_f4* vort, uint32_t n, const orc_consts* c);
void orc_cld_confinement(const orc_f4* pred, const orc_u2* se, const orc_f4* vort,
                         const orc_fluid_params* p, orc_f4* vel, uint32_t n, const orc_consts* c);
void orc_cld_xsph(const orc_f4* pred, const orc_u2* se, const orc_f4* velIn,
                  const orc_fluid_params* p, orc_f4* velOut, uint32_t n, const orc_consts* c);
void orc_cld_correct_temp(const float* corrT, float* temp, uint32_t n);
void orc_cld_update_pos(const orc_f4* pred, const orc_cloud_params* q, orc_f4* pos, uint32_t n,
                        const orc_consts* c);
void orc_cld_update_vel(const orc_f4* totCorr, const orc_fluid_params* p, orc_f4* vel, uint32_t n,
                        const orc_consts* c);
void orc_permute_f1(const uint32_t* idx, float* buf, float* tmp, uint32_t maxN);

typedef struct {
  orc_consts c;
  orc_fluid_params p;
  orc_cloud_params q;
  uint32_t maxN, n, iters;
  orc_f4 *pos, *vel, *pred, *totCorr, *corr, *vort, *velIn, *tmp;
  float *temp, *tempIn, *buoyancy, *vapor, *vaporIn, *cloud, *cloudIn, *gen, *partID;
  float *density, *lambda, *lap, *lambdaT, *corrT, *ftmp;
  uint32_t *key, *idx;
  orc_u2* se;
} orc_clouds;

/* Clouds::update (Clouds.cpp:503-610) without GL, colour, grid detector or the camera sort. */
int orc_clouds_update(orc_clouds* s);

/* Number of OpenMP threads the oracle uses (1 when built without OpenMP). */
int orc_num_threads(void);
void orc_set_num_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
