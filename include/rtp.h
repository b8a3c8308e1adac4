/*
 * rtp.h — C ABI of the MI355X (gfx950) particle-physics backend.
 *
 * This library replaces the OpenCL path behind the reference's Physics::Model interface
 * (physics/Model.hpp:81-216) for the two models on the hot path: position based fluids
 * (physics/ocl/Fluids.{hpp,cpp}) and 3-D boids (physics/ocl/Boids.{hpp,cpp}).  Each entry
 * point below names the reference interface it stands in for.
 *
 * Conventions
 *   - every function returns 0 on success or a negative rtp_status; rtp_last_error()
 *     returns a text for the last failure on the calling thread (the reference's LOG_ERROR);
 *   - no C++ exception crosses this boundary;
 *   - the library owns device memory; host arrays are caller-owned and copied;
 *   - particle arrays are float4 (x, y, z, w) laid out as 4 consecutive floats; the w lane is
 *     carried as 0, as the reference's initialisation sets it (Fluids.cpp:383-387);
 *   - one HIP stream per handle; a handle is not thread-safe (callers serialise per handle,
 *     as the reference's single SDL thread does).
 */
#ifndef RTP_H
#define RTP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RTP_ABI_VERSION 1

typedef enum {
  RTP_OK = 0,
  RTP_ERR_INVALID = -1,  /* bad argument / size */
  RTP_ERR_HIP = -2,      /* HIP runtime failure */
  RTP_ERR_NO_DEVICE = -3,
  RTP_ERR_STATE = -4     /* call not valid in the handle's current state */
} rtp_status;

/* physics/ocl/Fluids.hpp:17-32 FluidKernelInputs == define.cl:13-25 FluidParams (44 bytes). */
typedef struct {
  float restDensity;
  float relaxCFM;
  float timeStep;
  uint32_t dim;
  uint32_t isArtPressureEnabled;
  float artPressureRadius;
  float artPressureCoeff;
  uint32_t artPressureExp;  /* at most 64 (a longer power is rejected) */
  uint32_t isVorticityConfEnabled;
  float vorticityConfCoeff;
  float xsphViscosityCoeff;
} rtp_fluid_params;

/* physics/ocl/Boids.hpp:14-20 BoidsRuleKernelInputs == boids.cl:12-17 BoidsRuleParams. */
typedef struct {
  float velocityScale;
  float alignmentScale;
  float separationScale;
  float cohesionScale;
} rtp_boids_params;

/* physics/Model.hpp:60-73 ModelParams, minus the GL VBO ids (the library owns the buffers,
 * see rtp_*_device_state).  boxSize is float (the reference's BoxSize3D is size_t,
 * utils/Geometry.hpp:42-47); integral values reproduce the reference exactly. */
typedef struct {
  uint32_t maxNbParticles;  /* ModelParams.maxNbParticles */
  uint32_t currNbParticles; /* ModelParams.currNbParticles */
  float boxSize[3];         /* ModelParams.boxSize */
  int32_t gridRes[3];       /* ModelParams.gridRes */
  uint32_t maxPartsInCell;  /* m_maxNbPartsInCell: 0 = model default (100 fluids
                               Fluids.cpp:79, 3000 boids Boids.cpp:78) */
  float gravity;            /* ABS_GRAVITY_ACC_Y (define.cl:8); 0 = reference's 9.81 */
  int32_t device;           /* HIP device ordinal */
} rtp_model_desc;

/* Per-kernel timing record (CL::Context::runKernel profiling, Context.cpp:692-705). */
typedef struct {
  char name[48];
  uint32_t launches;
  double totalMs;
} rtp_kernel_time;

/* Launch tuning (no reference counterpart).  The reference fixes its launch shapes (local
 * size 512 for the sort, RadixSort.cpp:23-60, and the runtime's choice elsewhere).  Here
 * the block size, launch flags, sort digit width and survivor-queue depth are chosen per
 * workload by timing it (rtp_fluids_autotune, rtp_batch_autotune). */
typedef struct {
  int32_t blockSize;  /* threads per block of the per-particle passes: 64, 128, 192 or 256 */
  int32_t flags;      /* RTP_LAUNCH_* bits (a batch takes RTP_LAUNCH_RESORT and
                         RTP_LAUNCH_NB_LIST) */
  int32_t sortBits;   /* largest radix digit width, 4..11 (0 = default 9) */
  int32_t queueDepth; /* neighbour survivors queued per lane in LDS, 8..64 (0 = default 16);
                         queueDepth x blockSize x 4 bytes must stay within 64 KiB, 24 KiB
                         with RTP_LAUNCH_LDS_TILES */
  float msPerStep;    /* measured mean substep time of this choice (autotune output) */
} rtp_tuning;
#define RTP_LAUNCH_XCD_MAP 1   /* consecutive particle blocks on one XCD (shared L2) */
#define RTP_LAUNCH_LDS_TILES 2 /* neighbour passes stage candidates in LDS wave tiles */
#define RTP_LAUNCH_RESORT 4    /* after every correction, order the next passes' threads by the
                                  cells their particles now sit in (a counting sort) */
#define RTP_LAUNCH_NB_LIST 8   /* the density pass of each Jacobi iteration records every
                                  particle's neighbour list (up to 256) and the correction pass
                                  replays it; likewise vorticity for confinement and XSPH
                                  (1 KiB of device memory per particle; ignored with
                                  RTP_LAUNCH_LDS_TILES) */
#define RTP_LAUNCH_PERSISTENT 32 /* boids: each update runs its steps in one cooperative launch
                                    (hash, sort, reorder and rules as phases between grid
                                    barriers; results unchanged) when the sort fits one
                                    workgroup's LDS and the rendering steps are off */
#define RTP_LAUNCH_BRICKS 16   /* with RTP_LAUNCH_RESORT: the thread order runs along x through
                                  pencils of 8 x 8 cells in (y, z), pencil after pencil, instead
                                  of in cell index order, so an XCD's consecutive waves keep
                                  their neighbours' data in its L2 on large grids */
#define RTP_LAUNCH_QUAD 64     /* fluids, clouds, boids: the neighbour passes (of the plain solver)
                                  give each particle four lanes (sixteen particles per 64-lane
                                  workgroup) that split its walk and add the terms in the same
                                  order; for scenes too small to fill the GPU one lane per
                                  particle (results unchanged; not with RTP_LAUNCH_LDS_TILES).
                                  The default shape of a model of at most 262,144 particles
                                  (maxNbParticles) has it; set_tuning / autotune replace it */

/* Which buffers rtp_fluids_read_field copies. */
typedef enum {
  RTP_FIELD_DENSITY = 0,   /* p_density  (float, last Jacobi iteration) */
  RTP_FIELD_LAMBDA = 1,    /* p_constFactor (float) */
  RTP_FIELD_PRED_POS = 2,  /* p_predPos (float4, final) */
  RTP_FIELD_VORTICITY = 3, /* p_vort (float4; w = |vorticity|) */
  RTP_FIELD_ACC = 4,       /* boids p_acc (float4) */
  /* with the rendering steps enabled (rtp_*_set_render): */
  RTP_FIELD_COLOR = 5,       /* p_col (float4) */
  RTP_FIELD_CAMERA_DIST = 6, /* p_cameraDist (uint32 words, sorted by the last update) */
  RTP_FIELD_DETECTOR = 7,    /* c_partDetector (float8 per cell; count = cells) */
  /* the Clouds model (Clouds.cpp:161-188): */
  RTP_FIELD_TEMPERATURE = 8,     /* p_temp */
  RTP_FIELD_VAPOR_DENSITY = 9,   /* p_vaporDens */
  RTP_FIELD_CLOUD_DENSITY = 10,  /* p_cloudDens */
  RTP_FIELD_BUOYANCY = 11,       /* p_buoyancy */
  RTP_FIELD_PARTICLE_ID = 12,    /* p_partID (float) */
  RTP_FIELD_TOT_CORR_POS = 13,   /* p_totCorrPos (float4) */
  RTP_FIELD_LAPLACIAN_TEMP = 14, /* p_laplacianTemp */
  RTP_FIELD_LAMBDA_TEMP = 15,    /* p_constFactorTemp */
  /* the learned warm-start (no reference counterpart), per sorted slot of the last update: */
  RTP_FIELD_RESIDUAL_DENSITY = 16, /* post-correction density of the last residual pass (float) */
  RTP_FIELD_FALLBACK_SET = 17      /* fallback set (uint32 words: 1 = set, 2 = halo, 0 = outside) */
} rtp_field;

/* physics/ocl/Clouds.hpp:14-48 CloudKernelInputs == clouds.cl:18-32 CloudParams (52 bytes). */
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
} rtp_cloud_params;

/* ---- the JSON parameter channel ----
 * The reference's models keep their parameters in a JSON tree (initFluidsJson,
 * physics/ocl/Fluids.cpp:52-74; initBoidsJson, physics/ocl/Boids.cpp:41-73) that its
 * parameter widget edits and updateModelWithInputJson (physics/ocl/OclModel.hpp:48-54)
 * applies through transferJsonInputsToModel (Fluids.cpp:218-251, Boids.cpp:177-210).  These
 * read and write that tree; every numeric entry is [value, min, max].  Text out is
 * NUL-terminated; *len gets the length without the NUL, and a buffer too small fails with
 * RTP_ERR_INVALID after writing what fits. */
/* TargetParams (Boids.hpp:22-26) plus the target's on/visible switches (Target.hpp). */
typedef struct {
  int32_t active;   /* "Enable##Target" */
  int32_t visible;  /* "Show" while active */
  float radius;     /* "Radius" */
  int32_t sign;     /* "Attract" as an integer: true 1, false 0 */
} rtp_target_params;
int rtp_fluid_params_from_json(const char* json, rtp_fluid_params* p, uint32_t* nbJacobiIters);
int rtp_fluid_params_to_json(const rtp_fluid_params* p, uint32_t nbJacobiIters, char* buf,
                             size_t cap, size_t* len);
/* initCloudsJson (physics/ocl/Clouds.cpp:63-98): its "Fluids" tree as for the fluids, and
 * the "Clouds" tree into CloudParams, whose rest density, time step and relaxation come from
 * the "Fluids" tree (Clouds.cpp:283-302); dim and initVaporDensityCoeff are not in the tree
 * and keep the values passed in. */
int rtp_cloud_params_from_json(const char* json, rtp_fluid_params* p, rtp_cloud_params* q,
                               uint32_t* nbJacobiIters);
int rtp_cloud_params_to_json(const rtp_fluid_params* p, const rtp_cloud_params* q,
                             uint32_t nbJacobiIters, char* buf, size_t cap, size_t* len);
int rtp_boids_params_from_json(const char* json, rtp_boids_params* p, rtp_target_params* t);
int rtp_boids_params_to_json(const rtp_boids_params* p, const rtp_target_params* t, char* buf,
                             size_t cap, size_t* len);
/* The two tree operations of Model::updateInputJson (physics/Model.hpp:164-172), for a C++
 * Model that keeps its tree as text: rtp_json_equal sets *equal to 1 when json::diff(a, b)
 * would be empty (equal values: objects as maps, numbers by value), else 0;
 * rtp_json_merge_patch writes target.merge_patch(patch) (RFC 7386: an object patch merges
 * key by key and a null removes the key; any other patch replaces the target). */
int rtp_json_equal(const char* a, const char* b, int* equal);
int rtp_json_merge_patch(const char* target, const char* patch, char* buf, size_t cap, size_t* len);

/* ---- library ---- */
int rtp_abi_version(void);
/* Text of the last error on this thread ("" if none). */
const char* rtp_last_error(void);
/* Number of visible HIP devices. */
int rtp_device_count(int* count);

/* ======================= Position based fluids ======================= */
typedef struct rtp_fluids rtp_fluids;

/* Fluids::Fluids(ModelParams) (physics/ocl/Fluids.cpp:76-92) via Physics::CreateModel
 * (physics/Model.cpp:10-24): program constants (Fluids.cpp:97-126) and buffers (:128-150). */
int rtp_fluids_create(const rtp_model_desc* desc, rtp_fluids** out);
/* Fluids::~Fluids / OclModel::~OclModel (physics/ocl/OclModel.hpp:23-26). */
int rtp_fluids_destroy(rtp_fluids* h);
/* Fluids::transferKernelInputsToGPU (Fluids.cpp:253-271). */
int rtp_fluids_set_params(rtp_fluids* h, const rtp_fluid_params* p);
int rtp_fluids_get_params(rtp_fluids* h, rtp_fluid_params* p);
/* m_nbJacobiIters, set from "Nb Jacobi Iterations" (Fluids.cpp:227); at most 1000. */
int rtp_fluids_set_nb_jacobi_iters(rtp_fluids* h, uint32_t iters);
int rtp_fluids_get_nb_jacobi_iters(rtp_fluids* h, uint32_t* iters);
/* Model::setNbParticles (physics/Model.hpp:106). */
int rtp_fluids_set_nb_particles(rtp_fluids* h, uint32_t n);
int rtp_fluids_get_nb_particles(rtp_fluids* h, uint32_t* n);
/* Fluids::initFluidsParticles (Fluids.cpp:383-395) + resetCellIDs (:214): count float4
 * positions/velocities; entries count..maxN-1 become (inf,inf,inf,0) and 0. */
int rtp_fluids_load_state(rtp_fluids* h, const float* pos4, const float* vel4, uint32_t count);
/* Read back p_pos / p_vel (the GL VBO contents in the reference); either may be NULL. */
int rtp_fluids_read_state(rtp_fluids* h, float* pos4, float* vel4, uint32_t count);
/* Fluids::update (Fluids.cpp:400-471) repeated nsteps times, enqueued on the handle's
 * stream; returns once enqueued (rtp_fluids_sync waits, like Context::finishTasks). */
int rtp_fluids_update(rtp_fluids* h, uint32_t nsteps);
/* Model::pause (physics/Model.hpp:127): while paused, update() does nothing. */
int rtp_fluids_pause(rtp_fluids* h, int pause);
int rtp_fluids_sync(rtp_fluids* h);
/* p_cellID (maxN), RadixSortIndices (maxN) and c_startEndPartID (2*NC) after the last
 * update; any pointer may be NULL. */
int rtp_fluids_read_grid(rtp_fluids* h, uint32_t* cellIDs, uint32_t* sortIdx, uint32_t* startEnd);
/* Copy one per-particle field (see rtp_field) for the first count particles. */
int rtp_fluids_read_field(rtp_fluids* h, int field, float* out, uint32_t count);
/* Density error of the last update over active particles, from p_density:
 * out[0] = max |rho/rho0 - 1|, out[1] = mean |rho/rho0 - 1|, out[2] = max(rho/rho0 - 1, 0). */
int rtp_fluids_density_error(rtp_fluids* h, float out[3]);
/* Device pointers of the current float4 position/velocity buffers (the role of the
 * reference's GL VBOs, physics/ocl/Fluids.cpp:132-134); valid until the next update. */
int rtp_fluids_device_state(rtp_fluids* h, void** pos4, void** vel4);
/* The handle's HIP stream (hipStream_t). */
int rtp_fluids_stream(rtp_fluids* h, void** stream);
/* Model::enableProfiling / isProfilingEnabled (physics/Model.hpp:150-151): per-kernel
 * HIP-event timing on the handle's stream. */
int rtp_fluids_enable_profiling(rtp_fluids* h, int enable);
int rtp_fluids_profile(rtp_fluids* h, rtp_kernel_time* out, int capacity, int* count);
int rtp_fluids_reset_profile(rtp_fluids* h);
/* Device-side checkpoint of the model state (p_pos, p_vel, p_predPos, p_cellID and the
 * particle count; no reference counterpart — the reference re-creates its initial lattice in
 * reset(), Fluids.cpp:196-216).  restore enqueues device copies on the handle's stream, so
 * restore + update can be timed back to back without host transfers. */
int rtp_fluids_save_checkpoint(rtp_fluids* h);
int rtp_fluids_restore_checkpoint(rtp_fluids* h);
/* Replay pairs of substeps from a captured hipGraph (no reference counterpart; the
 * reference enqueues every kernel each frame).  Default on. */
int rtp_fluids_use_graph(rtp_fluids* h, int enable);
/* The rendering steps at the end of every update, Fluids.cpp:458-468: the grid detector
 * (fillGridDetector, grid.cl:51-60), the colour from p_density (fld_fillFluidColor,
 * fluids.cl:458-478) and, paused or not, the camera distances (fillCameraDist,
 * utils.cl:43-52) with the second RadixSort::sort that permutes p_pos, p_col, p_vel and
 * p_predPos.  The reference always runs them; here they are off by default, because they
 * only serve drawing and they change the next update's within-cell order.  Enabling resets
 * p_col, p_cameraDist and the detector as a model reset does (Fluids.cpp:209-215), and so
 * does every rtp_fluids_load_state while enabled.  Needs 48 bytes per particle of
 * maxNbParticles and 32 per cell, allocated the first time. */
int rtp_fluids_set_render(rtp_fluids* h, int enable);
/* u_cameraPos (Fluids.cpp:132): the point the following updates sort by distance to.
 * The origin until set. */
int rtp_fluids_set_camera(rtp_fluids* h, const float pos[3]);
/* Device addresses of the current p_col (float4 x maxN), p_cameraDist (uint32 x maxN) and
 * c_partDetector (float8 per cell), for an app that copies them into its GL buffers (the
 * role of cl::BufferGL, Context.cpp:517-541).  p_col and p_cameraDist each live in one of
 * two buffers, which an update may switch: ask again after every update.  Any pointer
 * argument may be NULL. */
int rtp_fluids_device_render(rtp_fluids* h, void** col4, void** cameraDist, void** detector);
/* Fluids::updateModelWithInputJson (OclModel.hpp:48-54, Fluids.cpp:218-251): the fluid
 * parameters and the Jacobi iteration count from the "Fluids" tree; FluidParams.dim keeps
 * the model's dimension. */
int rtp_fluids_update_with_json(rtp_fluids* h, const char* json);
/* The model's parameters as that tree (getInputJson's role). */
int rtp_fluids_params_json(rtp_fluids* h, char* buf, size_t cap, size_t* len);
/* Launch tuning (no reference counterpart), see rtp_tuning (msPerStep is ignored).  Results
 * do not depend on any of it. */
int rtp_fluids_set_tuning(rtp_fluids* h, const rtp_tuning* t);

/* ========================= Learned warm-start ========================= */
/* No reference counterpart: an 8 -> 64 -> 64 -> 1 MLP (GELU, BF16 weights, FP32
 * accumulation on the matrix cores) that predicts a particle's initial lambda.  Inputs per
 * particle, computed in the first density pass of a substep: C = rho/rho0 - 1, that
 * pass's lambda, neighbour count (r < h, self included), sum |grad C|^2, the particle's own
 * |grad_i C|^2, speed, and the distances to the two nearest walls of the boundary clamp
 * box.  They depend on the substep's input state only. */
typedef struct rtp_mlp rtp_mlp;

/* Network with deterministic N(0, 1/fan_in) weights from seed, zero biases. */
int rtp_mlp_create(int device, uint32_t seed, rtp_mlp** out);
int rtp_mlp_destroy(rtp_mlp* h);
/* Length of the fp32 parameter vector: W1[8][64] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3. */
int rtp_mlp_param_count(void);
int rtp_mlp_set_params(rtp_mlp* h, const float* params, uint32_t count);
int rtp_mlp_get_params(rtp_mlp* h, float* params, uint32_t count);
/* Inputs are standardised as (x - mu) / sigma; the output is multiplied by outScale. */
int rtp_mlp_set_normalization(rtp_mlp* h, const float* mu, const float* sigma, float outScale);
/* out[i] = network(feat[8 i .. 8 i + 7]) for n rows (host arrays). */
int rtp_mlp_forward(rtp_mlp* h, const float* feat, uint32_t n, float* out);
/* Gradient of the mean squared error on target / outScale over all n rows (parameter
 * layout above) and that loss; the parameters are not changed. */
int rtp_mlp_gradients(rtp_mlp* h, const float* feat, const float* target, uint32_t n, float* grad,
                      float* loss);
/* Adam (lr, betas 0.9/0.999) over shuffled mini-batches on the GPU; epochLoss[e] (optional)
 * receives each epoch's mean squared error. */
int rtp_mlp_train(rtp_mlp* h, const float* feat, const float* target, uint32_t n, uint32_t epochs,
                  uint32_t batch, float lr, uint32_t seed, float* epochLoss);
/* Use the network for the first of learnedIters Jacobi iterations of every substep (NULL
 * switches back to the plain solver).  With fallbackIters > 0, a density pass after those
 * iterations measures every particle's |rho/rho0 - 1|; the particles above tau, and every
 * particle within h of one of them, run fallbackIters plain iterations (the others keep
 * their positions and densities, and their lambda reads 0).  No particle above tau: the
 * substep equals one without fallback, bit for bit. */
int rtp_fluids_set_warm_start(rtp_fluids* h, rtp_mlp* mlp, uint32_t learnedIters,
                              uint32_t fallbackIters, float tau);
/* Particles above tau (seeds), in the fallback set (seeds and their neighbours within h) and
 * in its halo (the other particles within h of the set; may be NULL) of the last learned
 * substep with fallback; waits for the handle's stream. */
int rtp_fluids_fallback_count(rtp_fluids* h, uint32_t* seeds, uint32_t* particles, uint32_t* halo);
/* Run a post-correction density pass in every substep, plain or learned, after any
 * fallback iterations (diagnostics: the error the substep leaves). */
int rtp_fluids_set_residual_pass(rtp_fluids* h, int enable);
/* max, mean and max-compression |rho/rho0 - 1| of the last post-correction density pass,
 * the gate's or the diagnostic one (the same bits for the same state on every run). */
int rtp_fluids_residual_error(rtp_fluids* h, float out[3]);
/* One substep with teacherIters plain iterations that records, for every active particle
 * in the substep's sorted order, the network inputs (n x 8) and the sum of the lambdas of
 * all iterations (the training target).  The model state advances by that substep. */
int rtp_fluids_record_teacher(rtp_fluids* h, uint32_t teacherIters, float* feat, float* target);

/* ======================= Launch tuning (no reference counterpart) ================== */
/* Times fixed-input substeps (the current state restored before each) for every candidate
 * launch shape, keeps the fastest and restores the state.  trialSteps substeps per
 * candidate (after one untimed); best (optional) receives the choice. */
int rtp_fluids_autotune(rtp_fluids* h, uint32_t trialSteps, rtp_tuning* best);

/* ================= Batched fluids (no reference counterpart) =================
 * Many independent PBF scenes of uneven sizes advanced together: every pass of a substep
 * is one launch over all scenes (per-scene particle offsets, key ranges and constants), the
 * grid one sort over all of them.  Each scene evolves exactly as an rtp_fluids model of its
 * own whose particles are all active, with the batch's iteration count. */
typedef struct rtp_batch rtp_batch;

typedef struct {
  uint32_t nbParticles;     /* active particles of the scene (all of its slots) */
  float boxSize[3];         /* as rtp_model_desc */
  int32_t gridRes[3];
  uint32_t maxPartsInCell;  /* 0 = 100 */
  float gravity;            /* 0 = 9.81 */
} rtp_batch_scene;

int rtp_batch_create(int device, const rtp_batch_scene* scenes, uint32_t count, rtp_batch** out);
int rtp_batch_destroy(rtp_batch* h);
/* FluidKernelInputs of one scene, or of every scene with scene == UINT32_MAX.  The
 * vorticity switch (which decides the passes of a substep) must agree across scenes. */
int rtp_batch_set_params(rtp_batch* h, uint32_t scene, const rtp_fluid_params* p);
int rtp_batch_set_nb_jacobi_iters(rtp_batch* h, uint32_t iters);
/* Positions and velocities of one scene, nbParticles float4 each (w ignored). */
int rtp_batch_load_state(rtp_batch* h, uint32_t scene, const float* pos4, const float* vel4);
int rtp_batch_read_state(rtp_batch* h, uint32_t scene, float* pos4, float* vel4);
/* One scene's field (rtp_field, fluids fields), nbParticles entries. */
int rtp_batch_read_field(rtp_batch* h, uint32_t scene, int field, float* out);
/* One scene's grid in scene-local terms, as rtp_fluids_read_grid of a model of its own:
 * cell IDs and sort indices (nbParticles each), start/end (2 x its cell count). */
int rtp_batch_read_grid(rtp_batch* h, uint32_t scene, uint32_t* keys, uint32_t* idx,
                        uint32_t* startEnd);
int rtp_batch_update(rtp_batch* h, uint32_t nsteps);
int rtp_batch_sync(rtp_batch* h);
int rtp_batch_save_checkpoint(rtp_batch* h);
int rtp_batch_restore_checkpoint(rtp_batch* h);
int rtp_batch_use_graph(rtp_batch* h, int enable);
/* As rtp_fluids_set_tuning; of the flags RTP_LAUNCH_RESORT and RTP_LAUNCH_NB_LIST apply to a
 * batch (the lists take 1 KiB per particle plus 255 KiB per scene). */
int rtp_batch_set_tuning(rtp_batch* h, const rtp_tuning* t);
int rtp_batch_autotune(rtp_batch* h, uint32_t trialSteps, rtp_tuning* best);
int rtp_batch_enable_profiling(rtp_batch* h, int enable);
int rtp_batch_profile(rtp_batch* h, rtp_kernel_time* out, int capacity, int* count);
/* Total particles and cell-table entries of the batch. */
int rtp_batch_sizes(rtp_batch* h, uint32_t* particles, uint32_t* cells);

/* ============================== Boids ================================ */
typedef struct rtp_boids rtp_boids;

/* Boids::Boids(ModelParams) (physics/ocl/Boids.cpp:75-91), constants Boids.cpp:96-120. */
int rtp_boids_create(const rtp_model_desc* desc, rtp_boids** out);
int rtp_boids_destroy(rtp_boids* h);
/* Boids::transferKernelInputsToGPU (Boids.cpp:212-226). */
int rtp_boids_set_params(rtp_boids* h, const rtp_boids_params* p);
/* The update time step; the reference hard-codes 0.1 (Boids.cpp:334). */
int rtp_boids_set_time_step(rtp_boids* h, float dt);
int rtp_boids_set_nb_particles(rtp_boids* h, uint32_t n);
/* Boids::initBoidsParticles (Boids.cpp:310-318) + resetCellIDs (:271). */
int rtp_boids_load_state(rtp_boids* h, const float* pos4, const float* vel4, uint32_t count);
int rtp_boids_read_state(rtp_boids* h, float* pos4, float* vel4, uint32_t count);
/* Boids::update (Boids.cpp:323-384), 3-D rules and bouncing wall, nsteps times. */
int rtp_boids_update(rtp_boids* h, uint32_t nsteps);
int rtp_boids_pause(rtp_boids* h, int pause);
int rtp_boids_sync(rtp_boids* h);
int rtp_boids_read_grid(rtp_boids* h, uint32_t* cellIDs, uint32_t* sortIdx, uint32_t* startEnd);
int rtp_boids_read_field(rtp_boids* h, int field, float* out, uint32_t count);
int rtp_boids_device_state(rtp_boids* h, void** pos4, void** vel4);
/* Model::setDimension (physics/Model.hpp:109-113) for boids: 2 selects the rules of
 * bd_applyBoidsRulesWithGrid2D (boids.cl:137-221: the nine YZ cells of the x layer
 * GRID_RES_X / 2, Boids.cpp:346-349), 3 (the default) the 3-D rules. */
int rtp_boids_set_dimension(rtp_boids* h, int dim);
/* Launch shape of the boids step (no reference counterpart): the block size and, of the
 * flags, RTP_LAUNCH_XCD_MAP, RTP_LAUNCH_PERSISTENT and RTP_LAUNCH_QUAD (the step launch with
 * a quad of four lanes per boid, sixteen boids per 64-lane workgroup, which ignores the block
 * size).  Results do not depend on it.  Default 256 threads, with RTP_LAUNCH_QUAD for at most
 * 262,144 boids (maxNbParticles). */
int rtp_boids_set_tuning(rtp_boids* h, const rtp_tuning* t);
/* Model::setBoundary (physics/Model.hpp:116-119, Boundary :38-44): the wall of bd_updatePosAndApplyWallBC
 * (bouncing, the default) or of bd_updatePosAndApplyPeriodicBC (cyclic, boids.cl:289-315),
 * Boids.cpp:363-372. */
typedef enum { RTP_BOUNDARY_BOUNCING = 0, RTP_BOUNDARY_CYCLIC = 1 } rtp_boundary;
int rtp_boids_set_boundary(rtp_boids* h, int boundary);
/* The target rule, bd_addTargetRule (boids.cl:226-241, Boids.cpp:351-358): while active,
 * every step adds sign * (pos - p) * clamp(1.3 / |pos - p|, 0, 0.7) to the acceleration of
 * each boid p within radius of pos (TargetParams, Boids.hpp:22-26).  The reference moves
 * the target on the host every frame (Target::updatePos, physics/utils/Target.cpp:27-50);
 * here the caller passes the position of the frame.  A non-NULL pos is stored even while
 * inactive, so a target switched on later (here or by a JSON tree) starts there; radius
 * and sign are taken only with active set. */
int rtp_boids_set_target(rtp_boids* h, int active, const float pos[3], float radius, int sign);
/* Boids::targetPos / isTargetActivated / isTargetVisible (physics/ocl/Boids.hpp:38-42): the
 * target's switch, "Show" state, radius and sign, and its last position (pos may be NULL). */
int rtp_boids_get_target(rtp_boids* h, rtp_target_params* t, float pos[3]);
/* Boids.cpp:375-381: the grid detector after every step and, paused or not, the
 * camera-distance sort that permutes p_pos, p_vel and p_acc (p_col holds bd_fillBoidsColor's
 * single colour, boids.cl:38-41).  As rtp_fluids_set_render: off by default, enabling resets
 * p_col, p_cameraDist and the detector (Boids.cpp:265-272). */
int rtp_boids_set_render(rtp_boids* h, int enable);
/* u_cameraPos (Boids.cpp:126). */
int rtp_boids_set_camera(rtp_boids* h, const float pos[3]);
int rtp_boids_device_render(rtp_boids* h, void** col4, void** cameraDist, void** detector);
/* Boids::updateModelWithInputJson (Boids.cpp:177-210): the rules and the target's switch,
 * radius and sign from the "Boids" tree; the target keeps its position. */
int rtp_boids_update_with_json(rtp_boids* h, const char* json);
int rtp_boids_params_json(rtp_boids* h, char* buf, size_t cap, size_t* len);
int rtp_boids_enable_profiling(rtp_boids* h, int enable);
int rtp_boids_profile(rtp_boids* h, rtp_kernel_time* out, int capacity, int* count);

/* ================================ Clouds ================================ */
/* physics/ocl/Clouds.{hpp,cpp}: position based fluids (FluidParams) coupled to a
 * temperature, vapour and cloud density per particle (CloudParams), with periodic x/z walls,
 * a temperature-smoothing constraint and a wind; on the fluids model's grid and sort. */
typedef struct rtp_clouds rtp_clouds;
/* Clouds::Clouds (Clouds.cpp:100-119) + reset; the parameters start at FluidKernelInputs'
 * and CloudKernelInputs' defaults (Fluids.hpp:17-32, Clouds.hpp:14-48), 2 Jacobi iterations. */
int rtp_clouds_create(const rtp_model_desc* desc, rtp_clouds** out);
int rtp_clouds_destroy(rtp_clouds* h);
/* transferJsonInputsToModel's typed result (Clouds.cpp:253-303); either pointer may be NULL. */
int rtp_clouds_set_params(rtp_clouds* h, const rtp_fluid_params* p, const rtp_cloud_params* q);
int rtp_clouds_get_params(rtp_clouds* h, rtp_fluid_params* p, rtp_cloud_params* q);
/* Launch choice of the walk passes (no reference counterpart): t->flags RTP_LAUNCH_QUAD gives
 * each particle a quad of four lanes (the default when maxNbParticles <= 262,144), 0 one lane;
 * the other fields are ignored.  Results are the same either way. */
int rtp_clouds_set_tuning(rtp_clouds* h, const rtp_tuning* t);
/* m_nbJacobiIters (Clouds.hpp:71) */
int rtp_clouds_set_nb_jacobi_iters(rtp_clouds* h, uint32_t iters);
int rtp_clouds_get_nb_jacobi_iters(rtp_clouds* h, uint32_t* iters);
/* Model::setNbParticles (Model.hpp:106) */
int rtp_clouds_set_nb_particles(rtp_clouds* h, uint32_t n);
/* Clouds::initCloudsParticles (Clouds.cpp:401-500) with the caller's positions and
 * velocities: the tail at +inf, cloud density 0, particle ids, the temperature and vapour
 * density of every slot from its altitude (cld_initTemperature, cld_initVaporDensity), and
 * resetCellIDs.  count becomes the particle count. */
int rtp_clouds_load_state(rtp_clouds* h, const float* pos4, const float* vel4, uint32_t count);
int rtp_clouds_read_state(rtp_clouds* h, float* pos4, float* vel4, uint32_t count);
/* Clouds::update (Clouds.cpp:503-610) without the rendering steps, nsteps times. */
int rtp_clouds_update(rtp_clouds* h, uint32_t nsteps);
int rtp_clouds_pause(rtp_clouds* h, int pause);
int rtp_clouds_sync(rtp_clouds* h);
int rtp_clouds_read_grid(rtp_clouds* h, uint32_t* cellIDs, uint32_t* sortIdx, uint32_t* startEnd);
/* RTP_FIELD_DENSITY, _LAMBDA, _PRED_POS, _VORTICITY and the Clouds fields 8..15. */
int rtp_clouds_read_field(rtp_clouds* h, int field, float* out, uint32_t count);
/* Device addresses of p_pos and p_vel (float4 x maxN), valid until the next update. */
int rtp_clouds_device_state(rtp_clouds* h, void** pos4, void** vel4);
/* updateModelWithInputJson / inputJson of the Clouds model (OclModel.hpp:48-54,
 * Clouds.cpp:253-303). */
int rtp_clouds_update_with_json(rtp_clouds* h, const char* json);
int rtp_clouds_params_json(rtp_clouds* h, char* buf, size_t cap, size_t* len);
int rtp_clouds_enable_profiling(rtp_clouds* h, int enable);
int rtp_clouds_profile(rtp_clouds* h, rtp_kernel_time* out, int capacity, int* count);

/* ========================= Grid building blocks ======================== */
/* RadixSort::sort on one key buffer (physics/utils/RadixSort.cpp:122-161): stable
 * ascending sort of n 32-bit keys with maxKey >= every key; permOut[i] is the original
 * position of the key now at i.  digitBits: largest digit width of the multi-launch sort,
 * 4..11 (0 = 9; sorts of at most 8,192 keys run in one workgroup with 4-bit digits).  Host
 * arrays in and out (for tests and tools). */
int rtp_sort_pairs(int device, const uint32_t* keysIn, uint32_t n, uint32_t maxKey,
                   int32_t digitBits, uint32_t* keysOut, uint32_t* permOut);
/* The grid part of Fluids::update for one batch of positions (Fluids.cpp:415-424):
 * fillCellIDs over n, sort of all maxN keys (tail keys 2*NC+i as after resetCellIDs), cell
 * start/end with the cap.  keys/idx are maxN long, startEnd 2*NC. */
int rtp_grid_build(const rtp_model_desc* desc, const float* pos4, uint32_t* keys, uint32_t* idx,
                   uint32_t* startEnd);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc): the grid
 * build of rtp_grid_build, with the same results, on DEVICE buffers the caller owns -- pos4
 * (currNbParticles float4), keys and idx (maxNbParticles u32 each) and startEnd (2*NC u32) --
 * queued on `stream` (NULL: the null stream) without waiting for it, and with no allocation.
 * The sort's partner buffers and digit counts live in a device array the library holds on
 * each GPU it runs on until the process ends (about 3 x 4 B x RTP_NOALLOC_MAX_PARTICLES);
 * maxNbParticles above RTP_NOALLOC_MAX_PARTICLES is RTP_ERR_INVALID.  Every call on one GPU
 * uses that GPU's one array, so calls that overlap on one GPU (two streams or two threads)
 * race; calls in one stream are ordered and safe. */
#define RTP_NOALLOC_MAX_PARTICLES 2097152u
int rtp_grid_build_noalloc(const rtp_model_desc* desc, const float* pos4, uint32_t* keys,
                           uint32_t* idx, uint32_t* startEnd, void* stream);

/* ============================ Self-checks ============================== */
/* No counterpart in the reference.  The neighbour passes evaluate correctly rounded square
 * roots and divisions in short forms valid over the operand ranges they meet; these run
 * those forms on the device against the compiler's general correctly rounded sequences and
 * count the inputs whose results differ in any bit (NaN results compare equal).
 * rtp_selfcheck_sqrt: every bit pattern in [lo, hi) (pass x >= 0 patterns).
 * rtp_selfcheck_division: divisors 2^dExp * (1 + s/2^23) for s in [dSigLo, dSigHi), numerators
 * +-2^nExp * (1 + t/2^23) for t = 0, nSigStep, 2*nSigStep, ... and +-0.  firstBad: smallest
 * failing bit pattern (sqrt) or divisor significand (division), ~0 if none. */
int rtp_selfcheck_sqrt(int device, uint32_t lo, uint32_t hi, uint64_t* mismatches, uint32_t* firstBad);
int rtp_selfcheck_division(int device, int32_t nExp, int32_t dExp, uint32_t dSigLo, uint32_t dSigHi,
                           uint32_t nSigStep, uint64_t* mismatches, uint32_t* firstBad);
/* The Clouds model's pinned exponential on the device: y64 = exp(x64) (double), y32 =
 * exp(x32) (float) and svd32 = saturationVaporDensity(x32) (clouds.cl:91-96), n values each,
 * for comparison with the CPU's evaluation of the same algorithm. */
int rtp_selfcheck_exp(int device, const double* x64, const float* x32, uint32_t n, double* y64,
                      float* y32, float* svd32);

#ifdef __cplusplus
}
#endif
#endif /* RTP_H */
