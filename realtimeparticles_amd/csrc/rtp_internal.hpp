// Internal declarations shared by the HIP translation units of librtp.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "rtp.h"

namespace rtp {

// Records a failure for rtp_last_error() and returns the status.
int fail(int status, const std::string& msg);

#define RTP_HIP_TRY(expr)                                                                     \
  do {                                                                                        \
    hipError_t err_ = (expr);                                                                 \
    if (err_ != hipSuccess)                                                                   \
      return ::rtp::fail(RTP_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(err_));   \
  } while (0)

// The reference's compile-time "-D" constants of grid.cl (Fluids.cpp:107-116,
// Boids.cpp:104-113), passed to kernels by value instead of being baked into the source.
struct GridConsts {
  float absWall[3];   // ABS_WALL_X/Y/Z
  float cellSize;     // GRID_CELL_SIZE_XYZ
  int32_t res[3];     // GRID_RES_X/Y/Z
  uint32_t numCells;  // GRID_NUM_CELLS
  uint32_t cap;       // NUM_MAX_PARTS_IN_CELL
};

// Everything the PBF passes read besides buffers: FluidParams (define.cl:13-25), the
// remaining program constants (Fluids.cpp:104-119) and values the host folds once.
struct PbfConsts {
  GridConsts g;
  rtp_fluid_params p;
  float h;         // EFFECT_RADIUS
  float hh;        // EFFECT_RADIUS * EFFECT_RADIUS (rounded once, as the kernel's expression)
  float cullSq;    // smallest float >= h*h (exact); d2 >= cullSq implies sqrt(d2) >= h
  float poly6;     // POLY6_COEFF
  float spiky;     // SPIKY_COEFF
  float wdq;       // poly6L(artPressureRadius * h, h) (fluids.cl:56)
  float maxVel;    // MAX_VEL
  float gravity;   // ABS_GRAVITY_ACC_Y
  float rho0sq;    // restDensity * restDensity (fluids.cl:190)
  float dtEps;     // timeStep + FLOAT_EPS (fluids.cl:272)
  float gdt;       // -ABS_GRAVITY_ACC_Y * timeStep (fluids.cl:71)
  float lo[3];     // fld_applyBoundaryCondition lower bounds (fluids.cl:437)
  float hi[3];     // upper bounds (fluids.cl:438)
  // Candidate test on the fused squared distance fma(dz,dz, fma(dy,dy, dx*dx)), which is
  // within 4 ulp of the reference's unfused one: skipping at fused >= cullFused (cullSq
  // plus 2^-20 relative) and < farFused (kFarD2 less 2^-20 relative) skips only what the
  // unfused test would.
  float cullFused;
  float farFused;
  int32_t shortMath;  // kShortSqrtGrad | kShortArtDiv: which short correctly rounded forms apply
};

// Squared distances at and above this take the reference's exact expressions (its
// zero-weighted poly6 overflows to NaN well above it).
constexpr float kFarD2 = 1.0e12f;
enum ShortMath { kShortSqrtGrad = 1, kShortArtDiv = 2 };

struct BoidsConsts {
  GridConsts g;
  rtp_boids_params p;
  float radiusSq;  // EFFECT_RADIUS_SQUARED
  float dt;        // update time step
  float vmin;      // 0.2f * maxVelocity (boids.cl:256)
  // variants of Boids::update (Boids.cpp:346-372)
  int32_t dim2 = 0;      // bd_applyBoidsRulesWithGrid2D instead of the 3-D rules
  int32_t cyclic = 0;    // Boundary::CyclicWall: bd_updatePosAndApplyPeriodicBC
  int32_t targetOn = 0;  // bd_addTargetRule with the fields below
  float target[3] = { 0.0f, 0.0f, 0.0f };
  float targetRadius = 0.0f;
  float targetSign = 0.0f;  // TargetParams.targetSignEffect as the float the rule multiplies by
  int32_t shortDiv = 0;     // radiusSq <= 2^30: the pair divisions take the short form
};

// The persistent boids update (one cooperative launch for `steps` steps): the buffers of
// rtp_boids and the sort's pass count.  bar: 3 words, zeroed before the launch (arrival
// count, generation, timeout flag).
struct BoidsPersistentArgs {
  float4* pos[2];
  float4* vel[2];
  float4* unitVel;
  float4* acc;
  uint32_t* key[2];
  uint32_t* idx[2];
  uint2* startEnd;
  uint32_t* bar;
  uint32_t n, sortN, maxN, steps;
  int cur, keyCur, passes;
};
int launch_boids_persistent(const BoidsPersistentArgs& a, const BoidsConsts& c, uint32_t groups,
                            hipStream_t s);
// Largest co-resident workgroup count of the persistent update (0 if unknown).
uint32_t boids_persistent_max_groups(int dim2);

// Largest artificial-pressure exponent a parameter set may carry: the kernels raise to it
// with a left-to-right product, one multiplication per unit (the reference's range is 1-6).
constexpr uint32_t kMaxArtPressureExp = 64;

// Survivor-queue depth of the neighbour passes when a tuning gives 0.
constexpr int kDefaultQueue = 16;

// Launch shape shared by the per-particle kernels.
struct Launch {
  int block = 256;     // threads per block
  int xcdRemap = 0;    // 1: consecutive particle blocks on one XCD
  int tile = 0;        // 1: neighbour passes stage their candidates in LDS wave tiles
  int queue = kDefaultQueue;  // survivor-queue depth of the neighbour passes (entries per lane)
  int resort = 0;      // 1: after every correction, order the neighbour passes' threads by
                       // the cells their particles now sit in
  int nbList = 0;      // 1: the second pass of each pair over the same positions replays the
                       // first one's survivor lists
  int bricks = 0;      // 1 (with resort): thread order by brick_key instead of the cell index
  int quad = 0;        // 1: the plain solver's neighbour passes run a quad of lanes per particle
};
// Models of at most this many particles start with quads (Launch::quad): below it one lane
// per particle leaves most SIMDs with a single wave or none.
constexpr uint32_t kQuadDefaultMaxN = 1u << 18;

// Tile edge, in cells, of the (y, z) pencils of the brick thread order.
constexpr uint32_t kBrick = 8;

// ---------------------------------------------------------------------------- grid / sort
// Stable LSD radix sort of (key, index) pairs on the device.
struct SortPlan {
  uint32_t n = 0;          // keys to sort
  int passes = 0;
  int bits = 0;            // digit width
  int items = 32;          // keys per lane of a subtile (sort_items)
  uint32_t subtiles = 0;   // wave-owned subtiles of items*64 keys
  size_t countWords = 0;   // radix * subtiles
  int keyBits = 0;         // significant bits of the largest key
};
// The one-workgroup LDS sort (all passes in one launch) takes at most this many keys.
constexpr uint32_t kSmallSortMax = 8192;
constexpr int kSortItems = 32;  // keys per lane of a subtile, for 2^22 keys and more
// Keys per lane of a subtile for a sort of n keys: smaller subtiles below 2^22 keys, so a
// pass has about a thousand waves or more instead of n / 2048.
constexpr int sort_items(uint32_t n)
{
  return n < (1u << 18) ? 4 : (n < (1u << 20) ? 8 : (n < (1u << 22) ? 16 : kSortItems));
}
// Words of the count table one block of the multi-block scan owns (256 threads x 16 words).
constexpr uint32_t kScanSeg = 4096;
// Words of the counts buffer for n keys in subtiles of items*64 (any digit width): the
// digit-major count table (at most 2048 digits x subtiles) followed by its scan's segment sums.
constexpr size_t count_words(uint32_t n, int items)
{
  const size_t per = (size_t)items * 64;
  const size_t subtiles = n ? (n + per - 1) / per : 1;
  const size_t table = (size_t)2048 * subtiles;
  return table + table / kScanSeg + 64;
}
// ... for a sort of up to maxN keys with the subtiles sort_items picks: the largest table
// of each size class up to maxN.
constexpr size_t count_capacity(uint32_t maxN)
{
  size_t w = count_words(maxN, sort_items(maxN));
  const uint32_t a = maxN < (1u << 18) ? maxN : (1u << 18) - 1u;
  const uint32_t b = maxN < (1u << 20) ? maxN : (1u << 20) - 1u;
  const uint32_t d = maxN < (1u << 22) ? maxN : (1u << 22) - 1u;
  w = w > count_words(a, sort_items(a)) ? w : count_words(a, sort_items(a));
  w = w > count_words(b, sort_items(b)) ? w : count_words(b, sort_items(b));
  w = w > count_words(d, sort_items(d)) ? w : count_words(d, sort_items(d));
  return w;
}
// fixedItems: 0 for sort_items(n), else the keys per lane of every subtile (4, 8, 16 or 32).
SortPlan make_sort_plan(uint32_t n, uint32_t maxKey, int bitsOverride, int fixedItems = 0);
// keys/idx ping-pong through (keyA, idxA) <-> (keyB, idxB); returns whether the sorted
// result ended in the B buffers.  idxIn == nullptr means the identity permutation.
int sort_pairs(const SortPlan& plan, uint32_t* keyA, uint32_t* keyB, uint32_t* idxA,
               uint32_t* idxB, uint32_t* counts, bool* resultInB, hipStream_t s);
// Whether sort_pairs(plan, ...) on the current device leaves its result in the B buffers
// (known before the launch, so a caller can put its own output buffers on that side).
bool sort_result_in_b(const SortPlan& plan);
// Passes of the one-workgroup LDS sort when it takes this plan (n <= kSmallSortMax and its
// LDS fits the device), else 0.
int small_sort_pass_count(const SortPlan& plan);

void launch_fill_keys_and_reset_cells(const float4* pos, uint32_t n, uint32_t* key,
                                      uint2* startEnd, const GridConsts& g, hipStream_t s);
void launch_reset_cell_ids(uint32_t* key, uint32_t maxN, uint32_t numCells, hipStream_t s);
void launch_fill_identity(uint32_t* idx, uint32_t begin, uint32_t end, hipStream_t s);
// Gather float4 payloads by idx over [0, sortN) and fill cell start/end over [0, n);
// payload C of the active range [0, n) is clamped to the PBF walls when clampC is set.
struct ReorderArgs {
  const uint32_t* key;
  const uint32_t* idx;
  uint32_t sortN, n, maxN;
  const float4* srcA; float4* dstA;
  const float4* srcB; float4* dstB;
  const float4* srcC; float4* dstC;
  int clampC;
  float lo[3], hi[3];
  uint2* startEnd;
  GridConsts g;
};
void launch_reorder(const ReorderArgs& a, hipStream_t s);
void launch_copy_f4(float4* dst, const float4* src, uint32_t begin, uint32_t end, hipStream_t s);

// ------------------------------------------------------------------ rendering side of a frame
// fillCameraDist (utils.cl:43-52) over [0, n); cam points at u_cameraPos (xyz).
void launch_camera_key(const float4* pos, const float4* cam, uint32_t* key, uint32_t n,
                       hipStream_t s);
// fld_fillFluidColor (fluids.cl:458-478) over [0, n).
void launch_fluid_color(const float* density, float restDensity, float4* col, uint32_t n,
                        hipStream_t s);
void launch_fill_f4v(float4* dst, float4 v, uint32_t begin, uint32_t end, hipStream_t s);
// resetGridDetector + fillGridDetector (grid.cl:43-60); det holds two float4 per cell.
void launch_detector(const float4* pos, float4* det, uint32_t n, const GridConsts& g,
                     hipStream_t s);
// dst[k][i] = src[k][idx[i]] over [begin, end) for every k whose src is set.
struct GatherF4 {
  const float4* src[4] = { nullptr, nullptr, nullptr, nullptr };
  float4* dst[4] = { nullptr, nullptr, nullptr, nullptr };
};
void launch_gather_f4(const uint32_t* idx, uint32_t begin, uint32_t end, const GatherF4& g,
                      hipStream_t s);

// ---------------------------------------------------------------------------------- PBF
void launch_pbf_predict(const float4* pos, const float4* vel, float4* pred, uint32_t* key,
                        uint2* startEnd, uint32_t n, const PbfConsts& c, hipStream_t s);
// Optional buffers of the warm-start, teacher and fallback variants of the PBF passes.
struct LearnArgs {
  float* feat = nullptr;             // [n][8] features (kFeatures)
  const float4* vel = nullptr;       // sorted velocities (speed feature)
  float* lamAcc = nullptr;           // accumulated lambda (kAccum)
  int accumInit = 0;
};
// kNbRecord / kNbReplay: the pass records its threads' survivor lists / replays them
// (NbList below).
enum PassMode {
  kFeatures = 1, kAccum = 2, kDensOnly = 4, kClampIn = 8, kNbRecord = 32, kNbReplay = 64
};
// Survivor lists.  Two neighbour passes over the same positions, the same cell table and
// the same thread order visit the same candidates and run their bodies on the same
// survivors in the same order.  The first records, for its thread t, the index of every
// neighbour its body runs on -- the k-th (k < kNbCap) as component k % 4 of the 16-byte
// granule list[(k / 4) * stride + t], so each thread stores four at a time and a wave's
// stores and loads of one granule row are contiguous -- and their number to count[t].  The
// second runs its body over the recorded indices instead of walking the cells, or walks
// when count[t] > kNbCap.  256 entries (1 KiB per thread slot) hold the survivors of the
// clumps the dam-break coefficients form, whose cells fill to the cap.
constexpr uint32_t kNbCap = 256;
struct NbList {
  uint4* list = nullptr;      // kNbCap / 4 x stride granules
  uint32_t* count = nullptr;  // stride words
  uint32_t stride = 0;        // threads of the pass
};
// Which particle each thread of a neighbour pass handles, and (correction passes) where to
// write the cell key of every output position so the next pass can order its threads by
// the cells they will walk.  Results do not depend on the order.  nb: the survivor lists
// a kNbRecord / kNbReplay pass writes / reads.
struct PassOrder {
  const uint32_t* order = nullptr;  // thread t handles particle order[t]; identity when null
  uint32_t* keyOut = nullptr;       // correction passes: cell key of each output position
  uint32_t keyBase = 0;             //   plus this (a batched scene's key base)
  int bricks = 0;                   //   1: brick_key of the cell instead of its index
  NbList nb;
  __device__ uint32_t particle(uint32_t t) const { return order ? order[t] : t; }
};
// The thread order's two count tables of `words` words each, both all zero when allocated.
// A call counts in table[cur] (all zero on entry) and clears table[1 - cur] for the next
// call; the caller flips cur after every call that ran (a hipGraph replays its calls from
// the cur it was captured at, so a caller recaptures when cur differs).
struct CountTables {
  uint32_t* table[2] = { nullptr, nullptr };
  uint32_t words = 0;
  int cur = 0;
};
// Counting sort of particles by cell for the thread order: the count table (m16 <= words
// entries, a multiple of 16 covering every key) counts the particles of each cell in the
// previous thread order prev (the identity when null; ranks into rank[t]) and becomes its
// exclusive scan; then order[scan[key[i]] + rank[t]] = i for i = prev[t].  order must not
// be prev.  partial holds scan_partials_needed(m16) words.  Returns whether it launched
// (n and m16 nonzero): then the caller flips ct.cur.
void launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t s);
bool launch_counting_order(const CountTables& ct, uint32_t m16, const uint32_t* key,
                           const uint32_t* prev, uint32_t* rank, uint32_t n, uint32_t* partial,
                           uint32_t* order, hipStream_t s);
uint32_t scan_partials_needed(uint32_t m16);
enum LastMode { kNextIter = 0, kLastVel = 1, kLastOnly = 2 };

// ------------------------------------------------------------------------------ batches
// Many independent PBF scenes advanced by the same launches.  Scene s owns particles
// [pbase, pbase + n) of every particle array and the key range [kbase, kbase + keySpan) of
// the cell-key space and of the cell table (keySpan = the largest key its hash can produce
// + 1), so one global sort of the keys leaves every scene's particles in its own range, in
// its own order.  Cell tables and neighbour indices inside a scene are scene-local, as in a
// model of its own.
struct SceneSlot {
  uint32_t pbase, n, kbase, keySpan;
};
// blocks[b] = (scene, first scene-local particle) of block b of a batched launch.
struct BatchView {
  const SceneSlot* scenes;
  const uint2* blocks;
  const PbfConsts* consts;
  uint32_t numBlocks, block;
  uint32_t qdepth;                  // survivor-queue depth of the neighbour passes
  const uint32_t* order = nullptr;  // batch-wide thread order (scene slices stay in place)
  uint32_t* keyOut = nullptr;       // correction passes: batch-wide key of each output
  NbList nb;                        // survivor lists; a thread's slot is its launch-wide id
};
void launch_fill_start_end(uint2* startEnd, uint32_t count, hipStream_t s);
void launch_pbf_predict_batch(const BatchView& bv, const float4* pos, const float4* vel,
                              float4* pred, uint32_t* key, hipStream_t s);
// ReorderArgs with whole-batch pointers; per-scene counts, bounds and grid come from bv.
void launch_reorder_batch(const BatchView& bv, const ReorderArgs& a, hipStream_t s);
void launch_pbf_density_lambda_batch(const BatchView& bv, const float4* pred, const uint2* startEnd,
                                     float* density, float* lambda, hipStream_t s, int nb = 0);
void launch_pbf_correct_batch(const BatchView& bv, const float4* pred, const float* lambda,
                              const uint2* startEnd, float4* predOut, const float4* pos,
                              float4* vel, float4* posOut, int last, hipStream_t s, int nb = 0);
void launch_pbf_update_vel_batch(const BatchView& bv, const float4* pred, const float4* pos,
                                 float4* vel, float4* posOut, hipStream_t s);
void launch_pbf_vorticity_batch(const BatchView& bv, const float4* pred, const uint2* startEnd,
                                const float4* vel, float4* vort, hipStream_t s, int nb = 0);
void launch_pbf_confinement_batch(const BatchView& bv, const float4* pred, const uint2* startEnd,
                                  const float4* vort, const float4* vel, float4* velOut,
                                  hipStream_t s, int nb = 0);
void launch_pbf_xsph_batch(const BatchView& bv, const float4* pred, const uint2* startEnd,
                           const float4* velIn, float4* velOut, float4* posOut, hipStream_t s, int nb = 0);

void launch_pbf_density_lambda(const float4* pred, const uint2* startEnd, float* density,
                               float* lambda, uint32_t n, const PbfConsts& c, const Launch& L,
                               hipStream_t s, int mode = 0, const LearnArgs& la = LearnArgs(),
                        const PassOrder& po = PassOrder());
// Constraint correction + correctPosition, then per `last` (LastMode) the next iteration's
// boundary clamp, or updateVel (+ updatePosition when posOut), or nothing.
void launch_pbf_correct(const float4* pred, const float* lambda, const uint2* startEnd,
                        float4* predOut, const float4* pos, float4* vel, float4* posOut,
                        uint32_t n, int last, const PbfConsts& c, const Launch& L, hipStream_t s,
                        int mode = 0, const LearnArgs& la = LearnArgs(),
                        const PassOrder& po = PassOrder());
// The learned substep's per-particle fallback (no reference counterpart).  The set: the
// particles whose post-correction |rho/rho0 - 1| (resDensity) exceeds tau, and every
// particle within h of one of them.  The halo: every other particle within h of the set;
// its density and lambda are recomputed each fallback iteration, but it does not move.
// act (1 = set, 2 = halo), seeds, list (the set) and halo hold maxN entries; count[0] =
// seeds, count[1] = particles in the set, count[2] = particles in the halo.
struct FallbackSet {
  uint32_t* act = nullptr;
  uint32_t* seeds = nullptr;
  uint32_t* list = nullptr;
  uint32_t* halo = nullptr;
  uint32_t* count = nullptr;
};
void launch_fallback_set(const float4* pred, const uint2* startEnd, const float* resDensity,
                         float tau, uint32_t n, const PbfConsts& c, const FallbackSet& fs,
                         hipStream_t s);
// Plain Jacobi iteration f (fluids.cl:80-258) for the particles of the set, ending in pred:
// the first (when the set is not empty) clamps every particle to the walls, into pred and
// partner, and sets the lambda of every particle to 0; density and lambda are computed for
// the set and the halo, and the set's corrections written into the other buffer (iterations
// alternate between pred and partner); the last one of an odd count copies the set's final
// positions back into pred.  nb: survivor lists recorded by the density pass and replayed
// by the correction.
void launch_fallback_iteration(float4* pred, float4* partner, float* density, float* lambda,
                               const uint2* startEnd, uint32_t n, uint32_t f, bool last,
                               const PbfConsts& c, const Launch& L, const FallbackSet& fs,
                               const NbList* nb, hipStream_t s);
void launch_pbf_update_vel(const float4* pred, const float4* pos, float4* vel, float4* posOut,
                           uint32_t n, const PbfConsts& c, hipStream_t s);
// nb: 0, kNbRecord or kNbReplay (the post passes share their positions and thread order).
void launch_pbf_vorticity(const float4* pred, const uint2* startEnd, const float4* vel,
                          float4* vort, uint32_t n, const PbfConsts& c, const Launch& L,
                          hipStream_t s, const PassOrder& po = PassOrder(), int nb = 0);
void launch_pbf_confinement(const float4* pred, const uint2* startEnd, const float4* vort,
                            const float4* vel, float4* velOut, uint32_t n, const PbfConsts& c,
                            const Launch& L, hipStream_t s, const PassOrder& po = PassOrder(),
                            int nb = 0);
void launch_pbf_xsph(const float4* pred, const uint2* startEnd, const float4* velIn,
                     float4* velOut, float4* posOut, uint32_t n, const PbfConsts& c,
                     const Launch& L, hipStream_t s, const PassOrder& po = PassOrder(),
                     int nb = 0);
// Reduction of |rho/rho0 - 1| over n particles into metric[0..2] (max, sum, max
// compression), deterministic for a given n; metric holds kMetricFloats floats, the ones
// past metric[3] being per-block partials.
constexpr uint32_t kMetricFloats = 4 + 3 * 1024;
void launch_density_error(const float* density, uint32_t n, float rho0, float* metric,
                          hipStream_t s);

// -------------------------------------------------------------------------------- boids
void launch_boids_hash(const float4* pos, uint32_t* key, uint2* startEnd, uint32_t n,
                       const GridConsts& g, hipStream_t s);
// The rules read unit velocities from unitVel, written first over the sorted range sortN.
void launch_boids_step(const float4* pos, const float4* vel, float4* unitVel, const uint2* startEnd,
                       float4* posOut, float4* velOut, float4* acc, uint32_t n, uint32_t sortN,
                       const BoidsConsts& c, const Launch& L, hipStream_t s);

}  // namespace rtp
