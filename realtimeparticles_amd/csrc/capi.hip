// C ABI of librtp: model handles, device buffers, the per-substep launch sequence,
// hipGraph replay and per-kernel profiling.
//
// The launch sequence of one PBF substep (Fluids.cpp:400-471 without rendering):
//   k_predict -> radix sort (passes x {upsweep, scan, downsweep}) -> k_reorder
//   -> iters x {k_density_lambda, k_correct} -> [k_vorticity, k_confinement, k_xsph]
// and of one boids step (Boids.cpp:323-384):
//   k_boids_hash -> radix sort -> k_reorder -> k_boids_step
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "rtp_host.hpp"
#include "rtp_internal.hpp"
#include "rtp_mlp.hpp"

namespace rtp {

namespace {
thread_local std::string g_lastError;
}

int fail(int status, const std::string& msg)
{
  g_lastError = msg;
  return status;
}

namespace {


// The buffers of a frame's rendering steps (Fluids.cpp:132-164, Boids.cpp:128-153):
// u_cameraPos, p_col, p_cameraDist and c_partDetector, with gather / sort partners.
struct RenderState {
  bool on = false;
  float4* cam = nullptr;                    // u_cameraPos (xyz)
  float4 camHost = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
  float4* col[2] = { nullptr, nullptr };    // col[colCur] is p_col
  uint32_t* key[2] = { nullptr, nullptr };  // key[keyCur] is p_cameraDist
  float4* det = nullptr;                    // two float4 per cell
  int colCur = 0, keyCur = 0;
  // The largest particle count keyed since the reset: keys at or past it still hold
  // FAR_DIST, the largest key, so the stable sort leaves them where they are.
  uint32_t high = 0;

  int alloc(uint32_t maxN, uint32_t numCells, bool pairCol)
  {
    if (cam)
      return RTP_OK;
    int rc = 0;
    if ((rc = dev_alloc(&cam, 1)) || (rc = dev_alloc(&col[0], maxN)) ||
        (pairCol && (rc = dev_alloc(&col[1], maxN))) || (rc = dev_alloc(&key[0], maxN)) ||
        (rc = dev_alloc(&key[1], maxN)) || (rc = dev_alloc(&det, 2u * numCells)))
      return rc;
    return RTP_OK;
  }
  void release()
  {
    dev_free(cam);
    dev_free(col[0]);
    dev_free(col[1]);
    dev_free(key[0]);
    dev_free(key[1]);
    dev_free(det);
  }
  int set_camera(const float p[3], hipStream_t s)
  {
    camHost = make_float4(p[0], p[1], p[2], 0.0f);
    RTP_HIP_TRY(hipMemcpyAsync(cam, &camHost, sizeof(float4), hipMemcpyHostToDevice, s));
    return RTP_OK;
  }
  // The colour at model reset (Fluids.cpp:394 / bd_fillBoidsColor, Boids.cpp:267) and
  // resetCameraDist (Fluids.cpp:215, Boids.cpp:272) over maxN, in both partners.
  void reset(float4 colInit, uint32_t maxN, hipStream_t s)
  {
    for (int k = 0; k < 2; ++k) {
      if (col[k])
        launch_fill_f4v(col[k], colInit, 0, maxN, s);
      launch_fill_u32(key[k], maxN, 1000000u, s);
    }
    colCur = keyCur = 0;
    high = 0;
  }
  // Fluids.cpp:466-468 / Boids.cpp:379-381: camera distances of the n active particles
  // (from pos), then the second RadixSort::sort over [0, R) carrying the payloads of g
  // (sources set, destinations filled in here from dst).  R covers every key that is not
  // FAR_DIST and the grid's sorted range, past which every ping-pong pair agrees.
  int camera_sort(GridState& G, uint32_t n, const float4* pos, GatherF4 g, hipStream_t s,
                  Profiler& prof)
  {
    high = std::max(high, n);
    const uint32_t R = std::max(high, G.sortedN);
    RTP_PROF(prof, "camera_key", s, launch_camera_key(pos, cam, key[keyCur], n, s));
    // keys below 2^20: two passes of 10-bit digits whatever width the cell sort is tuned to
    const SortPlan plan = make_sort_plan(R, 1000000u, 10);
    bool inB = false;
    int rc = 0;
    RTP_PROF(prof, "camera_sort", s,
             rc = sort_pairs(plan, key[keyCur], key[1 - keyCur], G.idx[0], G.idx[1], G.counts, &inB, s));
    if (rc)
      return rc;
    if (inB)
      keyCur = 1 - keyCur;
    G.idxCur = inB ? 1 : 0;
    G.permN = R;
    RTP_PROF(prof, "camera_gather", s, launch_gather_f4(G.idx[G.idxCur], 0, R, g, s));
    return RTP_OK;
  }
};

// -------------------------------------------------------------- no-allocation grid build

// The sort's partner key and index buffers and its digit counts for rtp_grid_build_noalloc:
// one array per GPU, held until the process ends.  Host code never passes `scratch` itself
// (its name on the host is a stand-in, not a device address); it passes the address
// hipGetSymbolAddress returns on each device.
constexpr uint32_t kNoAllocMax = RTP_NOALLOC_MAX_PARTICLES;
// (the sort here keeps subtiles of kSortItems keys per lane at every size)
constexpr size_t kNoAllocWords = 2 * (size_t)kNoAllocMax + count_words(kNoAllocMax, kSortItems);
static __device__ __align__(256) uint32_t scratch[kNoAllocWords];

constexpr int kMaxDevices = 64;

// The current device's copy of `scratch`, looked up once per device (the device must be
// current).  Returns nullptr after recording the error.
uint32_t* scratch_on_current_device()
{
  static std::mutex mu;
  static uint32_t* ptr[kMaxDevices] = {};
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    fail(RTP_ERR_HIP, "hipGetDevice failed or device ordinal too large");
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(mu);
  if (!ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      fail(RTP_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      return nullptr;
    }
    ptr[dev] = static_cast<uint32_t*>(p);
  }
  return ptr[dev];
}

}  // namespace
}  // namespace rtp

using namespace rtp;

// =================================================================================== MLP

struct rtp_mlp {
  int device = 0;
  hipStream_t stream = nullptr;
  float* p = nullptr;     // fp32 master parameters
  float* grad = nullptr;
  float* m1 = nullptr;    // Adam moments
  float* m2 = nullptr;
  float* loss = nullptr;
  short* w1t = nullptr;
  short* w2t = nullptr;
  short* w2 = nullptr;
  MlpDev dev{};
  int step = 0;
  float* dFeat = nullptr;  // training / inference staging
  float* dTarget = nullptr;
  float* dOut = nullptr;
  uint32_t* dPerm = nullptr;
  uint32_t cap = 0;

  int reserve(uint32_t n)
  {
    if (n <= cap)
      return RTP_OK;
    dev_free(dFeat);
    dev_free(dTarget);
    dev_free(dOut);
    dev_free(dPerm);
    int rc = 0;
    if ((rc = dev_alloc(&dFeat, (size_t)n * 8)) || (rc = dev_alloc(&dTarget, n)) ||
        (rc = dev_alloc(&dOut, n)) || (rc = dev_alloc(&dPerm, n)))
      return rc;
    cap = n;
    return RTP_OK;
  }
  void release()
  {
    if (stream)
      (void)hipStreamSynchronize(stream);
    for (float** b : { &p, &grad, &m1, &m2, &loss, &dFeat, &dTarget, &dOut })
      dev_free(*b);
    dev_free(w1t);
    dev_free(w2t);
    dev_free(w2);
    dev_free(dPerm);
    if (stream)
      (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

// ================================================================================ fluids

struct rtp_fluids {
  int device = 0;
  hipStream_t stream = nullptr;
  GridState grid;
  PbfConsts c{};
  uint32_t n = 0;
  uint32_t iters = 2;  // m_nbJacobiIters default, Fluids.cpp:81
  bool paused = false;
  float4* pos[2] = { nullptr, nullptr };
  float4* vel[2] = { nullptr, nullptr };
  float4* pred[2] = { nullptr, nullptr };
  float4* vort = nullptr;
  float4* velTmp = nullptr;
  float* density = nullptr;
  float* lambda = nullptr;
  float* metric = nullptr;
  int cur = 0;        // pos[cur]/vel[cur] are p_pos/p_vel
  int predFinal = 1;  // pred[predFinal] is p_predPos after the last update
  Launch L;
  Profiler prof;
  bool useGraph = true;
  bool graphValid = false;
  int graphCur = -1, graphKeyCur = -1, graphPred = -1, graphCol = -1, graphCam = -1;
  int graphCount = -1;  // ocount.cur at the capture
  RenderState render;  // rtp_fluids_set_render
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;
  // learned warm-start, teacher recording and the post-correction residual pass
  float* feat = nullptr;      // [maxN][8]
  float* lamAcc = nullptr;
  // thread order of the neighbour passes (counting sort of the corrected positions' cells)
  uint32_t* okey = nullptr;
  uint32_t* orank = nullptr;
  uint32_t* oorder[2] = { nullptr, nullptr };  // thread orders, each counted in the other
  CountTables ocount;  // the thread order's count tables
  uint32_t* opart = nullptr;
  // survivor lists (RTP_LAUNCH_NB_LIST), allocated when first enabled
  uint4* nbList = nullptr;      // kNbCap / 4 x maxN granules
  uint32_t* nbCount = nullptr;  // maxN
  float* resDensity = nullptr;
  float* resMetric = nullptr; // max, sum, max compression of the residual pass + partials
  rtp_mlp* mlp = nullptr;
  uint32_t learnedIters = 0, fallbackIters = 0;
  float tau = 0.0f;  // per-particle fallback threshold on |rho/rho0 - 1|
  FallbackSet fb;    // allocated with the first fallback setting
  bool residualPass = false;
  uint32_t teacherIters = 0;  // > 0 while rtp_fluids_record_teacher runs its substep
  // checkpoint (rtp_fluids_save_checkpoint)
  float4* ckPos = nullptr;
  float4* ckVel = nullptr;
  float4* ckPred = nullptr;
  uint32_t* ckKey = nullptr;
  float4* ckCol = nullptr;
  uint32_t* ckCam = nullptr;
  bool ckValid = false, ckRender = false;
  uint32_t ckN = 0, ckDirty = 0, ckSorted = 0, ckHigh = 0;

  void drop_graph()
  {
    if (graphExec)
      (void)hipGraphExecDestroy(graphExec);
    if (graph)
      (void)hipGraphDestroy(graph);
    graphExec = nullptr;
    graph = nullptr;
    graphValid = false;
  }

  void refresh_consts()
  {
    refresh_pbf_consts(c);
    drop_graph();
  }

  int substep()
  {
    hipStream_t s = stream;
    GridState& G = grid;
    // p_predPos is pred[predFinal]; predict overwrites its active range only, so entries
    // past n keep the previous step's values, as in the reference's single buffer.
    const int P = predFinal;
    const bool learned = mlp && learnedIters > 0 && teacherIters == 0;
    const bool teacher = teacherIters > 0;
    const bool wantFeatures = learned || teacher;
    RTP_PROF(prof, "pbf_predict_hash", s,
             launch_pbf_predict(pos[cur], vel[cur], pred[P], G.key[G.keyCur], G.startEnd, n, c, s));
    int rc = G.sort(n, s, prof);
    if (rc)
      return rc;
    const uint32_t iters = teacher ? teacherIters : (learned ? learnedIters : this->iters);
    ReorderArgs a{};
    a.key = G.key[G.keyCur];
    a.idx = G.idx[G.idxCur];
    a.sortN = G.sortedN;
    a.n = n;
    a.maxN = G.maxN;
    a.srcA = pos[cur];
    a.dstA = pos[1 - cur];
    a.srcB = vel[cur];
    a.dstB = vel[1 - cur];
    a.srcC = pred[P];
    a.dstC = pred[1 - P];
    a.clampC = iters > 0;
    for (int k = 0; k < 3; ++k) {
      a.lo[k] = c.lo[k];
      a.hi[k] = c.hi[k];
    }
    a.startEnd = G.startEnd;
    a.g = c.g;
    RTP_PROF(prof, "reorder_cells", s, launch_reorder(a, s));
    cur = 1 - cur;
    const bool vortOn = c.p.isVorticityConfEnabled != 0;
    LearnArgs la;
    la.feat = feat;
    la.vel = vel[cur];
    la.lamAcc = lamAcc;
    int in = 1 - P;
    // Thread order of the neighbour passes: the grid's order for the first pass; with
    // L.resort, after every correction the particles sorted by the cells they now sit in.
    PassOrder po;
    const bool resort = L.resort != 0;
    PassOrder pc;  // of the correction passes: also write the cells of their outputs
    // survivor lists: recorded by each plain density pass (and by vorticity), replayed by
    // the correction pass of the same iteration (and by confinement and XSPH)
    const bool lists = L.nbList != 0 && L.tile == 0 && nbList != nullptr;
    if (lists) {
      po.nb = NbList{ nbList, nbCount, n };
      pc.nb = po.nb;
    }
    const uint32_t m16 = L.bricks ? brick_order_cells(c.g) : order_cells(c.g);
    pc.keyOut = resort ? okey : nullptr;
    pc.bricks = L.bricks;
    auto order_threads = [&]() -> int {
      if (!resort)
        return RTP_OK;
      uint32_t* next = po.order == oorder[0] ? oorder[1] : oorder[0];
      bool ran = false;
      RTP_PROF(prof, "thread_order", s,
               ran = launch_counting_order(ocount, m16, okey, po.order, orank, n, opart, next, s));
      if (ran)
        ocount.cur = 1 - ocount.cur;
      po.order = next;
      pc.order = next;
      return RTP_OK;
    };
    for (uint32_t k = 0; k < iters; ++k) {
      const bool lastIt = (k + 1 == iters);
      int mode = 0;
      if (k == 0 && wantFeatures)
        mode |= kFeatures;
      if (teacher) {
        mode |= kAccum;
        la.accumInit = k == 0;
      }
      // the plain and feature-writing density passes record; the teacher's accumulating
      // ones do not
      const bool replay = lists && (mode & ~kFeatures) == 0;
      RTP_PROF(prof, "pbf_density_lambda", s,
               launch_pbf_density_lambda(pred[in], G.startEnd, density, lambda, n, c, L, s,
                                         replay ? (mode | kNbRecord) : mode, la, po));
      if (learned && k == 0)
        RTP_PROF(prof, "mlp_warm_start", s, launch_mlp_forward(feat, n, mlp->dev, lambda, s));
      const int last = !lastIt ? kNextIter : (learned ? kLastOnly : kLastVel);
      RTP_PROF(prof, "pbf_correct", s,
               launch_pbf_correct(pred[in], lambda, G.startEnd, pred[1 - in], pos[cur], vel[cur],
                                  (last == kLastVel && !vortOn) ? pos[cur] : nullptr, n, last, c, L, s,
                                  replay ? kNbReplay : 0, LearnArgs(), pc));
      in = 1 - in;
      if ((rc = order_threads()))
        return rc;
    }
    // post-correction density: the residual the fallback gate and the diagnostics read
    auto residual = [&]() {
      RTP_PROF(prof, "pbf_residual", s,
               launch_pbf_density_lambda(pred[in], G.startEnd, resDensity, nullptr, n, c, L, s, kDensOnly, la, po));
      launch_density_error(resDensity, n, c.p.restDensity, resMetric, s);
    };
    // per-particle fallback: the particles the learned iterations left above tau, and their
    // neighbours within h, run plain iterations over a compacted list; the rest keep their
    // positions (the thread order of the later passes stays the last learned one's)
    if (learned && fallbackIters > 0) {
      residual();
      RTP_PROF(prof, "pbf_fallback_set", s,
               launch_fallback_set(pred[in], G.startEnd, resDensity, tau, n, c, fb, s));
      // ending in pred[in]; the iterations alternate between it and its partner buffer
      for (uint32_t f = 0; f < fallbackIters; ++f)
        RTP_PROF(prof, "pbf_fallback", s,
                 launch_fallback_iteration(pred[in], pred[1 - in], density, lambda, G.startEnd, n,
                                           f, f + 1 == fallbackIters, c, L, fb,
                                           lists ? &po.nb : nullptr, s));
    }
    if (residualPass)  // after any fallback iterations: the error the substep leaves
      residual();
    if (iters == 0 || learned)
      RTP_PROF(prof, "pbf_update_vel", s,
               launch_pbf_update_vel(pred[in], pos[cur], vel[cur], vortOn ? nullptr : pos[cur], n,
                                     c, s));
    // Permuted entries in [n, sortedN) live in the gather target; keep them in p_predPos.
    if (in != 1 - P && G.sortedN > n)
      RTP_PROF(prof, "copy_tail", s, launch_copy_f4(pred[in], pred[1 - P], n, G.sortedN, s));
    predFinal = in;
    if (vortOn) {
      RTP_PROF(prof, "pbf_vorticity", s,
               launch_pbf_vorticity(pred[in], G.startEnd, vel[cur], vort, n, c, L, s, po,
                                    lists ? kNbRecord : 0));
      RTP_PROF(prof, "pbf_confinement", s,
               launch_pbf_confinement(pred[in], G.startEnd, vort, vel[cur], velTmp, n, c, L, s, po,
                                      lists ? kNbReplay : 0));
      RTP_PROF(prof, "pbf_xsph", s,
               launch_pbf_xsph(pred[in], G.startEnd, velTmp, vel[cur], pos[cur], n, c, L, s, po,
                               lists ? kNbReplay : 0));
    }
    return render.on ? render_frame(true) : RTP_OK;
  }

  // Fluids.cpp:458-468: after the physics, the grid detector (from p_pos) and the colour
  // (from p_density); then, paused or not, the camera-distance sort permuting p_pos, p_col,
  // p_vel and p_predPos.  p_col rides the cell sort like the other payloads (:417), but
  // the colour fill overwrites its active range, so only its tail [n, sortedN) is gathered.
  int render_frame(bool physics)
  {
    hipStream_t s = stream;
    GridState& G = grid;
    RenderState& R = render;
    if (physics) {
      GatherF4 t;
      t.src[0] = R.col[R.colCur];
      t.dst[0] = R.col[1 - R.colCur];
      RTP_PROF(prof, "render_col_tail", s, launch_gather_f4(G.idx[G.idxCur], n, G.sortedN, t, s));
      RTP_PROF(prof, "render_detector", s, launch_detector(pos[cur], R.det, n, c.g, s));
      RTP_PROF(prof, "render_color", s,
               launch_fluid_color(density, c.p.restDensity, R.col[1 - R.colCur], n, s));
      R.colCur = 1 - R.colCur;
    }
    GatherF4 g;
    g.src[0] = pos[cur];
    g.dst[0] = pos[1 - cur];
    g.src[1] = vel[cur];
    g.dst[1] = vel[1 - cur];
    g.src[2] = pred[predFinal];
    g.dst[2] = pred[1 - predFinal];
    g.src[3] = R.col[R.colCur];
    g.dst[3] = R.col[1 - R.colCur];
    const int rc = R.camera_sort(G, n, pos[cur], g, s, prof);
    if (rc)
      return rc;
    cur = 1 - cur;
    predFinal = 1 - predFinal;
    R.colCur = 1 - R.colCur;
    return RTP_OK;
  }

  // The colour and camera distances as at model reset, and the detector of the current
  // positions (Fluids.cpp:209-215).
  void render_reset()
  {
    render.reset(make_float4(0.0f, 0.1f, 1.0f, 0.0f), grid.maxN, stream);
    launch_detector(pos[cur], render.det, n, c.g, stream);
  }

  int capture_pair()
  {
    drop_graph();
    graphCur = cur;
    graphKeyCur = grid.keyCur;
    graphPred = predFinal;
    graphCol = render.colCur;
    graphCam = render.keyCur;
    graphCount = ocount.cur;
    RTP_HIP_TRY(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
    int rc = substep();
    if (!rc)
      rc = substep();
    hipGraph_t gr = nullptr;
    const hipError_t e = hipStreamEndCapture(stream, &gr);
    if (rc)
      return rc;
    if (e != hipSuccess)
      return fail(RTP_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    graph = gr;
    RTP_HIP_TRY(hipGraphInstantiate(&graphExec, graph, nullptr, nullptr, 0));
    graphValid = true;
    // Capturing does not run the work; the buffer indices advanced twice and are back.
    return RTP_OK;
  }

  int update(uint32_t nsteps)
  {
    if (nsteps == 0)
      return RTP_OK;
    if (paused) {
      // A paused update still sorts by camera distance (Fluids.cpp:466-468).  The second
      // of several paused sorts finds its keys in order and leaves the identity, as every
      // later one does.
      if (!render.on)
        return RTP_OK;
      RTP_HIP_TRY(hipSetDevice(device));
      for (uint32_t k = 0; k < std::min(nsteps, 2u); ++k) {
        const int rc = render_frame(false);
        if (rc)
          return rc;
      }
      RTP_HIP_TRY(hipGetLastError());
      return RTP_OK;
    }
    MarkRange mark("rtp_fluids_update");
    RTP_HIP_TRY(hipSetDevice(device));
    uint32_t done = 0;
    if (useGraph && !prof.enabled && nsteps >= 2 && n > 0) {
      if (!graphValid || graphCur != cur || graphKeyCur != grid.keyCur || graphPred != predFinal ||
          graphCol != render.colCur || graphCam != render.keyCur || graphCount != ocount.cur) {
        const int rc = capture_pair();
        if (rc)
          return rc;
      }
      for (; done + 2 <= nsteps; done += 2) {
        MarkRange pair("rtp_fluids_substep_pair_graph");
        RTP_HIP_TRY(hipGraphLaunch(graphExec, stream));
      }
    }
    for (; done < nsteps; ++done) {
      MarkRange one("rtp_fluids_substep");
      const int rc = substep();
      if (rc)
        return rc;
    }
    RTP_HIP_TRY(hipGetLastError());
    return RTP_OK;
  }

  void release()
  {
    if (stream)
      (void)hipStreamSynchronize(stream);
    drop_graph();
    prof.destroy();
    grid.release();
    for (int k = 0; k < 2; ++k) {
      dev_free(pos[k]);
      dev_free(vel[k]);
      dev_free(pred[k]);
    }
    dev_free(vort);
    dev_free(velTmp);
    dev_free(density);
    dev_free(lambda);
    dev_free(metric);
    dev_free(ckPos);
    dev_free(ckVel);
    dev_free(ckPred);
    dev_free(ckKey);
    dev_free(ckCol);
    dev_free(ckCam);
    render.release();
    dev_free(feat);
    dev_free(fb.act);
    dev_free(fb.seeds);
    dev_free(fb.list);
    dev_free(fb.halo);
    dev_free(fb.count);
    dev_free(lamAcc);
    dev_free(resDensity);
    dev_free(resMetric);
    dev_free(okey);
    dev_free(orank);
    dev_free(oorder[0]);
    dev_free(oorder[1]);
    dev_free(ocount.table[0]);
    dev_free(ocount.table[1]);
    dev_free(opart);
    dev_free(nbList);
    dev_free(nbCount);
    if (stream)
      (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

// ================================================================================= boids

struct rtp_boids {
  int device = 0;
  hipStream_t stream = nullptr;
  GridState grid;
  BoidsConsts c{};
  uint32_t n = 0;
  bool paused = false;
  float4* pos[2] = { nullptr, nullptr };
  float4* vel[2] = { nullptr, nullptr };
  float4* acc = nullptr;     // p_acc
  float4* accAlt = nullptr;  // its gather partner, with the rendering steps on
  float4* unitVel = nullptr; // the sorted velocities normalised, read by the rules
  int cur = 0;
  Launch L;
  Profiler prof;
  RenderState render;  // rtp_boids_set_render
  bool targetVisible = false;  // the "Show" switch of the parameter tree (drawing only)
  bool persistent = false;     // RTP_LAUNCH_PERSISTENT
  uint32_t* bar = nullptr;      // the persistent launch's barrier words
  uint32_t* barHost = nullptr;  // pinned copy of its timeout flag, read at the next sync
  bool barPending = false;

  // The steps of one update in one cooperative launch, when the configuration allows it
  // (returns RTP_OK with *done = false otherwise).
  int update_persistent(uint32_t nsteps, bool* done)
  {
    *done = false;
    GridState& G = grid;
    const uint32_t sn = G.sort_range(n);
    const SortPlan plan = make_sort_plan(sn, G.max_key(n), G.sortBits);
    const int passes = small_sort_pass_count(plan);
    if (render.on || n == 0 || passes == 0)
      return RTP_OK;
    const uint32_t most = boids_persistent_max_groups(c.dim2);
    uint32_t groups = (sn + (uint32_t)L.block - 1u) / (uint32_t)L.block;
    groups = std::min(std::max(groups, 1u), most);
    if (groups == 0)
      return RTP_OK;
    int rc = 0;
    if (!bar) {
      if ((rc = dev_alloc(&bar, 4)))
        return rc;
      if (hipHostMalloc(reinterpret_cast<void**>(&barHost), sizeof(uint32_t)) != hipSuccess) {
        barHost = nullptr;
        return fail(RTP_ERR_HIP, "hipHostMalloc failed");
      }
      *barHost = 0u;
    }
    launch_fill_u32(bar, 4, 0u, stream);
    BoidsPersistentArgs a{};
    for (int k = 0; k < 2; ++k) {
      a.pos[k] = pos[k];
      a.vel[k] = vel[k];
      a.key[k] = G.key[k];
      a.idx[k] = G.idx[k];
    }
    a.unitVel = unitVel;
    a.acc = acc;
    a.startEnd = G.startEnd;
    a.bar = bar;
    a.n = n;
    a.sortN = sn;
    a.maxN = G.maxN;
    a.steps = nsteps;
    a.cur = cur;
    a.keyCur = G.keyCur;
    a.passes = passes;
    RTP_PROF(prof, "boids_persistent", stream, rc = launch_boids_persistent(a, c, groups, stream));
    if (rc)
      return rc;
    RTP_HIP_TRY(hipMemcpyAsync(barHost, bar + 2, sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    barPending = true;
    if ((passes & 1) && (nsteps & 1u))
      G.keyCur = 1 - G.keyCur;
    G.idxCur = passes & 1;
    G.sortedN = sn;
    G.permN = sn;
    *done = true;
    return RTP_OK;
  }

  // After a stream synchronisation: whether a persistent launch since the last check ended
  // on a timed-out grid barrier.
  int check_persistent()
  {
    if (!barPending)
      return RTP_OK;
    barPending = false;
    if (barHost && *barHost != 0u) {
      *barHost = 0u;
      return fail(RTP_ERR_HIP, "persistent boids update: a grid barrier timed out (results invalid)");
    }
    return RTP_OK;
  }

  int step()
  {
    hipStream_t s = stream;
    GridState& G = grid;
    RTP_PROF(prof, "boids_hash", s,
             launch_boids_hash(pos[cur], G.key[G.keyCur], G.startEnd, n, c.g, s));
    int rc = G.sort(n, s, prof);
    if (rc)
      return rc;
    ReorderArgs a{};
    a.key = G.key[G.keyCur];
    a.idx = G.idx[G.idxCur];
    a.sortN = G.sortedN;
    a.n = n;
    a.maxN = G.maxN;
    a.srcA = pos[cur];
    a.dstA = pos[1 - cur];
    a.srcB = vel[cur];
    a.dstB = vel[1 - cur];
    a.startEnd = G.startEnd;
    a.g = c.g;
    RTP_PROF(prof, "reorder_cells", s, launch_reorder(a, s));
    // With the rendering steps on, p_acc rides the cell sort too (Boids.cpp:337): the step
    // overwrites its active range, so only the tail [n, sortedN) is gathered, and the step
    // writes the partner.
    float4* accOut = acc;
    if (render.on) {
      GatherF4 t;
      t.src[0] = acc;
      t.dst[0] = accAlt;
      RTP_PROF(prof, "render_acc_tail", s, launch_gather_f4(G.idx[G.idxCur], n, G.sortedN, t, s));
      accOut = accAlt;
    }
    // Sorted state now lives in [1 - cur]; the fused step writes the next state to [cur].
    RTP_PROF(prof, "boids_step", s,
             launch_boids_step(pos[1 - cur], vel[1 - cur], unitVel, G.startEnd, pos[cur], vel[cur],
                               accOut, n, G.sortedN, c, L, s));
    if (render.on)
      std::swap(acc, accAlt);
    // The tail [n, sortedN) was only permuted; carry it over unchanged.
    if (G.sortedN > n) {
      RTP_HIP_TRY(hipMemcpyAsync(pos[cur] + n, pos[1 - cur] + n, sizeof(float4) * (G.sortedN - n),
                                 hipMemcpyDeviceToDevice, s));
      RTP_HIP_TRY(hipMemcpyAsync(vel[cur] + n, vel[1 - cur] + n, sizeof(float4) * (G.sortedN - n),
                                 hipMemcpyDeviceToDevice, s));
    }
    return render.on ? render_frame(true) : RTP_OK;
  }

  // Boids.cpp:375-381: after the step, the grid detector (from p_pos); then, paused or
  // not, the camera-distance sort permuting p_pos, p_vel and p_acc.  p_col holds one colour
  // everywhere (bd_fillBoidsColor), which every permutation keeps.
  int render_frame(bool physics)
  {
    hipStream_t s = stream;
    if (physics)
      RTP_PROF(prof, "render_detector", s, launch_detector(pos[cur], render.det, n, c.g, s));
    GatherF4 g;
    g.src[0] = pos[cur];
    g.dst[0] = pos[1 - cur];
    g.src[1] = vel[cur];
    g.dst[1] = vel[1 - cur];
    g.src[2] = acc;
    g.dst[2] = accAlt;
    const int rc = render.camera_sort(grid, n, pos[cur], g, s, prof);
    if (rc)
      return rc;
    cur = 1 - cur;
    std::swap(acc, accAlt);
    return RTP_OK;
  }

  // bd_fillBoidsColor over maxN, the camera distances reset and the detector of the
  // current positions (Boids.cpp:265-272); p_acc's partner made equal to it.
  void render_reset()
  {
    render.reset(make_float4(1.0f, 0.02f, 0.02f, 0.5f), grid.maxN, stream);
    launch_detector(pos[cur], render.det, n, c.g, stream);
    (void)hipMemcpyAsync(accAlt, acc, sizeof(float4) * grid.maxN, hipMemcpyDeviceToDevice, stream);
  }

  void release()
  {
    if (stream)
      (void)hipStreamSynchronize(stream);
    prof.destroy();
    grid.release();
    for (int k = 0; k < 2; ++k) {
      dev_free(pos[k]);
      dev_free(vel[k]);
    }
    dev_free(acc);
    dev_free(accAlt);
    dev_free(unitVel);
    dev_free(bar);
    if (barHost)
      (void)hipHostFree(barHost);
    barHost = nullptr;
    render.release();
    if (stream)
      (void)hipStreamDestroy(stream);
    stream = nullptr;
  }
};

// Device addresses of the current rendering buffers.
int render_pointers(const RenderState& R, void** col4, void** cameraDist, void** detector)
{
  if (!R.cam)
    return fail(RTP_ERR_STATE, "the rendering steps were never enabled");
  if (col4)
    *col4 = R.col[R.colCur];
  if (cameraDist)
    *cameraDist = R.key[R.keyCur];
  if (detector)
    *detector = R.det;
  return RTP_OK;
}

// p_col (float4), p_cameraDist (uint32) or c_partDetector (float8 per cell, count cells).
int read_render_field(RenderState& R, int field, uint32_t numCells, float* out, uint32_t count,
                      hipStream_t s)
{
  if (!R.cam)
    return fail(RTP_ERR_STATE, "the rendering steps were never enabled");
  const void* src = nullptr;
  size_t elem = 0;
  if (field == RTP_FIELD_COLOR) {
    src = R.col[R.colCur];
    elem = sizeof(float4);
  } else if (field == RTP_FIELD_CAMERA_DIST) {
    src = R.key[R.keyCur];
    elem = sizeof(uint32_t);
  } else {
    RTP_REQUIRE(count <= numCells, "count exceeds the number of cells");
    src = R.det;
    elem = 2 * sizeof(float4);
  }
  RTP_HIP_TRY(hipMemcpyAsync(out, src, elem * count, hipMemcpyDeviceToHost, s));
  RTP_HIP_TRY(hipStreamSynchronize(s));
  return RTP_OK;
}

// ================================================================================= C ABI

extern "C" {

int rtp_abi_version(void) { return RTP_ABI_VERSION; }

const char* rtp_last_error(void) { return g_lastError.c_str(); }

int rtp_device_count(int* count)
{
  RTP_REQUIRE(count, "count is NULL");
  int c = 0;
  if (hipGetDeviceCount(&c) != hipSuccess)
    c = 0;
  *count = c;
  return RTP_OK;
}

// ------------------------------------------------------------------------------- fluids

int rtp_fluids_create(const rtp_model_desc* d, rtp_fluids** out)
{
  RTP_REQUIRE(d && out, "NULL argument");
  *out = nullptr;
  RTP_REQUIRE(d->maxNbParticles > 0, "maxNbParticles must be positive");
  RTP_REQUIRE(d->currNbParticles <= d->maxNbParticles, "currNbParticles exceeds maxNbParticles");
  int rc = check_device(d->device);
  if (rc)
    return rc;
  auto* h = new rtp_fluids();
  h->device = d->device;
  if ((rc = make_grid(d, 100u, &h->grid.g))) {
    delete h;
    return rc;
  }
  init_pbf_consts(h->c, h->grid.g, d->boxSize[0], d->gridRes[0], d->gravity);
  h->n = d->currNbParticles;
  h->L.quad = d->maxNbParticles <= kQuadDefaultMaxN ? 1 : 0;  // the default shape's walks
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(RTP_ERR_HIP, "hipStreamCreate failed");
  }
  const uint32_t maxN = d->maxNbParticles;
  if ((rc = h->grid.alloc(maxN)) || (rc = dev_alloc(&h->pos[0], maxN)) ||
      (rc = dev_alloc(&h->pos[1], maxN)) || (rc = dev_alloc(&h->vel[0], maxN)) ||
      (rc = dev_alloc(&h->vel[1], maxN)) || (rc = dev_alloc(&h->pred[0], maxN)) ||
      (rc = dev_alloc(&h->pred[1], maxN)) || (rc = dev_alloc(&h->vort, maxN)) ||
      (rc = dev_alloc(&h->velTmp, maxN)) || (rc = dev_alloc(&h->density, maxN)) ||
      (rc = dev_alloc(&h->lambda, maxN)) || (rc = dev_alloc(&h->metric, kMetricFloats)) ||
      (rc = dev_alloc(&h->feat, (size_t)maxN * 8)) ||
      (rc = dev_alloc(&h->lamAcc, maxN)) || (rc = dev_alloc(&h->resDensity, maxN)) ||
      (rc = dev_alloc(&h->resMetric, kMetricFloats)) || (rc = dev_alloc(&h->okey, maxN)) ||
      (rc = dev_alloc(&h->orank, maxN)) || (rc = dev_alloc(&h->oorder[0], maxN)) ||
      (rc = dev_alloc(&h->oorder[1], maxN)) ||
      (rc = dev_alloc(&h->ocount.table[0], order_count_words(h->grid.g))) ||
      (rc = dev_alloc(&h->ocount.table[1], order_count_words(h->grid.g))) ||
      (rc = dev_alloc(&h->opart, scan_partials_needed(order_count_words(h->grid.g))))) {
    h->release();
    delete h;
    return rc;
  }
  h->ocount.words = order_count_words(h->grid.g);
  if (hipMemset(h->ocount.table[0], 0, sizeof(uint32_t) * h->ocount.words) != hipSuccess ||
      hipMemset(h->ocount.table[1], 0, sizeof(uint32_t) * h->ocount.words) != hipSuccess) {
    h->release();
    delete h;
    return fail(RTP_ERR_HIP, "hipMemset failed");
  }
  h->refresh_consts();
  // Zero state and the reset key tail, like the constructor's reset() (Fluids.cpp:91).
  std::vector<float> zeros(4, 0.0f);
  if ((rc = rtp_fluids_load_state(h, zeros.data(), zeros.data(), 0))) {
    h->release();
    delete h;
    return rc;
  }
  *out = h;
  return RTP_OK;
}

int rtp_fluids_destroy(rtp_fluids* h)
{
  if (!h)
    return RTP_OK;
  (void)hipSetDevice(h->device);
  h->release();
  delete h;
  return RTP_OK;
}

int rtp_fluids_set_params(rtp_fluids* h, const rtp_fluid_params* p)
{
  RTP_REQUIRE(h && p, "NULL argument");
  RTP_REQUIRE(p->restDensity != 0.0f, "restDensity must be non-zero");
  RTP_REQUIRE(p->artPressureExp <= kMaxArtPressureExp, "artPressureExp must be at most 64");
  h->c.p = *p;
  h->refresh_consts();
  return RTP_OK;
}

int rtp_fluids_get_params(rtp_fluids* h, rtp_fluid_params* p)
{
  RTP_REQUIRE(h && p, "NULL argument");
  *p = h->c.p;
  return RTP_OK;
}

int rtp_fluids_set_nb_jacobi_iters(rtp_fluids* h, uint32_t iters)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(iters <= 1000, "iteration count out of range");
  h->iters = iters;
  h->drop_graph();
  return RTP_OK;
}

int rtp_fluids_get_nb_jacobi_iters(rtp_fluids* h, uint32_t* iters)
{
  RTP_REQUIRE(h && iters, "NULL argument");
  *iters = h->iters;
  return RTP_OK;
}

int rtp_fluids_set_nb_particles(rtp_fluids* h, uint32_t n)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(n <= h->grid.maxN, "nb particles exceeds maxNbParticles");
  // Keys in [n, old n) keep what fillCellIDs wrote; they stay in the sort like the
  // reference's whole-buffer sort keeps them (RadixSort.cpp:134-161).
  h->grid.dirtyEnd = std::max(h->grid.dirtyEnd, h->n);
  if (h->grid.dirtyEnd <= n)
    h->grid.dirtyEnd = std::max(h->grid.dirtyEnd, h->grid.sortedN);
  h->n = n;
  h->drop_graph();
  return RTP_OK;
}

int rtp_fluids_get_nb_particles(rtp_fluids* h, uint32_t* n)
{
  RTP_REQUIRE(h && n, "NULL argument");
  *n = h->n;
  return RTP_OK;
}

int rtp_fluids_load_state(rtp_fluids* h, const float* pos4, const float* vel4, uint32_t count)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(count <= h->grid.maxN, "count exceeds maxNbParticles");
  RTP_REQUIRE(count == 0 || (pos4 && vel4), "NULL state arrays");
  RTP_HIP_TRY(hipSetDevice(h->device));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  const float inf = INFINITY;
  const uint32_t maxN = h->grid.maxN;
  int rc = 0;
  for (int k = 0; k < 2; ++k) {
    if ((rc = upload_f4(h->pos[k], pos4, count, maxN, make_float4(inf, inf, inf, 0.0f), h->stream)) ||
        (rc = upload_f4(h->vel[k], vel4, count, maxN, make_float4(0.0f, 0.0f, 0.0f, 0.0f), h->stream)) ||
        (rc = upload_f4(h->pred[k], pos4, count, maxN, make_float4(inf, inf, inf, 0.0f), h->stream)))
      return rc;
  }
  RTP_HIP_TRY(hipMemsetAsync(h->density, 0, sizeof(float) * maxN, h->stream));
  RTP_HIP_TRY(hipMemsetAsync(h->lambda, 0, sizeof(float) * maxN, h->stream));
  RTP_HIP_TRY(hipMemsetAsync(h->resMetric, 0, sizeof(float) * 4, h->stream));
  h->cur = 0;
  h->predFinal = 1;
  h->grid.reset_keys(h->stream);
  if (h->render.on)
    h->render_reset();
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  h->drop_graph();
  return RTP_OK;
}

int rtp_fluids_read_state(rtp_fluids* h, float* pos4, float* vel4, uint32_t count)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(count <= h->grid.maxN, "count exceeds maxNbParticles");
  RTP_HIP_TRY(hipSetDevice(h->device));
  if (pos4)
    RTP_HIP_TRY(hipMemcpyAsync(pos4, h->pos[h->cur], sizeof(float4) * count, hipMemcpyDeviceToHost, h->stream));
  if (vel4)
    RTP_HIP_TRY(hipMemcpyAsync(vel4, h->vel[h->cur], sizeof(float4) * count, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_fluids_update(rtp_fluids* h, uint32_t nsteps)
{
  RTP_REQUIRE(h, "NULL handle");
  return h->update(nsteps);
}

int rtp_fluids_pause(rtp_fluids* h, int pause)
{
  RTP_REQUIRE(h, "NULL handle");
  h->paused = pause != 0;
  return RTP_OK;
}

int rtp_fluids_sync(rtp_fluids* h)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_fluids_read_grid(rtp_fluids* h, uint32_t* cellIDs, uint32_t* sortIdx, uint32_t* startEnd)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  return h->grid.read(cellIDs, sortIdx, startEnd, h->stream);
}

int rtp_fluids_read_field(rtp_fluids* h, int field, float* out, uint32_t count)
{
  RTP_REQUIRE(h && out, "NULL argument");
  RTP_REQUIRE(count <= h->grid.maxN || field == RTP_FIELD_DETECTOR, "count exceeds maxNbParticles");
  RTP_HIP_TRY(hipSetDevice(h->device));
  const void* src = nullptr;
  size_t elem = sizeof(float);
  switch (field) {
  case RTP_FIELD_DENSITY: src = h->density; break;
  case RTP_FIELD_LAMBDA: src = h->lambda; break;
  case RTP_FIELD_PRED_POS: src = h->pred[h->predFinal]; elem = sizeof(float4); break;
  case RTP_FIELD_VORTICITY: src = h->vort; elem = sizeof(float4); break;
  case RTP_FIELD_RESIDUAL_DENSITY: src = h->resDensity; break;
  case RTP_FIELD_FALLBACK_SET:
    if (!h->fb.act)
      return fail(RTP_ERR_INVALID, "no fallback set: the warm start has no fallback iterations");
    src = h->fb.act;
    break;
  case RTP_FIELD_COLOR:
  case RTP_FIELD_CAMERA_DIST:
  case RTP_FIELD_DETECTOR:
    return read_render_field(h->render, field, h->grid.g.numCells, out, count, h->stream);
  default: return fail(RTP_ERR_INVALID, "unknown field for fluids");
  }
  RTP_HIP_TRY(hipMemcpyAsync(out, src, elem * count, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_fluids_density_error(rtp_fluids* h, float out[3])
{
  RTP_REQUIRE(h && out, "NULL argument");
  RTP_HIP_TRY(hipSetDevice(h->device));
  launch_density_error(h->density, h->n, h->c.p.restDensity, h->metric, h->stream);
  float m[3] = { 0.0f, 0.0f, 0.0f };
  RTP_HIP_TRY(hipMemcpyAsync(m, h->metric, sizeof(m), hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  out[0] = m[0];
  out[1] = h->n ? m[1] / (float)h->n : 0.0f;
  out[2] = m[2];
  return RTP_OK;
}

int rtp_fluids_device_state(rtp_fluids* h, void** pos4, void** vel4)
{
  RTP_REQUIRE(h, "NULL handle");
  if (pos4)
    *pos4 = h->pos[h->cur];
  if (vel4)
    *vel4 = h->vel[h->cur];
  return RTP_OK;
}

int rtp_fluids_stream(rtp_fluids* h, void** stream)
{
  RTP_REQUIRE(h && stream, "NULL argument");
  *stream = h->stream;
  return RTP_OK;
}

int rtp_fluids_enable_profiling(rtp_fluids* h, int enable)
{
  RTP_REQUIRE(h, "NULL handle");
  h->prof.enabled = enable != 0;
  return RTP_OK;
}

int rtp_fluids_profile(rtp_fluids* h, rtp_kernel_time* out, int capacity, int* count)
{
  RTP_REQUIRE(h, "NULL handle");
  return h->prof.report(out, capacity, count);
}

int rtp_fluids_reset_profile(rtp_fluids* h)
{
  RTP_REQUIRE(h, "NULL handle");
  h->prof.reset();
  return RTP_OK;
}

int rtp_fluids_save_checkpoint(rtp_fluids* h)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  const uint32_t maxN = h->grid.maxN;
  int rc = 0;
  if (!h->ckPos && ((rc = dev_alloc(&h->ckPos, maxN)) || (rc = dev_alloc(&h->ckVel, maxN)) ||
                    (rc = dev_alloc(&h->ckPred, maxN)) || (rc = dev_alloc(&h->ckKey, maxN))))
    return rc;
  RTP_HIP_TRY(hipMemcpyAsync(h->ckPos, h->pos[h->cur], sizeof(float4) * maxN, hipMemcpyDeviceToDevice, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(h->ckVel, h->vel[h->cur], sizeof(float4) * maxN, hipMemcpyDeviceToDevice, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(h->ckPred, h->pred[h->predFinal], sizeof(float4) * maxN, hipMemcpyDeviceToDevice, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(h->ckKey, h->grid.key[h->grid.keyCur], sizeof(uint32_t) * maxN, hipMemcpyDeviceToDevice, h->stream));
  h->ckRender = h->render.on;
  if (h->ckRender) {
    if (!h->ckCol && ((rc = dev_alloc(&h->ckCol, maxN)) || (rc = dev_alloc(&h->ckCam, maxN))))
      return rc;
    RTP_HIP_TRY(hipMemcpyAsync(h->ckCol, h->render.col[h->render.colCur], sizeof(float4) * maxN, hipMemcpyDeviceToDevice, h->stream));
    RTP_HIP_TRY(hipMemcpyAsync(h->ckCam, h->render.key[h->render.keyCur], sizeof(uint32_t) * maxN, hipMemcpyDeviceToDevice, h->stream));
    h->ckHigh = h->render.high;
  }
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  h->ckN = h->n;
  h->ckDirty = h->grid.dirtyEnd;
  h->ckSorted = h->grid.sortedN;
  h->ckValid = true;
  return RTP_OK;
}

int rtp_fluids_restore_checkpoint(rtp_fluids* h)
{
  RTP_REQUIRE(h, "NULL handle");
  if (!h->ckValid)
    return fail(RTP_ERR_STATE, "no checkpoint saved");
  RTP_HIP_TRY(hipSetDevice(h->device));
  const uint32_t maxN = h->grid.maxN;
  // Outside the sorted range both ping-pong partners of a buffer hold the same data.  When
  // that range is unchanged since the save, the current buffers alone need the copy;
  // otherwise both partners get it.
  const bool same = h->grid.dirtyEnd == h->ckDirty && h->grid.sortedN == h->ckSorted && h->n == h->ckN;
  for (int k = 0; k < 2; ++k) {
    const bool current = k == 0;
    if (!current && same)
      break;
    float4* pos = current ? h->pos[h->cur] : h->pos[1 - h->cur];
    float4* vel = current ? h->vel[h->cur] : h->vel[1 - h->cur];
    float4* pred = current ? h->pred[h->predFinal] : h->pred[1 - h->predFinal];
    uint32_t* key = current ? h->grid.key[h->grid.keyCur] : h->grid.key[1 - h->grid.keyCur];
    RTP_HIP_TRY(hipMemcpyAsync(pos, h->ckPos, sizeof(float4) * maxN, hipMemcpyDeviceToDevice, h->stream));
    RTP_HIP_TRY(hipMemcpyAsync(vel, h->ckVel, sizeof(float4) * maxN, hipMemcpyDeviceToDevice, h->stream));
    RTP_HIP_TRY(hipMemcpyAsync(pred, h->ckPred, sizeof(float4) * maxN, hipMemcpyDeviceToDevice, h->stream));
    RTP_HIP_TRY(hipMemcpyAsync(key, h->ckKey, sizeof(uint32_t) * maxN, hipMemcpyDeviceToDevice, h->stream));
  }
  // p_col and p_cameraDist into both partners: their gather and sort ranges need not be
  // those of the save
  if (h->ckRender && h->render.on) {
    for (int k = 0; k < 2; ++k) {
      RTP_HIP_TRY(hipMemcpyAsync(h->render.col[k], h->ckCol, sizeof(float4) * maxN, hipMemcpyDeviceToDevice, h->stream));
      RTP_HIP_TRY(hipMemcpyAsync(h->render.key[k], h->ckCam, sizeof(uint32_t) * maxN, hipMemcpyDeviceToDevice, h->stream));
    }
    h->render.high = h->ckHigh;
  }
  h->n = h->ckN;
  h->grid.dirtyEnd = h->ckDirty;
  h->grid.sortedN = h->ckSorted;
  if (!same)
    h->drop_graph();
  return RTP_OK;
}

int rtp_fluids_use_graph(rtp_fluids* h, int enable)
{
  RTP_REQUIRE(h, "NULL handle");
  h->useGraph = enable != 0;
  h->drop_graph();
  return RTP_OK;
}

int rtp_fluids_set_render(rtp_fluids* h, int enable)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  h->drop_graph();
  if (!enable) {
    h->render.on = false;
    return RTP_OK;
  }
  if (h->render.on)
    return RTP_OK;
  if (!h->render.cam) {
    const int rc = h->render.alloc(h->grid.maxN, h->grid.g.numCells, true);
    if (rc) {
      h->render.release();
      h->render = RenderState();
      return rc;
    }
    const float origin[3] = { 0.0f, 0.0f, 0.0f };
    const int rcc = h->render.set_camera(origin, h->stream);
    if (rcc)
      return rcc;
  }
  h->render.on = true;
  h->render_reset();
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_fluids_update_with_json(rtp_fluids* h, const char* json)
{
  RTP_REQUIRE(h, "NULL handle");
  rtp_fluid_params p = h->c.p;
  uint32_t iters = 0;
  int rc = rtp_fluid_params_from_json(json, &p, &iters);
  if (rc)
    return rc;
  p.dim = h->c.p.dim;  // Fluids.cpp:234: from the model's dimension, not from the tree
  // the parameters first: a rejected set (rest density 0) leaves the model as it was
  if ((rc = rtp_fluids_set_params(h, &p)))
    return rc;
  return rtp_fluids_set_nb_jacobi_iters(h, iters > 1000u ? 1000u : iters);
}

int rtp_fluids_params_json(rtp_fluids* h, char* buf, size_t cap, size_t* len)
{
  RTP_REQUIRE(h, "NULL handle");
  return rtp_fluid_params_to_json(&h->c.p, h->iters, buf, cap, len);
}

int rtp_fluids_device_render(rtp_fluids* h, void** col4, void** cameraDist, void** detector)
{
  RTP_REQUIRE(h, "NULL handle");
  return render_pointers(h->render, col4, cameraDist, detector);
}

int rtp_fluids_set_camera(rtp_fluids* h, const float pos[3])
{
  RTP_REQUIRE(h && pos, "NULL argument");
  RTP_REQUIRE(h->render.cam, "enable the rendering steps before setting the camera");
  RTP_HIP_TRY(hipSetDevice(h->device));
  const int rc = h->render.set_camera(pos, h->stream);
  if (rc)
    return rc;
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_fluids_set_tuning(rtp_fluids* h, const rtp_tuning* t)
{
  RTP_REQUIRE(h && t, "NULL argument");
  int rc = check_tuning(*t, kFluidsLaunchFlags);
  if (rc)
    return rc;
  if ((t->flags & RTP_LAUNCH_NB_LIST) && !h->nbList) {
    RTP_HIP_TRY(hipSetDevice(h->device));
    if ((rc = dev_alloc(&h->nbList, (size_t)(kNbCap / 4u) * h->grid.maxN)) ||
        (rc = dev_alloc(&h->nbCount, h->grid.maxN))) {
      dev_free(h->nbList);
      h->nbList = nullptr;
      return rc;
    }
  }
  h->L.block = t->blockSize;
  h->L.xcdRemap = t->flags & RTP_LAUNCH_XCD_MAP;
  h->L.tile = (t->flags & RTP_LAUNCH_LDS_TILES) ? 1 : 0;
  h->L.resort = (t->flags & RTP_LAUNCH_RESORT) ? 1 : 0;
  h->L.nbList = (t->flags & RTP_LAUNCH_NB_LIST) ? 1 : 0;
  h->L.bricks = (t->flags & RTP_LAUNCH_BRICKS) ? 1 : 0;
  h->L.quad = (t->flags & RTP_LAUNCH_QUAD) ? 1 : 0;
  h->L.queue = t->queueDepth ? t->queueDepth : kDefaultQueue;
  h->grid.sortBits = t->sortBits;
  h->drop_graph();
  return RTP_OK;
}

rtp_tuning current_tuning(const rtp_fluids* h)
{
  rtp_tuning t{};
  t.blockSize = h->L.block;
  t.flags = (h->L.xcdRemap ? RTP_LAUNCH_XCD_MAP : 0) | (h->L.tile ? RTP_LAUNCH_LDS_TILES : 0) |
            (h->L.resort ? RTP_LAUNCH_RESORT : 0) | (h->L.nbList ? RTP_LAUNCH_NB_LIST : 0) |
            (h->L.bricks ? RTP_LAUNCH_BRICKS : 0) | (h->L.quad ? RTP_LAUNCH_QUAD : 0);
  t.sortBits = h->grid.sortBits;
  t.queueDepth = h->L.queue;
  return t;
}

int rtp_fluids_autotune(rtp_fluids* h, uint32_t trialSteps, rtp_tuning* best)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(trialSteps > 0 && trialSteps <= 1000, "trialSteps must be 1..1000");
  RTP_HIP_TRY(hipSetDevice(h->device));
  int rc = rtp_fluids_save_checkpoint(h);
  if (rc)
    return rc;
  const bool profWas = h->prof.enabled;
  h->prof.enabled = false;
  const bool pausedWas = h->paused;
  h->paused = false;
  const rtp_tuning start = current_tuning(h);
  auto run = [&](const rtp_tuning& t) { return rtp_fluids_set_tuning(h, &t); };
  auto steps = [&](uint32_t k) {
    int r = 0;
    for (uint32_t s = 0; s < k && !r; ++s)
      if (!(r = rtp_fluids_restore_checkpoint(h)))
        r = h->update(1);
    return r;
  };
  rtp_tuning top{};
  rc = tune_search(h->stream, trialSteps, start, true, run, steps, &top);
  h->prof.enabled = profWas;
  h->paused = pausedWas;
  if (rc)
    return rc;
  if ((rc = run(top)) || (rc = rtp_fluids_restore_checkpoint(h)))
    return rc;
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  if (best)
    *best = top;
  return RTP_OK;
}

// -------------------------------------------------------------------------------- boids

int rtp_boids_create(const rtp_model_desc* d, rtp_boids** out)
{
  RTP_REQUIRE(d && out, "NULL argument");
  *out = nullptr;
  RTP_REQUIRE(d->maxNbParticles > 0, "maxNbParticles must be positive");
  RTP_REQUIRE(d->currNbParticles <= d->maxNbParticles, "currNbParticles exceeds maxNbParticles");
  int rc = check_device(d->device);
  if (rc)
    return rc;
  auto* h = new rtp_boids();
  h->device = d->device;
  if ((rc = make_grid(d, 3000u, &h->grid.g))) {
    delete h;
    return rc;
  }
  h->c.g = h->grid.g;
  // Boids.cpp:104: 1.0f * box.x * box.x / (res.x * res.x)
  h->c.radiusSq = float_roundtrip(1.0f * d->boxSize[0] * d->boxSize[0] /
                                  (float)(d->gridRes[0] * d->gridRes[0]));
  h->c.p = rtp_boids_params{ 0.5f, 1.6f, 1.6f, 1.45f };  // Boids.cpp:44-69 defaults
  h->c.dt = 0.1f;                                        // Boids.cpp:334
  h->c.targetRadius = 2.0f;  // Target::Target, physics/utils/Target.cpp:11-21
  h->c.targetSign = 1.0f;
  h->c.shortDiv = h->c.radiusSq <= 0x1p30f ? 1 : 0;
  h->c.vmin = 0.2f * h->c.p.velocityScale;
  h->n = d->currNbParticles;
  h->L.quad = d->maxNbParticles <= kQuadDefaultMaxN ? 1 : 0;  // the default shape's walk
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(RTP_ERR_HIP, "hipStreamCreate failed");
  }
  const uint32_t maxN = d->maxNbParticles;
  if ((rc = h->grid.alloc(maxN)) || (rc = dev_alloc(&h->pos[0], maxN)) ||
      (rc = dev_alloc(&h->pos[1], maxN)) || (rc = dev_alloc(&h->vel[0], maxN)) ||
      (rc = dev_alloc(&h->vel[1], maxN)) || (rc = dev_alloc(&h->acc, maxN)) ||
      (rc = dev_alloc(&h->unitVel, maxN))) {
    h->release();
    delete h;
    return rc;
  }
  std::vector<float> zeros(4, 0.0f);
  if ((rc = rtp_boids_load_state(h, zeros.data(), zeros.data(), 0))) {
    h->release();
    delete h;
    return rc;
  }
  *out = h;
  return RTP_OK;
}

int rtp_boids_destroy(rtp_boids* h)
{
  if (!h)
    return RTP_OK;
  (void)hipSetDevice(h->device);
  h->release();
  delete h;
  return RTP_OK;
}

int rtp_boids_set_params(rtp_boids* h, const rtp_boids_params* p)
{
  RTP_REQUIRE(h && p, "NULL argument");
  h->c.p = *p;
  h->c.vmin = 0.2f * p->velocityScale;
  return RTP_OK;
}

int rtp_boids_set_time_step(rtp_boids* h, float dt)
{
  RTP_REQUIRE(h, "NULL handle");
  h->c.dt = dt;
  return RTP_OK;
}

int rtp_boids_set_nb_particles(rtp_boids* h, uint32_t n)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(n <= h->grid.maxN, "nb particles exceeds maxNbParticles");
  h->grid.dirtyEnd = std::max(h->grid.dirtyEnd, h->n);
  if (h->grid.dirtyEnd <= n)
    h->grid.dirtyEnd = std::max(h->grid.dirtyEnd, h->grid.sortedN);
  h->n = n;
  return RTP_OK;
}

int rtp_boids_load_state(rtp_boids* h, const float* pos4, const float* vel4, uint32_t count)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(count <= h->grid.maxN, "count exceeds maxNbParticles");
  RTP_REQUIRE(count == 0 || (pos4 && vel4), "NULL state arrays");
  RTP_HIP_TRY(hipSetDevice(h->device));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  const float inf = INFINITY;
  int rc = 0;
  for (int k = 0; k < 2; ++k) {
    // Boids.cpp:310-318: inactive boids sit at +inf and their velocity buffer gets the
    // same values as the positions.
    if ((rc = upload_f4(h->pos[k], pos4, count, h->grid.maxN, make_float4(inf, inf, inf, 0.0f), h->stream)) ||
        (rc = upload_f4(h->vel[k], vel4, count, h->grid.maxN, make_float4(inf, inf, inf, 0.0f), h->stream)))
      return rc;
  }
  RTP_HIP_TRY(hipMemsetAsync(h->acc, 0, sizeof(float4) * h->grid.maxN, h->stream));
  h->cur = 0;
  h->grid.reset_keys(h->stream);
  if (h->render.on)
    h->render_reset();
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_boids_read_state(rtp_boids* h, float* pos4, float* vel4, uint32_t count)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(count <= h->grid.maxN, "count exceeds maxNbParticles");
  RTP_HIP_TRY(hipSetDevice(h->device));
  if (pos4)
    RTP_HIP_TRY(hipMemcpyAsync(pos4, h->pos[h->cur], sizeof(float4) * count, hipMemcpyDeviceToHost, h->stream));
  if (vel4)
    RTP_HIP_TRY(hipMemcpyAsync(vel4, h->vel[h->cur], sizeof(float4) * count, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return h->check_persistent();
}

int rtp_boids_update(rtp_boids* h, uint32_t nsteps)
{
  RTP_REQUIRE(h, "NULL handle");
  if (h->paused) {
    // a paused update still sorts by camera distance (Boids.cpp:379-381); from the second
    // paused sort on the keys are in order and the permutation is the identity
    if (!h->render.on)
      return RTP_OK;
    RTP_HIP_TRY(hipSetDevice(h->device));
    for (uint32_t k = 0; k < std::min(nsteps, 2u); ++k) {
      const int rc = h->render_frame(false);
      if (rc)
        return rc;
    }
    RTP_HIP_TRY(hipGetLastError());
    return RTP_OK;
  }
  RTP_HIP_TRY(hipSetDevice(h->device));
  bool done = false;
  if (h->persistent && nsteps > 0) {
    const int rc = h->update_persistent(nsteps, &done);
    if (rc)
      return rc;
  }
  for (uint32_t k = 0; k < nsteps && !done; ++k) {
    const int rc = h->step();
    if (rc)
      return rc;
  }
  RTP_HIP_TRY(hipGetLastError());
  return RTP_OK;
}

int rtp_boids_pause(rtp_boids* h, int pause)
{
  RTP_REQUIRE(h, "NULL handle");
  h->paused = pause != 0;
  return RTP_OK;
}

int rtp_boids_sync(rtp_boids* h)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return h->check_persistent();
}

int rtp_boids_read_grid(rtp_boids* h, uint32_t* cellIDs, uint32_t* sortIdx, uint32_t* startEnd)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  return h->grid.read(cellIDs, sortIdx, startEnd, h->stream);
}

int rtp_boids_read_field(rtp_boids* h, int field, float* out, uint32_t count)
{
  RTP_REQUIRE(h && out, "NULL argument");
  RTP_REQUIRE(count <= h->grid.maxN || field == RTP_FIELD_DETECTOR, "count exceeds maxNbParticles");
  RTP_HIP_TRY(hipSetDevice(h->device));
  if (field == RTP_FIELD_COLOR || field == RTP_FIELD_CAMERA_DIST || field == RTP_FIELD_DETECTOR)
    return read_render_field(h->render, field, h->grid.g.numCells, out, count, h->stream);
  RTP_REQUIRE(field == RTP_FIELD_ACC, "unknown field for boids");
  RTP_HIP_TRY(hipMemcpyAsync(out, h->acc, sizeof(float4) * count, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_boids_set_tuning(rtp_boids* h, const rtp_tuning* t)
{
  RTP_REQUIRE(h && t, "NULL argument");
  const int rc = check_tuning(*t, RTP_LAUNCH_XCD_MAP | RTP_LAUNCH_PERSISTENT | RTP_LAUNCH_QUAD);
  if (rc)
    return rc;
  h->L.block = t->blockSize;
  h->L.xcdRemap = t->flags & RTP_LAUNCH_XCD_MAP;
  h->L.quad = (t->flags & RTP_LAUNCH_QUAD) ? 1 : 0;
  h->persistent = (t->flags & RTP_LAUNCH_PERSISTENT) != 0;
  return RTP_OK;
}

int rtp_boids_set_dimension(rtp_boids* h, int dim)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(dim == 2 || dim == 3, "dimension must be 2 or 3");
  h->c.dim2 = dim == 2 ? 1 : 0;
  return RTP_OK;
}

int rtp_boids_set_boundary(rtp_boids* h, int boundary)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(boundary == RTP_BOUNDARY_BOUNCING || boundary == RTP_BOUNDARY_CYCLIC,
              "unknown boundary");
  h->c.cyclic = boundary == RTP_BOUNDARY_CYCLIC ? 1 : 0;
  return RTP_OK;
}

int rtp_boids_set_target(rtp_boids* h, int active, const float pos[3], float radius, int sign)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(!active || pos, "NULL target position");
  h->c.targetOn = active ? 1 : 0;
  if (pos)
    for (int k = 0; k < 3; ++k)
      h->c.target[k] = pos[k];
  if (active) {
    h->c.targetRadius = radius;
    h->c.targetSign = (float)sign;
  }
  return RTP_OK;
}

int rtp_boids_get_target(rtp_boids* h, rtp_target_params* t, float pos[3])
{
  RTP_REQUIRE(h && t, "NULL argument");
  t->active = h->c.targetOn;
  t->visible = h->targetVisible ? 1 : 0;
  t->radius = h->c.targetRadius;
  t->sign = (int32_t)h->c.targetSign;
  if (pos)
    for (int k = 0; k < 3; ++k)
      pos[k] = h->c.target[k];
  return RTP_OK;
}

int rtp_boids_update_with_json(rtp_boids* h, const char* json)
{
  RTP_REQUIRE(h, "NULL handle");
  rtp_boids_params p{};
  rtp_target_params t{};
  int rc = rtp_boids_params_from_json(json, &p, &t);
  if (rc)
    return rc;
  if ((rc = rtp_boids_set_params(h, &p)))
    return rc;
  h->c.targetOn = t.active ? 1 : 0;
  h->c.targetRadius = t.radius;
  h->c.targetSign = (float)t.sign;
  h->targetVisible = t.visible != 0;
  return RTP_OK;
}

int rtp_boids_params_json(rtp_boids* h, char* buf, size_t cap, size_t* len)
{
  RTP_REQUIRE(h, "NULL handle");
  rtp_target_params t{};
  t.active = h->c.targetOn;
  t.visible = h->targetVisible ? 1 : 0;
  t.radius = h->c.targetRadius;
  t.sign = (int32_t)h->c.targetSign;
  return rtp_boids_params_to_json(&h->c.p, &t, buf, cap, len);
}

int rtp_boids_set_render(rtp_boids* h, int enable)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  if (!enable) {
    h->render.on = false;
    return RTP_OK;
  }
  if (h->render.on)
    return RTP_OK;
  if (!h->render.cam) {
    int rc = h->render.alloc(h->grid.maxN, h->grid.g.numCells, false);
    if (!rc && !h->accAlt)
      rc = dev_alloc(&h->accAlt, h->grid.maxN);
    if (rc) {
      h->render.release();
      h->render = RenderState();
      return rc;
    }
    const float origin[3] = { 0.0f, 0.0f, 0.0f };
    if ((rc = h->render.set_camera(origin, h->stream)))
      return rc;
  }
  h->render.on = true;
  h->render_reset();
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_boids_device_render(rtp_boids* h, void** col4, void** cameraDist, void** detector)
{
  RTP_REQUIRE(h, "NULL handle");
  return render_pointers(h->render, col4, cameraDist, detector);
}

int rtp_boids_set_camera(rtp_boids* h, const float pos[3])
{
  RTP_REQUIRE(h && pos, "NULL argument");
  RTP_REQUIRE(h->render.cam, "enable the rendering steps before setting the camera");
  RTP_HIP_TRY(hipSetDevice(h->device));
  const int rc = h->render.set_camera(pos, h->stream);
  if (rc)
    return rc;
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_boids_device_state(rtp_boids* h, void** pos4, void** vel4)
{
  RTP_REQUIRE(h, "NULL handle");
  if (pos4)
    *pos4 = h->pos[h->cur];
  if (vel4)
    *vel4 = h->vel[h->cur];
  return RTP_OK;
}

int rtp_boids_enable_profiling(rtp_boids* h, int enable)
{
  RTP_REQUIRE(h, "NULL handle");
  h->prof.enabled = enable != 0;
  return RTP_OK;
}

int rtp_boids_profile(rtp_boids* h, rtp_kernel_time* out, int capacity, int* count)
{
  RTP_REQUIRE(h, "NULL handle");
  return h->prof.report(out, capacity, count);
}

// ----------------------------------------------------------------- learned warm-start

static void mlp_refresh_dev(rtp_mlp* h)
{
  h->dev.p = h->p;
  h->dev.w1t = h->w1t;
  h->dev.w2t = h->w2t;
  h->dev.w2 = h->w2;
  h->dev.w1tMut = h->w1t;
  h->dev.w2tMut = h->w2t;
  h->dev.w2Mut = h->w2;
}

int rtp_mlp_create(int device, uint32_t seed, rtp_mlp** out)
{
  RTP_REQUIRE(out, "NULL argument");
  *out = nullptr;
  int rc = check_device(device);
  if (rc)
    return rc;
  auto* h = new rtp_mlp();
  h->device = device;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(RTP_ERR_HIP, "hipStreamCreate failed");
  }
  if ((rc = dev_alloc(&h->p, kMlpParams)) || (rc = dev_alloc(&h->grad, kMlpParams)) ||
      (rc = dev_alloc(&h->m1, kMlpParams)) || (rc = dev_alloc(&h->m2, kMlpParams)) ||
      (rc = dev_alloc(&h->loss, 4)) || (rc = dev_alloc(&h->w1t, 64 * 16)) ||
      (rc = dev_alloc(&h->w2t, 64 * 64)) || (rc = dev_alloc(&h->w2, 64 * 64))) {
    h->release();
    delete h;
    return rc;
  }
  mlp_refresh_dev(h);
  for (int k = 0; k < 8; ++k) {
    h->dev.mu[k] = 0.0f;
    h->dev.invSigma[k] = 1.0f;
  }
  h->dev.outScale = 1.0f;
  // Deterministic initial weights: N(0, 1/fan_in) from SplitMix64 + Box-Muller, zero biases.
  std::vector<float> p(kMlpParams, 0.0f);
  uint64_t st = 0x9E3779B97F4A7C15ull * (seed + 1ull);
  auto next = [&]() {
    uint64_t z = (st += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    z ^= z >> 31;
    return ((z >> 11) + 0.5) * (1.0 / 9007199254740992.0);
  };
  auto normal = [&]() { return std::sqrt(-2.0 * std::log(next())) * std::cos(6.283185307179586 * next()); };
  for (int i = 0; i < 8 * 64; ++i)
    p[kMlpW1 + i] = (float)(normal() / std::sqrt(8.0));
  for (int i = 0; i < 64 * 64; ++i)
    p[kMlpW2 + i] = (float)(normal() / std::sqrt(64.0));
  for (int i = 0; i < 64; ++i)
    p[kMlpW3 + i] = (float)(normal() / std::sqrt(64.0));
  if ((rc = rtp_mlp_set_params(h, p.data(), kMlpParams))) {
    h->release();
    delete h;
    return rc;
  }
  *out = h;
  return RTP_OK;
}

int rtp_mlp_destroy(rtp_mlp* h)
{
  if (!h)
    return RTP_OK;
  (void)hipSetDevice(h->device);
  h->release();
  delete h;
  return RTP_OK;
}

int rtp_mlp_param_count(void) { return kMlpParams; }

int rtp_mlp_set_params(rtp_mlp* h, const float* params, uint32_t count)
{
  RTP_REQUIRE(h && params, "NULL argument");
  RTP_REQUIRE(count == (uint32_t)kMlpParams, "parameter count must be rtp_mlp_param_count()");
  RTP_HIP_TRY(hipSetDevice(h->device));
  RTP_HIP_TRY(hipMemcpyAsync(h->p, params, sizeof(float) * kMlpParams, hipMemcpyHostToDevice, h->stream));
  launch_zero(h->grad, kMlpParams, h->stream);
  launch_zero(h->m1, kMlpParams, h->stream);
  launch_zero(h->m2, kMlpParams, h->stream);
  launch_mlp_pack(h->p, h->dev, h->stream);
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  h->step = 0;
  return RTP_OK;
}

int rtp_mlp_get_params(rtp_mlp* h, float* params, uint32_t count)
{
  RTP_REQUIRE(h && params, "NULL argument");
  RTP_REQUIRE(count == (uint32_t)kMlpParams, "parameter count must be rtp_mlp_param_count()");
  RTP_HIP_TRY(hipSetDevice(h->device));
  RTP_HIP_TRY(hipMemcpyAsync(params, h->p, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_mlp_set_normalization(rtp_mlp* h, const float* mu, const float* sigma, float outScale)
{
  RTP_REQUIRE(h && mu && sigma, "NULL argument");
  RTP_REQUIRE(outScale != 0.0f, "outScale must be non-zero");
  for (int k = 0; k < 8; ++k) {
    RTP_REQUIRE(sigma[k] > 0.0f, "sigma must be positive");
    h->dev.mu[k] = mu[k];
    h->dev.invSigma[k] = 1.0f / sigma[k];
  }
  h->dev.outScale = outScale;
  return RTP_OK;
}

int rtp_mlp_forward(rtp_mlp* h, const float* feat, uint32_t n, float* out)
{
  RTP_REQUIRE(h && (n == 0 || (feat && out)), "NULL argument");
  if (n == 0)
    return RTP_OK;
  RTP_HIP_TRY(hipSetDevice(h->device));
  int rc = h->reserve(n);
  if (rc)
    return rc;
  RTP_HIP_TRY(hipMemcpyAsync(h->dFeat, feat, sizeof(float) * 8 * n, hipMemcpyHostToDevice, h->stream));
  launch_mlp_forward(h->dFeat, n, h->dev, h->dOut, h->stream);
  RTP_HIP_TRY(hipGetLastError());
  RTP_HIP_TRY(hipMemcpyAsync(out, h->dOut, sizeof(float) * n, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_mlp_gradients(rtp_mlp* h, const float* feat, const float* target, uint32_t n, float* grad,
                      float* loss)
{
  RTP_REQUIRE(h && feat && target && grad && loss && n > 0, "NULL argument or empty batch");
  RTP_HIP_TRY(hipSetDevice(h->device));
  int rc = h->reserve(n);
  if (rc)
    return rc;
  RTP_HIP_TRY(hipMemcpyAsync(h->dFeat, feat, sizeof(float) * 8 * n, hipMemcpyHostToDevice, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(h->dTarget, target, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
  launch_zero(h->grad, kMlpParams, h->stream);
  launch_zero(h->loss, 1, h->stream);
  launch_mlp_train(h->dFeat, h->dTarget, nullptr, n, h->dev, h->grad, h->loss, h->stream);
  RTP_HIP_TRY(hipGetLastError());
  RTP_HIP_TRY(hipMemcpyAsync(grad, h->grad, sizeof(float) * kMlpParams, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(loss, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
  launch_zero(h->grad, kMlpParams, h->stream);
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  *loss /= (float)n;
  return RTP_OK;
}

static uint64_t gcd64(uint64_t a, uint64_t b)
{
  while (b) {
    const uint64_t t = a % b;
    a = b;
    b = t;
  }
  return a;
}

int rtp_mlp_train(rtp_mlp* h, const float* feat, const float* target, uint32_t n, uint32_t epochs,
                  uint32_t batch, float lr, uint32_t seed, float* epochLoss)
{
  RTP_REQUIRE(h && feat && target && n > 0 && batch > 0, "NULL argument or empty data");
  RTP_HIP_TRY(hipSetDevice(h->device));
  int rc = h->reserve(n);
  if (rc)
    return rc;
  RTP_HIP_TRY(hipMemcpyAsync(h->dFeat, feat, sizeof(float) * 8 * n, hipMemcpyHostToDevice, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(h->dTarget, target, sizeof(float) * n, hipMemcpyHostToDevice, h->stream));
  uint64_t st = 0xD1B54A32D192ED03ull * (seed + 1ull);
  for (uint32_t ep = 0; ep < epochs; ++ep) {
    // a fresh affine permutation i -> (a i + c) mod n per epoch (a coprime with n)
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    uint64_t a = (st >> 33) % n;
    if (a == 0)
      a = 1;
    while (gcd64(a, n) != 1)
      a = (a + 1) % n == 0 ? 1 : (a + 1);
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    const uint32_t cc = (uint32_t)((st >> 33) % n);
    launch_affine_perm(h->dPerm, n, (uint32_t)a, cc, h->stream);
    launch_zero(h->loss, 1, h->stream);
    for (uint32_t b0 = 0; b0 < n; b0 += batch) {
      const uint32_t rows = std::min(batch, n - b0);
      launch_mlp_train(h->dFeat, h->dTarget, h->dPerm + b0, rows, h->dev, h->grad, h->loss, h->stream);
      launch_mlp_adam(h->p, h->grad, h->m1, h->m2, lr, ++h->step, h->dev, h->stream);
      launch_mlp_pack(h->p, h->dev, h->stream);
    }
    RTP_HIP_TRY(hipGetLastError());
    if (epochLoss) {
      float l = 0.0f;
      RTP_HIP_TRY(hipMemcpyAsync(&l, h->loss, sizeof(float), hipMemcpyDeviceToHost, h->stream));
      RTP_HIP_TRY(hipStreamSynchronize(h->stream));
      epochLoss[ep] = l / (float)n;
    }
  }
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_fluids_set_warm_start(rtp_fluids* h, rtp_mlp* mlp, uint32_t learnedIters,
                              uint32_t fallbackIters, float tau)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(!mlp || learnedIters > 0, "learnedIters must be positive with a network");
  RTP_REQUIRE(learnedIters <= 1000 && fallbackIters <= 1000, "iteration count out of range");
  if (mlp && mlp->device != h->device)
    return fail(RTP_ERR_INVALID, "network and model live on different devices");
  RTP_HIP_TRY(hipSetDevice(h->device));
  if (mlp && fallbackIters > 0 && !h->fb.act) {
    const uint32_t maxN = h->grid.maxN;
    FallbackSet& fb = h->fb;
    int rc = 0;
    if ((rc = dev_alloc(&fb.act, maxN)) || (rc = dev_alloc(&fb.seeds, maxN)) ||
        (rc = dev_alloc(&fb.list, maxN)) || (rc = dev_alloc(&fb.halo, maxN)) ||
        (rc = dev_alloc(&fb.count, 4))) {
      dev_free(fb.act);
      dev_free(fb.seeds);
      dev_free(fb.list);
      dev_free(fb.halo);
      dev_free(fb.count);
      fb = FallbackSet();
      return rc;
    }
    launch_fill_u32(fb.count, 4, 0u, h->stream);
  }
  h->mlp = mlp;
  h->learnedIters = mlp ? learnedIters : 0;
  h->fallbackIters = mlp ? fallbackIters : 0;
  h->tau = tau;
  h->drop_graph();
  return RTP_OK;
}

int rtp_fluids_fallback_count(rtp_fluids* h, uint32_t* seeds, uint32_t* particles, uint32_t* halo)
{
  RTP_REQUIRE(h && seeds && particles, "NULL argument");
  uint32_t v[3] = { 0u, 0u, 0u };
  if (h->fb.count) {
    RTP_HIP_TRY(hipSetDevice(h->device));
    RTP_HIP_TRY(hipMemcpyAsync(v, h->fb.count, sizeof(v), hipMemcpyDeviceToHost, h->stream));
    RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  }
  *seeds = v[0];
  *particles = v[1];
  if (halo)
    *halo = v[2];
  return RTP_OK;
}

int rtp_fluids_set_residual_pass(rtp_fluids* h, int enable)
{
  RTP_REQUIRE(h, "NULL handle");
  h->residualPass = enable != 0;
  h->drop_graph();
  return RTP_OK;
}

int rtp_fluids_residual_error(rtp_fluids* h, float out[3])
{
  RTP_REQUIRE(h && out, "NULL argument");
  RTP_HIP_TRY(hipSetDevice(h->device));
  float m[3] = { 0.0f, 0.0f, 0.0f };
  RTP_HIP_TRY(hipMemcpyAsync(m, h->resMetric, sizeof(m), hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  out[0] = m[0];
  out[1] = h->n ? m[1] / (float)h->n : 0.0f;
  out[2] = m[2];
  return RTP_OK;
}

int rtp_fluids_record_teacher(rtp_fluids* h, uint32_t teacherIters, float* feat, float* target)
{
  RTP_REQUIRE(h && feat && target, "NULL argument");
  RTP_REQUIRE(teacherIters > 0 && teacherIters <= 1000, "teacherIters out of range");
  RTP_HIP_TRY(hipSetDevice(h->device));
  if (h->paused)
    return fail(RTP_ERR_STATE, "model is paused");
  h->teacherIters = teacherIters;
  const int rc = h->substep();
  h->teacherIters = 0;
  if (rc)
    return rc;
  RTP_HIP_TRY(hipGetLastError());
  RTP_HIP_TRY(hipMemcpyAsync(feat, h->feat, sizeof(float) * 8 * h->n, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(target, h->lamAcc, sizeof(float) * h->n, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

// ------------------------------------------------------------------------ building blocks

int rtp_sort_pairs(int device, const uint32_t* keysIn, uint32_t n, uint32_t maxKey,
                   int32_t digitBits, uint32_t* keysOut, uint32_t* permOut)
{
  RTP_REQUIRE(n == 0 || (keysIn && keysOut && permOut), "NULL argument");
  RTP_REQUIRE(digitBits == 0 || (digitBits >= 4 && digitBits <= 11), "digitBits must be 0 or 4..11");
  if (n == 0)
    return RTP_OK;
  int rc = check_device(device);
  if (rc)
    return rc;
  hipStream_t s = nullptr;
  RTP_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  uint32_t *ka = nullptr, *kb = nullptr, *ia = nullptr, *ib = nullptr, *cnt = nullptr;
  const SortPlan plan = make_sort_plan(n, maxKey, digitBits);
  bool inB = false;
  rc = dev_alloc(&ka, n);
  if (!rc) rc = dev_alloc(&kb, n);
  if (!rc) rc = dev_alloc(&ia, n);
  if (!rc) rc = dev_alloc(&ib, n);
  if (!rc) rc = dev_alloc(&cnt, count_capacity(n));
  if (!rc && hipMemcpyAsync(ka, keysIn, sizeof(uint32_t) * n, hipMemcpyHostToDevice, s) != hipSuccess)
    rc = fail(RTP_ERR_HIP, "hipMemcpy keys");
  if (!rc)
    rc = sort_pairs(plan, ka, kb, ia, ib, cnt, &inB, s);
  if (!rc && hipGetLastError() != hipSuccess)
    rc = fail(RTP_ERR_HIP, "sort launch failed");
  if (!rc && (hipMemcpyAsync(keysOut, inB ? kb : ka, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
              hipMemcpyAsync(permOut, inB ? ib : ia, sizeof(uint32_t) * n, hipMemcpyDeviceToHost, s) != hipSuccess))
    rc = fail(RTP_ERR_HIP, "hipMemcpy results");
  if (!rc && hipStreamSynchronize(s) != hipSuccess)
    rc = fail(RTP_ERR_HIP, "sort failed");
  (void)hipStreamSynchronize(s);
  dev_free(ka);
  dev_free(kb);
  dev_free(ia);
  dev_free(ib);
  dev_free(cnt);
  (void)hipStreamDestroy(s);
  return rc;
}

int rtp_grid_build(const rtp_model_desc* d, const float* pos4, uint32_t* keys, uint32_t* idx,
                   uint32_t* startEnd)
{
  RTP_REQUIRE(d && (d->currNbParticles == 0 || pos4), "NULL argument");
  RTP_REQUIRE(d->maxNbParticles > 0 && d->currNbParticles <= d->maxNbParticles,
              "particle counts out of range");
  int rc = check_device(d->device);
  if (rc)
    return rc;
  GridState G;
  if ((rc = make_grid(d, 100u, &G.g)))
    return rc;
  const uint32_t n = d->currNbParticles, maxN = d->maxNbParticles;
  hipStream_t s = nullptr;
  RTP_HIP_TRY(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  float4* pos = nullptr;
  Profiler prof;
  rc = G.alloc(maxN);
  if (!rc)
    rc = dev_alloc(&pos, maxN);
  if (!rc)
    rc = upload_f4(pos, pos4, n, maxN, make_float4(INFINITY, INFINITY, INFINITY, 0.0f), s);
  if (!rc) {
    G.reset_keys(s);
    launch_fill_keys_and_reset_cells(pos, n, G.key[G.keyCur], G.startEnd, G.g, s);
    rc = G.sort(n, s, prof);
  }
  if (!rc) {
    ReorderArgs a{};
    a.key = G.key[G.keyCur];
    a.idx = G.idx[G.idxCur];
    a.sortN = G.sortedN;
    a.n = n;
    a.maxN = maxN;
    a.startEnd = G.startEnd;
    a.g = G.g;
    launch_reorder(a, s);
    if (hipGetLastError() != hipSuccess)
      rc = fail(RTP_ERR_HIP, "grid launch failed");
  }
  if (!rc)
    rc = G.read(keys, idx, startEnd, s);
  (void)hipStreamSynchronize(s);
  G.release();
  dev_free(pos);
  prof.destroy();
  (void)hipStreamDestroy(s);
  return rc;
}

int rtp_grid_build_noalloc(const rtp_model_desc* d, const float* pos4, uint32_t* keys,
                           uint32_t* idx, uint32_t* startEnd, void* stream)
{
  RTP_REQUIRE(d && keys && idx && startEnd && (d->currNbParticles == 0 || pos4), "NULL argument");
  RTP_REQUIRE(d->maxNbParticles > 0 && d->currNbParticles <= d->maxNbParticles,
              "particle counts out of range");
  if (d->maxNbParticles > kNoAllocMax)
    return fail(RTP_ERR_INVALID, "rtp_grid_build_noalloc: maxNbParticles " +
                                     std::to_string(d->maxNbParticles) + " exceeds the library's " +
                                     std::to_string(kNoAllocMax) + "-slot scratch array");
  int rc = check_device(d->device);
  if (rc)
    return rc;
  GridConsts g{};
  if ((rc = make_grid(d, 100u, &g)))
    return rc;
  uint32_t* const base = scratch_on_current_device();
  if (!base)
    return RTP_ERR_HIP;
  const uint32_t n = d->currNbParticles, maxN = d->maxNbParticles;
  hipStream_t s = static_cast<hipStream_t>(stream);
  uint32_t* const keyScratch = base;
  uint32_t* const idxScratch = base + kNoAllocMax;
  uint32_t* const counts = base + 2 * (size_t)kNoAllocMax;

  // resetCellIDs over maxN, the identity permutation past the sorted range, then the keys of
  // the n positions into whichever buffer the sort starts from, so that its result lands in
  // the caller's keys and idx.  No scratch word needs to start at zero: every count a sort
  // pass reads is written earlier in that pass, and the partners are read only where a pass
  // has written them.
  const SortPlan plan = make_sort_plan(n, max_active_key(g), 0, kSortItems);
  const bool inB = sort_result_in_b(plan);
  uint32_t* const keyA = inB ? keyScratch : keys;
  uint32_t* const keyB = inB ? keys : keyScratch;
  uint32_t* const idxA = inB ? idxScratch : idx;
  uint32_t* const idxB = inB ? idx : idxScratch;
  launch_reset_cell_ids(keys, maxN, g.numCells, s);
  launch_fill_identity(idx, n, maxN, s);
  launch_fill_keys_and_reset_cells(reinterpret_cast<const float4*>(pos4), n, keyA,
                                   reinterpret_cast<uint2*>(startEnd), g, s);
  bool landedInB = false;
  if ((rc = sort_pairs(plan, keyA, keyB, idxA, idxB, counts, &landedInB, s)))
    return rc;
  if (landedInB != inB)
    return fail(RTP_ERR_STATE, "rtp_grid_build_noalloc: sort result in an unexpected buffer");
  if (n) {
    ReorderArgs a{};
    a.key = keys;
    a.idx = idx;
    a.sortN = n;
    a.n = n;
    a.maxN = maxN;
    a.startEnd = reinterpret_cast<uint2*>(startEnd);
    a.g = g;
    launch_reorder(a, s);
  }
  if (hipGetLastError() != hipSuccess)
    return fail(RTP_ERR_HIP, "grid launch failed");
  return RTP_OK;
}

}  // extern "C"
