// Batched position based fluids: many independent scenes advanced by shared launches.
//
// Scene s owns particles [pbase_s, pbase_s + n_s) of every particle array and the key range
// [kbase_s, kbase_s + keySpan_s) of the cell keys and of the cell table, keySpan_s being the
// largest key its hash can produce + 1.  One stable sort of all keys therefore leaves each
// scene's particles in its own slice, ordered exactly as a sort of that scene alone would.
// Every pass of a substep is one launch whose blocks each serve one scene (a block table of
// (scene, first particle) pairs), with per-scene constants (grid, box, gravity, parameters)
// read from a device array.  The per-particle work is the single-model bodies of pbf.hip
// and grid_sort.hip on scene-local indices, so every scene's state is bit-identical to a
// model of its own.
#include "rtp_host.hpp"

namespace rtp {
namespace {

struct SceneHost {
  rtp_batch_scene desc;
  PbfConsts c;
  SceneSlot slot;
};

}  // namespace
}  // namespace rtp

using namespace rtp;

struct rtp_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  std::vector<SceneHost> scenes;
  uint32_t total = 0;     // particles of all scenes
  uint32_t keySpace = 0;  // cell keys (and cell-table entries) of all scenes
  uint32_t iters = 2;     // Fluids.cpp m_nbJacobiIters default
  float4* pos[2] = { nullptr, nullptr };
  float4* vel[2] = { nullptr, nullptr };
  float4* pred[2] = { nullptr, nullptr };
  float4* vort = nullptr;
  float4* velTmp = nullptr;
  float* density = nullptr;
  float* lambda = nullptr;
  uint32_t* key[2] = { nullptr, nullptr };
  uint32_t* idx[2] = { nullptr, nullptr };
  uint32_t* counts = nullptr;
  uint2* startEnd = nullptr;
  SceneSlot* dSlots = nullptr;
  PbfConsts* dConsts = nullptr;
  uint2* dBlocks = nullptr;
  uint32_t numBlocks = 0;
  int block = 256;
  int sortBits = 0;
  int queue = kDefaultQueue;
  bool resort = false;  // order the neighbour passes' threads by current cell
  uint32_t* okey = nullptr;
  uint32_t* orank = nullptr;
  uint32_t* oorder[2] = { nullptr, nullptr };  // thread orders, each counted in the other
  CountTables ocount;  // the thread order's count tables
  uint32_t* opart = nullptr;
  // survivor lists (RTP_LAUNCH_NB_LIST), allocated when first enabled, for nbThreads launch
  // threads (every block size's launch fits)
  bool lists = false;
  uint32_t nbThreads = 0;
  uint4* nbList = nullptr;
  uint32_t* nbCount = nullptr;
  bool constsDirty = true;
  int cur = 0;
  int predFinal = 1;
  int keyRes = 0, idxRes = 0;
  bool stepped = false;
  float4* ckPos = nullptr;
  float4* ckVel = nullptr;
  bool ckValid = false;
  Profiler prof;
  bool useGraph = true;
  bool graphValid = false;
  int graphCur = -1;
  int graphCount = -1;  // ocount.cur at the capture
  hipGraph_t graph = nullptr;
  hipGraphExec_t graphExec = nullptr;

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

  void release()
  {
    drop_graph();
    prof.destroy();
    for (int k = 0; k < 2; ++k) {
      dev_free(pos[k]);
      dev_free(vel[k]);
      dev_free(pred[k]);
      dev_free(key[k]);
      dev_free(idx[k]);
    }
    dev_free(vort);
    dev_free(velTmp);
    dev_free(density);
    dev_free(lambda);
    dev_free(counts);
    dev_free(startEnd);
    dev_free(dSlots);
    dev_free(dConsts);
    dev_free(dBlocks);
    dev_free(ckPos);
    dev_free(ckVel);
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

  BatchView view() const
  {
    BatchView bv;
    bv.scenes = dSlots;
    bv.blocks = dBlocks;
    bv.consts = dConsts;
    bv.numBlocks = numBlocks;
    bv.block = (uint32_t)block;
    bv.qdepth = (uint32_t)queue;
    return bv;
  }

  bool vorticity_on() const { return scenes.empty() ? false : scenes[0].c.p.isVorticityConfEnabled != 0; }

  // Block table for the current block size: each scene's particles in blocks of `block`.
  int build_blocks()
  {
    std::vector<uint2> t;
    for (uint32_t s = 0; s < (uint32_t)scenes.size(); ++s)
      for (uint32_t first = 0; first < scenes[s].slot.n; first += (uint32_t)block)
        t.push_back(make_uint2(s, first));
    numBlocks = (uint32_t)t.size();
    if (numBlocks)
      RTP_HIP_TRY(hipMemcpyAsync(dBlocks, t.data(), sizeof(uint2) * t.size(), hipMemcpyHostToDevice, stream));
    RTP_HIP_TRY(hipStreamSynchronize(stream));
    drop_graph();
    return RTP_OK;
  }

  int upload_consts()
  {
    if (!constsDirty)
      return RTP_OK;
    std::vector<PbfConsts> c(scenes.size());
    for (size_t s = 0; s < scenes.size(); ++s)
      c[s] = scenes[s].c;
    RTP_HIP_TRY(hipMemcpyAsync(dConsts, c.data(), sizeof(PbfConsts) * c.size(), hipMemcpyHostToDevice, stream));
    RTP_HIP_TRY(hipStreamSynchronize(stream));
    constsDirty = false;
    return RTP_OK;
  }

  // Fluids::update (Fluids.cpp:400-471) for every scene at once.
  int substep()
  {
    hipStream_t s = stream;
    const BatchView bv = view();
    const bool vortOn = vorticity_on();
    prof.begin("pbf_predict_hash", s);
    launch_fill_start_end(startEnd, keySpace, s);
    launch_pbf_predict_batch(bv, pos[cur], vel[cur], pred[0], key[0], s);
    prof.end(s);
    const SortPlan plan = make_sort_plan(total, keySpace - 1u, sortBits);
    bool inB = false;
    int rc = 0;
    RTP_PROF(prof, "radix_sort", s,
             rc = sort_pairs(plan, key[0], key[1], idx[0], idx[1], counts, &inB, s));
    if (rc)
      return rc;
    keyRes = idxRes = inB ? 1 : 0;
    ReorderArgs a{};
    a.key = key[keyRes];
    a.idx = idx[idxRes];
    a.srcA = pos[cur];
    a.dstA = pos[1 - cur];
    a.srcB = vel[cur];
    a.dstB = vel[1 - cur];
    a.srcC = pred[0];
    a.dstC = pred[1];
    a.clampC = iters > 0;
    a.startEnd = startEnd;
    RTP_PROF(prof, "reorder_cells", s, launch_reorder_batch(bv, a, s));
    cur = 1 - cur;
    int in = 1;
    // neighbour passes: threads in the grid's order first, then (resort) by current cell;
    // with lists the density / vorticity passes record survivors and the correction,
    // confinement and XSPH passes replay them
    BatchView nb = bv;
    if (lists)
      nb.nb = NbList{ nbList, nbCount, bv.numBlocks * bv.block };
    BatchView cb = nb;
    cb.keyOut = resort ? okey : nullptr;
    const int rec = lists ? kNbRecord : 0, rep = lists ? kNbReplay : 0;
    const uint32_t m16 = (keySpace + 15u) & ~15u;
    for (uint32_t k = 0; k < iters; ++k) {
      const int last = (k + 1 == iters) ? kLastVel : kNextIter;
      RTP_PROF(prof, "pbf_density_lambda", s,
               launch_pbf_density_lambda_batch(nb, pred[in], startEnd, density, lambda, s, rec));
      RTP_PROF(prof, "pbf_correct", s,
               launch_pbf_correct_batch(cb, pred[in], lambda, startEnd, pred[1 - in], pos[cur], vel[cur],
                                        (last == kLastVel && !vortOn) ? pos[cur] : nullptr, last, s,
                                        rep));
      in = 1 - in;
      if (resort) {
        uint32_t* next = nb.order == oorder[0] ? oorder[1] : oorder[0];
        bool ran = false;
        RTP_PROF(prof, "thread_order", s,
                 ran = launch_counting_order(ocount, m16, okey, nb.order, orank, total, opart, next, s));
        if (ran)
          ocount.cur = 1 - ocount.cur;
        nb.order = next;
        cb.order = next;
      }
    }
    if (iters == 0)
      RTP_PROF(prof, "pbf_update_vel", s,
               launch_pbf_update_vel_batch(bv, pred[in], pos[cur], vel[cur], vortOn ? nullptr : pos[cur], s));
    predFinal = in;
    if (vortOn) {
      RTP_PROF(prof, "pbf_vorticity", s,
               launch_pbf_vorticity_batch(nb, pred[in], startEnd, vel[cur], vort, s, rec));
      RTP_PROF(prof, "pbf_confinement", s,
               launch_pbf_confinement_batch(nb, pred[in], startEnd, vort, vel[cur], velTmp, s, rep));
      RTP_PROF(prof, "pbf_xsph", s,
               launch_pbf_xsph_batch(nb, pred[in], startEnd, velTmp, vel[cur], pos[cur], s, rep));
    }
    stepped = true;
    return RTP_OK;
  }

  int capture_pair()
  {
    drop_graph();
    graphCur = cur;
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
    return RTP_OK;
  }

  int update(uint32_t nsteps)
  {
    if (nsteps == 0 || total == 0)
      return RTP_OK;
    MarkRange mark("rtp_batch_update");
    RTP_HIP_TRY(hipSetDevice(device));
    int rc = upload_consts();
    if (rc)
      return rc;
    uint32_t done = 0;
    if (useGraph && !prof.enabled && nsteps >= 2) {
      if (!graphValid || graphCur != cur || graphCount != ocount.cur) {
        if ((rc = capture_pair()))
          return rc;
      }
      for (; done + 2 <= nsteps; done += 2) {
        MarkRange pair("rtp_batch_substep_pair_graph");
        RTP_HIP_TRY(hipGraphLaunch(graphExec, stream));
      }
      stepped = true;
    }
    for (; done < nsteps; ++done) {
      MarkRange one("rtp_batch_substep");
      if ((rc = substep()))
        return rc;
    }
    return RTP_OK;
  }

  int check_scene(uint32_t s) const
  {
    if (s >= scenes.size())
      return fail(RTP_ERR_INVALID, "scene index out of range");
    return RTP_OK;
  }
};

namespace {

int set_params_one(rtp_batch* h, uint32_t s, const rtp_fluid_params* p)
{
  h->scenes[s].c.p = *p;
  refresh_pbf_consts(h->scenes[s].c);
  return RTP_OK;
}

int read_f4_slice(rtp_batch* h, const float4* src, uint32_t s, float* out)
{
  const SceneSlot& sl = h->scenes[s].slot;
  RTP_HIP_TRY(hipMemcpyAsync(out, src + sl.pbase, sizeof(float4) * sl.n, hipMemcpyDeviceToHost, h->stream));
  return RTP_OK;
}

}  // namespace

extern "C" {

int rtp_batch_create(int device, const rtp_batch_scene* sc, uint32_t count, rtp_batch** out)
{
  RTP_REQUIRE(sc && out, "NULL argument");
  *out = nullptr;
  RTP_REQUIRE(count > 0, "a batch needs at least one scene");
  int rc = check_device(device);
  if (rc)
    return rc;
  auto* h = new rtp_batch();
  h->device = device;
  uint64_t total = 0, keys = 0;
  h->scenes.resize(count);
  for (uint32_t s = 0; s < count; ++s) {
    SceneHost& S = h->scenes[s];
    S.desc = sc[s];
    if (sc[s].nbParticles == 0) {
      delete h;
      return fail(RTP_ERR_INVALID, "every scene needs particles");
    }
    rtp_model_desc d{};
    d.maxNbParticles = d.currNbParticles = sc[s].nbParticles;
    for (int a = 0; a < 3; ++a) {
      d.boxSize[a] = sc[s].boxSize[a];
      d.gridRes[a] = sc[s].gridRes[a];
    }
    d.maxPartsInCell = sc[s].maxPartsInCell;
    d.gravity = sc[s].gravity;
    GridConsts g{};
    if ((rc = make_grid(&d, 100u, &g))) {
      delete h;
      return rc;
    }
    init_pbf_consts(S.c, g, d.boxSize[0], d.gridRes[0], d.gravity);
    refresh_pbf_consts(S.c);
    S.slot.pbase = (uint32_t)total;
    S.slot.n = sc[s].nbParticles;
    S.slot.kbase = (uint32_t)keys;
    S.slot.keySpan = max_active_key(g) + 1u;
    total += S.slot.n;
    keys += S.slot.keySpan;
    if (total > 0x7fffffffu || keys > 0xffffffffu) {
      delete h;
      return fail(RTP_ERR_INVALID, "batch exceeds 2^31 particles or 2^32 cell keys");
    }
  }
  h->total = (uint32_t)total;
  h->keySpace = (uint32_t)keys;
  if (hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking) != hipSuccess) {
    delete h;
    return fail(RTP_ERR_HIP, "hipStreamCreate failed");
  }
  uint64_t maxBlocks = 0;  // block table at the smallest block size
  for (auto& S : h->scenes)
    maxBlocks += (S.slot.n + 63u) / 64u;
  const uint32_t N = h->total;
  // ceil(n / b) * b < n + 256 for every block size b <= 256
  h->nbThreads = N + 255u * (uint32_t)h->scenes.size();
  if ((rc = dev_alloc(&h->pos[0], N)) || (rc = dev_alloc(&h->pos[1], N)) ||
      (rc = dev_alloc(&h->vel[0], N)) || (rc = dev_alloc(&h->vel[1], N)) ||
      (rc = dev_alloc(&h->pred[0], N)) || (rc = dev_alloc(&h->pred[1], N)) ||
      (rc = dev_alloc(&h->vort, N)) || (rc = dev_alloc(&h->velTmp, N)) ||
      (rc = dev_alloc(&h->density, N)) || (rc = dev_alloc(&h->lambda, N)) ||
      (rc = dev_alloc(&h->key[0], N)) || (rc = dev_alloc(&h->key[1], N)) ||
      (rc = dev_alloc(&h->idx[0], N)) || (rc = dev_alloc(&h->idx[1], N)) ||
      (rc = dev_alloc(&h->counts, count_capacity(N))) ||
      (rc = dev_alloc(&h->startEnd, h->keySpace)) ||
      (rc = dev_alloc(&h->dSlots, count)) || (rc = dev_alloc(&h->dConsts, count)) ||
      (rc = dev_alloc(&h->dBlocks, maxBlocks)) || (rc = dev_alloc(&h->okey, N)) ||
      (rc = dev_alloc(&h->orank, N)) || (rc = dev_alloc(&h->oorder[0], N)) ||
      (rc = dev_alloc(&h->oorder[1], N)) ||
      (rc = dev_alloc(&h->ocount.table[0], (h->keySpace + 15u) & ~15u)) ||
      (rc = dev_alloc(&h->ocount.table[1], (h->keySpace + 15u) & ~15u)) ||
      (rc = dev_alloc(&h->opart, scan_partials_needed((h->keySpace + 15u) & ~15u)))) {
    h->release();
    delete h;
    return rc;
  }
  h->ocount.words = (h->keySpace + 15u) & ~15u;
  if (hipMemset(h->ocount.table[0], 0, sizeof(uint32_t) * h->ocount.words) != hipSuccess ||
      hipMemset(h->ocount.table[1], 0, sizeof(uint32_t) * h->ocount.words) != hipSuccess) {
    h->release();
    delete h;
    return fail(RTP_ERR_HIP, "hipMemset failed");
  }
  std::vector<SceneSlot> slots(count);
  for (uint32_t s = 0; s < count; ++s)
    slots[s] = h->scenes[s].slot;
  if (hipMemcpyAsync(h->dSlots, slots.data(), sizeof(SceneSlot) * count, hipMemcpyHostToDevice, h->stream) != hipSuccess ||
      hipMemsetAsync(h->pos[0], 0, sizeof(float4) * N, h->stream) != hipSuccess ||
      hipMemsetAsync(h->vel[0], 0, sizeof(float4) * N, h->stream) != hipSuccess ||
      hipMemsetAsync(h->density, 0, sizeof(float) * N, h->stream) != hipSuccess ||
      hipMemsetAsync(h->lambda, 0, sizeof(float) * N, h->stream) != hipSuccess) {
    h->release();
    delete h;
    return fail(RTP_ERR_HIP, "batch initialisation failed");
  }
  if ((rc = h->build_blocks())) {
    h->release();
    delete h;
    return rc;
  }
  *out = h;
  return RTP_OK;
}

int rtp_batch_destroy(rtp_batch* h)
{
  if (!h)
    return RTP_OK;
  (void)hipSetDevice(h->device);
  if (h->stream)
    (void)hipStreamSynchronize(h->stream);
  h->release();
  delete h;
  return RTP_OK;
}

int rtp_batch_set_params(rtp_batch* h, uint32_t scene, const rtp_fluid_params* p)
{
  RTP_REQUIRE(h && p, "NULL argument");
  RTP_REQUIRE(p->restDensity != 0.0f, "restDensity must be non-zero");
  RTP_REQUIRE(p->artPressureExp <= kMaxArtPressureExp, "artPressureExp must be at most 64");
  const uint32_t count = (uint32_t)h->scenes.size();
  RTP_REQUIRE(scene == 0xffffffffu || scene < count, "scene index out of range");
  if (scene == 0xffffffffu) {
    for (uint32_t s = 0; s < count; ++s)
      set_params_one(h, s, p);
  } else {
    RTP_REQUIRE(count == 1 || (p->isVorticityConfEnabled != 0) ==
                                  (h->scenes[scene == 0 ? 1 : 0].c.p.isVorticityConfEnabled != 0),
                "isVorticityConfEnabled must be the same for every scene of a batch");
    set_params_one(h, scene, p);
  }
  h->constsDirty = true;
  h->drop_graph();
  return RTP_OK;
}

int rtp_batch_set_nb_jacobi_iters(rtp_batch* h, uint32_t iters)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(iters <= 1000, "iteration count out of range");
  h->iters = iters;
  h->drop_graph();
  return RTP_OK;
}

int rtp_batch_load_state(rtp_batch* h, uint32_t scene, const float* pos4, const float* vel4)
{
  RTP_REQUIRE(h && pos4 && vel4, "NULL argument");
  int rc = h->check_scene(scene);
  if (rc)
    return rc;
  RTP_HIP_TRY(hipSetDevice(h->device));
  const SceneSlot& sl = h->scenes[scene].slot;
  std::vector<float4> p(sl.n), v(sl.n);
  for (uint32_t i = 0; i < sl.n; ++i) {
    p[i] = make_float4(pos4[4 * i], pos4[4 * i + 1], pos4[4 * i + 2], 0.0f);
    v[i] = make_float4(vel4[4 * i], vel4[4 * i + 1], vel4[4 * i + 2], 0.0f);
  }
  RTP_HIP_TRY(hipMemcpyAsync(h->pos[h->cur] + sl.pbase, p.data(), sizeof(float4) * sl.n, hipMemcpyHostToDevice, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(h->vel[h->cur] + sl.pbase, v.data(), sizeof(float4) * sl.n, hipMemcpyHostToDevice, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_batch_read_state(rtp_batch* h, uint32_t scene, float* pos4, float* vel4)
{
  RTP_REQUIRE(h, "NULL handle");
  int rc = h->check_scene(scene);
  if (rc)
    return rc;
  RTP_HIP_TRY(hipSetDevice(h->device));
  if (pos4 && (rc = read_f4_slice(h, h->pos[h->cur], scene, pos4)))
    return rc;
  if (vel4 && (rc = read_f4_slice(h, h->vel[h->cur], scene, vel4)))
    return rc;
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_batch_read_field(rtp_batch* h, uint32_t scene, int field, float* out)
{
  RTP_REQUIRE(h && out, "NULL argument");
  int rc = h->check_scene(scene);
  if (rc)
    return rc;
  RTP_HIP_TRY(hipSetDevice(h->device));
  const SceneSlot& sl = h->scenes[scene].slot;
  switch (field) {
  case RTP_FIELD_DENSITY:
    RTP_HIP_TRY(hipMemcpyAsync(out, h->density + sl.pbase, sizeof(float) * sl.n, hipMemcpyDeviceToHost, h->stream));
    break;
  case RTP_FIELD_LAMBDA:
    RTP_HIP_TRY(hipMemcpyAsync(out, h->lambda + sl.pbase, sizeof(float) * sl.n, hipMemcpyDeviceToHost, h->stream));
    break;
  case RTP_FIELD_PRED_POS:
    if ((rc = read_f4_slice(h, h->pred[h->predFinal], scene, out)))
      return rc;
    break;
  case RTP_FIELD_VORTICITY:
    if ((rc = read_f4_slice(h, h->vort, scene, out)))
      return rc;
    break;
  default:
    return fail(RTP_ERR_INVALID, "unknown field for a fluids batch");
  }
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_batch_read_grid(rtp_batch* h, uint32_t scene, uint32_t* keys, uint32_t* idx, uint32_t* startEnd)
{
  RTP_REQUIRE(h, "NULL handle");
  int rc = h->check_scene(scene);
  if (rc)
    return rc;
  if (!h->stepped)
    return fail(RTP_ERR_STATE, "the batch has not stepped yet");
  RTP_HIP_TRY(hipSetDevice(h->device));
  const SceneHost& S = h->scenes[scene];
  const SceneSlot& sl = S.slot;
  if (keys)
    RTP_HIP_TRY(hipMemcpyAsync(keys, h->key[h->keyRes] + sl.pbase, sizeof(uint32_t) * sl.n, hipMemcpyDeviceToHost, h->stream));
  if (idx)
    RTP_HIP_TRY(hipMemcpyAsync(idx, h->idx[h->idxRes] + sl.pbase, sizeof(uint32_t) * sl.n, hipMemcpyDeviceToHost, h->stream));
  if (startEnd)
    RTP_HIP_TRY(hipMemcpyAsync(startEnd, h->startEnd + sl.kbase, sizeof(uint2) * S.c.g.numCells, hipMemcpyDeviceToHost, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  for (uint32_t i = 0; keys && i < sl.n; ++i)
    keys[i] -= sl.kbase;
  for (uint32_t i = 0; idx && i < sl.n; ++i)
    idx[i] -= sl.pbase;
  return RTP_OK;
}

int rtp_batch_update(rtp_batch* h, uint32_t nsteps)
{
  RTP_REQUIRE(h, "NULL handle");
  return h->update(nsteps);
}

int rtp_batch_sync(rtp_batch* h)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  return RTP_OK;
}

int rtp_batch_save_checkpoint(rtp_batch* h)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  int rc = 0;
  if (!h->ckPos && ((rc = dev_alloc(&h->ckPos, h->total)) || (rc = dev_alloc(&h->ckVel, h->total))))
    return rc;
  RTP_HIP_TRY(hipMemcpyAsync(h->ckPos, h->pos[h->cur], sizeof(float4) * h->total, hipMemcpyDeviceToDevice, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(h->ckVel, h->vel[h->cur], sizeof(float4) * h->total, hipMemcpyDeviceToDevice, h->stream));
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  h->ckValid = true;
  return RTP_OK;
}

// The state of a batch is its positions and velocities: every substep re-hashes and
// re-sorts all particles and predicts every position afresh.
int rtp_batch_restore_checkpoint(rtp_batch* h)
{
  RTP_REQUIRE(h, "NULL handle");
  if (!h->ckValid)
    return fail(RTP_ERR_STATE, "no checkpoint saved");
  RTP_HIP_TRY(hipSetDevice(h->device));
  RTP_HIP_TRY(hipMemcpyAsync(h->pos[h->cur], h->ckPos, sizeof(float4) * h->total, hipMemcpyDeviceToDevice, h->stream));
  RTP_HIP_TRY(hipMemcpyAsync(h->vel[h->cur], h->ckVel, sizeof(float4) * h->total, hipMemcpyDeviceToDevice, h->stream));
  return RTP_OK;
}

int rtp_batch_use_graph(rtp_batch* h, int enable)
{
  RTP_REQUIRE(h, "NULL handle");
  h->useGraph = enable != 0;
  h->drop_graph();
  return RTP_OK;
}

int rtp_batch_set_tuning(rtp_batch* h, const rtp_tuning* t)
{
  RTP_REQUIRE(h && t, "NULL argument");
  int rc = check_tuning(*t, kBatchLaunchFlags);
  if (rc)
    return rc;
  RTP_HIP_TRY(hipSetDevice(h->device));
  if ((t->flags & RTP_LAUNCH_NB_LIST) && !h->nbList) {
    if ((rc = dev_alloc(&h->nbList, (size_t)(kNbCap / 4u) * h->nbThreads)) ||
        (rc = dev_alloc(&h->nbCount, h->nbThreads))) {
      dev_free(h->nbList);
      h->nbList = nullptr;
      return rc;
    }
  }
  h->sortBits = t->sortBits;
  h->resort = (t->flags & RTP_LAUNCH_RESORT) != 0;
  h->lists = (t->flags & RTP_LAUNCH_NB_LIST) != 0;
  h->queue = t->queueDepth ? t->queueDepth : kDefaultQueue;
  if (h->block != t->blockSize) {
    h->block = t->blockSize;
    return h->build_blocks();
  }
  h->drop_graph();
  return RTP_OK;
}

int rtp_batch_autotune(rtp_batch* h, uint32_t trialSteps, rtp_tuning* best)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_REQUIRE(trialSteps > 0 && trialSteps <= 1000, "trialSteps must be 1..1000");
  RTP_HIP_TRY(hipSetDevice(h->device));
  int rc = rtp_batch_save_checkpoint(h);
  if (rc)
    return rc;
  const bool profWas = h->prof.enabled;
  h->prof.enabled = false;
  rtp_tuning start{};
  start.blockSize = h->block;
  start.sortBits = h->sortBits;
  start.queueDepth = h->queue;
  start.flags = (h->resort ? RTP_LAUNCH_RESORT : 0) | (h->lists ? RTP_LAUNCH_NB_LIST : 0);
  auto run = [&](const rtp_tuning& t) { return rtp_batch_set_tuning(h, &t); };
  auto steps = [&](uint32_t k) {
    int r = 0;
    for (uint32_t s = 0; s < k && !r; ++s)
      if (!(r = rtp_batch_restore_checkpoint(h)))
        r = h->update(1);
    return r;
  };
  rtp_tuning top{};
  rc = tune_search(h->stream, trialSteps, start, false, run, steps, &top);
  h->prof.enabled = profWas;
  if (rc)
    return rc;
  if ((rc = run(top)) || (rc = rtp_batch_restore_checkpoint(h)))
    return rc;
  RTP_HIP_TRY(hipStreamSynchronize(h->stream));
  if (best)
    *best = top;
  return RTP_OK;
}

int rtp_batch_enable_profiling(rtp_batch* h, int enable)
{
  RTP_REQUIRE(h, "NULL handle");
  h->prof.enabled = enable != 0;
  h->drop_graph();
  return RTP_OK;
}

int rtp_batch_profile(rtp_batch* h, rtp_kernel_time* out, int capacity, int* count)
{
  RTP_REQUIRE(h, "NULL handle");
  RTP_HIP_TRY(hipSetDevice(h->device));
  return h->prof.report(out, capacity, count);
}

int rtp_batch_sizes(rtp_batch* h, uint32_t* particles, uint32_t* cells)
{
  RTP_REQUIRE(h, "NULL handle");
  if (particles)
    *particles = h->total;
  if (cells)
    *cells = h->keySpace;
  return RTP_OK;
}

}  // extern "C"
