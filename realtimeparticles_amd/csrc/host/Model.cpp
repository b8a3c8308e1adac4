// Physics::Model and Physics::HIP::{Fluids,Boids} on top of the C ABI (rtp.h).
#include "Model.hpp"

#include <cmath>
#include <cstdlib>
#include <limits>

namespace {

void check(int rc, const char* what)
{
  if (rc != RTP_OK)
    throw std::runtime_error(std::string(what) + ": " + rtp_last_error());
}

// The initial parameter trees, initFluidsJson (physics/ocl/Fluids.cpp:52-74) and
// initBoidsJson (physics/ocl/Boids.cpp:41-73), with their float literals written so that
// they read back as the same floats.
const char* const kInitFluidsJson =
    "{\"Fluids\": {\"Rest Density\": [450, 10, 1000], \"Relax CFM\": [600, 100, 1000], "
    "\"Time Step\": [0.01, 0.0001, 0.02], \"Nb Jacobi Iterations\": [2, 1, 6], "
    "\"Artificial Pressure\": {\"Enable##Pressure\": true, \"Coefficient##Pressure\": [0.001, 0, 0.001], "
    "\"Radius\": [0.006, 0.001, 0.015], \"Exp\": [4, 1, 6]}, "
    "\"Vorticity Confinement\": {\"Enable##Vorticity\": true, \"Coefficient##Vorticity\": [0.0004, 0, 0.001], "
    "\"xSPH Viscosity Coefficient\": [0.0001, 0, 0.001]}}}";
// initCloudsJson (physics/ocl/Clouds.cpp:63-98)
const char* const kInitCloudsJson =
    "{\"Fluids\": {\"Rest Density\": [450, 10, 1000], \"Relax CFM\": [600, 100, 1000], "
    "\"Time Step\": [0.01, 0.0001, 0.02], \"Nb Jacobi Iterations\": [2, 1, 6], "
    "\"Artificial Pressure\": {\"Enable##Pressure\": true, \"Coefficient##Pressure\": [0.001, 0, 0.001], "
    "\"Radius\": [0.006, 0.001, 0.015], \"Exp\": [4, 1, 6]}, "
    "\"Vorticity Confinement\": {\"Enable##Vorticity\": true, \"Coefficient##Vorticity\": [0.0004, 0, 0.001], "
    "\"xSPH Viscosity Coefficient\": [0.0001, 0, 0.001]}}, "
    "\"Clouds\": {\"Enable Temperature Smoothing\": true, \"Ground Heat Coefficient\": [10, 0, 1000], "
    "\"Buoyancy Heat Coefficient\": [0.1, 0, 5], \"Gravity Coefficient\": [0.0005, 0, 0.1], "
    "\"Adiabatic Lapse Rate\": [5, 0, 20], \"Phase Transition Rate\": [0.3485, 0, 20], "
    "\"Latent Heat Coefficient\": [0.07, 0, 0.1], \"Wind Coefficient\": [1, 0, 1]}}";

const char* const kInitBoidsJson =
    "{\"Boids\": {\"Velocity\": [0.5, 0.01, 5], "
    "\"Target\": {\"Enable##Target\": false, \"Show\": true, \"Radius\": [2, 1, 20], \"Attract\": true}, "
    "\"Alignment\": {\"Enable##Alignment\": true, \"Scale##Alignment\": [1.6, 0, 3]}, "
    "\"Cohesion\": {\"Enable##Cohesion\": true, \"Scale##Cohesion\": [1.45, 0, 3]}, "
    "\"Separation\": {\"Enable##Separation\": true, \"Scale##Separation\": [1.6, 0, 3]}}}";

struct F3 {
  float x, y, z;
};

// utils/Geometry.cpp:198-229 GenerateBoxGrid (uniform): start + i * (end - start) / res.
std::vector<F3> box_grid(std::array<int, 3> res, F3 s, F3 e)
{
  std::vector<F3> v;
  v.reserve((size_t)res[0] * res[1] * res[2]);
  const F3 sp{ (e.x - s.x) / res[0], (e.y - s.y) / res[1], (e.z - s.z) / res[2] };
  for (int ix = 0; ix < res[0]; ++ix)
    for (int iy = 0; iy < res[1]; ++iy)
      for (int iz = 0; iz < res[2]; ++iz)
        v.push_back(F3{ s.x + ix * sp.x, s.y + iy * sp.y, s.z + iz * sp.z });
  return v;
}

// utils/Geometry.cpp:242-272 GenerateSphereGrid: shells of a sphere inscribed in [s, e].
std::vector<F3> sphere_grid(std::array<int, 3> res, F3 s, F3 e)
{
  const float PI_F = 3.1415927f;
  const F3 vec{ e.x - s.x, e.y - s.y, e.z - s.z };
  const F3 c{ s.x + vec.x / 2.0f, s.y + vec.y / 2.0f, s.z + vec.z / 2.0f };
  const float radius = std::sqrt(vec.x * vec.x + vec.y * vec.y + vec.z * vec.z) / 2.0f;
  const float dphi = PI_F / res[0], dtheta = 2.0f * PI_F / res[1], dr = radius / res[2];
  std::vector<F3> v;
  v.reserve((size_t)res[0] * res[1] * res[2]);
  for (int ip = 0; ip < res[0]; ++ip)
    for (int it = 0; it < res[1]; ++it)
      for (int ir = 0; ir < res[2]; ++ir) {
        const float r = (ir + 1) * dr;
        v.push_back(F3{ c.x + r * std::cos(it * dtheta) * std::sin(ip * dphi),
                        c.y + r * std::sin(it * dtheta) * std::sin(ip * dphi),
                        c.z + r * std::cos(ip * dphi) });
      }
  return v;
}

// utils/Geometry.cpp:71-133 GenerateRectangularGrid (uniform) in the YZ plane: res[0] rows
// along y, res[1] along z, start + i * (end - start) / res.
std::vector<F3> rect_grid_yz(std::array<int, 2> res, F3 s, F3 e)
{
  std::vector<F3> v;
  v.reserve((size_t)res[0] * res[1]);
  const F3 sp{ 0.0f, (e.y - s.y) / res[0], (e.z - s.z) / res[1] };
  for (int iy = 0; iy < res[0]; ++iy)
    for (int iz = 0; iz < res[1]; ++iz)
      v.push_back(F3{ s.x + 0 * sp.x, s.y + iy * sp.y, s.z + iz * sp.z });
  return v;
}

// utils/Geometry.cpp:136-196 GenerateCircularGrid in the YZ plane: res[0] angles times
// res[1] rings of the circle inscribed in [s, e], the innermost ring one spacing out.
std::vector<F3> circle_grid_yz(std::array<int, 2> res, F3 s, F3 e)
{
  const float PI_F = 3.1415927f;
  const F3 vec{ e.x - s.x, e.y - s.y, e.z - s.z };
  const F3 c{ s.x + vec.x / 2.0f, s.y + vec.y / 2.0f, s.z + vec.z / 2.0f };
  const float radius = std::sqrt(vec.x * vec.x + vec.y * vec.y + vec.z * vec.z) / 2.0f;
  const float da = 2.0f * PI_F / res[0], dr = radius / res[1];
  std::vector<F3> v;
  v.reserve((size_t)res[0] * res[1]);
  for (int io = 0; io < res[0]; ++io)
    for (int ir = 0; ir < res[1]; ++ir) {
      const float r = (ir + 1) * dr;
      v.push_back(F3{ c.x, c.y + r * std::cos(io * da), c.z + r * std::sin(io * da) });
    }
  return v;
}

// utils/Geometry.cpp:123-132 / 229-238 with Distribution::Random: every vertex takes three
// rand() draws, x then y then z, each rand() / RAND_MAX scaled into [s, e] in float.  The
// layout is the reference's when the process's rand() sequence is at the same point.
std::vector<F3> random_grid(size_t count, F3 s, F3 e)
{
  const F3 vec{ e.x - s.x, e.y - s.y, e.z - s.z };
  std::vector<F3> v;
  v.reserve(count);
  for (size_t i = 0; i < count; ++i) {
    const float x = (float)std::rand() / (float)RAND_MAX * vec.x + s.x;
    const float y = (float)std::rand() / (float)RAND_MAX * vec.y + s.y;
    const float z = (float)std::rand() / (float)RAND_MAX * vec.z + s.z;
    v.push_back(F3{ x, y, z });
  }
  return v;
}

std::vector<float> to_float4(const std::vector<F3>& v)
{
  std::vector<float> out(v.size() * 4, 0.0f);
  for (size_t i = 0; i < v.size(); ++i) {
    out[4 * i + 0] = v[i].x;
    out[4 * i + 1] = v[i].y;
    out[4 * i + 2] = v[i].z;
  }
  return out;
}

rtp_model_desc make_desc(const Physics::ModelParams& p)
{
  rtp_model_desc d{};
  d.maxNbParticles = (uint32_t)p.maxNbParticles;
  d.currNbParticles = (uint32_t)p.currNbParticles;
  d.boxSize[0] = p.boxSize.x;
  d.boxSize[1] = p.boxSize.y;
  d.boxSize[2] = p.boxSize.z;
  d.gridRes[0] = p.gridRes.x;
  d.gridRes[1] = p.gridRes.y;
  d.gridRes[2] = p.gridRes.z;
  d.maxPartsInCell = 0;
  d.gravity = p.gravity;
  d.device = p.device;
  return d;
}

}  // namespace

namespace Utils {
// utils/Parameters.hpp:89-98 (ALL_NB_PARTICLES, subdiv2D column)
std::array<int, 2> GetNbParticlesSubdiv2D(NbParticles nb)
{
  switch (nb) {
  case P512: return { 32, 16 };
  case P1K: return { 32, 32 };
  case P4K: return { 64, 64 };
  case P8K: return { 128, 64 };
  case P16K: return { 128, 128 };
  case P32K: return { 128, 256 };
  case P65K: return { 256, 256 };
  case P130K: return { 256, 512 };
  }
  return { 0, 0 };
}

std::array<int, 3> GetNbParticlesSubdiv3D(NbParticles nb)
{
  switch (nb) {
  case P512: return { 8, 8, 8 };
  case P1K: return { 16, 8, 8 };
  case P4K: return { 16, 16, 16 };
  case P8K: return { 32, 16, 16 };
  case P16K: return { 32, 32, 16 };
  case P32K: return { 32, 32, 32 };
  case P65K: return { 64, 32, 32 };
  case P130K: return { 64, 64, 32 };
  }
  return { 0, 0, 0 };
}
}  // namespace Utils

namespace Physics {

std::unique_ptr<Model> CreateModel(ModelType type, ModelParams params)
{
  switch ((int)type) {
  case ModelType::BOIDS: return std::make_unique<HIP::Boids>(params);
  case ModelType::FLUIDS: return std::make_unique<HIP::Fluids>(params);
  case ModelType::CLOUDS: return std::make_unique<HIP::Clouds>(params);
  default: return nullptr;
  }
}

Model::Model(const ModelParams& params)
    : m_maxNbParticles(params.maxNbParticles)
    , m_currNbParticles(params.currNbParticles)
    , m_boxSize(params.boxSize)
    , m_gridRes(params.gridRes)
    , m_nbCells((size_t)params.gridRes.x * params.gridRes.y * params.gridRes.z)
    , m_dimension(params.dimension)
    , m_case(params.pCase)
    , m_device(params.device)
{
}

void Model::setNbParticles(size_t nb)
{
  m_currNbParticles = nb;
  if (m_init)
    applyNbParticles();
}

void Model::pause(bool p)
{
  m_pause = p;
  if (m_init)
    applyPause();
}

void Model::enableProfiling(bool enable)
{
  m_profiling = enable;
  if (m_init)
    applyProfiling();
}

void Model::updateInputJson(const std::string& newJson)
{
  int equal = 0;
  check(rtp_json_equal(m_inputJson.c_str(), newJson.c_str(), &equal), "rtp_json_equal");
  if (equal)
    return;
  size_t n = 0;
  check(rtp_json_merge_patch(m_inputJson.c_str(), newJson.c_str(), nullptr, 0, &n),
        "rtp_json_merge_patch");
  std::string merged(n + 1, '\0');
  check(rtp_json_merge_patch(m_inputJson.c_str(), newJson.c_str(), &merged[0], merged.size(), &n),
        "rtp_json_merge_patch");
  merged.resize(n);
  m_inputJson = merged;
  updateModelWithInputJson(m_inputJson);
}

namespace HIP {

// ------------------------------------------------------------------------------ Fluids

Fluids::Fluids(const ModelParams& params) : Model(params)
{
  const rtp_model_desc d = make_desc(params);
  check(rtp_fluids_create(&d, &m_h), "rtp_fluids_create");
  check(rtp_fluids_set_nb_jacobi_iters(m_h, m_nbJacobiIters), "rtp_fluids_set_nb_jacobi_iters");
  resetInputJson(kInitFluidsJson);  // Fluids.cpp:76: the model starts with the initial tree
  m_init = true;
  if (m_case != Utils::CASE_INVALID)
    reset();
}

Fluids::~Fluids() { rtp_fluids_destroy(m_h); }

void Fluids::update()
{
  if (!m_init)
    return;
  check(rtp_fluids_update(m_h, 1), "rtp_fluids_update");
  check(rtp_fluids_sync(m_h), "rtp_fluids_sync");  // a frame ends synchronised (Context.cpp:746)
}

// Fluids::reset / initFluidsParticles (Fluids.cpp:196-216, 273-398): the 2-D cases lie in
// the YZ plane (:287-335), the 3-D ones fill boxes and a sphere.
void Fluids::reset()
{
  if (!m_init)
    return;
  const bool d2 = m_dimension == Geometry::Dimension::dim2D;
  // FluidParams.dim follows the model's dimension (Fluids.cpp:234), then the initial tree
  rtp_fluid_params p{};
  check(rtp_fluids_get_params(m_h, &p), "rtp_fluids_get_params");
  p.dim = d2 ? 2u : 3u;
  check(rtp_fluids_set_params(m_h, &p), "rtp_fluids_set_params");
  resetInputJson(kInitFluidsJson);
  updateModelWithInputJson(getInputJson());
  const F3 b{ m_boxSize.x, m_boxSize.y, m_boxSize.z };
  std::vector<F3> verts;
  if (d2) {
    switch (m_case) {
    case Utils::FLUIDS_DAM:
      m_currNbParticles = Utils::P4K;
      verts = rect_grid_yz(Utils::GetNbParticlesSubdiv2D(Utils::P4K), F3{ 0.0f, b.y / -2.0f, b.z / -2.0f },
                           F3{ 0.0f, 0.0f, 0.0f });
      break;
    case Utils::FLUIDS_BOMB:
      m_currNbParticles = Utils::P4K;
      verts = rect_grid_yz(Utils::GetNbParticlesSubdiv2D(Utils::P4K), F3{ 0.0f, b.y / -6.0f, b.z / -6.0f },
                           F3{ 0.0f, b.y / 6.0f, b.z / 6.0f });
      break;
    case Utils::FLUIDS_DROP: {
      m_currNbParticles = Utils::P512 + Utils::P4K;
      verts = rect_grid_yz(Utils::GetNbParticlesSubdiv2D(Utils::P512),
                           F3{ 0.0f, 2.0f * b.y / 10.0f, b.z / -10.0f }, F3{ 0.0f, 4.0f * b.y / 10.0f, b.z / 10.0f });
      // a 64 x 128 floor of which only the first 4,096 points are active, as in the reference
      const auto bottom = rect_grid_yz({ 64, 128 }, F3{ 0.0f, b.y / -2.0f, b.z / -2.0f },
                                       F3{ 0.0f, 0.0f, b.z / 2.0f });
      verts.insert(verts.end(), bottom.begin(), bottom.end());
      break;
    }
    default:
      throw std::runtime_error("Fluids::reset: unknown case");
    }
    if (verts.size() > m_maxNbParticles)
      verts.resize(m_maxNbParticles);
  } else switch (m_case) {
  case Utils::FLUIDS_DAM:
    m_currNbParticles = Utils::P130K;
    verts = box_grid(Utils::GetNbParticlesSubdiv3D(Utils::P130K), F3{ b.x / -2.0f, b.y / -2.0f, b.z / -2.0f },
                     F3{ b.x / 2.0f, 0.0f, 0.0f });
    break;
  case Utils::FLUIDS_BOMB:
    m_currNbParticles = Utils::P65K;
    verts = sphere_grid(Utils::GetNbParticlesSubdiv3D(Utils::P65K), F3{ b.x / -6.0f, b.y / -6.0f, b.z / -6.0f },
                        F3{ b.x / 6.0f, b.y / 6.0f, b.z / 6.0f });
    break;
  case Utils::FLUIDS_DROP: {
    m_currNbParticles = Utils::P4K + Utils::P65K;
    verts = box_grid(Utils::GetNbParticlesSubdiv3D(Utils::P4K),
                     F3{ b.x / -10.0f, 2.0f * b.y / 10.0f, b.z / -10.0f },
                     F3{ b.x / 10.0f, 4.0f * b.y / 10.0f, b.z / 10.0f });
    const auto bottom = box_grid({ 64, 16, 64 }, F3{ b.x / -2.0f, b.y / -2.0f, b.z / -2.0f },
                                 F3{ b.x / 2.0f, b.y / -2.55f, b.z / 2.0f });
    verts.insert(verts.end(), bottom.begin(), bottom.end());
    break;
  }
  default:
    throw std::runtime_error("Fluids::reset: unknown case");
  }
  if (m_currNbParticles > m_maxNbParticles)
    throw std::runtime_error("Fluids::reset: case needs more particles than maxNbParticles");
  const std::vector<float> pos = to_float4(verts);
  const std::vector<float> vel(pos.size(), 0.0f);
  check(rtp_fluids_set_nb_particles(m_h, (uint32_t)m_currNbParticles), "rtp_fluids_set_nb_particles");
  check(rtp_fluids_load_state(m_h, pos.data(), vel.data(), (uint32_t)verts.size()), "rtp_fluids_load_state");
}

void Fluids::setFluidParams(const rtp_fluid_params& p) { check(rtp_fluids_set_params(m_h, &p), "rtp_fluids_set_params"); }

rtp_fluid_params Fluids::fluidParams() const
{
  rtp_fluid_params p{};
  check(rtp_fluids_get_params(m_h, &p), "rtp_fluids_get_params");
  return p;
}

void Fluids::setNbJacobiIters(uint32_t iters)
{
  check(rtp_fluids_set_nb_jacobi_iters(m_h, iters), "rtp_fluids_set_nb_jacobi_iters");
  m_nbJacobiIters = iters;
}

void Fluids::loadState(const float* pos4, const float* vel4, size_t count)
{
  check(rtp_fluids_load_state(m_h, pos4, vel4, (uint32_t)count), "rtp_fluids_load_state");
}

void Fluids::readState(float* pos4, float* vel4, size_t count) const
{
  check(rtp_fluids_read_state(m_h, pos4, vel4, (uint32_t)count), "rtp_fluids_read_state");
}

void* Fluids::devicePositions() const
{
  void* p = nullptr;
  check(rtp_fluids_device_state(m_h, &p, nullptr), "rtp_fluids_device_state");
  return p;
}

void Fluids::updateModelWithInputJson(const std::string& json)
{
  check(rtp_fluids_update_with_json(m_h, json.c_str()), "rtp_fluids_update_with_json");
  // the count the model runs (the library caps it at 1000)
  uint32_t iters = 0;
  check(rtp_fluids_get_nb_jacobi_iters(m_h, &iters), "rtp_fluids_get_nb_jacobi_iters");
  m_nbJacobiIters = iters;
}

std::string Fluids::inputJson() const
{
  size_t n = 0;
  check(rtp_fluids_params_json(m_h, nullptr, 0, &n), "rtp_fluids_params_json");
  std::string s(n + 1, '\0');
  check(rtp_fluids_params_json(m_h, &s[0], s.size(), &n), "rtp_fluids_params_json");
  s.resize(n);
  return s;
}

void* Fluids::deviceColors() const
{
  void* p = nullptr;
  check(rtp_fluids_device_render(m_h, &p, nullptr, nullptr), "rtp_fluids_device_render");
  return p;
}

void Fluids::enableRendering(bool enable)
{
  check(rtp_fluids_set_render(m_h, enable ? 1 : 0), "rtp_fluids_set_render");
}

void Fluids::setCameraPos(float x, float y, float z)
{
  const float p[3] = { x, y, z };
  check(rtp_fluids_set_camera(m_h, p), "rtp_fluids_set_camera");
}

std::vector<rtp_kernel_time> Fluids::kernelTimes() const
{
  int count = 0;
  check(rtp_fluids_profile(m_h, nullptr, 0, &count), "rtp_fluids_profile");
  std::vector<rtp_kernel_time> out((size_t)count);
  check(rtp_fluids_profile(m_h, out.data(), count, &count), "rtp_fluids_profile");
  return out;
}

std::array<float, 3> Fluids::densityError() const
{
  std::array<float, 3> e{};
  check(rtp_fluids_density_error(m_h, e.data()), "rtp_fluids_density_error");
  return e;
}

void Fluids::applyNbParticles()
{
  check(rtp_fluids_set_nb_particles(m_h, (uint32_t)m_currNbParticles), "rtp_fluids_set_nb_particles");
}

void Fluids::applyPause() { check(rtp_fluids_pause(m_h, m_pause ? 1 : 0), "rtp_fluids_pause"); }

void Fluids::applyProfiling()
{
  check(rtp_fluids_enable_profiling(m_h, m_profiling ? 1 : 0), "rtp_fluids_enable_profiling");
}

// ------------------------------------------------------------------------------- Boids

Boids::Boids(const ModelParams& params) : Model(params)
{
  const rtp_model_desc d = make_desc(params);
  check(rtp_boids_create(&d, &m_h), "rtp_boids_create");
  resetInputJson(kInitBoidsJson);  // Boids.cpp:75
  m_init = true;
  if (m_case != Utils::CASE_INVALID)
    reset();
}

Boids::~Boids() { rtp_boids_destroy(m_h); }

void Boids::update()
{
  if (!m_init)
    return;
  check(rtp_boids_set_dimension(m_h, m_dimension == Geometry::Dimension::dim2D ? 2 : 3),
        "rtp_boids_set_dimension");
  check(rtp_boids_set_boundary(m_h, m_boundary == Boundary::CyclicWall ? RTP_BOUNDARY_CYCLIC
                                                                       : RTP_BOUNDARY_BOUNCING),
        "rtp_boids_set_boundary");
  check(rtp_boids_update(m_h, 1), "rtp_boids_update");
  check(rtp_boids_sync(m_h), "rtp_boids_sync");
}

// Boids::reset / initBoidsParticles (Boids.cpp:228-321): a circle lattice in the YZ plane
// (2-D) or a sphere lattice (3-D), with the velocity buffer initialised from the positions.
void Boids::reset()
{
  if (!m_init)
    return;
  resetInputJson(kInitBoidsJson);
  switch (m_case) {
  case Utils::BOIDS_SMALL: m_currNbParticles = Utils::P512; break;
  case Utils::BOIDS_MEDIUM: m_currNbParticles = Utils::P16K; break;
  case Utils::BOIDS_LARGE: m_currNbParticles = Utils::P65K; break;
  case Utils::BOIDS_XLARGE: m_currNbParticles = Utils::P130K; break;
  default: throw std::runtime_error("Boids::reset: unknown case");
  }
  updateModelWithInputJson(getInputJson());  // Boids.cpp:259
  if (m_currNbParticles > m_maxNbParticles)
    throw std::runtime_error("Boids::reset: case needs more particles than maxNbParticles");
  const F3 b{ m_boxSize.x, m_boxSize.y, m_boxSize.z };
  const auto nb = (Utils::NbParticles)m_currNbParticles;
  const auto verts = m_dimension == Geometry::Dimension::dim2D
                         ? circle_grid_yz(Utils::GetNbParticlesSubdiv2D(nb), F3{ 0.0f, b.y / -6.0f, b.z / -6.0f },
                                          F3{ 0.0f, b.y / 6.0f, b.z / 6.0f })
                         : sphere_grid(Utils::GetNbParticlesSubdiv3D(nb), F3{ b.x / -6.0f, b.y / -6.0f, b.z / -6.0f },
                                       F3{ b.x / 6.0f, b.y / 6.0f, b.z / 6.0f });
  const std::vector<float> pos = to_float4(verts);
  check(rtp_boids_set_nb_particles(m_h, (uint32_t)m_currNbParticles), "rtp_boids_set_nb_particles");
  check(rtp_boids_load_state(m_h, pos.data(), pos.data(), (uint32_t)verts.size()), "rtp_boids_load_state");
}

void Boids::setBoidsParams(const rtp_boids_params& p) { check(rtp_boids_set_params(m_h, &p), "rtp_boids_set_params"); }

void Boids::setTimeStep(float dt) { check(rtp_boids_set_time_step(m_h, dt), "rtp_boids_set_time_step"); }

void Boids::setTarget(bool active, std::array<float, 3> pos, float radius, int sign)
{
  check(rtp_boids_set_target(m_h, active ? 1 : 0, pos.data(), radius, sign), "rtp_boids_set_target");
  readTarget();
}

void Boids::readTarget()
{
  rtp_target_params t{};
  float pos[3] = { 0.0f, 0.0f, 0.0f };
  check(rtp_boids_get_target(m_h, &t, pos), "rtp_boids_get_target");
  m_targetActive = t.active != 0;
  m_targetVisible = t.visible != 0;
  m_targetPos = { pos[0], pos[1], pos[2] };
}

void Boids::loadState(const float* pos4, const float* vel4, size_t count)
{
  check(rtp_boids_load_state(m_h, pos4, vel4, (uint32_t)count), "rtp_boids_load_state");
}

void Boids::readState(float* pos4, float* vel4, size_t count) const
{
  check(rtp_boids_read_state(m_h, pos4, vel4, (uint32_t)count), "rtp_boids_read_state");
}

void* Boids::devicePositions() const
{
  void* p = nullptr;
  check(rtp_boids_device_state(m_h, &p, nullptr), "rtp_boids_device_state");
  return p;
}

void Boids::updateModelWithInputJson(const std::string& json)
{
  // the target keeps its last position (rtp_boids_set_target stores it even while off)
  check(rtp_boids_update_with_json(m_h, json.c_str()), "rtp_boids_update_with_json");
  readTarget();
}

std::string Boids::inputJson() const
{
  size_t n = 0;
  check(rtp_boids_params_json(m_h, nullptr, 0, &n), "rtp_boids_params_json");
  std::string s(n + 1, '\0');
  check(rtp_boids_params_json(m_h, &s[0], s.size(), &n), "rtp_boids_params_json");
  s.resize(n);
  return s;
}

void* Boids::deviceColors() const
{
  void* p = nullptr;
  check(rtp_boids_device_render(m_h, &p, nullptr, nullptr), "rtp_boids_device_render");
  return p;
}

void Boids::enableRendering(bool enable)
{
  check(rtp_boids_set_render(m_h, enable ? 1 : 0), "rtp_boids_set_render");
}

void Boids::setCameraPos(float x, float y, float z)
{
  const float p[3] = { x, y, z };
  check(rtp_boids_set_camera(m_h, p), "rtp_boids_set_camera");
}

std::vector<rtp_kernel_time> Boids::kernelTimes() const
{
  int count = 0;
  check(rtp_boids_profile(m_h, nullptr, 0, &count), "rtp_boids_profile");
  std::vector<rtp_kernel_time> out((size_t)count);
  check(rtp_boids_profile(m_h, out.data(), count, &count), "rtp_boids_profile");
  return out;
}

void Boids::applyNbParticles()
{
  check(rtp_boids_set_nb_particles(m_h, (uint32_t)m_currNbParticles), "rtp_boids_set_nb_particles");
}

void Boids::applyPause() { check(rtp_boids_pause(m_h, m_pause ? 1 : 0), "rtp_boids_pause"); }

void Boids::applyProfiling()
{
  check(rtp_boids_enable_profiling(m_h, m_profiling ? 1 : 0), "rtp_boids_enable_profiling");
}

// ------------------------------------------------------------------------------ Clouds

Clouds::Clouds(const ModelParams& params) : Model(params)
{
  const rtp_model_desc d = make_desc(params);
  check(rtp_clouds_create(&d, &m_h), "rtp_clouds_create");
  check(rtp_clouds_set_nb_jacobi_iters(m_h, m_nbJacobiIters), "rtp_clouds_set_nb_jacobi_iters");
  resetInputJson(kInitCloudsJson);
  m_init = true;
  if (m_case != Utils::CASE_INVALID)
    reset();
}

Clouds::~Clouds() { rtp_clouds_destroy(m_h); }

void Clouds::update()
{
  check(rtp_clouds_update(m_h, 1), "rtp_clouds_update");
  check(rtp_clouds_sync(m_h), "rtp_clouds_sync");  // each frame is synchronous (Context.cpp:746-747)
}

// Clouds::reset / initCloudsParticles (Clouds.cpp:379-500): the initial tree, then random
// particles over the lower part of the box (cumulus) or the whole box (homogeneous).
void Clouds::reset()
{
  if (!m_init)
    return;
  const bool d2 = m_dimension == Geometry::Dimension::dim2D;
  rtp_fluid_params p{};
  rtp_cloud_params q{};
  check(rtp_clouds_get_params(m_h, &p, &q), "rtp_clouds_get_params");
  p.dim = q.dim = d2 ? 2u : 3u;  // Clouds.cpp:260, 288
  check(rtp_clouds_set_params(m_h, &p, &q), "rtp_clouds_set_params");
  resetInputJson(kInitCloudsJson);
  updateModelWithInputJson(getInputJson());
  const F3 b{ m_boxSize.x, m_boxSize.y, m_boxSize.z };
  std::vector<F3> verts;
  if (d2) {
    // Geometry::Generate2DGrid(Rectangle, YZ, subdiv2D(P8K), ..., Random)
    const auto sub = Utils::GetNbParticlesSubdiv2D(Utils::P8K);
    m_currNbParticles = Utils::P8K;
    switch (m_case) {
    case Utils::CLOUDS_CUMULUS:
      verts = random_grid((size_t)sub[0] * sub[1], F3{ 0.0f, b.y / -2.0f, b.z / -2.0f },
                          F3{ 0.0f, 0.0f, b.z / 2.0f });
      break;
    case Utils::CLOUDS_HOMOGENEOUS:
      verts = random_grid((size_t)sub[0] * sub[1], F3{ 0.0f, b.y / -2.0f, b.z / -2.0f },
                          F3{ 0.0f, b.y / 2.0f, b.z / 2.0f });
      break;
    default:
      throw std::runtime_error("Clouds::reset: unknown case");
    }
  } else {
    const auto sub = Utils::GetNbParticlesSubdiv3D(Utils::P65K);
    m_currNbParticles = Utils::P65K;
    switch (m_case) {
    case Utils::CLOUDS_CUMULUS:
      verts = random_grid((size_t)sub[0] * sub[1] * sub[2], F3{ b.x / -2.0f, b.y / -2.0f, b.z / -2.0f },
                          F3{ b.x / 2.0f, b.y / -4.0f, b.z / 2.0f });
      break;
    case Utils::CLOUDS_HOMOGENEOUS:
      verts = random_grid((size_t)sub[0] * sub[1] * sub[2], F3{ b.x / -2.0f, b.y / -2.0f, b.z / -2.0f },
                          F3{ b.x / 2.0f, b.y / 2.0f, b.z / 2.0f });
      break;
    default:
      throw std::runtime_error("Clouds::reset: unknown case");
    }
  }
  if (m_currNbParticles > m_maxNbParticles)
    throw std::runtime_error("Clouds::reset: case needs more particles than maxNbParticles");
  const std::vector<float> pos = to_float4(verts);
  const std::vector<float> vel(pos.size(), 0.0f);
  check(rtp_clouds_load_state(m_h, pos.data(), vel.data(), (uint32_t)verts.size()), "rtp_clouds_load_state");
}

void Clouds::setFluidParams(const rtp_fluid_params& p)
{
  check(rtp_clouds_set_params(m_h, &p, nullptr), "rtp_clouds_set_params");
}

void Clouds::setCloudParams(const rtp_cloud_params& q)
{
  check(rtp_clouds_set_params(m_h, nullptr, &q), "rtp_clouds_set_params");
}

rtp_fluid_params Clouds::fluidParams() const
{
  rtp_fluid_params p{};
  check(rtp_clouds_get_params(m_h, &p, nullptr), "rtp_clouds_get_params");
  return p;
}

rtp_cloud_params Clouds::cloudParams() const
{
  rtp_cloud_params q{};
  check(rtp_clouds_get_params(m_h, nullptr, &q), "rtp_clouds_get_params");
  return q;
}

void Clouds::setNbJacobiIters(uint32_t iters)
{
  check(rtp_clouds_set_nb_jacobi_iters(m_h, iters), "rtp_clouds_set_nb_jacobi_iters");
  m_nbJacobiIters = iters;
}

std::vector<float> Clouds::quantity(const std::string& name) const
{
  int field = -1;
  if (name == "Particle ID")
    field = RTP_FIELD_PARTICLE_ID;
  else if (name == "Vapor Density")
    field = RTP_FIELD_VAPOR_DENSITY;
  else if (name == "Cloud Density")
    field = RTP_FIELD_CLOUD_DENSITY;
  else if (name == "Net Force")
    field = RTP_FIELD_BUOYANCY;
  else if (name == "Temperature")
    field = RTP_FIELD_TEMPERATURE;
  else
    throw std::runtime_error("Clouds::quantity: unknown quantity " + name);
  std::vector<float> out(m_currNbParticles);
  check(rtp_clouds_read_field(m_h, field, out.data(), (uint32_t)out.size()), "rtp_clouds_read_field");
  return out;
}

void Clouds::loadState(const float* pos4, const float* vel4, size_t count)
{
  check(rtp_clouds_load_state(m_h, pos4, vel4, (uint32_t)count), "rtp_clouds_load_state");
  m_currNbParticles = count;
}

void Clouds::readState(float* pos4, float* vel4, size_t count) const
{
  check(rtp_clouds_read_state(m_h, pos4, vel4, (uint32_t)count), "rtp_clouds_read_state");
}

void* Clouds::devicePositions() const
{
  void* p = nullptr;
  check(rtp_clouds_device_state(m_h, &p, nullptr), "rtp_clouds_device_state");
  return p;
}

std::vector<rtp_kernel_time> Clouds::kernelTimes() const
{
  int count = 0;
  check(rtp_clouds_profile(m_h, nullptr, 0, &count), "rtp_clouds_profile");
  std::vector<rtp_kernel_time> out((size_t)count);
  if (count)
    check(rtp_clouds_profile(m_h, out.data(), count, &count), "rtp_clouds_profile");
  return out;
}

void Clouds::enableRendering(bool enable)
{
  if (enable)
    throw std::runtime_error("Clouds: the rendering steps are not provided for this model");
}

void Clouds::setCameraPos(float, float, float)
{
  throw std::runtime_error("Clouds: the rendering steps are not provided for this model");
}

void* Clouds::deviceColors() const
{
  throw std::runtime_error("Clouds: the rendering steps are not provided for this model");
}

void Clouds::updateModelWithInputJson(const std::string& json)
{
  check(rtp_clouds_update_with_json(m_h, json.c_str()), "rtp_clouds_update_with_json");
  uint32_t iters = 0;
  check(rtp_clouds_get_nb_jacobi_iters(m_h, &iters), "rtp_clouds_get_nb_jacobi_iters");
  m_nbJacobiIters = iters;
}

std::string Clouds::inputJson() const
{
  size_t n = 0;
  check(rtp_clouds_params_json(m_h, nullptr, 0, &n), "rtp_clouds_params_json");
  std::string s(n + 1, '\0');
  check(rtp_clouds_params_json(m_h, &s[0], s.size(), &n), "rtp_clouds_params_json");
  s.resize(n);
  return s;
}

void Clouds::applyNbParticles()
{
  check(rtp_clouds_set_nb_particles(m_h, (uint32_t)m_currNbParticles), "rtp_clouds_set_nb_particles");
}

void Clouds::applyPause() { check(rtp_clouds_pause(m_h, m_pause ? 1 : 0), "rtp_clouds_pause"); }

void Clouds::applyProfiling()
{
  check(rtp_clouds_enable_profiling(m_h, m_profiling ? 1 : 0), "rtp_clouds_enable_profiling");
}

}  // namespace HIP
}  // namespace Physics
