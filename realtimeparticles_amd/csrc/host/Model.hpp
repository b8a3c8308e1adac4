// C++ host interface of the MI355X backend, shaped like the reference's Physics::Model
// (physics/Model.hpp:15-216) so the reference app's calls (CreateModel, update, reset,
// setNbParticles, setDimension, setBoundary, pause, enableProfiling) keep their meaning.
//
// Differences at this boundary:
//   * no GL VBOs: the models own their float4 buffers on the GPU; devicePositions() and
//     readState() stand in for the shared VBO;
//   * the JSON parameter tree is kept as text (getInputJson / resetInputJson /
//     updateInputJson, physics/Model.hpp:154-174, with the library's reader for the
//     reference's JSON library); typed setters (setFluidParams, setBoidsParams, ...) sit
//     beside it and change the model without touching the tree, and inputJson() writes the
//     model's current parameters as a tree;
//   * errors from the C ABI are thrown as std::runtime_error, like the reference's JSON and
//     program-build failures (Fluids.cpp:249, Context.cpp:279).
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "rtp.h"

namespace Geometry {
enum class Dimension { dim2D, dim3D };  // utils/Geometry.hpp:11-15
// utils/Geometry.hpp:42-47 (size_t there; float here so boxes need not be integral).
struct BoxSize3D {
  float x, y, z;
};
struct GridRes3D {
  int x, y, z;
};
static constexpr BoxSize3D BOX_SIZE_3D = { 10.0f, 10.0f, 10.0f };  // Geometry.hpp:50
static constexpr GridRes3D GRID_RES_3D = { 30, 30, 30 };          // Geometry.hpp:53
}  // namespace Geometry

namespace Utils {
// utils/Parameters.hpp:17-39
enum PhysicsCase {
  BOIDS_BEGIN, BOIDS_SMALL, BOIDS_MEDIUM, BOIDS_LARGE, BOIDS_XLARGE, BOIDS_END,
  FLUIDS_BEGIN, FLUIDS_DAM, FLUIDS_BOMB, FLUIDS_DROP, FLUIDS_END,
  CLOUDS_BEGIN, CLOUDS_CUMULUS, CLOUDS_HOMOGENEOUS, CLOUDS_END,
  CASE_INVALID = -1,
};
// utils/Parameters.hpp:62-72
enum NbParticles {
  P512 = 1 << 9, P1K = 1 << 10, P4K = 1 << 12, P8K = 1 << 13,
  P16K = 1 << 14, P32K = 1 << 15, P65K = 1 << 16, P130K = 1 << 17
};
// 2-D and 3-D lattice subdivisions of utils/Parameters.hpp:89-98.
std::array<int, 2> GetNbParticlesSubdiv2D(NbParticles nb);
std::array<int, 3> GetNbParticlesSubdiv3D(NbParticles nb);
}  // namespace Utils

namespace Physics {

enum ModelType { BOIDS = 0, FLUIDS = 1, CLOUDS = 2 };  // physics/Model.hpp:16-21

enum class Boundary { BouncingWall, CyclicWall };  // physics/Model.hpp:38-44

// physics/Model.hpp:60-73
struct ModelParams {
  size_t currNbParticles = 0;
  size_t maxNbParticles = 0;
  Geometry::BoxSize3D boxSize = { 0, 0, 0 };
  Geometry::GridRes3D gridRes = { 0, 0, 0 };
  float velocity = 0.0f;
  unsigned int particlePosVBO = 0;  // unused: buffers live on the GPU
  unsigned int particleColVBO = 0;  // unused
  unsigned int cameraVBO = 0;       // unused
  unsigned int gridVBO = 0;         // unused
  Geometry::Dimension dimension = Geometry::Dimension::dim3D;
  Utils::PhysicsCase pCase = Utils::PhysicsCase::CASE_INVALID;
  int device = 0;                   // HIP device ordinal
  float gravity = 9.81f;            // define.cl:8 (fluids)
};

class Model;
// physics/Model.cpp:10-24.
std::unique_ptr<Model> CreateModel(ModelType type, ModelParams params);

// physics/Model.hpp:81-216 minus the displayable-quantity registry.
class Model {
 public:
  explicit Model(const ModelParams& params);
  virtual ~Model() = default;

  size_t maxNbParticles() const { return m_maxNbParticles; }
  void setNbParticles(size_t nbSelParticles);
  size_t nbParticles() const { return m_currNbParticles; }

  void setDimension(Geometry::Dimension dim)
  {
    m_dimension = dim;
    reset();
  }
  Geometry::Dimension dimension() const { return m_dimension; }

  void setBoundary(Boundary boundary) { m_boundary = boundary; }
  Boundary boundary() const { return m_boundary; }

  virtual void update() = 0;
  virtual void reset() = 0;

  bool isInit() const { return m_init; }
  void pause(bool pause);
  bool onPause() const { return m_pause; }

  virtual std::array<float, 3> targetPos() const { return { 0.0f, 0.0f, 0.0f }; }
  virtual bool isTargetActivated() const { return false; }
  virtual bool isTargetVisible() const { return false; }

  virtual bool isProfilingEnabled() const { return m_profiling; }
  virtual void enableProfiling(bool enable);
  virtual bool isUsingIGPU() const { return false; }

  void setCase(Utils::PhysicsCase caseType) { m_case = caseType; }
  Utils::PhysicsCase getCase() const { return m_case; }

  // Replace the positions/velocities of the first `count` particles (float4 each) and
  // reset the grid, as the reference's init functions do.
  virtual void loadState(const float* pos4, const float* vel4, size_t count) = 0;
  virtual void readState(float* pos4, float* vel4, size_t count) const = 0;
  // Device pointer of the float4 position buffer (the GL VBO's role).
  virtual void* devicePositions() const = 0;
  // The rendering steps at the end of every update (grid detector, colour, camera-distance
  // sort; Fluids.cpp:458-468, Boids.cpp:375-381), off until enabled, and the camera
  // position they sort by (the reference's u_cameraPos buffer, Model.hpp:69).
  virtual void enableRendering(bool enable) = 0;
  virtual void setCameraPos(float x, float y, float z) = 0;
  // Device pointer of the float4 colour buffer p_col (the colour VBO's role), valid until
  // the next update; the rendering steps must be on.
  virtual void* deviceColors() const = 0;
  // The parameter tree of the JSON channel (physics/Model.hpp:154-172) as text: the
  // "Fluids" or "Boids" tree of the reference's widget (ui/PhysicsWidget.cpp:155-159 reads
  // it with getInputJson and sends edits with updateInputJson).  reset() puts the
  // reference's initial tree back and applies it, as Fluids::reset / Boids::reset do.
  std::string getInputJson() const { return m_inputJson; }
  void resetInputJson(const std::string& newJson) { m_inputJson = newJson; }
  // No-op when newJson equals the tree as a JSON value (json::diff empty); otherwise the
  // tree becomes tree.merge_patch(newJson) and is applied with updateModelWithInputJson.
  void updateInputJson(const std::string& newJson);
  // updateModelWithInputJson (physics/Model.hpp:174, OclModel.hpp:48-54): applies a whole
  // tree to the model (transferJsonInputsToModel) without storing it.
  virtual void updateModelWithInputJson(const std::string& json) = 0;
  // The model's current parameters written as a tree (typed setters included).
  virtual std::string inputJson() const = 0;
  // Per-kernel timings collected while profiling was enabled.
  virtual std::vector<rtp_kernel_time> kernelTimes() const = 0;

 protected:
  virtual void applyNbParticles() = 0;
  virtual void applyPause() = 0;
  virtual void applyProfiling() = 0;

  bool m_init = false;
  bool m_pause = false;
  bool m_profiling = false;
  size_t m_maxNbParticles;
  size_t m_currNbParticles;
  Geometry::BoxSize3D m_boxSize;
  Geometry::GridRes3D m_gridRes;
  size_t m_nbCells;
  Geometry::Dimension m_dimension;
  Utils::PhysicsCase m_case;
  Boundary m_boundary = Boundary::BouncingWall;
  int m_device = 0;

 private:
  std::string m_inputJson;  // the JSON channel's tree (physics/Model.hpp:214)
};

namespace HIP {

// Position based fluids (physics/ocl/Fluids.hpp:35-60) on librtp.
class Fluids : public Model {
 public:
  explicit Fluids(const ModelParams& params);
  ~Fluids() override;

  void update() override;
  void reset() override;

  void setFluidParams(const rtp_fluid_params& p);
  rtp_fluid_params fluidParams() const;
  void setNbJacobiIters(uint32_t iters);
  uint32_t nbJacobiIters() const { return m_nbJacobiIters; }

  void loadState(const float* pos4, const float* vel4, size_t count) override;
  void readState(float* pos4, float* vel4, size_t count) const override;
  void* devicePositions() const override;
  std::vector<rtp_kernel_time> kernelTimes() const override;
  // max |rho/rho0 - 1|, mean |rho/rho0 - 1|, max compression of the last update.
  std::array<float, 3> densityError() const;
  void enableRendering(bool enable) override;
  void setCameraPos(float x, float y, float z) override;
  void updateModelWithInputJson(const std::string& json) override;
  std::string inputJson() const override;
  void* deviceColors() const override;
  rtp_fluids* handle() const { return m_h; }

 protected:
  void applyNbParticles() override;
  void applyPause() override;
  void applyProfiling() override;

 private:
  rtp_fluids* m_h = nullptr;
  uint32_t m_nbJacobiIters = 2;  // Fluids.cpp:81
};

// 3-D boids (physics/ocl/Boids.hpp:35-67) on librtp.
class Boids : public Model {
 public:
  explicit Boids(const ModelParams& params);
  ~Boids() override;

  void update() override;
  void reset() override;

  void setBoidsParams(const rtp_boids_params& p);
  void setTimeStep(float dt);
  // The target rule (Boids.cpp:351-358, TargetParams Boids.hpp:22-26): active or not, where
  // the target is for the next updates, its radius of effect and sign (1 attracts, -1
  // repels).  The reference moves the target itself each frame (Target::updatePos); here
  // the caller does, by calling this again.
  void setTarget(bool active, std::array<float, 3> pos, float radius = 2.0f, int sign = 1);
  std::array<float, 3> targetPos() const override { return m_targetPos; }
  bool isTargetActivated() const override { return m_targetActive; }
  // Boids.hpp:42: the "Show" switch of the tree while the target is active (the app draws
  // the target when set, app/ParticleSystemApp.cpp:380).
  bool isTargetVisible() const override { return m_targetVisible; }

  void loadState(const float* pos4, const float* vel4, size_t count) override;
  void readState(float* pos4, float* vel4, size_t count) const override;
  void* devicePositions() const override;
  std::vector<rtp_kernel_time> kernelTimes() const override;
  void enableRendering(bool enable) override;
  void setCameraPos(float x, float y, float z) override;
  void updateModelWithInputJson(const std::string& json) override;
  std::string inputJson() const override;
  void* deviceColors() const override;
  rtp_boids* handle() const { return m_h; }

 protected:
  void applyNbParticles() override;
  void applyPause() override;
  void applyProfiling() override;

 private:
  void readTarget();

  rtp_boids* m_h = nullptr;
  bool m_targetActive = false;
  bool m_targetVisible = false;
  std::array<float, 3> m_targetPos{ 0.0f, 0.0f, 0.0f };
};

// The Clouds model (physics/ocl/Clouds.hpp:50-80) on librtp: position based fluids coupled
// to a temperature, vapour and cloud density per particle.  The app builds it with the y
// extent of the box and grid doubled (app/ParticleSystemApp.cpp:300-304); its cases are
// CLOUDS_CUMULUS and CLOUDS_HOMOGENEOUS.  No rendering steps: enableRendering(true),
// setCameraPos and deviceColors throw.
class Clouds : public Model {
 public:
  explicit Clouds(const ModelParams& params);
  ~Clouds() override;

  void update() override;
  void reset() override;

  void setFluidParams(const rtp_fluid_params& p);
  void setCloudParams(const rtp_cloud_params& q);
  rtp_fluid_params fluidParams() const;
  rtp_cloud_params cloudParams() const;
  void setNbJacobiIters(uint32_t iters);
  uint32_t nbJacobiIters() const { return m_nbJacobiIters; }
  // The displayable quantities of Clouds.cpp:161-173 by name ("Particle ID", "Vapor Density",
  // "Cloud Density", "Net Force", "Temperature"), one float per active particle.
  std::vector<float> quantity(const std::string& name) const;

  void loadState(const float* pos4, const float* vel4, size_t count) override;
  void readState(float* pos4, float* vel4, size_t count) const override;
  void* devicePositions() const override;
  std::vector<rtp_kernel_time> kernelTimes() const override;
  void enableRendering(bool enable) override;
  void setCameraPos(float x, float y, float z) override;
  void updateModelWithInputJson(const std::string& json) override;
  std::string inputJson() const override;
  void* deviceColors() const override;
  rtp_clouds* handle() const { return m_h; }

 protected:
  void applyNbParticles() override;
  void applyPause() override;
  void applyProfiling() override;

 private:
  rtp_clouds* m_h = nullptr;
  uint32_t m_nbJacobiIters = 1;  // Clouds.cpp:109, before reset() applies the tree (2)
};

}  // namespace HIP
}  // namespace Physics
