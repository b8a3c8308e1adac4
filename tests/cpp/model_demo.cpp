// Drives the C++ Physics::Model interface the way the reference app does
// (app/ParticleSystemApp.cpp:275-315 CreateModel, :377 update) and dumps the state before and
// after `steps` updates so a test can replay the same steps on the CPU oracle.
//
//   model_demo fluids|boids|fluids2d|boids2d|boidscyclic|boidstarget|fluidsjson|fluidsinput|boidsinput
//              |clouds|cloudsinput <out.bin> <steps>
//
// The 2d kinds switch the model to 2-D (Model::setDimension, which resets it); boidscyclic
// selects the cyclic wall (Model::setBoundary); boidstarget turns the target rule on with a
// target at (1, 0.5, 0), radius 3, attracting; fluidsjson applies a parameter tree with three
// Jacobi iterations and vorticity confinement off (Model::updateModelWithInputJson) and
// checks that inputJson() gives it back.  fluidsinput edits the model's tree the way the
// reference's widget does (getInputJson, then updateInputJson, ui/PhysicsWidget.cpp:155-159):
// an unchanged tree is a no-op, a partial tree is merged (three iterations, vorticity off),
// and reset() brings the initial tree back; the state dumped is after the edit.
// boidsinput parks the target at (1, 0.5, 0) while it is off and then switches it on
// through the tree (radius 2, attracting, shown).  clouds builds the Clouds model as the app
// does (box and grid doubled in y, the cumulus case, app/ParticleSystemApp.cpp:300-304);
// cloudsinput edits its tree (temperature smoothing off, three iterations) and checks the
// typed parameters and inputJson().
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "Model.hpp"

static void put(FILE* f, const void* p, size_t n) { std::fwrite(p, 1, n, f); }

int main(int argc, char** argv)
{
  if (argc < 4) {
    std::fprintf(stderr, "usage: model_demo fluids|boids|fluids2d|boids2d|boidscyclic|boidstarget|fluidsjson|fluidsinput|boidsinput|clouds|cloudsinput out.bin steps\n");
    return 2;
  }
  const bool fluids = std::strncmp(argv[1], "fluids", 6) == 0;
  const bool clouds = std::strncmp(argv[1], "clouds", 6) == 0;
  const bool cloudsInput = std::strcmp(argv[1], "cloudsinput") == 0;
  const bool dim2 = std::strstr(argv[1], "2d") != nullptr;
  const bool cyclic = std::strcmp(argv[1], "boidscyclic") == 0;
  const bool target = std::strcmp(argv[1], "boidstarget") == 0;
  const bool withJson = std::strcmp(argv[1], "fluidsjson") == 0;
  const bool fluidsInput = std::strcmp(argv[1], "fluidsinput") == 0;
  const bool boidsInput = std::strcmp(argv[1], "boidsinput") == 0;
  const int steps = std::atoi(argv[3]);
  Physics::ModelParams p;
  p.maxNbParticles = Utils::P130K;            // app: largest entry of ALL_NB_PARTICLES
  p.boxSize = Geometry::BOX_SIZE_3D;
  p.gridRes = Geometry::GRID_RES_3D;
  p.dimension = Geometry::Dimension::dim3D;
  p.pCase = fluids ? Utils::FLUIDS_DAM : Utils::BOIDS_LARGE;
  if (clouds) {
    p.pCase = Utils::CLOUDS_CUMULUS;
    p.boxSize.y *= 2;
    p.gridRes.y *= 2;
  }
  try {
    auto model = Physics::CreateModel(
        clouds ? Physics::CLOUDS : (fluids ? Physics::FLUIDS : Physics::BOIDS), p);
    if (dim2)
      model->setDimension(Geometry::Dimension::dim2D);  // app/ParticleSystemApp.cpp:428-431
    if (cyclic)
      model->setBoundary(Physics::Boundary::CyclicWall);
    if (withJson) {
      const std::string tree =
          "{\"Fluids\": {\"Rest Density\": [450.0, 10.0, 1000.0], \"Relax CFM\": [600.0, 100.0, 1000.0],"
          " \"Time Step\": [0.01, 0.0001, 0.02], \"Nb Jacobi Iterations\": [3, 1, 6],"
          " \"Artificial Pressure\": {\"Enable##Pressure\": true, \"Coefficient##Pressure\": [0.001, 0.0, 0.001],"
          " \"Radius\": [0.006, 0.001, 0.015], \"Exp\": [4, 1, 6]},"
          " \"Vorticity Confinement\": {\"Enable##Vorticity\": false, \"Coefficient##Vorticity\": [0.0004, 0.0, 0.001],"
          " \"xSPH Viscosity Coefficient\": [0.0001, 0.0, 0.001]}}}";
      model->updateModelWithInputJson(tree);
      const std::string back = model->inputJson();
      if (back.find("\"Nb Jacobi Iterations\": [3,") == std::string::npos ||
          back.find("\"Enable##Vorticity\": false") == std::string::npos)
        return 6;
    }
    if (withJson) {  // the rendering steps through the Model surface as well
      model->enableRendering(true);
      model->setCameraPos(0.0f, 2.0f, 30.0f);
      if (!model->deviceColors())
        return 7;
    }
    if (fluidsInput) {
      auto* fl = dynamic_cast<Physics::HIP::Fluids*>(model.get());
      if (!fl)
        return 10;
      const std::string init = model->getInputJson();
      if (init.find("\"Nb Jacobi Iterations\": [2, 1, 6]") == std::string::npos || fl->nbJacobiIters() != 2)
        return 11;
      // a change the tree does not carry, then the unchanged tree: updateInputJson must not
      // re-apply it (json::diff is empty), so the typed change stays
      fl->setNbJacobiIters(5);
      model->updateInputJson(init);
      if (fl->nbJacobiIters() != 5)
        return 12;
      // the same tree with its keys in another order is still equal
      model->updateInputJson(
          "{\"Fluids\": {\"Vorticity Confinement\": {\"xSPH Viscosity Coefficient\": [0.0001, 0, 0.001],"
          " \"Coefficient##Vorticity\": [0.0004, 0, 0.001], \"Enable##Vorticity\": true},"
          " \"Artificial Pressure\": {\"Exp\": [4, 1, 6], \"Radius\": [0.006, 0.001, 0.015],"
          " \"Coefficient##Pressure\": [0.001, 0, 0.001], \"Enable##Pressure\": true},"
          " \"Nb Jacobi Iterations\": [2, 1, 6], \"Time Step\": [0.01, 0.0001, 0.02],"
          " \"Relax CFM\": [600, 100, 1000], \"Rest Density\": [450, 10, 1000]}}");
      if (fl->nbJacobiIters() != 5)
        return 13;
      // a partial tree merges into the stored one and the merged tree is applied
      model->updateInputJson(
          "{\"Fluids\": {\"Nb Jacobi Iterations\": [3, 1, 6], \"Vorticity Confinement\": {\"Enable##Vorticity\": false}}}");
      const std::string merged = model->getInputJson();
      if (fl->nbJacobiIters() != 3 || fl->fluidParams().isVorticityConfEnabled != 0 ||
          fl->fluidParams().restDensity != 450.0f ||
          merged.find("\"Rest Density\": [450, 10, 1000]") == std::string::npos ||
          merged.find("\"Enable##Vorticity\": false") == std::string::npos)
        return 14;
      // reset brings the initial tree and its parameters back
      model->reset();
      if (fl->nbJacobiIters() != 2 || fl->fluidParams().isVorticityConfEnabled != 1 ||
          model->getInputJson() != init)
        return 15;
      model->updateInputJson(merged);
      if (fl->nbJacobiIters() != 3)
        return 16;
      // an iteration count past the library's cap reads back as the count that runs
      std::string big = merged;
      const size_t at = big.find("[3, 1, 6]");
      if (at == std::string::npos)
        return 17;
      fl->updateModelWithInputJson(big.replace(at, 9, "[2000000, 1, 6]"));
      if (fl->nbJacobiIters() != 1000)
        return 18;
      fl->setNbJacobiIters(3);
    }
    if (cloudsInput) {
      auto* cl = dynamic_cast<Physics::HIP::Clouds*>(model.get());
      if (!cl)
        return 30;
      if (cl->nbJacobiIters() != 2 || cl->cloudParams().isTempSmoothingEnabled != 1 ||
          cl->cloudParams().restDensity != 450.0f || cl->fluidParams().restDensity != 450.0f)
        return 31;
      model->updateInputJson(
          "{\"Fluids\": {\"Nb Jacobi Iterations\": [3, 1, 6]},"
          " \"Clouds\": {\"Enable Temperature Smoothing\": false, \"Wind Coefficient\": [0.5, 0, 1]}}");
      const rtp_cloud_params q = cl->cloudParams();
      if (cl->nbJacobiIters() != 3 || q.isTempSmoothingEnabled != 0 || q.windCoeff != 0.5f ||
          q.groundHeatCoeff != 10.0f)
        return 32;
      const std::string back = model->inputJson();
      if (back.find("\"Enable Temperature Smoothing\": false") == std::string::npos ||
          back.find("\"Wind Coefficient\": [0.5, 0, 1]") == std::string::npos)
        return 33;
      if (cl->quantity("Temperature").size() != model->nbParticles())
        return 34;
    }
    if (boidsInput) {
      auto* boids = dynamic_cast<Physics::HIP::Boids*>(model.get());
      if (!boids)
        return 20;
      boids->setTarget(false, { 1.0f, 0.5f, 0.0f }, 3.0f, -1);
      if (model->isTargetActivated() || model->isTargetVisible() || model->targetPos()[0] != 1.0f)
        return 21;
      model->updateInputJson("{\"Boids\": {\"Target\": {\"Enable##Target\": true}}}");
      if (!model->isTargetActivated() || !model->isTargetVisible() || model->targetPos()[1] != 0.5f)
        return 22;
      model->updateInputJson("{\"Boids\": {\"Target\": {\"Show\": false}}}");
      if (!model->isTargetActivated() || model->isTargetVisible())
        return 23;
    }
    if (target) {
      auto* boids = dynamic_cast<Physics::HIP::Boids*>(model.get());
      if (!boids)
        return 4;
      boids->setTarget(true, { 1.0f, 0.5f, 0.0f }, 3.0f, 1);
      if (!model->isTargetActivated() || model->targetPos()[1] != 0.5f)
        return 5;
    }
    const size_t n = model->nbParticles();
    std::vector<float> pos0(4 * n), vel0(4 * n), pos1(4 * n), vel1(4 * n);
    model->readState(pos0.data(), vel0.data(), n);
    model->enableProfiling(true);
    for (int s = 0; s < steps; ++s)
      model->update();
    model->readState(pos1.data(), vel1.data(), n);
    size_t launches = 0;
    for (const auto& k : model->kernelTimes())
      launches += k.launches;
    FILE* f = std::fopen(argv[2], "wb");
    if (!f)
      return 3;
    const unsigned int hdr[3] = { (unsigned int)n, (unsigned int)steps, (unsigned int)model->maxNbParticles() };
    put(f, hdr, sizeof(hdr));
    put(f, pos0.data(), pos0.size() * 4);
    put(f, vel0.data(), vel0.size() * 4);
    put(f, pos1.data(), pos1.size() * 4);
    put(f, vel1.data(), vel1.size() * 4);
    std::fclose(f);
    std::printf("model_demo %s: %zu particles, %d updates, %zu profiled launches\n", argv[1], n,
                steps, launches);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "model_demo: %s\n", e.what());
    return 1;
  }
  return 0;
}
