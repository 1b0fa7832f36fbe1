// cudaParticleDEM.hh -- DEM spheres with rotation (reference
// cudaParticleDEM.hh/.cu + cudaParticleRotation.hh/.cu) over the engine's
// contact kernels: contact CSR with tangential history, Hertz normal force,
// Coulomb-capped tangential force, rolling friction, Euler rotation.
#pragma once
#include <vector>
#include "cudaCutoffBlock.hh"
#include "cudaParticleLF.hh"

class cudaParticleRotation : public virtual cudaParticleBase {
public:
  // cudaParticleRotation::setInertia (cudaParticleRotation.cu:35-61)
  void setInertia(const std::vector<real> &r0) {
    std::vector<double> r(r0.begin(), r0.end());
    pbamd::check(pb_rot_set_inertia(sys, r.data(), 0.0), "setInertia");
  }
  void setInertia(real r0_all) {
    pbamd::check(pb_rot_set_inertia(sys, nullptr, r0_all), "setInertia");
  }
  // cudaParticleRotation::TimeEvolution (cudaParticleRotation.cu:29-33)
  void TimeEvolution(real dt) { pbamd::check(pb_rot_step(sys, dt), "Rotation::TimeEvolution"); }
};

class cudaParticleDEM : public cudaParticleLF,
                        public cudaParticleRotation,
                        virtual public cudaCutoffBlock {
protected:
  uint32_t lastNNZ = 0;

public:
  bool autotunetimestep = false;   //!< true when AdaptiveTime drives dt (no Cundall-Strack abort)

  cudaParticleDEM() { SingleParticleBlock = true; }   // cudaParticleDEM.hh:37

  void setup(int n) { cudaParticleBase::setup(n); }

  // the one-argument form is the DEM force (cudaParticleDEM.hh:53-54)
  void calcForce(void) override {}
  void calcForce(real dt) {
    pbamd::check(pb_dem_calc_force(sys, dt, autotunetimestep ? 1 : 0, &lastNNZ),
                 "cudaParticleDEM::calcForce");
  }
  uint32_t numContacts(void) const { return lastNNZ; }

  // cudaParticleDEM::TimeEvolution: translation (LF) then rotation
  void TimeEvolution(real dt) {
    cudaParticleLF::TimeEvolution(dt);
    cudaParticleRotation::TimeEvolution(dt);
  }

  // both overloads of setDEMProperties (cudaParticleDEM.cu:221-252)
  void setDEMProperties(real E, real mu, real sigma, real gamma, real mu_r, real r0_all) {
    pbamd::check(pb_dem_set_properties(sys, E, mu, sigma, gamma, mu_r, nullptr, r0_all),
                 "setDEMProperties");
  }
  void setDEMProperties(real E, real mu, real sigma, real gamma, real mu_r,
                        const std::vector<real> &r0) {
    std::vector<double> r(r0.begin(), r0.end());
    pbamd::check(pb_dem_set_properties(sys, E, mu, sigma, gamma, mu_r, r.data(), 0.0),
                 "setDEMProperties");
  }

  void selectBlocks(void) {}
  void calcBlockID(void) { cudaCutoffBlock::calcBlockID(); }

  // cudaParticleDEM::rollback / rollback2 (cudaParticleDEM.cu:302-309)
  void rollback(real dt) { cudaParticleLF::rollback(dt); }
  void rollback2(real dt) { cudaParticleLF::calcVinit(dt / 2); }
};
