// cudaParticleVV.hh -- velocity-Verlet mixin (reference cudaParticleVV.hh/.cu):
// Fold aliases the acceleration array a (cudaParticleVV.cu:7).
#pragma once
#include "cudaParticleBase.hh"

class cudaParticleVV : public virtual cudaParticleBase {
public:
  void setup(int n) override { cudaParticleBase::setup(n); }
  // v(t) = v(t-dt) + dt/2m (F(t)+F(t-dt)); r(t+dt) = r + dt v(t) + dt^2/2m F(t)
  void TimeEvolution(real dt) { pbamd::check(pb_vv_step(sys, dt), "cudaParticleVV::TimeEvolution"); }
  // clears F and Fold (cudaParticleVV.cu:10-15)
  void clearForce(void) { pbamd::check(pb_clear_force(sys, 1), "cudaParticleVV::clearForce"); }
};
