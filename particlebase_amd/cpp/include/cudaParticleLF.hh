// cudaParticleLF.hh -- leap-frog mixin (reference cudaParticleLF.hh/.cu).
#pragma once
#include "cudaParticleBase.hh"

class cudaParticleLF : public virtual cudaParticleBase {
public:
  // v(-dt/2) = v(0) - dt/2 F(0)/m after computing F(0) (cudaParticleLF.cu:4-12)
  void calcVinit(real dt) {
    cudaParticleBase::clearForce();
    calcForce();
    pbamd::check(pb_lf_vinit(sys, dt), "cudaParticleLF::calcVinit");
  }
  void TimeEvolution(real dt) { pbamd::check(pb_lf_step(sys, dt), "cudaParticleLF::TimeEvolution"); }
  void rollback(real dt) { pbamd::check(pb_lf_rollback(sys, dt), "cudaParticleLF::rollback"); }
};
