// GaussianThermo.hh -- Gaussian isokinetic thermostat for velocity Verlet
// (reference GaussianThermo.hh:14-90, GaussianThermo.cu).  The reference
// keeps an FP64 copy vd of the velocities; the engine's velocities are FP64
// already, so calcMV2 / scaleTemp / adjustVelocities of cudaParticleMD act
// on the same state.  TimeEvolution solves the implicit velocity equation on
// the device (pb_gt_step): its closed-form solution, or the Newton loop's
// one-pass result where the loop would stop after one pass under the
// threshold set by setE.
#pragma once
#include "cudaParticleMD.hh"

class GaussianThermo : public virtual cudaParticleMD {
protected:
  real _e = 1e-8f;   // Newton-Raphson threshold (GaussianThermo() : _e(1e-8))

public:
  GaussianThermo() = default;

  void setup(int n) override { cudaParticleMD::setup(n); }

  // GaussianThermo fields (GaussianThermo.hh:92-120): xi and _e; velocityD
  // is the engine's FP64 velocity, stored by cudaParticleBase
  void save(pbamd::OArchive &a) const override {
    cudaParticleMD::save(a);
    double xi = 0;
    pbamd::check(pb_gt_xi(sys, nullptr, &xi, nullptr), "GaussianThermo::save");
    a.scalar("GaussianThermo.xi", xi);
    a.scalar("GaussianThermo._e", _e);
  }
  void load(const pbamd::IArchive &a) override {
    cudaParticleMD::load(a);
    setXi(a.scalar<double>("GaussianThermo.xi"));
    _e = a.scalar<real>("GaussianThermo._e");
  }

  // v(t) = v(t-dt) + dt/2m (F(t)+F(t-dt)) - (v(t-dt) xi(t-dt) + v(t) xi(t)) dt/2
  // xi(t) = sum F(t).v(t) / sum m v(t)^2
  // r(t+dt) = r(t) + dt v(t) + dt^2/2m F(t) - v(t) xi(t) dt^2/2
  void TimeEvolution(real dt) { pbamd::check(pb_gt_step(sys, dt, _e), "GaussianThermo::TimeEvolution"); }

  void setE(real e0) { _e = e0; }

  // the thermostat variable xi (GaussianThermo.hh:16) for restarts
  double getXi(void) {
    double xi = 0;
    pbamd::check(pb_gt_xi(sys, nullptr, &xi, nullptr), "GaussianThermo::getXi");
    return xi;
  }
  void setXi(double xi) { pbamd::check(pb_gt_xi(sys, &xi, nullptr, nullptr), "GaussianThermo::setXi"); }
  // steps whose implicit equation had no solution (the reference exits there)
  uint32_t failedSteps(void) {
    uint32_t b = 0;
    pbamd::check(pb_gt_xi(sys, nullptr, nullptr, &b), "GaussianThermo::failedSteps");
    return b;
  }

  // scaleTemp with (3N-1)/3N in FP64 (GaussianThermo.cu:89-96)
  real scaleTemp(real Temp) {
    const real T0 = calcTemp();
    const double s = std::sqrt((double)(3 * N - 1) / (double)(3 * N) * Temp / T0);
    pbamd::check(pb_scale_velocities(sys, s), "GaussianThermo::scaleTemp");
    return (real)s;
  }
};
