// cudaParticleBase.hh -- particle state owner (reference cudaParticleBase.hh:
// 29-226, cudaParticleBase.cu).  The device state lives in a pb_sys (FP64
// r, v, a, F, 1/m, type, movable flag); the methods keep the reference's
// names, argument meaning and call order.
#pragma once
#include <algorithm>
#include <vector>
#include "ParticleBase.hh"
#include "cudaFrame.hh"

class cudaParticleBase : public cudaFrame {
protected:
  uint32_t N = 0;
  pb_sys *sys = nullptr;
  std::vector<double> hbuf;   // FP64 host staging

  void onCell() override {
    if (!sys) return;
    double c[6];
    for (int i = 0; i < 6; ++i) c[i] = cell[i];
    pbamd::check(pb_sys_set_cell(sys, c), "setCell");
  }

  void fetch(int field) {
    hbuf.resize(3 * (size_t)N);
    pbamd::check(pb_sys_get(sys, field, hbuf.data()), "get");
    std::lock_guard<std::mutex> lk(mutTMP);
    for (size_t i = 0; i < N; ++i) {
      TMP[i * 4] = (real)hbuf[3 * i];
      TMP[i * 4 + 1] = (real)hbuf[3 * i + 1];
      TMP[i * 4 + 2] = (real)hbuf[3 * i + 2];
      TMP[i * 4 + 3] = 0;
    }
  }

public:
  cudaParticleBase() = default;
  virtual ~cudaParticleBase() {
    waitPutTMP();
    if (sys) pb_sys_destroy(sys);
  }

  // cudaParticleBase::setup, cudaParticleBase.cu:18-66
  void setup(int n) {
    N = (uint32_t)n;
    TMP.resize(4 * (size_t)n);
    TMP2.resize(n);
    const uint32_t cnt = N;
    pbamd::check(pb_sys_create(pbamd::context(), 1, &cnt, &sys), "setup");
    if (cell[1] > cell[0]) onCell();
  }

  uint32_t size(void) const { return N; }
  pb_sys *handle(void) const { return sys; }

  // cudaParticleBase::import, cudaParticleBase.cu:68-117
  void import(const std::vector<ParticleBase> &P) {
    std::vector<double> r(3 * N), v(3 * N), a(3 * N), m(N);
    std::vector<uint32_t> t(N);
    std::vector<uint8_t> fx(N);
    for (size_t i = 0; i < N; ++i) {
      for (int k = 0; k < 3; ++k) {
        r[3 * i + k] = P[i].r[k];
        v[3 * i + k] = P[i].v[k];
        a[3 * i + k] = P[i].a[k];
      }
      m[i] = P[i].m;
      t[i] = P[i].type;
      fx[i] = P[i].isFixed ? 1 : 0;
    }
    pbamd::check(pb_sys_import(sys, r.data(), v.data(), a.data(), m.data(), t.data(), fx.data()),
                 "import");
  }

  virtual void calcForce(void) = 0;
  void TimeEvolution(real) {}

  void getPosition(void) { fetch(PB_FIELD_R); }
  void getAcceleration(void) { fetch(PB_FIELD_A); }
  void getForce(void) { fetch(PB_FIELD_F); }
  void getVelocity(void) { fetch(PB_FIELD_V); }
  void getTypeID(void) {
    std::vector<uint16_t> t(N);
    pbamd::check(pb_sys_get(sys, PB_FIELD_TYPE, t.data()), "getTypeID");
    for (size_t i = 0; i < N; ++i) TMP2[i] = t[i];
  }
  // FP64 copies of the state in the caller's particle order (no reference
  // counterpart: the reference exposes FP32 TMP only)
  std::vector<double> getState(int field) {
    std::vector<double> out(3 * (size_t)N);
    pbamd::check(pb_sys_get(sys, field, out.data()), "getState");
    return out;
  }
  const std::valarray<real> &tmp(void) const { return TMP; }

  void clearForce(void) { pbamd::check(pb_clear_force(sys, 0), "clearForce"); }
  void clearTmp3N(void) {}
  void calcAcceleration(void) { pbamd::check(pb_calc_acceleration(sys), "calcAcceleration"); }
  void treatPeriodicCondition(void) { pbamd::check(pb_treat_periodic(sys), "treatPeriodicCondition"); }
  void treatAbsoluteCondition(void) { pbamd::check(pb_treat_absolute(sys), "treatAbsoluteCondition"); }
  void treatRefrectCondition(void) { pbamd::check(pb_treat_reflect(sys), "treatRefrectCondition"); }
  void addForceX(real f) { pbamd::check(pb_add_force(sys, 0, f), "addForceX"); }
  void addForceY(real f) { pbamd::check(pb_add_force(sys, 1, f), "addForceY"); }
  void addForceZ(real f) { pbamd::check(pb_add_force(sys, 2, f), "addForceZ"); }
  void addAccelerationX(real a) { pbamd::check(pb_add_acceleration(sys, 0, a), "addAccelerationX"); }
  void addAccelerationY(real a) { pbamd::check(pb_add_acceleration(sys, 1, a), "addAccelerationY"); }
  void addAccelerationZ(real a) { pbamd::check(pb_add_acceleration(sys, 2, a), "addAccelerationZ"); }

  // cudaParticleBase::inspectVelocity, cudaParticleBase.cu:244-282
  int inspectVelocity(real vlim, real lim_u, real lim_l, real &_r, uint32_t &_r1,
                      bool debug = false) {
    int res = 0;
    double vr = 0;
    uint32_t im = 0;
    pbamd::check(pb_inspect_velocity(sys, vlim, lim_u, lim_l, &res, &vr, &im), "inspectVelocity");
    if (debug) {
      _r = (real)vr;
      _r1 = im;
    }
    return res;
  }
  int inspectVelocity(real vlim, real lim_u, real lim_l) {
    real r = 0;
    uint32_t i = 0;
    return inspectVelocity(vlim, lim_u, lim_l, r, i, false);
  }

  void dump3Narray(int field, std::ostream &o = std::cout) {
    fetch(field);
    o << std::endl << std::endl;
    for (size_t i = 0; i < N; ++i)
      o << i << "\t" << TMP[i * 4] << "\t" << TMP[i * 4 + 1] << "\t" << TMP[i * 4 + 2] << std::endl;
  }
};
