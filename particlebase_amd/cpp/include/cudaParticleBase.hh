// cudaParticleBase.hh -- particle state owner (reference cudaParticleBase.hh:
// 29-226, cudaParticleBase.cu).  The device state lives in a pb_sys (FP64
// r, v, a, F, 1/m, type, movable flag); the methods keep the reference's
// names, argument meaning and call order.
#pragma once
#include <algorithm>
#include <vector>
#include "ParticleBase.hh"
#include "Serialization.hh"
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

  // cudaParticleBase::setup, cudaParticleBase.cu:18-66 (virtual so that a
  // restart sets up the most derived class)
  virtual void setup(int n) {
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

  // checkpoint / restart (CUDAenv::writeSerialization / readSerialization,
  // CUDAenv.hh:268-285): every class writes its own fields after those of
  // its bases, in the reference's order
  virtual void save(pbamd::OArchive &a) const { saveBase(a); }
  virtual void load(const pbamd::IArchive &a) { loadBase(a); }
  void writeSerialization(const std::string &fn) const {
    pbamd::OArchive a(fn);
    save(a);
  }
  void readSerialization(const std::string &fn) {
    pbamd::IArchive a(fn);
    load(a);
  }

  virtual void calcForce(void) = 0;
  void TimeEvolution(real) {}

  // cudaFrame + cudaParticleBase fields (cudaFrame.hh:95-101,
  // cudaParticleBase.hh:228-354), FP64 as the engine keeps them
  void saveBase(pbamd::OArchive &a) const {
    a.put("cudaFrame.cellsize", cell, 9);
    a.scalar("cudaFrame.timestep", timestep);
    a.scalar("cudaParticleBase.particleNumN", N);
    std::vector<double> x(3 * (size_t)N);
    const std::pair<const char *, int> vecs[4] = {{"positionR", PB_FIELD_R},
                                                  {"velocityV", PB_FIELD_V},
                                                  {"accelerationA", PB_FIELD_A},
                                                  {"forceF", PB_FIELD_F}};
    for (const auto &f : vecs) {
      pbamd::check(pb_sys_get(sys, f.second, x.data()), "writeSerialization");
      a.put(std::string("cudaParticleBase.") + f.first, x);
    }
    std::vector<double> mi(N);
    pbamd::check(pb_sys_get(sys, PB_FIELD_MINV, mi.data()), "writeSerialization");
    a.put("cudaParticleBase.massInverseMinv", mi);
    std::vector<uint16_t> t(N);
    pbamd::check(pb_sys_get(sys, PB_FIELD_TYPE, t.data()), "writeSerialization");
    a.put("cudaParticleBase.typeID", t);
    std::vector<uint8_t> mv(N);
    pbamd::check(pb_sys_get(sys, PB_FIELD_MOVE, mv.data()), "writeSerialization");
    a.put("cudaParticleBase.moveFlag", mv);
  }
  void loadBase(const pbamd::IArchive &a) {
    const uint32_t n = a.scalar<uint32_t>("cudaParticleBase.particleNumN");
    if (!sys) setup((int)n);
    else if (n != N) throw std::runtime_error("readSerialization: particle count differs");
    const std::vector<real> c = a.get<real>("cudaFrame.cellsize");
    if (c.size() != 9) throw std::runtime_error("readSerialization: cellsize");
    setCell(c.data());
    timestep = a.scalar<uint32_t>("cudaFrame.timestep");
    const auto r = a.get<double>("cudaParticleBase.positionR");
    const auto v = a.get<double>("cudaParticleBase.velocityV");
    const auto ac = a.get<double>("cudaParticleBase.accelerationA");
    const auto f = a.get<double>("cudaParticleBase.forceF");
    const auto mi = a.get<double>("cudaParticleBase.massInverseMinv");
    const auto t = a.get<uint16_t>("cudaParticleBase.typeID");
    const auto mv = a.get<uint8_t>("cudaParticleBase.moveFlag");
    const size_t n3 = 3 * (size_t)N;
    if (r.size() != n3 || v.size() != n3 || ac.size() != n3 || f.size() != n3 || mi.size() != N ||
        t.size() != N || mv.size() != N)
      throw std::runtime_error("readSerialization: array sizes");
    std::vector<double> m(N);
    std::vector<uint32_t> t32(N);
    std::vector<uint8_t> fixed(N);
    for (size_t i = 0; i < N; ++i) {
      m[i] = 1.0 / mi[i];
      t32[i] = t[i];
      fixed[i] = mv[i] ? 0 : 1;
    }
    pbamd::check(pb_sys_import(sys, r.data(), v.data(), ac.data(), m.data(), t32.data(), fixed.data()),
                 "readSerialization");
    // 1/m exactly as saved, and the force of the last step
    pbamd::check(pb_sys_set(sys, PB_FIELD_MINV, mi.data()), "readSerialization");
    pbamd::check(pb_sys_set(sys, PB_FIELD_F, f.data()), "readSerialization");
  }

  void getPosition(void) { fetch(PB_FIELD_R); }
  void getAcceleration(void) { fetch(PB_FIELD_A); }
  void getForce(void) { fetch(PB_FIELD_F); }
  void getVelocity(void) { fetch(PB_FIELD_V); }
  void getTypeID(void) {
    std::vector<uint16_t> t(N);
    pbamd::check(pb_sys_get(sys, PB_FIELD_TYPE, t.data()), "getTypeID");
    std::lock_guard<std::mutex> lk(mutTMP);
    if (TMP2.size() != N) TMP2.resize(N);   // a selected-range dump may have shrunk it
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
