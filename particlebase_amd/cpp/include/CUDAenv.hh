// CUDAenv.hh -- single-GPU device environment (reference CUDAenv.hh:19-89).
// Every system here runs on one GPU with its full domain, so nDevices() is
// 1 and the exchange* calls of the replicated-data multi-GPU mode are
// refused.
#pragma once
#include <iostream>
#include <string>
#include "pbamd_env.hh"

template <typename T>
class CUDAenv {
public:
  CUDAenv() : P(new T()) { pbamd::context(); }
  ~CUDAenv() { delete P; }
  void setGPU(int i) {
    if (i != 0) {
      std::cerr << "CUDAenv: one GPU per process in this build" << std::endl;
      std::exit(1);
    }
  }
  std::string getName(void) const {
    char name[128];
    int cu = 0;
    pb_ctx_info(pbamd::context(), name, sizeof(name), &cu);
    return name;
  }
  uint32_t getNumMP(void) const {
    char name[8];
    int cu = 0;
    pb_ctx_info(pbamd::context(), name, sizeof(name), &cu);
    return (uint32_t)cu;
  }
  // launch shapes are chosen inside the engine; the fields are advisory
  template <typename U>
  void setThnum(U &p) {
    p.MPnum = (int)getNumMP();
    p.THnum1D = 256;
    p.THnum1DX = 256;
    p.THnum2D = 16;
    p.threadsMax = 1024;
    p.maxGrid = 0x7FFFFFFF;
  }
  int nDevices(void) { return 1; }
  T &operator[](int) { return *P; }
  void setup(void) { setThnum(*P); }
  // CUDAenv::readSerialization / writeSerialization (CUDAenv.hh:268-285)
  void readSerialization(const std::string &fn) { P->readSerialization(fn); }
  void writeSerialization(const std::string &fn) const { P->writeSerialization(fn); }
  void exchangeForce(void) {}
  void exchangeAccelerations(void) {}

private:
  T *P;
};
