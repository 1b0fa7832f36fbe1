// CUDAenv.hh -- single-GPU device environment (reference CUDAenv.hh:19-89).
// Every system here runs on one GPU with its full domain, so nDevices() is
// 1, setGPU accepts only device 0, and the exchange* calls of the
// replicated-data multi-GPU mode have nothing to exchange (no-ops).
#pragma once
#include <cstdint>
#include <iomanip>
#include <iostream>
#include <string>
#include "ExchangeMode.hh"
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
  // device limits (CUDAenv.hh:29-34).  getMemSize returns the full byte
  // count as 64 bits: the reference's uint32_t would truncate 288 GB.
  uint32_t getNumCpMP(void) const { return limits().lanes; }
  uint64_t getMemSize(void) const { return limits().mem; }
  uint32_t getNumTh(void) const { return limits().threads; }
  uint32_t getMaxGrid(void) const { return limits().grid; }
  uint32_t getShMem(void) const { return limits().lds; }
  // display (CUDAenv.hh:177-194): the same lines for the one device
  void display(std::ostream &o) const {
    const Limits L = limits();
    o << std::setw(5) << std::setprecision(5);
    o << getName() << std::endl
      << "  number of multiprocessor: " << getNumMP() << std::endl
      << "  number of cores / MP: " << L.lanes << std::endl
      << "  global memory size: " << static_cast<double>(L.mem) / 1024 / 1024 / 1024 << " [GB]"
      << std::endl
      << "  max threads per block: " << L.threads << std::endl
      << "  max block per grid: " << L.grid << std::endl
      << "  shared memory size: " << static_cast<double>(L.lds) / 1024 << " [KB]" << std::endl;
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
  // exchangeForceSelected (CUDAenv.hh:246-266)
  void exchangeForceSelected(const ExchangeMode = ExchangeMode::force) {}

private:
  struct Limits {
    uint64_t mem = 0;
    uint32_t threads = 0, grid = 0, lds = 0, lanes = 0;
  };
  Limits limits(void) const {
    Limits L;
    pbamd::check(pb_ctx_limits(pbamd::context(), &L.mem, &L.threads, &L.grid, &L.lds, &L.lanes),
                 "CUDAenv limits");
    return L;
  }
  T *P;
};
