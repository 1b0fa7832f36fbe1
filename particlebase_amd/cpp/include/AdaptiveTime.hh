// AdaptiveTime.hh -- adaptive time step with rollback (reference
// AdaptiveTime.hh:10-104): grow dt by growthratio after an all-slow step
// (then wait 1 step), on a too-fast step roll the leap-frog state back two
// half steps over two calls and divide dt by decayratio (then wait 3 steps).
#pragma once
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <iostream>
#include <vector>
#include "pbamd_env.hh"

template <class T>
class AdaptiveTime {
public:
  AdaptiveTime() : growthratio(std::sqrt(2.0)), decayratio(4.0) {}
  AdaptiveTime(double gr, double dc) : growthratio(gr), decayratio(dc) {}
  void init(real dt0) { deltaT = dt0; }
  real operator()(void) const { return thistime; }
  bool isRollbacking(void) const { return rollback_in_progress; }
  real currentDeltaT(void) const { return deltaT; }

  // res: inspectVelocity results of every device (1 fast, 0 ok, -1 slow)
  uint32_t Progress(T &particles, const std::vector<int> &res, uint32_t j) {
    const int ndev = particles.nDevices();
    if (!rollback_in_progress) {
      thistime += deltaT;
      ++j;
      const int resmax = *std::max_element(res.begin(), res.end());
      if (resmax == 1) {
        for (int i = 0; i < ndev; ++i) particles[i].rollback(deltaT);
        rollback_in_progress = true;
      } else if ((resmax == -1) && (wait_change_deltaT == 0)) {
        for (int i = 0; i < ndev; ++i) particles[i].calcVinit((3.0 - growthratio) * deltaT);
        deltaT *= growthratio;
        wait_change_deltaT = 1;
        if (statOutput)
          std::cerr << std::endl << "DeltaT(+):" << thistime << ":" << j << ":" << deltaT << std::endl;
      }
      if (wait_change_deltaT > 0) --wait_change_deltaT;
    } else {
      for (int i = 0; i < ndev; ++i) particles[i].rollback2(deltaT);
      rollback_in_progress = false;
      thistime += -deltaT;
      deltaT /= decayratio;
      wait_change_deltaT = 3;
      if (statOutput)
        std::cerr << std::endl << "DeltaT(-):" << thistime << ":" << j << ":" << deltaT << std::endl;
    }
    return j;
  }
  void PrintStat(uint32_t j = 0) {
    if (statOutput) std::cerr << "DeltaT(0):" << thistime << ":" << j << ":" << deltaT << std::endl;
  }
  bool statOutput = false;

private:
  real deltaT = 0.0;
  double thistime = 0.0;
  bool rollback_in_progress = false;
  uint32_t wait_change_deltaT = 0;
  const double growthratio;
  const double decayratio;
};
