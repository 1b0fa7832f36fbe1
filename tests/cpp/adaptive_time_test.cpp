// Host test of the AdaptiveTime state machine (AdaptiveTime.hh) with a mock
// particle set that records the calls it receives.  Prints one line per
// Progress call: "<j> <time> <dt> <rollbacking> <calls>".
#include <cstdio>
#include <string>
#include <vector>
#include "../../particlebase_amd/cpp/include/AdaptiveTime.hh"

struct Mock {
  std::string log;
  void rollback(real dt) { log += "R(" + std::to_string(dt) + ")"; }
  void rollback2(real dt) { log += "R2(" + std::to_string(dt) + ")"; }
  void calcVinit(real dt) { log += "V(" + std::to_string(dt) + ")"; }
};

struct Env {
  Mock m;
  int nDevices() { return 1; }
  Mock &operator[](int) { return m; }
};

int main() {
  Env env;
  AdaptiveTime<Env> t;
  t.init(1.0f);
  // scripted inspectVelocity results
  const int script[] = {0, -1, -1, -1, 1, 0, -1, -1, -1, -1, -1, 0};
  uint32_t j = 0;
  for (int r : script) {
    std::vector<int> res(1, r);
    env.m.log.clear();
    j = t.Progress(env, res, j);
    std::printf("%u %.9g %.9g %d %s\n", j, t(), (double)t.currentDeltaT(), t.isRollbacking() ? 1 : 0,
                env.m.log.empty() ? "-" : env.m.log.c_str());
  }
  return 0;
}
