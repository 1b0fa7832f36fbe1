// ParticleBase.hh -- host import record, field for field as the reference's
// ParticleBase (ParticleBase.hh:10-18).
#pragma once
#include <cstdint>

class ParticleBase {
public:
  double r[3];      //!< position
  double m;         //!< mass
  double v[3];      //!< velocity
  bool isFixed;     //!< fixed particle flag
  double a[3];      //!< acceleration
  uint32_t type;    //!< particle type
};
