// cudaSortedPositions.hh -- the sorted copy r_s / typeID_s of the reference
// (cudaSortedPositions.hh/.cu).  The engine keeps its own cell-ordered
// storage, so the copy is implicit; the methods remain for drivers.
#pragma once
#include "cudaCutoffBlock.hh"

class cudaSortedPositions : public virtual cudaCutoffBlock {
public:
  void setupSortedPositions(void) {}
  void calcBlockID(void) {}
};
