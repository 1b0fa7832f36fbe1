// cudaSortedPositions.hh -- the sorted copy r_s / typeID_s of the reference
// (cudaSortedPositions.hh/.cu).  The engine keeps its own cell-ordered
// storage, so the copy is implicit; the methods remain for drivers.
#pragma once
#include "cudaCutoffBlock.hh"

class cudaSortedPositions : public virtual cudaCutoffBlock {
public:
  void setupSortedPositions(void) {}
  // reset (cudaSortedPositions.hh:11-18) re-sizes r_s / typeID_s after a
  // load; the cell-ordered storage is sized with the system
  void reset(void) {}
  void calcBlockID(void) {}
};
