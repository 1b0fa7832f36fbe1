// pb_dem.hip -- DEM contact path (cudaParticleDEM); filled in with the DEM row.
#include "pb_internal.h"
