// pb_sph.hip -- SPH Navier-Stokes path (cudaParticleSPH_NS); filled in with the SPH row.
#include "pb_internal.h"
