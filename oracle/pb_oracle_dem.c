/* DEM restatement (filled in with the DEM path). TEST INFRASTRUCTURE ONLY. */
#include "pb_oracle.h"
