/* SPH restatement (filled in with the SPH path). TEST INFRASTRUCTURE ONLY. */
#include "pb_oracle.h"
