/* Oracle (test infrastructure only): integrator restatement -- filled in below. */
#include "oracle.h"
