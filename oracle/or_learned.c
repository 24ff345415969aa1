/* Oracle (test infrastructure only): learned offset/scale min-sum, filled in below. */
#include "oracle.h"
