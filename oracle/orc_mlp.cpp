// ORACLE - test infrastructure only (see orc_math.h).
#include "orc_mlp.h"
