// ORACLE - test infrastructure only (see orc_math.h).
//
// FP64 / FP32 restatement of the radiance-cache network (encoding, 5 x 64
// ReLU hidden layers, 3 exp outputs), its loss and Adam.
#pragma once

namespace orc {}
