// batch.cpp -- reserved for the batched multi-beam path.
#include "common.hpp"
