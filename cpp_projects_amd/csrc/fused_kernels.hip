// fused_kernels.hip -- reserved for fused FDMT/boxcar kernels.
#include "common.hpp"
