// Host execution of libse8's kernels for the CPU bounds checks of the tests: see ../simt_host.hpp.
#pragma once
#include "../simt_host.hpp"
