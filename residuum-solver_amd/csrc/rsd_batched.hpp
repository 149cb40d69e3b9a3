// rsd_batched.hpp -- batched PCG over independent problems.
#pragma once

#include "rsd_problem.hpp"
