// Iteration-graph builders: the lowered IR of the hot-path statements under
// the reference's GPU schedules.
#pragma once

#include <string>
#include <vector>

#include "ir.h"
#include "taco_hip.h"

namespace taco_hip {

// Builds the module's functions ("assemble", "compute", "evaluate") for a
// statement and schedule family; throws std::runtime_error on an
// unsupported combination.
std::vector<ir::Stmt> lower_spec(const taco_hip_gen_spec& spec);

}  // namespace taco_hip
