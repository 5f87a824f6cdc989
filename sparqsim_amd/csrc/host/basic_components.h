// Header that runtime-compiled operator plugins include, under the reference's
// header name (PySparQ/pysparq/dynamic_operator/compiler.py:41-105 generates
// `#include "basic_components.h"` + `using namespace qram_simulator;`).
// It exposes this engine's host layer -- System (register table and host-side
// entries), StateStorage, SparseState, BaseOperator / SelfAdjointOperator and the
// built-in operator classes -- under namespace qram_simulator.  A plugin class
// overrides operator()(std::vector<System>&) (and dag for a BaseOperator); the
// engine hands it a host copy of the device state and uploads the result.
#pragma once

#include <complex>
#include <vector>

#include "sparq.hpp"

namespace qram_simulator {
using namespace sparqsim;
}  // namespace qram_simulator
