// Umbrella header: the model/evolver API of CosmoLattice's CosmoInterface, backed by libclmi
// (include/clmi.h).  Model headers include "CosmoInterface/cosmointerface.h" as they do in the
// reference; add cosmolattice_amd/host and include/ to the include path and link libclmi.so.
#pragma once

#include "CosmoInterface/abstractmodel.h"
#include "CosmoInterface/evolvers.h"
#include "CosmoInterface/initializers.h"
#include "CosmoInterface/measurements.h"
#include "CosmoInterface/parameterparser.h"
#include "CosmoInterface/runparameters.h"
#include "CosmoInterface/simulation.h"
#include "CosmoInterface/symbolic.h"
