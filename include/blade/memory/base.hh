#ifndef BLADE_MEMORY_BASE_HH
#define BLADE_MEMORY_BASE_HH

#include "blade/memory/types.hh"
#include "blade/memory/shape.hh"
#include "blade/memory/profiler.hh"
#include "blade/memory/vector.hh"
#include "blade/memory/custom.hh"
#include "blade/memory/copy.hh"
#include "blade/memory/helper.hh"

#endif
