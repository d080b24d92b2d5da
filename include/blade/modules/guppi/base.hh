// GUPPI RAW reader and writer (reference include/blade/modules/guppi/base.hh).
#ifndef BLADE_MODULES_GUPPI_BASE_HH
#define BLADE_MODULES_GUPPI_BASE_HH

#include "blade/modules/guppi/reader.hh"
#include "blade/modules/guppi/writer.hh"

#endif
