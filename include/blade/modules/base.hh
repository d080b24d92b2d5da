#ifndef BLADE_MODULES_BASE_HH
#define BLADE_MODULES_BASE_HH

#include "blade/modules/beamformer/ata.hh"
#include "blade/modules/beamformer/batched.hh"
#include "blade/modules/cast.hh"
#include "blade/modules/channelizer/base.hh"
#include "blade/modules/channelizer/detector.hh"
#include "blade/modules/detector.hh"
#include "blade/modules/duplicate.hh"
#include "blade/modules/gather.hh"
#include "blade/modules/guppi/base.hh"
#include "blade/modules/permutation.hh"
#include "blade/modules/phasor/ata.hh"
#include "blade/modules/polarizer.hh"

#endif
