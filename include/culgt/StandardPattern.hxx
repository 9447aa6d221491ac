// culgt/StandardPattern.hxx -- lattice/access_pattern/StandardPattern.hxx:
// defined with the other patterns in culgt/GpuPattern.hxx.
#ifndef CULGT_STANDARDPATTERN_HXX_
#define CULGT_STANDARDPATTERN_HXX_
#include "GpuPattern.hxx"
#endif
