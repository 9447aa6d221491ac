// culgt/GpuPatternTimeslice.hxx -- lattice/access_pattern/GpuPatternTimeslice.hxx:
// defined with the other patterns in culgt/GpuPattern.hxx.
#ifndef CULGT_GPUPATTERNTIMESLICE_HXX_
#define CULGT_GPUPATTERNTIMESLICE_HXX_
#include "GpuPattern.hxx"
#endif
