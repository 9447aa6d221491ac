// culgt/PhiloxWrapper.hxx -- host side of the reference's
// lattice/rng/PhiloxWrapper.hxx:62-150: the process-wide launch counter that
// every RNG kernel launch takes (getNextCounter).  The device generator
// (Philox4x32-10 keyed by thread id and seed, PhiloxWrapper.hxx:93-140) lives
// inside the library's kernels.
#ifndef CULGT_PHILOXWRAPPER_HXX_
#define CULGT_PHILOXWRAPPER_HXX_

class PhiloxWrapper
{
public:
    static unsigned int getNextCounter() { return globalCounter++; }
    static unsigned int getCurrentCounter() { return globalCounter; }

private:
    static inline unsigned int globalCounter = 0;
};

#endif
