// Benchmark.h -- drop-in for attic/tConvolveMPI/Benchmark.h: the same header
// name, typedefs, run-size constants (:34-46) and `class Benchmark` in the
// global namespace, computed on one MI355X by libtcv (tcv::Benchmark).  Put
// this directory first on the include path and link -ltcv: the reference's
// attic drivers (tConvolveMPI.cc, tConvolveMT.cc) then build unchanged.
#ifndef TCV_ATTIC_BENCHMARK_H
#define TCV_ATTIC_BENCHMARK_H

#include "tcv_benchmark.hpp"

typedef tcv::Coord Coord;
typedef tcv::Real Real;
typedef tcv::Value Value;

// Change these if necessary to adjust run time (the reference's values)
const int nSamples = 160000;   // Number of data samples
const int wSize = 33;          // Number of lookup planes in w projection
const int nChan = 1;           // Number of spectral channels
const int gSize = 4096;        // Size of output grid in pixels
const Coord cellSize = 5.0;    // Cellsize of output grid in wavelengths
const int baseline = 2000;     // Maximum baseline in meters

inline tcv::Config tcvAtticConfig()
{
    tcv::Config c;
    c.nSamples = nSamples;
    c.wSize = wSize;
    c.nChan = nChan;
    c.gSize = gSize;
    c.cellSize = cellSize;
    c.baseline = baseline;
    return c;
}

class Benchmark : public tcv::Benchmark {
public:
    // the reference's default constructor (its run sizes above)
    Benchmark() : tcv::Benchmark(tcvAtticConfig()) {}
};

#endif
