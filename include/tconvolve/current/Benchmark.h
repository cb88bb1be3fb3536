// Benchmark.h -- drop-in for current/tConvolveMPI/Benchmark.h: the same header
// name, typedefs and `class Benchmark` in the global namespace (its problem is
// chosen by setRunType), computed on one MI355X by libtcv
// (tcv::CurrentBenchmark).  Put this directory first on the include path and
// link -ltcv: the reference's current driver builds unchanged.
#ifndef TCV_CURRENT_BENCHMARK_H
#define TCV_CURRENT_BENCHMARK_H

#include "tcv_benchmark.hpp"

typedef tcv::Coord Coord;
typedef tcv::Real Real;
typedef tcv::Value Value;

class Benchmark : public tcv::CurrentBenchmark {
public:
    // the reference's default constructor; the options pick the device and the
    // kernels (tiled by default)
    explicit Benchmark(const tcv::CurrentConfig &opts = tcv::CurrentConfig()) : tcv::CurrentBenchmark(opts) {}
};

#endif
