// tconvolve_gpu: the tConvolveMPI driver (attic/tConvolveMPI/tConvolveMPI.cc:62-127)
// run against the drop-in tcv::Benchmark on one MI355X.  Same report lines;
// times come from a steady clock around calls that wait for the device.
//
// usage: tconvolve_gpu [--plain] [--samples N] [--checksum]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "tcv_benchmark.hpp"

int main(int argc, char *argv[])
{
    tcv::Config cfg;
    bool checksum = false;
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--plain")) cfg.tiled = false;
        else if (!std::strcmp(argv[i], "--checksum")) checksum = true;
        else if (!std::strcmp(argv[i], "--samples") && i + 1 < argc) cfg.nSamples = std::atoi(argv[++i]);
        else {
            std::cerr << "usage: " << argv[0] << " [--plain] [--samples N] [--checksum]" << std::endl;
            return 1;
        }
    }
    try {
        tcv::Benchmark bmark(cfg);
        bmark.init();

        const int sSize = 2 * bmark.getSupport() + 1;
        const double griddings = double(cfg.nSamples) * cfg.nChan * double(sSize * sSize);
        using clk = std::chrono::steady_clock;

        std::cout << "+++++ Forward processing (" << (cfg.tiled ? "tiled" : "plain") << ") +++++" << std::endl;
        auto t0 = clk::now();
        bmark.runGrid();
        double time = std::chrono::duration<double>(clk::now() - t0).count();
        std::cout << "    Time " << time << " (s) " << std::endl;
        std::cout << "    Time per visibility spectral sample " << 1e6 * time / double(cfg.nSamples * cfg.nChan) << " (us) " << std::endl;
        std::cout << "    Time per gridding   " << 1e9 * time / griddings << " (ns) " << std::endl;
        std::cout << "    Gridding rate   " << (griddings / 1000000) / time << " (million grid points per second)" << std::endl;

        std::cout << "+++++ Reverse processing (" << (cfg.tiled ? "tiled" : "plain") << ") +++++" << std::endl;
        t0 = clk::now();
        bmark.runDegrid();
        time = std::chrono::duration<double>(clk::now() - t0).count();
        std::cout << "    Time " << time << " (s) " << std::endl;
        std::cout << "    Time per visibility spectral sample " << 1e6 * time / double(cfg.nSamples * cfg.nChan) << " (us) " << std::endl;
        std::cout << "    Time per degridding " << 1e9 * time / griddings << " (ns) " << std::endl;
        std::cout << "    Degridding rate " << (griddings / 1000000) / time << " (million grid points per second)" << std::endl;

        if (checksum) {
            const std::vector<tcv::Value> g = bmark.grid(), d = bmark.outdata();
            double gr = 0, gi = 0, dr = 0, di = 0;
            for (const auto &z : g) { gr += z.real(); gi += z.imag(); }
            for (const auto &z : d) { dr += z.real(); di += z.imag(); }
            std::printf("checksum grid %.9e %.9e outdata %.9e %.9e out_of_range %lld\n", gr, gi, dr, di,
                        bmark.nOutOfRange);
        }
        std::cout << "Done" << std::endl;
    } catch (const tcv::Error &e) {
        std::cerr << "tconvolve_gpu: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
