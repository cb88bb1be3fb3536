// tconvolve_current_gpu: the current tConvolveMPI driver
// (current/tConvolveMPI/tConvolveMPI.cc:63-190) on one MI355X, against the
// drop-in Benchmark class (include/tconvolve/current/Benchmark.h, compiled with
// that directory first on the include path).  One process, so the MPI calls of
// the reference become rank 0 of 1; the report lines are the reference's.
// Extensions for checks: --runtypes a,b,... (default 0,1 as the reference; a
// single run type starts from the generator's seed),
// --sort (setSort(1)), --plain (the one-workgroup-per-visibility kernels) and
// --checksum (FP64 sums of the grid and the degridded data after each run).
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "Benchmark.h"

int main(int argc, char *argv[])
{
    std::vector<int> runs = {0, 1};
    bool checksum = false, sorted = false, plain = false;
    for (int i = 1; i < argc; i++) {
        if (!std::strcmp(argv[i], "--checksum")) checksum = true;
        else if (!std::strcmp(argv[i], "--sort")) sorted = true;
        else if (!std::strcmp(argv[i], "--plain")) plain = true;
        else if (!std::strcmp(argv[i], "--runtypes") && i + 1 < argc) {
            runs.clear();
            std::stringstream ss(argv[++i]);
            std::string tok;
            while (std::getline(ss, tok, ',')) runs.push_back(std::atoi(tok.c_str()));
        } else {
            std::cerr << "usage: " << argv[0] << " [--runtypes a,b,...] [--sort] [--plain] [--checksum]"
                      << std::endl;
            return 1;
        }
    }
    const int numtasks = 1, rank = 0;
    using clk = std::chrono::steady_clock;
    try {
        tcv::CurrentConfig cfg;
        cfg.tiled = !plain;
        // one instance for every run, as the reference (its generator state
        // carries over from one run's init to the next)
        Benchmark bmark(cfg);
        // whether or not to sort visibilities. 0 = no sorting, 1 = sort by w-plane
        bmark.setSort(sorted ? 1 : 0);
        std::vector<float> rates;
        if (rank == 0) rates = bmark.requiredRate();
        for (int run : runs) {
            bmark.setMPIrank(rank);
            bmark.setRunType(run);
            if (rank == 0) std::cout << "+++++ Test " << bmark.getRunType() << " +++++" << std::endl;
            bmark.init();

            const double ngridvis = double(bmark.nVisibilitiesGridded());
            const double ngridpix = double(bmark.nPixelsGridded());
            const double tgridpix = ngridpix * double(numtasks);

            auto t0 = clk::now();
            bmark.runGrid();
            double time = std::chrono::duration<double>(clk::now() - t0).count();
            std::cout << "  Forward processing" << std::endl;
            std::cout << "    Number of processes: " << numtasks << std::endl;
            std::cout << "    Time " << time << " (s) " << std::endl;
            std::cout << "    Time per visibility spectral sample " << 1e6 * time / ngridvis << " (us) " << std::endl;
            std::cout << "    Time per gridding   " << 1e9 * time / ngridpix << " (ns) " << std::endl;
            std::cout << "    Gridding rate (per process)   " << (ngridvis / 1e6) / time << " (Mvis/sec)" << std::endl;
            std::cout << "    Gridding rate (per process)   " << (ngridpix / 1e6) / time << " (Mpix/sec)" << std::endl;
            std::cout << "    Gridding rate (total)      " << (tgridpix / 1e6) / time << " (Mpix/sec)" << std::endl;
            if (run == 0)
                std::cout << "    Continuum gridding performance (per process):   " << (ngridpix / 1e6) / time
                          << " (Mpix/sec)" << std::endl;
            if (run == 1)
                std::cout << "    Spectral gridding performance (per process):    " << (ngridpix / 1e6) / time
                          << " (Mpix/sec)" << std::endl;

            t0 = clk::now();
            bmark.runDegrid();
            time = std::chrono::duration<double>(clk::now() - t0).count();
            std::cout << "  Reverse processing" << std::endl;
            std::cout << "    Number of processes: " << numtasks << std::endl;
            std::cout << "    Time " << time << " (s) " << std::endl;
            std::cout << "    Time per visibility spectral sample " << 1e6 * time / ngridvis << " (us) " << std::endl;
            std::cout << "    Time per degridding " << 1e9 * time / ngridpix << " (ns) " << std::endl;
            std::cout << "    Degridding rate (per node) " << (ngridvis / 1e6) / time << " (Mvis/sec)" << std::endl;
            std::cout << "    Degridding rate (per node) " << (ngridpix / 1e6) / time << " (Mpix/sec)" << std::endl;
            std::cout << "    Degridding rate (total)    " << (tgridpix / 1e6) / time << " (Mpix/sec)" << std::endl;
            if (run == 0)
                std::cout << "    Continuum degridding performance (per process):   " << (ngridpix / 1e6) / time
                          << " (Mpix/sec)" << std::endl;
            if (run == 1)
                std::cout << "    Spectral degridding performance (per process):    " << (ngridpix / 1e6) / time
                          << " (Mpix/sec)" << std::endl;
            if (checksum) {
                const std::vector<Value> g = bmark.grid(), d = bmark.outdata();
                double gr = 0, gi = 0, dr = 0, di = 0;
                for (const auto &z : g) { gr += z.real(); gi += z.imag(); }
                for (const auto &z : d) { dr += z.real(); di += z.imag(); }
                std::printf("checksum run %d pixels %ld grid %.12e %.12e outdata %.12e %.12e "
                            "out_of_range %lld\n", run, (long)ngridpix, gr, gi, dr, di, bmark.nOutOfRange);
            }
            std::cout << "Done" << std::endl;
        }
    } catch (const tcv::Error &e) {
        std::cerr << "tconvolve_current_gpu: " << e.what() << std::endl;
        return 2;
    }
    return 0;
}
