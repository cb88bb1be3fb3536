// CPU baseline driver for the compiled reference (test and benchmark
// infrastructure only): times attic/tConvolveMPI/Benchmark.cc's own
// gridKernel and degridKernel (:125-195) on a workload read from files, one
// Benchmark per thread as tConvolveMT runs one per thread
// (attic/tConvolveMT/tConvolveMT.cc:80-149), the C table shared read-only.
// Thread t takes visibilities [t * perThread, (t + 1) * perThread) of the
// workload (wrapping), grids them into its own grid, then degrids them from
// it.  Prints one JSON line with the wall time of each phase.
//   ref_time_attic DIR THREADS PER_THREAD
// DIR holds iu.i32 iv.i32 cOffset.i32 data.c64 C.c64 and params.txt
// ("gSize support").
#include "Benchmark.h"
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <string>
#include <thread>
#include <vector>

template <typename T>
static std::vector<T> slurp(const std::string &path)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) { std::perror(path.c_str()); std::exit(1); }
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    std::vector<T> v(bytes / sizeof(T));
    if (bytes && std::fread(v.data(), 1, bytes, f) != (size_t)bytes) { std::perror("fread"); std::exit(1); }
    std::fclose(f);
    return v;
}

int main(int argc, char **argv)
{
    if (argc < 4) { std::fprintf(stderr, "usage: %s DIR THREADS PER_THREAD\n", argv[0]); return 1; }
    const std::string dir = argv[1];
    const int threads = std::atoi(argv[2]);
    const long perThread = std::atol(argv[3]);
    int g = 0, support = 0;
    {
        FILE *f = std::fopen((dir + "/params.txt").c_str(), "r");
        if (!f || std::fscanf(f, "%d %d", &g, &support) != 2) { std::fprintf(stderr, "bad params\n"); return 1; }
        std::fclose(f);
    }
    const auto iu = slurp<int>(dir + "/iu.i32");
    const auto iv = slurp<int>(dir + "/iv.i32");
    const auto co = slurp<int>(dir + "/cOffset.i32");
    const auto data = slurp<Value>(dir + "/data.c64");
    const auto C = slurp<Value>(dir + "/C.c64");
    const long n = (long)iu.size();
    if (threads < 1 || perThread < 1 || n < 1 || (long)iv.size() != n || (long)co.size() != n ||
        (long)data.size() != n) {
        std::fprintf(stderr, "bad inputs\n");
        return 1;
    }
    std::vector<Benchmark> b(threads);
    std::vector<std::vector<Value>> grids(threads), outs(threads);
    for (int t = 0; t < threads; ++t) {
        b[t].samples.resize(perThread);
        for (long i = 0; i < perThread; ++i) {
            const long k = (t * perThread + i) % n;
            b[t].samples[i].data = data[k];
            b[t].samples[i].iu = iu[k];
            b[t].samples[i].iv = iv[k];
            b[t].samples[i].cOffset = co[k];
        }
        grids[t].assign((size_t)g * g, Value(0.0f, 0.0f));
        outs[t].assign(perThread, Value(0.0f, 0.0f));
    }
    auto phase = [&](bool grid) {
        const auto t0 = std::chrono::steady_clock::now();
        std::vector<std::thread> pool;
        for (int t = 0; t < threads; ++t)
            pool.emplace_back([&, t] {
                if (grid) b[t].gridKernel(support, C, grids[t], g);
                else b[t].degridKernel(grids[t], g, support, C, outs[t]);
            });
        for (auto &th : pool) th.join();
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    };
    const double tg = phase(true), td = phase(false);
    double check = 0;
    for (int t = 0; t < threads; ++t) check += std::abs(outs[t][0]);
    std::printf("{\"grid_s\": %.6f, \"degrid_s\": %.6f, \"threads\": %d, \"per_thread\": %ld, "
                "\"sSize\": %d, \"check\": %.6e}\n", tg, td, threads, perThread, 2 * support + 1, check);
    return 0;
}
