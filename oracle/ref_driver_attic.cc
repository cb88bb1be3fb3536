// Checker driver for the compiled attic reference (attic/tConvolveMPI/Benchmark.cc,
// built with -Dprivate=public so the set-up arrays can be read back).
// Writes raw little-endian arrays into the directory given as argv[1]; the
// fixture script (tests/golden/make_golden.py) turns them into small fixtures.
// Test infrastructure only.
#include "Benchmark.h"
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

static void dump(const std::string& dir, const char* name, const void* p, size_t bytes) {
    std::string path = dir + "/" + name;
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(1); }
    if (bytes && std::fwrite(p, 1, bytes, f) != bytes) { std::perror("fwrite"); std::exit(1); }
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc < 2) { std::fprintf(stderr, "usage: %s outdir [nogrid]\n", argv[0]); return 1; }
    const std::string dir = argv[1];
    const bool doGrid = !(argc > 2 && std::string(argv[2]) == "nogrid");
    {
        Benchmark lcg;
        std::vector<int> draws(64);
        for (int i = 0; i < 64; i++) draws[i] = lcg.randomInt();
        dump(dir, "lcg_draws.i32", draws.data(), draws.size() * 4);
    }
    Benchmark b;
    b.init();
    const size_t n = b.samples.size();
    std::vector<int> iu(n), iv(n), co(n);
    for (size_t i = 0; i < n; i++) { iu[i] = b.samples[i].iu; iv[i] = b.samples[i].iv; co[i] = b.samples[i].cOffset; }
    dump(dir, "u.f64", b.u.data(), b.u.size() * 8);
    dump(dir, "v.f64", b.v.data(), b.v.size() * 8);
    dump(dir, "w.f64", b.w.data(), b.w.size() * 8);
    dump(dir, "iu.i32", iu.data(), n * 4);
    dump(dir, "iv.i32", iv.data(), n * 4);
    dump(dir, "cOffset.i32", co.data(), n * 4);
    dump(dir, "C.c64", b.C.data(), b.C.size() * 8);
    double sc[2] = { double(b.m_support), b.wCellSize };
    dump(dir, "scalars.f64", sc, sizeof(sc));
    if (doGrid) {
        b.runGrid();
        dump(dir, "grid.c64", b.grid.data(), b.grid.size() * 8);
        b.runDegrid();
        dump(dir, "outdata.c64", b.outdata.data(), b.outdata.size() * 8);
    }
    return 0;
}
