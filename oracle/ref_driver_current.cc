// Checker driver for the compiled current-generation reference
// (current/tConvolveMPI/Benchmark.cc, built with -Dprivate=public).
// usage: ref_current outdir runType [grid]
// Writes raw little-endian arrays into outdir.  Test infrastructure only.
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
    if (argc < 3) { std::fprintf(stderr, "usage: %s outdir runType [grid]\n", argv[0]); return 1; }
    const std::string dir = argv[1];
    const int rt = std::atoi(argv[2]);
    const bool doGrid = argc > 3 && std::string(argv[3]) == "grid";
    Benchmark b;
    b.setMPIrank(1);   // silence the rank-0 report
    b.setSort(0);
    b.setRunType(rt);
    b.init();
    const size_t n = b.iu.size();
    dump(dir, "u.f64", b.u.data(), b.u.size() * 8);
    dump(dir, "v.f64", b.v.data(), b.v.size() * 8);
    dump(dir, "w.f64", b.w.data(), b.w.size() * 8);
    dump(dir, "iu.i32", b.iu.data(), n * 4);
    dump(dir, "iv.i32", b.iv.data(), n * 4);
    dump(dir, "wPlane.i32", b.wPlane.data(), n * 4);
    dump(dir, "cOffset.i32", b.cOffset.data(), n * 4);
    dump(dir, "sSize.i32", b.sSize.data(), b.sSize.size() * 4);
    dump(dir, "cOffset0.i32", b.cOffset0.data(), b.cOffset0.size() * 4);
    dump(dir, "C.c64", b.C.data(), b.C.size() * 8);
    double sc[8] = { double(b.gSize), b.uvCellSize, double(b.wSize), double(b.m_support),
                     double(b.overSample), b.wCellSize, double(b.baseline), double(b.nPixelsGridded()) };
    dump(dir, "scalars.f64", sc, sizeof(sc));
    if (doGrid) {
        b.runGrid();
        dump(dir, "grid.c64", b.grid1.data(), b.grid1.size() * 8);
        b.runDegrid();
        dump(dir, "outdata.c64", b.outdata1.data(), b.outdata1.size() * 8);
    }
    return 0;
}
