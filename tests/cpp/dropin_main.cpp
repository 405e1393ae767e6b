// Drop-in check: the reference's call site shape (ElectroMag.cpp:236-306) with
// emtrace::CalcField_GPU in place of CalcField_CPU<float>.  Reads charges and
// seeds from a binary file, traces, writes the rows and the perf record.
//   dropin_main <in.bin> <out.bin> <resolution> <use_curvature> [aos]
// With "aos" the lines are an Array<Vector3<float>> and the call is
// emtrace::CalcField (the array-of-structures entry point).
// in.bin : uint64 p, uint64 n, uint64 len, p*4 float charges, n*3 float seeds
// out.bin: int32 status, double time, double performance, double progress,
//          uint32 steps, then 3 * len * n floats (x rows, y rows, z rows)
#include <cstdio>
#include <cstdint>
#include <cstdlib>
#include <vector>

#include "emtrace/calcfield.hpp"

using namespace emtrace::compat;

int main(int argc, char** argv)
{
    if (argc != 5 && argc != 6) return 64;
    const bool aos = argc == 6;
    FILE* f = std::fopen(argv[1], "rb");
    if (!f) return 65;
    uint64_t hdr[3];
    if (std::fread(hdr, sizeof(hdr), 1, f) != 1) return 66;
    const size_t p = hdr[0], n = hdr[1], len = hdr[2];
    Array<pointCharge<float> > charges;
    LinesSoA<float> lines;
    if (charges.AlignAlloc(p) || lines.AlignAlloc(n * len)) return 67;
    if (std::fread(charges.GetDataPointer(), 16, p, f) != p) return 68;
    std::vector<float> seeds(3 * n);
    if (std::fread(seeds.data(), sizeof(float), 3 * n, f) != 3 * n) return 69;
    std::fclose(f);
    for (size_t i = 0; i < n; ++i) lines.write({seeds[3 * i], seeds[3 * i + 1], seeds[3 * i + 2]}, i);
    perfPacket perf;
    int rc;
    if (aos) {
        Array<Vector3<float> > al;
        if (al.AlignAlloc(n * len)) return 71;
        for (size_t i = 0; i < n * len; ++i) al[i] = lines[i];
        rc = emtrace::CalcField(al, charges, n, std::atof(argv[3]), perf, std::atoi(argv[4]) != 0);
        for (size_t i = 0; i < n * len; ++i) lines.write(al[i], i);
    } else {
        rc = emtrace::CalcField_GPU(lines, charges, n, std::atof(argv[3]), perf,
                                    std::atoi(argv[4]) != 0);
    }
    FILE* o = std::fopen(argv[2], "wb");
    if (!o) return 70;
    int32_t st = rc;
    double t = perf.time, g = perf.performance, pr = perf.progress;
    uint32_t steps = (uint32_t)perf.stepTimes.size();
    std::fwrite(&st, 4, 1, o);
    std::fwrite(&t, 8, 1, o);
    std::fwrite(&g, 8, 1, o);
    std::fwrite(&pr, 8, 1, o);
    std::fwrite(&steps, 4, 1, o);
    auto ptr = lines.GetDataPointers();
    std::fwrite(ptr.x, 4, n * len, o);
    std::fwrite(ptr.y, 4, n * len, o);
    std::fwrite(ptr.z, 4, n * len, o);
    std::fclose(o);
    if (rc) std::fprintf(stderr, "CalcField_GPU: %d %s\n", rc, em_last_error());
    return 0;
}
