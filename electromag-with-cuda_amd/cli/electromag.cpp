// electromag.cpp -- the ElectroMag command line over the gfx950 tracer.
//
// Same flags, presets, problem and text outputs as the reference application
// (ElectroMag/src/ElectroMag.cpp:72-499); the GPU rows come from
// emtrace::CalcField_GPU (the C ABI of libemtrace.so), the optional CPU rows
// from the CPU tracer in cpu_tracer.hpp.
//
//   --cpu                 also run the CPU tracer (comparison mode)
//   --gpudisable          CPU tracer only
//   --save                write every point to data.txt:
//                           " CPUL [line][step] x: X y: Y z: Z" / " GPUL ..."
//   --autoregress         with --cpu: write to regression.txt the first point
//                         of each line where CPU and GPU are > 0.1 apart
//   --postrundebug        read "line step" pairs from stdin, print both points
//   --bogo --micro --enhanced --extreme --insane --fuckingInsane   presets
//   --randseed            charges from srand(timer) instead of srand(1)
//   --randfieldinit       random line seeds instead of the grid
//   --cpuprecision=double / --gpuprecision=double   FP64 CPU / GPU rows
//   --nodisp --GUI --clmode  accepted; there is no renderer or OpenCL path
// Additions (no reference counterpart):
//   --seed=N              the seed for --randseed (and N+1 for --randfieldinit)
//                         instead of the timer, so random runs can be repeated
//
// Differences from the reference, all deliberate: the reference snapshot
// forces CPU-only mode after parsing (ElectroMag.cpp:203-204) -- here the GPU
// runs as the flags say; --postrundebug ends at end of input and rejects
// out-of-range indices (the reference loops forever and indexes unchecked);
// the regression report skips the "previous point" of step 0.  Exit status:
// 0, or the failing tracer's status code.
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <iostream>
#include <limits>
#include <string>
#include <thread>
#include <vector>

#include "cpu_tracer.hpp"
#include "emtrace/calcfield.hpp"
#include "emtrace/legacy.hpp"

namespace {

using emtrace::compat::Array;
using emtrace::compat::LinesSoA;
using emtrace::compat::perfPacket;
using emtrace::compat::pointCharge;

struct SimulationParams {
    size_t nx, ny, nz;  // lines along x, y, z
    size_t pStatic;     // static point charges
    size_t pDynamic;    // dynamic charge elements (always 0)
    size_t len;         // rows per line
};

// ElectroMag.cpp:33-58, in the order of its ParamLevel enum.
enum ParamLevel { kBogo, kMicro, kCpuMode, kNormal, kEnhanced, kExtreme, kInsane, kFuckingInsane };
const SimulationParams kParams[] = {
    {16, 16, 1, 50, 0, 500},          // --bogo
    {16, 16, 1, 1000, 0, 1000},       // --micro
    {64, 64, 1, 1000, 0, 1000},       // CPU mode
    {128, 128, 1, 1024, 0, 2500},     // default
    {256, 112, 1, 2048, 0, 5000},     // --enhanced
    {256, 256, 1, 2048, 0, 5000},     // --extreme
    {512, 512, 1, 2048, 0, 5000},     // --insane
    {1024, 1024, 1, 5120, 0, 10000},  // --fuckingInsane
};

struct Options {
    ParamLevel level = kNormal;
    bool save = false, cpu = false, gpu = true, display = true;
    bool randseed = false, randfieldinit = false, debug = false, regress = false;
    bool cpuDP = false, gpuDP = false, clmode = false;
    bool haveSeed = false;
    uint32_t seed = 0;
};

long long timer_us()
{
    return std::chrono::duration_cast<std::chrono::microseconds>(
               std::chrono::system_clock::now().time_since_epoch())
        .count();
}

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void parse(int argc, char** argv, Options& o)
{
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        if (!std::strcmp(a, "--cpu")) {
            o.cpu = true;
        } else if (!std::strcmp(a, "--gpudisable")) {
            o.gpu = false;
            o.cpu = true;
        } else if (!std::strcmp(a, "--save")) {
            o.save = true;
        } else if (!std::strcmp(a, "--nodisp")) {
            o.display = false;
        } else if (!std::strcmp(a, "--bogo")) {
            if (o.level == kNormal) o.level = kBogo;
        } else if (!std::strcmp(a, "--micro")) {
            if (o.level == kNormal) o.level = kMicro;
        } else if (!std::strcmp(a, "--enhanced")) {
            if (o.level < kEnhanced) o.level = kEnhanced;
        } else if (!std::strcmp(a, "--extreme")) {
            if (o.level < kExtreme) o.level = kExtreme;
        } else if (!std::strcmp(a, "--insane")) {
            if (o.level < kInsane) o.level = kInsane;
        } else if (!std::strcmp(a, "--fuckingInsane")) {
            if (o.level < kFuckingInsane) o.level = kFuckingInsane;
        } else if (!std::strcmp(a, "--GUI")) {
            // progress window of the reference's GUI: no GUI here
        } else if (!std::strcmp(a, "--randseed")) {
            o.randseed = true;
        } else if (!std::strcmp(a, "--randfieldinit")) {
            o.randfieldinit = true;
        } else if (!std::strcmp(a, "--postrundebug")) {
            o.debug = true;
        } else if (!std::strcmp(a, "--autoregress")) {
            o.regress = true;
        } else if (!std::strcmp(a, "--cpuprecision=double")) {
            o.cpuDP = true;
        } else if (!std::strcmp(a, "--gpuprecision=double")) {
            o.gpuDP = true;
        } else if (!std::strcmp(a, "--clmode")) {
            o.clmode = true;
        } else if (!std::strncmp(a, "--seed=", 7)) {
            char* end = nullptr;
            const unsigned long v = std::strtoul(a + 7, &end, 10);
            if (end && *end == '\0' && a[7] != '\0') {
                o.haveSeed = true;
                o.seed = static_cast<uint32_t>(v);
            } else {
                std::cout << " Ignoring unknown argument: " << a << std::endl;
            }
        } else {
            std::cout << " Ignoring unknown argument: " << a << std::endl;
        }
    }
}

// The console progress bar of the reference (Electromag utils.h:133-151):
// 60 marks, one printed each time progress passes another 1/60.
class ConsoleMonitor {
public:
    explicit ConsoleMonitor(volatile double* progress) : progress_(progress), stop_(false)
    {
        th_ = std::thread([this] { run(); });
    }
    ~ConsoleMonitor()
    {
        stop_ = true;
        th_.join();
    }

private:
    void run()
    {
        const double step = 1.0 / 60;
        std::cout << "[__________________________________________________________]" << std::endl;
        for (double next = step; next < 1.0 - 1e-3; next += step) {
            while (*progress_ < next) {
                // stopped before the end: only when the tracer failed
                if (stop_ && *progress_ < next) return;
                std::this_thread::sleep_for(std::chrono::milliseconds(50));
            }
            std::cout << "." << std::flush;
        }
        std::cout << " Done" << std::endl;
    }
    volatile double* progress_;
    std::atomic<bool> stop_;
    std::thread th_;
};

template <class T>
struct Point {
    T x, y, z;
};

template <class T>
Point<T> at(LinesSoA<T>& l, size_t i)
{
    return {l.x[i], l.y[i], l.z[i]};
}

template <class A, class B>
float offset3d(const Point<A>& c, const Point<B>& g)
{
    // vec3Len(vec3(cpu, gpu)) (include/Vector.h:166-175) in the wider type
    typedef decltype(A() + B()) W;
    const W dx = (W)g.x - (W)c.x, dy = (W)g.y - (W)c.y, dz = (W)g.z - (W)c.z;
    return static_cast<float>(std::sqrt(dx * dx + dy * dy + dz * dz));
}

template <class T>
void print_point(std::ostream& os, const char* tag, size_t line, long long step, const Point<T>& p)
{
    os << " " << tag << " [" << line << "][" << step << "] x: " << p.x << " y: " << p.y
       << " z: " << p.z << std::endl;
}

template <class T>
void copy_charges(Array<pointCharge<T>>& dst, const std::vector<float>& q, size_t p)
{
    for (size_t i = 0; i < p; ++i) {
        dst[i].position.x = static_cast<T>(q[4 * i]);
        dst[i].position.y = static_cast<T>(q[4 * i + 1]);
        dst[i].position.z = static_cast<T>(q[4 * i + 2]);
        dst[i].magnitude = static_cast<T>(q[4 * i + 3]);
    }
}

template <class T>
void copy_seeds(LinesSoA<T>& dst, const std::vector<float>& s, size_t n)
{
    for (size_t i = 0; i < n; ++i) {
        dst.x[i] = static_cast<T>(s[i]);
        dst.y[i] = static_cast<T>(s[n + i]);
        dst.z[i] = static_cast<T>(s[2 * n + i]);
    }
}

template <class TC, class TG>
int run(const Options& o)
{
    bool cpu = o.cpu, gpu = o.gpu;
    const SimulationParams& sim = kParams[o.level];
    const size_t nw = sim.nx, nh = sim.ny, nd = sim.nz, n = nw * nh * nd, p = sim.pStatic,
                 len = sim.len;

    LinesSoA<TC> CPUlines;
    LinesSoA<TG> GPUlines;
    if (gpu && GPUlines.AlignAlloc(n * len)) gpu = false;
    if (cpu && CPUlines.AlignAlloc(n * len)) cpu = false;
    if (!cpu && !gpu) {
        std::cerr << " Could not allocate sufficient memory. Halting execution." << std::endl;
        std::cerr << " " << n * len * 3 * sizeof(TG) / 1024 / 1024
                  << " MB needed for initial allocation" << std::endl;
        return 666;  // the reference's status for this case (ElectroMag.cpp:255-261)
    }
    std::ofstream data;
    if (o.save) data.open("data.txt");

    // The reference's problem (Electromag utils.h:29-111), in its float
    // precision, then widened where a tracer runs in double.
    std::vector<float> q(4 * p), seeds(3 * n);
    uint32_t chargeSeed = 1;
    if (o.randseed) {
        chargeSeed = o.haveSeed ? o.seed
                                : static_cast<uint32_t>(timer_us() % emtrace::legacy::kRandMax);
        std::cout << " Charge seed:\t" << chargeSeed << std::endl;
    }
    emtrace::legacy::LibcRand g(chargeSeed);
    emtrace::legacy::init_charges(q.data(), p, g);
    if (o.randfieldinit) {
        const uint32_t lineSeed =
            o.haveSeed ? o.seed + 1u
                       : static_cast<uint32_t>(timer_us() % emtrace::legacy::kRandMax);
        std::cout << " Line seed:\t" << lineSeed << std::endl;
        g.reseed(lineSeed);
        emtrace::legacy::init_random_lines(seeds.data(), seeds.data() + n, seeds.data() + 2 * n,
                                           n, g);
    } else {
        emtrace::legacy::init_grid(seeds.data(), seeds.data() + n, seeds.data() + 2 * n, nw, nh,
                                   nd);
    }

    const double resolution = 1;  // ElectroMag.cpp:296
    const bool useCurvature = true;
    perfPacket CPUperf, GPUperf;
    double GPUtime = 0, CPUtime = 0;

    if (o.clmode)
        std::cout << " --clmode: the OpenCL test path is not part of this build;"
                  << " running the normal path" << std::endl;
    if (gpu) {
        Array<pointCharge<TG>> charges;
        charges.AlignAlloc(p);
        copy_charges(charges, q, p);
        copy_seeds(GPUlines, seeds, n);
        std::cout << " GPU" << std::endl;
        const auto t0 = std::chrono::steady_clock::now();
        const int rc = emtrace::CalcField_GPU(GPUlines, charges, n, resolution, GPUperf,
                                              useCurvature);
        GPUtime = seconds_since(t0);
        if (rc != EM_OK) {
            std::cerr << " GPU tracer failed with status " << rc << ": " << em_last_error()
                      << std::endl;
            return rc;
        }
        std::cout << " GPU kernel execution time:\t" << GPUperf.time << " seconds" << std::endl;
        std::cout << " Effective performance:\t\t" << GPUperf.performance << " GFLOP/s"
                  << std::endl;
        std::cout << " True kernel execution time:\t" << GPUtime << " seconds" << std::endl;
        for (const auto& t : GPUperf.stepTimes) {
            std::cout << t.message << ": " << t.time << std::endl;
            if (t.bandwidth != 0.0) std::cout << "  Bandwidth: " << t.bandwidth << " MB/s" << std::endl;
        }
    }
    if (cpu) {
        std::vector<TC> qc(4 * p);
        for (size_t i = 0; i < 4 * p; ++i) qc[i] = static_cast<TC>(q[i]);
        copy_seeds(CPUlines, seeds, n);
        auto ptr = CPUlines.GetDataPointers();
        int rc;
        const auto t0 = std::chrono::steady_clock::now();
        {
            ConsoleMonitor mon(&CPUperf.progress);
            rc = emcli::cpu_calc_lines<TC>(qc.data(), p, ptr.x, ptr.y, ptr.z, n, len,
                                           static_cast<TC>(resolution), useCurvature,
                                           &CPUperf.progress);
            CPUtime = seconds_since(t0);
            CPUperf.progress = 1.0;
        }
        if (rc != 0) {
            std::cerr << " CPU tracer failed with status " << rc << std::endl;
            return rc;
        }
        // perfData.time / performance as CPU_Implement.cpp:345-349 defines them
        CPUperf.time = CPUtime;
        CPUperf.performance = (n * ((len - 1) * (p * (15.0 + 3) + 13)) / CPUperf.time) / 1E9;
        std::cout << " CPU kernel execution time:\t" << CPUperf.time << " seconds" << std::endl;
        std::cout << " Effective performance:\t\t" << CPUperf.performance << " GFLOP/s"
                  << std::endl;
        std::cout << " True kernel execution time:\t" << CPUtime << " seconds" << std::endl;
        if (gpu) {
            std::cout << " Effective speedup:\t\t" << GPUperf.performance / CPUperf.performance
                      << " x" << std::endl;
            std::cout << " Realistic speedup:\t\t" << CPUtime / GPUtime << " x" << std::endl;
        }
    }

    if (o.display)
        std::cerr << " Could not load graphics module. Rendering disabled" << std::endl;
    std::cout << " Skipping display" << std::endl;

    if (o.save && (cpu || gpu)) {
        std::cout << " Beginning save procedure" << std::endl;
        for (size_t line = 0; line < n; ++line) {
            for (size_t step = 0; step < len; ++step) {
                const size_t i = step * n + line;
                if (cpu) print_point(data, "CPUL", line, (long long)step, at(CPUlines, i));
                if (gpu) print_point(data, "GPUL", line, (long long)step, at(GPUlines, i));
            }
            const float percent = (float)line / n * 100;
            std::cout << percent << " %complete" << std::endl;
        }
        std::cout << " Save procedure complete" << std::endl;
    }

    if (o.regress && cpu && gpu) {
        std::ofstream regress("regression.txt");
        std::cout << " Beginning verfication procedure" << std::endl;
        for (size_t line = 0; line < n; ++line) {
            for (size_t step = 0; step < len; ++step) {
                const size_t i = step * n + line;
                const float off = offset3d(at(CPUlines, i), at(GPUlines, i));
                if (off > 0.1f) {
                    if (step > 0) {
                        const size_t last = i - n;
                        print_point(regress, "CPUL", line, (long long)step - 1, at(CPUlines, last));
                        print_point(regress, "GPUL", line, (long long)step - 1, at(GPUlines, last));
                        regress << " 3D offset: " << offset3d(at(CPUlines, last), at(GPUlines, last))
                                << std::endl;
                    }
                    print_point(regress, "CPUL", line, (long long)step, at(CPUlines, i));
                    print_point(regress, "GPUL", line, (long long)step, at(GPUlines, i));
                    regress << " 3D offset: " << off << std::endl << std::endl;
                    break;  // one report per line
                }
            }
            if (!(line % 256)) std::cout << " " << (double)line / n * 100 << " % complete" << std::endl;
        }
        std::cout << " Verification complete" << std::endl;
    }

    while (o.debug) {
        size_t line, step;
        if (!(std::cin >> line >> step)) {
            if (std::cin.eof()) break;
            std::cin.clear();
            std::cin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
            continue;
        }
        std::cin.ignore(std::numeric_limits<std::streamsize>::max(), '\n');
        if (line >= n || step >= len) {
            std::cout << " Out of range: lines 0.." << n - 1 << ", steps 0.." << len - 1 << std::endl;
            continue;
        }
        const size_t i = step * n + line;
        if (cpu) print_point(std::cout, "CPUL", line, (long long)step, at(CPUlines, i));
        if (gpu) print_point(std::cout, "GPUL", line, (long long)step, at(GPUlines, i));
        if (cpu && gpu)
            std::cout << " 3D offset: " << offset3d(at(CPUlines, i), at(GPUlines, i)) << std::endl;
    }
    return 0;
}

std::string cpu_model()
{
    std::ifstream f("/proc/cpuinfo");
    std::string s;
    while (std::getline(f, s))
        if (s.compare(0, 10, "model name") == 0) {
            const size_t c = s.find(':');
            return c == std::string::npos ? s : s.substr(c + 2);
        }
    return "unknown";
}

}  // namespace

int main(int argc, char** argv)
{
    std::cout << " Electromagnetism simulation application" << std::endl;
    std::cout << " Compiled on " << __DATE__ << " at " << __TIME__ << std::endl;
    Options o;
    parse(argc, argv, o);
    std::clog << " Processor:\t" << cpu_model() << std::endl;
    if (o.gpu) std::clog << " Tracer:\t" << em_version() << ", " << em_device_count() << " device(s)"
                         << std::endl;
    if (o.cpuDP && o.gpuDP) return run<double, double>(o);
    if (o.cpuDP) return run<double, float>(o);
    if (o.gpuDP) return run<float, double>(o);
    return run<float, float>(o);
}
