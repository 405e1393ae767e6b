// emtrace/legacy.hpp -- the reference application's own problem generator.
//
// ElectroMag builds its default problem with libc rand() under srand(1)
// (ElectroMag/src/Electromag utils.h:91-111) and a grid of line seeds spaced
// 10 apart (Electromag utils.h:29-70), e.g. 1,024 charges and 128 x 128 lines
// for its default preset (ElectroMag.cpp:43).  This header restates both so
// that the reference's problems can be rebuilt bit for bit anywhere, without
// depending on the host C library:
//
//   LibcRand        -- glibc's rand()/random() (the TYPE_3 additive feedback
//                      generator: 31 words of state seeded by the Park-Miller
//                      minimal standard LCG, x_i = x_{i-3} + x_{i-31} mod 2^32,
//                      output x_i >> 1, first 310 outputs discarded), as
//                      documented for glibc's random_r.c.  RAND_MAX = 2^31 - 1.
//   init_charges    -- InitializePointChargeArray: per charge four draws,
//                      position (T)(r - (T)RAND_MAX/2) / RAND_MAX * 10000 and
//                      magnitude (T)(r - (T)RAND_MAX/10) / RAND_MAX, evaluated
//                      in T as the reference source orders them.
//   init_grid       -- InitializeFieldLineArray(random = false): coordinate
//                      10 * v with v = (T)(-(T)w/2 + 1e-5) + i per axis,
//                      index k*w*h + j*w + i.
//   init_random_lines -- InitializeFieldLineArray(random = true): three draws
//                      per line, (T)(r - RAND_MAX/2) / RAND_MAX * 10000 with
//                      the subtraction in int.
//
// The reference seeds its "random" modes from a timer (pseudoSeed % RAND_MAX);
// callers here pass the seed explicitly.
#pragma once

#include <cstddef>
#include <cstdint>

namespace emtrace {
namespace legacy {

static const int32_t kRandMax = 2147483647;

class LibcRand {
public:
    explicit LibcRand(uint32_t seed = 1) { reseed(seed); }

    // srand(seed): glibc stores the seed as a 32-bit signed word; 0 becomes 1.
    void reseed(uint32_t seed)
    {
        int32_t word = static_cast<int32_t>(seed == 0 ? 1u : seed);
        s_[0] = static_cast<uint32_t>(word);
        for (int i = 1; i < kDeg; ++i) {
            // 16807 * word mod (2^31 - 1) by Schrage's decomposition, exactly as
            // 32-bit signed arithmetic evaluates it (also for negative words).
            const int32_t hi = word / 127773;
            const int32_t lo = word % 127773;
            word = 16807 * lo - 2836 * hi;
            if (word < 0) word += 2147483647;
            s_[i] = static_cast<uint32_t>(word);
        }
        f_ = kSep;
        r_ = 0;
        for (int i = 0; i < kDeg * 10; ++i) next();
    }

    // rand(): 31-bit result in [0, RAND_MAX].
    int32_t next()
    {
        s_[f_] += s_[r_];
        const int32_t out = static_cast<int32_t>(s_[f_] >> 1);
        f_ = f_ + 1 == kDeg ? 0 : f_ + 1;
        r_ = r_ + 1 == kDeg ? 0 : r_ + 1;
        return out;
    }

private:
    static const int kDeg = 31, kSep = 3;
    uint32_t s_[kDeg];
    int f_, r_;
};

// Charges as p records {x, y, z, q} (the layout of electro::pointCharge<T>).
template <class T>
void init_charges(T* q, size_t p, LibcRand& g)
{
    const T half = static_cast<T>(kRandMax) / static_cast<T>(2);
    const T tenth = static_cast<T>(kRandMax) / static_cast<T>(10);
    const T scale = static_cast<T>(kRandMax);
    for (size_t i = 0; i < p; ++i) {
        for (int c = 0; c < 3; ++c) {
            const T d = static_cast<T>(g.next()) - half;
            q[4 * i + c] = d / scale * static_cast<T>(10000);
        }
        const T m = static_cast<T>(g.next()) - tenth;
        q[4 * i + 3] = m / scale;
    }
}

// Grid seeds, n = width * height * depth.
template <class T>
void init_grid(T* x, T* y, T* z, size_t width, size_t height, size_t depth)
{
    T zv = static_cast<T>(-static_cast<T>(depth) / static_cast<T>(2) + 1E-5);
    for (size_t k = 0; k < depth; ++k, zv += static_cast<T>(1)) {
        T yv = static_cast<T>(-static_cast<T>(height) / static_cast<T>(2) + 1E-5);
        for (size_t j = 0; j < height; ++j, yv += static_cast<T>(1)) {
            T xv = static_cast<T>(-static_cast<T>(width) / static_cast<T>(2) + 1E-5);
            for (size_t i = 0; i < width; ++i, xv += static_cast<T>(1)) {
                const size_t at = (k * height + j) * width + i;
                x[at] = static_cast<T>(10) * xv;
                y[at] = static_cast<T>(10) * yv;
                z[at] = static_cast<T>(10) * zv;
            }
        }
    }
}

// Random seeds: per line x, y, z draws in that order.
template <class T>
void init_random_lines(T* x, T* y, T* z, size_t n, LibcRand& g)
{
    const T scale = static_cast<T>(kRandMax);
    T* axis[3] = {x, y, z};
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 3; ++c)
            axis[c][i] = static_cast<T>(g.next() - kRandMax / 2) / scale * static_cast<T>(10000);
}

}  // namespace legacy
}  // namespace emtrace
