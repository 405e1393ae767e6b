// oracle/ref_init_driver.cpp -- TEST INFRASTRUCTURE ONLY.
//
// Flat extern "C" access to the reference's own problem initialisation,
// InitializePointChargeArray / InitializeFieldLineArray
// (ElectroMag/src/Electromag utils.h:29-111: libc rand() under srand(1), or a
// grid of seeds spaced 10 apart), compiled from the reference tree into
// oracle/_ref/libemref.so.  Used only to write golden vectors that pin the
// product's restatement of that generator.
#include <iostream>

#include "X-Compat/HPC Timing.h"
#include "SOA_utils.hpp"
#include "Electrostatics.h"
#include "Electromag utils.h"

extern "C" void ref_init_charges_f32(float* out, size_t p, int random)
{
    Array<electro::pointCharge<float> > q;
    q.AlignAlloc(p, 256);
    InitializePointChargeArray(q, p, random != 0);
    for (size_t i = 0; i < p; ++i) {
        out[4 * i + 0] = q[i].position.x;
        out[4 * i + 1] = q[i].position.y;
        out[4 * i + 2] = q[i].position.z;
        out[4 * i + 3] = q[i].magnitude;
    }
}

extern "C" void ref_init_lines_f32(float* x, float* y, float* z, size_t w, size_t h, size_t d)
{
    const size_t n = w * h * d;
    Vector3<Array<float> > lines;
    lines.AlignAlloc(n, 256);
    InitializeFieldLineArray(lines, n, w, h, d, false);
    for (size_t i = 0; i < n; ++i) {
        x[i] = lines.x[i];
        y[i] = lines.y[i];
        z[i] = lines.z[i];
    }
}
