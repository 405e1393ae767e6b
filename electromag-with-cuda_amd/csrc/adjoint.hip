// adjoint.hip -- placeholder until the adjoint kernels land.
#include "em_internal.hpp"

using namespace em;

extern "C" {

int em_adjoint_dev_f32(const float*, size_t, const float*, const float*, const float*, size_t,
                       size_t, size_t, const em_step*, const em_launch*, const float*, uint32_t,
                       const float*, const float*, const float*, float*, float*, void*)
{
    EM_FAIL(EM_ERR_UNSUPPORTED, "adjoint not built yet");
}

int em_batch_adjoint(em_batch*, const float*, const float*, const float*, const float*,
                     const float*, uint32_t, const float*, const float*, const float*, float*, void*)
{
    EM_FAIL(EM_ERR_UNSUPPORTED, "adjoint not built yet");
}

}
