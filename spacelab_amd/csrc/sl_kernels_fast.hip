// Fast path: placeholder until the Newton-3 tile kernels land.
#include "spacelab/sl_api.h"
#include "sl_internal.hpp"

namespace sl {

int fast_alloc(FastBuffers& fb, int n_max, int n_pad)
{
    fb.n_max = n_max;
    fb.n_pad = n_pad;
    return 0;
}
void fast_free(FastBuffers&) {}
int fast_max_blocks(int) { return 1; }
int fast_sync_from(FastBuffers&, State&, int, hipStream_t, std::string*) { return 0; }
int fast_sync_to(FastBuffers&, State&, int, hipStream_t, std::string*) { return 0; }
static int unavailable(std::string* err)
{
    if (err) *err = "SL_ORDER_FAST is not available in this build";
    return SL_E_ARG;
}
int fast_begin_init(FastBuffers&, State&, int, hipStream_t, std::string* err) { return unavailable(err); }
int fast_drift(FastBuffers&, State&, int, double, hipStream_t, std::string* err) { return unavailable(err); }
int fast_pairs(FastBuffers&, State&, const Phys&, int, int, int, unsigned long long*, hipStream_t,
               std::string* err)
{
    return unavailable(err);
}
int fast_ball(FastBuffers&, State&, const Phys&, int, int, int, int, double*, double*, unsigned long long*,
              int*, hipStream_t, int*, std::string* err)
{
    return unavailable(err);
}
hipError_t fast_reset_counters(FastBuffers&, hipStream_t) { return hipSuccess; }
long long fast_read_near_total(FastBuffers&, hipStream_t) { return 0; }

}  // namespace sl
