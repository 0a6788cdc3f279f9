// Mixed-precision implicit induction solve.
#include <stdexcept>

#include "ctx.h"

int mhdg_ctx::solve_mixed(const double *, const double *, const int8_t *, double, double *, double *)
{
    throw std::runtime_error("mixed-precision solve not available in this build");
}
