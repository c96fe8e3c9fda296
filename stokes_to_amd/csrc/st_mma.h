// Method of moving asymptotes (one constraint) on the device -- host interface.
#pragma once
#include "st_common.h"

namespace st {

struct MmaDev;
MmaDev *mma_create(long n, hipStream_t s);
void mma_destroy(MmaDev *);
// forget the asymptotes and previous iterates (the next update is a first one)
void mma_reset(MmaDev *);
// one update of the device design x (in place) from the objective gradient df,
// the constraint value g <= 0 and its gradient dg, inside the box [xmin, xmax]
// (all device arrays of n values); returns 0, or < 0 on a HIP failure / NaN
int mma_update(MmaDev *, double *x, const double *df, double g, const double *dg, const double *xmin,
               const double *xmax, int *newton_steps);
// xmin = max(lb, x - move), xmax = min(ub, x + move)
void launch_mma_box(long n, const double *x, double move, double lb, double ub, double *xmin, double *xmax,
                    hipStream_t s);

}  // namespace st
