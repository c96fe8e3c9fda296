// C entry points around the reference's MMASolver class (EXT/mma/MMASolver.h:56-106),
// so that ctypes can drive the reference implementation compiled from its own
// source file.  Test infrastructure only: built into oracle/_ref/ by `make ref`.
#include "MMASolver.h"

extern "C" {
void *mmaref_create(int n, int m) { return new MMASolver(n, m); }
void mmaref_free(void *h) { delete static_cast<MMASolver *>(h); }
void mmaref_update(void *h, double *x, const double *dfdx, const double *gx, const double *dgdx, const double *xmin,
                   const double *xmax)
{
    static_cast<MMASolver *>(h)->Update(x, dfdx, gx, dgdx, xmin, xmax);
}
}
