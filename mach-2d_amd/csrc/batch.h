// batch.h -- device descriptor of one system of a batched solve.
//
// A batch keeps every system's arrays in shared pools; the kernels find a
// system's arrays through its descriptor and their work item through a map
// of (system, band) or (system, block) pairs, so one launch per phase covers
// the whole batch whatever the grid sizes.
#pragma once

#include "m2d_internal.h"
#include "skew.h"

namespace m2d {

typedef unsigned long long u64;

template <typename T>
struct BatchSys {
    Skew sk;            // layout of the skewed arrays
    T *a;               // skewed copy of A (factorization input)
    T *dl, *du;         // skewed factor: dl1..dl5, 1/dl1; du1..du4
    T *r, *w;           // skewed residual and sweep vector
    u64 *bnd, *scratch; // boundary words and per-lane scratch words
    SolveState *st;     // state of this MSI solve (the inner one in refinement)
    const T *ra;        // A, reference layout, precision T
    const T *rrhs;      // right-hand side, reference layout
    T *rx;              // iterate, reference layout
    T *rx2;             // its alternate during an MSI run (odd iteration counts)
    double *partial;    // per-block partial sums of r.r (np of them)
    int np;
    int nitm;
    double gamma;
};

}  // namespace m2d
