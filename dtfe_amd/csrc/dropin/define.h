// Compile-time configuration of the DTFE drop-in, API-compatible with the
// reference's src/define.h:23-53 (Real, component counts, PI).
#ifndef DTFE_DROPIN_DEFINE_H
#define DTFE_DROPIN_DEFINE_H

#include <cstddef>

#ifndef NO_DIM
#define NO_DIM 3
#endif
#if NO_DIM != 3
#error "the MI355X DTFE path is 3D only"
#endif

#ifdef DOUBLE
typedef double Real;  // fields computed by the all-FP64 GPU path
#else
typedef float Real;   // fields computed by the mixed path (FP64 geometry, FP32 fields)
#endif

static const size_t noVelComp = NO_DIM;
static const size_t noGradComp = NO_DIM * NO_DIM;
static const size_t noShearComp = (NO_DIM * (NO_DIM + 1)) / 2 - 1;
static const size_t noVortComp = ((NO_DIM - 1) * NO_DIM) / 2;

#ifndef NO_SCALARS
#define NO_SCALARS 1
#endif
static const size_t noScalarComp = NO_SCALARS;
static const size_t noScalarGradComp = noScalarComp * NO_DIM;

// the reference's constant (its degree-to-radian conversion uses it)
#define PI 3.14159265

#ifndef MPC_UNIT
#define MPC_UNIT 1.
#endif

#endif
