// Exports the reference's own RANMAR generator and plane-normal routines so the
// oracle's restatement can be checked against them bit for bit.  The reference
// headers are compiled where they lie (REF/src/marsaglia.h, REF/src/RandomVector.h);
// nothing is copied.  Built by oracle/Makefile into oracle/_ref/ (test use only).
// Compiled as C++ like the reference host code (nvcc hands host code to g++), so
// sqrt/acos on float arguments pick the float overloads.
#include <math.h>
#include "marsaglia.h"
#include "RandomVector.h"

extern "C" {
void ref_rmarin(int ij, int kl) { rmarin(ij, kl); }
void ref_ranmar(float* v, int n) { ranmar(v, n); }
void ref_random_vector(float* n) { GetRandomVector(n); }
void ref_random_vector_basis(float* n, float* basis) { GetRandomVectorBasis(n, basis); }
}
