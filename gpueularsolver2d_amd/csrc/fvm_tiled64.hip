// FP64 instantiation of the tiled kernels.  Built with -ffp-contract=off:
// every operation is separately rounded, as in the FP64 oracle.
#define FVM_REAL double
#include "fvm_tiled_inst.inc"
