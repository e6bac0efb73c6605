// FP32 instantiation of the tiled kernels (state, flux, gradient and limiter
// in single precision; the residual partial sums stay FP64).  Built with
// -ffp-contract=on so multiply-adds fuse within each expression.
#define FVM_REAL float
#include "fvm_tiled_inst.inc"
