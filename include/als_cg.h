// The cuMF CG library's two host entry points with the reference's own
// signatures and C++ linkage (cumf_als_cg/cg.h:30,32), so a caller switches
// by including this header instead of cumf_als_cg/cg.h:
//   updateXWithCGHost          cg.h:30; host side cg.cu:683-688, kernel
//                              cg.cu:37-232
//   updateXWithCGHost_tt_fp16  cg.h:32; host side cg.cu:642-646, kernel
//                              cg.cu:236-430
// Arguments as in the reference: device pointers of the current device; A is
// batchSize x f x f (symmetric positive definite), x (warm start in,
// solution out) and b are batchSize x f; cgIter caps the iterations (the
// loop runs while iter < cgIter, so a fractional cap rounds up); the exit
// threshold on r.r is the reference's CG_ERROR, 1e-4 (cg.cu:32). Both run on
// the default stream and return once it has finished. On failure they print
// the error and exit(EXIT_FAILURE), as the reference's cudaCheckError does
// (cg.h:38-45).
// _tt_fp16: the reference reinterprets the FP32 A as half
// (cg.cu:642-646 called with als_model.cu:2003's FP32 buffer); here each
// FP32 entry is rounded to FP16 as it is loaded (the FP16-stored A that
// kernel is meant to work on). Batched calls on caller buffers without a
// default-stream sync: als_cg_update_x* in als.h.
#pragma once

void updateXWithCGHost(float *A, float *x, float *b, const int batchSize, const int f, const float cgIter);

void updateXWithCGHost_tt_fp16(float *A, float *x, float *b, const int batchSize, const int f, const float cgIter);
