// tapas_exafmm_gpu.h -- what an FMM application written for the reference
// includes in place of EX/examples/tapas_exafmm.h and the two kernel headers
// LaplaceSphericalCPU_tapas.h / LaplaceP2PCPU_tapas.h (EX =
// sample/exafmm-dev-13274dd4ac68): the types (real_t, vec3, Body, BodyAttr,
// CellAttr), TapasFMM and the kernel names P2M, M2M, M2L, L2L, L2P and P2P,
// all backed by the MI355X library (tapas_fmm.h).  With it, the reference's
// FMM_Upward / FMM_DTT / FMM_Downward functors (tapas_exafmm.cxx:87-286) and
// its main's TapasFMM::Map calls compile unchanged.
#ifndef TAPAS_GPU_TAPAS_EXAFMM_GPU_H_
#define TAPAS_GPU_TAPAS_EXAFMM_GPU_H_

#include "tapas_fmm.h"

using tapas_gpu::complex_t;
using tapas_gpu::index_t;
using tapas_gpu::kvec4;
using tapas_gpu::real_t;
using tapas_gpu::vec3;
using tapas_gpu::vecP;

using tapas_gpu::Body;
using tapas_gpu::BodyAttr;
using tapas_gpu::CellAttr;
typedef tapas_gpu::TapasFMM TapasFMM;

using tapas_gpu::L2L;
using tapas_gpu::L2P;
using tapas_gpu::M2L;
using tapas_gpu::M2M;
using tapas_gpu::P2M;
using tapas_gpu::P2P;

#endif  // TAPAS_GPU_TAPAS_EXAFMM_GPU_H_
