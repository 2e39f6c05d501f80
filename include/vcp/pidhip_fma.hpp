// pidhip_fma.hpp -- vcp::pidhip_fma, a drop-in for vcp::pidblas_fma
// (vcp/pidblas_fma.hpp:43-72): its interval x interval product is the
// inf-sup chain of interval FMAs (vcp/ifma.hpp:44-137), computed on an MI355X
// by libvcphip's VCPHIP_FORM_INFSUP; the other hooks are those of
// vcp::pidhip.
//
//   #include <vcp/pidhip_fma.hpp>
//   vcp::matrix< kv::interval<double>, vcp::pidhip_fma > A, B, C;
//   C = A * B;            // GPU inf-sup product, same bits as pidblas_fma
//                         // (up to the sign of zero ends)
//
// Products below the inf-sup crossover (vcphip_policy::crossover(
// CROSS_INFSUP), VCPHIP_CROSSOVER_FMA) run pidblas_fma's own host loop.
#pragma once

#ifndef VCP_PIDHIP_FMA_HPP
#define VCP_PIDHIP_FMA_HPP

#include <vcp/pidblas_fma.hpp>

#include "pidhip.hpp"

namespace vcp {

typedef pidhip_t<vcp::pidblas_fma, VCPHIP_FORM_INFSUP> pidhip_fma;

}  // namespace vcp

#endif  // VCP_PIDHIP_FMA_HPP
