// Fused middle of the fluid step:
//   curl (CalcVorticity.comp) -> confinement (ApplyVorticity.comp)
//   -> [viscosity (Viscosity.comp)] -> divergence (Divergence.comp)
//   -> first Jacobi sweep from the zero guess (ClearBuffer + Poisson.comp).
//
// One 64x16 output tile per 256-thread block.  The advected velocity is read
// once into LDS with a halo of 3 (4 with viscosity) and every intermediate
// field lives in LDS on a shrinking halo, so the kernel reads velocity + mask
// and writes velocity, divergence and p1: 8+1 -> 8+4+4 bytes per cell instead
// of the 12+20+(16)+13+4+13 bytes the separate passes move.
//
// Each LDS slot holds its stage's value at the CLAMPED grid coordinate of the
// slot, and neighbours are looked up at clamp(c +- 1) as common.sh:14-17 does,
// so border tiles need no special case and every value is bitwise the one
// the separate per-stage kernels produce.
#include "nx_fused_tile.h"
#include "nx_kernels.h"

namespace nx {

namespace {

using fused::TX;
using fused::TY;

__global__ void __launch_bounds__(256)
    k_fused_vort_div(const float2* __restrict__ uin,
                     const uint8_t* __restrict__ mask,
                     float2* __restrict__ uout, float* __restrict__ div,
                     float* __restrict__ p1, int W, int H, float dt,
                     float scale, int viscous, float alpha, float rbeta) {
  // 1D grid of TX x TY tiles in XCD order
  const int gx = (W + TX - 1) / TX;
  const int t = xcd_tile((int)blockIdx.x, gx * ((H + TY - 1) / TY));
  fused::fused_vort_div_tile((t % gx) * TX, (t / gx) * TY, uin, mask, uout, div, p1, W, H,
                             dt, scale, viscous, alpha, rbeta);
}

}  // namespace

hipError_t launch_fused_vort_div(const float2* uin, const uint8_t* mask,
                                 float2* uout, float* div, float* p1, int W,
                                 int H, float dt, float scale, int viscous,
                                 float alpha, float rbeta, hipStream_t s) {
  const dim3 grid(((W + TX - 1) / TX) * ((H + TY - 1) / TY), 1, 1);
  k_fused_vort_div<<<grid, dim3(64, 4), 0, s>>>(uin, mask, uout, div, p1, W,
                                                H, dt, scale, viscous, alpha,
                                                rbeta);
  return hipGetLastError();
}

}  // namespace nx
