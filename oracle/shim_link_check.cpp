// shim_link_check.cpp -- link check of host/bgw_reference_api.hpp (TEST INFRASTRUCTURE).
//
// A translation unit shaped like the reference's GPP/gpp.cpp and FF/ff.cpp once they call
// this library through the header shim: it includes the reference's common/Defines.h, then
// the shim, and calls every kernel of the reference API with the reference's ArrayMD types.
// oracle/Makefile (target shim-check) compiles and links it against libbgw in both of
// the reference's complex modes; it is never run (no GPU needed).
#include "common/Defines.h"
#include "bgw_reference_api.hpp"

int main() {
  const int nb = 2, ng = 4, nc = 8, nf = 3, nfe = 2;
  ARRAY1D_int inv(ng), indinv(nc + 1);
  ARRAY1D_DataType wx(3), vcoul(nc), ekq(nb), grid(nf);
  ARRAY2D wtilde(ng, nc), aqsm(nb, nc), aqsn(nb, nc), I_eps(ng, nc);
  ARRAY3D R(nf, ng, nc), A(nf, ng, nc);
  ARRAY1D achtemp(3), asx(nfe), cor(nfe);
  ComplexType achs, band;
  try {
    // declared in common/Defines.h:36-62
    noflagOCC_solver(nb, ng, nc, inv, indinv, wx, wtilde, aqsm, aqsn, I_eps, vcoul, achtemp);
    schDttt_corKernel1(band, inv, indinv, R, A, aqsm, aqsn, vcoul, nc, 0, ng, 0, 0.5, 0.5);
    schDttt_corKernel2(band, inv, indinv, R, A, aqsm, aqsn, vcoul, nc, 0, ng, 0, 0.5, 0.5);
    // defined in FF/ff.cpp:171-302 (no header declaration)
    achsDtemp_Kernel(nb, ng, nc, inv, indinv, aqsn, aqsm, R, vcoul, achs);
    asxDtemp_Kernel(1, nfe, nc, ng, nf, 0.0, 0.5, 1.0, ekq, grid, inv, indinv, aqsm, aqsn, vcoul, R, A, asx);
    achDtemp_cor_Kernel(nb, 1, nfe, nc, ng, nf, 0.0, 0.5, ekq, grid, inv, indinv, aqsm, aqsn, vcoul, R, A, cor);
  } catch (const bgw::Error&) {
    return 1;
  }
  return 0;
}
