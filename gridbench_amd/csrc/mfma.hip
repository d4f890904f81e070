// MFMA form of the multi-right-hand-side hopping term (placeholder until
// the matrix-core kernel lands: reports GBX_EUNSUPPORTED).
#include "gbx_internal.h"

int mfma_supported(const gbx_gauge_s *u, const gbx_spinor_s *in)
{
  (void)u;
  (void)in;
  return 0;
}

int mfma_launch(const gbx_gauge_s *u, gbx_spinor_s *out, const gbx_spinor_s *in, int dag, const gbx_kcfg &c, hipStream_t st)
{
  (void)u;
  (void)out;
  (void)in;
  (void)dag;
  (void)c;
  (void)st;
  return gbx_fail(GBX_EUNSUPPORTED, "MFMA kernel not available for this configuration");
}
