// skg_adjoint.hip -- density adjoint and frozen-field gradient at arbitrary points.
#include "skg_common.h"

int skg_gradient_points(skg_ctx *c, const float4 *q, int nq, float *out) {
	(void)c; (void)q; (void)nq; (void)out;
	return SKG_E_STATE;
}
extern "C" int skg_density_adjoint(skg_ctx *c, const float *gbar, int parts, float *dx, float *dh) {
	(void)gbar; (void)parts; (void)dx; (void)dh;
	return c ? SKG_E_STATE : SKG_E_PARAM;
}
