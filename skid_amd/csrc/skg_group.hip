// skg_group.hip -- group bookkeeping and unbinding (kd.c:969-1466, grav.c).
#include "skg_common.h"

void skg_group_free(skg_ctx *c) { (void)c; }

extern "C" int skg_centers(skg_ctx *c) { return c ? SKG_E_STATE : SKG_E_PARAM; }
extern "C" int skg_unbind(skg_ctx *c, int *nUnbound) { (void)nUnbound; return c ? SKG_E_STATE : SKG_E_PARAM; }
extern "C" int skg_too_small(skg_ctx *c) { return c ? SKG_E_STATE : SKG_E_PARAM; }
extern "C" int skg_get_groups(skg_ctx *c, int *grp, int *nGroup) {
	if (!c) return SKG_E_PARAM;
	if (c->stage < 5) return SKG_E_STATE;
	if (grp) { SKG_HIP(hipMemcpy(grp, c->d_grp, 4 * (size_t)c->n, hipMemcpyDeviceToHost)); }
	if (nGroup) *nGroup = c->nGroup;
	return SKG_OK;
}
extern "C" int skg_get_catalogue(skg_ctx *c, float *mass, float *center, float *vcm, float *rbound,
		float *radius, int *nMembers) {
	(void)mass; (void)center; (void)vcm; (void)rbound; (void)radius; (void)nMembers;
	return c ? SKG_E_STATE : SKG_E_PARAM;
}
