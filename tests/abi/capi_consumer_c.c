/*
 * The C side of the ABI consumer: include/mnl.h compiled as C99 and called
 * from C (no device work). Returns 0 when every check holds.
 */
#include <string.h>

#include "mnl.h"

int capi_consumer_c_checks(void) {
    mnl_request rq;
    memset(&rq, 0, sizeof rq);
    rq.backend = MNL_BACKEND_AUTO;
    rq.order = MNL_ORDER_LINEAR;
    if (mnl_neighbour_list_request(NULL, NULL) != MNL_ERROR) return 1;
    if (strlen(mnl_last_error()) == 0) return 2;
    if (MNL_Q_FIRST != 1 || MNL_Q_SHIFT != 16) return 3;
    if (mnl_plan_device(NULL) != -1) return 4;
    return 0;
}
