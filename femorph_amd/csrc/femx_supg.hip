// femx_supg.hip -- placeholder until the SUPG kernels land.
#include "femx_common.h"
using namespace femx;
struct femx_supg { int dummy; };
#define UNSUP return set_error("SUPG not built yet"), FEMX_ERR_UNSUPPORTED
extern "C" femx_status femx_supg_create(femx_mesh *, const femx_bc_desc *, const femx_supg_params *, femx_supg **) { UNSUP; }
extern "C" femx_status femx_supg_destroy(femx_supg *) { return FEMX_OK; }
extern "C" femx_status femx_supg_set_state(femx_supg *, const double *) { UNSUP; }
extern "C" femx_status femx_supg_step(femx_supg *, int, int *) { UNSUP; }
extern "C" femx_status femx_supg_get_state(femx_supg *, double *, double *, double *, double *, double *) { UNSUP; }
extern "C" femx_status femx_supg_get_mass(femx_supg *, double *) { UNSUP; }
extern "C" femx_status femx_supg_get_log(femx_supg *, double *) { UNSUP; }
extern "C" femx_status femx_supg_sync(femx_supg *) { UNSUP; }
extern "C" void *femx_supg_stream(femx_supg *) { return nullptr; }
extern "C" femx_status femx_supg_kernel_timing(femx_supg *, int) { UNSUP; }
extern "C" int femx_supg_kernel_times(femx_supg *, femx_kernel_time *, int) { return -1; }
