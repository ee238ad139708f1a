// Library state, error reporting and small C-ABI helpers.
#include <cstring>
#include <string>

#include "../../include/pmgpu.h"
#include "pm_common.h"

namespace pmgpu {

static thread_local std::string g_err;
static int g_lds_limit = 0;

void set_error(const std::string &msg) { g_err = msg; }
const std::string &get_error() { return g_err; }

int device_lds_limit() {
  if (g_lds_limit == 0) {
    int dev = 0, v = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
      g_lds_limit = v < 65536 ? v : 65536;
    else
      g_lds_limit = 65536;
  }
  return g_lds_limit;
}

}  // namespace pmgpu

using namespace pmgpu;

extern "C" {

int pmgpu_init(int32_t device) {
  int n = 0;
  PM_HIP(hipGetDeviceCount(&n));
  PM_CHECK(device >= 0 && device < n, "pmgpu_init: no such HIP device");
  PM_HIP(hipSetDevice(device));
  device_lds_limit();
  return pmgpu::penalise_noalloc_init();
}

int pmgpu_finalize(void) {
  PM_HIP(hipDeviceSynchronize());
  pmgpu::particles_release();
  pmgpu::penalise_release();
  pmgpu::diagnostics_release();
  pmgpu::m2m_release();
  pmgpu::penalisation_solve_release();
  return 0;
}

int pmgpu_last_error(char *buf, size_t len) {
  if (!buf || len == 0) return -1;
  std::strncpy(buf, g_err.c_str(), len - 1);
  buf[len - 1] = 0;
  return 0;
}

int pmgpu_mesh_extent(const pmgpu_geom *g, int32_t ext[3]) {
  PM_CHECK(g && ext, "pmgpu_mesh_extent: null argument");
  for (int d = 0; d < 3; ++d) {
    int ng = g->bound_cond[d] == 0 ? g->nghost : (g->interp_order == 4 ? 3 : 2);
    ext[d] = g->ncell[d] + 2 * ng;
  }
  return 0;
}

}  // extern "C"
