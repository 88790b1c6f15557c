// k_ext.hip -- placeholder for the strain-volume columns until that kernel lands.
#include "ucy_ext.h"
#include "../../include/ucy.h"
namespace ucy {
int ext_build_volume_columns(hipStream_t, double, double, int, const double *, const double *,
                             const double *, const double *, const double *, const Receiver *,
                             int, double *, int64_t, int64_t, std::string &msg)
{
    msg = "strain-volume sources not available in this build";
    return UCY_ERR_STATE;
}
}  // namespace ucy

extern "C" {
struct ucy_ctx;
int ucy_strainvolume_traction_kernels(ucy_ctx *, const double *, const double *, const double *,
                                      const double *, int, const double *, const double *,
                                      const double *, const double *, const double *, double,
                                      double, double, double, double, double, double, double,
                                      double *, double *, double *) { return UCY_ERR_STATE; }
int ucy_strainvolume_stress_kernels(ucy_ctx *, const double *, const double *, const double *,
                                    const double *, int, const double *, const double *,
                                    const double *, const double *, const double *, double, double,
                                    double, double, double, double, double, double, double *,
                                    double *, double *, double *, double *, double *) { return UCY_ERR_STATE; }
}
