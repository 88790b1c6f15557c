// k_ext.hip -- placeholder until the triangle and strain-volume kernels land.
#include "ucy_ext.h"
#include "../../include/ucy.h"
namespace ucy {
int ext_build_triangle_columns(hipStream_t, double, double, int, const int *, const int *,
                               const int *, int, const double *, const Receiver *, int, int,
                               double *, int64_t, int64_t, std::string &msg)
{
    msg = "triangle sources not available in this build";
    return UCY_ERR_STATE;
}
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
struct ucy_ensemble;
int ucy_nikkhoo15_traction_kernels(ucy_ctx *, const double *, const double *, const double *,
                                   const double *, int, const int *, const int *, const int *, int,
                                   const double *, double, double, double, double, double, double *,
                                   double *, double *) { return UCY_ERR_STATE; }
int ucy_nikkhoo15_stress_kernels(ucy_ctx *, const double *, const double *, const double *,
                                 const double *, int, const int *, const int *, const int *, int,
                                 const double *, double, double, double, double, double, double *,
                                 double *, double *, double *, double *, double *) { return UCY_ERR_STATE; }
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
ucy_ensemble *ucy_ensemble_create(ucy_ctx *, int) { return nullptr; }
void ucy_ensemble_destroy(ucy_ensemble *) {}
int ucy_ensemble_add_group(ucy_ensemble *, int, const double *, const double *, const double *,
                           const double *, const double *, double, double, int, const double *,
                           const double *, const double *, const double *, const double *,
                           const double *, const double *, const double *) { return UCY_ERR_STATE; }
int ucy_ensemble_build(ucy_ensemble *) { return UCY_ERR_STATE; }
int ucy_ensemble_run(ucy_ensemble *, double, int64_t, double, double, ucy_run_stats *) { return UCY_ERR_STATE; }
int ucy_ensemble_get_state(ucy_ensemble *, int, double *, double *, double *, int64_t *, int64_t *) { return UCY_ERR_STATE; }
}
