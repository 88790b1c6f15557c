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
}
