// ucy_ext.h -- assembly of triangle (Nikkhoo & Walter 2015) and strain-volume
// (Barbot, Moore & Lambert 2017) source columns of K.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>
#include "ucy_internal.h"

namespace ucy {
// Columns [col0, col0 + 2*ntri) of K for every receiver (np patches then nvol
// volumes, frames in d_R). Returns 0 or a UCY_ERR_* status with msg set.
int ext_build_triangle_columns(hipStream_t st, double G, double lambda, int ntri, const int *i1,
                               const int *i2, const int *i3, int nve, const double *vert,
                               const Receiver *d_R, int np, int nvol, double *K, int64_t ldk,
                               int64_t col0, std::string &msg);
// Columns [col0, col0 + 6*nvol) of K.
int ext_build_volume_columns(hipStream_t st, double G, double lambda, int nvol, const double *x,
                             const double *L, const double *T, const double *W,
                             const double *strike, const Receiver *d_R, int np, double *K,
                             int64_t ldk, int64_t col0, std::string &msg);
// Displacement kernels at npts observation points (d_pts (3,npts), device),
// buildO (greens.f90:302-462): rows 3k..3k+2 of O (row stride ldo); triangle
// columns [col0, col0 + 2*ntri), volume columns [col0, col0 + 6*nvol).
int ext_build_triangle_disp(hipStream_t st, double G, double lambda, int ntri, const int *i1,
                            const int *i2, const int *i3, const double *vert, const double *d_pts,
                            int npts, double *O, int64_t ldo, int64_t col0, std::string &msg);
int ext_build_volume_disp(hipStream_t st, double G, double lambda, int nvol, const double *x,
                          const double *L, const double *T, const double *W, const double *strike,
                          const double *pts, const double *d_pts, int npts, double *O,
                          int64_t ldo, int64_t col0, std::string &msg);
}  // namespace ucy
