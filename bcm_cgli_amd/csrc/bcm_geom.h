// Ray/triangle and closest-point tests of the IB setup, shared by the host
// setup and the device search (one definition, so both compute the same
// values operation by operation).
#pragma once

#include <hip/hip_runtime.h>
#include <math.h>

#include "bcm_layout.h"

namespace bcm {

// Möller–Trumbore ray/triangle test as used by the reference
// (BCM_Immersed_boundary.cpp:1437-1495): back faces and degenerate triangles
// are rejected by det < 1e-7, no sign test on t.
__host__ __device__ inline int intersect_triangle(const double orig[3], const double dir[3],
                                                  const double v0[3], const double v1[3],
                                                  const double v2[3], double* t, double* u,
                                                  double* v) {
  double e1[3], e2[3], tv[3], pv[3], qv[3];
  for (int d = 0; d < 3; ++d) { e1[d] = v1[d] - v0[d]; e2[d] = v2[d] - v0[d]; }
  pv[0] = dir[1] * e2[2] - dir[2] * e2[1];
  pv[1] = dir[2] * e2[0] - dir[0] * e2[2];
  pv[2] = dir[0] * e2[1] - dir[1] * e2[0];
  double det = e1[0] * pv[0] + e1[1] * pv[1] + e1[2] * pv[2];
  if (det < SMALLNUM) return 0;
  for (int d = 0; d < 3; ++d) tv[d] = orig[d] - v0[d];
  *u = tv[0] * pv[0] + tv[1] * pv[1] + tv[2] * pv[2];
  if (*u < 0.0 || *u > det) return 0;
  qv[0] = tv[1] * e1[2] - tv[2] * e1[1];
  qv[1] = tv[2] * e1[0] - tv[0] * e1[2];
  qv[2] = tv[0] * e1[1] - tv[1] * e1[0];
  *v = dir[0] * qv[0] + dir[1] * qv[1] + dir[2] * qv[2];
  if (*v < 0.0 || *u + *v > det) return 0;
  *t = e2[0] * qv[0] + e2[1] * qv[1] + e2[2] * qv[2];
  double inv_det = 1.0 / det;
  *t *= inv_det;
  *u *= inv_det;
  *v *= inv_det;
  return 1;
}

// Closest point on a triangle (Ericson's region tests) as restated from
// BCM_Immersed_boundary.cpp:1286-1432, including the interior-region
// expression w = vc + denom of :1425.
__host__ __device__ inline void closest_point(const double a[3], const double b[3],
                                              const double c[3], const double p[3], double out[3]) {
  double ab[3], ac[3], bc[3], ap[3], bp[3], cp[3];
  for (int d = 0; d < 3; ++d) {
    ab[d] = b[d] - a[d]; ac[d] = c[d] - a[d]; bc[d] = c[d] - b[d]; ap[d] = p[d] - a[d];
  }
  double d1 = ab[0] * ap[0] + ab[1] * ap[1] + ab[2] * ap[2];
  double d2 = ac[0] * ap[0] + ac[1] * ap[1] + ac[2] * ap[2];
  if (d1 <= 0. && d2 <= 0.) { for (int d = 0; d < 3; ++d) out[d] = a[d]; return; }
  for (int d = 0; d < 3; ++d) bp[d] = p[d] - b[d];
  double d3 = ab[0] * bp[0] + ab[1] * bp[1] + ab[2] * bp[2];
  double d4 = ac[0] * bp[0] + ac[1] * bp[1] + ac[2] * bp[2];
  if (d3 >= 0. && d4 <= d3) { for (int d = 0; d < 3; ++d) out[d] = b[d]; return; }
  double vc = d1 * d4 - d3 * d2;
  if (vc <= 0. && d1 >= 0. && d3 <= 0.) {
    double v = d1 / (d1 - d3);
    for (int d = 0; d < 3; ++d) out[d] = a[d] + v * ab[d];
    return;
  }
  for (int d = 0; d < 3; ++d) cp[d] = p[d] - c[d];
  double d5 = ab[0] * cp[0] + ab[1] * cp[1] + ab[2] * cp[2];
  double d6 = ac[0] * cp[0] + ac[1] * cp[1] + ac[2] * cp[2];
  if (d6 >= 0. && d5 <= d6) { for (int d = 0; d < 3; ++d) out[d] = c[d]; return; }
  double vb = d5 * d2 - d1 * d6;
  if (vb <= 0. && d2 >= 0. && d6 <= 0.) {
    double w = d2 / (d2 - d6);
    for (int d = 0; d < 3; ++d) out[d] = a[d] + w * ac[d];
    return;
  }
  double va = d3 * d6 - d5 * d4;
  if (va <= 0. && (d4 - d3) >= 0. && (d5 - d6) >= 0.) {
    double w = (d4 - d3) / ((d4 - d3) + (d5 - d6));
    for (int d = 0; d < 3; ++d) out[d] = b[d] + w * bc[d];
    return;
  }
  double denom = 1. / (va + vb + vc);
  double v = vb * denom;
  double w = vc + denom;
  for (int d = 0; d < 3; ++d) out[d] = a[d] + ab[d] * v + ac[d] * w;
}

}  // namespace bcm
