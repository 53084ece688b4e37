// IB setup on the device: the two quadratic passes of the reference's
// BCM_Immersed_boundary (triangles x cubes, ghost cells x triangles) as
// kernels.  They only choose: which triangles are kept, which triangle wins
// each ghost cell's search.  The host recomputes the winners' values with the
// same functions, so the IB tables are bitwise those of the host setup.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <string>
#include <vector>

#include "bcm_geom.h"
#include "bcm_setup.h"

namespace bcm {

namespace {

#define HIP_TRY(x)                                                             \
  do {                                                                         \
    hipError_t _e = (x);                                                       \
    if (_e != hipSuccess) {                                                    \
      err = std::string(#x) + ": " + hipGetErrorString(_e);                    \
      return 1;                                                                \
    }                                                                          \
  } while (0)

constexpr int SEARCH_WG = 64;       // ghost cells per workgroup (one wave)
constexpr int SEARCH_CHUNK = 512;   // triangles per workgroup column

// keep[t] = 1 when a vertex of triangle t lies in the box of some cube padded
// by 1.5 cells (BCM_Immersed_boundary.cpp:380-470).  vert: 9 doubles per
// triangle; box: X, Y, Z corner and size per cube (slot 0 unused).
__global__ __launch_bounds__(256) void k_tri_keep(const double* __restrict__ vert, int ntri,
                                                  const double* __restrict__ box, int ncube,
                                                  uint8_t* __restrict__ keep) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntri) return;
  bool k = false;
  for (int m = 0; m < 3 && !k; ++m) {
    const double px = vert[9 * t + 3 * m], py = vert[9 * t + 3 * m + 1],
                 pz = vert[9 * t + 3 * m + 2];
    for (int c = 1; c < ncube; ++c) {
      const double X = box[4 * c], Y = box[4 * c + 1], Z = box[4 * c + 2], S = box[4 * c + 3];
      const double dx = S / NC, dy = dx, dz = dx;
      if (px >= X - 1.5 * dx && px <= X + S + 1.5 * dx && py >= Y - 1.5 * dy &&
          py <= Y + S + 1.5 * dy && pz >= Z - 1.5 * dz && pz <= Z + S + 1.5 * dz) {
        k = true;
        break;
      }
    }
  }
  keep[t] = k ? 1 : 0;
}

// Ray test of every ghost cell against one chunk of triangles: the ray from
// the cell centre along the triangle's inward normal (intersect_triangle with
// the triangle-only terms taken from SurfaceTri).  Per (chunk, cell): whether
// any triangle was hit, and the first triangle of smallest |t| (dir . dir)
// below 1e8.
__global__ __launch_bounds__(SEARCH_WG) void k_hit_search(const SurfaceTri* __restrict__ tri,
                                                          int ntri,
                                                          const double* __restrict__ xyz, int ng,
                                                          double* __restrict__ best,
                                                          int* __restrict__ idx,
                                                          int* __restrict__ any) {
  const int g = blockIdx.x * SEARCH_WG + threadIdx.x;
  const int t0 = blockIdx.y * SEARCH_CHUNK;
  const int t1 = min(ntri, t0 + SEARCH_CHUNK);
  const bool live = g < ng;
  const double ox = live ? xyz[3 * g] : 0.0, oy = live ? xyz[3 * g + 1] : 0.0,
               oz = live ? xyz[3 * g + 2] : 0.0;
  double bmin = BIGNUM;
  int bi = -1, hit = 0;
  for (int t = t0; t < t1; ++t) {
    const SurfaceTri& T = tri[t];
    if (T.det < SMALLNUM) continue;
    const double tv0 = ox - T.v0[0], tv1 = oy - T.v0[1], tv2 = oz - T.v0[2];
    const double u = tv0 * T.pv[0] + tv1 * T.pv[1] + tv2 * T.pv[2];
    if (u < 0.0 || u > T.det) continue;
    const double qv0 = tv1 * T.e1[2] - tv2 * T.e1[1];
    const double qv1 = tv2 * T.e1[0] - tv0 * T.e1[2];
    const double qv2 = tv0 * T.e1[1] - tv1 * T.e1[0];
    const double v = T.dir[0] * qv0 + T.dir[1] * qv1 + T.dir[2] * qv2;
    if (v < 0.0 || u + v > T.det) continue;
    double tt = T.e2[0] * qv0 + T.e2[1] * qv1 + T.e2[2] * qv2;
    tt = tt * T.inv_det;
    const double ndis = fabs(tt) * T.dd;
    hit = 1;
    if (ndis < bmin) {
      bmin = ndis;
      bi = t;
    }
  }
  if (live) {
    const size_t o = (size_t)blockIdx.y * ng + g;
    best[o] = bmin;
    idx[o] = bi;
    any[o] = hit;
  }
}

// Closest point of each listed ghost cell on one chunk of triangles: the
// last triangle of smallest squared distance at or below 1e8.
__global__ __launch_bounds__(SEARCH_WG) void k_near_search(const SurfaceTri* __restrict__ tri,
                                                           int ntri,
                                                           const double* __restrict__ xyz,
                                                           int nn, double* __restrict__ best,
                                                           int* __restrict__ idx) {
  const int g = blockIdx.x * SEARCH_WG + threadIdx.x;
  const int t0 = blockIdx.y * SEARCH_CHUNK;
  const int t1 = min(ntri, t0 + SEARCH_CHUNK);
  const bool live = g < nn;
  const double o[3] = {live ? xyz[3 * g] : 0.0, live ? xyz[3 * g + 1] : 0.0,
                       live ? xyz[3 * g + 2] : 0.0};
  double bmin = BIGNUM;
  int bi = -1;
  for (int t = t0; t < t1; ++t) {
    const SurfaceTri& T = tri[t];
    double p[3];
    closest_point(T.v0, T.v1, T.v2, o, p);
    const double nd = (p[0] - o[0]) * (p[0] - o[0]) + (p[1] - o[1]) * (p[1] - o[1]) +
                      (p[2] - o[2]) * (p[2] - o[2]);
    if (nd <= bmin) {
      bmin = nd;
      bi = t;
    }
  }
  if (live) {
    const size_t q = (size_t)blockIdx.y * nn + g;
    best[q] = bmin;
    idx[q] = bi;
  }
}

template <typename T>
struct DevBuf {
  T* p = nullptr;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  int alloc(size_t n, std::string& err) {
    HIP_TRY(hipMalloc((void**)&p, std::max<size_t>(n, 1) * sizeof(T)));
    return 0;
  }
};

class GpuIb : public IbDevice {
 public:
  explicit GpuIb(int device) : device_(device) {}
  ~GpuIb() override {
    if (stream_) (void)hipStreamDestroy(stream_);
  }

  int init(std::string& err) {
    HIP_TRY(hipSetDevice(device_));
    HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    return 0;
  }

  int filter(const Case& cs, std::vector<uint8_t>& keep, std::string& err) override {
    HIP_TRY(hipSetDevice(device_));
    const int ntri = (int)cs.stl.size();
    keep.assign(ntri, 0);
    if (ntri == 0) return 0;
    std::vector<double> vert((size_t)9 * ntri), box((size_t)4 * cs.ncube, 0.0);
    for (int t = 0; t < ntri; ++t)
      for (int m = 0; m < 3; ++m)
        for (int d = 0; d < 3; ++d) vert[9 * t + 3 * m + d] = cs.stl[t].v[m][d];
    for (int c = 1; c < cs.ncube; ++c) {
      box[4 * c] = cs.Xcube[c];
      box[4 * c + 1] = cs.Ycube[c];
      box[4 * c + 2] = cs.Zcube[c];
      box[4 * c + 3] = cs.cube_size[c];
    }
    DevBuf<double> dv, db;
    DevBuf<uint8_t> dk;
    if (dv.alloc(vert.size(), err) || db.alloc(box.size(), err) || dk.alloc(ntri, err)) return 1;
    HIP_TRY(hipMemcpyAsync(dv.p, vert.data(), vert.size() * sizeof(double),
                           hipMemcpyHostToDevice, stream_));
    HIP_TRY(hipMemcpyAsync(db.p, box.data(), box.size() * sizeof(double), hipMemcpyHostToDevice,
                           stream_));
    k_tri_keep<<<(ntri + 255) / 256, 256, 0, stream_>>>(dv.p, ntri, db.p, cs.ncube, dk.p);
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipMemcpyAsync(keep.data(), dk.p, ntri, hipMemcpyDeviceToHost, stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    return 0;
  }

  int search(const std::vector<SurfaceTri>& tri, const std::vector<double>& xyz,
             std::vector<int>& hit_any, std::vector<int>& hit_tri, std::vector<int>& near_tri,
             std::string& err) override {
    HIP_TRY(hipSetDevice(device_));
    const int ntri = (int)tri.size();
    const int ng = (int)(xyz.size() / 3);
    hit_any.assign(ng, 0);
    hit_tri.assign(ng, -1);
    near_tri.assign(ng, -1);
    if (ntri == 0 || ng == 0) return 0;
    const int nchunk = (ntri + SEARCH_CHUNK - 1) / SEARCH_CHUNK;
    if (nchunk > 65535) {
      err = "too many surface triangles for the device search";
      return 1;
    }
    DevBuf<SurfaceTri> dt;
    DevBuf<double> dxyz, dbest;
    DevBuf<int> didx, dany;
    const size_t nres = (size_t)nchunk * ng;
    if (dt.alloc(ntri, err) || dxyz.alloc(xyz.size(), err) || dbest.alloc(nres, err) ||
        didx.alloc(nres, err) || dany.alloc(nres, err))
      return 1;
    HIP_TRY(hipMemcpyAsync(dt.p, tri.data(), (size_t)ntri * sizeof(SurfaceTri),
                           hipMemcpyHostToDevice, stream_));
    HIP_TRY(hipMemcpyAsync(dxyz.p, xyz.data(), xyz.size() * sizeof(double),
                           hipMemcpyHostToDevice, stream_));
    dim3 grid((ng + SEARCH_WG - 1) / SEARCH_WG, nchunk);
    k_hit_search<<<grid, SEARCH_WG, 0, stream_>>>(dt.p, ntri, dxyz.p, ng, dbest.p, didx.p,
                                                  dany.p);
    HIP_TRY(hipGetLastError());
    std::vector<double> best(nres);
    std::vector<int> idx(nres), any(nres);
    HIP_TRY(hipMemcpyAsync(best.data(), dbest.p, nres * sizeof(double), hipMemcpyDeviceToHost,
                           stream_));
    HIP_TRY(hipMemcpyAsync(idx.data(), didx.p, nres * sizeof(int), hipMemcpyDeviceToHost,
                           stream_));
    HIP_TRY(hipMemcpyAsync(any.data(), dany.p, nres * sizeof(int), hipMemcpyDeviceToHost,
                           stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    // Chunks in triangle order: the first strict minimum wins, as in the
    // sequential loop.
    std::vector<int> miss;
    for (int g = 0; g < ng; ++g) {
      double m = BIGNUM;
      for (int ch = 0; ch < nchunk; ++ch) {
        const size_t o = (size_t)ch * ng + g;
        hit_any[g] |= any[o];
        if (idx[o] >= 0 && best[o] < m) {
          m = best[o];
          hit_tri[g] = idx[o];
        }
      }
      if (!hit_any[g]) miss.push_back(g);
    }
    if (miss.empty()) return 0;
    const int nn = (int)miss.size();
    std::vector<double> mxyz((size_t)3 * nn);
    for (int n = 0; n < nn; ++n)
      for (int d = 0; d < 3; ++d) mxyz[3 * n + d] = xyz[3 * miss[n] + d];
    HIP_TRY(hipMemcpyAsync(dxyz.p, mxyz.data(), mxyz.size() * sizeof(double),
                           hipMemcpyHostToDevice, stream_));
    dim3 grid2((nn + SEARCH_WG - 1) / SEARCH_WG, nchunk);
    k_near_search<<<grid2, SEARCH_WG, 0, stream_>>>(dt.p, ntri, dxyz.p, nn, dbest.p, didx.p);
    HIP_TRY(hipGetLastError());
    const size_t nres2 = (size_t)nchunk * nn;
    HIP_TRY(hipMemcpyAsync(best.data(), dbest.p, nres2 * sizeof(double), hipMemcpyDeviceToHost,
                           stream_));
    HIP_TRY(hipMemcpyAsync(idx.data(), didx.p, nres2 * sizeof(int), hipMemcpyDeviceToHost,
                           stream_));
    HIP_TRY(hipStreamSynchronize(stream_));
    // The last minimum (at or below) wins.
    for (int n = 0; n < nn; ++n) {
      double m = BIGNUM;
      for (int ch = 0; ch < nchunk; ++ch) {
        const size_t o = (size_t)ch * nn + n;
        if (idx[o] >= 0 && best[o] <= m) {
          m = best[o];
          near_tri[miss[n]] = idx[o];
        }
      }
    }
    return 0;
  }

 private:
  int device_;
  hipStream_t stream_ = nullptr;
};

}  // namespace

int build_case_on_device(const CaseParams& p, int device, Case& cs, std::string& err) {
  GpuIb dev(device);
  if (dev.init(err)) return 1;
  return build_case(p, cs, err, &dev);
}

int read_case_on_device(const std::string& dir, const CaseParams& p, int device, Case& cs,
                        std::string& err) {
  GpuIb dev(device);
  if (dev.init(err)) return 1;
  return read_case(dir, p, cs, err, &dev);
}

}  // namespace bcm
