// Initial-condition kernels: fields given as sums of Gaussian blobs, set on
// every box of the tree, ghost layer included (as the drivers' set_init_cond
// does for the densities, src/streamer_2d.f90:218-248).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>

#include "a5g_internal.hpp"
#include "a5g_kernels.hpp"

namespace a5g {

// v(x) = base + sum_k amp_k * exp(-|x - c_k|^2 / w_k^2) at the cell centres
// r_min + (i - 0.5) dr (a$D_r_cc, m_aX_utils.f90:309-314). With bases set,
// a box of instance g takes base bases[g] and blobs g * nblob .. (g + 1) *
// nblob - 1.
template <int D>
__global__ __launch_bounds__(256) void k_gaussians(Dev a, const int* __restrict__ list, int n,
                                                   int nc, int iv, int nblob,
                                                   const double* __restrict__ blobs,
                                                   double base,
                                                   const double* __restrict__ bases) {
  const int li = blockIdx.x;
  if (li >= n) return;
  const int id = list[li];
  double* p = boxp(a, iv, id);
  if (bases != nullptr) {
    const int g = inst_of<D>(a, id);
    base = bases[g];
    blobs += (size_t)g * nblob * (D + 2);
  }
  const int m = nc + 2;
  const int bl = D == 2 ? m * m : m * m * m;
  const double dr = a.dr[id];
  for (int q = threadIdx.x; q < bl; q += blockDim.x) {
    const int i = q % m, j = (q / m) % m, k = D == 3 ? q / (m * m) : 0;
    double x[3];
    x[0] = a.rmin[id * 3 + 0] + (i - 0.5) * dr;
    x[1] = a.rmin[id * 3 + 1] + (j - 0.5) * dr;
    x[2] = D == 3 ? a.rmin[id * 3 + 2] + (k - 0.5) * dr : 0.0;
    double v = base;
    for (int b = 0; b < nblob; b++) {
      const double* g = blobs + b * (D + 2);
      double r2 = 0;
      for (int d = 0; d < D; d++) {
        const double t = x[d] - g[d];
        r2 = r2 + t * t;
      }
      const double w = g[D];
      v = v + g[D + 1] * exp(-r2 / (w * w));
    }
    p[cidx<D>(nc, i, j, k)] = v;
  }
}

void set_gaussians(Pool& P, int iv, int nblob, const double* blobs_host, double base,
                   const double* bases_host) {
  P.sync_topology();
  const int D = P.tree.dim;
  // per instance: n_inst blob sets and bases, stored after each other
  const int sets = bases_host ? P.n_inst : 1;
  const size_t n_blob_all = (size_t)nblob * sets;
  double* d_blobs = nullptr;
  const size_t bytes = (std::max(n_blob_all, (size_t)1) * (D + 2) + sets) * sizeof(double);
  if (hipMalloc((void**)&d_blobs, bytes) != hipSuccess) throw std::runtime_error("hipMalloc");
  double* d_bases = bases_host ? d_blobs + std::max(n_blob_all, (size_t)1) * (D + 2) : nullptr;
  if (n_blob_all > 0 &&
      hipMemcpyAsync(d_blobs, blobs_host, n_blob_all * (D + 2) * sizeof(double),
                     hipMemcpyHostToDevice, P.stream) != hipSuccess)
    throw std::runtime_error("copy blobs");
  if (d_bases && hipMemcpyAsync(d_bases, bases_host, sets * sizeof(double),
                                hipMemcpyHostToDevice, P.stream) != hipSuccess)
    throw std::runtime_error("copy bases");
  for (int lvl = P.tree.min_lvl; lvl <= P.tree.max_lvl; lvl++) {
    const int n = P.n_ids(lvl);
    if (n == 0) continue;
    const int nc = P.nc_lvl(lvl);
    if (D == 2)
      hipLaunchKernelGGL((k_gaussians<2>), dim3(n), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                         n, nc, iv, nblob, d_blobs, base, d_bases);
    else
      hipLaunchKernelGGL((k_gaussians<3>), dim3(n), dim3(256), 0, P.stream, P.dev(), P.ids(lvl),
                         n, nc, iv, nblob, d_blobs, base, d_bases);
    if (hipGetLastError() != hipSuccess) throw std::runtime_error("gaussians launch");
  }
  if (hipStreamSynchronize(P.stream) != hipSuccess) throw std::runtime_error("gaussians sync");
  (void)hipFree(d_blobs);
}

}  // namespace a5g
