// Gamma-point fold of the assembled blocks into a dense matrix: the consumer
// side of the path (build_Hmatrix, h.SOLVESH/M_kspace_Lowdin.f90:180-271,
// 630-724, at k = 0 every image phase is 1).  H = (T + Vna block) + Vna blocko,
// the in-scope terms; S = S.  Only the upper triangle is consumed by the
// eigensolver, as dsyev('V', 'U') reads it in the reference
// (M_diagonalization_Lowdin_gamma.f90:173-178).
#include "fb_internal.h"

// one wavefront per atom of the structure: the atom's rows are owned by its
// workgroup and lane e adds element e of each of its pair blocks in
// neighbour order, so the fold is race-free and deterministic
__global__ __launch_bounds__(64) void k_gamma_fold(DView v, int a0, int which, const int *orb_off, int n, double *A) {
  const int il = blockIdx.x;
  const int ia = a0 + il;
  const int e = threadIdx.x;
  const int nmu = v.sp.norb[v.imass[ia] - 1];
  const int ro = orb_off[il];
  for (int64_t p = v.noff[ia]; p < v.noff[ia + 1]; p++) {
    const int jl = v.neigh_j[p] - 1;
    const int nnu = v.sp.norb[v.imass[a0 + jl] - 1];
    if (e < nmu * nnu) {
      const int m = e % nmu, q = e / nmu;
      const int64_t b = v.boff[p] + e;
      const double val = which == 0 ? v.mat[FB_MAT_S][b]
                                    : (v.mat[FB_MAT_T][b] + v.mat[FB_MAT_VNA][b]) + v.mat[FB_MAT_VNAO][b];
      A[(int64_t)(orb_off[jl] + q) * n + ro + m] += val;
    }
  }
}

__global__ void k_fill_f64(int64_t n, double *a, double val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    a[i] = val;
}

extern "C" int fbk_gamma_fold(const DView &v, int a0, int natoms, int which, const int *orb_off, int n, double *A,
                              hipStream_t st) {
  const int64_t nn = (int64_t)n * n;
  hipLaunchKernelGGL(k_fill_f64, dim3(2048), dim3(256), 0, st, nn, A, 0.0);
  hipLaunchKernelGGL(k_gamma_fold, dim3(natoms), dim3(64), 0, st, v, a0, which, orb_off, n, A);
  return (int)hipGetLastError();
}
