// gfx950 kernels of the N-fold way (the body, shared with the run-time
// specialised kernels, is in clexmc_nfold.h).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "clexmc_nfold.h"

namespace clexmc {

size_t nfold_lds_bytes(int pad_sites, int n_sites, int nb_max, int nb_total) {
  const int n_groups = (n_sites + 63) / 64;
  return nfold_align16((size_t)image_bytes(pad_sites, false)) + nfold_align16((size_t)n_groups * 64 * 4) +
         nfold_align16((size_t)n_groups * 8) + 2 * nfold_align16((size_t)nb_max * 4) + 16 + kNfoldNbBeginBytes +
         nfold_align16((size_t)nb_total * 16) + (size_t)kMaxSublat * (1 + kMaxOcc) * 4;
}

template <bool BIN>
__global__ __launch_bounds__(256) void nfold_kernel(NfoldArgs na) {
  nfold_body<BIN, GenericField>(na);
}

hipError_t launch_nfold(const NfoldArgs &a, bool binary, int n_chains, int threads, size_t lds, hipStream_t st) {
  const void *fn = binary ? (const void *)nfold_kernel<true> : (const void *)nfold_kernel<false>;
  hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e != hipSuccess) return e;
  if (binary)
    hipLaunchKernelGGL(nfold_kernel<true>, dim3(n_chains), dim3(threads), lds, st, a);
  else
    hipLaunchKernelGGL(nfold_kernel<false>, dim3(n_chains), dim3(threads), lds, st, a);
  return hipGetLastError();
}

}  // namespace clexmc
