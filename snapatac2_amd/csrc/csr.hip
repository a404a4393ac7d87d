// Column reductions over a CSR count matrix.
#include "common.h"

namespace snapgpu {

template <typename IdxT>
__global__ void __launch_bounds__(256) col_sum_kernel(const IdxT* __restrict__ idx,
                                                       const uint32_t* __restrict__ val, int64_t nnz,
                                                       unsigned long long* __restrict__ out) {
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; e < nnz; e += stride)
    atomicAdd(&out[idx[e]], (unsigned long long)val[e]);
}

}  // namespace snapgpu

using namespace snapgpu;

extern "C" {

int snapgpu_csr_col_sum_u32(const void* d_indices, int32_t index_bits, const uint32_t* d_values,
                            int64_t nnz, uint64_t* d_out, void* stream) {
  clear_error();
  SG_ARG(nnz >= 0, "nnz < 0");
  SG_ARG(index_bits == 32 || index_bits == 64, "index_bits must be 32 or 64");
  if (nnz == 0) return SNAPGPU_OK;
  SG_ARG(d_indices && d_values && d_out, "null pointer");
  const int64_t blocks = std::min<int64_t>(ceil_div(nnz, 256), 256 * 64);
  auto* out = reinterpret_cast<unsigned long long*>(d_out);
  if (index_bits == 32)
    col_sum_kernel<int32_t><<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(
        static_cast<const int32_t*>(d_indices), d_values, nnz, out);
  else
    col_sum_kernel<int64_t><<<(unsigned)blocks, 256, 0, as_stream(stream)>>>(
        static_cast<const int64_t*>(d_indices), d_values, nnz, out);
  SG_LAUNCH_CHECK();
  return SNAPGPU_OK;
}

}  // extern "C"
