// internal.hpp -- internal interfaces between the translation units of
// libspira_amd (not part of the C-ABI).
#pragma once

#include "common.hpp"

namespace spira_amd {

// scan.hip: out[0..n) = exclusive prefix of in, out[n] = total.  in may equal out.
void scan_excl_u64(spira_dev *dev, const uint64_t *in, uint64_t *out, uint64_t n);

// spmv.hip
void spmv_build_plan(spira_dcsr *A);
void spmv_free_plan(spira_dcsr *A);
int spmv_heuristic_variant(const spira_dcsr *A);
void spmv_launch_f64(spira_dcsr *A, const double *x, double *y, int variant);

// lock.hip: runs the device lock pipeline on already-device-resident staged
// arrays; installs the new CSR in A.  Returns the new nnz.
uint64_t lock_device_f64(spira_dcsr *A, const uint64_t *d_stage_ptr, const uint32_t *d_stage_col,
                         const double *d_stage_val, uint64_t m, const uint8_t *d_dirty);

// validate.hip: load_csr checks on device arrays; throws status_error.
void validate_csr_device(spira_dcsr *A, uint64_t nnz, const uint64_t *d_off, const uint32_t *d_col,
                         const double *d_val, bool *has_zeros);

// Synchronise the handle's stream (throws on failure).
inline void dev_sync(spira_dev *dev) { SPIRA_HIP(hipStreamSynchronize(dev->stream)); }

// Read one device word.
template <class T>
inline T read_scalar(spira_dev *dev, const T *dptr)
{
    T v{};
    SPIRA_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, dev->stream));
    dev_sync(dev);
    return v;
}

void free_dcsr_arrays(spira_dcsr *A);

}  // namespace spira_amd
