// Pointers into a proof workspace, for proofs batched in lockstep.
//
// A batch of B proofs of one shape holds B copies of the workspace at a fixed
// stride S in one allocation, and every kernel of the proof is launched once
// for the whole batch with gridDim.z = B: workgroup z works on job z's copy.
// A kernel's pointer arguments point into job 0's copy and carry S in their
// top 16 bits (in 64 KiB units); JP<T> turns such a pointer into job
// blockIdx.z's address. A pointer without a tag (every pointer of an unbatched
// proof, and tables all jobs share: twiddles, the shape's claim descriptors)
// is used as it is. Device addresses are below 2^48, which the batch checks
// when it allocates.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace kb {

constexpr int JOB_TAG_SHIFT = 48;
constexpr uint64_t JOB_ADDR_MASK = (1ull << JOB_TAG_SHIFT) - 1;
constexpr int JOB_STRIDE_UNIT_LOG = 16;  // 64 KiB

// Host: job 0's pointer p tagged with the batch stride (a multiple of 64 KiB below 4 GiB).
template <class T>
inline T* job_tag(T* p, uint64_t stride) {
  if (!p) return p;
  return (T*)((uint64_t)p | ((stride >> JOB_STRIDE_UNIT_LOG) << JOB_TAG_SHIFT));
}
// Host: the address of job j's copy behind a (possibly tagged) pointer.
template <class T>
inline T* job_addr(T* p, uint32_t j) {
  const uint64_t v = (uint64_t)p;
  return (T*)((v & JOB_ADDR_MASK) + (uint64_t)j * ((v >> JOB_TAG_SHIFT) << JOB_STRIDE_UNIT_LOG));
}

template <class T>
struct JP {
  T* raw;
  __host__ __device__ JP(T* p = nullptr) : raw(p) {}
  template <class U>
  __host__ __device__ JP(const JP<U>& o) : raw(o.raw) {}
  // Device: this workgroup's job's address (blockIdx.z). Host: the pointer as
  // it is, tag included, to be passed on to a kernel (job_addr gives a job's
  // address to host code).
  // On the device the address is formed as a global-space pointer and then
  // converted to a generic one: the compiler then knows that every access
  // through it is global memory and emits global_load / global_store (waited
  // for with in-order vmcnt counts) instead of flat accesses (which also count
  // on lgkmcnt, so a wait on LDS or scalar loads waits for them too). Every
  // JP holds device (or host-mapped) memory, never LDS or scratch.
  __host__ __device__ __forceinline__ T* get() const {
#if defined(__HIP_DEVICE_COMPILE__)
    typedef __attribute__((address_space(1))) char gchar;
    const uint64_t v = (uint64_t)raw;
    gchar* g = (gchar*)(v & JOB_ADDR_MASK);
    return (T*)(g + (uint64_t)blockIdx.z * ((v >> JOB_TAG_SHIFT) << JOB_STRIDE_UNIT_LOG));
#else
    return raw;
#endif
  }
  __host__ __device__ __forceinline__ operator T*() const { return get(); }
  __host__ __device__ __forceinline__ T* operator->() const { return get(); }
};

}  // namespace kb
