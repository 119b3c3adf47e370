// Measured issue peaks of the resources that bound the enumeration kernel
// (mce_device_peaks): 32-bit integer VALU lane-operations per second and LDS
// read bytes per second, for the roofline t_roof = max(HBM bytes / HBM peak,
// OPS / VALU peak, LDS bytes / LDS peak). No counterpart in the reference.
//
// VALU: every lane runs eight independent chains of v_and / v_xor / v_bcnt
// (written as instructions, so the count per iteration is exact), enough
// waves to fill every SIMD four times over. LDS: every lane reads 16-byte
// words at addresses that put the 64 lanes of a wave in distinct banks.
#include <hip/hip_runtime.h>

#include <chrono>

#include "engine.hpp"
#include "internal.hpp"

namespace mce {
namespace {

constexpr int kValuIters = 4096;
constexpr int kValuPerIter = 24;  // instructions per iteration, per lane
constexpr int kLdsIters = 4096;
constexpr int kLdsWords = 2048;  // uint4 words of LDS per workgroup (32 KiB: five workgroups per CU)

__global__ __launch_bounds__(256) void valu_peak_kernel(const uint32_t *in, uint32_t *out) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  uint32_t a0 = in[0] ^ t, a1 = in[1] ^ t, a2 = in[2] ^ t, a3 = in[3] ^ t, a4 = in[4] ^ t, a5 = in[5] ^ t,
           a6 = in[6] ^ t, a7 = in[7] ^ t;
  const uint32_t b = in[8], c = in[9];
  for (int i = 0; i < kValuIters; ++i) {
#define MCE_STEP(a)                                                          \
  asm volatile("v_and_b32 %0, %0, %1\n\t"                                    \
               "v_xor_b32 %0, %0, %2\n\t"                                    \
               "v_bcnt_u32_b32 %0, %0, %0"                                   \
               : "+v"(a)                                                     \
               : "v"(b), "v"(c));
    MCE_STEP(a0) MCE_STEP(a1) MCE_STEP(a2) MCE_STEP(a3) MCE_STEP(a4) MCE_STEP(a5) MCE_STEP(a6) MCE_STEP(a7)
#undef MCE_STEP
  }
  const uint32_t r = a0 ^ a1 ^ a2 ^ a3 ^ a4 ^ a5 ^ a6 ^ a7;
  if (r == 0x9E3779B9u) out[t] = r;  // keeps the chains live
}

__global__ __launch_bounds__(256) void lds_peak_kernel(const uint32_t *in, uint32_t *out) {
  __shared__ uint4 buf[kLdsWords];
  for (int i = threadIdx.x; i < kLdsWords; i += blockDim.x)
    buf[i] = make_uint4(in[0] + i, in[1] ^ i, in[2] + 2 * i, in[3] ^ 3 * i);
  __syncthreads();
  // LDS byte address of this lane's first word; consecutive lanes read
  // consecutive 16-byte words. The reads are instructions (the compiler may
  // not hoist them out of the loop), four per block, waited for in the block.
  const uint32_t base = (uint32_t)(uintptr_t)buf + threadIdx.x * 16u;
  typedef uint32_t u32x4 __attribute__((ext_vector_type(4)));
  u32x4 acc = {0, 0, 0, 0};
  uint32_t k = 0;
  for (int i = 0; i < kLdsIters; i += 4) {
    u32x4 a, b, c, d;  // each a 4-register tuple: what one ds_read_b128 writes
    const uint32_t addr = base + k;
    asm volatile(
        "ds_read_b128 %0, %4\n\t"
        "ds_read_b128 %1, %4 offset:4096\n\t"
        "ds_read_b128 %2, %4 offset:8192\n\t"
        "ds_read_b128 %3, %4 offset:12288\n\t"
        "s_waitcnt lgkmcnt(0)"
        : "=v"(a), "=v"(b), "=v"(c), "=v"(d)
        : "v"(addr)
        : "memory");
    acc ^= a ^ b ^ c ^ d;
    k = (k + 16384u) & (kLdsWords * 16u - 1u);
  }
  const uint32_t r = acc.x ^ acc.y ^ acc.z ^ acc.w;
  if (r == 0x9E3779B9u) out[blockIdx.x * blockDim.x + threadIdx.x] = r;
}

}  // namespace
}  // namespace mce

using namespace mce;

extern "C" int mce_device_peaks(int device, double *valu_lane_ops_per_s, double *lds_bytes_per_s) {
  clear_error();
  if (!valu_lane_ops_per_s || !lds_bytes_per_s) return set_error(MCE_EINVAL, "null argument");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev)
    return set_error(MCE_EDEVICE, "no HIP device %d", device);
  MCE_HIP(hipSetDevice(device));
  hipDeviceProp_t prop;
  MCE_HIP(hipGetDeviceProperties(&prop, device));
  const uint32_t blocks = (uint32_t)prop.multiProcessorCount * 16;  // 16 waves of 4 per CU, 4 per SIMD
  uint32_t *in = nullptr, *out = nullptr;
  MCE_HIP(hipMalloc(&in, 64));
  if (hipMalloc(&out, sizeof(uint32_t) * blocks * 256) != hipSuccess) {
    (void)hipFree(in);
    return set_error(MCE_ENOMEM, "peak benchmark buffers");
  }
  const uint32_t hin[16] = {0x12345678u, 0x9abcdef0u, 0x0f1e2d3cu, 0x4b5a6978u, 0x87a5c3e1u, 0x13579bdfu,
                            0x2468ace0u, 0xfdb97531u, 0xf0f0f0ffu, 0x0ff00ff0u, 0, 0, 0, 0, 0, 0};
  int rc = MCE_OK;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  auto run = [&](auto kern, double &ms) -> int {
    hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, in, out);  // warm-up
    MCE_HIP(hipEventRecord(e0, 0));
    for (int r = 0; r < 3; ++r) hipLaunchKernelGGL(kern, dim3(blocks), dim3(256), 0, 0, in, out);
    MCE_HIP(hipEventRecord(e1, 0));
    MCE_HIP(hipEventSynchronize(e1));
    float f = 0;
    MCE_HIP(hipEventElapsedTime(&f, e0, e1));
    ms = f / 3.0;
    return MCE_OK;
  };
  double valu_ms = 0, lds_ms = 0;
  if (hipMemcpy(in, hin, sizeof(hin), hipMemcpyHostToDevice) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
      hipEventCreate(&e1) != hipSuccess)
    rc = set_error(MCE_EDEVICE, "peak benchmark set-up");
  if (!rc) rc = run(valu_peak_kernel, valu_ms);
  if (!rc) rc = run(lds_peak_kernel, lds_ms);
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  (void)hipFree(in);
  (void)hipFree(out);
  if (rc) return rc;
  const double lanes = (double)blocks * 256;
  *valu_lane_ops_per_s = lanes * kValuIters * kValuPerIter / (valu_ms * 1e-3);
  *lds_bytes_per_s = lanes * kLdsIters * 16.0 / (lds_ms * 1e-3);
  return MCE_OK;
}
