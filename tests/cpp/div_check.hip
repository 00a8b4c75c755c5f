// div_check.hip - checks the engine's correctly rounded division (rcp_refined, div_by_rcp,
// recip_rn, div3_rn, pl_div3_rn in tecs_amd/csrc/tecs_device.hpp) against IEEE division (hipcc with
// -fhip-fp32-correctly-rounded-divide-sqrt) on the GPU, bit for bit. Test infrastructure for
// tests/test_gpu_fast_div.py; built by tecs_amd/_build.py into tests/cpp/_build.
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../tecs_amd/csrc/tecs_device.hpp"

namespace {

using tecs_amd::div_by_rcp;
using tecs_amd::div3_rn;
using tecs_amd::len2;
using tecs_amd::pl_div3_rn;
using tecs_amd::rcp_refined;
using tecs_amd::recip_rn;

__device__ __forceinline__ bool same(float a, float b) {
  return __float_as_uint(a) == __float_as_uint(b) || (a != a && b != b);
}

__device__ __forceinline__ void note_bad(unsigned long long* bad, unsigned long long n,
                                         unsigned long long* first, unsigned long long tag) {
  if (n) {
    atomicAdd(bad, n);
    atomicMin(first, tag);
  }
}

// rcp_refined(d) == 1 / d for the d bit patterns [lo, lo + count)
__global__ void k_rcp(uint32_t lo, uint32_t count, unsigned long long* bad, unsigned long long* first) {
  unsigned long long n = 0, f = ~0ull;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    const float d = __uint_as_float(lo + i);
    if (!same(rcp_refined(d), 1.0f / d)) {
      ++n;
      f = min(f, (unsigned long long)(lo + i));
    }
  }
  note_bad(bad, n, first, f);
}

// div_by_rcp(x, d, rcp_refined(d)) == x / d for d = 1 + (d0 + t) * 2^-23 (one thread each,
// t < d_count) and every x significand in [1, 2); the first bad pair is (d bits << 32 | x bits)
__global__ void k_div(uint32_t d0, uint32_t d_count, unsigned long long* bad, unsigned long long* first) {
  const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= d_count) return;
  const uint32_t db = 0x3f800000u + d0 + t;
  const float d = __uint_as_float(db);
  const float y = rcp_refined(d);
  unsigned long long n = 0, f = ~0ull;
  for (uint32_t xb = 0x3f800000u; xb < 0x40000000u; ++xb) {
    const float x = __uint_as_float(xb);
    if (!same(div_by_rcp(x, d, y), x / d)) {
      ++n;
      f = min(f, ((unsigned long long)db << 32) | xb);
    }
  }
  note_bad(bad, n, first, f);
}

__device__ __forceinline__ uint32_t mix(uint64_t v) {
  v ^= v >> 33;
  v *= 0xff51afd7ed558ccdull;
  v ^= v >> 33;
  v *= 0xc4ceb9fe1a85ec53ull;
  v ^= v >> 33;
  return (uint32_t)v;
}

// The public functions on pseudo-random operands: mode 0 draws every bit pattern (all
// exponents, zeros, subnormals, infinities, NaNs), mode 1 draws x and d with exponents near the
// domain's edges and random signs; checks recip_rn(d) == 1 / d and div3_rn == x / d.
__global__ void k_random(uint64_t seed, uint32_t count, int mode, unsigned long long* bad,
                         unsigned long long* first) {
  unsigned long long n = 0, f = ~0ull;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    uint32_t xb[3], db = mix(seed * 0x9e3779b97f4a7c15ull + 4ull * i);
    for (int c = 0; c < 3; ++c) xb[c] = mix(seed * 0x9e3779b97f4a7c15ull + 4ull * i + 1 + c);
    if (mode == 1) {
      // exponent fields 96..160 for d, 20..240 for x; significand and sign random
      db = (db & 0x807fffffu) | ((96u + (db >> 23) % 65u) << 23);
      for (int c = 0; c < 3; ++c) {
        const uint32_t e = (xb[c] & 0x100u) ? 0u : (20u + (xb[c] >> 23) % 221u);
        xb[c] = (xb[c] & 0x807fffffu) | (e << 23);
        if (e == 0u) xb[c] &= 0x80000000u;  // a signed zero
      }
    }
    const float d = __uint_as_float(db);
    const float x[3] = {__uint_as_float(xb[0]), __uint_as_float(xb[1]), __uint_as_float(xb[2])};
    float q[3];
    div3_rn(x, d, q);
    bool ok = same(recip_rn(d), 1.0f / d);
    for (int c = 0; c < 3; ++c) ok = ok && same(q[c], x[c] / d);
    if (!ok) {
      ++n;
      f = min(f, (unsigned long long)i);
    }
  }
  note_bad(bad, n, first, f);
}

// pl_div3_rn(v, F, d) == (v * F) / d with d = |v| computed as the force computes it (len2 and
// sqrtf), for pseudo-random displacements and scales: each component and F get a random sign
// and an exponent field drawn from [0, 255] (zeros, subnormals, infinities and NaNs included)
// or, for half of the draws, from [100, 154]; pairs with d == 0 are skipped (the force is 0).
__global__ void k_random_pl(uint64_t seed, uint32_t count, unsigned long long* bad,
                            unsigned long long* first) {
  unsigned long long n = 0, f = ~0ull;
  for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
    uint32_t b[4];
    for (int c = 0; c < 4; ++c) {
      b[c] = mix(seed * 0x9e3779b97f4a7c15ull + 5ull * i + c);
      if (b[c] & 0x100u) b[c] = (b[c] & 0x807fffffu) | ((100u + (b[c] >> 23) % 55u) << 23);
      if ((b[c] & 0x3e00u) == 0u) b[c] &= 0x80000000u;  // a signed zero, 1 in 32
    }
    const float v[3] = {__uint_as_float(b[0]), __uint_as_float(b[1]), __uint_as_float(b[2])};
    const float F = __uint_as_float(b[3]);
    const float d = sqrtf(len2(v[0], v[1], v[2]));
    if (!(d > 0.0f)) continue;
    float q[3];
    pl_div3_rn(v, F, d, q);
    bool ok = true;
    for (int c = 0; c < 3; ++c) ok = ok && same(q[c], v[c] * F / d);
    if (!ok) {
      ++n;
      f = min(f, (unsigned long long)i);
    }
  }
  note_bad(bad, n, first, f);
}

struct Counters {
  unsigned long long bad, first;
};

int run(void (*launch)(Counters*), unsigned long long* bad, unsigned long long* first) {
  Counters* c = nullptr;
  if (hipMalloc(&c, sizeof(Counters)) != hipSuccess) return 1;
  const Counters init{0ull, ~0ull};
  hipMemcpy(c, &init, sizeof(init), hipMemcpyHostToDevice);
  launch(c);
  Counters h{};
  const hipError_t e = hipMemcpy(&h, c, sizeof(h), hipMemcpyDeviceToHost);
  hipFree(c);
  if (e != hipSuccess || hipGetLastError() != hipSuccess) return 2;
  *bad = h.bad;
  *first = h.first;
  return 0;
}

}  // namespace

extern "C" {

// Mismatches of rcp_refined over the d bit patterns [lo, lo + count).
int div_check_rcp(uint32_t lo, uint32_t count, unsigned long long* bad, unsigned long long* first) {
  static uint32_t s_lo, s_count;
  s_lo = lo;
  s_count = count;
  return run([](Counters* c) {
    hipLaunchKernelGGL(k_rcp, dim3(4096), dim3(256), 0, 0, s_lo, s_count, &c->bad, &c->first);
  }, bad, first);
}

// Mismatches of div_by_rcp over d significands [d0, d0 + d_count) x all x significands.
int div_check_div(uint32_t d0, uint32_t d_count, unsigned long long* bad, unsigned long long* first) {
  if (d0 + (uint64_t)d_count > (1u << 23)) return 3;
  static uint32_t s_d0, s_count;
  s_d0 = d0;
  s_count = d_count;
  return run([](Counters* c) {
    hipLaunchKernelGGL(k_div, dim3((s_count + 255) / 256), dim3(256), 0, 0, s_d0, s_count, &c->bad,
                       &c->first);
  }, bad, first);
}

// Mismatches of recip_rn / div3_rn over `count` pseudo-random operand sets.
int div_check_random(uint64_t seed, uint32_t count, int mode, unsigned long long* bad,
                     unsigned long long* first) {
  static uint64_t s_seed;
  static uint32_t s_count;
  static int s_mode;
  s_seed = seed;
  s_count = count;
  s_mode = mode;
  return run([](Counters* c) {
    hipLaunchKernelGGL(k_random, dim3(4096), dim3(256), 0, 0, s_seed, s_count, s_mode, &c->bad,
                       &c->first);
  }, bad, first);
}

// Mismatches of pl_div3_rn over `count` pseudo-random (displacement, scale) pairs.
int div_check_random_pl(uint64_t seed, uint32_t count, unsigned long long* bad,
                        unsigned long long* first) {
  static uint64_t s_seed;
  static uint32_t s_count;
  s_seed = seed;
  s_count = count;
  return run([](Counters* c) {
    hipLaunchKernelGGL(k_random_pl, dim3(4096), dim3(256), 0, 0, s_seed, s_count, &c->bad,
                       &c->first);
  }, bad, first);
}

}  // extern "C"
