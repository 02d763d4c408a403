// CPU stand-in for the HIP device API used by the engine's candidate and
// clip device code (hcs_query.h), for the CPU tests only: one wavefront is
// 64 std::threads; ballots, shuffles and barriers are std::barrier
// rendezvous of all 64, so every wave-level collective must be reached by
// all lanes, as on the GPU.
#pragma once
#include <math.h>
#include <stdint.h>

#include <barrier>
#include <cmath>
#include <cstdio>
#include <cstdlib>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline

struct int2 {
  int x, y;
};
inline int2 make_int2(int a, int b) { return int2{a, b}; }
struct int4 {
  int x, y, z, w;
};
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }
struct float4 {
  float x, y, z, w;
};
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
inline int min(int a, int b) { return a < b ? a : b; }

struct Dim3 {
  unsigned x, y, z;
};
extern thread_local Dim3 threadIdx;

struct WaveEmu {
  std::barrier<> bar{64};
  unsigned long long u[64];
  double d[64];
  long long l[64];
};
extern WaveEmu* g_wave;
extern thread_local int g_lane;

inline void __syncthreads() { g_wave->bar.arrive_and_wait(); }
inline unsigned long long __ballot(int p) {
  g_wave->u[g_lane] = p ? 1ull : 0ull;
  g_wave->bar.arrive_and_wait();
  unsigned long long m = 0;
  for (int i = 0; i < 64; ++i) m |= g_wave->u[i] << i;
  g_wave->bar.arrive_and_wait();
  return m;
}
inline int __any(int p) { return __ballot(p) != 0; }
inline unsigned long long clock64() { return 0; }
inline long long wall_clock64() { return 0; }
inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
inline double __shfl_xor(double v, int o, int = 64) {
  g_wave->d[g_lane] = v;
  g_wave->bar.arrive_and_wait();
  double r = g_wave->d[g_lane ^ o];
  g_wave->bar.arrive_and_wait();
  return r;
}
inline long long __shfl_xor(long long v, int o, int = 64) {
  g_wave->l[g_lane] = v;
  g_wave->bar.arrive_and_wait();
  long long r = g_wave->l[g_lane ^ o];
  g_wave->bar.arrive_and_wait();
  return r;
}
inline long __shfl_xor(long v, int o, int w = 64) { return (long)__shfl_xor((long long)v, o, w); }
inline unsigned long long __shfl(unsigned long long v, int src, int = 64) {
  g_wave->u[g_lane] = v;
  g_wave->bar.arrive_and_wait();
  unsigned long long r = g_wave->u[src];
  g_wave->bar.arrive_and_wait();
  return r;
}
inline unsigned __shfl(unsigned v, int src, int w = 64) { return (unsigned)__shfl((unsigned long long)v, src, w); }
inline unsigned __builtin_amdgcn_mbcnt_lo(unsigned m, unsigned x) {
  const int k = g_lane < 32 ? g_lane : 32;
  const unsigned mask = k == 32 ? 0xffffffffu : ((1u << k) - 1u);
  return x + __builtin_popcount(m & mask);
}
inline unsigned __builtin_amdgcn_mbcnt_hi(unsigned m, unsigned x) {
  const int k = g_lane < 32 ? 0 : g_lane - 32;
  const unsigned mask = (1u << k) - 1u;
  return x + __builtin_popcount(m & mask);
}
inline unsigned long long atomicAdd(unsigned long long* p, unsigned long long v) {
  return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST);
}
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline int atomicOr(int* p, int v) { return __atomic_fetch_or(p, v, __ATOMIC_SEQ_CST); }
