// CPU stand-in for the HIP device API used by the engine's candidate and
// clip device code (hcs_query.h), for the CPU tests only: one wavefront is
// 64 std::threads; ballots, shuffles and barriers are std::barrier
// rendezvous of all 64, so every wave-level collective must be reached by
// all lanes, as on the GPU.  Every collective also checks that all 64 lanes
// reached it from the same source site (file, line, column): lanes that meet
// at different ballots, shuffles or barriers (a collective under divergent
// control flow, which the GPU would run with a partial lane mask) abort the
// emulation instead of pairing up silently.
#pragma once
#include <math.h>
#include <stdint.h>

#include <barrier>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <source_location>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
/* one wave at a time: a function-local static is the wave's LDS */
#define __shared__ static

struct int2 {
  int x, y;
};
inline int2 make_int2(int a, int b) { return int2{a, b}; }
struct int4 {
  int x, y, z, w;
};
inline int4 make_int4(int a, int b, int c, int d) { return int4{a, b, c, d}; }
struct uint2 {
  unsigned x, y;
};
inline uint2 make_uint2(unsigned a, unsigned b) { return uint2{a, b}; }
struct float4 {
  float x, y, z, w;
};
inline float4 make_float4(float a, float b, float c, float d) { return float4{a, b, c, d}; }
inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }

struct Dim3 {
  unsigned x, y, z;
};
extern thread_local Dim3 threadIdx;

struct WaveEmu {
  std::barrier<> bar{64};
  unsigned long long u[64];
  double d[64];
  long long l[64];
  std::source_location site[64];
};
extern WaveEmu* g_wave;
extern thread_local int g_lane;

using emu_site = std::source_location;
/* First rendezvous of a collective: publish the lane's site, wait for all
 * lanes, and abort unless every lane came from the same site. */
inline void emu_meet(const emu_site& at) {
  g_wave->site[g_lane] = at;
  g_wave->bar.arrive_and_wait();
  for (int i = 0; i < 64; ++i) {
    const emu_site& o = g_wave->site[i];
    if (o.line() != at.line() || o.column() != at.column() || std::strcmp(o.file_name(), at.file_name()) != 0) {
      std::fprintf(stderr, "divergent collective: lane %d at %s:%u:%u, lane %d at %s:%u:%u\n", g_lane, at.file_name(),
                   at.line(), at.column(), i, o.file_name(), o.line(), o.column());
      std::abort();
    }
  }
}

inline void __syncthreads(const emu_site& at = emu_site::current()) {
  emu_meet(at);
  g_wave->bar.arrive_and_wait();
}
inline unsigned long long __ballot(int p, const emu_site& at = emu_site::current()) {
  g_wave->u[g_lane] = p ? 1ull : 0ull;
  emu_meet(at);
  unsigned long long m = 0;
  for (int i = 0; i < 64; ++i) m |= g_wave->u[i] << i;
  g_wave->bar.arrive_and_wait();
  return m;
}
inline int __any(int p, const emu_site& at = emu_site::current()) { return __ballot(p, at) != 0; }
inline unsigned long long clock64() { return 0; }
inline long long wall_clock64() { return 0; }
inline int __popcll(unsigned long long m) { return __builtin_popcountll(m); }
inline double __shfl_xor(double v, int o, int = 64, const emu_site& at = emu_site::current()) {
  g_wave->d[g_lane] = v;
  emu_meet(at);
  double r = g_wave->d[g_lane ^ o];
  g_wave->bar.arrive_and_wait();
  return r;
}
inline long long __shfl_xor(long long v, int o, int = 64, const emu_site& at = emu_site::current()) {
  g_wave->l[g_lane] = v;
  emu_meet(at);
  long long r = g_wave->l[g_lane ^ o];
  g_wave->bar.arrive_and_wait();
  return r;
}
inline long __shfl_xor(long v, int o, int w = 64, const emu_site& at = emu_site::current()) {
  return (long)__shfl_xor((long long)v, o, w, at);
}
inline unsigned long long __shfl(unsigned long long v, int src, int = 64, const emu_site& at = emu_site::current()) {
  g_wave->u[g_lane] = v;
  emu_meet(at);
  unsigned long long r = g_wave->u[src];
  g_wave->bar.arrive_and_wait();
  return r;
}
inline double __shfl(double v, int src, int = 64, const emu_site& at = emu_site::current()) {
  g_wave->d[g_lane] = v;
  emu_meet(at);
  double r = g_wave->d[src];
  g_wave->bar.arrive_and_wait();
  return r;
}
inline unsigned __shfl(unsigned v, int src, int w = 64, const emu_site& at = emu_site::current()) {
  return (unsigned)__shfl((unsigned long long)v, src, w, at);
}
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
inline unsigned long long atomicMin(unsigned long long* p, unsigned long long v) {
  unsigned long long o = __atomic_load_n(p, __ATOMIC_SEQ_CST);
  while (v < o && !__atomic_compare_exchange_n(p, &o, v, false, __ATOMIC_SEQ_CST, __ATOMIC_SEQ_CST)) {
  }
  return o;
}
inline long long __double_as_longlong(double x) {
  long long r;
  std::memcpy(&r, &x, 8);
  return r;
}
inline double __longlong_as_double(long long x) {
  double r;
  std::memcpy(&r, &x, 8);
  return r;
}
