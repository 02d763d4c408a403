// CPU build of the flat-sensor samplers (hcs_taxel_dev.h), test
// infrastructure only: every taxel of one sensor sampled by one emulated
// 64-lane wave, with the ray-major sampler and with the fan-major one, from
// the kept polygons of a contact step (oracle records, same layout as the
// device's).
#include <hip/hip_runtime.h>

#include <cmath>
#include <thread>
#include <vector>

#include "hcs_taxel_dev.h"

thread_local Dim3 threadIdx;
thread_local int g_lane;
WaveEmu* g_wave;

using namespace hcs;

extern "C" {

/* polys: np x HCS_PREC records; keys: np x 3 (pair, tet, triangle);
 * sensor: X (12), half sizes (3), pitch; S rays per side; window, sigma.
 * mode 0 ray-major, 1 fan-major.  Every polygon whose footprint covers a
 * taxel is in that taxel's list, in record order (chunk = 64 polygons, so
 * lists longer than that exercise the chunk loops).  out: cx x cy taxels,
 * index x + cx y. */
int emu_taxels(const double* polys, int64_t np, const int64_t* keys, const double* sensor, int S, int window,
               double sigma, int cx, int cy, int mode, double* out) {
  TaxelGridDev g;
  for (int k = 0; k < 12; ++k) g.X[k] = sensor[k];
  g.xs = sensor[12], g.ys = sensor[13], g.zs = sensor[14], g.res = sensor[15], g.sigma = sigma;
  g.S = S, g.window = window, g.cx = cx, g.cy = cy;
  if (mode == 1 && S * S > HCS_TX_RAYS) return -1;
  std::vector<std::vector<int>> bins((size_t)cx * cy);
  for (int64_t q = 0; q < np; ++q) {
    int x0, x1, y0, y1;
    footprint(polys + q * HCS_PREC, g, x0, x1, y0, y1);
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) bins[x + cx * y].push_back((int)q);
  }
  for (int cell = 0; cell < cx * cy; ++cell) {
    const std::vector<int>& list = bins[cell];
    const int x = cell % cx, y = cell / cx;
    WaveEmu wave;
    g_wave = &wave;
    double v[64];
    std::vector<std::thread> th;
    for (int lane = 0; lane < 64; ++lane)
      th.emplace_back([&, lane] {
        g_lane = lane;
        threadIdx.x = lane;
        const int* lp = list.empty() ? nullptr : list.data();
        v[lane] = mode == 1 ? sample_cell_scatter(polys, keys, g, 0, (unsigned)list.size(), lp, x, y)
                            : sample_cell<double>(polys, nullptr, keys, g, nullptr, 0, (unsigned)list.size(), lp, x, y);
      });
    for (auto& t : th) t.join();
    for (int lane = 1; lane < 64; ++lane)
      if (v[lane] != v[0]) return -2; /* every lane returns the wave's sum */
    out[cell] = v[0];
  }
  return 0;
}

}  // extern "C"
