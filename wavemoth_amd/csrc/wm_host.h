// Host-side internals shared by plan.hip (single transforms) and batch.hip
// (batched transforms): error reporting, device allocation bookkeeping and
// the builders of the per-geometry tables.  Not part of the public ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "wm_internal.h"

namespace wm {

// Per-thread last error (wavemoth_last_error).
void set_err(const std::string &s);
void clear_err();
const char *last_err();

#define WM_TRY(expr, what)                                                        \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      ::wm::set_err(std::string(what) + ": " + hipGetErrorString(e_));            \
      return false;                                                               \
    }                                                                             \
  } while (0)

// Device allocations owned by one plan or batch and freed with it.
struct DevArena {
  std::vector<void *> allocs;
  size_t bytes = 0;
  void release() {
    for (void *v : allocs) (void)hipFree(v);
    allocs.clear();
    bytes = 0;
  }
};

template <typename T>
bool dalloc(DevArena &a, T **ptr, size_t count, const char *what) {
  void *v = nullptr;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  const hipError_t e = hipMalloc(&v, bytes);
  if (e != hipSuccess) {
    set_err(std::string("hipMalloc ") + what + " (" + std::to_string(bytes) +
            " bytes): " + hipGetErrorString(e));
    return false;
  }
  a.allocs.push_back(v);
  a.bytes += bytes;
  *ptr = (T *)v;
  return true;
}

template <typename T>
bool upload(DevArena &a, T **ptr, const std::vector<T> &h, const char *what) {
  if (!dalloc(a, ptr, h.size(), what)) return false;
  if (!h.empty())
    WM_TRY(hipMemcpy(*ptr, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice), what);
  return true;
}

// Work counts of one (Nside, lmax, mmax) geometry, from its seeds.
struct LegCounts {
  int64_t T_dense = 0;         // (l, m, ring-pair) triples, l >= m
  int64_t T_eps = 0;           // triples at or after each column's first l
  std::vector<int64_t> elems;  // per (m, odd): (l, ring) elements from first l
};

// Tables that depend on (Nside, lmax, mmax) only: recurrence coefficients,
// ring cosines and the per-(m, ring) seeds.  Leaves RL/W/K and the per-map
// buffers unset.
bool build_leg_geometry(DevArena &a, LegTables &t, int N, int L, int M, hipStream_t s,
                        LegCounts *cnt);
// Wave bounds for rings-per-lane RL (t.RL, t.W, t.wstart, t.wseedmax) and
// the number of (l, ring) slots the sweep executes with them.
bool build_leg_waves(DevArena &a, LegTables &t, int RL, hipStream_t s, int64_t *T_exec);

// Rings of the long path (kRingLong), in groups of one transform length P
// (P = h for a power-of-two h, else the Bluestein length): entries
// vecs[first .. first + count) of (ring pair, hemisphere); with K maps a
// group is count K vectors, (entry e, map k) = vector e K + k.  u, v: two
// scratch buffers of elems complex entries each.
struct LongGroup {
  int first, count, P;
  bool blue;
};
struct LongRings {
  std::vector<LongGroup> groups;
  LongVec *vecs = nullptr;
  double2 *u = nullptr, *v = nullptr;
  int64_t elems = 0;
};

// Host copy of a geometry's ring-pair table with each pair's LDS need and
// cost estimate (used to order work).
struct RingHost {
  std::vector<RingPair> pairs;
  std::vector<int> need;     // complex LDS slots of the ring kernel
  std::vector<double> cost;  // relative cost of the pair
  std::vector<int> kind;     // RingKind of the pair
  int64_t npix = 0;
  LongRings lr;              // the long path's rings and scratch
};

// Long path (ring_long.hip): whether P is a supported four-step length, the
// kernels' LDS limits, the ring stage of the long rings, and the plan-time
// Bluestein tables of long lengths (nent entries of length P: h values,
// chirp and vhat offsets on the device).
bool long_size_ok(int P);
hipError_t long_set_lds_limits();
hipError_t launch_long_rings(const LegTables &t, const RingTables &rt, const LongRings &lr,
                             double *map, hipStream_t s);
hipError_t launch_long_bluestein_prep(const double2 *twiddle, double2 *chirp, double2 *vhat,
                                      const int *d_h, const int64_t *d_choff,
                                      const int64_t *d_vhoff, int nent, int P, double2 *u,
                                      double2 *v, int64_t elems, hipStream_t s);

// Ring geometry, Bluestein tables and the single-transform launch groups.
// wide: plan for kRingWideMaps or more maps (wrapping rings then take the
// register-FFT kernel, which needs many maps to fill the device).
bool build_ring_geometry(DevArena &a, RingTables &rt, RingHost &host, int N, int M,
                         bool phase_shifts, bool wide, hipStream_t s);

}  // namespace wm
