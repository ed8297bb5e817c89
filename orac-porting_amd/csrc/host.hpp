// host.hpp -- host runtime shared by the plain entry points (api.cpp) and the
// resident engine (engine.cpp): device buffers, PME plans, the context.
#pragma once
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/orac_hip.h"
#include "common.hpp"
#include "kernels.hpp"

namespace orac {

extern thread_local std::string g_last_error;

// ---------------------------------------------------------------- device buffers

struct DevBuf {
  void *p = nullptr;
  size_t bytes = 0;
  DevBuf() = default;
  DevBuf(const DevBuf &) = delete;
  DevBuf &operator=(const DevBuf &) = delete;
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void *ensure(size_t n) {
    if (n == 0) n = 8;
    if (n > bytes) {
      if (p) ORAC_HIP(hipFree(p));
      p = nullptr;
      size_t cap = std::max(n, bytes + bytes / 4);
      ORAC_HIP(hipMalloc(&p, cap));
      bytes = cap;
    }
    return p;
  }
};

// Named grow-only device buffers.  A buffer keeps its address until a larger
// size is asked for.
//
// Arena mode (arena != nullptr): get() carves consecutive 256-byte aligned
// pieces out of a fixed device region the caller owns, never allocates, and
// throws E_CAPACITY when the region is full; reset() starts over.  Uploads
// are never skipped in arena mode, since other calls reuse the region.
struct BufSet {
  std::map<std::string, std::unique_ptr<DevBuf>> bufs;
  std::map<std::string, std::vector<char>> cache;
  char *arena = nullptr;
  size_t arena_cap = 0, arena_used = 0;
  size_t bytes() const {
    size_t s = arena_used;
    for (auto &kv : bufs) s += kv.second->bytes;
    return s;
  }
  void reset() { arena_used = 0; }
  template <class T>
  T *get(const std::string &name, size_t n) {
    if (arena) {
      const size_t nb = std::max<size_t>(n * sizeof(T), 8), off = (arena_used + 255) & ~(size_t)255;
      ORAC_REQUIRE(off + nb <= arena_cap, E_CAPACITY,
                   "scratch region full: " + name + " needs " + std::to_string(off + nb) + " bytes of " +
                       std::to_string(arena_cap));
      arena_used = off + nb;
      return reinterpret_cast<T *>(arena + off);
    }
    auto &b = bufs[name];
    if (!b) b.reset(new DevBuf());
    return static_cast<T *>(b->ensure(n * sizeof(T)));
  }
  // upload n elements; cached uploads are skipped when the bytes are unchanged
  template <class T>
  T *put(const std::string &name, const T *host, size_t n, hipStream_t st, bool cached) {
    T *d = get<T>(name, n);
    size_t nb = n * sizeof(T);
    if (cached && !arena) {
      auto &c = cache[name];
      if (c.size() == nb && nb && std::memcmp(c.data(), host, nb) == 0) return d;
      c.assign(reinterpret_cast<const char *>(host), reinterpret_cast<const char *>(host) + nb);
    }
    if (nb) ORAC_HIP(hipMemcpyAsync(d, host, nb, hipMemcpyHostToDevice, st));
    return d;
  }
  template <class T>
  T *put_vec(const std::string &name, const std::vector<T> &v, hipStream_t st, bool cached) {
    return put<T>(name, v.data(), v.size(), st, cached);
  }
};

// ---------------------------------------------------------------- tables and plans

struct PmePlan {
  int K1 = 0, K2 = 0, K3 = 0, order = 0;
  FftAxis ax[3];
  std::vector<double> bsp[3];
};

void bsp_moduli(int order, int nfft, double *out);
bool lc_index(int indmax, int ncx, int ncy, int ncz, int32_t *nind, int32_t *indxi, int32_t *indxj, int32_t *indxk,
              double ctoff, const double *co);
double fix_scale_for(const double *q, int n);
void exclusions_from_mapnl(const int32_t *mapnl, int nato, const std::vector<int> &atomg, std::vector<int> &off,
                           std::vector<int> &idx, std::vector<unsigned char> &inter);
size_t mapnl_length(const int32_t *mapnl, int nato);
void build_pme_plan(BufSet &B, const std::string &pfx, PmePlan &P, int k1, int k2, int k3, int order,
                    hipStream_t st);
// descriptor of one system's reciprocal sum with its grids and partial-sum buffers
PmeDesc pme_desc_base(BufSet &B, const std::string &pfx, const PmePlan &P, int natoms, const double *oc,
                      double volume, double alpha, double rkcut, bool fp64, const Tuning &tu, hipStream_t st);
// apply the tuning's FFT tiling and spread path to a descriptor built by pme_desc_base
void pme_apply_tuning(PmeDesc &d, const Tuning &tu);
CellGrid make_cellgrid(BufSet &B, const std::string &pfx, int ncx, int ncy, int ncz, const int32_t *nind,
                       const int32_t *indxi, const int32_t *indxj, const int32_t *indxk, double rcut,
                       const double *co, int half, hipStream_t st);
// the A&S 7.1.26 erfc of force.f:123-125, 443-451
double erfc_as(double x);

struct Engine;  // resident systems and batches (engine.cpp)
void engine_destroy(Engine *e);

}  // namespace orac

struct orac_hip_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipStream_t side = nullptr;  // reciprocal sum beside the direct sum (same, default priority)
  orac::BufSet B;
  orac::BufSet S;  // arena over this device's static scratch region (the no-allocation entry point)
  orac::PmePlan plan;
  bool plan_ok = false;
  orac::Tuning tuning;
  orac::Engine *engine = nullptr;
};

namespace orac {

template <class F>
int guarded(F &&f) {
  try {
    f();
    return OK;
  } catch (const Error &e) {
    g_last_error = e.what();
    return e.code;
  } catch (const std::exception &e) {
    g_last_error = e.what();
    return E_HIP;
  } catch (...) {
    g_last_error = "unknown error";
    return E_HIP;
  }
}

// ---- alchemy (force_alchemy.f, alch_set in force_alchemy_shifted.f:732-790)
// Atoms with equal (rmin, epsm, ss_alchemy) share a class; the pair table
// holds {1 / sigma^6, epsmin, typeaij + 4, sigma} for every class pair.
struct AlchTables {
  std::vector<int> cls;  // class of each atom
  int ncls = 0;
  std::vector<double4> pair;
};
AlchTables alch_tables(int n, const int32_t *ss_alchemy, const double *rmin, const double *epsm, int opls);
// alch_fact(1:2, -4:4) as fact[2 * (typeaij + 4) + k]
void alch_set_host(double fact[18], double ladd, double lrm, bool intramol, bool ljintra, bool qq);
// the pair kernel's [9][8] factor rows from lam = ladd, lrm, ladd0, lrm0, laddq, lrmq, laddq0, lrmq0;
// (qx, ljx) are (qsoft, ljsoft), or (qshift, ljshift) for the shifted potential
void alch_factors(const double lam[8], bool intramol, bool ljintra, double qx, double ljx, bool shifted,
                  double out[72]);
// charges at (laddq, lrmq): (1 - |lambdaq|) q, end points exact (mts_furier_alchemy.f:98-111)
void alch_charges(int n, const double *q, const int32_t *ss_alchemy, double laddq, double lrmq, double *out);

// ---- ERF_CORR tables (int_corr_erf_spline.f, erf_corr_cutoff.f, cubspl)
void cubspl(const double *tau, double *c, int n, int ibcbeg, int ibcend);
void int_corr_erf_spline(double rmin, double rmax, int nbin, double alpha, double kcut, double *erf_arr_corr,
                         double *tau);
double erf_corr_cutoff(const double *oc, int nfft1, int nfft2, int nfft3);

inline void set_device(orac_hip_ctx *c) {
  ORAC_REQUIRE(c != nullptr, E_ARG, "null context");
  ORAC_HIP(hipSetDevice(c->device));
}

}  // namespace orac
