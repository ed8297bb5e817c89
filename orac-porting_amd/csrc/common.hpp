// common.hpp -- shared device/host definitions of the ORAC nonbonded path on gfx950.
//
// Conventions (the reference's): reduced coordinates xp = oc.x with the box
// spanning [-1,1); co/oc/recip are 3x3 column-major (m[i + 3*j] = m(i+1,j+1));
// grids are Q(K1,K2,K3), K1 fastest; the half-complex grid is FQ(K1/2+1,K2,K3).
// Every kernel reads a descriptor array and takes its system from blockIdx.y,
// so one launch serves one system or a batch of replicas.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

namespace orac {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define ORAC_HIP(call)                                                                             \
  do {                                                                                             \
    hipError_t e_ = (call);                                                                        \
    if (e_ != hipSuccess)                                                                          \
      throw ::orac::Error(3, std::string(#call " failed: ") + hipGetErrorString(e_));              \
  } while (0)

#define ORAC_REQUIRE(cond, code, msg)                                                              \
  do {                                                                                             \
    if (!(cond)) throw ::orac::Error((code), (msg));                                               \
  } while (0)

// status codes returned through the C ABI
enum Status { OK = 0, E_ARG = 1, E_CAPACITY = 2, E_HIP = 3, E_UNSUPPORTED = 4, E_STATE = 5 };

constexpr int MAX_ORDER = 8;       // B-spline order limit of the device kernels
constexpr int MAX_FFT_LEN = 256;   // longest pencil an FFT workgroup holds in LDS
constexpr int MAX_RADICES = 12;

// ---------------------------------------------------------------- PME

// Per-axis FFT plan: radix sequence and twiddle table tw[m] = exp(-2 pi i m / n).
struct FftAxis {
  int n;
  int nr;
  int radix[MAX_RADICES];
  const double2 *tw64;
  const float2 *tw32;
};

// One system's reciprocal-space state.
struct PmeDesc {
  int natoms;
  int K1, K2, K3, order;
  double recip[9];          // recip(i,j) = 0.5*oc(j,i)
  double volume, pi_f32, fac, rkcut2;
  double fixscale, fixinv;  // grid fixed point: value = int64 * fixinv
  const double *xp, *yp, *zp;   // reduced coordinates
  const double *co;             // 3x3 on device (cartesian = co . xp)
  const double *q;
  unsigned long long *Qfix;     // K1*K2*K3 int64 two's complement
  void *FQ;                     // complex half grid, double2 or float2
  void *Qr;                     // real grid after c2r, double or float
  const double *bsp1, *bsp2, *bsp3;
  FftAxis ax1, ax2, ax3;
  double *fx, *fy, *fz;         // PME forces
  int accumulate;               // 0: fx = f_pme, 1: fx += f_pme
  int tpx, tpy, tpz;            // pencils per FFT workgroup along K1, K2, K3
  int gx, gy;                   // workgroups of the K1 and K2 passes
  double *conv_part;            // per-block energy/virial partials [conv_nblk][8]
  int conv_nblk;                // workgroups of the K3 pass
  // molecular-virial correction (fft_pme.f:65-106)
  const int *atomp;             // 0-based molecule of each atom
  const double *xpcm, *ypcm, *zpcm;
  double *gather_part;          // per-block stc partials [nblk][9]
  int gather_nblk;
  // LDS-tiled spread: grid tiles of st1 x st2 x st3 points (st = 0: global-atomic spread)
  int st1, st2, st3, nb1, nb2, nb3;
  int *bin_cnt, *bin_start;     // [nb1*nb2*nb3] atoms per tile bin, exclusive prefix
  int2 *atom_bin;               // [natoms] {bin, position in bin}
  double4 *afr;                 // [natoms] {fr1, fr2, fr3, q}
  double4 *sfr;                 // [natoms] the same, sorted by bin
  double *swt;                  // [natoms][3][order] B-spline weights of the sorted atoms (exact T values)
  int tiled;                    // 1: LDS-tiled spread, 0: global-atomic spread
  double *out;                  // [20]: eer, vir (3x3 col-major), stc1..stc9
  int energy_only;              // 1: spread, forward FFT and energy only (a lambda0 charge grid)
  double *fx_dst, *fy_dst, *fz_dst;  // when set, launch_pme_combine adds fx/fy/fz into these
  const double *cfx, *cfy, *cfz;     // and then these (the system's correction forces), when set
};

// ---------------------------------------------------------------- helpers

// Global-memory view of a device pointer.  Pointers a kernel reads out of a
// descriptor are generic (flat) to the compiler; accesses through this view
// are global ones: no aperture check, a uniform base in SGPRs, and waits on
// the vector-memory counter only, so they overlap with LDS traffic.
template <typename T>
__device__ __forceinline__ __attribute__((address_space(1))) T *gmem(T *p) {
  return (__attribute__((address_space(1))) T *)p;
}

// HIP vector types (double4, int4, ...) through the global view: their
// member functions expect generic objects, so the native vector is moved.
template <typename V>
__device__ __forceinline__ V gload(const V *p) {
  typedef typename V::Native_vec_ N;
  const N n = *(__attribute__((address_space(1))) const N *)p;
  V v;
  __builtin_memcpy(&v, &n, sizeof(V));
  return v;
}
// Element i of an array of HIP vector types as a uniform base plus a 32-bit
// byte offset (i * sizeof(V) < 2^32), the form of a global load with a
// scalar base and one offset VGPR.
template <typename V>
__device__ __forceinline__ V gload_at(const V *base, unsigned i) {
  return gload(reinterpret_cast<const V *>(reinterpret_cast<const char *>(base) + (i * (unsigned)sizeof(V))));
}
template <typename V>
__device__ __forceinline__ void gstore(V *p, const V &v) {
  typedef typename V::Native_vec_ N;
  N n;
  __builtin_memcpy(&n, &v, sizeof(V));
  *(__attribute__((address_space(1))) N *)p = n;
}

__host__ __device__ inline int wrap_idx(int k0, int K) { return k0 >= 0 ? k0 : k0 + K; }

// Fortran ANINT/DNINT: round half away from zero.
__host__ __device__ inline double fnint(double x) { return round(x); }

// Cardinal B-spline weights and derivatives (Essmann recursion,
// pme_fftw3.f90:537-576): derivative taken before the last pass.
// 1 / (k - 1) for the passes k = 2..MAX_ORDER: the correctly rounded value,
// as the division gives it, in either precision
template <typename T>
__host__ __device__ inline T recip_km1(int k) {
  switch (k) {
    case 2: return T(1.0);
    case 3: return T(1.0 / 2.0);
    case 4: return T(1.0 / 3.0);
    case 5: return T(1.0 / 4.0);
    case 6: return T(1.0 / 5.0);
    case 7: return T(1.0 / 6.0);
    case 8: return T(1.0 / 7.0);
    default: return T(1) / T(k - 1);
  }
}

template <typename T>
__host__ __device__ inline void bspline_pass(T *c, T w, int k) {
  T div = recip_km1<T>(k);
  c[k - 1] = div * w * c[k - 2];
  for (int j = 1; j <= k - 2; ++j) c[k - j - 1] = div * ((w + T(j)) * c[k - j - 2] + (T(k - j) - w) * c[k - j - 1]);
  c[0] = div * (T(1) - w) * c[0];
}

template <typename T>
__host__ __device__ inline void fill_bspline(T w, int order, T *c, T *d) {
  c[order - 1] = T(0);
  c[1] = w;
  c[0] = T(1) - w;
  for (int k = 3; k <= order - 1; ++k) bspline_pass(c, w, k);
  d[0] = -c[0];
  for (int j = 2; j <= order; ++j) d[j - 1] = c[j - 2] - c[j - 1];
  bspline_pass(c, w, order);
}

// 64-bit fixed point: round-to-nearest of v*scale, two's complement.
__device__ inline unsigned long long to_fix(double v, double scale) {
  return (unsigned long long)__double2ll_rn(v * scale);
}

template <typename T>
__device__ inline T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Exclusive prefix sum of in[0..n) into out by one 1024-thread workgroup, in
// fixed order; returns the total to every thread.
__device__ inline int block_exclusive_scan(const int *in, int *out, int n) {
  __shared__ int wsum[16];
  __shared__ int carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  for (int base = 0; base < n; base += 1024) {
    const int i = base + threadIdx.x;
    const int v = i < n ? in[i] : 0;
    int x = v;
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      int y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) wsum[wid] = x;
    __syncthreads();
    if (threadIdx.x == 0) {
      int s = 0;
      for (int w = 0; w < 16; ++w) {
        int t = wsum[w];
        wsum[w] = s;
        s += t;
      }
    }
    __syncthreads();
    const int excl = carry + wsum[wid] + x - v;
    if (i < n) out[i] = excl;
    __syncthreads();
    if (threadIdx.x == 1023) carry = excl + v;
    __syncthreads();
  }
  const int total = carry;
  __syncthreads();  // every thread has read carry before a later call resets it
  return total;
}

// Workgroups are dealt to the 8 XCDs round robin (hardware id h runs on XCD
// h % 8).  Mapping h to logical block (h % 8) * (n / 8) + h / 8 gives each XCD
// one contiguous range of logical blocks, so a system's blocks share an L2.
__device__ inline int xcd_remap(int h, int n) {
  const int q = n >> 3;
  if (h >= (q << 3)) return h;
  return (h & 7) * q + (h >> 3);
}

}  // namespace orac
